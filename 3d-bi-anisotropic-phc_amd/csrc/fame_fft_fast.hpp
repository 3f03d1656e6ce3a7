// Compile-time specialised FFT passes for the hot grid sizes.
//
// Every line of length N = A*B is transformed by the four-step factorisation
// (Gentleman & Sande 1966; Bailey 1990):  with x[t + B j] (t < B, j < A),
//     X[k + A q] = sum_t W_N^{t k} W_B^{t q} sum_j x[t + B j] W_A^{j k}.
// Phase A: one thread per column t runs the A-point DFT over j in registers,
// multiplies by W_N^{t k} and writes back to the same LDS positions.  Phase B:
// one thread per row k runs the B-point DFT over t in registers and emits
// X[k + A q] directly in natural order.  The A- and B-point DFTs are themselves
// built recursively from radix-2/3/4/5 kernels in registers (rdft), so every
// index is a compile-time constant.  Twiddles come from an LDS copy of the
// line's table W_N^m = e^{-2 pi i m/N} (host-computed in long double).
#pragma once
#include "fame_kernels.hpp"

namespace fame {

template <int S>
constexpr int split_factor() {
  return (S % 4 == 0 && S > 4) ? 4 : (S % 2 == 0 && S > 2) ? 2 : (S % 3 == 0 && S > 3) ? 3 : (S % 5 == 0 && S > 5) ? 5 : S;
}

template <int DIR>
__device__ __forceinline__ double2 twmul(double2 v, double2 w) {
  return (DIR < 0) ? cmul(v, w) : cmulc(v, w);
}

// In-register DFT of length S (natural order in and out).  W_S^m = tw[m * ts].
template <int S, int DIR>
__device__ __forceinline__ void rdft(double2* v, const double2* __restrict__ tw, int ts) {
  if constexpr (S == 1) {
    return;
  } else if constexpr (S == 2) {
    dft2<DIR>(v[0], v[1]);
  } else if constexpr (S == 3) {
    dft3<DIR>(v[0], v[1], v[2]);
  } else if constexpr (S == 4) {
    dft4<DIR>(v[0], v[1], v[2], v[3]);
  } else if constexpr (S == 5) {
    dft5<DIR>(v[0], v[1], v[2], v[3], v[4]);
  } else {
    constexpr int P = split_factor<S>();
    constexpr int Q = S / P;
    static_assert(P > 1 && Q > 1, "rdft: S must factor into 2,3,4,5");
    double2 w[S];
#pragma unroll
    for (int t = 0; t < Q; ++t) {
      double2 u[P];
#pragma unroll
      for (int j = 0; j < P; ++j) u[j] = v[t + Q * j];
      rdft<P, DIR>(u, tw, ts * Q);
#pragma unroll
      for (int k = 0; k < P; ++k) w[t + Q * k] = (t * k == 0) ? u[k] : twmul<DIR>(u[k], tw[(t * k) * ts]);
    }
#pragma unroll
    for (int k = 0; k < P; ++k) {
      double2 u[Q];
#pragma unroll
      for (int t = 0; t < Q; ++t) u[t] = w[t + Q * k];
      rdft<Q, DIR>(u, tw, ts * P);
#pragma unroll
      for (int q = 0; q < Q; ++q) v[k + P * q] = u[q];
    }
  }
}

// ---------------------------------------------------------------------------
// xy-plane pass (N1 = N2 = N = A*B): one work-group per (z-plane, component, batch).
// LDS: the plane (N*N) + the twiddle table (N).
// ---------------------------------------------------------------------------
template <int N, int A, int B, int DIR, int TWIST, int NT>
__global__ void __launch_bounds__(NT) k_xyfast(Geo g, const double2* __restrict__ in, long long in_bstride,
                                               double2* __restrict__ out, long long out_bstride,
                                               const double2* __restrict__ tws, int tws_bstride,
                                               const double2* __restrict__ twg, double scale,
                                               const CgState* __restrict__ cgst) {
  static_assert(A * B == N, "k_xyfast: N = A*B");
  // odd row pitch: lanes that walk down a column hit distinct LDS banks
  constexpr int PT = (N % 2 == 0) ? N + 1 : N;
  extern __shared__ double2 sm[];
  double2* pl = sm;
  double2* tw = sm + N * PT;
  const int i3 = blockIdx.x, f = blockIdx.y, b = blockIdx.z;
  if (cgst && cgst[b].active == 0.0) return;  // inside a fused CG: item b has converged
  const long long base = (long long)f * g.N + (long long)i3 * (N * N);
  const double2* __restrict__ src = in + b * in_bstride + base;
  double2* __restrict__ dst = out + b * out_bstride + base;
  const double2* __restrict__ S = tws + (long long)b * tws_bstride;
  const double2* __restrict__ twn = twg + g.ax[0].tw_off;
  for (int t = threadIdx.x; t < N; t += NT) tw[t] = twn[t];
  double2 sz = make_double2(1, 0);
  if (TWIST) sz = S[2 * N + i3];
  __syncthreads();

  // x phase A: item (r, t), t fastest; global -> regs -> LDS (positions t + B k of row r)
  for (int it = threadIdx.x; it < N * B; it += NT) {
    const int r = it / B, t = it - (it / B) * B;
    double2 u[A];
    double2 syz = make_double2(1, 0);
    if (TWIST == 2) syz = cmul(S[N + r], sz);
#pragma unroll
    for (int j = 0; j < A; ++j) {
      double2 v = src[r * N + t + B * j];
      if (TWIST == 2) v = cmulc(v, cmul(S[t + B * j], syz));
      u[j] = v;
    }
    rdft<A, DIR>(u, tw, B);
    pl[r * PT + t] = u[0];
#pragma unroll
    for (int k = 1; k < A; ++k) pl[r * PT + t + B * k] = twmul<DIR>(u[k], tw[t * k]);
  }
  __syncthreads();

  // x phase B: item (r, k), row fastest; read B values, hold, barrier, write natural order k + A q
  constexpr int IB = (N * A + NT - 1) / NT;
  double2 ub[IB][B];
#pragma unroll
  for (int i = 0; i < IB; ++i) {
    const int it = threadIdx.x + i * NT;
    if (it < N * A) {
      const int k = it / N, r = it - (it / N) * N;
#pragma unroll
      for (int t = 0; t < B; ++t) ub[i][t] = pl[r * PT + t + B * k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < IB; ++i) {
    const int it = threadIdx.x + i * NT;
    if (it < N * A) {
      const int k = it / N, r = it - (it / N) * N;
      rdft<B, DIR>(ub[i], tw, A);
#pragma unroll
      for (int q = 0; q < B; ++q) pl[r * PT + k + A * q] = ub[i][q];
    }
  }
  __syncthreads();

  // y phase A: item (c, t), c fastest; rows t + B j of column c, in place
  for (int it = threadIdx.x; it < N * B; it += NT) {
    const int t = it / N, c = it - (it / N) * N;
    double2 u[A];
#pragma unroll
    for (int j = 0; j < A; ++j) u[j] = pl[(t + B * j) * PT + c];
    rdft<A, DIR>(u, tw, B);
    pl[t * PT + c] = u[0];
#pragma unroll
    for (int k = 1; k < A; ++k) pl[(t + B * k) * PT + c] = twmul<DIR>(u[k], tw[t * k]);
  }
  __syncthreads();

  // y phase B: item (c, k), c fastest; rows B k + t -> output rows k + A q in HBM
  for (int it = threadIdx.x; it < N * A; it += NT) {
    const int k = it / N, c = it - (it / N) * N;
    double2 u[B];
#pragma unroll
    for (int t = 0; t < B; ++t) u[t] = pl[(B * k + t) * PT + c];
    rdft<B, DIR>(u, tw, A);
    double2 sxz = make_double2(1, 0);
    if (TWIST == 1) sxz = cmul(S[c], sz);
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int row = k + A * q;
      double2 v = cscale(u[q], scale);
      if (TWIST == 1) v = cmul(v, cmul(S[N + row], sxz));
      dst[row * N + c] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// z-pencil pass (n3 = N3 = A*B), TP consecutive pencils per work-group.
// LDS: [3][N3][TP] (pencil index fastest) + twiddles (N3).
// ---------------------------------------------------------------------------
// CG = 1 fuses the conjugate-gradient iteration into the first and last pass of fncB:
//   MIX_IN  (iteration i >= 1) sums the previous MIX_OUT pass's partials of ||r||^2 into
//           rho_i, forms beta = rho_i / rho_{i-1} and tests convergence; every work-group
//           does this in the same order, so all reach the same values, and work-group 0
//           stores rho_i (slot i & 1), counts the iteration and clears the item's active flag
//           when it has converged (the item then returns at once).  Then it reads r (the
//           input), p and x at each point, forms p <- r + beta p and the deferred
//           x <- x + alpha p_old, writes p and x back and transforms p;
//   MIX_OUT sums the material pass's partials of p^H q, forms alpha = rho_i / p^H q (work-group
//           0 stores it for the next MIX_IN and the final x update), holds r in its output
//           array and updates it in place from each output q, r <- r - alpha q (q is not
//           stored), leaving one partial sum of ||r||^2 (in .x) per (item, slot, work-group)
//           in cf.part_r[item * cf.nparts + slot * gridDim.x + blockIdx.x].
// No separate scalar kernel runs between the passes.
// CG = 0 uses only cf.st, and only to skip items whose state is inactive (nullptr: none).
struct CgFuse {
  CgState* st;
  double2* x;
  double2* p;
  double2* part;    // material partials of p^H q, item stride nparts
  int nparts;
  double2* part_r;  // MIX_OUT partials of ||r||^2, item stride nparts
  int nparts_m;     // material partials per item in this launch
  int par;          // iteration parity: the rho slot of this iteration
  int first;        // iteration 0: beta = 0, no residual partials yet
};

template <int N3, int A, int B, int DIR, int MODE, int TP, int NT, int CG = 0>
__global__ void __launch_bounds__(NT) k_zfast(Geo g, const double2* __restrict__ in, long long in_bstride,
                                              double2* __restrict__ out, long long out_bstride,
                                              const double2* __restrict__ lam, int lam_bstride,
                                              const double2* __restrict__ twg, double scale, int grp_base,
                                              CgFuse cf) {
  static_assert(A * B == N3, "k_zfast: N3 = A*B");
  extern __shared__ double2 sm[];
  double2* tw = sm + 3 * N3 * TP;
  const int n12 = g.n12;
  const long long N = g.N;
  const int q0 = blockIdx.x * TP;
  const int slot = blockIdx.y;
  const int grp = grp_base + slot;
  const int b = blockIdx.z;
  // a converged CG item is left as it is: its last deferred x update is applied by k_cg_xfin;
  // outside the CG, cf.st (when given) marks items a caller's iteration has finished.
  // The flag is read once per work-group: in the CG's MIX_IN pass work-group 0 clears it when
  // the item converges while other work-groups of the same launch are starting, and waves that
  // read it on either side of that store would split their work-group, leaving the waves that
  // go on to reduce the partial sums below with slots of the exited waves unwritten.
  if (cf.st) {
    __shared__ int skip_wg;
    if (threadIdx.x == 0) skip_wg = cf.st[b].active == 0.0;
    __syncthreads();
    if (skip_wg) return;
  }
  const double2* __restrict__ lb = lam + (long long)b * lam_bstride;
  const double2* __restrict__ twn = twg + g.ax[2].tw_off;
  for (int t = threadIdx.x; t < N3; t += NT) tw[t] = twn[t];
  auto L = [&](int c, int z, int pp) -> double2& { return sm[(c * N3 + z) * TP + pp]; };

  if (MODE == ZM_MIX_IN) {
    // phase 0: load + Pi mix in natural order
    const double2* __restrict__ xin = in + b * in_bstride + (long long)(2 * slot) * N;
    double2* __restrict__ pf = nullptr;
    double2* __restrict__ xf = nullptr;
    double cal = 0, cbe = 0;
    if (CG) {
      pf = cf.p + b * in_bstride + (long long)(2 * slot) * N;
      xf = cf.x + b * in_bstride + (long long)(2 * slot) * N;
      CgState* __restrict__ st = cf.st + b;
      cal = st->alpha;
      if (!cf.first) {
        const double rprev = cf.par ? st->rho : st->rho1;  // rho_{i-1}, not written here
        const double bn2 = st->bnorm2, t2 = st->tol2, its = st->iters;
        const double rn = part_sum_all(cf.part_r + (long long)b * cf.nparts, gridDim.x * gridDim.y).x;
        cbe = rn / rprev;
        const bool conv = rn <= t2 * bn2;
        if (blockIdx.x == 0 && slot == 0 && threadIdx.x == 0) {
          if (cf.par) st->rho1 = rn;
          else st->rho = rn;
          st->iters = its + 1;
          if (conv) st->active = 0.0;
        }
        if (conv) return;  // uniform: every work-group reached the same rn
      }
    }
    for (int it = threadIdx.x; it < N3 * TP; it += NT) {
      const int k3 = it / TP, pp = it - (it / TP) * TP;
      const int q = q0 + pp;
      double2 v0 = make_double2(0, 0), v1 = v0, v2 = v0;
      if (q < n12) {
        const long long P = q + (long long)n12 * k3;
        double2 u1 = xin[P], u2 = xin[N + P];
        if (CG) {
          const double2 o1 = pf[P], o2 = pf[N + P];
          const double2 x1 = xf[P], x2 = xf[N + P];
          u1 = make_double2(fma(cbe, o1.x, u1.x), fma(cbe, o1.y, u1.y));
          u2 = make_double2(fma(cbe, o2.x, u2.x), fma(cbe, o2.y, u2.y));
          xf[P] = make_double2(fma(cal, o1.x, x1.x), fma(cal, o1.y, x1.y));
          xf[N + P] = make_double2(fma(cal, o2.x, x2.x), fma(cal, o2.y, x2.y));
          pf[P] = u1;
          pf[N + P] = u2;
        }
        const int k2 = q / g.n1, k1 = q - k2 * g.n1;
        double2 p1[3], p2[3];
        pi_vectors(lb[k1], lb[g.n1 + k2], lb[g.n1 + g.n2 + k3], p1, p2);
        if (grp == 0) {
          v0 = cfmac(cscale(cmulc(u1, p2[0]), -1.0), p1[0], u2);
          v1 = cfmac(cscale(cmulc(u1, p2[1]), -1.0), p1[1], u2);
          v2 = cfmac(cscale(cmulc(u1, p2[2]), -1.0), p1[2], u2);
        } else {
          v0 = cfma(cmul(p1[0], u1), p2[0], u2);
          v1 = cfma(cmul(p1[1], u1), p2[1], u2);
          v2 = cfma(cmul(p1[2], u1), p2[2], u2);
        }
      }
      L(0, k3, pp) = v0;
      L(1, k3, pp) = v1;
      L(2, k3, pp) = v2;
    }
    __syncthreads();
    // phase A in place: item (c, t, pp)
    for (int it = threadIdx.x; it < 3 * B * TP; it += NT) {
      const int pp = it % TP, r = it / TP;
      const int t = r % B, c = r / B;
      double2 u[A];
#pragma unroll
      for (int j = 0; j < A; ++j) u[j] = L(c, t + B * j, pp);
      rdft<A, DIR>(u, tw, B);
      L(c, t, pp) = u[0];
#pragma unroll
      for (int k = 1; k < A; ++k) L(c, t + B * k, pp) = twmul<DIR>(u[k], tw[t * k]);
    }
  } else {
    // phase A from HBM: item (c, t, pp)
    const double2* __restrict__ fin = in + b * in_bstride + (MODE == ZM_MIX_OUT ? (long long)(3 * slot) * N : 0);
    constexpr int IA = (3 * B * TP + NT - 1) / NT;
    double2 ua[IA][A];
#pragma unroll
    for (int i = 0; i < IA; ++i) {  // all loads first, before the twiddle barrier
      const int it = threadIdx.x + i * NT;
      const int pp = it % TP, r = it / TP;
      const int t = r % B, c = r / B;
      const int q = q0 + pp;
#pragma unroll
      for (int j = 0; j < A; ++j)
        ua[i][j] = (it < 3 * B * TP && q < n12) ? fin[c * N + q + (long long)n12 * (t + B * j)] : make_double2(0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IA; ++i) {
      const int it = threadIdx.x + i * NT;
      if (it >= 3 * B * TP) break;
      const int pp = it % TP, r = it / TP;
      const int t = r % B, c = r / B;
      double2* u = ua[i];
      rdft<A, DIR>(u, tw, B);
      L(c, t, pp) = u[0];
#pragma unroll
      for (int k = 1; k < A; ++k) L(c, t + B * k, pp) = twmul<DIR>(u[k], tw[t * k]);
    }
  }
  __syncthreads();

  if (MODE == ZM_MIX_OUT) {
    // phase B into LDS (hold, barrier, write natural order), then phase C: Pi^H mix per point
    constexpr int IB = (3 * A * TP + NT - 1) / NT;
    double2 ub[IB][B];
#pragma unroll
    for (int i = 0; i < IB; ++i) {
      const int it = threadIdx.x + i * NT;
      if (it < 3 * A * TP) {
        const int pp = it % TP, r = it / TP;
        const int k = r % A, c = r / A;
#pragma unroll
        for (int t = 0; t < B; ++t) ub[i][t] = L(c, B * k + t, pp);
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IB; ++i) {
      const int it = threadIdx.x + i * NT;
      if (it < 3 * A * TP) {
        const int pp = it % TP, r = it / TP;
        const int k = r % A, c = r / A;
        rdft<B, DIR>(ub[i], tw, A);
#pragma unroll
        for (int q = 0; q < B; ++q) L(c, k + A * q, pp) = ub[i][q];
      }
    }
    __syncthreads();
    double2* __restrict__ yout = out + b * out_bstride + (long long)(2 * slot) * N;
    double ca = 0.0;
    if (CG) {  // alpha = rho_i / p^H q from the material pass's partials (same in every work-group)
      const double pq = part_sum_all(cf.part + (long long)b * cf.nparts, cf.nparts_m).x;
      ca = (cf.par ? cf.st[b].rho1 : cf.st[b].rho) / pq;
      if (blockIdx.x == 0 && slot == 0 && threadIdx.x == 0) cf.st[b].alpha = ca;
    }
    double2 dacc = make_double2(0, 0);
    for (int it = threadIdx.x; it < N3 * TP; it += NT) {
      const int k3 = it / TP, pp = it - (it / TP) * TP;
      const int q = q0 + pp;
      if (q >= n12) continue;
      const double2 y0 = cscale(L(0, k3, pp), scale);
      const double2 y1 = cscale(L(1, k3, pp), scale);
      const double2 y2 = cscale(L(2, k3, pp), scale);
      const int k2 = q / g.n1, k1 = q - k2 * g.n1;
      double2 p1[3], p2[3];
      pi_vectors(lb[k1], lb[g.n1 + k2], lb[g.n1 + g.n2 + k3], p1, p2);
      double2 o1, o2;
      if (grp == 0) {
        o1 = cscale(cfma(cfma(cmul(p2[0], y0), p2[1], y1), p2[2], y2), -1.0);
        o2 = cfma(cfma(cmul(p1[0], y0), p1[1], y1), p1[2], y2);
      } else {
        o1 = cfmac(cfmac(cmulc(y0, p1[0]), p1[1], y1), p1[2], y2);
        o2 = cfmac(cfmac(cmulc(y0, p2[0]), p2[1], y1), p2[2], y2);
      }
      const long long P = q + (long long)n12 * k3;
      if (CG) {
        double2 r1 = yout[P], r2 = yout[N + P];
        r1 = make_double2(fma(-ca, o1.x, r1.x), fma(-ca, o1.y, r1.y));
        r2 = make_double2(fma(-ca, o2.x, r2.x), fma(-ca, o2.y, r2.y));
        yout[P] = r1;
        yout[N + P] = r2;
        dacc.x = fma(r1.x, r1.x, fma(r1.y, r1.y, dacc.x));
        dacc.x = fma(r2.x, r2.x, fma(r2.y, r2.y, dacc.x));
      } else {
        yout[P] = o1;
        yout[N + P] = o2;
      }
    }
    if (CG) {
      const double2 s = block_sum(dacc);
      if (threadIdx.x == 0) cf.part_r[(long long)b * cf.nparts + slot * gridDim.x + blockIdx.x] = s;
    }
  } else {
    // phase B straight to HBM: item (c, k, pp); output z = k + A q
    double2* __restrict__ fo = out + b * out_bstride + (MODE == ZM_MIX_IN ? (long long)(3 * slot) * N : 0);
    for (int it = threadIdx.x; it < 3 * A * TP; it += NT) {
      const int pp = it % TP, r = it / TP;
      const int k = r % A, c = r / A;
      const int q = q0 + pp;
      double2 u[B];
#pragma unroll
      for (int t = 0; t < B; ++t) u[t] = L(c, B * k + t, pp);
      rdft<B, DIR>(u, tw, A);
      if (q < n12) {
#pragma unroll
        for (int qq = 0; qq < B; ++qq) fo[c * N + q + (long long)n12 * (k + A * qq)] = cscale(u[qq], scale);
      }
    }
  }
}

}  // namespace fame
