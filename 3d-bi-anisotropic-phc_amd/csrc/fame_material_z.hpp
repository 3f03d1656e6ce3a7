// Material stencil, 2.5-D: each work-group owns a TX x TY tile of (x, y)
// columns and marches through a range of z-slices.  Thread (fx, fy) owns one
// column of the tile grown by one on every side, so
//   * z-neighbours (slice s-1 for the gathers, s+1 for the scatters) stay in
//     the thread's registers or in a two-slice vertex buffer,
//   * x/y neighbours are exchanged through small per-slice LDS arrays,
// and every field value is fetched from HBM once (plus the one-column halo).
// The arithmetic is the vertex form of k_material (same gathers, 6x6 vertex
// products and scatters).  With dpart, each work-group also leaves the sum over its
// outputs of conj(input) . output: summed over the grid this is u^H (Mat u), which for
// u = the transformed CG direction p equals p^H fncB p (the transforms are unitary).
#pragma once
#include <type_traits>

#include "fame_kernels.hpp"

namespace fame {

constexpr int kMzThreads = 256;

// Per-vertex bi-isotropic media: G = A^T A (row-major) and mu; the slots of vertex v are
// N_v = n G, M_v = m G, J_v = i j G with n = mu/d, m = eps/d, j = -gamma/d, d = eps mu - gamma^2.
struct PvCoef {
  double G[9];
  double mu;
};

// n, m, Im(J)/G of one vertex and their derivatives by eps and by gamma
struct PvSlots {
  double n, m, j;          // J = i * j * G
  double ne, me, je;       // d/d eps
  double ng, mg, jg;       // d/d gamma
};

__device__ __forceinline__ PvSlots pv_slots(double eps, double gam, double mu) {
  PvSlots s;
  const double d = eps * mu - gam * gam;
  const double id = 1.0 / d, id2 = id * id;
  s.n = mu * id;
  s.m = eps * id;
  s.j = -gam * id;
  s.ne = -mu * mu * id2;
  s.me = -gam * gam * id2;
  s.je = gam * mu * id2;
  s.ng = 2.0 * gam * mu * id2;
  s.mg = 2.0 * gam * eps * id2;
  s.jg = -(eps * mu + gam * gam) * id2;
  return s;
}

// LDS: (table media) the material table; the newest slice's exchange values (5 per
// column); vertex products read at neighbour positions: components 0,1 for two slices,
// components 2,3,4 for the current slice; material ids for two slices (table media) or
// the vertex scalars (n, m) for two slices (per-vertex media).
__host__ __device__ inline size_t material_z_lds(int TX, int TY, bool pv = false) {
  const int RF = (TX + 2) * (TY + 2);
  if (pv) return (size_t)(5 + 4 + 3 + 2) * RF * 16;
  return (size_t)kMaxMat * kMatWords * 16 + (size_t)(5 + 4 + 3) * RF * 16 + 2 * (size_t)RF;
}

// upper bound over all tiles with (TX+2)(TY+2) <= kMzThreads, either kind of medium
inline size_t material_z_lds_max() {
  return (size_t)kMaxMat * kMatWords * 16 + (size_t)(5 + 4 + 3) * kMzThreads * 16 + 2 * (size_t)kMzThreads * 16;
}

// PV = false: media from the table `mat` by the one-byte ids `mat_id`; PV = true: per-vertex
// media from veps, vgam and pc (the same arithmetic with N = n G, M = m G, J = i j G).
template <bool PV>
__global__ void __launch_bounds__(kMzThreads) k_material_z(Geo g, const double2* __restrict__ fin, long long in_bstride,
                                                           double2* __restrict__ fout, long long out_bstride,
                                                           const unsigned char* __restrict__ mat_id,
                                                           const double2* __restrict__ mat, int n_mat,
                                                           const double* __restrict__ veps,
                                                           const double* __restrict__ vgam, PvCoef pc,
                                                           const double2* __restrict__ omega, int TX, int TY, int TZ,
                                                           double2* __restrict__ dpart, int pstride,
                                                           const CgState* __restrict__ cgst) {
  extern __shared__ double2 sm[];
  const int RFX = TX + 2, RFY = TY + 2, RF = RFX * RFY;
  double2* tab = sm;
  double2* F = tab + (PV ? 0 : kMaxMat * kMatWords);  // [5][RF]: e_x, e_y, a_x, a_y, h_z of the newest slice
  double2* W01 = F + 5 * RF;                // [2 slices][2 comps][RF]: wh_x, wh_y
  double2* W234 = W01 + 4 * RF;             // [3][RF]: wh_z, we_x, we_y of the current slice
  unsigned char* IDS = (unsigned char*)(W234 + 3 * RF);  // [2][RF] (table media)
  double2* NMV = W234 + 3 * RF;                          // [2][RF] (n, m) (per-vertex media)
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  const long long N = g.N;
  const int ntx = (n1 + TX - 1) / TX;
  // XCD-aware order: work-groups are dealt round-robin to the 8 XCDs, so give each XCD a
  // contiguous range of (tile, slab, item) indices; neighbouring tiles then share that
  // XCD's L2 and their halo columns are L2 hits instead of second HBM reads.
  int tile, zc, b;
  {
    const unsigned G = gridDim.x * gridDim.y * gridDim.z;
    const unsigned L = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const unsigned xcd = L % 8u, slot = L / 8u, per = G / 8u, rem = G % 8u;
    const unsigned logical = xcd * per + min(xcd, rem) + slot;
    tile = (int)(logical % gridDim.x);
    zc = (int)((logical / gridDim.x) % gridDim.y);
    b = (int)(logical / (gridDim.x * gridDim.y));
  }
  if (cgst && cgst[b].active == 0.0) return;  // inside a fused CG: item b has converged
  const int x0 = (tile % ntx) * TX, y0 = (tile / ntx) * TY;
  const int z0 = zc * TZ;
  const int z1 = min(z0 + TZ, n3);
  const double2* __restrict__ H = fin + b * in_bstride;
  const double2* __restrict__ E = H + 3 * N;
  double2* __restrict__ HO = fout + b * out_bstride;
  double2* __restrict__ EO = HO + 3 * N;
  if (!PV)
    for (int t = threadIdx.x; t < n_mat * kMatWords; t += blockDim.x) tab[t] = mat[t];
  const double2 om0 = omega[3 * b + 0], om1 = omega[3 * b + 1], om2 = omega[3 * b + 2];

  const int t = threadIdx.x;
  const bool act = t < RF;
  const int fx = t % RFX, fy = t / RFX;
  const int gx = (x0 - 1 + fx + n1) % n1, gy = (y0 - 1 + fy + n2) % n2;
  const bool wact = act && fx >= 1 && fy >= 1;
  const bool oact = wact && fx <= TX && fy <= TY && (x0 + fx - 1 < n1) && (y0 + fy - 1 < n2);
  const long long col = gx + (long long)n1 * gy;
  const long long plane = (long long)n1 * n2;
  const double2 z2 = make_double2(0, 0);

  // the medium of a vertex: its id (table) or its (eps, gamma) (per vertex)
  using VM = typename std::conditional<PV, double2, int>::type;
  auto load = [&](int z, double2& a0, double2& a1, double2& a2, double2& c0, double2& c1, double2& c2, VM& id) {
    if (act) {
      const long long p = col + plane * z;
      a0 = E[p]; a1 = E[N + p]; a2 = E[2 * N + p];
      c0 = H[p]; c1 = H[N + p]; c2 = H[2 * N + p];
      if constexpr (PV) id = make_double2(veps[p], vgam[p]);
      else id = mat_id[p];
    }
  };
  // exchange values of a new slice (needs the previous slice's e_z, h_x, h_y for the z-partials)
  auto put_slice = [&](int buf, const double2& ce0, const double2& ce1, const double2& ch0, const double2& ch1,
                       const double2& ch2, const double2& pe0h, const double2& ph0, const double2& ph1, VM id) {
    (void)pe0h;
    if (act) {
      F[0 * RF + t] = ce0;
      F[1 * RF + t] = ce1;
      F[2 * RF + t] = cfmac(ch0, om2, ph0);
      F[3 * RF + t] = cfmac(ch1, om2, ph1);
      F[4 * RF + t] = ch2;
      if constexpr (PV) {
        const PvSlots sl = pv_slots(id.x, id.y, pc.mu);
        NMV[buf * RF + t] = make_double2(sl.n, sl.m);
      } else {
        IDS[buf * RF + t] = (unsigned char)id;
      }
    }
  };
  // vertex products of the newest slice; comps 0,1 go to LDS, 2..5 are returned
  auto products = [&](int buf, const double2& vez, VM id, double2& w2, double2& w3, double2& w4, double2& w5) {
    if (!wact) return;
    const double2 vex = cscale(cfmac(F[t], om0, F[t - 1]), 0.5);
    const double2 vey = cscale(cfmac(F[RF + t], om1, F[RF + t - RFX]), 0.5);
    const double2 vhx = cscale(cfmac(F[2 * RF + t], om1, F[2 * RF + t - RFX]), 0.25);
    const double2 vhy = cscale(cfmac(F[3 * RF + t], om0, F[3 * RF + t - 1]), 0.25);
    const double2 hz_a = cfmac(F[4 * RF + t], om0, F[4 * RF + t - 1]);
    const double2 hz_c = cfmac(F[4 * RF + t - RFX], om0, F[4 * RF + t - RFX - 1]);
    const double2 vhz = cscale(cfmac(hz_a, om1, hz_c), 0.25);
    // slots N (ep), M (mu), J (xi) of this vertex, entry k of the row-major 3x3
    double sn = 0, sm_ = 0, sj = 0;
    if constexpr (PV) {
      const PvSlots sl = pv_slots(id.x, id.y, pc.mu);
      sn = sl.n;
      sm_ = sl.m;
      sj = sl.j;
    }
    const double2* T = nullptr;
    if constexpr (!PV) T = tab + kMatWords * id;
    auto ep = [&](int k) { return PV ? make_double2(sn * pc.G[k], 0.0) : T[k]; };
    auto mu = [&](int k) { return PV ? make_double2(sm_ * pc.G[k], 0.0) : T[9 + k]; };
    auto xi = [&](int k) { return PV ? make_double2(0.0, sj * pc.G[k]) : T[18 + k]; };
    W01[(buf * 2 + 0) * RF + t] = cfmac(cfmac(cfmac(cfma(cmul(mu(1), vhy), mu(2), vhz), xi(0), vex), xi(3), vey), xi(6), vez);
    W01[(buf * 2 + 1) * RF + t] = cfmac(cfmac(cfmac(cfma(cmul(mu(3), vhx), mu(5), vhz), xi(1), vex), xi(4), vey), xi(7), vez);
    w2 = cfmac(cfmac(cfmac(cfma(cmul(mu(6), vhx), mu(7), vhy), xi(2), vex), xi(5), vey), xi(8), vez);
    w3 = cfma(cfma(cfma(cfma(cmul(ep(1), vey), ep(2), vez), xi(0), vhx), xi(1), vhy), xi(2), vhz);
    w4 = cfma(cfma(cfma(cfma(cmul(ep(3), vex), ep(5), vez), xi(3), vhx), xi(4), vhy), xi(5), vhz);
    w5 = cfma(cfma(cfma(cfma(cmul(ep(6), vex), ep(7), vey), xi(6), vhx), xi(7), vhy), xi(8), vhz);
  };

  __syncthreads();  // material table
  // prime: slice z0-1 (its e_z, h_x, h_y) and slice z0; products of z0
  double2 e0 = z2, e1 = z2, e2 = z2, h0 = z2, h1 = z2, h2 = z2;  // current slice fields
  double2 pe2 = z2, ph0 = z2, ph1 = z2;                           // previous slice partial inputs
  double2 hold2 = z2, hold3 = z2, hold4 = z2, cur5 = z2, nxt5 = z2;
  {
    VM id0{};
    if (act) {  // slice z0-1: only its e_z, h_x, h_y enter (the z-partials of slice z0)
      const long long p = col + plane * ((z0 - 1 + n3) % n3);
      pe2 = E[2 * N + p];
      ph0 = H[p];
      ph1 = H[N + p];
    }
    load(z0, e0, e1, e2, h0, h1, h2, id0);
    put_slice(0, e0, e1, h0, h1, h2, pe2, ph0, ph1, id0);
    const double2 vez = cscale(cfmac(e2, om2, pe2), 0.5);
    __syncthreads();
    products(0, vez, id0, hold2, hold3, hold4, cur5);
  }
  int buf = 0;
  double2 dacc = z2;  // sum of conj(input) . output over this thread's points (dpart only)
  double2 ne0 = z2, ne1 = z2, ne2 = z2, nh0 = z2, nh1 = z2, nh2 = z2;
  VM id_n{};
  load((z0 + 1) % n3, ne0, ne1, ne2, nh0, nh1, nh2, id_n);
  for (int s = z0; s < z1; ++s) {
    double2 fe0 = z2, fe1 = z2, fe2 = z2, fh0 = z2, fh1 = z2, fh2 = z2;
    VM id_f{};
    if (s + 1 < z1) load((s + 2) % n3, fe0, fe1, fe2, fh0, fh1, fh2, id_f);
    __syncthreads();  // readers of F, W234 and W01[buf ^ 1] from the previous step are done
    if (wact) {  // the held products of slice s become the current slice's neighbour values
      W234[0 * RF + t] = hold2;
      W234[1 * RF + t] = hold3;
      W234[2 * RF + t] = hold4;
    }
    put_slice(buf ^ 1, ne0, ne1, nh0, nh1, nh2, e2, h0, h1, id_n);
    const double2 vez = cscale(cfmac(ne2, om2, e2), 0.5);
    __syncthreads();
    products(buf ^ 1, vez, id_n, hold2, hold3, hold4, nxt5);
    __syncthreads();
    if (oact) {
      const double2* Wc = W01 + buf * 2 * RF;
      const double2* Wn = W01 + (buf ^ 1) * 2 * RF;
      // diagonal N_ii / M_ii of the vertex at LDS position q of slice buffer bi
      auto epd = [&](int bi, int q, int i) {
        if constexpr (PV) return NMV[bi * RF + q].x * pc.G[4 * i];
        else return tab[kMatWords * IDS[bi * RF + q] + 4 * i].x;
      };
      auto mud = [&](int bi, int q, int i) {
        if constexpr (PV) return NMV[bi * RF + q].y * pc.G[4 * i];
        else return tab[kMatWords * IDS[bi * RF + q] + 9 + 4 * i].x;
      };
      const int bc = buf, bn = buf ^ 1;  // current / next slice
      const long long p = col + plane * s;
      {  // edges
        const double d0 = 0.5 * (epd(bc, t, 0) + epd(bc, t + 1, 0));
        const double d1 = 0.5 * (epd(bc, t, 1) + epd(bc, t + RFX, 1));
        const double d2 = 0.5 * (epd(bc, t, 2) + epd(bn, t, 2));
        const double2 y0 = cscale(cfma(W234[RF + t], om0, W234[RF + t + 1]), 0.5);
        const double2 y1 = cscale(cfma(W234[2 * RF + t], om1, W234[2 * RF + t + RFX]), 0.5);
        const double2 y2 = cscale(cfma(cur5, om2, nxt5), 0.5);
        const double2 o0 = make_double2(y0.x + d0 * e0.x, y0.y + d0 * e0.y);
        const double2 o1 = make_double2(y1.x + d1 * e1.x, y1.y + d1 * e1.y);
        const double2 o2 = make_double2(y2.x + d2 * e2.x, y2.y + d2 * e2.y);
        EO[p] = o0;
        EO[N + p] = o1;
        EO[2 * N + p] = o2;
        if (dpart) dacc = cfmac(cfmac(cfmac(dacc, e0, o0), e1, o1), e2, o2);
      }
      {  // faces: h_x over (y, z), h_y over (x, z), h_z over (x, y)
        const double d0 = 0.25 * (mud(bc, t, 0) + mud(bc, t + RFX, 0) + mud(bn, t, 0) + mud(bn, t + RFX, 0));
        const double d1 = 0.25 * (mud(bc, t, 1) + mud(bc, t + 1, 1) + mud(bn, t, 1) + mud(bn, t + 1, 1));
        const double d2 = 0.25 * (mud(bc, t, 2) + mud(bc, t + 1, 2) + mud(bc, t + RFX, 2) + mud(bc, t + RFX + 1, 2));
        double2 a = cfma(Wc[t], om1, Wc[t + RFX]);
        double2 c = cfma(Wn[t], om1, Wn[t + RFX]);
        const double2 y0 = cscale(cfma(a, om2, c), 0.25);
        a = cfma(Wc[RF + t], om0, Wc[RF + t + 1]);
        c = cfma(Wn[RF + t], om0, Wn[RF + t + 1]);
        const double2 y1 = cscale(cfma(a, om2, c), 0.25);
        a = cfma(W234[t], om0, W234[t + 1]);
        c = cfma(W234[t + RFX], om0, W234[t + RFX + 1]);
        const double2 y2 = cscale(cfma(a, om1, c), 0.25);
        const double2 o0 = make_double2(y0.x + d0 * h0.x, y0.y + d0 * h0.y);
        const double2 o1 = make_double2(y1.x + d1 * h1.x, y1.y + d1 * h1.y);
        const double2 o2 = make_double2(y2.x + d2 * h2.x, y2.y + d2 * h2.y);
        HO[p] = o0;
        HO[N + p] = o1;
        HO[2 * N + p] = o2;
        if (dpart) dacc = cfmac(cfmac(cfmac(dacc, h0, o0), h1, o1), h2, o2);
      }
    }
    // advance: slice s+1 becomes current
    e0 = ne0; e1 = ne1; e2 = ne2;
    h0 = nh0; h1 = nh1; h2 = nh2;
    ne0 = fe0; ne1 = fe1; ne2 = fe2;
    nh0 = fh0; nh1 = fh1; nh2 = fh2;
    id_n = id_f;
    cur5 = nxt5;
    buf ^= 1;
  }
  if (dpart) {  // uniform over the grid: one partial sum per work-group, [item][slab][tile]
    dacc = block_sum(dacc);
    if (threadIdx.x == 0) dpart[(long long)b * pstride + zc * gridDim.x + tile] = dacc;
  }
}

}  // namespace fame
