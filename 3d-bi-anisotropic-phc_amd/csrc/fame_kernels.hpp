// Device kernels of the curl-SVD + bi-anisotropic material matvec path.
//
// Layout in HBM (complex128 interleaved, MATLAB column-major order so the
// C-ABI needs no permutation):  grid point p = i1 + n1*(i2 + n2*i3).
//   NFSEP vector  : [batch][4][N]  = [x_a1, x_a2, x_b1, x_b2]   (x_a, x_b each 2N)
//   field scratch : [batch][6][N]  = [h1, h2, h3, e1, e2, e3]
//   3-comp field  : [batch][3][N]
// Every kernel takes a batch index in blockIdx.z; batch items share the grid
// and the material and differ in the Bloch vector (theta).
//
// Fields inside fncB are kept in the periodic ("untwisted") representation
// u = conj(S) .* field: the Bloch twist S of Psi and conj(S) of Psis cancel,
// and every shift sigma_d of the material interpolation becomes a periodic
// shift times omega_d = exp(i theta_d / n_d).
#pragma once
#include "fame_fft.hpp"

namespace fame {

constexpr int kMaxMat = 16;      // materials held in LDS by the stencil kernel
constexpr int kMatWords = 27;    // complex words per material: eps(3x3), mu(3x3), xi(3x3)
constexpr int kZTile = 16;       // pencils per z-pass work-group
constexpr int kMatTX = 8, kMatTY = 8, kMatTZ = 4;  // stencil output tile

struct Geo {
  int n1, n2, n3;
  int n12;
  long long N;
  AxisPlan ax[3];
};

// Pi_1, Pi_2 of the null-space-free curl SVD at one Fourier point
// (SVD_Matrix_curl.m:45-57):  c = conj(lambda), w = (1,2,3)/sqrt(14),
// Pi_1 = c x w / sqrt(Lq - |w.lambda|^2),  Pi_2 = -(c x conj(Pi_1)) / sqrt(Lq).
__device__ __forceinline__ void pi_vectors(double2 lx, double2 ly, double2 lz, double2 p1[3], double2 p2[3]) {
  const double ws = 0.26726124191242438468;  // 1/sqrt(14)
  const double w1 = ws, w2 = 2.0 * ws, w3 = 3.0 * ws;
  const double lq = lx.x * lx.x + lx.y * lx.y + ly.x * ly.x + ly.y * ly.y + lz.x * lz.x + lz.y * lz.y;
  const double2 lw = make_double2(w1 * lx.x + w2 * ly.x + w3 * lz.x, w1 * lx.y + w2 * ly.y + w3 * lz.y);
  const double in0 = 1.0 / sqrt(lq);
  const double in1 = 1.0 / sqrt(lq - (lw.x * lw.x + lw.y * lw.y));
  const double2 cx = cconj(lx), cy = cconj(ly), cz = cconj(lz);
  p1[0] = cscale(make_double2(cy.x * w3 - cz.x * w2, cy.y * w3 - cz.y * w2), in1);
  p1[1] = cscale(make_double2(cz.x * w1 - cx.x * w3, cz.y * w1 - cx.y * w3), in1);
  p1[2] = cscale(make_double2(cx.x * w2 - cy.x * w1, cx.y * w2 - cy.y * w1), in1);
  const double2 q0 = cconj(p1[0]), q1 = cconj(p1[1]), q2 = cconj(p1[2]);
  p2[0] = cscale(csub(cmul(cy, q2), cmul(cz, q1)), -in0);
  p2[1] = cscale(csub(cmul(cz, q0), cmul(cx, q2)), -in0);
  p2[2] = cscale(csub(cmul(cx, q1), cmul(cy, q0)), -in0);
}

// z-pass modes
enum { ZM_PLAIN = 0, ZM_MIX_IN = 1, ZM_MIX_OUT = 2 };

// ---------------------------------------------------------------------------
// z-pencil pass.  A work-group owns kZTile consecutive (i1,i2) pencils and all
// of their n3 points, for three components, in LDS laid out [c][pencil][z].
//   ZM_MIX_IN  (DIR=+1): x (4 comps, Fourier) -> Pi-mix (Pr for group 0 from
//              x_a, Qr for group 1 from x_b) -> inverse z-DFT -> fields 3g..3g+2.
//   ZM_MIX_OUT (DIR=-1): fields 3g..3g+2 -> forward z-DFT -> Pr^H (group 0) or
//              Qr^H (group 1) -> y comps 2g, 2g+1.
//   ZM_PLAIN   : 3 components in -> z-DFT -> 3 components out.
// ---------------------------------------------------------------------------
template <int DIR, int MODE>
__global__ void __launch_bounds__(256) k_zpass(Geo g, const double2* __restrict__ in, long long in_bstride,
                                               double2* __restrict__ out, long long out_bstride,
                                               const double2* __restrict__ lam, int lam_bstride,
                                               const double2* __restrict__ tw, const int* __restrict__ perm,
                                               double scale, int grp_base) {
  extern __shared__ double2 sm[];
  const int n3 = g.n3, n12 = g.n12;
  const long long N = g.N;
  const int q0 = blockIdx.x * kZTile;
  const int slot = blockIdx.y;           // which half of x / which field triple
  const int grp = grp_base + slot;        // 0: Pr (h, from x_a), 1: Qr (e, from x_b)
  const int b = blockIdx.z;
  const double2* __restrict__ lb = lam + (long long)b * lam_bstride;
  const int* __restrict__ pz = perm + g.ax[2].perm_off;
  const int tot = kZTile * n3;

  if (MODE == ZM_MIX_IN) {
    const double2* __restrict__ xin = in + b * in_bstride + (long long)(2 * slot) * N;
    for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
      const int k3 = idx / kZTile, pp = idx - k3 * kZTile;
      const int q = q0 + pp;
      double2 v0 = make_double2(0, 0), v1 = v0, v2 = v0;
      if (q < n12) {
        const long long P = q + (long long)n12 * k3;
        const double2 u1 = xin[P], u2 = xin[N + P];
        const int k2 = q / g.n1, k1 = q - k2 * g.n1;
        double2 p1[3], p2[3];
        pi_vectors(lb[k1], lb[g.n1 + k2], lb[g.n1 + g.n2 + k3], p1, p2);
        if (grp == 0) {  // Pr = [-conj(Pi_2), conj(Pi_1)]
          v0 = cfmac(cscale(cmulc(u1, p2[0]), -1.0), p1[0], u2);
          v1 = cfmac(cscale(cmulc(u1, p2[1]), -1.0), p1[1], u2);
          v2 = cfmac(cscale(cmulc(u1, p2[2]), -1.0), p1[2], u2);
        } else {  // Qr = [Pi_1, Pi_2]
          v0 = cfma(cmul(p1[0], u1), p2[0], u2);
          v1 = cfma(cmul(p1[1], u1), p2[1], u2);
          v2 = cfma(cmul(p1[2], u1), p2[2], u2);
        }
      }
      sm[(0 * kZTile + pp) * n3 + k3] = v0;
      sm[(1 * kZTile + pp) * n3 + k3] = v1;
      sm[(2 * kZTile + pp) * n3 + k3] = v2;
    }
  } else {
    const double2* __restrict__ fin =
        in + b * in_bstride + (MODE == ZM_MIX_OUT ? (long long)(3 * slot) * N : 0);
    for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
      const int z = idx / kZTile, pp = idx - z * kZTile;
      const int q = q0 + pp;
      const long long P = q + (long long)n12 * z;
      for (int c = 0; c < 3; ++c)
        sm[(c * kZTile + pp) * n3 + z] = (q < n12) ? fin[c * N + P] : make_double2(0, 0);
    }
  }
  __syncthreads();
  lds_fft<DIR>(sm, 3 * kZTile, 1, n3, g.ax[2], tw);

  if (MODE == ZM_MIX_OUT) {
    double2* __restrict__ yout = out + b * out_bstride + (long long)(2 * slot) * N;
    for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
      const int k3 = idx / kZTile, pp = idx - k3 * kZTile;
      const int q = q0 + pp;
      if (q >= n12) continue;
      const int s = pz[k3];
      const double2 y0 = cscale(sm[(0 * kZTile + pp) * n3 + s], scale);
      const double2 y1 = cscale(sm[(1 * kZTile + pp) * n3 + s], scale);
      const double2 y2 = cscale(sm[(2 * kZTile + pp) * n3 + s], scale);
      const int k2 = q / g.n1, k1 = q - k2 * g.n1;
      double2 p1[3], p2[3];
      pi_vectors(lb[k1], lb[g.n1 + k2], lb[g.n1 + g.n2 + k3], p1, p2);
      double2 o1, o2;
      if (grp == 0) {  // Pr^H: o1 = -sum Pi_2 y, o2 = sum Pi_1 y
        o1 = cscale(cfma(cfma(cmul(p2[0], y0), p2[1], y1), p2[2], y2), -1.0);
        o2 = cfma(cfma(cmul(p1[0], y0), p1[1], y1), p1[2], y2);
      } else {  // Qr^H: o1 = sum conj(Pi_1) y, o2 = sum conj(Pi_2) y
        o1 = cfmac(cfmac(cmulc(y0, p1[0]), p1[1], y1), p1[2], y2);
        o2 = cfmac(cfmac(cmulc(y0, p2[0]), p2[1], y1), p2[2], y2);
      }
      const long long P = q + (long long)n12 * k3;
      yout[P] = o1;
      yout[N + P] = o2;
    }
  } else {
    double2* __restrict__ fo = out + b * out_bstride + (MODE == ZM_MIX_IN ? (long long)(3 * slot) * N : 0);
    for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
      const int z = idx / kZTile, pp = idx - z * kZTile;
      const int q = q0 + pp;
      if (q >= n12) continue;
      const int s = pz[z];
      const long long P = q + (long long)n12 * z;
      for (int c = 0; c < 3; ++c) fo[c * N + P] = cscale(sm[(c * kZTile + pp) * n3 + s], scale);
    }
  }
}

// ---------------------------------------------------------------------------
// xy-plane pass: one work-group transforms one z-plane of one component in
// LDS (x lines, then y lines).  TWIST = 1 multiplies the result by the Bloch
// twist S (Psi), TWIST = 2 multiplies the input by conj(S) (Psis).
// ---------------------------------------------------------------------------
template <int DIR, int TWIST>
__global__ void __launch_bounds__(256) k_xypass(Geo g, const double2* __restrict__ in, long long in_bstride,
                                                double2* __restrict__ out, long long out_bstride,
                                                const double2* __restrict__ tws, int tws_bstride,
                                                const double2* __restrict__ tw, const int* __restrict__ perm,
                                                double scale) {
  extern __shared__ double2 sm[];
  const int n1 = g.n1, n2 = g.n2, n12 = g.n12;
  const int i3 = blockIdx.x, f = blockIdx.y, b = blockIdx.z;
  const long long base = (long long)f * g.N + (long long)i3 * n12;
  const double2* __restrict__ src = in + b * in_bstride + base;
  double2* __restrict__ dst = out + b * out_bstride + base;
  const double2* __restrict__ S = tws + (long long)b * tws_bstride;
  double2 sz = make_double2(1, 0);
  if (TWIST) sz = S[n1 + n2 + i3];
  for (int idx = threadIdx.x; idx < n12; idx += blockDim.x) {
    double2 v = src[idx];
    if (TWIST == 2) {
      const int i2 = idx / n1, i1 = idx - i2 * n1;
      v = cmulc(v, cmul(cmul(S[i1], S[n1 + i2]), sz));
    }
    sm[idx] = v;
  }
  __syncthreads();
  lds_fft<DIR>(sm, n2, 1, n1, g.ax[0], tw);
  lds_fft<DIR>(sm, n1, n1, 1, g.ax[1], tw);
  const int* __restrict__ px = perm + g.ax[0].perm_off;
  const int* __restrict__ py = perm + g.ax[1].perm_off;
  for (int idx = threadIdx.x; idx < n12; idx += blockDim.x) {
    const int i2 = idx / n1, i1 = idx - i2 * n1;
    double2 v = cscale(sm[py[i2] * n1 + px[i1]], scale);
    if (TWIST == 1) v = cmul(v, cmul(cmul(S[i1], S[n1 + i2]), sz));
    dst[idx] = v;
  }
}

// ---------------------------------------------------------------------------
// Bi-anisotropic material apply  [y_h; y_e] = [M h + J^H e; J h + N e]
// (fun_Lanczos/fncB.m:4-5) on periodic fields, evaluated vertex-wise:
//   gather   ve_j = (e_j(p) + w_j^* e_j(p-e_j))/2,
//            vh_j = (I + w_a^* T_a^H)(I + w_b^* T_b^H) h_j / 4   ({a,b} = axes != j)
//   vertex   we_i = sum_{j!=i} eps_ij ve_j + sum_j xi_ij vh_j
//            wh_i = sum_{j!=i} mu_ij vh_j + sum_j conj(xi_ji) ve_j
//   scatter  ye_i = (we_i(p) + w_i we_i(p+e_i))/2 + (eps_ii(p)+eps_ii(p+e_i))/2 e_i(p)
//            yh_i = (I + w_a T_a)(I + w_b T_b) wh_i / 4 + (sum of mu_ii at the 4 face vertices)/4 h_i(p)
// with eps, mu the Hermitian-symmetrised vertex tensors and w_d = e^{i theta_d/n_d}.
// This equals the reference's interpolated sparse products
// (IntCovarForm{N,M,J}_vertex*.m + the symmetrisation in
// Interpolation_Material_Matrix_Mult_Point_vertex.m:22,24).
// One work-group: kMatTX x kMatTY x kMatTZ outputs; vertex products for the
// tile grown by one in +x,+y,+z are staged in LDS.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_material(Geo g, const double2* __restrict__ fin, long long in_bstride,
                                                  double2* __restrict__ fout, long long out_bstride,
                                                  const unsigned char* __restrict__ mat_id,
                                                  const double2* __restrict__ mat, int n_mat,
                                                  const double2* __restrict__ omega) {
  constexpr int WX = kMatTX + 1, WY = kMatTY + 1, WZ = kMatTZ + 1;
  constexpr int NW = WX * WY * WZ;
  __shared__ double2 wsm[6][NW];
  __shared__ double2 tab[kMaxMat * kMatWords];
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  const long long N = g.N;
  const int b = blockIdx.z;
  const int tiles_x = (n1 + kMatTX - 1) / kMatTX;
  const int x0 = (blockIdx.x % tiles_x) * kMatTX;
  const int y0 = (blockIdx.x / tiles_x) * kMatTY;
  const int z0 = blockIdx.y * kMatTZ;
  const double2* __restrict__ H = fin + b * in_bstride;
  const double2* __restrict__ E = H + 3 * N;
  for (int t = threadIdx.x; t < n_mat * kMatWords; t += blockDim.x) tab[t] = mat[t];
  const double2 om[3] = {omega[3 * b + 0], omega[3 * b + 1], omega[3 * b + 2]};
  __syncthreads();

  // phase 1: vertex products on the grown tile
  for (int v = threadIdx.x; v < NW; v += blockDim.x) {
    const int vz = v / (WX * WY), vr = v - vz * (WX * WY);
    const int vy = vr / WX, vx = vr - vy * WX;
    const int x = (x0 + vx) % n1, y = (y0 + vy) % n2, z = (z0 + vz) % n3;
    const int xm = (x == 0) ? n1 - 1 : x - 1;
    const int ym = (y == 0) ? n2 - 1 : y - 1;
    const int zm = (z == 0) ? n3 - 1 : z - 1;
    const long long r0 = (long long)n1 * (y + (long long)n2 * z);
    const long long rym = (long long)n1 * (ym + (long long)n2 * z);
    const long long rzm = (long long)n1 * (y + (long long)n2 * zm);
    const long long rymzm = (long long)n1 * (ym + (long long)n2 * zm);
    const long long p = x + r0;
    // edge gathers
    double2 ve0 = cscale(cfmac(E[p], om[0], E[xm + r0]), 0.5);
    double2 ve1 = cscale(cfmac(E[N + p], om[1], E[N + x + rym]), 0.5);
    double2 ve2 = cscale(cfmac(E[2 * N + p], om[2], E[2 * N + x + rzm]), 0.5);
    // face gathers: h_1 uses (y,z), h_2 uses (x,z), h_3 uses (x,y)
    double2 vh0, vh1, vh2;
    {
      const double2* h = H;
      double2 a = cfmac(h[p], om[1], h[x + rym]);
      double2 c = cfmac(h[x + rzm], om[1], h[x + rymzm]);
      vh0 = cscale(cfmac(a, om[2], c), 0.25);
    }
    {
      const double2* h = H + N;
      double2 a = cfmac(h[p], om[0], h[xm + r0]);
      double2 c = cfmac(h[x + rzm], om[0], h[xm + rzm]);
      vh1 = cscale(cfmac(a, om[2], c), 0.25);
    }
    {
      const double2* h = H + 2 * N;
      double2 a = cfmac(h[p], om[0], h[xm + r0]);
      double2 c = cfmac(h[x + rym], om[0], h[xm + rym]);
      vh2 = cscale(cfmac(a, om[1], c), 0.25);
    }
    const double2* T = tab + kMatWords * (int)mat_id[p];
    const double2* ep = T;        // eps (N slot)
    const double2* mu = T + 9;    // mu  (M slot)
    const double2* xi = T + 18;   // xi  (J slot)
    // we_i = sum_{j!=i} eps_ij ve_j + sum_j xi_ij vh_j
    double2 we0 = cfma(cfma(cfma(cfma(cmul(ep[1], ve1), ep[2], ve2), xi[0], vh0), xi[1], vh1), xi[2], vh2);
    double2 we1 = cfma(cfma(cfma(cfma(cmul(ep[3], ve0), ep[5], ve2), xi[3], vh0), xi[4], vh1), xi[5], vh2);
    double2 we2 = cfma(cfma(cfma(cfma(cmul(ep[6], ve0), ep[7], ve1), xi[6], vh0), xi[7], vh1), xi[8], vh2);
    // wh_i = sum_{j!=i} mu_ij vh_j + sum_j conj(xi_ji) ve_j
    double2 wh0 = cfmac(cfmac(cfmac(cfma(cmul(mu[1], vh1), mu[2], vh2), xi[0], ve0), xi[3], ve1), xi[6], ve2);
    double2 wh1 = cfmac(cfmac(cfmac(cfma(cmul(mu[3], vh0), mu[5], vh2), xi[1], ve0), xi[4], ve1), xi[7], ve2);
    double2 wh2 = cfmac(cfmac(cfmac(cfma(cmul(mu[6], vh0), mu[7], vh1), xi[2], ve0), xi[5], ve1), xi[8], ve2);
    wsm[0][v] = wh0;
    wsm[1][v] = wh1;
    wsm[2][v] = wh2;
    wsm[3][v] = we0;
    wsm[4][v] = we1;
    wsm[5][v] = we2;
  }
  __syncthreads();

  // phase 2: scatter to edges / faces, plus lumped diagonal blocks
  constexpr int NT = kMatTX * kMatTY * kMatTZ;
  double2* __restrict__ HO = fout + b * out_bstride;
  double2* __restrict__ EO = HO + 3 * N;
  for (int t = threadIdx.x; t < NT; t += blockDim.x) {
    const int tz = t / (kMatTX * kMatTY), tr = t - tz * (kMatTX * kMatTY);
    const int ty = tr / kMatTX, tx = tr - ty * kMatTX;
    const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
    if (x >= n1 || y >= n2 || z >= n3) continue;
    const int xp = (x + 1 == n1) ? 0 : x + 1;
    const int yp = (y + 1 == n2) ? 0 : y + 1;
    const int zp = (z + 1 == n3) ? 0 : z + 1;
    const long long p = x + (long long)n1 * (y + (long long)n2 * z);
    const int v = tx + WX * (ty + WY * tz);
    const int vx = v + 1, vy = v + WX, vz = v + WX * WY;
    const int id000 = mat_id[p];
    const int id100 = mat_id[xp + (long long)n1 * (y + (long long)n2 * z)];
    const int id010 = mat_id[x + (long long)n1 * (yp + (long long)n2 * z)];
    const int id001 = mat_id[x + (long long)n1 * (y + (long long)n2 * zp)];
    const int id110 = mat_id[xp + (long long)n1 * (yp + (long long)n2 * z)];
    const int id101 = mat_id[xp + (long long)n1 * (y + (long long)n2 * zp)];
    const int id011 = mat_id[x + (long long)n1 * (yp + (long long)n2 * zp)];
    auto epsd = [&](int id, int i) { return tab[kMatWords * id + 4 * i].x; };
    auto mud = [&](int id, int i) { return tab[kMatWords * id + 9 + 4 * i].x; };
    // e outputs (edges)
    {
      const double d0 = 0.5 * (epsd(id000, 0) + epsd(id100, 0));
      const double d1 = 0.5 * (epsd(id000, 1) + epsd(id010, 1));
      const double d2 = 0.5 * (epsd(id000, 2) + epsd(id001, 2));
      double2 y0v = cscale(cfma(wsm[3][v], om[0], wsm[3][vx]), 0.5);
      double2 y1v = cscale(cfma(wsm[4][v], om[1], wsm[4][vy]), 0.5);
      double2 y2v = cscale(cfma(wsm[5][v], om[2], wsm[5][vz]), 0.5);
      const double2 e0 = E[p], e1 = E[N + p], e2 = E[2 * N + p];
      EO[p] = make_double2(y0v.x + d0 * e0.x, y0v.y + d0 * e0.y);
      EO[N + p] = make_double2(y1v.x + d1 * e1.x, y1v.y + d1 * e1.y);
      EO[2 * N + p] = make_double2(y2v.x + d2 * e2.x, y2v.y + d2 * e2.y);
    }
    // h outputs (faces): h_1 over (y,z), h_2 over (x,z), h_3 over (x,y)
    {
      const double d0 = 0.25 * (mud(id000, 0) + mud(id010, 0) + mud(id001, 0) + mud(id011, 0));
      const double d1 = 0.25 * (mud(id000, 1) + mud(id100, 1) + mud(id001, 1) + mud(id101, 1));
      const double d2 = 0.25 * (mud(id000, 2) + mud(id100, 2) + mud(id010, 2) + mud(id110, 2));
      const double2* w0 = wsm[0];
      const double2* w1 = wsm[1];
      const double2* w2 = wsm[2];
      double2 a, c;
      a = cfma(w0[v], om[1], w0[vy]);
      c = cfma(w0[vz], om[1], w0[vy + WX * WY]);
      double2 y0v = cscale(cfma(a, om[2], c), 0.25);
      a = cfma(w1[v], om[0], w1[vx]);
      c = cfma(w1[vz], om[0], w1[vz + 1]);
      double2 y1v = cscale(cfma(a, om[2], c), 0.25);
      a = cfma(w2[v], om[0], w2[vx]);
      c = cfma(w2[vy], om[0], w2[vy + 1]);
      double2 y2v = cscale(cfma(a, om[1], c), 0.25);
      const double2 h0 = H[p], h1 = H[N + p], h2 = H[2 * N + p];
      HO[p] = make_double2(y0v.x + d0 * h0.x, y0v.y + d0 * h0.y);
      HO[N + p] = make_double2(y1v.x + d1 * h1.x, y1v.y + d1 * h1.y);
      HO[2 * N + p] = make_double2(y2v.x + d2 * h2.x, y2v.y + d2 * h2.y);
    }
  }
}

// fncA (fun_Lanczos/fncA.m:3-4): y = [ i inv_r .* x_b ; -i inv_r .* x_a ],
// inv_r = 1/Sigma_r, Sigma_r = [sqrt(Lq); sqrt(Lq)] recomputed from lambda.
__global__ void k_fncA(Geo g, const double2* __restrict__ x, double2* __restrict__ y, const double2* __restrict__ lam,
                       int lam_bstride, const int* __restrict__ act) {
  const long long N = g.N;
  const int b = blockIdx.y;
  if (act && act[b] < 0) return;  // item b skipped (act == nullptr: every item)
  const double2* __restrict__ lb = lam + (long long)b * lam_bstride;
  const double2* __restrict__ xb = x + (long long)b * 4 * N;
  double2* __restrict__ yb = y + (long long)b * 4 * N;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < N; p += (long long)gridDim.x * blockDim.x) {
    const int i1 = (int)(p % g.n1);
    const long long r = p / g.n1;
    const int i2 = (int)(r % g.n2), i3 = (int)(r / g.n2);
    const double2 lx = lb[i1], ly = lb[g.n1 + i2], lz = lb[g.n1 + g.n2 + i3];
    const double lq = lx.x * lx.x + lx.y * lx.y + ly.x * ly.x + ly.y * ly.y + lz.x * lz.x + lz.y * lz.y;
    const double ir = 1.0 / sqrt(lq);
    for (int c = 0; c < 2; ++c) {
      const double2 xa = xb[c * N + p], xbv = xb[(2 + c) * N + p];
      yb[c * N + p] = make_double2(-ir * xbv.y, ir * xbv.x);            //  i inv_r x_b
      yb[(2 + c) * N + p] = make_double2(ir * xa.y, -ir * xa.x);        // -i inv_r x_a
    }
  }
}

// ---------------------------------------------------------------------------
// Batched reductions (deterministic two-pass) and vector updates.
// Wavefronts are 64 lanes: reduce with DPP-backed shuffles, then across the
// block through LDS.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 wave_sum(double2 v) {
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_down(v.x, o, 64);
    v.y += __shfl_down(v.y, o, 64);
  }
  return v;
}

__device__ __forceinline__ double2 block_sum(double2 v) {
  __shared__ double2 red[16];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
  v = (threadIdx.x < nw) ? red[threadIdx.x] : make_double2(0, 0);
  if (wid == 0) v = wave_sum(v);
  __syncthreads();
  return v;  // valid in thread 0
}

// block_sum with the result in every thread (uniform control flow)
__device__ __forceinline__ double2 block_sum_all(double2 v) {
  __shared__ double2 bc;
  v = block_sum(v);
  if (threadIdx.x == 0) bc = v;
  __syncthreads();
  return bc;
}

// sum of n partial sums, in every thread: thread t adds entries t, t + blockDim.x, ... in order,
// then block_sum (with 256 threads the order of k_cg_beta, so the same bits)
__device__ __forceinline__ double2 part_sum_all(const double2* __restrict__ part, int n) {
  double2 acc = make_double2(0, 0);
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc = cadd(acc, part[i]);
  return block_sum_all(acc);
}

// part[b*gridDim.x + blk] = sum over this block's range of conj(x) .* y
__global__ void __launch_bounds__(256) k_dot_part(const double2* __restrict__ x, long long xb,
                                                  const double2* __restrict__ y, long long yb, long long len,
                                                  double2* __restrict__ part, const int* __restrict__ act) {
  const int b = blockIdx.y;
  if (act && act[b] < 0) return;  // uniform over the work-group
  const double2* xp = x + b * xb;
  const double2* yp = y + b * yb;
  double2 acc = make_double2(0, 0);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x) {
    const double2 a = xp[i], c = yp[i];
    acc.x = fma(a.x, c.x, fma(a.y, c.y, acc.x));
    acc.y = fma(a.x, c.y, fma(-a.y, c.x, acc.y));
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = acc;
}

// out[b] = sum of part[b][0:nblk]  (one block per batch item); items with act[b] < 0 get 0
__global__ void __launch_bounds__(256) k_reduce_part(const double2* __restrict__ part, int nblk,
                                                     double2* __restrict__ out, const int* __restrict__ act) {
  const int b = blockIdx.x;
  if (act && act[b] < 0) {
    if (threadIdx.x == 0) out[b] = make_double2(0, 0);
    return;
  }
  double2 acc = make_double2(0, 0);
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) acc = cadd(acc, part[b * nblk + i]);
  acc = block_sum(acc);
  if (threadIdx.x == 0) out[b] = acc;
}

// ---------------------------------------------------------------------------
// Conjugate gradients on B (fncinvB), all batch items in lock step.
// Scalars per batch item live on the device:
//   cg[b] = {rho, bnorm2, alpha, beta, active, iters, tol2, rho1}
// The fused CG keeps rho of even iterations in rho and of odd ones in rho1, so the pass that
// forms the new rho never overwrites the value its other work-groups are still reading.
// ---------------------------------------------------------------------------
struct CgState {
  double rho, bnorm2, alpha, beta;
  double active, iters, tol2, rho1;
};

__global__ void k_cg_init(const double2* __restrict__ rr, CgState* __restrict__ st, int nb, double tol) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  CgState s;
  s.rho = rr[b].x;
  s.bnorm2 = rr[b].x;
  s.alpha = 0;
  s.beta = 0;
  s.active = (s.bnorm2 > 0) ? 1.0 : 0.0;
  s.iters = 0;
  s.tol2 = tol * tol;
  s.rho1 = 0;
  st[b] = s;
}

// alpha = rho / Re(p^H q) for active items
__global__ void k_cg_alpha(const double2* __restrict__ pq, CgState* __restrict__ st, int nb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  CgState s = st[b];
  s.alpha = (s.active != 0.0) ? s.rho / pq[b].x : 0.0;
  st[b] = s;
}

// x += alpha p ; r -= alpha q ; partial ||r||^2
__global__ void __launch_bounds__(256) k_cg_update(double2* __restrict__ x, double2* __restrict__ r,
                                                   const double2* __restrict__ p, const double2* __restrict__ q,
                                                   long long len, const CgState* __restrict__ st,
                                                   double2* __restrict__ part) {
  const int b = blockIdx.y;
  const double a = st[b].alpha;
  const long long off = b * len;
  double acc = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x) {
    const double2 pv = p[off + i], qv = q[off + i];
    double2 xv = x[off + i], rv = r[off + i];
    xv.x = fma(a, pv.x, xv.x);
    xv.y = fma(a, pv.y, xv.y);
    rv.x = fma(-a, qv.x, rv.x);
    rv.y = fma(-a, qv.y, rv.y);
    x[off + i] = xv;
    r[off + i] = rv;
    acc = fma(rv.x, rv.x, fma(rv.y, rv.y, acc));
  }
  double2 s = block_sum(make_double2(acc, 0));
  if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = s;
}

// reduce ||r||^2, then beta, rho, convergence flag (one block per batch item)
// nblk partial sums of ||r||^2 per item, item stride `stride`
__global__ void __launch_bounds__(256) k_cg_beta(const double2* __restrict__ part, int nblk, int stride,
                                                 CgState* __restrict__ st) {
  const int b = blockIdx.x;
  double2 acc = make_double2(0, 0);
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) acc = cadd(acc, part[(long long)b * stride + i]);
  acc = block_sum(acc);
  if (threadIdx.x == 0) {
    CgState s = st[b];
    if (s.active != 0.0) {
      const double rn = acc.x;
      s.beta = rn / s.rho;
      s.rho = rn;
      s.iters += 1;
      if (rn <= s.tol2 * s.bnorm2) {
        s.active = 0.0;
        s.beta = 0.0;
      }
    } else {
      s.beta = 0.0;
    }
    st[b] = s;
  }
}

// p = r + beta p  (active items only)
__global__ void k_cg_p(double2* __restrict__ p, const double2* __restrict__ r, long long len,
                       const CgState* __restrict__ st) {
  const int b = blockIdx.y;
  if (st[b].active == 0.0) return;
  const double be = st[b].beta;
  const long long off = b * len;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x) {
    const double2 rv = r[off + i], pv = p[off + i];
    p[off + i] = make_double2(fma(be, pv.x, rv.x), fma(be, pv.y, rv.y));
  }
}

// End of a fused CG whose scalar steps ran inside the z passes: the MIX_IN pass of iteration
// i >= 1 forms rho_i from the previous MIX_OUT partials, tests convergence and counts the
// iteration; this does the same for the residual the last queued iteration left (items still
// active), and moves every item's latest rho (slot iters & 1) into rho for the host.
__global__ void __launch_bounds__(256) k_cg_fin2(const double2* __restrict__ part, int nblk, int stride,
                                                 CgState* __restrict__ st) {
  const int b = blockIdx.x;
  const double2 acc = part_sum_all(part + (long long)b * stride, nblk);
  if (threadIdx.x == 0) {
    CgState s = st[b];
    if (s.active != 0.0) {
      s.rho = acc.x;
      s.iters += 1;
      if (acc.x <= s.tol2 * s.bnorm2) s.active = 0.0;
    } else if (((long long)s.iters) & 1) {
      s.rho = s.rho1;
    }
    s.beta = 0.0;
    st[b] = s;
  }
}

// the deferred last update x += alpha p
__global__ void k_cg_xfin(double2* __restrict__ x, const double2* __restrict__ p, long long len,
                          const CgState* __restrict__ st) {
  const int b = blockIdx.y;
  const double a = st[b].alpha;
  if (a == 0.0) return;
  const long long off = b * len;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x) {
    const double2 pv = p[off + i];
    double2 xv = x[off + i];
    x[off + i] = make_double2(fma(a, pv.x, xv.x), fma(a, pv.y, xv.y));
  }
}

// out(:, 0:k) = Q(:, 0:m) * U(0:m, 0:k) for real U (column-major, leading dim ldu), per batch item.
// Q and out are [b][col][len]; out must not alias Q.
__global__ void k_tsgemm(const double2* __restrict__ Q, long long qbs, long long len, const double* __restrict__ U,
                         int ldu, int m, int k, double2* __restrict__ out, long long obs) {
  const int b = blockIdx.y;
  const double2* Qb = Q + b * qbs;
  double2* Ob = out + b * obs;
  const double* Ub = U + (long long)b * ldu * k;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x) {
    for (int j0 = 0; j0 < k; j0 += 8) {
      double2 acc[8];
      for (int jj = 0; jj < 8; ++jj) acc[jj] = make_double2(0, 0);
      for (int l = 0; l < m; ++l) {
        const double2 q = Qb[l * len + i];
        for (int jj = 0; jj < 8; ++jj) {
          if (j0 + jj < k) {
            const double u = Ub[(j0 + jj) * ldu + l];
            acc[jj].x = fma(u, q.x, acc[jj].x);
            acc[jj].y = fma(u, q.y, acc[jj].y);
          }
        }
      }
      for (int jj = 0; jj < 8; ++jj)
        if (j0 + jj < k) Ob[(j0 + jj) * len + i] = acc[jj];
    }
  }
}

// Q(:, 0:k) <- Q(:, 0:m) * U(0:m, 0:k) in place (real U, column-major, leading dim ldu; k <= m <= MM):
// each thread reads all m values of its row into registers before it writes any, so the product
// needs no scratch copy.  Same summation order as k_tsgemm, so the same bits.
template <int MM>
__global__ void __launch_bounds__(256) k_tsgemm_ip(double2* __restrict__ Q, long long len, const double* __restrict__ U,
                                                   int ldu, int m, int k) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x) {
    double2 q[MM];
#pragma unroll
    for (int l = 0; l < MM; ++l) q[l] = l < m ? Q[l * len + i] : make_double2(0, 0);
    for (int j = 0; j < k; ++j) {
      double2 acc = make_double2(0, 0);
#pragma unroll
      for (int l = 0; l < MM; ++l) {
        if (l < m) {
          const double u = U[j * ldu + l];
          acc.x = fma(u, q[l].x, acc.x);
          acc.y = fma(u, q[l].y, acc.y);
        }
      }
      Q[j * len + i] = acc;
    }
  }
}

// y = S .* x (conj_twist = 0) or conj(S) .* x (conj_twist = 1), ncomp components of N points.
__global__ void k_twist(Geo g, const double2* __restrict__ x, double2* __restrict__ y, int ncomp,
                        const double2* __restrict__ tws, int tws_bstride, int conj_twist) {
  const long long N = g.N;
  const int b = blockIdx.y;
  const double2* __restrict__ S = tws + (long long)b * tws_bstride;
  const double2* xb = x + (long long)b * ncomp * N;
  double2* yb = y + (long long)b * ncomp * N;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < N; p += (long long)gridDim.x * blockDim.x) {
    const int i1 = (int)(p % g.n1);
    const long long r = p / g.n1;
    const int i2 = (int)(r % g.n2), i3 = (int)(r / g.n2);
    const double2 s = cmul(cmul(S[i1], S[g.n1 + i2]), S[g.n1 + g.n2 + i3]);
    for (int c = 0; c < ncomp; ++c) yb[c * N + p] = conj_twist ? cmulc(xb[c * N + p], s) : cmul(xb[c * N + p], s);
  }
}

// y[i] = a fixed pseudo-random pattern in [-0.5, 0.5)^2 (integer hash of i): timing input
// whose data-dependent power draw is that of real fields, unlike zeros
__global__ void k_fill_pattern(double2* __restrict__ y, long long len) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x) {
    unsigned long long h = (unsigned long long)i * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 32;
    y[i] = make_double2((double)(h & 0xFFFFFFull) / 16777216.0 - 0.5, (double)((h >> 24) & 0xFFFFFFull) / 16777216.0 - 0.5);
  }
}

__global__ void k_fill_zero(double2* __restrict__ y, long long len) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x)
    y[i] = make_double2(0, 0);
}

__global__ void k_copy(const double2* __restrict__ x, double2* __restrict__ y, long long len) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x)
    y[i] = x[i];
}

}  // namespace fame
