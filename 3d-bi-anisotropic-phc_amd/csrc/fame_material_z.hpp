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
#include "fame_kernels.hpp"

namespace fame {

constexpr int kMzThreads = 256;

// LDS: material table, the newest slice's exchange values (5 per column),
// vertex products read at neighbour positions: components 0,1 for two slices,
// components 2,3,4 for the current slice; material ids for two slices.
__host__ __device__ inline size_t material_z_lds(int TX, int TY) {
  const int RF = (TX + 2) * (TY + 2);
  return (size_t)kMaxMat * kMatWords * 16 + (size_t)(5 + 4 + 3) * RF * 16 + 2 * (size_t)RF;
}

// upper bound over all tiles with (TX+2)(TY+2) <= kMzThreads
inline size_t material_z_lds_max() {
  return (size_t)kMaxMat * kMatWords * 16 + (size_t)(5 + 4 + 3) * kMzThreads * 16 + 2 * (size_t)kMzThreads;
}

__global__ void __launch_bounds__(kMzThreads) k_material_z(Geo g, const double2* __restrict__ fin, long long in_bstride,
                                                           double2* __restrict__ fout, long long out_bstride,
                                                           const unsigned char* __restrict__ mat_id,
                                                           const double2* __restrict__ mat, int n_mat,
                                                           const double2* __restrict__ omega, int TX, int TY, int TZ,
                                                           double2* __restrict__ dpart, int pstride) {
  extern __shared__ double2 sm[];
  const int RFX = TX + 2, RFY = TY + 2, RF = RFX * RFY;
  double2* tab = sm;
  double2* F = tab + kMaxMat * kMatWords;  // [5][RF]: e_x, e_y, a_x, a_y, h_z of the newest slice
  double2* W01 = F + 5 * RF;                // [2 slices][2 comps][RF]: wh_x, wh_y
  double2* W234 = W01 + 4 * RF;             // [3][RF]: wh_z, we_x, we_y of the current slice
  unsigned char* IDS = (unsigned char*)(W234 + 3 * RF);  // [2][RF]
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
  const int x0 = (tile % ntx) * TX, y0 = (tile / ntx) * TY;
  const int z0 = zc * TZ;
  const int z1 = min(z0 + TZ, n3);
  const double2* __restrict__ H = fin + b * in_bstride;
  const double2* __restrict__ E = H + 3 * N;
  double2* __restrict__ HO = fout + b * out_bstride;
  double2* __restrict__ EO = HO + 3 * N;
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

  auto load = [&](int z, double2& a0, double2& a1, double2& a2, double2& c0, double2& c1, double2& c2, int& id) {
    if (act) {
      const long long p = col + plane * z;
      a0 = E[p]; a1 = E[N + p]; a2 = E[2 * N + p];
      c0 = H[p]; c1 = H[N + p]; c2 = H[2 * N + p];
      id = mat_id[p];
    }
  };
  // exchange values of a new slice (needs the previous slice's e_z, h_x, h_y for the z-partials)
  auto put_slice = [&](int buf, const double2& ce0, const double2& ce1, const double2& ch0, const double2& ch1,
                       const double2& ch2, const double2& pe0h, const double2& ph0, const double2& ph1, int id) {
    (void)pe0h;
    if (act) {
      F[0 * RF + t] = ce0;
      F[1 * RF + t] = ce1;
      F[2 * RF + t] = cfmac(ch0, om2, ph0);
      F[3 * RF + t] = cfmac(ch1, om2, ph1);
      F[4 * RF + t] = ch2;
      IDS[buf * RF + t] = (unsigned char)id;
    }
  };
  // vertex products of the newest slice; comps 0,1 go to LDS, 2..5 are returned
  auto products = [&](int buf, const double2& vez, int id, double2& w2, double2& w3, double2& w4, double2& w5) {
    if (!wact) return;
    const double2 vex = cscale(cfmac(F[t], om0, F[t - 1]), 0.5);
    const double2 vey = cscale(cfmac(F[RF + t], om1, F[RF + t - RFX]), 0.5);
    const double2 vhx = cscale(cfmac(F[2 * RF + t], om1, F[2 * RF + t - RFX]), 0.25);
    const double2 vhy = cscale(cfmac(F[3 * RF + t], om0, F[3 * RF + t - 1]), 0.25);
    const double2 hz_a = cfmac(F[4 * RF + t], om0, F[4 * RF + t - 1]);
    const double2 hz_c = cfmac(F[4 * RF + t - RFX], om0, F[4 * RF + t - RFX - 1]);
    const double2 vhz = cscale(cfmac(hz_a, om1, hz_c), 0.25);
    const double2* T = tab + kMatWords * id;
    const double2* ep = T;
    const double2* mu = T + 9;
    const double2* xi = T + 18;
    W01[(buf * 2 + 0) * RF + t] = cfmac(cfmac(cfmac(cfma(cmul(mu[1], vhy), mu[2], vhz), xi[0], vex), xi[3], vey), xi[6], vez);
    W01[(buf * 2 + 1) * RF + t] = cfmac(cfmac(cfmac(cfma(cmul(mu[3], vhx), mu[5], vhz), xi[1], vex), xi[4], vey), xi[7], vez);
    w2 = cfmac(cfmac(cfmac(cfma(cmul(mu[6], vhx), mu[7], vhy), xi[2], vex), xi[5], vey), xi[8], vez);
    w3 = cfma(cfma(cfma(cfma(cmul(ep[1], vey), ep[2], vez), xi[0], vhx), xi[1], vhy), xi[2], vhz);
    w4 = cfma(cfma(cfma(cfma(cmul(ep[3], vex), ep[5], vez), xi[3], vhx), xi[4], vhy), xi[5], vhz);
    w5 = cfma(cfma(cfma(cfma(cmul(ep[6], vex), ep[7], vey), xi[6], vhx), xi[7], vhy), xi[8], vhz);
  };

  __syncthreads();  // material table
  // prime: slice z0-1 (its e_z, h_x, h_y) and slice z0; products of z0
  double2 e0 = z2, e1 = z2, e2 = z2, h0 = z2, h1 = z2, h2 = z2;  // current slice fields
  double2 pe2 = z2, ph0 = z2, ph1 = z2;                           // previous slice partial inputs
  double2 hold2 = z2, hold3 = z2, hold4 = z2, cur5 = z2, nxt5 = z2;
  {
    int id0 = 0;
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
  int id_n = 0;
  load((z0 + 1) % n3, ne0, ne1, ne2, nh0, nh1, nh2, id_n);
  for (int s = z0; s < z1; ++s) {
    double2 fe0 = z2, fe1 = z2, fe2 = z2, fh0 = z2, fh1 = z2, fh2 = z2;
    int id_f = 0;
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
      const unsigned char* Ic = IDS + buf * RF;
      const unsigned char* In = IDS + (buf ^ 1) * RF;
      auto epd = [&](int id, int i) { return tab[kMatWords * id + 4 * i].x; };
      auto mud = [&](int id, int i) { return tab[kMatWords * id + 9 + 4 * i].x; };
      const int i00 = Ic[t], i10 = Ic[t + 1], i01 = Ic[t + RFX], i11 = Ic[t + RFX + 1];
      const int j00 = In[t], j10 = In[t + 1], j01 = In[t + RFX];
      const long long p = col + plane * s;
      {  // edges
        const double d0 = 0.5 * (epd(i00, 0) + epd(i10, 0));
        const double d1 = 0.5 * (epd(i00, 1) + epd(i01, 1));
        const double d2 = 0.5 * (epd(i00, 2) + epd(j00, 2));
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
        const double d0 = 0.25 * (mud(i00, 0) + mud(i01, 0) + mud(j00, 0) + mud(j01, 0));
        const double d1 = 0.25 * (mud(i00, 1) + mud(i10, 1) + mud(j00, 1) + mud(j10, 1));
        const double d2 = 0.25 * (mud(i00, 2) + mud(i10, 2) + mud(i01, 2) + mud(i11, 2));
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
