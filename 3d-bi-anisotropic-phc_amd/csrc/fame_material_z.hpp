// Material stencil, 2.5-D: each work-group owns a TX x TY tile of (x, y)
// columns and marches through a range of z-slices.  Thread (fx, fy) owns one
// column of the tile grown by one on every side, so
//   * z-neighbours (slice s-1 for the gathers, s+1 for the scatters) stay in
//     the thread's registers or in a two-slice vertex buffer,
//   * x/y neighbours are exchanged through small per-slice LDS arrays,
// and every field value is fetched from HBM once (plus the one-column halo).
// The arithmetic is the vertex form of k_material (same gathers, 6x6 vertex
// products and scatters).
#pragma once
#include "fame_kernels.hpp"

namespace fame {

constexpr int kMzThreads = 256;

__host__ __device__ inline size_t material_z_lds(int TX, int TY) {
  const int RF = (TX + 2) * (TY + 2);
  return (size_t)kMaxMat * kMatWords * 16 + (size_t)(5 + 12) * RF * 16 + 2 * (size_t)RF;
}

// upper bound over all tiles with (TX+2)(TY+2) <= kMzThreads
inline size_t material_z_lds_max() {
  return (size_t)kMaxMat * kMatWords * 16 + (size_t)(5 + 12) * kMzThreads * 16 + 2 * (size_t)kMzThreads;
}

__global__ void __launch_bounds__(kMzThreads) k_material_z(Geo g, const double2* __restrict__ fin, long long in_bstride,
                                                           double2* __restrict__ fout, long long out_bstride,
                                                           const unsigned char* __restrict__ mat_id,
                                                           const double2* __restrict__ mat, int n_mat,
                                                           const double2* __restrict__ omega, int TX, int TY, int TZ) {
  extern __shared__ double2 sm[];
  const int RFX = TX + 2, RFY = TY + 2, RF = RFX * RFY;
  double2* tab = sm;
  double2* F = tab + kMaxMat * kMatWords;  // [5][RF]: e_x, e_y, a_x, a_y, h_z of the newest slice
  double2* W = F + 5 * RF;                  // [2][6][RF]: vertex products of two slices
  unsigned char* IDS = (unsigned char*)(W + 12 * RF);  // [2][RF]
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  const long long N = g.N;
  const int ntx = (n1 + TX - 1) / TX;
  const int x0 = (blockIdx.x % ntx) * TX, y0 = (blockIdx.x / ntx) * TY;
  const int z0 = blockIdx.y * TZ;
  const int z1 = min(z0 + TZ, n3);
  const int b = blockIdx.z;
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

  // registers: fields of the current slice and z-partials of the previous one
  double2 e0 = make_double2(0, 0), e1 = e0, e2 = e0, h0 = e0, h1 = e0, h2 = e0;
  double2 pe2 = e0, ph0 = e0, ph1 = e0;
  int id_cur = 0;

  auto load = [&](int z, double2& a0, double2& a1, double2& a2, double2& c0, double2& c1, double2& c2, int& id) {
    if (act) {
      const long long p = col + plane * z;
      a0 = E[p]; a1 = E[N + p]; a2 = E[2 * N + p];
      c0 = H[p]; c1 = H[N + p]; c2 = H[2 * N + p];
      id = mat_id[p];
    }
  };

  // vertex products of the slice whose fields are (e*, h*) given the previous slice's (pe2, ph0, ph1)
  auto vertex = [&](int buf, const double2& ce0, const double2& ce1, const double2& ce2, const double2& ch0,
                    const double2& ch1, const double2& ch2, int id) {
    const double2 vez = cscale(cfmac(ce2, om2, pe2), 0.5);
    if (act) {
      F[0 * RF + t] = ce0;
      F[1 * RF + t] = ce1;
      F[2 * RF + t] = cfmac(ch0, om2, ph0);
      F[3 * RF + t] = cfmac(ch1, om2, ph1);
      F[4 * RF + t] = ch2;
      IDS[buf * RF + t] = (unsigned char)id;
    }
    __syncthreads();
    if (wact) {
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
      double2* Wb = W + buf * 6 * RF;
      Wb[0 * RF + t] = cfmac(cfmac(cfmac(cfma(cmul(mu[1], vhy), mu[2], vhz), xi[0], vex), xi[3], vey), xi[6], vez);
      Wb[1 * RF + t] = cfmac(cfmac(cfmac(cfma(cmul(mu[3], vhx), mu[5], vhz), xi[1], vex), xi[4], vey), xi[7], vez);
      Wb[2 * RF + t] = cfmac(cfmac(cfmac(cfma(cmul(mu[6], vhx), mu[7], vhy), xi[2], vex), xi[5], vey), xi[8], vez);
      Wb[3 * RF + t] = cfma(cfma(cfma(cfma(cmul(ep[1], vey), ep[2], vez), xi[0], vhx), xi[1], vhy), xi[2], vhz);
      Wb[4 * RF + t] = cfma(cfma(cfma(cfma(cmul(ep[3], vex), ep[5], vez), xi[3], vhx), xi[4], vhy), xi[5], vhz);
      Wb[5 * RF + t] = cfma(cfma(cfma(cfma(cmul(ep[6], vex), ep[7], vey), xi[6], vhx), xi[7], vhy), xi[8], vhz);
    }
  };

  __syncthreads();  // material table
  // prime: previous slice z0-1 (only e_z, h_x, h_y are needed from it), then W(z0)
  {
    double2 a0, a1, c2;
    int idd = 0;
    load((z0 - 1 + n3) % n3, a0, a1, pe2, ph0, ph1, c2, idd);
    load(z0, e0, e1, e2, h0, h1, h2, id_cur);
    vertex(0, e0, e1, e2, h0, h1, h2, id_cur);
    pe2 = e2;
    ph0 = h0;
    ph1 = h1;
  }
  int buf = 0;
  // software pipeline: slice s+1 is already in registers when step s starts; slice s+2 is requested
  double2 ne0 = make_double2(0, 0), ne1 = ne0, ne2 = ne0, nh0 = ne0, nh1 = ne0, nh2 = ne0;
  int id_n = 0;
  load((z0 + 1) % n3, ne0, ne1, ne2, nh0, nh1, nh2, id_n);
  for (int s = z0; s < z1; ++s) {
    const int zf = (s + 2) % n3;
    double2 fe0 = make_double2(0, 0), fe1 = fe0, fe2 = fe0, fh0 = fe0, fh1 = fe0, fh2 = fe0;
    int id_f = 0;
    if (s + 1 < z1) load(zf, fe0, fe1, fe2, fh0, fh1, fh2, id_f);
    __syncthreads();  // all reads of F (and of W[buf ^ 1]) from the previous step are done
    vertex(buf ^ 1, ne0, ne1, ne2, nh0, nh1, nh2, id_n);
    __syncthreads();
    if (oact) {
      const double2* Wc = W + buf * 6 * RF;
      const double2* Wn = W + (buf ^ 1) * 6 * RF;
      const unsigned char* Ic = IDS + buf * RF;
      const unsigned char* In = IDS + (buf ^ 1) * RF;
      auto epd = [&](int id, int i) { return tab[kMatWords * id + 4 * i].x; };
      auto mud = [&](int id, int i) { return tab[kMatWords * id + 9 + 4 * i].x; };
      const int i00 = Ic[t], i10 = Ic[t + 1], i01 = Ic[t + RFX], i11 = Ic[t + RFX + 1];
      const int j00 = In[t], j10 = In[t + 1], j01 = In[t + RFX];
      const long long p = col + plane * s;
      // edges
      {
        const double d0 = 0.5 * (epd(i00, 0) + epd(i10, 0));
        const double d1 = 0.5 * (epd(i00, 1) + epd(i01, 1));
        const double d2 = 0.5 * (epd(i00, 2) + epd(j00, 2));
        const double2 y0 = cscale(cfma(Wc[3 * RF + t], om0, Wc[3 * RF + t + 1]), 0.5);
        const double2 y1 = cscale(cfma(Wc[4 * RF + t], om1, Wc[4 * RF + t + RFX]), 0.5);
        const double2 y2 = cscale(cfma(Wc[5 * RF + t], om2, Wn[5 * RF + t]), 0.5);
        EO[p] = make_double2(y0.x + d0 * e0.x, y0.y + d0 * e0.y);
        EO[N + p] = make_double2(y1.x + d1 * e1.x, y1.y + d1 * e1.y);
        EO[2 * N + p] = make_double2(y2.x + d2 * e2.x, y2.y + d2 * e2.y);
      }
      // faces: h_x over (y, z), h_y over (x, z), h_z over (x, y)
      {
        const int j11x = In[t + RFX];  // (x, y+1, s+1)
        const int j11y = In[t + 1];    // (x+1, y, s+1)
        const double d0 = 0.25 * (mud(i00, 0) + mud(i01, 0) + mud(j00, 0) + mud(j11x, 0));
        const double d1 = 0.25 * (mud(i00, 1) + mud(i10, 1) + mud(j00, 1) + mud(j11y, 1));
        const double d2 = 0.25 * (mud(i00, 2) + mud(i10, 2) + mud(i01, 2) + mud(i11, 2));
        double2 a = cfma(Wc[t], om1, Wc[t + RFX]);
        double2 c = cfma(Wn[t], om1, Wn[t + RFX]);
        const double2 y0 = cscale(cfma(a, om2, c), 0.25);
        a = cfma(Wc[RF + t], om0, Wc[RF + t + 1]);
        c = cfma(Wn[RF + t], om0, Wn[RF + t + 1]);
        const double2 y1 = cscale(cfma(a, om2, c), 0.25);
        a = cfma(Wc[2 * RF + t], om0, Wc[2 * RF + t + 1]);
        c = cfma(Wc[2 * RF + t + RFX], om0, Wc[2 * RF + t + RFX + 1]);
        const double2 y2 = cscale(cfma(a, om1, c), 0.25);
        HO[p] = make_double2(y0.x + d0 * h0.x, y0.y + d0 * h0.y);
        HO[N + p] = make_double2(y1.x + d1 * h1.x, y1.y + d1 * h1.y);
        HO[2 * N + p] = make_double2(y2.x + d2 * h2.x, y2.y + d2 * h2.y);
      }
      (void)j01;
      (void)j10;
    }
    // advance: the new slice becomes current; its z-partials feed the next vertex step
    pe2 = ne2;
    ph0 = nh0;
    ph1 = nh1;
    e0 = ne0; e1 = ne1; e2 = ne2;
    h0 = nh0; h1 = nh1; h2 = nh2;
    id_cur = id_n;
    ne0 = fe0; ne1 = fe1; ne2 = fe2;
    nh0 = fh0; nh1 = fh1; nh2 = fh2;
    id_n = id_f;
    buf ^= 1;
  }
}

}  // namespace fame
