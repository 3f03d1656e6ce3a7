// Mixed-radix complex FP64 FFT stages that run on data resident in LDS.
//
// Decimation in frequency, in place (Cooley-Tukey 1965; Gentleman & Sande
// 1966 for the DIF ordering): for n = r_0 r_1 ... r_{S-1}, stage s splits
// every block of length L_s = n / (r_0 ... r_{s-1}) into r_s interleaved
// sub-sequences of length m_s = L_s / r_s:
//     y_q[j] = W_{L_s}^{j q} * sum_p x[j + p m_s] W_{r_s}^{p q},
// stored back at position j + q m_s.  After the last stage the frequency k
// (mixed-radix digits k = q_0 + r_0 q_1 + r_0 r_1 q_2 + ...) sits at
// position sum_s q_s m_s; the host builds that permutation once per axis
// length and the kernels apply it when they move data from LDS to HBM.
//
// DIR = -1 is the forward transform (e^{-2 pi i jk/n}, MATLAB fft),
// DIR = +1 the inverse kernel without 1/n (MATLAB ifft * n).
#pragma once
#include <hip/hip_runtime.h>

namespace fame {

constexpr int kMaxStages = 16;
constexpr int kMaxRadix = 16;

// Per-axis transform description, passed by value to kernels.
struct AxisPlan {
  int n;                    // transform length
  int nstage;               // number of radix stages
  int radix[kMaxStages];    // r_s
  int blen[kMaxStages];     // L_s (block length entering stage s)
  int tw_off;               // offset of this length's twiddle table (n entries e^{-2 pi i t/n})
  int perm_off;             // offset of the digit-reversal table (n ints)
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// multiply by i*s (s = +-1 folded into the caller)
__device__ __forceinline__ double2 cmul_i(double2 a) { return make_double2(-a.y, a.x); }
// fused a + b*c
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
  return make_double2(fma(b.x, c.x, fma(-b.y, c.y, a.x)), fma(b.x, c.y, fma(b.y, c.x, a.y)));
}
// fused a + conj(b)*c
__device__ __forceinline__ double2 cfmac(double2 a, double2 b, double2 c) {
  return make_double2(fma(b.x, c.x, fma(b.y, c.y, a.x)), fma(b.x, c.y, fma(-b.y, c.x, a.y)));
}

// ---------------------------------------------------------------------------
// Small DFT kernels on registers.  Constants are the exact trigonometric
// values cos(2 pi/5) etc. written out to 20 significant digits.
// ---------------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ void dft2(double2& a, double2& b) {
  double2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR>
__device__ __forceinline__ void dft3(double2& a, double2& b, double2& c) {
  constexpr double h = 0.86602540378443864676;  // sqrt(3)/2
  double2 s = cadd(b, c), d = csub(b, c);
  double2 t = make_double2(a.x - 0.5 * s.x, a.y - 0.5 * s.y);
  a = cadd(a, s);
  // forward: y1 = t - i h d, y2 = t + i h d ; inverse swaps the signs
  double2 id = make_double2(-DIR * h * d.y, DIR * h * d.x);  // DIR * i * h * d
  b = cadd(t, id);
  c = csub(t, id);
}

template <int DIR>
__device__ __forceinline__ void dft4(double2& a, double2& b, double2& c, double2& d) {
  double2 s0 = cadd(a, c), d0 = csub(a, c);
  double2 s1 = cadd(b, d), d1 = csub(b, d);
  // forward W4 = -i: y1 = d0 - i d1, y3 = d0 + i d1
  double2 jd1 = make_double2(-DIR * d1.y, DIR * d1.x);  // DIR * i * d1
  a = cadd(s0, s1);
  c = csub(s0, s1);
  b = cadd(d0, jd1);
  d = csub(d0, jd1);
}

template <int DIR>
__device__ __forceinline__ void dft5(double2& a0, double2& a1, double2& a2, double2& a3, double2& a4) {
  constexpr double c1 = 0.30901699437494742410;   // cos(2 pi/5)
  constexpr double c2 = -0.80901699437494742410;  // cos(4 pi/5)
  constexpr double s1 = 0.95105651629515357212;   // sin(2 pi/5)
  constexpr double s2 = 0.58778525229247312917;   // sin(4 pi/5)
  double2 p14 = cadd(a1, a4), m14 = csub(a1, a4);
  double2 p23 = cadd(a2, a3), m23 = csub(a2, a3);
  double2 r1 = make_double2(a0.x + c1 * p14.x + c2 * p23.x, a0.y + c1 * p14.y + c2 * p23.y);
  double2 r2 = make_double2(a0.x + c2 * p14.x + c1 * p23.x, a0.y + c2 * p14.y + c1 * p23.y);
  double2 i1 = make_double2(s1 * m14.x + s2 * m23.x, s1 * m14.y + s2 * m23.y);
  double2 i2 = make_double2(s2 * m14.x - s1 * m23.x, s2 * m14.y - s1 * m23.y);
  // forward: y1 = r1 - i*i1, y4 = r1 + i*i1, y2 = r2 - i*i2, y3 = r2 + i*i2
  double2 j1 = make_double2(-DIR * i1.y, DIR * i1.x);
  double2 j2 = make_double2(-DIR * i2.y, DIR * i2.x);
  a0 = make_double2(a0.x + p14.x + p23.x, a0.y + p14.y + p23.y);
  a1 = cadd(r1, j1);
  a4 = csub(r1, j1);
  a2 = cadd(r2, j2);
  a3 = csub(r2, j2);
}

// Generic radix-r DFT (r <= kMaxRadix) using the length-n twiddle table:
// W_r^{pq} = W_n^{(n/r) pq mod n}.
template <int DIR>
__device__ __forceinline__ void dft_generic(double2* v, int r, int n, const double2* __restrict__ tw) {
  double2 y[kMaxRadix];
  const int step = n / r;
  for (int q = 0; q < r; ++q) {
    double2 acc = v[0];
    for (int p = 1; p < r; ++p) {
      int e = (step * ((p * q) % r));
      double2 w = tw[e];
      if (DIR > 0) w = cconj(w);
      acc = cfma(acc, v[p], w);
    }
    y[q] = acc;
  }
  for (int q = 0; q < r; ++q) v[q] = y[q];
}

// One DIF stage applied to every line in an LDS buffer.  Element i of line l
// is at buf[l * lstride + i * estride].  All threads of the block take part.
template <int DIR>
__device__ __forceinline__ void lds_stage(double2* buf, int nlines, int estride, int lstride, int n, int r, int L,
                                          const double2* __restrict__ tw) {
  const int m = L / r;
  const int tstep = n / L;
  const int nbf = n / r;  // butterflies per line
  const int total = nbf * nlines;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = t / nbf;
    const int b = t - line * nbf;
    const int blk = b / m;
    const int j = b - blk * m;
    double2* base = buf + line * lstride + (blk * L + j) * estride;
    const int es = m * estride;
    double2 v[kMaxRadix];
    for (int q = 0; q < r; ++q) v[q] = base[q * es];
    switch (r) {
      case 2: dft2<DIR>(v[0], v[1]); break;
      case 3: dft3<DIR>(v[0], v[1], v[2]); break;
      case 4: dft4<DIR>(v[0], v[1], v[2], v[3]); break;
      case 5: dft5<DIR>(v[0], v[1], v[2], v[3], v[4]); break;
      default: dft_generic<DIR>(v, r, n, tw); break;
    }
    base[0] = v[0];
    if (j == 0) {
      for (int q = 1; q < r; ++q) base[q * es] = v[q];
    } else {
      // j*q < L, so j*q*tstep < n: no reduction needed
      for (int q = 1; q < r; ++q) {
        double2 w = tw[j * q * tstep];
        base[q * es] = (DIR < 0) ? cmul(v[q], w) : cmulc(v[q], w);
      }
    }
  }
}

// Full transform of every line (all stages, barriers between stages).
template <int DIR>
__device__ __forceinline__ void lds_fft(double2* buf, int nlines, int estride, int lstride, const AxisPlan& ap,
                                        const double2* __restrict__ twbase) {
  const double2* tw = twbase + ap.tw_off;
  for (int s = 0; s < ap.nstage; ++s) {
    lds_stage<DIR>(buf, nlines, estride, lstride, ap.n, ap.radix[s], ap.blen[s], tw);
    __syncthreads();
  }
}

}  // namespace fame
