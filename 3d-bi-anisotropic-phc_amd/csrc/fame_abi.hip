// Host side of the MI355X curl-SVD + bi-anisotropic material library: plans,
// launch sequences, CG (fncinvB) and Lanczos, behind the C-ABI of
// include/fame_abi.h.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/fame_abi.h"
#include "fame_kernels.hpp"
#include "fame_fft_fast.hpp"
#include "fame_material_z.hpp"

using namespace fame;

namespace {

thread_local std::string g_err;

int set_err(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return set_err(FAME_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));         \
  } while (0)

#define CHECK_LAUNCH()                                                                         \
  do {                                                                                         \
    hipError_t e_ = hipGetLastError();                                                         \
    if (e_ != hipSuccess) return set_err(FAME_ERR_HIP, std::string("launch: ") + hipGetErrorString(e_)); \
  } while (0)

#define RET_IF(x)            \
  do {                       \
    int rc_ = (x);           \
    if (rc_ != FAME_OK) return rc_; \
  } while (0)

constexpr int kBlk = 256;

// Factor n into radices (4 first, then 2, 3, 5, 7, 11, 13).
bool factor_radices(int n, std::vector<int>& r) {
  r.clear();
  int m = n;
  while (m % 4 == 0) { r.push_back(4); m /= 4; }
  while (m % 2 == 0) { r.push_back(2); m /= 2; }
  for (int p : {3, 5, 7, 11, 13})
    while (m % p == 0) { r.push_back(p); m /= p; }
  if (m != 1) return false;
  if ((int)r.size() > kMaxStages) return false;
  if (r.empty()) r.push_back(1);
  return true;
}

}  // namespace

struct fame_plan {
  int n1 = 0, n2 = 0, n3 = 0, nb = 0;
  long long N = 0;
  int device = 0;                    // the HIP device current when the plan was created
  int ncu = 256;                     // compute units of that device
  hipStream_t stream = nullptr;      // stream in use
  hipStream_t own_stream = nullptr;  // stream created by the plan
  Geo geo{};
  double delta[3] = {0, 0, 0};
  std::vector<double> theta;  // [nb][3]
  int n_mat = 0;
  // device tables
  double2* d_tw = nullptr;    // all twiddle tables
  int* d_perm = nullptr;      // all digit-reversal tables
  double2* d_lam = nullptr;   // [nb][n1+n2+n3]
  double2* d_tws = nullptr;   // [nb][n1+n2+n3]  Bloch twist factors
  double2* d_omega = nullptr; // [nb][3]
  unsigned char* d_mat_id = nullptr;
  double2* d_mat = nullptr;   // [n_mat][27]
  // scratch
  double2* d_w1 = nullptr;    // [nb][6N]
  double2* d_w2 = nullptr;    // [nb][6N]
  double2* d_part = nullptr;  // reduction partials
  double2* d_part2 = nullptr; // fused-CG p^H q partials (material stencil or MIX_OUT pass)
  size_t part2_cap = 0;       // elements of d_part2
  double2* d_mr = nullptr;    // MINRES work vectors [6][nb][4N], allocated by the first fame_minres
  void* d_mrst = nullptr;     // MINRES per-item state [2][nb] (the iteration alternates halves)
  double2* d_mrp = nullptr;   // MINRES partial sums [3][nb][2048] and alpha [nb]
  double2* d_mrm = nullptr;   // MINRES: the material pass's partial sums of u^H A u [nb][mrm_stride]
  CgState* d_mrskip = nullptr;  // MINRES: per-item active flags for the fncB passes [nb]
  void* h_mr = nullptr;         // MINRES: pinned copy of the state for the lagged poll [nb]
  hipEvent_t ev_mr = nullptr;
  size_t mrm_cap = 0;
  double2* d_mgp = nullptr;   // fame_mgs partial sums [2][1024]
  double2* d_mgs = nullptr;   // fame_mgs coefficients when the caller passes none
  int mgs_cap = 0;
  double2* d_bdp = nullptr;   // fame_block_dot partial sums, grown to the largest m * work-groups seen
  size_t bdp_cap = 0;
  double2* d_jdy = nullptr;   // JD-SIRA step coefficients / projection columns
  size_t jdy_cap = 0;
  bool cg_fused = true;       // CG vector updates fused into fncB (specialised z kernels only)
  int cg_seq = 0;             // > 0: fused CG runs chunks of this many items to convergence one after another
  int mt[3] = {0, 0, 0};      // material stencil tile TX, TY and slab TZ (0: heuristic)
  bool cg_graph = false;      // fused CG iterations replayed from a captured HIP graph
  // per-vertex bi-isotropic material (eps_v, gamma_v at every vertex; replaces the id table)
  bool pv = false;
  double pv_G[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double pv_mu = 1.0;
  double* d_eps = nullptr;
  double* d_gam = nullptr;
  double2* d_pvw = nullptr;   // vertex products, [nb][6N]
  double2* d_red = nullptr;   // [nb] reduced values
  int nblk_red = 0;
  // CG scratch
  double2* d_r = nullptr;
  double2* d_p = nullptr;
  double2* d_q = nullptr;
  CgState* d_cg = nullptr;
  CgState* h_cg = nullptr;    // pinned host mirror of d_cg for the lagged convergence poll
  int h_cg_cap = 0;
  hipEvent_t ev_poll = nullptr;
  // extra streams for chunk groups: chunk k of fncB and of a lock-step CG iteration runs on
  // stream k % nstreams with its own part of the field scratch, so one chunk's kernels fill the
  // others' launch tails; fork / join by events on the caller's stream
  int nstreams = 1;
  hipStream_t xstream[2] = {nullptr, nullptr};  // extra streams (chunk k on stream k % nstreams)
  hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
  int w_off = 0;              // items into d_w1 / d_w2 where the current window's fields live
  // pinned host staging for the small device-to-host reads a call waits for (solver states,
  // inner products): a copy into pageable memory goes through the runtime's own staging
  void* h_stage = nullptr;
  size_t h_stage_cap = 0;
  int tab_len = 0;            // n1+n2+n3
  bool fast = true;           // use the specialised FFT kernels where a size has them
  // batch window used by the launch helpers: items [cb0, cb0 + cnb) of the current view
  int cb0 = 0, cnb = 0;
  // view: the batch the operators see.  By default all nb items with the plan's tables;
  // a packed view (set_view) holds m < nb items whose table rows were gathered into
  // d_*_v, so solvers can run only on the k-points that are still iterating.
  int vnb = 0;
  const double2* v_lam = nullptr;
  const double2* v_tws = nullptr;
  const double2* v_om = nullptr;
  double2* d_lam_v = nullptr;
  double2* d_tws_v = nullptr;
  double2* d_om_v = nullptr;
  int* d_idx = nullptr;       // [nb] item indices of the packed view
  int chunk = 0;              // fncB processes the batch in chunks of this many items (0: all)
  int xyv = 0;                // xy-pass launch shape: 0 = default threads, 1 = 512 threads (default from n1 >= 40),
                              // 2 = 1024 threads for planes of n1 >= 64 (512 below)
  int zv = 0;                 // z-pass launch shape: 0 = 16 pencils/256 thr, 1 = 8/128, 2 = 32/512, 3 = 16/320,
                              // 4 = 16/384
  // per-stage timing of fncB (HIP events between the five kernels), off by default
  bool prof = false;
  hipEvent_t ev[6] = {};
  double stage_ms[5] = {0, 0, 0, 0, 0};
  long long prof_calls = 0;
  hipEvent_t evk[2] = {};      // the material stencil's own launch events (profiling)
  bool evk_used = false;       // the last fncB window launched the stencil with them
  double mat_kernel_ms = 0;    // accumulated stencil kernel time (profiling)
  long long mat_launches = 0;
};

namespace {

inline const double2* wlam(const fame_plan* P) { return P->v_lam + (size_t)P->cb0 * P->tab_len; }
inline const double2* wtws(const fame_plan* P) { return P->v_tws + (size_t)P->cb0 * P->tab_len; }
inline const double2* wom(const fame_plan* P) { return P->v_om + (size_t)P->cb0 * 3; }

int build_axis(int n, std::vector<double2>& tw, std::vector<int>& perm, AxisPlan& ap) {
  std::vector<int> r;
  if (!factor_radices(n, r))
    return set_err(FAME_ERR_UNSUPPORTED, "grid length " + std::to_string(n) + " has a prime factor > 13");
  ap.n = n;
  ap.nstage = (n == 1) ? 0 : (int)r.size();
  int L = n;
  for (int s = 0; s < ap.nstage; ++s) {
    ap.radix[s] = r[s];
    ap.blen[s] = L;
    L /= r[s];
  }
  ap.tw_off = (int)tw.size();
  ap.perm_off = (int)perm.size();
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int t = 0; t < n; ++t) {
    const long double a = two_pi * (long double)t / (long double)n;
    tw.push_back(make_double2((double)cosl(a), (double)(-sinl(a))));
  }
  // frequency k sits at position sum_s q_s m_s (see fame_fft.hpp)
  for (int k = 0; k < n; ++k) {
    int kk = k, pos = 0, Ls = n;
    for (int s = 0; s < ap.nstage; ++s) {
      const int rs = ap.radix[s];
      const int q = kk % rs;
      kk /= rs;
      const int m = Ls / rs;
      pos += q * m;
      Ls = m;
    }
    perm.push_back(pos);
  }
  return FAME_OK;
}

// lambda_d(j) = (exp(i(2 theta(n,j+1) + theta_d/n)) - 1)/delta_d (lambda_QuasiPeriodic.m:3),
// S_d(j) = exp(i j theta_d/n) (S_sDFT.m:2), omega_d = exp(i theta_d/n).
void fill_k_tables(const fame_plan* P, std::vector<double2>& lam, std::vector<double2>& tws,
                   std::vector<double2>& om) {
  const int dims[3] = {P->n1, P->n2, P->n3};
  lam.assign((size_t)P->nb * P->tab_len, make_double2(0, 0));
  tws.assign((size_t)P->nb * P->tab_len, make_double2(0, 0));
  om.assign((size_t)P->nb * 3, make_double2(0, 0));
  for (int b = 0; b < P->nb; ++b) {
    int off = 0;
    for (int d = 0; d < 3; ++d) {
      const int m = dims[d];
      const double th = P->theta[3 * b + d];
      for (int j = 0; j < m; ++j) {
        const double thm = (double)j * M_PI / (double)m;  // DistOp.theta(m, j+1)
        const double arg = 2.0 * thm + th / (double)m;
        lam[(size_t)b * P->tab_len + off + j] =
            make_double2((std::cos(arg) - 1.0) / P->delta[d], std::sin(arg) / P->delta[d]);
        const double as = (double)j * th / (double)m;
        tws[(size_t)b * P->tab_len + off + j] = make_double2(std::cos(as), std::sin(as));
      }
      om[3 * b + d] = make_double2(std::cos(th / m), std::sin(th / m));
      off += m;
    }
  }
}

// Host check that the curl SVD is defined at every Fourier point of every item: Pi_0 and Pi_2
// divide by sqrt(Lambda_q) and Pi_1 by sqrt(Lambda_q - |w . lambda|^2) (SVD_Matrix_curl.m:45-57),
// and both vanish when lambda is parallel to w = (1,2,3)/sqrt(14), which includes lambda = 0 (the
// Gamma point: theta_d = 0 mod 2 pi on all three axes).  The reference avoids that point by
// nudging k by +1e-4 (band_structure_woodpile_Lanczos_n60_gamma0p2.m:50-52).  Flags a divisor below
// 1e-12 of the largest |lambda| (relative squared value 1e-24).
int check_theta(int n1, int n2, int n3, int nb, const double* theta, const double* delta) {
  const int dims[3] = {n1, n2, n3};
  double scale = 0;
  for (int d = 0; d < 3; ++d) {
    if (!(delta[d] > 0) || !std::isfinite(delta[d])) return set_err(FAME_ERR_ARG, "delta must be positive and finite");
    scale += 4.0 / (delta[d] * delta[d]);
  }
  const double ws = 0.26726124191242438468, w[3] = {ws, 2 * ws, 3 * ws};
  std::vector<double2> lam[3];
  for (int b = 0; b < nb; ++b) {
    for (int d = 0; d < 3; ++d) {
      const double th = theta[3 * b + d];
      if (!std::isfinite(th)) return set_err(FAME_ERR_ARG, "theta must be finite");
      lam[d].resize(dims[d]);
      for (int j = 0; j < dims[d]; ++j) {
        const double arg = 2.0 * ((double)j * M_PI / (double)dims[d]) + th / (double)dims[d];
        lam[d][j] = make_double2((std::cos(arg) - 1.0) / delta[d], std::sin(arg) / delta[d]);
      }
    }
    double worst = 1e300;
    for (int j3 = 0; j3 < n3; ++j3)
      for (int j2 = 0; j2 < n2; ++j2)
        for (int j1 = 0; j1 < n1; ++j1) {
          const double2 a = lam[0][j1], c = lam[1][j2], e = lam[2][j3];
          const double lq = a.x * a.x + a.y * a.y + c.x * c.x + c.y * c.y + e.x * e.x + e.y * e.y;
          const double wx = w[0] * a.x + w[1] * c.x + w[2] * e.x, wy = w[0] * a.y + w[1] * c.y + w[2] * e.y;
          worst = std::min(worst, lq - (wx * wx + wy * wy));
        }
    if (!(worst > 1e-24 * scale)) {
      char msg[320];
      std::snprintf(msg, sizeof msg,
                    "k-point %d (theta = %.17g, %.17g, %.17g): the curl SVD is singular at a Fourier point "
                    "(Lambda_q - |w.lambda|^2 = %.3g); at Gamma nudge k by +1e-4 as "
                    "band_structure_woodpile_Lanczos_n60_gamma0p2.m:50-52 does",
                    b, theta[3 * b], theta[3 * b + 1], theta[3 * b + 2], worst);
      return set_err(FAME_ERR_ARG, msg);
    }
  }
  return FAME_OK;
}

int upload_k_tables(fame_plan* P) {
  std::vector<double2> lam, tws, om;
  fill_k_tables(P, lam, tws, om);
  HIP_TRY(hipMemcpyAsync(P->d_lam, lam.data(), lam.size() * sizeof(double2), hipMemcpyHostToDevice, P->stream));
  HIP_TRY(hipMemcpyAsync(P->d_tws, tws.data(), tws.size() * sizeof(double2), hipMemcpyHostToDevice, P->stream));
  HIP_TRY(hipMemcpyAsync(P->d_omega, om.data(), om.size() * sizeof(double2), hipMemcpyHostToDevice, P->stream));
  HIP_TRY(hipStreamSynchronize(P->stream));
  return FAME_OK;
}

size_t zpass_lds(const fame_plan* P) { return (size_t)3 * kZTile * P->n3 * sizeof(double2); }
size_t xypass_lds(const fame_plan* P) { return (size_t)P->n1 * P->n2 * sizeof(double2); }

unsigned grid1d(long long len) {
  long long g = (len + kBlk - 1) / kBlk;
  if (g > 2048) g = 2048;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// ---- specialised sizes ----------------------------------------------------
// (N, A, B) with N = A*B: the four-step split used by k_xyfast / k_zfast.
#define FAME_FAST_SIZES(X) \
  X(8, 2, 4) X(12, 3, 4) X(16, 4, 4) X(20, 4, 5) X(24, 4, 6) X(32, 4, 8) X(40, 5, 8) X(48, 6, 8) \
  X(60, 6, 10) X(64, 8, 8) X(72, 8, 9) X(80, 8, 10) X(90, 9, 10) X(96, 8, 12)


template <int N>
constexpr int xy_threads() { return N >= 72 ? 512 : 256; }

bool fast_size(int n) {
#define X(N, A, B) if (n == N) return true;
  FAME_FAST_SIZES(X)
#undef X
  return false;
}

bool use_fast_xy(const fame_plan* P) { return P->n1 == P->n2 && fast_size(P->n1); }
bool use_fast_z(const fame_plan* P) { return fast_size(P->n3); }
size_t xyfast_lds(int n) { return (size_t)(n * (n % 2 == 0 ? n + 1 : n) + n) * sizeof(double2); }

template <int N, int A, int B, int NT>
int xy_fast_launch(fame_plan* P, int dir, int twist, const double2* in, long long ibs, double2* out, long long obs,
                   int nf, double sc, const CgState* cgst) {
  const dim3 grid(P->n3, nf, P->cnb);
  const size_t lds = xyfast_lds(N);
  if (dir > 0) {
    if (twist == 1)
      hipLaunchKernelGGL((k_xyfast<N, A, B, 1, 1, NT>), grid, dim3(NT), lds, P->stream, P->geo, in, ibs, out, obs,
                         wtws(P), P->tab_len, P->d_tw, sc, cgst);
    else
      hipLaunchKernelGGL((k_xyfast<N, A, B, 1, 0, NT>), grid, dim3(NT), lds, P->stream, P->geo, in, ibs, out, obs,
                         wtws(P), P->tab_len, P->d_tw, sc, cgst);
  } else {
    if (twist == 2)
      hipLaunchKernelGGL((k_xyfast<N, A, B, -1, 2, NT>), grid, dim3(NT), lds, P->stream, P->geo, in, ibs, out, obs,
                         wtws(P), P->tab_len, P->d_tw, sc, cgst);
    else
      hipLaunchKernelGGL((k_xyfast<N, A, B, -1, 0, NT>), grid, dim3(NT), lds, P->stream, P->geo, in, ibs, out, obs,
                         wtws(P), P->tab_len, P->d_tw, sc, cgst);
  }
  CHECK_LAUNCH();
  return FAME_OK;
}

template <int N, int A, int B, int kZTP, int NT>
int z_fast_launch(fame_plan* P, int dir, int mode, const double2* in, long long ibs, double2* out, long long obs,
                  int nslot, int grp_base, double sc, const CgFuse* cf, const CgState* skip) {
  const dim3 grid((P->geo.n12 + kZTP - 1) / kZTP, nslot, P->cnb);
  const size_t lds = (size_t)(3 * N * kZTP + N) * sizeof(double2);
  const CgFuse c0{const_cast<CgState*>(skip), nullptr, nullptr, nullptr, 0};
  if (dir > 0) {
    if (mode == ZM_MIX_IN && cf)
      hipLaunchKernelGGL((k_zfast<N, A, B, 1, ZM_MIX_IN, kZTP, NT, 1>), grid, dim3(NT), lds, P->stream, P->geo, in,
                         ibs, out, obs, wlam(P), P->tab_len, P->d_tw, sc, grp_base, *cf);
    else if (mode == ZM_MIX_IN)
      hipLaunchKernelGGL((k_zfast<N, A, B, 1, ZM_MIX_IN, kZTP, NT>), grid, dim3(NT), lds, P->stream, P->geo, in, ibs,
                         out, obs, wlam(P), P->tab_len, P->d_tw, sc, grp_base, c0);
    else
      hipLaunchKernelGGL((k_zfast<N, A, B, 1, ZM_PLAIN, kZTP, NT>), grid, dim3(NT), lds, P->stream, P->geo, in, ibs,
                         out, obs, wlam(P), P->tab_len, P->d_tw, sc, grp_base, c0);
  } else {
    if (mode == ZM_MIX_OUT && cf)
      hipLaunchKernelGGL((k_zfast<N, A, B, -1, ZM_MIX_OUT, kZTP, NT, 1>), grid, dim3(NT), lds, P->stream, P->geo, in,
                         ibs, out, obs, wlam(P), P->tab_len, P->d_tw, sc, grp_base, *cf);
    else if (mode == ZM_MIX_OUT)
      hipLaunchKernelGGL((k_zfast<N, A, B, -1, ZM_MIX_OUT, kZTP, NT>), grid, dim3(NT), lds, P->stream, P->geo, in,
                         ibs, out, obs, wlam(P), P->tab_len, P->d_tw, sc, grp_base, c0);
    else
      hipLaunchKernelGGL((k_zfast<N, A, B, -1, ZM_PLAIN, kZTP, NT>), grid, dim3(NT), lds, P->stream, P->geo, in, ibs,
                         out, obs, wlam(P), P->tab_len, P->d_tw, sc, grp_base, c0);
  }
  CHECK_LAUNCH();
  return FAME_OK;
}

template <int N, int A, int B, int NT>
int xy_attrs() {
  const int lx = (int)xyfast_lds(N);
  HIP_TRY(hipFuncSetAttribute((const void*)k_xyfast<N, A, B, 1, 0, NT>, hipFuncAttributeMaxDynamicSharedMemorySize, lx));
  HIP_TRY(hipFuncSetAttribute((const void*)k_xyfast<N, A, B, 1, 1, NT>, hipFuncAttributeMaxDynamicSharedMemorySize, lx));
  HIP_TRY(hipFuncSetAttribute((const void*)k_xyfast<N, A, B, -1, 0, NT>, hipFuncAttributeMaxDynamicSharedMemorySize, lx));
  HIP_TRY(hipFuncSetAttribute((const void*)k_xyfast<N, A, B, -1, 2, NT>, hipFuncAttributeMaxDynamicSharedMemorySize, lx));
  return FAME_OK;
}

template <int N, int A, int B, int TP, int NT>
int z_attrs() {
  const int lz = (int)((3 * N * TP + N) * sizeof(double2));
  HIP_TRY(hipFuncSetAttribute((const void*)k_zfast<N, A, B, 1, ZM_MIX_IN, TP, NT>, hipFuncAttributeMaxDynamicSharedMemorySize, lz));
  HIP_TRY(hipFuncSetAttribute((const void*)k_zfast<N, A, B, 1, ZM_PLAIN, TP, NT>, hipFuncAttributeMaxDynamicSharedMemorySize, lz));
  HIP_TRY(hipFuncSetAttribute((const void*)k_zfast<N, A, B, -1, ZM_MIX_OUT, TP, NT>, hipFuncAttributeMaxDynamicSharedMemorySize, lz));
  HIP_TRY(hipFuncSetAttribute((const void*)k_zfast<N, A, B, -1, ZM_PLAIN, TP, NT>, hipFuncAttributeMaxDynamicSharedMemorySize, lz));
  HIP_TRY(hipFuncSetAttribute((const void*)k_zfast<N, A, B, 1, ZM_MIX_IN, TP, NT, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, lz));
  HIP_TRY(hipFuncSetAttribute((const void*)k_zfast<N, A, B, -1, ZM_MIX_OUT, TP, NT, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, lz));
  return FAME_OK;
}

template <int N, int A, int B>
int fast_set_attrs() {
  RET_IF((xy_attrs<N, A, B, xy_threads<N>()>()));
  RET_IF((xy_attrs<N, A, B, 512>()));
  if constexpr (N >= 64) RET_IF((xy_attrs<N, A, B, 1024>()));
  RET_IF((z_attrs<N, A, B, 16, 256>()));
  RET_IF((z_attrs<N, A, B, 8, 128>()));
  RET_IF((z_attrs<N, A, B, 32, 512>()));
  RET_IF((z_attrs<N, A, B, 16, 320>()));
  RET_IF((z_attrs<N, A, B, 16, 384>()));
  return FAME_OK;
}

int set_fast_attrs(int n) {
#define X(N, A, B) if (n == N) return fast_set_attrs<N, A, B>();
  FAME_FAST_SIZES(X)
#undef X
  return FAME_OK;
}

// ---- launch helpers -------------------------------------------------------
// cf (fast z kernels only): fuse the CG updates into the MIX_IN / MIX_OUT pass; skip (fast
// kernels, without cf): items whose state is inactive are left alone
int zpass(fame_plan* P, int dir, int mode, const double2* in, long long ibs, double2* out, long long obs, int nslot,
          int grp_base, const CgFuse* cf = nullptr, const CgState* skip = nullptr) {
  const double sc = 1.0 / std::sqrt((double)P->n3);
  if (P->fast && use_fast_z(P)) {
#define X(N, A, B)                                                                                       \
    if (P->n3 == N) {                                                                                      \
      if (P->zv == 1) return z_fast_launch<N, A, B, 8, 128>(P, dir, mode, in, ibs, out, obs, nslot, grp_base, sc, cf, skip); \
      if (P->zv == 2) return z_fast_launch<N, A, B, 32, 512>(P, dir, mode, in, ibs, out, obs, nslot, grp_base, sc, cf, skip); \
      if (P->zv == 3) return z_fast_launch<N, A, B, 16, 320>(P, dir, mode, in, ibs, out, obs, nslot, grp_base, sc, cf, skip); \
      if (P->zv == 4) return z_fast_launch<N, A, B, 16, 384>(P, dir, mode, in, ibs, out, obs, nslot, grp_base, sc, cf, skip); \
      return z_fast_launch<N, A, B, 16, 256>(P, dir, mode, in, ibs, out, obs, nslot, grp_base, sc, cf, skip);        \
    }
    FAME_FAST_SIZES(X)
#undef X
  }
  if (cf) return set_err(FAME_ERR_UNSUPPORTED, "fused CG needs the specialised z kernels");
  const dim3 grid((P->geo.n12 + kZTile - 1) / kZTile, nslot, P->cnb);
  const size_t lds = zpass_lds(P);
  if (dir > 0) {
    if (mode == ZM_MIX_IN)
      hipLaunchKernelGGL((k_zpass<1, ZM_MIX_IN>), grid, dim3(kBlk), lds, P->stream, P->geo, in, ibs, out, obs,
                         wlam(P), P->tab_len, P->d_tw, P->d_perm, sc, grp_base);
    else
      hipLaunchKernelGGL((k_zpass<1, ZM_PLAIN>), grid, dim3(kBlk), lds, P->stream, P->geo, in, ibs, out, obs,
                         wlam(P), P->tab_len, P->d_tw, P->d_perm, sc, grp_base);
  } else {
    if (mode == ZM_MIX_OUT)
      hipLaunchKernelGGL((k_zpass<-1, ZM_MIX_OUT>), grid, dim3(kBlk), lds, P->stream, P->geo, in, ibs, out, obs,
                         wlam(P), P->tab_len, P->d_tw, P->d_perm, sc, grp_base);
    else
      hipLaunchKernelGGL((k_zpass<-1, ZM_PLAIN>), grid, dim3(kBlk), lds, P->stream, P->geo, in, ibs, out, obs,
                         wlam(P), P->tab_len, P->d_tw, P->d_perm, sc, grp_base);
  }
  CHECK_LAUNCH();
  return FAME_OK;
}

// cgst (fast kernels only, optional): inside a fused CG, items whose state is inactive are skipped
int xypass(fame_plan* P, int dir, int twist, const double2* in, long long ibs, double2* out, long long obs, int nf,
           const CgState* cgst = nullptr) {
  const double sc = 1.0 / std::sqrt((double)P->n1 * (double)P->n2);
  if (P->fast && use_fast_xy(P)) {
#define X(N, A, B)                                                                                    \
    if (P->n1 == N) {                                                                                   \
      if (P->xyv == 1 || (P->xyv == 2 && N < 64)) return xy_fast_launch<N, A, B, 512>(P, dir, twist, in, ibs, out, obs, nf, sc, cgst); \
      if (P->xyv == 2) return xy_fast_launch<N, A, B, (N >= 64 ? 1024 : 512)>(P, dir, twist, in, ibs, out, obs, nf, sc, cgst); \
      return xy_fast_launch<N, A, B, xy_threads<N>()>(P, dir, twist, in, ibs, out, obs, nf, sc, cgst);      \
    }
    FAME_FAST_SIZES(X)
#undef X
  }
  const dim3 grid(P->n3, nf, P->cnb);
  const size_t lds = xypass_lds(P);
#define XY_LAUNCH(D, T)                                                                                   \
  hipLaunchKernelGGL((k_xypass<D, T>), grid, dim3(kBlk), lds, P->stream, P->geo, in, ibs, out, obs, wtws(P), \
                     P->tab_len, P->d_tw, P->d_perm, sc)
  if (dir > 0) {
    if (twist == 1) XY_LAUNCH(1, 1); else XY_LAUNCH(1, 0);
  } else {
    if (twist == 2) XY_LAUNCH(-1, 2); else XY_LAUNCH(-1, 0);
  }
#undef XY_LAUNCH
  CHECK_LAUNCH();
  return FAME_OK;
}

// Slab length of the 2.5-D material kernel: as few slabs per tile column as fill one round of
// work-groups (tiles * nbc * nzc <= three 256-thread work-groups per compute unit, what fits at
// the kernel's registers and LDS), since every slab re-reads one slice below and one above it;
// at least one slab, slabs of at least three slices.
int slab_length(int n1, int n2, int n3, int TX, int TY, long long nbc, int ncu) {
  const long long cols = (long long)((n1 + TX - 1) / TX) * ((n2 + TY - 1) / TY) * nbc;
  long long nzc = std::max<long long>(1, 3LL * ncu / cols);
  nzc = std::min<long long>(nzc, std::max(1, n3 / 3));
  return (int)((n3 + nzc - 1) / nzc);
}

// Tile of the 2.5-D material kernel: maximise (useful columns) x (tile / tile-with-halo) under
// (TX+2)(TY+2) <= kMzThreads; the slab length as above.
void material_tile(int n1, int n2, int n3, int nbc, int ncu, int& TX, int& TY, int& TZ) {
  double best = -1;
  TX = TY = 1;
  for (int tx = 1; tx <= std::min(n1, kMzThreads); ++tx)
    for (int ty = 1; ty <= std::min(n2, kMzThreads); ++ty) {
      if ((tx + 2) * (ty + 2) > kMzThreads) break;
      const double pad = (double)n1 * n2 / ((double)((n1 + tx - 1) / tx * tx) * ((n2 + ty - 1) / ty * ty));
      const double eff = pad * tx * ty / (double)((tx + 2) * (ty + 2));
      if (eff > best + 1e-12) {
        best = eff;
        TX = tx;
        TY = ty;
      }
    }
  TZ = slab_length(n1, n2, n3, TX, TY, nbc, ncu);
}

int pv_material(fame_plan* P, const double2* in, long long ibs, double2* out, long long obs, double2* dpart,
                int pstride);

// tile, slab and grid of the 2.5-D material kernel for a launch of cnb items
void material_shape(const fame_plan* P, int cnb, int& TX, int& TY, int& TZ, int& ntile, int& nzc) {
  material_tile(P->n1, P->n2, P->n3, cnb, P->ncu, TX, TY, TZ);
  if (P->mt[0] > 0) {  // tile chosen by the autotuner or the caller
    TX = P->mt[0];
    TY = P->mt[1];
    if (P->mt[2] > 0) TZ = P->mt[2];
  }
  ntile = ((P->n1 + TX - 1) / TX) * ((P->n2 + TY - 1) / TY);
  nzc = (P->n3 + TZ - 1) / TZ;
}

// number of per-work-group partial sums the material pass leaves per item (dpart below)
int material_parts(const fame_plan* P, int cnb) {
  if (P->pv && !P->fast) return (int)grid1d(P->N);
  int TX, TY, TZ, ntile, nzc;
  material_shape(P, cnb, TX, TY, TZ, ntile, nzc);
  return ntile * nzc;
}

// dpart (specialised and per-vertex stencils): per-work-group sums of conj(in) . out, item
// stride pstride
// cgst (specialised stencil only, optional): inside a fused CG, items whose state is inactive are skipped
int material(fame_plan* P, const double2* in, long long ibs, double2* out, long long obs, double2* dpart = nullptr,
             int pstride = 0, const CgState* cgst = nullptr) {
  if (P->pv && !P->fast) return pv_material(P, in, ibs, out, obs, dpart, pstride);
  if (P->fast) {
    int TX, TY, TZ, ntile, nzc;
    material_shape(P, P->cnb, TX, TY, TZ, ntile, nzc);
    if (dpart && ntile * nzc > pstride) return set_err(FAME_ERR_ARG, "material: partial-sum stride too small");
    const dim3 grid(ntile, nzc, P->cnb);
    // profiling: the launch's own start / stop events time the kernel alone (as a kernel trace
    // does), without the queue gaps that events recorded between launches include
    const double* nulld = nullptr;
#define MZ_GO(PV_, EPS_, GAM_, COEF_)                                                                          \
  do {                                                                                                         \
    const std::uint32_t lds_ = (std::uint32_t)material_z_lds(TX, TY, PV_);                                   \
    if (P->prof)                                                                                               \
      hipExtLaunchKernelGGL(k_material_z<PV_>, grid, dim3(kMzThreads), lds_, P->stream, P->evk[0], P->evk[1], \
                            0u, P->geo, in, ibs, out, obs, (const unsigned char*)P->d_mat_id,                  \
                            (const double2*)P->d_mat, P->n_mat, EPS_, GAM_, COEF_, (const double2*)wom(P), TX, \
                            TY, TZ, dpart, pstride, cgst);                                                     \
    else                                                                                                       \
      hipLaunchKernelGGL(k_material_z<PV_>, grid, dim3(kMzThreads), lds_, P->stream, P->geo, in, ibs, out, obs, \
                         (const unsigned char*)P->d_mat_id, (const double2*)P->d_mat, P->n_mat, EPS_, GAM_,    \
                         COEF_, (const double2*)wom(P), TX, TY, TZ, dpart, pstride, cgst);                     \
  } while (0)
    if (P->pv) {
      PvCoef c;
      for (int i = 0; i < 9; ++i) c.G[i] = P->pv_G[i];
      c.mu = P->pv_mu;
      MZ_GO(true, (const double*)P->d_eps, (const double*)P->d_gam, c);
    } else {
      MZ_GO(false, nulld, nulld, PvCoef{});
    }
#undef MZ_GO
    P->evk_used = P->prof;
    CHECK_LAUNCH();
    return FAME_OK;
  }
  if (dpart) return set_err(FAME_ERR_ARG, "material: partial sums need the specialised stencil");
  const int tx = (P->n1 + kMatTX - 1) / kMatTX, ty = (P->n2 + kMatTY - 1) / kMatTY;
  const int tz = (P->n3 + kMatTZ - 1) / kMatTZ;
  hipLaunchKernelGGL(k_material, dim3(tx * ty, tz, P->cnb), dim3(kBlk), 0, P->stream, P->geo, in, ibs, out, obs,
                     P->d_mat_id, P->d_mat, P->n_mat, wom(P));
  CHECK_LAUNCH();
  return FAME_OK;
}

int twist(fame_plan* P, const double2* in, double2* out, int ncomp, int conj_twist) {
  hipLaunchKernelGGL(k_twist, dim3(grid1d(P->N), P->cnb), dim3(kBlk), 0, P->stream, P->geo, in, out, ncomp,
                     wtws(P), P->tab_len, conj_twist);
  CHECK_LAUNCH();
  return FAME_OK;
}

// One fncB over batch items [b0, b0 + nbc): z-pass + Pi mix, xy-pass, material, xy-pass, z-pass + Pi^H.
// With cf (fused CG), x is the CG residual r and the transformed vector is p <- r + beta p.
// mpart (optional, without cf): the material pass leaves its per-work-group partial sums of
// u^H (Mat u) there, item stride mpstride (for an isometric transform chain this is x^H fncB x)
int fncB_window(fame_plan* P, int b0, int nbc, const double2* x, double2* y, const CgFuse* cf = nullptr,
                double2* mpart = nullptr, int mpstride = 0, const CgState* skip = nullptr) {
  const long long N = P->N;
  struct Window {  // restores the full-batch window on every exit path
    fame_plan* P;
    ~Window() { P->cb0 = 0; P->cnb = P->vnb; }
  } guard{P};
  P->cb0 = b0;
  P->cnb = nbc;
  x += (size_t)b0 * 4 * N;
  y += (size_t)b0 * 4 * N;
  CgFuse wcf{};
  const CgFuse* pcf = nullptr;
  if (cf) {
    wcf = *cf;
    wcf.st = cf->st + b0;
    wcf.x = cf->x + (size_t)b0 * 4 * N;
    wcf.p = cf->p + (size_t)b0 * 4 * N;
    wcf.part = cf->part + (size_t)b0 * cf->nparts;
    wcf.part_r = cf->part_r + (size_t)b0 * cf->nparts;
    wcf.nparts_m = material_parts(P, nbc);
    pcf = &wcf;
  }
  const bool pr = P->prof;
  if (pr) HIP_TRY(hipEventRecord(P->ev[0], P->stream));
  // CG: the material pass sums p^H q (in real space), the MIX_OUT pass forms alpha from those
  // sums and updates r in place (y is r then)
  double2* w1 = P->d_w1 + (size_t)P->w_off * 6 * N;
  double2* w2 = P->d_w2 + (size_t)P->w_off * 6 * N;
  // items a caller's iteration has finished (skip, without cf) are left alone by every pass
  const CgState* wskip = (!pcf && skip) ? skip + b0 : nullptr;
  RET_IF(zpass(P, +1, ZM_MIX_IN, x, 4 * N, w1, 6 * N, 2, 0, pcf, wskip));
  if (pr) HIP_TRY(hipEventRecord(P->ev[1], P->stream));
  // converged items of a fused CG are skipped by every pass (their state is frozen)
  const CgState* cgst = pcf ? wcf.st : wskip;
  // the xy passes run in place (each work-group loads its whole plane before its first barrier
  // and stores it after its last): a chunk's fields then occupy two buffers that are each read
  // and rewritten on the same lines, which keeps more of them in the Infinity Cache between
  // passes (1.2 % faster fncB at n = 60, 40 k, same results bit for bit)
  RET_IF(xypass(P, +1, 0, w1, 6 * N, w1, 6 * N, 6, cgst));
  if (pr) HIP_TRY(hipEventRecord(P->ev[2], P->stream));
  double2* mp = pcf ? wcf.part : (mpart ? mpart + (size_t)b0 * mpstride : nullptr);
  RET_IF(material(P, w1, 6 * N, w2, 6 * N, mp, pcf ? wcf.nparts : mpstride, cgst));
  if (pr) HIP_TRY(hipEventRecord(P->ev[3], P->stream));
  RET_IF(xypass(P, -1, 0, w2, 6 * N, w2, 6 * N, 6, cgst));
  if (pr) HIP_TRY(hipEventRecord(P->ev[4], P->stream));
  RET_IF(zpass(P, -1, ZM_MIX_OUT, w2, 6 * N, y, 4 * N, 2, 0, pcf, wskip));
  if (pr) {
    HIP_TRY(hipEventRecord(P->ev[5], P->stream));
    HIP_TRY(hipEventSynchronize(P->ev[5]));
    if (P->evk_used) {
      float km = 0;
      HIP_TRY(hipEventElapsedTime(&km, P->evk[0], P->evk[1]));
      P->mat_kernel_ms += km;
      P->mat_launches += 1;
      P->evk_used = false;
    }
    for (int s = 0; s < 5; ++s) {
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, P->ev[s], P->ev[s + 1]));
      P->stage_ms[s] += ms;
    }
  }
  return FAME_OK;
}

// fncB on all nb items, in chunks so the intermediate fields of a chunk can stay
// in the device's last-level cache between the five kernels.
// Multi-stream chunk schedule: chunk k runs on the plan's stream (k % nstreams == 0) or on
// extra stream k % nstreams - 1, with fields in scratch items [(k % nstreams) c, ...); all
// streams start after the work queued so far (fork) and the caller's stream waits for them at
// the end (join).  Chunks touch disjoint items, so the only ordering they need is each chunk's
// own, which its stream keeps.  Off while profiling (the per-kernel events assume one stream)
// and for a single chunk.
struct ChunkStreams {
  fame_plan* P;
  hipStream_t main;
  int ns = 1;
  explicit ChunkStreams(fame_plan* p) : P(p), main(p->stream) {}
  int begin(int nchunks) {
    ns = (P->nstreams > 1 && !P->prof && nchunks > 1) ? std::min(P->nstreams, nchunks) : 1;
    if (ns == 1) return FAME_OK;
    HIP_TRY(hipEventRecord(P->ev_fork, main));
    for (int i = 0; i + 1 < ns; ++i) HIP_TRY(hipStreamWaitEvent(P->xstream[i], P->ev_fork, 0));
    return FAME_OK;
  }
  void select(int k, int c) {  // before chunk k (chunk size c) is queued
    if (ns == 1) return;
    const int j = k % ns;
    P->stream = j ? P->xstream[j - 1] : main;
    P->w_off = j * c;
  }
  int end() {  // restores the plan's stream; the caller's stream then waits for the others
    P->stream = main;
    P->w_off = 0;
    if (ns == 1) return FAME_OK;
    const int n = ns;
    ns = 1;
    for (int i = 0; i + 1 < n; ++i) {
      HIP_TRY(hipEventRecord(P->ev_join[i], P->xstream[i]));
      HIP_TRY(hipStreamWaitEvent(main, P->ev_join[i], 0));
    }
    return FAME_OK;
  }
  ~ChunkStreams() { (void)end(); }
};

// skip (optional, [vnb]): items whose .active is 0 are left alone (specialised kernels)
int do_fncB(fame_plan* P, const double2* x, double2* y, double2* mpart = nullptr, int mpstride = 0,
            const CgState* skip = nullptr) {
  const int nb = P->vnb;
  const int c = (P->chunk > 0 && P->chunk < nb) ? P->chunk : nb;
  ChunkStreams ps(P);
  RET_IF(ps.begin((nb + c - 1) / c));
  for (int b0 = 0, k = 0; b0 < nb; b0 += c, ++k) {
    ps.select(k, c);
    RET_IF(fncB_window(P, b0, std::min(c, nb - b0), x, y, nullptr, mpart, mpstride, skip));
  }
  RET_IF(ps.end());
  if (P->prof) P->prof_calls += 1;
  return FAME_OK;
}

// rows of `len` complex words: dst[j] = src[idx[j]] for j < m (blockIdx.y = j)
__global__ void k_gather_rows(const double2* __restrict__ src, double2* __restrict__ dst, const int* __restrict__ idx,
                              long long len, int scatter) {
  const int j = blockIdx.y;
  const long long si = scatter ? j : idx[j], di = scatter ? idx[j] : j;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len; i += (long long)gridDim.x * blockDim.x)
    dst[di * len + i] = src[si * len + i];
}

void reset_view(fame_plan* P) {
  P->vnb = P->nb;
  P->cnb = P->nb;
  P->cb0 = 0;
  P->v_lam = P->d_lam;
  P->v_tws = P->d_tws;
  P->v_om = P->d_omega;
}

// Operators see only `items` (ascending item indices) from now on, packed 0..m-1.
int set_view(fame_plan* P, const std::vector<int>& items) {
  const int m = (int)items.size();
  if (m < 1 || m > P->nb) return set_err(FAME_ERR_ARG, "set_view: bad item count");
  if (m == P->nb) {
    reset_view(P);
    return FAME_OK;
  }
  HIP_TRY(hipMemcpyAsync(P->d_idx, items.data(), m * sizeof(int), hipMemcpyHostToDevice, P->stream));
  hipLaunchKernelGGL(k_gather_rows, dim3(1, m), dim3(64), 0, P->stream, P->d_lam, P->d_lam_v, P->d_idx,
                     (long long)P->tab_len, 0);
  hipLaunchKernelGGL(k_gather_rows, dim3(1, m), dim3(64), 0, P->stream, P->d_tws, P->d_tws_v, P->d_idx,
                     (long long)P->tab_len, 0);
  hipLaunchKernelGGL(k_gather_rows, dim3(1, m), dim3(64), 0, P->stream, P->d_omega, P->d_om_v, P->d_idx, 3LL, 0);
  CHECK_LAUNCH();
  HIP_TRY(hipStreamSynchronize(P->stream));  // items' host buffer is reused by the caller
  P->vnb = m;
  P->cnb = m;
  P->v_lam = P->d_lam_v;
  P->v_tws = P->d_tws_v;
  P->v_om = P->d_om_v;
  return FAME_OK;
}

// pack (scatter = 0: dst[j] = src[items[j]]) or unpack (scatter = 1) vectors of `len` words
int pack_rows(fame_plan* P, const double2* src, double2* dst, long long len, int m, int scatter) {
  hipLaunchKernelGGL(k_gather_rows, dim3(grid1d(len), m), dim3(kBlk), 0, P->stream, src, dst, P->d_idx, len, scatter);
  CHECK_LAUNCH();
  return FAME_OK;
}

// act (optional, [vnb]): items with act[b] < 0 are skipped
int do_fncA(fame_plan* P, const double2* x, double2* y, const int* act = nullptr) {
  hipLaunchKernelGGL(k_fncA, dim3(grid1d(P->N), P->vnb), dim3(kBlk), 0, P->stream, P->geo, x, y, P->v_lam,
                     P->tab_len, act);
  CHECK_LAUNCH();
  return FAME_OK;
}

// out[b] = x[b]^H y[b] (device, P->d_red or given); items with act[b] < 0 (act optional) get 0
int do_dot(fame_plan* P, const double2* x, long long xbs, const double2* y, long long ybs, long long len, int nb,
           double2* out, const int* act = nullptr) {
  const int nblk = std::min<int>(P->nblk_red, (int)grid1d(len));
  hipLaunchKernelGGL(k_dot_part, dim3(nblk, nb), dim3(kBlk), 0, P->stream, x, xbs, y, ybs, len, P->d_part, act);
  CHECK_LAUNCH();
  hipLaunchKernelGGL(k_reduce_part, dim3(nb), dim3(kBlk), 0, P->stream, P->d_part, nblk, out, act);
  CHECK_LAUNCH();
  return FAME_OK;
}

int fill_zero(fame_plan* P, double2* y, long long len) {
  hipLaunchKernelGGL(k_fill_zero, dim3(grid1d(len)), dim3(kBlk), 0, P->stream, y, len);
  CHECK_LAUNCH();
  return FAME_OK;
}

int copy(fame_plan* P, const double2* x, double2* y, long long len) {
  hipLaunchKernelGGL(k_copy, dim3(grid1d(len)), dim3(kBlk), 0, P->stream, x, y, len);
  CHECK_LAUNCH();
  return FAME_OK;
}

// dst (host) <- src (device), `bytes` of them, through the plan's pinned staging buffer; waits
// for the plan's stream
int read_to_host(fame_plan* P, void* dst, const void* src, size_t bytes) {
  if (bytes > P->h_stage_cap) {
    if (P->h_stage) HIP_TRY(hipHostFree(P->h_stage));
    P->h_stage = nullptr;
    P->h_stage_cap = 0;
    const size_t cap = std::max<size_t>(bytes, 16384);
    HIP_TRY(hipHostMalloc(&P->h_stage, cap, hipHostMallocDefault));
    P->h_stage_cap = cap;
  }
  HIP_TRY(hipMemcpyAsync(P->h_stage, src, bytes, hipMemcpyDeviceToHost, P->stream));
  HIP_TRY(hipStreamSynchronize(P->stream));
  std::memcpy(dst, P->h_stage, bytes);
  return FAME_OK;
}

int cg_finish(fame_plan* P, std::vector<CgState>& hs, int* iters, double* relres) {
  const int nb = P->vnb;
  RET_IF(read_to_host(P, hs.data(), P->d_cg, nb * sizeof(CgState)));
  bool all = true;
  for (int i = 0; i < nb; ++i) {
    if (iters) iters[i] = (int)hs[i].iters;
    if (relres) relres[i] = hs[i].bnorm2 > 0 ? std::sqrt(hs[i].rho / hs[i].bnorm2) : 0.0;
    if (hs[i].active != 0.0) all = false;
  }
  if (!all) return set_err(FAME_ERR_NOT_CONVERGED, "fncinvB: CG reached maxit before tol");
  return FAME_OK;
}

int zfast_tp(const fame_plan* P) { return P->zv == 1 ? 8 : (P->zv == 2 ? 32 : 16); }  // zv 3, 4: 16 pencils

// CG with its vector updates and scalar steps fused into fncB (same iterates as the unfused
// loop below up to rounding): per chunk of items, fncB(p) where the first pass forms beta from
// the previous residual's partial sums, tests convergence, and does p <- r + beta p and the
// deferred x <- x + alpha p; the material pass sums p^H q; the last pass forms alpha and does
// r <- r - alpha q with the partials of ||r||^2 -- five kernels per chunk-iteration, while the
// chunk's vectors are recent in the last-level cache.  After the loop k_cg_fin2 tests the last
// residual and x receives its last alpha p.
int do_cg_fused(fame_plan* P, const double2* b, double2* x, double tol, int maxit, int* iters, double* relres) {
  const long long L = 4 * P->N;
  const long long tot = L * P->vnb;
  const int nb = P->vnb;
  // per-item partial sums: p^H q from the material pass (one per work-group of a launch of nbc
  // items) in the first area, ||r||^2 from the MIX_OUT pass (one per z work-group and slot) in
  // the second; the per-item stride is the largest count of any launch
  const int nparts_z = 2 * ((P->geo.n12 + zfast_tp(P) - 1) / zfast_tp(P));
  int nparts = nparts_z;
  for (int c = 1; c <= nb; ++c) nparts = std::max(nparts, material_parts(P, c));
  if (P->part2_cap < (size_t)2 * P->nb * nparts) {
    if (P->d_part2) HIP_TRY(hipFree(P->d_part2));
    P->d_part2 = nullptr;
    P->part2_cap = 0;
    HIP_TRY(hipMalloc(&P->d_part2, (size_t)2 * P->nb * nparts * sizeof(double2)));
    P->part2_cap = (size_t)2 * P->nb * nparts;
  }
  double2* part_r = P->d_part2 + (size_t)P->nb * nparts;
  RET_IF(fill_zero(P, x, tot));
  RET_IF(copy(P, b, P->d_r, tot));
  RET_IF(fill_zero(P, P->d_p, tot));
  RET_IF(do_dot(P, b, L, b, L, L, nb, P->d_red));
  hipLaunchKernelGGL(k_cg_init, dim3((nb + 63) / 64), dim3(64), 0, P->stream, P->d_red, P->d_cg, nb, tol);
  CHECK_LAUNCH();
  std::vector<CgState> hs(nb);
  // iteration `it` on the chunk [b0, b0 + nbc) (the output of fncB is r itself)
  auto step = [&](int b0, int nbc, int it) -> int {
    const CgFuse cf{P->d_cg, x, P->d_p, P->d_part2, nparts, part_r, 0, it & 1, it == 0 ? 1 : 0};
    return fncB_window(P, b0, nbc, P->d_r, P->d_r, &cf);
  };
  // all items of [b0, b0 + nbc) converged?
  auto finished = [&](int b0, int nbc, bool& done) -> int {
    RET_IF(read_to_host(P, hs.data() + b0, P->d_cg + b0, nbc * sizeof(CgState)));
    done = true;
    for (int i = b0; i < b0 + nbc; ++i)
      if (hs[i].active != 0.0) done = false;
    return FAME_OK;
  };
  const int c = (P->chunk > 0 && P->chunk < nb) ? P->chunk : nb;
  if (P->cg_graph && !P->prof && P->cg_seq == 0) {
    // iteration 0 eagerly, then iterations 1 and 2 over all chunks captured once as a HIP graph
    // on the plan's own stream (odd parity, then even) and replayed on the plan's stream: the
    // pointers and parities do not change within the solve
    int rc = FAME_OK;
    for (int b0 = 0; b0 < nb && rc == FAME_OK; b0 += c) rc = step(b0, std::min(c, nb - b0), 0);
    RET_IF(rc);
    int it = 1;
    bool done = false;
    if (maxit >= 3) {
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      hipStream_t saved = P->stream;
      HIP_TRY(hipStreamBeginCapture(P->own_stream, hipStreamCaptureModeThreadLocal));
      P->stream = P->own_stream;
      for (int k = 1; k <= 2 && rc == FAME_OK; ++k)
        for (int b0 = 0; b0 < nb && rc == FAME_OK; b0 += c) rc = step(b0, std::min(c, nb - b0), k);
      P->stream = saved;
      hipError_t ec = hipStreamEndCapture(P->own_stream, &graph);
      if (rc != FAME_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (ec != hipSuccess) return set_err(FAME_ERR_HIP, std::string("CG graph capture: ") + hipGetErrorString(ec));
      ec = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (ec != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return set_err(FAME_ERR_HIP, std::string("CG graph instantiate: ") + hipGetErrorString(ec));
      }
      while (it + 2 <= maxit && !done && rc == FAME_OK) {
        ec = hipGraphLaunch(exec, P->stream);
        if (ec != hipSuccess) rc = set_err(FAME_ERR_HIP, std::string("CG graph launch: ") + hipGetErrorString(ec));
        it += 2;
        if (rc == FAME_OK && ((it - 1) % 4 == 0 || it + 2 > maxit)) rc = finished(0, nb, done);
      }
      (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
      RET_IF(rc);
    }
    if (!done && it < maxit)  // an odd remainder of maxit
      for (int b0 = 0; b0 < nb; b0 += c) RET_IF(step(b0, std::min(c, nb - b0), it));
  } else if (P->cg_seq > 0) {
    // chunk by chunk to convergence: a chunk's vectors and fncB intermediates stay in the
    // last-level cache for all its iterations
    const int cs = std::min(P->cg_seq, nb);
    for (int b0 = 0; b0 < nb; b0 += cs) {
      const int nbc = std::min(cs, nb - b0);
      bool done = false;
      for (int it = 0; it < maxit && !done;) {
        RET_IF(step(b0, nbc, it));
        ++it;
        if (it % 4 == 0 || it == maxit) RET_IF(finished(b0, nbc, done));
      }
    }
  } else {
    // all items in lock step, chunk by chunk within an iteration.  The convergence poll runs
    // one batch of 4 iterations behind: after each batch the state goes to pinned host memory
    // and is read once the next batch is queued, so the device never waits for the host.
    // Iterations queued after the last item converged find every item inactive and return
    // at once (converged items are skipped by every pass).  Once every unconverged item's
    // ||r||^2 is within 1e4 of its threshold the batches shrink to one iteration, so fewer such
    // empty iterations are queued; results do not depend on the batch sizes (a converged item's
    // state is frozen on the device).
    if (P->h_cg_cap < nb) {
      if (P->h_cg) HIP_TRY(hipHostFree(P->h_cg));
      P->h_cg = nullptr;
      P->h_cg_cap = 0;
      HIP_TRY(hipHostMalloc((void**)&P->h_cg, (size_t)nb * sizeof(CgState), hipHostMallocDefault));
      P->h_cg_cap = nb;
    }
    if (!P->ev_poll) HIP_TRY(hipEventCreateWithFlags(&P->ev_poll, hipEventDisableTiming));
    bool done = false, pending = false;
    int bsize = 4;
    for (int it = 0; it < maxit && !done;) {
      const int batch = std::min(bsize, maxit - it);
      {
        // a chunk's iterations depend only on that chunk: with two streams each keeps its chunks
        // for the whole batch, joined before the state copy below
        ChunkStreams ps(P);
        RET_IF(ps.begin((nb + c - 1) / c));
        for (int k = 0; k < batch; ++k) {
          for (int b0 = 0, ck = 0; b0 < nb; b0 += c, ++ck) {
            ps.select(ck, c);
            RET_IF(step(b0, std::min(c, nb - b0), it + k));
          }
          if (P->prof) P->prof_calls += 1;
        }
        RET_IF(ps.end());
      }
      it += batch;
      if (pending) {  // the state after the previous batch
        HIP_TRY(hipEventSynchronize(P->ev_poll));
        done = true;
        bool near = true;
        for (int i = 0; i < nb; ++i) {
          const CgState& s = P->h_cg[i];
          if (s.active == 0.0) continue;
          done = false;
          // rho_i sits in the slot of i's parity, and iters = i after MIX_IN pass i
          const double rn = ((long long)s.iters & 1) ? s.rho1 : s.rho;
          if (!(rn <= 1e4 * s.tol2 * s.bnorm2)) near = false;
        }
        if (near) bsize = 1;
      }
      if (!done) {
        HIP_TRY(hipMemcpyAsync(P->h_cg, P->d_cg, (size_t)nb * sizeof(CgState), hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipEventRecord(P->ev_poll, P->stream));
        pending = true;
      }
    }
  }
  hipLaunchKernelGGL(k_cg_fin2, dim3(nb), dim3(kBlk), 0, P->stream, part_r, nparts_z, nparts, P->d_cg);
  CHECK_LAUNCH();
  hipLaunchKernelGGL(k_cg_xfin, dim3(grid1d(L), nb), dim3(kBlk), 0, P->stream, x, P->d_p, L, P->d_cg);
  CHECK_LAUNCH();
  return cg_finish(P, hs, iters, relres);
}

int do_cg_plain(fame_plan* P, const double2* b, double2* x, double tol, int maxit, int* iters, double* relres);

int do_cg(fame_plan* P, const double2* b, double2* x, double tol, int maxit, int* iters, double* relres) {
  if (P->cg_fused && P->fast && use_fast_z(P)) return do_cg_fused(P, b, x, tol, maxit, iters, relres);
  return do_cg_plain(P, b, x, tol, maxit, iters, relres);
}

// CG as in MATLAB pcg without preconditioner: separate dot, update and direction kernels.
int do_cg_plain(fame_plan* P, const double2* b, double2* x, double tol, int maxit, int* iters, double* relres) {
  const long long L = 4 * P->N;
  const long long tot = L * P->vnb;
  const int nb = P->vnb;
  RET_IF(fill_zero(P, x, tot));
  RET_IF(copy(P, b, P->d_r, tot));
  RET_IF(copy(P, b, P->d_p, tot));
  RET_IF(do_dot(P, b, L, b, L, L, nb, P->d_red));
  hipLaunchKernelGGL(k_cg_init, dim3((nb + 63) / 64), dim3(64), 0, P->stream, P->d_red, P->d_cg, nb, tol);
  CHECK_LAUNCH();
  const int nblk = std::min<int>(P->nblk_red, (int)grid1d(L));
  std::vector<CgState> hs(nb);
  int it = 0;
  bool done = false;
  const int check_every = 4;
  while (it < maxit && !done) {
    RET_IF(do_fncB(P, P->d_p, P->d_q));
    RET_IF(do_dot(P, P->d_p, L, P->d_q, L, L, nb, P->d_red));
    hipLaunchKernelGGL(k_cg_alpha, dim3((nb + 63) / 64), dim3(64), 0, P->stream, P->d_red, P->d_cg, nb);
    hipLaunchKernelGGL(k_cg_update, dim3(nblk, nb), dim3(kBlk), 0, P->stream, x, P->d_r, P->d_p, P->d_q, L, P->d_cg,
                       P->d_part);
    hipLaunchKernelGGL(k_cg_beta, dim3(nb), dim3(kBlk), 0, P->stream, P->d_part, nblk, nblk, P->d_cg);
    hipLaunchKernelGGL(k_cg_p, dim3(grid1d(L), nb), dim3(kBlk), 0, P->stream, P->d_p, P->d_r, L, P->d_cg);
    CHECK_LAUNCH();
    ++it;
    if (it % check_every == 0 || it == maxit) {
      RET_IF(read_to_host(P, hs.data(), P->d_cg, nb * sizeof(CgState)));
      done = true;
      for (int i = 0; i < nb; ++i)
        if (hs[i].active != 0.0) done = false;
    }
  }
  return cg_finish(P, hs, iters, relres);
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char* fame_version(void) { return "fame-mi355x 0.1 (gfx950, FP64)"; }

const char* fame_last_error(void) { return g_err.c_str(); }

int fame_plan_create(fame_plan** plan, int n1, int n2, int n3, int nb, const double* theta, const double* delta,
                     int n_mat, const double* mat_table, const unsigned char* mat_id) {
  if (!plan) return set_err(FAME_ERR_ARG, "plan pointer is NULL");
  *plan = nullptr;
  if (n1 < 1 || n2 < 1 || n3 < 1 || nb < 1) return set_err(FAME_ERR_ARG, "grid dims and nb must be >= 1");
  if (!theta || !delta) return set_err(FAME_ERR_ARG, "theta and delta are required");
  RET_IF(check_theta(n1, n2, n3, nb, theta, delta));
  if (n_mat < 1 || n_mat > kMaxMat || !mat_table)
    return set_err(FAME_ERR_ARG, "need 1 <= n_mat <= " + std::to_string(kMaxMat) + " and a material table");
  if ((long long)n1 * n2 * (long long)sizeof(double2) > 160 * 1024)
    return set_err(FAME_ERR_UNSUPPORTED, "n1*n2 too large for one xy-plane in LDS");
  if ((size_t)3 * kZTile * n3 * sizeof(double2) > 160 * 1024)
    return set_err(FAME_ERR_UNSUPPORTED, "n3 too large for the z-pencil tile in LDS");
  const long long N = (long long)n1 * n2 * n3;
  if (mat_id) {
    for (long long p = 0; p < N; ++p)
      if (mat_id[p] >= n_mat) return set_err(FAME_ERR_ARG, "mat_id entry >= n_mat");
  }
  fame_plan* P = new fame_plan();
  P->n1 = n1; P->n2 = n2; P->n3 = n3; P->nb = nb; P->N = N; P->n_mat = n_mat;
  P->cb0 = 0; P->cnb = nb; P->vnb = nb;
  P->xyv = n1 >= 40 ? 1 : 0;  // 512-thread xy kernels: faster from n = 48 up, batched or not
  P->tab_len = n1 + n2 + n3;
  P->theta.assign(theta, theta + 3 * nb);
  for (int d = 0; d < 3; ++d) P->delta[d] = delta[d];
  P->geo.n1 = n1; P->geo.n2 = n2; P->geo.n3 = n3; P->geo.n12 = n1 * n2; P->geo.N = N;
  std::vector<double2> tw;
  std::vector<int> perm;
  std::map<int, int> seen;
  const int dims[3] = {n1, n2, n3};
  for (int d = 0; d < 3; ++d) {
    if (seen.count(dims[d])) {
      P->geo.ax[d] = P->geo.ax[seen[dims[d]]];
      continue;
    }
    int rc = build_axis(dims[d], tw, perm, P->geo.ax[d]);
    if (rc != FAME_OK) { delete P; return rc; }
    seen[dims[d]] = d;
  }
  auto fail = [&](int rc) { fame_plan_destroy(P); return rc; };
  if (hipGetDevice(&P->device) != hipSuccess) return fail(set_err(FAME_ERR_HIP, "hipGetDevice failed"));
  if (hipDeviceGetAttribute(&P->ncu, hipDeviceAttributeMultiprocessorCount, P->device) != hipSuccess || P->ncu < 1)
    return fail(set_err(FAME_ERR_HIP, "compute unit count"));
#define PLAN_HIP(call)                                                                              \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) return fail(set_err(FAME_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_))); \
  } while (0)
  PLAN_HIP(hipStreamCreateWithFlags(&P->own_stream, hipStreamNonBlocking));
  P->stream = P->own_stream;
  PLAN_HIP(hipMalloc(&P->d_tw, tw.size() * sizeof(double2)));
  PLAN_HIP(hipMalloc(&P->d_perm, perm.size() * sizeof(int)));
  PLAN_HIP(hipMemcpy(P->d_tw, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
  PLAN_HIP(hipMemcpy(P->d_perm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice));
  PLAN_HIP(hipMalloc(&P->d_lam, (size_t)nb * P->tab_len * sizeof(double2)));
  PLAN_HIP(hipMalloc(&P->d_tws, (size_t)nb * P->tab_len * sizeof(double2)));
  PLAN_HIP(hipMalloc(&P->d_omega, (size_t)nb * 3 * sizeof(double2)));
  PLAN_HIP(hipMalloc(&P->d_lam_v, (size_t)nb * P->tab_len * sizeof(double2)));
  PLAN_HIP(hipMalloc(&P->d_tws_v, (size_t)nb * P->tab_len * sizeof(double2)));
  PLAN_HIP(hipMalloc(&P->d_om_v, (size_t)nb * 3 * sizeof(double2)));
  PLAN_HIP(hipMalloc(&P->d_idx, (size_t)nb * sizeof(int)));
  P->v_lam = P->d_lam;
  P->v_tws = P->d_tws;
  P->v_om = P->d_omega;
  // material: symmetrise eps and mu as 1/2 (X + X^H) (the symmetrisation of N_int, M_int at
  // Interpolation_Material_Matrix_Mult_Point_vertex.m:22,24 acts on the vertex tensors like this)
  std::vector<double2> mt((size_t)n_mat * kMatWords);
  for (int m = 0; m < n_mat; ++m) {
    const double* src = mat_table + (size_t)m * kMatWords * 2;
    auto at = [&](int blk, int i, int j) { return make_double2(src[2 * (9 * blk + 3 * i + j)], src[2 * (9 * blk + 3 * i + j) + 1]); };
    for (int blk = 0; blk < 3; ++blk)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double2 v = at(blk, i, j);
          if (blk < 2) {
            const double2 t = at(blk, j, i);
            v = make_double2(0.5 * (v.x + t.x), 0.5 * (v.y - t.y));
          }
          mt[(size_t)m * kMatWords + 9 * blk + 3 * i + j] = v;
        }
  }
  PLAN_HIP(hipMalloc(&P->d_mat, mt.size() * sizeof(double2)));
  PLAN_HIP(hipMemcpy(P->d_mat, mt.data(), mt.size() * sizeof(double2), hipMemcpyHostToDevice));
  PLAN_HIP(hipMalloc(&P->d_mat_id, (size_t)N));
  if (mat_id)
    PLAN_HIP(hipMemcpy(P->d_mat_id, mat_id, (size_t)N, hipMemcpyHostToDevice));
  else
    PLAN_HIP(hipMemset(P->d_mat_id, 0, (size_t)N));
  PLAN_HIP(hipMalloc(&P->d_w1, (size_t)nb * 6 * N * sizeof(double2)));
  PLAN_HIP(hipMalloc(&P->d_w2, (size_t)nb * 6 * N * sizeof(double2)));
  P->nblk_red = 1024;
  PLAN_HIP(hipMalloc(&P->d_part, (size_t)nb * P->nblk_red * sizeof(double2)));
  PLAN_HIP(hipMalloc(&P->d_red, (size_t)nb * 2 * sizeof(double2)));
  PLAN_HIP(hipMalloc(&P->d_cg, (size_t)nb * sizeof(CgState)));
  // LDS above 64 KB needs an explicit opt-in per kernel
  const int lds_max = (int)std::max(xypass_lds(P), zpass_lds(P));
  PLAN_HIP(hipFuncSetAttribute((const void*)k_xypass<1, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  PLAN_HIP(hipFuncSetAttribute((const void*)k_xypass<1, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  PLAN_HIP(hipFuncSetAttribute((const void*)k_xypass<-1, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  PLAN_HIP(hipFuncSetAttribute((const void*)k_xypass<-1, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  PLAN_HIP(hipFuncSetAttribute((const void*)k_zpass<1, ZM_MIX_IN>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  PLAN_HIP(hipFuncSetAttribute((const void*)k_zpass<1, ZM_PLAIN>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  PLAN_HIP(hipFuncSetAttribute((const void*)k_zpass<-1, ZM_MIX_OUT>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  PLAN_HIP(hipFuncSetAttribute((const void*)k_zpass<-1, ZM_PLAIN>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  PLAN_HIP(hipFuncSetAttribute((const void*)k_material_z<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)material_z_lds_max()));
  PLAN_HIP(hipFuncSetAttribute((const void*)k_material_z<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)material_z_lds_max()));
#undef PLAN_HIP
  {
    int rc = FAME_OK;
    if (use_fast_xy(P)) rc = set_fast_attrs(n1);
    if (rc == FAME_OK && use_fast_z(P)) rc = set_fast_attrs(n3);
    if (rc != FAME_OK) return fail(rc);
  }
  int rc = upload_k_tables(P);
  if (rc != FAME_OK) return fail(rc);
  *plan = P;
  return FAME_OK;
}

int fame_plan_set_theta(fame_plan* P, const double* theta) {
  if (!P || !theta) return set_err(FAME_ERR_ARG, "NULL argument");
  RET_IF(check_theta(P->n1, P->n2, P->n3, P->nb, theta, P->delta));
  P->theta.assign(theta, theta + 3 * P->nb);
  return upload_k_tables(P);
}

int fame_plan_destroy(fame_plan* P) {
  if (!P) return FAME_OK;
  if (P->stream) (void)hipStreamSynchronize(P->stream);
  void* bufs[] = {P->d_tw, P->d_perm, P->d_lam, P->d_tws, P->d_omega, P->d_mat_id, P->d_mat, P->d_w1, P->d_w2,
                  P->d_part, P->d_red, P->d_r, P->d_p, P->d_q, P->d_cg, P->d_lam_v, P->d_tws_v, P->d_om_v,
                  P->d_idx, P->d_part2, P->d_eps, P->d_gam, P->d_pvw, P->d_mr, P->d_mrst, P->d_mgs, P->d_bdp, P->d_jdy, P->d_mrp,
                  P->d_mgp, P->d_mrm, P->d_mrskip};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (hipEvent_t e : P->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : P->evk)
    if (e) (void)hipEventDestroy(e);
  if (P->ev_poll) (void)hipEventDestroy(P->ev_poll);
  if (P->ev_fork) (void)hipEventDestroy(P->ev_fork);
  for (int i = 0; i < 2; ++i) {
    if (P->ev_join[i]) (void)hipEventDestroy(P->ev_join[i]);
    if (P->xstream[i]) {
      (void)hipStreamSynchronize(P->xstream[i]);
      (void)hipStreamDestroy(P->xstream[i]);
    }
  }
  if (P->h_cg) (void)hipHostFree(P->h_cg);
  if (P->h_stage) (void)hipHostFree(P->h_stage);
  if (P->h_mr) (void)hipHostFree(P->h_mr);
  if (P->ev_mr) (void)hipEventDestroy(P->ev_mr);
  if (P->own_stream) {
    (void)hipStreamSynchronize(P->own_stream);
    (void)hipStreamDestroy(P->own_stream);
  }
  delete P;
  return FAME_OK;
}

int fame_sync(fame_plan* P) {
  if (!P) return set_err(FAME_ERR_ARG, "NULL plan");
  HIP_TRY(hipStreamSynchronize(P->stream));
  return FAME_OK;
}

void* fame_plan_stream(fame_plan* P) { return P ? (void*)P->stream : nullptr; }

int fame_plan_set_stream(fame_plan* P, void* stream) {
  if (!P) return set_err(FAME_ERR_ARG, "NULL plan");
  HIP_TRY(hipStreamSynchronize(P->stream));
  P->stream = (hipStream_t)stream;
  return FAME_OK;
}

int fame_plan_set_chunk(fame_plan* P, int chunk) {
  if (!P || chunk < 0) return set_err(FAME_ERR_ARG, "bad argument");
  P->chunk = chunk;
  return FAME_OK;
}

int fame_plan_get_chunk(fame_plan* P) { return P ? P->chunk : -1; }

int fame_plan_set_variant(fame_plan* P, int xy_variant, int z_variant) {
  if (!P || xy_variant < 0 || xy_variant > 2 || z_variant < 0 || z_variant > 4) return set_err(FAME_ERR_ARG, "bad variant");
  P->xyv = xy_variant;
  P->zv = z_variant;
  return FAME_OK;
}

int fame_plan_get_variant(fame_plan* P, int* xy_variant, int* z_variant) {
  if (!P) return set_err(FAME_ERR_ARG, "NULL plan");
  if (xy_variant) *xy_variant = P->xyv;
  if (z_variant) *z_variant = P->zv;
  return FAME_OK;
}

// Time fncB on the plan's own scratch for each candidate chunk and keep the fastest.
int fame_plan_autotune(fame_plan* P, double* best_ms) {
  if (!P) return set_err(FAME_ERR_ARG, "NULL plan");
  const size_t bytes = (size_t)P->nb * 4 * P->N * sizeof(double2);
  // on every exit: free the work buffers and events, restore profiling; on an error also
  // restore the configuration the plan had before
  struct Guard {
    fame_plan* P;
    double2 *x = nullptr, *y = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool prof, ok = false;
    int xyv, zv, chunk, mt[3], ns;
    explicit Guard(fame_plan* p)
        : P(p), prof(p->prof), xyv(p->xyv), zv(p->zv), chunk(p->chunk), mt{p->mt[0], p->mt[1], p->mt[2]},
          ns(p->nstreams) {}
    ~Guard() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (x || y) (void)hipStreamSynchronize(P->stream);
      if (x) (void)hipFree(x);
      if (y) (void)hipFree(y);
      P->prof = prof;
      if (!ok) {
        P->xyv = xyv;
        P->zv = zv;
        P->chunk = chunk;
        P->nstreams = ns;
        for (int i = 0; i < 3; ++i) P->mt[i] = mt[i];
      }
    }
  } G(P);
  HIP_TRY(hipMalloc(&G.x, bytes));
  HIP_TRY(hipMalloc(&G.y, bytes));
  double2* x = G.x;
  double2* y = G.y;
  hipLaunchKernelGGL(k_fill_pattern, dim3(grid1d((long long)P->nb * 4 * P->N)), dim3(kBlk), 0, P->stream, x,
                     (long long)P->nb * 4 * P->N);
  CHECK_LAUNCH();
  HIP_TRY(hipEventCreate(&G.e0));
  HIP_TRY(hipEventCreate(&G.e1));
  hipEvent_t e0 = G.e0, e1 = G.e1;
  P->prof = false;
  // the best of three timings of five calls each (a single timing varies by a few per cent)
  auto time_it = [&](double& ms_out) -> int {
    RET_IF(do_fncB(P, x, y));  // warm-up
    ms_out = 1e30;
    for (int trial = 0; trial < 3; ++trial) {
      HIP_TRY(hipEventRecord(e0, P->stream));
      const int reps = 5;
      for (int r = 0; r < reps; ++r) RET_IF(do_fncB(P, x, y));
      HIP_TRY(hipEventRecord(e1, P->stream));
      HIP_TRY(hipEventSynchronize(e1));
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      ms_out = std::min(ms_out, (double)ms / reps);
    }
    return FAME_OK;
  };
  double best = 1e30;
  int bestc = 0, bestxy = 0, bestz = 0;
  // launch shapes first (whole batch), then the chunk size with the best shapes
  P->chunk = 0;
  // the xy and z shapes belong to different kernels, whose times add: the xy shape with the
  // plan's z shape, then the z shape with the best xy shape
  const int z0 = P->zv;
  for (int xv = 0; xv < (P->n1 >= 64 ? 3 : 2); ++xv) {
    P->xyv = xv;
    P->zv = z0;
    double ms;
    RET_IF(time_it(ms));
    if (ms < best) { best = ms; bestxy = xv; bestz = z0; }
  }
  P->xyv = bestxy;
  for (int zv = 0; zv < 5; ++zv) {
    if (zv == z0) continue;
    P->zv = zv;
    double ms;
    RET_IF(time_it(ms));
    if (ms < best) { best = ms; bestz = zv; }
  }
  P->xyv = bestxy;
  P->zv = bestz;
  std::vector<int> cands;
  for (int c : {1, 2, 3, 4, 6, 8, 12, 16})
    if (c < P->nb) cands.push_back(c);
  // each chunk size on one, two or three streams (chunks on different streams overlap)
  P->nstreams = 1;
  int bestns = 1;
  if (!cands.empty()) RET_IF(fame_plan_set_streams(P, 1));
  // chunk sizes in increasing order, each on one, two or three streams; the fncB time is unimodal
  // in both (measured at n = 24-96: it falls to the best chunk and stream count and rises after),
  // so a stream count slower than the one before ends that chunk's stream search, and two chunk
  // sizes in a row more than 5 % slower than the best found end the chunk search
  int worse = 0;
  for (int c : cands) {
    double cbest = 1e30, prev = 1e30;
    for (int ns = 1; ns <= 3; ++ns) {
      if (ns > 1 && (ns - 1) * c >= P->nb) break;  // fewer chunks than streams
      RET_IF(fame_plan_set_streams(P, ns));
      P->chunk = c;
      double ms;
      RET_IF(time_it(ms));
      cbest = std::min(cbest, ms);
      if (ms < best) { best = ms; bestc = c; bestns = ns; }
      if (ms > prev) break;
      prev = ms;
    }
    worse = (cbest > 1.05 * best) ? worse + 1 : 0;
    if (worse >= 2) break;
  }
  P->chunk = bestc;
  P->nstreams = bestns;
  // material stencil tile and z-slab: the heuristic's best few shapes (and wide ones), four slab
  // lengths each
  if (P->fast) {
    struct Cand { double score; int tx, ty; };
    std::vector<Cand> tiles;
    for (int tx = 4; tx <= std::min(P->n1, 30); ++tx)
      for (int ty = 4; ty <= std::min(P->n2, 30); ++ty) {
        if ((tx + 2) * (ty + 2) > kMzThreads) continue;
        const double pad = (double)P->n1 * P->n2 /
                           ((double)((P->n1 + tx - 1) / tx * tx) * ((P->n2 + ty - 1) / ty * ty));
        tiles.push_back({pad * tx * ty / (double)((tx + 2) * (ty + 2)), tx, ty});
      }
    std::sort(tiles.begin(), tiles.end(), [](const Cand& a, const Cand& b) { return a.score > b.score; });
    if (tiles.size() > 6) tiles.resize(6);
    // wide tiles: longer contiguous runs per row (DRAM page locality) at a larger halo share
    for (int tx : {30, 40, 60})
      for (int ty = 1; ty <= 30; ++ty)
        if (tx <= P->n1 && ty <= P->n2 && (tx + 2) * (ty + 2) <= kMzThreads && (tx + 2) * (ty + 3) > kMzThreads) {
          bool dup = false;
          for (const Cand& c : tiles) dup |= (c.tx == tx && c.ty == ty);
          if (!dup) tiles.push_back({0.0, tx, ty});
        }
    const int nbc = bestc > 0 ? bestc : P->nb;
    // The tile and slab change only the stencil, so they are chosen by the stencil's own time
    // on one chunk (steadier than the whole fncB's, where the other four kernels add noise);
    // the fields it reads are those the fncB runs above left in the scratch.
    const long long N = P->N;
    auto time_mat = [&](double& ms_out) -> int {
      struct Win {
        fame_plan* P;
        ~Win() { P->cb0 = 0; P->cnb = P->vnb; }
      } win{P};
      P->cb0 = 0;
      P->cnb = nbc;
      RET_IF(material(P, P->d_w1, 6 * N, P->d_w2, 6 * N));  // warm-up
      ms_out = 1e30;
      for (int trial = 0; trial < 3; ++trial) {
        HIP_TRY(hipEventRecord(e0, P->stream));
        const int reps = 5;
        for (int r = 0; r < reps; ++r) RET_IF(material(P, P->d_w1, 6 * N, P->d_w2, 6 * N));
        HIP_TRY(hipEventRecord(e1, P->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        ms_out = std::min(ms_out, (double)ms / reps);
      }
      return FAME_OK;
    };
    int btx = 0, bty = 0, btz = 0;
    double bmat;
    for (int i = 0; i < 3; ++i) P->mt[i] = 0;
    RET_IF(time_mat(bmat));  // the heuristic's tile
    // slab lengths: about one, two or three rounds of work-groups, or half a round
    for (const Cand& c : tiles) {
      const int hz = slab_length(P->n1, P->n2, P->n3, c.tx, c.ty, nbc, P->ncu);
      for (int tz : {hz, std::max(1, hz / 2), std::max(1, hz / 3), std::min(P->n3, 2 * hz)}) {
        P->mt[0] = c.tx;
        P->mt[1] = c.ty;
        P->mt[2] = tz;
        double ms;
        RET_IF(time_mat(ms));
        if (ms < bmat) { bmat = ms; btx = c.tx; bty = c.ty; btz = tz; }
      }
    }
    P->mt[0] = btx;
    P->mt[1] = bty;
    P->mt[2] = btz;
    RET_IF(time_it(best));  // the whole fncB at the chosen configuration
  }
  G.ok = true;
  if (best_ms) *best_ms = best;
  return FAME_OK;
}

int fame_plan_set_generic(fame_plan* P, int generic) {
  if (!P) return set_err(FAME_ERR_ARG, "NULL plan");
  P->fast = (generic == 0);
  return FAME_OK;
}

int fame_plan_set_cg_graph(fame_plan* P, int on) {
  if (!P) return set_err(FAME_ERR_ARG, "NULL plan");
  P->cg_graph = (on != 0);
  return FAME_OK;
}

int fame_plan_set_streams(fame_plan* P, int nstreams) {
  if (!P) return set_err(FAME_ERR_ARG, "NULL plan");
  if (nstreams < 1 || nstreams > 3) return set_err(FAME_ERR_ARG, "streams: 1, 2 or 3");
  if (nstreams > 1 && !P->ev_fork) HIP_TRY(hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming));
  for (int i = 0; i + 1 < nstreams; ++i) {
    if (!P->xstream[i]) HIP_TRY(hipStreamCreateWithFlags(&P->xstream[i], hipStreamNonBlocking));
    if (!P->ev_join[i]) HIP_TRY(hipEventCreateWithFlags(&P->ev_join[i], hipEventDisableTiming));
  }
  P->nstreams = nstreams;
  return FAME_OK;
}

int fame_plan_get_streams(fame_plan* P, int* nstreams) {
  if (!P || !nstreams) return set_err(FAME_ERR_ARG, "NULL argument");
  *nstreams = P->nstreams;
  return FAME_OK;
}

int fame_plan_set_material_tile(fame_plan* P, int tx, int ty, int tz) {
  if (!P) return set_err(FAME_ERR_ARG, "NULL plan");
  if (tx == 0 && ty == 0) {
    P->mt[0] = P->mt[1] = P->mt[2] = 0;
    return FAME_OK;
  }
  if (tx < 1 || ty < 1 || tz < 0 || tx > P->n1 || ty > P->n2 || tz > P->n3 || (tx + 2) * (ty + 2) > kMzThreads)
    return set_err(FAME_ERR_ARG, "material tile: need (tx+2)(ty+2) <= 256, tx <= n1, ty <= n2, 0 <= tz <= n3");
  P->mt[0] = tx;
  P->mt[1] = ty;
  P->mt[2] = tz;
  return FAME_OK;
}

int fame_plan_get_material_tile(fame_plan* P, int* tx, int* ty, int* tz) {
  if (!P || !tx || !ty || !tz) return set_err(FAME_ERR_ARG, "NULL argument");
  *tx = P->mt[0];
  *ty = P->mt[1];
  *tz = P->mt[2];
  return FAME_OK;
}

int fame_plan_set_cg_fused(fame_plan* P, int fused) {
  if (!P) return set_err(FAME_ERR_ARG, "NULL plan");
  if (fused < 0) return set_err(FAME_ERR_ARG, "fused must be >= 0");
  P->cg_fused = (fused != 0);
  P->cg_seq = fused > 1 ? fused - 1 : 0;
  return FAME_OK;
}

int fame_plan_profile(fame_plan* P, int enable) {
  if (!P) return set_err(FAME_ERR_ARG, "NULL plan");
  if (enable && !P->ev[0]) {
    for (auto& e : P->ev) HIP_TRY(hipEventCreate(&e));
    for (auto& e : P->evk) HIP_TRY(hipEventCreate(&e));
  }
  P->prof = enable != 0;
  for (double& v : P->stage_ms) v = 0;
  P->prof_calls = 0;
  P->mat_kernel_ms = 0;
  P->mat_launches = 0;
  P->evk_used = false;
  return FAME_OK;
}

int fame_plan_material_kernel_time(fame_plan* P, double* ms, long long* launches) {
  if (!P || !ms) return set_err(FAME_ERR_ARG, "NULL argument");
  *ms = P->mat_kernel_ms;
  if (launches) *launches = P->mat_launches;
  return FAME_OK;
}

int fame_plan_stage_times(fame_plan* P, double* ms, long long* calls) {
  if (!P || !ms) return set_err(FAME_ERR_ARG, "NULL argument");
  for (int s = 0; s < 5; ++s) ms[s] = P->stage_ms[s];
  if (calls) *calls = P->prof_calls;
  return FAME_OK;
}

int fame_MtxVecMul_FFT(fame_plan* P, int istrans, const void* in, void* out) {
  if (!P || !in || !out) return set_err(FAME_ERR_ARG, "NULL argument");
  const long long N = P->N;
  const double2* x = (const double2*)in;
  double2* y = (double2*)out;
  if (istrans == 0) {
    RET_IF(zpass(P, +1, ZM_PLAIN, x, 3 * N, P->d_w1, 3 * N, 1, 0));
    RET_IF(xypass(P, +1, 1, P->d_w1, 3 * N, y, 3 * N, 3));
  } else if (istrans == 1) {
    RET_IF(xypass(P, -1, 2, x, 3 * N, P->d_w1, 3 * N, 3));
    RET_IF(zpass(P, -1, ZM_PLAIN, P->d_w1, 3 * N, y, 3 * N, 1, 0));
  } else {
    return set_err(FAME_ERR_ARG, "istrans must be 0 ('notranspose') or 1 ('transpose')");
  }
  return FAME_OK;
}

static int pr_qr(fame_plan* P, int grp, const void* in, void* out) {
  if (!P || !in || !out) return set_err(FAME_ERR_ARG, "NULL argument");
  const long long N = P->N;
  RET_IF(zpass(P, +1, ZM_MIX_IN, (const double2*)in, 2 * N, P->d_w1, 3 * N, 1, grp));
  RET_IF(xypass(P, +1, 1, P->d_w1, 3 * N, (double2*)out, 3 * N, 3));
  return FAME_OK;
}

static int prs_qrs(fame_plan* P, int grp, const void* in, void* out) {
  if (!P || !in || !out) return set_err(FAME_ERR_ARG, "NULL argument");
  const long long N = P->N;
  RET_IF(xypass(P, -1, 2, (const double2*)in, 3 * N, P->d_w1, 3 * N, 3));
  RET_IF(zpass(P, -1, ZM_MIX_OUT, P->d_w1, 3 * N, (double2*)out, 2 * N, 1, grp));
  return FAME_OK;
}

int fame_Pr(fame_plan* P, const void* in, void* out) { return pr_qr(P, 0, in, out); }
int fame_Qr(fame_plan* P, const void* in, void* out) { return pr_qr(P, 1, in, out); }
int fame_Prs(fame_plan* P, const void* in, void* out) { return prs_qrs(P, 0, in, out); }
int fame_Qrs(fame_plan* P, const void* in, void* out) { return prs_qrs(P, 1, in, out); }

int fame_fncA(fame_plan* P, const void* in, void* out) {
  if (!P || !in || !out) return set_err(FAME_ERR_ARG, "NULL argument");
  return do_fncA(P, (const double2*)in, (double2*)out);
}

int fame_fncB(fame_plan* P, const void* in, void* out) {
  if (!P || !in || !out) return set_err(FAME_ERR_ARG, "NULL argument");
  if (in == out) return set_err(FAME_ERR_ARG, "fncB: in and out must not alias");
  return do_fncB(P, (const double2*)in, (double2*)out);
}

int fame_material_apply(fame_plan* P, const void* in, void* out) {
  if (!P || !in || !out) return set_err(FAME_ERR_ARG, "NULL argument");
  const long long N = P->N;
  RET_IF(twist(P, (const double2*)in, P->d_w1, 6, 1));
  RET_IF(material(P, P->d_w1, 6 * N, P->d_w2, 6 * N));
  RET_IF(twist(P, P->d_w2, (double2*)out, 6, 0));
  return FAME_OK;
}

int fame_dot(fame_plan* P, const void* x, const void* y, long long len, void* out) {
  if (!P || !x || !y || !out) return set_err(FAME_ERR_ARG, "NULL argument");
  return do_dot(P, (const double2*)x, len, (const double2*)y, len, len, P->nb, (double2*)out);
}

static int ensure_cg(fame_plan* P) {
  const size_t bytes = (size_t)P->nb * 4 * P->N * sizeof(double2);
  if (!P->d_r) HIP_TRY(hipMalloc(&P->d_r, bytes));
  if (!P->d_p) HIP_TRY(hipMalloc(&P->d_p, bytes));
  if (!P->d_q) HIP_TRY(hipMalloc(&P->d_q, bytes));
  return FAME_OK;
}

int fame_fncinvB(fame_plan* P, const void* in, void* out, double tol, int maxit, int* iters, double* relres) {
  if (!P || !in || !out) return set_err(FAME_ERR_ARG, "NULL argument");
  if (in == out) return set_err(FAME_ERR_ARG, "fncinvB: in and out must not alias");
  if (!(tol > 0) || maxit < 1) return set_err(FAME_ERR_ARG, "need tol > 0 and maxit >= 1");
  RET_IF(ensure_cg(P));
  return do_cg(P, (const double2*)in, (double2*)out, tol, maxit, iters, relres);
}

}  // extern "C"

#include "fame_lanczos.inc"
#include "fame_minres.inc"
#include "fame_fcc.inc"
#include "fame_vertex.inc"
#include "fame_probe.inc"
#include "fame_noalloc.inc"
#include "fame_voxel.inc"
#include "fame_jd.inc"
