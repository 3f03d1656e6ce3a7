/* MATLAB MEX gateway over libfame (include/fame_abi.h).
 *
 * Source only: neither MATLAB nor mex.h exists in this build's environment, so
 * this file is not compiled or tested here.  Build inside MATLAB with
 *   mex -R2018a -I<repo>/include fame_mex.c -L<pkg> -lfame -lamdhip64
 * (interleaved complex API, R2018a+).
 *
 * MATLAB holds the vectors on the host (gpuArray is CUDA-only), so every
 * operator call copies its input to a device buffer owned by the plan slot and
 * its output back: PCIe traffic per call.  The solver entry points (fncinvB,
 * lanczos) keep all iterations on the device and copy only their operands.
 *
 *   h = fame_mex('plan_create', [n1 n2 n3], theta, delta, mat_table, mat_id)
 *        theta     nb x 3 double       (k . a_d per k-point)
 *        delta     1 x 3 double        (opt.comp.discrete.mesh_len)
 *        mat_table n_mat x 27 complex  (N-, M-, J-slot 3x3 tensors, row-major)
 *        mat_id    N x 1 uint8         (0-based material per vertex)
 *   y = fame_mex('fncA' | 'fncB', h, x)           x: 4N x nb complex
 *   y = fame_mex('Psi' | 'Psis', h, x)            x: 3N x nb complex (Psi == Phi for QP)
 *   y = fame_mex('Pr' | 'Qr', h, x)               x: 2N x nb complex -> 3N x nb
 *   y = fame_mex('Prs' | 'Qrs', h, x)             x: 3N x nb complex -> 2N x nb
 *   [y, iters, relres] = fame_mex('fncinvB', h, x, tol, maxit)
 *   [ev, ew, iters] = fame_mex('lanczos', h, k, tol, maxiter, p, v0)   v0: 4N x nb
 *   [y, iters, relres] = fame_mex('minres', h, x, sigma, tol, maxit)     (fncB - sigma fncA) y = x
 *   fame_mex('vertex_material', h, eps, gamma, mu, lattice)              eps, gamma: N x 1, lattice 3x3
 *   g = fame_mex('material_gradient', h, ev)   ev: 4N x nvec (nb = 1) -> N x 3 x nvec
 *   fame_mex('set_structure', h, kind, params, lattice)   device voxelisation (Material_Locate_Isofunction):
 *        kind 0 woodpile {width, a}, 1 sphere {r, c1, c2, c3}, 2 tetrahedron {edge, c1, c2, c3};
 *        lattice 3x3, lattice vectors as columns
 *   ids = fame_mex('material_ids', h)               N x 1 uint8
 *   fame_mex('plan_destroy', h)
 *
 *   f = fame_mex('fcc_create', [n1 n2 n3], vec_k, a1, a2, a3, delta)     SVDSingleCurl_FCC tables
 *   y = fame_mex('fcc_T' | 'fcc_TH', f, x)                               x: n x nvec complex
 *   fame_mex('fcc_set_material', f, edge_mask, face_mask, tensors)     mtx_Eps_Mu_Zeta_Xi: masks n x 1
 *        uint8, tensors 72 complex = [inside, air][eps, mu, zeta, xi] 3x3 row-major
 *   [xi_d, zeta_d, Phi] = fame_mex('fcc_nfgep_data', f)                 mtx_data from GEP B (coupling 1)
 *   y = fame_mex('fcc_nfgep', f, x, xi_d, zeta_d, Phi [, coupling])     x: 32n, coefficients 24n
 *   ev = fame_mex('fcc_eigenvector', f, x, xi_d, zeta_d, Phi [, coupling])   Compute_eigenvector, 48n
 *   fame_mex('fcc_destroy', f)
 */
#include <stdint.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "fame_abi.h"
#include "mex.h"

typedef struct {
  fame_plan* plan;
  long long N;
  int nb;
  void* din;   /* device staging, 6N x nb complex */
  void* dout;  /* device staging, 6N x nb complex */
} slot_t;

static slot_t* slot_of(const mxArray* a) {
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("fame:handle", "bad plan handle");
  return (slot_t*)(uintptr_t)(*(uint64_t*)mxGetData(a));
}

static void check(int rc) {
  if (rc != FAME_OK && rc != FAME_ERR_NOT_CONVERGED) mexErrMsgIdAndTxt("fame:call", "%s", fame_last_error());
}

static void hip_check(hipError_t e) {
  if (e != hipSuccess) mexErrMsgIdAndTxt("fame:hip", "%s", hipGetErrorString(e));
}

/* host complex (interleaved) -> device staging buffer */
static void to_dev(slot_t* s, const mxArray* x, long long len, void* dst) {
  if (!mxIsComplex(x)) mexErrMsgIdAndTxt("fame:arg", "complex input expected");
  if ((long long)mxGetNumberOfElements(x) != len * s->nb) mexErrMsgIdAndTxt("fame:arg", "wrong vector length");
  hip_check(hipMemcpy(dst, mxGetComplexDoubles(x), (size_t)len * s->nb * 16, hipMemcpyHostToDevice));
}

static mxArray* from_dev(slot_t* s, const void* src, long long len) {
  mxArray* y = mxCreateDoubleMatrix((mwSize)len, (mwSize)s->nb, mxCOMPLEX);
  check(fame_sync(s->plan));
  hip_check(hipMemcpy(mxGetComplexDoubles(y), src, (size_t)len * s->nb * 16, hipMemcpyDeviceToHost));
  return y;
}

typedef int (*op_t)(fame_plan*, const void*, void*);

static void apply(slot_t* s, op_t op, const mxArray* x, long long lin, long long lout, mxArray** y) {
  to_dev(s, x, lin, s->din);
  check(op(s->plan, s->din, s->dout));
  *y = from_dev(s, s->dout, lout);
}

static int psi(fame_plan* p, const void* a, void* b) { return fame_MtxVecMul_FFT(p, 0, a, b); }
static int psis(fame_plan* p, const void* a, void* b) { return fame_MtxVecMul_FFT(p, 1, a, b); }

/* FCC plans: handle + device staging sized on demand */
typedef struct {
  fame_fcc* plan;
  long long n;
  void* din;
  void* dout;
  size_t cap;  /* bytes of each staging buffer */
} fslot_t;

static fslot_t* fslot_of(const mxArray* a) {
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("fame:handle", "bad FCC plan handle");
  return (fslot_t*)(uintptr_t)(*(uint64_t*)mxGetData(a));
}

static void fstage(fslot_t* f, size_t bytes) {
  if (bytes <= f->cap) return;
  if (f->din) hipFree(f->din);
  if (f->dout) hipFree(f->dout);
  hip_check(hipMalloc(&f->din, bytes));
  hip_check(hipMalloc(&f->dout, bytes));
  f->cap = bytes;
}

static void* dev_copy(const mxArray* x, size_t count) {
  void* d = NULL;
  if (!mxIsComplex(x) || mxGetNumberOfElements(x) != count) mexErrMsgIdAndTxt("fame:arg", "complex array of the right length expected");
  hip_check(hipMalloc(&d, count * 16));
  hip_check(hipMemcpy(d, mxGetComplexDoubles(x), count * 16, hipMemcpyHostToDevice));
  return d;
}

static int fcc_command(const char* cmd, int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (!strcmp(cmd, "fcc_create")) {
    if (nrhs != 7) mexErrMsgIdAndTxt("fame:arg", "fcc_create(dims, vec_k, a1, a2, a3, delta)");
    const double* d = mxGetDoubles(prhs[1]);
    fslot_t* f = (fslot_t*)mxCalloc(1, sizeof(fslot_t));
    mexMakeMemoryPersistent(f);
    check(fame_fcc_create(&f->plan, (int)d[0], (int)d[1], (int)d[2], mxGetDoubles(prhs[2]), mxGetDoubles(prhs[3]),
                          mxGetDoubles(prhs[4]), mxGetDoubles(prhs[5]), mxGetDoubles(prhs[6])));
    f->n = (long long)d[0] * (long long)d[1] * (long long)d[2];
    plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
    *(uint64_t*)mxGetData(plhs[0]) = (uint64_t)(uintptr_t)f;
    return 1;
  }
  if (strncmp(cmd, "fcc_", 4)) return 0;
  if (nrhs < 2) mexErrMsgIdAndTxt("fame:arg", "missing FCC plan handle");
  fslot_t* f = fslot_of(prhs[1]);
  if (!strcmp(cmd, "fcc_destroy")) {
    fame_fcc_destroy(f->plan);
    if (f->din) hipFree(f->din);
    if (f->dout) hipFree(f->dout);
    mxFree(f);
  } else if (!strcmp(cmd, "fcc_T") || !strcmp(cmd, "fcc_TH")) {
    const size_t cnt = mxGetNumberOfElements(prhs[2]);
    if (!mxIsComplex(prhs[2]) || cnt % (size_t)f->n) mexErrMsgIdAndTxt("fame:arg", "x: n x nvec complex");
    fstage(f, cnt * 16);
    hip_check(hipMemcpy(f->din, mxGetComplexDoubles(prhs[2]), cnt * 16, hipMemcpyHostToDevice));
    const int nvec = (int)(cnt / (size_t)f->n);
    check(cmd[5] == 'H' ? fame_fcc_TH(f->plan, nvec, f->din, f->dout) : fame_fcc_T(f->plan, nvec, f->din, f->dout));
    check(fame_fcc_sync(f->plan));
    plhs[0] = mxCreateDoubleMatrix((mwSize)f->n, (mwSize)nvec, mxCOMPLEX);
    hip_check(hipMemcpy(mxGetComplexDoubles(plhs[0]), f->dout, cnt * 16, hipMemcpyDeviceToHost));
  } else if (!strcmp(cmd, "fcc_nfgep") || !strcmp(cmd, "fcc_eigenvector")) {
    if (nrhs != 6 && nrhs != 7) mexErrMsgIdAndTxt("fame:arg", "%s(f, x, xi_d, zeta_d, Phi [, coupling])", cmd);
    const int eig = cmd[4] == 'e';
    const size_t n32 = 32 * (size_t)f->n, n24 = 24 * (size_t)f->n, nout = eig ? 48 * (size_t)f->n : n32;
    const int coupling = nrhs == 7 ? (int)mxGetScalar(prhs[6]) : 0;
    fstage(f, nout * 16);
    void* c[3];
    for (int i = 0; i < 3; ++i) c[i] = dev_copy(prhs[3 + i], n24);
    if (!mxIsComplex(prhs[2]) || mxGetNumberOfElements(prhs[2]) != n32) mexErrMsgIdAndTxt("fame:arg", "x: 32n complex");
    hip_check(hipMemcpy(f->din, mxGetComplexDoubles(prhs[2]), n32 * 16, hipMemcpyHostToDevice));
    check(eig ? fame_fcc_eigenvector_ex(f->plan, f->din, f->dout, c[0], c[1], c[2], coupling)
              : fame_fcc_nfgep_A_ex(f->plan, f->din, f->dout, c[0], c[1], c[2], coupling));
    check(fame_fcc_sync(f->plan));
    plhs[0] = mxCreateDoubleMatrix((mwSize)nout, 1, mxCOMPLEX);
    hip_check(hipMemcpy(mxGetComplexDoubles(plhs[0]), f->dout, nout * 16, hipMemcpyDeviceToHost));
    for (int i = 0; i < 3; ++i) hipFree(c[i]);
  } else if (!strcmp(cmd, "fcc_set_material")) {
    if (nrhs != 5 || !mxIsComplex(prhs[4]) || mxGetNumberOfElements(prhs[4]) != 72 ||
        mxGetNumberOfElements(prhs[2]) != (size_t)f->n || mxGetNumberOfElements(prhs[3]) != (size_t)f->n)
      mexErrMsgIdAndTxt("fame:arg", "fcc_set_material(f, edge_mask (n uint8), face_mask (n uint8), tensors (72 complex))");
    check(fame_fcc_set_material(f->plan, (const unsigned char*)mxGetData(prhs[2]),
                                (const unsigned char*)mxGetData(prhs[3]),
                                (const double*)mxGetComplexDoubles(prhs[4])));
  } else if (!strcmp(cmd, "fcc_nfgep_data")) {
    const size_t n24 = 24 * (size_t)f->n;
    void* c[3];
    for (int i = 0; i < 3; ++i) hip_check(hipMalloc(&c[i], n24 * 16));
    check(fame_fcc_nfgep_data(f->plan, c[0], c[1], c[2]));
    check(fame_fcc_sync(f->plan));
    for (int i = 0; i < 3 && i < (nlhs > 0 ? nlhs : 1); ++i) {
      plhs[i] = mxCreateDoubleMatrix((mwSize)n24, 1, mxCOMPLEX);
      hip_check(hipMemcpy(mxGetComplexDoubles(plhs[i]), c[i], n24 * 16, hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < 3; ++i) hipFree(c[i]);
  } else {
    mexErrMsgIdAndTxt("fame:arg", "unknown command %s", cmd);
  }
  (void)nlhs;
  return 1;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  char cmd[32];
  if (nrhs < 1 || mxGetString(prhs[0], cmd, sizeof cmd)) mexErrMsgIdAndTxt("fame:arg", "first argument: command");
  if (fcc_command(cmd, nlhs, plhs, nrhs, prhs)) return;
  if (!strcmp(cmd, "plan_create")) {
    if (nrhs != 6) mexErrMsgIdAndTxt("fame:arg", "plan_create(dims, theta, delta, mat_table, mat_id)");
    const double* d = mxGetDoubles(prhs[1]);
    const int n1 = (int)d[0], n2 = (int)d[1], n3 = (int)d[2];
    const int nb = (int)mxGetM(prhs[2]);
    /* theta arrives column-major nb x 3; the ABI wants [nb][3] */
    double* th = (double*)mxMalloc(sizeof(double) * 3 * nb);
    const double* tm = mxGetDoubles(prhs[2]);
    for (int b = 0; b < nb; ++b)
      for (int c = 0; c < 3; ++c) th[3 * b + c] = tm[b + nb * c];
    const int n_mat = (int)mxGetM(prhs[4]);
    /* mat_table column-major n_mat x 27 -> [n_mat][27] interleaved */
    double* tab = (double*)mxMalloc(sizeof(double) * 54 * n_mat);
    const mxComplexDouble* tc = mxGetComplexDoubles(prhs[4]);
    for (int m = 0; m < n_mat; ++m)
      for (int w = 0; w < 27; ++w) {
        tab[2 * (27 * m + w)] = tc[m + n_mat * w].real;
        tab[2 * (27 * m + w) + 1] = tc[m + n_mat * w].imag;
      }
    slot_t* s = (slot_t*)mxCalloc(1, sizeof(slot_t));
    mexMakeMemoryPersistent(s);
    check(fame_plan_create(&s->plan, n1, n2, n3, nb, th, mxGetDoubles(prhs[3]), n_mat, tab,
                           (const unsigned char*)mxGetData(prhs[5])));
    s->N = (long long)n1 * n2 * n3;
    s->nb = nb;
    hip_check(hipMalloc(&s->din, (size_t)6 * s->N * nb * 16));
    hip_check(hipMalloc(&s->dout, (size_t)6 * s->N * nb * 16));
    mxFree(th);
    mxFree(tab);
    plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
    *(uint64_t*)mxGetData(plhs[0]) = (uint64_t)(uintptr_t)s;
    return;
  }
  if (nrhs < 2) mexErrMsgIdAndTxt("fame:arg", "missing plan handle");
  slot_t* s = slot_of(prhs[1]);
  const long long N = s->N;
  if (!strcmp(cmd, "plan_destroy")) {
    fame_plan_destroy(s->plan);
    hipFree(s->din);
    hipFree(s->dout);
    mxFree(s);
  } else if (!strcmp(cmd, "fncA")) {
    apply(s, fame_fncA, prhs[2], 4 * N, 4 * N, &plhs[0]);
  } else if (!strcmp(cmd, "fncB")) {
    apply(s, fame_fncB, prhs[2], 4 * N, 4 * N, &plhs[0]);
  } else if (!strcmp(cmd, "Psi") || !strcmp(cmd, "Phi")) {
    apply(s, psi, prhs[2], 3 * N, 3 * N, &plhs[0]);
  } else if (!strcmp(cmd, "Psis") || !strcmp(cmd, "Phis")) {
    apply(s, psis, prhs[2], 3 * N, 3 * N, &plhs[0]);
  } else if (!strcmp(cmd, "Pr")) {
    apply(s, fame_Pr, prhs[2], 2 * N, 3 * N, &plhs[0]);
  } else if (!strcmp(cmd, "Qr")) {
    apply(s, fame_Qr, prhs[2], 2 * N, 3 * N, &plhs[0]);
  } else if (!strcmp(cmd, "Prs")) {
    apply(s, fame_Prs, prhs[2], 3 * N, 2 * N, &plhs[0]);
  } else if (!strcmp(cmd, "Qrs")) {
    apply(s, fame_Qrs, prhs[2], 3 * N, 2 * N, &plhs[0]);
  } else if (!strcmp(cmd, "fncinvB")) {
    const double tol = mxGetScalar(prhs[3]);
    const int maxit = (int)mxGetScalar(prhs[4]);
    int* it = (int*)mxMalloc(sizeof(int) * s->nb);
    double* rr = (double*)mxMalloc(sizeof(double) * s->nb);
    to_dev(s, prhs[2], 4 * N, s->din);
    check(fame_fncinvB(s->plan, s->din, s->dout, tol, maxit, it, rr));
    plhs[0] = from_dev(s, s->dout, 4 * N);
    if (nlhs > 1) {
      plhs[1] = mxCreateDoubleMatrix(1, s->nb, mxREAL);
      for (int b = 0; b < s->nb; ++b) mxGetDoubles(plhs[1])[b] = it[b];
    }
    if (nlhs > 2) {
      plhs[2] = mxCreateDoubleMatrix(1, s->nb, mxREAL);
      memcpy(mxGetDoubles(plhs[2]), rr, sizeof(double) * s->nb);
    }
    mxFree(it);
    mxFree(rr);
  } else if (!strcmp(cmd, "lanczos")) {
    const int k = (int)mxGetScalar(prhs[2]);
    const double tol = mxGetScalar(prhs[3]);
    const int maxiter = (int)mxGetScalar(prhs[4]);
    const int p = (int)mxGetScalar(prhs[5]);
    void* dev_ev = NULL;
    hip_check(hipMalloc(&dev_ev, (size_t)k * 4 * N * s->nb * 16));
    to_dev(s, prhs[6], 4 * N, s->din);
    double* ew = (double*)mxMalloc(sizeof(double) * k * s->nb);
    int* it = (int*)mxMalloc(sizeof(int) * s->nb);
    check(fame_lanczos(s->plan, k, tol, maxiter, p, s->din, 1e-12, 10000, dev_ev, ew, it));
    plhs[0] = mxCreateDoubleMatrix((mwSize)(4 * N), (mwSize)(k * s->nb), mxCOMPLEX);
    hip_check(hipMemcpy(mxGetComplexDoubles(plhs[0]), dev_ev, (size_t)k * 4 * N * s->nb * 16, hipMemcpyDeviceToHost));
    hipFree(dev_ev);
    if (nlhs > 1) {
      plhs[1] = mxCreateDoubleMatrix((mwSize)k, (mwSize)s->nb, mxREAL);
      memcpy(mxGetDoubles(plhs[1]), ew, sizeof(double) * k * s->nb);
    }
    if (nlhs > 2) {
      plhs[2] = mxCreateDoubleMatrix(1, s->nb, mxREAL);
      for (int b = 0; b < s->nb; ++b) mxGetDoubles(plhs[2])[b] = it[b];
    }
    mxFree(ew);
    mxFree(it);
  } else if (!strcmp(cmd, "minres")) {
    const double sigma = mxGetScalar(prhs[3]), tol = mxGetScalar(prhs[4]);
    const int maxit = (int)mxGetScalar(prhs[5]);
    int* it = (int*)mxMalloc(sizeof(int) * s->nb);
    double* rr = (double*)mxMalloc(sizeof(double) * s->nb);
    to_dev(s, prhs[2], 4 * N, s->din);
    check(fame_minres(s->plan, s->din, s->dout, sigma, tol, maxit, it, rr));
    plhs[0] = from_dev(s, s->dout, 4 * N);
    if (nlhs > 1) {
      plhs[1] = mxCreateDoubleMatrix(1, s->nb, mxREAL);
      for (int b = 0; b < s->nb; ++b) mxGetDoubles(plhs[1])[b] = it[b];
    }
    if (nlhs > 2) {
      plhs[2] = mxCreateDoubleMatrix(1, s->nb, mxREAL);
      memcpy(mxGetDoubles(plhs[2]), rr, sizeof(double) * s->nb);
    }
    mxFree(it);
    mxFree(rr);
  } else if (!strcmp(cmd, "vertex_material")) {
    /* the 3x3 lattice arrives column-major; the ABI takes it row-major */
    const double* L = mxGetDoubles(prhs[5]);
    double A[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) A[3 * i + j] = L[i + 3 * j];
    check(fame_plan_set_vertex_material(s->plan, mxGetDoubles(prhs[2]), mxGetDoubles(prhs[3]), mxGetScalar(prhs[4]), A));
  } else if (!strcmp(cmd, "set_structure")) {
    if (nrhs != 5 || mxGetNumberOfElements(prhs[4]) != 9)
      mexErrMsgIdAndTxt("fame:arg", "set_structure(h, kind, params, lattice 3x3)");
    const double* L = mxGetDoubles(prhs[4]);
    double A[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) A[3 * i + j] = L[i + 3 * j];
    check(fame_plan_set_structure(s->plan, (int)mxGetScalar(prhs[2]), mxGetDoubles(prhs[3]), A));
  } else if (!strcmp(cmd, "material_ids")) {
    void* d = NULL;
    hip_check(hipMalloc(&d, (size_t)N));
    check(fame_plan_material_ids(s->plan, d));
    check(fame_sync(s->plan));
    plhs[0] = mxCreateNumericMatrix((mwSize)N, 1, mxUINT8_CLASS, mxREAL);
    hip_check(hipMemcpy(mxGetData(plhs[0]), d, (size_t)N, hipMemcpyDeviceToHost));
    hipFree(d);
  } else if (!strcmp(cmd, "material_gradient")) {
    if (s->nb != 1) mexErrMsgIdAndTxt("fame:arg", "material_gradient: plans with one k-point");
    const int nvec = (int)mxGetN(prhs[2]);
    void* dev_ev = dev_copy(prhs[2], (size_t)nvec * 4 * N);
    void* dev_g = NULL;
    hip_check(hipMalloc(&dev_g, (size_t)nvec * 3 * N * sizeof(double)));
    check(fame_material_gradient(s->plan, nvec, dev_ev, dev_g));
    check(fame_sync(s->plan));
    const mwSize dims[3] = {(mwSize)N, 3, (mwSize)nvec};
    plhs[0] = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
    hip_check(hipMemcpy(mxGetDoubles(plhs[0]), dev_g, (size_t)nvec * 3 * N * sizeof(double), hipMemcpyDeviceToHost));
    hipFree(dev_ev);
    hipFree(dev_g);
  } else {
    mexErrMsgIdAndTxt("fame:arg", "unknown command %s", cmd);
  }
}
