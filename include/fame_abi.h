/* C-ABI of the MI355X (gfx950) curl-SVD + bi-anisotropic material matvec library.
 *
 * Drop-in for the operator handles the reference's eigensolvers call through
 * MATLAB function handles (FAME-GPU/3D-Bi-Anisotropic-PhC).  Each entry point
 * names the reference function (file:line) it replaces.
 *
 * Conventions
 *   - Vectors are complex128, interleaved (re, im), in the reference's MATLAB
 *     column-major order: grid point p = i1 + n1*(i2 + n2*i3).  No permutation
 *     at the boundary.
 *   - All vector pointers are DEVICE pointers owned by the caller.  A plan holds
 *     `nb` Bloch vectors (k-points) on one grid and one material; vectors are
 *     batched as [nb][len] contiguous.  nb = 1 is the reference's plain path.
 *     Lengths: 3-component field 3N, Pr/Qr coefficient vector 2N, NFSEP vector
 *     4N = [x_a (2N); x_b (2N)], N = n1*n2*n3.
 *   - Calls are asynchronous on the plan's HIP stream; fame_sync() waits.  Solver
 *     calls (fame_fncinvB, fame_lanczos) return after their result is on the
 *     device.  Distinct plans may be used from distinct threads.
 *   - Return value: FAME_OK (0) or a negative error code; fame_last_error()
 *     returns the message of the calling thread's last error.
 *   - Grid lengths must factor into primes <= 13; every axis holds at most
 *     10240 points per xy-plane (n1*n2 <= 10240) and n3 <= 208.
 */
#ifndef FAME_ABI_H
#define FAME_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

#define FAME_OK 0
#define FAME_ERR_ARG (-1)
#define FAME_ERR_HIP (-2)
#define FAME_ERR_UNSUPPORTED (-3)
#define FAME_ERR_NOT_CONVERGED (-4)

typedef struct fame_plan fame_plan;

/* Version string of the library build. */
const char* fame_version(void);

/* Message of the last error raised on the calling thread ("" if none). */
const char* fame_last_error(void);

/* Create a plan: grid, Bloch phases, mesh lengths and material.
 *   theta      [nb][3]  theta_d = k . a_d
 *                       (band_structure_woodpile_Lanczos_n60_gamma0p2.m:54-56)
 *   delta      [3]      mesh lengths (Computational_settings_Lanczos.m:38)
 *   mat_table  [n_mat][27] complex (54 doubles): per material the 3x3 N-slot
 *              (inverse permittivity, covariant), M-slot (inverse permeability)
 *              and J-slot (inverse Xi) tensors, row-major, exactly the
 *              `par{2}`/`par{1}` tensors of DistMaterialMatrix
 *              (Public/Material_Matrix_Generate_Mult_Point.m:28-31,38-48).
 *   mat_id     [N] host array of material indices per vertex (NULL: all 0);
 *              the reference's `Standard_inner_idx` is mat_id == 1.
 * Replaces the per-k operand setup SVD_Matrix_Generate.m:24-30 +
 * Interpolation_Material_Matrix_Mult_Point_vertex.m:14-27 + Convert2gpuarray.m:2-20:
 * the device never stores the sparse N_int/M_int/J_int or Pi.Pr/Pi.Qr; it
 * evaluates them from these tables. */
int fame_plan_create(fame_plan** plan, int n1, int n2, int n3, int nb, const double* theta, const double* delta,
                     int n_mat, const double* mat_table, const unsigned char* mat_id);

/* Material ids computed on the device (replaces the host mat_id of fame_plan_create and the
 * host restatement of subroutines/Material_Locate_Isofunction.m:1-148): id 1 inside the
 * structure, 0 outside, at the vertices (i1/n1, i2/n2, i3/n3) (Grid_Generate_Mult_Point.m:7-17).
 *   kind 0: SC woodpile, params {rod width, a}: four layers of height a/4 with rods along x
 *           at y = a/4, along y at x = a/4, along x at y = 3a/4, along y at x = 3a/4
 *   kind 1: sphere, params {radius, c1, c2, c3} (centre in lattice coordinates)
 *   kind 2: regular tetrahedron, params {edge, c1, c2, c3}, vertices c + edge/sqrt(8) (+-1, +-1,
 *           +-1) with an even number of minus signs
 * Spheres and tetrahedra use `lattice` (row-major 3x3, lattice vectors as columns) and the
 * minimum image over the 27 neighbouring cells.  The plan's table needs n_mat >= 2. */
int fame_plan_set_structure(fame_plan* plan, int kind, const double* params, const double* lattice);
/* Copy the plan's per-vertex material ids (N bytes) to device memory `dst`. */
int fame_plan_material_ids(fame_plan* plan, void* dst);

/* Change the Bloch phases of all nb k-points (theta [nb][3]); grid and material stay. */
int fame_plan_set_theta(fame_plan* plan, const double* theta);

int fame_plan_destroy(fame_plan* plan);

/* Wait for all work queued on the plan's stream. */
int fame_sync(fame_plan* plan);

/* The plan's hipStream_t (as void*), for callers that order their own work after it. */
void* fame_plan_stream(fame_plan* plan);

/* Run the plan's work on the given hipStream_t from now on (e.g. the framework's
 * current stream); NULL is the device's default (null) stream.  The plan's own
 * stream is the value fame_plan_stream() returns right after creation.  Work
 * already queued on the previous stream is waited for first. */
int fame_plan_set_stream(fame_plan* plan, void* stream);

/* Batch chunking (no reference counterpart): fncB runs the nb k-points in chunks
 * of `chunk` items (0 = all at once) so a chunk's intermediate fields stay in the
 * last-level cache between kernels.  fame_plan_autotune times, on the plan, the FFT
 * launch shapes, then the chunk size with the stream count (whole fncB, best of three
 * timings), then the material-stencil tile and slab (the stencil's own time on one
 * chunk), keeps the fastest of each and returns the fncB time of the result in
 * best_ms (which may be NULL). */
int fame_plan_set_chunk(fame_plan* plan, int chunk);
int fame_plan_get_chunk(fame_plan* plan);
int fame_plan_autotune(fame_plan* plan, double* best_ms);

/* Launch shapes of the specialised FFT passes (no reference counterpart):
 * xy_variant 0 = default threads per plane work-group, 1 = 512 threads, 2 = 1024 threads for planes of
 * n1 >= 64 (512 below);
 * z_variant 0 = 16 pencils / 256 threads, 1 = 8 / 128, 2 = 32 / 512, 3 = 16 / 320, 4 = 16 / 384.
 * fame_plan_autotune picks these first, then the chunk size. */
int fame_plan_set_variant(fame_plan* plan, int xy_variant, int z_variant);
int fame_plan_get_variant(fame_plan* plan, int* xy_variant, int* z_variant);

/* Kernel selection (no reference counterpart): grid lengths with a compile-time
 * specialised FFT (8,12,16,20,24,32,40,48,60,64,72,80,90,96; xy needs n1 == n2)
 * use it by default, and the material stencil runs as a 2.5-D z-march;
 * generic != 0 forces the runtime-radix FFT kernels and the tiled 3-D stencil
 * that cover every supported size (used to cross-check the two). */
int fame_plan_set_generic(fame_plan* plan, int generic);

/* Material-stencil tile (no reference counterpart): TX x TY columns per work-group
 * ((TX+2)(TY+2) <= 256) marching TZ slices (0 = the host heuristic for TZ); tx = ty = 0
 * restores the heuristic.  fame_plan_autotune sets it. */
int fame_plan_set_material_tile(fame_plan* plan, int tx, int ty, int tz);
int fame_plan_get_material_tile(fame_plan* plan, int* tx, int* ty, int* tz);

/* CG schedule (no reference counterpart; same iterates up to rounding): fused = 1
 * (default) folds the CG vector updates and the p^H q reduction into the first and
 * last pass of fncB and advances all k-points in lock step, chunk by chunk;
 * fused = 1 + c runs chunks of c k-points to convergence one after another (their
 * working set stays in the last-level cache); 0 runs the textbook loop with separate
 * dot / update / direction kernels.  The fused forms need the specialised z kernels
 * and fall back to the textbook loop otherwise. */
int fame_plan_set_cg_fused(fame_plan* plan, int fused);

/* on != 0: the fused CG captures one iteration (all k-point chunks) as a HIP graph and
 * replays it between convergence polls; 0 (default) launches the kernels one by one.
 * Same kernels, same results (no reference counterpart). */
int fame_plan_set_cg_graph(fame_plan* plan, int on);

/* Chunk schedule on 1, 2 or 3 streams (no reference counterpart; same results bit for
 * bit): fncB and each lock-step fused CG iteration queue chunk k on stream k % nstreams
 * (the plan's stream or a plan-owned extra one) with its own part of the field scratch, so
 * one chunk's kernels fill the others' launch tails.  Every stream starts after the work
 * already queued on the plan's stream, which waits for all of them before the call returns.
 * Off while profiling.  fame_plan_autotune chooses it with the chunk size. */
int fame_plan_set_streams(fame_plan* plan, int nstreams);
int fame_plan_get_streams(fame_plan* plan, int* nstreams);

/* Measurement aid (no reference counterpart): with enable != 0 every fame_fncB
 * records HIP events on the plan's stream between its five kernels
 * (z-pass+Pi mix, xy-pass inverse, material stencil, xy-pass forward,
 * z-pass+Pi^H mix) and accumulates their durations; each call then waits for
 * its own completion.  fame_plan_stage_times returns the accumulated
 * milliseconds per stage (ms[5]) and the number of profiled calls. */
int fame_plan_profile(fame_plan* plan, int enable);
int fame_plan_stage_times(fame_plan* plan, double* ms, long long* calls);
/* Measurement aid (no reference counterpart): while profiling, the material stencil of the
 * specialised path is launched with its own start / stop events (hipExtLaunchKernelGGL), which
 * time the kernel alone as a kernel trace does; returns the accumulated milliseconds and the
 * number of timed launches. */
int fame_plan_material_kernel_time(fame_plan* plan, double* ms, long long* launches);

/* Psi / Phi ("notranspose", istrans = 0) and Psis / Phis ("transpose", istrans = 1)
 * on a 3-component field [nb][3N]:  y_c = S .* (ifftn(x_c) * sqrt(N)),
 * resp. y_c = fftn(conj(S) .* x_c) / sqrt(N).
 * Replaces subroutines/gMtxVecMul_FFT.m:1-61 (MtxVecMul_FFT.m:1-65) for QP
 * boundary conditions, where Psi == Phi. */
int fame_MtxVecMul_FFT(fame_plan* plan, int istrans, const void* in, void* out);

/* fnc.Pr / fnc.Qr : [nb][2N] -> [nb][3N];  fnc.Prs / fnc.Qrs : [nb][3N] -> [nb][2N].
 * Replace fun_Lanczos/Funhand_Generate_Lanzcos.m:79-82. */
int fame_Pr(fame_plan* plan, const void* in, void* out);
int fame_Qr(fame_plan* plan, const void* in, void* out);
int fame_Prs(fame_plan* plan, const void* in, void* out);
int fame_Qrs(fame_plan* plan, const void* in, void* out);

/* fncA: [nb][4N] -> [nb][4N].  Replaces fun_Lanczos/fncA.m:3-4. */
int fame_fncA(fame_plan* plan, const void* in, void* out);

/* fncB: [nb][4N] -> [nb][4N].  Replaces fun_Lanczos/fncB.m:3-6
 * (and fun_no_inv/fncB_Generate_no_inv.m:6-8). */
int fame_fncB(fame_plan* plan, const void* in, void* out);

/* Interpolated material sandwich on twisted real-space fields [nb][6N] = [h (3N); e (3N)]:
 * out = [M_int h + J_int^H e; J_int h + N_int e]  (fun_Lanczos/fncB.m:4-5). */
int fame_material_apply(fame_plan* plan, const void* in, void* out);

/* fncinvB: y = B \ x by unpreconditioned CG from x0 = 0 for all nb items in lock
 * step; stops when ||x - B y|| <= tol ||x|| for every item.  iters[nb] and
 * relres[nb] (may be NULL) receive per-item iteration counts and the final
 * recursive relative residuals.  Replaces fun_Lanczos/fncinvB.m:4 (MATLAB pcg).
 * Returns FAME_ERR_NOT_CONVERGED (result still written) if maxit is reached. */
int fame_fncinvB(fame_plan* plan, const void* in, void* out, double tol, int maxit, int* iters, double* relres);

/* Thick-restart B-orthogonal Lanczos on B^{-1} A (fun_Lanczos/Lanczos.m:1-156)
 * with fncA / fncB / fncinvB(cg_tol, cg_maxit) as operators, for all nb
 * k-points in lock step.
 *   k        wanted eigenvalues per item, p Krylov dimension (p = 2k in the reference)
 *   v0       [nb][4N] device start vectors (B-normalised inside, Lanczos.m:8)
 *   ev       [nb][k][4N] device output (eigenvectors, column-major per item)
 *   ew       [nb][k] host output: eigenvalues d of B^{-1}A (omega = 1/d)
 *   iters    [nb] host output: Lanczos steps taken (may be NULL); negated (-steps) for an item
 *            that did not converge (maxiter restarts reached, or an inner CG solve of a step it
 *            took part in reached cg_maxit)
 * Returns FAME_ERR_NOT_CONVERGED when any item did not converge; ev and ew are still written for
 * every item (the converged ones are final, the others hold the last Ritz pairs, as Lanczos.m
 * returns its iterates at mm == maxiter). */
int fame_lanczos(fame_plan* plan, int k, double tol, int maxiter, int p, const void* v0, double cg_tol, int cg_maxit,
                 void* ev, double* ew, int* iters);

/* MINRES on (fncB - sigma fncA) y = x for all nb items in lock step, from y0 = 0;
 * stops when the MINRES residual estimate |eta| <= tol ||x|| for every item.
 * Replaces the inner solve `minres1(@(vec)GEP_shift_mtx_prod_vec(vec, A, B, sigma), rhs, tol, maxit)`
 * of light_locked_Huang/SIRA_solve_residual_LS_GEP.m:54 (with A = fncB, B = fncA as
 * the JD-SIRA drivers pass them); sigma = 0 is the reference's operator
 * (Jacobi_Davidson/GEP_shift_mtx_prod_vec.m:28 applies A only).
 * Returns FAME_ERR_NOT_CONVERGED (result still written) if maxit is reached. */
int fame_minres(fame_plan* plan, const void* in, void* out, double sigma, double tol, int maxit, int* iters,
                double* relres);

/* Tall-skinny projections for Rayleigh-Ritz and Gram-Schmidt on one vector set
 * (light_locked_Huang/GEP_AB_Herm_JDSIRA_Locking4.m:93,97,172,326-340):
 *   fame_block_dot : out[j] = V_j^H w, j < m  (V: m columns of len, column-major; out on the device)
 *   fame_block_gemv: out = beta * out + V y   (y: m complex on the device) */
int fame_block_dot(fame_plan* plan, const void* V, int m, const void* w, long long len, void* out);
int fame_block_gemv(fame_plan* plan, const void* V, int m, long long len, const void* y, double beta, void* out);

/* fame_block_dot without any allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate (HIP graph capture, tight serving loops).  The per-work-group partial sums
 * that fame_block_dot allocates (a plan-owned buffer, on first use and whenever a larger
 * projection comes) live in a file-scope device array of 2^18 complex128
 * (4 MiB per GPU, held until the process ends), so m * ceil(len / 1024) <= 262144, else
 * FAME_ERR_ARG before any device work.  Calls that overlap on one GPU share that array and race. */
int fame_block_dot_noalloc(fame_plan* plan, const void* V, int m, const void* w, long long len, void* out);

/* Modified Gram-Schmidt of t (len) against the m columns of V (column-major, len each), in
 * column order, on the device: c_j = V_j^H t; t <- t - c_j V_j for j = 0..m-1
 * (light_locked_Huang/GEP_AB_Herm_JDSIRA_Locking4.m:326-330, the orthogonalisation of the
 * correction vector).  Each update is fused with the next inner product; the c_j stay on
 * the device (written to coef, m complex, when coef is not NULL).  Asynchronous on the
 * plan's stream, no host synchronisation. */
int fame_mgs(fame_plan* plan, const void* V, int m, long long len, void* t, void* coef);

/* One JD-SIRA outer step in two calls (GEP_AB_Herm_JDSIRA_Locking4.m, A = fncB, B = fncA,
 * one k-point plan, len = 4N, V: device columns of len, column-major):
 *   fame_jd_ritz_residual: u = V y / ||V y|| for the m host coefficients y (complex), and
 *     r = A u - theta B u with rnorm = ||r|| on the host (Locking4.m:170-178); u, r: device, len.
 *   fame_jd_expand: t = MINRES on (A - sigma B) t = r (tol, maxit; fame_minres), rescaled when
 *     ||t|| < 1e-9, orthogonalised against the m columns of V by MGS in column order and
 *     normalised into column m of V (Locking4.m:311-330); proj (host, 2(m+1) complex) receives
 *     V^H A t and V^H B t over the m + 1 columns (Locking4.m:335-341); iters: MINRES steps.
 * Each call synchronises the plan's stream once, for its host result. */
int fame_jd_ritz_residual(fame_plan* plan, const void* V, int m, long long len, const double* y, double theta_re,
                          double theta_im, void* u, void* r, double* rnorm);
int fame_jd_expand(fame_plan* plan, void* V, int m, long long len, const void* r, double sigma, double tol, int maxit,
                   int* iters, double* proj);

/* Eigen_Restoration_Chiral_no_inv (fun_no_inv/Eigen_Restoration_Chiral_no_inv.m:3-25):
 * for nvec NFSEP vectors per item ([nb][nvec][4N]) returns [nb][nvec][6N] =
 * [e (3N); h (3N)] with e = -i (N Qr(v_b) + J Pr(v_a)), h = -i (J^H Qr(v_b) + M Pr(v_a)). */
int fame_eigen_restoration(fame_plan* plan, int nvec, const void* ev, void* e_h);

/* Batched complex dot products out[b] = x[b]^H y[b] over len elements,
 * written to device memory (nb complex). */
int fame_dot(fame_plan* plan, const void* x, const void* y, long long len, void* out);

/* Per-vertex bi-isotropic material, replacing the id table of fame_plan_create:
 * eps_v I, mu I, xi = i gamma_v I, zeta = -i gamma_v I at every vertex (host arrays
 * eps[N], gamma[N], x fastest), lattice vectors as the columns of the row-major 3x3
 * `lattice`.  The slots are built as in Physical_settings_woodpile.m:91-96,184-194 and
 * Public/Material_Matrix_Generate_Mult_Point.m:28-31 (per-vertex A{i,j} instead of
 * inside/outside values).  Needs eps*mu - gamma^2 > 0 everywhere. */
int fame_plan_set_vertex_material(fame_plan* plan, const double* eps, const double* gamma, double mu,
                                  const double* lattice);

/* Adjoint of the material apply (no reference counterpart: the differentiation the
 * reference does not have).  For each of nvec NFSEP vectors x (device, [nb][nvec][4N])
 * writes grad[nb][nvec][3][N] (device, double):
 *   grad[.][j][0][v] = d(x^H fncB x)/d eps_v,  grad[.][j][1][v] = d(x^H fncB x)/d gamma_v,
 *   grad[.][j][2][v] = the vertex-v term of x^H fncB x (they sum to x^H fncB x).
 * For fncB x = w fncA x, dw/dp = grad / (x^H fncA x).  Needs the per-vertex material. */
int fame_material_gradient(fame_plan* plan, int nvec, const void* ev, void* grad);

/* HBM stream probe (no reference counterpart; a measurement aid): kind 0 reads `bytes` of src,
 * 1 writes `bytes` of dst, 2 copies src to dst, with plain 16-byte streaming kernels, `reps`
 * times after one warm-up on the given hipStream_t (NULL: the default stream); ms_per_rep gets
 * the mean HIP-event time per pass.  With buffers far beyond the 256 MB Infinity Cache the rate
 * is the memory system's streaming ceiling.  bytes: a positive multiple of 16. */
int fame_stream_probe(int kind, const void* src, void* dst, long long bytes, int reps, void* stream,
                      double* ms_per_rep);

/* ---------------------------------------------------------------------------
 * FCC skewed-grid operators (the reference's light_locked_Huang/ FCC path).
 *
 * One plan per (grid, k).  Grid vectors of length n = n1*n2*n3.  The input of T
 * (and output of T^H) is in frequency order, z fastest: p = l + n3*(nn + n2*m)
 * (p_mtx_T_prod_vec_FCC.m:9 reshapes it as n3 x n1*n2); the output of T is in grid
 * order, x fastest.  Sizes: Lambda_Pr/Qr map 16n -> 24n; the NFGEP vector is 32n;
 * the parameter matrices act on 12n; GEP B on 48n.  All vector pointers are
 * device pointers; calls are asynchronous on the plan's stream.
 * ------------------------------------------------------------------------- */
typedef struct fame_fcc fame_fcc;

/* Tables of SVDSingleCurl_FCC (light_locked_Huang/SVDSingleCurl_FCC.m:1-137):
 * vec_k[3], lattice vectors a1, a2, a3 (period_vec1..3), delta[3].  The explicit
 * difference matrices this factorises need n1 even and n2 divisible by 3
 * (Construct_GEP_mtx_A.m:451-505); T itself does not. */
int fame_fcc_create(fame_fcc** plan, int n1, int n2, int n3, const double* vec_k, const double* a1, const double* a2,
                    const double* a3, const double* delta);
int fame_fcc_destroy(fame_fcc* plan);
/* stream: a hipStream_t, NULL = the default stream */
int fame_fcc_set_stream(fame_fcc* plan, void* stream);
int fame_fcc_sync(fame_fcc* plan);
/* Kernel selection (no reference counterpart): n1 = n2 = n3 with a specialised FFT length
 * uses the compile-time four-step kernels by default; generic != 0 forces the runtime-radix
 * kernels that cover every size (used to cross-check the two). */
int fame_fcc_set_generic(fame_fcc* plan, int generic);

/* Launch shape of the fused NFGEP xy kernel (no reference counterpart; same results): split = 0
 * keeps the two planes of a work-group in LDS together, split = 1 transforms them one at a time
 * (one plane of LDS, so two work-groups share a CU; the default at lengths 40-48); lengths
 * above 64 always use the first. */
int fame_fcc_set_nfgep_variant(fame_fcc* plan, int split);

/* v = p_mtx_T_prod_vec_FCC(v, FFT_parameter) on nvec vectors [nvec][n]
 * (light_locked_Huang/p_mtx_T_prod_vec_FCC.m:1-37); in == out allowed. */
int fame_fcc_T(fame_fcc* plan, int nvec, const void* in, void* out);
/* v = p_mtx_TH_prod_vec_FCC(v, FFT_parameter_bar) (light_locked_Huang/p_mtx_TH_prod_vec_FCC.m:1-28). */
int fame_fcc_TH(fame_fcc* plan, int nvec, const void* in, void* out);
/* SVD_curl.Sigma_r (SVDSingleCurl_FCC.m:66-69): 8n doubles, device. */
int fame_fcc_Sigma_r(fame_fcc* plan, void* out);
/* (I_24 (x) T) SVD_curl.Lambda_Pr u and (I_24 (x) T) SVD_curl.Lambda_Qr v, 16n -> 24n
 * (mtx_prod_vec_A_NFGEP.m:12-24); Prs/Qrs are the adjoints 24n -> 16n (:36-50). */
int fame_fcc_Pr(fame_fcc* plan, const void* in, void* out);
int fame_fcc_Qr(fame_fcc* plan, const void* in, void* out);
int fame_fcc_Prs(fame_fcc* plan, const void* in, void* out);
int fame_fcc_Qrs(fame_fcc* plan, const void* in, void* out);
/* output_vec = mtx_prod_vec_A_NFGEP(n, input_vec, SVD_curl, mtx_data, TH, T)
 * (light_locked_Huang/mtx_prod_vec_A_NFGEP.m:1-52), 32n -> 32n, with
 * mtx_data.Phi_is_diag = 'yes'.  xi_d, zeta_d, Phi: 24n complex device arrays (the
 * diagonals of mtx_data.xi_d, zeta_d and the vector mtx_data.Phi). */
int fame_fcc_nfgep_A(fame_fcc* plan, const void* in, void* out, const void* xi_d, const void* zeta_d,
                     const void* Phi);
/* ev = Compute_eigenvector(GEP_mtx_B, n, input_vec, SVD_curl, mtx_data, T), 32n -> 48n
 * (light_locked_Huang/Compute_eigenvector.m:1-43), diagonal Phi. */
int fame_fcc_eigenvector(fame_fcc* plan, const void* in, void* out, const void* xi_d, const void* zeta_d,
                         const void* Phi);
/* The same two maps (mtx_prod_vec_A_NFGEP.m:12-50, Compute_eigenvector.m:9-41) with the
 * sparsity of mtx_data.xi_d and mtx_data.zeta_d (24n x 24n matrices in the reference, applied
 * as `mtx_data.xi_d * tmp1`) stated by `coupling`: one nonzero per row, given by the 24n arrays
 * xi_d / zeta_d, on the diagonal (coupling = 0: the calls above) or in the other 12n half
 * (coupling = 1: (xi_d t)[r] = xi_d[r] t[(r + 12n) mod 24n]), which is the form the physical
 * GEP B takes (fame_fcc_nfgep_data). */
int fame_fcc_nfgep_A_ex(fame_fcc* plan, const void* in, void* out, const void* xi_d, const void* zeta_d,
                        const void* Phi, int coupling);
int fame_fcc_eigenvector_ex(fame_fcc* plan, const void* in, void* out, const void* xi_d, const void* zeta_d,
                            const void* Phi, int coupling);
/* mtx_data (xi_d, zeta_d, Phi; 24n complex device arrays each, coupling = 1) for the material of
 * fame_fcc_set_material, derived from GEP B of Construct_GEP_mtx_B.m:1-69.  The reference never
 * builds mtx_data; this is the one for which Compute_eigenvector equals its commented check
 * `GEP_mtx_B \ [tmp1; tmp2]` (Compute_eigenvector.m:27,41) after the similarity diag(I, -iI) B
 * diag(I, iI), which leaves GEP A unchanged, so the NFGEP spectrum is the nonzero spectrum of
 * (GEP A, GEP B).  Needs mu = I and diagonal eps, zeta, xi (inside and air), else
 * FAME_ERR_UNSUPPORTED (Phi would not be diagonal: the LU branch of mtx_prod_vec_A_NFGEP.m:29-30). */
int fame_fcc_nfgep_data(fame_fcc* plan, void* xi_d, void* zeta_d, void* Phi);
/* Material of mtx_Eps_Mu_Zeta_Xi (light_locked_Huang/mtx_Eps_Mu_Zeta_Xi.m:1-38): host arrays.
 * edge_mask[n]: bit 0/1/2 = point in edge_idx_in.X/Y/Z, bit 3 = in Cubic_idx_in;
 * face_mask[n]: bit 0/1/2 = face_idx_in.X/Y/Z, bit 3 = vertex_idx_in.
 * tensors: [2][4][3][3] complex, [inside, air][eps, mu, zeta, xi], row-major 3x3
 * (mtx_par_* and val_*_air). */
int fame_fcc_set_material(fame_fcc* plan, const unsigned char* edge_mask, const unsigned char* face_mask,
                          const double* tensors);
/* out = M in for one of the eight matrices of mtx_Eps_Mu_Zeta_Xi.m:23-35
 * (Construct_parameter_mtx.m:1-318), 12n -> 12n; which = 2*q + f with q = 0 eps,
 * 1 mu, 2 zeta, 3 xi and f = 0 edge, 1 face. */
int fame_fcc_param_apply(fame_fcc* plan, int which, const void* in, void* out);
/* out = GEP_mtx_B in, 48n (light_locked_Huang/Construct_GEP_mtx_B.m:1-69). */
int fame_fcc_gep_B(fame_fcc* plan, const void* in, void* out);

#ifdef __cplusplus
}
#endif
#endif /* FAME_ABI_H */
