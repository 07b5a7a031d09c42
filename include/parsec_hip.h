/* parsec_hip.h -- C-ABI of libparsec_hip.so, the MI355X (gfx950) kernels of the
 * PARSEC.py SCF hot path.
 *
 * Conventions
 *   - Every entry point returns int: 0 = success, 1 = bad argument,
 *     2 = HIP runtime error, 3 = dense solver failure, 4 = size limit.
 *     pc_last_error() returns the text of the last failure (per host thread).
 *   - Device operands are plain device pointers owned by the caller; `stream`
 *     is a hipStream_t passed as void*.  Calls are asynchronous on that stream
 *     unless stated otherwise.
 *   - A "block" of k vectors on the domain is vector-major: element (p, v) at
 *     v*ld + p, p < npts (the reference's N x k array W in column-major order).
 *   - The domain is a set of x-rows of an nx*ny*nz grid, row (j,k) holding the
 *     contiguous points row_lo <= i < row_hi; points are numbered k-major, then
 *     j, then i (the reference's flat order i + nx*j + nx*ny*k restricted to the
 *     domain).
 *
 * Each declaration names the reference interface (zeyizhang-berkeley/
 * parsec_python, path:line) whose work it replaces.
 */
#ifndef PARSEC_HIP_H
#define PARSEC_HIP_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* pc_last_error(void);
int pc_version(void);

/* ---- grid and the fused stencil -------------------------------------------
 * replaces: src/Laplacian/fd3d.py:34-106 (fd3d) and every `H @ X` SpMM of
 * src/Eigensolvers/ch_filter.py:66-80, first_filt.py:94-101, chefsi1.py:46,
 * lanczosForChefsi1.py:41 (CPU) / chefsi1_gpu.py:51-57 (CuPy CSR SpMM).
 * weights[0..radius]: 1D weights of -d2/dx2 (centre first), radius in {1,2,3,4,6}.
 * out = sum_d scale*w_d*(6 neighbours) + 3*scale*w_0*y + (a*V + b)*y + c*x.
 * x may alias out; y must not.  dot_part (optional) receives per-workgroup
 * sums of out*y (see pc_stencil_num_partials). */
int pc_grid_create(int nx, int ny, int nz, const int* row_lo, const int* row_hi, int radius,
                   const double* weights, void** out_grid, long long* out_npts);
int pc_grid_destroy(void* grid);
int pc_stencil_f64(void* grid, const double* Y, const double* X, double* Out, const double* V, int nvec,
                   long long ld, double scale, double a, double b, double c, double* dot_part, void* stream);
int pc_stencil_f32(void* grid, const float* Y, const float* X, float* Out, const float* V, int nvec,
                   long long ld, double scale, double a, double b, double c, double* dot_part, void* stream);
/* stencil kernel for radius-4 grids of at most 128 points per row
 * (process-wide): 0 = fixed 64-point x-chunks per wavefront, 1 = the lanes
 * dealt the slab's actual columns, 2 = the latter for FP32 launches of fewer
 * than 128 vectors that fill the chip without z-chunks (default); the same
 * results.  No counterpart in the reference. */
int pc_stencil_set_variant(int variant);
int pc_stencil_direct_f64(void* grid, const double* Y, const double* X, double* Out, const double* V, int nvec,
                          long long ld, double scale, double a, double b, double c, double* dot_part, void* stream);
int pc_stencil_direct_f32(void* grid, const float* Y, const float* X, float* Out, const float* V, int nvec,
                          long long ld, double scale, double a, double b, double c, double* dot_part, void* stream);
int pc_stencil_num_partials(void* grid, int dtype_bytes, int nvec, int* out_count, int* out_per_vector);

/* ---- Kleinman-Bylander projectors (factored) --------------------------------
 * replaces: src/V_ion/pseudoNL_original.py:9-208 (V_nl as dense outer
 * products in a sparse matrix, applied inside every SpMM).
 * atom_ptr[natoms+1] groups entries; ent_pt[nent] point indices; ent_u[3*nent]
 * (u_x,u_y,u_z); scale[natoms] = h^3/Zvalue.  C is FP64, atom-major:
 * C[(a*nvec + v)*3 + m], natoms*nvec*3 doubles. */
int pc_proj_create(int natoms, const int* atom_ptr, const int* ent_pt, const double* ent_u, const double* scale,
                   long long npts, void** out_proj);
int pc_proj_destroy(void* proj);
int pc_proj_ncols(void* proj, int* out_ncols);
int pc_proj_gather_f64(void* proj, const double* Y, long long ld, int nvec, double* C, void* stream);
int pc_proj_gather_f32(void* proj, const float* Y, long long ld, int nvec, double* C, void* stream);
int pc_proj_scatter_f64(void* proj, const double* C, double* Out, long long ld, int nvec, double alpha,
                        void* stream);
int pc_proj_scatter_f32(void* proj, const double* C, float* Out, long long ld, int nvec, double alpha,
                        void* stream);

/* ---- Rayleigh-Ritz GEMMs on MFMA -------------------------------------------
 * replaces: the vdot double loop src/Eigensolvers/chefsi1.py:51-53,
 * first_filt.py:53-57 (CuPy: W.T @ Vin, chefsi1_gpu.py:61) and W @ Q
 * (chefsi1.py:67, chefsi1_gpu.py:66).
 * gram: S (kx x ky, column-major, lds) = X^T Y in FP64 (f32 inputs: FP32 MFMA
 * per chunk, FP64 across chunks); sym=1 computes the upper tiles of X^T X.
 * rotate: Z = X Q, Q FP64 kx x ky column-major (ldq); work >= kx*ky*sizeof(T). */
int pc_gram_workspace(long long n, int kx, int ky, int sym, size_t* out_bytes);
int pc_gram_f64(const double* X, long long ldx, const double* Y, long long ldy, long long n, int kx, int ky,
                double* S, long long lds, int sym, void* work, size_t work_bytes, void* stream);
int pc_gram_f32(const float* X, long long ldx, const float* Y, long long ldy, long long n, int kx, int ky,
                double* S, long long lds, int sym, void* work, size_t work_bytes, void* stream);
int pc_rotate_f64(const double* X, long long ldx, const double* Q, int ldq, double* Z, long long ldz, long long n,
                  int kx, int ky, void* work, size_t work_bytes, void* stream);
int pc_rotate_f32(const float* X, long long ldx, const double* Q, int ldq, float* Z, long long ldz, long long n,
                  int kx, int ky, void* work, size_t work_bytes, void* stream);
/* workspace of pc_rotate_f32 (fp32 = 1) or pc_rotate_f64, including the split
 * Q of the optional split-bf16 FP32 rotation.  No counterpart in the
 * reference (its W @ Q allocates). */
int pc_rotate_workspace(int kx, int ky, int fp32, size_t* out_bytes);
/* process-wide switch of the FP32 rotation on bf16 MFMA with three-way operand
 * splits (1 = on; off by default: faster, but it cost the Si87H76 SCF four
 * iterations where the rotation matrix cancels).  No counterpart in the
 * reference. */
int pc_rotate_set_split_bf16(int on);

/* ---- FP64 dense k x k stage on the GPU (rocSOLVER/rocBLAS) ------------------
 * replaces: scipy.linalg.qr + eigh in src/Eigensolvers/chefsi1.py:43,64 and
 * first_filt.py:43,59 (cuSOLVER qr/eigh in chefsi1_gpu.py:54,65).
 * pc_dense_rr: S = R^T R (potrf), G = R^-T Gy R^-1, G = Q L Q^T (syevd),
 * T = R^-1 Q; lam_host receives the ascending Ritz values.  Synchronises. */
int pc_dense_create(void* stream, int kmax, void** out_ctx);
int pc_dense_destroy(void* ctx);
int pc_dense_rr(void* ctx, int k, double* S, int lds, double* Gy, int ldg, double* T, int ldt, double* lam_dev,
                double* lam_host);
int pc_dense_syevd(void* ctx, int k, double* A, int lda, double* w_host);
/* eigensolver used by pc_dense_rr: 0 = rocSOLVER divide and conquer (syevd),
 * 1 = Jacobi (syevj), 2 = one-launch Householder tridiagonalization + the
 * tridiagonal eigensolver of this library (multisection bisection, inverse
 * iteration, Gram-Schmidt and Rayleigh-Ritz within clusters of close
 * eigenvalues; rocSOLVER stedc for a cluster of more than 32) + Q Z
 * (pc_dense_set_backtransform) (default; 32 workgroups exchanging through agent-scope atomics for
 * 128 <= k <= 320, one workgroup up to 512, larger k uses syevd), 3 = as 2
 * with the one-workgroup reduction always, 4 = as 2 with stedc always */
int pc_dense_set_method(void* ctx, int method);
/* the blocked Cholesky factor and inverse of S on the context's second stream,
 * after the work queued on the main stream; the next pc_dense_rr with this S
 * (unchanged until then) uses it, so the factor overlaps the work queued in
 * between.  Nothing is queued outside the blocked mode (k > 320, another
 * triangular mode).  No counterpart in the reference. */
int pc_dense_chol_async(void* ctx, int k, const double* S, int lds);
/* S = Y^T Y (pc_gram_f32 / pc_gram_f64, sym) and then the blocked Cholesky
 * factor and inverse, both on the context's second stream after the work
 * queued on the main stream; the next pc_dense_rr waits for them and uses the
 * factor.  work: a Gram workspace of its own (pc_gram_workspace); Y, S and
 * work unchanged until then.  No counterpart in the reference. */
int pc_dense_gram_chol_async(void* ctx, int fp32, const void* Y, long long ldy, long long n, int k, double* S,
                             int lds, void* work, size_t work_bytes);
/* how pc_dense_rr forms the Cholesky factor and applies R^-1: 0 = rocSOLVER
 * potrf + trtri + trmm, 1 = potrf + trsm in place, 2 = this library's blocked
 * Cholesky factor and inverse (two launches) + GEMMs (default; k > 320 uses 0).
 * No counterpart in the reference: its QR gives W directly. */
int pc_dense_set_triangular(void* ctx, int mode);
/* how pc_dense_rr (methods 2-4) applies the tridiagonalization's reflectors to
 * the tridiagonal eigenvectors, Q Z: 1 = this library's compact-WY kernels
 * (default; k > 320 uses 0), 0 = rocSOLVER's ormtr.  No counterpart in the
 * reference (its eigh is scipy's). */
int pc_dense_set_backtransform(void* ctx, int mode);
/* steps repeated with stedc because a cluster was too large (no counterpart
 * in the reference; for tests) */
int pc_dense_tri_fallbacks(void* ctx, int* out);
/* test control of the multi-workgroup reduction: spins (> 0) = polls of one
 * wait before it gives up and pc_dense_rr repeats the reduction with one
 * workgroup; reset = 1 re-enables it; out_aborts = reductions repeated so far.
 * No counterpart in the reference. */
int pc_dense_trd_control(void* ctx, int spins, int reset, int* out_aborts);
/* pc_dense_rr repeats a step whose overlap matrix is not numerically positive
 * definite with S + shift*max(diag S)*I, shift 1e-13, 1e-11, 1e-9, 1e-7 (the
 * reference's Householder QR in src/Eigensolvers/chefsi1.py:43 does not fail on
 * nearly dependent columns); out = steps of this context that needed it. */
int pc_dense_shift_count(void* ctx, int* out);

/* ---- point-wise SCF kernels --------------------------------------------------
 * density: rho = sum_v occ2[v] W_v^2           (src/rsdft_solver.py:324)
 * scf_post: LDA PZ81 v_xc, pot_new = ppot + 0.5 vxc + hpot + hpot0 and the
 *   sums [E_xc/h^3, |pot_new-pot|^2, |pot_new|^2, sum rho(hpot+hpot0),
 *   sum rho vxc]                                (src/V_xc/exc_nspn.py:8-103;
 *   src/rsdft_solver.py:128-153,350-352)
 * hartree_rhs: rhs = fac (rho - rho0)            (src/rsdft_solver.py:327)
 * dots / lincomb: the multisecant mixer's GEMV work (src/Mixer/msecant1.py:104-167)
 * rand: counter-based (splitmix64) uniform / normal start blocks
 *   (replaces np.random.rand/randn, src/Eigensolvers/first_filt.py:37,
 *   lancz_uppbnd.py:21, chefsi1.py:29). */
int pc_density_f64(const double* W, long long ld, long long n, int nev, const double* occ2, double* rho,
                   void* stream);
int pc_density_f32(const float* W, long long ld, long long n, int nev, const double* occ2, double* rho,
                   void* stream);
int pc_num_blocks(long long n, int* out);
int pc_scf_post(long long n, const double* rho_cell, double h3, const double* ppot, const double* hpot,
                const double* hpot0, const double* pot_old, double* rho_out, double* vxc_out, double* pot_new,
                double* part, double* sums5, void* stream);
int pc_xc(long long n, const double* rho, double* vxc, double* part, double* exc_sum, void* stream);
int pc_hartree_rhs(long long n, const double* rho_cell, const double* rho0, double fac, double* rhs, void* stream);
int pc_reduce_sets(const double* part, int nblk, int nsets, double* out, void* stream);
int pc_dots_nchunk(long long n, int* out);
int pc_dots(long long n, const double* A, long long lda, int m, const double* v, double* part, double* out,
            void* stream);
int pc_lincomb(long long n, const double* A, long long lda, int m, const double* coef_host, double alpha,
               const double* x, double beta, double* y, void* stream);
/* pc_lincomb with its scalars in device memory (pc_dots results): coef[i] =
 * coef_scale * coef_dev[i]; alpha = 1 / alpha_recip_dev[0] (0 when
 * |v| < alpha_tol |v|) when alpha_recip_dev is set; beta = beta_scale *
 * beta_dev[0] when beta_dev is set.  Lets the multisecant mixer
 * (src/Mixer/msecant1.py:40-175, whose numpy dots feed the next update
 * directly) run its update with one host readback instead of seven. */
int pc_lincomb_dev(long long n, const double* A, long long lda, int m, const double* coef_dev, double coef_scale,
                   double alpha, const double* alpha_recip_dev, double alpha_tol, const double* x, double beta,
                   const double* beta_dev, double beta_scale, double* y, void* stream);
/* Lanczos bound without host round trips (replaces the vdot / norm / axpy
 * sequence of src/Eigensolvers/lanczosForChefsi1.py:5-99): pc_lanczos_start
 * normalises v; after w = H v1, pc_lanczos_step(k) computes alpha_k = v1.w,
 * w -= alpha_k v1 + beta_{k-1} v0, beta_k = |w|, w /= beta_k and stores
 * coef[2k] = alpha_k, coef[2k+1] = beta_k on the device.  part: 2 *
 * pc_lanczos_nblocks(n) doubles. */
int pc_lanczos_nblocks(long long n, int* out);
int pc_lanczos_start(long long n, double* v, double* part, void* stream);
int pc_lanczos_step(long long n, const double* v0, const double* v1, double* w, double* coef, int k, double* part,
                    void* stream);
int pc_convert_f64_to_f32(long long n, int ncol, const double* src, long long lds, float* dst, long long ldd,
                          void* stream);
int pc_convert_f32_to_f64(long long n, int ncol, const float* src, long long lds, double* dst, long long ldd,
                          void* stream);
int pc_rand_f64(long long n, int ncol, long long ld, unsigned long long seed, int normal, double* out, void* stream);
int pc_rand_f32(long long n, int ncol, long long ld, unsigned long long seed, int normal, float* out, void* stream);

/* ---- Hartree CG ---------------------------------------------------------------
 * replaces: src/Eigensolvers/pcg.py:7-85 (CPU) / pcg_gpu.py:22-65 (CuPy).
 * Solves (-lap/h^2) x = rhs, FP64, warm start x, stop when r.r <= tol^2 r0.r0
 * or maxit; alias_first=1 reproduces the CPU reference's first-iteration p/r
 * aliasing.  Two launches per iteration (stencil + direction update fused;
 * x/r update), no separate reductions.  Synchronises every check_every
 * iterations.  res_out (host, may be NULL) receives {final r.r, tol^2 r0.r0}. */
int pc_pcg_workspace(void* grid, size_t* out_bytes);
int pc_pcg_solve(void* grid, double inv_h2, const double* rhs, double* x, int maxit, double tol, int alias_first,
                 int check_every, void* work, size_t work_bytes, int* its_out, double* res_out, void* stream);
/* The same solve without any allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The workspace is a fixed
 * device array inside the library (64 MiB per device, held until the process
 * ends), so the grid may hold at most pc_pcg_noalloc_max_points() points
 * (2^21 = 128^3); larger grids return 4 (size limit).  Results are those of
 * pc_pcg_solve, bit for bit.  Every call on one device shares that array:
 * calls that overlap on one device (other host threads or streams) race. */
int pc_pcg_solve_noalloc(void* grid, double inv_h2, const double* rhs, double* x, int maxit, double tol,
                         int alias_first, int check_every, int* its_out, double* res_out, void* stream);
int pc_pcg_noalloc_max_points(long long* out);

/* Multigrid-preconditioned CG on the same system and stop rule (replaces the
 * reference's optional preconditioned CG, cg_prec in src/Eigensolvers/pcg.py:7-85,
 * whose ILU preconditioner has no GPU form): one symmetric cell-centred V-cycle
 * per iteration (damped Jacobi on the domain's stencil and on 7-point cube
 * levels, Chebyshev on the coarsest cube).  Textbook PCG, no aliasing. */
int pc_mg_create(void* grid, double h, void** out_mg);
int pc_mg_destroy(void* mg);
int pc_mg_workspace(void* mg, size_t* out_bytes);
int pc_pcg_solve_mg(void* mg, const double* rhs, double* x, int maxit, double tol, int check_every, void* work,
                    size_t work_bytes, int* its_out, double* res_out, void* stream);
/* iterations before pc_pcg_solve_mg first reads its state back (0, the
 * default: its check_every); later readbacks follow every check_every.  A
 * caller that knows about how many iterations a solve takes (the previous SCF
 * iteration's) queues that many before the first host round trip instead of
 * queueing past convergence.  Results do not depend on it.  No counterpart in
 * the reference (its CG is a host loop). */
int pc_mg_set_first_check(void* mg, int first);

/* ---- ionic setup ------------------------------------------------------------------
 * replaces: the per-point spline loops of src/V_ion/pseudoDiag.py:199-278
 * (fsplevalIO, src/Splines/fsplevalIO.py:2-65). */
int pc_sad_eval(void* grid, double h, double rad, int natoms, const double* pos, const int* species,
                const int* tab_off, const int* tab_len, const double* xi, const double* z, const double* c,
                const double* d, double* pot, double* rho, double* hpot0, void* stream);
int pc_spline_eval(const double* xi, const double* z, const double* c, const double* d, int n, const double* r,
                   long long m, double* out, void* stream);

/* ---- learned density initialiser: 3D U-Net kernels -----------------------------
 * The reference has no network code: its ML start reads a density file
 * (src/V_ion/pseudoDiag_ML4Den_poisson.py:392-540).  These kernels train and
 * run the network that produces that density.  Tensors are BF16 (raw 16-bit
 * patterns) or FP32, channel-last: (voxel v, channel c) at v*ld + c,
 * v = (z*H + y)*W + x.  Weights W[co][t][ci], t = (dz+1)*9 + (dy+1)*3 + (dx+1).
 * conv3d_fwd: 3x3x3, zero padding, bf16 MFMA with FP32 accumulation;
 *   cin = 1 or a multiple of 16; mode 0 = FP32 out + bias, 1 = leaky ReLU ->
 *   BF16 out, 2 = FP32 out without bias.
 * conv3d_wflip: weights of the input-gradient convolution (taps reversed,
 *   channels transposed).
 * conv3d_wgrad: dW[co][t][ci] = sum_v dZ[v][co] X[v + t][ci] (FP32).
 * act_bwd: dZ = dA * leaky'(a) (a NULL: identity) rounded to BF16, and the
 *   bias gradient sum_v dZ.
 * pool2 / up2: 2x2x2 average pooling and nearest upsampling, with backward.
 * adam: bias-corrected Adam on FP32 master weights, writes the BF16 copy.
 * mse: dy = 2 (y - t) / n and sum (y - t)^2 (FP64, device). */
int pc_unet_conv3d_fwd(const void* X, int ldx, int cin, const void* W, const float* bias, int cout, int D, int H,
                       int Wd, int mode, float slope, void* Y, int ldy, void* stream);
int pc_unet_conv3d_wflip(const void* W, int cout, int cin, void* Wt, void* stream);
int pc_unet_wgrad_workspace(int cout, int cin, size_t* out_bytes);
int pc_unet_conv3d_wgrad(const void* dZ, int ldz, int cout, const void* X, int ldx, int cin, int D, int H, int Wd,
                         float* dW, void* work, size_t work_bytes, void* stream);
int pc_unet_act_bwd_workspace(int C, long long V, size_t* out_bytes);
int pc_unet_act_bwd(const float* dA, int lda, const void* a, int lda_a, int C, int D, int H, int Wd, float slope,
                    void* dZ, int ldz, float* db, void* work, size_t work_bytes, void* stream);
int pc_unet_pool2_fwd(const void* X, int ldx, int C, int D, int H, int Wd, void* Y, int ldy, void* stream);
int pc_unet_pool2_bwd(const float* dY, int ldy, const float* dS, int lds, int C, int D, int H, int Wd, float* dX,
                      int ldx, void* stream);
int pc_unet_up2_fwd(const void* X, int ldx, int C, int D, int H, int Wd, void* Y, int ldy, void* stream);
int pc_unet_up2_bwd(const float* dY, int ldy, int C, int D, int H, int Wd, float* dX, int ldx, void* stream);
int pc_unet_cast_bf16(const float* src, long long n, void* dst, void* stream);
int pc_unet_adam(float* w, const float* g, float* m, float* v, long long n, double lr, double b1, double b2,
                 double eps, double wd, int step, void* wbf, void* stream);
int pc_unet_mse_workspace(size_t* out_bytes);
int pc_unet_mse(const float* y, const float* t, long long n, float* dy, double* loss_dev, void* work,
                size_t work_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PARSEC_HIP_H */
