/* qmx — C ABI of the MI355X (gfx950) plane-wave band-solve kernels.
 *
 * Plain pointers, explicit int64 sizes and strides, an explicit HIP stream
 * (passed as void*, i.e. a hipStream_t), and an int status (0 = OK) on every
 * entry point. No C++ exception crosses this boundary; on failure
 * qmx_last_error() returns a message for the calling thread. The library
 * never frees caller memory; workspace is caller-provided and sized by
 * qmx_plan_workspace_bytes(). Kernels are re-entrant; one plan must not be
 * used by two streams at once.
 *
 * Layout conventions (those of QimPy's Wavefunction and Basis):
 *   - a wavefunction block is complex (interleaved re, im), element
 *     (k, b, g) at ptr[k * ldk + b * ldb + g], g < n_tot;
 *   - the FFT box is C-ordered (n0, n1, n2) with the last axis fastest;
 *   - fft_index[k * n_tot + g] is the linear box index of plane wave g of
 *     k-point k (dft/electrons/_basis.py:186-193); entries g >= n[k] are
 *     padding and are ignored on input and zeroed on output.
 */
#ifndef QMX_H
#define QMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element types of wavefunction data. */
enum { QMX_C64 = 0, QMX_C128 = 1 };

/* Status codes. */
enum {
  QMX_OK = 0,
  QMX_ERR_ARG = 1,    /* invalid argument (shape, null pointer, size) */
  QMX_ERR_HIP = 2,    /* HIP runtime error */
  QMX_ERR_NOMEM = 3,  /* workspace too small / allocation failed */
  QMX_ERR_UNSUPPORTED = 4
};

/* Geometry of one plane-wave basis (all host pointers). */
typedef struct qmx_geom {
  int32_t shape[3];          /* FFT box; each a product of 2, 3 and 5, <= 512 */
  int32_t nk;                /* number of k-points sharing the box */
  int32_t n_tot;             /* padded basis length */
  const int32_t* n;          /* [nk] in-sphere counts */
  const int64_t* fft_index;  /* [nk * n_tot] linear box indices */
  const int32_t* v_slot;     /* [nk] potential box used by each k, or NULL (all 0) */
  const int64_t* const* fft_rows; /* [nk] start of each k's n[k] indices, or NULL (rows
                                     of fft_index; with fft_rows, fft_index may be NULL) */
} qmx_geom;

typedef struct qmx_plan qmx_plan;

/* Host-only: the sparse-FFT index tables a plan uploads, for inspection
 * and CPU checks (no device calls). sizes[0] = nxa (active x-planes),
 * sizes[1] = ncol (active (x, y) columns). Any table pointer may be NULL;
 * call once with NULL tables to learn the sizes. Tables:
 *   x_list[nxa]            box x of each active plane (ascending)
 *   pcol_start[nxa + 1]    first column of each plane
 *   col_iy[ncol]           box y of each column
 *   cstart[nk * (ncol+1)]  first basis index of each column, per k
 *   t1idx[nk * n_tot]      column * shape[2] + z of each in-sphere point
 * The column grouping relies on fft_index increasing within each k, which
 * the reference's argsort construction guarantees (_basis.py:186-193). */
int qmx_geom_tables(const qmx_geom* geom, int64_t* sizes, int32_t* x_list,
                    int32_t* pcol_start, int32_t* col_iy, int32_t* cstart,
                    int32_t* t1idx);

/* Build the sparse-FFT index tables of a basis on the current device.
 * Replaces the index set-up of _KernelCommon.__init__
 * (dft/electrons/_basis_ops.py:119-141) and Basis.fft_index/pad_index
 * (_basis.py:186-204). */
int qmx_plan_create(const qmx_geom* geom, qmx_plan** plan);
int qmx_plan_destroy(qmx_plan* plan);

/* Workspace bytes needed by qmx_apply_local for n_bands bands per k.
 * Any smaller (nonzero) workspace is accepted as long as it holds one band
 * of every k; the bands are then processed in blocks. */
int64_t qmx_plan_workspace_bytes(const qmx_plan* plan, int dtype, int64_t n_bands);

/* Query plan geometry: [0] active x-planes, [1] active (x, y) columns. */
int qmx_plan_info(const qmx_plan* plan, int64_t* info);

/* Kernel sequence of qmx_apply_local. 0 (default): the fused plane pass
 * (one kernel for the y and x transforms and the potential, each x-y plane
 * in LDS; a kernel specialised for the plane size where one exists:
 * square planes of 24..64 points per side) wherever a plane fits, else
 * separate y, x and y passes; 1: always the separate passes; 2: same as 0;
 * 3: the plane pass, always the kernel for general sizes (for comparisons).
 * qmx_plan_local_path returns the path
 * taken for a precision (2 plane, 1 separate, -1 bad argument). Set it
 * before sizing the workspace. No reference counterpart (the reference
 * calls a library 3D FFT, _basis_ops.py:152-205). */
int qmx_plan_set_local_path(qmx_plan* plan, int path);
int qmx_plan_local_path(const qmx_plan* plan, int dtype);

/* Local-potential apply, optionally with the kinetic term:
 *   out(k, b, G) = [accumulate ? out : 0] + KE(k, G) C(k, b, G)
 *                  + (1/N) FFT[ V(r) IFFT[C](r) ](G)
 * i.e. _apply_potential (dft/electrons/_basis_ops.py:39-113, scatter,
 * ifftn(norm="forward"), *= V, fftn(norm="forward"), gather, constrain)
 * fused with _apply_ke (_basis_ops.py:29-37) when ke != NULL.
 * V: real boxes (float for QMX_C64, double for QMX_C128), N = box size
 * each; k uses box v_slot[k] (geom.v_slot; all k use box 0 when it was
 * NULL). ke: [nk * ke_ldk] real of the same precision, or NULL. out may
 * equal C. */
int qmx_apply_local(const qmx_plan* plan, int dtype, const void* C, int64_t ldb,
                    int64_t ldk, void* out, int64_t out_ldb, int64_t out_ldk,
                    const void* V, const void* ke, int64_t ke_ldk, int64_t n_bands,
                    int accumulate, void* workspace, int64_t workspace_bytes,
                    void* stream);

/* qmx_apply_local on the k-points [k0, k0 + nk) of the plan only: C, out
 * and ke start at k-point k0 (their k-strides as above), V is the same slot
 * array. A plan over the k-points of many cells (one potential slot per
 * cell) then serves both a batched apply over every cell and the apply for
 * one cell. Same semantics and reference as qmx_apply_local. */
int qmx_apply_local_k(const qmx_plan* plan, int64_t k0, int64_t nk, int dtype, const void* C,
                      int64_t ldb, int64_t ldk, void* out, int64_t out_ldb, int64_t out_ldk,
                      const void* V, const void* ke, int64_t ke_ldk, int64_t n_bands,
                      int accumulate, void* workspace, int64_t workspace_bytes,
                      void* stream);

/* Electron density of the bands C (k-points [k0, k0 + nk) of the plan):
 *   rho[x][y][z] (= or +=) sum_{k, b} prefac[k][b] |psi_kb(x, y, z)|^2,
 * psi = the unscaled inverse FFT of the sphere coefficients scattered into
 * the box (Basis.collect_density / _CollectDensityKernel with accum_norm_,
 * dft/electrons/_basis_ops.py:207-343 and math/_linalg.py:31-76, for one
 * spin channel and one spinor component; prefac = f w_sk / volume as in
 * _basis_ops.py:229). C, ldb, ldk as for qmx_apply_local_k; prefac: device
 * double, k stride ldp (bands with prefac 0 are skipped, like the
 * argwhere(f) walk of accum_norm_); rho: device double, the box in C order
 * (z fastest), accumulated in FP64. Boxes whose x-y planes fit the plane
 * pass only (every box up to 64 x 64 planes). Workspace:
 * qmx_density_workspace_bytes(plan, dtype, n_bands) bytes for one pass over
 * the bands (less runs them in blocks). */
int64_t qmx_density_workspace_bytes(const qmx_plan* plan, int dtype, int64_t n_bands);
int qmx_collect_density(const qmx_plan* plan, int64_t k0, int64_t nk, int dtype, const void* C,
                        int64_t ldb, int64_t ldk, int64_t n_bands, const double* prefac,
                        int64_t ldp, double* rho, int accumulate, void* workspace,
                        int64_t workspace_bytes, void* stream);

/* Spinor cross term of the density (Basis.collect_density with need_Mvec,
 * dft/electrons/_basis_ops.py:330-336: accum_prod_(f, psi_1, conj(psi_0)),
 * math/_linalg.py:77-107), summed over k and bands:
 *   rho[0][r] + i rho[N][r] (= or +=) sum_{k,b} prefac[k, b] psi1 conj(psi0)(r)
 * C0 and C1 hold the spinor components 0 and 1 of the same bands, with the
 * strides of qmx_collect_density; rho is 2 N doubles (real part box, then
 * imaginary part box), accumulated in FP64. The two x-y planes of a box must
 * fit LDS together (every box up to 64 x 64 planes in both precisions).
 * Workspace: qmx_density_cross_workspace_bytes(plan, dtype, n_bands). */
int64_t qmx_density_cross_workspace_bytes(const qmx_plan* plan, int dtype, int64_t n_bands);
int qmx_collect_density_cross(const qmx_plan* plan, int64_t k0, int64_t nk, int dtype,
                              const void* C0, const void* C1, int64_t ldb, int64_t ldk,
                              int64_t n_bands, const double* prefac, int64_t ldp, double* rho,
                              int accumulate, void* workspace, int64_t workspace_bytes,
                              void* stream);

/* Projectors of one species on the plane waves of a basis
 * (Ions.get_projectors, dft/ions/_ions_projectors.py:14-72):
 *   out[k * ldk + (atom * n_proj + p) * ldp + g] =
 *     R_{i_rf[p]}(|G+k|) (i^l r^l Y_lm)(G+k) exp(-2 pi i (iG+k) . tau_atom) scale
 * for g < n[k], 0 for the padding g in [n[k], n_tot). G + k = Gbasis (iG + k)
 * (Gbasis host double [3][3] row-major, Cartesian rows); iG device int32
 * [nk][n_tot][3]; k device double [nk][3] fractional; n device int32 [nk];
 * positions device double [n_atoms][3] fractional. Projector p has
 * angular momentum l[p] <= 6, -l <= m[p] <= l, radial function i_rf[p]
 * (host int32 arrays, n_proj <= 32), with real solid harmonics ordered as in
 * math/spherical_harmonics.py:53-137. R is the quintic blip spline of
 * quintic_spline.Interpolator (math/quintic_spline.py:56-128): coeff device
 * double [n_coeff][n_rf] (the (N + 4, n_rf) output of get_coeff) on node
 * spacing dG; every |G+k| <= G_max must lie on the table
 * (floor(G_max / dG) + 5 < n_coeff). scale: 1/sqrt(volume) (:48-49).
 * out: device complex128. */
int qmx_projectors(int64_t nk, int64_t n_tot, const int32_t* iG, const double* k,
                   const int32_t* n, const double* Gbasis, int64_t n_atoms,
                   const double* positions, int64_t n_proj, const int32_t* l, const int32_t* m,
                   const int32_t* i_rf, int64_t n_rf, int64_t n_coeff, const double* coeff,
                   double dG, double G_max, double scale, void* out, int64_t ldp, int64_t ldk,
                   void* stream);

/* The same projectors with Gaussian radial functions
 * R_i(G) = exp(-rc_i^2 G^2 / 2) (rc: host double [n_rf]) in place of the
 * splines: the analytic radials of the synthetic cells, built for each job
 * the batched engine admits. */
int qmx_gaussian_projectors(int64_t nk, int64_t n_tot, const int32_t* iG, const double* k,
                            const int32_t* n, const double* Gbasis, int64_t n_atoms,
                            const double* positions, int64_t n_proj, const int32_t* l,
                            const int32_t* m, const int32_t* i_rf, int64_t n_rf, const double* rc,
                            double scale, void* out, int64_t ldp, int64_t ldk, void* stream);

/* Gradient operator along Cartesian direction dir (Basis.apply_gradient,
 * dft/electrons/_basis_ops.py:15-26, with get_gradient = i (G + k),
 * _basis.py:240-254), scaled by alpha:
 *   out(k, b, g) = [accumulate ? out : 0] + alpha i G[k][g][dir] C(k, b, g)
 * G: Cartesian G + k of every basis entry, [nk][n_tot][3], real of the
 * working precision (float for QMX_C64, double for QMX_C128). C and out have
 * the strides of qmx_apply_local; out may equal C. With qmx_apply_local it
 * gives the meta-GGA term of H, -1/2 sum_dir grad(V_tau grad C)
 * (_hamiltonian.py:13-20). */
int qmx_apply_gradient(int dtype, const void* C, int64_t ldb, int64_t ldk, int64_t nk,
                       int64_t nb, int64_t n_tot, const void* G, int dir, double alpha, void* out,
                       int64_t out_ldb, int64_t out_ldk, int accumulate, void* stream);

/* Radial transforms of RadialFunction.transform (math/_radial_function.py:
 * 62-123) without the spline fit:
 *   out[i][iG] = sum_r w[r] f[i][r] r^(2 l_i) j_l(G r) / (G r)^l,  G = iG dG,
 * with f the real-space functions (r^l removed) on the radial grid r, w the
 * integration weights including 4 pi r^2 dr (the reference's Simpson rule,
 * :84-89), j_l(x) / x^l as in math/spherical_bessel.py:8-49. r, w, f, out:
 * device double ([n_r], [n_r], [n_f][n_r], [n_f][nG]); l: host int32 [n_f],
 * 0 <= l <= 6, n_f <= 64. */
int qmx_radial_transform(int64_t n_f, int64_t n_r, const double* r, const double* w,
                         const double* f, const int32_t* l, int64_t nG, double dG, double* out,
                         void* stream);

/* In-place batched dense c2c 3D FFT of `batch` contiguous boxes.
 * direction -1: forward, exp(-i G r); +1: inverse, exp(+i G r).
 * The result is multiplied by `scale` (1/N for QimPy's forward transform,
 * 1 for its inverse: grid/_grid.py:220-268, grid/_fft.py:228-337). */
int qmx_fft3d(int dtype, const int32_t* shape, void* data, int64_t batch,
              int direction, double scale, void* stream);

/* Per-band reductions norm2[k, b] = sum_g |C|^2 and (if ke != NULL)
 * kin[k, b] = sum_g |C|^2 KE(k, g), accumulated in double
 * (Wavefunction.band_norm / band_ke, _wavefunction.py:226-247).
 * ke is double [nk * ke_ldk]. Outputs are double, [nk * nb] row-major. */
int qmx_band_reduce(int dtype, const void* C, int64_t ldb, int64_t ldk, int64_t nk,
                    int64_t nb, int64_t n_tot, const double* ke, int64_t ke_ldk,
                    double* norm2, double* kin, void* stream);

/* Davidson residual + inverse-kinetic preconditioner (_davidson.py:93-99,
 * :182-184):  out = (HC - eig C) / x,  x = KE/KEref,  x <- x + exp(-x);
 * also writes norm2[k, b] = sum_g |out|^2 and, if rnorm2 != NULL, the
 * unpreconditioned residual norm rnorm2[k, b] = sum_g |HC - eig C|^2, both
 * accumulated in double. eig, KEref: double [nk * nb]; ke double
 * [nk * ke_ldk]. KEref may be NULL: the kernel then forms each band's
 * KEref = sum_g |C|^2 KE itself (Wavefunction.band_ke, _davidson.py:95).
 * With complex64 data the factor 1 / (x + exp(-x)) is evaluated in FP32
 * (its argument formed in FP64). */
int qmx_precond_residual(int dtype, const void* C, const void* HC, int64_t ldb,
                         int64_t ldk, const double* eig, const double* KEref,
                         const double* ke, int64_t ke_ldk, void* out, int64_t out_ldb,
                         int64_t out_ldk, int64_t nk, int64_t nb, int64_t n_tot,
                         double* norm2, double* rnorm2, void* stream);

/* The same residual, then each row scaled to unit norm (zero where its norm
 * is below 1e-30) and zeroed past n[k] (int32 [nk] on the device), in one
 * launch: the preconditioned, normalised and constrained expansion of a
 * Davidson step (_davidson.py:182-187 with Wavefunction.constrain,
 * _wavefunction.py:264-275). norm2 and rnorm2 are those before the
 * scaling. With complex64 data the preconditioner factor
 * 1 / (x + exp(-x)) is evaluated in FP32 (the residual and its norms stay
 * FP64). */
int qmx_precond_residual_normalize(int dtype, const void* C, const void* HC, int64_t ldb,
                                   int64_t ldk, const double* eig, const double* KEref,
                                   const double* ke, int64_t ke_ldk, void* out,
                                   int64_t out_ldb, int64_t out_ldk, int64_t nk, int64_t nb,
                                   int64_t n_tot, double* norm2, double* rnorm2,
                                   const int32_t* n, void* stream);

/* Linear combination of three wavefunction blocks of the same layout:
 *   out(k, b, g) = a[k, b] X(k, b, g) + b[k, b] Y(k, b, g) + c[k, b] Z(k, b, g)
 * with the coefficients device doubles at [k * cf_ldk + b * cf_ldb]
 * (cf_ldb = 0 for one coefficient per k-point); b or c NULL drops its term
 * (Y, Z may then be NULL); out may alias X, Y or Z. All blocks have the
 * strides ldb, ldk. The axpy-type wavefunction arithmetic
 * (_wavefunction_arithmetic.py:9-109) fused per step: one Chebyshev
 * filter step Y_{n+1} = (2 s'/e)(H Y_n - c Y_n) - s s' Y_{n-1} of CheFSI
 * (_chefsi.py:122-140) is a single call. */
int qmx_axpby(int dtype, int64_t nk, int64_t nb, int64_t n_tot, const double* a, const double* b,
              const double* c, int64_t cf_ldk, int64_t cf_ldb, const void* X, const void* Y,
              const void* Z, void* out, int64_t ldb, int64_t ldk, void* stream);

/* Scale each band by scale[k, b] (double, or 1 if NULL) and zero the
 * padding g >= n[k] (n: device int32 [nk]) — Wavefunction.constrain
 * (_wavefunction.py:264-275) fused with the band normalisation of
 * _davidson.py:186-187. */
int qmx_scale_constrain(int dtype, void* C, int64_t ldb, int64_t ldk, int64_t nk,
                        int64_t nb, int64_t n_tot, const double* scale,
                        const int32_t* n, void* stream);

/* Row blocks between device buffers in one launch (a batched strided copy
 * with zero padding). No counterpart in the reference: it moves the stacked
 * state of the batched engine when a bucket's jobs change (kept jobs'
 * wavefunctions, the rows of jobs that moved, their projectors and
 * potentials). blocks: device array of n qmx_copy_block; block i writes
 *   dst[i0 * dst_s0 + i1 * dst_s1 + c], i0 < n0, i1 < n1, c < dst_cols,
 * from src at the same (i0, i1) (src strides) for c < src_cols and zero
 * beyond it, or zero throughout when src is NULL. mode 0 copies 8-byte
 * words (counts and strides in words); mode 1 converts doubles to floats
 * (counts and strides in reals of each side). Workgroups are dealt out in
 * block order: block i takes ceil(n0 n1 / rows_per_wg) of them from
 * first_wg on, and n_wg is their total. Blocks must not overlap each other
 * or their sources. */
typedef struct qmx_copy_block {
  const void* src;
  void* dst;
  int64_t n0, n1, src_cols, dst_cols, src_s0, src_s1, dst_s0, dst_s1;
  int64_t rows_per_wg, first_wg, mode, reserved;
} qmx_copy_block;
int qmx_copy_blocks(const void* blocks, int64_t n, int64_t n_wg, void* stream);

/* Bandwidth-limited xorshift randomisation of bands [b_start, b_stop)
 * (Wavefunction.randomize, _wavefunction_init.py:27-125): state
 * 1 + seed + g + n_max (k_offset + k), b_start + 1 discarded draws, then
 * one complex normal per band, zero padding, times 1/(1+((4/3)KE)^6).
 * n: device int32 [nk]; ke: device double [nk * ke_ldk]. */
int qmx_randomize(int dtype, void* C, int64_t ldb, int64_t ldk, int64_t nk,
                  int64_t n_tot, int64_t n_max, int64_t k_offset, const int32_t* n,
                  const double* ke, int64_t ke_ldk, int64_t seed, int64_t b_start,
                  int64_t b_stop, void* stream);

/* Batched Hermitian eigensolver for small complex128 matrices (n <= 96) by
 * cyclic Jacobi, one workgroup per matrix: the subspace eigensolves
 * eigh(U^H H U) of eighg (math/_linalg.py:181) and Davidson's initial eigh
 * (_davidson.py:165) for all k-points / cells of a step in one launch.
 * Input: batch matrices, row-major with leading dimension lda, batch stride
 * stride_a, lower triangle read (UPLO = L). Output: E[batch * n] ascending,
 * V row-major (ldv, stride_v) with eigenvector j in column j. Workspace of
 * qmx_heevj_workspace_bytes(n, batch) bytes is needed for n > 70.
 * max_sweeps <= 0 selects 30. */
int64_t qmx_heevj_workspace_bytes(int64_t n, int64_t batch);
int qmx_heevj_batched(const void* A, int64_t n, int64_t lda, int64_t stride_a, int64_t batch,
                      double* E, void* V, int64_t ldv, int64_t stride_v, void* workspace,
                      int64_t workspace_bytes, int max_sweeps, void* stream);

/* Batched Hermitian eigensolver for n <= 256, one workgroup per matrix:
 * Householder tridiagonalisation, implicit-shift QL with eigenvectors, and the
 * back-transformation V = Q Z, all inside one launch (the same role as
 * qmx_heevj_batched, for larger n). Same input and output conventions;
 * info[b] != 0 marks a matrix whose QL iteration did not converge.
 * Workspace: qmx_heev_small_workspace_bytes(n, batch) bytes (the QL
 * eigenvector accumulator moves from LDS to the workspace above n = 128). */
int64_t qmx_heev_small_workspace_bytes(int64_t n, int64_t batch);
int qmx_heev_small_batched(const void* A, int64_t n, int64_t lda, int64_t stride_a,
                           int64_t batch, double* E, void* V, int64_t ldv, int64_t stride_v,
                           int* info, void* workspace, int64_t workspace_bytes, void* stream);

/* Many Hermitian matrices of different sizes (n <= 256) in one launch of
 * the qmx_heev_small_batched algorithm, one workgroup per descriptor; the
 * launch costs about the slowest matrix instead of the sum. `desc` is a
 * DEVICE array of `count` descriptors, every n <= n_max. Each matrix needs
 * its own work area of qmx_heev_small_work_bytes(n) bytes (256-byte
 * aligned). A is read (lower triangle, row stride lda), E gets the n
 * ascending eigenvalues, V the eigenvectors in columns (row stride ldv),
 * *info (if not NULL) is nonzero when QL did not converge. Used by the
 * batched cell solver (no counterpart in the reference, which solves one
 * cell at a time). */
typedef struct qmx_heev_desc {
  const void* A;
  double* E;
  void* V;
  int32_t* info;
  void* work;
  int64_t lda;
  int64_t ldv;
  int32_t n;
  int32_t n_want; /* eigenvectors wanted (qmx_heevx_grouped; 0 = all n) */
} qmx_heev_desc;
int64_t qmx_heev_small_work_bytes(int64_t n);
int qmx_heev_small_grouped(const qmx_heev_desc* desc, int64_t count, int64_t n_max, void* stream);

/* The same grouped eigenproblems (n <= 256, descriptors as above) with a
 * parallel tridiagonal stage: Householder reduction, all n eigenvalues by
 * bisection (one lane each, division-free Sturm counts), the eigenvectors
 * of the n_want lowest by inverse iteration (one lane each, re-orthogonalised
 * inside clusters of close eigenvalues) and the back-transformation. Replaces
 * the serial QL iteration of qmx_heev_small_grouped for eigh inside eighg
 * (math/_linalg.py:181) and the Davidson subspace solves
 * (_davidson.py:165,199-206). E gets all n eigenvalues ascending, V the
 * eigenvectors of the n_want lowest in columns 0..n_want-1 (the other columns
 * are not written). Work area per matrix: qmx_heevx_work_bytes(n, n_want)
 * bytes (256-byte aligned). *info is set to 0, or 1 for a non-finite input
 * (E and V are then NaN). */
int64_t qmx_heevx_work_bytes(int64_t n, int64_t n_want);
int qmx_heevx_grouped(const qmx_heev_desc* desc, int64_t count, int64_t n_max, void* stream);

/* One larger Hermitian complex128 matrix (2 <= n <= 2048): the subspace
 * eigensolve eigh(U^H H U) inside eighg (math/_linalg.py:144-183) for the
 * expanded Davidson subspace (_davidson.py:199-206) and the CheFSI
 * Rayleigh-Ritz (_chefsi.py:156) above n = 256. A (row stride lda,
 * lower triangle read, UPLO = L) is reduced to real tridiagonal form by one
 * cooperative launch whose workgroups own rows of A and exchange one column
 * per step; then all n eigenvalues by multisection, the eigenvectors of the
 * n_want lowest (0 = all) by inverse iteration with re-orthogonalisation in
 * clusters, and the back-transformation. E[n] ascending; eigenvector j goes
 * to V[i * vsi + j * vsj]; *info (device, may be NULL) is 0, or 1 for a
 * non-finite input (E and V NaN). Work area: qmx_heevx_large_work_bytes(n,
 * n_want) bytes, 256-byte aligned. n_groups > 0 sets the reduction's
 * workgroups (0 = chosen from n). */
int64_t qmx_heevx_large_work_bytes(int64_t n, int64_t n_want);
int qmx_heevx_large(const void* A, int64_t n, int64_t lda, int64_t n_want, double* E, void* V,
                    int64_t vsi, int64_t vsj, int* info, void* work, int64_t work_bytes,
                    int n_groups, void* stream);

/* Inverse Cholesky factor of one Hermitian positive definite complex128
 * matrix (1 <= n <= 2048): X = L^-1 with O = L L^H, the orthonormalising
 * U = X^H of ortho_matrix (math/_linalg.py:110-141: torch.linalg.cholesky
 * then solve_triangular against the identity) used by eighg and Davidson's
 * orthonormalisation. O (row stride ldo, lower triangle read); X row-major
 * n x n, lower triangular (zeros above). *info (device, may be NULL) is 0,
 * or k + 1 for the first non-positive or non-finite pivot k, as LAPACK potrf
 * and torch.linalg.cholesky_ex. One cooperative launch; work area of
 * qmx_potrf_inv_work_bytes(n) bytes, 256-byte aligned. */
int64_t qmx_potrf_inv_work_bytes(int64_t n);
int qmx_potrf_inv(const void* O, int64_t n, int64_t ldo, void* X, int* info, void* work,
                  int64_t work_bytes, int n_groups, void* stream);

/* Tuned launch choices (no counterpart in the reference, which has no
 * device kernels to tune): a process-wide table read by the launch code
 * before its own cost model, filled by qimpy_amd.autotune from timings on
 * the device (or its cache file). Keys and choices (int64):
 * - QMX_TUNE_DOT, qmx_dot: key (dtype, qmx_tune_batch_bucket(batch), M, N,
 *   qmx_tune_bucket(G)),
 *   choice (kernel 0 = fitted blocks / 1 = 64 x 64 tiles, bm16, bn16 block
 *   size in 16-row tiles (kernel 0), chunks of the basis axis);
 * - QMX_TUNE_ROTATE, qmx_gemm: key (dtype, m_dtype, variant, m, K,
 *   qmx_tune_bucket(G)), choice (16-row output tiles per block, MT);
 * - QMX_TUNE_PLANE, the fused plane pass of qmx_apply_local_k: key (dtype,
 *   nx, ny, nz), choice (0 the specialised kernel, 1 the general kernel on
 *   256 threads, 2 on 512).
 * A choice that does not fit the shape is ignored. Setting entries while
 * launches run on other threads is safe; the launches then see either. */
enum { QMX_TUNE_DOT = 1, QMX_TUNE_ROTATE = 2, QMX_TUNE_PLANE = 3 };
int64_t qmx_tune_bucket(int64_t G);
int64_t qmx_tune_batch_bucket(int64_t batch);
int qmx_tune_set(int op, const int64_t* key, int n_key, const int64_t* choice, int n_choice);
int qmx_tune_clear(void);
int64_t qmx_tune_count(void);

/* Fused steps of the batched Davidson iteration (no counterpart as kernels
 * in the reference, which runs them as tensor expressions):
 * - qmx_normalize_constrain: X[k][b] /= norm[k][b] (0 where norm <= 1e-30),
 *   zero padding g >= n[k] (_davidson.py:186-187 + constrain,
 *   _wavefunction.py:264-275); norm: double [nk][nb] on the device;
 * - qmx_canonical_basis: Ut[b][j][i] = W[b][i][j] / sqrt(lam[b][j]) where
 *   lam > tol, else 0 (W: complex128 eigenvectors in columns, m x m per
 *   batch; Ut in dtype): the canonical orthonormalisation of the residual
 *   block, transposed for the rotation Ut @ R;
 * - qmx_subspace_assemble: E = O - I and H of the expanded subspace
 *   [C, R] (2 nb x 2 nb per batch, complex128) from XO = [C^H R; R^H R] and
 *   XH = [C^H HR; R^H HR] (2 nb x nb each), eig (current Ritz values) and lam
 *   (residual-overlap eigenvalues: directions with lam <= tol are dropped,
 *   their O diagonal set to 1 and H diagonal shifted by `shift`); the
 *   TileExpansion layout of _davidson.py:196-206,250-264 with C^H C = I and
 *   C^H H C = diag(eig). */
int qmx_normalize_constrain(int dtype, void* C, int64_t ldb, int64_t ldk, int64_t nk, int64_t nb,
                            int64_t n_tot, const double* norm, const int32_t* n, void* stream);
int qmx_canonical_basis(int dtype, const void* W, const double* lam, int64_t batch, int64_t m,
                        double tol, void* Ut, void* stream);
int qmx_subspace_assemble(const void* XO, const void* XH, const double* eig, const double* lam,
                          int64_t batch, int64_t nb, double tol, double shift, void* E, void* H,
                          void* stream);

/* Batched overlap product over the basis axis on the matrix cores:
 *   out[b][m][n] = (accumulate ? out : 0) + sum_g conj(A[b][m][g]) B[b][n][g]
 * (_dot, dft/electrons/_wavefunction_dot.py:66-84: C1^H C2 per spin and k).
 * A: batch x M x G (row stride lda, batch stride sa, elements), B likewise
 * (N rows); out: complex128, row stride ldc, batch stride sc. QMX_C64
 * inputs run on FP32 MFMA, accumulated in FP32 within chunks of the basis
 * axis and in FP64 across them; QMX_C128 on FP64 MFMA. Workspace:
 * qmx_dot_workspace_bytes(dtype, batch, M, N, G) bytes. */
int64_t qmx_dot_workspace_bytes(int dtype, int64_t batch, int64_t M, int64_t N, int64_t G);
int qmx_dot(int dtype, const void* A, int64_t lda, int64_t sa, const void* B, int64_t ldb,
            int64_t sb, int64_t batch, int64_t M, int64_t N, int64_t G, void* out, int64_t ldc,
            int64_t sc, int accumulate, void* workspace, int64_t workspace_bytes, void* stream);

/* qmx_dot without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc); no counterpart in the reference. The
 * chunk partials live in a fixed device array of qmx_dot_noalloc_max_bytes()
 * (32 MiB) bytes that each GPU holds for the process lifetime; a product
 * whose partials need more is refused with QMX_ERR_NOMEM (qmx_dot_workspace_
 * bytes gives the need). It calls no hipMalloc, hipFree or allocating
 * library routine and records no profiling events. Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (concurrent
 * streams, threads) race; calls in one stream are ordered. Same arguments,
 * kernels and results as qmx_dot otherwise. */
int64_t qmx_dot_noalloc_max_bytes(void);
int qmx_dot_noalloc(int dtype, const void* A, int64_t lda, int64_t sa, const void* B, int64_t ldb,
                    int64_t sb, int64_t batch, int64_t M, int64_t N, int64_t G, void* out,
                    int64_t ldc, int64_t sc, int accumulate, void* stream);

/* Band-space rotation and back-projection on the matrix cores:
 *   out[b][i][g] = alpha sum_k op(M)[b][i][k] W[b][k][g] + beta out[b][i][g]
 * op(M) = M (m x K, row stride ldm) or, with trans_m, M^T (M stored K x m);
 * W: K rows of G plane waves (row stride ldw, unit stride in g); out: m rows
 * (row stride ldo); batch strides sm, sw, so in elements. dtype is the type
 * of W and out (QMX_C64, QMX_C128), m_dtype that of M (either; it is
 * rounded to the arithmetic type). variant 0: FP32 MFMA for complex64,
 * FP64 MFMA for complex128; variant 1 (complex64 only): BF16 x 3 split
 * precision (each FP32 operand as a BF16 hi + lo pair, three BF16 products
 * per real product, FP32 accumulation), relative error ~2^-16 per product.
 * out must not overlap M; it may be W itself (out == W with ldo == ldw and
 * so == sw: an in-place rotation), which takes one block of output rows, so
 * m <= 128 (complex64, variant 0), 96 (variant 1) or 80 (complex128); it must
 * not overlap W otherwise.
 * Replaces _matmul (the C @ M of
 * Wavefunction, dft/electrons/_wavefunction_dot.py:138-191), the nonlocal
 * back-projection beta (D beta^H C) (_hamiltonian.py:21-24) and the
 * Davidson rotations (_davidson.py:212-215). */
int qmx_gemm(int dtype, int m_dtype, int variant, int trans_m, int64_t batch, int64_t m,
             int64_t K, int64_t G, double alpha, const void* M, int64_t ldm, int64_t sm,
             const void* W, int64_t ldw, int64_t sw, double beta, void* out, int64_t ldo,
             int64_t so, void* stream);

/* The same product in BLAS shape (row-major, batched):
 *   C[b] (m x n) = alpha op(A[b]) (m x k) B[b] (k x n) + beta C[b],
 * op(A) = A^T when transA; B untransposed (transB must be 0: its rows of n
 * plane waves are the long axis). A, B and C share the entry point's type.
 * The qmx_gemm_{c64, c64_bf16x3, c128} family of the band algebra
 * (_wavefunction_dot.py:138-191, _hamiltonian.py:21-24). */
int qmx_gemm_c64(int transA, int transB, int64_t m, int64_t n, int64_t k, double alpha,
                 const void* A, int64_t lda, int64_t sA, const void* B, int64_t ldb, int64_t sB,
                 double beta, void* C, int64_t ldc, int64_t sC, int64_t batch, void* stream);
int qmx_gemm_c64_bf16x3(int transA, int transB, int64_t m, int64_t n, int64_t k, double alpha,
                        const void* A, int64_t lda, int64_t sA, const void* B, int64_t ldb,
                        int64_t sB, double beta, void* C, int64_t ldc, int64_t sC, int64_t batch,
                        void* stream);
int qmx_gemm_c128(int transA, int transB, int64_t m, int64_t n, int64_t k, double alpha,
                  const void* A, int64_t lda, int64_t sA, const void* B, int64_t ldb, int64_t sB,
                  double beta, void* C, int64_t ldc, int64_t sC, int64_t batch, void* stream);

/* Kernel ids reported by the profiler. */
enum {
  QMX_KERNEL_PASS_A = 0, /* sphere -> columns, inverse z */
  QMX_KERNEL_PASS_B = 1, /* columns -> planes, inverse y */
  QMX_KERNEL_PASS_C = 2, /* x pass: inverse x, * V, forward x */
  QMX_KERNEL_PASS_D = 3, /* planes -> columns, forward y */
  QMX_KERNEL_PASS_E = 4, /* columns -> sphere, forward z, KE, constrain */
  QMX_KERNEL_FFT3D = 5,
  QMX_KERNEL_GEMM = 6,
  QMX_KERNEL_PLANE = 7,   /* fused plane pass: inverse y, x, * V, forward x, y */
  QMX_KERNEL_VZMAJOR = 8, /* potential boxes to z-major order */
  QMX_KERNEL_APPLY = 9,   /* whole local apply, dense-transform byte convention */
  QMX_KERNEL_PROJ = 10,   /* qmx_projectors (complex128 output: reported as 10 + C128) */
  QMX_KERNEL_ROTATE = 11,         /* qmx_gemm, FP32 / FP64 MFMA */
  QMX_KERNEL_ROTATE_BF16X3 = 12,  /* qmx_gemm, BF16 x 3 on MFMA */
  QMX_KERNEL_C128 = 16,   /* added to the id of a complex128 launch */
  QMX_N_KERNEL_IDS = 32
};

/* Event-based profiling of this library's launches (no counterpart in the
 * reference, whose StopWatch (src/qimpy/profiler.py:15-117) times whole
 * Python calls). When enabled, each launch is bracketed by HIP events on its
 * own stream; qmx_profile_read() waits for them and returns, per kernel id,
 * the summed milliseconds, the summed compulsory HBM bytes and the launch
 * count, then clears the records. */
int qmx_profile_enable(int on);
int qmx_profile_read(double* ms, double* bytes, int64_t* count, int n_ids);
/* The same with the summed algorithmic flops per kernel id (8 real flops
 * per complex multiply-add of qmx_dot and qmx_gemm; 0 for the other
 * kernels); flops may be NULL. */
int qmx_profile_read_flops(double* ms, double* bytes, double* flops, int64_t* count, int n_ids);

const char* qmx_last_error(void);
const char* qmx_version(void);

#ifdef __cplusplus
}
#endif
#endif /* QMX_H */
