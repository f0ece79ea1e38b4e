/*
 * qumasun_hip.h -- C ABI of the MI355X-native QUMASUN LOBPCG H.psi hot path.
 *
 * Plain C types only.  Every array argument named d_* is a DEVICE pointer
 * (hipMalloc'd, or a torch CUDA tensor's data pointer); the others are host
 * pointers.  `stream` is a hipStream_t passed as void* (NULL = default stream).
 * All entry points return a status code (QM_OK = 0) instead of terminating the
 * process as the reference's gyCheckError does (gyield/gy_hip.h:95-100);
 * qm_last_error() gives the message of the last failure on the calling thread.
 *
 * Band-block layout (the reference WaveSet / SoAComplex, qumasun_base1.h:451-462):
 *   grid index i = ix + nx*(iy + ny*iz)   (x fastest, initialize.h:340)
 *   ncomp = 2 (SoA complex): band n real part at d_psi + 2n*Ng, imaginary at (2n+1)*Ng
 *   ncomp = 1 (real Gamma-point bands): band n at d_psi + n*Ng
 * Viewed as a real Ng x (ncomp*nb) column-major matrix for the Gram products,
 * exactly as the reference hands it to blas_DGEMM_t (lobpcg_z_multi_mpi.h:300-558).
 */
#ifndef QUMASUN_HIP_H
#define QUMASUN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QM_OK 0
#define QM_ERR_ARG (-1)      /* invalid argument or size */
#define QM_ERR_HIP (-2)      /* HIP / rocBLAS / rocSOLVER runtime error */
#define QM_ERR_STATE (-3)    /* call made before the state it needs was set */
#define QM_ERR_NUMERIC (-4)  /* numerical breakdown (e.g. overlap not positive definite) */
#define QM_ERR_NOMEM (-5)    /* device allocation failed */

typedef struct qm_ctx qm_ctx;
typedef struct qm_lobpcg qm_lobpcg;

/* Message of the last failing call on this thread ("" if none). */
const char *qm_last_error(void);

/* ---------------------------------------------------------------- context */

/* Grid and operator state on device `device`.  Replaces the QUMASUN_BASE1
 * members read by mHamiltonianMatrix_ddm_bundle (qumasun_base1.h:320,
 * qumasun_base1.matrix.h:141-254): grid sizes and spacing
 * (qumasun_base1.constructor.h:25-28).  One whole periodic grid, no DDM split. */
int qm_create(qm_ctx **out, int nx, int ny, int nz, double hx, double hy, double hz, int device);
void qm_destroy(qm_ctx *ctx);

/* Local potential V_tot(r), Ng host doubles.  Replaces ml_Vtot (qumasun_base1.h:482)
 * consumed by mPotentialMatrixAdd_ddm_bundle (qumasun_base1.matrix.h:329-337). */
int qm_set_vloc(qm_ctx *ctx, const double *vloc);

/* KB projector set (host arrays), replacing the tables that m_pp_SvF holds for
 * ProjectionPP_bundle (PseudoPotNonlocal_SvF2_kpoint.h:4239): per atom a,
 *   nproj[a] projectors (<= 32: two radial projectors for each l <= 3, as the
 *   reference's W.upf needs, num_proj_RY of ..._kpoint.h:940-945), voxels vox_off[a]..vox_off[a+1]-1 with UNWRAPPED
 *   integer grid coordinates vox_xyz[3v..3v+2] (the grid index is their periodic
 *   wrap; the Bloch phase uses the unwrapped value, ..._kpoint.h:3578-3596),
 *   table[tab_off[a] + w*V_a + v] (column-major V_a x M_a, ..._kpoint.h:2255-2272),
 *   energy[proj_off[a] + w] (projector_energy_RY, ..._kpoint.h:1069).
 * dtable (may be NULL) holds d p/d r_x, d p/d r_y, d p/d r_z in the same layout,
 * one block of tab_off[natoms] values per direction (FORCE_DIFF_PROJ tables,
 * ..._kpoint.h:2255-2267); it is required by qm_nonlocal_forces_*.
 * Overlapping spheres are allowed.  Every array is validated and the new set is
 * built before the old one is released: on any error the previous set (and its
 * k-point phases) stays in place. */
int qm_set_projectors(qm_ctx *ctx, int natoms, const int *nproj, const int64_t *vox_off,
                      const int *vox_xyz, const double *table, const double *dtable,
                      const int64_t *tab_off, const double *energy, const int *proj_off);

/* k-points in 1/bohr (kvec[3*ik..3*ik+2]); precomputes e^{ik.r} on every projector
 * voxel (mCreateBlochExp, ..._kpoint.h:3537-3618).  k-point 0 defaults to Gamma.
 * On error the previous k-points stay in place. */
int qm_set_kpoints(qm_ctx *ctx, int nk, const double *kvec);

/* ---------------------------------------------------------------- kinetic term */

#define QM_KINETIC_FD8 0  /* 8th-order finite differences (the default) */
#define QM_KINETIC_FFT 1  /* spectral, through 3-D FFTs */

/* Kinetic term of every later apply on the context (qm_hpsi_apply_*,
 * qm_local_apply_*, the LOBPCG solver, the Taylor propagator):
 * QM_KINETIC_FD8   -1/2 (Delta_FD8 + 2 i k.grad_FD8 - k^2), mKineticMatrixAdd_ddm_kpint
 *                  (qumasun_base1.matrix.h:298-317, GridDifferenceKpoint8th.h:17-151);
 * QM_KINETIC_FFT   IFFT(|G + k|^2 / 2 FFT psi) / N, the reference's default
 *                  (m_algorithm_Laplacian = 1, qumasun_input.h:117) in
 *                  mHamiltonianMatrix_ddm_bundle (qumasun_base1.matrix.h:171-228)
 *                  with OperateKineticInKspace_bundle (96-130): G = 2 pi / L_d times the
 *                  wrapped index i (i > n/2 -> i - n; the Nyquist index n/2 stays +n/2).
 * Spectral applies hold the band block's spectrum and rocFFT's work buffer on
 * the context (grown on demand); qm_hpsi_apply_f32_noalloc refuses them.  The
 * stencil autotuner only concerns QM_KINETIC_FD8. */
int qm_set_kinetic(qm_ctx *ctx, int mode);

/* ---------------------------------------------------------------- H.psi */

/* Stream capture.  An apply (qm_hpsi_apply_*, qm_local_apply_*, qm_nonlocal_add_*,
 * qm_hpsi_apply_f32_noalloc) may be recorded into a caller's HIP graph.  The graph
 * then holds the addresses of this context's projector tables, phases, V_loc
 * and (except for the noalloc entry point) its nonlocal work buffers.  From the
 * capture on, every call that would free or move them returns QM_ERR_STATE:
 * qm_set_projectors, qm_set_kpoints, and an apply that needs larger work
 * buffers (more bands, or FP64 after FP32).  qm_set_vloc overwrites V_loc in
 * place, so replays see the new potential.  Capture an apply after an eager
 * apply of the same shape (the work buffers and the stencil shape are then
 * settled).  No counterpart in the reference, which launches every kernel from
 * the host.
 * qm_release_captures: the caller has destroyed every graph holding an apply of
 * this context (or will not replay them again); the calls above work again. */
int qm_release_captures(qm_ctx *ctx);

/* d_hpsi = H(k_ik) d_psi on nb bands, overwriting d_hpsi:
 *   H = -1/2 (Delta_FD8 + 2 i k.grad_FD8 - k^2) + V_loc + sum_a |p> E dV <p|
 * This is the LOBPCG operator slot OpeA(Ax, x, num) (lobpcg_z_multi_mpi.h:1732-1742,
 * called at 1863, 2064, 2084) bound to mHamiltonianMatrix_ddm_bundle
 * (qumasun_kpoint.lobpcg.h:97-99) with the FD8 kinetic term of
 * mKineticMatrixAdd_ddm_kpint (qumasun_base1.matrix.h:298-317).
 * ncomp = 1 requires a Gamma k-point.  d_hpsi must not alias d_psi. */
int qm_hpsi_apply_f64(qm_ctx *ctx, double *d_hpsi, const double *d_psi, int nb, int ncomp, int ik,
                      void *stream);
/* FP32 storage and arithmetic; the projector inner products are reduced in FP64
 * across thread blocks. */
int qm_hpsi_apply_f32(qm_ctx *ctx, float *d_hpsi, const float *d_psi, int nb, int ncomp, int ik,
                      void *stream);

/* qm_hpsi_apply_f32 for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  It allocates nothing, creates no stream or
 * event, does not synchronise and does no shape timing: the nonlocal
 * coefficients and partial sums live in a fixed 64 MiB device array that the
 * library holds per GPU until the process ends, and the stencil takes the
 * shape already tuned for this call shape, else candidate 0 (same values), and
 * the nonlocal projection its tuned shape, else the size rule.
 * The work memory needed is 4*ptot*ncomp*nb (rounded up to 256) +
 * 8*nchunk*P*ncomp*nb bytes (ptot projectors, nchunk = sum over atoms of
 * ceil(V_a/512), P = 16, or 32 when an atom has more than 16 projectors);
 * above 64 MiB the call returns QM_ERR_ARG: apply fewer bands per call.  Every call on one GPU, from any context or stream, shares that
 * array, so calls that overlap in time on one GPU race; the array keeps the
 * last call's contents, and nothing in it needs clearing between calls.
 * Same contract and results as qm_hpsi_apply_f32 otherwise. */
int qm_hpsi_apply_f32_noalloc(qm_ctx *ctx, float *d_hpsi, const float *d_psi, int nb, int ncomp, int ik,
                              void *stream);

/* The two halves of qm_hpsi_apply_*, for callers that assemble H.psi
 * themselves (the reference's deprecated per-band path does, qumasun_base1.matrix.h:262-291):
 * qm_local_apply_*   d_out = -1/2 (Delta + 2 i k.grad - k^2) psi + V_loc psi (overwrites;
 *                    LaplacianKpoint8th_ddm + mPotentialMatrix_ddm, matrix.h:298-337);
 * qm_nonlocal_add_*  d_out += sum_a |p> E dV <p| psi (ProjectionPP_bundle,
 *                    PseudoPotNonlocal_SvF2_kpoint.h:4239-4247). */
int qm_local_apply_f64(qm_ctx *ctx, double *d_out, const double *d_psi, int nb, int ncomp, int ik, void *stream);
int qm_local_apply_f32(qm_ctx *ctx, float *d_out, const float *d_psi, int nb, int ncomp, int ik, void *stream);
int qm_nonlocal_add_f64(qm_ctx *ctx, double *d_out, const double *d_psi, int nb, int ncomp, int ik, void *stream);
int qm_nonlocal_add_f32(qm_ctx *ctx, float *d_out, const float *d_psi, int nb, int ncomp, int ik, void *stream);

/* ------------------------------------------------------ stencil autotuner */

/* Tile-shape autotuner of the fused stencil (no counterpart in the reference,
 * whose CPU loops have one shape).  The first qm_hpsi_apply_* / qm_local_apply_*
 * of a (precision, ncomp, k type, nb) shape on a context times every candidate
 * on the caller's buffers and keeps the fastest; it synchronises the stream
 * once.  Band counts in one power-of-two bucket (33..64, 65..128, ...) share
 * one choice.  Every candidate computes each point with the same arithmetic, so the
 * choice never changes a result.  Under stream capture an untuned shape uses
 * candidate 0.
 * qm_stencil_num_candidates  number of candidates;
 * qm_stencil_candidate_name  its description (NUL-terminated, truncated to buflen);
 * qm_stencil_force           use `candidate` for every later apply, -1 = autotune;
 * qm_stencil_choice          candidate in use for that shape, -1 if not tuned yet
 *                            (kt bit 0: a k-point with |k_d| >= 1e-8 in some d;
 *                            kt bit 1: the shape of full qm_hpsi_apply_* calls).
 * A full qm_hpsi_apply_* runs the stencil beside the nonlocal projection; its
 * shape is tuned separately, on full applies, the first time that shape of
 * call is seen (and synchronises the stream then). */
int qm_stencil_num_candidates(void);
int qm_stencil_candidate_name(int candidate, char *buf, int buflen);
int qm_stencil_force(qm_ctx *ctx, int candidate);
int qm_stencil_choice(qm_ctx *ctx, int precision, int ncomp, int kt, int nb, int *candidate);

/* ---------------------------------------------------------------- real-time propagation */

#define QM_TD_OP_FULL 3      /* H(k) = kinetic + V_loc + nonlocal */
#define QM_TD_OP_LOCAL 1     /* kinetic + V_loc */
#define QM_TD_OP_NONLOCAL 2  /* the nonlocal term alone */

/* Taylor propagator of nb SoA complex bands (ncomp = 2) in place:
 *   d_psi <- sum_{n=0..order} (-i dt Op)^n / n! d_psi,  Op as selected by `op`.
 * order 4 with QM_TD_OP_FULL is QUMASUN_TD::mTimeEvolutionH4 over
 * QUMASUN::TimeEvoH4_bundle (qumasun_td.time_evo_H4.h:70-204); order 4 with
 * QM_TD_OP_NONLOCAL is TimeEvoA4_bundle as TestEvolveNL4 calls it
 * (qumasun_symplectic.time_evo_K4.h:13-96, 305-335).  order <= 32.  Two work
 * blocks of the band block's size are held by the context (grown on demand).
 * Asynchronous on `stream`. */
int qm_td_taylor_f64(qm_ctx *ctx, double *d_psi, int nb, int ik, double dt, int order, int op, void *stream);
int qm_td_taylor_f32(qm_ctx *ctx, float *d_psi, int nb, int ik, double dt, int order, int op, void *stream);

/* Kinetic step of the split-operator propagator on nb SoA complex bands in place:
 *   d_psi <- IFFT(e^{-i dt |G + k|^2 / 2} FFT d_psi) / N
 * (EvolveKineticInKspace / mEvolveK_FFT, qumasun_symplectic.kinetic_fft.h:14-60, 65-140,
 * which take k = 0; k is the context's k-point ik here).  Asynchronous on `stream`. */
int qm_td_kinetic_fft_f64(qm_ctx *ctx, double *d_psi, int nb, int ik, double dt, void *stream);
int qm_td_kinetic_fft_f32(qm_ctx *ctx, float *d_psi, int nb, int ik, double dt, void *stream);

/* d_psi <- e^{-i V_loc dt} d_psi pointwise on nb SoA complex bands, the phase
 * evaluated in FP64 (mProductExpToV with exp_v = cos(-dt V) + i sin(-dt V),
 * qumasun_symplectic.evolution.h:80-100, 140-165).  Asynchronous on `stream`. */
int qm_td_expv_f64(qm_ctx *ctx, double *d_psi, int nb, double dt, void *stream);
int qm_td_expv_f32(qm_ctx *ctx, float *d_psi, int nb, double dt, void *stream);

/* ------------------------------------------------- nonlocal projection autotuner */

/* Shape autotuner of the nonlocal projection (no counterpart in the reference,
 * whose per-atom DGEMMs have one shape): column tiles of 16 per projection
 * workgroup, 1, 2 or 4 (at most 2 when some atom has more than 16 projectors).
 * The first projection of a (precision, ncomp, k type, padding, band-count
 * bucket) shape times the candidates on the caller's buffers and keeps the
 * fastest; it synchronises the stream once.  Every shape sums each column over
 * the same voxels in the same order, so the choice never changes a result.
 * Under stream capture, and in qm_hpsi_apply_f32_noalloc, an untuned shape
 * takes the size rule (4 tiles when that leaves >= 1024 workgroups, else 1).
 * qm_nonlocal_force   use `tiles` for every later projection, 0 = autotune;
 * qm_nonlocal_choice  tiles in use for that shape (kt bit 0: a k-point), -1 if not tuned yet. */
int qm_nonlocal_force(qm_ctx *ctx, int tiles);
int qm_nonlocal_choice(qm_ctx *ctx, int precision, int ncomp, int kt, int nb, int *tiles);

/* ---------------------------------------------------------------- adjoint */

/* d_rho[i] += sum_n occ[n] |psi_n(i)|^2 (FP64 accumulation).  dE_band/dV_loc(i) is
 * d_rho[i]*dV.  Replaces SetDensityOne (qumasun_SetDensityOne.h:9-41).  occ: host, nb. */
int qm_band_density_f64(qm_ctx *ctx, double *d_rho, const double *d_psi, const double *occ, int nb,
                        int ncomp, void *stream);
int qm_band_density_f32(qm_ctx *ctx, double *d_rho, const float *d_psi, const double *occ, int nb,
                        int ncomp, void *stream);

/* Nonlocal energy E_nl = sum_n occ_n sum_w E_w |<p_w|psi_n>|^2 and forces
 * F_a = -dE_nl/dR_a = 2 sum_n occ_n sum_w E_w Re(<p_w|psi_n>^* <d p_w/d r|psi_n>)
 * (host outputs: forces[3*natoms], e_nl[1]).  Replaces ForceNonlocal_bundle
 * (PseudoPotNonlocal_SvF2_kpoint.h:3629-3731) with mInnerNonlocalPsi_diff_bundle
 * (4775-4880).  Synchronises `stream`. */
int qm_nonlocal_forces_f64(qm_ctx *ctx, const double *d_psi, const double *occ, int nb, int ncomp,
                           int ik, double *forces, double *e_nl, void *stream);
int qm_nonlocal_forces_f32(qm_ctx *ctx, const float *d_psi, const double *occ, int nb, int ncomp,
                           int ik, double *forces, double *e_nl, void *stream);

/* ---------------------------------------------------------------- dense products */

/* Tall-skinny Gram product C = A^T B: A is k x m, B is k x n (column-major,
 * leading dimension k, device), C is m x n FP64 column-major (device, ldc = m).
 * FP32 inputs are multiplied on MFMA in FP32 and summed in FP64 across K blocks.
 * Replaces blas_DGEMM_t(M, N, K, A, B, C, ldc, 1, 0) (gyield/gy_w_dgemm.h:314-318,
 * rocblas_dgemm at 14-38) as used for the Rayleigh-Ritz blocks
 * (lobpcg_z_multi_mpi.h:300-558) and the projector inner products. */
int qm_gram_f64(qm_ctx *ctx, const double *d_A, int m, const double *d_B, int n, int64_t k, double *d_C,
                void *stream);
int qm_gram_f32(qm_ctx *ctx, const float *d_A, int m, const float *d_B, int n, int64_t k, double *d_C,
                void *stream);

/* Y = Z C: Z is k x m (column-major, ld k, device), C is m x n (column-major, ld m,
 * device, same type), Y is k x n (ld k, device).  Replaces blas_DGEMM_n
 * (gyield/gy_w_dgemm.h:356-360) as used by NextVector_z2 (lobpcg_z_multi_mpi.h:1515-1598). */
int qm_update_f64(qm_ctx *ctx, const double *d_Z, int m, const double *d_C, int n, int64_t k, double *d_Y,
                  void *stream);
int qm_update_f32(qm_ctx *ctx, const float *d_Z, int m, const float *d_C, int n, int64_t k, float *d_Y,
                  void *stream);

/* Generalized symmetric eigenproblem A z = l B z on the GPU in FP64, in place of
 * lapack_DSYGVD(UPLO='U', A, B, eigen_values, eigen_vectors, triangle_factor, N)
 * (w_dsygvd.h:14-72), following the Cholesky route of lapack_ZHEGVD_emu (w_zhegvd.h:227-272) in real arithmetic.
 * d_A, d_B: n x n column-major, upper triangles referenced, not modified.
 * d_w: all n eigenvalues, ascending.  d_Z: n x nev, eigenvectors of the nev
 * lowest, B-orthonormal.  d_U: n x n, upper triangle receives the Cholesky
 * factor of B (strictly lower part unspecified).  n <= 4096.  Synchronises the
 * stream.  QM_ERR_NUMERIC when B is not positive definite. */
int qm_sygvd_f64(qm_ctx *ctx, int n, const double *d_A, const double *d_B, double *d_w, double *d_Z, double *d_U,
                 int nev, void *stream);

/* ---------------------------------------------------------------- LOBPCG */

/* Block LOBPCG solver of nbands lowest states of H(k).  precision = 64: vectors,
 * H.psi and Gram products in FP64.  precision = 32: vectors, H.psi, Gram and
 * update products in FP32 (Gram partial sums reduced in FP64); the Cholesky,
 * generalized subspace eigensolve and residual/norm sums stay FP64.
 * Replaces LOBPCG::Eigen_z_multi_mpi (lobpcg_z_multi_mpi.h:1732-2852) with its
 * work arena (1779-1786, 2016-2022).  The solver keeps the P block between calls
 * (the reference's `keep`, qumasun_kpoint.h:83-100). */
int qm_lobpcg_create(qm_lobpcg **out, qm_ctx *ctx, int nbands, int ncomp, int precision);
void qm_lobpcg_destroy(qm_lobpcg *s);

/* Runs iter_max iterations on d_x (device, nbands bands in the layout above,
 * FP64 for precision 64, FP32 for 32; normalised so that sum |x|^2 dV = 1 on
 * return, as lobpcg_z_multi_mpi.h:2796-2800).  first_time != 0 starts with P = 0
 * (SCF_step == 0, lobpcg_z_multi_mpi.h:2052-2056); otherwise the kept P is reused
 * and A.P recomputed (2058-2065).  Host outputs: eigenvalues[nbands] (Ha),
 * residual_norms[nbands] (|| H x - e x ||^2 of the last residual, may be NULL).
 * For precision 32 the returned eigenvalues are the FP64 Rayleigh quotients
 * <x|H|x>/<x|x> of the returned vectors under the FP64 operator and the residual
 * norms are FP64 (one FP64 H.psi per call).
 * Returns the number of iterations done through *iters_done (may be NULL); the
 * reference's early exit on loss of normalisation (|1 - |x|^2| > 1e-3,
 * 2611-2666) is kept. */
int qm_lobpcg_solve(qm_lobpcg *s, void *d_x, double *eigenvalues, double *residual_norms, int iter_max,
                    int ik, int first_time, int *iters_done, void *stream);

/* qm_lobpcg_solve with the reference's locking of converged low bands: the
 * p_num_fix / residual_threshold arguments of Eigen_z_multi_mpi
 * (lobpcg_z_multi_mpi.h:1732-1742).
 * - *num_fix (in/out, 0 <= *num_fix <= nbands/2) is the caller's count of fixed
 *   lowest bands.  It is kept between calls and reset to 0 by the caller, as
 *   m_num_fix_state_sk is at each SCF step (qumasun_kpoint.scfloop.h:76-80).
 * - On entry H.x and the residuals are formed for the bands at or above
 *   *num_fix.  The new count is the first such band whose ||r||^2 (x scaled to
 *   sum |x|^2 = 1) exceeds residual_threshold^2, capped at nbands/2 (1962-1986).
 * - Only the bands above the count iterate.  Their R, and P when the count
 *   changed, first lose their components along the fixed bands:
 *   V - X_f (X_f^H V), VectorOrthogonalization (1637-1723).
 * - The fixed bands of d_x come back as passed, up to the rounding of the
 *   sqrt(dV) scaling round trip.  Their eigenvalues[] and residual_norms[]
 *   entries are left as the caller passed them (in/out arrays, 2846-2848).
 * - iter_max = 0 still updates *num_fix and, like the reference, leaves before
 *   touching P. */
int qm_lobpcg_solve_locking(qm_lobpcg *s, void *d_x, double *eigenvalues, double *residual_norms, int iter_max,
                            int ik, int first_time, int *num_fix, double residual_threshold, int *iters_done,
                            void *stream);

/* The subspace problem H c = l S c is solved as in lapack_ZHEGVD_emu (w_zhegvd.h:227-272):
 * Cholesky of S, two triangular solves, a standard eigensolve, back substitution.
 * When S is not positive definite, or the squared Cholesky pivots span less than
 * this ratio (min/max; defaults 1e-14 for precision 64, 1e-4 for precision 32), the
 * step is redone without the P block, and then on the eigenbasis of S with the
 * near-null directions dropped.  No counterpart in the reference, whose
 * LAPACK info is only printed (gy_w_zhegvd_hip.h:58-61). */
int qm_lobpcg_set_conditioning(qm_lobpcg *s, double min_pivot_ratio_fp64, double min_pivot_ratio_fp32);
/* Number of such fallbacks taken in the last qm_lobpcg_solve (or QM_ERR_ARG).
 * A ratio above 1 (at most 2) fails every pivot test, so every step takes the
 * S-eigenbasis path: a way to exercise the fallback. */
int qm_lobpcg_fallbacks(qm_lobpcg *s);
/* Subspace eigensolves of the last qm_lobpcg_solve that had to be redone: a
 * divide-and-conquer tridiagonal solve that did not converge (redone by
 * QL/QR), a heevd / syevd that did not converge (redone by Jacobi), a
 * tridiagonal reduction with non-finite output (redone by heevd), or Cholesky-
 * route Ritz vectors rejected because max_i |1 - c_i^H S c_i| exceeded 1e-4 (S
 * too ill-conditioned for its Cholesky factor; redone without P, then on the
 * S eigenbasis, and also counted by qm_lobpcg_fallbacks).  note
 * (may be NULL) receives a description of the last one, at most len bytes.
 * No counterpart in the reference, which prints LAPACK's info. */
int qm_lobpcg_eigen_retries(qm_lobpcg *s, char *note, int len);

/* The |1 - |x'|^2| limit above which an update is rolled back: P = R, the
 * previous vectors are returned (renormalised) with their Ritz values and
 * residual norms, and the call leaves its loop (LIMIT_Z_NORM_X_FOR_RESET_MATRIX,
 * lobpcg_z_multi_mpi.h:101, 2611-2666; default 1e-3).  For precision 32 the
 * returned values are FP64 Rayleigh quotients of the returned vectors in either
 * case.  limit = 0 rolls back the first update of a call. */
int qm_lobpcg_set_rollback_limit(qm_lobpcg *s, double limit);

/* Generalized Hermitian eigenproblem A z = l B z in complex FP64, in place of
 * lapack_ZHEGVD_emu (w_zhegvd.h:227-272): B = U^H U (rocSOLVER zpotrf),
 * C = U^-H A U^-1 (two ztrsm), the cooperative tridiagonal reduction of C,
 * stedc (steqr where stedc does not converge) and the back-transformations.
 * Matrices are n x n column-major, complex interleaved (re, im) doubles;
 * d_A and d_B: upper triangles referenced, not modified.  d_w: all n
 * eigenvalues, ascending.  d_Z: n x nev, the nev lowest eigenvectors,
 * B-orthonormal.  d_U: upper triangle receives the Cholesky factor of B.
 * n <= 4096.  Synchronises the stream.  QM_ERR_NUMERIC when B is not positive
 * definite.  Where the cooperative reduction does not fit (or
 * qm_subspace_force(0)), rocSOLVER heevd diagonalises C. */
int qm_hegvd_z64(qm_ctx *ctx, int n, const void *d_A, const void *d_B, double *d_w, void *d_Z, void *d_U, int nev,
                 void *stream);

/* Route of the subspace eigensolve's tridiagonal reduction, process-wide (no
 * reference counterpart: LAPACK / rocSOLVER's sytrd and hetrd are the
 * reference's).  By default (-1) the real and complex solves of
 * qm_lobpcg_solve and qm_sygvd_f64 reduce the subspace matrix with G
 * cooperating workgroups that hold it in LDS (one grid barrier per column),
 * where it fits (real n up to ~1700, complex n up to ~1200) and the device
 * accepts cooperative launches; otherwise the single-workgroup panel
 * reduction (real n <= 512) or rocSOLVER syevd / heevd.  0 never takes the
 * cooperative route; 1 takes it wherever it fits.  Results agree to rounding
 * (different summation orders). */
int qm_subspace_force(int route);

/* Bound of the grid-barrier waits in the cooperative kernels (the tridiagonal
 * reduction and the Cholesky), in polls, process-wide; no reference
 * counterpart (a diagnostic and test control).  A workgroup that waits past
 * it sets the launch's error word and stops waiting, so does every other
 * waiter within 256 polls, the kernel ends, and the solve reports it
 * (info -77: qm_sygvd_f64 / qm_hegvd_z64 fail with QM_ERR_NUMERIC,
 * qm_lobpcg_solve takes its fallback route).  The default,
 * qm_coop_spin_limit_default() (2^24, ~0.5 s), is far above any barrier's
 * wait; 0 makes every wait that is not over at its first poll fail, which
 * the tests use to exercise that path.  QM_ERR_ARG for polls < 0. */
int qm_coop_set_spin_limit(long polls);
long qm_coop_spin_limit_default(void);

/* HIP graphs for the iteration segments that hold no host decision (no
 * reference counterpart: the reference launches every kernel from the host).
 * Within one qm_lobpcg_solve call, A.R and the Gram products of an iteration
 * are captured the second time their block buffers recur and replayed after
 * that.  The first occurrence always runs eagerly, so buffer sizes and the
 * stencil shape have settled.  Replays launch the same kernels with the same
 * arguments, so results are bit-identical to eager execution.  The graphs are
 * destroyed when the call returns.  The Cholesky status, the pivot-ratio tests
 * and the normalisation rollback stay host decisions between segments, as in
 * the reference's control flow, and rocSOLVER calls are not captured.
 * enable = 0 runs everything eagerly.  Default: on. */
int qm_lobpcg_set_graphs(qm_lobpcg *s, int enable);
/* Graph replays in the last qm_lobpcg_solve (or QM_ERR_ARG). */
int qm_lobpcg_graph_replays(qm_lobpcg *s);

/* Milliseconds spent in the last qm_lobpcg_solve per phase:
 * t[0] H.psi, t[1] Gram, t[2] subspace eigensolve, t[3] updates, t[4] other, t[5] total. */
int qm_lobpcg_timings(qm_lobpcg *s, double *t6);

/* ------------------------------------------------------------ SCF glue */

/* The steps of QUMASUN_KPOINT::mSCF's loop (qumasun_kpoint.scfloop.h:96-222)
 * between one LOBPCG solve and the next: occupations, the band density
 * (qm_band_density_*), density mixing, and the new potential
 * V_tot = V_ion + V_H + V_xc (qm_set_vloc).  Fields are FP64 device arrays of
 * Ng points in the grid order. */

/* V_H from the FFT Poisson solve, V_H(G) = 4 pi rho(G) / |G|^2 with the G = 0
 * term dropped (SetPotentialByPoissonFFT, poisson_fft.h:16-120; the
 * reference's refined-grid up/down conversion is not done).  Uses the
 * context's rocFFT plans and spectrum buffer (ordered with its applies). */
int qm_hartree_f64(qm_ctx *ctx, const double *d_rho, double *d_vh, void *stream);

/* LDA exchange-correlation (Calc_XC_LDA, Vxc.h:17-76: Perdew-Zunger 1981):
 * d_vxc = V_x + V_c per point, and, if d_exc is not NULL, the energy density
 * rho (e_x + e_c).  Points with rho < 1e-8 get zeros, as in the reference. */
int qm_xc_lda_f64(qm_ctx *ctx, const double *d_rho, double *d_vxc, double *d_exc, void *stream);

/* *result = sum_i a_i b_i over the grid (no volume element), summed in a fixed
 * order (the same bits on every call).  Synchronises the stream. */
int qm_grid_dot_f64(qm_ctx *ctx, const double *d_a, const double *d_b, double *result, void *stream);

/* Density mixer on device fields: MixingSimple (mixing_simple.h:45-75) or
 * MixingDIIS (mixing_DIIS.h:48-229, the reference's default: a ring of
 * `history` densities and residuals, simple mixing until it is full, then the
 * coefficients from the bordered residual-overlap matrix inverted by
 * Gauss-Jordan as InverseM does).  qm_mixer_init takes the first density;
 * qm_mixer_predict replaces d_rho (the density from the bands) with the
 * mixed one and returns sum_i |rho_out - rho_in| in *diff (the reference
 * then multiplies by dV / N_e).  Synchronises the stream. */
#define QM_MIX_SIMPLE 0
#define QM_MIX_DIIS 1
typedef struct qm_mixer qm_mixer;
int qm_mixer_create(qm_mixer **out, qm_ctx *ctx, int mode, int history, double ratio);
void qm_mixer_destroy(qm_mixer *m);
int qm_mixer_init(qm_mixer *m, const double *d_rho, void *stream);
int qm_mixer_predict(qm_mixer *m, double *d_rho, double *diff, void *stream);

/* Occupations of nb bands at nk k-points (host arrays, eig[k*nb + n]
 * ascending per k-point; kweight[k] >= 0): below kT = 0.1 K the aufbau
 * filling of mSetOccupancyZero (qumasun_base1.occupancy.h:35-67; *mu = NaN),
 * else Fermi-Dirac with mu found by bisection to 1e-9 Ha as in
 * mSetOccupancyTemperature (occupancy.h:76-175, SetOccupancy.h).  Spin
 * unpolarised: each band holds 2 kweight[k] / sum(kweight) electrons at most. */
int qm_occupancy(int nk, int nb, const double *eig, const double *kweight, double nelec, double kbT, double *occ,
                 double *mu);

/* ------------------------------------------------- UPF pseudopotentials (host) */

/* A norm-conserving UPF v2 pseudopotential's nonlocal part, loaded on the host.
 * Replaces the projector fields LoadPseudoPotential_UPF_ONCV fills in
 * PseudoPot_MBK (upf_oncv_loader.h:355-530, pp_struct.h:5-96): number_of_proj
 * radial projectors beta_k with angular momentum l_k <= 3 and cutoff radius,
 * KB energies D_kk / 2 (Ry -> Ha), and z_valence.  The radial functions
 * f_k = beta_k / r are resampled as the reference does: onto a 1000-point
 * logarithmic grid from the uniform UPF mesh (UniformGrid::GetValue), then
 * evaluated by RadialGrid2::GetValueByXi, zero beyond the largest cutoff of
 * that l.  At most 32 projectors per atom (sum of 2 l_k + 1).  The mesh must be
 * uniform (ONCV meshes are).  These calls touch no GPU. */
typedef struct qm_upf qm_upf;
int qm_upf_load(qm_upf **out, const char *path);
void qm_upf_free(qm_upf *u);
/* radial projector count, projectors per atom (sum of 2 l + 1), valence
 * charge and the largest cutoff radius (bohr); any pointer may be NULL */
int qm_upf_info(const qm_upf *u, int *nradial, int *nproj, double *z_valence, double *cutoff);
/* projector k: l, its cutoff_radius and KB energy (Ha); pointers may be NULL */
int qm_upf_radial(const qm_upf *u, int k, int *l, double *cutoff, double *energy);
/* f_k(r) and df_k/dr (df may be NULL) at n radii r (bohr) */
int qm_upf_eval(const qm_upf *u, int k, int n, const double *r, double *f, double *df);
/* Projector tables in the qm_set_projectors layout for natoms atoms of this
 * species at positions[3a..3a+2] (bohr) on a grid of spacing hx, hy, hz: every
 * grid point (unwrapped integer coordinates ix*hx, ...) strictly within the
 * largest cutoff of an atom, ordered z, y, x; projectors in the reference's
 * order k, m = -l..l (num_proj_RY, PseudoPotNonlocal_SvF2_kpoint.h:940-966)
 * valued f_k(|r|) Y_lm(r/|r|) (real Y_lm without the Condon-Shortley sign,
 * RealSphericalHarmonics.h:148-162), r = grid point - atom, and the gradient
 * tables dtable (may be NULL).  counts[0..2] receive the voxel, table and
 * projector totals; with any output array NULL only counts are written, so a
 * caller sizes its arrays with a first call.  nx, ny, nz are checked but the
 * coordinates are not wrapped (qm_set_projectors wraps them).  The reference
 * filters its tables in Fourier space on a refined subgrid
 * (..._kpoint.h:911-1178, 1912-2208); these are the unfiltered point values. */
int qm_upf_tables(const qm_upf *u, int natoms, const double *positions, int nx, int ny, int nz, double hx,
                  double hy, double hz, int64_t *counts, int *nproj, int64_t *vox_off, int *vox_xyz,
                  double *table, double *dtable, int64_t *tab_off, double *energy, int *proj_off);

#ifdef __cplusplus
}
#endif

#endif /* QUMASUN_HIP_H */
