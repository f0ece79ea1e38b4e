// Internal declarations shared by the HIP translation units of libqumasun_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/qumasun_hip.h"
#include "projector_maps.h"

namespace qm {

void set_error(const std::string &msg);

// Records a HIP failure with the call site and returns QM_ERR_HIP from the
// enclosing function.
#define QM_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            ::qm::set_error(std::string(#call) + ": " + hipGetErrorString(e_) + " at " +     \
                            __FILE__ + ":" + std::to_string(__LINE__));                      \
            return QM_ERR_HIP;                                                               \
        }                                                                                    \
    } while (0)

#define QM_CHECK(cond, code, msg)                                                            \
    do {                                                                                     \
        if (!(cond)) {                                                                       \
            ::qm::set_error(msg);                                                            \
            return (code);                                                                   \
        }                                                                                    \
    } while (0)

#define QM_TRY(expr)                      \
    do {                                  \
        int s_ = (expr);                  \
        if (s_ != QM_OK) return s_;       \
    } while (0)

// Projectors per atom: at most 32 (two radial projectors for each l <= 3, as
// in W.upf: 2 x (1 + 3 + 5 + 7)).  Per-projector-set padding (Projectors::mpad)
// is 16 when every atom has at most 16, else 32: the MFMA row tiles of the
// projection and the table rows of the scatter are 16 wide.
constexpr int kMaxProj = 32;

// Device buffer with grow-only reallocation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int reserve(size_t n);
    void release();
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

// Projector set on device.  Voxels are stored atom-major; each atom's voxels
// are ordered z, y, x so consecutive voxels are mostly consecutive grid points.
struct Projectors {
    int natoms = 0;
    int ptot = 0;          // sum_a M_a
    int64_t vtot = 0;      // sum_a V_a
    int64_t ttot = 0;      // sum_a V_a M_a
    int max_m = 0;
    int mpad = 16;         // padded projector count of the voxel-major tables (16 or 32)
    bool has_grad = false;
    std::vector<int> nproj, proj_off;
    std::vector<int64_t> vox_off, tab_off;
    std::vector<int> vox_xyz;            // host copy of unwrapped coordinates (phases)
    std::vector<double> energy;
    // device arrays
    int *d_nproj = nullptr, *d_proj_off = nullptr;
    int64_t *d_vox_off = nullptr, *d_tab_off = nullptr;
    int *d_vox_grid = nullptr;           // wrapped grid index per voxel
    int *d_vox_atom = nullptr;           // owning atom per voxel
    double *d_energy = nullptr;
    double *d_table64 = nullptr, *d_dtable64 = nullptr;
    float *d_table32 = nullptr, *d_dtable32 = nullptr;
    // owner-computes scatter map: covered grid points, each with its list of voxels
    int ncov = 0;
    int *d_cov_grid = nullptr;           // [ncov] grid index
    int *d_cov_off = nullptr;            // [ncov+1]
    int *d_cov_vox = nullptr;            // [vtot] global voxel ids grouped by covered point
    // atom work list for the projection launch, longest atoms first
    int *d_atom_order = nullptr;
    // voxel-major tables padded to mpad projectors: tabp[v * mpad + w] (MFMA operand layout)
    double *d_tabp_64 = nullptr;
    float *d_tabp_32 = nullptr;
    // projection work items: chunks of kProjChunk voxels of one atom
    int nchunk = 0;
    int *d_chunk_atom = nullptr, *d_chunk_v0 = nullptr;
    int *d_item_order = nullptr;  // launch order of the projection work items (a permutation)
    int *d_atom_chunk_off = nullptr;  // [natoms + 1]
    int *d_proj_atom = nullptr;       // [ptot] owning atom of each projector
    // scatter blocks of kScatterBlock covered points: the atoms they touch
    // (block_atom_off / block_atoms) and each entry's slot in that list (-1 if
    // the block touches more than kScatterSlots atoms)
    int nsblock = 0;
    int *d_block_atom_off = nullptr, *d_block_atoms = nullptr, *d_cov_slot = nullptr;
};


struct FftState;  // rocFFT plans and buffers of the spectral kinetic term (kinetic_fft.hip)

struct KPoint {
    double k[3] = {0, 0, 0};
    bool gamma = true;
    double *d_phase64 = nullptr;  // [2*vtot]: cos, sin interleaved
    float *d_phase32 = nullptr;
};

}  // namespace qm

struct qm_ctx {
    int device = 0;
    int nx = 0, ny = 0, nz = 0;
    int64_t ng = 0;
    double hx = 0, hy = 0, hz = 0;
    bool has_vloc = false;
    double *d_vloc64 = nullptr;
    float *d_vloc32 = nullptr;
    qm::Projectors proj;
    std::vector<qm::KPoint> kpts;
    qm::DevBuf coef;     // projector coefficients for the nonlocal apply
    qm::DevBuf fwork;    // force partials
    qm::DevBuf gemm_part; // Gram partial sums of qm_gram_*
    qm::DevBuf nl_part;   // per-chunk projection partials (FP64)
    std::map<int64_t, int> stencil_choice;  // autotuned stencil candidate per shape
    int stencil_force = -1;                 // >= 0: use this candidate, no tuning
    std::map<int64_t, int> nl_choice;       // autotuned column tiles per projection workgroup per shape
    int nl_force = 0;                       // > 0: use this many column tiles, no tuning
    bool nl_notime = false;                 // untuned projection shapes take the size rule (no timing)
    void *blas = nullptr;                   // rocblas_handle of qm_sygvd_f64, created on first use
    qm::DevBuf eig_work;                    // qm_sygvd_f64 workspace
    qm::DevBuf td_work;                     // two band blocks of the Taylor propagator
    qm::DevBuf scf_part;                    // partial sums of qm_grid_dot_f64
    // second stream of a full H.psi apply: the nonlocal projection runs on it
    // beside the stencil, joined by events (created on first use)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // Set when an H.psi apply on this context was recorded into a caller's
    // graph (stream capture): the graph holds the addresses of the tables,
    // phases and work buffers, so calls that would free or move them refuse
    // (QM_ERR_STATE) until qm_release_captures.
    int capture_pins = 0;
    int kinetic = QM_KINETIC_FD8;  // kinetic term of every apply (qm_set_kinetic)
    qm::FftState *fft = nullptr;   // created on the first spectral apply
};

namespace qm {

int activate(const qm_ctx *ctx);
// H.psi of capi.hip (part: 3 = full, overwrite; 1 = local only, overwrite; 2 = nonlocal, accumulate)
template <typename T>
int hpsi_part(qm_ctx *ctx, T *out, const T *in, int nb, int ncomp, int ik, hipStream_t st, int part);
// QM_ERR_STATE (with a message naming `who`) while a caller's graph holds an
// apply of this context
int check_unpinned(const qm_ctx *ctx, const char *who);
// Nonzero while the library captures a graph of its own (the LOBPCG segment
// graphs, destroyed before their call returns): applies recorded then do not pin.
extern thread_local int g_internal_capture;

// stencil.hip
int launch_fd8_local_f64(qm_ctx *ctx, double *out, const double *in, int nb, int ncomp,
                         const double k[3], hipStream_t st);
int launch_fd8_local_f32(qm_ctx *ctx, float *out, const float *in, int nb, int ncomp,
                         const double k[3], hipStream_t st);
int stencil_num_candidates();
std::string stencil_candidate_name(int cand);
int stencil_choice(const qm_ctx *ctx, int precision, int ncomp, int kt, int nb);
// key of the tile shape a full H.psi apply (stencil beside the nonlocal
// projection) uses, tuned on full applies; and whether a shape can run
int64_t stencil_full_key(int precision, int ncomp, int kt, int nb);
bool stencil_feasible(const qm_ctx *ctx, int cand, int precision, int ncomp);

// eigen.hip: FP64 symmetric eigensolver pieces (rocblas_handle passed as void *)
int sytrd_lower(void *blas, int n, double *C, int lda, double *trd, double *VW, hipStream_t st);
int syev_lower(void *blas, int n, int nev, double *C, int lda, double *w, double *Z, int ldz, double *trd,
               double *VW, int *info, hipStream_t st);
int own_eig_max_n();
// A = U^T U, upper triangle of the column-major n x n A (FP64): with sync (512
// bytes of device memory) and n <= 512, one workgroup per 32-row panel
// (potrf_coop_f64); else one workgroup for n <= own_eig_max_n(), rocSOLVER
// above; *info (device) as LAPACK's dpotrf.
int potrf_upper(void *blas, int n, double *A, int lda, int *info, hipStream_t st, void *sync = nullptr);
// trd_coop.hip: the cooperative Cholesky (QM_ERR_STATE when it cannot run: the
// caller takes another kernel); *info as dpotrf, -77 if a barrier wait timed out
int potrf_coop_supported(int n);
int potrf_coop_f64(int n, double *A, int lda, int *info, void *sync, hipStream_t st);
// C := U^-T C U^-1 with Ui = U^-1 kept (n x n each, leading dimension n; T: n x n work),
// and Zout = Ui Z: the two halves of the Cholesky route around the eigensolve.
int chol_transform(void *blas, int n, const double *U, double *C, double *Ui, double *T, hipStream_t st);
int chol_back(void *blas, int n, int nev, const double *Ui, const double *Z, int ldz, double *Zout, int ldo);
constexpr int kTrdPanelCols = 96;  // VW workspace columns (n x 96) of sytrd_lower

// trd_coop.hip: tridiagonal reduction by cooperating workgroups (matrix in LDS),
// LAPACK's lower-storage output; reads A's upper triangle.  Returns
// QM_ERR_STATE when the order does not fit or the cooperative launch is
// refused (the caller then takes another route).  work: trd_coop_work_bytes;
// trd_coop_check merges the kernel's barrier error word into *info (call it
// after the routine that writes info).  Complex matrices are interleaved
// (re, im) FP64, tau likewise.
size_t trd_coop_work_bytes(int n, bool cplx);
int trd_coop_supported(int n, bool cplx);  // workgroups it would use, 0 if none
int trd_coop_f64(int n, double *A, int lda, double *d, double *e, double *tau, void *work, size_t work_bytes,
                 hipStream_t st);
int trd_coop_z64(int n, void *A, int lda, double *d, double *e, void *tau, void *work, size_t work_bytes,
                 hipStream_t st);
int trd_coop_check(void *work, int *info, hipStream_t st);
// Z := Q Z (first nev columns) for the Q of a LAPACK lower tridiagonal
// reduction (reflectors below A's subdiagonal, tau): dormtr / zunmtr with
// side L, uplo L, no transpose; n <= 2048.  work: the cooperative reduction's
// workspace (trd_coop_work_bytes), after trd_coop_check has read it
int ormtr_lower_f64(int n, int nev, const double *A, int lda, const double *tau, double *Z, int ldz, void *work,
                    hipStream_t st);
int unmtr_lower_z64(int n, int nev, const void *A, int lda, const void *tau, void *Z, int ldz, void *work,
                    hipStream_t st);
// route of the subspace eigensolve's tridiagonal reduction (qm_subspace_force):
// -1 automatic (cooperative where it fits), 0 never cooperative, 1 cooperative
// wherever it fits
int subspace_route();

// kinetic_fft.hip: out = IFFT(|G+k|^2/2 FFT psi)/N + V_loc psi (overwrites); context teardown
template <typename T>
int launch_fft_local(qm_ctx *ctx, T *out, const T *psi, int nb, int ncomp, const double k[3], hipStream_t st);
void fft_destroy(qm_ctx *ctx);

// nonlocal.hip
int launch_nonlocal_f64(qm_ctx *ctx, double *hpsi, const double *psi, int nb, int ncomp,
                        const KPoint &kp, hipStream_t st);
int launch_nonlocal_f32(qm_ctx *ctx, float *hpsi, const float *psi, int nb, int ncomp,
                        const KPoint &kp, hipStream_t st);
// the two halves of launch_nonlocal_*: projection coefficients (reads psi,
// writes the context's coefficient buffer), then the scatter (hpsi +=)
int nonlocal_project_f64(qm_ctx *ctx, const double *psi, int nb, int ncomp, const KPoint &kp, hipStream_t st);
int nonlocal_project_f32(qm_ctx *ctx, const float *psi, int nb, int ncomp, const KPoint &kp, hipStream_t st);
int nonlocal_scatter_f64(qm_ctx *ctx, double *hpsi, int nb, int ncomp, const KPoint &kp, hipStream_t st);
int nonlocal_scatter_f32(qm_ctx *ctx, float *hpsi, int nb, int ncomp, const KPoint &kp, hipStream_t st);
// Autotuned column tiles per projection workgroup for a shape (kt bit 0: a k-point), -1 if untuned
int nonlocal_choice(const qm_ctx *ctx, int precision, int ncomp, int kt, int nb);
// Bytes of the nonlocal work buffers for nb bands (coefficients, per-chunk partials)
void nonlocal_work_bytes(const qm_ctx *ctx, int nb, int ncomp, size_t tsize, size_t *coef_bytes,
                         size_t *part_bytes);
// launch_nonlocal_f32 with caller-provided work buffers (which: 1 project, 2 scatter, 3 both)
int nonlocal_apply_in_f32(qm_ctx *ctx, float *hpsi, const float *psi, int nb, int ncomp, const KPoint &kp,
                          int which, float *coef, double *part, hipStream_t st);
int nonlocal_forces_f64(qm_ctx *ctx, const double *psi, const double *occ, int nb, int ncomp,
                        const KPoint &kp, double *forces, double *e_nl, hipStream_t st);
int nonlocal_forces_f32(qm_ctx *ctx, const float *psi, const double *occ, int nb, int ncomp,
                        const KPoint &kp, double *forces, double *e_nl, hipStream_t st);

// blas.hip: logical matrix made of up to kMaxColBlocks column blocks, each with
// leading dimension Ng (column j of block b at ptr[b] + j*Ng)
constexpr int kMaxColBlocks = 6;
struct ColBlocks {
    const void *ptr[kMaxColBlocks] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int ncols[kMaxColBlocks] = {0, 0, 0, 0, 0, 0};
    int nblk = 0;
    void add(const void *p, int n) {
        ptr[nblk] = p;
        ncols[nblk] = n;
        ++nblk;
    }
    int total() const {
        int t = 0;
        for (int b = 0; b < nblk; ++b) t += ncols[b];
        return t;
    }
};
template <typename T>
int gram(DevBuf &partial, double *C, const ColBlocks &A, const ColBlocks &B, int m, int n, int64_t ng,
         hipStream_t st);
template <typename T>
int update(const ColBlocks &Z, const ColBlocks &Y, const T *Cm, int m, int n, int64_t ng, hipStream_t st);
template <typename T>
int band_reduce(DevBuf &part, int mode, double *out0, double *out1, T *r, const T *x, const T *ax, const double *lam,
                int nb, int ncomp, int64_t ng, hipStream_t st);
template <typename T>
int scale_bands(T *y, const T *x, const double *f, double fconst, int nb, int ncomp, int64_t ng, hipStream_t st);

// adjoint.hip
int launch_band_density_f64(const qm_ctx *ctx, double *rho, const double *psi, const double *d_occ,
                            int nb, int ncomp, hipStream_t st);
int launch_band_density_f32(const qm_ctx *ctx, double *rho, const float *psi, const double *d_occ,
                            int nb, int ncomp, hipStream_t st);

}  // namespace qm
