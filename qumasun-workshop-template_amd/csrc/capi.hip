// C ABI entry points: context, potential, projector set, k-points, H.psi and
// the adjoint.  See include/qumasun_hip.h for the contract of each function.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>

#include <rocblas/rocblas.h>

#include "qm_internal.h"

namespace qm {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int DevBuf::reserve(size_t n) {
    if (n <= bytes) return QM_OK;
    if (p) {
        QM_HIP(hipFree(p));
        p = nullptr;
        bytes = 0;
    }
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc of " + std::to_string(n) + " bytes failed: " + hipGetErrorString(e));
        return QM_ERR_NOMEM;
    }
    bytes = n;
    return QM_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int activate(const qm_ctx *ctx) {
    QM_HIP(hipSetDevice(ctx->device));
    return QM_OK;
}

thread_local int g_internal_capture = 0;

int check_unpinned(const qm_ctx *ctx, const char *who) {
    QM_CHECK(ctx->capture_pins == 0, QM_ERR_STATE,
             std::string(who) + ": a graph captured an H.psi apply of this context and holds the addresses of its "
             "tables and work buffers; destroy that graph and call qm_release_captures first");
    return QM_OK;
}

// Whether st is capturing a graph for the caller (not one of the library's own).
static int caller_capturing(hipStream_t st, bool *out) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    QM_HIP(hipStreamIsCapturing(st, &cs));
    *out = cs != hipStreamCaptureStatusNone && g_internal_capture == 0;
    return QM_OK;
}

template <typename T>
static int upload(T **dst, const T *src, size_t n) {
    if (*dst) {
        QM_HIP(hipFree(*dst));
        *dst = nullptr;
    }
    if (n == 0) return QM_OK;
    hipError_t e = hipMalloc((void **)dst, sizeof(T) * n);
    if (e != hipSuccess) {
        *dst = nullptr;
        set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
        return QM_ERR_NOMEM;
    }
    QM_HIP(hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
    return QM_OK;
}

template <typename T>
static void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

static void free_projectors(Projectors &p) {
    dfree(p.d_nproj);
    dfree(p.d_proj_off);
    dfree(p.d_vox_off);
    dfree(p.d_tab_off);
    dfree(p.d_vox_grid);
    dfree(p.d_vox_atom);
    dfree(p.d_energy);
    dfree(p.d_table64);
    dfree(p.d_dtable64);
    dfree(p.d_table32);
    dfree(p.d_dtable32);
    dfree(p.d_cov_grid);
    dfree(p.d_cov_off);
    dfree(p.d_cov_vox);
    dfree(p.d_atom_order);
    dfree(p.d_tabp_64);
    dfree(p.d_tabp_32);
    dfree(p.d_chunk_atom);
    dfree(p.d_chunk_v0);
    dfree(p.d_item_order);
    dfree(p.d_atom_chunk_off);
    dfree(p.d_proj_atom);
    dfree(p.d_block_atom_off);
    dfree(p.d_block_atoms);
    dfree(p.d_cov_slot);
    p = Projectors();
}

static void free_kpoints(qm_ctx *ctx) {
    for (auto &k : ctx->kpts) {
        dfree(k.d_phase64);
        dfree(k.d_phase32);
    }
    ctx->kpts.clear();
}


// e^{ik.r} on every voxel for one k-point (mCreateBlochExp, ..._kpoint.h:3537-3618):
// phase = k_x h_x ix + k_y h_y iy + k_z h_z iz on the unwrapped coordinates.
static int build_phase(const qm_ctx *ctx, const Projectors &p, KPoint &kp) {
    dfree(kp.d_phase64);
    dfree(kp.d_phase32);
    if (kp.gamma || p.vtot == 0) return QM_OK;
    std::vector<double> ph(2 * p.vtot);
    std::vector<float> ph32(2 * p.vtot);
    const double gx = kp.k[0] * ctx->hx, gy = kp.k[1] * ctx->hy, gz = kp.k[2] * ctx->hz;
    for (int64_t v = 0; v < p.vtot; ++v) {
        const double a = gx * (double)p.vox_xyz[3 * v] + gy * (double)p.vox_xyz[3 * v + 1] +
                         gz * (double)p.vox_xyz[3 * v + 2];
        ph[2 * v] = std::cos(a);
        ph[2 * v + 1] = std::sin(a);
        ph32[2 * v] = (float)ph[2 * v];
        ph32[2 * v + 1] = (float)ph[2 * v + 1];
    }
    QM_TRY(upload(&kp.d_phase64, ph.data(), ph.size()));
    QM_TRY(upload(&kp.d_phase32, ph32.data(), ph32.size()));
    return QM_OK;
}

// The context's second stream and its two events, created on first use
// (default priority, non-blocking).
int side_stream(qm_ctx *ctx) {
    if (ctx->side) return QM_OK;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    QM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    if (hipEventCreateWithFlags(&e0, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e1, hipEventDisableTiming) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        (void)hipStreamDestroy(s);
        QM_CHECK(false, QM_ERR_HIP, "hipEventCreateWithFlags failed");
    }
    ctx->side = s;
    ctx->ev_fork = e0;
    ctx->ev_join = e1;
    return QM_OK;
}

// Builds the device projector set p from validated host arrays: voxel maps,
// tables (column-major and voxel-major padded to p.mpad), the projection work
// list and the owner-computes scatter map.  On failure p holds whatever was
// uploaded so far (the caller frees it).
static int build_projectors(const qm_ctx *ctx, Projectors &p, int natoms, const int *nproj, const int64_t *vox_off,
                            const int *vox_xyz, const double *table, const double *dtable, const int64_t *tab_off,
                            const double *energy, const int *proj_off) {
    p.natoms = natoms;
    p.nproj.assign(nproj, nproj + natoms);
    p.vox_off.assign(vox_off, vox_off + natoms + 1);
    p.tab_off.assign(tab_off, tab_off + natoms + 1);
    p.proj_off.assign(proj_off, proj_off + natoms + 1);
    for (int a = 0; a < natoms; ++a) p.max_m = std::max(p.max_m, nproj[a]);
    p.mpad = p.max_m > 16 ? 32 : 16;
    p.vtot = vox_off[natoms];
    p.ttot = tab_off[natoms];
    p.ptot = proj_off[natoms];
    p.vox_xyz.assign(vox_xyz, vox_xyz + 3 * p.vtot);
    p.energy.assign(energy, energy + p.ptot);
    p.has_grad = dtable != nullptr;

    ProjectorMaps m;
    build_projector_maps(ctx->nx, ctx->ny, ctx->nz, natoms, nproj, vox_off, vox_xyz, table, tab_off, proj_off,
                         p.mpad, m);
    p.ncov = (int)m.cov_grid.size();
    p.nchunk = (int)m.chunk_atom.size();
    p.nsblock = (int)m.block_atom_off.size() - 1;

    std::vector<float> t32(table, table + p.ttot);
    QM_TRY(upload(&p.d_nproj, nproj, natoms));
    QM_TRY(upload(&p.d_proj_off, proj_off, natoms + 1));
    QM_TRY(upload(&p.d_vox_off, vox_off, natoms + 1));
    QM_TRY(upload(&p.d_tab_off, tab_off, natoms + 1));
    QM_TRY(upload(&p.d_vox_grid, m.grid.data(), p.vtot));
    QM_TRY(upload(&p.d_vox_atom, m.atom.data(), p.vtot));
    QM_TRY(upload(&p.d_energy, energy, p.ptot));
    QM_TRY(upload(&p.d_table64, table, p.ttot));
    QM_TRY(upload(&p.d_table32, t32.data(), p.ttot));
    if (dtable) {
        std::vector<float> d32(dtable, dtable + 3 * p.ttot);
        QM_TRY(upload(&p.d_dtable64, dtable, 3 * p.ttot));
        QM_TRY(upload(&p.d_dtable32, d32.data(), 3 * p.ttot));
    }
    QM_TRY(upload(&p.d_cov_grid, m.cov_grid.data(), m.cov_grid.size()));
    QM_TRY(upload(&p.d_cov_off, m.cov_off.data(), m.cov_off.size()));
    QM_TRY(upload(&p.d_cov_vox, m.cov_vox.data(), m.cov_vox.size()));
    QM_TRY(upload(&p.d_atom_order, m.atom_order.data(), m.atom_order.size()));
    QM_TRY(upload(&p.d_item_order, m.item_order.data(), m.item_order.size()));
    QM_TRY(upload(&p.d_tabp_64, m.tabp64.data(), m.tabp64.size()));
    QM_TRY(upload(&p.d_tabp_32, m.tabp32.data(), m.tabp32.size()));
    QM_TRY(upload(&p.d_chunk_atom, m.chunk_atom.data(), m.chunk_atom.size()));
    QM_TRY(upload(&p.d_chunk_v0, m.chunk_v0.data(), m.chunk_v0.size()));
    QM_TRY(upload(&p.d_atom_chunk_off, m.atom_chunk_off.data(), m.atom_chunk_off.size()));
    QM_TRY(upload(&p.d_proj_atom, m.proj_atom.data(), m.proj_atom.size()));
    QM_TRY(upload(&p.d_block_atom_off, m.block_atom_off.data(), m.block_atom_off.size()));
    QM_TRY(upload(&p.d_block_atoms, m.block_atoms.data(), m.block_atoms.size()));
    QM_TRY(upload(&p.d_cov_slot, m.cov_slot.data(), m.cov_slot.size()));
    return QM_OK;
}

// The local part (kinetic + V_loc, overwrite) with the context's kinetic term
template <typename T>
int launch_local(qm_ctx *ctx, T *out, const T *in, int nb, int ncomp, const double k[3], hipStream_t st) {
    if (ctx->kinetic == QM_KINETIC_FFT) return launch_fft_local<T>(ctx, out, in, nb, ncomp, k, st);
    if constexpr (sizeof(T) == 8) return launch_fd8_local_f64(ctx, out, in, nb, ncomp, k, st);
    else return launch_fd8_local_f32(ctx, out, in, nb, ncomp, k, st);
}

// part: 3 = local + nonlocal (overwrite), 1 = local only (overwrite), 2 = nonlocal only (accumulate)
template <typename T>
int hpsi_apply(qm_ctx *ctx, T *hpsi, const T *psi, int nb, int ncomp, int ik, void *stream, int part = 3) {
    QM_CHECK(ctx && hpsi && psi, QM_ERR_ARG, "qm_hpsi_apply: NULL argument");
    QM_CHECK(nb >= 0 && (ncomp == 1 || ncomp == 2), QM_ERR_ARG, "qm_hpsi_apply: bad nb / ncomp");
    QM_CHECK(ik >= 0 && ik < (int)ctx->kpts.size(), QM_ERR_ARG, "qm_hpsi_apply: k index out of range");
    QM_CHECK(ctx->has_vloc, QM_ERR_STATE, "qm_hpsi_apply: call qm_set_vloc first");
    QM_CHECK((const void *)hpsi != (const void *)psi, QM_ERR_ARG, "qm_hpsi_apply: output aliases input");
    if (nb == 0) return QM_OK;
    QM_TRY(activate(ctx));
    hipStream_t st = (hipStream_t)stream;
    bool capturing = false;
    QM_TRY(caller_capturing(st, &capturing));
    if (capturing) ctx->capture_pins = 1;
    const KPoint &kp = ctx->kpts[ik];
    const bool nl = ctx->proj.natoms > 0 && ctx->proj.ptot > 0;
    if (part == 3 && nl) {
        // The projection reads psi only and the stencil writes hpsi only: the
        // projection (and its coefficient sum) runs on the context's second
        // stream beside the stencil, and the scatter waits for both.  The fork
        // event is recorded after everything the caller queued before, which
        // includes the previous apply's scatter, the last reader of the
        // coefficient buffer.
        QM_TRY(side_stream(ctx));
        const int prec = sizeof(T) == 8 ? 64 : 32;
        const int kt = (std::fabs(kp.k[0]) >= 1.0e-8 || std::fabs(kp.k[1]) >= 1.0e-8 ||
                        std::fabs(kp.k[2]) >= 1.0e-8) ? 1 : 0;
        auto full = [&]() -> int {
            QM_HIP(hipEventRecord(ctx->ev_fork, st));
            QM_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
            if constexpr (sizeof(T) == 8) QM_TRY(nonlocal_project_f64(ctx, psi, nb, ncomp, kp, ctx->side));
            else QM_TRY(nonlocal_project_f32(ctx, psi, nb, ncomp, kp, ctx->side));
            QM_HIP(hipEventRecord(ctx->ev_join, ctx->side));
            QM_TRY(launch_local<T>(ctx, hpsi, psi, nb, ncomp, kp.k, st));
            QM_HIP(hipStreamWaitEvent(st, ctx->ev_join, 0));
            if constexpr (sizeof(T) == 8) QM_TRY(nonlocal_scatter_f64(ctx, hpsi, nb, ncomp, kp, st));
            else QM_TRY(nonlocal_scatter_f32(ctx, hpsi, nb, ncomp, kp, st));
            return QM_OK;
        };
        if (ctx->stencil_force >= 0 || ctx->kinetic == QM_KINETIC_FFT) return full();
        // The stencil's tile shape for full applies is tuned on full applies:
        // the shape that runs fastest alone is not always the one that leaves
        // the projection the most room beside it.  Every shape gives the same
        // values, so the timed applies also leave the right result.
        struct Force {  // the shape for this call; restored on every return
            qm_ctx *c;
            ~Force() { c->stencil_force = -1; }
        } force{ctx};
        const int64_t key = stencil_full_key(prec, ncomp, kt, nb);
        auto it = ctx->stencil_choice.find(key);
        int cand = 0;
        if (it != ctx->stencil_choice.end()) {
            cand = it->second;
        } else {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            QM_HIP(hipStreamIsCapturing(st, &cs));
            if (cs == hipStreamCaptureStatusNone) {
                hipEvent_t ev[2] = {nullptr, nullptr};
                QM_HIP(hipEventCreate(&ev[0]));
                if (hipEventCreate(&ev[1]) != hipSuccess) {
                    (void)hipEventDestroy(ev[0]);
                    QM_CHECK(false, QM_ERR_HIP, "hipEventCreate failed");
                }
                float best_ms = 1e30f;
                int status = QM_OK;
                for (int c = 0; c < stencil_num_candidates() && status == QM_OK; ++c) {
                    if (!stencil_feasible(ctx, c, prec, ncomp)) continue;
                    ctx->stencil_force = c;
                    // one warm-up apply, then the mean of 4: overlapped applies vary
                    // by ~10 % from run to run
                    if (status == QM_OK) status = full();
                    float ms = 0.0f;
                    if (status == QM_OK && hipEventRecord(ev[0], st) != hipSuccess) status = QM_ERR_HIP;
                    for (int rep = 0; rep < 4 && status == QM_OK; ++rep) status = full();
                    if (status == QM_OK && hipEventRecord(ev[1], st) != hipSuccess) status = QM_ERR_HIP;
                    if (status == QM_OK && hipEventSynchronize(ev[1]) != hipSuccess) status = QM_ERR_HIP;
                    if (status == QM_OK && hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) status = QM_ERR_HIP;
                    if (status == QM_OK && ms < best_ms) {
                        best_ms = ms;
                        cand = c;
                    }
                }
                (void)hipEventDestroy(ev[0]);
                (void)hipEventDestroy(ev[1]);
                if (status == QM_ERR_HIP) set_error("H.psi shape tuning: HIP event failure");
                if (status != QM_OK) return status;
                ctx->stencil_choice[key] = cand;
            }
        }
        ctx->stencil_force = cand;
        return full();
    }
    if (part & 1) QM_TRY(launch_local<T>(ctx, hpsi, psi, nb, ncomp, kp.k, st));
    if constexpr (sizeof(T) == 8) {
        if (part & 2) QM_TRY(launch_nonlocal_f64(ctx, hpsi, psi, nb, ncomp, kp, st));
    } else {
        if (part & 2) QM_TRY(launch_nonlocal_f32(ctx, hpsi, psi, nb, ncomp, kp, st));
    }
    return QM_OK;
}

template <typename T>
int hpsi_part(qm_ctx *ctx, T *out, const T *in, int nb, int ncomp, int ik, hipStream_t st, int part) {
    return hpsi_apply<T>(ctx, out, in, nb, ncomp, ik, (void *)st, part);
}
template int hpsi_part<double>(qm_ctx *, double *, const double *, int, int, int, hipStream_t, int);
template int hpsi_part<float>(qm_ctx *, float *, const float *, int, int, int, hipStream_t, int);

template <typename T>
int band_density_impl(qm_ctx *ctx, double *rho, const T *psi, const double *occ, int nb, int ncomp,
                        void *stream) {
    QM_CHECK(ctx && rho && psi && occ, QM_ERR_ARG, "qm_band_density: NULL argument");
    QM_CHECK(nb >= 0 && (ncomp == 1 || ncomp == 2), QM_ERR_ARG, "qm_band_density: bad nb / ncomp");
    if (nb == 0) return QM_OK;
    QM_TRY(activate(ctx));
    hipStream_t st = (hipStream_t)stream;
    QM_TRY(ctx->fwork.reserve(sizeof(double) * nb));
    double *d_occ = ctx->fwork.as<double>();
    QM_HIP(hipMemcpyAsync(d_occ, occ, sizeof(double) * nb, hipMemcpyHostToDevice, st));
    if constexpr (sizeof(T) == 8) QM_TRY(launch_band_density_f64(ctx, rho, psi, d_occ, nb, ncomp, st));
    else QM_TRY(launch_band_density_f32(ctx, rho, psi, d_occ, nb, ncomp, st));
    // the occupancy staging buffer is reused by the next call: finish first
    QM_HIP(hipStreamSynchronize(st));
    return QM_OK;
}

template <typename T>
int forces(qm_ctx *ctx, const T *psi, const double *occ, int nb, int ncomp, int ik, double *f, double *e,
                  void *stream) {
    QM_CHECK(ctx && psi && occ && f && e, QM_ERR_ARG, "qm_nonlocal_forces: NULL argument");
    QM_CHECK(nb >= 1 && nb <= 65535, QM_ERR_ARG, "qm_nonlocal_forces: nb must be in [1, 65535]");
    QM_CHECK(ik >= 0 && ik < (int)ctx->kpts.size(), QM_ERR_ARG, "qm_nonlocal_forces: k index out of range");
    QM_TRY(activate(ctx));
    if (ctx->proj.natoms == 0) {
        *e = 0.0;
        return QM_OK;
    }
    if constexpr (sizeof(T) == 8)
        return nonlocal_forces_f64(ctx, psi, occ, nb, ncomp, ctx->kpts[ik], f, e, (hipStream_t)stream);
    else
        return nonlocal_forces_f32(ctx, psi, occ, nb, ncomp, ctx->kpts[ik], f, e, (hipStream_t)stream);
}

template <typename T>
int gram_api(qm_ctx *ctx, const T *A, int m, const T *B, int n, int64_t k, double *C, void *stream) {
    QM_CHECK(ctx && A && B && C, QM_ERR_ARG, "qm_gram: NULL argument");
    QM_CHECK(m >= 1 && n >= 1 && k >= 1, QM_ERR_ARG, "qm_gram: sizes must be positive");
    QM_TRY(activate(ctx));
    ColBlocks a, b;
    a.add(A, m);
    b.add(B, n);
    hipStream_t st = (hipStream_t)stream;
    QM_TRY(gram<T>(ctx->gemm_part, C, a, b, m, n, k, st));
    // the partial buffer is reused by the next call: finish before returning
    QM_HIP(hipStreamSynchronize(st));
    return QM_OK;
}

template <typename T>
int update_api(qm_ctx *ctx, const T *Z, int m, const T *C, int n, int64_t k, T *Y, void *stream) {
    QM_CHECK(ctx && Z && C && Y, QM_ERR_ARG, "qm_update: NULL argument");
    QM_CHECK(m >= 1 && n >= 1 && k >= 1, QM_ERR_ARG, "qm_update: sizes must be positive");
    QM_TRY(activate(ctx));
    ColBlocks z, y;
    z.add(Z, m);
    y.add(Y, n);
    return update<T>(z, y, C, m, n, k, (hipStream_t)stream);
}

// ------------------------------------------------- work memory of the no-allocation apply
// qm_hpsi_apply_f32_noalloc keeps the nonlocal coefficients and per-chunk
// partial sums in this file-scope device array: one copy per GPU, held until
// the process ends, shared by every context and stream on that GPU.  In host
// code `scratch` names a host-side stand-in, not a device address, so host
// code never uses it: qm_create resolves the address of each device's copy
// with hipGetSymbolAddress once, with that device current, and keeps it in
// g_scratch_ptr[device]; kernels receive that pointer.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kScratchDevices] = {};

// Resolves the current device's copy of `scratch` (once per device).
static int resolve_scratch(int device) {
    QM_CHECK(device >= 0 && device < kScratchDevices, QM_ERR_ARG, "qm_create: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return QM_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    QM_CHECK(e == hipSuccess && ptr != nullptr, QM_ERR_HIP,
             std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return QM_OK;
}

// H.psi in FP32 with every work buffer in `scratch`: no allocation, no stream
// or event creation, no synchronisation, no shape timing.  The projection,
// its coefficient sums, the stencil and the scatter run in that order on the
// caller's stream; the stencil uses the shape already tuned for this call
// shape (full or local applies), else candidate 0, which computes the same
// values.
static int hpsi_apply_noalloc(qm_ctx *ctx, float *hpsi, const float *psi, int nb, int ncomp, int ik, void *stream) {
    QM_CHECK(ctx && hpsi && psi, QM_ERR_ARG, "qm_hpsi_apply_f32_noalloc: NULL argument");
    QM_CHECK(nb >= 0 && (ncomp == 1 || ncomp == 2), QM_ERR_ARG, "qm_hpsi_apply_f32_noalloc: bad nb / ncomp");
    QM_CHECK(ik >= 0 && ik < (int)ctx->kpts.size(), QM_ERR_ARG, "qm_hpsi_apply_f32_noalloc: k index out of range");
    QM_CHECK(ctx->has_vloc, QM_ERR_STATE, "qm_hpsi_apply_f32_noalloc: call qm_set_vloc first");
    QM_CHECK((const void *)hpsi != (const void *)psi, QM_ERR_ARG, "qm_hpsi_apply_f32_noalloc: output aliases input");
    QM_CHECK(ctx->kinetic == QM_KINETIC_FD8, QM_ERR_STATE,
             "qm_hpsi_apply_f32_noalloc: the spectral kinetic term needs work memory; select QM_KINETIC_FD8");
    if (nb == 0) return QM_OK;
    QM_TRY(activate(ctx));
    void *base = g_scratch_ptr[ctx->device];
    QM_CHECK(base != nullptr, QM_ERR_STATE, "qm_hpsi_apply_f32_noalloc: work memory not resolved for this device");
    hipStream_t st = (hipStream_t)stream;
    bool capturing = false;
    QM_TRY(caller_capturing(st, &capturing));
    if (capturing) ctx->capture_pins = 1;  // the graph reads the tables, phases and V_loc
    const KPoint &kp = ctx->kpts[ik];
    const bool nl = ctx->proj.natoms > 0 && ctx->proj.ptot > 0;
    size_t cb = 0, pb = 0;
    if (nl) nonlocal_work_bytes(ctx, nb, ncomp, sizeof(float), &cb, &pb);
    const size_t cb_al = (cb + 255) & ~(size_t)255;
    QM_CHECK(cb_al + pb <= kScratchBytes, QM_ERR_ARG,
             "qm_hpsi_apply_f32_noalloc: " + std::to_string(cb_al + pb) +
                 " bytes of nonlocal work memory needed, the limit is 67108864 (64 MiB): apply fewer bands per call");
    float *coef = static_cast<float *>(base);
    double *part = reinterpret_cast<double *>(static_cast<unsigned char *>(base) + cb_al);
    const int kt = (std::fabs(kp.k[0]) >= 1.0e-8 || std::fabs(kp.k[1]) >= 1.0e-8 ||
                    std::fabs(kp.k[2]) >= 1.0e-8) ? 1 : 0;
    int cand = ctx->stencil_force;
    if (cand < 0) cand = stencil_choice(ctx, 32, ncomp, kt | 2, nb);
    if (cand < 0) cand = stencil_choice(ctx, 32, ncomp, kt, nb);
    if (cand < 0) cand = 0;
    struct Restore {  // the caller's stencil setting and shape timing, on every return
        qm_ctx *c;
        int v;
        ~Restore() {
            c->stencil_force = v;
            c->nl_notime = false;
        }
    } restore{ctx, ctx->stencil_force};
    ctx->stencil_force = cand;
    ctx->nl_notime = true;  // an untuned projection shape takes the size rule
    if (nl) QM_TRY(nonlocal_apply_in_f32(ctx, nullptr, psi, nb, ncomp, kp, 1, coef, part, st));
    QM_TRY(launch_fd8_local_f32(ctx, hpsi, psi, nb, ncomp, kp.k, st));
    if (nl) QM_TRY(nonlocal_apply_in_f32(ctx, hpsi, nullptr, nb, ncomp, kp, 2, coef, nullptr, st));
    return QM_OK;
}

}  // namespace qm

using namespace qm;

extern "C" {

const char *qm_last_error(void) { return g_last_error.c_str(); }

int qm_create(qm_ctx **out, int nx, int ny, int nz, double hx, double hy, double hz, int device) {
    QM_CHECK(out != nullptr, QM_ERR_ARG, "qm_create: out is NULL");
    *out = nullptr;
    QM_CHECK(nx >= 4 && ny >= 4 && nz >= 4, QM_ERR_ARG, "qm_create: grid dimensions must be >= 4");
    QM_CHECK((int64_t)nx * ny * nz < ((int64_t)1 << 31), QM_ERR_ARG, "qm_create: grid must have < 2^31 points");
    QM_CHECK(hx > 0 && hy > 0 && hz > 0, QM_ERR_ARG, "qm_create: spacings must be positive");
    int ndev = 0;
    QM_HIP(hipGetDeviceCount(&ndev));
    QM_CHECK(device >= 0 && device < ndev, QM_ERR_ARG, "qm_create: no such device");
    QM_HIP(hipSetDevice(device));
    QM_TRY(resolve_scratch(device));
    qm_ctx *c = new qm_ctx();
    c->device = device;
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    c->ng = (int64_t)nx * ny * nz;
    c->hx = hx;
    c->hy = hy;
    c->hz = hz;
    c->kpts.resize(1);
    *out = c;
    return QM_OK;
}

void qm_destroy(qm_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    dfree(ctx->d_vloc64);
    dfree(ctx->d_vloc32);
    free_projectors(ctx->proj);
    free_kpoints(ctx);
    ctx->coef.release();
    ctx->fwork.release();
    ctx->gemm_part.release();
    ctx->nl_part.release();
    ctx->eig_work.release();
    ctx->td_work.release();
    ctx->scf_part.release();
    fft_destroy(ctx);
    if (ctx->blas) rocblas_destroy_handle((rocblas_handle)ctx->blas);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    delete ctx;
}

int qm_set_vloc(qm_ctx *ctx, const double *vloc) {
    QM_CHECK(ctx && vloc, QM_ERR_ARG, "qm_set_vloc: NULL argument");
    QM_TRY(activate(ctx));
    std::vector<float> v32(vloc, vloc + ctx->ng);
    if (ctx->d_vloc64 && ctx->d_vloc32) {
        // same size every time: overwrite in place (a captured apply keeps
        // reading the same buffers, and sees the new values); the copy is
        // ordered after the work queued on the null stream, and waits for it
        QM_HIP(hipDeviceSynchronize());
        QM_HIP(hipMemcpy(ctx->d_vloc64, vloc, sizeof(double) * ctx->ng, hipMemcpyHostToDevice));
        QM_HIP(hipMemcpy(ctx->d_vloc32, v32.data(), sizeof(float) * ctx->ng, hipMemcpyHostToDevice));
    } else {
        QM_TRY(upload(&ctx->d_vloc64, vloc, ctx->ng));
        QM_TRY(upload(&ctx->d_vloc32, v32.data(), ctx->ng));
    }
    ctx->has_vloc = true;
    return QM_OK;
}

int qm_set_projectors(qm_ctx *ctx, int natoms, const int *nproj, const int64_t *vox_off, const int *vox_xyz,
                      const double *table, const double *dtable, const int64_t *tab_off, const double *energy,
                      const int *proj_off) {
    QM_CHECK(ctx, QM_ERR_ARG, "qm_set_projectors: NULL context");
    QM_CHECK(natoms >= 0, QM_ERR_ARG, "qm_set_projectors: natoms < 0");
    QM_TRY(check_unpinned(ctx, "qm_set_projectors"));
    QM_TRY(activate(ctx));
    // Every argument is validated and the whole new set is built on the side;
    // the context's set and phases change only when everything succeeded.
    Projectors np;
    std::vector<KPoint> nk(ctx->kpts.size());
    struct Cleanup {  // frees the new set unless it was handed over
        Projectors *p;
        std::vector<KPoint> *k;
        ~Cleanup() {
            if (p) free_projectors(*p);
            if (k)
                for (auto &kp : *k) {
                    dfree(kp.d_phase64);
                    dfree(kp.d_phase32);
                }
        }
    } cleanup{&np, &nk};
    for (size_t i = 0; i < nk.size(); ++i) {
        for (int d = 0; d < 3; ++d) nk[i].k[d] = ctx->kpts[i].k[d];
        nk[i].gamma = ctx->kpts[i].gamma;
    }
    if (natoms > 0) {
        QM_CHECK(nproj && vox_off && vox_xyz && table && tab_off && energy && proj_off, QM_ERR_ARG,
                 "qm_set_projectors: NULL array");
        QM_CHECK(vox_off[0] == 0 && tab_off[0] == 0 && proj_off[0] == 0, QM_ERR_ARG,
                 "qm_set_projectors: offsets must start at 0");
        for (int a = 0; a < natoms; ++a) {
            const int64_t V = vox_off[a + 1] - vox_off[a];
            QM_CHECK(nproj[a] >= 0 && nproj[a] <= kMaxProj, QM_ERR_ARG, "qm_set_projectors: nproj must be in [0, 32]");
            QM_CHECK(V >= 0 && V < ((int64_t)1 << 31), QM_ERR_ARG, "qm_set_projectors: bad voxel range");
            QM_CHECK(tab_off[a + 1] - tab_off[a] == V * nproj[a], QM_ERR_ARG,
                     "qm_set_projectors: tab_off inconsistent with nproj * voxels");
            QM_CHECK(proj_off[a + 1] - proj_off[a] == nproj[a], QM_ERR_ARG,
                     "qm_set_projectors: proj_off inconsistent with nproj");
        }
        QM_CHECK(vox_off[natoms] < ((int64_t)1 << 31), QM_ERR_ARG, "qm_set_projectors: too many voxels");
        QM_TRY(build_projectors(ctx, np, natoms, nproj, vox_off, vox_xyz, table, dtable, tab_off, energy, proj_off));
    }
    for (auto &k : nk) QM_TRY(build_phase(ctx, np, k));
    // applies queued before this call may still read the old tables
    QM_HIP(hipDeviceSynchronize());
    free_projectors(ctx->proj);
    free_kpoints(ctx);
    ctx->proj = np;
    ctx->kpts = nk;
    cleanup.p = nullptr;
    cleanup.k = nullptr;
    return QM_OK;
}

int qm_set_kpoints(qm_ctx *ctx, int nk, const double *kvec) {
    QM_CHECK(ctx && kvec && nk >= 1, QM_ERR_ARG, "qm_set_kpoints: bad argument");
    QM_TRY(check_unpinned(ctx, "qm_set_kpoints"));
    QM_TRY(activate(ctx));
    std::vector<KPoint> kp(nk);
    struct Cleanup {  // frees the new phases unless they were handed over
        std::vector<KPoint> *k;
        ~Cleanup() {
            if (k)
                for (auto &x : *k) {
                    dfree(x.d_phase64);
                    dfree(x.d_phase32);
                }
        }
    } cleanup{&kp};
    for (int i = 0; i < nk; ++i) {
        KPoint &k = kp[i];
        for (int d = 0; d < 3; ++d) k.k[d] = kvec[3 * i + d];
        // the reference treats a k-point as Gamma only when flagged so
        // (m_is_gamma_point_list); here: exactly zero k
        k.gamma = (k.k[0] == 0.0 && k.k[1] == 0.0 && k.k[2] == 0.0);
        QM_TRY(build_phase(ctx, ctx->proj, k));
    }
    // applies queued before this call may still read the old phases
    QM_HIP(hipDeviceSynchronize());
    free_kpoints(ctx);
    ctx->kpts = kp;
    cleanup.k = nullptr;
    return QM_OK;
}

int qm_hpsi_apply_f64(qm_ctx *ctx, double *d_hpsi, const double *d_psi, int nb, int ncomp, int ik, void *stream) {
    return hpsi_apply<double>(ctx, d_hpsi, d_psi, nb, ncomp, ik, stream);
}

int qm_hpsi_apply_f32(qm_ctx *ctx, float *d_hpsi, const float *d_psi, int nb, int ncomp, int ik, void *stream) {
    return hpsi_apply<float>(ctx, d_hpsi, d_psi, nb, ncomp, ik, stream);
}

int qm_hpsi_apply_f32_noalloc(qm_ctx *ctx, float *d_hpsi, const float *d_psi, int nb, int ncomp, int ik,
                              void *stream) {
    return hpsi_apply_noalloc(ctx, d_hpsi, d_psi, nb, ncomp, ik, stream);
}

int qm_gram_f64(qm_ctx *ctx, const double *d_A, int m, const double *d_B, int n, int64_t k, double *d_C,
                void *stream) {
    return gram_api<double>(ctx, d_A, m, d_B, n, k, d_C, stream);
}

int qm_gram_f32(qm_ctx *ctx, const float *d_A, int m, const float *d_B, int n, int64_t k, double *d_C,
                void *stream) {
    return gram_api<float>(ctx, d_A, m, d_B, n, k, d_C, stream);
}

int qm_update_f64(qm_ctx *ctx, const double *d_Z, int m, const double *d_C, int n, int64_t k, double *d_Y,
                  void *stream) {
    return update_api<double>(ctx, d_Z, m, d_C, n, k, d_Y, stream);
}

int qm_update_f32(qm_ctx *ctx, const float *d_Z, int m, const float *d_C, int n, int64_t k, float *d_Y,
                  void *stream) {
    return update_api<float>(ctx, d_Z, m, d_C, n, k, d_Y, stream);
}

int qm_local_apply_f64(qm_ctx *ctx, double *d_out, const double *d_psi, int nb, int ncomp, int ik, void *stream) {
    return hpsi_apply<double>(ctx, d_out, d_psi, nb, ncomp, ik, stream, 1);
}

int qm_local_apply_f32(qm_ctx *ctx, float *d_out, const float *d_psi, int nb, int ncomp, int ik, void *stream) {
    return hpsi_apply<float>(ctx, d_out, d_psi, nb, ncomp, ik, stream, 1);
}

int qm_nonlocal_add_f64(qm_ctx *ctx, double *d_out, const double *d_psi, int nb, int ncomp, int ik, void *stream) {
    return hpsi_apply<double>(ctx, d_out, d_psi, nb, ncomp, ik, stream, 2);
}

int qm_nonlocal_add_f32(qm_ctx *ctx, float *d_out, const float *d_psi, int nb, int ncomp, int ik, void *stream) {
    return hpsi_apply<float>(ctx, d_out, d_psi, nb, ncomp, ik, stream, 2);
}

int qm_band_density_f64(qm_ctx *ctx, double *d_rho, const double *d_psi, const double *occ, int nb, int ncomp,
                        void *stream) {
    return band_density_impl<double>(ctx, d_rho, d_psi, occ, nb, ncomp, stream);
}

int qm_band_density_f32(qm_ctx *ctx, double *d_rho, const float *d_psi, const double *occ, int nb, int ncomp,
                        void *stream) {
    return band_density_impl<float>(ctx, d_rho, d_psi, occ, nb, ncomp, stream);
}

int qm_nonlocal_forces_f64(qm_ctx *ctx, const double *d_psi, const double *occ, int nb, int ncomp, int ik,
                           double *forces_out, double *e_nl, void *stream) {
    return forces<double>(ctx, d_psi, occ, nb, ncomp, ik, forces_out, e_nl, stream);
}

int qm_nonlocal_forces_f32(qm_ctx *ctx, const float *d_psi, const double *occ, int nb, int ncomp, int ik,
                           double *forces_out, double *e_nl, void *stream) {
    return forces<float>(ctx, d_psi, occ, nb, ncomp, ik, forces_out, e_nl, stream);
}

int qm_nonlocal_force(qm_ctx *ctx, int tiles) {
    QM_CHECK(ctx, QM_ERR_ARG, "qm_nonlocal_force: NULL context");
    QM_CHECK(tiles == 0 || tiles == 1 || tiles == 2 || tiles == 4, QM_ERR_ARG,
             "qm_nonlocal_force: tiles must be 0 (autotune), 1, 2 or 4");
    ctx->nl_force = tiles;
    return QM_OK;
}

int qm_nonlocal_choice(qm_ctx *ctx, int precision, int ncomp, int kt, int nb, int *tiles) {
    QM_CHECK(ctx && tiles, QM_ERR_ARG, "qm_nonlocal_choice: NULL argument");
    QM_CHECK(precision == 32 || precision == 64, QM_ERR_ARG, "qm_nonlocal_choice: precision must be 32 or 64");
    QM_CHECK(ncomp == 1 || ncomp == 2, QM_ERR_ARG, "qm_nonlocal_choice: ncomp must be 1 or 2");
    *tiles = nonlocal_choice(ctx, precision, ncomp, kt, nb);
    return QM_OK;
}

int qm_release_captures(qm_ctx *ctx) {
    QM_CHECK(ctx, QM_ERR_ARG, "qm_release_captures: NULL context");
    ctx->capture_pins = 0;
    return QM_OK;
}

int qm_stencil_num_candidates(void) { return stencil_num_candidates(); }

int qm_stencil_candidate_name(int candidate, char *buf, int buflen) {
    QM_CHECK(candidate >= 0 && candidate < stencil_num_candidates(), QM_ERR_ARG,
             "qm_stencil_candidate_name: candidate out of range");
    QM_CHECK(buf && buflen > 0, QM_ERR_ARG, "qm_stencil_candidate_name: NULL or empty buffer");
    const std::string name = stencil_candidate_name(candidate);
    const size_t n = std::min(name.size(), (size_t)buflen - 1);
    name.copy(buf, n);
    buf[n] = '\0';
    return QM_OK;
}

int qm_stencil_force(qm_ctx *ctx, int candidate) {
    QM_CHECK(ctx, QM_ERR_ARG, "qm_stencil_force: NULL context");
    QM_CHECK(candidate >= -1 && candidate < stencil_num_candidates(), QM_ERR_ARG,
             "qm_stencil_force: candidate must be -1 (autotune) or a candidate index");
    ctx->stencil_force = candidate;
    return QM_OK;
}

int qm_stencil_choice(qm_ctx *ctx, int precision, int ncomp, int kt, int nb, int *candidate) {
    QM_CHECK(ctx && candidate, QM_ERR_ARG, "qm_stencil_choice: NULL argument");
    QM_CHECK(precision == 32 || precision == 64, QM_ERR_ARG, "qm_stencil_choice: precision must be 32 or 64");
    QM_CHECK(ncomp == 1 || ncomp == 2, QM_ERR_ARG, "qm_stencil_choice: ncomp must be 1 or 2");
    *candidate = stencil_choice(ctx, precision, ncomp, kt, nb);
    return QM_OK;
}

}  // extern "C"
