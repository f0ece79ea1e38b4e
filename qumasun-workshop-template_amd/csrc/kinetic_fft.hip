// Spectral (FFT) kinetic term on gfx950: the reference's default kinetic
// operator in mHamiltonianMatrix_ddm_bundle (m_algorithm_Laplacian = 1,
// qumasun_base1.matrix.h:171-228; its hipFFT many-plan, gyield/gy_fft_executor_hip.h:188-248):
//   H psi = IFFT( |G + k|^2 / 2 . FFT psi ) / N + V_loc psi
// with G = 2 pi / L_d times the wrapped index (i > n/2 -> i - n, so the n/2
// Nyquist index is +n/2), OperateKineticInKspace_bundle (qumasun_base1.matrix.h:96-130).
// And the kinetic step of the split-operator propagator,
//   psi <- IFFT( e^{-i dt |G + k|^2 / 2} FFT psi ) / N
// (EvolveKineticInKspace / mEvolveK_FFT, qumasun_symplectic.kinetic_fft.h:14-60, 65-140).
//
// GPU structure: rocFFT batched 3-D transforms.  Complex bands are packed
// from the SoA (re, im) arrays into one interleaved block (the reference's
// GatherExchangeD2Z, without the DDM exchange) and transformed in place; the
// unpack after the inverse transform is fused with V_loc psi.  Real
// Gamma-point bands go through a real-to-Hermitian transform (half the
// spectrum) straight from the band block.  The k-space factor (and 1/N) is one
// elementwise kernel.  Plans are cached per (precision, ncomp, band count).
#include <rocfft/rocfft.h>

#include <cmath>
#include <map>
#include <mutex>

#include "qm_internal.h"

namespace qm {

namespace {

#define QM_FFT(call)                                                                               \
    do {                                                                                           \
        rocfft_status s_ = (call);                                                                 \
        if (s_ != rocfft_status_success) {                                                         \
            ::qm::set_error(std::string(#call) + ": rocFFT status " + std::to_string((int)s_) +    \
                            " at " + __FILE__ + ":" + std::to_string(__LINE__));                   \
            return QM_ERR_HIP;                                                                     \
        }                                                                                          \
    } while (0)

std::once_flag g_rocfft_once;

// One direction of a cached transform pair.
struct FftPlan {
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    size_t work = 0;
};

struct FftPair {
    FftPlan fwd, inv;
};

}  // namespace

// Transforms and buffers of one context (opaque in qm_ctx).
struct FftState {
    std::map<int64_t, FftPair> plans;
    DevBuf spec;  // spectra of the band block (complex interleaved)
    DevBuf work;  // rocFFT work buffer
};

namespace {

template <typename T>
struct C2 {
    T x, y;
};

// spec *= f(|G + k|^2 / 2) for every band's spectrum (nxe x ny x nz complex,
// x fastest; nxe = nx, or nx/2 + 1 for a real band's half spectrum).
//   mode 0: f = e2 * scale (kinetic operator)
//   mode 1: f = e^{-i dt e2} * scale (kinetic propagator)
//   mode 2: f = 4 pi / |G|^2 * scale, 0 at G = 0 (Poisson solve, k = 0)
template <typename T, int MODE>
__global__ __launch_bounds__(256) void kspace_kernel(C2<T> *__restrict__ spec, int64_t per_band, int64_t total,
                                                     int nxe, int nx, int ny, int nz, double gx, double gy, double gz,
                                                     double kx, double ky, double kz, double scale, double dt) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t i = e % per_band;
        const int ix = (int)(i % nxe);
        const int64_t r = i / nxe;
        const int iy = (int)(r % ny), iz = (int)(r / ny);
        const double qx = gx * (double)(ix * 2 > nx ? ix - nx : ix) + kx;
        const double qy = gy * (double)(iy * 2 > ny ? iy - ny : iy) + ky;
        const double qz = gz * (double)(iz * 2 > nz ? iz - nz : iz) + kz;
        const double e2 = 0.5 * (qx * qx + qy * qy + qz * qz);
        C2<T> v = spec[e];
        if constexpr (MODE == 0) {
            const T f = (T)(e2 * scale);
            v.x *= f;
            v.y *= f;
        } else if constexpr (MODE == 2) {
            const T f = (ix == 0 && iy == 0 && iz == 0) ? (T)0 : (T)(2.0 * M_PI / e2 * scale);
            v.x *= f;
            v.y *= f;
        } else {
            double sn, cs;
            sincos(-dt * e2, &sn, &cs);
            const T c = (T)(cs * scale), s = (T)(sn * scale);
            const T re = v.x, im = v.y;
            v.x = c * re - s * im;
            v.y = c * im + s * re;
        }
        spec[e] = v;
    }
}

// spec[n][i] = (psi_re, psi_im) of SoA complex band n
template <typename T>
__global__ __launch_bounds__(256) void pack_kernel(C2<T> *__restrict__ spec, const T *__restrict__ psi, int64_t ng,
                                                   int64_t total) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t n = e / ng, i = e - n * ng;
        spec[e] = C2<T>{psi[2 * n * ng + i], psi[(2 * n + 1) * ng + i]};
    }
}

// SoA complex out from the interleaved block: out = spec (+ V psi when ADDV)
template <typename T, bool ADDV>
__global__ __launch_bounds__(256) void unpack_kernel(T *__restrict__ out, const C2<T> *__restrict__ spec,
                                                     const T *__restrict__ psi, const T *__restrict__ v, int64_t ng,
                                                     int64_t total) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t n = e / ng, i = e - n * ng;
        const int64_t re = 2 * n * ng + i, im = re + ng;
        C2<T> z = spec[e];
        if constexpr (ADDV) {
            const T vi = v[i];
            z.x += vi * psi[re];
            z.y += vi * psi[im];
        }
        out[re] = z.x;
        out[im] = z.y;
    }
}

// out[col][i] += V[i] psi[col][i] over ncol columns of Ng
template <typename T>
__global__ __launch_bounds__(256) void vpsi_add_kernel(T *__restrict__ out, const T *__restrict__ psi,
                                                       const T *__restrict__ v, int64_t ng, int64_t total) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256)
        out[e] += v[e % ng] * psi[e];
}

inline unsigned grid_of(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)std::min<int64_t>(std::max<int64_t>(b, 1), 1 << 20);
}

int make_plan(FftPlan &p, rocfft_result_placement place, rocfft_transform_type type, rocfft_precision prec,
              const size_t len[3], size_t batch, rocfft_array_type in_t, rocfft_array_type out_t,
              const size_t in_str[3], size_t in_dist, const size_t out_str[3], size_t out_dist) {
    rocfft_plan_description d = nullptr;
    QM_FFT(rocfft_plan_description_create(&d));
    rocfft_status st = rocfft_plan_description_set_data_layout(d, in_t, out_t, nullptr, nullptr, 3, in_str, in_dist,
                                                               3, out_str, out_dist);
    if (st == rocfft_status_success) st = rocfft_plan_create(&p.plan, place, type, prec, 3, len, batch, d);
    (void)rocfft_plan_description_destroy(d);
    QM_FFT(st);
    QM_FFT(rocfft_plan_get_work_buffer_size(p.plan, &p.work));
    QM_FFT(rocfft_execution_info_create(&p.info));
    return QM_OK;
}

void destroy_pair(FftPair &q) {
    for (FftPlan *p : {&q.fwd, &q.inv}) {
        if (p->info) (void)rocfft_execution_info_destroy(p->info);
        if (p->plan) (void)rocfft_plan_destroy(p->plan);
        p->info = nullptr;
        p->plan = nullptr;
    }
}

// The forward / inverse pair for nb bands of this shape: complex bands are
// transformed in place in the packed interleaved block; real bands go real ->
// Hermitian half spectrum -> real, out of place.
template <typename T>
int get_pair(qm_ctx *ctx, int nb, int ncomp, FftPair **out) {
    std::call_once(g_rocfft_once, [] { (void)rocfft_setup(); });
    if (!ctx->fft) ctx->fft = new FftState();
    FftState &f = *ctx->fft;
    const int64_t key = ((int64_t)nb * 2 + (ncomp - 1)) * 2 + (sizeof(T) == 8 ? 1 : 0);
    auto it = f.plans.find(key);
    if (it != f.plans.end()) {
        *out = &it->second;
        return QM_OK;
    }
    const rocfft_precision prec = sizeof(T) == 8 ? rocfft_precision_double : rocfft_precision_single;
    const size_t nx = ctx->nx, ny = ctx->ny, nz = ctx->nz, ng = (size_t)ctx->ng;
    const size_t len[3] = {nx, ny, nz};  // fastest first
    const size_t str[3] = {1, nx, nx * ny};
    FftPair q;
    int rc = QM_OK;
    if (ncomp == 2) {
        rc = make_plan(q.fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward, prec, len, nb,
                       rocfft_array_type_complex_interleaved, rocfft_array_type_complex_interleaved, str, ng, str, ng);
        if (rc == QM_OK)
            rc = make_plan(q.inv, rocfft_placement_inplace, rocfft_transform_type_complex_inverse, prec, len, nb,
                           rocfft_array_type_complex_interleaved, rocfft_array_type_complex_interleaved, str, ng, str,
                           ng);
    } else {
        const size_t nxh = nx / 2 + 1;
        const size_t hstr[3] = {1, nxh, nxh * ny};
        rc = make_plan(q.fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward, prec, len, nb,
                       rocfft_array_type_real, rocfft_array_type_hermitian_interleaved, str, ng, hstr, nxh * ny * nz);
        if (rc == QM_OK)
            rc = make_plan(q.inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse, prec, len, nb,
                           rocfft_array_type_hermitian_interleaved, rocfft_array_type_real, hstr, nxh * ny * nz, str,
                           ng);
    }
    if (rc != QM_OK) {
        destroy_pair(q);
        return rc;
    }
    *out = &(f.plans[key] = q);
    return QM_OK;
}

int execute(FftState &f, FftPlan &p, void *in[], void *out[], hipStream_t st) {
    if (p.work > 0) {
        QM_TRY(f.work.reserve(p.work));
        QM_FFT(rocfft_execution_info_set_work_buffer(p.info, f.work.p, p.work));
    }
    QM_FFT(rocfft_execution_info_set_stream(p.info, st));
    QM_FFT(rocfft_execute(p.plan, in, out, p.info));
    return QM_OK;
}

// spec = FFT(psi) (nb bands, ncomp 1 or 2); returns the spectrum and its size per band
template <typename T>
int forward(qm_ctx *ctx, const T *psi, int nb, int ncomp, FftPair **pair, C2<T> **spec, int64_t *per_band,
            hipStream_t st) {
    QM_TRY(get_pair<T>(ctx, nb, ncomp, pair));
    FftState &f = *ctx->fft;
    const int64_t pb = ncomp == 2 ? ctx->ng : (int64_t)(ctx->nx / 2 + 1) * ctx->ny * ctx->nz;
    QM_CHECK(ctx->capture_pins == 0 || f.spec.bytes >= sizeof(C2<T>) * (size_t)pb * nb, QM_ERR_STATE,
             "FFT kinetic: larger spectrum buffer needed while a graph holds an apply of this context; destroy it "
             "and call qm_release_captures first");
    QM_TRY(f.spec.reserve(sizeof(C2<T>) * (size_t)pb * nb));
    C2<T> *s = f.spec.as<C2<T>>();
    if (ncomp == 2) {  // pack, then transform in place
        const int64_t total = (int64_t)nb * ctx->ng;
        hipLaunchKernelGGL(pack_kernel<T>, dim3(grid_of(total)), dim3(256), 0, st, s, psi, ctx->ng, total);
        QM_HIP(hipGetLastError());
        void *io[1] = {s};
        QM_TRY(execute(f, (*pair)->fwd, io, nullptr, st));
    } else {
        void *in[1] = {const_cast<T *>(psi)};
        void *out[1] = {s};
        QM_TRY(execute(f, (*pair)->fwd, in, out, st));
    }
    *spec = s;
    *per_band = pb;
    return QM_OK;
}

// Inverse transform into dst: complex bands in place and unpacked into the SoA
// block dst (plus V_loc psi when psi is given); real bands out of place into
// dst (V_loc psi added by the caller).
template <typename T>
int inverse(qm_ctx *ctx, FftPair *pair, C2<T> *spec, T *dst, int nb, int ncomp, const T *psi, hipStream_t st) {
    if (ncomp == 2) {
        void *io[1] = {spec};
        QM_TRY(execute(*ctx->fft, pair->inv, io, nullptr, st));
        const int64_t total = (int64_t)nb * ctx->ng;
        if (psi) {
            const T *v = sizeof(T) == 8 ? (const T *)ctx->d_vloc64 : (const T *)ctx->d_vloc32;
            hipLaunchKernelGGL((unpack_kernel<T, true>), dim3(grid_of(total)), dim3(256), 0, st, dst, spec, psi, v,
                               ctx->ng, total);
        } else {
            hipLaunchKernelGGL((unpack_kernel<T, false>), dim3(grid_of(total)), dim3(256), 0, st, dst, spec,
                               (const T *)nullptr, (const T *)nullptr, ctx->ng, total);
        }
        QM_HIP(hipGetLastError());
        return QM_OK;
    }
    void *in[1] = {spec};
    void *out[1] = {dst};
    return execute(*ctx->fft, pair->inv, in, out, st);
}

template <typename T, int MODE>
void launch_kspace(qm_ctx *ctx, C2<T> *spec, int64_t per_band, int nb, int ncomp, const double k[3], double dt,
                   hipStream_t st) {
    const int nxe = ncomp == 2 ? ctx->nx : ctx->nx / 2 + 1;
    const double gx = 2.0 * M_PI / (ctx->hx * ctx->nx), gy = 2.0 * M_PI / (ctx->hy * ctx->ny),
                 gz = 2.0 * M_PI / (ctx->hz * ctx->nz);
    const int64_t total = per_band * nb;
    hipLaunchKernelGGL((kspace_kernel<T, MODE>), dim3(grid_of(total)), dim3(256), 0, st, spec, per_band, total, nxe,
                       ctx->nx, ctx->ny, ctx->nz, gx, gy, gz, k[0], k[1], k[2], 1.0 / (double)ctx->ng, dt);
}

}  // namespace

void fft_destroy(qm_ctx *ctx) {
    if (!ctx->fft) return;
    for (auto &kv : ctx->fft->plans) destroy_pair(kv.second);
    ctx->fft->spec.release();
    ctx->fft->work.release();
    delete ctx->fft;
    ctx->fft = nullptr;
}

// out = IFFT(|G + k|^2 / 2 FFT psi) / N + V_loc psi (overwrites out)
template <typename T>
int launch_fft_local(qm_ctx *ctx, T *out, const T *psi, int nb, int ncomp, const double k[3], hipStream_t st) {
    const bool kt = k[0] != 0.0 || k[1] != 0.0 || k[2] != 0.0;
    QM_CHECK(ncomp == 2 || !kt, QM_ERR_ARG, "real (ncomp=1) bands need a Gamma k-point");
    FftPair *pair = nullptr;
    C2<T> *spec = nullptr;
    int64_t pb = 0;
    QM_TRY(forward<T>(ctx, psi, nb, ncomp, &pair, &spec, &pb, st));
    launch_kspace<T, 0>(ctx, spec, pb, nb, ncomp, k, 0.0, st);
    QM_HIP(hipGetLastError());
    QM_TRY(inverse<T>(ctx, pair, spec, out, nb, ncomp, psi, st));
    if (ncomp == 1) {
        const T *v = sizeof(T) == 8 ? (const T *)ctx->d_vloc64 : (const T *)ctx->d_vloc32;
        const int64_t total = (int64_t)nb * ctx->ng;
        hipLaunchKernelGGL(vpsi_add_kernel<T>, dim3(grid_of(total)), dim3(256), 0, st, out, psi, v, ctx->ng, total);
        QM_HIP(hipGetLastError());
    }
    return QM_OK;
}
template int launch_fft_local<double>(qm_ctx *, double *, const double *, int, int, const double *, hipStream_t);
template int launch_fft_local<float>(qm_ctx *, float *, const float *, int, int, const double *, hipStream_t);

// psi <- IFFT(e^{-i dt |G + k|^2 / 2} FFT psi) / N, SoA complex bands
template <typename T>
int fft_kinetic_step(qm_ctx *ctx, T *psi, int nb, const double k[3], double dt, hipStream_t st) {
    FftPair *pair = nullptr;
    C2<T> *spec = nullptr;
    int64_t pb = 0;
    QM_TRY(forward<T>(ctx, psi, nb, 2, &pair, &spec, &pb, st));
    launch_kspace<T, 1>(ctx, spec, pb, nb, 2, k, dt, st);
    QM_HIP(hipGetLastError());
    return inverse<T>(ctx, pair, spec, psi, nb, 2, nullptr, st);
}

// V_H = IFFT(4 pi / |G|^2 FFT rho) / N with the G = 0 term dropped: the FFT
// Poisson solve of SetPotentialByPoissonFFT (poisson_fft.h:16-120) and
// mSetPotentialVhart (qumasun_base1.potential.h:32-116), without the
// refined-grid up/down conversion.  rho and V_H are FP64 grid fields.
int fft_poisson(qm_ctx *ctx, const double *rho, double *vh, hipStream_t st) {
    FftPair *pair = nullptr;
    C2<double> *spec = nullptr;
    int64_t pb = 0;
    QM_TRY(forward<double>(ctx, rho, 1, 1, &pair, &spec, &pb, st));
    const double k0[3] = {0.0, 0.0, 0.0};
    launch_kspace<double, 2>(ctx, spec, pb, 1, 1, k0, 0.0, st);
    QM_HIP(hipGetLastError());
    return inverse<double>(ctx, pair, spec, vh, 1, 1, nullptr, st);
}

}  // namespace qm

using namespace qm;

extern "C" {

int qm_hartree_f64(qm_ctx *ctx, const double *d_rho, double *d_vh, void *stream) {
    QM_CHECK(ctx && d_rho && d_vh, QM_ERR_ARG, "qm_hartree_f64: NULL argument");
    QM_TRY(activate(ctx));
    return fft_poisson(ctx, d_rho, d_vh, (hipStream_t)stream);
}

int qm_set_kinetic(qm_ctx *ctx, int mode) {
    QM_CHECK(ctx, QM_ERR_ARG, "qm_set_kinetic: NULL context");
    QM_CHECK(mode == QM_KINETIC_FD8 || mode == QM_KINETIC_FFT, QM_ERR_ARG,
             "qm_set_kinetic: mode must be QM_KINETIC_FD8 or QM_KINETIC_FFT");
    ctx->kinetic = mode;
    return QM_OK;
}

int qm_td_kinetic_fft_f64(qm_ctx *ctx, double *d_psi, int nb, int ik, double dt, void *stream) {
    QM_CHECK(ctx && d_psi, QM_ERR_ARG, "qm_td_kinetic_fft: NULL argument");
    QM_CHECK(nb >= 0 && ik >= 0 && ik < (int)ctx->kpts.size(), QM_ERR_ARG, "qm_td_kinetic_fft: bad nb or k index");
    QM_CHECK(std::isfinite(dt), QM_ERR_ARG, "qm_td_kinetic_fft: dt is not finite");
    if (nb == 0) return QM_OK;
    QM_TRY(activate(ctx));
    return fft_kinetic_step<double>(ctx, d_psi, nb, ctx->kpts[ik].k, dt, (hipStream_t)stream);
}

int qm_td_kinetic_fft_f32(qm_ctx *ctx, float *d_psi, int nb, int ik, double dt, void *stream) {
    QM_CHECK(ctx && d_psi, QM_ERR_ARG, "qm_td_kinetic_fft: NULL argument");
    QM_CHECK(nb >= 0 && ik >= 0 && ik < (int)ctx->kpts.size(), QM_ERR_ARG, "qm_td_kinetic_fft: bad nb or k index");
    QM_CHECK(std::isfinite(dt), QM_ERR_ARG, "qm_td_kinetic_fft: dt is not finite");
    if (nb == 0) return QM_OK;
    QM_TRY(activate(ctx));
    return fft_kinetic_step<float>(ctx, d_psi, nb, ctx->kpts[ik].k, dt, (hipStream_t)stream);
}

}  // extern "C"
