// Real-time propagators of SoA complex band blocks on gfx950 (TDDFT).
//
//   qm_td_taylor_*: psi <- sum_{n=0..order} (-i dt Op)^n / n! psi
//     Op = the full H(k) (QUMASUN::TimeEvoH4_bundle, qumasun_td.time_evo_H4.h:70-160,
//     called by QUMASUN_TD::mTimeEvolutionH4 with order 4, 166-204), its local part
//     (kinetic + V_loc) or its nonlocal part alone (QUMASUN::TimeEvoA4_bundle,
//     qumasun_symplectic.time_evo_K4.h:13-96, called by TestEvolveNL4, 305-335).
//   qm_td_expv_*:   psi <- e^{-i V_loc dt} psi pointwise (mProductExpToV with
//     exp_v = cos(-dt V) + i sin(-dt V), qumasun_symplectic.evolution.h:80-100, 140-165).
//
// As in the reference, H^n psi is built one apply at a time from H^(n-1) psi,
// and each term is added to psi as soon as it exists, so two work blocks
// suffice.  The reference runs bundles of 16 bands through per-band CPU loops;
// here the whole block goes through each H.psi apply (one launch sequence per
// term) and each term is added by one fused elementwise kernel over the block.
#include <cmath>

#include "qm_internal.h"

namespace qm {
namespace {

// x += (alpha - i beta) h for SoA complex bands, per component:
//   x_re += alpha h_re + beta h_im,  x_im += alpha h_im - beta h_re.
// One thread per (band, grid point).
template <typename T>
__global__ __launch_bounds__(256) void td_axpy_kernel(T *__restrict__ x, const T *__restrict__ h, int64_t ng,
                                                      int64_t total, T alpha, T beta) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t n = e / ng, i = e - n * ng;
        const int64_t re = 2 * n * ng + i, im = re + ng;
        const T hr = h[re], hi = h[im];
        x[re] += alpha * hr + beta * hi;
        x[im] += alpha * hi - beta * hr;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void td_zero_kernel(T *__restrict__ p, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) p[e] = T(0);
}

// psi <- e^{-i V dt} psi: (cos(-dt V) + i sin(-dt V)) (re + i im), the phase in FP64
template <typename T>
__global__ __launch_bounds__(256) void td_expv_kernel(T *__restrict__ x, const double *__restrict__ v, int64_t ng,
                                                      int64_t total, double dt) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t n = e / ng, i = e - n * ng;
        const int64_t re = 2 * n * ng + i, im = re + ng;
        double sn, cs;
        sincos(-dt * v[i], &sn, &cs);
        const T c = (T)cs, s = (T)sn;
        const T pr = x[re], pi = x[im];
        x[re] = c * pr - s * pi;
        x[im] = c * pi + s * pr;
    }
}

inline unsigned grid_of(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)std::min<int64_t>(std::max<int64_t>(b, 1), 1 << 20);
}

template <typename T>
int td_apply(qm_ctx *ctx, T *out, const T *in, int nb, int ik, int op, hipStream_t st) {
    if (op == QM_TD_OP_NONLOCAL) {
        // Op psi = V_nl psi alone: clear, then the nonlocal term's "+="
        const int64_t n = 2 * (int64_t)nb * ctx->ng;
        hipLaunchKernelGGL(td_zero_kernel<T>, dim3(grid_of(n)), dim3(256), 0, st, out, n);
        QM_HIP(hipGetLastError());
    }
    const int part = op == QM_TD_OP_FULL ? 3 : (op == QM_TD_OP_LOCAL ? 1 : 2);
    return hpsi_part<T>(ctx, out, in, nb, 2, ik, st, part);
}

template <typename T>
int td_taylor(qm_ctx *ctx, T *psi, int nb, int ik, double dt, int order, int op, void *stream) {
    QM_CHECK(ctx && psi, QM_ERR_ARG, "qm_td_taylor: NULL argument");
    QM_CHECK(nb >= 0, QM_ERR_ARG, "qm_td_taylor: nb < 0");
    QM_CHECK(order >= 1 && order <= 32, QM_ERR_ARG, "qm_td_taylor: order must be in [1, 32]");
    QM_CHECK(op == QM_TD_OP_FULL || op == QM_TD_OP_LOCAL || op == QM_TD_OP_NONLOCAL, QM_ERR_ARG,
             "qm_td_taylor: op must be QM_TD_OP_FULL, QM_TD_OP_LOCAL or QM_TD_OP_NONLOCAL");
    QM_CHECK(std::isfinite(dt), QM_ERR_ARG, "qm_td_taylor: dt is not finite");
    QM_CHECK(ik >= 0 && ik < (int)ctx->kpts.size(), QM_ERR_ARG, "qm_td_taylor: k index out of range");
    if (nb == 0) return QM_OK;
    QM_TRY(activate(ctx));
    hipStream_t st = (hipStream_t)stream;
    const size_t blk = sizeof(T) * 2 * (size_t)nb * ctx->ng;
    QM_CHECK(ctx->capture_pins == 0 || ctx->td_work.bytes >= 2 * blk, QM_ERR_STATE,
             "qm_td_taylor: larger work blocks needed while a graph holds an apply of this context; destroy it "
             "and call qm_release_captures first");
    QM_TRY(ctx->td_work.reserve(2 * blk));
    T *work[2] = {ctx->td_work.as<T>(), reinterpret_cast<T *>(ctx->td_work.as<unsigned char>() + blk)};
    const int64_t total = (int64_t)nb * ctx->ng;
    const T *src = psi;
    double coef = 1.0;
    for (int k = 1; k <= order; ++k) {
        T *dst = work[(k - 1) & 1];
        QM_TRY(td_apply<T>(ctx, dst, src, nb, ik, op, st));
        coef *= dt / k;
        // (-i)^k coef: k = 1 mod 4 -> -i, 2 -> -1, 3 -> +i, 0 -> +1
        double alpha = 0.0, beta = 0.0;
        switch (k & 3) {
            case 1: beta = coef; break;
            case 2: alpha = -coef; break;
            case 3: beta = -coef; break;
            default: alpha = coef; break;
        }
        hipLaunchKernelGGL(td_axpy_kernel<T>, dim3(grid_of(total)), dim3(256), 0, st, psi, (const T *)dst, ctx->ng,
                           total, (T)alpha, (T)beta);
        QM_HIP(hipGetLastError());
        src = dst;
    }
    return QM_OK;
}

template <typename T>
int td_expv(qm_ctx *ctx, T *psi, int nb, double dt, void *stream) {
    QM_CHECK(ctx && psi, QM_ERR_ARG, "qm_td_expv: NULL argument");
    QM_CHECK(nb >= 0, QM_ERR_ARG, "qm_td_expv: nb < 0");
    QM_CHECK(ctx->has_vloc, QM_ERR_STATE, "qm_td_expv: call qm_set_vloc first");
    QM_CHECK(std::isfinite(dt), QM_ERR_ARG, "qm_td_expv: dt is not finite");
    if (nb == 0) return QM_OK;
    QM_TRY(activate(ctx));
    const int64_t total = (int64_t)nb * ctx->ng;
    hipLaunchKernelGGL(td_expv_kernel<T>, dim3(grid_of(total)), dim3(256), 0, (hipStream_t)stream, psi,
                       (const double *)ctx->d_vloc64, ctx->ng, total, dt);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

}  // namespace
}  // namespace qm

using namespace qm;

extern "C" {

int qm_td_taylor_f64(qm_ctx *ctx, double *d_psi, int nb, int ik, double dt, int order, int op, void *stream) {
    return td_taylor<double>(ctx, d_psi, nb, ik, dt, order, op, stream);
}

int qm_td_taylor_f32(qm_ctx *ctx, float *d_psi, int nb, int ik, double dt, int order, int op, void *stream) {
    return td_taylor<float>(ctx, d_psi, nb, ik, dt, order, op, stream);
}

int qm_td_expv_f64(qm_ctx *ctx, double *d_psi, int nb, double dt, void *stream) {
    return td_expv<double>(ctx, d_psi, nb, dt, stream);
}

int qm_td_expv_f32(qm_ctx *ctx, float *d_psi, int nb, double dt, void *stream) {
    return td_expv<float>(ctx, d_psi, nb, dt, stream);
}

}  // extern "C"
