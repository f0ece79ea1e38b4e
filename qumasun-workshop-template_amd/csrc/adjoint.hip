// Adjoint of H.psi with respect to the local potential: the band density.
//
//   rho(r) += sum_n occ_n |psi_n(r)|^2      (dE_band/dV_loc(r) = rho(r) dV)
//
// Mathematics: SetDensityOne (qumasun_SetDensityOne.h:9-41), which sums the
// SoA re and im rows of every band with the band's occupancy.  One thread per
// grid point walks all bands (coalesced along x) and accumulates in FP64.
#include "qm_internal.h"

namespace qm {
namespace {

template <typename T>
__global__ __launch_bounds__(256) void band_density_kernel(double *__restrict__ rho, const T *__restrict__ psi,
                                                            const double *__restrict__ occ, int64_t ng, int nb,
                                                            int ncomp) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= ng) return;
    double s = 0.0;
    for (int n = 0; n < nb; ++n) {
        double t = 0.0;
        for (int c = 0; c < ncomp; ++c) {
            const double v = (double)psi[((int64_t)n * ncomp + c) * ng + i];
            t += v * v;
        }
        s += occ[n] * t;
    }
    rho[i] += s;
}

template <typename T>
int launch_density(const qm_ctx *ctx, double *rho, const T *psi, const double *d_occ, int nb, int ncomp,
                   hipStream_t st) {
    QM_CHECK(ncomp == 1 || ncomp == 2, QM_ERR_ARG, "ncomp must be 1 or 2");
    const int64_t blocks = (ctx->ng + 255) / 256;
    hipLaunchKernelGGL((band_density_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, rho, psi, d_occ, ctx->ng,
                       nb, ncomp);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

}  // namespace

int launch_band_density_f64(const qm_ctx *ctx, double *rho, const double *psi, const double *d_occ, int nb,
                            int ncomp, hipStream_t st) {
    return launch_density<double>(ctx, rho, psi, d_occ, nb, ncomp, st);
}
int launch_band_density_f32(const qm_ctx *ctx, double *rho, const float *psi, const double *d_occ, int nb,
                            int ncomp, hipStream_t st) {
    return launch_density<float>(ctx, rho, psi, d_occ, nb, ncomp, st);
}

}  // namespace qm
