// SCF glue around the LOBPCG solve: the pieces of QUMASUN_KPOINT::mSCF's loop
// (qumasun_kpoint.scfloop.h:96-222) that sit between the eigensolver and the
// next H.psi, on the GPU where they touch grid fields:
//   - LDA exchange-correlation (Calc_XC_LDA, Vxc.h:17-76: Perdew-Zunger 1981),
//     one pass over the density writing V_xc and, optionally, rho e_xc;
//   - integrals over the grid, sum_i a_i b_i, in a fixed summation order (the
//     energy terms of mGetTotalEnergy, qumasun_base1.energy.h:19-160);
//   - density mixing (MixingSimple and MixingDIIS, mixing_simple.h:45-75,
//     mixing_DIIS.h:48-229) on device fields, the small DIIS matrix on the host;
//   - occupations (mSetOccupancyZero / mSetOccupancyTemperature,
//     qumasun_base1.occupancy.h:35-175 with SetOccupancy.h), host code.
// The Hartree potential (FFT Poisson solve) is in kinetic_fft.hip.
#include <cfloat>
#include <cmath>
#include <vector>

#include "qm_internal.h"

struct qm_mixer {
    qm_ctx *ctx = nullptr;
    int mode = 0;  // QM_MIX_SIMPLE or QM_MIX_DIIS
    int hist = 1;
    double ratio = 0.25;
    int step = 0;
    bool ready = false;
    int64_t n = 0;
    double *rho = nullptr;  // DIIS: hist fields; simple: the previous field
    double *R = nullptr;    // DIIS residual fields
    double *part = nullptr; // per-block partial sums (kMixBlocks x (hist + 1))
    double *coef = nullptr; // DIIS coefficients (hist)
    std::vector<double> M;  // (hist + 1)^2 bordered R.R matrix, column-major
};

namespace qm {
namespace {

constexpr int kRedBlocks = 512;  // partial sums per reduction (fixed, so the order is too)
constexpr int kRedThreads = 256;

__global__ __launch_bounds__(256) void xc_lda_kernel(const double *__restrict__ rho, double *__restrict__ vxc,
                                                     double *__restrict__ exc, int64_t n) {
    const double c2 = cbrt(3.0 / (4.0 * M_PI));
    const double c3 = -(3.0 / (4.0 * M_PI)) * cbrt(9.0 * M_PI / 4.0);
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double r = rho[i];
        double v = 0.0, e = 0.0;
        if (r >= 1.0e-8) {
            const double rs = c2 / cbrt(r);
            const double ex = c3 / rs;
            double ec, nd;
            if (rs >= 1.0) {  // gamma / (1 + b1 sqrt(rs) + b2 rs)
                const double g = -0.1423, b1 = 1.0529, b2 = 0.3334;
                const double sr = sqrt(rs);
                ec = g / (1.0 + b1 * sr + b2 * rs);
                nd = ec * ec * (b1 * 0.5 * sr + b2 * rs) / (3.0 * g);
            } else {  // A ln rs + B + C rs ln rs + D rs
                const double A = 0.0311, B = -0.0480, C = 0.0020, D = -0.0116;
                const double lr = log(rs);
                ec = B + A * lr + D * rs + C * rs * lr;
                nd = -(A + (D + C) * rs + C * rs * lr) / 3.0;
            }
            v = (ex + ex / 3.0) + (ec + nd);
            e = r * ex + r * ec;
        }
        vxc[i] = v;
        if (exc) exc[i] = e;
    }
}

// part[b] = sum over block b's strided share of a_i * b_i (MODE 0) or
// |a_i - b_i| (MODE 1), each block summing in a fixed order
template <int MODE>
__global__ __launch_bounds__(kRedThreads) void reduce_kernel(const double *__restrict__ a,
                                                             const double *__restrict__ b, int64_t n,
                                                             double *__restrict__ part) {
    __shared__ double red[kRedThreads / 64];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (int64_t)kRedBlocks * kRedThreads)
        s += MODE == 0 ? a[i] * b[i] : fabs(a[i] - b[i]);
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kRedThreads / 64; ++w) t += red[w];
        part[blockIdx.x] = t;
    }
}

// R = out - rho_k and part = per-block sums of |R|
__global__ __launch_bounds__(kRedThreads) void residual_kernel(double *__restrict__ R, const double *__restrict__ out,
                                                               const double *__restrict__ rk, int64_t n,
                                                               double *__restrict__ part) {
    __shared__ double red[kRedThreads / 64];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (int64_t)kRedBlocks * kRedThreads) {
        const double g = out[i] - rk[i];
        R[i] = g;
        s += fabs(g);
    }
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kRedThreads / 64; ++w) t += red[w];
        part[blockIdx.x] = t;
    }
}

// simple mixing: out = (out - prev) ratio + prev; prev = out
__global__ __launch_bounds__(256) void mix_simple_kernel(double *__restrict__ out, double *__restrict__ prev,
                                                         double ratio, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double d = out[i] - prev[i];
        const double v = d * ratio + prev[i];
        out[i] = v;
        prev[i] = v;
    }
}

// DIIS step: R_n = sum_m c_m R_m, rho_n = sum_m c_m rho_m + ratio R_n (slot nxt
// first, then the others in order, as the reference accumulates), out = rho_n
__global__ __launch_bounds__(256) void mix_diis_kernel(double *__restrict__ out, double *__restrict__ rho,
                                                       double *__restrict__ R, const double *__restrict__ c, int hist,
                                                       int nxt, double ratio, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        double r = R[(int64_t)nxt * n + i] * c[nxt];
        double x = rho[(int64_t)nxt * n + i] * c[nxt];
        for (int m = 0; m < hist; ++m) {
            if (m == nxt) continue;
            r += R[(int64_t)m * n + i] * c[m];
            x += rho[(int64_t)m * n + i] * c[m];
        }
        x += r * ratio;
        R[(int64_t)nxt * n + i] = r;
        rho[(int64_t)nxt * n + i] = x;
        out[i] = x;
    }
}

// rho_nxt = rho_k + ratio R_k; out = rho_nxt
__global__ __launch_bounds__(256) void mix_first_kernel(double *__restrict__ out, double *__restrict__ rnext,
                                                        const double *__restrict__ rk, const double *__restrict__ Rk,
                                                        double ratio, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double v = rk[i] + Rk[i] * ratio;
        rnext[i] = v;
        out[i] = v;
    }
}

inline unsigned grid_for(int64_t n) {
    return (unsigned)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 4096);
}

// sum of the kRedBlocks partials in order (on the host)
int sum_partials(const double *d_part, int count, double *out, hipStream_t st) {
    std::vector<double> h((size_t)count * kRedBlocks);
    QM_HIP(hipMemcpyAsync(h.data(), d_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    QM_HIP(hipStreamSynchronize(st));
    for (int c = 0; c < count; ++c) {
        double t = 0.0;
        for (int b = 0; b < kRedBlocks; ++b) t += h[(size_t)c * kRedBlocks + b];
        out[c] = t;
    }
    return QM_OK;
}

// Gauss-Jordan inverse without pivoting of the column-major n x n m (destroyed),
// as InverseM (inverse_m.h:17-60)
void inverse_m(std::vector<double> &m, int n, std::vector<double> &inv) {
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[i + (size_t)i * n] = 1.0;
    for (int i = 0; i < n; ++i) {
        const double b = 1.0 / m[i + (size_t)i * n];
        if (std::isfinite(b))
            for (int j = 0; j < n; ++j) {
                m[i + (size_t)j * n] *= b;
                inv[i + (size_t)j * n] *= b;
            }
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const double b2 = m[j + (size_t)i * n];
            for (int k = 0; k < n; ++k) {
                m[j + (size_t)k * n] -= m[i + (size_t)k * n] * b2;
                inv[j + (size_t)k * n] -= inv[i + (size_t)k * n] * b2;
            }
        }
    }
}

void mixer_free(qm_mixer *m) {
    for (double **p : {&m->rho, &m->R, &m->part, &m->coef})
        if (*p) (void)hipFree(*p);
}

}  // namespace
}  // namespace qm

using namespace qm;

extern "C" {

int qm_xc_lda_f64(qm_ctx *ctx, const double *d_rho, double *d_vxc, double *d_exc, void *stream) {
    QM_CHECK(ctx && d_rho && d_vxc, QM_ERR_ARG, "qm_xc_lda_f64: NULL argument");
    QM_TRY(activate(ctx));
    hipLaunchKernelGGL(xc_lda_kernel, dim3(grid_for(ctx->ng)), dim3(256), 0, (hipStream_t)stream, d_rho, d_vxc,
                       d_exc, ctx->ng);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

int qm_grid_dot_f64(qm_ctx *ctx, const double *d_a, const double *d_b, double *result, void *stream) {
    QM_CHECK(ctx && d_a && d_b && result, QM_ERR_ARG, "qm_grid_dot_f64: NULL argument");
    QM_TRY(activate(ctx));
    hipStream_t st = (hipStream_t)stream;
    QM_TRY(ctx->scf_part.reserve(sizeof(double) * kRedBlocks));
    hipLaunchKernelGGL(reduce_kernel<0>, dim3(kRedBlocks), dim3(kRedThreads), 0, st, d_a, d_b, ctx->ng,
                       ctx->scf_part.as<double>());
    QM_HIP(hipGetLastError());
    return sum_partials(ctx->scf_part.as<double>(), 1, result, st);
}

int qm_mixer_create(qm_mixer **out, qm_ctx *ctx, int mode, int history, double ratio) {
    QM_CHECK(out && ctx, QM_ERR_ARG, "qm_mixer_create: NULL argument");
    QM_CHECK(mode == QM_MIX_SIMPLE || mode == QM_MIX_DIIS, QM_ERR_ARG, "qm_mixer_create: unknown mode");
    QM_CHECK(mode == QM_MIX_SIMPLE || (history >= 2 && history <= 64), QM_ERR_ARG,
             "qm_mixer_create: DIIS history must be in [2, 64]");
    QM_CHECK(std::isfinite(ratio), QM_ERR_ARG, "qm_mixer_create: ratio is not finite");
    *out = nullptr;
    QM_TRY(activate(ctx));
    qm_mixer *m = new qm_mixer();
    m->ctx = ctx;
    m->mode = mode;
    m->hist = mode == QM_MIX_DIIS ? history : 1;
    m->ratio = ratio;
    m->n = ctx->ng;
    const size_t fb = sizeof(double) * (size_t)m->n * m->hist;
    hipError_t e = hipMalloc((void **)&m->rho, fb);
    if (e == hipSuccess && mode == QM_MIX_DIIS) e = hipMalloc((void **)&m->R, fb);
    if (e == hipSuccess) e = hipMalloc((void **)&m->part, sizeof(double) * kRedBlocks * (m->hist + 1));
    if (e == hipSuccess) e = hipMalloc((void **)&m->coef, sizeof(double) * (m->hist + 1));
    if (e != hipSuccess) {
        mixer_free(m);
        delete m;
        set_error("qm_mixer_create: hipMalloc failed");
        return QM_ERR_NOMEM;
    }
    m->M.assign((size_t)(m->hist + 1) * (m->hist + 1), 0.0);
    *out = m;
    return QM_OK;
}

void qm_mixer_destroy(qm_mixer *m) {
    if (!m) return;
    if (activate(m->ctx) == QM_OK) mixer_free(m);
    delete m;
}

int qm_mixer_init(qm_mixer *m, const double *d_rho, void *stream) {
    QM_CHECK(m && d_rho, QM_ERR_ARG, "qm_mixer_init: NULL argument");
    QM_TRY(activate(m->ctx));
    hipStream_t st = (hipStream_t)stream;
    const size_t fb = sizeof(double) * (size_t)m->n;
    if (m->mode == QM_MIX_DIIS) {
        QM_HIP(hipMemsetAsync(m->rho, 0, fb * m->hist, st));
        QM_HIP(hipMemsetAsync(m->R, 0, fb * m->hist, st));
    }
    QM_HIP(hipMemcpyAsync(m->rho, d_rho, fb, hipMemcpyDeviceToDevice, st));
    std::fill(m->M.begin(), m->M.end(), 0.0);
    m->step = 0;
    m->ready = true;
    return QM_OK;
}

int qm_mixer_predict(qm_mixer *m, double *d_rho, double *diff, void *stream) {
    QM_CHECK(m && d_rho, QM_ERR_ARG, "qm_mixer_predict: NULL argument");
    QM_CHECK(m->ready, QM_ERR_STATE, "qm_mixer_predict: qm_mixer_init has not been called");
    QM_TRY(activate(m->ctx));
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = m->n;
    double d = 0.0;
    if (m->mode == QM_MIX_SIMPLE) {
        hipLaunchKernelGGL(reduce_kernel<1>, dim3(kRedBlocks), dim3(kRedThreads), 0, st, d_rho, m->rho, n, m->part);
        QM_HIP(hipGetLastError());
        hipLaunchKernelGGL(mix_simple_kernel, dim3(grid_for(n)), dim3(256), 0, st, d_rho, m->rho, m->ratio, n);
        QM_HIP(hipGetLastError());
        QM_TRY(sum_partials(m->part, 1, &d, st));
        if (diff) *diff = d;
        ++m->step;
        return QM_OK;
    }
    const int H = m->hist, L = H + 1;
    auto idx = [H](int s) { return (s % H + H) % H; };
    const int k = idx(m->step);
    double *rk = m->rho + (int64_t)k * n, *Rk = m->R + (int64_t)k * n;
    hipLaunchKernelGGL(residual_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, st, Rk, d_rho, rk, n, m->part);
    QM_HIP(hipGetLastError());
    // R_k . R_m over the valid history
    const int m_end = std::max(m->step - H + 1, 0);
    std::vector<int> ms;
    for (int m1 = m->step; m1 >= m_end; --m1) ms.push_back(idx(m1));
    for (size_t q = 0; q < ms.size(); ++q) {
        hipLaunchKernelGGL(reduce_kernel<0>, dim3(kRedBlocks), dim3(kRedThreads), 0, st, Rk,
                           m->R + (int64_t)ms[q] * n, n, m->part + (size_t)(q + 1) * kRedBlocks);
        QM_HIP(hipGetLastError());
    }
    std::vector<double> sums(ms.size() + 1);
    QM_TRY(sum_partials(m->part, (int)ms.size() + 1, sums.data(), st));
    d = sums[0];
    for (size_t q = 0; q < ms.size(); ++q) m->M[ms[q] + (size_t)L * k] = sums[q + 1];
    for (int j = 0; j < H; ++j) {
        m->M[k + (size_t)L * j] = m->M[j + (size_t)L * k];
        m->M[H + (size_t)L * j] = 1.0;
        m->M[j + (size_t)L * H] = 1.0;
    }
    m->M[H + (size_t)L * H] = 0.0;
    const int nxt = idx(m->step + 1);
    if (m->step + 1 >= H) {
        std::vector<double> cp(m->M), inv;
        inverse_m(cp, L, inv);
        std::vector<double> c(L);
        for (int j = 0; j < L; ++j) c[j] = inv[H + (size_t)L * j];  // the bordering row
        QM_HIP(hipMemcpyAsync(m->coef, c.data(), sizeof(double) * L, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(mix_diis_kernel, dim3(grid_for(n)), dim3(256), 0, st, d_rho, m->rho, m->R, m->coef, H,
                           nxt, m->ratio, n);
    } else {
        hipLaunchKernelGGL(mix_first_kernel, dim3(grid_for(n)), dim3(256), 0, st, d_rho, m->rho + (int64_t)nxt * n,
                           rk, Rk, m->ratio, n);
    }
    QM_HIP(hipGetLastError());
    QM_HIP(hipStreamSynchronize(st));  // the coefficient upload reads host memory
    if (diff) *diff = d;
    ++m->step;
    return QM_OK;
}

int qm_occupancy(int nk, int nb, const double *eig, const double *kweight, double nelec, double kbT, double *occ,
                 double *mu) {
    QM_CHECK(nk >= 1 && nb >= 1 && eig && kweight && occ, QM_ERR_ARG, "qm_occupancy: bad argument");
    double wsum = 0.0;
    for (int k = 0; k < nk; ++k) {
        QM_CHECK(kweight[k] >= 0.0, QM_ERR_ARG, "qm_occupancy: negative k-point weight");
        wsum += kweight[k];
    }
    QM_CHECK(wsum > 0.0 && nelec >= 0.0, QM_ERR_ARG, "qm_occupancy: zero total weight or negative electron count");
    const double kb_hartree = 1.380649e-23 / 4.3597447222071e-18;  // Hartree per kelvin
    if (kbT < 0.1 * kb_hartree) {  // aufbau (mSetOccupancyZero)
        for (int k = 0; k < nk; ++k) {
            const double we = 2.0 * kweight[k] / wsum;
            double rem = nelec / 2.0;
            for (int n = 0; n < nb; ++n) {
                double o = 0.0;
                if (rem >= 1.0) {
                    o = we;
                    rem -= 1.0;
                } else if (rem > 0.0) {
                    o = we * rem;
                    rem = 0.0;
                }
                occ[(size_t)k * nb + n] = o;
            }
        }
        if (mu) *mu = NAN;
        return QM_OK;
    }
    // Fermi-Dirac: bisection on mu between the lowest and highest eigenvalue
    double lo = DBL_MAX, hi = -DBL_MAX;
    for (int k = 0; k < nk; ++k) {
        lo = std::min(lo, eig[(size_t)k * nb]);
        hi = std::max(hi, eig[(size_t)k * nb + nb - 1]);
    }
    double m = 0.5 * (lo + hi), used = m;
    for (size_t i = 0; i < (size_t)nk * nb; ++i) occ[i] = 0.0;
    bool go = std::fabs(hi - lo) > 1.0e-9;
    while (go) {
        double ne = 0.0;
        for (int k = 0; k < nk; ++k) {
            const double c = 2.0 * kweight[k] / wsum;
            for (int n = 0; n < nb; ++n) {
                const double p = std::min((eig[(size_t)k * nb + n] - m) / kbT, 80.0);
                const double o = c / (std::exp(p) + 1.0);
                occ[(size_t)k * nb + n] = o;
                ne += o;
            }
        }
        used = m;
        if (ne < nelec) {
            lo = m;
        } else if (ne > nelec) {
            hi = m;
        } else {
            break;
        }
        m = 0.5 * (lo + hi);
        go = std::fabs(hi - lo) > 1.0e-9;
    }
    if (mu) *mu = used;
    return QM_OK;
}

}  // extern "C"
