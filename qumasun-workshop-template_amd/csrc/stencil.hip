// Fused FD8 kinetic + local potential kernel for gfx950.
//
// out = -1/2 (Delta_FD8 + 2 i k.grad_FD8 - |k|^2) psi + V_loc psi   (overwrites out)
//
// Mathematics: LaplacianKpoint8th_halo_f / LaplacianKpoint8th_halo
// (GridDifferenceKpoint8th.h:17-151) with coef = -1/2
// (qumasun_base1.matrix.h:298-317) plus mPotentialMatrixAdd_ddm_bundle
// (qumasun_base1.matrix.h:329-337).  The periodic halo of PasteHalo_f4
// (GridDifference8th.h:18-78) becomes wrapped loads into LDS.
//
// Layout on the GPU: a block owns a TX x TY tile of the x-y plane of one band
// (x coalesced: one wave reads 64 consecutive x) and streams a chunk of z
// planes.  The +-4 z neighbours of each point live in a 9-deep register queue,
// so each grid value is read from HBM once per chunk; the x and y neighbours
// come from an LDS copy of the current plane with a 4-wide halo.  The local
// potential and the diagonal are folded into the same pass, so one H.psi
// reads psi once, V once per tile, and writes Hpsi once.
#include <cmath>

#include "qm_internal.h"

namespace qm {
namespace {

constexpr int TX = 64;   // tile width in x (= wave width)
constexpr int BY = 4;    // thread rows per block
constexpr int RY = 4;    // points per thread in y
constexpr int TY = BY * RY;
constexpr int HW = 4;    // FD8 half width
constexpr int LX = TX + 2 * HW;
constexpr int LY = TY + 2 * HW;

template <typename T>
struct FD8Coef {
    T c0;        // FP64: 14350/5040 * 1/2 * sum_d 1/h_d^2 + |k|^2/2; FP32: |k|^2/2
    T cx[4], cy[4], cz[4];   // -1/2 * {8064, -1008, 128, -9}/5040 / h_d^2
    T ax[4], ay[4], az[4];   // {672, -168, 32, -3}/840 * k_d / h_d
};

__device__ __forceinline__ int wrapi(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

template <typename T, int NC, bool KT>
__global__ __launch_bounds__(TX * BY) void fd8_local_kernel(T *__restrict__ out, const T *__restrict__ in,
                                                            const T *__restrict__ vloc, int nx, int ny, int nz,
                                                            int zchunk, int nzc, FD8Coef<T> c) {
    __shared__ T sm[NC][LY][LX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = ty * TX + tx;
    const int zc = blockIdx.x % nzc;
    const int band = blockIdx.x / nzc;
    const int y0 = blockIdx.y * TY;
    const int x0 = blockIdx.z * TX;
    const int zb = zc * zchunk;
    const int ze = min(zb + zchunk, nz);
    const int64_t plane = (int64_t)nx * ny;
    const int64_t ng = plane * nz;
    const int w = min(TX, nx - x0);   // valid tile width
    const int hgt = min(TY, ny - y0); // valid tile height

    const T *src[NC];
    T *dst[NC];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
        src[cc] = in + ((int64_t)band * NC + cc) * ng;
        dst[cc] = out + ((int64_t)band * NC + cc) * ng;
    }

    const bool xin = tx < w;
    int rowoff[RY];
    bool pin[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int ly = ty + BY * r;
        pin[r] = xin && (ly < hgt);
        rowoff[r] = pin[r] ? (x0 + tx) + nx * (y0 + ly) : 0;
    }
    // y halo: thread row ty loads halo rows ty and ty+BY (8 rows in total);
    // rows 0..3 sit above the tile, rows 4..7 below the last valid row.
    int yh_off[2];
    int yh_lds[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int hr = ty + BY * j;
        const int ly = hr < HW ? hr - HW : hgt + (hr - HW);
        yh_lds[j] = ly + HW;
        yh_off[j] = xin ? (x0 + tx) + nx * wrapi(y0 + ly, ny) : 0;
    }
    // x halo: 8 columns x TY rows, threads 0..8*TY-1
    const bool xh_on = tid < 8 * TY;
    const int xh_col = tid & 7, xh_row = tid >> 3;
    const int xh_lx = xh_col < HW ? xh_col - HW : w + (xh_col - HW);
    const bool xh_valid = xh_on && xh_row < hgt;
    const int xh_off = xh_valid ? wrapi(x0 + xh_lx, nx) + nx * (y0 + xh_row) : 0;

    T q[NC][RY][9];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int64_t zo = (int64_t)wrapi(zb - HW + j, nz) * plane;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int r = 0; r < RY; ++r) q[cc][r][j] = pin[r] ? src[cc][zo + rowoff[r]] : T(0);
    }

    for (int z = zb; z < ze; ++z) {
        const int64_t zo = (int64_t)z * plane;
        const int64_t zf = (int64_t)wrapi(z + HW, nz) * plane;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int r = 0; r < RY; ++r) q[cc][r][8] = pin[r] ? src[cc][zf + rowoff[r]] : T(0);
        T vv[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) vv[r] = pin[r] ? vloc[zo + rowoff[r]] : T(0);

        // plane z into LDS: centre from the register queue, halos from global
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
#pragma unroll
            for (int r = 0; r < RY; ++r)
                if (pin[r]) sm[cc][ty + BY * r + HW][tx + HW] = q[cc][r][4];
#pragma unroll
            for (int j = 0; j < 2; ++j)
                if (xin) sm[cc][yh_lds[j]][tx + HW] = src[cc][zo + yh_off[j]];
            if (xh_valid) sm[cc][xh_row + HW][xh_lx + HW] = src[cc][zo + xh_off];
        }
        __syncthreads();

#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int lx = tx + HW, ly = ty + BY * r + HW;
            T res[NC];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) {
                const T p = q[cc][r][4];
                T lap;
                if constexpr (sizeof(T) == 8) {
                    // reference grouping: c0 p + sum_j c_j (p[+j] + p[-j])
                    lap = c.c0 * p;
#pragma unroll
                    for (int j = 1; j <= 4; ++j) {
                        lap += c.cx[j - 1] * (sm[cc][ly][lx + j] + sm[cc][ly][lx - j]);
                        lap += c.cy[j - 1] * (sm[cc][ly + j][lx] + sm[cc][ly - j][lx]);
                        lap += c.cz[j - 1] * (q[cc][r][4 + j] + q[cc][r][4 - j]);
                    }
                } else {
                    // FP32: difference form sum_j c_j ((p[+j] - p) + (p[-j] - p)).  The
                    // kinetic diagonal (~14350/5040 / h^2 ~ 100 Ha) never appears, so
                    // rounding is relative to the kinetic energy and constants are
                    // annihilated exactly whatever the rounding of the taps.
                    lap = c.c0 * p;  // |k|^2 / 2 only in this form
#pragma unroll
                    for (int j = 1; j <= 4; ++j) {
                        lap += c.cx[j - 1] * ((sm[cc][ly][lx + j] - p) + (sm[cc][ly][lx - j] - p));
                        lap += c.cy[j - 1] * ((sm[cc][ly + j][lx] - p) + (sm[cc][ly - j][lx] - p));
                        lap += c.cz[j - 1] * ((q[cc][r][4 + j] - p) + (q[cc][r][4 - j] - p));
                    }
                }
                res[cc] = lap + vv[r] * p;
            }
            if constexpr (KT && NC == 2) {
                T d_re = T(0), d_im = T(0);
#pragma unroll
                for (int j = 1; j <= 4; ++j) {
                    d_re += c.ax[j - 1] * (sm[0][ly][lx + j] - sm[0][ly][lx - j]);
                    d_re += c.ay[j - 1] * (sm[0][ly + j][lx] - sm[0][ly - j][lx]);
                    d_re += c.az[j - 1] * (q[0][r][4 + j] - q[0][r][4 - j]);
                    d_im += c.ax[j - 1] * (sm[1][ly][lx + j] - sm[1][ly][lx - j]);
                    d_im += c.ay[j - 1] * (sm[1][ly + j][lx] - sm[1][ly - j][lx]);
                    d_im += c.az[j - 1] * (q[1][r][4 + j] - q[1][r][4 - j]);
                }
                res[0] += d_im;
                res[1] -= d_re;
            }
            if (pin[r]) {
#pragma unroll
                for (int cc = 0; cc < NC; ++cc) dst[cc][zo + rowoff[r]] = res[cc];
            }
        }
        __syncthreads();
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int r = 0; r < RY; ++r)
#pragma unroll
                for (int j = 0; j < 8; ++j) q[cc][r][j] = q[cc][r][j + 1];
    }
}

template <typename T>
FD8Coef<T> make_coef(const qm_ctx *ctx, const double k[3]) {
    const double h[3] = {ctx->hx, ctx->hy, ctx->hz};
    const double cj[4] = {8064.0, -1008.0, 128.0, -9.0};
    const double aj[4] = {672.0, -168.0, 32.0, -3.0};
    FD8Coef<T> c;
    const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    // FP64 keeps the reference's diagonal; the FP32 difference form folds the
    // -2 sum_j c_j part of it into the taps and keeps only |k|^2 / 2
    c.c0 = sizeof(T) == 8 ? T(14350.0 / 5040.0 * 0.5 * (1.0 / (h[0] * h[0]) + 1.0 / (h[1] * h[1]) + 1.0 / (h[2] * h[2])) +
                             0.5 * k2)
                          : T(0.5 * k2);
    T *cs[3] = {c.cx, c.cy, c.cz};
    T *as[3] = {c.ax, c.ay, c.az};
    for (int d = 0; d < 3; ++d)
        for (int j = 0; j < 4; ++j) {
            cs[d][j] = T(-0.5 * cj[j] / (5040.0 * h[d] * h[d]));
            // the reference drops the first-derivative term when |k_d| < 1e-8
            // (GridDifferenceKpoint8th.h:330-365)
            as[d][j] = std::fabs(k[d]) < 1.0e-8 ? T(0) : T(aj[j] * k[d] / (840.0 * h[d]));
        }
    return c;
}

template <typename T>
int launch_fd8(const qm_ctx *ctx, T *out, const T *in, const T *vloc, int nb, int ncomp, const double k[3],
               hipStream_t st) {
    QM_CHECK(ctx->nx >= HW && ctx->ny >= HW && ctx->nz >= HW, QM_ERR_ARG, "grid dimensions must be >= 4");
    const bool kt = !(std::fabs(k[0]) < 1.0e-8 && std::fabs(k[1]) < 1.0e-8 && std::fabs(k[2]) < 1.0e-8);
    QM_CHECK(ncomp == 2 || !kt, QM_ERR_ARG, "real (ncomp=1) bands need a Gamma k-point");
    const FD8Coef<T> c = make_coef<T>(ctx, k);
    const int ntx = (ctx->nx + TX - 1) / TX;
    const int nty = (ctx->ny + TY - 1) / TY;
    // z chunk: long enough to amortise the 8-plane queue warm-up, short enough
    // to put >= 4 blocks on each of the 256 CUs
    int zchunk = 32;
    while (zchunk > 8 && (int64_t)ntx * nty * nb * ((ctx->nz + zchunk - 1) / zchunk) < 1024) zchunk /= 2;
    const int nzc = (ctx->nz + zchunk - 1) / zchunk;
    QM_CHECK((int64_t)nb * nzc < (int64_t)INT32_MAX && nty <= 65535 && ntx <= 65535, QM_ERR_ARG,
             "grid too large for one launch");
    dim3 grid(nb * nzc, nty, ntx), block(TX, BY);
    if (ncomp == 1)
        hipLaunchKernelGGL((fd8_local_kernel<T, 1, false>), grid, block, 0, st, out, in, vloc, ctx->nx, ctx->ny,
                           ctx->nz, zchunk, nzc, c);
    else if (!kt)
        hipLaunchKernelGGL((fd8_local_kernel<T, 2, false>), grid, block, 0, st, out, in, vloc, ctx->nx, ctx->ny,
                           ctx->nz, zchunk, nzc, c);
    else
        hipLaunchKernelGGL((fd8_local_kernel<T, 2, true>), grid, block, 0, st, out, in, vloc, ctx->nx, ctx->ny,
                           ctx->nz, zchunk, nzc, c);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

}  // namespace

int launch_fd8_local_f64(const qm_ctx *ctx, double *out, const double *in, int nb, int ncomp, const double k[3],
                         hipStream_t st) {
    return launch_fd8<double>(ctx, out, in, ctx->d_vloc64, nb, ncomp, k, st);
}

int launch_fd8_local_f32(const qm_ctx *ctx, float *out, const float *in, int nb, int ncomp, const double k[3],
                         hipStream_t st) {
    return launch_fd8<float>(ctx, out, in, ctx->d_vloc32, nb, ncomp, k, st);
}

}  // namespace qm
