// Fused FD8 kinetic + local potential kernel for gfx950, with an autotuner
// over its tile shapes.
//
// out = -1/2 (Delta_FD8 + 2 i k.grad_FD8 - |k|^2) psi + V_loc psi   (overwrites out)
//
// Mathematics: LaplacianKpoint8th_halo_f / LaplacianKpoint8th_halo
// (GridDifferenceKpoint8th.h:17-151) with coef = -1/2
// (qumasun_base1.matrix.h:298-317) plus mPotentialMatrixAdd_ddm_bundle
// (qumasun_base1.matrix.h:329-337).  The periodic halo of PasteHalo_f4
// (GridDifference8th.h:18-78) becomes wrapped loads.
//
// Arithmetic: along each axis the 9-point second difference and the 8-point
// first difference are written in forward differences D(x) = p(x+1) - p(x):
//   sum_j c_j (p(x+j) + p(x-j) - 2 p(x)) = sum_{i=0..3} C_i (D(x+i) - D(x-1-i)),
//   sum_j a_j (p(x+j) - p(x-j))          = sum_{i=0..3} A_i (D(x+i) + D(x-1-i)),
// with C_i = sum_{j>i} c_j and A_i = sum_{j>i} a_j.  Each point computes one D
// per axis and reads the other seven, so a point costs ~30 VALU operations
// instead of ~50 for the term-by-term form, and the kinetic diagonal
// (14350/5040 / h^2 ~ 100 Ha) never appears: rounding stays relative to the
// kinetic energy in FP32 as well as FP64.
//
// Layout: each 64-lane wave is independent (no LDS tile, no barrier).  Lanes
// run along x; each lane owns RY consecutive y rows and streams a chunk of z
// planes.  Per row it keeps the next 8 planes of psi (prefetched 4 planes
// ahead of use) and the 8 z differences around the current plane in
// registers; the 8-step unrolled loop rotates both queues by renaming
// registers, not by moves.  The y differences come from the lane's rows plus
// 8 halo rows, the x differences from other lanes through ds_bpermute.  When
// nx divides 64 the wave holds whole periodic rows (64/nx side by side) and
// the x neighbours are a rotation inside each row; otherwise waves cover
// overlapping 64-wide x windows and write the middle 56 points.  Loads and
// stores are buffer operations whose per-plane offset is a wave-uniform
// scalar, so addressing costs no VALU.  Blocks are mapped so that consecutive
// tiles run on one XCD and share its L2 for the halos.
#include <cmath>

#include "qm_internal.h"

namespace qm {
namespace {

constexpr int HW = 4;              // FD8 half width
constexpr int WAVE = 64;
constexpr int XW = WAVE - 2 * HW;  // useful width of an overlapping x window
constexpr unsigned kRsrcWord3 = 0x00020000u;  // raw 32-bit buffer resource, gfx9 family

template <typename T>
struct FD8Coef {
    T c0;         // |k|^2 / 2
    T lw[3][4];   // C_i = sum_{j>i} c_j, c_j = -1/2 {8064, -1008, 128, -9} / (5040 h_d^2)
    T gw[3][4];   // A_i = sum_{j>i} a_j, a_j = {672, -168, 32, -3} k_d / (840 h_d)
};

using rsrc_t = __amdgpu_buffer_rsrc_t;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ rsrc_t make_rsrc(const void *base, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, (int)kRsrcWord3);
}
__device__ __forceinline__ void bload(float &v, rsrc_t r, int voff, int soff) {
    v = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void bload(double &v, rsrc_t r, int voff, int soff) {
    v = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void bstore(float v, rsrc_t r, int voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}
__device__ __forceinline__ void bstore(double v, rsrc_t r, int voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}

__device__ __forceinline__ int wrapi(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

// value of v held by the lane whose byte address (4 * lane) is addr
__device__ __forceinline__ float lane_read(float v, int addr) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}
__device__ __forceinline__ double lane_read(double v, int addr) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)(unsigned)(b & 0xffffffffull));
    const unsigned hi = (unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// SEG: nx divides 64 and the wave holds 64/nx whole rows side by side.
// Queue slots: at step phase P (plane w = zz + P), logical slot m of pq holds
// plane w + m and logical slot m of dz holds D(w - 4 + m); physical slot is
// (m + P) & 7.
template <typename T, int NC, bool KT, int RY, bool SEG>
__global__ __launch_bounds__(256) void fd8_wave_kernel(T *__restrict__ out, const T *__restrict__ in,
                                                       const T *__restrict__ vloc, int nx, int ny, int nz,
                                                       int zchunk, int nzc, int nyt, int nxt, int nwaves, int nblk,
                                                       int per_xcd, FD8Coef<T> c) {
    // hardware block b runs on XCD b % 8: give each XCD a contiguous range of tiles
    const int b = blockIdx.x;
    const int lb = (b & 7) * per_xcd + (b >> 3);
    if (lb >= nblk) return;
    const int lane = threadIdx.x & (WAVE - 1);
    // the wave index is uniform: keep everything derived from it (tile, band,
    // buffer resources, plane offsets) in scalar registers
    const int gw = lb * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (gw >= nwaves) return;
    int t = gw;
    const int xt = t % nxt;
    t /= nxt;
    const int yt = t % nyt;
    t /= nyt;
    const int zc = t % nzc;
    const int band = t / nzc;

    // lanes holding x+1, x+2, x+3 (la[0..2]) and x-1 .. x-4 (lb_[0..3])
    int x, y0, la[3], lb_[4];
    bool xout;
    if constexpr (SEG) {
        const int S = WAVE / nx;
        const int seg = lane / nx;
        x = lane - seg * nx;
        y0 = (yt * S + seg) * RY;
        xout = true;
#pragma unroll
        for (int i = 0; i < 3; ++i) la[i] = (seg * nx + wrapi(x + i + 1, nx)) * 4;
#pragma unroll
        for (int i = 0; i < 4; ++i) lb_[i] = (seg * nx + wrapi(x - i - 1, nx)) * 4;
    } else {
        const int x0 = xt * XW;
        x = wrapi(x0 - HW + lane, nx);
        y0 = yt * RY;
        xout = lane >= HW && lane < WAVE - HW && x0 + lane - HW < nx;
#pragma unroll
        for (int i = 0; i < 3; ++i) la[i] = min(lane + i + 1, WAVE - 1) * 4;
#pragma unroll
        for (int i = 0; i < 4; ++i) lb_[i] = max(lane - i - 1, 0) * 4;
    }
    const int64_t plane = (int64_t)nx * ny;
    const int64_t ng = plane * nz;
    const int pbytes = (int)(plane * sizeof(T));
    const int zb = zc * zchunk;
    const int ze = min(zb + zchunk, nz);

    rsrc_t rin[NC], rout[NC];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
        rin[cc] = make_rsrc(in + ((int64_t)band * NC + cc) * ng, ng * (int64_t)sizeof(T));
        rout[cc] = make_rsrc(out + ((int64_t)band * NC + cc) * ng, ng * (int64_t)sizeof(T));
    }
    const rsrc_t rv = make_rsrc(vloc, ng * (int64_t)sizeof(T));

    // byte offsets inside a plane; every load is wrapped into the grid, only
    // the stores are masked
    int rowoff[RY];
    bool rowok[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        rowoff[r] = (x + nx * wrapi(y0 + r, ny)) * (int)sizeof(T);
        rowok[r] = xout && (y0 + r < ny);
    }
    int hoff[8];  // rows y0-4..y0-1 and y0+RY..y0+RY+3
#pragma unroll
    for (int j = 0; j < 8; ++j)
        hoff[j] = (x + nx * wrapi(j < 4 ? y0 - 4 + j : y0 + RY + (j - 4), ny)) * (int)sizeof(T);

    T pq[NC][RY][8], dz[NC][RY][8];
    // planes zb .. zb+7 into pq; D(zb-4) .. D(zb+2) into dz
    {
        T lo[NC][RY][4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int so = wrapi(zb - 4 + m, nz) * pbytes;
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                for (int r = 0; r < RY; ++r) bload(lo[cc][r][m], rin[cc], rowoff[r], so);
        }
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int so = wrapi(zb + m, nz) * pbytes;
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                for (int r = 0; r < RY; ++r) bload(pq[cc][r][m], rin[cc], rowoff[r], so);
        }
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int r = 0; r < RY; ++r) {
#pragma unroll
                for (int m = 0; m < 3; ++m) dz[cc][r][m] = lo[cc][r][m + 1] - lo[cc][r][m];
                dz[cc][r][3] = pq[cc][r][0] - lo[cc][r][3];
#pragma unroll
                for (int m = 4; m < 7; ++m) dz[cc][r][m] = pq[cc][r][m - 3] - pq[cc][r][m - 4];
                dz[cc][r][7] = T(0);
            }
    }
    T hy[NC][8], vv[RY];
    {
        const int so = zb * pbytes;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int j = 0; j < 8; ++j) bload(hy[cc][j], rin[cc], hoff[j], so);
#pragma unroll
        for (int r = 0; r < RY; ++r) bload(vv[r], rv, rowoff[r], so);
    }

    for (int zz = zb; zz < ze; zz += 8) {
#pragma unroll
        for (int P = 0; P < 8; ++P) {
            const int w = zz + P;
            if (w < ze) {  // wave-uniform
                // halo rows and V of plane w + 1
                T nh[NC][8], nv[RY];
                if (w + 1 < ze) {
                    const int so = (w + 1) * pbytes;
#pragma unroll
                    for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                        for (int j = 0; j < 8; ++j) bload(nh[cc][j], rin[cc], hoff[j], so);
#pragma unroll
                    for (int r = 0; r < RY; ++r) bload(nv[r], rv, rowoff[r], so);
                }
                // newest z difference D(w + 3) = p(w + 4) - p(w + 3)
#pragma unroll
                for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                    for (int r = 0; r < RY; ++r)
                        dz[cc][r][(7 + P) & 7] = pq[cc][r][(4 + P) & 7] - pq[cc][r][(3 + P) & 7];
                // y differences of the lane's column: rows y0-4 .. y0+RY+3
                T dy[NC][RY + 7];
#pragma unroll
                for (int cc = 0; cc < NC; ++cc) {
                    T e[RY + 8];
#pragma unroll
                    for (int k = 0; k < RY + 8; ++k)
                        e[k] = k < 4 ? hy[cc][k] : (k < RY + 4 ? pq[cc][k - 4][P & 7] : hy[cc][k - RY]);
#pragma unroll
                    for (int k = 0; k < RY + 7; ++k) dy[cc][k] = e[k + 1] - e[k];
                }
                const int so = w * pbytes;
#pragma unroll
                for (int r = 0; r < RY; ++r) {
                    T lap[NC], grd[NC];
#pragma unroll
                    for (int cc = 0; cc < NC; ++cc) {
                        const T p = pq[cc][r][P & 7];
                        // x differences: own D(x), then D(x+1..x+3) and D(x-1..x-4) from lanes
                        const T dx0 = lane_read(p, la[0]) - p;
                        T dxp[4], dxm[4];
                        dxp[0] = dx0;
#pragma unroll
                        for (int i = 0; i < 3; ++i) dxp[i + 1] = lane_read(dx0, la[i]);
#pragma unroll
                        for (int i = 0; i < 4; ++i) dxm[i] = lane_read(dx0, lb_[i]);
                        T l = KT ? c.c0 * p : T(0);
                        T g = T(0);
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const T zp = dz[cc][r][(4 + i + P) & 7], zm = dz[cc][r][(3 - i + P) & 7];
                            const T yp = dy[cc][r + 4 + i], ym = dy[cc][r + 3 - i];
                            l = fmadd(c.lw[0][i], dxp[i] - dxm[i], l);
                            l = fmadd(c.lw[1][i], yp - ym, l);
                            l = fmadd(c.lw[2][i], zp - zm, l);
                            if constexpr (KT) {
                                g = fmadd(c.gw[0][i], dxp[i] + dxm[i], g);
                                g = fmadd(c.gw[1][i], yp + ym, g);
                                g = fmadd(c.gw[2][i], zp + zm, g);
                            }
                        }
                        lap[cc] = fmadd(vv[r], p, l);
                        grd[cc] = g;
                    }
                    if constexpr (KT && NC == 2) {
                        // re += k.grad im, im -= k.grad re
                        lap[0] += grd[1];
                        lap[1] -= grd[0];
                    }
                    if (rowok[r]) {
#pragma unroll
                        for (int cc = 0; cc < NC; ++cc) bstore(lap[cc], rout[cc], rowoff[r], so);
                    }
                }
                // plane w + 8 into the slot plane w leaves; the chunk's last step
                // reads D(ze + 2), so planes past ze + 3 are never needed
                if (w + 8 <= ze + 3) {
                    int pf = w + 8;
                    while (pf >= nz) pf -= nz;
                    const int sf = pf * pbytes;
#pragma unroll
                    for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                        for (int r = 0; r < RY; ++r) bload(pq[cc][r][P & 7], rin[cc], rowoff[r], sf);
                }
                if (w + 1 < ze) {
#pragma unroll
                    for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                        for (int j = 0; j < 8; ++j) hy[cc][j] = nh[cc][j];
#pragma unroll
                    for (int r = 0; r < RY; ++r) vv[r] = nv[r];
                }
            }
        }
    }
}

template <typename T>
FD8Coef<T> make_coef(const qm_ctx *ctx, const double k[3]) {
    const double h[3] = {ctx->hx, ctx->hy, ctx->hz};
    const double cj[4] = {8064.0, -1008.0, 128.0, -9.0};
    const double aj[4] = {672.0, -168.0, 32.0, -3.0};
    FD8Coef<T> c;
    c.c0 = T(0.5 * (k[0] * k[0] + k[1] * k[1] + k[2] * k[2]));
    for (int d = 0; d < 3; ++d) {
        // the reference drops the first-derivative term when |k_d| < 1e-8
        // (GridDifferenceKpoint8th.h:330-365)
        const bool grad = !(std::fabs(k[d]) < 1.0e-8);
        for (int i = 0; i < 4; ++i) {
            double cs = 0.0, as = 0.0;
            for (int j = i; j < 4; ++j) {
                cs += -0.5 * cj[j] / (5040.0 * h[d] * h[d]);
                as += aj[j] * k[d] / (840.0 * h[d]);
            }
            c.lw[d][i] = T(cs);
            c.gw[d][i] = grad ? T(as) : T(0);
        }
    }
    return c;
}

// ---------------------------------------------------------------------------
// candidates, autotuner and launch
// ---------------------------------------------------------------------------

struct Candidate {
    int ry;      // rows per lane
    int zchunk;  // z planes per wave (a multiple of 8)
};
constexpr Candidate kCandidates[] = {
    {4, 32}, {2, 16}, {2, 32}, {4, 16}, {4, 64}, {8, 16}, {8, 32}, {8, 64},
};
constexpr int kNumCandidates = sizeof(kCandidates) / sizeof(kCandidates[0]);

template <typename T, int NC, bool KT, int RY>
int launch_wave(const qm_ctx *ctx, T *out, const T *in, const T *vloc, int nb, int zchunk, const FD8Coef<T> &c,
                hipStream_t st) {
    const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
    const bool seg = WAVE % nx == 0;
    const int rows = seg ? (WAVE / nx) * RY : RY;  // y rows per wave
    const int nxt = seg ? 1 : (nx + XW - 1) / XW;
    const int nyt = (ny + rows - 1) / rows;
    const int nzc = (nz + zchunk - 1) / zchunk;
    const int64_t nwaves = (int64_t)nb * nzc * nyt * nxt;
    QM_CHECK(nwaves < (int64_t)INT32_MAX / 2, QM_ERR_ARG, "grid too large for one stencil launch");
    const int nblk = (int)((nwaves + 3) / 4);
    const int per_xcd = (nblk + 7) / 8;
    if (seg)
        hipLaunchKernelGGL((fd8_wave_kernel<T, NC, KT, RY, true>), dim3(8 * per_xcd), dim3(256), 0, st, out, in,
                           vloc, nx, ny, nz, zchunk, nzc, nyt, nxt, (int)nwaves, nblk, per_xcd, c);
    else
        hipLaunchKernelGGL((fd8_wave_kernel<T, NC, KT, RY, false>), dim3(8 * per_xcd), dim3(256), 0, st, out, in,
                           vloc, nx, ny, nz, zchunk, nzc, nyt, nxt, (int)nwaves, nblk, per_xcd, c);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

template <typename T, int NC, bool KT>
int launch_candidate(const qm_ctx *ctx, int cand, T *out, const T *in, const T *vloc, int nb, const FD8Coef<T> &c,
                     hipStream_t st) {
    const Candidate cd = kCandidates[cand];
    // 8 rows per lane would spill registers for FP64 or complex bands: 4 there
    constexpr int RY8 = (sizeof(T) == 8 || NC == 2) ? 4 : 8;
    switch (cd.ry) {
        case 2: return launch_wave<T, NC, KT, 2>(ctx, out, in, vloc, nb, cd.zchunk, c, st);
        case 4: return launch_wave<T, NC, KT, 4>(ctx, out, in, vloc, nb, cd.zchunk, c, st);
        default: return launch_wave<T, NC, KT, RY8>(ctx, out, in, vloc, nb, cd.zchunk, c, st);
    }
}

template <typename T>
int launch_any(const qm_ctx *ctx, int cand, int ncomp, bool kt, T *out, const T *in, const T *vloc, int nb,
               const FD8Coef<T> &c, hipStream_t st) {
    if (ncomp == 1) return launch_candidate<T, 1, false>(ctx, cand, out, in, vloc, nb, c, st);
    if (!kt) return launch_candidate<T, 2, false>(ctx, cand, out, in, vloc, nb, c, st);
    return launch_candidate<T, 2, true>(ctx, cand, out, in, vloc, nb, c, st);
}

int64_t tune_key(int precision, int ncomp, bool kt, int nb) {
    return ((((int64_t)nb * 2 + (ncomp - 1)) * 2 + (kt ? 1 : 0)) << 8) | precision;
}

// Times every candidate on the caller's buffers (each run overwrites out with
// the same values) and returns the fastest in *best.
template <typename T>
int autotune(const qm_ctx *ctx, int ncomp, bool kt, T *out, const T *in, const T *vloc, int nb, const FD8Coef<T> &c,
             hipStream_t st, int *best) {
    hipEvent_t ev[2];
    QM_HIP(hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        QM_CHECK(false, QM_ERR_HIP, "stencil autotune: hipEventCreate failed");
    }
    int status = QM_OK;
    float best_ms = 1e30f;
    *best = 0;
    for (int cand = 0; cand < kNumCandidates && status == QM_OK; ++cand) {
        float t_min = 1e30f;
        for (int rep = 0; rep < 3 && status == QM_OK; ++rep) {  // the first run warms up
            float ms = 0.0f;
            if (hipEventRecord(ev[0], st) != hipSuccess) status = QM_ERR_HIP;
            if (status == QM_OK) status = launch_any<T>(ctx, cand, ncomp, kt, out, in, vloc, nb, c, st);
            if (status == QM_OK && hipEventRecord(ev[1], st) != hipSuccess) status = QM_ERR_HIP;
            if (status == QM_OK && hipEventSynchronize(ev[1]) != hipSuccess) status = QM_ERR_HIP;
            if (status == QM_OK && hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) status = QM_ERR_HIP;
            if (status == QM_OK && rep > 0) t_min = std::min(t_min, ms);
        }
        if (status == QM_OK && t_min < best_ms) {
            best_ms = t_min;
            *best = cand;
        }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (status == QM_ERR_HIP) set_error("stencil autotune: HIP event failure");
    return status;
}

template <typename T>
int launch_fd8(qm_ctx *ctx, T *out, const T *in, const T *vloc, int nb, int ncomp, const double k[3],
               hipStream_t st) {
    QM_CHECK(ctx->nx >= HW && ctx->ny >= HW && ctx->nz >= HW, QM_ERR_ARG, "grid dimensions must be >= 4");
    QM_CHECK(ctx->ng * (int64_t)sizeof(T) < (int64_t)INT32_MAX, QM_ERR_ARG,
             "grid too large for 32-bit buffer offsets (Ng * sizeof(T) must be < 2 GiB)");
    const bool kt = !(std::fabs(k[0]) < 1.0e-8 && std::fabs(k[1]) < 1.0e-8 && std::fabs(k[2]) < 1.0e-8);
    QM_CHECK(ncomp == 2 || !kt, QM_ERR_ARG, "real (ncomp=1) bands need a Gamma k-point");
    const FD8Coef<T> c = make_coef<T>(ctx, k);
    int cand = ctx->stencil_force;
    if (cand < 0) {
        const int64_t key = tune_key(sizeof(T) == 8 ? 64 : 32, ncomp, kt, nb);
        auto it = ctx->stencil_choice.find(key);
        if (it != ctx->stencil_choice.end()) {
            cand = it->second;
        } else {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            QM_HIP(hipStreamIsCapturing(st, &cs));
            if (cs != hipStreamCaptureStatusNone) {
                cand = 0;  // no timing inside a capture: the default shape
            } else {
                QM_TRY(autotune<T>(ctx, ncomp, kt, out, in, vloc, nb, c, st, &cand));
                ctx->stencil_choice[key] = cand;
            }
        }
    }
    return launch_any<T>(ctx, cand, ncomp, kt, out, in, vloc, nb, c, st);
}

}  // namespace

int stencil_num_candidates() { return kNumCandidates; }

std::string stencil_candidate_name(int cand) {
    const Candidate cd = kCandidates[cand];
    return "fd8_wave_kernel: RY " + std::to_string(cd.ry) + " (4 for FP64 / complex when 8), z chunk " +
           std::to_string(cd.zchunk);
}

int stencil_choice(const qm_ctx *ctx, int precision, int ncomp, int kt, int nb) {
    if (ctx->stencil_force >= 0) return ctx->stencil_force;
    auto it = ctx->stencil_choice.find(tune_key(precision, ncomp, kt != 0, nb));
    return it == ctx->stencil_choice.end() ? -1 : it->second;
}

int launch_fd8_local_f64(qm_ctx *ctx, double *out, const double *in, int nb, int ncomp, const double k[3],
                         hipStream_t st) {
    return launch_fd8<double>(ctx, out, in, ctx->d_vloc64, nb, ncomp, k, st);
}

int launch_fd8_local_f32(qm_ctx *ctx, float *out, const float *in, int nb, int ncomp, const double k[3],
                         hipStream_t st) {
    return launch_fd8<float>(ctx, out, in, ctx->d_vloc32, nb, ncomp, k, st);
}

}  // namespace qm
