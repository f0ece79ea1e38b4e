// Fused FD8 kinetic + local potential kernels for gfx950, with an autotuner
// over their tile shapes.
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
// Layout common to both kernels: lanes run along x; each lane owns RY
// consecutive y rows and streams a chunk of z planes.  Per row it keeps the
// next 8 planes of psi (requested 4 planes ahead of first use) and the 8 z
// differences around the current plane in registers; the 8-step unrolled
// loop rotates both queues by renaming registers, not by moves.  The x
// differences come from other lanes through ds_bpermute.  When nx divides 64
// the wave holds whole periodic rows (64/nx side by side) and the x
// neighbours are a rotation inside each row; otherwise waves cover
// overlapping 64-wide x windows and write the middle 56 points.  Loads and
// stores are buffer operations whose per-plane offset is a wave-uniform
// scalar, so addressing costs no VALU; stores are non-temporal, so the
// written planes do not push psi out of L2.  Workgroups are mapped so that
// consecutive tiles run on one XCD and share its L2.
//
// fd8_wave_kernel: every wave is independent (no LDS tile, no barrier) and
// reads the 8 y-halo rows of its rows from memory.  When those re-reads miss
// L2 they cost up to (RY + 8) / RY times the psi traffic.
//
// fd8_slab_kernel: the waves of a workgroup stack in y over a slab of TY rows.
// Every step each wave writes its rows of the current plane into a
// double-buffered LDS tile, the workgroup meets at one barrier, and each lane
// reads its y-halo rows from the tile.  Only the 8 rows around the slab come
// from memory (requested 4 planes ahead), none when the slab is the whole
// periodic y range, so psi is read about (TY + 8) / TY times per z chunk.
//
// Both kernels do the same arithmetic in the same order for every point, so
// every candidate gives bit-identical results.
#include <cmath>

#include "qm_internal.h"

namespace qm {
namespace {

constexpr int HW = 4;              // FD8 half width
constexpr int WAVE = 64;
constexpr int XW = WAVE - 2 * HW;  // useful width of an overlapping x window
constexpr unsigned kRsrcWord3 = 0x00020000u;  // raw 32-bit buffer resource, gfx9 family
constexpr int kStoreNT = 2;                   // buffer cache policy: non-temporal
constexpr int kSlabLdsMax = 65536;            // bytes of LDS a slab workgroup may use
// planes between the request of a halo row (wave kernel) or of V and its use;
// longer leads cost more registers than they save in waiting
constexpr int kHLead = 1;
constexpr int kVLead = 1;
// x neighbours through DPP wave shifts where a wave holds one 64-point row or
// an x window (else ds_bpermute)
constexpr bool kXDpp = true;

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
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, kStoreNT);
}
__device__ __forceinline__ void bstore(double v, rsrc_t r, int voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, kStoreNT);
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

// x neighbours through DPP wave shifts, which move data between lanes inside
// the VALU (no LDS traffic).  XM = 1: the wave holds one whole periodic row of
// 64 points, rotate (wave_rol:1 / wave_ror:1); XM = 2: the wave holds a 64-wide
// x window, shift (wave_shl:1 / wave_shr:1; an edge lane keeps 0, and edge lanes
// are halo lanes, which write nothing).  x_next: the value of lane + 1.
template <int XM>
__device__ __forceinline__ int dpp_next(int v) {
    return XM == 1 ? __builtin_amdgcn_update_dpp(0, v, 0x134, 0xf, 0xf, false)
                   : __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false);
}
template <int XM>
__device__ __forceinline__ int dpp_prev(int v) {
    return XM == 1 ? __builtin_amdgcn_update_dpp(0, v, 0x13c, 0xf, 0xf, false)
                   : __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false);
}
template <int XM>
__device__ __forceinline__ float x_next(float v) { return __int_as_float(dpp_next<XM>(__float_as_int(v))); }
template <int XM>
__device__ __forceinline__ float x_prev(float v) { return __int_as_float(dpp_prev<XM>(__float_as_int(v))); }
template <int XM>
__device__ __forceinline__ double x_next(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)dpp_next<XM>((int)(unsigned)(b & 0xffffffffll));
    const unsigned hi = (unsigned)dpp_next<XM>((int)(unsigned)((unsigned long long)b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int XM>
__device__ __forceinline__ double x_prev(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)dpp_prev<XM>((int)(unsigned)(b & 0xffffffffll));
    const unsigned hi = (unsigned)dpp_prev<XM>((int)(unsigned)((unsigned long long)b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// A lane's place in the x direction: its grid column x, the lanes holding
// x+1 .. x+3 (la, as ds_bpermute byte addresses) and x-1 .. x-4 (lb_), and
// whether it writes its points.  SEG: nx divides 64 and the wave holds 64/nx
// whole rows side by side (segment seg); otherwise window xt of width XW.
template <bool SEG>
__device__ __forceinline__ void x_place(int lane, int nx, int xt, int &x, int &seg, int (&la)[3], int (&lb_)[4],
                                        bool &xout) {
    if constexpr (SEG) {
        seg = lane / nx;
        x = lane - seg * nx;
        xout = true;
#pragma unroll
        for (int i = 0; i < 3; ++i) la[i] = (seg * nx + wrapi(x + i + 1, nx)) * 4;
#pragma unroll
        for (int i = 0; i < 4; ++i) lb_[i] = (seg * nx + wrapi(x - i - 1, nx)) * 4;
    } else {
        const int x0 = xt * XW;
        seg = 0;
        x = wrapi(x0 - HW + lane, nx);
        xout = lane >= HW && lane < WAVE - HW && x0 + lane - HW < nx;
#pragma unroll
        for (int i = 0; i < 3; ++i) la[i] = min(lane + i + 1, WAVE - 1) * 4;
#pragma unroll
        for (int i = 0; i < 4; ++i) lb_[i] = max(lane - i - 1, 0) * 4;
    }
}

// Queue slots: at step phase P (plane w = zz + P), logical slot m of pq holds
// plane w + m and logical slot m of dz holds D(w - 4 + m); physical slot is
// (m + P) & 7.
//
// The chunk start: planes zb .. zb+7 into pq, D(zb-4) .. D(zb+2) into dz.
template <typename T, int NC, int RY>
__device__ __forceinline__ void fill_queues(T (&pq)[NC][RY][8], T (&dz)[NC][RY][8], const rsrc_t (&rin)[NC],
                                            const int (&rowoff)[RY], int zb, int nz, int pbytes) {
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

// One z step at plane w (queue phase P) for a lane's RY rows: the newest z
// difference D(w + 3), then out(w) at plane byte offset so.  hy holds the 8
// y-halo rows of plane w (rows y0-4 .. y0-1 and y0+RY .. y0+RY+3) and vv the
// potential of the lane's rows.  stoff[r] is the store offset of row r, out of
// the buffer's range for points the lane does not write: every lane stores.
template <typename T, int NC, bool KT, int RY, int XM>
__device__ __forceinline__ void fd8_plane(T (&pq)[NC][RY][8], T (&dz)[NC][RY][8], const T (&hy)[NC][8],
                                          const T (&vv)[RY], const int P, const int (&la)[3], const int (&lb_)[4],
                                          const int (&stoff)[RY], const rsrc_t (&rout)[NC], const int so,
                                          const FD8Coef<T> &c) {
    // newest z difference D(w + 3) = p(w + 4) - p(w + 3)
#pragma unroll
    for (int cc = 0; cc < NC; ++cc)
#pragma unroll
        for (int r = 0; r < RY; ++r) dz[cc][r][(7 + P) & 7] = pq[cc][r][(4 + P) & 7] - pq[cc][r][(3 + P) & 7];
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
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        T lap[NC], grd[NC];
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
            const T p = pq[cc][r][P & 7];
            // x differences: own D(x), then D(x+1..x+3) and D(x-1..x-4) from lanes
            T dxp[4], dxm[4];
            if constexpr (XM == 0) {
                const T dx0 = lane_read(p, la[0]) - p;
                dxp[0] = dx0;
#pragma unroll
                for (int i = 0; i < 3; ++i) dxp[i + 1] = lane_read(dx0, la[i]);
#pragma unroll
                for (int i = 0; i < 4; ++i) dxm[i] = lane_read(dx0, lb_[i]);
            } else {
                const T dx0 = x_next<XM>(p) - p;
                dxp[0] = dx0;
#pragma unroll
                for (int i = 0; i < 3; ++i) dxp[i + 1] = x_next<XM>(dxp[i]);
                dxm[0] = x_prev<XM>(dx0);
#pragma unroll
                for (int i = 0; i < 3; ++i) dxm[i + 1] = x_prev<XM>(dxm[i]);
            }
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
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) bstore(lap[cc], rout[cc], stoff[r], so);
    }
}

// plane w + 8 into the slot plane w leaves.  The chunk's last step reads
// D(ze + 2), so planes past ze + 3 are never needed: those requests go to the
// empty resource rnull, which returns 0 without a memory access.  Requests are
// never skipped by a branch: with the same vector memory operations on every
// path, each wait for a queued plane counts exactly the operations issued
// after it (which include the stores, on this counter).
template <typename T, int NC, int RY>
__device__ __forceinline__ void refill(T (&pq)[NC][RY][8], const rsrc_t (&rin)[NC], rsrc_t rnull,
                                       const int (&rowoff)[RY], int P, int w, int ze, int nz, int pbytes) {
    const bool need = w + 8 <= ze + 3;
    int pf = w + 8;
    while (pf >= nz) pf -= nz;
    const int sf = pf * pbytes;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
        const rsrc_t rr = need ? rin[cc] : rnull;
#pragma unroll
        for (int r = 0; r < RY; ++r) bload(pq[cc][r][P & 7], rr, rowoff[r], sf);
    }
}

template <typename T, int NC, bool KT, int RY, bool SEG, int XM>
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

    int x, seg, la[3], lb_[4];
    bool xout;
    x_place<SEG>(lane, nx, xt, x, seg, la, lb_, xout);
    const int y0 = SEG ? (yt * (WAVE / nx) + seg) * RY : yt * RY;
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
    const rsrc_t rnull = make_rsrc(in, 0);

    // byte offsets inside a plane; every load is wrapped into the grid, and a
    // point the lane does not write stores past the end of the buffer
    int rowoff[RY], stoff[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        rowoff[r] = (x + nx * wrapi(y0 + r, ny)) * (int)sizeof(T);
        stoff[r] = (xout && y0 + r < ny) ? rowoff[r] : (int)(ng * (int64_t)sizeof(T));
    }
    int hoff[8];  // rows y0-4..y0-1 and y0+RY..y0+RY+3
#pragma unroll
    for (int j = 0; j < 8; ++j)
        hoff[j] = (x + nx * wrapi(j < 4 ? y0 - 4 + j : y0 + RY + (j - 4), ny)) * (int)sizeof(T);

    T pq[NC][RY][8], dz[NC][RY][8];
    fill_queues<T, NC, RY>(pq, dz, rin, rowoff, zb, nz, pbytes);
    // halo rows and V, requested kHLead and kVLead planes ahead; slot (q - zb) & 3
    // holds plane q
    T hq[NC][8][4], vq[RY][4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int so = wrapi(zb + m, nz) * pbytes;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (m <= kHLead)
                    bload(hq[cc][j][m], rin[cc], hoff[j], so);
                else
                    hq[cc][j][m] = T(0);
            }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            if (m <= kVLead)
                bload(vq[r][m], rv, rowoff[r], so);
            else
                vq[r][m] = T(0);
        }
    }

    for (int zz = zb; zz < ze; zz += 8) {
#pragma unroll
        for (int P = 0; P < 8; ++P) {
            const int w = zz + P;
            // nothing follows the z loop: a finished wave leaves here, so that
            // the skipped steps do not keep the slots they would overwrite alive
            if (w >= ze) return;  // wave-uniform
            // requests past the chunk read the empty resource (see refill)
            {
                const bool need = w + kHLead < ze;
                const int so = need ? (w + kHLead) * pbytes : 0;
#pragma unroll
                for (int cc = 0; cc < NC; ++cc) {
                    const rsrc_t rr = need ? rin[cc] : rnull;
#pragma unroll
                    for (int j = 0; j < 8; ++j) bload(hq[cc][j][(P + kHLead) & 3], rr, hoff[j], so);
                }
            }
            {
                const bool need = w + kVLead < ze;
                const int so = need ? (w + kVLead) * pbytes : 0;
                const rsrc_t rr = need ? rv : rnull;
#pragma unroll
                for (int r = 0; r < RY; ++r) bload(vq[r][(P + kVLead) & 3], rr, rowoff[r], so);
            }
            T hy[NC][8], vv[RY];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                for (int j = 0; j < 8; ++j) hy[cc][j] = hq[cc][j][P & 3];
#pragma unroll
            for (int r = 0; r < RY; ++r) vv[r] = vq[r][P & 3];
            fd8_plane<T, NC, KT, RY, XM>(pq, dz, hy, vv, P, la, lb_, stoff, rout, w * pbytes, c);
            refill<T, NC, RY>(pq, rin, rnull, rowoff, P, w, ze, nz, pbytes);
        }
    }
}

// Workgroup = nw waves stacked in y over slab sl of ty rows (rows sy0 .. sy0+ty-1,
// x window xt), band `band`, z chunk zc.  LDS tile: 2 buffers x NC components x
// (ty + 8) rows x 64 values; tile row i holds slab row i - 4 of the plane being
// computed, at column x (whole rows, nx <= 64) or lane (x windows).  The fixed
// row pitch makes every tile offset of a lane a constant from one base address.
// All waves of a workgroup share band, chunk and slab, so they run the same
// number of steps and meet at every barrier.
template <typename T, int NC, bool KT, int RY, bool SEG, int XM, int MAXT>
__global__ __launch_bounds__(MAXT) void fd8_slab_kernel(T *__restrict__ out, const T *__restrict__ in,
                                                        const T *__restrict__ vloc, int nx, int ny, int nz,
                                                        int zchunk, int nzc, int nst, int nxt, int nwg, int per_xcd,
                                                        int ty, FD8Coef<T> c) {
    extern __shared__ __align__(16) unsigned char slab_lds[];
    T *tile = reinterpret_cast<T *>(slab_lds);
    const int b = blockIdx.x;
    const int lb = (b & 7) * per_xcd + (b >> 3);
    if (lb >= nwg) return;  // the whole workgroup
    const int lane = threadIdx.x & (WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int t = lb;
    const int xt = t % nxt;
    t /= nxt;
    const int sl = t % nst;
    t /= nst;
    const int zc = t % nzc;
    const int band = t / nzc;
    const int sy0 = sl * ty;

    int x, seg, la[3], lb_[4];
    bool xout;
    x_place<SEG>(lane, nx, xt, x, seg, la, lb_, xout);
    const int lx = SEG ? x : lane;                                   // column in the tile
    const int lr0 = SEG ? (wv * (WAVE / nx) + seg) * RY : wv * RY;  // first own row in the slab
    const int y0 = sy0 + lr0;
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
    const rsrc_t rnull = make_rsrc(in, 0);

    int rowoff[RY], stoff[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        rowoff[r] = (x + nx * wrapi(y0 + r, ny)) * (int)sizeof(T);
        stoff[r] = (xout && y0 + r < ny) ? rowoff[r] : (int)(ng * (int64_t)sizeof(T));
    }

    const int csz = (ty + 8) * WAVE;        // values per component in one buffer
    const int own = (lr0 + 4) * WAVE + lx;  // tile index of the lane's first row
    // The 8 rows around the slab.  When the slab is the whole periodic y range
    // (ty == ny) they are rows of the slab itself, which their owners also
    // write to the halo rows of the tile.  Otherwise waves 0 .. 7 request one
    // row each from memory, 4 planes ahead into hq (slot (q - zb) & 3 holds
    // plane q).
    const bool wrapy = ty == ny;
    const bool hwave = !wrapy && wv < 8;  // wave-uniform
    const bool hlane = hwave && (!SEG || lane < nx);
    int hoffg = 0, hidx = 0;
    if (hlane) {
        const int h = wv;
        const int gy = wrapi(h < 4 ? sy0 - 4 + h : sy0 + ty + h - 4, ny);
        const int gx = SEG ? lane : wrapi(xt * XW - HW + lane, nx);
        hoffg = (gx + nx * gy) * (int)sizeof(T);
        hidx = (h < 4 ? h : ty + h) * WAVE + lane;
    }
    T hq[NC][4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int so = wrapi(zb + m, nz) * pbytes;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) bload(hq[cc][m], hwave ? rin[cc] : rnull, hoffg, so);
    }

    T pq[NC][RY][8], dz[NC][RY][8];
    fill_queues<T, NC, RY>(pq, dz, rin, rowoff, zb, nz, pbytes);
    T vq[RY][4];  // V, requested kVLead planes ahead; slot (q - zb) & 3 holds plane q
#pragma unroll
    for (int m = 0; m <= kVLead; ++m) {
#pragma unroll
        for (int r = 0; r < RY; ++r) bload(vq[r][m], rv, rowoff[r], wrapi(zb + m, nz) * pbytes);
    }

    for (int zz = zb; zz < ze; zz += 8) {
#pragma unroll
        for (int P = 0; P < 8; ++P) {
            const int w = zz + P;
            if (w >= ze) return;  // uniform over the workgroup
            T *buf = tile + (P & 1) * (NC * csz);
            // publish plane w: the lane's rows, and the rows around the slab
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                for (int r = 0; r < RY; ++r) {
                    const T v = pq[cc][r][P & 7];
                    buf[cc * csz + own + r * WAVE] = v;
                    if (wrapy) {
                        if (lr0 + r < 4) buf[cc * csz + own + ty * WAVE + r * WAVE] = v;
                        if (lr0 + r >= ty - 4) buf[cc * csz + own - ty * WAVE + r * WAVE] = v;
                    }
                }
            if (hlane) {
#pragma unroll
                for (int cc = 0; cc < NC; ++cc) buf[cc * csz + hidx] = hq[cc][P & 3];
            }
            // requests past the chunk, and those of waves that load no halo row,
            // read the empty resource (see refill)
            {
                const bool need = hwave && w + 4 < ze;
                const int so = need ? (w + 4) * pbytes : 0;
#pragma unroll
                for (int cc = 0; cc < NC; ++cc) bload(hq[cc][P & 3], need ? rin[cc] : rnull, hoffg, so);
            }
            {
                const bool need = w + kVLead < ze;
                const int so = need ? (w + kVLead) * pbytes : 0;
                const rsrc_t rr = need ? rv : rnull;
#pragma unroll
                for (int r = 0; r < RY; ++r) bload(vq[r][(P + kVLead) & 3], rr, rowoff[r], so);
            }
            __syncthreads();
            T hy[NC][8];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                for (int j = 0; j < 8; ++j) hy[cc][j] = buf[cc * csz + own + (j < 4 ? j - 4 : RY + j - 4) * WAVE];
            T vv[RY];
#pragma unroll
            for (int r = 0; r < RY; ++r) vv[r] = vq[r][P & 3];
            fd8_plane<T, NC, KT, RY, XM>(pq, dz, hy, vv, P, la, lb_, stoff, rout, w * pbytes, c);
            refill<T, NC, RY>(pq, rin, rnull, rowoff, P, w, ze, nz, pbytes);
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
    bool slab;   // fd8_slab_kernel (else fd8_wave_kernel)
    int ry;      // rows per lane
    int nw;      // waves per slab workgroup
    int zchunk;  // z planes per wave (a multiple of 8); 0: the whole z range
};
constexpr Candidate kCandidates[] = {
    {false, 4, 0, 32}, {false, 2, 0, 16}, {false, 2, 0, 32}, {false, 4, 0, 16},
    {false, 4, 0, 64}, {false, 8, 0, 16}, {false, 8, 0, 32}, {false, 8, 0, 64},
    {true, 4, 16, 32}, {true, 4, 16, 64}, {true, 4, 16, 0},  {true, 2, 16, 0},
    {true, 4, 8, 0},   {true, 4, 10, 0},  {true, 4, 12, 0},  {true, 2, 16, 32},
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
    if (seg && nx == WAVE && kXDpp)
        hipLaunchKernelGGL((fd8_wave_kernel<T, NC, KT, RY, true, 1>), dim3(8 * per_xcd), dim3(256), 0, st, out, in,
                           vloc, nx, ny, nz, zchunk, nzc, nyt, nxt, (int)nwaves, nblk, per_xcd, c);
    else if (seg)
        hipLaunchKernelGGL((fd8_wave_kernel<T, NC, KT, RY, true, 0>), dim3(8 * per_xcd), dim3(256), 0, st, out, in,
                           vloc, nx, ny, nz, zchunk, nzc, nyt, nxt, (int)nwaves, nblk, per_xcd, c);
    else
        hipLaunchKernelGGL((fd8_wave_kernel<T, NC, KT, RY, false, kXDpp ? 2 : 0>), dim3(8 * per_xcd), dim3(256), 0,
                           st, out, in, vloc, nx, ny, nz, zchunk, nzc, nyt, nxt, (int)nwaves, nblk, per_xcd, c);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

// Slab shape of a candidate for this grid: rows per slab and LDS bytes.
template <typename T, int NC>
void slab_shape(const qm_ctx *ctx, int ry, int nw, int *ty, int64_t *lds) {
    const bool seg = WAVE % ctx->nx == 0;
    *ty = nw * (seg ? WAVE / ctx->nx : 1) * ry;
    *lds = 2 * (int64_t)NC * (*ty + 8) * WAVE * (int64_t)sizeof(T);
}

template <typename T, int NC, bool KT, int RY, int MAXT>
int launch_slab(const qm_ctx *ctx, T *out, const T *in, const T *vloc, int nb, int nw, int zchunk,
                const FD8Coef<T> &c, hipStream_t st) {
    const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
    const bool seg = WAVE % nx == 0;
    int ty;
    int64_t lds;
    slab_shape<T, NC>(ctx, RY, nw, &ty, &lds);
    QM_CHECK(nw >= 8 && nw * WAVE <= MAXT && lds <= kSlabLdsMax, QM_ERR_ARG,
             "stencil slab shape does not fit its workgroup");
    const int nst = (ny + ty - 1) / ty;
    const int nxt = seg ? 1 : (nx + XW - 1) / XW;
    const int zc = zchunk > 0 ? zchunk : nz;
    const int nzc = (nz + zc - 1) / zc;
    const int64_t nwg = (int64_t)nb * nzc * nst * nxt;
    QM_CHECK(nwg < (int64_t)INT32_MAX / 16, QM_ERR_ARG, "grid too large for one stencil launch");
    const int per_xcd = (int)((nwg + 7) / 8);
    if (seg && nx == WAVE && kXDpp)
        hipLaunchKernelGGL((fd8_slab_kernel<T, NC, KT, RY, true, 1, MAXT>), dim3(8 * per_xcd), dim3(nw * WAVE),
                           (size_t)lds, st, out, in, vloc, nx, ny, nz, zc, nzc, nst, nxt, (int)nwg, per_xcd, ty, c);
    else if (seg)
        hipLaunchKernelGGL((fd8_slab_kernel<T, NC, KT, RY, true, 0, MAXT>), dim3(8 * per_xcd), dim3(nw * WAVE),
                           (size_t)lds, st, out, in, vloc, nx, ny, nz, zc, nzc, nst, nxt, (int)nwg, per_xcd, ty, c);
    else
        hipLaunchKernelGGL((fd8_slab_kernel<T, NC, KT, RY, false, kXDpp ? 2 : 0, MAXT>), dim3(8 * per_xcd),
                           dim3(nw * WAVE), (size_t)lds, st, out, in, vloc, nx, ny, nz, zc, nzc, nst, nxt, (int)nwg,
                           per_xcd, ty, c);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

// A candidate's shape for one element type.  The wave kernel's 8 rows per
// lane would spill registers for FP64 or complex bands, so 4 there.  Slab
// workgroups over 8 waves leave 128 registers per lane, which only FP32 real
// bands fit: FP64 real and FP32 complex slabs run 2 rows per lane (128-142
// registers, 3 waves per SIMD), 8 waves, or 12 for the 12-wave candidate
// (one 8-wave workgroup per CU leaves a third of the wave slots empty; 12
// waves fill them), and FP64 complex bands have no slab shape.
template <typename T, int NC>
Candidate effective(Candidate cd) {
    constexpr int bytes = (int)sizeof(T) * NC;
    if (!cd.slab) {
        if (cd.ry == 8 && bytes > 4) cd.ry = 4;
    } else if (bytes == 8) {
        cd.nw = cd.nw == 12 ? 12 : 8;
        cd.ry = 2;
    }
    return cd;
}

// True when candidate cand can run on this grid and element type; otherwise
// it runs as candidate 0, which gives the same values.
template <typename T, int NC>
bool feasible(const qm_ctx *ctx, int cand) {
    const Candidate cd = effective<T, NC>(kCandidates[cand]);
    if (!cd.slab) return true;
    if ((int)sizeof(T) * NC > 8) return false;
    int ty;
    int64_t lds;
    slab_shape<T, NC>(ctx, cd.ry, cd.nw, &ty, &lds);
    return lds <= kSlabLdsMax && cd.nw >= 8;
}

template <typename T, int NC, bool KT>
int launch_candidate(const qm_ctx *ctx, int cand, T *out, const T *in, const T *vloc, int nb, const FD8Coef<T> &c,
                     hipStream_t st) {
    if (!feasible<T, NC>(ctx, cand)) cand = 0;
    const Candidate cd = effective<T, NC>(kCandidates[cand]);
    if (cd.slab) {
        if constexpr (sizeof(T) * NC == 4) {
            if (cd.nw > 8) {
                if (cd.ry == 2)
                    return launch_slab<T, NC, KT, 2, 1024>(ctx, out, in, vloc, nb, cd.nw, cd.zchunk, c, st);
                return launch_slab<T, NC, KT, 4, 1024>(ctx, out, in, vloc, nb, cd.nw, cd.zchunk, c, st);
            }
            if (cd.ry == 2) return launch_slab<T, NC, KT, 2, 512>(ctx, out, in, vloc, nb, cd.nw, cd.zchunk, c, st);
            return launch_slab<T, NC, KT, 4, 512>(ctx, out, in, vloc, nb, cd.nw, cd.zchunk, c, st);
        } else if constexpr (sizeof(T) * NC == 8) {
            if (cd.nw > 8) return launch_slab<T, NC, KT, 2, 768>(ctx, out, in, vloc, nb, cd.nw, cd.zchunk, c, st);
            return launch_slab<T, NC, KT, 2, 512>(ctx, out, in, vloc, nb, cd.nw, cd.zchunk, c, st);
        }
    }
    constexpr int RY8 = ((int)sizeof(T) * NC > 4) ? 4 : 8;
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

template <typename T>
bool feasible_any(const qm_ctx *ctx, int cand, int ncomp) {
    return ncomp == 1 ? feasible<T, 1>(ctx, cand) : feasible<T, 2>(ctx, cand);
}

// Tuning key of a call shape.  Band counts share a key within each power-of-two
// bucket (33..64, 65..128, ...): the fastest shape depends on the band count
// through the launch size only, and locking, which shrinks the iterated block
// a few bands at a time, then does not re-tune at every new count.
int64_t tune_key(int precision, int ncomp, bool kt, int nb) {
    int64_t bucket = 1;
    while (bucket < nb) bucket <<= 1;
    return (((bucket * 2 + (ncomp - 1)) * 2 + (kt ? 1 : 0)) << 8) | precision;
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
        if (!feasible_any<T>(ctx, cand, ncomp)) continue;
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
    const std::string z = cd.zchunk > 0 ? std::to_string(cd.zchunk) : std::string("all");
    if (!cd.slab)
        return "fd8_wave_kernel: RY " + std::to_string(cd.ry) + " (4 for FP64 / complex when 8), z chunk " + z;
    return "fd8_slab_kernel: RY " + std::to_string(cd.ry) + ", " + std::to_string(cd.nw) +
           " waves (FP64 or complex: RY 2, " + (cd.nw == 12 ? "12" : "8") +
           " waves; FP64 complex: as candidate 0), z chunk " + z;
}

int64_t stencil_full_key(int precision, int ncomp, int kt, int nb) {
    return tune_key(precision, ncomp, kt != 0, nb) | ((int64_t)1 << 60);
}

bool stencil_feasible(const qm_ctx *ctx, int cand, int precision, int ncomp) {
    return precision == 64 ? feasible_any<double>(ctx, cand, ncomp) : feasible_any<float>(ctx, cand, ncomp);
}

int stencil_choice(const qm_ctx *ctx, int precision, int ncomp, int kt, int nb) {
    if (ctx->stencil_force >= 0) return ctx->stencil_force;
    const int64_t key = (kt & 2) ? stencil_full_key(precision, ncomp, kt & 1, nb)
                                 : tune_key(precision, ncomp, (kt & 1) != 0, nb);
    auto it = ctx->stencil_choice.find(key);
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
