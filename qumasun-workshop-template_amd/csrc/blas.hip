// Tall-skinny MFMA products and band reductions for the LOBPCG driver.
//
// gram:   C (m x n, FP64) = A^T B with A = Ng x m, B = Ng x n, K = Ng (the
//         Rayleigh-Ritz Gram blocks of MakeMatrix_z_blas, lobpcg_z_multi_mpi.h:300-558,
//         there one blas_DGEMM_t per block on the Ng x 2N real view of SoA bands).
//         One launch covers every block of [X R P]^T [R AR] (or the full
//         recompute) by chaining column blocks.  K is split into chunks across
//         workgroups; each wave accumulates with MFMA (FP32: v_mfma_f32_16x16x4f32,
//         FP64: v_mfma_f64_16x16x4f64), flushes the FP32 accumulators into FP64
//         registers every 256 K values, and the per-chunk FP64 partials are summed
//         in a fixed order by a second kernel (deterministic FP64 cross-block reduction).
// update: Y (Ng x n) = Z (Ng x m) C (m x n) with small C (NextVector_z2,
//         lobpcg_z_multi_mpi.h:1515-1598), computed as Y^T = C^T Z^T so the MFMA
//         B operand and the result are 16 consecutive grid points per lane group.
// band reductions: per-band sums over Ng (both SoA components) in FP64 for
//         eps = Re <x|Ax>, |r|^2 with r = Ax - eps x, |x|^2 (innerprod_hermite.h:12-133).
#include <algorithm>
#include <cstdint>

#include "qm_internal.h"
#include "qm_mfma.h"

namespace qm {
namespace {

template <typename T>
__device__ __forceinline__ const T *col_ptr(const ColBlocks &c, int j, int64_t ng) {
#pragma unroll
    for (int b = 0; b < kMaxColBlocks; ++b) {
        if (b >= c.nblk) break;
        if (j < c.ncols[b]) return static_cast<const T *>(c.ptr[b]) + (int64_t)j * ng;
        j -= c.ncols[b];
    }
    return nullptr;
}

template <typename T>
__device__ __forceinline__ T *col_ptr_out(const ColBlocks &c, int j, int64_t ng) {
    return const_cast<T *>(col_ptr<T>(c, j, ng));
}

constexpr int GRAM_FLUSH = 16;  // MFMA K-steps between FP32 -> FP64 flushes
constexpr int kGramTN = 4;      // 16-column MFMA tiles per wave along n (workgroup tile 64 x 32 kGramTN)

// Wave tile 32 x 16 TN (2 x TN MFMA tiles), workgroup tile 64 x 32 TN (4 waves).
// partial[chunk][j * m + i] = sum_{k in chunk} A[k][i] B[k][j]
// Hardware workgroup b runs on XCD b % 8.  Workgroups are renumbered so that
// consecutive numbers (the tiles of one K chunk) share an XCD and read the
// chunk's rows of A and B from that XCD's L2 instead of once per XCD.
__device__ __forceinline__ int xcd_block(int nblk, int per_xcd) {
    const int b = blockIdx.x;
    const int lb = (b & 7) * per_xcd + (b >> 3);
    return lb < nblk ? lb : -1;
}

template <typename T, bool VECLOAD, int TN>
__global__ __launch_bounds__(256) void gram_partial_kernel(ColBlocks A, ColBlocks B, int m, int n, int64_t ng,
                                                           int64_t kchunk, int tiles_n, int tiles, int nblk,
                                                           int per_xcd, double *__restrict__ partial) {
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    using vec_t = typename M::vec_t;
    constexpr int VEC = M::VEC;
    constexpr int KSTEP = 4 * VEC;
    const int lb = xcd_block(nblk, per_xcd);
    if (lb < 0) return;
    const int tile = lb % tiles;
    const int ti = tile / tiles_n, tj = tile % tiles_n;
    const int chunk = lb / tiles;
    const int64_t k0 = (int64_t)chunk * kchunk;
    const int64_t k1 = min(k0 + kchunk, ng);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    const int i0 = ti * 64 + (wave >> 1) * 32;
    const int j0 = tj * (32 * TN) + (wave & 1) * (16 * TN);

    const T *pa[2], *pb[TN];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int ia = i0 + 16 * s + li;
        pa[s] = ia < m ? col_ptr<T>(A, ia, ng) : nullptr;
    }
#pragma unroll
    for (int t = 0; t < TN; ++t) {
        const int jb = j0 + 16 * t + li;
        pb[t] = jb < n ? col_ptr<T>(B, jb, ng) : nullptr;
    }
    acc_t acc[2][TN];
    double acc64[2][TN][4];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < TN; ++t) {
            acc[s][t] = acc_t{0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < 4; ++r) acc64[s][t][r] = 0.0;
        }

    int steps = 0;
    int64_t k = k0;
    for (; k + KSTEP <= k1; k += KSTEP) {
        T va[2][VEC], vb[TN][VEC];
        if constexpr (VECLOAD) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                vec_t a = pa[s] ? *reinterpret_cast<const vec_t *>(pa[s] + k + VEC * g) : vec_t{};
#pragma unroll
                for (int c = 0; c < VEC; ++c) va[s][c] = a[c];
            }
#pragma unroll
            for (int t = 0; t < TN; ++t) {
                vec_t bb = pb[t] ? *reinterpret_cast<const vec_t *>(pb[t] + k + VEC * g) : vec_t{};
#pragma unroll
                for (int c = 0; c < VEC; ++c) vb[t][c] = bb[c];
            }
        } else {
#pragma unroll
            for (int c = 0; c < VEC; ++c) {
#pragma unroll
                for (int s = 0; s < 2; ++s) va[s][c] = pa[s] ? pa[s][k + VEC * g + c] : T(0);
#pragma unroll
                for (int t = 0; t < TN; ++t) vb[t][c] = pb[t] ? pb[t][k + VEC * g + c] : T(0);
            }
        }
#pragma unroll
        for (int c = 0; c < VEC; ++c)
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < TN; ++t) acc[s][t] = M::mma(va[s][c], vb[t][c], acc[s][t]);
        if constexpr (sizeof(T) == 4) {
            if (++steps == GRAM_FLUSH) {
                steps = 0;
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int t = 0; t < TN; ++t) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) acc64[s][t][r] += (double)acc[s][t][r];
                        acc[s][t] = acc_t{0, 0, 0, 0};
                    }
            }
        }
    }
    if (k < k1) {  // tail: guarded scalar loads
        T va[2][VEC], vb[TN][VEC];
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
            const int64_t kk = k + VEC * g + c;
#pragma unroll
            for (int s = 0; s < 2; ++s) va[s][c] = (pa[s] && kk < k1) ? pa[s][kk] : T(0);
#pragma unroll
            for (int t = 0; t < TN; ++t) vb[t][c] = (pb[t] && kk < k1) ? pb[t][kk] : T(0);
        }
#pragma unroll
        for (int c = 0; c < VEC; ++c)
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < TN; ++t) acc[s][t] = M::mma(va[s][c], vb[t][c], acc[s][t]);
    }
    double *out = partial + (int64_t)chunk * m * n;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < TN; ++t) {
            const int j = j0 + 16 * t + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 16 * s + M::row(g, r);
                if (i < m && j < n) out[(int64_t)j * m + i] = acc64[s][t][r] + (double)acc[s][t][r];
            }
        }
}

// Large Gram products: workgroup tile 128 x 128 of C = A^T B over a K range,
// panels of KT rows of the 256 columns staged through LDS (double-buffered),
// WGM x WGN waves, each (128 / WGM) x (128 / WGN) outputs in 16 x 16 MFMA
// tiles.  FP32 accumulators are flushed into FP64 registers every 256 K
// values (FP32) ; FP64 accumulates in FP64.
constexpr int GT_M = 128, GT_N = 128;
constexpr int GT_FLUSH_K = 256;  // FP32 K values between FP32 -> FP64 flushes

template <typename T, int WGM, int WGN, int KTV>
struct GramTile {
    static constexpr int WM = WGM, WN = WGN;             // waves along m and n
    static constexpr int THREADS = 64 * WGM * WGN;
    static constexpr int KT = KTV;                       // K rows per panel
    static constexpr int VEC = 16 / sizeof(T);           // values per 16-byte load
    static constexpr int PAD = VEC;                      // LDS row padding (keeps 16-byte alignment)
    static constexpr int LD = KT + PAD;                  // LDS stride of one column
    static constexpr int LOADS = (GT_M + GT_N) * KT / VEC / THREADS;  // 16-byte loads per thread per panel
    static constexpr int TA = GT_M / WGM / 16;           // MFMA tiles per wave along m
    static constexpr int TB = GT_N / WGN / 16;           // MFMA tiles per wave along n
    static constexpr int FLUSH = GT_FLUSH_K / KT;        // panels between flushes (FP32)
    static_assert(LOADS * VEC * THREADS == (GT_M + GT_N) * KT, "whole 16-byte loads per panel");
    static_assert(GT_FLUSH_K % KT == 0, "flushes on panel boundaries");
};
// the shapes in use: FP32 16 waves (4 x 4, 32 x 32 per wave: 88 VGPRs, 4
// waves per SIMD), 32 K rows per panel; FP64 8 waves (2 x 4, 64 x 32 per wave:
// 126 VGPRs, two workgroups per CU), 16 K rows per panel (128 B per column).
// On the 1536 x 1024 x 884,736 Gram the FP32 4 x 4 shape ran 23.8 ms against
// 24.3 for 2 x 4 (same chunks), FP64 4 x 4 47.8 against 43.7.  Both give the
// same bits: every output is the same MFMA chain over K.
template <typename T>
using GramShape = GramTile<T, sizeof(T) == 4 ? 4 : 2, 4, sizeof(T) == 4 ? 32 : 16>;

// partial[chunk][j * m + i] = sum_{k in chunk} A[k][i] B[k][j] for the tile's
// 128 x 128 outputs.  Requires ng % KT == 0 and kchunk % KT == 0 (16-byte
// aligned columns and whole panels).
template <typename T, typename P>
__global__ __launch_bounds__(P::THREADS) void gram_tile_kernel(ColBlocks A, ColBlocks B, int m, int n, int64_t ng,
                                                               int64_t kchunk, int tiles_n, int tiles, int nblk,
                                                               int per_xcd, double *__restrict__ partial) {
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    using vec_t = typename M::vec_t;
    constexpr int KT = P::KT, VEC = P::VEC, LD = P::LD, TA = P::TA, TB = P::TB, THREADS = P::THREADS;
    __shared__ __align__(16) T lds[2][GT_M + GT_N][LD];
    const int lb = xcd_block(nblk, per_xcd);
    if (lb < 0) return;
    // tj fastest: the tiles of one row of tiles (sharing the A panel) are consecutive on an XCD
    const int tile = lb % tiles;
    const int ti = tile / tiles_n, tj = tile % tiles_n;
    const int chunk = lb / tiles;
    const int64_t k0 = (int64_t)chunk * kchunk;
    const int64_t k1 = min(k0 + kchunk, ng);
    const int npanel = (int)((k1 - k0) / KT);

    // loader: thread t loads 16-byte pieces (VEC values) of columns; KT / VEC
    // consecutive threads cover one column's panel rows
    constexpr int PER_COL = KT / VEC;
    const T *src[P::LOADS];
    int dcol[P::LOADS];
    const int piece = threadIdx.x % PER_COL;
#pragma unroll
    for (int q = 0; q < P::LOADS; ++q) {
        const int c = threadIdx.x / PER_COL + q * (THREADS / PER_COL);  // 0 .. 255
        dcol[q] = c;
        const T *p = nullptr;
        if (c < GT_M) {
            const int i = ti * GT_M + c;
            if (i < m) p = col_ptr<T>(A, i, ng);
        } else {
            const int j = tj * GT_N + (c - GT_M);
            if (j < n) p = col_ptr<T>(B, j, ng);
        }
        src[q] = p ? p + k0 + piece * VEC : nullptr;
    }
    vec_t stage[P::LOADS];
    auto fetch = [&](int panel) {
#pragma unroll
        for (int q = 0; q < P::LOADS; ++q)
            stage[q] = src[q] ? *reinterpret_cast<const vec_t *>(src[q] + (int64_t)panel * KT) : vec_t{};
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < P::LOADS; ++q) *reinterpret_cast<vec_t *>(&lds[buf][dcol[q]][piece * VEC]) = stage[q];
    };

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    const int wm = (wave / P::WN) * 16 * TA;  // wave's first output row in the tile
    const int wn = (wave % P::WN) * 16 * TB;  // wave's first output column in the tile
    acc_t acc[TA][TB];
    double acc64[TA][TB][4];
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
        for (int b = 0; b < TB; ++b) {
            acc[a][b] = acc_t{0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < 4; ++r) acc64[a][b][r] = 0.0;
        }

    if (npanel > 0) {
        fetch(0);
        store(0);
    }
    __syncthreads();
    for (int p = 0; p < npanel; ++p) {
        const int cur = p & 1;
        if (p + 1 < npanel) fetch(p + 1);  // next panel's loads in flight during this panel's MFMAs
        // K steps of 4 VEC values: lane group g holds k = 4g .. (VEC values per lane)
#pragma unroll
        for (int ks = 0; ks < KT; ks += 4 * VEC) {
            vec_t av[TA], bv[TB];
#pragma unroll
            for (int a = 0; a < TA; ++a)
                av[a] = *reinterpret_cast<const vec_t *>(&lds[cur][wm + 16 * a + li][ks + VEC * g]);
#pragma unroll
            for (int b = 0; b < TB; ++b)
                bv[b] = *reinterpret_cast<const vec_t *>(&lds[cur][GT_M + wn + 16 * b + li][ks + VEC * g]);
#pragma unroll
            for (int c = 0; c < VEC; ++c)
#pragma unroll
                for (int a = 0; a < TA; ++a)
#pragma unroll
                    for (int b = 0; b < TB; ++b) acc[a][b] = M::mma(av[a][c], bv[b][c], acc[a][b]);
        }
        if constexpr (sizeof(T) == 4) {
            if ((p + 1) % P::FLUSH == 0) {
#pragma unroll
                for (int a = 0; a < TA; ++a)
#pragma unroll
                    for (int b = 0; b < TB; ++b) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) acc64[a][b][r] += (double)acc[a][b][r];
                        acc[a][b] = acc_t{0, 0, 0, 0};
                    }
            }
        }
        if (p + 1 < npanel) store(cur ^ 1);  // the other buffer: last read one panel ago, before the barrier
        __syncthreads();
    }
    double *out = partial + (int64_t)chunk * m * n;
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
        for (int b = 0; b < TB; ++b) {
            const int j = tj * GT_N + wn + 16 * b + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = ti * GT_M + wm + 16 * a + M::row(g, r);
                if (i < m && j < n) out[(int64_t)j * m + i] = acc64[a][b][r] + (double)acc[a][b][r];
            }
        }
}

__global__ void gram_reduce_kernel(double *__restrict__ C, const double *__restrict__ partial, int64_t mn,
                                   int nchunk) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= mn) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += partial[(int64_t)c * mn + e];
    C[e] = s;
}

// Y^T (n x Ng) = C^T (n x m) Z^T (m x Ng).  Workgroup: 64 output columns x 128
// grid points; each wave 64 columns x 32 points (4 x 2 MFMA tiles).  C tiles of
// UPD_KC K rows x 64 columns are staged in LDS (16 measured 4 % faster than 32 or
// 64 on the FP32 384 -> 256 column update, FP64 equal).
constexpr int UPD_KC = 16;
template <typename T>
__global__ __launch_bounds__(256) void update_kernel(ColBlocks Z, ColBlocks Y, const T *__restrict__ Cm, int m,
                                                     int n, int64_t ng, int tiles_n, int nblk, int per_xcd) {
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    __shared__ T cs[UPD_KC][64];
    const int lb = xcd_block(nblk, per_xcd);
    if (lb < 0) return;
    const int tj = lb % tiles_n;   // output-column tile (fastest: neighbours share Z rows in one XCD's L2)
    const int64_t gt = lb / tiles_n;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    const int c0 = tj * 64;
    const int64_t p0 = gt * 128 + wave * 32;
    acc_t acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    const bool pin[2] = {p0 + li < ng, p0 + 16 + li < ng};

    for (int kc = 0; kc < m; kc += UPD_KC) {
        __syncthreads();
        for (int e = threadIdx.x; e < UPD_KC * 64; e += 256) {
            const int kk = e % UPD_KC, cc = e / UPD_KC;  // consecutive threads: consecutive K (column-major C)
            cs[kk][cc] = (kc + kk < m && c0 + cc < n) ? Cm[(int64_t)(c0 + cc) * m + kc + kk] : T(0);
        }
        __syncthreads();
        // K steps unrolled twice for FP32, fully for FP64 (each measured fastest:
        // full unrolling is 10 % slower in FP32 and 6 % faster in FP64)
        constexpr int kUnroll = sizeof(T) == 8 ? UPD_KC / 4 : 2;
#pragma unroll kUnroll
        for (int ks = 0; ks < UPD_KC; ks += 4) {
            const int kcol = kc + ks + g;
            const T *zc = kcol < m ? col_ptr<T>(Z, kcol, ng) : nullptr;
            T bv[2];
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = (zc && pin[b]) ? zc[p0 + 16 * b + li] : T(0);
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const T av = cs[ks + g][16 * a + li];
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = M::mma(av, bv[b], acc[a][b]);
            }
        }
    }
    // D[row = out column 16a + row(g, r)][col = grid point 16b + li]
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int col = c0 + 16 * a + M::row(g, r);
            if (col >= n) continue;
            T *yc = col_ptr_out<T>(Y, col, ng);
#pragma unroll
            for (int b = 0; b < 2; ++b)
                if (pin[b]) yc[p0 + 16 * b + li] = acc[a][b][r];
        }
}

// Large updates Y (Ng x n) = Z (Ng x m) C (m x n): workgroup tile of 128 output
// columns x 128 grid points, 8 waves in a 2 x 4 grid (64 columns x 32 points
// each, 4 x 2 MFMA tiles), K (= m) in panels of UT_KT rows: the panel's Z rows
// (UT_KT columns of Z x 128 points) and C rows (UT_KT x 128) staged through LDS
// (double-buffered), so each Z value is read from memory once per 128 output
// columns instead of once per 64.  As the untiled kernel: Y^T = C^T Z^T, the
// MFMA B operand and the result run along 16 consecutive grid points per lane
// group.
constexpr int UT_N = 128, UT_P = 128, UT_THREADS = 512;

template <typename T, int KTV, int PADV>
struct UpdTile {
    static constexpr int KT = KTV;                                   // K rows (Z columns) per panel
    static constexpr int VEC = 16 / sizeof(T);
    static constexpr int ZLOADS = KT * UT_P / VEC / UT_THREADS;  // 16-byte Z loads per thread per panel
    static constexpr int CLOADS = KT * UT_N / VEC / UT_THREADS;  // 16-byte C loads per thread per panel
    static constexpr int PAD = PADV;                             // LDS row padding (values)
    static_assert(ZLOADS * VEC * UT_THREADS == KT * UT_P && CLOADS * VEC * UT_THREADS == KT * UT_N,
                  "whole 16-byte loads per panel");
};
// the shape in use: 16 (FP32) / 8 (FP64) K rows per panel, rows padded by 4.
// Shallow panels leave room for more resident workgroups (FP32: 64 VGPRs,
// 34 KB of LDS): the 1536 -> 1024 column update over 884,736 points ran
// 21.4 ms against 23.2 with 32 K rows (FP64 44.6 against 45.1 with 16), the
// 384 -> 256 one over 262,144 points 0.49 against 0.57 ms (FP64 0.93 against
// 1.02); 64 K rows and wider padding were slower.  Same bits for every depth.
template <typename T>
using UpdShape = UpdTile<T, sizeof(T) == 4 ? 16 : 8, 4>;

// Requires ng % UT_P == 0 and m % KT == 0 (whole tiles and panels; the caller
// checks) and 16-byte aligned column blocks.
template <typename T, typename P>
__global__ __launch_bounds__(UT_THREADS) void update_tile_kernel(ColBlocks Z, ColBlocks Y, const T *__restrict__ Cm,
                                                                 int m, int n, int64_t ng, int tiles_n, int nblk,
                                                                 int per_xcd) {
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    using vec_t = typename M::vec_t;
    constexpr int KT = P::KT, VEC = P::VEC;
    __shared__ __align__(16) T zs[2][KT][UT_P + P::PAD];
    __shared__ __align__(16) T cs[2][KT][UT_N + P::PAD];
    const int lb = xcd_block(nblk, per_xcd);
    if (lb < 0) return;
    const int tj = lb % tiles_n;  // output-column tile fastest: neighbours share Z panels in one XCD's L2
    const int64_t gt = lb / tiles_n;
    const int c0 = tj * UT_N;
    const int64_t p0 = gt * UT_P;
    const int npanel = m / KT;

    // Z loader: UT_P / VEC consecutive threads cover one Z column's 128 points
    constexpr int ZPER = UT_P / VEC;
    const int zpiece = threadIdx.x % ZPER;
    // C loader: KT / VEC consecutive threads cover KT consecutive rows of one C column
    constexpr int CPER = KT / VEC;
    const int cpiece = threadIdx.x % CPER;
    vec_t zst[P::ZLOADS], cst[P::CLOADS];
    auto fetch = [&](int panel) {
        const int k0 = panel * KT;
#pragma unroll
        for (int q = 0; q < P::ZLOADS; ++q) {
            const int kk = threadIdx.x / ZPER + q * (UT_THREADS / ZPER);
            const T *zc = col_ptr<T>(Z, k0 + kk, ng);
            zst[q] = *reinterpret_cast<const vec_t *>(zc + p0 + zpiece * VEC);
        }
#pragma unroll
        for (int q = 0; q < P::CLOADS; ++q) {
            const int cc = threadIdx.x / CPER + q * (UT_THREADS / CPER);
            cst[q] = (c0 + cc < n) ? *reinterpret_cast<const vec_t *>(Cm + (int64_t)(c0 + cc) * m + k0 + cpiece * VEC)
                                   : vec_t{};
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < P::ZLOADS; ++q) {
            const int kk = threadIdx.x / ZPER + q * (UT_THREADS / ZPER);
            *reinterpret_cast<vec_t *>(&zs[buf][kk][zpiece * VEC]) = zst[q];
        }
#pragma unroll
        for (int q = 0; q < P::CLOADS; ++q) {
            const int cc = threadIdx.x / CPER + q * (UT_THREADS / CPER);
#pragma unroll
            for (int v = 0; v < VEC; ++v) cs[buf][cpiece * VEC + v][cc] = cst[q][v];
        }
    };

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    const int wc = (wave >> 2) * 64;  // wave's first output column in the tile
    const int wp = (wave & 3) * 32;   // wave's first grid point in the tile
    acc_t acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

    fetch(0);
    store(0);
    __syncthreads();
    for (int p = 0; p < npanel; ++p) {
        const int cur = p & 1;
        if (p + 1 < npanel) fetch(p + 1);
#pragma unroll
        for (int ks = 0; ks < KT; ks += 4) {
            T av[4], bv[2];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = cs[cur][ks + g][wc + 16 * a + li];  // A[col][k = g]
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = zs[cur][ks + g][wp + 16 * b + li];  // B[k = g][point]
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = M::mma(av[a], bv[b], acc[a][b]);
        }
        if (p + 1 < npanel) store(cur ^ 1);
        __syncthreads();
    }
    // D[row = out column 16a + row(g, r)][col = grid point 16b + li]
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int col = c0 + wc + 16 * a + M::row(g, r);
            if (col >= n) continue;
            T *yc = col_ptr_out<T>(Y, col, ng);
#pragma unroll
            for (int b = 0; b < 2; ++b) yc[p0 + wp + 16 * b + li] = acc[a][b][r];
        }
}

// Per-band FP64 partial sums over a chunk of the grid.  mode 0: sum x*ax;
// mode 1: r = ax - lam x (written), sums |r|^2 and |x|^2; mode 2: sum x*x.
constexpr int RED_THREADS = 256;
template <typename T, int MODE>
__global__ __launch_bounds__(RED_THREADS) void band_reduce_kernel(double *__restrict__ part, T *__restrict__ r,
                                                                   const T *__restrict__ x, const T *__restrict__ ax,
                                                                   const double *__restrict__ lam, int64_t ng, int ncomp,
                                                                   int64_t chunk, int nchunk) {
    __shared__ double sm[2][RED_THREADS / 64];
    const int n = blockIdx.y;
    const int64_t b0 = (int64_t)blockIdx.x * chunk;
    const int64_t b1 = min(b0 + chunk, ng);
    double s0 = 0.0, s1 = 0.0;
    const T l = MODE == 1 ? (T)lam[n] : T(0);
    for (int c = 0; c < ncomp; ++c) {
        const int64_t off = ((int64_t)n * ncomp + c) * ng;
        for (int64_t i = b0 + threadIdx.x; i < b1; i += RED_THREADS) {
            const T xv = x[off + i];
            if constexpr (MODE == 0) {
                s0 += (double)xv * (double)ax[off + i];
            } else if constexpr (MODE == 1) {
                const T rv = ax[off + i] - l * xv;
                r[off + i] = rv;
                s0 += (double)rv * (double)rv;
                s1 += (double)xv * (double)xv;
            } else {
                s0 += (double)xv * (double)xv;
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off, 64);
        s1 += __shfl_xor(s1, off, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sm[0][wave] = s0;
        sm[1][wave] = s1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0.0, t1 = 0.0;
        for (int w = 0; w < RED_THREADS / 64; ++w) {
            t0 += sm[0][w];
            t1 += sm[1][w];
        }
        part[((int64_t)n * nchunk + blockIdx.x) * 2] = t0;
        part[((int64_t)n * nchunk + blockIdx.x) * 2 + 1] = t1;
    }
}

__global__ void band_reduce_finish(double *__restrict__ out0, double *__restrict__ out1, const double *__restrict__ part,
                                   int nb, int nchunk) {
    const int n = blockIdx.x * 64 + threadIdx.x;
    if (n >= nb) return;
    double t0 = 0.0, t1 = 0.0;
    for (int c = 0; c < nchunk; ++c) {
        t0 += part[((int64_t)n * nchunk + c) * 2];
        t1 += part[((int64_t)n * nchunk + c) * 2 + 1];
    }
    if (out0) out0[n] = t0;
    if (out1) out1[n] = t1;
}

template <typename T>
__global__ void scale_bands_kernel(T *__restrict__ y, const T *__restrict__ x, const double *__restrict__ f,
                                   double fconst, int64_t ng, int ncomp) {
    const int n = blockIdx.y;
    const double s = f ? f[n] : fconst;
    for (int c = 0; c < ncomp; ++c) {
        const int64_t off = ((int64_t)n * ncomp + c) * ng;
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < ng; i += (int64_t)gridDim.x * 256)
            y[off + i] = (T)(s * (double)x[off + i]);
    }
}

int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

}  // namespace

// Gram products with both sides >= 128 columns over whole panels take the
// LDS-tiled 128 x 128 kernel: each panel row is read once per workgroup tile
// and shared by its 8 waves, and the tiles of one tile row (the same A panel)
// run consecutively on one XCD.  K is split so that about 1024 workgroups run.
// K rows per chunk are a multiple of this (whole flush intervals for FP32)
template <typename T>
constexpr int gram_kround() { return sizeof(T) == 4 ? GT_FLUSH_K : 128; }

// K chunks of the tiled Gram, from a cost model fitted to measured launches:
//   time = per-CU rounds x chunk length x (per-workgroup time per K row)
//        + the partials' write and reduction (16 B per element and chunk, ~2.6 TB/s).
// Each CU takes b = ceil(blocks / CUs) workgroups, per_cu of them at a time; a
// last, partly filled round costs its share plus 14 % (a lone FP64 workgroup
// hides less latency than a pair: measured 0.57 of a pair's time).  A
// workgroup-round advances one K row of a 128 x 128 tile (32,768 flops) at ~0.8
// of the CU's MFMA peak.  The cheapest chunk count is taken, the smallest one
// on ties, with at most 256 chunks and 512 MB of partials.  Measured on the
// LOBPCG shapes: 384 x 256 over 64^3 points, FP32 128 chunks (0.52 ms; the
// previous 147: 0.66 ms), FP64 128 (0.87 ms; 158: 1.01 ms); 1536 x 1024 over
// 96^3, FP32 8 (23.8 ms; 10: 25.3), FP64 16 (40.8 ms; 10: 43.7).
inline int64_t gram_chunks(int tiles, int64_t ng, int64_t kround, int cus, int per_cu, double cu_flops_us,
                           int64_t mn, int64_t *kchunk_out) {
    const double t_k = per_cu * 32768.0 / (0.8 * cu_flops_us);  // us per K row of a full round
    int64_t best_nc = 1, best_k = round_up(ng, kround);
    double best_cost = 1e300;
    for (int64_t nc = 1; nc <= 256; ++nc) {
        const int64_t kchunk = round_up((ng + nc - 1) / nc, kround);
        const int64_t act = (ng + kchunk - 1) / kchunk;
        if (act < nc) continue;  // the same partition as a smaller count
        if (act > 1 && act * mn * (int64_t)sizeof(double) > ((int64_t)512 << 20)) break;
        const int64_t b = (tiles * act + cus - 1) / cus;
        const double rounds = (double)(b / per_cu) + std::min(1.0, 1.14 * (double)(b % per_cu) / per_cu);
        const double cost = rounds * (double)kchunk * t_k + (double)act * mn * 16.0 / 2.6e6;
        if (cost < best_cost) {
            best_cost = cost;
            best_nc = act;
            best_k = kchunk;
        }
    }
    *kchunk_out = best_k;
    return best_nc;
}

template <typename T, typename P = GramShape<T>>
int gram_tiled(DevBuf &partial, double *C, const ColBlocks &A, const ColBlocks &B, int m, int n, int64_t ng,
               hipStream_t st) {
    static_assert(gram_kround<T>() % P::KT == 0, "chunks of whole panels");
    const int tiles_m = (m + GT_M - 1) / GT_M, tiles_n = (n + GT_N - 1) / GT_N;
    const int tiles = tiles_m * tiles_n;
    int dev = 0, cus = 0, per_cu = 0;
    QM_HIP(hipGetDevice(&dev));
    QM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    QM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, gram_tile_kernel<T, P>, P::THREADS, 0));
    // dense MFMA peak per CU and microsecond: FP32 157.3, FP64 78.6 TFLOP/s over 256 CUs
    const double cu_flops_us = (sizeof(T) == 4 ? 157.3e6 : 78.6e6) / 256.0;
    int64_t kchunk = 0;
    const int64_t nchunk = gram_chunks(tiles, ng, gram_kround<T>(), std::max(cus, 1), std::max(per_cu, 1),
                                       cu_flops_us, (int64_t)m * n, &kchunk);
    QM_TRY(partial.reserve(sizeof(double) * (size_t)nchunk * m * n));
    const int64_t nblk = (int64_t)tiles * nchunk;
    QM_CHECK(nblk < ((int64_t)1 << 30), QM_ERR_ARG, "gram: grid too large");
    const int per_xcd = (int)((nblk + 7) / 8);
    hipLaunchKernelGGL((gram_tile_kernel<T, P>), dim3((unsigned)(8 * per_xcd)), dim3(P::THREADS), 0, st, A, B, m, n,
                       ng, kchunk, tiles_n, tiles, (int)nblk, per_xcd, partial.as<double>());
    QM_HIP(hipGetLastError());
    const int64_t mn = (int64_t)m * n;
    hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, st, C,
                       partial.as<double>(), mn, (int)nchunk);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

// every column block starts on a 16-byte boundary (vector loads of whole columns)
bool aligned16(const ColBlocks &c) {
    for (int b = 0; b < c.nblk; ++b)
        if (reinterpret_cast<uintptr_t>(c.ptr[b]) % 16 != 0) return false;
    return true;
}

template <typename T>
int gram(DevBuf &partial, double *C, const ColBlocks &A, const ColBlocks &B, int m, int n, int64_t ng,
         hipStream_t st) {
    const bool al = aligned16(A) && aligned16(B);
    if (al && m >= GT_M && n >= GT_N && ng % GramShape<T>::KT == 0)
        return gram_tiled<T>(partial, C, A, B, m, n, ng, st);
    constexpr int KSTEP = 4 * Mfma<T>::VEC;
    constexpr int TN = kGramTN;  // 16-column MFMA tiles per wave along n
    const int tiles_m = (m + 63) / 64, tiles_n = (n + 32 * TN - 1) / (32 * TN);
    const int tiles = tiles_m * tiles_n;
    // enough workgroups to fill the 256 CUs a few times (up to 128 K chunks: the
    // 384 x 256 LOBPCG Gram over 64^3 points has 12 tiles, and 128 chunks instead
    // of 64 measured 969 -> 830 us FP32, 2311 -> 1800 us FP64); K chunks of >= 1024
    int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(128, 2048 / tiles));
    int64_t kchunk = round_up((ng + nchunk - 1) / nchunk, (int64_t)KSTEP * GRAM_FLUSH);
    kchunk = std::max<int64_t>(kchunk, 1024);
    nchunk = (ng + kchunk - 1) / kchunk;
    QM_TRY(partial.reserve(sizeof(double) * (size_t)nchunk * m * n));
    const bool vec = al && (ng % Mfma<T>::VEC) == 0;
    const int64_t nblk = (int64_t)tiles * nchunk;
    QM_CHECK(nblk < ((int64_t)1 << 30), QM_ERR_ARG, "gram: grid too large");
    const int per_xcd = (int)((nblk + 7) / 8);
    const dim3 grid((unsigned)(8 * per_xcd));
    if (vec)
        hipLaunchKernelGGL((gram_partial_kernel<T, true, TN>), grid, dim3(256), 0, st, A, B, m, n, ng, kchunk, tiles_n,
                           tiles, (int)nblk, per_xcd, partial.as<double>());
    else
        hipLaunchKernelGGL((gram_partial_kernel<T, false, TN>), grid, dim3(256), 0, st, A, B, m, n, ng, kchunk, tiles_n,
                           tiles, (int)nblk, per_xcd, partial.as<double>());
    QM_HIP(hipGetLastError());
    const int64_t mn = (int64_t)m * n;
    hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, st, C,
                       partial.as<double>(), mn, (int)nchunk);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

template <typename T, typename P = UpdShape<T>>
int update_tiled(const ColBlocks &Z, const ColBlocks &Y, const T *Cm, int m, int n, int64_t ng, hipStream_t st) {
    const int tiles_n = (n + UT_N - 1) / UT_N;
    const int64_t nblk = (ng / UT_P) * tiles_n;
    QM_CHECK(nblk < ((int64_t)1 << 30), QM_ERR_ARG, "update: grid too large");
    const int per_xcd = (int)((nblk + 7) / 8);
    hipLaunchKernelGGL((update_tile_kernel<T, P>), dim3((unsigned)(8 * per_xcd)), dim3(UT_THREADS), 0, st, Z, Y, Cm, m,
                       n, ng, tiles_n, (int)nblk, per_xcd);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

template <typename T>
int update(const ColBlocks &Z, const ColBlocks &Y, const T *Cm, int m, int n, int64_t ng, hipStream_t st) {
    if (n >= UT_N && ng % UT_P == 0 && m % UpdShape<T>::KT == 0 && aligned16(Z) && aligned16(Y) &&
        reinterpret_cast<uintptr_t>(Cm) % 16 == 0)
        return update_tiled<T>(Z, Y, Cm, m, n, ng, st);
    const int tiles_n = (n + 63) / 64;
    const int64_t gtiles = (ng + 127) / 128;
    const int64_t nblk = gtiles * tiles_n;
    QM_CHECK(nblk < ((int64_t)1 << 30), QM_ERR_ARG, "update: grid too large");
    const int per_xcd = (int)((nblk + 7) / 8);
    hipLaunchKernelGGL((update_kernel<T>), dim3((unsigned)(8 * per_xcd)), dim3(256), 0, st, Z, Y, Cm, m, n, ng,
                       tiles_n, (int)nblk, per_xcd);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

template <typename T>
int band_reduce(DevBuf &part, int mode, double *out0, double *out1, T *r, const T *x, const T *ax, const double *lam,
                int nb, int ncomp, int64_t ng, hipStream_t st) {
    const int64_t chunk = std::max<int64_t>(4096, round_up((ng + 63) / 64, 256));
    const int nchunk = (int)((ng + chunk - 1) / chunk);
    QM_CHECK(nb <= 65535, QM_ERR_ARG, "band_reduce: too many bands");
    QM_TRY(part.reserve(sizeof(double) * 2 * (size_t)nb * nchunk));
    dim3 grid(nchunk, nb);
    if (mode == 0)
        hipLaunchKernelGGL((band_reduce_kernel<T, 0>), grid, dim3(RED_THREADS), 0, st, part.as<double>(), r, x, ax, lam,
                           ng, ncomp, chunk, nchunk);
    else if (mode == 1)
        hipLaunchKernelGGL((band_reduce_kernel<T, 1>), grid, dim3(RED_THREADS), 0, st, part.as<double>(), r, x, ax, lam,
                           ng, ncomp, chunk, nchunk);
    else
        hipLaunchKernelGGL((band_reduce_kernel<T, 2>), grid, dim3(RED_THREADS), 0, st, part.as<double>(), r, x, ax, lam,
                           ng, ncomp, chunk, nchunk);
    QM_HIP(hipGetLastError());
    hipLaunchKernelGGL(band_reduce_finish, dim3((nb + 63) / 64), dim3(64), 0, st, out0, out1, part.as<double>(), nb,
                       nchunk);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

template <typename T>
int scale_bands(T *y, const T *x, const double *f, double fconst, int nb, int ncomp, int64_t ng, hipStream_t st) {
    const int64_t bx = std::min<int64_t>((ng + 255) / 256, 1024);
    hipLaunchKernelGGL((scale_bands_kernel<T>), dim3((unsigned)bx, nb), dim3(256), 0, st, y, x, f, fconst, ng, ncomp);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

template int gram<float>(DevBuf &, double *, const ColBlocks &, const ColBlocks &, int, int, int64_t, hipStream_t);
template int gram<double>(DevBuf &, double *, const ColBlocks &, const ColBlocks &, int, int, int64_t, hipStream_t);
template int update<float>(const ColBlocks &, const ColBlocks &, const float *, int, int, int64_t, hipStream_t);
template int update<double>(const ColBlocks &, const ColBlocks &, const double *, int, int, int64_t, hipStream_t);
template int band_reduce<float>(DevBuf &, int, double *, double *, float *, const float *, const float *,
                                const double *, int, int, int64_t, hipStream_t);
template int band_reduce<double>(DevBuf &, int, double *, double *, double *, const double *, const double *,
                                 const double *, int, int, int64_t, hipStream_t);
template int scale_bands<float>(float *, const float *, const double *, double, int, int, int64_t, hipStream_t);
template int scale_bands<double>(double *, const double *, const double *, double, int, int, int64_t, hipStream_t);

}  // namespace qm
