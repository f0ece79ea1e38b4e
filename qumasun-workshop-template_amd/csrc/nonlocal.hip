// Nonlocal Kleinman-Bylander apply and its position derivative for gfx950.
//
//   Hpsi += sum_a e^{-ik.r} P_a diag(E_a dV) P_a^T e^{+ik.r} psi
//
// Mathematics: ProjectionPP_bundle_1 / _2 (PseudoPotNonlocal_SvF2_kpoint.h:3910-4236)
// with mInnerNonlocalPsi_v4/_v6_bundle (4463-4758), the Bloch phases of
// mCreateBlochExp (3537-3618) and the force sums of ForceNonlocal_bundle
// (3629-3731, inner products of mInnerNonlocalPsi_diff_bundle 4775-4880).
//
// GPU structure (replaces per-atom DGEMM_t / DGEMM_n calls, host Bloch loops,
// host-side MPI reductions and one AddSubgrid launch per atom colour):
//   1. project: one launch over (voxel chunk, 16-column band tile) work items
//      of every atom, so spheres of any size balance by construction.  Each
//      wave gathers 256 voxels x 16 SoA columns straight from the grid, applies
//      e^{+ik.r}, and multiplies by the voxel-major projector table on MFMA
//      (16 projectors x 16 columns, K = voxels); the four waves of a chunk and
//      then the chunks of an atom are summed in FP64 in a fixed order, and
//      E_w dV is folded in (coef).
//   2. scatter: owner-computes over the grid points covered by any sphere.
//      Each covered point sums the contributions of all atoms whose sphere
//      contains it for a tile of columns, so overlapping spheres need no
//      atomics or colouring, tables are read once per tile, and the result is
//      deterministic.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "qm_internal.h"
#include "qm_mfma.h"

namespace qm {
namespace {

constexpr int NL_THREADS = 256;

// Workgroup numbering: hardware block b runs on XCD b % 8, so giving each XCD a
// contiguous range of logical blocks puts consecutive logical blocks (the
// column groups of one item or point block) on one XCD, where they share its
// L2 for the projector table rows and entry lists.  Returns -1 past the end.
__device__ __forceinline__ int xcd_block(int nblk, int per_xcd) {
    const int b = blockIdx.x;
    const int lb = (b & 7) * per_xcd + (b >> 3);
    return lb < nblk ? lb : -1;
}

template <typename T>
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// part[(item * 16 MT + w) * ncol + col] = <p_w | e^{ik.r} psi>_col summed over the
// item's voxels (col = real SoA column n * NC + c), for the GT 16-column tiles
// of column group g (logical block = item * groups + g) and the MT 16-row
// projector tiles of the table (rows padded to 16 MT).  Each wave takes kProjChunk / 4 voxels in
// sub-chunks of 64.  Per sub-chunk it reads its block of the voxel-major
// projector table once, as the MFMA A operand of all GT tiles; per tile the
// 64 x 16 block of (phase-rotated) psi values is gathered with lanes along
// voxels (coalesced along x runs) into the wave's own LDS stage and fed to
// MFMA as the B operand.  The stage is private to the wave, so no workgroup
// barrier is needed until the final reduction over the four waves.
constexpr int PJ_VPW = kProjChunk / 4;  // voxels per wave
constexpr int PJ_SUB = 64;              // voxels per LDS sub-chunk
template <typename T, int NC, bool PH, int GT, int MT>
__global__ __launch_bounds__(256) void nl_project_kernel(double *__restrict__ part, const T *__restrict__ psi,
                                                         int64_t ng, int ncol, const int *__restrict__ chunk_atom,
                                                         const int *__restrict__ chunk_v0,
                                                         const int64_t *__restrict__ vox_off,
                                                         const int *__restrict__ vox_grid, const T *__restrict__ tabp,
                                                         const T *__restrict__ phase,
                                                         const int *__restrict__ item_order, int groups, int nblk,
                                                         int per_xcd) {
    constexpr int MPAD = 16 * MT;
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    __shared__ T stage[4][PJ_SUB][16 + 1];
    __shared__ double red[4][16 * 16];
    const int lb = xcd_block(nblk, per_xcd);
    if (lb < 0) return;  // the whole workgroup
    const int slot = lb / groups;
    const int item = item_order[slot];
    const int c0 = (lb - slot * groups) * 16 * GT;
    const int a = chunk_atom[item];
    const int64_t vg0 = vox_off[a];
    const int V = (int)(vox_off[a + 1] - vg0);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = lane & 15, g = lane >> 4;
    const int vb = chunk_v0[item] + wave * PJ_VPW;
    acc_t acc[GT][MT];
#pragma unroll
    for (int t = 0; t < GT; ++t)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[t][mt] = acc_t{0, 0, 0, 0};
    for (int sc = 0; sc < PJ_VPW / PJ_SUB; ++sc) {
        const int vs = vb + sc * PJ_SUB;
        if (vs >= V) break;  // wave-uniform: the rest of the wave's voxels are past the sphere
        // A[w = 16 mt + j][k = g] of the 16 k-steps of this sub-chunk
        T av[MT][PJ_SUB / 4];
#pragma unroll
        for (int s = 0; s < PJ_SUB / 4; ++s) {
            const int v = vs + 4 * s + g;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) av[mt][s] = v < V ? tabp[(vg0 + v) * MPAD + 16 * mt + j] : T(0);
        }
        // the lane's voxel for the gathers
        const int v = vs + lane;
        const bool in = v < V;
        const int64_t vid = vg0 + (in ? v : 0);
        const int gi = in ? vox_grid[vid] : 0;
        T cs = T(1), sn = T(0);
        if constexpr (PH) {
            cs = phase[2 * vid];
            sn = phase[2 * vid + 1];
        }
        // all GT x 16 gathers first, so that they are in flight together
        T gv[GT][16];
#pragma unroll
        for (int t = 0; t < GT; ++t)
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const int col = c0 + 16 * t + jj;
                gv[t][jj] = (in && col < ncol) ? psi[(int64_t)col * ng + gi] : T(0);
            }
#pragma unroll
        for (int t = 0; t < GT; ++t) {
            const int cb = c0 + 16 * t;
            if (cb < ncol) {  // uniform
#pragma unroll
                for (int jj = 0; jj < 16; jj += NC) {
                    const T re = gv[t][jj];
                    T im = T(0);
                    if constexpr (NC == 2) im = gv[t][jj + 1];
                    if constexpr (NC == 2) {
                        if constexpr (PH) {
                            stage[wave][lane][jj] = cs * re - sn * im;
                            stage[wave][lane][jj + 1] = cs * im + sn * re;
                        } else {
                            stage[wave][lane][jj] = re;
                            stage[wave][lane][jj + 1] = im;
                        }
                    } else {
                        stage[wave][lane][jj] = re;
                    }
                }
                // the stage is written and read by this wave only: its LDS
                // operations complete in order, and the fence keeps the
                // compiler from moving them across
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
                for (int s = 0; s < PJ_SUB / 4; ++s) {
                    const T b = stage[wave][4 * s + g][j];  // B[k = g][col = j]
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt) acc[t][mt] = M::mma(av[mt][s], b, acc[t][mt]);
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
        }
    }
    // sum of the four waves, tile by tile, in a fixed order
#pragma unroll
    for (int t = 0; t < GT; ++t) {
        const int cb = c0 + 16 * t;
        if (cb >= ncol) break;  // uniform over the workgroup
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
            for (int r = 0; r < 4; ++r) red[wave][M::row(g, r) * 16 + j] = (double)acc[t][mt][r];
            __syncthreads();
            {
                const int e = threadIdx.x;  // 256 = 16 projectors x 16 columns
                const int w = e >> 4, cl = e & 15;
                if (cb + cl < ncol) {
                    const double sum = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
                    part[((int64_t)item * MPAD + 16 * mt + w) * ncol + cb + cl] = sum;
                }
            }
            __syncthreads();
        }
    }
}

// coef[p * ncol + col] = E_p dV sum over the atom's chunks (fixed order, FP64)
template <typename T>
__global__ void nl_coef_kernel(T *__restrict__ coef, const double *__restrict__ part, int ncol, int ptot, int mpad,
                               const int *__restrict__ proj_atom, const int *__restrict__ proj_off,
                               const int *__restrict__ atom_chunk_off, const double *__restrict__ energy, double dV) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)ptot * ncol) return;
    const int p = (int)(e / ncol), col = (int)(e % ncol);
    const int a = proj_atom[p];
    const int w = p - proj_off[a];
    double s = 0.0;
    for (int it = atom_chunk_off[a]; it < atom_chunk_off[a + 1]; ++it)
        s += part[((int64_t)it * mpad + w) * ncol + col];
    coef[e] = (T)(s * dV * energy[p]);
}

// Hpsi[col][g] += sum over (atom, voxel) entries covering g of (e^{-ik.r} sum_w p_w c_w)_col
// for the NG = NBS x GT real columns (NBS a multiple of NC) of column group
// g (logical block = point block * groups + g).  The coefficients of the (up to kScatterSlots) atoms a workgroup
// touches are staged in LDS; each thread reads its point's entry list and
// table rows once for the whole group.  MP: projector slots per entry (9 when
// every atom has at most 9 projectors, l <= 2, else the table padding 16 or
// 32); table rows are MS = 16 (MP <= 16) or 32 values apart.
template <typename T, int NC, bool PH, int NBS, int GT, int MP>
__global__ __launch_bounds__(kScatterBlock) void nl_scatter_kernel(
    T *__restrict__ hpsi, const T *__restrict__ coef, int64_t ng, int ncol, int ncov,
    const int *__restrict__ cov_grid, const int *__restrict__ cov_off, const int *__restrict__ cov_vox,
    const int *__restrict__ cov_slot, const int *__restrict__ block_atom_off, const int *__restrict__ block_atoms,
    const int *__restrict__ vox_atom, const int *__restrict__ nproj, const int *__restrict__ proj_off,
    const T *__restrict__ tabp, const T *__restrict__ phase, int groups, int nblk, int per_xcd) {
    constexpr int NG = NBS * GT;  // columns of the group
    constexpr int MS = MP > 16 ? 32 : 16;
    __shared__ T cf_s[kScatterSlots][MP][NG];
    const int lb = xcd_block(nblk, per_xcd);
    if (lb < 0) return;  // the whole workgroup
    const int pb = lb / groups;  // point block
    const int cg = (lb - pb * groups) * NG;
    {
        const int b0 = block_atom_off[pb], nat = block_atom_off[pb + 1] - b0;
        for (int e = threadIdx.x; e < nat * MP * NG; e += kScatterBlock) {
            const int sl = e / (MP * NG), w = (e / NG) % MP, q = e % NG;
            const int a = block_atoms[b0 + sl];
            cf_s[sl][w][q] = (w < nproj[a] && cg + q < ncol) ? coef[(int64_t)(proj_off[a] + w) * ncol + cg + q] : T(0);
        }
    }
    __syncthreads();
    const int i = pb * kScatterBlock + threadIdx.x;
    if (i >= ncov) return;
    const int g = cov_grid[i];
    const int e0 = cov_off[i], e1 = cov_off[i + 1];  // e1 > e0: a covered point has an entry
    // An entry's voxel, slot, table row and phase are requested before the
    // point's NG values: the table row depends on the voxel index, so this
    // dependent round trip overlaps the NG loads instead of following them.
    int vid = cov_vox[e0], sl = cov_slot[e0];
    T pw[MP];
#pragma unroll
    for (int w = 0; w < MP; ++w) pw[w] = tabp[(int64_t)vid * MS + w];
    T cs = T(1), sn = T(0);
    if constexpr (PH) {
        cs = phase[2 * (int64_t)vid];
        sn = phase[2 * (int64_t)vid + 1];
    }
    // the point's values of all NG columns, requested together, then updated
    // in registers and written back once
    T hv[NG];
#pragma unroll
    for (int q = 0; q < NG; ++q) hv[q] = cg + q < ncol ? hpsi[(int64_t)(cg + q) * ng + g] : T(0);
    for (int e = e0;;) {
        if (sl >= 0) {
#pragma unroll
            for (int q = 0; q < NG; q += NC) {
                T tr = T(0), ti = T(0);
#pragma unroll
                for (int w = 0; w < MP; ++w) {
                    tr += pw[w] * cf_s[sl][w][q];
                    if constexpr (NC == 2) ti += pw[w] * cf_s[sl][w][q + 1];
                }
                if constexpr (PH) {
                    hv[q] += cs * tr + sn * ti;
                    hv[q + 1] += cs * ti - sn * tr;
                } else {
                    hv[q] += tr;
                    if constexpr (NC == 2) hv[q + 1] += ti;
                }
            }
        } else {  // more atoms in this workgroup than LDS slots: coefficients from global memory
            const int a = vox_atom[vid];
            const int M = nproj[a];
            const T *cf = coef + (int64_t)proj_off[a] * ncol + cg;
#pragma unroll
            for (int q = 0; q < NG; q += NC) {
                T tr = T(0), ti = T(0);
                if (cg + q < ncol) {
                    for (int w = 0; w < M; ++w) {
                        tr += pw[w] * cf[(int64_t)w * ncol + q];
                        if constexpr (NC == 2) ti += pw[w] * cf[(int64_t)w * ncol + q + 1];
                    }
                }
                if constexpr (PH) {
                    hv[q] += cs * tr + sn * ti;
                    hv[q + 1] += cs * ti - sn * tr;
                } else {
                    hv[q] += tr;
                    if constexpr (NC == 2) hv[q + 1] += ti;
                }
            }
        }
        if (++e >= e1) break;
        vid = cov_vox[e];
        sl = cov_slot[e];
#pragma unroll
        for (int w = 0; w < MP; ++w) pw[w] = tabp[(int64_t)vid * MS + w];
        if constexpr (PH) {
            cs = phase[2 * (int64_t)vid];
            sn = phase[2 * (int64_t)vid + 1];
        }
    }
#pragma unroll
    for (int q = 0; q < NG; ++q)
        if (cg + q < ncol) hpsi[(int64_t)(cg + q) * ng + g] = hv[q];
}

// Per (atom, band group, group of MM projectors): E_nl and
// F = 2 sum occ E Re(conj(<p|psi>) <dp/dr|psi>) over the projectors
// MM blockIdx.z .. MM blockIdx.z + MM - 1 of the atom.
// out[((blockIdx.z * gridDim.y + blockIdx.y) * natoms + a) * 4 + {0: E, 1..3: F}]
template <typename T, int MM, int NBG, int NC, bool PH>
__global__ __launch_bounds__(NL_THREADS) void nl_force_kernel(
    double *__restrict__ out, const T *__restrict__ psi, int64_t ng, int nb, int natoms, const double *__restrict__ occ,
    const int *__restrict__ nproj, const int *__restrict__ proj_off, const int64_t *__restrict__ vox_off,
    const int64_t *__restrict__ tab_off, int64_t ttot, const int *__restrict__ vox_grid, const T *__restrict__ table,
    const T *__restrict__ dtable, const T *__restrict__ phase, const double *__restrict__ energy, double dV) {
    constexpr int NV = 4 * MM * NBG * NC;
    __shared__ double red[NL_THREADS / 64][NV];
    __shared__ double tot[4];
    const int a = blockIdx.x;
    const int n0 = blockIdx.y * NBG;
    const int64_t v0 = vox_off[a];
    const int V = (int)(vox_off[a + 1] - v0);
    const int w0 = blockIdx.z * MM;
    const int M = min(nproj[a] - w0, MM);  // projectors of this block (may be <= 0)
    const int64_t t0 = tab_off[a] + (int64_t)w0 * V;
    const T *tabs[4] = {table + t0, dtable + t0, dtable + ttot + t0, dtable + 2 * ttot + t0};

    T acc[4][MM][NBG][NC];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int w = 0; w < MM; ++w)
#pragma unroll
            for (int b = 0; b < NBG; ++b)
#pragma unroll
                for (int c = 0; c < NC; ++c) acc[q][w][b][c] = T(0);

    for (int v = threadIdx.x; v < V; v += NL_THREADS) {
        const int g = vox_grid[v0 + v];
        T cs = T(1), sn = T(0);
        if constexpr (PH) {
            cs = phase[2 * (v0 + v)];
            sn = phase[2 * (v0 + v) + 1];
        }
        T br[NBG], bi[NBG];
#pragma unroll
        for (int b = 0; b < NBG; ++b) {
            const int n = min(n0 + b, nb - 1);
            T re = psi[(int64_t)n * NC * ng + g];
            T im = T(0);
            if constexpr (NC == 2) im = psi[((int64_t)n * NC + 1) * ng + g];
            br[b] = PH ? cs * re - sn * im : re;
            bi[b] = PH ? cs * im + sn * re : im;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int w = 0; w < MM; ++w) {
                const T pv = (w < M) ? tabs[q][(int64_t)w * V + v] : T(0);
#pragma unroll
                for (int b = 0; b < NBG; ++b) {
                    acc[q][w][b][0] += pv * br[b];
                    if constexpr (NC == 2) acc[q][w][b][1] += pv * bi[b];
                }
            }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int w = 0; w < MM; ++w)
#pragma unroll
            for (int b = 0; b < NBG; ++b)
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const double s = wave_sum<T>((double)acc[q][w][b][c]);
                    if (lane == 0) red[wid][((q * MM + w) * NBG + b) * NC + c] = s;
                }
    if (threadIdx.x < 4) tot[threadIdx.x] = 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
        double e = 0.0, f[3] = {0.0, 0.0, 0.0};
        for (int b = 0; b < NBG; ++b) {
            if (n0 + b >= nb) break;
            const double o = occ[n0 + b];
            for (int w = 0; w < M; ++w) {
                double s[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
                for (int q = 0; q < 4; ++q)
                    for (int c = 0; c < NC; ++c) {
                        double t = 0.0;
                        for (int k = 0; k < NL_THREADS / 64; ++k) t += red[k][((q * MM + w) * NBG + b) * NC + c];
                        s[q][c] = t * dV;
                    }
                const double eo = energy[proj_off[a] + w0 + w] * o;
                e += (s[0][0] * s[0][0] + s[0][1] * s[0][1]) * eo;
                for (int d = 0; d < 3; ++d) f[d] += 2.0 * (s[0][0] * s[d + 1][0] + s[0][1] * s[d + 1][1]) * eo;
            }
        }
        double *o4 = out + (((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * natoms + a) * 4;
        o4[0] = e;
        o4[1] = f[0];
        o4[2] = f[1];
        o4[3] = f[2];
    }
}

template <typename T>
const T *table_of(const Projectors &p);
template <>
const double *table_of<double>(const Projectors &p) { return p.d_table64; }
template <>
const float *table_of<float>(const Projectors &p) { return p.d_table32; }
template <typename T>
const T *dtable_of(const Projectors &p);
template <>
const double *dtable_of<double>(const Projectors &p) { return p.d_dtable64; }
template <>
const float *dtable_of<float>(const Projectors &p) { return p.d_dtable32; }
template <typename T>
const T *phase_of(const KPoint &k);
template <>
const double *phase_of<double>(const KPoint &k) { return k.d_phase64; }
template <>
const float *phase_of<float>(const KPoint &k) { return k.d_phase32; }

template <typename T>
const T *tabp_of(const Projectors &p);
template <>
const double *tabp_of<double>(const Projectors &p) { return p.d_tabp_64; }
template <>
const float *tabp_of<float>(const Projectors &p) { return p.d_tabp_32; }

constexpr int64_t kProjMinBlocks = 1024;  // 4 workgroups per CU

// Tuning key of a projection shape; band counts share a key within a
// power-of-two bucket (as the stencil's).
int64_t nl_tune_key(int precision, int nc, bool ph, int mpad, int nb) {
    int64_t bucket = 1;
    while (bucket < nb) bucket <<= 1;
    return ((((bucket * 4 + (nc - 1) * 2 + (ph ? 1 : 0)) * 64) + mpad) << 8) | precision;
}

template <typename T, int NC, bool PH, int GT, int MT>
void launch_project(qm_ctx *ctx, double *part, const T *psi, int ncol, const T *phase, hipStream_t st) {
    const Projectors &p = ctx->proj;
    const int groups = (ncol + 16 * GT - 1) / (16 * GT);
    const int64_t nblk = (int64_t)p.nchunk * groups;
    const int per_xcd = (int)((nblk + 7) / 8);
    hipLaunchKernelGGL((nl_project_kernel<T, NC, PH, GT, MT>), dim3(8 * per_xcd), dim3(256), 0, st, part, psi, ctx->ng,
                       ncol, p.d_chunk_atom, p.d_chunk_v0, p.d_vox_off, p.d_vox_grid, tabp_of<T>(p), phase,
                       p.d_item_order, groups, (int)nblk, per_xcd);
}

// Projection coefficients of nb bands into coef (ptot x ncol, T) through the
// per-chunk partial sums part (nchunk x mpad x ncol, FP64).  Both buffers
// are written in full before they are read, so they need no clearing.
template <typename T, int NC, bool PH>
int nonlocal_project(qm_ctx *ctx, T *coef, double *part, const T *psi, int nb, const T *phase, hipStream_t st) {
    const Projectors &p = ctx->proj;
    const int ncol = nb * NC;
    // 4 tiles of 16 columns per projection workgroup share each table read;
    // the columns beyond split over more workgroups.  A launch that would then
    // have fewer than kProjMinBlocks workgroups (a small grid or few bands:
    // 64^3 x 128 has 640) takes 1 tile per workgroup instead, for four times
    // the workgroups in flight on this latency-bound gather (2 tiles measured
    // 1-2 % slower there).  Each column's sum runs over the same voxels in the
    // same order either way, so every shape gives the same values.  The first
    // projection of a shape (precision, ncomp, phases, padding, band-count
    // bucket) times the candidates on the caller's buffers (the projection only
    // writes its own partial sums) and keeps the fastest; tables padded to 32
    // projectors (two MFMA row tiles per table read) take at most 2 column
    // tiles, for the registers.
    const int64_t g4 = (ncol + 63) / 64;
    QM_CHECK((int64_t)p.nchunk * g4 * 4 < (int64_t)INT32_MAX / 16, QM_ERR_ARG,
             "nonlocal apply: too many bands per call");
    auto launch = [&](int gt) {
        if (p.mpad == 16) {
            if (gt == 4) launch_project<T, NC, PH, 4, 1>(ctx, part, psi, ncol, phase, st);
            else if (gt == 2) launch_project<T, NC, PH, 2, 1>(ctx, part, psi, ncol, phase, st);
            else launch_project<T, NC, PH, 1, 1>(ctx, part, psi, ncol, phase, st);
        } else {
            if (gt >= 2) launch_project<T, NC, PH, 2, 2>(ctx, part, psi, ncol, phase, st);
            else launch_project<T, NC, PH, 1, 2>(ctx, part, psi, ncol, phase, st);
        }
    };
    const int max_gt = p.mpad == 16 ? 4 : 2;
    int gt = ctx->nl_force > 0 ? std::min(ctx->nl_force, max_gt) : 0;
    if (gt == 0) {
        const int64_t key = nl_tune_key(sizeof(T) == 8 ? 64 : 32, NC, PH, p.mpad, nb);
        auto it = ctx->nl_choice.find(key);
        if (it != ctx->nl_choice.end()) {
            gt = it->second;
        } else {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            QM_HIP(hipStreamIsCapturing(st, &cs));
            if (cs != hipStreamCaptureStatusNone || g_internal_capture || ctx->nl_notime) {
                gt = (int64_t)p.nchunk * g4 >= kProjMinBlocks ? max_gt : 1;  // no timing inside a capture
            } else {
                hipEvent_t ev[2] = {nullptr, nullptr};
                QM_HIP(hipEventCreate(&ev[0]));
                if (hipEventCreate(&ev[1]) != hipSuccess) {
                    (void)hipEventDestroy(ev[0]);
                    QM_CHECK(false, QM_ERR_HIP, "nonlocal autotune: hipEventCreate failed");
                }
                float best = 1e30f;
                int status = QM_OK;
                for (int cand = 1; cand <= max_gt && status == QM_OK; cand *= 2) {
                    launch(cand);  // warm-up
                    float ms = 0.0f;
                    if (hipEventRecord(ev[0], st) != hipSuccess) status = QM_ERR_HIP;
                    for (int rep = 0; rep < 3 && status == QM_OK; ++rep) launch(cand);
                    if (status == QM_OK && hipEventRecord(ev[1], st) != hipSuccess) status = QM_ERR_HIP;
                    if (status == QM_OK && hipEventSynchronize(ev[1]) != hipSuccess) status = QM_ERR_HIP;
                    if (status == QM_OK && hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) status = QM_ERR_HIP;
                    if (status == QM_OK && hipGetLastError() != hipSuccess) status = QM_ERR_HIP;
                    if (status == QM_OK && ms < best) {
                        best = ms;
                        gt = cand;
                    }
                }
                (void)hipEventDestroy(ev[0]);
                (void)hipEventDestroy(ev[1]);
                QM_CHECK(status == QM_OK, QM_ERR_HIP, "nonlocal autotune: HIP failure");
                ctx->nl_choice[key] = gt;
            }
        }
    }
    launch(gt);
    QM_HIP(hipGetLastError());
    const int64_t ncoef = (int64_t)p.ptot * ncol;
    hipLaunchKernelGGL((nl_coef_kernel<T>), dim3((unsigned)((ncoef + 255) / 256)), dim3(256), 0, st, coef, part, ncol,
                       p.ptot, p.mpad, p.d_proj_atom, p.d_proj_off, p.d_atom_chunk_off, p.d_energy,
                       ctx->hx * ctx->hy * ctx->hz);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

template <typename T, int NC, bool PH>
int nonlocal_scatter(qm_ctx *ctx, T *hpsi, const T *coef, int nb, const T *phase, hipStream_t st) {
    const Projectors &p = ctx->proj;
    const int ncol = nb * NC;
    // columns per scatter tile (16; 64 per thread measured 1.9x slower on
    // 64^3 x 128) and tiles per workgroup (the staged coefficients, entry lists
    // and table rows serve all of them)
    constexpr int NBS = 16;
    constexpr int SGT = sizeof(T) == 8 ? 1 : 2;
    const int sgroups = (ncol + NBS * SGT - 1) / (NBS * SGT);
    const int64_t nblk = (int64_t)p.nsblock * sgroups;
    QM_CHECK(nblk < (int64_t)INT32_MAX / 16, QM_ERR_ARG, "nonlocal apply: too many bands per call");
    const int per_xcd = (int)((nblk + 7) / 8);
    if (p.max_m <= 9)
        hipLaunchKernelGGL((nl_scatter_kernel<T, NC, PH, NBS, SGT, 9>), dim3(8 * per_xcd), dim3(kScatterBlock), 0, st,
                           hpsi, coef, ctx->ng, ncol, p.ncov, p.d_cov_grid, p.d_cov_off, p.d_cov_vox, p.d_cov_slot,
                           p.d_block_atom_off, p.d_block_atoms, p.d_vox_atom, p.d_nproj, p.d_proj_off,
                           tabp_of<T>(p), phase, sgroups, (int)nblk, per_xcd);
    else if (p.max_m <= 16)
        hipLaunchKernelGGL((nl_scatter_kernel<T, NC, PH, NBS, SGT, 16>), dim3(8 * per_xcd), dim3(kScatterBlock),
                           0, st, hpsi, coef, ctx->ng, ncol, p.ncov, p.d_cov_grid, p.d_cov_off, p.d_cov_vox,
                           p.d_cov_slot, p.d_block_atom_off, p.d_block_atoms, p.d_vox_atom, p.d_nproj, p.d_proj_off,
                           tabp_of<T>(p), phase, sgroups, (int)nblk, per_xcd);
    else
        hipLaunchKernelGGL((nl_scatter_kernel<T, NC, PH, NBS, SGT, 32>), dim3(8 * per_xcd), dim3(kScatterBlock),
                           0, st, hpsi, coef, ctx->ng, ncol, p.ncov, p.d_cov_grid, p.d_cov_off, p.d_cov_vox,
                           p.d_cov_slot, p.d_block_atom_off, p.d_block_atoms, p.d_vox_atom, p.d_nproj, p.d_proj_off,
                           tabp_of<T>(p), phase, sgroups, (int)nblk, per_xcd);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

// which: 1 = projection coefficients, 2 = scatter, 3 = both.  coef / part are
// the work buffers (nonlocal_work_bytes); NULL takes the context's own, grown
// as needed.
template <typename T>
int launch_nonlocal(qm_ctx *ctx, T *hpsi, const T *psi, int nb, int ncomp, const KPoint &kp, hipStream_t st,
                    int which = 3, T *coef = nullptr, double *part = nullptr) {
    const Projectors &p = ctx->proj;
    if (p.natoms == 0 || p.ptot == 0) return QM_OK;
    const bool ph = !kp.gamma;
    QM_CHECK(ncomp == 2 || !ph, QM_ERR_ARG, "real (ncomp=1) bands need a Gamma k-point");
    const T *phase = phase_of<T>(kp);
    size_t cb = 0, pb = 0;
    nonlocal_work_bytes(ctx, nb, ncomp, sizeof(T), &cb, &pb);
    if (!coef && ctx->capture_pins) {  // a caller's graph holds the current work buffers
        QM_CHECK(ctx->coef.bytes >= cb && ((which & 1) == 0 || ctx->nl_part.bytes >= pb), QM_ERR_STATE,
                 "H.psi apply: larger nonlocal work buffers needed (more bands, or FP64 after FP32) while a "
                 "graph holds an apply of this context; destroy it and call qm_release_captures first");
    }
    if (!coef) {
        if (which & 1) QM_TRY(ctx->coef.reserve(cb));
        QM_CHECK(ctx->coef.bytes >= cb, QM_ERR_STATE,
                 "nonlocal scatter: no projection coefficients for this band count");
        coef = ctx->coef.as<T>();
    }
    if ((which & 1) && !part) {
        QM_TRY(ctx->nl_part.reserve(pb));
        part = ctx->nl_part.as<double>();
    }
    if (which & 1) {
        if (ncomp == 1) QM_TRY((nonlocal_project<T, 1, false>(ctx, coef, part, psi, nb, phase, st)));
        else if (!ph) QM_TRY((nonlocal_project<T, 2, false>(ctx, coef, part, psi, nb, phase, st)));
        else QM_TRY((nonlocal_project<T, 2, true>(ctx, coef, part, psi, nb, phase, st)));
    }
    if (which & 2) {
        if (ncomp == 1) QM_TRY((nonlocal_scatter<T, 1, false>(ctx, hpsi, coef, nb, phase, st)));
        else if (!ph) QM_TRY((nonlocal_scatter<T, 2, false>(ctx, hpsi, coef, nb, phase, st)));
        else QM_TRY((nonlocal_scatter<T, 2, true>(ctx, hpsi, coef, nb, phase, st)));
    }
    return QM_OK;
}

template <typename T, int MM, int NC, bool PH>
void launch_force(qm_ctx *ctx, double *out, const T *psi, int nb, const double *d_occ, const T *phase, int nz,
                  hipStream_t st) {
    constexpr int NBG = 1;
    const Projectors &p = ctx->proj;
    dim3 grid(p.natoms, (nb + NBG - 1) / NBG, nz);
    hipLaunchKernelGGL((nl_force_kernel<T, MM, NBG, NC, PH>), grid, dim3(NL_THREADS), 0, st, out, psi, ctx->ng, nb,
                       p.natoms, d_occ, p.d_nproj, p.d_proj_off, p.d_vox_off, p.d_tab_off, p.ttot, p.d_vox_grid,
                       table_of<T>(p), dtable_of<T>(p), phase, p.d_energy, ctx->hx * ctx->hy * ctx->hz);
}

template <typename T>
int nonlocal_forces(qm_ctx *ctx, const T *psi, const double *occ, int nb, int ncomp, const KPoint &kp,
                    double *forces, double *e_nl, hipStream_t st) {
    const Projectors &p = ctx->proj;
    QM_CHECK(p.has_grad, QM_ERR_STATE, "nonlocal forces need the projector gradient tables (dtable)");
    QM_CHECK(ncomp == 1 || ncomp == 2, QM_ERR_ARG, "ncomp must be 1 or 2");
    const bool ph = !kp.gamma;
    QM_CHECK(ncomp == 2 || !ph, QM_ERR_ARG, "real (ncomp=1) bands need a Gamma k-point");
    // one band per force block (see launch_force); projectors in groups of 16
    // over blockIdx.z when an atom has more than 16
    const int nz = p.max_m > 16 ? 2 : 1;
    const int ngrp = nb * nz;
    const size_t out_bytes = sizeof(double) * (size_t)ngrp * p.natoms * 4;
    QM_TRY(ctx->fwork.reserve(out_bytes + sizeof(double) * nb));
    double *d_out = ctx->fwork.as<double>();
    double *d_occ = d_out + (size_t)ngrp * p.natoms * 4;
    QM_HIP(hipMemcpyAsync(d_occ, occ, sizeof(double) * nb, hipMemcpyHostToDevice, st));
    const T *phase = phase_of<T>(kp);
    const bool small = p.max_m <= 9;
    if (ncomp == 1) {
        if (small) launch_force<T, 9, 1, false>(ctx, d_out, psi, nb, d_occ, phase, nz, st);
        else launch_force<T, 16, 1, false>(ctx, d_out, psi, nb, d_occ, phase, nz, st);
    } else if (!ph) {
        if (small) launch_force<T, 9, 2, false>(ctx, d_out, psi, nb, d_occ, phase, nz, st);
        else launch_force<T, 16, 2, false>(ctx, d_out, psi, nb, d_occ, phase, nz, st);
    } else {
        if (small) launch_force<T, 9, 2, true>(ctx, d_out, psi, nb, d_occ, phase, nz, st);
        else launch_force<T, 16, 2, true>(ctx, d_out, psi, nb, d_occ, phase, nz, st);
    }
    QM_HIP(hipGetLastError());
    std::vector<double> h((size_t)ngrp * p.natoms * 4);
    QM_HIP(hipMemcpyAsync(h.data(), d_out, out_bytes, hipMemcpyDeviceToHost, st));
    QM_HIP(hipStreamSynchronize(st));
    double e = 0.0;
    for (int a = 0; a < p.natoms; ++a) {
        double f[3] = {0, 0, 0};
        for (int g = 0; g < ngrp; ++g) {
            const double *o4 = &h[((size_t)g * p.natoms + a) * 4];
            e += o4[0];
            f[0] += o4[1];
            f[1] += o4[2];
            f[2] += o4[3];
        }
        forces[3 * a] = f[0];
        forces[3 * a + 1] = f[1];
        forces[3 * a + 2] = f[2];
    }
    *e_nl = e;
    return QM_OK;
}

}  // namespace

int launch_nonlocal_f64(qm_ctx *ctx, double *hpsi, const double *psi, int nb, int ncomp, const KPoint &kp,
                        hipStream_t st) {
    return launch_nonlocal<double>(ctx, hpsi, psi, nb, ncomp, kp, st);
}
int launch_nonlocal_f32(qm_ctx *ctx, float *hpsi, const float *psi, int nb, int ncomp, const KPoint &kp,
                        hipStream_t st) {
    return launch_nonlocal<float>(ctx, hpsi, psi, nb, ncomp, kp, st);
}
int nonlocal_project_f64(qm_ctx *ctx, const double *psi, int nb, int ncomp, const KPoint &kp, hipStream_t st) {
    return launch_nonlocal<double>(ctx, nullptr, psi, nb, ncomp, kp, st, 1);
}
int nonlocal_project_f32(qm_ctx *ctx, const float *psi, int nb, int ncomp, const KPoint &kp, hipStream_t st) {
    return launch_nonlocal<float>(ctx, nullptr, psi, nb, ncomp, kp, st, 1);
}
int nonlocal_scatter_f64(qm_ctx *ctx, double *hpsi, int nb, int ncomp, const KPoint &kp, hipStream_t st) {
    return launch_nonlocal<double>(ctx, hpsi, nullptr, nb, ncomp, kp, st, 2);
}
int nonlocal_scatter_f32(qm_ctx *ctx, float *hpsi, int nb, int ncomp, const KPoint &kp, hipStream_t st) {
    return launch_nonlocal<float>(ctx, hpsi, nullptr, nb, ncomp, kp, st, 2);
}
// Bytes of the two nonlocal work buffers for nb bands: the coefficients
// (ptot x ncol, T) and the per-chunk partial sums (nchunk x mpad x ncol, FP64).
void nonlocal_work_bytes(const qm_ctx *ctx, int nb, int ncomp, size_t tsize, size_t *coef_bytes,
                         size_t *part_bytes) {
    const Projectors &p = ctx->proj;
    const size_t ncol = (size_t)nb * ncomp;
    *coef_bytes = tsize * (size_t)p.ptot * ncol;
    *part_bytes = sizeof(double) * (size_t)p.nchunk * p.mpad * ncol;
}

int nonlocal_choice(const qm_ctx *ctx, int precision, int ncomp, int kt, int nb) {
    if (ctx->nl_force > 0) return ctx->nl_force;
    const int64_t key = nl_tune_key(precision, ncomp, (kt & 1) != 0 && ncomp == 2, ctx->proj.mpad, nb);
    auto it = ctx->nl_choice.find(key);
    return it == ctx->nl_choice.end() ? -1 : it->second;
}

int nonlocal_apply_in_f32(qm_ctx *ctx, float *hpsi, const float *psi, int nb, int ncomp, const KPoint &kp,
                          int which, float *coef, double *part, hipStream_t st) {
    return launch_nonlocal<float>(ctx, hpsi, psi, nb, ncomp, kp, st, which, coef, part);
}
int nonlocal_forces_f64(qm_ctx *ctx, const double *psi, const double *occ, int nb, int ncomp, const KPoint &kp,
                        double *forces, double *e_nl, hipStream_t st) {
    return nonlocal_forces<double>(ctx, psi, occ, nb, ncomp, kp, forces, e_nl, st);
}
int nonlocal_forces_f32(qm_ctx *ctx, const float *psi, const double *occ, int nb, int ncomp, const KPoint &kp,
                        double *forces, double *e_nl, hipStream_t st) {
    return nonlocal_forces<float>(ctx, psi, occ, nb, ncomp, kp, forces, e_nl, st);
}

}  // namespace qm
