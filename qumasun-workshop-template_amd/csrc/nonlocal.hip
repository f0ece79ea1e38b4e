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
//   1. project: one launch covers every atom (uneven sphere sizes, heaviest
//      atoms scheduled first) x every band group.  Each block gathers the
//      atom's voxels of NBG bands straight from the grid, applies e^{+ik.r},
//      accumulates <p_w|psi_n> per thread, and reduces across the block in
//      FP64; E_w dV is folded in before the coefficient is stored.
//   2. scatter: owner-computes over the grid points covered by any sphere.
//      Each covered point sums the contributions of all atoms whose sphere
//      contains it, so overlapping spheres need no atomics or colouring and
//      the result is deterministic.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "qm_internal.h"

namespace qm {
namespace {

constexpr int NL_THREADS = 256;

template <typename T>
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// coef[(proj_off[a] + w) * ncol + n*NC + c] = E_w dV <p_w | e^{ik.r} psi_n>_c
template <typename T, int MM, int NBG, int NC, bool PH>
__global__ __launch_bounds__(NL_THREADS) void nl_project_kernel(
    T *__restrict__ coef, const T *__restrict__ psi, int64_t ng, int nb, const int *__restrict__ atom_order,
    const int *__restrict__ nproj, const int *__restrict__ proj_off, const int64_t *__restrict__ vox_off,
    const int64_t *__restrict__ tab_off, const int *__restrict__ vox_grid, const T *__restrict__ table,
    const T *__restrict__ phase, const double *__restrict__ energy, double dV) {
    constexpr int NV = MM * NBG * NC;
    __shared__ double red[NL_THREADS / 64][NV];
    const int a = atom_order[blockIdx.x];
    const int n0 = blockIdx.y * NBG;
    const int64_t v0 = vox_off[a];
    const int V = (int)(vox_off[a + 1] - v0);
    const int M = nproj[a];
    const T *tab = table + tab_off[a];
    const int ncol = nb * NC;

    T acc[MM][NBG][NC];
#pragma unroll
    for (int w = 0; w < MM; ++w)
#pragma unroll
        for (int b = 0; b < NBG; ++b)
#pragma unroll
            for (int c = 0; c < NC; ++c) acc[w][b][c] = T(0);

    const T *col[NBG][NC];
#pragma unroll
    for (int b = 0; b < NBG; ++b)
#pragma unroll
        for (int c = 0; c < NC; ++c) col[b][c] = psi + ((int64_t)min(n0 + b, nb - 1) * NC + c) * ng;

    for (int v = threadIdx.x; v < V; v += NL_THREADS) {
        const int g = vox_grid[v0 + v];
        T pw[MM];
#pragma unroll
        for (int w = 0; w < MM; ++w) pw[w] = (w < M) ? tab[(int64_t)w * V + v] : T(0);
        T cs = T(1), sn = T(0);
        if constexpr (PH) {
            cs = phase[2 * (v0 + v)];
            sn = phase[2 * (v0 + v) + 1];
        }
#pragma unroll
        for (int b = 0; b < NBG; ++b) {
            T re = col[b][0][g];
            T im = T(0);
            if constexpr (NC == 2) im = col[b][1][g];
            if constexpr (PH) {
                const T br = cs * re - sn * im;
                const T bi = cs * im + sn * re;
                re = br;
                im = bi;
            }
#pragma unroll
            for (int w = 0; w < MM; ++w) {
                acc[w][b][0] += pw[w] * re;
                if constexpr (NC == 2) acc[w][b][1] += pw[w] * im;
            }
        }
    }

    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int w = 0; w < MM; ++w)
#pragma unroll
        for (int b = 0; b < NBG; ++b)
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const double s = wave_sum<T>((double)acc[w][b][c]);
                if (lane == 0) red[wid][(w * NBG + b) * NC + c] = s;
            }
    __syncthreads();
    for (int i = threadIdx.x; i < NV; i += NL_THREADS) {
        const int w = i / (NBG * NC);
        const int b = (i / NC) % NBG;
        const int c = i % NC;
        if (w < M && n0 + b < nb) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < NL_THREADS / 64; ++k) s += red[k][i];
            coef[(int64_t)(proj_off[a] + w) * ncol + (n0 + b) * NC + c] = (T)(s * dV * energy[proj_off[a] + w]);
        }
    }
}

// Hpsi[g] += sum over (atom, voxel) entries covering g of e^{-ik.r} sum_w p_w c_w
template <typename T, int NC, bool PH>
__global__ __launch_bounds__(NL_THREADS) void nl_scatter_kernel(
    T *__restrict__ hpsi, const T *__restrict__ coef, int64_t ng, int nb, int ncov, const int *__restrict__ cov_grid,
    const int *__restrict__ cov_off, const int *__restrict__ cov_vox, const int *__restrict__ vox_atom,
    const int *__restrict__ nproj, const int *__restrict__ proj_off, const int64_t *__restrict__ vox_off,
    const int64_t *__restrict__ tab_off, const T *__restrict__ table, const T *__restrict__ phase) {
    const int i = blockIdx.x * NL_THREADS + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= ncov) return;
    const int ncol = nb * NC;
    const int g = cov_grid[i];
    T sr = T(0), si = T(0);
    for (int e = cov_off[i]; e < cov_off[i + 1]; ++e) {
        const int vid = cov_vox[e];
        const int a = vox_atom[vid];
        const int64_t va = vox_off[a];
        const int V = (int)(vox_off[a + 1] - va);
        const int v = (int)(vid - va);
        const int M = nproj[a];
        const T *tab = table + tab_off[a] + v;
        const T *cf = coef + (int64_t)proj_off[a] * ncol + n * NC;
        T tr = T(0), ti = T(0);
        for (int w = 0; w < M; ++w) {
            const T pv = tab[(int64_t)w * V];
            tr += pv * cf[(int64_t)w * ncol];
            if constexpr (NC == 2) ti += pv * cf[(int64_t)w * ncol + 1];
        }
        if constexpr (PH) {
            const T cs = phase[2 * (int64_t)vid], sn = phase[2 * (int64_t)vid + 1];
            sr += cs * tr + sn * ti;
            si += cs * ti - sn * tr;
        } else {
            sr += tr;
            si += ti;
        }
    }
    hpsi[(int64_t)n * NC * ng + g] += sr;
    if constexpr (NC == 2) hpsi[((int64_t)n * NC + 1) * ng + g] += si;
}

// Per (atom, band group): E_nl and F = 2 sum occ E Re(conj(<p|psi>) <dp/dr|psi>).
// out[(blockIdx.y * natoms + a) * 4 + {0: E, 1..3: F}]
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
    const int M = nproj[a];
    const T *tabs[4] = {table + tab_off[a], dtable + tab_off[a], dtable + ttot + tab_off[a],
                        dtable + 2 * ttot + tab_off[a]};

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
                const double eo = energy[proj_off[a] + w] * o;
                e += (s[0][0] * s[0][0] + s[0][1] * s[0][1]) * eo;
                for (int d = 0; d < 3; ++d) f[d] += 2.0 * (s[0][0] * s[d + 1][0] + s[0][1] * s[d + 1][1]) * eo;
            }
        }
        double *o4 = out + ((int64_t)blockIdx.y * natoms + a) * 4;
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

template <typename T, int MM, int NC, bool PH>
void launch_project(qm_ctx *ctx, T *coef, const T *psi, int nb, const T *phase, hipStream_t st) {
    constexpr int NBG = (NC == 1) ? 8 : 4;
    const Projectors &p = ctx->proj;
    dim3 grid(p.natoms, (nb + NBG - 1) / NBG);
    hipLaunchKernelGGL((nl_project_kernel<T, MM, NBG, NC, PH>), grid, dim3(NL_THREADS), 0, st, coef, psi, ctx->ng,
                       nb, p.d_atom_order, p.d_nproj, p.d_proj_off, p.d_vox_off, p.d_tab_off, p.d_vox_grid,
                       table_of<T>(p), phase, p.d_energy, ctx->hx * ctx->hy * ctx->hz);
}

template <typename T, int NC, bool PH>
void launch_scatter(qm_ctx *ctx, T *hpsi, const T *coef, int nb, const T *phase, hipStream_t st) {
    const Projectors &p = ctx->proj;
    dim3 grid((p.ncov + NL_THREADS - 1) / NL_THREADS, nb);
    hipLaunchKernelGGL((nl_scatter_kernel<T, NC, PH>), grid, dim3(NL_THREADS), 0, st, hpsi, coef, ctx->ng, nb, p.ncov,
                       p.d_cov_grid, p.d_cov_off, p.d_cov_vox, p.d_vox_atom, p.d_nproj, p.d_proj_off, p.d_vox_off,
                       p.d_tab_off, table_of<T>(p), phase);
}

template <typename T>
int launch_nonlocal(qm_ctx *ctx, T *hpsi, const T *psi, int nb, int ncomp, const KPoint &kp, hipStream_t st) {
    const Projectors &p = ctx->proj;
    if (p.natoms == 0 || p.ptot == 0) return QM_OK;
    QM_CHECK(nb <= 65535, QM_ERR_ARG, "nonlocal apply: at most 65535 bands per call");
    QM_TRY(ctx->coef.reserve(sizeof(T) * (size_t)p.ptot * nb * ncomp));
    T *coef = ctx->coef.as<T>();
    const bool ph = !kp.gamma;
    QM_CHECK(ncomp == 2 || !ph, QM_ERR_ARG, "real (ncomp=1) bands need a Gamma k-point");
    const T *phase = phase_of<T>(kp);
    const bool small = p.max_m <= 9;
    if (ncomp == 1) {
        if (small) launch_project<T, 9, 1, false>(ctx, coef, psi, nb, phase, st);
        else launch_project<T, 16, 1, false>(ctx, coef, psi, nb, phase, st);
        QM_HIP(hipGetLastError());
        launch_scatter<T, 1, false>(ctx, hpsi, coef, nb, phase, st);
    } else if (!ph) {
        if (small) launch_project<T, 9, 2, false>(ctx, coef, psi, nb, phase, st);
        else launch_project<T, 16, 2, false>(ctx, coef, psi, nb, phase, st);
        QM_HIP(hipGetLastError());
        launch_scatter<T, 2, false>(ctx, hpsi, coef, nb, phase, st);
    } else {
        if (small) launch_project<T, 9, 2, true>(ctx, coef, psi, nb, phase, st);
        else launch_project<T, 16, 2, true>(ctx, coef, psi, nb, phase, st);
        QM_HIP(hipGetLastError());
        launch_scatter<T, 2, true>(ctx, hpsi, coef, nb, phase, st);
    }
    QM_HIP(hipGetLastError());
    return QM_OK;
}

template <typename T, int MM, int NC, bool PH>
void launch_force(qm_ctx *ctx, double *out, const T *psi, int nb, const double *d_occ, const T *phase,
                  hipStream_t st) {
    constexpr int NBG = 1;
    const Projectors &p = ctx->proj;
    dim3 grid(p.natoms, (nb + NBG - 1) / NBG);
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
    const int ngrp = nb;  // one band per force block (see launch_force)
    const size_t out_bytes = sizeof(double) * (size_t)ngrp * p.natoms * 4;
    QM_TRY(ctx->fwork.reserve(out_bytes + sizeof(double) * nb));
    double *d_out = ctx->fwork.as<double>();
    double *d_occ = d_out + (size_t)ngrp * p.natoms * 4;
    QM_HIP(hipMemcpyAsync(d_occ, occ, sizeof(double) * nb, hipMemcpyHostToDevice, st));
    const T *phase = phase_of<T>(kp);
    const bool small = p.max_m <= 9;
    if (ncomp == 1) {
        if (small) launch_force<T, 9, 1, false>(ctx, d_out, psi, nb, d_occ, phase, st);
        else launch_force<T, 16, 1, false>(ctx, d_out, psi, nb, d_occ, phase, st);
    } else if (!ph) {
        if (small) launch_force<T, 9, 2, false>(ctx, d_out, psi, nb, d_occ, phase, st);
        else launch_force<T, 16, 2, false>(ctx, d_out, psi, nb, d_occ, phase, st);
    } else {
        if (small) launch_force<T, 9, 2, true>(ctx, d_out, psi, nb, d_occ, phase, st);
        else launch_force<T, 16, 2, true>(ctx, d_out, psi, nb, d_occ, phase, st);
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
int nonlocal_forces_f64(qm_ctx *ctx, const double *psi, const double *occ, int nb, int ncomp, const KPoint &kp,
                        double *forces, double *e_nl, hipStream_t st) {
    return nonlocal_forces<double>(ctx, psi, occ, nb, ncomp, kp, forces, e_nl, st);
}
int nonlocal_forces_f32(qm_ctx *ctx, const float *psi, const double *occ, int nb, int ncomp, const KPoint &kp,
                        double *forces, double *e_nl, hipStream_t st) {
    return nonlocal_forces<float>(ctx, psi, occ, nb, ncomp, kp, forces, e_nl, st);
}

}  // namespace qm
