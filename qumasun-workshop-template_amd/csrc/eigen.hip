// Dense FP64 symmetric eigensolver for the LOBPCG subspace problem.
//
// The generalized problem A z = l B z of lapack_DSYGVD (w_dsygvd.h:14-72) and
// lapack_ZHEGVD_emu (w_zhegvd.h:227-272) is solved by the Cholesky route:
// B = U^T U, C = U^-T A U^-1, C = Q T Q^T (tridiagonal), T = Y L Y^T, z = U^-1 Q Y.
//
// The tridiagonal reduction is this file's own: rocSOLVER's sytrd launches
// about ten small kernels per column, which at the subspace sizes of LOBPCG
// (n = 2N or 3N, a few hundred) costs ~15 us per column in launch latency
// alone.  Here one workgroup of 1024 threads reduces a panel of 16 columns in
// one launch (the blocked scheme of LAPACK's dlatrd: each column is updated
// with the panel's earlier reflectors, its Householder vector is formed, and
// w = tau (A22 v - V W^T v - W V^T v) is accumulated against the trailing
// matrix as it stood at the panel's start), with the panel's V and W held in
// LDS; one rocBLAS GEMM then applies the rank-32 update A22 -= V W^T + W V^T,
// and the last <= 128 columns are reduced unblocked with the trailing block in
// LDS.  Output follows LAPACK's lower-storage convention (d, e, tau,
// reflectors below the subdiagonal), so rocSOLVER's stedc (divide and
// conquer on T) and ormtr (apply Q) consume it.  The matrix is kept symmetric
// in both triangles (the upper one is mirrored first), so the matrix-vector
// product is a plain column sweep.  Matrices above 512 go to rocSOLVER syevd.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>

#include "qm_internal.h"

namespace qm {
namespace {

constexpr int kNB = 16;         // panel width
constexpr int kTailMax = 128;   // columns left for the LDS kernel (128^2 doubles = 131072 bytes)
constexpr int kThreads = 1024;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxN = 512;      // the panel's V and W (n x 16 each) live in LDS

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the block in a fixed order; every thread receives the total.  One
// barrier: `red` may be written again only after another barrier, so the two
// sums of a column (Householder norm, alpha) use separate buffers.
__device__ double block_sum(double v, double *red) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < kWaves; ++i) t += red[i];
    return t;
}

// Householder reflector of LAPACK's dlarfg for x = xs[0..m-1]: H x = beta e_1,
// H = I - tau v v^T, v_0 = 1.  Returns tau and beta, overwrites xs with v.
__device__ void householder(double *xs, int m, double *red, double &tau, double &beta) {
    const double alpha = xs[0];
    double loc = 0.0;
    for (int k = 1 + threadIdx.x; k < m; k += kThreads) loc += xs[k] * xs[k];
    const double xn2 = block_sum(loc, red);  // all reads of xs[0] happen before its sync
    double scal = 1.0;
    if (xn2 == 0.0) {
        tau = 0.0;
        beta = alpha;
    } else {
        beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
        tau = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
    }
    for (int k = threadIdx.x; k < m; k += kThreads) xs[k] = k == 0 ? 1.0 : xs[k] * scal;
    __syncthreads();
}

// Columns i0 .. i0+kNB-1 of the symmetric n x n matrix A (column-major, both
// triangles valid on entry), n - i0 <= kMaxN.  The panel's V and W stay in LDS
// (column q of V at Vs[q][r - i0]); on exit VW holds [V | W | V] (n x kNB
// blocks, rows i0 .. n-1) for the trailing update.
__global__ __launch_bounds__(kThreads) void trd_panel_kernel(double *__restrict__ A, int lda, int n, int i0,
                                                             double *__restrict__ d, double *__restrict__ e,
                                                             double *__restrict__ tau, double *__restrict__ VW,
                                                             int ldp) {
    // w = A22 v is split over kCG column groups x kRQ row quarters (one wave
    // each); a lane keeps kRC row accumulators and loads 8 columns per round
    constexpr int kCG = 4;
    constexpr int kRQ = kWaves / kCG;
    constexpr int kRC = kMaxN / (64 * kRQ);
    __shared__ double Vs[kNB][kMaxN], Ws[kNB][kMaxN];
    __shared__ double vs[kMaxN];
    __shared__ double part[kCG][kMaxN];
    __shared__ double tt[2 * kNB];
    __shared__ double red[kWaves], red2[kWaves];
    const int tid = threadIdx.x;
    const int wv = tid >> 6, ln = tid & 63;
    const int np = n - i0;
    for (int idx = tid; idx < kNB * kMaxN; idx += kThreads) {
        (&Vs[0][0])[idx] = 0.0;
        (&Ws[0][0])[idx] = 0.0;
    }
    __syncthreads();
    for (int j = 0; j < kNB; ++j) {
        const int c = i0 + j;
        const int m = n - c - 1;
        // A(c:n, c) -= V(c:n, :j) W(c, :j)^T + W(c:n, :j) V(c, :j)^T
        for (int rl = j + tid; rl < np; rl += kThreads) {
            double a = A[(i0 + rl) + (size_t)c * lda];
            for (int q = 0; q < j; ++q) a = fma(-Vs[q][rl], Ws[q][j], fma(-Ws[q][rl], Vs[q][j], a));
            if (rl == j) d[c] = a;
            else vs[rl - j - 1] = a;
        }
        __syncthreads();
        double t, beta;
        householder(vs, m, red, t, beta);
        for (int k = tid; k < m; k += kThreads) {
            Vs[j][j + 1 + k] = vs[k];
            A[c + 1 + k + (size_t)c * lda] = k == 0 ? beta : vs[k];
        }
        if (tid == 0) {
            e[c] = beta;
            tau[c] = t;
        }
        // w = A22 v, A22 = A(c+1:n, c+1:n) as at the panel start
        {
            const int g = wv % kCG, rq = wv / kCG;
            const int k0 = (int)((int64_t)m * g / kCG), k1 = (int)((int64_t)m * (g + 1) / kCG);
            const int rbeg = (int)((int64_t)m * rq / kRQ), rend = (int)((int64_t)m * (rq + 1) / kRQ);
            const double *base = A + (c + 1) + (size_t)(c + 1) * lda;
            double acc[kRC];
            bool act[kRC];
#pragma unroll
            for (int i = 0; i < kRC; ++i) {
                acc[i] = 0.0;
                act[i] = rbeg + ln + 64 * i < rend;
            }
            int k = k0;
            for (; k + 7 < k1; k += 8) {
                double x[8], a[8][kRC];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    x[u] = vs[k + u];
                    const double *col = base + (size_t)(k + u) * lda + rbeg + ln;
#pragma unroll
                    for (int i = 0; i < kRC; ++i) {
                        a[u][i] = 0.0;
                        if (act[i]) a[u][i] = col[64 * i];  // rows past the quarter are not read
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int i = 0; i < kRC; ++i) acc[i] = fma(a[u][i], x[u], acc[i]);
            }
            for (; k < k1; ++k) {
                const double x0 = vs[k];
                const double *col = base + (size_t)k * lda + rbeg + ln;
#pragma unroll
                for (int i = 0; i < kRC; ++i)
                    if (act[i]) acc[i] = fma(col[64 * i], x0, acc[i]);
            }
#pragma unroll
            for (int i = 0; i < kRC; ++i)
                if (act[i]) part[g][rbeg + ln + 64 * i] = acc[i];
        }
        // t1 = W(c+1:n, :j)^T v, t2 = V(c+1:n, :j)^T v: one wave per dot product
        for (int q2 = wv; q2 < 2 * j; q2 += kWaves) {
            const double *col = (q2 < j ? &Ws[q2][0] : &Vs[q2 - j][0]) + j + 1;
            double a0 = 0.0;
            for (int k = ln; k < m; k += 64) a0 = fma(col[k], vs[k], a0);
            const double acc = wave_sum(a0);
            if (ln == 0) tt[q2] = acc;
        }
        __syncthreads();
        double loc = 0.0, wr[kMaxN / kThreads + 1];
        int nr = 0;
        for (int r = tid; r < m; r += kThreads) {
            double w = 0.0;
#pragma unroll
            for (int g = 0; g < kCG; ++g) w += part[g][r];
            for (int q = 0; q < j; ++q) w = fma(-Vs[q][j + 1 + r], tt[q], fma(-Ws[q][j + 1 + r], tt[j + q], w));
            w *= t;
            wr[nr++] = w;
            loc = fma(w, vs[r], loc);
        }
        const double alpha = -0.5 * t * block_sum(loc, red2);
        nr = 0;
        for (int r = tid; r < m; r += kThreads) Ws[j][j + 1 + r] = fma(alpha, vs[r], wr[nr++]);
        __syncthreads();
    }
    // [V | W | V] for the trailing rank-2 kNB update
    for (int idx = tid; idx < kNB * np; idx += kThreads) {
        const int q = idx / np, rl = idx % np;
        const size_t row = (size_t)(i0 + rl);
        VW[row + (size_t)q * ldp] = Vs[q][rl];
        VW[row + (size_t)(kNB + q) * ldp] = Ws[q][rl];
        VW[row + (size_t)(2 * kNB + q) * ldp] = Vs[q][rl];
    }
}

// Cholesky A = U^T U (upper triangle, column-major, n <= kMaxN) in one
// 1024-thread workgroup, right-looking over blocks of kCB columns:
//   - the diagonal block is factored in LDS (two barriers per column);
//   - the block row right of it is solved, U_kk^T X = A(k, k+1:), one thread
//     per column, and X is kept in LDS;
//   - A(k+1:, k+1:) -= X^T X on the upper triangle, 4 x 4 tiles per thread.
// rocSOLVER's potrf spends ~1 ms on the subspace sizes of LOBPCG (n = N and
// 3N) in small single-block kernels; this takes one launch.  The strictly
// lower triangle is never touched.  *info = 0, or the order of the first
// leading minor whose pivot is not positive (or NaN), where the factorisation
// stops as LAPACK's dpotrf (upper) does.
constexpr int kCB = 32;
__global__ __launch_bounds__(kThreads) void potrf_upper_kernel(double *__restrict__ A, int lda, int n,
                                                               int *__restrict__ info) {
    __shared__ double Ds[kCB][kCB + 1];  // diagonal block, Ds[row][col], upper triangle
    __shared__ double rinv[kCB];         // 1 / U_jj of the block
    __shared__ double Xs[kCB][kMaxN];    // block row of U right of the diagonal block
    const int tid = threadIdx.x;
    if (tid == 0) *info = 0;
    for (int k0 = 0; k0 < n; k0 += kCB) {
        const int b = min(kCB, n - k0);
        const int nr = n - k0 - b;
        for (int e = tid; e < b * b; e += kThreads) {
            const int r = e % b, c = e / b;
            if (r <= c) Ds[r][c] = A[(k0 + r) + (size_t)(k0 + c) * lda];
        }
        __syncthreads();
        for (int j = 0; j < b; ++j) {
            const double d = Ds[j][j];  // every thread reads the same pivot: one branch for all
            if (!(d > 0.0)) {
                if (tid == 0) *info = k0 + j + 1;
                return;
            }
            const double ujj = sqrt(d), inv = 1.0 / ujj;
            if (j + 1 + tid < b) Ds[j][j + 1 + tid] *= inv;  // row j of U
            __syncthreads();
            const int m = b - j - 1;
            if (tid < m * m) {
                const int i = j + 1 + tid % m, l = j + 1 + tid / m;
                if (i <= l) Ds[i][l] = fma(-Ds[j][i], Ds[j][l], Ds[i][l]);
            }
            if (tid == 0) {
                Ds[j][j] = ujj;
                rinv[j] = inv;
            }
            __syncthreads();
        }
        for (int e = tid; e < b * b; e += kThreads) {
            const int r = e % b, c = e / b;
            if (r <= c) A[(k0 + r) + (size_t)(k0 + c) * lda] = Ds[r][c];
        }
        if (nr == 0) break;
        // U_kk^T X = A(k0:k0+b, k0+b:n): forward substitution per column
        if (tid < nr) {
            double *col = A + k0 + (size_t)(k0 + b + tid) * lda;
            double x[kCB];
#pragma unroll
            for (int p = 0; p < kCB; ++p) x[p] = p < b ? col[p] : 0.0;
#pragma unroll
            for (int p = 0; p < kCB; ++p) {
                if (p < b) {
                    x[p] *= rinv[p];
#pragma unroll
                    for (int q = p + 1; q < kCB; ++q)
                        if (q < b) x[q] = fma(-Ds[p][q], x[p], x[q]);
                }
            }
#pragma unroll
            for (int p = 0; p < kCB; ++p)
                if (p < b) {
                    col[p] = x[p];
                    Xs[p][tid] = x[p];
                }
        }
        __syncthreads();
        // A(k0+b:n, k0+b:n) -= X^T X, upper triangle, 4 x 4 tiles (ti <= tl)
        const int nt = (nr + 3) / 4;
        const int ntiles = nt * (nt + 1) / 2;
        double *T = A + (k0 + b) + (size_t)(k0 + b) * lda;
        for (int q = tid; q < ntiles; q += kThreads) {
            int tl = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
            while (tl * (tl + 1) / 2 > q) --tl;
            while ((tl + 1) * (tl + 2) / 2 <= q) ++tl;
            const int ti = q - tl * (tl + 1) / 2;
            double acc[4][4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;
            for (int p = 0; p < b; ++p) {
                double xi[4], xl[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    xi[a] = 4 * ti + a < nr ? Xs[p][4 * ti + a] : 0.0;
                    xl[a] = 4 * tl + a < nr ? Xs[p][4 * tl + a] : 0.0;
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[a][c] = fma(xi[a], xl[c], acc[a][c]);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int l = 4 * tl + c;
                if (l >= nr) continue;
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int i = 4 * ti + a;
                    if (i <= l) T[i + (size_t)l * lda] -= acc[a][c];
                }
            }
        }
        __syncthreads();
    }
}

// Unblocked reduction of the trailing block A(it:n, it:n), nt = n - it <= kTailMax,
// held in LDS (LAPACK's dsytd2, lower).
__global__ __launch_bounds__(kThreads) void trd_tail_kernel(double *__restrict__ A, int lda, int n, int it,
                                                            double *__restrict__ d, double *__restrict__ e,
                                                            double *__restrict__ tau) {
    __shared__ double L[kTailMax * kTailMax];
    __shared__ double vs[kTailMax], ws[kTailMax];
    __shared__ double red[kWaves], red2[kWaves];
    constexpr int kKG = kThreads / 128;  // k groups of the matrix-vector product
    __shared__ double wpart[kKG][128];
    const int tid = threadIdx.x;
    const int nt = n - it;
    for (int idx = tid; idx < nt * nt; idx += kThreads) {
        const int r = idx % nt, k = idx / nt;
        L[idx] = A[(it + r) + (size_t)(it + k) * lda];
    }
    __syncthreads();
    for (int j = 0; j < nt - 1; ++j) {
        const int c = it + j;
        const int m = nt - j - 1;
        for (int k = tid; k < m; k += kThreads) vs[k] = L[(j + 1 + k) + j * nt];
        if (tid == 0) d[c] = L[j + j * nt];
        __syncthreads();
        double t, beta;
        householder(vs, m, red, t, beta);
        for (int k = tid; k < m; k += kThreads) A[c + 1 + k + (size_t)c * lda] = k == 0 ? beta : vs[k];
        if (tid == 0) {
            e[c] = beta;
            tau[c] = t;
        }
        // w = tau A22 v, alpha = -1/2 tau w.v, w += alpha v, A22 -= v w^T + w v^T.
        // The product: rows over 128 threads, the k range split over kKG groups
        // of them (a dependent chain of m / kKG instead of m), partials summed in
        // a fixed order.
        {
            const int r = tid & 127, grp = tid >> 7;
            const int k0 = (m * grp) / kKG, k1 = (m * (grp + 1)) / kKG;
            double acc = 0.0;
            if (r < m)
                for (int k = k0; k < k1; ++k) acc = fma(L[(j + 1 + r) + (j + 1 + k) * nt], vs[k], acc);
            wpart[grp][r] = acc;
        }
        __syncthreads();
        double loc = 0.0;
        for (int r = tid; r < m; r += kThreads) {
            double acc = 0.0;
#pragma unroll
            for (int g = 0; g < kKG; ++g) acc += wpart[g][r];
            acc *= t;
            ws[r] = acc;
            loc = fma(acc, vs[r], loc);
        }
        const double alpha = -0.5 * t * block_sum(loc, red2);
        for (int r = tid; r < m; r += kThreads) ws[r] = fma(alpha, vs[r], ws[r]);
        __syncthreads();
        for (int idx = tid; idx < m * m; idx += kThreads) {
            const int r = idx % m, k = idx / m;
            double &a = L[(j + 1 + r) + (j + 1 + k) * nt];
            a = fma(-vs[r], ws[k], a);
            a = fma(-ws[r], vs[k], a);
        }
        __syncthreads();
    }
    if (tid == 0) {
        d[n - 1] = L[(nt - 1) + (nt - 1) * nt];
        tau[n - 1] = 0.0;
        e[n - 1] = 0.0;
    }
}

// C(i, j) = C(j, i) for i > j: the upper triangle defines the matrix, as in
// LAPACK's uplo = 'U' (a product like U^-T H U^-1 is symmetric only up to
// rounding, and the reduction below reads both triangles)
// complex n x n column-major (interleaved re, im): the strictly lower triangle
// := the conjugate of the upper one, the diagonal's imaginary part := 0
__global__ void hermitian_from_upper_kernel(double *__restrict__ C, int n) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e % n), j = (int)(e / n);
    if (i > j) {
        C[2 * e] = C[2 * ((int64_t)j + (int64_t)i * n)];
        C[2 * e + 1] = -C[2 * ((int64_t)j + (int64_t)i * n) + 1];
    } else if (i == j) {
        C[2 * e + 1] = 0.0;
    }
}

// strictly lower triangle of the n x n column-major M set to zero
__global__ void zero_lower_kernel(double *__restrict__ M, int n, int ld) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int r = (int)(e % n), c = (int)(e / n);
    if (r > c) M[r + (size_t)c * ld] = 0.0;
}

__global__ void mirror_upper_kernel(double *__restrict__ C, int n, int lda) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)n * n) return;
    const int i = (int)(idx % n), j = (int)(idx / n);
    if (i > j) C[i + (int64_t)j * lda] = C[j + (int64_t)i * lda];
}

#define QM_EIG_BLAS(call)                                                                                   \
    do {                                                                                                    \
        rocblas_status s_ = (call);                                                                         \
        if (s_ != rocblas_status_success) {                                                                 \
            ::qm::set_error(std::string(#call) + " failed with rocblas status " + std::to_string((int)s_)); \
            return QM_ERR_HIP;                                                                              \
        }                                                                                                   \
    } while (0)

}  // namespace

// C = Q T Q^T for the symmetric n x n matrix C (column-major, lda >= n, both
// triangles; overwritten with the reflectors).  trd holds d, e, tau (3 n);
// VW is n x 48 workspace.  n <= 512.  Runs on the handle's stream.
int sytrd_lower(void *blas, int n, double *C, int lda, double *trd, double *VW, hipStream_t st) {
    rocblas_handle h = (rocblas_handle)blas;
    static_assert(3 * kNB <= kTrdPanelCols, "VW workspace width");
    QM_CHECK(n >= 1 && n <= kMaxN, QM_ERR_ARG, "tridiagonal reduction: n must be in [1, 512]");
    double *d = trd, *e = trd + n, *tau = trd + 2 * n;
    int i0 = 0;
    const double minus1 = -1.0, one = 1.0;
    while (n - i0 > kTailMax) {
        hipLaunchKernelGGL(trd_panel_kernel, dim3(1), dim3(kThreads), 0, st, C, lda, n, i0, d, e, tau, VW, n);
        QM_HIP(hipGetLastError());
        const int m2 = n - i0 - kNB;
        QM_EIG_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, m2, m2, 2 * kNB, &minus1,
                                  VW + i0 + kNB, n, VW + (size_t)kNB * n + i0 + kNB, n, &one,
                                  C + (i0 + kNB) + (size_t)(i0 + kNB) * lda, lda));
        i0 += kNB;
    }
    hipLaunchKernelGGL(trd_tail_kernel, dim3(1), dim3(kThreads), 0, st, C, lda, n, i0, d, e, tau);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

// Eigenvalues (all, ascending, into w) and the eigenvectors of the nev lowest
// into Z[:, :nev] (ldz >= n) of the symmetric matrix whose upper triangle C
// holds (C overwritten).
// Where the cooperative reduction fits (trd_coop.hip): that reduction,
// rocSOLVER stedc + ormtr; else n <= kMaxN: the single-workgroup panel
// reduction below, stedc + ormtr (Z must hold n x n in both); larger n:
// rocSOLVER syevd.  info receives rocSOLVER's status (or the cooperative
// reduction's barrier error, -77).  trd holds 4 n doubles (d, e, tau, and a
// copy of e for the steqr retry after a stedc failure).
int syev_lower(void *blas, int n, int nev, double *C, int lda, double *w, double *Z, int ldz, double *trd,
               double *VW, int *info, hipStream_t st) {
    rocblas_handle h = (rocblas_handle)blas;
    if (subspace_route() != 0 && trd_coop_supported(n, false) &&
        trd_coop_work_bytes(n, false) <= sizeof(double) * (size_t)kTrdPanelCols * n) {
        const int r = trd_coop_f64(n, C, lda, trd, trd + n, trd + 2 * n, VW,
                                   sizeof(double) * (size_t)kTrdPanelCols * n, st);
        if (r == QM_OK) {
            double *e_copy = trd + 3 * n;  // stedc overwrites e
            QM_HIP(hipMemcpyAsync(e_copy, trd + n, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
            QM_HIP(hipMemcpyAsync(w, trd, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
            QM_EIG_BLAS(rocsolver_dstedc(h, rocblas_evect_tridiagonal, n, w, trd + n, Z, ldz, info));
            int h_info = 0;
            QM_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, st));
            QM_HIP(hipStreamSynchronize(st));
            if (h_info != 0) {
                // divide and conquer did not converge (it can on nearly degenerate
                // spectra): implicit QL/QR on the same tridiagonal instead
                QM_HIP(hipMemcpyAsync(w, trd, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
                QM_EIG_BLAS(rocsolver_dsteqr(h, rocblas_evect_tridiagonal, n, w, e_copy, Z, ldz, info));
            }
            QM_TRY(trd_coop_check(VW, info, st));
            QM_TRY(ormtr_lower_f64(n, nev, C, lda, trd + 2 * n, Z, ldz, VW, st));
            return QM_OK;
        }
        if (r != QM_ERR_STATE) return r;
    }
    if (n > kMaxN) {
        QM_EIG_BLAS(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, C, lda, w, trd, info));
        QM_HIP(hipMemcpy2DAsync(Z, sizeof(double) * ldz, C, sizeof(double) * lda, sizeof(double) * n, nev,
                                hipMemcpyDeviceToDevice, st));
        return QM_OK;
    }
    hipLaunchKernelGGL(mirror_upper_kernel, dim3((unsigned)(((int64_t)n * n + 255) / 256)), dim3(256), 0, st, C, n,
                       lda);
    QM_HIP(hipGetLastError());
    QM_TRY(sytrd_lower(blas, n, C, lda, trd, VW, st));
    QM_HIP(hipMemcpyAsync(w, trd, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    QM_EIG_BLAS(rocsolver_dstedc(h, rocblas_evect_tridiagonal, n, w, trd + n, Z, ldz, info));
    QM_EIG_BLAS(rocsolver_dormtr(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, n, nev, C, lda,
                                 trd + 2 * n, Z, ldz));
    return QM_OK;
}

// C := U^-T C U^-1 for the upper Cholesky factor U (n x n, all column-major,
// leading dimension n) through Ui = U^-1 (rocBLAS trtri, its strictly lower
// part cleared) and two GEMMs: at the LOBPCG subspace size (n = 384) this is
// 0.18 ms against 0.25 ms for two triangular solves.  Ui is kept for
// chol_back; T is n x n work.
int chol_transform(void *blas, int n, const double *U, double *C, double *Ui, double *T, hipStream_t st) {
    rocblas_handle h = (rocblas_handle)blas;
    const double one = 1.0, zero = 0.0;
    QM_EIG_BLAS(rocblas_dtrtri(h, rocblas_fill_upper, rocblas_diagonal_non_unit, n, U, n, Ui, n));
    hipLaunchKernelGGL(zero_lower_kernel, dim3((unsigned)(((int64_t)n * n + 255) / 256)), dim3(256), 0, st, Ui, n, n);
    QM_HIP(hipGetLastError());
    QM_EIG_BLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, n, n, n, &one, Ui, n, C, n,
                              &zero, T, n));
    QM_EIG_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, T, n, Ui, n, &zero,
                              C, n));
    return QM_OK;
}

// Zout = Ui Z (n x nev; Z with leading dimension ldz, Zout with ldo)
int chol_back(void *blas, int n, int nev, const double *Ui, const double *Z, int ldz, double *Zout, int ldo) {
    rocblas_handle h = (rocblas_handle)blas;
    const double one = 1.0, zero = 0.0;
    QM_EIG_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, nev, n, &one, Ui, n, Z, ldz,
                              &zero, Zout, ldo));
    return QM_OK;
}

int potrf_upper(void *blas, int n, double *A, int lda, int *info, hipStream_t st, void *sync) {
    // one workgroup per 32-row panel where the caller gives the barrier words
    // (and the subspace route is not forced onto the single-workgroup kernels)
    if (sync && subspace_route() < 0 && potrf_coop_supported(n)) {
        const int rc = potrf_coop_f64(n, A, lda, info, sync, st);
        if (rc != QM_ERR_STATE) return rc;
    }
    if (n > kMaxN) {
        QM_EIG_BLAS(rocsolver_dpotrf((rocblas_handle)blas, rocblas_fill_upper, n, A, lda, info));
        return QM_OK;
    }
    hipLaunchKernelGGL(potrf_upper_kernel, dim3(1), dim3(kThreads), 0, st, A, lda, n, info);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

int own_eig_max_n() { return kMaxN; }

}  // namespace qm

namespace qm {
namespace {

// C = A with the upper triangle mirrored into the lower one
__global__ void symmetrize_upper_kernel(double *__restrict__ C, const double *__restrict__ A, int n) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)n * n) return;
    const int r = (int)(idx % n), k = (int)(idx / n);
    C[idx] = r <= k ? A[idx] : A[k + (int64_t)r * n];
}

}  // namespace
}  // namespace qm

using namespace qm;

namespace qm {
namespace {
int g_subspace_route = -1;
}
int subspace_route() { return g_subspace_route; }
}  // namespace qm

extern "C" {

int qm_subspace_force(int route) {
    QM_CHECK(route >= -1 && route <= 1, QM_ERR_ARG, "qm_subspace_force: route must be -1, 0 or 1");
    g_subspace_route = route;
    return QM_OK;
}

int qm_sygvd_f64(qm_ctx *ctx, int n, const double *d_A, const double *d_B, double *d_w, double *d_Z, double *d_U,
                 int nev, void *stream) {
    QM_CHECK(ctx && d_A && d_B && d_w && d_Z && d_U, QM_ERR_ARG, "qm_sygvd_f64: NULL argument");
    QM_CHECK(n >= 1 && n <= 4096, QM_ERR_ARG, "qm_sygvd_f64: n must be in [1, 4096]");
    QM_CHECK(nev >= 1 && nev <= n, QM_ERR_ARG, "qm_sygvd_f64: nev must be in [1, n]");
    QM_TRY(activate(ctx));
    hipStream_t st = (hipStream_t)stream;
    if (!ctx->blas) {
        rocblas_handle h = nullptr;
        QM_CHECK(rocblas_create_handle(&h) == rocblas_status_success, QM_ERR_HIP,
                 "qm_sygvd_f64: rocblas_create_handle failed");
        ctx->blas = h;
    }
    rocblas_handle h = (rocblas_handle)ctx->blas;
    QM_EIG_BLAS(rocblas_set_stream(h, st));
    // workspace: C (n^2), Z (n^2), d e tau + e copy (4n), VW (96 n), U^-1 and T (n^2 each), info
    const size_t nn = (size_t)n * n;
    const size_t need = sizeof(double) * (4 * nn + 4 * (size_t)n + (size_t)kTrdPanelCols * n) + 256 + 512;
    QM_TRY(ctx->eig_work.reserve(need));
    double *C = ctx->eig_work.as<double>(), *Z = C + nn, *trd = Z + nn, *VW = trd + 4 * n;
    double *Ui = VW + (size_t)kTrdPanelCols * n, *T = Ui + nn;
    int *info = (int *)(T + nn);
    void *psync = reinterpret_cast<char *>(T + nn) + 256;  // the cooperative Cholesky's barrier words
    hipLaunchKernelGGL(symmetrize_upper_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, C, d_A, n);
    QM_HIP(hipGetLastError());
    QM_HIP(hipMemcpyAsync(d_U, d_B, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
    QM_TRY(potrf_upper(h, n, d_U, n, info, st, psync));
    int h_info = 0;
    QM_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, st));
    QM_HIP(hipStreamSynchronize(st));
    QM_CHECK(h_info != -77, QM_ERR_NUMERIC, "qm_sygvd_f64: the cooperative Cholesky's grid barrier timed out (info -77)");
    QM_CHECK(h_info == 0, QM_ERR_NUMERIC,
             "qm_sygvd_f64: B is not positive definite (leading minor " + std::to_string(h_info) + ")");
    QM_TRY(chol_transform(h, n, d_U, C, Ui, T, st));
    QM_TRY(syev_lower(h, n, nev, C, n, d_w, Z, n, trd, VW, info, st));
    QM_TRY(chol_back(h, n, nev, Ui, Z, n, d_Z, n));
    QM_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, st));
    QM_HIP(hipStreamSynchronize(st));
    QM_CHECK(h_info == 0, QM_ERR_NUMERIC,
             "qm_sygvd_f64: tridiagonal eigensolve failed (info " + std::to_string(h_info) +
                 ": rocSOLVER stedc, or -77 where the cooperative reduction's grid barrier timed out)");
    return QM_OK;
}

int qm_hegvd_z64(qm_ctx *ctx, int n, const void *d_A, const void *d_B, double *d_w, void *d_Z, void *d_U, int nev,
                 void *stream) {
    QM_CHECK(ctx && d_A && d_B && d_w && d_Z && d_U, QM_ERR_ARG, "qm_hegvd_z64: NULL argument");
    QM_CHECK(n >= 1 && n <= 4096, QM_ERR_ARG, "qm_hegvd_z64: n must be in [1, 4096]");
    QM_CHECK(nev >= 1 && nev <= n, QM_ERR_ARG, "qm_hegvd_z64: nev must be in [1, n]");
    QM_TRY(activate(ctx));
    hipStream_t st = (hipStream_t)stream;
    if (!ctx->blas) {
        rocblas_handle h = nullptr;
        QM_CHECK(rocblas_create_handle(&h) == rocblas_status_success, QM_ERR_HIP,
                 "qm_hegvd_z64: rocblas_create_handle failed");
        ctx->blas = h;
    }
    rocblas_handle h = (rocblas_handle)ctx->blas;
    QM_EIG_BLAS(rocblas_set_stream(h, st));
    using zc = rocblas_double_complex;
    // workspace: C and Z (n^2 complex each), d, e, tau (complex), e copy (5n doubles),
    // the cooperative reduction's work, info
    const size_t nn = (size_t)n * n;
    const size_t coopb = trd_coop_work_bytes(n, true);
    const size_t need = sizeof(zc) * 2 * nn + sizeof(double) * 5 * (size_t)n + coopb + 512;
    QM_TRY(ctx->eig_work.reserve(need));
    zc *C = ctx->eig_work.as<zc>(), *Z = C + nn;
    double *d = reinterpret_cast<double *>(Z + nn), *e = d + n, *e_copy = d + 4 * (size_t)n;
    zc *tau = reinterpret_cast<zc *>(d + 2 * (size_t)n);
    void *coop = d + 5 * (size_t)n;
    int *info = reinterpret_cast<int *>(static_cast<char *>(coop) + coopb);
    zc *U = static_cast<zc *>(d_U);
    QM_HIP(hipMemcpyAsync(C, d_A, sizeof(zc) * nn, hipMemcpyDeviceToDevice, st));
    QM_HIP(hipMemcpyAsync(U, d_B, sizeof(zc) * nn, hipMemcpyDeviceToDevice, st));
    QM_EIG_BLAS(rocsolver_zpotrf(h, rocblas_fill_upper, n, U, n, info));
    int h_info = 0;
    QM_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, st));
    QM_HIP(hipStreamSynchronize(st));
    QM_CHECK(h_info == 0, QM_ERR_NUMERIC,
             "qm_hegvd_z64: B is not positive definite (leading minor " + std::to_string(h_info) + ")");
    const zc one{1.0, 0.0};
    // C = U^-H A U^-1 (the upper triangle defines A)
    hipLaunchKernelGGL(hermitian_from_upper_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<double *>(C), n);
    QM_HIP(hipGetLastError());
    QM_EIG_BLAS(rocblas_ztrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_conjugate_transpose,
                              rocblas_diagonal_non_unit, n, n, &one, U, n, C, n));
    QM_EIG_BLAS(rocblas_ztrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                              rocblas_diagonal_non_unit, n, n, &one, U, n, C, n));
    int r = QM_ERR_STATE;
    if (subspace_route() != 0 && trd_coop_supported(n, true))
        r = trd_coop_z64(n, C, n, d, e, tau, coop, coopb, st);
    if (r == QM_OK) {
        QM_HIP(hipMemcpyAsync(e_copy, e, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        QM_HIP(hipMemcpyAsync(d_w, d, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        QM_EIG_BLAS(rocsolver_zstedc(h, rocblas_evect_tridiagonal, n, d_w, e, Z, n, info));
        QM_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, st));
        QM_HIP(hipStreamSynchronize(st));
        if (h_info != 0) {  // divide and conquer did not converge: QL/QR instead
            QM_HIP(hipMemcpyAsync(d_w, d, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
            QM_EIG_BLAS(rocsolver_zsteqr(h, rocblas_evect_tridiagonal, n, d_w, e_copy, Z, n, info));
        }
        QM_TRY(trd_coop_check(coop, info, st));
        QM_TRY(unmtr_lower_z64(n, nev, C, n, tau, Z, n, coop, st));
    } else {
        if (r != QM_ERR_STATE) return r;
        QM_EIG_BLAS(rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, n, C, n, d_w, e, info));
        QM_HIP(hipMemcpyAsync(Z, C, sizeof(zc) * n * nev, hipMemcpyDeviceToDevice, st));
    }
    // z = U^-1 y for the nev lowest
    QM_EIG_BLAS(rocblas_ztrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                              rocblas_diagonal_non_unit, n, nev, &one, U, n, Z, n));
    QM_HIP(hipMemcpyAsync(d_Z, Z, sizeof(zc) * n * nev, hipMemcpyDeviceToDevice, st));
    QM_HIP(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, st));
    QM_HIP(hipStreamSynchronize(st));
    QM_CHECK(h_info == 0, QM_ERR_NUMERIC,
             "qm_hegvd_z64: tridiagonal eigensolve failed (rocSOLVER info " + std::to_string(h_info) + ")");
    return QM_OK;
}

}  // extern "C"
