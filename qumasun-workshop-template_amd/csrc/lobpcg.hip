// Block LOBPCG on one GPU (replaces LOBPCG::Eigen_z_multi_mpi,
// lobpcg_z_multi_mpi.h:1732-2852).
//
// Per iteration (all on the GPU, FP64 unless noted):
//   A R               H.psi of the residual block (T = FP64 or FP32)
//   G1 = [X R P]^T [R AR], and on recompute iterations G2 = [X P]^T [X P AX AP]
//                     tall-skinny MFMA Gram products (T inputs, FP64 result)
//   S, H              assembled from G1, G2 or the subspace recurrence, with R
//                     normalised by the residual norms; complex blocks packed from
//                     the real Gram of the SoA columns (AoSComplexFrom2by2Real,
//                     matrix_2x2real_to_complex.h:8-56)
//   H c = l S c       the Cholesky route of lapack_ZHEGVD_emu (w_zhegvd.h:227-272;
//                     lobpcg_z_multi_mpi.h:2254-2263, 2492-2511): real n <= 512 by
//                     the own cooperative Cholesky, trtri + GEMMs; real and complex
//                     by the own cooperative tridiagonal reduction (trd_coop.hip),
//                     rocSOLVER stedc and the own back-transformation of the
//                     wanted eigenvectors; complex Cholesky and the problems the
//                     cooperative reduction does not fit by rocSOLVER potrf /
//                     trsm / syevd or heevd
//   recurrence        S' = Cb^H S Cb, H' = Cb^H H Cb with Cb = [C | C_P] (rocBLAS;
//                     NextMatrix_z_blasmp, 1104-1436); P' normalised by its norm
//   update            [X' P'] = [X R P] U, [AX' AP'] = [AX AR AP] U (MFMA, T)
//   residual          R' = AX' - X' L with |R'|^2, |X'|^2 (FP64 sums); the
//                     |1 - |x'|^2| > 1e-3 rollback of 2611-2666
// Gram blocks are recomputed from the vectors at iteration 0 and every 6th
// iteration (S_MAT_REST_LIMIT, 1808); R blocks every iteration.
// Locking (qm_lobpcg_solve_locking; 1962-1986, 1637-1723): the lowest num_fix
// bands are held fixed and everything above runs on the remaining na bands,
// with R (and P when num_fix changes) orthogonalised against the fixed bands.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <tuple>
#include <utility>

#include <cstdio>

#include "qm_internal.h"

namespace qm {
namespace {

constexpr int kRestLimit = 6;          // S_MAT_REST_LIMIT
constexpr double kLimitNormX = 1.0e-3; // LIMIT_Z_NORM_X_FOR_RESET_MATRIX
// Ritz vectors of the Cholesky route are accepted when max_i |1 - c_i^H S c_i|
// stays below this (S-orthonormality; an S too ill-conditioned for its
// Cholesky factor breaks it long before the pivot ratio shows it)
constexpr double kRitzOrthTol = 1.0e-4;

// ------------------------------------------------------------------ small FP64 matrices
// Complex elements are interleaved (re, im) pairs, as rocblas_double_complex.

struct Cplx {
    double re, im;
};

template <bool CPLX>
struct Sm;
template <>
struct Sm<false> {
    typedef double t;
    static __device__ __forceinline__ t zero() { return 0.0; }
    static __device__ __forceinline__ t conj(t a) { return a; }
    static __device__ __forceinline__ t scale(t a, double s) { return a * s; }
    // packed element of a real Gram: band rows a, b -> G[r0 + a][c0 + b]
    static __device__ __forceinline__ t pack(const double *G, int ldg, int r0, int c0, int a, int b) {
        return G[(int64_t)(c0 + b) * ldg + r0 + a];
    }
    static __device__ __forceinline__ double realpart(t a) { return a; }
};
template <>
struct Sm<true> {
    typedef Cplx t;
    static __device__ __forceinline__ t zero() { return {0.0, 0.0}; }
    static __device__ __forceinline__ t conj(t a) { return {a.re, -a.im}; }
    static __device__ __forceinline__ t scale(t a, double s) { return {a.re * s, a.im * s}; }
    // (Z^H W)_{ab} from the real Gram of SoA columns (re = 2a, im = 2a+1)
    static __device__ __forceinline__ t pack(const double *G, int ldg, int r0, int c0, int a, int b) {
        auto g = [&](int i, int j) { return G[(int64_t)(c0 + j) * ldg + r0 + i]; };
        return {g(2 * a, 2 * b) + g(2 * a + 1, 2 * b + 1), g(2 * a, 2 * b + 1) - g(2 * a + 1, 2 * b)};
    }
    static __device__ __forceinline__ double realpart(t a) { return a.re; }
};

// Blocks of the subspace matrices kept between iterations (N x N each, column-major):
// 0 XX, 1 XP, 2 PP, 3 XAX, 4 XAP, 5 PAP
enum { B_XX = 0, B_XP, B_PP, B_XAX, B_XAP, B_PAP, B_COUNT };

// Extract X / P blocks from G2 = [X P]^T [X P AX AP] (real, m2 = c*N*(1+hp) rows).
template <bool CPLX>
__global__ void extract_blocks_kernel(typename Sm<CPLX>::t *blk, const double *G2, int N, int c, bool hp) {
    using S = Sm<CPLX>;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= N * N) return;
    const int a = e % N, b = e / N;
    const int nc = c * N;
    const int m2 = hp ? 2 * nc : nc;
    const int64_t nn = (int64_t)N * N;
    // column blocks of G2: X at 0, P at nc (if hp), AX at nc*(1+hp), AP after it
    const int cX = 0, cP = nc, cAX = hp ? 2 * nc : nc, cAP = 3 * nc;
    blk[B_XX * nn + e] = S::pack(G2, m2, 0, cX, a, b);
    blk[B_XAX * nn + e] = S::pack(G2, m2, 0, cAX, a, b);
    if (hp) {
        blk[B_XP * nn + e] = S::pack(G2, m2, 0, cP, a, b);
        blk[B_PP * nn + e] = S::pack(G2, m2, nc, cP, a, b);
        blk[B_XAP * nn + e] = S::pack(G2, m2, 0, cAP, a, b);
        blk[B_PAP * nn + e] = S::pack(G2, m2, nc, cAP, a, b);
    }
}

// S and H (n3 = nbk*N, full Hermitian, column-major) from the kept blocks and
// G1 = [X R P]^T [R AR] (real, ld1 rows, columns 2*c*N).  R rows/columns are
// scaled by dr = 1/|r|.
// Diagonal given to dropped directions (decoupled R columns, SVQB null
// directions) so that they sort above the N lowest Ritz values: twice a
// Gershgorin bound of the X block's H (above its largest Ritz value, which
// bounds the N-th Ritz value of any subspace containing X) plus one.  A fixed
// huge value (1e30) would be exact in exact arithmetic, but a dense eigensolver's
// backward error (eps |M|, 1e14 for 1e30) then lands on every eigenvalue.
template <bool CPLX>
__global__ __launch_bounds__(64) void row_abs_sum_kernel(double *rs, const typename Sm<CPLX>::t *hxx, int N) {
    const int i = blockIdx.x;
    double r = 0.0;
    for (int j = threadIdx.x; j < N; j += 64) {
        const auto v = hxx[i + (int64_t)j * N];
        if constexpr (CPLX) r += sqrt(v.re * v.re + v.im * v.im);
        else r += fabs(v);
    }
    for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off, 64);
    if (threadIdx.x == 0) rs[i] = r;
}

// pen[0] = 2 max(rs[0 .. n-1]) + 1 (one block)
__global__ __launch_bounds__(256) void penalty_kernel(double *pen, const double *rs, int n) {
    __shared__ double red[256];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, rs[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) pen[0] = 2.0 * red[0] + 1.0;
}

template <bool CPLX>
__global__ void assemble_kernel(typename Sm<CPLX>::t *Sm_, typename Sm<CPLX>::t *Hm_,
                                const typename Sm<CPLX>::t *blk, const double *G1, const double *dr, int N, int c,
                                int nbk, int ld1, const double *rkeep, const double *pen) {
    using S = Sm<CPLX>;
    using t = typename S::t;
    const int n3 = nbk * N;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n3 * n3) return;
    const int I = (int)(e % n3), J = (int)(e / n3);
    const int bi = I / N, bj = J / N;
    const bool lower = bi > bj;
    // evaluate the upper block (u, v) at (a, b); lower blocks are its conjugate transpose
    const int u = lower ? bj : bi, v = lower ? bi : bj;
    const int a = lower ? J % N : I % N, b = lower ? I % N : J % N;
    const int nc = c * N;
    const int64_t nn = (int64_t)N * N;
    t s, h;
    if (u == 0 && v == 0) {
        s = blk[B_XX * nn + a + (int64_t)b * N];
        h = blk[B_XAX * nn + a + (int64_t)b * N];
    } else if (u == 0 && v == 2) {
        s = blk[B_XP * nn + a + (int64_t)b * N];
        h = blk[B_XAP * nn + a + (int64_t)b * N];
    } else if (u == 2 && v == 2) {
        s = blk[B_PP * nn + a + (int64_t)b * N];
        h = blk[B_PAP * nn + a + (int64_t)b * N];
    } else if (v == 1) {  // X R or R R: G1 rows of block u, columns R / AR
        s = S::scale(S::pack(G1, ld1, u * nc, 0, a, b), dr[b] * (u == 1 ? dr[a] : 1.0));
        h = S::scale(S::pack(G1, ld1, u * nc, nc, a, b), dr[b] * (u == 1 ? dr[a] : 1.0));
    } else {  // (1, 2): R^H P = (P^H R)^H from G1 rows P
        s = S::conj(S::scale(S::pack(G1, ld1, 2 * nc, 0, b, a), dr[a]));
        h = S::conj(S::scale(S::pack(G1, ld1, 2 * nc, nc, b, a), dr[a]));
    }
    if (lower) {
        s = S::conj(s);
        h = S::conj(h);
    }
    if (rkeep && I == J && bi == 1 && rkeep[I - N] == 0.0) {  // R column dropped by the orthonormalisation
        if constexpr (CPLX) {
            s = {1.0, 0.0};
            h = {*pen, 0.0};
        } else {
            s = 1.0;
            h = *pen;
        }
    }
    Sm_[e] = s;
    Hm_[e] = h;
}

// Cb = [C | C_P] (n3 x 2N): C = first N eigenvectors, C_P = C with the X rows zeroed.
template <bool CPLX>
__global__ void build_cb_kernel(typename Sm<CPLX>::t *Cb, const typename Sm<CPLX>::t *evec, int N, int n3) {
    using S = Sm<CPLX>;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n3 * 2 * N) return;
    const int row = (int)(e % n3), col = (int)(e / n3);
    const int src = col % N;
    auto v = evec[(int64_t)src * n3 + row];
    if (col >= N && row < N) v = S::zero();
    Cb[e] = v;
}

// From SS = Cb^H S Cb and HH = Cb^H H Cb (2N x 2N): P' norms, the kept blocks of
// the next iteration, and the real update matrix U (T) for the SoA columns:
//   [X' P'] = [X R P] U, U = [C | C_P diag(dp)] with R rows scaled by dr.
template <bool CPLX, typename T>
__global__ void next_blocks_kernel(typename Sm<CPLX>::t *blk, double *dp, const typename Sm<CPLX>::t *SS,
                                   const typename Sm<CPLX>::t *HH, int N) {
    using S = Sm<CPLX>;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= N * N) return;
    const int a = e % N, b = e / N;
    const int n2 = 2 * N;
    auto ss = [&](int i, int j) { return SS[(int64_t)j * n2 + i]; };
    auto hh = [&](int i, int j) { return HH[(int64_t)j * n2 + i]; };
    auto inv = [&](int i) {
        const double d = S::realpart(ss(N + i, N + i));
        return d > 0.0 ? 1.0 / sqrt(d) : 1.0;
    };
    const double da = inv(a), db = inv(b);
    const int64_t nn = (int64_t)N * N;
    blk[B_XX * nn + e] = ss(a, b);
    blk[B_XAX * nn + e] = hh(a, b);
    blk[B_XP * nn + e] = S::scale(ss(a, N + b), db);
    blk[B_XAP * nn + e] = S::scale(hh(a, N + b), db);
    blk[B_PP * nn + e] = S::scale(ss(N + a, N + b), da * db);
    blk[B_PAP * nn + e] = S::scale(hh(N + a, N + b), da * db);
    if (a == 0) dp[b] = db;
}

template <bool CPLX, typename T>
__global__ void build_update_kernel(T *U, const typename Sm<CPLX>::t *Cb, const double *dr, const double *dp, int N,
                                    int n3) {
    using S = Sm<CPLX>;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n3 * 2 * N) return;
    const int k = (int)(e % n3), n = (int)(e / n3);
    auto u = Cb[e];
    double sc = (k >= N && k < 2 * N) ? dr[k - N] : 1.0;
    if (n >= N) sc *= dp[n - N];
    u = S::scale(u, sc);
    if constexpr (!CPLX) {
        U[(int64_t)n * n3 + k] = (T)u;
    } else {
        const int ld = 2 * n3;
        U[(int64_t)(2 * n) * ld + 2 * k] = (T)u.re;
        U[(int64_t)(2 * n) * ld + 2 * k + 1] = (T)(-u.im);
        U[(int64_t)(2 * n + 1) * ld + 2 * k] = (T)u.im;
        U[(int64_t)(2 * n + 1) * ld + 2 * k + 1] = (T)u.re;
    }
}

__global__ void inv_sqrt_kernel(double *out, const double *in, int n, double scale) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const double v = in[i] * scale;
    out[i] = v > 0.0 ? 1.0 / sqrt(v) : 1.0;
}

// R orthonormalisation against X.  From G0 = [X R]^T R (real, 2nc x nc):
//   XRn = X^H R diag(dr),  M = diag(dr) R^H R diag(dr)   (both N x N; M gets -XRn^H XRn later)
template <bool CPLX>
__global__ void rorth_pack_kernel(typename Sm<CPLX>::t *XRn, typename Sm<CPLX>::t *M, const double *G0,
                                  const double *dr, int N, int c) {
    using S = Sm<CPLX>;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= N * N) return;
    const int a = e % N, b = e / N;
    const int nc = c * N;
    XRn[e] = S::scale(S::pack(G0, 2 * nc, 0, 0, a, b), dr[b]);
    M[e] = S::scale(S::pack(G0, 2 * nc, nc, 0, a, b), dr[a] * dr[b]);
}

// W = [-XRn; I] (2N x N) before the right triangular solve W U^-1
template <bool CPLX>
__global__ void rorth_w_kernel(typename Sm<CPLX>::t *W, const typename Sm<CPLX>::t *XRn, int N) {
    using S = Sm<CPLX>;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)2 * N * N) return;
    const int row = (int)(e % (2 * N)), col = (int)(e / (2 * N));
    typename S::t v;
    if (row < N) v = S::scale(XRn[(int64_t)col * N + row], -1.0);
    else v = (row - N == col) ? S::scale(S::zero(), 0.0) : S::zero();
    if (row >= N && row - N == col) {
        if constexpr (CPLX) v = {1.0, 0.0};
        else v = 1.0;
    }
    W[e] = v;
}

// SVQB variant: V columns scaled by s^-1/2 (0 when s <= tau s_max), keep mask per column
template <bool CPLX>
__global__ void rorth_svqb_kernel(typename Sm<CPLX>::t *V, double *keep, const double *sv, int N, double tau) {
    using S = Sm<CPLX>;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= N * N) return;
    const int j = e / N;
    const double smax = sv[N - 1];
    const double wj = sv[j] > tau * smax ? 1.0 / sqrt(sv[j]) : 0.0;
    V[e] = S::scale(V[e], wj);
    if (e % N == 0) keep[j] = wj > 0.0 ? 1.0 : 0.0;
}

// rows N..2N-1 of W (2N x N) scaled by dr (R rows); the X rows pass through
template <bool CPLX>
__global__ void rorth_scale_kernel(typename Sm<CPLX>::t *W, const double *dr, int N) {
    using S = Sm<CPLX>;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)2 * N * N) return;
    const int row = (int)(e % (2 * N));
    if (row >= N) W[e] = S::scale(W[e], dr[row - N]);
}

// real update matrix (T) from a complex / real coefficient block W (m x n)
template <bool CPLX, typename T>
__global__ void real_coef_kernel(T *U, const typename Sm<CPLX>::t *W, int m, int n) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)m * n) return;
    const int k = (int)(e % m), j = (int)(e / m);
    const auto u = W[e];
    if constexpr (!CPLX) {
        U[(int64_t)j * m + k] = (T)u;
    } else {
        const int ld = 2 * m;
        U[(int64_t)(2 * j) * ld + 2 * k] = (T)u.re;
        U[(int64_t)(2 * j) * ld + 2 * k + 1] = (T)(-u.im);
        U[(int64_t)(2 * j + 1) * ld + 2 * k] = (T)u.im;
        U[(int64_t)(2 * j + 1) * ld + 2 * k + 1] = (T)u.re;
    }
}

// U = [-C; I] ((c f + c na) x c na, column-major, T) with C the real form of
// X_f^H V from the real Gram G = X_f^T V of the SoA columns (c f x c na), so
// that [X_f V] U = V - X_f (X_f^H V) (VectorOrthogonalization,
// lobpcg_z_multi_mpi.h:1637-1723; the complex unpacking is that of 1653-1670).
template <bool CPLX, typename T>
__global__ void fixed_coef_kernel(T *U, const double *G, int f, int na, int c) {
    const int m = c * (f + na), n = c * na, mf = c * f;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)m * n) return;
    const int row = (int)(e % m), col = (int)(e / m);
    double v;
    if (row >= mf) {
        v = (row - mf == col) ? 1.0 : 0.0;
    } else if constexpr (!CPLX) {
        v = -G[(int64_t)col * mf + row];
    } else {
        // (X_f^H V)_{jk} = re + i im; rows 2j / 2j+1 (re / im of x_j), columns 2k / 2k+1
        const int j = row >> 1, k = col >> 1;
        auto g = [&](int a, int b) { return G[(int64_t)b * mf + a]; };
        const double re = g(2 * j, 2 * k) + g(2 * j + 1, 2 * k + 1);
        const double im = g(2 * j, 2 * k + 1) - g(2 * j + 1, 2 * k);
        const bool xre = (row & 1) == 0, vre = (col & 1) == 0;
        v = xre ? (vre ? -re : -im) : (vre ? im : -re);
    }
    U[e] = (T)v;
}

__global__ void cast_kernel(double *__restrict__ y, const float *__restrict__ x, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) y[i] = x[i];
}

__global__ void ratio_kernel(double *out, const double *num, const double *den, int n) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < n) out[i] = den[i] > 0.0 ? num[i] / den[i] : num[i];
}

__global__ void fill_kernel(double *p, double v, int n) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < n) p[i] = v;
}

// min / max of the squared Cholesky pivots |U_ii|^2 (a cheap conditioning probe of S)
template <bool CPLX>
__global__ void pivot_ratio_kernel(double *out, const typename Sm<CPLX>::t *U, int n3) {
    __shared__ double mn[256], mx[256];
    double lo = 1.0e300, hi = 0.0;
    for (int i = threadIdx.x; i < n3; i += 256) {
        const double d = Sm<CPLX>::realpart(U[(int64_t)i * n3 + i]);
        lo = fmin(lo, d * d);
        hi = fmax(hi, d * d);
    }
    mn[threadIdx.x] = lo;
    mx[threadIdx.x] = hi;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + st]);
            mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = mx[0] > 0.0 ? mn[0] / mx[0] : 0.0;
}

// dev[i] = |1 - c_i^H t_i| for the N lowest Ritz vectors c_i (block i) and
// t_i = S c_i (columns of C and T, leading dimension n3); non-finite -> 1e300
template <bool CPLX>
__global__ __launch_bounds__(256) void ritz_orth_kernel(double *dev, const typename Sm<CPLX>::t *C,
                                                        const typename Sm<CPLX>::t *T, int n3) {
    __shared__ double red[256];
    const auto *c = C + (int64_t)blockIdx.x * n3;
    const auto *t = T + (int64_t)blockIdx.x * n3;
    double acc = 0.0;
    for (int j = threadIdx.x; j < n3; j += 256) {
        if constexpr (CPLX) acc += c[j].re * t[j].re + c[j].im * t[j].im;  // Re(conj(c_j) t_j)
        else acc += c[j] * t[j];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double d = fabs(1.0 - red[0]);
        dev[blockIdx.x] = d == d && d < 1.0e300 ? d : 1.0e300;
    }
}

// dev[0] = max(dev[0 .. n-1]) (one block)
__global__ __launch_bounds__(256) void max_reduce_kernel(double *dev, int n) {
    __shared__ double red[256];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, dev[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) dev[0] = red[0];
}

// SVQB fallback: columns of the S eigenvectors scaled by s^-1/2, near-null
// directions (s <= tau * s_max) zeroed; w[j] records the scale (0 = dropped).
template <bool CPLX>
__global__ void svqb_basis_kernel(typename Sm<CPLX>::t *V, double *w, const double *sv, int n3, double tau) {
    using S = Sm<CPLX>;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n3 * n3) return;
    const int j = (int)(e / n3);
    const double smax = sv[n3 - 1];
    const double wj = sv[j] > tau * smax ? 1.0 / sqrt(sv[j]) : 0.0;
    V[e] = S::scale(V[e], wj);
    if (e % n3 == 0) w[j] = wj;
}

// dropped directions of the reduced H (zero rows and columns) get the
// penalty diagonal so they sort above the N lowest
template <bool CPLX>
__global__ void svqb_penalty_kernel(typename Sm<CPLX>::t *Hr, const double *w, int n3, const double *pen) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n3) return;
    if (w[i] == 0.0) {
        if constexpr (CPLX) Hr[(int64_t)i * n3 + i].re += *pen;
        else Hr[(int64_t)i * n3 + i] += *pen;
    }
}

}  // namespace
}  // namespace qm

struct qm_lobpcg {
    qm_ctx *ctx = nullptr;
    int N = 0, c = 1, prec = 64;
    int na = 0;  // bands iterated by the current solve (N minus the fixed bands)
    int64_t ng = 0;
    size_t tsz = 8;
    // vector blocks (Ng x c*N each, T)
    void *X = nullptr, *R = nullptr, *P = nullptr, *AX = nullptr, *AR = nullptr, *AP = nullptr;
    void *W[4] = {nullptr, nullptr, nullptr, nullptr};
    bool have_p = false;  // P kept from an earlier call
    // small FP64 buffers
    double *G1 = nullptr, *G2 = nullptr;
    void *S = nullptr, *H = nullptr, *Sw = nullptr, *Hw = nullptr, *blk = nullptr, *Cb = nullptr, *T1 = nullptr,
         *SS = nullptr, *HH = nullptr;
    void *U = nullptr;  // update matrix (T)
    void *Ev = nullptr; // eigenvectors of the SVQB fallback (n3 x n3)
    int fallbacks = 0;  // subspace-solve fallbacks (P dropped or SVQB) in the last solve
    int rdrops = 0;     // R orthonormalisations that dropped near-dependent directions
    int stedc_retries = 0;  // tridiagonal eigensolves redone by steqr after stedc failed
    std::string eig_note;   // what the last failed eigensolve reported (for error messages)
    double cond64 = 1.0e-14, cond32 = 1.0e-4;  // minimum squared Cholesky pivot ratio of S
    double limit_normx = qm::kLimitNormX;       // rollback when max |1 - |x'|^2| exceeds this
    double rdrop64 = 1.0e-13, rdrop32 = 1.0e-6; // R directions kept above this eigenvalue ratio
    void *XRn = nullptr, *Mr = nullptr, *Wr = nullptr, *Ur = nullptr;  // R orthonormalisation
    double *X64 = nullptr, *AX64 = nullptr;  // FP64 Rayleigh-quotient refinement of the mixed solve
    double *G0 = nullptr, *rkeep = nullptr;
    double *evals = nullptr, *ework = nullptr, *eps = nullptr, *dr = nullptr, *dp = nullptr, *rnorm = nullptr,
           *xnorm = nullptr, *svals = nullptr, *svw = nullptr;
    int *info = nullptr;
    double *trd = nullptr, *VW = nullptr;  // tridiagonal reduction workspace (real subspace problems)
    bool evec_t1 = false;                  // last chol_solve left the eigenvectors in T1 (else Hw)
    int ortho_rejects = 0;                 // Cholesky-route solves rejected for S-orthonormality (last call)
    qm::DevBuf gram_part, red_part;
    rocblas_handle blas = nullptr;
    double t[6] = {0, 0, 0, 0, 0, 0};
    // phase timing without stream synchronisation: events recorded at phase
    // ends, read once when the solve has synchronised (reused between solves)
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_phase;
    size_t ev_used = 0;
    double *evk = nullptr;   // Ritz values of the last accepted iteration (device, N)
    double *pen = nullptr;   // diagonal of dropped directions in the subspace problem (device, 1)
    void *rt = nullptr;      // S times the N lowest Ritz vectors (device, 3N x N subspace elements)
    void *psync = nullptr;   // barrier words of the cooperative Cholesky (device, 512 bytes)
    bool graphs = true;      // replay A.R + Gram of an iteration as a HIP graph
    hipStream_t gstream = nullptr;  // capture stream (graphs launch on the caller's stream)
    int graph_replays = 0;   // replays in the last solve
};

namespace qm {
namespace {

void lfree(qm_lobpcg *s) {
    void **ptrs[] = {&s->X, &s->R, &s->P, &s->AX, &s->AR, &s->AP, &s->W[0], &s->W[1], &s->W[2], &s->W[3],
                     &s->S, &s->H, &s->Sw, &s->Hw, &s->blk, &s->Cb, &s->T1, &s->SS, &s->HH, &s->U, &s->Ev,
                     &s->XRn, &s->Mr, &s->Wr, &s->Ur, (void **)&s->X64, (void **)&s->AX64};
    for (void **p : ptrs)
        if (*p) {
            (void)hipFree(*p);
            *p = nullptr;
        }
    double **dptrs[] = {&s->G1, &s->G2, &s->evals, &s->ework, &s->eps, &s->dr, &s->dp, &s->rnorm, &s->xnorm,
                        &s->svals, &s->svw, &s->G0, &s->rkeep, &s->trd, &s->VW};
    for (double **p : dptrs)
        if (*p) {
            (void)hipFree(*p);
            *p = nullptr;
        }
    if (s->info) (void)hipFree(s->info);
    s->info = nullptr;
    if (s->evk) (void)hipFree(s->evk);
    s->evk = nullptr;
    if (s->pen) (void)hipFree(s->pen);
    s->pen = nullptr;
    if (s->rt) (void)hipFree(s->rt);
    s->rt = nullptr;
    if (s->psync) (void)hipFree(s->psync);
    s->psync = nullptr;
    if (s->gstream) (void)hipStreamDestroy(s->gstream);
    s->gstream = nullptr;
    for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
    s->ev.clear();
    s->ev_phase.clear();
    s->gram_part.release();
    s->red_part.release();
    if (s->blas) rocblas_destroy_handle(s->blas);
    s->blas = nullptr;
}

int dalloc(void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) {
        *p = nullptr;
        set_error("qm_lobpcg_create: hipMalloc of " + std::to_string(bytes) + " bytes failed");
        return QM_ERR_NOMEM;
    }
    return QM_OK;
}

#define QM_BLAS(call)                                                                                       \
    do {                                                                                                    \
        rocblas_status s_ = (call);                                                                         \
        if (s_ != rocblas_status_success) {                                                                 \
            ::qm::set_error(std::string(#call) + " failed with rocblas status " + std::to_string((int)s_)); \
            return QM_ERR_HIP;                                                                              \
        }                                                                                                   \
    } while (0)

template <typename T>
int hpsi(qm_lobpcg *s, T *out, const T *in, int nb, int ik, hipStream_t st) {
    if constexpr (sizeof(T) == 8) return qm_hpsi_apply_f64(s->ctx, out, in, nb, s->c, ik, st);
    else return qm_hpsi_apply_f32(s->ctx, out, in, nb, s->c, ik, st);
}

// H c = l S c as in lapack_ZHEGVD_emu (w_zhegvd.h:227-272): S = U^H U (potrf),
// C = U^-H H U^-1 (two trsm), eigh(C) (syevd), c = U^-1 y (trsm, first N columns).
// Real n <= own_eig_max_n() instead factors S in one launch of the own
// Cholesky, forms U^-1 once (trtri) and applies it by GEMMs, and diagonalises
// C by the own tridiagonal reduction + stedc / ormtr.
// Works on copies Sw, Hw; eigenvalues in s->evals, eigenvectors in Hw.  After
// potrf the squared pivot ratio is written to s->svals[0] for the caller's
// conditioning check; on potrf failure s->info > 0 and the rest is skipped by
// the caller.
template <bool CPLX>
int chol_factor(qm_lobpcg *s, int n3, hipStream_t st) {
    using SmT = typename Sm<CPLX>::t;
    QM_HIP(hipMemcpyAsync(s->Sw, s->S, sizeof(SmT) * n3 * n3, hipMemcpyDeviceToDevice, st));
    QM_HIP(hipMemcpyAsync(s->Hw, s->H, sizeof(SmT) * n3 * n3, hipMemcpyDeviceToDevice, st));
    if constexpr (CPLX)
        QM_BLAS(rocsolver_zpotrf(s->blas, rocblas_fill_upper, n3, (rocblas_double_complex *)s->Sw, n3, s->info));
    else
        QM_TRY(potrf_upper(s->blas, n3, (double *)s->Sw, n3, s->info, st, s->psync));
    hipLaunchKernelGGL((pivot_ratio_kernel<CPLX>), dim3(1), dim3(256), 0, st, s->svals, (const SmT *)s->Sw, n3);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

template <bool CPLX>
int chol_solve(qm_lobpcg *s, int n3, hipStream_t st) {
    const int N = s->na;
    if constexpr (CPLX) {
        const rocblas_double_complex one{1.0, 0.0};
        auto *U = (rocblas_double_complex *)s->Sw, *C = (rocblas_double_complex *)s->Hw;
        QM_BLAS(rocblas_ztrsm(s->blas, rocblas_side_left, rocblas_fill_upper, rocblas_operation_conjugate_transpose,
                              rocblas_diagonal_non_unit, n3, n3, &one, U, n3, C, n3));
        QM_BLAS(rocblas_ztrsm(s->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                              rocblas_diagonal_non_unit, n3, n3, &one, U, n3, C, n3));
        if (subspace_route() != 0 && trd_coop_supported(n3, true) &&
            trd_coop_work_bytes(n3, true) <= sizeof(double) * (size_t)kTrdPanelCols * n3) {
            // own cooperative reduction C = Q T Q^H, stedc on T, Z = Q Y for the N
            // lowest, c = U^-1 Z; eigenvectors in T1
            double *d = s->trd, *e = s->trd + n3;
            auto *tau = (rocblas_double_complex *)(s->trd + 2 * n3);
            auto *Z = (rocblas_double_complex *)s->T1;
            const int r = trd_coop_z64(n3, C, n3, d, e, tau, s->VW, sizeof(double) * (size_t)kTrdPanelCols * n3, st);
            if (r == QM_OK) {
                double *e_copy = s->trd + 4 * n3;  // stedc overwrites e
                QM_HIP(hipMemcpyAsync(e_copy, e, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
                QM_HIP(hipMemcpyAsync(s->evals, d, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
                std::vector<double> hde(2 * (size_t)n3);
                QM_HIP(hipMemcpyAsync(hde.data(), d, sizeof(double) * 2 * n3, hipMemcpyDeviceToHost, st));
                QM_BLAS(rocsolver_zstedc(s->blas, rocblas_evect_tridiagonal, n3, s->evals, e, Z, n3, s->info));
                int h_info = 0;
                QM_HIP(hipMemcpyAsync(&h_info, s->info, sizeof(int), hipMemcpyDeviceToHost, st));
                QM_HIP(hipStreamSynchronize(st));
                bool finite = true;
                for (size_t i = 0; i + 1 < hde.size(); ++i) finite = finite && std::isfinite(hde[i]);
                if (!finite) {
                    // the tridiagonal is not finite: say whether its input was, then
                    // solve by heevd from C's upper triangle (left intact)
                    std::vector<Cplx> hc((size_t)n3 * n3);
                    QM_HIP(hipMemcpyAsync(hc.data(), C, sizeof(Cplx) * hc.size(), hipMemcpyDeviceToHost, st));
                    QM_HIP(hipStreamSynchronize(st));
                    bool cfin = true;
                    for (int jj = 0; jj < n3; ++jj)
                        for (int ii = 0; ii <= jj; ++ii) {
                            const Cplx &x = hc[ii + (size_t)jj * n3];
                            cfin = cfin && std::isfinite(x.re) && std::isfinite(x.im);
                        }
                    ++s->stedc_retries;
                    s->eig_note = std::string("the tridiagonal reduction of order ") + std::to_string(n3) +
                                  " gave non-finite values from a " + (cfin ? "finite" : "non-finite") + " matrix";
                    QM_BLAS(rocsolver_zheevd(s->blas, rocblas_evect_original, rocblas_fill_upper, n3, C, n3,
                                             s->evals, s->ework, s->info));
                    QM_BLAS(rocblas_ztrsm(s->blas, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                          rocblas_diagonal_non_unit, n3, N, &one, U, n3, C, n3));
                    s->evec_t1 = false;
                    return QM_OK;
                }
                if (h_info != 0) {
                    // divide and conquer did not converge (it can on nearly degenerate
                    // spectra): implicit QL/QR on the same tridiagonal instead
                    ++s->stedc_retries;
                    s->eig_note = "stedc of order " + std::to_string(n3) + " info " + std::to_string(h_info);
                    QM_HIP(hipMemcpyAsync(s->evals, d, sizeof(double) * n3, hipMemcpyDeviceToDevice, st));
                    QM_BLAS(rocsolver_zsteqr(s->blas, rocblas_evect_tridiagonal, n3, s->evals, e_copy, Z, n3,
                                             s->info));
                }
                QM_TRY(trd_coop_check(s->VW, s->info, st));
                QM_TRY(unmtr_lower_z64(n3, N, C, n3, tau, Z, n3, s->VW, st));
                QM_BLAS(rocblas_ztrsm(s->blas, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                      rocblas_diagonal_non_unit, n3, N, &one, U, n3, Z, n3));
                s->evec_t1 = true;
                return QM_OK;
            }
            if (r != QM_ERR_STATE) return r;
        }
        QM_BLAS(rocsolver_zheevd(s->blas, rocblas_evect_original, rocblas_fill_upper, n3, C, n3, s->evals, s->ework,
                                 s->info));
        QM_BLAS(rocblas_ztrsm(s->blas, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                              rocblas_diagonal_non_unit, n3, N, &one, U, n3, C, n3));
    } else {
        const double one = 1.0;
        auto *U = (double *)s->Sw, *C = (double *)s->Hw;
        if (n3 <= own_eig_max_n()) {
            // C = U^-T H U^-1 through U^-1 and GEMMs, own tridiagonal reduction +
            // rocSOLVER stedc / ormtr (eigenvectors in T1), then c = U^-1 y into Hw
            auto *Ui = (double *)s->Ev, *Z = (double *)s->T1;
            QM_TRY(chol_transform(s->blas, n3, U, C, Ui, Z, st));
            QM_TRY(syev_lower(s->blas, n3, N, C, n3, s->evals, Z, n3, s->trd, s->VW, s->info, st));
            QM_TRY(chol_back(s->blas, n3, N, Ui, Z, n3, C, n3));
            s->evec_t1 = false;
            return QM_OK;
        }
        QM_BLAS(rocblas_dtrsm(s->blas, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                              rocblas_diagonal_non_unit, n3, n3, &one, U, n3, C, n3));
        QM_BLAS(rocblas_dtrsm(s->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                              rocblas_diagonal_non_unit, n3, n3, &one, U, n3, C, n3));
        QM_BLAS(rocsolver_dsyevd(s->blas, rocblas_evect_original, rocblas_fill_upper, n3, C, n3, s->evals, s->ework,
                                 s->info));
        QM_BLAS(rocblas_dtrsm(s->blas, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                              rocblas_diagonal_non_unit, n3, N, &one, U, n3, C, n3));
    }
    s->evec_t1 = false;
    return QM_OK;
}

// All eigenpairs of the Hermitian n x n matrix A (upper triangle, leading
// dimension n): eigenvalues ascending into w, eigenvectors over A, by rocSOLVER
// heevd / syevd.  Where that reports non-convergence (its divide and conquer
// can fail on nearly degenerate spectra), the same problem again from a copy by
// rocSOLVER's Jacobi solver (heevj / syevj), a different algorithm.  Uses Ev
// (the copy) and svw (Jacobi residual, sweeps); synchronises the stream;
// s->info holds the final status.
template <bool CPLX>
int heev_all(qm_lobpcg *s, int n, void *A, double *w, hipStream_t st) {
    using SmT = typename Sm<CPLX>::t;
    QM_HIP(hipMemcpyAsync(s->Ev, A, sizeof(SmT) * n * n, hipMemcpyDeviceToDevice, st));
    if constexpr (CPLX)
        QM_BLAS(rocsolver_zheevd(s->blas, rocblas_evect_original, rocblas_fill_upper, n,
                                 (rocblas_double_complex *)A, n, w, s->ework, s->info));
    else
        QM_BLAS(rocsolver_dsyevd(s->blas, rocblas_evect_original, rocblas_fill_upper, n, (double *)A, n, w, s->ework,
                                 s->info));
    int h_info = 0;
    QM_HIP(hipMemcpyAsync(&h_info, s->info, sizeof(int), hipMemcpyDeviceToHost, st));
    QM_HIP(hipStreamSynchronize(st));
    if (h_info == 0) return QM_OK;
    ++s->stedc_retries;
    QM_HIP(hipMemcpyAsync(A, s->Ev, sizeof(SmT) * n * n, hipMemcpyDeviceToDevice, st));
    std::vector<SmT> h((size_t)n * n);
    QM_HIP(hipMemcpyAsync(h.data(), A, sizeof(SmT) * h.size(), hipMemcpyDeviceToHost, st));
    QM_HIP(hipStreamSynchronize(st));
    bool finite = true;
    for (const SmT &x : h) {
        if constexpr (CPLX) finite = finite && std::isfinite(x.re) && std::isfinite(x.im);
        else finite = finite && std::isfinite(x);
    }
    s->eig_note = "heevd/syevd of order " + std::to_string(n) + " info " + std::to_string(h_info) +
                  (finite ? " on a finite matrix" : " on a matrix with NaN/Inf");
    QM_CHECK(finite, QM_ERR_NUMERIC, "qm_lobpcg_solve: subspace eigensolve failed (" + s->eig_note + ")");
    int *sweeps = reinterpret_cast<int *>(s->svw + 1);
    if constexpr (CPLX)
        QM_BLAS(rocsolver_zheevj(s->blas, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_upper, n,
                                 (rocblas_double_complex *)A, n, 0.0, s->svw, 100, sweeps, w, s->info));
    else
        QM_BLAS(rocsolver_dsyevj(s->blas, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_upper, n,
                                 (double *)A, n, 0.0, s->svw, 100, sweeps, w, s->info));
    QM_HIP(hipMemcpyAsync(&h_info, s->info, sizeof(int), hipMemcpyDeviceToHost, st));
    QM_HIP(hipStreamSynchronize(st));
    s->eig_note += ", then Jacobi info " + std::to_string(h_info);
    return QM_OK;
}

// Fallback when S is not numerically positive definite: B = V s^-1/2 over the
// kept S eigendirections, eigh(B^H H B), eigenvectors C = B y into s->Ev.
template <bool CPLX>
int svqb_eig(qm_lobpcg *s, int n3, double tau, hipStream_t st) {
    using SmT = typename Sm<CPLX>::t;
    QM_HIP(hipMemcpyAsync(s->Sw, s->S, sizeof(SmT) * n3 * n3, hipMemcpyDeviceToDevice, st));
    QM_TRY(heev_all<CPLX>(s, n3, s->Sw, s->svals, st));
    hipLaunchKernelGGL((svqb_basis_kernel<CPLX>), dim3((unsigned)(((int64_t)n3 * n3 + 255) / 256)), dim3(256), 0, st,
                       (SmT *)s->Sw, s->svw, s->svals, n3, tau);
    QM_HIP(hipGetLastError());
    if constexpr (CPLX) {
        const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
        auto *B = (rocblas_double_complex *)s->Sw, *T1 = (rocblas_double_complex *)s->T1,
             *Hr = (rocblas_double_complex *)s->Hw;
        QM_BLAS(rocblas_zgemm(s->blas, rocblas_operation_none, rocblas_operation_none, n3, n3, n3, &one,
                              (rocblas_double_complex *)s->H, n3, B, n3, &zero, T1, n3));
        QM_BLAS(rocblas_zgemm(s->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, n3, n3, n3,
                              &one, B, n3, T1, n3, &zero, Hr, n3));
        hipLaunchKernelGGL((svqb_penalty_kernel<CPLX>), dim3((n3 + 63) / 64), dim3(64), 0, st, (SmT *)s->Hw, s->svw,
                           n3, s->pen);
        QM_HIP(hipGetLastError());
        QM_TRY(heev_all<CPLX>(s, n3, Hr, s->evals, st));
        QM_BLAS(rocblas_zgemm(s->blas, rocblas_operation_none, rocblas_operation_none, n3, n3, n3, &one, B, n3, Hr,
                              n3, &zero, (rocblas_double_complex *)s->Ev, n3));
    } else {
        const double one = 1.0, zero = 0.0;
        auto *B = (double *)s->Sw, *T1 = (double *)s->T1, *Hr = (double *)s->Hw;
        QM_BLAS(rocblas_dgemm(s->blas, rocblas_operation_none, rocblas_operation_none, n3, n3, n3, &one,
                              (double *)s->H, n3, B, n3, &zero, T1, n3));
        QM_BLAS(rocblas_dgemm(s->blas, rocblas_operation_transpose, rocblas_operation_none, n3, n3, n3, &one, B, n3,
                              T1, n3, &zero, Hr, n3));
        hipLaunchKernelGGL((svqb_penalty_kernel<CPLX>), dim3((n3 + 63) / 64), dim3(64), 0, st, (SmT *)s->Hw, s->svw,
                           n3, s->pen);
        QM_HIP(hipGetLastError());
        QM_TRY(heev_all<CPLX>(s, n3, Hr, s->evals, st));
        QM_BLAS(rocblas_dgemm(s->blas, rocblas_operation_none, rocblas_operation_none, n3, n3, n3, &one, B, n3, Hr,
                              n3, &zero, (double *)s->Ev, n3));
    }
    return QM_OK;
}

// R' = [X R] W: R orthogonalised against X (X^H X = I) and orthonormalised by a
// Cholesky QR of M = Rn^H Rn - (X^H Rn)^H (X^H Rn), Rn = R diag(dr); when M is not
// well conditioned, by its eigenbasis with the near-null directions dropped
// (their columns of R' are zero and s->rkeep marks them).  R' spans the same
// space as R minus its X component, so the Rayleigh-Ritz step is unchanged in
// exact arithmetic; H and S of the R block are then computed from a
// well-conditioned basis.  Writes R' into Rout.
template <typename T, bool CPLX>
int orthonormalise_r(qm_lobpcg *s, T *Rout, const T *X, const T *R, hipStream_t st) {
    using SmT = typename Sm<CPLX>::t;
    const int N = s->na, c = s->c, nc = c * N;
    const int64_t ng = s->ng;
    ColBlocks A0, B0;
    A0.add(X, nc);
    A0.add(R, nc);
    B0.add(R, nc);
    QM_TRY(gram<T>(s->gram_part, s->G0, A0, B0, 2 * nc, nc, ng, st));
    hipLaunchKernelGGL((rorth_pack_kernel<CPLX>), dim3((unsigned)(((int64_t)N * N + 255) / 256)), dim3(256), 0, st,
                       (SmT *)s->XRn, (SmT *)s->Mr, s->G0, s->dr, N, c);
    QM_HIP(hipGetLastError());
    // M -= XRn^H XRn
    if constexpr (CPLX) {
        const rocblas_double_complex mone{-1.0, 0.0}, one{1.0, 0.0};
        QM_BLAS(rocblas_zgemm(s->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, N, N, N, &mone,
                              (rocblas_double_complex *)s->XRn, N, (rocblas_double_complex *)s->XRn, N, &one,
                              (rocblas_double_complex *)s->Mr, N));
    } else {
        const double mone = -1.0, one = 1.0;
        QM_BLAS(rocblas_dgemm(s->blas, rocblas_operation_transpose, rocblas_operation_none, N, N, N, &mone,
                              (double *)s->XRn, N, (double *)s->XRn, N, &one, (double *)s->Mr, N));
    }
    // Cholesky attempt on a copy (Sw is free at this point)
    QM_HIP(hipMemcpyAsync(s->Sw, s->Mr, sizeof(SmT) * N * N, hipMemcpyDeviceToDevice, st));
    if constexpr (CPLX)
        QM_BLAS(rocsolver_zpotrf(s->blas, rocblas_fill_upper, N, (rocblas_double_complex *)s->Sw, N, s->info));
    else
        QM_TRY(potrf_upper(s->blas, N, (double *)s->Sw, N, s->info, st, s->psync));
    hipLaunchKernelGGL((pivot_ratio_kernel<CPLX>), dim3(1), dim3(256), 0, st, s->svals, (const SmT *)s->Sw, N);
    QM_HIP(hipGetLastError());
    int h_info = 0;
    double h_ratio = 0.0;
    QM_HIP(hipMemcpyAsync(&h_info, s->info, sizeof(int), hipMemcpyDeviceToHost, st));
    QM_HIP(hipMemcpyAsync(&h_ratio, s->svals, sizeof(double), hipMemcpyDeviceToHost, st));
    QM_HIP(hipStreamSynchronize(st));
    const double drop = sizeof(T) == 8 ? s->rdrop64 : s->rdrop32;
    const unsigned gW = (unsigned)(((int64_t)2 * N * N + 255) / 256);
    if (h_info == 0 && h_ratio >= drop) {
        hipLaunchKernelGGL((rorth_w_kernel<CPLX>), dim3(gW), dim3(256), 0, st, (SmT *)s->Wr, (const SmT *)s->XRn, N);
        QM_HIP(hipGetLastError());
        if constexpr (CPLX) {
            const rocblas_double_complex one{1.0, 0.0};
            QM_BLAS(rocblas_ztrsm(s->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                  rocblas_diagonal_non_unit, 2 * N, N, &one, (rocblas_double_complex *)s->Sw, N,
                                  (rocblas_double_complex *)s->Wr, 2 * N));
        } else {
            const double one = 1.0;
            QM_BLAS(rocblas_dtrsm(s->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                  rocblas_diagonal_non_unit, 2 * N, N, &one, (double *)s->Sw, N, (double *)s->Wr,
                                  2 * N));
        }
        hipLaunchKernelGGL(fill_kernel, dim3((N + 63) / 64), dim3(64), 0, st, s->rkeep, 1.0, N);
        QM_HIP(hipGetLastError());
    } else {
        // eigenbasis of M with near-null directions dropped
        QM_HIP(hipMemcpyAsync(s->Sw, s->Mr, sizeof(SmT) * N * N, hipMemcpyDeviceToDevice, st));
        QM_TRY(heev_all<CPLX>(s, N, s->Sw, s->svals, st));
        hipLaunchKernelGGL((rorth_svqb_kernel<CPLX>), dim3((unsigned)(((int64_t)N * N + 255) / 256)), dim3(256), 0,
                           st, (SmT *)s->Sw, s->rkeep, s->svals, N, drop);
        QM_HIP(hipGetLastError());
        // W = [-XRn Vs ; Vs]
        if constexpr (CPLX) {
            const rocblas_double_complex mone{-1.0, 0.0}, zero{0.0, 0.0};
            QM_BLAS(rocblas_zgemm(s->blas, rocblas_operation_none, rocblas_operation_none, N, N, N, &mone,
                                  (rocblas_double_complex *)s->XRn, N, (rocblas_double_complex *)s->Sw, N, &zero,
                                  (rocblas_double_complex *)s->Wr, 2 * N));
            QM_HIP(hipMemcpy2DAsync((rocblas_double_complex *)s->Wr + N, sizeof(SmT) * 2 * N, s->Sw, sizeof(SmT) * N,
                                    sizeof(SmT) * N, N, hipMemcpyDeviceToDevice, st));
        } else {
            const double mone = -1.0, zero = 0.0;
            QM_BLAS(rocblas_dgemm(s->blas, rocblas_operation_none, rocblas_operation_none, N, N, N, &mone,
                                  (double *)s->XRn, N, (double *)s->Sw, N, &zero, (double *)s->Wr, 2 * N));
            QM_HIP(hipMemcpy2DAsync((double *)s->Wr + N, sizeof(SmT) * 2 * N, s->Sw, sizeof(SmT) * N, sizeof(SmT) * N,
                                    N, hipMemcpyDeviceToDevice, st));
        }
        ++s->rdrops;
    }
    hipLaunchKernelGGL((rorth_scale_kernel<CPLX>), dim3(gW), dim3(256), 0, st, (SmT *)s->Wr, s->dr, N);
    QM_HIP(hipGetLastError());
    hipLaunchKernelGGL((real_coef_kernel<CPLX, T>), dim3(gW), dim3(256), 0, st, (T *)s->Ur, (const SmT *)s->Wr, 2 * N,
                       N);
    QM_HIP(hipGetLastError());
    ColBlocks Y;
    Y.add(Rout, nc);
    QM_TRY(update<T>(A0, Y, (const T *)s->Ur, 2 * nc, nc, ng, st));
    // R' columns are orthonormal: unit scale for the Rayleigh-Ritz assembly
    hipLaunchKernelGGL(fill_kernel, dim3((N + 63) / 64), dim3(64), 0, st, s->dr, 1.0, N);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

// Y = V - X_f (X_f^H V) for the na active bands V against the f fixed bands X_f
// (VectorOrthogonalization, lobpcg_z_multi_mpi.h:1637-1723): one Gram product
// (FP64 result) and one update.  Y must not alias X_f or V.
template <typename T, bool CPLX>
int orthogonalise_fixed(qm_lobpcg *s, T *Y, const T *Xf, int f, const T *V, hipStream_t st) {
    const int c = s->c, mf = c * f, nc = c * s->na;
    ColBlocks A, B;
    A.add(Xf, mf);
    B.add(V, nc);
    QM_TRY(gram<T>(s->gram_part, s->G0, A, B, mf, nc, s->ng, st));
    hipLaunchKernelGGL((fixed_coef_kernel<CPLX, T>), dim3((unsigned)(((int64_t)(mf + nc) * nc + 255) / 256)),
                       dim3(256), 0, st, (T *)s->U, s->G0, f, s->na, c);
    QM_HIP(hipGetLastError());
    ColBlocks Z, Yb;
    Z.add(Xf, mf);
    Z.add(V, nc);
    Yb.add(Y, nc);
    return update<T>(Z, Yb, (const T *)s->U, mf + nc, nc, s->ng, st);
}

// SS = Cb^H M Cb for M = S, H (n3 x n3), Cb n3 x 2N.
template <bool CPLX>
int recurrence(qm_lobpcg *s, int n3) {
    const int n2 = 2 * s->na;
    if constexpr (CPLX) {
        const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
        auto *Cb = (rocblas_double_complex *)s->Cb, *T1 = (rocblas_double_complex *)s->T1;
        QM_BLAS(rocblas_zgemm(s->blas, rocblas_operation_none, rocblas_operation_none, n3, n2, n3, &one,
                              (rocblas_double_complex *)s->S, n3, Cb, n3, &zero, T1, n3));
        QM_BLAS(rocblas_zgemm(s->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, n2, n2, n3,
                              &one, Cb, n3, T1, n3, &zero, (rocblas_double_complex *)s->SS, n2));
        QM_BLAS(rocblas_zgemm(s->blas, rocblas_operation_none, rocblas_operation_none, n3, n2, n3, &one,
                              (rocblas_double_complex *)s->H, n3, Cb, n3, &zero, T1, n3));
        QM_BLAS(rocblas_zgemm(s->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, n2, n2, n3,
                              &one, Cb, n3, T1, n3, &zero, (rocblas_double_complex *)s->HH, n2));
    } else {
        const double one = 1.0, zero = 0.0;
        auto *Cb = (double *)s->Cb, *T1 = (double *)s->T1;
        QM_BLAS(rocblas_dgemm(s->blas, rocblas_operation_none, rocblas_operation_none, n3, n2, n3, &one,
                              (double *)s->S, n3, Cb, n3, &zero, T1, n3));
        QM_BLAS(rocblas_dgemm(s->blas, rocblas_operation_transpose, rocblas_operation_none, n2, n2, n3, &one, Cb, n3,
                              T1, n3, &zero, (double *)s->SS, n2));
        QM_BLAS(rocblas_dgemm(s->blas, rocblas_operation_none, rocblas_operation_none, n3, n2, n3, &one,
                              (double *)s->H, n3, Cb, n3, &zero, T1, n3));
        QM_BLAS(rocblas_dgemm(s->blas, rocblas_operation_transpose, rocblas_operation_none, n2, n2, n3, &one, Cb, n3,
                              T1, n3, &zero, (double *)s->HH, n2));
    }
    return QM_OK;
}

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A captured segment of one iteration (replayed while its key recurs in a call).
struct SegKey {
    int seg;
    const void *X, *R, *P, *AX, *AR, *AP, *part;
    bool hp, rec;
    bool operator<(const SegKey &o) const {
        return std::tie(seg, X, R, P, AX, AR, AP, part, hp, rec) <
               std::tie(o.seg, o.X, o.R, o.P, o.AX, o.AR, o.AP, o.part, o.hp, o.rec);
    }
};
struct SegGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

// Captures fn(cs) on the capture stream cs into g (nothing runs).  A capture
// that fails, or whose graph holds a memset node (such nodes have replayed
// wrongly on this hardware), leaves g.exec null; fn then runs eagerly.
template <typename F>
int capture_segment(SegGraph &g, F &fn, hipStream_t cs) {
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed) != hipSuccess) {
        (void)hipGetLastError();
        return QM_OK;
    }
    ++g_internal_capture;  // this graph is destroyed before the call returns: no pin
    const int rc = fn(cs);
    --g_internal_capture;
    hipStream_t st = cs;
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    (void)hipGetLastError();
    if (rc != QM_OK || ec != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        return QM_OK;
    }
    size_t nn = 0;
    bool ok = hipGraphGetNodes(graph, nullptr, &nn) == hipSuccess && nn > 0;
    if (ok) {
        std::vector<hipGraphNode_t> nodes(nn);
        ok = hipGraphGetNodes(graph, nodes.data(), &nn) == hipSuccess;
        for (size_t i = 0; ok && i < nn; ++i) {
            hipGraphNodeType t;
            ok = hipGraphNodeGetType(nodes[i], &t) == hipSuccess && t != hipGraphNodeTypeMemset;
        }
    }
    if (ok) ok = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (!ok) {
        g.exec = nullptr;
        (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return QM_OK;
    }
    g.graph = graph;
    return QM_OK;
}

// Runs a segment: replays its graph if this key was captured, captures it on
// the key's second occurrence (while ok), else runs it eagerly.  A segment
// that cannot be captured turns graphs off for the rest of the call.
template <typename F>
int run_segment(std::map<SegKey, SegGraph> &graphs, const SegKey &key, F &fn, bool &ok, hipStream_t st,
                hipStream_t cs, int &replays) {
    auto g = graphs.find(key);
    if (g != graphs.end() && g->second.exec) {
        QM_HIP(hipGraphLaunch(g->second.exec, st));
        ++replays;
        return QM_OK;
    }
    if (ok && g != graphs.end()) {
        QM_TRY(capture_segment(g->second, fn, cs));
        if (g->second.exec) {
            QM_HIP(hipGraphLaunch(g->second.exec, st));
            ++replays;
            return QM_OK;
        }
        ok = false;
    }
    graphs[key];
    return fn(st);
}

// num_fix == nullptr: no locking (every band iterates).  Otherwise *num_fix is
// the caller's fixed-band count on entry and receives the new one.
template <typename T, bool CPLX>
int solve_impl(qm_lobpcg *s, T *x_user, double *eigenvalues, double *residual_norms, int iter_max, int ik,
               int first_time, int *num_fix, double res_threshold, int *iters_done, hipStream_t st) {
    using SmT = typename Sm<CPLX>::t;
    qm_ctx *ctx = s->ctx;
    const int Nall = s->N, c = s->c;
    const int64_t ng = s->ng;
    const int f0 = num_fix ? *num_fix : 0;
    // first column of band b of a block (SoA: c columns of ng per band)
    auto band = [&](T *p, int b) { return p + (int64_t)c * b * ng; };
    const double dV = ctx->hx * ctx->hy * ctx->hz;
    QM_BLAS(rocblas_set_stream(s->blas, st));
    for (double &v : s->t) v = 0.0;
    s->fallbacks = 0;
    s->rdrops = 0;
    s->stedc_retries = 0;
    s->ortho_rejects = 0;
    s->eig_note.clear();
    const double t_start = now_ms();
    // end of a phase: t[ph] receives the stream time since the previous mark
    s->ev_used = 0;
    s->graph_replays = 0;
    std::map<SegKey, SegGraph> graphs;  // this call's segment graphs
    struct GraphsGuard {                 // destroyed when the call returns (the stream is idle by then)
        std::map<SegKey, SegGraph> &m;
        hipStream_t st;
        ~GraphsGuard() {
            if (m.empty()) return;
            (void)hipStreamSynchronize(st);
            for (auto &kv : m) {
                if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
                if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
            }
        }
    } graphs_guard{graphs, st};
    bool graphs_ok = s->graphs;
    auto mark = [&](int ph) -> int {
        if (s->ev_used == s->ev.size()) {
            hipEvent_t e;
            QM_HIP(hipEventCreate(&e));
            s->ev.push_back(e);
            s->ev_phase.push_back(0);
        }
        QM_HIP(hipEventRecord(s->ev[s->ev_used], st));
        s->ev_phase[s->ev_used++] = ph;
        return QM_OK;
    };
    QM_TRY(mark(-1));

    // Block bases (all N bands) and, from here on, views of the iterated bands:
    // every block is viewed from the same first band, so rotating views rotates
    // the blocks.
    T *Xb = (T *)s->X, *Rb = (T *)s->R, *Pb = (T *)s->P, *AXb = (T *)s->AX;

    // x := sqrt(dV) x, all bands (lobpcg_z_multi_mpi.h:1851); for the bands not
    // fixed on entry: A x, eps = Re x^H A x, r = Ax - eps x, |r|^2 (1863-1918)
    QM_TRY(scale_bands<T>(Xb, x_user, nullptr, std::sqrt(dV), Nall, c, ng, st));
    QM_TRY(mark(4));
    const int na0 = Nall - f0;
    QM_TRY(hpsi<T>(s, band(AXb, f0), band(Xb, f0), na0, ik, st));
    QM_TRY(mark(0));
    QM_TRY(band_reduce<T>(s->red_part, 0, s->eps, nullptr, nullptr, band(Xb, f0), band(AXb, f0), nullptr, na0, c, ng,
                          st));
    QM_TRY(band_reduce<T>(s->red_part, 1, s->rnorm, s->xnorm, band(Rb, f0), band(Xb, f0), band(AXb, f0), s->eps, na0,
                          c, ng, st));
    std::vector<double> h_norm(Nall), h_eps(Nall), h_rn0(na0);
    QM_TRY(mark(4));
    if (num_fix) {  // the selection below reads the entry residuals
        QM_HIP(hipMemcpyAsync(h_rn0.data(), s->rnorm, sizeof(double) * na0, hipMemcpyDeviceToHost, st));
        QM_HIP(hipStreamSynchronize(st));
    }

    // num_fix: the first band at or above the entry count whose |r|^2 exceeds
    // threshold^2, at most N/2 (1962-1986); it never decreases within a call
    int f = f0;
    if (num_fix) {
        const double thr2 = res_threshold * res_threshold;
        int pos = Nall;
        for (int k = f0; k < Nall; ++k)
            if (thr2 < h_rn0[k - f0]) {
                pos = k;
                break;
            }
        f = std::min(pos, Nall / 2);
        *num_fix = f;
    }
    const int d = f - f0;  // entry results of the iterated bands start at index d
    const int N = Nall - f, nc = c * N;
    s->na = N;
    int rn_off = d;  // s->rnorm index of the first iterated band
    // evk: the Ritz values returned (entry Rayleigh quotients until an iteration is accepted)
    QM_HIP(hipMemcpyAsync(s->evk, s->eps + d, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    T *X = band(Xb, f), *R = band(Rb, f), *P = band(Pb, f), *AX = band(AXb, f), *AR = band((T *)s->AR, f),
      *AP = band((T *)s->AP, f);
    T *W[4] = {band((T *)s->W[0], f), band((T *)s->W[1], f), band((T *)s->W[2], f), band((T *)s->W[3], f)};
    hipLaunchKernelGGL(inv_sqrt_kernel, dim3(nblk(N, 64)), dim3(64), 0, st, s->dr, s->rnorm + d, N, 1.0);
    QM_HIP(hipGetLastError());
    bool hp = false;
    if (iter_max > 0) {
        if (f > 0) {
            // both X buffers of the rotation carry the fixed bands
            QM_HIP(hipMemcpyAsync(s->W[0], Xb, s->tsz * (size_t)c * f * ng, hipMemcpyDeviceToDevice, st));
        }
        if (!first_time && s->have_p) {
            // P of the iterated bands loses its fixed-band components when the
            // count changed (2058-2063), then A P
            if (f != f0) {
                QM_TRY((orthogonalise_fixed<T, CPLX>(s, W[1], Xb, f, P, st)));
                std::swap(P, W[1]);
            }
            QM_TRY(mark(1));
            QM_TRY(hpsi<T>(s, AP, P, N, ik, st));
            QM_TRY(mark(0));
            hp = true;
        }
    }

    int it_done = 0;
    bool rolled_back = false;
    for (int it = 0; it < iter_max; ++it) {
        it_done = it + 1;
        const int nbk = hp ? 3 : 2;
        // R loses its fixed-band components (2080); R <- orthonormal basis of (1 - X X^H) R
        if (f > 0) {
            QM_TRY((orthogonalise_fixed<T, CPLX>(s, AR, Xb, f, R, st)));
            std::swap(R, AR);
        }
        QM_TRY((orthonormalise_r<T, CPLX>(s, AR, X, R, st)));
        std::swap(R, AR);
        QM_TRY(mark(1));
        // A R, then the Gram products: neither holds a host decision or a library
        // call, so each is a HIP graph from the second time its block pointers
        // recur in this call (the buffer rotation gives two such sets; the first
        // run is eager, so buffers and the stencil shape have settled by then)
        const bool rec = it % kRestLimit == 0;
        auto seg_hpsi = [&](hipStream_t sq) -> int { return hpsi<T>(s, AR, R, N, ik, sq); };
        auto seg_gram = [&](hipStream_t sq) -> int {
            // Gram products
            {
                ColBlocks A, B;
                A.add(X, nc);
                A.add(R, nc);
                if (hp) A.add(P, nc);
                B.add(R, nc);
                B.add(AR, nc);
                QM_TRY(gram<T>(s->gram_part, s->G1, A, B, A.total(), B.total(), ng, sq));
                if (it % kRestLimit == 0) {
                    ColBlocks A2, B2;
                    A2.add(X, nc);
                    B2.add(X, nc);
                    if (hp) {
                        A2.add(P, nc);
                        B2.add(P, nc);
                    }
                    B2.add(AX, nc);
                    if (hp) B2.add(AP, nc);
                    QM_TRY(gram<T>(s->gram_part, s->G2, A2, B2, A2.total(), B2.total(), ng, sq));
                    hipLaunchKernelGGL((extract_blocks_kernel<CPLX>), dim3(nblk((int64_t)N * N)), dim3(256), 0, sq,
                                       (SmT *)s->blk, s->G2, N, c, hp);
                    QM_HIP(hipGetLastError());
                }
            }
            return QM_OK;
        };
        QM_TRY(run_segment(graphs, SegKey{0, X, R, P, AX, AR, AP, nullptr, hp, false}, seg_hpsi, graphs_ok, st,
                           s->gstream, s->graph_replays));
        QM_TRY(mark(0));
        QM_TRY(run_segment(graphs, SegKey{1, X, R, P, AX, AR, AP, s->gram_part.p, hp, rec}, seg_gram, graphs_ok, st,
                           s->gstream, s->graph_replays));
        QM_TRY(mark(1));
        // subspace eigenproblem: Cholesky route; if S is not numerically positive
        // definite or its pivots span more than cond_limit, retry without P
        // ([X R], as on the first iteration), then fall back to the S eigenbasis
        int nbk_used = nbk;
        const void *evec = s->Hw;
        const double cond_limit = sizeof(T) == 8 ? s->cond64 : s->cond32;
        hipLaunchKernelGGL((row_abs_sum_kernel<CPLX>), dim3(N), dim3(64), 0, st, s->svw,
                           (const SmT *)s->blk + (int64_t)B_XAX * N * N, N);
        QM_HIP(hipGetLastError());
        hipLaunchKernelGGL(penalty_kernel, dim3(1), dim3(256), 0, st, s->pen, (const double *)s->svw, N);
        QM_HIP(hipGetLastError());
        for (int attempt = 0;; ++attempt) {
            const int n3a = nbk_used * N;
            hipLaunchKernelGGL((assemble_kernel<CPLX>), dim3(nblk((int64_t)n3a * n3a)), dim3(256), 0, st,
                               (SmT *)s->S, (SmT *)s->H, (const SmT *)s->blk, s->G1, s->dr, N, c, nbk_used,
                               nbk * nc, s->rkeep, s->pen);
            QM_HIP(hipGetLastError());
            QM_TRY(chol_factor<CPLX>(s, n3a, st));
            int h_info = 0;
            double h_ratio = 0.0;
            QM_HIP(hipMemcpyAsync(&h_info, s->info, sizeof(int), hipMemcpyDeviceToHost, st));
            QM_HIP(hipMemcpyAsync(&h_ratio, s->svals, sizeof(double), hipMemcpyDeviceToHost, st));
            QM_HIP(hipStreamSynchronize(st));
            if (h_info == 0 && h_ratio >= cond_limit) {
                QM_TRY(chol_solve<CPLX>(s, n3a, st));
                // a tridiagonal eigensolve that does not converge (a barely
                // definite S makes U^-H H U^-1 badly scaled), or Ritz vectors
                // that have lost their S-orthonormality (S too ill-conditioned
                // for its Cholesky factor: the update would fail its norm check
                // and its Ritz values can be spurious), are handled like a
                // failed Cholesky: without P, then the S eigenbasis
                const SmT *cv = (const SmT *)(s->evec_t1 ? s->T1 : s->Hw);
                if constexpr (CPLX) {
                    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
                    QM_BLAS(rocblas_zgemm(s->blas, rocblas_operation_none, rocblas_operation_none, n3a, N, n3a, &one,
                                          (const rocblas_double_complex *)s->S, n3a,
                                          (const rocblas_double_complex *)cv, n3a, &zero,
                                          (rocblas_double_complex *)s->rt, n3a));
                } else {
                    const double one = 1.0, zero = 0.0;
                    QM_BLAS(rocblas_dgemm(s->blas, rocblas_operation_none, rocblas_operation_none, n3a, N, n3a, &one,
                                          (const double *)s->S, n3a, (const double *)cv, n3a, &zero,
                                          (double *)s->rt, n3a));
                }
                hipLaunchKernelGGL((ritz_orth_kernel<CPLX>), dim3(N), dim3(256), 0, st, s->svw, cv,
                                   (const SmT *)s->rt, n3a);
                QM_HIP(hipGetLastError());
                hipLaunchKernelGGL(max_reduce_kernel, dim3(1), dim3(256), 0, st, s->svw, N);
                QM_HIP(hipGetLastError());
                double h_dev = 0.0;
                QM_HIP(hipMemcpyAsync(&h_info, s->info, sizeof(int), hipMemcpyDeviceToHost, st));
                QM_HIP(hipMemcpyAsync(&h_dev, s->svw, sizeof(double), hipMemcpyDeviceToHost, st));
                QM_HIP(hipStreamSynchronize(st));
                if (h_info == 0 && h_dev <= kRitzOrthTol) {
                    evec = cv;
                    break;
                }
                if (h_info == 0) {
                    ++s->ortho_rejects;
                    s->eig_note += " (Cholesky-route Ritz vectors of order " + std::to_string(n3a) +
                                   " off S-orthonormal by " + std::to_string(h_dev) + ")";
                } else {
                    s->eig_note += " (Cholesky-route solve of order " + std::to_string(n3a) + " info " +
                                   std::to_string(h_info) + ")";
                }
            }
            if (nbk_used == 3) {  // drop P and retry
                nbk_used = 2;
                ++s->fallbacks;
                continue;
            }
            QM_TRY(svqb_eig<CPLX>(s, n3a, sizeof(T) == 8 ? 1.0e-12 : cond_limit, st));
            QM_HIP(hipMemcpyAsync(&h_info, s->info, sizeof(int), hipMemcpyDeviceToHost, st));
            QM_HIP(hipStreamSynchronize(st));
            QM_CHECK(h_info == 0, QM_ERR_NUMERIC,
                     "qm_lobpcg_solve: subspace eigensolve failed (rocSOLVER info " + std::to_string(h_info) +
                         "; " + s->eig_note + ")");
            evec = s->Ev;
            ++s->fallbacks;
            break;
        }
        const int n3 = nbk_used * N;
        const bool hp_used = nbk_used == 3;
        QM_TRY(mark(2));
        // recurrence, update matrix, vector updates
        hipLaunchKernelGGL((build_cb_kernel<CPLX>), dim3(nblk((int64_t)n3 * 2 * N)), dim3(256), 0, st, (SmT *)s->Cb,
                           (const SmT *)evec, N, n3);
        QM_HIP(hipGetLastError());
        QM_TRY(recurrence<CPLX>(s, n3));
        hipLaunchKernelGGL((next_blocks_kernel<CPLX, T>), dim3(nblk((int64_t)N * N)), dim3(256), 0, st, (SmT *)s->blk,
                           s->dp, (const SmT *)s->SS, (const SmT *)s->HH, N);
        QM_HIP(hipGetLastError());
        hipLaunchKernelGGL((build_update_kernel<CPLX, T>), dim3(nblk((int64_t)n3 * 2 * N)), dim3(256), 0, st,
                           (T *)s->U, (const SmT *)s->Cb, s->dr, s->dp, N, n3);
        QM_HIP(hipGetLastError());
        {
            ColBlocks Z, Y, AZ, AY;
            Z.add(X, nc);
            Z.add(R, nc);
            if (hp_used) Z.add(P, nc);
            AZ.add(AX, nc);
            AZ.add(AR, nc);
            if (hp_used) AZ.add(AP, nc);
            Y.add(W[0], nc);
            Y.add(W[1], nc);
            AY.add(W[2], nc);
            AY.add(W[3], nc);
            QM_TRY(update<T>(Z, Y, (const T *)s->U, c * n3, 2 * nc, ng, st));
            QM_TRY(update<T>(AZ, AY, (const T *)s->U, c * n3, 2 * nc, ng, st));
        }
        // the residual norms of the current x, returned if the update is rolled back
        QM_HIP(hipMemcpyAsync(s->rnorm + 2 * Nall, s->rnorm + rn_off, sizeof(double) * N, hipMemcpyDeviceToDevice,
                              st));
        // R' = AX' - X' L, |R'|^2, |X'|^2
        QM_TRY(band_reduce<T>(s->red_part, 1, s->rnorm, s->xnorm, AR, W[0], W[2], s->evals, N, c, ng, st));
        rn_off = 0;
        QM_TRY(mark(3));
        QM_HIP(hipMemcpyAsync(h_norm.data(), s->xnorm, sizeof(double) * N, hipMemcpyDeviceToHost, st));
        QM_HIP(hipStreamSynchronize(st));
        double maxdev = 0.0;
        for (int n = 0; n < N; ++n)  // a non-finite norm (a failed update) rolls back too
            maxdev = std::isfinite(h_norm[n]) ? std::max(maxdev, std::fabs(1.0 - h_norm[n])) : HUGE_VAL;
        if (maxdev > s->limit_normx) {
            // rollback and escape (lobpcg_z_multi_mpi.h:2611-2666): P = R and leave the
            // loop.  The reference rescales the old x by the norms of the rejected
            // update (2639); the old x is renormalised with its own norms here, and
            // its Ritz values and residual norms are the ones returned.
            QM_HIP(hipMemcpyAsync(s->rnorm, s->rnorm + 2 * Nall, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
            QM_HIP(hipMemcpyAsync(P, R, s->tsz * (size_t)nc * ng, hipMemcpyDeviceToDevice, st));
            s->have_p = true;
            QM_TRY(band_reduce<T>(s->red_part, 2, s->xnorm, nullptr, nullptr, X, nullptr, nullptr, N, c, ng, st));
            hipLaunchKernelGGL(inv_sqrt_kernel, dim3(nblk(N, 64)), dim3(64), 0, st, s->dp, s->xnorm, N, dV);
            QM_HIP(hipGetLastError());
            QM_TRY(scale_bands<T>(band(x_user, f), X, s->dp, 0.0, N, c, ng, st));
            QM_TRY(mark(4));
            rolled_back = true;
            break;
        }
        // the residual block written by band_reduce sits in AR (free after the update); rotate buffers
        std::swap(X, W[0]);
        std::swap(P, W[1]);
        std::swap(AX, W[2]);
        std::swap(AP, W[3]);
        std::swap(R, AR);
        hp = true;
        QM_HIP(hipMemcpyAsync(s->evk, s->evals, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(inv_sqrt_kernel, dim3(nblk(N, 64)), dim3(64), 0, st, s->dr, s->rnorm, N, 1.0);
        QM_HIP(hipGetLastError());
    }
    // fixed bands: x_out = x / sqrt(dV) (2651-2652, 2796-2797); with no
    // iteration every band leaves this way (1992-2004)
    const int nunscale = iter_max > 0 ? f : Nall;
    if (nunscale > 0) QM_TRY(scale_bands<T>(x_user, Xb, nullptr, 1.0 / std::sqrt(dV), nunscale, c, ng, st));
    if (!rolled_back) {
        // x_out = x / sqrt(|x|^2 dV)  (2796-2800)
        if (iter_max > 0) {
            hipLaunchKernelGGL(inv_sqrt_kernel, dim3(nblk(N, 64)), dim3(64), 0, st, s->dp, s->xnorm, N, dV);
            QM_HIP(hipGetLastError());
            QM_TRY(scale_bands<T>(band(x_user, f), X, s->dp, 0.0, N, c, ng, st));
        }
        s->have_p = iter_max > 0 ? hp : (!first_time && s->have_p);
    }
    if constexpr (sizeof(T) == 4) {
        if (iter_max > 0) {  // also after a rollback: the returned (old) vectors
            // mixed precision: band energies as FP64 Rayleigh quotients of the returned
            // FP32 vectors with the FP64 operator, and FP64 residual norms
            // (second order in the vector error, so the FP32 operator's rounding drops out)
            const int64_t nel = (int64_t)nc * ng;
            hipLaunchKernelGGL(cast_kernel, dim3((unsigned)std::min<int64_t>((nel + 255) / 256, 8192)), dim3(256), 0,
                               st, s->X64, (const float *)band(x_user, f), nel);
            QM_HIP(hipGetLastError());
            QM_TRY(hpsi<double>(s, s->AX64, s->X64, N, ik, st));
            QM_TRY(band_reduce<double>(s->red_part, 0, s->eps, nullptr, nullptr, s->X64, s->AX64, nullptr, N, c, ng,
                                       st));
            QM_TRY(band_reduce<double>(s->red_part, 2, s->xnorm, nullptr, nullptr, s->X64, nullptr, nullptr, N, c, ng,
                                       st));
            hipLaunchKernelGGL(ratio_kernel, dim3(nblk(N, 64)), dim3(64), 0, st, s->evals, s->eps, s->xnorm, N);
            QM_HIP(hipGetLastError());
            // |H x - e x|^2 dV (x is dV-normalised: scaled-unit residual)
            QM_TRY(band_reduce<double>(s->red_part, 1, s->rnorm, nullptr, s->AX64, s->X64, s->AX64, s->evals, N, c, ng,
                                       st));
            hipLaunchKernelGGL(ratio_kernel, dim3(nblk(N, 64)), dim3(64), 0, st, s->rnorm, s->rnorm, s->xnorm, N);
            QM_HIP(hipGetLastError());
            QM_HIP(hipMemcpyAsync(s->evk, s->evals, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        }
    }
    // keep the rotated buffers (block bases)
    auto base = [&](T *p) { return p - (int64_t)c * f * ng; };
    s->X = base(X);
    s->R = base(R);
    s->P = base(P);
    s->AX = base(AX);
    s->AR = base(AR);
    s->AP = base(AP);
    for (int i = 0; i < 4; ++i) s->W[i] = base(W[i]);
    std::vector<double> h_rn(N);
    QM_TRY(mark(4));
    QM_HIP(hipMemcpyAsync(h_eps.data(), s->evk, sizeof(double) * N, hipMemcpyDeviceToHost, st));
    QM_HIP(hipMemcpyAsync(h_rn.data(), s->rnorm + rn_off, sizeof(double) * N, hipMemcpyDeviceToHost, st));
    QM_HIP(hipStreamSynchronize(st));
    for (size_t i = 1; i < s->ev_used; ++i) {
        float ms = 0.0f;
        QM_HIP(hipEventElapsedTime(&ms, s->ev[i - 1], s->ev[i]));
        s->t[s->ev_phase[i]] += ms;
    }
    // entries of the fixed bands are left as the caller passed them (2846-2848)
    for (int n = 0; n < N; ++n) {
        eigenvalues[f + n] = h_eps[n];
        if (residual_norms) residual_norms[f + n] = h_rn[n];
    }
    if (iters_done) *iters_done = it_done;
    s->t[5] = now_ms() - t_start;
    return QM_OK;
}

}  // namespace
}  // namespace qm

using namespace qm;

extern "C" {

int qm_lobpcg_create(qm_lobpcg **out, qm_ctx *ctx, int nbands, int ncomp, int precision) {
    QM_CHECK(out && ctx, QM_ERR_ARG, "qm_lobpcg_create: NULL argument");
    *out = nullptr;
    QM_CHECK(nbands >= 1 && nbands <= 2048, QM_ERR_ARG, "qm_lobpcg_create: nbands must be in [1, 2048]");
    QM_CHECK(ncomp == 1 || ncomp == 2, QM_ERR_ARG, "qm_lobpcg_create: ncomp must be 1 or 2");
    QM_CHECK(precision == 32 || precision == 64, QM_ERR_ARG, "qm_lobpcg_create: precision must be 32 or 64");
    QM_TRY(activate(ctx));
    qm_lobpcg *s = new qm_lobpcg();
    s->ctx = ctx;
    s->N = nbands;
    s->c = ncomp;
    s->prec = precision;
    s->ng = ctx->ng;
    s->tsz = precision == 64 ? 8 : 4;
    const int N = nbands, nc = ncomp * N, n3 = 3 * N;
    const size_t blk = s->tsz * (size_t)nc * s->ng;
    const size_t esz = ncomp == 2 ? 16 : 8;  // small-matrix element
    int st = QM_OK;
    void **vec[] = {&s->X, &s->R, &s->P, &s->AX, &s->AR, &s->AP, &s->W[0], &s->W[1], &s->W[2], &s->W[3]};
    for (void **p : vec)
        if (st == QM_OK) st = dalloc(p, blk);
    if (st == QM_OK) st = dalloc((void **)&s->G1, sizeof(double) * (size_t)(3 * nc) * (2 * nc));
    if (st == QM_OK) st = dalloc((void **)&s->G2, sizeof(double) * (size_t)(2 * nc) * (4 * nc));
    void **sm[] = {&s->S, &s->H, &s->Sw, &s->Hw};
    for (void **p : sm)
        if (st == QM_OK) st = dalloc(p, esz * (size_t)n3 * n3);
    if (st == QM_OK) st = dalloc(&s->blk, esz * (size_t)B_COUNT * N * N);
    if (st == QM_OK) st = dalloc(&s->Cb, esz * (size_t)n3 * 2 * N);
    if (st == QM_OK) st = dalloc(&s->T1, esz * (size_t)n3 * n3);
    if (st == QM_OK) st = dalloc(&s->Ev, esz * (size_t)n3 * n3);
    if (st == QM_OK) st = dalloc(&s->SS, esz * (size_t)4 * N * N);
    if (st == QM_OK) st = dalloc(&s->HH, esz * (size_t)4 * N * N);
    if (st == QM_OK) st = dalloc(&s->U, s->tsz * (size_t)(ncomp * n3) * (2 * nc));
    if (st == QM_OK) st = dalloc((void **)&s->G0, sizeof(double) * (size_t)(2 * nc) * nc);
    if (precision == 32) {
        if (st == QM_OK) st = dalloc((void **)&s->X64, sizeof(double) * (size_t)nc * s->ng);
        if (st == QM_OK) st = dalloc((void **)&s->AX64, sizeof(double) * (size_t)nc * s->ng);
    }
    if (st == QM_OK) st = dalloc(&s->XRn, esz * (size_t)N * N);
    if (st == QM_OK) st = dalloc(&s->Mr, esz * (size_t)N * N);
    if (st == QM_OK) st = dalloc(&s->Wr, esz * (size_t)2 * N * N);
    if (st == QM_OK) st = dalloc(&s->Ur, s->tsz * (size_t)(2 * nc) * nc);
    double **dv[] = {&s->evals, &s->ework, &s->eps, &s->dr, &s->dp, &s->rnorm, &s->xnorm, &s->svals, &s->svw,
                     &s->rkeep};
    for (double **p : dv)
        if (st == QM_OK) st = dalloc((void **)p, sizeof(double) * n3);
    if (st == QM_OK) st = dalloc((void **)&s->info, sizeof(int));
    if (st == QM_OK) st = dalloc((void **)&s->evk, sizeof(double) * N);
    if (st == QM_OK) st = dalloc((void **)&s->pen, sizeof(double));
    if (st == QM_OK) st = dalloc(&s->rt, esz * (size_t)n3 * N);
    if (st == QM_OK) st = dalloc(&s->psync, 512);
    if (st == QM_OK && hipStreamCreateWithFlags(&s->gstream, hipStreamNonBlocking) != hipSuccess) {
        s->gstream = nullptr;
        set_error("qm_lobpcg_create: hipStreamCreateWithFlags failed");
        st = QM_ERR_HIP;
    }
    if (st == QM_OK) st = dalloc((void **)&s->trd, sizeof(double) * 5 * (size_t)n3);  // d, e, tau (complex), e copy
    if (st == QM_OK) st = dalloc((void **)&s->VW, sizeof(double) * (size_t)kTrdPanelCols * n3);
    if (st == QM_OK && rocblas_create_handle(&s->blas) != rocblas_status_success) {
        set_error("qm_lobpcg_create: rocblas_create_handle failed");
        st = QM_ERR_HIP;
    }
    if (st != QM_OK) {
        lfree(s);
        delete s;
        return st;
    }
    *out = s;
    return QM_OK;
}

void qm_lobpcg_destroy(qm_lobpcg *s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipDeviceSynchronize();
    lfree(s);
    delete s;
}

}  // extern "C"

namespace qm {
namespace {
int solve_dispatch(qm_lobpcg *s, void *d_x, double *eigenvalues, double *residual_norms, int iter_max, int ik,
                   int first_time, int *num_fix, double res_threshold, int *iters_done, void *stream) {
    QM_CHECK(s && d_x && eigenvalues, QM_ERR_ARG, "qm_lobpcg_solve: NULL argument");
    QM_CHECK(iter_max >= 0, QM_ERR_ARG, "qm_lobpcg_solve: iter_max < 0");
    QM_CHECK(ik >= 0 && ik < (int)s->ctx->kpts.size(), QM_ERR_ARG, "qm_lobpcg_solve: k index out of range");
    QM_CHECK(s->c == 2 || s->ctx->kpts[ik].gamma, QM_ERR_ARG, "qm_lobpcg_solve: real bands need a Gamma k-point");
    QM_TRY(activate(s->ctx));
    hipStream_t st = (hipStream_t)stream;
    if (s->prec == 64) {
        if (s->c == 2)
            return solve_impl<double, true>(s, (double *)d_x, eigenvalues, residual_norms, iter_max, ik, first_time,
                                            num_fix, res_threshold, iters_done, st);
        return solve_impl<double, false>(s, (double *)d_x, eigenvalues, residual_norms, iter_max, ik, first_time,
                                         num_fix, res_threshold, iters_done, st);
    }
    if (s->c == 2)
        return solve_impl<float, true>(s, (float *)d_x, eigenvalues, residual_norms, iter_max, ik, first_time,
                                       num_fix, res_threshold, iters_done, st);
    return solve_impl<float, false>(s, (float *)d_x, eigenvalues, residual_norms, iter_max, ik, first_time, num_fix,
                                    res_threshold, iters_done, st);
}
}  // namespace
}  // namespace qm

extern "C" {

int qm_lobpcg_solve(qm_lobpcg *s, void *d_x, double *eigenvalues, double *residual_norms, int iter_max, int ik,
                    int first_time, int *iters_done, void *stream) {
    return solve_dispatch(s, d_x, eigenvalues, residual_norms, iter_max, ik, first_time, nullptr, 0.0, iters_done,
                          stream);
}

int qm_lobpcg_solve_locking(qm_lobpcg *s, void *d_x, double *eigenvalues, double *residual_norms, int iter_max,
                            int ik, int first_time, int *num_fix, double residual_threshold, int *iters_done,
                            void *stream) {
    QM_CHECK(s && num_fix, QM_ERR_ARG, "qm_lobpcg_solve_locking: NULL argument");
    QM_CHECK(*num_fix >= 0 && *num_fix <= s->N / 2, QM_ERR_ARG,
             "qm_lobpcg_solve_locking: num_fix must be in [0, nbands/2]");
    QM_CHECK(residual_threshold >= 0.0, QM_ERR_ARG, "qm_lobpcg_solve_locking: residual_threshold < 0");
    return solve_dispatch(s, d_x, eigenvalues, residual_norms, iter_max, ik, first_time, num_fix, residual_threshold,
                          iters_done, stream);
}

int qm_lobpcg_set_conditioning(qm_lobpcg *s, double min_pivot_ratio_fp64, double min_pivot_ratio_fp32) {
    QM_CHECK(s, QM_ERR_ARG, "qm_lobpcg_set_conditioning: NULL solver");
    QM_CHECK(min_pivot_ratio_fp64 >= 0.0 && min_pivot_ratio_fp32 >= 0.0 && min_pivot_ratio_fp64 <= 2.0 &&
                 min_pivot_ratio_fp32 <= 2.0,
             QM_ERR_ARG, "qm_lobpcg_set_conditioning: ratios must be in [0, 2]");
    s->cond64 = min_pivot_ratio_fp64;
    s->cond32 = min_pivot_ratio_fp32;
    return QM_OK;
}

int qm_lobpcg_fallbacks(qm_lobpcg *s) { return s ? s->fallbacks : QM_ERR_ARG; }

int qm_lobpcg_eigen_retries(qm_lobpcg *s, char *note, int len) {
    QM_CHECK(s, QM_ERR_ARG, "qm_lobpcg_eigen_retries: NULL solver");
    if (note && len > 0) {
        std::snprintf(note, (size_t)len, "%s", s->eig_note.c_str());
    }
    return s->stedc_retries + s->ortho_rejects;
}

int qm_lobpcg_set_rollback_limit(qm_lobpcg *s, double limit) {
    QM_CHECK(s, QM_ERR_ARG, "qm_lobpcg_set_rollback_limit: NULL solver");
    QM_CHECK(limit >= 0.0, QM_ERR_ARG, "qm_lobpcg_set_rollback_limit: limit must be >= 0");
    s->limit_normx = limit;
    return QM_OK;
}

int qm_lobpcg_set_graphs(qm_lobpcg *s, int enable) {
    QM_CHECK(s, QM_ERR_ARG, "qm_lobpcg_set_graphs: NULL solver");
    s->graphs = enable != 0;
    return QM_OK;
}

int qm_lobpcg_graph_replays(qm_lobpcg *s) { return s ? s->graph_replays : QM_ERR_ARG; }

int qm_lobpcg_timings(qm_lobpcg *s, double *t6) {
    QM_CHECK(s && t6, QM_ERR_ARG, "qm_lobpcg_timings: NULL argument");
    for (int i = 0; i < 6; ++i) t6[i] = s->t[i];
    return QM_OK;
}

}  // extern "C"
