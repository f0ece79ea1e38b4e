// Householder tridiagonal reduction of the LOBPCG subspace matrix by a group of
// cooperating workgroups that keep the whole matrix in LDS.
//
// LAPACK's dsytd2 / zhetd2 (uplo = 'L') reduce one column at a time: for
// column j, the reflector H_j = I - tau v v^H that zeroes A(j+2:n, j), then
// p = tau A22 v, w = p - (tau/2)(p^H v) v and A22 -= v w^H + w v^H.  The
// matrix-vector product touches the whole trailing matrix once per column, so
// a single workgroup streaming it from L2 is bound by one CU's bandwidth
// (~4 ms for n = 384).  Here G workgroups (one per CU) hold the rows
// r = g, g + G, g + 2G, ... of the matrix in LDS: p and the rank-2 update of
// a row are local to its owner, and only vectors cross workgroups through
// global memory.
//
// One grid barrier per column.  Before it, each workgroup writes its rows of
// p and its part of p^H v; after it, every workgroup forms w from all of p,
// updates its rows, and ALSO updates row j+1 itself from the copy of that row
// its owner published one column earlier: row j+1 after the update is the
// next column's x (A(j+2:n, j+1) is its conjugate), so every workgroup can
// form the next reflector without waiting for the owner.  The owner of row
// j+2 then publishes its updated row for the next column's look-ahead.  The
// global vectors are double-buffered by column parity; each reader of a
// buffer passes the barrier after the column that wrote it and before the
// column that overwrites it.  Every workgroup computes the reflector from the
// same data with the same operation order, so all hold the same v and tau.
//
// Output follows LAPACK's convention so rocSOLVER's stedc and ormtr / unmtr
// consume it: d, e (real), tau, and v (v_{j+1} = 1 implied) below the
// subdiagonal of column j of A.  The input is read from A's upper triangle
// (uplo 'U' defines the matrix; the diagonal's imaginary part is dropped).
//
// The barrier is a counter and a generation word in device memory, with
// agent-scope release / acquire ordering on the arrival and the wait; its
// waits are bounded: past
// kCoopSpinLimit polls a workgroup sets the error word and goes on, so the
// kernel always ends (the caller turns the error word into a failed solve).
// hipLaunchCooperativeKernel guarantees that all G workgroups are resident.
#include "qm_internal.h"

namespace qm {
namespace {

constexpr int kCoopThreads = 256;
constexpr int kCoopWaves = kCoopThreads / 64;
constexpr int kCoopLdsBytes = 150 * 1024;
constexpr long kCoopSpinLimit = 1L << 24;  // ~0.5 s of polling
constexpr int kCoopMaxPer = 8;              // vector elements per thread: n <= 2048
constexpr int kCoopMaxGroups = 256;
constexpr int kCoopRowsPerWave = 4;         // own rows per wave: nr <= 16

struct zdbl {
    double re, im;
};

__device__ inline double conj_(double a) { return a; }
__device__ inline zdbl conj_(zdbl a) { return {a.re, -a.im}; }
__device__ inline double mul(double a, double b) { return a * b; }
__device__ inline zdbl mul(zdbl a, zdbl b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline double add(double a, double b) { return a + b; }
__device__ inline zdbl add(zdbl a, zdbl b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline double sub(double a, double b) { return a - b; }
__device__ inline zdbl sub(zdbl a, zdbl b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline double abs2(double a) { return a * a; }
__device__ inline double abs2(zdbl a) { return a.re * a.re + a.im * a.im; }
__device__ inline double re_(double a) { return a; }
__device__ inline double re_(zdbl a) { return a.re; }
__device__ inline double im_(double) { return 0.0; }
__device__ inline double im_(zdbl a) { return a.im; }
template <typename T>
__device__ inline T make(double re, double im);
template <>
__device__ inline double make<double>(double re, double) {
    return re;
}
template <>
__device__ inline zdbl make<zdbl>(double re, double im) {
    return {re, im};
}
__device__ inline double shfl_xor(double a, int m) { return __shfl_xor(a, m); }
__device__ inline zdbl shfl_xor(zdbl a, int m) { return {__shfl_xor(a.re, m), __shfl_xor(a.im, m)}; }

// a - x conj(y) - y' conj(x'): the rank-2 update of one element
template <typename T>
__device__ inline T rank2(T a, T vr, T wc, T wr, T vc) {
    return sub(sub(a, mul(vr, conj_(wc))), mul(wr, conj_(vc)));
}

// Every wave waits for its own stores to complete (they are then in this
// XCD's L2); thread 0's agent-scope release writes the L2 back and its
// acquire invalidates this CU's caches, so after the closing workgroup
// barrier every wave reads the other workgroups' data from memory.
__device__ void coop_barrier(unsigned *count, unsigned *gen, unsigned nblk, unsigned &lg, int *err) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned g = lg;
        const unsigned arrived = __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nblk - 1) {
            __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(gen, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long spins = 0;
            while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (++spins > kCoopSpinLimit) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        lg = g + 1;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// sync[0]: arrival count, sync[32]: generation (separate 128-byte lines),
// sync[64]: error word.  All zero at launch.
template <typename T, int PER>
__global__ __launch_bounds__(kCoopThreads) void trd_coop_kernel(T *__restrict__ A, int lda, int n, int nr,
                                                                double *__restrict__ dout,
                                                                double *__restrict__ eout, T *__restrict__ tauout,
                                                                T *__restrict__ rowbuf, T *__restrict__ pbuf,
                                                                T *__restrict__ dots, unsigned *sync) {
    __shared__ __align__(16) unsigned char lds_raw[kCoopLdsBytes];
    __shared__ T wsum[kCoopWaves];
    __shared__ T dsh[kCoopMaxGroups];
    T *L = reinterpret_cast<T *>(lds_raw);  // own rows: row g + q G at L + q n
    T *xr = L + (size_t)nr * n;             // the current column's row (conjugate of x)
    T *v = xr + n;                          // reflector, global row index
    T *w = v + n;
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    unsigned *count = sync, *gen = sync + 32;
    int *err = reinterpret_cast<int *>(sync + 64);
    unsigned lg = 0;
    const int nr_own = g < n ? min(nr, (n - g + G - 1) / G) : 0;  // own rows r = g + q G < n

    for (int q = 0; q < nr; ++q) {
        const int r = g + q * G;
        if (r >= n) break;
        T *Lq = L + (size_t)q * n;
        for (int c = tid; c < n; c += kCoopThreads) {
            T a = c >= r ? A[r + (size_t)c * lda] : conj_(A[c + (size_t)r * lda]);
            if (c == r) a = make<T>(re_(a), 0.0);
            Lq[c] = a;
        }
    }
    __syncthreads();
    for (int r = 0; r < 2 && r < n; ++r)
        if (r % G == g) {
            const T *Lq = L + (size_t)(r / G) * n;
            for (int c = tid; c < n; c += kCoopThreads) rowbuf[(size_t)r * n + c] = Lq[c];
        }
    coop_barrier(count, gen, G, lg, err);
    for (int c = tid; c < n; c += kCoopThreads) xr[c] = rowbuf[c];
    __syncthreads();

    for (int j = 0; j < n - 1; ++j) {
        const int par = j & 1;
        // ---- reflector from x = conj(xr[j+1 .. n-1]) (LAPACK dlarfg / zlarfg).
        // Every wave forms |x_1..|^2 itself, in the same order, and takes lane
        // 0's total, so all waves of all workgroups hold the same scalars.
        double s = 0.0;
        for (int c = j + 2 + lane; c < n; c += 64) s += abs2(xr[c]);
        for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
        s = __shfl(s, 0);
        const T alpha = conj_(xr[j + 1]);
        T tau = make<T>(0.0, 0.0), scal = make<T>(0.0, 0.0);
        double beta = re_(alpha);
        if (s != 0.0 || im_(alpha) != 0.0) {
            beta = -copysign(sqrt(abs2(alpha) + s), re_(alpha));
            tau = make<T>((beta - re_(alpha)) / beta, -im_(alpha) / beta);
            const T den = sub(alpha, make<T>(beta, 0.0));
            const double dd = abs2(den);
            scal = make<T>(re_(den) / dd, -im_(den) / dd);  // 1 / (alpha - beta)
        }
        auto vof = [&](int c) { return c == j + 1 ? make<T>(1.0, 0.0) : mul(conj_(xr[c]), scal); };
        // ---- p = tau A22 v on own rows (a wave takes rows wave, wave + 4, ...,
        // all at once), partial p^H v
        T acc[kCoopRowsPerWave];
#pragma unroll
        for (int i = 0; i < kCoopRowsPerWave; ++i) acc[i] = make<T>(0.0, 0.0);
        // every slot loads (finished or missing rows from a clamped valid row,
        // their sums unused), so the loads of a column issue together
        const int qmax = nr_own > 0 ? nr_own - 1 : 0;
        for (int c = j + 1 + lane; c < n; c += 64) {
            const T vc = vof(c);
            T lv[kCoopRowsPerWave];
#pragma unroll
            for (int i = 0; i < kCoopRowsPerWave; ++i)
                lv[i] = L[(size_t)min(wave + i * kCoopWaves, qmax) * n + c];
#pragma unroll
            for (int i = 0; i < kCoopRowsPerWave; ++i) acc[i] = add(acc[i], mul(lv[i], vc));
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
            for (int i = 0; i < kCoopRowsPerWave; ++i) acc[i] = add(acc[i], shfl_xor(acc[i], m));
        if (lane == 0) {
            T dot = make<T>(0.0, 0.0);
#pragma unroll
            for (int i = 0; i < kCoopRowsPerWave; ++i) {
                const int q = wave + i * kCoopWaves;
                const int r = g + q * G;
                if (q < nr_own && r > j) {
                    const T p = mul(tau, acc[i]);
                    pbuf[(size_t)par * n + r] = p;
                    dot = add(dot, mul(conj_(p), vof(r)));
                }
            }
            wsum[wave] = dot;
        }
        for (int c = j + 1 + tid; c < n; c += kCoopThreads) v[c] = vof(c);
        if (g == 0) {
            if (tid == 0) {
                dout[j] = re_(xr[j]);
                eout[j] = beta;
                tauout[j] = tau;
            }
            for (int c = j + 2 + tid; c < n; c += kCoopThreads) A[c + (size_t)j * lda] = vof(c);
        }
        __syncthreads();
        if (tid == 0) {
            T t = make<T>(0.0, 0.0);
            for (int i = 0; i < kCoopWaves; ++i) t = add(t, wsum[i]);
            dots[(size_t)par * G + g] = t;
        }
        coop_barrier(count, gen, G, lg, err);
        // ---- one round of loads: the partial dots, p and the old row j+1
        const T *old = rowbuf + (size_t)((j + 1) & 1) * n;
        T pr[PER], orow[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int c = j + 1 + tid + k * kCoopThreads;
            if (c < n) {
                pr[k] = pbuf[(size_t)par * n + c];
                orow[k] = old[c];
            }
        }
        if (tid < G) dsh[tid] = dots[(size_t)par * G + tid];
        __syncthreads();
        // ---- w = p - (tau / 2) (p^H v) v; every wave sums the G partials in
        // the same order
        T t = make<T>(0.0, 0.0);
        for (int i = lane; i < G; i += 64) t = add(t, dsh[i]);
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) t = add(t, shfl_xor(t, m));
        if constexpr (sizeof(T) == 8) {
            t = make<T>(__shfl(re_(t), 0), 0.0);
        } else {
            t = make<T>(__shfl(re_(t), 0), __shfl(im_(t), 0));
        }
        const T a2 = mul(make<T>(-0.5, 0.0), mul(tau, t));
        // this thread's columns of v and w stay in registers for the updates
        T vc[PER], wc[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int c = j + 1 + tid + k * kCoopThreads;
            if (c < n) {
                vc[k] = v[c];
                wc[k] = add(pr[k], mul(a2, vc[k]));
                w[c] = wc[k];
            }
        }
        __syncthreads();
        // ---- look-ahead: row j+1 after this update, from its owner's copy
        const T vj = v[j + 1], wj = w[j + 1];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int c = j + 1 + tid + k * kCoopThreads;
            if (c < n) xr[c] = rank2(orow[k], vj, wc[k], wj, vc[k]);
        }
        // ---- A22 -= v w^H + w v^H on own rows r > j, four rows per pass (each
        // pass loads all four rows' columns, then stores them)
        const int q_lo = j >= g ? (j - g) / G + 1 : 0;  // first own row below j
        for (int q0 = q_lo; q0 < nr_own; q0 += 4) {
            T a[4][PER], vr[4], wr[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = min(q0 + u, nr_own - 1);
                const int r = g + q * G;
                vr[u] = v[r];
                wr[u] = w[r];
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const int c = j + 1 + tid + k * kCoopThreads;
                    if (c < n) a[u][k] = L[(size_t)q * n + c];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (q0 + u >= nr_own) break;
                T *Lq = L + (size_t)(q0 + u) * n;
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const int c = j + 1 + tid + k * kCoopThreads;
                    if (c < n) Lq[c] = rank2(a[u][k], vr[u], wc[k], wr[u], vc[k]);
                }
            }
        }
        __syncthreads();
        // ---- the owner of row j+2 publishes it for the next column's look-ahead
        if (j + 2 < n && (j + 2) % G == g) {
            const T *Lq = L + (size_t)((j + 2) / G) * n;
            T *dst = rowbuf + (size_t)((j + 2) & 1) * n;
            for (int c = j + 2 + tid; c < n; c += kCoopThreads) dst[c] = Lq[c];
        }
    }
    if (g == 0 && tid == 0) dout[n - 1] = re_(xr[n - 1]);
}

__global__ void coop_clear_kernel(unsigned *sync) {
    for (int i = threadIdx.x; i < 128; i += blockDim.x) sync[i] = 0u;
}

// info := -77 where the reduction's error word is set (and info is 0)
__global__ void coop_err_kernel(const unsigned *sync, int *info) {
    if (threadIdx.x == 0 && sync[64] != 0u && *info == 0) *info = -77;
}

int coop_cus() {
    static int cus[16] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return 0;
    if (cus[dev] == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
        int coop = 0;
        if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop) v = -1;
        cus[dev] = v;
    }
    return cus[dev];
}

// workgroups for order n: the smallest multiple of 8 whose row share and three
// vectors fit the LDS budget; 0 when none does within the device's CUs
template <typename T>
int coop_groups(int n) {
    const int cus = coop_cus();
    if (cus <= 0 || n < 2 || n > kCoopMaxPer * kCoopThreads) return 0;
    for (int G = 8; G <= cus && G <= kCoopMaxGroups; G += 8) {
        const int nr = (n + G - 1) / G;
        if (nr <= kCoopRowsPerWave * kCoopWaves && (size_t)(nr + 3) * n * sizeof(T) <= (size_t)kCoopLdsBytes)
            return G;
    }
    return 0;
}

template <typename T>
int trd_coop_impl(int n, T *A, int lda, double *d, double *e, T *tau, void *work, size_t work_bytes, int *info,
                  hipStream_t st) {
    const int G = coop_groups<T>(n);
    if (G == 0) return QM_ERR_STATE;
    const int nr = (n + G - 1) / G;
    unsigned char *wb = static_cast<unsigned char *>(work);
    unsigned *sync = reinterpret_cast<unsigned *>(wb);
    T *rowbuf = reinterpret_cast<T *>(wb + 512);
    T *pbuf = rowbuf + 2 * (size_t)n;
    T *dots = pbuf + 2 * (size_t)n;
    QM_CHECK(512 + (4 * (size_t)n + 2 * (size_t)G) * sizeof(T) <= work_bytes, QM_ERR_ARG,
             "tridiagonal reduction: work buffer too small");
    hipLaunchKernelGGL(coop_clear_kernel, dim3(1), dim3(128), 0, st, sync);
    QM_HIP(hipGetLastError());
    void *args[] = {&A, &lda, &n, (void *)&nr, &d, &e, &tau, &rowbuf, &pbuf, &dots, &sync};
    const int per = (n + kCoopThreads - 1) / kCoopThreads;  // columns per thread
    const void *fn = per <= 2   ? reinterpret_cast<const void *>(&trd_coop_kernel<T, 2>)
                     : per <= 4 ? reinterpret_cast<const void *>(&trd_coop_kernel<T, 4>)
                                : reinterpret_cast<const void *>(&trd_coop_kernel<T, 8>);
    const hipError_t le = hipLaunchCooperativeKernel(fn, dim3(G), dim3(kCoopThreads), args, 0, st);
    if (le != hipSuccess) {
        (void)hipGetLastError();
        return QM_ERR_STATE;  // the caller takes the single-workgroup route
    }
    if (info) {
        hipLaunchKernelGGL(coop_err_kernel, dim3(1), dim3(64), 0, st, sync, info);
        QM_HIP(hipGetLastError());
    }
    return QM_OK;
}

// *info := the Cholesky's failing minor (sync[96]), or -77 where a barrier
// wait ran past its limit
__global__ void potrf_coop_info_kernel(const unsigned *sync, int *info) {
    if (threadIdx.x == 0) *info = sync[96] != 0u ? (int)sync[96] : (sync[64] != 0u ? -77 : 0);
}

// ---------------------------------------------------------------- Cholesky
// Upper Cholesky factor A = U^T U (LAPACK dpotrf, uplo 'U') by cooperating
// workgroups, one per 32-row panel.  Thread c of a workgroup holds column
// r0 + c of its panel's rows in registers (32 values).  Step k: the owner of
// panel k factors it, one column per workgroup barrier (the row being
// eliminated is shared through LDS, double-buffered; rows stay unscaled until
// the panel is done, so a thread needs no other thread's updated values), and
// writes it to A; one grid barrier; every later panel subtracts the rank-32
// product of panel k's rows from its own (thread c: panel k's column c from
// memory into registers, the 32 x 32 block over the panel's own columns from
// LDS), and the owner of panel k + 1 goes on to factor it.  One grid barrier
// per panel.  *info = 0, or the order of the first leading minor whose pivot
// is not positive (or NaN), as dpotrf reports it; the strictly lower triangle
// is never written.  sync (zeroed by coop_clear_kernel): as the reduction's,
// plus sync[96] = info.
constexpr int kPcNB = 32;
constexpr int kPcMaxN = 512;
constexpr int kPcThreads = kPcMaxN;  // one thread per column

__global__ __launch_bounds__(kPcThreads) void potrf_coop_kernel(double *__restrict__ A, int lda, int n,
                                                                unsigned *sync) {
    __shared__ double rowj[2][kPcMaxN];          // the row being eliminated (by parity of j)
    __shared__ double piv[kPcNB];                // the panel's pivots
    __shared__ double Ux[kPcNB][kPcNB + 1];      // panel k's rows over this panel's columns
    const int g = blockIdx.x, tid = threadIdx.x;
    unsigned *count = sync, *gen = sync + 32;
    int *err = reinterpret_cast<int *>(sync + 64);
    int *info = reinterpret_cast<int *>(sync + 96);
    unsigned lg = 0;
    const int P = gridDim.x;
    const int r0 = g * kPcNB, nrow = min(kPcNB, n - r0), w = n - r0;
    const bool mine = tid < w;  // column c = r0 + tid of the panel's rows
    const int c = r0 + tid;
    rowj[0][tid] = 0.0;  // columns past n stay zero
    rowj[1][tid] = 0.0;
    double a[kPcNB];
#pragma unroll
    for (int i = 0; i < kPcNB; ++i) a[i] = (mine && i < nrow && tid >= i) ? A[(r0 + i) + (size_t)c * lda] : 0.0;
    for (int k = 0; k < P; ++k) {
        if (k == g) {
            int bad = 0;
            // nrow and bad are uniform over the workgroup, so every thread
            // takes the same branches around the barriers
#pragma unroll
            for (int j = 0; j < kPcNB; ++j) {
                if (j < nrow && bad == 0) {
                    if (mine) rowj[j & 1][tid] = a[j];
                    __syncthreads();
                    const double d = rowj[j & 1][j];  // the same pivot for every thread
                    if (!(d > 0.0)) {
                        bad = r0 + j + 1;
                    } else {
                        if (tid == 0) piv[j] = d;
                        const double inv = 1.0 / d;
                        // rows below the diagonal (i > tid) or past the panel
                        // (i >= nrow) take values that are never read or written out
                        if (mine && tid > j) {
                            const double f = a[j] * inv;
#pragma unroll
                            for (int i = j + 1; i < kPcNB; ++i) a[i] = fma(-rowj[j & 1][i], f, a[i]);
                        }
                    }
                }
            }
            __syncthreads();  // piv complete
            if (bad) {
                if (tid == 0) __hip_atomic_store(info, bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else if (mine) {
#pragma unroll
                for (int i = 0; i < kPcNB; ++i) {
                    if (i >= nrow || i > tid) continue;
                    const double u = sqrt(piv[i]);
                    A[(r0 + i) + (size_t)c * lda] = i == tid ? u : a[i] / u;
                }
            }
        }
        coop_barrier(count, gen, P, lg, err);
        if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;  // all workgroups
        if (g <= k) continue;
        // own rows -= U(panel k rows, own columns)^T U(panel k rows, column c)
        const int k0 = k * kPcNB;
        for (int e = tid; e < kPcNB * kPcNB; e += kPcThreads) {
            const int t = e % kPcNB, i = e / kPcNB;
            Ux[t][i] = i < nrow ? A[(k0 + t) + (size_t)(r0 + i) * lda] : 0.0;
        }
        double uk[kPcNB];
#pragma unroll
        for (int t = 0; t < kPcNB; ++t) uk[t] = mine ? A[(k0 + t) + (size_t)c * lda] : 0.0;
        __syncthreads();
        if (mine) {  // (rows below the diagonal and past the panel: as above)
#pragma unroll
            for (int t = 0; t < kPcNB; ++t) {
                const double ut = uk[t];
#pragma unroll
                for (int i = 0; i < kPcNB; ++i) a[i] = fma(-Ux[t][i], ut, a[i]);
            }
        }
        __syncthreads();  // Ux is rewritten at the next step
    }
}

}  // namespace

size_t trd_coop_work_bytes(int n, bool cplx) { return 512 + (4 * (size_t)n + 512) * (cplx ? 16 : 8); }

int trd_coop_supported(int n, bool cplx) { return cplx ? coop_groups<zdbl>(n) : coop_groups<double>(n); }

int trd_coop_f64(int n, double *A, int lda, double *d, double *e, double *tau, void *work, size_t work_bytes,
                 hipStream_t st) {
    return trd_coop_impl<double>(n, A, lda, d, e, tau, work, work_bytes, nullptr, st);
}

int trd_coop_z64(int n, void *A, int lda, double *d, double *e, void *tau, void *work, size_t work_bytes,
                 hipStream_t st) {
    return trd_coop_impl<zdbl>(n, static_cast<zdbl *>(A), lda, d, e, static_cast<zdbl *>(tau), work, work_bytes,
                               nullptr, st);
}

int trd_coop_check(void *work, int *info, hipStream_t st) {
    hipLaunchKernelGGL(coop_err_kernel, dim3(1), dim3(64), 0, st, static_cast<const unsigned *>(work), info);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

}  // namespace qm

namespace qm {

int potrf_coop_supported(int n) {
    const int cus = coop_cus();
    const int P = (n + kPcNB - 1) / kPcNB;
    return n >= 1 && n <= kPcMaxN && cus > 0 && P <= cus;
}

int potrf_coop_f64(int n, double *A, int lda, int *info, void *sync, hipStream_t st) {
    if (!potrf_coop_supported(n)) return QM_ERR_STATE;
    const int P = (n + kPcNB - 1) / kPcNB;
    unsigned *sy = static_cast<unsigned *>(sync);
    hipLaunchKernelGGL(coop_clear_kernel, dim3(1), dim3(128), 0, st, sy);
    QM_HIP(hipGetLastError());
    void *args[] = {&A, &lda, &n, &sy};
    const hipError_t le = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&potrf_coop_kernel), dim3(P),
                                                     dim3(kPcThreads), args, 0, st);
    if (le != hipSuccess) {
        (void)hipGetLastError();
        return QM_ERR_STATE;  // the caller takes the single-workgroup kernel
    }
    hipLaunchKernelGGL(potrf_coop_info_kernel, dim3(1), dim3(64), 0, st, sy, info);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

}  // namespace qm
