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
// Output follows LAPACK's convention so rocSOLVER's stedc and the
// back-transformation below (ormtr / unmtr's operation) consume it: d, e
// (real), tau, and v (v_{j+1} = 1 implied) below the
// subdiagonal of column j of A.  The input is read from A's upper triangle
// (uplo 'U' defines the matrix; the diagonal's imaginary part is dropped).
//
// The barrier is an arrival counter in device memory that only grows within
// a launch, with agent-scope release / acquire ordering on the arrival and
// the wait; its waits are bounded: past
// kCoopSpinLimit polls a workgroup sets the error word and goes on, so the
// kernel always ends (the caller turns the error word into a failed solve).
// hipLaunchCooperativeKernel guarantees that all G workgroups are resident.
#include <atomic>

#include "qm_internal.h"

namespace qm {
namespace {

constexpr int kCoopThreads = 256;
constexpr int kCoopWaves = kCoopThreads / 64;
constexpr int kCoopLdsBytes = 150 * 1024;
constexpr long kCoopSpinLimit = 1L << 24;  // ~0.5 s of polling
// the bound the launches use (qm_coop_set_spin_limit), process-wide
std::atomic<long> g_coop_spin_limit{kCoopSpinLimit};
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
// XCD's L2); thread 0's agent-scope release increment writes the L2 back.
// Barrier b (b = 0, 1, ..) is passed when the counter reads (b + 1) nblk:
// each waiter polls the counter itself, so no second store (a generation
// word written by the last arrival) sits between the last arrival and the
// waiters.  The polls are relaxed loads; the fence after the loop takes the
// acquire once (every write to the counter is a release increment, so the
// fence synchronizes with all of them) and invalidates this CU's caches, so
// after the closing workgroup barrier every wave reads the other
// workgroups' data from memory.  An acquire load would invalidate the
// caches at every poll.
__device__ void coop_barrier(unsigned *count, unsigned nblk, unsigned &lg, int *err, long limit) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned target = (lg + 1) * nblk;
        __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        long spins = 0;
        while (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins > limit) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            // once any workgroup has timed out the launch has failed: stop
            // waiting (every 256 polls), so a failed launch ends after about
            // one limit in all, not one per barrier
            if ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            __builtin_amdgcn_s_sleep(1);
        }
        lg = lg + 1;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// sync[0]: arrival count (grows by G per barrier; at most n G < 2^32),
// sync[64]: error word.  All zero at launch.
template <typename T, int PER>
__global__ __launch_bounds__(kCoopThreads) void trd_coop_kernel(T *__restrict__ A, int lda, int n, int nr,
                                                                double *__restrict__ dout,
                                                                double *__restrict__ eout, T *__restrict__ tauout,
                                                                T *__restrict__ rowbuf, T *__restrict__ pbuf,
                                                                T *__restrict__ dots, unsigned *sync, long limit) {
    __shared__ __align__(16) unsigned char lds_raw[kCoopLdsBytes];
    __shared__ T wsum[kCoopWaves];
    __shared__ T dsh[kCoopMaxGroups];
    T *L = reinterpret_cast<T *>(lds_raw);  // own rows: row g + q G at L + q n
    T *xr = L + (size_t)nr * n;             // the current column's row (conjugate of x)
    T *v = xr + n;                          // reflector, global row index
    T *w = v + n;
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    unsigned *count = sync;
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
    coop_barrier(count, G, lg, err, limit);
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
        coop_barrier(count, G, lg, err, limit);
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

// d, e := 0 (n values each) where the reduction's error word is set
__global__ void coop_guard_kernel(const unsigned *sync, double *d, double *e, int n) {
    if (sync[64] == 0u) return;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        d[i] = 0.0;
        e[i] = 0.0;
    }
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
    long limit = g_coop_spin_limit.load();
    void *args[] = {&A, &lda, &n, (void *)&nr, &d, &e, &tau, &rowbuf, &pbuf, &dots, &sync, &limit};
    const int per = (n + kCoopThreads - 1) / kCoopThreads;  // columns per thread
    const void *fn = per <= 2   ? reinterpret_cast<const void *>(&trd_coop_kernel<T, 2>)
                     : per <= 4 ? reinterpret_cast<const void *>(&trd_coop_kernel<T, 4>)
                                : reinterpret_cast<const void *>(&trd_coop_kernel<T, 8>);
    const hipError_t le = hipLaunchCooperativeKernel(fn, dim3(G), dim3(kCoopThreads), args, 0, st);
    if (le != hipSuccess) {
        (void)hipGetLastError();
        return QM_ERR_STATE;  // the caller takes the single-workgroup route
    }
    // a launch whose barrier timed out leaves d and e undefined: they are
    // zeroed (a valid tridiagonal) before any eigensolver reads them, and
    // the caller's check of the error word rejects the solve
    hipLaunchKernelGGL(coop_guard_kernel, dim3(1), dim3(256), 0, st, (const unsigned *)sync, d, e, n);
    QM_HIP(hipGetLastError());
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
                                                                unsigned *sync, long limit) {
    __shared__ double rowj[2][kPcMaxN];          // the row being eliminated (by parity of j)
    __shared__ double piv[kPcNB];                // the panel's pivots
    __shared__ double Ux[kPcNB][kPcNB + 1];      // panel k's rows over this panel's columns
    const int g = blockIdx.x, tid = threadIdx.x;
    unsigned *count = sync;
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
        coop_barrier(count, P, lg, err, limit);
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

// ------------------------------------------------- back-transformation (Q Z)
// Z := Q Z for the first nev columns of Z, Q = H(0) H(1) ... H(n-2) from the
// reflectors a LAPACK-convention lower tridiagonal reduction leaves (H(j) =
// I - tau_j v_j v_j^H, v_j = (0 .. 0, 1 at j + 1, A(j+2 .. n-1, j)); dormtr /
// zunmtr with side L, uplo L, no transpose).  The reflectors are applied last
// to first in groups of QB: group g holds r_k = n - 2 - (g QB + k), applied
// in the order k = 0 .. QB-1.  With d_k = v_{r_k}^H z (z on entry to the
// group) and M_km = v_{r_k}^H v_{r_m}, the k-th reflector sees
// s_k = d_k - sum_{m<k} M_km w_m, w_k = tau_{r_k} s_k, and the group leaves
// z - sum_k w_k v_{r_k}.  The QB dot products of a group share one round of
// reductions, which cuts the serial chain of loads and reductions QB-fold;
// M depends on the reflectors only and is formed once for all columns by
// reflector_gram_kernel.  A wave per column left each SIMD a single
// dependent instruction stream (nev waves on the chip): a workgroup per
// column spreads the column over four SIMDs and nev CUs.

template <typename T>
__device__ inline T refl_elem(const T *aj, int i, int j, int n) {
    return i == j + 1 ? make<T>(1.0, 0.0) : (i >= j + 2 && i < n ? aj[i] : make<T>(0.0, 0.0));
}

// M[g][k][m] = v_{r_k}^H v_{r_m}, m < k (r_k < r_m: the product's support
// starts at r_m + 1, where v_{r_m} is 1 and v_{r_k} an element of A); zero
// when r_k < 0 (the last group is short).  One workgroup per group, one wave
// per pair.
template <typename T, int QB>
__global__ __launch_bounds__(256) void reflector_gram_kernel(const T *__restrict__ A, int lda, int n,
                                                             T *__restrict__ M) {
    const int g = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int p = wave; p < QB * (QB - 1) / 2; p += 4) {
        int k = 1;
        while (k * (k + 1) / 2 <= p) ++k;
        const int m = p - k * (k - 1) / 2;
        const int rk = n - 2 - (g * QB + k), rm = n - 2 - (g * QB + m);
        T sum = make<T>(0.0, 0.0);
        if (rk >= 0) {
            const T *ak = A + (size_t)rk * lda, *am = A + (size_t)rm * lda;
            for (int i = rm + 1 + lane; i < n; i += 64) sum = add(sum, mul(conj_(ak[i]), refl_elem(am, i, rm, n)));
        }
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) sum = add(sum, shfl_xor(sum, s));
        if (lane == 0) M[((size_t)g * QB + k) * QB + m] = sum;
    }
}

// Raw element for row i of reflector r: an unconditional load from a row
// clamped into r + 1 .. n - 1 of column max(r, 0), always inside the matrix.
// The value is masked (refl_mask) only where it is used, a group later, so
// that the load stays in flight meanwhile: a select next to the load would
// make the compiler wait for it there.
template <typename T>
__device__ inline T refl_raw(const T *__restrict__ A, int lda, int i, int r, int n) {
    const int rc = r < 0 ? 0 : r;
    const int ic = i < rc + 1 ? rc + 1 : (i > n - 1 ? n - 1 : i);
    return A[(size_t)rc * lda + ic];
}

template <typename T>
__device__ inline T refl_mask(T raw, int i, int r, int n) {
    return r < 0 || i < r + 1 || i > n - 1 ? make<T>(0.0, 0.0) : (i == r + 1 ? make<T>(1.0, 0.0) : raw);
}

template <typename T, int PER, int QB>
__device__ inline void load_group(const T *__restrict__ A, int lda, int n, int g, int tid, T (&v)[QB][PER]) {
#pragma unroll
    for (int k = 0; k < QB; ++k) {
#pragma unroll
        for (int q = 0; q < PER; ++q) v[k][q] = refl_raw(A, lda, tid + 256 * q, n - 2 - (g * QB + k), n);
    }
}

// The group's coefficient for thread tid, raw: an M entry (tid < QB^2), else
// the tau of reflector k = tid - QB^2 (address clamped into tau); coef_mask
// zeroes the tau past the last reflector.
template <typename T, int QB>
__device__ inline T coef_raw(const T *__restrict__ M, const T *__restrict__ tau, int n, int g, int tid) {
    const int r = n - 2 - (g * QB + tid - QB * QB);
    const T *p = tid < QB * QB ? M + (size_t)g * QB * QB + tid : tau + (r < 0 || r > n - 2 ? 0 : r);
    return *p;
}

template <typename T, int QB>
__device__ inline T coef_mask(T raw, int n, int g, int tid) {
    return tid >= QB * QB && n - 2 - (g * QB + tid - QB * QB) < 0 ? make<T>(0.0, 0.0) : raw;
}

// Wave sums of QB values at once by recursive halving: at the exchange with
// lane ^ s (s = 32, 16, ..) each lane keeps the half of its list selected by
// its lane bit s and adds the partner's copy of that half, so the list halves
// with every exchange; the last value is then summed over the lanes that
// still share it: QB - 1 + 6 - log2(QB) exchanges in all (10 for QB = 8)
// where separate reductions take 6 QB.  Lane l ends with the sum of value
// multi_index<QB>(l).
template <int QB>
__device__ inline int multi_index(int lane) {
    int k = 0;
#pragma unroll
    for (int m = QB, s = 32; m > 1; m >>= 1, s >>= 1) k += (lane & s) ? m / 2 : 0;
    return k;
}

// M values left, next exchange distance S (compile-time, so that d stays in
// registers: a run-time index into d would turn into select chains)
template <typename T, int QB, int M, int S>
__device__ inline void multi_halve(T (&d)[QB], int lane) {
    if constexpr (M > 1) {
        const bool up = (lane & S) != 0;
#pragma unroll
        for (int j = 0; j < M / 2; ++j) {
            const T send = up ? d[j] : d[j + M / 2];
            const T keep = up ? d[j + M / 2] : d[j];
            d[j] = add(keep, shfl_xor(send, S));
        }
        multi_halve<T, QB, M / 2, S / 2>(d, lane);
    } else {
#pragma unroll
        for (int s = S; s >= 1; s >>= 1) d[0] = add(d[0], shfl_xor(d[0], s));
    }
}

template <typename T, int QB>
__device__ inline T wave_multi_sum(T (&d)[QB], int lane) {
    multi_halve<T, QB, QB, 32>(d, lane);
    return d[0];
}

// Workgroup barrier that orders LDS only: global loads still in flight (the
// next group's prefetch) are not waited for, as the full __syncthreads fence
// would.
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One workgroup per column of Z, row i of the column in thread i mod 256
// (PER rows a thread).  Per group of QB reflectors: QB partial dot products a
// thread, a wave reduction of each, the four waves' partials through LDS (one
// barrier), the triangular recurrence for w (every thread, uniform), the
// update.  The next group's reflector elements, M and tau are loaded while
// the current group is applied; M and tau reach every thread through LDS.
template <typename T, int PER, int QB>
__global__ __launch_bounds__(256) void apply_q_kernel(const T *__restrict__ A, int lda, const T *__restrict__ tau,
                                                      const T *__restrict__ M, T *__restrict__ Z, int ldz, int n) {
    static_assert(QB * QB + QB <= 256, "group coefficients: one a thread");
    // by group parity, so that one barrier a group suffices
    __shared__ T part[2][QB][4];
    __shared__ T coef[2][QB * QB + QB];  // M (QB x QB), then tau of the group
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    T *z = Z + (size_t)blockIdx.x * ldz;
    T zr[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = tid + 256 * q;
        zr[q] = i < n ? z[i] : make<T>(0.0, 0.0);
    }
    const int groups = (n - 1 + QB - 1) / QB;
    T vn[QB][PER];
    load_group<T, PER, QB>(A, lda, n, 0, tid, vn);
    T cn = coef_raw<T, QB>(M, tau, n, 0, tid);
    for (int g = 0; g < groups; ++g) {
        T vr[QB][PER], d[QB];
#pragma unroll
        for (int k = 0; k < QB; ++k) {
#pragma unroll
            for (int q = 0; q < PER; ++q) vr[k][q] = refl_mask(vn[k][q], tid + 256 * q, n - 2 - (g * QB + k), n);
        }
        const T cc = coef_mask<T, QB>(cn, n, g, tid);
        // next group (the last group reloads itself: no branch around the loads)
        const int gn = g + 1 < groups ? g + 1 : g;
        load_group<T, PER, QB>(A, lda, n, gn, tid, vn);
        cn = coef_raw<T, QB>(M, tau, n, gn, tid);
#pragma unroll
        for (int k = 0; k < QB; ++k) {
            d[k] = make<T>(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < PER; ++q) d[k] = add(d[k], mul(conj_(vr[k][q]), zr[q]));
        }
        const T dk = wave_multi_sum<T, QB>(d, lane);
        const int par = g & 1;
        if ((lane & (64 / QB - 1)) == 0) part[par][multi_index<QB>(lane)][wave] = dk;
        if (tid < QB * QB + QB) coef[par][tid] = cc;
        lds_barrier();
        T w[QB];
#pragma unroll
        for (int k = 0; k < QB; ++k) {
            T t = add(add(part[par][k][0], part[par][k][1]), add(part[par][k][2], part[par][k][3]));
#pragma unroll
            for (int m = 0; m < k; ++m) t = sub(t, mul(coef[par][k * QB + m], w[m]));
            w[k] = mul(coef[par][QB * QB + k], t);
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
#pragma unroll
            for (int k = 0; k < QB; ++k) zr[q] = sub(zr[q], mul(vr[k][q], w[k]));
        }
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = tid + 256 * q;
        if (i < n) z[i] = zr[q];
    }
}

// reflectors per group: the two register copies of a group's reflector
// elements (PER x QB each) stay within 32 values a thread
template <typename T>
constexpr int apply_q_group(int per) {
    return 16 / per < 8 ? 16 / per : 8;
}

size_t trd_core_bytes(int n, size_t esz) { return 512 + (4 * (size_t)n + 512) * esz; }

size_t apply_q_m_offset(int n, size_t esz) { return (trd_core_bytes(n, esz) + 255) / 256 * 256; }

template <typename T, int PER>
int apply_q_launch(int n, int nev, const T *A, int lda, const T *tau, T *Z, int ldz, T *M, hipStream_t st) {
    constexpr int QB = apply_q_group<T>(PER);
    const int groups = (n - 1 + QB - 1) / QB;
    if (QB > 1) {
        hipLaunchKernelGGL((reflector_gram_kernel<T, QB>), dim3((unsigned)groups), dim3(256), 0, st, A, lda, n, M);
        QM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL((apply_q_kernel<T, PER, QB>), dim3((unsigned)nev), dim3(256), 0, st, A, lda, tau,
                       static_cast<const T *>(M), Z, ldz, n);
    QM_HIP(hipGetLastError());
    return QM_OK;
}

template <typename T>
int apply_q_impl(int n, int nev, const T *A, int lda, const T *tau, T *Z, int ldz, void *work, hipStream_t st) {
    if (n < 2 || nev < 1) return QM_OK;  // Q = I
    QM_CHECK(n <= 64 * 32, QM_ERR_ARG, "back-transformation: order above 2048");
    T *M = reinterpret_cast<T *>(static_cast<char *>(work) + apply_q_m_offset(n, sizeof(T)));
    const int per = (n + 255) / 256;
    if (per <= 1) return apply_q_launch<T, 1>(n, nev, A, lda, tau, Z, ldz, M, st);
    if (per <= 2) return apply_q_launch<T, 2>(n, nev, A, lda, tau, Z, ldz, M, st);
    if (per <= 4) return apply_q_launch<T, 4>(n, nev, A, lda, tau, Z, ldz, M, st);
    return apply_q_launch<T, 8>(n, nev, A, lda, tau, Z, ldz, M, st);
}

}  // namespace

size_t trd_coop_work_bytes(int n, bool cplx) {
    // the reduction's own words, then the reflector group products of the
    // back-transformation (at most 8 reflectors a group)
    const size_t esz = cplx ? 16 : 8;
    return apply_q_m_offset(n, esz) + ((size_t)n / 8 + 1) * 64 * esz;
}

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

int ormtr_lower_f64(int n, int nev, const double *A, int lda, const double *tau, double *Z, int ldz, void *work,
                    hipStream_t st) {
    return apply_q_impl<double>(n, nev, A, lda, tau, Z, ldz, work, st);
}

int unmtr_lower_z64(int n, int nev, const void *A, int lda, const void *tau, void *Z, int ldz, void *work,
                    hipStream_t st) {
    return apply_q_impl<zdbl>(n, nev, static_cast<const zdbl *>(A), lda, static_cast<const zdbl *>(tau),
                              static_cast<zdbl *>(Z), ldz, work, st);
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
    long limit = g_coop_spin_limit.load();
    void *args[] = {&A, &lda, &n, &sy, &limit};
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

extern "C" int qm_coop_set_spin_limit(long polls) {
    QM_CHECK(polls >= 0, QM_ERR_ARG, "qm_coop_set_spin_limit: polls must be >= 0 (a negative value resets)");
    qm::g_coop_spin_limit.store(polls);
    return QM_OK;
}

extern "C" long qm_coop_spin_limit_default(void) { return qm::kCoopSpinLimit; }
