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
constexpr int kCoopPer = 8;                 // vector elements per thread: n <= 2048
constexpr int kCoopMaxGroups = 256;

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

// total of one double per thread, the same bits in every thread (fixed order)
__device__ inline double block_sum(double s, double *red) {
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < kCoopWaves; ++i) t += red[i];
    return t;
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
template <typename T>
__global__ __launch_bounds__(kCoopThreads) void trd_coop_kernel(T *__restrict__ A, int lda, int n, int nr,
                                                                double *__restrict__ dout,
                                                                double *__restrict__ eout, T *__restrict__ tauout,
                                                                T *__restrict__ rowbuf, T *__restrict__ pbuf,
                                                                T *__restrict__ dots, unsigned *sync) {
    __shared__ __align__(16) unsigned char lds_raw[kCoopLdsBytes];
    __shared__ double red[kCoopWaves];
    __shared__ T wsum[kCoopWaves];
    __shared__ T dsh[kCoopMaxGroups];
    __shared__ T bcast;
    T *L = reinterpret_cast<T *>(lds_raw);  // own rows: row g + q G at L + q n
    T *xr = L + (size_t)nr * n;             // the current column's row (conjugate of x)
    T *v = xr + n;                          // reflector, global row index
    T *w = v + n;
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    unsigned *count = sync, *gen = sync + 32;
    int *err = reinterpret_cast<int *>(sync + 64);
    unsigned lg = 0;

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
        // ---- reflector from x = conj(xr[j+1 .. n-1]) (LAPACK dlarfg / zlarfg)
        double s = 0.0;
        for (int c = j + 2 + tid; c < n; c += kCoopThreads) s += abs2(xr[c]);
        s = block_sum(s, red);
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
        for (int c = j + 1 + tid; c < n; c += kCoopThreads)
            v[c] = c == j + 1 ? make<T>(1.0, 0.0) : mul(conj_(xr[c]), scal);
        __syncthreads();
        if (g == 0) {
            if (tid == 0) {
                dout[j] = re_(xr[j]);
                eout[j] = beta;
                tauout[j] = tau;
            }
            for (int c = j + 2 + tid; c < n; c += kCoopThreads) A[c + (size_t)j * lda] = v[c];
        }
        // ---- p = tau A22 v on own rows, partial p^H v
        T dot = make<T>(0.0, 0.0);
        for (int q = wave; q < nr; q += kCoopWaves) {
            const int r = g + q * G;
            if (r >= n) break;
            if (r <= j) continue;
            const T *Lq = L + (size_t)q * n;
            T acc = make<T>(0.0, 0.0);
            for (int c = j + 1 + lane; c < n; c += 64) acc = add(acc, mul(Lq[c], v[c]));
            for (int m = 32; m >= 1; m >>= 1) acc = add(acc, shfl_xor(acc, m));
            const T p = mul(tau, acc);
            if (lane == 0) {
                pbuf[(size_t)par * n + r] = p;
                dot = add(dot, mul(conj_(p), v[r]));
            }
        }
        if (lane == 0) wsum[wave] = dot;
        __syncthreads();
        if (tid == 0) {
            T t = make<T>(0.0, 0.0);
            for (int i = 0; i < kCoopWaves; ++i) t = add(t, wsum[i]);
            dots[(size_t)par * G + g] = t;
        }
        coop_barrier(count, gen, G, lg, err);
        // ---- one round of loads: the partial dots, p and the old row j+1
        const T *old = rowbuf + (size_t)((j + 1) & 1) * n;
        T pr[kCoopPer], orow[kCoopPer];
#pragma unroll
        for (int k = 0; k < kCoopPer; ++k) {
            const int c = j + 1 + tid + k * kCoopThreads;
            if (c < n) {
                pr[k] = pbuf[(size_t)par * n + c];
                orow[k] = old[c];
            }
        }
        if (tid < G) dsh[tid] = dots[(size_t)par * G + tid];
        __syncthreads();
        // ---- w = p - (tau / 2) (p^H v) v
        if (tid == 0) {
            T t = make<T>(0.0, 0.0);
            for (int i = 0; i < G; ++i) t = add(t, dsh[i]);
            bcast = mul(make<T>(-0.5, 0.0), mul(tau, t));
        }
        __syncthreads();
        const T a2 = bcast;
#pragma unroll
        for (int k = 0; k < kCoopPer; ++k) {
            const int c = j + 1 + tid + k * kCoopThreads;
            if (c < n) w[c] = add(pr[k], mul(a2, v[c]));
        }
        __syncthreads();
        // ---- look-ahead: row j+1 after this update, from its owner's copy
        const T vj = v[j + 1], wj = w[j + 1];
#pragma unroll
        for (int k = 0; k < kCoopPer; ++k) {
            const int c = j + 1 + tid + k * kCoopThreads;
            if (c < n) xr[c] = rank2(orow[k], vj, w[c], wj, v[c]);
        }
        // ---- A22 -= v w^H + w v^H on own rows
        for (int q = 0; q < nr; ++q) {
            const int r = g + q * G;
            if (r >= n) break;
            if (r <= j) continue;
            T *Lq = L + (size_t)q * n;
            const T vr = v[r], wr = w[r];
            for (int c = j + 1 + tid; c < n; c += kCoopThreads) Lq[c] = rank2(Lq[c], vr, w[c], wr, v[c]);
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
    if (cus <= 0 || n < 2 || n > kCoopPer * kCoopThreads) return 0;
    for (int G = 8; G <= cus && G <= kCoopMaxGroups; G += 8) {
        const int nr = (n + G - 1) / G;
        if ((size_t)(nr + 3) * n * sizeof(T) <= (size_t)kCoopLdsBytes) return G;
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
    const hipError_t le = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&trd_coop_kernel<T>), dim3(G),
                                                     dim3(kCoopThreads), args, 0, st);
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
