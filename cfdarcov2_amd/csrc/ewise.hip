// Elementwise kernels and the two reductions the time loop needs (CFL minimum,
// non-finite count).  Reductions are two-stage (per-block partials, then one
// block) so results are deterministic and no atomics are involved.
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>

#include "common.hpp"
#include "kernels.hpp"

namespace cfd {

// exp from the device libm in the operand's precision; abs as the sign bit
// cleared (exact, -0 -> +0, NaN stays NaN)
template <class T> __device__ inline T ew_exp(T a) {
    if constexpr (sizeof(T) == 4)
        return expf(a);
    else
        return ::exp(a);
}
template <class T> __device__ inline T ew_abs(T a) {
    if constexpr (sizeof(T) == 4)
        return fabsf(a);
    else
        return ::fabs(a);
}

template <class T>
__global__ void __launch_bounds__(256) ewise_k(int op, const T* __restrict__ a, const T* __restrict__ b,
                                               const T* __restrict__ aux, T s, T* __restrict__ out, int64_t n) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        T r;
        switch (op) {
            case CFD_EW_ADD: r = a[i] + b[i]; break;
            case CFD_EW_SUB: r = a[i] - b[i]; break;
            case CFD_EW_MUL: r = a[i] * b[i]; break;
            case CFD_EW_DIV: r = a[i] / b[i]; break;
            case CFD_EW_NEG: r = -a[i]; break;
            case CFD_EW_MULS: r = a[i] * s; break;
            case CFD_EW_DIVS: r = a[i] / s; break;
            case CFD_EW_SMUL: r = s * a[i]; break;
            case CFD_EW_AXPY: r = s * a[i] + b[i]; break;
            case CFD_EW_LEAP: r = ((s * s) * a[i] + T(2) * b[i]) - aux[i]; break;
            case CFD_EW_FILL: r = s; break;
            case CFD_EW_ADDS: r = a[i] + s; break;
            case CFD_EW_SADD: r = s + a[i]; break;
            case CFD_EW_SUBS: r = a[i] - s; break;
            case CFD_EW_SSUB: r = s - a[i]; break;
            case CFD_EW_SDIV: r = s / a[i]; break;
            case CFD_EW_EXP: r = ew_exp(a[i]); break;
            case CFD_EW_ABS: r = ew_abs(a[i]); break;
            default: r = a[i]; break;  // CFD_EW_COPY
        }
        out[i] = r;
    }
}

// One element of an elementwise op (the operation orders of the reference's
// kernels, src/cuda_operators.cu:27-136, and the DtVar / D2tVar updates).
template <int OP, class T> __device__ inline T ew_one(T a, T b, T aux, T s) {
    switch (OP) {
        case CFD_EW_ADD: return a + b;
        case CFD_EW_SUB: return a - b;
        case CFD_EW_MUL: return a * b;
        case CFD_EW_DIV: return a / b;
        case CFD_EW_NEG: return -a;
        case CFD_EW_MULS: return a * s;
        case CFD_EW_DIVS: return a / s;
        case CFD_EW_SMUL: return s * a;
        case CFD_EW_AXPY: return s * a + b;
        case CFD_EW_LEAP: return ((s * s) * a + T(2) * b) - aux;
        case CFD_EW_FILL: return s;
        case CFD_EW_ADDS: return a + s;
        case CFD_EW_SADD: return s + a;
        case CFD_EW_SUBS: return a - s;
        case CFD_EW_SSUB: return s - a;
        case CFD_EW_SDIV: return s / a;
        case CFD_EW_EXP: return ew_exp(a);
        case CFD_EW_ABS: return ew_abs(a);
        default: return a;  // CFD_EW_COPY
    }
}

constexpr bool ew_reads_a(int op) { return op != CFD_EW_FILL; }
constexpr bool ew_reads_b(int op) {
    return op == CFD_EW_ADD || op == CFD_EW_SUB || op == CFD_EW_MUL || op == CFD_EW_DIV || op == CFD_EW_AXPY ||
           op == CFD_EW_LEAP;
}

// Vector form: 16 bytes per access (4 floats or 2 doubles), one vector per
// thread, the op fixed at compile time; the tail of n that does not fill a
// vector is handled by the last vector's threads one element at a time.
// Fields from the pool are 256-byte aligned; callers with other alignments
// take ewise_k.
template <int OP, class T>
__global__ void __launch_bounds__(256) ewise_vec_k(const T* __restrict__ a, const T* __restrict__ b,
                                                   const T* __restrict__ aux, T s, T* __restrict__ out, int64_t n) {
    constexpr int W = 16 / sizeof(T);
    using V = T __attribute__((ext_vector_type(W)));
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nv = n / W;
    if (i < nv) {
        V va = {}, vb = {}, vx = {}, r;
        if constexpr (ew_reads_a(OP)) va = reinterpret_cast<const V*>(a)[i];
        if constexpr (ew_reads_b(OP)) vb = reinterpret_cast<const V*>(b)[i];
        if constexpr (OP == CFD_EW_LEAP) vx = reinterpret_cast<const V*>(aux)[i];
#pragma unroll
        for (int k = 0; k < W; ++k) r[k] = ew_one<OP, T>(va[k], vb[k], vx[k], s);
        reinterpret_cast<V*>(out)[i] = r;
    } else if (i == nv) {
        for (int64_t j = nv * W; j < n; ++j) {
            const T ea = ew_reads_a(OP) ? a[j] : T(0);
            const T eb = ew_reads_b(OP) ? b[j] : T(0);
            const T ex = OP == CFD_EW_LEAP ? aux[j] : T(0);
            out[j] = ew_one<OP, T>(ea, eb, ex, s);
        }
    }
}

template <class T, int OP>
static void launch_vec(const void* a, const void* b, const void* aux, double s, void* out, int64_t n, hipStream_t st) {
    constexpr int W = 16 / sizeof(T);
    const int64_t threads = n / W + 1;  // one more for the tail
    const int64_t blocks = (threads + 255) / 256;
    CFD_REQUIRE(blocks < (int64_t(1) << 31), "elementwise launch too large");
    ewise_vec_k<OP, T><<<dim3(unsigned(blocks)), 256, 0, st>>>((const T*)a, (const T*)b, (const T*)aux, T(s),
                                                                 (T*)out, n);
}

template <class T>
static void dispatch_vec(int op, const void* a, const void* b, const void* aux, double s, void* out, int64_t n,
                         hipStream_t st) {
    switch (op) {
        case CFD_EW_ADD: launch_vec<T, CFD_EW_ADD>(a, b, aux, s, out, n, st); break;
        case CFD_EW_SUB: launch_vec<T, CFD_EW_SUB>(a, b, aux, s, out, n, st); break;
        case CFD_EW_MUL: launch_vec<T, CFD_EW_MUL>(a, b, aux, s, out, n, st); break;
        case CFD_EW_DIV: launch_vec<T, CFD_EW_DIV>(a, b, aux, s, out, n, st); break;
        case CFD_EW_NEG: launch_vec<T, CFD_EW_NEG>(a, b, aux, s, out, n, st); break;
        case CFD_EW_MULS: launch_vec<T, CFD_EW_MULS>(a, b, aux, s, out, n, st); break;
        case CFD_EW_DIVS: launch_vec<T, CFD_EW_DIVS>(a, b, aux, s, out, n, st); break;
        case CFD_EW_SMUL: launch_vec<T, CFD_EW_SMUL>(a, b, aux, s, out, n, st); break;
        case CFD_EW_AXPY: launch_vec<T, CFD_EW_AXPY>(a, b, aux, s, out, n, st); break;
        case CFD_EW_LEAP: launch_vec<T, CFD_EW_LEAP>(a, b, aux, s, out, n, st); break;
        case CFD_EW_FILL: launch_vec<T, CFD_EW_FILL>(a, b, aux, s, out, n, st); break;
        case CFD_EW_ADDS: launch_vec<T, CFD_EW_ADDS>(a, b, aux, s, out, n, st); break;
        case CFD_EW_SADD: launch_vec<T, CFD_EW_SADD>(a, b, aux, s, out, n, st); break;
        case CFD_EW_SUBS: launch_vec<T, CFD_EW_SUBS>(a, b, aux, s, out, n, st); break;
        case CFD_EW_SSUB: launch_vec<T, CFD_EW_SSUB>(a, b, aux, s, out, n, st); break;
        case CFD_EW_SDIV: launch_vec<T, CFD_EW_SDIV>(a, b, aux, s, out, n, st); break;
        case CFD_EW_EXP: launch_vec<T, CFD_EW_EXP>(a, b, aux, s, out, n, st); break;
        case CFD_EW_ABS: launch_vec<T, CFD_EW_ABS>(a, b, aux, s, out, n, st); break;
        default: launch_vec<T, CFD_EW_COPY>(a, b, aux, s, out, n, st); break;
    }
}

static int blocks_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return int(b < 1 ? 1 : b);
}

template <class T>
void launch_ewise(int op, const void* a, const void* b, const void* aux, double s, void* out, int64_t n,
                  hipStream_t st) {
    if (n <= 0) return;
    auto aligned = [](const void* p) { return p == nullptr || (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    if (aligned(a) && aligned(b) && aligned(aux) && aligned(out))
        dispatch_vec<T>(op, a, b, aux, s, out, n, st);
    else
        ewise_k<T><<<blocks_for(n), 256, 0, st>>>(op, (const T*)a, (const T*)b, (const T*)aux, T(s), (T*)out, n);
    CFD_HIP(hipGetLastError());
}

// wave64 min via cross-lane shuffles, then across the block's waves through LDS
template <class T> __device__ inline T block_min(T v) {
    __shared__ T part[16];
    for (int off = 32; off > 0; off >>= 1) {
        const T o = __shfl_down(v, off, 64);
        v = (o < v) ? o : v;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) part[wid] = v;
    __syncthreads();
    if (wid == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        v = lane < nw ? part[lane] : std::numeric_limits<T>::infinity();
        for (int off = 32; off > 0; off >>= 1) {
            const T o = __shfl_down(v, off, 64);
            v = (o < v) ? o : v;
        }
    }
    return v;
}

// per-cell dl * (1 / (sqrt((gamma*p)/rho) + sqrt((u*u + v*v) + w*w)))
// (UpdatePolicies::CourantFriedrichsLewy3D/1D, src/fvm3d.cpp:512-555)
template <class T>
__global__ void __launch_bounds__(256) cfl_partial_k(const T* __restrict__ u, const T* __restrict__ v,
                                                     const T* __restrict__ w, const T* __restrict__ p,
                                                     const T* __restrict__ rho, T dl, T gamma, int64_t n,
                                                     T* __restrict__ partials) {
    T mn = std::numeric_limits<T>::infinity();
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const T a = sqrt((gamma * p[i]) / rho[i]);
        T sq = u[i] * u[i];
        if (v) sq = sq + v[i] * v[i];
        if (w) sq = sq + w[i] * w[i];
        const T val = dl * (T(1) / (a + sqrt(sq)));
        mn = (val < mn) ? val : mn;
    }
    mn = block_min(mn);
    if (threadIdx.x == 0) partials[blockIdx.x] = mn;
}

template <class T>
__global__ void __launch_bounds__(256) min_final_k(const T* __restrict__ partials, int nb, T* __restrict__ out) {
    T mn = std::numeric_limits<T>::infinity();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) mn = (partials[i] < mn) ? partials[i] : mn;
    mn = block_min(mn);
    if (threadIdx.x == 0) out[0] = mn;
}

template <class T>
void launch_cfl(const void* u, const void* v, const void* w, const void* p, const void* rho, T dl, T gamma,
                int64_t n, void* partials, int nblocks, void* dev_out, hipStream_t s) {
    cfl_partial_k<T><<<nblocks, 256, 0, s>>>((const T*)u, (const T*)v, (const T*)w, (const T*)p, (const T*)rho, dl,
                                             gamma, n, (T*)partials);
    CFD_HIP(hipGetLastError());
    min_final_k<T><<<1, 256, 0, s>>>((const T*)partials, nblocks, (T*)dev_out);
    CFD_HIP(hipGetLastError());
}

template <class T>
__global__ void __launch_bounds__(256) nonfinite_partial_k(const T* __restrict__ x, int64_t n,
                                                           unsigned long long* __restrict__ partials) {
    __shared__ unsigned long long part[256];
    unsigned long long cnt = 0;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        cnt += isfinite(x[i]) ? 0ull : 1ull;
    part[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (int(threadIdx.x) < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = part[0];
}

__global__ void __launch_bounds__(256) sum_final_k(unsigned long long* partials, int nb) {
    __shared__ unsigned long long part[256];
    unsigned long long cnt = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) cnt += partials[i];
    part[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (int(threadIdx.x) < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[nb] = part[0];
}

// dev_count must hold nblocks + 1 entries; the total lands in the last one.
template <class T> void launch_nonfinite(const void* x, int64_t n, unsigned long long* dev_count, hipStream_t s) {
    const int nb = 1024;
    nonfinite_partial_k<T><<<nb, 256, 0, s>>>((const T*)x, n, dev_count);
    CFD_HIP(hipGetLastError());
    sum_final_k<<<1, 256, 0, s>>>(dev_count, nb);
    CFD_HIP(hipGetLastError());
}

template void launch_ewise<float>(int, const void*, const void*, const void*, double, void*, int64_t, hipStream_t);
template void launch_ewise<double>(int, const void*, const void*, const void*, double, void*, int64_t, hipStream_t);
template void launch_cfl<float>(const void*, const void*, const void*, const void*, const void*, float, float,
                                int64_t, void*, int, void*, hipStream_t);
template void launch_cfl<double>(const void*, const void*, const void*, const void*, const void*, double, double,
                                 int64_t, void*, int, void*, hipStream_t);
template void launch_nonfinite<float>(const void*, int64_t, unsigned long long*, hipStream_t);
template void launch_nonfinite<double>(const void*, int64_t, unsigned long long*, hipStream_t);

}  // namespace cfd
