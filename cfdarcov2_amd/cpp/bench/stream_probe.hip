// Streaming-bandwidth probe for the benchmark's ceiling figure: the access
// pattern of one acoustic step without its arithmetic.  Four fields are read
// and four written with 16-byte vector accesses, and consecutive passes swap
// the two sets (a pass reads what the previous one wrote), with every other
// pass walking its blocks in reverse.  policy bit 0: non-temporal stores,
// bit 1: non-temporal loads.
#include <hip/hip_runtime.h>

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

template <bool NTL, bool NTS>
__global__ void __launch_bounds__(256) pass_k(const f4* __restrict__ a, const f4* __restrict__ b,
                                              const f4* __restrict__ c, const f4* __restrict__ d, f4* __restrict__ e,
                                              f4* __restrict__ f, f4* __restrict__ g, f4* __restrict__ h, long n,
                                              int rev) {
    const long blk = rev ? long(gridDim.x) - 1 - blockIdx.x : long(blockIdx.x);
    const long stride = long(gridDim.x) * 256;
    for (long i = blk * 256 + threadIdx.x; i < n; i += stride) {
        f4 x, y, z, w;
        if (NTL) {
            x = __builtin_nontemporal_load(&a[i]);
            y = __builtin_nontemporal_load(&b[i]);
            z = __builtin_nontemporal_load(&c[i]);
            w = __builtin_nontemporal_load(&d[i]);
        } else {
            x = a[i];
            y = b[i];
            z = c[i];
            w = d[i];
        }
        if (NTS) {
            __builtin_nontemporal_store(x + y, &e[i]);
            __builtin_nontemporal_store(y + z, &f[i]);
            __builtin_nontemporal_store(z + w, &g[i]);
            __builtin_nontemporal_store(w + x, &h[i]);
        } else {
            e[i] = x + y;
            f[i] = y + z;
            g[i] = z + w;
            h[i] = w + x;
        }
    }
}

template <bool NTL, bool NTS>
void launch(f4* const* in, f4* const* out, long n4, int grid, int rev) {
    pass_k<NTL, NTS><<<grid, 256>>>(in[0], in[1], in[2], in[3], out[0], out[1], out[2], out[3], n4, rev);
}

}  // namespace

// Average microseconds per pass over `reps` passes (after 6 warm-up passes)
// of fields holding `nfloat` floats each (a multiple of 4).  Returns 0, or -1
// on a HIP error.
extern "C" int stream_probe_pingpong(long nfloat, int reps, int grid, int policy, double* us_per_pass) {
    if (nfloat <= 0 || nfloat % 4 != 0 || reps < 1 || grid < 1 || !us_per_pass) return -1;
    const long n4 = nfloat / 4;
    f4* buf[8] = {};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    for (int i = 0; i < 8 && rc == 0; ++i)
        if (hipMalloc(&buf[i], size_t(nfloat) * 4) != hipSuccess || hipMemset(buf[i], 0, size_t(nfloat) * 4) != hipSuccess)
            rc = -1;
    if (rc == 0 && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = -1;
    if (rc == 0) {
        for (int r = 0; r < reps + 6; ++r) {
            if (r == 6) hipEventRecord(e0);
            f4* const* in = (r & 1) ? buf + 4 : buf;
            f4* const* out = (r & 1) ? buf : buf + 4;
            const int rev = r & 1;
            switch (policy & 3) {
                case 0: launch<false, false>(in, out, n4, grid, rev); break;
                case 1: launch<false, true>(in, out, n4, grid, rev); break;
                case 2: launch<true, false>(in, out, n4, grid, rev); break;
                default: launch<true, true>(in, out, n4, grid, rev); break;
            }
        }
        hipEventRecord(e1);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess ||
            hipGetLastError() != hipSuccess)
            rc = -1;
        else
            *us_per_pass = double(ms) * 1e3 / reps;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    for (int i = 0; i < 8; ++i)
        if (buf[i]) hipFree(buf[i]);
    return rc;
}
