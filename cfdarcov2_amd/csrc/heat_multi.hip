// Several heat steps per launch on an exactly uniform mesh whose state fits
// the caches (no reference counterpart; the form for small meshes, where a
// single step lasts about as long as a launch).
//
// A workgroup owns an output box of B x B x BZ cells (BZ = min(Z, 64 - 2K)).
// It loads its box widened by K cells on every side that is not a wall into
// an LDS region (at most (B + 2K) x (B + 2K) x 64 cells), then takes `ns` <= K
// steps there: step s computes the region narrowed by s on its non-wall
// sides, which is all that step s + 1 reads, ping-ponging between two LDS
// buffers; at the walls a cell's outer neighbour is the cell itself, as in
// the single-step kernel.  Finally it stores its output box.  Cells near a
// box's sides are computed by several workgroups (the redundant work that
// buys K - 1 fewer launches).  Each cell's arithmetic is heat_uniform_cell's
// (stencil.hip), so ns steps here equal ns single steps bit for bit.
//
// Threads: 64 lanes along z (one region row per wave pass) x 16 waves over the
// region's (x, y) rows (enough waves per CU to hide the LDS latency: the boxes
// are few, about one workgroup per CU at 64^3).  A launch lasts a few memory
// latencies (the region's loads, issued together, and the stores) plus ns
// passes over LDS.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace cfd {

constexpr int kMultiWaves = 16;

template <class T, int B, int K, bool PW>
__global__ void __launch_bounds__(64 * kMultiWaves) heat_multi_k(UniformGeo<T> g, const T* __restrict__ Tin, T* __restrict__ Tout,
                                                    T alpha, T dt, int ns, int bz, int nzt, int nyt) {
    constexpr int R = B + 2 * K;   // region extent along x and y
    __shared__ T buf[2][R * R * 64];
    const int64_t X = g.X, Y = g.Y, Z = g.Z;
    const int lin = int(blockIdx.x);
    const int zt = lin % nzt, yt = (lin / nzt) % nyt, xt = lin / (nzt * nyt);
    const int lane = int(threadIdx.x), wy = int(threadIdx.y);
    // output box and region (clipped to the mesh)
    const int ox0 = xt * B, ox1 = int(min<int64_t>(ox0 + B, X));
    const int oy0 = yt * B, oy1 = int(min<int64_t>(oy0 + B, Y));
    const int oz0 = zt * bz, oz1 = int(min<int64_t>(oz0 + bz, Z));
    const int gx0 = max(ox0 - K, 0), gx1 = int(min<int64_t>(ox1 + K, X));
    const int gy0 = max(oy0 - K, 0), gy1 = int(min<int64_t>(oy1 + K, Y));
    const int gz0 = max(oz0 - K, 0), gz1 = int(min<int64_t>(oz1 + K, Z));
    const int nry = gy1 - gy0;
    auto slot = [&](int x, int y, int z) { return ((x - gx0) * R + (y - gy0)) * 64 + (z - gz0); };

    // every load of the region is issued before the first LDS store, so the
    // wave waits for the memory latency once rather than once per row
    constexpr int NR = (R * R + kMultiWaves - 1) / kMultiWaves;
    const int nrows = (gx1 - gx0) * nry;
    T v[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int r = wy + i * kMultiWaves;
        const int x = gx0 + r / nry, y = gy0 + r % nry, z = gz0 + lane;
        v[i] = (r < nrows && z < gz1) ? Tin[(int64_t(x) * Y + y) * Z + z] : T(0);
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int r = wy + i * kMultiWaves;
        const int x = gx0 + r / nry, y = gy0 + r % nry, z = gz0 + lane;
        if (r < nrows && z < gz1) buf[0][slot(x, y, z)] = v[i];
    }
    __syncthreads();

    const int64_t ext[3] = {X, Y, Z};
    int cur = 0;
    for (int s = 1; s <= ns; ++s) {
        // the region narrowed by s on its non-wall sides
        const int bx0 = gx0 == 0 ? 0 : gx0 + s, bx1 = gx1 == X ? gx1 : gx1 - s;
        const int by0 = gy0 == 0 ? 0 : gy0 + s, by1 = gy1 == Y ? gy1 : gy1 - s;
        const int bz0 = gz0 == 0 ? 0 : gz0 + s, bz1 = gz1 == Z ? gz1 : gz1 - s;
        const int nby = by1 - by0;
        const T* P = buf[cur];
        T* Q = buf[cur ^ 1];
        for (int r = wy; r < (bx1 - bx0) * nby; r += kMultiWaves) {
            const int x = bx0 + r / nby, y = by0 + r % nby, z = bz0 + lane;
            if (z >= bz1) continue;
            const int c = slot(x, y, z);
            const T pc = P[c];
            const int crd[3] = {x, y, z};
            const int str[3] = {R * 64, 64, 1};
            T L = T(0);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const bool lo = crd[a] == 0, hi = crd[a] == ext[a] - 1;
                const T al_lo = lo ? g.alpha_b[a] : g.alpha[a];
                const T al_hi = hi ? g.alpha_b[a] : g.alpha[a];
                const T s_lo = (-pc + (lo ? pc : P[c - str[a]])) * al_lo;
                const T s_hi = (-pc + (hi ? pc : P[c + str[a]])) * al_hi;
                const T num = s_lo * g.area[a] + s_hi * g.area[a];
                // a power-of-two volume divides exactly: the product is the same number
                if constexpr (PW) L = L + num * g.inv_vol;
                else L = L + num / g.vol;
            }
            Q[c] = dt * (alpha * L) + pc;
        }
        __syncthreads();
        cur ^= 1;
    }

    const int noy = oy1 - oy0;
    for (int r = wy; r < (ox1 - ox0) * noy; r += kMultiWaves) {
        const int x = ox0 + r / noy, y = oy0 + r % noy, z = oz0 + lane;
        if (z < oz1) Tout[(int64_t(x) * Y + y) * Z + z] = buf[cur][slot(x, y, z)];
    }
}

bool heat_multi_supported(int B, int K) {
    return (B == 4 || B == 8) && (K == 2 || K == 4);
}

template <class T, int B, int K>
static void launch_multi(const UniformGeo<T>& g, const void* in, void* out, double alpha, double dt, int ns,
                         hipStream_t s) {
    const int bz = int(g.Z <= 64 ? g.Z : 64 - 2 * K);
    const int nzt = int((g.Z + bz - 1) / bz), nyt = int((g.Y + B - 1) / B), nxt = int((g.X + B - 1) / B);
    const int64_t nwg = int64_t(nzt) * nyt * nxt;
    CFD_REQUIRE(nwg < (int64_t(1) << 31), "multi-step heat grid too large");
    if (g.vol_pow2)
        heat_multi_k<T, B, K, true><<<dim3(unsigned(nwg)), dim3(64, kMultiWaves), 0, s>>>(g, (const T*)in, (T*)out, T(alpha),
                                                                               T(dt), ns, bz, nzt, nyt);
    else
        heat_multi_k<T, B, K, false><<<dim3(unsigned(nwg)), dim3(64, kMultiWaves), 0, s>>>(g, (const T*)in, (T*)out, T(alpha),
                                                                                T(dt), ns, bz, nzt, nyt);
    CFD_HIP(hipGetLastError());
}

template <class T>
void launch_heat_multi(const UniformGeo<T>& g, const void* in, void* out, double alpha, double dt, int B, int K,
                       int ns, hipStream_t s) {
    CFD_REQUIRE(ns >= 1 && ns <= K, "multi-step heat launch: 1 to K steps");
    switch (B * 8 + K) {
        case 4 * 8 + 2: launch_multi<T, 4, 2>(g, in, out, alpha, dt, ns, s); break;
        case 4 * 8 + 4: launch_multi<T, 4, 4>(g, in, out, alpha, dt, ns, s); break;
        case 8 * 8 + 2: launch_multi<T, 8, 2>(g, in, out, alpha, dt, ns, s); break;
        case 8 * 8 + 4:
            if constexpr (sizeof(T) == 4) {
                launch_multi<T, 8, 4>(g, in, out, alpha, dt, ns, s);
                break;
            }
            [[fallthrough]];
        default: throw Error(CFD_E_INVALID, "multi-step heat boxes: B 4 or 8, K 2 or 4 (B 8 with K 4 in FP32 only)");
    }
}

template void launch_heat_multi<float>(const UniformGeo<float>&, const void*, void*, double, double, int, int, int,
                                       hipStream_t);
template void launch_heat_multi<double>(const UniformGeo<double>&, const void*, void*, double, double, int, int, int,
                                        hipStream_t);

}  // namespace cfd
