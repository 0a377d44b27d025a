// Several heat steps per launch on an exactly uniform mesh whose state fits
// the caches (no reference counterpart; the form for small meshes, where a
// single step lasts about as long as a launch).
//
// A workgroup owns an output box of B x B x BZ cells (BZ = min(Z, 64 - 2K)).
// Its region is the box widened by K cells on every side that is not a wall
// (at most (B + 2K) x (B + 2K) rows of at most 64 z-cells).  Step s computes
// the region narrowed by s on its non-wall sides, which is all that step
// s + 1 reads; at the walls a cell's outer neighbour is the cell itself, as in
// the single-step kernel.  Cells near a box's sides are computed by several
// workgroups (the redundant work that buys K - 1 fewer launches).  Each cell's
// arithmetic is heat_uniform_cell's (stencil.hip), in the same order, so ns
// steps here equal ns single steps bit for bit.
//
// Threads: 64 lanes along z x 16 waves.  Every wave owns a fixed set of the
// region's (x, y) rows for the whole launch and keeps their cells in
// registers; the region is mirrored in one LDS buffer, from which a cell
// reads its x and y neighbours.  z neighbours are the adjacent lanes' values
// (DPP lane shifts).  Per step: compute every owned row of the narrowed
// region into registers, barrier, write them to LDS, barrier.  The region's
// loads are issued together before the first LDS store, and the output box
// is stored from the registers.  Row ownership and the x/y wall tests are
// wave-uniform; only the z walls are tested per lane.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"
#include "row.hpp"

namespace cfd {

constexpr int kMultiWaves = 16;

template <class T, int B, int K, bool PW>
__global__ void __launch_bounds__(64 * kMultiWaves)
    heat_multi_k(UniformGeo<T> g, const T* __restrict__ Tin, T* __restrict__ Tout, T alpha, T dt, int ns, int bz,
                 int nzt, int nyt) {
    constexpr int R = B + 2 * K;                                   // region extent along x and y
    constexpr int NR = (R * R + kMultiWaves - 1) / kMultiWaves;    // rows per wave
    __shared__ T buf[R * R * 64];
    const int X = int(g.X), Y = int(g.Y), Z = int(g.Z);
    const int lin = int(blockIdx.x);
    const int zt = lin % nzt, yt = (lin / nzt) % nyt, xt = lin / (nzt * nyt);
    const int lane = int(threadIdx.x);
    const int wy = __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    // output box and region (clipped to the mesh)
    const int ox0 = xt * B, ox1 = min(ox0 + B, X);
    const int oy0 = yt * B, oy1 = min(oy0 + B, Y);
    const int oz0 = zt * bz, oz1 = min(oz0 + bz, Z);
    const int gx0 = max(ox0 - K, 0), gx1 = min(ox1 + K, X);
    const int gy0 = max(oy0 - K, 0), gy1 = min(oy1 + K, Y);
    const int gz0 = max(oz0 - K, 0), gz1 = min(oz1 + K, Z);
    const int nry = gy1 - gy0, nrows = (gx1 - gx0) * nry;
    const int z = gz0 + lane;
    const bool zin = z < gz1;
    const bool zlo = z == 0, zhi = z == Z - 1;

    // the wave's rows (x = -1: no row) and their cells
    int rx[NR], ry[NR];
    T cur[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int r = wy + i * kMultiWaves;
        rx[i] = r < nrows ? gx0 + r / nry : -1;
        ry[i] = r < nrows ? gy0 + r % nry : -1;
        cur[i] = (rx[i] >= 0 && zin) ? Tin[(int64_t(rx[i]) * Y + ry[i]) * Z + z] : T(0);
    }
    auto slot = [&](int x, int y) { return ((x - gx0) * R + (y - gy0)) * 64 + lane; };
#pragma unroll
    for (int i = 0; i < NR; ++i)
        if (rx[i] >= 0) buf[slot(rx[i], ry[i])] = cur[i];
    __syncthreads();

    const T A0 = g.area[0], A1 = g.area[1], A2 = g.area[2];
    auto axis = [&](T L, T pc, T lo_v, T hi_v, bool lo, bool hi, int a, T A) {
        const T al_lo = lo ? g.alpha_b[a] : g.alpha[a];
        const T al_hi = hi ? g.alpha_b[a] : g.alpha[a];
        const T s_lo = (-pc + (lo ? pc : lo_v)) * al_lo;
        const T s_hi = (-pc + (hi ? pc : hi_v)) * al_hi;
        const T num = s_lo * A + s_hi * A;
        // a power-of-two volume divides exactly: the product is the same number
        if constexpr (PW) return L + num * g.inv_vol;
        else return L + num / g.vol;
    };
    for (int s = 1; s <= ns; ++s) {
        // the region narrowed by s on its non-wall sides
        const int bx0 = gx0 == 0 ? 0 : gx0 + s, bx1 = gx1 == X ? gx1 : gx1 - s;
        const int by0 = gy0 == 0 ? 0 : gy0 + s, by1 = gy1 == Y ? gy1 : gy1 - s;
        const int bz0 = gz0 == 0 ? 0 : gz0 + s, bz1 = gz1 == Z ? gz1 : gz1 - s;
        const bool zact = z >= bz0 && z < bz1;
        T nv[NR];
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            nv[i] = cur[i];
            const int x = rx[i], y = ry[i];
            // z neighbours from the adjacent lanes (lanes 0 and 63 keep their
            // own value; they are walls or outside the narrowed region)
            const T pc = cur[i];
            const T zl = row::lane_shift<row::kWaveShr1>(pc, pc);
            const T zh = row::lane_shift<row::kWaveShl1>(pc, pc);
            if (x < bx0 || x >= bx1 || y < by0 || y >= by1) continue;   // wave-uniform; x = -1 fails
            const int c = slot(x, y);
            const bool xlo = x == 0, xhi = x == X - 1, ylo = y == 0, yhi = y == Y - 1;
            const T xl = xlo ? pc : buf[c - R * 64], xh = xhi ? pc : buf[c + R * 64];
            const T yl = ylo ? pc : buf[c - 64], yh = yhi ? pc : buf[c + 64];
            T L = T(0);
            L = axis(L, pc, xl, xh, xlo, xhi, 0, A0);
            L = axis(L, pc, yl, yh, ylo, yhi, 1, A1);
            L = axis(L, pc, zl, zh, zlo, zhi, 2, A2);
            if (zact) nv[i] = dt * (alpha * L) + pc;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int x = rx[i], y = ry[i];
            if (x >= bx0 && x < bx1 && y >= by0 && y < by1 && zact) buf[slot(x, y)] = nv[i];
            cur[i] = nv[i];
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int x = rx[i], y = ry[i];
        if (x >= ox0 && x < ox1 && y >= oy0 && y < oy1 && z >= oz0 && z < oz1)
            Tout[(int64_t(x) * Y + y) * Z + z] = cur[i];
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
        case 8 * 8 + 4: launch_multi<T, 8, 4>(g, in, out, alpha, dt, ns, s); break;
        default: throw Error(CFD_E_INVALID, "multi-step heat boxes: B 4 or 8, K 2 or 4");
    }
}

template void launch_heat_multi<float>(const UniformGeo<float>&, const void*, void*, double, double, int, int, int,
                                       hipStream_t);
template void launch_heat_multi<double>(const UniformGeo<double>&, const void*, void*, double, double, int, int, int,
                                        hipStream_t);

}  // namespace cfd
