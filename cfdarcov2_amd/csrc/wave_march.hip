// Fused acoustic step on an exactly uniform mesh, 2.5D form.
//
// A workgroup of 64 x TYT threads owns a TY x 64 (y, z) tile of columns,
// TY = TYT * R (each thread carries R interleaved rows, row = r*TYT + ly), and
// marches along x over a chunk of L planes.  Per plane every field is read and
// written once (32 B per cell-update in FP32):
//   * x-direction: each (row, column) slot keeps p(x-1..x+2) and u'(x-1..x+1)
//     in registers;
//   * y/z-directions: the p plane with a 2-cell halo goes to LDS, the tile
//     computes v' (y-halo 1) and w' (z-halo 1) into LDS, and the pressure update
//     reads its neighbours from there.
// Taller tiles (R > 1) cut the re-read y-halo of p and v from 4 and 2 rows per
// TYT rows to the same per TY rows without more threads per workgroup.
// LDS planes are double-buffered (two barriers per plane) and the next plane's
// global loads are issued before the current plane's arithmetic.  The
// arithmetic is the sequence of the reference evaluation (see wave_uniform_k in
// stencil.hip), so results equal the oracle bit for bit.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"
#include "march.hpp"

namespace cfd {

template <class T, int TYT, int R, bool NT = false>
__global__ void __launch_bounds__(64 * TYT) wave_march_k(UniformGeo<T> g, const T* __restrict__ p,
                                                         const T* __restrict__ u, const T* __restrict__ v,
                                                         const T* __restrict__ w, T* __restrict__ po,
                                                         T* __restrict__ uo, T* __restrict__ vo,
                                                         T* __restrict__ wo, T K, const T* __restrict__ Kf, T dt,
                                                         int L, int nzt, int nyt, int remap) {
    __shared__ MarchLds<T, TYT, R> lds;
    // workgroup -> tile; with remap the 8 XCDs (round-robin dispatch) each get a
    // contiguous run of tiles, so tiles sharing y/z halos share an L2
    const int64_t nwg = gridDim.x;
    int64_t lin_id = blockIdx.x;
    if (remap) lin_id = (lin_id % 8) * (nwg / 8) + lin_id / 8;
    const int64_t zt = lin_id % nzt;
    const int64_t yt = (lin_id / nzt) % nyt;
    const int64_t xc = lin_id / (int64_t(nzt) * nyt);
    if (g.vol_pow2) wave_march_tile<T, TYT, R, NT, true>(g, p, u, v, w, po, uo, vo, wo, K, Kf, dt, L, zt, yt, xc, lds);
    else wave_march_tile<T, TYT, R, NT, false>(g, p, u, v, w, po, uo, vo, wo, K, Kf, dt, L, zt, yt, xc, lds);
}

template <class T, int TYT, int R, bool NT = false>
static void launch_one(const UniformGeo<T>& g, void* const* in, void* const* out, double K, const void* Kf,
                       double dt, int L, int remap, hipStream_t s) {
    constexpr int TY = TYT * R;
    const int nzt = int((g.Z + 63) / 64);
    const int nyt = int((g.Y + TY - 1) / TY);
    const int nxc = int((g.X + L - 1) / L);
    const int64_t nwg = int64_t(nzt) * nyt * nxc;
    CFD_REQUIRE(nwg < (int64_t(1) << 31), "march grid too large");
    const int rm = (remap && nwg % 8 == 0) ? 1 : 0;
    wave_march_k<T, TYT, R, NT><<<dim3(unsigned(nwg)), dim3(64, TYT), 0, s>>>(
        g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3], (T*)out[0], (T*)out[1], (T*)out[2],
        (T*)out[3], T(K), (const T*)Kf, T(dt), L, nzt, nyt, rm);
    CFD_HIP(hipGetLastError());
}

// Variant 1: tile_rows selects (threads in y) x (rows per thread): 4 = 4x1,
// 8 = 4x2, 16 = 4x4, 32 = 8x4.  Variants 7 to 9 take the row form
// (wave_row.hip) and 11 and 12 the two-step row form (wave_row2.hip).
template <class T>
void launch_wave_march(const cfd_mesh_s* m, void* const a[4], void* const b[4], double K, const void* Kf, double dt,
                       int64_t nsteps, int variant, int ty, int L, int remap, hipStream_t s) {
    const UniformGeo<T>& g = (sizeof(T) == 4) ? *(const UniformGeo<T>*)&m->ug32 : *(const UniformGeo<T>*)&m->ug64;
    CFD_REQUIRE(L >= 1, "march chunk must be positive");
    int64_t k0 = 0;
    if (variant >= 11) {
        // two steps per launch, in pairs (a -> b -> a) so the state ends where
        // single steps leave it; constant K.  The rest of the steps, and every
        // step of a case the two-step form cannot take, run the single-step
        // row form.
        if (!Kf) {
            for (; k0 + 4 <= nsteps; k0 += 4) {
                if (!launch_wave_row2<T>(g, a, b, K, dt, ty, L, remap & 1, variant == 12, s)) break;
                launch_wave_row2<T>(g, b, a, K, dt, ty, L, (remap & 1) | (remap & 2), variant == 12, s);
            }
        }
        variant = 9;
        ty = ty == 10 ? 8 : ty == 2 ? 4 : ty;
    }
    for (int64_t k = k0; k < nsteps; ++k) {
        void* const* in = (k & 1) ? b : a;
        void* const* out = (k & 1) ? a : b;
        if (variant >= 7) {  // row form; misaligned fields take the 4-row march tile
            // remap bit 1: odd steps walk the tiles in reverse.  Cache policy:
            // 7 non-temporal stores, 8 plain, 9 non-temporal loads
            const int order = (remap & 1) | ((remap & 2) && (k & 1) ? 2 : 0);
            const int cache = variant == 7 ? 1 : variant == 8 ? 0 : 2;
            if (!launch_wave_row<T>(g, in, out, K, Kf, dt, ty, L, order, cache, s))
                launch_one<T, 4, 1, true>(g, in, out, K, Kf, dt, L, remap & 1, s);
            continue;
        }
        switch (ty) {
            case 4: launch_one<T, 4, 1>(g, in, out, K, Kf, dt, L, remap, s); break;
            case 8: launch_one<T, 4, 2>(g, in, out, K, Kf, dt, L, remap, s); break;
            case 16: launch_one<T, 4, 4>(g, in, out, K, Kf, dt, L, remap, s); break;
            case 32: launch_one<T, 8, 4>(g, in, out, K, Kf, dt, L, remap, s); break;
            default: throw Error(CFD_E_INVALID, "march tile rows must be 4, 8, 16 or 32");
        }
    }
}

template void launch_wave_march<float>(const cfd_mesh_s*, void* const*, void* const*, double, const void*, double,
                                       int64_t, int, int, int, int, hipStream_t);
template void launch_wave_march<double>(const cfd_mesh_s*, void* const*, void* const*, double, const void*, double,
                                        int64_t, int, int, int, int, hipStream_t);

}  // namespace cfd
