// Two heat steps per launch in row form (heat variant 1) on an exactly uniform
// mesh whose z extent is a multiple of 4 and at most 256, constant alpha.
//
// The heat step moves 2 x 4 (FP32) or 2 x 8 (FP64) bytes per cell and small
// meshes are bound by the launch itself, so halving the launches and the bytes
// per step is the whole gain.  A workgroup of NW waves holds a band of rows
// (256/Z rows per wave, four cells per lane) and marches x over a chunk of
// planes; the tile code is row::heat_row2_tile (row.hpp).  Two launches of two
// steps each leave the state where four single steps would, so runs go in
// pairs and the remaining steps take the per-thread kernel.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"
#include "march.hpp"
#include "row.hpp"

namespace cfd {

template <class T, int NW, bool PW>
__global__ void __launch_bounds__(64 * NW) heat_row2_k(UniformGeo<T> g, const T* __restrict__ Tin,
                                                      T* __restrict__ Tout, T alpha, T dt, int L, int nyt) {
    using V4 = typename row::Row4<T>::type;
    __shared__ V4 s1[NW][64];
    __shared__ V4 s2[NW][64];
    const int64_t lin = blockIdx.x;
    const int yt = int(lin % nyt);
    const int64_t xc = lin / nyt;
    const int wv = __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    const int64_t X = g.X;
    const int64_t x0 = xc * L;
    const int64_t x1 = (x0 + L < X) ? x0 + L : X;
    const int y0 = yt * (NW * row::rows_per_wave(g.Z) - 4);
    // chunks two planes away from the x walls need no x-boundary tests
    if (x0 >= 2 && x1 + 2 <= X)
        row::heat_row2_tile<T, NW, PW, false, true>(g, Tin, Tout, alpha, dt, x0, x1, y0, wv, s1, s2);
    else
        row::heat_row2_tile<T, NW, PW, true, true>(g, Tin, Tout, alpha, dt, x0, x1, y0, wv, s1, s2);
}

template <class T, int NW>
static void launch_nw(const UniformGeo<T>& g, const void* in, void* out, double alpha, double dt, int L,
                      hipStream_t s) {
    const int64_t rows = int64_t(NW) * row::rows_per_wave(g.Z) - 4;  // output rows per tile
    const int nyt = int((g.Y + rows - 1) / rows);
    const int64_t nxc = (g.X + L - 1) / L;
    const int64_t nwg = int64_t(nyt) * nxc;
    CFD_REQUIRE(nwg < (int64_t(1) << 31), "heat row grid too large");
    if (g.vol_pow2)
        heat_row2_k<T, NW, true><<<dim3(unsigned(nwg)), dim3(64, NW), 0, s>>>(g, (const T*)in, (T*)out, T(alpha),
                                                                           T(dt), L, nyt);
    else
        heat_row2_k<T, NW, false><<<dim3(unsigned(nwg)), dim3(64, NW), 0, s>>>(g, (const T*)in, (T*)out, T(alpha),
                                                                            T(dt), L, nyt);
    CFD_HIP(hipGetLastError());
}

bool heat_row_supported(int64_t Z, int nw) {
    return wave_row_supported(Z) && int64_t(nw) * row::rows_per_wave(Z) > 4;
}

template <class T>
bool launch_heat_row2(const UniformGeo<T>& g, const void* in, void* out, double alpha, double dt, int nw, int L,
                      hipStream_t s) {
    if (!heat_row_supported(g.Z, nw)) return false;
    const uintptr_t al = 4 * sizeof(T) - 1;
    if ((reinterpret_cast<uintptr_t>(in) & al) || (reinterpret_cast<uintptr_t>(out) & al)) return false;
    switch (nw) {
        case 8: launch_nw<T, 8>(g, in, out, alpha, dt, L, s); break;
        case 12: launch_nw<T, 12>(g, in, out, alpha, dt, L, s); break;
        case 16: launch_nw<T, 16>(g, in, out, alpha, dt, L, s); break;
        default: throw Error(CFD_E_INVALID, "two-step heat tiles have 8, 12 or 16 waves");
    }
    return true;
}

template bool launch_heat_row2<float>(const UniformGeo<float>&, const void*, void*, double, double, int, int,
                                      hipStream_t);
template bool launch_heat_row2<double>(const UniformGeo<double>&, const void*, void*, double, double, int, int,
                                       hipStream_t);

}  // namespace cfd
