// Two time steps per pass of the row form (variants 11 and 12): constant K, z
// extent a multiple of 4 and at most 256; FP32 tiles of 4, 8 or 10 rows, FP64
// tiles of 2 rows (registers).
//
// The single-step row form (wave_row.hip) reads and writes every field once per
// step and runs at the memory's streaming rate.  This kernel advances the
// state two steps per pass, so each field crosses HBM once per two steps; the
// price is recomputing the first step on a wider band of rows.
//
// A workgroup owns output rows y0 .. y0+TY-1 and holds one wave per row of
// y0-3 .. y0+TY+2 (TY + 6 waves).  With d the distance of a wave's row from the
// output band:
//   * d <= 2: the first step in full (u', v', w', p'), as the single-step kernel;
//   * d == 3: the first step's v' only (the y-halo of the d == 2 rows' p');
//   * d <= 1: the second step's v'' and w'';
//   * d == 0: the second step in full, and the outputs.
// Marching along x, iteration x computes the first step at plane x and the
// second step at plane x-2, with every quantity the second step needs kept in
// per-lane register queues (p' of three planes, u', v', w' of two, u'' of three).
// Four LDS planes carry the y-neighbours: p and v' of the first step, p' and v''
// of the second.  Per iteration:
//   A. u'(x+1); p(x) and p'(x-2) into LDS;
//   B. (barrier) v'(x), w'(x) and v''(x-2), w''(x-2); v'(x), v''(x-2) into LDS;
//   C. (barrier) p'(x), then u''(x-1), then p''(x-2) and the stores.
// A plane is rewritten one iteration later, after a barrier every wave passes
// only once its reads of the plane are done, so single buffers suffice.
//
// Each value is computed by the operation sequence of the single-step kernels,
// so two passes of this kernel equal four single steps bit for bit.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"
#include "march.hpp"
#include "row.hpp"
#include "row2.hpp"

namespace cfd {


template <class T, int TY, bool PW, bool NTL>
__global__ void __launch_bounds__(64 * (TY + 6))
wave_row2_k(UniformGeo<T> g, const T* __restrict__ p, const T* __restrict__ u, const T* __restrict__ v,
            const T* __restrict__ w, T* __restrict__ po, T* __restrict__ uo, T* __restrict__ vo, T* __restrict__ wo,
            T K, T dt, int L, int nyt, int order, int nt) {
    using V4 = row2::V4T<T>;
    __shared__ V4 P1[TY + 6][64];
    __shared__ V4 V1[TY + 6][64];
    __shared__ V4 P2[TY + 6][64];
    __shared__ V4 V2[TY + 6][64];
    const int64_t nwg = gridDim.x;
    int64_t lin = blockIdx.x;
    if (order & 2) lin = nwg - 1 - lin;
    if (order & 1) lin = (lin % 8) * (nwg / 8) + lin / 8;
    const int yt = int(lin % nyt);
    const int64_t xc = lin / nyt;
    const int j = __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    const int64_t X = g.X;
    const int64_t x0 = xc * L;
    const int64_t x1 = (x0 + L < X) ? x0 + L : X;
    const int y0 = yt * TY;
    // chunks four planes away from the x walls need no x-boundary tests
    if (x0 >= 4 && x1 + 4 <= X)
        row2::tile2<T, TY, PW, false, NTL>(g, p, u, v, w, po, uo, vo, wo, K, dt, x0, x1, y0, j, nt != 0, P1, V1, P2,
                                           V2);
    else
        row2::tile2<T, TY, PW, true, NTL>(g, p, u, v, w, po, uo, vo, wo, K, dt, x0, x1, y0, j, nt != 0, P1, V1, P2,
                                          V2);
}

template <class T, int TY, bool PW>
static void launch_row2_k(const UniformGeo<T>& g, void* const* in, void* const* out, double K, double dt, int L,
                          int order, bool nt, hipStream_t s) {
    const int nyt = int((g.Y + TY - 1) / TY);
    const int64_t nxc = (g.X + L - 1) / L;
    const int64_t nwg = int64_t(nyt) * nxc;
    CFD_REQUIRE(nwg < (int64_t(1) << 31), "row march grid too large");
    const int ord = ((order & 1) && nwg % 8 == 0 ? 1 : 0) | (order & 2);
    wave_row2_k<T, TY, PW, true><<<dim3(unsigned(nwg)), dim3(64, TY + 6), 0, s>>>(
        g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3], (T*)out[0], (T*)out[1], (T*)out[2],
        (T*)out[3], T(K), T(dt), L, nyt, ord, nt ? 1 : 0);
    CFD_HIP(hipGetLastError());
}

template <class T, int TY>
static void launch_row2_ty(const UniformGeo<T>& g, void* const* in, void* const* out, double K, double dt, int L,
                           int order, bool nt, hipStream_t s) {
    if (g.vol_pow2) launch_row2_k<T, TY, true>(g, in, out, K, dt, L, order, nt, s);
    else launch_row2_k<T, TY, false>(g, in, out, K, dt, L, order, nt, s);
}

template <class T, int NW, bool PW, bool NTL>
__global__ void __launch_bounds__(64 * NW)
wave_row2m_k(UniformGeo<T> g, const T* __restrict__ p, const T* __restrict__ u, const T* __restrict__ v,
             const T* __restrict__ w, T* __restrict__ po, T* __restrict__ uo, T* __restrict__ vo, T* __restrict__ wo,
             T K, T dt, int L, int nyt, int order, int nt) {
    using V4 = row2::V4T<T>;
    __shared__ V4 P1[NW][64];
    __shared__ V4 V1[NW][64];
    __shared__ V4 P2[NW][64];
    __shared__ V4 V2[NW][64];
    const int64_t nwg = gridDim.x;
    int64_t lin = blockIdx.x;
    if (order & 2) lin = nwg - 1 - lin;
    if (order & 1) lin = (lin % 8) * (nwg / 8) + lin / 8;
    const int yt = int(lin % nyt);
    const int64_t xc = lin / nyt;
    const int wv = __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    const int64_t X = g.X;
    const int64_t x0 = xc * L;
    const int64_t x1 = (x0 + L < X) ? x0 + L : X;
    const int y0 = yt * (NW * row::rows_per_wave(g.Z) - 8);
    if (x0 >= 4 && x1 + 4 <= X)
        row2::tile2m<T, NW, PW, false, NTL>(g, p, u, v, w, po, uo, vo, wo, K, dt, x0, x1, y0, wv, nt != 0, P1, V1,
                                            P2, V2);
    else
        row2::tile2m<T, NW, PW, true, NTL>(g, p, u, v, w, po, uo, vo, wo, K, dt, x0, x1, y0, wv, nt != 0, P1, V1,
                                           P2, V2);
}

template <class T, int NW>
static void launch_row2m(const UniformGeo<T>& g, void* const* in, void* const* out, double K, double dt, int L,
                         int order, bool nt, hipStream_t s) {
    const int64_t rows = int64_t(NW) * row::rows_per_wave(g.Z) - 8;  // output rows per tile
    CFD_REQUIRE(rows >= 1, "two-step tile without output rows");
    const int nyt = int((g.Y + rows - 1) / rows);
    const int64_t nxc = (g.X + L - 1) / L;
    const int64_t nwg = int64_t(nyt) * nxc;
    CFD_REQUIRE(nwg < (int64_t(1) << 31), "row march grid too large");
    const int ord = ((order & 1) && nwg % 8 == 0 ? 1 : 0) | (order & 2);
    const T* const* fi = (const T* const*)in;
    T* const* fo = (T* const*)out;
    if (g.vol_pow2)
        wave_row2m_k<T, NW, true, true><<<dim3(unsigned(nwg)), dim3(64, NW), 0, s>>>(
            g, fi[0], fi[1], fi[2], fi[3], fo[0], fo[1], fo[2], fo[3], T(K), T(dt), L, nyt, ord, nt ? 1 : 0);
    else
        wave_row2m_k<T, NW, false, true><<<dim3(unsigned(nwg)), dim3(64, NW), 0, s>>>(
            g, fi[0], fi[1], fi[2], fi[3], fo[0], fo[1], fo[2], fo[3], T(K), T(dt), L, nyt, ord, nt ? 1 : 0);
    CFD_HIP(hipGetLastError());
}

template <class T>
bool launch_wave_row2(const UniformGeo<T>& g, void* const* in, void* const* out, double K, double dt, int ty, int L,
                      int order, bool nt, hipStream_t s) {
    if (!wave_row_supported(g.Z)) return false;
    const uintptr_t al = 4 * sizeof(T) - 1;
    for (int i = 0; i < 4; ++i)
        if ((reinterpret_cast<uintptr_t>(in[i]) & al) || (reinterpret_cast<uintptr_t>(out[i]) & al)) return false;
    const bool multi = row::rows_per_wave(g.Z) > 1;
    if constexpr (sizeof(T) == 8) {
        // FP64 keeps twice the registers per queue entry: only the tiles that
        // fit without spilling, 8 waves per workgroup (2 output rows of full
        // rows, or 8 waves of several rows each)
        if (ty != 2) throw Error(CFD_E_INVALID, "FP64 two-step row-form tiles have 2 rows");
        if (multi) launch_row2m<T, 8>(g, in, out, K, dt, L, order, nt, s);
        else launch_row2_ty<T, 2>(g, in, out, K, dt, L, order, nt, s);
        return true;
    } else {
        if (multi) {  // several rows per wave: 8, 10 or 12 waves per workgroup
            switch (ty) {
                case 4: launch_row2m<T, 8>(g, in, out, K, dt, L, order, nt, s); break;
                case 8: launch_row2m<T, 10>(g, in, out, K, dt, L, order, nt, s); break;
                case 10: launch_row2m<T, 12>(g, in, out, K, dt, L, order, nt, s); break;
                default: throw Error(CFD_E_INVALID, "two-step row-form tiles have 4, 8 or 10 rows");
            }
            return true;
        }
        switch (ty) {
            case 4: launch_row2_ty<T, 4>(g, in, out, K, dt, L, order, nt, s); break;
            case 8: launch_row2_ty<T, 8>(g, in, out, K, dt, L, order, nt, s); break;
            case 10: launch_row2_ty<T, 10>(g, in, out, K, dt, L, order, nt, s); break;
            default: throw Error(CFD_E_INVALID, "two-step row-form tiles have 4, 8 or 10 rows");
        }
        return true;
    }
}

template bool launch_wave_row2<float>(const UniformGeo<float>&, void* const*, void* const*, double, double, int, int,
                                      int, bool, hipStream_t);
template bool launch_wave_row2<double>(const UniformGeo<double>&, void* const*, void* const*, double, double, int,
                                       int, int, bool, hipStream_t);

}  // namespace cfd
