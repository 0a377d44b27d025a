// Row form of the fused acoustic step (variants 7 to 9) on an exactly uniform
// mesh whose z extent is a multiple of 4 and at most 256.
//
// One wave holds whole z rows of a plane: RW = 256 / Z consecutive rows, one
// contiguous span of at most 256 cells, four consecutive cells per lane.  Every
// load and store is a 16-byte vector access (32 bytes in FP64) at a scalar
// base plus a fixed lane offset.  The z-neighbours of a lane's cells are its
// own registers or the adjacent lane's, reached with DPP wave shifts; at the z
// walls (a row's first and last lane) the neighbour is the cell itself (the
// reference's neighbour = self rule) and the wall flux is zero, so no z halo
// is loaded or computed.
//
// A workgroup holds TY output waves (rows y0 .. y0+TY*RW-1) and two halo waves
// (rows y0-1 and y0+TY*RW).  It marches along x over a chunk of L planes; the
// output waves keep p(x), p(x+1) and u'(x-1), u'(x) per lane in registers,
// as the 2.5D tile form does.  Per plane:
//   1. the output waves compute u'(x+1) and put their p rows into the LDS
//      plane; the halo waves put the halo p rows there;
//   2. barrier; the output waves compute v' (y-neighbours from LDS, self at the
//      y walls) and w' (z-neighbours from registers), the halo waves the halo
//      rows' v';
//   3. barrier; the output waves read the v' neighbours and update p.
// One p plane and one v' plane live in LDS: a wave writes plane x+1 only after
// the second barrier of plane x, which every wave passes after its last read
// of plane x, so a single buffer is race-free.
//
// The arithmetic per cell is the 2.5D tile's (march.hpp), which is the
// reference evaluation's sequence, so results equal the oracle bit for bit.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"
#include "march.hpp"
#include "row.hpp"

namespace cfd {


template <class T, int TY, bool PW, bool KF, bool NTL>
__global__ void __launch_bounds__(64 * (TY + 2))
wave_row_k(UniformGeo<T> g, const T* __restrict__ p, const T* __restrict__ u, const T* __restrict__ v,
           const T* __restrict__ w, T* __restrict__ po, T* __restrict__ uo, T* __restrict__ vo, T* __restrict__ wo,
           T K, const T* __restrict__ Kf, T dt, int L, int nyt, int order, int nt) {
    using V4 = typename row::Row4<T>::type;
    __shared__ V4 sp[TY + 2][64];
    __shared__ V4 sv[TY + 2][64];
    // workgroup -> (y tile, x chunk).  order bit 1: walk the tiles last to
    // first (alternate steps run in opposite directions, so a step starts on
    // the planes the previous one wrote last, which the caches still hold);
    // order bit 0: each XCD (round-robin dispatch) gets a contiguous run of y
    // tiles, so tiles sharing halo rows share an L2
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
    const int y0 = yt * TY * row::rows_per_wave(g.Z);
    // chunks away from the x walls need no x-boundary tests
    const bool inner = x0 >= 1 && x1 + 2 <= X;
#define CFD_ROW(XB, MR) \
    row::row_tile<T, TY, PW, KF, XB, NTL, MR>(g, p, u, v, w, po, uo, vo, wo, K, Kf, dt, x0, x1, y0, wv, nt != 0, sp, sv)
    if (row::rows_per_wave(g.Z) == 1) {  // one row per wave: uniform rows
        if (inner) CFD_ROW(false, false);
        else CFD_ROW(true, false);
    } else {
        if (inner) CFD_ROW(false, true);
        else CFD_ROW(true, true);
    }
#undef CFD_ROW
}

template <class T, int TY, bool PW, bool KF, bool NTL>
static void launch_row_k(const UniformGeo<T>& g, void* const* in, void* const* out, double K, const void* Kf,
                         double dt, int L, int order, bool nt, hipStream_t s) {
    const int64_t rows = int64_t(TY) * row::rows_per_wave(g.Z);  // rows per tile
    const int nyt = int((g.Y + rows - 1) / rows);
    const int64_t nxc = (g.X + L - 1) / L;
    const int64_t nwg = int64_t(nyt) * nxc;
    CFD_REQUIRE(nwg < (int64_t(1) << 31), "row march grid too large");
    const int ord = ((order & 1) && nwg % 8 == 0 ? 1 : 0) | (order & 2);
    wave_row_k<T, TY, PW, KF, NTL><<<dim3(unsigned(nwg)), dim3(64, TY + 2), 0, s>>>(
        g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3], (T*)out[0], (T*)out[1], (T*)out[2],
        (T*)out[3], T(K), (const T*)Kf, T(dt), L, nyt, ord, nt ? 1 : 0);
    CFD_HIP(hipGetLastError());
}

template <class T, int TY, bool NTL>
static void launch_row_ty(const UniformGeo<T>& g, void* const* in, void* const* out, double K, const void* Kf,
                          double dt, int L, int order, bool nt, hipStream_t s) {
    if constexpr (sizeof(T) == 8 && TY == 8) {
        // The FP64 8-row tile with a K field would spill registers, and the
        // tile reads other lanes' registers (the z neighbours), which a reload
        // under a partial exec mask leaves stale in the inactive lanes: such
        // runs take the 4-row tile (same results, every form is bit-exact).
        if (Kf) {
            launch_row_ty<T, 4, NTL>(g, in, out, K, Kf, dt, L, order, nt, s);
            return;
        }
        if (g.vol_pow2) launch_row_k<T, TY, true, false, NTL>(g, in, out, K, Kf, dt, L, order, nt, s);
        else launch_row_k<T, TY, false, false, NTL>(g, in, out, K, Kf, dt, L, order, nt, s);
    } else {
        if (g.vol_pow2) {
            if (Kf) launch_row_k<T, TY, true, true, NTL>(g, in, out, K, Kf, dt, L, order, nt, s);
            else launch_row_k<T, TY, true, false, NTL>(g, in, out, K, Kf, dt, L, order, nt, s);
        } else {
            if (Kf) launch_row_k<T, TY, false, true, NTL>(g, in, out, K, Kf, dt, L, order, nt, s);
            else launch_row_k<T, TY, false, false, NTL>(g, in, out, K, Kf, dt, L, order, nt, s);
        }
    }
}

template <class T, int TY>
static void launch_row_c(const UniformGeo<T>& g, void* const* in, void* const* out, double K, const void* Kf,
                         double dt, int L, int order, int cache, hipStream_t s) {
    if (cache & 2) launch_row_ty<T, TY, true>(g, in, out, K, Kf, dt, L, order, (cache & 1) != 0, s);
    else launch_row_ty<T, TY, false>(g, in, out, K, Kf, dt, L, order, (cache & 1) != 0, s);
}

bool wave_row_supported(int64_t Z) { return Z >= 4 && Z <= 256 && Z % 4 == 0; }

template <class T>
bool launch_wave_row(const UniformGeo<T>& g, void* const* in, void* const* out, double K, const void* Kf, double dt,
                     int ty, int L, int order, int cache, hipStream_t s) {
    // rows start on 4-cell boundaries; the vector accesses need the field
    // pointers aligned to four cells as well
    if (!wave_row_supported(g.Z)) return false;
    const uintptr_t al = 4 * sizeof(T) - 1;
    for (int i = 0; i < 4; ++i)
        if ((reinterpret_cast<uintptr_t>(in[i]) & al) || (reinterpret_cast<uintptr_t>(out[i]) & al)) return false;
    if (Kf && (reinterpret_cast<uintptr_t>(Kf) & al)) return false;
    switch (ty) {
        case 4: launch_row_c<T, 4>(g, in, out, K, Kf, dt, L, order, cache, s); break;
        case 8: launch_row_c<T, 8>(g, in, out, K, Kf, dt, L, order, cache, s); break;
        case 14:  // FP32 only: the FP64 form of this tile spills registers (see launch_row_ty)
            if constexpr (sizeof(T) == 4) {
                launch_row_c<T, 14>(g, in, out, K, Kf, dt, L, order, cache, s);
                break;
            }
            [[fallthrough]];
        default: throw Error(CFD_E_INVALID, "row-form tiles have 4 or 8 rows (or 14 in FP32)");
    }
    return true;
}

template bool launch_wave_row<float>(const UniformGeo<float>&, void* const*, void* const*, double, const void*, double,
                                     int, int, int, int, hipStream_t);
template bool launch_wave_row<double>(const UniformGeo<double>&, void* const*, void* const*, double, const void*,
                                      double, int, int, int, int, hipStream_t);

}  // namespace cfd
