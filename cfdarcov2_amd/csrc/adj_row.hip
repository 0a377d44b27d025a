// Row form of the adjoint's two per-step kernels on an exactly uniform mesh
// whose z extent is a multiple of 4 and at most 256 (no reference
// counterpart):
//   MODE 1, the forward step that records p' at the sensor cells and, for the
//           K gradient's recomputation, the divergence S of every cell;
//   MODE 2, the transposed step lam_t -> lam_{t-1} with constant K, adding
//           the trace gradients at the sensors;
//   MODE 3, MODE 2 that also adds the step's contribution (lam_p dt) S_t to
//           the K gradient.
// Both run the row tile of the fused step (row.hpp): one wave per z row, four
// cells per lane, TY output waves and two halo waves per workgroup marching x
// over L planes.  MODE 1 is the fused step's arithmetic, so it equals the
// oracle bit for bit; MODES 2 and 3 are the per-thread transposed kernel's
// (adjoint.hip) operation for operation, so the two forms agree bit for bit.
#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "march.hpp"
#include "row.hpp"

namespace cfd {

template <class T, int TY, bool PW, bool KF, int MODE>
__global__ void __launch_bounds__(64 * (TY + 2))
wave_row_aux_k(UniformGeo<T> g, const T* __restrict__ p, const T* __restrict__ u, const T* __restrict__ v,
               const T* __restrict__ w, T* __restrict__ po, T* __restrict__ uo, T* __restrict__ vo,
               T* __restrict__ wo, T K, const T* __restrict__ Kf, T dt, int L, int nyt, int order,
               const SensorList* __restrict__ sens, row::RowAux<T> aux) {
    using V4 = typename row::Row4<T>::type;
    __shared__ V4 sp[TY + 2][64];
    __shared__ V4 sv[TY + 2][64];
    // tile order as the fused row kernel's: bit 1 walks the tiles last to
    // first (alternate steps, so a step starts where the previous one ended),
    // bit 0 gives each XCD a contiguous run of tiles
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
    const bool inner = x0 >= 1 && x1 + 2 <= X;
#define CFD_AUX(XB, MR)                                                                                         \
    row::row_tile<T, TY, PW, KF, XB, false, MR, MODE>(g, p, u, v, w, po, uo, vo, wo, K, Kf, dt, x0, x1, y0, wv, \
                                                      false, sp, sv, sens, &aux, lin)
    if (row::rows_per_wave(g.Z) == 1) {
        if (inner) CFD_AUX(false, false);
        else CFD_AUX(true, false);
    } else {
        if (inner) CFD_AUX(false, true);
        else CFD_AUX(true, true);
    }
#undef CFD_AUX
}

namespace {

template <class T, int TY, bool PW, bool KF, int MODE>
void launch_aux(const UniformGeo<T>& g, const void* const in[4], void* const out[4], T K, const void* Kf, T dt,
                int L, int reverse, const SensorList* sens, const row::RowAux<T>& aux, hipStream_t s) {
    const int64_t rows = int64_t(TY) * row::rows_per_wave(g.Z);
    const int nyt = int((g.Y + rows - 1) / rows);
    const int64_t nwg = int64_t(nyt) * ((g.X + L - 1) / L);
    CFD_REQUIRE(nwg < (int64_t(1) << 31), "adjoint row grid too large");
    const int order = (nwg % 8 == 0 ? 1 : 0) | (reverse ? 2 : 0);
    wave_row_aux_k<T, TY, PW, KF, MODE><<<dim3(unsigned(nwg)), dim3(64, TY + 2), 0, s>>>(
        g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3], (T*)out[0], (T*)out[1], (T*)out[2],
        (T*)out[3], K, (const T*)Kf, dt, L, nyt, order, sens, aux);
    CFD_HIP(hipGetLastError());
}

template <class T, int TY, int MODE>
void launch_aux_ty(const UniformGeo<T>& g, const void* const in[4], void* const out[4], T K, const void* Kf, T dt,
                   int L, int rev, const SensorList* sens, const row::RowAux<T>& aux, hipStream_t s) {
    if constexpr (MODE == 2) {
        if (aux.gK) {
            if (g.vol_pow2) launch_aux<T, TY, true, false, 3>(g, in, out, K, nullptr, dt, L, rev, sens, aux, s);
            else launch_aux<T, TY, false, false, 3>(g, in, out, K, nullptr, dt, L, rev, sens, aux, s);
        } else {
            if (g.vol_pow2) launch_aux<T, TY, true, false, 2>(g, in, out, K, nullptr, dt, L, rev, sens, aux, s);
            else launch_aux<T, TY, false, false, 2>(g, in, out, K, nullptr, dt, L, rev, sens, aux, s);
        }
    } else {
        if (g.vol_pow2) {
            if (Kf) launch_aux<T, TY, true, true, MODE>(g, in, out, K, Kf, dt, L, rev, sens, aux, s);
            else launch_aux<T, TY, true, false, MODE>(g, in, out, K, Kf, dt, L, rev, sens, aux, s);
        } else {
            if (Kf) launch_aux<T, TY, false, true, MODE>(g, in, out, K, Kf, dt, L, rev, sens, aux, s);
            else launch_aux<T, TY, false, false, MODE>(g, in, out, K, Kf, dt, L, rev, sens, aux, s);
        }
    }
}

template <class T> const UniformGeo<T>& ugeo_of(const cfd_mesh_s* m) {
    if constexpr (sizeof(T) == 4) return m->ug32; else return m->ug64;
}

bool aligned4(const void* p, size_t elem) { return (reinterpret_cast<uintptr_t>(p) & (4 * elem - 1)) == 0; }

}  // namespace

// Output rows of the row tile for `m`'s adjoint shape
template <class T> int adj_row_tile_rows(const cfd_mesh_s* m) { return m->adj_ty; }

// Whether the row form takes this mesh and these fields (16-byte aligned rows
// in FP32, 32-byte in FP64).
template <class T>
bool adj_row_usable(const cfd_mesh_s* m, const void* const* fields, int n) {
    if (m->adj_form != 1 || !m->uniform || !wave_row_supported(m->Z)) return false;
    for (int i = 0; i < n; ++i)
        if (fields[i] && !aligned4(fields[i], sizeof(T))) return false;
    return true;
}

// Per-tile sensor masks for the row-form launches of `m`'s adjoint shape:
// tile = x chunk * (y tiles) + y tile, as the kernel numbers its tiles.
template <class T> std::vector<uint64_t> adj_row_tile_sensors(const cfd_mesh_s* m, const SensorList& sens) {
    const int ty = adj_row_tile_rows<T>(m);
    const int64_t rows = int64_t(ty) * row::rows_per_wave(m->Z);
    const int64_t nyt = (m->Y + rows - 1) / rows, nxc = (m->X + m->adj_L - 1) / m->adj_L;
    std::vector<uint64_t> mask(size_t(nyt * nxc), 0);
    for (int j = 0; j < sens.n; ++j)
        mask[size_t((sens.x[j] / m->adj_L) * nyt + sens.y[j] / rows)] |= uint64_t(1) << j;
    return mask;
}

template <class T>
void launch_wave_row_aux(const cfd_mesh_s* m, int mode, const void* const in[4], void* const out[4], double K,
                         const void* Kf, double dt, const SensorList* sens, const row::RowAux<T>& aux, int rev,
                         hipStream_t s) {
    const UniformGeo<T>& g = ugeo_of<T>(m);
    const int L = m->adj_L;
    CFD_REQUIRE(mode == 1 || (mode == 2 && Kf == nullptr), "transposed row step: constant K only");
    const int ty = adj_row_tile_rows<T>(m);
    if constexpr (sizeof(T) == 8) {
        // FP64 takes the 4-row tile only: the 8-row builds spill registers,
        // and these kernels read other lanes' registers (z neighbours, the
        // sensors), which a reload under a partial exec mask leaves stale
        CFD_REQUIRE(ty == 4, "FP64 adjoint row tiles have 4 output waves");
        if (mode == 1) launch_aux_ty<T, 4, 1>(g, in, out, T(K), Kf, T(dt), L, rev, sens, aux, s);
        else launch_aux_ty<T, 4, 2>(g, in, out, T(K), Kf, T(dt), L, rev, sens, aux, s);
    } else {
        switch (ty * 4 + mode) {
            case 4 * 4 + 1: launch_aux_ty<T, 4, 1>(g, in, out, T(K), Kf, T(dt), L, rev, sens, aux, s); break;
            case 4 * 4 + 2: launch_aux_ty<T, 4, 2>(g, in, out, T(K), Kf, T(dt), L, rev, sens, aux, s); break;
            case 8 * 4 + 1: launch_aux_ty<T, 8, 1>(g, in, out, T(K), Kf, T(dt), L, rev, sens, aux, s); break;
            case 8 * 4 + 2: launch_aux_ty<T, 8, 2>(g, in, out, T(K), Kf, T(dt), L, rev, sens, aux, s); break;
            default: throw Error(CFD_E_INVALID, "adjoint row tiles have 4 or 8 output waves");
        }
    }
}

#define CFD_AUX_INSTANTIATE(T)                                                                                   \
    template bool adj_row_usable<T>(const cfd_mesh_s*, const void* const*, int);                                \
    template std::vector<uint64_t> adj_row_tile_sensors<T>(const cfd_mesh_s*, const SensorList&);              \
    template void launch_wave_row_aux<T>(const cfd_mesh_s*, int, const void* const[4], void* const[4], double,  \
                                         const void*, double, const SensorList*, const row::RowAux<T>&, int,     \
                                         hipStream_t);
CFD_AUX_INSTANTIATE(float)
CFD_AUX_INSTANTIATE(double)

}  // namespace cfd
