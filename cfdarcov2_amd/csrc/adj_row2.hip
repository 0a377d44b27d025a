// Two-step row form of the adjoint's per-step kernels (adjoint form 2; no
// reference counterpart): one launch advances two steps, so every field
// crosses HBM once per two steps.
//   MODE 1, two forward steps recording p' at the sensor cells after each;
//   MODE 2, two transposed steps lam_t -> lam_{t-2} with constant K, adding
//           the trace gradients at the sensors after each.
// Both run the two-step tiles of the fused step (row2.hpp: tile2 for z rows of
// 256 cells, tile2m for several rows per wave) with an epilogue selected at
// compile time.  Each value is computed by the operation sequence of the
// single-step row tile (row.hpp), so one launch equals two single-step
// launches bit for bit.  Runs with a K gradient keep the single-step row form
// (its forward recomputation stores a divergence field per step).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"
#include "march.hpp"
#include "row.hpp"
#include "row2.hpp"

namespace cfd {

template <class T, int TY, bool PW, int MODE>
__global__ void __launch_bounds__(64 * (TY + 6))
wave_row2_aux_k(UniformGeo<T> g, const T* __restrict__ p, const T* __restrict__ u, const T* __restrict__ v,
                const T* __restrict__ w, T* __restrict__ po, T* __restrict__ uo, T* __restrict__ vo,
                T* __restrict__ wo, T K, T dt, int L, int nyt, int order, const SensorList* __restrict__ sens,
                row2::Aux2<T> aux) {
    using V4 = row2::V4T<T>;
    __shared__ V4 P1[TY + 6][64];
    __shared__ V4 V1[TY + 6][64];
    __shared__ V4 P2[TY + 6][64];
    __shared__ V4 V2[TY + 6][64];
    // tile order as the fused two-step kernel's: bit 1 walks the tiles last to
    // first, bit 0 gives each XCD a contiguous run of tiles
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
    if (x0 >= 4 && x1 + 4 <= X)
        row2::tile2<T, TY, PW, false, true, MODE>(g, p, u, v, w, po, uo, vo, wo, K, dt, x0, x1, y0, j, false, P1, V1,
                                                  P2, V2, sens, &aux);
    else
        row2::tile2<T, TY, PW, true, true, MODE>(g, p, u, v, w, po, uo, vo, wo, K, dt, x0, x1, y0, j, false, P1, V1,
                                                 P2, V2, sens, &aux);
}

template <class T, int NW, bool PW, int MODE>
__global__ void __launch_bounds__(64 * NW)
wave_row2m_aux_k(UniformGeo<T> g, const T* __restrict__ p, const T* __restrict__ u, const T* __restrict__ v,
                 const T* __restrict__ w, T* __restrict__ po, T* __restrict__ uo, T* __restrict__ vo,
                 T* __restrict__ wo, T K, T dt, int L, int nyt, int order, const SensorList* __restrict__ sens,
                 row2::Aux2<T> aux) {
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
        row2::tile2m<T, NW, PW, false, true, MODE>(g, p, u, v, w, po, uo, vo, wo, K, dt, x0, x1, y0, wv, false, P1,
                                                   V1, P2, V2, sens, &aux);
    else
        row2::tile2m<T, NW, PW, true, true, MODE>(g, p, u, v, w, po, uo, vo, wo, K, dt, x0, x1, y0, wv, false, P1,
                                                  V1, P2, V2, sens, &aux);
}

namespace {

bool aligned4_2(const void* p, size_t elem) { return (reinterpret_cast<uintptr_t>(p) & (4 * elem - 1)) == 0; }

// grid order: XCD-contiguous when the grid divides by 8, reversed on
// alternate launches
int grid_order(int64_t nwg, int reverse) { return (nwg % 8 == 0 ? 1 : 0) | (reverse ? 2 : 0); }

template <class T, int TY, bool PW, int MODE>
void launch_full(const UniformGeo<T>& g, const void* const in[4], void* const out[4], T K, T dt, int L, int rev,
                 const SensorList* sens, const row2::Aux2<T>& aux, hipStream_t s) {
    const int nyt = int((g.Y + TY - 1) / TY);
    const int64_t nwg = int64_t(nyt) * ((g.X + L - 1) / L);
    CFD_REQUIRE(nwg < (int64_t(1) << 31), "adjoint two-step grid too large");
    wave_row2_aux_k<T, TY, PW, MODE><<<dim3(unsigned(nwg)), dim3(64, TY + 6), 0, s>>>(
        g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3], (T*)out[0], (T*)out[1], (T*)out[2],
        (T*)out[3], K, dt, L, nyt, grid_order(nwg, rev), sens, aux);
    CFD_HIP(hipGetLastError());
}

template <class T, int NW, bool PW, int MODE>
void launch_multi(const UniformGeo<T>& g, const void* const in[4], void* const out[4], T K, T dt, int L, int rev,
                  const SensorList* sens, const row2::Aux2<T>& aux, hipStream_t s) {
    const int64_t rows = int64_t(NW) * row::rows_per_wave(g.Z) - 8;
    CFD_REQUIRE(rows >= 1, "two-step tile without output rows");
    const int nyt = int((g.Y + rows - 1) / rows);
    const int64_t nwg = int64_t(nyt) * ((g.X + L - 1) / L);
    CFD_REQUIRE(nwg < (int64_t(1) << 31), "adjoint two-step grid too large");
    wave_row2m_aux_k<T, NW, PW, MODE><<<dim3(unsigned(nwg)), dim3(64, NW), 0, s>>>(
        g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3], (T*)out[0], (T*)out[1], (T*)out[2],
        (T*)out[3], K, dt, L, nyt, grid_order(nwg, rev), sens, aux);
    CFD_HIP(hipGetLastError());
}

// tile_rows -> the tile: one z row per wave (z extents above 128) takes TY = 4
// output waves (the 8- and 10-row tiles of the epilogue spill registers at 14
// and 16 waves); several rows per wave take NW = 8, 10 or 12 waves for 4, 8
// or 10
template <bool PW, int MODE>
void launch_pw(const UniformGeo<float>& g, int ty, const void* const in[4], void* const out[4], float K, float dt,
               int L, int rev, const SensorList* sens, const row2::Aux2<float>& aux, hipStream_t s) {
    if (row::rows_per_wave(g.Z) == 1) {
        launch_full<float, 4, PW, MODE>(g, in, out, K, dt, L, rev, sens, aux, s);
        return;
    }
    switch (ty) {
        case 4: launch_multi<float, 8, PW, MODE>(g, in, out, K, dt, L, rev, sens, aux, s); break;
        case 8: launch_multi<float, 10, PW, MODE>(g, in, out, K, dt, L, rev, sens, aux, s); break;
        default: launch_multi<float, 12, PW, MODE>(g, in, out, K, dt, L, rev, sens, aux, s); break;
    }
}

}  // namespace

// Whether the two-step form takes this mesh and these fields: adjoint form 2,
// an FP32, exactly uniform mesh whose z extent the row form takes, 16-byte
// aligned rows.
template <class T>
bool adj_row2_usable(const cfd_mesh_s* m, const void* const* fields, int n) {
    if (sizeof(T) != 4 || m->adj_form != 2 || !m->uniform || !wave_row_supported(m->Z)) return false;
    for (int i = 0; i < n; ++i)
        if (fields[i] && !aligned4_2(fields[i], sizeof(T))) return false;
    return true;
}

template <class T>
void launch_wave_row2_aux(const cfd_mesh_s* m, int mode, const void* const in[4], void* const out[4], double K,
                          double dt, const SensorList* sens, const row2::Aux2<T>& aux, int rev, hipStream_t s) {
    if constexpr (sizeof(T) == 8) {
        // the FP64 transposed tiles spill registers at every tile size
        throw Error(CFD_E_UNSUPPORTED, "the two-step adjoint form is FP32 only");
    } else {
        const UniformGeo<float>& g = m->ug32;
        const int ty = m->adj_ty, L = m->adj_L;
        CFD_REQUIRE(ty == 4 || (row::rows_per_wave(g.Z) > 1 && (ty == 8 || ty == 10)),
                    "two-step adjoint tiles have 4 rows, or 4, 8 or 10 rows for z extents up to 128");
        CFD_REQUIRE(mode == 1 || mode == 2, "two-step adjoint mode must be 1 or 2");
        if (g.vol_pow2) {
            if (mode == 1) launch_pw<true, 1>(g, ty, in, out, float(K), float(dt), L, rev, sens, aux, s);
            else launch_pw<true, 2>(g, ty, in, out, float(K), float(dt), L, rev, sens, aux, s);
        } else {
            if (mode == 1) launch_pw<false, 1>(g, ty, in, out, float(K), float(dt), L, rev, sens, aux, s);
            else launch_pw<false, 2>(g, ty, in, out, float(K), float(dt), L, rev, sens, aux, s);
        }
    }
}

#define CFD_AUX2_INSTANTIATE(T)                                                                                  \
    template bool adj_row2_usable<T>(const cfd_mesh_s*, const void* const*, int);                               \
    template void launch_wave_row2_aux<T>(const cfd_mesh_s*, int, const void* const[4], void* const[4], double, \
                                          double, const SensorList*, const row2::Aux2<T>&, int, hipStream_t);
CFD_AUX2_INSTANTIATE(float)
CFD_AUX2_INSTANTIATE(double)

}  // namespace cfd
