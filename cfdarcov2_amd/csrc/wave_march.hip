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
    if (g.vol_pow2) wave_march_tile<T, TYT, R, NT, 1, true>(g, p, u, v, w, po, uo, vo, wo, K, Kf, dt, L, zt, yt, xc, lds);
    else wave_march_tile<T, TYT, R, NT, 1, false>(g, p, u, v, w, po, uo, vo, wo, K, Kf, dt, L, zt, yt, xc, lds);
}

// Register form (variant 3): every wave is independent and owns R consecutive
// rows x 64 z-columns x L planes.  Rows live in the thread's registers, so the
// y-neighbours of the velocity and pressure stencils are register reads; the
// z-neighbours come from the adjacent lanes (cross-lane shuffles), with lanes 0
// and 63 loading the one-column halo themselves; x uses the same register
// queue as the LDS form.  No LDS and no barriers.  Same arithmetic, same bits.
template <class T, int R>
__global__ void __launch_bounds__(256) wave_reg_k(UniformGeo<T> g, const T* __restrict__ p, const T* __restrict__ u,
                                                  const T* __restrict__ v, const T* __restrict__ w,
                                                  T* __restrict__ po, T* __restrict__ uo, T* __restrict__ vo,
                                                  T* __restrict__ wo, T K, const T* __restrict__ Kf, T dt, int L,
                                                  int nzt, int nyb, int64_t nwaves, int remap) {
    const int lane = threadIdx.x;
    int64_t blk = blockIdx.x;
    if (remap) blk = (blk % 8) * (int64_t(gridDim.x) / 8) + blk / 8;
    // the wave index is uniform across the wave: keep it (and everything derived
    // from it: tile coordinates, plane base pointers) in scalar registers
    const int64_t gw = __builtin_amdgcn_readfirstlane(int(blk * 4 + threadIdx.y));
    if (gw >= nwaves) return;  // whole wave; waves never synchronise
    const int yb = int(gw % nyb);
    const int zt = int((gw / nyb) % nzt);
    const int64_t xc = gw / (int64_t(nyb) * nzt);
    const int64_t X = g.X, SX = g.Y * g.Z;
    const int Y = int(g.Y), Z = int(g.Z);  // a plane holds fewer than 2^31 cells
    const T Ax = g.area[0], Ay = g.area[1], Az = g.area[2], V = g.vol;
    const int z = zt * 64 + lane;
    const bool zval = z < Z;
    const int zc = zval ? z : Z - 1;  // in-bounds column for the idle lanes' loads
    const int y0 = yb * R;
    const int64_t x0 = xc * L;
    const int64_t x1 = (x0 + L < X) ? x0 + L : X;
    // z-halo lanes: lane 0 reads column z0-1 (and z0-2), lane 63 column z0+64 (and z0+65)
    const bool hlo = lane == 0 && zt * 64 - 1 >= 0;
    const bool hhi = lane == 63 && zt * 64 + 64 < Z;
    const int hz1 = hlo ? zt * 64 - 1 : (hhi ? zt * 64 + 64 : zc);
    const int hz2 = hlo ? zt * 64 - 2 : (hhi ? (zt * 64 + 65 < Z ? zt * 64 + 65 : zt * 64 + 64) : zc);
    auto rowy = [&](int yy) { return yy < 0 ? 0 : (yy >= Y ? Y - 1 : yy); };
    // in-plane offsets: own rows, halo p rows (-2, -1, R, R+1), v rows (-1..R)
    unsigned col[R], hcol1[R], hcol2[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        col[r] = unsigned(rowy(y0 + r) * Z + zc);
        hcol1[r] = unsigned(rowy(y0 + r) * Z + hz1);
        hcol2[r] = unsigned(rowy(y0 + r) * Z + hz2);
    }
    unsigned hrow[4];
#pragma unroll
    for (int h = 0; h < 4; ++h) hrow[h] = unsigned(rowy(h < 2 ? y0 - 2 + h : y0 + R + (h - 2)) * Z + zc);
    const unsigned vlo = unsigned(rowy(y0 - 1) * Z + zc), vhi = unsigned(rowy(y0 + R) * Z + zc);

    T P0[R], Pp1[R], Um1[R], U0[R];
    {  // warm-up: p(x0-1..x0+1), u'(x0-1), u'(x0)
        const T* p0 = p + x0 * SX;
        const T* pm1 = x0 >= 1 ? p0 - SX : p0;
        const T* pm2 = x0 >= 2 ? p0 - 2 * SX : pm1;
        const T* pp1 = x0 + 1 < X ? p0 + SX : p0;
        const T* u0 = u + x0 * SX;
        const T* um1 = x0 >= 1 ? u0 - SX : u0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            P0[r] = p0[col[r]];
            const T Pm1 = pm1[col[r]];
            Pp1[r] = pp1[col[r]];
            Um1[r] = T(0);
            if (x0 >= 1) Um1[r] = mvel(pm2[col[r]], Pm1, P0[r], um1[col[r]], Ax, V, dt);
            U0[r] = mvel(Pm1, P0[r], Pp1[r], u0[col[r]], Ax, V, dt);
        }
    }
#pragma unroll 1
    for (int64_t x = x0; x < x1; ++x) {
        const int64_t base = x * SX;
        const T* px = p + base;
        const T* vx = v + base;
        const T* wx = w + base;
        const T* ux1 = u + (x + 1 < X ? base + SX : base);
        const T* px2 = p + (x + 2 < X ? base + 2 * SX : base);
        // p rows y0-2, y0-1 and y0+R, y0+R+1 of plane x (walls: the row itself)
        T pr[R + 4];
        pr[0] = px[hrow[0]];
        pr[1] = px[hrow[1]];
        pr[R + 2] = px[hrow[2]];
        pr[R + 3] = px[hrow[3]];
#pragma unroll
        for (int r = 0; r < R; ++r) pr[2 + r] = P0[r];
        // v rows y0-1 .. y0+R
        T vr[R + 2];
        vr[0] = vx[vlo];
        vr[R + 1] = vx[vhi];
#pragma unroll
        for (int r = 0; r < R; ++r) vr[r + 1] = vx[col[r]];
        T Up1[R], Pp2[R], wv[R], whl[R], pz1[R], pz2[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Pp2[r] = x + 2 < X ? px2[col[r]] : Pp1[r];
            Up1[r] = x + 1 < X ? mvel(P0[r], Pp1[r], Pp2[r], ux1[col[r]], Ax, V, dt) : T(0);
            wv[r] = wx[col[r]];
            pz1[r] = pz2[r] = whl[r] = T(0);
            if (hlo || hhi) {
                pz1[r] = px[hcol1[r]];
                pz2[r] = px[hcol2[r]];
                whl[r] = wx[hcol1[r]];
            }
        }
        // v' on rows -1..R (only read where the face to the tile is not a wall)
        T vn[R + 2];
#pragma unroll
        for (int r = 0; r < R + 2; ++r) {
            const int yy = y0 - 1 + r;
            const T pl = (yy - 1 >= 0) ? pr[r] : pr[r + 1];
            const T ph_ = (yy + 1 < Y) ? pr[r + 2] : pr[r + 1];
            vn[r] = mvel(pl, pr[r + 1], ph_, vr[r], Ay, V, dt);
        }
        T* pox = po + base;
        T* uox = uo + base;
        T* vox = vo + base;
        T* wox = wo + base;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            // w' on the row, neighbours in z from the adjacent lanes
            const T pc = P0[r];
            T pzl = __shfl_up(pc, 1, 64);
            T pzh = __shfl_down(pc, 1, 64);
            if (lane == 0) pzl = pz1[r];
            if (lane == 63) pzh = pz1[r];
            if (z == 0) pzl = pc;
            if (z == Z - 1) pzh = pc;
            const T wn = mvel(pzl, pc, pzh, wv[r], Az, V, dt);
            // halo w' at z0-1 (lane 0) / z0+64 (lane 63)
            T whh = T(0);
            if (hlo) whh = mvel(pz2[r], pz1[r], pc, whl[r], Az, V, dt);
            if (hhi) whh = mvel(pc, pz1[r], pz2[r], whl[r], Az, V, dt);
            T wl = __shfl_up(wn, 1, 64);
            T wh = __shfl_down(wn, 1, 64);
            if (lane == 0) wl = whh;
            if (lane == 63) wh = whh;
            const int y = y0 + r;
            if (zval && y < Y) {
                const T fx_lo = (x > 0) ? mlin(U0[r], Um1[r]) : T(0);
                const T fx_hi = (x < X - 1) ? mlin(U0[r], Up1[r]) : T(0);
                const T fy_lo = (y == 0) ? T(0) : mlin(vn[r + 1], vn[r]);
                const T fy_hi = (y == Y - 1) ? T(0) : mlin(vn[r + 1], vn[r + 2]);
                const T fz_lo = (z > 0) ? mlin(wn, wl) : T(0);
                const T fz_hi = (z < Z - 1) ? mlin(wn, wh) : T(0);
                T S = mggp(fx_lo, fx_hi, Ax, V);
                S = S + mggp(fy_lo, fy_hi, Ay, V);
                S = S + mggp(fz_lo, fz_hi, Az, V);
                const unsigned o = unsigned(y * Z + z);
                const T k = Kf ? Kf[base + o] : K;
                pox[o] = dt * (k * S) + pc;
                uox[o] = U0[r];
                vox[o] = vn[r + 1];
                wox[o] = wn;
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            P0[r] = Pp1[r];
            Pp1[r] = Pp2[r];
            Um1[r] = U0[r];
            U0[r] = Up1[r];
        }
    }
}

template <class T, int R>
static void launch_reg(const UniformGeo<T>& g, void* const* in, void* const* out, double K, const void* Kf, double dt,
                       int L, int remap, hipStream_t s) {
    const int nzt = int((g.Z + 63) / 64);
    const int nyb = int((g.Y + R - 1) / R);
    const int64_t nxc = (g.X + L - 1) / L;
    const int64_t nwaves = int64_t(nzt) * nyb * nxc;
    const int64_t nblk = (nwaves + 3) / 4;
    CFD_REQUIRE(nwaves < (int64_t(1) << 31), "register march grid too large");
    const int rm = (remap && nblk % 8 == 0) ? 1 : 0;
    wave_reg_k<T, R><<<dim3(unsigned(nblk)), dim3(64, 4), 0, s>>>(
        g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3], (T*)out[0], (T*)out[1], (T*)out[2],
        (T*)out[3], T(K), (const T*)Kf, T(dt), L, nzt, nyb, nwaves, rm);
    CFD_HIP(hipGetLastError());
}

// The variant-4 tile at full occupancy: at most 64 VGPRs, so eight waves fit
// each SIMD (variant 5).
template <class T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
wave_march_occ_k(UniformGeo<T> g, const T* __restrict__ p, const T* __restrict__ u, const T* __restrict__ v,
                 const T* __restrict__ w, T* __restrict__ po, T* __restrict__ uo, T* __restrict__ vo,
                 T* __restrict__ wo, T K, const T* __restrict__ Kf, T dt, int L, int nzt, int nyt, int remap) {
    __shared__ MarchLds<T, 4, 1> lds;
    const int64_t nwg = gridDim.x;
    int64_t lin_id = blockIdx.x;
    if (remap) lin_id = (lin_id % 8) * (nwg / 8) + lin_id / 8;
    const int64_t zt = lin_id % nzt;
    const int64_t yt = (lin_id / nzt) % nyt;
    const int64_t xc = lin_id / (int64_t(nzt) * nyt);
    if (g.vol_pow2) wave_march_tile<T, 4, 1, true, 1, true>(g, p, u, v, w, po, uo, vo, wo, K, Kf, dt, L, zt, yt, xc, lds);
    else wave_march_tile<T, 4, 1, true, 1, false>(g, p, u, v, w, po, uo, vo, wo, K, Kf, dt, L, zt, yt, xc, lds);
}

// Variant 6: variant-4 tiles (4 rows) with the loads of two planes in flight.
template <class T>
__global__ void __launch_bounds__(256) wave_march_pf2_k(UniformGeo<T> g, const T* __restrict__ p,
                                                        const T* __restrict__ u, const T* __restrict__ v,
                                                        const T* __restrict__ w, T* __restrict__ po,
                                                        T* __restrict__ uo, T* __restrict__ vo, T* __restrict__ wo,
                                                        T K, const T* __restrict__ Kf, T dt, int L, int nzt, int nyt,
                                                        int remap) {
    __shared__ MarchLds<T, 4, 1> lds;
    const int64_t nwg = gridDim.x;
    int64_t lin_id = blockIdx.x;
    if (remap) lin_id = (lin_id % 8) * (nwg / 8) + lin_id / 8;
    const int64_t zt = lin_id % nzt;
    const int64_t yt = (lin_id / nzt) % nyt;
    const int64_t xc = lin_id / (int64_t(nzt) * nyt);
    if (g.vol_pow2) wave_march_tile<T, 4, 1, true, 2, true>(g, p, u, v, w, po, uo, vo, wo, K, Kf, dt, L, zt, yt, xc, lds);
    else wave_march_tile<T, 4, 1, true, 2, false>(g, p, u, v, w, po, uo, vo, wo, K, Kf, dt, L, zt, yt, xc, lds);
}

template <class T>
static void launch_occ(const UniformGeo<T>& g, void* const* in, void* const* out, double K, const void* Kf, double dt,
                       int L, int remap, hipStream_t s, bool pf2 = false) {
    const int nzt = int((g.Z + 63) / 64);
    const int nyt = int((g.Y + 3) / 4);
    const int nxc = int((g.X + L - 1) / L);
    const int64_t nwg = int64_t(nzt) * nyt * nxc;
    CFD_REQUIRE(nwg < (int64_t(1) << 31), "march grid too large");
    const int rm = (remap && nwg % 8 == 0) ? 1 : 0;
    if (pf2)
        wave_march_pf2_k<T><<<dim3(unsigned(nwg)), dim3(64, 4), 0, s>>>(
            g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3], (T*)out[0], (T*)out[1],
            (T*)out[2], (T*)out[3], T(K), (const T*)Kf, T(dt), L, nzt, nyt, rm);
    else
        wave_march_occ_k<T><<<dim3(unsigned(nwg)), dim3(64, 4), 0, s>>>(
            g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3], (T*)out[0], (T*)out[1],
            (T*)out[2], (T*)out[3], T(K), (const T*)Kf, T(dt), L, nzt, nyt, rm);
    CFD_HIP(hipGetLastError());
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
// 8 = 4x2, 16 = 4x4, 32 = 8x4.  Variant 2: one row per thread, tile_rows
// threads in y (4, 8 or 16: 256 to 1024 threads per workgroup).
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
        if (variant >= 7) {  // row form; misaligned fields take the variant-4 tile
            // remap bit 1: odd steps walk the tiles in reverse.  Cache policy:
            // 7 non-temporal stores, 8 plain, 9 non-temporal loads, 10 both
            const int order = (remap & 1) | ((remap & 2) && (k & 1) ? 2 : 0);
            const int cache = variant == 7 ? 1 : variant == 8 ? 0 : variant == 9 ? 2 : 3;
            if (!launch_wave_row<T>(g, in, out, K, Kf, dt, ty, L, order, cache, s))
                launch_one<T, 4, 1, true>(g, in, out, K, Kf, dt, L, remap & 1, s);
            continue;
        }
        if (variant == 5 || variant == 6) {
            CFD_REQUIRE(ty == 4, "variant 5 and 6 tiles have 4 rows");
            launch_occ<T>(g, in, out, K, Kf, dt, L, remap, s, variant == 6);
            continue;
        }
        if (variant == 4) {  // variant 1 tiles with non-temporal output stores
            switch (ty) {
                case 4: launch_one<T, 4, 1, true>(g, in, out, K, Kf, dt, L, remap, s); break;
                case 8: launch_one<T, 4, 2, true>(g, in, out, K, Kf, dt, L, remap, s); break;
                default: throw Error(CFD_E_INVALID, "variant 4 tiles have 4 or 8 rows");
            }
            continue;
        }
        if (variant == 3) {
            switch (ty) {
                case 2: launch_reg<T, 2>(g, in, out, K, Kf, dt, L, remap, s); break;
                case 4: launch_reg<T, 4>(g, in, out, K, Kf, dt, L, remap, s); break;
                case 8: launch_reg<T, 8>(g, in, out, K, Kf, dt, L, remap, s); break;
                default: throw Error(CFD_E_INVALID, "register march tiles have 2, 4 or 8 rows");
            }
            continue;
        }
        if (variant == 2) {
            switch (ty) {
                case 4: launch_one<T, 4, 1>(g, in, out, K, Kf, dt, L, remap, s); break;
                case 8: launch_one<T, 8, 1>(g, in, out, K, Kf, dt, L, remap, s); break;
                case 16: launch_one<T, 16, 1>(g, in, out, K, Kf, dt, L, remap, s); break;
                default: throw Error(CFD_E_INVALID, "one-row march tiles have 4, 8 or 16 rows");
            }
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
