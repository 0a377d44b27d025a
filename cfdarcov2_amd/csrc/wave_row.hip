// Row form of the fused acoustic step (variant 7) on an exactly uniform mesh
// whose z extent is a multiple of 4 and at most 256.
//
// One wave holds a whole z row of a plane, four consecutive cells per lane, so
// every load and store is a 16-byte vector access (32 bytes in FP64) with the
// address in scalar registers plus one fixed lane offset.  The z-neighbours of
// a lane's cells are its own registers or the adjacent lane's, reached with
// DPP wave shifts; at the z walls the shift's out-of-range lane keeps its own
// value (the reference's neighbour = self rule) and the wall flux is zero, so
// no z halo is loaded or computed.
//
// A workgroup holds TY output waves (rows y0 .. y0+TY-1) and two halo waves
// (rows y0-1 and y0+TY).  It marches along x over a chunk of L planes; the
// output waves keep p(x), p(x+1) and u'(x-1), u'(x) per lane in registers,
// as the 2.5D tile form does.  Per plane:
//   1. the output waves compute u'(x+1) and put their p row into the LDS plane;
//      the halo waves put the (clamped) halo p rows there;
//   2. barrier; the output waves compute v' (y-neighbours from LDS) and w'
//      (z-neighbours from registers), the halo waves the halo rows of v';
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

namespace row {

template <class T, int TY, bool PW, bool KF, bool XB, bool NTL>
__device__ inline void row_tile(const UniformGeo<T>& g, const T* __restrict__ p, const T* __restrict__ u,
                                const T* __restrict__ v, const T* __restrict__ w, T* __restrict__ po,
                                T* __restrict__ uo, T* __restrict__ vo, T* __restrict__ wo, T K,
                                const T* __restrict__ Kf, T dt, int64_t x0, int64_t x1, int y0, int wv, bool nt,
                                typename Row4<T>::type (*sp)[64], typename Row4<T>::type (*sv)[64]) {
    using V4 = typename Row4<T>::type;
    const int64_t X = g.X;
    const int Y = int(g.Y), Z = int(g.Z);
    const int64_t SX = int64_t(Y) * Z;
    const T Ax = g.area[0], Ay = g.area[1], Az = g.area[2];
    const VolDiv<T, PW> Vd{g.vol, g.inv_vol};
    const int lane = threadIdx.x;
    const int nl = Z / 4;
    const bool lv = lane < nl;

    // ---- the wave's role (uniform across the wave) ---------------------------
    // out: a real output row; prow: the row whose p this wave puts into LDS row
    // `lrow` (-1: none); halo_v: a halo wave that also computes v' of its row
    const bool is_out = wv < TY;
    const bool lower = wv == TY;
    const int y = is_out ? y0 + wv : (lower ? y0 - 1 : y0 + TY);
    const bool out = is_out && y < Y;
    const int lrow = is_out ? wv + 1 : (lower ? 0 : TY + 1);
    const int prow = (y >= 0 && y < Y) ? y : (y == -1 ? 0 : (y == Y ? Y - 1 : -1));
    const bool pcopy = !out && prow >= 0;        // halo waves and the row just past Y
    const bool halo_v = !is_out && y >= 0 && y < Y;
    // halo: the outer p row (clamped) and the LDS row of the inner neighbour
    const int orow = lower ? (y - 1 >= 0 ? y - 1 : 0) : (y + 1 < Y ? y + 1 : Y - 1);
    const int irow = lower ? 1 : TY;
    const bool first = y == 0, last = y == Y - 1;

    // per-plane row bases (element offsets of the row start, uniform)
    const int64_t rb = int64_t(prow >= 0 ? prow : 0) * Z;
    const int64_t ob = int64_t(orow) * Z;

    // Per-plane loads, as (field, plane offset, row) slots fixed per wave:
    //   output rows: a = p(xx+2), b = u(xx+1), c = v(xx), d = w(xx), k = K(xx);
    //   halo rows:   a = p(xx) of the row, b = p(xx) of the outer row, c = v(xx);
    //   the row just past Y: a = p(xx) of row Y-1.
    const T* const fa = p;
    const T* const fb = out ? u : p;
    const int da = out ? 2 : 0, db = out ? 1 : 0;
    const int64_t rbb = out ? rb : ob;
    const bool ea = out || pcopy, eb = out || halo_v, ec = out || halo_v, ed = out;
    struct Fetch {
        V4 a, b, c, d, k;
    };
    const V4 zero4 = {T(0), T(0), T(0), T(0)};
    // NTL: non-temporal loads (streamed past the caches' retention, which then
    // keep what this step writes for the next step to read)
    auto ld = [&](const T* base, int64_t off) -> V4 {
        const V4* a = reinterpret_cast<const V4*>(base + off) + lane;
        if constexpr (NTL) return __builtin_nontemporal_load(a);
        else return *a;
    };
    auto fetch = [&](int64_t xx, Fetch& f) {
        if (!lv) return;
        const int64_t pb = xx * SX;
        if (ea && (!XB || xx + da < X)) f.a = ld(fa, pb + da * SX + rb);
        if (eb && (!XB || xx + db < X)) f.b = ld(fb, pb + db * SX + rbb);
        if (ec) f.c = ld(v, pb + rb);
        if (ed) f.d = ld(w, pb + rb);
        if constexpr (KF) {
            if (ed) f.k = ld(Kf, pb + rb);
        }
    };

    // ---- warm-up of the output rows' x queue ---------------------------------
    V4 P0 = zero4, Pp1 = zero4, Um1 = zero4, U0 = zero4;
    if (out && lv) {
        V4 Pm1 = zero4;
        P0 = ld(p, x0 * SX + rb);
        if (x0 >= 1) Pm1 = ld(p, (x0 - 1) * SX + rb);
        if (x0 + 1 < X) Pp1 = ld(p, (x0 + 1) * SX + rb);
        if (x0 >= 1) {
            const V4 Pm2 = (x0 >= 2) ? ld(p, (x0 - 2) * SX + rb) : Pm1;
            const V4 um = ld(u, (x0 - 1) * SX + rb);
#pragma unroll
            for (int i = 0; i < 4; ++i) Um1[i] = mvel(Pm2[i], Pm1[i], P0[i], um[i], Ax, Vd, dt);
        }
        const V4 uc = ld(u, x0 * SX + rb);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            U0[i] = mvel(x0 >= 1 ? Pm1[i] : P0[i], P0[i], x0 + 1 < X ? Pp1[i] : P0[i], uc[i], Ax, Vd, dt);
    }

    Fetch nx{zero4, zero4, zero4, zero4, zero4};
    fetch(x0, nx);

    for (int64_t x = x0; x < x1; ++x) {
        const Fetch cur = nx;
        if (x + 1 < x1) fetch(x + 1, nx);

        // 1. u'(x+1); p rows into LDS
        V4 Up1 = zero4;
        if (out) {
            if (!XB || x + 1 < X) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    Up1[i] = mvel(P0[i], Pp1[i], (!XB || x + 2 < X) ? cur.a[i] : Pp1[i], cur.b[i], Ax, Vd, dt);
            }
            sp[lrow][lane] = P0;
        } else if (pcopy) {
            sp[lrow][lane] = cur.a;
        }
        __syncthreads();

        // 2. v' and w' of the plane; halo rows of v'
        V4 vn = zero4, wn = zero4;
        if (out) {
            const V4 pyl = sp[lrow - 1][lane];
            const V4 pyh = sp[lrow + 1][lane];
            V4 pzl, pzh;
            pzl[0] = lane_shift<kWaveShr1>(P0[0], P0[3]);
            pzl[1] = P0[0];
            pzl[2] = P0[1];
            pzl[3] = P0[2];
            pzh[0] = P0[1];
            pzh[1] = P0[2];
            pzh[2] = P0[3];
            pzh[3] = lane_shift<kWaveShl1>(P0[3], P0[0]);
            if (lane == 0) pzl[0] = P0[0];       // z walls: neighbour = self
            if (lane == nl - 1) pzh[3] = P0[3];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                vn[i] = mvel(pyl[i], P0[i], pyh[i], cur.c[i], Ay, Vd, dt);
                wn[i] = mvel(pzl[i], P0[i], pzh[i], cur.d[i], Az, Vd, dt);
            }
            sv[lrow][lane] = vn;
        } else if (halo_v) {
            const V4 in = sp[irow][lane];
            const V4 pl = lower ? cur.b : in;
            const V4 ph = lower ? in : cur.b;
            V4 vh;
#pragma unroll
            for (int i = 0; i < 4; ++i) vh[i] = mvel(pl[i], cur.a[i], ph[i], cur.c[i], Ay, Vd, dt);
            sv[lrow][lane] = vh;
        }
        __syncthreads();

        // 3. pressure update and the four outputs
        if (out) {
            const V4 vyl = sv[lrow - 1][lane];
            const V4 vyh = sv[lrow + 1][lane];
            V4 wzl, wzh;
            wzl[0] = lane_shift<kWaveShr1>(wn[0], wn[3]);
            wzl[1] = wn[0];
            wzl[2] = wn[1];
            wzl[3] = wn[2];
            wzh[0] = wn[1];
            wzh[1] = wn[2];
            wzh[2] = wn[3];
            wzh[3] = lane_shift<kWaveShl1>(wn[3], wn[0]);
            V4 pn;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const T fx_lo = (XB && x == 0) ? T(0) : mlin(U0[i], Um1[i]);
                const T fx_hi = (XB && x == X - 1) ? T(0) : mlin(U0[i], Up1[i]);
                const T fy_lo = first ? T(0) : mlin(vn[i], vyl[i]);
                const T fy_hi = last ? T(0) : mlin(vn[i], vyh[i]);
                const T fz_lo = (i == 0 && lane == 0) ? T(0) : mlin(wn[i], wzl[i]);
                const T fz_hi = (i == 3 && lane == nl - 1) ? T(0) : mlin(wn[i], wzh[i]);
                T S = mggp(fx_lo, fx_hi, Ax, Vd);
                S = S + mggp(fy_lo, fy_hi, Ay, Vd);
                S = S + mggp(fz_lo, fz_hi, Az, Vd);
                const T k = KF ? cur.k[i] : K;
                pn[i] = dt * (k * S) + P0[i];
            }
            if (lv) {
                const int64_t o = x * SX + rb;
                put(reinterpret_cast<V4*>(po + o) + lane, pn, nt);
                put(reinterpret_cast<V4*>(uo + o) + lane, U0, nt);
                put(reinterpret_cast<V4*>(vo + o) + lane, vn, nt);
                put(reinterpret_cast<V4*>(wo + o) + lane, wn, nt);
            }
            P0 = Pp1;
            Pp1 = cur.a;
            Um1 = U0;
            U0 = Up1;
        }
    }
}

}  // namespace row

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
    const int y0 = yt * TY;
    // chunks away from the x walls need no x-boundary tests
    if (x0 >= 1 && x1 + 2 <= X)
        row::row_tile<T, TY, PW, KF, false, NTL>(g, p, u, v, w, po, uo, vo, wo, K, Kf, dt, x0, x1, y0, wv, nt != 0, sp, sv);
    else
        row::row_tile<T, TY, PW, KF, true, NTL>(g, p, u, v, w, po, uo, vo, wo, K, Kf, dt, x0, x1, y0, wv, nt != 0, sp, sv);
}

template <class T, int TY, bool PW, bool KF, bool NTL>
static void launch_row_k(const UniformGeo<T>& g, void* const* in, void* const* out, double K, const void* Kf,
                         double dt, int L, int order, bool nt, hipStream_t s) {
    const int nyt = int((g.Y + TY - 1) / TY);
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
    if (g.vol_pow2) {
        if (Kf) launch_row_k<T, TY, true, true, NTL>(g, in, out, K, Kf, dt, L, order, nt, s);
        else launch_row_k<T, TY, true, false, NTL>(g, in, out, K, Kf, dt, L, order, nt, s);
    } else {
        if (Kf) launch_row_k<T, TY, false, true, NTL>(g, in, out, K, Kf, dt, L, order, nt, s);
        else launch_row_k<T, TY, false, false, NTL>(g, in, out, K, Kf, dt, L, order, nt, s);
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
        case 14:  // FP32 only: the FP64 form of this tile spills registers
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
