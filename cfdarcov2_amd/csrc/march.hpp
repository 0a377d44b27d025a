// 2.5D march tile of the fused acoustic step on a uniform mesh (see
// wave_march.hip), shared by the single-mesh kernel and the batched ensemble
// kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace cfd {

namespace march {

// In-domain index for a halo coordinate r in [-2, n+1]: r inside the domain,
// the boundary cell for r == -1 or r == n (the reference's neighbour = self
// rule), -1 where no real cell ever reads the slot.
__device__ inline int64_t clamp_halo(int64_t r, int64_t n) {
    if (r >= 0 && r < n) return r;
    if (r == -1) return 0;
    if (r == n) return n - 1;
    return -1;
}

template <class T> __device__ inline T mlin(T c, T nb) { return (c + nb) * T(0.5); }
template <class T> __device__ inline T mggp(T f_lo, T f_hi, T A, T V) {
    const T t_lo = (-f_lo) * A;
    const T t_hi = f_hi * A;
    return (t_lo + t_hi) / V;
}
// velocity update from the three pressures along the velocity's axis
template <class T> __device__ inline T mvel(T pl, T pc, T ph, T q, T A, T V, T dt) {
    return dt * (-mggp(mlin(pc, pl), mlin(pc, ph), A, V)) + q;
}

// Division by the cell volume.  When the volume is an exact power of two the
// quotient equals the product with its reciprocal bit for bit (both are the
// correctly rounded value of the same real number), and the product is one
// instruction instead of a division sequence.
template <class T, bool PW> struct VolDiv {
    T V, inv;
    __device__ T operator()(T x) const {
        if constexpr (PW) return x * inv;
        else return x / V;
    }
};
template <class T, bool PW> __device__ inline T mggp(T f_lo, T f_hi, T A, const VolDiv<T, PW>& d) {
    const T t_lo = (-f_lo) * A;
    const T t_hi = f_hi * A;
    return d(t_lo + t_hi);
}
template <class T, bool PW> __device__ inline T mvel(T pl, T pc, T ph, T q, T A, const VolDiv<T, PW>& d, T dt) {
    return dt * (-mggp(mlin(pc, pl), mlin(pc, ph), A, d)) + q;
}

}  // namespace march
using namespace march;

// LDS planes of one march tile (double-buffered)
template <class T, int TYT, int R> struct MarchLds {
    static constexpr int TY = TYT * R;
    static constexpr int TZ = 64;
    static constexpr int PZ = TZ + 4;          // p plane row length, halo 2 in z
    static constexpr int PS = (TY + 4) * PZ;   // p plane, halo 2 in y
    static constexpr int VS = (TY + 2) * TZ;   // v' plane, halo 1 in y
    static constexpr int WZ = TZ + 2;
    static constexpr int WS = TY * WZ;         // w' plane, halo 1 in z
    T sp[2][PS];
    T sv[2][VS];
    T sw[2][WS];
};

// One (zt, yt, xc) tile of the fused step: columns [yt*TY, +TY) x [zt*64, +64),
// planes [xc*L, +L).  Called by the single-mesh kernel below and by the
// batched ensemble kernel.
// NT: stream the outputs past the L2 (non-temporal stores), leaving its
// capacity to the halo rows that neighbouring tiles read again.
template <class T, int TYT, int R, bool NT = false, bool PW = false>
__device__ inline void wave_march_tile(const UniformGeo<T>& g, const T* __restrict__ p, const T* __restrict__ u,
                                       const T* __restrict__ v, const T* __restrict__ w, T* __restrict__ po,
                                       T* __restrict__ uo, T* __restrict__ vo, T* __restrict__ wo, T K,
                                       const T* __restrict__ Kf, T dt, int L, int64_t zt, int64_t yt, int64_t xc,
                                       MarchLds<T, TYT, R>& lds) {
    static_assert(TYT >= 4, "four halo-row threads are needed");
    using M = MarchLds<T, TYT, R>;
    constexpr int TY = M::TY;
    constexpr int TZ = M::TZ;
    constexpr int PZ = M::PZ;
    constexpr int WZ = M::WZ;
    auto& sp = lds.sp;
    auto& sv = lds.sv;
    auto& sw = lds.sw;

    const int64_t X = g.X, Y = g.Y, Z = g.Z, SX = Y * Z;
    const T Ax = g.area[0], Ay = g.area[1], Az = g.area[2];
    const VolDiv<T, PW> V{g.vol, g.inv_vol};

    const int lz = threadIdx.x, ly = threadIdx.y;
    const int64_t z = zt * TZ + lz;
    const int64_t y0 = yt * TY;
    const int64_t x0 = xc * L;
    const int64_t x1 = (x0 + L < X) ? x0 + L : X;
    if (x0 >= X) return;  // uniform across the workgroup
    const int64_t zc = clamp_halo(z, Z);

    // per-slot geometry
    int64_t col[R];
    bool valid[R], slot[R], first[R], last[R];
    int own_p[R], own_v[R], own_w[R];
    int64_t ch_col[R];
    int ch_lds[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = r * TYT + ly;
        const int64_t y = y0 + row;
        valid[r] = z < Z && y < Y;
        const int64_t yc = clamp_halo(y, Y);
        slot[r] = yc >= 0 && zc >= 0;
        col[r] = slot[r] ? yc * Z + zc : 0;
        first[r] = y == 0;
        last[r] = y == Y - 1;
        own_p[r] = (row + 2) * PZ + (lz + 2);
        own_v[r] = (row + 1) * TZ + lz;
        own_w[r] = row * WZ + (lz + 1);
        ch_lds[r] = -1;
        ch_col[r] = -1;
        if (lz < 2 || lz >= TZ - 2) {  // p-plane column halo: cols -2,-1 / 64,65
            const int c = lz < 2 ? lz - 2 : lz + 2;
            ch_lds[r] = (row + 2) * PZ + (c + 2);
            const int64_t gz = clamp_halo(zt * TZ + c, Z);
            if (gz >= 0 && yc >= 0) ch_col[r] = yc * Z + gz;
        }
    }
    // p-plane row halo: rows -2, -1, TY, TY+1 by threads ly = 0..3
    int rh_lds = -1;
    int64_t rh_col = -1;
    if (ly < 4) {
        const int hr = ly < 2 ? ly - 2 : TY + (ly - 2);
        rh_lds = (hr + 2) * PZ + (lz + 2);
        const int64_t gy = clamp_halo(y0 + hr, Y);
        if (gy >= 0 && zc >= 0) rh_col = gy * Z + zc;
    }
    // v' halo rows: row -1 by ly 0, row TY by ly 1 (only where they are real
    // cells); the w' halo columns -1 / 64 are computed per slot by lanes 0 / 63
    const bool vh_lo = ly == 0 && y0 >= 1 && z < Z;
    const bool vh_hi = ly == 1 && y0 + TY < Y && z < Z;
    const int64_t vh_col = vh_lo ? (y0 - 1) * Z + z : vh_hi ? (y0 + TY) * Z + z : 0;
    const int vh_lds = vh_lo ? lz : (TY + 1) * TZ + lz;
    const int vh_p = vh_lo ? 1 * PZ + (lz + 2) : (TY + 2) * PZ + (lz + 2);  // p row of the halo cell

    // ---- warm-up: p(x0-1..x0+1) and u'(x0-1), u'(x0) per slot ---------------
    T Pm1[R], P0[R], Pp1[R], Um1[R], U0[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        Pm1[r] = P0[r] = Pp1[r] = Um1[r] = U0[r] = T(0);
        if (slot[r]) {
            P0[r] = p[x0 * SX + col[r]];
            if (x0 >= 1) Pm1[r] = p[(x0 - 1) * SX + col[r]];
            if (x0 + 1 < X) Pp1[r] = p[(x0 + 1) * SX + col[r]];
            if (valid[r]) {
                if (x0 >= 1) {
                    const T Pm2 = (x0 >= 2) ? p[(x0 - 2) * SX + col[r]] : Pm1[r];
                    Um1[r] = mvel(Pm2, Pm1[r], P0[r], u[(x0 - 1) * SX + col[r]], Ax, V, dt);
                }
                U0[r] = mvel(x0 >= 1 ? Pm1[r] : P0[r], P0[r], x0 + 1 < X ? Pp1[r] : P0[r], u[x0 * SX + col[r]], Ax,
                             V, dt);
            }
        }
    }

    // ---- prefetch registers: the loads of plane x + 1 are issued before the
    // arithmetic of plane x
    struct Fetch {
        T pp2[R], un[R], vn[R], wn[R], ch[R], wlo[R], whi[R];
        T rh, vh;
    };
    auto fetch = [&](int64_t x, Fetch& f) {
        const int64_t base = x * SX;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (slot[r] && x + 2 < X) f.pp2[r] = p[base + 2 * SX + col[r]];
            if (valid[r] && x + 1 < X) f.un[r] = u[base + SX + col[r]];
            if (valid[r]) {
                f.vn[r] = v[base + col[r]];
                f.wn[r] = w[base + col[r]];
            }
            if (ch_col[r] >= 0) f.ch[r] = p[base + ch_col[r]];
            if (valid[r] && lz == 0 && z >= 1) f.wlo[r] = w[base + col[r] - 1];
            if (valid[r] && lz == TZ - 1 && z + 1 < Z) f.whi[r] = w[base + col[r] + 1];
        }
        if (rh_col >= 0) f.rh = p[base + rh_col];
        if (vh_lo || vh_hi) f.vh = v[base + vh_col];
    };
    auto clear = [&](Fetch& f) {
#pragma unroll
        for (int r = 0; r < R; ++r) f.pp2[r] = f.un[r] = f.vn[r] = f.wn[r] = f.ch[r] = f.wlo[r] = f.whi[r] = T(0);
        f.rh = f.vh = T(0);
    };
    Fetch fa;
    clear(fa);
    fetch(x0, fa);

    int b = 0;
    for (int64_t x = x0; x < x1; ++x, b ^= 1) {
        T Pp2[R], un[R], vold[R], wold[R], ch[R], wlo[R], whi[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Pp2[r] = fa.pp2[r]; un[r] = fa.un[r]; vold[r] = fa.vn[r]; wold[r] = fa.wn[r];
            ch[r] = fa.ch[r]; wlo[r] = fa.wlo[r]; whi[r] = fa.whi[r];
        }
        const T rh = fa.rh, vh = fa.vh;
        if (x + 1 < x1) fetch(x + 1, fa);

        T* P = sp[b];
        T* SV = sv[b];
        T* SW = sw[b];
        // u'(x+1) from p(x..x+2); p plane x into LDS
        T Up1[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Up1[r] = T(0);
            if (valid[r] && x + 1 < X) Up1[r] = mvel(P0[r], Pp1[r], (x + 2 < X) ? Pp2[r] : Pp1[r], un[r], Ax, V, dt);
            if (slot[r]) P[own_p[r]] = P0[r];
            if (ch_lds[r] >= 0) P[ch_lds[r]] = ch_col[r] >= 0 ? ch[r] : T(0);
        }
        if (rh_lds >= 0) P[rh_lds] = rh_col >= 0 ? rh : T(0);
        __syncthreads();

        // v' (y-halo 1) and w' (z-halo 1) of plane x into LDS
        T vnew[R], wnew[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            vnew[r] = wnew[r] = T(0);
            if (valid[r]) {
                const int op = own_p[r];
                vnew[r] = mvel(P[op - PZ], P0[r], P[op + PZ], vold[r], Ay, V, dt);
                wnew[r] = mvel(P[op - 1], P0[r], P[op + 1], wold[r], Az, V, dt);
                SV[own_v[r]] = vnew[r];
                SW[own_w[r]] = wnew[r];
                if (lz == 0 && z >= 1) SW[own_w[r] - 1] = mvel(P[op - 2], P[op - 1], P0[r], wlo[r], Az, V, dt);
                if (lz == TZ - 1 && z + 1 < Z) SW[own_w[r] + 1] = mvel(P0[r], P[op + 1], P[op + 2], whi[r], Az, V, dt);
            }
        }
        if (vh_lo || vh_hi) SV[vh_lds] = mvel(P[vh_p - PZ], P[vh_p], P[vh_p + PZ], vh, Ay, V, dt);
        __syncthreads();

#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (valid[r]) {
                const T fx_lo = (x > 0) ? mlin(U0[r], Um1[r]) : T(0);
                const T fx_hi = (x < X - 1) ? mlin(U0[r], Up1[r]) : T(0);
                const T fy_lo = first[r] ? T(0) : mlin(vnew[r], SV[own_v[r] - TZ]);
                const T fy_hi = last[r] ? T(0) : mlin(vnew[r], SV[own_v[r] + TZ]);
                const T fz_lo = (z > 0) ? mlin(wnew[r], SW[own_w[r] - 1]) : T(0);
                const T fz_hi = (z < Z - 1) ? mlin(wnew[r], SW[own_w[r] + 1]) : T(0);
                T S = mggp(fx_lo, fx_hi, Ax, V);
                S = S + mggp(fy_lo, fy_hi, Ay, V);
                S = S + mggp(fz_lo, fz_hi, Az, V);
                const int64_t o = x * SX + col[r];
                const T k = Kf ? Kf[o] : K;
                const T pn = dt * (k * S) + P0[r];
                if constexpr (NT) {
                    __builtin_nontemporal_store(pn, &po[o]);
                    __builtin_nontemporal_store(U0[r], &uo[o]);
                    __builtin_nontemporal_store(vnew[r], &vo[o]);
                    __builtin_nontemporal_store(wnew[r], &wo[o]);
                } else {
                    po[o] = pn;
                    uo[o] = U0[r];
                    vo[o] = vnew[r];
                    wo[o] = wnew[r];
                }
            }
            Pm1[r] = P0[r];
            P0[r] = Pp1[r];
            Pp1[r] = Pp2[r];
            Um1[r] = U0[r];
            U0[r] = Up1[r];
        }
    }
}

// LDS plane of one heat march tile (halo 1 in y and z, double-buffered)
template <class T, int TYT, int R> struct HeatLds {
    static constexpr int TY = TYT * R;
    static constexpr int TZ = 64;
    static constexpr int PZ = TZ + 2;
    static constexpr int PS = (TY + 2) * PZ;
    T s[2][PS];
};

// One (zt, yt, xc) tile of the heat step dt * (alpha * lapl(T)) + T on a
// uniform mesh, marching along x with T(x-1), T(x), T(x+1) in registers and the
// plane's y/z neighbours in LDS.  The arithmetic is heat_uniform_k's
// (Grad2Var with neighbour = self on the walls), so the two agree bit for bit.
template <class T, int TYT, int R, bool PW = false>
__device__ inline void heat_march_tile(const UniformGeo<T>& g, const T* __restrict__ Tin, T* __restrict__ Tout,
                                       T alpha, T dt, int L, int64_t zt, int64_t yt, int64_t xc,
                                       HeatLds<T, TYT, R>& lds) {
    static_assert(TYT >= 2, "two halo-row threads are needed");
    using H = HeatLds<T, TYT, R>;
    constexpr int TY = H::TY;
    constexpr int TZ = H::TZ;
    constexpr int PZ = H::PZ;
    const int64_t X = g.X, Y = g.Y, Z = g.Z, SX = Y * Z;
    const VolDiv<T, PW> Vd{g.vol, g.inv_vol};
    const int lz = threadIdx.x, ly = threadIdx.y;
    const int64_t z = zt * TZ + lz;
    const int64_t y0 = yt * TY;
    const int64_t x0 = xc * L;
    const int64_t x1 = (x0 + L < X) ? x0 + L : X;
    if (x0 >= X) return;  // uniform across the workgroup
    const int64_t zc = clamp_halo(z, Z);

    int64_t col[R];
    bool valid[R];
    int own[R];
    int64_t zh_col[R];  // z-halo column of lanes 0 / 63 (-1: none)
    int zh_lds[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = r * TYT + ly;
        const int64_t y = y0 + row;
        valid[r] = z < Z && y < Y;
        col[r] = valid[r] ? y * Z + z : 0;
        own[r] = (row + 1) * PZ + (lz + 1);
        zh_col[r] = -1;
        zh_lds[r] = -1;
        if (y < Y && (lz == 0 || lz == TZ - 1)) {
            const int c = lz == 0 ? -1 : TZ;
            const int64_t gz = clamp_halo(zt * TZ + c, Z);
            if (gz >= 0) {
                zh_col[r] = y * Z + gz;
                zh_lds[r] = (row + 1) * PZ + (c + 1);
            }
        }
    }
    // y-halo rows -1 (ly 0) and TY (ly 1), neighbour = self at the walls
    int64_t yh_col = -1;
    int yh_lds = -1;
    if ((ly == 0 || ly == 1) && zc >= 0 && z < Z) {
        const int hr = ly == 0 ? -1 : TY;
        const int64_t gy = clamp_halo(y0 + hr, Y);
        if (gy >= 0) {
            yh_col = gy * Z + z;
            yh_lds = (hr + 1) * PZ + (lz + 1);
        }
    }
    T Tm1[R], T0[R], Tp1[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        Tm1[r] = T0[r] = Tp1[r] = T(0);
        if (valid[r]) {
            T0[r] = Tin[x0 * SX + col[r]];
            if (x0 >= 1) Tm1[r] = Tin[(x0 - 1) * SX + col[r]];
            if (x0 + 1 < X) Tp1[r] = Tin[(x0 + 1) * SX + col[r]];
        }
    }
    int b = 0;
    for (int64_t x = x0; x < x1; ++x, b ^= 1) {
        const int64_t base = x * SX;
        T* P = lds.s[b];
        T nxt[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            nxt[r] = T(0);
            if (valid[r] && x + 2 < X) nxt[r] = Tin[base + 2 * SX + col[r]];
            if (valid[r]) P[own[r]] = T0[r];
            if (zh_lds[r] >= 0) P[zh_lds[r]] = Tin[base + zh_col[r]];
        }
        if (yh_lds >= 0) P[yh_lds] = Tin[base + yh_col];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (valid[r]) {
                const int row = r * TYT + ly;
                const int64_t y = y0 + row;
                const T pc = T0[r];
                // walls: the neighbour is the cell itself (reference rule); the
                // LDS slots read here are all written this plane
                const bool lo_w[3] = {x == 0, y == 0, z == 0};
                const bool hi_w[3] = {x == X - 1, y == Y - 1, z == Z - 1};
                const T nb_lo[3] = {lo_w[0] ? pc : Tm1[r], lo_w[1] ? pc : P[own[r] - PZ], lo_w[2] ? pc : P[own[r] - 1]};
                const T nb_hi[3] = {hi_w[0] ? pc : Tp1[r], hi_w[1] ? pc : P[own[r] + PZ], hi_w[2] ? pc : P[own[r] + 1]};
                T Lp = T(0);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const T al_lo = lo_w[a] ? g.alpha_b[a] : g.alpha[a];
                    const T al_hi = hi_w[a] ? g.alpha_b[a] : g.alpha[a];
                    const T s_lo = (-pc + nb_lo[a]) * al_lo;
                    const T s_hi = (-pc + nb_hi[a]) * al_hi;
                    Lp = Lp + Vd(s_lo * g.area[a] + s_hi * g.area[a]);
                }
                Tout[base + col[r]] = dt * (alpha * Lp) + pc;
            }
            Tm1[r] = T0[r];
            T0[r] = Tp1[r];
            Tp1[r] = nxt[r];
        }
    }
}

}  // namespace cfd
