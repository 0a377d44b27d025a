// Tile code of the two-step row form (wave_row2.hip) and of the batched
// two-step ensemble step (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "march.hpp"
#include "row.hpp"

namespace cfd {

namespace row2 {

// the four consecutive cells of a lane, as one vector
template <class T> using V4T = typename row::Row4<T>::type;
using row::kWaveShl1;
using row::kWaveShr1;
using row::lane_shift;

// z-neighbours of a row vector, neighbour = self at the row's two ends
template <class V4>
__device__ inline void z_nbrs(const V4& c, int lane, int nl, V4& lo, V4& hi) {
    lo[0] = lane_shift<kWaveShr1>(c[0], c[3]);
    lo[1] = c[0];
    lo[2] = c[1];
    lo[3] = c[2];
    hi[0] = c[1];
    hi[1] = c[2];
    hi[2] = c[3];
    hi[3] = lane_shift<kWaveShl1>(c[3], c[0]);
    if (lane == 0) lo[0] = c[0];
    if (lane == nl - 1) hi[3] = c[3];
}

template <class T, bool PW>
__device__ inline V4T<T> vel4(const V4T<T>& lo, const V4T<T>& c, const V4T<T>& hi, const V4T<T>& q, T A,
                              const VolDiv<T, PW>& Vd, T dt) {
    V4T<T> r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = mvel(lo[i], c[i], hi[i], q[i], A, Vd, dt);
    return r;
}

// pressure update of a row vector: the single-step kernels' flux sums
template <class T, bool PW>
__device__ inline V4T<T> pres4(const V4T<T>& P, const V4T<T>& Um, const V4T<T>& U, const V4T<T>& Up, bool xlo,
                               bool xhi, const V4T<T>& vn, const V4T<T>& vyl, const V4T<T>& vyh, bool first, bool last,
                               const V4T<T>& wn, int lane, int nl, T Ax, T Ay, T Az, const VolDiv<T, PW>& Vd, T K,
                               T dt) {
    V4T<T> wzl, wzh;
    z_nbrs(wn, lane, nl, wzl, wzh);
    V4T<T> pn;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const T fx_lo = xlo ? mlin(U[i], Um[i]) : T(0);
        const T fx_hi = xhi ? mlin(U[i], Up[i]) : T(0);
        const T fy_lo = first ? T(0) : mlin(vn[i], vyl[i]);
        const T fy_hi = last ? T(0) : mlin(vn[i], vyh[i]);
        const T fz_lo = (i == 0 && lane == 0) ? T(0) : mlin(wn[i], wzl[i]);
        const T fz_hi = (i == 3 && lane == nl - 1) ? T(0) : mlin(wn[i], wzh[i]);
        T S = mggp(fx_lo, fx_hi, Ax, Vd);
        S = S + mggp(fy_lo, fy_hi, Ay, Vd);
        S = S + mggp(fz_lo, fz_hi, Az, Vd);
        pn[i] = dt * (K * S) + P[i];
    }
    return pn;
}

template <class T, int TY, bool PW, bool XB, bool NTL>
__device__ inline void tile2(const UniformGeo<T>& g, const T* __restrict__ p, const T* __restrict__ u,
                             const T* __restrict__ v, const T* __restrict__ w, T* __restrict__ po, T* __restrict__ uo,
                             T* __restrict__ vo, T* __restrict__ wo, T K, T dt, int64_t x0, int64_t x1, int y0, int j,
                             bool nt, V4T<T> (*P1)[64], V4T<T> (*V1)[64], V4T<T> (*P2)[64], V4T<T> (*V2)[64]) {
    const int64_t X = g.X;
    const int Y = int(g.Y), Z = int(g.Z);
    const int64_t SX = int64_t(Y) * Z;
    const T Ax = g.area[0], Ay = g.area[1], Az = g.area[2];
    const VolDiv<T, PW> Vd{g.vol, g.inv_vol};
    const int lane = threadIdx.x;
    const int nl = Z / 4;
    const bool lv = lane < nl;

    // ---- the wave's roles (uniform across the wave) --------------------------
    const int y = y0 - 3 + j;
    const bool real = y >= 0 && y < Y;
    const int d = y < y0 ? y0 - y : (y > y0 + TY - 1 ? y - (y0 + TY - 1) : 0);
    const bool rA = real && d <= 2;  // first step in full
    const bool rH = real && d == 3;  // first step's v' only
    const bool rB = real && d <= 1;  // second step's v'', w''
    const bool rO = real && d == 0;  // second step in full, outputs
    const bool first = y == 0, last = y == Y - 1;
    const int jl = first ? j : j - 1, jh = last ? j : j + 1;  // LDS rows of the y-neighbours (self at walls)
    const bool lowerH = y < y0;
    const int orow = lowerH ? (y >= 1 ? y - 1 : y) : (y + 1 < Y ? y + 1 : y);  // halo wave's outer row
    const int jin = lowerH ? j + 1 : j - 1;                                    // and its inner LDS row
    const int64_t rb = int64_t(real ? y : 0) * Z;
    const int64_t ob = int64_t(real ? orow : 0) * Z;

    // per-plane loads as (field, plane offset, row) slots:
    //   d <= 2: a = p(xx+2), b = u(xx+1), c = v(xx), d = w(xx)
    //   d == 3: a = p(xx), b = p(xx) of the outer row, c = v(xx)
    const T* const fbp = rA ? u : p;
    const int da = rA ? 2 : 0, db = rA ? 1 : 0;
    const int64_t rbb = rA ? rb : ob;
    const bool ea = rA || rH;
    struct Fetch {
        V4T<T> a, b, c, d;
    };
    const V4T<T> zero4 = {T(0), T(0), T(0), T(0)};
    auto ld = [&](const T* base, int64_t off) -> V4T<T> {
        const V4T<T>* q = reinterpret_cast<const V4T<T>*>(base + off) + lane;
        if constexpr (NTL) return __builtin_nontemporal_load(q);
        else return *q;
    };
    auto fetch = [&](int64_t xx, Fetch& f) {
        if (!lv || !ea) return;
        const int64_t pb = xx * SX;
        if (!XB || xx + da < X) f.a = ld(p, pb + da * SX + rb);
        if (!XB || xx + db < X) f.b = ld(fbp, pb + db * SX + rbb);
        f.c = ld(v, pb + rb);
        if (rA) f.d = ld(w, pb + rb);
    };

    const int64_t xs = x0 >= 2 ? x0 - 2 : 0;  // first plane of the first step
    const int64_t xe = x1 + 2;                // end of the iterations
    const int64_t xs1 = xe < X ? xe : X;      // end of the first step's planes

    // ---- warm-up of the first step's x queue (the single-step kernel's) ------
    // every load of the warm-up and of the first plane issued before any is
    // used (planes clamped into the mesh, the wall cases chosen by selects),
    // as the single-step row tile's warm-up
    V4T<T> P0 = zero4, Pp1 = zero4, Um1 = zero4, U0 = zero4;
    V4T<T> a0 = zero4, am1 = zero4, ap1 = zero4, am2 = zero4, um = zero4, uc = zero4;
    if (rA && lv) {
        const int64_t xm1 = xs >= 1 ? xs - 1 : xs, xm2 = xs >= 2 ? xs - 2 : xm1;
        const int64_t xp1 = xs + 1 < X ? xs + 1 : xs;
        a0 = ld(p, xs * SX + rb);
        am1 = ld(p, xm1 * SX + rb);
        ap1 = ld(p, xp1 * SX + rb);
        am2 = ld(p, xm2 * SX + rb);
        um = ld(u, xm1 * SX + rb);
        uc = ld(u, xs * SX + rb);
    }
    Fetch nx{zero4, zero4, zero4, zero4};
    fetch(xs, nx);
    if (rA && lv) {
        P0 = a0;
        const V4T<T> Pm1 = xs >= 1 ? am1 : zero4;
        Pp1 = xs + 1 < X ? ap1 : zero4;
        if (xs >= 1) {
            const V4T<T> Pm2 = (xs >= 2) ? am2 : Pm1;
            Um1 = vel4(Pm2, Pm1, P0, um, Ax, Vd, dt);
        }
        V4T<T> lo = P0, hi = P0;
        if (xs >= 1) lo = Pm1;
        if (xs + 1 < X) hi = Pp1;
        U0 = vel4(lo, P0, hi, uc, Ax, Vd, dt);
    }
    // second-step queues: p'(x-2), p'(x-1); v', w' of x-2, x-1; u''(x-3), u''(x-2)
    V4T<T> Q2 = zero4, Q1 = zero4, Va2 = zero4, Va1 = zero4, Wa2 = zero4, Wa1 = zero4, R3 = zero4, R2 = zero4;

    for (int64_t x = xs; x < xe; ++x) {
        const bool s1 = !XB || x < X;  // first step at plane x
        const int64_t xp = x - 2;      // second step's plane
        const bool s2 = xp >= x0 && xp < x1;
        const int64_t m = x - 1;       // u'' computed at plane m
        const bool su = m >= x0 - 1 && m >= 0 && m < X;
        const Fetch cur = nx;
        if (x + 1 < xs1) fetch(x + 1, nx);

        // A. u'(x+1); p(x) and p'(x-2) into LDS
        V4T<T> Up1 = zero4;
        if (rA) {
            if (s1) {
                if (!XB || x + 1 < X) {
                    const V4T<T> hi = (!XB || x + 2 < X) ? cur.a : Pp1;
                    Up1 = vel4(P0, Pp1, hi, cur.b, Ax, Vd, dt);
                }
                P1[j][lane] = P0;
            }
            if (s2) P2[j][lane] = Q2;
        } else if (rH && s1) {
            P1[j][lane] = cur.a;
        }
        __syncthreads();

        // B. v'(x), w'(x); v''(x-2), w''(x-2)
        V4T<T> vn = zero4, wn = zero4, v2 = zero4, w2 = zero4;
        if (rA) {
            if (s1) {
                const V4T<T> pyl = P1[jl][lane];
                const V4T<T> pyh = P1[jh][lane];
                V4T<T> pzl, pzh;
                z_nbrs(P0, lane, nl, pzl, pzh);
                vn = vel4(pyl, P0, pyh, cur.c, Ay, Vd, dt);
                wn = vel4(pzl, P0, pzh, cur.d, Az, Vd, dt);
                V1[j][lane] = vn;
            }
            if (rB && s2) {
                const V4T<T> qyl = P2[jl][lane];
                const V4T<T> qyh = P2[jh][lane];
                V4T<T> qzl, qzh;
                z_nbrs(Q2, lane, nl, qzl, qzh);
                v2 = vel4(qyl, Q2, qyh, Va2, Ay, Vd, dt);
                if (rO) w2 = vel4(qzl, Q2, qzh, Wa2, Az, Vd, dt);
                V2[j][lane] = v2;
            }
        } else if (rH && s1) {
            const V4T<T> in = P1[jin][lane];
            const V4T<T> pl = lowerH ? cur.b : in;
            const V4T<T> ph = lowerH ? in : cur.b;
            V1[j][lane] = vel4(pl, cur.a, ph, cur.c, Ay, Vd, dt);
        }
        __syncthreads();

        // C. p'(x); u''(x-1); p''(x-2) and the outputs
        if (rA) {
            V4T<T> Q0 = zero4;
            if (s1) {
                const V4T<T> vyl = first ? vn : V1[j - 1][lane];
                const V4T<T> vyh = last ? vn : V1[j + 1][lane];
                Q0 = pres4(P0, Um1, U0, Up1, !XB || x > 0, !XB || x < X - 1, vn, vyl, vyh, first, last, wn, lane,
                           nl, Ax, Ay, Az, Vd, K, dt);
            }
            V4T<T> R1 = zero4;
            if (rO) {
                if (su) {
                    const V4T<T> lo = (!XB || m >= 1) ? Q2 : Q1;
                    const V4T<T> hi = (!XB || x < X) ? Q0 : Q1;
                    R1 = vel4(lo, Q1, hi, Um1, Ax, Vd, dt);
                }
                if (s2) {
                    const V4T<T> vyl = first ? v2 : V2[j - 1][lane];
                    const V4T<T> vyh = last ? v2 : V2[j + 1][lane];
                    const V4T<T> pn = pres4(Q2, R3, R2, R1, !XB || xp > 0, !XB || xp < X - 1, v2, vyl, vyh, first, last,
                                        w2, lane, nl, Ax, Ay, Az, Vd, K, dt);
                    if (lv) {
                        const int64_t o = xp * SX + rb;
                        row::put(reinterpret_cast<V4T<T>*>(po + o) + lane, pn, nt);
                        row::put(reinterpret_cast<V4T<T>*>(uo + o) + lane, R2, nt);
                        row::put(reinterpret_cast<V4T<T>*>(vo + o) + lane, v2, nt);
                        row::put(reinterpret_cast<V4T<T>*>(wo + o) + lane, w2, nt);
                    }
                }
            }
            P0 = Pp1;
            Pp1 = cur.a;
            Um1 = U0;
            U0 = Up1;
            Q2 = Q1;
            Q1 = Q0;
            Va2 = Va1;
            Va1 = vn;
            Wa2 = Wa1;
            Wa1 = wn;
            R3 = R2;
            R2 = R1;
        }
    }
}

// Pressure update of a lane's four cells when a wave holds several rows: the
// z walls are a row's first and last lane (zlo, zhi), first / last are the
// lane's own y walls.
template <class T, bool PW>
__device__ inline V4T<T> pres4m(const V4T<T>& P, const V4T<T>& Um, const V4T<T>& U, const V4T<T>& Up, bool xlo,
                                bool xhi, const V4T<T>& vn, const V4T<T>& vyl, const V4T<T>& vyh, bool first, bool last,
                                const V4T<T>& wn, bool zlo, bool zhi, T Ax, T Ay, T Az, const VolDiv<T, PW>& Vd, T K,
                                T dt) {
    V4T<T> wzl, wzh;
    row::z_nbrs_row(wn, zlo, zhi, wzl, wzh);
    V4T<T> pn;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const T fx_lo = xlo ? mlin(U[i], Um[i]) : T(0);
        const T fx_hi = xhi ? mlin(U[i], Up[i]) : T(0);
        const T fy_lo = first ? T(0) : mlin(vn[i], vyl[i]);
        const T fy_hi = last ? T(0) : mlin(vn[i], vyh[i]);
        const T fz_lo = (i == 0 && zlo) ? T(0) : mlin(wn[i], wzl[i]);
        const T fz_hi = (i == 3 && zhi) ? T(0) : mlin(wn[i], wzh[i]);
        T S = mggp(fx_lo, fx_hi, Ax, Vd);
        S = S + mggp(fy_lo, fy_hi, Ay, Vd);
        S = S + mggp(fz_lo, fz_hi, Az, Vd);
        pn[i] = dt * (K * S) + P[i];
    }
    return pn;
}

// Two steps per pass with several rows per wave (z extent below 256).  A
// workgroup of NW waves holds a band of NW*RW consecutive rows (RW = 256/Z per
// wave); the output rows are the band without its four outer rows on each
// side.  Every lane evaluates both steps on its cells: the first step is
// correct on all but the band's two outer rows on each side, the second on the
// output rows, which alone are stored.  The x pipeline (first step at plane x,
// second at x-2, register queues) and the four LDS planes are those of tile2,
// with row j of the band at element j*Z.
template <class T, int NW, bool PW, bool XB, bool NTL>
__device__ inline void tile2m(const UniformGeo<T>& g, const T* __restrict__ p, const T* __restrict__ u,
                              const T* __restrict__ v, const T* __restrict__ w, T* __restrict__ po, T* __restrict__ uo,
                              T* __restrict__ vo, T* __restrict__ wo, T K, T dt, int64_t x0, int64_t x1, int y0, int wv,
                              bool nt, V4T<T> (*sP1)[64], V4T<T> (*sV1)[64], V4T<T> (*sP2)[64], V4T<T> (*sV2)[64]) {
    const int64_t X = g.X;
    const int Y = int(g.Y), Z = int(g.Z);
    const int64_t SX = int64_t(Y) * Z;
    const T Ax = g.area[0], Ay = g.area[1], Az = g.area[2];
    const VolDiv<T, PW> Vd{g.vol, g.inv_vol};
    const int lane = threadIdx.x;
    const int RW = row::rows_per_wave(Z);
    const int zq = Z / 4;
    const int band = NW * RW;
    const int TR = band - 8;
    V4T<T>* const P1 = &sP1[0][0];
    V4T<T>* const V1 = &sV1[0][0];
    V4T<T>* const P2 = &sP2[0][0];
    V4T<T>* const V2 = &sV2[0][0];
    const row::LanePlace lp = row::lane_place(lane, Z, RW);
    const int bi = wv * RW + lp.r;
    const int y = y0 - 4 + bi;
    const bool real = lp.span && y >= 0 && y < Y;
    const bool out = real && y >= y0 && y < y0 + TR;
    const bool live = __any(real);
    const int own = wv * RW * zq + lane;
    const bool first = y == 0, last = y == Y - 1;
    // LDS neighbours: self at the y walls and at the band's ends; lanes past
    // the wave's span read their own slot (their results are never used)
    const int ilo = (!lp.span || first || bi == 0) ? own : own - zq;
    const int ihi = (!lp.span || last || bi == band - 1) ? own : own + zq;
    const bool zlo = lp.zc == 0, zhi = lp.zc + 4 == Z;
    const int64_t rb = int64_t(real ? y : 0) * Z + lp.zc;

    struct Fetch {
        V4T<T> a, b, c, d;
    };
    const V4T<T> zero4 = {T(0), T(0), T(0), T(0)};
    auto ld = [&](const T* base, int64_t off) -> V4T<T> {
        const V4T<T>* q = reinterpret_cast<const V4T<T>*>(base + off);
        if constexpr (NTL) return __builtin_nontemporal_load(q);
        else return *q;
    };
    // loads of plane xx: a = p(xx+2), b = u(xx+1), c = v(xx), d = w(xx)
    auto fetch = [&](int64_t xx, Fetch& f) {
        if (!real) return;
        const int64_t pb = xx * SX;
        if (!XB || xx + 2 < X) f.a = ld(p, pb + 2 * SX + rb);
        if (!XB || xx + 1 < X) f.b = ld(u, pb + SX + rb);
        f.c = ld(v, pb + rb);
        f.d = ld(w, pb + rb);
    };

    const int64_t xs = x0 >= 2 ? x0 - 2 : 0;
    const int64_t xe = x1 + 2;
    const int64_t xs1 = xe < X ? xe : X;

    // warm-up: every load of it and of the first plane issued before any is
    // used, as tile2's
    V4T<T> P0 = zero4, Pp1 = zero4, Um1 = zero4, U0 = zero4;
    V4T<T> a0 = zero4, am1 = zero4, ap1 = zero4, am2 = zero4, um = zero4, uc = zero4;
    if (real) {
        const int64_t xm1 = xs >= 1 ? xs - 1 : xs, xm2 = xs >= 2 ? xs - 2 : xm1;
        const int64_t xp1 = xs + 1 < X ? xs + 1 : xs;
        a0 = ld(p, xs * SX + rb);
        am1 = ld(p, xm1 * SX + rb);
        ap1 = ld(p, xp1 * SX + rb);
        am2 = ld(p, xm2 * SX + rb);
        um = ld(u, xm1 * SX + rb);
        uc = ld(u, xs * SX + rb);
    }
    Fetch nx{zero4, zero4, zero4, zero4};
    fetch(xs, nx);
    if (real) {
        P0 = a0;
        const V4T<T> Pm1 = xs >= 1 ? am1 : zero4;
        Pp1 = xs + 1 < X ? ap1 : zero4;
        if (xs >= 1) {
            const V4T<T> Pm2 = (xs >= 2) ? am2 : Pm1;
            Um1 = vel4(Pm2, Pm1, P0, um, Ax, Vd, dt);
        }
        V4T<T> lo = P0, hi = P0;
        if (xs >= 1) lo = Pm1;
        if (xs + 1 < X) hi = Pp1;
        U0 = vel4(lo, P0, hi, uc, Ax, Vd, dt);
    }
    V4T<T> Q2 = zero4, Q1 = zero4, Va2 = zero4, Va1 = zero4, Wa2 = zero4, Wa1 = zero4, R3 = zero4, R2 = zero4;

    for (int64_t x = xs; x < xe; ++x) {
        const bool s1 = !XB || x < X;
        const int64_t xp = x - 2;
        const bool s2 = xp >= x0 && xp < x1;
        const int64_t m = x - 1;
        const bool su = m >= x0 - 1 && m >= 0 && m < X;
        const Fetch cur = nx;
        if (x + 1 < xs1) fetch(x + 1, nx);

        // A. u'(x+1); p(x) and p'(x-2) into LDS
        V4T<T> Up1 = zero4;
        if (live) {
            if (s1) {
                if (!XB || x + 1 < X) {
                    const V4T<T> hi = (!XB || x + 2 < X) ? cur.a : Pp1;
                    Up1 = vel4(P0, Pp1, hi, cur.b, Ax, Vd, dt);
                }
                if (lp.span) P1[own] = P0;
            }
            if (s2 && lp.span) P2[own] = Q2;
        }
        __syncthreads();

        // B. v'(x), w'(x); v''(x-2), w''(x-2)
        V4T<T> vn = zero4, wn = zero4, v2 = zero4, w2 = zero4;
        if (live) {
            if (s1) {
                V4T<T> pzl, pzh;
                row::z_nbrs_row(P0, zlo, zhi, pzl, pzh);
                vn = vel4(P1[ilo], P0, P1[ihi], cur.c, Ay, Vd, dt);
                wn = vel4(pzl, P0, pzh, cur.d, Az, Vd, dt);
                if (lp.span) V1[own] = vn;
            }
            if (s2) {
                V4T<T> qzl, qzh;
                row::z_nbrs_row(Q2, zlo, zhi, qzl, qzh);
                v2 = vel4(P2[ilo], Q2, P2[ihi], Va2, Ay, Vd, dt);
                w2 = vel4(qzl, Q2, qzh, Wa2, Az, Vd, dt);
                if (lp.span) V2[own] = v2;
            }
        }
        __syncthreads();

        // C. p'(x); u''(x-1); p''(x-2) and the outputs
        if (live) {
            V4T<T> Q0 = zero4;
            if (s1)
                Q0 = pres4m(P0, Um1, U0, Up1, !XB || x > 0, !XB || x < X - 1, vn, V1[ilo], V1[ihi], first, last, wn,
                            zlo, zhi, Ax, Ay, Az, Vd, K, dt);
            V4T<T> R1 = zero4;
            if (su) {
                const V4T<T> lo = (!XB || m >= 1) ? Q2 : Q1;
                const V4T<T> hi = (!XB || x < X) ? Q0 : Q1;
                R1 = vel4(lo, Q1, hi, Um1, Ax, Vd, dt);
            }
            if (s2) {
                const V4T<T> pn = pres4m(Q2, R3, R2, R1, !XB || xp > 0, !XB || xp < X - 1, v2, V2[ilo], V2[ihi], first,
                                     last, w2, zlo, zhi, Ax, Ay, Az, Vd, K, dt);
                if (out) {
                    const int64_t o = xp * SX + rb;
                    row::put(reinterpret_cast<V4T<T>*>(po + o), pn, nt);
                    row::put(reinterpret_cast<V4T<T>*>(uo + o), R2, nt);
                    row::put(reinterpret_cast<V4T<T>*>(vo + o), v2, nt);
                    row::put(reinterpret_cast<V4T<T>*>(wo + o), w2, nt);
                }
            }
            P0 = Pp1;
            Pp1 = cur.a;
            Um1 = U0;
            U0 = Up1;
            Q2 = Q1;
            Q1 = Q0;
            Va2 = Va1;
            Va1 = vn;
            Wa2 = Wa1;
            Wa1 = wn;
            R3 = R2;
            R2 = R1;
        }
    }
}

}  // namespace row2

}  // namespace cfd
