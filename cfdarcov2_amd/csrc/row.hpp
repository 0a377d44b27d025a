// Helpers of the row-form acoustic kernels (wave_row.hip, wave_row2.hip): one
// wave holds a z row, four cells per lane.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"
#include "march.hpp"

namespace cfd {

namespace row {

template <class T> struct Row4;
template <> struct Row4<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <> struct Row4<double> {
    typedef double type __attribute__((ext_vector_type(4)));
};

// DPP wave shifts (CDNA): wave_shr:1 gives lane i the value of lane i-1,
// wave_shl:1 the value of lane i+1; the lane whose source is outside the wave
// keeps `own` (bound_ctrl off).
constexpr int kWaveShr1 = 0x138;
constexpr int kWaveShl1 = 0x130;

template <int CTRL> __device__ inline float lane_shift(float own, float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, own), __builtin_bit_cast(int, v), CTRL, 0xf, 0xf,
                                           false));
}
template <int CTRL> __device__ inline double lane_shift(double own, double v) {
    const unsigned long long o = __builtin_bit_cast(unsigned long long, own);
    const unsigned long long s = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = unsigned(__builtin_amdgcn_update_dpp(int(unsigned(o)), int(unsigned(s)), CTRL, 0xf, 0xf, false));
    const unsigned hi =
        unsigned(__builtin_amdgcn_update_dpp(int(unsigned(o >> 32)), int(unsigned(s >> 32)), CTRL, 0xf, 0xf, false));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// lane j's value of v (j wave-uniform)
__device__ inline int lane_value(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ inline float lane_value(float v, int j) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), j));
}
__device__ inline double lane_value(double v, int j) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(unsigned(b)), j));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(unsigned(b >> 32)), j));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

template <class V4> __device__ inline void put(V4* dst, const V4& v, bool nt) {
    if (nt) __builtin_nontemporal_store(v, dst);
    else *dst = v;
}

// Rows per wave: a wave holds RW = 256 / Z consecutive z rows (one contiguous
// span of at most 256 cells, four per lane), so narrow meshes keep their lanes
// busy.
__host__ __device__ inline int rows_per_wave(int64_t Z) { return Z >= 256 ? 1 : int(256 / Z); }

// Per-lane place of a wave's cells: the lane's first cell is (row r, column
// zc) of the wave's span of RW rows; `span` says the lane holds cells at all.
struct LanePlace {
    int r, zc;
    bool span;
};
__device__ inline LanePlace lane_place(int lane, int Z, int RW) {
    const int c = 4 * lane;
    return {c / Z, c % Z, c < RW * Z};
}

// z-neighbours of a lane's four cells, neighbour = self at the z walls
template <class V4>
__device__ inline void z_nbrs_row(const V4& c, bool zlo_wall, bool zhi_wall, V4& lo, V4& hi) {
    lo[0] = lane_shift<kWaveShr1>(c[0], c[3]);
    lo[1] = c[0];
    lo[2] = c[1];
    lo[3] = c[2];
    hi[0] = c[1];
    hi[1] = c[2];
    hi[2] = c[3];
    hi[3] = lane_shift<kWaveShl1>(c[3], c[0]);
    if (zlo_wall) lo[0] = c[0];
    if (zhi_wall) hi[3] = c[3];
}

// Extra inputs and outputs of the row tile's trace-recording forward step
// (MODE 1) and transposed step (MODE 2, and MODE 3 with the K gradient).
template <class T> struct RowAux {
    T dK;              // MODE 2/3: dt * K; the transposed step's q = dK * lam_p
    T* trace;          // MODE 1: p' at the sensor cells (nullable)
    T* s_out;          // MODE 1: the velocity divergence S of every cell (nullable)
    const T* src;      // MODE 2/3: added to nu_p at the sensor cells (nullable)
    const T* s_in;     // MODE 3: the forward step's S, read with gK
    T* gK;             // MODE 3: gK += (lam_p dt) S
    // MODE 1-3: per tile (the kernel's tile index: x chunk * y tiles + y tile)
    // the sensors inside it (bit j = sensor j), built on the host; a tile
    // without sensors reads nothing else of the list
    const uint64_t* tile_sensors;
};

// One (y tile, x chunk) of the single-step acoustic row form: TY output waves
// of RW rows each (rows y0 .. y0+TY*RW-1) and two halo waves (rows y0-1 and
// y0+TY*RW, one row each), marching x over [x0, x1).  The LDS planes hold the
// tile's rows with the halo, row j at element j*Z.  Used by the single-mesh
// kernel (wave_row.hip), the batched ensemble kernel and the adjoint kernels
// (adj_row.hip).
//
// MODE 0 is the fused step.  MODE 1 is the same step that also stores p' at
// the sensor cells and, when aux->s_out is set, the divergence S (the value
// that multiplies dt K).  MODE 2 is the transposed step of the adjoint: with
// (p, u, v, w) read as (lam_p, lam_u, lam_v, lam_w) and q = dK lam_p,
//   mu_a = lam_a - G_a(q),  nu_p = dt (Dx mu_x + Dy mu_y + Dz mu_z) + lam_p,
// which is the fused step with q in place of p, the velocity update's
// dt * (-G) + u replaced by lam - G(q), and the pressure update's
// dt * (K * S) + p by dt * S + lam_p (KF must be false); aux->src is added to
// nu_p at the sensors.  MODE 3 is MODE 2 whose aux->gK also gains
// (lam_p dt) S_t from the forward step's aux->s_in (a mode of its own, so the
// plain transposed step keeps no registers for it).
template <class T, int TY, bool PW, bool KF, bool XB, bool NTL, bool MR = true, int MODE = 0>
__device__ inline void row_tile(const UniformGeo<T>& g, const T* __restrict__ p, const T* __restrict__ u,
                                const T* __restrict__ v, const T* __restrict__ w, T* __restrict__ po,
                                T* __restrict__ uo, T* __restrict__ vo, T* __restrict__ wo, T K,
                                const T* __restrict__ Kf, T dt, int64_t x0, int64_t x1, int y0, int wv, bool nt,
                                typename Row4<T>::type (*sp)[64], typename Row4<T>::type (*sv)[64],
                                const SensorList* sens = nullptr, const RowAux<T>* aux = nullptr,
                                int64_t tile = 0) {
    constexpr bool ADJ = MODE >= 2;
    constexpr bool GKM = MODE == 3;
    static_assert(!(ADJ && KF), "the transposed row step takes a constant K");
    using V4 = typename Row4<T>::type;
    const int64_t X = g.X;
    const int Y = int(g.Y), Z = int(g.Z);
    const int64_t SX = int64_t(Y) * Z;
    const T Ax = g.area[0], Ay = g.area[1], Az = g.area[2];
    const VolDiv<T, PW> Vd{g.vol, g.inv_vol};
    const int lane = threadIdx.x;
    const int RW = rows_per_wave(Z);
    const int zq = Z / 4;                 // V4 entries per row
    V4* const SP = &sp[0][0];
    V4* const SV = &sv[0][0];

    // ---- the wave's role (uniform) and the lane's cells ------------------------
    const bool is_out = wv < TY;
    const bool lower = wv == TY;
    const LanePlace lp = is_out ? lane_place(lane, Z, RW) : LanePlace{0, 4 * lane, lane < zq};
    // MR false: one row per wave (the caller has RW == 1), so the row is
    // uniform across the wave and its wall tests are scalar
    const int y = MR ? (is_out ? y0 + wv * RW + lp.r : (lower ? y0 - 1 : y0 + TY * RW))
                     : __builtin_amdgcn_readfirstlane(is_out ? y0 + wv : (lower ? y0 - 1 : y0 + TY));
    const bool real = lp.span && y >= 0 && y < Y;
    const bool out = is_out && real;             // a cell of an output row
    // an output wave with no cell inside the mesh (a tile past the last row)
    // skips the arithmetic; no real cell reads its LDS slots (the last real
    // row takes itself as its wall neighbour)
    const bool live = is_out && __any(out);
    const bool halo_v = !is_out && real;         // a cell of a halo row: v' of the row
    const int own = is_out ? (1 + wv * RW) * zq + lane : (lower ? lane : (1 + TY * RW) * zq + lane);
    const int ilo = y == 0 ? own : own - zq;     // LDS neighbours, self at the y walls
    const int ihi = y == Y - 1 ? own : own + zq;
    const bool first = y == 0, last = y == Y - 1;
    const bool zlo = lp.zc == 0, zhi = lp.zc + 4 == Z;
    // halo rows: the outer p row (clamped) and the LDS entry of the inner one
    const int orow = lower ? (y >= 1 ? y - 1 : 0) : (y + 1 < Y ? y + 1 : Y - 1);
    const int iin = lower ? own + zq : own - zq;
    const int64_t rb = int64_t(real ? y : 0) * Z + lp.zc;   // first cell of the lane's run
    const int64_t ob = int64_t(real ? orow : 0) * Z + lp.zc;

    // MODE 2: pressure-slot loads become q = dK * lam_p, and the velocity
    // update is lam - G(q)
    T dK = T(0);
    if constexpr (ADJ) dK = aux->dK;
    auto qv = [&](const V4& a) -> V4 {
        if constexpr (ADJ) {
            V4 r;
#pragma unroll
            for (int i = 0; i < 4; ++i) r[i] = dK * a[i];
            return r;
        } else {
            return a;
        }
    };
    auto vel = [&](T pl, T pc, T ph, T q, T A) -> T {
        if constexpr (ADJ) return q - mggp(mlin(pc, pl), mlin(pc, ph), A, Vd);
        else return mvel(pl, pc, ph, q, A, Vd, dt);
    };
    // MODE 1-3: the tile's sensors come from a per-tile mask (one scalar load;
    // most tiles have none and read nothing more).  In a tile with sensors,
    // lane j loads sensor j's x, y, z (and in MODE 2/3 its trace gradient)
    // and one ballot gives `mine`, the sensors in this output wave's rows
    // (bit j = sensor j, wave-uniform); the planes then read a sensor's values
    // from its lane (readlane), so a wave holding sensors waits on no memory
    // for them.
    const int wy0 = y0 + wv * RW;
    uint64_t mine = 0;
    int sxl = -1, syl = -1, szl = 0;
    T srcl = T(0);
    if constexpr (MODE != 0) {
        const bool want = is_out && (ADJ ? aux->src != nullptr : aux->trace != nullptr);
        const uint64_t tmask = want ? aux->tile_sensors[tile] : 0;
        if (tmask) {
            if (lane < sens->n) {
                sxl = sens->x[lane];
                syl = sens->y[lane];
                szl = sens->z[lane];
                if constexpr (ADJ) srcl = aux->src[lane];
            }
            mine = __ballot(((tmask >> lane) & 1) && syl >= wy0 && syl < wy0 + RW);
        }
    }

    // Per-plane loads as (field, plane offset, row) slots:
    //   output cells: a = p(xx+2), b = u(xx+1), c = v(xx), d = w(xx), k = K(xx);
    //   halo cells:   a = p(xx), b = p(xx) of the outer row, c = v(xx).
    const T* const fb = out ? u : p;
    const int da = out ? 2 : 0, db = out ? 1 : 0;
    const int64_t rbb = out ? rb : ob;
    struct Fetch {
        V4 a, b, c, d, k, gk;
    };
    const V4 zero4 = {T(0), T(0), T(0), T(0)};
    // NTL: non-temporal loads (streamed past the caches' retention, which then
    // keep what this step writes for the next step to read)
    auto ld = [&](const T* base, int64_t off) -> V4 {
        const V4* a = reinterpret_cast<const V4*>(base + off);
        if constexpr (NTL) return __builtin_nontemporal_load(a);
        else return *a;
    };
    auto fetch = [&](int64_t xx, Fetch& f) {
        if (!real) return;
        const int64_t pb = xx * SX;
        if (!XB || xx + da < X) f.a = ld(p, pb + da * SX + rb);
        if (!XB || xx + db < X) f.b = ld(fb, pb + db * SX + rbb);
        f.c = ld(v, pb + rb);
        if (out) f.d = ld(w, pb + rb);
        if constexpr (KF) {
            if (out) f.k = ld(Kf, pb + rb);
        }
        if constexpr (GKM) {  // the K gradient's inputs: S_t and gK
            if (out) {
                f.k = ld(aux->s_in, pb + rb);
                f.gk = ld(aux->gK, pb + rb);
            }
        }
    };

    // ---- warm-up of the output cells' x queue ---------------------------------
    // Every load of the warm-up and of the first plane is issued before any
    // is used (planes clamped into the mesh, the wall cases chosen by
    // selects), so a workgroup starts after one memory round trip, not one
    // per plane: with short chunks the start is a large part of its run.
    V4 P0 = zero4, Pp1 = zero4, Um1 = zero4, U0 = zero4;
    V4 R0 = zero4, Rp1 = zero4;  // MODE 2: lam_p(x), lam_p(x+1) as loaded (P0, Pp1 hold q)
    V4 a0 = zero4, am1 = zero4, ap1 = zero4, am2 = zero4, um = zero4, uc = zero4;
    if (out) {
        const int64_t xm1 = x0 >= 1 ? x0 - 1 : x0, xm2 = x0 >= 2 ? x0 - 2 : xm1;
        const int64_t xp1 = x0 + 1 < X ? x0 + 1 : x0;
        a0 = ld(p, x0 * SX + rb);
        am1 = ld(p, xm1 * SX + rb);
        ap1 = ld(p, xp1 * SX + rb);
        am2 = ld(p, xm2 * SX + rb);
        um = ld(u, xm1 * SX + rb);
        uc = ld(u, x0 * SX + rb);
    }
    Fetch nx{zero4, zero4, zero4, zero4, zero4, zero4};
    fetch(x0, nx);
    if (out) {
        P0 = a0;
        V4 Pm1 = x0 >= 1 ? am1 : zero4;
        Pp1 = x0 + 1 < X ? ap1 : zero4;
        if constexpr (ADJ) {
            R0 = P0;
            Rp1 = Pp1;
            P0 = qv(P0);
            Pp1 = qv(Pp1);
            Pm1 = qv(Pm1);
        }
        if (x0 >= 1) {
            const V4 Pm2 = (x0 >= 2) ? qv(am2) : Pm1;
#pragma unroll
            for (int i = 0; i < 4; ++i) Um1[i] = vel(Pm2[i], Pm1[i], P0[i], um[i], Ax);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
            U0[i] = vel(x0 >= 1 ? Pm1[i] : P0[i], P0[i], x0 + 1 < X ? Pp1[i] : P0[i], uc[i], Ax);
    }

    for (int64_t x = x0; x < x1; ++x) {
        const Fetch cur = nx;
        if (x + 1 < x1) fetch(x + 1, nx);
        const V4 Qa = qv(cur.a);  // p(x+2) (output cells) or p(x) (halo cells), as q in MODE 2

        // 1. u'(x+1); p into the LDS plane
        V4 Up1 = zero4;
        if (live) {
            if (!XB || x + 1 < X) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    Up1[i] = vel(P0[i], Pp1[i], (!XB || x + 2 < X) ? Qa[i] : Pp1[i], cur.b[i], Ax);
            }
            if (lp.span) SP[own] = P0;
        } else if (halo_v) {
            SP[own] = Qa;
        }
        __syncthreads();

        // 2. v' and w' of the plane; the halo rows' v'
        V4 vn = zero4, wn = zero4;
        if (live) {
            const V4 pyl = SP[ilo];
            const V4 pyh = SP[ihi];
            V4 pzl, pzh;
            z_nbrs_row(P0, zlo, zhi, pzl, pzh);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                vn[i] = vel(pyl[i], P0[i], pyh[i], cur.c[i], Ay);
                wn[i] = vel(pzl[i], P0[i], pzh[i], cur.d[i], Az);
            }
            if (lp.span) SV[own] = vn;
        } else if (halo_v) {
            const V4 in = SP[iin];
            const V4 Qb = qv(cur.b);
            const V4 pl = lower ? Qb : in;
            const V4 ph = lower ? in : Qb;
            V4 vh;
#pragma unroll
            for (int i = 0; i < 4; ++i) vh[i] = vel(pl[i], Qa[i], ph[i], cur.c[i], Ay);
            SV[own] = vh;
        }
        __syncthreads();

        // 3. pressure update and the four outputs
        if (live) {
            const V4 vyl = SV[ilo];
            const V4 vyh = SV[ihi];
            V4 wzl, wzh;
            z_nbrs_row(wn, zlo, zhi, wzl, wzh);
            V4 pn, Sv;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const T fx_lo = (XB && x == 0) ? T(0) : mlin(U0[i], Um1[i]);
                const T fx_hi = (XB && x == X - 1) ? T(0) : mlin(U0[i], Up1[i]);
                const T fy_lo = first ? T(0) : mlin(vn[i], vyl[i]);
                const T fy_hi = last ? T(0) : mlin(vn[i], vyh[i]);
                const T fz_lo = (i == 0 && zlo) ? T(0) : mlin(wn[i], wzl[i]);
                const T fz_hi = (i == 3 && zhi) ? T(0) : mlin(wn[i], wzh[i]);
                T S = mggp(fx_lo, fx_hi, Ax, Vd);
                S = S + mggp(fy_lo, fy_hi, Ay, Vd);
                S = S + mggp(fz_lo, fz_hi, Az, Vd);
                Sv[i] = S;
                if constexpr (ADJ) {
                    pn[i] = dt * S + R0[i];
                } else {
                    const T k = KF ? cur.k[i] : K;
                    pn[i] = dt * (k * S) + P0[i];
                }
            }
            const int64_t o = x * SX + rb;
            if constexpr (MODE != 0) {
                // sensors of this plane: record p' (MODE 1) or add the trace
                // gradient (MODE 2) in the lane that holds the cell
                // (every lane_value is read here, with the whole wave active:
                // a value reloaded from scratch under a narrower exec mask
                // would be stale in the lanes outside it)
                uint64_t m = mine;
                while (m) {
                    const int j = __builtin_ctzll(m);
                    m &= m - 1;
                    const int sx = lane_value(sxl, j), sy = lane_value(syl, j), zz = lane_value(szl, j);
                    T sv = T(0);
                    if constexpr (ADJ) sv = lane_value(srcl, j);
                    if (sx != x) continue;
                    if (out && lp.r == sy - wy0 && zz >= lp.zc && zz < lp.zc + 4) {
                        const int i = zz - lp.zc;
                        if constexpr (ADJ) pn[i] = pn[i] + sv;
                        else aux->trace[j] = pn[i];
                    }
                }
                if constexpr (GKM) {
                    if (out) {
                        V4 gk;
#pragma unroll
                        for (int i = 0; i < 4; ++i) gk[i] = cur.gk[i] + (R0[i] * dt) * cur.k[i];
                        *reinterpret_cast<V4*>(aux->gK + o) = gk;
                    }
                } else if constexpr (MODE == 1) {
                    if (out && aux->s_out) *reinterpret_cast<V4*>(aux->s_out + o) = Sv;
                }
            }
            if (out) {
                put(reinterpret_cast<V4*>(po + o), pn, nt);
                put(reinterpret_cast<V4*>(uo + o), U0, nt);
                put(reinterpret_cast<V4*>(vo + o), vn, nt);
                put(reinterpret_cast<V4*>(wo + o), wn, nt);
            }
            P0 = Pp1;
            Pp1 = Qa;
            if constexpr (ADJ) {
                R0 = Rp1;
                Rp1 = cur.a;
            }
            Um1 = U0;
            U0 = Up1;
        }
    }
}

// One (y tile, x chunk) of the heat step dt * (alpha * lapl(T)) + T in row
// form: TY output waves of RW rows each and two halo waves (rows y0-1 and
// y0+TY*RW) that only put their T row into the LDS plane; x neighbours in a
// per-lane register queue, z neighbours in the lane's registers or the adjacent
// lane's.  The arithmetic is heat_uniform_k's (Grad2Var with neighbour = self
// on the walls).
template <int TY, bool PW, bool XB, bool NTL>
__device__ inline void heat_row_tile(const UniformGeo<float>& g, const float* __restrict__ Tin,
                                     float* __restrict__ Tout, float alpha, float dt, int64_t x0, int64_t x1, int y0,
                                     int wv, typename Row4<float>::type (*sp)[64]) {
    using V4 = typename Row4<float>::type;
    const int64_t X = g.X;
    const int Y = int(g.Y), Z = int(g.Z);
    const int64_t SX = int64_t(Y) * Z;
    const VolDiv<float, PW> Vd{g.vol, g.inv_vol};
    const int lane = threadIdx.x;
    const int RW = rows_per_wave(Z);
    const int zq = Z / 4;
    V4* const SP = &sp[0][0];
    const bool is_out = wv < TY;
    const bool lower = wv == TY;
    const LanePlace lp = is_out ? lane_place(lane, Z, RW) : LanePlace{0, 4 * lane, lane < zq};
    const int y = is_out ? y0 + wv * RW + lp.r : (lower ? y0 - 1 : y0 + TY * RW);
    const bool real = lp.span && y >= 0 && y < Y;
    const bool out = is_out && real;
    const int own = is_out ? (1 + wv * RW) * zq + lane : (lower ? lane : (1 + TY * RW) * zq + lane);
    const bool first = y == 0, last = y == Y - 1;
    const bool zlo = lp.zc == 0, zhi = lp.zc + 4 == Z;
    const int ilo = first ? own : own - zq, ihi = last ? own : own + zq;
    const int64_t rb = int64_t(real ? y : 0) * Z + lp.zc;
    const V4 zero4 = {0.f, 0.f, 0.f, 0.f};
    auto ld = [&](int64_t off) -> V4 {
        const V4* q = reinterpret_cast<const V4*>(Tin + off);
        if constexpr (NTL) return __builtin_nontemporal_load(q);
        else return *q;
    };
    // x queue of the output cells: T(x-1), T(x), T(x+1); the halo cells keep
    // T(x) in T0 only.  nx: the next plane's load (T(x+2), or T(x+1) for halo
    // cells)
    V4 Tm1 = zero4, T0 = zero4, Tp1 = zero4, nx = zero4;
    if (real) {
        T0 = ld(x0 * SX + rb);
        if (out) {
            if (x0 >= 1) Tm1 = ld((x0 - 1) * SX + rb);
            if (x0 + 1 < X) Tp1 = ld((x0 + 1) * SX + rb);
        }
    }
    auto fetch = [&](int64_t xx) {  // the loads for plane xx
        if (!real) return;
        if (out) {
            if (!XB || xx + 1 < X) nx = ld((xx + 1) * SX + rb);
        } else {
            nx = ld(xx * SX + rb);
        }
    };
    if (x0 + 1 < x1) fetch(x0 + 1);
    for (int64_t x = x0; x < x1; ++x) {
        const V4 cur = nx;
        if (x + 2 < x1) fetch(x + 2);
        if (real || (is_out && lp.span)) SP[own] = T0;
        __syncthreads();
        if (out) {
            const V4 ylo = SP[ilo];
            const V4 yhi = SP[ihi];
            V4 zl, zh;
            z_nbrs_row(T0, zlo, zhi, zl, zh);
            V4 r;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float pc = T0[i];
                const bool lo_w[3] = {XB && x == 0, first, i == 0 && zlo};
                const bool hi_w[3] = {XB && x == X - 1, last, i == 3 && zhi};
                const float nb_lo[3] = {lo_w[0] ? pc : Tm1[i], lo_w[1] ? pc : ylo[i], lo_w[2] ? pc : zl[i]};
                const float nb_hi[3] = {hi_w[0] ? pc : Tp1[i], hi_w[1] ? pc : yhi[i], hi_w[2] ? pc : zh[i]};
                float Lp = 0.f;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float al_lo = lo_w[a] ? g.alpha_b[a] : g.alpha[a];
                    const float al_hi = hi_w[a] ? g.alpha_b[a] : g.alpha[a];
                    const float s_lo = (-pc + nb_lo[a]) * al_lo;
                    const float s_hi = (-pc + nb_hi[a]) * al_hi;
                    Lp = Lp + Vd(s_lo * g.area[a] + s_hi * g.area[a]);
                }
                r[i] = dt * (alpha * Lp) + pc;
            }
            *reinterpret_cast<V4*>(Tout + x * SX + rb) = r;
        }
        __syncthreads();  // the plane is rewritten next iteration
        if (out) {
            Tm1 = T0;
            T0 = Tp1;
            Tp1 = cur;
        } else {
            T0 = cur;
        }
    }
}


// Two heat steps per pass in row form.  A workgroup of NW waves holds a band
// of NW*RW consecutive rows; the tile's output rows are the band without its
// two outer rows on each side (rows y0 .. y0+NW*RW-5).  Every lane evaluates
// both steps; the first step is correct on all but the band's outer rows and
// the second on the output rows, which alone are stored.  Iteration x takes
// the first step at plane x and the second at plane x-1, the second reading
// the first's planes x-2 .. x from a register queue and its y-neighbours from
// an LDS plane.  The arithmetic per cell is heat_uniform_k's.
template <class T, int NW, bool PW, bool XB, bool NTL>
__device__ inline void heat_row2_tile(const UniformGeo<T>& g, const T* __restrict__ Tin, T* __restrict__ Tout,
                                      T alpha, T dt, int64_t x0, int64_t x1, int y0, int wv,
                                      typename Row4<T>::type (*sp1)[64], typename Row4<T>::type (*sp2)[64]) {
    using V4 = typename Row4<T>::type;
    const int64_t X = g.X;
    const int Y = int(g.Y), Z = int(g.Z);
    const int64_t SX = int64_t(Y) * Z;
    const VolDiv<T, PW> Vd{g.vol, g.inv_vol};
    const int lane = threadIdx.x;
    const int RW = rows_per_wave(Z);
    const int zq = Z / 4;
    const int band = NW * RW;
    const int TR = band - 4;  // output rows
    V4* const S1 = &sp1[0][0];
    V4* const S2 = &sp2[0][0];
    const LanePlace lp = lane_place(lane, Z, RW);
    const int bi = wv * RW + lp.r;  // row within the band
    const int y = y0 - 2 + bi;
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
    const V4 zero4 = {T(0), T(0), T(0), T(0)};
    auto ld = [&](int64_t off) -> V4 {
        const V4* q = reinterpret_cast<const V4*>(Tin + off);
        if constexpr (NTL) return __builtin_nontemporal_load(q);
        else return *q;
    };
    // one step of a lane's four cells: centre c, x-neighbours (xm, xp), y
    // neighbours (ym, yp) and the walls of the plane
    auto cell4 = [&](const V4& c, const V4& xm, const V4& xpn, const V4& ym, const V4& yp, bool xlo_w, bool xhi_w) {
        V4 zl, zh;
        z_nbrs_row(c, zlo, zhi, zl, zh);
        V4 r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const T pc = c[i];
            const bool lo_w[3] = {xlo_w, first, i == 0 && zlo};
            const bool hi_w[3] = {xhi_w, last, i == 3 && zhi};
            const T nb_lo[3] = {lo_w[0] ? pc : xm[i], lo_w[1] ? pc : ym[i], lo_w[2] ? pc : zl[i]};
            const T nb_hi[3] = {hi_w[0] ? pc : xpn[i], hi_w[1] ? pc : yp[i], hi_w[2] ? pc : zh[i]};
            T Lp = T(0);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const T al_lo = lo_w[a] ? g.alpha_b[a] : g.alpha[a];
                const T al_hi = hi_w[a] ? g.alpha_b[a] : g.alpha[a];
                const T s_lo = (-pc + nb_lo[a]) * al_lo;
                const T s_hi = (-pc + nb_hi[a]) * al_hi;
                Lp = Lp + Vd(s_lo * g.area[a] + s_hi * g.area[a]);
            }
            r[i] = dt * (alpha * Lp) + pc;
        }
        return r;
    };

    const int64_t xs = x0 >= 1 ? x0 - 1 : 0;  // first plane of the first step
    const int64_t xe = x1 + 1;                // end of the iterations
    const int64_t xf = xe < X ? xe : X;       // end of the first step's planes
    // first step's x queue T(x-1), T(x), T(x+1) and the next load T(x+2)
    V4 Tm1 = zero4, T0 = zero4, Tp1 = zero4, nx = zero4;
    if (real) {
        T0 = ld(xs * SX + rb);
        if (xs >= 1) Tm1 = ld((xs - 1) * SX + rb);
        if (xs + 1 < X) Tp1 = ld((xs + 1) * SX + rb);
        if (xs + 1 < xf && (!XB || xs + 2 < X)) nx = ld((xs + 2) * SX + rb);
    }
    // second step's queue: T'(x-2), T'(x-1)
    V4 Q2 = zero4, Q1 = zero4;
    for (int64_t x = xs; x < xe; ++x) {
        const bool s1 = !XB || x < X;
        const int64_t xq = x - 1;  // second step's plane
        const bool s2 = xq >= x0 && xq < x1;
        const V4 cur = nx;
        if (real && x + 2 < xf && (!XB || x + 3 < X)) nx = ld((x + 3) * SX + rb);
        if (live && lp.span) {
            if (s1) S1[own] = T0;
            if (s2) S2[own] = Q1;
        }
        __syncthreads();
        V4 Q0 = zero4;
        if (live) {
            if (s1) Q0 = cell4(T0, Tm1, Tp1, S1[ilo], S1[ihi], XB && x == 0, XB && x == X - 1);
            if (s2) {
                const V4 hi = (!XB || x < X) ? Q0 : Q1;  // T'(xq + 1), self at the wall
                const V4 r = cell4(Q1, Q2, hi, S2[ilo], S2[ihi], XB && xq == 0, XB && xq == X - 1);
                if (out) *reinterpret_cast<V4*>(Tout + xq * SX + rb) = r;
            }
        }
        __syncthreads();  // the planes are rewritten next iteration
        Tm1 = T0;
        T0 = Tp1;
        Tp1 = cur;
        Q2 = Q1;
        Q1 = Q0;
    }
}

}  // namespace row

}  // namespace cfd
