// Compressible Euler with a Rusanov face flux on the hex mesh (no reference
// counterpart; the reference's Sod driver, examples/main_3d_sod_shock.cpp:150-167,
// runs a central predictor with stabilisation instead).
//
// State: five SoA fields (rho, rho u, rho v, rho w, E).  One thread per cell;
// each thread rebuilds the primitive state of itself and its six neighbours,
// evaluates the six face fluxes in the lower-cell -> upper-cell orientation (so
// both cells of a face compute bit-identical fluxes and the scheme stays
// conservative without scatter or atomics on the fields), and updates
//   U <- U - (dt * net) * (1 / V),  net = ((Gx+ - Gx-) + (Gy+ - Gy-)) + (Gz+ - Gz-),
// every product that is added to something rounded once as a fused
// multiply-add, in the same nesting as the oracle's stepper.
//
// Time step, all on the device: the step kernel reads dt from a rotating slot,
// and from the cell's new state computes V / sum_f A_f (|u.n_f| + a), reduces
// the block minimum and folds it into the next slot with one atomic minimum
// per block (non-negative floats order like their bit patterns).  A third slot
// is reset to +inf for the step after, so the loop needs no host round trip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "hexmesh.hpp"
#include "kernels.hpp"
#include "umesh.hpp"

namespace cfd {

template <class T> using Bits = std::conditional_t<sizeof(T) == 4, unsigned int, unsigned long long>;

__device__ inline float from_bits(unsigned int b) { return __uint_as_float(b); }
__device__ inline double from_bits(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ inline unsigned int to_bits(float v) { return __float_as_uint(v); }
__device__ inline unsigned long long to_bits(double v) { return (unsigned long long)__double_as_longlong(v); }
template <class T> __device__ inline Bits<T> inf_bits() { return to_bits(std::numeric_limits<T>::infinity()); }

template <class T> struct EIn { const T* u[5]; };
template <class T> struct EOut { T* u[5]; };

template <class T> struct Prim { T ux, uy, uz, p, a; };

// a * b + c rounded once (the fused multiply-add the scheme's rounding is
// defined with; the oracle uses std::fma in the same places)
template <class T> __device__ inline T fmul_add(T a, T b, T c) {
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}

// u . n of a velocity and a face normal
template <class T> __device__ inline T un_of(T ux, T uy, T uz, T nx, T ny, T nz) {
    return fmul_add(uz, nz, fmul_add(uy, ny, ux * nx));
}

// Primitive state with one division: the reciprocal density multiplies the
// momenta and the sound-speed argument (the oracle's euler_prim, operation
// for operation).
template <class T> __device__ inline Prim<T> prim(T gamma, const T U[5]) {
    Prim<T> q;
    const T ir = T(1) / U[0];
    q.ux = U[1] * ir;
    q.uy = U[2] * ir;
    q.uz = U[3] * ir;
    const T q2 = fmul_add(q.uz, q.uz, fmul_add(q.uy, q.uy, q.ux * q.ux));
    q.p = (gamma - T(1)) * fmul_add(-(T(0.5) * U[0]), q2, U[4]);
    q.a = sqrt((gamma * q.p) * ir);
    return q;
}

// Rusanov flux through a face of unit +axis normal n and area A, L = lower cell
// (the oracle's rusanov_flux, operation for operation)
template <class T>
__device__ inline void rusanov(const T UL[5], const Prim<T>& L, const T UR[5], const Prim<T>& R, T nx, T ny, T nz,
                               T A, T G[5]) {
    const T unL = un_of(L.ux, L.uy, L.uz, nx, ny, nz);
    const T unR = un_of(R.ux, R.uy, R.uz, nx, ny, nz);
    const T FL[5] = {UL[0] * unL, fmul_add(UL[1], unL, L.p * nx), fmul_add(UL[2], unL, L.p * ny),
                     fmul_add(UL[3], unL, L.p * nz), (UL[4] + L.p) * unL};
    const T FR[5] = {UR[0] * unR, fmul_add(UR[1], unR, R.p * nx), fmul_add(UR[2], unR, R.p * ny),
                     fmul_add(UR[3], unR, R.p * nz), (UR[4] + R.p) * unR};
    const T lam = fmax(fabs(unL) + L.a, fabs(unR) + R.a);
    const T hA = T(0.5) * A;
#pragma unroll
    for (int k = 0; k < 5; ++k) G[k] = fmul_add(-lam, UR[k] - UL[k], FL[k] + FR[k]) * hA;
}

// family and family-local index of local face fc of cell (x, y, z)
template <class T>
__device__ inline int64_t fv_face(const FvGeo<T>& g, int64_t x, int64_t y, int64_t z, int fc, int& fam) {
    switch (fc) {
        case 0: fam = 0; return x * g.Y * (g.Z + 1) + y * (g.Z + 1) + z;
        case 1: fam = 0; return x * g.Y * (g.Z + 1) + y * (g.Z + 1) + z + 1;
        case 2: fam = 1; return x * (g.Y + 1) * g.Z + y * g.Z + z;
        case 3: fam = 1; return x * (g.Y + 1) * g.Z + (y + 1) * g.Z + z;
        case 4: fam = 2; return x * g.Y * g.Z + y * g.Z + z;
        default: fam = 2; return (x + 1) * g.Y * g.Z + y * g.Z + z;
    }
}

// V / sum over fc = 0..5 of A_f (|u.n_f| + a)
template <class T>
__device__ inline T local_dt(const FvGeo<T>& g, int64_t x, int64_t y, int64_t z, int64_t c, const Prim<T>& q) {
    T s = T(0);
#pragma unroll
    for (int fc = 0; fc < 6; ++fc) {
        int fam;
        const int64_t f = fv_face(g, x, y, z, fc, fam);
        const T un = un_of(q.ux, q.uy, q.uz, g.n[fam][0][f], g.n[fam][1][f], g.n[fam][2][f]);
        s = fmul_add(g.A[fam][f], fabs(un) + q.a, s);
    }
    return g.vol[c] / s;
}

// The same value from the six faces' geometry already at hand, in the local
// face order z-, z+, y-, y+, x-, x+ (f[fc] = {nx, ny, nz, A})
template <class T> __device__ inline T local_dt_faces(T V, const Prim<T>& q, const T (&f)[6][4]) {
    T s = T(0);
#pragma unroll
    for (int fc = 0; fc < 6; ++fc) {
        const T un = un_of(q.ux, q.uy, q.uz, f[fc][0], f[fc][1], f[fc][2]);
        s = fmul_add(f[fc][3], fabs(un) + q.a, s);
    }
    return V / s;
}

// minimum over the block (64 x blockDim.y threads, one wave per y row), folded
// into *slot by thread (0, 0) with one atomic per block
template <class T> __device__ inline void block_min_into(T v, Bits<T>* slot) {
    __shared__ T red[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    const int lane = threadIdx.x, wave = threadIdx.y;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (lane == 0 && wave == 0) {
        T m = red[0];
        for (int w = 1; w < int(blockDim.y); ++w) m = fmin(m, red[w]);
        if (m < std::numeric_limits<T>::infinity()) atomicMin(slot, to_bits(m));
    }
}

template <class T>
__global__ void __launch_bounds__(256) euler_dt_k(FvGeo<T> g, EIn<T> in, T gamma, Bits<T>* slot) {
    const int64_t z = int64_t(blockIdx.x) * 64 + threadIdx.x;
    const int64_t y = int64_t(blockIdx.y) * 4 + threadIdx.y;
    const int64_t x = blockIdx.z;
    T loc = std::numeric_limits<T>::infinity();
    if (z < g.Z && y < g.Y) {
        const int64_t c = (x * g.Y + y) * g.Z + z;
        T U[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) U[k] = in.u[k][c];
        loc = local_dt(g, x, y, z, c, prim(gamma, U));
    }
    block_min_into(loc, slot);
}

template <class T>
__global__ void __launch_bounds__(256) euler_step_k(FvGeo<T> g, EIn<T> in, EOut<T> out, T gamma, T cfl,
                                                   const Bits<T>* dt_cur, Bits<T>* dt_next, Bits<T>* dt_reset,
                                                   T* dt_hist, int64_t step) {
    const int64_t z = int64_t(blockIdx.x) * 64 + threadIdx.x;
    const int64_t y = int64_t(blockIdx.y) * 4 + threadIdx.y;
    const int64_t x = blockIdx.z;
    const T dt = cfl * from_bits(*dt_cur);
    if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
        *dt_reset = inf_bits<T>();
        if (dt_hist) dt_hist[step] = dt;
    }
    T loc = std::numeric_limits<T>::infinity();
    if (z < g.Z && y < g.Y) {
        const int64_t c = (x * g.Y + y) * g.Z + z;
        T Ui[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) Ui[k] = in.u[k][c];
        const Prim<T> qi = prim(gamma, Ui);
        const int64_t crd[3] = {z, y, x};           // fc >> 1 = 0 z, 1 y, 2 x
        const int64_t ext[3] = {g.Z, g.Y, g.X};
        const int64_t str[3] = {1, g.Z, g.Y * g.Z};
        T net[5] = {T(0), T(0), T(0), T(0), T(0)};
        // axis pairs in the order x, y, z: net = ((x+ - x-) + (y+ - y-)) + (z+ - z-)
#pragma unroll
        for (int pair = 2; pair >= 0; --pair) {
            T Gp[2][5];
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const int fc = 2 * pair + side;
                const bool upper = side == 1;
                const bool wall = upper ? crd[pair] == ext[pair] - 1 : crd[pair] == 0;
                T Un[5];
                if (wall) {
#pragma unroll
                    for (int k = 0; k < 5; ++k) Un[k] = Ui[k];
                    Un[3 - pair] = -Ui[3 - pair];
                } else {
                    const int64_t nb = upper ? c + str[pair] : c - str[pair];
#pragma unroll
                    for (int k = 0; k < 5; ++k) Un[k] = in.u[k][nb];
                }
                const Prim<T> qn = prim(gamma, Un);
                int fam;
                const int64_t f = fv_face(g, x, y, z, fc, fam);
                const T nx = g.n[fam][0][f], ny = g.n[fam][1][f], nz = g.n[fam][2][f], A = g.A[fam][f];
                if (upper) rusanov(Ui, qi, Un, qn, nx, ny, nz, A, Gp[side]);
                else rusanov(Un, qn, Ui, qi, nx, ny, nz, A, Gp[side]);
            }
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const T d = Gp[1][k] - Gp[0][k];
                net[k] = pair == 2 ? d : net[k] + d;
            }
        }
        const T V = g.vol[c];
        const T iV = T(1) / V;
        T Un[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            Un[k] = fmul_add(-(dt * net[k]), iV, Ui[k]);
            out.u[k][c] = Un[k];
        }
        loc = local_dt(g, x, y, z, c, prim(gamma, Un));
    }
    block_min_into(loc, dt_next);
}

// The same step as a 2.5D march: a workgroup owns TY rows x 64 z-columns and
// walks `L` x-planes.  Each cell's state and primitives go to LDS once per
// plane (with a one-cell y/z halo loaded by edge threads), the y/z face fluxes
// read the neighbours from there, and the x-face flux between planes x and x+1
// is carried in registers to the next plane, so each cell converts its state
// to primitives once instead of seven times.  Every face flux is evaluated with
// the lower cell as L, exactly as in euler_step_k, so both give the same bits.
template <class T> __device__ inline void mirror(const T U[5], const Prim<T>& q, int axis, T Um[5], Prim<T>& qm) {
#pragma unroll
    for (int k = 0; k < 5; ++k) Um[k] = U[k];
    Um[1 + axis] = -U[1 + axis];  // axis 0 x, 1 y, 2 z
    qm = q;
    if (axis == 0) qm.ux = -q.ux;
    else if (axis == 1) qm.uy = -q.uy;
    else qm.uz = -q.uz;
}

template <class T, int TY>
__global__ void __launch_bounds__(64 * TY) euler_march_k(FvGeo<T> g, EIn<T> in, EOut<T> out, T gamma, T cfl,
                                                        const Bits<T>* dt_cur, Bits<T>* dt_next, Bits<T>* dt_reset,
                                                        T* dt_hist, int64_t step, int L, int nzt, int nyt) {
    constexpr int TZ = 64;
    constexpr int PZ = TZ + 2;
    constexpr int PS = (TY + 2) * PZ;
    __shared__ T sh[2][10][PS];  // U0..U4, ux, uy, uz, p, a per slot, double-buffered

    const int64_t X = g.X, Y = g.Y, Z = g.Z, SX = Y * Z;
    const int64_t lin = blockIdx.x;
    const int64_t zt = lin % nzt;
    const int64_t yt = (lin / nzt) % nyt;
    const int64_t xc = lin / (int64_t(nzt) * nyt);
    const int lz = threadIdx.x, ly = threadIdx.y;
    const int64_t z = zt * TZ + lz, y = yt * TY + ly;
    const bool valid = z < Z && y < Y;
    const int64_t col = valid ? y * Z + z : 0;
    const int64_t x0 = xc * L;
    const int64_t x1 = (x0 + L < X) ? x0 + L : X;
    const T dt = cfl * from_bits(*dt_cur);
    if (lin == 0 && lz == 0 && ly == 0) {
        *dt_reset = inf_bits<T>();
        if (dt_hist) dt_hist[step] = dt;
    }
    const int own = (ly + 1) * PZ + (lz + 1);
    // halo cells: row -1 (ly 0) / row TY (ly 1) at this column, column -1
    // (lane 0) / column 64 (lane 63) at this row; -1 where outside the domain
    int64_t hy_col = -1, hz_col = -1;
    int hy_slot = 0, hz_slot = 0;
    if (ly < 2 && z < Z) {
        const int64_t yy = ly == 0 ? yt * TY - 1 : yt * TY + TY;
        if (yy >= 0 && yy < Y) {
            hy_col = yy * Z + z;
            hy_slot = (ly == 0 ? 0 : TY + 1) * PZ + (lz + 1);
        }
    }
    if ((lz == 0 || lz == TZ - 1) && y < Y) {
        const int64_t zz = lz == 0 ? zt * TZ - 1 : zt * TZ + TZ;
        if (zz >= 0 && zz < Z) {
            hz_col = y * Z + zz;
            hz_slot = (ly + 1) * PZ + (lz == 0 ? 0 : TZ + 1);
        }
    }
    // Addresses: a pointer uniform over the workgroup (the field plus the
    // plane's offset) and a 32-bit byte offset per lane (the launcher checks
    // that a plane fits 32-bit byte offsets)
    auto at = [](const T* base, uint32_t boff) {
        return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
    };
    const uint32_t bcol = uint32_t(col) * uint32_t(sizeof(T));
    const uint32_t bzf = valid ? uint32_t(y * (Z + 1) + z) * uint32_t(sizeof(T)) : 0;  // z-face of (y, z)
    auto load = [&](int64_t plane, uint32_t boff, T U[5]) {
#pragma unroll
        for (int k = 0; k < 5; ++k) U[k] = at(in.u[k] + plane * SX, boff);
    };
    auto put = [&](T* const* S, int slot, const T U[5], const Prim<T>& q) {
#pragma unroll
        for (int k = 0; k < 5; ++k) S[k][slot] = U[k];
        S[5][slot] = q.ux; S[6][slot] = q.uy; S[7][slot] = q.uz; S[8][slot] = q.p; S[9][slot] = q.a;
    };
    auto get = [&](T* const* S, int slot, T U[5], Prim<T>& q) {
#pragma unroll
        for (int k = 0; k < 5; ++k) U[k] = S[k][slot];
        q.ux = S[5][slot]; q.uy = S[6][slot]; q.uz = S[7][slot]; q.p = S[8][slot]; q.a = S[9][slot];
    };
    // face geometry: family fam, faces from `first` (uniform), lane offset boff
    auto face4 = [&](int fam, int64_t first, uint32_t boff, T (&o)[4]) {
        o[0] = at(g.n[fam][0] + first, boff);
        o[1] = at(g.n[fam][1] + first, boff);
        o[2] = at(g.n[fam][2] + first, boff);
        o[3] = at(g.A[fam] + first, boff);
    };

    T Uc[5] = {}, Un[5] = {}, Gxl[5] = {};
    T fg[6][4] = {};  // this cell's face geometry in local face order; fg[4] carried along x
    Prim<T> qc{}, qn{};
    if (valid) {  // warm-up: state of plane x0 and the flux through its lower x-face
        load(x0, bcol, Uc);
        qc = prim(gamma, Uc);
        face4(2, x0 * SX, bcol, fg[4]);
        T Ul[5];
        Prim<T> ql;
        if (x0 == 0) mirror(Uc, qc, 0, Ul, ql);
        else {
            load(x0 - 1, bcol, Ul);
            ql = prim(gamma, Ul);
        }
        rusanov(Ul, ql, Uc, qc, fg[4][0], fg[4][1], fg[4][2], fg[4][3], Gxl);
    }
    T loc = std::numeric_limits<T>::infinity();
    int b = 0;
#pragma unroll 1
    for (int64_t x = x0; x < x1; ++x, b ^= 1) {
        T* S[10];
#pragma unroll
        for (int k = 0; k < 10; ++k) S[k] = sh[b][k];
        if (valid && x + 1 < X) {
            load(x + 1, bcol, Un);
            qn = prim(gamma, Un);
        }
        if (hy_col >= 0) {
            T Uh[5];
            load(x, uint32_t(hy_col) * uint32_t(sizeof(T)), Uh);
            put(S, hy_slot, Uh, prim(gamma, Uh));
        }
        if (hz_col >= 0) {
            T Uh[5];
            load(x, uint32_t(hz_col) * uint32_t(sizeof(T)), Uh);
            put(S, hz_slot, Uh, prim(gamma, Uh));
        }
        if (valid) put(S, own, Uc, qc);
        __syncthreads();
        if (valid) {
            T Um[5], G1[5], G2[5], acc[5];
            Prim<T> qm;
            // net = ((Gx+ - Gx-) + (Gy+ - Gy-)) + (Gz+ - Gz-), accumulated one
            // axis at a time; the upper x flux becomes the next plane's lower.
            // The six faces' geometry stays in registers for the next dt.
            face4(2, (x + 1) * SX, bcol, fg[5]);
            if (x == X - 1) {
                mirror(Uc, qc, 0, Um, qm);
                rusanov(Uc, qc, Um, qm, fg[5][0], fg[5][1], fg[5][2], fg[5][3], G1);
            } else {
                rusanov(Uc, qc, Un, qn, fg[5][0], fg[5][1], fg[5][2], fg[5][3], G1);
            }
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                acc[k] = G1[k] - Gxl[k];
                Gxl[k] = G1[k];
            }
            // y faces
            face4(1, x * (Y + 1) * Z, bcol, fg[2]);
            if (y == 0) mirror(Uc, qc, 1, Um, qm);
            else get(S, own - PZ, Um, qm);
            rusanov(Um, qm, Uc, qc, fg[2][0], fg[2][1], fg[2][2], fg[2][3], G1);
            face4(1, x * (Y + 1) * Z + Z, bcol, fg[3]);
            if (y == Y - 1) mirror(Uc, qc, 1, Um, qm);
            else get(S, own + PZ, Um, qm);
            rusanov(Uc, qc, Um, qm, fg[3][0], fg[3][1], fg[3][2], fg[3][3], G2);
#pragma unroll
            for (int k = 0; k < 5; ++k) acc[k] = acc[k] + (G2[k] - G1[k]);
            // z faces
            const int64_t fz = x * Y * (Z + 1);
            face4(0, fz, bzf, fg[0]);
            if (z == 0) mirror(Uc, qc, 2, Um, qm);
            else get(S, own - 1, Um, qm);
            rusanov(Um, qm, Uc, qc, fg[0][0], fg[0][1], fg[0][2], fg[0][3], G1);
            face4(0, fz + 1, bzf, fg[1]);
            if (z == Z - 1) mirror(Uc, qc, 2, Um, qm);
            else get(S, own + 1, Um, qm);
            rusanov(Uc, qc, Um, qm, fg[1][0], fg[1][1], fg[1][2], fg[1][3], G2);
            const T V = at(g.vol + x * SX, bcol);
            const T iV = T(1) / V;
            T Unew[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const T net = acc[k] + (G2[k] - G1[k]);
                Unew[k] = fmul_add(-(dt * net), iV, Uc[k]);
                *reinterpret_cast<T*>(reinterpret_cast<char*>(out.u[k] + x * SX) + bcol) = Unew[k];
            }
            loc = fmin(loc, local_dt_faces(V, prim(gamma, Unew), fg));
#pragma unroll
            for (int k = 0; k < 5; ++k) Uc[k] = Un[k];
            qc = qn;
#pragma unroll
            for (int k = 0; k < 4; ++k) fg[4][k] = fg[5][k];
        }
    }
    block_min_into(loc, dt_next);
}

template <class B> __global__ void fill_bits_k(B* p, int n, B v) {
    if (int(threadIdx.x) < n) p[threadIdx.x] = v;
}

// Device geometry of the stepper, built on first use from the mesh's vertices.
template <class T> static FvGeo<T>& fv_geometry(cfd_mesh_s* m, hipStream_t s) {
    FvGeo<T>& fv = [&]() -> FvGeo<T>& {
        if constexpr (sizeof(T) == 4) return m->fv32; else return m->fv64;
    }();
    if (m->fv_block) return fv;
    HexGeometry<T> g;
    build_hex_geometry<T>(g, m->X, m->Y, m->Z, T(m->dx), T(m->dy), T(m->dz),
                          m->has_offsets ? m->vertex_offsets.data() : nullptr);
    std::vector<T> host;
    int64_t off_fam[3], off_vol;
    build_fv_geometry<T>(g, host, off_fam, off_vol);
    const int64_t bytes = int64_t(host.size() * sizeof(T));
    void* dev = pool_alloc(m->ctx, bytes, s);
    hipError_t e = hipMemcpyAsync(dev, host.data(), size_t(bytes), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        pool_free(m->ctx, dev, s);
        CFD_HIP(e);
    }
    T* d = (T*)dev;
    fv.X = m->X; fv.Y = m->Y; fv.Z = m->Z;
    for (int fam = 0; fam < 3; ++fam) {
        const int64_t nf = g.nf[fam];
        for (int k = 0; k < 3; ++k) fv.n[fam][k] = d + off_fam[fam] + k * nf;
        fv.A[fam] = d + off_fam[fam] + 3 * nf;
    }
    fv.vol = d + off_vol;
    m->fv_block = dev;
    m->fv_bytes = bytes;
    return fv;
}

template <class T>
void launch_euler(cfd_mesh_s* m, void* const a[5], void* const b[5], double gamma, double cfl, int64_t nsteps,
                  void* dt_hist, hipStream_t s) {
    const FvGeo<T>& g = fv_geometry<T>(m, s);
    Bits<T>* slots = (Bits<T>*)pool_alloc(m->ctx, 3 * sizeof(Bits<T>), s);
    try {
        const dim3 block(64, 4, 1);
        const dim3 grid(unsigned((m->Z + 63) / 64), unsigned((m->Y + 3) / 4), unsigned(m->X));
        const Bits<T> inf = sizeof(T) == 4 ? Bits<T>(0x7f800000u) : Bits<T>(0x7ff0000000000000ull);
        fill_bits_k<Bits<T>><<<1, 64, 0, s>>>(slots, 3, inf);
        EIn<T> in{};
        EOut<T> out{};
        for (int k = 0; k < 5; ++k) in.u[k] = (const T*)a[k];
        euler_dt_k<T><<<grid, block, 0, s>>>(g, in, T(gamma), slots);
        CFD_HIP(hipGetLastError());
        const int ty = m->euler_ty, L = m->euler_L;
        const int nzt = int((m->Z + 63) / 64), nyt = int((m->Y + ty - 1) / ty);
        CFD_REQUIRE(m->euler_variant == 0 || m->Y * (m->Z + 1) * int64_t(sizeof(T)) < (int64_t(1) << 32),
                    "Euler march: a plane of this mesh exceeds 32-bit byte offsets (use the per-thread form)");
        const int64_t nwg = int64_t(nzt) * nyt * ((m->X + L - 1) / L);
        CFD_REQUIRE(nwg < (int64_t(1) << 31), "euler march grid too large");
        for (int64_t st = 0; st < nsteps; ++st) {
            void* const* src = (st & 1) ? b : a;
            void* const* dst = (st & 1) ? a : b;
            for (int k = 0; k < 5; ++k) {
                in.u[k] = (const T*)src[k];
                out.u[k] = (T*)dst[k];
            }
            Bits<T>* cur = slots + (st % 3);
            Bits<T>* nxt = slots + ((st + 1) % 3);
            Bits<T>* rst = slots + ((st + 2) % 3);
            if (m->euler_variant == 0) {
                euler_step_k<T><<<grid, block, 0, s>>>(g, in, out, T(gamma), T(cfl), cur, nxt, rst, (T*)dt_hist,
                                                       st);
                continue;
            }
            const dim3 mgrid{unsigned(nwg), 1u, 1u};
#define CFD_EULER_MARCH(TYV)                                                                                    \
    euler_march_k<T, TYV><<<mgrid, dim3(64, TYV), 0, s>>>(g, in, out, T(gamma), T(cfl), cur, nxt, rst, (T*)dt_hist, \
                                                         st, L, nzt, nyt)
            switch (ty) {
                case 2: CFD_EULER_MARCH(2); break;
                case 4: CFD_EULER_MARCH(4); break;
                case 8:
                    // FP32 only: the FP64 8-row build spills registers, and the
                    // kernel's dt reduction reads other lanes' registers
                    if constexpr (sizeof(T) == 4) {
                        CFD_EULER_MARCH(8);
                        break;
                    }
                    [[fallthrough]];
                default: throw Error(CFD_E_INVALID, "euler march tile rows must be 2, 4 or 8 (2 or 4 in FP64)");
            }
#undef CFD_EULER_MARCH
        }
        CFD_HIP(hipGetLastError());
    } catch (...) {
        pool_free(m->ctx, slots, s);
        throw;
    }
    pool_free(m->ctx, slots, s);
}

// ---------------------------------------------------------------------------
// The same stepper on a mesh given by its connectivity (cfd_mesh_create_
// unstructured): one thread per cell in the mesh's reverse Cuthill-McKee
// numbering, the neighbours through the (local face, cell) ids, the
// stepper's geometry per face (build_ufv_geometry).  Each cell evaluates its
// six face fluxes with the lower cell of the face's axis pair as L, in the
// per-thread kernel's order, so results equal the oracle's on the same
// connectivity.
// ---------------------------------------------------------------------------

__device__ inline int64_t ucell() { return int64_t(blockIdx.x) * 256 + threadIdx.y * 64 + threadIdx.x; }

// V / sum over fc = 0..5 of A_f (|u.n_f| + a), the faces through the ids
template <class T> __device__ inline T ulocal_dt(const UnstructGeo<T>& g, int64_t i, const Prim<T>& q) {
    T s = T(0);
#pragma unroll
    for (int fc = 0; fc < 6; ++fc) {
        const int32_t f = g.face[int64_t(fc) * g.N + i];
        const T un = un_of(q.ux, q.uy, q.uz, g.sn[0][f], g.sn[1][f], g.sn[2][f]);
        s = fmul_add(g.sa[f], fabs(un) + q.a, s);
    }
    return g.svol[i] / s;
}

template <class T>
__global__ void __launch_bounds__(256) ueuler_dt_k(UnstructGeo<T> g, EIn<T> in, T gamma, Bits<T>* slot) {
    const int64_t i = ucell();
    T loc = std::numeric_limits<T>::infinity();
    if (i < g.N) {
        T U[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) U[k] = in.u[k][i];
        loc = ulocal_dt(g, i, prim(gamma, U));
    }
    block_min_into(loc, slot);
}

template <class T>
__global__ void __launch_bounds__(256) ueuler_step_k(UnstructGeo<T> g, EIn<T> in, EOut<T> out, T gamma, T cfl,
                                                    const Bits<T>* dt_cur, Bits<T>* dt_next, Bits<T>* dt_reset,
                                                    T* dt_hist, int64_t step) {
    const int64_t i = ucell();
    const T dt = cfl * from_bits(*dt_cur);
    if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
        *dt_reset = inf_bits<T>();
        if (dt_hist) dt_hist[step] = dt;
    }
    T loc = std::numeric_limits<T>::infinity();
    if (i < g.N) {
        T Ui[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) Ui[k] = in.u[k][i];
        const Prim<T> qi = prim(gamma, Ui);
        T net[5] = {T(0), T(0), T(0), T(0), T(0)};
        // axis pairs in the order x, y, z: net = ((x+ - x-) + (y+ - y-)) + (z+ - z-)
#pragma unroll
        for (int pair = 2; pair >= 0; --pair) {
            T Gp[2][5];
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const int fc = 2 * pair + side;
                const int64_t k6 = int64_t(fc) * g.N + i;
                const int32_t nb = g.nbr[k6];
                T Un[5];
                if (nb < 0) {  // wall: the mirrored state
#pragma unroll
                    for (int k = 0; k < 5; ++k) Un[k] = Ui[k];
                    Un[3 - pair] = -Ui[3 - pair];
                } else {
#pragma unroll
                    for (int k = 0; k < 5; ++k) Un[k] = in.u[k][nb];
                }
                const Prim<T> qn = prim(gamma, Un);
                const int32_t f = g.face[k6];
                const T nx = g.sn[0][f], ny = g.sn[1][f], nz = g.sn[2][f], A = g.sa[f];
                if (side == 1) rusanov(Ui, qi, Un, qn, nx, ny, nz, A, Gp[side]);
                else rusanov(Un, qn, Ui, qi, nx, ny, nz, A, Gp[side]);
            }
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const T d = Gp[1][k] - Gp[0][k];
                net[k] = pair == 2 ? d : net[k] + d;
            }
        }
        const T iV = T(1) / g.svol[i];
        T Un[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            Un[k] = fmul_add(-(dt * net[k]), iV, Ui[k]);
            out.u[k][i] = Un[k];
        }
        loc = ulocal_dt(g, i, prim(gamma, Un));
    }
    block_min_into(loc, dt_next);
}

namespace {

template <class T> UnstructGeo<T>& ugeo_e(cfd_mesh_s* m) {
    if constexpr (sizeof(T) == 4) return m->uu32; else return m->uu64;
}

// the stepper's geometry of an unstructured mesh, built and uploaded on first
// use into one pool block (faces and cells in the mesh's numbering)
template <class T> const UnstructGeo<T>& ufv_geometry(cfd_mesh_s* m, hipStream_t s) {
    UnstructGeo<T>& u = ugeo_e<T>(m);
    if (u.svol) return u;
    const UConn& c = m->uconn;
    UGeometry<T> g;
    build_ugeometry<T>(g, c);
    std::vector<T> sv, svol;
    build_ufv_geometry<T>(c, g, sv, svol);
    const int64_t N = c.N, F = c.F;
    std::vector<T> host(size_t(4 * F + N));
    for (int64_t n = 0; n < F; ++n) {
        const int64_t f = m->face_perm[n];
        for (int k = 0; k < 4; ++k) host[size_t(k * F + n)] = sv[size_t(k * F + f)];
    }
    for (int64_t n = 0; n < N; ++n) host[size_t(4 * F + n)] = svol[size_t(m->cell_perm[n])];
    const int64_t bytes = int64_t(host.size() * sizeof(T));
    void* dev = pool_alloc(m->ctx, bytes, s);
    hipError_t e = hipMemcpyAsync(dev, host.data(), size_t(bytes), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        pool_free(m->ctx, dev, s);
        CFD_HIP(e);
    }
    const T* d = static_cast<const T*>(dev);
    for (int k = 0; k < 3; ++k) u.sn[k] = d + k * F;
    u.sa = d + 3 * F;
    u.svol = d + 4 * F;
    m->fv_block = dev;
    m->fv_bytes = bytes;
    return u;
}

}  // namespace

template <class T>
void launch_ueuler(cfd_mesh_s* m, void* const a[5], void* const b[5], double gamma, double cfl, int64_t nsteps,
                   void* dt_hist, hipStream_t s) {
    const UnstructGeo<T>& g = ufv_geometry<T>(m, s);
    Bits<T>* slots = (Bits<T>*)pool_alloc(m->ctx, 3 * sizeof(Bits<T>), s);
    try {
        const dim3 block(64, 4, 1);
        const dim3 grid(unsigned((m->N + 255) / 256));
        const Bits<T> inf = sizeof(T) == 4 ? Bits<T>(0x7f800000u) : Bits<T>(0x7ff0000000000000ull);
        fill_bits_k<Bits<T>><<<1, 64, 0, s>>>(slots, 3, inf);
        EIn<T> in{};
        EOut<T> out{};
        for (int k = 0; k < 5; ++k) in.u[k] = (const T*)a[k];
        ueuler_dt_k<T><<<grid, block, 0, s>>>(g, in, T(gamma), slots);
        CFD_HIP(hipGetLastError());
        for (int64_t st = 0; st < nsteps; ++st) {
            void* const* src = (st & 1) ? b : a;
            void* const* dst = (st & 1) ? a : b;
            for (int k = 0; k < 5; ++k) {
                in.u[k] = (const T*)src[k];
                out.u[k] = (T*)dst[k];
            }
            ueuler_step_k<T><<<grid, block, 0, s>>>(g, in, out, T(gamma), T(cfl), slots + (st % 3),
                                                    slots + ((st + 1) % 3), slots + ((st + 2) % 3), (T*)dt_hist, st);
            CFD_HIP(hipGetLastError());
        }
    } catch (...) {
        pool_free(m->ctx, slots, s);
        throw;
    }
    pool_free(m->ctx, slots, s);
}

template void launch_ueuler<float>(cfd_mesh_s*, void* const[5], void* const[5], double, double, int64_t, void*,
                                   hipStream_t);
template void launch_ueuler<double>(cfd_mesh_s*, void* const[5], void* const[5], double, double, int64_t, void*,
                                    hipStream_t);

template void launch_euler<float>(cfd_mesh_s*, void* const[5], void* const[5], double, double, int64_t, void*,
                                  hipStream_t);
template void launch_euler<double>(cfd_mesh_s*, void* const[5], void* const[5], double, double, int64_t, void*,
                                   hipStream_t);

}  // namespace cfd
