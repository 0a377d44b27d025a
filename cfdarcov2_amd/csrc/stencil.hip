// Stencil kernels of the explicit finite-volume step on the structured hex
// mesh: Green-Gauss gradient, Laplacian, stabilisation, and the fused heat and
// acoustic steps.  Two geometry forms:
//   uniform  - exactly uniform meshes; geometry is a handful of constants and
//              neighbours are implicit, so only state crosses HBM;
//   general  - per-face SoA geometry (face families z, y, x) for perturbed
//              meshes; neighbours are still implicit (structured topology).
// All arithmetic follows the reference operation order (see the oracle and
// src/fvm3d.cpp / include/grad_utils.hpp cited per kernel) and the library is
// built with -ffp-contract=off, so FP64 results are bit-identical to a CPU
// evaluation in the same order.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace cfd {

// ---------------------------------------------------------------------------
// indexing helpers
// ---------------------------------------------------------------------------
struct Cell {
    int64_t x, y, z, id;
    bool valid;
};

template <class G> __device__ inline Cell cell_of_thread(const G& g) {
    Cell c;
    c.z = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    c.y = int64_t(blockIdx.y) * blockDim.y + threadIdx.y;
    c.x = int64_t(blockIdx.z);
    c.valid = c.z < g.Z && c.y < g.Y && c.x < g.X;
    c.id = (c.x * g.Y + c.y) * g.Z + c.z;
    return c;
}

// Neighbour of cell c across its lower / upper face on axis a (0 = x, 1 = y,
// 2 = z); the cell itself on a boundary face (reference rule).
template <class G> __device__ inline int64_t nb_lo(const G& g, const Cell& c, int a) {
    if (a == 0) return c.x > 0 ? c.id - g.Y * g.Z : c.id;
    if (a == 1) return c.y > 0 ? c.id - g.Z : c.id;
    return c.z > 0 ? c.id - 1 : c.id;
}
template <class G> __device__ inline int64_t nb_hi(const G& g, const Cell& c, int a) {
    if (a == 0) return c.x < g.X - 1 ? c.id + g.Y * g.Z : c.id;
    if (a == 1) return c.y < g.Y - 1 ? c.id + g.Z : c.id;
    return c.z < g.Z - 1 ? c.id + 1 : c.id;
}
template <class G> __device__ inline bool at_lo(const G& g, const Cell& c, int a) {
    return a == 0 ? c.x == 0 : a == 1 ? c.y == 0 : c.z == 0;
}
template <class G> __device__ inline bool at_hi(const G& g, const Cell& c, int a) {
    return a == 0 ? c.x == g.X - 1 : a == 1 ? c.y == g.Y - 1 : c.z == g.Z - 1;
}

// Face-family index of the lower / upper face of cell c on axis a.
template <class T> __device__ inline int64_t face_lo(const GeneralGeo<T>& g, const Cell& c, int a) {
    if (a == 2) return c.x * g.Y * (g.Z + 1) + c.y * (g.Z + 1) + c.z;
    if (a == 1) return c.x * (g.Y + 1) * g.Z + c.y * g.Z + c.z;
    return c.x * g.Y * g.Z + c.y * g.Z + c.z;
}
template <class T> __device__ inline int64_t face_hi(const GeneralGeo<T>& g, const Cell& c, int a) {
    if (a == 2) return face_lo(g, c, a) + 1;
    if (a == 1) return face_lo(g, c, a) + g.Z;
    return face_lo(g, c, a) + g.Y * g.Z;
}
// local face fc -> axis: fc 0,1 z; 2,3 y; 4,5 x
__device__ inline int fc_axis(int fc) { return 2 - (fc >> 1); }
__device__ inline int axis_family(int a) { return 2 - a; }  // family 0 = z-faces, 1 = y, 2 = x

// ---------------------------------------------------------------------------
// uniform-mesh building blocks
// ---------------------------------------------------------------------------

// x / V on a uniform mesh.  For a power-of-two volume the quotient equals the
// product with the exact reciprocal bit for bit (both are the correctly
// rounded value of the same real number), which is one instruction instead of
// a division sequence; the branch is uniform across the launch.
template <class T> __device__ inline T udiv(const UniformGeo<T>& g, T x) {
    return g.vol_pow2 ? x * g.inv_vol : x / g.vol;
}

// Green-Gauss component along axis a from the two face values on that axis:
// acc = (f_lo * -1) * A + (f_hi * +1) * A, then / V (gauss_grad_cu_k order,
// include/custom_cuda_functions.hpp:277-295; the other four faces add zeros).
template <class T> __device__ inline T gg_pair_u(const UniformGeo<T>& g, T f_lo, T f_hi, T A) {
    const T t_lo = (-f_lo) * A;
    const T t_hi = f_hi * A;
    return udiv(g, t_lo + t_hi);
}

// Linear face interpolation, (phi_c + phi_nb) * 0.5 (grad_utils.hpp:23-36).
template <class T> __device__ inline T lin(T c, T nb) { return (c + nb) * T(0.5); }

// ---------------------------------------------------------------------------
// GradVar / Grad2Var / StabVar operator kernels
// ---------------------------------------------------------------------------

template <class T>
__global__ void __launch_bounds__(256) grad_uniform_k(UniformGeo<T> g, const T* __restrict__ phi,
                                                      T* __restrict__ out, int mask, int wall_axis) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    const T pc = phi[c.id];
    T res = T(0);
    for (int a = 0; a < 3; ++a) {
        if (!(mask & (1 << a))) continue;
        T f_lo = lin(pc, phi[nb_lo(g, c, a)]);
        T f_hi = lin(pc, phi[nb_hi(g, c, a)]);
        if (wall_axis == a) {
            if (at_lo(g, c, a)) f_lo = T(0);
            if (at_hi(g, c, a)) f_hi = T(0);
        }
        res = res + gg_pair_u(g, f_lo, f_hi, g.area[a]);
    }
    out[c.id] = res;
}

template <class T>
__global__ void __launch_bounds__(256) lapl_uniform_k(UniformGeo<T> g, const T* __restrict__ phi,
                                                      T* __restrict__ out, int mask) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    const T pc = phi[c.id];
    T res = T(0);
    for (int a = 0; a < 3; ++a) {
        if (!(mask & (1 << a))) continue;
        const T al_lo = at_lo(g, c, a) ? g.alpha_b[a] : g.alpha[a];
        const T al_hi = at_hi(g, c, a) ? g.alpha_b[a] : g.alpha[a];
        const T s_lo = (-pc + phi[nb_lo(g, c, a)]) * al_lo;
        const T s_hi = (-pc + phi[nb_hi(g, c, a)]) * al_hi;
        // axis normals are +1 on both faces of the pair
        res = res + udiv(g, s_lo * g.area[a] + s_hi * g.area[a]);
    }
    out[c.id] = res;
}

// One axis d of StabVar on a uniform mesh (src/fvm3d.cpp:909-936), added to
// res.  Second stage: upwind face values phi + (len * g_d) * n_out_d, jump
// against the neighbour's opposite face (self's opposite face on the
// boundary), 0.5 * jump, Gauss with axis normals for component d.  The first
// stage's gradient components it reads (own and the two axis neighbours') are
// recomputed here from phi at five points on the axis: pc at the cell, pl / ph
// at its lower / upper neighbour, pll / plh at the lower neighbour's own lower
// / upper neighbour, phl / phh likewise for the upper one (self on walls).
template <class T>
__device__ inline T stab_axis_u(const UniformGeo<T>& g, T res, int d, T pc, T pl, T ph, T pll, T plh, T phl,
                                T phh) {
    const T L = g.len[d];
    const T gc = gg_pair_u(g, lin(pc, pl), lin(pc, ph), g.area[d]);
    const T glo = gg_pair_u(g, lin(pl, pll), lin(pl, plh), g.area[d]);
    const T ghi = gg_pair_u(g, lin(ph, phl), lin(ph, phh), g.area[d]);
    const T own_lo = pc + (L * gc) * T(-1);
    const T col_lo = pl + (L * glo) * T(1);
    const T own_hi = pc + (L * gc) * T(1);
    const T col_hi = ph + (L * ghi) * T(-1);
    const T f_lo = T(0.5) * (own_lo - col_lo);
    const T f_hi = T(0.5) * (own_hi - col_hi);
    return res + udiv(g, f_lo * g.area[d] + f_hi * g.area[d]);
}

// The five points of one axis from a fetch function P(offset) (offset in
// cells along the axis, -2..2), coordinate k on an axis of n cells.
template <class T, class F>
__device__ inline T stab_axis_at(const UniformGeo<T>& g, T res, int d, int64_t k, int64_t n, T pc, F P) {
    const bool wl = k == 0, wh = k == n - 1;
    const T pl = wl ? pc : P(-1);
    const T ph = wh ? pc : P(1);
    const T pll = wl ? pc : (k == 1 ? pl : P(-2));
    const T plh = wl ? ph : pc;
    const T phh = wh ? pc : (k == n - 2 ? ph : P(2));
    const T phl = wh ? pl : pc;
    return stab_axis_u(g, res, d, pc, pl, ph, pll, plh, phl, phh);
}

// StabVar on a uniform mesh in one pass, as a march along x: a workgroup
// holds TY rows of 64 z-cells and walks `L` planes.  Each thread keeps its
// column's values at x-2 .. x+2 in registers; the plane x goes to LDS with
// two halo rows and columns on each side, so the y and z points come from
// there; the halo values of the next plane are loaded into registers while
// the current one is computed.  Every value passed to the arithmetic is the
// one the cell-indexed rule picks (self on walls), so the result is the
// per-cell evaluation's.
template <class T, int TY>
__global__ void __launch_bounds__(64 * TY) stab_march_uniform_k(UniformGeo<T> g, const T* __restrict__ phi,
                                                                T* __restrict__ out, int mask, int L, int nzt,
                                                                int nyt) {
    static_assert(TY >= 4, "halo rows come from two distinct leading and trailing rows");
    __shared__ T sh[TY + 4][64 + 4];
    const int64_t X = g.X, Y = g.Y, Z = g.Z, SX = Y * Z;
    const int64_t lin_ = blockIdx.x;
    const int64_t zt = lin_ % nzt, yt = (lin_ / nzt) % nyt, xc = lin_ / (int64_t(nzt) * nyt);
    const int lz = threadIdx.x, ly = threadIdx.y;
    const int64_t z = zt * 64 + lz, y = yt * TY + ly;
    const bool valid = z < Z && y < Y;
    const int64_t col = valid ? y * Z + z : 0;
    const int64_t x0 = xc * L, x1 = (x0 + L < X) ? x0 + L : X;
    // halo rows (threads of the first / last two rows) and columns (lanes
    // 0, 1, 62, 63) of this thread, -1 where outside the mesh
    int64_t hy = -1, hz = -1;
    int hy_row = 0, hz_col = 0;
    if (z < Z) {
        int64_t yy = -1;
        if (ly < 2) { yy = yt * TY - 2 + ly; hy_row = ly; }
        else if (ly >= TY - 2) { yy = yt * TY + TY + (ly - (TY - 2)); hy_row = ly + 4; }
        if (yy >= 0 && yy < Y) hy = yy * Z + z;
    }
    if (y < Y) {
        int64_t zz = -1;
        if (lz < 2) { zz = zt * 64 - 2 + lz; hz_col = lz; }
        else if (lz >= 62) { zz = zt * 64 + 64 + (lz - 62); hz_col = lz + 4; }
        if (zz >= 0 && zz < Z) hz = y * Z + zz;
    }
    T w[5] = {T(0), T(0), T(0), T(0), T(0)};  // column values at x-2 .. x+2
    if (valid)
        for (int k = 0; k < 5; ++k) {
            const int64_t xx = x0 - 2 + k;
            if (xx >= 0 && xx < X) w[k] = phi[xx * SX + col];
        }
    // halo values of plane x (x0 < x1 <= X whenever the block has planes)
    T hyv = (hy >= 0 && x0 < x1) ? phi[x0 * SX + hy] : T(0);
    T hzv = (hz >= 0 && x0 < x1) ? phi[x0 * SX + hz] : T(0);
    for (int64_t x = x0; x < x1; ++x) {
        const int64_t pb = x * SX;
        if (valid) sh[ly + 2][lz + 2] = w[2];
        if (hy >= 0) sh[hy_row][lz + 2] = hyv;
        if (hz >= 0) sh[ly + 2][hz_col] = hzv;
        if (x + 1 < x1) {
            if (hy >= 0) hyv = phi[pb + SX + hy];
            if (hz >= 0) hzv = phi[pb + SX + hz];
        }
        __syncthreads();
        if (valid) {
            const T pc = w[2];
            T res = T(0);
            if (mask & 1) res = stab_axis_at(g, res, 0, x, X, pc, [&](int o) { return w[2 + o]; });
            if (mask & 2) res = stab_axis_at(g, res, 1, y, Y, pc, [&](int o) { return sh[ly + 2 + o][lz + 2]; });
            if (mask & 4) res = stab_axis_at(g, res, 2, z, Z, pc, [&](int o) { return sh[ly + 2][lz + 2 + o]; });
            out[pb + col] = res;
            w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = w[4];
            w[4] = (x + 3 < X) ? phi[(x + 3) * SX + col] : T(0);
        }
        __syncthreads();
    }
}

// --- general (per-face geometry) forms -------------------------------------

// value of face fc's stored axis normal component d and the outward sign
template <class T>
__device__ inline void face_geo(const GeneralGeo<T>& g, const Cell& c, int fc, int64_t& f, int& fam) {
    const int a = fc_axis(fc);
    fam = axis_family(a);
    f = (fc & 1) ? face_hi(g, c, a) : face_lo(g, c, a);
}
template <class G> __device__ inline int64_t nb_of(const G& g, const Cell& c, int fc) {
    const int a = fc_axis(fc);
    return (fc & 1) ? nb_hi(g, c, a) : nb_lo(g, c, a);
}
template <class G> __device__ inline bool is_bnd_face(const G& g, const Cell& c, int fc) {
    const int a = fc_axis(fc);
    return (fc & 1) ? at_hi(g, c, a) : at_lo(g, c, a);
}

// Gauss sum of six face values for component d, fc = 0..5 from 0, outward
// (use_alt=false in the reference's inverted flag) or axis normals.
template <class T>
__device__ inline T gauss_general(const GeneralGeo<T>& g, const Cell& c, const T fv[6], int d, bool axis_normals) {
    T acc = T(0);
    for (int fc = 0; fc < 6; ++fc) {
        int64_t f; int fam;
        face_geo(g, c, fc, f, fam);
        T n = g.fn[fam][d][f];
        if (!axis_normals && !(fc & 1)) n = -n;
        T t = fv[fc] * n;
        t = t * g.farea[fam][f];
        acc = acc + t;
    }
    return acc / g.vol[c.id];
}

template <class T>
__global__ void __launch_bounds__(256) grad_general_k(GeneralGeo<T> g, const T* __restrict__ phi,
                                                      T* __restrict__ out, int mask, int wall_axis) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    const T pc = phi[c.id];
    T fv[6];
    for (int fc = 0; fc < 6; ++fc) {
        fv[fc] = lin(pc, phi[nb_of(g, c, fc)]);
        if (wall_axis == fc_axis(fc) && is_bnd_face(g, c, fc)) fv[fc] = T(0);
    }
    T res = T(0);
    for (int d = 0; d < 3; ++d)
        if (mask & (1 << d)) res = res + gauss_general(g, c, fv, d, false);
    out[c.id] = res;
}

template <class T>
__global__ void __launch_bounds__(256) lapl_general_k(GeneralGeo<T> g, const T* __restrict__ phi,
                                                      T* __restrict__ out, int mask) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    const T pc = phi[c.id];
    T fv[6];
    for (int fc = 0; fc < 6; ++fc) {
        int64_t f; int fam;
        face_geo(g, c, fc, f, fam);
        fv[fc] = (-pc + phi[nb_of(g, c, fc)]) * g.falpha[fam][f];
    }
    T res = T(0);
    for (int d = 0; d < 3; ++d)
        if (mask & (1 << d)) res = res + gauss_general(g, c, fv, d, true);
    out[c.id] = res;
}

template <class T>
__global__ void __launch_bounds__(256) grad3_general_k(GeneralGeo<T> g, const T* __restrict__ phi,
                                                       T* __restrict__ gx, T* __restrict__ gy, T* __restrict__ gz) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    const T pc = phi[c.id];
    T fv[6];
    for (int fc = 0; fc < 6; ++fc) fv[fc] = lin(pc, phi[nb_of(g, c, fc)]);
    gx[c.id] = gauss_general(g, c, fv, 0, false);
    gy[c.id] = gauss_general(g, c, fv, 1, false);
    gz[c.id] = gauss_general(g, c, fv, 2, false);
}

template <class T>
__global__ void __launch_bounds__(256) stab_general_k(GeneralGeo<T> g, const T* __restrict__ phi,
                                                      const T* __restrict__ gx, const T* __restrict__ gy,
                                                      const T* __restrict__ gz, T* __restrict__ out, int mask) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    const T* gs[3] = {gx, gy, gz};
    const int64_t N = g.X * g.Y * g.Z;
    T res = T(0);
    for (int d = 0; d < 3; ++d) {
        if (!(mask & (1 << d))) continue;
        T fv[6];
        for (int fc = 0; fc < 6; ++fc) {
            int64_t f; int fam;
            face_geo(g, c, fc, f, fam);
            const T n = g.fn[fam][d][f];
            const int ofc = fc ^ 1;
            const int64_t nb = nb_of(g, c, fc);
            // outward normal of (cell, fc) = (fc odd ? +n : -n); the neighbour
            // sees the same face as its opposite local face.  On a boundary the
            // neighbour is the cell itself and the opposite face is its own.
            const T n_own = (fc & 1) ? n : -n;
            T n_col;
            if (nb == c.id) {
                int64_t f2; int fam2;
                face_geo(g, c, ofc, f2, fam2);
                const T n2 = g.fn[fam2][d][f2];
                n_col = (ofc & 1) ? n2 : -n2;
            } else {
                n_col = (ofc & 1) ? n : -n;
            }
            const T own = phi[c.id] + (g.len[int64_t(fc) * N + c.id] * gs[d][c.id]) * n_own;
            const T col = phi[nb] + (g.len[int64_t(ofc) * N + nb] * gs[d][nb]) * n_col;
            fv[fc] = T(0.5) * (own - col);
        }
        res = res + gauss_general(g, c, fv, d, true);
    }
    out[c.id] = res;
}

// ---------------------------------------------------------------------------
// Fused heat step: T' = dt * (alpha * lapl(T)) + T   (DtVar::extract order,
// src/fvm3d.cpp:690-692, with lapl = ((0 + L_x) + L_y) + L_z).
// ---------------------------------------------------------------------------
template <class T, bool PW>
__device__ inline void heat_uniform_cell(const UniformGeo<T>& g, const T* __restrict__ Tin, T* __restrict__ Tout,
                                         T alpha, const T* __restrict__ alpha_f, T dt) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    const T pc = Tin[c.id];
    T L = T(0);
    for (int a = 0; a < 3; ++a) {
        const T al_lo = at_lo(g, c, a) ? g.alpha_b[a] : g.alpha[a];
        const T al_hi = at_hi(g, c, a) ? g.alpha_b[a] : g.alpha[a];
        const T s_lo = (-pc + Tin[nb_lo(g, c, a)]) * al_lo;
        const T s_hi = (-pc + Tin[nb_hi(g, c, a)]) * al_hi;
        const T num = s_lo * g.area[a] + s_hi * g.area[a];
        // a power-of-two volume divides exactly, so the product is the same number
        if constexpr (PW) L = L + num * g.inv_vol;
        else L = L + num / g.vol;
    }
    const T k = alpha_f ? alpha_f[c.id] : alpha;
    Tout[c.id] = dt * (k * L) + pc;
}

template <class T>
__global__ void __launch_bounds__(256) heat_uniform_k(UniformGeo<T> g, const T* __restrict__ Tin,
                                                      T* __restrict__ Tout, T alpha, const T* __restrict__ alpha_f,
                                                      T dt) {
    if (g.vol_pow2) heat_uniform_cell<T, true>(g, Tin, Tout, alpha, alpha_f, dt);
    else heat_uniform_cell<T, false>(g, Tin, Tout, alpha, alpha_f, dt);
}

template <class T>
__global__ void __launch_bounds__(256) heat_general_k(GeneralGeo<T> g, const T* __restrict__ Tin,
                                                      T* __restrict__ Tout, T alpha, const T* __restrict__ alpha_f,
                                                      T dt) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    const T pc = Tin[c.id];
    T fv[6];
    for (int fc = 0; fc < 6; ++fc) {
        int64_t f; int fam;
        face_geo(g, c, fc, f, fam);
        fv[fc] = (-pc + Tin[nb_of(g, c, fc)]) * g.falpha[fam][f];
    }
    T L = T(0);
    for (int d = 0; d < 3; ++d) L = L + gauss_general(g, c, fv, d, true);
    const T k = alpha_f ? alpha_f[c.id] : alpha;
    Tout[c.id] = dt * (k * L) + pc;
}

// ---------------------------------------------------------------------------
// Fused acoustic step (uniform mesh), single pass per step.
//   u' = dt*(-Gx(p)) + u, v' = dt*(-Gy(p)) + v, w' = dt*(-Gz(p)) + w
//   p' = dt*(K*((Gx(u') + Gy(v')) + Gz(w'))) + p
// p' needs u' at x+-1, v' at y+-1, w' at z+-1, which each thread recomputes
// from p (distance 2) and the old velocities; the recomputation is bitwise
// identical to the neighbour's own update, so the result equals the
// sequential (Gauss-Seidel across equations) evaluation.  Walls: the normal
// velocity's face value is zero; p is zero-gradient (neighbour = self).
// ---------------------------------------------------------------------------
template <class T>
__device__ inline T vel_update(const UniformGeo<T>& g, const T* __restrict__ p, const T* __restrict__ q,
                               int64_t j, int64_t stride, bool lo_wall, bool hi_wall, int a, T dt) {
    const T pc = p[j];
    const T pl = lo_wall ? pc : p[j - stride];
    const T ph = hi_wall ? pc : p[j + stride];
    const T G = gg_pair_u(g, lin(pc, pl), lin(pc, ph), g.area[a]);
    return dt * (-G) + q[j];
}

template <class T>
__global__ void __launch_bounds__(256) wave_uniform_k(UniformGeo<T> g, const T* __restrict__ p,
                                                      const T* __restrict__ u, const T* __restrict__ v,
                                                      const T* __restrict__ w, T* __restrict__ po,
                                                      T* __restrict__ uo, T* __restrict__ vo, T* __restrict__ wo,
                                                      T K, const T* __restrict__ Kf, T dt) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    const int64_t sx = g.Y * g.Z, sy = g.Z, sz = 1;
    const T* vel[3] = {u, v, w};
    T* velo[3] = {uo, vo, wo};
    const int64_t crd[3] = {c.x, c.y, c.z};
    const int64_t ext[3] = {g.X, g.Y, g.Z};
    const int64_t str[3] = {sx, sy, sz};
    T S = T(0);
    for (int a = 0; a < 3; ++a) {
        const bool lo = crd[a] == 0, hi = crd[a] == ext[a] - 1;
        const T qc = vel_update(g, p, vel[a], c.id, str[a], lo, hi, a, dt);
        velo[a][c.id] = qc;
        T f_lo = T(0), f_hi = T(0);  // wall faces carry zero normal velocity
        if (!lo) {
            const int64_t j = c.id - str[a];
            const T ql = vel_update(g, p, vel[a], j, str[a], crd[a] - 1 == 0, false, a, dt);
            f_lo = lin(qc, ql);
        }
        if (!hi) {
            const int64_t j = c.id + str[a];
            const T qh = vel_update(g, p, vel[a], j, str[a], false, crd[a] + 1 == ext[a] - 1, a, dt);
            f_hi = lin(qc, qh);
        }
        const T G = gg_pair_u(g, f_lo, f_hi, g.area[a]);
        S = (a == 0) ? G : S + G;
    }
    const T k = Kf ? Kf[c.id] : K;
    po[c.id] = dt * (k * S) + p[c.id];
}

// General-mesh acoustic step.  Skewed face normals make every gradient
// component depend on all six faces, so the step runs as two launches: the
// three velocity equations (reading p), then the pressure equation (reading the
// new velocities) - the same sequential semantics as the uniform kernel.
template <class T>
__global__ void __launch_bounds__(256) wave_vel_general_k(GeneralGeo<T> g, const T* __restrict__ p,
                                                          const T* __restrict__ u, const T* __restrict__ v,
                                                          const T* __restrict__ w, T* __restrict__ uo,
                                                          T* __restrict__ vo, T* __restrict__ wo, T dt) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    const T pc = p[c.id];
    T fv[6];
    for (int fc = 0; fc < 6; ++fc) fv[fc] = lin(pc, p[nb_of(g, c, fc)]);
    uo[c.id] = dt * (-gauss_general(g, c, fv, 0, false)) + u[c.id];
    vo[c.id] = dt * (-gauss_general(g, c, fv, 1, false)) + v[c.id];
    wo[c.id] = dt * (-gauss_general(g, c, fv, 2, false)) + w[c.id];
}

template <class T>
__global__ void __launch_bounds__(256) wave_p_general_k(GeneralGeo<T> g, const T* __restrict__ uo,
                                                        const T* __restrict__ vo, const T* __restrict__ wo,
                                                        const T* __restrict__ p, T* __restrict__ po, T K,
                                                        const T* __restrict__ Kf, T dt) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    const T* vel[3] = {uo, vo, wo};
    T S = T(0);
    for (int a = 0; a < 3; ++a) {
        const T qc = vel[a][c.id];
        T fv[6];
        for (int fc = 0; fc < 6; ++fc) {
            fv[fc] = lin(qc, vel[a][nb_of(g, c, fc)]);
            if (fc_axis(fc) == a && is_bnd_face(g, c, fc)) fv[fc] = T(0);
        }
        const T G = gauss_general(g, c, fv, a, false);
        S = (a == 0) ? G : S + G;
    }
    const T k = Kf ? Kf[c.id] : K;
    po[c.id] = dt * (k * S) + p[c.id];
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
template <class G> inline dim3 grid_for(const G& g, dim3 blk) {
    return dim3(unsigned((g.Z + blk.x - 1) / blk.x), unsigned((g.Y + blk.y - 1) / blk.y), unsigned(g.X));
}
static const dim3 kBlock(64, 4, 1);

template <class T> const UniformGeo<T>& ugeo(const cfd_mesh_s* m);
template <> const UniformGeo<float>& ugeo<float>(const cfd_mesh_s* m) { return m->ug32; }
template <> const UniformGeo<double>& ugeo<double>(const cfd_mesh_s* m) { return m->ug64; }
template <class T> const GeneralGeo<T>& ggeo(const cfd_mesh_s* m);
template <> const GeneralGeo<float>& ggeo<float>(const cfd_mesh_s* m) { return m->gg32; }
template <> const GeneralGeo<double>& ggeo<double>(const cfd_mesh_s* m) { return m->gg64; }

inline void check_launch() { CFD_HIP(hipGetLastError()); }

template <class T> void launch_grad(const cfd_mesh_s* m, const void* phi, void* out, int mask, int wall, hipStream_t s) {
    if (m->uniform) {
        const auto& g = ugeo<T>(m);
        grad_uniform_k<T><<<grid_for(g, kBlock), kBlock, 0, s>>>(g, (const T*)phi, (T*)out, mask, wall);
    } else {
        const auto& g = ggeo<T>(m);
        grad_general_k<T><<<grid_for(g, kBlock), kBlock, 0, s>>>(g, (const T*)phi, (T*)out, mask, wall);
    }
    check_launch();
}

template <class T> void launch_lapl(const cfd_mesh_s* m, const void* phi, void* out, int mask, hipStream_t s) {
    if (m->uniform) {
        const auto& g = ugeo<T>(m);
        lapl_uniform_k<T><<<grid_for(g, kBlock), kBlock, 0, s>>>(g, (const T*)phi, (T*)out, mask);
    } else {
        const auto& g = ggeo<T>(m);
        lapl_general_k<T><<<grid_for(g, kBlock), kBlock, 0, s>>>(g, (const T*)phi, (T*)out, mask);
    }
    check_launch();
}

// scratch3 (3N entries, the first stage's gradients) is needed by the
// general form only: uniform meshes run the one-pass kernel
template <class T>
void launch_stab(const cfd_mesh_s* m, const void* phi, void* out, int mask, void* scratch3, hipStream_t s) {
    if (m->uniform) {
        const auto& g = ugeo<T>(m);
#ifndef CFD_STAB_TY
#define CFD_STAB_TY 4
#endif
#ifndef CFD_STAB_PLANES
#define CFD_STAB_PLANES 8
#endif
        // workgroup rows and x-planes per workgroup (build-time tunables)
        constexpr int TY = CFD_STAB_TY;
        const int L = CFD_STAB_PLANES;
        const int nzt = int((g.Z + 63) / 64), nyt = int((g.Y + TY - 1) / TY);
        const int64_t nwg = int64_t(nzt) * nyt * ((g.X + L - 1) / L);
        CFD_REQUIRE(nwg < (int64_t(1) << 31), "stabilisation grid too large");
        stab_march_uniform_k<T, TY><<<dim3(unsigned(nwg)), dim3(64, TY), 0, s>>>(g, (const T*)phi, (T*)out, mask,
                                                                                 L, nzt, nyt);
    } else {
        CFD_REQUIRE(scratch3 != nullptr, "stabilisation on a general mesh needs 3N scratch entries");
        T* gx = (T*)scratch3;
        T* gy = gx + m->N;
        T* gz = gy + m->N;
        const auto& g = ggeo<T>(m);
        grad3_general_k<T><<<grid_for(g, kBlock), kBlock, 0, s>>>(g, (const T*)phi, gx, gy, gz);
        check_launch();
        stab_general_k<T><<<grid_for(g, kBlock), kBlock, 0, s>>>(g, (const T*)phi, gx, gy, gz, (T*)out, mask);
    }
    check_launch();
}

template <class T>
void launch_heat(const cfd_mesh_s* m, void* Ta, void* Tb, double alpha, const void* af, double dt, int64_t nsteps,
                 hipStream_t s) {
    int64_t k0 = 0;
    if (m->uniform && !af && m->heat_variant == 2 && nsteps > 0) {
        // up to K steps per launch in LDS, reading one buffer and writing the
        // other: the number of launches has the parity of nsteps, so the
        // state ends where single steps leave it, and the steps are spread
        // evenly over the launches
        const auto& g = ugeo<T>(m);
        const int K = m->heat_L;
        int64_t launches = (nsteps + K - 1) / K;
        if ((launches - nsteps) % 2 != 0) ++launches;
        for (int64_t l = 0; l < launches; ++l) {
            const int64_t done = nsteps * l / launches, next = nsteps * (l + 1) / launches;
            const void* in = (l & 1) ? Tb : Ta;
            void* out = (l & 1) ? Ta : Tb;
            launch_heat_multi<T>(g, in, out, alpha, dt, m->heat_nw, K, int(next - done), s);
        }
        return;
    }
    if (m->uniform && !af && m->heat_variant == 1) {
        // two steps per launch, in pairs (Ta -> Tb -> Ta) so the state ends
        // where single steps leave it; the remaining steps run one per launch
        const auto& g = ugeo<T>(m);
        for (; k0 + 4 <= nsteps; k0 += 4) {
            if (!launch_heat_row2<T>(g, Ta, Tb, alpha, dt, m->heat_nw, m->heat_L, s)) break;
            launch_heat_row2<T>(g, Tb, Ta, alpha, dt, m->heat_nw, m->heat_L, s);
        }
    }
    for (int64_t k = k0; k < nsteps; ++k) {
        const T* in = (const T*)((k & 1) ? Tb : Ta);
        T* out = (T*)((k & 1) ? Ta : Tb);
        if (m->uniform) {
            const auto& g = ugeo<T>(m);
            heat_uniform_k<T><<<grid_for(g, kBlock), kBlock, 0, s>>>(g, in, out, T(alpha), (const T*)af, T(dt));
        } else {
            const auto& g = ggeo<T>(m);
            heat_general_k<T><<<grid_for(g, kBlock), kBlock, 0, s>>>(g, in, out, T(alpha), (const T*)af, T(dt));
        }
        check_launch();
    }
}

template <class T>
void launch_wave(const cfd_mesh_s* m, void* const a[4], void* const b[4], double K, const void* Kf, double dt,
                 int64_t nsteps, hipStream_t s) {
    if (!m->uniform) {
        const auto& g = ggeo<T>(m);
        for (int64_t k = 0; k < nsteps; ++k) {
            void* const* in = (k & 1) ? b : a;
            void* const* out = (k & 1) ? a : b;
            wave_vel_general_k<T><<<grid_for(g, kBlock), kBlock, 0, s>>>(
                g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3], (T*)out[1], (T*)out[2],
                (T*)out[3], T(dt));
            check_launch();
            wave_p_general_k<T><<<grid_for(g, kBlock), kBlock, 0, s>>>(
                g, (const T*)out[1], (const T*)out[2], (const T*)out[3], (const T*)in[0], (T*)out[0], T(K),
                (const T*)Kf, T(dt));
            check_launch();
        }
        return;
    }
    if (m->wave_variant >= 1) {
        launch_wave_march<T>(m, a, b, K, Kf, dt, nsteps, m->wave_variant, m->wave_ty, m->wave_L, m->wave_remap, s);
        return;
    }
    const auto& g = ugeo<T>(m);
    for (int64_t k = 0; k < nsteps; ++k) {
        void* const* in = (k & 1) ? b : a;
        void* const* out = (k & 1) ? a : b;
        wave_uniform_k<T><<<grid_for(g, kBlock), kBlock, 0, s>>>(
            g, (const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3], (T*)out[0], (T*)out[1],
            (T*)out[2], (T*)out[3], T(K), (const T*)Kf, T(dt));
        check_launch();
    }
}

// ---------------------------------------------------------------------------
// Adjoint of the heat step on uniform meshes (no reference counterpart).
// The step is H = I + dt K L, K = diag(k) with k the constant alpha or the
// alpha field, L the Laplacian of heat_uniform_cell.  On a uniform mesh L is
// symmetric (an interior face couples its two cells with the same
// alpha_d A / V, a wall face couples nothing), so H^T = I + dt L K:
//   constant alpha:  H^T = H, the forward kernels themselves;
//   alpha field:     (H^T mu)_i = dt L(k o mu)_i + mu_i   (heat_adj_field_k).
// dL/dk_i = sum over steps t of mu_t,i dt L(T_{t-1})_i (heat_grad_alpha_k),
// the forward states T_{t-1} recomputed from checkpoints.
// ---------------------------------------------------------------------------

// L(k o f) at cell c (L(f) when kf is null), heat_uniform_cell's operation
// order; on a wall the neighbour is the cell itself and its term is zero
template <class T, bool PW>
__device__ inline T lapl_scaled(const UniformGeo<T>& g, const Cell& c, const T* __restrict__ f,
                                const T* __restrict__ kf) {
    auto val = [&](int64_t j) { return kf ? kf[j] * f[j] : f[j]; };
    const T pc = val(c.id);
    T L = T(0);
    for (int a = 0; a < 3; ++a) {
        const T al_lo = at_lo(g, c, a) ? g.alpha_b[a] : g.alpha[a];
        const T al_hi = at_hi(g, c, a) ? g.alpha_b[a] : g.alpha[a];
        const T s_lo = (-pc + val(nb_lo(g, c, a))) * al_lo;
        const T s_hi = (-pc + val(nb_hi(g, c, a))) * al_hi;
        const T num = s_lo * g.area[a] + s_hi * g.area[a];
        if constexpr (PW) L = L + num * g.inv_vol;
        else L = L + num / g.vol;
    }
    return L;
}

template <class T, bool PW>
__global__ void __launch_bounds__(256) heat_adj_field_k(UniformGeo<T> g, const T* __restrict__ mu,
                                                        T* __restrict__ out, const T* __restrict__ kf, T dt) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    out[c.id] = dt * lapl_scaled<T, PW>(g, c, mu, kf) + mu[c.id];
}

template <class T, bool PW>
__global__ void __launch_bounds__(256) heat_grad_alpha_k(UniformGeo<T> g, const T* __restrict__ mu,
                                                         const T* __restrict__ Tprev, T* __restrict__ gk, T dt) {
    const Cell c = cell_of_thread(g);
    if (!c.valid) return;
    gk[c.id] = gk[c.id] + (dt * lapl_scaled<T, PW>(g, c, Tprev, nullptr)) * mu[c.id];
}

template <class T> __global__ void gather_sensors_k(const T* __restrict__ f, SensorList s, T* __restrict__ out) {
    const int j = int(threadIdx.x);
    if (j < s.n) out[j] = f[s.cell[j]];
}

// nsteps heat steps from Ta, ping-pong with Tb; after step t (1-based) T at
// the sensors goes to traces[(t-1) * n]; before steps 0, every, 2 every, ...
// the state is copied to checkpoint k.
template <class T>
void run_heat_traces(const cfd_mesh_s* m, void* Ta, void* Tb, double alpha, const void* af, double dt,
                     int64_t nsteps, const SensorList& sens, void* traces, void* ckpt, int64_t every,
                     hipStream_t s) {
    const size_t field = size_t(m->N) * sizeof(T);
    for (int64_t t = 0; t < nsteps; ++t) {
        void* in = (t & 1) ? Tb : Ta;
        void* out = (t & 1) ? Ta : Tb;
        if (ckpt && t % every == 0)
            CFD_HIP(hipMemcpyAsync((char*)ckpt + size_t(t / every) * field, in, field, hipMemcpyDeviceToDevice, s));
        launch_heat<T>(m, in, out, alpha, af, dt, 1, s);
        if (traces && sens.n) {
            gather_sensors_k<T><<<1, 64, 0, s>>>((const T*)out, sens, (T*)traces + t * sens.n);
            check_launch();
        }
    }
}

// The transposed steps in reverse order from adj_a (the adjoint of the final
// state), the trace gradients of step t added at the sensors before the
// transposed step t; the result is in adj_a (nsteps even) or adj_b (odd).
template <class T>
void run_heat_adjoint(const cfd_mesh_s* m, void* la, void* lb, double alpha, const void* af, double dt,
                      int64_t nsteps, const SensorList& sens, const void* tgrad, const void* ckpt, int64_t every,
                      void* gA, hipStream_t s) {
    const auto& g = ugeo<T>(m);
    const size_t field = size_t(m->N) * sizeof(T);
    void* ws = nullptr;  // the forward states of one checkpoint segment
    if (gA) ws = pool_alloc(m->ctx, int64_t(every) * int64_t(field), s);
    try {
        int64_t done = 0;
        auto step = [&](int64_t t, const void* Tprev) {
            void* in = (done & 1) ? lb : la;
            void* out = (done & 1) ? la : lb;
            if (tgrad) launch_scatter_add<T>(in, sens, (const T*)tgrad + (t - 1) * sens.n, s);
            if (gA) {
                if (g.vol_pow2)
                    heat_grad_alpha_k<T, true><<<grid_for(g, kBlock), kBlock, 0, s>>>(g, (const T*)in, (const T*)Tprev,
                                                                                   (T*)gA, T(dt));
                else
                    heat_grad_alpha_k<T, false><<<grid_for(g, kBlock), kBlock, 0, s>>>(g, (const T*)in,
                                                                                    (const T*)Tprev, (T*)gA, T(dt));
                check_launch();
            }
            if (af) {
                if (g.vol_pow2)
                    heat_adj_field_k<T, true><<<grid_for(g, kBlock), kBlock, 0, s>>>(g, (const T*)in, (T*)out,
                                                                                  (const T*)af, T(dt));
                else
                    heat_adj_field_k<T, false><<<grid_for(g, kBlock), kBlock, 0, s>>>(g, (const T*)in, (T*)out,
                                                                                   (const T*)af, T(dt));
                check_launch();
            } else {
                launch_heat<T>(m, in, out, alpha, nullptr, dt, 1, s);
            }
            ++done;
        };
        int64_t t = nsteps;
        while (t >= 1) {
            if (!gA) {
                step(t, nullptr);
                --t;
                continue;
            }
            // states T_{t0} .. T_{t-1} of this segment, from its checkpoint
            const int64_t seg = (t - 1) / every, t0 = seg * every, len = t - t0;
            CFD_HIP(hipMemcpyAsync(ws, (const char*)ckpt + size_t(seg) * field, field, hipMemcpyDeviceToDevice, s));
            for (int64_t k = 1; k < len; ++k)
                launch_heat<T>(m, (char*)ws + size_t(k - 1) * field, (char*)ws + size_t(k) * field, alpha, af, dt, 1,
                               s);
            for (int64_t tt = t; tt > t0; --tt) step(tt, (const char*)ws + size_t(tt - 1 - t0) * field);
            t = t0;
        }
    } catch (...) {
        if (ws) pool_free(m->ctx, ws, s);
        throw;
    }
    if (ws) pool_free(m->ctx, ws, s);
}

#define CFD_INSTANTIATE(T)                                                                                       \
    template void run_heat_traces<T>(const cfd_mesh_s*, void*, void*, double, const void*, double, int64_t,     \
                                     const SensorList&, void*, void*, int64_t, hipStream_t);                    \
    template void run_heat_adjoint<T>(const cfd_mesh_s*, void*, void*, double, const void*, double, int64_t,    \
                                      const SensorList&, const void*, const void*, int64_t, void*, hipStream_t); \
    template void launch_grad<T>(const cfd_mesh_s*, const void*, void*, int, int, hipStream_t);                \
    template void launch_lapl<T>(const cfd_mesh_s*, const void*, void*, int, hipStream_t);                     \
    template void launch_stab<T>(const cfd_mesh_s*, const void*, void*, int, void*, hipStream_t);              \
    template void launch_heat<T>(const cfd_mesh_s*, void*, void*, double, const void*, double, int64_t,         \
                                 hipStream_t);                                                                   \
    template void launch_wave<T>(const cfd_mesh_s*, void* const*, void* const*, double, const void*, double,    \
                                 int64_t, hipStream_t);
CFD_INSTANTIATE(float)
CFD_INSTANTIATE(double)

}  // namespace cfd
