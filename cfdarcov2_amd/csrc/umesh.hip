// Unstructured hex meshes (cfd_mesh_create_unstructured): the reference's
// operators on a mesh given by its connectivity (no lattice), cells reordered
// for locality, as cell-centred gathers.
//
// Device layout (cells and faces in the mesh's reordered numbering):
//   per (local face fc, cell), fc-major, int32: the neighbour across the face
//   (-1 on the boundary) and the face id; int8: the sign that turns the face
//   normal into the cell's outward normal;
//   per face: the normal (first three vertices), area and alpha_d;
//   per cell: the volume; per (fc, cell): the centre-to-face distance (the
//   stabilisation's upwind extrapolation).
// Each cell reads its own six faces and neighbours and writes only itself:
// no atomics and no scatter.  A face's value is formed by each of its two
// cells in that cell's own operation order, which is the reference's per
// (cell, fc) evaluation (include/grad_utils.hpp:23-125), so the results equal
// the reference semantics bit for bit whatever the numbering.  Reverse
// Cuthill-McKee numbering keeps the six neighbours of a cell, and so the
// gathered values, within a few cache lines of the cell's own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "umesh.hpp"

namespace cfd {

namespace {

constexpr int kUBlock = 256;

template <class T> __device__ inline T ugauss(const UnstructGeo<T>& g, const T (&fv)[6], const int32_t (&f)[6],
                                              const int8_t (&sg)[6], int d, int64_t i, bool axis) {
    // sum over fc = 0..5 of (fv * n) * area from 0, then / V; n is the axis
    // normal (+n on fc 2, 3, -n on the others) or the outward normal
    T acc = T(0);
#pragma unroll
    for (int fc = 0; fc < 6; ++fc) {
        const T n = g.fn[d][f[fc]];
        const bool neg = axis ? !(fc == 2 || fc == 3) : sg[fc] < 0;
        T t = fv[fc] * (neg ? -n : n);
        t = t * g.area[f[fc]];
        acc = acc + t;
    }
    return acc / g.vol[i];
}

template <class T> __device__ inline void uload(const UnstructGeo<T>& g, int64_t i, int32_t (&nb)[6], int32_t (&f)[6],
                                                int8_t (&sg)[6]) {
#pragma unroll
    for (int fc = 0; fc < 6; ++fc) {
        const int64_t k = int64_t(fc) * g.N + i;
        nb[fc] = g.nbr[k];
        f[fc] = g.face[k];
        sg[fc] = g.osign[k];
    }
}

}  // namespace

// GradVar: face value (phi_c + phi_nb) / 2 (neighbour = self on the boundary,
// or 0 on boundary faces normal to wall_axis), outward Gauss sum per selected
// axis, res = 0 + g_x (+ g_y) (+ g_z).
template <class T>
__global__ void __launch_bounds__(kUBlock) ugrad_k(UnstructGeo<T> g, const T* __restrict__ phi, T* __restrict__ out,
                                                   int mask, int wall) {
    const int64_t i = int64_t(blockIdx.x) * kUBlock + threadIdx.x;
    if (i >= g.N) return;
    int32_t nb[6], f[6];
    int8_t sg[6];
    uload(g, i, nb, f, sg);
    const T pc = phi[i];
    T fv[6];
#pragma unroll
    for (int fc = 0; fc < 6; ++fc) {
        const bool bnd = nb[fc] < 0;
        fv[fc] = (wall >= 0 && bnd && (fc >> 1) == 2 - wall) ? T(0) : (pc + (bnd ? pc : phi[nb[fc]])) * T(0.5);
    }
    T res = T(0);
    for (int d = 0; d < 3; ++d)
        if (mask & (1 << d)) res = res + ugauss(g, fv, f, sg, d, i, false);
    out[i] = res;
}

// Grad2Var: (-phi_c + phi_nb) * alpha_d, axis-normal Gauss sum per selected axis.
template <class T> __device__ inline T ulapl_cell(const UnstructGeo<T>& g, const T* __restrict__ phi, int64_t i,
                                                  int mask) {
    int32_t nb[6], f[6];
    int8_t sg[6];
    uload(g, i, nb, f, sg);
    const T pc = phi[i];
    T fv[6];
#pragma unroll
    for (int fc = 0; fc < 6; ++fc) fv[fc] = (-pc + (nb[fc] < 0 ? pc : phi[nb[fc]])) * g.alpha[f[fc]];
    T res = T(0);
    for (int d = 0; d < 3; ++d)
        if (mask & (1 << d)) res = res + ugauss(g, fv, f, sg, d, i, true);
    return res;
}

template <class T>
__global__ void __launch_bounds__(kUBlock) ulapl_k(UnstructGeo<T> g, const T* __restrict__ phi, T* __restrict__ out,
                                                   int mask) {
    const int64_t i = int64_t(blockIdx.x) * kUBlock + threadIdx.x;
    if (i >= g.N) return;
    out[i] = ulapl_cell(g, phi, i, mask);
}

// d1t(T) = alpha * lapl(T): T <- dt * (alpha * L) + T
template <class T>
__global__ void __launch_bounds__(kUBlock) uheat_k(UnstructGeo<T> g, const T* __restrict__ Tin, T* __restrict__ Tout,
                                                   T alpha, const T* __restrict__ alpha_f, T dt) {
    const int64_t i = int64_t(blockIdx.x) * kUBlock + threadIdx.x;
    if (i >= g.N) return;
    const T L = ulapl_cell(g, Tin, i, 7);
    const T k = alpha_f ? alpha_f[i] : alpha;
    Tout[i] = dt * (k * L) + Tin[i];
}

namespace {

template <class T> const UnstructGeo<T>& ugeo_u(const cfd_mesh_s* m) {
    if constexpr (sizeof(T) == 4) return m->uu32; else return m->uu64;
}
template <class T> UnstructGeo<T>& ugeo_u(cfd_mesh_s* m) {
    if constexpr (sizeof(T) == 4) return m->uu32; else return m->uu64;
}
dim3 ugrid(int64_t N) { return dim3(unsigned((N + kUBlock - 1) / kUBlock)); }

}  // namespace

template <class T> void launch_ugrad(const cfd_mesh_s* m, const void* phi, void* out, int mask, int wall, hipStream_t s) {
    ugrad_k<T><<<ugrid(m->N), kUBlock, 0, s>>>(ugeo_u<T>(m), (const T*)phi, (T*)out, mask, wall);
    CFD_HIP(hipGetLastError());
}
template <class T> void launch_ulapl(const cfd_mesh_s* m, const void* phi, void* out, int mask, hipStream_t s) {
    ulapl_k<T><<<ugrid(m->N), kUBlock, 0, s>>>(ugeo_u<T>(m), (const T*)phi, (T*)out, mask);
    CFD_HIP(hipGetLastError());
}
// StabVar, first stage: the three outward-Gauss gradient components of phi
// with the reference face rule (neighbour = self on the boundary), each
// written as computed (no 0 + g), as the reference keeps them per axis.
template <class T>
__global__ void __launch_bounds__(kUBlock) ustab_grad_k(UnstructGeo<T> g, const T* __restrict__ phi,
                                                        T* __restrict__ gx, T* __restrict__ gy, T* __restrict__ gz) {
    const int64_t i = int64_t(blockIdx.x) * kUBlock + threadIdx.x;
    if (i >= g.N) return;
    int32_t nb[6], f[6];
    int8_t sg[6];
    uload(g, i, nb, f, sg);
    const T pc = phi[i];
    T fv[6];
#pragma unroll
    for (int fc = 0; fc < 6; ++fc) fv[fc] = (pc + (nb[fc] < 0 ? pc : phi[nb[fc]])) * T(0.5);
    gx[i] = ugauss(g, fv, f, sg, 0, i, false);
    gy[i] = ugauss(g, fv, f, sg, 1, i, false);
    gz[i] = ugauss(g, fv, f, sg, 2, i, false);
}

// StabVar, second stage (include/grad_utils.hpp:38-80, src/fvm3d.cpp:915-936):
// per selected axis d, the upwind face value phi + len * g_d * n_out,d of the
// cell on face fc against the neighbour's on its opposite local face fc ^ 1
// (self on the boundary), 0.5 * (own - neighbour's), axis-normal Gauss sum,
// res = 0 + S_x (+ S_y) (+ S_z).  The neighbour's value is recomputed from its
// phi, gradient and face geometry, as its own cell forms it.
template <class T>
__global__ void __launch_bounds__(kUBlock) ustab_k(UnstructGeo<T> g, const T* __restrict__ phi,
                                                   const T* __restrict__ gx, const T* __restrict__ gy,
                                                   const T* __restrict__ gz, T* __restrict__ out, int mask) {
    const int64_t i = int64_t(blockIdx.x) * kUBlock + threadIdx.x;
    if (i >= g.N) return;
    int32_t nb[6], f[6];
    int8_t sg[6];
    uload(g, i, nb, f, sg);
    const T* const gd[3] = {gx, gy, gz};
    auto upwind = [&](int d, int64_t c, int fc) {
        const int64_t k = int64_t(fc) * g.N + c;
        const T n = g.fn[d][g.face[k]];
        const T na = g.osign[k] < 0 ? -n : n;
        return phi[c] + g.len[k] * gd[d][c] * na;
    };
    T res = T(0);
    for (int d = 0; d < 3; ++d) {
        if (!(mask & (1 << d))) continue;
        T fv[6];
#pragma unroll
        for (int fc = 0; fc < 6; ++fc) {
            const int64_t c = nb[fc] < 0 ? i : int64_t(nb[fc]);
            const T own = upwind(d, i, fc);
            const T col = upwind(d, c, fc ^ 1);
            fv[fc] = T(0.5) * (own - col);
        }
        res = res + ugauss(g, fv, f, sg, d, i, true);
    }
    out[i] = res;
}

// The fused acoustic step (the reference operators' composition, as the
// structured kernels run it): first every cell's velocities
//   u_d' = dt * (-(0 + G_d(p))) + u_d   (face values (p_c + p_nb) / 2, neighbour = self)
// then the pressure from the new velocities
//   p' = dt * (K * ((D_x(u') + D_y(v')) + D_z(w'))) + p
// with D_a = 0 + the outward Gauss sum of velocity a under the reflective-wall
// face rule (zero on boundary faces of the axis-a pair of local faces).
template <class T>
__global__ void __launch_bounds__(kUBlock) uwave_vel_k(UnstructGeo<T> g, const T* __restrict__ p,
                                                       const T* __restrict__ u, const T* __restrict__ v,
                                                       const T* __restrict__ w, T* __restrict__ uo,
                                                       T* __restrict__ vo, T* __restrict__ wo, T dt) {
    const int64_t i = int64_t(blockIdx.x) * kUBlock + threadIdx.x;
    if (i >= g.N) return;
    int32_t nb[6], f[6];
    int8_t sg[6];
    uload(g, i, nb, f, sg);
    const T pc = p[i];
    T fv[6];
#pragma unroll
    for (int fc = 0; fc < 6; ++fc) fv[fc] = (pc + (nb[fc] < 0 ? pc : p[nb[fc]])) * T(0.5);
    const T* const vin[3] = {u, v, w};
    T* const vout[3] = {uo, vo, wo};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T gd = T(0) + ugauss(g, fv, f, sg, d, i, false);
        vout[d][i] = dt * (-gd) + vin[d][i];
    }
}

template <class T>
__global__ void __launch_bounds__(kUBlock) uwave_pres_k(UnstructGeo<T> g, const T* __restrict__ p,
                                                        const T* __restrict__ u, const T* __restrict__ v,
                                                        const T* __restrict__ w, T* __restrict__ po, T K,
                                                        const T* __restrict__ Kf, T dt) {
    const int64_t i = int64_t(blockIdx.x) * kUBlock + threadIdx.x;
    if (i >= g.N) return;
    int32_t nb[6], f[6];
    int8_t sg[6];
    uload(g, i, nb, f, sg);
    const T* const vel[3] = {u, v, w};
    T div[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T* q = vel[d];
        const T qc = q[i];
        T fv[6];
#pragma unroll
        for (int fc = 0; fc < 6; ++fc) {
            const bool bnd = nb[fc] < 0;
            fv[fc] = (bnd && (fc >> 1) == 2 - d) ? T(0) : (qc + (bnd ? qc : q[nb[fc]])) * T(0.5);
        }
        div[d] = T(0) + ugauss(g, fv, f, sg, d, i, false);
    }
    const T S = (div[0] + div[1]) + div[2];
    const T k = Kf ? Kf[i] : K;
    po[i] = dt * (k * S) + p[i];
}

template <class T>
void launch_uwave(const cfd_mesh_s* m, void* const a[4], void* const b[4], double K, const void* Kf, double dt,
                  int64_t nsteps, hipStream_t s) {
    for (int64_t t = 0; t < nsteps; ++t) {
        void* const* in = (t & 1) ? b : a;
        void* const* out = (t & 1) ? a : b;
        uwave_vel_k<T><<<ugrid(m->N), kUBlock, 0, s>>>(ugeo_u<T>(m), (const T*)in[0], (const T*)in[1],
                                                       (const T*)in[2], (const T*)in[3], (T*)out[1], (T*)out[2],
                                                       (T*)out[3], T(dt));
        CFD_HIP(hipGetLastError());
        uwave_pres_k<T><<<ugrid(m->N), kUBlock, 0, s>>>(ugeo_u<T>(m), (const T*)in[0], (const T*)out[1],
                                                        (const T*)out[2], (const T*)out[3], (T*)out[0], T(K),
                                                        (const T*)Kf, T(dt));
        CFD_HIP(hipGetLastError());
    }
}

template <class T>
void launch_ustab(const cfd_mesh_s* m, const void* phi, void* out, int mask, void* work, hipStream_t s) {
    T* w = static_cast<T*>(work);
    const int64_t N = m->N;
    ustab_grad_k<T><<<ugrid(N), kUBlock, 0, s>>>(ugeo_u<T>(m), (const T*)phi, w, w + N, w + 2 * N);
    CFD_HIP(hipGetLastError());
    ustab_k<T><<<ugrid(N), kUBlock, 0, s>>>(ugeo_u<T>(m), (const T*)phi, w, w + N, w + 2 * N, (T*)out, mask);
    CFD_HIP(hipGetLastError());
}

template <class T>
void launch_uheat(const cfd_mesh_s* m, void* Ta, void* Tb, double alpha, const void* af, double dt, int64_t nsteps,
                  hipStream_t s) {
    for (int64_t k = 0; k < nsteps; ++k) {
        const T* in = (const T*)((k & 1) ? Tb : Ta);
        T* out = (T*)((k & 1) ? Ta : Tb);
        uheat_k<T><<<ugrid(m->N), kUBlock, 0, s>>>(ugeo_u<T>(m), in, out, T(alpha), (const T*)af, T(dt));
        CFD_HIP(hipGetLastError());
    }
}

// Geometry, numbering and upload of an unstructured mesh.  Device block:
// nbr (6N int32), face (6N int32), then per face fn (3F), area (F), alpha (F),
// then vol (N), len (6N), then osign (6N int8).
template <class T> void create_unstructured(cfd_mesh_s* m, int reorder) {
    const UConn& c = m->uconn;
    check_uconn(c);
    UGeometry<T> g;
    build_ugeometry<T>(g, c);
    const int64_t N = c.N, F = c.F;
    if (reorder) {
        m->cell_perm = rcm_order(c);
    } else {
        m->cell_perm.resize(N);
        for (int64_t i = 0; i < N; ++i) m->cell_perm[i] = i;
    }
    m->cell_iperm.assign(N, 0);
    for (int64_t n = 0; n < N; ++n) m->cell_iperm[m->cell_perm[n]] = n;
    m->face_perm = face_order(c, m->cell_iperm);
    std::vector<int64_t> face_iperm(F);
    for (int64_t n = 0; n < F; ++n) face_iperm[m->face_perm[n]] = n;

    const size_t ib = 6 * size_t(N) * sizeof(int32_t);
    const size_t tb = (5 * size_t(F) + 7 * size_t(N)) * sizeof(T);
    const size_t sb = 6 * size_t(N);
    const int64_t bytes = int64_t(2 * ib + tb + sb);
    std::vector<char> host(static_cast<size_t>(bytes));
    int32_t* nbr = reinterpret_cast<int32_t*>(host.data());
    int32_t* fid = nbr + 6 * N;
    T* tv = reinterpret_cast<T*>(host.data() + 2 * ib);
    int8_t* sgn = reinterpret_cast<int8_t*>(host.data() + 2 * ib + tb);
    for (int64_t n = 0; n < N; ++n) {
        const int64_t i = m->cell_perm[n];
        for (int fc = 0; fc < 6; ++fc) {
            const int64_t f = c.cell_faces[6 * i + fc];
            const int64_t a = c.face_cells[2 * f], b = c.face_cells[2 * f + 1];
            const int64_t other = b < 0 ? -1 : (a == i ? b : a);
            nbr[fc * N + n] = other < 0 ? -1 : int32_t(m->cell_iperm[other]);
            fid[fc * N + n] = int32_t(face_iperm[f]);
            sgn[fc * N + n] = g.out_sign[6 * i + fc];
        }
    }
    for (int64_t n = 0; n < F; ++n) {
        const int64_t f = m->face_perm[n];
        tv[n] = g.fnx[f];
        tv[F + n] = g.fny[f];
        tv[2 * F + n] = g.fnz[f];
        tv[3 * F + n] = g.farea[f];
        tv[4 * F + n] = g.falpha[f];
    }
    for (int64_t n = 0; n < N; ++n) tv[5 * F + n] = g.vol[m->cell_perm[n]];
    T* lv = tv + 5 * F + N;
    for (int64_t n = 0; n < N; ++n)
        for (int fc = 0; fc < 6; ++fc) lv[fc * N + n] = g.len[6 * m->cell_perm[n] + fc];

    hipStream_t s = nullptr;
    m->dev_block = pool_alloc(m->ctx, bytes, s);
    m->device_bytes = bytes;
    CFD_HIP(hipMemcpyAsync(m->dev_block, host.data(), size_t(bytes), hipMemcpyHostToDevice, s));
    CFD_HIP(hipStreamSynchronize(s));
    char* d = static_cast<char*>(m->dev_block);
    UnstructGeo<T>& u = ugeo_u<T>(m);
    u.N = N;
    u.F = F;
    u.nbr = reinterpret_cast<const int32_t*>(d);
    u.face = reinterpret_cast<const int32_t*>(d + ib);
    const T* dt = reinterpret_cast<const T*>(d + 2 * ib);
    for (int k = 0; k < 3; ++k) u.fn[k] = dt + k * F;
    u.area = dt + 3 * F;
    u.alpha = dt + 4 * F;
    u.vol = dt + 5 * F;
    u.len = dt + 5 * F + N;
    u.osign = reinterpret_cast<const int8_t*>(d + 2 * ib + tb);
}

// Reference-layout tables of an unstructured mesh (N x 6 column-major), cells
// in the mesh's numbering when `perm` is given (perm[new] = old), else in the
// caller's.
template <class T>
void unstructured_table(const UConn& c, const int64_t* perm, const int64_t* iperm, int which, void* out) {
    check_uconn(c);
    UGeometry<T> g;
    build_ugeometry<T>(g, c);
    const int64_t N = c.N;
    T* ot = static_cast<T*>(out);
    int64_t* oi = static_cast<int64_t*>(out);
    for (int64_t n = 0; n < N; ++n) {
        const int64_t i = perm ? perm[n] : n;
        switch (which) {
            case CFD_T_VOL: ot[n] = g.vol[i]; continue;
            case CFD_T_CCX: ot[n] = g.ccx[i]; continue;
            case CFD_T_CCY: ot[n] = g.ccy[i]; continue;
            case CFD_T_CCZ: ot[n] = g.ccz[i]; continue;
            default: break;
        }
        bool bnd = false;
        for (int fc = 0; fc < 6; ++fc) {
            const int64_t f = c.cell_faces[6 * i + fc];
            const int64_t a = c.face_cells[2 * f], b = c.face_cells[2 * f + 1];
            const bool fb = b < 0;
            bnd = bnd || fb;
            const int64_t other = fb ? i : (a == i ? b : a);
            const int64_t k = int64_t(fc) * N + n;
            const P3<T> nf{g.fnx[f], g.fny[f], g.fnz[f]};
            const P3<T> ax = HexGeometry<T>::axis_normal(nf, fc);
            const T so = T(g.out_sign[6 * i + fc]);
            switch (which) {
                case CFD_T_IDS: oi[k] = iperm ? iperm[other] : other; break;
                case CFD_T_IDS_BF: oi[k] = fb ? -1 : (iperm ? iperm[other] : other); break;
                case CFD_T_NX: ot[k] = ax.x; break;
                case CFD_T_NY: ot[k] = ax.y; break;
                case CFD_T_NZ: ot[k] = ax.z; break;
                case CFD_T_NAX: ot[k] = so > 0 ? nf.x : -nf.x; break;
                case CFD_T_NAY: ot[k] = so > 0 ? nf.y : -nf.y; break;
                case CFD_T_NAZ: ot[k] = so > 0 ? nf.z : -nf.z; break;
                case CFD_T_AREA: ot[k] = g.farea[f]; break;
                case CFD_T_ALPHA: ot[k] = g.falpha[f]; break;
                case CFD_T_LEN: ot[k] = g.len[6 * i + fc]; break;
                default: break;
            }
        }
        if (which == CFD_T_IS_BND) ot[n] = bnd ? T(1) : T(0);
    }
    CFD_REQUIRE(which >= CFD_T_IDS && which <= CFD_T_CCZ, "unstructured meshes have the reference tables 0 to 15");
}

#define CFD_U_INSTANTIATE(T)                                                                                   \
    template void launch_ugrad<T>(const cfd_mesh_s*, const void*, void*, int, int, hipStream_t);              \
    template void launch_ulapl<T>(const cfd_mesh_s*, const void*, void*, int, hipStream_t);                   \
    template void launch_ustab<T>(const cfd_mesh_s*, const void*, void*, int, void*, hipStream_t);            \
    template void launch_uwave<T>(const cfd_mesh_s*, void* const[4], void* const[4], double, const void*,      \
                                  double, int64_t, hipStream_t);                                               \
    template void launch_uheat<T>(const cfd_mesh_s*, void*, void*, double, const void*, double, int64_t,       \
                                  hipStream_t);                                                                \
    template void create_unstructured<T>(cfd_mesh_s*, int);                                                    \
    template void unstructured_table<T>(const UConn&, const int64_t*, const int64_t*, int, void*);
CFD_U_INSTANTIATE(float)
CFD_U_INSTANTIATE(double)

}  // namespace cfd
