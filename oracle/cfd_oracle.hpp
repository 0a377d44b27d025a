// ORACLE — test infrastructure only.
//
// CPU restatement of cfdARCO's explicit finite-volume path (mesh topology and
// geometry, face interpolation, Green-Gauss gradient, Laplacian, stabilisation
// and the explicit time integrators), templated on the scalar type so the same
// code is the FP32 reference-semantics checker and the FP64 oracle.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
// this code, and only as the checker or as the timed CPU baseline.  The product
// library (cfdarcov2_amd/) never includes, links or calls it.
//
// Every function cites the reference file:line it restates.  Arithmetic is
// written in the reference's operation order; compile with -ffp-contract=off so
// no multiply-add is fused behind our back.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

namespace cfdo {

template <class T> struct V3 { T x, y, z; };

template <class T> inline V3<T> vsub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

// Eigen's 3-vector cross product component formulas.
template <class T> inline V3<T> vcross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Fixed-size (3) reductions in Eigen unroll as a0 + (a1 + a2); dynamic-size
// vectors reduce left to right, (a0 + a1) + a2.  The reference uses fixed-size
// vectors in Face3D/Quadrangle3D::compute and dynamic ones in Mesh3D::compute.
template <class T> inline T sum3_fixed(T a, T b, T c) { return a + (b + c); }
template <class T> inline T sum3_dyn(T a, T b, T c) { return (a + b) + c; }
template <class T> inline T dot_fixed(V3<T> a, V3<T> b) { return sum3_fixed(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> inline T dot_dyn(V3<T> a, V3<T> b) { return sum3_dyn(a.x * b.x, a.y * b.y, a.z * b.z); }

// opposite local face: src/mesh3d.cpp:140-148
inline int opposite_fc(int fc) { return fc ^ 1; }

// Hex mesh with the reference's numbering and N x 6 column-major tables
// (element (i, fc) at fc * N + i, include/decls.hpp:25-28).
template <class T> struct HexMesh {
    int64_t X = 0, Y = 0, Z = 0;
    T dx = 0, dy = 0, dz = 0;
    int64_t N = 0, F = 0, V = 0;

    std::vector<T> vx, vy, vz;              // vertex coordinates
    std::vector<int64_t> face_verts;        // F x 4
    std::vector<int64_t> face_cells;        // F x 2 (-1 when absent)
    std::vector<int32_t> face_ncells;       // 1 on the boundary, 2 inside
    std::vector<T> fcx, fcy, fcz;           // face centre
    std::vector<T> fnx, fny, fnz;           // face normal (from the first three vertices)
    std::vector<T> farea;                   // face area (sum-of-components form)
    std::vector<int64_t> cell_faces;        // N x 6 (row-major per cell)
    std::vector<int64_t> cell_verts;        // N x 8
    std::vector<T> ccx, ccy, ccz;           // cell centre

    // N x 6 column-major tables
    std::vector<int64_t> ids, ids_bf;
    std::vector<T> nx, ny, nz;              // _normals (axis direction)
    std::vector<T> nax, nay, naz;           // _normals_alt (outward)
    std::vector<T> area6, alpha_d, len6;
    std::vector<T> vol, is_bnd, is_bnd_rev;

    // finite-volume geometry of the Euler stepper (per global face / per cell)
    std::vector<T> svx, svy, svz, sva;      // unit +axis normal and vector-area magnitude
    std::vector<T> svol;                    // cell volume from the vector areas

    // Mesh3D::square_node_coord_to_idx, src/mesh3d.cpp:387-389
    int64_t cell_idx(int64_t x, int64_t y, int64_t z) const { return Z * Y * x + Z * y + z; }

    // Mesh3D::square_node_coord_to_face_idx, src/mesh3d.cpp:391-403
    void face_idx(int64_t x, int64_t y, int64_t z, int64_t out[6]) const {
        const int64_t nzf = X * Y * (Z + 1);
        const int64_t nyf = X * (Y + 1) * Z;
        out[0] = (Z + 1) * Y * x + (Z + 1) * y + z;
        out[1] = (Z + 1) * Y * x + (Z + 1) * y + (z + 1);
        out[2] = nzf + (X * (Y + 1) * z + (Y + 1) * x + y);
        out[3] = nzf + (X * (Y + 1) * z + (Y + 1) * x + (y + 1));
        out[4] = nzf + nyf + ((X + 1) * Z * y + (X + 1) * z + x);
        out[5] = nzf + nyf + ((X + 1) * Z * y + (X + 1) * z + (x + 1));
    }

    // Mesh3D::square_node_coord_to_vertex_idx, src/mesh3d.cpp:405-416
    void vertex_idx(int64_t x, int64_t y, int64_t z, int64_t out[8]) const {
        const int64_t sx = (Z + 1) * (Y + 1), sy = (Z + 1);
        int k = 0;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c) out[k++] = (x + a) * sx + (y + b) * sy + (z + c);
    }
};

// Mesh3D::init_basic_internals, src/mesh3d.cpp:323-385.  Vertex coordinates are
// x_ * dx in T (:328).  `perturb` (nullable, V x 3, in T) is added to the
// coordinates before any geometry is computed: the seeded non-uniform-mesh hook.
template <class T>
void build_topology(HexMesh<T>& m, int64_t X, int64_t Y, int64_t Z, T dx, T dy, T dz, const T* perturb) {
    m.X = X; m.Y = Y; m.Z = Z; m.dx = dx; m.dy = dy; m.dz = dz;
    m.N = X * Y * Z;
    m.V = (X + 1) * (Y + 1) * (Z + 1);
    m.F = X * Y * (Z + 1) + Z * Y * (X + 1) + X * Z * (Y + 1);
    m.vx.resize(m.V); m.vy.resize(m.V); m.vz.resize(m.V);
    int64_t vid = 0;
    for (int64_t x = 0; x < X + 1; ++x)
        for (int64_t y = 0; y < Y + 1; ++y)
            for (int64_t z = 0; z < Z + 1; ++z) {
                m.vx[vid] = T(x) * dx; m.vy[vid] = T(y) * dy; m.vz[vid] = T(z) * dz;
                if (perturb) {
                    m.vx[vid] = m.vx[vid] + perturb[3 * vid + 0];
                    m.vy[vid] = m.vy[vid] + perturb[3 * vid + 1];
                    m.vz[vid] = m.vz[vid] + perturb[3 * vid + 2];
                }
                ++vid;
            }
    m.face_verts.assign(4 * m.F, 0);
    m.face_cells.assign(2 * m.F, -1);
    m.face_ncells.assign(m.F, 0);
    m.cell_faces.resize(6 * m.N);
    m.cell_verts.resize(8 * m.N);
    // per-cell local face -> vertex quads, src/mesh3d.cpp:361-374
    static const int quad[6][4] = {{0, 2, 6, 4}, {1, 3, 7, 5}, {0, 1, 5, 4},
                                   {2, 3, 7, 6}, {1, 3, 2, 0}, {5, 7, 6, 4}};
    for (int64_t x = 0; x < X; ++x)
        for (int64_t y = 0; y < Y; ++y)
            for (int64_t z = 0; z < Z; ++z) {
                const int64_t c = m.cell_idx(x, y, z);
                int64_t fid[6], vids[8];
                m.face_idx(x, y, z, fid);
                m.vertex_idx(x, y, z, vids);
                for (int j = 0; j < 6; ++j) m.cell_faces[6 * c + j] = fid[j];
                for (int j = 0; j < 8; ++j) m.cell_verts[8 * c + j] = vids[j];
                for (int j = 0; j < 6; ++j)
                    for (int q = 0; q < 4; ++q) m.face_verts[4 * fid[j] + q] = vids[quad[j][q]];
                for (int j = 0; j < 6; ++j) {
                    const int64_t f = fid[j];
                    m.face_cells[2 * f + m.face_ncells[f]] = c;
                    m.face_ncells[f] += 1;
                }
            }
}

// A hex mesh from its connectivity (the mesh.json layout written by
// store_mesh, include/io_operators_tmpl.hpp:161-212, and read back by the -m
// readers, :215-247 and legacy/io_operators.cpp:170-205): the same objects
// init_basic_internals builds, taken as given.  face_cells holds the face's
// cells in the order they list it, -1 when absent.
template <class T>
void build_from_connectivity(HexMesh<T>& m, int64_t N, int64_t F, int64_t V, const double* verts,
                             const int64_t* face_verts, const int64_t* face_cells, const int64_t* cell_faces,
                             const int64_t* cell_verts) {
    m.X = m.Y = m.Z = 0;
    m.N = N; m.F = F; m.V = V;
    m.vx.resize(V); m.vy.resize(V); m.vz.resize(V);
    for (int64_t v = 0; v < V; ++v) {
        m.vx[v] = T(verts[3 * v]); m.vy[v] = T(verts[3 * v + 1]); m.vz[v] = T(verts[3 * v + 2]);
    }
    m.face_verts.assign(face_verts, face_verts + 4 * F);
    m.face_cells.assign(face_cells, face_cells + 2 * F);
    m.face_ncells.assign(F, 0);
    for (int64_t f = 0; f < F; ++f) m.face_ncells[f] = face_cells[2 * f + 1] >= 0 ? 2 : 1;
    m.cell_faces.assign(cell_faces, cell_faces + 6 * N);
    m.cell_verts.assign(cell_verts, cell_verts + 8 * N);
}

// Face3D::compute (src/mesh3d.cpp:43-62), Quadrangle3D::compute (:69-105) and the
// table assembly of Mesh3D::compute (:151-321).
template <class T> void compute_geometry(HexMesh<T>& m) {
    const int64_t N = m.N, F = m.F;
    m.fcx.resize(F); m.fcy.resize(F); m.fcz.resize(F);
    m.fnx.resize(F); m.fny.resize(F); m.fnz.resize(F); m.farea.resize(F);
    auto vert = [&](int64_t v) { return V3<T>{m.vx[v], m.vy[v], m.vz[v]}; };
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < F; ++f) {
        V3<T> p[4];
        V3<T> c{0, 0, 0};
        for (int q = 0; q < 4; ++q) {
            p[q] = vert(m.face_verts[4 * f + q]);
            c = {c.x + p[q].x, c.y + p[q].y, c.z + p[q].z};
        }
        const V3<T> AB = vsub(p[1], p[0]), BC = vsub(p[2], p[1]), CD = vsub(p[3], p[2]);
        const V3<T> AD = vsub(p[3], p[0]), CB = vsub(p[1], p[2]);
        V3<T> n = vcross(AB, BC);
        const T sq = dot_fixed(n, n);  // normalized(): divide by sqrt(squaredNorm) when > 0
        if (sq > T(0)) {
            const T s = std::sqrt(sq);
            n = {n.x / s, n.y / s, n.z / s};
        }
        m.fnx[f] = n.x; m.fny[f] = n.y; m.fnz[f] = n.z;
        const T four = T(4);
        m.fcx[f] = c.x / four; m.fcy[f] = c.y / four; m.fcz[f] = c.z / four;
        const V3<T> c1 = vcross(AB, AD), c2 = vcross(CB, CD);
        const T s1 = std::abs(sum3_fixed(c1.x, c1.y, c1.z));
        const T s2 = std::abs(sum3_fixed(c2.x, c2.y, c2.z));
        m.farea[f] = T(0.5 * (s1 + s2));  // |sum of components|, src/mesh3d.cpp:61
    }

    m.ccx.resize(N); m.ccy.resize(N); m.ccz.resize(N); m.vol.resize(N);
    const size_t n6 = size_t(6) * size_t(N);
    m.nx.resize(n6); m.ny.resize(n6); m.nz.resize(n6);
    m.nax.resize(n6); m.nay.resize(n6); m.naz.resize(n6);
    m.ids.resize(n6); m.ids_bf.resize(n6); m.area6.resize(n6); m.alpha_d.resize(n6); m.len6.resize(n6);
    m.is_bnd.resize(N); m.is_bnd_rev.resize(N);

#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        V3<T> c{0, 0, 0};
        for (int q = 0; q < 8; ++q) {
            const V3<T> p = vert(m.cell_verts[8 * i + q]);
            c = {c.x + p.x, c.y + p.y, c.z + p.z};
        }
        const T eight = T(8);
        c = {c.x / eight, c.y / eight, c.z / eight};
        m.ccx[i] = c.x; m.ccy[i] = c.y; m.ccz[i] = c.z;
        T volume = 0;
        for (int j = 0; j < 6; ++j) {
            const int64_t f = m.cell_faces[6 * i + j];
            const V3<T> n{m.fnx[f], m.fny[f], m.fnz[f]};
            const V3<T> fc{m.fcx[f], m.fcy[f], m.fcz[f]};
            const T a = m.farea[f];
            V3<T> good = n;
            if (!(dot_fixed(n, vsub(fc, c)) >= T(0))) good = {-n.x, -n.y, -n.z};
            const size_t k = size_t(j) * size_t(N) + size_t(i);
            // axis-direction normals, src/mesh3d.cpp:92-96
            if (j == 4 || j == 5 || j == 0 || j == 1) {
                m.nx[k] = -n.x; m.ny[k] = -n.y; m.nz[k] = -n.z;
            } else {
                m.nx[k] = n.x; m.ny[k] = n.y; m.nz[k] = n.z;
            }
            m.nax[k] = good.x; m.nay[k] = good.y; m.naz[k] = good.z;
            volume += dot_fixed(fc, good) * a;  // :101
        }
        m.vol[i] = volume / T(3);
    }

#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        bool bnd = false;
        for (int j = 0; j < 6; ++j) bnd = bnd || (m.face_ncells[m.cell_faces[6 * i + j]] == 1);
        m.is_bnd[i] = bnd ? T(1) : T(0);
        m.is_bnd_rev[i] = bnd ? T(0) : T(1);
        const V3<T> c1{m.ccx[i], m.ccy[i], m.ccz[i]};
        for (int j = 0; j < 6; ++j) {
            const int64_t f = m.cell_faces[6 * i + j];
            const size_t k = size_t(j) * size_t(N) + size_t(i);
            m.area6[k] = m.farea[f];
            const V3<T> fn{m.nx[k], m.ny[k], m.nz[k]};
            const V3<T> fcen{m.fcx[f], m.fcy[f], m.fcz[f]};
            const V3<T> to_face = vsub(fcen, c1);
            if (m.face_ncells[f] > 1) {  // :240-265
                const int64_t a0 = m.face_cells[2 * f + 0], a1 = m.face_cells[2 * f + 1];
                const int64_t nb = (a0 == i) ? a1 : a0;
                m.ids[k] = nb;
                m.ids_bf[k] = nb;
                const V3<T> d = vsub(V3<T>{m.ccx[nb], m.ccy[nb], m.ccz[nb]}, c1);
                const T dn = std::sqrt(dot_dyn(d, d));
                const T cos_phi = dot_dyn(fn, d) / (std::sqrt(dot_dyn(fn, fn)) * dn);
                const T alpha = T(1) / cos_phi;
                m.alpha_d[k] = std::abs(alpha / dn);
                m.len6[k] = std::abs(std::sqrt(dot_dyn(to_face, to_face)));
            } else {  // boundary: neighbour = self, bound-free id = -1 (:266-312)
                m.ids[k] = i;
                m.ids_bf[k] = -1;
                const V3<T> d = to_face;
                const T dn = std::sqrt(dot_dyn(d, d));
                const T cos_phi = dot_dyn(fn, d) / (std::sqrt(dot_dyn(fn, fn)) * dn);
                const T alpha = T(1) / cos_phi;
                m.alpha_d[k] = std::abs(alpha / (dn * T(2)));
                m.len6[k] = std::abs(std::sqrt(dot_dyn(to_face, to_face)));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Operators (include/grad_utils.hpp and src/fvm3d.cpp evaluate())
// ---------------------------------------------------------------------------

// Face value of `phi` on face fc of cell i under linear interpolation
// (grad_utils.hpp:23-36).  `wall_axis` >= 0 selects the reflective-wall rule used
// by the first-order acoustic system: on boundary faces normal to that axis the
// face value is zero (antisymmetric mirror).  wall_axis < 0 is the reference
// rule, neighbour = self, which gives phi_f = phi_c.
template <class T>
inline T face_linear(const HexMesh<T>& m, const T* phi, int64_t i, int fc, int wall_axis) {
    const size_t k = size_t(fc) * size_t(m.N) + size_t(i);
    if (wall_axis >= 0 && m.ids_bf[k] < 0 && (fc >> 1) == (2 - wall_axis)) return T(0);
    return T((phi[i] + phi[m.ids[k]]) * 0.5);
}

// Green-Gauss sum for component d over the six faces of cell i
// (grad_utils.hpp:82-99).  `axis_normals` = the reference's use_alt_normals=true
// (that flag selects the axis normals).  The row sum runs fc = 0..5 from 0.
template <class T>
inline T gauss_component(const HexMesh<T>& m, const T* face_val6_i /*6 values*/, int64_t i, int d, bool axis_normals) {
    const std::vector<T>* nrm;
    if (axis_normals) nrm = d == 0 ? &m.nx : d == 1 ? &m.ny : &m.nz;
    else nrm = d == 0 ? &m.nax : d == 1 ? &m.nay : &m.naz;
    T acc = T(0);
    for (int fc = 0; fc < 6; ++fc) {
        const size_t k = size_t(fc) * size_t(m.N) + size_t(i);
        T t = face_val6_i[fc] * (*nrm)[k];
        t = t * m.area6[k];
        acc = acc + t;
    }
    return acc / m.vol[i];
}

// GradVar::evaluate (src/fvm3d.cpp:808-823): linear interpolation, Gauss sum with
// the outward normals, then res = 0 + g_x (+ g_y) (+ g_z) over the selected axes.
template <class T>
void grad(const HexMesh<T>& m, const T* phi, int axis_mask, int wall_axis, T* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m.N; ++i) {
        T fv[6];
        for (int fc = 0; fc < 6; ++fc) fv[fc] = face_linear(m, phi, i, fc, wall_axis);
        T res = T(0);
        for (int d = 0; d < 3; ++d)
            if (axis_mask & (1 << d)) res = res + gauss_component(m, fv, i, d, false);
        out[i] = res;
    }
}

// Grad2Var::evaluate (src/fvm3d.cpp:880-900) with corrected_surface_normal_grad
// (grad_utils.hpp:101-125): (-phi_c + phi_nb) * alpha_d, Gauss sum with the axis
// normals, res = 0 + L_x (+ L_y) (+ L_z).
template <class T>
void lapl(const HexMesh<T>& m, const T* phi, int axis_mask, T* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m.N; ++i) {
        T fv[6];
        for (int fc = 0; fc < 6; ++fc) {
            const size_t k = size_t(fc) * size_t(m.N) + size_t(i);
            fv[fc] = (-phi[i] + phi[m.ids[k]]) * m.alpha_d[k];
        }
        T res = T(0);
        for (int d = 0; d < 3; ++d)
            if (axis_mask & (1 << d)) res = res + gauss_component(m, fv, i, d, true);
        out[i] = res;
    }
}

// StabVar::evaluate (src/fvm3d.cpp:915-936): linear interpolation, outward Gauss
// gradient, upwind face extrapolation phi + len * g_d * n_out_d
// (grad_utils.hpp:38-59), the neighbour's opposite-face value
// (collect_vals_neigh_faces<.., false>, grad_utils.hpp:61-80; boundary reads
// self's opposite face), 0.5 * (own - collected), Gauss with axis normals,
// component d, summed over the selected axes from 0.
template <class T>
void stab(const HexMesh<T>& m, const T* phi, int axis_mask, T* out) {
    const int64_t N = m.N;
    std::vector<T> g[3];
    for (int d = 0; d < 3; ++d) g[d].resize(N);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        T fv[6];
        for (int fc = 0; fc < 6; ++fc) fv[fc] = face_linear(m, phi, i, fc, -1);
        for (int d = 0; d < 3; ++d) g[d][i] = gauss_component(m, fv, i, d, false);
    }
    auto upw = [&](int d, int64_t c, int fc) {
        const size_t k = size_t(fc) * size_t(N) + size_t(c);
        const std::vector<T>& na = d == 0 ? m.nax : d == 1 ? m.nay : m.naz;
        return phi[c] + m.len6[k] * g[d][c] * na[k];
    };
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        T res = T(0);
        for (int d = 0; d < 3; ++d) {
            if (!(axis_mask & (1 << d))) continue;
            T fv[6];
            for (int fc = 0; fc < 6; ++fc) {
                const size_t k = size_t(fc) * size_t(N) + size_t(i);
                const int64_t nb = m.ids[k];
                const T own = upw(d, i, fc);
                const T col = upw(d, nb, opposite_fc(fc));
                fv[fc] = T(0.5) * (own - col);
            }
            res = res + gauss_component(m, fv, i, d, true);
        }
        out[i] = res;
    }
}

// ---------------------------------------------------------------------------
// Time integration (src/fvm3d.cpp:690-765, 944-1069)
// ---------------------------------------------------------------------------

// DtVar::extract (src/fvm3d.cpp:690-692): dt * rhs + current.
template <class T> inline T euler_update(T dt, T rhs, T cur) { return dt * rhs + cur; }

// D2tVar::extract (src/fvm3d.cpp:735-740) once two history entries exist, and
// the fused form's second stage (:719-731): dt * dt * rhs + 2 * current - previous.
template <class T> inline T leapfrog_update(T dt, T rhs, T cur, T prev) {
    return ((dt * dt) * rhs + T(2) * cur) - prev;
}

// When does D2tVar leapfrog at step t = 0, 1, ... of Equation::evaluate?
// Before that it takes the Euler form dt * rhs + current (the same
// DtVar-shaped update for a second-order equation).
//  * fused (KernelBuilder) equations: set_dt shifts
//    second_stage <- second_stage_next <- second_stage_next_next <- 1 before
//    every launch (include/fvm3d.hpp:290-295), all three starting at 0
//    (:306-308), so the
//    kernel leapfrogs from its third launch on;
//  * per-operator equations with stored history (--store): extract leapfrogs
//    once the history holds more than one entry (src/fvm3d.cpp:736), and
//    Equation::evaluate appends one entry per step after the boundary
//    callbacks (:1048, add_history :301-306), so at step t it holds t entries;
//  * per-operator equations without history (skip_history, the default when
//    --store is absent, include/utils3d.hpp:126): never.
// In both leapfrogging forms "previous" is the state at the start of step
// t - 1: the fused kernel saved it as pre_current during that launch, and the
// history's second-to-last entry is the state after step t - 2's callbacks.
enum D2tRule { D2T_FUSED = 0, D2T_HISTORY = 1, D2T_NO_HISTORY = 2 };

inline bool d2t_leapfrogs(int64_t t, int rule) {
    if (rule == D2T_FUSED) {
        int64_t stage = 0, next = 0, next_next = 0;  // D2tVar members, zero-initialised
        for (int64_t k = 0; k <= t; ++k) {           // set_dt before launch k
            stage = next;
            next = next_next;
            next_next = 1;
        }
        return stage != 0;
    }
    if (rule == D2T_HISTORY) return t > 1;  // history.size() == t
    return false;
}

// One step of {d2t(u), '=', k * (d2dx(u) + d2dy(u) [+ d2dz(u)])}: Grad2Var of
// each selected axis (src/fvm3d.cpp:880-900), summed x, y, z from the left,
// times k (the drivers' c * c), then the D2tVar update of the rule above.
template <class T>
void d2t_step(const HexMesh<T>& m, const T* u, const T* prev, T k, int axis_mask, T dt, bool leapfrog, T* out) {
    const int64_t N = m.N;
    std::vector<T> L[3], sum(N);
    bool first = true;
    for (int d = 0; d < 3; ++d) {
        if (!(axis_mask & (1 << d))) continue;
        L[d].resize(N);
        lapl(m, u, 1 << d, L[d].data());
        for (int64_t i = 0; i < N; ++i) sum[i] = first ? L[d][i] : sum[i] + L[d][i];
        first = false;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const T rhs = k * sum[i];
        out[i] = leapfrog ? leapfrog_update(dt, rhs, u[i], prev[i]) : euler_update(dt, rhs, u[i]);
    }
}

// UpdatePolicies::CourantFriedrichsLewy3D (src/fvm3d.cpp:512-527): gamma = 5/3
// regardless of the driver, dl = min(dx, dy, dz).
template <class T>
T cfl3d(T CFL, const HexMesh<T>& m, const T* u, const T* v, const T* w, const T* p, const T* rho) {
    const T gamma = T(5. / 3.);
    const T dl = std::min({m.dx, m.dy, m.dz});
    T mn = std::numeric_limits<T>::infinity();
    for (int64_t i = 0; i < m.N; ++i) {
        const T a = std::sqrt((gamma * p[i]) / rho[i]);
        const T s = std::sqrt((u[i] * u[i] + v[i] * v[i]) + w[i] * w[i]);
        const T val = dl * (T(1) / (a + s));
        mn = std::min(mn, val);
    }
    return CFL * mn;
}

// UpdatePolicies::CourantFriedrichsLewy1D (src/fvm3d.cpp:543-555).
template <class T>
T cfl1d(T CFL, const HexMesh<T>& m, const T* u, const T* p, const T* rho) {
    const T gamma = T(5. / 3.);
    const T dl = std::min({m.dx, m.dy, m.dz});
    T mn = std::numeric_limits<T>::infinity();
    for (int64_t i = 0; i < m.N; ++i) {
        const T a = std::sqrt((gamma * p[i]) / rho[i]);
        const T s = std::sqrt(u[i] * u[i]);
        const T val = dl * (T(1) / (a + s));
        mn = std::min(mn, val);
    }
    return CFL * mn;
}

// ---------------------------------------------------------------------------
// Whole-step solvers for the benchmark systems, written as the reference's
// Equation::evaluate would run them on the CPU: each equation's right-hand
// side is evaluated over the whole mesh and then assigned (Jacobi within an
// equation), the equations run in listed order (Gauss-Seidel across
// equations), boundary callbacks after all equations (src/fvm3d.cpp:1008-1049).
// ---------------------------------------------------------------------------

// Heat: {d1t(T), '=', alpha * lapl(T)} with no boundary callback; zero-gradient
// walls come from neighbour = self.  alpha_field (nullable) replaces the
// constant by a per-cell coefficient.  Operators are materialised per node, as the
// reference CPU path does.
template <class T>
void heat_run(const HexMesh<T>& m, T* Tf, T alpha, const T* alpha_field, T dt, int64_t steps) {
    const int64_t N = m.N;
    std::vector<T> L(N);
    for (int64_t s = 0; s < steps; ++s) {
        lapl(m, Tf, 7, L.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            const T a = alpha_field ? alpha_field[i] : alpha;
            Tf[i] = euler_update(dt, a * L[i], Tf[i]);
        }
    }
}

// First-order acoustic system, equations in order
//   d1t(u) = -d1dx(p); d1t(v) = -d1dy(p); d1t(w) = -d1dz(p);
//   d1t(p) = K * (d1dx(u) + d1dy(v) + d1dz(w)),  K = -c^2 (constant or per cell)
// with reflective walls: the normal velocity's face value is zero on the walls
// normal to it, p is zero-gradient.  Sequential equations make this the
// symplectic Euler scheme.
template <class T>
void wave_run(const HexMesh<T>& m, T* p, T* u, T* v, T* w, T K, const T* K_field, T dt, int64_t steps) {
    const int64_t N = m.N;
    std::vector<T> g(N), gx(N), gy(N), gz(N);
    T* vel[3] = {u, v, w};
    for (int64_t s = 0; s < steps; ++s) {
        for (int d = 0; d < 3; ++d) {
            grad(m, p, 1 << d, -1, g.data());
            T* q = vel[d];
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < N; ++i) q[i] = euler_update(dt, -g[i], q[i]);
        }
        grad(m, u, 1, 0, gx.data());
        grad(m, v, 2, 1, gy.data());
        grad(m, w, 4, 2, gz.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            const T div = (gx[i] + gy[i]) + gz[i];
            const T k = K_field ? K_field[i] : K;
            p[i] = euler_update(dt, k * div, p[i]);
        }
    }
}

// The same steps as wave_run in one fused pass per stage: each cell forms its
// own face values and Gauss sums in the operators' order (grad's res = 0 + g),
// so the results equal wave_run bit for bit without the whole-mesh
// temporaries (fewer face values than the operator form: 24 per cell and step
// instead of 36).  Timed beside wave_run as a CPU baseline.
template <class T>
void wave_run_fused(const HexMesh<T>& m, T* p, T* u, T* v, T* w, T K, const T* K_field, T dt, int64_t steps) {
    const int64_t N = m.N;
    T* vel[3] = {u, v, w};
    for (int64_t s = 0; s < steps; ++s) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            T fv[6];
            for (int fc = 0; fc < 6; ++fc) fv[fc] = face_linear(m, p, i, fc, -1);
            for (int d = 0; d < 3; ++d) {
                const T g = T(0) + gauss_component(m, fv, i, d, false);
                vel[d][i] = euler_update(dt, -g, vel[d][i]);
            }
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            T gd[3];
            for (int d = 0; d < 3; ++d) {
                T fv[6];
                for (int fc = 0; fc < 6; ++fc) fv[fc] = face_linear(m, vel[d], i, fc, d);
                gd[d] = T(0) + gauss_component(m, fv, i, d, false);
            }
            const T div = (gd[0] + gd[1]) + gd[2];
            const T k = K_field ? K_field[i] : K;
            p[i] = euler_update(dt, k * div, p[i]);
        }
    }
}

// ---------------------------------------------------------------------------
// Compressible Euler with a Rusanov face flux.  No reference counterpart: the
// reference's Sod driver (examples/main_3d_sod_shock.cpp:150-167) uses a
// central predictor plus stabilisation; this stepper is the finite-volume
// scheme the Sod workload asks for, written on the reference mesh geometry
// (Face3D normal from the first three vertices, sum-of-components area, cell
// volume, src/mesh3d.cpp:43-105).  Walls are reflective: the state across a
// wall face is the cell's own with the momentum component along the wall's
// axis negated (the walls are axis planes: only interior vertices move), so no
// mass or energy crosses a wall.
//
// State per cell U = (rho, rho u, rho v, rho w, E).  On the face between the
// lower cell L and upper cell R along an axis, n is the unit axis-direction
// normal (the reference's _normals) and A the face area:
//   un = (u n_x + v n_y) + w n_z,  F(U) = (rho un, rho u un + p n_x,
//        rho v un + p n_y, rho w un + p n_z, (E + p) un),
//   lam = max(|un_L| + a_L, |un_R| + a_R),
//   G_k = ((F_L,k + F_R,k) - lam (U_R,k - U_L,k)) (A / 2).
// Update: net_k = ((Gx+ - Gx-) + (Gy+ - Gy-)) + (Gz+ - Gz-),
//         U_k <- U_k - (dt net_k) (1 / V)  (one division per cell; the
//         primitive velocities and sound speed use 1 / rho the same way).
// Time step, taken from the state at the start of each step:
//   dt = CFL min_i V_i / s_i,  s_i = sum over fc = 0..5 of A_f (|un_f| + a_i).
// Rounding: every product that is added to something is one fused
// multiply-add, fma(a, b, c) = a b + c rounded once (std::fma here, the same
// instruction on the GPU), in the nesting written out in euler_prim,
// rusanov_flux, euler_dt and euler_step; the scheme is this project's own,
// and its rounding is defined here for the kernels and this restatement
// alike.
// ---------------------------------------------------------------------------

// Geometry of the stepper.  The reference's sum-of-components area and
// the volume built on it (src/mesh3d.cpp:61,101-104) are exact only for
// axis-aligned faces and give negative cell volumes on a mesh whose vertices
// move by 20% of dx, so the stepper uses the vector area of each quad face,
//   S = 0.5 (p2 - p0) x (p3 - p1),  oriented along +axis,  A = |S|, n = S / A,
// (exact for a bilinear face; the six vector areas of a cell sum to zero, so a
// uniform state is preserved) and the cell volume
//   V = (1/3) sum over fc = 0..5 of (x_f - x_c) . S_out,f
// with x_f the face centre and x_c the cell centre (mean of the vertices).  A
// face's axis is that of its local face pair in its first cell (fc >> 1 = 0 z,
// 1 y, 2 x), which on a lattice is its face family, so the same definition
// holds for a mesh given by its connectivity.
template <class T> void compute_fv_geometry(HexMesh<T>& m) {
    const int64_t F = m.F, N = m.N;
    std::vector<int8_t> axis(F, 0);  // 0 x, 1 y, 2 z
    for (int64_t i = 0; i < N; ++i)
        for (int fc = 0; fc < 6; ++fc) {
            const int64_t f = m.cell_faces[6 * i + fc];
            if (m.face_cells[2 * f] == i) axis[f] = int8_t(2 - (fc >> 1));
        }
    m.svx.resize(F); m.svy.resize(F); m.svz.resize(F); m.sva.resize(F); m.svol.resize(N);
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < F; ++f) {
        V3<T> p[4];
        for (int q = 0; q < 4; ++q) {
            const int64_t v = m.face_verts[4 * f + q];
            p[q] = {m.vx[v], m.vy[v], m.vz[v]};
        }
        const V3<T> c = vcross(vsub(p[2], p[0]), vsub(p[3], p[1]));
        V3<T> S{T(0.5) * c.x, T(0.5) * c.y, T(0.5) * c.z};
        const T along = axis[f] == 2 ? S.z : (axis[f] == 1 ? S.y : S.x);
        if (along < T(0)) S = {-S.x, -S.y, -S.z};
        const T A = std::sqrt(dot_fixed(S, S));
        m.svx[f] = S.x / A; m.svy[f] = S.y / A; m.svz[f] = S.z / A; m.sva[f] = A;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const V3<T> xc{m.ccx[i], m.ccy[i], m.ccz[i]};
        T acc = T(0);
        for (int fc = 0; fc < 6; ++fc) {
            const int64_t f = m.cell_faces[6 * i + fc];
            const T sgn = (fc & 1) ? T(1) : T(-1);
            const T A = m.sva[f];
            const V3<T> So{sgn * (m.svx[f] * A), sgn * (m.svy[f] * A), sgn * (m.svz[f] * A)};
            acc = acc + dot_fixed(vsub(V3<T>{m.fcx[f], m.fcy[f], m.fcz[f]}, xc), So);
        }
        m.svol[i] = acc / T(3);
    }
}

template <class T> struct EulerPrim { T ux, uy, uz, p, a; };

// One division per state: the reciprocal density multiplies the momenta and
// the sound-speed argument.
template <class T> inline EulerPrim<T> euler_prim(T gamma, T r, T mx, T my, T mz, T E) {
    EulerPrim<T> q;
    const T ir = T(1) / r;
    q.ux = mx * ir;
    q.uy = my * ir;
    q.uz = mz * ir;
    const T q2 = std::fma(q.uz, q.uz, std::fma(q.uy, q.uy, q.ux * q.ux));
    q.p = (gamma - T(1)) * std::fma(-(T(0.5) * r), q2, E);
    q.a = std::sqrt((gamma * q.p) * ir);
    return q;
}

// u . n of a velocity and a face normal
template <class T> inline T euler_un(T ux, T uy, T uz, T nx, T ny, T nz) {
    return std::fma(uz, nz, std::fma(uy, ny, ux * nx));
}

template <class T>
inline void rusanov_flux(const T UL[5], const EulerPrim<T>& L, const T UR[5], const EulerPrim<T>& R, T nx, T ny,
                         T nz, T A, T G[5]) {
    const T unL = euler_un(L.ux, L.uy, L.uz, nx, ny, nz);
    const T unR = euler_un(R.ux, R.uy, R.uz, nx, ny, nz);
    const T FL[5] = {UL[0] * unL, std::fma(UL[1], unL, L.p * nx), std::fma(UL[2], unL, L.p * ny),
                     std::fma(UL[3], unL, L.p * nz), (UL[4] + L.p) * unL};
    const T FR[5] = {UR[0] * unR, std::fma(UR[1], unR, R.p * nx), std::fma(UR[2], unR, R.p * ny),
                     std::fma(UR[3], unR, R.p * nz), (UR[4] + R.p) * unR};
    const T lam = std::max(std::abs(unL) + L.a, std::abs(unR) + R.a);
    const T hA = T(0.5) * A;
    for (int k = 0; k < 5; ++k) G[k] = std::fma(-lam, UR[k] - UL[k], FL[k] + FR[k]) * hA;
}

// +axis unit normal and area of local face fc of cell i
template <class T> inline void face_na(const HexMesh<T>& m, int64_t i, int fc, T& nx, T& ny, T& nz, T& A) {
    const int64_t f = m.cell_faces[6 * i + fc];
    nx = m.svx[f]; ny = m.svy[f]; nz = m.svz[f]; A = m.sva[f];
}

// CFL time step of the Rusanov scheme (formula above)
template <class T> T euler_dt(const HexMesh<T>& m, const T* const U[5], T gamma, T CFL) {
    T mn = std::numeric_limits<T>::infinity();
#pragma omp parallel for schedule(static) reduction(min : mn)
    for (int64_t i = 0; i < m.N; ++i) {
        const EulerPrim<T> q = euler_prim(gamma, U[0][i], U[1][i], U[2][i], U[3][i], U[4][i]);
        T s = T(0);
        for (int fc = 0; fc < 6; ++fc) {
            T nx, ny, nz, A;
            face_na(m, i, fc, nx, ny, nz, A);
            const T un = euler_un(q.ux, q.uy, q.uz, nx, ny, nz);
            s = std::fma(A, std::abs(un) + q.a, s);
        }
        mn = std::min(mn, m.svol[i] / s);
    }
    return CFL * mn;
}

// One Rusanov step with time step dt; U is updated in place (through a copy).
template <class T> void euler_step(const HexMesh<T>& m, T* const U[5], T gamma, T dt) {
    const int64_t N = m.N;
    std::vector<T> old(5 * size_t(N));
    for (int k = 0; k < 5; ++k) std::memcpy(old.data() + size_t(k) * N, U[k], sizeof(T) * size_t(N));
    auto st = [&](int64_t c, T out[5]) {
        for (int k = 0; k < 5; ++k) out[k] = old[size_t(k) * N + c];
    };
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        T Ui[5];
        st(i, Ui);
        const EulerPrim<T> qi = euler_prim(gamma, Ui[0], Ui[1], Ui[2], Ui[3], Ui[4]);
        T Gf[6][5];
        for (int fc = 0; fc < 6; ++fc) {
            const size_t kf = size_t(fc) * size_t(N) + size_t(i);
            T Un[5];
            if (m.ids_bf[kf] >= 0) {
                st(m.ids_bf[kf], Un);
            } else {  // wall: mirrored state (momentum index 1 = x, 2 = y, 3 = z)
                for (int k = 0; k < 5; ++k) Un[k] = Ui[k];
                Un[3 - (fc >> 1)] = -Ui[3 - (fc >> 1)];
            }
            const EulerPrim<T> qn = euler_prim(gamma, Un[0], Un[1], Un[2], Un[3], Un[4]);
            T nx, ny, nz, A;
            face_na(m, i, fc, nx, ny, nz, A);
            if (fc & 1) rusanov_flux(Ui, qi, Un, qn, nx, ny, nz, A, Gf[fc]);   // i is the lower cell
            else rusanov_flux(Un, qn, Ui, qi, nx, ny, nz, A, Gf[fc]);          // i is the upper cell
        }
        const T iV = T(1) / m.svol[i];
        for (int k = 0; k < 5; ++k) {
            // fc 4/5 = x-/x+, 2/3 = y-/y+, 0/1 = z-/z+
            const T net = ((Gf[5][k] - Gf[4][k]) + (Gf[3][k] - Gf[2][k])) + (Gf[1][k] - Gf[0][k]);
            U[k][i] = std::fma(-(dt * net), iV, Ui[k]);
        }
    }
}

// `steps` Rusanov steps with the CFL time step recomputed from the state at the
// start of each step; dt_hist (nullable) receives every step's dt.
template <class T> void euler_run(const HexMesh<T>& m, T* const U[5], T gamma, T CFL, int64_t steps, T* dt_hist) {
    for (int64_t s = 0; s < steps; ++s) {
        const T dt = euler_dt(m, U, gamma, CFL);
        if (dt_hist) dt_hist[s] = dt;
        euler_step(m, U, gamma, dt);
    }
}

}  // namespace cfdo
