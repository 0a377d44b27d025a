// Host side of the unstructured hex mesh: a mesh given by its connectivity
// (the layout of the reference's mesh.json, include/io_operators_tmpl.hpp:161-212:
// vertex coordinates, each face's vertex quad and cells, each cell's six faces
// in the local order z-, z+, y-, y+, x-, x+ and its eight vertices).
//
// The geometry follows the same reference formulas as the structured builder
// (hexmesh.hpp; src/mesh3d.cpp:43-105,151-321), now keyed by the given
// connectivity instead of lattice positions:
//   face: centre = mean of its 4 vertices, normal = normalised AB x BC (first
//         three vertices), area = 0.5 (|sum(AB x AD)| + |sum(CB x CD)|);
//   cell: centre = mean of its 8 vertices, outward normal of face fc = +-n by
//         the sign of n . (face centre - cell centre), volume = sum over fc of
//         (face centre . outward n) * area / 3, axis normal of fc = +n on
//         fc 2, 3 and -n on fc 0, 1, 4, 5;
//   alpha_d of a face = |1 / (cos(phi) |d|)| with d the centre-to-centre
//         vector (interior) or |1 / (cos(phi) 2 |d|)| with d centre-to-face
//         (boundary); the same number from both cells of a face.
//
// Cells can be renumbered by reverse Cuthill-McKee (BFS from a cell of least
// degree, neighbours by increasing degree, order reversed) so that a cell's
// neighbours sit close to it in memory; faces are then numbered by their
// lowest new cell id.  perm[new] = old for both.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <queue>
#include <string>
#include <vector>

#include "common.hpp"
#include "hexmesh.hpp"

namespace cfd {


// Connectivity checks: ids in range, every face of a cell lists the cell, a
// face has one or two cells.
inline void check_uconn(const UConn& c) {
    CFD_REQUIRE(c.N > 0 && c.F > 0 && c.V > 0, "unstructured mesh: empty mesh");
    CFD_REQUIRE(c.N < (int64_t(1) << 31) && c.F < (int64_t(1) << 31), "unstructured mesh: more than 2^31 cells or faces");
    for (int64_t f = 0; f < c.F; ++f) {
        for (int q = 0; q < 4; ++q)
            CFD_REQUIRE(c.face_verts[4 * f + q] >= 0 && c.face_verts[4 * f + q] < c.V,
                        "unstructured mesh: face vertex id out of range");
        const int64_t a = c.face_cells[2 * f], b = c.face_cells[2 * f + 1];
        CFD_REQUIRE(a >= 0 && a < c.N, "unstructured mesh: a face without a first cell");
        CFD_REQUIRE(b == -1 || (b >= 0 && b < c.N && b != a), "unstructured mesh: bad second cell of a face");
    }
    for (int64_t i = 0; i < c.N; ++i) {
        for (int q = 0; q < 8; ++q)
            CFD_REQUIRE(c.cell_verts[8 * i + q] >= 0 && c.cell_verts[8 * i + q] < c.V,
                        "unstructured mesh: cell vertex id out of range");
        for (int fc = 0; fc < 6; ++fc) {
            const int64_t f = c.cell_faces[6 * i + fc];
            CFD_REQUIRE(f >= 0 && f < c.F, "unstructured mesh: cell face id out of range");
            CFD_REQUIRE(c.face_cells[2 * f] == i || c.face_cells[2 * f + 1] == i,
                        "unstructured mesh: a cell's face does not list the cell");
        }
    }
}

template <class T> struct UGeometry {
    std::vector<T> fcx, fcy, fcz, fnx, fny, fnz, farea, falpha;  // per face
    std::vector<T> ccx, ccy, ccz, vol;                           // per cell
    std::vector<int8_t> out_sign;                                // N x 6: outward normal = out_sign * n
    std::vector<T> len;                                          // N x 6 centre-to-face distance
};

template <class T> void build_ugeometry(UGeometry<T>& g, const UConn& c) {
    const int64_t N = c.N, F = c.F;
    auto vert = [&](int64_t v) { return P3<T>{T(c.verts[3 * v]), T(c.verts[3 * v + 1]), T(c.verts[3 * v + 2])}; };
    g.fcx.resize(F); g.fcy.resize(F); g.fcz.resize(F);
    g.fnx.resize(F); g.fny.resize(F); g.fnz.resize(F);
    g.farea.resize(F); g.falpha.resize(F);
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < F; ++f) {
        const P3<T> p0 = vert(c.face_verts[4 * f]), p1 = vert(c.face_verts[4 * f + 1]);
        const P3<T> p2 = vert(c.face_verts[4 * f + 2]), p3 = vert(c.face_verts[4 * f + 3]);
        P3<T> s{0, 0, 0};
        s = {s.x + p0.x, s.y + p0.y, s.z + p0.z};
        s = {s.x + p1.x, s.y + p1.y, s.z + p1.z};
        s = {s.x + p2.x, s.y + p2.y, s.z + p2.z};
        s = {s.x + p3.x, s.y + p3.y, s.z + p3.z};
        const P3<T> AB = sub3(p1, p0), BC = sub3(p2, p1), CD = sub3(p3, p2), AD = sub3(p3, p0), CB = sub3(p1, p2);
        P3<T> n = cross3(AB, BC);
        const T sq = dotf(n, n);
        if (sq > T(0)) {
            const T r = std::sqrt(sq);
            n = {n.x / r, n.y / r, n.z / r};
        }
        const P3<T> a1 = cross3(AB, AD), a2 = cross3(CB, CD);
        const T s1 = std::abs(a1.x + (a1.y + a1.z));
        const T s2 = std::abs(a2.x + (a2.y + a2.z));
        g.fcx[f] = s.x / T(4); g.fcy[f] = s.y / T(4); g.fcz[f] = s.z / T(4);
        g.fnx[f] = n.x; g.fny[f] = n.y; g.fnz[f] = n.z;
        g.farea[f] = T(0.5 * (s1 + s2));
    }
    g.ccx.resize(N); g.ccy.resize(N); g.ccz.resize(N); g.vol.resize(N);
    g.out_sign.resize(6 * N);
    g.len.resize(6 * N);
    auto fnorm = [&](int64_t f) { return P3<T>{g.fnx[f], g.fny[f], g.fnz[f]}; };
    auto fcen = [&](int64_t f) { return P3<T>{g.fcx[f], g.fcy[f], g.fcz[f]}; };
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        P3<T> s{0, 0, 0};
        for (int q = 0; q < 8; ++q) {
            const P3<T> p = vert(c.cell_verts[8 * i + q]);
            s = {s.x + p.x, s.y + p.y, s.z + p.z};
        }
        const P3<T> cc{s.x / T(8), s.y / T(8), s.z / T(8)};
        g.ccx[i] = cc.x; g.ccy[i] = cc.y; g.ccz[i] = cc.z;
        T volume = 0;
        for (int fc = 0; fc < 6; ++fc) {
            const int64_t f = c.cell_faces[6 * i + fc];
            const P3<T> n = fnorm(f), fce = fcen(f);
            const bool keep = dotf(n, sub3(fce, cc)) >= T(0);
            g.out_sign[6 * i + fc] = keep ? 1 : -1;
            const P3<T> good = keep ? n : P3<T>{-n.x, -n.y, -n.z};
            volume += dotf(fce, good) * g.farea[f];
            const P3<T> tf = sub3(fce, cc);
            g.len[6 * i + fc] = std::abs(std::sqrt(dotd(tf, tf)));
        }
        g.vol[i] = volume / T(3);
    }
    // alpha_d, from the face's first cell (its axis normal and the vector to
    // the other cell's centre, or to the face centre on the boundary)
    auto alpha_of = [](P3<T> ax, P3<T> d, bool boundary) {
        const T dn = std::sqrt(dotd(d, d));
        const T cos_phi = dotd(ax, d) / (std::sqrt(dotd(ax, ax)) * dn);
        const T alpha = T(1) / cos_phi;
        return boundary ? std::abs(alpha / (dn * T(2))) : std::abs(alpha / dn);
    };
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < F; ++f) {
        const int64_t a = c.face_cells[2 * f], b = c.face_cells[2 * f + 1];
        int fc = 0;
        while (fc < 6 && c.cell_faces[6 * a + fc] != f) ++fc;
        const P3<T> ax = HexGeometry<T>::axis_normal(fnorm(f), fc);
        const P3<T> ca{g.ccx[a], g.ccy[a], g.ccz[a]};
        if (b < 0) g.falpha[f] = alpha_of(ax, sub3(fcen(f), ca), true);
        else g.falpha[f] = alpha_of(ax, sub3(P3<T>{g.ccx[b], g.ccy[b], g.ccz[b]}, ca), false);
    }
}

// Geometry of the Rusanov stepper on a mesh given by its connectivity, in the
// caller's numbering (the structured builder's formulas, build_fv_geometry):
// per face the vector area S = 0.5 (p2 - p0) x (p3 - p1) oriented along +axis,
// the axis being that of the face's local pair in its first cell (fc >> 1 =
// 0 z, 1 y, 2 x), stored as unit normal and area (sv: nx F, ny F, nz F, A F);
// per cell V = (1/3) sum over fc of (x_f - x_c) . S_out with S_out = -S on
// the lower face of a pair and +S on the upper.
template <class T>
void build_ufv_geometry(const UConn& c, const UGeometry<T>& g, std::vector<T>& sv, std::vector<T>& svol) {
    const int64_t N = c.N, F = c.F;
    std::vector<int8_t> axis(F, 0);  // 0 x, 1 y, 2 z
    for (int64_t i = 0; i < N; ++i)
        for (int fc = 0; fc < 6; ++fc) {
            const int64_t f = c.cell_faces[6 * i + fc];
            if (c.face_cells[2 * f] == i) axis[f] = int8_t(2 - (fc >> 1));
        }
    auto vert = [&](int64_t v) { return P3<T>{T(c.verts[3 * v]), T(c.verts[3 * v + 1]), T(c.verts[3 * v + 2])}; };
    sv.assign(4 * size_t(F), T(0));
    svol.assign(size_t(N), T(0));
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < F; ++f) {
        const P3<T> p0 = vert(c.face_verts[4 * f]), p1 = vert(c.face_verts[4 * f + 1]);
        const P3<T> p2 = vert(c.face_verts[4 * f + 2]), p3 = vert(c.face_verts[4 * f + 3]);
        const P3<T> cr = cross3(sub3(p2, p0), sub3(p3, p1));
        P3<T> S{T(0.5) * cr.x, T(0.5) * cr.y, T(0.5) * cr.z};
        const T along = axis[f] == 2 ? S.z : (axis[f] == 1 ? S.y : S.x);
        if (along < T(0)) S = {-S.x, -S.y, -S.z};
        const T A = std::sqrt(dotf(S, S));
        sv[f] = S.x / A;
        sv[F + f] = S.y / A;
        sv[2 * F + f] = S.z / A;
        sv[3 * F + f] = A;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const P3<T> xc{g.ccx[i], g.ccy[i], g.ccz[i]};
        T acc = T(0);
        for (int fc = 0; fc < 6; ++fc) {
            const int64_t f = c.cell_faces[6 * i + fc];
            const T sgn = (fc & 1) ? T(1) : T(-1);
            const T A = sv[3 * F + f];
            const P3<T> So{sgn * (sv[f] * A), sgn * (sv[F + f] * A), sgn * (sv[2 * F + f] * A)};
            acc = acc + dotf(sub3(P3<T>{g.fcx[f], g.fcy[f], g.fcz[f]}, xc), So);
        }
        svol[i] = acc / T(3);
    }
}

// Reverse Cuthill-McKee order of the cells (perm[new] = old), every connected
// component in turn, each started from an unvisited cell of least degree.
inline std::vector<int64_t> rcm_order(const UConn& c) {
    const int64_t N = c.N;
    std::vector<int> deg(N, 0);
    for (int64_t i = 0; i < N; ++i)
        for (int fc = 0; fc < 6; ++fc) {
            const int64_t f = c.cell_faces[6 * i + fc];
            if (c.face_cells[2 * f + 1] >= 0) ++deg[i];
        }
    auto nbr = [&](int64_t i, int fc) -> int64_t {
        const int64_t f = c.cell_faces[6 * i + fc];
        const int64_t a = c.face_cells[2 * f], b = c.face_cells[2 * f + 1];
        if (b < 0) return -1;
        return a == i ? b : a;
    };
    std::vector<int64_t> order;
    order.reserve(N);
    std::vector<char> seen(N, 0);
    std::vector<int64_t> by_degree(N);
    std::iota(by_degree.begin(), by_degree.end(), 0);
    std::stable_sort(by_degree.begin(), by_degree.end(), [&](int64_t a, int64_t b) { return deg[a] < deg[b]; });
    for (int64_t start : by_degree) {
        if (seen[start]) continue;
        std::queue<int64_t> q;
        q.push(start);
        seen[start] = 1;
        while (!q.empty()) {
            const int64_t i = q.front();
            q.pop();
            order.push_back(i);
            int64_t nb[6];
            int k = 0;
            for (int fc = 0; fc < 6; ++fc) {
                const int64_t j = nbr(i, fc);
                if (j >= 0 && !seen[j]) {
                    seen[j] = 1;
                    nb[k++] = j;
                }
            }
            std::stable_sort(nb, nb + k, [&](int64_t a, int64_t b) { return deg[a] < deg[b]; });
            for (int t = 0; t < k; ++t) q.push(nb[t]);
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// Face order for a cell order: faces by (lowest new cell id, other new cell
// id), boundary faces after the interior faces of the same cell.
inline std::vector<int64_t> face_order(const UConn& c, const std::vector<int64_t>& cell_iperm) {
    std::vector<int64_t> f(c.F);
    std::iota(f.begin(), f.end(), 0);
    auto key = [&](int64_t x) {
        const int64_t a = cell_iperm[c.face_cells[2 * x]];
        const int64_t b = c.face_cells[2 * x + 1] < 0 ? INT64_MAX : cell_iperm[c.face_cells[2 * x + 1]];
        return std::make_pair(std::min(a, b), std::max(a, b));
    };
    std::stable_sort(f.begin(), f.end(), [&](int64_t x, int64_t y) { return key(x) < key(y); });
    return f;
}

}  // namespace cfd
