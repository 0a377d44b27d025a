// Host-side geometry of the structured hex mesh, computed in the mesh dtype the
// way the reference computes it, but stored per face family (structure of
// arrays) instead of per-entity shared_ptr objects and N x 6 tables.
//
// Reference: Mesh3D::init_basic_internals (src/mesh3d.cpp:323-385),
// Face3D::compute (:43-62), Quadrangle3D::compute (:69-105) and the table
// assembly of Mesh3D::compute (:151-321).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace cfd {

template <class T> struct P3 { T x, y, z; };
template <class T> inline P3<T> sub3(P3<T> a, P3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline P3<T> cross3(P3<T> a, P3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Eigen reduces fixed-size 3-vectors as a0 + (a1 + a2) (used by Face3D and
// Quadrangle3D, whose vectors are Vector3f) and dynamic-size ones left to right
// (used by Mesh3D::compute, whose vectors are VectorXf).
template <class T> inline T dotf(P3<T> a, P3<T> b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
template <class T> inline T dotd(P3<T> a, P3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

template <class T> struct HexGeometry {
    int64_t X = 0, Y = 0, Z = 0, N = 0, V = 0;
    std::vector<T> vx, vy, vz;
    // face families: 0 = z-faces, 1 = y-faces, 2 = x-faces
    int64_t nf[3] = {0, 0, 0};
    std::vector<T> fcx[3], fcy[3], fcz[3];   // face centre
    std::vector<T> fn[3][3];                 // [family][comp] face normal (Face3D::_normal)
    std::vector<T> farea[3];
    std::vector<T> falpha[3];                // alpha_d of the face (same from both sides)
    std::vector<T> ccx, ccy, ccz, vol;
    std::vector<T> len;                      // 6 x N, fc-major

    int64_t vid(int64_t x, int64_t y, int64_t z) const { return x * (Y + 1) * (Z + 1) + y * (Z + 1) + z; }
    int64_t zface(int64_t x, int64_t y, int64_t zf) const { return x * Y * (Z + 1) + y * (Z + 1) + zf; }
    int64_t yface(int64_t x, int64_t yf, int64_t z) const { return x * (Y + 1) * Z + yf * Z + z; }
    int64_t xface(int64_t xf, int64_t y, int64_t z) const { return xf * Y * Z + y * Z + z; }

    // (family, index) of local face fc of cell (x, y, z): fc 0/1 z-, 2/3 y-, 4/5 x-faces.
    void cell_face(int64_t x, int64_t y, int64_t z, int fc, int& fam, int64_t& idx) const {
        switch (fc) {
            case 0: fam = 0; idx = zface(x, y, z); break;
            case 1: fam = 0; idx = zface(x, y, z + 1); break;
            case 2: fam = 1; idx = yface(x, y, z); break;
            case 3: fam = 1; idx = yface(x, y + 1, z); break;
            case 4: fam = 2; idx = xface(x, y, z); break;
            default: fam = 2; idx = xface(x + 1, y, z); break;
        }
    }
    P3<T> vert(int64_t v) const { return {vx[v], vy[v], vz[v]}; }
    P3<T> fnorm(int fam, int64_t f) const { return {fn[fam][0][f], fn[fam][1][f], fn[fam][2][f]}; }
    P3<T> fcen(int fam, int64_t f) const { return {fcx[fam][f], fcy[fam][f], fcz[fam][f]}; }
    P3<T> ccen(int64_t c) const { return {ccx[c], ccy[c], ccz[c]}; }
    // Reference _normals (axis direction): -n on fc 0,1,4,5 and +n on 2,3.
    static P3<T> axis_normal(P3<T> n, int fc) {
        if (fc == 2 || fc == 3) return n;
        return {-n.x, -n.y, -n.z};
    }
};

// Builds vertex coordinates (x*dx in T, plus the optional offsets) and every
// geometric quantity the kernels or the reference tables need.
template <class T>
void build_hex_geometry(HexGeometry<T>& g, int64_t X, int64_t Y, int64_t Z, T dx, T dy, T dz, const double* offsets) {
    g.X = X; g.Y = Y; g.Z = Z; g.N = X * Y * Z; g.V = (X + 1) * (Y + 1) * (Z + 1);
    g.vx.resize(g.V); g.vy.resize(g.V); g.vz.resize(g.V);
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x <= X; ++x)
        for (int64_t y = 0; y <= Y; ++y)
            for (int64_t z = 0; z <= Z; ++z) {
                const int64_t v = g.vid(x, y, z);
                T px = T(x) * dx, py = T(y) * dy, pz = T(z) * dz;
                if (offsets) {
                    px = px + T(offsets[3 * v + 0]);
                    py = py + T(offsets[3 * v + 1]);
                    pz = pz + T(offsets[3 * v + 2]);
                }
                g.vx[v] = px; g.vy[v] = py; g.vz[v] = pz;
            }

    g.nf[0] = X * Y * (Z + 1);
    g.nf[1] = X * (Y + 1) * Z;
    g.nf[2] = (X + 1) * Y * Z;
    for (int fam = 0; fam < 3; ++fam) {
        const int64_t n = g.nf[fam];
        g.fcx[fam].resize(n); g.fcy[fam].resize(n); g.fcz[fam].resize(n);
        for (int c = 0; c < 3; ++c) g.fn[fam][c].resize(n);
        g.farea[fam].resize(n);
        g.falpha[fam].resize(n);
    }

    // Face3D::compute over each family; the vertex quads follow the per-cell
    // lists of src/mesh3d.cpp:361-374 (every cell touching a face writes the
    // same order).
    auto face_compute = [&](int fam, int64_t f, int64_t q0, int64_t q1, int64_t q2, int64_t q3) {
        const P3<T> p0 = g.vert(q0), p1 = g.vert(q1), p2 = g.vert(q2), p3 = g.vert(q3);
        P3<T> c{0, 0, 0};
        c = {c.x + p0.x, c.y + p0.y, c.z + p0.z};
        c = {c.x + p1.x, c.y + p1.y, c.z + p1.z};
        c = {c.x + p2.x, c.y + p2.y, c.z + p2.z};
        c = {c.x + p3.x, c.y + p3.y, c.z + p3.z};
        const P3<T> AB = sub3(p1, p0), BC = sub3(p2, p1), CD = sub3(p3, p2), AD = sub3(p3, p0), CB = sub3(p1, p2);
        P3<T> n = cross3(AB, BC);
        const T sq = dotf(n, n);
        if (sq > T(0)) {
            const T s = std::sqrt(sq);
            n = {n.x / s, n.y / s, n.z / s};
        }
        const P3<T> a1 = cross3(AB, AD), a2 = cross3(CB, CD);
        const T s1 = std::abs(a1.x + (a1.y + a1.z));
        const T s2 = std::abs(a2.x + (a2.y + a2.z));
        g.fcx[fam][f] = c.x / T(4); g.fcy[fam][f] = c.y / T(4); g.fcz[fam][f] = c.z / T(4);
        g.fn[fam][0][f] = n.x; g.fn[fam][1][f] = n.y; g.fn[fam][2][f] = n.z;
        g.farea[fam][f] = T(0.5 * (s1 + s2));
    };
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < X; ++x)
        for (int64_t y = 0; y < Y; ++y)
            for (int64_t zf = 0; zf <= Z; ++zf)
                face_compute(0, g.zface(x, y, zf), g.vid(x, y, zf), g.vid(x, y + 1, zf), g.vid(x + 1, y + 1, zf),
                             g.vid(x + 1, y, zf));
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < X; ++x)
        for (int64_t yf = 0; yf <= Y; ++yf)
            for (int64_t z = 0; z < Z; ++z)
                face_compute(1, g.yface(x, yf, z), g.vid(x, yf, z), g.vid(x, yf, z + 1), g.vid(x + 1, yf, z + 1),
                             g.vid(x + 1, yf, z));
#pragma omp parallel for schedule(static)
    for (int64_t xf = 0; xf <= X; ++xf)
        for (int64_t y = 0; y < Y; ++y)
            for (int64_t z = 0; z < Z; ++z)
                face_compute(2, g.xface(xf, y, z), g.vid(xf, y, z + 1), g.vid(xf, y + 1, z + 1), g.vid(xf, y + 1, z),
                             g.vid(xf, y, z));

    // Quadrangle3D::compute: centre, outward-sign test, volume.
    g.ccx.resize(g.N); g.ccy.resize(g.N); g.ccz.resize(g.N); g.vol.resize(g.N);
    int bad_orientation = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad_orientation)
    for (int64_t x = 0; x < X; ++x)
        for (int64_t y = 0; y < Y; ++y)
            for (int64_t z = 0; z < Z; ++z) {
                const int64_t cid = Z * Y * x + Z * y + z;
                P3<T> c{0, 0, 0};
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        for (int cc = 0; cc < 2; ++cc) {
                            const P3<T> p = g.vert(g.vid(x + a, y + b, z + cc));
                            c = {c.x + p.x, c.y + p.y, c.z + p.z};
                        }
                c = {c.x / T(8), c.y / T(8), c.z / T(8)};
                g.ccx[cid] = c.x; g.ccy[cid] = c.y; g.ccz[cid] = c.z;
                T volume = 0;
                for (int fc = 0; fc < 6; ++fc) {
                    int fam; int64_t f;
                    g.cell_face(x, y, z, fc, fam, f);
                    const P3<T> n = g.fnorm(fam, f), fcen = g.fcen(fam, f);
                    const bool keep = dotf(n, sub3(fcen, c)) >= T(0);
                    const P3<T> good = keep ? n : P3<T>{-n.x, -n.y, -n.z};
                    // The kernels use outward = (fc odd ? +1 : -1) * axis normal;
                    // confirm the reference's per-cell sign test agrees.
                    const P3<T> ax = HexGeometry<T>::axis_normal(n, fc);
                    const P3<T> rule = (fc & 1) ? ax : P3<T>{-ax.x, -ax.y, -ax.z};
                    if (!(rule.x == good.x && rule.y == good.y && rule.z == good.z)) bad_orientation += 1;
                    volume += dotf(fcen, good) * g.farea[fam][f];
                }
                g.vol[cid] = volume / T(3);
            }
    if (bad_orientation)
        throw Error(CFD_E_INVALID, "hex mesh too distorted: a face normal points against its cell's outward side");

    // alpha_d per face (identical from both cells) and centre-to-face lengths.
    g.len.resize(6 * g.N);
    auto alpha_of = [](P3<T> ax, P3<T> d, bool boundary) {
        const T dn = std::sqrt(dotd(d, d));
        const T cos_phi = dotd(ax, d) / (std::sqrt(dotd(ax, ax)) * dn);
        const T alpha = T(1) / cos_phi;
        return boundary ? std::abs(alpha / (dn * T(2))) : std::abs(alpha / dn);
    };
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < X; ++x)
        for (int64_t y = 0; y < Y; ++y)
            for (int64_t z = 0; z < Z; ++z) {
                const int64_t cid = Z * Y * x + Z * y + z;
                const P3<T> c = g.ccen(cid);
                const int64_t crd[3] = {x, y, z};
                const int64_t ext[3] = {X, Y, Z};
                for (int fc = 0; fc < 6; ++fc) {
                    int fam; int64_t f;
                    g.cell_face(x, y, z, fc, fam, f);
                    const P3<T> fcen = g.fcen(fam, f);
                    const P3<T> tf = sub3(fcen, c);
                    g.len[fc * g.N + cid] = std::abs(std::sqrt(dotd(tf, tf)));
                    const int axis = 2 - (fc >> 1);
                    const bool upper = fc & 1;
                    const bool boundary = upper ? (crd[axis] == ext[axis] - 1) : (crd[axis] == 0);
                    const P3<T> ax = HexGeometry<T>::axis_normal(g.fnorm(fam, f), fc);
                    if (boundary) {
                        g.falpha[fam][f] = alpha_of(ax, tf, true);
                    } else if (upper) {  // interior faces are written once, from the lower cell
                        int64_t nb = cid;
                        if (axis == 0) nb += Y * Z;
                        else if (axis == 1) nb += Z;
                        else nb += 1;
                        g.falpha[fam][f] = alpha_of(ax, sub3(g.ccen(nb), c), false);
                    }
                }
            }
}

// Geometry of the Euler stepper (no reference counterpart).  The reference's
// sum-of-components face area and the volume built on it (src/mesh3d.cpp:61,
// 101-104) are exact only for axis-aligned faces and turn negative on a mesh
// whose interior vertices move by 20% of dx, so the stepper uses each quad's
// vector area S = 0.5 (p2 - p0) x (p3 - p1) oriented along +axis, A = |S|,
// n = S / A, and V = (1/3) sum over fc = 0..5 of (x_f - x_c) . S_out,f with
// x_f the face centre and x_c the cell centre.  Output layout: per family
// [nx, ny, nz, A] * nf, then V (N).
template <class T> void build_fv_geometry(const HexGeometry<T>& g, std::vector<T>& out, int64_t off_fam[3],
                                          int64_t& off_vol) {
    int64_t total = 0;
    for (int fam = 0; fam < 3; ++fam) {
        off_fam[fam] = total;
        total += 4 * g.nf[fam];
    }
    off_vol = total;
    total += g.N;
    out.assign(size_t(total), T(0));
    const int64_t X = g.X, Y = g.Y, Z = g.Z;
    auto face = [&](int fam, int64_t f, int64_t q0, int64_t q1, int64_t q2, int64_t q3) {
        const P3<T> c = cross3(sub3(g.vert(q2), g.vert(q0)), sub3(g.vert(q3), g.vert(q1)));
        P3<T> S{T(0.5) * c.x, T(0.5) * c.y, T(0.5) * c.z};
        const T along = fam == 0 ? S.z : (fam == 1 ? S.y : S.x);
        if (along < T(0)) S = {-S.x, -S.y, -S.z};
        const T A = std::sqrt(dotf(S, S));
        const int64_t nf = g.nf[fam], o = off_fam[fam];
        out[o + f] = S.x / A;
        out[o + nf + f] = S.y / A;
        out[o + 2 * nf + f] = S.z / A;
        out[o + 3 * nf + f] = A;
    };
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < X; ++x)
        for (int64_t y = 0; y < Y; ++y)
            for (int64_t zf = 0; zf <= Z; ++zf)
                face(0, g.zface(x, y, zf), g.vid(x, y, zf), g.vid(x, y + 1, zf), g.vid(x + 1, y + 1, zf),
                     g.vid(x + 1, y, zf));
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < X; ++x)
        for (int64_t yf = 0; yf <= Y; ++yf)
            for (int64_t z = 0; z < Z; ++z)
                face(1, g.yface(x, yf, z), g.vid(x, yf, z), g.vid(x, yf, z + 1), g.vid(x + 1, yf, z + 1),
                     g.vid(x + 1, yf, z));
#pragma omp parallel for schedule(static)
    for (int64_t xf = 0; xf <= X; ++xf)
        for (int64_t y = 0; y < Y; ++y)
            for (int64_t z = 0; z < Z; ++z)
                face(2, g.xface(xf, y, z), g.vid(xf, y, z + 1), g.vid(xf, y + 1, z + 1), g.vid(xf, y + 1, z),
                     g.vid(xf, y, z));
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < X; ++x)
        for (int64_t y = 0; y < Y; ++y)
            for (int64_t z = 0; z < Z; ++z) {
                const int64_t cid = Z * Y * x + Z * y + z;
                const P3<T> xc = g.ccen(cid);
                T acc = T(0);
                for (int fc = 0; fc < 6; ++fc) {
                    int fam; int64_t f;
                    g.cell_face(x, y, z, fc, fam, f);
                    const int64_t nf = g.nf[fam], o = off_fam[fam];
                    const T sgn = (fc & 1) ? T(1) : T(-1);
                    const T A = out[o + 3 * nf + f];
                    const P3<T> So{sgn * (out[o + f] * A), sgn * (out[o + nf + f] * A), sgn * (out[o + 2 * nf + f] * A)};
                    acc = acc + dotf(sub3(g.fcen(fam, f), xc), So);
                }
                out[off_vol + cid] = acc / T(3);
            }
}

}  // namespace cfd
