// -m/--mesh input of the drivers: reads a mesh.json file (the layout the
// stepping output writes, include/io_operators_tmpl.hpp:161-212 of the
// reference, and its -m readers take, include/io_operators_tmpl.hpp:215-247,
// legacy/io_operators.cpp:170-205) into a Mesh3D.  The operator API builds its
// kernels for the lattice numbering, so the file's connectivity must be that
// of Mesh3D(x, y, z, ...) (src/mesh3d.cpp:323-416); its vertex coordinates may
// be anywhere, and become the mesh's vertex offsets.  Meshes of any other
// connectivity go through the C-ABI's cfd_mesh_create_unstructured.
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "io_operators_tmpl.hpp"
#include "mesh3d.hpp"

namespace cfdarco::io {

namespace {

// A JSON value: enough of RFC 8259 for mesh files (numbers kept as double).
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<JVal> arr;
    std::map<std::string, JVal> obj;
};

class JParser {
public:
    explicit JParser(const std::string& s) : s_(s) {}
    JVal parse() {
        JVal v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error("mesh.json: " + std::string(what) + " at byte " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\n' || s_[i_] == '\r' || s_[i_] == '\t')) ++i_;
    }
    bool eat(char c) {
        ws();
        if (i_ < s_.size() && s_[i_] == c) {
            ++i_;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail("unexpected character");
    }
    std::string string() {
        expect('"');
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            if (s_[i_] == '\\') {
                ++i_;
                if (i_ >= s_.size()) fail("bad escape");
                const char e = s_[i_];
                out += (e == 'n') ? '\n' : (e == 't') ? '\t' : e;
            } else {
                out += s_[i_];
            }
            ++i_;
        }
        if (i_ >= s_.size()) fail("unterminated string");
        ++i_;
        return out;
    }
    JVal value() {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        JVal v;
        const char c = s_[i_];
        if (c == '{') {
            ++i_;
            v.kind = JVal::Obj;
            if (eat('}')) return v;
            do {
                ws();
                std::string k = string();
                expect(':');
                v.obj[k] = value();
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            ++i_;
            v.kind = JVal::Arr;
            if (eat(']')) return v;
            do v.arr.push_back(value());
            while (eat(','));
            expect(']');
        } else if (c == '"') {
            v.kind = JVal::Str;
            v.str = string();
        } else if (s_.compare(i_, 4, "true") == 0) {
            v.kind = JVal::Bool; v.b = true; i_ += 4;
        } else if (s_.compare(i_, 5, "false") == 0) {
            v.kind = JVal::Bool; i_ += 5;
        } else if (s_.compare(i_, 4, "null") == 0) {
            i_ += 4;
        } else {
            const char* begin = s_.c_str() + i_;
            char* end = nullptr;
            v.num = std::strtod(begin, &end);
            if (end == begin) fail("bad value");
            v.kind = JVal::Num;
            i_ += size_t(end - begin);
        }
        return v;
    }
    const std::string& s_;
    size_t i_ = 0;
};

const JVal& member(const JVal& o, const char* k) {
    auto it = o.obj.find(k);
    if (o.kind != JVal::Obj || it == o.obj.end()) throw std::runtime_error(std::string("mesh.json: no \"") + k + "\"");
    return it->second;
}
size_t as_size(const JVal& v) {
    if (v.kind != JVal::Num || v.num < 0 || v.num != std::floor(v.num)) throw std::runtime_error("mesh.json: bad id");
    return size_t(v.num);
}
std::vector<size_t> ids(const JVal& a, size_t n) {
    if (a.kind != JVal::Arr || (n && a.arr.size() != n)) throw std::runtime_error("mesh.json: bad id list");
    std::vector<size_t> out;
    for (const auto& v : a.arr) out.push_back(as_size(v));
    return out;
}

}  // namespace

LatticeMeshFile read_lattice_mesh_json(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("cannot open mesh file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    const JVal root = JParser(text).parse();
    const size_t X = as_size(member(root, "x")), Y = as_size(member(root, "y")), Z = as_size(member(root, "z"));
    if (!X || !Y || !Z) throw std::runtime_error("mesh.json: empty extents");
    const JVal& verts = member(root, "vertexes");
    const JVal& faces = member(root, "face");
    const JVal& nodes = member(root, "nodes");
    const size_t V = (X + 1) * (Y + 1) * (Z + 1), N = X * Y * Z;
    const size_t F = X * Y * (Z + 1) + X * (Y + 1) * Z + (X + 1) * Y * Z;
    const std::string other =
        "mesh.json: the connectivity is not that of Mesh3D(x, y, z) (the operator API takes lattice numbering; "
        "other hex meshes go through cfd_mesh_create_unstructured)";
    if (verts.arr.size() != V || faces.arr.size() != F || nodes.arr.size() != N) throw std::runtime_error(other);
    auto coord = [&](size_t v, int k) {
        const JVal& p = verts.arr[v];
        if (p.kind != JVal::Arr || p.arr.size() != 3 || p.arr[k].kind != JVal::Num)
            throw std::runtime_error("mesh.json: bad vertex");
        return float(p.arr[k].num);
    };
    // spacing from the first vertex along each axis (the lattice walls are not
    // moved by the interior perturbation)
    const float dx = coord((Y + 1) * (Z + 1), 0), dy = coord(Z + 1, 1), dz = coord(1, 2);
    auto mesh = std::make_unique<Mesh3D>(X, Y, Z, dx, dy, dz);  // for the closed-form numbering
    // connectivity: the closed forms of src/mesh3d.cpp:387-416 and the quads of
    // :361-374
    for (size_t x = 0; x < X; ++x)
        for (size_t y = 0; y < Y; ++y)
            for (size_t z = 0; z < Z; ++z) {
                const size_t c = mesh->square_node_coord_to_idx(x, y, z);
                const auto fc = mesh->square_node_coord_to_face_idx(x, y, z);
                const auto vx = mesh->square_node_coord_to_vertex_idx(x, y, z);
                const JVal& n = nodes.arr[c];
                const auto nf = ids(member(n, "face"), 6), nv = ids(member(n, "vertexes"), 8);
                for (int k = 0; k < 6; ++k)
                    if (nf[k] != fc[k]) throw std::runtime_error(other);
                for (int k = 0; k < 8; ++k)
                    if (nv[k] != vx[k]) throw std::runtime_error(other);
                static const int quad[6][4] = {{0, 2, 6, 4}, {1, 3, 7, 5}, {0, 1, 5, 4},
                                               {2, 3, 7, 6}, {1, 3, 2, 0}, {5, 7, 6, 4}};
                for (int k = 0; k < 6; ++k) {
                    const auto q = ids(member(faces.arr[fc[k]], "vertexes_id"), 4);
                    for (int j = 0; j < 4; ++j)
                        if (q[j] != vx[quad[k][j]]) throw std::runtime_error(other);
                }
            }
    // vertex offsets: coordinate - x * dx, exact in float for moves within a
    // factor of two of the lattice position; any coordinate they cannot
    // reproduce is refused
    std::vector<double> off(3 * V, 0.0);
    const float d[3] = {dx, dy, dz};
    bool any = false;
    size_t v = 0;
    for (size_t x = 0; x <= X; ++x)
        for (size_t y = 0; y <= Y; ++y)
            for (size_t z = 0; z <= Z; ++z, ++v) {
                const size_t lat[3] = {x, y, z};
                for (int k = 0; k < 3; ++k) {
                    const float base = float(lat[k]) * d[k];
                    const float want = coord(v, k);
                    const double o = double(want) - double(base);
                    if (base + float(o) != want)
                        throw std::runtime_error("mesh.json: a vertex coordinate is not reproducible as lattice + offset");
                    off[3 * v + k] = o;
                    any = any || o != 0.0;
                }
            }
    LatticeMeshFile out;
    out.x = X; out.y = Y; out.z = Z;
    out.dx = dx; out.dy = dy; out.dz = dz;
    if (any) out.vertex_offsets = std::move(off);
    return out;
}

}  // namespace cfdarco::io
