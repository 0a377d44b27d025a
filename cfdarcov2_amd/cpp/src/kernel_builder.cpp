// KernelBuilder: one fused device kernel per equation, generated from the
// right-hand side tree and compiled by hipRTC through the library
// (cfd_program_compile).  Replaces the reference's string-fragment OCCA
// builder (src/fvm3d.cpp:1071-1394) with a code generator over expression
// nodes:
//   * pointwise nodes are emitted as C++ expressions in the reference
//     evaluation order (one parenthesised operation per node, no contraction);
//   * a gradient or Laplacian whose operand is pointwise is fused: the operand
//     is evaluated at the cell and its six neighbours (neighbour = self on a
//     boundary) and the stencil runs in registers, with the same arithmetic as
//     the library's operator kernels;
//   * stencils nested inside stencil operands, and the two-stage StabVar, are
//     evaluated operator by operator before the launch and read as buffers.
// The new state goes to a separate buffer that is swapped in after the launch
// (Jacobi within the equation), fixing the reference's in-place JIT update.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <sstream>
#include <stdexcept>

#include "cfdarcho_main_3d.hpp"
#include "fvm3d.hpp"

namespace cfdarco {
void check(int rc, const char* what);
}

using cfdarco::Node;
using cfdarco::NodeKind;

namespace {

// Geometry argument of the generated kernels; the device prelude declares the
// same layout.
struct JGeo {
    long long X, Y, Z;
    float area[3], alpha[3], alpha_b[3], len[3], vol, inv_vol;
    int uniform, vol_pow2;
    const float* fn[3][3];
    const float* farea[3];
    const float* falpha[3];
    const float* fvol;
};
static_assert(sizeof(JGeo) == 216, "JGeo layout");

// hipRTC declares the HIP device built-ins itself, so the prelude includes no
// runtime header.
const char* kPrelude = R"PRELUDE(
struct JGeo {
    long long X, Y, Z;
    float area[3], alpha[3], alpha_b[3], len[3], vol, inv_vol;
    int uniform, vol_pow2;
    const float* fn[3][3];
    const float* farea[3];
    const float* falpha[3];
    const float* fvol;
};
static_assert(sizeof(JGeo) == 216, "JGeo layout");
// linear face interpolation (phi_c + phi_nb) * 0.5
__device__ inline float jlin(float c, float n) { return (c + n) * 0.5f; }
// x / V of a uniform mesh: for a power-of-two V the product with its exact
// reciprocal, which rounds to the same value as the quotient
__device__ inline float jdivv(const JGeo& g, float x) { return g.vol_pow2 ? x * g.inv_vol : x / g.vol; }
// uniform-mesh Green-Gauss pair of one axis: ((-f_lo) * A + f_hi * A) / V
__device__ inline float jgg(const JGeo& g, float f_lo, float f_hi, float A) {
    const float t_lo = (-f_lo) * A;
    const float t_hi = f_hi * A;
    return jdivv(g, t_lo + t_hi);
}
// one axis of StabVar on a uniform mesh, added to acc: p at the cell (pc),
// its axis neighbours (pl, ph), and the neighbours' own neighbours (pll, plh
// of the lower one, phl, phh of the upper one; self on walls)
__device__ inline float jstab_axis(const JGeo& g, float acc, float pc, float pl, float ph, float pll, float plh,
                                   float phl, float phh, float A, float L) {
    const float gc = jgg(g, jlin(pc, pl), jlin(pc, ph), A);
    const float glo = jgg(g, jlin(pl, pll), jlin(pl, plh), A);
    const float ghi = jgg(g, jlin(ph, phl), jlin(ph, phh), A);
    const float own_lo = pc + (L * gc) * -1.0f;
    const float col_lo = pl + (L * glo) * 1.0f;
    const float own_hi = pc + (L * gc) * 1.0f;
    const float col_hi = ph + (L * ghi) * -1.0f;
    const float f_lo = 0.5f * (own_lo - col_lo);
    const float f_hi = 0.5f * (own_hi - col_hi);
    return acc + jdivv(g, f_lo * A + f_hi * A);
}
// v[0] = cell value, v[1 + fc] = value across local face fc (z-, z+, y-, y+, x-, x+)
__device__ inline bool jat(const JGeo& g, long long x, long long y, long long z, int fc) {
    switch (fc) {
        case 0: return z == 0;
        case 1: return z == g.Z - 1;
        case 2: return y == 0;
        case 3: return y == g.Y - 1;
        case 4: return x == 0;
        default: return x == g.X - 1;
    }
}
__device__ inline long long jface(const JGeo& g, long long x, long long y, long long z, int fc, int& fam) {
    switch (fc) {
        case 0: fam = 0; return x * g.Y * (g.Z + 1) + y * (g.Z + 1) + z;
        case 1: fam = 0; return x * g.Y * (g.Z + 1) + y * (g.Z + 1) + z + 1;
        case 2: fam = 1; return x * (g.Y + 1) * g.Z + y * g.Z + z;
        case 3: fam = 1; return x * (g.Y + 1) * g.Z + (y + 1) * g.Z + z;
        case 4: fam = 2; return x * g.Y * g.Z + y * g.Z + z;
        default: fam = 2; return (x + 1) * g.Y * g.Z + y * g.Z + z;
    }
}
// Green-Gauss sum over the six faces for component d (fc = 0..5 from 0),
// outward (axis_normals = false) or axis-direction normals
__device__ inline float jgauss(const JGeo& g, long long x, long long y, long long z, long long c, const float fv[6],
                               int d, bool axis_normals) {
    float acc = 0.0f;
    for (int fc = 0; fc < 6; ++fc) {
        int fam;
        const long long f = jface(g, x, y, z, fc, fam);
        float n = g.fn[fam][d][f];
        if (!axis_normals && !(fc & 1)) n = -n;
        float t = fv[fc] * n;
        t = t * g.farea[fam][f];
        acc = acc + t;
    }
    return acc / g.fvol[c];
}
// GradVar: uniform form sums the two faces of each selected axis
__device__ inline float jgrad(const JGeo& g, long long x, long long y, long long z, long long c, const float v[7],
                              int mask) {
    float res = 0.0f;
    if (g.uniform) {
        for (int a = 0; a < 3; ++a) {
            if (!(mask & (1 << a))) continue;
            const int lo = 1 + 2 * (2 - a), hi = lo + 1;
            const float f_lo = jlin(v[0], v[lo]), f_hi = jlin(v[0], v[hi]);
            const float t_lo = (-f_lo) * g.area[a];
            const float t_hi = f_hi * g.area[a];
            res = res + jdivv(g, t_lo + t_hi);
        }
        return res;
    }
    float fv[6];
    for (int fc = 0; fc < 6; ++fc) fv[fc] = jlin(v[0], v[1 + fc]);
    for (int d = 0; d < 3; ++d)
        if (mask & (1 << d)) res = res + jgauss(g, x, y, z, c, fv, d, false);
    return res;
}
// Grad2Var: (-phi_c + phi_nb) * alpha_d, Gauss sum with axis normals
__device__ inline float jlapl(const JGeo& g, long long x, long long y, long long z, long long c, const float v[7],
                              int mask) {
    float res = 0.0f;
    if (g.uniform) {
        for (int a = 0; a < 3; ++a) {
            if (!(mask & (1 << a))) continue;
            const int flo = 2 * (2 - a), fhi = flo + 1;
            const float al_lo = jat(g, x, y, z, flo) ? g.alpha_b[a] : g.alpha[a];
            const float al_hi = jat(g, x, y, z, fhi) ? g.alpha_b[a] : g.alpha[a];
            const float s_lo = (-v[0] + v[1 + flo]) * al_lo;
            const float s_hi = (-v[0] + v[1 + fhi]) * al_hi;
            res = res + jdivv(g, s_lo * g.area[a] + s_hi * g.area[a]);
        }
        return res;
    }
    float fv[6];
    for (int fc = 0; fc < 6; ++fc) {
        int fam;
        const long long f = jface(g, x, y, z, fc, fam);
        fv[fc] = (-v[0] + v[1 + fc]) * g.falpha[fam][f];
    }
    for (int d = 0; d < 3; ++d)
        if (mask & (1 << d)) res = res + jgauss(g, x, y, z, c, fv, d, true);
    return res;
}
)PRELUDE";

std::string literal(float v) {
    if (std::isnan(v)) return "(__builtin_nanf(\"\"))";
    if (std::isinf(v)) return v > 0 ? "(__builtin_inff())" : "(-__builtin_inff())";
    char buf[64];
    std::snprintf(buf, sizeof(buf), "(%af)", double(v));
    return buf;
}

}  // namespace

struct KernelBuilder::Impl {
    enum class Target { Assign, Euler, Leapfrog };

    Target target = Target::Assign;
    Variable* state = nullptr;          // variable whose state the equation updates
    DtVar* dtvar = nullptr;
    D2tVar* d2tvar = nullptr;
    const Node* root = nullptr;
    Mesh3D* mesh = nullptr;

    std::vector<const Node*> buffers;           // buffer inputs in slot order
    std::map<const void*, int> buffer_slot;     // Variable* for leaves, Node* otherwise
    std::vector<const Node*> scalars;           // DtScalar / Pointer nodes in slot order
    std::map<const void*, int> scalar_slot;
    std::vector<const Node*> materialized;      // evaluated operator by operator
    std::vector<const Node*> fused;             // stencils computed in registers
    std::map<const Node*, int> fused_id;
    std::vector<const Node*> fused_stab;        // StabVar computed in registers (uniform meshes)
    std::map<const Node*, int> stab_id;
    // StabVar whose pointwise operand is computed by one generated kernel
    // (operand_src / operand_name / operand_prog), then stabilised in one pass
    std::vector<const Node*> operand_stab;
    std::vector<std::string> operand_src, operand_name;
    std::vector<cfd_program_t> operand_prog;
    bool uniform_mesh = false;                  // exactly uniform mesh on the device
    bool stab_in_kernel = false;                // CFDARCO_STAB_IN_KERNEL=1: the 13-point in-kernel form
    std::string idx_t = "long long";            // cell index type of the generated code

    std::string source;
    std::string kname;
    cfd_program_t prog = nullptr;
    JGeo geo{};
    CudaDataMatrixD next;

    static const void* buffer_key(const Node& n) {
        return n.kind == NodeKind::Leaf ? static_cast<const void*>(n.leaf) : static_cast<const void*>(&n);
    }
    static const void* scalar_key(const Node& n) {
        return n.kind == NodeKind::DtScalar ? static_cast<const void*>(n.dt) : static_cast<const void*>(n.ptr);
    }
    void add_buffer(const Node& n) {
        const void* k = buffer_key(n);
        if (buffer_slot.count(k)) return;
        buffer_slot[k] = int(buffers.size());
        buffers.push_back(&n);
    }
    void add_scalar(const Node& n) {
        const void* k = scalar_key(n);
        if (scalar_slot.count(k)) return;
        scalar_slot[k] = int(scalars.size());
        scalars.push_back(&n);
    }
    void materialize(const Node& n) {
        if (buffer_slot.count(buffer_key(n))) return;
        materialized.push_back(&n);
        add_buffer(n);
    }

    void walk(const Node& n, bool in_stencil) {
        switch (n.kind) {
            case NodeKind::Leaf:
            case NodeKind::Array:
                add_buffer(n);
                return;
            case NodeKind::Const:
                return;
            case NodeKind::DtScalar:
            case NodeKind::Pointer:
                add_scalar(n);
                return;
            case NodeKind::Binary:
                walk(*n.a, in_stencil);
                walk(*n.b, in_stencil);
                return;
            case NodeKind::Neg:
            case NodeKind::Exp:
            case NodeKind::Abs:
                walk(*n.a, in_stencil);
                return;
            case NodeKind::Grad:
            case NodeKind::Grad2:
                if (in_stencil) {
                    materialize(n);
                } else if (!fused_id.count(&n)) {
                    fused_id[&n] = int(fused.size());
                    fused.push_back(&n);
                    walk(*n.a, true);
                }
                return;
            case NodeKind::Stab:
                // A stabilisation is evaluated before the launch and read as a
                // buffer.  A pointwise operand with two or more operators is
                // computed by one generated kernel of its own (one read of
                // each field, one write) and stabilised by the one-pass
                // kernel; other operands run operator by operator.
                //
                // CFDARCO_STAB_IN_KERNEL=1 selects the in-kernel form instead:
                // on a uniform mesh an operand of three or more fields is
                // evaluated at 13 points inside the generated kernel.  On the
                // Euler driver at 200^3, stab_tot(E) of five fields took 210 us
                // that way against about 250 us operator by operator (eight
                // elementwise launches, the one-pass stabilisation and the
                // fused update), while stab_tot(rho * u) took 130 us in-kernel
                // against about 100 us operator by operator.
                if (!stab_in_kernel && !has_stencil(*n.a) && count_ops(*n.a) >= 2) {
                    if (!buffer_slot.count(buffer_key(n))) {
                        operand_stab.push_back(&n);
                        walk(*n.a, true);
                        materialize(n);
                    }
                } else if (stab_in_kernel && uniform_mesh && !in_stencil && !has_stencil(*n.a) &&
                           distinct_leaves(*n.a) >= 3) {
                    if (!stab_id.count(&n)) {
                        stab_id[&n] = int(fused_stab.size());
                        fused_stab.push_back(&n);
                        walk(*n.a, true);
                    }
                } else {
                    materialize(n);
                }
                return;
        }
    }

    static size_t distinct_leaves(const Node& n) {
        std::vector<const void*> seen;
        std::function<void(const Node&)> visit = [&](const Node& m) {
            if (m.kind == NodeKind::Leaf || m.kind == NodeKind::Array) {
                const void* k = buffer_key(m);
                if (std::find(seen.begin(), seen.end(), k) == seen.end()) seen.push_back(k);
            }
            if (m.a) visit(*m.a);
            if (m.b) visit(*m.b);
        };
        visit(n);
        return seen.size();
    }

    static size_t count_ops(const Node& n) {
        const size_t own = (n.kind == NodeKind::Binary || n.kind == NodeKind::Neg || n.kind == NodeKind::Exp ||
                            n.kind == NodeKind::Abs)
                               ? 1
                               : 0;
        return own + (n.a ? count_ops(*n.a) : 0) + (n.b ? count_ops(*n.b) : 0);
    }

    static bool has_stencil(const Node& n) {
        if (n.kind == NodeKind::Grad || n.kind == NodeKind::Grad2 || n.kind == NodeKind::Stab) return true;
        return (n.a && has_stencil(*n.a)) || (n.b && has_stencil(*n.b));
    }

    std::string render(const Node& n, const std::string& idx) const {
        switch (n.kind) {
            case NodeKind::Leaf:
            case NodeKind::Array:
                return "in" + std::to_string(buffer_slot.at(buffer_key(n))) + "[" + idx + "]";
            case NodeKind::Const:
                return literal(n.value);
            case NodeKind::DtScalar:
            case NodeKind::Pointer:
                return "s" + std::to_string(scalar_slot.at(scalar_key(n)));
            case NodeKind::Binary:
                return "(" + render(*n.a, idx) + " " + n.op + " " + render(*n.b, idx) + ")";
            case NodeKind::Neg:
                return "(-" + render(*n.a, idx) + ")";
            case NodeKind::Exp:
                // the device libm's exp, as the operator-by-operator path
                return "expf(" + render(*n.a, idx) + ")";
            case NodeKind::Abs:
                return "fabsf(" + render(*n.a, idx) + ")";
            case NodeKind::Grad:
            case NodeKind::Grad2:
            case NodeKind::Stab: {
                auto it = buffer_slot.find(buffer_key(n));
                if (it != buffer_slot.end()) return "in" + std::to_string(it->second) + "[" + idx + "]";
                if (idx != "c") throw std::logic_error("fused stencil read away from its cell");
                if (n.kind == NodeKind::Stab) return "T" + std::to_string(stab_id.at(&n));
                return "S" + std::to_string(fused_id.at(&n));
            }
            default:
                throw std::runtime_error("Not implemented");
        }
    }

    void generate() {
        std::ostringstream k;
        k << kPrelude;
        std::ostringstream sig;
        sig << "JGeo g";
        for (size_t i = 0; i < buffers.size(); ++i) sig << ", const float* __restrict__ in" << i;
        for (size_t i = 0; i < scalars.size(); ++i) sig << ", float s" << i;
        const std::string sig_inputs = sig.str();
        sig << ", float* __restrict__ out_, const float* __restrict__ cur_, const float* __restrict__ prev_, float dt_, "
               "int stage_";
        // kernel name from the source body (stable across runs)
        std::ostringstream body;
        // the right-hand side and the update of cell c
        auto emit_update = [&](const std::string& ind) {
            body << ind << "const float rhs = " << render(*root, "c") << ";\n";
            switch (target) {
                case Target::Assign:
                    body << ind << "out_[c] = rhs;\n";
                    break;
                case Target::Euler:
                    // DtVar: dt * rhs + current
                    body << ind << "out_[c] = dt_ * rhs + cur_[c];\n";
                    break;
                case Target::Leapfrog:
                    // D2tVar: Euler for two steps, then dt * dt * rhs + 2 * current - previous
                    body << ind << "if (stage_) out_[c] = dt_ * dt_ * rhs + 2.0f * cur_[c] - prev_[c];\n"
                         << ind << "else out_[c] = dt_ * rhs + cur_[c];\n";
                    break;
            }
        };
        auto emit_stencil = [&](size_t i, const std::string& ind, const std::string& xm, const std::string& xp) {
            const Node& n = *fused[i];
            body << ind << "float v" << i << "[7];\n";
            body << ind << "v" << i << "[0] = " << render(*n.a, "c") << ";\n";
            for (int fc = 0; fc < 4; ++fc)
                body << ind << "v" << i << "[" << fc + 1 << "] = " << render(*n.a, "nb" + std::to_string(fc)) << ";\n";
            body << ind << "v" << i << "[5] = " << xm << ";\n" << ind << "v" << i << "[6] = " << xp << ";\n";
            body << ind << "const float S" << i << " = " << (n.kind == NodeKind::Grad ? "jgrad" : "jlapl")
                 << "(g, x, y, z, c, v" << i << ", " << n.mask << ");\n";
        };
        // StabVar of a pointwise operand on a uniform mesh: per selected
        // axis, the operand at the cell, its two neighbours and theirs (self
        // on walls; each point evaluated only when it is not one of the
        // others), then the arithmetic of stab_axis_u in csrc/stencil.hip
        auto emit_stab = [&](size_t i) {
            const Node& n = *fused_stab[i];
            const std::string t = "T" + std::to_string(i), pc = "P" + std::to_string(i);
            body << "    float " << t << " = 0.0f;\n"
                 << "    const float " << pc << " = " << render(*n.a, "c") << ";\n";
            static const char* stride[3] = {"g.Y * g.Z", "g.Z", "1"};
            static const char* coord[3] = {"x", "y", "z"};
            static const char* extent[3] = {"g.X", "g.Y", "g.Z"};
            for (int a = 0; a < 3; ++a) {
                if (!(n.mask & (1 << a))) continue;
                const std::string sa = stride[a], ca = coord[a], na = extent[a];
                body << "    {\n"
                     << "        const " << idx_t << " sa = " << sa << ";\n"
                     << "        const bool wl = " << ca << " == 0, wh = " << ca << " == " << na << " - 1;\n"
                     << "        const " << idx_t << " il = wl ? c : c - sa;\n"
                     << "        const " << idx_t << " ih = wh ? c : c + sa;\n"
                     << "        const float pl = wl ? " << pc << " : " << render(*n.a, "il") << ";\n"
                     << "        const float ph = wh ? " << pc << " : " << render(*n.a, "ih") << ";\n"
                     << "        const float pll = wl ? " << pc << " : (" << ca << " == 1 ? pl : "
                     << render(*n.a, "(il - sa)") << ");\n"
                     << "        const float plh = wl ? ph : " << pc << ";\n"
                     << "        const float phh = wh ? " << pc << " : (" << ca << " == " << na << " - 2 ? ph : "
                     << render(*n.a, "(ih + sa)") << ");\n"
                     << "        const float phl = wh ? pl : " << pc << ";\n"
                     << "        " << t << " = jstab_axis(g, " << t << ", " << pc
                     << ", pl, ph, pll, plh, phl, phh, g.area[" << a << "], g.len[" << a << "]);\n"
                     << "    }\n";
            }
        };
        // one thread per cell
        // cell indices in 32 bits when the mesh allows (fewer VALU
        // instructions for the address arithmetic), else 64
        const std::string I = idx_t;
        body << "    const " << I << " Xg = " << I << "(g.X), Yg = " << I << "(g.Y), Zg = " << I << "(g.Z);\n"
             << "    const " << I << " z = " << I << "(blockIdx.x) * " << I << "(blockDim.x) + " << I
             << "(threadIdx.x);\n"
             << "    const " << I << " y = " << I << "(blockIdx.y) * " << I << "(blockDim.y) + " << I
             << "(threadIdx.y);\n"
             << "    const " << I << " x = " << I << "(blockIdx.z);\n"
             << "    if (z >= Zg || y >= Yg || x >= Xg) return;\n"
             << "    const " << I << " c = (x * Yg + y) * Zg + z;\n"
             << "    const " << I << " nb0 = z > 0 ? c - 1 : c;\n"
             << "    const " << I << " nb1 = z < Zg - 1 ? c + 1 : c;\n"
             << "    const " << I << " nb2 = y > 0 ? c - Zg : c;\n"
             << "    const " << I << " nb3 = y < Yg - 1 ? c + Zg : c;\n"
             << "    const " << I << " nb4 = x > 0 ? c - Yg * Zg : c;\n"
             << "    const " << I << " nb5 = x < Xg - 1 ? c + Yg * Zg : c;\n";
        for (size_t i = 0; i < fused.size(); ++i)
            emit_stencil(i, "    ", render(*fused[i]->a, "nb4"), render(*fused[i]->a, "nb5"));
        for (size_t i = 0; i < fused_stab.size(); ++i) emit_stab(i);
        emit_update("    ");
        const std::string b = body.str();
        uint64_t h = 1469598103934665603ull;
        for (unsigned char ch : sig.str() + b) {
            h ^= ch;
            h *= 1099511628211ull;
        }
        char nm[64];
        std::snprintf(nm, sizeof(nm), "cfd_fused_%016llx", (unsigned long long)h);
        kname = nm;
        k << "extern \"C\" __global__ void __launch_bounds__(256) " << kname << "(" << sig.str() << ") {\n"
          << b << "}\n";
        source = k.str();
        // operand kernels: the main kernel's inputs (a stabilised buffer is
        // not ready yet and is not read), one output, one thread per cell
        for (const Node* st : operand_stab) {
            const std::string ob = "    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;\n"
                                   "    if (c >= n_) return;\n"
                                   "    res_[c] = " + render(*st->a, "c") + ";\n";
            const std::string os = sig_inputs + ", float* __restrict__ res_, long long n_";
            uint64_t ho = 1469598103934665603ull;
            for (unsigned char ch : os + ob) {
                ho ^= ch;
                ho *= 1099511628211ull;
            }
            char on[64];
            std::snprintf(on, sizeof(on), "cfd_operand_%016llx", (unsigned long long)ho);
            std::ostringstream ok;
            ok << kPrelude << "extern \"C\" __global__ void __launch_bounds__(256) " << on << "(" << os << ") {\n"
               << ob << "}\n";
            operand_name.push_back(on);
            operand_src.push_back(ok.str());
        }
    }
};

KernelBuilder::KernelBuilder(Variable* left_val_, Variable* expression_)
    : expression{expression_}, left_val{left_val_}, impl_{std::make_unique<Impl>()} {}
KernelBuilder::KernelBuilder(KernelBuilder&&) noexcept = default;
KernelBuilder& KernelBuilder::operator=(KernelBuilder&&) noexcept = default;
KernelBuilder::~KernelBuilder() = default;

const std::string& KernelBuilder::source() const { return impl_->source; }
const std::string& KernelBuilder::kernel_name() const { return impl_->kname; }
const std::vector<std::string>& KernelBuilder::operand_sources() const { return impl_->operand_src; }
const std::vector<std::string>& KernelBuilder::operand_kernel_names() const { return impl_->operand_name; }

void KernelBuilder::generate() {
    Impl& I = *impl_;
    if (!I.source.empty()) return;
    if (auto* d = dynamic_cast<DtVar*>(left_val)) {
        I.target = Impl::Target::Euler;
        I.dtvar = d;
        I.state = d->var.get();
    } else if (auto* d2 = dynamic_cast<D2tVar*>(left_val)) {
        I.target = Impl::Target::Leapfrog;
        I.d2tvar = d2;
        I.state = d2->var.get();
    } else {
        I.target = Impl::Target::Assign;
        I.state = left_val;
    }
    if (I.state == nullptr || I.state->node->kind != NodeKind::Leaf)
        throw std::runtime_error("equation target must be a state Variable");
    I.root = expression->node.get();
    I.mesh = I.state->mesh;
    // stabilisation runs in the kernel only on an exactly uniform mesh, known
    // once the mesh is on the device (generation without a device treats the
    // mesh as general)
    if (auto* cm = dynamic_cast<CudaMesh3D*>(I.mesh); cm != nullptr && cm->device_mesh() != nullptr) {
        cfd_geometry cg{};
        if (cfd_mesh_geometry(cm->device_mesh(), &cg) == CFD_OK) I.uniform_mesh = cg.uniform != 0;
    }
    // 32-bit cell indices when every index the kernel forms (up to two planes
    // past a cell, and the face numbers of the general form) stays below 2^31
    if (I.mesh != nullptr && uint64_t(I.mesh->_num_nodes) < (uint64_t(1) << 28)) I.idx_t = "int";
    const char* sik = std::getenv("CFDARCO_STAB_IN_KERNEL");
    I.stab_in_kernel = sik != nullptr && std::strcmp(sik, "1") == 0;
    I.walk(*I.root, false);
    I.generate();
}

void KernelBuilder::build(bool announce) {
    generate();
    Impl& I = *impl_;
    auto* cm = dynamic_cast<CudaMesh3D*>(I.mesh);
    if (cm == nullptr || cm->device_mesh() == nullptr)
        throw std::runtime_error("KernelBuilder needs a CudaMesh3D on the device");
    cfd_geometry cg{};
    cfdarco::check(cfd_mesh_geometry(cm->device_mesh(), &cg), "cfd_mesh_geometry");
    JGeo& g = I.geo;
    g.X = cg.nx;
    g.Y = cg.ny;
    g.Z = cg.nz;
    for (int a = 0; a < 3; ++a) {
        g.area[a] = float(cg.area[a]);
        g.alpha[a] = float(cg.alpha[a]);
        g.alpha_b[a] = float(cg.alpha_b[a]);
    }
    g.vol = float(cg.vol);
    {
        int e = 0;
        const float m = std::frexp(g.vol, &e);
        const float inv = 1.0f / g.vol;
        // a power of two whose reciprocal is a normal float
        g.vol_pow2 = (m == 0.5f && std::isnormal(inv) && std::isnormal(g.vol)) ? 1 : 0;
        g.inv_vol = g.vol_pow2 ? inv : 0.0f;
    }
    for (int a = 0; a < 3; ++a) g.len[a] = float(cg.len[a]);
    g.uniform = cg.uniform;
    for (int f = 0; f < 3; ++f) {
        for (int d = 0; d < 3; ++d) g.fn[f][d] = static_cast<const float*>(cg.face_normal[f][d]);
        g.farea[f] = static_cast<const float*>(cg.face_area[f]);
        g.falpha[f] = static_cast<const float*>(cg.face_alpha[f]);
    }
    g.fvol = static_cast<const float*>(cg.cell_vol);
    if (std::getenv("CFDARCO_JIT_PRINT")) {
        std::cout << "Building kernel:\n" << I.source << std::endl;
        for (const std::string& src : I.operand_src) std::cout << "Building kernel:\n" << src << std::endl;
    }
    cfdarco::check(cfd_program_compile(CFDArcoGlobalInit::context(), I.source.c_str(), I.kname.c_str(), &I.prog),
                   "KernelBuilder::build");
    I.operand_prog.assign(I.operand_src.size(), nullptr);
    for (size_t j = 0; j < I.operand_src.size(); ++j)
        cfdarco::check(cfd_program_compile(CFDArcoGlobalInit::context(), I.operand_src[j].c_str(),
                                           I.operand_name[j].c_str(), &I.operand_prog[j]),
                       "KernelBuilder::build (stabilised operand)");
    if (announce) std::cout << "Kernel build" << std::endl;
}

void KernelBuilder::run() {
    Impl& I = *impl_;
    if (I.prog == nullptr) throw std::runtime_error("KernelBuilder::run before build");
    Variable* st = I.state;
    const size_t n = st->current_cu._size;
    std::vector<const float*> bufs(I.buffers.size(), nullptr);
    for (size_t i = 0; i < I.buffers.size(); ++i) {
        const Node& b = *I.buffers[i];
        if (b.kind == NodeKind::Leaf) bufs[i] = b.leaf->current_cu.data.get();
        else if (b.kind == NodeKind::Array) bufs[i] = b.array.data.get();
    }
    std::vector<float> scal(I.scalars.size());
    for (size_t i = 0; i < I.scalars.size(); ++i) {
        const Node& s = *I.scalars[i];
        scal[i] = s.kind == NodeKind::DtScalar ? s.dt->_dt : *s.ptr;
    }
    // sub-expressions evaluated before the launch: operator by operator, or a
    // stabilised operand from its generated kernel
    std::vector<CudaDataMatrixD> mats;
    mats.reserve(I.materialized.size());
    for (const Node* m : I.materialized) {
        const auto it = std::find(I.operand_stab.begin(), I.operand_stab.end(), m);
        if (it == I.operand_stab.end()) {
            mats.push_back(cfdarco::evaluate_node(*m));
        } else {
            const size_t j = size_t(it - I.operand_stab.begin());
            auto* cm = dynamic_cast<CudaMesh3D*>(I.mesh);
            if (cm == nullptr || j >= I.operand_prog.size() || I.operand_prog[j] == nullptr)
                throw std::runtime_error("KernelBuilder::run: stabilised operand kernel not built");
            CudaDataMatrixD operand{n};
            float* res = operand.data.get();
            long long cells = static_cast<long long>(n);
            std::vector<void*> oargs;
            oargs.reserve(3 + bufs.size() + scal.size());
            oargs.push_back(&I.geo);
            for (auto& p : bufs) oargs.push_back(&p);
            for (auto& s : scal) oargs.push_back(&s);
            oargs.push_back(&res);
            oargs.push_back(&cells);
            const uint32_t ogrid[3] = {uint32_t((n + 255) / 256), 1, 1};
            const uint32_t oblock[3] = {256, 1, 1};
            cfdarco::check(cfd_program_launch(I.operand_prog[j], ogrid, oblock, oargs.data(), nullptr),
                           "KernelBuilder::run (stabilised operand)");
            mats.push_back(eval_stab(cm, operand, m->mask & 1, m->mask & 2, m->mask & 4));
        }
        bufs[size_t(I.buffer_slot.at(Impl::buffer_key(*m)))] = mats.back().data.get();
    }
    if (I.next._size != n || I.next.data.use_count() > 1) I.next = CudaDataMatrixD{n};
    float* out = I.next.data.get();
    const float* cur = st->current_cu.data.get();
    const float* prev = cur;
    float dtv = 0.0f;
    int stage = 0;
    if (I.target == Impl::Target::Euler) {
        dtv = I.dtvar->dt_;
    } else if (I.target == Impl::Target::Leapfrog) {
        dtv = I.d2tvar->dt_;
        stage = int(I.d2tvar->second_stage);
        prev = I.d2tvar->pre_current_cu.data.get();
    }
    std::vector<void*> args;
    args.reserve(4 + bufs.size() + scal.size() + 5);
    args.push_back(&I.geo);
    for (auto& p : bufs) args.push_back(&p);
    for (auto& s : scal) args.push_back(&s);
    args.push_back(&out);
    args.push_back(&cur);
    args.push_back(&prev);
    args.push_back(&dtv);
    args.push_back(&stage);
    // workgroups are dealt round-robin to the 8 XCDs; with a multiple of 8
    // per x-plane the same (y, z) tile of every plane runs on one XCD, whose
    // L2 then already holds the x-neighbours (rows past Y exit at once)
    // (padding is skipped where it would push gridDim.y past 65535)
    const uint32_t gz = uint32_t((I.geo.Z + 63) / 64);
    uint32_t gy = uint32_t((I.geo.Y + 3) / 4);
    uint32_t gyp = gy;
    while ((uint64_t(gz) * gyp) % 8 != 0) ++gyp;
    if (gyp <= 65535u) gy = gyp;
    const uint32_t grid[3] = {gz, gy, uint32_t(I.geo.X)};
    const uint32_t block[3] = {64, 4, 1};
    cfdarco::check(cfd_program_launch(I.prog, grid, block, args.data(), nullptr), "KernelBuilder::run");
    // new state in, old state becomes the next output (or the previous state of
    // d2t); the handle is shared, and a buffer someone else still holds is not
    // reused as an output (use count above)
    CudaDataMatrixD old;
    old = st->current_cu;
    st->current_cu = I.next;
    if (I.target == Impl::Target::Leapfrog) {
        I.next = I.d2tvar->pre_current_cu;
        I.d2tvar->pre_current_cu = old;
    } else {
        I.next = old;
    }
}
