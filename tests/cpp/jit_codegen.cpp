// Emits the fused-kernel sources KernelBuilder generates for the equation
// systems of fvm3d_parity.cpp, on the host only (no device), one file per
// equation: <outdir>/<kernel_name>.hip, and one file per kernel that computes
// a stabilised operand.  tests/test_jit_codegen.py compiles them with hipRTC.
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "fvm3d.hpp"
#include "mesh3d.hpp"

namespace {
Eigen::Matrix<float, -1, 1> keep(Mesh3D*, Eigen::Matrix<float, -1, 1>& arr, const DT*) { return arr; }

int emit(const std::string& dir, EquationTemplate& eqs) {
    int n = 0;
    for (auto& eq : eqs) {
        KernelBuilder kb(std::get<0>(eq), &std::get<2>(eq));
        kb.generate();
        std::ofstream f(dir + "/" + kb.kernel_name() + ".hip");
        f << kb.source();
        std::cout << kb.kernel_name() << "\n";
        for (size_t j = 0; j < kb.operand_sources().size(); ++j) {
            std::ofstream o(dir + "/" + kb.operand_kernel_names()[j] + ".hip");
            o << kb.operand_sources()[j];
            std::cout << kb.operand_kernel_names()[j] << "\n";
        }
        ++n;
    }
    return n;
}
}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: jit_codegen <outdir>\n";
        return 2;
    }
    const std::string dir = argv[1];
    Mesh3D mesh(6, 5, 4, 0.25f, 0.25f, 0.25f);
    mesh.init_basic_internals();
    mesh.compute();
    Eigen::Matrix<float, -1, 1> one = Eigen::Matrix<float, -1, 1>::Constant(Eigen::Index(mesh._num_nodes), 1.0f);

    auto T = Variable(&mesh, one, keep, "T");
    auto p = Variable(&mesh, one, keep, "p");
    auto u = Variable(&mesh, one, keep, "u");
    auto v = Variable(&mesh, one, keep, "v");
    auto w = Variable(&mesh, one, keep, "w");
    auto rho = Variable(&mesh, one, keep, "rho");
    auto mass = Variable(&mesh, one, keep, "mass");
    auto mass_tmp = Variable(&mesh, one, keep, "mass_tmp");
    auto volumes_var = Variable(&mesh, mesh._volumes, keep, "volumes_var");
    std::vector<Variable*> sv{&u, &p, &rho};
    auto dt = DT(&mesh, UpdatePolicies::constant_dt, UpdatePolicies::constant_dt_cu, 0.1f, sv);
    const float gamma = 1.4f, c = 1.25f;
    auto E = p / (gamma - 1) + 0.5 * rho * ((u * u));
    float scalar = 2.5f;
    PointerVariable pv(&mesh, &scalar);

    EquationTemplate eqs = {
        {d1t(T), '=', 3.0 * lapl(T), true},
        {d1t(u), '=', -d1dx(p), true},
        {d1t(p), '=', (-c * c) * (d1dx(u) + d1dy(v) + d1dz(w)), true},
        {d2t(u), '=', c * c * (d2dx(u) + d2dy(u)), true},
        {&rho, '=', mass / volumes_var, true},
        {&p, '=', (p / volumes_var - 0.5 * rho * (u * u)) * (gamma - 1), true},
        {&rho, '=', rho - 0.5 * dt * (u * d1dx(rho) + rho * d1dx(u)), true},
        {&mass_tmp, '=', 0 * mass_tmp - (d1dx((E + p) * u)), true},
        {d1t(mass), '=', mass_tmp - stabx(rho * u) * 2.0f, true},
        {d1t(mass), '=', mass_tmp - stab_tot(E) * 2.0f, true},
        {&mass_tmp, '=', lapl(d1dx(rho)) + d2dy(mass) + pv * rho, true},
        {&T, '=', -(-T) / (1.0f / u) - 0.0f * T, true},
    };
    return emit(dir, eqs) == int(eqs.size()) ? 0 : 1;
}
