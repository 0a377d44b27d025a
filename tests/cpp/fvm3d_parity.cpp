// Parity program for the C++ operator API: runs an equation system through
// Variable / Equation::evaluate on the device and writes the final fields, for
// comparison with the CPU oracle composed in Python (tests/test_gpu_fvm3d.py).
//
//   fvm3d_parity <case> <input.bin> <output.bin>
//
// input.bin : int64 nx, ny, nz, steps; float64 dx, dy, dz, dt; int64 nfields;
//             nfields x N float32 initial fields
// output.bin: the case's fields after `steps` steps, N float32 each
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cfdarcho_main_3d.hpp"
#include "fvm3d.hpp"
#include "mesh3d.hpp"

namespace {

struct Input {
    int64_t nx, ny, nz, steps;
    double dx, dy, dz, dt;
    std::vector<Eigen::Matrix<float, -1, 1>> fields;
};

Input read_input(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    Input in{};
    int64_t hdr[4];
    double sp[4];
    int64_t nf = 0;
    f.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    f.read(reinterpret_cast<char*>(sp), sizeof(sp));
    f.read(reinterpret_cast<char*>(&nf), sizeof(nf));
    in.nx = hdr[0]; in.ny = hdr[1]; in.nz = hdr[2]; in.steps = hdr[3];
    in.dx = sp[0]; in.dy = sp[1]; in.dz = sp[2]; in.dt = sp[3];
    const size_t N = size_t(in.nx * in.ny * in.nz);
    for (int64_t k = 0; k < nf; ++k) {
        Eigen::Matrix<float, -1, 1> v{N};
        f.read(reinterpret_cast<char*>(v.data()), std::streamsize(N * sizeof(float)));
        in.fields.push_back(v);
    }
    if (!f) throw std::runtime_error("short input file");
    return in;
}

void write_output(const std::string& path, const std::vector<Variable*>& vars) {
    std::ofstream f(path, std::ios::binary);
    for (auto* v : vars) {
        auto h = v->current_cu.to_eigen();
        f.write(reinterpret_cast<const char*>(h.data()), std::streamsize(h.size() * sizeof(float)));
    }
}

Eigen::Matrix<float, -1, 1> keep(Mesh3D*, Eigen::Matrix<float, -1, 1>& arr, const DT*) { return arr; }
CudaDataMatrixD keep_cu(Mesh3D*, CudaDataMatrixD& arr, const DT*) { return arr; }

// Dirichlet hold of the boundary layer, as examples/main_3d_heat.cpp:41-60
// (device callback built from the operator algebra)
CudaDataMatrixD hold_cu(Mesh3D* m, CudaDataMatrixD& arr, const CudaDataMatrixD& fixed) {
    auto* mesh = dynamic_cast<CudaMesh3D*>(m);
    auto a1 = arr * mesh->_node_is_boundary_reverse_cu;
    auto b1 = fixed * mesh->_node_is_boundary_cu;
    return a1 + b1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::cerr << "usage: fvm3d_parity <case> <input.bin> <output.bin>\n";
        return 2;
    }
    const std::string kase = argv[1];
    Input in = read_input(argv[2]);
    CFDArcoGlobalInit::initialize(argc, argv, kase.find("store") == std::string::npos);
    auto mesh = std::make_shared<CudaMesh3D>(size_t(in.nx), size_t(in.ny), size_t(in.nz), float(in.dx), float(in.dy),
                                             float(in.dz));
    mesh->init_basic_internals();
    if (kase.find("perturbed") != std::string::npos) mesh->perturb_interior_vertices(5, 0.02);
    mesh->compute();
    CFDArcoGlobalInit::enable_cuda(1, true, false);
    mesh->host_to_device();
    const float dtv = float(in.dt);
    const bool jit = kase.find("nojit") == std::string::npos;
    auto& F = in.fields;

    if (kase.rfind("heat", 0) == 0) {
        // examples/main_3d_heat.cpp: d1t(T) = 3 * lapl(T), boundary layer held
        Eigen::Matrix<float, -1, 1> T0 = F.at(0);
        CudaDataMatrixD fixed = CudaDataMatrixD::from_eigen(T0);
        auto T = Variable(mesh.get(), F.at(0), keep,
                          [fixed](Mesh3D* m, CudaDataMatrixD& arr, const DT*) { return hold_cu(m, arr, fixed); }, "T");
        std::vector<Variable*> vars{&T};
        auto dt = DT(mesh.get(), UpdatePolicies::constant_dt, UpdatePolicies::constant_dt_cu, dtv, vars);
        EquationTemplate eqs = {{d1t(T), '=', 3.0 * lapl(T), jit}};
        Equation(size_t(in.steps)).evaluate(vars, eqs, &dt, false, vars);
        write_output(argv[3], vars);
    } else if (kase.rfind("wave", 0) == 0) {
        // first-order acoustic system through the operator API (reference
        // boundary rule: neighbour = self on the walls)
        auto p = Variable(mesh.get(), F.at(0), keep, keep_cu, "p");
        auto u = Variable(mesh.get(), F.at(1), keep, keep_cu, "u");
        auto v = Variable(mesh.get(), F.at(2), keep, keep_cu, "v");
        auto w = Variable(mesh.get(), F.at(3), keep, keep_cu, "w");
        std::vector<Variable*> vars{&p, &u, &v, &w};
        auto dt = DT(mesh.get(), UpdatePolicies::constant_dt, UpdatePolicies::constant_dt_cu, dtv, vars);
        const float c = 1.25f;
        EquationTemplate eqs = {
            {d1t(u), '=', -d1dx(p), jit},
            {d1t(v), '=', -d1dy(p), jit},
            {d1t(w), '=', -d1dz(p), jit},
            {d1t(p), '=', (-c * c) * (d1dx(u) + d1dy(v) + d1dz(w)), jit},
        };
        Equation(size_t(in.steps)).evaluate(vars, eqs, &dt, false, vars);
        write_output(argv[3], vars);
    } else if (kase.rfind("expabs", 0) == 0) {
        // exp and abs on the device (the reference's exp/abs, src/fvm3d.cpp:
        // 472-494, run on the host only), inside fused stencil expressions
        // and as assignment operands
        auto x = Variable(mesh.get(), F.at(0), keep, keep_cu, "x");
        auto y = Variable(mesh.get(), F.at(1), keep, keep_cu, "y");
        auto z = Variable(mesh.get(), F.at(2), keep, keep_cu, "z");
        std::vector<Variable*> vars{&x, &y, &z};
        auto dt = DT(mesh.get(), UpdatePolicies::constant_dt, UpdatePolicies::constant_dt_cu, dtv, vars);
        EquationTemplate eqs = {
            {&y, '=', exp(-abs(x) * 0.5f) * y, jit},
            {&z, '=', abs(d1dx(x)) - abs(x) * z, jit},
            {d1t(x), '=', -abs(d1dy(x)), jit},
        };
        Equation(size_t(in.steps)).evaluate(vars, eqs, &dt, false, vars);
        write_output(argv[3], vars);
    } else if (kase.rfind("d2t", 0) == 0) {
        // examples/main_3d_wave.cu: d2t(u) = c*c*(d2dx(u) + d2dy(u)); host
        // boundary callback (host round trip every step)
        auto u = Variable(mesh.get(), F.at(0), keep, "u");
        std::vector<Variable*> vars{&u};
        auto dt = DT(mesh.get(), UpdatePolicies::constant_dt, UpdatePolicies::constant_dt_cu, dtv, vars);
        const float c = 0.3f;
        EquationTemplate eqs = {{d2t(u), '=', c * c * (d2dx(u) + d2dy(u)), jit}};
        Equation(size_t(in.steps)).evaluate(vars, eqs, &dt, false, vars);
        write_output(argv[3], vars);
    } else if (kase.rfind("euler1d", 0) == 0) {
        // the structure of examples/main_3d_sod_shock.cpp: assignment
        // equations, predictor with DT in the expression, nested products in
        // stencil operands, stabilisation, CFL time step (1D policy on device)
        auto rho = Variable(mesh.get(), F.at(0), keep, keep_cu, "rho");
        auto u = Variable(mesh.get(), F.at(1), keep, keep_cu, "u");
        auto p = Variable(mesh.get(), F.at(2), keep, keep_cu, "p");
        auto mass = Variable(mesh.get(), F.at(3), keep, keep_cu, "mass");
        auto rho_u = Variable(mesh.get(), F.at(4), keep, keep_cu, "rho_u");
        auto rho_e = Variable(mesh.get(), F.at(5), keep, keep_cu, "rho_e");
        auto mass_tmp = Variable(mesh.get(), F.at(3), keep, keep_cu, "mass_tmp");
        auto rho_u_tmp = Variable(mesh.get(), F.at(4), keep, keep_cu, "rho_u_tmp");
        auto rho_e_tmp = Variable(mesh.get(), F.at(5), keep, keep_cu, "rho_e_tmp");
        auto volumes_var = Variable(mesh.get(), mesh->_volumes, keep, keep_cu, "volumes_var");
        const float gamma = 1.4f;
        auto E = p / (gamma - 1) + 0.5 * rho * ((u * u));
        std::vector<Variable*> space{&u, &p, &rho};
        auto dt = DT(mesh.get(), UpdatePolicies::CourantFriedrichsLewy1D, UpdatePolicies::CourantFriedrichsLewy1DCu,
                     dtv, space);
        const float dissip = 2;
        EquationTemplate eqs = {
            {&rho, '=', mass / volumes_var, jit},
            {&u, '=', rho_u / rho / volumes_var, jit},
            {&p, '=', (rho_e / volumes_var - 0.5 * rho * (u * u)) * (gamma - 1), jit},
            {&rho, '=', rho - 0.5 * dt * (u * d1dx(rho) + rho * d1dx(u)), jit},
            {&u, '=', u - 0.5 * dt * (u * d1dx(u) + (1 / rho) * d1dx(p)), jit},
            {&p, '=', p - 0.5 * dt * (gamma * p * (d1dx(u)) + u * d1dx(p)), jit},
            {&mass_tmp, '=', 0 * mass_tmp - (d1dx(rho * u)), jit},
            {&rho_u_tmp, '=', 0 * rho_u_tmp - (d1dx(rho * u * u + p)), jit},
            {&rho_e_tmp, '=', 0 * rho_e_tmp - (d1dx((E + p) * u)), jit},
            {d1t(mass), '=', mass_tmp - stabx(rho) * dissip, jit},
            {d1t(rho_u), '=', rho_u_tmp - stabx(rho * u) * dissip, jit},
            {d1t(rho_e), '=', rho_e_tmp - stab_tot(E) * dissip, jit},
            {&mass_tmp, '=', lapl(d1dx(rho)) + d2dy(mass), jit},
        };
        std::vector<Variable*> all{&rho, &u, &p, &mass, &rho_u, &rho_e};
        Equation(size_t(in.steps)).evaluate(all, eqs, &dt, false, all);
        std::vector<Variable*> out{&rho, &u, &p, &mass, &rho_u, &rho_e, &mass_tmp};
        write_output(argv[3], out);
    } else {
        std::cerr << "unknown case " << kase << "\n";
        return 2;
    }
    CFDArcoGlobalInit::finalize();
    return 0;
}
