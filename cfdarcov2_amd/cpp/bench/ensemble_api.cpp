// Sequential plain solves of an ensemble through the operator API: every case
// is set up and advanced on its own with Variable / Equation::evaluate (fused
// hipRTC kernels, the way a reference driver solves one case), and the time
// of all solves together is reported.  Used by bench.py as the operator-API
// baseline of the batched ensemble.
//
//   ensemble_api <cases.bin>
//
// cases.bin: int64 ncases, steps; per case int64 kind (0 heat, 1 acoustic), n;
// float64 spacing, coef (alpha, or c^2 for the acoustic system), dt; then n^3
// float32 initial values (T, or p with u = v = w = 0).
#include <chrono>
#include <cstdint>
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

struct CaseSpec {
    int64_t kind = 0, n = 0;
    double h = 0, coef = 0, dt = 0;
    Eigen::Matrix<float, -1, 1> f0;
};

Eigen::Matrix<float, -1, 1> keep(Mesh3D*, Eigen::Matrix<float, -1, 1>& arr, const DT*) { return arr; }
CudaDataMatrixD keep_cu(Mesh3D*, CudaDataMatrixD& arr, const DT*) { return arr; }

struct Solver {
    float coef = 0.0f;  // read through a PointerVariable, so every case shares one compiled kernel
    std::unique_ptr<PointerVariable> coef_var;
    std::shared_ptr<CudaMesh3D> mesh;
    std::vector<std::unique_ptr<Variable>> vars;
    std::unique_ptr<DT> dt;
    EquationTemplate eqs;
    std::vector<Variable*> all;
};

std::unique_ptr<Solver> make_solver(CaseSpec& c) {
    auto s = std::make_unique<Solver>();
    s->mesh = std::make_shared<CudaMesh3D>(size_t(c.n), size_t(c.n), size_t(c.n), float(c.h), float(c.h), float(c.h));
    s->mesh->init_basic_internals();
    s->mesh->compute();
    s->mesh->host_to_device();
    Mesh3D* m = s->mesh.get();
    s->coef = c.kind == 0 ? float(c.coef) : -float(c.coef);
    s->coef_var = std::make_unique<PointerVariable>(m, &s->coef);
    if (c.kind == 0) {
        auto T = std::make_unique<Variable>(m, c.f0, keep, keep_cu, "T");
        s->all = {T.get()};
        s->dt = std::make_unique<DT>(m, UpdatePolicies::constant_dt, UpdatePolicies::constant_dt_cu, float(c.dt),
                                     s->all);
        s->eqs = {{d1t(*T), '=', *s->coef_var * lapl(*T), true}};
        s->vars.push_back(std::move(T));
    } else {
        Eigen::Matrix<float, -1, 1> zero = Eigen::Matrix<float, -1, 1>::Constant(Eigen::Index(m->_num_nodes), 0.0f);
        auto p = std::make_unique<Variable>(m, c.f0, keep, keep_cu, "p");
        auto u = std::make_unique<Variable>(m, zero, keep, keep_cu, "u");
        auto v = std::make_unique<Variable>(m, zero, keep, keep_cu, "v");
        auto w = std::make_unique<Variable>(m, zero, keep, keep_cu, "w");
        s->all = {p.get(), u.get(), v.get(), w.get()};
        s->dt = std::make_unique<DT>(m, UpdatePolicies::constant_dt, UpdatePolicies::constant_dt_cu, float(c.dt),
                                     s->all);
        s->eqs = {
            {d1t(*u), '=', -d1dx(*p), true},
            {d1t(*v), '=', -d1dy(*p), true},
            {d1t(*w), '=', -d1dz(*p), true},
            {d1t(*p), '=', *s->coef_var * (d1dx(*u) + d1dy(*v) + d1dz(*w)), true},
        };
        s->vars.push_back(std::move(p));
        s->vars.push_back(std::move(u));
        s->vars.push_back(std::move(v));
        s->vars.push_back(std::move(w));
    }
    return s;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: ensemble_api <cases.bin>\n";
        return 2;
    }
    std::ifstream f(argv[1], std::ios::binary);
    if (!f) throw std::runtime_error("cannot open case file");
    int64_t hdr[2];
    f.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    const int64_t ncases = hdr[0], steps = hdr[1];
    std::vector<CaseSpec> cases(static_cast<size_t>(ncases));
    for (auto& c : cases) {
        int64_t ki[2];
        double d[3];
        f.read(reinterpret_cast<char*>(ki), sizeof(ki));
        f.read(reinterpret_cast<char*>(d), sizeof(d));
        c.kind = ki[0];
        c.n = ki[1];
        c.h = d[0];
        c.coef = d[1];
        c.dt = d[2];
        const size_t N = size_t(c.n * c.n * c.n);
        c.f0 = Eigen::Matrix<float, -1, 1>{N};
        f.read(reinterpret_cast<char*>(c.f0.data()), std::streamsize(N * sizeof(float)));
    }
    if (!f) throw std::runtime_error("short case file");
    CFDArcoGlobalInit::initialize(argc, argv, true);
    CFDArcoGlobalInit::enable_cuda(1, true, false);
    // set up every case and compile its kernels (one warm-up step each)
    std::vector<std::unique_ptr<Solver>> solvers;
    for (auto& c : cases) {
        solvers.push_back(make_solver(c));
        auto& s = *solvers.back();
        Equation(1).evaluate(s.all, s.eqs, s.dt.get(), false, s.all);
    }
    sync_device();
    const auto t0 = std::chrono::steady_clock::now();
    for (auto& sp : solvers) Equation(size_t(steps)).evaluate(sp->all, sp->eqs, sp->dt.get(), false, sp->all);
    sync_device();
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "{\"ensemble_api_seconds\": " << sec << ", \"cases\": " << ncases << ", \"steps\": " << steps << "}"
              << std::endl;
    CFDArcoGlobalInit::finalize();
    return 0;
}
