// The second-order wave example (examples/main_3d_wave.cu:61-90 of the
// reference) with its two point sources applied on the device.
//
// The example as shipped registers only the host form of its boundary
// callback, so every step brings the whole field to the host, copies it in the
// callback and sends it back.  This program solves the same problem
// (d2t(u) = c^2 (d2dx(u) + d2dy(u)), c = 0.3, u = 0, dt = 0.1 constant, sources
// sin(0.2 * step) at 10% and 90% of the box) with a device callback as well, so
// the solver loop stays on the GPU.  It takes the reference harness's
// arguments (-Lx -Ly -Lz -dx -dy -dz -t -c ...) and prints the same
// "Time difference = ...[microseconds]" line.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <iostream>

#include "fvm3d.hpp"
#include "mesh3d.hpp"
#include "utils3d.hpp"

namespace {

__global__ void point_sources_k(float* a, int64_t i1, int64_t i2, float value) {
    if (threadIdx.x == 0) a[i1] = value;
    if (threadIdx.x == 1) a[i2] = value;
}

struct Sources {
    size_t i1, i2;
};

Sources sources_of(Mesh3D* mesh) {
    return {mesh->square_node_coord_to_idx(size_t(mesh->_x * 0.1), size_t(mesh->_y * 0.1), size_t(mesh->_z * 0.1)),
            mesh->square_node_coord_to_idx(size_t(mesh->_x * 0.9), size_t(mesh->_y * 0.9), size_t(mesh->_z * 0.9))};
}

float source_value(const DT* dt) { return std::sin(static_cast<float>(dt->_current_time_step_int) * 0.2f); }

Eigen::Matrix<float, -1, 1> sources_host(Mesh3D* mesh, Eigen::Matrix<float, -1, 1>& arr, const DT* dt) {
    const Sources s = sources_of(mesh);
    Eigen::Matrix<float, -1, 1> ret{arr};
    ret[Eigen::Index(s.i1)] = source_value(dt);
    ret[Eigen::Index(s.i2)] = source_value(dt);
    return ret;
}

// writes the two source cells in place (the returned matrix shares the
// argument's storage)
CudaDataMatrixD sources_device(Mesh3D* mesh, CudaDataMatrixD& arr, const DT* dt) {
    const Sources s = sources_of(mesh);
    if (s.i1 >= arr._size || s.i2 >= arr._size) throw std::runtime_error("point source outside the field");
    point_sources_k<<<1, 64>>>(arr.data.get(), int64_t(s.i1), int64_t(s.i2), source_value(dt));
    if (hipGetLastError() != hipSuccess) throw std::runtime_error("point source launch failed");
    return arr;
}

}  // namespace

int main(int argc, char** argv) {
    SingleLibInitializer3D initializer{argc, argv};
    auto mesh = initializer.mesh;
    auto timesteps = initializer.timesteps;

    auto initial_zero = initial_with_val(mesh.get(), 0);
    auto u = Variable(mesh.get(), initial_zero, sources_host, sources_device, "u");
    std::vector<Variable*> space_vars{&u};
    auto dt = DT(mesh.get(), UpdatePolicies::constant_dt, UpdatePolicies::constant_dt_cu, 0.1, space_vars);
    const float c = 0.3f;
    EquationTemplate equation_system = {
        {d2t(u), '=', c * c * (d2dx(u) + d2dy(u)), true},
    };
    std::vector<Variable*> all_vars{&u};
    auto equation = Equation(timesteps);
    initializer.init_store(all_vars);

    const auto begin = std::chrono::steady_clock::now();
    equation.evaluate(all_vars, equation_system, &dt, initializer.visualize, all_vars);
    const auto end = std::chrono::steady_clock::now();
    std::cout << std::endl
              << "Time difference = " << std::chrono::duration_cast<std::chrono::microseconds>(end - begin).count()
              << "[microseconds]" << std::endl;
    initializer.finalize();
    return 0;
}
