// Command line, mesh construction and run set-up of the drivers
// (include/utils3d.hpp:31-223 of the reference).  Flags keep the reference's
// names; -c/--cuda_enable and -hip/--hip_enable both select the MI355X path.
#pragma once

#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "argparse/argparse.hpp"
#include "cfdarcho_main_3d.hpp"
#include "fvm3d.hpp"
#include "io_operators_tmpl.hpp"
#include "mesh3d.hpp"

inline argparse::ArgumentParser parse_args_base(int argc, char** argv) {
    argparse::ArgumentParser program("cfdARCO");
    program.add_argument("-v", "--visualize").default_value(false).implicit_value(true);
    program.add_argument("--create_plot").default_value(false).implicit_value(true);
    program.add_argument("-Lx").help("cells along x").default_value(200).scan<'i', int>();
    program.add_argument("-Ly").help("cells along y").default_value(200).scan<'i', int>();
    program.add_argument("-Lz").help("cells along z").default_value(200).scan<'i', int>();
    program.add_argument("-dx").help("cell spacing along x").default_value(1.0).scan<'g', float>();
    program.add_argument("-dy").help("cell spacing along y").default_value(1.0).scan<'g', float>();
    program.add_argument("-dz").help("cell spacing along z").default_value(1.0).scan<'g', float>();
    program.add_argument("-t", "--timesteps").help("number of time steps").default_value(1000).scan<'i', int>();
    program.add_argument("-dt").help("time step, or the CFL number where a driver sets a CFL policy").default_value(0.005).scan<'g', float>();
    program.add_argument("-c", "--cuda_enable").default_value(false).implicit_value(true);
    program.add_argument("-hip", "--hip_enable").default_value(false).implicit_value(true);
    program.add_argument("--cuda_ranks").default_value(1).scan<'i', int>();
    program.add_argument("-s", "--store").default_value(false).implicit_value(true);
    program.add_argument("-st", "--store_stepping").default_value(false).implicit_value(true);
    program.add_argument("-sl", "--store_last").default_value(false).implicit_value(true);
    program.add_argument("--skip_history").default_value(false).implicit_value(true);
    program.add_argument("-d", "--dist").default_value(std::string("cl"));
    program.add_argument("-p", "--priorities")
        .nargs(argparse::nargs_pattern::any)
        .default_value(std::vector<size_t>{})
        .scan<'i', size_t>();
    program.add_argument("--strange_mesh").default_value(false).implicit_value(true);
    program.add_argument("-m", "--mesh").default_value(std::string(""));
    // seeded interior-vertex perturbation (fraction of the spacing), 0 = lattice
    program.add_argument("--perturb").default_value(0.0).scan<'g', double>();
    program.add_argument("--perturb_seed").default_value(0).scan<'i', int>();
    try {
        program.parse_args(argc, argv);
    } catch (const std::runtime_error& err) {
        std::cerr << err.what() << std::endl;
        std::cerr << program;
        std::exit(1);
    }
    return program;
}

class SingleLibInitializer3D {
public:
    bool visualize;
    bool create_plot;
    size_t Lx, Ly, Lz;
    float dx, dy, dz;
    int timesteps;
    float dt;
    bool cuda_enable;
    bool hip_enable;
    bool store_stepping;
    bool store_last;
    bool store;
    std::vector<size_t> priorities;
    std::shared_ptr<Mesh3D> mesh;
    std::vector<Variable*> store_vars;

    SingleLibInitializer3D(int argc, char** argv) {
        auto program = parse_args_base(argc, argv);
        CFDArcoGlobalInit::initialize(argc, argv, !(program.get<bool>("store")));
        visualize = program.get<bool>("visualize");
        create_plot = program.get<bool>("create_plot");
        Lx = program.get<int>("Lx");
        Ly = program.get<int>("Ly");
        Lz = program.get<int>("Lz");
        dx = program.get<float>("dx");
        dy = program.get<float>("dy");
        dz = program.get<float>("dz");
        timesteps = program.get<int>("timesteps");
        dt = program.get<float>("dt");
        cuda_enable = program.get<bool>("cuda_enable");
        hip_enable = program.get<bool>("hip_enable");
        store_stepping = program.get<bool>("store_stepping");
        store_last = program.get<bool>("store_last");
        store = program.get<bool>("store");
        if (cuda_enable && hip_enable) throw std::runtime_error{"-c/--cuda_enable and -hip/--hip_enable exclude each other"};
        if (!cuda_enable && !hip_enable)
            throw std::runtime_error{
                "this build evaluates on the MI355X only: pass -c/--cuda_enable or -hip/--hip_enable"};
        const std::string mesh_file = program.get<std::string>("mesh");
        std::vector<double> file_offsets;
        if (!mesh_file.empty()) {  // -m: extents, spacing and vertices from a mesh.json
            auto mf = cfdarco::io::read_lattice_mesh_json(mesh_file);
            Lx = mf.x; Ly = mf.y; Lz = mf.z;
            dx = mf.dx; dy = mf.dy; dz = mf.dz;
            file_offsets = std::move(mf.vertex_offsets);
        }
        mesh = std::make_shared<CudaMesh3D>(Lx, Ly, Lz, dx, dy, dz);
        mesh->init_basic_internals();
        if (!file_offsets.empty()) mesh->set_vertex_offsets(std::move(file_offsets));
        const double perturb = program.get<double>("perturb");
        if (perturb > 0) mesh->perturb_interior_vertices(uint64_t(program.get<int>("perturb_seed")), perturb);
        mesh->compute();
        priorities = program.get<std::vector<size_t>>("priorities");
        CFDArcoGlobalInit::enable_cuda(program.get<int>("cuda_ranks"), cuda_enable, hip_enable);
        dynamic_cast<CudaMesh3D*>(mesh.get())->host_to_device();
    }

    // With -s, the stored variables' history goes to .vti files at finalize();
    // with -s -st, mesh.json is written here and every step's state to .bin
    // files as the run goes (the reference's stepping writers,
    // include/io_operators_tmpl.hpp:105-159, whose calls it leaves commented
    // out at include/utils3d.hpp:182,197-198 and src/fvm3d.cpp:1057)
    void init_store(const std::vector<Variable*>& vars_to_store) {
        store_vars = vars_to_store;
        if (store && store_stepping) init_store_history_stepping(store_vars, mesh.get());
    }

    void finalize() {
        if (store) {
            if (store_stepping) finalize_history_stepping();
            else store_history(store_vars, mesh.get());
        }
        CFDArcoGlobalInit::finalize();
    }
};

// Stores val at cell (x, y, z) of a field in the reference cell order: `rows`
// is the z extent and `cols` the y extent, so the cell id is
// (x * cols + y) * rows + z.
inline void sett(Eigen::Matrix<float, -1, 1>& v, int rows, int cols, int x, int y, int z, float val) {
    v[(size_t(x) * size_t(cols) + size_t(y)) * size_t(rows) + size_t(z)] = val;
}

template <typename MeshClass>
inline Eigen::Matrix<float, -1, 1> initial_with_val(MeshClass* mesh, float val) {
    auto ret = Eigen::Matrix<float, -1, 1>{mesh->_num_nodes};
    ret.setConstant(val);
    return ret;
}
