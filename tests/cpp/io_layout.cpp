// Writes the on-disk output files of the reference's layout on the host only
// (no device): mesh.json of a perturbed 3 x 2 x 4 mesh, a .vti and a .bin of a
// field whose value encodes each cell's lattice position, and the run_latest
// link.  tests/test_io_layout.py reads them back.
#include <cstdio>
#include <iostream>
#include <vector>

#include "io_operators_tmpl.hpp"
#include "mesh3d.hpp"

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: io_layout <outdir>\n";
        return 2;
    }
    const fs::path out = argv[1];
    const fs::path run = out / "run_test";
    fs::create_directories(run / "T");
    Mesh3D mesh(3, 2, 4, 0.5f, 0.25f, 0.125f);
    mesh.init_basic_internals();
    mesh.perturb_interior_vertices(11, 0.2);
    store_mesh(&mesh, run);
    // value of cell (x, y, z) = 100 x + 10 y + z, stored in the mesh's cell order
    std::vector<float> f(mesh._num_nodes);
    for (size_t x = 0; x < mesh._x; ++x)
        for (size_t y = 0; y < mesh._y; ++y)
            for (size_t z = 0; z < mesh._z; ++z) f[mesh.square_node_coord_to_idx(x, y, z)] = float(100 * x + 10 * y + z);
    cfdarco::io::write_vti(run / "T" / "res_0.vti", mesh._x, mesh._y, mesh._z, f.data());
    cfdarco::io::write_bin(run / "T" / "0.bin", f.data(), f.size());
    cfdarco::io::link_latest(run, out / "run_latest");
    // the offsets the vertex coordinates carry, for the reader's check
    const auto& off = mesh.vertex_offsets();
    FILE* o = std::fopen((out / "offsets.f64").c_str(), "wb");
    std::fwrite(off.data(), sizeof(double), off.size(), o);
    std::fclose(o);
    // -m input: the same file read back gives the extents, spacing and vertices
    const auto mf = cfdarco::io::read_lattice_mesh_json((run / "mesh.json").string());
    if (mf.x != mesh._x || mf.y != mesh._y || mf.z != mesh._z || mf.dx != mesh._dx || mf.dy != mesh._dy ||
        mf.dz != mesh._dz || mf.vertex_offsets.size() != off.size()) {
        std::cerr << "mesh.json read back with other extents or spacing\n";
        return 1;
    }
    const float d[3] = {mesh._dx, mesh._dy, mesh._dz};
    size_t v = 0;
    for (size_t x = 0; x <= mesh._x; ++x)
        for (size_t y = 0; y <= mesh._y; ++y)
            for (size_t z = 0; z <= mesh._z; ++z, ++v) {
                const size_t lat[3] = {x, y, z};
                for (int k = 0; k < 3; ++k) {
                    const float base = float(lat[k]) * d[k];
                    if (base + float(mf.vertex_offsets[3 * v + k]) != base + float(off[3 * v + k])) {
                        std::cerr << "vertex " << v << " read back elsewhere\n";
                        return 1;
                    }
                }
            }
    std::cout << "ok\n";
    return 0;
}
