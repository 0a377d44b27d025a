// On-disk output of the drivers in the reference's layout
// (include/io_operators_tmpl.hpp:44-212 of the reference), written by hand (no
// VTK or JSON library):
//   store_history:  <store_dir>/<var>/res_<i>.vti for every history entry but
//                   the last, VTK XML ImageData of X x Y x Z points with the
//                   cell values as one Float32 point array (x fastest), and
//                   the symlink <store_dir>/../run_latest -> <store_dir>;
//   stepping:       <store_dir>/mesh.json (vertexes, faces, nodes and the
//                   extents), then <store_dir>/<var>/<i>.bin (the raw float
//                   field after step i) every step, and run_latest at the end.
// The stepping snapshots are asynchronous: each step's fields are copied to
// page-locked host buffers on a side stream and written by a writer thread
// while the next steps run.
#pragma once

#include <filesystem>
#include <string>
#include <vector>

class Variable;
class Mesh3D;

namespace fs = std::filesystem;

void store_history(const std::vector<Variable*>& vars_to_store, const Mesh3D* mesh);
void store_mesh(const Mesh3D* mesh, const fs::path& store_path);
void init_store_history_stepping(const std::vector<Variable*>& vars_to_store, const Mesh3D* mesh,
                                 const fs::path& store_path = "./dumps");
void store_history_stepping(const std::vector<Variable*>& vars_to_store, const Mesh3D* mesh, int i);
void finalize_history_stepping(const fs::path& store_path = "./dumps");

namespace cfdarco::io {
// A mesh.json of lattice connectivity (the -m/--mesh input): the extents, the
// spacing and the vertex offsets (empty when the vertices sit on the lattice)
// of the Mesh3D it describes.  Throws for other connectivity.
struct LatticeMeshFile {
    size_t x = 0, y = 0, z = 0;
    float dx = 0, dy = 0, dz = 0;
    std::vector<double> vertex_offsets;
};
LatticeMeshFile read_lattice_mesh_json(const std::string& path);

// One field on an X x Y x Z lattice (values in the mesh's cell order) as VTK
// XML ImageData point data named "ImageScalars", inline base64 binary.
void write_vti(const fs::path& file, size_t X, size_t Y, size_t Z, const float* cells);
void write_bin(const fs::path& file, const float* values, size_t n);
// link -> absolute(target), replacing whatever link was there
void link_latest(const fs::path& target, const fs::path& link);
// The time loop's hook: when stepping output is active, queues the snapshot of
// step `step` (Equation::evaluate calls it after each step's boundary update).
bool stepping_active();
void stepping_snapshot(int step);
}  // namespace cfdarco::io
