// History output (include/io_operators_tmpl.hpp:44-212 of the reference):
// VTK XML ImageData (.vti, ASCII) per stored step with one cell array per
// variable, and mesh.json with the lattice extents and spacing.  Unlike the
// reference's writer, every history entry is written.
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>

#include "cfdarcho_main_3d.hpp"
#include "fvm3d.hpp"
#include "io_operators_tmpl.hpp"

void store_mesh(Mesh3D* mesh) {
    fs::create_directories(CFDArcoGlobalInit::store_dir);
    std::ofstream f(CFDArcoGlobalInit::store_dir / "mesh.json");
    f << "{\"x\": " << mesh->_x << ", \"y\": " << mesh->_y << ", \"z\": " << mesh->_z << ", \"dx\": " << mesh->_dx
      << ", \"dy\": " << mesh->_dy << ", \"dz\": " << mesh->_dz << ", \"num_nodes\": " << mesh->_num_nodes << "}\n";
}

void store_history(const std::vector<Variable*>& vars_to_store, Mesh3D* mesh) {
    if (vars_to_store.empty()) return;
    fs::create_directories(CFDArcoGlobalInit::store_dir);
    store_mesh(mesh);
    size_t steps = vars_to_store.front()->history.size();
    for (auto* v : vars_to_store) steps = std::min(steps, v->history.size());
    for (size_t t = 0; t < steps; ++t) {
        std::ofstream f(CFDArcoGlobalInit::store_dir / ("step_" + std::to_string(t) + ".vti"));
        f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
        // cell data is indexed with x fastest in VTK; the mesh numbers cells z fastest
        f << "<ImageData WholeExtent=\"0 " << mesh->_x << " 0 " << mesh->_y << " 0 " << mesh->_z
          << "\" Origin=\"0 0 0\" Spacing=\"" << mesh->_dx << " " << mesh->_dy << " " << mesh->_dz << "\">\n";
        f << "<Piece Extent=\"0 " << mesh->_x << " 0 " << mesh->_y << " 0 " << mesh->_z << "\">\n<CellData>\n";
        f << std::setprecision(9);
        for (auto* v : vars_to_store) {
            const auto& h = v->history[t];
            f << "<DataArray type=\"Float32\" Name=\"" << v->name << "\" format=\"ascii\">\n";
            for (size_t z = 0; z < mesh->_z; ++z)
                for (size_t y = 0; y < mesh->_y; ++y)
                    for (size_t x = 0; x < mesh->_x; ++x) f << h(Eigen::Index(mesh->square_node_coord_to_idx(x, y, z))) << " ";
            f << "\n</DataArray>\n";
        }
        f << "</CellData>\n</Piece>\n</ImageData>\n</VTKFile>\n";
    }
    std::cout << "Stored " << steps << " steps in " << CFDArcoGlobalInit::store_dir << std::endl;
}
