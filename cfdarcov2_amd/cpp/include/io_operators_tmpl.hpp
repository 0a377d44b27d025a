// History output of the drivers (include/io_operators_tmpl.hpp:44-212 of the
// reference): one VTK XML ImageData file per stored step with every stored
// variable as a cell array, plus mesh.json with the mesh extents, under
// CFDArcoGlobalInit::store_dir.  Written by hand (no VTK dependency).
#pragma once

#include <vector>

class Variable;
class Mesh3D;

void store_history(const std::vector<Variable*>& vars_to_store, Mesh3D* mesh);
void store_mesh(Mesh3D* mesh);
