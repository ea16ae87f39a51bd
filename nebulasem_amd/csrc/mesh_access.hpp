// Internal accessors of the opaque nsem_mesh for the other C-ABI units.
#pragma once
#include <string>
#include <vector>

#include "host/mesh.hpp"
#include "nebulasem_amd.h"

const nsem::Geometry& nsem_mesh_geometry(const nsem_mesh* m);
const nsem::LoadedMesh& nsem_mesh_loaded(const nsem_mesh* m);
const std::vector<std::string>& nsem_mesh_boundary_names(const nsem_mesh* m);
