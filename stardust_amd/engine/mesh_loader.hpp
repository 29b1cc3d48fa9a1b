// STL reading for DEMMesh (reference: src/IO/mesh_loader.cpp:12-44, which
// calls the third-party stl_reader::ReadStlFile of the external/stl_reader
// submodule, not vendored). Restated here from that library's published
// behaviour: binary or ASCII STL; every triangle corner is merged with the
// identical corners of other triangles, the unique vertices in lexicographic
// (x, y, z) order; three vertex indices and the file's normal per triangle.
#ifndef STARDUST_AMD_ENGINE_MESH_LOADER_HPP
#define STARDUST_AMD_ENGINE_MESH_LOADER_HPP

#include <vector>

#include "types.hpp"

namespace STARDUST {

// Throws std::runtime_error for a file that cannot be read or parsed.
void loadAndConvertMesh(const char* filename, std::vector<float4>& out_vert, std::vector<int>& out_idx,
                        std::vector<float4>& out_norm);

}  // namespace STARDUST

#endif
