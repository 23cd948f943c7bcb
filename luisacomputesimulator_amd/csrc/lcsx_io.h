// Scene JSON / OBJ input.
#pragma once
#include <string>
#include <vector>
#include "../../include/lcsx.h"
#include "lcsx_setup.h"

namespace lcsx
{
bool read_obj(const std::string& path, std::vector<float>& verts, std::vector<int32_t>& tris, std::string& err);
bool load_scene_json(const std::string& path, const std::string& mesh_dir, lcsx_params& prm, std::vector<MeshDesc>& meshes,
                     std::string& err);
}  // namespace lcsx

