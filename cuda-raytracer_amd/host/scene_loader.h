// scene_loader.h -- the reference's scene JSON (+ OBJ import) flattened into
// the kr_scene_desc arrays of include/kr_api.h.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "json.h"
#include "kr_api.h"

namespace koishi {

struct ObjMesh {
    std::string object, material;
    std::vector<float> positions, normals;  // 3 per vertex
    std::vector<uint32_t> indices;          // 3 per triangle
};

// Reads an OBJ file into one mesh per (object, material run), the grouping
// assimp gives an OBJ import (src/core/primitive/mesh.cu:332-402 walks it).
// Polygons are fanned into triangles; vertices without normals get smooth
// area-weighted normals. Materials are named by `usemtl`; faces before any
// `usemtl` use "DefaultMaterial". material_names receives "DefaultMaterial"
// first (when faces use it) and then every `usemtl` name in order of first
// mention, with or without faces.
std::vector<ObjMesh> load_obj(const std::string &path, std::vector<std::string> &material_names);
// PLY import (ascii and binary); one mesh with the material "DefaultMaterial".
std::vector<ObjMesh> load_ply(const std::string &path, std::vector<std::string> &material_names);
// By extension: ".ply" (any case) reads PLY, anything else OBJ.
std::vector<ObjMesh> load_mesh_file(const std::string &path, std::vector<std::string> &material_names);

struct LoadedScene {
    std::vector<float> positions, normals;
    std::vector<uint32_t> indices;
    std::vector<kr_prim_desc> prims;
    std::vector<std::string> shader_names;   // index = shader id
    std::vector<kr_shader_desc> shaders;
    std::vector<bool> shader_defined;        // false: no config in "shaders" (named shaders only;
                                             // Mixed children follow them in `shaders`)
    std::vector<kr_point_light> point_lights; // "pointLight" assets in file order
    bool lights_enabled = true;              // false when a shader config is missing
    std::vector<kr_camera> cameras;
    std::string tracer = "CPUMulticore";
    Json tracer_props = Json::object();
    std::string kernel = "Radiance";
    uint32_t max_bounce = 16;
    uint64_t light_seed = 0x4B4F49534849ull;
    // "profiler": Renderer::Configuration's key, absent = disabled
    // (renderer.hpp:19-20), with Profiler::Configuration's fields and
    // defaults (kernel.hpp:37-43) and the slice's "depth" (radiance.hpp:27-30)
    bool profiler_enable = false;
    std::string profiler_output = "a.prof";
    uint32_t profiler_area[4] = {0u, 0u, 0xFFFFFFFFu, 0xFFFFFFFFu};
    uint32_t profiler_depth = 8;

    kr_scene_desc desc(uint32_t bvh_width = 0, uint32_t leaf_size = 0) const;
};

// Scene::Scene (src/core/meta/scene.cu:18-219) + Renderer::Configuration
// (src/core/renderer/renderer.hpp:17-23): assets in order ("import" once,
// "sphere", lights), the shader table by name, the camera list and the
// tracer / kernel selection. Errors throw std::logic_error with the
// reference's messages.
LoadedScene load_scene_file(const std::string &path);
LoadedScene load_scene_json(const Json &doc, const std::string &base_dir);

// Shader config [type, props] -> kr_shader_desc with the reference defaults.
kr_shader_desc shader_from_config(const Json &cfg);

}  // namespace koishi
