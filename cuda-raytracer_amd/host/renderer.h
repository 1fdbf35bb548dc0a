// renderer.h -- host orchestration of the drop-in `cr`: core::Renderer and
// the tracer / kernel plugin registry.
//
// Mirrors core::Renderer (src/core/renderer/renderer.hpp:15-32) and the
// Factory<Tracer> / Factory<Kernel> registries
// (src/core/tracer/tracer.impl.cu:8-11, src/core/kernel/kernel.impl.cu:10-15).
#pragma once
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "kr_api.h"
#include "scene_loader.h"

namespace koishi {
namespace core {

struct Image {
    uint32_t w = 0, h = 0;
    std::vector<double> rgb;     // sum of valid samples, 3 per pixel
    std::vector<uint32_t> valid; // valid samples per pixel
    double pixel(uint32_t i, int c) const { return rgb[3 * size_t(i) + c] / double(valid[i]); }
    void dump_png(const std::string &path) const;
    void dump_raw(const std::string &path) const;
};

struct RenderOptions {
    std::string tracer, kernel;  // empty: from the scene file
    uint64_t seed = 1;
    std::string dump_raw;
    int device = 0;
    uint32_t bvh_width = 0;
};

// Tracer::execute (src/core/tracer/tracer.hpp:20-21) over the C ABI.
struct Tracer {
    virtual ~Tracer() = default;
    virtual void execute(Image &image, const kr_camera &camera, const LoadedScene &scene, uint32_t spp,
                         const RenderOptions &opt) = 0;
};

using TracerCtor = std::function<std::unique_ptr<Tracer>(const Json &props)>;

// The registry: "HIPWavefront" (the gfx950 wavefront integrator),
// "HIPBatched" (the same integrator as a one-job kr_batch replayed from a
// HIP graph) and "CPUMulticore", the reference's only tracer and its
// default, served by the wavefront integrator so reference scene files
// render unchanged.
// Unknown names throw "no such type:" (src/core/basic/factory.hpp:18-21).
std::unique_ptr<Tracer> create_tracer(const std::string &name, const Json &props);
std::vector<std::string> tracer_names();
std::vector<std::string> kernel_names();

class Renderer {
public:
    Renderer(uint32_t w, uint32_t h) : w_(w), h_(h) {}
    // Renderer::render (src/core/renderer/renderer.cu:13-52)
    void render(const std::string &path, const std::string &dest, uint32_t spp, const RenderOptions &opt = {});

private:
    uint32_t w_, h_;
};

}  // namespace core
}  // namespace koishi
