// renderer.cpp -- Renderer::render and the tracer registry over the C ABI.
#include "renderer.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <stdexcept>

#include "log.h"
#include "png.h"

namespace koishi {
namespace core {

namespace {

void kr_check(int rc, const char *what) {
    if (rc == KR_OK) return;
    char buf[1024] = {0};
    kr_last_error(buf, sizeof buf);
    throw std::logic_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + buf);
}

// Factory<Kernel> names (src/core/kernel/kernel.impl.cu:10-15)
uint32_t kernel_type(const std::string &name) {
    if (name == "Radiance") return KR_KERNEL_RADIANCE;
    if (name == "Normal") return KR_KERNEL_NORMAL;
    if (name == "Visibility") return KR_KERNEL_VISIBILITY;
    if (name == "HalfVector") return KR_KERNEL_HALFVECTOR;
    throw std::logic_error("no such type: " + name);
}

// RAII over the kr_* handles
struct DeviceScene {
    kr_device *dev = nullptr;
    kr_scene *scene = nullptr;
    kr_render *render = nullptr;
    ~DeviceScene() {
        if (render) kr_render_free(render);
        if (scene) kr_scene_free(scene);
        if (dev) kr_device_close(dev);
    }
};

// Profiler::~Profiler (src/core/kernel/kernel.cu:53-73) writing every slice
// with RadianceKernel::Slice::writeSlice (src/core/kernel/radiance.hpp:31-46):
// "at [ x, y, k ]: " for every path of the area, then, for paths whose
// radiance reaches 100 in all channels, the loop counter and the recorded
// vertices. Vectors print as the reference's "[ a, b, c ]", floats with the
// stream's default format.
void write_profile(kr_render *r, const std::string &path, const uint32_t area[4], uint32_t spp) {
    std::ofstream of(path);
    if (!of) throw std::logic_error("cannot write '" + path + "'");
    uint64_t n = 0;
    uint32_t depth = 0;
    kr_check(kr_render_profiler_size(r, &n, &depth), "kr_render_profiler_size");
    const uint32_t aw = area[2] - area[0];
    auto v3 = [&](const float *v) { of << "[ " << v[0] << ", " << v[1] << ", " << v[2] << " ]"; };
    const uint64_t kRead = 1u << 16;
    std::vector<float> L;
    std::vector<uint32_t> bounce;
    std::vector<kr_slice_record> recs;
    for (uint64_t first = 0; first < n; first += kRead) {
        uint64_t cnt = std::min(kRead, n - first);
        L.resize(3 * cnt);
        bounce.resize(cnt);
        recs.resize(std::max<uint64_t>(1, cnt * depth));
        kr_check(kr_render_read_slices(r, first, cnt, L.data(), bounce.data(), recs.data()), "kr_render_read_slices");
        for (uint64_t i = 0; i < cnt; ++i) {
            uint64_t sl = first + i, px = sl / spp;
            of << "at [ " << area[0] + px % aw << ", " << area[1] + px / aw << ", " << sl % spp << " ]: " << std::endl;
            const float *l = &L[3 * i];
            if (!(l[0] >= 1e2f && l[1] >= 1e2f && l[2] >= 1e2f)) continue;
            of << "bounce: " << bounce[i] << " / " << depth << std::endl;
            for (uint32_t j = 0; j < std::min(bounce[i], depth); ++j) {
                const kr_slice_record &b = recs[i * depth + j];
                of << "ray: ";
                v3(b.o);
                v3(b.d);
                of << std::endl << "f: ";
                v3(b.f);
                of << std::endl << "theta: " << b.theta << std::endl << "L: ";
                v3(b.L);
                of << std::endl << "shader: " << b.shader << std::endl;
            }
        }
    }
}

// The wavefront integrator behind Tracer::execute: uploads the flat scene,
// renders every pixel's samples in FP64 accumulators on the device and reads
// back sums and valid counts (src/core/tracer/cpuMulticore.cu:47-64 semantics).
class HIPWavefrontTracer : public Tracer {
public:
    explicit HIPWavefrontTracer(const Json &props) {
        if (const Json *k = props.find("kernel")) kernel_ = (*k)[0].str();
    }
    void execute(Image &image, const kr_camera &camera, const LoadedScene &scene, uint32_t spp,
                 const RenderOptions &opt) override {
        const std::string kernel = opt.kernel.empty() ? scene.kernel : opt.kernel;
        uint32_t ktype = kernel_type(kernel);
        DeviceScene ds;
        kr_check(kr_device_open(opt.device, &ds.dev), "kr_device_open");
        kr_scene_desc d = scene.desc(opt.bvh_width);
        kr_check(kr_scene_upload(ds.dev, &d, &ds.scene), "kr_scene_upload");
        kr_scene_info info{};
        kr_scene_get_info(ds.scene, &info);
        KLOG("Scene on device:", info.num_triangles, "triangles,", info.num_spheres, "spheres,", info.num_lights,
             "lights, BVH" + std::to_string(info.bvh_width), info.bvh_nodes, "nodes");
        KLOG(info.num_lights, "lights in scene");
        kr_params p{};
        p.width = image.w;
        p.height = image.h;
        p.spp = spp;
        p.max_bounce = scene.max_bounce;
        p.seed = opt.seed;
        p.kernel = ktype;
        kr_check(kr_render_create(ds.scene, &camera, &p, &ds.render), "kr_render_create");
        // Profiler::Profiler (src/core/kernel/kernel.cu:26-51): the area
        // clamped to the image, one slice per path of the area
        uint32_t area[4] = {std::min(scene.profiler_area[0], image.w), std::min(scene.profiler_area[1], image.h),
                            std::min(scene.profiler_area[2], image.w), std::min(scene.profiler_area[3], image.h)};
        if (scene.profiler_enable) {
            KINFO("profiler", "Enabled in", "[ " + std::to_string(area[0]) + ", " + std::to_string(area[1]) + ", " +
                                                std::to_string(area[2]) + ", " + std::to_string(area[3]) + " ]");
            kr_profiler_config pc{{area[0], area[1], area[2], area[3]}, scene.profiler_depth};
            kr_check(kr_render_set_profiler(ds.render, &pc), "kr_render_set_profiler");
            uint64_t slices = 0;
            kr_check(kr_render_profiler_size(ds.render, &slices, nullptr), "kr_render_profiler_size");
            KINFO("profiler", "Allocoated ", slices, "profile slices");
        } else {
            KINFO("profiler", "Disabled");
        }
        KINFO("tracer", "Tracing start");
        auto t0 = std::chrono::steady_clock::now();
        kr_check(kr_render_run(ds.render, 0, spp), "kr_render_run");
        kr_check(kr_render_sync(ds.render), "kr_render_sync");
        double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        KINFO("tracer", "Tracer joint in", sec, "seconds");
        image.rgb.assign(3 * size_t(image.w) * image.h, 0.0);
        image.valid.assign(size_t(image.w) * image.h, 0);
        kr_check(kr_render_read(ds.render, image.rgb.data(), image.valid.data()), "kr_render_read");
        kr_render_stats st{};
        kr_render_get_stats(ds.render, &st);
        KLOG(double(st.paths) / sec, "paths/s,", double(st.closest_rays + st.shadow_rays) / sec * 1e-6, "Mrays/s,",
             st.invalid_samples, "samples rejected (NaN/Inf)");
        if (scene.profiler_enable) {
            KINFO("profiler", "Writting profile");
            write_profile(ds.render, scene.profiler_output, area, spp);
            KINFO("profiler", "Written to '" + scene.profiler_output + "'");
        }
    }

private:
    std::string kernel_ = "Radiance";
};

// The same integrator through kr_batch: the render is a batch of one job,
// captured into a HIP graph and replayed with one launch.
class HIPBatchedTracer : public Tracer {
public:
    explicit HIPBatchedTracer(const Json &) {}
    void execute(Image &image, const kr_camera &camera, const LoadedScene &scene, uint32_t spp,
                 const RenderOptions &opt) override {
        const std::string kernel = opt.kernel.empty() ? scene.kernel : opt.kernel;
        if (kernel_type(kernel) != KR_KERNEL_RADIANCE)
            throw std::logic_error("tracer HIPBatched renders the Radiance kernel only");
        if (scene.profiler_enable) throw std::logic_error("tracer HIPBatched has no profiler");
        struct Handles {
            kr_device *dev = nullptr;
            kr_batch *batch = nullptr;
            ~Handles() {
                if (batch) kr_batch_free(batch);
                if (dev) kr_device_close(dev);
            }
        } h;
        kr_check(kr_device_open(opt.device, &h.dev), "kr_device_open");
        kr_scene_desc d = scene.desc(opt.bvh_width);
        kr_batch_job job{};
        job.scene = 0;
        job.camera = camera;
        job.params.width = image.w;
        job.params.height = image.h;
        job.params.spp = spp;
        job.params.max_bounce = scene.max_bounce;
        job.params.seed = opt.seed;
        kr_batch_tuning tu{opt.bvh_width ? opt.bvh_width : 4u, 32u << 20, 64, 0, 0.0};
        kr_check(kr_batch_create(h.dev, &d, 1, &job, 1, &tu, &h.batch), "kr_batch_create");
        KINFO("tracer", "Tracing start");
        auto t0 = std::chrono::steady_clock::now();
        kr_check(kr_batch_run(h.batch), "kr_batch_run");
        kr_check(kr_batch_sync(h.batch), "kr_batch_sync");
        double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        KINFO("tracer", "Tracer joint in", sec, "seconds");
        image.rgb.assign(3 * size_t(image.w) * image.h, 0.0);
        image.valid.assign(size_t(image.w) * image.h, 0);
        kr_check(kr_batch_read(h.batch, 0, image.rgb.data(), image.valid.data()), "kr_batch_read");
        kr_batch_info info{};
        kr_batch_get_info(h.batch, &info);
        KLOG(double(info.total_paths) / sec, "paths/s,", info.graph_nodes, "graph nodes");
    }
};

std::map<std::string, TracerCtor> &registry() {
    static std::map<std::string, TracerCtor> r = {
        {"HIPWavefront", [](const Json &p) { return std::unique_ptr<Tracer>(new HIPWavefrontTracer(p)); }},
        {"CPUMulticore", [](const Json &p) { return std::unique_ptr<Tracer>(new HIPWavefrontTracer(p)); }},
        {"HIPBatched", [](const Json &p) { return std::unique_ptr<Tracer>(new HIPBatchedTracer(p)); }},
    };
    return r;
}

}  // namespace

std::unique_ptr<Tracer> create_tracer(const std::string &name, const Json &props) {
    auto &r = registry();
    auto it = r.find(name);
    if (it == r.end()) throw std::logic_error("no such type: " + name);
    if (name == "CPUMulticore") KLOG("Tracer CPUMulticore is served by the HIPWavefront integrator");
    return it->second(props);
}

std::vector<std::string> tracer_names() {
    std::vector<std::string> v;
    for (auto &kv : registry()) v.push_back(kv.first);
    return v;
}

std::vector<std::string> kernel_names() { return {"Radiance", "Normal", "Visibility", "HalfVector"}; }

void Image::dump_png(const std::string &path) const {
    std::vector<uint8_t> bytes(3 * size_t(w) * h);
    for (size_t i = 0; i < size_t(w) * h; ++i)
        for (int c = 0; c < 3; ++c) bytes[3 * i + c] = to_byte(rgb[3 * i + c] / double(valid[i]));
    write_png(path, w, h, bytes.data());
}

// Raw dump: "KRRAW1\0\0", uint32 w, uint32 h, w*h*3 float64 pixel means
// (sum / valid), w*h uint32 valid counts; little endian.
void Image::dump_raw(const std::string &path) const {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::logic_error("cannot write '" + path + "'");
    const char magic[8] = {'K', 'R', 'R', 'A', 'W', '1', 0, 0};
    f.write(magic, 8);
    f.write(reinterpret_cast<const char *>(&w), 4);
    f.write(reinterpret_cast<const char *>(&h), 4);
    std::vector<double> mean(rgb.size());
    for (size_t i = 0; i < valid.size(); ++i)
        for (int c = 0; c < 3; ++c) mean[3 * i + c] = rgb[3 * i + c] / double(valid[i]);
    f.write(reinterpret_cast<const char *>(mean.data()), std::streamsize(mean.size() * 8));
    f.write(reinterpret_cast<const char *>(valid.data()), std::streamsize(valid.size() * 4));
}

void Renderer::render(const std::string &path, const std::string &dest, uint32_t spp, const RenderOptions &opt) {
    KINFO("renderer", "Render start");
    KINFO("renderer", "Loading config '" + path + "'");
    LoadedScene scene = load_scene_file(path);
    if (scene.cameras.empty()) throw std::logic_error("no camera in the scene");
    const kr_camera &camera = scene.cameras[0];
    Image image;
    image.w = w_;
    image.h = h_;
    KLOG("Target resolution:", w_, "x", h_, ",", spp, "spp");
    std::string tname = opt.tracer.empty() ? scene.tracer : opt.tracer;
    auto tracer = create_tracer(tname, scene.tracer_props);
    KLOG("Start intergrating");
    tracer->execute(image, camera, scene, spp, opt);
    KLOG("Finished intergrating");
    KINFO("renderer", "Writting image to file");
    image.dump_png(dest);
    if (!opt.dump_raw.empty()) image.dump_raw(opt.dump_raw);
    KINFO("renderer", "Written to '" + dest + "'");
    KINFO("renderer", "Render finished successfully");
}

}  // namespace core
}  // namespace koishi
