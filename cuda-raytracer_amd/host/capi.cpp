// capi.cpp -- C entry points of libkoishi.so (include/koishi.h).
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "koishi.h"
#include "log.h"
#include "png.h"
#include "renderer.h"
#include "scene_loader.h"

struct koishi_scene {
    koishi::LoadedScene s;
};

namespace {
thread_local std::string g_err;

int catch_all(const std::exception &e) {
    g_err = e.what();
    return KR_E_INVALID;
}

void copy_str(const std::string &s, char *buf, size_t len) {
    if (buf && len) std::snprintf(buf, len, "%s", s.c_str());
}
}  // namespace

#define KOISHI_EXPORT extern "C" __attribute__((visibility("default")))

KOISHI_EXPORT int koishi_last_error(char *buf, size_t len) {
    if (!buf || !len) return KR_E_INVALID;
    copy_str(g_err, buf, len);
    return KR_OK;
}

KOISHI_EXPORT void koishi_set_logging(int on) { koishi::log_enabled() = on != 0; }

KOISHI_EXPORT int koishi_scene_load(const char *path, koishi_scene **out) {
    if (!path || !out) return KR_E_INVALID;
    *out = nullptr;
    try {
        auto *s = new koishi_scene();
        try {
            s->s = koishi::load_scene_file(path);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return KR_OK;
    } catch (const std::exception &e) {
        return catch_all(e);
    }
}

KOISHI_EXPORT void koishi_scene_free(koishi_scene *s) { delete s; }

KOISHI_EXPORT int koishi_scene_desc(const koishi_scene *s, kr_scene_desc *out) {
    if (!s || !out) return KR_E_INVALID;
    *out = s->s.desc();
    return KR_OK;
}

KOISHI_EXPORT int koishi_scene_summary(const koishi_scene *s, uint32_t *num_cameras, uint32_t *max_bounce,
                                       char *tracer, size_t tracer_len, char *kernel, size_t kernel_len) {
    if (!s) return KR_E_INVALID;
    if (num_cameras) *num_cameras = uint32_t(s->s.cameras.size());
    if (max_bounce) *max_bounce = s->s.max_bounce;
    copy_str(s->s.tracer, tracer, tracer_len);
    copy_str(s->s.kernel, kernel, kernel_len);
    return KR_OK;
}

KOISHI_EXPORT int koishi_scene_camera(const koishi_scene *s, uint32_t index, kr_camera *out) {
    if (!s || !out || index >= s->s.cameras.size()) return KR_E_INVALID;
    *out = s->s.cameras[index];
    return KR_OK;
}

KOISHI_EXPORT int koishi_scene_shader_name(const koishi_scene *s, uint32_t index, char *buf, size_t len) {
    if (!s || index >= s->s.shader_names.size()) return KR_E_INVALID;
    copy_str(s->s.shader_names[index], buf, len);
    return KR_OK;
}

KOISHI_EXPORT int koishi_render_file(const char *scene_path, const char *png_path, uint32_t width, uint32_t height,
                                     uint32_t spp, const koishi_render_options *o) {
    if (!scene_path || !png_path) return KR_E_INVALID;
    try {
        koishi::core::RenderOptions opt;
        if (o) {
            if (o->tracer) opt.tracer = o->tracer;
            if (o->kernel) opt.kernel = o->kernel;
            opt.seed = o->seed;
            if (o->dump_raw) opt.dump_raw = o->dump_raw;
            opt.device = o->device;
            opt.bvh_width = o->bvh_width;
        }
        koishi::core::Renderer(width, height).render(scene_path, png_path, spp, opt);
        return KR_OK;
    } catch (const std::exception &e) {
        return catch_all(e);
    }
}

KOISHI_EXPORT int koishi_write_png(const char *path, uint32_t w, uint32_t h, const double *rgb, const uint32_t *valid) {
    if (!path || !rgb || !valid) return KR_E_INVALID;
    try {
        koishi::core::Image im;
        im.w = w;
        im.h = h;
        im.rgb.assign(rgb, rgb + 3 * size_t(w) * h);
        im.valid.assign(valid, valid + size_t(w) * h);
        im.dump_png(path);
        return KR_OK;
    } catch (const std::exception &e) {
        return catch_all(e);
    }
}
