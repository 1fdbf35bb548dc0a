/*
 * koishi.h -- C entry points of the host side (libkoishi.so): the reference's
 * scene-JSON + OBJ loading and the `cr` render flow, above the device C ABI
 * of kr_api.h.
 *
 * The reference drives this path from C++ (core::Renderer::render,
 * src/core/renderer/renderer.cu:13-52, called by main, src/cr.cu:53-60);
 * libkoishi.so exposes the same steps to C callers and to the tests.
 */
#ifndef KOISHI_H
#define KOISHI_H

#include <stddef.h>
#include <stdint.h>

#include "kr_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct koishi_scene koishi_scene;

/* Last error message of the calling thread (std::logic_error text). */
int koishi_last_error(char *buf, size_t len);

/* Turns the reference-style log lines on stdout on (1) or off (0). */
void koishi_set_logging(int on);

/* Parses a scene JSON (and its "import"ed OBJ) as Scene::Scene does
 * (src/core/meta/scene.cu:18-219). Returns KR_E_INVALID with the message in
 * koishi_last_error on any error. */
int koishi_scene_load(const char *path, koishi_scene **out);
void koishi_scene_free(koishi_scene *scene);

/* Flat description for kr_scene_upload; pointers live as long as the scene. */
int koishi_scene_desc(const koishi_scene *scene, kr_scene_desc *out);

/* Number of cameras, tracer / kernel names and the Radiance "maxBounce". */
int koishi_scene_summary(const koishi_scene *scene, uint32_t *num_cameras, uint32_t *max_bounce, char *tracer,
                         size_t tracer_len, char *kernel, size_t kernel_len);
int koishi_scene_camera(const koishi_scene *scene, uint32_t index, kr_camera *out);
int koishi_scene_shader_name(const koishi_scene *scene, uint32_t index, char *buf, size_t len);

typedef struct koishi_render_options {
    const char *tracer;   /* NULL: the scene's; overrides the JSON like `cr -t` */
    const char *kernel;   /* NULL: the scene's; overrides like `cr -k`         */
    uint64_t seed;        /* RNG key                                            */
    const char *dump_raw; /* NULL or path of the raw FP64 image dump            */
    int device;           /* HIP device ordinal                                 */
    uint32_t bvh_width;   /* 0 = default                                        */
} koishi_render_options;

/* The whole `cr` flow: load, render w x h at spp, write the 8-bit PNG
 * (clamped, no gamma; src/util/image.hpp:29-34). */
int koishi_render_file(const char *scene_path, const char *png_path, uint32_t width, uint32_t height,
                       uint32_t spp, const koishi_render_options *options);

/* Writes rgb / valid as the reference's 8-bit PNG. */
int koishi_write_png(const char *path, uint32_t width, uint32_t height, const double *rgb_sum,
                     const uint32_t *valid);

#ifdef __cplusplus
}
#endif

#endif /* KOISHI_H */
