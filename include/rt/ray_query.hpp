// ray_query.hpp - header-only C++ mirror of the reference's render-task
// shape over the C ABI (rt_c.h).
//
// The reference's host API for this path is a pair of igx::rt render tasks:
// RaygenTask (include/rt/task/raygen_task.hpp:7-33: prepareCommandList,
// update, resize, switchToScene; dispatch at src/rt/task/raygen_task.cpp:
// 88-94) and ShadowTask (include/rt/task/shadow_task.hpp:54-91, ballot
// buffer sized ceil(W/16) * ceil(H/4) * samples * 8 bytes at
// src/rt/task/shadow_task.cpp:146-165).  These classes keep those method
// names and meanings; "prepareCommandList" records nothing and instead
// launches on a HIP stream, which is how the GPU work is submitted here.
// Errors throw std::runtime_error with the library's message (the reference
// builds with exceptions enabled, CMakeLists.txt:46).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rt_c.h"

namespace igx::rt::hip {

inline void check(int rc, rt_ctx ctx = nullptr) {
    if (rc != RT_OK)
        throw std::runtime_error(std::string("rt: ") + (ctx ? rt_last_error(ctx) : rt_last_error_global()));
}

// Owns a device context (one per GPU).
class Context {
   public:
    explicit Context(int device = 0) { check(rt_ctx_create(device, &ctx_)); }
    ~Context() { rt_ctx_destroy(ctx_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    rt_ctx get() const { return ctx_; }

   private:
    rt_ctx ctx_ = nullptr;
};

// A device-resident scene: the counterpart of the triangles[] + SceneInfo
// bindings a SceneGraph provides (res/shaders/scene.glsl:6-28), plus the BVH.
class Scene {
   public:
    Scene(Context &ctx, const float *tri_xyz, uint32_t n, uint32_t node_width = RT_NODE_BVH4,
          uint32_t builder = RT_BUILD_SAH_CPU)
        : ctx_(ctx.get()) {
        rt_build_opts o;
        rt_build_opts_default(&o);
        o.node_width = node_width;
        o.builder = builder;
        check(rt_scene_create(ctx_, tri_xyz, n, &o, &scene_), ctx_);
    }
    ~Scene() { rt_scene_destroy(scene_); }
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;
    rt_scene get() const { return scene_; }
    rt_ctx ctx() const { return ctx_; }

    // From the reference's own 48-byte Triangle records (primitive.glsl:
    // 28-38); the scene keeps their spheremap normals for RaygenTask.
    static std::unique_ptr<Scene> fromPacked(Context &ctx, const void *triangles, uint32_t n,
                                             uint32_t node_width = RT_NODE_BVH4,
                                             uint32_t builder = RT_BUILD_SAH_CPU) {
        std::unique_ptr<Scene> s(new Scene(ctx.get()));
        rt_build_opts o;
        rt_build_opts_default(&o);
        o.node_width = node_width;
        o.builder = builder;
        check(rt_scene_create_packed(s->ctx_, triangles, n, &o, &s->scene_), s->ctx_);
        return s;
    }

    // The scene's spheres (centre, radius), cubes (start, end) and planes
    // (normal, offset), object ids after the triangles (trace.glsl:22-46).
    void setAnalytic(const float *spheres, uint32_t ns, const float *cubes, uint32_t nc, const float *planes,
                     uint32_t np) {
        check(rt_scene_set_analytic(ctx_, scene_, spheres, ns, cubes, nc, planes, np), ctx_);
    }

   private:
    explicit Scene(rt_ctx ctx) : ctx_(ctx) {}
    rt_ctx ctx_ = nullptr;
    rt_scene scene_ = nullptr;
};

// RaygenTask: one primary ray per pixel, closest hit, the dirObject and
// uvNormal images (raygen.comp:39-51).
class RaygenTask {
   public:
    explicit RaygenTask(Context &ctx) : ctx_(ctx.get()) {}
    ~RaygenTask() { release(); }
    RaygenTask(const RaygenTask &) = delete;
    RaygenTask &operator=(const RaygenTask &) = delete;

    void switchToScene(Scene *scene) { scene_ = scene; }
    void update(double) {}
    // The screen camera with a fixed pixel offset (rt_raygen_closest).
    void setCamera(const rt_camera &cam) {
        cam_ = cam;
        primary_ = false;
    }
    // calculatePrimary in full (camera.glsl:113-138): projections and the
    // per-pixel rand(loc + seed) offset (rt_raygen_primary).  The reference
    // refreshes Seed every frame (raygen_task.cpp update); call it per frame.
    void setPrimaryCamera(const rt_primary_camera &cam) {
        pcam_ = cam;
        primary_ = true;
    }
    void resize(uint32_t width, uint32_t height) {
        release();
        w_ = width;
        h_ = height;
        const size_t bytes = static_cast<size_t>(w_) * h_ * 4 * sizeof(float);
        if (bytes && (hipMalloc(&dir_object_, bytes) != hipSuccess || hipMalloc(&uv_normal_, bytes) != hipSuccess))
            throw std::runtime_error("rt: image allocation failed");
    }
    // Launches the pass on `stream` (the reference records Dispatch(size())).
    void prepareCommandList(hipStream_t stream, const uint32_t *d_normals = nullptr) {
        if (!scene_) throw std::runtime_error("rt: no scene");
        if (primary_)
            check(rt_raygen_primary(ctx_, scene_->get(), &pcam_, w_, h_, d_normals, dir_object_, uv_normal_, stream),
                  ctx_);
        else
            check(rt_raygen_closest(ctx_, scene_->get(), &cam_, w_, h_, d_normals, dir_object_, uv_normal_, stream),
                  ctx_);
    }
    // getTexture(0) / getTexture(1) of the reference: float4 [h][w] images.
    float *getTexture(int i) const { return i == 0 ? dir_object_ : uv_normal_; }

   private:
    void release() {
        if (dir_object_) (void)hipFree(dir_object_);
        if (uv_normal_) (void)hipFree(uv_normal_);
        dir_object_ = uv_normal_ = nullptr;
    }
    rt_ctx ctx_;
    Scene *scene_ = nullptr;
    rt_camera cam_{};
    rt_primary_camera pcam_{};
    bool primary_ = false;
    uint32_t w_ = 0, h_ = 0;
    float *dir_object_ = nullptr, *uv_normal_ = nullptr;
};

// ShadowTask: occlusion of one ray per (sample, pixel) into 64-bit ballots
// per 16x4 tile (nv_all.shadow.comp:128-141); the caller fills the rays.
class ShadowTask {
   public:
    explicit ShadowTask(Context &ctx, uint32_t samples = 2) : ctx_(ctx.get()), samples_(samples) {}
    ~ShadowTask() {
        if (out_) (void)hipFree(out_);
    }
    ShadowTask(const ShadowTask &) = delete;
    ShadowTask &operator=(const ShadowTask &) = delete;

    void switchToScene(Scene *scene) { scene_ = scene; }
    void update(double) {}
    void resize(uint32_t width, uint32_t height) {
        if (out_) (void)hipFree(out_);
        out_ = nullptr;
        w_ = width;
        h_ = height;
        // shadow_task.cpp:146-165: ceil(W/16) * ceil(H/4) * samples * 8 bytes.
        const size_t words = static_cast<size_t>((w_ + 15) / 16) * ((h_ + 3) / 4) * samples_;
        if (words && hipMalloc(&out_, words * sizeof(uint64_t)) != hipSuccess)
            throw std::runtime_error("rt: ballot buffer allocation failed");
    }
    bool needsCommandUpdate() const { return false; }
    void prepareCommandList(const rt_ray *d_rays, hipStream_t stream) {
        if (!scene_) throw std::runtime_error("rt: no scene");
        check(rt_shadow_tiles(ctx_, scene_->get(), d_rays, w_, h_, samples_, out_, stream), ctx_);
    }
    // nv_all.shadow.comp in full: the shadow rays are made from raygen's
    // dirObject image (RaygenTask::getTexture(0)), the light and the frame
    // seed, as shadow_task.cpp:194-215 dispatches it after the raygen pass.
    void prepareCommandList(const RaygenTask &raygen, const float eye[3], const rt_light &light, const float seed[2],
                            hipStream_t stream) {
        if (!scene_) throw std::runtime_error("rt: no scene");
        check(rt_shadow_pass(ctx_, scene_->get(), raygen.getTexture(0), eye, &light, seed, w_, h_, samples_, out_,
                             stream),
              ctx_);
    }
    const uint64_t *getOutput() const { return out_; }
    // Bytes of the ballot buffer resize() sized (shadow_task.cpp:161-165).
    size_t outputBytes() const { return static_cast<size_t>((w_ + 15) / 16) * ((h_ + 3) / 4) * samples_ * 8; }

   private:
    rt_ctx ctx_;
    uint32_t samples_;
    Scene *scene_ = nullptr;
    uint32_t w_ = 0, h_ = 0;
    uint64_t *out_ = nullptr;
};

// The query calls themselves, for callers that bring their own rays:
// traceGeometry (trace.glsl:9) and traceOcclusion (trace.glsl:70).
inline void traceGeometry(const Scene &s, const rt_ray *d_rays, uint32_t n, rt_hit *d_hits, hipStream_t stream) {
    check(rt_trace_closest(s.ctx(), s.get(), d_rays, n, d_hits, stream), s.ctx());
}
inline void traceOcclusion(const Scene &s, const rt_ray *d_rays, uint32_t n, uint8_t *d_flags, hipStream_t stream) {
    check(rt_trace_occluded(s.ctx(), s.get(), d_rays, n, d_flags, stream), s.ctx());
}

}  // namespace igx::rt::hip
