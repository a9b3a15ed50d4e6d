// The two callers of the query path, as kernels:
//   * raygen: one primary ray per pixel by calculatePrimary (screen,
//     omnidirectional and stereo projections, camera.glsl:49-138), closest
//     hit (triangles, then the scene's
//     analytic primitives), and the reference's packing
//     of the result into the dirObject / uvNormal images (raygen.comp:29-51)
//     with the shading normal of the hit (trace.glsl:48-57: spheremap-packed
//     vertex normals decoded and interpolated, primitive.glsl:95-104,
//     146-151; encodeNormal :84-87), or the geometric normal
//     (primitive.glsl:281) for scenes without vertex normals;
//   * shadow tiles: occlusion of one ray per (sample, pixel) with the
//     results as wave64 ballots, one 64-bit word per 16x4 pixel tile per
//     sample at index tile.x + tile.y * tilesX + sample * tilesX * tilesY
//     and bit (x & 15) | (y & 3) << 4 (nv_all.shadow.comp:23-28,128-141,
//     indexToLight light.glsl:221-229).
// Both map one wave to one 16x4 pixel tile, so a wave's rays are coherent.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include "callers.hpp"
#include "traverse.hpp"

namespace rtx {
namespace {

using namespace dev;

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return (ax * bx + ay * by) + az * bz;
}

// v / |v| with |v| = sqrt(dot(v, v)).
__device__ __forceinline__ void normalize3(float &x, float &y, float &z) {
    const float l = sqrtf(dot3(x, y, z, x, y, z));
    x = x / l;
    y = y / l;
    z = z / l;
}

__device__ __forceinline__ float half_lo(uint32_t v) { return __half2float(__ushort_as_half(static_cast<unsigned short>(v & 0xFFFFu))); }
__device__ __forceinline__ float half_hi(uint32_t v) { return __half2float(__ushort_as_half(static_cast<unsigned short>(v >> 16))); }

// decodeSpheremap, primitive.glsl:95-104.
__device__ __forceinline__ void decode_spheremap(uint32_t n, float &x, float &y, float &z) {
    float nx = half_lo(n), ny = half_hi(n);
    const float l = dot3(nx, ny, 1.f, -nx, -ny, 1.f);  // dot(vec3(nn,1), -vec3(nn,-1))
    const float r = sqrtf(l);
    nx = nx * r;
    ny = ny * r;
    x = nx * 2.f + 0.f;
    y = ny * 2.f + 0.f;
    z = l * 2.f + -1.f;
}

// encodeNormal, primitive.glsl:84-87.
__device__ __forceinline__ uint2 encode_normal(float x, float y, float z) {
    normalize3(x, y, z);
    const uint32_t hx = static_cast<uint32_t>((x * 0.5f + 0.5f) * 65535.f);
    const uint32_t hy = static_cast<uint32_t>((y * 0.5f + 0.5f) * 65535.f);
    const uint32_t hz = static_cast<uint32_t>((z * 0.5f + 0.5f) * 65535.f);
    return make_uint2((hx << 16) | hy, hz);
}

// GLSL sin/cos are driver-defined; here they are the FP32 rounding of the
// FP64 sine/cosine of the FP32 argument (a correctly rounded sinf/cosf
// up to ties of the double rounding), which the oracle evaluates the same
// way.
__device__ __forceinline__ float sin_r(float x) { return static_cast<float>(sin(static_cast<double>(x))); }
__device__ __forceinline__ float cos_r(float x) { return static_cast<float>(cos(static_cast<double>(x))); }
__device__ __forceinline__ float fract_f(float x) { return x - floorf(x); }

// rand1(vec2), rand_util.glsl:123-125: fract(sin(dot(co, (12.9898, 78.233))) * 43758.5453).
__device__ __forceinline__ float rand1(float px, float py) {
    return fract_f(sin_r(px * 12.9898f + py * 78.233f) * 43758.5453f);
}

// calculateScreen, camera.glsl:72-87 (left screen p0..p2, right p3..p5).
__device__ __forceinline__ void screen_ray(const CameraDev &c, float cx, float cy, bool right, float org[3],
                                           float dir[3]) {
    const float *p0 = right ? c.p3 : c.p0, *p1 = right ? c.p4 : c.p1, *p2 = right ? c.p5 : c.p2;
    for (int k = 0; k < 3; ++k) {
        const float r = p1[k] - p0[k], u = p2[k] - p0[k];
        dir[k] = (p0[k] + cx * r) + cy * u - c.eye[k];
        org[k] = c.eye[k];
    }
    normalize3(dir[0], dir[1], dir[2]);
}

// calculateOmni, camera.glsl:49-70.
__device__ __forceinline__ void omni_ray(const CameraDev &c, float cx, float cy, bool left, float org[3],
                                         float dir[3]) {
    constexpr float kPi = 3.1415927410125732421875f;  // rand_util.glsl:13
    const float sx = (cx - 0.5f) * (2.f * kPi), sy = (0.5f - cy) * kPi;
    const float sinx = sin_r(sx), siny = sin_r(sy), cosx = cos_r(sx), cosy = cos_r(sy);
    const float e = c.ipd * 5e-4f;
    const float sg = left ? -1.f : 1.f;
    org[0] = c.eye[0] + (cosx * e) * sg;
    org[1] = c.eye[1] + (0.f * e) * sg;
    org[2] = c.eye[2] + (sinx * e) * sg;
    dir[0] = sinx * cosy;
    dir[1] = siny;
    dir[2] = -cosx * cosy;
}

// calculatePrimary, camera.glsl:113-138: the pixel's offset, the flipped y,
// then the projection.
__device__ __forceinline__ void primary_ray(const CameraDev &c, uint32_t x, uint32_t y, float org[3],
                                            float dir[3]) {
    float jx = c.jitter[0], jy = c.jitter[1];
    if (c.rand_jitter) {
        // rand(loc + randLoc), rand_util.glsl:131-133: (rand1(p), rand1(p * 1103515245 + 12345)).
        const float px = static_cast<float>(x) + c.seed[0], py = static_cast<float>(y) + c.seed[1];
        jx = rand1(px, py);
        jy = rand1(px * 1103515245.f + 12345.f, py * 1103515245.f + 12345.f);
    }
    const float cx = (static_cast<float>(x) + jx) * c.inv_res[0];
    float cy = (static_cast<float>(y) + jy) * c.inv_res[1];
    cy = 1.f - cy;
    switch (c.projection) {
        case RT_PROJ_OMNI: omni_ray(c, cx, cy, false, org, dir); break;
        case RT_PROJ_OMNI_STEREO_TB: omni_ray(c, cx, fract_f(cy * 2.f), cy < 0.5f, org, dir); break;
        case RT_PROJ_OMNI_STEREO_LR: omni_ray(c, fract_f(cx * 2.f), cy, cx < 0.5f, org, dir); break;
        case RT_PROJ_STEREO_TB:
            if (cy < 0.5f) screen_ray(c, cx, cy * 2.f, false, org, dir);
            else screen_ray(c, cx, cy * 2.f - 1.f, true, org, dir);
            break;
        case RT_PROJ_STEREO_LR:
            if (cx < 0.5f) screen_ray(c, cx * 2.f, cy, false, org, dir);
            else screen_ray(c, cx * 2.f - 1.f, cy, true, org, dir);
            break;
        default: screen_ray(c, cx, cy, false, org, dir); break;
    }
}

template <int W, int S>
__global__ void __launch_bounds__(256) k_raygen(const DevScene s, const CallerLaunch a) {
    using TR = Tracer<W, 0, false, S, 256>;
    __shared__ typename TR::Entry lds_stack[S * 256];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t total_waves = gridDim.x * 4u;
    const uint32_t wave = blockIdx.x * 4u + threadIdx.x / 64u;
    const uint32_t tiles_x = (a.width + 15u) / 16u, tiles_y = (a.height + 3u) / 4u;
    const uint32_t ntiles = tiles_x * tiles_y;
    TR t;
    init_stack(t, lds_stack, a.gstack, a.gstack_depth);
    const CameraDev &c = a.cam;
    for (uint32_t tile = wave; tile < ntiles; tile += total_waves) {
        const uint32_t x = (tile % tiles_x) * 16u + (lane & 15u);
        const uint32_t y = (tile / tiles_x) * 4u + (lane >> 4);
        if (x >= a.width || y >= a.height) continue;  // raygen.comp:22
        float org[3], dir[3];
        primary_ray(c, x, y, org, dir);  // raygen.comp:29
        const float dx = dir[0], dy = dir[1], dz = dir[2];
        t.begin_ray(RayD{org[0], org[1], org[2], dx, dy, dz}, kNoHit, RT_MISS, s.root);  // raygen.comp:31
        t.run(s.nodes, s.tris);
        if (s.ana.count) t.analytic(s.ana, s.num_tris);
        if (t.overflow) atomicOr(a.err, 1u);
        const size_t p = static_cast<size_t>(y) * a.width + x;
        const bool hit = t.best_t < t.tlim;
        // raygen.comp:39-44
        a.dir_object[p] = hit ? make_float4(dx * t.best_t, dy * t.best_t, dz * t.best_t, __uint_as_float(t.best_id))
                              : make_float4(dx, dy, dz, __uint_as_float(RT_MISS));
        if (!hit) continue;  // uvNormal is written on hits only (raygen.comp:49-50)
        float nx, ny, nz;
        if (t.best_id >= s.num_tris) {
            // trace.glsl:58-59: an analytic object's normal is its geometry normal.
            analytic_normal(s.ana, t.best_id - s.num_tris, AnaRay{org[0], org[1], org[2], dx, dy, dz}, nx, ny, nz);
        } else if (a.normals) {
            // trace.glsl:52-56: interpolate(n0, n1, n2, uv) = bary.x*n1 + bary.y*n2 + bary.z*n0
            float n[3][3];
            for (int k = 0; k < 3; ++k)
                decode_spheremap(a.normals[3 * static_cast<size_t>(t.best_id) + k], n[k][0], n[k][1], n[k][2]);
            const float bz = (1.f - t.bu) - t.bv;
            nx = (t.bu * n[1][0] + t.bv * n[2][0]) + bz * n[0][0];
            ny = (t.bu * n[1][1] + t.bv * n[2][1]) + bz * n[0][1];
            nz = (t.bu * n[1][2] + t.bv * n[2][2]) + bz * n[0][2];
        } else {
            // primitive.glsl:281: cross(normalize(p1 - p0), normalize(p2 - p0)) * -sign(a)
            const float4 *tp = s.tris + static_cast<size_t>(t.best_rec) * 3;
            const float4 B = tp[1], C = tp[2];
            const float hx = dy * C.z - C.y * dz, hy = dz * C.x - C.z * dx, hz = dx * C.y - C.x * dy;
            const float av = dot3(B.x, B.y, B.z, hx, hy, hz);
            const float sg = av > 0.f ? -1.f : (av < 0.f ? 1.f : -0.f);
            float e1x = B.x, e1y = B.y, e1z = B.z, e2x = C.x, e2y = C.y, e2z = C.z;
            normalize3(e1x, e1y, e1z);
            normalize3(e2x, e2y, e2z);
            nx = (e1y * e2z - e2y * e1z) * sg;
            ny = (e1z * e2x - e2z * e1x) * sg;
            nz = (e1x * e2y - e2x * e1y) * sg;
        }
        const uint2 enc = encode_normal(nx, ny, nz);
        a.uv_normal[p] = make_float4(t.bu, t.bv, __uint_as_float(enc.x), __uint_as_float(enc.y));
    }
}

// ---- nv_all.shadow.comp's ray setup -----------------------------------------

__device__ __forceinline__ float acos_r(float x) { return static_cast<float>(acos(static_cast<double>(x))); }
// GLSL max/min as the specification defines them (NaN and signed zeros
// included): max(x, y) = x < y ? y : x, min(x, y) = y < x ? y : x.
__device__ __forceinline__ float glsl_max(float x, float y) { return x < y ? y : x; }
__device__ __forceinline__ float glsl_min(float x, float y) { return y < x ? y : x; }

// rand(vec2), rand_util.glsl:131-133.
__device__ __forceinline__ void rand2(float px, float py, float &rx, float &ry) {
    rx = rand1(px, py);
    ry = rand1(px * 1103515245.f + 12345.f, py * 1103515245.f + 12345.f);
}

// hammersley(i, n), rand_util.glsl:96-111: (i / n, radical inverse of i).
__device__ __forceinline__ void hammersley(uint32_t i, uint32_t n, float &hx, float &hy) {
    uint32_t v = (i << 16u) | (i >> 16u);
    v = ((v & 0x55555555u) << 1u) | ((v & 0xAAAAAAAAu) >> 1u);
    v = ((v & 0x33333333u) << 2u) | ((v & 0xCCCCCCCCu) >> 2u);
    v = ((v & 0x0F0F0F0Fu) << 4u) | ((v & 0xF0F0F0F0u) >> 4u);
    v = ((v & 0x00FF00FFu) << 8u) | ((v & 0xFF00FF00u) >> 8u);
    hx = static_cast<float>(i) / static_cast<float>(n);
    hy = static_cast<float>(v) * 2.3283064365386963e-10f;
}

__device__ __forceinline__ void cross3(const float *x, const float *y, float *o) {
    o[0] = x[1] * y[2] - y[1] * x[2];
    o[1] = x[2] * y[0] - y[2] * x[0];
    o[2] = x[0] * y[1] - y[0] * x[1];
}

// getDirToLight, light.glsl:98-133: the unit direction from a light sample
// to `pos`, and dist (the distance to a point light's centre, -1 for a
// directional light).
__device__ __forceinline__ void dir_to_light(const LightDev &L, const float pos[3], float rx, float ry, float l[3],
                                             float &dist) {
    constexpr float kPi = 3.1415927410125732421875f;
    dist = -1.f;
    float ro_x = glsl_max(half_lo(L.rad_origin), 0.f), ro_y = glsl_max(half_hi(L.rad_origin), 0.f);
    ro_y = glsl_min(ro_y, ro_x);
    if ((L.color_type[1] >> 16) == 2u) {  // LightType_Point
        for (int k = 0; k < 3; ++k) l[k] = pos[k] - L.pos[k];
        dist = sqrtf(dot3(l[0], l[1], l[2], l[0], l[1], l[2]));
        // randomPointOnHemisphere -> randomPointOnUnitSphere (rand_util.glsl:32-51);
        // its `n *= 1` flip is a no-op.
        const float px = (2.f * kPi) * rx, py = acos_r(1.f - 2.f * ry);
        const float spx = sin_r(px), spy = sin_r(py), cpx = cos_r(px), cpy = cos_r(py);
        const float n[3] = {spx * cpy, spx * spy, cpx};
        for (int k = 0; k < 3; ++k) l[k] = pos[k] - (L.pos[k] + n[k] * ro_y);
    } else {
        // getSunDirection(random, normalize(decodeNormal(dir)), radius), rand_util.glsl:55-86.
        float d[3] = {((static_cast<float>(L.dir[0] >> 16) / 65535.f) * 2.f) - 1.f,
                      ((static_cast<float>(L.dir[0] & 65535u) / 65535.f) * 2.f) - 1.f,
                      ((static_cast<float>(L.dir[1]) / 65535.f) * 2.f) - 1.f};
        normalize3(d[0], d[1], d[2]);
        const float h = cos_r(ro_x);
        const float phi = (2.f * kPi) * rx;
        const float z = h + (1.f - h) * ry;
        const float sin_t = sqrtf(1.f - z * z);
        const float x = cos_r(phi) * sin_t, y = sin_r(phi) * sin_t;
        const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
        const uint32_t xm = (ax - ay < 0.f && ax - az < 0.f) ? 1u : 0u;
        const uint32_t ym = ay - az < 0.f ? (1u ^ xm) : 0u;
        const uint32_t zm = 1u ^ (xm | ym);
        const float m[3] = {static_cast<float>(xm), static_cast<float>(ym), static_cast<float>(zm)};
        float b[3], t[3];
        cross3(d, m, b);
        cross3(b, d, t);
        for (int k = 0; k < 3; ++k) l[k] = (b[k] * x + t[k] * y) + d[k] * z;
    }
    normalize3(l[0], l[1], l[2]);
}

// nv_all.shadow.comp:50-143 as one kernel: per 16x4 tile and sample, each
// lane reads its pixel's dirObject, builds the shadow ray toward a sample of
// the light and traces it with prevHit = the pixel's object; the tile's
// ballot is its word.  A tile whose pixels all missed is written as 0 (the
// reference leaves it unwritten outside DEBUG builds, :69-82).
template <int W, int S>
__global__ void __launch_bounds__(256) k_shadow_pass(const DevScene s, const CallerLaunch a) {
    using TR = Tracer<W, 1, false, S, 256>;
    __shared__ typename TR::Entry lds_stack[S * 256];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t total_waves = gridDim.x * 4u;
    const uint32_t wave = blockIdx.x * 4u + threadIdx.x / 64u;
    const uint32_t tiles_x = (a.width + 15u) / 16u, tiles_y = (a.height + 3u) / 4u;
    const uint32_t per_sample = tiles_x * tiles_y;
    const uint32_t nwords = per_sample * a.samples;
    TR t;
    init_stack(t, lds_stack, a.gstack, a.gstack_depth);
    const CameraDev &c = a.cam;
    for (uint32_t word = wave; word < nwords; word += total_waves) {
        const uint32_t sample = word / per_sample, tile = word % per_sample;
        const uint32_t x = (tile % tiles_x) * 16u + (lane & 15u);
        const uint32_t y = (tile / tiles_x) * 4u + (lane >> 4);
        bool occluded = false;
        if (x < a.width && y < a.height) {  // :57
            const float4 dob = a.dir_in[static_cast<size_t>(y) * a.width + x];  // :64-65
            const uint32_t object = __float_as_uint(dob.w);
            if (object != RT_MISS) {
                const float pos[3] = {c.eye[0] + dob.x, c.eye[1] + dob.y, c.eye[2] + dob.z};  // :84
                // :88-94: uv over 128x128 tiles plus the sample's Hammersley point
                float jx, jy;
                rand2(static_cast<float>(x) + c.seed[0], static_cast<float>(y) + c.seed[1], jx, jy);
                float ux = (static_cast<float>(x) + jx) / 128.f, uy = (static_cast<float>(y) + jy) / 128.f;
                float hx, hy;
                hammersley(sample, a.samples, hx, hy);
                ux = ux + hx;
                uy = uy + hy;
                float rx, ry;
                rand2(ux, uy, rx, ry);
                float l[3], dist;
                dir_to_light(a.light, pos, rx, ry, l, dist);
                float tmax = -1.f;  // no trace
                if (dist >= 0.f) {
                    // :104-117: sphere lights trace only between origin and radius
                    const float rad = half_lo(a.light.rad_origin), org = half_hi(a.light.rad_origin);
                    if (dist >= org && dist < rad) tmax = dist - org;
                } else {
                    tmax = kNoHit;  // :121-125 directional
                }
                if (tmax >= 0.f) {
                    t.begin_ray(RayD{pos[0], pos[1], pos[2], -l[0], -l[1], -l[2]}, tmax, object, s.root);
                    t.run(s.nodes, s.tris);
                    if (s.ana.count) t.analytic(s.ana, s.num_tris);
                    if (t.overflow) atomicOr(a.err, 1u);
                    occluded = t.occluded;
                }
            }
        }
        const unsigned long long m = __ballot(occluded);  // :128
        if (lane == 0) a.tiles[word] = m;                 // :133-140
    }
}

template <int W, int S>
__global__ void __launch_bounds__(256) k_shadow_tiles(const DevScene s, const CallerLaunch a) {
    using TR = Tracer<W, 1, false, S, 256>;
    __shared__ typename TR::Entry lds_stack[S * 256];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t total_waves = gridDim.x * 4u;
    const uint32_t wave = blockIdx.x * 4u + threadIdx.x / 64u;
    const uint32_t tiles_x = (a.width + 15u) / 16u, tiles_y = (a.height + 3u) / 4u;
    const uint32_t per_sample = tiles_x * tiles_y;
    const uint32_t nwords = per_sample * a.samples;
    TR t;
    init_stack(t, lds_stack, a.gstack, a.gstack_depth);
    for (uint32_t word = wave; word < nwords; word += total_waves) {
        const uint32_t sample = word / per_sample, tile = word % per_sample;
        const uint32_t x = (tile % tiles_x) * 16u + (lane & 15u);
        const uint32_t y = (tile / tiles_x) * 4u + (lane >> 4);
        bool occluded = false;
        if (x < a.width && y < a.height) {  // nv_all.shadow.comp:57
            const size_t i = (static_cast<size_t>(sample) * a.height + y) * a.width + x;
            t.begin(load_ray(a.rays + i), s.root);
            t.run(s.nodes, s.tris);
            if (s.ana.count) t.analytic(s.ana, s.num_tris);
            if (t.overflow) atomicOr(a.err, 1u);
            occluded = t.occluded;
        }
        const unsigned long long m = __ballot(occluded);  // nv_all.shadow.comp:128
        if (lane == 0) a.tiles[word] = m;                 // :133-140
    }
}

template <int W>
CallerFn pick_w(CallerKind k) {
    if (k == CallerKind::Raygen) return &k_raygen<W, 16>;
    if (k == CallerKind::ShadowPass) return &k_shadow_pass<W, 8>;
    return &k_shadow_tiles<W, 8>;
}

}  // namespace

CallerFn caller_kernel(uint32_t width, CallerKind k) {
    switch (width) {
        case 2: return pick_w<2>(k);
        case 4: return pick_w<4>(k);
        case 8: return pick_w<8>(k);
        default: return nullptr;
    }
}

uint32_t caller_lds_stack(CallerKind k) { return k == CallerKind::Raygen ? 16u : 8u; }  // shadow kinds: 8

int launch_caller(uint32_t width, CallerKind k, const DevScene &s, const CallerLaunch &a, uint32_t blocks,
                  hipStream_t stream) {
    CallerFn fn = caller_kernel(width, k);
    if (!fn) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(256), 0, stream, s, a);
    return hipGetLastError();
}

}  // namespace rtx
