// Analytic primitives of a scene on the device: spheres, cubes and planes in
// the object-index space after the triangles (trace.glsl:22-46), tested per
// ray in index order after the triangle walk, with the reference's per-type
// accept rules (primitive.glsl:173-237, :286-333).
//
// The triangle walk leaves the state the reference's triangle loop leaves
// (the (t, id)-lexicographic minimum over triangles), and the reference
// tests spheres, cubes and planes after every triangle, so continuing its
// scan from that state reproduces the reference's result exactly, including
// its order-dependent tie rules: a sphere or plane at the same t as the
// current best loses (hitT < hit.hitT), a cube wins (its reject is
// tmin > hit.hitT, :302).
//
// Every statement is the reference's, in FP32 without contraction (the
// translation units that include this are built with -ffp-contract=off):
// dot(x,y) = (x.x*y.x + x.y*y.y) + x.z*y.z, normalize(v) = v / sqrt(dot(v,v)),
// GLSL cross, and GLSL min(x,y) = y < x ? y : x, max(x,y) = x < y ? y : x.
// asin and atan (sphere uv, :198-199) are the Cephes single-precision
// polynomials (S. L. Moshier, Cephes Math Library asinf.c / atanf.c) in
// FP32, so uv is reproducible on any IEEE machine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace rtx {
namespace dev {

struct AnaRay {
    float ox, oy, oz, dx, dy, dz;
};

__device__ __forceinline__ float ana_dot(float ax, float ay, float az, float bx, float by, float bz) {
    return (ax * bx + ay * by) + az * bz;
}

__device__ __forceinline__ void ana_normalize(float &x, float &y, float &z) {
    const float l = sqrtf(ana_dot(x, y, z, x, y, z));
    x = x / l;
    y = y / l;
    z = z / l;
}

__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }

constexpr float kPio2f = 1.5707963267948966192f;
constexpr float kPio4f = 0.7853981633974483096f;

__device__ __forceinline__ float ana_asinf(float xx) {
    float a, sign;
    if (xx < 0.f) {
        sign = -1.f;
        a = -xx;
    } else {
        sign = 1.f;
        a = xx;
    }
    if (a > 1.f) return __builtin_nanf("");
    float z;
    if (a < 1.0e-4f) {
        z = a;
    } else {
        float x;
        bool flag;
        if (a > 0.5f) {
            z = 0.5f * (1.f - a);
            x = sqrtf(z);
            flag = true;
        } else {
            x = a;
            z = x * x;
            flag = false;
        }
        float p = 4.2163199048e-2f * z + 2.4181311049e-2f;
        p = p * z + 4.5470025998e-2f;
        p = p * z + 7.4953002686e-2f;
        p = p * z + 1.6666752422e-1f;
        z = ((p * z) * x) + x;
        if (flag) {
            z = z + z;
            z = kPio2f - z;
        }
    }
    return sign < 0.f ? -z : z;
}

__device__ __forceinline__ float ana_atanf(float xx) {
    float x, y, sign;
    if (xx < 0.f) {
        sign = -1.f;
        x = -xx;
    } else {
        sign = 1.f;
        x = xx;
    }
    if (x > 2.414213562373095f) {
        y = kPio2f;
        x = -(1.f / x);
    } else if (x > 0.4142135623730950f) {
        y = kPio4f;
        x = (x - 1.f) / (x + 1.f);
    } else {
        y = 0.f;
    }
    const float z = x * x;
    float p = 8.05374449538e-2f * z - 1.38776856032e-1f;
    p = p * z + 1.99777106478e-1f;
    p = p * z - 3.33329491539e-1f;
    y = y + (((p * z) * x) + x);
    return sign < 0.f ? -y : y;
}

// rayIntersectSphere, primitive.glsl:173-210.  s = centre.xyz, radius.
__device__ __forceinline__ bool sphere_test(const AnaRay &r, const float4 s, uint32_t obj, uint32_t prev,
                                            float &best_t, float &u, float &v, float &nx, float &ny, float &nz) {
    const float difx = s.x - r.ox, dify = s.y - r.oy, difz = s.z - r.oz;  // :175
    const float t = ana_dot(difx, dify, difz, r.dx, r.dy, r.dz);          // :176
    const float qx = difx - t * r.dx, qy = dify - t * r.dy, qz = difz - t * r.dz;  // :178
    const float q2 = ana_dot(qx, qy, qz, qx, qy, qz);                     // :179
    const float r2 = s.w * s.w;                                           // :180
    const bool out = q2 > r2;                                             // :182
    const float hit_t = t - sqrtf(r2 - q2);                               // :183
    if (!(!out && obj != prev && hit_t >= 0.f && hit_t < best_t)) return false;  // :185
    best_t = hit_t;                                                       // :187
    const float px = hit_t * r.dx + r.ox, py = hit_t * r.dy + r.oy, pz = hit_t * r.dz + r.oz;  // :189
    nx = s.x - px;                                                        // :190
    ny = s.y - py;
    nz = s.z - pz;
    ana_normalize(nx, ny, nz);
    const float lat = ana_asinf(nz);                                      // :198
    float lon = ana_atanf(ny / nx);                                       // :199
    if (lon != lon) lon = 0.f;                                            // :201-202
    const float c = 0.636619746685f * 0.5f;                               // :204
    u = lat * c + 0.5f;
    v = lon * c + 0.5f;
    return true;
}

// rayIntersectPlane, primitive.glsl:212-237.  p = normal.xyz, offset.
__device__ __forceinline__ bool plane_test(const AnaRay &r, const float4 p, uint32_t obj, uint32_t prev,
                                           float &best_t, float &u, float &v, float &nx, float &ny, float &nz) {
    float dx = p.x, dy = p.y, dz = p.z;
    ana_normalize(dx, dy, dz);                                            // :214
    const float ndx = -dx, ndy = -dy, ndz = -dz;
    const float dif = ana_dot(r.dx, r.dy, r.dz, ndx, ndy, ndz);           // :215
    const float hit_t = -(ana_dot(r.ox, r.oy, r.oz, ndx, ndy, ndz) + p.w) / dif;  // :219
    if (!(hit_t >= 0.f && obj != prev && hit_t < best_t)) return false;   // :221
    best_t = hit_t;                                                       // :223
    nx = dif > 0.f ? ndx : dx;                                            // :224
    ny = dif > 0.f ? ndy : dy;
    nz = dif > 0.f ? ndz : dz;
    const float hx = hit_t * r.dx + r.ox, hy = hit_t * r.dy + r.oy, hz = hit_t * r.dz + r.oz;  // :226
    // cross(p.xyz, (0,0,1)) and cross(p.xyz, (1,0,0)), GLSL component order (:228-229).
    const float zx = 0.f, zy = 0.f, zz = 1.f, xx = 1.f, xy = 0.f, xz = 0.f;
    const float pxx = p.y * zz - zy * p.z, pxy = p.z * zx - zz * p.x, pxz = p.x * zy - zx * p.y;
    const float pzx = p.y * xz - xy * p.z, pzy = p.z * xx - xz * p.x, pzz = p.x * xy - xx * p.y;
    u = ana_dot(hx, hy, hz, pxx, pxy, pxz);                               // :231
    v = ana_dot(hx, hy, hz, pzx, pzy, pzz);
    return true;
}

// rayIntersectCube, primitive.glsl:286-333.  c = start.xyz, end.xyz.
__device__ __forceinline__ bool cube_test(const AnaRay &r, const float *c, uint32_t obj, uint32_t prev, float &best_t,
                                          float &u, float &v, float &nx, float &ny, float &nz) {
    const float rx = 1.f / r.dx, ry = 1.f / r.dy, rz = 1.f / r.dz;       // :288
    const float sx = (c[0] - r.ox) * rx, sy = (c[1] - r.oy) * ry, sz = (c[2] - r.oz) * rz;  // :293
    const float ex = (c[3] - r.ox) * rx, ey = (c[4] - r.oy) * ry, ez = (c[5] - r.oz) * rz;  // :294
    const float mix = gmin(sx, ex), miy = gmin(sy, ey), miz = gmin(sz, ez);  // :296
    const float max_ = gmax(sx, ex), may = gmax(sy, ey), maz = gmax(sz, ez);  // :297
    const float tmin = gmax(gmax(mix, miy), miz);                         // :299
    const float tmax = gmin(gmin(max_, may), maz);                        // :300
    if (tmax < 0.f || tmin > tmax || tmin > best_t || obj == prev) return false;  // :302
    const float px = ((r.dx * tmin + r.ox) - c[0]) / c[3];                // :308-309
    const float py = ((r.dy * tmin + r.oy) - c[1]) / c[4];
    const float pz = ((r.dz * tmin + r.oz) - c[2]) / c[5];
    // :311-326: the face is the first axis whose entry distance is tmin; the
    // normal's sign says whether that entry was the start plane.  Written as
    // selects so no output is addressed through a pointer.
    const int axis = tmin == mix ? 0 : (tmin == miy ? 1 : 2);
    const bool start_face = axis == 0 ? mix == sx : (axis == 1 ? miy == sy : miz == sz);
    const float sgn = start_face ? 1.f : -1.f;
    nx = axis == 0 ? sgn : 0.f;
    ny = axis == 1 ? sgn : 0.f;
    nz = axis == 2 ? sgn : 0.f;
    u = axis == 0 ? py : px;
    v = axis == 2 ? py : pz;
    best_t = tmin;                                                        // :331
    return true;
}

// The analytic loops of traceGeometry / traceOcclusion (trace.glsl:30-46,
// :82-95), continuing from the state the triangle loop left.  Object ids
// continue after `base` triangles: spheres, then cubes, then planes.  The
// loop bounds and the primitive addresses are the same for every lane, so
// the loads are wave-uniform.
__device__ __forceinline__ void analytic_scan(const AnaDev &a, uint32_t base, const AnaRay &r, uint32_t prev,
                                              float &best_t, uint32_t &best_id, float &bu, float &bv) {
    uint32_t j = base;
    float u, v, nx, ny, nz;
    for (uint32_t i = 0; i < a.num_spheres; ++i, ++j)
        if (sphere_test(r, a.spheres[i], j, prev, best_t, u, v, nx, ny, nz)) {
            best_id = j;
            bu = u;
            bv = v;
        }
    for (uint32_t i = 0; i < a.num_cubes; ++i, ++j)
        if (cube_test(r, a.cubes + 6 * static_cast<size_t>(i), j, prev, best_t, u, v, nx, ny, nz)) {
            best_id = j;
            bu = u;
            bv = v;
        }
    for (uint32_t i = 0; i < a.num_planes; ++i, ++j)
        if (plane_test(r, a.planes[i], j, prev, best_t, u, v, nx, ny, nz)) {
            best_id = j;
            bu = u;
            bv = v;
        }
}

// Geometry normal of analytic object k (0-based after the triangles) that
// the ray hit: its test re-run with nothing to beat sets the normal from the
// same operations as when it won the scan.
__device__ __forceinline__ void analytic_normal(const AnaDev &a, uint32_t k, const AnaRay &r, float &nx, float &ny,
                                                float &nz) {
    float t = __builtin_inff(), u, v;
    nx = ny = nz = 0.f;
    if (k < a.num_spheres) {
        sphere_test(r, a.spheres[k], 0u, RT_MISS, t, u, v, nx, ny, nz);
        return;
    }
    k -= a.num_spheres;
    if (k < a.num_cubes) {
        cube_test(r, a.cubes + 6 * static_cast<size_t>(k), 0u, RT_MISS, t, u, v, nx, ny, nz);
        return;
    }
    k -= a.num_cubes;
    if (k < a.num_planes) plane_test(r, a.planes[k], 0u, RT_MISS, t, u, v, nx, ny, nz);
}

}  // namespace dev
}  // namespace rtx
