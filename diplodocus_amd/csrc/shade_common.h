// Small FP32 vector helpers shared by the shading kernels (render.hip, whitted.hip), with the
// operation order of the reference's Vec3 (src/util/vec3.h) and RayEpsilon (src/scene/ray.h).
#pragma once

#include <hip/hip_runtime.h>

#include "dpl_internal.h"

namespace dpl {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 Add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 Sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 Mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 Mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 Cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// src/util/vec3.h:109-117
__device__ __forceinline__ V3 Normalize(V3 v) {
    const float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    if (len <= 1e-6f) return {0.0f, 0.0f, 0.0f};
    return {v.x / len, v.y / len, v.z / len};
}
// src/scene/ray.h:15-24
__device__ __forceinline__ float RayEpsilon(V3 p, float t) {
    const float ps = Max2(Max2(Max2(fabsf(p.x), fabsf(p.y)), fabsf(p.z)), 1.0f);
    const float ts = Max2(t, 1.0f);
    const float bias = Max2(1e-6f * ps, 1e-7f * ts);
    return bias < 1e-7f ? 1e-7f : (1e-3f < bias ? 1e-3f : bias);
}

}  // namespace dpl
