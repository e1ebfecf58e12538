// fg_quat.hpp -- quaternion and 3x3 helpers with the arithmetic of
// GPU_Kernels/K_MathUtils.cu:57-236 (x,y,z vector part, w scalar part).
// Contraction is off so results match a host evaluation operation by
// operation.
#pragma once
#include <hip/hip_runtime.h>

namespace fg {
#pragma clang fp contract(off)

// K_MathUtils.cu:57-62
__device__ __forceinline__ float4 vec3_cross(float4 a, float4 b) {
    return make_float4((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x), 0.0f);
}

// K_MathUtils.cu:67-70
__device__ __forceinline__ float vec3_dot(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// K_MathUtils.cu:160-169
__device__ __forceinline__ float4 quat_product(float4 q1, float4 q2) {
    float w = q1.w * q2.w - vec3_dot(q1, q2);
    float4 c = vec3_cross(q1, q2);
    return make_float4((q1.w * q2.x) + (q2.w * q1.x) + c.x, (q1.w * q2.y) + (q2.w * q1.y) + c.y,
                       (q1.w * q2.z) + (q2.w * q1.z) + c.z, w);
}

// K_MathUtils.cu:87-90
__device__ __forceinline__ float4 quat_conjugate(float4 q) { return make_float4(-q.x, -q.y, -q.z, q.w); }

// K_MathUtils.cu:175-181: conj(q) (p,0) q
__device__ __forceinline__ float4 quat_transform_point(float4 p, float4 q) {
    p.w = 0.0f;
    return quat_product(quat_product(quat_conjugate(q), p), q);
}

__device__ __forceinline__ float4 quat_add(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// K_HashGrid.cu:52-57: a body sphere in world space, conj(q) (p,0) q + cm,
// keeping the body index in w.
__device__ __forceinline__ float4 world_sphere(float4 rel, float4 q, float4 cm) {
    float4 a = quat_add(quat_transform_point(rel, q), cm);
    a.w = rel.w;
    return a;
}

__device__ __forceinline__ float4 f4_scale(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }

// K_MathUtils.cu:72-82
__device__ __forceinline__ float4 quat_normalize(float4 q) {
    float norm = sqrtf(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    q.x /= norm;
    q.y /= norm;
    q.z /= norm;
    q.w /= norm;
    return q;
}

// K_MathUtils.cu:183-196 (rows r[0..2])
__device__ __forceinline__ void quat_to_matrix(float4 q, float4* r) {
    r[0].x = 1 - (2 * q.y * q.y) - (2 * q.z * q.z);
    r[0].y = (2 * q.x * q.y) + (2 * q.w * q.z);
    r[0].z = (2 * q.x * q.z) - (2 * q.w * q.y);
    r[1].x = (2 * q.x * q.y) - (2 * q.w * q.z);
    r[1].y = 1 - (2 * q.x * q.x) - (2 * q.z * q.z);
    r[1].z = (2 * q.y * q.z) + (2 * q.w * q.x);
    r[2].x = (2 * q.x * q.z) + (2 * q.w * q.y);
    r[2].y = (2 * q.y * q.z) - (2 * q.w * q.x);
    r[2].z = 1 - (2 * q.x * q.x) - (2 * q.y * q.y);
}

// K_MathUtils.cu:20-33
__device__ __forceinline__ void mat3_mult(const float4* a, const float4* b, float4* r) {
    for (int i = 0; i < 3; i++) {
        r[i].x = a[i].x * b[0].x + a[i].y * b[1].x + a[i].z * b[2].x;
        r[i].y = a[i].x * b[0].y + a[i].y * b[1].y + a[i].z * b[2].y;
        r[i].z = a[i].x * b[0].z + a[i].y * b[1].z + a[i].z * b[2].z;
    }
}

// K_MathUtils.cu:201-206
__device__ __forceinline__ void mat3_transpose(const float4* m, float4* r) {
    r[0].x = m[0].x; r[0].y = m[1].x; r[0].z = m[2].x;
    r[1].x = m[0].y; r[1].y = m[1].y; r[1].z = m[2].y;
    r[2].x = m[0].z; r[2].y = m[1].z; r[2].z = m[2].z;
}

// K_MathUtils.cu:211-236: R I^-1 R^T L
__device__ __forceinline__ float4 angular_velocity(const float4* inv_inertia, float4 L, float4 q) {
    float4 R[3], RT[3], T[3], G[3];
    quat_to_matrix(q, R);
    mat3_transpose(R, RT);
    mat3_mult(R, inv_inertia, T);
    mat3_mult(T, RT, G);
    float4 w;
    w.x = G[0].x * L.x + G[0].y * L.y + G[0].z * L.z;
    w.y = G[1].x * L.x + G[1].y * L.y + G[1].z * L.z;
    w.z = G[2].x * L.x + G[2].y * L.y + G[2].z * L.z;
    w.w = 0.0f;
    return w;
}

#pragma clang fp contract(fast)
}  // namespace fg
