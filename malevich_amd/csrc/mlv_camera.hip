// mlv_camera.hip -- per-view constant buffers on the device (gfx950).
//
// The reference builds its PerFrameCB on the host every frame: the camera's projection at start-up
// (main.c:1422-1441), then update() (main.c:1480-1562) wraps yaw, clamps pitch, forms
// world_from_view from the yaw and pitch rotations and the view-to-world change of basis, inverts
// it and the projection with m4x4f32_inverse (math.h:295-320) and multiplies with
// m4x4f32_mul_m4x4f32 (math.h:188-197). Batched views need one such buffer per view; one lane
// builds one, in the reference's f32/f64 operation order so that the matrices are the ones the
// reference would compute for the same camera state.
#include <cmath>

#include "mlv_internal.h"

namespace mlv {

namespace {

constexpr float kPi = 3.141592654f, kTau = 6.283185307f, kPiOverTwo = 1.570796326f;  // math.h:34-36

// Cofactor expansions of the 4x4 determinant and inverse (math.h:282-320) as (sign, entries)
// tables over row-major indices. The terms are summed in this order, each product taken left to
// right, so the roundings are those of the expressions they restate.
__constant__ int8_t kDet[24][5] = {
    {1, 3, 6, 9, 12},
    {-1, 2, 7, 9, 12},
    {-1, 3, 5, 10, 12},
    {1, 1, 7, 10, 12},
    {1, 2, 5, 11, 12},
    {-1, 1, 6, 11, 12},
    {-1, 3, 6, 8, 13},
    {1, 2, 7, 8, 13},
    {1, 3, 4, 10, 13},
    {-1, 0, 7, 10, 13},
    {-1, 2, 4, 11, 13},
    {1, 0, 6, 11, 13},
    {1, 3, 5, 8, 14},
    {-1, 1, 7, 8, 14},
    {-1, 3, 4, 9, 14},
    {1, 0, 7, 9, 14},
    {1, 1, 4, 11, 14},
    {-1, 0, 5, 11, 14},
    {-1, 2, 5, 8, 15},
    {1, 1, 6, 8, 15},
    {1, 2, 4, 9, 15},
    {-1, 0, 6, 9, 15},
    {-1, 1, 4, 10, 15},
    {1, 0, 5, 10, 15}
};
__constant__ int8_t kInv[16][6][4] = {
    {{1, 6, 11, 13}, {-1, 7, 10, 13}, {1, 7, 9, 14}, {-1, 5, 11, 14}, {-1, 6, 9, 15}, {1, 5, 10, 15}},
    {{1, 3, 10, 13}, {-1, 2, 11, 13}, {-1, 3, 9, 14}, {1, 1, 11, 14}, {1, 2, 9, 15}, {-1, 1, 10, 15}},
    {{1, 2, 7, 13}, {-1, 3, 6, 13}, {1, 3, 5, 14}, {-1, 1, 7, 14}, {-1, 2, 5, 15}, {1, 1, 6, 15}},
    {{1, 3, 6, 9}, {-1, 2, 7, 9}, {-1, 3, 5, 10}, {1, 1, 7, 10}, {1, 2, 5, 11}, {-1, 1, 6, 11}},
    {{1, 7, 10, 12}, {-1, 6, 11, 12}, {-1, 7, 8, 14}, {1, 4, 11, 14}, {1, 6, 8, 15}, {-1, 4, 10, 15}},
    {{1, 2, 11, 12}, {-1, 3, 10, 12}, {1, 3, 8, 14}, {-1, 0, 11, 14}, {-1, 2, 8, 15}, {1, 0, 10, 15}},
    {{1, 3, 6, 12}, {-1, 2, 7, 12}, {-1, 3, 4, 14}, {1, 0, 7, 14}, {1, 2, 4, 15}, {-1, 0, 6, 15}},
    {{1, 2, 7, 8}, {-1, 3, 6, 8}, {1, 3, 4, 10}, {-1, 0, 7, 10}, {-1, 2, 4, 11}, {1, 0, 6, 11}},
    {{1, 5, 11, 12}, {-1, 7, 9, 12}, {1, 7, 8, 13}, {-1, 4, 11, 13}, {-1, 5, 8, 15}, {1, 4, 9, 15}},
    {{1, 3, 9, 12}, {-1, 1, 11, 12}, {-1, 3, 8, 13}, {1, 0, 11, 13}, {1, 1, 8, 15}, {-1, 0, 9, 15}},
    {{1, 1, 7, 12}, {-1, 3, 5, 12}, {1, 3, 4, 13}, {-1, 0, 7, 13}, {-1, 1, 4, 15}, {1, 0, 5, 15}},
    {{1, 3, 5, 8}, {-1, 1, 7, 8}, {-1, 3, 4, 9}, {1, 0, 7, 9}, {1, 1, 4, 11}, {-1, 0, 5, 11}},
    {{1, 6, 9, 12}, {-1, 5, 10, 12}, {-1, 6, 8, 13}, {1, 4, 10, 13}, {1, 5, 8, 14}, {-1, 4, 9, 14}},
    {{1, 1, 10, 12}, {-1, 2, 9, 12}, {1, 2, 8, 13}, {-1, 0, 10, 13}, {-1, 1, 8, 14}, {1, 0, 9, 14}},
    {{1, 2, 5, 12}, {-1, 1, 6, 12}, {-1, 2, 4, 13}, {1, 0, 6, 13}, {1, 1, 4, 14}, {-1, 0, 5, 14}},
    {{1, 1, 6, 8}, {-1, 2, 5, 8}, {1, 2, 4, 9}, {-1, 0, 6, 9}, {-1, 1, 4, 10}, {1, 0, 5, 10}}
};

__device__ __forceinline__ float dot4_seq(const float *a, float b0, float b1, float b2, float b3) {
    return ((a[0] * b0 + a[1] * b1) + a[2] * b2) + a[3] * b3;  // v4f32_dot, math.h:137-140
}

// c = a * b (math.h:188-197): entry (i, j) is the sequential dot of row i of a and column j of b
__device__ void mul4(const float *a, const float *b, float *c) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) c[4 * i + j] = dot4_seq(a + 4 * i, b[j], b[4 + j], b[8 + j], b[12 + j]);
}

__device__ float det4(const float *m) {
    float v = 0.0f;
    for (int k = 0; k < 24; ++k) {
        const float t = ((m[kDet[k][1]] * m[kDet[k][2]]) * m[kDet[k][3]]) * m[kDet[k][4]];
        v = k == 0 ? t : (kDet[k][0] > 0 ? v + t : v - t);
    }
    return v;
}

__device__ void inv4(const float *m, float *r) {
    const float scale = (float)(1.0 / (double)det4(m));  // f32 scale = 1.0 / det
    for (int e = 0; e < 16; ++e) {
        float v = 0.0f;
        for (int k = 0; k < 6; ++k) {
            const int8_t *t = kInv[e][k];
            const float p = (m[t[1]] * m[t[2]]) * m[t[3]];
            v = k == 0 ? p : (t[0] > 0 ? v + p : v - p);
        }
        r[e] = v * scale;
    }
}

__global__ __launch_bounds__(64) void k_camera_cbs(const CameraRec *cams, uint32_t n, float *cbs) {
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= n) return;
    const CameraRec c = cams[i];
    // projection (main.c:1431-1441): left-handed reversed-Z infinite
    const float fov_rad = c.fov_y_deg * (kPi / 180.0f);
    const float aspect = (float)c.width / (float)c.height;
    const float sy = (float)(1.0 / tan((double)fov_rad / 2.0));
    const float sx = sy / aspect;
    const float clip_from_view[16] = {sx, 0, 0, 0, 0, sy, 0, 0, 0, 0, 0, c.near_plane, 0, 0, 1, 0};
    // update() with no input: yaw wrapped, pitch clamped (main.c:1496-1506)
    float yaw = c.yaw_rad;
    if (yaw > kPi) yaw -= kTau;
    else if (yaw <= -kPi) yaw += kTau;
    float pitch = kPiOverTwo < c.pitch_rad ? kPiOverTwo : c.pitch_rad;  // min(PI_OVER_TWO, pitch)
    pitch = -kPiOverTwo > pitch ? -kPiOverTwo : pitch;                  // max(-PI_OVER_TWO, pitch)
    const float cp = (float)cos(-(double)pitch), sp = (float)sin(-(double)pitch);
    const float cy = (float)cos(-(double)yaw), syw = (float)sin(-(double)yaw);
    const float rot_pitch[16] = {1, 0, 0, 0, 0, cp, sp, 0, 0, -sp, cp, 0, 0, 0, 0, 1};
    const float rot_yaw[16] = {cy, 0, -syw, 0, 0, 1, 0, 0, syw, 0, cy, 0, 0, 0, 0, 1};
    const float change_of_basis[16] = {0, 0, -1, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1};
    float t[16], world_from_view[16], view_from_world[16];
    mul4(rot_yaw, rot_pitch, t);
    mul4(change_of_basis, t, world_from_view);
    world_from_view[3] = c.position[0];
    world_from_view[7] = c.position[1];
    world_from_view[11] = c.position[2];
    inv4(world_from_view, view_from_world);
    float *out = cbs + 48u * i;
    mul4(clip_from_view, view_from_world, out);  // clip_from_world
    inv4(clip_from_view, out + 16);              // view_from_clip
    for (int e = 0; e < 16; ++e) out[32 + e] = world_from_view[e];
}

}  // namespace

hipError_t launch_camera_cbs(const CameraRec *cams, uint32_t n, float *cbs, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_camera_cbs, dim3((n + 63u) / 64u), dim3(64), 0, s, cams, n, cbs);
    return hipGetLastError();
}

}  // namespace mlv
