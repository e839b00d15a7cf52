/* malevich_camera.c -- the oracle's per-view constant buffer: the reference's camera set-up
 * (main.c:1431-1441) and update() with no input (main.c:1480-1562), with m4x4f32_mul_m4x4f32
 * (math.h:188-197), m4x4f32_determinant (math.h:282-293) and m4x4f32_inverse (math.h:295-320) in
 * their operation order. Test infrastructure only. */
#include <math.h>
#include <stdint.h>

#include "malevich_oracle.h"

static const float PI_F = 3.141592654f, TAU_F = 6.283185307f, PI_OVER_TWO_F = 1.570796326f; /* math.h:34-36 */

static const signed char DET[24][5] = {
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
static const signed char INV[16][6][4] = {
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

static float dot_seq(const float *a, float b0, float b1, float b2, float b3) {
    return ((a[0] * b0 + a[1] * b1) + a[2] * b2) + a[3] * b3; /* math.h:137-140 */
}

static void mul4(const float *a, const float *b, float *c) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) c[4 * i + j] = dot_seq(a + 4 * i, b[j], b[4 + j], b[8 + j], b[12 + j]);
}

float orc_det4(const float m[16]) {
    float v = 0.0f;
    for (int k = 0; k < 24; ++k) {
        float t = ((m[DET[k][1]] * m[DET[k][2]]) * m[DET[k][3]]) * m[DET[k][4]];
        v = k == 0 ? t : (DET[k][0] > 0 ? v + t : v - t);
    }
    return v;
}

void orc_inv4(const float m[16], float r[16]) {
    float scale = (float)(1.0 / (double)orc_det4(m));
    for (int e = 0; e < 16; ++e) {
        float v = 0.0f;
        for (int k = 0; k < 6; ++k) {
            const signed char *t = INV[e][k];
            float p = (m[t[1]] * m[t[2]]) * m[t[3]];
            v = k == 0 ? p : (t[0] > 0 ? v + p : v - p);
        }
        r[e] = v * scale;
    }
}

void orc_camera_cb(const orc_camera *c, float out[48]) {
    float fov_rad = c->fov_y_deg * (PI_F / 180.0f);
    float aspect = (float)c->width / (float)c->height;
    float sy = (float)(1.0 / tan((double)fov_rad / 2.0));
    float sx = sy / aspect;
    float clip_from_view[16] = {sx, 0, 0, 0, 0, sy, 0, 0, 0, 0, 0, c->near_plane, 0, 0, 1, 0};
    float yaw = c->yaw_rad;
    if (yaw > PI_F) yaw -= TAU_F;
    else if (yaw <= -PI_F) yaw += TAU_F;
    float pitch = PI_OVER_TWO_F < c->pitch_rad ? PI_OVER_TWO_F : c->pitch_rad; /* min(PI_OVER_TWO, pitch) */
    pitch = -PI_OVER_TWO_F > pitch ? -PI_OVER_TWO_F : pitch;                  /* max(-PI_OVER_TWO, pitch) */
    float cp = (float)cos(-(double)pitch), sp = (float)sin(-(double)pitch);
    float cy = (float)cos(-(double)yaw), syw = (float)sin(-(double)yaw);
    float rot_pitch[16] = {1, 0, 0, 0, 0, cp, sp, 0, 0, -sp, cp, 0, 0, 0, 0, 1};
    float rot_yaw[16] = {cy, 0, -syw, 0, 0, 1, 0, 0, syw, 0, cy, 0, 0, 0, 0, 1};
    float cob[16] = {0, 0, -1, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1};
    float t[16], wfv[16], vfw[16];
    mul4(rot_yaw, rot_pitch, t);
    mul4(cob, t, wfv);
    wfv[3] = c->position[0];
    wfv[7] = c->position[1];
    wfv[11] = c->position[2];
    orc_inv4(wfv, vfw);
    mul4(clip_from_view, vfw, out);
    orc_inv4(clip_from_view, out + 16);
    for (int e = 0; e < 16; ++e) out[32 + e] = wfv[e];
}
