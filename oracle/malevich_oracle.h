/*
 * malevich_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Scalar CPU restatement of the Malevich draw pipeline (reference:
 * source/main.c, source/math.h, source/common_shader_core.h and the seven
 * shader files). It is the checker for the HIP library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. It is
 * never the product path.
 *
 * Semantics follow the reference source with IEEE single precision, no FP
 * contraction (-ffp-contract=off) and two's-complement wrap for i32
 * (-fwrapv). Deviations, each because the reference's behaviour is not
 * defined by its source:
 *   - _mm256_rsqrt_ps (math.h:277-280) is an x86 approximation; the oracle
 *     uses 1.0f/sqrtf(x). Parity unpinned at that point.
 *   - SVML _mm256_acos_ps/_exp_ps/_pow_ps (math.h:20-23) are replaced by
 *     libm acosf/expf/powf. Parity unpinned at that point; RGBA8 results
 *     are compared within 1/255.
 *   - Primitive assembly runs in input order (the reference's OpenMP atomic
 *     slot at main.c:874-875 makes the order thread-dependent; single-thread
 *     order is the one reproduced).
 *   - Unwritten pad lanes of the VS output (basic_vs.c:13 etc.) are 0.
 */
#ifndef MALEVICH_ORACLE_H
#define MALEVICH_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_VS_BASIC = 0, ORC_VS_VERTEX_LIGHTING = 1, ORC_VS_PASSTHROUGH = 2, ORC_VS_FULLSCREEN = 3 };
/* ORC_PS_BASIC_LAMBERT and ORC_PS_ENV_TEXTURED restate the library's extension shaders (the
 * reference has no mips, lights or second texture): mip level by footprint, Lambert light,
 * albedo texture. ORC_PS_NONE restates its depth-only pass: a fragment that passes early-Z
 * writes depth and visibility, and no colour. */
enum { ORC_PS_BASIC = 0, ORC_PS_ENV_LIGHTING = 1, ORC_PS_PASSTHROUGH = 2, ORC_PS_BASIC_LAMBERT = 3,
       ORC_PS_ENV_TEXTURED = 4, ORC_PS_NONE = 5 };
enum { ORC_FMT_RGBA8 = 0, ORC_FMT_RGBA32F = 1 };
enum { ORC_DEPTH_GREATER_EQUAL = 0, ORC_DEPTH_LESS_EQUAL = 1 };

enum {
    ORC_OK = 0,
    ORC_ERR_INVALID_ARG = -1,
    ORC_ERR_CAPACITY = -2,
    ORC_ERR_NOMEM = -3,
};

typedef struct orc_texture {
    const void *data; /* RGBA8 as u32 (R in the low byte) or float4; `levels` levels, level 0 first */
    uint32_t width, height, format;
    uint32_t levels;  /* 0 or 1: one level; level l is max(w >> l, 1) x max(h >> l, 1) */
} orc_texture;

typedef struct orc_viewport {
    float top_left_x, top_left_y, width, height, min_depth, max_depth;
} orc_viewport;

typedef struct orc_draw_desc {
    const uint32_t *index_buffer;
    uint32_t index_count, start_index;
    int32_t base_vertex;
    const void *vertex_buffer;
    uint32_t vertex_stride; /* bytes */
    uint32_t vertex_count;
    int32_t vs, ps;
    const float *cb; /* 48 floats: clip_from_world, view_from_clip, world_from_view (row-major) */
    orc_texture vs_srv0, ps_srv0;
    orc_viewport viewport;
    int32_t depth_func;
    orc_texture ps_srv1; /* extension shaders: slot 1 */
    const float *ps_cb;  /* extension shaders: 12 floats */
    int32_t cull_mode;   /* extension: ORC_CULL_BACK (main.c:856), ORC_CULL_NONE or ORC_CULL_FRONT */
    int32_t fill_rule;   /* extension: ORC_FILL_INCLUSIVE (main.c:1015-1028) or ORC_FILL_TOP_LEFT */
} orc_draw_desc;
enum { ORC_CULL_BACK = 0, ORC_CULL_NONE = 1, ORC_CULL_FRONT = 2 };
enum { ORC_FILL_INCLUSIVE = 0, ORC_FILL_TOP_LEFT = 1 };

typedef struct orc_target {
    uint32_t *colors;
    float *depth;
    float *hiz;          /* one float per 8x8 bin, (width/8)*(height/8) */
    uint32_t *vis_tri;   /* optional: assembled-triangle id of the last writer, 0xFFFFFFFF = none */
    uint32_t *vis_draw;  /* optional: draw counter of the last writer */
    uint32_t width, height;
    uint32_t draw_counter;
} orc_target;

typedef struct orc_stats {
    uint64_t vertex_count, input_triangle_count, assembled_triangle_count;
    uint64_t active_bin_count, total_triangle_count_in_bins, oob_triangle_count;
} orc_stats;

/* One assembled triangle as produced by primitive assembly (main.c:127-139). */
typedef struct orc_triangle {
    int32_t edge[3][3];   /* {a,b,c} for edge slots 0,1,2 */
    float rw[3];
    float one_over_area;
    float max_depth;
    int32_t min_bounds[2], max_bounds[2];
    uint32_t prim;        /* input triangle index within the draw */
    float attr[3][12];    /* 3 vertices x 3 registers x 4 floats; reg0 = screen position */
    uint32_t excl;        /* extension, top-left rule: bit k set = edge slot k is neither top nor left */
} orc_triangle;

int orc_clear_render_target_view(orc_target *rt, const float rgba[4]);
int orc_clear_depth_stencil_view(orc_target *rt, float depth, int32_t depth_func);
int orc_draw_indexed(orc_target *rt, const orc_draw_desc *d, orc_stats *stats);

/* Stage-level entry points for unit checks. */
int orc_vertex_shader(const orc_draw_desc *d, uint32_t vertex_index, float out12[12]);
int orc_primitive_assembly(const orc_draw_desc *d, orc_triangle *out, uint32_t capacity,
                           uint32_t *out_count, uint64_t *oob_count);
uint64_t orc_coverage_mask(const int32_t edge[3][3], int32_t origin_x, int32_t origin_y);
/* extension: the same under the top-left rule (a corner on an edge of slot k in excl is not covered) */
uint64_t orc_coverage_mask_excl(const int32_t edge[3][3], uint32_t excl, int32_t origin_x, int32_t origin_y);
void orc_bilinear(const orc_texture *t, float u, float v, float out[4]);
uint32_t orc_nearest_level(float rho2, uint32_t levels); /* extension: level k when rho2 >= 2^(2k-1) */
void orc_bilinear_level(const orc_texture *t, uint32_t level, float u, float v, float out[4]);
void orc_latlon(const orc_texture *t, const float dir[3], float out[4]);
uint32_t orc_encode_output(const float rgb[3]);
uint32_t orc_encode_clear(const float rgba[4]);
void orc_srgb_to_linear_rgba8(const uint32_t *in, uint32_t *out, size_t n); /* main.c:546-558 */

/* camera state (main.c:1422-1430) -> PerFrameCB (main.c:1480-1562); malevich_camera.c */
typedef struct orc_camera {
    float position[3];
    float yaw_rad, pitch_rad, fov_y_deg, near_plane;
    uint32_t width, height;
} orc_camera;
void orc_camera_cb(const orc_camera *c, float out[48]);
float orc_det4(const float m[16]);
void orc_inv4(const float m[16], float r[16]);
float orc_srgb_from_linear_approx(float c);
float orc_tonemap(float c);
void orc_normalize3(const float in[3], float out[3]);
int32_t orc_cvt_rne(float f);
void orc_clip_planes(float out[6][4]);
int orc_threads(void); /* OpenMP threads the assembly and bin loops use */
int orc_clip_polygon(const float *in_vertices /* n x 12 */, int n, float *out_vertices /* 16 x 12 */);

/* ---- differentiable shading, FP64 (malevich_diff.c). The f32 copies of positions, uvs and
 * the matrix feed only the discrete mip choice, which is computed bit-identically to the GPU. */
typedef struct orc_diff_scene {
    const double *positions, *colors, *uvs, *normals; /* [V][3], [V][3], [V][2], [V][3] */
    const float *positions_f, *uvs_f;                 /* f32 originals of positions and uvs */
    const uint32_t *indices;                          /* [T][3] */
    uint32_t vertex_count, triangle_count;
    const double *texture;                            /* RGBA mip chain, level 0 first */
    uint32_t tex_width, tex_height, tex_levels;
    const double *sh;                                 /* [9][3] */
    double clip_from_world[16];
    float clip_from_world_f[16];
    uint32_t width, height;
} orc_diff_scene;

typedef struct orc_diff_grads {
    double *positions, *colors, *uvs, *normals, *texture, *sh;
} orc_diff_grads;

/* The real SH basis of bands 0..2 the shading uses, index l(l+1)+m, at a unit direction n. */
void orc_sh9(const double n[3], double out[9]);
int orc_diff_forward(const orc_diff_scene *s, const uint32_t *prim, double *rgba);
int orc_diff_backward(const orc_diff_scene *s, const uint32_t *prim, const double *grad_rgba, orc_diff_grads *g);

#ifdef __cplusplus
}
#endif
#endif
