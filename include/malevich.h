/*
 * malevich.h -- C ABI of the MI355X (gfx950) draw pipeline.
 *
 * Drop-in for the D3D11-style raster path of Malevich. The reference keeps
 * one global `Pipeline graphics_pipeline` (source/main.c:109-115, global at
 * main.c:222) that the caller writes field by field (main.c:1276-1294), and
 * three commands: clear_render_target_view (main.c:1191),
 * clear_depth_stencil_view (main.c:1204) and draw_indexed (main.c:1219).
 * Here the same state lives in an mlv_context; each state field has a
 * setter, each command an entry point. Shaders are chosen by enum because
 * the reference's shader function pointers (csc.h:10-18) cannot cross to
 * the device; the enum set is the reference's compiled-in shader set
 * (main.c:46-52).
 *
 * Conventions: every call returns an mlv_status (0 = ok) and never aborts.
 * Commands are stream-ordered on the context's HIP stream; a readback or
 * mlv_synchronize waits for them. One host thread per context.
 */
#ifndef MALEVICH_H
#define MALEVICH_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define MLV_API __attribute__((visibility("default")))
#else
#define MLV_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlv_context mlv_context;
typedef struct mlv_buffer mlv_buffer;
typedef struct mlv_texture2d mlv_texture2d;
typedef struct mlv_render_target mlv_render_target;
typedef struct mlv_depth_stencil mlv_depth_stencil;

typedef enum mlv_status {
    MLV_OK = 0,
    MLV_ERROR_INVALID_ARGUMENT = -1,
    MLV_ERROR_OUT_OF_MEMORY = -2,
    MLV_ERROR_DEVICE = -3,
    MLV_ERROR_INVALID_STATE = -4,
    MLV_ERROR_CAPACITY = -5,
    MLV_ERROR_NO_DEVICE = -6
} mlv_status;

typedef enum mlv_buffer_kind {
    MLV_BUFFER_VERTEX = 1,
    MLV_BUFFER_INDEX_U32 = 2, /* IA.p_index_buffer is u32 (main.c:72) */
    MLV_BUFFER_CONSTANT = 3
} mlv_buffer_kind;

typedef enum mlv_format {
    MLV_FORMAT_RGBA8_UNORM = 1, /* Texture2D u32 texels, R in the low byte (csc.h:30-36, math.h:336-344) */
    MLV_FORMAT_RGBA32F = 2,     /* Texture2D float4 texels (csc.h:42-53) */
    MLV_FORMAT_RGBA16F = 3      /* extension: half4 texels, widened to float on fetch */
} mlv_format;

typedef enum mlv_topology {
    MLV_TOPOLOGY_UNDEFINED = 0,    /* main.c:67 */
    MLV_TOPOLOGY_TRIANGLELIST = 1  /* main.c:68 */
} mlv_topology;

/* The reference's vertex shaders (main.c:46-52): basic_vs.c, vertex_lighting_vs.c,
 * passthrough_vs.c, fullscreen_vs.c. */
typedef enum mlv_vertex_shader {
    MLV_VS_BASIC = 1,
    MLV_VS_VERTEX_LIGHTING = 2,
    MLV_VS_PASSTHROUGH = 3,
    MLV_VS_FULLSCREEN = 4
} mlv_vertex_shader;

/* The reference's pixel shaders: basic_ps.c, env_lighting_ps.c, passthrough_ps.c; and two
 * extensions that sample mip chains on the level nearest the pixel's footprint (the reference's
 * Texture2D has one level, csc.h:20-24):
 *   MLV_PS_BASIC_LAMBERT  basic_ps (basic_ps.c:16-25) times a directional light: texture (slot 0)
 *                         at UV * (light_colour * max(dot(normalize(NORMAL), L), 0) + ambient),
 *                         then basic_ps's sRGB curve. PS constant buffer slot 0, 12 floats:
 *                         L.xyz _, light_colour.rgb _, ambient.rgb _ (L points at the light).
 *   MLV_PS_ENV_TEXTURED   env_lighting_ps (env_lighting_ps.c:13-22) times an albedo texture: the
 *                         float equirect map (slot 0) at normalize(NORMAL) times the texture
 *                         (slot 1) at UV, then env_lighting_ps's tone curve.
 * The level: rho^2 = the larger squared distance, in level-0 texels, between the sampled
 * coordinate at the pixel corner and at the next corners right and below (for the equirect map:
 * the chord between the unit directions times width / 2 pi); level k is taken when
 * rho^2 >= 2^(2k-1). */
typedef enum mlv_pixel_shader {
    MLV_PS_BASIC = 1,
    MLV_PS_ENV_LIGHTING = 2,
    MLV_PS_PASSTHROUGH = 3,
    MLV_PS_BASIC_LAMBERT = 4, /* extension */
    MLV_PS_ENV_TEXTURED = 5,  /* extension */
    MLV_PS_NONE = 6           /* extension: no pixel shader (a depth-only pass, D3D's null pixel
                                 shader): depth, hierarchical Z and the visibility buffers are
                                 written as by any shader, the colour target is left as it is, and
                                 the vertex stage keeps only the position (register 0) */
} mlv_pixel_shader;

typedef enum mlv_depth_func {
    MLV_DEPTH_GREATER_EQUAL = 0, /* reference: reversed Z, _CMP_GE_OQ (main.c:1166) */
    MLV_DEPTH_LESS_EQUAL = 1     /* extension: mirrored test and hierarchical-Z */
} mlv_depth_func;

typedef enum mlv_cull_mode {
    MLV_CULL_BACK = 0,  /* reference: a triangle of signed area > 0 is culled (main.c:856) */
    MLV_CULL_NONE = 1,  /* extension: both windings are drawn */
    MLV_CULL_FRONT = 2  /* extension: a triangle of signed area < 0 is culled */
} mlv_cull_mode;

typedef enum mlv_fill_rule {
    MLV_FILL_INCLUSIVE = 0, /* reference: a pixel corner on any edge is covered (main.c:1015-1028) */
    MLV_FILL_TOP_LEFT = 1   /* extension, D3D: a corner on an edge is covered only by a top or left edge */
} mlv_fill_rule;

/* Viewport, field for field (main.c:85-92). */
typedef struct mlv_viewport {
    float top_left_x;
    float top_left_y;
    float width;
    float height;
    float min_depth;
    float max_depth;
} mlv_viewport;

/* Stats (main.c:199-206), accumulated per draw like main.c:1228-1246. */
typedef struct mlv_stats {
    uint64_t vertex_count;
    uint64_t input_triangle_count;
    uint64_t assembled_triangle_count;
    uint64_t active_bin_count;             /* non-empty 8x8 bins, summed over draws */
    uint64_t total_triangle_count_in_bins; /* (triangle, 8x8 bin) pairs, summed over draws */
    uint64_t oob_triangle_count;           /* triangles dropped for an index past the vertex buffer */
    uint64_t draw_count;
    uint64_t overflow_count;               /* draws that ran the slow path after a bin-list overflow */
} mlv_stats;

/* ---- lifecycle (no reference counterpart: the reference state is a global, main.c:222) */
MLV_API int mlv_context_create(int device, mlv_context **out);
MLV_API int mlv_context_destroy(mlv_context *ctx);
MLV_API const char *mlv_get_last_error(const mlv_context *ctx);
MLV_API int mlv_synchronize(mlv_context *ctx);
MLV_API void *mlv_context_stream(mlv_context *ctx); /* the hipStream_t commands are queued on */
MLV_API int mlv_device_count(int *out);

/* ---- resources. Replaces the caller-owned arrays the reference binds by pointer:
 * Mesh VB/IB (main.c:60-64), PerFrameCB (main.c:169-173), Texture2D (csc.h:20-24),
 * frame_buffer / depth_buffer (main.c:35-36). */
MLV_API int mlv_buffer_create(mlv_context *ctx, int kind, size_t bytes, const void *host_init, mlv_buffer **out);
MLV_API int mlv_buffer_wrap(mlv_context *ctx, int kind, void *device_ptr, size_t bytes, mlv_buffer **out);
MLV_API int mlv_buffer_update(mlv_context *ctx, mlv_buffer *buf, size_t offset, size_t bytes, const void *host_src);
MLV_API int mlv_buffer_readback(mlv_context *ctx, const mlv_buffer *buf, size_t offset, size_t bytes, void *host_dst);
MLV_API int mlv_buffer_destroy(mlv_context *ctx, mlv_buffer *buf);
MLV_API void *mlv_buffer_device_ptr(const mlv_buffer *buf);

/* mip_levels (0 means 1) levels are allocated, level 0 first, level l of max(width >> l, 1) x
 * max(height >> l, 1) texels, at most down to 1x1; host_data, if given, holds all of them. */
MLV_API int mlv_texture2d_create(mlv_context *ctx, uint32_t width, uint32_t height, int format, uint32_t mip_levels,
                                 const void *host_data, mlv_texture2d **out);
/* Extension: fills levels 1.. of the texture's chain from level 0 on the device, each texel the
 * average of its 2x2 window of the level above (2x1 or 1x2 once a side is 1; an odd side drops its
 * last row or column): RGBA8 per byte in integers rounding halves up, the float formats summed in
 * row order in f32 and divided by the window's area (RGBA16F rounded to half, nearest-even). */
MLV_API int mlv_texture2d_generate_mips(mlv_context *ctx, mlv_texture2d *tex);
MLV_API int mlv_texture2d_destroy(mlv_context *ctx, mlv_texture2d *tex);
/* load_texture(..., is_in_srgb = true) (main.c:538-559): the in-place sRGB -> linear pre-pass of
 * an RGBA8 texture (all four channels, truncating re-encode; every level), run on the device. */
MLV_API int mlv_texture2d_srgb_to_linear(mlv_context *ctx, mlv_texture2d *tex);
/* Copies the texture's texels, every level (level 0 first), to the host. */
MLV_API int mlv_texture2d_readback(mlv_context *ctx, const mlv_texture2d *tex, void *host_dst);
MLV_API void *mlv_texture2d_device_ptr(const mlv_texture2d *tex);

MLV_API int mlv_render_target_create(mlv_context *ctx, uint32_t width, uint32_t height, mlv_render_target **out);
MLV_API int mlv_render_target_destroy(mlv_context *ctx, mlv_render_target *rt);
MLV_API void *mlv_render_target_device_ptr(const mlv_render_target *rt);
MLV_API int mlv_depth_stencil_create(mlv_context *ctx, uint32_t width, uint32_t height, mlv_depth_stencil **out);
MLV_API int mlv_depth_stencil_destroy(mlv_context *ctx, mlv_depth_stencil *ds);
MLV_API void *mlv_depth_stencil_device_ptr(const mlv_depth_stencil *ds);

/* ---- IA (main.c:71-76) */
MLV_API int mlv_ia_set_index_buffer(mlv_context *ctx, mlv_buffer *ib, uint32_t offset_bytes);
/* stride = IA.input_layout, the per-vertex byte size (main.c:74, main.c:1286) */
MLV_API int mlv_ia_set_vertex_buffer(mlv_context *ctx, mlv_buffer *vb, uint32_t stride, uint32_t offset_bytes);
MLV_API int mlv_ia_set_primitive_topology(mlv_context *ctx, int topology);

/* ---- VS (main.c:78-83; shader and register count set per draw at main.c:1287-1288) */
MLV_API int mlv_vs_set_shader(mlv_context *ctx, int vertex_shader);
MLV_API int mlv_vs_set_constant_buffer(mlv_context *ctx, uint32_t slot, mlv_buffer *cb);
MLV_API int mlv_vs_set_shader_resource(mlv_context *ctx, uint32_t slot, mlv_texture2d *tex);

/* ---- RS (main.c:94-96) */
MLV_API int mlv_rs_set_viewport(mlv_context *ctx, const mlv_viewport *vp);
/* Extension: the reference hard-codes back-face culling (main.c:856). Kept triangles of either
 * winding are set up alike (edge functions oriented by the area's sign, main.c:563-575). */
MLV_API int mlv_rs_set_cull(mlv_context *ctx, int cull_mode);
/* Extension: the reference's coverage is inclusive and leaves the top-left rule a TODO
 * (main.c:1027). Under MLV_FILL_TOP_LEFT an edge whose inward normal (a, b) has a > 0 (left) or
 * a == 0, b > 0 (top; screen y grows downward) keeps the corners on it, any other edge drops
 * them, so a mesh's shared edges cover each corner once. Depth and barycentrics are unchanged. */
MLV_API int mlv_rs_set_fill_rule(mlv_context *ctx, int fill_rule);

/* ---- PS (main.c:98-101) */
MLV_API int mlv_ps_set_shader(mlv_context *ctx, int pixel_shader);
MLV_API int mlv_ps_set_shader_resource(mlv_context *ctx, uint32_t slot, mlv_texture2d *tex);
/* Extension: the reference's PS stage has shader resources only (main.c:98-101); the extension
 * shaders read constant buffer slot 0. */
MLV_API int mlv_ps_set_constant_buffer(mlv_context *ctx, uint32_t slot, mlv_buffer *cb);

/* ---- OM (main.c:103-107) */
MLV_API int mlv_om_set_render_targets(mlv_context *ctx, mlv_render_target *rt, mlv_depth_stencil *ds);
MLV_API int mlv_om_set_depth_func(mlv_context *ctx, int depth_func);
/* Extension: per-pixel visibility (assembled-triangle id and draw tag of the last writer).
 * Two device arrays of width*height u32 each, caller-owned; NULL disables. */
MLV_API int mlv_om_set_visibility_buffer(mlv_context *ctx, uint32_t *device_tri_ids, uint32_t *device_draw_tags);
/* Extension: per-pixel primitive id (the input triangle index within its draw, counted from
 * start_index, like SV_PrimitiveID) of the visible triangle; 0xFFFFFFFF where nothing was drawn
 * since the last clear of the bound render target. One device array of width*height u32,
 * caller-owned; NULL disables. The differentiable shading entry points below read it. */
MLV_API int mlv_om_set_primitive_id_buffer(mlv_context *ctx, uint32_t *device_prim_ids);

/* ---- commands */
/* main.c:1191 clear_render_target_view(const f32 *p_clear_color) */
MLV_API int mlv_clear_render_target_view(mlv_context *ctx, mlv_render_target *rt, const float rgba[4]);
/* main.c:1204 clear_depth_stencil_view(const f32 depth); also resets the 8x8 tile minima (main.c:1212-1214) */
MLV_API int mlv_clear_depth_stencil_view(mlv_context *ctx, mlv_depth_stencil *ds, float depth);
/* main.c:1219 draw_indexed(UINT index_count); start_index/base_vertex are the D3D11 arguments
 * the reference leaves as a TODO (main.c:1219). (n, 0, 0) is the reference call. */
MLV_API int mlv_draw_indexed(mlv_context *ctx, uint32_t index_count, uint32_t start_index, int32_t base_vertex);
/* Copies the render target's u32 values (row-major) to the host: the values the reference's
 * frame_buffer holds (main.c:35), shaded pixels R<<16|G<<8|B (main.c:1176-1178), clears
 * A<<24|B<<16|G<<8|R (math.h:322-324). */
MLV_API int mlv_render_target_readback(mlv_context *ctx, mlv_render_target *rt, void *host_dst);
/* main.c:1301 present() / paint_window (main.c:286-311), headless: the image the window shows.
 * The frame buffer is read as a top-down 32-bit DIB with masks R 0x00FF0000, G 0x0000FF00,
 * B 0x000000FF and written to host_rgb as width*height*3 bytes, row-major, top row first (a
 * PPM/PNG-ready RGB8 image). The conversion runs on the device, so a quarter fewer bytes cross PCIe. */
MLV_API int mlv_present(mlv_context *ctx, mlv_render_target *rt, uint8_t *host_rgb);
MLV_API int mlv_depth_stencil_readback(mlv_context *ctx, mlv_depth_stencil *ds, float *host_dst);

/* Recorded draws run together when their results are needed (readback, synchronize,
 * a clear of a target they draw into, a buffer/texture/target update or destroy).
 * mlv_flush enqueues them now without waiting. No reference counterpart: the
 * reference executes each draw_indexed to completion (main.c:1219-1261). */
MLV_API int mlv_flush(mlv_context *ctx);

/* Extension: per-stage GPU time measured with HIP events on the context stream.
 * Stages: 0 vertex, 1 setup (cull/clip/ordered compaction), 2 binning, 3 tile
 * (raster + early-Z + shading + output merger). Timing adds two event records per
 * stage per flush; mlv_get_profile synchronizes, returns totals since the last call
 * and restarts them. */
typedef struct mlv_profile {
    double stage_ms[4];
    uint64_t draws;
    uint64_t flushes;
    double tile_kernel_ms; /* the raster/shade kernel of stage 3 alone (kb_tile) */
} mlv_profile;
MLV_API int mlv_context_set_profiling(mlv_context *ctx, int enable);
MLV_API int mlv_get_profile(mlv_context *ctx, mlv_profile *out);

/* Extension: footprint counting, for measurement (no reference counterpart). While enabled, the
 * tile stage marks every assembled triangle that wins a pixel and counts the pixels written, and
 * the shaders (and the differentiable passes) mark every texel their bilinear footprints read, in
 * one bitmap per texture (allocated when a counted draw first samples it). mlv_get_footprint
 * flushes, returns the totals since its last call and clears them: visible triangles (summed over
 * flushes), pixels written, distinct texels per mip level and their bytes. Counting adds atomics
 * to the shading: leave it off when timing. */
typedef struct mlv_footprint {
    uint64_t visible_triangles;
    uint64_t shaded_pixels;
    uint64_t texels;
    uint64_t texel_bytes;
} mlv_footprint;
MLV_API int mlv_context_set_footprint_counting(mlv_context *ctx, int enable);
MLV_API int mlv_get_footprint(mlv_context *ctx, mlv_footprint *out);

/* Stats (main.c:199-206), reset like render() does at main.c:1268-1269. The reset also
 * restarts the draw tags written to the visibility buffer at 0. */
MLV_API int mlv_get_stats(mlv_context *ctx, mlv_stats *out);
MLV_API int mlv_reset_stats(mlv_context *ctx);

/* ---- per-view constant buffers (extension of the reference's update(), main.c:1480-1562)
 *
 * The camera state the reference keeps (main.c:1422-1430) for each view; width/height give the
 * aspect as (float)width / height. mlv_camera_constant_buffers writes one PerFrameCB per camera
 * (48 floats: clip_from_world, view_from_clip, world_from_view; main.c:169-173, 1557-1559) to
 * device memory, computed on the device as update() does with no input: yaw wrapped into
 * (-pi, pi], pitch clamped to [-pi/2, pi/2], the yaw and pitch rotations, the view-to-world change
 * of basis, the position, and the reference's 4x4 inverse (math.h:295-320) and products
 * (math.h:188-197) in their operation order. Stream-ordered before later draws. */
typedef struct mlv_camera {
    float position[3];
    float yaw_rad, pitch_rad;
    float fov_y_deg, near_plane;
    uint32_t width, height;
} mlv_camera;
MLV_API int mlv_camera_constant_buffers(mlv_context *ctx, const mlv_camera *cameras, uint32_t count,
                                        float *device_cbs);

/* ---- graphs (extension; no reference counterpart: the reference runs every draw_indexed to
 * completion, main.c:1219-1261)
 *
 * Between mlv_graph_capture_begin and mlv_graph_capture_end, clears and draws are recorded
 * instead of run, with the state they read resolved as they are issued; only state setters,
 * clears, draws, mlv_diff_batch_forward / mlv_diff_batch_backward and mlv_diff_l2_grad may be
 * called meanwhile (anything that flushes returns MLV_ERROR_INVALID_STATE; a recorded batched
 * pass runs after the draws recorded before it). capture_end runs the recorded commands once (sizing every scratch
 * arena; this is the frame's first execution) and then captures the same commands into a HIP
 * graph. mlv_graph_launch replays the whole frame as one graph launch on the context stream:
 * the same targets, buffers and textures, whose contents the caller may change between
 * launches (a new camera through mlv_buffer_update, new vertex data). Resources a graph names
 * must outlive it; destroy graphs before their context. A replay whose geometry outgrows what
 * was sized at capture reports MLV_ERROR_CAPACITY at the next synchronising call: capture
 * again. */
typedef struct mlv_graph mlv_graph;
MLV_API int mlv_graph_capture_begin(mlv_context *ctx);
MLV_API int mlv_graph_capture_end(mlv_context *ctx, mlv_graph **out);
MLV_API int mlv_graph_launch(mlv_context *ctx, mlv_graph *graph);
MLV_API int mlv_graph_destroy(mlv_context *ctx, mlv_graph *graph);

/* ---- launch tuning (extension; no reference counterpart; the reference's tile is fixed at
 * 8x8 bins, main.c:24-26)
 *
 * The per-context launch parameters of the setup, binning and tile stages. mlv_autotune_begin /
 * mlv_autotune_end record a frame exactly as a graph capture does (only state, clears and
 * draws in between), then run it repeatedly under candidate parameters (coordinate descent, in
 * the order of the fields below: skip_cornerless, tile_edge, tile_stage, tile_waves_per_cu,
 * setup_waves_per_cu, heavy_tile_pairs, heavy_blocks, pair_blocks). Each candidate is timed as
 * the fastest of `repeats` runs (HIP events on the context stream, after one warm-up run) and
 * replaces the incumbent only if it is more than 1% faster. Results (frames and stats) are the
 * same under every candidate; the targets hold the recorded frame afterwards. The bin-list
 * capacity follows the largest pair count measured (with a quarter of headroom) and is reported
 * in pair_capacity. */
typedef struct mlv_tuning {
    uint32_t heavy_tile_pairs; /* a 16x16 tile with a longer (triangle, tile) list is walked by 16 waves */
    uint32_t heavy_blocks;     /* blocks (1024 threads) of the persistent heavy-tile kernel */
    uint32_t pair_blocks;      /* blocks (256 threads) of the pair count and emit kernels */
    uint32_t tile_stage;       /* bin capacity: records of a tile's list held in LDS at a time, 128 or 256 */
    uint32_t pair_capacity;    /* out: bin-list capacity (pairs) of the last flush; ignored by set */
    uint32_t tile_edge;        /* screen tile edge of the binning, 16 or 32 pixels (one wave per 8x8 bin) */
    uint32_t tile_waves_per_cu;  /* 0, or about this many waves of the tile kernel per CU at most (an LDS cap) */
    uint32_t setup_waves_per_cu; /* likewise for the setup kernels */
    uint32_t skip_cornerless;  /* 1: an assembled triangle whose snapped bounding box holds no pixel corner
                                  covers nothing; it is counted in the stats but not binned */
} mlv_tuning;
MLV_API int mlv_get_tuning(mlv_context *ctx, mlv_tuning *out);
MLV_API int mlv_set_tuning(mlv_context *ctx, const mlv_tuning *tuning);
MLV_API int mlv_autotune_begin(mlv_context *ctx);
MLV_API int mlv_autotune_end(mlv_context *ctx, uint32_t repeats, mlv_tuning *best);

/* ---- differentiable shading (extension; the reference renders forward only)
 *
 * Shading of a mesh with per-vertex colours, UVs and normals, an RGBA32F mipmapped texture and
 * a 9-coefficient RGB spherical-harmonics light, on top of the visibility the raster pipeline
 * produced (mlv_om_set_primitive_id_buffer). Per covered pixel (x, y), sampled at the integer
 * corner the coverage test uses:
 *   lambda   perspective-correct barycentrics of the pixel in the visible input triangle,
 *            from the clip-space positions clip = clip_from_world * (position, 1)
 *   albedo   bilinear fetch (basic_ps's filter, csc.h:143-161) of the texture at uv on the mip
 *            level nearest the pixel's uv footprint
 *   E        sum_k sh[k] * Y_k(normalize(n)), real SH basis of bands 0..2
 *   out.rgb  srgb(albedo * colour * E) with basic_ps's curve max(1.055 c^(1/2.4) - 0.055, 0)
 *            (math.h:419-422) of max(c, 0); out.a = 1. Uncovered pixels are (0, 0, 0, 0).
 * The backward pass recomputes this from the primitive ids (nothing else is kept from the
 * forward pass) and returns the gradient of sum(out * grad_out) with respect to positions,
 * colours, UVs, normals, texels (every mip level) and SH coefficients, at fixed visibility and
 * fixed mip choice. Position gradients flow through the barycentrics only. */
typedef struct mlv_diff_scene {
    const float *positions;    /* device, [vertex_count][3], world space */
    const float *colors;       /* device, [vertex_count][3] */
    const float *uvs;          /* device, [vertex_count][2] */
    const float *normals;      /* device, [vertex_count][3] */
    const uint32_t *indices;   /* device, [triangle_count][3] */
    uint32_t vertex_count, triangle_count;
    const float *texture;      /* device RGBA32F mip chain, level 0 first, level l is max(w>>l,1) x max(h>>l,1) */
    uint32_t tex_width, tex_height, tex_levels;
    const float *sh;           /* device, [9][3] */
    float clip_from_world[16]; /* row-major, the first matrix of basic_vs's constant buffer (basic_vs.c:16-20) */
    uint32_t width, height;    /* image size; the viewport covers it */
} mlv_diff_scene;

typedef struct mlv_diff_grads {
    float *positions, *colors, *uvs, *normals; /* device, shapes as in mlv_diff_scene; NULL skips */
    float *texture;                            /* device, the whole mip chain */
    float *sh;                                 /* device, [9][3] */
} mlv_diff_grads;

/* The mip chain the differentiable pass samples, built on the device: level 0 is the texture
 * ((height, width) RGBA32F), each next level its 2x2 box filter (2x1 or 1x2 once a side is 1; an
 * odd side drops its last row or column), all levels concatenated. mlv_mip_chain_adjoint maps a
 * gradient on the whole chain to the level-0 texture. Extension: no counterpart in the reference,
 * whose textures are single-level (load_texture, main.c:538-560); they replace the per-level
 * framework calls a host would otherwise make twice per training step. Both run on the context's
 * stream. */
MLV_API int mlv_mip_chain_size(uint32_t width, uint32_t height, uint32_t *texels, uint32_t *levels);
MLV_API int mlv_mip_chain(mlv_context *ctx, const float *device_texture, uint32_t width, uint32_t height,
                          float *device_chain);
MLV_API int mlv_mip_chain_adjoint(mlv_context *ctx, const float *device_grad_chain, uint32_t width, uint32_t height,
                                  float *device_grad_texture);

/* device_rgba_out: width*height float4. */
MLV_API int mlv_diff_forward(mlv_context *ctx, const mlv_diff_scene *scene, const uint32_t *device_prim_ids,
                             float *device_rgba_out);
/* Overwrites every non-NULL gradient array. deterministic = 0 accumulates with FP32 atomics;
 * deterministic = 1 sorts the per-pixel contributions by texel / vertex and sums each segment
 * in pixel order, so repeated calls give bit-identical gradients. */
MLV_API int mlv_diff_backward(mlv_context *ctx, const mlv_diff_scene *scene, const uint32_t *device_prim_ids,
                              const float *device_grad_rgba, const mlv_diff_grads *grads, int deterministic);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): mlv_diff_backward
 * with deterministic = 0, whose scratch (the FP64 position sums and the per-block SH partials) is
 * a fixed 32 MiB device array of the library, one per GPU, held until the process ends. It calls
 * no hipMalloc or hipFree and runs no flush: recorded draws must have been flushed
 * (MLV_ERROR_INVALID_STATE otherwise). A scene needing more than 32 MiB of scratch (24 B per
 * vertex + 216 B per 16x16 pixel tile of the image) returns MLV_ERROR_CAPACITY. Every call on a
 * GPU uses that GPU's array: calls that overlap on one GPU (other streams or contexts) race. */
MLV_API int mlv_diff_backward_noalloc(mlv_context *ctx, const mlv_diff_scene *scene, const uint32_t *device_prim_ids,
                                      const float *device_grad_rgba, const mlv_diff_grads *grads);

/* ---- batched differentiable shading (extension): many views of many meshes in one launch
 *
 * The shading above for every covered pixel of every view, with each pixel's mesh found from
 * the draw that wrote it. Record the views' draws with mlv_om_set_primitive_id_buffer and
 * mlv_om_set_visibility_buffer bound per view; draw tags count draws from the last
 * mlv_reset_stats. Pixel (x, y) of a view whose primitive id is p and draw tag is t is shaded
 * with mesh draw_mesh[t - first_tag] when t - first_tag < draw_count, triangle p of it, the
 * view's clip_from_world and one SH light shared by all views. */
typedef struct mlv_diff_mesh {
    const float *positions;    /* device, [vertex_count][3] */
    const float *colors;       /* device, [vertex_count][3] */
    const float *uvs;          /* device, [vertex_count][2] */
    const float *normals;      /* device, [vertex_count][3] */
    const uint32_t *indices;   /* device, [triangle_count][3] */
    uint32_t vertex_count, triangle_count;
    const float *texture;      /* device RGBA32F mip chain as in mlv_diff_scene */
    uint32_t tex_width, tex_height, tex_levels;
    /* backward outputs, device, shapes as the inputs above; NULL skips */
    float *grad_positions, *grad_colors, *grad_uvs, *grad_normals, *grad_texture;
    /* optional device scratch of [vertex_count][3] doubles carrying the FP64 position sums;
       needed (with grad_positions) by accumulating backward calls, NULL uses library scratch */
    double *pos_accum;
} mlv_diff_mesh;

typedef struct mlv_diff_view {
    float clip_from_world[16]; /* row-major */
    uint32_t width, height;
    const uint32_t *prim_ids;  /* device, width*height: the view's primitive id buffer */
    const uint32_t *draw_tags; /* device, width*height: the view's visibility-buffer draw tags */
    uint32_t first_tag;        /* draw tag of draw_mesh[0] */
    uint32_t draw_count;
    const uint32_t *draw_mesh; /* host, draw_count mesh indices; 0xFFFFFFFF (or any index past
                                  the mesh count) leaves that draw's pixels out */
    float *rgba;               /* device, width*height float4: forward output */
    const float *grad_rgba;    /* device, width*height float4: backward input */
} mlv_diff_view;

/* Writes every view's rgba: (0, 0, 0, 0) where no primitive was drawn, the shaded colour where
 * the pixel's draw maps to a mesh; pixels of draws mapped to no mesh are left unchanged. */
MLV_API int mlv_diff_batch_forward(mlv_context *ctx, const mlv_diff_mesh *meshes, uint32_t mesh_count,
                                   const mlv_diff_view *views, uint32_t view_count, const float *device_sh);
/* Gradients of sum over views of sum(rgba * grad_rgba). accumulate = 0 overwrites every
 * non-NULL gradient array, pos_accum and device_grad_sh ([9][3], if not NULL); accumulate = 1
 * adds this call's gradients to their contents (positions: to the FP64 sums in pos_accum, then
 * grad_positions is rewritten from them). Texel, colour, UV and normal gradients accumulate with
 * FP32 atomics, positions with FP64 atomics. */
MLV_API int mlv_diff_batch_backward(mlv_context *ctx, const mlv_diff_mesh *meshes, uint32_t mesh_count,
                                    const mlv_diff_view *views, uint32_t view_count, const float *device_sh,
                                    float *device_grad_sh, int accumulate);
/* out[i] = (a[i] - b[i]) * scale for i < count (device arrays): the gradient of
 * (scale / 2) * sum((a - b)^2) with respect to a, i.e. of the mean squared error for
 * scale = 2 / count. No reference counterpart: the loss of the differentiable step. Like the two
 * batched passes, it may be recorded into a graph (mlv_graph_capture_begin/end), so that one
 * graph holds a whole step: the raster frame, the shading pass, the loss gradient and the adjoint;
 * a recorded pass keeps copies of its descriptor arrays and draw maps, and the device buffers
 * they name must outlive the graph. */
MLV_API int mlv_diff_l2_grad(mlv_context *ctx, const float *device_a, const float *device_b, uint64_t count,
                             float scale, float *device_out);

#ifdef __cplusplus
}
#endif
#endif
