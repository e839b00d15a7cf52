// mlv_internal.h -- structures shared by the host runtime and the gfx950 kernels.
//
// HBM layout of one draw (all arrays are library-owned arenas reused across draws):
//   vpos[V]          float4   clip-space position per vertex (VS register 0)
//   vattr[2*(V+C)]   float4   VS registers 1..2 per vertex row; rows >= V hold vertices made by the clipper
//   cnt[T]           u32      per input triangle: (#assembled) | (#clip vertices << 16)
//   tvis[N]          TriVis   80 B per assembled triangle: edge functions, depth plane, bin rect
//   tshade[N]        TriShade 32 B per assembled triangle: 1/w, attribute rows
//   tile_cnt/start   u32      per 16x16 screen tile
//   tile_list[P]     u32      assembled-triangle ids per tile
// The render target is a row-major u32 colour array plus a row-major f32 depth
// array (the reference's frame_buffer/depth_buffer, main.c:35-36), and one f32
// per 8x8 bin for the hierarchical-Z minima (a_tile_min_depths, main.c:230).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlv {

constexpr int kTile = 16;          // screen tile edge for binning (pixels)
constexpr int kBin = 8;            // reference bin edge (main.c:24-25): hierarchical-Z granularity
constexpr int kSetupBlock = 256;   // threads per block of the setup kernels
constexpr int kMaxClip = 16;       // MAX_NUM_CLIP_VERTICES (main.c:38)
constexpr int kBigRect = 32;       // a triangle spanning more tiles is binned by the whole block

enum VsKind { VS_BASIC = 1, VS_VERTEX_LIGHTING = 2, VS_PASSTHROUGH = 3, VS_FULLSCREEN = 4 };
enum PsKind { PS_BASIC = 1, PS_ENV_LIGHTING = 2, PS_PASSTHROUGH = 3 };
enum TexFormat { FMT_RGBA8 = 1, FMT_RGBA32F = 2, FMT_RGBA16F = 3 };
enum DepthFunc { DEPTH_GE = 0, DEPTH_LE = 1 };

enum OverflowBits : uint32_t {
    OVF_TRIANGLES = 1u,   // more assembled triangles than the triangle arena holds
    OVF_CLIPVERTS = 2u,   // more clipper vertices than the attribute arena holds
    OVF_PAIRS = 4u,       // more (triangle, tile) pairs than the bin-list arena holds (slow path ran)
};

struct TexDesc {
    const void *data;
    uint32_t width, height, format, mip_levels;
};

// One assembled triangle as the tile kernel reads it (80 B).
// a16/b16 are the edge coefficients pre-multiplied by 16 (mod 2^32), so that
// E = a16*x + b16*y + c equals ((a*x)<<4) + ((b*y)<<4) + c of main.c:1015-1025.
struct alignas(16) TriVis {
    int32_t a16[3], b16[3], c[3];
    float ooa;               // one_over_area (main.c:866)
    float z0, dz1, dz2;      // screen z of vertex 0 and the differences (v1-v0), (v2-v0)
    float hiz_depth;         // max_depth (main.c:871), or the minimum under LESS_EQUAL
    float cull_bound;        // a bound no interpolated depth of this triangle can pass
    uint32_t rect0, rect1;   // 8x8 bin rect: bx0 | by0<<16, bx1 | by1<<16 (main.c:930-931)
    uint32_t id;             // assembled index (submission order)
    uint32_t pad[2];
};
static_assert(sizeof(TriVis) == 80, "TriVis layout");

struct alignas(16) TriShade {
    float rw[3];             // a_reciprocal_ws (main.c:868-870)
    uint32_t row[3];         // attribute rows of the three vertices in vattr
    uint32_t prim;           // input triangle index
    uint32_t sub;            // fan index within a clipped polygon
};
static_assert(sizeof(TriShade) == 32, "TriShade layout");

struct DrawCounters {
    uint32_t n_assembled;
    uint32_t n_clipverts;
    uint32_t n_clip_list;
    uint32_t n_pairs;
    uint32_t overflow;
    uint32_t oob;
    uint32_t active8;
    uint32_t pad0;
    unsigned long long pairs8;
};

struct DevStats {
    unsigned long long assembled, active8, pairs8, oob, overflow_draws, capacity_errors, max_pairs, max_clip_list;
};

struct DrawParams {
    // input assembler
    const uint32_t *ib;
    uint32_t index_count, start_index;
    int32_t base_vertex;
    uint32_t tri_count;          // index_count / 3
    const uint8_t *vb;
    uint32_t stride, vertex_count;
    // vertex stage
    const float *cb;             // 48 floats on the device: clip_from_world, view_from_clip, world_from_view
    TexDesc vs_tex, ps_tex;
    // rasteriser state
    float screen[4][4];          // screen_from_ndc (main.c:826-831) as f32 rows
    float planes[6][4];          // clipper planes (main.c:652-657)
    int32_t W, H;                // render target = viewport extent
    int32_t tiles_x, tiles_y, bins_x, bins_y;
    int32_t depth_func;
    int32_t depth_cull;          // 1 when the in-draw conservative depth cull is exact-safe
    uint32_t draw_tag;
    // targets
    uint32_t *color;
    float *depth;
    float *hiz;
    uint32_t *vis_tri, *vis_draw;
    // arenas
    float4 *vpos, *vattr;
    uint32_t vattr_rows;         // capacity in rows (V + clip rows)
    uint32_t *cnt;
    uint32_t *blocksum;          // packed per setup block
    uint32_t *blockoff_tri, *blockoff_vtx;
    uint32_t nblocks;
    uint32_t *clip_list;         // capacity tri_count
    uint2 *clip_off;             // per input triangle (only clipped entries used)
    TriVis *tvis;
    TriShade *tshade;
    uint32_t tri_cap;
    uint32_t *tile_cnt, *tile_start, *tile_fill, *tile_list;
    uint32_t pair_cap;
    DrawCounters *ctr;
    DevStats *stats;
};

// Launchers (mlv_kernels.hip). All enqueue on `s`.
hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, size_t n, hipStream_t s);
hipError_t launch_fill_f32(float *dst, float value, size_t n, hipStream_t s);
hipError_t launch_draw(const DrawParams &p, int vs, int ps, hipStream_t s);

}  // namespace mlv
