// mlv_internal.h -- structures shared by the host runtime and the gfx950 kernels.
//
// Draws are recorded by mlv_draw_indexed and executed in batches (a flush): one pass of
// each stage covers every recorded draw of every view.
//
// HBM layout of one flush (library-owned arenas, reused across flushes):
//   draws[D]         DrawRec   per recorded draw (shaders, buffers, viewport, offsets)
//   views[NV]        ViewRec   per render target (colour/depth/hi-Z/visibility pointers, tile grid)
//   vpos[Vt]         float4    clip-space position per vertex slot (VS register 0)
//   vattr[2*(Vt+C)]  float4    VS registers 1..2 per vertex slot; rows >= Vt hold clipper vertices
//   cnt[Tt]          u32       per input-triangle slot: #binned | #assembled << 8 | #clip vertices << 16
//   tvis[N]          TriVis    80 B per binned triangle (edge functions, depth plane, bin rect)
//   tshade[N]        TriShade  32 B per binned triangle (1/w, attribute rows, numbers)
//   tbin[N]          uint4     16 B per binned triangle (bin rect and its view's tile grid)
//   cmark            u32       bitmap, one bit per (draw, 8x8 bin of its view): bins met by a
//                              triangle that was assembled but not binned (zero between flushes)
// An assembled triangle whose snapped bounding box holds no pixel corner covers no pixel
// (coverage is sampled at pixel corners only), so it is not binned: it counts in the stats, its
// bins in the bitmap. Binned triangles are numbered in draw order, then submission order; each
// also carries its number among the draw's assembled triangles (the reference's triangle id).
//   pair keys/vals   u32       (global tile, triangle) pairs, stably sorted by tile
//   tile_start/end   u32       per global tile (tiles of all views, view-major): its pair range
// Slots of draw d: vertices [vbase, vbase+V_d), input triangles [tbase, tbase+T_d); both
// start on a 256 boundary so every block of the vertex and setup stages belongs to one draw.
// The render target is the reference's row-major u32 frame_buffer and f32 depth_buffer
// (main.c:35-36) plus one f32 per 8x8 bin of hierarchical-Z minima (main.c:230).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlv {

constexpr int kTile = 16;          // default screen tile edge for binning (pixels); 32 is the other choice
constexpr int kBin = 8;            // reference bin edge (main.c:24-25): hierarchical-Z granularity
constexpr int kBlock = 256;        // threads per block of the vertex and setup stages
constexpr int kMaxClip = 16;       // MAX_NUM_CLIP_VERTICES (main.c:38)
constexpr int kBigRect = 32;       // a triangle spanning more tiles is binned by the whole block
constexpr int kStages = 4;         // vertex, setup, binning, tile
// Launch parameters the autotuner picks (defaults below): a tile whose list is longer than
// heavy_min has its bins walked by 16 waves each (4 per bin); the persistent heavy-tile kernel
// runs heavy_blocks blocks of 1024 threads; the pair count/emit kernels run pair_blocks blocks.
// The 16-wave path is off by default: with the per-bin screening, one wave per bin was faster
// on every measured workload (1080p frames, the binning stress, 512^2 views of dense meshes).
constexpr uint32_t kHeavyMinDefault = 0xFFFFFFFFu, kHeavyBlocksDefault = 1024, kPairBlocksDefault = 2048;
// The tile kernel stages its list tile_stage records at a time: 256 (22.5 KB of LDS, 7 blocks per
// CU) or 128 (half the LDS: 8 blocks per CU, more barriers for long lists).
constexpr uint32_t kTileStageDefault = 256;
constexpr uint32_t kLdsPerCu = 160u * 1024u;   // gfx950
constexpr uint32_t kLdsPerBlockMax = 64u * 1024u;

constexpr uint32_t kFirstOfDraw = 0x80000000u;  // blk_draw flag: the block is its draw's first
enum VsKind { VS_BASIC = 1, VS_VERTEX_LIGHTING = 2, VS_PASSTHROUGH = 3, VS_FULLSCREEN = 4 };
// PS_BASIC_LAMBERT and PS_ENV_TEXTURED are extensions (mip-mapped fetch, a light, a second
// texture): they read the registers at the next pixel corners too (footprints for the mip level).
enum PsKind { PS_BASIC = 1, PS_ENV_LIGHTING = 2, PS_PASSTHROUGH = 3, PS_BASIC_LAMBERT = 4, PS_ENV_TEXTURED = 5, PS_NONE = 6 };
enum TexFormat { FMT_RGBA8 = 1, FMT_RGBA32F = 2, FMT_RGBA16F = 3 };
enum DepthFunc { DEPTH_GE = 0, DEPTH_LE = 1 };
enum CullMode { CULL_BACK = 0, CULL_NONE = 1, CULL_FRONT = 2 };  // mlv_cull_mode
enum FillRule { FILL_INCLUSIVE = 0, FILL_TOP_LEFT = 1 };         // mlv_fill_rule

enum OverflowBits : uint32_t {
    OVF_TRIANGLES = 1u,   // more assembled triangles than the triangle arena holds
    OVF_CLIPVERTS = 2u,   // more clipper vertices than the attribute arena holds
    OVF_PAIRS = 4u,       // more (triangle, tile) pairs than the bin-list arena holds (slow path ran)
};

// A texture and its mip chain: mip_levels levels concatenated, level 0 first, level l of
// max(width >> l, 1) x max(height >> l, 1) texels.
struct TexDesc {
    const void *data;
    uint32_t *mark;               // footprint counting: one bit per texel of the chain (NULL: off)
    uint64_t mark_base;           // bit of this level's first texel in `mark`
    uint32_t width, height, format, mip_levels;
};
constexpr uint32_t kPsConstants = 12;  // floats of the extension pixel shaders' constant buffer

struct DrawRec {
    const uint32_t *ib;           // index buffer + IA offset + start_index
    const uint8_t *vb;            // vertex buffer + IA offset
    const float *cb;              // constant buffer slot 0 (48 floats)
    TexDesc vs_tex, ps_tex;
    TexDesc ps_tex1;              // pixel shader resource slot 1 (extension shaders)
    const float *ps_cb;           // pixel shader constant buffer slot 0 (extension shaders, kPsConstants floats)
    float screen[4][4];           // screen_from_ndc of the draw's viewport (main.c:826-831)
    uint32_t stride, vertex_count, tri_count;
    int32_t base_vertex;
    int32_t vs, ps, depth_func, depth_cull;
    int32_t cull, fill;           // mlv_cull_mode, mlv_fill_rule
    uint32_t view;                // index into the view table
    uint32_t vbase, tbase;        // first vertex slot / first input-triangle slot
    // First attribute row of the draw's vertices (VS registers 1-2). Registers 1-2 of basic_vs,
    // vertex_lighting_vs and passthrough_vs depend on the vertex alone (and vertex_lighting_vs's
    // texture), not on the constant buffer, so a draw of the same vertices with the same shader
    // earlier in the flush (another view of the same mesh) lends its rows: abase is then that
    // draw's vbase and the vertex stage writes positions only. Otherwise abase = vbase.
    uint32_t abase;
    uint32_t blk_v0, blk_t0;      // first block of the vertex / setup stage
    uint32_t draw_tag;            // written to the visibility buffer
    uint32_t cbase;               // first bit of this draw's bins in BatchParams::cmark
};

struct ViewRec {
    uint32_t *color;
    float *depth;
    float *hiz;
    uint32_t *vis_tri, *vis_draw;
    uint32_t *vis_prim;           // input triangle index of the visible triangle (optional)
    int32_t W, H, tiles_x, tiles_y, bins_x, bins_y;
    uint32_t tile_base;           // first global tile of this view
    int32_t depth_func;
};

// One binned triangle as the tile kernel reads it (80 B). Its index in the binned array is its
// order key (draw order, then submission order); the reference's per-draw number is
// TriShade::num.
// a16/b16 are the edge coefficients pre-multiplied by 16 (mod 2^32), so that
// E = a16*x + b16*y + c equals ((a*x)<<4) + ((b*y)<<4) + c of main.c:1015-1025.
struct alignas(16) TriVis {
    int32_t a16[3], b16[3], c[3];
    float ooa;               // one_over_area (main.c:866)
    float z0, dz1, dz2;      // screen z of vertex 0 and the differences (v1-v0), (v2-v0)
    float hiz_depth;         // max_depth (main.c:871), or the minimum under LESS_EQUAL
    float cull_bound;        // a bound no interpolated depth of this triangle can pass
    uint32_t rect0, rect1;   // 8x8 bin rect: bx0 | by0<<16, bx1 | by1<<16 (main.c:930-931)
    uint32_t draw;           // index of the draw in the flush; bits 16-18: edge slots the top-left rule excludes
    // The pixel corners inside the snapped bounding box, clamped to the view: x in
    // [lat_x & 0xFFFF, lat_x >> 16], y likewise (as int16). Coverage is sampled only at pixel
    // corners (main.c:1015-1025), so a triangle with no corner in a bin covers nothing there.
    uint32_t lat_x, lat_y;
};
static_assert(sizeof(TriVis) == 80, "TriVis layout");

struct alignas(16) TriShade {
    float rw[3];             // a_reciprocal_ws (main.c:868-870)
    uint32_t row[3];         // attribute rows of the three vertices in vattr
    uint32_t prim;           // input triangle index within the draw
    uint32_t num;            // number among the draw's assembled triangles (the visibility id)
};
static_assert(sizeof(TriShade) == 32, "TriShade layout");

constexpr size_t kMaxBatchDraws = 1u << 16;
constexpr uint32_t kBlockDrawTable = 256;  // draws per flush from which blocks look their draw up in a table  // draws per flush; the tile kernels keep draw indices in 16 bits

struct BatchCounters {
    unsigned long long scan_total;         // setup scan total: binned | assembled << 32
    uint32_t n_assembled;                  // binned triangles (the tile stage's triangles)
    uint32_t n_clipverts;                  // clipper vertex rows taken
    uint32_t n_clip_list;
    uint32_t n_pairs;
    uint32_t overflow;
    uint32_t radix_blocks;                 // blocks of the radix passes over n_pairs
    uint32_t hist_n;                       // radix_blocks * 256 digit counters
    uint32_t n_heavy;                      // tiles whose list is longer than heavy_min
    uint32_t class_cnt[8], class_cur[8];   // tiles per list-length class; slots taken so far (tile_order)
};

// Footprint counting (mlv_context_set_footprint_counting): totals since the last read.
struct FootCounts {
    unsigned long long shaded_px;    // pixels a draw wrote (winners shaded)
    unsigned long long vis_tris;     // assembled triangles that won a pixel, summed over flushes
    unsigned long long tex_bytes;    // bytes of the distinct texels the counted bitmaps hold
    unsigned long long texels;       // their count
};

struct DevStats {
    unsigned long long assembled, active8, pairs8, oob, overflow_draws, capacity_errors, max_pairs, max_clip_list;
    uint32_t scan_errors;  // scans whose look-back gave up waiting (never expected)
    uint32_t pad;
    // kb_tile's active-bin counts, one atomic per tile spread over 64 counters (so that no
    // single address takes thousands of atomics); the host adds them to active8
    unsigned long long active8_part[64];
    unsigned long long pairs8_part[64];    // (triangle, 8x8 bin) pairs of unbinned triangles, spread likewise
};

struct BatchParams {
    const DrawRec *draws;
    const uint32_t *draw_v0, *draw_t0;    // DrawRec::blk_v0 / blk_t0 of every draw, dense
    const ViewRec *views;
    uint32_t n_draws, n_views;
    uint32_t n_vblocks, n_tblocks;        // blocks of the vertex / setup stages
    uint32_t total_tiles;
    float planes[6][4];                   // clipper planes (main.c:652-657)
    // arenas
    float4 *vpos, *vattr;
    uint32_t vslots;                      // vertex slots (rows below this are VS outputs)
    uint32_t vattr_rows;                  // capacity in rows
    uint32_t *cnt;
    unsigned long long *blocksum, *blockoff;  // per setup block: assembled | clipverts << 32
    uint32_t *asm_base;                   // per draw: its first binned triangle's index
    uint32_t *blk_draw;                   // per setup block: its draw | kFirstOfDraw (kb_setup_count writes it)
    // per vertex / setup block: its draw (kb_block_draws writes them), for flushes of at least
    // kBlockDrawTable draws; null for fewer, where a binary search over the draws is short
    uint32_t *vblk_draw, *tblk_draw;
    uint32_t *cmark;                      // (draw, bin) bits of unbinned triangles (NULL: all are binned)
    uint32_t cmark_words;
    uint2 *clip_list;       // (triangle slot, draw) of each triangle the clipper takes
    uint2 *clip_off;                      // per clipped slot: first binned index, first number in its draw
    uint32_t *clip_rows;                  // per clipped slot: its first clipper vertex row (minus vslots)
    TriVis *tvis;
    TriShade *tshade;
    uint4 *tbin;                          // per assembled triangle: rect0, rect1, tile_base, tiles_x of its view
    uint32_t tri_cap;
    uint32_t *pair_cnt, *pair_off;        // per assembled triangle
    uint32_t *keys0, *vals0, *keys1, *vals1;
    uint32_t pair_cap;
    uint32_t *tile_cnt, *tile_start;      // per global tile: end and start of its range in the sorted pairs
    uint32_t *heavy;                      // tiles with more than heavy_min pairs (n_heavy of them)
    uint4 *tile_order;                    // the order the tile kernel takes tiles in, long lists first: per
                                          // slot (global tile, list start, list length, view)
    uint32_t heavy_min, heavy_blocks, pair_blocks;
    uint32_t tile_stage;                  // records the tile kernel stages in LDS at a time (128 or 256)
    uint32_t tile_shift;                  // tiles of 2^tile_shift x 2^tile_shift bins: 1 (16x16 px) or 2 (32x32)
    uint32_t tile_lds_pad, setup_lds_pad; // dynamic LDS of the tile / setup kernels (caps their blocks per CU)
    uint32_t *scan_status;                // the scans' status words, cleared by the first stage
    uint32_t scan_status_words;
    BatchCounters *ctr;
    DevStats *stats;
    uint32_t *tri_mark;                   // footprint counting: one bit per assembled triangle (NULL: off)
    FootCounts *foot;
};

// Differentiable shading over the raster's primitive ids (mlv_diff.hip).
constexpr uint32_t kMaxMips = 16;
// One mesh: per-vertex attributes, triangles and its RGBA32F mip chain.
struct DiffMesh {
    const float *pos, *col, *uv, *nrm;
    const uint32_t *idx;
    uint32_t V, T;
    const float4 *tex;
    uint32_t tex_total;                   // texels of the whole chain
    uint32_t levels;
    uint32_t lw[kMaxMips], lh[kMaxMips], loff[kMaxMips];
    uint32_t *tex_mark;                   // footprint counting: one bit per texel of the chain (NULL: off)
};
// One view: the clip_from_world matrix and the image size (the viewport covers the image).
struct DiffView {
    float M[16];
    uint32_t W, H;
};
struct DiffParams {
    DiffMesh m;
    DiffView v;
    const float *sh;
    const uint32_t *prim;
};
struct DiffGrads {
    float *pos, *col, *uv, *nrm, *tex, *sh;
};
// Batched shading of many views of many meshes in one launch. The grid covers every view's
// pixels, one 16x16 tile per block, each view starting on a block of its own (view_block0). A pixel's mesh
// is draw_mesh[draw_off + tag - first_tag] for the draw tag the visibility buffer holds there.
struct DiffBatchMesh {
    DiffMesh m;
    DiffGrads g;              // per-mesh gradient outputs (g.sh unused); NULL skips
    double *posacc;           // [V][3] FP64 position accumulator
};
struct DiffBatchView {
    DiffView v;
    const uint32_t *prim, *tags;
    uint32_t first_tag, n_draws, draw_off;
    uint32_t identity_map;    // draw d of the view shades mesh d (d < n_meshes): no draw_mesh lookup
    float4 *out;              // forward output
    const float4 *gout;       // backward input
};
struct DiffZeroSeg {
    uint32_t *p;              // words to clear
    uint32_t words;
    uint32_t pad;
};
struct DiffBatchParams {
    const DiffBatchMesh *meshes;
    const DiffBatchView *views;
    const uint32_t *view_block0;          // [n_views] first pixel block of each view
    const uint32_t *draw_mesh;            // mesh of each draw, per view from draw_off
    const uint32_t *mesh_block0;          // [n_meshes] first block of each mesh's 3V position values
    const DiffZeroSeg *zero;              // arrays cleared before the backward pass
    const uint32_t *zero_block0;          // [n_zero] first block of each segment (1024 words a block)
    uint32_t n_meshes, n_views, n_blocks, n_pos_blocks, n_zero, n_zero_blocks;
    const float *sh;
    uint32_t *defer;                      // forward: [0] waves listed for the fix-up pass, [1] its finished blocks
    uint32_t *defer_list;                 // forward: (block << 2 | wave) of each listed wave
};
// Scratch of the deterministic backward: per-pixel contributions, their sort keys and order.
struct DiffScratch {
    uint32_t *tkey, *tval, *tkey2, *tval2;   // 4 per pixel (bilinear taps)
    float4 *tpay;
    uint32_t *vkey, *vval, *vkey2, *vval2;   // 3 per pixel (triangle corners)
    float4 *vpay;                            // 2 float4 per record: dcol.xyz duv.x | duv.y dn.xyz
    double *vposd;                           // 4 doubles per record: dpos.xyz -
    double *posacc;                          // [V][3] position gradient accumulator (atomic mode)
    double *sh_part;                         // 27 per block
    void *part;                              // scan scratch (scan_scratch_bytes)
    uint32_t *hist;                          // radix digit counts
    uint32_t *epoch;                         // host: scan epoch counter
    uint32_t *err;                           // device: scan error counter
};
hipError_t launch_diff_forward(const DiffParams &p, float4 *out, hipStream_t s);
constexpr uint32_t kMaxMipLevels = 16;  // textures up to 32768 texels on a side
uint32_t mip_chain_texels(uint32_t w, uint32_t h, uint32_t *levels);
hipError_t launch_mip_chain(const float4 *tex, uint32_t w, uint32_t h, float4 *chain, hipStream_t s);
hipError_t launch_mip_adjoint(const float4 *gchain, uint32_t w, uint32_t h, float4 *gtex, hipStream_t s);
hipError_t launch_diff_backward(const DiffParams &p, const float4 *gout, const DiffGrads &g, const DiffScratch &sc,
                                bool deterministic, hipStream_t s);
// Blocks of the per-pixel launches: 256 pixels in row order, or 16x16 tiles (the atomic
// backward passes, which aggregate per tile, and the batched launches).
uint32_t diff_blocks(uint32_t W, uint32_t H);
uint32_t diff_tile_blocks(uint32_t W, uint32_t H);
hipError_t launch_diff_batch_forward(const DiffBatchParams &p, hipStream_t s);
hipError_t launch_l2_grad(const float *a, const float *b, uint64_t n, float scale, float *out, hipStream_t s);
// accumulate: keep the gradients and position sums of the previous call and add to them.
hipError_t launch_diff_batch_backward(const DiffBatchParams &p, double *sh_part, float *grad_sh, bool accumulate,
                                      hipStream_t s);
// Scratch of the one-launch exclusive scans and of the radix sort histograms.
struct ScanScratch {
    void *part;          // scan_scratch_bytes(n_max): a ticket counter, then one status per partition;
                         // zeroed when allocated
    size_t part_bytes;
    uint32_t *hist;      // radix digit histograms: 256 x blocks
    size_t hist_elems;
    uint32_t *epoch;     // host-side counter, advanced by every scan launch
    uint32_t *err;       // device counter of scans whose look-back gave up waiting
    size_t wide_off;     // byte offset in `part` of the 64-bit scans' statuses (0: none run)
};

// Stable LSD sort of n (key, value) pairs by the low `bits` of the key (8 bits per pass).
// Returns in `*in_alt` whether the sorted data ended in k1/v1 (else k0/v0).
hipError_t sort_pairs(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, uint32_t n, uint32_t bits,
                      const ScanScratch &sc, bool *in_alt, hipStream_t s);

// A camera state as the reference keeps it (main.c:1422-1430), for per-view constant buffers
// (mlv_camera.hip); the layout of mlv_camera.
struct CameraRec {
    float position[3];
    float yaw_rad, pitch_rad, fov_y_deg, near_plane;
    uint32_t width, height;
};
// cbs: n PerFrameCBs of 48 floats (clip_from_world, view_from_clip, world_from_view)
hipError_t launch_camera_cbs(const CameraRec *cams, uint32_t n, float *cbs, hipStream_t s);

// Launchers (mlv_kernels.hip). All enqueue on `s`.
// Buffers of 32-bit words to fill with one value each, in one launch.
constexpr uint32_t kMaxFills = 8;
struct FillJobs {
    uint32_t *d[kMaxFills];
    uint32_t v[kMaxFills];
    size_t n[kMaxFills];
    uint32_t count;
    void add(void *dst, uint32_t value, size_t words) {  // a null dst or no words: nothing to fill
        if (!dst || !words || count >= kMaxFills) return;
        d[count] = (uint32_t *)dst;
        v[count] = value;
        n[count] = words;
        ++count;
    }
};
hipError_t launch_fill(const FillJobs &f, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, size_t n, hipStream_t s);
// n_pixels u32 frame-buffer values to 3 * n_pixels bytes of RGB; rgb holds 12 * ceil(n_pixels / 4)
// bytes and is 4-byte aligned
hipError_t launch_present_rgb(const uint32_t *fb, uint8_t *rgb, size_t n_pixels, hipStream_t s);
hipError_t launch_srgb_to_linear_rgba8(uint32_t *texels, size_t n, hipStream_t s);
// levels 1 .. mip_levels-1 of the texture's chain from level 0, in order
hipError_t launch_generate_mips(const TexDesc &t, hipStream_t s);
// adds mul * (set bits of words[0..n)) to *out
hipError_t launch_count_clear(uint32_t *words, size_t n, unsigned long long *out, hipStream_t s);
uint32_t tile_kernel_lds(uint32_t tile_stage, uint32_t shift);
uint32_t setup_kernel_lds();
hipError_t launch_popcount(const uint32_t *words, size_t n, uint32_t mul, unsigned long long *out_bits,
                           unsigned long long *out_bytes, hipStream_t s);
hipError_t launch_fill_f32(float *dst, float value, size_t n, hipStream_t s);
// The flush in four stages, in this order: vertex, setup (cull/clip/ordered compaction),
// binning (ordered tile lists), tile (raster + early-Z + shading + output merger).
// tile_ev, if given, gets two events recorded around the kb_tile launch of stage 3.
hipError_t launch_stage(int stage, const BatchParams &p, const ScanScratch &sc, uint32_t radix_passes,
                        hipStream_t s, hipEvent_t *tile_ev = nullptr);
// Bytes of scan scratch for scans over up to n_max elements; elements of the radix histograms.
size_t scan_scratch_bytes(size_t n_max);
size_t scan_wide_offset(size_t n_max);
size_t radix_hist_elems(size_t n_max);

}  // namespace mlv
