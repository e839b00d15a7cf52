// mlv_runtime.hip -- host side of the C ABI (include/malevich.h).
//
// The context owns the pipeline state that the reference keeps in its global
// `Pipeline graphics_pipeline` (main.c:109-115, 222), one HIP stream, and the
// transient arenas that draw_indexed mallocs and frees on every call in the
// reference (main.c:675-746, 947, 965, 986, 1253-1259). Arenas here grow
// geometrically and are reused, so a steady-state frame allocates nothing.
//
// mlv_draw_indexed validates the state and records the draw; recorded draws run
// together in one pass of each stage (a flush) when their results are needed: a
// readback, mlv_synchronize, a clear of a target they draw into, a change to a
// buffer, texture or target they read, or when the batch reaches its size limit.
// Stream order then gives the same results as running each draw on its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "malevich.h"
#include "mlv_internal.h"

#include <rocprofiler-sdk-roctx/roctx.h>

using namespace mlv;

struct mlv_buffer {
    void *ptr;
    size_t bytes;
    int kind;
    bool owned;
};

struct mlv_texture2d {
    void *ptr;
    uint32_t width, height, format, mip_levels;
    size_t bytes;
    uint32_t *mark = nullptr;  // footprint counting: one bit per texel of the chain
    size_t mark_words = 0;
};

struct mlv_render_target {
    uint32_t *ptr;
    uint32_t width, height;
};

struct mlv_depth_stencil {
    float *ptr;
    float *hiz;
    uint32_t width, height, bins_x, bins_y;
    int hiz_mode;
};

namespace {

struct Arena {
    void *p = nullptr;
    size_t cap = 0;
};

// What VS registers 1-2 of a draw's vertices depend on (see DrawRec::abase).
struct AttrKey {
    const void *vb;
    uint32_t stride, count;
    int32_t vs;
    const void *tex;
    bool operator<(const AttrKey &o) const {
        if (vb != o.vb) return vb < o.vb;
        if (stride != o.stride) return stride < o.stride;
        if (count != o.count) return count < o.count;
        if (vs != o.vs) return vs < o.vs;
        return tex < o.tex;
    }
};

struct ViewKey {
    mlv_render_target *rt;
    mlv_depth_stencil *ds;
    uint32_t *vis_tri, *vis_draw, *vis_prim;
    int depth_func;
    bool operator==(const ViewKey &o) const {
        return rt == o.rt && ds == o.ds && vis_tri == o.vis_tri && vis_draw == o.vis_draw && vis_prim == o.vis_prim &&
               depth_func == o.depth_func;
    }
};

struct PendingDraw {
    DrawRec rec;
    ViewKey key;
};

// A batched differentiable pass or a loss gradient recorded for a graph: its arguments, with the
// host arrays they point to (the views' draw maps) copied.
struct DiffCmd {
    std::vector<mlv_diff_mesh> meshes;
    std::vector<mlv_diff_view> views;
    std::vector<uint32_t> draw_mesh;  // the views' draw maps, concatenated
    const float *sh = nullptr;
    float *grad_sh = nullptr;
    int accumulate = 0;
    const float *a = nullptr, *b = nullptr;  // loss gradient: out = (a - b) * scale
    float *out = nullptr;
    uint64_t n = 0;
    float scale = 0.0f;
};

// A command recorded for a graph: a colour clear, a depth clear, a draw, a batched
// differentiable pass or a loss gradient, with every piece of state it reads resolved when it
// was issued.
enum CmdKind { kCmdClearColor = 0, kCmdClearDepth = 1, kCmdDraw = 2, kCmdDiffForward = 3, kCmdDiffBackward = 4,
               kCmdL2Grad = 5 };
struct Cmd {
    int kind = 0;
    mlv_render_target *rt = nullptr;
    mlv_depth_stencil *ds = nullptr;
    uint32_t value = 0;       // encoded clear colour
    float depth = 0.0f;
    int depth_func = 0;
    uint32_t *vis_tri = nullptr, *vis_draw = nullptr, *vis_prim = nullptr;  // cleared with the colour
    PendingDraw pd{};
    uint32_t T = 0, V = 0;
    std::shared_ptr<DiffCmd> diff;
};

// One flush never exceeds these (indices are 32-bit; the arenas stay bounded).
// Default flush size, and a draw's limit: 96M input triangles (256 views of 64 meshes, 1.28G
// triangle draws, measured 85.2 / 81.6 / 83.3 ms of forward raster with 48M / 96M / 192M flushes).
constexpr uint64_t kMaxBatchTriangles = 96ull << 20;
constexpr uint64_t kMaxBatchVertices = 96ull << 20;
constexpr uint64_t kBatchCeiling = 256ull << 20;      // MLV_MAX_BATCH_TRIANGLES may raise flushes to this
constexpr int kStagingSlots = 16;

}  // namespace

struct mlv_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // Pipeline state (main.c:66-115)
    mlv_buffer *ib = nullptr;
    uint32_t ib_off = 0;
    mlv_buffer *vb = nullptr;
    uint32_t vb_stride = 0, vb_off = 0;
    int topology = MLV_TOPOLOGY_UNDEFINED;
    int vs = 0, ps = 0;
    mlv_buffer *vs_cb[16] = {};
    mlv_texture2d *vs_srv[16] = {};
    mlv_texture2d *ps_srv[16] = {};
    mlv_buffer *ps_cb[16] = {};  // extension: the reference's PS stage has no constant buffers (main.c:98-101)
    mlv_viewport vp{};
    bool vp_set = false;
    mlv_render_target *rt = nullptr;
    mlv_depth_stencil *ds = nullptr;
    int depth_func = MLV_DEPTH_GREATER_EQUAL;
    int cull = MLV_CULL_BACK, fill = MLV_FILL_INCLUSIVE;
    uint32_t *vis_tri = nullptr, *vis_draw = nullptr, *vis_prim = nullptr;
    // recorded draws
    std::vector<PendingDraw> pending;
    uint64_t pending_tris = 0, pending_verts = 0;
    // arenas
    Arena blk_draw;  // per setup block: its draw (written by the count pass for the emit pass)
    Arena vblk_draw, tblk_draw;  // per vertex / setup block: its draw (the flush's first kernel)
    Arena draws_dev, views_dev, vpos, vattr, cnt, blocksum, blockoff, asm_base, clip_list, clip_off, clip_rows, tvis,
        tshade, tbin;
    Arena pair_cnt, pair_off, keys0, vals0, keys1, vals1, tile_cnt, tile_start, scan_part, radix_hist, heavy, tile_order;
    // differentiable shading scratch
    Arena d_shpart, d_tkey, d_tval, d_tkey2, d_tval2, d_tpay, d_vkey, d_vval, d_vkey2, d_vval2, d_vpay, d_part, d_hist;
    Arena d_vposd, d_posacc;
    // batched differentiable shading: device tables, FP64 position sums, SH partials
    Arena b_tables, b_posacc, b_shpart;
    Arena b_defer;  // batched forward: the fix-up pass's counts (zero between calls) and wave list
    Arena present_rgb;  // packed RGB of the last present
    // footprint counting (mlv_context_set_footprint_counting)
    bool counting = false;
    FootCounts *foot = nullptr;
    Arena tri_mark, diff_mark;
    Arena cmark;  // (draw, bin) bitmap of assembled triangles that are not binned (zero between flushes)
    bool skip_cornerless = true;  // assembled triangles whose box holds no pixel corner are not binned
    size_t diff_mark_words = 0;
    std::vector<mlv_texture2d *> marked;  // textures holding a footprint bitmap
    Arena cameras;      // camera table of the last mlv_camera_constant_buffers
    BatchCounters *ctr = nullptr;
    DevStats *dstats = nullptr;
    uint32_t scan_epoch = 0;  // advanced by every scan launch (tags the scans' status words)
    // pinned staging for the draw/view tables, one slot per flush in flight
    void *staging[kStagingSlots] = {};
    size_t staging_cap[kStagingSlots] = {};
    hipEvent_t staging_ev[kStagingSlots] = {};
    int staging_next = 0;
    // host-side stats (main.c:1228-1246)
    uint64_t vertex_count = 0, input_tris = 0, draws = 0;
    uint32_t draw_counter = 0;
    uint64_t capacity_errors = 0;
    size_t clip_hint = 4096, pair_hint = 0;
    size_t pair_limit = 0;  // MLV_PAIR_CAPACITY: caps the bin-list arena (exercises the overflow path)
    // a flush takes recorded draws until one more would pass either limit (MLV_MAX_BATCH_TRIANGLES
    // and MLV_MAX_BATCH_DRAWS lower them; a single larger draw still runs, alone)
    uint64_t batch_tris = kMaxBatchTriangles, batch_verts = kMaxBatchVertices, batch_draws = kMaxBatchDraws;
    // stage profiling: event pairs recorded per stage, resolved at mlv_get_profile
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double stage_ms[kStages] = {};
    double tile_kernel_ms = 0.0;  // the kb_tile launches alone (bucket kStages of ev_pending)
    uint64_t prof_draws = 0, prof_flushes = 0;
    // graphs: commands recorded between capture_begin and capture_end, replayed once to size
    // every arena and then under stream capture
    bool recording = false;          // clears and draws are recorded, not run
    bool capturing = false;          // the stream is being captured: no allocation, no host waits
    bool sizing = false;             // the pass before capture: note the staging sizes
    std::vector<Cmd> rec;
    std::vector<size_t> cap_sizes;   // staging bytes of each table upload of the sizing pass
    std::vector<void *> *cap_bufs = nullptr;  // pinned tables owned by the graph being captured
    size_t cap_next = 0;
    int graphs_alive = 0;
    std::vector<void *> retired;     // arenas replaced while a graph may still read them
    // launch parameters (mlv_tuning)
    uint32_t heavy_min = kHeavyMinDefault, heavy_blocks = kHeavyBlocksDefault, pair_blocks = kPairBlocksDefault;
    uint32_t tile_stage = kTileStageDefault;
    uint32_t tile_edge = kTile, tile_waves = 0, setup_waves = 0;  // 0 waves: no occupancy cap
    size_t last_pair_cap = 0;
};

struct mlv_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void *> pinned;      // table uploads the graph's copy nodes read
    uint64_t vertex_count = 0, input_tris = 0, draws = 0;
};

static int fail(mlv_context *ctx, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
static int fail(mlv_context *ctx, int code, const char *fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

#define HIP_OR_FAIL(ctx, call)                                                                                \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            return fail((ctx), e_ == hipErrorOutOfMemory ? MLV_ERROR_OUT_OF_MEMORY : MLV_ERROR_DEVICE,        \
                        "%s failed: %s", #call, hipGetErrorString(e_));                                       \
    } while (0)

#define RC_OR_RETURN(expr)             \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != MLV_OK) return rc_; \
    } while (0)

// A roctx range for the life of a scope, named like the reference's Remotery samples
// (rmt_BeginCPUSample, main.c:663-1305) so that profiler timelines carry the same stage names.
struct Marker {
    explicit Marker(const char *name) { roctxRangePushA(name); }
    ~Marker() { roctxRangePop(); }
    Marker(const Marker &) = delete;
    Marker &operator=(const Marker &) = delete;
};

static int ensure(mlv_context *ctx, Arena &a, size_t bytes) {
    if (bytes <= a.cap) return MLV_OK;
    if (ctx->capturing)
        return fail(ctx, MLV_ERROR_CAPACITY, "graph capture needed more scratch than its sizing pass; capture again");
    size_t want = std::max(bytes, a.cap + a.cap / 2);
    want = (want + 255) & ~(size_t)255;
    if (a.p) {
        HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->graphs_alive) ctx->retired.push_back(a.p);  // a graph's kernels may still name it
        else HIP_OR_FAIL(ctx, hipFree(a.p));
        a.p = nullptr;
        a.cap = 0;
    }
    HIP_OR_FAIL(ctx, hipMalloc(&a.p, want));
    a.cap = want;
    return MLV_OK;
}

static void release(Arena &a) {
    if (a.p) (void)hipFree(a.p);
    a.p = nullptr;
    a.cap = 0;
}

static void free_retired(mlv_context *ctx) {
    for (void *p : ctx->retired) (void)hipFree(p);
    ctx->retired.clear();
}

static int set_device(mlv_context *ctx) {
    HIP_OR_FAIL(ctx, hipSetDevice(ctx->device));
    return MLV_OK;
}

static hipEvent_t take_event(mlv_context *ctx) {
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

static void resolve_events(mlv_context *ctx) {
    for (auto &pe : ctx->ev_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
            if (pe.first < kStages) ctx->stage_ms[pe.first] += ms;
            else ctx->tile_kernel_ms += ms;
        }
        ctx->ev_pool.push_back(pe.second.first);
        ctx->ev_pool.push_back(pe.second.second);
    }
    ctx->ev_pending.clear();
}

// ------------------------------------------------------------------ flush

static uint32_t align256(uint64_t v) { return (uint32_t)((v + 255u) & ~255ull); }

// screen_from_ndc, main.c:826-831: entries computed in double and stored as f32
static void screen_matrix(const mlv_viewport &vp, float S[4][4]) {
    const float M[4][4] = {
        {(float)((double)vp.width * 0.5), 0.0f, 0.0f, (float)((double)vp.width * 0.5 + (double)vp.top_left_x)},
        {0.0f, (float)(-(double)vp.height * 0.5), 0.0f, (float)((double)vp.height * 0.5 + (double)vp.top_left_y)},
        {0.0f, 0.0f, vp.max_depth - vp.min_depth, vp.min_depth},
        {0.0f, 0.0f, 0.0f, 1.0f}};
    memcpy(S, M, sizeof M);
}

// Pinned staging for tables copied to the device: a ring of slots, each reused only after the
// copies queued from it have completed (staging_commit records that point on the stream).
static int staging_acquire(mlv_context *ctx, size_t bytes, void **host, int *slot_out) {
    if (ctx->capturing) {  // the graph's own buffer, filled once and read by every replay
        if (!ctx->cap_bufs || ctx->cap_next >= ctx->cap_bufs->size() || ctx->cap_sizes[ctx->cap_next] < bytes)
            return fail(ctx, MLV_ERROR_CAPACITY, "graph capture uploads differ from its sizing pass; capture again");
        *host = (*ctx->cap_bufs)[ctx->cap_next++];
        *slot_out = -1;
        return MLV_OK;
    }
    if (ctx->sizing) ctx->cap_sizes.push_back(bytes);
    const int slot = ctx->staging_next;
    ctx->staging_next = (slot + 1) % kStagingSlots;
    if (ctx->staging_ev[slot]) HIP_OR_FAIL(ctx, hipEventSynchronize(ctx->staging_ev[slot]));
    if (ctx->staging_cap[slot] < bytes) {
        if (ctx->staging[slot]) (void)hipHostFree(ctx->staging[slot]);
        ctx->staging[slot] = nullptr;
        ctx->staging_cap[slot] = 0;
        size_t want = std::max<size_t>(bytes, 64 * 1024);
        HIP_OR_FAIL(ctx, hipHostMalloc(&ctx->staging[slot], want, hipHostMallocDefault));
        ctx->staging_cap[slot] = want;
    }
    if (!ctx->staging_ev[slot])
        HIP_OR_FAIL(ctx, hipEventCreateWithFlags(&ctx->staging_ev[slot], hipEventDisableTiming));
    *host = ctx->staging[slot];
    *slot_out = slot;
    return MLV_OK;
}

static int staging_commit(mlv_context *ctx, int slot) {
    if (slot >= 0) HIP_OR_FAIL(ctx, hipEventRecord(ctx->staging_ev[slot], ctx->stream));
    return MLV_OK;
}

// Dynamic LDS that holds a kernel of `wpb` waves a block, `lds` static LDS bytes, to about
// `waves` waves per CU: LDS per CU over the blocks that many waves make, less the static part
// (within the per-block limit). 0 waves: no cap.
static uint32_t lds_pad(uint32_t lds, uint32_t wpb, uint32_t waves) {
    if (!waves) return 0u;
    const uint32_t blocks = std::max(1u, waves / wpb);
    const uint32_t per = std::min(kLdsPerCu / blocks, kLdsPerBlockMax);
    return per > lds ? (per - lds) & ~255u : 0u;
}

static int flush(mlv_context *ctx) {
    if (ctx->recording)
        return fail(ctx, MLV_ERROR_INVALID_STATE,
                    "only state, clears, draws and batched differentiable passes may be issued while recording a graph");
    if (ctx->pending.empty()) return MLV_OK;
    RC_OR_RETURN(set_device(ctx));
    std::vector<PendingDraw> draws;
    draws.swap(ctx->pending);
    ctx->pending_tris = ctx->pending_verts = 0;

    // view table (first appearance order) and slot offsets
    std::vector<ViewKey> keys;
    std::vector<ViewRec> views;
    std::map<AttrKey, uint32_t> attr_owner;  // vertices whose registers 1-2 a draw of this flush writes
    uint64_t total_tiles = 0, vslots = 0, tslots = 0;
    for (auto &pd : draws) {
        size_t vi = 0;
        while (vi < keys.size() && !(keys[vi] == pd.key)) ++vi;
        if (vi == keys.size()) {
            keys.push_back(pd.key);
            ViewRec v{};
            v.color = pd.key.rt->ptr;
            v.depth = pd.key.ds->ptr;
            v.hiz = pd.key.ds->hiz;
            v.vis_tri = pd.key.vis_tri;
            v.vis_draw = pd.key.vis_draw;
            v.vis_prim = pd.key.vis_prim;
            v.W = (int32_t)pd.key.rt->width;
            v.H = (int32_t)pd.key.rt->height;
            v.tiles_x = (v.W + (int32_t)ctx->tile_edge - 1) / (int32_t)ctx->tile_edge;
            v.tiles_y = (v.H + (int32_t)ctx->tile_edge - 1) / (int32_t)ctx->tile_edge;
            v.bins_x = (int32_t)pd.key.ds->bins_x;
            v.bins_y = (int32_t)pd.key.ds->bins_y;
            v.tile_base = (uint32_t)total_tiles;
            v.depth_func = pd.key.depth_func;
            total_tiles += (uint64_t)v.tiles_x * v.tiles_y;
            views.push_back(v);
        }
        DrawRec &r = pd.rec;
        r.view = (uint32_t)vi;
        r.vbase = (uint32_t)vslots;
        r.abase = r.vbase;
        if (r.ps != PS_NONE && r.vs != VS_FULLSCREEN) {  // registers 1-2 from the vertex alone
            const AttrKey ak{r.vb, r.stride, r.vertex_count, r.vs, r.vs == VS_VERTEX_LIGHTING ? r.vs_tex.data : nullptr};
            auto it = attr_owner.find(ak);
            if (it != attr_owner.end()) r.abase = it->second;
            else attr_owner.emplace(ak, r.vbase);
        }
        r.blk_v0 = (uint32_t)(vslots / 256);
        vslots += align256(r.vertex_count);
        r.tbase = (uint32_t)tslots;
        r.blk_t0 = (uint32_t)(tslots / 256);
        tslots += align256(r.tri_count);
    }
    const uint32_t n_draws = (uint32_t)draws.size();
    const uint32_t n_views = (uint32_t)views.size();
    const uint64_t T = tslots, Vs = vslots;
    // one bit per (draw, 8x8 bin of its view) for the triangles that are assembled but not binned
    uint64_t cbits = 0;
    for (auto &pd : draws) {
        pd.rec.cbase = (uint32_t)cbits;
        const ViewRec &v = views[pd.rec.view];
        cbits += (uint64_t)v.bins_x * (uint64_t)v.bins_y;
    }
    const bool cull = ctx->skip_cornerless && cbits < (1ull << 31);  // else every assembled triangle is binned
    // Clipped triangles are unknown until setup runs: start from 1/32 of the batch (triangles
    // straddling the screen edges of a dense scene) and grow from the measured maximum.
    const size_t clip_tris = std::max<size_t>({ctx->clip_hint, (size_t)4096, (size_t)(T / 32)});
    const uint64_t tri_cap64 = T + 6ull * clip_tris;
    const uint64_t rows64 = Vs + 9ull * clip_tris;
    if (tri_cap64 > 0xF0000000ull || rows64 > 0xF0000000ull || total_tiles > 0xF0000000ull)
        return fail(ctx, MLV_ERROR_CAPACITY, "flush too large for 32-bit indexing");
    const uint32_t tri_cap = (uint32_t)tri_cap64;
    const uint32_t vattr_rows = (uint32_t)rows64;
    // (triangle, tile) pairs cost 16 B each; a generous first guess avoids the exact but slow
    // overflow path (every tile walking every triangle), and later flushes size from the
    // measured maximum.
    size_t pair_cap = std::max<size_t>(ctx->pair_hint, 8ull * T + 16ull * total_tiles + 1024);
    pair_cap = std::min<size_t>(pair_cap, 0xF0000000ull);
    if (ctx->pair_limit) pair_cap = std::min(pair_cap, ctx->pair_limit);
    const uint32_t n_tblocks = (uint32_t)(T / 256), n_vblocks = (uint32_t)(Vs / 256);
    const size_t scan_n =
        std::max<size_t>({(size_t)n_tblocks, (size_t)tri_cap, (size_t)total_tiles, radix_hist_elems(pair_cap)});

    RC_OR_RETURN(ensure(ctx, ctx->draws_dev, (sizeof(DrawRec) + 8) * n_draws));  // + the first blocks, dense
    RC_OR_RETURN(ensure(ctx, ctx->views_dev, sizeof(ViewRec) * n_views));
    RC_OR_RETURN(ensure(ctx, ctx->vpos, std::max<size_t>(Vs, 1) * 16));
    RC_OR_RETURN(ensure(ctx, ctx->vattr, (size_t)vattr_rows * 32));
    RC_OR_RETURN(ensure(ctx, ctx->cnt, std::max<size_t>(T, 1) * 4));
    RC_OR_RETURN(ensure(ctx, ctx->blocksum, (size_t)(n_tblocks + 1) * 8));
    RC_OR_RETURN(ensure(ctx, ctx->blockoff, (size_t)(n_tblocks + 1) * 8));
    RC_OR_RETURN(ensure(ctx, ctx->asm_base, (size_t)n_draws * 4));
    RC_OR_RETURN(ensure(ctx, ctx->blk_draw, (size_t)(n_tblocks + 1) * 4));
    RC_OR_RETURN(ensure(ctx, ctx->vblk_draw, (size_t)(n_vblocks + 1) * 4));
    RC_OR_RETURN(ensure(ctx, ctx->tblk_draw, (size_t)(n_tblocks + 1) * 4));
    RC_OR_RETURN(ensure(ctx, ctx->clip_list, std::max<size_t>(T, 1) * 8));
    RC_OR_RETURN(ensure(ctx, ctx->clip_off, std::max<size_t>(T, 1) * 8));
    RC_OR_RETURN(ensure(ctx, ctx->clip_rows, std::max<size_t>(T, 1) * 4));
    RC_OR_RETURN(ensure(ctx, ctx->tvis, (size_t)tri_cap * sizeof(TriVis)));
    RC_OR_RETURN(ensure(ctx, ctx->tshade, (size_t)tri_cap * sizeof(TriShade)));
    RC_OR_RETURN(ensure(ctx, ctx->tbin, (size_t)tri_cap * 16));
    RC_OR_RETURN(ensure(ctx, ctx->pair_cnt, (size_t)tri_cap * 4));
    RC_OR_RETURN(ensure(ctx, ctx->pair_off, (size_t)tri_cap * 4));
    RC_OR_RETURN(ensure(ctx, ctx->keys0, pair_cap * 4));
    RC_OR_RETURN(ensure(ctx, ctx->vals0, pair_cap * 4));
    RC_OR_RETURN(ensure(ctx, ctx->keys1, pair_cap * 4));
    RC_OR_RETURN(ensure(ctx, ctx->vals1, pair_cap * 4));
    RC_OR_RETURN(ensure(ctx, ctx->tile_cnt, (size_t)(total_tiles + 1) * 4));
    RC_OR_RETURN(ensure(ctx, ctx->tile_start, (size_t)(total_tiles + 1) * 4));
    RC_OR_RETURN(ensure(ctx, ctx->heavy, (size_t)(total_tiles + 1) * 4));
    RC_OR_RETURN(ensure(ctx, ctx->tile_order, (size_t)(total_tiles + 1) * 16));
    RC_OR_RETURN(ensure(ctx, ctx->scan_part, scan_scratch_bytes(scan_n)));  // cleared by the first stage
    RC_OR_RETURN(ensure(ctx, ctx->radix_hist, radix_hist_elems(pair_cap) * 4));
    const size_t cwords = (size_t)((cbits + 31) / 32);
    if (cull) {
        const void *old = ctx->cmark.p;
        RC_OR_RETURN(ensure(ctx, ctx->cmark, std::max<size_t>(cwords, 1) * 4));
        if (ctx->cmark.p != old)  // new memory starts at zero; each flush clears the words it used
            HIP_OR_FAIL(ctx, launch_fill_u32((uint32_t *)ctx->cmark.p, 0u, ctx->cmark.cap / 4, ctx->stream));
    }

    // tables through pinned staging
    // the draw records, then each draw's first vertex-stage and setup block as two dense arrays
    // (the blocks' binary searches read those: a few cache lines instead of one per probe)
    const size_t bytes_r = sizeof(DrawRec) * n_draws, bytes_d = bytes_r + 8u * n_draws;
    const size_t bytes_v = sizeof(ViewRec) * n_views;
    void *stage = nullptr;
    int slot = 0;
    RC_OR_RETURN(staging_acquire(ctx, bytes_d + bytes_v, &stage, &slot));
    DrawRec *stage_d = (DrawRec *)stage;
    uint32_t *stage_b = (uint32_t *)((uint8_t *)stage + bytes_r);
    for (uint32_t i = 0; i < n_draws; ++i) {
        stage_d[i] = draws[i].rec;
        stage_b[i] = draws[i].rec.blk_v0;
        stage_b[n_draws + i] = draws[i].rec.blk_t0;
    }
    memcpy((uint8_t *)stage + bytes_d, views.data(), bytes_v);
    HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->draws_dev.p, stage, bytes_d, hipMemcpyHostToDevice, ctx->stream));
    HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->views_dev.p, (uint8_t *)stage + bytes_d, bytes_v, hipMemcpyHostToDevice,
                                    ctx->stream));
    RC_OR_RETURN(staging_commit(ctx, slot));

    BatchParams p{};
    p.draws = (const DrawRec *)ctx->draws_dev.p;
    p.draw_v0 = (const uint32_t *)((const uint8_t *)ctx->draws_dev.p + bytes_r);
    p.draw_t0 = p.draw_v0 + n_draws;
    p.views = (const ViewRec *)ctx->views_dev.p;
    p.n_draws = n_draws;
    p.n_views = n_views;
    p.n_vblocks = n_vblocks;
    p.n_tblocks = n_tblocks;
    p.total_tiles = (uint32_t)total_tiles;
    // clipper planes, main.c:652-657 via v4f32_normalize (math.h:267-270)
    static const float raw[6][4] = {{1, 0, 0, 1}, {-1, 0, 0, 1}, {0, 1, 0, 1}, {0, -1, 0, 1}, {0, 0, 1, 1}, {0, 0, -1, 1}};
    for (int i = 0; i < 6; ++i) {
        float d = raw[i][0] * raw[i][0];
        d = d + raw[i][1] * raw[i][1];
        d = d + raw[i][2] * raw[i][2];
        d = d + raw[i][3] * raw[i][3];
        float len = (float)std::sqrt((double)d);
        float ool = (float)(1.0 / (double)len);
        for (int k = 0; k < 4; ++k) p.planes[i][k] = raw[i][k] * ool;
    }
    p.vpos = (float4 *)ctx->vpos.p;
    p.vattr = (float4 *)ctx->vattr.p;
    p.vslots = (uint32_t)Vs;
    p.vattr_rows = vattr_rows;
    p.cnt = (uint32_t *)ctx->cnt.p;
    p.blocksum = (unsigned long long *)ctx->blocksum.p;
    p.blockoff = (unsigned long long *)ctx->blockoff.p;
    p.asm_base = (uint32_t *)ctx->asm_base.p;
    p.blk_draw = (uint32_t *)ctx->blk_draw.p;
    p.vblk_draw = n_draws >= kBlockDrawTable ? (uint32_t *)ctx->vblk_draw.p : nullptr;
    p.tblk_draw = n_draws >= kBlockDrawTable ? (uint32_t *)ctx->tblk_draw.p : nullptr;
    p.cmark = cull ? (uint32_t *)ctx->cmark.p : nullptr;
    p.cmark_words = cull ? (uint32_t)cwords : 0u;
    p.clip_list = (uint2 *)ctx->clip_list.p;
    p.clip_off = (uint2 *)ctx->clip_off.p;
    p.clip_rows = (uint32_t *)ctx->clip_rows.p;
    p.tvis = (TriVis *)ctx->tvis.p;
    p.tshade = (TriShade *)ctx->tshade.p;
    p.tbin = (uint4 *)ctx->tbin.p;
    p.tri_cap = tri_cap;
    p.pair_cnt = (uint32_t *)ctx->pair_cnt.p;
    p.pair_off = (uint32_t *)ctx->pair_off.p;
    p.keys0 = (uint32_t *)ctx->keys0.p;
    p.vals0 = (uint32_t *)ctx->vals0.p;
    p.keys1 = (uint32_t *)ctx->keys1.p;
    p.vals1 = (uint32_t *)ctx->vals1.p;
    p.pair_cap = (uint32_t)pair_cap;
    p.tile_cnt = (uint32_t *)ctx->tile_cnt.p;
    p.tile_start = (uint32_t *)ctx->tile_start.p;
    p.heavy = (uint32_t *)ctx->heavy.p;
    p.tile_order = (uint4 *)ctx->tile_order.p;
    p.tile_shift = ctx->tile_edge == 32u ? 2u : 1u;
    p.heavy_min = p.tile_shift == 1u ? ctx->heavy_min : 0xFFFFFFFFu;  // the 16-wave path needs 16x16 tiles
    p.tile_lds_pad = lds_pad(tile_kernel_lds(ctx->tile_stage, p.tile_shift), p.tile_shift == 2u ? 16u : 4u,
                             ctx->tile_waves);
    p.setup_lds_pad = lds_pad(setup_kernel_lds(), 4u, ctx->setup_waves);
    p.heavy_blocks = ctx->heavy_blocks;
    p.pair_blocks = ctx->pair_blocks;
    p.tile_stage = ctx->tile_stage;
    ctx->last_pair_cap = pair_cap;
    p.ctr = ctx->ctr;
    p.stats = ctx->dstats;
    if (ctx->counting) {
        const size_t words = ((size_t)tri_cap + 31) / 32;
        RC_OR_RETURN(ensure(ctx, ctx->tri_mark, words * 4));
        HIP_OR_FAIL(ctx, launch_fill_u32((uint32_t *)ctx->tri_mark.p, 0u, words, ctx->stream));
        p.tri_mark = (uint32_t *)ctx->tri_mark.p;
        p.foot = ctx->foot;
    }
    // The first stage clears the scan status words, so epochs only need to differ between the
    // scans of one flush: they count from 1 in every flush, and a replayed graph reuses them.
    uint32_t flush_epoch = 0;
    p.scan_status = (uint32_t *)ctx->scan_part.p;
    p.scan_status_words = (uint32_t)(scan_scratch_bytes(scan_n) / 4);
    ScanScratch sc{ctx->scan_part.p, ctx->scan_part.cap, (uint32_t *)ctx->radix_hist.p, ctx->radix_hist.cap / 4,
                   &flush_epoch, &ctx->dstats->scan_errors, scan_wide_offset(scan_n)};

    // keys are global tile indices < total_tiles: one 8-bit pass per byte they use
    uint32_t bits = 0;
    while (bits < 32 && (1ull << bits) < total_tiles) ++bits;
    const uint32_t passes = std::max<uint32_t>(1u, (bits + 7u) / 8u);

    const bool timing = ctx->profiling && !ctx->capturing;
    // the reference's sample names for the work each stage covers (main.c:663, 699, 737, 916, 984, 1047)
    static const char *const kStageMarker[kStages] = {"input_assambler_stage+vertex_shader_stage",
                                                       "primitive_assembly_stage", "binner",
                                                       "rasterizer_stage+pixel_shader_stage"};
    Marker flush_marker("flush");
    for (int st = 0; st < kStages; ++st) {
        Marker stage_marker(kStageMarker[st]);
        hipEvent_t a = nullptr, b = nullptr;
        if (timing) {
            a = take_event(ctx);
            b = take_event(ctx);
            if (!a || !b) return fail(ctx, MLV_ERROR_DEVICE, "hipEventCreate failed");
            HIP_OR_FAIL(ctx, hipEventRecord(a, ctx->stream));
        }
        hipEvent_t tk[2] = {nullptr, nullptr};
        if (timing && st == kStages - 1) {
            tk[0] = take_event(ctx);
            tk[1] = take_event(ctx);
            if (!tk[0] || !tk[1]) return fail(ctx, MLV_ERROR_DEVICE, "hipEventCreate failed");
        }
        HIP_OR_FAIL(ctx, launch_stage(st, p, sc, passes, ctx->stream, tk[0] ? tk : nullptr));
        if (timing) {
            HIP_OR_FAIL(ctx, hipEventRecord(b, ctx->stream));
            ctx->ev_pending.push_back({st, {a, b}});
            if (tk[0]) ctx->ev_pending.push_back({kStages, {tk[0], tk[1]}});
        }
    }
    if (p.tri_mark)
        HIP_OR_FAIL(ctx, launch_popcount(p.tri_mark, ((size_t)tri_cap + 31) / 32, 0u, &ctx->foot->vis_tris, nullptr,
                                         ctx->stream));
    if (timing) {
        ctx->prof_draws += n_draws;
        ctx->prof_flushes++;
        if (ctx->ev_pending.size() > 4096) {
            HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
            resolve_events(ctx);
        }
    }
    return MLV_OK;
}

// Flush when a recorded draw targets the render target or depth target given.
static int flush_if_targets(mlv_context *ctx, const mlv_render_target *rt, const mlv_depth_stencil *ds) {
    for (auto &pd : ctx->pending)
        if ((rt && pd.key.rt == rt) || (ds && pd.key.ds == ds)) return flush(ctx);
    return MLV_OK;
}

// Reads the device-side counters and turns capacity overflows into an error and a
// larger arena for the next flushes.
static int check_device_stats(mlv_context *ctx) {
    DevStats s{};
    HIP_OR_FAIL(ctx, hipMemcpyAsync(&s, ctx->dstats, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    ctx->pair_hint = std::max<size_t>(ctx->pair_hint, (size_t)(s.max_pairs + s.max_pairs / 4));
    ctx->clip_hint = std::max<size_t>(ctx->clip_hint, (size_t)(s.max_clip_list + s.max_clip_list / 4));
    if (s.scan_errors) {
        (void)hipMemsetAsync(&ctx->dstats->scan_errors, 0, sizeof(uint32_t), ctx->stream);
        return fail(ctx, MLV_ERROR_DEVICE, "a scan's look-back gave up waiting; the flush's results are not valid");
    }
    if (s.capacity_errors > ctx->capacity_errors) {
        ctx->capacity_errors = s.capacity_errors;
        ctx->clip_hint *= 4;
        return fail(ctx, MLV_ERROR_CAPACITY, "a flush exceeded the clipper arena; capacity was raised, redraw");
    }
    return MLV_OK;
}

// ------------------------------------------------------------------ fixed scratch
//
// The scratch of mlv_diff_backward_noalloc: one fixed device array, which every GPU the library
// runs on holds (each its own copy at its own address) until the process ends. In host code the
// array's name is not a device address: each device's address is read once with
// hipGetSymbolAddress when the first context on that device is created, and only that address
// is passed to kernels.
constexpr size_t kFixedScratchBytes = 32ull << 20;
static __device__ __align__(256) uint8_t scratch[kFixedScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];

// With ctx->device current: the device's address of `scratch`, looked up once per device.
static void bind_fixed_scratch(mlv_context *ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[ctx->device]) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_err[ctx->device] = e;
    if (e == hipSuccess) g_scratch_ptr[ctx->device] = ptr;
}

extern "C" {

MLV_API int mlv_device_count(int *out) {
    if (!out) return MLV_ERROR_INVALID_ARGUMENT;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *out = n;
    return MLV_OK;
}

MLV_API int mlv_context_create(int device, mlv_context **out) {
    if (!out) return MLV_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MLV_ERROR_NO_DEVICE;
    if (device < 0 || device >= n) return MLV_ERROR_INVALID_ARGUMENT;
    mlv_context *ctx = new (std::nothrow) mlv_context();
    if (!ctx) return MLV_ERROR_OUT_OF_MEMORY;
    ctx->device = device;
    if (const char *lim = getenv("MLV_PAIR_CAPACITY")) ctx->pair_limit = (size_t)strtoull(lim, nullptr, 10);
    // flushes of another size than the default (smaller: a batch split exercised on small scenes)
    if (const char *lim = getenv("MLV_MAX_BATCH_TRIANGLES")) {
        const uint64_t v = strtoull(lim, nullptr, 10);
        if (v) ctx->batch_tris = ctx->batch_verts = std::min<uint64_t>(v, kBatchCeiling);
    }
    if (const char *lim = getenv("MLV_MAX_BATCH_DRAWS")) {
        const uint64_t v = strtoull(lim, nullptr, 10);
        if (v) ctx->batch_draws = std::min<uint64_t>(v, kMaxBatchDraws);
    }
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&ctx->ctr, sizeof(BatchCounters)) != hipSuccess ||
        hipMalloc(&ctx->dstats, sizeof(DevStats)) != hipSuccess ||
        hipMemset(ctx->dstats, 0, sizeof(DevStats)) != hipSuccess ||
        hipMemset(ctx->ctr, 0, sizeof(BatchCounters)) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        if (ctx->ctr) (void)hipFree(ctx->ctr);
        if (ctx->dstats) (void)hipFree(ctx->dstats);
        delete ctx;
        return MLV_ERROR_DEVICE;
    }
    bind_fixed_scratch(ctx);  // the device is current
    *out = ctx;
    return MLV_OK;
}

MLV_API int mlv_context_destroy(mlv_context *ctx) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    (void)flush(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    Arena *all[] = {&ctx->blk_draw,  &ctx->vblk_draw, &ctx->tblk_draw, &ctx->draws_dev, &ctx->views_dev,
                    &ctx->vpos,      &ctx->vattr,     &ctx->cnt,
                    &ctx->blocksum,  &ctx->blockoff,  &ctx->asm_base, &ctx->clip_list, &ctx->clip_off, &ctx->clip_rows,
                    &ctx->tvis,      &ctx->tshade,    &ctx->tbin,    &ctx->pair_cnt, &ctx->pair_off,  &ctx->keys0,
                    &ctx->vals0,     &ctx->keys1,     &ctx->vals1,    &ctx->tile_cnt,  &ctx->tile_start,
                    &ctx->scan_part, &ctx->radix_hist, &ctx->heavy, &ctx->tile_order, &ctx->d_shpart, &ctx->d_tkey, &ctx->d_tval,
                    &ctx->d_tkey2,   &ctx->d_tval2,   &ctx->d_tpay,   &ctx->d_vkey,    &ctx->d_vval,
                    &ctx->d_vkey2,   &ctx->d_vval2,   &ctx->d_vpay,   &ctx->d_part,    &ctx->d_hist,
                    &ctx->d_vposd,   &ctx->d_posacc,  &ctx->b_tables, &ctx->b_posacc, &ctx->b_shpart, &ctx->present_rgb, &ctx->cameras,
                    &ctx->tri_mark, &ctx->diff_mark, &ctx->b_defer, &ctx->cmark};
    for (Arena *a : all) release(*a);
    free_retired(ctx);
    resolve_events(ctx);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (int i = 0; i < kStagingSlots; ++i) {
        if (ctx->staging[i]) (void)hipHostFree(ctx->staging[i]);
        if (ctx->staging_ev[i]) (void)hipEventDestroy(ctx->staging_ev[i]);
    }
    (void)hipFree(ctx->ctr);
    (void)hipFree(ctx->dstats);
    if (ctx->foot) (void)hipFree(ctx->foot);
    for (mlv_texture2d *t : ctx->marked) {  // textures outliving their context keep no bitmap
        (void)hipFree(t->mark);
        t->mark = nullptr;
        t->mark_words = 0;
    }
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MLV_OK;
}

MLV_API const char *mlv_get_last_error(const mlv_context *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

MLV_API void *mlv_context_stream(mlv_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

MLV_API int mlv_synchronize(mlv_context *ctx) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    return check_device_stats(ctx);
}

// ------------------------------------------------------------------ resources

MLV_API int mlv_buffer_create(mlv_context *ctx, int kind, size_t bytes, const void *host_init, mlv_buffer **out) {
    if (!ctx || !out || bytes == 0 || kind < MLV_BUFFER_VERTEX || kind > MLV_BUFFER_CONSTANT)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "mlv_buffer_create: bad argument");
    RC_OR_RETURN(set_device(ctx));
    void *p = nullptr;
    HIP_OR_FAIL(ctx, hipMalloc(&p, (bytes + 15) & ~(size_t)15));
    if (host_init) {
        hipError_t e = hipMemcpyAsync(p, host_init, bytes, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            (void)hipFree(p);
            return fail(ctx, MLV_ERROR_DEVICE, "mlv_buffer_create upload: %s", hipGetErrorString(e));
        }
    }
    mlv_buffer *b = new (std::nothrow) mlv_buffer{p, bytes, kind, true};
    if (!b) {
        (void)hipFree(p);
        return MLV_ERROR_OUT_OF_MEMORY;
    }
    *out = b;
    return MLV_OK;
}

MLV_API int mlv_buffer_wrap(mlv_context *ctx, int kind, void *device_ptr, size_t bytes, mlv_buffer **out) {
    if (!ctx || !out || !device_ptr || bytes == 0 || kind < MLV_BUFFER_VERTEX || kind > MLV_BUFFER_CONSTANT)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "mlv_buffer_wrap: bad argument");
    mlv_buffer *b = new (std::nothrow) mlv_buffer{device_ptr, bytes, kind, false};
    if (!b) return MLV_ERROR_OUT_OF_MEMORY;
    *out = b;
    return MLV_OK;
}

MLV_API int mlv_buffer_update(mlv_context *ctx, mlv_buffer *buf, size_t offset, size_t bytes, const void *host_src) {
    if (!ctx || !buf || !host_src || offset + bytes > buf->bytes)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "mlv_buffer_update: bad argument");
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));  // recorded draws read the old contents
    HIP_OR_FAIL(ctx, hipMemcpyAsync((uint8_t *)buf->ptr + offset, host_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    return MLV_OK;
}

MLV_API int mlv_buffer_readback(mlv_context *ctx, const mlv_buffer *buf, size_t offset, size_t bytes, void *host_dst) {
    if (!ctx || !buf || !host_dst || offset + bytes > buf->bytes)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "mlv_buffer_readback: bad argument");
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    HIP_OR_FAIL(ctx, hipMemcpyAsync(host_dst, (const uint8_t *)buf->ptr + offset, bytes, hipMemcpyDeviceToHost,
                                    ctx->stream));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    return check_device_stats(ctx);
}

MLV_API int mlv_buffer_destroy(mlv_context *ctx, mlv_buffer *buf) {
    if (!ctx || !buf) return MLV_ERROR_INVALID_ARGUMENT;
    (void)flush(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    if (buf->owned) (void)hipFree(buf->ptr);
    for (auto &c : ctx->vs_cb)
        if (c == buf) c = nullptr;
    for (auto &c : ctx->ps_cb)
        if (c == buf) c = nullptr;
    if (ctx->ib == buf) ctx->ib = nullptr;
    if (ctx->vb == buf) ctx->vb = nullptr;
    delete buf;
    return MLV_OK;
}

MLV_API void *mlv_buffer_device_ptr(const mlv_buffer *buf) { return buf ? buf->ptr : nullptr; }

static size_t texel_bytes(int fmt) {
    switch (fmt) {
    case MLV_FORMAT_RGBA8_UNORM: return 4;
    case MLV_FORMAT_RGBA32F: return 16;
    case MLV_FORMAT_RGBA16F: return 8;
    default: return 0;
    }
}

static bool is_float_tex(const mlv_texture2d *t) {
    return t && (t->format == MLV_FORMAT_RGBA32F || t->format == MLV_FORMAT_RGBA16F);
}

static TexDesc tex_desc(const mlv_texture2d *t) {
    TexDesc d{};
    if (t) {
        d.data = t->ptr;
        d.width = t->width;
        d.height = t->height;
        d.format = t->format;
        d.mip_levels = t->mip_levels;
        d.mark = t->mark;
    }
    return d;
}

// Footprint counting: the texture's bitmap (one bit per texel of the chain), allocated cleared
// the first time a counted draw samples it.
static int ensure_mark(mlv_context *ctx, mlv_texture2d *t) {
    if (!t || t->mark) return MLV_OK;
    const size_t words = (t->bytes / texel_bytes((int)t->format) + 31) / 32;
    HIP_OR_FAIL(ctx, hipMalloc(&t->mark, words * 4));
    t->mark_words = words;
    HIP_OR_FAIL(ctx, launch_fill_u32(t->mark, 0u, words, ctx->stream));
    ctx->marked.push_back(t);
    return MLV_OK;
}

// Texels of the first `levels` levels of a width x height chain (level l is max(w >> l, 1) x
// max(h >> l, 1)); 0 when levels exceeds the full chain's count.
static size_t chain_texels(uint32_t width, uint32_t height, uint32_t levels) {
    size_t n = 0;
    uint32_t w = width, h = height;
    for (uint32_t l = 0; l < levels; ++l) {
        n += (size_t)w * h;
        if (w == 1u && h == 1u && l + 1u < levels) return 0;  // no level after 1x1
        w = w > 1u ? w >> 1 : 1u;
        h = h > 1u ? h >> 1 : 1u;
    }
    return n;
}

MLV_API int mlv_texture2d_create(mlv_context *ctx, uint32_t width, uint32_t height, int format, uint32_t mip_levels,
                                 const void *host_data, mlv_texture2d **out) {
    const size_t tb = texel_bytes(format);
    if (!ctx || !out || !width || !height || !tb || width > 65536u || height > 65536u)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "mlv_texture2d_create: bad argument");
    if (mip_levels == 0) mip_levels = 1;
    const size_t texels = mip_levels <= kMaxMipLevels ? chain_texels(width, height, mip_levels) : 0;
    if (!texels)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "mlv_texture2d_create: %u levels do not fit a %ux%u texture",
                    mip_levels, width, height);
    RC_OR_RETURN(set_device(ctx));
    const size_t bytes = texels * tb;
    void *p = nullptr;
    HIP_OR_FAIL(ctx, hipMalloc(&p, bytes));
    if (host_data) {
        hipError_t e = hipMemcpyAsync(p, host_data, bytes, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            (void)hipFree(p);
            return fail(ctx, MLV_ERROR_DEVICE, "texture upload: %s", hipGetErrorString(e));
        }
    }
    mlv_texture2d *t = new (std::nothrow) mlv_texture2d{p, width, height, (uint32_t)format, mip_levels, bytes};
    if (!t) {
        (void)hipFree(p);
        return MLV_ERROR_OUT_OF_MEMORY;
    }
    *out = t;
    return MLV_OK;
}

MLV_API int mlv_texture2d_generate_mips(mlv_context *ctx, mlv_texture2d *tex) {
    if (!ctx || !tex) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "texture2d_generate_mips: bad argument");
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));  // recorded draws read the texture as it was
    TexDesc d = tex_desc(tex);
    d.mark = nullptr;
    HIP_OR_FAIL(ctx, launch_generate_mips(d, ctx->stream));
    return MLV_OK;
}

MLV_API int mlv_texture2d_srgb_to_linear(mlv_context *ctx, mlv_texture2d *tex) {
    if (!ctx || !tex) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "texture2d_srgb_to_linear: bad argument");
    if (tex->format != MLV_FORMAT_RGBA8_UNORM)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "texture2d_srgb_to_linear: RGBA8 textures only (main.c:546)");
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));  // recorded draws read the texture as it was
    HIP_OR_FAIL(ctx, launch_srgb_to_linear_rgba8((uint32_t *)tex->ptr, tex->bytes / 4, ctx->stream));  // every level
    return MLV_OK;
}

MLV_API int mlv_texture2d_readback(mlv_context *ctx, const mlv_texture2d *tex, void *host_dst) {
    if (!ctx || !tex || !host_dst) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "texture2d_readback: bad argument");
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    HIP_OR_FAIL(ctx, hipMemcpyAsync(host_dst, tex->ptr, tex->bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    return MLV_OK;
}

MLV_API int mlv_texture2d_destroy(mlv_context *ctx, mlv_texture2d *tex) {
    if (!ctx || !tex) return MLV_ERROR_INVALID_ARGUMENT;
    (void)flush(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    if (tex->mark && ctx->foot) {  // a counted texture's distinct texels join the totals before it goes
        (void)launch_popcount(tex->mark, tex->mark_words, (uint32_t)texel_bytes((int)tex->format), &ctx->foot->texels,
                              &ctx->foot->tex_bytes, ctx->stream);
        (void)hipStreamSynchronize(ctx->stream);
    }
    (void)hipFree(tex->ptr);
    if (tex->mark) (void)hipFree(tex->mark);
    ctx->marked.erase(std::remove(ctx->marked.begin(), ctx->marked.end(), tex), ctx->marked.end());
    for (auto &s : ctx->vs_srv)
        if (s == tex) s = nullptr;
    for (auto &s : ctx->ps_srv)
        if (s == tex) s = nullptr;
    delete tex;
    return MLV_OK;
}

MLV_API void *mlv_texture2d_device_ptr(const mlv_texture2d *tex) { return tex ? tex->ptr : nullptr; }

MLV_API int mlv_render_target_create(mlv_context *ctx, uint32_t width, uint32_t height, mlv_render_target **out) {
    if (!ctx || !out || !width || !height || width > 16384u || height > 16384u)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "mlv_render_target_create: bad size");
    RC_OR_RETURN(set_device(ctx));
    uint32_t *p = nullptr;
    HIP_OR_FAIL(ctx, hipMalloc(&p, (size_t)width * height * 4));
    mlv_render_target *rt = new (std::nothrow) mlv_render_target{p, width, height};
    if (!rt) {
        (void)hipFree(p);
        return MLV_ERROR_OUT_OF_MEMORY;
    }
    *out = rt;
    return MLV_OK;
}

MLV_API int mlv_render_target_destroy(mlv_context *ctx, mlv_render_target *rt) {
    if (!ctx || !rt) return MLV_ERROR_INVALID_ARGUMENT;
    (void)flush(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(rt->ptr);
    if (ctx->rt == rt) ctx->rt = nullptr;
    delete rt;
    return MLV_OK;
}

MLV_API void *mlv_render_target_device_ptr(const mlv_render_target *rt) { return rt ? rt->ptr : nullptr; }

MLV_API int mlv_depth_stencil_create(mlv_context *ctx, uint32_t width, uint32_t height, mlv_depth_stencil **out) {
    if (!ctx || !out || !width || !height || width > 16384u || height > 16384u)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "mlv_depth_stencil_create: bad size");
    RC_OR_RETURN(set_device(ctx));
    const uint32_t bx = (width + kBin - 1) / kBin, by = (height + kBin - 1) / kBin;
    float *p = nullptr, *h = nullptr;
    HIP_OR_FAIL(ctx, hipMalloc(&p, (size_t)width * height * 4));
    if (hipMalloc(&h, (size_t)bx * by * 4) != hipSuccess) {
        (void)hipFree(p);
        return fail(ctx, MLV_ERROR_OUT_OF_MEMORY, "depth hi-z allocation failed");
    }
    mlv_depth_stencil *ds = new (std::nothrow) mlv_depth_stencil{p, h, width, height, bx, by, MLV_DEPTH_GREATER_EQUAL};
    if (!ds) {
        (void)hipFree(p);
        (void)hipFree(h);
        return MLV_ERROR_OUT_OF_MEMORY;
    }
    // a fresh depth target behaves as if cleared: hierarchical-Z starts at the reset value
    if (launch_fill_f32(h, 0.0f, (size_t)bx * by, ctx->stream) != hipSuccess) {
        (void)hipFree(p);
        (void)hipFree(h);
        delete ds;
        return fail(ctx, MLV_ERROR_DEVICE, "hi-z init failed");
    }
    *out = ds;
    return MLV_OK;
}

MLV_API int mlv_depth_stencil_destroy(mlv_context *ctx, mlv_depth_stencil *ds) {
    if (!ctx || !ds) return MLV_ERROR_INVALID_ARGUMENT;
    (void)flush(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ds->ptr);
    (void)hipFree(ds->hiz);
    if (ctx->ds == ds) ctx->ds = nullptr;
    delete ds;
    return MLV_OK;
}

MLV_API void *mlv_depth_stencil_device_ptr(const mlv_depth_stencil *ds) { return ds ? ds->ptr : nullptr; }

// ------------------------------------------------------------------ state

MLV_API int mlv_ia_set_index_buffer(mlv_context *ctx, mlv_buffer *ib, uint32_t offset_bytes) {
    if (!ctx || (ib && (ib->kind != MLV_BUFFER_INDEX_U32 || offset_bytes % 4 || offset_bytes > ib->bytes)))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "mlv_ia_set_index_buffer: bad argument");
    ctx->ib = ib;
    ctx->ib_off = offset_bytes;
    return MLV_OK;
}

MLV_API int mlv_ia_set_vertex_buffer(mlv_context *ctx, mlv_buffer *vb, uint32_t stride, uint32_t offset_bytes) {
    if (!ctx || (vb && (vb->kind != MLV_BUFFER_VERTEX || stride < 32 || stride % 4 || offset_bytes % 4 ||
                        offset_bytes > vb->bytes)))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT,
                    "mlv_ia_set_vertex_buffer: stride must be >= 32 and a multiple of 4 (the reference layouts read 8 floats)");
    ctx->vb = vb;
    ctx->vb_stride = stride;
    ctx->vb_off = offset_bytes;
    return MLV_OK;
}

MLV_API int mlv_ia_set_primitive_topology(mlv_context *ctx, int topology) {
    if (!ctx || (topology != MLV_TOPOLOGY_UNDEFINED && topology != MLV_TOPOLOGY_TRIANGLELIST))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "unknown topology");
    ctx->topology = topology;
    return MLV_OK;
}

MLV_API int mlv_vs_set_shader(mlv_context *ctx, int vertex_shader) {
    if (!ctx || vertex_shader < MLV_VS_BASIC || vertex_shader > MLV_VS_FULLSCREEN)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "unknown vertex shader");
    ctx->vs = vertex_shader;
    return MLV_OK;
}

MLV_API int mlv_vs_set_constant_buffer(mlv_context *ctx, uint32_t slot, mlv_buffer *cb) {
    if (!ctx || slot >= 16 || (cb && cb->kind != MLV_BUFFER_CONSTANT))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "mlv_vs_set_constant_buffer: bad argument");
    ctx->vs_cb[slot] = cb;
    return MLV_OK;
}

MLV_API int mlv_vs_set_shader_resource(mlv_context *ctx, uint32_t slot, mlv_texture2d *tex) {
    if (!ctx || slot >= 16) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "bad shader resource slot");
    ctx->vs_srv[slot] = tex;
    return MLV_OK;
}

MLV_API int mlv_rs_set_viewport(mlv_context *ctx, const mlv_viewport *vp) {
    if (!ctx || !vp || !(vp->width > 0.0f) || !(vp->height > 0.0f))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "bad viewport");
    ctx->vp = *vp;
    ctx->vp_set = true;
    return MLV_OK;
}

MLV_API int mlv_ps_set_shader(mlv_context *ctx, int pixel_shader) {
    if (!ctx || pixel_shader < MLV_PS_BASIC || pixel_shader > MLV_PS_NONE)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "unknown pixel shader");
    ctx->ps = pixel_shader;
    return MLV_OK;
}

MLV_API int mlv_ps_set_shader_resource(mlv_context *ctx, uint32_t slot, mlv_texture2d *tex) {
    if (!ctx || slot >= 16) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "bad shader resource slot");
    ctx->ps_srv[slot] = tex;
    return MLV_OK;
}

MLV_API int mlv_ps_set_constant_buffer(mlv_context *ctx, uint32_t slot, mlv_buffer *cb) {
    if (!ctx || slot >= 16) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "bad constant buffer slot");
    ctx->ps_cb[slot] = cb;
    return MLV_OK;
}

MLV_API int mlv_om_set_render_targets(mlv_context *ctx, mlv_render_target *rt, mlv_depth_stencil *ds) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    if (rt && ds && (rt->width != ds->width || rt->height != ds->height))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "render target and depth target sizes differ");
    ctx->rt = rt;
    ctx->ds = ds;
    return MLV_OK;
}

MLV_API int mlv_rs_set_cull(mlv_context *ctx, int cull_mode) {
    if (!ctx || (cull_mode != MLV_CULL_BACK && cull_mode != MLV_CULL_NONE && cull_mode != MLV_CULL_FRONT))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "unknown cull mode");
    ctx->cull = cull_mode;
    return MLV_OK;
}

MLV_API int mlv_rs_set_fill_rule(mlv_context *ctx, int fill_rule) {
    if (!ctx || (fill_rule != MLV_FILL_INCLUSIVE && fill_rule != MLV_FILL_TOP_LEFT))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "unknown fill rule");
    ctx->fill = fill_rule;
    return MLV_OK;
}

MLV_API int mlv_om_set_depth_func(mlv_context *ctx, int depth_func) {
    if (!ctx || (depth_func != MLV_DEPTH_GREATER_EQUAL && depth_func != MLV_DEPTH_LESS_EQUAL))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "unknown depth func");
    ctx->depth_func = depth_func;
    return MLV_OK;
}

MLV_API int mlv_om_set_visibility_buffer(mlv_context *ctx, uint32_t *device_tri_ids, uint32_t *device_draw_tags) {
    if (!ctx || ((device_tri_ids == nullptr) != (device_draw_tags == nullptr)))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "visibility buffer needs both arrays or neither");
    ctx->vis_tri = device_tri_ids;
    ctx->vis_draw = device_draw_tags;
    return MLV_OK;
}

MLV_API int mlv_om_set_primitive_id_buffer(mlv_context *ctx, uint32_t *device_prim_ids) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    ctx->vis_prim = device_prim_ids;
    return MLV_OK;
}

// ------------------------------------------------------------------ commands

// encode_color_as_u32 (math.h:322-324): truncating conversion per channel
static uint32_t trunc_u32(float f) {
    if (!(f > -9.2e18f && f < 9.2e18f)) return 0u;
    return (uint32_t)(int64_t)f;
}

static int exec_clear_color(mlv_context *ctx, const Cmd &c) {
    RC_OR_RETURN(flush_if_targets(ctx, c.rt, nullptr));
    const size_t n = (size_t)c.rt->width * c.rt->height;
    FillJobs f{};  // colour and the visibility buffers in one launch
    f.add(c.rt->ptr, c.value, n);
    if (c.vis_tri) {
        f.add(c.vis_tri, 0xFFFFFFFFu, n);
        f.add(c.vis_draw, 0xFFFFFFFFu, n);
    }
    if (c.vis_prim) f.add(c.vis_prim, 0xFFFFFFFFu, n);
    HIP_OR_FAIL(ctx, launch_fill(f, ctx->stream));
    return MLV_OK;
}

static int exec_clear_depth(mlv_context *ctx, const Cmd &c) {
    RC_OR_RETURN(flush_if_targets(ctx, nullptr, c.ds));
    // main.c:1212-1214: every tile minimum back to 0.0 (1.0 for the LESS_EQUAL mirror); depth and
    // hierarchical-Z in one launch
    const float h = c.depth_func == MLV_DEPTH_LESS_EQUAL ? 1.0f : 0.0f;
    FillJobs f{};
    f.add(c.ds->ptr, __builtin_bit_cast(uint32_t, c.depth), (size_t)c.ds->width * c.ds->height);
    f.add(c.ds->hiz, __builtin_bit_cast(uint32_t, h), (size_t)c.ds->bins_x * c.ds->bins_y);
    HIP_OR_FAIL(ctx, launch_fill(f, ctx->stream));
    c.ds->hiz_mode = c.depth_func;
    return MLV_OK;
}

MLV_API int mlv_clear_render_target_view(mlv_context *ctx, mlv_render_target *rt, const float rgba[4]) {
    if (!ctx || !rt || !rgba) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "clear_render_target_view: bad argument");
    Marker marker("clear_render_target_view");
    RC_OR_RETURN(set_device(ctx));
    Cmd c;
    c.kind = kCmdClearColor;
    c.rt = rt;
    c.value = (trunc_u32(rgba[3] * 255.f) << 24) + (trunc_u32(rgba[2] * 255.f) << 16) +
              (trunc_u32(rgba[1] * 255.f) << 8) + trunc_u32(rgba[0] * 255.f);
    if (ctx->vis_tri && ctx->rt == rt) {
        c.vis_tri = ctx->vis_tri;
        c.vis_draw = ctx->vis_draw;
    }
    if (ctx->vis_prim && ctx->rt == rt) c.vis_prim = ctx->vis_prim;
    if (ctx->recording) {
        ctx->rec.push_back(c);
        return MLV_OK;
    }
    return exec_clear_color(ctx, c);
}

MLV_API int mlv_clear_depth_stencil_view(mlv_context *ctx, mlv_depth_stencil *ds, float depth) {
    if (!ctx || !ds) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "clear_depth_stencil_view: bad argument");
    Marker marker("clear_depth_stencil_view");
    RC_OR_RETURN(set_device(ctx));
    Cmd c;
    c.kind = kCmdClearDepth;
    c.ds = ds;
    c.depth = depth;
    c.depth_func = ctx->depth_func;
    if (ctx->recording) {
        ctx->rec.push_back(c);
        return MLV_OK;
    }
    return exec_clear_depth(ctx, c);
}

// Adds a draw to the pending batch, flushing first where the batch cannot take it.
static int submit_draw(mlv_context *ctx, const Cmd &c) {
    const PendingDraw &pd = c.pd;
    mlv_depth_stencil *ds = pd.key.ds;
    // one depth function per target within a flush, and hierarchical-Z built for it
    for (auto &q : ctx->pending)
        if ((q.key.rt == pd.key.rt || q.key.ds == pd.key.ds) && !(q.key == pd.key)) {
            RC_OR_RETURN(flush(ctx));
            break;
        }
    // A flush shades each pixel's final winner only. A depth-only draw's winner writes no colour,
    // so the colour a pixel keeps is that of the last shaded fragment before it: a depth-only draw
    // after a shaded draw on the same target starts a new flush (depth-only draws before shaded
    // ones need none, nor do flushes of depth-only draws alone).
    if (pd.rec.ps == PS_NONE)
        for (auto &q : ctx->pending)
            if (q.key.rt == pd.key.rt && q.rec.ps != PS_NONE) {
                RC_OR_RETURN(flush(ctx));
                break;
            }
    if (ds->hiz_mode != pd.key.depth_func) {
        RC_OR_RETURN(flush_if_targets(ctx, nullptr, ds));
        const float h = pd.key.depth_func == MLV_DEPTH_LESS_EQUAL ? HUGE_VALF : -HUGE_VALF;  // rejects nothing
        HIP_OR_FAIL(ctx, launch_fill_f32(ds->hiz, h, (size_t)ds->bins_x * ds->bins_y, ctx->stream));
        ds->hiz_mode = pd.key.depth_func;
    }
    if (ctx->pending.size() >= ctx->batch_draws || ctx->pending_tris + align256(c.T) > ctx->batch_tris ||
        ctx->pending_verts + align256(c.V) > ctx->batch_verts)
        RC_OR_RETURN(flush(ctx));
    ctx->pending.push_back(pd);
    ctx->pending_tris += align256(c.T);
    ctx->pending_verts += align256(c.V);
    return MLV_OK;
}

MLV_API int mlv_draw_indexed(mlv_context *ctx, uint32_t index_count, uint32_t start_index, int32_t base_vertex) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    Marker marker("draw_indexed");
    if (ctx->topology != MLV_TOPOLOGY_TRIANGLELIST)
        return fail(ctx, MLV_ERROR_INVALID_STATE, "draw_indexed: topology must be TRIANGLELIST (main.c:666)");
    if (index_count % 3u)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "draw_indexed: index_count must be a multiple of 3 (main.c:1230)");
    if (!ctx->ib || !ctx->vb) return fail(ctx, MLV_ERROR_INVALID_STATE, "draw_indexed: index/vertex buffer not bound");
    if (!ctx->vs || !ctx->ps) return fail(ctx, MLV_ERROR_INVALID_STATE, "draw_indexed: shaders not set");
    if (!ctx->rt || !ctx->ds) return fail(ctx, MLV_ERROR_INVALID_STATE, "draw_indexed: render targets not set");
    if (!ctx->vp_set) return fail(ctx, MLV_ERROR_INVALID_STATE, "draw_indexed: viewport not set");
    const uint32_t W = ctx->rt->width, H = ctx->rt->height;
    // read_tile/write_tile use viewport.width as the row pitch (main.c:582,596): the viewport
    // must cover the render target exactly.
    if (ctx->vp.width != (float)W || ctx->vp.height != (float)H)
        return fail(ctx, MLV_ERROR_INVALID_STATE, "draw_indexed: viewport must equal the render target size");
    if ((uint64_t)ctx->ib_off + 4ull * ((uint64_t)start_index + index_count) > ctx->ib->bytes)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "draw_indexed: index range past the index buffer");
    const bool needs_cb = ctx->vs != MLV_VS_PASSTHROUGH;
    if (needs_cb && (!ctx->vs_cb[0] || ctx->vs_cb[0]->bytes < 48 * sizeof(float)))
        return fail(ctx, MLV_ERROR_INVALID_STATE, "draw_indexed: constant buffer 0 (3 float4x4) not bound");
    if (ctx->vs == MLV_VS_VERTEX_LIGHTING && !is_float_tex(ctx->vs_srv[0]))
        return fail(ctx, MLV_ERROR_INVALID_STATE, "vertex_lighting_vs needs a float RGBA texture in slot 0");
    if (ctx->ps == MLV_PS_ENV_LIGHTING && !is_float_tex(ctx->ps_srv[0]))
        return fail(ctx, MLV_ERROR_INVALID_STATE, "env_lighting_ps needs a float RGBA texture in slot 0");
    if (ctx->ps == MLV_PS_BASIC && !ctx->ps_srv[0])
        return fail(ctx, MLV_ERROR_INVALID_STATE, "basic_ps needs a texture in slot 0");
    if (ctx->ps == MLV_PS_BASIC_LAMBERT &&
        (!ctx->ps_srv[0] || !ctx->ps_cb[0] || ctx->ps_cb[0]->bytes < kPsConstants * sizeof(float)))
        return fail(ctx, MLV_ERROR_INVALID_STATE,
                    "the Lambert shader needs a texture in slot 0 and a 12-float constant buffer in slot 0");
    if (ctx->ps == MLV_PS_ENV_TEXTURED && (!is_float_tex(ctx->ps_srv[0]) || !ctx->ps_srv[1]))
        return fail(ctx, MLV_ERROR_INVALID_STATE,
                    "the textured environment shader needs a float RGBA map in slot 0 and a texture in slot 1");
    RC_OR_RETURN(set_device(ctx));

    const uint32_t T = index_count / 3u;
    const uint64_t vbytes = ctx->vb->bytes - ctx->vb_off;
    const uint32_t V = (uint32_t)std::min<uint64_t>(vbytes / ctx->vb_stride, 0x7FFFFFFFull);
    if (V > kMaxBatchVertices || T > kMaxBatchTriangles)
        return fail(ctx, MLV_ERROR_CAPACITY, "draw_indexed: at most %llu triangles and vertices per draw",
                    (unsigned long long)kMaxBatchTriangles);
    ctx->vertex_count += index_count;
    ctx->input_tris += T;
    ctx->draws++;
    const uint32_t draw_tag = ctx->draw_counter++;
    if (T == 0) return MLV_OK;

    mlv_depth_stencil *ds = ctx->ds;
    Cmd c;
    c.kind = kCmdDraw;
    c.T = T;
    c.V = V;
    PendingDraw &pd = c.pd;
    pd.key = ViewKey{ctx->rt, ds, ctx->vis_tri, ctx->vis_draw, ctx->vis_prim, ctx->depth_func};
    DrawRec &r = pd.rec;
    r.ib = (const uint32_t *)((const uint8_t *)ctx->ib->ptr + ctx->ib_off) + start_index;
    r.vb = (const uint8_t *)ctx->vb->ptr + ctx->vb_off;
    r.cb = ctx->vs_cb[0] ? (const float *)ctx->vs_cb[0]->ptr : nullptr;
    if (ctx->counting) {
        RC_OR_RETURN(ensure_mark(ctx, ctx->vs_srv[0]));
        RC_OR_RETURN(ensure_mark(ctx, ctx->ps_srv[0]));
        RC_OR_RETURN(ensure_mark(ctx, ctx->ps_srv[1]));
    }
    r.vs_tex = tex_desc(ctx->vs_srv[0]);
    r.ps_tex = tex_desc(ctx->ps_srv[0]);
    r.ps_tex1 = tex_desc(ctx->ps_srv[1]);
    if (!ctx->counting) r.vs_tex.mark = r.ps_tex.mark = r.ps_tex1.mark = nullptr;
    r.ps_cb = ctx->ps_cb[0] ? (const float *)ctx->ps_cb[0]->ptr : nullptr;
    screen_matrix(ctx->vp, r.screen);
    r.stride = ctx->vb_stride;
    r.vertex_count = V;
    r.tri_count = T;
    r.base_vertex = base_vertex;
    r.vs = ctx->vs;
    r.ps = ctx->ps;
    r.depth_func = ctx->depth_func;
    r.cull = ctx->cull;
    r.fill = ctx->fill;
    // the conservative in-draw depth cull and the per-bin edge screening rely on exact edge
    // values: |E| < 2^31 at every pixel of every bin
    // (bins reach up to 7 pixels past the right and bottom edges)
    r.depth_cull = (ctx->vp.top_left_x == 0.0f && ctx->vp.top_left_y == 0.0f &&
                    2.0 * (16.0 * (W + 8) + 2.0) * (16.0 * (H + 8) + 2.0) < 2147483648.0 * 0.99)
                       ? 1
                       : 0;
    r.draw_tag = draw_tag;
    if (ctx->recording) {
        ctx->rec.push_back(c);
        return MLV_OK;
    }
    return submit_draw(ctx, c);
}

// ------------------------------------------------------------------ graphs

// The device's additive stat counters (what mlv_get_stats reports from the device) around runs
// of a recorded frame that are not frames of the caller's: the sizing retries of a capture and
// the timing runs of the autotuner. The frame counts once, like an eager frame.
static int read_dstats(mlv_context *ctx, DevStats &s) {
    HIP_OR_FAIL(ctx, hipMemcpyAsync(&s, ctx->dstats, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    return MLV_OK;
}

// Sets the additive counters to base + (after - before): the counts from before the runs plus
// those of the one run between `before` and `after`. High-water marks and error counters are
// left as the runs left them.
static int keep_one_run_stats(mlv_context *ctx, const DevStats &base, const DevStats &before, const DevStats &after) {
    DevStats s{};
    RC_OR_RETURN(read_dstats(ctx, s));
    s.assembled = base.assembled + (after.assembled - before.assembled);
    s.active8 = base.active8 + (after.active8 - before.active8);
    s.pairs8 = base.pairs8 + (after.pairs8 - before.pairs8);
    s.oob = base.oob + (after.oob - before.oob);
    s.overflow_draws = base.overflow_draws + (after.overflow_draws - before.overflow_draws);
    for (int i = 0; i < 64; ++i) {
        s.active8_part[i] = base.active8_part[i] + (after.active8_part[i] - before.active8_part[i]);
        s.pairs8_part[i] = base.pairs8_part[i] + (after.pairs8_part[i] - before.pairs8_part[i]);
    }
    HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->dstats, &s, sizeof s, hipMemcpyHostToDevice, ctx->stream));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    return MLV_OK;
}

static int exec_diff_cmd(mlv_context *ctx, const Cmd &c);

static int run_recorded(mlv_context *ctx, const std::vector<Cmd> &rec) {
    for (const Cmd &c : rec) {
        switch (c.kind) {
        case kCmdClearColor: RC_OR_RETURN(exec_clear_color(ctx, c)); break;
        case kCmdClearDepth: RC_OR_RETURN(exec_clear_depth(ctx, c)); break;
        case kCmdDraw: RC_OR_RETURN(submit_draw(ctx, c)); break;
        default:
            RC_OR_RETURN(flush(ctx));  // the passes read what the recorded draws wrote
            RC_OR_RETURN(exec_diff_cmd(ctx, c));
            break;
        }
    }
    return flush(ctx);
}

MLV_API int mlv_graph_capture_begin(mlv_context *ctx) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    if (ctx->recording) return fail(ctx, MLV_ERROR_INVALID_STATE, "graph_capture_begin: already recording");
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));  // earlier draws stay outside the graph
    ctx->rec.clear();
    ctx->recording = true;
    return MLV_OK;
}

MLV_API int mlv_graph_capture_end(mlv_context *ctx, mlv_graph **out) {
    if (!ctx || !out) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "graph_capture_end: bad argument");
    *out = nullptr;
    if (!ctx->recording) return fail(ctx, MLV_ERROR_INVALID_STATE, "graph_capture_end without graph_capture_begin");
    ctx->recording = false;
    std::vector<Cmd> rec;
    rec.swap(ctx->rec);
    RC_OR_RETURN(set_device(ctx));
    // 1. run the commands once: sizes every arena and yields this frame's results
    int rc = MLV_OK;
    DevStats base{}, before{}, after{};
    RC_OR_RETURN(read_dstats(ctx, base));
    for (int attempt = 0; attempt < 3; ++attempt) {
        RC_OR_RETURN(read_dstats(ctx, before));
        ctx->sizing = true;
        ctx->cap_sizes.clear();
        rc = run_recorded(ctx, rec);
        ctx->sizing = false;
        if (rc == MLV_OK) {
            HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
            rc = check_device_stats(ctx);
        }
        if (rc != MLV_ERROR_CAPACITY) break;  // a capacity miss raised the capacity: size again
    }
    if (rc != MLV_OK) return rc;
    RC_OR_RETURN(read_dstats(ctx, after));
    RC_OR_RETURN(keep_one_run_stats(ctx, base, before, after));  // the frame's first execution counts once
    // 2. run them again under stream capture, uploading tables from buffers the graph owns
    mlv_graph *g = new (std::nothrow) mlv_graph();
    if (!g) return MLV_ERROR_OUT_OF_MEMORY;
    auto drop = [&](int code) {
        for (void *h : g->pinned) (void)hipHostFree(h);
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return code;
    };
    for (size_t b : ctx->cap_sizes) {
        void *h = nullptr;
        if (hipHostMalloc(&h, std::max<size_t>(b, 64), hipHostMallocDefault) != hipSuccess)
            return drop(fail(ctx, MLV_ERROR_OUT_OF_MEMORY, "graph_capture_end: pinned table buffer"));
        g->pinned.push_back(h);
    }
    if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return drop(fail(ctx, MLV_ERROR_DEVICE, "hipStreamBeginCapture failed"));
    ctx->capturing = true;
    ctx->cap_bufs = &g->pinned;
    ctx->cap_next = 0;
    rc = run_recorded(ctx, rec);
    ctx->capturing = false;
    ctx->cap_bufs = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g->graph);
    if (rc != MLV_OK) return drop(rc);
    if (e != hipSuccess) return drop(fail(ctx, MLV_ERROR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e)));
    const hipError_t ei = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) return drop(fail(ctx, MLV_ERROR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei)));
    for (const Cmd &c : rec)
        if (c.kind == kCmdDraw) {
            g->draws++;
            g->input_tris += c.T;
            g->vertex_count += 3ull * c.T;
        }
    ctx->graphs_alive++;
    *out = g;
    return MLV_OK;
}

MLV_API int mlv_graph_launch(mlv_context *ctx, mlv_graph *g) {
    if (!ctx || !g) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "graph_launch: bad argument");
    Marker marker("render (graph)");
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    HIP_OR_FAIL(ctx, hipGraphLaunch(g->exec, ctx->stream));
    ctx->vertex_count += g->vertex_count;
    ctx->input_tris += g->input_tris;
    ctx->draws += g->draws;
    return MLV_OK;
}

MLV_API int mlv_graph_destroy(mlv_context *ctx, mlv_graph *g) {
    if (!ctx || !g) return MLV_ERROR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    for (void *h : g->pinned) (void)hipHostFree(h);
    delete g;
    if (--ctx->graphs_alive <= 0) {
        ctx->graphs_alive = 0;
        free_retired(ctx);
    }
    return MLV_OK;
}

// ------------------------------------------------------------------ cameras

static_assert(sizeof(mlv_camera) == sizeof(CameraRec), "mlv_camera layout");

MLV_API int mlv_camera_constant_buffers(mlv_context *ctx, const mlv_camera *cameras, uint32_t count,
                                        float *device_cbs) {
    if (!ctx || (count && (!cameras || !device_cbs)))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "camera_constant_buffers: bad argument");
    for (uint32_t i = 0; i < count; ++i)
        if (!cameras[i].width || !cameras[i].height)
            return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "camera_constant_buffers: camera %u has an empty view", i);
    if (!count) return MLV_OK;
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));  // recorded draws read the buffers as they were
    const size_t bytes = sizeof(CameraRec) * count;
    RC_OR_RETURN(ensure(ctx, ctx->cameras, bytes));
    void *stage = nullptr;
    int slot = 0;
    RC_OR_RETURN(staging_acquire(ctx, bytes, &stage, &slot));
    memcpy(stage, cameras, bytes);
    HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->cameras.p, stage, bytes, hipMemcpyHostToDevice, ctx->stream));
    RC_OR_RETURN(staging_commit(ctx, slot));
    HIP_OR_FAIL(ctx, launch_camera_cbs((const CameraRec *)ctx->cameras.p, count, device_cbs, ctx->stream));
    return MLV_OK;
}

// ------------------------------------------------------------------ tuning

MLV_API int mlv_get_tuning(mlv_context *ctx, mlv_tuning *out) {
    if (!ctx || !out) return MLV_ERROR_INVALID_ARGUMENT;
    out->heavy_tile_pairs = ctx->heavy_min;
    out->heavy_blocks = ctx->heavy_blocks;
    out->pair_blocks = ctx->pair_blocks;
    out->tile_stage = ctx->tile_stage;
    out->pair_capacity = (uint32_t)std::min<size_t>(ctx->last_pair_cap, 0xFFFFFFFFu);
    out->tile_edge = ctx->tile_edge;
    out->tile_waves_per_cu = ctx->tile_waves;
    out->setup_waves_per_cu = ctx->setup_waves;
    out->skip_cornerless = ctx->skip_cornerless ? 1u : 0u;
    return MLV_OK;
}

MLV_API int mlv_set_tuning(mlv_context *ctx, const mlv_tuning *t) {
    if (!ctx || !t) return MLV_ERROR_INVALID_ARGUMENT;
    if (!t->heavy_blocks || t->heavy_blocks > 65536u || !t->pair_blocks || t->pair_blocks > 65536u)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "set_tuning: block counts must be 1..65536");
    if (t->tile_stage != 128u && t->tile_stage != 256u)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "set_tuning: tile_stage must be 128 or 256");
    if (t->tile_edge != 16u && t->tile_edge != 32u)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "set_tuning: tile_edge must be 16 or 32");
    if (t->tile_waves_per_cu > 32u || t->setup_waves_per_cu > 32u || t->skip_cornerless > 1u)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT,
                    "set_tuning: waves per CU must be 0..32 and skip_cornerless 0 or 1");
    RC_OR_RETURN(flush(ctx));  // recorded draws keep the parameters they were recorded under
    ctx->heavy_min = t->heavy_tile_pairs;
    ctx->heavy_blocks = t->heavy_blocks;
    ctx->pair_blocks = t->pair_blocks;
    ctx->tile_stage = t->tile_stage;
    ctx->tile_edge = t->tile_edge;
    ctx->tile_waves = t->tile_waves_per_cu;
    ctx->setup_waves = t->setup_waves_per_cu;
    ctx->skip_cornerless = t->skip_cornerless != 0u;
    return MLV_OK;
}

MLV_API int mlv_autotune_begin(mlv_context *ctx) { return mlv_graph_capture_begin(ctx); }

// Milliseconds per run of the recorded commands under the context's current parameters.
// The fastest of `repeats` timed runs of a recorded frame (after a warm-up run): the minimum is
// robust to the occasional slow run that an average would carry.
static int time_recorded(mlv_context *ctx, const std::vector<Cmd> &rec, uint32_t repeats, double *ms) {
    RC_OR_RETURN(run_recorded(ctx, rec));  // warm-up under these parameters
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    hipEvent_t a = nullptr, b = nullptr;
    HIP_OR_FAIL(ctx, hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(ctx, MLV_ERROR_DEVICE, "hipEventCreate failed");
    }
    int rc = MLV_OK;
    float best = 0.0f;
    for (uint32_t r = 0; r < repeats && rc == MLV_OK; ++r) {
        float t = 0.0f;
        if (hipEventRecord(a, ctx->stream) != hipSuccess) rc = MLV_ERROR_DEVICE;
        if (rc == MLV_OK) rc = run_recorded(ctx, rec);
        if (rc == MLV_OK && (hipEventRecord(b, ctx->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                             hipEventElapsedTime(&t, a, b) != hipSuccess))
            rc = fail(ctx, MLV_ERROR_DEVICE, "autotune timing failed");
        if (rc == MLV_OK && (r == 0 || t < best)) best = t;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    RC_OR_RETURN(rc);
    RC_OR_RETURN(check_device_stats(ctx));
    *ms = (double)best;
    return MLV_OK;
}

MLV_API int mlv_autotune_end(mlv_context *ctx, uint32_t repeats, mlv_tuning *best) {
    if (!ctx || !best || !repeats) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "autotune_end: bad argument");
    if (!ctx->recording) return fail(ctx, MLV_ERROR_INVALID_STATE, "autotune_end without autotune_begin");
    ctx->recording = false;
    std::vector<Cmd> rec;
    rec.swap(ctx->rec);
    RC_OR_RETURN(set_device(ctx));
    DevStats base{}, before{}, after{};
    RC_OR_RETURN(read_dstats(ctx, base));
    // size the arenas (a capacity miss raises the capacity; run again)
    int rc = MLV_OK;
    for (int attempt = 0; attempt < 3; ++attempt) {
        rc = run_recorded(ctx, rec);
        if (rc == MLV_OK) {
            HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
            rc = check_device_stats(ctx);
        }
        if (rc != MLV_ERROR_CAPACITY) break;
    }
    RC_OR_RETURN(rc);
    static const uint32_t skip[] = {1u, 0u};
    static const uint32_t tile_edge[] = {16u, 32u};
    static const uint32_t tile_stage[] = {128u, 256u};
    static const uint32_t tile_waves[] = {0u, 8u, 16u, 24u};
    static const uint32_t setup_waves[] = {0u, 16u, 24u};
    static const uint32_t heavy_min[] = {0u, 32u, 64u, 192u, 384u, 768u, 1536u, 0xFFFFFFFFu};
    static const uint32_t heavy_blocks[] = {256u, 512u, 1024u, 2048u};
    static const uint32_t pair_blocks[] = {512u, 1024u, 2048u, 4096u};
    double best_ms = 0.0;
    RC_OR_RETURN(time_recorded(ctx, rec, repeats, &best_ms));
    uint32_t skip_now = ctx->skip_cornerless ? 1u : 0u;
    constexpr int kKnobs = 8;
    uint32_t *knob[kKnobs] = {&skip_now,         &ctx->tile_edge, &ctx->tile_stage,   &ctx->tile_waves,
                              &ctx->setup_waves, &ctx->heavy_min, &ctx->heavy_blocks, &ctx->pair_blocks};
    const uint32_t *cand[kKnobs] = {skip, tile_edge, tile_stage, tile_waves, setup_waves, heavy_min, heavy_blocks,
                                    pair_blocks};
    const size_t ncand[kKnobs] = {2, 2, 2, 4, 3, 8, 4, 4};
    for (int k = 0; k < kKnobs; ++k) {
        if ((k == 5 || k == 6) && ctx->tile_edge != 16u) continue;  // the heavy-tile path needs 16x16 tiles
        uint32_t keep = *knob[k];
        for (size_t i = 0; i < ncand[k]; ++i) {
            if (cand[k][i] == keep) continue;
            *knob[k] = cand[k][i];
            ctx->skip_cornerless = skip_now != 0u;
            double ms = 0.0;
            RC_OR_RETURN(time_recorded(ctx, rec, repeats, &ms));
            if (ms < best_ms * 0.99) {  // a candidate must win by more than the runs' spread
                best_ms = ms;
                keep = cand[k][i];
            }
        }
        *knob[k] = keep;
        ctx->skip_cornerless = skip_now != 0u;
    }
    RC_OR_RETURN(read_dstats(ctx, before));
    RC_OR_RETURN(run_recorded(ctx, rec));  // leave the frame in its targets under the chosen parameters
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    RC_OR_RETURN(check_device_stats(ctx));
    RC_OR_RETURN(read_dstats(ctx, after));
    RC_OR_RETURN(keep_one_run_stats(ctx, base, before, after));  // the tuned frame counts once
    return mlv_get_tuning(ctx, best);
}

MLV_API int mlv_context_set_footprint_counting(mlv_context *ctx, int enable) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    if (ctx->recording)
        return fail(ctx, MLV_ERROR_INVALID_STATE, "footprint counting cannot change while recording");
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));  // recorded draws keep the mode they were recorded under
    if (enable && !ctx->foot) {
        HIP_OR_FAIL(ctx, hipMalloc(&ctx->foot, sizeof(FootCounts)));
        HIP_OR_FAIL(ctx, launch_fill_u32((uint32_t *)ctx->foot, 0u, sizeof(FootCounts) / 4, ctx->stream));
    }
    ctx->counting = enable != 0;
    return MLV_OK;
}

MLV_API int mlv_get_footprint(mlv_context *ctx, mlv_footprint *out) {
    if (!ctx || !out) return MLV_ERROR_INVALID_ARGUMENT;
    *out = mlv_footprint{};
    if (!ctx->foot) return MLV_OK;
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    for (mlv_texture2d *t : ctx->marked) {
        HIP_OR_FAIL(ctx, launch_popcount(t->mark, t->mark_words, (uint32_t)texel_bytes((int)t->format),
                                         &ctx->foot->texels, &ctx->foot->tex_bytes, ctx->stream));
        HIP_OR_FAIL(ctx, launch_fill_u32(t->mark, 0u, t->mark_words, ctx->stream));
    }
    FootCounts f{};
    HIP_OR_FAIL(ctx, hipMemcpyAsync(&f, ctx->foot, sizeof f, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OR_FAIL(ctx, launch_fill_u32((uint32_t *)ctx->foot, 0u, sizeof(FootCounts) / 4, ctx->stream));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    out->shaded_pixels = f.shaded_px;
    out->visible_triangles = f.vis_tris;
    out->texels = f.texels;
    out->texel_bytes = f.tex_bytes;
    return MLV_OK;
}

MLV_API int mlv_context_set_profiling(mlv_context *ctx, int enable) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    RC_OR_RETURN(flush(ctx));
    ctx->profiling = enable != 0;
    return MLV_OK;
}

MLV_API int mlv_get_profile(mlv_context *ctx, mlv_profile *out) {
    if (!ctx || !out) return MLV_ERROR_INVALID_ARGUMENT;
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    resolve_events(ctx);
    for (int i = 0; i < kStages; ++i) {
        out->stage_ms[i] = ctx->stage_ms[i];
        ctx->stage_ms[i] = 0.0;
    }
    out->tile_kernel_ms = ctx->tile_kernel_ms;
    ctx->tile_kernel_ms = 0.0;
    out->draws = ctx->prof_draws;
    out->flushes = ctx->prof_flushes;
    ctx->prof_draws = ctx->prof_flushes = 0;
    return MLV_OK;
}

MLV_API int mlv_flush(mlv_context *ctx) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    return flush(ctx);
}

MLV_API int mlv_render_target_readback(mlv_context *ctx, mlv_render_target *rt, void *host_dst) {
    if (!ctx || !rt || !host_dst) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "readback: bad argument");
    Marker marker("present");
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    HIP_OR_FAIL(ctx, hipMemcpyAsync(host_dst, rt->ptr, (size_t)rt->width * rt->height * 4, hipMemcpyDeviceToHost,
                                    ctx->stream));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    return check_device_stats(ctx);
}

MLV_API int mlv_present(mlv_context *ctx, mlv_render_target *rt, uint8_t *host_rgb) {
    if (!ctx || !rt || !host_rgb) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "present: bad argument");
    Marker marker("present");
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    const size_t n = (size_t)rt->width * rt->height;
    RC_OR_RETURN(ensure(ctx, ctx->present_rgb, 12 * ((n + 3) / 4)));
    HIP_OR_FAIL(ctx, launch_present_rgb(rt->ptr, (uint8_t *)ctx->present_rgb.p, n, ctx->stream));
    HIP_OR_FAIL(ctx, hipMemcpyAsync(host_rgb, ctx->present_rgb.p, 3 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    return check_device_stats(ctx);
}

MLV_API int mlv_depth_stencil_readback(mlv_context *ctx, mlv_depth_stencil *ds, float *host_dst) {
    if (!ctx || !ds || !host_dst) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "readback: bad argument");
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    HIP_OR_FAIL(ctx, hipMemcpyAsync(host_dst, ds->ptr, (size_t)ds->width * ds->height * 4, hipMemcpyDeviceToHost,
                                    ctx->stream));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    return check_device_stats(ctx);
}

MLV_API int mlv_get_stats(mlv_context *ctx, mlv_stats *out) {
    if (!ctx || !out) return MLV_ERROR_INVALID_ARGUMENT;
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    DevStats s{};
    HIP_OR_FAIL(ctx, hipMemcpyAsync(&s, ctx->dstats, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    out->vertex_count = ctx->vertex_count;
    out->input_triangle_count = ctx->input_tris;
    out->assembled_triangle_count = s.assembled;
    out->active_bin_count = s.active8;
    for (unsigned long long a : s.active8_part) out->active_bin_count += a;
    out->total_triangle_count_in_bins = s.pairs8;
    for (unsigned long long a : s.pairs8_part) out->total_triangle_count_in_bins += a;
    out->oob_triangle_count = s.oob;
    out->draw_count = ctx->draws;
    out->overflow_count = s.overflow_draws;
    return MLV_OK;
}

MLV_API int mlv_reset_stats(mlv_context *ctx) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
    HIP_OR_FAIL(ctx, hipMemset(ctx->dstats, 0, sizeof(DevStats)));
    HIP_OR_FAIL(ctx, hipDeviceSynchronize());
    ctx->vertex_count = ctx->input_tris = ctx->draws = 0;
    ctx->capacity_errors = 0;
    ctx->draw_counter = 0;
    return MLV_OK;
}

// ------------------------------------------------------------------ differentiable shading

static int diff_mesh(mlv_context *ctx, const float *pos, const float *col, const float *uv, const float *nrm,
                     const uint32_t *idx, uint32_t V, uint32_t T, const float *tex, uint32_t tw, uint32_t th,
                     uint32_t levels, DiffMesh &m) {
    if (!pos || !col || !uv || !nrm || !idx || !tex)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff: a mesh array is NULL");
    if (!V || !T || V > 0x3FFFFFFFu || T > 0x3FFFFFFFu || !tw || !th || tw > 16384u || th > 16384u)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff: bad mesh or texture sizes");
    uint32_t max_levels = 1;
    while (max_levels < kMaxMips && ((tw >> max_levels) || (th >> max_levels))) ++max_levels;
    if (levels < 1 || levels > max_levels)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff: tex_levels must be 1..%u for a %ux%u texture", max_levels,
                    tw, th);
    m = DiffMesh{};
    m.pos = pos;
    m.col = col;
    m.uv = uv;
    m.nrm = nrm;
    m.idx = idx;
    m.V = V;
    m.T = T;
    m.tex = (const float4 *)tex;
    m.levels = levels;
    uint64_t off = 0;
    for (uint32_t l = 0; l < levels; ++l) {
        m.lw[l] = std::max(tw >> l, 1u);
        m.lh[l] = std::max(th >> l, 1u);
        m.loff[l] = (uint32_t)off;
        off += (uint64_t)m.lw[l] * m.lh[l];
    }
    m.tex_total = (uint32_t)off;  // <= 4/3 * 2^28 texels: 4 * texel index fits 32 bits
    return MLV_OK;
}

static int diff_view(mlv_context *ctx, const float *clip_from_world, uint32_t W, uint32_t H, DiffView &v) {
    if (!W || !H || W > 16384u || H > 16384u) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff: bad image size");
    memcpy(v.M, clip_from_world, sizeof v.M);
    v.W = W;
    v.H = H;
    return MLV_OK;
}

static int diff_params(mlv_context *ctx, const mlv_diff_scene *sc, const uint32_t *prim, DiffParams &p) {
    if (!sc || !prim) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff: scene and primitive ids are required");
    if (!sc->sh) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff: sh is NULL");
    p = DiffParams{};
    RC_OR_RETURN(diff_mesh(ctx, sc->positions, sc->colors, sc->uvs, sc->normals, sc->indices, sc->vertex_count,
                           sc->triangle_count, sc->texture, sc->tex_width, sc->tex_height, sc->tex_levels, p.m));
    RC_OR_RETURN(diff_view(ctx, sc->clip_from_world, sc->width, sc->height, p.v));
    p.sh = sc->sh;
    p.prim = prim;
    return MLV_OK;
}

MLV_API int mlv_diff_forward(mlv_context *ctx, const mlv_diff_scene *scene, const uint32_t *device_prim_ids,
                             float *device_rgba_out) {
    if (!ctx || !device_rgba_out) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff_forward: bad argument");
    DiffParams p;
    RC_OR_RETURN(diff_params(ctx, scene, device_prim_ids, p));
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));  // the primitive ids come from recorded draws
    size_t words = 0;
    if (ctx->counting) {  // footprint counting: a bitmap over the chain
        words = ((size_t)p.m.tex_total + 31) / 32;
        RC_OR_RETURN(ensure(ctx, ctx->diff_mark, words * 4));
        HIP_OR_FAIL(ctx, launch_fill_u32((uint32_t *)ctx->diff_mark.p, 0u, words, ctx->stream));
        p.m.tex_mark = (uint32_t *)ctx->diff_mark.p;
    }
    HIP_OR_FAIL(ctx, launch_diff_forward(p, (float4 *)device_rgba_out, ctx->stream));
    if (words)
        HIP_OR_FAIL(ctx, launch_popcount(p.m.tex_mark, words, 16u, &ctx->foot->texels, &ctx->foot->tex_bytes,
                                         ctx->stream));
    return MLV_OK;
}

MLV_API int mlv_mip_chain_size(uint32_t width, uint32_t height, uint32_t *texels, uint32_t *levels) {
    if (!width || !height || width > 32768u || height > 32768u || (uint64_t)width * height > 0x40000000ull || !texels)
        return MLV_ERROR_INVALID_ARGUMENT;
    *texels = mip_chain_texels(width, height, levels);
    return MLV_OK;
}

MLV_API int mlv_mip_chain(mlv_context *ctx, const float *device_texture, uint32_t width, uint32_t height,
                          float *device_chain) {
    uint32_t n = 0;
    if (!ctx || !device_texture || !device_chain || mlv_mip_chain_size(width, height, &n, nullptr) != MLV_OK)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "mip_chain: bad argument");
    RC_OR_RETURN(set_device(ctx));
    HIP_OR_FAIL(ctx, launch_mip_chain((const float4 *)device_texture, width, height, (float4 *)device_chain, ctx->stream));
    return MLV_OK;
}

MLV_API int mlv_mip_chain_adjoint(mlv_context *ctx, const float *device_grad_chain, uint32_t width, uint32_t height,
                                  float *device_grad_texture) {
    uint32_t n = 0;
    if (!ctx || !device_grad_chain || !device_grad_texture || mlv_mip_chain_size(width, height, &n, nullptr) != MLV_OK)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "mip_chain_adjoint: bad argument");
    RC_OR_RETURN(set_device(ctx));
    HIP_OR_FAIL(ctx, launch_mip_adjoint((const float4 *)device_grad_chain, width, height, (float4 *)device_grad_texture,
                                        ctx->stream));
    return MLV_OK;
}

MLV_API int mlv_diff_backward(mlv_context *ctx, const mlv_diff_scene *scene, const uint32_t *device_prim_ids,
                              const float *device_grad_rgba, const mlv_diff_grads *grads, int deterministic) {
    if (!ctx || !device_grad_rgba || !grads) return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff_backward: bad argument");
    DiffParams p;
    RC_OR_RETURN(diff_params(ctx, scene, device_prim_ids, p));
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    const uint32_t nb = std::max(diff_blocks(p.v.W, p.v.H), diff_tile_blocks(p.v.W, p.v.H));
    const size_t P = (size_t)p.v.W * p.v.H;
    DiffScratch sc{};
    RC_OR_RETURN(ensure(ctx, ctx->d_shpart, (size_t)std::max<uint32_t>(nb, 1u) * 27 * 8));
    sc.sh_part = (double *)ctx->d_shpart.p;
    RC_OR_RETURN(ensure(ctx, ctx->d_posacc, (size_t)p.m.V * 3 * 8));
    sc.posacc = (double *)ctx->d_posacc.p;
    if (deterministic) {
        const size_t nt = 4 * P, nv = 3 * P, nmax = std::max(nt, nv);
        RC_OR_RETURN(ensure(ctx, ctx->d_tkey, nt * 4));
        RC_OR_RETURN(ensure(ctx, ctx->d_tval, nt * 4));
        RC_OR_RETURN(ensure(ctx, ctx->d_tkey2, nt * 4));
        RC_OR_RETURN(ensure(ctx, ctx->d_tval2, nt * 4));
        RC_OR_RETURN(ensure(ctx, ctx->d_tpay, nt * 16));
        RC_OR_RETURN(ensure(ctx, ctx->d_vkey, nv * 4));
        RC_OR_RETURN(ensure(ctx, ctx->d_vval, nv * 4));
        RC_OR_RETURN(ensure(ctx, ctx->d_vkey2, nv * 4));
        RC_OR_RETURN(ensure(ctx, ctx->d_vval2, nv * 4));
        RC_OR_RETURN(ensure(ctx, ctx->d_vpay, nv * 32));
        RC_OR_RETURN(ensure(ctx, ctx->d_vposd, nv * 32));
        RC_OR_RETURN(ensure(ctx, ctx->d_hist, radix_hist_elems(nmax) * 4));
        {
            const void *before = ctx->d_part.p;
            RC_OR_RETURN(ensure(ctx, ctx->d_part, scan_scratch_bytes(radix_hist_elems(nmax))));
            if (ctx->d_part.p != before)
                HIP_OR_FAIL(ctx, hipMemsetAsync(ctx->d_part.p, 0, ctx->d_part.cap, ctx->stream));
        }
        sc.tkey = (uint32_t *)ctx->d_tkey.p;
        sc.tval = (uint32_t *)ctx->d_tval.p;
        sc.tkey2 = (uint32_t *)ctx->d_tkey2.p;
        sc.tval2 = (uint32_t *)ctx->d_tval2.p;
        sc.tpay = (float4 *)ctx->d_tpay.p;
        sc.vkey = (uint32_t *)ctx->d_vkey.p;
        sc.vval = (uint32_t *)ctx->d_vval.p;
        sc.vkey2 = (uint32_t *)ctx->d_vkey2.p;
        sc.vval2 = (uint32_t *)ctx->d_vval2.p;
        sc.vpay = (float4 *)ctx->d_vpay.p;
        sc.vposd = (double *)ctx->d_vposd.p;
        sc.hist = (uint32_t *)ctx->d_hist.p;
        sc.part = ctx->d_part.p;
        sc.epoch = &ctx->scan_epoch;
        sc.err = &ctx->dstats->scan_errors;
    }
    DiffGrads g{grads->positions, grads->colors, grads->uvs, grads->normals, grads->texture, grads->sh};
    HIP_OR_FAIL(ctx, launch_diff_backward(p, (const float4 *)device_grad_rgba, g, sc, deterministic != 0, ctx->stream));
    return MLV_OK;
}

MLV_API int mlv_diff_backward_noalloc(mlv_context *ctx, const mlv_diff_scene *scene, const uint32_t *device_prim_ids,
                                      const float *device_grad_rgba, const mlv_diff_grads *grads) {
    if (!ctx || !device_grad_rgba || !grads)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff_backward_noalloc: bad argument");
    // nothing here may allocate: recorded draws would need a flush, whose arenas may grow
    if (ctx->recording || !ctx->pending.empty())
        return fail(ctx, MLV_ERROR_INVALID_STATE,
                    "diff_backward_noalloc: recorded draws are pending; call mlv_flush first");
    DiffParams p;
    RC_OR_RETURN(diff_params(ctx, scene, device_prim_ids, p));
    const uint32_t nb = diff_tile_blocks(p.v.W, p.v.H);
    const size_t pos_bytes = ((size_t)p.m.V * 3 * 8 + 255) & ~(size_t)255;
    const size_t need = pos_bytes + (size_t)std::max<uint32_t>(nb, 1u) * 27 * 8;
    if (need > kFixedScratchBytes)
        return fail(ctx, MLV_ERROR_CAPACITY,
                    "diff_backward_noalloc: %zu bytes of scratch needed (24 B per vertex + 216 B per 16x16 pixel "
                    "tile), the fixed array holds %zu",
                    need, kFixedScratchBytes);
    if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch_ptr[ctx->device])
        return fail(ctx, MLV_ERROR_DEVICE, "diff_backward_noalloc: no scratch address for device %d (%s)", ctx->device,
                    (ctx->device >= 0 && ctx->device < kMaxDevices) ? hipGetErrorString(g_scratch_err[ctx->device])
                                                                    : "device index out of range");
    RC_OR_RETURN(set_device(ctx));
    uint8_t *base = (uint8_t *)g_scratch_ptr[ctx->device];
    DiffScratch sc{};
    sc.posacc = (double *)base;  // cleared by the launch before use
    sc.sh_part = (double *)(base + pos_bytes);  // every block writes its 27 partials
    DiffGrads g{grads->positions, grads->colors, grads->uvs, grads->normals, grads->texture, grads->sh};
    HIP_OR_FAIL(ctx, launch_diff_backward(p, (const float4 *)device_grad_rgba, g, sc, false, ctx->stream));
    return MLV_OK;
}

// Batched shading: validates the meshes and views, lays out the device tables (meshes, views,
// per-view first blocks, draw -> mesh map, per-mesh first position blocks, the ranges the
// backward pass clears) and copies them through pinned staging.
static int diff_batch_setup(mlv_context *ctx, const mlv_diff_mesh *meshes, uint32_t n_meshes,
                            const mlv_diff_view *views, uint32_t n_views, const float *sh, bool backward,
                            bool accumulate, DiffBatchParams &P) {
    if ((n_meshes && !meshes) || (n_views && !views) || !sh)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff batch: meshes, views and sh are required");
    std::vector<DiffBatchMesh> bm(n_meshes);
    uint64_t pos_total = 0;  // library-held position sums
    for (uint32_t k = 0; k < n_meshes; ++k) {
        const mlv_diff_mesh &m = meshes[k];
        RC_OR_RETURN(diff_mesh(ctx, m.positions, m.colors, m.uvs, m.normals, m.indices, m.vertex_count,
                               m.triangle_count, m.texture, m.tex_width, m.tex_height, m.tex_levels, bm[k].m));
        bm[k].g = DiffGrads{m.grad_positions, m.grad_colors, m.grad_uvs, m.grad_normals, m.grad_texture, nullptr};
        if (backward && accumulate && m.grad_positions && !m.pos_accum)
            return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff batch: accumulating positions needs pos_accum (mesh %u)",
                        k);
        if (!m.pos_accum) pos_total += 3ull * m.vertex_count;
    }
    std::vector<DiffBatchView> bv(n_views);
    std::vector<uint32_t> view_block0(std::max<uint32_t>(n_views, 1u), 0u), draw_mesh;
    uint64_t blocks = 0;
    for (uint32_t v = 0; v < n_views; ++v) {
        const mlv_diff_view &w = views[v];
        DiffBatchView &o = bv[v];
        o = DiffBatchView{};
        RC_OR_RETURN(diff_view(ctx, w.clip_from_world, w.width, w.height, o.v));
        if (!w.prim_ids || !w.draw_tags || (w.draw_count && !w.draw_mesh))
            return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff batch: view %u lacks its id buffers or draw map", v);
        if (backward ? !w.grad_rgba : !w.rgba)
            return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff batch: view %u lacks its %s image", v,
                        backward ? "gradient" : "output");
        o.prim = w.prim_ids;
        o.tags = w.draw_tags;
        o.first_tag = w.first_tag;
        o.n_draws = w.draw_count;
        o.draw_off = (uint32_t)draw_mesh.size();
        o.identity_map = 1u;
        for (uint32_t d = 0; d < w.draw_count; ++d) {
            draw_mesh.push_back(w.draw_mesh[d]);
            if (w.draw_mesh[d] != d || d >= n_meshes) o.identity_map = 0u;
        }
        o.out = (float4 *)w.rgba;
        o.gout = (const float4 *)w.grad_rgba;
        view_block0[v] = (uint32_t)blocks;
        blocks += diff_tile_blocks(w.width, w.height);
        if (blocks > 0x7FFFFFFFull || draw_mesh.size() > 0x7FFFFFFFull)
            return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff batch: too many pixels or draws");
    }
    if (draw_mesh.empty()) draw_mesh.push_back(0xFFFFFFFFu);
    if (!backward && ctx->counting) {  // footprint counting: one bitmap per mesh chain
        size_t words = 0;
        for (auto &b : bm) words += ((size_t)b.m.tex_total + 31) / 32;
        RC_OR_RETURN(ensure(ctx, ctx->diff_mark, std::max<size_t>(words, 1) * 4));
        HIP_OR_FAIL(ctx, launch_fill_u32((uint32_t *)ctx->diff_mark.p, 0u, words, ctx->stream));
        size_t off = 0;
        for (auto &b : bm) {
            b.m.tex_mark = (uint32_t *)ctx->diff_mark.p + off;
            off += ((size_t)b.m.tex_total + 31) / 32;
        }
        ctx->diff_mark_words = words;
    }

    // FP64 position sums: the caller's pos_accum, else one [V][3] range of library scratch
    if (backward) RC_OR_RETURN(ensure(ctx, ctx->b_posacc, std::max<uint64_t>(pos_total, 1) * 8));
    std::vector<uint32_t> mesh_block0(std::max<uint32_t>(n_meshes, 1u), 0u);
    std::vector<DiffZeroSeg> zero;
    std::vector<uint32_t> zero_block0;
    uint64_t pos_blocks = 0, zero_blocks = 0, pos_off = 0;
    auto clear = [&](void *p, uint64_t words) {
        for (uint64_t w0 = 0; p && w0 < words; w0 += (1ull << 30)) {
            const uint32_t n = (uint32_t)std::min<uint64_t>(words - w0, 1ull << 30);
            zero.push_back(DiffZeroSeg{(uint32_t *)p + w0, n, 0});
            zero_block0.push_back((uint32_t)zero_blocks);
            zero_blocks += (n + 1023u) / 1024u;
        }
    };
    for (uint32_t k = 0; k < n_meshes; ++k) {
        DiffBatchMesh &b = bm[k];
        const uint64_t V = b.m.V;
        b.posacc = nullptr;
        if (backward) {
            if (meshes[k].pos_accum) {
                b.posacc = meshes[k].pos_accum;
            } else {
                b.posacc = (double *)ctx->b_posacc.p + pos_off;
                pos_off += 3 * V;
            }
        }
        mesh_block0[k] = (uint32_t)pos_blocks;
        pos_blocks += (3 * V + 255) / 256;
        if (backward && !accumulate) {
            clear(b.g.pos ? b.posacc : nullptr, 6 * V);
            clear(b.g.pos, 3 * V);
            clear(b.g.col, 3 * V);
            clear(b.g.uv, 2 * V);
            clear(b.g.nrm, 3 * V);
            clear(b.g.tex, 4ull * b.m.tex_total);
        }
    }
    if (pos_blocks > 0x7FFFFFFFull || zero_blocks > 0x7FFFFFFFull)
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff batch: meshes too large");
    if (zero.empty()) {
        zero.push_back(DiffZeroSeg{nullptr, 0, 0});
        zero_block0.push_back(0);
    }

    // one blob: [meshes][views][view_block0][draw_mesh][mesh_block0][zero][zero_block0]
    auto a256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_mesh = 0, s_mesh = sizeof(DiffBatchMesh) * std::max<uint32_t>(n_meshes, 1u);
    const size_t o_view = a256(o_mesh + s_mesh), s_view = sizeof(DiffBatchView) * std::max<uint32_t>(n_views, 1u);
    const size_t o_vb = a256(o_view + s_view), s_vb = 4 * view_block0.size();
    const size_t o_dm = a256(o_vb + s_vb), s_dm = 4 * draw_mesh.size();
    const size_t o_mb = a256(o_dm + s_dm), s_mb = 4 * mesh_block0.size();
    const size_t o_z = a256(o_mb + s_mb), s_z = sizeof(DiffZeroSeg) * zero.size();
    const size_t o_zb = a256(o_z + s_z), s_zb = 4 * zero_block0.size();
    const size_t bytes = a256(o_zb + s_zb);
    RC_OR_RETURN(ensure(ctx, ctx->b_tables, bytes));
    void *stage = nullptr;
    int slot = 0;
    RC_OR_RETURN(staging_acquire(ctx, bytes, &stage, &slot));
    uint8_t *h = (uint8_t *)stage;
    if (n_meshes) memcpy(h + o_mesh, bm.data(), sizeof(DiffBatchMesh) * n_meshes);
    if (n_views) memcpy(h + o_view, bv.data(), sizeof(DiffBatchView) * n_views);
    memcpy(h + o_vb, view_block0.data(), s_vb);
    memcpy(h + o_dm, draw_mesh.data(), s_dm);
    memcpy(h + o_mb, mesh_block0.data(), s_mb);
    memcpy(h + o_z, zero.data(), s_z);
    memcpy(h + o_zb, zero_block0.data(), s_zb);
    HIP_OR_FAIL(ctx, hipMemcpyAsync(ctx->b_tables.p, stage, bytes, hipMemcpyHostToDevice, ctx->stream));
    RC_OR_RETURN(staging_commit(ctx, slot));
    uint8_t *d = (uint8_t *)ctx->b_tables.p;
    P = DiffBatchParams{};
    P.meshes = (const DiffBatchMesh *)(d + o_mesh);
    P.views = (const DiffBatchView *)(d + o_view);
    P.view_block0 = (const uint32_t *)(d + o_vb);
    P.draw_mesh = (const uint32_t *)(d + o_dm);
    P.mesh_block0 = (const uint32_t *)(d + o_mb);
    P.zero = (const DiffZeroSeg *)(d + o_z);
    P.zero_block0 = (const uint32_t *)(d + o_zb);
    P.n_meshes = n_meshes;
    P.n_views = n_views;
    P.n_blocks = n_views ? (uint32_t)blocks : 0u;
    P.n_pos_blocks = (backward && n_meshes) ? (uint32_t)pos_blocks : 0u;
    P.n_zero = (uint32_t)zero.size();
    P.n_zero_blocks = (backward && !accumulate) ? (uint32_t)zero_blocks : 0u;
    P.sh = sh;
    return MLV_OK;
}

static int diff_batch_forward_now(mlv_context *ctx, const mlv_diff_mesh *meshes, uint32_t mesh_count,
                                  const mlv_diff_view *views, uint32_t view_count, const float *device_sh) {
    DiffBatchParams P;
    RC_OR_RETURN(diff_batch_setup(ctx, meshes, mesh_count, views, view_count, device_sh, false, false, P));
    const void *old = ctx->b_defer.p;
    RC_OR_RETURN(ensure(ctx, ctx->b_defer, 64u + (size_t)P.n_blocks * 16u));
    if (ctx->b_defer.p != old)  // new memory; afterwards the fix-up pass leaves its counts at zero
        HIP_OR_FAIL(ctx, launch_fill_u32((uint32_t *)ctx->b_defer.p, 0u, 16u, ctx->stream));
    P.defer = (uint32_t *)ctx->b_defer.p;
    P.defer_list = P.defer + 16;
    HIP_OR_FAIL(ctx, launch_diff_batch_forward(P, ctx->stream));
    if (ctx->counting)  // RGBA32F chains: 16 B a texel
        HIP_OR_FAIL(ctx, launch_popcount((const uint32_t *)ctx->diff_mark.p, ctx->diff_mark_words, 16u,
                                         &ctx->foot->texels, &ctx->foot->tex_bytes, ctx->stream));
    return MLV_OK;
}

static int diff_batch_backward_now(mlv_context *ctx, const mlv_diff_mesh *meshes, uint32_t mesh_count,
                                   const mlv_diff_view *views, uint32_t view_count, const float *device_sh,
                                   float *device_grad_sh, int accumulate) {
    DiffBatchParams P;
    RC_OR_RETURN(diff_batch_setup(ctx, meshes, mesh_count, views, view_count, device_sh, true, accumulate != 0, P));
    RC_OR_RETURN(ensure(ctx, ctx->b_shpart, (size_t)std::max<uint32_t>(P.n_blocks, 1u) * 27 * 8));
    HIP_OR_FAIL(ctx, launch_diff_batch_backward(P, (double *)ctx->b_shpart.p, device_grad_sh, accumulate != 0,
                                                ctx->stream));
    return MLV_OK;
}

// While a graph is recorded, a batched pass or a loss gradient is recorded with copies of its
// descriptor arrays and draw maps (the device buffers they name must outlive the graph).
static int record_diff(mlv_context *ctx, int kind, const mlv_diff_mesh *meshes, uint32_t mesh_count,
                       const mlv_diff_view *views, uint32_t view_count, std::shared_ptr<DiffCmd> d) {
    if ((mesh_count && !meshes) || (view_count && !views))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff batch: meshes and views are required");
    d->meshes.assign(meshes, meshes + mesh_count);
    d->views.assign(views, views + view_count);
    for (const mlv_diff_view &v : d->views) {
        if (v.draw_count && !v.draw_mesh)
            return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "diff batch: a view lacks its draw map");
        if (v.draw_count) d->draw_mesh.insert(d->draw_mesh.end(), v.draw_mesh, v.draw_mesh + v.draw_count);
    }
    Cmd c;
    c.kind = kind;
    c.diff = std::move(d);
    ctx->rec.push_back(std::move(c));
    return MLV_OK;
}

static int exec_diff_cmd(mlv_context *ctx, const Cmd &c) {
    const DiffCmd &d = *c.diff;
    if (c.kind == kCmdL2Grad) {
        HIP_OR_FAIL(ctx, launch_l2_grad(d.a, d.b, d.n, d.scale, d.out, ctx->stream));
        return MLV_OK;
    }
    std::vector<mlv_diff_view> views = d.views;
    size_t off = 0;
    for (mlv_diff_view &v : views) {
        v.draw_mesh = v.draw_count ? d.draw_mesh.data() + off : nullptr;
        off += v.draw_count;
    }
    if (c.kind == kCmdDiffForward)
        return diff_batch_forward_now(ctx, d.meshes.data(), (uint32_t)d.meshes.size(), views.data(),
                                      (uint32_t)views.size(), d.sh);
    return diff_batch_backward_now(ctx, d.meshes.data(), (uint32_t)d.meshes.size(), views.data(),
                                   (uint32_t)views.size(), d.sh, d.grad_sh, d.accumulate);
}

MLV_API int mlv_diff_batch_forward(mlv_context *ctx, const mlv_diff_mesh *meshes, uint32_t mesh_count,
                                   const mlv_diff_view *views, uint32_t view_count, const float *device_sh) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    if (ctx->recording) {
        auto d = std::make_shared<DiffCmd>();
        d->sh = device_sh;
        return record_diff(ctx, kCmdDiffForward, meshes, mesh_count, views, view_count, std::move(d));
    }
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));  // the id buffers come from recorded draws
    return diff_batch_forward_now(ctx, meshes, mesh_count, views, view_count, device_sh);
}

MLV_API int mlv_diff_batch_backward(mlv_context *ctx, const mlv_diff_mesh *meshes, uint32_t mesh_count,
                                    const mlv_diff_view *views, uint32_t view_count, const float *device_sh,
                                    float *device_grad_sh, int accumulate) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    if (ctx->recording) {
        auto d = std::make_shared<DiffCmd>();
        d->sh = device_sh;
        d->grad_sh = device_grad_sh;
        d->accumulate = accumulate;
        return record_diff(ctx, kCmdDiffBackward, meshes, mesh_count, views, view_count, std::move(d));
    }
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    return diff_batch_backward_now(ctx, meshes, mesh_count, views, view_count, device_sh, device_grad_sh,
                                   accumulate);
}

MLV_API int mlv_diff_l2_grad(mlv_context *ctx, const float *device_a, const float *device_b, uint64_t count,
                             float scale, float *device_out) {
    if (!ctx) return MLV_ERROR_INVALID_ARGUMENT;
    if (count && (!device_a || !device_b || !device_out))
        return fail(ctx, MLV_ERROR_INVALID_ARGUMENT, "l2_grad: a, b and out are required");
    if (ctx->recording) {
        auto d = std::make_shared<DiffCmd>();
        d->a = device_a;
        d->b = device_b;
        d->out = device_out;
        d->n = count;
        d->scale = scale;
        return record_diff(ctx, kCmdL2Grad, nullptr, 0, nullptr, 0, std::move(d));
    }
    RC_OR_RETURN(set_device(ctx));
    RC_OR_RETURN(flush(ctx));
    HIP_OR_FAIL(ctx, launch_l2_grad(device_a, device_b, count, scale, device_out, ctx->stream));
    return MLV_OK;
}

}  // extern "C"
