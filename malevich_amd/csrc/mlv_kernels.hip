// mlv_kernels.hip -- the draw pipeline as gfx950 kernels, one pass per stage per flush.
//
// Stage 0, vertex:   kb_block_draws      each vertex / setup block's draw (flushes of many draws)
//                    kb_vertex           one lane per vertex slot of every draw (main.c:662-734),
//                                        and the flush's counters cleared
// Stage 1, setup:    kb_setup_count      one lane per input triangle: cull/classify, output count
//                    kb_clip_count       triangles that need the clipper (main.c:609-660)
//                    scan (u64)          ordered compaction offsets over all setup blocks
//                    kb_setup_emit       TriVis/TriShade at the ordered slot (main.c:751-908); totals,
//                                        capacity checks, first assembled index per draw
//                    kb_clip_emit        clipped fan triangles + clipper vertices
// Stage 2, binning:  kb_pairs<false>     (triangle, 16x16 tile) pair count per triangle
//                    scan                pair offsets (triangle order)
//                    kb_pairs<true>      pairs in triangle order
//                    radix passes        stable LSD sort by tile: lists stay in submission order
//                    kb_tile_ranges      each tile's [start, end) read off the sorted keys
// Stage 3, tile:     kb_tile             one block per 16x16 tile of every view, one wave per
//                                        8x8 bin: hierarchical-Z, corner-sampled coverage,
//                                        early-Z in registers (tiny triangles one per lane, merged
//                                        by LDS maxima), the pixel shader of the surviving
//                                        fragment and the output merger (main.c:983-1189)
//
// Visibility per pixel is the maximum of (depth, submission order) under the reference's
// GREATER_EQUAL test (main.c:1166): running "z >= d then write" over a list in submission
// order leaves the fragment of maximal depth and, among equal depths, the latest. The tile
// lists are also kept in submission order, so each wave meets the draws of its 8x8 bin in
// order and can refresh that bin's hierarchical-Z minimum between draws exactly as the
// reference's write_tile does at the end of every draw (main.c:589-603).
#include <algorithm>
#include <type_traits>
#include <utility>

#include "mlv_device.h"

namespace mlv {

// ------------------------------------------------------------ helpers

__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T x) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// Whole-wave reductions (every lane of the wave active).
__device__ __forceinline__ uint32_t wave_min(uint32_t x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, d, 64));
    return x;
}

// Block-wide exclusive scan for NT threads (NT a multiple of 64, <= 1024). s_wave holds NT/64+1.
template <int NT, class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *s_wave, T &total) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    T inc = wave_inclusive_scan(v);
    if (lane == 63u) s_wave[w] = inc;
    __syncthreads();
    if (threadIdx.x < 64u) {
        T s = threadIdx.x < (uint32_t)(NT / 64) ? s_wave[threadIdx.x] : T(0);
        T si = wave_inclusive_scan(s);
        if (threadIdx.x < (uint32_t)(NT / 64)) s_wave[threadIdx.x] = si - s;
        if (threadIdx.x == 63u) s_wave[NT / 64] = si;
    }
    __syncthreads();
    T r = s_wave[w] + inc - v;
    total = s_wave[NT / 64];
    __syncthreads();
    return r;
}

// Largest index lo in [0, n) with key(lo) <= x, for ascending keys with key(0) <= x.
template <class F>
__device__ __forceinline__ uint32_t upper_index(uint32_t n, uint32_t x, F key) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (key(mid) <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// The three indices are loaded together and checked afterwards, so one memory round trip
// brings all three.
__device__ __forceinline__ bool load_tri_indices(const DrawRec &d, uint32_t t, uint32_t vi[3]) {
    const uint32_t *ib = d.ib + 3u * t;
    const uint32_t r0 = load_global(ib), r1 = load_global(ib + 1), r2 = load_global(ib + 2);
    const long long bv = (long long)d.base_vertex, nv = (long long)d.vertex_count;
    const long long v0 = (long long)r0 + bv, v1 = (long long)r1 + bv, v2 = (long long)r2 + bv;
    vi[0] = (uint32_t)v0;
    vi[1] = (uint32_t)v1;
    vi[2] = (uint32_t)v2;
    return v0 >= 0 && v0 < nv && v1 >= 0 && v1 < nv && v2 >= 0 && v2 < nv;
}

// ------------------------------------------------------------ the clipper
//
// Sutherland-Hodgman against the six planes (main.c:609-660), one triangle per group of 16
// lanes: lane i of the group holds polygon vertex i (NC floats in registers; NC = 4 is the
// clip-space position, which decides the polygon's shape, NC = 12 adds the attributes). A plane
// is one parallel step: each vertex's signed distance, the next vertex's by a shuffle, the
// outputs' slots by a prefix count over two ballots, and one pass through LDS to put output k in
// lane k. The outputs, their order and their arithmetic are those of the serial loop, so the
// polygons are bit-identical to it.
constexpr uint32_t kClipGroup = 16;  // lanes per triangle: a clipped polygon has at most kMaxClip vertices
static_assert(kMaxClip <= (int)kClipGroup, "one lane per polygon vertex");

struct ClipLanes {
    uint32_t lane;   // lane in the wave
    uint32_t li;     // lane in the group = the vertex it holds
    uint32_t g0;     // the group's first lane
    uint64_t gmask;  // the group's lanes
    __device__ __forceinline__ ClipLanes() {
        lane = threadIdx.x & 63u;
        li = lane & (kClipGroup - 1u);
        g0 = lane & ~(kClipGroup - 1u);
        gmask = 0xFFFFull << g0;
    }
    __device__ __forceinline__ uint64_t below() const { return (1ull << lane) - 1ull; }
};

// One clip_by_plane (main.c:609-647, plane_d = 0) of the group's n-vertex polygon. `s` is the
// group's LDS, NC x 16 floats. Returns the new vertex count, or -1 where the serial loop gives up
// (an output slot beyond MAX_NUM_CLIP_VERTICES: its check at the top of each vertex and before
// each intersection fails exactly when more than kMaxClip vertices come out or kMaxClip are out
// before the last input vertex).
template <int NC>
__device__ __forceinline__ int clip_plane_group(const ClipLanes &g, float (&x)[NC], int n, const float *pl, float *s) {
    const bool act = g.li < (uint32_t)n;
    const float cur = dot4_seq(pl[0], pl[1], pl[2], pl[3], x[0], x[1], x[2], x[3]);
    const bool cin = cur > -0.0f;
    const uint64_t b_in = __ballot(act && cin) & g.gmask;
    const uint64_t b_act = (((1ull << (uint32_t)n) - 1ull) << g.g0);  // n <= 16
    if (b_in == b_act) return n;  // every vertex inside: the plane returns the polygon unchanged
    const uint32_t nx = g.g0 + (g.li + 1u < (uint32_t)n ? g.li + 1u : 0u);
    const float nd = __shfl(cur, (int)nx, 64);
    float xn[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) xn[c] = __shfl(x[c], (int)nx, 64);
    const bool nin = nd > -0.0f;
    const bool keep = act && cin, cross = act && (cin != nin);
    const uint64_t b1 = __ballot(keep) & g.gmask, b2 = __ballot(cross) & g.gmask;
    const uint32_t total = (uint32_t)(__popcll(b1) + __popcll(b2));
    const uint64_t before_last = (1ull << (g.g0 + (uint32_t)n - 1u)) - 1ull;
    const uint32_t k_last = (uint32_t)(__popcll(b1 & before_last) + __popcll(b2 & before_last));
    if (total > (uint32_t)kMaxClip || k_last >= (uint32_t)kMaxClip) return -1;
    const uint32_t k = (uint32_t)(__popcll(b1 & g.below()) + __popcll(b2 & g.below()));
    if (keep) {
#pragma unroll
        for (int c = 0; c < NC; ++c) s[c * kClipGroup + k] = x[c];
    }
    if (cross) {
        const float t = (0.0f + cur) / (cur - nd);
        const float omt = 1.0f - t;
        const uint32_t kc = k + (keep ? 1u : 0u);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const float a = x[c] * omt;
            const float b = xn[c] * t;
            s[c * kClipGroup + kc] = a + b;
        }
    }
    // a wave's LDS operations complete in issue order; the barrier keeps the compiler from
    // moving the reads above the writes (or the next plane's writes above these reads)
    __builtin_amdgcn_wave_barrier();
    if (g.li < total) {
#pragma unroll
        for (int c = 0; c < NC; ++c) x[c] = s[c * kClipGroup + g.li];
    }
    __builtin_amdgcn_wave_barrier();
    return (int)total;
}

// The clipper (main.c:649-660) on triangle t of draw d: loads the three vertices into the
// group's first three lanes and runs the six planes. Returns the polygon's vertex count (0 when
// it is empty or the clipper gave up); vertex i is left in lane i of the group.
template <int NC>
__device__ __forceinline__ int clip_group(const BatchParams &p, const DrawRec &d, uint32_t t, const ClipLanes &g,
                                          float (&x)[NC], float *s) {
    uint32_t vi[3];
    if (!load_tri_indices(d, t, vi)) return 0;  // the same for the whole group
    const uint32_t v = g.li == 1u ? vi[1] : g.li == 2u ? vi[2] : vi[0];
    const float4 q = p.vpos[d.vbase + v];
    x[0] = q.x; x[1] = q.y; x[2] = q.z; x[3] = q.w;
    if constexpr (NC > 4) {
        const uint32_t row = d.abase + v;  // registers 1-2 (DrawRec::abase)
        const float4 y = p.vattr[2u * row], z = p.vattr[2u * row + 1u];
        x[4] = y.x; x[5] = y.y; x[6] = y.z; x[7] = y.w;
        x[8] = z.x; x[9] = z.y; x[10] = z.z; x[11] = z.w;
    }
    int n = 3;
    for (int pl = 0; pl < 6; ++pl) {
        n = clip_plane_group<NC>(g, x, n, p.planes[pl], s);
        if (n < 0) return 0;
    }
    return n;
}

// Fan triangle (0, li, li + 1) of the group's polygon (main.c:795-798): the first vertex's
// and the next vertex's positions, read from their lanes.
template <int NC>
__device__ __forceinline__ void fan_positions(const ClipLanes &g, const float (&x)[NC], float4 &p0, float4 &p1,
                                              float4 &p2) {
    const uint32_t nx = g.g0 + (g.li + 1u < kClipGroup ? g.li + 1u : g.li);
    p0 = make_float4(__shfl(x[0], (int)g.g0, 64), __shfl(x[1], (int)g.g0, 64), __shfl(x[2], (int)g.g0, 64),
                     __shfl(x[3], (int)g.g0, 64));
    p1 = make_float4(x[0], x[1], x[2], x[3]);
    p2 = make_float4(__shfl(x[0], (int)nx, 64), __shfl(x[1], (int)nx, 64), __shfl(x[2], (int)nx, 64),
                     __shfl(x[3], (int)nx, 64));
}

// ------------------------------------------------------------ fills

// The image paint_window shows (main.c:286-311): the frame buffer as a top-down 32-bit DIB with
// red mask 0x00FF0000, green 0x0000FF00 and blue 0x000000FF, as packed 8-bit RGB. Each thread
// converts 4 pixels into 12 bytes (three words of an output padded to whole words).
__global__ void k_present_rgb(const uint32_t *fb, uint32_t *rgb_words, size_t n_pixels) {
    const size_t n_quads = (n_pixels + 3u) / 4u;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_quads; q += (size_t)gridDim.x * blockDim.x) {
        uint8_t b[12];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const size_t i = 4u * q + (size_t)k;
            const uint32_t v = i < n_pixels ? fb[i] : 0u;
            b[3 * k + 0] = (uint8_t)(v >> 16);
            b[3 * k + 1] = (uint8_t)(v >> 8);
            b[3 * k + 2] = (uint8_t)v;
        }
#pragma unroll
        for (int w = 0; w < 3; ++w)
            rgb_words[3 * q + w] = (uint32_t)b[4 * w] | ((uint32_t)b[4 * w + 1] << 8) | ((uint32_t)b[4 * w + 2] << 16) |
                                   ((uint32_t)b[4 * w + 3] << 24);
    }
}

// load_texture's pre-pass for sRGB RGBA8 textures (main.c:546-558), in place. A texel's four bytes
// each map through the same 256-entry table; every block first builds it in LDS, one byte per
// thread, with the reference's arithmetic: decode byte * f32(1/255) (math.h:326-334),
// srgb_to_linear in double (math.h:386-394) stored as f32, truncating re-encode (math.h:322-324).
__global__ __launch_bounds__(256) void k_srgb_to_linear_rgba8(uint32_t *texels, size_t n) {
    __shared__ uint32_t lut[256];
    {
        const float v = (float)threadIdx.x * (float)(1.0 / 255.0);
        const float lin = (double)v <= 0.04045 ? (float)((double)v / 12.92) : (float)pow(((double)v + 0.055) / 1.055, 2.4);
        const float x = lin * 255.f;
        lut[threadIdx.x] = (x > 0.0f && x < 4294967296.0f) ? (uint32_t)x : 0u;
    }
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t t = texels[i];
        texels[i] = lut[t & 255u] | (lut[(t >> 8) & 255u] << 8) | (lut[(t >> 16) & 255u] << 16) | (lut[t >> 24] << 24);
    }
}

// Fill of n 32-bit words: words up to the first 16-byte boundary, then 16-byte stores, then the
// remaining words.
__device__ __forceinline__ void fill_one(uint32_t *d, uint32_t v, size_t n, size_t tid, size_t stride) {
    const size_t head = min(n, (size_t)(((16u - ((uintptr_t)d & 15u)) & 15u) / 4u));
    if (tid < head) d[tid] = v;
    uint4 *body = (uint4 *)(d + head);
    const size_t nb = (n - head) / 4u;
    const uint4 v4 = make_uint4(v, v, v, v);
    for (size_t i = tid; i < nb; i += stride) body[i] = v4;
    const size_t tail = head + nb * 4u;
    if (tid < n - tail) d[tail + tid] = v;
}

// Up to kMaxFills buffers filled by one launch (a clear's colour and visibility buffers, or its
// depth and hierarchical-Z values).
__global__ __launch_bounds__(256) void k_fill32(FillJobs f) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    for (uint32_t k = 0; k < f.count; ++k) fill_one(f.d[k], f.v[k], f.n[k], tid, stride);
}

// ------------------------------------------------------------ scans
// Exclusive scan over n elements (n read from the device when n_dev is given) in one launch:
// chained partitions with decoupled look-back. Block b scans its 4096 elements, publishes its
// aggregate, reads its predecessors' published aggregates 64 at a time (one per lane of its
// first wave) back to the nearest inclusive prefix, publishes its own inclusive prefix and
// writes its outputs. A block waits only for blocks with smaller indices, which the dispatcher
// starts first and which publish their aggregate without waiting, so every wait ends. Status
// words carry the epoch of the scan they belong to, so the status array never needs clearing.
// The wait is still bounded: a predecessor that never published would be recorded in the
// error word and counted as zero rather than hold the wave. (A shared ticket counter would
// make the order explicit, but hundreds of blocks hitting one address serialise in L2.)

constexpr uint32_t kScanItems = 16;
constexpr uint32_t kScanChunk = 256 * kScanItems;
constexpr uint32_t kScanHeader = 256;            // bytes before the status array: the ticket
constexpr uint32_t kFlagAgg = 1u, kFlagPrefix = 2u;
constexpr uint32_t kSpinLimit = 1u << 20;

// Status of one partition: its aggregate (flag kFlagAgg) or its inclusive prefix (kFlagPrefix),
// tagged with the scan's epoch. A 32-bit scan packs epoch, flag and value into one 64-bit word
// that is published and polled with single relaxed atomics at device scope: no cache-wide
// release/acquire fences, which would write back or invalidate the whole L2 at every step.
// A 64-bit scan keeps value and flag apart and orders them with release/acquire (its arrays are
// short: one value per 256 triangles).
template <class T>
struct ScanPub;

template <>
struct ScanPub<uint32_t> {
    static constexpr size_t kBytes = 8;
    static __device__ __forceinline__ void publish(uint8_t *st, uint32_t i, uint32_t tagflag, uint32_t v) {
        __hip_atomic_store((unsigned long long *)st + i, ((unsigned long long)tagflag << 32) | v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    static __device__ __forceinline__ uint32_t poll(uint8_t *st, uint32_t i, uint32_t &v) {
        const unsigned long long w =
            __hip_atomic_load((unsigned long long *)st + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v = (uint32_t)w;
        return (uint32_t)(w >> 32);
    }
};

template <>
struct ScanPub<unsigned long long> {
    struct Rec {
        uint32_t flag;
        uint32_t pad;
        unsigned long long agg, prefix;  // apart: a reader that saw kFlagAgg may still be reading agg
    };
    static constexpr size_t kBytes = sizeof(Rec);
    static __device__ __forceinline__ void publish(uint8_t *st, uint32_t i, uint32_t tagflag, unsigned long long v) {
        Rec *r = (Rec *)st + i;
        if ((tagflag & 3u) == kFlagPrefix) r->prefix = v;
        else r->agg = v;
        __hip_atomic_store(&r->flag, tagflag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    static __device__ __forceinline__ uint32_t poll(uint8_t *st, uint32_t i, unsigned long long &v) {
        Rec *r = (Rec *)st + i;
        const uint32_t f = __hip_atomic_load(&r->flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        v = ((f & 3u) == kFlagPrefix) ? r->prefix : r->agg;
        return f;
    }
};

// NT threads per partition of NT * 16 elements: 256 in general, 1024 when one partition holds
// the whole input (then nothing looks back).
template <class T, uint32_t NT>
__global__ __launch_bounds__(NT) void k_scan(const T *in, T *out, size_t n_host, const uint32_t *n_dev, T *total,
                                             uint8_t *st, uint32_t epoch, uint32_t nparts, uint32_t *err) {
    constexpr uint32_t kChunk = NT * kScanItems;
    __shared__ T s_prefix;
    __shared__ T s_w[NT / 64u + 1u];
    const uint32_t part = blockIdx.x;
    const size_t n = n_dev ? min(n_host, (size_t)*n_dev) : n_host;
    const uint32_t nd = (uint32_t)((n + kChunk - 1) / kChunk);  // partitions holding data
    if (part >= nd) {
        // nothing to scan here, and no partition with data looks back this far
        if (n == 0 && part == 0u && threadIdx.x == 0 && total) *total = T(0);
        return;
    }
    const size_t i0 = (size_t)part * kChunk + (size_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    T sum = T(0);
    // a thread's 16 items are contiguous: they go as 16-byte loads where all are in range
    const bool whole = i0 + kScanItems <= n;
    if constexpr (sizeof(T) == 8) {
        if (whole) {
#pragma unroll
            for (uint32_t q = 0; q < kScanItems / 2u; ++q) {
                const ulonglong2 u = ((const ulonglong2 *)(in + i0))[q];
                v[2 * q] = u.x;
                v[2 * q + 1] = u.y;
            }
        }
    } else {
        if (whole) {
#pragma unroll
            for (uint32_t q = 0; q < kScanItems / 4u; ++q) {
                const uint4 u = ((const uint4 *)(in + i0))[q];
                v[4 * q] = u.x;
                v[4 * q + 1] = u.y;
                v[4 * q + 2] = u.z;
                v[4 * q + 3] = u.w;
            }
        }
    }
    if (!whole) {
#pragma unroll
        for (uint32_t k = 0; k < kScanItems; ++k) v[k] = (i0 + k < n) ? in[i0 + k] : T(0);
    }
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; ++k) sum += v[k];
    T agg;
    const T ex = block_exclusive_scan<NT, T>(sum, s_w, agg);
    if (threadIdx.x < 64u) {  // wave 0 publishes and looks back, 64 predecessors at a time
        const uint32_t lane = threadIdx.x;
        const uint32_t tag = epoch << 2;
        T prefix = T(0);
        if (part == 0u) {
            // with one partition nobody looks back (and a 64-bit publish is a release, which
            // writes the L2 back)
            if (lane == 0u && nd > 1u) ScanPub<T>::publish(st, 0u, tag | kFlagPrefix, agg);
        } else {
            if (lane == 0u) ScanPub<T>::publish(st, part, tag | kFlagAgg, agg);
            int64_t j = (int64_t)part - 1;
            for (;;) {
                const int64_t idx = j - (int64_t)lane;  // lane 0 is the nearest predecessor
                uint32_t f = 0u;
                T val = T(0);
                if (idx >= 0) {
                    uint32_t spins = 0u;
                    for (;;) {
                        f = ScanPub<T>::poll(st, (uint32_t)idx, val);
                        if ((f & ~3u) == tag && (f & 3u) != 0u) break;
                        if (++spins > kSpinLimit) {
                            atomicAdd(err, 1u);
                            f = tag | kFlagPrefix;  // give up: count nothing from here on
                            val = T(0);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                // partition 0 always holds a prefix, so a window reaching it finds one
                const uint64_t pre = __ballot(idx >= 0 && (f & 3u) == kFlagPrefix);
                const uint32_t first = pre ? (uint32_t)__builtin_ctzll(pre) : 64u;
                prefix += wave_sum((idx >= 0 && lane <= first) ? val : T(0));
                if (pre) break;
                j -= 64;
            }
            if (lane == 0u) ScanPub<T>::publish(st, part, tag | kFlagPrefix, prefix + agg);
        }
        if (lane == 0u) {
            s_prefix = prefix;
            if (part == nd - 1u && total) *total = prefix + agg;
        }
    }
    __syncthreads();
    T e = s_prefix + ex;
    if constexpr (sizeof(T) == 8) {
        if (whole) {
#pragma unroll
            for (uint32_t q = 0; q < kScanItems / 2u; ++q) {
                ulonglong2 u;
                u.x = e;
                e += v[2 * q];
                u.y = e;
                e += v[2 * q + 1];
                ((ulonglong2 *)(out + i0))[q] = u;
            }
        }
    } else {
        if (whole) {
#pragma unroll
            for (uint32_t q = 0; q < kScanItems / 4u; ++q) {
                uint4 u;
                u.x = e;
                e += v[4 * q];
                u.y = e;
                e += v[4 * q + 1];
                u.z = e;
                e += v[4 * q + 2];
                u.w = e;
                e += v[4 * q + 3];
                ((uint4 *)(out + i0))[q] = u;
            }
        }
    }
    if (!whole) {
#pragma unroll
        for (uint32_t k = 0; k < kScanItems; ++k) {
            if (i0 + k < n) out[i0 + k] = e;
            e += v[k];
        }
    }
}

template <class T>
static hipError_t launch_scan(const T *in, T *out, size_t n_host, const uint32_t *n_dev, const ScanScratch &sc,
                              T *total, hipStream_t s) {
    if (n_host == 0) return hipSuccess;
    const uint32_t nparts = (uint32_t)((n_host + kScanChunk - 1) / kScanChunk);
    constexpr uint32_t kOneBlock = 1024u * kScanItems;  // one 1024-thread partition: no look-back
    // 64-bit statuses keep their own region: a 32-bit scan reading one of their value words
    // could take its upper half for a status of its own epoch
    const size_t off = sizeof(T) == 4 ? kScanHeader : sc.wide_off;
    if (off == 0) return hipErrorInvalidValue;
    const uint32_t epoch = (++*sc.epoch) & 0x3FFFFFFFu;
    if (n_host <= kOneBlock)
        hipLaunchKernelGGL((k_scan<T, 1024>), dim3(1), dim3(1024), 0, s, in, out, n_host, n_dev, total,
                           (uint8_t *)sc.part + off, epoch, 1u, sc.err);
    else
        hipLaunchKernelGGL((k_scan<T, 256>), dim3(nparts), dim3(256), 0, s, in, out, n_host, n_dev, total,
                           (uint8_t *)sc.part + off, epoch, nparts, sc.err);
    return hipGetLastError();
}

// Bytes of scan scratch for scans over up to n_max elements: the header, one 32-bit scan status
// per partition, then one 64-bit scan status per partition (from scan_wide_offset).
size_t scan_scratch_bytes(size_t n_max) {
    return scan_wide_offset(n_max) + ((n_max + kScanChunk - 1) / kScanChunk + 1) * ScanPub<unsigned long long>::kBytes;
}

size_t scan_wide_offset(size_t n_max) {
    return kScanHeader + ((n_max + kScanChunk - 1) / kScanChunk + 1) * ScanPub<uint32_t>::kBytes;
}


// ------------------------------------------------------------ stage 0: vertex

// A draw's inputs to the vertex stage, read by kb_vertex before its lanes diverge and before its
// first store, so that every field and constant is a scalar load.
struct VsDraw {
    const uint8_t *vb;
    uint32_t stride, vbase, vertex_count;
    bool attrs;    // registers 1-2 are written: the draw's pixel shader reads them (all but MLV_PS_NONE)
                   // and no earlier draw of the flush wrote them for the same vertices
    float cb[44];  // constant buffer slot 0, floats 0..43 (those the draw's shader reads)
    TexDesc tex;
};

template <int VS>
__device__ __forceinline__ void run_vs(const BatchParams &p, const VsDraw &u, uint32_t i) {
    if (i >= u.vertex_count) return;
    VsConstants<VS> k;
#pragma unroll
    for (int j = 0; j < VsConstants<VS>::kFloats; ++j) k.cb[j] = u.cb[j];
    k.tex = u.tex;
    float in[8];
    const uint8_t *src = u.vb + (size_t)i * u.stride;
    if ((((uintptr_t)u.vb) & 15u) == 0u && (u.stride & 15u) == 0u) {
        // a draw that writes positions only reads the first 16 bytes of each vertex (u.attrs is
        // uniform over the draw)
        const float4 a = load_global((const float4 *)src);
        const float4 b = u.attrs ? load_global((const float4 *)src + 1) : make_float4(0.f, 0.f, 0.f, 0.f);
        in[0] = a.x; in[1] = a.y; in[2] = a.z; in[3] = a.w; in[4] = b.x; in[5] = b.y; in[6] = b.z; in[7] = b.w;
    } else {
        for (int k2 = 0; k2 < 8; ++k2) in[k2] = load_global((const float *)src + k2);
    }
    float4 r0, r1, r2;
    const uint32_t slot = u.vbase + i;
    if (u.attrs) {
        vertex_shader<VS, true>(k, in, r0, r1, r2);
        p.vattr[2u * slot] = r1;
        p.vattr[2u * slot + 1u] = r2;
    } else {
        vertex_shader<VS, false>(k, in, r0, r1, r2);
    }
    p.vpos[slot] = r0;
}

// Each vertex-stage and setup block's draw, one workgroup per draw: the vertex and setup kernels
// then find their draw with one load instead of a binary search over the draws' first blocks
// (a chain of dependent loads at the head of every block: eleven for 2,048 draws). Launched for
// flushes of kBlockDrawTable draws or more; below, the search is shorter than the extra launch.
__global__ __launch_bounds__(256) void kb_block_draws(BatchParams p) {
    const uint32_t d = blockIdx.x;
    const DrawRec &r = p.draws[d];
    const uint32_t nv = (r.vertex_count + 255u) / 256u, nt = (r.tri_count + 255u) / 256u;
    for (uint32_t i = threadIdx.x; i < nv; i += 256u) p.vblk_draw[r.blk_v0 + i] = d;
    for (uint32_t i = threadIdx.x; i < nt; i += 256u) p.tblk_draw[r.blk_t0 + i] = d;
}

// A flush's counters, tile ranges and scan status words start at zero.
__device__ __forceinline__ void reset_flush(const BatchParams &p) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        BatchCounters z = {};
        *p.ctr = z;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.total_tiles; i += gridDim.x * blockDim.x) {
        p.tile_start[i] = 0u;
        p.tile_cnt[i] = 0u;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.scan_status_words; i += gridDim.x * blockDim.x)
        p.scan_status[i] = 0u;
}

// Stage 0 in one launch: blocks below n_vblocks take one lane per vertex slot of every draw,
// and every block then takes a share of the flush's reset. The reset's stores come last: a
// load that no store of the kernel can precede is read through the scalar cache (the draw
// search, the draw record and the constant buffer), and after a store it would be a vector
// load on the wave's critical path.
template <bool kTable>  // the block's draw from BatchParams::vblk_draw, else by a binary search
__global__ __launch_bounds__(256) void kb_vertex(BatchParams p) {
    const uint32_t b = blockIdx.x;
    if (b < p.n_vblocks) {
        const uint32_t di = kTable ? p.vblk_draw[b]
                                   : upper_index(p.n_draws, b, [&](uint32_t i) { return p.draw_v0[i]; });
        const DrawRec &d = p.draws[di];
        const uint32_t i = (b - d.blk_v0) * 256u + threadIdx.x;
        const int vs = d.vs;
        VsDraw u;
        u.vb = d.vb;
        u.stride = d.stride;
        u.vbase = d.vbase;
        u.vertex_count = d.vertex_count;
        u.attrs = d.ps != PS_NONE && d.abase == d.vbase;
        u.tex = d.vs_tex;
        const float *cb = d.cb;  // unbound (null) for the passthrough shader
        if (vs == VS_FULLSCREEN) {
#pragma unroll
            for (int j = 0; j < 44; ++j) u.cb[j] = load_global(cb + j);
        } else if (vs != VS_PASSTHROUGH) {
#pragma unroll
            for (int j = 0; j < 16; ++j) u.cb[j] = load_global(cb + j);
        }
        switch (vs) {
        case VS_BASIC: run_vs<VS_BASIC>(p, u, i); break;
        case VS_VERTEX_LIGHTING: run_vs<VS_VERTEX_LIGHTING>(p, u, i); break;
        case VS_PASSTHROUGH: run_vs<VS_PASSTHROUGH>(p, u, i); break;
        default: run_vs<VS_FULLSCREEN>(p, u, i); break;
        }
    }
    reset_flush(p);
}

// ------------------------------------------------------------ stage 1: setup


// kCountPer 256-slot blocks per workgroup: each lane takes one triangle of each, and all the
// triangles' loads are issued before any is set up, so the two dependent round trips (indices,
// then positions) are paid once for all of them. Two on small flushes; four on flushes of many
// draws (kBlockDrawTable or more), which have millions of triangles: there the count pass is
// bound by its arithmetic, and fewer, longer workgroups run it 6 % faster (the 1080p frame's
// count pass is 3 % slower that way).
template <bool kBig>
constexpr uint32_t kCountPer = kBig ? 4u : 2u;

// The bins of the workgroup's assembled triangles that are not binned (rect of triangle h of this
// lane in draw d[h] when unb[h]): counted for the stats and marked in the (draw, bin) bitmap. The
// workgroup merges its bits in LDS first, in a window of kMarkWindow words from its lowest marked
// word (bits beyond it, rare for consecutive triangles of one mesh, go to memory directly), so
// that each word it touches takes one atomic: neighbouring triangles meet the same few words, and
// atomics on one word serialise (one atomic per wave and word instead, on 32 views of 64 meshes:
// 2.7 ms per setup launch, against 1.18 ms with the merge and 0.96 ms without any marks; one
// atomic per lane and bin: 11.5 ms). Every thread of the 256-thread workgroup must call it.
constexpr uint32_t kMarkWindow = 256;
static_assert(kMarkWindow == (uint32_t)kBlock, "one window word per thread of a setup workgroup");

template <uint32_t N>
__device__ __forceinline__ void mark_unbinned(const BatchParams &p, const uint32_t d[N], const bool unb[N],
                                              const uint32_t r0[N], const uint32_t r1[N]) {
    __shared__ uint32_t s_win[kMarkWindow];
    __shared__ uint32_t s_lo[4], s_p8[4];
    uint32_t lo_t = 0xFFFFFFFFu, nb_t = 0u, cbase[N], bins_x[N];
#pragma unroll
    for (uint32_t h = 0; h < N; ++h) {
        const DrawRec &D = p.draws[d[h]];
        cbase[h] = D.cbase;
        bins_x[h] = (uint32_t)p.views[D.view].bins_x;
        if (unb[h]) {
            int32_t bx0, by0, bx1, by1;
            rect_decode(r0[h], r1[h], bx0, by0, bx1, by1);
            if (bx1 >= bx0 && by1 >= by0) {
                nb_t += (uint32_t)(bx1 - bx0 + 1) * (uint32_t)(by1 - by0 + 1);
                lo_t = min(lo_t, (cbase[h] + (uint32_t)by0 * bins_x[h] + (uint32_t)bx0) >> 5);
            }
        }
    }
    const uint32_t lo_w = wave_min(lo_t);
    const uint32_t n_w = wave_sum(nb_t);
    s_win[threadIdx.x] = 0u;
    if ((threadIdx.x & 63u) == 0u) {
        s_lo[threadIdx.x >> 6] = lo_w;
        s_p8[threadIdx.x >> 6] = n_w;
    }
    __syncthreads();
    const uint32_t lo = min(min(s_lo[0], s_lo[1]), min(s_lo[2], s_lo[3]));
    if (lo == 0xFFFFFFFFu) return;  // nothing to mark in this workgroup (uniform)
#pragma unroll
    for (uint32_t h = 0; h < N; ++h) {
        if (!unb[h]) continue;
        int32_t bx0, by0, bx1, by1;
        rect_decode(r0[h], r1[h], bx0, by0, bx1, by1);
        for (int32_t by = by0; by <= by1; ++by)
            for (int32_t bx = bx0; bx <= bx1; ++bx) {
                const uint32_t bit = cbase[h] + (uint32_t)by * bins_x[h] + (uint32_t)bx;
                const uint32_t w = (bit >> 5) - lo;
                if (w < kMarkWindow) atomicOr(&s_win[w], 1u << (bit & 31u));
                else atomicOr(p.cmark + (bit >> 5), 1u << (bit & 31u));
            }
    }
    __syncthreads();
    const uint32_t m = s_win[threadIdx.x];
    if (m) atomicOr(p.cmark + lo + threadIdx.x, m);
    if (threadIdx.x == 0u)
        atomicAdd(&p.stats->pairs8_part[blockIdx.x & 63u], (unsigned long long)(s_p8[0] + s_p8[1] + s_p8[2] + s_p8[3]));
}

// kTable: the blocks' draws from BatchParams::tblk_draw (and four blocks a workgroup), else by a
// binary search (and two)
template <bool kTable>
__global__ __launch_bounds__(256) void kb_setup_count(BatchParams p) {
    constexpr uint32_t kPer = kCountPer<kTable>;
    __shared__ uint32_t s_w[kPer][4];
    const uint32_t hb = xcd_contiguous(blockIdx.x, gridDim.x);
    uint32_t di[kPer], slot[kPer], vi[kPer][3];
    bool inb[kPer], live[kPer], ok[kPer];
    // every load comes before the first atomic or store (see kb_vertex)
#pragma unroll
    for (uint32_t h = 0; h < kPer; ++h) {
        const uint32_t b = hb * kPer + h;
        inb[h] = b < p.n_tblocks;  // the grid's last workgroup may hold one block
        const uint32_t bc = inb[h] ? b : p.n_tblocks - 1u;
        di[h] = kTable ? p.tblk_draw[bc]
                       : upper_index(p.n_draws, bc, [&](uint32_t i) { return p.draw_t0[i]; });
        const DrawRec &d = p.draws[di[h]];
        const uint32_t t = (bc - d.blk_t0) * 256u + threadIdx.x;
        slot[h] = bc * 256u + threadIdx.x;
        live[h] = inb[h] && t < d.tri_count;
        // the raw indices only, and from every lane (a lane past the draw reads its last
        // triangle, which a block of the draw always has, and ignores it): a check here, or a
        // load under a branch, would wait for them before the next block's loads go out
        const uint32_t *ib = d.ib + 3u * min(t, d.tri_count - 1u);
        vi[h][0] = load_global(ib);
        vi[h][1] = load_global(ib + 1);
        vi[h][2] = load_global(ib + 2);
    }
#pragma unroll
    for (uint32_t h = 0; h < kPer; ++h) {
        // the range check of load_tri_indices
        const DrawRec &d = p.draws[di[h]];
        const long long bv = (long long)d.base_vertex, nv = (long long)d.vertex_count;
        bool in_range = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const long long v = (long long)vi[h][k] + bv;
            in_range = in_range && v >= 0 && v < nv;
            vi[h][k] = (uint32_t)v;
        }
        ok[h] = live[h] && in_range;
    }
    float4 q[kPer][3];
#pragma unroll
    for (uint32_t h = 0; h < kPer; ++h) {
        // from every lane as well (slot 0 where the triangle is not set up)
        const uint32_t vb = p.draws[di[h]].vbase;
#pragma unroll
        for (int k = 0; k < 3; ++k) q[h][k] = p.vpos[ok[h] ? vb + vi[h][k] : 0u];
    }
    uint32_t packed[kPer];  // binned | assembled << 16
    int cls[kPer];
    bool unb[kPer];         // assembled without the clipper and not binned: its bins are marked here
    uint32_t r0[kPer], r1[kPer];
#pragma unroll
    for (uint32_t h = 0; h < kPer; ++h) {
        packed[h] = 0u;
        cls[h] = 0;
        unb[h] = false;
        r0[h] = r1[h] = 0u;
        if (ok[h]) {
            const DrawRec &d = p.draws[di[h]];
            cls[h] = classify(q[h][0], q[h][1], q[h][2]);
            if (cls[h] == 1) {
                TriVis tv;
                float rw[3];
                if (setup_triangle<true>(d, p.views[d.view], q[h][0], q[h][1], q[h][2], tv, rw)) {
                    unb[h] = p.cmark && !has_corner(tv);
                    packed[h] = unb[h] ? 0x10000u : 0x10001u;
                    r0[h] = tv.rect0;
                    r1[h] = tv.rect1;
                }
            }
        }
    }
#pragma unroll
    for (uint32_t h = 0; h < kPer; ++h) {
        if (live[h] && !ok[h]) atomicAdd(&p.stats->oob, 1ull);
        if (cls[h] == 2) {
            uint32_t k = atomicAdd(&p.ctr->n_clip_list, 1u);
            p.clip_list[k] = make_uint2(slot[h], di[h]);
        }
        if (inb[h]) p.cnt[slot[h]] = (packed[h] & 0xFFu) | ((packed[h] >> 16) << 8);
        uint32_t sum = wave_sum(packed[h]);
        if ((threadIdx.x & 63u) == 0u) s_w[h][threadIdx.x >> 6] = sum;
    }
    __syncthreads();
    if (threadIdx.x < kPer && hb * kPer + threadIdx.x < p.n_tblocks) {
        const uint32_t sum = s_w[threadIdx.x][0] + s_w[threadIdx.x][1] + s_w[threadIdx.x][2] + s_w[threadIdx.x][3];
        p.blocksum[hb * kPer + threadIdx.x] =
            (unsigned long long)(sum & 0xFFFFu) | ((unsigned long long)(sum >> 16) << 32);
        const uint32_t bk = hb * kPer + threadIdx.x, dk = di[threadIdx.x];
        p.blk_draw[bk] = dk | (p.draws[dk].blk_t0 == bk ? kFirstOfDraw : 0u);
    }
    if (p.cmark) mark_unbinned<kPer>(p, di, unb, r0, r1);
}

// Triangles that need the clipper, one per 16-lane group: the polygon's fan triangles that
// survive setup are counted, and those of them to bin (packed with the polygon's vertex count).
__global__ __launch_bounds__(256) void kb_clip_count(BatchParams p) {
    __shared__ float s_poly[256 / kClipGroup][4 * kClipGroup];
    const ClipLanes g;
    const uint32_t grp = threadIdx.x / kClipGroup;
    const uint32_t n = p.ctr->n_clip_list;
    for (uint32_t s = blockIdx.x * (256u / kClipGroup) + grp; s < n; s += gridDim.x * (256u / kClipGroup)) {
        const uint2 e = p.clip_list[s];  // (slot, draw)
        const DrawRec &d = p.draws[e.y];
        const ViewRec &v = p.views[d.view];
        float x[4];
        const int nv = clip_group<4>(p, d, e.x - d.tbase, g, x, s_poly[grp]);  // positions decide count and cull
        float4 q0, q1, q2;
        fan_positions<4>(g, x, q0, q1, q2);
        bool live = false, binned = false;
        if (g.li >= 1u && (int)g.li < nv - 1) {
            TriVis tv;
            float rw[3];
            live = setup_triangle(d, v, q0, q1, q2, tv, rw);
            binned = live && (!p.cmark || has_corner(tv));
        }
        const uint32_t n_asm = (uint32_t)__popcll(__ballot(live) & g.gmask);
        const uint32_t n_bin = (uint32_t)__popcll(__ballot(binned) & g.gmask);
        if (g.li == 0u) {
            const uint32_t packed = n_asm ? (n_bin | (n_asm << 8) | ((uint32_t)nv << 16)) : 0u;
            p.cnt[e.x] = packed;
            if (packed)
                atomicAdd(&p.blocksum[e.x >> 8], (unsigned long long)n_bin | ((unsigned long long)n_asm << 32));
        }
    }
}


__device__ __forceinline__ void write_triangle(const BatchParams &p, uint32_t idx, uint32_t draw, const ViewRec &v,
                                               TriVis &tv, const float rw[3], uint32_t r0, uint32_t r1,
                                               uint32_t r2, uint32_t prim, uint32_t num) {
    if (idx >= p.tri_cap) return;
    // every load before the first store, which the compiler cannot move loads across
    const uint32_t tile_base = v.tile_base, tiles_x = (uint32_t)v.tiles_x;
    tv.draw = (tv.draw & 0x70000u) | draw;  // keeps the top-left rule's edge bits
    p.tvis[idx] = tv;
    TriShade sh;
    sh.rw[0] = rw[0]; sh.rw[1] = rw[1]; sh.rw[2] = rw[2];
    sh.row[0] = r0; sh.row[1] = r1; sh.row[2] = r2;
    sh.prim = prim;
    sh.num = num;
    p.tshade[idx] = sh;
    p.tbin[idx] = make_uint4(tv.rect0, tv.rect1, tile_base, tiles_x);
}

// Adds a block's count of unbinned (triangle, 8x8 bin) pairs to the stats (spread over 64
// counters: one shared counter hit by every block would serialise in L2).
__device__ __forceinline__ void add_unbinned_pairs(const BatchParams &p, uint32_t nb) {
    __shared__ uint32_t s_p8[4];
    const uint32_t s = wave_sum(nb);
    if ((threadIdx.x & 63u) == 0u) s_p8[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0u) {
        const uint32_t t = s_p8[0] + s_p8[1] + s_p8[2] + s_p8[3];
        if (t) atomicAdd(&p.stats->pairs8_part[blockIdx.x & 63u], (unsigned long long)t);
    }
}

__global__ __launch_bounds__(256) void kb_setup_emit(BatchParams p) {
    __shared__ unsigned long long s_w[5];
    __shared__ unsigned long long s_ex[256];
    __shared__ uint32_t s_nw[4];
    __shared__ uint8_t s_list[256];
    const uint32_t b = xcd_contiguous(blockIdx.x, gridDim.x);
    const uint32_t slot = b * 256u + threadIdx.x;
    const uint32_t bd = p.blk_draw[b];  // the count pass found the block's draw
    const uint32_t blk_binned = (uint32_t)p.blocksum[b];  // binned triangles of the block (clipped ones too)
    // a block without binned triangles writes nothing: no records, no clipper rows (a clipped
    // polygon takes rows only when one of its fan triangles is binned) and no clipper offsets
    // (kb_clip_emit reads them only for binned fan triangles); only a draw's first block and the
    // last block have per-draw and per-flush values to write
    if (blk_binned == 0u && !(bd & kFirstOfDraw) && b != gridDim.x - 1u) return;
    const uint32_t di = bd & ~kFirstOfDraw;
    const DrawRec &d = p.draws[di];
    const ViewRec v = p.views[d.view];  // a copy: read here, before the first store
    const unsigned long long bo = p.blockoff[b];                       // binned | assembled << 32
    const uint32_t draw_asm0 = (uint32_t)(p.blockoff[d.blk_t0] >> 32);  // the draw's first assembled
    const uint32_t packed = p.cnt[slot];
    const uint32_t n_bin = packed & 0xFFu, n_asm = (packed >> 8) & 0xFFu;
    const bool plain = n_asm == 1u && (packed >> 16) == 0u;  // assembled without the clipper
    const uint32_t n_rows = (!plain && n_bin) ? packed >> 16 : 0u;  // a clipped polygon's vertex rows
    // only binned triangles are written (kb_setup_count marked the bins of the others)
    const bool emit = plain && n_bin != 0u;
    if (threadIdx.x == 0u) {
        // a draw's first binned index; empty draws start where the next one does
        if (b == d.blk_t0)
            for (int32_t k = (int32_t)di; k >= 0 && p.draws[k].blk_t0 == b; --k) p.asm_base[k] = (uint32_t)bo;
        if (b == gridDim.x - 1u) {
            // the flush's totals and capacity check, and the empty draws after the last triangle
            const unsigned long long tot = p.ctr->scan_total;
            const uint32_t n_binned = (uint32_t)(tot & 0xFFFFFFFFull), n_assembled = (uint32_t)(tot >> 32);
            for (uint32_t k = di + 1u; k < p.n_draws; ++k) p.asm_base[k] = n_binned;
            p.ctr->n_assembled = n_binned;
            if (n_binned > p.tri_cap) {
                atomicOr(&p.ctr->overflow, OVF_TRIANGLES);
                atomicAdd(&p.stats->capacity_errors, 1ull);
            }
            atomicAdd(&p.stats->assembled, (unsigned long long)n_assembled);
            atomicMax(&p.stats->max_clip_list, (unsigned long long)p.ctr->n_clip_list);
        }
    }
    if (blk_binned == 0u) return;
    unsigned long long total;
    // within one block the counts (<= 14 triangles, <= 16 rows each) cannot carry out of 16 bits
    const unsigned long long ex = block_exclusive_scan<256, unsigned long long>(
        (unsigned long long)(n_bin | (n_asm << 16)) | ((unsigned long long)n_rows << 32), s_w, total);
    if (total >> 32) {  // the block's clipped polygons take their vertex rows with one atomic
        __shared__ uint32_t s_rows;
        if (threadIdx.x == 0u) s_rows = atomicAdd(&p.ctr->n_clipverts, (uint32_t)(total >> 32));
        __syncthreads();
        if (n_rows) p.clip_rows[slot] = s_rows + (uint32_t)(ex >> 32);
    }
    if (!plain && n_asm)
        p.clip_off[slot] = make_uint2((uint32_t)(bo & 0xFFFFFFFFull) + (uint32_t)(ex & 0xFFFFu),
                                      (uint32_t)(bo >> 32) + (uint32_t)((ex >> 16) & 0xFFFFu) - draw_asm0);
    // The triangles to write, packed into the block's first waves: where few of a block's
    // triangles are binned (small triangles without a pixel corner, back faces), the set-up work
    // runs on one or two waves instead of four.
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint64_t bal = __ballot(emit);
    if (lane == 0u) s_nw[w] = (uint32_t)__popcll(bal);
    s_ex[threadIdx.x] = ex;
    __syncthreads();
    uint32_t before = 0u;
    for (uint32_t k = 0; k < w; ++k) before += s_nw[k];
    const uint32_t n_emit = s_nw[0] + s_nw[1] + s_nw[2] + s_nw[3];
    if (emit) s_list[before + mbcnt(bal)] = (uint8_t)threadIdx.x;
    __syncthreads();
    if (threadIdx.x >= n_emit) return;
    const uint32_t tl = s_list[threadIdx.x];
    const uint32_t tt = b * 256u + tl - d.tbase;  // the triangle within its draw
    const unsigned long long ext = s_ex[tl];
    uint32_t vi[3];
    load_tri_indices(d, tt, vi);
    TriVis tv;
    float rw[3];
    setup_triangle<true>(d, v, p.vpos[d.vbase + vi[0]], p.vpos[d.vbase + vi[1]], p.vpos[d.vbase + vi[2]], tv, rw);
    const uint32_t off_bin = (uint32_t)(bo & 0xFFFFFFFFull) + (uint32_t)(ext & 0xFFFFu);
    const uint32_t num = (uint32_t)(bo >> 32) + (uint32_t)((ext >> 16) & 0xFFFFu) - draw_asm0;
    // the attribute rows the pixel shader reads (DrawRec::abase)
    write_triangle(p, off_bin, di, v, tv, rw, d.abase + vi[0], d.abase + vi[1], d.abase + vi[2], tt, num);
}

// kb_setup_emit for flushes of many draws (kBlockDrawTable or more: millions of small triangles,
// few of them binned): one wave per 256-slot block, four independent waves a workgroup and no
// workgroup barrier: lane
// l holds slots j * 64 + l (j < 4), the block's scan is four wave scans, and the binned triangles
// are packed into rounds of 64 through a per-wave LDS list. The pass is bound by each block's
// chain of dependent loads (its draw, the draw's view and first offset, then indices and
// corners), so what pays is many blocks in flight: a wave ends as soon as its block is done
// (batched views 663 -> 579 us per flush). Where many of a block's triangles are binned (the
// 1080p frame) one wave setting them all up in turn is slower than kb_setup_emit's four.
constexpr uint32_t kEmitWaves = 4;

__global__ __launch_bounds__(256) void kb_setup_emit_waves(BatchParams p) {
    __shared__ uint8_t s_list[kEmitWaves][256];
    __shared__ unsigned long long s_ex[kEmitWaves][256];
    const uint32_t lane = threadIdx.x & 63u, wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t b = xcd_contiguous(blockIdx.x, gridDim.x) * kEmitWaves + wv;
    if (b >= p.n_tblocks) return;
    const uint32_t bd = p.blk_draw[b];  // the count pass found the block's draw
    const uint32_t blk_binned = (uint32_t)p.blocksum[b];  // binned triangles of the block (clipped ones too)
    const bool last = b == p.n_tblocks - 1u;
    // a block without binned triangles writes nothing: no records, no clipper rows (a clipped
    // polygon takes rows only when one of its fan triangles is binned) and no clipper offsets
    // (kb_clip_emit reads them only for binned fan triangles); only a draw's first block and the
    // last block have per-draw and per-flush values to write
    if (blk_binned == 0u && !(bd & kFirstOfDraw) && !last) return;
    const uint32_t di = bd & ~kFirstOfDraw;
    const DrawRec &d = p.draws[di];
    const ViewRec v = p.views[d.view];  // a copy: read here, before the first store
    const unsigned long long bo = p.blockoff[b];                       // binned | assembled << 32
    const uint32_t draw_asm0 = (uint32_t)(p.blockoff[d.blk_t0] >> 32);  // the draw's first assembled
    uint32_t packed[4];
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) packed[j] = p.cnt[b * 256u + j * 64u + lane];
    if (lane == 0u) {
        // a draw's first binned index; empty draws start where the next one does
        if (b == d.blk_t0)
            for (int32_t k = (int32_t)di; k >= 0 && p.draws[k].blk_t0 == b; --k) p.asm_base[k] = (uint32_t)bo;
        if (last) {
            // the flush's totals and capacity check, and the empty draws after the last triangle
            const unsigned long long tot = p.ctr->scan_total;
            const uint32_t n_binned = (uint32_t)(tot & 0xFFFFFFFFull), n_assembled = (uint32_t)(tot >> 32);
            for (uint32_t k = di + 1u; k < p.n_draws; ++k) p.asm_base[k] = n_binned;
            p.ctr->n_assembled = n_binned;
            if (n_binned > p.tri_cap) {
                atomicOr(&p.ctr->overflow, OVF_TRIANGLES);
                atomicAdd(&p.stats->capacity_errors, 1ull);
            }
            atomicAdd(&p.stats->assembled, (unsigned long long)n_assembled);
            atomicMax(&p.stats->max_clip_list, (unsigned long long)p.ctr->n_clip_list);
        }
    }
    if (blk_binned == 0u) return;
    // the block's exclusive scan in slot order (within one block the counts, <= 14 triangles and
    // <= 16 rows a slot, cannot carry out of 16 bits)
    unsigned long long ex[4], carry = 0ull;
    bool emit[4], plain[4];
    uint32_t n_rows[4], n_asm[4];
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) {
        const uint32_t n_bin = packed[j] & 0xFFu;
        n_asm[j] = (packed[j] >> 8) & 0xFFu;
        plain[j] = n_asm[j] == 1u && (packed[j] >> 16) == 0u;  // assembled without the clipper
        n_rows[j] = (!plain[j] && n_bin) ? packed[j] >> 16 : 0u;  // a clipped polygon's vertex rows
        emit[j] = plain[j] && n_bin != 0u;  // only binned triangles are written (count marked the others' bins)
        const unsigned long long x =
            (unsigned long long)(n_bin | (n_asm[j] << 16)) | ((unsigned long long)n_rows[j] << 32);
        const unsigned long long inc = wave_inclusive_scan(x);
        ex[j] = carry + inc - x;
        carry += (unsigned long long)__builtin_amdgcn_readlane((int)(uint32_t)inc, 63) |
                 ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(inc >> 32), 63) << 32);
    }
    if (carry >> 32) {  // the block's clipped polygons take their vertex rows with one atomic
        uint32_t rows0 = 0u;
        if (lane == 0u) rows0 = atomicAdd(&p.ctr->n_clipverts, (uint32_t)(carry >> 32));
        rows0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)rows0);
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j)
            if (n_rows[j]) p.clip_rows[b * 256u + j * 64u + lane] = rows0 + (uint32_t)(ex[j] >> 32);
    }
    // clipped polygons' offsets for kb_clip_emit; the binned triangles packed into s_list
    uint32_t n_emit = 0u;
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) {
        const uint32_t sl = j * 64u + lane;
        if (!plain[j] && n_asm[j])
            p.clip_off[b * 256u + sl] = make_uint2((uint32_t)(bo & 0xFFFFFFFFull) + (uint32_t)(ex[j] & 0xFFFFu),
                                                   (uint32_t)(bo >> 32) + (uint32_t)((ex[j] >> 16) & 0xFFFFu) -
                                                       draw_asm0);
        const uint64_t bal = __ballot(emit[j]);
        if (emit[j]) {
            const uint32_t r = n_emit + mbcnt(bal);
            s_list[wv][r] = (uint8_t)sl;
            s_ex[wv][r] = ex[j];
        }
        n_emit += (uint32_t)__popcll(bal);
    }
    __builtin_amdgcn_wave_barrier();  // the wave's own LDS writes come before its reads
    for (uint32_t r0 = 0; r0 < n_emit; r0 += 64u) {
        const uint32_t i = r0 + lane;
        if (i >= n_emit) break;
        const uint32_t tt = b * 256u + s_list[wv][i] - d.tbase;  // the triangle within its draw
        const unsigned long long ext = s_ex[wv][i];
        uint32_t vi[3];
        load_tri_indices(d, tt, vi);
        TriVis tv;
        float rw[3];
        setup_triangle<true>(d, v, p.vpos[d.vbase + vi[0]], p.vpos[d.vbase + vi[1]], p.vpos[d.vbase + vi[2]], tv,
                             rw);
        const uint32_t off_bin = (uint32_t)(bo & 0xFFFFFFFFull) + (uint32_t)(ext & 0xFFFFu);
        const uint32_t num = (uint32_t)(bo >> 32) + (uint32_t)((ext >> 16) & 0xFFFFu) - draw_asm0;
        // the attribute rows the pixel shader reads (DrawRec::abase)
        write_triangle(p, off_bin, di, v, tv, rw, d.abase + vi[0], d.abase + vi[1], d.abase + vi[2], tt, num);
    }
}

// The clipped triangles' output, one per 16-lane group: lane k writes polygon vertex k's attribute
// row (when any fan triangle is binned: the rows kb_setup_emit gave it), and lane ci the fan
// triangle (0, ci, ci + 1) if it survives setup, at the offsets kb_setup_emit scanned.
// Unbinned fan triangles go to the stats and the bitmap, one atomic per bin (the clipper is rare).
__global__ __launch_bounds__(256) void kb_clip_emit(BatchParams p) {
    __shared__ float s_poly[256 / kClipGroup][12 * kClipGroup];
    const ClipLanes g;
    const uint32_t grp = threadIdx.x / kClipGroup;
    const uint32_t n = p.ctr->n_clip_list;
    uint32_t p8 = 0u;
    for (uint32_t s = blockIdx.x * (256u / kClipGroup) + grp; s < n; s += gridDim.x * (256u / kClipGroup)) {
        const uint2 e = p.clip_list[s];  // (slot, draw)
        const uint32_t c = p.cnt[e.x];
        if (((c >> 8) & 0xFFu) == 0u) continue;  // nothing assembled
        const uint32_t di = e.y;
        const DrawRec &d = p.draws[di];
        const ViewRec &v = p.views[d.view];
        const uint2 off = p.clip_off[e.x];
        const uint32_t t = e.x - d.tbase;
        float x[12];
        const int nv = clip_group<12>(p, d, t, g, x, s_poly[grp]);
        uint32_t row0 = 0u;
        bool rows = false;
        if (c & 0xFFu) {
            const uint32_t r = p.clip_rows[e.x];
            rows = (unsigned long long)p.vslots + r + (uint32_t)nv <= (unsigned long long)p.vattr_rows;
            if (!rows && g.li == 0u) {
                const uint32_t old = atomicOr(&p.ctr->overflow, OVF_CLIPVERTS);
                if (!(old & (OVF_TRIANGLES | OVF_CLIPVERTS))) atomicAdd(&p.stats->capacity_errors, 1ull);
            }
            row0 = p.vslots + r;
        }
        const uint32_t row = row0 + g.li;
        if (rows && (int)g.li < nv) {
            p.vattr[2u * row] = make_float4(x[4], x[5], x[6], x[7]);
            p.vattr[2u * row + 1u] = make_float4(x[8], x[9], x[10], x[11]);
        }
        float4 q0, q1, q2;
        fan_positions<12>(g, x, q0, q1, q2);
        bool live = false;
        TriVis tv;
        float rw[3];
        if (g.li >= 1u && (int)g.li < nv - 1) live = setup_triangle(d, v, q0, q1, q2, tv, rw);
        const bool binned = live && (!p.cmark || has_corner(tv));
        const uint32_t j_asm = (uint32_t)__popcll(__ballot(live) & g.gmask & g.below());
        const uint32_t j_bin = (uint32_t)__popcll(__ballot(binned) & g.gmask & g.below());
        if (binned && rows) write_triangle(p, off.x + j_bin, di, v, tv, rw, row0, row, row + 1u, t, off.y + j_asm);
        if (live && !binned) {
            int32_t bx0, by0, bx1, by1;
            rect_of(tv, bx0, by0, bx1, by1);
            for (int32_t by = by0; by <= by1; ++by)
                for (int32_t bx = bx0; bx <= bx1; ++bx) {
                    const uint32_t bit = d.cbase + (uint32_t)by * (uint32_t)v.bins_x + (uint32_t)bx;
                    atomicOr(p.cmark + (bit >> 5), 1u << (bit & 31u));
                    ++p8;
                }
        }
    }
    if (p.cmark) add_unbinned_pairs(p, p8);
}

// ------------------------------------------------------------ stage 2: binning

// Tile rect (tiles of 2^shift x 2^shift bins) of a TriVis 8x8-bin rect; returns the tile count
// (0 if empty).
__device__ __forceinline__ uint32_t tile_rect(uint32_t rect0, uint32_t rect1, uint32_t shift, int32_t &tx0, int32_t &ty0,
                                              int32_t &tw, uint32_t &pairs8) {
    int32_t bx0 = (int32_t)(rect0 & 0xFFFFu), by0 = (int32_t)(rect0 >> 16);
    int32_t bx1 = (int32_t)(int16_t)(rect1 & 0xFFFFu), by1 = (int32_t)(int16_t)(rect1 >> 16);
    if (bx1 < bx0 || by1 < by0) {
        pairs8 = 0u;
        tx0 = ty0 = 0;
        tw = 1;
        return 0u;
    }
    pairs8 = (uint32_t)(bx1 - bx0 + 1) * (uint32_t)(by1 - by0 + 1);
    tx0 = bx0 >> shift;
    ty0 = by0 >> shift;
    tw = (bx1 >> shift) - tx0 + 1;
    return (uint32_t)tw * (uint32_t)((by1 >> shift) - ty0 + 1);
}

constexpr uint32_t kRadixRounds = 16;
constexpr uint32_t kRadixItems = 256 * kRadixRounds;

// EMIT = false: per-triangle pair counts. EMIT = true: the pairs in triangle order at the
// scanned offsets. Triangles spanning more than kBigRect tiles are
// expanded by the whole block.
// The pair count's totals, by block 0 of the emitting pass: the (triangle, 8x8 bin) count for the
// stats from the counting blocks' partials, the capacity check and the radix passes' sizes.
__device__ __forceinline__ void pair_totals(const BatchParams &p, uint32_t total, bool over) {
    __shared__ unsigned long long s_w[4];
    unsigned long long acc = 0ull;
    for (uint32_t i = threadIdx.x; i < gridDim.x; i += 256u) acc += p.blocksum[i];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63u) == 0u) s_w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        p.stats->pairs8 += s_w[0] + s_w[1] + s_w[2] + s_w[3];
        atomicMax(&p.stats->max_pairs, (unsigned long long)total);
        uint32_t nb = (total + kRadixItems - 1u) / kRadixItems;
        if (over) {
            atomicOr(&p.ctr->overflow, OVF_PAIRS);
            atomicAdd(&p.stats->overflow_draws, 1ull);
            nb = 0u;
        }
        p.ctr->radix_blocks = nb;
        p.ctr->hist_n = nb * 256u;
    }
}

// Both passes run the same grid: the counting pass leaves one partial per block in blocksum.
template <bool EMIT>
__global__ __launch_bounds__(256) void kb_pairs(BatchParams p) {
    __shared__ uint32_t s_big[256];
    __shared__ uint32_t s_nbig;
    // the counters are read before block 0's stores (scalar loads)
    const uint32_t n = min(p.ctr->n_assembled, p.tri_cap);
    if (EMIT) {
        const uint32_t total = p.ctr->n_pairs;
        const bool tri_over = (p.ctr->overflow & OVF_TRIANGLES) != 0u;
        const bool over = total > p.pair_cap;  // past the arena: the tile stage walks every triangle
        if (blockIdx.x == 0) pair_totals(p, total, over);
        if (over || tri_over) return;
    }
    unsigned long long pairs8_acc = 0ull;
    for (uint32_t base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
        if (threadIdx.x == 0) s_nbig = 0u;
        __syncthreads();
        const uint32_t i = base + threadIdx.x;
        if (i < n) {
            const uint4 tb = p.tbin[i];
            const uint32_t off = EMIT ? p.pair_off[i] : 0u;  // in flight with the rect
            int32_t tx0, ty0, tw;
            uint32_t p8;
            const uint32_t nt = tile_rect(tb.x, tb.y, p.tile_shift, tx0, ty0, tw, p8);
            if (!EMIT) {
                p.pair_cnt[i] = nt;
                pairs8_acc += p8;
            }
            if (!EMIT) {
                // counting pass: nothing more per triangle
            } else if (nt <= (uint32_t)kBigRect) {
                for (uint32_t k = 0; k < nt; ++k) {
                    const uint32_t tile = tb.z + (uint32_t)(ty0 + (int32_t)(k / (uint32_t)tw)) * tb.w +
                                          (uint32_t)(tx0 + (int32_t)(k % (uint32_t)tw));
                    p.keys0[off + k] = tile;
                    p.vals0[off + k] = i;
                }
            } else {
                s_big[atomicAdd(&s_nbig, 1u)] = i;
            }
        }
        __syncthreads();
        const uint32_t nbig = s_nbig;
        for (uint32_t bi = 0; bi < nbig; ++bi) {
            const uint32_t ti = s_big[bi];
            const uint4 tb = p.tbin[ti];
            int32_t tx0, ty0, tw;
            uint32_t p8;
            const uint32_t nt = tile_rect(tb.x, tb.y, p.tile_shift, tx0, ty0, tw, p8);
            const uint32_t off = p.pair_off[ti];
            for (uint32_t k = threadIdx.x; k < nt; k += 256u) {
                const uint32_t tile = tb.z + (uint32_t)(ty0 + (int32_t)(k / (uint32_t)tw)) * tb.w +
                                      (uint32_t)(tx0 + (int32_t)(k % (uint32_t)tw));
                p.keys0[off + k] = tile;
                p.vals0[off + k] = ti;
            }
        }
        __syncthreads();
    }
    if (!EMIT) {
        // one partial per block (blocksum is free once setup has emitted): a single shared
        // counter hit by every wave serialises in L2
        __shared__ unsigned long long s_p8[4];
        unsigned long long s = wave_sum(pairs8_acc);
        if ((threadIdx.x & 63u) == 0u) s_p8[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) p.blocksum[blockIdx.x] = s_p8[0] + s_p8[1] + s_p8[2] + s_p8[3];
    }
}


// Stable LSD radix sort of (tile, triangle) pairs by tile, 8 bits per pass. A block owns
// kRadixItems consecutive pairs; within a round of 256, a pair's rank among equal digits
// comes from eight ballots (the lanes sharing its digit) and mbcnt; per-wave digit counts
// are combined across waves and rounds in LDS, so equal keys keep their input order.
// The grids are sized for the arena; blocks past the device's pair count return at once.

size_t radix_hist_elems(size_t n_max) { return 256 * ((n_max + kRadixItems - 1) / kRadixItems + 1); }
constexpr uint32_t kRadixGrid = 4096;  // blocks of a flush's radix passes (each walks chunks)

// Digit counts of one block's kRadixItems keys, stored digit-major (hist[digit * nb + block])
// so that one exclusive scan over hist gives every (digit, block) its output offset.
template <uint32_t NT>
__device__ __forceinline__ void radix_hist_block(const uint32_t *keys, uint32_t n, uint32_t nb, uint32_t b,
                                                 uint32_t shift, uint32_t *hist) {
    constexpr uint32_t R = kRadixItems / NT;
    // Four copies of the counters, lane & 3 choosing one: neighbouring keys (pairs of neighbouring
    // triangles) often share a digit, and one LDS atomic serialises its lanes per address. The
    // copies sit kHistPitch words apart, so one digit's four counters fall in four banks.
    constexpr uint32_t kHistPitch = 257u;
    __shared__ uint32_t h[4 * kHistPitch];
    for (uint32_t i = threadIdx.x; i < 4u * kHistPitch; i += NT) h[i] = 0u;
    __syncthreads();
    const uint32_t base = b * kRadixItems;
    // all of the thread's loads are issued before the first count, so their latencies overlap
    uint32_t k[R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        const uint32_t i = base + r * NT + threadIdx.x;
        k[r] = i < n ? keys[i] : 0xFFFFFFFFu;
    }
    uint32_t *mine = h + (threadIdx.x & 3u) * kHistPitch;
#pragma unroll
    for (uint32_t r = 0; r < R; ++r)
        if (base + r * NT + threadIdx.x < n) atomicAdd(&mine[(k[r] >> shift) & 255u], 1u);
    __syncthreads();
    if (threadIdx.x < 256u)
        hist[threadIdx.x * nb + b] = (h[threadIdx.x] + h[kHistPitch + threadIdx.x]) +
                                     (h[2u * kHistPitch + threadIdx.x] + h[3u * kHistPitch + threadIdx.x]);
}

// Stable scatter of one block's keys and values to the scanned offsets, by NT threads: rounds of
// NT items in index order; within a round each wave ranks its items by digit with ballots, and
// the waves' counts per digit are prefixed in wave order.
template <uint32_t NT>
__device__ __forceinline__ void radix_scatter_block(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                                                    uint32_t *vout, uint32_t n, uint32_t nb, uint32_t b,
                                                    uint32_t shift, const uint32_t *hist) {
    constexpr uint32_t NW = NT / 64u, R = kRadixItems / NT;
    static_assert(NT >= 256 && NT % 64 == 0 && kRadixItems % NT == 0, "radix block shape");
    __shared__ uint32_t wc[NW][256];
    __shared__ uint32_t loff[256];  // per digit: next free slot of the block's items in LDS
    __shared__ uint32_t lbeg[256];  // per digit: first slot of the block's items in LDS
    __shared__ uint32_t goff[256];  // per digit: first slot of the block's items in the output
    __shared__ uint32_t s_w[NW + 1];
    __shared__ uint32_t s_key[kRadixItems], s_val[kRadixItems];
    const uint32_t base = b * kRadixItems;
    const uint32_t tid = threadIdx.x, w = tid >> 6;
    // the block's count per digit is the step of the scanned digit-major histogram
    {
        uint32_t here = 0u, next = 0u;
        if (tid < 256u) {
            const uint32_t idx = tid * nb + b;
            here = hist[idx];
            next = (idx + 1u < 256u * nb) ? hist[idx + 1u] : n;
        }
        uint32_t total;
        const uint32_t ex = block_exclusive_scan<NT, uint32_t>(next - here, s_w, total);
        if (tid < 256u) {
            goff[tid] = here;
            loff[tid] = ex;
            lbeg[tid] = ex;
        }
    }
    // the thread's items, loaded before the first round so the loads' latencies overlap
    uint32_t keys[R], vals[R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        const uint32_t i = base + r * NT + tid;
        keys[r] = i < n ? kin[i] : 0u;
        vals[r] = i < n ? vin[i] : 0u;
    }
    // rank every item among the block's items of its digit (stable) and place it in LDS
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        for (uint32_t e = tid; e < NW * 256u; e += NT) (&wc[0][0])[e] = 0u;
        __syncthreads();
        const uint32_t i = base + r * NT + tid;
        const bool valid = i < n;
        const uint32_t key = keys[r];
        const uint32_t val = vals[r];
        const uint32_t dg = (key >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (dg >> bit) & 1u;
            const uint64_t bb = __ballot(on);
            m &= on ? bb : ~bb;
        }
        const uint32_t rank = mbcnt(m);
        if (valid && rank == 0u) wc[w][dg] = (uint32_t)__popcll(m);
        __syncthreads();
        if (tid < 256u) {
            uint32_t e = loff[tid];
#pragma unroll
            for (uint32_t q = 0; q < NW; ++q) {
                const uint32_t c = wc[q][tid];
                wc[q][tid] = e;
                e += c;
            }
            loff[tid] = e;
        }
        __syncthreads();
        if (valid) {
            const uint32_t lp = wc[w][dg] + rank;
            s_key[lp] = key;
            s_val[lp] = val;
        }
        __syncthreads();  // every wave has read wc before the next round clears it
    }
    // write the block's items digit run by digit run, so consecutive lanes write consecutive slots
    const uint32_t cnt = n > base ? min(kRadixItems, n - base) : 0u;
    for (uint32_t j = tid; j < cnt; j += NT) {
        const uint32_t key = s_key[j];
        const uint32_t dg = (key >> shift) & 255u;
        const uint32_t pos = goff[dg] + (j - lbeg[dg]);
        kout[pos] = key;
        vout[pos] = s_val[j];
    }
}

// Binning passes: the pair count and the block count are read from the device counters.
// The flush's passes: the grid is capped (kRadixGrid) and each block walks chunks b, b + grid,
// ... up to the pair count the device measured, so a flush with few pairs launches few idle
// blocks whatever the arena's capacity.
constexpr uint32_t kRadixThreads = 1024;  // the flush's radix blocks: 16 waves per 4096 pairs

__global__ __launch_bounds__(kRadixThreads) void kb_radix_hist(BatchParams p, const uint32_t *keys, uint32_t shift,
                                                               uint32_t *hist) {
    if (p.ctr->overflow & (OVF_PAIRS | OVF_TRIANGLES)) return;
    const uint32_t nb = p.ctr->radix_blocks;
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        radix_hist_block<kRadixThreads>(keys, p.ctr->n_pairs, nb, b, shift, hist);
        __syncthreads();  // the block's LDS counts are reused by its next chunk
    }
}

__global__ __launch_bounds__(kRadixThreads) void kb_radix_scatter(BatchParams p, const uint32_t *kin,
                                                                  const uint32_t *vin, uint32_t *kout,
                                                                  uint32_t *vout, uint32_t shift,
                                                                  const uint32_t *hist) {
    if (p.ctr->overflow & (OVF_PAIRS | OVF_TRIANGLES)) return;
    const uint32_t nb = p.ctr->radix_blocks;
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        radix_scatter_block<kRadixThreads>(kin, vin, kout, vout, p.ctr->n_pairs, nb, b, shift, hist);
        __syncthreads();  // the block's LDS staging is reused by its next chunk
    }
}

// Host-sized passes (n known when launching): one block per kRadixItems keys.
__global__ __launch_bounds__(256) void k_sort_hist(const uint32_t *keys, uint32_t n, uint32_t shift, uint32_t *hist) {
    radix_hist_block<256>(keys, n, gridDim.x, blockIdx.x, shift, hist);
}

__global__ __launch_bounds__(256) void k_sort_scatter(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                                                      uint32_t *vout, uint32_t n, uint32_t shift,
                                                      const uint32_t *hist) {
    radix_scatter_block<256>(kin, vin, kout, vout, n, gridDim.x, blockIdx.x, shift, hist);
}

hipError_t sort_pairs(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, uint32_t n, uint32_t bits,
                      const ScanScratch &sc, bool *in_alt, hipStream_t s) {
    *in_alt = false;
    if (n == 0) return hipSuccess;
    const uint32_t nb = (n + kRadixItems - 1) / kRadixItems;
    const uint32_t passes = std::max<uint32_t>(1u, (bits + 7u) / 8u);
    uint32_t *kin = k0, *vin = v0, *kout = k1, *vout = v1;
    for (uint32_t pass = 0; pass < passes; ++pass) {
        const uint32_t shift = 8u * pass;
        hipLaunchKernelGGL(k_sort_hist, dim3(nb), dim3(256), 0, s, (const uint32_t *)kin, n, shift, sc.hist);
        hipError_t e = launch_scan<uint32_t>(sc.hist, sc.hist, (size_t)nb * 256u, nullptr, sc, nullptr, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_sort_scatter, dim3(nb), dim3(256), 0, s, (const uint32_t *)kin, (const uint32_t *)vin,
                           kout, vout, n, shift, (const uint32_t *)sc.hist);
        std::swap(kin, kout);
        std::swap(vin, vout);
        *in_alt = !*in_alt;
    }
    return hipGetLastError();
}

// Tile ranges from the sorted keys: the first pair of a tile writes its start, the last its
// end (tile_cnt holds the end). Tiles without pairs keep start = end = 0.
__global__ __launch_bounds__(256) void kb_tile_ranges(BatchParams p, const uint32_t *keys) {
    if (p.ctr->overflow & (OVF_PAIRS | OVF_TRIANGLES)) return;
    const uint32_t n = p.ctr->n_pairs;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const uint32_t k = keys[i];
        if (i == 0u || keys[i - 1u] != k) p.tile_start[k] = i;
        if (i + 1u == n || keys[i + 1u] != k) p.tile_cnt[k] = i + 1u;
    }
}

// ------------------------------------------------------------ stage 3: tile

// DPP lane moves within a row of 16 lanes (GFX9 encodings): quad_perm [1,0,3,2] and
// [2,3,0,1] swap neighbours at distance 1 and 2; row_ror 4 and 8 rotate the row.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Reduce over the 64 lanes of a wave: minimum (GE) or maximum (LE) of the in-range depths,
// ignoring NaN. Register-only (DPP within rows, then the four row results); uniform result.
template <bool GE>
__device__ __forceinline__ float wave_depth_extreme(bool inside, float d) {
    float v = inside ? d : (GE ? __builtin_inff() : -__builtin_inff());
    if (GE) {
        v = fminf(v, dpp_move<0xB1>(v));
        v = fminf(v, dpp_move<0x4E>(v));
        v = fminf(v, dpp_move<0x124>(v));
        v = fminf(v, dpp_move<0x128>(v));
    } else {
        v = fmaxf(v, dpp_move<0xB1>(v));
        v = fmaxf(v, dpp_move<0x4E>(v));
        v = fmaxf(v, dpp_move<0x124>(v));
        v = fmaxf(v, dpp_move<0x128>(v));
    }
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return GE ? fminf(fminf(r0, r1), fminf(r2, r3)) : fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

// write_tile's tile minimum (main.c:589-603), folded from 1.0 (0.0 for the LESS_EQUAL mirror)
template <bool GE>
__device__ __forceinline__ float tile_hiz(bool inside, float d) {
    float m = wave_depth_extreme<GE>(inside, d);
    return GE ? min_ref(1.0f, m) : max_ref(0.0f, m);
}

__device__ __forceinline__ float uniform_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ uint32_t uniform_u(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// Per-wave state of one 8x8 bin while it walks its triangles in submission order.
struct BinState {
    float d;            // this lane's depth
    int32_t best;       // this lane's winning assembled triangle (-1: none yet)
    float hz;           // the bin's hierarchical-Z value after the previous draw (uniform)
    float wext;         // a bound on the bin's current depths for the conservative cull (uniform)
    int32_t cur_draw;   // draw of the last triangle seen (uniform)
    uint32_t n_draws;   // draws seen (uniform)
    int32_t depth_cull; // the current draw allows the conservative cull (uniform)
    uint64_t stale;     // lanes whose depth changed since wext was last recomputed (uniform)
    uint64_t inside_m;  // lanes whose pixel is inside the view (uniform)
    uint32_t bin;       // the bin's index in its view (uniform)
    uint32_t seen;      // lane j: the unbinned-triangle bit of the j-th draw met (bitmap on, j < 64)
};

// Up to N staged records of a tile's list in LDS, one array per field: the triangle-parallel
// screening reads a different record per lane, and with one array per field consecutive records
// sit in consecutive banks (an 80-byte record stride would put 64 lanes on a handful of banks).
// The rasteriser reads one record at a time as broadcasts. A record's order key is its position
// in the tile's list (the list is in submission order); the winner's position becomes its
// triangle index once, after the walk.
template <uint32_t N>
struct TileLdsN {
    static constexpr uint32_t kN = N;
    int32_t a16[3][N], b16[3][N], c[3][N];
    float ooa[N], z0[N], dz1[N], dz2[N], hiz[N], cull[N];
    uint32_t box[N];  // the snapped box's pixel corners relative to the tile origin (tile_box)
    uint16_t draw[N];
    uint8_t excl[N];  // edge slots the top-left rule excludes (0 under the reference's rule)
    static_assert(kMaxBatchDraws <= 65536, "draw indices of a flush fit 16 bits");
};
using TileLds = TileLdsN<256>;

// The pixel corners of TriVis::lat_x/lat_y relative to the tile origin (ox, oy), each clamped to
// [-1, 127] (the tile is at most 32 pixels wide, so the clamp keeps every comparison with its
// bins), as bytes x0, x1, y0, y1.
__device__ __forceinline__ uint32_t tile_box(const TriVis &t, int32_t ox, int32_t oy) {
    const auto rel = [](int32_t v, int32_t o) { return (uint32_t)(uint8_t)(int8_t)imin(imax(v - o, -1), 127); };
    return rel((int16_t)(t.lat_x & 0xFFFFu), ox) | (rel((int16_t)(t.lat_x >> 16), ox) << 8) |
           (rel((int16_t)(t.lat_y & 0xFFFFu), oy) << 16) | (rel((int16_t)(t.lat_y >> 16), oy) << 24);
}

template <class Lds>
__device__ __forceinline__ void stage_record(Lds &L, uint32_t k, const TriVis &t, int32_t ox, int32_t oy) {
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        L.a16[e][k] = t.a16[e];
        L.b16[e][k] = t.b16[e];
        L.c[e][k] = t.c[e];
    }
    L.ooa[k] = t.ooa;
    L.z0[k] = t.z0;
    L.dz1[k] = t.dz1;
    L.dz2[k] = t.dz2;
    L.hiz[k] = t.hiz_depth;
    L.cull[k] = t.cull_bound;
    L.box[k] = tile_box(t, ox, oy);
    L.draw[k] = (uint16_t)t.draw;
    L.excl[k] = (uint8_t)(t.draw >> 16);
}

// Entering a new draw: the bin's hierarchical-Z value becomes the tile minimum of its depths
// after the previous draw (write_tile, main.c:589-603).
// The (bin, draw) pair is counted here as active, so its bit in the bitmap of unbinned
// triangles is to be cleared (that bitmap's remaining bits are the pairs met only by unbinned
// ones): lane j keeps the j-th draw's bit and clear_seen clears them after the walk, off its path.
template <bool GE>
__device__ __forceinline__ void bin_draw_change(const BatchParams &p, int32_t draw, BinState &st, bool inside) {
    if (st.cur_draw >= 0) st.hz = uniform_f(tile_hiz<GE>(inside, st.d));
    st.cur_draw = draw;
    const DrawRec &D = p.draws[draw];
    st.depth_cull = (int32_t)uniform_u((uint32_t)D.depth_cull);
    if (p.cmark) {
        const uint32_t bit = D.cbase + st.bin, lane = threadIdx.x & 63u;
        if (st.n_draws < 64u) {
            if (lane == st.n_draws) st.seen = bit;
        } else if (lane == 0u) {
            atomicAnd(p.cmark + (bit >> 5), ~(1u << (bit & 31u)));
        }
    }
    st.n_draws++;
}

__device__ __forceinline__ void clear_seen(const BatchParams &p, const BinState &st) {
    if (p.cmark && (threadIdx.x & 63u) < min(st.n_draws, 64u))
        atomicAnd(p.cmark + (st.seen >> 5), ~(1u << (st.seen & 31u)));  // marks precede this kernel
}

// What the rasteriser reads of one staged record: the edge functions, the depth plane and the
// order key.
struct CoverRec {
    int32_t a16[3], b16[3], c[3];
    float ooa, z0, dz1, dz2;
    int32_t id;  // order key: the position in the tile's list
    uint32_t excl;
};

template <class Lds>
__device__ __forceinline__ CoverRec load_cover(const Lds &L, uint32_t k, uint32_t base) {
    CoverRec r;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        r.a16[e] = L.a16[e][k];
        r.b16[e] = L.b16[e][k];
        r.c[e] = L.c[e][k];
    }
    r.ooa = L.ooa[k];
    r.z0 = L.z0[k];
    r.dz1 = L.dz1[k];
    r.dz2 = L.dz2[k];
    r.id = (int32_t)(base + k);  // the list position
    r.excl = L.excl[k];
    return r;
}

// c + a * x (mod 2^32), a and x taken as signed 24-bit values (__mul24): one v_mad_i32_i24, where
// the compiler would emit a multiply per term and a three-input add (three instructions per edge
// function instead of two).
__device__ __forceinline__ uint32_t mad24(int32_t a, int32_t x, uint32_t c) {
    uint32_t r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(x), "v"(c));
    return r;
}

// Corner-sampled coverage of one triangle over the bin's 64 pixels and the early-Z key update
// (main.c:1015-1028, 1083-1168); the triangle has passed hierarchical-Z.
template <bool GE>
__device__ __forceinline__ void bin_cover(const CoverRec &r, BinState &st, bool inside, int32_t px, int32_t py) {
    const uint32_t e0 = mad24(r.a16[0], px, mad24(r.b16[0], py, (uint32_t)r.c[0]));
    const uint32_t e1 = mad24(r.a16[1], px, mad24(r.b16[1], py, (uint32_t)r.c[1]));
    const uint32_t e2 = mad24(r.a16[2], px, mad24(r.b16[2], py, (uint32_t)r.c[2]));
    const int32_t eo = (int32_t)(e0 | e1 | e2);
    bool cov = inside && eo > 0;  // main.c:1028
    if (r.excl)  // the top-left rule (extension): a corner on an excluded edge is not covered
        cov = cov && !((r.excl & 1u) && e0 == 0u) && !((r.excl & 2u) && e1 == 0u) && !((r.excl & 4u) && e2 == 0u);
    // the covered lanes straight from the compares' lane masks (38: signed greater, 33: not
    // equal); a ballot of cov would first turn the mask into a 0/1 register and compare again
    if ((__builtin_amdgcn_sicmp(eo, 0, 38) & st.inside_m) == 0ull) return;
    const float b1 = (float)((int32_t)e1 >> 8) * r.ooa;  // main.c:1089-1102
    const float b2 = (float)((int32_t)e2 >> 8) * r.ooa;
    float z = r.z0 + r.dz1 * b1;  // register 0 is affine (main.c:1118-1154)
    z = z + r.dz2 * b2;
    // the greater (depth, order) key wins (GE; the lesser depth under LESS_EQUAL), without branches
    const bool nearer = GE ? z > st.d : z < st.d;
    const bool tie = z == st.d;
    const bool upd = cov & (nearer | (tie & (r.id > st.best)));
    const float d_old = st.d;
    st.d = upd ? z : st.d;
    st.best = upd ? r.id : st.best;
    // the cull bound follows the depths, so only a changed depth makes it stale (14: unordered
    // or not equal, a compare's lane mask)
    st.stale |= __builtin_amdgcn_fcmpf(st.d, d_old, 14);
}

// The record index of the next survivor (the lowest set bit of S, which it clears); lane j of
// k holds entry j's record index.
__device__ __forceinline__ uint32_t next_survivor(uint32_t k, uint64_t &S) {
    const uint32_t jj = (uint32_t)__builtin_ctzll(S);
    S &= S - 1ull;
    return (uint32_t)__builtin_amdgcn_readlane((int)k, (int)jj);
}

// Triangle-parallel rasterisation of tiny survivors. Within one draw the walk's result does not
// depend on the order of its triangles: each pixel keeps the greatest (depth, order) key (GE; the
// least depth under LESS_EQUAL) among the triangles covering it, ties in depth going to the later
// triangle, and the draw's hierarchical-Z value and the screening do not change while it lasts.
// So a wave can take up to 64 survivors of one draw at once, a lane each, every lane evaluating
// only the pixel corners inside its triangle's snapped box within the bin (coverage lies inside
// that box when edge values cannot wrap: DrawRec::depth_cull), and merge them per pixel with
// 64-bit LDS maxima of keys: high word the depth in an order-preserving code, low word
// (lane + 1) << 1 (0 for the pixel's state before the batch, so a triangle wins a tie with it,
// and among the batch the later entry wins) plus a bit that marks a depth of -0.0, which the key
// orders as +0.0 (they compare equal) but the depth buffer keeps as written.
constexpr uint32_t kTpCorners = 16u;  // at most this many corners of the bin in the box
constexpr uint32_t kTpMin = 8u;      // at least this many such survivors in the batch

template <bool GE>
__device__ __forceinline__ uint32_t depth_key(float z) {
    uint32_t u = __float_as_uint(z);
    if ((u << 1) == 0u) u = 0u;                                   // -0.0 as +0.0
    const uint32_t o = u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);  // unsigned order = float order
    return GE ? o : ~o;
}

template <bool GE>
__device__ __forceinline__ float depth_of_key(uint32_t hi, bool neg_zero) {
    const uint32_t o = GE ? hi : ~hi;
    const uint32_t u = (o & 0x80000000u) ? (o ^ 0x80000000u) : ~o;
    return neg_zero ? -0.0f : __uint_as_float(u);
}

// The batch's tiny survivors (lanes of T; lane j holds entry j's record slot k and the corners
// [cx0, cx1] x [cy0, cy1] of its box, relative to the bin origin x0, y0) into the bin's state.
// s_key holds this wave's 64 keys; slot k is list position base + k.
template <bool GE, class Lds>
__device__ __forceinline__ void tp_rasterise(const Lds &L, uint64_t *s_key, uint64_t T, uint32_t k, uint32_t base,
                                             int32_t cx0, int32_t cx1, int32_t cy0, int32_t cy1, BinState &st,
                                             int32_t x0, int32_t y0, uint32_t lane) {
    // the pixel's state before the batch; a NaN depth takes no update, so nothing may beat it
    s_key[lane] = (st.d != st.d) ? ~0ull : ((uint64_t)depth_key<GE>(st.d) << 32);
    __builtin_amdgcn_wave_barrier();
    const bool mine = (T >> lane) & 1ull;
    if (mine) {
        const CoverRec r = load_cover(L, k, base);
        const uint32_t tag = (lane + 1u) << 1;
        int32_t x = cx0, y = cy0;
        for (;;) {
            const int32_t px = x0 + x, py = y0 + y;
            const uint32_t e0 = mad24(r.a16[0], px, mad24(r.b16[0], py, (uint32_t)r.c[0]));
            const uint32_t e1 = mad24(r.a16[1], px, mad24(r.b16[1], py, (uint32_t)r.c[1]));
            const uint32_t e2 = mad24(r.a16[2], px, mad24(r.b16[2], py, (uint32_t)r.c[2]));
            bool cov = (int32_t)(e0 | e1 | e2) > 0;  // main.c:1028 (the box lies inside the view)
            if (r.excl)
                cov = cov && !((r.excl & 1u) && e0 == 0u) && !((r.excl & 2u) && e1 == 0u) &&
                      !((r.excl & 4u) && e2 == 0u);
            if (cov) {
                // the depth exactly as bin_cover computes it
                const float b1 = (float)((int32_t)e1 >> 8) * r.ooa;
                const float b2 = (float)((int32_t)e2 >> 8) * r.ooa;
                float z = r.z0 + r.dz1 * b1;
                z = z + r.dz2 * b2;
                if (z == z) {  // a NaN depth never wins (its comparisons are false)
                    const uint64_t key = ((uint64_t)depth_key<GE>(z) << 32) |
                                         (uint64_t)(tag | (__float_as_uint(z) == 0x80000000u ? 1u : 0u));
                    atomicMax((unsigned long long *)&s_key[y * 8 + x], (unsigned long long)key);
                }
            }
            if (++x > cx1) {
                x = cx0;
                if (++y > cy1) break;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    const uint64_t kk = s_key[lane];
    const uint32_t lo = (uint32_t)kk;
    const bool won = lo != 0u && lo != 0xFFFFFFFFu;  // a triangle of the batch took this pixel
    const uint32_t j = won ? (lo >> 1) - 1u : 0u;
    const uint32_t kw = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(j << 2), (int)k);  // lane j's slot
    if (won) {
        st.d = depth_of_key<GE>((uint32_t)(kk >> 32), (lo & 1u) != 0u);
        st.best = (int32_t)(base + kw);
    }
    st.stale |= __ballot(won);
    __builtin_amdgcn_wave_barrier();  // s_key is rewritten by the next batch
}

// Largest value an edge function takes over the bin's pixels (x0..x0+7, y0..y0+7), exact when
// the draw guarantees no i32 wrap at any pixel (DrawRec::depth_cull).
__device__ __forceinline__ int32_t edge_max(int32_t a16, int32_t b16, int32_t c, int32_t x0, int32_t y0) {
    const int32_t x = a16 > 0 ? x0 + 7 : x0, y = b16 > 0 ? y0 + 7 : y0;
    return (int32_t)mad24(a16, x, mad24(b16, y, (uint32_t)c));
}

// Walk entries list[first + i * stride], i < n, of this wave's bin in order (staged slot k is
// list position base + k; the bin's origin is (x0, y0), (rx, ry) relative to the tile). Triangle-parallel
// screening first: lane j takes entry i0 + j and tests, at once for up to 64 entries of the
// current draw, the reference's hierarchical-Z reject (main.c:1005-1010), the conservative depth
// cull and an edge that is negative over the whole bin (then no pixel of the bin is covered).
// The survivors are then rasterised pixel-parallel, in order. Every test only removes triangles
// that could not have written a pixel, so the result equals the one-by-one walk.
template <bool GE, class Lds>
__device__ __forceinline__ void walk_bin(const BatchParams &p, const Lds &L, const uint8_t *list, uint32_t first,
                                         uint32_t stride, uint32_t n, uint32_t base, BinState &st, bool inside,
                                         int32_t px, int32_t py, int32_t x0, int32_t y0, int32_t rx, int32_t ry,
                                         uint32_t lane, uint64_t *s_key) {
    for (uint32_t i0 = 0; i0 < n;) {
        const uint32_t i = i0 + lane;
        const bool in = i < n;
        const uint32_t k = list[first + (in ? i : i0) * stride];
        const int32_t draw = (int32_t)L.draw[k];
        const int32_t d0 = (int32_t)uniform_u((uint32_t)draw);
        if (d0 != st.cur_draw) bin_draw_change<GE>(p, d0, st, inside);
        const uint64_t same = __ballot(in && draw == d0);
        const uint32_t run = (~same == 0ull) ? 64u : (uint32_t)__builtin_ctzll(~same);
        if (st.depth_cull && st.stale) {
            st.wext = uniform_f(wave_depth_extreme<GE>(inside, st.d));
            st.stale = 0ull;
        }
        bool live = lane < run;
        // a pixel corner of the bin inside the triangle's bounding box: the bin's corners
        // [bx0, bx1] x [by0, by1] in it (relative to the tile)
        const uint32_t bw = L.box[k];
        const int32_t bx0 = imax((int8_t)(bw & 0xFFu), rx), bx1 = imin((int8_t)((bw >> 8) & 0xFFu), rx + 7);
        const int32_t by0 = imax((int8_t)((bw >> 16) & 0xFFu), ry), by1 = imin((int8_t)(bw >> 24), ry + 7);
        live = live && bx0 <= bx1 && by0 <= by1;
        const float hzd = L.hiz[k];
        live = live && !(GE ? (hzd < st.hz) : (hzd > st.hz));
        if (st.depth_cull) {
            const float cb = L.cull[k];
            live = live && !(GE ? (cb < st.wext) : (cb > st.wext));
            live = live && edge_max(L.a16[0][k], L.b16[0][k], L.c[0][k], x0, y0) >= 0 &&
                   edge_max(L.a16[1][k], L.b16[1][k], L.c[1][k], x0, y0) >= 0 &&
                   edge_max(L.a16[2][k], L.b16[2][k], L.c[2][k], x0, y0) >= 0;
        }
        // survivors in order; lane jj already holds entry i0 + jj's record index, and the next
        // survivor's record is read from LDS while the current one is rasterised. Two records
        // alternate (the loop is unrolled by two), so the prefetch needs no register copies.
        uint64_t S = __ballot(live);
        if (s_key && st.depth_cull && (uint32_t)__popcll(S) >= kTpMin) {
            // tiny survivors triangle-parallel when there are enough of them; the rest below
            const uint64_t T = __ballot(live && (uint32_t)((bx1 - bx0 + 1) * (by1 - by0 + 1)) <= kTpCorners);
            if ((uint32_t)__popcll(T) >= kTpMin) {
                tp_rasterise<GE>(L, s_key, T, k, base, bx0 - rx, bx1 - rx, by0 - ry, by1 - ry, st, x0, y0, lane);
                S &= ~T;
            }
        }
        if (S) {
            CoverRec ra = load_cover(L, next_survivor(k, S), base), rb = ra;
            for (;;) {
                bool more = S != 0ull;
                if (more) rb = load_cover(L, next_survivor(k, S), base);
                bin_cover<GE>(ra, st, inside, px, py);
                if (!more) break;
                more = S != 0ull;
                if (more) ra = load_cover(L, next_survivor(k, S), base);
                bin_cover<GE>(rb, st, inside, px, py);
                if (!more) break;
            }
        }
        i0 += run;
    }
}

constexpr uint32_t kHeavyDraws = 1;

// Perspective-correct barycentrics of pixel corner (x, y) in an assembled triangle: the quantised
// affine barycentrics (E >> 8) * one_over_area (main.c:1083-1102), then main.c:1106-1115.
__device__ __forceinline__ void persp_bary(const TriVis &T, const TriShade &S, int32_t x, int32_t y, float &pbx,
                                           float &pby) {
    const uint32_t e1 = (uint32_t)T.c[1] + (uint32_t)__mul24(T.a16[1], x) + (uint32_t)__mul24(T.b16[1], y);
    const uint32_t e2 = (uint32_t)T.c[2] + (uint32_t)__mul24(T.a16[2], x) + (uint32_t)__mul24(T.b16[2], y);
    const float bx = (float)((int32_t)e1 >> 8) * T.ooa;
    const float by = (float)((int32_t)e2 >> 8) * T.ooa;
    float den = 1.0f - (bx + by);
    den = den * S.rw[0];
    den = den + bx * S.rw[1];
    den = den + by * S.rw[2];
    den = 1.0f / den;
    pbx = (bx * S.rw[1]) * den;
    pby = (by * S.rw[2]) * den;
}

// v0 + (v1 - v0) * pbx + (v2 - v0) * pby per component, main.c:1118-1160
__device__ __forceinline__ float4 interp_reg(float4 c0, float4 c1, float4 c2, float pbx, float pby) {
#define MLV_INTERP(q) (((c0.q) + ((c1.q) - (c0.q)) * pbx) + ((c2.q) - (c0.q)) * pby)
    return make_float4(MLV_INTERP(x), MLV_INTERP(y), MLV_INTERP(z), MLV_INTERP(w));
#undef MLV_INTERP
}

// The pixel shader stage for one pixel's winner (main.c:1106-1181): interpolated registers 1-2,
// the draw's pixel shader, the output encode, and the depth and visibility writes. The extension
// shaders also get the registers at the next corners right and below (x + 1, y + 1 on the same
// triangle's planes, like the neighbouring lanes of a 2x2 quad) for their mip footprints; the
// corners of an integer pixel grid need no edge or bounds test to be evaluated there.
__device__ __forceinline__ void shade_winner(const BatchParams &p, const ViewRec &V, uint32_t best, float depth,
                                             int32_t px, int32_t py, size_t pix) {
    const TriVis T = p.tvis[best];
    const TriShade S = p.tshade[best];
    const DrawRec &D = p.draws[T.draw & 0xFFFFu];
    if (D.ps == PS_NONE) {  // depth-only: no shader, no colour write (the vertex stage kept no attributes)
        V.depth[pix] = depth;
        if (p.tri_mark) atomicOr(p.tri_mark + (best >> 5), 1u << (best & 31u));  // footprint counting
        if (V.vis_tri) {
            V.vis_tri[pix] = S.num;
            V.vis_draw[pix] = D.draw_tag;
        }
        if (V.vis_prim) V.vis_prim[pix] = S.prim;
        return;
    }
    float pbx, pby;
    persp_bary(T, S, px, py, pbx, pby);
    const float4 a0 = p.vattr[2u * S.row[0]], b0 = p.vattr[2u * S.row[0] + 1u];
    const float4 a1 = p.vattr[2u * S.row[1]], b1 = p.vattr[2u * S.row[1] + 1u];
    const float4 a2 = p.vattr[2u * S.row[2]], b2 = p.vattr[2u * S.row[2] + 1u];
    const float4 r1 = interp_reg(a0, a1, a2, pbx, pby);
    const float4 r2 = interp_reg(b0, b1, b2, pbx, pby);
    float cr, cg, cb;
    if (D.ps >= PS_BASIC_LAMBERT) {
        float qx, qy, sx, sy;
        persp_bary(T, S, px + 1, py, qx, qy);
        persp_bary(T, S, px, py + 1, sx, sy);
        ext_pixel_shader(D, r1, r2, interp_reg(a0, a1, a2, qx, qy), interp_reg(b0, b1, b2, qx, qy),
                         interp_reg(a0, a1, a2, sx, sy), interp_reg(b0, b1, b2, sx, sy), cr, cg, cb);
    } else {
        pixel_shader(D.ps, D.ps_tex, r1, r2, cr, cg, cb);
    }
    V.color[pix] = encode_output(cr, cg, cb);
    V.depth[pix] = depth;
    if (p.tri_mark) atomicOr(p.tri_mark + (best >> 5), 1u << (best & 31u));  // footprint counting
    if (V.vis_tri) {
        V.vis_tri[pix] = S.num;  // the reference's number within the draw
        V.vis_draw[pix] = D.draw_tag;
    }
    if (V.vis_prim) V.vis_prim[pix] = S.prim;
}

__device__ __forceinline__ bool is_heavy(const BatchParams &p, const uint32_t *sorted_vals, uint32_t a, uint32_t len) {
    if (len <= p.heavy_min) return false;
    const uint32_t e = a + len;
    const uint32_t d0 = upper_index(p.n_draws, sorted_vals[a], [&](uint32_t i) { return p.asm_base[i]; });
    const uint32_t d1 = upper_index(p.n_draws, sorted_vals[e - 1u], [&](uint32_t i) { return p.asm_base[i]; });
    return d1 - d0 < kHeavyDraws;
}

// TB x TB bins per tile (TB = 2: 16x16 tiles, TB = 4: 32x32), one wave per bin.
template <bool GE, uint32_t N, uint32_t TB, class Mask>
__device__ __forceinline__ void tile_body(const BatchParams &p, const uint32_t *sorted_vals, uint32_t g, uint32_t vi,
                                          uint32_t lstart, uint32_t lcount, TileLdsN<N> &s_tri, Mask *s_mask,
                                          uint8_t (*s_list)[N], uint64_t (*s_key)[64]) {
    // the wave index is uniform per wave; saying so keeps everything derived from it (the bin,
    // its Hi-Z value, the walk's branches) in scalar registers
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = uniform_u(tid >> 6);
    const ViewRec &V = p.views[vi];
    const uint32_t tile = g - V.tile_base;
    const int32_t tx = (int32_t)(tile % (uint32_t)V.tiles_x), ty = (int32_t)(tile / (uint32_t)V.tiles_x);
    const int32_t sbx = (int32_t)TB * tx + (int32_t)(w % TB), sby = (int32_t)TB * ty + (int32_t)(w / TB);
    const int32_t px = sbx * kBin + (int32_t)(lane & 7u), py = sby * kBin + (int32_t)(lane >> 3);
    const bool sub_valid = sbx < V.bins_x && sby < V.bins_y;
    const bool inside = px < V.W && py < V.H;
    const size_t pix = (size_t)py * (size_t)V.W + (size_t)px;

    // the list range comes from kb_tile_order (every triangle of the flush when the pairs overflowed)
    const bool ovf = (p.ctr->overflow & OVF_PAIRS) != 0u;
    if (!ovf && is_heavy(p, sorted_vals, lstart, lcount)) return;  // kb_tile_heavy has drawn it
    if (lcount == 0u) return;

    BinState st;
    st.d = inside ? V.depth[pix] : 0.0f;
    st.best = -1;
    st.hz = uniform_f(sub_valid ? V.hiz[sby * V.bins_x + sbx] : 0.0f);
    st.cur_draw = -1;
    st.n_draws = 0u;
    st.depth_cull = 0;
    st.wext = uniform_f(wave_depth_extreme<GE>(inside, st.d));
    st.stale = 0ull;
    st.inside_m = __ballot(inside);
    st.bin = (uint32_t)(sby * V.bins_x + sbx);
    st.seen = 0u;

    for (uint32_t base = 0; base < lcount; base += N) {
        const uint32_t n = min(N, lcount - base);
        if (tid < n) {
            const uint32_t t = ovf ? base + tid : sorted_vals[lstart + base + tid];
            const TriVis tv = p.tvis[t];
            stage_record(s_tri, tid, tv, (int32_t)(TB * kBin) * tx, (int32_t)(TB * kBin) * ty);
            int32_t bx0, by0, bx1, by1;
            rect_of(tv, bx0, by0, bx1, by1);
            uint32_t m = 0u;
            if (!ovf || p.draws[tv.draw & 0xFFFFu].view == vi) {
#pragma unroll
                for (int k = 0; k < (int)(TB * TB); ++k) {
                    int32_t gx = (int32_t)TB * tx + (k % (int)TB), gy = (int32_t)TB * ty + (k / (int)TB);
                    if (gx >= bx0 && gx <= bx1 && gy >= by0 && gy <= by1) m |= 1u << k;
                }
            }
            s_mask[tid] = (Mask)m;
        }
        __syncthreads();
        uint32_t cnt = 0u;
#pragma unroll
        for (int c = 0; c < (int)(N / 64u); ++c) {
            const uint32_t i = (uint32_t)c * 64u + lane;
            const bool b = i < n && ((s_mask[i] >> w) & 1u);
            const uint64_t bal = __ballot(b);
            if (b) s_list[w][cnt + mbcnt(bal)] = (uint8_t)i;
            cnt += (uint32_t)__popcll(bal);
        }
        if (sub_valid && cnt)
            walk_bin<GE>(p, s_tri, s_list[w], 0u, 1u, cnt, base, st, inside, px, py, sbx * kBin, sby * kBin,
                         (int32_t)((w % TB) * kBin), (int32_t)((w / TB) * kBin), lane, s_key[w]);
        __syncthreads();
    }

    // the winner's list position to its assembled triangle
    if (st.best >= 0) st.best = ovf ? st.best : (int32_t)sorted_vals[lstart + (uint32_t)st.best];
    if (st.best >= 0) shade_winner(p, V, (uint32_t)st.best, st.d, px, py, pix);
    if (p.foot) {  // footprint counting: pixels written
        const uint64_t m = __ballot(st.best >= 0);
        if (lane == 0u && m) atomicAdd(&p.foot->shaded_px, (unsigned long long)__popcll(m));
    }

    if (sub_valid && st.cur_draw >= 0) {
        const float m = tile_hiz<GE>(inside, st.d);
        if (lane == 0u) {
            V.hiz[sby * V.bins_x + sbx] = m;
            // (bin, draw) activity for the stats, spread over 64 counters
            atomicAdd(&p.stats->active8_part[blockIdx.x & 63u], (unsigned long long)st.n_draws);
        }
    }
    clear_seen(p, st);  // last: no later load of the wave waits for these atomics
}

// A heavy tile: a sorted list longer than heavy_min whose triangles all belong to one draw
// (kHeavyDraws), e.g. where many small triangles of one mesh converge or a stress draw covers
// the screen. Lists that span draws stay with kb_tile, whose single wave per bin keeps the
// Hi-Z refresh between draws in order.
__global__ __launch_bounds__(256) void kb_heavy_tiles(BatchParams p, const uint32_t *sorted_vals) {
    if (p.ctr->overflow & (OVF_PAIRS | OVF_TRIANGLES | OVF_CLIPVERTS)) return;
    for (uint32_t g = blockIdx.x * 256u + threadIdx.x; g < p.total_tiles; g += gridDim.x * 256u)
        if (is_heavy(p, sorted_vals, p.tile_start[g], p.tile_cnt[g] - p.tile_start[g]))
            p.heavy[atomicAdd(&p.ctr->n_heavy, 1u)] = g;
}

// A heavy tile per block iteration, its four 8x8 bins walked by four waves each. Heavy tiles
// hold one draw, so within a bin the result is the maximum of (depth, submission order) keys
// over the triangles that pass, which does not depend on the order they are visited in, and the
// hierarchical-Z value stays the one from before the flush (it changes only between draws,
// main.c:589-603). Wave q of a bin takes the bin's entries q, q+4, ...; at the end the four
// waves' keys are merged in LDS and the first wave of the bin writes it.
constexpr uint32_t kHeavyWaves = 16;
constexpr uint32_t kWavesPerBin = kHeavyWaves / 4;

template <bool GE>
__device__ __forceinline__ void heavy_tile(const BatchParams &p, const uint32_t *sorted_vals, uint32_t g,
                                           uint32_t vi, TileLds &s_tri, uint8_t *s_mask, uint8_t (*s_list)[256],
                                           uint32_t *s_cnt, float (*s_md)[64], int32_t (*s_mb)[64],
                                           uint64_t (*s_key)[64]) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = uniform_u(tid >> 6);
    const uint32_t bin = w / kWavesPerBin, q = w % kWavesPerBin;
    const ViewRec &V = p.views[vi];
    const uint32_t tile = g - V.tile_base;
    const int32_t tx = (int32_t)(tile % (uint32_t)V.tiles_x), ty = (int32_t)(tile / (uint32_t)V.tiles_x);
    const int32_t sbx = 2 * tx + (int32_t)(bin & 1u), sby = 2 * ty + (int32_t)(bin >> 1);
    const int32_t px = sbx * kBin + (int32_t)(lane & 7u), py = sby * kBin + (int32_t)(lane >> 3);
    const bool sub_valid = sbx < V.bins_x && sby < V.bins_y;
    const bool inside = px < V.W && py < V.H;
    const size_t pix = (size_t)py * (size_t)V.W + (size_t)px;
    const uint32_t lstart = p.tile_start[g], lcount = p.tile_cnt[g] - lstart;

    BinState st;
    st.d = inside ? V.depth[pix] : 0.0f;
    st.best = -1;
    st.hz = uniform_f(sub_valid ? V.hiz[sby * V.bins_x + sbx] : 0.0f);
    st.cur_draw = -1;
    st.n_draws = 0u;
    st.depth_cull = 0;
    st.wext = uniform_f(wave_depth_extreme<GE>(inside, st.d));
    st.stale = 0ull;
    st.inside_m = __ballot(inside);
    st.bin = (uint32_t)(sby * V.bins_x + sbx);
    st.seen = 0u;
    bool seen = false;

    for (uint32_t base = 0; base < lcount; base += 256u) {
        const uint32_t n = min(256u, lcount - base);
        if (tid < n) {
            const uint32_t t = sorted_vals[lstart + base + tid];
            const TriVis tv = p.tvis[t];
            stage_record(s_tri, tid, tv, 16 * tx, 16 * ty);
            int32_t bx0, by0, bx1, by1;
            rect_of(tv, bx0, by0, bx1, by1);
            uint32_t m = 0u;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                int32_t gx = 2 * tx + (k & 1), gy = 2 * ty + (k >> 1);
                if (gx >= bx0 && gx <= bx1 && gy >= by0 && gy <= by1) m |= 1u << k;
            }
            s_mask[tid] = m;
        }
        __syncthreads();
        if (w < 4u) {  // wave b builds bin b's list, in order
            uint32_t cnt = 0u;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint32_t i = (uint32_t)c * 64u + lane;
                const bool b = i < n && ((s_mask[i] >> w) & 1u);
                const uint64_t bal = __ballot(b);
                if (b) s_list[w][cnt + mbcnt(bal)] = (uint8_t)i;
                cnt += (uint32_t)__popcll(bal);
            }
            if (lane == 0u) s_cnt[w] = cnt;
        }
        __syncthreads();
        const uint32_t cnt = s_cnt[bin];
        if (cnt) seen = true;
        if (sub_valid && q < cnt)
            walk_bin<GE>(p, s_tri, s_list[bin], q, kWavesPerBin, (cnt - q + kWavesPerBin - 1u) / kWavesPerBin, base,
                         st, inside, px, py, sbx * kBin, sby * kBin, (int32_t)((bin & 1u) * kBin),
                         (int32_t)((bin >> 1) * kBin), lane, s_key[w]);
        __syncthreads();
    }
    // merge the bin's four keys: the greater (depth, order) key wins, as in bin_cover
    s_md[w][lane] = st.d;
    s_mb[w][lane] = st.best;
    __syncthreads();
    clear_seen(p, st);
    if (q != 0u || !sub_valid) return;
    for (uint32_t k = 1; k < kWavesPerBin; ++k) {
        const float dq = s_md[w + k][lane];
        const int32_t bq = s_mb[w + k][lane];
        const bool better = GE ? (dq > st.d || (dq == st.d && bq > st.best)) : (dq < st.d || (dq == st.d && bq > st.best));
        if (better) {
            st.d = dq;
            st.best = bq;
        }
    }
    if (st.best >= 0) st.best = (int32_t)sorted_vals[lstart + (uint32_t)st.best];  // list position to triangle
    if (st.best >= 0) shade_winner(p, V, (uint32_t)st.best, st.d, px, py, pix);
    if (p.foot) {  // footprint counting: pixels written
        const uint64_t m = __ballot(st.best >= 0);
        if (lane == 0u && m) atomicAdd(&p.foot->shaded_px, (unsigned long long)__popcll(m));
    }
    if (seen) {
        const float m = tile_hiz<GE>(inside, st.d);
        if (lane == 0u) {
            V.hiz[sby * V.bins_x + sbx] = m;
            atomicAdd(&p.stats->active8, 1ull);  // one draw met this bin
        }
    }
}

__global__ __launch_bounds__(1024) void kb_tile_heavy(BatchParams p, const uint32_t *sorted_vals) {
    __shared__ TileLds s_tri;
    __shared__ uint8_t s_mask[256];
    __shared__ __attribute__((aligned(8))) uint8_t s_list[4][256];
    __shared__ uint32_t s_cnt[4];
    __shared__ float s_md[kHeavyWaves][64];
    __shared__ int32_t s_mb[kHeavyWaves][64];
    __shared__ uint64_t s_key[kHeavyWaves][64];
    if (p.ctr->overflow & (OVF_PAIRS | OVF_TRIANGLES | OVF_CLIPVERTS)) return;
    const uint32_t items = p.ctr->n_heavy;
    for (uint32_t h = blockIdx.x; h < items; h += gridDim.x) {
        const uint32_t g = p.heavy[h];
        const uint32_t vi = upper_index(p.n_views, g, [&](uint32_t i) { return p.views[i].tile_base; });
        if (p.views[vi].depth_func == DEPTH_GE)
            heavy_tile<true>(p, sorted_vals, g, vi, s_tri, s_mask, s_list, s_cnt, s_md, s_mb, s_key);
        else
            heavy_tile<false>(p, sorted_vals, g, vi, s_tri, s_mask, s_list, s_cnt, s_md, s_mb, s_key);
        __syncthreads();  // LDS is reused by the next tile
    }
}

// The order kb_tile takes tiles in: longest lists first (largest-first scheduling). Per-tile work
// varies tenfold where meshes crowd; started last, the longest tiles would decide when the launch
// ends. Tiles fall into eight classes by list length against the frame's average (>= 32x, 16x,
// 8x, 4x, 2x, 1x, shorter, empty); one pass counts the classes, the next places every tile after
// the longer classes, each wave claiming its slots with one atomic per class.
__device__ __forceinline__ uint32_t tile_class(const BatchParams &p, uint32_t g, bool ovf, uint32_t avg) {
    if (ovf) return 0u;  // the slow path walks every triangle in every tile: one class
    const uint32_t len = p.tile_cnt[g] - p.tile_start[g];
    if (len == 0u) return 7u;
    uint32_t c = 6u;
    for (uint32_t k = 0; k < 6u; ++k)
        if (len >= (avg << (5u - k))) {
            c = k;
            break;
        }
    return c;
}

template <bool PLACE>
__global__ __launch_bounds__(256) void kb_tile_order(BatchParams p) {
    // a block's four waves combine their class counts in LDS, so each block makes one global
    // atomic per class (one per wave and class contended for the eight counters)
    __shared__ uint32_t s_wc[4][8];  // per wave and class: its count, then its offset in the block
    __shared__ uint32_t s_base[8];   // per class: the block's first slot
    const uint32_t n = p.total_tiles;
    const bool ovf = (p.ctr->overflow & OVF_PAIRS) != 0u;
    const uint32_t avg = max(1u, n ? p.ctr->n_pairs / n : 1u);
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t start[8];
    const uint32_t n_asm = min(p.ctr->n_assembled, p.tri_cap);
    if (PLACE) {
        uint32_t acc = 0u;
        for (int c = 0; c < 8; ++c) {
            start[c] = acc;
            acc += p.ctr->class_cnt[c];
        }
    }
    for (uint32_t base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
        const uint32_t g = base + threadIdx.x;
        const uint32_t cls = g < n ? tile_class(p, g, ovf, avg) : 8u;
#pragma unroll
        for (uint32_t c = 0; c < 8u; ++c) {
            const uint64_t b = __ballot(cls == c);
            if (lane == 0u) s_wc[w][c] = (uint32_t)__popcll(b);
        }
        __syncthreads();
        if (threadIdx.x < 8u) {
            const uint32_t c = threadIdx.x;
            uint32_t e = 0u;
#pragma unroll
            for (uint32_t q = 0; q < 4u; ++q) {
                const uint32_t t = s_wc[q][c];
                s_wc[q][c] = e;
                e += t;
            }
            if (e) {
                if (!PLACE) atomicAdd(&p.ctr->class_cnt[c], e);
                else s_base[c] = start[c] + atomicAdd(&p.ctr->class_cur[c], e);
            }
        }
        __syncthreads();
        if (PLACE) {
#pragma unroll
            for (uint32_t c = 0; c < 8u; ++c) {
                const uint64_t b = __ballot(cls == c);
                if (cls == c) {
                    // the tile's list range and view go with it: the tile kernel's first load
                    const uint32_t a = ovf ? 0u : p.tile_start[g], len = ovf ? n_asm : p.tile_cnt[g] - a;
                    const uint32_t vi = upper_index(p.n_views, g, [&](uint32_t i) { return p.views[i].tile_base; });
                    p.tile_order[s_base[c] + s_wc[w][c] + mbcnt(b)] = make_uint4(g, a, len, vi);
                }
            }
        }
        __syncthreads();  // s_wc and s_base are reused by the next round
    }
}

// One block per tile of every view (TB x TB bins of 8x8: 16x16 or 32x32 pixels), one wave per
// bin. The block stages up to N records of the tile's sorted list in LDS; each wave compacts the
// ones touching its bin and walks them in submission order with its pixels' depths in registers;
// afterwards each pixel's winner is shaded once and colour, depth, visibility and the bin's Hi-Z
// are written.
// 16x16 tiles run at 8 blocks per CU: 64 VGPRs and 20 KiB of LDS a block.
template <uint32_t N, uint32_t TB>
__global__ __launch_bounds__(64 * TB * TB) __attribute__((amdgpu_waves_per_eu(TB == 2u ? 8u : 1u)))
void kb_tile(BatchParams p, const uint32_t *sorted_vals) {
    using Mask = typename std::conditional<TB * TB <= 8u, uint8_t, uint16_t>::type;  // a bit per bin
    __shared__ TileLdsN<N> s_tri;
    __shared__ Mask s_mask[N];
    __shared__ __attribute__((aligned(8))) uint8_t s_list[TB * TB][N];
    __shared__ uint64_t s_key[TB * TB][64];  // per wave: the triangle-parallel keys

    if (p.ctr->overflow & (OVF_TRIANGLES | OVF_CLIPVERTS)) return;  // flush failed; reported on the host
    // Tiles in kb_tile_order's order; the dispatcher deals consecutive blocks to the 8 XCDs in
    // turn, so the long tiles at the front spread over all XCDs.
    const uint4 e = p.tile_order[blockIdx.x];  // (global tile, list start, list length, view)
    if (e.z == 0u) return;                      // an empty tile: nothing drawn, nothing counted
    const uint32_t g = e.x, vi = e.w;
    if (p.views[vi].depth_func == DEPTH_GE)
        tile_body<true, N, TB>(p, sorted_vals, g, vi, e.y, e.z, s_tri, s_mask, s_list, s_key);
    else
        tile_body<false, N, TB>(p, sorted_vals, g, vi, e.y, e.z, s_tri, s_mask, s_list, s_key);
}

// Sum of the per-tile activity counts into the stats (one atomic per block).

// ------------------------------------------------------------ launchers

// Level l + 1 of a texture's chain from level l: each texel the average of its 2x2 window (2x1
// or 1x2 once a side of level l is 1; an odd side drops its last row or column). RGBA8 averages
// each byte in integers, rounding halves up; the float formats sum the window in row order in
// f32 and divide by its area (RGBA16F then rounds to half, nearest-even).
__global__ __launch_bounds__(256) void k_mip_level(TexDesc t, uint32_t l) {
    const TexDesc src = tex_level(t, l), dst = tex_level(t, l + 1u);
    const uint32_t W = dst.width, H = dst.height, sw = src.width;
    const uint32_t kw = src.width > 1u ? 2u : 1u, kh = src.height > 1u ? 2u : 1u, n = kw * kh;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < W * H; i += gridDim.x * 256u) {
        const uint32_t x = i % W, y = i / W;
        if (t.format == FMT_RGBA8) {
            const uint32_t *q = (const uint32_t *)src.data;
            uint32_t acc[4] = {0u, 0u, 0u, 0u};
            for (uint32_t dy = 0; dy < kh; ++dy)
                for (uint32_t dx = 0; dx < kw; ++dx) {
                    const uint32_t c = q[(size_t)(kh * y + dy) * sw + kw * x + dx];
#pragma unroll
                    for (int k = 0; k < 4; ++k) acc[k] += (c >> (8 * k)) & 255u;
                }
            uint32_t o = 0u;
#pragma unroll
            for (int k = 0; k < 4; ++k) o |= ((acc[k] + n / 2u) / n) << (8 * k);
            ((uint32_t *)dst.data)[i] = o;
        } else {
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
            for (uint32_t dy = 0; dy < kh; ++dy)
                for (uint32_t dx = 0; dx < kw; ++dx) {
                    const size_t j = (size_t)(kh * y + dy) * sw + kw * x + dx;
                    const float4 c = t.format == FMT_RGBA32F ? ((const float4 *)src.data)[j]
                                                             : decode_rgba16f(((const ushort4 *)src.data)[j]);
                    acc.x += c.x;
                    acc.y += c.y;
                    acc.z += c.z;
                    acc.w += c.w;
                }
            const float a = (float)n;
            const float4 o = make_float4(acc.x / a, acc.y / a, acc.z / a, acc.w / a);
            if (t.format == FMT_RGBA32F) {
                ((float4 *)dst.data)[i] = o;
            } else {
                ((ushort4 *)dst.data)[i] = make_ushort4(__half_as_ushort(__float2half_rn(o.x)),
                                                        __half_as_ushort(__float2half_rn(o.y)),
                                                        __half_as_ushort(__float2half_rn(o.z)),
                                                        __half_as_ushort(__float2half_rn(o.w)));
            }
        }
    }
}

// Footprint counting: adds the set bits of words[0..n) to *out_bits and mul times them to *out_bytes.
__global__ __launch_bounds__(256) void k_popcount(const uint32_t *words, size_t n, uint32_t mul,
                                                  unsigned long long *out_bits, unsigned long long *out_bytes) {
    unsigned long long c = 0;
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (size_t)gridDim.x * 256u)
        c += (unsigned long long)__popc(words[i]);
    c = wave_sum(c);
    if ((threadIdx.x & 63u) == 0u && c) {
        if (out_bits) atomicAdd(out_bits, c);
        if (out_bytes) atomicAdd(out_bytes, c * mul);
    }
}

// Adds the set bits of words[0, n) to *out and clears them.
__global__ __launch_bounds__(256) void k_count_clear(uint32_t *words, size_t n, unsigned long long *out) {
    unsigned long long c = 0;
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (size_t)gridDim.x * 256u) {
        const uint32_t w = words[i];
        if (w) {
            c += (unsigned long long)__popc(w);
            words[i] = 0u;
        }
    }
    c = wave_sum(c);
    if ((threadIdx.x & 63u) == 0u && c) atomicAdd(out, c);
}

hipError_t launch_count_clear(uint32_t *words, size_t n, unsigned long long *out, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_count_clear, dim3((unsigned)std::min<size_t>((n + 255) / 256, 1024)), dim3(256), 0, s, words,
                       n, out);
    return hipGetLastError();
}

hipError_t launch_popcount(const uint32_t *words, size_t n, uint32_t mul, unsigned long long *out_bits,
                           unsigned long long *out_bytes, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_popcount, dim3((unsigned)std::min<size_t>((n + 255) / 256, 2048)), dim3(256), 0, s, words,
                       n, mul, out_bits, out_bytes);
    return hipGetLastError();
}

hipError_t launch_generate_mips(const TexDesc &t, hipStream_t s) {
    uint32_t w = t.width, h = t.height;
    for (uint32_t l = 0; l + 1u < t.mip_levels; ++l) {
        w = w > 1u ? w >> 1 : 1u;
        h = h > 1u ? h >> 1 : 1u;
        const uint32_t n = w * h;
        hipLaunchKernelGGL(k_mip_level, dim3(std::min<uint32_t>((n + 255u) / 256u, 4096u)), dim3(256), 0, s, t, l);
    }
    return hipGetLastError();
}

hipError_t launch_present_rgb(const uint32_t *fb, uint8_t *rgb, size_t n_pixels, hipStream_t s) {
    const size_t quads = (n_pixels + 3u) / 4u;
    const unsigned grid = (unsigned)std::min<size_t>((quads + 255) / 256, 4096);
    if (grid) hipLaunchKernelGGL(k_present_rgb, dim3(grid), dim3(256), 0, s, fb, (uint32_t *)rgb, n_pixels);
    return hipGetLastError();
}

hipError_t launch_srgb_to_linear_rgba8(uint32_t *texels, size_t n, hipStream_t s) {
    const unsigned grid = (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 2048));
    hipLaunchKernelGGL(k_srgb_to_linear_rgba8, dim3(grid), dim3(256), 0, s, texels, n);
    return hipGetLastError();
}

hipError_t launch_fill(const FillJobs &f, hipStream_t s) {
    size_t n = 0;
    for (uint32_t k = 0; k < f.count; ++k) n = std::max(n, f.n[k]);
    if (!n) return hipSuccess;
    const unsigned grid = (unsigned)std::min<size_t>((n / 4u + 255u) / 256u + 1u, 2048u);
    hipLaunchKernelGGL(k_fill32, dim3(grid), dim3(256), 0, s, f);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, size_t n, hipStream_t s) {
    FillJobs f{};
    f.add(dst, value, n);
    return launch_fill(f, s);
}

hipError_t launch_fill_f32(float *dst, float value, size_t n, hipStream_t s) {
    return launch_fill_u32((uint32_t *)dst, __builtin_bit_cast(uint32_t, value), n, s);
}

// Static LDS bytes of the tile kernel variant (tile_stage records, 2^shift bins a side) and of the
// setup kernels (the larger of count and emit), for the occupancy caps.
uint32_t tile_kernel_lds(uint32_t tile_stage, uint32_t shift) {
    hipFuncAttributes a{};
    hipError_t e;
    if (shift == 2u)
        e = tile_stage == 128u ? hipFuncGetAttributes(&a, (const void *)kb_tile<128, 4>)
                               : hipFuncGetAttributes(&a, (const void *)kb_tile<256, 4>);
    else
        e = tile_stage == 128u ? hipFuncGetAttributes(&a, (const void *)kb_tile<128, 2>)
                               : hipFuncGetAttributes(&a, (const void *)kb_tile<256, 2>);
    return e == hipSuccess ? (uint32_t)a.sharedSizeBytes : 0u;
}

uint32_t setup_kernel_lds() {
    hipFuncAttributes a{}, b{};
    if (hipFuncGetAttributes(&a, (const void *)kb_setup_count<false>) != hipSuccess) return 0u;
    hipFuncAttributes c{};
    if (hipFuncGetAttributes(&b, (const void *)kb_setup_emit) != hipSuccess) return 0u;
    if (hipFuncGetAttributes(&c, (const void *)kb_setup_emit_waves) != hipSuccess) return 0u;
    return (uint32_t)std::max({a.sharedSizeBytes, b.sharedSizeBytes, c.sharedSizeBytes});
}

hipError_t launch_stage(int stage, const BatchParams &p, const ScanScratch &sc, uint32_t radix_passes,
                        hipStream_t s, hipEvent_t *tile_ev) {
    hipError_t e;
    switch (stage) {
    case 0:
    {
        const dim3 grid(std::max(p.n_vblocks, std::min((p.total_tiles + 255u) / 256u + 1u, 1024u)));
        if (p.vblk_draw) {
            hipLaunchKernelGGL(kb_block_draws, dim3(p.n_draws), dim3(256), 0, s, p);
            hipLaunchKernelGGL(kb_vertex<true>, grid, dim3(256), 0, s, p);
        } else {
            hipLaunchKernelGGL(kb_vertex<false>, grid, dim3(256), 0, s, p);
        }
        break;
    }
    case 1:
        if (!p.n_tblocks) break;
        if (p.tblk_draw)
            hipLaunchKernelGGL(kb_setup_count<true>, dim3((p.n_tblocks + kCountPer<true> - 1u) / kCountPer<true>),
                               dim3(256), p.setup_lds_pad, s, p);
        else
            hipLaunchKernelGGL(kb_setup_count<false>, dim3((p.n_tblocks + kCountPer<false> - 1u) / kCountPer<false>),
                               dim3(256), p.setup_lds_pad, s, p);
        hipLaunchKernelGGL(kb_clip_count, dim3(1024), dim3(256), 0, s, p);
        if ((e = launch_scan<unsigned long long>(p.blocksum, p.blockoff, p.n_tblocks, nullptr, sc,
                                                 &p.ctr->scan_total, s)) != hipSuccess)
            return e;
        if (p.tblk_draw)
            hipLaunchKernelGGL(kb_setup_emit_waves, dim3((p.n_tblocks + kEmitWaves - 1u) / kEmitWaves), dim3(256),
                               p.setup_lds_pad, s, p);
        else
            hipLaunchKernelGGL(kb_setup_emit, dim3(p.n_tblocks), dim3(256), p.setup_lds_pad, s, p);
        hipLaunchKernelGGL(kb_clip_emit, dim3(1024), dim3(256), 0, s, p);
        break;
    case 2: {
        if (!p.n_tblocks) break;
        const uint32_t grid = std::max(1u, std::min<uint32_t>(p.n_tblocks, p.pair_blocks));
        hipLaunchKernelGGL(kb_pairs<false>, dim3(grid), dim3(256), 0, s, p);
        if ((e = launch_scan<uint32_t>(p.pair_cnt, p.pair_off, p.tri_cap, &p.ctr->n_assembled, sc,
                                       &p.ctr->n_pairs, s)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(kb_pairs<true>, dim3(grid), dim3(256), 0, s, p);
        const uint32_t nb = (p.pair_cap + kRadixItems - 1) / kRadixItems;  // chunks the arena can hold
        const uint32_t rgrid = std::max(1u, std::min(nb, kRadixGrid));
        uint32_t *kin = p.keys0, *vin = p.vals0, *kout = p.keys1, *vout = p.vals1;
        for (uint32_t pass = 0; pass < radix_passes; ++pass) {
            const uint32_t shift = 8u * pass;
            hipLaunchKernelGGL(kb_radix_hist, dim3(rgrid), dim3(kRadixThreads), 0, s, p, (const uint32_t *)kin, shift,
                               sc.hist);
            if ((e = launch_scan<uint32_t>(sc.hist, sc.hist, (size_t)nb * 256u, &p.ctr->hist_n, sc, nullptr,
                                           s)) != hipSuccess)
                return e;
            hipLaunchKernelGGL(kb_radix_scatter, dim3(rgrid), dim3(kRadixThreads), 0, s, p, (const uint32_t *)kin,
                               (const uint32_t *)vin, kout, vout, shift, (const uint32_t *)sc.hist);
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
        hipLaunchKernelGGL(kb_tile_ranges, dim3(std::min<uint32_t>(nb * 8u, 2048u)), dim3(256), 0, s, p,
                           (const uint32_t *)kin);
        break;
    }
    case 3: {
        if (!p.n_tblocks || !p.total_tiles) break;
        const uint32_t *sorted = (radix_passes & 1u) ? p.vals1 : p.vals0;
        if (p.heavy_min != 0xFFFFFFFFu && p.tile_shift == 1u) {  // the 16-wave path is on (16x16 tiles)
            hipLaunchKernelGGL(kb_heavy_tiles, dim3(std::min<uint32_t>((p.total_tiles + 255) / 256, 256u)), dim3(256),
                               0, s, p, sorted);
            hipLaunchKernelGGL(kb_tile_heavy, dim3(std::max(1u, std::min<uint32_t>(p.total_tiles, p.heavy_blocks))),
                               dim3(1024), 0, s, p, sorted);
        }
        const uint32_t ogrid = std::min<uint32_t>((p.total_tiles + 255) / 256, 1024u);
        hipLaunchKernelGGL(kb_tile_order<false>, dim3(ogrid), dim3(256), 0, s, p);
        hipLaunchKernelGGL(kb_tile_order<true>, dim3(ogrid), dim3(256), 0, s, p);
        if (tile_ev && (e = hipEventRecord(tile_ev[0], s)) != hipSuccess) return e;
        const uint32_t pad = p.tile_lds_pad;  // dynamic LDS that caps the blocks per CU (0: none)
        if (p.tile_shift == 2u) {
            if (p.tile_stage == 128u)
                hipLaunchKernelGGL((kb_tile<128, 4>), dim3(p.total_tiles), dim3(1024), pad, s, p, sorted);
            else
                hipLaunchKernelGGL((kb_tile<256, 4>), dim3(p.total_tiles), dim3(1024), pad, s, p, sorted);
        } else if (p.tile_stage == 128u) {
            hipLaunchKernelGGL((kb_tile<128, 2>), dim3(p.total_tiles), dim3(256), pad, s, p, sorted);
        } else {
            hipLaunchKernelGGL((kb_tile<256, 2>), dim3(p.total_tiles), dim3(256), pad, s, p, sorted);
        }
        if (tile_ev && (e = hipEventRecord(tile_ev[1], s)) != hipSuccess) return e;
        // (bin, draw) pairs met only by unbinned triangles: active bins of the stats
        if (p.cmark && (e = launch_count_clear(p.cmark, p.cmark_words, &p.stats->active8, s)) != hipSuccess) return e;
        break;
    }
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mlv
