// mlv_kernels.hip -- the draw pipeline as gfx950 kernels.
//
// Per draw (launch_draw):
//   k_reset            zero the per-draw counters and the tile histogram
//   k_vertex<VS>       one lane per vertex: VS registers 0..2 -> vpos / vattr   (main.c:698-734)
//   k_setup_count      one lane per input triangle: cull/classify, output count (main.c:751-908)
//   k_clip_count       one lane per triangle that needs the clipper           (main.c:609-660)
//   k_scan_blocks      exclusive scan of per-block counts (order-preserving compaction)
//   k_setup_emit       one lane per input triangle: write TriVis/TriShade at its ordered slot
//   k_clip_emit        clipped triangles: write fan triangles and clipper vertices
//   k_bin_count        (triangle, 16x16 tile) histogram; large triangles binned block-wide
//   k_scan_tiles       tile offsets
//   k_bin_scatter      tile lists
//   k_tile<PS>         one block per 16x16 tile, one wave per 8x8 bin: hierarchical-Z,
//                      corner-sampled coverage, early-Z visibility in registers, then the
//                      pixel shader for the surviving fragment and the output merger
//                      (main.c:983-1189).
//
// Visibility is resolved per pixel as the maximum of (depth, submission order) under the
// reference's GREATER_EQUAL test (main.c:1166): sequential "z >= d then write" over a bin
// list in submission order leaves the fragment of maximal depth, and among equal depths the
// latest one. The tile lists therefore need no ordering for the result to equal the
// reference's ordered loop, which lets binning use atomics.
#include "mlv_device.h"

namespace mlv {

// ------------------------------------------------------------ helpers

__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// Block-wide exclusive scan for NT threads (NT a multiple of 64, <= 1024).
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *s_wave, uint32_t &total) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63u) s_wave[w] = inc;
    __syncthreads();
    if (threadIdx.x < 64u) {
        uint32_t s = threadIdx.x < (uint32_t)(NT / 64) ? s_wave[threadIdx.x] : 0u;
        uint32_t si = wave_inclusive_scan(s);
        if (threadIdx.x < (uint32_t)(NT / 64)) s_wave[threadIdx.x] = si - s;
        if (threadIdx.x == 63u) s_wave[NT / 64] = si;
    }
    __syncthreads();
    uint32_t r = s_wave[w] + inc - v;
    total = s_wave[NT / 64];
    __syncthreads();
    return r;
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of
// tiles so neighbouring tiles (which share triangles and texels) meet in the same L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
    const uint32_t X = 8u;
    if (n < X) return b;
    uint32_t per = n / X, rem = n % X, x = b % X, loc = b / X;
    return x * per + (x < rem ? x : rem) + loc;
}

__device__ __forceinline__ bool load_tri_indices(const DrawParams &p, uint32_t t, uint32_t vi[3]) {
    const uint32_t *ib = p.ib + p.start_index + 3u * t;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        long long v = (long long)ib[k] + (long long)p.base_vertex;
        if (v < 0 || v >= (long long)p.vertex_count) return false;
        vi[k] = (uint32_t)v;
    }
    return true;
}

// Polygon buffer for the clipper: the three vertices' registers, then the six planes.
__device__ __noinline__ int clip_triangle(const DrawParams &p, uint32_t t, float (*poly)[12], float (*tmp)[12],
                                          uint32_t vi[3]) {
    if (!load_tri_indices(p, t, vi)) return 0;
    for (int k = 0; k < 3; ++k) {
        float4 a = p.vpos[vi[k]], b = p.vattr[2u * vi[k]], c = p.vattr[2u * vi[k] + 1u];
        poly[k][0] = a.x; poly[k][1] = a.y; poly[k][2] = a.z; poly[k][3] = a.w;
        poly[k][4] = b.x; poly[k][5] = b.y; poly[k][6] = b.z; poly[k][7] = b.w;
        poly[k][8] = c.x; poly[k][9] = c.y; poly[k][10] = c.z; poly[k][11] = c.w;
    }
    int n = 3;
    for (int pl = 0; pl < 6; pl += 2) {
        n = clip_by_plane(poly, n, p.planes[pl], tmp);
        if (n < 0) return 0;
        n = clip_by_plane(tmp, n, p.planes[pl + 1], poly);
        if (n < 0) return 0;
    }
    return n;
}

__device__ __forceinline__ float4 poly_pos(const float *v) { return make_float4(v[0], v[1], v[2], v[3]); }

// ------------------------------------------------------------ kernels

__global__ void k_fill_u32(uint32_t *d, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = v;
}

__global__ void k_fill_f32(float *d, float v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = v;
}

__global__ void k_reset(DrawParams p) {
    const uint32_t ntiles = (uint32_t)(p.tiles_x * p.tiles_y);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        DrawCounters z = {};
        *p.ctr = z;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ntiles; i += gridDim.x * blockDim.x)
        p.tile_cnt[i] = 0u;
}

template <int VS>
__global__ __launch_bounds__(256) void k_vertex(DrawParams p) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= p.vertex_count) return;
    float in[8];
    const uint8_t *src = p.vb + (size_t)i * p.stride;
    if ((((uintptr_t)p.vb) & 15u) == 0u && (p.stride & 15u) == 0u) {
        float4 a = ((const float4 *)src)[0], b = ((const float4 *)src)[1];
        in[0] = a.x; in[1] = a.y; in[2] = a.z; in[3] = a.w; in[4] = b.x; in[5] = b.y; in[6] = b.z; in[7] = b.w;
    } else {
        for (int k = 0; k < 8; ++k) in[k] = ((const float *)src)[k];
    }
    float4 r0, r1, r2;
    vertex_shader<VS>(p, in, r0, r1, r2);
    p.vpos[i] = r0;
    p.vattr[2u * i] = r1;
    p.vattr[2u * i + 1u] = r2;
}

__global__ __launch_bounds__(256) void k_setup_count(DrawParams p) {
    __shared__ uint32_t s_w[4];
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    uint32_t packed = 0u;
    if (t < p.tri_count) {
        uint32_t vi[3];
        if (!load_tri_indices(p, t, vi)) {
            atomicAdd(&p.ctr->oob, 1u);
            atomicAdd(&p.stats->oob, 1ull);
        } else {
            float4 a = p.vpos[vi[0]], b = p.vpos[vi[1]], c = p.vpos[vi[2]];
            int cls = classify(a, b, c);
            if (cls == 1) {
                TriVis tv;
                float rw[3];
                packed = setup_triangle(p, a, b, c, tv, rw) ? 1u : 0u;
            } else if (cls == 2) {
                uint32_t slot = atomicAdd(&p.ctr->n_clip_list, 1u);
                p.clip_list[slot] = t;
            }
        }
        p.cnt[t] = packed;
    }
    uint32_t s = wave_sum(packed);
    if ((threadIdx.x & 63u) == 0u) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) p.blocksum[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ __launch_bounds__(64) void k_clip_count(DrawParams p) {
    const uint32_t n = p.ctr->n_clip_list;
    float poly[kMaxClip][12], tmp[kMaxClip][12];
    for (uint32_t s = blockIdx.x * 64u + threadIdx.x; s < n; s += gridDim.x * 64u) {
        const uint32_t t = p.clip_list[s];
        uint32_t vi[3];
        int nv = clip_triangle(p, t, poly, tmp, vi);
        uint32_t ntri = 0u;
        for (int ci = 1; ci < nv - 1; ++ci) {
            TriVis tv;
            float rw[3];
            if (setup_triangle(p, poly_pos(poly[0]), poly_pos(poly[ci]), poly_pos(poly[ci + 1]), tv, rw)) ntri++;
        }
        uint32_t packed = ntri ? (ntri | ((uint32_t)nv << 16)) : 0u;
        p.cnt[t] = packed;
        if (packed) atomicAdd(&p.blocksum[t / 256u], packed);
    }
}

// Exclusive scan of the packed block counts, one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan_blocks(DrawParams p) {
    __shared__ uint32_t s_w[17];
    const uint32_t n = p.nblocks;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t st = 0u, sv = 0u;
    for (uint32_t i = 0; i < per; ++i)
        if (b0 + i < n) {
            uint32_t v = p.blocksum[b0 + i];
            st += v & 0xFFFFu;
            sv += v >> 16;
        }
    uint32_t tt, tv_;
    uint32_t et = block_exclusive_scan<1024>(st, s_w, tt);
    uint32_t ev = block_exclusive_scan<1024>(sv, s_w, tv_);
    for (uint32_t i = 0; i < per; ++i)
        if (b0 + i < n) {
            uint32_t v = p.blocksum[b0 + i];
            p.blockoff_tri[b0 + i] = et;
            p.blockoff_vtx[b0 + i] = ev;
            et += v & 0xFFFFu;
            ev += v >> 16;
        }
    if (threadIdx.x == 0) {
        p.ctr->n_assembled = tt;
        p.ctr->n_clipverts = tv_;
        uint32_t ovf = 0u;
        if (tt > p.tri_cap) ovf |= OVF_TRIANGLES;
        if ((unsigned long long)p.vertex_count + tv_ > (unsigned long long)p.vattr_rows) ovf |= OVF_CLIPVERTS;
        if (ovf) {
            atomicOr(&p.ctr->overflow, ovf);
            atomicAdd(&p.stats->capacity_errors, 1ull);
        }
        atomicAdd(&p.stats->assembled, (unsigned long long)tt);
        atomicMax(&p.stats->max_clip_list, (unsigned long long)p.ctr->n_clip_list);
    }
}

__device__ __forceinline__ void write_triangle(const DrawParams &p, uint32_t idx, TriVis &tv, const float rw[3],
                                               uint32_t r0, uint32_t r1, uint32_t r2, uint32_t prim, uint32_t sub) {
    if (idx >= p.tri_cap) return;
    tv.id = idx;
    tv.pad[0] = 0u;
    tv.pad[1] = 0u;
    p.tvis[idx] = tv;
    TriShade sh;
    sh.rw[0] = rw[0]; sh.rw[1] = rw[1]; sh.rw[2] = rw[2];
    sh.row[0] = r0; sh.row[1] = r1; sh.row[2] = r2;
    sh.prim = prim;
    sh.sub = sub;
    p.tshade[idx] = sh;
}

__global__ __launch_bounds__(256) void k_setup_emit(DrawParams p) {
    __shared__ uint32_t s_w[5];
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    uint32_t packed = t < p.tri_count ? p.cnt[t] : 0u;
    uint32_t total;
    uint32_t ex = block_exclusive_scan<256>(packed, s_w, total);
    if (packed == 0u) return;
    const uint32_t off_tri = p.blockoff_tri[blockIdx.x] + (ex & 0xFFFFu);
    const uint32_t off_vtx = p.blockoff_vtx[blockIdx.x] + (ex >> 16);
    if (packed == 1u) {
        uint32_t vi[3];
        load_tri_indices(p, t, vi);
        float4 a = p.vpos[vi[0]], b = p.vpos[vi[1]], c = p.vpos[vi[2]];
        TriVis tv;
        float rw[3];
        setup_triangle(p, a, b, c, tv, rw);
        write_triangle(p, off_tri, tv, rw, vi[0], vi[1], vi[2], t, 0u);
    } else {
        p.clip_off[t] = make_uint2(off_tri, off_vtx);
    }
}

__global__ __launch_bounds__(64) void k_clip_emit(DrawParams p) {
    const uint32_t n = p.ctr->n_clip_list;
    float poly[kMaxClip][12], tmp[kMaxClip][12];
    for (uint32_t s = blockIdx.x * 64u + threadIdx.x; s < n; s += gridDim.x * 64u) {
        const uint32_t t = p.clip_list[s];
        if (p.cnt[t] == 0u) continue;
        const uint2 off = p.clip_off[t];
        uint32_t vi[3];
        int nv = clip_triangle(p, t, poly, tmp, vi);
        const uint32_t row0 = p.vertex_count + off.y;
        for (int k = 0; k < nv; ++k) {
            uint32_t row = row0 + (uint32_t)k;
            if (row >= p.vattr_rows) break;
            p.vattr[2u * row] = make_float4(poly[k][4], poly[k][5], poly[k][6], poly[k][7]);
            p.vattr[2u * row + 1u] = make_float4(poly[k][8], poly[k][9], poly[k][10], poly[k][11]);
        }
        uint32_t j = 0u;
        for (int ci = 1; ci < nv - 1; ++ci) {
            TriVis tv;
            float rw[3];
            if (setup_triangle(p, poly_pos(poly[0]), poly_pos(poly[ci]), poly_pos(poly[ci + 1]), tv, rw)) {
                write_triangle(p, off.x + j, tv, rw, row0, row0 + (uint32_t)ci, row0 + (uint32_t)ci + 1u, t,
                               (uint32_t)(ci - 1));
                j++;
            }
        }
    }
}

// Tile rect (16x16 tiles) from a TriVis 8x8-bin rect. Returns the tile count (0 if empty).
__device__ __forceinline__ uint32_t tile_rect(const DrawParams &p, uint32_t rect0, uint32_t rect1, int32_t &tx0,
                                              int32_t &ty0, int32_t &tw, uint32_t &pairs8) {
    int32_t bx0 = (int32_t)(rect0 & 0xFFFFu), by0 = (int32_t)(rect0 >> 16);
    int32_t bx1 = (int32_t)(int16_t)(rect1 & 0xFFFFu), by1 = (int32_t)(int16_t)(rect1 >> 16);
    if (bx1 < bx0 || by1 < by0) {
        pairs8 = 0u;
        return 0u;
    }
    pairs8 = (uint32_t)(bx1 - bx0 + 1) * (uint32_t)(by1 - by0 + 1);
    tx0 = bx0 >> 1;
    ty0 = by0 >> 1;
    int32_t tx1 = bx1 >> 1, ty1 = by1 >> 1;
    tw = tx1 - tx0 + 1;
    return (uint32_t)tw * (uint32_t)(ty1 - ty0 + 1);
}

template <bool SCATTER>
__global__ __launch_bounds__(256) void k_bin(DrawParams p) {
    __shared__ uint32_t s_big[256];
    __shared__ uint32_t s_nbig;
    if (SCATTER && (p.ctr->overflow & (OVF_PAIRS | OVF_TRIANGLES))) return;
    const uint32_t n = min(p.ctr->n_assembled, p.tri_cap);
    unsigned long long pairs8_acc = 0ull;
    for (uint32_t base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
        if (threadIdx.x == 0) s_nbig = 0u;
        __syncthreads();
        const uint32_t i = base + threadIdx.x;
        if (i < n) {
            const uint2 r = make_uint2(p.tvis[i].rect0, p.tvis[i].rect1);
            int32_t tx0, ty0, tw;
            uint32_t p8;
            uint32_t nt = tile_rect(p, r.x, r.y, tx0, ty0, tw, p8);
            pairs8_acc += p8;
            if (nt <= (uint32_t)kBigRect) {
                for (uint32_t k = 0; k < nt; ++k) {
                    uint32_t tile = (uint32_t)(ty0 + (int32_t)(k / (uint32_t)tw)) * (uint32_t)p.tiles_x +
                                    (uint32_t)(tx0 + (int32_t)(k % (uint32_t)tw));
                    if (SCATTER) {
                        uint32_t slot = atomicAdd(&p.tile_fill[tile], 1u);
                        p.tile_list[p.tile_start[tile] + slot] = i;
                    } else {
                        atomicAdd(&p.tile_cnt[tile], 1u);
                    }
                }
            } else {
                uint32_t k = atomicAdd(&s_nbig, 1u);
                s_big[k] = i;
            }
        }
        __syncthreads();
        const uint32_t nbig = s_nbig;
        for (uint32_t b = 0; b < nbig; ++b) {
            const uint32_t ti = s_big[b];
            const uint2 r = make_uint2(p.tvis[ti].rect0, p.tvis[ti].rect1);
            int32_t tx0, ty0, tw;
            uint32_t p8;
            uint32_t nt = tile_rect(p, r.x, r.y, tx0, ty0, tw, p8);
            for (uint32_t k = threadIdx.x; k < nt; k += 256u) {
                uint32_t tile = (uint32_t)(ty0 + (int32_t)(k / (uint32_t)tw)) * (uint32_t)p.tiles_x +
                                (uint32_t)(tx0 + (int32_t)(k % (uint32_t)tw));
                if (SCATTER) {
                    uint32_t slot = atomicAdd(&p.tile_fill[tile], 1u);
                    p.tile_list[p.tile_start[tile] + slot] = ti;
                } else {
                    atomicAdd(&p.tile_cnt[tile], 1u);
                }
            }
        }
        __syncthreads();
    }
    if (!SCATTER) {
        unsigned long long s = wave_sum64(pairs8_acc);
        if ((threadIdx.x & 63u) == 0u && s) {
            atomicAdd(&p.ctr->pairs8, s);
            atomicAdd(&p.stats->pairs8, s);
        }
    }
}

__global__ __launch_bounds__(1024) void k_scan_tiles(DrawParams p) {
    __shared__ uint32_t s_w[17];
    const uint32_t n = (uint32_t)(p.tiles_x * p.tiles_y);
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t s = 0u;
    for (uint32_t i = 0; i < per; ++i)
        if (b0 + i < n) s += p.tile_cnt[b0 + i];
    uint32_t total;
    uint32_t e = block_exclusive_scan<1024>(s, s_w, total);
    for (uint32_t i = 0; i < per; ++i)
        if (b0 + i < n) {
            p.tile_start[b0 + i] = e;
            p.tile_fill[b0 + i] = 0u;
            e += p.tile_cnt[b0 + i];
        }
    if (threadIdx.x == 0) {
        p.ctr->n_pairs = total;
        atomicMax(&p.stats->max_pairs, (unsigned long long)total);
        if (total > p.pair_cap) {
            atomicOr(&p.ctr->overflow, OVF_PAIRS);
            atomicAdd(&p.stats->overflow_draws, 1ull);
        }
    }
}

// Reduce over the 64 lanes of a wave: minimum (GE) or maximum (LE) of the in-range depths,
// ignoring NaN. Equal on every lane.
__device__ __forceinline__ float wave_depth_extreme(bool inside, float d, bool ge) {
    float v = inside ? d : (ge ? __builtin_inff() : -__builtin_inff());
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        float u = __shfl_xor(v, o, 64);
        v = ge ? fminf(v, u) : fmaxf(v, u);
    }
    return v;
}

template <int PS>
__global__ __launch_bounds__(256) void k_tile(DrawParams p) {
    __shared__ TriVis s_tri[256];
    __shared__ uint32_t s_mask[256];
    __shared__ uint16_t s_list[4][256];

    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t ntiles = (uint32_t)(p.tiles_x * p.tiles_y);
    const uint32_t tile = xcd_remap(blockIdx.x, ntiles);
    const int32_t tx = (int32_t)(tile % (uint32_t)p.tiles_x), ty = (int32_t)(tile / (uint32_t)p.tiles_x);
    const int32_t sbx = 2 * tx + (int32_t)(w & 1u), sby = 2 * ty + (int32_t)(w >> 1);
    const int32_t px = sbx * kBin + (int32_t)(lane & 7u), py = sby * kBin + (int32_t)(lane >> 3);
    const bool sub_valid = sbx < p.bins_x && sby < p.bins_y;
    const bool inside = px < p.W && py < p.H;
    const size_t pix = (size_t)py * (size_t)p.W + (size_t)px;
    const bool ge = p.depth_func == DEPTH_GE;

    if (p.ctr->overflow & (OVF_TRIANGLES | OVF_CLIPVERTS)) return;  // draw failed; reported on the host

    float d = inside ? p.depth[pix] : 0.0f;
    int32_t best = -1;
    const float hz = sub_valid ? p.hiz[sby * p.bins_x + sbx] : 0.0f;
    bool touched = false;
    float wext = wave_depth_extreme(inside, d, ge);

    const uint32_t n_asm = min(p.ctr->n_assembled, p.tri_cap);
    const bool ovf = (p.ctr->overflow & OVF_PAIRS) != 0u;
    uint32_t lstart = 0u, lcount = n_asm;
    if (!ovf) {
        lstart = p.tile_start[tile];
        lcount = p.tile_cnt[tile];
    }

    for (uint32_t base = 0; base < lcount; base += 256u) {
        const uint32_t n = min(256u, lcount - base);
        if (tid < n) {
            const uint32_t t = ovf ? base + tid : p.tile_list[lstart + base + tid];
            const TriVis tv = p.tvis[t];
            s_tri[tid] = tv;
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
        uint32_t cnt = 0u;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint32_t i = (uint32_t)c * 64u + lane;
            const bool b = i < n && ((s_mask[i] >> w) & 1u);
            const uint64_t bal = __ballot(b);
            if (b) s_list[w][cnt + mbcnt(bal)] = (uint16_t)i;
            cnt += (uint32_t)__popcll(bal);
        }
        if (sub_valid && cnt) {
            touched = true;
            for (uint32_t k = 0; k < cnt; ++k) {
                const uint32_t i = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_list[w][k]);
                const TriVis T = s_tri[i];
                // reference hierarchical-Z (main.c:1005-1010), per (triangle, 8x8 bin)
                bool rej = ge ? (T.hiz_depth < hz) : (T.hiz_depth > hz);
                // conservative in-draw cull: no fragment of T can pass the current depths
                if (p.depth_cull) rej = rej || (ge ? (T.cull_bound < wext) : (T.cull_bound > wext));
                if (__builtin_amdgcn_readfirstlane((int)rej)) continue;
                const uint32_t e0 = (uint32_t)T.c[0] + (uint32_t)__mul24(T.a16[0], px) + (uint32_t)__mul24(T.b16[0], py);
                const uint32_t e1 = (uint32_t)T.c[1] + (uint32_t)__mul24(T.a16[1], px) + (uint32_t)__mul24(T.b16[1], py);
                const uint32_t e2 = (uint32_t)T.c[2] + (uint32_t)__mul24(T.a16[2], px) + (uint32_t)__mul24(T.b16[2], py);
                const bool cov = inside && (int32_t)(e0 | e1 | e2) > 0;  // main.c:1028
                const float b1 = (float)((int32_t)e1 >> 8) * T.ooa;     // main.c:1089-1102
                const float b2 = (float)((int32_t)e2 >> 8) * T.ooa;
                float z = T.z0 + T.dz1 * b1;                             // main.c:1118-1154, register 0 (affine)
                z = z + T.dz2 * b2;
                const int32_t id = (int32_t)T.id;
                const bool pass = ge ? (z > d || (z == d && id > best)) : (z < d || (z == d && id > best));
                const bool upd = cov && pass;
                if (upd) {
                    d = z;
                    best = id;
                }
                if (p.depth_cull && __ballot(upd)) wext = wave_depth_extreme(inside, d, ge);
            }
        }
        __syncthreads();
    }

    if (best >= 0) {
        const TriVis T = p.tvis[best];
        const TriShade S = p.tshade[best];
        const uint32_t e1 = (uint32_t)T.c[1] + (uint32_t)__mul24(T.a16[1], px) + (uint32_t)__mul24(T.b16[1], py);
        const uint32_t e2 = (uint32_t)T.c[2] + (uint32_t)__mul24(T.a16[2], px) + (uint32_t)__mul24(T.b16[2], py);
        const float bx = (float)((int32_t)e1 >> 8) * T.ooa;
        const float by = (float)((int32_t)e2 >> 8) * T.ooa;
        // perspective correction, main.c:1106-1115
        float den = 1.0f - (bx + by);
        den = den * S.rw[0];
        den = den + bx * S.rw[1];
        den = den + by * S.rw[2];
        den = 1.0f / den;
        const float pbx = (bx * S.rw[1]) * den;
        const float pby = (by * S.rw[2]) * den;
        const float4 a0 = p.vattr[2u * S.row[0]], b0 = p.vattr[2u * S.row[0] + 1u];
        const float4 a1 = p.vattr[2u * S.row[1]], b1 = p.vattr[2u * S.row[1] + 1u];
        const float4 a2 = p.vattr[2u * S.row[2]], b2 = p.vattr[2u * S.row[2] + 1u];
#define MLV_INTERP(c0, c1, c2) (((c0) + ((c1) - (c0)) * pbx) + ((c2) - (c0)) * pby)
        const float4 r1 = make_float4(MLV_INTERP(a0.x, a1.x, a2.x), MLV_INTERP(a0.y, a1.y, a2.y),
                                      MLV_INTERP(a0.z, a1.z, a2.z), MLV_INTERP(a0.w, a1.w, a2.w));
        const float4 r2 = make_float4(MLV_INTERP(b0.x, b1.x, b2.x), MLV_INTERP(b0.y, b1.y, b2.y),
                                      MLV_INTERP(b0.z, b1.z, b2.z), MLV_INTERP(b0.w, b1.w, b2.w));
#undef MLV_INTERP
        float cr, cg, cb;
        pixel_shader<PS>(p, r1, r2, cr, cg, cb);
        p.color[pix] = encode_output(cr, cg, cb);
        p.depth[pix] = d;
        if (p.vis_tri) {
            p.vis_tri[pix] = (uint32_t)best;
            p.vis_draw[pix] = p.draw_tag;
        }
    }

    // write_tile's tile minimum (main.c:589-603) for every 8x8 bin the draw touched
    if (sub_valid && touched) {
        float m = wave_depth_extreme(inside, d, ge);
        m = ge ? min_ref(1.0f, m) : max_ref(0.0f, m);
        if (lane == 0u) {
            p.hiz[sby * p.bins_x + sbx] = m;
            atomicAdd(&p.stats->active8, 1ull);
        }
    }
}

// ------------------------------------------------------------ launchers

hipError_t launch_fill_u32(uint32_t *dst, uint32_t value, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    unsigned grid = (unsigned)((n + 255) / 256);
    if (grid > 65535u * 8u) grid = 65535u * 8u;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, s, dst, value, n);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float *dst, float value, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    unsigned grid = (unsigned)((n + 255) / 256);
    if (grid > 65535u * 8u) grid = 65535u * 8u;
    hipLaunchKernelGGL(k_fill_f32, dim3(grid), dim3(256), 0, s, dst, value, n);
    return hipGetLastError();
}

hipError_t launch_draw(const DrawParams &p, int vs, int ps, hipStream_t s) {
    const uint32_t ntiles = (uint32_t)(p.tiles_x * p.tiles_y);
    hipLaunchKernelGGL(k_reset, dim3((ntiles + 255) / 256), dim3(256), 0, s, p);
    if (p.vertex_count) {
        dim3 g((p.vertex_count + 255u) / 256u);
        switch (vs) {
        case VS_BASIC: hipLaunchKernelGGL(k_vertex<VS_BASIC>, g, dim3(256), 0, s, p); break;
        case VS_VERTEX_LIGHTING: hipLaunchKernelGGL(k_vertex<VS_VERTEX_LIGHTING>, g, dim3(256), 0, s, p); break;
        case VS_PASSTHROUGH: hipLaunchKernelGGL(k_vertex<VS_PASSTHROUGH>, g, dim3(256), 0, s, p); break;
        case VS_FULLSCREEN: hipLaunchKernelGGL(k_vertex<VS_FULLSCREEN>, g, dim3(256), 0, s, p); break;
        default: return hipErrorInvalidValue;
        }
    }
    if (p.tri_count) {
        hipLaunchKernelGGL(k_setup_count, dim3(p.nblocks), dim3(256), 0, s, p);
        hipLaunchKernelGGL(k_clip_count, dim3(512), dim3(64), 0, s, p);
        hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, p);
        hipLaunchKernelGGL(k_setup_emit, dim3(p.nblocks), dim3(256), 0, s, p);
        hipLaunchKernelGGL(k_clip_emit, dim3(512), dim3(64), 0, s, p);
        const uint32_t bgrid = p.nblocks < 8192u ? p.nblocks : 8192u;
        hipLaunchKernelGGL(k_bin<false>, dim3(bgrid), dim3(256), 0, s, p);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, p);
        hipLaunchKernelGGL(k_bin<true>, dim3(bgrid), dim3(256), 0, s, p);
        switch (ps) {
        case PS_BASIC: hipLaunchKernelGGL(k_tile<PS_BASIC>, dim3(ntiles), dim3(256), 0, s, p); break;
        case PS_ENV_LIGHTING: hipLaunchKernelGGL(k_tile<PS_ENV_LIGHTING>, dim3(ntiles), dim3(256), 0, s, p); break;
        case PS_PASSTHROUGH: hipLaunchKernelGGL(k_tile<PS_PASSTHROUGH>, dim3(ntiles), dim3(256), 0, s, p); break;
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

}  // namespace mlv
