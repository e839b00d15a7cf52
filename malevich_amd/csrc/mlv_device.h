// mlv_device.h -- per-lane arithmetic of the draw pipeline for gfx950.
//
// Every function reproduces the reference's IEEE f32 operation sequence (the
// library is compiled with -ffp-contract=off, so a*b+c stays a multiply and an
// add). Integer edge arithmetic wraps mod 2^32 like the reference's AVX2 code.
// Transcendentals (acosf/expf/powf) and 1/sqrt stand in for SVML and the x86
// rsqrt approximation; their results agree to a few ulp, inside the 1/255
// RGBA8 tolerance.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "mlv_internal.h"

namespace mlv {

// The dispatcher deals workgroups to the 8 XCDs in turn (workgroup b to XCD b % 8), and each
// XCD has its own L2. Numbering a grid of n blocks so that XCD x runs the x-th contiguous share
// keeps neighbouring blocks, which gather the same vertices, under one L2.
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t b, uint32_t n) {
    const uint32_t x = b & 7u, q = b >> 3, per = n >> 3, rem = n & 7u;
    return x * per + min(x, rem) + q;
}


// ---------------------------------------------------------------- math.h

// v4f32_dot, math.h:137-140: ((x*x' + y*y') + z*z') + w*w'
__device__ __forceinline__ float dot4_seq(float ax, float ay, float az, float aw, float bx, float by, float bz,
                                          float bw) {
    float r = ax * bx;
    r = r + ay * by;
    r = r + az * bz;
    r = r + aw * bw;
    return r;
}

// A load through a pointer that the kernel read from memory, as a global rather than a flat
// access: a flat load also counts against the scalar-memory counter, so waits for it and for
// scalar loads become one wait. A uniform address with no store of the kernel before it becomes
// a scalar load.
template <class T>
__device__ __forceinline__ T load_global(const T *q) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "a load of 4, 8 or 16 bytes");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) uint32_t *)q);
    } else {
        typedef uint32_t W __attribute__((ext_vector_type(sizeof(T) / 4)));
        return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) W *)q);
    }
}

// v4f256_dot against a matrix row (m4x4f32_mul_v4f256, math.h:142-146,174-177)
__device__ __forceinline__ float dot4_pair(const float *row, float x, float y, float z, float w) {
    float xy = row[0] * x + row[1] * y;
    float zw = row[2] * z + row[3] * w;
    return xy + zw;
}

// v3f256_dot, math.h:153-157
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    float r = ax * bx + ay * by;
    return r + az * bz;
}

// v3f256_normalize, math.h:277-280 (rsqrt approximated by the correctly rounded 1/sqrt)
__device__ __forceinline__ void normalize3(float x, float y, float z, float &ox, float &oy, float &oz) {
    float r = 1.0f / sqrtf(dot3(x, y, z, x, y, z));
    ox = x * r;
    oy = y * r;
    oz = z * r;
}

// _mm256_cvtps_epi32: nearest-even; NaN or out of i32 range gives 0x80000000
__device__ __forceinline__ int32_t cvt_rne(float f) {
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int32_t)rintf(f);
}

__device__ __forceinline__ float lerp_ref(float a, float b, float f) {  // math.h:381-384
    float p = a * (1.0f - f);
    float q = b * f;
    return p + q;
}

// The shaders' power and exponential on the hardware's base-2 logarithm and exponential
// (v_log_f32, v_exp_f32, 1 ulp each): x^y = 2^(y log2 x) and e^x = 2^(x log2 e). The reference
// calls SVML's pow and exp (the oracle, libm's); for the tone and sRGB curves (|y log2 x| of a few
// units) the results differ from a correctly rounded pow by a few parts in 10^7, and the 8-bit
// encodes by at most one step. Zero, infinities, negative bases and NaN map as pow maps them.
__device__ __forceinline__ float shader_pow(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
__device__ __forceinline__ float shader_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }

// f256_srgb_from_linear_approx, math.h:419-422
__device__ __forceinline__ float srgb_approx(float c) {
    float p = shader_pow(c, 0.416666667f);
    float s = 1.055f * p;
    s = s + -0.055f;
    return (s > 0.0f) ? s : 0.0f;
}

// env_lighting_ps.c:20-21, vertex_lighting_vs.c:34-35
__device__ __forceinline__ float tonemap(float c) {
    float e = shader_exp(c * -1.0f);
    return shader_pow(1.0f - e, (float)(1.0 / 2.2));
}

// Output merger encode, main.c:1176-1178
__device__ __forceinline__ uint32_t encode_output(float r, float g, float b) {
    uint32_t ir = (uint32_t)cvt_rne(r * 255.0f);
    uint32_t ig = (uint32_t)cvt_rne(g * 255.0f);
    uint32_t ib = (uint32_t)cvt_rne(b * 255.0f);
    return (ir << 16) + (ig << 8) + ib;
}

// ----------------------------------------------------- texture sampling

__device__ __forceinline__ uint32_t texel_index(const TexDesc &t, int32_t s, int32_t tt) {
    int32_t wm1 = (int32_t)(t.width - 1u), hm1 = (int32_t)(t.height - 1u);
    s = s < wm1 ? s : wm1;
    s = s > 0 ? s : 0;
    tt = tt < hm1 ? tt : hm1;
    tt = tt > 0 ? tt : 0;
    return (uint32_t)tt * t.width + (uint32_t)s;
}

// decode_u32_as_color_x8 (math.h:336-344)
__device__ __forceinline__ float4 decode_rgba8(uint32_t c) {
    const float k = (float)(1.0 / 255.0);
    return make_float4((float)(c & 0xFFu) * k, (float)((c >> 8) & 0xFFu) * k, (float)((c >> 16) & 0xFFu) * k,
                       (float)((c >> 24) & 0xFFu) * k);
}

__device__ __forceinline__ float4 decode_rgba16f(ushort4 h) {
    return make_float4(__half2float(__ushort_as_half(h.x)), __half2float(__ushort_as_half(h.y)),
                       __half2float(__ushort_as_half(h.z)), __half2float(__ushort_as_half(h.w)));
}

// Footprint counting: sets the texel's bit in the texture's bitmap.
__device__ __forceinline__ void mark_texel(const TexDesc &t, uint32_t i) {
    const uint64_t b = t.mark_base + i;
    atomicOr(t.mark + (b >> 5), 1u << (uint32_t)(b & 31u));
}

// get_texel_u_x8 / get_texel_f_x8 (csc.h:30-53) for the four texels of a bilinear footprint:
// one branch on the format, then the four loads in flight together.
__device__ __forceinline__ void fetch4(const TexDesc &t, int32_t s0, int32_t t0, int32_t s1, int32_t t1,
                                       float4 &f00, float4 &f10, float4 &f01, float4 &f11) {
    const uint32_t i00 = texel_index(t, s0, t0), i10 = texel_index(t, s1, t0);
    const uint32_t i01 = texel_index(t, s0, t1), i11 = texel_index(t, s1, t1);
    if (t.mark) {  // footprint counting only
        mark_texel(t, i00);
        mark_texel(t, i10);
        mark_texel(t, i01);
        mark_texel(t, i11);
    }
    if (t.format == FMT_RGBA8) {
        const uint32_t *q = (const uint32_t *)t.data;
        const uint32_t c00 = load_global(q + i00), c10 = load_global(q + i10);
        const uint32_t c01 = load_global(q + i01), c11 = load_global(q + i11);
        f00 = decode_rgba8(c00);
        f10 = decode_rgba8(c10);
        f01 = decode_rgba8(c01);
        f11 = decode_rgba8(c11);
    } else if (t.format == FMT_RGBA32F) {
        const float4 *q = (const float4 *)t.data;
        f00 = load_global(q + i00);
        f10 = load_global(q + i10);
        f01 = load_global(q + i01);
        f11 = load_global(q + i11);
    } else {
        const ushort4 *q = (const ushort4 *)t.data;
        const ushort4 h00 = load_global(q + i00), h10 = load_global(q + i10);
        const ushort4 h01 = load_global(q + i01), h11 = load_global(q + i11);
        f00 = decode_rgba16f(h00);
        f10 = decode_rgba16f(h10);
        f01 = decode_rgba16f(h01);
        f11 = decode_rgba16f(h11);
    }
}

// bilinear_u_x8 / bilinear_f_x8, csc.h:103-121 and csc.h:143-161
__device__ __forceinline__ float4 bilinear(const TexDesc &t, float u, float v) {
    float sf = (float)t.width * u;
    sf = sf + -0.5f;
    float tf = (float)t.height * (1.0f - v);
    tf = tf + -0.5f;
    int32_t s = cvt_rne(floorf(sf));
    int32_t tt = cvt_rne(floorf(tf));
    float fs = sf - (float)s;
    float ft = tf - (float)tt;
    int32_t s1 = (int32_t)((uint32_t)s + 1u), t1 = (int32_t)((uint32_t)tt + 1u);
    float4 t00, t10, t01, t11;
    fetch4(t, s, tt, s1, t1, t00, t10, t01, t11);
    float4 a = make_float4(lerp_ref(t00.x, t10.x, fs), lerp_ref(t00.y, t10.y, fs), lerp_ref(t00.z, t10.z, fs),
                           lerp_ref(t00.w, t10.w, fs));
    float4 b = make_float4(lerp_ref(t01.x, t11.x, fs), lerp_ref(t01.y, t11.y, fs), lerp_ref(t01.z, t11.z, fs),
                           lerp_ref(t01.w, t11.w, fs));
    return make_float4(lerp_ref(a.x, b.x, ft), lerp_ref(a.y, b.y, ft), lerp_ref(a.z, b.z, ft),
                       lerp_ref(a.w, b.w, ft));
}

// sample_2D_latlon_x8, csc.h:226-244
__device__ __forceinline__ float4 latlon(const TexDesc &t, float dx, float dy, float dz) {
    const float TAU_F = 6.283185307f, PI_F = 3.141592654f;  // math.h:33-34
    float kx0, kx1, kx2, ky0, ky1, ky2, kz0, kz1, kz2, cx, cy, cz;
    normalize3(0.0f, 0.0f, 1.0f, kz0, kz1, kz2);
    float cos_theta = dot3(kz0, kz1, kz2, dx, dy, dz);
    normalize3(dx, dy, 0.0f, cx, cy, cz);
    normalize3(1.0f, 0.0f, 0.0f, kx0, kx1, kx2);
    normalize3(0.0f, 1.0f, 0.0f, ky0, ky1, ky2);
    float cos_x = dot3(kx0, kx1, kx2, cx, cy, cz);
    float cos_y = dot3(ky0, ky1, ky2, cx, cy, cz);
    float ax = acosf(cos_x) * (float)(1.0 / (double)TAU_F);
    float uv_x = (cos_y >= 0.0f) ? ax : (1.0f - ax);
    float uv_y = acosf(cos_theta) * (float)(1.0 / (double)PI_F);
    bool pos = cos_theta > 0.999f, neg = cos_theta < -0.999f;
    if (pos || neg) uv_x = 0.5f;
    if (pos) uv_y = 0.0f;
    if (neg) uv_y = 1.0f;
    uv_y = 1.0f - uv_y;
    return bilinear(t, uv_x, uv_y);
}

// ------------------------------------------------------------ mip levels
//
// Extension: the reference's Texture2D has one level (csc.h:20-24). A texture here may hold a
// chain of levels (level 0 first, level l of max(w >> l, 1) x max(h >> l, 1) texels), and the
// extension shaders pick one level per pixel, nearest to the pixel's footprint, then filter it
// bilinearly as csc.h:103-161 does. The footprint compares the sampled coordinate at the pixel
// corner with those at the next corners right and below (the registers interpolated there with
// the same triangle's planes), in level-0 texels; level k is taken when rho^2 >= 2^(2k-1), i.e.
// rho >= 2^(k - 1/2), which needs no logarithm, so every step is an IEEE f32 operation.

__device__ __forceinline__ uint32_t texel_bytes(uint32_t format) {
    return format == FMT_RGBA8 ? 4u : (format == FMT_RGBA16F ? 8u : 16u);
}

// Level l of a texture's chain as a one-level texture.
__device__ __forceinline__ TexDesc tex_level(const TexDesc &t, uint32_t l) {
    uint32_t w = t.width, h = t.height;
    size_t off = 0;
    for (uint32_t k = 0; k < l; ++k) {
        off += (size_t)w * h;
        w = w > 1u ? w >> 1 : 1u;
        h = h > 1u ? h >> 1 : 1u;
    }
    TexDesc r = t;
    r.data = (const uint8_t *)t.data + off * texel_bytes(t.format);
    r.mark_base = t.mark_base + off;
    r.width = w;
    r.height = h;
    r.mip_levels = 1u;
    return r;
}

__device__ __forceinline__ uint32_t nearest_level(float rho2, uint32_t levels) {
    uint32_t lvl = 0u;
    float thr = 2.0f;  // 2^(2k-1) for k = 1
    for (uint32_t k = 1; k < levels; ++k) {
        if (rho2 >= thr) lvl = k;
        thr *= 4.0f;
    }
    return lvl;
}

// rho^2 of a uv footprint: uv at the pixel corner (u, v) and at the next corners right (ux, vx)
// and below (uy, vy), in level-0 texels.
__device__ __forceinline__ float uv_footprint(const TexDesc &t, float u, float v, float ux, float vx, float uy,
                                              float vy) {
    const float tw = (float)t.width, th = (float)t.height;
    const float dux = (ux - u) * tw, dvx = (vx - v) * th, duy = (uy - u) * tw, dvy = (vy - v) * th;
    const float rx = dux * dux + dvx * dvx, ry = duy * duy + dvy * dvy;
    return rx > ry ? rx : ry;
}

// rho^2 of a direction footprint on an equirect map (2 pi radians across its width): the chord
// between the unit directions at the pixel corner and at the next corners, in level-0 texels.
__device__ __forceinline__ float dir_footprint(const TexDesc &t, float dx, float dy, float dz, float xx, float xy,
                                               float xz, float yx, float yy, float yz) {
    const float ax = xx - dx, ay = xy - dy, az = xz - dz, bx = yx - dx, by = yy - dy, bz = yz - dz;
    const float k = (float)t.width * 0.159154943f;  // texels per radian
    const float k2 = k * k;
    const float rx = dot3(ax, ay, az, ax, ay, az) * k2, ry = dot3(bx, by, bz, bx, by, bz) * k2;
    return rx > ry ? rx : ry;
}

// ------------------------------------------------------------- shaders

// Floats of constant buffer slot 0 that each vertex shader reads: the world-view-projection
// rows, or (fullscreen) the view-from-clip and world-from-view rows at floats 16..43.
template <int VS>
struct VsConstants {
    static constexpr int kFloats = VS == VS_FULLSCREEN ? 44 : (VS == VS_PASSTHROUGH ? 0 : 16);
    float cb[kFloats > 0 ? kFloats : 1];
    TexDesc tex;
};

// Vertex shaders (basic_vs.c:22-33, vertex_lighting_vs.c:22-39, passthrough_vs.c:16-25,
// fullscreen_vs.c:22-39). Writes the three output registers.
// kAttrs = false: register 0 only (a draw whose registers 1-2 another draw of the flush writes).
template <int VS, bool kAttrs = true>
__device__ __forceinline__ void vertex_shader(const VsConstants<VS> &k, const float *in, float4 &r0, float4 &r1,
                                              float4 &r2) {
    if constexpr (VS == VS_BASIC || VS == VS_VERTEX_LIGHTING) {
        const float *cfw = k.cb;
        float x = in[0], y = in[1], z = in[2];
        r0 = make_float4(dot4_pair(cfw, x, y, z, 1.0f), dot4_pair(cfw + 4, x, y, z, 1.0f),
                         dot4_pair(cfw + 8, x, y, z, 1.0f), dot4_pair(cfw + 12, x, y, z, 1.0f));
        if constexpr (!kAttrs) return;
        float nx, ny, nz;
        normalize3(in[3], in[4], in[5], nx, ny, nz);
        if constexpr (VS == VS_BASIC) {
            r1 = make_float4(nx, ny, nz, in[6]);
        } else {
            float4 c = latlon(k.tex, nx, ny, nz);
            r1 = make_float4(tonemap(c.x), tonemap(c.y), tonemap(c.z), in[6]);
        }
        r2 = make_float4(in[7], 0.0f, 0.0f, 0.0f);
    } else if constexpr (VS == VS_PASSTHROUGH) {
        r0 = make_float4(in[0] * 2.0f - 1.0f, in[1] * 2.0f - 1.0f, in[2], in[3]);
        r1 = make_float4(in[4], in[5], in[6], 0.0f);
        r2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {  // VS_FULLSCREEN
        float px = in[0] * 2.0f - 1.0f, py = in[1] * 2.0f - 1.0f, pz = in[2], pw = in[3];
        const float *vfc = k.cb + 16, *wfv = k.cb + 32;
        float d0 = dot4_pair(vfc, px, py, pz, pw), d1 = dot4_pair(vfc + 4, px, py, pz, pw);
        float d2 = dot4_pair(vfc + 8, px, py, pz, pw);
        r0 = make_float4(px, py, pz, pw);
        r1 = make_float4(dot4_pair(wfv, d0, d1, d2, 0.0f), dot4_pair(wfv + 4, d0, d1, d2, 0.0f),
                         dot4_pair(wfv + 8, d0, d1, d2, 0.0f), 0.0f);
        r2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// Pixel shaders (basic_ps.c:16-25, env_lighting_ps.c:13-22, passthrough_ps.c:13-19).
// a1/a2 are interpolated registers 1 and 2.
__device__ __forceinline__ void pixel_shader(int ps, const TexDesc &tex, float4 a1, float4 a2, float &r, float &g,
                                             float &b) {
    if (ps == PS_BASIC) {
        float4 c = bilinear(tex, a1.w, a2.x);  // UV = (reg1.w, reg2.x)
        r = srgb_approx(c.x);
        g = srgb_approx(c.y);
        b = srgb_approx(c.z);
    } else if (ps == PS_ENV_LIGHTING) {
        float nx, ny, nz;
        normalize3(a1.x, a1.y, a1.z, nx, ny, nz);
        float4 c = latlon(tex, nx, ny, nz);
        r = tonemap(c.x);
        g = tonemap(c.y);
        b = tonemap(c.z);
    } else {
        r = a1.x;
        g = a1.y;
        b = a1.z;
    }
}

// Extension pixel shaders. a1/a2 are registers 1 and 2 at the pixel corner, x1/x2 and y1/y2 the
// same registers at the next corners right and below (for the footprints).
//  PS_BASIC_LAMBERT (basic_ps.c:16-25 with a directional light): the texture (slot 0) on the level
//   nearest the UV footprint; c = albedo * (light_colour * max(dot(normalize(NORMAL), L), 0) +
//   ambient); basic_ps's sRGB curve. Constant buffer: L.xyz _, light_colour.rgb _, ambient.rgb _.
//   With light_colour 0 and ambient 1 it is basic_ps exactly.
//  PS_ENV_TEXTURED (env_lighting_ps.c:13-22 with an albedo texture): the equirect map (slot 0) at
//   normalize(NORMAL) on the level nearest the direction footprint, times the albedo texture
//   (slot 1) at UV on the level nearest the UV footprint; env_lighting_ps's tone curve. With a
//   one-level map and an albedo of 1 it is env_lighting_ps exactly.
__device__ __forceinline__ void ext_pixel_shader(const DrawRec &D, float4 a1, float4 a2, float4 x1, float4 x2,
                                                 float4 y1, float4 y2, float &r, float &g, float &b) {
    const float u = a1.w, v = a2.x;
    if (D.ps == PS_BASIC_LAMBERT) {
        const TexDesc &t = D.ps_tex;
        const uint32_t lvl =
            t.mip_levels > 1u ? nearest_level(uv_footprint(t, u, v, x1.w, x2.x, y1.w, y2.x), t.mip_levels) : 0u;
        const float4 c = bilinear(lvl ? tex_level(t, lvl) : t, u, v);
        float nx, ny, nz;
        normalize3(a1.x, a1.y, a1.z, nx, ny, nz);
        const float *k = D.ps_cb;
        float ndl = dot3(nx, ny, nz, k[0], k[1], k[2]);
        ndl = ndl > 0.0f ? ndl : 0.0f;
        float lr = k[4] * ndl, lg = k[5] * ndl, lb = k[6] * ndl;
        lr = lr + k[8];
        lg = lg + k[9];
        lb = lb + k[10];
        r = srgb_approx(c.x * lr);
        g = srgb_approx(c.y * lg);
        b = srgb_approx(c.z * lb);
    } else {  // PS_ENV_TEXTURED
        const TexDesc &m = D.ps_tex, &t = D.ps_tex1;
        float dx, dy, dz;
        normalize3(a1.x, a1.y, a1.z, dx, dy, dz);
        uint32_t lm = 0u;
        if (m.mip_levels > 1u) {
            float ax, ay, az, bx, by, bz;
            normalize3(x1.x, x1.y, x1.z, ax, ay, az);
            normalize3(y1.x, y1.y, y1.z, bx, by, bz);
            lm = nearest_level(dir_footprint(m, dx, dy, dz, ax, ay, az, bx, by, bz), m.mip_levels);
        }
        const float4 e = latlon(lm ? tex_level(m, lm) : m, dx, dy, dz);
        const uint32_t lt =
            t.mip_levels > 1u ? nearest_level(uv_footprint(t, u, v, x1.w, x2.x, y1.w, y2.x), t.mip_levels) : 0u;
        const float4 c = bilinear(lt ? tex_level(t, lt) : t, u, v);
        r = tonemap(c.x * e.x);
        g = tonemap(c.y * e.y);
        b = tonemap(c.z * e.z);
    }
}

// ---------------------------------------------------- primitive assembly

// main.c:842-848: floor(v*16 + 0.5) in double, here in float and exact: with x = v*16 (exact),
// floor(x + 0.5) = floor(x) + (x - floor(x) >= 0.5). x - floor(x) is exact except for negative x
// above -0.5, where it rounds monotonically from a value above 0.5 and so stays at or above it;
// x at or beyond 2^24 is an integer (difference 0). Out of the int32 range (or NaN): INT32_MIN.
__device__ __forceinline__ int32_t snap28_4(float v) {
    const float x = v * 16.0f;
    const float fl = floorf(x);
    const float r = (x - fl >= 0.5f) ? fl + 1.0f : fl;
    if (!(r >= -2147483648.0f && r < 2147483648.0f)) return INT32_MIN;
    return (int32_t)r;
}

// set_edge_function, main.c:563-575, u32 wrap
__device__ __forceinline__ void set_edge(int32_t area, int32_t x0, int32_t y0, int32_t x1, int32_t y1, int32_t &ea,
                                         int32_t &eb, int32_t &ec) {
    uint32_t a = (uint32_t)y0 - (uint32_t)y1;
    uint32_t b = (uint32_t)x1 - (uint32_t)x0;
    if (area < 0) {
        a = 0u - a;
        b = 0u - b;
    }
    uint32_t c = (0u - a) * (uint32_t)x0 - b * (uint32_t)y0;
    ea = (int32_t)a;
    eb = (int32_t)b;
    ec = (int32_t)c;
}

__device__ __forceinline__ float max_ref(float x, float y) { return (x > y) ? x : y; }  // math.h:28-30
__device__ __forceinline__ float min_ref(float x, float y) { return (x < y) ? x : y; }  // math.h:25-27
__device__ __forceinline__ int32_t imin(int32_t x, int32_t y) { return (x < y) ? x : y; }
__device__ __forceinline__ int32_t imax(int32_t x, int32_t y) { return (x > y) ? x : y; }

// Trivial frustum reject (main.c:764-772) and the "no clipping needed" test (main.c:775-781).
// Returns 0 = dropped, 1 = inside, 2 = needs clipping.
__device__ __forceinline__ int classify(float4 a, float4 b, float4 c) {
    if (a.w == 0.0f || b.w == 0.0f || c.w == 0.0f) return 0;  // main.c:759
    if ((a.x < -a.w && b.x < -b.w && c.x < -c.w) || (a.x > +a.w && b.x > +b.w && c.x > +c.w) ||
        (a.y < -a.w && b.y < -b.w && c.y < -c.w) || (a.y > +a.w && b.y > +b.w && c.y > +c.w) ||
        (a.z < 0.f && b.z < 0.f && c.z < 0.f) || (a.z > +a.w && b.z > +b.w && c.z > +c.w))
        return 0;
    bool inside = (a.x >= -a.w && b.x >= -b.w && c.x >= -c.w) && (a.x <= +a.w && b.x <= +b.w && c.x <= +c.w) &&
                  (a.y >= -a.w && b.y >= -b.w && c.y >= -c.w) && (a.y <= +a.w && b.y <= +b.w && c.y <= +c.w) &&
                  (a.z >= 0.f && b.z >= 0.f && c.z >= 0.f) && (a.z <= +a.w && b.z <= +b.w && c.z <= +c.w);
    return inside ? 1 : 2;
}

// Projection, viewport transform, snapping, area, cull and setup for one fan triangle
// (main.c:799-898). Returns false when back-face culled (signed_area > 0).
// kInside: the triangle passed the frustum's inside test (main.c:775-781), so w > 0 and the
// normalised x, y, z are finite whenever 1/w is. The viewport matrix has zeros off its diagonal
// and last column (screen_matrix), and a product with a zero entry is then a signed zero, which
// leaves a nonzero running sum unchanged; so x and y are formed from their two nonzero terms in
// the same order, and can differ from the four-term sum only in the sign of a zero result, which
// snaps to 0 either way (an infinite 1/w gives NaN or infinity both ways, which snap alike). z
// keeps all four terms: its sign of zero reaches the depth buffer.
template <bool kInside = false>
__device__ __forceinline__ bool setup_triangle(const DrawRec &p, const ViewRec &v, float4 q0, float4 q1, float4 q2,
                                               TriVis &tv, float rw[3]) {
    float4 q[3] = {q0, q1, q2};
    float sx[3], sy[3], sz[3];
    int32_t x[3], y[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        rw[k] = 1.0f / q[k].w;  // (f32)(1.0/(f64)w) equals the f32 quotient
        float X = q[k].x * rw[k], Y = q[k].y * rw[k], Z = q[k].z * rw[k], Wn = q[k].w * rw[k];
        if (kInside) {
            sx[k] = p.screen[0][0] * X + p.screen[0][3] * Wn;
            sy[k] = p.screen[1][1] * Y + p.screen[1][3] * Wn;
        } else {
            sx[k] = dot4_seq(p.screen[0][0], p.screen[0][1], p.screen[0][2], p.screen[0][3], X, Y, Z, Wn);
            sy[k] = dot4_seq(p.screen[1][0], p.screen[1][1], p.screen[1][2], p.screen[1][3], X, Y, Z, Wn);
        }
        sz[k] = dot4_seq(p.screen[2][0], p.screen[2][1], p.screen[2][2], p.screen[2][3], X, Y, Z, Wn);
        x[k] = snap28_4(sx[k]);
        y[k] = snap28_4(sy[k]);
    }
    int32_t area = (int32_t)(((uint32_t)x[1] - (uint32_t)x[0]) * ((uint32_t)y[2] - (uint32_t)y[0]) -
                             ((uint32_t)x[2] - (uint32_t)x[0]) * ((uint32_t)y[1] - (uint32_t)y[0]));
    // main.c:856 culls signed_area > 0; the other modes are extensions (set_edge orients the edge
    // functions by the area's sign, so a kept triangle of either winding is set up alike)
    if (p.cull == CULL_BACK ? area > 0 : (p.cull == CULL_FRONT ? area < 0 : false)) return false;
    int32_t ea, eb, ec;
    uint32_t excl = 0u;  // top-left rule (extension): edges neither left (a > 0) nor top (a == 0, b > 0)
    set_edge(area, x[1], y[1], x[2], y[2], ea, eb, ec);  // slot 0 = (v1, v2)
    tv.a16[0] = (int32_t)((uint32_t)ea << 4); tv.b16[0] = (int32_t)((uint32_t)eb << 4); tv.c[0] = ec;
    excl |= (ea > 0 || (ea == 0 && eb > 0)) ? 0u : 1u;
    set_edge(area, x[2], y[2], x[0], y[0], ea, eb, ec);  // slot 1 = (v2, v0)
    tv.a16[1] = (int32_t)((uint32_t)ea << 4); tv.b16[1] = (int32_t)((uint32_t)eb << 4); tv.c[1] = ec;
    excl |= (ea > 0 || (ea == 0 && eb > 0)) ? 0u : 2u;
    set_edge(area, x[0], y[0], x[1], y[1], ea, eb, ec);  // slot 2 = (v0, v1)
    tv.a16[2] = (int32_t)((uint32_t)ea << 4); tv.b16[2] = (int32_t)((uint32_t)eb << 4); tv.c[2] = ec;
    excl |= (ea > 0 || (ea == 0 && eb > 0)) ? 0u : 4u;
    tv.draw = p.fill == FILL_TOP_LEFT ? excl << 16 : 0u;  // the draw index is added when written
    float af = (float)(area >> 8);
    if (af == 0.0f) af = 1.0f;
    tv.ooa = fabsf(1.f / af);
    tv.z0 = sz[0];
    tv.dz1 = sz[1] - sz[0];
    tv.dz2 = sz[2] - sz[0];
    float mx = max_ref(sz[0], max_ref(sz[1], sz[2]));
    float mn = min_ref(sz[0], min_ref(sz[1], sz[2]));
    float S = fabsf(sz[0]) + fabsf(sz[1]) + fabsf(sz[2]);
    if (p.depth_func == DEPTH_GE) {
        tv.hiz_depth = mx;
        tv.cull_bound = (mx + S * 3.814697265625e-06f) + 1e-30f;  // 2^-18 relative margin
    } else {
        tv.hiz_depth = mn;
        tv.cull_bound = (mn - S * 3.814697265625e-06f) - 1e-30f;
    }
    // bounds (main.c:888-898) then the 8x8 bin rect (main.c:930-931)
    int32_t mnx = imin(x[0], imin(x[1], x[2])) >> 4, mny = imin(y[0], imin(y[1], y[2])) >> 4;
    mnx = imin(imax(mnx, 0), v.W - 1);
    mny = imin(imax(mny, 0), v.H - 1);
    int32_t mxx = imax(x[0], imax(x[1], x[2])) >> 4, mxy = imax(y[0], imax(y[1], y[2])) >> 4;
    mxx = imin(mxx + 1, v.W - 1);
    mxy = imin(mxy + 1, v.H - 1);
    int32_t bx0 = imax(mnx / kBin, 0), by0 = imax(mny / kBin, 0);
    int32_t bx1 = imin(mxx / kBin, v.bins_x - 1), by1 = imin(mxy / kBin, v.bins_y - 1);
    tv.rect0 = (uint32_t)(uint16_t)bx0 | ((uint32_t)(uint16_t)by0 << 16);
    tv.rect1 = (uint32_t)(uint16_t)(int16_t)bx1 | ((uint32_t)(uint16_t)(int16_t)by1 << 16);
    // pixel corners in the snapped bounding box: x from ceil(min/16) to floor(max/16)
    const int32_t lx0 = imax((imin(x[0], imin(x[1], x[2])) + 15) >> 4, 0);
    const int32_t lx1 = imin(imax(x[0], imax(x[1], x[2])) >> 4, v.W - 1);
    const int32_t ly0 = imax((imin(y[0], imin(y[1], y[2])) + 15) >> 4, 0);
    const int32_t ly1 = imin(imax(y[0], imax(y[1], y[2])) >> 4, v.H - 1);
    tv.lat_x = (uint32_t)(uint16_t)(int16_t)imin(lx0, 32767) | ((uint32_t)(uint16_t)(int16_t)imax(lx1, -1) << 16);
    tv.lat_y = (uint32_t)(uint16_t)(int16_t)imin(ly0, 32767) | ((uint32_t)(uint16_t)(int16_t)imax(ly1, -1) << 16);
    return true;
}

// The 8x8-bin rect packed in TriVis::rect0/rect1 (empty when bx1 < bx0 or by1 < by0).
__device__ __forceinline__ void rect_decode(uint32_t rect0, uint32_t rect1, int32_t &bx0, int32_t &by0, int32_t &bx1,
                                            int32_t &by1) {
    bx0 = (int32_t)(rect0 & 0xFFFFu);
    by0 = (int32_t)(rect0 >> 16);
    bx1 = (int32_t)(int16_t)(rect1 & 0xFFFFu);
    by1 = (int32_t)(int16_t)(rect1 >> 16);
}
__device__ __forceinline__ void rect_of(const TriVis &tv, int32_t &bx0, int32_t &by0, int32_t &bx1, int32_t &by1) {
    rect_decode(tv.rect0, tv.rect1, bx0, by0, bx1, by1);
}

// Whether the snapped bounding box of a set-up triangle holds a pixel corner of the view. Coverage
// is sampled at pixel corners only (main.c:1015-1025), so a triangle without one covers nothing.
__device__ __forceinline__ bool has_corner(const TriVis &tv) {
    return (int32_t)(int16_t)(tv.lat_x & 0xFFFFu) <= (int32_t)(int16_t)(tv.lat_x >> 16) &&
           (int32_t)(int16_t)(tv.lat_y & 0xFFFFu) <= (int32_t)(int16_t)(tv.lat_y >> 16);
}

}  // namespace mlv
