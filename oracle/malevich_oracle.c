/*
 * malevich_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Scalar CPU restatement of the Malevich D3D11-style draw pipeline. Every
 * function cites the reference lines it restates. Build with
 *   gcc -O2 -ffp-contract=off -fwrapv -fno-fast-math -fPIC -shared
 * (see oracle/Makefile). Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library.
 *
 * Parity pins (checks under tests/ that would fail on a plausible mistake):
 *   coverage / edge functions / snapping  -- exhaustive brute force against an
 *       exact-rational point-in-closed-triangle test, the E0+E1+E2 = |area|
 *       invariant, shared-edge double coverage, zero-area => no coverage.
 *   binning            -- brute-force rect enumeration of (triangle, bin) pairs.
 *   depth interpolation -- constant-z triangles give z exactly; vertex corners
 *       reproduce vertex z; against an FP64 evaluation within 4 ulp.
 *   perspective correction -- against FP64 ray/plane intersection.
 *   clipping           -- clipped polygon vs an FP64 Sutherland-Hodgman on the
 *       same planes (vertex positions within tolerance, all inside).
 *   bilinear           -- constant textures, texel-centre exactness, linear ramps.
 *   encode/clear       -- integer tables (k/255 round trip, truncation).
 *   latlon             -- closed form with acos, pole special cases.
 *   full pipeline      -- Suprematism scene against an exact-rational scene
 *       render (visibility + colour) built independently in the test.
 * Parity unpinned: rsqrt (x86 approximation) and SVML transcendentals, see
 * malevich_oracle.h.
 */
#include "malevich_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

#define NREG 3                  /* out_vertex_size/(16*8) of every reference VS, main.c:1287 */
#define MAX_CLIP_VERTICES 16    /* main.c:38 */
#define SUBPIX_BITS 4           /* main.c:39 */
#define BIN 8                   /* main.c:24-25 */

typedef struct { float x, y, z, w; } v4;

/* ---------------------------------------------------------------- math.h */

/* math.h:137-140 v4f32_dot: ((x*x' + y*y') + z*z') + w*w' */
static inline float dot4_seq(v4 a, v4 b) {
    float r = a.x * b.x;
    r = r + a.y * b.y;
    r = r + a.z * b.z;
    r = r + a.w * b.w;
    return r;
}

/* math.h:142-146 v4f256_dot with a matrix row (m4x4f32_mul_v4f256, math.h:174-177):
 * (r0*x + r1*y) + (r2*z + r3*w) */
static inline float dot4_pair(const float *row, float x, float y, float z, float w) {
    float xy = row[0] * x + row[1] * y;
    float zw = row[2] * z + row[3] * w;
    return xy + zw;
}

/* math.h:153-157 v3f256_dot: (x*x' + y*y') + z*z' */
static inline float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    float r = ax * bx + ay * by;
    return r + az * bz;
}

/* math.h:277-280 v3f256_normalize; rsqrt stand-in is 1/sqrt (parity unpinned). */
static inline void normalize3(float x, float y, float z, float *o) {
    float r = 1.0f / sqrtf(dot3(x, y, z, x, y, z));
    o[0] = x * r;
    o[1] = y * r;
    o[2] = z * r;
}

void orc_normalize3(const float in[3], float out[3]) { normalize3(in[0], in[1], in[2], out); }

/* _mm256_cvtps_epi32: round to nearest even; out of range or NaN -> 0x80000000. */
static inline int32_t cvt_rne(float f) {
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int32_t)rintf(f);
}
int32_t orc_cvt_rne(float f) { return cvt_rne(f); }

/* math.h:336-344 decode_u32_as_color_x8: byte * (float)(1/255), R in the low byte. */
static inline void decode_rgba8(uint32_t c, float o[4]) {
    const float k = (float)(1.0 / 255.0);
    o[0] = (float)(c & 0xFFu) * k;
    o[1] = (float)((c >> 8) & 0xFFu) * k;
    o[2] = (float)((c >> 16) & 0xFFu) * k;
    o[3] = (float)((c >> 24) & 0xFFu) * k;
}

/* math.h:381-384 v4f256_lerp: v0*(1-f) + v1*f */
static inline float lerp(float a, float b, float f) {
    float p = a * (1.0f - f);
    float q = b * f;
    return p + q;
}

/* math.h:419-422 f256_srgb_from_linear_approx: max(1.055*pow(c,0.416666667) - 0.055, 0).
 * maxps returns its second operand (0) for NaN and for equal zeros. */
float orc_srgb_from_linear_approx(float c) {
    float p = powf(c, 0.416666667f);
    float s = 1.055f * p;
    s = s + -0.055f;
    return (s > 0.0f) ? s : 0.0f;
}

/* env_lighting_ps.c:20-21 / vertex_lighting_vs.c:34-35: pow(1 - exp(-c), 1/2.2) */
float orc_tonemap(float c) {
    float e = expf(c * -1.0f);
    return powf(1.0f - e, (float)(1.0 / 2.2));
}

/* math.h:322-324 encode_color_as_u32 (clear path): truncating float->u32 per channel,
 * A<<24 | B<<16 | G<<8 | R. Conversion follows x86-64 cvttss2si (64-bit) low word. */
static inline uint32_t trunc_u32(float f) {
    if (!(f > -9.2e18f && f < 9.2e18f)) return 0u;
    return (uint32_t)(int64_t)f;
}
uint32_t orc_encode_clear(const float rgba[4]) {
    uint32_t r = trunc_u32(rgba[0] * 255.f), g = trunc_u32(rgba[1] * 255.f);
    uint32_t b = trunc_u32(rgba[2] * 255.f), a = trunc_u32(rgba[3] * 255.f);
    return (a << 24) + (b << 16) + (g << 8) + r;
}

/* load_texture's pre-pass for sRGB textures (main.c:546-558): each RGBA8 texel is decoded as
 * byte * f32(1/255) (math.h:326-334), every channel including alpha goes through srgb_to_linear
 * (math.h:386-394, evaluated in double as written and stored as f32), and the result is
 * re-encoded with the truncating encode_color_as_u32 (math.h:322-324). */
static float srgb_to_linear_ref(float v) {
    if ((double)v <= 0.04045) return (float)((double)v / 12.92);
    return (float)pow(((double)v + 0.055) / 1.055, 2.4);
}

void orc_srgb_to_linear_rgba8(const uint32_t *in, uint32_t *out, size_t n) {
    const float normalizer = (float)(1.0 / 255.0);
    for (size_t i = 0; i < n; ++i) {
        float c[4];
        for (int k = 0; k < 4; ++k) c[k] = srgb_to_linear_ref((float)((in[i] >> (8 * k)) & 255u) * normalizer);
        out[i] = orc_encode_clear(c);
    }
}

/* main.c:1176-1178 output merger: (rne(r*255)<<16) + (rne(g*255)<<8) + rne(b*255), no clamp. */
uint32_t orc_encode_output(const float rgb[3]) {
    uint32_t r = (uint32_t)cvt_rne(rgb[0] * 255.0f);
    uint32_t g = (uint32_t)cvt_rne(rgb[1] * 255.0f);
    uint32_t b = (uint32_t)cvt_rne(rgb[2] * 255.0f);
    return (r << 16) + (g << 8) + b;
}

/* ------------------------------------------------- common_shader_core.h */

/* csc.h:30-36 get_texel_u_x8 and csc.h:42-53 get_texel_f_x8: clamp to edge, t*w + s. */
static inline uint32_t texel_index(const orc_texture *t, int32_t s, int32_t tt) {
    int32_t wm1 = (int32_t)(t->width - 1u), hm1 = (int32_t)(t->height - 1u);
    s = s < wm1 ? s : wm1;
    s = s > 0 ? s : 0;
    tt = tt < hm1 ? tt : hm1;
    tt = tt > 0 ? tt : 0;
    return (uint32_t)tt * t->width + (uint32_t)s;
}

static inline void fetch(const orc_texture *t, int32_t s, int32_t tt, float o[4]) {
    uint32_t i = texel_index(t, s, tt);
    if (t->format == ORC_FMT_RGBA8) {
        decode_rgba8(((const uint32_t *)t->data)[i], o);
    } else {
        const float *p = (const float *)t->data + 4u * i;
        o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; o[3] = p[3];
    }
}

/* csc.h:103-121 bilinear_u_x8 and csc.h:143-161 bilinear_f_x8. */
void orc_bilinear(const orc_texture *t, float u, float v, float out[4]) {
    float sf = (float)t->width * u;
    sf = sf + -0.5f;
    float tf = (float)t->height * (1.0f - v);
    tf = tf + -0.5f;
    int32_t s = cvt_rne(floorf(sf));
    int32_t tt = cvt_rne(floorf(tf));
    float fs = sf - (float)s;
    float ft = tf - (float)tt;
    float t00[4], t10[4], t01[4], t11[4];
    fetch(t, s, tt, t00);
    fetch(t, (int32_t)((uint32_t)s + 1u), tt, t10);
    fetch(t, s, (int32_t)((uint32_t)tt + 1u), t01);
    fetch(t, (int32_t)((uint32_t)s + 1u), (int32_t)((uint32_t)tt + 1u), t11);
    for (int c = 0; c < 4; ++c) {
        float a = lerp(t00[c], t10[c], fs);
        float b = lerp(t01[c], t11[c], fs);
        out[c] = lerp(a, b, ft);
    }
}

/* csc.h:226-244 sample_2D_latlon_x8. */
void orc_latlon(const orc_texture *t, const float dir[3], float out[4]) {
    const float TAU_F = 6.283185307f, PI_F = 3.141592654f; /* math.h:33-34 */
    float kz[3], kx[3], ky[3], cxy[3];
    normalize3(0.0f, 0.0f, 1.0f, kz);
    float cos_theta = dot3(kz[0], kz[1], kz[2], dir[0], dir[1], dir[2]);
    normalize3(dir[0], dir[1], 0.0f, cxy);
    normalize3(1.0f, 0.0f, 0.0f, kx);
    normalize3(0.0f, 1.0f, 0.0f, ky);
    float cos_x = dot3(kx[0], kx[1], kx[2], cxy[0], cxy[1], cxy[2]);
    float cos_y = dot3(ky[0], ky[1], ky[2], cxy[0], cxy[1], cxy[2]);
    float ax = acosf(cos_x) * (float)(1.0 / (double)TAU_F);
    float uv_x = (cos_y >= 0.0f) ? ax : (1.0f - ax);
    float uv_y = acosf(cos_theta) * (float)(1.0 / (double)PI_F);
    int pos = cos_theta > 0.999f, neg = cos_theta < -0.999f;
    if (pos || neg) uv_x = 0.5f;
    if (pos) uv_y = 0.0f;
    if (neg) uv_y = 1.0f;
    uv_y = 1.0f - uv_y;
    orc_bilinear(t, uv_x, uv_y, out);
}

/* ------------------------------------------------ mip levels (extension)
 * The library's extension shaders pick one level of a texture's chain per pixel, nearest to the
 * footprint of the sampled coordinate between the pixel corner and the next corners right and
 * below; level k when rho^2 >= 2^(2k-1). Every step is an IEEE f32 operation in this order. */

static orc_texture tex_level(const orc_texture *t, uint32_t l) {
    uint32_t w = t->width, h = t->height;
    size_t off = 0;
    for (uint32_t k = 0; k < l; ++k) {
        off += (size_t)w * h;
        w = w > 1u ? w >> 1 : 1u;
        h = h > 1u ? h >> 1 : 1u;
    }
    const size_t tb = t->format == ORC_FMT_RGBA8 ? 4 : 16;
    orc_texture r = *t;
    r.data = (const uint8_t *)t->data + off * tb;
    r.width = w;
    r.height = h;
    r.levels = 1;
    return r;
}

static uint32_t nearest_level(float rho2, uint32_t levels) {
    uint32_t lvl = 0;
    float thr = 2.0f;
    for (uint32_t k = 1; k < levels; ++k) {
        if (rho2 >= thr) lvl = k;
        thr *= 4.0f;
    }
    return lvl;
}

static float uv_footprint(const orc_texture *t, float u, float v, float ux, float vx, float uy, float vy) {
    const float tw = (float)t->width, th = (float)t->height;
    const float dux = (ux - u) * tw, dvx = (vx - v) * th, duy = (uy - u) * tw, dvy = (vy - v) * th;
    const float rx = dux * dux + dvx * dvx, ry = duy * duy + dvy * dvy;
    return rx > ry ? rx : ry;
}

/* chord between unit directions, in level-0 texels of an equirect map (2 pi across its width) */
static float dir_footprint(const orc_texture *t, const float d[3], const float x[3], const float y[3]) {
    const float ax = x[0] - d[0], ay = x[1] - d[1], az = x[2] - d[2];
    const float bx = y[0] - d[0], by = y[1] - d[1], bz = y[2] - d[2];
    const float k = (float)t->width * 0.159154943f;
    const float k2 = k * k;
    const float rx = dot3(ax, ay, az, ax, ay, az) * k2, ry = dot3(bx, by, bz, bx, by, bz) * k2;
    return rx > ry ? rx : ry;
}

uint32_t orc_nearest_level(float rho2, uint32_t levels) { return nearest_level(rho2, levels); }

/* bilinear_u_x8 / bilinear_f_x8 on level `level` of a texture's chain. */
void orc_bilinear_level(const orc_texture *t, uint32_t level, float u, float v, float out[4]) {
    const orc_texture tl = tex_level(t, level);
    orc_bilinear(&tl, u, v, out);
}

/* ------------------------------------------------------------- shaders */

static const float *vertex_ptr(const orc_draw_desc *d, uint32_t vi) {
    return (const float *)((const uint8_t *)d->vertex_buffer + (size_t)vi * d->vertex_stride);
}

/* basic_vs.c:22-33, vertex_lighting_vs.c:22-39, passthrough_vs.c:16-25, fullscreen_vs.c:22-39.
 * Output is the 12-float AoS record written by vertex_shader_stage (main.c:714-727). */
int orc_vertex_shader(const orc_draw_desc *d, uint32_t vi, float o[12]) {
    const float *in = vertex_ptr(d, vi);
    const float *cfw = d->cb, *vfc = d->cb + 16, *wfv = d->cb + 32;
    memset(o, 0, 12 * sizeof(float));
    switch (d->vs) {
    case ORC_VS_BASIC:
    case ORC_VS_VERTEX_LIGHTING: {
        for (int r = 0; r < 4; ++r) o[r] = dot4_pair(cfw + 4 * r, in[0], in[1], in[2], 1.0f);
        float n[3];
        normalize3(in[3], in[4], in[5], n);
        if (d->vs == ORC_VS_BASIC) {
            o[4] = n[0]; o[5] = n[1]; o[6] = n[2];
        } else {
            float c[4];
            orc_latlon(&d->vs_srv0, n, c);
            o[4] = orc_tonemap(c[0]); o[5] = orc_tonemap(c[1]); o[6] = orc_tonemap(c[2]);
        }
        o[7] = in[6];
        o[8] = in[7];
        return ORC_OK;
    }
    case ORC_VS_PASSTHROUGH:
        o[0] = in[0] * 2.0f - 1.0f; /* _mm256_fmadd_ps(x, 2, -1): 2x is exact */
        o[1] = in[1] * 2.0f - 1.0f;
        o[2] = in[2];
        o[3] = in[3];
        o[4] = in[4]; o[5] = in[5]; o[6] = in[6];
        return ORC_OK;
    case ORC_VS_FULLSCREEN: {
        float p[4] = {in[0] * 2.0f - 1.0f, in[1] * 2.0f - 1.0f, in[2], in[3]};
        float dv[4];
        for (int r = 0; r < 4; ++r) dv[r] = dot4_pair(vfc + 4 * r, p[0], p[1], p[2], p[3]);
        o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; o[3] = p[3];
        for (int r = 0; r < 3; ++r) o[4 + r] = dot4_pair(wfv + 4 * r, dv[0], dv[1], dv[2], 0.0f);
        return ORC_OK;
    }
    default:
        return ORC_ERR_INVALID_ARG;
    }
}

/* basic_ps.c:16-25, env_lighting_ps.c:13-22, passthrough_ps.c:13-19. in = 12 interpolated floats;
 * inx / iny: the same at the next pixel corners right and below (extension shaders only). */
static void pixel_shader(const orc_draw_desc *d, const float in[12], const float inx[12], const float iny[12],
                         float rgb[3]) {
    switch (d->ps) {
    case ORC_PS_BASIC_LAMBERT: {
        /* basic_ps with a directional light: albedo * (light_colour * max(N.L, 0) + ambient) */
        const orc_texture *t = &d->ps_srv0;
        const uint32_t lvl =
            t->levels > 1 ? nearest_level(uv_footprint(t, in[7], in[8], inx[7], inx[8], iny[7], iny[8]), t->levels) : 0;
        const orc_texture tl = tex_level(t, lvl);
        float c[4], n[3];
        orc_bilinear(&tl, in[7], in[8], c);
        normalize3(in[4], in[5], in[6], n);
        const float *k = d->ps_cb;
        float ndl = dot3(n[0], n[1], n[2], k[0], k[1], k[2]);
        ndl = ndl > 0.0f ? ndl : 0.0f;
        for (int j = 0; j < 3; ++j) {
            float l = k[4 + j] * ndl;
            l = l + k[8 + j];
            rgb[j] = orc_srgb_from_linear_approx(c[j] * l);
        }
        break;
    }
    case ORC_PS_ENV_TEXTURED: {
        /* env_lighting_ps times an albedo texture, both on the level nearest their footprints */
        const orc_texture *m = &d->ps_srv0, *t = &d->ps_srv1;
        float dir[3], e[4], c[4];
        normalize3(in[4], in[5], in[6], dir);
        uint32_t lm = 0;
        if (m->levels > 1) {
            float dx[3], dy[3];
            normalize3(inx[4], inx[5], inx[6], dx);
            normalize3(iny[4], iny[5], iny[6], dy);
            lm = nearest_level(dir_footprint(m, dir, dx, dy), m->levels);
        }
        const orc_texture ml = tex_level(m, lm);
        orc_latlon(&ml, dir, e);
        const uint32_t lt =
            t->levels > 1 ? nearest_level(uv_footprint(t, in[7], in[8], inx[7], inx[8], iny[7], iny[8]), t->levels) : 0;
        const orc_texture tl = tex_level(t, lt);
        orc_bilinear(&tl, in[7], in[8], c);
        for (int j = 0; j < 3; ++j) rgb[j] = orc_tonemap(c[j] * e[j]);
        break;
    }
    case ORC_PS_BASIC: {
        float c[4];
        orc_bilinear(&d->ps_srv0, in[7], in[8], c);
        for (int k = 0; k < 3; ++k) rgb[k] = orc_srgb_from_linear_approx(c[k]);
        break;
    }
    case ORC_PS_ENV_LIGHTING: {
        float n[3], c[4];
        normalize3(in[4], in[5], in[6], n);
        orc_latlon(&d->ps_srv0, n, c);
        for (int k = 0; k < 3; ++k) rgb[k] = orc_tonemap(c[k]);
        break;
    }
    default:
        rgb[0] = in[4]; rgb[1] = in[5]; rgb[2] = in[6];
        break;
    }
}

/* ------------------------------------------------------------ clipping */

/* main.c:652-657 plane normals v4f32_normalize((+-1,0,0,1)) etc.; v4f32_normalize
 * (math.h:267-270) takes a double sqrt of the f32 dot, rounds to f32, then a
 * double 1.0/len rounded to f32. */
void orc_clip_planes(float out[6][4]) {
    static const float raw[6][4] = {{1, 0, 0, 1}, {-1, 0, 0, 1}, {0, 1, 0, 1},
                                    {0, -1, 0, 1}, {0, 0, 1, 1}, {0, 0, -1, 1}};
    for (int p = 0; p < 6; ++p) {
        v4 v = {raw[p][0], raw[p][1], raw[p][2], raw[p][3]};
        float len = (float)sqrt((double)dot4_seq(v, v));
        float ool = (float)(1.0 / (double)len);
        out[p][0] = v.x * ool; out[p][1] = v.y * ool; out[p][2] = v.z * ool; out[p][3] = v.w * ool;
    }
}

/* main.c:609-647 clip_by_plane with plane_d = 0: inside is dot > -0; t = (0 + d0)/(d0 - d1);
 * every register lerped as a*(1-t) + b*t. Vertices are 12 floats (3 registers). */
static int clip_by_plane(float *v, int n, const float plane[4]) {
    float out[MAX_CLIP_VERTICES][12];
    int k = 0;
    const float plane_d = 0.0f;
    v4 pn = {plane[0], plane[1], plane[2], plane[3]};
    if (n <= 0) return 0;
    v4 p0 = {v[0], v[1], v[2], v[3]};
    float cur = dot4_seq(pn, p0);
    int cur_in = cur > -plane_d;
    for (int i = 0; i < n; ++i) {
        if (k >= MAX_CLIP_VERTICES) return -1; /* main.c:621 assert */
        int nx = (i + 1) % n;
        if (cur_in) memcpy(out[k++], v + 12 * i, 12 * sizeof(float));
        v4 pnx = {v[12 * nx], v[12 * nx + 1], v[12 * nx + 2], v[12 * nx + 3]};
        float nd = dot4_seq(pn, pnx);
        int nx_in = nd > -plane_d;
        if (cur_in != nx_in) {
            if (k >= MAX_CLIP_VERTICES) return -1;
            float t = (plane_d + cur) / (cur - nd);
            float omt = 1.0f - t;
            for (int c = 0; c < 4 * NREG; ++c) {
                float a = v[12 * i + c] * omt;
                float b = v[12 * nx + c] * t;
                out[k][c] = a + b;
            }
            k++;
        }
        cur = nd;
        cur_in = nx_in;
    }
    memcpy(v, out, (size_t)k * 12 * sizeof(float));
    return k;
}

/* main.c:649-660 clipper: the six planes in order. */
int orc_clip_polygon(const float *in, int n, float *out) {
    float planes[6][4];
    orc_clip_planes(planes);
    memcpy(out, in, (size_t)n * 12 * sizeof(float));
    for (int p = 0; p < 6; ++p) {
        n = clip_by_plane(out, n, planes[p]);
        if (n < 0) return -1;
    }
    return n;
}

/* ------------------------------------------------- primitive assembly */

#define MIN_(x, y) (((x) < (y)) ? (x) : (y)) /* math.h:25-27 */
#define MAX_(x, y) (((x) > (y)) ? (x) : (y)) /* math.h:28-30 */

/* main.c:563-575 set_edge_function, in u32 arithmetic (two's-complement wrap). */
static void set_edge(int32_t e[3], int32_t area, int32_t x0, int32_t y0, int32_t x1, int32_t y1) {
    uint32_t a = (uint32_t)y0 - (uint32_t)y1;
    uint32_t b = (uint32_t)x1 - (uint32_t)x0;
    if (area < 0) { a = 0u - a; b = 0u - b; }
    uint32_t c = (0u - a) * (uint32_t)x0 - b * (uint32_t)y0;
    e[0] = (int32_t)a; e[1] = (int32_t)b; e[2] = (int32_t)c;
}

static int is_valid_vs_ps(const orc_draw_desc *d) {
    if (!(d->vs >= ORC_VS_BASIC && d->vs <= ORC_VS_FULLSCREEN && d->ps >= ORC_PS_BASIC && d->ps <= ORC_PS_NONE))
        return 0;
    if (d->ps == ORC_PS_BASIC_LAMBERT && (!d->ps_cb || !d->ps_srv0.data)) return 0;
    if (d->ps == ORC_PS_ENV_TEXTURED && (!d->ps_srv0.data || !d->ps_srv1.data)) return 0;
    return 1;
}

/* main.c:662-734 (IA + VS, de-indexed) and main.c:736-913 (primitive assembly) in input order. */
/* Primitive assembly of input triangles [t0, t1) in input order (main.c:736-913): up to
 * `capacity` triangles go to out[0..] (out may be NULL to only count them); *n counts them. */
static int pa_range(const orc_draw_desc *d, const float planes[6][4], const float M[4][4], int32_t W, int32_t H,
                    uint32_t t0, uint32_t t1, orc_triangle *out, uint32_t capacity, uint32_t *n, uint64_t *oob) {
    for (uint32_t t = t0; t < t1; ++t) {
        float cv[MAX_CLIP_VERTICES][12];
        int bad = 0;
        for (int k = 0; k < 3; ++k) {
            int64_t vi = (int64_t)d->index_buffer[d->start_index + 3u * t + (uint32_t)k] + d->base_vertex;
            if (vi < 0 || vi >= (int64_t)d->vertex_count) { bad = 1; break; }
            orc_vertex_shader(d, (uint32_t)vi, cv[k]);
        }
        if (bad) { (*oob)++; continue; }
        v4 p[3];
        for (int k = 0; k < 3; ++k) { p[k].x = cv[k][0]; p[k].y = cv[k][1]; p[k].z = cv[k][2]; p[k].w = cv[k][3]; }

        /* main.c:758-761 */
        if (p[0].w == 0 || p[1].w == 0 || p[2].w == 0) continue;
        /* main.c:764-772 trivial frustum reject */
        if ((p[0].x < -p[0].w && p[1].x < -p[1].w && p[2].x < -p[2].w) ||
            (p[0].x > +p[0].w && p[1].x > +p[1].w && p[2].x > +p[2].w) ||
            (p[0].y < -p[0].w && p[1].y < -p[1].w && p[2].y < -p[2].w) ||
            (p[0].y > +p[0].w && p[1].y > +p[1].w && p[2].y > +p[2].w) ||
            (p[0].z < 0.f && p[1].z < 0.f && p[2].z < 0.f) ||
            (p[0].z > +p[0].w && p[1].z > +p[1].w && p[2].z > +p[2].w))
            continue;
        /* main.c:775-781 */
        int inside = (p[0].x >= -p[0].w && p[1].x >= -p[1].w && p[2].x >= -p[2].w) &&
                     (p[0].x <= +p[0].w && p[1].x <= +p[1].w && p[2].x <= +p[2].w) &&
                     (p[0].y >= -p[0].w && p[1].y >= -p[1].w && p[2].y >= -p[2].w) &&
                     (p[0].y <= +p[0].w && p[1].y <= +p[1].w && p[2].y <= +p[2].w) &&
                     (p[0].z >= 0.f && p[1].z >= 0.f && p[2].z >= 0.f) &&
                     (p[0].z <= +p[0].w && p[1].z <= +p[1].w && p[2].z <= +p[2].w);
        int nv = 3;
        if (!inside) {
            for (int pl = 0; pl < 6; ++pl) {
                nv = clip_by_plane(&cv[0][0], nv, planes[pl]);
                if (nv < 0) return ORC_ERR_CAPACITY;
            }
        }
        /* main.c:797 fan */
        for (int ci = 1; ci < nv - 1; ++ci) {
            const int vid[3] = {0, ci, ci + 1};
            v4 q[3];
            float rw[3];
            for (int k = 0; k < 3; ++k) {
                const float *s = cv[vid[k]];
                q[k].x = s[0]; q[k].y = s[1]; q[k].z = s[2]; q[k].w = s[3];
                /* main.c:803-820: 1.0/w in double, stored f32 (== f32 division) */
                rw[k] = (float)(1.0 / (double)q[k].w);
                q[k].x *= rw[k]; q[k].y *= rw[k]; q[k].z *= rw[k]; q[k].w *= rw[k];
            }
            v4 ss[3];
            int32_t x[3], y[3];
            for (int k = 0; k < 3; ++k) {
                /* main.c:833-839 sequential dot (math.h:159-162) */
                v4 r0 = {M[0][0], M[0][1], M[0][2], M[0][3]}, r1 = {M[1][0], M[1][1], M[1][2], M[1][3]};
                v4 r2 = {M[2][0], M[2][1], M[2][2], M[2][3]}, r3 = {M[3][0], M[3][1], M[3][2], M[3][3]};
                ss[k].x = dot4_seq(r0, q[k]); ss[k].y = dot4_seq(r1, q[k]);
                ss[k].z = dot4_seq(r2, q[k]); ss[k].w = dot4_seq(r3, q[k]);
                /* main.c:842-848 snap to 28.4: floor(x*16 + 0.5) in double */
                x[k] = (int32_t)floor((double)(ss[k].x * (float)(1 << SUBPIX_BITS)) + 0.5);
                y[k] = (int32_t)floor((double)(ss[k].y * (float)(1 << SUBPIX_BITS)) + 0.5);
            }
            /* main.c:851 */
            int32_t area = (int32_t)(((uint32_t)x[1] - (uint32_t)x[0]) * ((uint32_t)y[2] - (uint32_t)y[0]) -
                                     ((uint32_t)x[2] - (uint32_t)x[0]) * ((uint32_t)y[1] - (uint32_t)y[0]));
            /* main.c:856 back-face cull; the other modes are extensions (edge functions are
             * oriented by area's sign, so a kept triangle of either winding is set up alike) */
            if (d->cull_mode == ORC_CULL_BACK ? area > 0 : d->cull_mode == ORC_CULL_FRONT ? area < 0 : 0) continue;
            if (*n >= capacity) return ORC_ERR_CAPACITY;
            orc_triangle scratch_tr;
            orc_triangle *tr = out ? &out[*n] : &scratch_tr;
            (*n)++;
            /* main.c:859-861 edge slots */
            set_edge(tr->edge[2], area, x[0], y[0], x[1], y[1]);
            set_edge(tr->edge[0], area, x[1], y[1], x[2], y[2]);
            set_edge(tr->edge[1], area, x[2], y[2], x[0], y[0]);
            /* extension, D3D top-left rule: an edge whose inward normal (a, b) points right (a > 0:
             * a left edge) or straight down (a == 0, b > 0: a top edge; screen y grows down) keeps
             * its corners; any other edge excludes the corners that lie on it */
            tr->excl = 0u;
            if (d->fill_rule == ORC_FILL_TOP_LEFT)
                for (int k = 0; k < 3; ++k) {
                    const int32_t a = tr->edge[k][0], b = tr->edge[k][1];
                    if (!(a > 0 || (a == 0 && b > 0))) tr->excl |= 1u << k;
                }
            /* main.c:863-866 */
            float af = (float)(area >> (SUBPIX_BITS * 2));
            if (af == 0.0f) af = 1.0f;
            tr->one_over_area = fabsf(1.f / af);
            tr->rw[0] = rw[0]; tr->rw[1] = rw[1]; tr->rw[2] = rw[2];
            /* main.c:871 MAX3 */
            tr->max_depth = MAX_(ss[0].z, MAX_(ss[1].z, ss[2].z));
            tr->prim = t;
            /* main.c:877-883: attributes copied, reg0 replaced by the screen position */
            for (int k = 0; k < 3; ++k) {
                memcpy(tr->attr[k], cv[vid[k]], 12 * sizeof(float));
                tr->attr[k][0] = ss[k].x; tr->attr[k][1] = ss[k].y;
                tr->attr[k][2] = ss[k].z; tr->attr[k][3] = ss[k].w;
            }
            /* main.c:888-898 bounds */
            int32_t mnx = MIN_(x[0], MIN_(x[1], x[2])) >> SUBPIX_BITS;
            int32_t mny = MIN_(y[0], MIN_(y[1], y[2])) >> SUBPIX_BITS;
            mnx = MIN_(MAX_(mnx, 0), W - 1);
            mny = MIN_(MAX_(mny, 0), H - 1);
            int32_t mxx = MAX_(x[0], MAX_(x[1], x[2])) >> SUBPIX_BITS;
            int32_t mxy = MAX_(y[0], MAX_(y[1], y[2])) >> SUBPIX_BITS;
            mxx = MIN_(mxx + 1, W - 1);
            mxy = MIN_(mxy + 1, H - 1);
            tr->min_bounds[0] = mnx; tr->min_bounds[1] = mny;
            tr->max_bounds[0] = mxx; tr->max_bounds[1] = mxy;
        }
    }
    return ORC_OK;
}

int orc_primitive_assembly(const orc_draw_desc *d, orc_triangle *out, uint32_t capacity,
                           uint32_t *out_count, uint64_t *oob_count) {
    if (!d || !out_count || d->index_count % 3u != 0u || !is_valid_vs_ps(d)) return ORC_ERR_INVALID_ARG;
    const uint32_t T = d->index_count / 3u;
    const orc_viewport vp = d->viewport;
    float planes[6][4];
    orc_clip_planes(planes);
    /* main.c:826-831 screen_from_ndc, rows stored as f32 */
    const float M[4][4] = {
        {(float)((double)vp.width * 0.5), 0.0f, 0.0f, (float)((double)vp.width * 0.5 + (double)vp.top_left_x)},
        {0.0f, (float)(-(double)vp.height * 0.5), 0.0f, (float)((double)vp.height * 0.5 + (double)vp.top_left_y)},
        {0.0f, 0.0f, vp.max_depth - vp.min_depth, vp.min_depth},
        {0.0f, 0.0f, 0.0f, 1.0f}};
    const int32_t W = (int32_t)vp.width, H = (int32_t)vp.height;
    /* Chunks of input triangles are assembled in parallel twice: once to count each chunk's
     * output, then again into its place after an exclusive scan, so the output is in input
     * order (the single-thread order of main.c:874-875) for any thread count. */
    const uint32_t CH = 4096u, nch = (T + CH - 1u) / CH;
    uint32_t *cnt = (uint32_t *)calloc((size_t)nch + 1u, sizeof(uint32_t));
    if (!cnt) return ORC_ERR_NOMEM;
    uint64_t oob = 0;
    int err = ORC_OK;
#pragma omp parallel for schedule(dynamic) reduction(+ : oob)
    for (int64_t c = 0; c < (int64_t)nch; ++c) {
        const uint32_t t0 = (uint32_t)c * CH, t1 = t0 + CH < T ? t0 + CH : T;
        uint32_t n = 0;
        uint64_t o = 0;
        const int rc = pa_range(d, planes, M, W, H, t0, t1, NULL, UINT32_MAX, &n, &o);
        if (rc) {
#pragma omp critical
            err = rc;
        }
        cnt[c] = n;
        oob += o;
    }
    uint64_t total = 0;
    for (uint32_t c = 0; c < nch; ++c) {
        const uint32_t k = cnt[c];
        cnt[c] = (uint32_t)total;
        total += k;
    }
    if (!err && total > capacity) err = ORC_ERR_CAPACITY;
    if (!err && out) {
#pragma omp parallel for schedule(dynamic)
        for (int64_t c = 0; c < (int64_t)nch; ++c) {
            const uint32_t t0 = (uint32_t)c * CH, t1 = t0 + CH < T ? t0 + CH : T;
            uint32_t n = 0;
            uint64_t o = 0;
            pa_range(d, planes, M, W, H, t0, t1, out + cnt[c], capacity - cnt[c], &n, &o);
        }
    }
    free(cnt);
    if (err) return err;
    *out_count = (uint32_t)total;
    if (oob_count) *oob_count = oob;
    return ORC_OK;
}

int orc_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------- rasterisation */

static inline int32_t edge_at(const int32_t e[3], int32_t x, int32_t y);

/* main.c:1083-1160 at pixel corner (x, y): quantised barycentrics, perspective correction,
 * register 0 affine and the others perspective-correct. */
static void interp_fragment(const orc_triangle *tr, int32_t x, int32_t y, float frag[12]) {
    int32_t e1 = edge_at(tr->edge[1], x, y) >> (SUBPIX_BITS * 2);
    int32_t e2 = edge_at(tr->edge[2], x, y) >> (SUBPIX_BITS * 2);
    float bx = (float)e1 * tr->one_over_area;
    float by = (float)e2 * tr->one_over_area;
    float den = 1.0f - (bx + by);
    den = den * tr->rw[0];
    den = den + bx * tr->rw[1];
    den = den + by * tr->rw[2];
    den = 1.0f / den;
    float pbx = (bx * tr->rw[1]) * den;
    float pby = (by * tr->rw[2]) * den;
    for (int reg = 0; reg < NREG; ++reg) {
        float u = reg == 0 ? bx : pbx, w = reg == 0 ? by : pby;
        for (int c = 0; c < 4; ++c) {
            float v0 = tr->attr[0][4 * reg + c], v1 = tr->attr[1][4 * reg + c], v2 = tr->attr[2][4 * reg + c];
            float tv0 = v0 + (v1 - v0) * u;
            frag[4 * reg + c] = tv0 + (v2 - v0) * w;
        }
    }
}

static inline int32_t edge_at(const int32_t e[3], int32_t x, int32_t y) {
    /* main.c:1015-1025: ((a*x)<<4) + ((b*y)<<4) + c in i32 wrap */
    uint32_t v = ((uint32_t)e[0] * (uint32_t)x) << SUBPIX_BITS;
    v += ((uint32_t)e[1] * (uint32_t)y) << SUBPIX_BITS;
    v += (uint32_t)e[2];
    return (int32_t)v;
}

/* main.c:1012-1041: 8x8 corner-sampled coverage, bit row*8+col; covered = (E0|E1|E2) > 0. */
uint64_t orc_coverage_mask_excl(const int32_t e[3][3], uint32_t excl, int32_t ox, int32_t oy) {
    uint64_t m = 0;
    for (int r = 0; r < 8; ++r)
        for (int c = 0; c < 8; ++c) {
            int32_t x = ox + c, y = oy + r;
            int32_t a = edge_at(e[0], x, y), b = edge_at(e[1], x, y), g = edge_at(e[2], x, y);
            int cov = (int32_t)((uint32_t)a | (uint32_t)b | (uint32_t)g) > 0;
            if (((excl & 1u) && a == 0) || ((excl & 2u) && b == 0) || ((excl & 4u) && g == 0)) cov = 0;
            if (cov) m |= 1ull << (r * 8 + c);
        }
    return m;
}

uint64_t orc_coverage_mask(const int32_t e[3][3], int32_t ox, int32_t oy) {
    return orc_coverage_mask_excl(e, 0u, ox, oy);
}

int orc_clear_render_target_view(orc_target *rt, const float rgba[4]) {
    if (!rt || !rt->colors) return ORC_ERR_INVALID_ARG;
    uint32_t v = orc_encode_clear(rgba); /* main.c:1191-1202 */
    size_t n = (size_t)rt->width * rt->height;
    for (size_t i = 0; i < n; ++i) rt->colors[i] = v;
    return ORC_OK;
}

int orc_clear_depth_stencil_view(orc_target *rt, float depth, int32_t depth_func) {
    if (!rt || !rt->depth || !rt->hiz) return ORC_ERR_INVALID_ARG;
    size_t n = (size_t)rt->width * rt->height;
    for (size_t i = 0; i < n; ++i) rt->depth[i] = depth; /* main.c:1204-1210 */
    size_t nb = (size_t)(rt->width / BIN) * (rt->height / BIN);
    /* main.c:1212-1214 resets every tile minimum to 0.0; the LESS_EQUAL mirror resets to 1.0 */
    float h = depth_func == ORC_DEPTH_LESS_EQUAL ? 1.0f : 0.0f;
    for (size_t i = 0; i < nb; ++i) rt->hiz[i] = h;
    if (rt->vis_tri)
        for (size_t i = 0; i < n; ++i) rt->vis_tri[i] = 0xFFFFFFFFu;
    if (rt->vis_draw)
        for (size_t i = 0; i < n; ++i) rt->vis_draw[i] = 0xFFFFFFFFu;
    return ORC_OK;
}

/* main.c:1219-1261 draw_indexed: IA, VS, PA, binner (main.c:915-981), rasterizer
 * (main.c:983-1044) and pixel shader stage (main.c:1046-1189). The rasterizer
 * and pixel stage are fused per bin; hiz[b] is read before it is rewritten for b
 * only, so the result equals the two-pass order. */
int orc_draw_indexed(orc_target *rt, const orc_draw_desc *d, orc_stats *st) {
    if (!rt || !d || !rt->colors || !rt->depth || !rt->hiz) return ORC_ERR_INVALID_ARG;
    if (rt->width % BIN || rt->height % BIN) return ORC_ERR_INVALID_ARG;
    if ((uint32_t)d->viewport.width != rt->width || (uint32_t)d->viewport.height != rt->height)
        return ORC_ERR_INVALID_ARG;
    if (d->index_count % 3u != 0u || !is_valid_vs_ps(d)) return ORC_ERR_INVALID_ARG;
    const int le = d->depth_func == ORC_DEPTH_LESS_EQUAL;
    const uint32_t T = d->index_count / 3u;
    const uint32_t cap = T + MAX_(2u * T, 512u); /* main.c:739-740 */
    orc_triangle *tris = (orc_triangle *)malloc(sizeof(orc_triangle) * (size_t)cap);
    if (!tris) return ORC_ERR_NOMEM;
    uint32_t n = 0;
    uint64_t oob = 0;
    int rc = orc_primitive_assembly(d, tris, cap, &n, &oob);
    if (rc) { free(tris); return rc; }

    const int32_t WT = (int32_t)rt->width / BIN, HT = (int32_t)rt->height / BIN;
    const size_t NB = (size_t)WT * HT;
    uint32_t *cnt = (uint32_t *)calloc(NB + 1, sizeof(uint32_t));
    uint32_t *start = (uint32_t *)calloc(NB + 1, sizeof(uint32_t));
    if (!cnt || !start) { free(tris); free(cnt); free(start); return ORC_ERR_NOMEM; }
    /* main.c:924-936 count; minimum tile index MAX(min/8,0), maximum MIN(max/8, WT-1) */
    for (uint32_t t = 0; t < n; ++t) {
        const orc_triangle *tr = &tris[t];
        int32_t bx0 = MAX_(tr->min_bounds[0] / BIN, 0), by0 = MAX_(tr->min_bounds[1] / BIN, 0);
        int32_t bx1 = MIN_(tr->max_bounds[0] / BIN, WT - 1), by1 = MIN_(tr->max_bounds[1] / BIN, HT - 1);
        for (int32_t y = by0; y <= by1; ++y)
            for (int32_t x = bx0; x <= bx1; ++x) cnt[(size_t)y * WT + x]++;
    }
    /* main.c:938-946 exclusive scan */
    uint64_t total = 0, active = 0;
    for (size_t b = 0; b < NB; ++b) { start[b] = (uint32_t)total; total += cnt[b]; active += cnt[b] != 0; }
    uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(total ? total : 1));
    uint32_t *fill = (uint32_t *)calloc(NB + 1, sizeof(uint32_t));
    if (!ids || !fill) { free(tris); free(cnt); free(start); free(ids); free(fill); return ORC_ERR_NOMEM; }
    /* main.c:950-962 order-preserving scatter */
    for (uint32_t t = 0; t < n; ++t) {
        const orc_triangle *tr = &tris[t];
        int32_t bx0 = MAX_(tr->min_bounds[0] / BIN, 0), by0 = MAX_(tr->min_bounds[1] / BIN, 0);
        int32_t bx1 = MIN_(tr->max_bounds[0] / BIN, WT - 1), by1 = MIN_(tr->max_bounds[1] / BIN, HT - 1);
        for (int32_t y = by0; y <= by1; ++y)
            for (int32_t x = bx0; x <= bx1; ++x) {
                size_t b = (size_t)y * WT + x;
                ids[start[b] + fill[b]++] = t;
            }
    }

    const uint32_t W = rt->width;
    /* bins of one draw are independent (each reads and writes only its own 64 pixels) */
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t bi = 0; bi < (int64_t)NB; ++bi) {
        const size_t b = (size_t)bi;
        if (!cnt[b]) continue;
        const int32_t ox = BIN * (int32_t)(b % WT), oy = BIN * (int32_t)(b / WT);
        uint32_t tc[64], tv[64], td[64];
        float tz[64];
        /* main.c:577-587 read_tile */
        for (int j = 0; j < 8; ++j)
            for (int i = 0; i < 8; ++i) {
                size_t p = (size_t)(oy + j) * W + (size_t)(ox + i);
                tc[j * 8 + i] = rt->colors[p];
                tz[j * 8 + i] = rt->depth[p];
                tv[j * 8 + i] = rt->vis_tri ? rt->vis_tri[p] : 0u;
                td[j * 8 + i] = rt->vis_draw ? rt->vis_draw[p] : 0u;
            }
        const float hz = rt->hiz[b]; /* main.c:996 get_tile_minimum_depth */
        for (uint32_t k = 0; k < cnt[b]; ++k) {
            const uint32_t tid = ids[start[b] + k];
            const orc_triangle *tr = &tris[tid];
            /* main.c:1005-1010 hierarchical-Z (LESS_EQUAL mirrors it with the minimum depth) */
            if (!le) {
                if (tr->max_depth < hz) continue;
            } else {
                float mn = MIN_(tr->attr[0][2], MIN_(tr->attr[1][2], tr->attr[2][2]));
                if (mn > hz) continue;
            }
            uint64_t mask = orc_coverage_mask_excl(tr->edge, tr->excl, ox, oy);
            if (!mask) continue;
            for (int r = 0; r < 8; ++r) {
                uint32_t m8 = (uint32_t)(mask >> (8 * r)) & 0xFFu;
                if (!m8) continue;
                const int32_t y = oy + r;
                float frag[8][12];
                int pass[8];
                int any = 0;
                for (int l = 0; l < 8; ++l) {
                    const int32_t x = ox + l;
                    interp_fragment(tr, x, y, frag[l]);
                    /* main.c:1164-1168 early-Z, GE ordered (LE mirror) */
                    float z = frag[l][2];
                    int ok = le ? (z <= tz[r * 8 + l]) : (z >= tz[r * 8 + l]);
                    pass[l] = ((m8 >> l) & 1u) && ok;
                    any |= pass[l];
                }
                if (!any) continue;
                for (int l = 0; l < 8; ++l) {
                    if (!pass[l]) continue;
                    float rgb[3], fx[12], fy[12];
                    if (d->ps == ORC_PS_BASIC_LAMBERT || d->ps == ORC_PS_ENV_TEXTURED) {
                        interp_fragment(tr, ox + l + 1, y, fx);
                        interp_fragment(tr, ox + l, y + 1, fy);
                    }
                    if (d->ps != ORC_PS_NONE) { /* the depth-only extension writes no colour */
                        pixel_shader(d, frag[l], fx, fy, rgb); /* main.c:1172 */
                        tc[r * 8 + l] = orc_encode_output(rgb);
                    }
                    tz[r * 8 + l] = frag[l][2];
                    tv[r * 8 + l] = tid;
                    td[r * 8 + l] = rt->draw_counter;
                }
            }
        }
        /* main.c:589-603 write_tile: tile minimum folded from 1.0 in pixel order */
        float m = le ? 0.0f : 1.0f;
        for (int j = 0; j < 8; ++j)
            for (int i = 0; i < 8; ++i) {
                size_t p = (size_t)(oy + j) * W + (size_t)(ox + i);
                float z = tz[j * 8 + i];
                rt->colors[p] = tc[j * 8 + i];
                rt->depth[p] = z;
                if (rt->vis_tri) rt->vis_tri[p] = tv[j * 8 + i];
                if (rt->vis_draw) rt->vis_draw[p] = td[j * 8 + i];
                m = le ? MAX_(m, z) : MIN_(m, z);
            }
        rt->hiz[b] = m;
    }
    if (st) {
        st->vertex_count += d->index_count;
        st->input_triangle_count += T;
        st->assembled_triangle_count += n;
        st->active_bin_count += active;
        st->total_triangle_count_in_bins += total;
        st->oob_triangle_count += oob;
    }
    rt->draw_counter++;
    free(tris); free(cnt); free(start); free(ids); free(fill);
    return ORC_OK;
}
