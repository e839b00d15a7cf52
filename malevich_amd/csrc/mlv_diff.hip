// mlv_diff.hip -- differentiable shading over the raster's primitive ids (gfx950).
//
// Forward: one lane per pixel. The visible input triangle (from the primitive-id buffer the
// tile kernel writes) is re-projected with the first constant-buffer matrix, exactly as
// basic_vs computes clip positions (basic_vs.c:22-35, pairwise dot math.h:142-146).
// Perspective-correct barycentrics come from the homogeneous 2D form
//   a_k = X_k - u W_k,  b_k = Y_k - v W_k,  l = a x b,  lambda = l / sum(l)
// at the pixel's NDC (u, v): the point of the triangle's plane seen through that pixel. For
// the reference's own interpolants this is the continuous limit of its quantised
// perspective correction (main.c:1106-1154); unlike those, it is differentiable in the vertex
// positions.
//
// Backward: the same lane recomputes the pixel (only the primitive ids survive the forward
// pass) and applies the chain rule in reverse: sRGB curve, SH light, bilinear filter,
// normalisation, interpolation, barycentrics, projection. The recompute and the adjoint run
// in FP64: position gradients of small triangles are sums of large, nearly cancelling
// per-pixel terms, and f32 terms would lose about three digits there. The mip level is the
// forward pass's f32 choice. Per-vertex and per-texel sums either use atomics (FP64 for
// positions, FP32 for the rest) or, in the deterministic mode, are emitted as records, sorted
// by target (stable radix sort) and summed per segment in pixel order.
#include <algorithm>

#include "mlv_device.h"

namespace mlv {

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
// real spherical harmonics of bands 0..2 (f32 constants; the FP64 path widens the same values)
constexpr float kY0 = 0.282094792f, kY1 = 0.488602512f, kY2 = 1.092548431f, kY3 = 0.315391565f,
                kY4 = 0.546274215f;
constexpr float kSrgbP = 0.416666667f;

__device__ __forceinline__ float r_pow(float x, float y) { return powf(x, y); }
__device__ __forceinline__ double r_pow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float r_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double r_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float r_floor(float x) { return floorf(x); }
__device__ __forceinline__ double r_floor(double x) { return floor(x); }

// basic_ps's curve (math.h:419-422) on max(c, 0): f32 in the reference's operation order,
// FP64 for the adjoint's recompute
__device__ __forceinline__ float srgb_of(float c) { return c > 0.0f ? srgb_approx(c) : 0.0f; }
__device__ __forceinline__ double srgb_of(double c) {
    if (!(c > 0.0)) return 0.0;
    const double s = 1.055 * pow(c, (double)kSrgbP) - 0.055;
    return s > 0.0 ? s : 0.0;
}

template <class R>
__device__ __forceinline__ void sh_basis(R x, R y, R z, R Y[9]) {
    Y[0] = (R)kY0;
    Y[1] = (R)kY1 * y;
    Y[2] = (R)kY1 * z;
    Y[3] = (R)kY1 * x;
    Y[4] = (R)kY2 * (x * y);
    Y[5] = (R)kY2 * (y * z);
    Y[6] = (R)kY3 * ((R)3 * (z * z) - (R)1);
    Y[7] = (R)kY2 * (x * z);
    Y[8] = (R)kY4 * (x * x - y * y);
}

template <class R>
struct Clip {
    R X[3], Y[3], W[3];
};

// A triangle's three corners as the shading reads them: loaded once, each vertex's position,
// colour and normal as one 12-byte load and its uv as one 8-byte load (the attribute arrays are
// tightly packed float triples and pairs, 4-byte aligned), so a pixel issues 12 vector loads for
// its triangle instead of 33 single-float gathers. Positions and uvs come first (they place the
// texel taps); colours and normals are loaded with the taps, so they hold no registers while the
// barycentrics and the mip level are computed.
struct Corners {
    float p[3][3], c[3][3], uv[3][2], n[3][3];
};
typedef const __attribute__((address_space(1))) float *gfloat;
typedef const __attribute__((address_space(1))) uint32_t *guint;

// n consecutive floats through a global pointer: adjacent 4-byte loads, merged by the compiler
// into one 8- or 12-byte load
template <int N>
__device__ __forceinline__ void ld_floats(const float *q, float (&o)[N]) {
    const gfloat g = (gfloat)q;
#pragma unroll
    for (int j = 0; j < N; ++j) o[j] = g[j];
}

__device__ __forceinline__ void load_placement(const DiffMesh &m, const uint32_t vi[3], Corners &q) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ld_floats<3>(m.pos + 3u * vi[k], q.p[k]);
        ld_floats<2>(m.uv + 2u * vi[k], q.uv[k]);
    }
}
__device__ __forceinline__ void load_surface(const DiffMesh &m, const uint32_t vi[3], Corners &q) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ld_floats<3>(m.col + 3u * vi[k], q.c[k]);
        ld_floats<3>(m.nrm + 3u * vi[k], q.n[k]);
    }
}

__device__ __forceinline__ void clip_of(const Corners &q, const DiffView &v, Clip<float> &c) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float px = q.p[k][0], py = q.p[k][1], pz = q.p[k][2];
        c.X[k] = dot4_pair(v.M + 0, px, py, pz, 1.0f);
        c.Y[k] = dot4_pair(v.M + 4, px, py, pz, 1.0f);
        c.W[k] = dot4_pair(v.M + 12, px, py, pz, 1.0f);
    }
}

__device__ __forceinline__ void clip_of(const Corners &q, const DiffView &v, Clip<double> &c) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double px = q.p[k][0], py = q.p[k][1], pz = q.p[k][2];
        const float *M = v.M;
        c.X[k] = (double)M[0] * px + (double)M[1] * py + (double)M[2] * pz + (double)M[3];
        c.Y[k] = (double)M[4] * px + (double)M[5] * py + (double)M[6] * pz + (double)M[7];
        c.W[k] = (double)M[12] * px + (double)M[13] * py + (double)M[14] * pz + (double)M[15];
    }
}

// pixel corner (x, y) in NDC, the inverse of screen_from_ndc (main.c:826-831)
__device__ __forceinline__ float ndc_u(float x, uint32_t W) { return (2.0f * x) / (float)W - 1.0f; }
__device__ __forceinline__ float ndc_v(float y, uint32_t H) { return 1.0f - (2.0f * y) / (float)H; }

template <class R>
struct Bary {
    R a[3], b[3], l[3], S, lam[3];
};

template <class R>
__device__ __forceinline__ void bary(const Clip<R> &c, R u, R v, Bary<R> &B) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        B.a[k] = c.X[k] - u * c.W[k];
        B.b[k] = c.Y[k] - v * c.W[k];
    }
    B.l[0] = B.a[1] * B.b[2] - B.a[2] * B.b[1];
    B.l[1] = B.a[2] * B.b[0] - B.a[0] * B.b[2];
    B.l[2] = B.a[0] * B.b[1] - B.a[1] * B.b[0];
    B.S = (B.l[0] + B.l[1]) + B.l[2];
    const R inv = (R)1 / B.S;  // one division, three products
#pragma unroll
    for (int k = 0; k < 3; ++k) B.lam[k] = B.l[k] * inv;
}

// Whether the f32 barycentrics of B are accurate to a few ulp of 1: the cross products' terms,
// summed in magnitude, stay within 64 times |S| (inside a triangle seen at a normal angle they sum
// to about |S|). Slivers seen edge-on and corners far outside the float triangle fail it (the
// terms nearly cancel), and so do NaN and infinities.
__device__ __forceinline__ bool well_conditioned(const Bary<float> &B) {
    const float mag = (fabsf(B.a[1] * B.b[2]) + fabsf(B.a[2] * B.b[1])) +
                      (fabsf(B.a[2] * B.b[0]) + fabsf(B.a[0] * B.b[2])) +
                      (fabsf(B.a[0] * B.b[1]) + fabsf(B.a[1] * B.b[0]));
    return mag < 64.0f * fabsf(B.S);
}

__device__ __forceinline__ void uv_at(const Corners &q, const float lam[3], float &u, float &v) {
    u = (lam[0] * q.uv[0][0] + lam[1] * q.uv[1][0]) + lam[2] * q.uv[2][0];
    v = (lam[0] * q.uv[0][1] + lam[1] * q.uv[1][1]) + lam[2] * q.uv[2][1];
}

// Mip level nearest the pixel's footprint, in f32: the uv differences to the next pixel right
// and below, in level-0 texels; level k is taken when rho >= 2^(k - 1/2), i.e.
// rho^2 >= 2^(2k-1), which needs no logarithm.
// b0: the barycentrics at the pixel corner itself.
__device__ __forceinline__ uint32_t mip_level(const DiffMesh &m, const DiffView &v, const Corners &q, const Clip<float> &c,
                                              const Bary<float> &b0, float x, float y) {
    if (m.levels <= 1u) return 0u;
    Bary<float> bx, by;
    bary(c, ndc_u(x + 1.0f, v.W), ndc_v(y, v.H), bx);
    bary(c, ndc_u(x, v.W), ndc_v(y + 1.0f, v.H), by);
    float u0, v0, ux, vx, uy, vy;
    uv_at(q, b0.lam, u0, v0);
    uv_at(q, bx.lam, ux, vx);
    uv_at(q, by.lam, uy, vy);
    const float tw = (float)m.lw[0], th = (float)m.lh[0];
    const float dux = (ux - u0) * tw, dvx = (vx - v0) * th, duy = (uy - u0) * tw, dvy = (vy - v0) * th;
    const float rx = dux * dux + dvx * dvx, ry = duy * duy + dvy * dvy;
    const float r2 = rx > ry ? rx : ry;
    uint32_t lvl = 0u;
    float thr = 2.0f;  // 2^(2k-1) for k = 1
    for (uint32_t k = 1; k < m.levels; ++k) {
        if (r2 >= thr) lvl = k;
        thr *= 4.0f;
    }
    return lvl;
}

// Everything one pixel computes, kept for the backward pass.
template <class R>
struct Eval {
    uint32_t vi[3];
    Bary<R> B;
    R u, v;                    // NDC of the pixel corner
    R col[3], uvs[2], n[3];
    uint32_t lvl;
    int32_t x0, y0, x1, y1;    // clamped taps
    R fs, ft;
    float t00[3], t10[3], t01[3], t11[3];  // the taps' texels (RGBA32F: exact in either precision)
    R alb[3];
    R nh[3], inv_len;
    R Y[9], E[3], lin[3], out[3];
};

template <class R>
__device__ __forceinline__ R lerp_r(R a, R b, R f) {
    return a * ((R)1 - f) + b * f;
}

// The texture filter, the SH light and the curve, from the interpolated attributes in e.
template <class R>
__device__ __forceinline__ bool shade_tail(const DiffMesh &m, const float *sh, Eval<R> &e) {
    // bilinear_f_x8 (csc.h:143-161) on the chosen level
    const uint32_t lw = m.lw[e.lvl], lh = m.lh[e.lvl];
    const R s = (R)lw * e.uvs[0] - (R)0.5;
    const R tt = (R)lh * ((R)1 - e.uvs[1]) - (R)0.5;
    const R fx = r_floor(s), fy = r_floor(tt);
    e.fs = s - fx;
    e.ft = tt - fy;
    const int32_t ix = (fx > (R)-1e9 && fx < (R)1e9) ? (int32_t)fx : 0;
    const int32_t iy = (fy > (R)-1e9 && fy < (R)1e9) ? (int32_t)fy : 0;
    const int32_t wm = (int32_t)lw - 1, hm = (int32_t)lh - 1;
    e.x0 = min(max(ix, 0), wm);
    e.x1 = min(max(ix + 1, 0), wm);
    e.y0 = min(max(iy, 0), hm);
    e.y1 = min(max(iy + 1, 0), hm);
    const float4 *T = m.tex + m.loff[e.lvl];
    if constexpr (sizeof(R) == 4) {
        if (m.tex_mark) {  // footprint counting (forward passes only): the four taps' texels
            const uint32_t b0 = m.loff[e.lvl];
            const uint32_t taps[4] = {b0 + (uint32_t)e.y0 * lw + (uint32_t)e.x0, b0 + (uint32_t)e.y0 * lw + (uint32_t)e.x1,
                                      b0 + (uint32_t)e.y1 * lw + (uint32_t)e.x0, b0 + (uint32_t)e.y1 * lw + (uint32_t)e.x1};
#pragma unroll
            for (int k = 0; k < 4; ++k) atomicOr(m.tex_mark + (taps[k] >> 5), 1u << (taps[k] & 31u));
        }
    }
    const float4 q00 = load_global(T + (uint32_t)e.y0 * lw + (uint32_t)e.x0),
                 q10 = load_global(T + (uint32_t)e.y0 * lw + (uint32_t)e.x1),
                 q01 = load_global(T + (uint32_t)e.y1 * lw + (uint32_t)e.x0),
                 q11 = load_global(T + (uint32_t)e.y1 * lw + (uint32_t)e.x1);
    e.t00[0] = q00.x; e.t00[1] = q00.y; e.t00[2] = q00.z;
    e.t10[0] = q10.x; e.t10[1] = q10.y; e.t10[2] = q10.z;
    e.t01[0] = q01.x; e.t01[1] = q01.y; e.t01[2] = q01.z;
    e.t11[0] = q11.x; e.t11[1] = q11.y; e.t11[2] = q11.z;
#pragma unroll
    for (int j = 0; j < 3; ++j)
        e.alb[j] = lerp_r(lerp_r((R)e.t00[j], (R)e.t10[j], e.fs), lerp_r((R)e.t01[j], (R)e.t11[j], e.fs), e.ft);
    const R len = r_sqrt((e.n[0] * e.n[0] + e.n[1] * e.n[1]) + e.n[2] * e.n[2]);
    e.inv_len = (R)1 / len;
    e.nh[0] = e.n[0] * e.inv_len;
    e.nh[1] = e.n[1] * e.inv_len;
    e.nh[2] = e.n[2] * e.inv_len;
    sh_basis(e.nh[0], e.nh[1], e.nh[2], e.Y);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        R acc = (R)0;
#pragma unroll
        for (int k = 0; k < 9; ++k) acc = acc + (R)sh[3 * k + j] * e.Y[k];
        e.E[j] = acc;
        e.lin[j] = (e.alb[j] * e.col[j]) * e.E[j];
        e.out[j] = srgb_of(e.lin[j]);
    }
    return true;
}

// 0: nothing to shade (invalid triangle), 1: shaded into e. With kDefer (the f32 forward), a pixel
// whose f32 barycentrics are ill-conditioned returns kDeferred instead of running the FP64 path.
constexpr int kDeferred = 2;
template <class R, bool kDefer = false>
__device__ __forceinline__ int evaluate(const DiffMesh &m, const DiffView &v, const float *sh, uint32_t t, uint32_t x,
                                        uint32_t y, Eval<R> &e) {
    if (t >= m.T) return 0;
    const guint tri = (guint)(m.idx + 3u * t);
    e.vi[0] = tri[0];
    e.vi[1] = tri[1];
    e.vi[2] = tri[2];
    if (e.vi[0] >= m.V || e.vi[1] >= m.V || e.vi[2] >= m.V) return 0;
    Corners q;
    load_placement(m, e.vi, q);
    Clip<float> cf;
    clip_of(q, v, cf);
    Bary<float> bf;
    bary(cf, ndc_u((float)x, v.W), ndc_v((float)y, v.H), bf);
    e.lvl = mip_level(m, v, q, cf, bf, (float)x, (float)y);
    if constexpr (sizeof(R) == 4) {
        // forward: the f32 barycentrics where they are accurate (nearly every pixel)
        if (well_conditioned(bf)) {
            load_surface(m, e.vi, q);
            e.u = ndc_u((float)x, v.W);
            e.v = ndc_v((float)y, v.H);
            e.B = bf;
            const float *L = bf.lam;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                e.col[j] = (L[0] * q.c[0][j] + L[1] * q.c[1][j]) + L[2] * q.c[2][j];
                e.n[j] = (L[0] * q.n[0][j] + L[1] * q.n[1][j]) + L[2] * q.n[2][j];
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) e.uvs[j] = (L[0] * q.uv[0][j] + L[1] * q.uv[1][j]) + L[2] * q.uv[2][j];
            return shade_tail(m, sh, e) ? 1 : 0;
        }
        if constexpr (kDefer) return kDeferred;
    }
    // Otherwise (and always for the adjoint's recompute) the pixel corner's NDC, the barycentrics
    // and the interpolated attributes in FP64: the homogeneous form is ill-conditioned for slivers
    // seen edge-on (the sum of the three cross products nearly cancels), where f32 barycentrics
    // drift far from the mathematical value. Only the texture filter and the curves run in R.
    load_surface(m, e.vi, q);
    Clip<double> c;
    clip_of(q, v, c);
    const double ud = (2.0 * (double)x) / (double)v.W - 1.0, vd = 1.0 - (2.0 * (double)y) / (double)v.H;
    Bary<double> B;
    bary(c, ud, vd, B);
    e.u = (R)ud;
    e.v = (R)vd;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e.B.a[k] = (R)B.a[k];
        e.B.b[k] = (R)B.b[k];
        e.B.l[k] = (R)B.l[k];
        e.B.lam[k] = (R)B.lam[k];
    }
    e.B.S = (R)B.S;
    const double *L = B.lam;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        e.col[j] = (R)((L[0] * (double)q.c[0][j] + L[1] * (double)q.c[1][j]) + L[2] * (double)q.c[2][j]);
        e.n[j] = (R)((L[0] * (double)q.n[0][j] + L[1] * (double)q.n[1][j]) + L[2] * (double)q.n[2][j]);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
        e.uvs[j] = (R)((L[0] * (double)q.uv[0][j] + L[1] * (double)q.uv[1][j]) + L[2] * (double)q.uv[2][j]);
    return shade_tail(m, sh, e) ? 1 : 0;
}

__global__ __launch_bounds__(256) void kd_forward(DiffParams p, float4 *out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= p.v.W * p.v.H) return;
    Eval<float> e;
    const uint32_t x = i % p.v.W, y = i / p.v.W;
    out[i] = evaluate(p.m, p.v, p.sh, p.prim[i], x, y, e) ? make_float4(e.out[0], e.out[1], e.out[2], 1.0f)
                                                          : make_float4(0.f, 0.f, 0.f, 0.f);
}

// Per-pixel adjoint (FP64). Outputs: per-corner vertex gradients (dpos, dcol, duv, dn),
// per-tap texel gradients (rgb) and dE, from which the pixel's SH gradient is dE[j] * Y[k] for
// coefficient k, channel j. The gradients that are summed in f32 (all but positions) are kept
// rounded to f32 here, as they would be when summed.
struct Adj {
    double dpos[3][3];
    float dcol[3][3], duv[3][2], dn[3][3];
    float dtex[4][3];
    uint32_t tap[4];
    double dE[3];
};

__device__ __forceinline__ void adjoint(const DiffMesh &m, const DiffView &v, const float *sh, const Eval<double> &e, float4 g,
                                        Adj &a) {
    const double P = (double)kSrgbP;
    double dlin[3];
    const double gg[3] = {g.x, g.y, g.z};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const bool on = e.lin[j] > 0.0 && e.out[j] > 0.0;
        dlin[j] = on ? gg[j] * (1.055 * P) * pow(e.lin[j], P - 1.0) : 0.0;
    }
    double dalb[3], dci[3], dE[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        dalb[j] = dlin[j] * e.col[j] * e.E[j];
        dci[j] = dlin[j] * e.alb[j] * e.E[j];
        dE[j] = dlin[j] * e.alb[j] * e.col[j];
    }
    double dY[9];
    a.dE[0] = dE[0];
    a.dE[1] = dE[1];
    a.dE[2] = dE[2];
#pragma unroll
    for (int k = 0; k < 9; ++k)
        dY[k] = dE[0] * (double)sh[3 * k] + dE[1] * (double)sh[3 * k + 1] + dE[2] * (double)sh[3 * k + 2];
    const double x = e.nh[0], y = e.nh[1], z = e.nh[2];
    const double c1 = kY1, c2 = kY2, c3 = kY3, c4 = kY4;
    double dnh[3];
    dnh[0] = c1 * dY[3] + c2 * (y * dY[4] + z * dY[7]) + 2.0 * c4 * x * dY[8];
    dnh[1] = c1 * dY[1] + c2 * (x * dY[4] + z * dY[5]) - 2.0 * c4 * y * dY[8];
    dnh[2] = c1 * dY[2] + c2 * (y * dY[5] + x * dY[7]) + 6.0 * c3 * z * dY[6];
    const double pr = x * dnh[0] + y * dnh[1] + z * dnh[2];
    const double dni[3] = {(dnh[0] - x * pr) * e.inv_len, (dnh[1] - y * pr) * e.inv_len,
                           (dnh[2] - z * pr) * e.inv_len};
    // bilinear taps and the uv derivative (s = w u - 0.5, t = h (1 - v) - 0.5)
    const double w00 = (1.0 - e.fs) * (1.0 - e.ft), w10 = e.fs * (1.0 - e.ft), w01 = (1.0 - e.fs) * e.ft,
                 w11 = e.fs * e.ft;
    const uint32_t lw = m.lw[e.lvl], base = m.loff[e.lvl];
    a.tap[0] = base + (uint32_t)e.y0 * lw + (uint32_t)e.x0;
    a.tap[1] = base + (uint32_t)e.y0 * lw + (uint32_t)e.x1;
    a.tap[2] = base + (uint32_t)e.y1 * lw + (uint32_t)e.x0;
    a.tap[3] = base + (uint32_t)e.y1 * lw + (uint32_t)e.x1;
    double ds = 0.0, dt = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        a.dtex[0][j] = (float)(w00 * dalb[j]);
        a.dtex[1][j] = (float)(w10 * dalb[j]);
        a.dtex[2][j] = (float)(w01 * dalb[j]);
        a.dtex[3][j] = (float)(w11 * dalb[j]);
        const double t00 = e.t00[j], t10 = e.t10[j], t01 = e.t01[j], t11 = e.t11[j];
        const double dAs = (1.0 - e.ft) * (t10 - t00) + e.ft * (t11 - t01);
        const double dAt = lerp_r(t01, t11, e.fs) - lerp_r(t00, t10, e.fs);
        ds += dalb[j] * dAs;
        dt += dalb[j] * dAt;
    }
    const double duvi[2] = {ds * (double)lw, -dt * (double)m.lh[e.lvl]};
    // interpolation: attributes to corners and barycentric adjoints
    const double *L = e.B.lam;
    double dlam[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t v = e.vi[k];
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            a.dcol[k][j] = (float)(L[k] * dci[j]);
            a.dn[k][j] = (float)(L[k] * dni[j]);
            acc += dci[j] * (double)m.col[3u * v + j] + dni[j] * (double)m.nrm[3u * v + j];
        }
        a.duv[k][0] = (float)(L[k] * duvi[0]);
        a.duv[k][1] = (float)(L[k] * duvi[1]);
        acc += duvi[0] * (double)m.uv[2u * v] + duvi[1] * (double)m.uv[2u * v + 1u];
        dlam[k] = acc;
    }
    // lambda = l / S
    const double mean = L[0] * dlam[0] + L[1] * dlam[1] + L[2] * dlam[2];
    double dl[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) dl[k] = (dlam[k] - mean) / e.B.S;
    // l = a x b
    const double *A = e.B.a, *Bb = e.B.b;
    double da[3], db[3];
    da[0] = dl[2] * Bb[1] - dl[1] * Bb[2];
    da[1] = dl[0] * Bb[2] - dl[2] * Bb[0];
    da[2] = dl[1] * Bb[0] - dl[0] * Bb[1];
    db[0] = dl[1] * A[2] - dl[2] * A[1];
    db[1] = dl[2] * A[0] - dl[0] * A[2];
    db[2] = dl[0] * A[1] - dl[1] * A[0];
    // a = X - u W, b = Y - v W; clip = M (p, 1)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double dX = da[k], dYc = db[k], dW = -e.u * da[k] - e.v * db[k];
#pragma unroll
        for (int j = 0; j < 3; ++j)
            a.dpos[k][j] = (double)v.M[j] * dX + (double)v.M[4 + j] * dYc + (double)v.M[12 + j] * dW;
    }
}

__device__ __forceinline__ void atomic_add(float *dst, double v) {
    if (v != 0.0) unsafeAtomicAdd(dst, (float)v);
}

__device__ __forceinline__ void atomic_add(double *dst, double v) {
    if (v != 0.0) unsafeAtomicAdd(dst, v);
}

// Sum over the 64 lanes of the wave (every lane active), returned in every lane: DPP steps within
// quads, half-rows and rows, then the row broadcasts leave the total in lane 63.
template <int Ctrl, int RowMask>
__device__ __forceinline__ int dpp_move(int v) {
    return __builtin_amdgcn_update_dpp(0, v, Ctrl, RowMask, 0xF, false);
}
template <int Ctrl, int RowMask>
__device__ __forceinline__ void dpp_add(float &v) {
    v += __int_as_float(dpp_move<Ctrl, RowMask>(__float_as_int(v)));
}
template <int Ctrl, int RowMask>
__device__ __forceinline__ void dpp_add(double &v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_move<Ctrl, RowMask>((int)(b & 0xFFFFFFFFll)), hi = dpp_move<Ctrl, RowMask>((int)(b >> 32));
    v += __longlong_as_double((long long)(((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo));
}
template <class T>
__device__ __forceinline__ T wave_total(T v) {
    dpp_add<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
    dpp_add<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
    dpp_add<0x141, 0xF>(v);  // row_half_mirror
    dpp_add<0x140, 0xF>(v);  // row_mirror
    dpp_add<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    dpp_add<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    if constexpr (sizeof(T) == 8) {
        const long long b = __double_as_longlong(v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(b & 0xFFFFFFFFll), 63);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(b >> 32), 63);
        return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    } else {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
    }
}


// Block sum of the 27 SH gradients (dE[j] * Y[k] of the live pixels, Y the basis at the unit
// normal nh, evaluated again here rather than kept through the adjoint) into
// sh_part[coefficient][block] (fixed order: deterministic; coefficient-major so that the final
// sum reads each coefficient's partials contiguously).
__device__ __forceinline__ void sh_block_sum(bool live, const double dE[3], const double nh[3], double *sh_part) {
    __shared__ double s_w[4][27];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    double Y[9];
    sh_basis(nh[0], nh[1], nh[2], Y);
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const double v = wave_total(live ? dE[k % 3] * Y[k / 3] : 0.0);
        if (lane == 0u) s_w[w][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 27u)
        sh_part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] =
            ((s_w[0][threadIdx.x] + s_w[1][threadIdx.x]) + s_w[2][threadIdx.x]) + s_w[3][threadIdx.x];
}

// ---- per-block aggregation of the atomic backward
//
// A block shades a 16x16 pixel tile. Neighbouring pixels share vertices and texels, so the
// block first sums its contributions per (mesh, vertex) and per (mesh, texel) in LDS hash
// tables (open addressing, keys inserted with compare-and-swap) and then adds each distinct
// entry to global memory once. Device-wide atomics cost far more than LDS ones; a contribution
// that finds no free slot within kProbe probes goes straight to global memory.
constexpr uint32_t kTileSide = 16;
// 256 vertex and 512 texel slots: 27 KB of LDS, four blocks a CU.
constexpr uint32_t kVtxSlots = 256, kTexSlots = 512, kProbe = 16;
constexpr unsigned long long kEmptyKey = ~0ull;

struct VtxTable {
    unsigned long long key[kVtxSlots];
    double pos[3][kVtxSlots];
    float att[8][kVtxSlots];  // colour xyz, normal xyz, uv xy
};
struct TexTable {
    unsigned long long key[kTexSlots];
    float rgb[3][kTexSlots];
};

// Global destinations of one mesh's gradients.
struct GradSink {
    double *posacc;
    float *col, *nrm, *uv, *tex;
};

__device__ __forceinline__ uint32_t slot_hash(unsigned long long k) {
    return (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> 40);
}

// Slot of key k in a table of N slots: its hash slot, or one of the next kProbe - 1 (linear
// probing); -1 when none is free or holds k. first_old is what the compare-and-swap at the hash
// slot returned (gather_pixel issues the first probes of all its keys together).
template <uint32_t N>
__device__ __forceinline__ int lds_slot_after(unsigned long long *keys, unsigned long long k,
                                              unsigned long long first_old) {
    uint32_t s = slot_hash(k) & (N - 1u);
    if (first_old == kEmptyKey || first_old == k) return (int)s;
    for (uint32_t i = 1; i < kProbe; ++i) {
        s = (s + 1u) & (N - 1u);
        const unsigned long long old = atomicCAS(&keys[s], kEmptyKey, k);
        if (old == kEmptyKey || old == k) return (int)s;
    }
    return -1;
}
template <uint32_t N>
__device__ __forceinline__ unsigned long long first_probe(unsigned long long *keys, unsigned long long k) {
    return atomicCAS(&keys[slot_hash(k) & (N - 1u)], kEmptyKey, k);
}
template <uint32_t N>
__device__ __forceinline__ int lds_slot(unsigned long long *keys, unsigned long long k) {
    return lds_slot_after<N>(keys, k, first_probe<N>(keys, k));
}

__device__ __forceinline__ void tables_clear(VtxTable &vt, TexTable &tt) {
    for (uint32_t s = threadIdx.x; s < kVtxSlots; s += 256u) {
        vt.key[s] = kEmptyKey;
        for (int c = 0; c < 3; ++c) vt.pos[c][s] = 0.0;
        for (int c = 0; c < 8; ++c) vt.att[c][s] = 0.0f;
    }
    for (uint32_t s = threadIdx.x; s < kTexSlots; s += 256u) {
        tt.key[s] = kEmptyKey;
        for (int c = 0; c < 3; ++c) tt.rgb[c][s] = 0.0f;
    }
}

__device__ __forceinline__ void vertex_to_global(const GradSink &g, uint32_t v, const double dpos[3],
                                                 const float att[8]) {
    for (int j = 0; j < 3; ++j) {
        if (g.posacc) atomic_add(&g.posacc[3u * v + j], dpos[j]);
        if (g.col) atomic_add(&g.col[3u * v + j], (double)att[j]);
        if (g.nrm) atomic_add(&g.nrm[3u * v + j], (double)att[3 + j]);
    }
    if (g.uv) {
        atomic_add(&g.uv[2u * v], (double)att[6]);
        atomic_add(&g.uv[2u * v + 1u], (double)att[7]);
    }
}

// One vertex's contribution into the block's table slot s (from lds_slot on the key of mesh m,
// vertex v), or into global memory when the table had no slot for it (s < 0).
__device__ __forceinline__ void vertex_to_slot(VtxTable &vt, const GradSink &g, int s, uint32_t v,
                                               const double dpos[3], const float att[8]) {
    if (s < 0) {
        vertex_to_global(g, v, dpos, att);
        return;
    }
    for (int j = 0; j < 3; ++j)
        if (dpos[j] != 0.0) unsafeAtomicAdd(&vt.pos[j][s], dpos[j]);
    for (int j = 0; j < 8; ++j)
        if (att[j] != 0.0f) unsafeAtomicAdd(&vt.att[j][s], att[j]);
}

// Lanes whose pixels show the same triangle of the same mesh (on large triangles, most of a
// wave) add their vertex contributions up across the wave first, so that the triangle's three
// vertices take one set of table atomics instead of one per lane: atomics on one LDS address
// serialise. Up to kMaxGroups triangles of the wave, each shown by at least kMinGroup lanes, are
// taken this way; returns whether this lane's vertex contributions were. Every lane of the wave
// must call it.
constexpr uint32_t kMinGroup = 12;
constexpr int kMaxGroups = 4;

__device__ __forceinline__ bool group_vertices(VtxTable &vt, const GradSink &g, bool live, uint32_t m,
                                               const Eval<double> &e, const Adj &a) {
    uint64_t pend = __ballot(live);
    const uint32_t lane = threadIdx.x & 63u;
    bool done = false;
    for (int it = 0; it < kMaxGroups && pend; ++it) {
        const int lead = (int)__builtin_ctzll(pend);
        const uint32_t km = (uint32_t)__builtin_amdgcn_readlane((int)m, lead);
        const uint32_t k0 = (uint32_t)__builtin_amdgcn_readlane((int)e.vi[0], lead);
        const uint32_t k1 = (uint32_t)__builtin_amdgcn_readlane((int)e.vi[1], lead);
        const uint32_t k2 = (uint32_t)__builtin_amdgcn_readlane((int)e.vi[2], lead);
        const bool mine = live && m == km && e.vi[0] == k0 && e.vi[1] == k1 && e.vi[2] == k2;
        const uint64_t grp = __ballot(mine);
        pend &= ~grp;
        if ((uint32_t)__popcll(grp) < kMinGroup) continue;
        done = done || mine;
        const uint32_t kv[3] = {k0, k1, k2};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double dpos[3];
            float att[8];
#pragma unroll
            for (int j = 0; j < 3; ++j) dpos[j] = wave_total(mine ? a.dpos[k][j] : 0.0);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                att[j] = wave_total(mine ? a.dcol[k][j] : 0.0f);
                att[3 + j] = wave_total(mine ? a.dn[k][j] : 0.0f);
            }
            att[6] = wave_total(mine ? a.duv[k][0] : 0.0f);
            att[7] = wave_total(mine ? a.duv[k][1] : 0.0f);
            if (lane == (uint32_t)lead)
                vertex_to_slot(vt, g, lds_slot<kVtxSlots>(vt.key, ((unsigned long long)km << 32) | kv[k]), kv[k],
                               dpos, att);
        }
    }
    return done;
}

// One pixel's contributions into the block's tables (mesh index m in the keys' high word); the
// vertex contributions only when group_vertices did not take them.
__device__ __forceinline__ void gather_pixel(VtxTable &vt, TexTable &tt, const GradSink &g, uint32_t m,
                                             const Eval<double> &e, const Adj &a, bool vertices_done) {
    // the first probes of the seven keys go out together, so their round trips overlap
    const unsigned long long hi = (unsigned long long)m << 32;
    unsigned long long ov[3], ot[4];
#pragma unroll
    for (int k = 0; k < 3; ++k) ov[k] = vertices_done ? 0ull : first_probe<kVtxSlots>(vt.key, hi | e.vi[k]);
#pragma unroll
    for (int t = 0; t < 4; ++t) ot[t] = g.tex ? first_probe<kTexSlots>(tt.key, hi | a.tap[t]) : 0ull;
    if (!vertices_done) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float att[8] = {a.dcol[k][0], a.dcol[k][1], a.dcol[k][2], a.dn[k][0],
                                  a.dn[k][1],   a.dn[k][2],   a.duv[k][0], a.duv[k][1]};
            vertex_to_slot(vt, g, lds_slot_after<kVtxSlots>(vt.key, hi | e.vi[k], ov[k]), e.vi[k], a.dpos[k], att);
        }
    }
    if (!g.tex) return;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int s = lds_slot_after<kTexSlots>(tt.key, hi | a.tap[t], ot[t]);
        for (int j = 0; j < 3; ++j) {
            if (a.dtex[t][j] == 0.0f) continue;
            if (s < 0) unsafeAtomicAdd(&g.tex[4u * a.tap[t] + j], a.dtex[t][j]);
            else unsafeAtomicAdd(&tt.rgb[j][s], a.dtex[t][j]);
        }
    }
}

// After a barrier: every filled slot adds its sums to global memory once.
template <class SinkOf>
__device__ __forceinline__ void tables_flush(const VtxTable &vt, const TexTable &tt, SinkOf sink_of) {
    for (uint32_t s = threadIdx.x; s < kVtxSlots; s += 256u) {
        const unsigned long long k = vt.key[s];
        if (k == kEmptyKey) continue;
        const GradSink g = sink_of((uint32_t)(k >> 32));
        const double dpos[3] = {vt.pos[0][s], vt.pos[1][s], vt.pos[2][s]};
        float att[8];
        for (int c = 0; c < 8; ++c) att[c] = vt.att[c][s];
        vertex_to_global(g, (uint32_t)k, dpos, att);
    }
    for (uint32_t s = threadIdx.x; s < kTexSlots; s += 256u) {
        const unsigned long long k = tt.key[s];
        if (k == kEmptyKey) continue;
        const GradSink g = sink_of((uint32_t)(k >> 32));
        for (int j = 0; j < 3; ++j) atomic_add(&g.tex[4u * (uint32_t)k + j], (double)tt.rgb[j][s]);
    }
}

// Pixel of thread t in tile `tile` of a W-wide image (16x16 tiles, row-major).
__device__ __forceinline__ void tile_pixel(uint32_t tile, uint32_t W, uint32_t tid, uint32_t &x, uint32_t &y) {
    const uint32_t tx = (W + kTileSide - 1u) / kTileSide;
    x = (tile % tx) * kTileSide + (tid & (kTileSide - 1u));
    y = (tile / tx) * kTileSide + tid / kTileSide;
}
__device__ __forceinline__ void tile_pixel(uint32_t tile, uint32_t W, uint32_t &x, uint32_t &y) {
    tile_pixel(tile, W, threadIdx.x, x, y);
}

__global__ __launch_bounds__(256) void kd_backward_atomic(DiffParams p, const float4 *gout, DiffGrads g,
                                                          double *posacc, double *sh_part) {
    __shared__ VtxTable vt;
    __shared__ TexTable tt;
    tables_clear(vt, tt);
    __syncthreads();
    uint32_t x, y;
    tile_pixel(blockIdx.x, p.v.W, x, y);
    const GradSink sink{posacc, g.col, g.nrm, g.uv, g.tex};
    Eval<double> e;
    Adj a;
    bool live = false;
    if (x < p.v.W && y < p.v.H) {
        const uint32_t i = y * p.v.W + x;
        live = evaluate(p.m, p.v, p.sh, p.prim[i], x, y, e);
        if (live) adjoint(p.m, p.v, p.sh, e, gout[i], a);
    }
    sh_block_sum(live, a.dE, e.nh, sh_part);  // first: dE and nh hold no registers past it
    const bool grouped = group_vertices(vt, sink, live, 0u, e, a);
    if (live) gather_pixel(vt, tt, sink, 0u, e, a, grouped);
    __syncthreads();
    tables_flush(vt, tt, [&](uint32_t) { return sink; });
}

__global__ __launch_bounds__(256) void kd_backward_records(DiffParams p, const float4 *gout, DiffScratch sc) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t P = p.v.W * p.v.H;
    Eval<double> e;
    Adj a;
    bool live = false;
    if (i < P) {
        live = evaluate(p.m, p.v, p.sh, p.prim[i], i % p.v.W, i / p.v.W, e);
        if (live) adjoint(p.m, p.v, p.sh, e, gout[i], a);
    }
    sh_block_sum(live, a.dE, e.nh, sc.sh_part);
    if (i < P) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t r = 4u * i + (uint32_t)t;
            sc.tkey[r] = live ? a.tap[t] : p.m.tex_total;
            sc.tval[r] = r;
            sc.tpay[r] = live ? make_float4(a.dtex[t][0], a.dtex[t][1], a.dtex[t][2], 0.0f)
                              : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t r = 3u * i + (uint32_t)k;
            sc.vkey[r] = live ? e.vi[k] : p.m.V;
            sc.vval[r] = r;
            if (live) {
                sc.vposd[4u * r] = a.dpos[k][0];
                sc.vposd[4u * r + 1u] = a.dpos[k][1];
                sc.vposd[4u * r + 2u] = a.dpos[k][2];
                sc.vpay[2u * r] = make_float4(a.dcol[k][0], a.dcol[k][1], a.dcol[k][2], a.duv[k][0]);
                sc.vpay[2u * r + 1u] = make_float4(a.duv[k][1], a.dn[k][0], a.dn[k][1], a.dn[k][2]);
            }
        }
    }
}

// Sum of the per-block SH partials, one block per coefficient: thread t sums blocks t, t+256, ...
// in order, then a fixed tree over the 256 partial sums (the same order on every call).
__device__ __forceinline__ void sh_sum_block(const double *sh_part, uint32_t nblocks, float *dsh, uint32_t accumulate,
                                             uint32_t k) {
    __shared__ double s[256];
    const uint32_t t = threadIdx.x;
    double acc = 0.0;
    for (uint32_t b = t; b < nblocks; b += 256u) acc += sh_part[(size_t)k * nblocks + b];
    s[t] = acc;
    __syncthreads();
    for (uint32_t h = 128u; h >= 1u; h >>= 1) {
        if (t < h) s[t] = s[t] + s[t + h];
        __syncthreads();
    }
    if (t == 0u) dsh[k] = accumulate ? dsh[k] + (float)s[0] : (float)s[0];
}

// SH coefficient k = blockIdx.x: the sum of its per-block partials in a fixed tree order.
__global__ __launch_bounds__(256) void kd_sh_sum(const double *sh_part, uint32_t nblocks, float *dsh,
                                                 uint32_t accumulate) {
    sh_sum_block(sh_part, nblocks, dsh, accumulate, blockIdx.x);
}

// The atomic backward's tail in one launch: the first n_sh blocks sum the SH partials (as
// kd_sh_sum), the others turn the FP64 position sums into the float gradient.
__global__ __launch_bounds__(256) void kd_backward_tail(const double *sh_part, uint32_t nblocks, float *dsh,
                                                        uint32_t n_sh, const double *acc, float *out, uint32_t n) {
    if (blockIdx.x < n_sh) {
        sh_sum_block(sh_part, nblocks, dsh, 0u, blockIdx.x);
        return;
    }
    const uint32_t i = (blockIdx.x - n_sh) * 256u + threadIdx.x;
    if (i < n) out[i] = (float)acc[i];
}

// Segments of the sorted texel records: the first record of each texel sums its segment in
// pixel order (the sort is stable) and writes the texel's gradient.
__global__ __launch_bounds__(256) void kd_segment_texel(const uint32_t *key, const uint32_t *val, const float4 *pay,
                                                        uint32_t n, uint32_t none, float *dtex) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = key[i];
    if (k >= none || (i > 0u && key[i - 1u] == k)) return;
    double r = 0.0, g = 0.0, b = 0.0;
    for (uint32_t j = i; j < n && key[j] == k; ++j) {
        const float4 v = pay[val[j]];
        r += v.x;
        g += v.y;
        b += v.z;
    }
    dtex[4u * k] = (float)r;
    dtex[4u * k + 1u] = (float)g;
    dtex[4u * k + 2u] = (float)b;
}

__global__ __launch_bounds__(256) void kd_segment_vertex(const uint32_t *key, const uint32_t *val, const float4 *pay,
                                                         const double *posd, uint32_t n, uint32_t none, DiffGrads g) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = key[i];
    if (k >= none || (i > 0u && key[i - 1u] == k)) return;
    double s[11] = {};
    for (uint32_t j = i; j < n && key[j] == k; ++j) {
        const uint32_t r = val[j];
        s[0] += posd[4u * r];
        s[1] += posd[4u * r + 1u];
        s[2] += posd[4u * r + 2u];
        const float4 a = pay[2u * r], b = pay[2u * r + 1u];
        s[3] += a.x; s[4] += a.y; s[5] += a.z; s[6] += a.w;
        s[7] += b.x; s[8] += b.y; s[9] += b.z; s[10] += b.w;
    }
    if (g.pos) { g.pos[3u * k] = (float)s[0]; g.pos[3u * k + 1u] = (float)s[1]; g.pos[3u * k + 2u] = (float)s[2]; }
    if (g.col) { g.col[3u * k] = (float)s[3]; g.col[3u * k + 1u] = (float)s[4]; g.col[3u * k + 2u] = (float)s[5]; }
    if (g.uv) { g.uv[2u * k] = (float)s[6]; g.uv[2u * k + 1u] = (float)s[7]; }
    if (g.nrm) { g.nrm[3u * k] = (float)s[8]; g.nrm[3u * k + 1u] = (float)s[9]; g.nrm[3u * k + 2u] = (float)s[10]; }
}

template <class T>
__global__ void kd_zero(T *d, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = T(0);
}

template <class T>
hipError_t zero(T *d, size_t n, hipStream_t s) {
    if (!d || !n) return hipSuccess;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(kd_zero<T>, dim3(grid), dim3(256), 0, s, d, n);
    return hipGetLastError();
}

uint32_t bits_for(uint32_t n) {
    uint32_t b = 0;
    while (b < 32 && (1ull << b) <= n) ++b;
    return b;
}

// ---- batched views of many meshes

// Largest i < n with start[i] <= b (start[0] == 0, ascending): the owner of block b. The block
// index is wave-uniform, so the search runs on scalar loads.
// The last i with start[i] <= b (start ascending, start[0] = 0), by the whole wave at once: each
// lane compares one entry of each 64-entry chunk and the ballots are counted, so the loads are
// independent (a binary search would chain log2(n) dependent loads at the start of every block).
// Every lane of the wave must call it with the same b.
__device__ __forceinline__ uint32_t block_owner(const uint32_t *start, uint32_t n, uint32_t b) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t cnt = 0u;
    for (uint32_t c = 0; c < n; c += 64u) {
        const uint32_t j = c + lane;
        cnt += (uint32_t)__popcll(__ballot(j < n && start[j] <= b));
    }
    return cnt - 1u;
}

// Mesh shading pixel i of view V, or kNone when its draw is not one this call shades.
// A view whose draws map to meshes 0, 1, ... in order (identity_map) needs no table load.
__device__ __forceinline__ uint32_t pixel_mesh(const DiffBatchParams &P, const DiffBatchView &V, uint32_t i) {
    const uint32_t d = V.tags[i] - V.first_tag;
    if (d >= V.n_draws) return kNone;
    return V.identity_map ? d : P.draw_mesh[V.draw_off + d];
}

// Pixels with no primitive get (0, 0, 0, 0); pixels of draws mapped to no mesh are left as they are.
// Each block covers one 16x16 tile of one view; pixel `tid` of block `blk`. kDefer: the main pass,
// which leaves the pixels whose f32 barycentrics are ill-conditioned (slivers seen edge-on; rare)
// to the fix-up pass: each wave with such pixels appends its (block, wave) to a list. Without the
// FP64 path in it the kernel holds 70 VGPRs instead of 88 (7 waves per SIMD instead of 5).
template <bool kDefer>
__device__ __forceinline__ void batch_forward_pixel(const DiffBatchParams &P, uint32_t blk, uint32_t tid) {
    const uint32_t vi = block_owner(P.view_block0, P.n_views, blk);
    const DiffBatchView &V = P.views[vi];
    uint32_t x, y;
    tile_pixel(blk - P.view_block0[vi], V.v.W, tid, x, y);
    if (x >= V.v.W || y >= V.v.H) return;
    const uint32_t i = y * V.v.W + x;
    const uint32_t t = V.prim[i], tag = V.tags[i];  // both loads in flight together
    if (t == kNone) {
        V.out[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const uint32_t d = tag - V.first_tag;
    const uint32_t mi = d >= V.n_draws ? kNone : (V.identity_map ? d : P.draw_mesh[V.draw_off + d]);
    if (mi >= P.n_meshes) return;
    Eval<float> e;
    // A wave's pixels mostly show one mesh: its record (pointers, sizes, level tables) is then
    // read with scalar loads through a wave-uniform index, off the vector-memory path.
    const uint32_t m0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)mi);
    const int r = __ballot(mi != m0) == 0ull ? evaluate<float, kDefer>(P.meshes[m0].m, V.v, P.sh, t, x, y, e)
                                             : evaluate<float, kDefer>(P.meshes[mi].m, V.v, P.sh, t, x, y, e);
    if (kDefer && r == kDeferred) {
        // the active lanes here are the deferred ones: the lowest appends the wave
        if ((tid & 63u) == (uint32_t)__ffsll((unsigned long long)__ballot(1)) - 1u)
            P.defer_list[atomicAdd(&P.defer[0], 1u)] = blk * 4u + (tid >> 6);
        return;
    }
    V.out[i] = r ? make_float4(e.out[0], e.out[1], e.out[2], 1.0f) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// One block per tile; XCD x takes the x-th contiguous share of the tiles, so neighbouring tiles,
// which gather the same corners, share an L2. (A resident grid looping over tiles measured slower:
// the loop costs VGPRs, 97 against 72.)
__global__ __launch_bounds__(256) void kd_batch_forward(DiffBatchParams P) {
    batch_forward_pixel<true>(P, xcd_contiguous(blockIdx.x, gridDim.x), threadIdx.x);
}

// The fix-up pass: each wave takes listed waves of the main pass and shades them in full (their
// f32 pixels come out as before). The last block to finish zeroes the list's count and its own
// finish count, so both are zero between calls.
__global__ __launch_bounds__(256) void kd_batch_forward_fix(DiffBatchParams P) {
    const uint32_t n = P.defer[0];
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t k = blockIdx.x * 4u + (threadIdx.x >> 6); k < n; k += gridDim.x * 4u) {
        const uint32_t w = P.defer_list[k];
        batch_forward_pixel<false>(P, w >> 2, (w & 3u) * 64u + lane);
    }
    __syncthreads();  // every wave of the block has read the count
    if (threadIdx.x == 0u) {
        __threadfence();
        if (atomicAdd(&P.defer[1], 1u) == gridDim.x - 1u) {
            P.defer[0] = 0u;
            P.defer[1] = 0u;
        }
    }
}

__device__ __forceinline__ GradSink mesh_sink(const DiffBatchMesh &B) {
    return GradSink{B.g.pos ? B.posacc : nullptr, B.g.col, B.g.nrm, B.g.uv, B.g.tex};
}

__global__ __launch_bounds__(256) void kd_batch_backward(DiffBatchParams P, double *sh_part) {
    __shared__ VtxTable vt;
    __shared__ TexTable tt;
    tables_clear(vt, tt);
    __syncthreads();
    const uint32_t blk = xcd_contiguous(blockIdx.x, gridDim.x);  // neighbouring tiles under one L2
    const uint32_t vi = block_owner(P.view_block0, P.n_views, blk);
    const DiffBatchView &V = P.views[vi];
    uint32_t x, y;
    tile_pixel(blk - P.view_block0[vi], V.v.W, x, y);
    Eval<double> e;
    Adj a;
    bool live = false;
    uint32_t mi = kNone;
    if (x < V.v.W && y < V.v.H) {
        const uint32_t i = y * V.v.W + x;
        const uint32_t t = V.prim[i];
        if (t != kNone) {
            mi = pixel_mesh(P, V, i);
            if (mi < P.n_meshes) {
                live = evaluate(P.meshes[mi].m, V.v, P.sh, t, x, y, e);
                if (live) adjoint(P.meshes[mi].m, V.v, P.sh, e, V.gout[i], a);
            }
        }
    }
    sh_block_sum(live, a.dE, e.nh, sh_part);  // first: dE and nh hold no registers past it
    const GradSink sink = live ? mesh_sink(P.meshes[mi]) : GradSink{};
    const bool grouped = group_vertices(vt, sink, live, mi, e, a);
    if (live) gather_pixel(vt, tt, sink, mi, e, a, grouped);
    __syncthreads();
    tables_flush(vt, tt, [&](uint32_t m) { return mesh_sink(P.meshes[m]); });
}

// Clears the listed word ranges, 1024 words a block.
__global__ __launch_bounds__(256) void kd_batch_zero(DiffBatchParams P) {
    const uint32_t si = block_owner(P.zero_block0, P.n_zero, blockIdx.x);
    const DiffZeroSeg &z = P.zero[si];
    const uint32_t base = (blockIdx.x - P.zero_block0[si]) * 1024u;
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k) {
        const uint32_t w = base + k * 256u + threadIdx.x;
        if (w < z.words) z.p[w] = 0u;
    }
}

// Position gradients of every mesh from the FP64 sums.
__global__ __launch_bounds__(256) void kd_batch_pos_finalize(DiffBatchParams P) {
    const uint32_t mi = block_owner(P.mesh_block0, P.n_meshes, blockIdx.x);
    const DiffBatchMesh &B = P.meshes[mi];
    const uint32_t i = (blockIdx.x - P.mesh_block0[mi]) * 256u + threadIdx.x;
    if (B.g.pos && i < 3u * B.m.V) B.g.pos[i] = (float)B.posacc[i];
}

}  // namespace

uint32_t diff_blocks(uint32_t W, uint32_t H) { return (W * H + 255u) / 256u; }
uint32_t diff_tile_blocks(uint32_t W, uint32_t H) {
    return ((W + kTileSide - 1u) / kTileSide) * ((H + kTileSide - 1u) / kTileSide);
}

// ------------------------------------------------------------ mip chain and its adjoint
//
// The texture's mip chain by 2x2 box filtering (a 2x1 or 1x2 box once a side is 1; an odd side
// drops its last row or column), every level concatenated after level 0. Each texel of level l + 1
// is the sum of its window in level l, in row order, divided by the window's area. The adjoint
// maps a gradient on the whole chain to level 0: a level-0 texel receives its own entry plus, for
// each coarser level whose window holds its ancestor, that ancestor's entry divided by the
// product of the window areas in between.

struct MipDims {
    uint32_t w[kMaxMipLevels], h[kMaxMipLevels], off[kMaxMipLevels];
    uint32_t levels;
};

__host__ __device__ inline MipDims mip_dims(uint32_t w, uint32_t h) {
    MipDims d{};
    uint32_t off = 0u, l = 0u;
    for (;;) {
        d.w[l] = w;
        d.h[l] = h;
        d.off[l] = off;
        off += w * h;
        ++l;
        if ((w == 1u && h == 1u) || l == kMaxMipLevels) break;
        w = w > 1u ? w / 2u : 1u;
        h = h > 1u ? h / 2u : 1u;
    }
    d.levels = l;
    return d;
}

// Levels src + 1 .. src + kMipTileLevels of the chain, one block per 32x32 tile of level src
// (read from tex when src = 0, the block then also copying its tile to level 0). Each level is
// made in LDS from the one before, each texel the sum of its window in row order divided by the
// window's area, so the chain is bit-identical to a level-by-level build (which took one launch
// per level: 11 for a 1024^2 texture, now 2). A tile's texels of level k + 1 are those whose
// windows lie in its texels of level k: the tile corners are multiples of 2^5, so every window of
// an inner tile is whole, and the last tile of an odd side drops its last row or column as the
// whole level does.
constexpr uint32_t kMipTile = 32, kMipTileLevels = 5;

__global__ __launch_bounds__(256) void kd_mip_tiles(const float4 *tex, float4 *chain, MipDims d, uint32_t src) {
    __shared__ float4 s_a[kMipTile * kMipTile];              // levels src, src + 2, src + 4 (pitch 32)
    __shared__ float4 s_b[(kMipTile / 2) * (kMipTile / 2)];  // levels src + 1, src + 3, src + 5 (pitch 16)
    const uint32_t ntx = (d.w[src] + kMipTile - 1u) / kMipTile;
    uint32_t xs = (blockIdx.x % ntx) * kMipTile, ys = (blockIdx.x / ntx) * kMipTile;
    uint32_t rw = min(kMipTile, d.w[src] - xs), rh = min(kMipTile, d.h[src] - ys);
    const float4 *in = src == 0u ? tex : chain + d.off[src];
    for (uint32_t i = threadIdx.x; i < rw * rh; i += 256u) {
        const uint32_t x = i % rw, y = i / rw;
        const size_t at = (size_t)(ys + y) * d.w[src] + xs + x;
        const float4 t = in[at];
        s_a[y * kMipTile + x] = t;
        if (src == 0u) chain[at] = t;
    }
    __syncthreads();
    for (uint32_t k = 1; k <= kMipTileLevels && src + k < d.levels; ++k) {
        const uint32_t l = src + k - 1u, nl = src + k;
        const uint32_t kw = d.w[l] > 1u ? 2u : 1u, kh = d.h[l] > 1u ? 2u : 1u;
        const float4 *prev = (k & 1u) ? s_a : s_b;
        float4 *next = (k & 1u) ? s_b : s_a;
        const uint32_t pp = (k & 1u) ? kMipTile : kMipTile / 2u, np = (k & 1u) ? kMipTile / 2u : kMipTile;
        const uint32_t nxs = xs / kw, nys = ys / kh;
        const uint32_t nrw = min(rw / kw, d.w[nl] - nxs), nrh = min(rh / kh, d.h[nl] - nys);
        const float a = (float)(kw * kh);
        for (uint32_t i = threadIdx.x; i < nrw * nrh; i += 256u) {
            const uint32_t x = i % nrw, y = i / nrw;
            float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
            for (uint32_t dy = 0; dy < kh; ++dy)
                for (uint32_t dx = 0; dx < kw; ++dx) {
                    const float4 t = prev[(kh * y + dy) * pp + kw * x + dx];
                    s.x += t.x;
                    s.y += t.y;
                    s.z += t.z;
                    s.w += t.w;
                }
            const float4 o = make_float4(s.x / a, s.y / a, s.z / a, s.w / a);
            next[y * np + x] = o;
            chain[d.off[nl] + (size_t)(nys + y) * d.w[nl] + nxs + x] = o;
        }
        __syncthreads();
        xs = nxs;
        ys = nys;
        rw = nrw;
        rh = nrh;
    }
}

__global__ __launch_bounds__(256) void kd_mip_adjoint(const float4 *gchain, MipDims d, float4 *gtex) {
    const uint32_t n = d.w[0] * d.h[0];
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        // walk up the ancestors while each is inside the next level's windows; each level's entry
        // counts with the product of the window areas below it
        uint32_t x = i % d.w[0], y = i / d.w[0];
        float4 acc = gchain[i];
        float scale = 1.0f;
        for (uint32_t l = 0; l + 1u < d.levels; ++l) {
            const uint32_t kw = d.w[l] > 1u ? 2u : 1u, kh = d.h[l] > 1u ? 2u : 1u;
            x /= kw;
            y /= kh;
            if (x >= d.w[l + 1u] || y >= d.h[l + 1u]) break;  // a dropped odd row or column
            scale = scale / (float)(kw * kh);
            const float4 g = gchain[d.off[l + 1u] + y * d.w[l + 1u] + x];
            acc = make_float4(acc.x + g.x * scale, acc.y + g.y * scale, acc.z + g.z * scale, acc.w + g.w * scale);
        }
        gtex[i] = acc;
    }
}

uint32_t mip_chain_texels(uint32_t w, uint32_t h, uint32_t *levels) {
    const MipDims d = mip_dims(w, h);
    if (levels) *levels = d.levels;
    return d.off[d.levels - 1u] + d.w[d.levels - 1u] * d.h[d.levels - 1u];
}

hipError_t launch_mip_chain(const float4 *tex, uint32_t w, uint32_t h, float4 *chain, hipStream_t s) {
    const MipDims d = mip_dims(w, h);
    if (w == 0u || h == 0u) return hipSuccess;
    // level 0 is copied by the first pass; a 1x1 texture has no other level
    for (uint32_t src = 0; src == 0u || src + 1u < d.levels; src += kMipTileLevels) {
        const uint32_t tiles = ((d.w[src] + kMipTile - 1u) / kMipTile) * ((d.h[src] + kMipTile - 1u) / kMipTile);
        hipLaunchKernelGGL(kd_mip_tiles, dim3(tiles), dim3(256), 0, s, tex, chain, d, src);
    }
    return hipGetLastError();
}

hipError_t launch_mip_adjoint(const float4 *gchain, uint32_t w, uint32_t h, float4 *gtex, hipStream_t s) {
    const MipDims d = mip_dims(w, h);
    const uint32_t n = w * h;
    hipLaunchKernelGGL(kd_mip_adjoint, dim3(std::min<uint32_t>((n + 255u) / 256u, 4096u)), dim3(256), 0, s, gchain, d,
                       gtex);
    return hipGetLastError();
}

hipError_t launch_diff_forward(const DiffParams &p, float4 *out, hipStream_t s) {
    const uint32_t nb = diff_blocks(p.v.W, p.v.H);
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(kd_forward, dim3(nb), dim3(256), 0, s, p, out);
    return hipGetLastError();
}

hipError_t launch_diff_backward(const DiffParams &p, const float4 *gout, const DiffGrads &g, const DiffScratch &sc,
                                bool deterministic, hipStream_t s) {
    const uint32_t nb = deterministic ? diff_blocks(p.v.W, p.v.H) : diff_tile_blocks(p.v.W, p.v.H);
    hipError_t e;
    double *posacc = (!deterministic && g.pos) ? sc.posacc : nullptr;
    // every gradient array (and the FP64 position sums) cleared by one launch
    FillJobs z{};
    z.add(g.pos, 0u, 3ull * p.m.V);
    z.add(g.col, 0u, 3ull * p.m.V);
    z.add(g.uv, 0u, 2ull * p.m.V);
    z.add(g.nrm, 0u, 3ull * p.m.V);
    z.add(g.tex, 0u, 4ull * p.m.tex_total);
    z.add(posacc, 0u, 6ull * p.m.V);
    if (nb == 0) z.add(g.sh, 0u, 27);
    if ((e = launch_fill(z, s)) != hipSuccess || nb == 0) return e;
    if (!deterministic) {
        hipLaunchKernelGGL(kd_backward_atomic, dim3(nb), dim3(256), 0, s, p, gout, g, posacc, sc.sh_part);
        const uint32_t n_sh = g.sh ? 27u : 0u, n_pos = g.pos ? 3u * p.m.V : 0u;
        if (n_sh + n_pos)
            hipLaunchKernelGGL(kd_backward_tail, dim3(n_sh + (n_pos + 255u) / 256u), dim3(256), 0, s,
                               (const double *)sc.sh_part, nb, g.sh, n_sh, (const double *)posacc, g.pos, n_pos);
        return hipGetLastError();
    } else {
        hipLaunchKernelGGL(kd_backward_records, dim3(nb), dim3(256), 0, s, p, gout, sc);
        const uint32_t P = p.v.W * p.v.H;
        bool alt = false;
        const ScanScratch ss{sc.part, 0, sc.hist, 0, sc.epoch, sc.err};
        if (g.tex) {
            if ((e = sort_pairs(sc.tkey, sc.tval, sc.tkey2, sc.tval2, 4u * P, bits_for(p.m.tex_total), ss, &alt, s)) !=
                hipSuccess)
                return e;
            hipLaunchKernelGGL(kd_segment_texel, dim3((4u * P + 255u) / 256u), dim3(256), 0, s,
                               (const uint32_t *)(alt ? sc.tkey2 : sc.tkey), (const uint32_t *)(alt ? sc.tval2 : sc.tval),
                               (const float4 *)sc.tpay, 4u * P, p.m.tex_total, g.tex);
        }
        if (g.pos || g.col || g.uv || g.nrm) {
            if ((e = sort_pairs(sc.vkey, sc.vval, sc.vkey2, sc.vval2, 3u * P, bits_for(p.m.V), ss, &alt, s)) !=
                hipSuccess)
                return e;
            hipLaunchKernelGGL(kd_segment_vertex, dim3((3u * P + 255u) / 256u), dim3(256), 0, s,
                               (const uint32_t *)(alt ? sc.vkey2 : sc.vkey), (const uint32_t *)(alt ? sc.vval2 : sc.vval),
                               (const float4 *)sc.vpay, (const double *)sc.vposd, 3u * P, p.m.V, g);
        }
    }
    if (g.sh) hipLaunchKernelGGL(kd_sh_sum, dim3(27), dim3(256), 0, s, (const double *)sc.sh_part, nb, g.sh, 0u);
    return hipGetLastError();
}

// out = (a - b) * scale, four floats a lane where the arrays allow.
__global__ __launch_bounds__(256) void kd_l2_grad(const float *a, const float *b, uint64_t n, float scale, float *out) {
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    const uint64_t n4 = ((((uintptr_t)a | (uintptr_t)b | (uintptr_t)out) & 15u) == 0u) ? n / 4u : 0u;
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n4; i += stride) {
        const float4 x = ((const float4 *)a)[i], y = ((const float4 *)b)[i];
        ((float4 *)out)[i] = make_float4((x.x - y.x) * scale, (x.y - y.y) * scale, (x.z - y.z) * scale,
                                         (x.w - y.w) * scale);
    }
    for (uint64_t i = 4u * n4 + (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += stride)
        out[i] = (a[i] - b[i]) * scale;
}

hipError_t launch_l2_grad(const float *a, const float *b, uint64_t n, float scale, float *out, hipStream_t s) {
    if (!n) return hipSuccess;
    const unsigned grid = (unsigned)std::min<uint64_t>((n / 4u + 255u) / 256u + 1u, 8192u);
    hipLaunchKernelGGL(kd_l2_grad, dim3(grid), dim3(256), 0, s, a, b, n, scale, out);
    return hipGetLastError();
}

constexpr uint32_t kFixBlocks = 512;  // the fix-up pass's grid (its list is short)

hipError_t launch_diff_batch_forward(const DiffBatchParams &p, hipStream_t s) {
    if (p.n_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(kd_batch_forward, dim3(p.n_blocks), dim3(256), 0, s, p);
    hipLaunchKernelGGL(kd_batch_forward_fix, dim3(std::min(p.n_blocks, kFixBlocks)), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_diff_batch_backward(const DiffBatchParams &p, double *sh_part, float *grad_sh, bool accumulate,
                                      hipStream_t s) {
    if (!accumulate && p.n_zero_blocks)
        hipLaunchKernelGGL(kd_batch_zero, dim3(p.n_zero_blocks), dim3(256), 0, s, p);
    if (p.n_blocks) hipLaunchKernelGGL(kd_batch_backward, dim3(p.n_blocks), dim3(256), 0, s, p, sh_part);
    if (p.n_pos_blocks) hipLaunchKernelGGL(kd_batch_pos_finalize, dim3(p.n_pos_blocks), dim3(256), 0, s, p);
    if (grad_sh) {
        if (p.n_blocks)
            hipLaunchKernelGGL(kd_sh_sum, dim3(27), dim3(256), 0, s, (const double *)sh_part, p.n_blocks, grad_sh,
                               accumulate ? 1u : 0u);
        else if (!accumulate)
            return zero(grad_sh, 27, s);
    }
    return hipGetLastError();
}

}  // namespace mlv
