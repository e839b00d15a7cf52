/*
 * malevich_diff.c -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 restatement of the differentiable shading the library adds on top of
 * the reference's raster path (the reference renders forward only). It is
 * the checker for the GPU kernels' forward image and gradients. The shading
 * model follows basic_ps (basic_ps.c:16-29: bilinear_f, csc.h:143-161, then
 * the sRGB curve of math.h:419-422) with per-vertex colour and a 9-coefficient
 * RGB SH light, at the pixel corners the reference's coverage test samples
 * (main.c:1015-1028):
 *
 *   clip_k   = clip_from_world * (position_k, 1)
 *   a_k      = X_k - u W_k,  b_k = Y_k - v W_k     (u, v: the pixel corner in NDC)
 *   lambda   = (a x b) / sum(a x b)                 (perspective-correct barycentrics)
 *   colour, uv, normal = sum_k lambda_k attr_k
 *   albedo   = bilinear(texture level L, uv)
 *   E        = sum_k sh[k] Y_k(normalize(normal))
 *   out      = max(1.055 c^(5/12) - 0.055, 0) with c = albedo * colour * E (0 where c <= 0)
 *
 * Everything continuous is computed in double. The one discrete choice other
 * than visibility, the mip level L, is computed in f32 exactly as the GPU
 * computes it (the same IEEE operation sequence; no library function is
 * involved), so both pick the same level bit-for-bit; the gradient is defined
 * at that fixed level. Backward: the exact adjoint of the forward in double,
 * summed sequentially in pixel order.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "malevich_oracle.h"

#define NONE 0xFFFFFFFFu

static const double SHC[5] = {(double)0.282094792f, (double)0.488602512f, (double)1.092548431f,
                              (double)0.315391565f, (double)0.546274215f};
static const double SRGB_P = (double)0.416666667f;

static void sh_basis(double x, double y, double z, double Y[9]) {
    Y[0] = SHC[0];
    Y[1] = SHC[1] * y;
    Y[2] = SHC[1] * z;
    Y[3] = SHC[1] * x;
    Y[4] = SHC[2] * x * y;
    Y[5] = SHC[2] * y * z;
    Y[6] = SHC[3] * (3.0 * z * z - 1.0);
    Y[7] = SHC[2] * x * z;
    Y[8] = SHC[4] * (x * x - y * y);
}

void orc_sh9(const double n[3], double out[9]) { sh_basis(n[0], n[1], n[2], out); }

/* ---- the f32 mip choice (same operation sequence as the GPU kernel) */

static float dot4_pair_f(const float *row, float x, float y, float z, float w) {
    float xy = row[0] * x + row[1] * y;
    float zw = row[2] * z + row[3] * w;
    return xy + zw;
}

static float ndc_u_f(float x, uint32_t W) { return (2.0f * x) / (float)W - 1.0f; }
static float ndc_v_f(float y, uint32_t H) { return 1.0f - (2.0f * y) / (float)H; }

static void bary_f(const float X[3], const float Y[3], const float Wc[3], float u, float v, float lam[3]) {
    float a[3], b[3], l[3];
    for (int k = 0; k < 3; ++k) {
        a[k] = X[k] - u * Wc[k];
        b[k] = Y[k] - v * Wc[k];
    }
    l[0] = a[1] * b[2] - a[2] * b[1];
    l[1] = a[2] * b[0] - a[0] * b[2];
    l[2] = a[0] * b[1] - a[1] * b[0];
    float S = (l[0] + l[1]) + l[2];
    const float inv = 1.0f / S; /* one division, three products */
    for (int k = 0; k < 3; ++k) lam[k] = l[k] * inv;
}

static void uv_at_f(const float *uv, const uint32_t vi[3], const float lam[3], float *u, float *v) {
    *u = (lam[0] * uv[2 * vi[0]] + lam[1] * uv[2 * vi[1]]) + lam[2] * uv[2 * vi[2]];
    *v = (lam[0] * uv[2 * vi[0] + 1] + lam[1] * uv[2 * vi[1] + 1]) + lam[2] * uv[2 * vi[2] + 1];
}

static uint32_t mip_level_f(const orc_diff_scene *s, const uint32_t vi[3], uint32_t x, uint32_t y) {
    if (s->tex_levels <= 1) return 0;
    float X[3], Y[3], Wc[3];
    for (int k = 0; k < 3; ++k) {
        const float *p = s->positions_f + 3 * vi[k];
        X[k] = dot4_pair_f(s->clip_from_world_f + 0, p[0], p[1], p[2], 1.0f);
        Y[k] = dot4_pair_f(s->clip_from_world_f + 4, p[0], p[1], p[2], 1.0f);
        Wc[k] = dot4_pair_f(s->clip_from_world_f + 12, p[0], p[1], p[2], 1.0f);
    }
    float l0[3], lx[3], ly[3], u0, v0, ux, vx, uy, vy;
    bary_f(X, Y, Wc, ndc_u_f((float)x, s->width), ndc_v_f((float)y, s->height), l0);
    bary_f(X, Y, Wc, ndc_u_f((float)x + 1.0f, s->width), ndc_v_f((float)y, s->height), lx);
    bary_f(X, Y, Wc, ndc_u_f((float)x, s->width), ndc_v_f((float)y + 1.0f, s->height), ly);
    uv_at_f(s->uvs_f, vi, l0, &u0, &v0);
    uv_at_f(s->uvs_f, vi, lx, &ux, &vx);
    uv_at_f(s->uvs_f, vi, ly, &uy, &vy);
    const float tw = (float)s->tex_width, th = (float)s->tex_height;
    const float dux = (ux - u0) * tw, dvx = (vx - v0) * th, duy = (uy - u0) * tw, dvy = (vy - v0) * th;
    const float rx = dux * dux + dvx * dvx, ry = duy * duy + dvy * dvy;
    const float r2 = rx > ry ? rx : ry;
    uint32_t lvl = 0;
    float thr = 2.0f;
    for (uint32_t k = 1; k < s->tex_levels; ++k) {
        if (r2 >= thr) lvl = k;
        thr *= 4.0f;
    }
    return lvl;
}

/* ---- FP64 evaluation */

typedef struct {
    uint32_t vi[3];
    double u, v, a[3], b[3], S, lam[3];
    double col[3], uv[2], n[3];
    uint32_t lw, lh, off;
    int32_t x0, y0, x1, y1;
    double fs, ft, t00[4], t10[4], t01[4], t11[4];
    double alb[3], nh[3], len, Y[9], E[3], lin[3], out[3];
} eval_t;

static void level_dims(const orc_diff_scene *s, uint32_t l, uint32_t *w, uint32_t *h, uint32_t *off) {
    uint32_t o = 0;
    for (uint32_t k = 0; k < l; ++k) {
        uint32_t wk = s->tex_width >> k, hk = s->tex_height >> k;
        o += (wk ? wk : 1) * (hk ? hk : 1);
    }
    *w = (s->tex_width >> l) ? (s->tex_width >> l) : 1;
    *h = (s->tex_height >> l) ? (s->tex_height >> l) : 1;
    *off = o;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static int evaluate(const orc_diff_scene *s, const uint32_t *prim, uint32_t x, uint32_t y, eval_t *e) {
    const uint32_t t = prim[(size_t)y * s->width + x];
    if (t == NONE || t >= s->triangle_count) return 0;
    for (int k = 0; k < 3; ++k) e->vi[k] = s->indices[3 * t + k];
    if (e->vi[0] >= s->vertex_count || e->vi[1] >= s->vertex_count || e->vi[2] >= s->vertex_count) return 0;
    e->u = 2.0 * x / s->width - 1.0;
    e->v = 1.0 - 2.0 * y / s->height;
    const double *M = s->clip_from_world;
    double l[3];
    for (int k = 0; k < 3; ++k) {
        const double *p = s->positions + 3 * e->vi[k];
        const double X = M[0] * p[0] + M[1] * p[1] + M[2] * p[2] + M[3];
        const double Y = M[4] * p[0] + M[5] * p[1] + M[6] * p[2] + M[7];
        const double W = M[12] * p[0] + M[13] * p[1] + M[14] * p[2] + M[15];
        e->a[k] = X - e->u * W;
        e->b[k] = Y - e->v * W;
    }
    l[0] = e->a[1] * e->b[2] - e->a[2] * e->b[1];
    l[1] = e->a[2] * e->b[0] - e->a[0] * e->b[2];
    l[2] = e->a[0] * e->b[1] - e->a[1] * e->b[0];
    e->S = l[0] + l[1] + l[2];
    for (int k = 0; k < 3; ++k) e->lam[k] = l[k] / e->S;
    for (int j = 0; j < 3; ++j) {
        e->col[j] = e->n[j] = 0.0;
        for (int k = 0; k < 3; ++k) {
            e->col[j] += e->lam[k] * s->colors[3 * e->vi[k] + j];
            e->n[j] += e->lam[k] * s->normals[3 * e->vi[k] + j];
        }
    }
    for (int j = 0; j < 2; ++j) {
        e->uv[j] = 0.0;
        for (int k = 0; k < 3; ++k) e->uv[j] += e->lam[k] * s->uvs[2 * e->vi[k] + j];
    }
    const uint32_t lvl = mip_level_f(s, e->vi, x, y);
    level_dims(s, lvl, &e->lw, &e->lh, &e->off);
    const double ss = e->lw * e->uv[0] - 0.5, tt = e->lh * (1.0 - e->uv[1]) - 0.5;
    const double fx = floor(ss), fy = floor(tt);
    e->fs = ss - fx;
    e->ft = tt - fy;
    const int ix = (fx > -1e9 && fx < 1e9) ? (int)fx : 0, iy = (fy > -1e9 && fy < 1e9) ? (int)fy : 0;
    e->x0 = clampi(ix, 0, (int)e->lw - 1);
    e->x1 = clampi(ix + 1, 0, (int)e->lw - 1);
    e->y0 = clampi(iy, 0, (int)e->lh - 1);
    e->y1 = clampi(iy + 1, 0, (int)e->lh - 1);
    const double *T = s->texture + 4 * (size_t)e->off;
    memcpy(e->t00, T + 4 * ((size_t)e->y0 * e->lw + e->x0), 32);
    memcpy(e->t10, T + 4 * ((size_t)e->y0 * e->lw + e->x1), 32);
    memcpy(e->t01, T + 4 * ((size_t)e->y1 * e->lw + e->x0), 32);
    memcpy(e->t11, T + 4 * ((size_t)e->y1 * e->lw + e->x1), 32);
    for (int j = 0; j < 3; ++j) {
        const double top = e->t00[j] * (1.0 - e->fs) + e->t10[j] * e->fs;
        const double bot = e->t01[j] * (1.0 - e->fs) + e->t11[j] * e->fs;
        e->alb[j] = top * (1.0 - e->ft) + bot * e->ft;
    }
    e->len = sqrt(e->n[0] * e->n[0] + e->n[1] * e->n[1] + e->n[2] * e->n[2]);
    for (int j = 0; j < 3; ++j) e->nh[j] = e->n[j] / e->len;
    sh_basis(e->nh[0], e->nh[1], e->nh[2], e->Y);
    for (int j = 0; j < 3; ++j) {
        double acc = 0.0;
        for (int k = 0; k < 9; ++k) acc += s->sh[3 * k + j] * e->Y[k];
        e->E[j] = acc;
        e->lin[j] = e->alb[j] * e->col[j] * e->E[j];
        const double o = e->lin[j] > 0.0 ? 1.055 * pow(e->lin[j], SRGB_P) - 0.055 : 0.0;
        e->out[j] = o > 0.0 ? o : 0.0;
    }
    return 1;
}

int orc_diff_forward(const orc_diff_scene *s, const uint32_t *prim, double *rgba) {
    if (!s || !prim || !rgba) return ORC_ERR_INVALID_ARG;
    for (uint32_t y = 0; y < s->height; ++y)
        for (uint32_t x = 0; x < s->width; ++x) {
            eval_t e;
            double *o = rgba + 4 * ((size_t)y * s->width + x);
            if (evaluate(s, prim, x, y, &e)) {
                o[0] = e.out[0];
                o[1] = e.out[1];
                o[2] = e.out[2];
                o[3] = 1.0;
            } else {
                o[0] = o[1] = o[2] = o[3] = 0.0;
            }
        }
    return ORC_OK;
}

int orc_diff_backward(const orc_diff_scene *s, const uint32_t *prim, const double *grad_rgba, orc_diff_grads *g) {
    if (!s || !prim || !grad_rgba || !g) return ORC_ERR_INVALID_ARG;
    const size_t V = s->vertex_count;
    uint32_t lw, lh, total;
    level_dims(s, s->tex_levels, &lw, &lh, &total);
    memset(g->positions, 0, V * 3 * sizeof(double));
    memset(g->colors, 0, V * 3 * sizeof(double));
    memset(g->uvs, 0, V * 2 * sizeof(double));
    memset(g->normals, 0, V * 3 * sizeof(double));
    memset(g->texture, 0, (size_t)total * 4 * sizeof(double));
    memset(g->sh, 0, 27 * sizeof(double));
    const double *M = s->clip_from_world;
    for (uint32_t y = 0; y < s->height; ++y)
        for (uint32_t x = 0; x < s->width; ++x) {
            eval_t e;
            if (!evaluate(s, prim, x, y, &e)) continue;
            const double *gr = grad_rgba + 4 * ((size_t)y * s->width + x);
            double dlin[3], dalb[3], dci[3], dE[3];
            for (int j = 0; j < 3; ++j) {
                const int on = e.lin[j] > 0.0 && e.out[j] > 0.0;
                dlin[j] = on ? gr[j] * 1.055 * SRGB_P * pow(e.lin[j], SRGB_P - 1.0) : 0.0;
                dalb[j] = dlin[j] * e.col[j] * e.E[j];
                dci[j] = dlin[j] * e.alb[j] * e.E[j];
                dE[j] = dlin[j] * e.alb[j] * e.col[j];
            }
            double dY[9];
            for (int k = 0; k < 9; ++k) {
                dY[k] = 0.0;
                for (int j = 0; j < 3; ++j) {
                    g->sh[3 * k + j] += dE[j] * e.Y[k];
                    dY[k] += dE[j] * s->sh[3 * k + j];
                }
            }
            const double nx = e.nh[0], ny = e.nh[1], nz = e.nh[2];
            double dnh[3];
            dnh[0] = SHC[1] * dY[3] + SHC[2] * (ny * dY[4] + nz * dY[7]) + 2.0 * SHC[4] * nx * dY[8];
            dnh[1] = SHC[1] * dY[1] + SHC[2] * (nx * dY[4] + nz * dY[5]) - 2.0 * SHC[4] * ny * dY[8];
            dnh[2] = SHC[1] * dY[2] + SHC[2] * (ny * dY[5] + nx * dY[7]) + 6.0 * SHC[3] * nz * dY[6];
            const double pr = nx * dnh[0] + ny * dnh[1] + nz * dnh[2];
            double dni[3];
            for (int j = 0; j < 3; ++j) dni[j] = (dnh[j] - e.nh[j] * pr) / e.len;
            /* bilinear */
            const double w[4] = {(1.0 - e.fs) * (1.0 - e.ft), e.fs * (1.0 - e.ft), (1.0 - e.fs) * e.ft, e.fs * e.ft};
            const size_t tap[4] = {e.off + (size_t)e.y0 * e.lw + e.x0, e.off + (size_t)e.y0 * e.lw + e.x1,
                                   e.off + (size_t)e.y1 * e.lw + e.x0, e.off + (size_t)e.y1 * e.lw + e.x1};
            double ds = 0.0, dt = 0.0;
            for (int j = 0; j < 3; ++j) {
                for (int q = 0; q < 4; ++q) g->texture[4 * tap[q] + j] += w[q] * dalb[j];
                const double dAs = (1.0 - e.ft) * (e.t10[j] - e.t00[j]) + e.ft * (e.t11[j] - e.t01[j]);
                const double top = e.t00[j] * (1.0 - e.fs) + e.t10[j] * e.fs;
                const double bot = e.t01[j] * (1.0 - e.fs) + e.t11[j] * e.fs;
                ds += dalb[j] * dAs;
                dt += dalb[j] * (bot - top);
            }
            const double duvi[2] = {ds * e.lw, -dt * e.lh};
            /* interpolation */
            double dlam[3];
            for (int k = 0; k < 3; ++k) {
                const uint32_t v = e.vi[k];
                double acc = 0.0;
                for (int j = 0; j < 3; ++j) {
                    g->colors[3 * v + j] += e.lam[k] * dci[j];
                    g->normals[3 * v + j] += e.lam[k] * dni[j];
                    acc += dci[j] * s->colors[3 * v + j] + dni[j] * s->normals[3 * v + j];
                }
                for (int j = 0; j < 2; ++j) {
                    g->uvs[2 * v + j] += e.lam[k] * duvi[j];
                    acc += duvi[j] * s->uvs[2 * v + j];
                }
                dlam[k] = acc;
            }
            const double mean = e.lam[0] * dlam[0] + e.lam[1] * dlam[1] + e.lam[2] * dlam[2];
            double dl[3], da[3], db[3];
            for (int k = 0; k < 3; ++k) dl[k] = (dlam[k] - mean) / e.S;
            da[0] = dl[2] * e.b[1] - dl[1] * e.b[2];
            da[1] = dl[0] * e.b[2] - dl[2] * e.b[0];
            da[2] = dl[1] * e.b[0] - dl[0] * e.b[1];
            db[0] = dl[1] * e.a[2] - dl[2] * e.a[1];
            db[1] = dl[2] * e.a[0] - dl[0] * e.a[2];
            db[2] = dl[0] * e.a[1] - dl[1] * e.a[0];
            for (int k = 0; k < 3; ++k) {
                const double dX = da[k], dYc = db[k], dW = -e.u * da[k] - e.v * db[k];
                for (int j = 0; j < 3; ++j) g->positions[3 * e.vi[k] + j] += M[j] * dX + M[4 + j] * dYc + M[12 + j] * dW;
            }
        }
    return ORC_OK;
}
