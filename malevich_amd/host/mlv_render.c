/*
 * mlv_render.c -- a C host of the draw pipeline: the counterpart of the
 * reference's render() (source/main.c:1265-1299) written against the C ABI
 * of include/malevich.h, with no Python and no HIP headers.
 *
 * It draws the reference's built-in Suprematism scene (main.c:232-254) with
 * passthrough_vs / passthrough_ps after clearing to render()'s clear colour
 * (main.c:1271) and depth 0, reads back the frame and writes it as raw
 * little-endian u32 pixels (R in the low byte) to the file named on the
 * command line.
 *
 * usage: mlv_render <out.raw> [width height]
 */
#include <stdio.h>
#include <stdlib.h>

#include "malevich.h"

/* pos4 + colour3 + pad per vertex, as main.c:232-250 lays them out */
static const float kVertices[11][8] = {
    {0.34107f, 0.12215f, 0.5f, 1.0f, 0.07500f, 0.08200f, 0.06300f, 0.0f},
    {0.95357f, 0.12500f, 0.5f, 1.0f, 0.07500f, 0.08200f, 0.06300f, 0.0f},
    {0.96250f, 0.86931f, 0.5f, 1.0f, 0.07500f, 0.08200f, 0.06300f, 0.0f},
    {0.33928f, 0.86505f, 0.5f, 1.0f, 0.07500f, 0.08200f, 0.06300f, 0.0f},
    {0.09464f, 0.12500f, 0.75f, 1.0f, 0.14100f, 0.29000f, 0.60800f, 0.0f},
    {0.69285f, 0.39772f, 0.75f, 1.0f, 0.14100f, 0.29000f, 0.60800f, 0.0f},
    {0.09107f, 0.60937f, 0.75f, 1.0f, 0.14100f, 0.29000f, 0.60800f, 0.0f},
    {0.00000f, 0.00000f, 0.25f, 1.0f, 0.96100f, 0.96100f, 0.92900f, 0.0f},
    {1.00000f, 0.00000f, 0.25f, 1.0f, 0.96100f, 0.96100f, 0.92900f, 0.0f},
    {1.00000f, 1.00000f, 0.25f, 1.0f, 0.96100f, 0.96100f, 0.92900f, 0.0f},
    {0.00000f, 1.00000f, 0.25f, 1.0f, 0.96100f, 0.96100f, 0.92900f, 0.0f},
};
/* padded to a multiple of 8 like the reference's index buffer (main.c:251-254) */
static const uint32_t kIndices[24] = {0, 1, 2, 2, 3, 0, 4, 5, 6, 7, 8, 9, 9, 10, 7, 0, 0, 0, 0, 0, 0, 0, 0, 0};

#define CHECK(call)                                                                              \
    do {                                                                                         \
        int rc_ = (call);                                                                        \
        if (rc_ != MLV_OK) {                                                                     \
            fprintf(stderr, "%s failed: %d (%s)\n", #call, rc_, ctx ? mlv_get_last_error(ctx) : ""); \
            return 1;                                                                            \
        }                                                                                        \
    } while (0)

int main(int argc, char **argv) {
    if (argc != 2 && argc != 4) {
        fprintf(stderr, "usage: %s <out.raw> [width height]\n", argv[0]);
        return 2;
    }
    const uint32_t W = argc == 4 ? (uint32_t)atoi(argv[2]) : 1200u;
    const uint32_t H = argc == 4 ? (uint32_t)atoi(argv[3]) : 720u;
    mlv_context *ctx = NULL;
    CHECK(mlv_context_create(0, &ctx));

    mlv_buffer *vb = NULL, *ib = NULL;
    mlv_render_target *rt = NULL;
    mlv_depth_stencil *ds = NULL;
    CHECK(mlv_buffer_create(ctx, MLV_BUFFER_VERTEX, sizeof kVertices, kVertices, &vb));
    CHECK(mlv_buffer_create(ctx, MLV_BUFFER_INDEX_U32, sizeof kIndices, kIndices, &ib));
    CHECK(mlv_render_target_create(ctx, W, H, &rt));
    CHECK(mlv_depth_stencil_create(ctx, W, H, &ds));

    /* render(): clears, then the pipeline state field by field, then the draw */
    const float clear[4] = {227.0f / 255.0f, 223.0f / 255.0f, 216.0f / 255.0f, 0.0f};
    CHECK(mlv_clear_render_target_view(ctx, rt, clear));
    CHECK(mlv_clear_depth_stencil_view(ctx, ds, 0.0f));
    const mlv_viewport vp = {0.0f, 0.0f, (float)W, (float)H, 0.0f, 1.0f};
    CHECK(mlv_rs_set_viewport(ctx, &vp));
    CHECK(mlv_om_set_render_targets(ctx, rt, ds));
    CHECK(mlv_ia_set_primitive_topology(ctx, MLV_TOPOLOGY_TRIANGLELIST));
    CHECK(mlv_ia_set_vertex_buffer(ctx, vb, sizeof kVertices[0], 0));
    CHECK(mlv_ia_set_index_buffer(ctx, ib, 0));
    CHECK(mlv_vs_set_shader(ctx, MLV_VS_PASSTHROUGH));
    CHECK(mlv_ps_set_shader(ctx, MLV_PS_PASSTHROUGH));
    CHECK(mlv_draw_indexed(ctx, 24, 0, 0));

    uint32_t *pixels = (uint32_t *)malloc((size_t)W * H * 4);
    if (!pixels) return 1;
    CHECK(mlv_render_target_readback(ctx, rt, pixels));
    FILE *f = fopen(argv[1], "wb");
    if (!f || fwrite(pixels, 4, (size_t)W * H, f) != (size_t)W * H) {
        fprintf(stderr, "cannot write %s\n", argv[1]);
        return 1;
    }
    fclose(f);
    uint64_t h = 1469598103934665603ull; /* FNV-1a over the pixels */
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        h ^= pixels[i];
        h *= 1099511628211ull;
    }
    mlv_stats st;
    CHECK(mlv_get_stats(ctx, &st));
    printf("mlv_render: %ux%u, %llu assembled triangles, fnv1a %016llx\n", W, H,
           (unsigned long long)st.assembled_triangle_count, (unsigned long long)h);
    free(pixels);
    CHECK(mlv_buffer_destroy(ctx, vb));
    CHECK(mlv_buffer_destroy(ctx, ib));
    CHECK(mlv_render_target_destroy(ctx, rt));
    CHECK(mlv_depth_stencil_destroy(ctx, ds));
    CHECK(mlv_context_destroy(ctx));
    return 0;
}
