/*
 * CPU oracle for the polygon rendering path -- TEST INFRASTRUCTURE ONLY.
 *
 * A scalar C++14 restatement of the reference renderer's arithmetic, used by
 * tests/ as the checker of the GPU library and by bench.py as a CPU baseline.
 * The product path (libdfpsr_hip.so) never links, loads or calls it.
 * Each function cites the reference file:line it restates (paths relative to
 * the reference's Source/ folder).
 */
#ifndef DFPSR_ORACLE_H
#define DFPSR_ORACLE_H

#include <stdint.h>
#include "scene_desc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Render one frame.  threads <= 1: every command over the whole canvas in order
 * (model_render, modelAPI.cpp:197-201).  threads > 1: 12 row slices like
 * CommandQueue::execute (renderCore.cpp:449-480), shared by min(threads, 12)
 * worker threads.  Both give identical images.  Returns 0 on success. */
int32_t oracle_render_scene(const scene_desc *scene, int32_t threads, uint32_t *out_color, float *out_depth);

/* model_renderDepth (modelAPI.cpp:202-206) of every instance into the depth canvas alone,
 * starting from initial_depth (or zeros). */
int32_t oracle_render_depth_scene(const scene_desc *scene, float *out_depth);

/* Median seconds of the begin..end frame over `iterations` (after `warmup`). */
int32_t oracle_time_scene(const scene_desc *scene, int32_t threads, int32_t warmup, int32_t iterations, double *out_seconds);

/* Stage-level entry points used to pin the restatement piece by piece. */
void oracle_reciprocal(int32_t n, const float *in, float *out);
/* Texture layout probes on a TextureRgbaU8(log2w, log2h, maxMip) after `writes` (x, y, mip, value)
 * texture_writePixel calls.  op: 0 startOffset, 1 maxLevelMask, 2 layer offset(mip),
 * 3 pixel offset(x, y, mip), 4 readPixel(x, y, mip), 5 sample_nearest(u, v, mip), 6 max mip. */
uint32_t oracle_texture_query(int32_t log2w, int32_t log2h, int32_t max_mip, int32_t write_count, const uint32_t *writes,
                              int32_t op, uint32_t x, uint32_t y, uint32_t mip, float u, float v);
/* texture_interpolate_color_linear of one pixel pair (textureAPI.h:302-315). */
uint32_t oracle_mix_colors(uint32_t a, uint32_t b, uint32_t weight);
/* Differentiable float shader on the same raster (no reference counterpart): colour.rgb =
 * interpolated per-point RGB multipliers x bilinear(texture), colour.a = texture alpha.
 * Writes colour (W*H*4), visibility (command * 4 + lane, -1 elsewhere), depth and the weights
 * wa, wb, wc, u, v of each drawn pixel (W*H*5).  Returns the command count. */
int32_t oracle_diff_forward(const scene_desc *scene, const float *texture, int32_t log2w, int32_t log2h,
                            const float *multipliers, float *out_color, int32_t *out_visibility, float *out_depth,
                            float *out_weights);
/* Its adjoint in double precision: adds dL/dtexture and dL/dmultipliers from grad_color. */
void oracle_diff_backward(const scene_desc *scene, const float *texture, int32_t log2w, int32_t log2h,
                          const float *multipliers, const int32_t *visibility, const float *weights,
                          const float *grad_color, double *grad_texture, double *grad_multipliers);
/* packOrder_floatToSaturatedByte for RGBA (PackOrder.h:191-212). */
uint32_t oracle_pack_saturated(float r, float g, float b, float a);
void oracle_rasterize(const int64_t *corners, int32_t left, int32_t top, int32_t width, int32_t height, int32_t *out_rows);
int32_t oracle_texture_pyramid(int32_t w, int32_t h, int32_t res, const uint32_t *level0, uint32_t *out_buffer,
                               int32_t capacity, int32_t *out_log2w, int32_t *out_log2h, int32_t *out_max_mip);
int32_t oracle_sample_quads(int32_t w, int32_t h, int32_t res, const uint32_t *level0, int32_t quad_count,
                            const float *u, const float *v, uint32_t *out_color, uint32_t *out_mip);
void oracle_project(const scene_desc *scene, const float *transform, int32_t n, const float *points,
                    float *out_cs, float *out_is, int64_t *out_flat);

/* The occlusion grid a frame ends with (renderer_end's completeOcclusion reads it):
 * ceil(W/16) x ceil(H/16) cells into out_grid; returns 1 if the frame had occluder calls. */
int32_t oracle_occlusion_grid(const scene_desc *scene, float *out_grid, int32_t *out_width, int32_t *out_height);

/* Draw-command list of a frame, as renderer_end would execute it: per command the
 * source triangle (instance-major, then polygon fan order), the three projected
 * corners (flat sub-pixel x,y) and the pixel bound.  Returns the command count
 * (or the required capacity if larger than `capacity`). */
int32_t oracle_commands(const scene_desc *scene, int32_t capacity, int32_t *out_triangle, int64_t *out_flat, int32_t *out_bound);

#ifdef __cplusplus
}
#endif

#endif
