/*
 * dfpsr_hip.h -- C ABI of the MI355X (gfx950) polygon renderer, libdfpsr_hip.so.
 *
 * A drop-in for the reference renderer's polygon path: renderer_create / renderer_begin /
 * renderer_giveTask / renderer_end over modelAPI models, drawing into an OrderedImageRgbaU8
 * colour canvas and an AlignedImageF32 depth canvas.  Each entry point below names the
 * reference interface it replaces (paths relative to the reference's Source/ folder).
 *
 * Conventions
 *   - Every function returns DFPSR_HIP_OK (0) or a negative error code; no C++ exception
 *     crosses the ABI.  dfpsr_hip_last_error() returns a thread-local message for the last
 *     failure on the calling thread.
 *   - Plain pointers and sizes only.  Arrays are tightly packed, little-endian.
 *   - A renderer is used from one host thread at a time, like the reference's Renderer.
 *   - A model may be drawn from several threads at once (renderers or batches on different
 *     threads, dfpsr_hip_model_render_depth into different canvases) while no thread changes
 *     it; model_set_points and model_add_part must not overlap any draw of that model.
 *   - Colours are packed RGBA8 (red in the lowest byte), the OrderedImageRgbaU8 pack order.
 *   - Arithmetic follows the reference operation for operation: coverage and depth are
 *     bit-exact; RGBA8 colours may differ by 1 LSB where the reference's SSE reciprocal
 *     estimate (CPU-vendor specific) differs from the GPU's.
 */
#ifndef DFPSR_HIP_H
#define DFPSR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	DFPSR_HIP_OK = 0,
	DFPSR_HIP_ERROR_ARGUMENT = -1, /* null handle, bad size or index                      */
	DFPSR_HIP_ERROR_STATE = -2,    /* call out of order (begin twice, end without begin)   */
	DFPSR_HIP_ERROR_DEVICE = -3,   /* a HIP runtime call failed                            */
	DFPSR_HIP_ERROR_MEMORY = -4    /* device or host allocation failed                     */
};

/* Filters (implementation/render/constants.h:33). */
enum { DFPSR_HIP_FILTER_SOLID = 0, DFPSR_HIP_FILTER_ALPHA = 1 };

typedef struct dfpsr_hip_renderer dfpsr_hip_renderer;
typedef struct dfpsr_hip_model dfpsr_hip_model;
typedef struct dfpsr_hip_texture dfpsr_hip_texture;

/* A canvas: row 0 at `pixels`, rows `stride_bytes` apart.  pixels == NULL means no canvas
 * (the reference's empty image handle).  on_device = 1: a device pointer that stays where it
 * is (the renderer draws in place); 0: host memory, copied to the device at begin and back
 * at end. */
typedef struct {
	void *pixels;
	int32_t width, height;
	int32_t stride_bytes;
	int32_t on_device;
} dfpsr_hip_image;

/* Camera::createPerspective / Camera::createOrthogonal parameters (implementation/render/
 * Camera.h:144-156).  location = Transform3D: position xyz, then xAxis, yAxis, zAxis.
 * For an orthogonal camera width_slope is the half width and near/far are ignored. */
typedef struct {
	int32_t perspective;
	float location[12];
	float image_width, image_height;
	float width_slope, near_clip, far_clip;
} dfpsr_hip_camera;

/* Message of the last failed call on this thread ("" if none). */
const char *dfpsr_hip_last_error(void);

/* Library version string. */
const char *dfpsr_hip_version(void);

/* ---- Textures (api/textureAPI.h) -------------------------------------------------------- */

/* texture_create_RgbaU8(width, height, resolutions) (api/textureAPI.h:456, textureAPI.cpp:65-78):
 * a pyramid of power-of-two levels, smallest first, resident on the device. */
int dfpsr_hip_texture_create(dfpsr_hip_texture **out, int32_t width, int32_t height, int32_t resolutions);
/* Writes the highest resolution, as texture_getSafePointer(texture, 0, y) row writes do
 * (api/textureAPI.h:514-517).  pixels: (1 << log2h) rows of (1 << log2w) texels, row stride
 * in texels. */
int dfpsr_hip_texture_write_level0(dfpsr_hip_texture *texture, const uint32_t *pixels, int32_t stride_texels);
/* texture_generatePyramid (api/textureAPI.h:70, textureAPI.cpp:82-87), run on the GPU. */
int dfpsr_hip_texture_generate_pyramid(dfpsr_hip_texture *texture);
/* Layout: log2 width/height, smallest mip level index and the pyramid's texel count. */
int dfpsr_hip_texture_info(const dfpsr_hip_texture *texture, int32_t *log2w, int32_t *log2h, int32_t *max_mip,
                           int32_t *texel_count);
/* Copies the whole pyramid buffer (texel_count texels, smallest level first) to host memory. */
int dfpsr_hip_texture_read(const dfpsr_hip_texture *texture, uint32_t *out);
/* Device address of the pyramid buffer (for callers that keep data resident). */
int dfpsr_hip_texture_device_pointer(const dfpsr_hip_texture *texture, uint32_t **out);
void dfpsr_hip_texture_destroy(dfpsr_hip_texture *texture);

/* ---- Models (api/modelAPI.h) ------------------------------------------------------------ */

/* model_create + model_setFilter (api/modelAPI.h:62, :76). */
int dfpsr_hip_model_create(dfpsr_hip_model **out, int32_t filter);
/* model_addPoint for `count` points (api/modelAPI.h:124); replaces any previous points.
 * The model's bound starts at the origin and grows with each point (Model.cpp:281-311). */
int dfpsr_hip_model_set_points(dfpsr_hip_model *model, int32_t count, const float *xyz);
/* model_addEmptyPart + model_addTriangle/model_addQuad + model_setVertexColor +
 * model_setTexCoord + model_setDiffuseMap/model_setLightMap (api/modelAPI.h:93-258) for one
 * part: indices = count x 4 point indices (index 3 = -1 for a triangle), colors = count x 4
 * corners x RGBA, texcoords = count x 4 corners x (u1, v1, u2, v2).  Textures may be NULL
 * and must outlive the renderer_end of every frame that draws the model (as in the
 * reference, rendererAPI.h:104-107). */
int dfpsr_hip_model_add_part(dfpsr_hip_model *model, int32_t count, const int32_t *indices, const float *colors,
                             const float *texcoords, const dfpsr_hip_texture *diffuse, const dfpsr_hip_texture *light);
/* importFromContent_DMF1(content, pool, detail_level) (api/modelAPI.h:319,
 * implementation/render/model/format/dmf1.cpp): a new model from the text of a DMF1 file
 * (length bytes, ASCII or UTF-8).  Parts outside [MinDetailLevel, MaxDetailLevel] are left out;
 * equal points (closer than 0.00001) are merged; each triangle becomes a polygon.  Shader
 * M_Diffuse_1Tex takes Texture[0] as the diffuse map and M_Diffuse_2Tex also Texture[1] as
 * the light map, fetched from the caller's pool: texture_names[i] -> textures[i], names
 * compared ignoring ASCII case, first match; a name the pool lacks leaves the part without
 * that texture, as a texture the reference's pool fails to load does.  The model is Solid
 * (the reference's import does not apply FilterType).  Fails with DFPSR_HIP_ERROR_ARGUMENT
 * when the text does not start with "DMF1". */
int dfpsr_hip_model_import_dmf1(dfpsr_hip_model **out, const char *content, int64_t length, int32_t detail_level,
                                int32_t texture_count, const char *const *texture_names,
                                const dfpsr_hip_texture *const *textures);
/* model_getNumberOfPoints + model_getPoint (api/modelAPI.h:106-109): *count points into xyz
 * (count x 3 floats; xyz may be NULL to ask for the count). */
int dfpsr_hip_model_get_points(const dfpsr_hip_model *model, int32_t *count, float *xyz);
/* model_getNumberOfParts (api/modelAPI.h:96). */
int dfpsr_hip_model_part_count(const dfpsr_hip_model *model, int32_t *out);
/* model_getNumberOfPolygons, model_getVertexPointIndex, model_getVertexColor, model_getTexCoord,
 * model_getDiffuseMap, model_getLightMap (api/modelAPI.h:143-239) of one part: *polygons, then
 * (each output may be NULL) polygons x 4 indices (-1 for a triangle's fourth corner),
 * polygons x 16 colour and texture-coordinate floats, and the part's textures. */
int dfpsr_hip_model_get_part(const dfpsr_hip_model *model, int32_t part, int32_t *polygons, int32_t *indices, float *colors,
                             float *texcoords, const dfpsr_hip_texture **diffuse, const dfpsr_hip_texture **light);
/* model_getPartName (api/modelAPI.h:102): the name's length in *length, and the name,
 * NUL-terminated and cut to capacity - 1 bytes, in out (may be NULL). */
int dfpsr_hip_model_part_name(const dfpsr_hip_model *model, int32_t part, char *out, int32_t capacity, int32_t *length);
/* model_getBoundingBox (api/modelAPI.h:138). */
int dfpsr_hip_model_bound(const dfpsr_hip_model *model, float minimum[3], float maximum[3]);
/* model_renderDepth (api/modelAPI.h:291, modelAPI.cpp:202-206; renderCore.cpp:344-443): draws
 * the model's front-facing triangles into a depth canvas alone, keeping the greater inverse
 * depth for a perspective camera and the smaller depth for an orthogonal one.  A NULL or
 * pixel-less canvas draws nothing.  Synchronous: the canvas holds the result on return.
 * Rows that the reference's even-rounded row bound puts past an odd canvas height are not
 * drawn (the reference writes them past its image).  Like the reference's const
 * model_renderDepth, it may run on several threads at once for one model: its working memory
 * belongs to the calling thread.  Fails with DFPSR_HIP_ERROR_ARGUMENT, before any device
 * work, when a polygon names a point that model_set_points has since removed (as do every
 * other draw of a model: give_model, batch_render, diff_forward). */
int dfpsr_hip_model_render_depth(const dfpsr_hip_model *model, const float transform[12], const dfpsr_hip_image *depth,
                                 const dfpsr_hip_camera *camera);
/* No counterpart in the reference: dfpsr_hip_model_render_depth for callers that cannot
 * allocate device memory (code under HIP graph capture elsewhere in the process, tight
 * serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc/hipFree: the
 * projected points go to a fixed device array of 2^18 entries (10 MiB per GPU, held until the
 * process ends), and each fan triangle is drawn by its own group of lanes with a
 * compare-and-swap per pixel instead of the binned tiles of dfpsr_hip_model_render_depth (whose
 * bins are working memory).  The row and depth arithmetic is the same, and so is the canvas,
 * except where two drawn depths of +0 and -0 meet on a pixel: the tiles keep the first, as the
 * reference does, this entry either (a drawn depth of -0 needs weight-plane products that
 * underflow to zero).
 *   - The canvas must be device memory (on_device = 1); a host canvas fails with
 *     DFPSR_HIP_ERROR_ARGUMENT.
 *   - The model must already be resident: a model with changes not yet uploaded (never
 *     drawn, or changed since its last draw) fails with DFPSR_HIP_ERROR_STATE; any allocating
 *     draw of it uploads them.
 *   - A model of more than 262 144 points fails with DFPSR_HIP_ERROR_ARGUMENT.
 * All calls on one GPU share the array and would race on it if they overlapped; the library
 * serialises them with a process-wide lock.  Synchronous, like model_render_depth. */
int dfpsr_hip_model_render_depth_noalloc(const dfpsr_hip_model *model, const float transform[12],
                                         const dfpsr_hip_image *depth, const dfpsr_hip_camera *camera);
void dfpsr_hip_model_destroy(dfpsr_hip_model *model);

/* ---- Renderer (api/rendererAPI.h) ------------------------------------------------------- */

/* renderer_create (api/rendererAPI.h:56, rendererAPI.cpp:490-492). */
int dfpsr_hip_renderer_create(dfpsr_hip_renderer **out);
void dfpsr_hip_renderer_destroy(dfpsr_hip_renderer *renderer);
/* renderer_begin (api/rendererAPI.h:65, RendererImpl::beginFrame rendererAPI.cpp:151-168).
 * Either canvas may be absent; when both exist they must have the same size.  Fails with
 * DFPSR_HIP_ERROR_STATE when called twice without renderer_end. */
int dfpsr_hip_renderer_begin(dfpsr_hip_renderer *renderer, const dfpsr_hip_image *color, const dfpsr_hip_image *depth);
/* renderer_giveTask(renderer, model, modelToWorld, camera) = model_render_threaded
 * (api/modelAPI.h:299-301, modelAPI.cpp:214-281): box culling, projection of every point and
 * the quad fan (0,1,2),(0,2,3).  transform = Transform3D as in dfpsr_hip_camera.location. */
int dfpsr_hip_renderer_give_model(dfpsr_hip_renderer *renderer, const dfpsr_hip_model *model, const float transform[12],
                                  const dfpsr_hip_camera *camera);
/* renderer_giveTask_triangle for `count` triangles (api/rendererAPI.h:123-128,
 * rendererAPI.cpp:503-519).  Corners are given in camera space (count x 3 x xyz); the device
 * projects them with Camera::cameraToScreen, which is what ProjectedPoint holds.
 * colors / texcoords: count x 3 corners x 4 floats. */
int dfpsr_hip_renderer_give_triangles(dfpsr_hip_renderer *renderer, int32_t count, const float *camera_space,
                                      const float *colors, const float *texcoords, const dfpsr_hip_texture *diffuse,
                                      const dfpsr_hip_texture *light, int32_t filter, const dfpsr_hip_camera *camera);
/* ProjectedPoint (implementation/render/ProjectedPoint.h:33-49) in the reference's own 40-byte
 * layout: the camera-space position, the image-space position in pixels and the position in
 * 1/256 sub-pixels (truncated toward zero by Camera::cameraToScreen, Camera.h:160-187). */
typedef struct {
	float cs[3];
	float is[2];
	int64_t flat[2];
} dfpsr_hip_projected_point;

/* renderer_giveTask_triangle for `count` triangles exactly as the reference takes them
 * (api/rendererAPI.h:123-128, rendererAPI.cpp:503-519 -> renderTriangleFromData,
 * renderCore.cpp:284-341): corners = count x 3 of the caller's ProjectedPoints, used as given
 * (cs for culling and clipping, is for the weight planes, flat for coverage and the
 * front-face test), so a caller whose flat or is is not cameraToScreen(cs) draws as the
 * reference draws it.  Sub-triangles of a clipped triangle are projected again from their
 * camera-space corners, as in the reference (renderCore.cpp:220-240).
 * colors / texcoords: count x 3 corners x 4 floats. */
int dfpsr_hip_renderer_give_projected_triangles(dfpsr_hip_renderer *renderer, int32_t count,
                                                const dfpsr_hip_projected_point *corners, const float *colors,
                                                const float *texcoords, const dfpsr_hip_texture *diffuse,
                                                const dfpsr_hip_texture *light, int32_t filter,
                                                const dfpsr_hip_camera *camera);
/* ---- Occlusion (api/rendererAPI.h:77-93, :131, rendererAPI.cpp:36-477) ----------------------
 * Between renderer_begin and renderer_end a renderer keeps an occlusion grid of one cell per
 * 16 x 16 pixels (rounded up), each holding the largest camera-space depth at which something
 * in the cell can still be seen; it starts at +inf with the frame's first occluder call.
 * With occluders, renderer_giveTask skips a model whose bounding box the grid hides
 * (modelAPI.cpp:229-233), and renderer_end drops every draw command whose nearest corner is
 * not nearer than the grid plus 0.001 in every cell under its bound (completeOcclusion).
 * All calls fail with DFPSR_HIP_ERROR_STATE outside renderer_begin .. renderer_end. */

/* renderer_occludeFromBox (rendererAPI.h:77, rendererAPI.cpp:270-295): the convex hull of the
 * box's projected corners lowers the cells it covers whole to the box's farthest corner
 * depth.  A box with any corner outside the cull frustum (x and y halved) occludes nothing.
 * The reference's debugSilhouette lines belong to its debug drawing, which is not built. */
int dfpsr_hip_renderer_occlude_from_box(dfpsr_hip_renderer *renderer, const float minimum[3], const float maximum[3],
                                        const float transform[12], const dfpsr_hip_camera *camera);
/* renderer_occludeFromTopRows (rendererAPI.h:83, rendererAPI.cpp:403-474): each cell takes the
 * farthest depth of the first pixel row of its cell row in the depth canvas (as the canvas is
 * now: nothing of this frame is drawn before renderer_end), if nearer.  Runs on the GPU over
 * the device (or staged) canvas.  Fails with DFPSR_HIP_ERROR_STATE without a depth canvas. */
int dfpsr_hip_renderer_occlude_from_top_rows(dfpsr_hip_renderer *renderer, const dfpsr_hip_camera *camera);
/* renderer_occludeFromExistingTriangles (rendererAPI.h:131, rendererAPI.cpp:241-258): the draw
 * commands of every task given so far with the Solid filter lower the cells their triangles
 * cover whole.  Projects and sets up those tasks on the GPU; waits for it. */
int dfpsr_hip_renderer_occlude_from_existing_triangles(dfpsr_hip_renderer *renderer);
/* renderer_isBoxVisible (rendererAPI.h:93, rendererAPI.cpp:298-343, :536-541): *out = 0 when the
 * box is outside the cull frustum or hidden by the grid, else 1.  As in the reference the
 * grid outlives the frame: before the frame's first occluder call a query reads the grid an
 * earlier frame left, or 0 in every cell if this renderer never had one. */
int dfpsr_hip_renderer_is_box_visible(const dfpsr_hip_renderer *renderer, const float minimum[3], const float maximum[3],
                                      const float transform[12], const dfpsr_hip_camera *camera, int32_t *out);
/* renderer_hasOccluders (rendererAPI.h:85): *out = 1 after any occluder call of the frame. */
int dfpsr_hip_renderer_has_occluders(const dfpsr_hip_renderer *renderer, int32_t *out);
/* The frame's occlusion cells as they are now (no counterpart in the reference, which keeps
 * its grid private): *width x *height cells, row by row, into out (NULL: the size only);
 * 0 everywhere while the renderer never had a grid. */
int dfpsr_hip_renderer_occlusion_grid(const dfpsr_hip_renderer *renderer, float *out, int32_t *width, int32_t *height);

/* renderer_takesTriangles (api/rendererAPI.h:73). */
int dfpsr_hip_renderer_takes_triangles(const dfpsr_hip_renderer *renderer, int32_t *out);
/* renderer_end (api/rendererAPI.h:135, RendererImpl::endFrame rendererAPI.cpp:352-402):
 * draws every given triangle and returns when the canvases hold the result (host canvases
 * are copied back).  Fails with DFPSR_HIP_ERROR_STATE without renderer_begin. */
int dfpsr_hip_renderer_end(dfpsr_hip_renderer *renderer);
/* Statistics of the last renderer_end: draw commands, (tile, command) pairs, tiles, passes. */
int dfpsr_hip_renderer_stats(const dfpsr_hip_renderer *renderer, int64_t out[4]);
/* Per-stage device timing with HIP events on the renderer's stream (no counterpart in the
 * reference, which has no profiler hooks).  When enabled, renderer_end records the device
 * milliseconds of its last pass: [0] projection, [1] triangle setup, [2] tile binning,
 * [3] tile raster, [4] the whole pass.  With graph replay the frame is launched as two graphs
 * (everything before the raster, then the raster) and [0], [1] read 0 while [2] covers all
 * stages before the raster. */
int dfpsr_hip_renderer_set_timing(dfpsr_hip_renderer *renderer, int32_t enable);
int dfpsr_hip_renderer_timing(const dfpsr_hip_renderer *renderer, float out[5]);

/* ---- Tuning and batched rendering (no counterpart in the reference) --------------------- */

/* How a frame is executed on the device.  tile_config picks the screen tile of the raster
 * kernel (0: 16x16, 1: 32x16, 2: 32x32 pixels); persistent_per_cu = 0 gives every non-empty
 * tile its own workgroup, k > 0 launches k workgroups per CU that pull tiles from a queue;
 * use_graph = 1 captures the launch sequence into a hipGraph and replays it while the work
 * is unchanged; clear = 1 fills the colour canvas with clear_color and the depth canvas with
 * clear_depth first (the caller's image_fill).  raster_path = 0 draws frames whose commands are
 * all Solid and perspective, on canvases of even height, with the visibility raster (each
 * pixel's winning command is found with atomic max over (depth, command) keys and shaded
 * once); raster_path = 1 always walks every tile's commands in order.  binning = 0 lets the
 * setup kernel append each command to fixed-size per-tile lists (grown when a tile overflows
 * them, while they stay within a memory bound); binning = 1 always counts, scans and then fills
 * packed lists.  Results never depend on these settings. */
typedef struct {
	int32_t tile_config;
	int32_t persistent_per_cu;
	int32_t use_graph;
	int32_t clear;
	uint32_t clear_color;
	float clear_depth;
	int32_t raster_path;
	int32_t binning;
} dfpsr_hip_tuning;

/* Tuning used by renderer_end (default: 16x16 tiles, one workgroup per tile, no graph). */
int dfpsr_hip_renderer_set_tuning(dfpsr_hip_renderer *renderer, const dfpsr_hip_tuning *tuning);

/* One independent render job: what renderer_begin(color, depth) + renderer_giveTask(model,
 * transform, camera) + renderer_end does, on device canvases (on_device = 1). */
typedef struct {
	dfpsr_hip_image color, depth;
	const dfpsr_hip_model *model;
	float transform[12];
	dfpsr_hip_camera camera;
} dfpsr_hip_job;

typedef struct dfpsr_hip_batch dfpsr_hip_batch;

int dfpsr_hip_batch_create(dfpsr_hip_batch **out);
void dfpsr_hip_batch_destroy(dfpsr_hip_batch *batch);
/* Draws every job in one launch sequence over all jobs' triangles and tiles, returning when
 * every canvas holds its image.  Each image is identical to the one the plain renderer draws
 * for that job alone.  A call with the same job array as the previous call reuses the
 * uploaded tables (and, with use_graph, replays the captured graph). */
int dfpsr_hip_batch_render(dfpsr_hip_batch *batch, const dfpsr_hip_job *jobs, int32_t job_count,
                           const dfpsr_hip_tuning *tuning);
/* Statistics and timing of the last dfpsr_hip_batch_render, as for the renderer. */
int dfpsr_hip_batch_stats(const dfpsr_hip_batch *batch, int64_t out[4]);
int dfpsr_hip_batch_set_timing(dfpsr_hip_batch *batch, int32_t enable);
int dfpsr_hip_batch_timing(const dfpsr_hip_batch *batch, float out[5]);

/* Autotuner of the batched form.  Renders the jobs with every tile shape (tile_config 0..2)
 * and persistent_per_cu in {0, 1, 2, 4, 8}, graph replay on, and writes the fastest setting
 * to *best (clear, clear_color and clear_depth are taken from base, which may be NULL).
 * Each setting is timed as the median host wall-clock time of `repeats` whole
 * dfpsr_hip_batch_render calls after two untimed ones.  ms_out (may be NULL) receives the 15
 * medians in milliseconds at [tile_config * 5 + persistent step].  The canvases end up
 * holding the jobs' images, which are the same for every setting.  The bins are sized
 * exactly by a count pass and a scan, so there is no bin depth to tune. */
#define DFPSR_HIP_AUTOTUNE_TILE_CONFIGS 3
#define DFPSR_HIP_AUTOTUNE_PERSISTENT_STEPS 5
int dfpsr_hip_batch_autotune(dfpsr_hip_batch *batch, const dfpsr_hip_job *jobs, int32_t job_count, int32_t repeats,
                             const dfpsr_hip_tuning *base, dfpsr_hip_tuning *best, float *ms_out);

/* ---- Differentiable rendering (no counterpart in the reference) ------------------------- */

/* One view of a differentiable frame.  All buffers are device memory with stride_pixels
 * pixels per row: color = RGBA float32, visibility = the draw command * 4 + quad lane that
 * wrote each pixel (-1 where nothing was drawn), depth = the depth canvas, grad_color =
 * dL/dcolor (read by dfpsr_hip_diff_backward, may be NULL for the forward). */
typedef struct {
	float *color;
	int32_t *visibility;
	float *depth;
	const float *grad_color;
	int32_t width, height;
	int32_t stride_pixels;
	float transform[12];
	dfpsr_hip_camera camera;
} dfpsr_hip_view;

typedef struct dfpsr_hip_diff dfpsr_hip_diff;

int dfpsr_hip_diff_create(dfpsr_hip_diff **out);
void dfpsr_hip_diff_destroy(dfpsr_hip_diff *diff);
/* Forward: draws the model's triangles into every view with the reference's coverage, depth
 * test and perspective-correct weights, shading colour.rgb = (per-point RGB multipliers
 * interpolated by the vertex weights) x bilinear(texture, u1, v1) and colour.a = the
 * texture's alpha.  The bilinear taps and 8-bit sub-texel weights are those of
 * texture_sample_bilinear at the highest resolution (api/textureAPI.h:342-438).
 * texture: device RGBA float32, (1 << log2h) rows of (1 << log2w) texels; multipliers:
 * device, point_count x 3 floats (NULL = all 1).  With tuning->clear the views are cleared
 * first (colour 0, depth clear_depth, visibility -1). */
int dfpsr_hip_diff_forward(dfpsr_hip_diff *diff, const dfpsr_hip_model *model, const dfpsr_hip_view *views,
                           int32_t view_count, const float *texture, int32_t log2w, int32_t log2h,
                           const float *multipliers, const dfpsr_hip_tuning *tuning);
/* Adjoint of the last forward with its visibility fixed: adds dL/dtexture (texture layout) and
 * dL/dmultipliers (point_count x 3) from every view's grad_color, recomputing each pixel's
 * interpolants.  Either output may be NULL.  Float atomics: sums are order-dependent. */
int dfpsr_hip_diff_backward(dfpsr_hip_diff *diff, float *grad_texture, float *grad_multipliers,
                            const dfpsr_hip_tuning *tuning);
int dfpsr_hip_diff_stats(const dfpsr_hip_diff *diff, int64_t out[4]);
int dfpsr_hip_diff_set_timing(dfpsr_hip_diff *diff, int32_t enable);
int dfpsr_hip_diff_timing(const dfpsr_hip_diff *diff, float out[5]);

/* ---- Test hook --------------------------------------------------------------------------
 * No counterpart in the reference.  For each i, the column the device's row intervals take
 * from one edge: clamp((int32_t)(trunc(n[i] / d[i]) + 1), L[i], R[i]), the truncating int64
 * division of ITriangle2D.cpp:86-148 followed by the row's clamp to its clip columns.  Lets
 * the tests check the device arithmetic against exact integers.  d[i] must not be 0. */
int dfpsr_hip_probe_edge_side(const int64_t *n, const int64_t *d, const int32_t *L, const int32_t *R, int32_t count,
                              int32_t *out);

#ifdef __cplusplus
}
#endif

#endif
