/*
 * Test infrastructure only: the scene description shared by the CPU oracle
 * (oracle.cpp) and the reference driver (ref_driver.cpp).  Neither the
 * product library nor its callers include this header.
 *
 * A scene is what one renderer_begin / renderer_giveTask(model)* / renderer_end
 * frame consumes in the reference (Source/SDK/terrain/main.cpp:381-421):
 *   - one model (points + parts; a part is textures + polygons),
 *     modelAPI.h:62-319, Model.h:39-133;
 *   - one or more instance transforms, each given to renderer_giveTask;
 *   - one camera made by Camera::createPerspective / createOrthogonal
 *     (Camera.h:144-156);
 *   - a colour canvas (OrderedImageRgbaU8) and a depth canvas (AlignedImageF32).
 */
#ifndef DFPSR_ORACLE_SCENE_DESC_H
#define DFPSR_ORACLE_SCENE_DESC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
	/* Canvas. Either canvas may be absent (has_color / has_depth = 0). */
	int32_t width, height;
	int32_t has_color, has_depth;
	/* Optional initial canvas content (width*height, tightly packed); NULL = zero. */
	const uint32_t *initial_color;
	const float *initial_depth;
	/* Camera location: position xyz, then xAxis, yAxis, zAxis (Transform3D.h:30-31). */
	float cam_location[12];
	/* Perspective: width_slope, near_clip, far_clip as Camera::createPerspective.
	 * Orthogonal (perspective = 0): width_slope is the half width (Camera::createOrthogonal). */
	int32_t perspective;
	float width_slope, near_clip, far_clip;
	/* Model. filter: 0 = Solid, 1 = Alpha (constants.h:33). */
	int32_t filter;
	int32_t point_count;
	const float *points;               /* point_count x 3 */
	int32_t part_count;
	const int32_t *part_polygon_start; /* part_count + 1 offsets into the polygon arrays */
	const int32_t *part_diffuse;       /* texture index per part, -1 = none */
	const int32_t *part_light;         /* texture index per part, -1 = none */
	const int32_t *indices;            /* polygon_count x 4, [3] = -1 for triangles */
	const float *colors;               /* polygon_count x 4 corners x RGBA */
	const float *texcoords;            /* polygon_count x 4 corners x (u1, v1, u2, v2) */
	/* Textures: per texture width, height, resolutions; level 0 pixels (width*height RGBA8). */
	int32_t texture_count;
	const int32_t *texture_dims;       /* texture_count x 3 */
	const uint32_t *const *texture_level0;
	/* Instances: each is one renderer_giveTask(model, transform, camera) call. */
	int32_t instance_count;
	const float *instance_transforms;  /* instance_count x 12, same layout as cam_location */
	/* 1 = renderer_create/begin/giveTask/end (12 row slices), 0 = model_render (one thread). */
	int32_t threaded;
	/* Optional: the polygons (triangles only) are given as renderer_giveTask_triangle calls
	 * (rendererAPI.h:123-128) with the caller's ProjectedPoints instead of renderer_giveTask of
	 * the model: polygon_count x 3 records of 40 bytes in the ProjectedPoint layout (cs xyz and
	 * is xy as float, flat xy as int64).  Colours, texture coordinates and each part's textures
	 * come from the model's arrays; points and instances are not used.  NULL: the model path. */
	const void *projected;
	/* Optional occlusion calls of the renderer (rendererAPI.h:77-93, :131), in order, each as 20
	 * floats: [0] kind (1 renderer_occludeFromBox, 2 renderer_occludeFromTopRows,
	 * 3 renderer_occludeFromExistingTriangles, 4 renderer_isBoxVisible), [1] the instance the
	 * call comes before (instance_count: after the last), [2..4] box minimum, [5..7] box
	 * maximum, [8..19] box transform (kinds 1 and 4).  The answers of the kind-4 calls go to
	 * visible_out, in order (1 = visible).  Scenes with occluder calls use the renderer. */
	int32_t occluder_count;
	const float *occluders;
	int32_t *visible_out;
} scene_desc;

#ifdef __cplusplus
}
#endif

#endif
