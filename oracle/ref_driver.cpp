// Test infrastructure only.  A small driver of the reference renderer's own
// public API, compiled against the reference sources where they lie (see
// Makefile.ref).  It turns a scene_desc into exactly the calls an application
// makes (Source/SDK/terrain/main.cpp:120-176 for model building,
// :381-421 for the frame), and exposes a few of the reference's internal
// functions so the oracle's restatement can be pinned piece by piece.
// Nothing here is shipped: the GPU parity tests check against the oracle's
// restatement, and the benchmark uses this build only to time the reference's
// CPU renderer.

#include "DFPSR/includeFramework.h"
#include "DFPSR/implementation/render/ITriangle2D.h"
#include "DFPSR/implementation/render/renderCore.h"
#include "scene_desc.h"
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

using namespace dsr;

static Transform3D makeTransform(const float *t) {
	return Transform3D(FVector3D(t[0], t[1], t[2]),
	                   FMatrix3x3(FVector3D(t[3], t[4], t[5]), FVector3D(t[6], t[7], t[8]), FVector3D(t[9], t[10], t[11])));
}

static Camera makeCamera(const scene_desc *s) {
	Transform3D location = makeTransform(s->cam_location);
	if (s->perspective) {
		return Camera::createPerspective(location, float(s->width), float(s->height), s->width_slope, s->near_clip, s->far_clip);
	} else {
		return Camera::createOrthogonal(location, float(s->width), float(s->height), s->width_slope);
	}
}

static TextureRgbaU8 makeTexture(int32_t w, int32_t h, int32_t res, const uint32_t *level0) {
	TextureRgbaU8 tex = texture_create_RgbaU8(w, h, res);
	int32_t tw = texture_getMaxWidth(tex);
	int32_t th = texture_getMaxHeight(tex);
	for (int32_t y = 0; y < th; y++) {
		SafePointer<uint32_t> row = texture_getSafePointer(tex, 0, y);
		for (int32_t x = 0; x < tw; x++) {
			row[x] = level0[(y % h) * w + (x % w)];
		}
	}
	texture_generatePyramid(tex);
	return tex;
}

struct BuiltScene {
	Model model;
	std::vector<TextureRgbaU8> textures;
	Camera camera;
	OrderedImageRgbaU8 color;
	AlignedImageF32 depth;
};

static void buildScene(const scene_desc *s, BuiltScene &b) {
	for (int32_t t = 0; t < s->texture_count; t++) {
		b.textures.push_back(makeTexture(s->texture_dims[t * 3], s->texture_dims[t * 3 + 1], s->texture_dims[t * 3 + 2], s->texture_level0[t]));
	}
	b.model = model_create();
	model_setFilter(b.model, s->filter ? Filter::Alpha : Filter::Solid);
	for (int32_t p = 0; p < s->point_count; p++) {
		model_addPoint(b.model, FVector3D(s->points[p * 3], s->points[p * 3 + 1], s->points[p * 3 + 2]));
	}
	for (int32_t part = 0; part < s->part_count; part++) {
		int32_t partIndex = model_addEmptyPart(b.model, U"part");
		if (s->part_diffuse && s->part_diffuse[part] >= 0) model_setDiffuseMap(b.model, partIndex, b.textures[s->part_diffuse[part]]);
		if (s->part_light && s->part_light[part] >= 0) model_setLightMap(b.model, partIndex, b.textures[s->part_light[part]]);
		for (int32_t poly = s->part_polygon_start[part]; poly < s->part_polygon_start[part + 1]; poly++) {
			const int32_t *idx = s->indices + poly * 4;
			int32_t polyIndex;
			int32_t corners;
			if (idx[3] < 0) {
				polyIndex = model_addTriangle(b.model, partIndex, idx[0], idx[1], idx[2]);
				corners = 3;
			} else {
				polyIndex = model_addQuad(b.model, partIndex, idx[0], idx[1], idx[2], idx[3]);
				corners = 4;
			}
			for (int32_t c = 0; c < corners; c++) {
				const float *col = s->colors + poly * 16 + c * 4;
				const float *tc = s->texcoords + poly * 16 + c * 4;
				model_setVertexColor(b.model, partIndex, polyIndex, c, FVector4D(col[0], col[1], col[2], col[3]));
				model_setTexCoord(b.model, partIndex, polyIndex, c, FVector4D(tc[0], tc[1], tc[2], tc[3]));
			}
		}
	}
	b.camera = makeCamera(s);
	if (s->has_color) {
		b.color = image_create_RgbaU8(s->width, s->height);
	}
	if (s->has_depth) {
		b.depth = image_create_F32(s->width, s->height);
	}
}

static void loadCanvas(const scene_desc *s, BuiltScene &b) {
	for (int32_t y = 0; y < s->height; y++) {
		if (s->has_color) {
			SafePointer<uint32_t> row = image_getSafePointer(b.color, y);
			for (int32_t x = 0; x < s->width; x++) row[x] = s->initial_color ? s->initial_color[y * s->width + x] : 0u;
		}
		if (s->has_depth) {
			SafePointer<float> row = image_getSafePointer(b.depth, y);
			for (int32_t x = 0; x < s->width; x++) row[x] = s->initial_depth ? s->initial_depth[y * s->width + x] : 0.0f;
		}
	}
}

// renderer_giveTask_triangle per polygon with the scene's own ProjectedPoints.
static void giveProjectedTriangles(const scene_desc *s, BuiltScene &b, Renderer &renderer) {
	const uint8_t *records = (const uint8_t *)s->projected;
	TextureRgbaU8 none;
	for (int32_t part = 0; part < s->part_count; part++) {
		const TextureRgbaU8 &diffuse = (s->part_diffuse && s->part_diffuse[part] >= 0) ? b.textures[s->part_diffuse[part]] : none;
		const TextureRgbaU8 &light = (s->part_light && s->part_light[part] >= 0) ? b.textures[s->part_light[part]] : none;
		for (int32_t poly = s->part_polygon_start[part]; poly < s->part_polygon_start[part + 1]; poly++) {
			ProjectedPoint p[3];
			FVector4D col[3], tc[3];
			for (int32_t c = 0; c < 3; c++) {
				const uint8_t *r = records + ((size_t)poly * 3 + c) * 40;
				float f[5];
				int64_t flat[2];
				memcpy(f, r, sizeof(f));
				memcpy(flat, r + 24, sizeof(flat));
				p[c] = ProjectedPoint(FVector3D(f[0], f[1], f[2]), FVector2D(f[3], f[4]), LVector2D(flat[0], flat[1]));
				const float *cc = s->colors + poly * 16 + c * 4;
				const float *tt = s->texcoords + poly * 16 + c * 4;
				col[c] = FVector4D(cc[0], cc[1], cc[2], cc[3]);
				tc[c] = FVector4D(tt[0], tt[1], tt[2], tt[3]);
			}
			renderer_giveTask_triangle(renderer, p[0], p[1], p[2], col[0], col[1], col[2], tc[0], tc[1], tc[2], diffuse, light,
			                           s->filter ? Filter::Alpha : Filter::Solid, b.camera);
		}
	}
}

// The scene's occluder calls that come before instance `before`, through the renderer API.
static void applyOccluders(const scene_desc *s, BuiltScene &b, Renderer &renderer, int32_t before, int32_t &query) {
	for (int32_t k = 0; k < s->occluder_count; k++) {
		const float *op = s->occluders + (size_t)k * 20;
		if ((int32_t)op[1] != before) continue;
		FVector3D mn(op[2], op[3], op[4]), mx(op[5], op[6], op[7]);
		Transform3D t = makeTransform(op + 8);
		switch ((int32_t)op[0]) {
		case 1: renderer_occludeFromBox(renderer, mn, mx, t, b.camera); break;
		case 2: renderer_occludeFromTopRows(renderer, b.camera); break;
		case 3: renderer_occludeFromExistingTriangles(renderer); break;
		case 4: {
			bool visible = renderer_isBoxVisible(renderer, mn, mx, t, b.camera);
			if (s->visible_out) s->visible_out[query] = visible ? 1 : 0;
			query++;
			break;
		}
		default: break;
		}
	}
}

static void renderFrame(const scene_desc *s, BuiltScene &b) {
	if (s->occluder_count > 0) {
		Renderer renderer = renderer_create();
		ImageRgbaU8 color = b.color;
		ImageF32 depth = b.depth;
		renderer_begin(renderer, color, depth);
		int32_t query = 0;
		for (int32_t i = 0; i < s->instance_count; i++) {
			applyOccluders(s, b, renderer, i, query);
			renderer_giveTask(renderer, b.model, makeTransform(s->instance_transforms + i * 12), b.camera);
		}
		applyOccluders(s, b, renderer, s->instance_count, query);
		renderer_end(renderer);
	} else if (s->projected) {
		Renderer renderer = renderer_create();
		ImageRgbaU8 color = b.color;
		ImageF32 depth = b.depth;
		renderer_begin(renderer, color, depth);
		giveProjectedTriangles(s, b, renderer);
		renderer_end(renderer);
	} else if (s->threaded) {
		Renderer renderer = renderer_create();
		ImageRgbaU8 color = b.color;
		ImageF32 depth = b.depth;
		renderer_begin(renderer, color, depth);
		for (int32_t i = 0; i < s->instance_count; i++) {
			renderer_giveTask(renderer, b.model, makeTransform(s->instance_transforms + i * 12), b.camera);
		}
		renderer_end(renderer);
	} else {
		ImageRgbaU8 color = b.color;
		ImageF32 depth = b.depth;
		for (int32_t i = 0; i < s->instance_count; i++) {
			model_render(b.model, makeTransform(s->instance_transforms + i * 12), color, depth, b.camera);
		}
	}
}

static void readCanvas(const scene_desc *s, BuiltScene &b, uint32_t *outColor, float *outDepth) {
	for (int32_t y = 0; y < s->height; y++) {
		if (s->has_color && outColor) {
			SafePointer<uint32_t> row = image_getSafePointer(b.color, y);
			for (int32_t x = 0; x < s->width; x++) outColor[y * s->width + x] = row[x];
		}
		if (s->has_depth && outDepth) {
			SafePointer<float> row = image_getSafePointer(b.depth, y);
			for (int32_t x = 0; x < s->width; x++) outDepth[y * s->width + x] = row[x];
		}
	}
}

// A resource pool that loads nothing: every non-empty name gets a 1x1 texture whose one texel
// holds the name's number (1, 2, ... in order of first request), so that the imported model's
// part textures can be traced back to the names the file asked for.
struct RecordingPool : public ResourcePool {
	List<String> names;
	List<TextureRgbaU8> textures;
	const ImageRgbaU8 fetchImageRgba(const ReadableString &) override { return ImageRgbaU8(); }
	const TextureRgbaU8 fetchTextureRgba(const ReadableString &name, int32_t) override {
		if (string_length(name) == 0) return TextureRgbaU8();
		for (int32_t i = 0; i < names.length(); i++) {
			if (string_caseInsensitiveMatch(names[i], name)) return textures[i];
		}
		names.push(name);
		TextureRgbaU8 t = texture_create_RgbaU8(1, 1, 1);
		texture_writePixel(t, 0, 0, 0, (uint32_t)names.length());
		textures.push(t);
		return t;
	}
};

static int32_t textureNumber(const TextureRgbaU8 &t) { return texture_exists(t) ? (int32_t)texture_readPixel(t, 0u, 0u, 0u) : 0; }

extern "C" {

// The reference's own end-of-program step (heap.cpp heap_terminatingApplication): after it the
// heap's static pool is torn down quietly.  Called once, when the loading process exits.
void ref_terminate() { heap_terminatingApplication(); }

// importFromContent_DMF1 of an ASCII text at one detail level.  Writes the points, per part
// (polygon count, diffuse texture number, light texture number), each polygon's three corners
// (point index, RGBA, u1 v1 u2 v2) and the part names and requested texture names, each
// ending in a line feed.  Returns the polygon count, or -1 when a capacity is too small.
int ref_import_dmf1(const char *text, int32_t detail, int32_t pointCapacity, float *points, int32_t *pointCount, int32_t partCapacity,
                    int32_t *partInfo, int32_t *partCount, int32_t polygonCapacity, int32_t *indices, float *colors, float *texcoords,
                    int32_t namesCapacity, char *names) {
	try {
		RecordingPool pool;
		Model model = importFromContent_DMF1(String(text), pool, detail);
		*pointCount = model_getNumberOfPoints(model);
		*partCount = model_getNumberOfParts(model);
		if (*pointCount > pointCapacity || *partCount > partCapacity) return -1;
		for (int32_t p = 0; p < *pointCount; p++) {
			FVector3D v = model_getPoint(model, p);
			points[p * 3] = v.x; points[p * 3 + 1] = v.y; points[p * 3 + 2] = v.z;
		}
		int32_t polygon = 0;
		std::string nameText;
		for (int32_t part = 0; part < *partCount; part++) {
			int32_t count = model_getNumberOfPolygons(model, part);
			partInfo[part * 3] = count;
			partInfo[part * 3 + 1] = textureNumber(model_getDiffuseMap(model, part));
			partInfo[part * 3 + 2] = textureNumber(model_getLightMap(model, part));
			String partName = model_getPartName(model, part);
			for (int32_t c = 0; c < string_length(partName); c++) nameText.push_back((char)partName[c]);
			nameText.push_back('\n');
			for (int32_t i = 0; i < count; i++, polygon++) {
				if (polygon >= polygonCapacity) return -1;
				for (int32_t v = 0; v < 3; v++) {
					indices[polygon * 3 + v] = model_getVertexPointIndex(model, part, i, v);
					FVector4D col = model_getVertexColor(model, part, i, v);
					FVector4D tc = model_getTexCoord(model, part, i, v);
					float *cd = colors + (polygon * 3 + v) * 4;
					float *td = texcoords + (polygon * 3 + v) * 4;
					cd[0] = col.x; cd[1] = col.y; cd[2] = col.z; cd[3] = col.w;
					td[0] = tc.x; td[1] = tc.y; td[2] = tc.z; td[3] = tc.w;
				}
			}
		}
		for (int32_t i = 0; i < pool.names.length(); i++) {
			for (int32_t c = 0; c < string_length(pool.names[i]); c++) nameText.push_back((char)pool.names[i][c]);
			nameText.push_back('\n');
		}
		if ((int32_t)nameText.size() + 1 > namesCapacity) return -1;
		memcpy(names, nameText.c_str(), nameText.size() + 1);
		return polygon;
	} catch (...) {
		return -2;
	}
}

int ref_render_scene(const scene_desc *s, uint32_t *outColor, float *outDepth) {
	try {
		BuiltScene b;
		buildScene(s, b);
		loadCanvas(s, b);
		renderFrame(s, b);
		readCanvas(s, b, outColor, outDepth);
		return 0;
	} catch (...) {
		return -1;
	}
}

// model_renderDepth (modelAPI.cpp:202-206) of every instance into the depth canvas alone.
int ref_render_depth_scene(const scene_desc *s, float *outDepth) {
	try {
		if (!s->has_depth) return -1;
		BuiltScene b;
		buildScene(s, b);
		loadCanvas(s, b);
		ImageF32 depth = b.depth;
		for (int32_t i = 0; i < s->instance_count; i++) {
			model_renderDepth(b.model, makeTransform(s->instance_transforms + i * 12), depth, b.camera);
		}
		readCanvas(s, b, nullptr, outDepth);
		return 0;
	} catch (...) {
		return -1;
	}
}

// Median wall time of renderer_begin..renderer_end over `iterations` frames,
// with the canvas reloaded (untimed) before each.
int ref_time_scene(const scene_desc *s, int32_t warmup, int32_t iterations, double *outSeconds) {
	try {
		BuiltScene b;
		buildScene(s, b);
		std::vector<double> times;
		for (int32_t i = 0; i < warmup + iterations; i++) {
			loadCanvas(s, b);
			auto t0 = std::chrono::steady_clock::now();
			renderFrame(s, b);
			auto t1 = std::chrono::steady_clock::now();
			if (i >= warmup) times.push_back(std::chrono::duration<double>(t1 - t0).count());
		}
		std::sort(times.begin(), times.end());
		*outSeconds = times.empty() ? 0.0 : times[times.size() / 2];
		return 0;
	} catch (...) {
		return -1;
	}
}

// The full pyramid buffer (smallest level first) after texture_generatePyramid.
int ref_texture_pyramid(int32_t w, int32_t h, int32_t res, const uint32_t *level0, uint32_t *outBuffer, int32_t capacity,
                        int32_t *outLog2w, int32_t *outLog2h, int32_t *outMaxMip) {
	TextureRgbaU8 tex = makeTexture(w, h, res, level0);
	int32_t levels = texture_getMipLevelCount(tex);
	int32_t total = 0;
	for (int32_t l = 0; l < levels; l++) total += texture_getWidth(tex, l) * texture_getHeight(tex, l);
	if (total > capacity) return -1;
	SafePointer<uint32_t> data = tex.impl_buffer.getSafe<uint32_t>("pyramid");
	for (int32_t i = 0; i < total; i++) outBuffer[i] = data[i];
	*outLog2w = tex.impl_log2width;
	*outLog2h = tex.impl_log2height;
	*outMaxMip = tex.impl_maxMipLevel;
	return total;
}

// Mip selection + bilinear sampling of 2x2 quads, as the "with mipmap" shaders
// call them (shaderMethods.h:59-83).  u, v: quadCount x 4 lanes.
int ref_sample_quads(int32_t w, int32_t h, int32_t res, const uint32_t *level0, int32_t quadCount,
                     const float *u, const float *v, uint32_t *outColor, uint32_t *outMip) {
	TextureRgbaU8 tex = makeTexture(w, h, res, level0);
	for (int32_t q = 0; q < quadCount; q++) {
		F32x4 vu(u[q * 4], u[q * 4 + 1], u[q * 4 + 2], u[q * 4 + 3]);
		F32x4 vv(v[q * 4], v[q * 4 + 1], v[q * 4 + 2], v[q * 4 + 3]);
		uint32_t mip = texture_getMipLevelIndex<F32x4>(tex, vu, vv);
		U32x4 c = texture_sample_bilinear<false, false, true, false>(tex, vu, vv, mip);
		UVector4D cv = c.get();
		outColor[q * 4] = cv.x; outColor[q * 4 + 1] = cv.y; outColor[q * 4 + 2] = cv.z; outColor[q * 4 + 3] = cv.w;
		outMip[q] = mip;
	}
	return 0;
}

// 1 / x as the perspective filler computes it (simd.h:4045-4068).
void ref_reciprocal(int32_t n, const float *in, float *out) {
	for (int32_t i = 0; i + 3 < n; i += 4) {
		FVector4D r = reciprocal(F32x4(in[i], in[i + 1], in[i + 2], in[i + 3])).get();
		out[i] = r.x; out[i + 1] = r.y; out[i + 2] = r.z; out[i + 3] = r.w;
	}
}

// Row intervals of one triangle in sub-pixel coordinates (ITriangle2D.cpp:150-169).
void ref_rasterize(const int64_t *corners, int32_t left, int32_t top, int32_t width, int32_t height, int32_t *outRows) {
	std::vector<RowInterval> rows(height);
	rasterizeTriangle(LVector2D(corners[0], corners[1]), LVector2D(corners[2], corners[3]), LVector2D(corners[4], corners[5]), rows.data(), IRect(left, top, width, height));
	for (int32_t r = 0; r < height; r++) {
		outRows[r * 2] = rows[r].left;
		outRows[r * 2 + 1] = rows[r].right;
	}
}

// Camera::worldToScreen of points (Camera.h:157-187): cs xyz, is xy, flat xy.
void ref_project(const scene_desc *s, const float *transform, int32_t n, const float *points, float *outCs, float *outIs, int64_t *outFlat) {
	Camera camera = makeCamera(s);
	Transform3D t = makeTransform(transform);
	for (int32_t i = 0; i < n; i++) {
		ProjectedPoint p = camera.worldToScreen(t.transformPoint(FVector3D(points[i * 3], points[i * 3 + 1], points[i * 3 + 2])));
		outCs[i * 3] = p.cs.x; outCs[i * 3 + 1] = p.cs.y; outCs[i * 3 + 2] = p.cs.z;
		outIs[i * 2] = p.is.x; outIs[i * 2 + 1] = p.is.y;
		outFlat[i * 2] = p.flat.x; outFlat[i * 2 + 1] = p.flat.y;
	}
}

}
