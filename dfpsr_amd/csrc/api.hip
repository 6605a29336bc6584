// C ABI of libdfpsr_hip.so: textures, models and renderers (include/dfpsr_hip.h).
//
// Host side of the polygon path.  renderer_giveTask only records tasks (and culls whole
// models against the camera, as model_render_threaded does, modelAPI.cpp:227); all per-point,
// per-triangle and per-pixel work runs in the kernels of kernels.hip at renderer_end.
#include "../../include/dfpsr_hip.h"
#include "frame.h"
#include "occlusion.h"
#include "dmf1.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

using namespace dfpsr;

namespace dfpsr {
thread_local std::string g_lastError;

int fail(int code, const std::string &message) {
	g_lastError = message;
	return code;
}

hipError_t DevBuf::ensure(size_t need) {
	if (need <= bytes && ptr != nullptr) return hipSuccess;
	release();
	size_t n = need < 256 ? 256 : need;
	hipError_t e = hipMalloc(&ptr, n);
	if (e == hipSuccess) bytes = n; else ptr = nullptr;
	generation++;
	return e;
}

void DevBuf::release() {
	if (ptr) (void)hipFree(ptr);
	ptr = nullptr;
	bytes = 0;
}
}  // namespace dfpsr

// ---------------------------------------------------------------------------------------
// Textures
// ---------------------------------------------------------------------------------------

struct dfpsr_hip_texture {
	TextureDesc desc;  // desc.texels is the device pyramid
	DevBuf buffer;
};

static uint32_t log2Ceil(uint32_t size) {  // findLog2Size (textureAPI.cpp:31-41)
	for (uint32_t l = 0; l < 15; l++)
		if ((1u << l) >= size) return l;
	return 15;
}

extern "C" const char *dfpsr_hip_last_error(void) { return g_lastError.c_str(); }

extern "C" const char *dfpsr_hip_version(void) { return "dfpsr_hip 0.1 gfx950"; }

extern "C" int dfpsr_hip_texture_create(dfpsr_hip_texture **out, int32_t width, int32_t height, int32_t resolutions) {
	if (!out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "texture_create: null output");
	*out = nullptr;
	if (resolutions < 1) return fail(DFPSR_HIP_ERROR_ARGUMENT, "texture_create: a texture needs at least one resolution");
	if (width < 1 || height < 1 || width > 32768 || height > 32768)
		return fail(DFPSR_HIP_ERROR_ARGUMENT, "texture_create: dimensions must be 1..32768");
	std::unique_ptr<dfpsr_hip_texture> t(new (std::nothrow) dfpsr_hip_texture());
	if (!t) return fail(DFPSR_HIP_ERROR_MEMORY, "texture_create: out of host memory");
	t->desc = textureLayout(log2Ceil((uint32_t)width), log2Ceil((uint32_t)height), (uint32_t)(resolutions - 1));
	size_t bytes = (size_t)t->desc.texelCount * 4;
	if (t->buffer.ensure(bytes) != hipSuccess) return fail(DFPSR_HIP_ERROR_MEMORY, "texture_create: device allocation failed");
	HIP_TRY(hipMemset(t->buffer.ptr, 0, bytes));
	t->desc.texels = (const uint32_t *)t->buffer.ptr;
	*out = t.release();
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_texture_write_level0(dfpsr_hip_texture *t, const uint32_t *pixels, int32_t stride) {
	if (!t || !pixels) return fail(DFPSR_HIP_ERROR_ARGUMENT, "texture_write_level0: null argument");
	uint32_t w = 1u << t->desc.log2w, h = 1u << t->desc.log2h;
	if (stride < (int32_t)w) return fail(DFPSR_HIP_ERROR_ARGUMENT, "texture_write_level0: stride below the width");
	uint32_t *dst = (uint32_t *)t->buffer.ptr + t->desc.startOffset;
	HIP_TRY(hipMemcpy2D(dst, w * 4, pixels, (size_t)stride * 4, w * 4, h, hipMemcpyHostToDevice));
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_texture_generate_pyramid(dfpsr_hip_texture *t) {
	if (!t) return fail(DFPSR_HIP_ERROR_ARGUMENT, "texture_generate_pyramid: null texture");
	for (uint32_t level = 1; level <= t->desc.maxMip; level++) {
		uint32_t n = (1u << (t->desc.log2w - level)) * (1u << (t->desc.log2h - level));
		hipLaunchKernelGGL(k_downsample, dim3((n + 255) / 256), dim3(256), 0, 0, (uint32_t *)t->buffer.ptr, t->desc, level);
		HIP_TRY(hipGetLastError());
	}
	HIP_TRY(hipDeviceSynchronize());
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_texture_info(const dfpsr_hip_texture *t, int32_t *log2w, int32_t *log2h, int32_t *maxMip, int32_t *texels) {
	if (!t) return fail(DFPSR_HIP_ERROR_ARGUMENT, "texture_info: null texture");
	if (log2w) *log2w = (int32_t)t->desc.log2w;
	if (log2h) *log2h = (int32_t)t->desc.log2h;
	if (maxMip) *maxMip = (int32_t)t->desc.maxMip;
	if (texels) *texels = (int32_t)t->desc.texelCount;
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_texture_read(const dfpsr_hip_texture *t, uint32_t *out) {
	if (!t || !out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "texture_read: null argument");
	HIP_TRY(hipMemcpy(out, t->buffer.ptr, (size_t)t->desc.texelCount * 4, hipMemcpyDeviceToHost));
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_texture_device_pointer(const dfpsr_hip_texture *t, uint32_t **out) {
	if (!t || !out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "texture_device_pointer: null argument");
	*out = (uint32_t *)t->buffer.ptr;
	return DFPSR_HIP_OK;
}

extern "C" void dfpsr_hip_texture_destroy(dfpsr_hip_texture *t) { delete t; }

// ---------------------------------------------------------------------------------------
// Models
// ---------------------------------------------------------------------------------------

struct dfpsr_hip_model {
	int32_t filter = 0;
	std::vector<float> points;
	int32_t pointFloats = 3;  // floats per point: 3 (x, y, z), or 10 for 40-byte ProjectedPoint records
	int32_t pointCount() const { return (int32_t)(points.size() / (size_t)pointFloats); }
	f3 minBound = make3(0.0f, 0.0f, 0.0f), maxBound = make3(0.0f, 0.0f, 0.0f);
	std::vector<int32_t> indices;
	int32_t maxIndex = -1;  // the largest entry of indices, checked against the point count per draw
	std::vector<float> colors, texcoords;
	std::vector<int32_t> fanPolygon, polygonPart;
	std::vector<const TextureDesc *> partDiffuse, partLight;
	std::vector<const dfpsr_hip_texture *> partDiffuseHandle, partLightHandle;
	std::vector<std::string> partNames;
	std::vector<FanRecord> fans;  // rebuilt from indices and fanPolygon at each upload
	DevBuf dPoints, dIndices, dColors, dTexcoords, dFans, dPolygonPart;
	std::atomic<bool> dirty{true};
	// Guards the upload of changed arrays: several threads may draw one model at once (the
	// reference's model_renderDepth is const and documented for background threads,
	// modelAPI.h:280-291), and each draw first makes the model's arrays resident.
	std::mutex uploadLock;

	ModelView view() const {
		ModelView v;
		v.points = dPoints.as<float>();
		v.indices = dIndices.as<int32_t>();
		v.colors = dColors.as<float>();
		v.texcoords = dTexcoords.as<float>();
		v.fans = dFans.as<FanRecord>();
		v.polygonPart = dPolygonPart.as<int32_t>();
		v.pointCount = pointCount();
		v.fanCount = (int32_t)fanPolygon.size();
		v.partCount = (int32_t)partDiffuse.size();
		v.filter = filter;
		v.maxIndex = maxIndex;
		v.partDiffuse = partDiffuse.data();
		v.partLight = partLight.data();
		return v;
	}

	int upload() {
		std::lock_guard<std::mutex> hold(uploadLock);
		if (!dirty) return DFPSR_HIP_OK;
		fans.resize(fanPolygon.size());
		for (size_t g = 0; g < fanPolygon.size(); g++) {
			const int32_t fp = fanPolygon[g], poly = fp >> 1, k = fp & 1;
			const int32_t *ix = &indices[(size_t)poly * 4];
			fans[g] = FanRecord{{ix[0], ix[1 + k], ix[2 + k]}, fp};
		}
		struct { DevBuf *buf; const void *src; size_t bytes; } items[] = {
			{&dPoints, points.data(), points.size() * 4}, {&dIndices, indices.data(), indices.size() * 4},
			{&dColors, colors.data(), colors.size() * 4}, {&dTexcoords, texcoords.data(), texcoords.size() * 4},
			{&dFans, fans.data(), fans.size() * sizeof(FanRecord)}, {&dPolygonPart, polygonPart.data(), polygonPart.size() * 4}};
		for (auto &it : items) {
			if (it.buf->ensure(it.bytes) != hipSuccess) return fail(DFPSR_HIP_ERROR_MEMORY, "model upload: device allocation failed");
			if (it.bytes) HIP_TRY(hipMemcpy(it.buf->ptr, it.src, it.bytes, hipMemcpyHostToDevice));
		}
		dirty = false;
		return DFPSR_HIP_OK;
	}
};

extern "C" int dfpsr_hip_model_create(dfpsr_hip_model **out, int32_t filter) {
	if (!out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_create: null output");
	if (filter != DFPSR_HIP_FILTER_SOLID && filter != DFPSR_HIP_FILTER_ALPHA) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_create: unknown filter");
	*out = new (std::nothrow) dfpsr_hip_model();
	if (!*out) return fail(DFPSR_HIP_ERROR_MEMORY, "model_create: out of host memory");
	(*out)->filter = filter;
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_model_set_points(dfpsr_hip_model *m, int32_t count, const float *xyz) {
	if (!m || count < 0 || (count > 0 && !xyz)) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_set_points: bad argument");
	m->points.assign(xyz, xyz + (size_t)count * 3);
	// ModelImpl::expandBound from a bound that starts at the origin (Model.cpp:281-311).
	f3 mn = make3(0.0f, 0.0f, 0.0f), mx = make3(0.0f, 0.0f, 0.0f);
	for (int32_t i = 0; i < count; i++) {
		const float *p = xyz + (size_t)i * 3;
		if (mn.x > p[0]) mn.x = p[0];
		if (mn.y > p[1]) mn.y = p[1];
		if (mn.z > p[2]) mn.z = p[2];
		if (mx.x < p[0]) mx.x = p[0];
		if (mx.y < p[1]) mx.y = p[1];
		if (mx.z < p[2]) mx.z = p[2];
	}
	m->minBound = mn;
	m->maxBound = mx;
	m->dirty = true;
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_model_add_part(dfpsr_hip_model *m, int32_t count, const int32_t *indices, const float *colors,
                                        const float *texcoords, const dfpsr_hip_texture *diffuse, const dfpsr_hip_texture *light) {
	if (!m || count < 0 || (count > 0 && (!indices || !colors || !texcoords)))
		return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_add_part: bad argument");
	int32_t points = m->pointCount();
	for (int32_t i = 0; i < count; i++) {
		const int32_t *ix = indices + (size_t)i * 4;
		int32_t corners = ix[3] < 0 ? 3 : 4;
		for (int32_t c = 0; c < corners; c++)
			if (ix[c] < 0 || ix[c] >= points) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_add_part: point index out of range");
	}
	int32_t part = (int32_t)m->partDiffuse.size();
	int32_t polyBase = (int32_t)(m->indices.size() / 4);
	m->indices.insert(m->indices.end(), indices, indices + (size_t)count * 4);
	for (size_t i = 0; i < (size_t)count * 4; i++) m->maxIndex = std::max(m->maxIndex, indices[i]);
	m->colors.insert(m->colors.end(), colors, colors + (size_t)count * 16);
	m->texcoords.insert(m->texcoords.end(), texcoords, texcoords + (size_t)count * 16);
	for (int32_t i = 0; i < count; i++) {
		int32_t poly = polyBase + i;
		m->polygonPart.push_back(part);
		m->fanPolygon.push_back(poly * 2);
		if (indices[(size_t)i * 4 + 3] >= 0) m->fanPolygon.push_back(poly * 2 + 1);
	}
	m->partDiffuse.push_back(diffuse ? &diffuse->desc : nullptr);
	m->partLight.push_back(light ? &light->desc : nullptr);
	m->partDiffuseHandle.push_back(diffuse);
	m->partLightHandle.push_back(light);
	m->partNames.emplace_back();
	m->dirty = true;
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_model_import_dmf1(dfpsr_hip_model **out, const char *content, int64_t length, int32_t detailLevel,
                                           int32_t textureCount, const char *const *textureNames,
                                           const dfpsr_hip_texture *const *textures) {
	if (!out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_import_dmf1: null output");
	*out = nullptr;
	if (length < 0 || (length > 0 && !content) || textureCount < 0 || (textureCount > 0 && (!textureNames || !textures)))
		return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_import_dmf1: bad argument");
	dmf1::Imported imported;
	if (!dmf1::importModel(content, (size_t)length, detailLevel, imported))
		return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_import_dmf1: the content does not start with \"DMF1\"");
	// fetchTextureRgba from the caller's pool: the first entry whose name matches, ignoring
	// ASCII case; an empty name or a name the pool lacks leaves the part without that texture.
	auto fetch = [&](const std::string &name) -> const dfpsr_hip_texture * {
		if (name.empty()) return nullptr;
		for (int32_t i = 0; i < textureCount; i++)
			if (textureNames[i] && dmf1::sameTextureName(name, textureNames[i])) return textures[i];
		return nullptr;
	};
	std::unique_ptr<dfpsr_hip_model> m(new (std::nothrow) dfpsr_hip_model());
	if (!m) return fail(DFPSR_HIP_ERROR_MEMORY, "model_import_dmf1: out of host memory");
	int rc = dfpsr_hip_model_set_points(m.get(), (int32_t)(imported.points.size() / 3), imported.points.data());
	if (rc) return rc;
	for (const dmf1::ImportedPart &part : imported.parts) {
		rc = dfpsr_hip_model_add_part(m.get(), (int32_t)(part.indices.size() / 4), part.indices.data(), part.colors.data(),
		                              part.texcoords.data(), fetch(part.diffuseName), fetch(part.lightName));
		if (rc) return rc;
		m->partNames.back() = part.name;
	}
	*out = m.release();
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_model_get_points(const dfpsr_hip_model *m, int32_t *count, float *xyz) {
	if (!m || !count) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_get_points: null argument");
	if (m->pointFloats != 3) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_get_points: not a model of positions");
	*count = m->pointCount();
	if (xyz && !m->points.empty()) std::memcpy(xyz, m->points.data(), m->points.size() * sizeof(float));
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_model_part_count(const dfpsr_hip_model *m, int32_t *out) {
	if (!m || !out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_part_count: null argument");
	*out = (int32_t)m->partDiffuse.size();
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_model_get_part(const dfpsr_hip_model *m, int32_t part, int32_t *polygons, int32_t *indices, float *colors,
                                        float *texcoords, const dfpsr_hip_texture **diffuse, const dfpsr_hip_texture **light) {
	if (!m || !polygons) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_get_part: null argument");
	if (part < 0 || part >= (int32_t)m->partDiffuse.size()) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_get_part: part index out of range");
	// Parts own consecutive polygons (model_add_part appends them).
	int32_t first = 0;
	while (first < (int32_t)m->polygonPart.size() && m->polygonPart[(size_t)first] != part) first++;
	int32_t count = 0;
	while (first + count < (int32_t)m->polygonPart.size() && m->polygonPart[(size_t)(first + count)] == part) count++;
	*polygons = count;
	if (count > 0) {  // an empty part has no polygon to point into
		if (indices) std::memcpy(indices, m->indices.data() + (size_t)first * 4, (size_t)count * 4 * sizeof(int32_t));
		if (colors) std::memcpy(colors, m->colors.data() + (size_t)first * 16, (size_t)count * 16 * sizeof(float));
		if (texcoords) std::memcpy(texcoords, m->texcoords.data() + (size_t)first * 16, (size_t)count * 16 * sizeof(float));
	}
	if (diffuse) *diffuse = m->partDiffuseHandle[(size_t)part];
	if (light) *light = m->partLightHandle[(size_t)part];
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_model_part_name(const dfpsr_hip_model *m, int32_t part, char *out, int32_t capacity, int32_t *length) {
	if (!m) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_part_name: null model");
	if (part < 0 || part >= (int32_t)m->partNames.size()) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_part_name: part index out of range");
	const std::string &name = m->partNames[(size_t)part];
	if (length) *length = (int32_t)name.size();
	if (out && capacity > 0) {
		const size_t n = std::min(name.size(), (size_t)capacity - 1);
		std::memcpy(out, name.data(), n);
		out[n] = '\0';
	}
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_model_bound(const dfpsr_hip_model *m, float mn[3], float mx[3]) {
	if (!m || !mn || !mx) return fail(DFPSR_HIP_ERROR_ARGUMENT, "model_bound: null argument");
	mn[0] = m->minBound.x; mn[1] = m->minBound.y; mn[2] = m->minBound.z;
	mx[0] = m->maxBound.x; mx[1] = m->maxBound.y; mx[2] = m->maxBound.z;
	return DFPSR_HIP_OK;
}

extern "C" void dfpsr_hip_model_destroy(dfpsr_hip_model *m) { delete m; }

// ---------------------------------------------------------------------------------------
// Renderer
// ---------------------------------------------------------------------------------------

struct dfpsr_hip_renderer {
	Frame frame;
	bool receiving = false;
	dfpsr_hip_image color{}, depth{};
	int32_t width = 0, height = 0;
	DevBuf colorStage, depthStage;
	std::vector<std::unique_ptr<dfpsr_hip_model>> triangleModels;  // giveTask_triangle batches of this frame
	RenderOptions options;
	// Occlusion grid (RendererImpl's depthGrid, rendererAPI.cpp:140-188).  Like the reference's
	// grid image it outlives the frame: it is made at the first occluder call that finds none
	// large enough, and refilled with +inf by the first occluder call of each frame, so a box
	// query before that call reads what an earlier frame left (or 0 while there is none).
	// The frame's cells are the top-left gridW x gridH of the image (imageW cells per row).
	std::vector<float> gridImage;
	int32_t imageW = 0, imageH = 0;
	int32_t gridW = 0, gridH = 0;
	bool occluded = false;  // an occluder call since renderer_begin (renderer_hasOccluders)
	std::vector<float> cells;  // the frame's cells packed for the device
	DevBuf dGrid;

	const float *gridRead() const { return gridImage.empty() ? nullptr : gridImage.data(); }
	// prepareForOcclusion (rendererAPI.cpp:179-188).
	void prepareForOcclusion() {
		if (!occluded) {
			if (!(!gridImage.empty() && imageW >= gridW && imageH >= gridH)) {
				imageW = gridW;
				imageH = gridH;
				gridImage.assign((size_t)imageW * imageH, 0.0f);
			}
			std::fill(gridImage.begin(), gridImage.end(), __builtin_huge_valf());
		}
		occluded = true;
	}
	float *deviceCanvasDepth() const {
		return depth.pixels ? (float *)(depth.on_device ? depth.pixels : depthStage.ptr) : nullptr;
	}
	// The frame's cells to the device (packed gridW x gridH) and back into the image.
	int uploadGrid() {
		cells.resize((size_t)gridW * gridH);
		for (int32_t y = 0; y < gridH; y++)
			std::memcpy(&cells[(size_t)y * gridW], &gridImage[(size_t)y * imageW], (size_t)gridW * sizeof(float));
		const size_t bytes = cells.size() * sizeof(float);
		if (dGrid.ensure(bytes) != hipSuccess) return fail(DFPSR_HIP_ERROR_MEMORY, "occlusion grid: device allocation failed");
		if (bytes) HIP_TRY(hipMemcpyAsync(dGrid.ptr, cells.data(), bytes, hipMemcpyHostToDevice, frame.stream));
		return DFPSR_HIP_OK;
	}
	int downloadGrid() {
		const size_t bytes = cells.size() * sizeof(float);
		if (bytes) HIP_TRY(hipMemcpyAsync(cells.data(), dGrid.ptr, bytes, hipMemcpyDeviceToHost, frame.stream));
		HIP_TRY(hipStreamSynchronize(frame.stream));
		for (int32_t y = 0; y < gridH; y++)
			std::memcpy(&gridImage[(size_t)y * imageW], &cells[(size_t)y * gridW], (size_t)gridW * sizeof(float));
		return DFPSR_HIP_OK;
	}
};

static RenderOptions optionsFrom(const dfpsr_hip_tuning *t) {
	RenderOptions o;
	o.tileConfig = t->tile_config;
	o.persistentPerCU = t->persistent_per_cu;
	o.graph = t->use_graph != 0;
	o.clear = t->clear != 0;
	o.clearColor = t->clear_color;
	o.clearDepth = t->clear_depth;
	o.rasterPath = t->raster_path;
	o.binning = t->binning;
	return o;
}

static int checkTuning(const dfpsr_hip_tuning *t) {
	if (!t) return fail(DFPSR_HIP_ERROR_ARGUMENT, "null tuning");
	if (t->tile_config < 0 || t->tile_config >= TILE_CONFIGS) return fail(DFPSR_HIP_ERROR_ARGUMENT, "tile_config must be 0, 1 or 2");
	if (t->persistent_per_cu < 0 || t->persistent_per_cu > 64) return fail(DFPSR_HIP_ERROR_ARGUMENT, "persistent_per_cu must be 0..64");
	if (t->raster_path < 0 || t->raster_path > 1) return fail(DFPSR_HIP_ERROR_ARGUMENT, "raster_path must be 0 or 1");
	if (t->binning < 0 || t->binning > 1) return fail(DFPSR_HIP_ERROR_ARGUMENT, "binning must be 0 or 1");
	return DFPSR_HIP_OK;
}

static CameraParams cameraFrom(const dfpsr_hip_camera *c) {
	return deriveCamera(c->perspective, c->location, c->image_width, c->image_height, c->width_slope, c->near_clip, c->far_clip);
}

static int checkCamera(const dfpsr_hip_camera *c) {
	if (!c) return fail(DFPSR_HIP_ERROR_ARGUMENT, "null camera");
	if (!(c->image_width > 0.0f) || !(c->image_height > 0.0f) || c->width_slope == 0.0f)
		return fail(DFPSR_HIP_ERROR_ARGUMENT, "camera needs a positive image size and a non-zero slope");
	return DFPSR_HIP_OK;
}

static bool hasImage(const dfpsr_hip_image *im) { return im && im->pixels; }

// model_set_points may replace the points with fewer after polygons were added; drawing such a
// model would read points past the model's own.  Checked before any device work.
static int checkModelIndices(const dfpsr_hip_model *m, const char *who) {
	if (m->maxIndex >= m->pointCount())
		return fail(DFPSR_HIP_ERROR_ARGUMENT, std::string(who) + ": a polygon names a point the model does not have");
	return DFPSR_HIP_OK;
}

// Lanes per fan triangle in the depth-only draw, as a power-of-two exponent: 32 lanes (fastest
// on the 1080p terrain: 0.160 ms per call against 0.234, 0.198 and 0.187 ms for 8, 16 and 64)
// unless DFPSR_HIP_DEPTH_LANES names 8, 16 or 64 (a tuning override, read once).
static int32_t depthGroupShift() {
	static const int32_t shift = [] {
		const char *v = getenv("DFPSR_HIP_DEPTH_LANES");
		const int lanes = v ? atoi(v) : 32;
		return lanes == 8 ? 3 : lanes == 16 ? 4 : lanes == 64 ? 6 : 5;
	}();
	return shift;
}

// Working memory of the scratch-free depth-only draw: projected points of one model, in a
// device array of fixed size that every call on a GPU shares (each GPU holds its own copy
// until the process ends).  Host code never passes the array's name to a kernel: the address
// comes from hipGetSymbolAddress, once per device, with that device current.
static constexpr int32_t NOALLOC_MAX_POINTS = 1 << 18;  // 10 MiB of 40-byte projected points
static __device__ __align__(256) unsigned char scratch[(size_t)NOALLOC_MAX_POINTS * sizeof(Projected)];
static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratchLock;  // serialises set-up and every scratch-free call in the process
static void *g_scratchAddress[NOALLOC_MAX_DEVICES] = {};

// The calling thread's current device's copy of `scratch` (g_scratchLock held).
static int scratchAddress(Projected **out) {
	int device = 0;
	HIP_TRY(hipGetDevice(&device));
	if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(DFPSR_HIP_ERROR_DEVICE, "model_render_depth_noalloc: device index beyond 63");
	if (!g_scratchAddress[device]) {
		void *ptr = nullptr;
		HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
		if (!ptr) return fail(DFPSR_HIP_ERROR_DEVICE, "model_render_depth_noalloc: no device address for the scratch array");
		g_scratchAddress[device] = ptr;
	}
	*out = (Projected *)g_scratchAddress[device];
	return DFPSR_HIP_OK;
}

// Tile shape of the depth-only frames: 32 x 16 pixels.
static constexpr int32_t DEPTH_TILE_CONFIG = 1;

// model_renderDepth through the tiled pipeline: one frame of one task (the model, its
// transform and the camera) on the depth canvas alone, drawn by the depth-only tile kernel.
// The frame, its buffers and its stream belong to the calling thread: threads drawing one model
// into their own canvases never share working memory, and a call waits for its frame before
// the same thread's next call reuses it.
static int renderDepthTiled(dfpsr_hip_model *m, const float transform[12], const dfpsr_hip_image *depth, const CameraParams &cam) {
	struct DepthFrame {
		Frame frame;
		DevBuf stage;  // a host canvas's device copy
		bool ready = false;
	};
	static thread_local DepthFrame own;
	if (!own.ready) {
		int rc = own.frame.init();
		if (rc) return rc;
		own.ready = true;
	}
	const size_t canvasBytes = (size_t)depth->stride_bytes * depth->height;
	float *pixels = (float *)depth->pixels;
	if (!depth->on_device) {
		if (own.stage.ensure(canvasBytes) != hipSuccess) return fail(DFPSR_HIP_ERROR_MEMORY, "model_render_depth: device allocation failed");
		pixels = own.stage.as<float>();
		HIP_TRY(hipMemcpy(pixels, depth->pixels, canvasBytes, hipMemcpyHostToDevice));
	}
	Frame &fr = own.frame;
	fr.reset(DEPTH_TILE_CONFIG);
	CanvasHost c;
	c.color = nullptr;
	c.depth = pixels;
	c.width = depth->width;
	c.height = depth->height;
	c.colorStride = 0;
	c.depthStride = depth->stride_bytes / 4;
	const int32_t slot = fr.addCanvas(c);
	ModelView v = m->view();
	v.filter = 0;  // every polygon is drawn whatever the model's filter (Model.cpp:164-210)
	int rc = fr.addTask(v, transform, cam, slot, POINTS_MODEL);
	if (rc) return rc;
	RenderOptions o;
	o.tileConfig = DEPTH_TILE_CONFIG;
	o.mode = 2;
	rc = fr.render(o);
	if (rc) return rc;
	if (!depth->on_device) HIP_TRY(hipMemcpy(depth->pixels, pixels, canvasBytes, hipMemcpyDeviceToHost));
	return DFPSR_HIP_OK;
}

// model_renderDepth (modelAPI.cpp:202-206): box cull, project the points, draw every fan
// triangle's depth.  Synchronous like the reference: the canvas is final on return.
// noAlloc = false: the tiled pipeline (renderDepthTiled).  noAlloc = true: the projected
// points go to the shared scratch array (at most NOALLOC_MAX_POINTS), the canvas must be
// device memory and the model resident, nothing is allocated, and each fan triangle is drawn
// by its own lanes with a compare-and-swap per pixel; the lock keeps two calls in one
// process from using the array at once.
static int renderDepth(const dfpsr_hip_model *cm, const float transform[12], const dfpsr_hip_image *depth,
                       const dfpsr_hip_camera *camera, const char *who, bool noAlloc) {
	if (!cm || !transform) return fail(DFPSR_HIP_ERROR_ARGUMENT, std::string(who) + ": null argument");
	if (!hasImage(depth)) return DFPSR_HIP_OK;  // no target buffer: nothing to draw (renderCore.cpp:409)
	if (depth->width < 1 || depth->height < 1 || depth->stride_bytes < depth->width * 4 || (depth->stride_bytes & 3))
		return fail(DFPSR_HIP_ERROR_ARGUMENT, std::string(who) + ": bad canvas size or stride");
	if (noAlloc && !depth->on_device)
		return fail(DFPSR_HIP_ERROR_ARGUMENT, std::string(who) + ": the canvas must be device memory (a host canvas needs a staging allocation)");
	int rc = checkCamera(camera);
	if (rc) return rc;
	CameraParams cam = cameraFrom(camera);
	const Transform modelToWorld = transformFrom(transform);
	rc = checkModelIndices(cm, who);
	if (rc) return rc;
	const int32_t pointCount = cm->pointCount(), fanCount = (int32_t)cm->fanPolygon.size();
	if (noAlloc && pointCount > NOALLOC_MAX_POINTS)
		return fail(DFPSR_HIP_ERROR_ARGUMENT, std::string(who) + ": the model has " + std::to_string(pointCount) +
		                                          " points, more than the scratch array's " + std::to_string(NOALLOC_MAX_POINTS));
	if (boxSeen(cam, cm->minBound, cm->maxBound, modelToWorld) == 0) return DFPSR_HIP_OK;
	if (pointCount == 0 || fanCount == 0) return DFPSR_HIP_OK;
	dfpsr_hip_model *m = const_cast<dfpsr_hip_model *>(cm);
	if (!noAlloc) {
		rc = m->upload();
		if (rc) return rc;
		return renderDepthTiled(m, transform, depth, cam);
	}
	if (m->dirty)
		return fail(DFPSR_HIP_ERROR_STATE, std::string(who) + ": the model has changes that are not on the device yet "
		                                                      "(any allocating draw of it uploads them)");
	const int32_t groupShift = depthGroupShift();
	const int64_t drawThreads = (int64_t)fanCount << groupShift;
	if (drawThreads > INT32_MAX - 255) return fail(DFPSR_HIP_ERROR_ARGUMENT, std::string(who) + ": too many polygons for one launch");
	std::lock_guard<std::mutex> hold(g_scratchLock);
	Projected *projected = nullptr;
	rc = scratchAddress(&projected);
	if (rc) return rc;
	DepthCanvas cv;
	cv.depth = (float *)depth->pixels;
	cv.width = depth->width;
	cv.height = depth->height;
	cv.stride = depth->stride_bytes / 4;
	// Every projected point the draw reads ([0, pointCount)) is written first: nothing in the
	// shared array needs clearing between calls.
	hipLaunchKernelGGL(k_depth_project, dim3((pointCount + 255) / 256), dim3(256), 0, 0, m->dPoints.as<float>(), pointCount,
	                   modelToWorld, cam, projected);
	HIP_TRY(hipGetLastError());
	hipLaunchKernelGGL(k_depth_draw, dim3((uint32_t)((drawThreads + 255) / 256)), dim3(256), 0, 0, projected,
	                   m->dFans.as<FanRecord>(), fanCount, groupShift, cam, cv);
	HIP_TRY(hipGetLastError());
	HIP_TRY(hipDeviceSynchronize());
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_model_render_depth(const dfpsr_hip_model *model, const float transform[12], const dfpsr_hip_image *depth,
                                            const dfpsr_hip_camera *camera) {
	return renderDepth(model, transform, depth, camera, "model_render_depth", false);
}

extern "C" int dfpsr_hip_model_render_depth_noalloc(const dfpsr_hip_model *model, const float transform[12],
                                                    const dfpsr_hip_image *depth, const dfpsr_hip_camera *camera) {
	return renderDepth(model, transform, depth, camera, "model_render_depth_noalloc", true);
}

extern "C" int dfpsr_hip_renderer_create(dfpsr_hip_renderer **out) {
	if (!out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_create: null output");
	std::unique_ptr<dfpsr_hip_renderer> r(new (std::nothrow) dfpsr_hip_renderer());
	if (!r) return fail(DFPSR_HIP_ERROR_MEMORY, "renderer_create: out of host memory");
	int rc = r->frame.init();
	if (rc != DFPSR_HIP_OK) return rc;
	*out = r.release();
	return DFPSR_HIP_OK;
}

extern "C" void dfpsr_hip_renderer_destroy(dfpsr_hip_renderer *r) { delete r; }

extern "C" int dfpsr_hip_renderer_begin(dfpsr_hip_renderer *r, const dfpsr_hip_image *color, const dfpsr_hip_image *depth) {
	if (!r) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_begin: null renderer");
	if (r->receiving) return fail(DFPSR_HIP_ERROR_STATE, "Called renderer_begin on the same renderer twice without ending the previous batch!");
	bool hc = hasImage(color), hd = hasImage(depth);
	for (const dfpsr_hip_image *im : {color, depth}) {
		if (hasImage(im) && (im->width < 1 || im->height < 1 || im->stride_bytes < im->width * 4 || (im->stride_bytes & 3)))
			return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_begin: bad canvas size or stride");
	}
	if (hc && hd && (color->width != depth->width || color->height != depth->height))
		return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_begin: colour and depth canvases differ in size");
	r->color = hc ? *color : dfpsr_hip_image{};
	r->depth = hd ? *depth : dfpsr_hip_image{};
	r->width = hc ? color->width : (hd ? depth->width : 0);
	r->height = hc ? color->height : (hd ? depth->height : 0);
	// Host canvases are staged on the device for the frame.
	if (hc && !color->on_device) {
		size_t bytes = (size_t)color->stride_bytes * color->height;
		if (r->colorStage.ensure(bytes) != hipSuccess) return fail(DFPSR_HIP_ERROR_MEMORY, "renderer_begin: device allocation failed");
		HIP_TRY(hipMemcpy(r->colorStage.ptr, color->pixels, bytes, hipMemcpyHostToDevice));
	}
	if (hd && !depth->on_device) {
		size_t bytes = (size_t)depth->stride_bytes * depth->height;
		if (r->depthStage.ensure(bytes) != hipSuccess) return fail(DFPSR_HIP_ERROR_MEMORY, "renderer_begin: device allocation failed");
		HIP_TRY(hipMemcpy(r->depthStage.ptr, depth->pixels, bytes, hipMemcpyHostToDevice));
	}
	CanvasHost cv;
	cv.color = hc ? (uint32_t *)(color->on_device ? color->pixels : r->colorStage.ptr) : nullptr;
	cv.depth = hd ? (float *)(depth->on_device ? depth->pixels : r->depthStage.ptr) : nullptr;
	cv.width = r->width;
	cv.height = r->height;
	cv.colorStride = hc ? color->stride_bytes / 4 : 0;
	cv.depthStride = hd ? depth->stride_bytes / 4 : 0;
	r->frame.reset(r->options.tileConfig);
	r->frame.addCanvas(cv);
	r->triangleModels.clear();
	// One grid cell per 16 x 16 pixels, rounded up (rendererAPI.cpp:162-163).
	r->gridW = (r->width + (OCCLUSION_CELL - 1)) / OCCLUSION_CELL;
	r->gridH = (r->height + (OCCLUSION_CELL - 1)) / OCCLUSION_CELL;
	r->occluded = false;
	r->receiving = true;
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_renderer_give_model(dfpsr_hip_renderer *r, const dfpsr_hip_model *m, const float transform[12],
                                             const dfpsr_hip_camera *camera) {
	if (!r || !m || !transform) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_give_model: null argument");
	if (!r->receiving) return fail(DFPSR_HIP_ERROR_STATE, "Cannot call renderer_giveTask before renderer_begin!");
	int rc = checkCamera(camera);
	if (rc) return rc;
	rc = checkModelIndices(m, "renderer_give_model");
	if (rc) return rc;
	CameraParams cam = cameraFrom(camera);
	// Camera::isBoxSeen on the model bound (modelAPI.cpp:227).
	if (boxSeen(cam, m->minBound, m->maxBound, transformFrom(transform)) == 0) return DFPSR_HIP_OK;
	// Whole models behind the occluders given so far are skipped (modelAPI.cpp:229-233).
	if (r->occluded && boxOccluded(r->gridRead(), r->imageW, r->gridW, r->gridH, m->minBound, m->maxBound, transformFrom(transform), cam))
		return DFPSR_HIP_OK;
	rc = const_cast<dfpsr_hip_model *>(m)->upload();
	if (rc) return rc;
	return r->frame.addTask(m->view(), transform, cam, 0, POINTS_MODEL);
}

// The per-call model behind both triangle entries: `count` triangles of corners (3t, 3t + 1,
// 3t + 2) in one part, with the given corner colours and texture coordinates (3 x 4 floats per
// triangle).  The points are set by the caller.
static int triangleModel(int32_t count, const float *colors, const float *texcoords, const dfpsr_hip_texture *diffuse,
                         const dfpsr_hip_texture *light, int32_t filter, std::unique_ptr<dfpsr_hip_model> &out) {
	std::unique_ptr<dfpsr_hip_model> m(new (std::nothrow) dfpsr_hip_model());
	if (!m) return fail(DFPSR_HIP_ERROR_MEMORY, "renderer_give_triangles: out of host memory");
	m->filter = filter;
	m->indices.resize((size_t)count * 4);
	m->colors.assign((size_t)count * 16, 0.0f);
	m->texcoords.assign((size_t)count * 16, 0.0f);
	for (int32_t t = 0; t < count; t++) {
		int32_t *ix = &m->indices[(size_t)t * 4];
		ix[0] = t * 3; ix[1] = t * 3 + 1; ix[2] = t * 3 + 2; ix[3] = -1;
		std::memcpy(&m->colors[(size_t)t * 16], colors + (size_t)t * 12, 12 * sizeof(float));
		std::memcpy(&m->texcoords[(size_t)t * 16], texcoords + (size_t)t * 12, 12 * sizeof(float));
		m->fanPolygon.push_back(t * 2);
		m->polygonPart.push_back(0);
	}
	m->maxIndex = count * 3 - 1;
	m->partDiffuse.push_back(diffuse ? &diffuse->desc : nullptr);
	m->partLight.push_back(light ? &light->desc : nullptr);
	out = std::move(m);
	return DFPSR_HIP_OK;
}

static int checkTriangleCall(dfpsr_hip_renderer *r, int32_t count, const void *corners, const float *colors,
                             const float *texcoords, int32_t filter, const dfpsr_hip_camera *camera, const char *who) {
	if (!r || count < 0 || (count > 0 && (!corners || !colors || !texcoords)))
		return fail(DFPSR_HIP_ERROR_ARGUMENT, std::string(who) + ": bad argument");
	if (count > INT32_MAX / 8) return fail(DFPSR_HIP_ERROR_ARGUMENT, std::string(who) + ": too many triangles for one call");
	if (!r->receiving) return fail(DFPSR_HIP_ERROR_STATE, "Cannot call renderer_giveTask_triangle before renderer_begin!");
	if (filter != DFPSR_HIP_FILTER_SOLID && filter != DFPSR_HIP_FILTER_ALPHA) return fail(DFPSR_HIP_ERROR_ARGUMENT, "unknown filter");
	return checkCamera(camera);
}

extern "C" int dfpsr_hip_renderer_give_triangles(dfpsr_hip_renderer *r, int32_t count, const float *cs, const float *colors,
                                                 const float *texcoords, const dfpsr_hip_texture *diffuse,
                                                 const dfpsr_hip_texture *light, int32_t filter, const dfpsr_hip_camera *camera) {
	int rc = checkTriangleCall(r, count, cs, colors, texcoords, filter, camera, "renderer_give_triangles");
	if (rc || count == 0) return rc;
	std::unique_ptr<dfpsr_hip_model> m;
	rc = triangleModel(count, colors, texcoords, diffuse, light, filter, m);
	if (rc) return rc;
	m->points.assign(cs, cs + (size_t)count * 9);
	rc = m->upload();
	if (rc) return rc;
	static const float identity[12] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1};
	rc = r->frame.addTask(m->view(), identity, cameraFrom(camera), 0, POINTS_CAMERA_SPACE);
	if (rc) return rc;
	r->triangleModels.push_back(std::move(m));
	return DFPSR_HIP_OK;
}

static_assert(sizeof(dfpsr_hip_projected_point) == sizeof(Projected), "ProjectedPoint layout");
static_assert(offsetof(dfpsr_hip_projected_point, is) == offsetof(Projected, is), "ProjectedPoint layout");
static_assert(offsetof(dfpsr_hip_projected_point, flat) == offsetof(Projected, flat), "ProjectedPoint layout");

extern "C" int dfpsr_hip_renderer_give_projected_triangles(dfpsr_hip_renderer *r, int32_t count,
                                                           const dfpsr_hip_projected_point *corners, const float *colors,
                                                           const float *texcoords, const dfpsr_hip_texture *diffuse,
                                                           const dfpsr_hip_texture *light, int32_t filter,
                                                           const dfpsr_hip_camera *camera) {
	int rc = checkTriangleCall(r, count, corners, colors, texcoords, filter, camera, "renderer_give_projected_triangles");
	if (rc || count == 0) return rc;
	std::unique_ptr<dfpsr_hip_model> m;
	rc = triangleModel(count, colors, texcoords, diffuse, light, filter, m);
	if (rc) return rc;
	// The 40-byte records travel as they are: the projection kernel copies them in place of
	// projecting (POINTS_PROJECTED), so the setup sees exactly the caller's cs, is and flat.
	const float *raw = reinterpret_cast<const float *>(corners);
	m->points.assign(raw, raw + (size_t)count * 3 * (sizeof(Projected) / sizeof(float)));
	m->pointFloats = sizeof(Projected) / sizeof(float);
	rc = m->upload();
	if (rc) return rc;
	static const float identity[12] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1};
	rc = r->frame.addTask(m->view(), identity, cameraFrom(camera), 0, POINTS_PROJECTED);
	if (rc) return rc;
	r->triangleModels.push_back(std::move(m));
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_renderer_takes_triangles(const dfpsr_hip_renderer *r, int32_t *out) {
	if (!r || !out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_takes_triangles: null argument");
	*out = r->receiving ? 1 : 0;
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_renderer_end(dfpsr_hip_renderer *r) {
	if (!r) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_end: null renderer");
	if (!r->receiving) return fail(DFPSR_HIP_ERROR_STATE, "Called renderer_end without renderer_begin!");
	r->receiving = false;
	// completeOcclusion (rendererAPI.cpp:190-215) runs inside the setup kernels: with occluders,
	// each draw command is tested against the frame's final grid.
	if (r->occluded) {
		int rc = r->uploadGrid();
		if (rc) return rc;
		r->frame.setOcclusion(r->dGrid.as<float>(), r->gridW, r->gridH);
	}
	int rc = r->frame.render(r->options);
	if (rc) return rc;
	if (hasImage(&r->color) && !r->color.on_device)
		HIP_TRY(hipMemcpy(r->color.pixels, r->colorStage.ptr, (size_t)r->color.stride_bytes * r->color.height, hipMemcpyDeviceToHost));
	if (hasImage(&r->depth) && !r->depth.on_device)
		HIP_TRY(hipMemcpy(r->depth.pixels, r->depthStage.ptr, (size_t)r->depth.stride_bytes * r->depth.height, hipMemcpyDeviceToHost));
	r->triangleModels.clear();
	return DFPSR_HIP_OK;
}

// ---- Occlusion (rendererAPI.h:77-95) ----

static int checkBoxCall(const dfpsr_hip_renderer *r, const float *mn, const float *mx, const float *transform,
                        const dfpsr_hip_camera *camera, const char *who) {
	if (!r || !mn || !mx || !transform) return fail(DFPSR_HIP_ERROR_ARGUMENT, std::string(who) + ": null argument");
	if (!r->receiving) return fail(DFPSR_HIP_ERROR_STATE, std::string("Cannot call ") + who + " without first calling renderer_begin!");
	return checkCamera(camera);
}

extern "C" int dfpsr_hip_renderer_occlude_from_box(dfpsr_hip_renderer *r, const float minimum[3], const float maximum[3],
                                                   const float transform[12], const dfpsr_hip_camera *camera) {
	int rc = checkBoxCall(r, minimum, maximum, transform, camera, "renderer_occludeFromBox");
	if (rc) return rc;
	r->prepareForOcclusion();
	occludeFromBox(r->gridImage.data(), r->imageW, r->gridW, r->gridH, make3(minimum[0], minimum[1], minimum[2]),
	               make3(maximum[0], maximum[1], maximum[2]), transformFrom(transform), cameraFrom(camera));
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_renderer_occlude_from_top_rows(dfpsr_hip_renderer *r, const dfpsr_hip_camera *camera) {
	if (!r) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_occludeFromTopRows: null renderer");
	if (!r->receiving) return fail(DFPSR_HIP_ERROR_STATE, "Cannot call renderer_occludeFromTopRows without first calling renderer_begin!");
	if (!hasImage(&r->depth))
		return fail(DFPSR_HIP_ERROR_STATE, "Cannot call renderer_occludeFromTopRows without having given a depth buffer in renderer_begin!");
	int rc = checkCamera(camera);
	if (rc) return rc;
	r->prepareForOcclusion();
	const int32_t cells = r->gridW * r->gridH;
	if (cells == 0) return DFPSR_HIP_OK;
	rc = r->uploadGrid();
	if (rc) return rc;
	hipLaunchKernelGGL(k_occlude_top_rows, dim3((cells + 255) / 256), dim3(256), 0, r->frame.stream, r->deviceCanvasDepth(),
	                   r->depth.width, r->depth.height, r->depth.stride_bytes / 4, r->dGrid.as<float>(), r->gridW, r->gridH,
	                   camera->perspective ? 1 : 0);
	HIP_TRY(hipGetLastError());
	return r->downloadGrid();
}

extern "C" int dfpsr_hip_renderer_occlude_from_existing_triangles(dfpsr_hip_renderer *r) {
	if (!r) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_occludeFromExistingTriangles: null renderer");
	if (!r->receiving)
		return fail(DFPSR_HIP_ERROR_STATE, "Cannot call renderer_occludeFromExistingTriangles without first calling renderer_begin!");
	r->prepareForOcclusion();
	if (r->gridW * r->gridH == 0) return DFPSR_HIP_OK;
	int rc = r->uploadGrid();
	if (rc) return rc;
	rc = r->frame.occludeExisting(r->dGrid.as<float>(), r->gridW, r->gridH);
	if (rc) return rc;
	return r->downloadGrid();
}

extern "C" int dfpsr_hip_renderer_is_box_visible(const dfpsr_hip_renderer *r, const float minimum[3], const float maximum[3],
                                                 const float transform[12], const dfpsr_hip_camera *camera, int32_t *out) {
	int rc = checkBoxCall(r, minimum, maximum, transform, camera, "renderer_isBoxVisible");
	if (rc) return rc;
	if (!out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_isBoxVisible: null output");
	*out = boxOccluded(r->gridRead(), r->imageW, r->gridW, r->gridH, make3(minimum[0], minimum[1], minimum[2]),
	                   make3(maximum[0], maximum[1], maximum[2]), transformFrom(transform), cameraFrom(camera)) ? 0 : 1;
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_renderer_has_occluders(const dfpsr_hip_renderer *r, int32_t *out) {
	if (!r || !out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_hasOccluders: null argument");
	*out = r->occluded ? 1 : 0;
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_renderer_occlusion_grid(const dfpsr_hip_renderer *r, float *out, int32_t *width, int32_t *height) {
	if (!r || !width || !height) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_occlusion_grid: null argument");
	*width = r->gridW;
	*height = r->gridH;
	if (out)
		for (int32_t y = 0; y < r->gridH; y++)
			for (int32_t x = 0; x < r->gridW; x++)
				out[(size_t)y * r->gridW + x] = r->gridImage.empty() ? 0.0f : r->gridImage[(size_t)y * r->imageW + x];
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_renderer_stats(const dfpsr_hip_renderer *r, int64_t out[4]) {
	if (!r || !out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_stats: null argument");
	for (int i = 0; i < 4; i++) out[i] = r->frame.stats[i];
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_renderer_set_timing(dfpsr_hip_renderer *r, int32_t enable) {
	if (!r) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_set_timing: null renderer");
	r->frame.timing = enable != 0;
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_renderer_timing(const dfpsr_hip_renderer *r, float out[5]) {
	if (!r || !out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_timing: null argument");
	for (int i = 0; i < 5; i++) out[i] = r->frame.stageMs[i];
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_renderer_set_tuning(dfpsr_hip_renderer *r, const dfpsr_hip_tuning *t) {
	if (!r) return fail(DFPSR_HIP_ERROR_ARGUMENT, "renderer_set_tuning: null renderer");
	int rc = checkTuning(t);
	if (rc) return rc;
	if (r->receiving) return fail(DFPSR_HIP_ERROR_STATE, "renderer_set_tuning: cannot change tuning between begin and end");
	r->options = optionsFrom(t);
	return DFPSR_HIP_OK;
}

// ---------------------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------------------

struct dfpsr_hip_batch {
	Frame frame;
	std::vector<dfpsr_hip_job> lastJobs;
	int32_t lastTileConfig = -1;
};

extern "C" int dfpsr_hip_batch_create(dfpsr_hip_batch **out) {
	if (!out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "batch_create: null output");
	std::unique_ptr<dfpsr_hip_batch> b(new (std::nothrow) dfpsr_hip_batch());
	if (!b) return fail(DFPSR_HIP_ERROR_MEMORY, "batch_create: out of host memory");
	int rc = b->frame.init();
	if (rc) return rc;
	*out = b.release();
	return DFPSR_HIP_OK;
}

extern "C" void dfpsr_hip_batch_destroy(dfpsr_hip_batch *b) { delete b; }

extern "C" int dfpsr_hip_batch_render(dfpsr_hip_batch *b, const dfpsr_hip_job *jobs, int32_t count, const dfpsr_hip_tuning *tuning) {
	if (!b || count < 0 || (count > 0 && !jobs)) return fail(DFPSR_HIP_ERROR_ARGUMENT, "batch_render: bad argument");
	int rc = checkTuning(tuning);
	if (rc) return rc;
	RenderOptions o = optionsFrom(tuning);
	bool same = (int32_t)b->lastJobs.size() == count && b->lastTileConfig == o.tileConfig &&
	            (count == 0 || std::memcmp(b->lastJobs.data(), jobs, (size_t)count * sizeof(dfpsr_hip_job)) == 0);
	if (same) {
		for (int32_t i = 0; i < count; i++)
			if (jobs[i].model->dirty) { same = false; break; }
	}
	if (!same) {
		b->frame.reset(o.tileConfig);
		for (int32_t i = 0; i < count; i++) {
			const dfpsr_hip_job &j = jobs[i];
			bool hc = hasImage(&j.color), hd = hasImage(&j.depth);
			if ((hc && !j.color.on_device) || (hd && !j.depth.on_device))
				return fail(DFPSR_HIP_ERROR_ARGUMENT, "batch_render: job canvases must be device memory");
			if (hc && hd && (j.color.width != j.depth.width || j.color.height != j.depth.height))
				return fail(DFPSR_HIP_ERROR_ARGUMENT, "batch_render: a job's colour and depth canvases differ in size");
			for (const dfpsr_hip_image *im : {&j.color, &j.depth})
				if (hasImage(im) && (im->width < 1 || im->height < 1 || im->stride_bytes < im->width * 4 || (im->stride_bytes & 3)))
					return fail(DFPSR_HIP_ERROR_ARGUMENT, "batch_render: bad canvas size or stride");
			if (!j.model) return fail(DFPSR_HIP_ERROR_ARGUMENT, "batch_render: null model");
			rc = checkModelIndices(j.model, "batch_render");
			if (rc) return rc;
			rc = checkCamera(&j.camera);
			if (rc) return rc;
			CanvasHost cv;
			cv.color = hc ? (uint32_t *)j.color.pixels : nullptr;
			cv.depth = hd ? (float *)j.depth.pixels : nullptr;
			cv.width = hc ? j.color.width : (hd ? j.depth.width : 0);
			cv.height = hc ? j.color.height : (hd ? j.depth.height : 0);
			cv.colorStride = hc ? j.color.stride_bytes / 4 : 0;
			cv.depthStride = hd ? j.depth.stride_bytes / 4 : 0;
			if (!hc && !hd) continue;
			int32_t slot = b->frame.addCanvas(cv);
			CameraParams cam = cameraFrom(&j.camera);
			if (boxSeen(cam, j.model->minBound, j.model->maxBound, transformFrom(j.transform)) == 0) continue;
			rc = const_cast<dfpsr_hip_model *>(j.model)->upload();
			if (rc) return rc;
			rc = b->frame.addTask(j.model->view(), j.transform, cam, slot, POINTS_MODEL);
			if (rc) return rc;
		}
		rc = b->frame.prepare();
		if (rc) return rc;
		b->lastJobs.assign(jobs, jobs + count);
		b->lastTileConfig = o.tileConfig;
	}
	rc = b->frame.run(o);
	if (rc) b->lastJobs.clear();
	return rc;
}

extern "C" int dfpsr_hip_batch_stats(const dfpsr_hip_batch *b, int64_t out[4]) {
	if (!b || !out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "batch_stats: null argument");
	for (int i = 0; i < 4; i++) out[i] = b->frame.stats[i];
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_batch_set_timing(dfpsr_hip_batch *b, int32_t enable) {
	if (!b) return fail(DFPSR_HIP_ERROR_ARGUMENT, "batch_set_timing: null batch");
	b->frame.timing = enable != 0;
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_batch_timing(const dfpsr_hip_batch *b, float out[5]) {
	if (!b || !out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "batch_timing: null argument");
	for (int i = 0; i < 5; i++) out[i] = b->frame.stageMs[i];
	return DFPSR_HIP_OK;
}

static const int32_t AUTOTUNE_PERSISTENT[DFPSR_HIP_AUTOTUNE_PERSISTENT_STEPS] = {0, 1, 2, 4, 8};

extern "C" int dfpsr_hip_batch_autotune(dfpsr_hip_batch *b, const dfpsr_hip_job *jobs, int32_t count, int32_t repeats,
                                        const dfpsr_hip_tuning *base, dfpsr_hip_tuning *best, float *ms_out) {
	if (!b || !best || repeats < 1 || repeats > 1000) return fail(DFPSR_HIP_ERROR_ARGUMENT, "batch_autotune: bad argument");
	int rc = checkTuning(base);
	if (rc) return rc;
	dfpsr_hip_tuning t = {};
	if (base) t = *base;
	t.use_graph = 1;
	float bestMs = 0.0f;
	std::vector<double> samples((size_t)repeats);
	for (int32_t tc = 0; tc < DFPSR_HIP_AUTOTUNE_TILE_CONFIGS; tc++) {
		for (int32_t p = 0; p < DFPSR_HIP_AUTOTUNE_PERSISTENT_STEPS; p++) {
			t.tile_config = tc;
			t.persistent_per_cu = AUTOTUNE_PERSISTENT[p];
			// Two untimed renders: the first (re)builds the tables and captures the graph.
			for (int w = 0; w < 2; w++) {
				rc = dfpsr_hip_batch_render(b, jobs, count, &t);
				if (rc) return rc;
			}
			for (int32_t r = 0; r < repeats; r++) {
				auto start = std::chrono::steady_clock::now();
				rc = dfpsr_hip_batch_render(b, jobs, count, &t);
				if (rc) return rc;
				samples[(size_t)r] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
			}
			std::sort(samples.begin(), samples.end());
			const float ms = (float)samples[(size_t)repeats / 2];
			if (ms_out) ms_out[tc * DFPSR_HIP_AUTOTUNE_PERSISTENT_STEPS + p] = ms;
			if ((tc == 0 && p == 0) || ms < bestMs) {
				bestMs = ms;
				*best = t;
			}
		}
	}
	return DFPSR_HIP_OK;
}

// ---------------------------------------------------------------------------------------
// Differentiable rendering
// ---------------------------------------------------------------------------------------

struct dfpsr_hip_diff {
	Frame frame;
	std::vector<dfpsr_hip_view> lastViews;
	const dfpsr_hip_model *lastModel = nullptr;
	int32_t lastTileConfig = -1;
};

extern "C" int dfpsr_hip_diff_create(dfpsr_hip_diff **out) {
	if (!out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "diff_create: null output");
	std::unique_ptr<dfpsr_hip_diff> d(new (std::nothrow) dfpsr_hip_diff());
	if (!d) return fail(DFPSR_HIP_ERROR_MEMORY, "diff_create: out of host memory");
	int rc = d->frame.init();
	if (rc) return rc;
	*out = d.release();
	return DFPSR_HIP_OK;
}

extern "C" void dfpsr_hip_diff_destroy(dfpsr_hip_diff *d) { delete d; }

extern "C" int dfpsr_hip_diff_forward(dfpsr_hip_diff *d, const dfpsr_hip_model *model, const dfpsr_hip_view *views, int32_t count,
                                      const float *texture, int32_t log2w, int32_t log2h, const float *multipliers,
                                      const dfpsr_hip_tuning *tuning) {
	if (!d || !model || count < 0 || (count > 0 && !views) || !texture)
		return fail(DFPSR_HIP_ERROR_ARGUMENT, "diff_forward: bad argument");
	if (log2w < 0 || log2w > 15 || log2h < 0 || log2h > 15) return fail(DFPSR_HIP_ERROR_ARGUMENT, "diff_forward: texture size");
	int rc = checkTuning(tuning);
	if (rc) return rc;
	RenderOptions o = optionsFrom(tuning);
	o.mode = 1;
	bool same = d->lastModel == model && !model->dirty && d->lastTileConfig == o.tileConfig && (int32_t)d->lastViews.size() == count &&
	            (count == 0 || std::memcmp(d->lastViews.data(), views, (size_t)count * sizeof(dfpsr_hip_view)) == 0);
	d->frame.setDifferentiable(texture, log2w, log2h, multipliers);
	if (!same) {
		rc = checkModelIndices(model, "diff_forward");
		if (rc) return rc;
		d->frame.reset(o.tileConfig);
		rc = const_cast<dfpsr_hip_model *>(model)->upload();
		if (rc) return rc;
		for (int32_t i = 0; i < count; i++) {
			const dfpsr_hip_view &v = views[i];
			if (!v.color || !v.visibility || !v.depth || v.width < 1 || v.height < 1 || v.stride_pixels < v.width)
				return fail(DFPSR_HIP_ERROR_ARGUMENT, "diff_forward: a view needs colour, visibility and depth buffers");
			rc = checkCamera(&v.camera);
			if (rc) return rc;
			CanvasHost cv;
			cv.color = nullptr;
			cv.depth = v.depth;
			cv.width = v.width;
			cv.height = v.height;
			cv.colorStride = 0;
			cv.depthStride = v.stride_pixels;
			cv.diffColor = v.color;
			cv.visibility = v.visibility;
			cv.gradColor = v.grad_color;
			cv.diffStride = v.stride_pixels;
			int32_t slot = d->frame.addCanvas(cv);
			CameraParams cam = cameraFrom(&v.camera);
			if (boxSeen(cam, model->minBound, model->maxBound, transformFrom(v.transform)) == 0) continue;
			rc = d->frame.addTask(model->view(), v.transform, cam, slot, POINTS_MODEL);
			if (rc) return rc;
		}
		rc = d->frame.prepare();
		if (rc) return rc;
		d->lastViews.assign(views, views + count);
		d->lastModel = model;
		d->lastTileConfig = o.tileConfig;
	}
	rc = d->frame.run(o);
	if (rc) d->lastViews.clear();
	return rc;
}

extern "C" int dfpsr_hip_diff_backward(dfpsr_hip_diff *d, float *gradTexture, float *gradMultipliers, const dfpsr_hip_tuning *tuning) {
	if (!d) return fail(DFPSR_HIP_ERROR_ARGUMENT, "diff_backward: null argument");
	int rc = checkTuning(tuning);
	if (rc) return rc;
	for (const dfpsr_hip_view &v : d->lastViews)
		if (!v.grad_color) return fail(DFPSR_HIP_ERROR_ARGUMENT, "diff_backward: every view needs grad_color");
	return d->frame.backward(gradTexture, gradMultipliers, optionsFrom(tuning));
}

extern "C" int dfpsr_hip_diff_stats(const dfpsr_hip_diff *d, int64_t out[4]) {
	if (!d || !out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "diff_stats: null argument");
	for (int i = 0; i < 4; i++) out[i] = d->frame.stats[i];
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_diff_set_timing(dfpsr_hip_diff *d, int32_t enable) {
	if (!d) return fail(DFPSR_HIP_ERROR_ARGUMENT, "diff_set_timing: null argument");
	d->frame.timing = enable != 0;
	return DFPSR_HIP_OK;
}

extern "C" int dfpsr_hip_diff_timing(const dfpsr_hip_diff *d, float out[5]) {
	if (!d || !out) return fail(DFPSR_HIP_ERROR_ARGUMENT, "diff_timing: null argument");
	for (int i = 0; i < 5; i++) out[i] = d->frame.stageMs[i];
	return DFPSR_HIP_OK;
}

// ---------------------------------------------------------------------------------------
// Test hook
// ---------------------------------------------------------------------------------------

extern "C" int dfpsr_hip_probe_edge_side(const int64_t *n, const int64_t *d, const int32_t *L, const int32_t *R, int32_t count,
                                         int32_t *out) {
	if (count < 0 || (count > 0 && (!n || !d || !L || !R || !out))) return fail(DFPSR_HIP_ERROR_ARGUMENT, "probe_edge_side: bad argument");
	for (int32_t i = 0; i < count; i++)
		if (d[i] == 0) return fail(DFPSR_HIP_ERROR_ARGUMENT, "probe_edge_side: zero divisor");
	if (count == 0) return DFPSR_HIP_OK;
	DevBuf buf;
	const size_t c = (size_t)count;
	if (buf.ensure(c * 28) != hipSuccess) return fail(DFPSR_HIP_ERROR_MEMORY, "probe_edge_side: device allocation failed");
	uint8_t *base = (uint8_t *)buf.ptr;
	int64_t *dn = (int64_t *)base, *dd = (int64_t *)(base + c * 8);
	int32_t *dl = (int32_t *)(base + c * 16), *dr = (int32_t *)(base + c * 20), *dout = (int32_t *)(base + c * 24);
	HIP_TRY(hipMemcpy(dn, n, c * 8, hipMemcpyHostToDevice));
	HIP_TRY(hipMemcpy(dd, d, c * 8, hipMemcpyHostToDevice));
	HIP_TRY(hipMemcpy(dl, L, c * 4, hipMemcpyHostToDevice));
	HIP_TRY(hipMemcpy(dr, R, c * 4, hipMemcpyHostToDevice));
	hipLaunchKernelGGL(k_probe_edge_side, dim3((count + 255) / 256), dim3(256), 0, 0, dn, dd, dl, dr, count, dout);
	HIP_TRY(hipGetLastError());
	HIP_TRY(hipMemcpy(out, dout, c * 4, hipMemcpyDeviceToHost));
	return DFPSR_HIP_OK;
}
