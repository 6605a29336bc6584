// C++14 host-side mirror of DFPSR's rendererAPI / modelAPI / textureAPI for the polygon path,
// over the C ABI of libdfpsr_hip.so (include/dfpsr_hip.h).  Header only.
//
// Names, argument meaning and call order follow the reference (api/rendererAPI.h,
// api/modelAPI.h, api/textureAPI.h, implementation/render/Camera.h), so application code
// like Source/SDK/terrain/main.cpp:120-176, :381-421 reads the same with `dsr_hip::` in
// place of `dsr::`.  Handles are reference counted like dsr::Handle; a failing call throws
// dsr_hip::Error (a std::runtime_error), as the reference's default throwError does.
#ifndef DFPSR_HIP_HPP
#define DFPSR_HIP_HPP

#include "../../include/dfpsr_hip.h"

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace dsr_hip {

struct Error : public std::runtime_error {
	int code;
	Error(int code, const std::string &message) : std::runtime_error(message), code(code) {}
};

inline void check(int code) {
	if (code != DFPSR_HIP_OK) throw Error(code, dfpsr_hip_last_error());
}

struct FVector3D {
	float x, y, z;
	FVector3D() : x(0.0f), y(0.0f), z(0.0f) {}
	FVector3D(float x, float y, float z) : x(x), y(y), z(z) {}
};
struct FVector4D {
	float x, y, z, w;
	FVector4D() : x(0.0f), y(0.0f), z(0.0f), w(0.0f) {}
	FVector4D(float x, float y, float z, float w) : x(x), y(y), z(z), w(w) {}
};
struct FMatrix3x3 {
	FVector3D xAxis, yAxis, zAxis;
	FMatrix3x3() : xAxis(1.0f, 0.0f, 0.0f), yAxis(0.0f, 1.0f, 0.0f), zAxis(0.0f, 0.0f, 1.0f) {}
	FMatrix3x3(const FVector3D &x, const FVector3D &y, const FVector3D &z) : xAxis(x), yAxis(y), zAxis(z) {}
};
// Transform3D (math/Transform3D.h:30-35): position and axis system.
struct Transform3D {
	FVector3D position;
	FMatrix3x3 transform;
	Transform3D() {}
	Transform3D(const FVector3D &position, const FMatrix3x3 &transform) : position(position), transform(transform) {}
	void pack(float out[12]) const {
		const float v[12] = {position.x, position.y, position.z, transform.xAxis.x, transform.xAxis.y, transform.xAxis.z,
		                     transform.yAxis.x, transform.yAxis.y, transform.yAxis.z, transform.zAxis.x, transform.zAxis.y,
		                     transform.zAxis.z};
		for (int i = 0; i < 12; i++) out[i] = v[i];
	}
};

enum class Filter { Solid = DFPSR_HIP_FILTER_SOLID, Alpha = DFPSR_HIP_FILTER_ALPHA };

// Camera::createPerspective / createOrthogonal (Camera.h:144-156); the library derives the
// frustums exactly as the reference does.
struct Camera {
	dfpsr_hip_camera c;
	static Camera createPerspective(const Transform3D &location, float imageWidth, float imageHeight, float widthSlope = 1.0f,
	                                float nearClip = 0.01f, float farClip = 1000.0f) {
		Camera cam;
		cam.c.perspective = 1;
		location.pack(cam.c.location);
		cam.c.image_width = imageWidth;
		cam.c.image_height = imageHeight;
		cam.c.width_slope = widthSlope;
		cam.c.near_clip = nearClip;
		cam.c.far_clip = farClip;
		return cam;
	}
	static Camera createOrthogonal(const Transform3D &location, float imageWidth, float imageHeight, float halfWidth) {
		Camera cam = createPerspective(location, imageWidth, imageHeight, halfWidth);
		cam.c.perspective = 0;
		return cam;
	}
};

// Images: a canvas in host or device memory (OrderedImageRgbaU8 / AlignedImageF32).
template <typename T>
struct Image {
	std::shared_ptr<std::vector<T>> host;  // owned host pixels, or null
	T *device = nullptr;                   // caller-owned device pixels, or null
	int32_t width = 0, height = 0, stride = 0;  // stride in elements
	T *data() const { return host ? host->data() : device; }
	dfpsr_hip_image abi() const {
		dfpsr_hip_image im;
		im.pixels = (void *)data();
		im.width = width;
		im.height = height;
		im.stride_bytes = stride * (int32_t)sizeof(T);
		im.on_device = device ? 1 : 0;
		return im;
	}
};
using ImageRgbaU8 = Image<uint32_t>;
using ImageF32 = Image<float>;

// image_create_RgbaU8 / image_create_F32 (api/imageAPI.cpp:40-72), zero-filled host canvases.
inline ImageRgbaU8 image_create_RgbaU8(int32_t width, int32_t height) {
	ImageRgbaU8 im;
	im.width = width;
	im.height = height;
	im.stride = width;
	im.host = std::make_shared<std::vector<uint32_t>>((size_t)width * height, 0u);
	return im;
}
inline ImageF32 image_create_F32(int32_t width, int32_t height) {
	ImageF32 im;
	im.width = width;
	im.height = height;
	im.stride = width;
	im.host = std::make_shared<std::vector<float>>((size_t)width * height, 0.0f);
	return im;
}
// A canvas over device memory the caller owns; the renderer draws in place.
template <typename T>
inline Image<T> image_fromDevice(T *pixels, int32_t width, int32_t height, int32_t strideElements) {
	Image<T> im;
	im.device = pixels;
	im.width = width;
	im.height = height;
	im.stride = strideElements;
	return im;
}

// ---- Textures ------------------------------------------------------------------------------

using TextureRgbaU8 = std::shared_ptr<dfpsr_hip_texture>;

inline TextureRgbaU8 texture_create_RgbaU8(int32_t width, int32_t height, int32_t resolutions) {
	dfpsr_hip_texture *t = nullptr;
	check(dfpsr_hip_texture_create(&t, width, height, resolutions));
	return TextureRgbaU8(t, dfpsr_hip_texture_destroy);
}
inline bool texture_exists(const TextureRgbaU8 &texture) { return texture != nullptr; }
// Writes the highest resolution, like texture_getSafePointer(texture, 0, y) row writes.
inline void texture_writeLevel0(const TextureRgbaU8 &texture, const uint32_t *pixels, int32_t strideTexels) {
	check(dfpsr_hip_texture_write_level0(texture.get(), pixels, strideTexels));
}
inline void texture_generatePyramid(const TextureRgbaU8 &texture) { check(dfpsr_hip_texture_generate_pyramid(texture.get())); }

// ---- Models --------------------------------------------------------------------------------

// The model is built on the host with the reference's per-polygon calls and uploaded to the
// device when first drawn (and again after a change).
struct ModelImpl {
	Filter filter = Filter::Solid;
	std::vector<float> points;
	struct Part {
		std::vector<int32_t> indices;
		std::vector<float> colors, texcoords;
		TextureRgbaU8 diffuse, light;
	};
	std::vector<Part> parts;
	std::shared_ptr<dfpsr_hip_model> device;
	bool dirty = true;

	dfpsr_hip_model *sync() {
		if (dirty || !device) {
			dfpsr_hip_model *m = nullptr;
			check(dfpsr_hip_model_create(&m, (int32_t)filter));
			device = std::shared_ptr<dfpsr_hip_model>(m, dfpsr_hip_model_destroy);
			check(dfpsr_hip_model_set_points(m, (int32_t)(points.size() / 3), points.data()));
			for (const Part &p : parts)
				check(dfpsr_hip_model_add_part(m, (int32_t)(p.indices.size() / 4), p.indices.data(), p.colors.data(),
				                               p.texcoords.data(), p.diffuse.get(), p.light.get()));
			dirty = false;
		}
		return device.get();
	}
};
using Model = std::shared_ptr<ModelImpl>;

inline Model model_create() { return std::make_shared<ModelImpl>(); }
inline void model_setFilter(const Model &model, Filter filter) { model->filter = filter; model->dirty = true; }
inline int32_t model_addPoint(const Model &model, const FVector3D &p) {
	model->points.push_back(p.x);
	model->points.push_back(p.y);
	model->points.push_back(p.z);
	model->dirty = true;
	return (int32_t)(model->points.size() / 3) - 1;
}
inline int32_t model_addEmptyPart(Model &model, const std::string &) {
	model->parts.emplace_back();
	model->dirty = true;
	return (int32_t)model->parts.size() - 1;
}
// Polygon defaults of the reference (Model.cpp:74-104): white, texture corners (0,0), (1,0), (1,1), (0,1).
inline int32_t addPolygon(Model &model, int32_t part, int32_t a, int32_t b, int32_t c, int32_t d) {
	ModelImpl::Part &p = model->parts.at((size_t)part);
	const int32_t idx[4] = {a, b, c, d};
	static const float tc[16] = {0, 0, 0, 0, 1, 0, 1, 0, 1, 1, 1, 1, 0, 1, 0, 1};
	for (int i = 0; i < 4; i++) p.indices.push_back(idx[i]);
	for (int i = 0; i < 16; i++) p.colors.push_back(1.0f);
	for (int i = 0; i < 16; i++) p.texcoords.push_back(tc[i]);
	model->dirty = true;
	return (int32_t)(p.indices.size() / 4) - 1;
}
inline int32_t model_addTriangle(Model &model, int32_t part, int32_t a, int32_t b, int32_t c) { return addPolygon(model, part, a, b, c, -1); }
inline int32_t model_addQuad(Model &model, int32_t part, int32_t a, int32_t b, int32_t c, int32_t d) { return addPolygon(model, part, a, b, c, d); }
inline void model_setVertexColor(Model &model, int32_t part, int32_t polygon, int32_t vertex, const FVector4D &color) {
	float *dst = &model->parts.at((size_t)part).colors.at((size_t)polygon * 16 + (size_t)vertex * 4);
	dst[0] = color.x; dst[1] = color.y; dst[2] = color.z; dst[3] = color.w;
	model->dirty = true;
}
inline void model_setTexCoord(Model &model, int32_t part, int32_t polygon, int32_t vertex, const FVector4D &texCoord) {
	float *dst = &model->parts.at((size_t)part).texcoords.at((size_t)polygon * 16 + (size_t)vertex * 4);
	dst[0] = texCoord.x; dst[1] = texCoord.y; dst[2] = texCoord.z; dst[3] = texCoord.w;
	model->dirty = true;
}
inline void model_setDiffuseMap(Model &model, int32_t part, const TextureRgbaU8 &texture) { model->parts.at((size_t)part).diffuse = texture; model->dirty = true; }
inline void model_setLightMap(Model &model, int32_t part, const TextureRgbaU8 &texture) { model->parts.at((size_t)part).light = texture; model->dirty = true; }

// ---- Renderer ------------------------------------------------------------------------------

struct RendererImpl {
	std::shared_ptr<dfpsr_hip_renderer> impl;
	std::vector<Model> drawn;  // models given this frame stay alive until renderer_end
};
using Renderer = std::shared_ptr<RendererImpl>;

inline Renderer renderer_create() {
	dfpsr_hip_renderer *r = nullptr;
	check(dfpsr_hip_renderer_create(&r));
	Renderer out = std::make_shared<RendererImpl>();
	out->impl = std::shared_ptr<dfpsr_hip_renderer>(r, dfpsr_hip_renderer_destroy);
	return out;
}
inline bool renderer_exists(const Renderer &renderer) { return renderer != nullptr; }
inline void renderer_begin(Renderer &renderer, ImageRgbaU8 &colorBuffer, ImageF32 &depthBuffer) {
	dfpsr_hip_image c = colorBuffer.abi(), d = depthBuffer.abi();
	check(dfpsr_hip_renderer_begin(renderer->impl.get(), &c, &d));
}
inline bool renderer_takesTriangles(const Renderer &renderer) {
	int32_t v = 0;
	check(dfpsr_hip_renderer_takes_triangles(renderer->impl.get(), &v));
	return v != 0;
}
// renderer_giveTask(renderer, model, modelToWorld, camera) (api/modelAPI.h:299-301).
inline void renderer_giveTask(Renderer &renderer, const Model &model, const Transform3D &modelToWorld, const Camera &camera) {
	float t[12];
	modelToWorld.pack(t);
	check(dfpsr_hip_renderer_give_model(renderer->impl.get(), model->sync(), t, &camera.c));
	renderer->drawn.push_back(model);
}
inline void renderer_end(Renderer &renderer) {
	check(dfpsr_hip_renderer_end(renderer->impl.get()));
	renderer->drawn.clear();
}

}  // namespace dsr_hip

#endif
