// Host-side frame of the polygon path: the tasks given between renderer_begin and
// renderer_end, their canvases and textures, the device buffers of the pipeline and the
// launch sequence (frame.hip).
#pragma once

#include "pipeline.h"

#include <string>
#include <vector>

namespace dfpsr {

extern thread_local std::string g_lastError;
int fail(int code, const std::string &message);

#define HIP_TRY(call)                                                                                    \
	do {                                                                                                 \
		hipError_t e_ = (call);                                                                          \
		if (e_ != hipSuccess) return ::dfpsr::fail(-3, std::string(#call) + ": " + hipGetErrorString(e_)); \
	} while (0)

// A growable device allocation.
struct DevBuf {
	void *ptr = nullptr;
	size_t bytes = 0;
	DevBuf() = default;
	DevBuf(const DevBuf &) = delete;
	DevBuf &operator=(const DevBuf &) = delete;
	~DevBuf() { release(); }
	hipError_t ensure(size_t need);  // keeps the allocation when large enough; contents are not kept
	void release();
	template <typename T> T *as() const { return (T *)ptr; }
};

struct CanvasHost {
	uint32_t *color;
	float *depth;
	int32_t width, height;
	int32_t colorStride, depthStride;  // in elements
};

// What a task needs of a model: device arrays and the textures of each part.
struct ModelView {
	const float *points;
	const int32_t *indices;
	const float *colors;
	const float *texcoords;
	const int32_t *fanPolygon;
	const int32_t *polygonPart;
	int32_t pointCount, fanCount, partCount, filter;
	const TextureDesc *const *partDiffuse;  // per part, null when absent
	const TextureDesc *const *partLight;
};

class Frame {
public:
	hipStream_t stream = nullptr;
	int64_t stats[4] = {0, 0, 0, 0};  // commands, tile pairs, tiles, passes

	~Frame();
	int init();
	void reset();
	int32_t addCanvas(const CanvasHost &canvas);
	int addTask(const ModelView &model, const float *transform, const CameraParams &camera, int32_t canvas, bool cameraSpacePoints);
	// Runs the whole pipeline on the frame's stream and waits for it.
	int render();

private:
	std::vector<TaskDev> tasks_;
	std::vector<int32_t> taskFanEnd_, taskPointEnd_;
	std::vector<CanvasDev> canvases_;
	std::vector<int32_t> canvasTileEnd_;
	std::vector<TextureDesc> textures_;
	std::vector<const TextureDesc *> textureKeys_;
	std::vector<int32_t> partTexture_;
	std::vector<size_t> partTextureOffset_;
	bool alphaAny_ = false;
	int32_t commandCapacity_ = 0, pairCapacity_ = 0;
	DevBuf dTasks_, dTaskFanEnd_, dTaskPointEnd_, dCanvases_, dCanvasTileEnd_, dTextures_, dPartTexture_;
	DevBuf dProjected_, dFanCount_, dBlockSums_, dCommands_, dCounters_, dTileCount_, dTileOffset_, dTileFill_, dTileList_;
	int32_t *hostCounters_ = nullptr;

	int32_t textureSlot(const TextureDesc *desc);
	int upload(const void *src, size_t bytes, DevBuf &dst);
};

}  // namespace dfpsr
