// Host-side frame of the polygon path: the tasks given between renderer_begin and
// renderer_end (or the jobs of a batch), their canvases and textures, the device buffers of
// the pipeline and the launch sequence (frame.hip).
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
	uint64_t generation = 0;  // counts the allocations: a new one holds nothing of the old
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
	// Differentiable canvases (null otherwise): float RGBA, visibility ids, incoming gradient.
	float *diffColor = nullptr;
	int32_t *visibility = nullptr;
	const float *gradColor = nullptr;
	int32_t diffStride = 0;
};

// What a task needs of a model: device arrays and the textures of each part.
struct ModelView {
	const float *points;
	const int32_t *indices;
	const float *colors;
	const float *texcoords;
	const FanRecord *fans;
	const int32_t *polygonPart;
	int32_t pointCount, fanCount, partCount, filter;
	int32_t maxIndex;  // the largest point index of any polygon (-1 without polygons)
	const TextureDesc *const *partDiffuse;  // per part, null when absent
	const TextureDesc *const *partLight;
};

// How a frame is run.
struct RenderOptions {
	int32_t tileConfig = 0;          // tile shape, see tileConfigW/H
	int32_t persistentPerCU = 0;     // 0: one raster workgroup per tile; k: k persistent workgroups per CU
	bool graph = false;              // capture the launch sequence once and replay it while nothing changes
	bool clear = false;              // fill the canvases before drawing (image_fill)
	uint32_t clearColor = 0;
	float clearDepth = 0.0f;
	int32_t mode = 0;                // 0: the reference's RGBA8 shaders; 1: the differentiable float shader;
	                                 // 2: depth only (model_renderDepth), one task, depth canvases only
	int32_t rasterPath = 0;          // 0: visibility raster where the frame allows it; 1: always in order
	int32_t binning = 0;             // 0: direct per-tile bins while they fit; 1: counted, scanned and filled lists
};

class Frame {
public:
	hipStream_t stream = nullptr;
	int64_t stats[4] = {0, 0, 0, 0};     // commands, tile pairs, tiles, passes
	float stageMs[5] = {0, 0, 0, 0, 0};  // projection, setup, binning, raster, whole pass (graph replay:
	                                     // 0, 0, everything before the raster, raster, whole)
	bool timing = false;

	~Frame();
	int init();
	// Starts a new frame whose canvases are cut into tiles of the given configuration.
	void reset(int32_t tileConfig = 0);
	int32_t addCanvas(const CanvasHost &canvas);
	int addTask(const ModelView &model, const float *transform, const CameraParams &camera, int32_t canvas, int32_t pointKind);
	// Packs and uploads the frame's tables, then runs the pipeline and waits for it.
	int render(const RenderOptions &options) {
		int rc = prepare();
		return rc ? rc : run(options);
	}
	// Packs the tables and uploads them when they changed since the last upload.
	int prepare();
	// Runs the pipeline over the prepared tables on the frame's stream and waits for it.
	int run(const RenderOptions &options);
	int32_t canvasCount() const { return (int32_t)canvases_.size(); }
	// Inputs of the differentiable shader: float RGBA texture (level 0, device) and per-point RGB
	// multipliers (device, null = all 1).
	void setDifferentiable(const float *texels, int32_t log2w, int32_t log2h, const float *multipliers) {
		ftex_ = texels;
		ftexLog2w_ = log2w;
		ftexLog2h_ = log2h;
		multipliers_ = multipliers;
	}
	// Adjoint of the last differentiable run: adds into the gradient buffers (either may be null).
	int backward(float *gradTexture, float *gradMultipliers, const RenderOptions &options);
	// Occlusion grid (device, gridW x gridH cells) the next run drops covered commands with;
	// null turns the test off.  reset() turns it off.
	void setOcclusion(float *deviceGrid, int32_t gridW, int32_t gridH) {
		occlusionGrid_ = deviceGrid;
		gridW_ = gridW;
		gridH_ = gridH;
	}
	// renderer_occludeFromExistingTriangles: projects the tasks given so far and lowers the device
	// grid under the hulls of their Solid draw commands.  Waits for the device.
	int occludeExisting(float *deviceGrid, int32_t gridW, int32_t gridH);

private:
	int32_t tileConfig_ = 0;
	std::vector<TaskDev> tasks_;
	std::vector<int32_t> taskFanEnd_, taskPointEnd_;
	std::vector<CanvasDev> canvases_;
	std::vector<int32_t> canvasTileEnd_;
	std::vector<TextureDesc> textures_;
	std::vector<const TextureDesc *> textureKeys_;
	std::vector<int32_t> partTexture_;
	std::vector<size_t> partTextureOffset_;
	bool alphaAny_ = false;
	bool orthogonalAny_ = false;  // a task with an orthogonal camera
	bool oddCanvas_ = false;      // a canvas of odd height
	bool wideCanvas_ = false;     // a canvas wider than the visibility raster's 16-bit row columns
	int32_t commandCapacity_ = 0, pairCapacity_ = 0;
	// Entries per tile of the direct bins; grows to the longest list seen.  Bins are used while
	// tiles x binCapacity_ stays within BIN_ENTRY_LIMIT, the scanned lists otherwise.  The bound
	// also sends frames with very long tile lists to the packed lists, which measured faster for
	// them (256 views of 256x256 with lists of several hundred commands: a forward pass of
	// 2.7 ms with bins against 1.7 ms packed).
	int32_t binCapacity_ = 64;
	// Single-pass setup: command slots per setup block (every block may write that many commands
	// from blockIdx * blockCapacity_ on).  Grows to the most commands a block has made; frames
	// whose slots would pass SINGLE_SLOT_LIMIT, and differentiable frames (whose visibility ids
	// count commands densely), take the two-pass setup.
	int32_t blockCapacity_ = SCAN_BLOCK;
	static constexpr int64_t SINGLE_SLOT_LIMIT = (int64_t)1 << 21;
	bool singlePass(const RenderOptions &o, int32_t fanBlocks) const {
		return o.mode != 1 && (int64_t)fanBlocks * blockCapacity_ <= SINGLE_SLOT_LIMIT;
	}
	static constexpr int64_t BIN_ENTRY_LIMIT = (int64_t)1 << 25;
	// All per-frame tables in one upload, skipped when identical to the previous frame's.
	std::vector<uint8_t> tables_, uploadedTables_;
	DevBuf dTables_;
	size_t offTasks_ = 0, offFanEnd_ = 0, offPointEnd_ = 0, offCanvases_ = 0, offTileEnd_ = 0, offTextures_ = 0;
	bool prepared_ = false;
	DevBuf dShaders_, dTriangleRows_;
	DevBuf dCanvasBase_, dFanCount_, dBlockSums_, dCommands_, dCounters_, dTileCount_, dTileOffset_, dTileFill_, dTileList_, dTileQueue_, dPairRecords_;
	int32_t *hostCounters_ = nullptr;
	int32_t *hostCountersDevice_ = nullptr;
	hipEvent_t events_[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
	hipGraph_t graph_ = nullptr;  // the whole sequence, or (timing on) everything before the raster
	hipGraphExec_t graphExec_ = nullptr;
	hipGraph_t rasterGraph_ = nullptr;  // timing on: the raster alone
	hipGraphExec_t rasterGraphExec_ = nullptr;
	std::vector<uint8_t> graphKey_;
	int32_t deviceCUs_ = 0;
	const float *ftex_ = nullptr;
	int32_t ftexLog2w_ = 0, ftexLog2h_ = 0;
	const float *multipliers_ = nullptr;
	float *occlusionGrid_ = nullptr;
	int32_t gridW_ = 0, gridH_ = 0;
	FrameDev lastRun_;
	bool lastRunValid_ = false;
	hipGraph_t bwdGraph_ = nullptr;
	hipGraphExec_t bwdGraphExec_ = nullptr;
	std::vector<uint8_t> bwdGraphKey_;

	int32_t textureSlot(const TextureDesc *desc);
	// parts: 1 = everything before the raster, 2 = the raster, 3 = both.
	int enqueue(const FrameDev &f, const RenderOptions &o, int32_t fanBlocks, int32_t tileBlocks, bool events, int parts,
	            bool skipBegin = false);
	// The last run left its counts zeroed (a self-cleaning raster that finished), for buffers of
	// these generations, this many tiles and canvases.
	bool clean_ = false;
	int32_t parity_ = 0;  // the set of counters and first-command indices the next frame uses
	uint64_t cleanGenerations_[3] = {0, 0, 0};
	int32_t cleanTiles_ = 0, cleanCanvases_ = 0;
	void dropGraph();
	// The visibility raster reproduces the in-order result when every command is Solid and
	// perspective and no canvas has an odd height.
	bool deferredRaster(const RenderOptions &o) const {
		return o.rasterPath == 0 && !alphaAny_ && !orthogonalAny_ && !oddCanvas_ && !wideCanvas_;
	}
};

}  // namespace dfpsr
