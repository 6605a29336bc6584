// Device-side data of one frame of the polygon path and the kernel entry points.
#pragma once

#include "shading.h"

namespace dfpsr {

// Screen tiles are TW x TH pixels with even sides, so a row pair and a 2x2 quad never
// straddle two tiles; the raster kernel gives each tile one workgroup of (TW/2)*(TH/2)
// lanes, one lane per quad.  Instantiated shapes (the autotuner picks among them):
//   0: 16x16 (64 lanes, one wavefront)   1: 32x16 (128 lanes)   2: 32x32 (256 lanes)
constexpr int32_t TILE_CONFIGS = 3;
DH int32_t tileConfigW(int32_t c) { return c == 0 ? 16 : 32; }
DH int32_t tileConfigH(int32_t c) { return c == 2 ? 32 : 16; }
// Commands a tile sorts in LDS at once; longer lists are processed in ordered rounds.
constexpr int32_t TILE_SORT_CAP = 1024;

// One renderer_giveTask(model, transform, camera) call, as the kernels see it.
// What a task's points hold: model-space positions (renderer_giveTask of a model), camera-space
// corners (renderer_giveTask_triangle with the corners' cs), or the caller's ProjectedPoint
// records of 40 bytes (renderer_giveTask_triangle as the reference takes it).
enum { POINTS_MODEL = 0, POINTS_CAMERA_SPACE = 1, POINTS_PROJECTED = 2 };

// A fan triangle of a model as the setup reads it: the point indices of its corners A, B, C
// (polygon corners 0, 1 + k, 2 + k) and polygon * 2 + k, in one 16-byte load.
struct __attribute__((aligned(16))) FanRecord {
	int32_t corner[3];
	int32_t polygonFan;
};

struct TaskDev {
	const float *points;       // pointCount x 3
	const int32_t *indices;    // polygons x 4
	const float *colors;       // polygons x 16
	const float *texcoords;    // polygons x 16
	const FanRecord *fans;     // per fan triangle
	const int32_t *polygonPart;
	const int32_t *partTexture;  // frame texture slot per part: diffuse, light (-1 = none)
	int32_t pointCount, fanCount;
	int32_t pointBase, fanBase;  // offsets of the task's points and fan triangles in the frame's numbering
	int32_t filter;              // 0 Solid, 1 Alpha
	int32_t canvas;              // canvas slot (the batched form renders many canvases)
	int32_t pointKind;           // POINTS_MODEL, POINTS_CAMERA_SPACE or POINTS_PROJECTED
	float transform[12];
	CameraParams camera;
};

// A colour and depth canvas pair.  Either pointer may be null.
struct CanvasDev {
	uint32_t *color;
	float *depth;
	int32_t width, height;
	int32_t colorStride, depthStride;  // in elements
	int32_t tileBase;                  // first tile of this canvas in the frame's tile numbering
	int32_t tilesX, tilesY;
	// Differentiable frames: float RGBA colour, visibility id (command * 4 + lane or -1) and the
	// incoming gradient of the colour, all with diffStride pixels per row.
	float *diffColor;
	int32_t *visibility;
	const float *gradColor;
	int32_t diffStride;
};

// One draw command: a (sub-)triangle that passed culling, clipping and the front-face test.
struct __attribute__((aligned(16))) Command {
	int64_t fx[3], fy[3];  // corners in sub-pixels (256 per pixel)
	f3 start, dx, dy;      // weight planes
	int32_t task, fan;     // source: task and fan triangle within the task's model
	int32_t canvas, flags; // canvas slot; bit 0 = affine (orthogonal camera)
};

// A (tile, command) pair of the packed lists: the command's slot in the tile's list, from the
// tile count's atomic in the setup kernel.
struct PairRecord { int32_t tile, slot, command; };

struct FrameDev {
	const TaskDev *tasks;
	int32_t taskCount;
	const int32_t *taskFanEnd;    // inclusive prefix of fan triangles per task
	const int32_t *taskPointEnd;  // inclusive prefix of points per task
	int32_t fanTotal, pointTotal;
	const CanvasDev *canvases;
	int32_t canvasCount;
	const int32_t *canvasTileEnd; // inclusive prefix of tiles per canvas
	int32_t tileTotal;
	const TextureDesc *textures;
	int32_t textureCount;
	int32_t *fanCount;            // fanTotal: commands made by each fan triangle
	int32_t *blockSums;           // scan scratch (per block: int32 for triangles, int64 for tiles)
	Command *commands;
	ShaderInputs *shaders;        // per command: the finished shader inputs of its source triangle
	TriangleRows *triangleRows;   // per command: edges and bounds of its rows
	int32_t commandCapacity;
	int32_t *counters;            // [0] commands, [1] pairs, [2] overflow flags, [3] tiles with commands,
	                              // [5] backward head, [6] longest tile list when a bin overflowed,
	                              // [7] pair records, [8] most commands of one setup block when a
	                              // single-pass block overflowed, [QUEUE_HEADS + 32 x] XCD x's raster head
	int32_t blockCapacity;        // single-pass setup: command slots per setup block (0: two passes)
	int32_t selfClean;            // the raster leaves zeroed counts for the next frame (resetNextCounters)
	int32_t *nextCounters;        // selfClean: the counter set of the next frame
	int32_t *nextCanvasBase;      // selfClean: the next frame's canvasCount first-command indices
	int32_t *tileCount;           // tileTotal + 1
	// Tile lists.  With binCapacity > 0 tile t owns tileList[t * binCapacity, ...) and the setup
	// kernel appends to it directly; otherwise the lists are packed at tileOffset (an exclusive
	// scan of tileCount) by the scan and fill kernels, pairCapacity entries in all, from the
	// setup kernel's pair records.
	int32_t binCapacity;
	int32_t *tileOffset;          // tileTotal + 1 (exclusive scan of tileCount)
	int32_t *tileFill;            // tileTotal
	int32_t *tileList;
	int32_t pairCapacity;
	PairRecord *pairRecords;      // pairCapacity (packed lists only)
	int32_t *tileQueue;           // tiles with at least one command, in tile order (counters[3] of them)
	int32_t *canvasCommandBase;   // per canvas: index of its first draw command (visibility ids count from it)
	int32_t *hostCounters;        // pinned host copy of counters[0..7], written by the raster kernel
	int32_t tileW, tileH;
	int32_t tileShiftW, tileShiftH;  // log2 of tileW and tileH (the tile sides are powers of two)
	int32_t alphaAny;             // any task uses the Alpha filter: tiles read colour first
	// The visibility raster fills the canvases itself (image_fill) instead of a separate clear:
	// it then visits every tile, and empty ones only take the fill values.
	int32_t clearInRaster;
	uint32_t clearColor;
	float clearDepth;
	// Differentiable frames (mode 1): float RGBA texture (level 0), per-point RGB multipliers
	// and the gradient buffers of the adjoint.
	int32_t mode;
	const float *ftexTexels;
	int32_t ftexLog2w, ftexLog2h;
	const float *multipliers;
	float *gradTexture;
	float *gradMultipliers;
	// Occlusion grid of a renderer frame (one canvas), gridW x gridH cells of 16 x 16 pixels;
	// null when the frame has no occluders.  The setup drops each command it covers
	// (completeOcclusion); k_occlude_existing lowers it.
	float *occlusionGrid;
	int32_t gridW, gridH;
};

// First entry of tile t's command list.
DH const int32_t *tileListOf(const FrameDev &f, int32_t tile) {
	return f.binCapacity > 0 ? f.tileList + (size_t)tile * f.binCapacity : f.tileList + f.tileOffset[tile];
}

DH int32_t findSegment(const int32_t *inclusiveEnd, int32_t count, int32_t index) {
	int32_t lo = 0, hi = count - 1;
	while (lo < hi) {
		int32_t mid = (lo + hi) >> 1;
		if (inclusiveEnd[mid] > index) hi = mid; else lo = mid + 1;
	}
	return lo;
}

// Kernel entry points (kernels.hip).
__global__ void k_frame_begin(FrameDev f);
__global__ void k_setup_count(FrameDev f);
__global__ void k_scan_block_sums(int32_t *blockSums, int32_t blocks, int32_t *total);
__global__ void k_scan_block_sums64(long long *blockSums, int32_t blocks, int32_t *totalLow, int32_t *totalHigh);
__global__ void k_setup_emit(FrameDev f, int32_t sumBlocks);
__global__ void k_setup_single(FrameDev f);
__global__ void k_occlude_existing(FrameDev f);
__global__ void k_occlude_top_rows(const float *depth, int32_t width, int32_t height, int32_t stride, float *grid, int32_t gridW,
                                   int32_t gridH, int32_t perspective);
constexpr int32_t SETUP_SUM_BLOCKS = 2048;  // up to this many setup blocks, each sums its offset itself
__global__ void k_scan_tiles_local(FrameDev f);
__global__ void k_scan_tiles_add(FrameDev f);
__global__ void k_bin_fill(FrameDev f);
template <int TW, int TH> __global__ void k_raster(FrameDev f, int32_t persistent);
template <int TW, int TH, int P> __global__ void k_raster_deferred(FrameDev f);
template <int TW, int TH> __global__ void k_raster_diff(FrameDev f, int32_t persistent);
template <int TW, int TH, int P> __global__ void k_raster_deferred_diff(FrameDev f);
template <int TW, int TH> __global__ void k_raster_depth(FrameDev f);
__global__ void k_probe_edge_side(const int64_t *n, const int64_t *d, const int32_t *L, const int32_t *R, int32_t count, int32_t *out);
__global__ void k_backward(FrameDev f, int32_t *head);
__global__ void k_fill(int32_t *p, int32_t n, int32_t value);
__global__ void k_clear_canvases(FrameDev f, uint32_t color, float depth);
__global__ void k_downsample(uint32_t *texels, TextureDesc t, uint32_t targetLevel);

// Depth-only drawing (model_renderDepth): a depth canvas alone, stride in elements.
struct DepthCanvas {
	float *depth;
	int32_t width, height, stride;
};
// Projects a model's points; then 2^groupShift lanes (3 <= groupShift <= 6) per fan triangle,
// in workgroups of 256 threads.
__global__ void k_depth_project(const float *points, int32_t n, Transform modelToWorld, CameraParams cam, Projected *out);
__global__ void k_depth_draw(const Projected *projected, const FanRecord *fans, int32_t fanCount,
                             int32_t groupShift, CameraParams cam, DepthCanvas cv);

// Workgroups are dispatched to gfx950's eight XCDs round-robin (workgroup b runs on XCD
// b % 8, each XCD with its own L2).  The ordered tile queue is cut into segments of
// XCD_SEGMENT consecutive slots dealt to the XCDs in turn: XCD x takes segments x, x + 8,
// x + 16, ...  Neighbouring tiles of a segment share commands and texels in one L2, and every
// XCD gets a share of every part of the canvas, so a band of costly tiles (a horizon, say)
// is spread over all eight instead of landing on one.
constexpr int32_t XCDS = 8;

constexpr int32_t SCAN_BLOCK = 256;
// The persistent raster's per-XCD queue heads, each on a 128-byte line of its own after the
// frame's other counters, so that the eight XCDs' tickets do not contend for one line.
constexpr int32_t QUEUE_HEADS = 32;
constexpr int32_t QUEUE_HEAD_STRIDE = 32;
constexpr int32_t COUNTERS = QUEUE_HEADS + QUEUE_HEAD_STRIDE * XCDS;
constexpr int32_t HOST_COUNTERS = 16;  // counters[0..15] are copied to pinned host memory
#ifndef DFPSR_XCD_SEGMENT
#define DFPSR_XCD_SEGMENT 4
#endif
constexpr int32_t XCD_SEGMENT = DFPSR_XCD_SEGMENT;
// Queue slot of the i-th tile of XCD x.  Increasing in i: the first slot >= queued ends x's share.
DH int32_t xcdSlot(int32_t x, int32_t i) {
	return ((i / XCD_SEGMENT) * XCDS + x) * XCD_SEGMENT + i % XCD_SEGMENT;
}
// Workgroups a one-tile-per-workgroup launch needs so that every slot below tileTotal has one.
DH int32_t xcdGrid(int32_t tileTotal) {
	const int32_t round = XCDS * XCD_SEGMENT;
	return (tileTotal + round - 1) / round * round;
}

}  // namespace dfpsr
