// The launch sequence of one frame (renderer_end): projection -> triangle setup (count,
// scan, emit) -> tile binning (count, scan, fill) -> tile raster.  Buffers that depend on
// data (commands, tile pairs) start from an estimate; a pass that runs out of room raises a
// flag, leaves the canvases untouched, and the frame is run again with the exact sizes.
#include "frame.h"

#include <cstring>

namespace dfpsr {

Frame::~Frame() {
	if (hostCounters_) (void)hipHostFree(hostCounters_);
	if (stream) (void)hipStreamDestroy(stream);
}

int Frame::init() {
	HIP_TRY(hipStreamCreate(&stream));  // blocking: ordered after work on the legacy default stream
	HIP_TRY(hipHostMalloc((void **)&hostCounters_, 4 * sizeof(int32_t), hipHostMallocDefault));
	return 0;
}

void Frame::reset() {
	tasks_.clear();
	taskFanEnd_.clear();
	taskPointEnd_.clear();
	canvases_.clear();
	canvasTileEnd_.clear();
	textures_.clear();
	textureKeys_.clear();
	partTexture_.clear();
	partTextureOffset_.clear();
	alphaAny_ = false;
}

int32_t Frame::addCanvas(const CanvasHost &c) {
	CanvasDev d;
	d.color = c.color;
	d.depth = c.depth;
	d.width = c.width;
	d.height = c.height;
	d.colorStride = c.colorStride;
	d.depthStride = c.depthStride;
	d.tilesX = (c.width + TILE_W - 1) / TILE_W;
	d.tilesY = (c.height + TILE_H - 1) / TILE_H;
	d.tileBase = canvasTileEnd_.empty() ? 0 : canvasTileEnd_.back();
	canvases_.push_back(d);
	canvasTileEnd_.push_back(d.tileBase + d.tilesX * d.tilesY);
	return (int32_t)canvases_.size() - 1;
}

int32_t Frame::textureSlot(const TextureDesc *desc) {
	if (!desc) return -1;
	for (size_t i = 0; i < textureKeys_.size(); i++)
		if (textureKeys_[i] == desc) return (int32_t)i;
	textureKeys_.push_back(desc);
	textures_.push_back(*desc);
	return (int32_t)textures_.size() - 1;
}

int Frame::addTask(const ModelView &m, const float *transform, const CameraParams &camera, int32_t canvas, bool cameraSpacePoints) {
	if (m.fanCount == 0 || m.pointCount == 0) return 0;
	TaskDev t;
	std::memset(&t, 0, sizeof(t));
	t.points = m.points;
	t.indices = m.indices;
	t.colors = m.colors;
	t.texcoords = m.texcoords;
	t.fanPolygon = m.fanPolygon;
	t.polygonPart = m.polygonPart;
	t.partTexture = nullptr;  // patched at upload
	t.pointCount = m.pointCount;
	t.fanCount = m.fanCount;
	t.pointBase = taskPointEnd_.empty() ? 0 : taskPointEnd_.back();
	t.fanBase = taskFanEnd_.empty() ? 0 : taskFanEnd_.back();
	t.filter = m.filter;
	t.canvas = canvas;
	t.cameraSpacePoints = cameraSpacePoints ? 1 : 0;
	std::memcpy(t.transform, transform, sizeof(t.transform));
	t.camera = camera;
	partTextureOffset_.push_back(partTexture_.size());
	for (int32_t p = 0; p < m.partCount; p++) {
		partTexture_.push_back(textureSlot(m.partDiffuse[p]));
		partTexture_.push_back(textureSlot(m.partLight[p]));
	}
	if (m.filter != 0) alphaAny_ = true;
	tasks_.push_back(t);
	taskPointEnd_.push_back(t.pointBase + m.pointCount);
	taskFanEnd_.push_back(t.fanBase + m.fanCount);
	return 0;
}

int Frame::upload(const void *src, size_t bytes, DevBuf &dst) {
	if (dst.ensure(bytes) != hipSuccess) return fail(-4, "device allocation failed");
	if (bytes) HIP_TRY(hipMemcpyAsync(dst.ptr, src, bytes, hipMemcpyHostToDevice, stream));
	return 0;
}

static inline int32_t blocksOf(int64_t n, int32_t size) { return (int32_t)((n + size - 1) / size); }

int Frame::render() {
	stats[0] = stats[1] = stats[2] = stats[3] = 0;
	if (tasks_.empty() || canvases_.empty()) return 0;
	const int32_t pointTotal = taskPointEnd_.back();
	const int32_t fanTotal = taskFanEnd_.back();
	const int32_t tileTotal = canvasTileEnd_.back();
	if (fanTotal == 0 || tileTotal == 0) return 0;

	// Per-frame tables.
	if (partTexture_.empty()) partTexture_.push_back(-1);
	if (dPartTexture_.ensure(partTexture_.size() * 4) != hipSuccess) return fail(-4, "device allocation failed");
	for (size_t i = 0; i < tasks_.size(); i++) tasks_[i].partTexture = dPartTexture_.as<int32_t>() + partTextureOffset_[i];
	int rc;
	if ((rc = upload(partTexture_.data(), partTexture_.size() * 4, dPartTexture_))) return rc;
	if ((rc = upload(tasks_.data(), tasks_.size() * sizeof(TaskDev), dTasks_))) return rc;
	if ((rc = upload(taskFanEnd_.data(), taskFanEnd_.size() * 4, dTaskFanEnd_))) return rc;
	if ((rc = upload(taskPointEnd_.data(), taskPointEnd_.size() * 4, dTaskPointEnd_))) return rc;
	if ((rc = upload(canvases_.data(), canvases_.size() * sizeof(CanvasDev), dCanvases_))) return rc;
	if ((rc = upload(canvasTileEnd_.data(), canvasTileEnd_.size() * 4, dCanvasTileEnd_))) return rc;
	TextureDesc none;
	std::memset(&none, 0, sizeof(none));
	if (textures_.empty()) textures_.push_back(none);
	if ((rc = upload(textures_.data(), textures_.size() * sizeof(TextureDesc), dTextures_))) return rc;

	// Working buffers.
	const int32_t fanBlocks = blocksOf(fanTotal, SCAN_BLOCK);
	const int32_t tileBlocks = blocksOf(tileTotal, SCAN_BLOCK);
	if (commandCapacity_ < fanTotal + 1024) commandCapacity_ = fanTotal + 1024;
	if (pairCapacity_ < 4 * commandCapacity_) pairCapacity_ = 4 * commandCapacity_;
	if (dProjected_.ensure((size_t)pointTotal * sizeof(Projected)) != hipSuccess ||
	    dFanCount_.ensure((size_t)fanTotal * 4) != hipSuccess ||
	    dBlockSums_.ensure((size_t)(fanBlocks > tileBlocks ? fanBlocks : tileBlocks) * 4) != hipSuccess ||
	    dCounters_.ensure(4 * 4) != hipSuccess || dTileCount_.ensure(((size_t)tileTotal + 1) * 4) != hipSuccess ||
	    dTileOffset_.ensure(((size_t)tileTotal + 1) * 4) != hipSuccess || dTileFill_.ensure((size_t)tileTotal * 4) != hipSuccess)
		return fail(-4, "device allocation failed");

	for (int pass = 1; pass <= 3; pass++) {
		if (dCommands_.ensure((size_t)commandCapacity_ * sizeof(Command)) != hipSuccess ||
		    dTileList_.ensure((size_t)pairCapacity_ * 4) != hipSuccess)
			return fail(-4, "device allocation failed");
		FrameDev f;
		f.tasks = dTasks_.as<TaskDev>();
		f.taskCount = (int32_t)tasks_.size();
		f.taskFanEnd = dTaskFanEnd_.as<int32_t>();
		f.taskPointEnd = dTaskPointEnd_.as<int32_t>();
		f.fanTotal = fanTotal;
		f.pointTotal = pointTotal;
		f.canvases = dCanvases_.as<CanvasDev>();
		f.canvasCount = (int32_t)canvases_.size();
		f.canvasTileEnd = dCanvasTileEnd_.as<int32_t>();
		f.tileTotal = tileTotal;
		f.textures = dTextures_.as<TextureDesc>();
		f.projected = dProjected_.as<Projected>();
		f.fanCount = dFanCount_.as<int32_t>();
		f.blockSums = dBlockSums_.as<int32_t>();
		f.commands = dCommands_.as<Command>();
		f.commandCapacity = commandCapacity_;
		f.counters = dCounters_.as<int32_t>();
		f.tileCount = dTileCount_.as<int32_t>();
		f.tileOffset = dTileOffset_.as<int32_t>();
		f.tileFill = dTileFill_.as<int32_t>();
		f.tileList = dTileList_.as<int32_t>();
		f.pairCapacity = pairCapacity_;
		f.alphaAny = alphaAny_ ? 1 : 0;

		HIP_TRY(hipMemsetAsync(f.counters, 0, 16, stream));
		HIP_TRY(hipMemsetAsync(f.tileCount, 0, ((size_t)tileTotal + 1) * 4, stream));
		HIP_TRY(hipMemsetAsync(f.tileFill, 0, (size_t)tileTotal * 4, stream));
		hipLaunchKernelGGL(k_project, dim3(blocksOf(pointTotal, 256)), dim3(256), 0, stream, f);
		hipLaunchKernelGGL(k_setup_count, dim3(fanBlocks), dim3(SCAN_BLOCK), 0, stream, f);
		hipLaunchKernelGGL(k_scan_block_sums, dim3(1), dim3(SCAN_BLOCK), 0, stream, f.blockSums, fanBlocks, f.counters + 0);
		hipLaunchKernelGGL(k_setup_emit, dim3(fanBlocks), dim3(SCAN_BLOCK), 0, stream, f);
		hipLaunchKernelGGL(k_bin_count, dim3(blocksOf(commandCapacity_, 256)), dim3(256), 0, stream, f);
		hipLaunchKernelGGL(k_scan_tiles_local, dim3(tileBlocks), dim3(SCAN_BLOCK), 0, stream, f);
		hipLaunchKernelGGL(k_scan_block_sums, dim3(1), dim3(SCAN_BLOCK), 0, stream, f.blockSums, tileBlocks, f.counters + 1);
		hipLaunchKernelGGL(k_scan_tiles_add, dim3(tileBlocks), dim3(SCAN_BLOCK), 0, stream, f);
		hipLaunchKernelGGL(k_bin_fill, dim3(blocksOf(commandCapacity_, 256)), dim3(256), 0, stream, f);
		hipLaunchKernelGGL(k_raster, dim3(tileTotal), dim3(TILE_LANES), 0, stream, f);
		HIP_TRY(hipGetLastError());
		HIP_TRY(hipMemcpyAsync(hostCounters_, f.counters, 16, hipMemcpyDeviceToHost, stream));
		HIP_TRY(hipStreamSynchronize(stream));
		stats[0] = hostCounters_[0];
		stats[1] = hostCounters_[1];
		stats[2] = tileTotal;
		stats[3] = pass;
		int32_t overflow = hostCounters_[2];
		if (overflow == 0) return 0;
		if (overflow & 1) commandCapacity_ = hostCounters_[0] + 1024;
		if (overflow & 2) pairCapacity_ = hostCounters_[1] + 1024;
		if ((overflow & 1) && pairCapacity_ < 4 * commandCapacity_) pairCapacity_ = 4 * commandCapacity_;
	}
	return fail(-4, "renderer_end: buffers did not converge");
}

}  // namespace dfpsr
