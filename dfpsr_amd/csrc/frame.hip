// The launch sequence of one frame (renderer_end, or one batch of jobs):
//   [clear canvases, unless the raster fills them] -> [frame begin: counters and tile counts
//   zeroed, unless the previous frame's raster left them so] -> triangle setup (the corners
//   projected, count, emit with per-command row setup, shader inputs and tile counts, in one
//   pass or in count / scan / emit) -> [tile offsets (scan + ordered tile queue) -> list fill,
//   for packed lists] -> tile raster.
// Buffers that depend on data (draw commands, tile pairs) start from an estimate; a pass that
// runs out of room raises a flag, leaves the canvases untouched, and the frame is run again
// with the exact sizes.  With RenderOptions::graph the sequence is captured once into a
// hipGraph and replayed while the frame's tables and buffers stay the same.
#include "frame.h"

#include <algorithm>
#include <cstring>

namespace dfpsr {

Frame::~Frame() {
	dropGraph();
	for (hipEvent_t e : events_)
		if (e) (void)hipEventDestroy(e);
	if (hostCounters_) (void)hipHostFree(hostCounters_);
	if (stream) (void)hipStreamDestroy(stream);
}

int Frame::init() {
	HIP_TRY(hipStreamCreate(&stream));  // blocking: ordered after work on the legacy default stream
	HIP_TRY(hipHostMalloc((void **)&hostCounters_, HOST_COUNTERS * sizeof(int32_t), hipHostMallocDefault));
	HIP_TRY(hipHostGetDevicePointer((void **)&hostCountersDevice_, hostCounters_, 0));
	for (hipEvent_t &e : events_) HIP_TRY(hipEventCreate(&e));
	int device = 0;
	HIP_TRY(hipGetDevice(&device));
	HIP_TRY(hipDeviceGetAttribute(&deviceCUs_, hipDeviceAttributeMultiprocessorCount, device));
	return 0;
}

void Frame::dropGraph() {
	if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
	if (graph_) (void)hipGraphDestroy(graph_);
	if (rasterGraphExec_) (void)hipGraphExecDestroy(rasterGraphExec_);
	if (rasterGraph_) (void)hipGraphDestroy(rasterGraph_);
	graphExec_ = nullptr;
	graph_ = nullptr;
	rasterGraphExec_ = nullptr;
	rasterGraph_ = nullptr;
	graphKey_.clear();
	if (bwdGraphExec_) (void)hipGraphExecDestroy(bwdGraphExec_);
	if (bwdGraph_) (void)hipGraphDestroy(bwdGraph_);
	bwdGraphExec_ = nullptr;
	bwdGraph_ = nullptr;
	bwdGraphKey_.clear();
}

void Frame::reset(int32_t tileConfig) {
	tileConfig_ = (tileConfig >= 0 && tileConfig < TILE_CONFIGS) ? tileConfig : 0;
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
	orthogonalAny_ = false;
	occlusionGrid_ = nullptr;
	gridW_ = gridH_ = 0;
	oddCanvas_ = false;
	wideCanvas_ = false;
}

int32_t Frame::addCanvas(const CanvasHost &c) {
	const int32_t tw = tileConfigW(tileConfig_), th = tileConfigH(tileConfig_);
	CanvasDev d;
	d.color = c.color;
	d.depth = c.depth;
	d.width = c.width;
	d.height = c.height;
	d.colorStride = c.colorStride;
	d.depthStride = c.depthStride;
	d.tilesX = (c.width + tw - 1) / tw;
	d.tilesY = (c.height + th - 1) / th;
	d.tileBase = canvasTileEnd_.empty() ? 0 : canvasTileEnd_.back();
	d.diffColor = c.diffColor;
	d.visibility = c.visibility;
	d.gradColor = c.gradColor;
	d.diffStride = c.diffStride;
	if (c.height & 1) oddCanvas_ = true;
	if (c.width > 32767) wideCanvas_ = true;
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

int Frame::addTask(const ModelView &m, const float *transform, const CameraParams &camera, int32_t canvas, int32_t pointKind) {
	if (m.fanCount == 0 || m.pointCount == 0) return 0;
	// The points may have been replaced by fewer after the polygons were added: a polygon
	// naming a missing point would send the setup kernel past the model's points.
	if (m.maxIndex >= m.pointCount) return fail(-1, "giveTask: a polygon names a point the model does not have");
	TaskDev t;
	std::memset(&t, 0, sizeof(t));
	t.points = m.points;
	t.indices = m.indices;
	t.colors = m.colors;
	t.texcoords = m.texcoords;
	t.fans = m.fans;
	t.polygonPart = m.polygonPart;
	t.partTexture = nullptr;  // patched when the tables are packed
	t.pointCount = m.pointCount;
	t.fanCount = m.fanCount;
	t.pointBase = taskPointEnd_.empty() ? 0 : taskPointEnd_.back();
	t.fanBase = taskFanEnd_.empty() ? 0 : taskFanEnd_.back();
	t.filter = m.filter;
	t.canvas = canvas;
	t.pointKind = pointKind;
	std::memcpy(t.transform, transform, sizeof(t.transform));
	t.camera = camera;
	partTextureOffset_.push_back(partTexture_.size());
	for (int32_t p = 0; p < m.partCount; p++) {
		partTexture_.push_back(textureSlot(m.partDiffuse[p]));
		partTexture_.push_back(textureSlot(m.partLight[p]));
	}
	if (m.filter != 0) alphaAny_ = true;
	if (!camera.perspective) orthogonalAny_ = true;
	tasks_.push_back(t);
	taskPointEnd_.push_back(t.pointBase + m.pointCount);
	taskFanEnd_.push_back(t.fanBase + m.fanCount);
	return 0;
}

static inline int32_t blocksOf(int64_t n, int32_t size) { return (int32_t)((n + size - 1) / size); }
static inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

int Frame::enqueue(const FrameDev &f, const RenderOptions &o, int32_t fanBlocks, int32_t tileBlocks, bool events, int parts,
                   bool skipBegin) {
	const int32_t tileTotal = f.tileTotal;
	if (parts & 1) {
	if (o.clear && !f.clearInRaster) hipLaunchKernelGGL(k_clear_canvases, dim3(tileTotal), dim3(256), 0, stream, f, o.clearColor, o.clearDepth);
	if (events) HIP_TRY(hipEventRecord(events_[0], stream));
	int64_t beginThreads = 2 * (int64_t)tileTotal + 1;
	if (beginThreads < f.canvasCount) beginThreads = f.canvasCount;
	if (beginThreads < COUNTERS) beginThreads = COUNTERS;
	if (!skipBegin) hipLaunchKernelGGL(k_frame_begin, dim3(blocksOf(beginThreads, 256)), dim3(256), 0, stream, f);
	if (events) HIP_TRY(hipEventRecord(events_[1], stream));
	if (f.blockCapacity > 0) {
		hipLaunchKernelGGL(k_setup_single, dim3(fanBlocks), dim3(SCAN_BLOCK), 0, stream, f);
	} else {
		hipLaunchKernelGGL(k_setup_count, dim3(fanBlocks), dim3(SCAN_BLOCK), 0, stream, f);
		const int32_t sumBlocks = fanBlocks <= SETUP_SUM_BLOCKS ? 1 : 0;
		if (!sumBlocks) hipLaunchKernelGGL(k_scan_block_sums, dim3(1), dim3(SCAN_BLOCK), 0, stream, f.blockSums, fanBlocks, f.counters + 0);
		hipLaunchKernelGGL(k_setup_emit, dim3(fanBlocks), dim3(SCAN_BLOCK), 0, stream, f, sumBlocks);
	}
	if (events) HIP_TRY(hipEventRecord(events_[2], stream));
	if (f.binCapacity == 0) {
	hipLaunchKernelGGL(k_scan_tiles_local, dim3(tileBlocks), dim3(SCAN_BLOCK), 0, stream, f);
	hipLaunchKernelGGL(k_scan_block_sums64, dim3(1), dim3(SCAN_BLOCK), 0, stream, (long long *)f.blockSums, tileBlocks,
	                   f.counters + 1, f.counters + 3);
	hipLaunchKernelGGL(k_scan_tiles_add, dim3(tileBlocks), dim3(SCAN_BLOCK), 0, stream, f);
	hipLaunchKernelGGL(k_bin_fill, dim3(blocksOf(f.pairCapacity, 256)), dim3(256), 0, stream, f);
	}
	if (events) HIP_TRY(hipEventRecord(events_[3], stream));
	}
	if (!(parts & 2)) {
		HIP_TRY(hipGetLastError());
		return 0;
	}
	const int32_t persistent = o.persistentPerCU > 0 ? 1 : 0;
	const int32_t grid = persistent ? deviceCUs_ * o.persistentPerCU : xcdGrid(tileTotal);
	if (o.mode == 2) {
		switch (tileConfig_) {
		case 1: hipLaunchKernelGGL((k_raster_depth<32, 16>), dim3(xcdGrid(tileTotal)), dim3(32 * 16 / 4), 0, stream, f); break;
		case 2: hipLaunchKernelGGL((k_raster_depth<32, 32>), dim3(xcdGrid(tileTotal)), dim3(32 * 32 / 4), 0, stream, f); break;
		default: hipLaunchKernelGGL((k_raster_depth<16, 16>), dim3(xcdGrid(tileTotal)), dim3(16 * 16 / 4), 0, stream, f); break;
		}
	} else if (o.mode == 1) {
		const bool deferred = deferredRaster(o);
		switch (tileConfig_) {
		case 1:
			if (!deferred) hipLaunchKernelGGL((k_raster_diff<32, 16>), dim3(grid), dim3(128), 0, stream, f, persistent);
			else if (persistent) hipLaunchKernelGGL((k_raster_deferred_diff<32, 16, 1>), dim3(grid), dim3(128), 0, stream, f);
			else hipLaunchKernelGGL((k_raster_deferred_diff<32, 16, 0>), dim3(grid), dim3(128), 0, stream, f);
			break;
		case 2:
			if (!deferred) hipLaunchKernelGGL((k_raster_diff<32, 32>), dim3(grid), dim3(256), 0, stream, f, persistent);
			else if (persistent) hipLaunchKernelGGL((k_raster_deferred_diff<32, 32, 1>), dim3(grid), dim3(256), 0, stream, f);
			else hipLaunchKernelGGL((k_raster_deferred_diff<32, 32, 0>), dim3(grid), dim3(256), 0, stream, f);
			break;
		default:
			if (!deferred) hipLaunchKernelGGL((k_raster_diff<16, 16>), dim3(grid), dim3(64), 0, stream, f, persistent);
			else if (persistent) hipLaunchKernelGGL((k_raster_deferred_diff<16, 16, 1>), dim3(grid), dim3(64), 0, stream, f);
			else hipLaunchKernelGGL((k_raster_deferred_diff<16, 16, 0>), dim3(grid), dim3(64), 0, stream, f);
			break;
		}
	} else if (deferredRaster(o)) {
		switch (tileConfig_) {
		case 1:
			if (persistent) hipLaunchKernelGGL((k_raster_deferred<32, 16, 1>), dim3(grid), dim3(32 * 16 / 4), 0, stream, f);
			else hipLaunchKernelGGL((k_raster_deferred<32, 16, 0>), dim3(grid), dim3(32 * 16 / 4), 0, stream, f);
			break;
		case 2:
			if (persistent) hipLaunchKernelGGL((k_raster_deferred<32, 32, 1>), dim3(grid), dim3(32 * 32 / 4), 0, stream, f);
			else hipLaunchKernelGGL((k_raster_deferred<32, 32, 0>), dim3(grid), dim3(32 * 32 / 4), 0, stream, f);
			break;
		default:
			if (persistent) hipLaunchKernelGGL((k_raster_deferred<16, 16, 1>), dim3(grid), dim3(16 * 16 / 4), 0, stream, f);
			else hipLaunchKernelGGL((k_raster_deferred<16, 16, 0>), dim3(grid), dim3(16 * 16 / 4), 0, stream, f);
			break;
		}
	} else switch (tileConfig_) {
	case 1: hipLaunchKernelGGL((k_raster<32, 16>), dim3(grid), dim3(128), 0, stream, f, persistent); break;
	case 2: hipLaunchKernelGGL((k_raster<32, 32>), dim3(grid), dim3(256), 0, stream, f, persistent); break;
	default: hipLaunchKernelGGL((k_raster<16, 16>), dim3(grid), dim3(64), 0, stream, f, persistent); break;
	}
	if (events) HIP_TRY(hipEventRecord(events_[4], stream));
	HIP_TRY(hipGetLastError());
	return 0;
}

int Frame::prepare() {
	prepared_ = false;
	if (tasks_.empty() || canvases_.empty()) return 0;
	if (partTexture_.empty()) partTexture_.push_back(-1);
	if (textures_.empty()) {
		TextureDesc none;
		std::memset(&none, 0, sizeof(none));
		textures_.push_back(none);
	}
	const size_t nT = tasks_.size(), nC = canvases_.size();
	offTasks_ = 0;
	offFanEnd_ = align256(offTasks_ + nT * sizeof(TaskDev));
	offPointEnd_ = align256(offFanEnd_ + nT * 4);
	offCanvases_ = align256(offPointEnd_ + nT * 4);
	offTileEnd_ = align256(offCanvases_ + nC * sizeof(CanvasDev));
	offTextures_ = align256(offTileEnd_ + nC * 4);
	const size_t offParts = align256(offTextures_ + textures_.size() * sizeof(TextureDesc));
	const size_t total = align256(offParts + partTexture_.size() * 4);
	if (dTables_.ensure(total) != hipSuccess) return fail(-4, "device allocation failed");
	uint8_t *base = (uint8_t *)dTables_.ptr;
	tables_.assign(total, 0);
	for (size_t i = 0; i < nT; i++) tasks_[i].partTexture = (const int32_t *)(base + offParts) + partTextureOffset_[i];
	std::memcpy(&tables_[offTasks_], tasks_.data(), nT * sizeof(TaskDev));
	std::memcpy(&tables_[offFanEnd_], taskFanEnd_.data(), nT * 4);
	std::memcpy(&tables_[offPointEnd_], taskPointEnd_.data(), nT * 4);
	std::memcpy(&tables_[offCanvases_], canvases_.data(), nC * sizeof(CanvasDev));
	std::memcpy(&tables_[offTileEnd_], canvasTileEnd_.data(), nC * 4);
	std::memcpy(&tables_[offTextures_], textures_.data(), textures_.size() * sizeof(TextureDesc));
	std::memcpy(&tables_[offParts], partTexture_.data(), partTexture_.size() * 4);
	if (tables_ != uploadedTables_) {
		HIP_TRY(hipMemcpyAsync(dTables_.ptr, tables_.data(), total, hipMemcpyHostToDevice, stream));
		uploadedTables_ = tables_;
	}
	prepared_ = true;
	return 0;
}

int Frame::run(const RenderOptions &o) {
	stats[0] = stats[1] = stats[2] = stats[3] = 0;
	for (float &m : stageMs) m = 0.0f;
	if (!prepared_ || tasks_.empty() || canvases_.empty()) return 0;
	const int32_t pointTotal = taskPointEnd_.back();
	const int32_t fanTotal = taskFanEnd_.back();
	const int32_t tileTotal = canvasTileEnd_.back();
	if (fanTotal == 0 || tileTotal == 0) return 0;
	const size_t nT = tasks_.size(), nC = canvases_.size();
	uint8_t *base = (uint8_t *)dTables_.ptr;

	const int32_t fanBlocks = blocksOf(fanTotal, SCAN_BLOCK);
	const int32_t tileBlocks = blocksOf(tileTotal, SCAN_BLOCK);
	if (commandCapacity_ < fanTotal + 1024) commandCapacity_ = fanTotal + 1024;
	if (pairCapacity_ < 4 * commandCapacity_) pairCapacity_ = 4 * commandCapacity_;
	if (dFanCount_.ensure((size_t)fanTotal * 4) != hipSuccess ||
	    dBlockSums_.ensure((size_t)(fanBlocks > tileBlocks ? fanBlocks : tileBlocks) * 8) != hipSuccess ||
	    dCounters_.ensure(2 * COUNTERS * 4) != hipSuccess || dTileCount_.ensure((2 * (size_t)tileTotal + 1) * 4) != hipSuccess ||
	    dTileOffset_.ensure(((size_t)tileTotal + 1) * 4) != hipSuccess || dTileQueue_.ensure((size_t)tileTotal * 4) != hipSuccess ||
	    dCanvasBase_.ensure(2 * nC * 4) != hipSuccess)
		return fail(-4, "device allocation failed");

	for (int pass = 1; pass <= 3; pass++) {
		const bool single = singlePass(o, fanBlocks);
		if (single && commandCapacity_ < fanBlocks * blockCapacity_) commandCapacity_ = fanBlocks * blockCapacity_;
		if (pairCapacity_ < 4 * commandCapacity_) pairCapacity_ = 4 * commandCapacity_;
		const bool direct = o.binning == 0 && (int64_t)tileTotal * binCapacity_ <= BIN_ENTRY_LIMIT;
		const size_t listEntries = direct ? (size_t)tileTotal * binCapacity_ : (size_t)pairCapacity_;
		if (dCommands_.ensure((size_t)commandCapacity_ * sizeof(Command)) != hipSuccess ||
		    (o.mode == 0 && dShaders_.ensure((size_t)commandCapacity_ * sizeof(ShaderInputs)) != hipSuccess) ||
		    dTriangleRows_.ensure((size_t)commandCapacity_ * sizeof(TriangleRows)) != hipSuccess ||
		    dTileList_.ensure(listEntries * 4) != hipSuccess ||
		    (!direct && dPairRecords_.ensure((size_t)pairCapacity_ * sizeof(PairRecord)) != hipSuccess))
			return fail(-4, "device allocation failed");
		FrameDev f;
		std::memset(&f, 0, sizeof(f));
		f.tasks = (const TaskDev *)(base + offTasks_);
		f.taskCount = (int32_t)nT;
		f.taskFanEnd = (const int32_t *)(base + offFanEnd_);
		f.taskPointEnd = (const int32_t *)(base + offPointEnd_);
		f.fanTotal = fanTotal;
		f.pointTotal = pointTotal;
		f.canvases = (const CanvasDev *)(base + offCanvases_);
		f.canvasCount = (int32_t)nC;
		f.canvasTileEnd = (const int32_t *)(base + offTileEnd_);
		f.tileTotal = tileTotal;
		f.textures = (const TextureDesc *)(base + offTextures_);
		f.textureCount = (int32_t)textures_.size();
		f.fanCount = dFanCount_.as<int32_t>();
		f.blockSums = dBlockSums_.as<int32_t>();
		f.commands = dCommands_.as<Command>();
		f.shaders = dShaders_.as<ShaderInputs>();
		f.triangleRows = dTriangleRows_.as<TriangleRows>();
		f.commandCapacity = commandCapacity_;
		// Two sets of counters and first-command indices, used by alternate self-cleaning frames.
		f.counters = dCounters_.as<int32_t>() + parity_ * COUNTERS;
		f.hostCounters = hostCountersDevice_;
		f.tileCount = dTileCount_.as<int32_t>();
		f.tileFill = dTileCount_.as<int32_t>() + tileTotal + 1;
		f.tileOffset = dTileOffset_.as<int32_t>();
		f.tileList = dTileList_.as<int32_t>();
		f.tileQueue = dTileQueue_.as<int32_t>();
		f.canvasCommandBase = dCanvasBase_.as<int32_t>() + parity_ * (int32_t)nC;
		f.pairCapacity = pairCapacity_;
		f.pairRecords = direct ? nullptr : dPairRecords_.as<PairRecord>();
		f.binCapacity = direct ? binCapacity_ : 0;
		f.tileW = tileConfigW(tileConfig_);
		f.tileH = tileConfigH(tileConfig_);
		f.tileShiftW = f.tileW == 32 ? 5 : 4;
		f.tileShiftH = f.tileH == 32 ? 5 : 4;
		f.alphaAny = alphaAny_ ? 1 : 0;
		// The differentiable frames fill in the raster only with 32-pixel-wide tiles: with 16x16
		// tiles the separate fill kernel streams the float canvases faster than the one-wave
		// workgroups' write-back of every pixel (256-view batch forward 1.61 -> 1.98 ms fused).
		f.clearInRaster = (o.clear && (o.mode == 0 || (o.mode == 1 && f.tileW == 32)) && deferredRaster(o)) ? 1 : 0;
		f.clearColor = o.clearColor;
		f.clearDepth = o.clearDepth;
		f.mode = o.mode;
		f.ftexTexels = ftex_;
		f.ftexLog2w = ftexLog2w_;
		f.ftexLog2h = ftexLog2h_;
		f.multipliers = multipliers_;
		f.blockCapacity = single ? blockCapacity_ : 0;
		// The raster leaves the counts zeroed for the next frame on the common path: one pass of
		// setup, direct bins, one workgroup per tile, no graph.  A frame may then skip the begin
		// kernel when the frame before it cleaned the same buffers and it uses no more tiles or
		// canvases than that frame did.
		const bool selfClean = !o.graph && single && direct && o.persistentPerCU <= 0 &&
		                       (o.mode == 2 || (o.mode == 0 && deferredRaster(o)));
		const uint64_t gens[3] = {dCounters_.generation, dTileCount_.generation, dCanvasBase_.generation};
		const bool skipBegin = selfClean && clean_ && std::equal(gens, gens + 3, cleanGenerations_) &&
		                       tileTotal <= cleanTiles_ && (int32_t)nC == cleanCanvases_;
		f.selfClean = selfClean ? 1 : 0;
		f.nextCounters = selfClean ? dCounters_.as<int32_t>() + (1 - parity_) * COUNTERS : nullptr;
		f.nextCanvasBase = selfClean ? dCanvasBase_.as<int32_t>() + (1 - parity_) * (int32_t)nC : nullptr;
		clean_ = false;  // this frame uses the counts; they are clean again only if its raster says so
		f.occlusionGrid = occlusionGrid_;
		f.gridW = gridW_;
		f.gridH = gridH_;
		lastRun_ = f;
		lastRunValid_ = false;
		// The raster kernel copies counters[0..7] into pinned host memory as it starts; the
		// sentinel, written before anything of the frame is launched, shows whether it did.
		*(volatile int32_t *)&hostCounters_[2] = -1;

		if (o.graph) {
			// The graphs stay valid while the launch arguments, the options and the timing switch
			// are the same.  Events recorded inside a graph cannot be timed, so with timing on the
			// sequence is captured as two graphs -- everything before the raster, and the raster --
			// with the events recorded between their launches.
			const uint8_t split = timing ? 1 : 0;
			std::vector<uint8_t> key(sizeof(FrameDev) + sizeof(RenderOptions) + sizeof(int32_t) + 1);
			std::memcpy(key.data(), &f, sizeof(FrameDev));
			std::memcpy(key.data() + sizeof(FrameDev), &o, sizeof(RenderOptions));
			std::memcpy(key.data() + sizeof(FrameDev) + sizeof(RenderOptions), &tileConfig_, sizeof(int32_t));
			key.back() = split;
			if (!graphExec_ || key != graphKey_) {
				dropGraph();
				for (int g = 0; g < (split ? 2 : 1); g++) {
					HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
					int rc = enqueue(f, o, fanBlocks, tileBlocks, false, split ? (g == 0 ? 1 : 2) : 3);
					hipGraph_t captured = nullptr;
					hipError_t e = hipStreamEndCapture(stream, &captured);
					if (rc) {
						if (captured) (void)hipGraphDestroy(captured);
						return rc;
					}
					if (e != hipSuccess) return fail(-3, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
					hipGraph_t &slotGraph = g == 0 ? graph_ : rasterGraph_;
					hipGraphExec_t &slotExec = g == 0 ? graphExec_ : rasterGraphExec_;
					slotGraph = captured;
					HIP_TRY(hipGraphInstantiate(&slotExec, slotGraph, nullptr, nullptr, 0));
				}
				graphKey_ = key;
			}
			if (split) {
				HIP_TRY(hipEventRecord(events_[0], stream));
				HIP_TRY(hipGraphLaunch(graphExec_, stream));
				HIP_TRY(hipEventRecord(events_[3], stream));
				HIP_TRY(hipGraphLaunch(rasterGraphExec_, stream));
				HIP_TRY(hipEventRecord(events_[4], stream));
			} else {
				HIP_TRY(hipGraphLaunch(graphExec_, stream));
			}
		} else {
			int rc = enqueue(f, o, fanBlocks, tileBlocks, timing, 3, skipBegin);
			if (rc) return rc;
		}
		HIP_TRY(hipStreamSynchronize(stream));
		if (*(volatile int32_t *)&hostCounters_[2] == -1) return fail(-3, "renderer_end: the frame did not report its counters");
		stats[0] = hostCounters_[0];
		stats[1] = hostCounters_[1];
		stats[2] = hostCounters_[3];
		stats[3] = pass;
		if (timing) {
			if (o.graph) {
				// The graph's stages before the raster are timed as one (reported as binning).
				stageMs[0] = stageMs[1] = 0.0f;
				HIP_TRY(hipEventElapsedTime(&stageMs[2], events_[0], events_[3]));
				HIP_TRY(hipEventElapsedTime(&stageMs[3], events_[3], events_[4]));
			} else {
				for (int k = 0; k < 4; k++) HIP_TRY(hipEventElapsedTime(&stageMs[k], events_[k], events_[k + 1]));
			}
			HIP_TRY(hipEventElapsedTime(&stageMs[4], events_[0], events_[4]));
		}
		int32_t overflow = hostCounters_[2];
		if (overflow == 0) {
			lastRunValid_ = true;
			if (selfClean) {
				clean_ = true;  // the other set is reset and every tile count zeroed: the next frame uses that set
				parity_ = 1 - parity_;
				std::copy(gens, gens + 3, cleanGenerations_);
				cleanTiles_ = tileTotal;
				cleanCanvases_ = (int32_t)nC;
			}
			return 0;
		}
		if (overflow & 1) {
			if (single) {
				// A setup block made more commands than its slots: every block gets room for that many.
				const int32_t most = hostCounters_[8];
				blockCapacity_ = std::max(blockCapacity_, ((most + 63) / 64) * 64);
			} else {
				commandCapacity_ = hostCounters_[0] + 1024;
			}
		}
		if (overflow & 2) {
			// Tile pairs no longer fit: the packed lists grow to the frame's pair count, the
			// direct bins to the longest tile list with a quarter to spare.
			pairCapacity_ = hostCounters_[1] + 1024;
			if (direct) {
				const int64_t longest = hostCounters_[6];
				binCapacity_ = (int32_t)std::min<int64_t>(((longest + longest / 4 + 15) / 16) * 16, BIN_ENTRY_LIMIT);
			}
		}
		if ((overflow & 1) && pairCapacity_ < 4 * commandCapacity_) pairCapacity_ = 4 * commandCapacity_;
		dropGraph();
	}
	return fail(-4, "renderer_end: buffers did not converge");
}

int Frame::occludeExisting(float *deviceGrid, int32_t gridW, int32_t gridH) {
	int rc = prepare();
	if (rc || !prepared_) return rc;
	const int32_t pointTotal = taskPointEnd_.back();
	const int32_t fanTotal = taskFanEnd_.back();
	const int32_t tileTotal = canvasTileEnd_.back();
	if (fanTotal == 0) return 0;
	const size_t nC = canvases_.size();
	uint8_t *base = (uint8_t *)dTables_.ptr;
	FrameDev f;
	std::memset(&f, 0, sizeof(f));
	f.tasks = (const TaskDev *)(base + offTasks_);
	f.taskCount = (int32_t)tasks_.size();
	f.taskFanEnd = (const int32_t *)(base + offFanEnd_);
	f.taskPointEnd = (const int32_t *)(base + offPointEnd_);
	f.fanTotal = fanTotal;
	f.pointTotal = pointTotal;
	f.canvases = (const CanvasDev *)(base + offCanvases_);
	f.canvasCount = (int32_t)nC;
	f.canvasTileEnd = (const int32_t *)(base + offTileEnd_);
	f.tileTotal = tileTotal;
	f.textures = (const TextureDesc *)(base + offTextures_);
	f.textureCount = (int32_t)textures_.size();
	f.occlusionGrid = deviceGrid;
	f.gridW = gridW;
	f.gridH = gridH;
	// The setup's occlusion mode projects each fan triangle's corners itself and touches no
	// counters, so nothing runs before it.
	hipLaunchKernelGGL(k_occlude_existing, dim3(blocksOf(fanTotal, 256)), dim3(256), 0, stream, f);
	HIP_TRY(hipGetLastError());
	HIP_TRY(hipStreamSynchronize(stream));
	return 0;
}

int Frame::backward(float *gradTexture, float *gradMultipliers, const RenderOptions &o) {
	if (!lastRunValid_ || lastRun_.mode != 1) return fail(-2, "backward needs a differentiable forward pass first");
	FrameDev f = lastRun_;
	f.gradTexture = gradTexture;
	f.gradMultipliers = gradMultipliers;
	int32_t *head = f.counters + 5;
	const int32_t grid = o.persistentPerCU > 0 ? deviceCUs_ * o.persistentPerCU : f.tileTotal;
	auto enqueueBackward = [&]() -> int {
		hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, stream, head, 1, 0);
		hipLaunchKernelGGL(k_backward, dim3(grid), dim3(256), 0, stream, f, head);
		HIP_TRY(hipGetLastError());
		return 0;
	};
	if (timing) HIP_TRY(hipEventRecord(events_[0], stream));
	if (o.graph) {
		std::vector<uint8_t> key(sizeof(FrameDev) + sizeof(int32_t));
		std::memcpy(key.data(), &f, sizeof(FrameDev));
		std::memcpy(key.data() + sizeof(FrameDev), &grid, sizeof(int32_t));
		if (!bwdGraphExec_ || key != bwdGraphKey_) {
			if (bwdGraphExec_) (void)hipGraphExecDestroy(bwdGraphExec_);
			if (bwdGraph_) (void)hipGraphDestroy(bwdGraph_);
			bwdGraphExec_ = nullptr;
			bwdGraph_ = nullptr;
			HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
			int rc = enqueueBackward();
			hipGraph_t g = nullptr;
			hipError_t e = hipStreamEndCapture(stream, &g);
			if (rc) return rc;
			if (e != hipSuccess) return fail(-3, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
			bwdGraph_ = g;
			HIP_TRY(hipGraphInstantiate(&bwdGraphExec_, bwdGraph_, nullptr, nullptr, 0));
			bwdGraphKey_ = key;
		}
		HIP_TRY(hipGraphLaunch(bwdGraphExec_, stream));
	} else {
		int rc = enqueueBackward();
		if (rc) return rc;
	}
	if (timing) HIP_TRY(hipEventRecord(events_[4], stream));
	HIP_TRY(hipStreamSynchronize(stream));
	if (timing) HIP_TRY(hipEventElapsedTime(&stageMs[4], events_[0], events_[4]));
	return 0;
}

}  // namespace dfpsr
