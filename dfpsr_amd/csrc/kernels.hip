// gfx950 kernels of the polygon path: projection, triangle setup and clipping, ordered tile
// binning and the tile raster/shade kernel that reproduces the reference's fill loops.
//
// Reference files (relative to its Source/ folder): model projection modelAPI.cpp:214-281,
// setup and clipping renderCore.cpp:172-341, row intervals ITriangle2D.cpp:70-176, weight
// planes ITriangle2D.cpp:182-300, the 2x2 filler shader/fillerTemplates.h:92-385, shading
// shader/RgbaMultiply.h:74-175.
#include "pipeline.h"
#include "occlusion.h"

#include <type_traits>

namespace dfpsr {

// ---------------------------------------------------------------------------------------
// Projection of a task's point (modelAPI.cpp:237-242): a model point transformed to the world,
// to the camera and onto the screen; a camera-space corner projected; or the caller's
// ProjectedPoint as it is.  The setup projects the corners of each fan triangle it reads, so
// no pass over the points comes before it.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ Projected projectedPoint(const TaskDev &t, int32_t local) {
	if (t.pointKind == POINTS_PROJECTED) return reinterpret_cast<const Projected *>(t.points)[local];
	const f3 p = make3(t.points[local * 3], t.points[local * 3 + 1], t.points[local * 3 + 2]);
	if (t.pointKind == POINTS_CAMERA_SPACE) return cameraToScreen(t.camera, p);
	return cameraToScreen(t.camera, worldToCamera(t.camera.location, transformPoint(transformFrom(t.transform), p)));
}

// ---------------------------------------------------------------------------------------
// Triangle setup: renderTriangleFromData -> renderTriangleWithShader -> drawClippedTriangle
// -> drawSubTriangle (renderCore.cpp:220-341).  Returns the number of draw commands the fan
// triangle makes; with EMIT they are written to out[0..count).
// ---------------------------------------------------------------------------------------
DH Rect canvasRect(const CanvasDev &cv) { return rectSides(0, 0, cv.width, cv.height); }

// n / d for 0 <= n < 2^24 and d >= 1 with a quotient below 2^21 (a command's pair index over
// its tile columns: at most a canvas's tiles, and its tile rows): both operands are exact in
// float and the float quotient is within one of the integer one, which one step corrects -- no
// integer division sequence.
__device__ __forceinline__ int32_t divSmall(int32_t n, int32_t d) {
	int32_t q = (int32_t)((float)n * __builtin_amdgcn_rcpf((float)d));
	if (q * d > n) q--;
	else if ((q + 1) * d <= n) q++;
	return q;
}

// (command, tile) pairs and newly used tiles counted by one thread of the setup kernel.
struct BinTally { int32_t pairs, tiles; };

// A command's tiles as the setup block bins them: the first tile, the width of the tile
// rectangle, the canvas's tiles per row, and the command's first pair in the block's pairs.
struct BinJob { int32_t index, firstTile, width, tilesX, pairStart; };
constexpr int32_t BIN_JOBS = 512;  // commands a setup block bins together (more: each thread bins its own)
constexpr int32_t TILE_TABLE_LOG2 = 10, TILE_TABLE = 1 << TILE_TABLE_LOG2, TILE_TABLE_PROBES = 8;
static_assert(TILE_TABLE % SCAN_BLOCK == 0, "the tile table's entries split evenly over a setup block");

// Enters pair slot[k] of tile[k] into the tile's list (direct bins).
__device__ __forceinline__ void binPair(const FrameDev &f, int32_t tile, int32_t slot, int32_t index, BinTally &tally) {
	if (slot < f.binCapacity) {
		f.tileList[(size_t)tile * f.binCapacity + slot] = index;
	} else {
		atomicOr(&f.counters[2], 2);
		atomicMax(&f.counters[6], slot + 1);
	}
	tally.tiles += slot == 0 ? 1 : 0;
}

// Keeps pair record r (packed lists): the command's slot in the tile's list.
__device__ __forceinline__ void recordPair(const FrameDev &f, int32_t r, int32_t tile, int32_t slot, int32_t index) {
	if (r < f.pairCapacity) f.pairRecords[r] = PairRecord{tile, slot, index};
	else atomicOr(&f.counters[2], 2);
}

// Writes draw command `index` with everything the raster needs of it -- the command, its row
// setup and (for the RGBA8 shaders) its finished shader inputs -- and enters it into every tile
// it can touch: the columns of cut(wholeBound, canvas), the rows of the aligned raster bound
// (ITriangle2D.cpp:70-84).  With direct bins the index goes into the tile's list at once (in
// any order: the raster sorts each list); otherwise the tile's count grows here and a pair
// record (tile, slot, command) keeps the slot until the counts are scanned and the list is
// filled from the records.
__device__ void emitCommand(const FrameDev &f, int32_t index, const Command &cmd, const CanvasDev &cv,
                            const ShaderInputs &si, BinTally &tally, BinJob *jobs, int32_t jobBase) {
	f.commands[index] = cmd;
	const TriangleRows tr = triangleRows(l2{cmd.fx[0], cmd.fy[0]}, l2{cmd.fx[1], cmd.fy[1]}, l2{cmd.fx[2], cmd.fy[2]}, canvasRect(cv));
	f.triangleRows[index] = tr;
	if (f.mode == 0) f.shaders[index] = si;
	int32_t tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;  // no tiles when the bound misses the canvas
	if (tr.clipR > tr.clipL) {
		// The columns and rows are at least 0 here, so shifts divide.
		tx0 = tr.clipL >> f.tileShiftW;
		tx1 = (tr.clipR - 1) >> f.tileShiftW;
		ty0 = tr.alT >> f.tileShiftH;
		ty1 = tmin((tr.alB - 2) >> f.tileShiftH, cv.tilesY - 1);
	}
	const int32_t w = tx1 - tx0 + 1, count = tmax(w, 0) * tmax(ty1 - ty0 + 1, 0);
	if (jobs) {
		// Binned later by the whole block, pairs spread over its threads.  Every command writes
		// its job, with no pairs when it touches no tile.
		jobs[index - jobBase] = BinJob{index, cv.tileBase + ty0 * cv.tilesX + tx0, tmax(w, 1), cv.tilesX, tally.pairs};
		tally.pairs += count;
		return;
	}
	if (f.binCapacity == 0) {
		// Packed lists, this thread alone: records reserved for the command's pairs at once.
		const int32_t recordBase = count > 0 ? atomicAdd(&f.counters[7], count) : 0;
		for (int32_t i = 0; i < count; i++) {
			const int32_t tile = cv.tileBase + (ty0 + i / w) * cv.tilesX + tx0 + i % w;
			recordPair(f, recordBase + i, tile, atomicAdd(&f.tileCount[tile], 1), index);
		}
		return;
	}
	// Direct bins of this thread alone: the slot atomics of up to eight tiles are in flight
	// together before their list entries are written.
	constexpr int32_t BATCH = 8;
	for (int32_t base = 0; base < count; base += BATCH) {
		int32_t tile[BATCH], slot[BATCH];
#pragma unroll
		for (int k = 0; k < BATCH; k++) {
			const int32_t i = base + k;
			tile[k] = i < count ? cv.tileBase + (ty0 + i / w) * cv.tilesX + tx0 + i % w : -1;
			slot[k] = tile[k] >= 0 ? atomicAdd(&f.tileCount[tile[k]], 1) : 0;
		}
#pragma unroll
		for (int k = 0; k < BATCH; k++) {
			if (tile[k] < 0) continue;
			binPair(f, tile[k], slot[k], index, tally);
			tally.pairs++;
		}
	}
}

// The finished shader inputs of fan triangle (poly, k) of a task (Shader.h:37-57,
// RgbaMultiply.h:37-58): corner colours and texture coordinates, texture slots, the shader
// variant flags.
__device__ ShaderInputs triangleShader(const TaskDev &t, int32_t poly, int32_t k) {
	const float *col = t.colors + poly * 16;
	const float *tc = t.texcoords + poly * 16;
	const int32_t ca = 0, cb = (1 + k) * 4, cc = (2 + k) * 4;
	ShaderInputs si;
	si.red = make3(col[ca], col[cb], col[cc]);
	si.green = make3(col[ca + 1], col[cb + 1], col[cc + 1]);
	si.blue = make3(col[ca + 2], col[cb + 2], col[cc + 2]);
	si.alpha = make3(col[ca + 3], col[cb + 3], col[cc + 3]);
	si.u1 = make3(tc[ca], tc[cb], tc[cc]);
	si.v1 = make3(tc[ca + 1], tc[cb + 1], tc[cc + 1]);
	si.u2 = make3(tc[ca + 2], tc[cb + 2], tc[cc + 2]);
	si.v2 = make3(tc[ca + 3], tc[cb + 3], tc[cc + 3]);
	int32_t part = t.polygonPart[poly];
	si.diffuse = t.partTexture[part * 2];
	si.light = t.partTexture[part * 2 + 1];
	finishShaderInputs(si);
	return si;
}

// The projected corners of fan triangle g (of task ti) and its polygon.
struct TriangleCorners {
	Projected P[3];
	int32_t poly, k;
};

__device__ __forceinline__ TriangleCorners loadCorners(const FrameDev &f, int32_t g, int32_t ti) {
	const TaskDev &t = f.tasks[ti];
	const FanRecord fan = t.fans[g - t.fanBase];
	TriangleCorners tc;
	tc.poly = fan.polygonFan >> 1;
	tc.k = fan.polygonFan & 1;
	for (int c = 0; c < 3; c++) tc.P[c] = projectedPoint(t, fan.corner[c]);
	return tc;
}

// What setupTriangle does with the draw commands it makes: count them, write and bin them, or
// lower the occlusion grid under their hulls (renderer_occludeFromExistingTriangles).
enum { SETUP_COUNT = 0, SETUP_EMIT = 1, SETUP_OCCLUDE = 2 };

// occludeFromSortedHull of one Solid draw command's triangle (rendererAPI.cpp:218-239, with
// the triangle's wholeBound as the pixel bound, :246-258): every cell whose four corners are
// inside the triangle takes the larger corner depth if that is nearer.  The depth is +0 or
// more, whose float bits order as integers, so an integer atomic min keeps the nearest and
// leaves any negative cell value (sign bit set) in place, as the reference's '<' does.
__device__ void occludeFromCommand(const FrameDev &f, const Projected *P) {
	const Rect bound = triangleBound(P[0].flat, P[1].flat, P[2].flat);
	if (!hullCanOcclude(bound)) return;
	const float z[3] = {P[0].cs.z, P[1].cs.z, P[2].cs.z};
	const int bits = __float_as_int(hullDistance(z, 3));
	const l2 hull[3] = {P[0].flat, P[1].flat, P[2].flat};
	const Rect cells = outerCellBound(bound, f.gridW, f.gridH);
	for (int32_t cy = cells.t; cy < cells.bottom(); cy++)
		for (int32_t cx = cells.l; cx < cells.right(); cx++)
			if (cellInsideOfHull(hull, 3, cx, cy)) atomicMin(reinterpret_cast<int *>(f.occlusionGrid) + cy * f.gridW + cx, bits);
}

// Whether a command survives completeOcclusion (rendererAPI.cpp:190-215): always without a
// grid; otherwise when some cell under its bound is not nearer than its nearest corner.
__device__ __forceinline__ bool survivesOcclusion(const FrameDev &f, const Projected *P) {
	if (f.occlusionGrid == nullptr) return true;
	return commandVisibleInGrid(f.occlusionGrid, f.gridW, f.gridH, triangleBound(P[0].flat, P[1].flat, P[2].flat), P[0].cs.z,
	                            P[1].cs.z, P[2].cs.z);
}

// corners: the triangle's corners when the caller has loaded them already.  clipPoly: the
// calling wavefront's clipped polygon in LDS.  Triangles that need clipping are rare; the
// lanes holding one take the wavefront's polygon in turn, so that the polygon (nine
// sub-vertices indexed at run time) lives in LDS instead of private scratch memory.
// single: the caller has counted this triangle already and found one draw command, inside the
// clip frustum and front-facing (the tests are not repeated).  With SETUP_COUNT, *single is set
// when that is the case.  haveShader: shader holds the triangle's shader inputs already.
template <int MODE>
__device__ __forceinline__ int32_t setupTriangle(const FrameDev &f, int32_t g, int32_t ti, int32_t offset, BinTally &tally,
                                                 ClipPolygon *clipPoly, BinJob *jobs = nullptr, int32_t jobBase = 0,
                                                 const TriangleCorners *corners = nullptr, bool *single = nullptr,
                                                 bool haveShader = false, ShaderInputs shader = ShaderInputs()) {
	constexpr bool EMIT = MODE == SETUP_EMIT;
	const TaskDev &t = f.tasks[ti];
	const CanvasDev &cv = f.canvases[t.canvas];
	if (cv.color == nullptr && cv.depth == nullptr) return 0;
	int32_t local = g - t.fanBase;
	const TriangleCorners tc = corners ? *corners : loadCorners(f, g, ti);
	const int32_t poly = tc.poly, k = tc.k;
	const int32_t corner[3] = {0, 1 + k, 2 + k};
	const Projected *P = tc.P;
	const CameraParams &cam = t.camera;
	const bool known = EMIT && single != nullptr && *single;
	if (!known) {
		if (triangleVisibility(P[0].cs, P[1].cs, P[2].cs, cam.cull) == VIS_HIDDEN) return 0;
		if (t.filter != 0) {
			const float *col = t.colors + poly * 16;
			if (almostZero(col[corner[0] * 4 + 3]) && almostZero(col[corner[1] * 4 + 3]) && almostZero(col[corner[2] * 4 + 3])) return 0;
		}
	}
	int32_t count = 0;
	const bool perspective = cam.perspective != 0;
	ShaderInputs si;
	if (EMIT && f.mode == 0) si = haveShader ? shader : triangleShader(t, poly, k);
	if (known || triangleVisibility(P[0].cs, P[1].cs, P[2].cs, cam.clip) == VIS_FULL) {
		if (known || frontFacing(P[0].flat, P[1].flat, P[2].flat)) {
			if (MODE == SETUP_OCCLUDE) {
				if (t.filter == 0) occludeFromCommand(f, P);
				return 0;
			}
			if (!known && !survivesOcclusion(f, P)) return 0;
			if (MODE == SETUP_COUNT && single != nullptr) *single = true;
			if (EMIT) {
				Command c;
				const f2 is[3] = {P[0].is, P[1].is, P[2].is};
				Projection pr = weightPlanes(is, make3(P[0].cs.z, P[1].cs.z, P[2].cs.z), make3(0.0f, 1.0f, 0.0f), make3(0.0f, 0.0f, 1.0f), perspective);
				for (int q = 0; q < 3; q++) { c.fx[q] = P[q].flat.x; c.fy[q] = P[q].flat.y; }
				c.start = pr.start; c.dx = pr.dx; c.dy = pr.dy;
				c.task = ti; c.fan = local; c.canvas = t.canvas; c.flags = perspective ? 0 : 1;
				emitCommand(f, offset, c, cv, si, tally, jobs, jobBase);
			}
			count = 1;
		}
	} else {
		// The lanes here are exactly those that clip: one at a time, in lane order.
		const uint64_t clipping = __ballot(1);
		const int32_t me = (int32_t)(threadIdx.x & 63);
		for (uint64_t pending = clipping; pending != 0; pending &= pending - 1) {
		if (me != __builtin_ctzll(pending)) continue;
		ClipPolygon &poly9 = *clipPoly;
		poly9.init(P[0].cs, P[1].cs, P[2].cs);
		for (int32_t s = 0; s < cam.clip.count; s++) poly9.clip(cam.clip.planes[s]);
		for (int32_t i = 0; i < poly9.count - 2; i++) {
			const SubVertex &va = poly9.v[0], &vb = poly9.v[1 + i], &vc = poly9.v[2 + i];
			Projected S[3] = {cameraToScreen(cam, va.cs), cameraToScreen(cam, vb.cs), cameraToScreen(cam, vc.cs)};
			if (frontFacing(S[0].flat, S[1].flat, S[2].flat)) {
				if (MODE == SETUP_OCCLUDE) {
					if (t.filter == 0) occludeFromCommand(f, S);
					continue;
				}
				if (!survivesOcclusion(f, S)) continue;
				if (EMIT) {
					Command c;
					const f2 is[3] = {S[0].is, S[1].is, S[2].is};
					Projection pr = weightPlanes(is, make3(S[0].cs.z, S[1].cs.z, S[2].cs.z), make3(va.subB, vb.subB, vc.subB),
					                             make3(va.subC, vb.subC, vc.subC), perspective);
					for (int q = 0; q < 3; q++) { c.fx[q] = S[q].flat.x; c.fy[q] = S[q].flat.y; }
					c.start = pr.start; c.dx = pr.dx; c.dy = pr.dy;
					c.task = ti; c.fan = local; c.canvas = t.canvas; c.flags = perspective ? 0 : 1;
					emitCommand(f, offset + count, c, cv, si, tally, jobs, jobBase);
				}
				count++;
			}
		}
		}
	}
	return count;
}

// Exclusive prefix over one value per thread of an NT-thread workgroup (NT a multiple of 64);
// returns the total.  Each wavefront scans its 64 values with lane shuffles, then the
// wavefront totals meet in waveTotals[NT / 64]: two barriers in all.
template <int NT>
__device__ __forceinline__ int32_t workgroupPrefix(int32_t value, int32_t &before, int32_t *waveTotals) {
	const int32_t lane = (int32_t)(threadIdx.x & 63), wave = (int32_t)(threadIdx.x >> 6);
	int32_t x = value;
#pragma unroll
	for (int32_t offset = 1; offset < 64; offset <<= 1) {
		const int32_t y = __shfl_up(x, offset, 64);
		if (lane >= offset) x += y;
	}
	if (NT == 64) {
		__syncthreads();  // a barrier like the multi-wavefront form
		before = x - value;
		return __shfl(x, 63, 64);
	}
	if (lane == 63) waveTotals[wave] = x;
	__syncthreads();
	int32_t waveBase = 0, total = 0;
#pragma unroll
	for (int32_t w = 0; w < NT / 64; w++) {
		const int32_t v = waveTotals[w];
		waveBase += w < wave ? v : 0;
		total += v;
	}
	__syncthreads();  // waveTotals may be written again right after
	before = waveBase + x - value;
	return total;
}

// Block-wide exclusive scan of one int32 per thread of a SCAN_BLOCK-thread block; returns the
// block total.
__device__ int32_t blockExclusiveScan(int32_t value, int32_t &exclusive) {
	__shared__ int32_t waveTotals[SCAN_BLOCK / 64];
	return workgroupPrefix<SCAN_BLOCK>(value, exclusive, waveTotals);
}

// The task of fan triangle g, and whether every active lane of the wavefront has the same one.
// A wavefront's triangles nearly always come from one task; the task index is then made
// wave-uniform (readfirstlane) so that the task's fields -- camera planes, array bases -- are
// read with scalar loads instead of one vector load per lane.
__device__ __forceinline__ int32_t taskOf(const FrameDev &f, int32_t g, bool &uniform) {
	const int32_t ti = findSegment(f.taskFanEnd, f.taskCount, g);
	const int32_t first = __builtin_amdgcn_readfirstlane(ti);
	uniform = __all(ti == first);
	return ti;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_setup_count(FrameDev f) {
	__shared__ ClipPolygon clipPolys[SCAN_BLOCK / 64];
	int32_t g = blockIdx.x * SCAN_BLOCK + threadIdx.x;
	BinTally unused{0, 0};
	int32_t n = 0;
	if (g < f.fanTotal) {
		bool uniform;
		const int32_t ti = taskOf(f, g, uniform);
		ClipPolygon *poly = &clipPolys[threadIdx.x >> 6];
		if (uniform) n = setupTriangle<SETUP_COUNT>(f, g, __builtin_amdgcn_readfirstlane(ti), 0, unused, poly);
		else n = setupTriangle<SETUP_COUNT>(f, g, ti, 0, unused, poly);
	}
	if (g < f.fanTotal) f.fanCount[g] = n;
	int32_t ex;
	int32_t total = blockExclusiveScan(n, ex);
	if (threadIdx.x == 0) f.blockSums[blockIdx.x] = total;
}

// Exclusive scan of the per-block sums by one workgroup; writes the grand total.  Each thread
// takes a run of PER consecutive sums per step, so a step covers SCAN_BLOCK * PER of them.
__global__ void k_scan_block_sums(int32_t *blockSums, int32_t blocks, int32_t *total) {
	constexpr int32_t PER = 16;
	__shared__ int32_t carry;
	if (threadIdx.x == 0) carry = 0;
	__syncthreads();
	for (int32_t base = 0; base < blocks; base += SCAN_BLOCK * PER) {
		const int32_t first = base + (int32_t)threadIdx.x * PER;
		int32_t v[PER], run = 0;
#pragma unroll
		for (int k = 0; k < PER; k++) {
			v[k] = first + k < blocks ? blockSums[first + k] : 0;
			run += v[k];
		}
		int32_t ex;
		const int32_t sum = blockExclusiveScan(run, ex);
		const int32_t c = carry;
		int32_t at = c + ex;
#pragma unroll
		for (int k = 0; k < PER; k++) {
			if (first + k < blocks) blockSums[first + k] = at;
			at += v[k];
		}
		__syncthreads();
		if (threadIdx.x == 0) carry = c + sum;
		__syncthreads();
	}
	if (threadIdx.x == 0) *total = carry;
}

// The same over int64 values.
__device__ long long blockExclusiveScan64(long long value, long long &exclusive) {
	__shared__ long long s[SCAN_BLOCK];
	int32_t tid = threadIdx.x;
	s[tid] = value;
	__syncthreads();
	for (int32_t offset = 1; offset < SCAN_BLOCK; offset <<= 1) {
		long long v = (tid >= offset) ? s[tid - offset] : 0;
		__syncthreads();
		s[tid] += v;
		__syncthreads();
	}
	long long inclusive = s[tid];
	long long total = s[SCAN_BLOCK - 1];
	__syncthreads();
	exclusive = inclusive - value;
	return total;
}

// Exclusive scan of int64 block sums by one workgroup; writes the two 32-bit halves of the
// grand total.
__global__ void k_scan_block_sums64(long long *blockSums, int32_t blocks, int32_t *totalLow, int32_t *totalHigh) {
	__shared__ long long carry;
	if (threadIdx.x == 0) carry = 0;
	__syncthreads();
	for (int32_t base = 0; base < blocks; base += SCAN_BLOCK) {
		int32_t i = base + threadIdx.x;
		long long v = (i < blocks) ? blockSums[i] : 0;
		long long ex;
		long long sum = blockExclusiveScan64(v, ex);
		long long c = carry;
		if (i < blocks) blockSums[i] = c + ex;
		__syncthreads();
		if (threadIdx.x == 0) carry = c + sum;
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		*totalLow = (int32_t)(carry & 0xFFFFFFFFll);
		*totalHigh = (int32_t)(carry >> 32);
	}
}

// sumBlocks = 1: each block adds up the totals of the blocks before it (and the last block writes
// the command total) instead of reading offsets scanned by k_scan_block_sums.
// Writes the draw commands of one block of SCAN_BLOCK fan triangles.  Two-pass form (SINGLE =
// false): k_setup_count has stored each triangle's command count, and the commands take dense
// indices from the prefix over all blocks.  Single-pass form (SINGLE = true): the block counts
// its own commands and writes them from index blockIdx.x * blockCapacity on, so no block waits
// for another and no count kernel runs before; indices keep submission order but leave gaps
// (the raster only orders commands by index).  A block with more commands than blockCapacity
// writes nothing and reports its count, and the frame runs again with room for it.
template <bool SINGLE>
__device__ __forceinline__ void setupEmit(const FrameDev &f, int32_t sumBlocks) {
	__shared__ ClipPolygon clipPolys[SCAN_BLOCK / 64];
	__shared__ int32_t blockTally[2];
	__shared__ int32_t recordBaseLds;
	__shared__ BinJob jobs[BIN_JOBS];
	__shared__ int32_t tableTile[TILE_TABLE];
	__shared__ int32_t tableCount[TILE_TABLE];
	int32_t g = blockIdx.x * SCAN_BLOCK + threadIdx.x;
	for (int32_t e = threadIdx.x; e < TILE_TABLE; e += SCAN_BLOCK) {  // empty before the prefix sums' barriers
		tableTile[e] = -1;
		tableCount[e] = 0;
	}
	const int32_t ti = g < f.fanTotal ? findSegment(f.taskFanEnd, f.taskCount, g) : 0;
	TriangleCorners corners;
	ShaderInputs shader;
	int32_t n;
	bool single = false;  // counted as one unclipped, front-facing command
	if (SINGLE) {
		// The corners and the shader inputs of a triangle that makes commands stay in registers
		// from the count to the emit; the shader inputs' loads are issued before the block's
		// prefix sums, so that their latency overlaps them.
		n = 0;
		if (g < f.fanTotal) {
			BinTally unused{0, 0};
			corners = loadCorners(f, g, ti);
			n = setupTriangle<SETUP_COUNT>(f, g, ti, 0, unused, &clipPolys[threadIdx.x >> 6], nullptr, 0, &corners, &single);
			if (n > 0 && f.mode == 0) shader = triangleShader(f.tasks[ti], corners.poly, corners.k);
		}
	} else {
		n = (g < f.fanTotal) ? f.fanCount[g] : 0;
		// The corners are loaded before the block's prefix sums, so their latency overlaps them.
		if (n > 0) corners = loadCorners(f, g, ti);
	}
	int32_t ex;
	const int32_t blockCommands = blockExclusiveScan(n, ex);
	if (threadIdx.x < 2) blockTally[threadIdx.x] = 0;
	int32_t jobBase;
	bool room;
	if (SINGLE) {
		jobBase = (int32_t)blockIdx.x * f.blockCapacity;
		room = blockCommands <= f.blockCapacity;
		if (threadIdx.x == 0 && blockCommands > 0) {
			atomicAdd(&f.counters[0], blockCommands);
			if (!room) {
				atomicOr(&f.counters[2], 1);
				atomicMax(&f.counters[8], blockCommands);
			}
		}
	} else {
		if (sumBlocks) {
			int32_t partial = 0, unusedPrefix;
			for (int32_t i = threadIdx.x; i < (int32_t)blockIdx.x; i += SCAN_BLOCK) partial += f.blockSums[i];
			jobBase = blockExclusiveScan(partial, unusedPrefix);
			if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) f.counters[0] = jobBase + blockCommands;
		} else {
			jobBase = f.blockSums[blockIdx.x];
		}
		room = jobBase + blockCommands <= f.commandCapacity;
	}
	// The block's commands are binned together (below) when they fit the table and the command
	// arrays (a block past the capacity leaves its jobs unwritten).
	const bool together = blockCommands <= BIN_JOBS && room;
	BinTally tally{0, 0};
	if (g < f.fanTotal) {
		const int32_t offset = jobBase + ex;
		// The commands of a canvas's tasks are consecutive, so the offset at the first fan
		// triangle of each task bounds its canvas's first command: one atomic per task.
		if (g == f.tasks[ti].fanBase) atomicMin(&f.canvasCommandBase[f.tasks[ti].canvas], offset);
		if (n > 0) {
			if (!SINGLE && offset + n > f.commandCapacity) atomicOr(&f.counters[2], 1);
			else if (room) setupTriangle<SETUP_EMIT>(f, g, ti, offset, tally, &clipPolys[threadIdx.x >> 6], together ? jobs : nullptr, jobBase, &corners,
			                                         SINGLE ? &single : nullptr, SINGLE && f.mode == 0, shader);
		}
	}
	if (together) {
		// Pairs of the block's commands, one per thread and round: job j owns pairs
		// [pairStart, pairStart + width * height) in command order.  A command whose setup ran
		// out of room (capacity) wrote no job; then no command of the frame is drawn anyway.
		int32_t pairBase;
		const int32_t pairs = blockExclusiveScan(tally.pairs, pairBase);
		for (int32_t k = 0; k < n; k++) jobs[ex + k].pairStart += pairBase;
		if (threadIdx.x == 0 && f.binCapacity == 0) recordBaseLds = pairs > 0 ? atomicAdd(&f.counters[7], pairs) : 0;
		__syncthreads();
		const int32_t recordBase = recordBaseLds;
		const int32_t jobCount = blockCommands;
		BinTally binned{0, 0};
		// Rounds of up to UNROLL pairs per thread.  The pairs of a round that go to one tile take
		// their slots with one global atomic: each pair draws a rank from the tile's entry in an
		// LDS table (TILE_TABLE entries, open addressing), one thread per entry then adds the
		// tile's count to its global count, and the pair's slot is the returned base plus its
		// rank.  A pair that finds no entry takes its slot with a global atomic of its own.
		constexpr int32_t UNROLL = 4;
		for (int32_t round = 0; round < pairs; round += UNROLL * SCAN_BLOCK) {
			int32_t tile[UNROLL], slot[UNROLL], index[UNROLL], entry[UNROLL];
#pragma unroll
			for (int u = 0; u < UNROLL; u++) {
				const int32_t p = round + (int32_t)threadIdx.x + u * SCAN_BLOCK;
				tile[u] = -1;
				entry[u] = -1;
				if (p < pairs) {
					int32_t lo = 0, hi = jobCount - 1;
					while (lo < hi) {
						const int32_t mid = (lo + hi + 1) >> 1;
						if (jobs[mid].pairStart <= p) lo = mid; else hi = mid - 1;
					}
					const BinJob &j = jobs[lo];
					const int32_t i = p - j.pairStart;
					const int32_t row = divSmall(i, j.width);
					tile[u] = j.firstTile + row * j.tilesX + (i - row * j.width);
					index[u] = j.index;
					const int32_t h = (int32_t)(((uint32_t)tile[u] * 2654435761u) >> (32 - TILE_TABLE_LOG2));
					for (int32_t q = 0; q < TILE_TABLE_PROBES; q++) {
						const int32_t e = (h + q) & (TILE_TABLE - 1);
						const int32_t old = atomicCAS(&tableTile[e], -1, tile[u]);
						if (old == -1 || old == tile[u]) {
							entry[u] = e;
							break;
						}
					}
					slot[u] = entry[u] >= 0 ? atomicAdd(&tableCount[entry[u]], 1) : atomicAdd(&f.tileCount[tile[u]], 1);
				}
			}
			__syncthreads();
			{
				// Every entry's global atomic is issued before any of the returned bases is used.
				constexpr int32_t PER = TILE_TABLE / SCAN_BLOCK;
				int32_t entryTile[PER], entryBase[PER];
#pragma unroll
				for (int q = 0; q < PER; q++) entryTile[q] = tableTile[(int32_t)threadIdx.x + q * SCAN_BLOCK];
#pragma unroll
				for (int q = 0; q < PER; q++)
					entryBase[q] = entryTile[q] >= 0 ? atomicAdd(&f.tileCount[entryTile[q]], tableCount[(int32_t)threadIdx.x + q * SCAN_BLOCK]) : 0;
#pragma unroll
				for (int q = 0; q < PER; q++)
					if (entryTile[q] >= 0) tableCount[(int32_t)threadIdx.x + q * SCAN_BLOCK] = entryBase[q];
			}
			__syncthreads();
#pragma unroll
			for (int u = 0; u < UNROLL; u++) {
				if (tile[u] < 0) continue;
				const int32_t s = entry[u] >= 0 ? tableCount[entry[u]] + slot[u] : slot[u];
				if (f.binCapacity > 0) binPair(f, tile[u], s, index[u], binned);
				else recordPair(f, recordBase + round + (int32_t)threadIdx.x + u * SCAN_BLOCK, tile[u], s, index[u]);
			}
			if (round + UNROLL * SCAN_BLOCK >= pairs) break;  // the last round: the table is not needed again
			__syncthreads();
			for (int32_t e = threadIdx.x; e < TILE_TABLE; e += SCAN_BLOCK) {
				tableTile[e] = -1;
				tableCount[e] = 0;
			}
			__syncthreads();
		}
		tally.tiles = binned.tiles;
		tally.pairs = threadIdx.x == 0 ? pairs : 0;
	}
	// The pair and used-tile totals, one global atomic per block.
	__syncthreads();
	if (tally.pairs) atomicAdd(&blockTally[0], tally.pairs);
	if (tally.tiles) atomicAdd(&blockTally[1], tally.tiles);
	__syncthreads();
	if (threadIdx.x == 0 && blockTally[0]) atomicAdd(&f.counters[1], blockTally[0]);
	if (threadIdx.x == 1 && blockTally[1]) atomicAdd(&f.counters[3], blockTally[1]);
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_setup_emit(FrameDev f, int32_t sumBlocks) { setupEmit<false>(f, sumBlocks); }
__global__ void __launch_bounds__(SCAN_BLOCK) k_setup_single(FrameDev f) { setupEmit<true>(f, 0); }

// ---------------------------------------------------------------------------------------
// Binning into screen tiles.  List slots come from atomics on the tile counts, so a list is
// in no particular order; the raster kernel sorts it by command index, which restores
// submission order per pixel.
// ---------------------------------------------------------------------------------------

// Tile offsets and the ordered queue of non-empty tiles come from one scan of
// (count | nonEmpty << 32); the block-local queue position waits in tileFill in between.
__global__ void k_scan_tiles_local(FrameDev f) {
	int32_t i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
	int32_t v = (i < f.tileTotal) ? f.tileCount[i] : 0;
	long long ex;
	long long total = blockExclusiveScan64((long long)v | ((long long)(v > 0) << 32), ex);
	if (i < f.tileTotal) {
		f.tileOffset[i] = (int32_t)(ex & 0xFFFFFFFFll);
		f.tileFill[i] = (int32_t)(ex >> 32);
	}
	if (threadIdx.x == 0) ((long long *)f.blockSums)[blockIdx.x] = total;
}

__global__ void k_scan_tiles_add(FrameDev f) {
	int32_t i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
	if (i < f.tileTotal) {
		const long long base = ((const long long *)f.blockSums)[blockIdx.x];
		f.tileOffset[i] += (int32_t)(base & 0xFFFFFFFFll);
		if (f.tileCount[i] > 0) f.tileQueue[(int32_t)(base >> 32) + f.tileFill[i]] = i;
	}
	if (i == f.tileTotal - 1) {
		int32_t end = f.tileOffset[i] + f.tileCount[i];
		f.tileOffset[f.tileTotal] = end;
		if (end > f.pairCapacity) atomicOr(&f.counters[2], 2);
	}
}

// The packed lists from the pair records: each command goes to its slot after its tile's offset.
__global__ void k_bin_fill(FrameDev f) {
	const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
	if (f.counters[2] || r >= f.counters[7] || r >= f.pairCapacity) return;
	const PairRecord rec = f.pairRecords[r];
	f.tileList[f.tileOffset[rec.tile] + rec.slot] = rec.command;
}

// ---------------------------------------------------------------------------------------
// Tile raster / shade.
// ---------------------------------------------------------------------------------------

// Sort of n distinct ints in LDS by the workgroup's NL lanes (n <= TILE_SORT_CAP).  Up to 4 NL
// values are ranked (each lane counts the values smaller than each of its own and writes them
// to those places: two barriers); longer lists take a bitonic sort.
template <int NL>
__device__ void sortLds(int32_t *a, int32_t n) {
	if (n <= NL) {
		const int32_t i = (int32_t)threadIdx.x;
		const int32_t v = i < n ? a[i] : 0;
		int32_t rank = 0;
		for (int32_t j = 0; j < n; j++) rank += a[j] < v ? 1 : 0;
		__syncthreads();
		if (i < n) a[rank] = v;
		__syncthreads();
		return;
	}
	if (n <= 4 * NL) {
		const int32_t i = (int32_t)threadIdx.x;
		int32_t v[4], rank[4] = {0, 0, 0, 0};
		for (int k = 0; k < 4; k++) v[k] = i + k * NL < n ? a[i + k * NL] : 0x7FFFFFFF;
		for (int32_t j = 0; j < n; j++) {
			const int32_t aj = a[j];
			for (int k = 0; k < 4; k++) rank[k] += aj < v[k] ? 1 : 0;
		}
		__syncthreads();
		for (int k = 0; k < 4; k++)
			if (i + k * NL < n) a[rank[k]] = v[k];
		__syncthreads();
		return;
	}
	int32_t size = 1;
	while (size < n) size <<= 1;
	for (int32_t i = n + (int32_t)threadIdx.x; i < size; i += NL) a[i] = 0x7FFFFFFF;
	__syncthreads();
	for (int32_t k = 2; k <= size; k <<= 1) {
		for (int32_t j = k >> 1; j > 0; j >>= 1) {
			for (int32_t i = threadIdx.x; i < size; i += NL) {
				int32_t ixj = i ^ j;
				if (ixj > i) {
					int32_t x = a[i], y = a[ixj];
					bool up = (i & k) == 0;
					if ((x > y) == up) { a[i] = y; a[ixj] = x; }
				}
			}
			__syncthreads();
		}
	}
}

// The four pixels of a lane's quad, held in registers for the whole tile.  MODE 0 holds the
// RGBA8 colours of the reference's shaders; MODE 1 (differentiable) holds float RGBA and the
// visibility id (draw command * 4 + quad lane) of the last write.
template <int MODE> struct QuadPixels;
template <> struct QuadPixels<0> {
	uint32_t color[4];
	float depth[4];
	uint32_t colorDirty, depthDirty;
};
template <> struct QuadPixels<1> {
	float r[4], g[4], b[4], a[4];
	int32_t vis[4];
	float depth[4];
	uint32_t colorDirty, depthDirty;
};

template <typename T> __device__ __forceinline__ T pick4(const T *v, int s) {
	return s == 0 ? v[0] : (s == 1 ? v[1] : (s == 2 ? v[2] : v[3]));
}
template <typename T> __device__ __forceinline__ void put4(T *v, int s, T x) {
	if (s == 0) v[0] = x; else if (s == 1) v[1] = x; else if (s == 2) v[2] = x; else v[3] = x;
}

// One quad of one command: which lanes the command covers and their weight-plane values,
// stepped exactly as the reference's loops step them -- the row-pair logic of fillShapeSuper
// (fillerTemplates.h:246-385), the incremental stepping of fillRowSuper (:195-244) and the
// side clipping of clipPixels (:92-108).
struct QuadEval {
	float l0[4], l1[4], l2[4];  // depth (1/W or W) and the two depth-divided vertex weights per lane
	int slot[4];                // pixel written by each lane (a missing row aliases the other row)
	bool clip[4];
	bool clipSides;
};

// n steps of both rows' weights by d, one rounded addition at a time as the reference steps
// them (n <= 0: none).  Four steps per trip keep the loop control of lanes with different n
// off most steps; the additions and their order are the same.
__device__ __forceinline__ void stepPair(f3 &wU, f3 &wL, const f3 &d, int32_t n) {
	int32_t i = 0;
#pragma clang loop unroll(disable)
	for (; i + 4 <= n; i += 4) {
		wU = add3(wU, d); wL = add3(wL, d);
		wU = add3(wU, d); wL = add3(wL, d);
		wU = add3(wU, d); wL = add3(wL, d);
		wU = add3(wU, d); wL = add3(wL, d);
	}
#pragma clang loop unroll(disable)
	for (; i < n; i++) { wU = add3(wU, d); wL = add3(wL, d); }
}

// upper/lower: row intervals of rows py and py + 1 (py even).  Returns false when the command
// does not touch the quad at px (even).
__device__ __forceinline__ bool evalQuad(const Projection &cmd, RowInterval upper, RowInterval lower, int32_t px, int32_t py,
                                         int32_t canvasHeight, QuadEval &e) {
	int32_t outerStart = tmin(upper.left, lower.left);
	int32_t outerEnd = tmax(upper.right, lower.right);
	int32_t innerStart = tmax(upper.left, lower.left);
	int32_t innerEnd = tmin(upper.right, lower.right);
	int32_t outerBlockStart = (int32_t)((uint32_t)outerStart & ~1u);
	int32_t outerBlockEnd = (int32_t)(((uint32_t)outerEnd + 1u) & ~1u);
	int32_t innerBlockStart = (int32_t)(((uint32_t)innerStart + 1u) & ~1u);
	int32_t innerBlockEnd = (int32_t)((uint32_t)innerEnd & ~1u);
	if (py + 1 >= canvasHeight) lower.right = lower.left;  // the missing last row of an odd canvas
	bool hasTop = upper.right > upper.left;
	bool hasBottom = lower.right > lower.left;
	if (!(hasTop || hasBottom)) return false;
	if (px < outerBlockStart || px >= outerBlockEnd) return false;

	const f3 dx = cmd.dx;
	const f3 dx2 = scale3(dx, 2.0f);
	f3 wU = weightAtPixel(cmd, outerBlockStart, py);
	f3 wL = add3(wU, cmd.dy);
	const bool splitSpan = innerBlockEnd > innerBlockStart;
	const bool inner = splitSpan && px >= innerBlockStart && px < innerBlockEnd;  // unclipped inner span
	const bool right = splitSpan && px >= innerBlockEnd;                         // clipped, right of it
	// The lanes of a wavefront take the same steps in the same places whatever part of the
	// row pair their quads are in: the pair chain to the quad (left part) or to the inner
	// span's start; past the inner span, its whole length at once and the chain to the quad.
	stepPair(wU, wL, dx2, (((inner || right) ? innerBlockStart : px) - outerBlockStart) >> 1);
	if (right) {
		float quadCount = (float)((innerBlockEnd - innerBlockStart) / 2);
		wU = add3(wU, scale3(dx2, quadCount));
		wL = add3(wL, scale3(dx2, quadCount));
		stepPair(wU, wL, dx2, (px - innerBlockEnd) >> 1);
	}
	e.clipSides = !inner;
	e.l0[0] = wU.x; e.l0[1] = wU.x + dx.x; e.l0[2] = wL.x; e.l0[3] = wL.x + dx.x;
	e.l1[0] = wU.y; e.l1[1] = wU.y + dx.y; e.l1[2] = wL.y; e.l1[3] = wL.y + dx.y;
	e.l2[0] = wU.z; e.l2[1] = wU.z + dx.z; e.l2[2] = wL.z; e.l2[3] = wL.z + dx.z;
	if (inner) {
		// Inside the span the four lanes step on their own, a quad at a time.
		int32_t x = innerBlockStart;
#pragma clang loop unroll(disable)
		for (; x + 4 <= px; x += 4) {
			for (int l = 0; l < 4; l++) { e.l0[l] = e.l0[l] + dx2.x; e.l1[l] = e.l1[l] + dx2.y; e.l2[l] = e.l2[l] + dx2.z; }
			for (int l = 0; l < 4; l++) { e.l0[l] = e.l0[l] + dx2.x; e.l1[l] = e.l1[l] + dx2.y; e.l2[l] = e.l2[l] + dx2.z; }
		}
#pragma clang loop unroll(disable)
		for (; x < px; x += 2) {
			for (int l = 0; l < 4; l++) { e.l0[l] = e.l0[l] + dx2.x; e.l1[l] = e.l1[l] + dx2.y; e.l2[l] = e.l2[l] + dx2.z; }
		}
	}
	e.slot[0] = hasTop ? 0 : 2;
	e.slot[1] = hasTop ? 1 : 3;
	e.slot[2] = hasBottom ? 2 : 0;
	e.slot[3] = hasBottom ? 3 : 1;
	if (e.clipSides) {
		int32_t x2 = px + 1;
		e.clip[0] = px >= upper.left && px < upper.right;
		e.clip[1] = x2 >= upper.left && x2 < upper.right;
		e.clip[2] = px >= lower.left && px < lower.right;
		e.clip[3] = x2 >= lower.left && x2 < lower.right;
	} else {
		e.clip[0] = e.clip[1] = e.clip[2] = e.clip[3] = true;
	}
	return true;
}

__device__ __forceinline__ bool evalQuad(const Command &cmd, RowInterval upper, RowInterval lower, int32_t px, int32_t py,
                                         int32_t canvasHeight, QuadEval &e) {
	return evalQuad(Projection{cmd.start, cmd.dx, cmd.dy}, upper, lower, px, py, canvasHeight, e);
}

// The integer part of evalQuad: whether the command touches the quad at (px, py) and, per lane,
// the pixel it writes and whether the row's sides keep it -- without the weight chains.
struct QuadLanes {
	int slot[4];
	bool clip[4];
	bool clipSides;
};

__device__ __forceinline__ bool quadLanes(RowInterval upper, RowInterval lower, int32_t px, int32_t py, int32_t canvasHeight,
                                          QuadLanes &q) {
	const int32_t outerStart = tmin(upper.left, lower.left);
	const int32_t outerEnd = tmax(upper.right, lower.right);
	const int32_t innerStart = tmax(upper.left, lower.left);
	const int32_t innerEnd = tmin(upper.right, lower.right);
	const int32_t outerBlockStart = (int32_t)((uint32_t)outerStart & ~1u);
	const int32_t outerBlockEnd = (int32_t)(((uint32_t)outerEnd + 1u) & ~1u);
	const int32_t innerBlockStart = (int32_t)(((uint32_t)innerStart + 1u) & ~1u);
	const int32_t innerBlockEnd = (int32_t)((uint32_t)innerEnd & ~1u);
	if (py + 1 >= canvasHeight) lower.right = lower.left;
	const bool hasTop = upper.right > upper.left;
	const bool hasBottom = lower.right > lower.left;
	if (!(hasTop || hasBottom)) return false;
	if (px < outerBlockStart || px >= outerBlockEnd) return false;
	const bool inner = innerBlockEnd > innerBlockStart && px >= innerBlockStart && px < innerBlockEnd;
	q.clipSides = !inner;
	q.slot[0] = hasTop ? 0 : 2;
	q.slot[1] = hasTop ? 1 : 3;
	q.slot[2] = hasBottom ? 2 : 0;
	q.slot[3] = hasBottom ? 3 : 1;
	const int32_t x2 = px + 1;
	q.clip[0] = px >= upper.left && px < upper.right;
	q.clip[1] = x2 >= upper.left && x2 < upper.right;
	q.clip[2] = px >= lower.left && px < lower.right;
	q.clip[3] = x2 >= lower.left && x2 < lower.right;
	return true;
}

// Vertex weights of a lane (fillRowSuper, fillerTemplates.h:222-238): perspective divides the
// depth-divided weights by 1/W.
__device__ __forceinline__ void laneWeights(const QuadEval &e, int l, bool affine, float &wa, float &wb, float &wc) {
	if (affine) {
		wb = e.l1[l];
		wc = e.l2[l];
	} else {
		float w = refinedReciprocal(e.l0[l]);
		wb = e.l1[l] * w;
		wc = e.l2[l] * w;
	}
	wa = 1.0f - (wb + wc);
}

// Depth test of getVisibility (fillerTemplates.h:110-152).
template <int MODE>
__device__ __forceinline__ bool quadVisibility(const QuadEval &e, const QuadPixels<MODE> &q, bool depthRead, bool affine, bool *vis) {
	bool anyVis = false;
	for (int l = 0; l < 4; l++) {
		bool front = true;
		if (depthRead) {
			if (e.clipSides && !e.clip[l]) {
				front = false;
			} else {
				float old = pick4(q.depth, e.slot[l]);
				front = affine ? (e.l0[l] < old) : (e.l0[l] > old);
			}
		}
		vis[l] = e.clip[l] && front;
		anyVis |= vis[l];
	}
	return anyVis;
}

// Shader of one quad of command c for all four lanes: RgbaMultiply's getPixels_2x2 with
// the vertex weights of the lanes (RgbaMultiply.h:74-175, shaderMethods.h:39-94).  The
// shader inputs were finished once per source triangle at setup.
__device__ __forceinline__ void shadeCommand(const FrameDev &f, int32_t c, const QuadEval &e, bool affine, uint32_t mask,
                                             float *r, float *g, float *b, float *a) {
	float wa[4], wb[4], wc[4];
	for (int l = 0; l < 4; l++) laneWeights(e, l, affine, wa[l], wb[l], wc[l]);
	const ShaderInputs si = f.shaders[c];
	shadeQuad(si, f.textures, wa, wb, wc, mask, r, g, b, a);
}

// The reference's shading of one quad (fillQuadSuper, fillerTemplates.h:154-187, with the
// RgbaMultiply shaders of RgbaMultiply.h:74-175).
__device__ void drawCommand(const FrameDev &f, const Command &cmd, int32_t c, const QuadEval &e,
                            QuadPixels<0> &q, bool colorWrite, bool depthRead, bool depthWrite, bool alphaFilter) {
	const bool affine = (cmd.flags & 1) != 0;
	bool vis[4];
	if (!quadVisibility<0>(e, q, depthRead, affine, vis)) return;
	if (colorWrite) {
		float r[4], g[4], b[4], a[4];
		const uint32_t mask = (vis[0] ? 1u : 0u) | (vis[1] ? 2u : 0u) | (vis[2] ? 4u : 0u) | (vis[3] ? 8u : 0u);
		shadeCommand(f, c, e, affine, mask, r, g, b, a);
		if (alphaFilter) {
			float tr[4], tg[4], tb[4], ta[4];
			for (int l = 0; l < 4; l++) {
				uint32_t old = (!e.clipSides || vis[l]) ? pick4(q.color, e.slot[l]) : 0u;
				unpackColor(old, tr[l], tg[l], tb[l], ta[l]);
			}
			for (int l = 0; l < 4; l++) {
				float opacity = a[l] * (1.0f / 255.0f);
				float inv = 1.0f - opacity;
				r[l] = (r[l] * opacity) + (tr[l] * inv);
				g[l] = (g[l] * opacity) + (tg[l] * inv);
				b[l] = (b[l] * opacity) + (tb[l] * inv);
				a[l] = (a[l] * opacity) + (ta[l] * inv);
			}
		}
		for (int l = 0; l < 4; l++) {
			if (vis[l]) {
				put4(q.color, e.slot[l], packSaturated(r[l], g[l], b[l], a[l]));
				q.colorDirty |= 1u << e.slot[l];
			}
		}
	}
	if (depthWrite) {
		for (int l = 0; l < 4; l++) {
			if (vis[l]) {
				put4(q.depth, e.slot[l], e.l0[l]);
				q.depthDirty |= 1u << e.slot[l];
			}
		}
	}
}

// ---------------------------------------------------------------------------------------
// Differentiable shading: colour = interpolated per-point RGB multipliers x bilinear sample
// of a float RGBA texture (level 0), alpha = the texture's alpha.  The texel choice and the
// 8-bit sub-texel weights are those of texture_sample_bilinear at the highest resolution
// (textureAPI.h:342-438); the adjoint below scatters through the same weights.
// ---------------------------------------------------------------------------------------

struct BilinearTaps {
	uint32_t index[4];  // texel indices: upper left, upper right, lower left, lower right
	float weight[4];
};

DH BilinearTaps bilinearTaps(int32_t log2w, int32_t log2h, float u, float v) {
	uint32_t scaleU = 256u << log2w, scaleV = 256u << log2h;
	uint32_t maskX = (1u << log2w) - 1u, maskY = (1u << log2h) - 1u;
	uint32_t subX = truncU32((u + 256.0f) * fromU32(scaleU)) - 128u;
	uint32_t subY = truncU32((v + 256.0f) * fromU32(scaleV)) - 128u;
	float fx = (float)(subX & 255u) * (1.0f / 256.0f), fy = (float)(subY & 255u) * (1.0f / 256.0f);
	uint32_t left = (subX >> 8) & maskX, top = (subY >> 8) & maskY;
	uint32_t right = (left + 1u) & maskX, bottom = (top + 1u) & maskY;
	BilinearTaps t;
	t.index[0] = (top << log2w) | left;
	t.index[1] = (top << log2w) | right;
	t.index[2] = (bottom << log2w) | left;
	t.index[3] = (bottom << log2w) | right;
	t.weight[0] = (1.0f - fx) * (1.0f - fy);
	t.weight[1] = fx * (1.0f - fy);
	t.weight[2] = (1.0f - fx) * fy;
	t.weight[3] = fx * fy;
	return t;
}

// Inputs of the differentiable shader for one command's source triangle.
struct DiffInputs {
	f3 u, v;               // u1 and v1 at the three corners
	int32_t point[3];      // model point of each corner (per-point multipliers)
};

__device__ __forceinline__ DiffInputs diffInputs(const TaskDev &task, int32_t fan) {
	const FanRecord rec = task.fans[fan];
	int32_t poly = rec.polygonFan >> 1, k = rec.polygonFan & 1;
	const float *tc = task.texcoords + poly * 16;
	const int32_t ca = 0, cb = (1 + k) * 4, cc = (2 + k) * 4;
	DiffInputs d;
	d.u = make3(tc[ca], tc[cb], tc[cc]);
	d.v = make3(tc[ca + 1], tc[cb + 1], tc[cc + 1]);
	d.point[0] = rec.corner[0];
	d.point[1] = rec.corner[1];
	d.point[2] = rec.corner[2];
	return d;
}

__device__ __forceinline__ float multiplier(const FrameDev &f, int32_t point, int ch) {
	return f.multipliers ? f.multipliers[point * 3 + ch] : 1.0f;
}

__device__ void drawCommandDiff(const FrameDev &f, const Command &cmd, int32_t commandIndex, const TaskDev &task,
                                const QuadEval &e, QuadPixels<1> &q, bool depthRead, bool depthWrite) {
	const bool affine = (cmd.flags & 1) != 0;
	bool vis[4];
	if (!quadVisibility<1>(e, q, depthRead, affine, vis)) return;
	const DiffInputs d = diffInputs(task, cmd.fan);
	float m[3][3];
	for (int c = 0; c < 3; c++)
		for (int ch = 0; ch < 3; ch++) m[c][ch] = multiplier(f, d.point[c], ch);
	for (int l = 0; l < 4; l++) {
		if (!vis[l]) continue;
		float wa, wb, wc;
		laneWeights(e, l, affine, wa, wb, wc);
		float u = interp(d.u, wa, wb, wc), v = interp(d.v, wa, wb, wc);
		BilinearTaps taps = bilinearTaps(f.ftexLog2w, f.ftexLog2h, u, v);
		float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
		for (int k = 0; k < 4; k++) {
			const float4 texel = *(const float4 *)(f.ftexTexels + (size_t)taps.index[k] * 4);
			t[0] += taps.weight[k] * texel.x;
			t[1] += taps.weight[k] * texel.y;
			t[2] += taps.weight[k] * texel.z;
			t[3] += taps.weight[k] * texel.w;
		}
		const int s = e.slot[l];
		put4(q.r, s, (wa * m[0][0] + wb * m[1][0] + wc * m[2][0]) * t[0]);
		put4(q.g, s, (wa * m[0][1] + wb * m[1][1] + wc * m[2][1]) * t[1]);
		put4(q.b, s, (wa * m[0][2] + wb * m[1][2] + wc * m[2][2]) * t[2]);
		put4(q.a, s, t[3]);
		put4(q.vis, s, commandIndex * 4 + l);
		q.colorDirty |= 1u << s;
		if (depthWrite) {
			put4(q.depth, s, e.l0[l]);
			q.depthDirty |= 1u << s;
		}
	}
}

// One tile: load the lanes' pixels, draw every binned command in increasing command index
// (submission order), write back the pixels that changed.
// Loads the next round of a tile's command list into LDS, sorted by command index: the
// commands in (last, threshold] for the largest threshold that keeps the round within
// TILE_SORT_CAP (all of them in one round when the list fits).  Every one of the workgroup's
// NT threads calls it; returns the round's size.
template <int NT, int CAP = TILE_SORT_CAP>
__device__ int32_t nextRound(const int32_t *src, int32_t n, int32_t done, int32_t &last, int32_t *list, int32_t *roundInfo) {
	const int32_t t = threadIdx.x;
	int32_t take;
	if (done == 0 && n <= CAP) {
		for (int32_t i = t; i < n; i += NT) list[i] = src[i];
		take = n;
		last = 0x7FFFFFFE;
	} else {
		if (t == 0) { roundInfo[0] = last + 1; roundInfo[1] = 0x7FFFFFFE; }
		while (true) {
			__syncthreads();
			int32_t lo = roundInfo[0], hi = roundInfo[1];
			if (lo >= hi) break;
			int32_t mid = lo + ((hi - lo + 1) >> 1);
			__syncthreads();
			if (t == 0) roundInfo[2] = 0;
			__syncthreads();
			int32_t cnt = 0;
			for (int32_t i = t; i < n; i += NT) { int32_t v = src[i]; cnt += (v > last && v <= mid) ? 1 : 0; }
			if (cnt) atomicAdd(&roundInfo[2], cnt);
			__syncthreads();
			if (t == 0) { if (roundInfo[2] <= CAP) roundInfo[0] = mid; else roundInfo[1] = mid - 1; }
		}
		__syncthreads();
		int32_t threshold = roundInfo[0];
		__syncthreads();
		if (t == 0) roundInfo[2] = 0;
		__syncthreads();
		for (int32_t i = t; i < n; i += NT) {
			int32_t v = src[i];
			if (v > last && v <= threshold) list[atomicAdd(&roundInfo[2], 1)] = v;
		}
		__syncthreads();
		take = roundInfo[2];
		last = threshold;
	}
	__syncthreads();
	sortLds<NT>(list, take);
	return take;
}

// The three places where the in-order tile walk differs between the RGBA8 shaders (MODE 0)
// and the differentiable float shader (MODE 1), chosen by overload on the quad's pixel type.
__device__ __forceinline__ void loadQuadColor(QuadPixels<0> &q, int s, bool in, const CanvasDev &cv, const FrameDev &f, int32_t x, int32_t y) {
	q.color[s] = (in && cv.color && f.alphaAny) ? cv.color[(size_t)y * cv.colorStride + x] : 0u;
}
__device__ __forceinline__ void loadQuadColor(QuadPixels<1> &, int, bool, const CanvasDev &, const FrameDev &, int32_t, int32_t) {}

__device__ __forceinline__ void drawQuadCommand(const FrameDev &f, const Command &cmd, int32_t c, int32_t ci, const TaskDev &task,
                                                const QuadEval &e, QuadPixels<0> &q, bool colorWrite, bool depthRead, bool depthWrite,
                                                bool alphaFilter) {
	(void)ci;
	(void)task;
	drawCommand(f, cmd, c, e, q, colorWrite, depthRead, depthWrite, alphaFilter);
}
__device__ __forceinline__ void drawQuadCommand(const FrameDev &f, const Command &cmd, int32_t c, int32_t ci, const TaskDev &task,
                                                const QuadEval &e, QuadPixels<1> &q, bool, bool depthRead, bool depthWrite, bool) {
	drawCommandDiff(f, cmd, c - f.canvasCommandBase[ci], task, e, q, depthRead, depthWrite);
}

__device__ __forceinline__ void storeQuadColor(const QuadPixels<0> &q, int s, const CanvasDev &cv, int32_t x, int32_t y) {
	cv.color[(size_t)y * cv.colorStride + x] = q.color[s];
}
__device__ __forceinline__ void storeQuadColor(const QuadPixels<1> &q, int s, const CanvasDev &cv, int32_t x, int32_t y) {
	size_t p = (size_t)y * cv.diffStride + x;
	*(float4 *)(cv.diffColor + p * 4) = make_float4(q.r[s], q.g[s], q.b[s], q.a[s]);
	cv.visibility[p] = q.vis[s];
}

template <int TW, int TH, int MODE>
__device__ void rasterTile(const FrameDev &f, int32_t tile, int32_t *list, RowInterval *rows, int32_t *roundInfo) {
	constexpr int32_t NL = (TW / 2) * (TH / 2);
	const int32_t n = f.tileCount[tile];
	const int32_t ci = findSegment(f.canvasTileEnd, f.canvasCount, tile);
	const CanvasDev cv = f.canvases[ci];
	const int32_t local = tile - cv.tileBase;
	const int32_t tileX0 = (local % cv.tilesX) * TW;
	const int32_t tileY0 = (local / cv.tilesX) * TH;
	const int32_t lane = threadIdx.x;
	const int32_t px = tileX0 + 2 * (lane % (TW / 2));
	const int32_t py = tileY0 + 2 * (lane / (TW / 2));
	const int32_t *src = tileListOf(f, tile);

	QuadPixels<MODE> q;
	q.colorDirty = q.depthDirty = 0;
	for (int s = 0; s < 4; s++) {
		int32_t x = px + (s & 1), y = py + (s >> 1);
		bool in = x < cv.width && y < cv.height;
		q.depth[s] = (in && cv.depth) ? cv.depth[(size_t)y * cv.depthStride + x] : 0.0f;
		loadQuadColor(q, s, in, cv, f, x, y);
	}

	int32_t last = -1;
	int32_t done = 0;
	while (done < n) {
		const int32_t take = nextRound<NL>(src, n, done, last, list, roundInfo);
		for (int32_t i = 0; i < take; i++) {
			const int32_t c = __builtin_amdgcn_readfirstlane(list[i]);
			const Command cmd = f.commands[c];
			const TaskDev &task = f.tasks[cmd.task];
			const bool solid = task.filter == 0;
			const bool colorWrite = MODE == 1 || cv.color != nullptr;
			const bool depthRead = cv.depth != nullptr;
			const bool depthWrite = cv.depth != nullptr && (solid || !colorWrite);
			const bool alphaFilter = colorWrite && !solid;
			// Rows of this command inside the tile: rasterizeTriangle over the aligned bound.
			const TriangleRows &tr = f.triangleRows[c];
			const int32_t alT = tr.alT, alB = tr.alB;
			__syncthreads();
			if (lane < TH) rows[lane] = triangleRow(tr, tileY0 + lane);
			__syncthreads();
			QuadEval e;
			if (py >= alT && py < alB && evalQuad(cmd, rows[py - tileY0], rows[py + 1 - tileY0], px, py, cv.height, e)) {
				drawQuadCommand(f, cmd, c, ci, task, e, q, colorWrite, depthRead, depthWrite, alphaFilter);
			}
		}
		done += take;
		if (take == 0) break;
	}

	for (int s = 0; s < 4; s++) {
		int32_t x = px + (s & 1), y = py + (s >> 1);
		if (x >= cv.width || y >= cv.height) continue;
		if ((q.colorDirty >> s) & 1u) storeQuadColor(q, s, cv, x, y);
		if ((q.depthDirty >> s) & 1u) cv.depth[(size_t)y * cv.depthStride + x] = q.depth[s];
	}
}

// The raster kernel over the queue of non-empty tiles.  persistent = 0: workgroup b takes
// queue slot b (the grid covers every tile; surplus workgroups exit).  persistent = 1: a fixed
// grid whose workgroups pull slots from an atomic head until the queue is empty, so every
// workgroup reaches the exit once all tiles are handed out.
// The raster kernels start after every counter is final: workgroup 0 copies them to the
// frame's pinned host memory, which the host reads after the stream completes.
__device__ __forceinline__ void reportCounters(const FrameDev &f) {
	if (blockIdx.x == 0 && threadIdx.x < HOST_COUNTERS) f.hostCounters[threadIdx.x] = f.counters[threadIdx.x];
}

// A raster that leaves the next frame what k_frame_begin would (selfClean): the counters and the
// canvases' first-command indices come in two sets used by alternate frames, and workgroup 0
// resets the set this frame does not use; each tile's workgroup zeroes the tile's count once it
// has read it (tiles no workgroup visits have a count of zero).  The next frame then needs no
// begin kernel.  A frame whose raster returns early on an overflow leaves counts behind; the
// host runs the begin kernel after it.
__device__ __forceinline__ void resetNextCounters(const FrameDev &f) {
	if (!f.selfClean || blockIdx.x != 0) return;
	for (int32_t i = threadIdx.x; i < COUNTERS; i += blockDim.x) f.nextCounters[i] = 0;
	for (int32_t c = threadIdx.x; c < f.canvasCount; c += blockDim.x) f.nextCanvasBase[c] = 0x7FFFFFFF;
}

// Runs body(tile) over the tiles with commands -- every tile when the raster fills the
// canvases -- XCD by XCD.  The slots are the ordered queue of non-empty tiles after the scan,
// or the tiles themselves with direct bins (empty ones skipped).  Without persistence the grid
// has xcdGrid(tiles) workgroups and workgroup b takes the (b / 8)-th slot of XCD b % 8.
// Persistent workgroups pull slots from their own XCD's head and then from the other XCDs'
// heads; a head only grows and every XCD's slots end below the slot count, so all of them exit.
template <typename Body>
__device__ void forEachQueuedTile(const FrameDev &f, int32_t persistent, int32_t *slotLds, Body body) {
	const bool allTiles = f.clearInRaster || f.binCapacity > 0;
	const int32_t queued = allTiles ? f.tileTotal : f.counters[3];
	const int32_t x = (int32_t)(blockIdx.x % XCDS);
	auto visit = [&](int32_t slot) {
		const int32_t tile = allTiles ? slot : f.tileQueue[slot];
		if (f.clearInRaster || f.tileCount[tile] > 0) body(tile);
	};
	if (!persistent) {
		const int32_t slot = xcdSlot(x, (int32_t)(blockIdx.x / XCDS));
		if (slot < queued) visit(slot);
		return;
	}
	// Persistent workgroups walk their own XCD's share first, then help the others.
	for (int32_t r = 0; r < XCDS; r++) {
		const int32_t xr = (x + r) % XCDS;
		while (true) {
			if (threadIdx.x == 0) *slotLds = xcdSlot(xr, atomicAdd(&f.counters[QUEUE_HEADS + QUEUE_HEAD_STRIDE * xr], 1));
			__syncthreads();
			const int32_t slot = *slotLds;
			__syncthreads();
			if (slot >= queued) break;
			visit(slot);
			__syncthreads();
		}
	}
}

template <int TW, int TH>
__global__ void __launch_bounds__((TW / 2) * (TH / 2)) k_raster(FrameDev f, int32_t persistent) {
	if (f.mode == 1) return;  // the differentiable frames use k_raster_diff
	__shared__ int32_t list[TILE_SORT_CAP];
	__shared__ RowInterval rows[TH];
	__shared__ int32_t roundInfo[3];
	__shared__ int32_t slotLds;
	reportCounters(f);
	if (f.counters[2]) return;  // an earlier stage ran out of room: leave the canvases untouched
	forEachQueuedTile(f, persistent, &slotLds, [&](int32_t tile) { rasterTile<TW, TH, 0>(f, tile, list, rows, roundInfo); });
}

template __global__ void k_raster<16, 16>(FrameDev, int32_t);
template __global__ void k_raster<32, 16>(FrameDev, int32_t);
template __global__ void k_raster<32, 32>(FrameDev, int32_t);

// ---------------------------------------------------------------------------------------
// Visibility raster for frames whose commands are all Solid and perspective on canvases of
// even height.
//
// In the reference's order a pixel ends with the colour and depth of the first command
// reaching the largest depth above the pixel's starting depth: a command writes where its
// 1/W is strictly greater than the depth held at that moment (fillerTemplates.h:110-152),
// so the held depth is always the largest value written so far and a later equal value
// loses.  That winner is found without walking the commands in order: every covered lane of
// every command offers the key (depth order << 32 | ~command) to the pixel with an LDS
// atomic max, and only the winning command is shaded for each pixel, with the same quad
// evaluation, weights and mip selection as the in-order walk.  Without a depth canvas the key
// is the command alone and the last covering command wins.  Commands run in chunks per tile;
// each chunk starts from the depths the previous one left.  Work is spread densely: one
// thread per (command, row) for the exact rows, per (command, covered quad) for the keys and
// per (command, covered quad) again for the shading of the lanes each command won.
//
// Frames with the Alpha filter (blending reads the colour held at that moment), orthogonal
// cameras (a lower depth wins), canvases wider than 32767 pixels (rows are kept in 16 bits) or
// an odd-height canvas (the missing last row aliases lanes
// onto the row above within one command) take the in-order kernel above.
// ---------------------------------------------------------------------------------------


// Keys order depths as floats compare: larger depth, larger key; -0 equals +0.
__device__ __forceinline__ uint32_t depthOrder(float d) {
	uint32_t u = __float_as_uint(d == 0.0f ? 0.0f : d);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// A row interval in LDS: columns fit 16 bits (the visibility raster takes canvases up to
// 32767 pixels wide).
struct RowSpan { int16_t left, right; };
__device__ __forceinline__ RowSpan rowSpan(RowInterval r) { return RowSpan{(int16_t)r.left, (int16_t)r.right}; }
__device__ __forceinline__ RowInterval rowInterval(RowSpan r) { return RowInterval{r.left, r.right}; }

// CHUNK: commands staged and drawn at once.  Sixteen keep a 32x16 workgroup of the RGBA8 shaders
// within the LDS of ten workgroups per CU, the occupancy of five waves per SIMD.  The persistent
// grid's 32-pixel-wide tiles take 32 (at most eight workgroups per CU run anyway), and so do all
// tiles of the differentiable shader, whose scenes are denser, so that fewer tiles need a
// second chunk.
template <int TW, int TH, int P>
constexpr int32_t colorChunk() { return P != 0 && TW * TH >= 512 ? 32 : 16; }
constexpr int32_t DIFF_CHUNK = 32;

// The thread's index as a value the compiler cannot see through.  Taken at the top of a loop's
// body, the values derived from it there are computed in each trip instead of being held in
// registers across the loop (the visibility raster's per-chunk work would otherwise keep a
// dozen of them live through a whole tile, in scratch memory at five waves per SIMD).
__device__ __forceinline__ int32_t opaqueThreadIndex() {
	int32_t t = (int32_t)threadIdx.x;
	asm volatile("" : "+v"(t));
	return t;
}

// Threads of a visibility-raster workgroup: four pixels of the tile per thread.
template <int TW, int TH, int CH> struct DeferShape {
	static constexpr int32_t NT = TW * TH / 4;
	static constexpr int32_t CHUNK = CH;                   // commands per chunk
	static constexpr int32_t LIST = 256;                    // commands sorted per round
};

constexpr int32_t LDS_TEXTURES = 16;  // texture tables up to this size are read from LDS

// The bytes of a draw command from its weight planes to its end (Command::start ..).
struct __attribute__((aligned(16))) CommandTail {
	Projection plane;
	int32_t task, fan, canvas, flags;
	int32_t pad[3];
};
constexpr int32_t COMMAND_TAIL_OFFSET = 48;  // bytes of Command before start
static_assert(offsetof(Command, start) == COMMAND_TAIL_OFFSET && offsetof(Command, flags) == COMMAND_TAIL_OFFSET + 48 &&
              sizeof(CommandTail) == sizeof(Command) - COMMAND_TAIL_OFFSET, "CommandTail mirrors the tail of Command");
static_assert(sizeof(TriangleRows) % 16 == 0 && sizeof(ShaderInputs) % 16 == 0 && sizeof(CommandTail) % 16 == 0,
              "staged records are whole 16-byte units");

// The differentiable shader (drawCommandDiff) for the lanes in mask of one command's quad at
// (px, py), written straight to the float colour canvas with the visibility id base + lane.
__device__ void shadeDiffLanes(const FrameDev &f, const CommandTail &ct, const QuadEval &e, uint32_t mask, uint32_t slots,
                               int32_t idBase, const CanvasDev &cv, int32_t px, int32_t py) {
	const bool affine = (ct.flags & 1) != 0;
	const TaskDev &task = f.tasks[ct.task];
	const DiffInputs d = diffInputs(task, ct.fan);
	float m[3][3];
	for (int c = 0; c < 3; c++)
		for (int ch = 0; ch < 3; ch++) m[c][ch] = multiplier(f, d.point[c], ch);
	for (int l = 0; l < 4; l++) {
		if (!((mask >> l) & 1u)) continue;
		const int32_t sl = (int32_t)((slots >> (2 * l)) & 3u);
		const int32_t x = px + (sl & 1), y = py + (sl >> 1);
		if (x >= cv.width || y >= cv.height) continue;
		float wa, wb, wc;
		laneWeights(e, l, affine, wa, wb, wc);
		const float u = interp(d.u, wa, wb, wc), v = interp(d.v, wa, wb, wc);
		const BilinearTaps taps = bilinearTaps(f.ftexLog2w, f.ftexLog2h, u, v);
		float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
		for (int k = 0; k < 4; k++) {
			const float4 texel = *(const float4 *)(f.ftexTexels + (size_t)taps.index[k] * 4);
			t[0] += taps.weight[k] * texel.x;
			t[1] += taps.weight[k] * texel.y;
			t[2] += taps.weight[k] * texel.z;
			t[3] += taps.weight[k] * texel.w;
		}
		const size_t p = (size_t)y * cv.diffStride + x;
		*(float4 *)(cv.diffColor + p * 4) = make_float4((wa * m[0][0] + wb * m[1][0] + wc * m[2][0]) * t[0],
		                                                (wa * m[0][1] + wb * m[1][1] + wc * m[2][1]) * t[1],
		                                                (wa * m[0][2] + wb * m[1][2] + wc * m[2][2]) * t[2], t[3]);
		cv.visibility[p] = idBase + l;
	}
}

// LDS of one visibility-raster workgroup.  The differentiable shader (MODE 1) writes its colours
// straight to the float canvas and reads no shader inputs or texture table, so it keeps none of
// them here, and more of its workgroups fit a CU.
template <int TW, int TH, int CH, int MODE> struct DeferLds {
	using S = DeferShape<TW, TH, CH>;
	static constexpr bool RGBA = MODE == 0;
	static constexpr int32_t PAIRS = S::CHUNK * (TH / 2);  // (command, row pair) slots of a chunk
	// Pass-1 items are found by table up to one per pixel per chunk (by binary search beyond).
	static constexpr int32_t ITEM_TABLE = TW * TH;
	using PairIndex = typename std::conditional<(PAIRS <= 256), uint8_t, uint16_t>::type;
	int32_t list[S::LIST];                       // the round's commands, sorted
	union {
		TriangleRows tr[S::CHUNK];               // the chunk's commands: row setup (until the rows are made),
		unsigned long long keys[TW * TH];        // then the pixels' (depth, command) keys
	};
	ShaderInputs si[RGBA ? S::CHUNK : 1];        // shader inputs,
	CommandTail tail[S::CHUNK];                  // weight planes and flags
	RowSpan rows[S::CHUNK * TH];                 // chunk command j, tile row r
	union {
		struct {                                 // during the rows:
			int32_t rowFirst[S::CHUNK];          // each command's first tile row inside its aligned bound
			int32_t rowStart[S::CHUNK + 1];      // exclusive prefix of those rows per command
		};
		struct {                                 // from the pairs to pass 2:
			uint8_t pairStart[PAIRS];            // first quad column of each pair, from the tile's left
			int16_t pairOffset[PAIRS + 1];       // exclusive prefix of quads per pair
			PairIndex pairOfItem[ITEM_TABLE];    // the pair of each of the first ITEM_TABLE pass-1 items
		};
	};
	uint32_t color[RGBA ? TW * TH : 1];          // the tile's canvas pixels as they stand
	float depth[TW * TH];
	uint32_t dirty[TW * TH / 32];                // pixels written, one bit each
	uint32_t chunkWon[RGBA ? TW * TH / 32 : 1];  // pixels the current chunk's commands won
	int32_t scan[S::NT / 64];                    // wavefront totals of the prefix sums
	TextureDesc textures[RGBA ? LDS_TEXTURES : 1];
	int32_t roundInfo[4];
	int32_t slot;
};

// Copies the 16-byte units [offset, offset + sizeof(T)) of element list[c0 + j] of a global
// array with `stride` bytes per element into dst[j], for every chunk command j, with loads that
// write LDS directly (no registers hold the data).  Each wave instruction fills 64
// consecutive units of dst.  The caller waits for them (stageWait) before reading dst.
template <int NT, typename T>
__device__ __forceinline__ void stageToLds(T *dst, const uint8_t *src, int32_t stride, int32_t offset,
                                           const int32_t *list, int32_t c0, int32_t cn) {
	constexpr int32_t UNITS = (int32_t)(sizeof(T) / 16);
	const int32_t total = cn * UNITS;
	const int32_t lane = (int32_t)(threadIdx.x & 63);
	for (int32_t base = (int32_t)(threadIdx.x & ~63u); base < total; base += NT) {
		const int32_t q = base + lane;
		if (q < total) {
			const int32_t j = q / UNITS, u = q % UNITS;
			const uint8_t *g = src + (size_t)list[c0 + j] * stride + offset + u * 16;
			__builtin_amdgcn_global_load_lds((const void *)g, (void *)((uint8_t *)dst + (size_t)base * 16), 16, 0, 0);
		}
	}
}

// Waits for this thread's LDS-direct loads; a barrier then publishes them to the workgroup.
__device__ __forceinline__ void stageWait() { __builtin_amdgcn_s_waitcnt(0); }


// The chunk command that won a pixel's key, or -1.
__device__ __forceinline__ int32_t keyWinner(unsigned long long key, bool depthRead) {
	if (!depthRead) return (int32_t)key - 1;
	const uint32_t low = (uint32_t)key;
	return low == 0xFFFFFFFFu ? -1 : (int32_t)(0xFFFFFFFEu - low);
}

// Index of the segment holding item it in an exclusive prefix of count + 1 entries.
template <typename T>
__device__ __forceinline__ int32_t segmentOf(const T *offset, int32_t count, int32_t it) {
	int32_t lo = 0, hi = count - 1;
	while (lo < hi) {
		const int32_t mid = (lo + hi + 1) >> 1;
		if (offset[mid] <= it) lo = mid; else hi = mid - 1;
	}
	return lo;
}

// MODE 0: the RGBA8 shaders into the colour canvas.  MODE 1: the differentiable float shader
// into the float colour canvas plus the visibility id of each pixel's winning lane.
template <int TW, int TH, int MODE = 0, int CH>
__device__ void rasterTileDeferred(const FrameDev &f, int32_t tile, DeferLds<TW, TH, CH, MODE> &s, const TextureDesc *textures) {
	using S = DeferShape<TW, TH, CH>;
	constexpr int32_t NT = S::NT, CHUNK = S::CHUNK;
	constexpr int32_t NP = TW * TH, RP = TH / 2;
	constexpr int32_t PAIRS = DeferLds<TW, TH, CH, MODE>::PAIRS;
	constexpr int32_t ITEM_TABLE = DeferLds<TW, TH, CH, MODE>::ITEM_TABLE;
	constexpr int32_t PAIRS_PER_THREAD = PAIRS / NT > 0 ? PAIRS / NT : 1;
	static_assert(PAIRS * (TW / 2) <= 32 * NT, "a thread's items fit the 32 four-bit masks of pass 2a");
	const int32_t n = f.tileCount[tile];
	const int32_t ci = findSegment(f.canvasTileEnd, f.canvasCount, tile);
	const CanvasDev cv = f.canvases[ci];
	const int32_t local = tile - cv.tileBase;
	const int32_t tileX0 = (local % cv.tilesX) * TW;
	const int32_t tileY0 = (local / cv.tilesX) * TH;
	const int32_t t = threadIdx.x;
	const int32_t *src = tileListOf(f, tile);
	const bool colorWrite = MODE == 0 && cv.color != nullptr;
	const bool depthRead = cv.depth != nullptr;

	for (int32_t p = t; p < NP; p += NT) {
		if (f.clearInRaster) {
			s.depth[p] = f.clearDepth;
			if (MODE == 0) s.color[p] = f.clearColor;
		} else {
			const int32_t x = tileX0 + p % TW, y = tileY0 + p / TW;
			s.depth[p] = (depthRead && x < cv.width && y < cv.height) ? cv.depth[(size_t)y * cv.depthStride + x] : 0.0f;
		}
	}
	// With the fill fused in, every pixel is written back.  The differentiable shader writes its
	// colours and ids itself, so there the dirty bits keep marking the drawn pixels: the others
	// take the fill values at the write-back.
	const bool fillAll = f.clearInRaster && MODE == 0;
	const bool fillUndrawn = f.clearInRaster && MODE == 1;
	for (int32_t w = t; w < NP / 32; w += NT) s.dirty[w] = fillAll ? ~0u : 0u;

	int32_t last = -1;
	int32_t done = 0;
	while (done < n) {
		const int32_t take = nextRound<NT, S::LIST>(src, n, done, last, s.list, s.roundInfo);
		for (int32_t c0 = 0; c0 < take; c0 += CHUNK) {
			const int32_t t = opaqueThreadIndex();
			const int32_t cn = take - c0 < CHUNK ? take - c0 : CHUNK;
			// The chunk's row setups, shader inputs and weight planes into LDS.
			stageToLds<NT>(s.tr, (const uint8_t *)f.triangleRows, (int32_t)sizeof(TriangleRows), 0, s.list, c0, cn);
			if (colorWrite) stageToLds<NT>(s.si, (const uint8_t *)f.shaders, (int32_t)sizeof(ShaderInputs), 0, s.list, c0, cn);
			stageToLds<NT>(s.tail, (const uint8_t *)f.commands, (int32_t)sizeof(Command), COMMAND_TAIL_OFFSET, s.list, c0, cn);
			stageWait();
			__syncthreads();
			// Rows of every chunk command in the tile: empty outside the command's aligned bound,
			// and inside it the exact interval of rasterizeTriangle, one (command, row) per thread.
			for (int32_t idx = t; idx < cn * TH; idx += NT) s.rows[idx] = RowSpan{0, 0};
			if (t < 64) {
				int32_t first = 0, count = 0;
				if (t < cn) {
					first = tmin(tmax(s.tr[t].alT - tileY0, 0), TH);
					count = tmax(tmin(tmax(s.tr[t].alB - tileY0, 0), TH) - first, 0);
				}
				int32_t x = count;
#pragma unroll
				for (int32_t offset = 1; offset < 64; offset <<= 1) {
					const int32_t y = __shfl_up(x, offset, 64);
					if (t >= offset) x += y;
				}
				if (t < cn) {
					s.rowFirst[t] = first;
					s.rowStart[t] = x - count;
				}
				if (t == cn - 1) s.rowStart[cn] = x;
			}
			__syncthreads();
			const int32_t rowItems = s.rowStart[cn];
			for (int32_t it = t; it < rowItems; it += NT) {
				const int32_t j = segmentOf(s.rowStart, cn, it);
				const int32_t r = s.rowFirst[j] + (it - s.rowStart[j]);
				s.rows[j * TH + r] = rowSpan(triangleRow(s.tr[j], tileY0 + r));
			}
			__syncthreads();
			// Keys start at the held depth with the lowest command part above every command's,
			// so only a strictly greater depth replaces it; a NaN depth is never replaced.
			for (int32_t p = t; p < NP; p += NT) {
				const float d = s.depth[p];
				s.keys[p] = !depthRead ? 0ull : (d != d ? ~0ull : (((unsigned long long)depthOrder(d) << 32) | 0xFFFFFFFFull));
			}
			if (MODE == 0)
				for (int32_t w = t; w < NP / 32; w += NT) s.chunkWon[w] = 0u;
			// Quads of each (command, row pair) inside the tile: the outer span of the row pair
			// (fillShapeSuper), cut to the tile's columns.
			int32_t pairCount[PAIRS_PER_THREAD];
			int32_t threadSum = 0;
			for (int32_t k = 0; k < PAIRS_PER_THREAD; k++) {
				const int32_t pr = t * PAIRS_PER_THREAD + k;
				const int32_t j = pr / RP, rp = pr % RP;
				int32_t count = 0, x0 = 0;
				if (pr < PAIRS && j < cn) {
					const RowInterval upper = rowInterval(s.rows[j * TH + 2 * rp]), lower = rowInterval(s.rows[j * TH + 2 * rp + 1]);
					if (upper.right > upper.left || lower.right > lower.left) {
						const int32_t outerBlockStart = (int32_t)((uint32_t)tmin(upper.left, lower.left) & ~1u);
						const int32_t outerBlockEnd = (int32_t)(((uint32_t)tmax(upper.right, lower.right) + 1u) & ~1u);
						x0 = tmax(outerBlockStart, tileX0);
						const int32_t x1 = tmin(outerBlockEnd, tileX0 + TW);
						count = x1 > x0 ? (x1 - x0) >> 1 : 0;
					}
				}
				if (pr < PAIRS) s.pairStart[pr] = (uint8_t)(x0 - tileX0);
				pairCount[k] = count;
				threadSum += count;
			}
			int32_t before;
			const int32_t items = workgroupPrefix<NT>(threadSum, before, s.scan);
			for (int32_t k = 0; k < PAIRS_PER_THREAD; k++) {
				const int32_t pr = t * PAIRS_PER_THREAD + k;
				if (pr < PAIRS) s.pairOffset[pr] = (int16_t)before;
				for (int32_t i = before; i < before + pairCount[k] && i < ITEM_TABLE; i++)
					s.pairOfItem[i] = (typename DeferLds<TW, TH, CH, MODE>::PairIndex)pr;
				before += pairCount[k];
			}
			if (t == 0) s.pairOffset[PAIRS] = (int16_t)items;
			__syncthreads();
			// Pass 1: one (command, quad) item per thread offers the covered lanes' keys.
			for (int32_t it = t; it < items; it += NT) {
				const int32_t pr = it < ITEM_TABLE ? (int32_t)s.pairOfItem[it] : segmentOf(s.pairOffset, PAIRS, it);
				const int32_t j = pr / RP, rp = pr % RP;
				const int32_t px = tileX0 + s.pairStart[pr] + 2 * (it - s.pairOffset[pr]);
				const int32_t py = tileY0 + 2 * rp;
				QuadEval e;
				const bool ok = evalQuad(s.tail[j].plane, rowInterval(s.rows[j * TH + 2 * rp]), rowInterval(s.rows[j * TH + 2 * rp + 1]),
				                         px, py, cv.height, e);
				if (!ok) continue;
				for (int l = 0; l < 4; l++) {
					if (e.clipSides && !e.clip[l]) continue;
					const int32_t sl = e.slot[l];
					const int32_t p = (py - tileY0 + (sl >> 1)) * TW + (px - tileX0 + (sl & 1));
					if (depthRead) {
						const float d = e.l0[l];
						if (d == d) atomicMax(&s.keys[p], ((unsigned long long)depthOrder(d) << 32) | (unsigned long long)(0xFFFFFFFEu - (uint32_t)j));
					} else {
						atomicMax(&s.keys[p], (unsigned long long)(j + 1));
					}
				}
			}
			__syncthreads();
			if (MODE == 0) {
				// Pass 2a: the lanes each item's command won, read from the keys before pass 2b
				// overwrites them; four bits per item of this thread (at most 32 items).
				uint64_t wonLow = 0, wonHigh = 0;
				for (int32_t it = t, k = 0; it < items; it += NT, k++) {
					const int32_t pr = it < ITEM_TABLE ? (int32_t)s.pairOfItem[it] : segmentOf(s.pairOffset, PAIRS, it);
					const int32_t j = pr / RP, rp = pr % RP;
					const int32_t qx = s.pairStart[pr] + 2 * (it - s.pairOffset[pr]), qy = 2 * rp;
					QuadLanes q;
					uint64_t mask = 0;
					if (quadLanes(rowInterval(s.rows[j * TH + qy]), rowInterval(s.rows[j * TH + qy + 1]), tileX0 + qx, tileY0 + qy, cv.height, q)) {
						for (int l = 0; l < 4; l++) {
							if (q.clipSides && !q.clip[l]) continue;
							const int32_t sl = q.slot[l];
							if (keyWinner(s.keys[(qy + (sl >> 1)) * TW + qx + (sl & 1)], depthRead) == j) mask |= 1ull << l;
						}
					}
					if (k < 16) wonLow |= mask << (4 * k);
					else wonHigh |= mask << (4 * (k - 16));
				}
				__syncthreads();
				// Pass 2b: each item that won a lane evaluates its quad again (keeping pass 1's
				// evaluations would cost more registers than the evaluation), writes the won lanes'
				// depths, and leaves for every won pixel its weights (in the pixel's key, no longer
				// read) and its command and the quad's mip level (in its colour, about to be
				// replaced), so that pass 2c shades each pixel once with all threads busy.
				for (int32_t it = t, k = 0; it < items; it += NT, k++) {
					const uint32_t mask = (uint32_t)((k < 16 ? wonLow >> (4 * k) : wonHigh >> (4 * (k - 16))) & 0xFu);
					if (!mask) continue;
					const int32_t pr = it < ITEM_TABLE ? (int32_t)s.pairOfItem[it] : segmentOf(s.pairOffset, PAIRS, it);
					const int32_t j = pr / RP, rp = pr % RP;
					const int32_t qx = s.pairStart[pr] + 2 * (it - s.pairOffset[pr]), qy = 2 * rp;
					QuadEval e;
					if (!evalQuad(s.tail[j].plane, rowInterval(s.rows[j * TH + qy]), rowInterval(s.rows[j * TH + qy + 1]), tileX0 + qx,
					              tileY0 + qy, cv.height, e))
						continue;
					int32_t pixel[4];
					for (int l = 0; l < 4; l++) {
						const int32_t sl = e.slot[l];
						pixel[l] = (qy + (sl >> 1)) * TW + qx + (sl & 1);
						if (!((mask >> l) & 1u)) continue;
						s.depth[pixel[l]] = e.l0[l];
						if (!fillAll) atomicOr(&s.dirty[pixel[l] >> 5], 1u << (pixel[l] & 31));
					}
					if (colorWrite) {
						const bool affine = (s.tail[j].flags & 1) != 0;
						float wa[4], wb[4], wc[4];
						for (int l = 0; l < 4; l++) laneWeights(e, l, affine, wa[l], wb[l], wc[l]);
						const uint32_t mip = quadMip(s.si[j], textures, wa, wb, wc);
						for (int l = 0; l < 4; l++) {
							if (!((mask >> l) & 1u)) continue;
							s.keys[pixel[l]] = ((unsigned long long)__float_as_uint(wc[l]) << 32) | __float_as_uint(wb[l]);
							s.color[pixel[l]] = (uint32_t)j | (mip << 8);
							atomicOr(&s.chunkWon[pixel[l] >> 5], 1u << (pixel[l] & 31));
						}
					}
				}
				__syncthreads();
				// Pass 2c: every pixel won in this chunk, shaded once.
				if (colorWrite) {
					for (int32_t p = t; p < NP; p += NT) {
						if (!((s.chunkWon[p >> 5] >> (p & 31)) & 1u)) continue;
						const uint32_t meta = s.color[p];
						const unsigned long long weights = s.keys[p];
						const float wb = __uint_as_float((uint32_t)weights), wc = __uint_as_float((uint32_t)(weights >> 32));
						s.color[p] = shadePixel(s.si[meta & 0xFFu], textures, 1.0f - (wb + wc), wb, wc, meta >> 8);
					}
				}
			} else {
				// Pass 2 of the differentiable shader: the same items, evaluated again; each writes
				// the lanes whose pixels its command won, with the quad's weights as in the in-order
				// walk.  A pixel has one winner, so exactly one item writes it.
				for (int32_t it = t; it < items; it += NT) {
					const int32_t pr = it < ITEM_TABLE ? (int32_t)s.pairOfItem[it] : segmentOf(s.pairOffset, PAIRS, it);
					const int32_t j = pr / RP, rp = pr % RP;
					const int32_t qx = s.pairStart[pr] + 2 * (it - s.pairOffset[pr]), qy = 2 * rp;
					QuadEval e;
					if (!evalQuad(s.tail[j].plane, rowInterval(s.rows[j * TH + qy]), rowInterval(s.rows[j * TH + qy + 1]), tileX0 + qx,
					              tileY0 + qy, cv.height, e))
						continue;
					uint32_t mask = 0;
					for (int l = 0; l < 4; l++) {
						if (e.clipSides && !e.clip[l]) continue;
						const int32_t sl = e.slot[l];
						if (keyWinner(s.keys[(qy + (sl >> 1)) * TW + qx + (sl & 1)], depthRead) == j) mask |= 1u << l;
					}
					if (!mask) continue;
					// Depths first, so that only the weights and the won lanes' pixels outlive the quad
					// evaluation while the colours are made.
					uint32_t slots = 0;
					for (int l = 0; l < 4; l++) {
						if (!((mask >> l) & 1u)) continue;
						const int32_t sl = e.slot[l];
						const int32_t p = (qy + (sl >> 1)) * TW + qx + (sl & 1);
						s.depth[p] = e.l0[l];
						atomicOr(&s.dirty[p >> 5], 1u << (p & 31));
						slots |= (uint32_t)sl << (2 * l);
					}
					const int32_t id = (s.list[c0 + j] - f.canvasCommandBase[ci]) * 4;
					shadeDiffLanes(f, s.tail[j], e, mask, slots, id, cv, tileX0 + qx, tileY0 + qy);
				}
			}
			__syncthreads();
		}
		done += take;
		if (take == 0) break;
	}

	for (int32_t p = t; p < NP; p += NT) {
		const bool drawn = ((s.dirty[p >> 5] >> (p & 31)) & 1u) != 0;
		if (!drawn && !fillUndrawn) continue;
		const int32_t x = tileX0 + p % TW, y = tileY0 + p / TW;
		if (x >= cv.width || y >= cv.height) continue;
		if (colorWrite) cv.color[(size_t)y * cv.colorStride + x] = s.color[p];
		if (depthRead) cv.depth[(size_t)y * cv.depthStride + x] = s.depth[p];
		if (MODE == 1 && !drawn) {  // the fill values of k_clear_canvases
			const size_t q = (size_t)y * cv.diffStride + x;
			*(float4 *)(cv.diffColor + q * 4) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
			cv.visibility[q] = -1;
		}
	}
	if (f.selfClean && t == 0) f.tileCount[tile] = 0;  // read by this workgroup alone, long before
}

// Five waves per SIMD: the raster is bound by latency, and the fifth wave hides more of it than a
// larger register budget would save.  P = 1: persistent workgroups, at four waves per SIMD (a
// persistent grid holds at most eight workgroups per CU).  The two forms are separate kernels so
// that each gets its own register allocation (the persistent loop keeps more values live).
template <int TW, int TH, int P>
__global__ void __launch_bounds__(TW * TH / 4) __attribute__((amdgpu_waves_per_eu(P ? 4 : 5, P ? 4 : 5))) k_raster_deferred(FrameDev f) {
	__shared__ DeferLds<TW, TH, colorChunk<TW, TH, P>(), 0> s;
	reportCounters(f);
	resetNextCounters(f);
	if (f.counters[2]) return;  // an earlier stage ran out of room: leave the canvases untouched
	// Small texture tables are read from LDS by the shading.
	const bool ldsTextures = f.textureCount <= LDS_TEXTURES;
	if (ldsTextures) {
		constexpr int32_t W = (int32_t)(sizeof(TextureDesc) / 4);
		for (int32_t i = threadIdx.x; i < f.textureCount * W; i += TW * TH / 4)
			((uint32_t *)s.textures)[i] = ((const uint32_t *)f.textures)[i];
		__syncthreads();
	}
	const TextureDesc *textures = ldsTextures ? s.textures : f.textures;
	forEachQueuedTile(f, P, &s.slot, [&](int32_t tile) { rasterTileDeferred<TW, TH>(f, tile, s, textures); });
}

// Waves per SIMD of the differentiable visibility raster: five for 32x16 tiles of one workgroup
// each (20 B of scratch, 6 % faster on the 8-view fit), four for the others, where a fifth costs
// more in scratch than it gains.
template <int TW, int TH, int P>
constexpr int32_t diffWaves() { return P == 0 && TW * TH == 512 ? 5 : 4; }

// The differentiable forward by the visibility raster.
template <int TW, int TH, int P>
__global__ void __launch_bounds__(TW * TH / 4) __attribute__((amdgpu_waves_per_eu(diffWaves<TW, TH, P>(), diffWaves<TW, TH, P>()))) k_raster_deferred_diff(FrameDev f) {
	__shared__ DeferLds<TW, TH, DIFF_CHUNK, 1> s;
	reportCounters(f);
	if (f.counters[2]) return;
	forEachQueuedTile(f, P, &s.slot, [&](int32_t tile) { rasterTileDeferred<TW, TH, 1>(f, tile, s, nullptr); });
}

template __global__ void k_raster_deferred<16, 16, 0>(FrameDev);
template __global__ void k_raster_deferred<16, 16, 1>(FrameDev);
template __global__ void k_raster_deferred<32, 16, 0>(FrameDev);
template __global__ void k_raster_deferred<32, 16, 1>(FrameDev);
template __global__ void k_raster_deferred<32, 32, 0>(FrameDev);
template __global__ void k_raster_deferred<32, 32, 1>(FrameDev);
template __global__ void k_raster_deferred_diff<16, 16, 0>(FrameDev);
template __global__ void k_raster_deferred_diff<16, 16, 1>(FrameDev);
template __global__ void k_raster_deferred_diff<32, 16, 0>(FrameDev);
template __global__ void k_raster_deferred_diff<32, 16, 1>(FrameDev);
template __global__ void k_raster_deferred_diff<32, 32, 0>(FrameDev);
template __global__ void k_raster_deferred_diff<32, 32, 1>(FrameDev);

// The differentiable forward: the same tile walk with the float shader, persistent or not.
template <int TW, int TH>
__global__ void __launch_bounds__((TW / 2) * (TH / 2)) k_raster_diff(FrameDev f, int32_t persistent) {
	__shared__ int32_t list[TILE_SORT_CAP];
	__shared__ RowInterval rows[TH];
	__shared__ int32_t roundInfo[3];
	__shared__ int32_t slotLds;
	reportCounters(f);
	if (f.counters[2]) return;
	forEachQueuedTile(f, persistent, &slotLds, [&](int32_t tile) { rasterTile<TW, TH, 1>(f, tile, list, rows, roundInfo); });
}

template __global__ void k_raster_diff<16, 16>(FrameDev, int32_t);
template __global__ void k_raster_diff<32, 16>(FrameDev, int32_t);
template __global__ void k_raster_diff<32, 32>(FrameDev, int32_t);

// ---------------------------------------------------------------------------------------
// Depth-only tiles (model_renderDepth: executeTriangleDrawingDepth, renderCore.cpp:344-387).
//
// Each row of a draw command starts at the depth of its first pixel from the weight plane and
// adds the plane's x step once per pixel; a pixel takes the value when it is strictly greater
// than the depth held (perspective, 1/W) or strictly smaller (orthogonal, W).  The tile's
// depths live in LDS as 64-bit keys: the depth's order in the high half (reversed for
// orthogonal cameras, +0 and -0 equal), and below it the command's priority -- the held value
// first, then the tile's commands in submission order -- and the sign of a zero depth.  An
// atomic max over the keys of every covered pixel of every command leaves, per pixel, the
// first command to reach the best depth: the value the in-order walk keeps, including which of
// +0 and -0 when they tie.  NaN is never written and a held NaN is never replaced.  One thread
// per (command, row): the row is exact (rasterizeTriangle), and its depth chain runs from the
// row's first pixel, so columns left of the tile are stepped over without being written.
// ---------------------------------------------------------------------------------------
constexpr uint32_t DEPTH_HELD_PRIORITY = 0x7FFFFFFFu;

__device__ __forceinline__ uint32_t depthOnlyOrder(float d, bool affine) {
	const uint32_t o = depthOrder(d);
	return affine ? ~o : o;
}

__device__ __forceinline__ unsigned long long depthOnlyKey(float d, bool affine, uint32_t priority) {
	return ((unsigned long long)depthOnlyOrder(d, affine) << 32) | (unsigned long long)((priority << 1) | (__float_as_uint(d) >> 31));
}

// The depth a key holds.
__device__ __forceinline__ float depthOfKey(unsigned long long key, bool affine) {
	uint32_t o = (uint32_t)(key >> 32);
	if (affine) o = ~o;
	const uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
	return __uint_as_float(u == 0u ? (((uint32_t)key & 1u) << 31) : u);
}

template <int TW, int TH> struct DepthLds {
	static constexpr int32_t NT = TW * TH / 4;
	static constexpr int32_t CHUNK = 32;   // commands staged at once
	static constexpr int32_t LIST = 128;   // commands sorted per round (longer lists take more rounds)
	int32_t list[LIST];
	unsigned long long keys[TW * TH];
	TriangleRows tr[CHUNK];
	CommandTail tail[CHUNK];
	int32_t rowFirst[CHUNK];
	int32_t rowStart[CHUNK + 1];
	int32_t roundInfo[4];
	int32_t slot;
};

template <int TW, int TH>
__device__ void depthTile(const FrameDev &f, int32_t tile, DepthLds<TW, TH> &s) {
	using S = DepthLds<TW, TH>;
	constexpr int32_t NT = S::NT, NP = TW * TH, CHUNK = S::CHUNK;
	const int32_t n = f.tileCount[tile];
	const int32_t ci = findSegment(f.canvasTileEnd, f.canvasCount, tile);
	const CanvasDev cv = f.canvases[ci];
	const int32_t local = tile - cv.tileBase;
	const int32_t tileX0 = (local % cv.tilesX) * TW;
	const int32_t tileY0 = (local / cv.tilesX) * TH;
	const int32_t t = threadIdx.x;
	const int32_t *src = tileListOf(f, tile);
	// One projection kind per depth-only frame (one model, one camera).
	const bool affine = f.tasks[0].camera.perspective == 0;

	for (int32_t p = t; p < NP; p += NT) {
		const int32_t x = tileX0 + p % TW, y = tileY0 + p / TW;
		const float d = (x < cv.width && y < cv.height) ? cv.depth[(size_t)y * cv.depthStride + x] : 0.0f;
		s.keys[p] = d != d ? ~0ull : depthOnlyKey(d, affine, DEPTH_HELD_PRIORITY);
	}

	int32_t last = -1;
	int32_t done = 0;
	while (done < n) {
		const int32_t take = nextRound<NT, S::LIST>(src, n, done, last, s.list, s.roundInfo);
		for (int32_t c0 = 0; c0 < take; c0 += CHUNK) {
			const int32_t cn = take - c0 < CHUNK ? take - c0 : CHUNK;
			stageToLds<NT>(s.tr, (const uint8_t *)f.triangleRows, (int32_t)sizeof(TriangleRows), 0, s.list, c0, cn);
			stageToLds<NT>(s.tail, (const uint8_t *)f.commands, (int32_t)sizeof(Command), COMMAND_TAIL_OFFSET, s.list, c0, cn);
			stageWait();
			__syncthreads();
			// Each command's rows inside the tile and inside the canvas, and their prefix.
			if (t < 64) {
				int32_t first = 0, count = 0;
				if (t < cn) {
					const int32_t bottom = tmin(tmin(s.tr[t].alB, cv.height) - tileY0, TH);
					first = tmin(tmax(s.tr[t].alT - tileY0, 0), TH);
					count = tmax(bottom - first, 0);
				}
				int32_t x = count;
#pragma unroll
				for (int32_t offset = 1; offset < 64; offset <<= 1) {
					const int32_t y = __shfl_up(x, offset, 64);
					if (t >= offset) x += y;
				}
				if (t < cn) {
					s.rowFirst[t] = first;
					s.rowStart[t] = x - count;
				}
				if (t == cn - 1) s.rowStart[cn] = x;
			}
			__syncthreads();
			const int32_t rowItems = s.rowStart[cn];
			for (int32_t it = t; it < rowItems; it += NT) {
				const int32_t j = segmentOf(s.rowStart, cn, it);
				const int32_t y = tileY0 + s.rowFirst[j] + (it - s.rowStart[j]);
				const RowInterval row = triangleRow(s.tr[j], y);
				const int32_t x1 = tmin(row.right, tileX0 + TW);
				if (row.left >= x1 || row.right <= tileX0) continue;
				const Projection &pl = s.tail[j].plane;
				const float dx = pl.dx.x;
				float value = weightAtPixel(pl, row.left, y).x;
				int32_t x = row.left;
#pragma clang loop unroll(disable)
				for (; x < tileX0; x++) value += dx;
				const uint32_t priority = DEPTH_HELD_PRIORITY - 1u - (uint32_t)(done + c0 + j);
				unsigned long long *key = &s.keys[(y - tileY0) * TW + (x - tileX0)];
#pragma clang loop unroll(disable)
				for (; x < x1; x++) {
					if (value == value) atomicMax(key, depthOnlyKey(value, affine, priority));
					value += dx;
					key++;
				}
			}
			__syncthreads();
		}
		done += take;
		if (take == 0) break;
	}

	for (int32_t p = t; p < NP; p += NT) {
		const unsigned long long key = s.keys[p];
		if (key == ~0ull || (uint32_t)key >> 1 == DEPTH_HELD_PRIORITY) continue;
		const int32_t x = tileX0 + p % TW, y = tileY0 + p / TW;
		if (x >= cv.width || y >= cv.height) continue;
		cv.depth[(size_t)y * cv.depthStride + x] = depthOfKey(key, affine);
	}
	if (f.selfClean && t == 0) f.tileCount[tile] = 0;  // read by this workgroup alone, long before
}

template <int TW, int TH>
__global__ void __launch_bounds__(TW * TH / 4) k_raster_depth(FrameDev f) {
	__shared__ DepthLds<TW, TH> s;
	reportCounters(f);
	resetNextCounters(f);
	if (f.counters[2]) return;  // an earlier stage ran out of room: leave the canvas untouched
	forEachQueuedTile(f, 0, &s.slot, [&](int32_t tile) { depthTile<TW, TH>(f, tile, s); });
}

template __global__ void k_raster_depth<16, 16>(FrameDev);
template __global__ void k_raster_depth<32, 16>(FrameDev);
template __global__ void k_raster_depth<32, 32>(FrameDev);

// Orders one wave's LDS writes before its later reads of what other lanes of the wave wrote.
__device__ __forceinline__ void waveSync() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The adjoint of the differentiable forward with visibility fixed: each pixel re-derives the
// interpolants of the (command, lane) that drew it -- the same rows and the same weight
// stepping as the forward -- and scatters dL/dcolour into the four bilinear texels and the
// three corners' RGB multipliers with float atomics.  Workgroups pull tiles from a queue
// (all tiles of all views) until it is empty.  In tiles of 512 pixels and more a thread takes
// two horizontally adjacent pixels of one quad: drawn by the same command they share the
// quad's evaluation and one set of multiplier terms, and with the same four texels one set of
// taps.  The multiplier
// gradients of the pixels of one command are first added up in LDS (a small open-addressed
// table of the tile's commands), so a tile makes one global atomic per (command, corner,
// channel) instead of one per pixel; commands beyond the table's room add theirs directly.
// Texture gradients are gathered the same way per texel (neighbouring pixels of a magnified
// texture share their taps); taps that find no room are staged per wave in LDS and added with
// one lane per channel, so that one atomic instruction covers the four channels of 16 texels
// (L2 serves float atomics per cache line and request: four channels in one request instead
// of four requests).
constexpr int32_t BACKWARD_SLOTS = 64;
constexpr int32_t TEXEL_SLOTS = 256;
constexpr int32_t TEXEL_PROBES = 8;

// One pixel's share of the adjoint: its command (-1: none), the corners' points, its bilinear
// taps, dL/dtexel per channel before the tap weights, and dL/dmultiplier per corner and channel.
struct AdjointTerms {
	int32_t c;
	int32_t point[3];
	uint32_t texel[4];
	float weight[4];
	float gt[4];
	float mul[9];
};

// Three waves per SIMD (at most 168 VGPRs): the compiler's own choice of 189 registers left two,
// and the third wave hides more of the adjoint's gather and atomic latency than the few values it
// then keeps in scratch (and some private arrays it moves to LDS) cost.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) k_backward(FrameDev f, int32_t *head) {
	__shared__ int32_t tileLds;
	__shared__ int32_t slotCommand[BACKWARD_SLOTS];
	__shared__ int32_t slotPoint[BACKWARD_SLOTS][3];
	__shared__ float slotGrad[BACKWARD_SLOTS][9];
	__shared__ int32_t texelKey[TEXEL_SLOTS];
	__shared__ float4 texelGrad[TEXEL_SLOTS];
	__shared__ int32_t stageTexel[256 * 4];
	__shared__ float stageGrad[256 * 4 * 4];
	const int32_t t = (int32_t)threadIdx.x;

	// The taps of P (plus Q's, which then hit the same texels) into the tile's texel table, or
	// staged and added one lane per channel.  Every lane of the wave calls it.
	auto addTaps = [&](const AdjointTerms &P, const AdjointTerms &Q, bool any, bool withQ) {
		for (int k = 0; k < 4; k++) stageTexel[t * 4 + k] = -1;
		bool staged = false;
		if (any)
			for (int k = 0; k < 4; k++) {
				if (P.weight[k] == 0.0f && !(withQ && Q.weight[k] != 0.0f)) continue;
				float value[4];
				for (int ch = 0; ch < 4; ch++) value[ch] = P.gt[ch] * P.weight[k] + (withQ ? Q.gt[ch] * Q.weight[k] : 0.0f);
				const int32_t texel = (int32_t)P.texel[k];
				int32_t slot = -1;
				const int32_t h = (int32_t)(((uint32_t)texel * 2654435761u) >> 24);
				for (int32_t q = 0; q < TEXEL_PROBES; q++) {
					const int32_t i = (h + q) & (TEXEL_SLOTS - 1);
					const int32_t old = atomicCAS(&texelKey[i], -1, texel);
					if (old == -1 || old == texel) {
						slot = i;
						break;
					}
				}
				if (slot >= 0) {
					float *acc = (float *)&texelGrad[slot];
					for (int ch = 0; ch < 4; ch++) atomicAdd(acc + ch, value[ch]);
				} else {
					stageTexel[t * 4 + k] = texel;
					staged = true;
					for (int ch = 0; ch < 4; ch++) stageGrad[(t * 4 + k) * 4 + ch] = value[ch];
				}
			}
		if (__any(staged)) {
			// The wave's 256 staged taps in 16 rounds of 16 taps x 4 channels; the stage is
			// the wave's own, so a wave-scope barrier orders its writes and reads.
			waveSync();
			const int32_t lane = t & 63, ch = lane & 3, waveBase = (t & ~63) * 4;
			for (int32_t r = 0; r < 16; r++) {
				const int32_t e = waveBase + r * 16 + (lane >> 2);
				const int32_t texel = stageTexel[e];
				if (texel >= 0) unsafeAtomicAdd(f.gradTexture + (size_t)texel * 4 + ch, stageGrad[e * 4 + ch]);
			}
		}
		waveSync();
	};
	auto addMultipliers = [&](const AdjointTerms &P) {
		int32_t slot = -1;
		const int32_t h = (int32_t)(((uint32_t)P.c * 2654435761u) >> 26);
		for (int32_t k = 0; k < BACKWARD_SLOTS; k++) {
			const int32_t i = (h + k) & (BACKWARD_SLOTS - 1);
			const int32_t old = atomicCAS(&slotCommand[i], -1, P.c);
			if (old == -1) {
				for (int c3 = 0; c3 < 3; c3++) slotPoint[i][c3] = P.point[c3];
				slot = i;
				break;
			}
			if (old == P.c) {
				slot = i;
				break;
			}
		}
		for (int c3 = 0; c3 < 3; c3++)
			for (int ch = 0; ch < 3; ch++) {
				if (slot >= 0) atomicAdd(&slotGrad[slot][c3 * 3 + ch], P.mul[c3 * 3 + ch]);
				else unsafeAtomicAdd(f.gradMultipliers + P.point[c3] * 3 + ch, P.mul[c3 * 3 + ch]);
			}
	};

	while (true) {
		const int32_t t = opaqueThreadIndex();  // values derived from it are computed per tile
		if (t == 0) tileLds = atomicAdd(head, 1);
		if (t < BACKWARD_SLOTS) slotCommand[t] = -1;
		for (int32_t i = t; i < BACKWARD_SLOTS * 9; i += 256) (&slotGrad[0][0])[i] = 0.0f;
		for (int32_t i = t; i < TEXEL_SLOTS; i += 256) {
			texelKey[i] = -1;
			texelGrad[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		}
		__syncthreads();
		const int32_t tile = tileLds;
		if (tile >= f.tileTotal) break;
		const int32_t ci = findSegment(f.canvasTileEnd, f.canvasCount, tile);
		const CanvasDev cv = f.canvases[ci];
		const int32_t local = tile - cv.tileBase;
		// Pixel pairs (x even, x + 1) per thread and round, or single pixels in tiles of 256
		// pixels (every thread busy): 256 or 512 items a tile, whole rounds for every wave.
		const int32_t span = f.tileW * f.tileH >= 512 ? 2 : 1;
		const int32_t across = f.tileW / span;
		for (int32_t item = t; item < across * f.tileH; item += 256) {
			const int32_t x = (local % cv.tilesX) * f.tileW + span * (item % across);
			const int32_t y = (local / cv.tilesX) * f.tileH + item / across;
			AdjointTerms a[2];
			a[0].c = -1;
			a[1].c = -1;
			QuadEval e;
			int32_t evaluated = -1;  // the command whose quad e holds
#pragma unroll
			for (int i = 0; i < 2; i++) {
				if (i >= span || x + i >= cv.width || y >= cv.height) continue;
				const size_t p = (size_t)y * cv.diffStride + x + i;
				const int32_t id = cv.visibility[p];
				if (id < 0) continue;
				const int32_t c = (id >> 2) + f.canvasCommandBase[ci], lane = id & 3;
				const Command cmd = f.commands[c];
				if (c != evaluated) {
					evaluated = -1;
					const TriangleRows &tr = f.triangleRows[c];
					const int32_t py = y & ~1;
					RowInterval upper = triangleRow(tr, py);
					RowInterval lower = triangleRow(tr, py + 1);
					if (!evalQuad(cmd, upper, lower, x & ~1, py, cv.height, e)) continue;
					evaluated = c;
				}
				const bool affine = (cmd.flags & 1) != 0;
				float wa, wbw, wc;
				laneWeights(e, lane, affine, wa, wbw, wc);
				const DiffInputs d = diffInputs(f.tasks[cmd.task], cmd.fan);
				const float u = interp(d.u, wa, wbw, wc), v = interp(d.v, wa, wbw, wc);
				const BilinearTaps taps = bilinearTaps(f.ftexLog2w, f.ftexLog2h, u, v);
				float tx[4] = {0.0f, 0.0f, 0.0f, 0.0f};
				for (int k = 0; k < 4; k++) {
					const float4 texel = *(const float4 *)(f.ftexTexels + (size_t)taps.index[k] * 4);
					tx[0] += taps.weight[k] * texel.x;
					tx[1] += taps.weight[k] * texel.y;
					tx[2] += taps.weight[k] * texel.z;
					tx[3] += taps.weight[k] * texel.w;
				}
				const float4 g = *(const float4 *)(cv.gradColor + p * 4);
				const float w3[3] = {wa, wbw, wc};
				float mixed[3];
				for (int ch = 0; ch < 3; ch++)
					mixed[ch] = wa * multiplier(f, d.point[0], ch) + wbw * multiplier(f, d.point[1], ch) + wc * multiplier(f, d.point[2], ch);
				AdjointTerms &o = a[i];
				o.c = c;
				for (int c3 = 0; c3 < 3; c3++) o.point[c3] = d.point[c3];
				for (int k = 0; k < 4; k++) {
					o.texel[k] = taps.index[k];
					o.weight[k] = taps.weight[k];
				}
				o.gt[0] = g.x * mixed[0];
				o.gt[1] = g.y * mixed[1];
				o.gt[2] = g.z * mixed[2];
				o.gt[3] = g.w;
				const float gm[3] = {g.x * tx[0], g.y * tx[1], g.z * tx[2]};
				for (int c3 = 0; c3 < 3; c3++)
					for (int ch = 0; ch < 3; ch++) o.mul[c3 * 3 + ch] = gm[ch] * w3[c3];
			}
			const bool v0 = a[0].c >= 0, v1 = a[1].c >= 0;
			if (f.gradTexture) {
				bool same = v0 && v1;
				for (int k = 0; k < 4; k++) same = same && a[0].texel[k] == a[1].texel[k];
				// First the left pixel's taps (with the right one's when they hit the same
				// texels) or the right pixel's alone, then the right pixel's when still due.
				AdjointTerms first = a[1];
				if (v0) first = a[0];
				addTaps(first, a[1], v0 || v1, same);
				const bool second = v0 && v1 && !same;
				if (__any(second)) addTaps(a[1], a[1], second, false);
			}
			if (f.gradMultipliers) {
				if (v0 && v1 && a[0].c == a[1].c) {
					for (int m = 0; m < 9; m++) a[0].mul[m] += a[1].mul[m];
					addMultipliers(a[0]);
				} else {
					if (v0) addMultipliers(a[0]);
					if (v1) addMultipliers(a[1]);
				}
			}
		}
		__syncthreads();
		if (f.gradTexture)
			for (int32_t i = t; i < TEXEL_SLOTS * 4; i += 256) {
				const int32_t texel = texelKey[i >> 2];
				if (texel >= 0) unsafeAtomicAdd(f.gradTexture + (size_t)texel * 4 + (i & 3), (&texelGrad[i >> 2].x)[i & 3]);
			}
		if (f.gradMultipliers)
			for (int32_t i = t; i < BACKWARD_SLOTS * 9; i += 256) {
				const int32_t slot = i / 9, m = i % 9;
				if (slotCommand[slot] >= 0) unsafeAtomicAdd(f.gradMultipliers + slotPoint[slot][m / 3] * 3 + m % 3, slotGrad[slot][m]);
			}
		__syncthreads();
	}
}

// The start of a frame: zeroed counters, tile counts and fill cursors, unset canvas command
// bases, and the projection of every model point (modelAPI.cpp:237-242), in one launch.
__global__ void k_frame_begin(FrameDev f) {
	const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < COUNTERS) f.counters[i] = 0;
	if (i < 2 * f.tileTotal + 1) f.tileCount[i] = 0;
	if (i < f.canvasCount) f.canvasCommandBase[i] = 0x7FFFFFFF;
}

// renderer_occludeFromExistingTriangles (rendererAPI.cpp:241-258): the Solid draw commands of
// every task given so far lower the grid under their hulls.  One thread per fan triangle.
__global__ void __launch_bounds__(256) k_occlude_existing(FrameDev f) {
	__shared__ ClipPolygon clipPolys[256 / 64];
	const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= f.fanTotal) return;
	BinTally unused{0, 0};
	(void)setupTriangle<SETUP_OCCLUDE>(f, g, findSegment(f.taskFanEnd, f.taskCount, g), 0, unused, &clipPolys[threadIdx.x >> 6]);
}

// renderer_occludeFromTopRows (rendererAPI.cpp:403-474), one thread per cell: the first row of
// the cell's pixel rows, over the columns the reference's scan gives that cell (it starts one
// column early after the first cell and stops at the canvas width), reduced to the farthest
// depth -- 1 / the smallest inverse depth for a perspective camera, the largest depth for an
// orthogonal one -- which replaces the cell's depth if nearer.
__global__ void k_occlude_top_rows(const float *depth, int32_t width, int32_t height, int32_t stride, float *grid, int32_t gridW,
                                   int32_t gridH, int32_t perspective) {
	const int32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
	if (cell >= gridW * gridH) return;
	const int32_t gx = cell % gridW, gy = cell / gridW;
	const int32_t y = gy * OCCLUSION_CELL;
	if (y >= height) return;
	const int32_t x0 = gx == 0 ? 0 : tmin(gx * OCCLUSION_CELL - 1, width);
	const int32_t x1 = tmin(gx * OCCLUSION_CELL + OCCLUSION_CELL - 1, width);
	const float *row = depth + (size_t)y * stride;
	float farthest;
	if (perspective) {
		float smallestInverse = __builtin_huge_valf();
		for (int32_t x = x0; x < x1; x++)
			if (row[x] < smallestInverse) smallestInverse = row[x];
		farthest = 1.0f / smallestInverse;
	} else {
		farthest = 0.0f;
		for (int32_t x = x0; x < x1; x++)
			if (row[x] > farthest) farthest = row[x];
	}
	if (farthest < grid[cell]) grid[cell] = farthest;
}

// The side a row takes from one edge (edgeSide) for count (n, d, L, R) inputs: the check of
// the row intervals' quotient against exact integer arithmetic.
__global__ void k_probe_edge_side(const int64_t *n, const int64_t *d, const int32_t *L, const int32_t *R, int32_t count, int32_t *out) {
	const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < count) out[i] = edgeSide(n[i], d[i], L[i], R[i]);
}

__global__ void k_fill(int32_t *p, int32_t n, int32_t value) {
	int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = value;
}

// image_fill of every canvas (colour and depth), one thread per pixel of the frame's tiles.
__global__ void k_clear_canvases(FrameDev f, uint32_t color, float depth) {
	const int32_t tile = blockIdx.x;
	if (tile >= f.tileTotal) return;
	const int32_t ci = findSegment(f.canvasTileEnd, f.canvasCount, tile);
	const CanvasDev cv = f.canvases[ci];
	const int32_t local = tile - cv.tileBase;
	const int32_t x0 = (local % cv.tilesX) * f.tileW, y0 = (local / cv.tilesX) * f.tileH;
	for (int32_t i = threadIdx.x; i < f.tileW * f.tileH; i += blockDim.x) {
		int32_t x = x0 + i % f.tileW, y = y0 + i / f.tileW;
		if (x >= cv.width || y >= cv.height) continue;
		if (cv.color) cv.color[(size_t)y * cv.colorStride + x] = color;
		if (cv.depth) cv.depth[(size_t)y * cv.depthStride + x] = depth;
		if (cv.diffColor) {
			size_t p = (size_t)y * cv.diffStride + x;
			*(float4 *)(cv.diffColor + p * 4) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
			cv.visibility[p] = -1;
		}
	}
}

// ---------------------------------------------------------------------------------------
// texture_generatePyramid on the GPU: one level from the next larger, truncating the mean
// of each 2x2 block per channel (textureAPI.cpp:44-63).
// ---------------------------------------------------------------------------------------
__global__ void k_downsample(uint32_t *texels, TextureDesc t, uint32_t targetLevel) {
	uint32_t w = 1u << (t.log2w - targetLevel), h = 1u << (t.log2h - targetLevel);
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= w * h) return;
	uint32_t x = i % w, y = i / w;
	uint32_t src = targetLevel - 1;
	uint32_t a = texels[pixelOffset(t, x * 2, y * 2, src)];
	uint32_t b = texels[pixelOffset(t, x * 2 + 1, y * 2, src)];
	uint32_t c = texels[pixelOffset(t, x * 2, y * 2 + 1, src)];
	uint32_t d = texels[pixelOffset(t, x * 2 + 1, y * 2 + 1, src)];
	uint32_t out = 0;
	for (uint32_t ch = 0; ch < 32; ch += 8) {
		uint32_t sum = ((a >> ch) & 255u) + ((b >> ch) & 255u) + ((c >> ch) & 255u) + ((d >> ch) & 255u);
		out |= (sum / 4u) << ch;
	}
	texels[pixelOffset(t, x, y, targetLevel)] = out;
}

}  // namespace dfpsr
