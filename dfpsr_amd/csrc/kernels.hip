// gfx950 kernels of the polygon path: projection, triangle setup and clipping, ordered tile
// binning and the tile raster/shade kernel that reproduces the reference's fill loops.
//
// Reference files (relative to its Source/ folder): model projection modelAPI.cpp:214-281,
// setup and clipping renderCore.cpp:172-341, row intervals ITriangle2D.cpp:70-176, weight
// planes ITriangle2D.cpp:182-300, the 2x2 filler shader/fillerTemplates.h:92-385, shading
// shader/RgbaMultiply.h:74-175.
#include "pipeline.h"

namespace dfpsr {

// ---------------------------------------------------------------------------------------
// Projection of every model point once per task (modelAPI.cpp:237-242).
// ---------------------------------------------------------------------------------------
__global__ void k_project(FrameDev f) {
	int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= f.pointTotal) return;
	int32_t ti = findSegment(f.taskPointEnd, f.taskCount, i);
	const TaskDev &t = f.tasks[ti];
	int32_t local = i - t.pointBase;
	f3 p = make3(t.points[local * 3], t.points[local * 3 + 1], t.points[local * 3 + 2]);
	if (t.cameraSpacePoints) {
		f.projected[i] = cameraToScreen(t.camera, p);
	} else {
		Transform m = transformFrom(t.transform);
		f.projected[i] = cameraToScreen(t.camera, worldToCamera(t.camera.location, transformPoint(m, p)));
	}
}

// ---------------------------------------------------------------------------------------
// Triangle setup: renderTriangleFromData -> renderTriangleWithShader -> drawClippedTriangle
// -> drawSubTriangle (renderCore.cpp:220-341).  Returns the number of draw commands the fan
// triangle makes; with EMIT they are written to out[0..count).
// ---------------------------------------------------------------------------------------
template <bool EMIT>
__device__ int32_t setupTriangle(const FrameDev &f, int32_t g, Command *out) {
	int32_t ti = findSegment(f.taskFanEnd, f.taskCount, g);
	const TaskDev &t = f.tasks[ti];
	const CanvasDev &cv = f.canvases[t.canvas];
	if (cv.color == nullptr && cv.depth == nullptr) return 0;
	int32_t local = g - t.fanBase;
	int32_t fp = t.fanPolygon[local];
	int32_t poly = fp >> 1, k = fp & 1;
	const int32_t corner[3] = {0, 1 + k, 2 + k};
	Projected P[3];
	for (int c = 0; c < 3; c++) P[c] = f.projected[t.pointBase + t.indices[poly * 4 + corner[c]]];
	const CameraParams &cam = t.camera;
	if (triangleVisibility(P[0].cs, P[1].cs, P[2].cs, cam.cull) == VIS_HIDDEN) return 0;
	if (t.filter != 0) {
		const float *col = t.colors + poly * 16;
		if (almostZero(col[corner[0] * 4 + 3]) && almostZero(col[corner[1] * 4 + 3]) && almostZero(col[corner[2] * 4 + 3])) return 0;
	}
	int32_t count = 0;
	const bool perspective = cam.perspective != 0;
	if (triangleVisibility(P[0].cs, P[1].cs, P[2].cs, cam.clip) == VIS_FULL) {
		if (frontFacing(P[0].flat, P[1].flat, P[2].flat)) {
			if (EMIT) {
				Command &c = out[0];
				const f2 is[3] = {P[0].is, P[1].is, P[2].is};
				Projection pr = weightPlanes(is, make3(P[0].cs.z, P[1].cs.z, P[2].cs.z), make3(0.0f, 1.0f, 0.0f), make3(0.0f, 0.0f, 1.0f), perspective);
				for (int q = 0; q < 3; q++) { c.fx[q] = P[q].flat.x; c.fy[q] = P[q].flat.y; }
				c.start = pr.start; c.dx = pr.dx; c.dy = pr.dy;
				c.task = ti; c.fan = local; c.canvas = t.canvas; c.flags = perspective ? 0 : 1;
			}
			count = 1;
		}
	} else {
		ClipPolygon poly9;
		poly9.init(P[0].cs, P[1].cs, P[2].cs);
		for (int32_t s = 0; s < cam.clip.count; s++) poly9.clip(cam.clip.planes[s]);
		for (int32_t i = 0; i < poly9.count - 2; i++) {
			const SubVertex &va = poly9.v[0], &vb = poly9.v[1 + i], &vc = poly9.v[2 + i];
			Projected S[3] = {cameraToScreen(cam, va.cs), cameraToScreen(cam, vb.cs), cameraToScreen(cam, vc.cs)};
			if (frontFacing(S[0].flat, S[1].flat, S[2].flat)) {
				if (EMIT) {
					Command &c = out[count];
					const f2 is[3] = {S[0].is, S[1].is, S[2].is};
					Projection pr = weightPlanes(is, make3(S[0].cs.z, S[1].cs.z, S[2].cs.z), make3(va.subB, vb.subB, vc.subB),
					                             make3(va.subC, vb.subC, vc.subC), perspective);
					for (int q = 0; q < 3; q++) { c.fx[q] = S[q].flat.x; c.fy[q] = S[q].flat.y; }
					c.start = pr.start; c.dx = pr.dx; c.dy = pr.dy;
					c.task = ti; c.fan = local; c.canvas = t.canvas; c.flags = perspective ? 0 : 1;
				}
				count++;
			}
		}
	}
	return count;
}

// Block-wide exclusive scan of one int32 per thread in LDS; returns the block total.
__device__ int32_t blockExclusiveScan(int32_t value, int32_t &exclusive) {
	__shared__ int32_t s[SCAN_BLOCK];
	int32_t tid = threadIdx.x;
	s[tid] = value;
	__syncthreads();
	for (int32_t offset = 1; offset < SCAN_BLOCK; offset <<= 1) {
		int32_t v = (tid >= offset) ? s[tid - offset] : 0;
		__syncthreads();
		s[tid] += v;
		__syncthreads();
	}
	int32_t inclusive = s[tid];
	int32_t total = s[SCAN_BLOCK - 1];
	__syncthreads();
	exclusive = inclusive - value;
	return total;
}

__global__ void k_setup_count(FrameDev f) {
	int32_t g = blockIdx.x * SCAN_BLOCK + threadIdx.x;
	int32_t n = (g < f.fanTotal) ? setupTriangle<false>(f, g, nullptr) : 0;
	if (g < f.fanTotal) f.fanCount[g] = n;
	int32_t ex;
	int32_t total = blockExclusiveScan(n, ex);
	if (threadIdx.x == 0) f.blockSums[blockIdx.x] = total;
}

// Exclusive scan of the per-block sums by one workgroup; writes the grand total.
__global__ void k_scan_block_sums(int32_t *blockSums, int32_t blocks, int32_t *total) {
	__shared__ int32_t carry;
	if (threadIdx.x == 0) carry = 0;
	__syncthreads();
	for (int32_t base = 0; base < blocks; base += SCAN_BLOCK) {
		int32_t i = base + threadIdx.x;
		int32_t v = (i < blocks) ? blockSums[i] : 0;
		int32_t ex;
		int32_t sum = blockExclusiveScan(v, ex);
		int32_t c = carry;
		if (i < blocks) blockSums[i] = c + ex;
		__syncthreads();
		if (threadIdx.x == 0) carry = c + sum;
		__syncthreads();
	}
	if (threadIdx.x == 0) *total = carry;
}

__global__ void k_setup_emit(FrameDev f) {
	int32_t g = blockIdx.x * SCAN_BLOCK + threadIdx.x;
	int32_t n = (g < f.fanTotal) ? f.fanCount[g] : 0;
	int32_t ex;
	blockExclusiveScan(n, ex);
	if (n == 0) return;
	int32_t offset = f.blockSums[blockIdx.x] + ex;
	if (offset + n > f.commandCapacity) {
		atomicOr(&f.counters[2], 1);
		return;
	}
	setupTriangle<true>(f, g, f.commands + offset);
}

// ---------------------------------------------------------------------------------------
// Binning into screen tiles.  Lists are filled with atomics and sorted by command index in
// the raster kernel, which restores submission order per pixel.
// ---------------------------------------------------------------------------------------

struct TileRange { int32_t tx0, tx1, ty0, ty1; bool any; };

DH Rect canvasRect(const CanvasDev &cv) { return rectSides(0, 0, cv.width, cv.height); }

// Pixel region a command can touch: columns of cut(wholeBound, canvas) rounded out to quads,
// rows of the aligned raster bound (ITriangle2D.cpp:70-84).
DH TileRange commandTiles(const Command &c, const CanvasDev &cv, int32_t tw, int32_t th) {
	TileRange r;
	Rect wb = triangleBound(l2{c.fx[0], c.fy[0]}, l2{c.fx[1], c.fy[1]}, l2{c.fx[2], c.fy[2]});
	Rect full = canvasRect(cv);
	r.any = overlaps(wb, full);
	if (!r.any) { r.tx0 = r.tx1 = r.ty0 = r.ty1 = 0; return r; }
	Rect clip = cutRect(wb, full);
	Rect al = alignedRasterBound(wb, full);
	r.tx0 = clip.l / tw;
	r.tx1 = (clip.right() - 1) / tw;
	r.ty0 = al.t / th;
	r.ty1 = (al.bottom() - 2) / th;
	if (r.ty1 > cv.tilesY - 1) r.ty1 = cv.tilesY - 1;
	return r;
}

__global__ void k_bin_count(FrameDev f) {
	int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= f.counters[0] || f.counters[2]) return;
	const Command &cmd = f.commands[c];
	const CanvasDev &cv = f.canvases[cmd.canvas];
	TileRange r = commandTiles(cmd, cv, f.tileW, f.tileH);
	if (!r.any) return;
	for (int32_t ty = r.ty0; ty <= r.ty1; ty++)
		for (int32_t tx = r.tx0; tx <= r.tx1; tx++) atomicAdd(&f.tileCount[cv.tileBase + ty * cv.tilesX + tx], 1);
}

__global__ void k_scan_tiles_local(FrameDev f) {
	int32_t i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
	int32_t v = (i < f.tileTotal) ? f.tileCount[i] : 0;
	int32_t ex;
	int32_t total = blockExclusiveScan(v, ex);
	if (i < f.tileTotal) f.tileOffset[i] = ex;
	if (threadIdx.x == 0) f.blockSums[blockIdx.x] = total;
}

__global__ void k_scan_tiles_add(FrameDev f) {
	int32_t i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
	if (i < f.tileTotal) {
		f.tileOffset[i] += f.blockSums[blockIdx.x];
		if (f.tileCount[i] > 0) f.tileQueue[atomicAdd(&f.counters[3], 1)] = i;
	}
	if (i == f.tileTotal - 1) {
		int32_t end = f.tileOffset[i] + f.tileCount[i];
		f.tileOffset[f.tileTotal] = end;
		if (end > f.pairCapacity) atomicOr(&f.counters[2], 2);
	}
}

__global__ void k_bin_fill(FrameDev f) {
	int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= f.counters[0] || f.counters[2]) return;
	const Command &cmd = f.commands[c];
	const CanvasDev &cv = f.canvases[cmd.canvas];
	TileRange r = commandTiles(cmd, cv, f.tileW, f.tileH);
	if (!r.any) return;
	for (int32_t ty = r.ty0; ty <= r.ty1; ty++) {
		for (int32_t tx = r.tx0; tx <= r.tx1; tx++) {
			int32_t t = cv.tileBase + ty * cv.tilesX + tx;
			int32_t slot = atomicAdd(&f.tileFill[t], 1);
			f.tileList[f.tileOffset[t] + slot] = c;
		}
	}
}

// ---------------------------------------------------------------------------------------
// Tile raster / shade.
// ---------------------------------------------------------------------------------------

// Bitonic sort of n ints in LDS by the workgroup's NL lanes (n <= TILE_SORT_CAP).
template <int NL>
__device__ void sortLds(int32_t *a, int32_t n) {
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

// The four pixels of a lane's quad, held in registers for the whole tile.
struct QuadPixels {
	uint32_t color[4];
	float depth[4];
	uint32_t colorDirty, depthDirty;
};

__device__ __forceinline__ uint32_t getColor(const QuadPixels &q, int s) {
	return s == 0 ? q.color[0] : (s == 1 ? q.color[1] : (s == 2 ? q.color[2] : q.color[3]));
}
__device__ __forceinline__ float getDepth(const QuadPixels &q, int s) {
	return s == 0 ? q.depth[0] : (s == 1 ? q.depth[1] : (s == 2 ? q.depth[2] : q.depth[3]));
}
__device__ __forceinline__ void setColor(QuadPixels &q, int s, uint32_t v) {
	if (s == 0) q.color[0] = v; else if (s == 1) q.color[1] = v; else if (s == 2) q.color[2] = v; else q.color[3] = v;
	q.colorDirty |= 1u << s;
}
__device__ __forceinline__ void setDepth(QuadPixels &q, int s, float v) {
	if (s == 0) q.depth[0] = v; else if (s == 1) q.depth[1] = v; else if (s == 2) q.depth[2] = v; else q.depth[3] = v;
	q.depthDirty |= 1u << s;
}

// Draws one command into the lane's quad: the row-pair logic of fillShapeSuper
// (fillerTemplates.h:246-385), the incremental weight stepping of fillRowSuper (:195-244)
// and the visibility / shading / writes of fillQuadSuper (:154-187).
__device__ void drawCommand(const FrameDev &f, const Command &cmd, const TaskDev &task, const CanvasDev &cv,
                            const RowInterval *rows, int32_t rowTop, int32_t rowBottom, int32_t tileY0,
                            int32_t px, int32_t py, QuadPixels &q, bool colorWrite, bool depthRead, bool depthWrite,
                            bool alphaFilter) {
	if (py < rowTop || py >= rowBottom) return;
	RowInterval upper = rows[py - tileY0];
	RowInterval lower = rows[py + 1 - tileY0];
	int32_t outerStart = tmin(upper.left, lower.left);
	int32_t outerEnd = tmax(upper.right, lower.right);
	int32_t innerStart = tmax(upper.left, lower.left);
	int32_t innerEnd = tmin(upper.right, lower.right);
	int32_t outerBlockStart = (int32_t)((uint32_t)outerStart & ~1u);
	int32_t outerBlockEnd = (int32_t)(((uint32_t)outerEnd + 1u) & ~1u);
	int32_t innerBlockStart = (int32_t)(((uint32_t)innerStart + 1u) & ~1u);
	int32_t innerBlockEnd = (int32_t)((uint32_t)innerEnd & ~1u);
	if (py + 1 >= cv.height) lower.right = lower.left;  // the missing last row of an odd canvas
	bool hasTop = upper.right > upper.left;
	bool hasBottom = lower.right > lower.left;
	if (!(hasTop || hasBottom)) return;
	if (px < outerBlockStart || px >= outerBlockEnd) return;

	// Weights of this quad, stepped exactly as the reference's loops step them.
	const f3 dx = cmd.dx;
	const f3 dx2 = scale3(dx, 2.0f);
	f3 wU = weightAtPixel(Projection{cmd.start, cmd.dx, cmd.dy}, outerBlockStart, py);
	f3 wL = add3(wU, cmd.dy);
	bool clipSides = true;
	float l0[4], l1[4], l2v[4];
	bool splitSpan = innerBlockEnd > innerBlockStart;
	if (!splitSpan || px < innerBlockStart) {
		for (int32_t x = outerBlockStart; x < px; x += 2) { wU = add3(wU, dx2); wL = add3(wL, dx2); }
		l0[0] = wU.x; l0[1] = wU.x + dx.x; l0[2] = wL.x; l0[3] = wL.x + dx.x;
		l1[0] = wU.y; l1[1] = wU.y + dx.y; l1[2] = wL.y; l1[3] = wL.y + dx.y;
		l2v[0] = wU.z; l2v[1] = wU.z + dx.z; l2v[2] = wL.z; l2v[3] = wL.z + dx.z;
	} else {
		for (int32_t x = outerBlockStart; x < innerBlockStart; x += 2) { wU = add3(wU, dx2); wL = add3(wL, dx2); }
		if (px < innerBlockEnd) {
			clipSides = false;
			l0[0] = wU.x; l0[1] = wU.x + dx.x; l0[2] = wL.x; l0[3] = wL.x + dx.x;
			l1[0] = wU.y; l1[1] = wU.y + dx.y; l1[2] = wL.y; l1[3] = wL.y + dx.y;
			l2v[0] = wU.z; l2v[1] = wU.z + dx.z; l2v[2] = wL.z; l2v[3] = wL.z + dx.z;
			for (int32_t x = innerBlockStart; x < px; x += 2) {
				for (int l = 0; l < 4; l++) { l0[l] = l0[l] + dx2.x; l1[l] = l1[l] + dx2.y; l2v[l] = l2v[l] + dx2.z; }
			}
		} else {
			float quadCount = (float)((innerBlockEnd - innerBlockStart) / 2);
			wU = add3(wU, scale3(dx2, quadCount));
			wL = add3(wL, scale3(dx2, quadCount));
			for (int32_t x = innerBlockEnd; x < px; x += 2) { wU = add3(wU, dx2); wL = add3(wL, dx2); }
			l0[0] = wU.x; l0[1] = wU.x + dx.x; l0[2] = wL.x; l0[3] = wL.x + dx.x;
			l1[0] = wU.y; l1[1] = wU.y + dx.y; l1[2] = wL.y; l1[3] = wL.y + dx.y;
			l2v[0] = wU.z; l2v[1] = wU.z + dx.z; l2v[2] = wL.z; l2v[3] = wL.z + dx.z;
		}
	}

	// Lanes 0-1 use the upper row and 2-3 the lower row; a missing row aliases the other one.
	const int s0 = hasTop ? 0 : 2, s1 = hasTop ? 1 : 3, s2 = hasBottom ? 2 : 0, s3 = hasBottom ? 3 : 1;
	const int slot[4] = {s0, s1, s2, s3};
	bool clip[4];
	if (clipSides) {
		int32_t x2 = px + 1;
		clip[0] = px >= upper.left && px < upper.right;
		clip[1] = x2 >= upper.left && x2 < upper.right;
		clip[2] = px >= lower.left && px < lower.right;
		clip[3] = x2 >= lower.left && x2 < lower.right;
	} else {
		clip[0] = clip[1] = clip[2] = clip[3] = true;
	}
	const bool affine = (cmd.flags & 1) != 0;
	bool vis[4];
	bool anyVis = false;
	for (int l = 0; l < 4; l++) {
		bool front = true;
		if (depthRead) {
			if (clipSides && !clip[l]) {
				front = false;
			} else {
				float old = getDepth(q, slot[l]);
				front = affine ? (l0[l] < old) : (l0[l] > old);
			}
		}
		vis[l] = clip[l] && front;
		anyVis |= vis[l];
	}
	if (!anyVis) return;
	if (colorWrite) {
		float wa[4], wb[4], wc[4];
		for (int l = 0; l < 4; l++) {
			if (affine) {
				wb[l] = l1[l];
				wc[l] = l2v[l];
			} else {
				float w = refinedReciprocal(l0[l]);
				wb[l] = l1[l] * w;
				wc[l] = l2v[l] * w;
			}
			wa[l] = 1.0f - (wb[l] + wc[l]);
		}
		// Shader inputs of the source triangle (Shader.h:37-57, RgbaMultiply.h:37-58).
		int32_t fp = task.fanPolygon[cmd.fan];
		int32_t poly = fp >> 1, k = fp & 1;
		const float *col = task.colors + poly * 16;
		const float *tc = task.texcoords + poly * 16;
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
		int32_t part = task.polygonPart[poly];
		si.diffuse = task.partTexture[part * 2];
		si.light = task.partTexture[part * 2 + 1];
		finishShaderInputs(si);
		float r[4], g[4], b[4], a[4];
		shadeQuad(si, f.textures, wa, wb, wc, r, g, b, a);
		if (alphaFilter) {
			float tr[4], tg[4], tb[4], ta[4];
			for (int l = 0; l < 4; l++) {
				uint32_t old = (!clipSides || vis[l]) ? getColor(q, slot[l]) : 0u;
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
			if (vis[l]) setColor(q, slot[l], packSaturated(r[l], g[l], b[l], a[l]));
		}
	}
	if (depthWrite) {
		for (int l = 0; l < 4; l++) {
			if (vis[l]) setDepth(q, slot[l], l0[l]);
		}
	}
}

// One tile: load the lanes' pixels, draw every binned command in increasing command index
// (submission order), write back the pixels that changed.
template <int TW, int TH>
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
	const int32_t *src = f.tileList + f.tileOffset[tile];

	QuadPixels q;
	q.colorDirty = q.depthDirty = 0;
	for (int s = 0; s < 4; s++) {
		int32_t x = px + (s & 1), y = py + (s >> 1);
		bool in = x < cv.width && y < cv.height;
		q.depth[s] = (in && cv.depth) ? cv.depth[(size_t)y * cv.depthStride + x] : 0.0f;
		q.color[s] = (in && cv.color && f.alphaAny) ? cv.color[(size_t)y * cv.colorStride + x] : 0u;
	}

	// Commands in increasing index, in rounds of at most TILE_SORT_CAP: each round takes the
	// commands in (last, threshold] with the largest threshold that keeps the round in LDS.
	int32_t last = -1;
	int32_t done = 0;
	while (done < n) {
		int32_t take;
		if (done == 0 && n <= TILE_SORT_CAP) {
			for (int32_t i = lane; i < n; i += NL) list[i] = src[i];
			take = n;
		} else {
			if (lane == 0) { roundInfo[0] = last + 1; roundInfo[1] = 0x7FFFFFFE; }
			while (true) {
				__syncthreads();
				int32_t lo = roundInfo[0], hi = roundInfo[1];
				if (lo >= hi) break;
				int32_t mid = lo + ((hi - lo + 1) >> 1);
				__syncthreads();
				if (lane == 0) roundInfo[2] = 0;
				__syncthreads();
				int32_t cnt = 0;
				for (int32_t i = lane; i < n; i += NL) { int32_t v = src[i]; cnt += (v > last && v <= mid) ? 1 : 0; }
				if (cnt) atomicAdd(&roundInfo[2], cnt);
				__syncthreads();
				if (lane == 0) { if (roundInfo[2] <= TILE_SORT_CAP) roundInfo[0] = mid; else roundInfo[1] = mid - 1; }
			}
			__syncthreads();
			int32_t threshold = roundInfo[0];
			__syncthreads();
			if (lane == 0) roundInfo[2] = 0;
			__syncthreads();
			for (int32_t i = lane; i < n; i += NL) {
				int32_t v = src[i];
				if (v > last && v <= threshold) list[atomicAdd(&roundInfo[2], 1)] = v;
			}
			__syncthreads();
			take = roundInfo[2];
			last = threshold;
		}
		__syncthreads();
		sortLds<NL>(list, take);
		for (int32_t i = 0; i < take; i++) {
			const int32_t c = __builtin_amdgcn_readfirstlane(list[i]);
			const Command cmd = f.commands[c];
			const TaskDev &task = f.tasks[cmd.task];
			const bool solid = task.filter == 0;
			const bool colorWrite = cv.color != nullptr;
			const bool depthRead = cv.depth != nullptr;
			const bool depthWrite = cv.depth != nullptr && (solid || cv.color == nullptr);
			const bool alphaFilter = colorWrite && !solid;
			// Rows of this command inside the tile: rasterizeTriangle over the aligned bound.
			Rect wb = triangleBound(l2{cmd.fx[0], cmd.fy[0]}, l2{cmd.fx[1], cmd.fy[1]}, l2{cmd.fx[2], cmd.fy[2]});
			Rect clipR = cutRect(wb, canvasRect(cv));
			Rect al = alignedRasterBound(wb, canvasRect(cv));
			Rect rowClip = rectSides(clipR.l, al.t, clipR.right(), al.bottom());
			__syncthreads();
			if (lane < TH) {
				int32_t y = tileY0 + lane;
				if (y >= al.t && y < al.bottom()) {
					l2 A{cmd.fx[0], cmd.fy[0]}, B{cmd.fx[1], cmd.fy[1]}, C{cmd.fx[2], cmd.fy[2]};
					bool degenerate = sameCorner(A, B) || sameCorner(B, C) || sameCorner(C, A);
					EdgeSetup e[3] = {edgeSetup(A, B, rowClip), edgeSetup(B, C, rowClip), edgeSetup(C, A, rowClip)};
					rows[lane] = rowInterval(e, degenerate, rowClip, y - al.t);
				}
			}
			__syncthreads();
			drawCommand(f, cmd, task, cv, rows, al.t, al.bottom(), tileY0, px, py, q, colorWrite, depthRead, depthWrite, alphaFilter);
		}
		done += take;
		if (take == 0) break;
	}

	for (int s = 0; s < 4; s++) {
		int32_t x = px + (s & 1), y = py + (s >> 1);
		if (x >= cv.width || y >= cv.height) continue;
		if ((q.colorDirty >> s) & 1u) cv.color[(size_t)y * cv.colorStride + x] = q.color[s];
		if ((q.depthDirty >> s) & 1u) cv.depth[(size_t)y * cv.depthStride + x] = q.depth[s];
	}
}

// The raster kernel over the queue of non-empty tiles.  persistent = 0: workgroup b takes
// queue slot b (the grid covers every tile; surplus workgroups exit).  persistent = 1: a fixed
// grid whose workgroups pull slots from an atomic head until the queue is empty, so every
// workgroup reaches the exit once all tiles are handed out.
template <int TW, int TH>
__global__ void __launch_bounds__((TW / 2) * (TH / 2)) k_raster(FrameDev f, int32_t persistent) {
	__shared__ int32_t list[TILE_SORT_CAP];
	__shared__ RowInterval rows[TH];
	__shared__ int32_t roundInfo[3];
	__shared__ int32_t slotLds;
	if (f.counters[2]) return;  // an earlier stage ran out of room: leave the canvases untouched
	const int32_t queued = f.counters[3];
	int32_t slot = blockIdx.x;
	while (true) {
		if (persistent) {
			if (threadIdx.x == 0) slotLds = atomicAdd(f.queueHead, 1);
			__syncthreads();
			slot = slotLds;
			__syncthreads();
		}
		if (slot >= queued) break;
		rasterTile<TW, TH>(f, f.tileQueue[slot], list, rows, roundInfo);
		if (!persistent) break;
		__syncthreads();
	}
}

template __global__ void k_raster<16, 16>(FrameDev, int32_t);
template __global__ void k_raster<32, 16>(FrameDev, int32_t);
template __global__ void k_raster<32, 32>(FrameDev, int32_t);

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
