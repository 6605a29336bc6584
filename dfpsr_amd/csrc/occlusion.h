// Occlusion grid of the renderer (api/rendererAPI.cpp:36-477): one cell per 16 x 16 pixels of
// the canvas, holding the largest camera-space depth at which something in the cell can still
// be seen (+inf where nothing occludes).  Occluders lower cells; whole models are skipped when
// every cell their box touches is nearer than the box, and at renderer_end each draw command
// is dropped when every cell of its bound is nearer than its nearest corner.
//
// The cell arithmetic is shared by the host side of the library (box occluders and box
// queries, a few dozen cells per call) and the kernels (the per-command test, the hulls of
// queued triangles, the top rows of the depth canvas).  Integer work is in the reference's
// int32/int64 types and float work in its order.
#pragma once

#include "geometry.h"

namespace dfpsr {

constexpr int32_t OCCLUSION_CELL = 16;  // cellSize (rendererAPI.cpp:36)

// getOuterCellBound (rendererAPI.cpp:166-177): cells under a pixel bound, clamped to the grid.
// The width or height is negative when the bound misses the grid; loops then run no times.
DH Rect outerCellBound(const Rect &pixel, int32_t gridW, int32_t gridH) {
	int32_t minX = pixel.l / OCCLUSION_CELL, maxX = pixel.right() / OCCLUSION_CELL + 1;
	int32_t minY = pixel.t / OCCLUSION_CELL, maxY = pixel.bottom() / OCCLUSION_CELL + 1;
	if (minX < 0) minX = 0;
	if (minY < 0) minY = 0;
	if (maxX > gridW) maxX = gridW;
	if (maxY > gridH) maxY = gridH;
	Rect r;
	r.l = minX; r.t = minY; r.w = maxX - minX; r.h = maxY - minY;
	return r;
}

// pointInsideOfEdge (rendererAPI.cpp:97-101) in sub-pixels.
DH bool insideOfEdge(const l2 &a, const l2 &b, int64_t px, int64_t py) {
	const int64_t nx = b.y - a.y, ny = a.x - b.x;
	return (nx * (px - a.x)) + (ny * (py - a.y)) <= 0;
}

// pointInsideOfHull (rendererAPI.cpp:105-121): inside every edge of a clockwise convex hull.
DH bool insideOfHull(const l2 *hull, int32_t n, int64_t px, int64_t py) {
	for (int32_t c = 0; c < n; c++) {
		const int32_t nc = (c + 1 == n) ? 0 : c + 1;
		if (!insideOfEdge(hull[c], hull[nc], px, py)) return false;
	}
	return true;
}

// rectangleInsideOfHull of cell (cx, cy) (rendererAPI.cpp:124-129, :227-228): its four corners
// in sub-pixels, the right and bottom ones one cell past the left and top.
DH bool cellInsideOfHull(const l2 *hull, int32_t n, int32_t cx, int32_t cy) {
	const int32_t l = cx * OCCLUSION_CELL * 256, t = cy * OCCLUSION_CELL * 256;
	const int32_t r = l + OCCLUSION_CELL * 256, b = t + OCCLUSION_CELL * 256;
	return insideOfHull(hull, n, l, t) && insideOfHull(hull, n, r, t) && insideOfHull(hull, n, l, b) && insideOfHull(hull, n, r, b);
}

// The depth an occluding hull writes (occludeFromSortedHull, rendererAPI.cpp:218-224): the
// largest camera-space depth of its corners, at least +0.
DH float hullDistance(const float *z, int32_t n) {
	float distance = 0.0f;
	for (int32_t c = 0; c < n; c++)
		if (z[c] > distance) distance = z[c];
	return distance;
}

// Whether the hull's pixel bound is large enough to occlude anything (rendererAPI.cpp:220).
DH bool hullCanOcclude(const Rect &pixelBound) { return pixelBound.w > OCCLUSION_CELL && pixelBound.h > OCCLUSION_CELL; }

// completeOcclusion's test of one draw command (rendererAPI.cpp:190-215): visible when, in any
// cell under its bound, its nearest corner is nearer than the cell's depth plus 0.001 (the sum
// and the comparison in double, as the reference's double literal makes them).
DH bool commandVisibleInGrid(const float *grid, int32_t gridW, int32_t gridH, const Rect &wholeBound, float z0, float z1, float z2) {
	float nearest = z0;
	if (z1 < nearest) nearest = z1;
	if (z2 < nearest) nearest = z2;
	const Rect cells = outerCellBound(wholeBound, gridW, gridH);
	for (int32_t cy = cells.t; cy < cells.bottom(); cy++)
		for (int32_t cx = cells.l; cx < cells.right(); cx++)
			if ((double)nearest < (double)grid[cy * gridW + cx] + 0.001) return true;
	return false;
}

// ---------------------------------------------------------------------------------------
// Host side: box occluders and box queries (a handful of corners and cells per call).
// ---------------------------------------------------------------------------------------

// The eight corners of a box in the reference's order (GENERATE_BOX_CORNERS,
// rendererAPI.cpp:259-267): x changes slowest and z fastest.  The order decides where the
// convex hull starts.
__host__ inline void boxCorners(const f3 &mn, const f3 &mx, f3 *out) {
	for (int32_t i = 0; i < 8; i++) out[i] = make3((i & 4) ? mx.x : mn.x, (i & 2) ? mx.y : mn.y, (i & 1) ? mx.z : mn.z);
}

// counterClockwise (rendererAPI.cpp:38-40) on the sub-pixel positions.
__host__ inline bool counterClockwise(const l2 &p, const l2 &q, const l2 &r) {
	return (q.y - p.y) * (r.x - q.x) - (q.x - p.x) * (r.y - q.y) < 0;
}

// jarvisConvexHullAlgorithm (rendererAPI.cpp:44-72): gift wrapping from the first corner with
// the smallest x, taking at each step the last corner that the current candidate turns
// counter-clockwise to; fewer than three corners are copied, and a walk that does not close
// within n steps stops there.
__host__ inline void convexHull(const Projected *in, int32_t n, Projected *out, int32_t &outCount) {
	if (n < 3) {
		outCount = n;
		for (int32_t p = 0; p < n; p++) out[p] = in[p];
		return;
	}
	int32_t leftmost = 0;
	for (int32_t i = 1; i < n; i++)
		if (in[i].flat.x < in[leftmost].flat.x) leftmost = i;
	outCount = 0;
	int32_t p = leftmost;
	do {
		if (outCount >= n) return;
		out[outCount++] = in[p];
		int32_t q = (p + 1) % n;
		for (int32_t i = 0; i < n; i++)
			if (counterClockwise(in[p].flat, in[i].flat, in[q].flat)) q = i;
		p = q;
	} while (p != leftmost);
}

// getPixelBoundFromProjection (rendererAPI.cpp:88-94): the union of each corner's whole pixel
// (sub-pixels divided by 256, truncating toward zero).
__host__ inline Rect pixelBoundOf(const Projected *pts, int32_t n) {
	int32_t l = (int32_t)(pts[0].flat.x / 256), t = (int32_t)(pts[0].flat.y / 256);
	int32_t r = l + 1, b = t + 1;
	for (int32_t p = 1; p < n; p++) {
		const int32_t x = (int32_t)(pts[p].flat.x / 256), y = (int32_t)(pts[p].flat.y / 256);
		l = tmin(l, x); t = tmin(t, y); r = tmax(r, x + 1); b = tmax(b, y + 1);
	}
	return rectSides(l, t, r, b);
}

// occludeFromSortedHull (rendererAPI.cpp:218-239) into a host grid of `stride` cells per row.
__host__ inline void occludeHull(float *grid, int32_t stride, int32_t gridW, int32_t gridH, const Projected *hull, int32_t n,
                                 const Rect &pixelBound) {
	if (!hullCanOcclude(pixelBound)) return;
	float z[8];
	l2 flat[8];
	for (int32_t c = 0; c < n; c++) { z[c] = hull[c].cs.z; flat[c] = hull[c].flat; }
	const float distance = hullDistance(z, n);
	const Rect cells = outerCellBound(pixelBound, gridW, gridH);
	for (int32_t cy = cells.t; cy < cells.bottom(); cy++)
		for (int32_t cx = cells.l; cx < cells.right(); cx++)
			if (cellInsideOfHull(flat, n, cx, cy) && distance < grid[cy * stride + cx]) grid[cy * stride + cx] = distance;
}

// occludeFromBox (rendererAPI.cpp:270-295) with projectHull (:76-90): every corner, with x and y
// halved in camera space, must be inside the cull frustum, or the box occludes nothing; then
// the convex hull of the projected corners lowers the grid.
__host__ inline void occludeFromBox(float *grid, int32_t stride, int32_t gridW, int32_t gridH, const f3 &mn, const f3 &mx,
                                    const Transform &modelToWorld, const CameraParams &cam) {
	f3 local[8];
	boxCorners(mn, mx, local);
	Projected projected[8], hull[8];
	for (int32_t p = 0; p < 8; p++) {
		const f3 cameraPoint = worldToCamera(cam.location, transformPoint(modelToWorld, local[p]));
		const f3 narrow = make3(cameraPoint.x * 0.5f, cameraPoint.y * 0.5f, cameraPoint.z * 1.0f);
		for (int32_t s = 0; s < cam.cull.count; s++)
			if (!insidePlane(cam.cull.planes[s], narrow)) return;
		projected[p] = cameraToScreen(cam, cameraPoint);
	}
	int32_t hullCount = 0;
	convexHull(projected, 8, hull, hullCount);
	occludeHull(grid, stride, gridW, gridH, hull, hullCount, pixelBoundOf(hull, hullCount));
}

// isBoxOccluded / isHullOccluded (rendererAPI.cpp:298-343): occluded when all corners are
// outside one plane of the cull frustum, or when the box's nearest corner is not nearer than
// any cell under the pixel bound of its eight projected corners.  The grid is read as
// image_readPixel_clamp reads the renderer's grid image (api/imageAPI.h:276-282): cell (x, y)
// at grid[y * stride + x] (the image may be wider than the frame's grid), and 0 for every cell
// while no grid image exists (grid == null).
__host__ inline bool boxOccluded(const float *grid, int32_t stride, int32_t gridW, int32_t gridH, const f3 &mn, const f3 &mx,
                                 const Transform &modelToWorld, const CameraParams &cam) {
	f3 local[8], cameraPoints[8];
	Projected projected[8];
	boxCorners(mn, mx, local);
	for (int32_t p = 0; p < 8; p++) {
		cameraPoints[p] = worldToCamera(cam.location, transformPoint(modelToWorld, local[p]));
		projected[p] = cameraToScreen(cam, cameraPoints[p]);
	}
	for (int32_t s = 0; s < cam.cull.count; s++) {
		bool allOutside = true;
		for (int32_t p = 0; p < 8; p++)
			if (insidePlane(cam.cull.planes[s], cameraPoints[p])) { allOutside = false; break; }
		if (allOutside) return true;
	}
	const Rect pixelBound = pixelBoundOf(projected, 8);
	float closest = __builtin_huge_valf();
	for (int32_t c = 0; c < 8; c++)
		if (projected[c].cs.z < closest) closest = projected[c].cs.z;
	const Rect cells = outerCellBound(pixelBound, gridW, gridH);
	for (int32_t cy = cells.t; cy < cells.bottom(); cy++)
		for (int32_t cx = cells.l; cx < cells.right(); cx++)
			if (closest < (grid ? grid[cy * stride + cx] : 0.0f)) return false;
	return true;
}

}  // namespace dfpsr
