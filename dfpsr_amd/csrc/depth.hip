// Depth-only drawing of a model: model_renderDepth (modelAPI.cpp:202-206), which the
// reference runs on the calling thread through ModelImpl::renderDepth / Part::renderDepth
// (Model.cpp:164-210) and renderTriangleFromDataDepth (renderCore.cpp:344-443).
//
// A group of 2^groupShift lanes (8 to 64, within one wavefront) per fan triangle.  Every lane
// repeats the triangle's culling, clipping and weight-plane setup (values uniform over the
// group), then the group's lanes take the rows of each front-facing (sub-)triangle in turn.
// A row keeps the reference's arithmetic: the depth at the row's first pixel from the plane,
// then one float add of the x step per pixel, so each lane walks its row in order.  The keep-nearest write (greater inverse depth for perspective, smaller depth for
// orthogonal) is a compare-and-swap loop with the reference's strict comparison; the result
// does not depend on the order of the triangles except between +0 and -0.
#include "pipeline.h"

namespace dfpsr {

__global__ void k_depth_project(const float *points, int32_t n, Transform modelToWorld, CameraParams cam, Projected *out) {
	const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const f3 p = make3(points[i * 3], points[i * 3 + 1], points[i * 3 + 2]);
	out[i] = cameraToScreen(cam, worldToCamera(cam.location, transformPoint(modelToWorld, p)));
}

// if (value > old) old = value (or < for orthogonal), as one atomic step per attempt.
__device__ __forceinline__ void keepNearest(float *cell, float value, bool greater) {
	unsigned int *word = reinterpret_cast<unsigned int *>(cell);
	unsigned int seen = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	while (true) {
		const float old = __uint_as_float(seen);
		if (!(greater ? (value > old) : (value < old))) return;
		const unsigned int prior = atomicCAS(word, seen, __float_as_uint(value));
		if (prior == seen) return;
		seen = prior;
	}
}

// executeTriangleDrawingDepth + drawTriangleDepth (renderCore.cpp:344-394) for one lane's rows.
// Rows from the aligned bound's even bottom past an odd canvas height are skipped.
__device__ void drawDepthTriangle(const Projected &A, const Projected &B, const Projected &C, bool perspective,
                                  const DepthCanvas &cv, int32_t lane, int32_t lanes) {
	if (!frontFacing(A.flat, B.flat, C.flat)) return;
	const Rect canvas = rectSides(0, 0, cv.width, cv.height);
	if (!overlaps(triangleBound(A.flat, B.flat, C.flat), canvas)) return;
	const TriangleRows tr = triangleRows(A.flat, B.flat, C.flat, canvas);
	const f2 is[3] = {A.is, B.is, C.is};
	const f3 zero = make3(0.0f, 0.0f, 0.0f);
	const Projection pr = weightPlanes(is, make3(A.cs.z, B.cs.z, C.cs.z), zero, zero, perspective);
	const int32_t bottom = tmin(tr.alB, cv.height);
	for (int32_t y = tr.alT + lane; y < bottom; y += lanes) {
		const RowInterval row = triangleRow(tr, y);
		if (row.left >= row.right) continue;
		float value = weightAtPixel(pr, row.left, y).x;
		float *cell = cv.depth + (size_t)y * cv.stride + row.left;
		for (int32_t x = row.left; x < row.right; x++) {
			keepNearest(cell, value, perspective);
			value += pr.dx.x;
			cell++;
		}
	}
}

// renderTriangleFromDataDepth (renderCore.cpp:405-443) for fan triangle (poly, k) of each lane
// group.  The grid holds fanCount << groupShift threads or more, in workgroups of 256.
__global__ void __launch_bounds__(256) k_depth_draw(const Projected *projected, const FanRecord *fans, int32_t fanCount,
                                                    int32_t groupShift, CameraParams cam, DepthCanvas cv) {
	// One clipped polygon per lane group (at most 32 groups of 8 lanes in a workgroup), in LDS
	// rather than in each lane's private scratch memory.
	__shared__ ClipPolygon groupPolys[256 / 8];
	const int32_t thread = blockIdx.x * blockDim.x + threadIdx.x;
	const int32_t g = thread >> groupShift;
	const int32_t lanes = 1 << groupShift;
	const int32_t lane = thread & (lanes - 1);
	if (g >= fanCount) return;
	const FanRecord fan = fans[g];
	const Projected A = projected[fan.corner[0]];
	const Projected B = projected[fan.corner[1]];
	const Projected C = projected[fan.corner[2]];
	const bool perspective = cam.perspective != 0;
	if (triangleVisibility(A.cs, B.cs, C.cs, cam.cull) == VIS_HIDDEN) return;
	if (triangleVisibility(A.cs, B.cs, C.cs, cam.clip) == VIS_FULL) {
		drawDepthTriangle(A, B, C, perspective, cv, lane, lanes);
		return;
	}
	// The group's first lane clips; the group's lanes (all in one wavefront, whose LDS accesses
	// complete in order) then read the polygon.
	ClipPolygon &poly9 = groupPolys[threadIdx.x >> groupShift];
	if (lane == 0) {
		poly9.init(A.cs, B.cs, C.cs);
		for (int32_t s = 0; s < cam.clip.count; s++) poly9.clip(cam.clip.planes[s]);
	}
	__builtin_amdgcn_wave_barrier();
	for (int32_t i = 0; i < poly9.count - 2; i++) {
		const Projected a = cameraToScreen(cam, poly9.v[0].cs);
		const Projected b = cameraToScreen(cam, poly9.v[1 + i].cs);
		const Projected c = cameraToScreen(cam, poly9.v[2 + i].cs);
		drawDepthTriangle(a, b, c, perspective, cv, lane, lanes);
	}
}

}  // namespace dfpsr
