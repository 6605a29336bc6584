// CPU oracle for the polygon rendering path -- TEST INFRASTRUCTURE ONLY.
//
// A scalar restatement of the reference renderer (Dawoodoz/DFPSR) written for
// checking the GPU library.  It follows the reference's order of floating-point
// operations exactly (compile with -ffp-contract=off, no fast-math), because
// depth and coverage must match bit for bit.  The only vector instruction used
// is the SSE approximate reciprocal, which the reference's perspective filler
// refines with one Newton step (base/simd.h:4045-4068); its table is
// CPU-vendor specific, so colours pinned through it are vendor tagged.
//
// File:line citations are relative to the reference's Source/ folder.

#include "oracle.h"

#include <xmmintrin.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Scalar helpers with the reference's semantics.
// ---------------------------------------------------------------------------

struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct L2 { int64_t x, y; };

struct Rect {
	int32_t l = 0, t = 0, w = 0, h = 0;
	int32_t right() const { return l + w; }
	int32_t bottom() const { return t + h; }
};

inline Rect rectFromSides(int32_t l, int32_t t, int32_t r, int32_t b) { Rect x; x.l = l; x.t = t; x.w = r - l; x.h = b - t; return x; }

// IRect::overlaps / IRect::cut (math/IRect.h:56-66, :76-78)
inline bool overlaps(const Rect &a, const Rect &b) {
	return a.l < b.right() && a.right() > b.l && a.t < b.bottom() && a.bottom() > b.t;
}
inline Rect cut(const Rect &a, const Rect &b) {
	if (!overlaps(a, b)) return Rect();
	return rectFromSides(std::max(a.l, b.l), std::max(a.t, b.t), std::min(a.right(), b.right()), std::min(a.bottom(), b.bottom()));
}

// signedModulo / roundUp / roundDown (implementation/math/scalar.h:30-49)
inline int32_t signedModulo(int32_t a, int32_t b) { return a >= 0 ? a % b : (b - (-a % b)) % b; }
inline int32_t roundUpTo(int32_t size, int32_t alignment) { return size + (alignment - 1) - signedModulo(size - 1, alignment); }
inline int32_t roundDownTo(int32_t size, int32_t alignment) { return size - signedModulo(size, alignment); }

// _mm_cvttps_epi32 on one lane: truncation, 0x80000000 when out of range or NaN (base/simd.h:157).
inline uint32_t truncU32(float f) { return (uint32_t)_mm_cvtt_ss2si(_mm_set_ss(f)); }
// _mm_cvtepi32_ps on one lane (base/simd.h:159).
inline float fromU32(uint32_t u) { return (float)(int32_t)u; }
// _mm_min_ps(a, b): a < b ? a : b, so NaN in a gives b (base/simd.h:203, noSimd.h:183).
inline float minps(float a, float b) { return a < b ? a : b; }

// reciprocal(F32x4) on SSE: rcpps estimate + one Newton step (base/simd.h:4045-4053).
inline float reciprocalLane(float x) {
	float r = _mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(x)));
	return (r + r) - ((x * r) * r);
}

inline bool almostZero(float v) { return v > -0.001f && v < 0.001f; }  // fillerTemplates.h:37-39
inline bool almostOne(float v) { return v > 0.999f && v < 1.001f; }    // fillerTemplates.h:45-47
inline bool almostZero3(const V3 &c) { return almostZero(c.x) && almostZero(c.y) && almostZero(c.z); }
inline bool almostOne3(const V3 &c) { return almostOne(c.x) && almostOne(c.y) && almostOne(c.z); }
inline bool almostSame3(const V3 &c) { return almostZero(c.x - c.y) && almostZero(c.x - c.z) && almostZero(c.y - c.z); }

// ---------------------------------------------------------------------------
// Camera (implementation/render/Camera.h).
// ---------------------------------------------------------------------------

struct Plane { V3 n; float offset; };

// FPlane3D(normal, offset) normalizes its normal (math/FPlane3D.h:36, math/FVector.h:91-119).
Plane makePlane(V3 n, float offset) {
	float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
	Plane p;
	if (len == 0.0f) {
		p.n = V3{0.0f, 0.0f, 1.0f};
	} else {
		p.n = V3{n.x / len, n.y / len, n.z / len};
	}
	p.offset = offset;
	return p;
}
// FPlane3D::signedDistance / inside (FPlane3D.h:39-44)
inline float signedDistance(const Plane &p, const V3 &v) { return ((p.n.x * v.x) + (p.n.y * v.y)) + (p.n.z * v.z) - p.offset; }
inline bool insidePlane(const Plane &p, const V3 &v) { return signedDistance(p, v) <= 0.0f; }

struct Frustum { Plane planes[6]; int32_t count = 0; };

// ViewFrustum perspective constructor (Camera.h:64-74)
Frustum perspectiveFrustum(float nearClip, float farClip, float ws, float hs) {
	Frustum f;
	f.count = (farClip == std::numeric_limits<float>::infinity()) ? 5 : 6;
	f.planes[0] = makePlane(V3{-1.0f, 0.0f, -ws}, 0.0f);
	f.planes[1] = makePlane(V3{1.0f, 0.0f, -ws}, 0.0f);
	f.planes[2] = makePlane(V3{0.0f, 1.0f, -hs}, 0.0f);
	f.planes[3] = makePlane(V3{0.0f, -1.0f, -hs}, 0.0f);
	f.planes[4] = makePlane(V3{0.0f, 0.0f, -1.0f}, -nearClip);
	f.planes[5] = makePlane(V3{0.0f, 0.0f, 1.0f}, farClip);
	return f;
}
// ViewFrustum orthogonal constructor (Camera.h:56-62)
Frustum orthogonalFrustum(float halfWidth, float halfHeight) {
	Frustum f;
	f.count = 4;
	f.planes[0] = makePlane(V3{-1.0f, 0.0f, 0.0f}, halfWidth);
	f.planes[1] = makePlane(V3{1.0f, 0.0f, 0.0f}, halfWidth);
	f.planes[2] = makePlane(V3{0.0f, 1.0f, 0.0f}, halfHeight);
	f.planes[3] = makePlane(V3{0.0f, -1.0f, 0.0f}, halfHeight);
	return f;
}

struct Transform { V3 pos, xAxis, yAxis, zAxis; };

Transform transformFrom(const float *t) {
	return Transform{V3{t[0], t[1], t[2]}, V3{t[3], t[4], t[5]}, V3{t[6], t[7], t[8]}, V3{t[9], t[10], t[11]}};
}
// Transform3D::transformPoint = FMatrix3x3::transform(p) + position (Transform3D.h:41-43, FMatrix3x3.h:55-61)
inline V3 transformPoint(const Transform &m, const V3 &p) {
	V3 r;
	r.x = ((p.x * m.xAxis.x + p.y * m.yAxis.x) + p.z * m.zAxis.x) + m.pos.x;
	r.y = ((p.x * m.xAxis.y + p.y * m.yAxis.y) + p.z * m.zAxis.y) + m.pos.y;
	r.z = ((p.x * m.xAxis.z + p.y * m.yAxis.z) + p.z * m.zAxis.z) + m.pos.z;
	return r;
}
// transformPointTransposedInverse (Transform3D.h:51-53, FMatrix3x3.h:66-72)
inline V3 transformPointTransposedInverse(const Transform &m, const V3 &p) {
	V3 d{p.x - m.pos.x, p.y - m.pos.y, p.z - m.pos.z};
	V3 r;
	r.x = (d.x * m.xAxis.x + d.y * m.xAxis.y) + d.z * m.xAxis.z;
	r.y = (d.x * m.yAxis.x + d.y * m.yAxis.y) + d.z * m.yAxis.z;
	r.z = (d.x * m.zAxis.x + d.y * m.zAxis.y) + d.z * m.zAxis.z;
	return r;
}

struct Camera {
	bool perspective;
	Transform location;
	float widthSlope, heightSlope, invWidthSlope, invHeightSlope, imageWidth, imageHeight, nearClip, farClip;
	Frustum cull, clip;
};

// Camera::createPerspective / createOrthogonal and the Camera constructor (Camera.h:139-156)
Camera makeCamera(const scene_desc *s) {
	Camera c;
	c.location = transformFrom(s->cam_location);
	c.imageWidth = float(s->width);
	c.imageHeight = float(s->height);
	const float cullRatio = 1.0001f, clipRatio = 2.0f;  // Camera.h:113-118
	if (s->perspective) {
		float ws = s->width_slope;
		float hs = ws * c.imageHeight / c.imageWidth;
		c.perspective = true;
		c.widthSlope = ws; c.heightSlope = hs;
		c.nearClip = s->near_clip; c.farClip = s->far_clip;
		c.cull = perspectiveFrustum(s->near_clip, s->far_clip, ws * cullRatio, hs * cullRatio);
		c.clip = perspectiveFrustum(s->near_clip, s->far_clip, ws * clipRatio, hs * clipRatio);
	} else {
		float hw = s->width_slope;
		float hh = hw * c.imageHeight / c.imageWidth;
		c.perspective = false;
		c.widthSlope = hw; c.heightSlope = hh;
		c.nearClip = -std::numeric_limits<float>::max();
		c.farClip = std::numeric_limits<float>::infinity();
		c.cull = orthogonalFrustum(hw * cullRatio, hh * cullRatio);
		c.clip = orthogonalFrustum(hw * clipRatio, hh * clipRatio);
	}
	c.invWidthSlope = 0.5f / c.widthSlope;
	c.invHeightSlope = 0.5f / c.heightSlope;
	return c;
}

struct Projected { V3 cs; V2 is; L2 flat; };

// float -> int64 truncation as cvttss2si with a 64-bit destination.
inline int64_t truncI64(float f) {
	if (!(f > -9223372036854775808.0f && f < 9223372036854775808.0f)) return INT64_MIN;
	return (int64_t)f;
}

// Camera::cameraToScreen (Camera.h:160-187)
Projected cameraToScreen(const Camera &c, const V3 &cs) {
	Projected p;
	p.cs = cs;
	if (c.perspective) {
		float invDepth = (cs.z > 0.0f) ? (1.0f / cs.z) : 0.0f;
		float centerShear = cs.z * 0.5f;
		float px = (cs.x * c.invWidthSlope + centerShear) * c.imageWidth;
		float py = (-cs.y * c.invHeightSlope + centerShear) * c.imageHeight;
		p.is.x = px * invDepth;
		p.is.y = py * invDepth;
	} else {
		p.is.x = (cs.x * c.invWidthSlope + 0.5f) * c.imageWidth;
		p.is.y = (-cs.y * c.invHeightSlope + 0.5f) * c.imageHeight;
	}
	p.flat.x = truncI64(p.is.x * 256.0f);
	p.flat.y = truncI64(p.is.y * 256.0f);
	return p;
}

// ViewFrustum::isConvexHullSeen (Camera.h:80-100)
int32_t convexHullSeen(const Frustum &f, const V3 *pts, int32_t n) {
	bool anyOutside = false;
	for (int32_t s = 0; s < f.count; s++) {
		bool anyInside = false;
		for (int32_t p = 0; p < n; p++) {
			if (insidePlane(f.planes[s], pts[p])) anyInside = true; else anyOutside = true;
		}
		if (!anyInside) return 0;
	}
	return anyOutside ? 1 : 2;
}

// Camera::isBoxSeen (Camera.h:202-217)
int32_t boxSeen(const Camera &c, const V3 &mn, const V3 &mx, const Transform &modelToWorld) {
	V3 corners[8] = {
		{mn.x, mn.y, mn.z}, {mx.x, mn.y, mn.z}, {mn.x, mx.y, mn.z}, {mx.x, mx.y, mn.z},
		{mn.x, mn.y, mx.z}, {mx.x, mn.y, mx.z}, {mn.x, mx.y, mx.z}, {mx.x, mx.y, mx.z}};
	V3 cs[8];
	for (int i = 0; i < 8; i++) cs[i] = transformPointTransposedInverse(c.location, transformPoint(modelToWorld, corners[i]));
	return convexHullSeen(c.cull, cs, 8);
}

// ---------------------------------------------------------------------------
// Triangle setup (implementation/render/ITriangle2D.cpp, renderCore.cpp).
// ---------------------------------------------------------------------------

struct Tri {
	Projected p[3];
	Rect wholeBound;
};

// getTriangleBound (ITriangle2D.cpp:31-43): round to the nearest pixel, then widen by one.
Rect triangleBound(const L2 &a, const L2 &b, const L2 &c) {
	int32_t x1 = (int32_t)((a.x + 128) / 256), y1 = (int32_t)((a.y + 128) / 256);
	int32_t x2 = (int32_t)((b.x + 128) / 256), y2 = (int32_t)((b.y + 128) / 256);
	int32_t x3 = (int32_t)((c.x + 128) / 256), y3 = (int32_t)((c.y + 128) / 256);
	int32_t l = std::min(std::min(x1, x2), x3) - 1;
	int32_t t = std::min(std::min(y1, y2), y3) - 1;
	int32_t r = std::max(std::max(x1, x2), x3) + 1;
	int32_t b2 = std::max(std::max(y1, y2), y3) + 1;
	return rectFromSides(l, t, r, b2);
}

Tri makeTri(const Projected &a, const Projected &b, const Projected &c) {
	Tri t;
	t.p[0] = a; t.p[1] = b; t.p[2] = c;
	t.wholeBound = triangleBound(a.flat, b.flat, c.flat);
	return t;
}

// ITriangle2D::isFrontfacing (ITriangle2D.cpp:55-60): clockwise on screen.
bool frontFacing(const Tri &t) {
	const L2 &a = t.p[0].flat, &b = t.p[1].flat, &c = t.p[2].flat;
	return ((c.x - a.x) * (b.y - a.y)) + ((c.y - a.y) * (a.x - b.x)) < 0;
}

enum Visibility { Hidden, Full, Partial };

// getTriangleVisibility (renderCore.cpp:172-198)
Visibility triangleVisibility(const Tri &t, const Frustum &f) {
	bool outside[3][6];
	for (int c = 0; c < 3; c++)
		for (int s = 0; s < f.count; s++)
			outside[c][s] = !insidePlane(f.planes[s], t.p[c].cs);
	for (int s = 0; s < f.count; s++)
		if (outside[0][s] && outside[1][s] && outside[2][s]) return Hidden;
	for (int c = 0; c < 3; c++)
		for (int s = 0; s < f.count; s++)
			if (outside[c][s]) return Partial;
	return Full;
}

// SubVertex and its interpolating constructor (renderCore.cpp:33-48)
struct SubVertex {
	V3 cs{0.0f, 0.0f, 0.0f};
	float subB = 0.0f, subC = 0.0f;
	int32_t state = 0;
	float value = 0.0f;
};
SubVertex lerpVertex(const SubVertex &a, const SubVertex &b, float ratio) {
	SubVertex r;
	float inv = 1.0f - ratio;
	r.cs.x = a.cs.x * inv + b.cs.x * ratio;
	r.cs.y = a.cs.y * inv + b.cs.y * ratio;
	r.cs.z = a.cs.z * inv + b.cs.z * ratio;
	r.subB = a.subB * inv + b.subB * ratio;
	r.subC = a.subC * inv + b.subC * ratio;
	return r;
}
// inverseLerp (renderCore.cpp:52-59)
inline float inverseLerp(float a, float b, float value) {
	float c = b - a;
	return (c == 0.0f) ? 0.5f : (value - a) / c;
}

// ClippedTriangle (renderCore.cpp:62-170): polygon of at most 9 sub-vertices,
// cut plane by plane.  The vertex order it produces decides the fan, so the
// same insert/replace/delete steps are taken.
struct Polygon9 {
	static const int32_t maxPoints = 9;
	SubVertex v[maxPoints];
	int32_t count = 0;
	void remove(int32_t i) {
		if (count > 0 && count <= maxPoints) {
			for (int32_t k = i; k < count - 1; k++) v[k] = v[k + 1];
			count--;
		}
	}
	void insert(int32_t i, const SubVertex &x) {
		if (count < maxPoints) {
			for (int32_t k = count - 1; k >= i; k--) v[k + 1] = v[k];
			v[i] = x;
			count++;
		}
	}
	void clip(const Plane &plane) {
		const int32_t use = 0, del = 1, modified = 2;
		if (!(count >= 3 && count < maxPoints)) return;
		int32_t outsideCount = 0, lastOutside = 0;
		for (int32_t k = 0; k < count; k++) {
			float d = signedDistance(plane, v[k].cs);
			v[k].value = d;
			if (d > 0.0f) { outsideCount++; lastOutside = k; v[k].state = del; } else { v[k].state = use; }
		}
		if (outsideCount == 0) return;
		if (outsideCount >= count) {
			count = 0;
		} else if (outsideCount == 1) {
			int32_t cur = lastOutside;
			int32_t prev = (lastOutside - 1 + count) % count;
			int32_t next = (lastOutside + 1) % count;
			float r1 = inverseLerp(v[prev].value, v[cur].value, 0.0f);
			float r2 = inverseLerp(v[cur].value, v[next].value, 0.0f);
			SubVertex cutStart = lerpVertex(v[prev], v[cur], r1);
			SubVertex cutEnd = lerpVertex(v[cur], v[next], r2);
			v[lastOutside] = cutStart;
			insert(next, cutEnd);
		} else {
			for (int32_t cur = 0; cur < count; cur++) {
				int32_t prev = (cur - 1 + count) % count;
				int32_t next = (cur + 1) % count;
				if (v[cur].state == del) {
					if (v[prev].state == use) {
						float r = inverseLerp(v[prev].value, v[cur].value, 0.0f);
						v[cur] = lerpVertex(v[prev], v[cur], r);
						v[cur].state = modified;
					} else if (v[next].state == use) {
						float r = inverseLerp(v[cur].value, v[next].value, 0.0f);
						v[cur] = lerpVertex(v[cur], v[next], r);
						v[cur].state = modified;
					}
				}
			}
			if (outsideCount > 2) {
				for (int32_t k = count - 1; k >= 0; k--) {
					if (v[k].state == del) remove(k);
				}
			}
		}
	}
};

// Per-triangle shading inputs: TriangleTexCoords / TriangleColors (shader/Shader.h:37-57)
struct TriangleInput {
	V3 u1, v1, u2, v2;
	V3 red, green, blue, alpha;
	int32_t diffuse, light;  // texture index or -1
	int32_t point[3];        // model points of the corners (differentiable shader's multipliers)
};

struct Command {
	Tri tri;
	V3 subB, subC;
	int32_t input;        // index into the TriangleInput list
	int32_t sourceTriangle;
};

// ---------------------------------------------------------------------------
// Textures (implementation/image/Texture.h, api/textureAPI.{h,cpp}).
// ---------------------------------------------------------------------------

struct Texture {
	std::vector<uint32_t> buffer;  // all levels, smallest first
	uint32_t log2w = 0, log2h = 0, maxMip = 0;
	uint32_t startOffset = 0, maxLevelMask = 0;
	uint32_t minWidthOrMask = 0, minHeightOrMask = 0, maxWidthAndMask = 0, maxHeightAndMask = 0;
	float floatMaxWidth = 0.0f, floatMaxHeight = 0.0f;
};

// findLog2Size (textureAPI.cpp:31-41)
uint32_t log2Ceil(uint32_t size) {
	for (uint32_t l = 0; l < 15; l++) if ((1u << l) >= size) return l;
	return 15;
}

inline uint32_t pixelOffsetToLayer(const Texture &t, uint32_t mip) {  // textureAPI.h:79-85
	return t.startOffset & (t.maxLevelMask >> (mip << 1));
}
inline uint32_t levelWidth(const Texture &t, uint32_t mip) { return 1u << (t.log2w - mip); }
inline uint32_t levelHeight(const Texture &t, uint32_t mip) { return 1u << (t.log2h - mip); }

// texture_getPixelOffset without the optimization flags (textureAPI.h:104-170).  For a mip
// level past log2w the stride log2w - mip wraps around as a uint32 and the reference's scalar
// shift takes it modulo 32 (x86 shl); the shift here says so, which keeps it defined.
inline uint32_t pixelOffset(const Texture &t, uint32_t x, uint32_t y, uint32_t mip) {
	uint32_t mx = (t.maxWidthAndMask >> mip) | t.minWidthOrMask;
	uint32_t my = (t.maxHeightAndMask >> mip) | t.minHeightOrMask;
	return pixelOffsetToLayer(t, mip) + (((y & my) << ((t.log2w - mip) & 31u)) | (x & mx));
}

// texture_create_RgbaU8 + the Texture constructor (textureAPI.cpp:65-78, Texture.h:66-101),
// level 0 filled from `level0`, then texture_generatePyramid's box filter (textureAPI.cpp:44-87).
// The TextureRgbaU8 constructor's layout (Texture.h:66-101): levels stored smallest first.
Texture layoutTexture(uint32_t log2w, uint32_t log2h, uint32_t requestedMaxMip) {
	Texture t;
	t.log2w = log2w;
	t.log2h = log2h;
	if (requestedMaxMip > 15u) requestedMaxMip = 15u;
	t.maxMip = std::min(std::min(t.log2w, t.log2h), requestedMaxMip);
	uint32_t highest = 1u << (t.log2w + t.log2h);
	uint64_t pixelCount = 0;
	uint32_t levelCount = highest;
	for (int32_t level = (int32_t)t.maxMip; level >= 0; level--) {
		pixelCount |= levelCount;
		levelCount >>= 2;
	}
	t.startOffset = (uint32_t)pixelCount & ~highest;
	t.maxLevelMask = highest - 1;
	t.minWidthOrMask = (1u << (t.log2w - t.maxMip)) - 1;
	t.minHeightOrMask = (1u << (t.log2h - t.maxMip)) - 1;
	t.maxWidthAndMask = (1u << t.log2w) - 1;
	t.maxHeightAndMask = (1u << t.log2h) - 1;
	t.floatMaxWidth = float(1u << t.log2w);
	t.floatMaxHeight = float(1u << t.log2h);
	t.buffer.assign((size_t)pixelCount, 0u);
	return t;
}

// box filter of texture_generatePyramid / downsample (textureAPI.cpp:44-87)
void generatePyramid(Texture &t) {
	for (uint32_t level = 1; level <= t.maxMip; level++) {
		uint32_t lw = levelWidth(t, level), lh = levelHeight(t, level);
		for (uint32_t y = 0; y < lh; y++) {
			for (uint32_t x = 0; x < lw; x++) {
				uint32_t a = t.buffer[pixelOffset(t, x * 2, y * 2, level - 1)];
				uint32_t b = t.buffer[pixelOffset(t, x * 2 + 1, y * 2, level - 1)];
				uint32_t c = t.buffer[pixelOffset(t, x * 2, y * 2 + 1, level - 1)];
				uint32_t d = t.buffer[pixelOffset(t, x * 2 + 1, y * 2 + 1, level - 1)];
				uint32_t out = 0;
				for (uint32_t ch = 0; ch < 32; ch += 8) {
					uint32_t sum = ((a >> ch) & 255u) + ((b >> ch) & 255u) + ((c >> ch) & 255u) + ((d >> ch) & 255u);
					out |= (sum / 4u) << ch;
				}
				t.buffer[pixelOffset(t, x, y, level)] = out;
			}
		}
	}
}

// texture_create_RgbaU8 (textureAPI.cpp:65-78), level 0 filled from `level0`, then the pyramid.
Texture makeTexture(int32_t w, int32_t h, int32_t resolutions, const uint32_t *level0) {
	Texture t = layoutTexture(log2Ceil((uint32_t)w), log2Ceil((uint32_t)h), (uint32_t)(resolutions - 1));
	uint32_t tw = 1u << t.log2w, th = 1u << t.log2h;
	for (uint32_t y = 0; y < th; y++)
		for (uint32_t x = 0; x < tw; x++)
			t.buffer[t.startOffset + y * tw + x] = level0[(y % (uint32_t)h) * (uint32_t)w + (x % (uint32_t)w)];
	generatePyramid(t);
	return t;
}

// texture_sample_nearest (textureAPI.h:232-244); not used by the shaders, kept to pin the
// wrap offset and truncation that the bilinear sampler shares.
uint32_t sampleNearest(const Texture &t, float u, float v, uint32_t mip) {
	uint32_t scaleU = (1u << t.log2w) >> mip, scaleV = (1u << t.log2h) >> mip;
	uint32_t x = truncU32((u + 256.0f) * fromU32(scaleU));
	uint32_t y = truncU32((v + 256.0f) * fromU32(scaleV));
	return t.buffer[pixelOffset(t, x, y, mip)];
}

// texture_getMipLevelIndex (textureAPI.h:472-495) from quad lanes 0, 1 and 2.
uint32_t mipLevelIndex(const Texture &t, const float *u, const float *v) {
	float offsetUX = fabsf(u[0] - u[1]);
	float offsetUY = fabsf(u[0] - u[2]);
	float offsetVX = fabsf(v[0] - v[1]);
	float offsetVY = fabsf(v[0] - v[2]);
	float offsetU = (offsetUX > offsetUY ? offsetUX : offsetUY) * t.floatMaxWidth;
	float offsetV = (offsetVX > offsetVY ? offsetVX : offsetVY) * t.floatMaxHeight;
	float offset = offsetU > offsetV ? offsetU : offsetV;
	int32_t result = 0;
	if (offset > 2.0f) result = 1;
	if (offset > 4.0f) result = 2;
	if (offset > 8.0f) result = 3;
	if (offset > 16.0f) result = 4;
	if (result > (int32_t)t.maxMip) result = (int32_t)t.maxMip;
	return (uint32_t)result;
}

// One channel pair of weightColors: (a * (256 - w) + b * w) >> 8 per byte (textureAPI.h:248-300).
inline uint32_t mixColors(uint32_t a, uint32_t b, uint32_t w) {
	uint32_t out = 0;
	for (uint32_t ch = 0; ch < 32; ch += 8) {
		uint32_t ca = (a >> ch) & 255u, cb = (b >> ch) & 255u;
		out |= (((ca * (256u - w) + cb * w) >> 8) & 255u) << ch;
	}
	return out;
}

// texture_sample_bilinear for MIP_INSIDE (textureAPI.h:342-438).  highest = HIGHEST_RESOLUTION.
uint32_t sampleBilinear(const Texture &t, float u, float v, uint32_t mip, bool highest) {
	uint32_t scaleU = 256u << t.log2w, scaleV = 256u << t.log2h;
	uint32_t maskX = t.maxWidthAndMask, maskY = t.maxHeightAndMask;
	uint32_t stride = t.log2w;
	uint32_t layer;
	if (!highest) {
		scaleU >>= mip; scaleV >>= mip;
		maskX >>= mip; maskY >>= mip;
		stride -= mip;
		layer = pixelOffsetToLayer(t, mip);
	} else {
		layer = t.startOffset;
	}
	uint32_t subX = truncU32((u + 256.0f) * fromU32(scaleU)) - 128u;
	uint32_t subY = truncU32((v + 256.0f) * fromU32(scaleV)) - 128u;
	uint32_t wx = subX & 255u, wy = subY & 255u;
	uint32_t left = subX >> 8, top = subY >> 8;
	uint32_t right = (left + 1u) & maskX, bottom = (top + 1u) & maskY;
	left &= maskX; top &= maskY;
	uint32_t ul = t.buffer[((top << stride) | left) + layer];
	uint32_t ur = t.buffer[((top << stride) | right) + layer];
	uint32_t bl = t.buffer[((bottom << stride) | left) + layer];
	uint32_t br = t.buffer[((bottom << stride) | right) + layer];
	return mixColors(mixColors(ul, ur, wx), mixColors(bl, br, wx), wy);
}

// ---------------------------------------------------------------------------
// Rasterization (ITriangle2D.cpp:70-176).
// ---------------------------------------------------------------------------

struct Row { int32_t left, right; };

// cutConvexEdge (ITriangle2D.cpp:86-148)
void cutEdge(const L2 &s, const L2 &e, Row *rows, const Rect &clip) {
	const int32_t L = clip.l, T = clip.t, R = clip.right(), B = clip.bottom();
	int64_t originX = 128 + (int64_t)L * 256;
	int64_t originY = 128 + (int64_t)T * 256;
	int64_t threshold = (s.x > e.x || (s.x == e.x && s.y > e.y)) ? -1 : 0;
	int64_t normalX = e.y - s.y;
	int64_t normalY = s.x - e.x;
	int64_t offsetX = normalX * 256;
	int64_t offsetY = normalY * 256;
	int64_t valueOrigin = ((originX - s.x) * normalX) + ((originY - s.y) * normalY);
	if (normalX != 0) {
		int64_t limit = threshold - valueOrigin + offsetX * L;
		if (normalX < 0) {
			for (int32_t y = T; y < B; y++) {
				int32_t side = std::min(std::max(L, (int32_t)((limit + 1) / offsetX + 1)), R);
				Row &r = rows[y - T];
				r.left = std::max(r.left, side);
				limit -= offsetY;
			}
		} else {
			for (int32_t y = T; y < B; y++) {
				int32_t side = std::min(std::max(L, (int32_t)(limit / offsetX + 1)), R);
				Row &r = rows[y - T];
				r.right = std::min(r.right, side);
				limit -= offsetY;
			}
		}
	} else if (normalY != 0) {
		int64_t value = valueOrigin;
		for (int32_t y = T; y < B; y++) {
			if (value > threshold) { rows[y - T].left = R; rows[y - T].right = L; }
			value += offsetY;
		}
	}
}

// rasterizeTriangle (ITriangle2D.cpp:150-169)
void rasterize(const L2 &a, const L2 &b, const L2 &c, Row *rows, const Rect &clip) {
	bool degenerate = (a.x == b.x && a.y == b.y) || (b.x == c.x && b.y == c.y) || (c.x == a.x && c.y == a.y);
	for (int32_t r = 0; r < clip.h; r++) {
		rows[r].left = degenerate ? clip.right() : clip.l;
		rows[r].right = degenerate ? clip.l : clip.right();
	}
	if (!degenerate) {
		const L2 p[3] = {a, b, c};
		for (int i = 0; i < 3; i++) cutEdge(p[i], p[(i + 1) % 3], rows, clip);
	}
}

// getAlignedRasterBound (ITriangle2D.cpp:70-75): rows rounded out to even numbers.
Rect alignedRasterBound(const Tri &t, const Rect &clip) {
	Rect u = cut(t.wholeBound, clip);
	int32_t top = roundDownTo(u.t, 2);
	int32_t bottom = roundUpTo(u.bottom(), 2);
	return rectFromSides(u.l, top, u.right(), bottom);
}

// ---------------------------------------------------------------------------
// Weight planes (ITriangle2D.cpp:182-300).
// ---------------------------------------------------------------------------

struct Projection { bool affine; V3 start, dx, dy; };

inline float &at(V3 &v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
inline float get(const V3 &v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

Projection getProjection(const Tri &t, const V3 &subB, const V3 &subC, bool perspective) {
	V3 offX, offY;
	for (int i = 0; i < 3; i++) {
		int j = (i + 1) % 3;
		at(offX, i) = t.p[j].is.y - t.p[i].is.y;
		at(offY, i) = t.p[i].is.x - t.p[j].is.x;
	}
	V3 mult;
	for (int i = 0; i < 3; i++) {
		int o = (i + 2) % 3;
		float other = ((t.p[o].is.x - t.p[i].is.x) * get(offX, i)) + ((t.p[o].is.y - t.p[i].is.y) * get(offY, i));
		at(mult, o) = (other == 0.0f) ? 0.0f : 1.0f / other;
	}
	V3 nX, nY;
	for (int i = 0; i < 3; i++) {
		at(nX, i) = get(offX, i) * get(mult, i);
		at(nY, i) = get(offY, i) * get(mult, i);
	}
	V3 tw;
	for (int i = 0; i < 3; i++) {
		int o = (i + 2) % 3;
		at(tw, o) = t.p[i].is.x * -get(nX, i) + t.p[i].is.y * -get(nY, i);
	}
	V3 adx{nX.y, nX.z, nX.x};
	V3 ady{nY.y, nY.z, nY.x};
	Projection p;
	if (!perspective) {
		V3 W{t.p[0].cs.z, t.p[1].cs.z, t.p[2].cs.z};
		p.affine = true;
		p.start.x = W.x * tw.x + W.y * tw.y + W.z * tw.z;
		p.start.y = tw.x * subB.x + tw.y * subB.y + tw.z * subB.z;
		p.start.z = tw.x * subC.x + tw.y * subC.y + tw.z * subC.z;
		p.dx.x = W.x * adx.x + W.y * adx.y + W.z * adx.z;
		p.dx.y = adx.x * subB.x + adx.y * subB.y + adx.z * subB.z;
		p.dx.z = adx.x * subC.x + adx.y * subC.y + adx.z * subC.z;
		p.dy.x = W.x * ady.x + W.y * ady.y + W.z * ady.z;
		p.dy.y = ady.x * subB.x + ady.y * subB.y + ady.z * subB.z;
		p.dy.z = ady.x * subC.x + ady.y * subC.y + ady.z * subC.z;
	} else {
		V3 IW{1.0f / t.p[0].cs.z, 1.0f / t.p[1].cs.z, 1.0f / t.p[2].cs.z};
		p.affine = false;
		p.start.x = IW.x * tw.x + IW.y * tw.y + IW.z * tw.z;
		p.start.y = IW.x * tw.x * subB.x + IW.y * tw.y * subB.y + IW.z * tw.z * subB.z;
		p.start.z = IW.x * tw.x * subC.x + IW.y * tw.y * subC.y + IW.z * tw.z * subC.z;
		p.dx.x = IW.x * adx.x + IW.y * adx.y + IW.z * adx.z;
		p.dx.y = IW.x * adx.x * subB.x + IW.y * adx.y * subB.y + IW.z * adx.z * subB.z;
		p.dx.z = IW.x * adx.x * subC.x + IW.y * adx.y * subC.y + IW.z * adx.z * subC.z;
		p.dy.x = IW.x * ady.x + IW.y * ady.y + IW.z * ady.z;
		p.dy.y = IW.x * ady.x * subB.x + IW.y * ady.y * subB.y + IW.z * ady.z * subB.z;
		p.dy.z = IW.x * ady.x * subC.x + IW.y * ady.y * subC.y + IW.z * ady.z * subC.z;
	}
	return p;
}

inline V3 add(const V3 &a, const V3 &b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scale(const V3 &a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }

// Projection::getDepthDividedWeight_perspective / getWeight_affine at a pixel (ITriangle2D.h:80-100)
inline V3 weightAtPixel(const Projection &p, int32_t x, int32_t y) {
	float fx = (float)x + 0.5f, fy = (float)y + 0.5f;
	return V3{(p.start.x + p.dx.x * fx) + p.dy.x * fy, (p.start.y + p.dx.y * fx) + p.dy.y * fy, (p.start.z + p.dx.z * fx) + p.dy.z * fy};
}

// ---------------------------------------------------------------------------
// Canvas and the per-quad shading (shader/fillerTemplates.h, RgbaMultiply.h).
// ---------------------------------------------------------------------------

struct Canvas {
	uint32_t *color;  // tightly packed rows, or null
	float *depth;     // tightly packed rows, or null
	int32_t width, height;
};

struct ShaderChoice {
	bool diffuse, light, fading, colorless;
	V3 red, green, blue, alpha;  // scaled corner colours (RgbaMultiply.h:37-58)
};

// RgbaMultiply_data and the template choice in processTriangle_RgbaMultiply (RgbaMultiply.h:37-58, :110-175)
ShaderChoice chooseShader(const TriangleInput &in) {
	ShaderChoice s;
	s.diffuse = in.diffuse >= 0;
	s.light = in.light >= 0;
	float k = 255.0f;
	if (s.diffuse) k *= 1.0f / 255.0f;
	if (s.light) k *= 1.0f / 255.0f;
	s.red = scale(in.red, k); s.green = scale(in.green, k); s.blue = scale(in.blue, k); s.alpha = scale(in.alpha, k);
	s.fading = !(almostSame3(s.red) && almostSame3(s.green) && almostSame3(s.blue) && almostSame3(s.alpha));
	// The "without normalization" (texture-only) branches sit under !hasVertexFade
	// (RgbaMultiply.h:137-149, :157-162): a colourless triangle that fades takes the
	// texture x interpolated vertex colour shader instead.
	s.colorless = !s.fading && almostOne3(s.red) && almostOne3(s.green) && almostOne3(s.blue) && almostOne3(s.alpha);
	return s;
}

// shaderMethods::interpolate (shaderMethods.h:39-44)
inline float interp(const V3 &d, float wa, float wb, float wc) { return ((d.x * wa) + (d.y * wb)) + (d.z * wc); }

struct Rgba4 { float r[4], g[4], b[4], a[4]; };

inline void unpack(uint32_t c, int lane, Rgba4 &o) {
	o.r[lane] = fromU32(c & 255u); o.g[lane] = fromU32((c >> 8) & 255u);
	o.b[lane] = fromU32((c >> 16) & 255u); o.a[lane] = fromU32(c >> 24);
}

// getPixels_2x2 (RgbaMultiply.h:74-106) for the four lanes of one quad.
void shadeQuad(const ShaderChoice &s, const TriangleInput &in, const std::vector<Texture> &tex,
               const float *wa, const float *wb, const float *wc, Rgba4 &out) {
	if (s.diffuse && !s.light && s.colorless) {
		float u[4], v[4];
		for (int l = 0; l < 4; l++) { u[l] = interp(in.u1, wa[l], wb[l], wc[l]); v[l] = interp(in.v1, wa[l], wb[l], wc[l]); }
		const Texture &t = tex[in.diffuse];
		uint32_t mip = mipLevelIndex(t, u, v);
		for (int l = 0; l < 4; l++) unpack(sampleBilinear(t, u[l], v[l], mip, false), l, out);
		return;
	}
	if (s.light && !s.diffuse && s.colorless) {
		const Texture &t = tex[in.light];
		for (int l = 0; l < 4; l++) {
			float u = interp(in.u2, wa[l], wb[l], wc[l]);
			float v = interp(in.v2, wa[l], wb[l], wc[l]);
			unpack(sampleBilinear(t, u, v, 0u, true), l, out);
		}
		return;
	}
	for (int l = 0; l < 4; l++) {
		if (s.fading) {
			out.r[l] = interp(s.red, wa[l], wb[l], wc[l]);
			out.g[l] = interp(s.green, wa[l], wb[l], wc[l]);
			out.b[l] = interp(s.blue, wa[l], wb[l], wc[l]);
			out.a[l] = interp(s.alpha, wa[l], wb[l], wc[l]);
		} else {
			out.r[l] = s.red.x; out.g[l] = s.green.x; out.b[l] = s.blue.x; out.a[l] = s.alpha.x;
		}
	}
	if (s.diffuse) {
		float u[4], v[4];
		for (int l = 0; l < 4; l++) { u[l] = interp(in.u1, wa[l], wb[l], wc[l]); v[l] = interp(in.v1, wa[l], wb[l], wc[l]); }
		const Texture &t = tex[in.diffuse];
		uint32_t mip = mipLevelIndex(t, u, v);
		Rgba4 smp;
		for (int l = 0; l < 4; l++) unpack(sampleBilinear(t, u[l], v[l], mip, false), l, smp);
		for (int l = 0; l < 4; l++) { out.r[l] *= smp.r[l]; out.g[l] *= smp.g[l]; out.b[l] *= smp.b[l]; out.a[l] *= smp.a[l]; }
	}
	if (s.light) {
		const Texture &t = tex[in.light];
		Rgba4 smp;
		for (int l = 0; l < 4; l++) {
			float u = interp(in.u2, wa[l], wb[l], wc[l]);
			float v = interp(in.v2, wa[l], wb[l], wc[l]);
			unpack(sampleBilinear(t, u, v, 0u, true), l, smp);
		}
		for (int l = 0; l < 4; l++) { out.r[l] *= smp.r[l]; out.g[l] *= smp.g[l]; out.b[l] *= smp.b[l]; out.a[l] *= smp.a[l]; }
	}
}

// packOrder_floatToSaturatedByte for RGBA (PackOrder.h:191-212)
inline uint32_t packSaturated(float r, float g, float b, float a) {
	return truncU32(minps(r, 255.1f)) | (truncU32(minps(g, 255.1f)) << 8) | (truncU32(minps(b, 255.1f)) << 16) | (truncU32(minps(a, 255.1f)) << 24);
}

// Inputs and outputs of the differentiable float shader (no reference counterpart): colour =
// interpolated per-point RGB multipliers x bilinear(float RGBA texture, level 0), alpha = the
// texture's alpha; taps and 8-bit sub-texel weights as texture_sample_bilinear at the highest
// resolution (textureAPI.h:342-438).
struct DiffTarget {
	const float *texture;
	int32_t log2w, log2h;
	const float *multipliers;  // point x 3, or null for all 1
	float *color;              // W*H*4
	int32_t *visibility;       // W*H: command * 4 + lane, or -1
	float *weights;            // W*H*5: wa, wb, wc, u, v of the lane that drew the pixel
};

struct FillContext {
	const Canvas *canvas;
	const std::vector<Texture> *textures;
	const TriangleInput *input;
	ShaderChoice shader;
	bool colorWrite, depthRead, depthWrite, alphaFilter, affine;
	int32_t maxHeight;
	const DiffTarget *diff;
	int32_t commandIndex;
};

struct Taps { uint32_t index[4]; float weight[4]; };

Taps bilinearTaps(int32_t log2w, int32_t log2h, float u, float v) {
	uint32_t scaleU = 256u << log2w, scaleV = 256u << log2h;
	uint32_t maskX = (1u << log2w) - 1u, maskY = (1u << log2h) - 1u;
	uint32_t subX = truncU32((u + 256.0f) * fromU32(scaleU)) - 128u;
	uint32_t subY = truncU32((v + 256.0f) * fromU32(scaleV)) - 128u;
	float fx = (float)(subX & 255u) * (1.0f / 256.0f), fy = (float)(subY & 255u) * (1.0f / 256.0f);
	uint32_t left = (subX >> 8) & maskX, top = (subY >> 8) & maskY;
	uint32_t right = (left + 1u) & maskX, bottom = (top + 1u) & maskY;
	Taps t;
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

inline float multiplierOf(const DiffTarget &d, int32_t point, int ch) { return d.multipliers ? d.multipliers[point * 3 + ch] : 1.0f; }

// fillQuadSuper (fillerTemplates.h:154-187) for the quad whose upper-left pixel is (x, yUpper).
// rowUpper/rowLower are where the lanes 0-1 and 2-3 read and write (the lower lanes alias
// the upper row when the lower row is missing, fillerTemplates.h:297-323).
void fillQuad(const FillContext &f, bool clipSides, int32_t x, int32_t rowUpper, int32_t rowLower,
              const Row &upper, const Row &lower, const float *depth, const float *wb, const float *wc) {
	const Canvas &cv = *f.canvas;
	bool clip[4];
	if (clipSides) {
		int32_t x2 = x + 1;
		clip[0] = x >= upper.left && x < upper.right;
		clip[1] = x2 >= upper.left && x2 < upper.right;
		clip[2] = x >= lower.left && x < lower.right;
		clip[3] = x2 >= lower.left && x2 < lower.right;
	} else {
		clip[0] = clip[1] = clip[2] = clip[3] = true;
	}
	const int32_t px[4] = {x, x + 1, x, x + 1};
	const int32_t py[4] = {rowUpper, rowUpper, rowLower, rowLower};
	bool vis[4];
	for (int l = 0; l < 4; l++) {
		bool front = true;
		if (f.depthRead) {
			if (clipSides && !clip[l]) {
				front = false;
			} else {
				float old = cv.depth[(size_t)py[l] * cv.width + px[l]];
				front = f.affine ? (depth[l] < old) : (depth[l] > old);
			}
		}
		vis[l] = clip[l] && front;
	}
	if (!(vis[0] || vis[1] || vis[2] || vis[3])) return;
	if (f.diff) {
		const DiffTarget &d = *f.diff;
		const TriangleInput &in = *f.input;
		for (int l = 0; l < 4; l++) {
			if (!vis[l]) continue;
			float wa = 1.0f - (wb[l] + wc[l]);
			float u = interp(in.u1, wa, wb[l], wc[l]), v = interp(in.v1, wa, wb[l], wc[l]);
			Taps taps = bilinearTaps(d.log2w, d.log2h, u, v);
			float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
			for (int k = 0; k < 4; k++)
				for (int ch = 0; ch < 4; ch++) t[ch] += taps.weight[k] * d.texture[(size_t)taps.index[k] * 4 + ch];
			size_t p = (size_t)py[l] * cv.width + px[l];
			for (int ch = 0; ch < 3; ch++)
				d.color[p * 4 + ch] = (wa * multiplierOf(d, in.point[0], ch) + wb[l] * multiplierOf(d, in.point[1], ch) +
				                       wc[l] * multiplierOf(d, in.point[2], ch)) * t[ch];
			d.color[p * 4 + 3] = t[3];
			d.visibility[p] = f.commandIndex * 4 + l;
			d.weights[p * 5 + 0] = wa; d.weights[p * 5 + 1] = wb[l]; d.weights[p * 5 + 2] = wc[l];
			d.weights[p * 5 + 3] = u; d.weights[p * 5 + 4] = v;
			if (f.depthWrite) cv.depth[p] = depth[l];
		}
		return;
	}
	if (f.colorWrite) {
		float wa[4];
		for (int l = 0; l < 4; l++) wa[l] = 1.0f - (wb[l] + wc[l]);
		Rgba4 c;
		shadeQuad(f.shader, *f.input, *f.textures, wa, wb, wc, c);
		if (f.alphaFilter) {
			Rgba4 target;
			for (int l = 0; l < 4; l++) {
				uint32_t old = (!clipSides || vis[l]) ? cv.color[(size_t)py[l] * cv.width + px[l]] : 0u;
				unpack(old, l, target);
			}
			for (int l = 0; l < 4; l++) {
				float opacity = c.a[l] * (1.0f / 255.0f);
				float inv = 1.0f - opacity;
				c.r[l] = (c.r[l] * opacity) + (target.r[l] * inv);
				c.g[l] = (c.g[l] * opacity) + (target.g[l] * inv);
				c.b[l] = (c.b[l] * opacity) + (target.b[l] * inv);
				c.a[l] = (c.a[l] * opacity) + (target.a[l] * inv);
			}
		}
		for (int l = 0; l < 4; l++) {
			if (vis[l]) cv.color[(size_t)py[l] * cv.width + px[l]] = packSaturated(c.r[l], c.g[l], c.b[l], c.a[l]);
		}
	}
	if (f.depthWrite) {
		for (int l = 0; l < 4; l++) {
			if (vis[l]) cv.depth[(size_t)py[l] * cv.width + px[l]] = depth[l];
		}
	}
}

// fillRowSuper (fillerTemplates.h:195-244): quads from startX to endX, lanes built from the
// upper and lower row weights and advanced by 2*dx per quad.
void fillRow(const FillContext &f, bool clipSides, int32_t rowUpper, int32_t rowLower, V3 wUpper, V3 wLower,
             const V3 &dx, int32_t startX, int32_t endX, const Row &upper, const Row &lower) {
	V3 dx2 = scale(dx, 2.0f);
	float l0[4] = {wUpper.x, wUpper.x + dx.x, wLower.x, wLower.x + dx.x};
	float l1[4] = {wUpper.y, wUpper.y + dx.y, wLower.y, wLower.y + dx.y};
	float l2[4] = {wUpper.z, wUpper.z + dx.z, wLower.z, wLower.z + dx.z};
	for (int32_t x = startX; x < endX; x += 2) {
		float depth[4], wb[4], wc[4];
		for (int l = 0; l < 4; l++) {
			depth[l] = l0[l];
			if (f.affine) {
				wb[l] = l1[l];
				wc[l] = l2[l];
			} else {
				float w = reciprocalLane(l0[l]);
				wb[l] = l1[l] * w;
				wc[l] = l2[l] * w;
			}
		}
		fillQuad(f, clipSides, x, rowUpper, rowLower, upper, lower, depth, wb, wc);
		for (int l = 0; l < 4; l++) { l0[l] = l0[l] + dx2.x; l1[l] = l1[l] + dx2.y; l2[l] = l2[l] + dx2.z; }
	}
}

inline uint32_t roundUpEven(uint32_t x) { return (x + 1u) & ~1u; }  // fillerTemplates.h:57-59
inline uint32_t roundDownEven(uint32_t x) { return x & ~1u; }       // fillerTemplates.h:61-63

// fillShapeSuper (fillerTemplates.h:246-385)
void fillShape(const FillContext &f, const Projection &proj, int32_t startRow, int32_t rowCount, const Row *rows) {
	const V3 dx2 = scale(proj.dx, 2.0f);
	for (int32_t y1 = startRow; y1 < startRow + rowCount; y1 += 2) {
		int32_t y2 = y1 + 1;
		Row upper = rows[y1 - startRow];
		Row lower = rows[y2 - startRow];
		int32_t outerStart = std::min(upper.left, lower.left);
		int32_t outerEnd = std::max(upper.right, lower.right);
		int32_t innerStart = std::max(upper.left, lower.left);
		int32_t innerEnd = std::min(upper.right, lower.right);
		int32_t outerBlockStart = (int32_t)roundDownEven((uint32_t)outerStart);
		int32_t outerBlockEnd = (int32_t)roundUpEven((uint32_t)outerEnd);
		int32_t innerBlockStart = (int32_t)roundUpEven((uint32_t)innerStart);
		int32_t innerBlockEnd = (int32_t)roundDownEven((uint32_t)innerEnd);
		if (y2 >= f.maxHeight) lower.right = lower.left;
		bool hasTop = upper.right > upper.left;
		bool hasBottom = lower.right > lower.left;
		if (!(hasTop || hasBottom)) continue;
		int32_t rowUpper = hasTop ? y1 : y2;
		int32_t rowLower = hasBottom ? y2 : y1;
		V3 wUpper = weightAtPixel(proj, outerBlockStart, y1);
		V3 wLower = add(wUpper, proj.dy);
		if (innerBlockEnd <= innerBlockStart) {
			for (int32_t x = outerBlockStart; x < outerBlockEnd; x += 2) {
				fillRow(f, true, rowUpper, rowLower, wUpper, wLower, proj.dx, x, x + 2, upper, lower);
				wUpper = add(wUpper, dx2); wLower = add(wLower, dx2);
			}
		} else {
			for (int32_t x = outerBlockStart; x < innerBlockStart; x += 2) {
				fillRow(f, true, rowUpper, rowLower, wUpper, wLower, proj.dx, x, x + 2, upper, lower);
				wUpper = add(wUpper, dx2); wLower = add(wLower, dx2);
			}
			int32_t quadCount = (innerBlockEnd - innerBlockStart) / 2;
			fillRow(f, false, rowUpper, rowLower, wUpper, wLower, proj.dx, innerBlockStart, innerBlockEnd, Row{0, 0}, Row{0, 0});
			wUpper = add(wUpper, scale(dx2, (float)quadCount)); wLower = add(wLower, scale(dx2, (float)quadCount));
			for (int32_t x = innerBlockEnd; x < outerBlockEnd; x += 2) {
				fillRow(f, true, rowUpper, rowLower, wUpper, wLower, proj.dx, x, x + 2, upper, lower);
				wUpper = add(wUpper, dx2); wLower = add(wLower, dx2);
			}
		}
	}
}

struct Frame {
	Camera camera;
	Canvas canvas;
	const DiffTarget *diff = nullptr;
	int32_t filter;
	std::vector<Texture> textures;
	std::vector<TriangleInput> inputs;
	std::vector<Command> commands;
	// Occlusion grid (RendererImpl::depthGrid, rendererAPI.cpp:140-188): 16 x 16 pixel cells.
	std::vector<float> grid;
	int32_t gridW = 0, gridH = 0;
	bool occluded = false;
};

// executeTriangleDrawing (renderCore.cpp:203-217) with the shader dispatch of fillShape (fillerTemplates.h:387-441).
void executeCommand(const Frame &fr, const Command &cmd, int32_t commandIndex, const Rect &slice) {
	const Canvas &cv = fr.canvas;
	Rect full = rectFromSides(0, 0, cv.width, cv.height);
	Rect clip = cut(full, slice);
	if (!overlaps(cmd.tri.wholeBound, clip)) return;
	Rect bound = alignedRasterBound(cmd.tri, clip);
	int32_t rowCount = bound.h;
	if (rowCount <= 0) return;
	std::vector<Row> rows((size_t)rowCount);
	rasterize(cmd.tri.p[0].flat, cmd.tri.p[1].flat, cmd.tri.p[2].flat, rows.data(), bound);
	Projection proj = getProjection(cmd.tri, cmd.subB, cmd.subC, fr.camera.perspective);
	FillContext f;
	f.canvas = &cv;
	f.textures = &fr.textures;
	f.input = &fr.inputs[cmd.input];
	f.shader = chooseShader(*f.input);
	f.affine = proj.affine;
	bool hasColor = cv.color != nullptr, hasDepth = cv.depth != nullptr;
	bool solid = fr.filter == 0;
	f.colorWrite = hasColor;
	f.depthRead = hasDepth;
	f.depthWrite = hasDepth && (solid || !hasColor);
	f.alphaFilter = hasColor && !solid;
	f.maxHeight = cv.height;
	f.diff = fr.diff;
	f.commandIndex = commandIndex;
	if (fr.diff) {
		f.colorWrite = false;
		f.alphaFilter = false;
		f.depthWrite = hasDepth;
	} else if (!hasColor && !hasDepth) {
		return;
	}
	fillShape(f, proj, bound.t, rowCount, rows.data());
}

// renderTriangleFromData + renderTriangleWithShader + drawClippedTriangle + drawSubTriangle
// (renderCore.cpp:220-341): culling, clipping and the commands a triangle adds to the queue.
void giveTriangle(Frame &fr, const Projected &a, const Projected &b, const Projected &c, const TriangleInput &in, int32_t source) {
	if (fr.canvas.color == nullptr && fr.canvas.depth == nullptr && fr.diff == nullptr) return;
	Tri tri = makeTri(a, b, c);
	if (triangleVisibility(tri, fr.camera.cull) == Hidden) return;
	if (fr.filter != 0 && almostZero3(in.alpha)) return;
	int32_t inputIndex = (int32_t)fr.inputs.size();
	bool used = false;
	if (triangleVisibility(tri, fr.camera.clip) == Full) {
		if (frontFacing(tri)) {
			fr.commands.push_back(Command{tri, V3{0.0f, 1.0f, 0.0f}, V3{0.0f, 0.0f, 1.0f}, inputIndex, source});
			used = true;
		}
	} else {
		Polygon9 poly;
		poly.v[0].cs = tri.p[0].cs; poly.v[0].subB = 0.0f; poly.v[0].subC = 0.0f;
		poly.v[1].cs = tri.p[1].cs; poly.v[1].subB = 1.0f; poly.v[1].subC = 0.0f;
		poly.v[2].cs = tri.p[2].cs; poly.v[2].subB = 0.0f; poly.v[2].subC = 1.0f;
		poly.count = 3;
		for (int32_t s = 0; s < fr.camera.clip.count; s++) poly.clip(fr.camera.clip.planes[s]);
		for (int32_t i = 0; i < poly.count - 2; i++) {
			const SubVertex &va = poly.v[0], &vb = poly.v[1 + i], &vc = poly.v[2 + i];
			Tri sub = makeTri(cameraToScreen(fr.camera, va.cs), cameraToScreen(fr.camera, vb.cs), cameraToScreen(fr.camera, vc.cs));
			if (frontFacing(sub)) {
				fr.commands.push_back(Command{sub, V3{va.subB, vb.subB, vc.subB}, V3{va.subC, vb.subC, vc.subC}, inputIndex, source});
				used = true;
			}
		}
	}
	if (used) fr.inputs.push_back(in);
}

inline V3 cornerOf(const float *p4, int comp) { (void)comp; return V3{p4[0], p4[1], p4[2]}; }

// ---------------------------------------------------------------------------
// Occlusion grid (api/rendererAPI.cpp:36-477).
// ---------------------------------------------------------------------------
const int32_t cellSize = 16;  // rendererAPI.cpp:36

// prepareForOcclusion (rendererAPI.cpp:179-188)
void prepareForOcclusion(Frame &fr) {
	if (!fr.occluded) fr.grid.assign((size_t)fr.gridW * (size_t)fr.gridH, std::numeric_limits<float>::infinity());
	fr.occluded = true;
}

// getOuterCellBound (rendererAPI.cpp:166-177)
Rect outerCellBound(const Frame &fr, const Rect &pixel) {
	int32_t minX = pixel.l / cellSize, maxX = pixel.right() / cellSize + 1;
	int32_t minY = pixel.t / cellSize, maxY = pixel.bottom() / cellSize + 1;
	minX = std::max(minX, 0); minY = std::max(minY, 0);
	maxX = std::min(maxX, fr.gridW); maxY = std::min(maxY, fr.gridH);
	Rect r; r.l = minX; r.t = minY; r.w = maxX - minX; r.h = maxY - minY;
	return r;
}

// counterClockwise (rendererAPI.cpp:38-40)
bool counterClockwise(const Projected &p, const Projected &q, const Projected &r) {
	return (q.flat.y - p.flat.y) * (r.flat.x - q.flat.x) - (q.flat.x - p.flat.x) * (r.flat.y - q.flat.y) < 0;
}

// jarvisConvexHullAlgorithm (rendererAPI.cpp:44-72)
void jarvisHull(const Projected *in, int32_t n, Projected *out, int32_t &outCount) {
	if (n < 3) {
		outCount = n;
		for (int32_t i = 0; i < n; i++) out[i] = in[i];
		return;
	}
	int32_t l = 0;
	outCount = 0;
	for (int32_t i = 1; i < n; i++)
		if (in[i].flat.x < in[l].flat.x) l = i;
	int32_t p = l;
	do {
		if (outCount >= n) return;
		out[outCount++] = in[p];
		int32_t q = (p + 1) % n;
		for (int32_t i = 0; i < n; i++)
			if (counterClockwise(in[p], in[i], in[q])) q = i;
		p = q;
	} while (p != l);
}

// getPixelBoundFromProjection (rendererAPI.cpp:88-94) with IRect::merge (math/IRect.h:68-74)
Rect pixelBoundFromProjection(const Projected *pts, int32_t n) {
	int32_t l = (int32_t)(pts[0].flat.x / 256), t = (int32_t)(pts[0].flat.y / 256), r = l + 1, b = t + 1;
	for (int32_t i = 1; i < n; i++) {
		int32_t x = (int32_t)(pts[i].flat.x / 256), y = (int32_t)(pts[i].flat.y / 256);
		l = std::min(l, x); t = std::min(t, y); r = std::max(r, x + 1); b = std::max(b, y + 1);
	}
	return rectFromSides(l, t, r, b);
}

// pointInsideOfEdge / pointInsideOfHull / rectangleInsideOfHull (rendererAPI.cpp:97-129)
bool pointInsideOfEdge(const L2 &a, const L2 &b, const L2 &point) {
	L2 dir{b.y - a.y, a.x - b.x};
	L2 rel{point.x - a.x, point.y - a.y};
	return (dir.x * rel.x) + (dir.y * rel.y) <= 0;
}
bool pointInsideOfHull(const Projected *hull, int32_t n, const L2 &point) {
	for (int32_t c = 0; c < n; c++) {
		int32_t nc = c + 1;
		if (nc == n) nc = 0;
		if (!pointInsideOfEdge(hull[c].flat, hull[nc].flat, point)) return false;
	}
	return true;
}
bool rectangleInsideOfHull(const Projected *hull, int32_t n, const Rect &r) {
	return pointInsideOfHull(hull, n, L2{r.l, r.t}) && pointInsideOfHull(hull, n, L2{r.right(), r.t}) &&
	       pointInsideOfHull(hull, n, L2{r.l, r.bottom()}) && pointInsideOfHull(hull, n, L2{r.right(), r.bottom()});
}

// occludeFromSortedHull (rendererAPI.cpp:218-239)
void occludeFromSortedHull(Frame &fr, const Projected *hull, int32_t n, const Rect &pixelBound) {
	if (!(pixelBound.w > cellSize && pixelBound.h > cellSize)) return;
	float distance = 0.0f;
	for (int32_t c = 0; c < n; c++)
		if (hull[c].cs.z > distance) distance = hull[c].cs.z;
	Rect cells = outerCellBound(fr, pixelBound);
	for (int32_t cy = cells.t; cy < cells.bottom(); cy++) {
		for (int32_t cx = cells.l; cx < cells.right(); cx++) {
			Rect sub;
			sub.l = cx * cellSize * 256; sub.t = cy * cellSize * 256; sub.w = cellSize * 256; sub.h = cellSize * 256;
			if (rectangleInsideOfHull(hull, n, sub)) {
				float &cell = fr.grid[(size_t)cy * (size_t)fr.gridW + (size_t)cx];
				if (distance < cell) cell = distance;
			}
		}
	}
}

// The corners of a box (GENERATE_BOX_CORNERS, rendererAPI.cpp:259-267)
void boxCornersOf(const V3 &mn, const V3 &mx, V3 *out) {
	out[0] = V3{mn.x, mn.y, mn.z}; out[1] = V3{mn.x, mn.y, mx.z}; out[2] = V3{mn.x, mx.y, mn.z}; out[3] = V3{mn.x, mx.y, mx.z};
	out[4] = V3{mx.x, mn.y, mn.z}; out[5] = V3{mx.x, mn.y, mx.z}; out[6] = V3{mx.x, mx.y, mn.z}; out[7] = V3{mx.x, mx.y, mx.z};
}

// occludeFromBox with projectHull (rendererAPI.cpp:76-90, :270-295)
void occludeFromBox(Frame &fr, const V3 &mn, const V3 &mx, const Transform &modelToWorld) {
	prepareForOcclusion(fr);
	V3 local[8];
	boxCornersOf(mn, mx, local);
	Projected proj[8], hull[8];
	for (int32_t p = 0; p < 8; p++) {
		V3 cameraPoint = transformPointTransposedInverse(fr.camera.location, transformPoint(modelToWorld, local[p]));
		V3 narrow{cameraPoint.x * 0.5f, cameraPoint.y * 0.5f, cameraPoint.z * 1.0f};
		for (int32_t s = 0; s < fr.camera.cull.count; s++)
			if (!insidePlane(fr.camera.cull.planes[s], narrow)) return;
		proj[p] = cameraToScreen(fr.camera, cameraPoint);
	}
	int32_t count = 0;
	jarvisHull(proj, 8, hull, count);
	occludeFromSortedHull(fr, hull, count, pixelBoundFromProjection(hull, count));
}

// isBoxOccluded / isHullOccluded (rendererAPI.cpp:298-343).  Before the first occluder call a
// fresh renderer has no grid image, which image_readPixel_clamp reads as 0 (imageAPI.h:276-282).
bool isBoxOccluded(const Frame &fr, const V3 &mn, const V3 &mx, const Transform &modelToWorld) {
	V3 local[8], cam[8];
	Projected proj[8];
	boxCornersOf(mn, mx, local);
	for (int32_t p = 0; p < 8; p++) {
		cam[p] = transformPointTransposedInverse(fr.camera.location, transformPoint(modelToWorld, local[p]));
		proj[p] = cameraToScreen(fr.camera, cam[p]);
	}
	for (int32_t s = 0; s < fr.camera.cull.count; s++) {
		bool allOutside = true;
		for (int32_t p = 0; p < 8; p++)
			if (insidePlane(fr.camera.cull.planes[s], cam[p])) { allOutside = false; break; }
		if (allOutside) return true;
	}
	Rect pixelBound = pixelBoundFromProjection(proj, 8);
	float closest = std::numeric_limits<float>::infinity();
	for (int32_t c = 0; c < 8; c++)
		if (proj[c].cs.z < closest) closest = proj[c].cs.z;
	Rect cells = outerCellBound(fr, pixelBound);
	for (int32_t cy = cells.t; cy < cells.bottom(); cy++)
		for (int32_t cx = cells.l; cx < cells.right(); cx++) {
			float cell = fr.occluded ? fr.grid[(size_t)cy * (size_t)fr.gridW + (size_t)cx] : 0.0f;
			if (closest < cell) return false;
		}
	return true;
}

// occludeFromTopRows (rendererAPI.cpp:403-474): one scan along the first row of each cell row,
// where each cell after the first starts one column early (right starts at cellSize - 1).
void occludeFromTopRows(Frame &fr) {
	prepareForOcclusion(fr);
	const int32_t width = fr.canvas.width, height = fr.canvas.height;
	for (int32_t y = 0, gy = 0; y < height; y += cellSize, gy++) {
		const float *depthRow = fr.canvas.depth + (size_t)y * (size_t)width;
		int32_t x = 0, right = cellSize - 1;
		for (int32_t gx = 0; gx < fr.gridW; gx++) {
			if (right >= width) right = width;
			float farthest;
			if (fr.camera.perspective) {
				float smallestInverse = std::numeric_limits<float>::infinity();
				while (x < right) {
					if (depthRow[x] < smallestInverse) smallestInverse = depthRow[x];
					x++;
				}
				farthest = 1.0f / smallestInverse;
			} else {
				farthest = 0.0f;
				while (x < right) {
					if (depthRow[x] > farthest) farthest = depthRow[x];
					x++;
				}
			}
			float &cell = fr.grid[(size_t)gy * (size_t)fr.gridW + (size_t)gx];
			if (farthest < cell) cell = farthest;
			right += cellSize;
		}
	}
}

// occludeFromExistingTriangles (rendererAPI.cpp:241-258): the Solid commands so far.
void occludeFromExistingTriangles(Frame &fr) {
	prepareForOcclusion(fr);
	if (fr.filter != 0) return;  // one model per scene: every command has the scene's filter
	for (const Command &c : fr.commands) occludeFromSortedHull(fr, c.tri.p, 3, c.tri.wholeBound);
}

// completeOcclusion (rendererAPI.cpp:190-215): commands whose nearest corner is not nearer than
// 0.001 past any cell under their bound are removed before execution.
void completeOcclusion(Frame &fr) {
	if (!fr.occluded) return;
	std::vector<Command> kept;
	for (const Command &c : fr.commands) {
		bool anyVisible = false;
		Rect cells = outerCellBound(fr, c.tri.wholeBound);
		for (int32_t cy = cells.t; cy < cells.bottom(); cy++) {
			for (int32_t cx = cells.l; cx < cells.right(); cx++) {
				float background = fr.grid[(size_t)cy * (size_t)fr.gridW + (size_t)cx];
				float nearest = c.tri.p[0].cs.z;
				if (c.tri.p[1].cs.z < nearest) nearest = c.tri.p[1].cs.z;
				if (c.tri.p[2].cs.z < nearest) nearest = c.tri.p[2].cs.z;
				if (nearest < background + 0.001) anyVisible = true;
			}
		}
		if (anyVisible) kept.push_back(c);
	}
	fr.commands.swap(kept);
}

// model_render_threaded (modelAPI.cpp:214-281): box cull, project every point once,
// then fan each polygon into (0,1,2) and (0,2,3).
void giveModel(Frame &fr, const scene_desc *s, const Transform &modelToWorld, int32_t instance, int32_t &sourceCounter) {
	// The model bound starts at the origin (ModelImpl's default-constructed minBound/maxBound,
	// Model.cpp:281-287), then grows with every point.
	V3 mn{0.0f, 0.0f, 0.0f}, mx{0.0f, 0.0f, 0.0f};
	for (int32_t i = 0; i < s->point_count; i++) {
		const float *p = s->points + i * 3;
		if (mn.x > p[0]) mn.x = p[0];
		if (mn.y > p[1]) mn.y = p[1];
		if (mn.z > p[2]) mn.z = p[2];
		if (mx.x < p[0]) mx.x = p[0];
		if (mx.y < p[1]) mx.y = p[1];
		if (mx.z < p[2]) mx.z = p[2];
	}
	int32_t polygonTriangles = 0;
	for (int32_t part = 0; part < s->part_count; part++)
		for (int32_t poly = s->part_polygon_start[part]; poly < s->part_polygon_start[part + 1]; poly++)
			polygonTriangles += (s->indices[poly * 4 + 3] >= 0) ? 2 : 1;
	(void)instance;
	if (boxSeen(fr.camera, mn, mx, modelToWorld) == 0 || (fr.occluded && isBoxOccluded(fr, mn, mx, modelToWorld))) {
		sourceCounter += polygonTriangles;
		return;
	}
	std::vector<Projected> proj((size_t)s->point_count);
	for (int32_t i = 0; i < s->point_count; i++) {
		const float *p = s->points + i * 3;
		proj[(size_t)i] = cameraToScreen(fr.camera, transformPointTransposedInverse(fr.camera.location, transformPoint(modelToWorld, V3{p[0], p[1], p[2]})));
	}
	for (int32_t part = 0; part < s->part_count; part++) {
		int32_t diffuse = s->part_diffuse ? s->part_diffuse[part] : -1;
		int32_t light = s->part_light ? s->part_light[part] : -1;
		for (int32_t poly = s->part_polygon_start[part]; poly < s->part_polygon_start[part + 1]; poly++) {
			const int32_t *idx = s->indices + poly * 4;
			const float *col = s->colors + poly * 16;
			const float *tc = s->texcoords + poly * 16;
			int32_t fan = (idx[3] >= 0) ? 2 : 1;
			for (int32_t k = 0; k < fan; k++) {
				int32_t ca = 0, cb = 1 + k, cc = 2 + k;
				TriangleInput in;
				const float *A = col + ca * 4, *B = col + cb * 4, *Cc = col + cc * 4;
				in.red = V3{A[0], B[0], Cc[0]}; in.green = V3{A[1], B[1], Cc[1]};
				in.blue = V3{A[2], B[2], Cc[2]}; in.alpha = V3{A[3], B[3], Cc[3]};
				const float *TA = tc + ca * 4, *TB = tc + cb * 4, *TC = tc + cc * 4;
				in.u1 = V3{TA[0], TB[0], TC[0]}; in.v1 = V3{TA[1], TB[1], TC[1]};
				in.u2 = V3{TA[2], TB[2], TC[2]}; in.v2 = V3{TA[3], TB[3], TC[3]};
				in.diffuse = diffuse; in.light = light;
				in.point[0] = idx[ca]; in.point[1] = idx[cb]; in.point[2] = idx[cc];
				giveTriangle(fr, proj[(size_t)idx[ca]], proj[(size_t)idx[cb]], proj[(size_t)idx[cc]], in, sourceCounter);
				sourceCounter++;
			}
		}
	}
}

void buildFrame(const scene_desc *s, Frame &fr, std::vector<uint32_t> &color, std::vector<float> &depth) {
	fr.camera = makeCamera(s);
	fr.filter = s->filter;
	for (int32_t t = 0; t < s->texture_count; t++)
		fr.textures.push_back(makeTexture(s->texture_dims[t * 3], s->texture_dims[t * 3 + 1], s->texture_dims[t * 3 + 2], s->texture_level0[t]));
	size_t n = (size_t)s->width * (size_t)s->height;
	if (s->has_color) color.assign(n, 0u);
	if (s->has_depth) depth.assign(n, 0.0f);
	fr.canvas.color = s->has_color ? color.data() : nullptr;
	fr.canvas.depth = s->has_depth ? depth.data() : nullptr;
	fr.canvas.width = s->width;
	fr.canvas.height = s->height;
}

void loadCanvas(const scene_desc *s, Frame &fr) {
	size_t n = (size_t)s->width * (size_t)s->height;
	if (fr.canvas.color) {
		if (s->initial_color) memcpy(fr.canvas.color, s->initial_color, n * 4); else memset(fr.canvas.color, 0, n * 4);
	}
	if (fr.canvas.depth) {
		if (s->initial_depth) memcpy(fr.canvas.depth, s->initial_depth, n * 4); else memset(fr.canvas.depth, 0, n * 4);
	}
}

// renderer_giveTask_triangle (rendererAPI.cpp:503-519) per polygon with the caller's
// ProjectedPoints: no box cull and no projection, straight into renderTriangleFromData.
void giveProjectedTriangles(Frame &fr, const scene_desc *s) {
	const uint8_t *records = (const uint8_t *)s->projected;
	int32_t source = 0;
	for (int32_t part = 0; part < s->part_count; part++) {
		for (int32_t poly = s->part_polygon_start[part]; poly < s->part_polygon_start[part + 1]; poly++) {
			Projected p[3];
			for (int32_t c = 0; c < 3; c++) {
				const uint8_t *r = records + ((size_t)poly * 3 + c) * 40;
				float f[5];
				int64_t flat[2];
				memcpy(f, r, sizeof(f));
				memcpy(flat, r + 24, sizeof(flat));
				p[c].cs = V3{f[0], f[1], f[2]};
				p[c].is = V2{f[3], f[4]};
				p[c].flat = L2{flat[0], flat[1]};
			}
			const float *col = s->colors + poly * 16;
			const float *tc = s->texcoords + poly * 16;
			TriangleInput in;
			in.red = V3{col[0], col[4], col[8]}; in.green = V3{col[1], col[5], col[9]};
			in.blue = V3{col[2], col[6], col[10]}; in.alpha = V3{col[3], col[7], col[11]};
			in.u1 = V3{tc[0], tc[4], tc[8]}; in.v1 = V3{tc[1], tc[5], tc[9]};
			in.u2 = V3{tc[2], tc[6], tc[10]}; in.v2 = V3{tc[3], tc[7], tc[11]};
			in.diffuse = s->part_diffuse ? s->part_diffuse[part] : -1;
			in.light = s->part_light ? s->part_light[part] : -1;
			in.point[0] = poly * 3; in.point[1] = poly * 3 + 1; in.point[2] = poly * 3 + 2;
			giveTriangle(fr, p[0], p[1], p[2], in, source++);
		}
	}
}

// The scene's occluder calls that come before instance `before` (before == instance_count: after
// the last one), in order.
void applyOccluders(const scene_desc *s, Frame &fr, int32_t before, int32_t &query) {
	for (int32_t k = 0; k < s->occluder_count; k++) {
		const float *op = s->occluders + (size_t)k * 20;
		if ((int32_t)op[1] != before) continue;
		V3 mn{op[2], op[3], op[4]}, mx{op[5], op[6], op[7]};
		Transform t = transformFrom(op + 8);
		switch ((int32_t)op[0]) {
		case 1: occludeFromBox(fr, mn, mx, t); break;
		case 2: occludeFromTopRows(fr); break;
		case 3: occludeFromExistingTriangles(fr); break;
		case 4:
			if (s->visible_out) s->visible_out[query] = isBoxOccluded(fr, mn, mx, t) ? 0 : 1;
			query++;
			break;
		default: break;
		}
	}
}

void giveAll(const scene_desc *s, Frame &fr) {
	fr.commands.clear();
	fr.inputs.clear();
	fr.grid.clear();
	fr.occluded = false;
	fr.gridW = (fr.canvas.width + (cellSize - 1)) / cellSize;
	fr.gridH = (fr.canvas.height + (cellSize - 1)) / cellSize;
	if (s->projected) {
		giveProjectedTriangles(fr, s);
		return;
	}
	int32_t source = 0, query = 0;
	for (int32_t i = 0; i < s->instance_count; i++) {
		applyOccluders(s, fr, i, query);
		giveModel(fr, s, transformFrom(s->instance_transforms + i * 12), i, source);
	}
	applyOccluders(s, fr, s->instance_count, query);
	completeOcclusion(fr);
}

// CommandQueue::execute (renderCore.cpp:449-480): 12 slices, boundaries on even rows except
// the last; every slice walks every command in order.
void executeAll(const Frame &fr, int32_t threads) {
	Rect full = rectFromSides(0, 0, fr.canvas.width, fr.canvas.height);
	if (threads <= 1) {
		for (size_t i = 0; i < fr.commands.size(); i++) executeCommand(fr, fr.commands[i], (int32_t)i, full);
		return;
	}
	const int32_t jobCount = 12;
	Rect slices[jobCount];
	int32_t y1 = full.t;
	for (int32_t j = 0; j < jobCount; j++) {
		int32_t y2 = full.t + ((full.bottom() * (j + 1)) / jobCount);
		if (j < jobCount - 1) y2 = (y2 / 2) * 2;
		slices[j] = rectFromSides(full.l, y1, full.right(), y2);
		y1 = y2;
	}
	int32_t workers = std::min(threads, jobCount);
	std::vector<std::thread> pool;
	for (int32_t w = 0; w < workers; w++) {
		pool.emplace_back([&, w]() {
			for (int32_t j = w; j < jobCount; j += workers)
				for (size_t i = 0; i < fr.commands.size(); i++) executeCommand(fr, fr.commands[i], (int32_t)i, slices[j]);
		});
	}
	for (auto &t : pool) t.join();
}

// executeTriangleDrawingDepth + drawTriangleDepth (renderCore.cpp:344-394): rows of the aligned
// bound over the whole canvas, one depth value per row start, then one add of the x step per
// pixel; perspective keeps the greater depth, orthogonal the smaller.  Rows at or below the
// canvas height are skipped: the reference's aligned bound rounds the bottom up to an even row,
// which on an odd-height canvas lies outside the image.
void drawTriangleDepth(const Camera &cam, float *depth, int32_t width, int32_t height, const Tri &t) {
	if (!frontFacing(t)) return;
	Rect clip = rectFromSides(0, 0, width, height);
	if (!overlaps(t.wholeBound, clip)) return;
	Rect bound = alignedRasterBound(t, clip);
	if (bound.h <= 0) return;
	std::vector<Row> rows((size_t)bound.h);
	rasterize(t.p[0].flat, t.p[1].flat, t.p[2].flat, rows.data(), bound);
	Projection proj = getProjection(t, V3{0.0f, 0.0f, 0.0f}, V3{0.0f, 0.0f, 0.0f}, cam.perspective);
	for (int32_t y = bound.t; y < bound.t + bound.h && y < height; y++) {
		Row r = rows[(size_t)(y - bound.t)];
		float value = weightAtPixel(proj, r.left, y).x;
		float *row = depth + (size_t)y * (size_t)width;
		for (int32_t x = r.left; x < r.right; x++) {
			if (proj.affine ? (value < row[x]) : (value > row[x])) row[x] = value;
			value += proj.dx.x;
		}
	}
}

// renderTriangleFromDataDepth (renderCore.cpp:405-443): cull, draw whole or clip into a fan.
void giveTriangleDepth(const Camera &cam, float *depth, int32_t width, int32_t height, const Projected &a,
                       const Projected &b, const Projected &c) {
	Tri tri = makeTri(a, b, c);
	if (triangleVisibility(tri, cam.cull) == Hidden) return;
	if (triangleVisibility(tri, cam.clip) == Full) {
		if (frontFacing(tri)) drawTriangleDepth(cam, depth, width, height, tri);
		return;
	}
	Polygon9 poly;
	poly.v[0].cs = tri.p[0].cs; poly.v[0].subB = 0.0f; poly.v[0].subC = 0.0f;
	poly.v[1].cs = tri.p[1].cs; poly.v[1].subB = 1.0f; poly.v[1].subC = 0.0f;
	poly.v[2].cs = tri.p[2].cs; poly.v[2].subB = 0.0f; poly.v[2].subC = 1.0f;
	poly.count = 3;
	for (int32_t s = 0; s < cam.clip.count; s++) poly.clip(cam.clip.planes[s]);
	for (int32_t i = 0; i < poly.count - 2; i++) {
		Tri sub = makeTri(cameraToScreen(cam, poly.v[0].cs), cameraToScreen(cam, poly.v[1 + i].cs), cameraToScreen(cam, poly.v[2 + i].cs));
		drawTriangleDepth(cam, depth, width, height, sub);
	}
}

// ModelImpl::renderDepth + Part::renderDepth (Model.cpp:164-210): box cull, project every point,
// then each polygon as the fan (0,1,2), (0,2,3) in polygon order.
void renderModelDepth(const scene_desc *s, const Camera &cam, const Transform &modelToWorld, float *depth) {
	V3 mn{0.0f, 0.0f, 0.0f}, mx{0.0f, 0.0f, 0.0f};
	for (int32_t i = 0; i < s->point_count; i++) {
		const float *p = s->points + i * 3;
		mn.x = std::min(mn.x, p[0]); mn.y = std::min(mn.y, p[1]); mn.z = std::min(mn.z, p[2]);
		mx.x = std::max(mx.x, p[0]); mx.y = std::max(mx.y, p[1]); mx.z = std::max(mx.z, p[2]);
	}
	if (boxSeen(cam, mn, mx, modelToWorld) == 0) return;
	std::vector<Projected> proj((size_t)s->point_count);
	for (int32_t i = 0; i < s->point_count; i++) {
		const float *p = s->points + i * 3;
		proj[(size_t)i] = cameraToScreen(cam, transformPointTransposedInverse(cam.location, transformPoint(modelToWorld, V3{p[0], p[1], p[2]})));
	}
	for (int32_t part = 0; part < s->part_count; part++) {
		for (int32_t poly = s->part_polygon_start[part]; poly < s->part_polygon_start[part + 1]; poly++) {
			const int32_t *idx = s->indices + poly * 4;
			int32_t fan = (idx[3] >= 0) ? 2 : 1;
			for (int32_t k = 0; k < fan; k++)
				giveTriangleDepth(cam, depth, s->width, s->height, proj[(size_t)idx[0]], proj[(size_t)idx[1 + k]], proj[(size_t)idx[2 + k]]);
		}
	}
}

}  // namespace

extern "C" {

int32_t oracle_render_depth_scene(const scene_desc *s, float *outDepth) {
	if (s->width < 1 || s->height < 1 || !outDepth) return -1;
	Camera cam = makeCamera(s);
	size_t n = (size_t)s->width * (size_t)s->height;
	if (s->initial_depth) memcpy(outDepth, s->initial_depth, n * 4); else memset(outDepth, 0, n * 4);
	for (int32_t i = 0; i < s->instance_count; i++)
		renderModelDepth(s, cam, transformFrom(s->instance_transforms + i * 12), outDepth);
	return 0;
}

int32_t oracle_render_scene(const scene_desc *s, int32_t threads, uint32_t *outColor, float *outDepth) {
	if (s->width < 1 || s->height < 1) return -1;
	Frame fr;
	std::vector<uint32_t> color;
	std::vector<float> depth;
	buildFrame(s, fr, color, depth);
	loadCanvas(s, fr);
	giveAll(s, fr);
	executeAll(fr, threads);
	size_t n = (size_t)s->width * (size_t)s->height;
	if (outColor && fr.canvas.color) memcpy(outColor, color.data(), n * 4);
	if (outDepth && fr.canvas.depth) memcpy(outDepth, depth.data(), n * 4);
	return 0;
}

int32_t oracle_time_scene(const scene_desc *s, int32_t threads, int32_t warmup, int32_t iterations, double *outSeconds) {
	if (s->width < 1 || s->height < 1) return -1;
	Frame fr;
	std::vector<uint32_t> color;
	std::vector<float> depth;
	buildFrame(s, fr, color, depth);
	std::vector<double> times;
	for (int32_t i = 0; i < warmup + iterations; i++) {
		loadCanvas(s, fr);
		auto t0 = std::chrono::steady_clock::now();
		giveAll(s, fr);
		executeAll(fr, threads);
		auto t1 = std::chrono::steady_clock::now();
		if (i >= warmup) times.push_back(std::chrono::duration<double>(t1 - t0).count());
	}
	std::sort(times.begin(), times.end());
	*outSeconds = times.empty() ? 0.0 : times[times.size() / 2];
	return 0;
}

uint32_t oracle_texture_query(int32_t log2w, int32_t log2h, int32_t maxMip, int32_t writeCount, const uint32_t *writes,
                              int32_t op, uint32_t x, uint32_t y, uint32_t mip, float u, float v) {
	Texture t = layoutTexture((uint32_t)log2w, (uint32_t)log2h, (uint32_t)maxMip);
	for (int32_t i = 0; i < writeCount; i++) {
		const uint32_t *wr = writes + i * 4;
		t.buffer[pixelOffset(t, wr[0], wr[1], wr[2])] = wr[3];
	}
	switch (op) {
	case 0: return t.startOffset;
	case 1: return t.maxLevelMask;
	case 2: return pixelOffsetToLayer(t, mip);
	case 3: return pixelOffset(t, x, y, mip);
	case 4: return t.buffer[pixelOffset(t, x, y, mip)];
	case 5: return sampleNearest(t, u, v, mip);
	case 6: return t.maxMip;
	default: return 0xFFFFFFFFu;
	}
}

uint32_t oracle_mix_colors(uint32_t a, uint32_t b, uint32_t weight) { return mixColors(a, b, weight); }

int32_t oracle_diff_forward(const scene_desc *s, const float *texture, int32_t log2w, int32_t log2h, const float *multipliers,
                            float *outColor, int32_t *outVisibility, float *outDepth, float *outWeights) {
	if (s->width < 1 || s->height < 1 || !s->has_depth) return -1;
	size_t n = (size_t)s->width * (size_t)s->height;
	DiffTarget d;
	d.texture = texture; d.log2w = log2w; d.log2h = log2h; d.multipliers = multipliers;
	d.color = outColor; d.visibility = outVisibility; d.weights = outWeights;
	for (size_t i = 0; i < n; i++) {
		outVisibility[i] = -1;
		for (int ch = 0; ch < 4; ch++) outColor[i * 4 + ch] = 0.0f;
		for (int k = 0; k < 5; k++) outWeights[i * 5 + k] = 0.0f;
	}
	Frame fr;
	std::vector<uint32_t> color;
	std::vector<float> depth;
	buildFrame(s, fr, color, depth);
	fr.canvas.color = nullptr;
	fr.diff = &d;
	loadCanvas(s, fr);
	giveAll(s, fr);
	executeAll(fr, 1);
	memcpy(outDepth, depth.data(), n * 4);
	return (int32_t)fr.commands.size();
}

// Adjoint of oracle_diff_forward in double precision, from the recorded weights of each pixel.
void oracle_diff_backward(const scene_desc *s, const float *texture, int32_t log2w, int32_t log2h, const float *multipliers,
                          const int32_t *visibility, const float *weights, const float *gradColor, double *gradTexture,
                          double *gradMultipliers) {
	Frame fr;
	std::vector<uint32_t> color;
	std::vector<float> depth;
	buildFrame(s, fr, color, depth);
	DiffTarget d;
	d.texture = texture; d.log2w = log2w; d.log2h = log2h; d.multipliers = multipliers;
	fr.diff = &d;
	giveAll(s, fr);
	size_t n = (size_t)s->width * (size_t)s->height;
	for (size_t p = 0; p < n; p++) {
		int32_t id = visibility[p];
		if (id < 0) continue;
		const TriangleInput &in = fr.inputs[fr.commands[(size_t)(id >> 2)].input];
		double w[3] = {weights[p * 5], weights[p * 5 + 1], weights[p * 5 + 2]};
		Taps taps = bilinearTaps(log2w, log2h, weights[p * 5 + 3], weights[p * 5 + 4]);
		double t[4] = {0, 0, 0, 0};
		for (int k = 0; k < 4; k++)
			for (int ch = 0; ch < 4; ch++) t[ch] += (double)taps.weight[k] * texture[(size_t)taps.index[k] * 4 + ch];
		double mixed[3];
		for (int ch = 0; ch < 3; ch++) {
			mixed[ch] = 0.0;
			for (int c = 0; c < 3; c++) mixed[ch] += w[c] * multiplierOf(d, in.point[c], ch);
		}
		const float *g = gradColor + p * 4;
		if (gradTexture) {
			for (int k = 0; k < 4; k++) {
				double *dst = gradTexture + (size_t)taps.index[k] * 4;
				for (int ch = 0; ch < 3; ch++) dst[ch] += (double)g[ch] * mixed[ch] * taps.weight[k];
				dst[3] += (double)g[3] * taps.weight[k];
			}
		}
		if (gradMultipliers) {
			for (int c = 0; c < 3; c++)
				for (int ch = 0; ch < 3; ch++) gradMultipliers[(size_t)in.point[c] * 3 + ch] += (double)g[ch] * t[ch] * w[c];
		}
	}
}

uint32_t oracle_pack_saturated(float r, float g, float b, float a) { return packSaturated(r, g, b, a); }

void oracle_reciprocal(int32_t n, const float *in, float *out) {
	for (int32_t i = 0; i < n; i++) out[i] = reciprocalLane(in[i]);
}

void oracle_rasterize(const int64_t *corners, int32_t left, int32_t top, int32_t width, int32_t height, int32_t *outRows) {
	std::vector<Row> rows((size_t)height);
	Rect clip; clip.l = left; clip.t = top; clip.w = width; clip.h = height;
	rasterize(L2{corners[0], corners[1]}, L2{corners[2], corners[3]}, L2{corners[4], corners[5]}, rows.data(), clip);
	for (int32_t r = 0; r < height; r++) { outRows[r * 2] = rows[(size_t)r].left; outRows[r * 2 + 1] = rows[(size_t)r].right; }
}

int32_t oracle_texture_pyramid(int32_t w, int32_t h, int32_t res, const uint32_t *level0, uint32_t *outBuffer, int32_t capacity,
                               int32_t *outLog2w, int32_t *outLog2h, int32_t *outMaxMip) {
	Texture t = makeTexture(w, h, res, level0);
	if ((int64_t)t.buffer.size() > capacity) return -1;
	memcpy(outBuffer, t.buffer.data(), t.buffer.size() * 4);
	*outLog2w = (int32_t)t.log2w; *outLog2h = (int32_t)t.log2h; *outMaxMip = (int32_t)t.maxMip;
	return (int32_t)t.buffer.size();
}

int32_t oracle_sample_quads(int32_t w, int32_t h, int32_t res, const uint32_t *level0, int32_t quadCount,
                            const float *u, const float *v, uint32_t *outColor, uint32_t *outMip) {
	Texture t = makeTexture(w, h, res, level0);
	for (int32_t q = 0; q < quadCount; q++) {
		uint32_t mip = mipLevelIndex(t, u + q * 4, v + q * 4);
		for (int l = 0; l < 4; l++) outColor[q * 4 + l] = sampleBilinear(t, u[q * 4 + l], v[q * 4 + l], mip, false);
		outMip[q] = mip;
	}
	return 0;
}

void oracle_project(const scene_desc *s, const float *transform, int32_t n, const float *points, float *outCs, float *outIs, int64_t *outFlat) {
	Camera c = makeCamera(s);
	Transform m = transformFrom(transform);
	for (int32_t i = 0; i < n; i++) {
		Projected p = cameraToScreen(c, transformPointTransposedInverse(c.location, transformPoint(m, V3{points[i * 3], points[i * 3 + 1], points[i * 3 + 2]})));
		outCs[i * 3] = p.cs.x; outCs[i * 3 + 1] = p.cs.y; outCs[i * 3 + 2] = p.cs.z;
		outIs[i * 2] = p.is.x; outIs[i * 2 + 1] = p.is.y;
		outFlat[i * 2] = p.flat.x; outFlat[i * 2 + 1] = p.flat.y;
	}
}

int32_t oracle_occlusion_grid(const scene_desc *s, float *outGrid, int32_t *outWidth, int32_t *outHeight) {
	Frame fr;
	std::vector<uint32_t> color;
	std::vector<float> depth;
	buildFrame(s, fr, color, depth);
	loadCanvas(s, fr);
	giveAll(s, fr);
	*outWidth = fr.gridW;
	*outHeight = fr.gridH;
	size_t n = (size_t)fr.gridW * (size_t)fr.gridH;
	for (size_t i = 0; i < n; i++) outGrid[i] = fr.occluded ? fr.grid[i] : 0.0f;
	return fr.occluded ? 1 : 0;
}

int32_t oracle_commands(const scene_desc *s, int32_t capacity, int32_t *outTriangle, int64_t *outFlat, int32_t *outBound) {
	Frame fr;
	std::vector<uint32_t> color;
	std::vector<float> depth;
	buildFrame(s, fr, color, depth);
	loadCanvas(s, fr);
	giveAll(s, fr);
	int32_t n = (int32_t)fr.commands.size();
	if (n > capacity) return n;
	for (int32_t i = 0; i < n; i++) {
		const Command &c = fr.commands[(size_t)i];
		outTriangle[i] = c.sourceTriangle;
		for (int k = 0; k < 3; k++) { outFlat[i * 6 + k * 2] = c.tri.p[k].flat.x; outFlat[i * 6 + k * 2 + 1] = c.tri.p[k].flat.y; }
		outBound[i * 4] = c.tri.wholeBound.l; outBound[i * 4 + 1] = c.tri.wholeBound.t;
		outBound[i * 4 + 2] = c.tri.wholeBound.right(); outBound[i * 4 + 3] = c.tri.wholeBound.bottom();
	}
	return n;
}

}
