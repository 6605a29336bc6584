// Triangle setup arithmetic of the polygon path, shared by the host side of the library
// (camera derivation, model box culling) and the gfx950 kernels (projection, culling,
// clipping, weight planes, exact row intervals).
//
// Every expression keeps the reference's order of operations: the library is compiled
// with -ffp-contract=off and without fast-math, so each float result here rounds exactly
// as in the reference renderer.  File:line citations are relative to the reference's
// Source/ folder.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DH __host__ __device__ __forceinline__

namespace dfpsr {

struct f2 { float x, y; };
struct f3 { float x, y, z; };
struct l2 { int64_t x, y; };

DH f3 make3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
DH f3 add3(const f3 &a, const f3 &b) { return make3(a.x + b.x, a.y + b.y, a.z + b.z); }
DH f3 scale3(const f3 &a, float s) { return make3(a.x * s, a.y * s, a.z * s); }
DH float comp(const f3 &v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
DH void setComp(f3 &v, int i, float value) { if (i == 0) v.x = value; else if (i == 1) v.y = value; else v.z = value; }

template <typename T> DH T tmin(T a, T b) { return a < b ? a : b; }
template <typename T> DH T tmax(T a, T b) { return a > b ? a : b; }

// Integer rectangle as IRect (math/IRect.h:40-78): left, top, width, height.
struct Rect {
	int32_t l, t, w, h;
	DH int32_t right() const { return l + w; }
	DH int32_t bottom() const { return t + h; }
};
DH Rect rectSides(int32_t l, int32_t t, int32_t r, int32_t b) { Rect x; x.l = l; x.t = t; x.w = r - l; x.h = b - t; return x; }
DH bool overlaps(const Rect &a, const Rect &b) {
	return a.l < b.right() && a.right() > b.l && a.t < b.bottom() && a.bottom() > b.t;
}
DH Rect cutRect(const Rect &a, const Rect &b) {
	if (!overlaps(a, b)) return rectSides(0, 0, 0, 0);
	return rectSides(tmax(a.l, b.l), tmax(a.t, b.t), tmin(a.right(), b.right()), tmin(a.bottom(), b.bottom()));
}
// signedModulo, roundUp, roundDown (implementation/math/scalar.h:30-49)
DH int32_t signedMod(int32_t a, int32_t b) { return a >= 0 ? a % b : (b - (-a % b)) % b; }
DH int32_t roundUpTo(int32_t size, int32_t alignment) { return size + (alignment - 1) - signedMod(size - 1, alignment); }
DH int32_t roundDownTo(int32_t size, int32_t alignment) { return size - signedMod(size, alignment); }

// ---------------------------------------------------------------------------------------
// Camera (implementation/render/Camera.h)
// ---------------------------------------------------------------------------------------

struct Plane { f3 n; float offset; };

// FPlane3D keeps a normalized normal (math/FPlane3D.h:36; normalize in math/FVector.h:113-119).
__host__ inline Plane makePlane(f3 n, float offset) {
	float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
	Plane p;
	p.n = (len == 0.0f) ? make3(0.0f, 0.0f, 1.0f) : make3(n.x / len, n.y / len, n.z / len);
	p.offset = offset;
	return p;
}
// FPlane3D::signedDistance and inside (FPlane3D.h:39-44)
DH float signedDistance(const Plane &p, const f3 &v) { return ((p.n.x * v.x) + (p.n.y * v.y)) + (p.n.z * v.z) - p.offset; }
DH bool insidePlane(const Plane &p, const f3 &v) { return signedDistance(p, v) <= 0.0f; }

struct Frustum { Plane planes[6]; int32_t count; };

// Transform3D: position then the x, y, z axes (math/Transform3D.h:30-31).
struct Transform { f3 pos, xAxis, yAxis, zAxis; };

DH Transform transformFrom(const float *t) {
	Transform m;
	m.pos = make3(t[0], t[1], t[2]);
	m.xAxis = make3(t[3], t[4], t[5]);
	m.yAxis = make3(t[6], t[7], t[8]);
	m.zAxis = make3(t[9], t[10], t[11]);
	return m;
}
// Transform3D::transformPoint (Transform3D.h:41-43 with FMatrix3x3::transform, FMatrix3x3.h:55-61)
DH f3 transformPoint(const Transform &m, const f3 &p) {
	return make3(((p.x * m.xAxis.x + p.y * m.yAxis.x) + p.z * m.zAxis.x) + m.pos.x,
	             ((p.x * m.xAxis.y + p.y * m.yAxis.y) + p.z * m.zAxis.y) + m.pos.y,
	             ((p.x * m.xAxis.z + p.y * m.yAxis.z) + p.z * m.zAxis.z) + m.pos.z);
}
// Transform3D::transformPointTransposedInverse (Transform3D.h:51-53, FMatrix3x3.h:66-72)
DH f3 worldToCamera(const Transform &m, const f3 &p) {
	f3 d = make3(p.x - m.pos.x, p.y - m.pos.y, p.z - m.pos.z);
	return make3((d.x * m.xAxis.x + d.y * m.xAxis.y) + d.z * m.xAxis.z,
	             (d.x * m.yAxis.x + d.y * m.yAxis.y) + d.z * m.yAxis.z,
	             (d.x * m.zAxis.x + d.y * m.zAxis.y) + d.z * m.zAxis.z);
}

// Everything the device needs of a Camera (Camera.h:125-142), derived on the host.
struct CameraParams {
	Transform location;
	float invWidthSlope, invHeightSlope, imageWidth, imageHeight;
	int32_t perspective;
	Frustum cull, clip;
};

__host__ inline Frustum perspectiveFrustum(float nearClip, float farClip, float ws, float hs) {  // Camera.h:64-74
	Frustum f;
	f.count = (farClip == __builtin_huge_valf()) ? 5 : 6;
	f.planes[0] = makePlane(make3(-1.0f, 0.0f, -ws), 0.0f);
	f.planes[1] = makePlane(make3(1.0f, 0.0f, -ws), 0.0f);
	f.planes[2] = makePlane(make3(0.0f, 1.0f, -hs), 0.0f);
	f.planes[3] = makePlane(make3(0.0f, -1.0f, -hs), 0.0f);
	f.planes[4] = makePlane(make3(0.0f, 0.0f, -1.0f), -nearClip);
	f.planes[5] = makePlane(make3(0.0f, 0.0f, 1.0f), farClip);
	return f;
}
__host__ inline Frustum orthogonalFrustum(float halfWidth, float halfHeight) {  // Camera.h:56-62
	Frustum f;
	f.count = 4;
	f.planes[0] = makePlane(make3(-1.0f, 0.0f, 0.0f), halfWidth);
	f.planes[1] = makePlane(make3(1.0f, 0.0f, 0.0f), halfWidth);
	f.planes[2] = makePlane(make3(0.0f, 1.0f, 0.0f), halfHeight);
	f.planes[3] = makePlane(make3(0.0f, -1.0f, 0.0f), halfHeight);
	f.planes[4] = f.planes[5] = makePlane(make3(0.0f, 0.0f, 1.0f), 0.0f);
	return f;
}

// Camera::createPerspective / createOrthogonal (Camera.h:144-156) with cullRatio 1.0001 and
// clipRatio 2 (Camera.h:113-118).  For an orthogonal camera widthSlope is the half width.
__host__ inline CameraParams deriveCamera(int32_t perspective, const float *location, float imageWidth, float imageHeight,
                                          float widthSlope, float nearClip, float farClip) {
	CameraParams c;
	c.location = transformFrom(location);
	c.imageWidth = imageWidth;
	c.imageHeight = imageHeight;
	c.perspective = perspective;
	const float cullRatio = 1.0001f, clipRatio = 2.0f;
	float heightSlope = widthSlope * imageHeight / imageWidth;
	if (perspective) {
		c.cull = perspectiveFrustum(nearClip, farClip, widthSlope * cullRatio, heightSlope * cullRatio);
		c.clip = perspectiveFrustum(nearClip, farClip, widthSlope * clipRatio, heightSlope * clipRatio);
	} else {
		c.cull = orthogonalFrustum(widthSlope * cullRatio, heightSlope * cullRatio);
		c.clip = orthogonalFrustum(widthSlope * clipRatio, heightSlope * clipRatio);
	}
	c.invWidthSlope = 0.5f / widthSlope;
	c.invHeightSlope = 0.5f / heightSlope;
	return c;
}

// Projected corner: ProjectedPoint (implementation/render/ProjectedPoint.h:33-49).
struct Projected { f3 cs; f2 is; l2 flat; };

// float -> int64 as cvttss2si with a 64-bit destination: out of range and NaN give INT64_MIN.
DH int64_t truncI64(float f) {
	if (!(f > -9223372036854775808.0f && f < 9223372036854775808.0f)) return (int64_t)0x8000000000000000ull;
	return (int64_t)f;
}

// Camera::cameraToScreen (Camera.h:160-187)
DH Projected cameraToScreen(const CameraParams &c, const f3 &cs) {
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
DH int32_t convexHullSeen(const Frustum &f, const f3 *pts, int32_t n) {
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
DH int32_t boxSeen(const CameraParams &c, const f3 &mn, const f3 &mx, const Transform &modelToWorld) {
	f3 cs[8];
	for (int i = 0; i < 8; i++) {
		f3 corner = make3((i & 1) ? mx.x : mn.x, (i & 2) ? mx.y : mn.y, (i & 4) ? mx.z : mn.z);
		cs[i] = worldToCamera(c.location, transformPoint(modelToWorld, corner));
	}
	return convexHullSeen(c.cull, cs, 8);
}

// ---------------------------------------------------------------------------------------
// Triangles (implementation/render/ITriangle2D.cpp, renderCore.cpp)
// ---------------------------------------------------------------------------------------

// getTriangleBound (ITriangle2D.cpp:31-43): nearest pixel of each corner, widened by one.
DH Rect triangleBound(const l2 &a, const l2 &b, const l2 &c) {
	int32_t x1 = (int32_t)((a.x + 128) / 256), y1 = (int32_t)((a.y + 128) / 256);
	int32_t x2 = (int32_t)((b.x + 128) / 256), y2 = (int32_t)((b.y + 128) / 256);
	int32_t x3 = (int32_t)((c.x + 128) / 256), y3 = (int32_t)((c.y + 128) / 256);
	return rectSides(tmin(tmin(x1, x2), x3) - 1, tmin(tmin(y1, y2), y3) - 1, tmax(tmax(x1, x2), x3) + 1, tmax(tmax(y1, y2), y3) + 1);
}

// ITriangle2D::isFrontfacing (ITriangle2D.cpp:55-60): clockwise on screen.
DH bool frontFacing(const l2 &a, const l2 &b, const l2 &c) {
	return ((c.x - a.x) * (b.y - a.y)) + ((c.y - a.y) * (a.x - b.x)) < 0;
}

enum Visibility { VIS_HIDDEN = 0, VIS_FULL = 1, VIS_PARTIAL = 2 };

// getTriangleVisibility (renderCore.cpp:172-198)
DH int32_t triangleVisibility(const f3 &a, const f3 &b, const f3 &c, const Frustum &f) {
	uint32_t outA = 0, outB = 0, outC = 0;
	for (int32_t s = 0; s < f.count; s++) {
		if (!insidePlane(f.planes[s], a)) outA |= 1u << s;
		if (!insidePlane(f.planes[s], b)) outB |= 1u << s;
		if (!insidePlane(f.planes[s], c)) outC |= 1u << s;
	}
	if (outA & outB & outC) return VIS_HIDDEN;
	if (outA | outB | outC) return VIS_PARTIAL;
	return VIS_FULL;
}

// SubVertex (renderCore.cpp:33-48): camera-space position and the weights of the parent's
// second and third corners.  state/value are per-plane scratch of the clipper.
struct SubVertex { f3 cs; float subB, subC; int32_t state; float value; };

DH SubVertex lerpVertex(const SubVertex &a, const SubVertex &b, float ratio) {
	SubVertex r;
	float inv = 1.0f - ratio;
	r.cs = make3(a.cs.x * inv + b.cs.x * ratio, a.cs.y * inv + b.cs.y * ratio, a.cs.z * inv + b.cs.z * ratio);
	r.subB = a.subB * inv + b.subB * ratio;
	r.subC = a.subC * inv + b.subC * ratio;
	r.state = 0;
	r.value = 0.0f;
	return r;
}
// inverseLerp (renderCore.cpp:52-59)
DH float inverseLerp(float a, float b, float value) {
	float c = b - a;
	return (c == 0.0f) ? 0.5f : (value - a) / c;
}

// The clipped polygon of at most 9 sub-vertices (renderCore.cpp:62-170).  The insert,
// replace and delete steps decide the order of the vertices and therefore the fan, so the
// same steps are taken here.
struct ClipPolygon {
	SubVertex v[9];
	int32_t count;

	DH void init(const f3 &a, const f3 &b, const f3 &c) {
		v[0].cs = a; v[0].subB = 0.0f; v[0].subC = 0.0f; v[0].state = 0; v[0].value = 0.0f;
		v[1].cs = b; v[1].subB = 1.0f; v[1].subC = 0.0f; v[1].state = 0; v[1].value = 0.0f;
		v[2].cs = c; v[2].subB = 0.0f; v[2].subC = 1.0f; v[2].state = 0; v[2].value = 0.0f;
		count = 3;
	}
	DH void removeAt(int32_t i) {
		if (count > 0 && count <= 9) {
			for (int32_t k = i; k < count - 1; k++) v[k] = v[k + 1];
			count--;
		}
	}
	DH void insertAt(int32_t i, const SubVertex &x) {
		if (count < 9) {
			for (int32_t k = count - 1; k >= i; k--) v[k + 1] = v[k];
			v[i] = x;
			count++;
		}
	}
	DH void clip(const Plane &plane) {
		const int32_t USE = 0, DEL = 1, MODIFIED = 2;
		if (!(count >= 3 && count < 9)) return;
		int32_t outsideCount = 0, lastOutside = 0;
		for (int32_t k = 0; k < count; k++) {
			float d = signedDistance(plane, v[k].cs);
			v[k].value = d;
			if (d > 0.0f) { outsideCount++; lastOutside = k; v[k].state = DEL; } else { v[k].state = USE; }
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
			insertAt(next, cutEnd);
		} else {
			for (int32_t cur = 0; cur < count; cur++) {
				int32_t prev = (cur - 1 + count) % count;
				int32_t next = (cur + 1) % count;
				if (v[cur].state == DEL) {
					if (v[prev].state == USE) {
						float r = inverseLerp(v[prev].value, v[cur].value, 0.0f);
						v[cur] = lerpVertex(v[prev], v[cur], r);
						v[cur].state = MODIFIED;
					} else if (v[next].state == USE) {
						float r = inverseLerp(v[cur].value, v[next].value, 0.0f);
						v[cur] = lerpVertex(v[cur], v[next], r);
						v[cur].state = MODIFIED;
					}
				}
			}
			if (outsideCount > 2) {
				for (int32_t k = count - 1; k >= 0; k--) {
					if (v[k].state == DEL) removeAt(k);
				}
			}
		}
	}
};

// ---------------------------------------------------------------------------------------
// Weight planes: ITriangle2D::getProjection (ITriangle2D.cpp:182-300)
// ---------------------------------------------------------------------------------------

struct Projection { f3 start, dx, dy; };

DH Projection weightPlanes(const f2 *is, const f3 &csZ, const f3 &subB, const f3 &subC, bool perspective) {
	f3 offX, offY;
	for (int i = 0; i < 3; i++) {
		int j = (i + 1) % 3;
		setComp(offX, i, is[j].y - is[i].y);
		setComp(offY, i, is[i].x - is[j].x);
	}
	f3 mult;
	for (int i = 0; i < 3; i++) {
		int o = (i + 2) % 3;
		float other = ((is[o].x - is[i].x) * comp(offX, i)) + ((is[o].y - is[i].y) * comp(offY, i));
		setComp(mult, o, (other == 0.0f) ? 0.0f : 1.0f / other);
	}
	f3 nX, nY;
	for (int i = 0; i < 3; i++) {
		setComp(nX, i, comp(offX, i) * comp(mult, i));
		setComp(nY, i, comp(offY, i) * comp(mult, i));
	}
	f3 tw;
	for (int i = 0; i < 3; i++) {
		int o = (i + 2) % 3;
		setComp(tw, o, is[i].x * -comp(nX, i) + is[i].y * -comp(nY, i));
	}
	f3 adx = make3(nX.y, nX.z, nX.x);
	f3 ady = make3(nY.y, nY.z, nY.x);
	Projection p;
	if (!perspective) {
		f3 W = csZ;
		p.start = make3(W.x * tw.x + W.y * tw.y + W.z * tw.z, tw.x * subB.x + tw.y * subB.y + tw.z * subB.z,
		                tw.x * subC.x + tw.y * subC.y + tw.z * subC.z);
		p.dx = make3(W.x * adx.x + W.y * adx.y + W.z * adx.z, adx.x * subB.x + adx.y * subB.y + adx.z * subB.z,
		             adx.x * subC.x + adx.y * subC.y + adx.z * subC.z);
		p.dy = make3(W.x * ady.x + W.y * ady.y + W.z * ady.z, ady.x * subB.x + ady.y * subB.y + ady.z * subB.z,
		             ady.x * subC.x + ady.y * subC.y + ady.z * subC.z);
	} else {
		f3 IW = make3(1.0f / csZ.x, 1.0f / csZ.y, 1.0f / csZ.z);
		p.start = make3(IW.x * tw.x + IW.y * tw.y + IW.z * tw.z,
		                IW.x * tw.x * subB.x + IW.y * tw.y * subB.y + IW.z * tw.z * subB.z,
		                IW.x * tw.x * subC.x + IW.y * tw.y * subC.y + IW.z * tw.z * subC.z);
		p.dx = make3(IW.x * adx.x + IW.y * adx.y + IW.z * adx.z,
		             IW.x * adx.x * subB.x + IW.y * adx.y * subB.y + IW.z * adx.z * subB.z,
		             IW.x * adx.x * subC.x + IW.y * adx.y * subC.y + IW.z * adx.z * subC.z);
		p.dy = make3(IW.x * ady.x + IW.y * ady.y + IW.z * ady.z,
		             IW.x * ady.x * subB.x + IW.y * ady.y * subB.y + IW.z * ady.z * subB.z,
		             IW.x * ady.x * subC.x + IW.y * ady.y * subC.y + IW.z * ady.z * subC.z);
	}
	return p;
}

// Weights at the centre of pixel (x, y): Projection::getDepthDividedWeight_perspective /
// getWeight_affine (ITriangle2D.h:80-100).
DH f3 weightAtPixel(const Projection &p, int32_t x, int32_t y) {
	float fx = (float)x + 0.5f, fy = (float)y + 0.5f;
	return make3((p.start.x + p.dx.x * fx) + p.dy.x * fy, (p.start.y + p.dx.y * fx) + p.dy.y * fy,
	             (p.start.z + p.dx.z * fx) + p.dy.z * fy);
}

// ---------------------------------------------------------------------------------------
// Exact row intervals (ITriangle2D.cpp:70-176)
// ---------------------------------------------------------------------------------------

// getAlignedRasterBound (ITriangle2D.cpp:70-75): rows rounded out to even numbers.
DH Rect alignedRasterBound(const Rect &wholeBound, const Rect &clip) {
	Rect u = cutRect(wholeBound, clip);
	return rectSides(u.l, roundDownTo(u.t, 2), u.right(), roundUpTo(u.bottom(), 2));
}

// One edge's cut of one row (cutConvexEdge, ITriangle2D.cpp:86-148).  The row value is a
// function of y alone: the clip's left/top terms cancel, so rows can be computed one by one.
struct EdgeSetup {
	int64_t offsetX, offsetY, limitTop, valueTop, threshold;
	int32_t kind;  // 0: no effect, 1: left edge, 2: right edge, 3: horizontal
};

DH EdgeSetup edgeSetup(const l2 &s, const l2 &e, const Rect &clip) {
	EdgeSetup st;
	const int32_t L = clip.l, T = clip.t;
	int64_t originX = 128 + (int64_t)L * 256;
	int64_t originY = 128 + (int64_t)T * 256;
	st.threshold = (s.x > e.x || (s.x == e.x && s.y > e.y)) ? -1 : 0;
	int64_t normalX = e.y - s.y;
	int64_t normalY = s.x - e.x;
	st.offsetX = normalX * 256;
	st.offsetY = normalY * 256;
	int64_t valueOrigin = ((originX - s.x) * normalX) + ((originY - s.y) * normalY);
	st.valueTop = valueOrigin;
	st.limitTop = st.threshold - valueOrigin + st.offsetX * L;
	st.kind = (normalX != 0) ? (normalX < 0 ? 1 : 2) : (normalY != 0 ? 3 : 0);
	return st;
}

struct RowInterval { int32_t left, right; };

// Truncating int64 division.  Both operands are exact in double when |n| < 2^52 (edge values
// of triangles inside the clip frustum stay far below that), so the double quotient is within
// one of the truncated quotient and one correction makes it exact; larger operands take the
// integer division.
DH int64_t divTrunc64(int64_t n, int64_t d) {
	const int64_t lim = (int64_t)1 << 52;
	if (n >= lim || n <= -lim || d >= lim || d <= -lim) return n / d;
	int64_t q = (int64_t)((double)n / (double)d);
	const int64_t dir = ((n < 0) == (d < 0)) ? 1 : -1;  // sign of the true quotient
	const int64_t ad = d < 0 ? -d : d;
	for (int k = 0; k < 2; k++) {
		const int64_t r = n - q * d;
		const int64_t ar = r < 0 ? -r : r;
		if (r != 0 && ((r < 0) != (n < 0))) q -= dir;       // one step too far from zero
		else if (ar >= ad) q += dir;                         // one step short
		else break;
	}
	return q;
}

DH bool sameCorner(const l2 &a, const l2 &b) { return a.x == b.x && a.y == b.y; }

// An edge as the rows read it: a left or right edge (kind 1, 2) keeps a = offsetX,
// b = offsetY, c = limitTop; a horizontal edge (kind 3) keeps a = threshold, b = offsetY,
// c = valueTop; an edge without effect (kind 0) keeps nothing.
struct EdgeRows { int64_t a, b, c; };

DH EdgeRows edgeRows(const EdgeSetup &st) {
	EdgeRows e;
	e.a = st.kind == 3 ? st.threshold : st.offsetX;
	e.b = st.offsetY;
	e.c = st.kind == 3 ? st.valueTop : st.limitTop;
	return e;
}

// What every row of a triangle needs (ITriangle2D.cpp:70-176): the three edges of
// rasterizeTriangle, the clipped columns and the aligned row bound.  Set up once per draw
// command; rows are then independent of each other.  flags: the edges' kinds in bits 0-1,
// 2-3 and 4-5, bit 6 = two equal corners.
struct __attribute__((aligned(16))) TriangleRows {
	EdgeRows edges[3];
	int32_t clipL, clipR, alT, alB;
	int32_t flags;
};

DH TriangleRows triangleRows(const l2 &A, const l2 &B, const l2 &C, const Rect &canvas) {
	const Rect wb = triangleBound(A, B, C);
	const Rect clipR = cutRect(wb, canvas);
	const Rect al = alignedRasterBound(wb, canvas);
	const Rect rowClip = rectSides(clipR.l, al.t, clipR.right(), al.bottom());
	const EdgeSetup e0 = edgeSetup(A, B, rowClip), e1 = edgeSetup(B, C, rowClip), e2 = edgeSetup(C, A, rowClip);
	TriangleRows tr;
	tr.edges[0] = edgeRows(e0);
	tr.edges[1] = edgeRows(e1);
	tr.edges[2] = edgeRows(e2);
	tr.clipL = clipR.l;
	tr.clipR = clipR.right();
	tr.alT = al.t;
	tr.alB = al.bottom();
	const bool degenerate = sameCorner(A, B) || sameCorner(B, C) || sameCorner(C, A);
	tr.flags = e0.kind | (e1.kind << 2) | (e2.kind << 4) | (degenerate ? 64 : 0);
	return tr;
}

// The float nearest to an integer below 2^46 in magnitude: both 23-bit halves are exact in
// float, so their sum is rounded once.
__device__ __forceinline__ float nearestFloat46(int64_t v) {
	return (float)(int32_t)(v >> 23) * 8388608.0f + (float)(int32_t)(v & 0x7FFFFF);
}

// clamp((int32_t)(trunc(n / d) + 1), L, R) for d != 0, as the rows take an edge's side.
// Operands below 2^46 with L and R within +-65536 take a float estimate of the quotient: its
// relative error stays below 2^-21, so an estimate at least 4 above R (or 6 below L) decides
// the clamp, and one nearer is within one of trunc(n / d), which one exact remainder step
// (trunc semantics: the remainder has the dividend's sign) corrects.  Estimates of 2^30 and
// more, and larger operands, take the exact division, which also keeps the int32 conversion
// of huge quotients as it is.
__device__ __forceinline__ int32_t edgeSide(int64_t n, int64_t d, int32_t L, int32_t R) {
	const int64_t big = (int64_t)1 << 46;
	if (n < big && n > -big && d < big && d > -big && L >= -65536 && R <= 65536) {
		const float est = nearestFloat46(n) * __builtin_amdgcn_rcpf(nearestFloat46(d));
		if (est < 1073741824.0f && est > -1073741824.0f) {
			if (est >= (float)R + 4.0f) return R;
			if (est <= (float)L - 6.0f) return tmin(L, R);
			int64_t q = (int64_t)(int32_t)est;
			const int64_t r = n - q * d;
			const int64_t dir = ((n < 0) == (d < 0)) ? 1 : -1;  // sign of the true quotient
			const int64_t ad = d < 0 ? -d : d;
			if (r != 0 && ((r < 0) != (n < 0))) q -= dir;       // one step too far from zero
			else if ((r < 0 ? -r : r) >= ad) q += dir;           // one step short
			return tmin(tmax(L, (int32_t)(q + 1)), R);
		}
	}
	return tmin(tmax(L, (int32_t)(divTrunc64(n, d) + 1)), R);
}

// Row y of the triangle (rasterizeTriangle, ITriangle2D.cpp:150-169, row y - alT of the
// aligned bound); empty outside the aligned bound.
__device__ __forceinline__ RowInterval triangleRow(const TriangleRows &tr, int32_t y) {
	RowInterval row{0, 0};
	if (y < tr.alT || y >= tr.alB) return row;
	const int32_t L = tr.clipL, R = tr.clipR;
	if (tr.flags & 64) { row.left = R; row.right = L; return row; }
	const int64_t r = y - tr.alT;
	row.left = L;
	row.right = R;
	// Every edge takes the same steps whatever its kind (lanes of one wavefront hold edges of
	// different kinds); the kind picks which result is kept.
	for (int k = 0; k < 3; k++) {
		const EdgeRows &e = tr.edges[k];
		const int32_t kind = (tr.flags >> (2 * k)) & 3;
		const int64_t along = e.b * r;
		const bool sloped = kind == 1 || kind == 2;
		const int32_t side = edgeSide(e.c - along + (kind == 1 ? 1 : 0), sloped ? e.a : 1, L, R);
		if (kind == 1) row.left = tmax(row.left, side);
		if (kind == 2) row.right = tmin(row.right, side);
		if (kind == 3 && e.c + along > e.a) { row.left = R; row.right = L; }
	}
	return row;
}

}  // namespace dfpsr
