// Per-quad shading of the polygon path on gfx950: the RgbaMultiply shader family
// (implementation/render/shader/RgbaMultiply.h), mip selection and 8-bit fixed-point
// bilinear sampling of the reference's texture pyramid (api/textureAPI.h) and the
// saturating byte packing (implementation/image/PackOrder.h).
#pragma once

#include "geometry.h"

namespace dfpsr {

// A texture as the kernels see it: the pyramid buffer (smallest level first, the layout of
// implementation/image/Texture.h:42-96) plus the bit masks the sampler needs.
struct TextureDesc {
	const uint32_t *texels;
	uint32_t log2w, log2h, maxMip;
	uint32_t startOffset, maxLevelMask;
	uint32_t maxWidthAndMask, maxHeightAndMask;
	uint32_t minWidthOrMask, minHeightOrMask;
	float floatMaxWidth, floatMaxHeight;
	uint32_t texelCount;
};

// The layout fields of the TextureRgbaU8 constructor (Texture.h:66-101).
__host__ inline TextureDesc textureLayout(uint32_t log2w, uint32_t log2h, uint32_t requestedMaxMip) {
	TextureDesc t;
	t.texels = nullptr;
	t.log2w = log2w;
	t.log2h = log2h;
	if (requestedMaxMip > 15u) requestedMaxMip = 15u;
	t.maxMip = tmin(tmin(log2w, log2h), requestedMaxMip);
	uint32_t highest = 1u << (log2w + log2h);
	uint64_t pixelCount = 0;
	uint32_t levelCount = highest;
	for (int32_t level = (int32_t)t.maxMip; level >= 0; level--) {
		pixelCount |= levelCount;
		levelCount >>= 2;
	}
	t.startOffset = (uint32_t)pixelCount & ~highest;
	t.maxLevelMask = highest - 1;
	t.minWidthOrMask = (1u << (log2w - t.maxMip)) - 1;
	t.minHeightOrMask = (1u << (log2h - t.maxMip)) - 1;
	t.maxWidthAndMask = (1u << log2w) - 1;
	t.maxHeightAndMask = (1u << log2h) - 1;
	t.floatMaxWidth = (float)(1u << log2w);
	t.floatMaxHeight = (float)(1u << log2h);
	t.texelCount = (uint32_t)pixelCount;
	return t;
}

// texture_getPixelOffsetToLayer (textureAPI.h:79-85)
DH uint32_t layerOffset(const TextureDesc &t, uint32_t mip) { return t.startOffset & (t.maxLevelMask >> (mip << 1)); }

// texture_getPixelOffset without optimization flags (textureAPI.h:104-170).  Past log2w the
// stride wraps around as a uint32 and the shift takes it modulo 32, as the reference's scalar
// shift (x86 shl) and v_lshlrev_b32 both do.
DH uint32_t pixelOffset(const TextureDesc &t, uint32_t x, uint32_t y, uint32_t mip) {
	uint32_t mx = (t.maxWidthAndMask >> mip) | t.minWidthOrMask;
	uint32_t my = (t.maxHeightAndMask >> mip) | t.minHeightOrMask;
	return layerOffset(t, mip) + (((y & my) << ((t.log2w - mip) & 31u)) | (x & mx));
}

// _mm_cvttps_epi32 of one lane: truncation, 0x80000000 when out of range or NaN.
DH uint32_t truncU32(float f) {
	if (!(f >= -2147483648.0f && f < 2147483648.0f)) return 0x80000000u;
	return (uint32_t)(int32_t)f;
}
DH float fromU32(uint32_t u) { return (float)(int32_t)u; }  // _mm_cvtepi32_ps

// texture_getMipLevelIndex (textureAPI.h:472-495) from lanes 0, 1 and 2 of the quad.
DH uint32_t mipLevelIndex(const TextureDesc &t, float u0, float u1, float u2, float v0, float v1, float v2) {
	float offsetUX = fabsf(u0 - u1);
	float offsetUY = fabsf(u0 - u2);
	float offsetVX = fabsf(v0 - v1);
	float offsetVY = fabsf(v0 - v2);
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

// (a * (256 - w) + b * w) >> 8 on the four bytes at once, two 16-bit lanes per 32-bit half
// like weightColors (textureAPI.h:248-266).
DH uint32_t mixColors(uint32_t a, uint32_t b, uint32_t w) {
	uint32_t iw = 256u - w;
	uint32_t lo = ((a & 0x00FF00FFu) * iw + (b & 0x00FF00FFu) * w) >> 8;
	uint32_t hi = (((a >> 8) & 0x00FF00FFu) * iw + ((b >> 8) & 0x00FF00FFu) * w);
	return (lo & 0x00FF00FFu) | (hi & 0xFF00FF00u);
}

// texture_sample_bilinear with MIP_INSIDE (textureAPI.h:342-438); highest = HIGHEST_RESOLUTION.
// The four texel offsets (from t.texels) and 8-bit weights of one bilinear sample.
struct BilinearSample {
	uint32_t offset[4];  // upper left, upper right, lower left, lower right
	uint32_t wx, wy;
};

DH BilinearSample bilinearSample(const TextureDesc &t, float u, float v, uint32_t mip, bool highest) {
	uint32_t scaleU = 256u << t.log2w, scaleV = 256u << t.log2h;
	uint32_t maskX = t.maxWidthAndMask, maskY = t.maxHeightAndMask;
	uint32_t stride = t.log2w;
	uint32_t layer = t.startOffset;
	if (!highest) {
		scaleU >>= mip; scaleV >>= mip;
		maskX >>= mip; maskY >>= mip;
		stride -= mip;
		layer = layerOffset(t, mip);
	}
	uint32_t subX = truncU32((u + 256.0f) * fromU32(scaleU)) - 128u;
	uint32_t subY = truncU32((v + 256.0f) * fromU32(scaleV)) - 128u;
	BilinearSample b;
	b.wx = subX & 255u;
	b.wy = subY & 255u;
	uint32_t left = subX >> 8, top = subY >> 8;
	uint32_t right = (left + 1u) & maskX, bottom = (top + 1u) & maskY;
	left &= maskX;
	top &= maskY;
	b.offset[0] = layer + ((top << stride) | left);
	b.offset[1] = layer + ((top << stride) | right);
	b.offset[2] = layer + ((bottom << stride) | left);
	b.offset[3] = layer + ((bottom << stride) | right);
	return b;
}

// Texels live in global memory; loading them through a global pointer keeps the loads off the
// LDS counter, so a quad's sixteen loads can all be in flight at once.
__device__ __forceinline__ uint32_t loadTexel(const uint32_t *texels, uint32_t offset) {
	return ((const __attribute__((address_space(1))) uint32_t *)texels)[offset];
}

DH uint32_t mixSample(const BilinearSample &b, uint32_t ul, uint32_t ur, uint32_t bl, uint32_t br) {
	return mixColors(mixColors(ul, ur, b.wx), mixColors(bl, br, b.wx), b.wy);
}

// The bilinear samples of the four lanes of a quad from one texture level: all sixteen texels
// are loaded before any is mixed.
__device__ __forceinline__ void sampleQuad(const TextureDesc &t, const float *u, const float *v, uint32_t mip, bool highest,
                                           uint32_t *out) {
	BilinearSample b[4];
	for (int l = 0; l < 4; l++) b[l] = bilinearSample(t, u[l], v[l], mip, highest);
	uint32_t px[16];
	for (int l = 0; l < 4; l++)
		for (int k = 0; k < 4; k++) px[l * 4 + k] = loadTexel(t.texels, b[l].offset[k]);
	for (int l = 0; l < 4; l++) out[l] = mixSample(b[l], px[l * 4], px[l * 4 + 1], px[l * 4 + 2], px[l * 4 + 3]);
}

DH uint32_t sampleBilinear(const TextureDesc &t, float u, float v, uint32_t mip, bool highest) {
	uint32_t scaleU = 256u << t.log2w, scaleV = 256u << t.log2h;
	uint32_t maskX = t.maxWidthAndMask, maskY = t.maxHeightAndMask;
	uint32_t stride = t.log2w;
	uint32_t layer = t.startOffset;
	if (!highest) {
		scaleU >>= mip; scaleV >>= mip;
		maskX >>= mip; maskY >>= mip;
		stride -= mip;
		layer = layerOffset(t, mip);
	}
	uint32_t subX = truncU32((u + 256.0f) * fromU32(scaleU)) - 128u;
	uint32_t subY = truncU32((v + 256.0f) * fromU32(scaleV)) - 128u;
	uint32_t wx = subX & 255u, wy = subY & 255u;
	uint32_t left = subX >> 8, top = subY >> 8;
	uint32_t right = (left + 1u) & maskX, bottom = (top + 1u) & maskY;
	left &= maskX;
	top &= maskY;
	const uint32_t *base = t.texels + layer;
	uint32_t ul = __ldg(base + ((top << stride) | left));
	uint32_t ur = __ldg(base + ((top << stride) | right));
	uint32_t bl = __ldg(base + ((bottom << stride) | left));
	uint32_t br = __ldg(base + ((bottom << stride) | right));
	return mixColors(mixColors(ul, ur, wx), mixColors(bl, br, wx), wy);
}

// clampUpper + truncateToU32 of one channel: the clamp maps NaN to 255.1 (as minps does with the
// limit as second operand), so only values below -2^31 remain for truncU32's 0x80000000, and
// raising them to -2^31 gives exactly that.
DH uint32_t saturatedChannel(float x) {
	const float clamped = x < 255.1f ? x : 255.1f;
	return (uint32_t)(int32_t)fmaxf(clamped, -2147483648.0f);
}

// clampUpper + truncateToU32 + byte packing for RGBA (PackOrder.h:191-212).
DH uint32_t packSaturated(float r, float g, float b, float a) {
	return saturatedChannel(r) | (saturatedChannel(g) << 8) | (saturatedChannel(b) << 16) | (saturatedChannel(a) << 24);
}

DH void unpackColor(uint32_t c, float &r, float &g, float &b, float &a) {
	r = fromU32(c & 255u);
	g = fromU32((c >> 8) & 255u);
	b = fromU32((c >> 16) & 255u);
	a = fromU32(c >> 24);
}

// 1 / x for the perspective weights.  The reference refines an approximate reciprocal with one
// Newton step, r = (r0 + r0) - x * r0 * r0 (base/simd.h:4045-4053); the same step is applied
// here to the hardware estimate v_rcp_f32, which keeps the result within about an ulp of the
// reference's (whose estimate table is CPU-vendor specific).
__device__ __forceinline__ float refinedReciprocal(float x) {
	float r0 = __builtin_amdgcn_rcpf(x);
	return (r0 + r0) - ((x * r0) * r0);
}

// The shader a triangle selects (RgbaMultiply.h:37-58, :110-175): colours pre-scaled by 255
// and by 1/255 per bound texture, "vertex fading" when a channel differs across corners,
// "colorless" when every channel is about 1.
struct __attribute__((aligned(16))) ShaderInputs {
	f3 red, green, blue, alpha;  // scaled corner colours
	f3 u1, v1, u2, v2;
	int32_t diffuse, light;      // texture slots, -1 when absent
	bool fading, colorless;
};

DH bool almostZero(float v) { return v > -0.001f && v < 0.001f; }
DH bool almostOne(float v) { return v > 0.999f && v < 1.001f; }
DH bool almostSame3(const f3 &c) { return almostZero(c.x - c.y) && almostZero(c.x - c.z) && almostZero(c.y - c.z); }
DH bool almostOne3(const f3 &c) { return almostOne(c.x) && almostOne(c.y) && almostOne(c.z); }

DH void finishShaderInputs(ShaderInputs &s) {
	float k = 255.0f;
	if (s.diffuse >= 0) k *= 1.0f / 255.0f;
	if (s.light >= 0) k *= 1.0f / 255.0f;
	s.red = scale3(s.red, k);
	s.green = scale3(s.green, k);
	s.blue = scale3(s.blue, k);
	s.alpha = scale3(s.alpha, k);
	s.fading = !(almostSame3(s.red) && almostSame3(s.green) && almostSame3(s.blue) && almostSame3(s.alpha));
	// The texture-only shaders are chosen only for triangles without vertex fading: corners
	// within 0.001 of 1 that still differ by 0.001 or more take the texture x interpolated
	// vertex colour shader (DiffuseVertex / LightVertex, RgbaMultiply.h:130-136, :155-156).
	s.colorless = !s.fading && almostOne3(s.red) && almostOne3(s.green) && almostOne3(s.blue) && almostOne3(s.alpha);
}

// shaderMethods::interpolate (shaderMethods.h:39-44)
DH float interp(const f3 &d, float wa, float wb, float wc) { return ((d.x * wa) + (d.y * wb)) + (d.z * wc); }

struct Quad4 { float v[4]; };

// getPixels_2x2 (RgbaMultiply.h:74-106) for the lanes of one quad in mask (bit l = lane l;
// the others are left unset).  The mip level always comes from lanes 0, 1 and 2.
DH void shadeQuad(const ShaderInputs &s, const TextureDesc *textures, const float *wa, const float *wb, const float *wc,
                  uint32_t mask, float *r, float *g, float *b, float *a) {
	if (s.diffuse >= 0 && s.light < 0 && s.colorless) {
		const TextureDesc &t = textures[s.diffuse];
		float u[4], v[4];
		for (int l = 0; l < 4; l++) { u[l] = interp(s.u1, wa[l], wb[l], wc[l]); v[l] = interp(s.v1, wa[l], wb[l], wc[l]); }
		uint32_t mip = mipLevelIndex(t, u[0], u[1], u[2], v[0], v[1], v[2]);
		if (mask == 0xFu) {
			uint32_t texel[4];
			sampleQuad(t, u, v, mip, false, texel);
			for (int l = 0; l < 4; l++) unpackColor(texel[l], r[l], g[l], b[l], a[l]);
			return;
		}
		for (int l = 0; l < 4; l++)
			if ((mask >> l) & 1u) unpackColor(sampleBilinear(t, u[l], v[l], mip, false), r[l], g[l], b[l], a[l]);
		return;
	}
	if (s.light >= 0 && s.diffuse < 0 && s.colorless) {
		const TextureDesc &t = textures[s.light];
		for (int l = 0; l < 4; l++) {
			if (!((mask >> l) & 1u)) continue;
			float u = interp(s.u2, wa[l], wb[l], wc[l]);
			float v = interp(s.v2, wa[l], wb[l], wc[l]);
			unpackColor(sampleBilinear(t, u, v, 0u, true), r[l], g[l], b[l], a[l]);
		}
		return;
	}
	for (int l = 0; l < 4; l++) {
		if (s.fading) {
			r[l] = interp(s.red, wa[l], wb[l], wc[l]);
			g[l] = interp(s.green, wa[l], wb[l], wc[l]);
			b[l] = interp(s.blue, wa[l], wb[l], wc[l]);
			a[l] = interp(s.alpha, wa[l], wb[l], wc[l]);
		} else {
			r[l] = s.red.x; g[l] = s.green.x; b[l] = s.blue.x; a[l] = s.alpha.x;
		}
	}
	if (s.diffuse >= 0) {
		const TextureDesc &t = textures[s.diffuse];
		float u[4], v[4];
		for (int l = 0; l < 4; l++) { u[l] = interp(s.u1, wa[l], wb[l], wc[l]); v[l] = interp(s.v1, wa[l], wb[l], wc[l]); }
		uint32_t mip = mipLevelIndex(t, u[0], u[1], u[2], v[0], v[1], v[2]);
		uint32_t texel[4];
		if (mask == 0xFu) sampleQuad(t, u, v, mip, false, texel);
		for (int l = 0; l < 4; l++) {
			if (!((mask >> l) & 1u)) continue;
			float sr, sg, sb, sa;
			unpackColor(mask == 0xFu ? texel[l] : sampleBilinear(t, u[l], v[l], mip, false), sr, sg, sb, sa);
			r[l] = r[l] * sr; g[l] = g[l] * sg; b[l] = b[l] * sb; a[l] = a[l] * sa;
		}
	}
	if (s.light >= 0) {
		const TextureDesc &t = textures[s.light];
		for (int l = 0; l < 4; l++) {
			if (!((mask >> l) & 1u)) continue;
			float u = interp(s.u2, wa[l], wb[l], wc[l]);
			float v = interp(s.v2, wa[l], wb[l], wc[l]);
			float sr, sg, sb, sa;
			unpackColor(sampleBilinear(t, u, v, 0u, true), sr, sg, sb, sa);
			r[l] = r[l] * sr; g[l] = g[l] * sg; b[l] = b[l] * sb; a[l] = a[l] * sa;
		}
	}
}

// The mip level a quad samples its diffuse texture at (0 without one): texture_getMipLevelIndex
// of lanes 0, 1 and 2's texture coordinates, as shadeQuad takes it.
__device__ __forceinline__ uint32_t quadMip(const ShaderInputs &s, const TextureDesc *textures, const float *wa, const float *wb,
                                            const float *wc) {
	if (s.diffuse < 0) return 0u;
	float u[3], v[3];
	for (int l = 0; l < 3; l++) { u[l] = interp(s.u1, wa[l], wb[l], wc[l]); v[l] = interp(s.v1, wa[l], wb[l], wc[l]); }
	return mipLevelIndex(textures[s.diffuse], u[0], u[1], u[2], v[0], v[1], v[2]);
}

// One lane of shadeQuad followed by packSaturated, with the quad's mip level given: the same
// operations in the same order for one pixel.
__device__ __forceinline__ uint32_t shadePixel(const ShaderInputs &s, const TextureDesc *textures, float wa, float wb, float wc,
                                               uint32_t mip) {
	const bool textureOnly = s.colorless && ((s.diffuse >= 0) != (s.light >= 0));
	float r = 0.0f, g = 0.0f, b = 0.0f, a = 0.0f;
	if (!textureOnly) {
		if (s.fading) {
			r = interp(s.red, wa, wb, wc);
			g = interp(s.green, wa, wb, wc);
			b = interp(s.blue, wa, wb, wc);
			a = interp(s.alpha, wa, wb, wc);
		} else {
			r = s.red.x; g = s.green.x; b = s.blue.x; a = s.alpha.x;
		}
	}
	if (s.diffuse >= 0) {
		float sr, sg, sb, sa;
		unpackColor(sampleBilinear(textures[s.diffuse], interp(s.u1, wa, wb, wc), interp(s.v1, wa, wb, wc), mip, false), sr, sg, sb, sa);
		if (textureOnly) {
			r = sr; g = sg; b = sb; a = sa;
		} else {
			r = r * sr; g = g * sg; b = b * sb; a = a * sa;
		}
	}
	if (s.light >= 0) {
		float sr, sg, sb, sa;
		unpackColor(sampleBilinear(textures[s.light], interp(s.u2, wa, wb, wc), interp(s.v2, wa, wb, wc), 0u, true), sr, sg, sb, sa);
		if (textureOnly) {
			r = sr; g = sg; b = sb; a = sa;
		} else {
			r = r * sr; g = g * sg; b = b * sb; a = a * sa;
		}
	}
	return packSaturated(r, g, b, a);
}

// Lane l of shadeQuad: the same operations for one pixel of the quad.  The mip level still
// comes from the texture coordinates of lanes 0, 1 and 2 (texture_getMipLevelIndex).
DH float pickLane(const float *v, int l) { return l == 0 ? v[0] : (l == 1 ? v[1] : (l == 2 ? v[2] : v[3])); }

DH void shadeLane(const ShaderInputs &s, const TextureDesc *textures, const float *wa, const float *wb, const float *wc, int l,
                  float &r, float &g, float &b, float &a) {
	const float la = pickLane(wa, l), lb = pickLane(wb, l), lc = pickLane(wc, l);
	if (s.diffuse >= 0 && s.light < 0 && s.colorless) {
		const TextureDesc &t = textures[s.diffuse];
		float u[3], v[3];
		for (int k = 0; k < 3; k++) { u[k] = interp(s.u1, wa[k], wb[k], wc[k]); v[k] = interp(s.v1, wa[k], wb[k], wc[k]); }
		uint32_t mip = mipLevelIndex(t, u[0], u[1], u[2], v[0], v[1], v[2]);
		unpackColor(sampleBilinear(t, interp(s.u1, la, lb, lc), interp(s.v1, la, lb, lc), mip, false), r, g, b, a);
		return;
	}
	if (s.light >= 0 && s.diffuse < 0 && s.colorless) {
		const TextureDesc &t = textures[s.light];
		unpackColor(sampleBilinear(t, interp(s.u2, la, lb, lc), interp(s.v2, la, lb, lc), 0u, true), r, g, b, a);
		return;
	}
	if (s.fading) {
		r = interp(s.red, la, lb, lc);
		g = interp(s.green, la, lb, lc);
		b = interp(s.blue, la, lb, lc);
		a = interp(s.alpha, la, lb, lc);
	} else {
		r = s.red.x; g = s.green.x; b = s.blue.x; a = s.alpha.x;
	}
	if (s.diffuse >= 0) {
		const TextureDesc &t = textures[s.diffuse];
		float u[3], v[3];
		for (int k = 0; k < 3; k++) { u[k] = interp(s.u1, wa[k], wb[k], wc[k]); v[k] = interp(s.v1, wa[k], wb[k], wc[k]); }
		uint32_t mip = mipLevelIndex(t, u[0], u[1], u[2], v[0], v[1], v[2]);
		float sr, sg, sb, sa;
		unpackColor(sampleBilinear(t, interp(s.u1, la, lb, lc), interp(s.v1, la, lb, lc), mip, false), sr, sg, sb, sa);
		r = r * sr; g = g * sg; b = b * sb; a = a * sa;
	}
	if (s.light >= 0) {
		const TextureDesc &t = textures[s.light];
		float sr, sg, sb, sa;
		unpackColor(sampleBilinear(t, interp(s.u2, la, lb, lc), interp(s.v2, la, lb, lc), 0u, true), sr, sg, sb, sa);
		r = r * sr; g = g * sg; b = b * sb; a = a * sa;
	}
}

}  // namespace dfpsr
