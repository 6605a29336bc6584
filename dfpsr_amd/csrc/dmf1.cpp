// DMF1 model files: importFromContent_DMF1 (api/modelAPI.h:319,
// implementation/render/model/format/dmf1.cpp:25-361) as plain C++14 host code.
//
// The text is read as the reference reads it: a state machine over tokens -- identifiers,
// (property values), [indices] and <namespaces> -- split at white space and at the opening
// brackets, with numbers parsed by the reference's own digit-by-digit rule (string_toDouble,
// api/stringAPI.cpp:1586-1618) rather than strtod, so every coordinate rounds to the same
// float.  The converted model keeps the reference's point merging (ModelImpl::findPoint with
// a threshold of 0.00001, Model.cpp:289-321) and part filtering by detail level.  Malformed
// input never stops the parse: what the reference reports and skips is skipped here too.
#include "dmf1.h"

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace dfpsr {
namespace dmf1 {
namespace {

struct Vertex {
	float position[3] = {0.0f, 0.0f, 0.0f};
	float texCoord[4] = {0.0f, 0.0f, 0.0f, 0.0f};
	float color[4] = {1.0f, 1.0f, 1.0f, 1.0f};
};

struct Triangle {
	Vertex vertices[3];
};

struct Part {
	std::string textures[16];
	std::string shaderZero;
	int32_t minDetailLevel = 0, maxDetailLevel = 2;
	std::vector<Triangle> triangles;
	std::string name;
};

enum ParserState { WaitForStatement = 0, WaitForIndexOrProperty = 1, WaitForProperty = 2 };
enum ParserSpace { SpaceMain, SpacePart, SpaceTriangle, SpaceBone, SpaceShape, SpacePoint, SpaceUnhandled };

struct Parser {
	std::vector<Part> parts;
	int32_t state = WaitForStatement, space = SpaceMain, propertyIndex = 0;
	std::string lastPropertyName;
};

char lowerAscii(char c) { return (c >= 'A' && c <= 'Z') ? (char)(c - 'A' + 'a') : c; }

// string_caseInsensitiveMatch for the ASCII names the format uses.
bool sameName(const std::string &a, const char *b) {
	size_t i = 0;
	for (; i < a.size(); i++) {
		if (b[i] == '\0' || lowerAscii(a[i]) != lowerAscii(b[i])) return false;
	}
	return b[i] == '\0';
}
bool sameNameCI(const std::string &a, const std::string &b) {
	if (a.size() != b.size()) return false;
	for (size_t i = 0; i < a.size(); i++)
		if (lowerAscii(a[i]) != lowerAscii(b[i])) return false;
	return true;
}

// string_toInteger (stringAPI.cpp:1562-1583): '-' and '~' flip the sign anywhere, digits
// accumulate, a ',' or '.' ends the number, anything else is skipped.
int64_t toInteger(const std::string &s) {
	int64_t result = 0;
	bool negated = false;
	for (char c : s) {
		if (c == '-' || c == '~') {
			negated = !negated;
		} else if (c >= '0' && c <= '9') {
			result = (result * 10) + (int64_t)(c - '0');
		} else if (c == ',' || c == '.') {
			break;
		}
	}
	return negated ? -result : result;
}

// string_toDouble (stringAPI.cpp:1586-1618): whole digits as result * 10 + digit, each decimal
// digit added as digit / 10^k, the exponent after 'e' applied with pow(10, n).
double toDouble(const std::string &s) {
	double result = 0.0;
	bool negated = false, reachedDecimal = false;
	int64_t digitDivider = 1;
	for (size_t i = 0; i < s.size(); i++) {
		const char c = s[i];
		if (c == '-' || c == '~') {
			negated = !negated;
		} else if (c >= '0' && c <= '9') {
			if (reachedDecimal) {
				digitDivider = digitDivider * 10;
				result = result + ((double)(c - '0') / (double)digitDivider);
			} else {
				result = (result * 10) + (double)(c - '0');
			}
		} else if (c == ',' || c == '.') {
			reachedDecimal = true;
		} else if (c == 'e' || c == 'E') {
			result *= std::pow(10.0, (double)toInteger(s.substr(i + 1)));
			break;
		}
	}
	return negated ? -result : result;
}

// roundIndex (dmf1.cpp:107-110)
int32_t roundIndex(double value) { return (int32_t)std::round(value); }

// setProperty (dmf1.cpp:112-214)
void setProperty(Parser &p, const std::string &name, int32_t index, const std::string &content) {
	const float value = (float)toDouble(content);
	if (p.space == SpacePart) {
		if (p.parts.empty()) return;
		Part &part = p.parts.back();
		if (sameName(name, "Name")) {
			part.name = content;
		} else if (sameName(name, "Texture")) {
			if (index >= 0 && index < 16) part.textures[index] = content;
		} else if (sameName(name, "Shader")) {
			if (index == 0) part.shaderZero = content;
		} else if (sameName(name, "MinDetailLevel")) {
			part.minDetailLevel = roundIndex(value);
		} else if (sameName(name, "MaxDetailLevel")) {
			part.maxDetailLevel = roundIndex(value);
		}
	} else if (p.space == SpaceTriangle) {
		if (p.parts.empty() || p.parts.back().triangles.empty() || index < 0 || index > 2) return;
		Vertex &v = p.parts.back().triangles.back().vertices[index];
		static const char *const names[11] = {"X", "Y", "Z", "CR", "CG", "CB", "CA", "U1", "V1", "U2", "V2"};
		float *const slots[11] = {&v.position[0], &v.position[1], &v.position[2], &v.color[0], &v.color[1], &v.color[2],
		                          &v.color[3], &v.texCoord[0], &v.texCoord[1], &v.texCoord[2], &v.texCoord[3]};
		for (int k = 0; k < 11; k++) {
			if (sameName(name, names[k])) {
				*slots[k] = value;
				break;
			}
		}
	}
	// The main space's FilterType, CullingType and BoundMultiplier do not reach the converted
	// model (convertFromDMF1 makes a Solid model), and the other spaces hold nothing drawn.
}

// changeNamespace (dmf1.cpp:216-238)
void changeNamespace(Parser &p, const std::string &space) {
	if (sameName(space, "Part")) {
		p.parts.emplace_back();
		p.space = SpacePart;
	} else if (sameName(space, "Triangle")) {
		if (p.space == SpacePart || p.space == SpaceTriangle) {
			p.parts.back().triangles.emplace_back();
			p.space = SpaceTriangle;
		}
	} else if (sameName(space, "Bone")) {
		p.space = SpaceBone;
	} else if (sameName(space, "Shape")) {
		p.space = SpaceShape;
	} else if (sameName(space, "Point")) {
		p.space = SpacePoint;
	} else {
		p.space = SpaceUnhandled;
	}
}

// readToken (dmf1.cpp:242-289): text[start..end] inclusive.
void readToken(Parser &p, const std::string &text, int64_t start, int64_t end) {
	if (end < start) return;
	const char first = text[(size_t)start], last = text[(size_t)end];
	if (first == '(' && last == ')') {
		if (p.state == WaitForProperty || p.state == WaitForIndexOrProperty) {
			setProperty(p, p.lastPropertyName, p.propertyIndex, text.substr((size_t)start + 1, (size_t)(end - start - 1)));
			p.state = WaitForStatement;
			p.propertyIndex = 0;
		}
	} else if (first == '[' && last == ']') {
		if (p.state == WaitForIndexOrProperty) p.propertyIndex = roundIndex(toDouble(text.substr((size_t)start + 1, (size_t)(end - start - 1))));
	} else if (first == '<' && last == '>') {
		if (p.state == WaitForStatement && end - start <= 258) changeNamespace(p, text.substr((size_t)start + 1, (size_t)(end - start - 1)));
	} else if (p.state == WaitForStatement && end - start <= 258) {
		p.lastPropertyName = text.substr((size_t)start, (size_t)(end - start + 1));
		p.state = WaitForIndexOrProperty;
	}
}

bool isSpace(char c) { return c == 9 || c == 32 || c == 10 || c == 13; }

}  // namespace

// loadNative_DMF1 (dmf1.cpp:292-339) followed by convertFromDMF1 (:341-355).
bool importModel(const char *content, size_t length, int32_t detailLevel, Imported &out) {
	out = Imported();
	std::string text(content, length);
	// A UTF-8 byte order mark is not part of the text (string_load decodes it away).
	if (text.size() >= 3 && (unsigned char)text[0] == 0xEF && (unsigned char)text[1] == 0xBB && (unsigned char)text[2] == 0xBF) text.erase(0, 3);
	Parser p;
	if (text.size() >= 4 && text.compare(0, 4, "DMF1") == 0) {
		int64_t tokenStart = 4, readIndex = 4;
		char firstCharOfToken = '\0';
		const int64_t n = (int64_t)text.size();
		for (readIndex = 4; readIndex < n; readIndex++) {
			const char c = text[(size_t)readIndex];
			if (firstCharOfToken == '\0' && isSpace(c)) {
				readToken(p, text, tokenStart, readIndex - 1);
				tokenStart = readIndex + 1;
			} else if (c == '<' || c == '(' || c == '[') {
				readToken(p, text, tokenStart, readIndex - 1);
				tokenStart = readIndex;
				firstCharOfToken = c;
			} else if ((firstCharOfToken == '<' && c == '>') || (firstCharOfToken == '(' && c == ')') ||
			           (firstCharOfToken == '[' && c == ']')) {
				readToken(p, text, tokenStart, readIndex);
				tokenStart = readIndex + 1;
				firstCharOfToken = '\0';
			}
		}
		readToken(p, text, tokenStart, readIndex - 1);
	} else {
		out.valid = false;  // the reference reports the missing "DMF1" and returns an empty model
		return false;
	}
	out.valid = true;
	for (const Part &in : p.parts) {
		if (detailLevel < in.minDetailLevel || detailLevel > in.maxDetailLevel) continue;
		ImportedPart part;
		part.name = in.name;
		if (sameName(in.shaderZero, "M_Diffuse_1Tex")) {
			part.diffuseName = in.textures[0];
		} else if (sameName(in.shaderZero, "M_Diffuse_2Tex")) {
			part.diffuseName = in.textures[0];
			part.lightName = in.textures[1];
		}
		for (const Triangle &t : in.triangles) {
			for (int c = 0; c < 3; c++) {
				const Vertex &v = t.vertices[c];
				// addPointIfNeeded: the nearest existing point closer than the threshold, else a new one.
				const float threshold = 0.00001f;
				float bestDistance = threshold;
				int32_t best = -1;
				const int32_t count = (int32_t)(out.points.size() / 3);
				for (int32_t i = 0; i < count; i++) {
					const float dx = v.position[0] - out.points[(size_t)i * 3], dy = v.position[1] - out.points[(size_t)i * 3 + 1],
					            dz = v.position[2] - out.points[(size_t)i * 3 + 2];
					const float distance = std::sqrt(dx * dx + dy * dy + dz * dz);
					if (distance < bestDistance) {
						bestDistance = distance;
						best = i;
					}
				}
				if (best < 0) {
					best = count;
					out.points.insert(out.points.end(), v.position, v.position + 3);
				}
				part.indices.push_back(best);
				part.colors.insert(part.colors.end(), v.color, v.color + 4);
				part.texcoords.insert(part.texcoords.end(), v.texCoord, v.texCoord + 4);
			}
			// The fourth corner of a triangle polygon is unused: index -1, colour and texture
			// coordinates zero.
			part.indices.push_back(-1);
			for (int k = 0; k < 4; k++) {
				part.colors.push_back(0.0f);
				part.texcoords.push_back(0.0f);
			}
		}
		out.parts.push_back(part);
	}
	return true;
}

bool sameTextureName(const std::string &a, const std::string &b) { return sameNameCI(a, b); }

}  // namespace dmf1
}  // namespace dfpsr
