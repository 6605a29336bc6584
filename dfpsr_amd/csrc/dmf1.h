// DMF1 model import (plain C++14 host code, dmf1.cpp): the geometry and texture names of a
// DMF1 text, converted as importFromContent_DMF1 converts it
// (implementation/render/model/format/dmf1.cpp:341-361).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace dfpsr {
namespace dmf1 {

struct ImportedPart {
	std::string name;
	std::string diffuseName, lightName;  // texture names to fetch from the caller's pool ("" = none)
	std::vector<int32_t> indices;        // polygons x 4 point indices (triangles: the fourth is -1)
	std::vector<float> colors;           // polygons x 4 corners x RGBA
	std::vector<float> texcoords;        // polygons x 4 corners x (u1, v1, u2, v2)
};

struct Imported {
	bool valid = false;                  // the text starts with "DMF1"
	std::vector<float> points;           // merged points x 3
	std::vector<ImportedPart> parts;     // parts visible at the detail level, in file order
};

// Parses `length` bytes of DMF1 text; returns false (and an empty model) when it does not
// start with "DMF1", as the reference then returns an empty model.
bool importModel(const char *content, size_t length, int32_t detailLevel, Imported &out);

// BasicResourcePool::findTextureRgba's name comparison (ResourcePool.cpp:43-50): ASCII case-insensitive.
bool sameTextureName(const std::string &a, const std::string &b);

}  // namespace dmf1
}  // namespace dfpsr
