// STL file reader (STLReader::read_from_file, src/geo/stlreader.cpp:57-150):
// binary and text STL, float32 coordinates promoted to double, written as 9
// doubles per triangle straight into the array the mesh upload takes.
// Host code only.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

namespace ocl_amd {

inline void stl_push(std::vector<double>& out, const float v[3][3]) {
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) out.push_back((double)v[a][c]);
}

// Text STL: "vertex" lines fill up to three vertex slots (a fourth
// overwrites the third), "facet" starts a facet, "endfacet" adds it once the
// third slot was reached; keywords are matched on their first five
// non-blank characters (stlreader.cpp:92-146).
inline void stl_read_text(std::ifstream& in, std::vector<double>& out) {
    std::string line;
    std::getline(in, line);  // the rest of the "solid" line
    float v[3][3] = {};
    int k = 0;
    while (std::getline(in, line)) {
        char key[6] = {0, 0, 0, 0, 0, 0};
        size_t j = 0;
        int i = 0;
        for (; i < 5; ++i, ++j) {
            if (j >= line.size()) break;
            while (j < line.size() && (line[j] == ' ' || line[j] == '\t')) ++j;
            if (j >= line.size()) break;
            key[i] = line[j];
        }
        if (i < 5) continue;
        if (!std::strcmp(key, "verte")) {
            const size_t at = line.find("vertex");
            if (at != std::string::npos) std::sscanf(line.c_str() + at + 6, "%f %f %f", &v[k][0], &v[k][1], &v[k][2]);
            k = k + 1 > 2 ? 2 : k + 1;
        } else if (!std::strcmp(key, "facet")) {
            k = 0;
        } else if (!std::strcmp(key, "endfa")) {
            if (k == 2) stl_push(out, v);
        }
    }
}

// Binary STL: 80-byte header, uint32 facet count, then per facet 12 float32
// (normal, three vertices) and a 2-byte attribute (stlreader.cpp:70-89).  A
// truncated file stops at its last whole facet.
inline void stl_read_binary(std::ifstream& in, std::vector<double>& out) {
    char header[75];
    in.read(header, 75);
    uint32_t count = 0;
    in.read(reinterpret_cast<char*>(&count), 4);
    for (uint32_t i = 0; i < count && in; ++i) {
        unsigned char rec[50];
        in.read(reinterpret_cast<char*>(rec), 50);
        if (!in) break;
        float v[3][3];
        std::memcpy(v, rec + 12, 36);
        stl_push(out, v);
    }
}

// Appends the file's triangles to out.  False if the file cannot be opened
// or is shorter than the 5-byte "solid" probe (the reference then reads no
// triangles).
inline bool stl_read(const char* path, std::vector<double>& out) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    char head[5];
    in.read(head, 5);
    if (in.gcount() < 5) return false;
    if (std::memcmp(head, "solid", 5) == 0)
        stl_read_text(in, out);
    else
        stl_read_binary(in, out);
    return true;
}

}  // namespace ocl_amd
