// STL reading for DEMMesh (reference: src/IO/mesh_loader.cpp:12-44).
#include "mesh_loader.hpp"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>
#include <string>

namespace STARDUST {

namespace {

struct Corner {
    float p[3];
    int k;  // corner number: 3 * triangle + 0..2
};

bool lex_less(const Corner& a, const Corner& b) {
    for (int i = 0; i < 3; ++i) {
        if (a.p[i] < b.p[i]) return true;
        if (b.p[i] < a.p[i]) return false;
    }
    return false;
}

bool read_binary(const std::string& raw, std::vector<float>& corners, std::vector<float>& normals) {
    if (raw.size() < 84) return false;
    uint32_t n = 0;
    std::memcpy(&n, raw.data() + 80, 4);
    if (raw.size() != 84 + 50 * (size_t)n) return false;
    corners.resize(9 * (size_t)n);
    normals.resize(3 * (size_t)n);
    for (uint32_t t = 0; t < n; ++t) {
        const char* r = raw.data() + 84 + 50 * (size_t)t;  // normal, 3 corners, attribute count
        std::memcpy(&normals[3 * (size_t)t], r, 12);
        std::memcpy(&corners[9 * (size_t)t], r + 12, 36);
    }
    return true;
}

void read_ascii(const std::string& raw, std::vector<float>& corners, std::vector<float>& normals) {
    std::istringstream in(raw);
    std::string w;
    int pending = 0;
    while (in >> w) {
        if (w == "facet") {
            std::string kw;
            float n[3];
            if (!(in >> kw >> n[0] >> n[1] >> n[2]) || kw != "normal")
                throw std::runtime_error("STL: malformed facet");
            if (pending) throw std::runtime_error("STL: facet with fewer than three vertices");
            normals.insert(normals.end(), n, n + 3);
            pending = 3;
        } else if (w == "vertex") {
            float v[3];
            if (!(in >> v[0] >> v[1] >> v[2]) || pending == 0) throw std::runtime_error("STL: malformed vertex");
            corners.insert(corners.end(), v, v + 3);
            --pending;
        }
    }
    if (pending) throw std::runtime_error("STL: truncated facet");
}

}  // namespace

void loadAndConvertMesh(const char* filename, std::vector<float4>& out_vert, std::vector<int>& out_idx,
                        std::vector<float4>& out_norm) {
    std::ifstream f(filename ? filename : "", std::ios::binary);
    if (!f) throw std::runtime_error(std::string("STL: cannot open ") + (filename ? filename : "null"));
    const std::string raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<float> corners, normals;
    if (!read_binary(raw, corners, normals)) {
        if (raw.compare(0, 5, "solid") != 0) throw std::runtime_error(std::string("STL: not an STL file: ") + filename);
        read_ascii(raw, corners, normals);
    }
    const size_t nc = corners.size() / 3;
    std::vector<Corner> c(nc);
    for (size_t k = 0; k < nc; ++k) {
        std::memcpy(c[k].p, &corners[3 * k], 12);
        c[k].k = (int)k;
    }
    std::stable_sort(c.begin(), c.end(), lex_less);
    out_vert.clear();
    out_idx.assign(nc, -1);
    for (size_t k = 0; k < nc; ++k) {
        if (k == 0 || lex_less(c[k - 1], c[k])) out_vert.push_back(make_float4(c[k].p[0], c[k].p[1], c[k].p[2], 0.0f));
        out_idx[c[k].k] = (int)out_vert.size() - 1;
    }
    out_norm.clear();
    for (size_t t = 0; t < normals.size() / 3; ++t)
        out_norm.push_back(make_float4(normals[3 * t], normals[3 * t + 1], normals[3 * t + 2], 0.0f));
}

}  // namespace STARDUST
