// Host-side reading of a 3DGS PLY header: the rules and messages of the reference loader
// (src/ply_loader.cpp:20-66 header, :86-110 columns and SH degree). Accepted: "ply" first line,
// "format binary_little_endian", one "vertex" element whose properties are all float/float32
// (other elements must be empty; properties of other elements are skipped), "end_header".
// Columns are found by name (a repeated name resolves to its last column, as the reference's
// name map does); f_rest_0.. are counted while present and must match an SH degree 0..4.
// Normals and any other columns are ignored.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

namespace ans {

struct PlyError : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

struct PlyLayout {
    int64_t n = 0;       // vertices (Gaussians)
    int stride = 0;      // float32 properties per vertex
    int degree = 0;      // SH degree of the file, from the f_rest count
    size_t body = 0;     // byte offset of the vertex block
    int pos[3] = {}, opacity = 0, scale[3] = {}, rot[4] = {}, dc[3] = {};
    std::vector<int> rest;  // column of f_rest_k, k = 0 .. 3 * ((degree + 1)^2 - 1) - 1
};

namespace ply_detail {

inline std::vector<std::string_view> words(std::string_view line) {
    std::vector<std::string_view> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == '\v' || line[i] == '\f')) ++i;
        const size_t j = i;
        while (i < line.size() && !(line[i] == ' ' || line[i] == '\t' || line[i] == '\v' || line[i] == '\f')) ++i;
        if (i > j) out.push_back(line.substr(j, i - j));
    }
    return out;
}

}  // namespace ply_detail

// Parses the header of the file image data[bytes] (name is used in messages). If bytes covers
// only the header, pass the file's full size as fileBytes so the length check sees the body.
inline PlyLayout parsePlyHeader(const char* data, size_t bytes, const std::string& name, size_t fileBytes) {
    using ply_detail::words;
    PlyLayout L;
    std::vector<std::string> props;
    size_t at = 0;
    auto nextLine = [&](std::string_view& line) {
        if (at >= bytes) return false;
        const void* nl = std::memchr(data + at, '\n', bytes - at);
        const size_t end = nl ? static_cast<size_t>(static_cast<const char*>(nl) - data) : bytes;
        line = std::string_view(data + at, end - at);
        at = nl ? end + 1 : bytes;
        if (!line.empty() && line.back() == '\r') line.remove_suffix(1);
        return true;
    };
    std::string_view line;
    if (!nextLine(line) || line.substr(0, 3) != "ply") throw PlyError(name + ": not a PLY file");
    bool inVertex = false, sawFormat = false, ended = false;
    while (nextLine(line)) {
        const auto w = words(line);
        if (w.empty()) continue;
        if (w[0] == "format") {
            const std::string format(w.size() > 1 ? w[1] : std::string_view());
            if (format != "binary_little_endian")
                throw PlyError(name + ": only binary_little_endian PLY is supported, got " + format);
            sawFormat = true;
        } else if (w[0] == "element") {
            const std::string element(w.size() > 1 ? w[1] : std::string_view());
            const uint64_t count = w.size() > 2 ? std::strtoull(std::string(w[2]).c_str(), nullptr, 10) : 0;
            inVertex = element == "vertex";
            if (inVertex) L.n = static_cast<int64_t>(count);
            else if (count != 0) throw PlyError(name + ": unexpected element '" + element + "'");
        } else if (w[0] == "property") {
            if (!inVertex) continue;
            const std::string type(w.size() > 1 ? w[1] : std::string_view());
            const std::string prop(w.size() > 2 ? w[2] : std::string_view());
            if (type != "float" && type != "float32")
                throw PlyError(name + ": property '" + prop + "' has type " + type + ", expected float");
            props.push_back(prop);
        } else if (w[0] == "end_header") {
            if (!sawFormat) throw PlyError(name + ": missing format line");
            L.body = at;
            ended = true;
            break;
        }
    }
    if (!ended) throw PlyError(name + ": header has no end_header");

    std::unordered_map<std::string, int> column;
    for (size_t i = 0; i < props.size(); ++i) column[props[i]] = static_cast<int>(i);
    auto col = [&](const std::string& p) {
        const auto it = column.find(p);
        if (it == column.end()) throw PlyError(name + ": missing property '" + p + "'");
        return it->second;
    };
    for (int k = 0; k < 3; ++k) L.pos[k] = col(std::string(1, "xyz"[k]));
    L.opacity = col("opacity");
    for (int k = 0; k < 3; ++k) L.scale[k] = col("scale_" + std::to_string(k));
    for (int k = 0; k < 4; ++k) L.rot[k] = col("rot_" + std::to_string(k));
    for (int k = 0; k < 3; ++k) L.dc[k] = col("f_dc_" + std::to_string(k));
    size_t restCount = 0;
    while (column.count("f_rest_" + std::to_string(restCount))) ++restCount;
    L.degree = -1;
    for (int d = 0; d <= 4; ++d)
        if (restCount == static_cast<size_t>(3 * ((d + 1) * (d + 1) - 1))) L.degree = d;
    if (L.degree < 0)
        throw PlyError(name + ": " + std::to_string(restCount) + " f_rest properties do not match any SH degree");
    for (size_t k = 0; k < restCount; ++k) L.rest.push_back(col("f_rest_" + std::to_string(k)));
    L.stride = static_cast<int>(props.size());
    const uint64_t rowBytes = static_cast<uint64_t>(L.stride) * 4u;
    if (fileBytes < L.body || static_cast<uint64_t>(L.n) > (fileBytes - L.body) / rowBytes)
        throw PlyError(name + ": file is shorter than the header promises (" + std::to_string(L.n) + " Gaussians)");
    return L;
}

}  // namespace ans
