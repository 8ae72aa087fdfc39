// Trained-model export and import in the 3D Gaussian Splatting .ply layout (the layout of
// writeSplatPly / readSplatPly, SplatPly.cpp:261-342): binary little endian vertices with
// x y z nx ny nz f_dc_0..2 [f_rest_*] opacity scale_0..2 rot_0..3. Spherical-harmonic
// coefficients beyond DC (an extension; the reference trains DC only) are written as the
// usual f_rest_0 .. f_rest_{3K-1}, channel-major (all red coefficients, then green, then blue)
// for K = 3, 8 or 15 coefficients per channel. Host-only code.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "trainer.h"

using namespace u3d;

namespace {

int shCoefficients(int degree) { return degree <= 0 ? 0 : (degree + 1) * (degree + 1) - 1; }

int degreeOfRest(int restCount)
{
    for (int d = 1; d <= 3; ++d)
        if (restCount == 3 * shCoefficients(d))
            return d;
    return restCount == 0 ? 0 : -1;
}

void putF32(std::vector<char> &buf, float v)
{
    char b[4];
    std::memcpy(b, &v, 4); // the host is little endian (x86-64)
    buf.insert(buf.end(), b, b + 4);
}

struct PlyProp {
    std::string name;
    int size = 0;   // bytes in a binary file
    char kind = 0;  // 'f' float, 'd' double, 'i' integer (signed), 'u' unsigned
};

bool typeInfo(const std::string &t, PlyProp &p)
{
    if (t == "float" || t == "float32") { p.size = 4; p.kind = 'f'; return true; }
    if (t == "double" || t == "float64") { p.size = 8; p.kind = 'd'; return true; }
    if (t == "uchar" || t == "uint8") { p.size = 1; p.kind = 'u'; return true; }
    if (t == "char" || t == "int8") { p.size = 1; p.kind = 'i'; return true; }
    if (t == "ushort" || t == "uint16") { p.size = 2; p.kind = 'u'; return true; }
    if (t == "short" || t == "int16") { p.size = 2; p.kind = 'i'; return true; }
    if (t == "uint" || t == "uint32") { p.size = 4; p.kind = 'u'; return true; }
    if (t == "int" || t == "int32") { p.size = 4; p.kind = 'i'; return true; }
    return false;
}

double decode(const char *b, const PlyProp &p)
{
    switch (p.kind) {
    case 'f': { float v; std::memcpy(&v, b, 4); return v; }
    case 'd': { double v; std::memcpy(&v, b, 8); return v; }
    case 'u': { uint32_t v = 0; std::memcpy(&v, b, size_t(p.size)); return double(v); }
    default: {
        int32_t v = 0;
        std::memcpy(&v, b, size_t(p.size));
        const int shift = 32 - 8 * p.size; // sign-extend
        return double((v << shift) >> shift);
    }
    }
}

} // namespace

extern "C" {

int u3d_write_splat_ply(const char *path, const float *means, const float *scales,
                        const float *quats, const float *opacities, const float *colors, int64_t n,
                        const float *sh_rest, int sh_degree)
{
    U3D_CHECK(path != nullptr && n >= 0, U3D_ERR_INVALID, "bad argument");
    U3D_CHECK(n == 0 || (means && scales && quats && opacities && colors), U3D_ERR_INVALID, "null model");
    U3D_CHECK(sh_degree >= 0 && sh_degree <= 3, U3D_ERR_INVALID, "sh_degree must be 0..3");
    U3D_CHECK(sh_degree == 0 || sh_rest != nullptr || n == 0, U3D_ERR_INVALID, "null sh_rest");
    const int K = shCoefficients(sh_degree);
    std::ostringstream h;
    h << "ply\nformat binary_little_endian 1.0\nelement vertex " << n << "\n";
    for (const char *p : {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"})
        h << "property float " << p << "\n";
    for (int k = 0; k < 3 * K; ++k)
        h << "property float f_rest_" << k << "\n";
    for (const char *p : {"opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"})
        h << "property float " << p << "\n";
    h << "end_header\n";
    std::ofstream out(path, std::ios::binary);
    U3D_CHECK(bool(out), U3D_ERR_INVALID, std::string("cannot write ") + path);
    const std::string header = h.str();
    out.write(header.data(), std::streamsize(header.size()));
    std::vector<char> buf;
    const size_t row = size_t(17 + 3 * K) * 4;
    for (int64_t i0 = 0; i0 < n; i0 += 65536) {
        const int64_t i1 = std::min<int64_t>(n, i0 + 65536);
        buf.clear();
        buf.reserve(size_t(i1 - i0) * row);
        for (int64_t i = i0; i < i1; ++i) {
            for (int k = 0; k < 3; ++k)
                putF32(buf, means[i * 3 + k]);
            for (int k = 0; k < 3; ++k)
                putF32(buf, 0.0f);
            for (int k = 0; k < 3; ++k)
                putF32(buf, colors[i * 3 + k]);
            for (int ch = 0; ch < 3; ++ch) // channel-major f_rest
                for (int c = 0; c < K; ++c)
                    putF32(buf, sh_rest[i * 45 + c * 3 + ch]);
            putF32(buf, opacities[i]);
            for (int k = 0; k < 3; ++k)
                putF32(buf, scales[i * 3 + k]);
            for (int k = 0; k < 4; ++k)
                putF32(buf, quats[i * 4 + k]);
        }
        out.write(buf.data(), std::streamsize(buf.size()));
    }
    U3D_CHECK(bool(out), U3D_ERR_INVALID, std::string("write error: ") + path);
    return U3D_OK;
}

int u3d_read_splat_ply(const char *path, int64_t capacity, float *means, float *scales,
                       float *quats, float *opacities, float *colors, float *sh_rest,
                       int64_t *n_out, int *sh_degree_out)
{
    U3D_CHECK(path != nullptr && n_out != nullptr, U3D_ERR_INVALID, "bad argument");
    std::ifstream in(path, std::ios::binary);
    U3D_CHECK(bool(in), U3D_ERR_INVALID, std::string("cannot read ") + path);
    std::string line;
    U3D_CHECK(std::getline(in, line) && line.rfind("ply", 0) == 0, U3D_ERR_INVALID, "not a PLY file");
    bool binary = false, vertexSeen = false, inVertex = false;
    int64_t count = 0;
    std::vector<PlyProp> props;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r')
            line.pop_back();
        std::istringstream ls(line);
        std::string w;
        ls >> w;
        if (w == "format") {
            std::string f;
            ls >> f;
            U3D_CHECK(f == "binary_little_endian" || f == "ascii", U3D_ERR_UNSUPPORTED,
                      "PLY format " + f + " not supported");
            binary = f != "ascii";
        } else if (w == "element") {
            std::string name;
            int64_t c = 0;
            ls >> name >> c;
            U3D_CHECK(!vertexSeen || name != "vertex", U3D_ERR_INVALID, "two vertex elements");
            U3D_CHECK(vertexSeen || name == "vertex", U3D_ERR_UNSUPPORTED,
                      "elements before the vertex element are not supported");
            inVertex = name == "vertex";
            if (inVertex) {
                vertexSeen = true;
                count = c;
            }
        } else if (w == "property" && inVertex) {
            std::string t, name;
            ls >> t;
            U3D_CHECK(t != "list", U3D_ERR_UNSUPPORTED, "list properties in vertices");
            ls >> name;
            PlyProp p;
            p.name = name;
            U3D_CHECK(typeInfo(t, p), U3D_ERR_UNSUPPORTED, "PLY property type " + t);
            props.push_back(p);
        } else if (w == "end_header") {
            break;
        }
    }
    U3D_CHECK(vertexSeen && count >= 0, U3D_ERR_INVALID, "PLY file has no vertex element");
    auto col = [&](const std::string &name) {
        for (size_t k = 0; k < props.size(); ++k)
            if (props[k].name == name)
                return int(k);
        return -1;
    };
    int rest = 0;
    while (col("f_rest_" + std::to_string(rest)) >= 0)
        ++rest;
    const int degree = degreeOfRest(rest);
    U3D_CHECK(degree >= 0, U3D_ERR_INVALID, "f_rest count is not 9, 24 or 45");
    const char *need[] = {"x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0",
                          "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"};
    int idx[14];
    for (int k = 0; k < 14; ++k) {
        idx[k] = col(need[k]);
        U3D_CHECK(idx[k] >= 0, U3D_ERR_INVALID, std::string("PLY vertices lack ") + need[k]);
    }
    *n_out = count;
    if (sh_degree_out)
        *sh_degree_out = degree;
    if (capacity == 0)
        return U3D_OK;
    U3D_CHECK(capacity >= count, U3D_ERR_CAPACITY, "arrays smaller than the vertex count");
    U3D_CHECK(count == 0 || (means && scales && quats && opacities && colors), U3D_ERR_INVALID,
              "null output");
    U3D_CHECK(degree == 0 || sh_rest || count == 0, U3D_ERR_INVALID, "null sh_rest");
    const int K = shCoefficients(degree);
    std::vector<int> restIdx(size_t(3 * K));
    for (int k = 0; k < 3 * K; ++k)
        restIdx[size_t(k)] = col("f_rest_" + std::to_string(k));
    size_t rowBytes = 0;
    std::vector<size_t> offset(props.size());
    for (size_t k = 0; k < props.size(); ++k) {
        offset[k] = rowBytes;
        rowBytes += size_t(props[k].size);
    }
    std::vector<double> v(props.size());
    std::vector<char> row(rowBytes);
    for (int64_t i = 0; i < count; ++i) {
        if (binary) {
            U3D_CHECK(bool(in.read(row.data(), std::streamsize(rowBytes))), U3D_ERR_INVALID,
                      "PLY file truncated");
            for (size_t k = 0; k < props.size(); ++k)
                v[k] = decode(row.data() + offset[k], props[k]);
        } else {
            for (size_t k = 0; k < props.size(); ++k)
                U3D_CHECK(bool(in >> v[k]), U3D_ERR_INVALID, "PLY file truncated");
        }
        for (int k = 0; k < 3; ++k) {
            means[i * 3 + k] = float(v[size_t(idx[k])]);
            colors[i * 3 + k] = float(v[size_t(idx[3 + k])]);
            scales[i * 3 + k] = float(v[size_t(idx[7 + k])]);
        }
        opacities[i] = float(v[size_t(idx[6])]);
        for (int k = 0; k < 4; ++k)
            quats[i * 4 + k] = float(v[size_t(idx[10 + k])]);
        for (int ch = 0; ch < 3; ++ch)
            for (int c = 0; c < K; ++c)
                sh_rest[i * 45 + c * 3 + ch] = float(v[size_t(restIdx[size_t(ch * K + c)])]);
        if (degree > 0)
            for (int c = K; c < 15; ++c)
                for (int ch = 0; ch < 3; ++ch)
                    sh_rest[i * 45 + c * 3 + ch] = 0.0f;
    }
    return U3D_OK;
}

int u3d_trainer_export_ply(u3d_trainer *t, const char *path)
{
    U3D_CHECK(t != nullptr && path != nullptr, U3D_ERR_INVALID, "null argument");
    const int64_t n = t->params.n;
    std::vector<float> h(size_t(n) * 14), sh(t->shDegree > 0 ? size_t(n) * 45 : 0);
    float *b = h.data();
    U3D_TRY(u3d_trainer_get_model(t, b, b + 3 * n, b + 6 * n, b + 10 * n, b + 11 * n));
    if (t->shDegree > 0)
        U3D_TRY(u3d_trainer_get_sh_rest(t, sh.data()));
    return u3d_write_splat_ply(path, b, b + 3 * n, b + 6 * n, b + 10 * n, b + 11 * n, n,
                               t->shDegree > 0 ? sh.data() : nullptr, t->shDegree);
}

} // extern "C"
