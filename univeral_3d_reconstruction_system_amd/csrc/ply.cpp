// Model and point-cloud files in the PLY layouts of engine/src/splat/SplatPly.cpp: Gaussian
// splats (writeSplatPly / readSplatPly, :261-342: x y z nx ny nz f_dc_0..2 opacity scale_0..2
// rot_0..3, binary little endian) and coloured points (writePointsPly / readPointsPly,
// :344-386: x y z red green blue). Spherical-harmonic coefficients beyond DC (an extension; the
// reference trains DC only) are written as the usual f_rest_0 .. f_rest_{3K-1}, channel-major
// (all red coefficients, then green, then blue) for K = 3, 8 or 15 per channel.
//
// The reader takes what the reference's reader takes: ASCII and binary files of either byte
// order, any elements before or after the vertex element (skipped, list properties included),
// list properties inside the vertex element (skipped), any scalar type per property. Vertices
// whose x, y or z is not finite are dropped, colours stored as uchar / ushort are scaled to
// 0..1, and a property the file lacks reads as NaN. Host-only code.
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "status.h"

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

// ------------------------------------------------------------------------------ reading
enum class Encoding { Ascii, Little, Big };

struct Scalar {
    int bytes = 0;
    char kind = 0; // 'f' float, 'i' signed integer, 'u' unsigned integer
};

bool scalarOf(const std::string &t, Scalar &s)
{
    struct Named {
        const char *a, *b;
        int bytes;
        char kind;
    };
    static const Named table[] = {
        {"char", "int8", 1, 'i'},   {"uchar", "uint8", 1, 'u'},   {"short", "int16", 2, 'i'},
        {"ushort", "uint16", 2, 'u'}, {"int", "int32", 4, 'i'},   {"uint", "uint32", 4, 'u'},
        {"float", "float32", 4, 'f'}, {"double", "float64", 8, 'f'}};
    for (const Named &n : table)
        if (t == n.a || t == n.b) {
            s.bytes = n.bytes;
            s.kind = n.kind;
            return true;
        }
    return false;
}

struct Field {
    std::string name;
    Scalar value;
    bool list = false;
    Scalar count; // list length type
};

struct Element {
    std::string name;
    int64_t count = 0;
    std::vector<Field> fields;
};

struct Header {
    Encoding enc = Encoding::Ascii;
    std::vector<Element> elements;
    size_t body = 0; // offset of the first byte after "end_header\n"
};

int parseHeader(const std::vector<char> &file, Header &h)
{
    const size_t probe = std::min<size_t>(file.size(), size_t(1) << 16);
    const std::string text(file.data(), probe);
    U3D_CHECK(text.compare(0, 3, "ply") == 0, U3D_ERR_INVALID, "not a PLY file");
    const size_t end = text.find("end_header");
    U3D_CHECK(end != std::string::npos, U3D_ERR_INVALID, "PLY header has no end_header");
    const size_t nl = text.find('\n', end);
    U3D_CHECK(nl != std::string::npos, U3D_ERR_INVALID, "truncated PLY header");
    h.body = nl + 1;
    std::istringstream lines(text.substr(0, end));
    std::string line;
    while (std::getline(lines, line)) {
        std::istringstream ls(line);
        std::string key;
        ls >> key;
        if (key == "format") {
            std::string f;
            ls >> f;
            if (f == "ascii")
                h.enc = Encoding::Ascii;
            else if (f == "binary_little_endian")
                h.enc = Encoding::Little;
            else if (f == "binary_big_endian")
                h.enc = Encoding::Big;
            else
                U3D_CHECK(false, U3D_ERR_UNSUPPORTED, "unknown PLY format " + f);
        } else if (key == "element") {
            Element e;
            ls >> e.name >> e.count;
            U3D_CHECK(!ls.fail() && e.count >= 0, U3D_ERR_INVALID, "bad PLY element line: " + line);
            h.elements.push_back(e);
        } else if (key == "property") {
            U3D_CHECK(!h.elements.empty(), U3D_ERR_INVALID, "PLY property before any element");
            Field f;
            std::string t;
            ls >> t;
            if (t == "list") {
                std::string ct, vt;
                ls >> ct >> vt >> f.name;
                f.list = true;
                U3D_CHECK(scalarOf(ct, f.count) && f.count.kind != 'f' && scalarOf(vt, f.value),
                          U3D_ERR_UNSUPPORTED, "unsupported PLY property: " + line);
            } else {
                ls >> f.name;
                U3D_CHECK(scalarOf(t, f.value), U3D_ERR_UNSUPPORTED, "unsupported PLY property: " + line);
            }
            h.elements.back().fields.push_back(f);
        }
    }
    return U3D_OK;
}

// Sequential reader of the body's values.
class Body
{
public:
    Body(const std::vector<char> &file, size_t offset, Encoding enc)
        : m_p(file.data() + offset), m_end(file.data() + file.size()), m_enc(enc)
    {
    }

    // Next value of type s as a double; false at the end of the data.
    bool next(const Scalar &s, double &out)
    {
        if (m_enc == Encoding::Ascii)
            return token(out);
        if (m_end - m_p < s.bytes)
            return false;
        unsigned char b[8];
        std::memcpy(b, m_p, size_t(s.bytes));
        m_p += s.bytes;
        if (m_enc == Encoding::Big) // this host is little endian
            for (int i = 0; i < s.bytes / 2; ++i)
                std::swap(b[i], b[s.bytes - 1 - i]);
        if (s.kind == 'f') {
            if (s.bytes == 4) {
                float v;
                std::memcpy(&v, b, 4);
                out = v;
            } else {
                std::memcpy(&out, b, 8);
            }
            return true;
        }
        uint64_t u = 0;
        std::memcpy(&u, b, size_t(s.bytes)); // low bytes of a little-endian word
        if (s.kind == 'u') {
            out = double(u);
        } else {
            const int shift = 64 - 8 * s.bytes;
            out = double(int64_t(u << shift) >> shift);
        }
        return true;
    }

    bool skip(const Field &f)
    {
        double v;
        if (!f.list)
            return m_enc == Encoding::Ascii ? token(v) : advance(f.value.bytes);
        double n;
        if (!next(f.count, n) || !(n >= 0.0 && n <= 4294967295.0)) // a 32-bit count at most
            return false;
        for (int64_t k = 0; k < int64_t(n); ++k)
            if (!(m_enc == Encoding::Ascii ? token(v) : advance(f.value.bytes)))
                return false;
        return true;
    }

private:
public:
    size_t remaining() const { return size_t(m_end - m_p); }

    // Fewest bytes one row of e can take: every binary value its size, every ASCII value one
    // character (0 when e has no properties).
    size_t minRowBytes(const Element &e) const
    {
        size_t b = 0;
        for (const Field &f : e.fields)
            b += m_enc == Encoding::Ascii ? 1 : size_t(f.list ? f.count.bytes : f.value.bytes);
        return b;
    }

private:
    bool advance(int bytes)
    {
        if (m_end - m_p < bytes)
            return false;
        m_p += bytes;
        return true;
    }

    bool token(double &out)
    {
        while (m_p < m_end && std::isspace(static_cast<unsigned char>(*m_p)))
            ++m_p;
        const char *start = m_p;
        while (m_p < m_end && !std::isspace(static_cast<unsigned char>(*m_p)))
            ++m_p;
        if (m_p == start)
            return false;
        const char *s = start;
        if (*s == '+') // from_chars does not take a leading '+'
            ++s;
        // numbers only, as a C++ stream reads them: "nan" / "inf" and values out of the double
        // range make the file unreadable, as in the reference
        const char *d = (*s == '-') ? s + 1 : s;
        if (d < m_p && std::isalpha(static_cast<unsigned char>(*d)))
            return false;
        const std::from_chars_result r = std::from_chars(s, m_p, out);
        return r.ec == std::errc() && r.ptr == m_p;
    }

    const char *m_p;
    const char *m_end;
    Encoding m_enc;
};

// double -> float as IEEE round-to-nearest gives it, also beyond the float range (where a C++
// conversion is undefined): up to half an ulp (2^103) above the largest float rounds to it, from
// there on (the tie rounds to even) to infinity.
float toFloat(double v)
{
    constexpr float kMaxF = std::numeric_limits<float>::max();
    constexpr double kMax = double(kMaxF), kInf = kMax + 0x1p103;
    if (v >= kInf || v <= -kInf)
        return v > 0 ? std::numeric_limits<float>::infinity() : -std::numeric_limits<float>::infinity();
    if (v > kMax || v < -kMax)
        return v > 0 ? kMaxF : -kMaxF;
    return float(v);
}

// The vertex element's columns `names`: values [rows x names] (float, NaN where the file has no
// such property), present[k] whether the file has names[k]. Rows with a non-finite x, y or z
// (names 0..2 must be x y z) are dropped.
struct VertexColumns {
    int64_t rows = 0;
    std::vector<bool> present;
    std::vector<float> values;
};

int readFile(const char *path, std::vector<char> &file)
{
    std::ifstream in(path, std::ios::binary);
    U3D_CHECK(bool(in), U3D_ERR_INVALID, std::string("cannot read ") + path);
    file.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    return U3D_OK;
}

const Element *vertexElement(const Header &h)
{
    for (const Element &e : h.elements)
        if (e.name == "vertex")
            return &e;
    return nullptr;
}

int readVertexColumns(const char *path, const std::vector<std::string> &names, VertexColumns &out)
{
    std::vector<char> file;
    U3D_TRY(readFile(path, file));
    Header h;
    U3D_TRY(parseHeader(file, h));
    U3D_CHECK(vertexElement(h) != nullptr, U3D_ERR_INVALID, "PLY file has no vertex element");
    Body body(file, h.body, h.enc);
    const size_t W = names.size();
    for (const Element &e : h.elements) {
        // a count the rest of the file cannot hold is refused before anything is sized by it
        const size_t minRow = body.minRowBytes(e);
        U3D_CHECK(minRow == 0 || uint64_t(e.count) <= body.remaining() / minRow, U3D_ERR_INVALID,
                  "PLY data truncated: element " + e.name + " claims more rows than the file holds");
        if (e.name != "vertex") {
            if (minRow == 0) // rows of no properties take no bytes
                continue;
            for (int64_t i = 0; i < e.count; ++i)
                for (const Field &f : e.fields)
                    U3D_CHECK(body.skip(f), U3D_ERR_INVALID, "PLY data truncated in element " + e.name);
            continue;
        }
        // which wanted column each field feeds (-1: none), and the colour scale of that field
        std::vector<int> target(e.fields.size(), -1);
        std::vector<double> scale(e.fields.size(), 1.0);
        out.present.assign(W, false);
        for (size_t f = 0; f < e.fields.size(); ++f) {
            if (e.fields[f].list)
                continue;
            for (size_t k = 0; k < W; ++k)
                if (e.fields[f].name == names[k]) {
                    target[f] = int(k);
                    out.present[k] = true;
                    const bool colour = names[k] == "red" || names[k] == "green" || names[k] == "blue";
                    if (colour && e.fields[f].value.kind == 'u' && e.fields[f].value.bytes == 1)
                        scale[f] = 255.0;
                    else if (colour && e.fields[f].value.kind == 'u' && e.fields[f].value.bytes == 2)
                        scale[f] = 65535.0;
                }
        }
        U3D_CHECK(minRow > 0, U3D_ERR_INVALID, "PLY vertices have no properties");
        const float nan = std::numeric_limits<float>::quiet_NaN();
        out.values.assign(size_t(e.count) * W, nan);
        std::vector<float> row(W);
        int64_t kept = 0;
        for (int64_t i = 0; i < e.count; ++i) {
            std::fill(row.begin(), row.end(), nan);
            for (size_t f = 0; f < e.fields.size(); ++f) {
                if (target[f] < 0) {
                    U3D_CHECK(body.skip(e.fields[f]), U3D_ERR_INVALID,
                              "PLY data truncated at vertex " + std::to_string(i));
                    continue;
                }
                double v;
                U3D_CHECK(body.next(e.fields[f].value, v), U3D_ERR_INVALID,
                          "PLY data truncated at vertex " + std::to_string(i));
                if (scale[f] != 1.0)
                    v /= scale[f];
                row[size_t(target[f])] = toFloat(v);
            }
            if (!std::isfinite(row[0]) || !std::isfinite(row[1]) || !std::isfinite(row[2]))
                continue;
            std::copy(row.begin(), row.end(), out.values.begin() + ptrdiff_t(size_t(kept) * W));
            ++kept;
        }
        out.rows = kept;
        out.values.resize(size_t(kept) * W);
        return U3D_OK;
    }
    return U3D_OK; // not reached: the vertex element exists
}

bool isSplatField(const std::string &name)
{
    // the fields whose presence makes a file a splat file (readSplatPly, SplatPly.cpp:295)
    return name == "f_dc_0" || name == "opacity" || name == "scale_0" || name == "rot_0";
}

// ------------------------------------------------------------------------------ writing
void putF32(std::vector<char> &buf, float v)
{
    char b[4];
    std::memcpy(b, &v, 4); // the host is little endian (x86-64)
    buf.insert(buf.end(), b, b + 4);
}

int writeBody(const char *path, const std::string &header, const std::vector<char> &rows)
{
    std::ofstream out(path, std::ios::binary);
    U3D_CHECK(bool(out), U3D_ERR_INVALID, std::string("cannot write ") + path);
    out.write(header.data(), std::streamsize(header.size()));
    out.write(rows.data(), std::streamsize(rows.size()));
    U3D_CHECK(bool(out), U3D_ERR_INVALID, std::string("write error: ") + path);
    return U3D_OK;
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
    std::vector<char> buf;
    buf.reserve(size_t(n) * size_t(17 + 3 * K) * 4);
    for (int64_t i = 0; i < n; ++i) {
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
    return writeBody(path, h.str(), buf);
}

int u3d_ply_info(const char *path, int64_t *vertices, int *is_splat, int *sh_degree)
{
    U3D_CHECK(path != nullptr, U3D_ERR_INVALID, "null path");
    std::vector<char> file;
    U3D_TRY(readFile(path, file));
    Header h;
    U3D_TRY(parseHeader(file, h));
    const Element *v = vertexElement(h);
    U3D_CHECK(v != nullptr, U3D_ERR_INVALID, "PLY file has no vertex element");
    int splatFields = 0, rest = 0;
    bool x = false, y = false, z = false;
    for (const Field &f : v->fields) {
        if (f.list)
            continue;
        splatFields += isSplatField(f.name) ? 1 : 0;
        x |= f.name == "x";
        y |= f.name == "y";
        z |= f.name == "z";
    }
    U3D_CHECK(x && y && z, U3D_ERR_INVALID, "PLY vertices have no x/y/z");
    auto has = [&](const std::string &name) {
        for (const Field &f : v->fields)
            if (!f.list && f.name == name)
                return true;
        return false;
    };
    while (has("f_rest_" + std::to_string(rest)))
        ++rest;
    if (vertices)
        *vertices = v->count;
    if (is_splat)
        *is_splat = splatFields == 4 ? 1 : 0;
    if (sh_degree)
        *sh_degree = std::max(0, degreeOfRest(rest));
    return U3D_OK;
}

int u3d_read_splat_ply(const char *path, int64_t capacity, float *means, float *scales,
                       float *quats, float *opacities, float *colors, float *sh_rest,
                       int64_t *n_out, int *sh_degree_out)
{
    U3D_CHECK(path != nullptr && n_out != nullptr, U3D_ERR_INVALID, "bad argument");
    int64_t vertices = 0;
    int splat = 0, degree = 0;
    U3D_TRY(u3d_ply_info(path, &vertices, &splat, &degree));
    U3D_CHECK(splat, U3D_ERR_INVALID, "not a Gaussian splat file (no f_dc_0 / opacity / scale_0 / rot_0)");
    const int K = shCoefficients(degree);
    std::vector<std::string> names = {"x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity",
                                      "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"};
    for (int k = 0; k < 3 * K; ++k)
        names.push_back("f_rest_" + std::to_string(k));
    VertexColumns t;
    U3D_TRY(readVertexColumns(path, names, t));
    *n_out = t.rows;
    if (sh_degree_out)
        *sh_degree_out = degree;
    if (capacity == 0)
        return U3D_OK;
    U3D_CHECK(capacity >= t.rows, U3D_ERR_CAPACITY, "arrays smaller than the vertex count");
    U3D_CHECK(t.rows == 0 || (means && scales && quats && opacities && colors), U3D_ERR_INVALID,
              "null output");
    U3D_CHECK(degree == 0 || sh_rest || t.rows == 0, U3D_ERR_INVALID, "null sh_rest");
    const size_t W = names.size();
    for (int64_t i = 0; i < t.rows; ++i) {
        const float *r = &t.values[size_t(i) * W];
        for (int k = 0; k < 3; ++k) {
            means[i * 3 + k] = r[k];
            colors[i * 3 + k] = r[3 + k];
            scales[i * 3 + k] = r[7 + k];
        }
        opacities[i] = r[6];
        for (int k = 0; k < 4; ++k)
            quats[i * 4 + k] = r[10 + k];
        if (degree > 0) {
            for (int c = 0; c < 15; ++c)
                for (int ch = 0; ch < 3; ++ch)
                    sh_rest[i * 45 + c * 3 + ch] = c < K ? r[14 + ch * K + c] : 0.0f;
        }
    }
    return U3D_OK;
}

int u3d_write_points_ply(const char *path, const float *xyz_host, const float *rgb_host,
                         int64_t rgb_count, int64_t n)
{
    // writePointsPly, SplatPly.cpp:344-362: colours as uchar round(clamp(c, 0, 1) * 255), 0.8
    // where rgb has no entry for a point
    U3D_CHECK(path != nullptr && n >= 0 && rgb_count >= 0, U3D_ERR_INVALID, "bad argument");
    U3D_CHECK(n == 0 || xyz_host, U3D_ERR_INVALID, "null points");
    U3D_CHECK(rgb_count == 0 || rgb_host, U3D_ERR_INVALID, "null colours");
    std::ostringstream h;
    h << "ply\nformat binary_little_endian 1.0\nelement vertex " << n << "\n"
      << "property float x\nproperty float y\nproperty float z\n"
      << "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n";
    std::vector<char> buf;
    buf.reserve(size_t(n) * 15);
    for (int64_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k)
            putF32(buf, xyz_host[i * 3 + k]);
        for (int k = 0; k < 3; ++k) {
            const float c = rgb_count >= (i + 1) * 3 ? rgb_host[i * 3 + k] : 0.8f;
            const float clamped = std::min(1.0f, std::max(0.0f, c));
            buf.push_back(char(static_cast<unsigned char>(std::lround(clamped * 255.0f))));
        }
    }
    return writeBody(path, h.str(), buf);
}

int u3d_read_points_ply(const char *path, int64_t capacity, float *xyz_host, float *rgb_host,
                        int64_t *n_out)
{
    // readPointsPly, SplatPly.cpp:364-386: rgb 0..1 (0.8 when the file has no colours)
    U3D_CHECK(path != nullptr && n_out != nullptr, U3D_ERR_INVALID, "bad argument");
    VertexColumns t;
    U3D_TRY(readVertexColumns(path, {"x", "y", "z", "red", "green", "blue"}, t));
    U3D_CHECK(t.present[0] && t.present[1] && t.present[2], U3D_ERR_INVALID, "PLY vertices have no x/y/z");
    *n_out = t.rows;
    if (capacity == 0)
        return U3D_OK;
    U3D_CHECK(capacity >= t.rows, U3D_ERR_CAPACITY, "arrays smaller than the point count");
    U3D_CHECK(t.rows == 0 || (xyz_host && rgb_host), U3D_ERR_INVALID, "null output");
    const bool colours = t.present[3];
    for (int64_t i = 0; i < t.rows; ++i) {
        const float *r = &t.values[size_t(i) * 6];
        for (int k = 0; k < 3; ++k) {
            xyz_host[i * 3 + k] = r[k];
            rgb_host[i * 3 + k] = colours ? r[3 + k] : 0.8f;
        }
    }
    return U3D_OK;
}

} // extern "C"
