// The PLY readers (csrc/ply.cpp) on hostile input, built with AddressSanitizer and
// UndefinedBehaviorSanitizer: valid files of every layout the readers take (the writers' binary
// files with and without SH coefficients, point clouds, ASCII and big-endian files with other
// elements and list properties), then deterministic mutations of each -- flipped bytes,
// truncations, vertex / list counts far beyond the data, changed property types, injected
// header lines. Every file goes through u3d_ply_info, u3d_read_splat_ply and
// u3d_read_points_ply; a reader may refuse a file but must not touch memory out of bounds,
// overflow, allocate by an unchecked count or loop without end. Exits 0 when every valid seed
// reads back and every call returned.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../include/u3d_hip.h"

namespace u3d {
void setError(const std::string &) {}
} // namespace u3d

namespace {

std::vector<char> slurp(const std::string &path)
{
    std::ifstream in(path, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

void spit(const std::string &path, const std::vector<char> &b)
{
    std::ofstream out(path, std::ios::binary);
    out.write(b.data(), std::streamsize(b.size()));
}

void spit(const std::string &path, const std::string &s) { spit(path, std::vector<char>(s.begin(), s.end())); }

// Reads a file every way; returns the rows read by the splat reader (-1 when refused).
int64_t readAll(const std::string &path)
{
    int64_t vertices = 0, n = 0;
    int splat = 0, degree = 0;
    u3d_ply_info(path.c_str(), &vertices, &splat, &degree);
    int64_t rows = -1;
    if (u3d_read_splat_ply(path.c_str(), 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &n,
                           &degree) == U3D_OK && n >= 0 && n < (int64_t(1) << 22)) {
        const size_t m = size_t(n > 0 ? n : 1);
        std::vector<float> means(3 * m), scales(3 * m), quats(4 * m), op(m), col(3 * m), rest(45 * m);
        int64_t got = 0;
        if (u3d_read_splat_ply(path.c_str(), int64_t(m), means.data(), scales.data(), quats.data(),
                               op.data(), col.data(), rest.data(), &got, &degree) == U3D_OK)
            rows = got;
    }
    if (u3d_read_points_ply(path.c_str(), 0, nullptr, nullptr, &n) == U3D_OK && n >= 0 && n < (int64_t(1) << 22)) {
        const size_t m = size_t(n > 0 ? n : 1);
        std::vector<float> xyz(3 * m), rgb(3 * m);
        int64_t got = 0;
        u3d_read_points_ply(path.c_str(), int64_t(m), xyz.data(), rgb.data(), &got);
        // a capacity below the count is refused, not overrun
        if (n > 1)
            u3d_read_points_ply(path.c_str(), n - 1, xyz.data(), rgb.data(), &got);
    }
    return rows;
}

std::vector<std::string> seeds(const std::string &dir, std::mt19937 &rng)
{
    std::vector<std::string> out;
    std::uniform_real_distribution<float> u(-2.0f, 2.0f);
    const int n = 37;
    std::vector<float> means(3 * n), scales(3 * n), quats(4 * n), op(n), col(3 * n), rest(45 * n);
    for (auto *v : {&means, &scales, &quats, &op, &col, &rest})
        for (float &x : *v)
            x = u(rng);
    for (int degree : {0, 1, 3}) {
        const std::string p = dir + "/seed_splat" + std::to_string(degree) + ".ply";
        u3d_write_splat_ply(p.c_str(), means.data(), scales.data(), quats.data(), op.data(), col.data(), n,
                            rest.data(), degree);
        out.push_back(p);
    }
    {
        const std::string p = dir + "/seed_points.ply";
        u3d_write_points_ply(p.c_str(), means.data(), col.data(), 3 * n, n);
        out.push_back(p);
    }
    {
        const std::string p = dir + "/seed_ascii.ply";
        std::string s = "ply\nformat ascii 1.0\ncomment hand made\nelement face 2\n"
                        "property list uchar int vertex_indices\nelement vertex 3\nproperty double x\n"
                        "property float y\nproperty float z\nproperty uchar red\nproperty uchar green\n"
                        "property uchar blue\nproperty list uchar float extra\nproperty float f_dc_0\n"
                        "property float f_dc_1\nproperty float f_dc_2\nproperty float opacity\n"
                        "property float scale_0\nproperty float scale_1\nproperty float scale_2\n"
                        "property float rot_0\nproperty float rot_1\nproperty float rot_2\nproperty float rot_3\n"
                        "element tail 1\nproperty int marker\nend_header\n"
                        "3 0 1 2\n3 2 1 0\n";
        for (int i = 0; i < 3; ++i)
            s += "0.5 -1 2 255 0 7 2 1.5 2.5 0.1 0.2 0.3 -1 -2 -3 -4 1 0 0 0\n";
        s += "42\n";
        spit(p, s);
        out.push_back(p);
    }
    {
        const std::string p = dir + "/seed_big.ply";
        std::string head = "ply\nformat binary_big_endian 1.0\nelement vertex 4\nproperty float x\n"
                           "property float y\nproperty float z\nproperty list uchar short idx\n"
                           "property ushort red\nproperty short green\nproperty uint blue\nend_header\n";
        std::vector<char> b(head.begin(), head.end());
        for (int i = 0; i < 4; ++i) {
            for (int k = 0; k < 3; ++k) {
                float v = float(i + k);
                unsigned char w[4];
                std::memcpy(w, &v, 4);
                for (int j = 3; j >= 0; --j)
                    b.push_back(char(w[j]));
            }
            b.push_back(char(2));
            for (int j = 0; j < 4; ++j)
                b.push_back(char(j));
            for (int j = 0; j < 2 + 2 + 4; ++j)
                b.push_back(char(j * 17));
        }
        spit(p, b);
        out.push_back(p);
    }
    return out;
}

std::vector<char> mutate(const std::vector<char> &in, std::mt19937 &rng)
{
    std::vector<char> b = in;
    const std::string text(b.begin(), b.end());
    const size_t headEnd = std::min(text.find("end_header"), b.size());
    auto pick = [&](size_t n) { return n ? size_t(rng() % n) : 0; };
    switch (rng() % 9) {
    case 0: // flip bytes anywhere
        for (int k = 0; k < 1 + int(rng() % 8); ++k)
            if (!b.empty())
                b[pick(b.size())] ^= char(1 << (rng() % 8));
        break;
    case 1: // truncate
        b.resize(pick(b.size() + 1));
        break;
    case 2: { // a vertex count far beyond the data
        static const char *counts[] = {"4294967296", "9223372036854775807", "-1", "99999999999", "0", "1e9"};
        std::string t(b.begin(), b.end());
        const size_t at = t.find("element vertex ");
        if (at != std::string::npos) {
            const size_t s = at + 15, e = t.find('\n', s);
            t.replace(s, e - s, counts[rng() % 6]);
        }
        b.assign(t.begin(), t.end());
        break;
    }
    case 3: { // change a property type
        static const char *types[] = {"double", "uchar", "char", "int", "list uint float", "list uchar uchar", "bogus"};
        std::string t(b.begin(), b.end());
        size_t at = t.find("property float");
        for (int k = int(rng() % 6); k > 0 && at != std::string::npos; --k)
            at = t.find("property float", at + 1);
        if (at != std::string::npos && at < headEnd)
            t.replace(at + 9, 5, types[rng() % 7]);
        b.assign(t.begin(), t.end());
        break;
    }
    case 4: { // inject a header line
        static const char *lines[] = {"element face 4294967295\nproperty list uint int v\n",
                                      "element empty 99999999999999\n", "property float orphan\n",
                                      "format binary_little_endian 1.0\n", "element vertex 1\n"};
        std::string t(b.begin(), b.end());
        const size_t nl = t.find('\n');
        if (nl != std::string::npos)
            t.insert(nl + 1, lines[rng() % 5]);
        b.assign(t.begin(), t.end());
        break;
    }
    case 5: // garbage body
        for (size_t k = headEnd; k < b.size(); ++k)
            if (rng() % 4 == 0)
                b[k] = char(rng());
        break;
    case 6: // huge list counts: set every byte after the header to 0xff
        for (size_t k = headEnd; k < b.size(); ++k)
            if (rng() % 3 == 0)
                b[k] = char(0xff);
        break;
    case 7: { // ASCII numbers out of range
        static const char *vals[] = {"1e400", "-1e39", "nan", "inf", "+", "-", "1e-400", "0x10", "4294967296"};
        std::string t(b.begin(), b.end());
        const size_t at = t.find("end_header\n");
        if (at != std::string::npos && at + 11 < t.size())
            t.insert(at + 11 + pick(t.size() - at - 11), std::string(" ") + vals[rng() % 9] + " ");
        b.assign(t.begin(), t.end());
        break;
    }
    default: // duplicate the body
        b.insert(b.end(), b.begin() + ptrdiff_t(headEnd), b.end());
        break;
    }
    return b;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: ply_fuzz DIR ITERATIONS\n");
        return 2;
    }
    const std::string dir = argv[1];
    const int iterations = std::atoi(argv[2]);
    std::mt19937 rng(12345);
    const std::vector<std::string> files = seeds(dir, rng);
    const int64_t want[] = {37, 37, 37, -1, 3, -1};
    for (size_t k = 0; k < files.size(); ++k) {
        const int64_t rows = readAll(files[k]);
        if (rows != want[k]) {
            std::fprintf(stderr, "seed %s read %lld rows, expected %lld\n", files[k].c_str(),
                         (long long)rows, (long long)want[k]);
            return 1;
        }
    }
    const std::string scratch = dir + "/mutant.ply";
    for (int it = 0; it < iterations; ++it) {
        const std::vector<char> base = slurp(files[size_t(it) % files.size()]);
        std::vector<char> b = mutate(base, rng);
        if (rng() % 3 == 0)
            b = mutate(b, rng);
        spit(scratch, b);
        readAll(scratch);
    }
    std::printf("ply_fuzz: %zu seeds, %d mutants\n", files.size(), iterations);
    return 0;
}
