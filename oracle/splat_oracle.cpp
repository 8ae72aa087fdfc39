// CPU oracle for the Gaussian Splatting trainer path -- TEST INFRASTRUCTURE ONLY (never part
// of the product; see splat_oracle.h). Serial C++17 restatement of engine/src/splat, written
// from the reference's maths with the same floating-point operation order; compile with
// -ffp-contract=off and without -ffast-math.
//
// How each function is pinned (tests/test_oracle_*.py):
//   orc_project          compiled reference (oracle/_ref) radii + images, golden fixtures,
//                        closed-form ellipse moments (test_splat.cpp:153-191), rect formula
//   orc_bin_tiles        golden tile lists from the compiled reference; brute force per tile
//   orc_raster_forward   compiled reference images bit-exact; FP64 brute-force compositing
//   orc_raster_backward  compiled reference (1 thread) bit-exact; FP64 finite differences
//   orc_render_f64       closed-form single Gaussian; agreement with the FP32 path
//   orc_photometric_loss compiled reference bit-exact; FP64 direct 11x11 convolution SSIM;
//                        finite differences; SSIM(x, x) = 1
//   orc_adam_step        compiled reference bit-exact; closed form of the first step
//   orc_densify          compiled reference densifyAndPrune with injected statistics
//   orc_init_from_points compiled reference bit-exact; brute-force 3-NN on separated points
//   orc_trainer_*        compiled reference Trainer (1 thread) bit-exact over many steps
#include "splat_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <unordered_map>
#include <vector>

namespace {

constexpr float kShC0 = 0.28209479177387814f;  // GaussianModel.h:10
constexpr float kMaxAlpha = 0.99f;             // Rasterizer.cpp:13
constexpr int kG2 = 9;                         // Rasterizer.cpp:14 (mx my A B C r g b alpha)

inline float sigmoidf_(float x) { return 1.0f / (1.0f + std::exp(-x)); } // GaussianModel.h:11
inline float logitf_(float p) { return std::log(p / (1.0f - p)); }        // GaussianModel.h:12

template <typename T> inline T ex(T x) { return std::exp(x); }

template <typename T> struct Cam {
    int width, height;
    T fx, fy, cx, cy, R[9], t[3];
};
template <typename T> Cam<T> camOf(const orc_camera &c)
{
    Cam<T> r;
    r.width = c.width;
    r.height = c.height;
    r.fx = T(c.fx); r.fy = T(c.fy); r.cx = T(c.cx); r.cy = T(c.cy);
    for (int k = 0; k < 9; ++k) r.R[k] = T(c.R[k]);
    for (int k = 0; k < 3; ++k) r.t[k] = T(c.t[k]);
    return r;
}

// Everything one Gaussian's projection computes (Rasterizer.cpp:31-45).
template <typename T> struct Proj {
    T pc[3], q[4], qNorm, Rq[9], s[3], M[9], Sigma[9];
    T tx, ty, limTx, limTy;
    bool clampX, clampY;
    T T0[3], T1[3];
    T a, b, c;
};

// Rasterizer.cpp:24-29
template <typename T> inline void rotOfQuat(T w, T x, T y, T z, T R[9])
{
    R[0] = T(1) - T(2) * (y * y + z * z);
    R[1] = T(2) * (x * y - w * z);
    R[2] = T(2) * (x * z + w * y);
    R[3] = T(2) * (x * y + w * z);
    R[4] = T(1) - T(2) * (x * x + z * z);
    R[5] = T(2) * (y * z - w * x);
    R[6] = T(2) * (x * z - w * y);
    R[7] = T(2) * (y * z + w * x);
    R[8] = T(1) - T(2) * (x * x + y * y);
}

// EWA projection of one Gaussian, Rasterizer.cpp:48-110. False if behind the near plane.
template <typename T>
bool projectOne(const T *mu, const T *qr, const T *logScale, const Cam<T> &cam, T nearPlane,
                T eps2d, Proj<T> &p)
{
    const T *R = cam.R;
    p.pc[0] = R[0] * mu[0] + R[1] * mu[1] + R[2] * mu[2] + cam.t[0];
    p.pc[1] = R[3] * mu[0] + R[4] * mu[1] + R[5] * mu[2] + cam.t[1];
    p.pc[2] = R[6] * mu[0] + R[7] * mu[1] + R[8] * mu[2] + cam.t[2];
    const T z = p.pc[2];
    if (z < nearPlane)
        return false;
    p.qNorm = std::sqrt(qr[0] * qr[0] + qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
    if (p.qNorm < T(1e-12f)) {
        p.qNorm = T(1);
        p.q[0] = T(1);
        p.q[1] = p.q[2] = p.q[3] = T(0);
    } else {
        for (int k = 0; k < 4; ++k)
            p.q[k] = qr[k] / p.qNorm;
    }
    rotOfQuat(p.q[0], p.q[1], p.q[2], p.q[3], p.Rq);
    for (int k = 0; k < 3; ++k)
        p.s[k] = ex(logScale[k]);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            p.M[r * 3 + c] = p.Rq[r * 3 + c] * p.s[c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            p.Sigma[r * 3 + c] = p.M[r * 3] * p.M[c * 3] + p.M[r * 3 + 1] * p.M[c * 3 + 1]
                                 + p.M[r * 3 + 2] * p.M[c * 3 + 2];
    const T guard = T(0.3f) * T(0.5f);
    const T W = T(cam.width), H = T(cam.height);
    const T limXPos = (W - cam.cx) / cam.fx + guard * W / cam.fx;
    const T limXNeg = cam.cx / cam.fx + guard * W / cam.fx;
    const T limYPos = (H - cam.cy) / cam.fy + guard * H / cam.fy;
    const T limYNeg = cam.cy / cam.fy + guard * H / cam.fy;
    const T xz = p.pc[0] / z, yz = p.pc[1] / z;
    const T cxz = std::min(limXPos, std::max(-limXNeg, xz));
    const T cyz = std::min(limYPos, std::max(-limYNeg, yz));
    p.clampX = cxz != xz;
    p.clampY = cyz != yz;
    p.limTx = cxz;
    p.limTy = cyz;
    p.tx = z * cxz;
    p.ty = z * cyz;
    const T J00 = cam.fx / z, J02 = -cam.fx * p.tx / (z * z);
    const T J11 = cam.fy / z, J12 = -cam.fy * p.ty / (z * z);
    for (int k = 0; k < 3; ++k) {
        p.T0[k] = J00 * R[k] + J02 * R[6 + k];
        p.T1[k] = J11 * R[3 + k] + J12 * R[6 + k];
    }
    T S0[3], S1[3];
    for (int r = 0; r < 3; ++r) {
        S0[r] = p.Sigma[r * 3] * p.T0[0] + p.Sigma[r * 3 + 1] * p.T0[1] + p.Sigma[r * 3 + 2] * p.T0[2];
        S1[r] = p.Sigma[r * 3] * p.T1[0] + p.Sigma[r * 3 + 1] * p.T1[1] + p.Sigma[r * 3 + 2] * p.T1[2];
    }
    p.a = p.T0[0] * S0[0] + p.T0[1] * S0[1] + p.T0[2] * S0[2] + eps2d;
    p.b = p.T0[0] * S1[0] + p.T0[1] * S1[1] + p.T0[2] * S1[2];
    p.c = p.T1[0] * S1[0] + p.T1[1] * S1[1] + p.T1[2] * S1[2] + eps2d;
    return true;
}

// Per-Gaussian screen-space state (Rasterizer.h:66-76).
template <typename T> struct Screen {
    std::vector<T> mean2d, conic, cov2d, depth, rgb, alpha;
    std::vector<int> radius, rect;
    std::vector<uint8_t> flags;
    void resize(size_t n)
    {
        mean2d.resize(n * 2);
        conic.resize(n * 3);
        cov2d.resize(n * 3);
        depth.resize(n);
        rgb.resize(n * 3);
        alpha.resize(n);
        radius.assign(n, 0);
        rect.resize(n * 4);
        flags.assign(n, 0);
    }
};

// ---- Spherical harmonics of degree 1..3 (build-side extension; the reference trains the DC
// term only, GaussianModel.h:21). Real SH basis of a unit direction with the standard
// normalisation constants sqrt((2l+1)/(4 pi) (l-|m|)!/(l+|m|)!) (signs as in the common 3DGS
// convention), and its partial derivatives, written out from the polynomials.
constexpr double kShC1 = 0.4886025119029199;
constexpr double kShC2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                             -1.0925484305920792, 0.5462742152960396};
constexpr double kShC3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
                             0.3731763325901154, -0.4570457994644658, 1.445305721320277,
                             -0.5900435899266435};

inline int shCount(int degree) { return degree <= 0 ? 0 : (degree == 1 ? 3 : (degree == 2 ? 8 : 15)); }

template <typename T>
void shBasis(T x, T y, T z, int degree, T b[15], T gx[15], T gy[15], T gz[15])
{
    const T c1 = T(float(kShC1));
    const T xx = x * x, yy = y * y, zz = z * z;
    if (degree >= 1) {
        b[0] = -c1 * y; b[1] = c1 * z; b[2] = -c1 * x;
        gx[0] = T(0); gy[0] = -c1; gz[0] = T(0);
        gx[1] = T(0); gy[1] = T(0); gz[1] = c1;
        gx[2] = -c1; gy[2] = T(0); gz[2] = T(0);
    }
    if (degree >= 2) {
        const T k0 = T(float(kShC2[0])), k1 = T(float(kShC2[1])), k2 = T(float(kShC2[2])),
                k3 = T(float(kShC2[3])), k4 = T(float(kShC2[4]));
        b[3] = k0 * (x * y);
        b[4] = k1 * (y * z);
        b[5] = k2 * (T(2) * zz - xx - yy);
        b[6] = k3 * (x * z);
        b[7] = k4 * (xx - yy);
        gx[3] = k0 * y; gy[3] = k0 * x; gz[3] = T(0);
        gx[4] = T(0); gy[4] = k1 * z; gz[4] = k1 * y;
        gx[5] = k2 * (T(-2) * x); gy[5] = k2 * (T(-2) * y); gz[5] = k2 * (T(4) * z);
        gx[6] = k3 * z; gy[6] = T(0); gz[6] = k3 * x;
        gx[7] = k4 * (T(2) * x); gy[7] = k4 * (T(-2) * y); gz[7] = T(0);
    }
    if (degree >= 3) {
        T k[7];
        for (int j = 0; j < 7; ++j)
            k[j] = T(float(kShC3[j]));
        b[8] = k[0] * (y * (T(3) * xx - yy));
        b[9] = k[1] * (x * y * z);
        b[10] = k[2] * (y * (T(4) * zz - xx - yy));
        b[11] = k[3] * (z * (T(2) * zz - T(3) * xx - T(3) * yy));
        b[12] = k[4] * (x * (T(4) * zz - xx - yy));
        b[13] = k[5] * (z * (xx - yy));
        b[14] = k[6] * (x * (xx - T(3) * yy));
        gx[8] = k[0] * (T(6) * x * y); gy[8] = k[0] * (T(3) * xx - T(3) * yy); gz[8] = T(0);
        gx[9] = k[1] * (y * z); gy[9] = k[1] * (x * z); gz[9] = k[1] * (x * y);
        gx[10] = k[2] * (T(-2) * x * y); gy[10] = k[2] * (T(4) * zz - xx - T(3) * yy); gz[10] = k[2] * (T(8) * y * z);
        gx[11] = k[3] * (T(-6) * x * z); gy[11] = k[3] * (T(-6) * y * z); gz[11] = k[3] * (T(6) * zz - T(3) * xx - T(3) * yy);
        gx[12] = k[4] * (T(4) * zz - T(3) * xx - yy); gy[12] = k[4] * (T(-2) * x * y); gz[12] = k[4] * (T(8) * x * z);
        gx[13] = k[5] * (T(2) * x * z); gy[13] = k[5] * (T(-2) * y * z); gz[13] = k[5] * (xx - yy);
        gx[14] = k[6] * (T(3) * xx - T(3) * yy); gy[14] = k[6] * (T(-6) * x * y); gz[14] = T(0);
    }
}

// Unit view direction from the camera centre to the mean, and the length before normalising.
template <typename T> void viewDir(const T *mu, const T campos[3], T d[3], T *len)
{
    const T vx = mu[0] - campos[0], vy = mu[1] - campos[1], vz = mu[2] - campos[2];
    *len = std::sqrt(vx * vx + vy * vy + vz * vz);
    d[0] = vx / *len;
    d[1] = vy / *len;
    d[2] = vz / *len;
}

template <typename T> void cameraCentre(const Cam<T> &cam, T c[3])
{
    for (int k = 0; k < 3; ++k)
        c[k] = -(cam.R[k] * cam.t[0] + cam.R[3 + k] * cam.t[1] + cam.R[6 + k] * cam.t[2]);
}

// Rasterizer.cpp:137-204 for one Gaussian; `rest` (15 x 3 per Gaussian, coefficient-major) and
// `degree` add the view-dependent SH terms (degree 0 is the reference's colour exactly).
template <typename T>
void projectScreen(size_t i, const T *means, const T *scales, const T *quats, const T *opac,
                   const T *colors, const Cam<T> &cam, const orc_settings &st, int tilesX,
                   int tilesY, Screen<T> &s, const T *rest = nullptr, int degree = 0)
{
    Proj<T> p;
    if (!projectOne(&means[i * 3], &quats[i * 4], &scales[i * 3], cam, T(st.near_plane),
                    T(st.eps2d), p))
        return;
    const T det = p.a * p.c - p.b * p.b;
    if (!(det > T(0)))
        return;
    const T invDet = T(1) / det;
    const T z = p.pc[2];
    const T u = cam.fx * p.pc[0] / z + cam.cx;
    const T v = cam.fy * p.pc[1] / z + cam.cy;
    const T mid = T(0.5f) * (p.a + p.c);
    const T lambda = mid + std::sqrt(std::max(T(0.1f), mid * mid - det));
    const int radius = int(std::ceil(T(3) * std::sqrt(lambda)));
    const T rr = T(radius);
    if (u + rr < T(0) || v + rr < T(0) || u - rr >= T(cam.width) || v - rr >= T(cam.height))
        return;
    const T ts = T(st.tile_size);
    const int tx0 = std::max(0, int(std::floor((u - rr) / ts)));
    const int ty0 = std::max(0, int(std::floor((v - rr) / ts)));
    const int tx1 = std::min(tilesX, int(std::floor((u + rr) / ts)) + 1);
    const int ty1 = std::min(tilesY, int(std::floor((v + rr) / ts)) + 1);
    if (tx0 >= tx1 || ty0 >= ty1)
        return;
    s.mean2d[i * 2] = u;
    s.mean2d[i * 2 + 1] = v;
    s.cov2d[i * 3] = p.a;
    s.cov2d[i * 3 + 1] = p.b;
    s.cov2d[i * 3 + 2] = p.c;
    s.conic[i * 3] = p.c * invDet;
    s.conic[i * 3 + 1] = -p.b * invDet;
    s.conic[i * 3 + 2] = p.a * invDet;
    s.depth[i] = z;
    uint8_t flags = uint8_t((p.clampX ? 1 : 0) | (p.clampY ? 2 : 0));
    const int nb = rest ? shCount(degree) : 0;
    T basis[15], gx[15], gy[15], gz[15];
    if (nb) {
        T campos[3], d[3], len;
        cameraCentre(cam, campos);
        viewDir(&means[i * 3], campos, d, &len);
        shBasis(d[0], d[1], d[2], degree, basis, gx, gy, gz);
    }
    for (int ch = 0; ch < 3; ++ch) {
        T col;
        if (nb == 0) {
            col = T(0.5f) + T(kShC0) * colors[i * 3 + ch];
        } else {
            T sum = T(kShC0) * colors[i * 3 + ch];
            for (int k = 0; k < nb; ++k)
                sum += basis[k] * rest[i * 45 + k * 3 + ch];
            col = T(0.5f) + sum;
        }
        if (col < T(0))
            flags |= uint8_t(4 << ch);
        s.rgb[i * 3 + ch] = std::max(col, T(0));
    }
    s.flags[i] = flags;
    s.alpha[i] = T(1) / (T(1) + ex(-opac[i]));
    s.rect[i * 4] = tx0;
    s.rect[i * 4 + 1] = ty0;
    s.rect[i * 4 + 2] = tx1;
    s.rect[i * 4 + 3] = ty1;
    s.radius[i] = radius;
}

// Rasterizer.cpp:206-261: visible list, stable LSD radix sort of depth bits (2 x 16 bits),
// per-tile count / scan / stable scatter. Depth sort for doubles uses the float bits of the
// value rounded to float (only the FP64 finite-difference path uses it; it keeps the same
// per-tile order as the FP32 path for the same model).
template <typename T>
void binTiles(size_t n, const std::vector<int> &radius, const std::vector<int> &rect,
              const std::vector<T> &depth, int tilesX, int tilesY, std::vector<int> &offsets,
              std::vector<int> &list)
{
    std::vector<int> vis;
    for (size_t i = 0; i < n; ++i)
        if (radius[i] > 0)
            vis.push_back(int(i));
    std::vector<uint32_t> key(vis.size()), key2(vis.size());
    std::vector<int> tmp(vis.size());
    for (size_t k = 0; k < vis.size(); ++k) {
        const float d = float(depth[size_t(vis[k])]);
        std::memcpy(&key[k], &d, 4);
    }
    std::vector<uint32_t> hist(65536);
    for (int pass = 0; pass < 2; ++pass) {
        const int shift = 16 * pass;
        std::fill(hist.begin(), hist.end(), 0u);
        for (uint32_t k : key)
            ++hist[(k >> shift) & 0xffffu];
        uint32_t run = 0;
        for (uint32_t &h : hist) {
            const uint32_t c = h;
            h = run;
            run += c;
        }
        for (size_t k = 0; k < vis.size(); ++k) {
            const uint32_t at = hist[(key[k] >> shift) & 0xffffu]++;
            tmp[at] = vis[k];
            key2[at] = key[k];
        }
        vis.swap(tmp);
        key.swap(key2);
    }
    const int tiles = tilesX * tilesY;
    offsets.assign(size_t(tiles) + 1, 0);
    for (int g : vis) {
        const int *r = &rect[size_t(g) * 4];
        for (int ty = r[1]; ty < r[3]; ++ty)
            for (int tx = r[0]; tx < r[2]; ++tx)
                ++offsets[size_t(ty * tilesX + tx) + 1];
    }
    for (int t = 0; t < tiles; ++t)
        offsets[size_t(t) + 1] += offsets[size_t(t)];
    list.resize(size_t(offsets.back()));
    std::vector<int> cursor(offsets.begin(), offsets.end() - 1);
    for (int g : vis) {
        const int *r = &rect[size_t(g) * 4];
        for (int ty = r[1]; ty < r[3]; ++ty)
            for (int tx = r[0]; tx < r[2]; ++tx)
                list[size_t(cursor[size_t(ty * tilesX + tx)]++)] = g;
    }
}

// Front-to-back compositing of one tile, Rasterizer.cpp:281-329.
template <typename T>
void rasterTile(int tile, int tilesX, int W, int H, const orc_settings &st, const Screen<T> &s,
                const std::vector<int> &offsets, const std::vector<int> &list, T *rgb, T *finalT,
                int *nContrib)
{
    const int ts = st.tile_size;
    const int tx = tile % tilesX, ty = tile / tilesX;
    const int begin = offsets[size_t(tile)], count = offsets[size_t(tile) + 1] - begin;
    const T minAlpha = T(st.min_alpha), minT = T(st.min_transmittance);
    const T bg[3] = {T(st.background[0]), T(st.background[1]), T(st.background[2])};
    const int x0 = tx * ts, y0 = ty * ts;
    const int x1 = std::min(W, x0 + ts), y1 = std::min(H, y0 + ts);
    for (int py = y0; py < y1; ++py)
        for (int px = x0; px < x1; ++px) {
            T Tr = T(1), cr = T(0), cg = T(0), cb = T(0);
            int last = 0;
            for (int j = 0; j < count; ++j) {
                const size_t g = size_t(list[size_t(begin + j)]);
                const T dx = s.mean2d[g * 2] - T(px), dy = s.mean2d[g * 2 + 1] - T(py);
                const T A = s.conic[g * 3], B = s.conic[g * 3 + 1], C = s.conic[g * 3 + 2];
                const T power = T(-0.5f) * (A * dx * dx + C * dy * dy) - B * dx * dy;
                if (power > T(0))
                    continue;
                const T alpha = std::min(T(kMaxAlpha), s.alpha[g] * ex(power));
                if (alpha < minAlpha)
                    continue;
                const T nextT = Tr * (T(1) - alpha);
                if (nextT < minT)
                    break;
                const T w = alpha * Tr;
                cr += s.rgb[g * 3] * w;
                cg += s.rgb[g * 3 + 1] * w;
                cb += s.rgb[g * 3 + 2] * w;
                Tr = nextT;
                last = j + 1;
            }
            const size_t pix = size_t(py) * size_t(W) + size_t(px);
            rgb[pix * 3] = cr + Tr * bg[0];
            rgb[pix * 3 + 1] = cg + Tr * bg[1];
            rgb[pix * 3 + 2] = cb + Tr * bg[2];
            if (finalT)
                finalT[pix] = Tr;
            if (nContrib)
                nContrib[pix] = last;
        }
}

} // namespace

// ------------------------------------------------------------------------------------ C API

extern "C" {

float orc_expf(float x) { return std::exp(x); }
void orc_expf_array(const float *x, float *y, long n)
{
    for (long k = 0; k < n; ++k)
        y[k] = std::exp(x[k]);
}

void orc_default_settings(orc_settings *s)
{
    s->tile_size = 16;
    s->near_plane = 0.01f;
    s->eps2d = 0.3f;
    s->background[0] = s->background[1] = s->background[2] = 0.0f;
    s->min_alpha = 1.0f / 255.0f;
    s->min_transmittance = 1e-4f;
}

// TrainOptions defaults (Trainer.h:19-37) with the compressed schedule of
// TrainOptions::forIterations (Trainer.cpp:26-35).
void orc_options_for_iterations(int iterations, orc_train_options *o)
{
    o->iterations = std::max(1, iterations);
    o->max_gaussians = 200000;
    o->ssim_lambda = 0.2f;
    o->lr = {1.6e-4f, 5e-3f, 1e-3f, 5e-2f, 2.5e-3f};
    o->means_lr_final = 1.6e-6f;
    o->densify_grad_threshold = 2e-4f;
    o->percent_dense = 0.01f;
    o->refine_start = std::min(500, o->iterations / 10);
    o->refine_stop = o->iterations * 6 / 10;
    o->refine_interval = 100;
    o->opacity_reset_interval = o->iterations >= 6000 ? 3000 : 0;
    o->prune_opacity = 0.005f;
    o->seed = 42;
    orc_default_settings(&o->raster);
}

void orc_project(const orc_model *m, const orc_camera *camera, const orc_settings *st,
                 float *mean2d, float *conic, float *cov2d, float *depth, float *rgb,
                 float *alpha, int *radius, int *rect, uint8_t *flags)
{
    const size_t n = size_t(m->n);
    const Cam<float> cam = camOf<float>(*camera);
    const int ts = st->tile_size;
    const int tilesX = (cam.width + ts - 1) / ts, tilesY = (cam.height + ts - 1) / ts;
    Screen<float> s;
    s.resize(n);
    for (size_t i = 0; i < n; ++i)
        projectScreen<float>(i, m->means, m->scales, m->quats, m->opacities, m->colors, cam, *st,
                             tilesX, tilesY, s);
    for (size_t i = 0; i < n; ++i) {
        radius[i] = s.radius[i];
        flags[i] = s.flags[i];
        if (s.radius[i] <= 0)
            continue;
        std::memcpy(&mean2d[i * 2], &s.mean2d[i * 2], 8);
        std::memcpy(&conic[i * 3], &s.conic[i * 3], 12);
        std::memcpy(&cov2d[i * 3], &s.cov2d[i * 3], 12);
        depth[i] = s.depth[i];
        std::memcpy(&rgb[i * 3], &s.rgb[i * 3], 12);
        alpha[i] = s.alpha[i];
        std::memcpy(&rect[i * 4], &s.rect[i * 4], 16);
    }
}

long orc_bin_tiles(int n, const int *radius, const int *rect, const float *depth, int tiles_x,
                   int tiles_y, int *tile_offsets, int *tile_list, long capacity)
{
    std::vector<int> rad(radius, radius + n), rc(rect, rect + size_t(n) * 4);
    std::vector<float> dep(depth, depth + n);
    std::vector<int> off, list;
    binTiles<float>(size_t(n), rad, rc, dep, tiles_x, tiles_y, off, list);
    if (long(list.size()) > capacity)
        return -1;
    std::copy(off.begin(), off.end(), tile_offsets);
    std::copy(list.begin(), list.end(), tile_list);
    return long(list.size());
}

struct orc_raster {
    orc_camera cam{};
    orc_settings st{};
    size_t n = 0;
    int tilesX = 0, tilesY = 0;
    Screen<float> s;
    std::vector<int> offsets, list;
    std::vector<float> finalT;
    std::vector<int> nContrib;
    std::vector<float> grad2d;
};

orc_raster *orc_raster_new(void) { return new orc_raster(); }
void orc_raster_free(orc_raster *r) { delete r; }

// Rasterizer::forward, Rasterizer.cpp:263-330.
void orc_raster_forward_sh(orc_raster *r, const orc_model *m, const float *rest, int degree,
                           const orc_camera *camera, const orc_settings *st, float *rgb)
{
    r->cam = *camera;
    r->st = *st;
    r->n = size_t(m->n);
    const Cam<float> cam = camOf<float>(*camera);
    const int ts = st->tile_size, W = cam.width, H = cam.height;
    r->tilesX = (W + ts - 1) / ts;
    r->tilesY = (H + ts - 1) / ts;
    r->s.resize(r->n);
    for (size_t i = 0; i < r->n; ++i)
        projectScreen<float>(i, m->means, m->scales, m->quats, m->opacities, m->colors, cam, *st,
                             r->tilesX, r->tilesY, r->s, rest, degree);
    binTiles<float>(r->n, r->s.radius, r->s.rect, r->s.depth, r->tilesX, r->tilesY, r->offsets,
                    r->list);
    r->finalT.assign(size_t(W) * size_t(H), 0.0f);
    r->nContrib.assign(size_t(W) * size_t(H), 0);
    for (int tile = 0; tile < r->tilesX * r->tilesY; ++tile)
        rasterTile<float>(tile, r->tilesX, W, H, *st, r->s, r->offsets, r->list, rgb,
                          r->finalT.data(), r->nContrib.data());
}

void orc_raster_forward(orc_raster *r, const orc_model *m, const orc_camera *camera,
                        const orc_settings *st, float *rgb)
{
    orc_raster_forward_sh(r, m, nullptr, 0, camera, st, rgb);
}

// Rasterizer::backward, Rasterizer.cpp:334-545, with the per-tile merge done serially in tile
// order (the reference's order when it runs on one thread).
} // extern "C"

namespace {
// Stage 2 of the backward (Rasterizer.cpp:432-544): per visible Gaussian of the last forward,
// the chain rule from its 2D gradient (grad2d [9n]) to the parameters, ADDED into the outputs,
// and the NDC statistic into ndc (nullable). T = float is the reference's arithmetic (same
// operation order); T = double evaluates the same expressions in FP64 from the same inputs.
template <typename T>
void chainRule(const orc_raster *r, const orc_model *m, const float *rest, int degree,
               const float *grad2d, T *gMeans, T *gScales, T *gQuats, T *gOpac, T *gColors,
               T *gRest, float *ndc)
{
    const Screen<float> &s = r->s;
    const Cam<T> cam = camOf<T>(r->cam);
    const size_t n = r->n;
    for (size_t i = 0; i < n; ++i) {
        if (s.radius[i] == 0)
            continue;
        const float *g = &grad2d[i * kG2];
        if (ndc) {
            const float nx = g[0] * 0.5f * float(cam.width), ny = g[1] * 0.5f * float(cam.height);
            ndc[i] += std::sqrt(nx * nx + ny * ny);
        }
        const T dmx = T(g[0]), dmy = T(g[1]), dA = T(g[2]), dB = T(g[3]), dC = T(g[4]);
        for (int ch = 0; ch < 3; ++ch)
            if (!(s.flags[i] & (4 << ch)))
                gColors[i * 3 + ch] += T(kShC0) * T(g[5 + ch]);
        T mu[3], qr[4], ls[3];
        for (int k = 0; k < 3; ++k) {
            mu[k] = T(m->means[i * 3 + k]);
            ls[k] = T(m->scales[i * 3 + k]);
        }
        for (int k = 0; k < 4; ++k)
            qr[k] = T(m->quats[i * 4 + k]);
        T dDir[3] = {0, 0, 0}, viewD[3] = {0, 0, 0}, viewLen = 1;
        const int nb = rest ? shCount(degree) : 0;
        if (nb) {
            // View-dependent colour: coefficient gradients and the direction -> mean term.
            T campos[3], b[15], bx[15], by[15], bz[15];
            cameraCentre(cam, campos);
            viewDir(mu, campos, viewD, &viewLen);
            shBasis(viewD[0], viewD[1], viewD[2], degree, b, bx, by, bz);
            for (int ch = 0; ch < 3; ++ch) {
                if (s.flags[i] & (4 << ch))
                    continue;
                const T gc = T(g[5 + ch]);
                for (int k = 0; k < nb; ++k) {
                    const T coef = T(rest[i * 45 + k * 3 + ch]);
                    if (gRest)
                        gRest[i * 45 + k * 3 + ch] += b[k] * gc;
                    dDir[0] += coef * bx[k] * gc;
                    dDir[1] += coef * by[k] * gc;
                    dDir[2] += coef * bz[k] * gc;
                }
            }
        }
        const T op = T(s.alpha[i]);
        gOpac[i] += T(g[8]) * op * (T(1) - op);

        Proj<T> p;
        projectOne<T>(mu, qr, ls, cam, T(r->st.near_plane), T(r->st.eps2d), p);
        const T z = p.pc[2];
        const T a = p.a, b = p.b, c = p.c;
        const T det = a * c - b * b;
        const T det2 = det * det;
        const T da = (dA * -c * c + dB * b * c + dC * -b * b) / det2;
        const T db = (dA * T(2) * b * c + dC * T(2) * a * b) / det2 + dB * (T(-1) / det - T(2) * b * b / det2);
        const T dc = (dA * -b * b + dB * a * b + dC * -a * a) / det2;
        const T G00 = da, G01 = T(0.5) * db, G11 = dc;
        T dSigma[9];
        for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l)
                dSigma[k * 3 + l] = p.T0[k] * (G00 * p.T0[l] + G01 * p.T1[l])
                                    + p.T1[k] * (G01 * p.T0[l] + G11 * p.T1[l]);
        T TS0[3], TS1[3];
        for (int col = 0; col < 3; ++col) {
            TS0[col] = p.T0[0] * p.Sigma[col] + p.T0[1] * p.Sigma[3 + col] + p.T0[2] * p.Sigma[6 + col];
            TS1[col] = p.T1[0] * p.Sigma[col] + p.T1[1] * p.Sigma[3 + col] + p.T1[2] * p.Sigma[6 + col];
        }
        T dT0[3], dT1[3];
        for (int col = 0; col < 3; ++col) {
            dT0[col] = T(2) * (G00 * TS0[col] + G01 * TS1[col]);
            dT1[col] = T(2) * (G01 * TS0[col] + G11 * TS1[col]);
        }
        T dM[9];
        for (int rr = 0; rr < 3; ++rr)
            for (int col = 0; col < 3; ++col)
                dM[rr * 3 + col] = T(2) * (dSigma[rr * 3] * p.M[col] + dSigma[rr * 3 + 1] * p.M[3 + col]
                                           + dSigma[rr * 3 + 2] * p.M[6 + col]);
        T dRq[9];
        for (int col = 0; col < 3; ++col) {
            T ds = 0;
            for (int rr = 0; rr < 3; ++rr) {
                ds += dM[rr * 3 + col] * p.Rq[rr * 3 + col];
                dRq[rr * 3 + col] = dM[rr * 3 + col] * p.s[col];
            }
            gScales[i * 3 + col] += ds * p.s[col];
        }
        const T w = p.q[0], x = p.q[1], y = p.q[2], zq = p.q[3];
        const T two = T(2);
        const T dw = two * (-zq * dRq[1] + y * dRq[2] + zq * dRq[3] - x * dRq[5] - y * dRq[6] + x * dRq[7]);
        const T dx = two * (y * dRq[1] + zq * dRq[2] + y * dRq[3] - two * x * dRq[4] - w * dRq[5]
                            + zq * dRq[6] + w * dRq[7] - two * x * dRq[8]);
        const T dy = two * (-two * y * dRq[0] + x * dRq[1] + w * dRq[2] + x * dRq[3] + zq * dRq[5]
                            - w * dRq[6] + zq * dRq[7] - two * y * dRq[8]);
        const T dz = two * (-two * zq * dRq[0] - w * dRq[1] + x * dRq[2] + w * dRq[3] - two * zq * dRq[4]
                            + y * dRq[5] + x * dRq[6] + y * dRq[7]);
        const T dq[4] = {dw, dx, dy, dz};
        const T qdot = dq[0] * p.q[0] + dq[1] * p.q[1] + dq[2] * p.q[2] + dq[3] * p.q[3];
        for (int k = 0; k < 4; ++k)
            gQuats[i * 4 + k] += (dq[k] - qdot * p.q[k]) / p.qNorm;
        const T *R = cam.R;
        const T dJ00 = dT0[0] * R[0] + dT0[1] * R[1] + dT0[2] * R[2];
        const T dJ02 = dT0[0] * R[6] + dT0[1] * R[7] + dT0[2] * R[8];
        const T dJ11 = dT1[0] * R[3] + dT1[1] * R[4] + dT1[2] * R[5];
        const T dJ12 = dT1[0] * R[6] + dT1[1] * R[7] + dT1[2] * R[8];
        const T z2 = z * z, z3 = z2 * z;
        T dpc[3] = {0, 0, 0};
        dpc[2] += dJ00 * -cam.fx / z2 + dJ02 * two * cam.fx * p.tx / z3 + dJ11 * -cam.fy / z2
                  + dJ12 * two * cam.fy * p.ty / z3;
        const T dtx = dJ02 * -cam.fx / z2;
        const T dty = dJ12 * -cam.fy / z2;
        if (p.clampX)
            dpc[2] += dtx * p.limTx;
        else
            dpc[0] += dtx;
        if (p.clampY)
            dpc[2] += dty * p.limTy;
        else
            dpc[1] += dty;
        dpc[0] += dmx * cam.fx / z;
        dpc[1] += dmy * cam.fy / z;
        dpc[2] += -dmx * cam.fx * p.pc[0] / z2 - dmy * cam.fy * p.pc[1] / z2;
        for (int k = 0; k < 3; ++k)
            gMeans[i * 3 + k] += R[k] * dpc[0] + R[3 + k] * dpc[1] + R[6 + k] * dpc[2];
        if (nb) {
            const T dot = dDir[0] * viewD[0] + dDir[1] * viewD[1] + dDir[2] * viewD[2];
            for (int k = 0; k < 3; ++k)
                gMeans[i * 3 + k] += (dDir[k] - dot * viewD[k]) / viewLen;
        }
    }
}
} // namespace

extern "C" {

void orc_raster_backward_sh(orc_raster *r, const orc_model *m, const float *rest, int degree,
                            const float *dl, orc_model *grads, float *grad_rest, float *ndc)
{
    const size_t n = r->n;
    const Screen<float> &s = r->s;
    r->grad2d.assign(n * kG2, 0.0f);
    const int W = r->cam.width, H = r->cam.height, ts = r->st.tile_size;
    const float *bg = r->st.background;
    const float minAlpha = r->st.min_alpha;
    std::vector<float> acc;
    for (int tile = 0; tile < r->tilesX * r->tilesY; ++tile) {
        const int tx = tile % r->tilesX, ty = tile / r->tilesX;
        const int begin = r->offsets[size_t(tile)];
        const int count = r->offsets[size_t(tile) + 1] - begin;
        if (count == 0)
            continue;
        acc.assign(size_t(count) * kG2, 0.0f);
        const int x0 = tx * ts, y0 = ty * ts;
        const int x1 = std::min(W, x0 + ts), y1 = std::min(H, y0 + ts);
        for (int py = y0; py < y1; ++py)
            for (int px = x0; px < x1; ++px) {
                const size_t pix = size_t(py) * size_t(W) + size_t(px);
                const float *dp = &dl[pix * 3];
                const float finalT = r->finalT[pix];
                const float bgDot = bg[0] * dp[0] + bg[1] * dp[1] + bg[2] * dp[2];
                float T = finalT;
                float accum[3] = {0, 0, 0}, lastColor[3] = {0, 0, 0};
                float lastAlpha = 0;
                for (int j = r->nContrib[pix] - 1; j >= 0; --j) {
                    const size_t g = size_t(r->list[size_t(begin + j)]);
                    const float A = s.conic[g * 3], B = s.conic[g * 3 + 1], C = s.conic[g * 3 + 2];
                    const float op = s.alpha[g];
                    const float dx = s.mean2d[g * 2] - float(px), dy = s.mean2d[g * 2 + 1] - float(py);
                    const float power = -0.5f * (A * dx * dx + C * dy * dy) - B * dx * dy;
                    if (power > 0.0f)
                        continue;
                    const float G = std::exp(power);
                    const float rawAlpha = op * G;
                    const float alpha = std::min(kMaxAlpha, rawAlpha);
                    if (alpha < minAlpha)
                        continue;
                    T = T / (1.0f - alpha);
                    const float w = alpha * T;
                    float *a = &acc[size_t(j) * kG2];
                    float dAlpha = 0.0f;
                    for (int ch = 0; ch < 3; ++ch) {
                        accum[ch] = lastAlpha * lastColor[ch] + (1.0f - lastAlpha) * accum[ch];
                        lastColor[ch] = s.rgb[g * 3 + ch];
                        dAlpha += (s.rgb[g * 3 + ch] - accum[ch]) * dp[ch];
                        a[5 + ch] += w * dp[ch];
                    }
                    dAlpha *= T;
                    lastAlpha = alpha;
                    dAlpha += -finalT / (1.0f - alpha) * bgDot;
                    if (rawAlpha > kMaxAlpha)
                        continue;
                    a[8] += G * dAlpha;
                    const float dPower = op * dAlpha * G;
                    a[0] += dPower * -(A * dx + B * dy);
                    a[1] += dPower * -(C * dy + B * dx);
                    a[2] += dPower * -0.5f * dx * dx;
                    a[3] += dPower * -dx * dy;
                    a[4] += dPower * -0.5f * dy * dy;
                }
            }
        for (int j = 0; j < count; ++j) {
            const float *a = &acc[size_t(j) * kG2];
            float *dst = &r->grad2d[size_t(r->list[size_t(begin + j)]) * kG2];
            for (int k = 0; k < kG2; ++k)
                if (a[k] != 0.0f)
                    dst[k] = dst[k] + a[k];
        }
    }

    chainRule<float>(r, m, rest, degree, r->grad2d.data(), grads->means, grads->scales,
                     grads->quats, grads->opacities, grads->colors, grad_rest, ndc);
}

// Rasterizer.cpp:137-204 in FP64 for n Gaussians (double arrays); outputs only where
// radius[i] > 0 (radius 0 elsewhere). rest may be NULL (degree 0).
void orc_project_f64(int n, const double *means, const double *scales, const double *quats,
                     const double *opacities, const double *colors, const double *rest, int degree,
                     const orc_camera *camera, const orc_settings *st, double *mean2d,
                     double *conic, double *rgb, double *alpha, int *radius)
{
    const Cam<double> cam = camOf<double>(*camera);
    const int ts = st->tile_size;
    const int tilesX = (camera->width + ts - 1) / ts, tilesY = (camera->height + ts - 1) / ts;
    Screen<double> s;
    s.resize(size_t(n));
    for (size_t i = 0; i < size_t(n); ++i)
        projectScreen<double>(i, means, scales, quats, opacities, colors, cam, *st, tilesX, tilesY, s,
                              rest, degree);
    for (size_t i = 0; i < size_t(n); ++i) {
        radius[i] = s.radius[i];
        for (int k = 0; k < 2; ++k)
            mean2d[i * 2 + k] = s.mean2d[i * 2 + k];
        for (int k = 0; k < 3; ++k) {
            conic[i * 3 + k] = s.conic[i * 3 + k];
            rgb[i * 3 + k] = s.rgb[i * 3 + k];
        }
        alpha[i] = s.alpha[i];
    }
}

// The chain-rule stage in FP64 from given 2D gradients (e.g. orc_raster_get_grad2d after a
// backward) and the last forward's state; outputs (zeroed by the caller) are added to.
void orc_chain_rule_f64(const orc_raster *r, const orc_model *m, const float *rest, int degree,
                        const float *grad2d, double *g_means, double *g_scales, double *g_quats,
                        double *g_opacities, double *g_colors, double *g_rest)
{
    chainRule<double>(r, m, rest, degree, grad2d, g_means, g_scales, g_quats, g_opacities, g_colors,
                      g_rest, nullptr);
}

void orc_raster_backward(orc_raster *r, const orc_model *m, const float *dl, orc_model *grads,
                         float *ndc)
{
    orc_raster_backward_sh(r, m, nullptr, 0, dl, grads, nullptr, ndc);
}

long orc_raster_num_pairs(const orc_raster *r) { return long(r->list.size()); }
int orc_raster_num_tiles(const orc_raster *r) { return r->tilesX * r->tilesY; }

void orc_raster_get_state(const orc_raster *r, int *radius, int *tile_offsets, int *tile_list,
                          float *final_t, int *n_contrib)
{
    if (radius)
        std::copy(r->s.radius.begin(), r->s.radius.end(), radius);
    if (tile_offsets)
        std::copy(r->offsets.begin(), r->offsets.end(), tile_offsets);
    if (tile_list)
        std::copy(r->list.begin(), r->list.end(), tile_list);
    if (final_t)
        std::copy(r->finalT.begin(), r->finalT.end(), final_t);
    if (n_contrib)
        std::copy(r->nContrib.begin(), r->nContrib.end(), n_contrib);
}

void orc_raster_get_projection(const orc_raster *r, float *mean2d, float *conic, float *depth,
                               float *rgb, float *alpha, int *rect, uint8_t *flags)
{
    const Screen<float> &s = r->s;
    if (mean2d) std::copy(s.mean2d.begin(), s.mean2d.end(), mean2d);
    if (conic) std::copy(s.conic.begin(), s.conic.end(), conic);
    if (depth) std::copy(s.depth.begin(), s.depth.end(), depth);
    if (rgb) std::copy(s.rgb.begin(), s.rgb.end(), rgb);
    if (alpha) std::copy(s.alpha.begin(), s.alpha.end(), alpha);
    if (rect) std::copy(s.rect.begin(), s.rect.end(), rect);
    if (flags) std::copy(s.flags.begin(), s.flags.end(), flags);
}

void orc_raster_get_grad2d(const orc_raster *r, float *grad2d)
{
    std::copy(r->grad2d.begin(), r->grad2d.end(), grad2d);
}

void orc_render_f64_sh(int n, const double *means, const double *scales, const double *quats,
                       const double *opacities, const double *colors, const double *rest,
                       int degree, const orc_camera *camera, const orc_settings *st, double *rgb)
{
    const Cam<double> cam = camOf<double>(*camera);
    const int ts = st->tile_size, W = cam.width, H = cam.height;
    const int tilesX = (W + ts - 1) / ts, tilesY = (H + ts - 1) / ts;
    Screen<double> s;
    s.resize(size_t(n));
    for (size_t i = 0; i < size_t(n); ++i)
        projectScreen<double>(i, means, scales, quats, opacities, colors, cam, *st, tilesX, tilesY, s,
                              rest, degree);
    std::vector<int> off, list;
    binTiles<double>(size_t(n), s.radius, s.rect, s.depth, tilesX, tilesY, off, list);
    for (int tile = 0; tile < tilesX * tilesY; ++tile)
        rasterTile<double>(tile, tilesX, W, H, *st, s, off, list, rgb, nullptr, nullptr);
}

void orc_render_f64(int n, const double *means, const double *scales, const double *quats,
                    const double *opacities, const double *colors, const orc_camera *camera,
                    const orc_settings *st, double *rgb)
{
    orc_render_f64_sh(n, means, scales, quats, opacities, colors, nullptr, 0, camera, st, rgb);
}

// SH basis values (degree 1..3) of a unit direction, for the orthonormality check.
void orc_sh_basis(float x, float y, float z, int degree, double *out15)
{
    double b[15] = {}, gx[15], gy[15], gz[15];
    shBasis<double>(x, y, z, degree, b, gx, gy, gz);
    for (int k = 0; k < 15; ++k)
        out15[k] = b[k];
}

// One Adam step over a single array (used for the SH coefficient group).
void orc_adam_array(float *p, const float *g, float *m, float *v, long len, float lr, int step)
{
    const float bc1 = 1.0f - std::pow(0.9f, float(step));
    const float bc2 = 1.0f - std::pow(0.999f, float(step));
    const float stepSize = lr / bc1;
    const float invSqrtBc2 = 1.0f / std::sqrt(bc2);
    for (long k = 0; k < len; ++k) {
        const float gk = g[k];
        m[k] = 0.9f * m[k] + (1.0f - 0.9f) * gk;
        v[k] = 0.999f * v[k] + (1.0f - 0.999f) * gk * gk;
        p[k] -= stepSize * m[k] / (std::sqrt(v[k]) * invSqrtBc2 + 1e-15f);
    }
}

// photometricLoss, Loss.cpp:47-185: four separable passes over planar scratch.
orc_loss_value orc_photometric_loss(const float *render, const float *target, int W, int H,
                                    float lambda, float *grad)
{
    constexpr int kR = 5;
    constexpr float kC1 = 0.01f * 0.01f, kC2 = 0.03f * 0.03f;
    float wk[2 * kR + 1];
    {
        float sum = 0;
        for (int k = -kR; k <= kR; ++k) {
            wk[k + kR] = std::exp(-float(k * k) / (2.0f * 1.5f * 1.5f));
            sum += wk[k + kR];
        }
        for (float &v : wk)
            v /= sum;
    }
    const size_t P = size_t(W) * size_t(H);
    const float invCount = 1.0f / float(P * 3);
    std::vector<float> h[5], abc[3], h2[3];
    for (auto &v : h) v.assign(3 * P, 0.0f);
    for (auto &v : abc) v.assign(3 * P, 0.0f);
    for (auto &v : h2) v.assign(3 * P, 0.0f);
    double l1 = 0;
    for (int r = 0; r < 3 * H; ++r) {
        const int c = r / H, y = r % H;
        const size_t base = size_t(c) * P + size_t(y) * size_t(W);
        for (int x = 0; x < W; ++x) {
            const size_t idx = (size_t(y) * size_t(W) + size_t(x)) * 3 + size_t(c);
            const float d = render[idx] - target[idx];
            l1 += std::fabs(d);
            grad[idx] = (1.0f - lambda) * invCount * (d > 0 ? 1.0f : (d < 0 ? -1.0f : 0.0f));
        }
        if (lambda <= 0.0f)
            continue;
        for (int x = 0; x < W; ++x) {
            float sx = 0, sy = 0, sxx = 0, syy = 0, sxy = 0;
            const int k0 = std::max(-kR, -x), k1 = std::min(kR, W - 1 - x);
            for (int k = k0; k <= k1; ++k) {
                const size_t idx = (size_t(y) * size_t(W) + size_t(x + k)) * 3 + size_t(c);
                const float a = render[idx], b = target[idx], w = wk[k + kR];
                sx += w * a;
                sy += w * b;
                sxx += w * a * a;
                syy += w * b * b;
                sxy += w * a * b;
            }
            h[0][base + size_t(x)] = sx;
            h[1][base + size_t(x)] = sy;
            h[2][base + size_t(x)] = sxx;
            h[3][base + size_t(x)] = syy;
            h[4][base + size_t(x)] = sxy;
        }
    }
    orc_loss_value res{0, 0, 0};
    res.l1 = float(l1 * invCount);
    if (lambda <= 0.0f) {
        res.total = res.l1;
        return res;
    }
    const float dLds = -lambda * invCount;
    double ssim = 0;
    for (int r = 0; r < 3 * H; ++r) {
        const int c = r / H, y = r % H;
        const size_t cbase = size_t(c) * P;
        const int k0 = std::max(-kR, -y), k1 = std::min(kR, H - 1 - y);
        for (int x = 0; x < W; ++x) {
            float m[5] = {0, 0, 0, 0, 0};
            for (int k = k0; k <= k1; ++k) {
                const size_t i = cbase + size_t(y + k) * size_t(W) + size_t(x);
                const float w = wk[k + kR];
                for (int j = 0; j < 5; ++j)
                    m[j] += w * h[j][i];
            }
            const float mx = m[0], my = m[1];
            const float sx = m[2] - mx * mx, sy = m[3] - my * my, sxy = m[4] - mx * my;
            const float a1 = 2 * mx * my + kC1, a2 = 2 * sxy + kC2;
            const float b1 = mx * mx + my * my + kC1, b2 = sx + sy + kC2;
            const float v = (a1 * a2) / (b1 * b2);
            ssim += v;
            const float dmu = (2 * my * a2 - 2 * my * a1) / (b1 * b2) - v * (2 * mx / b1 - 2 * mx / b2);
            const size_t o = cbase + size_t(y) * size_t(W) + size_t(x);
            abc[0][o] = dLds * dmu;
            abc[1][o] = dLds * (-v / b2);
            abc[2][o] = dLds * (2 * a1 / (b1 * b2));
        }
    }
    for (int r = 0; r < 3 * H; ++r) {
        const size_t base = size_t(r / H) * P + size_t(r % H) * size_t(W);
        for (int x = 0; x < W; ++x) {
            float acc[3] = {0, 0, 0};
            const int k0 = std::max(-kR, -x), k1 = std::min(kR, W - 1 - x);
            for (int k = k0; k <= k1; ++k) {
                const float w = wk[k + kR];
                for (int j = 0; j < 3; ++j)
                    acc[j] += w * abc[j][base + size_t(x + k)];
            }
            for (int j = 0; j < 3; ++j)
                h2[j][base + size_t(x)] = acc[j];
        }
    }
    for (int r = 0; r < 3 * H; ++r) {
        const int c = r / H, y = r % H;
        const size_t cbase = size_t(c) * P;
        const int k0 = std::max(-kR, -y), k1 = std::min(kR, H - 1 - y);
        for (int x = 0; x < W; ++x) {
            float g[3] = {0, 0, 0};
            for (int k = k0; k <= k1; ++k) {
                const size_t i = cbase + size_t(y + k) * size_t(W) + size_t(x);
                const float w = wk[k + kR];
                for (int j = 0; j < 3; ++j)
                    g[j] += w * h2[j][i];
            }
            const size_t idx = (size_t(y) * size_t(W) + size_t(x)) * 3 + size_t(c);
            grad[idx] += g[0] + 2.0f * render[idx] * g[1] + target[idx] * g[2];
        }
    }
    res.ssim = float(ssim * invCount);
    res.total = (1.0f - lambda) * res.l1 + lambda * (1.0f - res.ssim);
    return res;
}

float orc_psnr(const float *a, const float *b, long count)
{
    double mse = 0;
    for (long k = 0; k < count; ++k) {
        const double d = double(a[k]) - double(b[k]);
        mse += d * d;
    }
    mse /= double(std::max<long>(count, 1));
    return mse <= 1e-12 ? 99.0f : float(-10.0 * std::log10(mse));
}

// Adam::step + adamArray, Adam.cpp:13-28, 68-78. `step` is the 1-based step number.
void orc_adam_step(orc_model *p, const orc_model *g, orc_model *m, orc_model *v, const orc_lrs *lr,
                   int step)
{
    const float bc1 = 1.0f - std::pow(0.9f, float(step));
    const float bc2 = 1.0f - std::pow(0.999f, float(step));
    auto run = [&](float *pp, const float *gg, float *mm, float *vv, size_t len, float rate) {
        const float stepSize = rate / bc1;
        const float invSqrtBc2 = 1.0f / std::sqrt(bc2);
        for (size_t k = 0; k < len; ++k) {
            const float gk = gg[k];
            mm[k] = 0.9f * mm[k] + (1.0f - 0.9f) * gk;
            vv[k] = 0.999f * vv[k] + (1.0f - 0.999f) * gk * gk;
            pp[k] -= stepSize * mm[k] / (std::sqrt(vv[k]) * invSqrtBc2 + 1e-15f);
        }
    };
    const size_t n = size_t(p->n);
    run(p->means, g->means, m->means, v->means, n * 3, lr->means);
    run(p->scales, g->scales, m->scales, v->scales, n * 3, lr->scales);
    run(p->quats, g->quats, m->quats, v->quats, n * 4, lr->quats);
    run(p->opacities, g->opacities, m->opacities, v->opacities, n, lr->opacities);
    run(p->colors, g->colors, m->colors, v->colors, n * 3, lr->colors);
}

} // extern "C"

namespace {

struct HostModel {
    std::vector<float> means, scales, quats, opacities, colors;
    size_t size() const { return opacities.size(); }
    void resize(size_t n)
    {
        means.resize(n * 3);
        scales.resize(n * 3);
        quats.resize(n * 4);
        opacities.resize(n);
        colors.resize(n * 3);
    }
    void zero()
    {
        for (auto *v : {&means, &scales, &quats, &opacities, &colors})
            std::fill(v->begin(), v->end(), 0.0f);
    }
    orc_model view()
    {
        return {means.data(), scales.data(), quats.data(), opacities.data(), colors.data(), int(size())};
    }
    void append(const float *mu, const float *sc, const float *q, float o, const float *col)
    {
        means.insert(means.end(), mu, mu + 3);
        scales.insert(scales.end(), sc, sc + 3);
        quats.insert(quats.end(), q, q + 4);
        opacities.push_back(o);
        colors.insert(colors.end(), col, col + 3);
    }
    void appendFrom(const HostModel &o, size_t i)
    {
        append(&o.means[i * 3], &o.scales[i * 3], &o.quats[i * 4], o.opacities[i], &o.colors[i * 3]);
    }
    void compact(const std::vector<uint8_t> &keep)
    {
        size_t out = 0;
        for (size_t i = 0; i < keep.size(); ++i) {
            if (!keep[i])
                continue;
            if (out != i) {
                std::copy_n(&means[i * 3], 3, &means[out * 3]);
                std::copy_n(&scales[i * 3], 3, &scales[out * 3]);
                std::copy_n(&quats[i * 4], 4, &quats[out * 4]);
                opacities[out] = opacities[i];
                std::copy_n(&colors[i * 3], 3, &colors[out * 3]);
            }
            ++out;
        }
        resize(out);
    }
};

// The greedy densification loop of Trainer.cpp:105-169; normal draws come from `draw`.
template <typename Draw>
void densifyLoop(const orc_model &m, const float *accum, const int *count, float extent,
                 int maxG, float thr, float percentDense, float pruneOpacity, bool wasReset,
                 Draw draw, std::vector<uint8_t> &keep, std::vector<uint8_t> &kind, HostModel &extra)
{
    const size_t n = size_t(m.n);
    const float splitScale = percentDense * extent;
    const size_t budget = maxG > int(n) ? size_t(maxG) - n : 0;
    keep.assign(n, 1);
    kind.assign(n, 0);
    size_t added = 0;
    const float shrink = std::log(1.6f);
    for (size_t i = 0; i < n; ++i) {
        const float maxLog = std::max({m.scales[i * 3], m.scales[i * 3 + 1], m.scales[i * 3 + 2]});
        const float maxScale = std::exp(maxLog);
        const float alpha = sigmoidf_(m.opacities[i]);
        if (alpha < pruneOpacity || (wasReset && maxScale > 0.1f * extent)) {
            keep[i] = 0;
            continue;
        }
        if (count[i] == 0)
            continue;
        const float avg = accum[i] / float(count[i]);
        if (avg < thr)
            continue;
        if (maxScale <= splitScale) {
            if (added + 1 > budget)
                continue;
            extra.append(&m.means[i * 3], &m.scales[i * 3], &m.quats[i * 4], m.opacities[i], &m.colors[i * 3]);
            kind[i] = 1;
            added += 1;
        } else {
            if (added + 2 > budget)
                continue;
            const float *q = &m.quats[i * 4];
            const float qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) + 1e-12f;
            const float w = q[0] / qn, x = q[1] / qn, y = q[2] / qn, z = q[3] / qn;
            float R[9];
            rotOfQuat(w, x, y, z, R);
            for (int c = 0; c < 2; ++c) {
                const size_t j = extra.size();
                extra.append(&m.means[i * 3], &m.scales[i * 3], &m.quats[i * 4], m.opacities[i], &m.colors[i * 3]);
                float local[3];
                for (int k = 0; k < 3; ++k)
                    local[k] = draw() * std::exp(m.scales[i * 3 + k]);
                for (int r = 0; r < 3; ++r)
                    extra.means[j * 3 + r] += R[r * 3] * local[0] + R[r * 3 + 1] * local[1] + R[r * 3 + 2] * local[2];
                for (int k = 0; k < 3; ++k)
                    extra.scales[j * 3 + k] -= shrink;
            }
            keep[i] = 0;
            kind[i] = 2;
            added += 2;
        }
    }
}

} // namespace

extern "C" {

long orc_densify(const orc_model *m, const float *grad_accum, const int *grad_count, float extent,
                 int max_gaussians, float grad_threshold, float percent_dense, float prune_opacity,
                 int opacity_was_reset, const float *normals, long normals_len, uint8_t *keep,
                 uint8_t *kind, orc_model *extra, long extra_cap, long *normals_used)
{
    long used = 0;
    auto draw = [&]() -> float { return used < normals_len ? normals[used++] : (used++, 0.0f); };
    std::vector<uint8_t> kp, kd;
    HostModel ex;
    densifyLoop(*m, grad_accum, grad_count, extent, max_gaussians, grad_threshold, percent_dense,
                prune_opacity, opacity_was_reset != 0, draw, kp, kd, ex);
    std::copy(kp.begin(), kp.end(), keep);
    std::copy(kd.begin(), kd.end(), kind);
    if (normals_used)
        *normals_used = used;
    const long cnt = long(ex.size());
    if (extra && cnt <= extra_cap) {
        std::copy(ex.means.begin(), ex.means.end(), extra->means);
        std::copy(ex.scales.begin(), ex.scales.end(), extra->scales);
        std::copy(ex.quats.begin(), ex.quats.end(), extra->quats);
        std::copy(ex.opacities.begin(), ex.opacities.end(), extra->opacities);
        std::copy(ex.colors.begin(), ex.colors.end(), extra->colors);
    }
    return cnt;
}

// initGaussiansFromPoints, GaussianModel.cpp:68-144 (hash grid of 2x the mean spacing, 27
// neighbouring cells, 3 smallest squared distances).
void orc_init_from_points(int n, const float *xyz, const float *rgb, int rgb_count, orc_model *out)
{
    if (n <= 0)
        return;
    float lo[3] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                   std::numeric_limits<float>::max()};
    float hi[3] = {-lo[0], -lo[1], -lo[2]};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], xyz[i * 3 + k]);
            hi[k] = std::max(hi[k], xyz[i * 3 + k]);
        }
    const float extent = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-6f});
    const float cell = std::max(extent / std::cbrt(float(n)) * 2.0f, 1e-6f);
    auto key = [](int x, int y, int z) {
        return (int64_t(x) * 73856093) ^ (int64_t(y) * 19349663) ^ (int64_t(z) * 83492791);
    };
    auto cellOf = [&](int i, int c[3]) {
        for (int k = 0; k < 3; ++k)
            c[k] = int(std::floor((xyz[i * 3 + k] - lo[k]) / cell));
    };
    std::unordered_map<int64_t, std::vector<uint32_t>> grid;
    grid.reserve(size_t(n));
    for (int i = 0; i < n; ++i) {
        int c[3];
        cellOf(i, c);
        grid[key(c[0], c[1], c[2])].push_back(uint32_t(i));
    }
    const float dc = 0.5f;
    for (int i = 0; i < n; ++i) {
        float best[3] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                         std::numeric_limits<float>::max()};
        int c[3];
        cellOf(i, c);
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    auto it = grid.find(key(c[0] + dx, c[1] + dy, c[2] + dz));
                    if (it == grid.end())
                        continue;
                    for (uint32_t j : it->second) {
                        if (int(j) == i)
                            continue;
                        const float ex_ = xyz[j * 3] - xyz[i * 3], ey = xyz[j * 3 + 1] - xyz[i * 3 + 1],
                                    ez = xyz[j * 3 + 2] - xyz[i * 3 + 2];
                        float d2 = ex_ * ex_ + ey * ey + ez * ez;
                        for (int k = 0; k < 3; ++k)
                            if (d2 < best[k])
                                std::swap(d2, best[k]);
                    }
                }
        float sum = 0;
        int cnt = 0;
        for (float b : best)
            if (b < std::numeric_limits<float>::max()) {
                sum += b;
                ++cnt;
            }
        const float meanD2 = cnt ? sum / float(cnt) : cell * cell;
        const float logScale = std::log(std::sqrt(std::max(meanD2, 1e-14f)));
        for (int k = 0; k < 3; ++k) {
            out->means[i * 3 + k] = xyz[i * 3 + k];
            out->scales[i * 3 + k] = logScale;
            const float col = (rgb && rgb_count >= (i + 1) * 3) ? rgb[i * 3 + k] : 0.5f;
            out->colors[i * 3 + k] = (col - dc) / kShC0;
        }
        out->quats[i * 4] = 1.0f;
        out->quats[i * 4 + 1] = out->quats[i * 4 + 2] = out->quats[i * 4 + 3] = 0.0f;
        out->opacities[i] = logitf_(0.1f);
    }
}

// Trainer, Trainer.cpp:37-197 (one thread).
struct orc_trainer {
    HostModel model, grads, adamM, adamV;
    int adamStep = 0;
    std::vector<orc_camera> cams;
    std::vector<std::vector<uint8_t>> images;
    float extent = 1.0f;
    orc_train_options opt{};
    orc_raster raster;
    std::mt19937 rng;
    std::vector<int> order;
    size_t orderPos = 0;
    int iteration = 0;
    bool opacityWasReset = false;
    std::vector<float> target, render, dLoss, gradAccum;
    std::vector<int> gradCount;
    int statIteration = 0;
    float statLoss = 0, statPsnr = 0;
    long statGaussians = 0;
    // SH extension (no reference counterpart): degree, rate scale, coefficients [n][15][3] and
    // their gradient and Adam moments.
    int shDegree = 0;
    float shLrScale = 0.05f;
    std::vector<float> shRest, shGrad, shM, shV;

    void loadTarget(size_t k)
    {
        const auto &img = images[k];
        target.resize(img.size());
        constexpr float inv = 1.0f / 255.0f;
        for (size_t i = 0; i < img.size(); ++i)
            target[i] = float(img[i]) * inv;
    }
};

orc_trainer *orc_trainer_new(const orc_model *initial, const orc_camera *cams, const uint8_t *rgb8,
                             int n_images, float extent, const orc_train_options *opt)
{
    auto *t = new orc_trainer();
    const size_t n = size_t(initial->n);
    t->model.means.assign(initial->means, initial->means + n * 3);
    t->model.scales.assign(initial->scales, initial->scales + n * 3);
    t->model.quats.assign(initial->quats, initial->quats + n * 4);
    t->model.opacities.assign(initial->opacities, initial->opacities + n);
    t->model.colors.assign(initial->colors, initial->colors + n * 3);
    size_t off = 0;
    for (int k = 0; k < n_images; ++k) {
        t->cams.push_back(cams[k]);
        const size_t len = size_t(cams[k].width) * size_t(cams[k].height) * 3;
        t->images.emplace_back(rgb8 + off, rgb8 + off + len);
        off += len;
    }
    t->extent = std::max(extent, 1e-6f);
    t->opt = *opt;
    t->rng.seed(opt->seed);
    t->adamM.resize(n);
    t->adamV.resize(n);
    t->adamM.zero();
    t->adamV.zero();
    t->gradAccum.assign(n, 0.0f);
    t->gradCount.assign(n, 0);
    t->statGaussians = long(n);
    return t;
}

void orc_trainer_set_sh(orc_trainer *t, int degree, float lr_scale, const float *rest)
{
    t->shDegree = degree;
    t->shLrScale = lr_scale > 0.0f ? lr_scale : 1.0f / 20.0f;
    const size_t len = t->model.size() * 45;
    t->shRest.assign(len, 0.0f);
    if (rest)
        t->shRest.assign(rest, rest + len);
    t->shM.assign(len, 0.0f);
    t->shV.assign(len, 0.0f);
}

void orc_trainer_get_sh(const orc_trainer *t, float *rest)
{
    std::copy(t->shRest.begin(), t->shRest.end(), rest);
}

void orc_trainer_free(orc_trainer *t) { delete t; }
int orc_trainer_done(const orc_trainer *t) { return t->iteration >= t->opt.iterations; }
int orc_trainer_iteration(const orc_trainer *t) { return t->iteration; }
int orc_trainer_size(const orc_trainer *t) { return int(t->model.size()); }

// Rows of a [n][width] array after densification: kept rows in order, then one row per clone
// and two per split in index order, copied from the parent (or zero when zeroExtras).
static void densifyRows(std::vector<float> &a, size_t width, const std::vector<uint8_t> &keep,
                        const std::vector<uint8_t> &kind, bool zeroExtras)
{
    std::vector<float> out, extras;
    for (size_t i = 0; i < keep.size(); ++i) {
        const float *row = &a[i * width];
        if (keep[i])
            out.insert(out.end(), row, row + width);
        for (int c = 0; c < (kind[i] == 1 ? 1 : kind[i] == 2 ? 2 : 0); ++c) {
            if (zeroExtras)
                extras.insert(extras.end(), width, 0.0f);
            else
                extras.insert(extras.end(), row, row + width);
        }
    }
    out.insert(out.end(), extras.begin(), extras.end());
    a.swap(out);
}

static void trainerDensify(orc_trainer *t)
{
    std::normal_distribution<float> normal(0.0f, 1.0f);
    std::vector<uint8_t> keep, kind;
    HostModel extra;
    const orc_model mv = t->model.view();
    densifyLoop(mv, t->gradAccum.data(), t->gradCount.data(), t->extent, t->opt.max_gaussians,
                t->opt.densify_grad_threshold, t->opt.percent_dense, t->opt.prune_opacity,
                t->opacityWasReset, [&]() { return normal(t->rng); }, keep, kind, extra);
    if (t->shDegree > 0) {
        densifyRows(t->shRest, 45, keep, kind, false);
        densifyRows(t->shM, 45, keep, kind, true);
        densifyRows(t->shV, 45, keep, kind, true);
    }
    t->model.compact(keep);
    t->adamM.compact(keep);
    t->adamV.compact(keep);
    for (size_t j = 0; j < extra.size(); ++j)
        t->model.appendFrom(extra, j);
    const size_t n = t->model.size();
    const size_t old = t->adamM.size();
    t->adamM.resize(n);
    t->adamV.resize(n);
    for (HostModel *h : {&t->adamM, &t->adamV}) {
        std::fill(h->means.begin() + long(old * 3), h->means.end(), 0.0f);
        std::fill(h->scales.begin() + long(old * 3), h->scales.end(), 0.0f);
        std::fill(h->quats.begin() + long(old * 4), h->quats.end(), 0.0f);
        std::fill(h->opacities.begin() + long(old), h->opacities.end(), 0.0f);
        std::fill(h->colors.begin() + long(old * 3), h->colors.end(), 0.0f);
    }
    t->gradAccum.assign(n, 0.0f);
    t->gradCount.assign(n, 0);
}

void orc_trainer_step(orc_trainer *t)
{
    if (t->images.empty() || orc_trainer_done(t))
        return;
    if (t->orderPos >= t->order.size()) {
        t->order.resize(t->images.size());
        for (size_t k = 0; k < t->order.size(); ++k)
            t->order[k] = int(k);
        std::shuffle(t->order.begin(), t->order.end(), t->rng);
        t->orderPos = 0;
    }
    const size_t view = size_t(t->order[t->orderPos++]);
    const orc_camera &cam = t->cams[view];
    const int W = cam.width, H = cam.height;
    t->loadTarget(view);
    t->render.resize(size_t(W) * size_t(H) * 3);
    orc_model mv = t->model.view();
    if (t->shDegree > 0)
        orc_raster_forward_sh(&t->raster, &mv, t->shRest.data(), t->shDegree, &cam, &t->opt.raster,
                              t->render.data());
    else
        orc_raster_forward(&t->raster, &mv, &cam, &t->opt.raster, t->render.data());
    t->dLoss.resize(t->render.size());
    const orc_loss_value loss = orc_photometric_loss(t->render.data(), t->target.data(), W, H,
                                                     t->opt.ssim_lambda, t->dLoss.data());
    t->grads.resize(t->model.size());
    t->grads.zero();
    orc_model gv = t->grads.view();
    if (t->shDegree > 0) {
        t->shGrad.assign(t->shRest.size(), 0.0f);
        orc_raster_backward_sh(&t->raster, &mv, t->shRest.data(), t->shDegree, t->dLoss.data(), &gv,
                               t->shGrad.data(), t->gradAccum.data());
    } else {
        orc_raster_backward(&t->raster, &mv, t->dLoss.data(), &gv, t->gradAccum.data());
    }
    for (size_t i = 0; i < t->raster.s.radius.size(); ++i)
        if (t->raster.s.radius[i] > 0)
            ++t->gradCount[i];
    orc_lrs lr = t->opt.lr;
    const float tt = float(t->iteration) / float(std::max(1, t->opt.iterations - 1));
    lr.means = std::exp(std::log(t->opt.lr.means) * (1 - tt) + std::log(t->opt.means_lr_final) * tt) * t->extent;
    orc_model am = t->adamM.view(), av = t->adamV.view();
    orc_adam_step(&mv, &gv, &am, &av, &lr, ++t->adamStep);
    if (t->shDegree > 0)
        orc_adam_array(t->shRest.data(), t->shGrad.data(), t->shM.data(), t->shV.data(),
                       long(t->shRest.size()), t->opt.lr.colors * t->shLrScale, t->adamStep);
    ++t->iteration;
    const float p = orc_psnr(t->render.data(), t->target.data(), long(t->render.size()));
    const float k = t->statIteration == 0 ? 1.0f : 0.05f;
    t->statLoss += k * (loss.total - t->statLoss);
    t->statPsnr += k * (p - t->statPsnr);
    t->statIteration = t->iteration;
    const orc_train_options &o = t->opt;
    if (t->iteration >= o.refine_start && t->iteration < o.refine_stop && t->iteration % o.refine_interval == 0)
        trainerDensify(t);
    if (o.opacity_reset_interval > 0 && t->iteration < o.refine_stop
        && t->iteration % o.opacity_reset_interval == 0) {
        const float cap = logitf_(0.01f);
        for (float &v : t->model.opacities)
            v = std::min(v, cap);
        std::fill(t->adamM.opacities.begin(), t->adamM.opacities.end(), 0.0f);
        std::fill(t->adamV.opacities.begin(), t->adamV.opacities.end(), 0.0f);
        t->opacityWasReset = true;
    }
    t->statGaussians = long(t->model.size());
}

void orc_trainer_get_model(const orc_trainer *t, orc_model *out)
{
    std::copy(t->model.means.begin(), t->model.means.end(), out->means);
    std::copy(t->model.scales.begin(), t->model.scales.end(), out->scales);
    std::copy(t->model.quats.begin(), t->model.quats.end(), out->quats);
    std::copy(t->model.opacities.begin(), t->model.opacities.end(), out->opacities);
    std::copy(t->model.colors.begin(), t->model.colors.end(), out->colors);
}

void orc_trainer_stats(const orc_trainer *t, int *iteration, float *loss, float *psnr, long *gaussians)
{
    *iteration = t->statIteration;
    *loss = t->statLoss;
    *psnr = t->statPsnr;
    *gaussians = t->statGaussians;
}

float orc_trainer_evaluate_psnr(orc_trainer *t)
{
    if (t->images.empty())
        return 0;
    double sum = 0;
    std::vector<float> rgb;
    orc_model mv = t->model.view();
    for (size_t k = 0; k < t->images.size(); ++k) {
        t->loadTarget(k);
        rgb.resize(t->target.size());
        if (t->shDegree > 0)
            orc_raster_forward_sh(&t->raster, &mv, t->shRest.data(), t->shDegree, &t->cams[k],
                                  &t->opt.raster, rgb.data());
        else
            orc_raster_forward(&t->raster, &mv, &t->cams[k], &t->opt.raster, rgb.data());
        sum += orc_psnr(rgb.data(), t->target.data(), long(rgb.size()));
    }
    return float(sum / double(t->images.size()));
}

// Mean absolute error over all training views (the time-to-target metric's L1).
float orc_trainer_evaluate_l1(orc_trainer *t)
{
    if (t->images.empty())
        return 0;
    double sum = 0;
    std::vector<float> rgb;
    orc_model mv = t->model.view();
    for (size_t k = 0; k < t->images.size(); ++k) {
        t->loadTarget(k);
        rgb.resize(t->target.size());
        if (t->shDegree > 0)
            orc_raster_forward_sh(&t->raster, &mv, t->shRest.data(), t->shDegree, &t->cams[k],
                                  &t->opt.raster, rgb.data());
        else
            orc_raster_forward(&t->raster, &mv, &t->cams[k], &t->opt.raster, rgb.data());
        double s = 0;
        for (size_t i = 0; i < rgb.size(); ++i)
            s += std::fabs(double(rgb[i]) - double(t->target[i]));
        sum += s / double(rgb.size());
    }
    return float(sum / double(t->images.size()));
}

void orc_trainer_render(orc_trainer *t, const orc_camera *cam, float *rgb)
{
    orc_model mv = t->model.view();
    orc_raster_forward(&t->raster, &mv, cam, &t->opt.raster, rgb);
}

// ---- viewer (app/src/SplatView.cpp): display attributes at load (:117-180) and the
// front-to-back order of a view (:349-397), with the reference's float expressions.
void orc_view_attributes(const orc_model *m, uint8_t *rgba, float *cov6)
{
    for (size_t i = 0; i < size_t(m->n); ++i) {
        for (int k = 0; k < 3; ++k) {
            const float c = std::clamp(0.5f + kShC0 * m->colors[i * 3 + k], 0.0f, 1.0f);
            rgba[i * 4 + k] = static_cast<uint8_t>(std::lround(c * 255.0f));
        }
        rgba[i * 4 + 3] = static_cast<uint8_t>(std::lround(sigmoidf_(m->opacities[i]) * 255.0f));
        const float *q = &m->quats[i * 4];
        const float qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) + 1e-12f;
        float R[9];
        rotOfQuat(q[0] / qn, q[1] / qn, q[2] / qn, q[3] / qn, R);
        float M[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                M[r * 3 + c] = R[r * 3 + c] * std::exp(m->scales[i * 3 + c]);
        auto sigma = [&](int r, int c) {
            return M[r * 3] * M[c * 3] + M[r * 3 + 1] * M[c * 3 + 1] + M[r * 3 + 2] * M[c * 3 + 2];
        };
        float *o = &cov6[i * 6];
        o[0] = sigma(0, 0);
        o[1] = sigma(0, 1);
        o[2] = sigma(0, 2);
        o[3] = sigma(1, 1);
        o[4] = sigma(1, 2);
        o[5] = sigma(2, 2);
    }
}

void orc_view_depth_order(const float *c, long n, const float *row2, uint32_t *order)
{
    std::vector<float> depth(static_cast<size_t>(n));
    float lo = std::numeric_limits<float>::max(), hi = -lo;
    for (size_t i = 0; i < size_t(n); ++i) {
        const float z = -(row2[0] * c[i * 3] + row2[1] * c[i * 3 + 1] + row2[2] * c[i * 3 + 2] + row2[3]);
        depth[i] = z;
        lo = std::min(lo, z);
        hi = std::max(hi, z);
    }
    const float scale = hi > lo ? 65535.0f / (hi - lo) : 0.0f;
    std::vector<uint32_t> counts(65537, 0);
    std::vector<uint16_t> keys(static_cast<size_t>(n));
    for (size_t i = 0; i < size_t(n); ++i) {
        keys[i] = uint16_t((depth[i] - lo) * scale);
        ++counts[size_t(keys[i]) + 1];
    }
    for (size_t k = 1; k < counts.size(); ++k)
        counts[k] += counts[k - 1];
    for (size_t i = 0; i < size_t(n); ++i)
        order[counts[keys[i]]++] = uint32_t(i);
}

} // extern "C"
