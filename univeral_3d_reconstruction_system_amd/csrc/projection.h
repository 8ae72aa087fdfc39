// EWA projection of one Gaussian (Rasterizer.cpp:24-110), shared by the projection kernel and
// the chain-rule kernel of the backward pass. Compiled with floating-point contraction off so
// every value (and thus depth keys, radii and tile rectangles) equals the reference's.
#pragma once
#include "common.h"

namespace u3d {

struct Projection {
    float pc[3];
    float q[4];
    float qNorm;
    float Rq[9];
    float s[3];
    float M[9];
    float Sigma[9];
    float tx, ty, limTx, limTy;
    bool clampX, clampY;
    float T0[3], T1[3];
    float a, b, c;
};

// Rotation matrix of a normalised quaternion (w, x, y, z), row-major (Rasterizer.cpp:24-29).
__device__ __forceinline__ void rotOfQuat(float w, float x, float y, float z, float R[9])
{
    R[0] = 1.0f - 2.0f * (y * y + z * z);
    R[1] = 2.0f * (x * y - w * z);
    R[2] = 2.0f * (x * z + w * y);
    R[3] = 2.0f * (x * y + w * z);
    R[4] = 1.0f - 2.0f * (x * x + z * z);
    R[5] = 2.0f * (y * z - w * x);
    R[6] = 2.0f * (x * z - w * y);
    R[7] = 2.0f * (y * z + w * x);
    R[8] = 1.0f - 2.0f * (x * x + y * y);
}

// Returns false when the Gaussian is in front of the near plane (Rasterizer.cpp:48-110).
__device__ __forceinline__ bool projectGaussian(const float mu[3], const float qr[4],
                                                const float logScale[3], const u3d_camera &cam,
                                                float nearPlane, float eps2d, Projection &p)
{
    const float *R = cam.R;
    p.pc[0] = R[0] * mu[0] + R[1] * mu[1] + R[2] * mu[2] + cam.t[0];
    p.pc[1] = R[3] * mu[0] + R[4] * mu[1] + R[5] * mu[2] + cam.t[1];
    p.pc[2] = R[6] * mu[0] + R[7] * mu[1] + R[8] * mu[2] + cam.t[2];
    const float z = p.pc[2];
    if (z < nearPlane)
        return false;
    // A near-zero quaternion becomes the identity with norm 1. Written as selects: a branch
    // with a chained assignment kept two components in private memory.
    const float qn = sqrtf(qr[0] * qr[0] + qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
    const bool tiny = qn < 1e-12f;
    p.qNorm = tiny ? 1.0f : qn;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        p.q[k] = tiny ? (k == 0 ? 1.0f : 0.0f) : qr[k] / qn;
    rotOfQuat(p.q[0], p.q[1], p.q[2], p.q[3], p.Rq);
#pragma unroll
    for (int k = 0; k < 3; ++k)
        p.s[k] = expfExact(logScale[k]);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            p.M[r * 3 + c] = p.Rq[r * 3 + c] * p.s[c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            p.Sigma[r * 3 + c] = p.M[r * 3] * p.M[c * 3] + p.M[r * 3 + 1] * p.M[c * 3 + 1]
                                 + p.M[r * 3 + 2] * p.M[c * 3 + 2];
    const float guard = 0.3f * 0.5f;
    const float W = float(cam.width), H = float(cam.height);
    const float limXPos = (W - cam.cx) / cam.fx + guard * W / cam.fx;
    const float limXNeg = cam.cx / cam.fx + guard * W / cam.fx;
    const float limYPos = (H - cam.cy) / cam.fy + guard * H / cam.fy;
    const float limYNeg = cam.cy / cam.fy + guard * H / cam.fy;
    const float xz = p.pc[0] / z, yz = p.pc[1] / z;
    // std::min(a, b) = (b < a) ? b : a and std::max(a, b) = (a < b) ? b : a.
    const float lo = -limXNeg, loy = -limYNeg;
    const float mx = (lo < xz) ? xz : lo;
    const float cxz = (mx < limXPos) ? mx : limXPos;
    const float my = (loy < yz) ? yz : loy;
    const float cyz = (my < limYPos) ? my : limYPos;
    p.clampX = cxz != xz;
    p.clampY = cyz != yz;
    p.limTx = cxz;
    p.limTy = cyz;
    p.tx = z * cxz;
    p.ty = z * cyz;
    const float J00 = cam.fx / z, J02 = -cam.fx * p.tx / (z * z);
    const float J11 = cam.fy / z, J12 = -cam.fy * p.ty / (z * z);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        p.T0[k] = J00 * R[k] + J02 * R[6 + k];
        p.T1[k] = J11 * R[3 + k] + J12 * R[6 + k];
    }
    float S0[3], S1[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        S0[r] = p.Sigma[r * 3] * p.T0[0] + p.Sigma[r * 3 + 1] * p.T0[1] + p.Sigma[r * 3 + 2] * p.T0[2];
        S1[r] = p.Sigma[r * 3] * p.T1[0] + p.Sigma[r * 3 + 1] * p.T1[1] + p.Sigma[r * 3 + 2] * p.T1[2];
    }
    p.a = p.T0[0] * S0[0] + p.T0[1] * S0[1] + p.T0[2] * S0[2] + eps2d;
    p.b = p.T0[0] * S1[0] + p.T0[1] * S1[1] + p.T0[2] * S1[2];
    p.c = p.T1[0] * S1[0] + p.T1[1] * S1[1] + p.T1[2] * S1[2] + eps2d;
    return true;
}

// ------------------------------------------------------------------ spherical harmonics
// View-dependent colour of degree 1..3 (an extension: the reference trains the DC term only,
// GaussianModel.h:21). Real SH basis of the unit direction from the camera centre to the mean,
// with the standard normalisation constants; coefficients are stored 15 x 3 per Gaussian,
// coefficient-major. Degree 0 keeps the reference's colour expression exactly.
constexpr float kShC1 = 0.4886025119029199f;
constexpr float kShC2_0 = 1.0925484305920792f, kShC2_1 = -1.0925484305920792f,
                kShC2_2 = 0.31539156525252005f, kShC2_3 = -1.0925484305920792f,
                kShC2_4 = 0.5462742152960396f;
constexpr float kShC3_0 = -0.5900435899266435f, kShC3_1 = 2.890611442640554f,
                kShC3_2 = -0.4570457994644658f, kShC3_3 = 0.3731763325901154f,
                kShC3_4 = -0.4570457994644658f, kShC3_5 = 1.445305721320277f,
                kShC3_6 = -0.5900435899266435f;

__host__ __device__ inline int shCount(int degree)
{
    return degree <= 0 ? 0 : (degree == 1 ? 3 : (degree == 2 ? 8 : 15));
}

// Basis values b[k] and, when g != nullptr, their partial derivatives g[k] = (d/dx, d/dy, d/dz).
__device__ __forceinline__ void shBasis(float x, float y, float z, int degree, float b[15],
                                        float3 *g)
{
    const float xx = x * x, yy = y * y, zz = z * z;
    if (degree >= 1) {
        b[0] = -kShC1 * y;
        b[1] = kShC1 * z;
        b[2] = -kShC1 * x;
        if (g) {
            g[0] = make_float3(0.0f, -kShC1, 0.0f);
            g[1] = make_float3(0.0f, 0.0f, kShC1);
            g[2] = make_float3(-kShC1, 0.0f, 0.0f);
        }
    }
    if (degree >= 2) {
        b[3] = kShC2_0 * (x * y);
        b[4] = kShC2_1 * (y * z);
        b[5] = kShC2_2 * (2.0f * zz - xx - yy);
        b[6] = kShC2_3 * (x * z);
        b[7] = kShC2_4 * (xx - yy);
        if (g) {
            g[3] = make_float3(kShC2_0 * y, kShC2_0 * x, 0.0f);
            g[4] = make_float3(0.0f, kShC2_1 * z, kShC2_1 * y);
            g[5] = make_float3(kShC2_2 * (-2.0f * x), kShC2_2 * (-2.0f * y), kShC2_2 * (4.0f * z));
            g[6] = make_float3(kShC2_3 * z, 0.0f, kShC2_3 * x);
            g[7] = make_float3(kShC2_4 * (2.0f * x), kShC2_4 * (-2.0f * y), 0.0f);
        }
    }
    if (degree >= 3) {
        b[8] = kShC3_0 * (y * (3.0f * xx - yy));
        b[9] = kShC3_1 * (x * y * z);
        b[10] = kShC3_2 * (y * (4.0f * zz - xx - yy));
        b[11] = kShC3_3 * (z * (2.0f * zz - 3.0f * xx - 3.0f * yy));
        b[12] = kShC3_4 * (x * (4.0f * zz - xx - yy));
        b[13] = kShC3_5 * (z * (xx - yy));
        b[14] = kShC3_6 * (x * (xx - 3.0f * yy));
        if (g) {
            g[8] = make_float3(kShC3_0 * (6.0f * x * y), kShC3_0 * (3.0f * xx - 3.0f * yy), 0.0f);
            g[9] = make_float3(kShC3_1 * (y * z), kShC3_1 * (x * z), kShC3_1 * (x * y));
            g[10] = make_float3(kShC3_2 * (-2.0f * x * y), kShC3_2 * (4.0f * zz - xx - 3.0f * yy),
                                kShC3_2 * (8.0f * y * z));
            g[11] = make_float3(kShC3_3 * (-6.0f * x * z), kShC3_3 * (-6.0f * y * z),
                                kShC3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy));
            g[12] = make_float3(kShC3_4 * (4.0f * zz - 3.0f * xx - yy), kShC3_4 * (-2.0f * x * y),
                                kShC3_4 * (8.0f * x * z));
            g[13] = make_float3(kShC3_5 * (2.0f * x * z), kShC3_5 * (-2.0f * y * z), kShC3_5 * (xx - yy));
            g[14] = make_float3(kShC3_6 * (3.0f * xx - 3.0f * yy), kShC3_6 * (-6.0f * x * y), 0.0f);
        }
    }
}

// Unit direction from the camera centre to mean `mu`, and the distance.
__device__ __forceinline__ void viewDir(const float mu[3], float3 campos, float d[3], float *len)
{
    const float vx = mu[0] - campos.x, vy = mu[1] - campos.y, vz = mu[2] - campos.z;
    *len = sqrtf(vx * vx + vy * vy + vz * vz);
    d[0] = vx / *len;
    d[1] = vy / *len;
    d[2] = vz / *len;
}

// Camera centre -R^T t in float (host and device compute it the same way).
__host__ __device__ inline float3 cameraCentre(const u3d_camera &c)
{
    float v[3];
    for (int k = 0; k < 3; ++k)
        v[k] = -(c.R[k] * c.t[0] + c.R[3 + k] * c.t[1] + c.R[6 + k] * c.t[2]);
    return make_float3(v[0], v[1], v[2]);
}

} // namespace u3d
