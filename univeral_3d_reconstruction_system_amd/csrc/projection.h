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
    p.qNorm = sqrtf(qr[0] * qr[0] + qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
    if (p.qNorm < 1e-12f) {
        p.qNorm = 1.0f;
        p.q[0] = 1.0f;
        p.q[1] = p.q[2] = p.q[3] = 0.0f;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            p.q[k] = qr[k] / p.qNorm;
    }
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

} // namespace u3d
