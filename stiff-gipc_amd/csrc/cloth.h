// cloth.h -- constants shared by the cloth kernels and the host set-up.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace sgipc
{

// per-triangle record in the element buffer: 6 upper 3x3 blocks + 9 gradient
// entries; per-hinge record: 10 upper blocks + 12 gradient entries
constexpr int kTriStride   = 63;
constexpr int kHingeStride = 102;
constexpr int kTriPairA[6] = {0, 0, 0, 1, 1, 2};
constexpr int kTriPairB[6] = {0, 1, 2, 1, 2, 2};

// Signed dihedral angle of the hinge (q0, q1 | q2, q3), the angle
// LibShell::edgeTheta (femEnergy.cu:142-260) measures between
// n0 = (q0-q2) x (q1-q2) and n1 = (q1-q3) x (q0-q3) about q1 - q0, written
// in the edge vectors e_k = q_k - q0:
//   theta = atan2(-V |e1|, (e1.e2)(e1.e3) - |e1|^2 (e2.e3)),  V = e1.(e2 x e3).
__host__ __device__ inline double hinge_theta(const double* q0, const double* q1,
                                              const double* q2, const double* q3)
{
    double e1[3], e2[3], e3[3];
    for(int i = 0; i < 3; i++)
    {
        e1[i] = q1[i] - q0[i];
        e2[i] = q2[i] - q0[i];
        e3[i] = q3[i] - q0[i];
    }
    double V = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) + e1[1] * (e2[2] * e3[0] - e2[0] * e3[2])
               + e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
    double s = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    double p = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
    double q = e1[0] * e3[0] + e1[1] * e3[1] + e1[2] * e3[2];
    double t = e2[0] * e3[0] + e2[1] * e3[1] + e2[2] * e3[2];
    return atan2(-V * sqrt(s), p * q - s * t);
}

}  // namespace sgipc
