// pairvec_dev.h -- the periodic pair vector r_i - r_j - R n in FP64 (device).
#pragma once
#include "common.h"

namespace ann {

// r_i - r_j - R n in FP64, evaluated exactly as the oracle does (no
// contraction) so that inclusion decisions agree bit for bit.
__device__ __forceinline__ void pair_vec(const double* __restrict__ pos, int i, int j,
                                         const double* __restrict__ R, int n0, int n1, int n2,
                                         double& x, double& y, double& z, double& d2) {
#pragma clang fp contract(off)
    const double m0 = n0, m1 = n1, m2 = n2;
    x = __dsub_rn(__dsub_rn(pos[3 * i], pos[3 * j]),
                  __dadd_rn(__dadd_rn(__dmul_rn(R[0], m0), __dmul_rn(R[1], m1)), __dmul_rn(R[2], m2)));
    y = __dsub_rn(__dsub_rn(pos[3 * i + 1], pos[3 * j + 1]),
                  __dadd_rn(__dadd_rn(__dmul_rn(R[3], m0), __dmul_rn(R[4], m1)), __dmul_rn(R[5], m2)));
    z = __dsub_rn(__dsub_rn(pos[3 * i + 2], pos[3 * j + 2]),
                  __dadd_rn(__dadd_rn(__dmul_rn(R[6], m0), __dmul_rn(R[7], m1)), __dmul_rn(R[8], m2)));
    d2 = __dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z));
}

// the same expression on positions already in registers
__device__ __forceinline__ void pair_vec_v(double pix, double piy, double piz, double pjx, double pjy, double pjz,
                                           const double* __restrict__ R, int n0, int n1, int n2,
                                           double& x, double& y, double& z, double& d2) {
#pragma clang fp contract(off)
    const double m0 = n0, m1 = n1, m2 = n2;
    x = __dsub_rn(__dsub_rn(pix, pjx), __dadd_rn(__dadd_rn(__dmul_rn(R[0], m0), __dmul_rn(R[1], m1)), __dmul_rn(R[2], m2)));
    y = __dsub_rn(__dsub_rn(piy, pjy), __dadd_rn(__dadd_rn(__dmul_rn(R[3], m0), __dmul_rn(R[4], m1)), __dmul_rn(R[5], m2)));
    z = __dsub_rn(__dsub_rn(piz, pjz), __dadd_rn(__dadd_rn(__dmul_rn(R[6], m0), __dmul_rn(R[7], m1)), __dmul_rn(R[8], m2)));
    d2 = __dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z));
}

}  // namespace ann
