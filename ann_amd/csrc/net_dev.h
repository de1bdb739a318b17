// net_dev.h -- FP32 activation functions with first and second derivatives
// (AFFPBP2, src/ml/nn.cpp:149-813 for the neurons supported here; softplus
// keeps the reference's truncated ln2, src/math/const.hpp:15).
#pragma once
#include "common.h"

namespace ann {

enum { NEU_LINEAR = 1, NEU_SIGMOID = 2, NEU_TANH = 3, NEU_ARCTAN = 5, NEU_RELU = 6, NEU_ELU = 7,
       NEU_SWISH = 10, NEU_SOFTPLUS = 16, NEU_SQPLUS = 17 };

__host__ inline bool neuron_supported(int n) {
    return n == NEU_LINEAR || n == NEU_SIGMOID || n == NEU_TANH || n == NEU_ARCTAN || n == NEU_RELU ||
           n == NEU_ELU || n == NEU_SWISH || n == NEU_SOFTPLUS || n == NEU_SQPLUS;
}

template <bool D2>
__device__ __forceinline__ void act_f(int neuron, float c, float z, float& a, float& d, float& d2) {
    switch (neuron) {
        case NEU_SIGMOID: {
            const float e = __expf(-c * fabsf(z)), f = 1.0f / (1.0f + e);
            a = z >= 0.0f ? f : e * f;
            d = c * e * f * f;
            if (D2) d2 = (z >= 0.0f ? -1.0f : 1.0f) * c * c * e * (1.0f - e) * f * f * f;
        } break;
        case NEU_TANH: {
            const float t = tanhf(c * z);
            a = t;
            d = c * (1.0f - t * t);
            if (D2) d2 = -2.0f * c * c * t * (1.0f - t * t);
        } break;
        case NEU_ARCTAN: {
            const float zz = c * z * 1.5707963267948966f, f = 1.0f / (1.0f + zz * zz);
            a = 0.6366197723675814f * atanf(zz);
            d = c * f;
            if (D2) d2 = -2.0f * c * zz * f * f * (c * 1.5707963267948966f);  // true d2 (see oracle)
        } break;
        case NEU_RELU:
            a = z > 0.0f ? z : 0.0f;
            d = z > 0.0f ? 1.0f : 0.0f;
            if (D2) d2 = 0.0f;
            break;
        case NEU_ELU:
            if (z >= 0.0f) { a = z; d = 1.0f; if (D2) d2 = 0.0f; }
            else { const float e = __expf(c * z); a = e - 1.0f; d = c * e; if (D2) d2 = c * c * e; }
            break;
        case NEU_SWISH: {
            const float zz = c * z;
            const float e = __expf(-fabsf(zz)), den = 1.0f / (1.0f + e);
            if (zz > 0.0f) {
                a = z * den;
                d = (1.0f + e * (1.0f + zz)) * den * den;
                if (D2) d2 = c * e * (2.0f - zz + e * (2.0f + zz)) * den * den * den;
            } else {
                a = z * e * den;
                d = e * (1.0f + zz + e) * den * den;
                if (D2) d2 = c * e * (2.0f + zz + e * (2.0f - zz)) * den * den * den;
            }
        } break;
        case NEU_SOFTPLUS: {
            const float ln2 = 0.693147180559945f;
            const float e = __expf(-c * fabsf(z)), den = 1.0f / (e + 1.0f);
            a = (z >= 0.0f ? z : 0.0f) + (log1pf(e) - ln2) / c;
            d = z >= 0.0f ? den : e * den;
            if (D2) d2 = c * e * den * den;
        } break;
        case NEU_SQPLUS: {
            const float zz = c * z, s = sqrtf(1.0f + zz * zz);
            a = 0.5f / c * (zz - 1.0f + s);
            d = 0.5f * (zz / s + 1.0f);
            if (D2) d2 = 0.5f * c / (s * s * s);
        } break;
        default:  // linear
            a = z;
            d = 1.0f;
            if (D2) d2 = 0.0f;
            break;
    }
}

}  // namespace ann
