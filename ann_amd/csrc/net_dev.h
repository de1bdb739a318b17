// net_dev.h -- FP32 activation functions with first and second derivatives
// (AFFPBP / AFFPBP2, src/ml/nn.cpp:149-780, every neuron of NN::Neuron;
// softplus and logish keep the reference's truncated ln2, src/math/const.hpp:15).
// Values and first derivatives follow the reference, including the neurons
// that ignore the sharpness c (sqre, serf, logish, atish).  Second derivatives
// (used only by the force-loss second backward) are the true ones, where the
// reference's are zero (gelu, mish, serf, logish, atish), carry an extra c
// (tanhre) or the wrong sign (sqre).
#pragma once
#include "common.h"

namespace ann {

enum { NEU_LINEAR = 1, NEU_SIGMOID = 2, NEU_TANH = 3, NEU_ISRU = 4, NEU_ARCTAN = 5, NEU_RELU = 6, NEU_ELU = 7,
       NEU_TANHRE = 8, NEU_SQRE = 9, NEU_SWISH = 10, NEU_GELU = 11, NEU_MISH = 12, NEU_PFLU = 13, NEU_SERF = 14,
       NEU_LOGISH = 15, NEU_SOFTPLUS = 16, NEU_SQPLUS = 17, NEU_ATISH = 18 };

// every neuron of NN::Neuron (src/ml/nn.hpp:83-107) except UNKNOWN and TEST
__host__ inline bool neuron_supported(int n) { return n >= NEU_LINEAR && n <= NEU_ATISH; }

// logistic sigmoid without overflow
__device__ __forceinline__ float sigm(float u) {
    const float e = __expf(-fabsf(u)), f = 1.0f / (1.0f + e);
    return u >= 0.0f ? f : e * f;
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
        case NEU_ISRU: {
            const float zz = c * z, isr = rsqrtf(1.0f + zz * zz), isr3 = isr * isr * isr;
            a = zz * isr;
            d = c * isr3;
            if (D2) d2 = -3.0f * c * c * zz * isr3 * isr * isr;
        } break;
        case NEU_TANHRE:
            if (z >= 0.0f) { a = z; d = 1.0f; if (D2) d2 = 0.0f; }
            else {
                const float t = tanhf(c * z);
                a = t;
                d = c * (1.0f - t * t);
                if (D2) d2 = -2.0f * c * t * d;
            }
            break;
        case NEU_SQRE:
            if (z >= 0.0f) { a = z; d = 1.0f; if (D2) d2 = 0.0f; }
            else {
                const float den = rsqrtf(1.0f + z * z), den3 = den * den * den;
                a = z * den;
                d = den3;
                if (D2) d2 = -3.0f * z * den3 * den * den;
            }
            break;
        case NEU_GELU: {   // 0.5 z (1 + erf(sqrt(pi)/2 c z))
            const float zz = 0.886226925452758f * c * z, g = __expf(-zz * zz), ep1 = 1.0f + erff(zz);
            a = 0.5f * z * ep1;
            d = 0.5f * (ep1 + c * z * g);
            if (D2) d2 = c * g * (1.0f - zz * zz);
        } break;
        case NEU_MISH: {   // z tanh(softplus(c z))
            const float zz = c * z;
            float f;
            if (zz > 0.0f) {
                const float e = __expf(-zz), e2 = e * e;
                f = 1.0f - 2.0f * e2 / (1.0f + 2.0f * (e + e2));
            } else {
                const float e = __expf(zz), ff = e * (e + 2.0f);
                f = ff / (ff + 2.0f);
            }
            const float s = sigm(zz), fp = (1.0f - f * f) * s;
            a = z * f;
            d = f + zz * fp;
            if (D2) d2 = 2.0f * c * fp + c * zz * fp * ((1.0f - s) - 2.0f * f * s);
        } break;
        case NEU_PFLU: {
            const float zz = c * z, z2 = zz * zz, q = rsqrtf(z2 + 1.0f), q3 = q * q * q;
            a = 0.5f * z * (1.0f + zz * q);
            d = 0.5f * (zz * (z2 + 2.0f) * q3 + 1.0f);
            if (D2) d2 = -0.5f * c * (z2 - 2.0f) * q3 * q * q;
        } break;
        case NEU_SERF: {   // z erf(softplus(z)) (no sharpness)
            const float sp = fmaxf(z, 0.0f) + log1pf(__expf(-fabsf(z))), s = sigm(z);
            const float g = 1.1283791670955126f * __expf(-sp * sp) * s, ef = erff(sp);
            a = z * ef;
            d = ef + z * g;
            if (D2) d2 = 2.0f * g + z * g * ((1.0f - s) - 2.0f * sp * s);
        } break;
        case NEU_LOGISH: {  // z log(1 + sigmoid(z)) / ln2 (no sharpness; the reference's truncated ln2)
            const float li = 1.0f / 0.693147180559945f, s = sigm(z), L = log1pf(s);
            const float h = s * (1.0f - s) / (1.0f + s);
            a = z * L * li;
            d = (L + z * h) * li;
            if (D2) {
                const float hp = s * (1.0f - s) * ((1.0f - 2.0f * s) / (1.0f + s) - h / (1.0f + s));
                d2 = (2.0f * h + z * hp) * li;
            }
        } break;
        case NEU_ATISH: {   // z (1 + 2/pi atan(pi z / 2)) / 2 (no sharpness)
            const float zz = 1.5707963267948966f * z, q = 1.0f / (1.0f + zz * zz), at = atanf(zz);
            a = z * 0.5f * (1.0f + 0.6366197723675814f * at);
            d = 0.5f * (1.0f + z * q) + 0.3183098861837907f * at;
            if (D2) d2 = q * q;
        } break;
        default:  // linear
            a = z;
            d = 1.0f;
            if (D2) d2 = 0.0f;
            break;
    }
}

}  // namespace ann
