// net_dev.h -- FP32 and FP64 activation functions with first and second derivatives
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

// Layout of a hidden layer's weights for the vector-ALU network (net_vec.hip):
// rows k = 0 .. ni-1 of W[k][j], j contiguous, with an odd row stride, so that
// lanes over the neurons (forward: one row) and lanes over the inputs
// (backward: one column) both read distinct LDS banks.
__host__ __device__ inline int vec_row_stride(int no) { return no | 1; }
// floats of a region of n, rounded to whole 16-byte units
__host__ __device__ inline int vec_pad4(int n) { return (n + 3) / 4 * 4; }

// every neuron of NN::Neuron (src/ml/nn.hpp:83-107) except UNKNOWN and TEST
__host__ inline bool neuron_supported(int n) { return n >= NEU_LINEAR && n <= NEU_ATISH; }

// the math of one precision: FP32 uses the fast device intrinsics of the
// network kernels, FP64 the full-precision library functions
__device__ __forceinline__ float mexp(float x) { return __expf(x); }
__device__ __forceinline__ double mexp(double x) { return exp(x); }
__device__ __forceinline__ float mtanh(float x) { return tanhf(x); }
__device__ __forceinline__ double mtanh(double x) { return tanh(x); }
__device__ __forceinline__ float matan(float x) { return atanf(x); }
__device__ __forceinline__ double matan(double x) { return atan(x); }
__device__ __forceinline__ float mlog1p(float x) { return log1pf(x); }
__device__ __forceinline__ double mlog1p(double x) { return log1p(x); }
__device__ __forceinline__ float msqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double msqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float mrsqrt(float x) { return rsqrtf(x); }
__device__ __forceinline__ double mrsqrt(double x) { return 1.0 / sqrt(x); }
__device__ __forceinline__ float merf(float x) { return erff(x); }
__device__ __forceinline__ double merf(double x) { return erf(x); }
__device__ __forceinline__ float mabs(float x) { return fabsf(x); }
__device__ __forceinline__ double mabs(double x) { return fabs(x); }
__device__ __forceinline__ float mmax(float x, float y) { return fmaxf(x, y); }
__device__ __forceinline__ double mmax(double x, double y) { return fmax(x, y); }

// logistic sigmoid without overflow
template <typename T>
__device__ __forceinline__ T sigm(T u) {
    const T e = mexp(-mabs(u)), f = T(1.0) / (T(1.0) + e);
    return u >= T(0.0) ? f : e * f;
}

template <bool D2, typename T>
__device__ __forceinline__ void act_f(int neuron, T c, T z, T& a, T& d, T& d2) {
    switch (neuron) {
        case NEU_SIGMOID: {
            const T e = mexp(-c * mabs(z)), f = T(1.0) / (T(1.0) + e);
            a = z >= T(0.0) ? f : e * f;
            d = c * e * f * f;
            if (D2) d2 = (z >= T(0.0) ? -T(1.0) : T(1.0)) * c * c * e * (T(1.0) - e) * f * f * f;
        } break;
        case NEU_TANH: {
            const T t = mtanh(c * z);
            a = t;
            d = c * (T(1.0) - t * t);
            if (D2) d2 = -T(2.0) * c * c * t * (T(1.0) - t * t);
        } break;
        case NEU_ARCTAN: {
            const T zz = c * z * T(1.5707963267948966), f = T(1.0) / (T(1.0) + zz * zz);
            a = T(0.6366197723675814) * matan(zz);
            d = c * f;
            if (D2) d2 = -T(2.0) * c * zz * f * f * (c * T(1.5707963267948966));  // true d2 (see oracle)
        } break;
        case NEU_RELU:
            a = z > T(0.0) ? z : T(0.0);
            d = z > T(0.0) ? T(1.0) : T(0.0);
            if (D2) d2 = T(0.0);
            break;
        case NEU_ELU:
            if (z >= T(0.0)) { a = z; d = T(1.0); if (D2) d2 = T(0.0); }
            else { const T e = mexp(c * z); a = e - T(1.0); d = c * e; if (D2) d2 = c * c * e; }
            break;
        case NEU_SWISH: {
            const T zz = c * z;
            const T e = mexp(-mabs(zz)), den = T(1.0) / (T(1.0) + e);
            if (zz > T(0.0)) {
                a = z * den;
                d = (T(1.0) + e * (T(1.0) + zz)) * den * den;
                if (D2) d2 = c * e * (T(2.0) - zz + e * (T(2.0) + zz)) * den * den * den;
            } else {
                a = z * e * den;
                d = e * (T(1.0) + zz + e) * den * den;
                if (D2) d2 = c * e * (T(2.0) + zz + e * (T(2.0) - zz)) * den * den * den;
            }
        } break;
        case NEU_SOFTPLUS: {
            const T ln2 = T(0.693147180559945);
            const T e = mexp(-c * mabs(z)), den = T(1.0) / (e + T(1.0));
            a = (z >= T(0.0) ? z : T(0.0)) + (mlog1p(e) - ln2) / c;
            d = z >= T(0.0) ? den : e * den;
            if (D2) d2 = c * e * den * den;
        } break;
        case NEU_SQPLUS: {
            const T zz = c * z, s = msqrt(T(1.0) + zz * zz);
            a = T(0.5) / c * (zz - T(1.0) + s);
            d = T(0.5) * (zz / s + T(1.0));
            if (D2) d2 = T(0.5) * c / (s * s * s);
        } break;
        case NEU_ISRU: {
            const T zz = c * z, isr = mrsqrt(T(1.0) + zz * zz), isr3 = isr * isr * isr;
            a = zz * isr;
            d = c * isr3;
            if (D2) d2 = -T(3.0) * c * c * zz * isr3 * isr * isr;
        } break;
        case NEU_TANHRE:
            if (z >= T(0.0)) { a = z; d = T(1.0); if (D2) d2 = T(0.0); }
            else {
                const T t = mtanh(c * z);
                a = t;
                d = c * (T(1.0) - t * t);
                if (D2) d2 = -T(2.0) * c * t * d;
            }
            break;
        case NEU_SQRE:
            if (z >= T(0.0)) { a = z; d = T(1.0); if (D2) d2 = T(0.0); }
            else {
                const T den = mrsqrt(T(1.0) + z * z), den3 = den * den * den;
                a = z * den;
                d = den3;
                if (D2) d2 = -T(3.0) * z * den3 * den * den;
            }
            break;
        case NEU_GELU: {   // 0.5 z (1 + erf(sqrt(pi)/2 c z))
            const T zz = T(0.886226925452758) * c * z, g = mexp(-zz * zz), ep1 = T(1.0) + merf(zz);
            a = T(0.5) * z * ep1;
            d = T(0.5) * (ep1 + c * z * g);
            if (D2) d2 = c * g * (T(1.0) - zz * zz);
        } break;
        case NEU_MISH: {   // z tanh(softplus(c z))
            const T zz = c * z;
            T f;
            if (zz > T(0.0)) {
                const T e = mexp(-zz), e2 = e * e;
                f = T(1.0) - T(2.0) * e2 / (T(1.0) + T(2.0) * (e + e2));
            } else {
                const T e = mexp(zz), ff = e * (e + T(2.0));
                f = ff / (ff + T(2.0));
            }
            const T s = sigm(zz), fp = (T(1.0) - f * f) * s;
            a = z * f;
            d = f + zz * fp;
            if (D2) d2 = T(2.0) * c * fp + c * zz * fp * ((T(1.0) - s) - T(2.0) * f * s);
        } break;
        case NEU_PFLU: {
            const T zz = c * z, z2 = zz * zz, q = mrsqrt(z2 + T(1.0)), q3 = q * q * q;
            a = T(0.5) * z * (T(1.0) + zz * q);
            d = T(0.5) * (zz * (z2 + T(2.0)) * q3 + T(1.0));
            if (D2) d2 = -T(0.5) * c * (z2 - T(2.0)) * q3 * q * q;
        } break;
        case NEU_SERF: {   // z erf(softplus(z)) (no sharpness)
            const T sp = mmax(z, T(0.0)) + mlog1p(mexp(-mabs(z))), s = sigm(z);
            const T g = T(1.1283791670955126) * mexp(-sp * sp) * s, ef = merf(sp);
            a = z * ef;
            d = ef + z * g;
            if (D2) d2 = T(2.0) * g + z * g * ((T(1.0) - s) - T(2.0) * sp * s);
        } break;
        case NEU_LOGISH: {  // z log(1 + sigmoid(z)) / ln2 (no sharpness; the reference's truncated ln2)
            const T li = T(1.0) / T(0.693147180559945), s = sigm(z), L = mlog1p(s);
            const T h = s * (T(1.0) - s) / (T(1.0) + s);
            a = z * L * li;
            d = (L + z * h) * li;
            if (D2) {
                const T hp = s * (T(1.0) - s) * ((T(1.0) - T(2.0) * s) / (T(1.0) + s) - h / (T(1.0) + s));
                d2 = (T(2.0) * h + z * hp) * li;
            }
        } break;
        case NEU_ATISH: {   // z (1 + 2/pi atan(pi z / 2)) / 2 (no sharpness)
            const T zz = T(1.5707963267948966) * z, q = T(1.0) / (T(1.0) + zz * zz), at = matan(zz);
            a = z * T(0.5) * (T(1.0) + T(0.6366197723675814) * at);
            d = T(0.5) * (T(1.0) + z * q) + T(0.3183098861837907) * at;
            if (D2) d2 = q * q;
        } break;
        default:  // linear
            a = z;
            d = T(1.0);
            if (D2) d2 = T(0.0);
            break;
    }
}

// act_f over M values with the neuron's switch taken once: each case holds the
// M evaluations of one neuron, unrolled.  Inlining the per-value switch M
// times instead makes the hot loop of a network kernel tens of kilobytes long
// (past the instruction cache) and branches through the case chain per value.
template <bool D2, int M>
__device__ __forceinline__ void act_many(int neuron, float c, const float (&z)[M], float (&a)[M], float (&d)[M],
                                         float (&d2)[M]) {
#define ANN_ACT_CASE(NEU)                                                        \
    case NEU:                                                                    \
        _Pragma("unroll") for (int q = 0; q < M; ++q) act_f<D2>(NEU, c, z[q], a[q], d[q], d2[q]); \
        break;
    switch (neuron) {
        ANN_ACT_CASE(NEU_SIGMOID)
        ANN_ACT_CASE(NEU_TANH)
        ANN_ACT_CASE(NEU_ISRU)
        ANN_ACT_CASE(NEU_ARCTAN)
        ANN_ACT_CASE(NEU_RELU)
        ANN_ACT_CASE(NEU_ELU)
        ANN_ACT_CASE(NEU_TANHRE)
        ANN_ACT_CASE(NEU_SQRE)
        ANN_ACT_CASE(NEU_SWISH)
        ANN_ACT_CASE(NEU_GELU)
        ANN_ACT_CASE(NEU_MISH)
        ANN_ACT_CASE(NEU_PFLU)
        ANN_ACT_CASE(NEU_SERF)
        ANN_ACT_CASE(NEU_LOGISH)
        ANN_ACT_CASE(NEU_SOFTPLUS)
        ANN_ACT_CASE(NEU_SQPLUS)
        ANN_ACT_CASE(NEU_ATISH)
        default:
            ANN_ACT_CASE(NEU_LINEAR)
    }
#undef ANN_ACT_CASE
}

}  // namespace ann
