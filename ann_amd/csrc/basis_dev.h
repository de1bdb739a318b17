// basis_dev.h -- FP32 device forms of the cutoff, radial and angular
// symmetry functions (generic over every family the reference defines).
//
// Cutoff::compute       src/nnp/cutoff.cpp:115-143
// BasisR symm / symmd   src/nnp/basis_radial.cpp:175-309
// BasisA symmf / symmd  src/nnp/basis_angular.cpp:212-364
#pragma once
#include "common.h"

namespace ann {

enum { CUT_STEP = 1, CUT_COS = 2, CUT_TANH = 3, CUT_POLY3 = 4 };
enum { RAD_GAUSSIAN = 1, RAD_SECH = 2, RAD_LOGISTIC = 3, RAD_TANH = 4, RAD_LOGCOSH = 5, RAD_LOGCOSH2 = 6 };
enum { ANG_GAUSS = 1, ANG_GAUSS2 = 2, ANG_SECH = 3, ANG_STUDENT3 = 4, ANG_STUDENT4 = 5, ANG_STUDENT5 = 6 };

__device__ __forceinline__ void cutoff_fg(int name, float rc, float rci, float pirc, float r, float& f,
                                          float& g) {
    if (!(r < rc)) { f = (name == CUT_STEP && r <= rc) ? 1.0f : 0.0f; g = 0.0f; return; }
    switch (name) {
        case CUT_COS: {
            float s, c;
            sincosf(r * pirc, &s, &c);
            f = 0.5f * (c + 1.0f);
            g = -0.5f * pirc * s;
        } break;
        case CUT_TANH: {
            const float t = tanhf(1.0f - r * rci);
            f = t * t * t;
            g = -3.0f * t * t * (1.0f - t * t) * rci;
        } break;
        case CUT_POLY3: {
            const float x = r * rci;
            f = (2.0f * x - 3.0f) * x * x + 1.0f;
            g = 6.0f * x * (x - 1.0f) * rci;
        } break;
        default: f = 1.0f; g = 0.0f; break;
    }
}

// sech and tanh of y from exp(-|y|), no overflow
__device__ __forceinline__ void sech_tanh(float y, float& s, float& t) {
    const float e = __expf(-fabsf(y)), e2 = e * e, den = 1.0f / (1.0f + e2);
    s = 2.0f * e * den;
    t = copysignf((1.0f - e2) * den, y);
}

// value and d/dr of one radial function at distance r (fc, gc: cutoff, cutoff')
__device__ __forceinline__ void radial_fd(int name, float eta, float rs, float r, float fc, float gc,
                                          float& v, float& d) {
    const float x = r - rs;
    switch (name) {
        case RAD_GAUSSIAN: {
            const float a = eta * x, e = __expf(-a * a);
            v = e * fc;
            d = e * (-2.0f * eta * a * fc + gc);
        } break;
        case RAD_SECH: {  // sech(eta x)
            float s, t;
            sech_tanh(eta * x, s, t);
            v = s * fc;
            d = s * (-eta * t * fc + gc);
        } break;
        case RAD_LOGISTIC: {  // sech^2(eta x)
            float s, t;
            sech_tanh(eta * x, s, t);
            v = s * s * fc;
            d = s * s * (-2.0f * eta * t * fc + gc);
        } break;
        case RAD_TANH: {  // (tanh(-eta x) + 1)/2
            const float t = tanhf(-eta * x);
            v = 0.5f * (t + 1.0f) * fc;
            d = 0.5f * (-eta * (1.0f - t * t) * fc + (1.0f + t) * gc);
        } break;
        case RAD_LOGCOSH: {  // log(1 + exp(-2 eta x))/2
            const float y = -2.0f * eta * x;
            const float lp = fmaxf(y, 0.0f) + log1pf(__expf(-fabsf(y)));  // log(1+e^y)
            const float sg = 1.0f / (1.0f + __expf(-y));                    // e^y/(1+e^y)
            v = 0.5f * lp * fc;
            d = -eta * sg * fc + 0.5f * lp * gc;
        } break;
        case RAD_LOGCOSH2: {  // a e/(e - 1), a = -eta x, e = exp(2 sqrt2 a)
            const float k2 = 1.414213562373095f;
            const float a = -eta * x;
            if (a == 0.0f) {
                // the reference's value at exactly a == 0 (src/nnp/basis_radial.cpp:250,
                // :367), kept as is: it is not the a -> 0 limit of the branch below
                v = 0.5f * fc / k2;
                d = 0.5f * (-eta * gc + fc / k2);
            } else {
                // e - 1 from expm1 and e - 1 - y by its series where the difference
                // cancels, so that small |a| keeps FP32 accuracy
                const float y = 2.0f * k2 * a;
                const float em1 = expm1f(y), e = em1 + 1.0f, den = 1.0f / em1;
                const float c = fabsf(y) < 0.25f
                                    ? y * y * (0.5f + y * (1.0f / 6.0f + y * (1.0f / 24.0f + y * (1.0f / 120.0f +
                                                           y * (1.0f / 720.0f + y * (1.0f / 5040.0f))))))
                                    : em1 - y;
                v = a * e * den * fc;
                d = e * (-eta * c * den * fc + a * gc) * den;
            }
        } break;
        default: v = 0.0f; d = 0.0f; break;
    }
}

// distance factor fr(r2s) of one angular function and dfr/dr2s
__device__ __forceinline__ void ang_radial(int name, float ieta2, float r2s, float& fr, float& dfr) {
    const float hp = 1.5707963267948966f;
    switch (name) {
        case ANG_GAUSS: {
            const float e = __expf(-ieta2 * r2s);
            fr = e;
            dfr = -ieta2 * e;
        } break;
        case ANG_GAUSS2: {
            const float e = __expf(-hp * ieta2 * r2s), den = 1.0f / (1.0f + e);
            fr = 2.0f * e * den;
            dfr = fr * (-hp * ieta2 * den);
        } break;
        case ANG_SECH: {
            const float e = __expf(-hp * ieta2 * r2s), e2 = e * e, den = 1.0f / (1.0f + e2);
            fr = 2.0f * e * den;
            dfr = fr * (-hp * ieta2 * (1.0f - e2) * den);
        } break;
        default: {  // STUDENT3/4/5: (1 + ieta2 r2s)^(-p/2), p = 4, 5, 6
            const int p = name == ANG_STUDENT3 ? 4 : (name == ANG_STUDENT4 ? 5 : 6);
            const float den = rsqrtf(1.0f + ieta2 * r2s);
            float v = 1.0f;
            for (int q = 0; q < p; ++q) v *= den;
            fr = v;
            dfr = v * (-0.5f * p * ieta2 * den * den);
        } break;
    }
}

// one angular function: value, d/dcos, d/dr0, d/dr1 (reference symmf/symmd)
__device__ __forceinline__ void angular_fd(int name, float ieta2, float zeta, float lamh, float r0,
                                           float r1, float c, float f0, float f1, float g0, float g1,
                                           float& G, float& dc, float& d0, float& d1) {
    float fr, dfr;
    ang_radial(name, ieta2, r0 * r0 + r1 * r1, fr, dfr);
    const float cw = fabsf(0.5f + lamh * c);
    const float wz1 = powf(cw, zeta - 1.0f);
    const float wz = wz1 * cw;
    G = wz * fr * f0 * f1;
    dc = zeta * lamh * wz1 * fr * f0 * f1;
    d0 = wz * (2.0f * r0 * dfr * f0 + fr * g0) * f1;
    d1 = wz * (2.0f * r1 * dfr * f1 + fr * g1) * f0;
}

__device__ __forceinline__ int pair_idx(int j, int k) {
    return j > k ? j * (j + 1) / 2 + k : k * (k + 1) / 2 + j;
}

// w^z for a wave-uniform exponent z (integer when zi >= 0)
__device__ __forceinline__ float upow(float w, int zi, float z) {
    if (zi >= 0) {
        float r = 1.0f, b = w;
        for (int e = zi; e; e >>= 1) {
            if (e & 1) r *= b;
            b *= b;
        }
        return r;
    }
    return w > 0.0f ? exp2f(z * __log2f(w)) : 0.0f;
}

}  // namespace ann
