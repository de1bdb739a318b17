/*
 * basis.cpp -- oracle: special functions, cutoff functions, radial and angular
 * symmetry functions (FP64).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows:
 *   fmexp (Pade exp2)       src/math/special.cpp:225-275
 *   coscut / sincut         src/math/special.cpp:22-42
 *   Cutoff::cutf/cutg       src/nnp/cutoff.cpp:62-113
 *   BasisR::symm / force    src/nnp/basis_radial.cpp:263-376 (symmd :210-257)
 *   BasisA::init            src/nnp/basis_angular.cpp:178-205
 *   BasisA::symm/symmd/force src/nnp/basis_angular.cpp:212-563
 *
 * Parity pins (tests/test_oracle_basis.py):
 *   fmexp, coscut, sincut, cutoffs: bit-compared with the reference's own
 *     cutoff.cpp + special.cpp compiled from source into oracle/_ref, plus the
 *     closed form of each cutoff and int_0^rc fc = rc/2 (COS, POLY3).
 *   radial / angular values: closed forms (numpy) for GAUSSIAN/GAUSS and the
 *     other families' defining formulas; derivatives by central differences.
 */
#include "oracle.h"
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {
const double kPI = 3.141592653589793;   // src/math/const.hpp:9
const double kRad2 = 1.414213562373095; // src/math/const.hpp:11

// Pade approximant of 2^f on [-0.5,0.5] combined with an exponent-field
// construction of 2^i; the three+two coefficients are the published
// "fastermath" exp2 table the reference uses (kept identical so the oracle
// matches the reference bit for bit).
const double kExp2Q[2] = {2.33184211722314911771e2, 4.36821166879210612817e3};
const double kExp2P[3] = {2.30933477057345225087e-2, 2.02020656693165307700e1,
                          1.51390680115615096133e3};
const double kLog2e = 1.4426950408889634074;

double exp2_pade(double x) {
    const double ip = std::floor(x + 0.5);
    const double fp = x - ip;
    union { double f; uint64_t u; } ep;
    ep.u = (uint64_t)(uint32_t)(((int)ip + 1023) << 20) << 32;
    const double x2 = fp * fp;
    double px = kExp2P[0];
    px = px * x2 + kExp2P[1];
    double qx = x2 + kExp2Q[0];
    px = px * x2 + kExp2P[2];
    qx = qx * x2 + kExp2Q[1];
    px = px * fp;
    return ep.f * (1.0 + 2.0 * (px / (qx - px)));
}

double sech_ref(double x) {  // src/math/special.cpp:87-95
    if (x >= 0) { const double e = std::exp(-x); return 2.0 * e / (1.0 + e * e); }
    const double e = std::exp(x); return 2.0 * e / (e * e + 1.0);
}
double tanh_ref(double x) {  // src/math/special.cpp:68-76
    if (x >= 0) { const double e = std::exp(-2.0 * x); return (1.0 - e) / (1.0 + e); }
    const double e = std::exp(2.0 * x); return (e - 1.0) / (e + 1.0);
}
double powint(double x, int n) {  // src/math/special.cpp:140-149
    if (n == 0) return 1.0;
    if (x == 0.0) return 0.0;
    int nn = n > 0 ? n : -n;
    double y = 1.0, w = x;
    for (; nn != 0; nn >>= 1, w *= w) if (nn & 1) y *= w;
    return n > 0 ? y : 1.0 / y;
}
}  // namespace

extern "C" double orc_fmexp(double x) {
    if (x < -1022.0 / kLog2e) return 0.0;
    if (x > 1023.0 / kLog2e) return INFINITY;
    return exp2_pade(kLog2e * x);
}

extern "C" double orc_coscut(double x) {
    // degree-15 Taylor series of cos about pi/2, written in the shifted
    // variable t = x - pi/2: cos(x) = -sin(t) = sum_k (-1)^(k+1) t^(2k+1)/(2k+1)!
    const double t = x - 0.5 * kPI;
    const double t2 = t * t;
    return t * (-1.0 + t2 * (1.0 / 6.0 + t2 * (-1.0 / 120.0 + t2 * (1.0 / 5040.0 +
           t2 * (-1.0 / 362880.0 + t2 * (1.0 / 39916800.0 + t2 * (-1.0 / 6227020800.0 +
           t2 * 1.0 / 1307674368000.0)))))));
}

extern "C" double orc_sincut(double x) {
    // degree-16 Taylor series of sin about pi/2: sin(x) = cos(t)
    const double t = x - 0.5 * kPI;
    const double t2 = t * t;
    return 1.0 + t2 * (-1.0 / 2.0 + t2 * (1.0 / 24.0 + t2 * (-1.0 / 720.0 + t2 * (1.0 / 40320.0 +
           t2 * (-1.0 / 3628800.0 + t2 * (1.0 / 479001600.0 + t2 * (-1.0 / 87178291200.0 +
           t2 * 1.0 / 20922789888000.0)))))));
}

// Cutoff::compute, src/nnp/cutoff.cpp:115-143 (value and derivative)
extern "C" void orc_cutoff(int name, double rc, double r, double* f, double* g) {
    const double rci = 1.0 / rc;
    const double pirci = kPI / rc;
    switch (name) {
        case ORC_CUT_STEP: *f = (r > rc) ? 0.0 : 1.0; *g = 0.0; break;
        case ORC_CUT_COS:
            if (r < rc) {
                const double arg = r * pirci;
                *f = 0.5 * (orc_coscut(arg) + 1.0);
                *g = -0.5 * pirci * orc_sincut(arg);
            } else { *f = 0.0; *g = 0.0; }
            break;
        case ORC_CUT_TANH:
            if (r < rc) {
                const double t = std::tanh(1.0 - r * rci);
                *f = t * t * t;
                *g = -3.0 * t * t * (1.0 - t * t) * rci;
            } else { *f = 0.0; *g = 0.0; }
            break;
        case ORC_CUT_POLY3:
            if (r < rc) {
                const double x = r * rci;
                *f = (2.0 * x - 3.0) * x * x + 1.0;
                *g = 6.0 * x * (x - 1.0) * rci;
            } else { *f = 0.0; *g = 0.0; }
            break;
        default: throw std::invalid_argument("orc_cutoff: invalid cutoff name");
    }
}

// BasisR::symm (:263-309) and BasisR::symmd (:210-257) for all functions
extern "C" void orc_radial(int name, int cut, double rc, int n, const double* rs, const double* eta,
                           double r, double* G, double* dG) {
    double cf, cg;
    orc_cutoff(cut, rc, r, &cf, &cg);
    for (int i = 0; i < n; ++i) {
        const double e = eta[i], x = r - rs[i];
        double v = 0, d = 0;
        switch (name) {
            case ORC_RAD_GAUSSIAN: {
                const double a = e * x;
                const double ex = orc_fmexp(-a * a);
                v = ex * cf;
                d = ex * (-2.0 * e * a * cf + cg);
            } break;
            case ORC_RAD_SECH: {
                const double ex = orc_fmexp(-e * x);
                v = 2.0 * ex / (1.0 + ex * ex) * cf;
                const double ex2 = ex * ex, den = 1.0 / (1.0 + ex2);
                d = 2.0 * ex * den * (-1.0 * e * (1.0 - ex2) * den * cf + cg);
            } break;
            case ORC_RAD_LOGISTIC: {
                const double ex = orc_fmexp(-2.0 * e * x);
                v = 4.0 * ex / ((1.0 + ex) * (1.0 + ex)) * cf;
                const double sf = sech_ref(e * x), tf = tanh_ref(e * x);
                d = sf * sf * (-2.0 * e * tf * cf + cg);
            } break;
            case ORC_RAD_TANH: {
                const double t = std::tanh(-e * x);
                v = 0.5 * (t + 1.0) * cf;
                const double tf = tanh_ref(-e * x);
                d = 0.5 * (-e * (1.0 - tf * tf) * cf + (1.0 + tf) * cg);
            } break;
            case ORC_RAD_LOGCOSH: {
                const double ex = orc_fmexp(-2.0 * e * x);
                v = 0.5 * std::log1p(ex) * cf;
                d = -e * ex / (1.0 + ex) * cf + 0.5 * std::log1p(ex) * cg;
            } break;
            case ORC_RAD_LOGCOSH2: {
                const double a = -e * x;
                const double ex = orc_fmexp(2.0 * kRad2 * a);
                v = (a == 0.0) ? 0.5 * cf / kRad2 : a * ex / (ex - 1.0) * cf;
                const double den = 1.0 / (ex - 1.0);
                d = (a == 0.0) ? 0.5 * (-e * cg + cf / kRad2)
                               : ex * (e * (1.0 + 2.0 * kRad2 * a - ex) * den * cf + a * cg) * den;
            } break;
            default: throw std::invalid_argument("orc_radial: invalid radial function");
        }
        G[i] = v;
        if (dG) dG[i] = d;
    }
}

// BasisR::force, src/nnp/basis_radial.cpp:316-376: amp = -sum dEdG_i dG_i/dr
extern "C" double orc_radial_force(int name, int cut, double rc, int n, const double* rs,
                                   const double* eta, double r, const double* dEdG) {
    std::vector<double> G(n), dG(n);
    orc_radial(name, cut, rc, n, rs, eta, r, G.data(), dG.data());
    double amp = 0;
    for (int i = 0; i < n; ++i) amp -= dEdG[i] * dG[i];
    return amp;
}

// Distance factor of one angular function and its derivative w.r.t. r2s.
// GAUSS: exp(-r2s/eta^2); GAUSS2: 2e/(1+e); SECH: 2e/(1+e^2) with
// e = exp(-pi/2 r2s/eta^2); STUDENTk: (1+r2s/eta^2)^(-(k+1)/2)
// (src/nnp/basis_angular.cpp:212-249, 259-364)
static void ang_radial(int name, double ieta2, double r2s, double* fr, double* dfr_dr2s) {
    switch (name) {
        case ORC_ANG_GAUSS: {
            const double e = orc_fmexp(-ieta2 * r2s);
            *fr = e; *dfr_dr2s = -ieta2 * e;
        } break;
        case ORC_ANG_GAUSS2: {
            const double e = orc_fmexp(-0.5 * kPI * ieta2 * r2s);
            const double den = 1.0 / (1.0 + e);
            *fr = 2.0 * e * den;
            *dfr_dr2s = *fr * (-0.5 * kPI * ieta2 * den);
        } break;
        case ORC_ANG_SECH: {
            const double e = orc_fmexp(-0.5 * kPI * ieta2 * r2s);
            const double e2 = e * e, den = 1.0 / (1.0 + e2);
            *fr = 2.0 * e * den;
            *dfr_dr2s = *fr * (-0.5 * kPI * ieta2 * (1.0 - e2) * den);
        } break;
        case ORC_ANG_STUDENT3: case ORC_ANG_STUDENT4: case ORC_ANG_STUDENT5: {
            const int p = (name == ORC_ANG_STUDENT3) ? 4 : (name == ORC_ANG_STUDENT4 ? 5 : 6);
            const double den = 1.0 / std::sqrt(1.0 + ieta2 * r2s);
            *fr = powint(den, p);
            *dfr_dr2s = *fr * (-0.5 * p * ieta2 * den * den);
        } break;
        default: throw std::invalid_argument("orc_angular: invalid angular function");
    }
}

// BasisA::symm (:371-420) and BasisA::symmd (:259-364) for one triplet:
// G = |1/2 + lambda/2 cos|^zeta * fr(r0^2+r1^2) * fc(r0) fc(r1)
extern "C" void orc_angular(int name, int cut, double rc, int n, const double* eta,
                            const double* zeta, const int* lambda, double r0, double r1,
                            double cos, double* G, double* dGdcos, double* dGdr0, double* dGdr1) {
    double c0, g0, c1, g1;
    orc_cutoff(cut, rc, r0, &c0, &g0);
    orc_cutoff(cut, rc, r1, &c1, &g1);
    const double r2s = r0 * r0 + r1 * r1;
    for (int i = 0; i < n; ++i) {
        const double ieta2 = 1.0 / (eta[i] * eta[i]);
        const double lf = 0.5 * lambda[i];
        double fr, dfr;
        ang_radial(name, ieta2, r2s, &fr, &dfr);
        const double cw = std::fabs(0.5 + lf * cos);
        const double wz = std::pow(cw, zeta[i]);
        G[i] = wz * fr * c0 * c1;
        if (dGdcos) {
            dGdcos[i] = zeta[i] * lf * std::pow(cw, zeta[i] - 1.0) * fr * c0 * c1;
            dGdr0[i] = wz * (2.0 * r0 * dfr * c0 + fr * g0) * c1;
            dGdr1[i] = wz * (2.0 * r1 * dfr * c1 + fr * g1) * c0;
        }
    }
}

// BasisA::force, src/nnp/basis_angular.cpp:430-563:
// phi = -sum dEdG dG/dcos, eta = -sum dEdG dG/dr
extern "C" void orc_angular_force(int name, int cut, double rc, int n, const double* eta,
                                  const double* zeta, const int* lambda, double r0, double r1,
                                  double cos, const double* dEdG, double* phi, double* eta_out) {
    std::vector<double> G(n), dc(n), d0(n), d1(n);
    orc_angular(name, cut, rc, n, eta, zeta, lambda, r0, r1, cos, G.data(), dc.data(), d0.data(),
                d1.data());
    double p = 0, e0 = 0, e1 = 0;
    for (int i = 0; i < n; ++i) {
        p -= dEdG[i] * dc[i];
        e0 -= dEdG[i] * d0[i];
        e1 -= dEdG[i] * d1[i];
    }
    *phi = p;
    eta_out[0] = e0;
    eta_out[1] = e1;
}
