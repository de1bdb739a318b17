/*
 * ref_wrap.cpp -- C entry points over the reference's own Cutoff and
 * math::special sources (src/nnp/cutoff.cpp, src/math/special.cpp), compiled
 * from /root/reference by oracle/Makefile into oracle/_ref/libref_cutoff.so.
 * TEST INFRASTRUCTURE ONLY: used by tests to check the oracle's restatement
 * of the cutoff functions and of fmexp bit for bit.  These are the only parts
 * of the reference path that compile without Eigen and without the absent
 * src/struc/ tree.
 */
#include "nnp/cutoff.hpp"
#include "math/special.hpp"

extern "C" double ref_cutf(int name, double rc, double r) {
    Cutoff c((Cutoff::Name::Type)name, rc);
    return c.cutf(r);
}
extern "C" double ref_cutg(int name, double rc, double r) {
    Cutoff c((Cutoff::Name::Type)name, rc);
    return c.cutg(r);
}
extern "C" void ref_cut_compute(int name, double rc, double r, double* v, double* g) {
    Cutoff c((Cutoff::Name::Type)name, rc);
    c.compute(r, *v, *g);
}
extern "C" double ref_fmexp(double x) { return math::special::fmexp(x); }
extern "C" double ref_coscut(double x) { return math::special::coscut(x); }
extern "C" double ref_sincut(double x) { return math::special::sincut(x); }
