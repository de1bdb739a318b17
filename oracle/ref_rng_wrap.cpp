/*
 * ref_rng_wrap.cpp -- C entry points over the reference's own random
 * distributions (src/rand/rng.cpp, with src/math/special.cpp), compiled from
 * /root/reference by oracle/Makefile into oracle/_ref/libref_rng.so, plus the
 * standard library's std::mt19937 and std::uniform_real_distribution<double>
 * that the reference trainer draws from (src/nnp/nnptefr2.cpp:2160-2168,
 * 2409-2466).  TEST INFRASTRUCTURE ONLY: the Taylor-augmentation tests check
 * the package's restatement of these draws bit for bit against it.
 */
#include <random>

#include "rand/rng.hpp"

// n draws of distribution dist (rng::dist::Name numbering: UNIFORM, LAPLACE,
// NORMAL, LOGISTIC, SECH, COSINE as below) with location 0 and scale s from a
// std::mt19937 seeded with seed
extern "C" void ref_rng_draw(int dist, double s, unsigned seed, int n, double* out) {
    std::mt19937 gen(seed);
    rng::dist::Uniform u(0.0, s);
    rng::dist::Laplace l(0.0, s);
    rng::dist::Normal g(0.0, s);
    rng::dist::Logistic lo(0.0, s);
    rng::dist::Sech se(0.0, s);
    rng::dist::Cosine co(0.0, s);
    for (int i = 0; i < n; ++i) {
        switch (dist) {
            case 0: out[i] = u.rand(gen); break;
            case 1: out[i] = l.rand(gen); break;
            case 2: out[i] = g.rand(gen); break;
            case 3: out[i] = lo.rand(gen); break;
            case 4: out[i] = se.rand(gen); break;
            default: out[i] = co.rand(gen); break;
        }
    }
}

// n raw outputs of std::mt19937(seed)
extern "C" void ref_mt_raw(unsigned seed, int n, unsigned* out) {
    std::mt19937 gen(seed);
    for (int i = 0; i < n; ++i) out[i] = (unsigned)gen();
}

// n draws of std::uniform_real_distribution<double>(a, b) from std::mt19937(seed)
extern "C" void ref_uniform(unsigned seed, double a, double b, int n, double* out) {
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> d(a, b);
    for (int i = 0; i < n; ++i) out[i] = d(gen);
}

// interleaved draws in the order of one displaced atom of the Taylor
// augmentation (src/nnp/nnptefr2.cpp:2447-2466): length from dist, then two
// U(-1, 1) draws; the trainer's seed chain first (:2160-2168, one MPI rank):
// rngen = mt19937(seed_global); seed_local = rngen(); rngen = mt19937(seed_local)
extern "C" void ref_taylor_draws(int dist, double s, unsigned seed_global, int natoms, double* out) {
    std::mt19937 g0(seed_global);
    const int seed_local = (int)g0();
    std::mt19937 gen(seed_local);
    std::uniform_real_distribution<double> dista(-1.0, 1.0);
    rng::dist::Uniform u(0.0, s);
    rng::dist::Laplace l(0.0, s);
    rng::dist::Normal g(0.0, s);
    rng::dist::Logistic lo(0.0, s);
    rng::dist::Sech se(0.0, s);
    rng::dist::Cosine co(0.0, s);
    for (int j = 0; j < natoms; ++j) {
        double vr;
        switch (dist) {
            case 0: vr = u.rand(gen); break;
            case 1: vr = l.rand(gen); break;
            case 2: vr = g.rand(gen); break;
            case 3: vr = lo.rand(gen); break;
            case 4: vr = se.rand(gen); break;
            default: vr = co.rand(gen); break;
        }
        const double ct = dista(gen);
        const double ph = dista(gen);
        out[3 * j] = vr;
        out[3 * j + 1] = ct;
        out[3 * j + 2] = ph;
    }
}
