// test_nnp_gpu.cpp -- the reference's force test (src/nnp/test_nnp_force.cpp:11-185)
// restated over the drop-in classes of ann_amd/cpp: a random 20-atom
// structure, forces from NNP::compute against central differences of
// NNP::symm + NNP::energy, plus the reference relations between the entry
// points (energy = compute + vacuum energies, force = compute, sum F = 0).
//
//   test_nnp_gpu roundtrip <in.nnp> <out.nnp>   read + write (no GPU)
//   test_nnp_gpu force <pot.nnp> <seed>         GPU checks, prints errors, exit 0 on pass
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>

#include "nnp_gpu.hpp"

using namespace ann_gpu;

static Structure make_random(int ntypes, unsigned seed) {
    const int natoms = 20;
    const double a0 = 9.0;
    Structure s(natoms);
    s.init({a0, 0, 0, 0, a0, 0, 0, 0, a0});
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> u(-a0, a0);
    for (int i = 0; i < natoms; ++i) {
        s.type(i) = (int)(gen() % ntypes);
        for (int x = 0; x < 3; ++x) s.posn(i)[x] = u(gen);
    }
    return s;
}

static double energy_of(NNP& nnp, Structure s, double rc) {
    NeighborList nl(s, rc);
    NNP::init(nnp, s);
    NNP::symm(nnp, s, nl);
    return NNP::energy(nnp, s);
}

int main(int argc, char** argv) {
    if (argc < 4) { std::fprintf(stderr, "usage: test_nnp_gpu roundtrip|force ...\n"); return 2; }
    const std::string mode = argv[1];
    NNP nnp;
    NNP::read(argv[2], nnp);
    if (mode == "roundtrip") {
        NNP::write(argv[3], nnp);
        std::printf("types %d rc %f params %d\n", nnp.ntypes(), nnp.rc(), nnp.size());
        return 0;
    }
    Structure s = make_random(nnp.ntypes(), (unsigned)std::atoi(argv[3]));
    NeighborList nl(s, nnp.rc());
    NNP::init(nnp, s);
    NNP::symm(nnp, s, nl);
    const double e = NNP::energy(nnp, s);
    NNP::force(nnp, s, nl);
    std::vector<Vec3> f1(s.nAtoms());
    for (int i = 0; i < s.nAtoms(); ++i) f1[i] = s.force(i);
    NNP::compute(nnp, s, nl);
    double evac = 0, fmax = 0, dcf = 0, fsum[3] = {0, 0, 0};
    for (int i = 0; i < s.nAtoms(); ++i) {
        evac += nnp.type(s.type(i)).energy;
        for (int x = 0; x < 3; ++x) {
            fmax = std::fmax(fmax, std::fabs(s.force(i)[x]));
            dcf = std::fmax(dcf, std::fabs(s.force(i)[x] - f1[i][x]));
            fsum[x] += s.force(i)[x];
        }
    }
    const double de = std::fabs(e - (s.pe() + evac));
    // central differences of the FP32 energy (eps 1e-3 A; FP32 rounding makes smaller eps noisy)
    const double eps = 1e-3;
    double err = 0;
    for (int i = 0; i < s.nAtoms(); ++i)
        for (int x = 0; x < 3; ++x) {
            Structure sp = s, sm = s;
            sp.posn(i)[x] += eps;
            sm.posn(i)[x] -= eps;
            const double fd = -(energy_of(nnp, sp, nnp.rc()) - energy_of(nnp, sm, nnp.rc())) / (2 * eps);
            err = std::fmax(err, std::fabs(fd - s.force(i)[x]));
        }
    int nmin = 0, nlist = 0;
    for (int i = 0; i < s.nAtoms(); ++i)
        for (int j = 0; j < nl.size(i); ++j) { ++nlist; nmin += nl.neigh(i, j).min(); }
    std::printf("pairs %d (min images %d) E %.9f |E - (pe + Evac)| %.3e |F(force) - F(compute)| %.3e "
                "max|F| %.4e max|F - FD| %.3e |sum F| %.3e\n",
                nlist, nmin, e, de, dcf, fmax, err, std::fabs(fsum[0]) + std::fabs(fsum[1]) + std::fabs(fsum[2]));
    const bool ok = de < 1e-4 && dcf < 1e-6 * std::fmax(1.0, fmax) && err < 2e-3 * std::fmax(1.0, fmax) &&
                    std::fabs(fsum[0]) + std::fabs(fsum[1]) + std::fabs(fsum[2]) < 1e-6 * std::fmax(1.0, fmax);
    return ok ? 0 : 1;
}
