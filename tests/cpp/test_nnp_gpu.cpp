// test_nnp_gpu.cpp -- the reference's force test (src/nnp/test_nnp_force.cpp:11-185)
// restated over the drop-in classes of ann_amd/cpp: a random 20-atom
// structure, forces from NNP::compute against central differences of
// NNP::symm + NNP::energy, plus the reference relations between the entry
// points (energy = compute + vacuum energies, force = compute, sum F = 0).
//
//   test_nnp_gpu roundtrip <in.nnp> <out.nnp>   read + write (no GPU)
//   test_nnp_gpu basis <in.nnp> <basis.txt> <name> <out.nnp>   read_basis + write (no GPU)
//   test_nnp_gpu force <pot.nnp> <seed>         GPU checks, prints errors, exit 0 on pass
//   test_nnp_gpu dump <pot.nnp> <struc.txt> <out.txt> <skin> [<basis.txt> <name>]
//       every entry point on one structure, written out for comparison with
//       the oracle: NeighborList overloads at positions 0, then verlet::List
//       overloads (list built at positions 0 with rc + skin, evaluated at
//       positions 1).  With a basis file, NNP::read_basis replaces the basis
//       of species <name> first.  struc.txt: N, the 9 entries of R (columns
//       are lattice vectors), then N lines "type x0 y0 z0 x1 y1 z1".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
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

static void put_symm(FILE* o, const char* tag, const Structure& s) {
    std::fprintf(o, "%s\n", tag);
    for (int i = 0; i < s.nAtoms(); ++i) {
        for (double g : s.symm(i)) std::fprintf(o, "%.17g ", g);
        std::fprintf(o, "\n");
    }
}

static void put_force(FILE* o, const char* tag, const Structure& s) {
    std::fprintf(o, "%s\n", tag);
    for (int i = 0; i < s.nAtoms(); ++i)
        std::fprintf(o, "%.17g %.17g %.17g\n", s.force(i)[0], s.force(i)[1], s.force(i)[2]);
}

static int dump(NNP& nnp, int argc, char** argv) {
    if (argc < 6) { std::fprintf(stderr, "usage: test_nnp_gpu dump pot struc out skin [basis name]\n"); return 2; }
    if (argc >= 8) NNP::read_basis(argv[6], nnp, argv[7]);
    std::ifstream in(argv[3]);
    int N = 0;
    in >> N;
    std::array<double, 9> R;
    for (auto& r : R) in >> r;
    Structure s(N);
    s.init(R);
    std::vector<Vec3> x0(N), x1(N);
    for (int i = 0; i < N; ++i) {
        in >> s.type(i);
        for (int x = 0; x < 3; ++x) in >> x0[i][x];
        for (int x = 0; x < 3; ++x) in >> x1[i][x];
        s.posn(i) = x0[i];
    }
    if (!in) { std::fprintf(stderr, "bad structure file\n"); return 2; }
    FILE* o = std::fopen(argv[4], "w");
    // NeighborList overloads at positions 0
    NeighborList nl(s, nnp.rc());
    NNP::init(nnp, s);
    NNP::symm(nnp, s, nl);
    put_symm(o, "G0", s);
    std::fprintf(o, "E0\n%.17g\n", NNP::energy(nnp, s));
    NNP::force(nnp, s, nl);
    put_force(o, "Fforce0", s);
    NNP::compute(nnp, s, nl);
    std::fprintf(o, "pe0\n%.17g\n", s.pe());
    put_force(o, "Fcompute0", s);
    // verlet::List overloads: list at positions 0 (rc + skin), evaluation at positions 1
    verlet::List vl;
    vl.rc() = nnp.rc() + std::atof(argv[5]);
    vl.build(s);
    NNP::compute(nnp, s, vl);                 // takes the list over (positions 0)
    std::fprintf(o, "pev0\n%.17g\n", s.pe());
    for (int i = 0; i < N; ++i) s.posn(i) = x1[i];
    NNP::symm(nnp, s, vl);                    // positions-only refresh
    put_symm(o, "G1", s);
    std::fprintf(o, "E1\n%.17g\n", NNP::energy(nnp, s));
    NNP::force(nnp, s, vl);
    put_force(o, "Fforce1", s);
    NNP::compute(nnp, s, vl);
    std::fprintf(o, "pe1\n%.17g\n", s.pe());
    put_force(o, "Fcompute1", s);
    int P = 0;
    for (int i = 0; i < N; ++i) P += vl.size(i);
    std::fprintf(o, "vpairs\n%d\n", P);
    std::fclose(o);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 4) { std::fprintf(stderr, "usage: test_nnp_gpu roundtrip|force ...\n"); return 2; }
    const std::string mode = argv[1];
    NNP nnp;
    NNP::read(argv[2], nnp);
    if (mode == "dump") return dump(nnp, argc, argv);
    if (mode == "basis") {        // read_basis + write, no GPU call: basis <pot> <basis> <name> <out>
        if (argc < 6) return 2;
        NNP::read_basis(argv[3], nnp, argv[4]);
        NNP::write(argv[5], nnp);
        std::printf("types %d params %d\n", nnp.ntypes(), nnp.size());
        return 0;
    }
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
