// test_classes_gpu.cpp -- the per-function classes of ann_amd/cpp
// (Cutoff, BasisR, BasisA, NN::ANN, Cost, DODZ, D2ODZDI with the reference's
// member names) driven one point at a time, as the reference's callers do,
// and with the batched overloads; every result is written out for
// comparison with the oracle.
//
//   test_classes_gpu <spec.txt> <out.txt>
// spec: cut rc | radname n rs[n] eta[n] | angname n eta[n] zeta[n] lambda[n] |
//       npts r[npts] | ntrip (r0 r1 cos)[ntrip] | dEdG_r[npts*n] | dEdG_a[ntrip*n] |
//       nin nlayer nodes[nlayer] neuron c outw outb | np params[np] inpw[nin] inpb[nin] |
//       nrow x[nrow*nin] dcdo[nrow] v[nrow*nin]
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "classes_gpu.hpp"

using namespace ann_gpu;

template <typename T>
static std::vector<T> readv(std::istream& in, size_t n) {
    std::vector<T> v(n);
    for (auto& x : v) in >> x;
    return v;
}

static void put(FILE* o, const char* tag, const std::vector<double>& v) {
    std::fprintf(o, "%s %zu", tag, v.size());
    for (double x : v) std::fprintf(o, " %.17g", x);
    std::fprintf(o, "\n");
}

int main(int argc, char** argv) {
    if (argc < 3) { std::fprintf(stderr, "usage: test_classes_gpu spec out\n"); return 2; }
    std::ifstream in(argv[1]);
    int cut, radname, nr, angname, na, npts, ntrip;
    double rc;
    in >> cut >> rc >> radname >> nr;
    const auto rs = readv<double>(in, nr), reta = readv<double>(in, nr);
    in >> angname >> na;
    const auto aeta = readv<double>(in, na), zeta = readv<double>(in, na);
    const auto lam = readv<int>(in, na);
    in >> npts;
    const auto r = readv<double>(in, npts);
    in >> ntrip;
    const auto trip = readv<double>(in, 3 * (size_t)ntrip);
    const auto dEr = readv<double>(in, (size_t)npts * nr), dEa = readv<double>(in, (size_t)ntrip * na);
    int nin, nlayer;
    in >> nin >> nlayer;
    const auto nodes = readv<int>(in, nlayer);
    int neuron, np, nrow;
    double c, outw, outb;
    in >> neuron >> c >> outw >> outb >> np;
    const auto params = readv<double>(in, np), inpw = readv<double>(in, nin), inpb = readv<double>(in, nin);
    in >> nrow;
    const auto x = readv<double>(in, (size_t)nrow * nin), dcdo = readv<double>(in, nrow),
               v = readv<double>(in, (size_t)nrow * nin);
    if (!in) { std::fprintf(stderr, "bad spec\n"); return 2; }
    FILE* o = std::fopen(argv[2], "w");

    // ---- Cutoff
    const Cutoff cf((Cutoff::Name)cut, rc);
    std::vector<double> f1, g1, fb, gb;
    for (double d : r) { f1.push_back(cf.cutf(d)); g1.push_back(cf.cutg(d)); }
    cf.compute(r, fb, gb);
    put(o, "cutf", f1); put(o, "cutg", g1); put(o, "cutf_b", fb); put(o, "cutg_b", gb);

    // ---- BasisR
    BasisR br(rc, (Cutoff::Name)cut, nr, (BasisR::Name)radname);
    br.rs() = rs;
    br.eta() = reta;
    std::vector<double> rsy, rfo, rsf, rsd, rsyb, rfob;
    for (int p = 0; p < npts; ++p) {
        const std::vector<double> dr(1, r[p]);
        br.symm(dr);
        rsy.insert(rsy.end(), br.symm().begin(), br.symm().end());
        rfo.push_back(br.force(dr, dEr.data() + (size_t)p * nr));
        rsf.push_back(br.symmf(r[p], reta[0], rs[0]));
        rsd.push_back(br.symmd(r[p], reta[0], rs[0]));
    }
    br.symm(r, rsyb);
    br.force(r, dEr, rfob);
    put(o, "rsymm", rsy); put(o, "rforce", rfo); put(o, "rsymmf", rsf); put(o, "rsymmd", rsd);
    put(o, "rsymm_b", rsyb); put(o, "rforce_b", rfob);

    // ---- BasisA
    BasisA ba(rc, (Cutoff::Name)cut, na, (BasisA::Name)angname);
    ba.eta() = aeta;
    ba.zeta() = zeta;
    ba.lambda() = lam;
    std::vector<double> asy, afo, asf, asd, legs, coss, asyb, phib, etab;
    for (int t = 0; t < ntrip; ++t) {
        const std::vector<double> dr = {trip[3 * t], trip[3 * t + 1]}, cs(1, trip[3 * t + 2]);
        legs.insert(legs.end(), dr.begin(), dr.end());
        coss.push_back(cs[0]);
        ba.symm(dr, cs);
        asy.insert(asy.end(), ba.symm().begin(), ba.symm().end());
        double phi, eta[2];
        ba.force(dr, cs, phi, eta, dEa.data() + (size_t)t * na);
        afo.insert(afo.end(), {phi, eta[0], eta[1]});
        asf.push_back(ba.symmf(cs[0], dr.data(), aeta[0], zeta[0], lam[0]));
        double fphi, feta[2];
        ba.symmd(fphi, feta, cs[0], dr.data(), aeta[0], zeta[0], lam[0]);
        asd.insert(asd.end(), {fphi, feta[0], feta[1]});
    }
    ba.symm(legs, coss, asyb);
    ba.force(legs, coss, dEa, phib, etab);
    put(o, "asymm", asy); put(o, "aforce", afo); put(o, "asymmf", asf); put(o, "asymmd", asd);
    put(o, "asymm_b", asyb); put(o, "aphi_b", phib); put(o, "aeta_b", etab);

    // ---- NN::ANN and its gradients
    NN::ANN nn;
    nn.resize(nin, std::vector<int>(nodes.begin(), nodes.end() - 1), nodes.back());
    nn.neuron() = (NN::Neuron)neuron;
    nn.c() = c;
    nn.outw() = outw;
    nn.outb() = outb;
    nn.params() = params;
    nn.inpw() = inpw;
    nn.inpb() = inpb;
    std::vector<double> outs, dodis;
    NN::DODZ dodz(nn);
    for (int q = 0; q < nrow; ++q) {
        const std::vector<double> row(x.begin() + (size_t)q * nin, x.begin() + (size_t)(q + 1) * nin);
        outs.push_back(nn.fpbp(row)[0]);
        dodz.grad(nn);
        dodis.insert(dodis.end(), dodz.dodi().begin(), dodz.dodi().end());
    }
    put(o, "out", outs); put(o, "dodi", dodis);
    std::vector<double> outb_v, dodib;
    nn.fp(x, outb_v, &dodib);
    put(o, "out_b", outb_v); put(o, "dodi_b", dodib);
    // Cost and D2ODZDI at row 0 (the network's current input after fpbp2)
    nn.fpbp2(std::vector<double>(x.begin(), x.begin() + nin));
    NN::Cost cost(nn);
    put(o, "cost0", cost.grad(nn, std::vector<double>(1, dcdo[0])));
    NN::D2ODZDI d2(nn);
    d2.grad(nn);
    std::vector<double> flat;
    for (const auto& row : d2.d2odpdi()) flat.insert(flat.end(), row.begin(), row.end());
    put(o, "d2odpdi0", flat);
    put(o, "cost_b", cost.grad(nn, x, dcdo));
    std::vector<double> rop;
    NN::D2ODZDI::grad_rop(nn, x, dcdo, v, rop);
    put(o, "rop_b", rop);
    // a parameter change reaches the device (operator<< of the reference)
    nn.params()[0] += 0.25;
    put(o, "out_shift", std::vector<double>(1, nn.fp(std::vector<double>(x.begin(), x.begin() + nin))[0]));
    // the reference's errors for invalid names and sizes
    int errors = 0;
    try { Cutoff bad(Cutoff::Name::NONE, rc); } catch (const std::invalid_argument&) { ++errors; }
    try { Cutoff bad(Cutoff::Name::COS, -1.0); } catch (const std::invalid_argument&) { ++errors; }
    try { BasisR bad; bad.symm(std::vector<double>(1, 1.0)); } catch (const std::invalid_argument&) { ++errors; }
    try { NN::ANN bad; bad.resize(0, {4}); } catch (const std::invalid_argument&) { ++errors; }
    try { NN::ANN bad; bad.resize(3, {4}); bad.neuron() = (NN::Neuron)99; bad.fp(std::vector<double>(3, 0.0)); }
    catch (const std::invalid_argument&) { ++errors; }
    put(o, "errors", std::vector<double>(1, errors));
    std::fclose(o);
    return 0;
}
