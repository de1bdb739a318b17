/*
 * oracle_impl.h -- internal C++ helpers shared by the oracle sources.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#pragma once
#include "oracle.h"
#include <vector>

namespace orc {

const double kPI = 3.141592653589793;  // src/math/const.hpp:9

void activation(int neuron, double c, double z, double& a, double& d, double& d2);

// G-vector layout of a centre species (NNH::init_input, src/nnp/nnh.cpp:94-112)
inline int npairs(int S) { return S * (S + 1) / 2; }
inline int pair_index(int j, int k) { return j > k ? j * (j + 1) / 2 + k : k * (k + 1) / 2 + j; }
int ninput(const orc_model* m, int a);
int offset_r(const orc_model* m, int a, int b);
int offset_a(const orc_model* m, int a, int p);  // includes nInputR

struct Net {
    int nin = 0, nl = 0, neuron = 0, nparam = 0;
    double c = 1, outw = 1, outb = 0;
    std::vector<int> nodes;
    std::vector<const double*> W, b;
    const double* inpw = nullptr;
    const double* inpb = nullptr;
    struct State {
        std::vector<double> s;
        std::vector<std::vector<double>> z, a, d, d2;
    };
    Net(const orc_model* m, int sp);
    int nprev(int l) const;
    double forward(const double* G, State& st) const;
    void dedg(const State& st, double* out) const;
    void grad_rop(const State& st, const double* v, double gE, double* grad) const;
};

struct NList {
    int natoms = 0;
    std::vector<int> off, idx, img, rev, minimg;
    std::vector<double> r, d;
};

}  // namespace orc
