// classes_gpu.hpp -- the reference's per-function classes over the FP64
// evaluators of libann_amd.so (include/ann_gpu.h, "FP64 evaluators"):
//
//   Cutoff::cutf / cutg / compute        src/nnp/cutoff.hpp:59-61
//   BasisR::symmf / symmd / symm / force src/nnp/basis_radial.hpp:84-88
//   BasisA::symmf / symmd / symm / force src/nnp/basis_angular.hpp:98-102
//   NN::ANN::fp / fpbp / fpbp2           src/ml/nn.hpp:357-362
//   NN::Cost::grad                       src/ml/nn.hpp:453
//   NN::DODZ::grad (dodi)                src/ml/nn.hpp:485
//   NN::D2ODZDI::grad (d2odpdi)          src/ml/nn.hpp:588
//
// Same names, argument meaning and errors (std::invalid_argument for invalid
// names and sizes); vectors are std::vector instead of Eigen types.  A call
// with one distance, triplet or input vector is one GPU launch; each class
// also takes many at once (the overloads with vectors of points), which is
// the form that uses the GPU well.  Networks have one output.
#pragma once
#include <array>
#include <memory>
#include <vector>

struct ann_mlp;

namespace ann_gpu {

class Cutoff {
public:
    enum class Name { NONE = 0, STEP = 1, COS = 2, TANH = 3, POLY3 = 4 };
    Cutoff() = default;
    Cutoff(Name name, double rc);
    Name name() const { return name_; }
    double rc() const { return rc_; }
    double cutf(double dr) const;
    double cutg(double dr) const;
    void compute(double dr, double& v, double& g) const;
    // many distances in one launch
    void compute(const std::vector<double>& dr, std::vector<double>& v, std::vector<double>& g) const;

private:
    Name name_ = Name::NONE;
    double rc_ = 0;
};

class BasisR {
public:
    enum class Name { NONE = 0, GAUSSIAN = 1, SECH = 2, LOGISTIC = 3, TANH = 4, LOGCOSH = 5, LOGCOSH2 = 6 };
    BasisR() = default;
    BasisR(double rc, Cutoff::Name cutname, int nf, Name name);
    Name& name() { return name_; }
    Name name() const { return name_; }
    const Cutoff& cutoff() const { return cutoff_; }
    int size() const { return (int)rs_.size(); }
    void resize(int n);
    double& rs(int i) { return rs_[i]; }
    double rs(int i) const { return rs_[i]; }
    double& eta(int i) { return eta_[i]; }
    double eta(int i) const { return eta_[i]; }
    std::vector<double>& rs() { return rs_; }
    std::vector<double>& eta() { return eta_; }
    std::vector<double>& symm() { return symm_; }
    const std::vector<double>& symm() const { return symm_; }
    // one function at distance dr: value, d/dr
    double symmf(double dr, double eta, double rs) const;
    double symmd(double dr, double eta, double rs) const;
    // every function at distance dr[0] into symm()
    void symm(const std::vector<double>& dr);
    // -sum_i dEdG[i] dG_i/dr at distance dr[0]
    double force(const std::vector<double>& dr, const double* dEdG) const;
    // many distances: G [n][size()], and amp[p] for dEdG [n][size()]
    void symm(const std::vector<double>& r, std::vector<double>& G) const;
    void force(const std::vector<double>& r, const std::vector<double>& dEdG, std::vector<double>& amp) const;

private:
    void eval(int n, const double* r, double* G, double* dG, const double* dEdG, double* amp) const;
    Name name_ = Name::NONE;
    Cutoff cutoff_;
    std::vector<double> rs_, eta_, symm_;
};

class BasisA {
public:
    enum class Name { NONE = 0, GAUSS = 1, GAUSS2 = 2, SECH = 3, STUDENT3 = 4, STUDENT4 = 5, STUDENT5 = 6 };
    BasisA() = default;
    BasisA(double rc, Cutoff::Name cutname, int nf, Name name);
    Name& name() { return name_; }
    Name name() const { return name_; }
    const Cutoff& cutoff() const { return cutoff_; }
    int size() const { return (int)eta_.size(); }
    void resize(int n);
    double& eta(int i) { return eta_[i]; }
    double& zeta(int i) { return zeta_[i]; }
    int& lambda(int i) { return lambda_[i]; }
    std::vector<double>& eta() { return eta_; }
    std::vector<double>& zeta() { return zeta_; }
    std::vector<int>& lambda() { return lambda_; }
    std::vector<double>& symm() { return symm_; }
    const std::vector<double>& symm() const { return symm_; }
    // one function of a triplet (cos, legs dr[0], dr[1])
    double symmf(double cos, const double dr[2], double eta, double zeta, int lambda) const;
    // its derivatives: fphi = dG/dcos, feta = (dG/dr0, dG/dr1)
    void symmd(double& fphi, double* feta, double cos, const double dr[2], double eta, double zeta, int lambda) const;
    // every function of the triplet (dr[0], dr[1], cos[0]) into symm()
    void symm(const std::vector<double>& dr, const std::vector<double>& cos);
    // phi = -sum dEdG dG/dcos, eta = -sum dEdG (dG/dr0, dG/dr1)
    void force(const std::vector<double>& dr, const std::vector<double>& cos, double& phi, double* eta,
               const double* dEdG) const;
    // many triplets: r [n][2], cos [n] -> G [n][size()]; phi [n], eta [n][2]
    void symm(const std::vector<double>& r, const std::vector<double>& cos, std::vector<double>& G) const;
    void force(const std::vector<double>& r, const std::vector<double>& cos, const std::vector<double>& dEdG,
               std::vector<double>& phi, std::vector<double>& eta) const;

private:
    void eval(int n, const double* r, const double* cos, double* G, double* dG, const double* dEdG, double* phi,
              double* eta) const;
    Name name_ = Name::NONE;
    Cutoff cutoff_;
    std::vector<double> eta_, zeta_, symm_;
    std::vector<int> lambda_;
};

namespace NN {

// NN::Neuron numbering (src/ml/nn.hpp:83-107)
enum class Neuron {
    LINEAR = 1, SIGMOID, TANH, ISRU, ARCTAN, RELU, ELU, TANHRE, SQRE, SWISH, GELU, MISH, PFLU, SERF, LOGISH,
    SOFTPLUS, SQPLUS, ATISH
};

class ANN {
public:
    ANN();
    ~ANN();
    ANN(const ANN& o);
    ANN& operator=(const ANN& o);
    // nInput inputs, hidden layers nNodes, then the output layer of nOutput (= 1)
    // nodes; parameters zero, input scaling identity
    void resize(int nInput, const std::vector<int>& nNodes, int nOutput = 1);
    int nlayer() const { return (int)nodes_.size(); }
    int nInp() const { return nin_; }
    int nOut() const { return nodes_.empty() ? 0 : nodes_.back(); }
    int nNodes(int l) const { return nodes_[l]; }
    int size() const { return (int)params_.size(); }
    int nBias() const;
    int nWeight() const { return size() - nBias(); }
    double& c() { return c_; }
    Neuron& neuron() { return neuron_; }
    std::vector<double>& inpw() { return inpw_; }
    std::vector<double>& inpb() { return inpb_; }
    double& outw() { return outw_; }
    double& outb() { return outb_; }
    std::vector<double>& inp() { return inp_; }
    const std::vector<double>& inp() const { return inp_; }
    const std::vector<double>& out() const { return out_; }
    // reference packing: biases layer by layer, then weights column-major
    std::vector<double>& params() { return params_; }
    const std::vector<double>& params() const { return params_; }
    const std::vector<double>& fp();
    const std::vector<double>& fpbp() { return fp(); }
    const std::vector<double>& fpbp2() { return fp(); }
    const std::vector<double>& fp(const std::vector<double>& inp) { inp_ = inp; return fp(); }
    const std::vector<double>& fpbp(const std::vector<double>& inp) { inp_ = inp; return fp(); }
    const std::vector<double>& fpbp2(const std::vector<double>& inp) { inp_ = inp; return fp(); }
    // many input rows [n][nInp] in one launch: out [n], dodi [n][nInp] (optional)
    void fp(const std::vector<double>& in, std::vector<double>& out, std::vector<double>* dodi = nullptr) const;
    // the evaluator of the current parameters and scaling
    ann_mlp* handle() const;

private:
    int nin_ = 0;
    std::vector<int> nodes_;
    Neuron neuron_ = Neuron::TANH;
    double c_ = 1.0, outw_ = 1.0, outb_ = 0.0;
    std::vector<double> params_, inpw_, inpb_, inp_, out_;
    struct Dev;
    mutable std::unique_ptr<Dev> dev_;
};

// gradient of a cost w.r.t. the parameters, for dc/do at the network's input
class Cost {
public:
    Cost() = default;
    explicit Cost(const ANN& nn) { resize(nn); }
    void resize(const ANN& nn) { grad_.assign(nn.size(), 0.0); }
    const std::vector<double>& grad() const { return grad_; }
    const std::vector<double>& grad(const ANN& nn, const std::vector<double>& dcdo);
    // many rows: sum over rows of dcdo[p] do_p/dparams
    const std::vector<double>& grad(const ANN& nn, const std::vector<double>& in, const std::vector<double>& dcdo);

private:
    std::vector<double> grad_;
};

// do/di at the network's input (the dodz matrices of the reference are internal here)
class DODZ {
public:
    DODZ() = default;
    explicit DODZ(const ANN& nn) { resize(nn); }
    void resize(const ANN& nn) { dodi_.assign(nn.nInp(), 0.0); }
    void grad(const ANN& nn);
    const std::vector<double>& dodi() const { return dodi_; }

private:
    std::vector<double> dodi_;
};

// d2o/dp di at the network's input: d2odpdi()[p][k]
class D2ODZDI {
public:
    D2ODZDI() = default;
    explicit D2ODZDI(const ANN& nn) { resize(nn); }
    void resize(const ANN& nn) { d2odpdi_.assign(nn.size(), std::vector<double>(nn.nInp(), 0.0)); }
    void grad(const ANN& nn);
    const std::vector<std::vector<double>>& d2odpdi() const { return d2odpdi_; }
    // many rows: sum over rows of dcdo[p] do_p/dparams + sum_k v[p][k] d2o_p/dparams di_k
    static void grad_rop(const ANN& nn, const std::vector<double>& in, const std::vector<double>& dcdo,
                         const std::vector<double>& v, std::vector<double>& grad);

private:
    std::vector<std::vector<double>> d2odpdi_;
};

}  // namespace NN
}  // namespace ann_gpu
