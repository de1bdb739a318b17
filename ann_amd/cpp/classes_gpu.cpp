// classes_gpu.cpp -- see classes_gpu.hpp.  Host code only: each evaluation
// copies its points to the device, calls the FP64 evaluator of libann_amd.so
// and copies the results back.
#include "classes_gpu.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "ann_gpu.h"

namespace ann_gpu {

namespace {

void check(int rc, const char* what) {
    if (rc != ANN_OK) {
        const std::string msg = std::string(what) + ": " + ann_last_error();
        if (rc == ANN_ERR_ARG || rc == ANN_ERR_UNSUPPORTED) throw std::invalid_argument(msg);
        throw std::runtime_error(msg);
    }
}

void hipcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// device staging of one evaluation (per host thread, grown on demand)
struct Buf {
    double* p = nullptr;
    size_t n = 0;
    double* get(size_t want) {
        if (want > n) {
            if (p) (void)hipFree(p);
            p = nullptr;
            n = 0;
            hipcheck(hipMalloc(&p, sizeof(double) * std::max<size_t>(want, 1)), "hipMalloc");
            n = want;
        }
        return p;
    }
    ~Buf() { if (p) (void)hipFree(p); }
};

struct Staging {
    Buf b[8];
};
thread_local Staging stage;

double* up(int slot, const double* h, size_t n) {
    double* d = stage.b[slot].get(n);
    if (n) hipcheck(hipMemcpy(d, h, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    return d;
}

void down(double* h, const double* d, size_t n) {
    if (n) hipcheck(hipMemcpy(h, d, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
}

}  // namespace

// ------------------------------------------------------------------ Cutoff

// Cutoff(Name, double) (src/nnp/cutoff.cpp:46-53)
Cutoff::Cutoff(Name name, double rc) : name_(name), rc_(rc) {
    if (name == Name::NONE) throw std::invalid_argument("Cutoff::Cutoff(Name,double): invalid cutoff name.");
    if (!(rc > 0)) throw std::invalid_argument("Cutoff::Cutoff(Name,double): invalid cutoff radius.");
}

void Cutoff::compute(const std::vector<double>& dr, std::vector<double>& v, std::vector<double>& g) const {
    const size_t n = dr.size();
    v.resize(n);
    g.resize(n);
    double* r = up(0, dr.data(), n);
    double* f = stage.b[1].get(n);
    double* d = stage.b[2].get(n);
    check(ann_cutoff_eval((int)name_, rc_, (int)n, r, f, d), "Cutoff::compute");
    down(v.data(), f, n);
    down(g.data(), d, n);
}

void Cutoff::compute(double dr, double& v, double& g) const {
    std::vector<double> r(1, dr), a, b;
    compute(r, a, b);
    v = a[0];
    g = b[0];
}

double Cutoff::cutf(double dr) const {
    double v, g;
    compute(dr, v, g);
    return v;
}

double Cutoff::cutg(double dr) const {
    double v, g;
    compute(dr, v, g);
    return g;
}

// ------------------------------------------------------------------ BasisR

BasisR::BasisR(double rc, Cutoff::Name cutname, int nf, Name name) : name_(name), cutoff_(cutname, rc) {
    if (nf < 0) throw std::invalid_argument("BasisR::BasisR(double,Cutoff::Name,int,Name): invalid size.");
    resize(nf);
}

void BasisR::resize(int n) {
    rs_.resize(n, 0.0);
    eta_.resize(n, 0.0);
    symm_.assign(n, 0.0);
}

void BasisR::eval(int n, const double* r, double* G, double* dG, const double* dEdG, double* amp) const {
    if (name_ == Name::NONE) throw std::invalid_argument("BasisR::symm(double): Invalid symmetry function.");
    const int nf = size();
    ann_radial_desc b{(int)cutoff_.name(), cutoff_.rc(), (int)name_, nf, rs_.data(), eta_.data()};
    double* dr = up(0, r, n);
    double* dGo = G ? stage.b[1].get((size_t)n * nf) : nullptr;
    double* ddo = dG ? stage.b[2].get((size_t)n * nf) : nullptr;
    double* de = dEdG ? up(3, dEdG, (size_t)n * nf) : nullptr;
    double* am = amp ? stage.b[4].get(n) : nullptr;
    check(ann_radial_eval(&b, n, dr, dGo, ddo, de, am), "BasisR");
    if (G) down(G, dGo, (size_t)n * nf);
    if (dG) down(dG, ddo, (size_t)n * nf);
    if (amp) down(amp, am, n);
}

double BasisR::symmf(double dr, double eta, double rs) const {
    BasisR one(cutoff_.rc(), cutoff_.name(), 1, name_);
    one.rs_[0] = rs;
    one.eta_[0] = eta;
    double G;
    one.eval(1, &dr, &G, nullptr, nullptr, nullptr);
    return G;
}

double BasisR::symmd(double dr, double eta, double rs) const {
    BasisR one(cutoff_.rc(), cutoff_.name(), 1, name_);
    one.rs_[0] = rs;
    one.eta_[0] = eta;
    double d;
    one.eval(1, &dr, nullptr, &d, nullptr, nullptr);
    return d;
}

void BasisR::symm(const std::vector<double>& dr) {
    symm_.resize(size());
    eval(1, dr.data(), symm_.data(), nullptr, nullptr, nullptr);
}

double BasisR::force(const std::vector<double>& dr, const double* dEdG) const {
    double amp;
    eval(1, dr.data(), nullptr, nullptr, dEdG, &amp);
    return amp;
}

void BasisR::symm(const std::vector<double>& r, std::vector<double>& G) const {
    G.resize(r.size() * size());
    eval((int)r.size(), r.data(), G.data(), nullptr, nullptr, nullptr);
}

void BasisR::force(const std::vector<double>& r, const std::vector<double>& dEdG, std::vector<double>& amp) const {
    if (dEdG.size() != r.size() * size()) throw std::invalid_argument("BasisR::force: dEdG size mismatch");
    amp.resize(r.size());
    eval((int)r.size(), r.data(), nullptr, nullptr, dEdG.data(), amp.data());
}

// ------------------------------------------------------------------ BasisA

BasisA::BasisA(double rc, Cutoff::Name cutname, int nf, Name name) : name_(name), cutoff_(cutname, rc) {
    if (nf < 0) throw std::invalid_argument("BasisA::BasisA(double,Cutoff::Name,int,Name): invalid size.");
    resize(nf);
}

void BasisA::resize(int n) {
    eta_.resize(n, 0.0);
    zeta_.resize(n, 0.0);
    lambda_.resize(n, 0);
    symm_.assign(n, 0.0);
}

void BasisA::eval(int n, const double* r, const double* cos, double* G, double* dG, const double* dEdG, double* phi,
                  double* eta) const {
    if (name_ == Name::NONE) throw std::invalid_argument("BasisA::symm(double): Invalid symmetry function.");
    const int nf = size();
    ann_angular_desc b{(int)cutoff_.name(), cutoff_.rc(), (int)name_, nf, eta_.data(), zeta_.data(), lambda_.data()};
    double* dr = up(0, r, 2 * (size_t)n);
    double* dc = up(1, cos, n);
    double* dGo = G ? stage.b[2].get((size_t)n * nf) : nullptr;
    double* ddo = dG ? stage.b[3].get(3 * (size_t)n * nf) : nullptr;
    double* de = dEdG ? up(4, dEdG, (size_t)n * nf) : nullptr;
    double* dp = phi ? stage.b[5].get(n) : nullptr;
    double* dn = eta ? stage.b[6].get(2 * (size_t)n) : nullptr;
    check(ann_angular_eval(&b, n, dr, dc, dGo, ddo, de, dp, dn), "BasisA");
    if (G) down(G, dGo, (size_t)n * nf);
    if (dG) down(dG, ddo, 3 * (size_t)n * nf);
    if (phi) down(phi, dp, n);
    if (eta) down(eta, dn, 2 * (size_t)n);
}

double BasisA::symmf(double cos, const double dr[2], double eta, double zeta, int lambda) const {
    BasisA one(cutoff_.rc(), cutoff_.name(), 1, name_);
    one.eta_[0] = eta;
    one.zeta_[0] = zeta;
    one.lambda_[0] = lambda;
    double G;
    one.eval(1, dr, &cos, &G, nullptr, nullptr, nullptr, nullptr);
    return G;
}

void BasisA::symmd(double& fphi, double* feta, double cos, const double dr[2], double eta, double zeta,
                   int lambda) const {
    BasisA one(cutoff_.rc(), cutoff_.name(), 1, name_);
    one.eta_[0] = eta;
    one.zeta_[0] = zeta;
    one.lambda_[0] = lambda;
    double d[3];
    one.eval(1, dr, &cos, nullptr, d, nullptr, nullptr, nullptr);
    fphi = d[0];
    feta[0] = d[1];
    feta[1] = d[2];
}

void BasisA::symm(const std::vector<double>& dr, const std::vector<double>& cos) {
    symm_.resize(size());
    eval(1, dr.data(), cos.data(), symm_.data(), nullptr, nullptr, nullptr, nullptr);
}

void BasisA::force(const std::vector<double>& dr, const std::vector<double>& cos, double& phi, double* eta,
                   const double* dEdG) const {
    eval(1, dr.data(), cos.data(), nullptr, nullptr, dEdG, &phi, eta);
}

void BasisA::symm(const std::vector<double>& r, const std::vector<double>& cos, std::vector<double>& G) const {
    if (r.size() != 2 * cos.size()) throw std::invalid_argument("BasisA::symm: r must hold two legs per cosine");
    G.resize(cos.size() * size());
    eval((int)cos.size(), r.data(), cos.data(), G.data(), nullptr, nullptr, nullptr, nullptr);
}

void BasisA::force(const std::vector<double>& r, const std::vector<double>& cos, const std::vector<double>& dEdG,
                   std::vector<double>& phi, std::vector<double>& eta) const {
    if (r.size() != 2 * cos.size()) throw std::invalid_argument("BasisA::force: r must hold two legs per cosine");
    if (dEdG.size() != cos.size() * size()) throw std::invalid_argument("BasisA::force: dEdG size mismatch");
    phi.resize(cos.size());
    eta.resize(2 * cos.size());
    eval((int)cos.size(), r.data(), cos.data(), nullptr, nullptr, dEdG.data(), phi.data(), eta.data());
}

// ------------------------------------------------------------------ NN

namespace NN {

// the evaluator and the host state it was built from
struct ANN::Dev {
    ann_mlp* h = nullptr;
    int nin = -1, neuron = 0;
    double c = 0;
    std::vector<int> nodes;
    std::vector<double> params, inpw, inpb;
    double outw = 0, outb = 0;
    ~Dev() { ann_mlp_destroy(h); }
};

ANN::ANN() = default;
ANN::~ANN() = default;
ANN::ANN(const ANN& o)
    : nin_(o.nin_), nodes_(o.nodes_), neuron_(o.neuron_), c_(o.c_), outw_(o.outw_), outb_(o.outb_),
      params_(o.params_), inpw_(o.inpw_), inpb_(o.inpb_), inp_(o.inp_), out_(o.out_) {}
ANN& ANN::operator=(const ANN& o) {
    if (this != &o) {
        nin_ = o.nin_; nodes_ = o.nodes_; neuron_ = o.neuron_; c_ = o.c_; outw_ = o.outw_; outb_ = o.outb_;
        params_ = o.params_; inpw_ = o.inpw_; inpb_ = o.inpb_; inp_ = o.inp_; out_ = o.out_;
        dev_.reset();
    }
    return *this;
}

void ANN::resize(int nInput, const std::vector<int>& nNodes, int nOutput) {
    if (nInput <= 0) throw std::invalid_argument("ANN::resize(const ANNP&,int,const std::vector<int>&,int): Invalid input size.");
    if (nOutput != 1) throw std::invalid_argument("ANN::resize: only single-output networks are supported");
    for (int n : nNodes)
        if (n <= 0) throw std::invalid_argument("ANN::resize(const ANNP&,const std::vector<int>&): Invalid layer size.");
    nin_ = nInput;
    nodes_ = nNodes;
    nodes_.push_back(nOutput);
    int cnt = 0, prev = nInput;
    for (int n : nodes_) { cnt += n + n * prev; prev = n; }
    params_.assign(cnt, 0.0);
    inpw_.assign(nInput, 1.0);
    inpb_.assign(nInput, 0.0);
    inp_.assign(nInput, 0.0);
    out_.assign(nOutput, 0.0);
    outw_ = 1.0;
    outb_ = 0.0;
    dev_.reset();
}

int ANN::nBias() const {
    int n = 0;
    for (int x : nodes_) n += x;
    return n;
}

ann_mlp* ANN::handle() const {
    if (nodes_.empty()) throw std::invalid_argument("ANN: network not initialised (resize first)");
    if ((int)inpw_.size() != nin_ || (int)inpb_.size() != nin_)
        throw std::invalid_argument("ANN: input scaling size mismatch");
    int cnt = 0, prev = nin_;
    for (int n : nodes_) { cnt += n + n * prev; prev = n; }
    if ((int)params_.size() != cnt) throw std::invalid_argument("ANN: parameter count mismatch");
    Dev* d = dev_.get();
    const bool shape = !d || d->nin != nin_ || d->nodes != nodes_ || d->neuron != (int)neuron_ || d->c != c_;
    if (shape) {
        dev_.reset(new Dev);
        d = dev_.get();
        ann_mlp_desc desc{nin_, (int)nodes_.size(), nodes_.data(), (int)neuron_, c_, params_.data(), inpw_.data(),
                          inpb_.data(), outw_, outb_};
        check(ann_mlp_create(&desc, &d->h), "ANN");
        d->nin = nin_;
        d->nodes = nodes_;
        d->neuron = (int)neuron_;
        d->c = c_;
        d->params = params_;
        d->inpw = inpw_;
        d->inpb = inpb_;
        d->outw = outw_;
        d->outb = outb_;
        return d->h;
    }
    if (d->params != params_) {
        check(ann_mlp_set_params(d->h, params_.data(), (int)params_.size()), "ANN");
        d->params = params_;
    }
    if (d->inpw != inpw_ || d->inpb != inpb_ || d->outw != outw_ || d->outb != outb_) {
        check(ann_mlp_set_scaling(d->h, inpw_.data(), inpb_.data(), outw_, outb_), "ANN");
        d->inpw = inpw_;
        d->inpb = inpb_;
        d->outw = outw_;
        d->outb = outb_;
    }
    return d->h;
}

void ANN::fp(const std::vector<double>& in, std::vector<double>& out, std::vector<double>* dodi) const {
    if (nin_ <= 0 || in.size() % nin_) throw std::invalid_argument("ANN::fp: input rows of nInp values expected");
    const int n = (int)(in.size() / nin_);
    ann_mlp* h = handle();
    double* di = up(0, in.data(), in.size());
    double* dout = stage.b[1].get(n);
    double* dd = dodi ? stage.b[2].get(in.size()) : nullptr;
    check(ann_mlp_eval(h, n, di, dout, dd), "ANN::fp");
    out.resize(n);
    down(out.data(), dout, n);
    if (dodi) {
        dodi->resize(in.size());
        down(dodi->data(), dd, in.size());
    }
}

const std::vector<double>& ANN::fp() {
    if ((int)inp_.size() != nin_) throw std::invalid_argument("ANN::fp: input size mismatch");
    std::vector<double> o;
    fp(inp_, o, nullptr);
    out_ = o;
    return out_;
}

const std::vector<double>& Cost::grad(const ANN& nn, const std::vector<double>& dcdo) {
    return grad(nn, nn.inp(), dcdo);
}

const std::vector<double>& Cost::grad(const ANN& nn, const std::vector<double>& in, const std::vector<double>& dcdo) {
    if (nn.nInp() <= 0 || in.size() % nn.nInp()) throw std::invalid_argument("Cost::grad: input rows expected");
    const size_t n = in.size() / nn.nInp();
    if (dcdo.size() != n) throw std::invalid_argument("Cost::grad: one dc/do per input row expected");
    ann_mlp* h = nn.handle();
    double* di = up(0, in.data(), in.size());
    double* dc = up(1, dcdo.data(), n);
    double* dg = stage.b[2].get(nn.size());
    check(ann_mlp_grad(h, (int)n, di, dc, nullptr, 0, dg), "Cost::grad");
    grad_.resize(nn.size());
    down(grad_.data(), dg, nn.size());
    return grad_;
}

void DODZ::grad(const ANN& nn) {
    std::vector<double> out;
    nn.fp(nn.inp(), out, &dodi_);
}

void D2ODZDI::grad(const ANN& nn) {
    const int nin = nn.nInp(), np = nn.size();
    std::vector<double> in((size_t)nin * nin), v((size_t)nin * nin, 0.0);
    for (int k = 0; k < nin; ++k) {
        std::copy(nn.inp().begin(), nn.inp().end(), in.begin() + (size_t)k * nin);
        v[(size_t)k * nin + k] = 1.0;
    }
    ann_mlp* h = nn.handle();
    double* di = up(0, in.data(), in.size());
    double* dv = up(1, v.data(), v.size());
    double* dg = stage.b[2].get((size_t)nin * np);
    check(ann_mlp_grad(h, nin, di, nullptr, dv, 1, dg), "D2ODZDI::grad");
    std::vector<double> g((size_t)nin * np);
    down(g.data(), dg, g.size());
    d2odpdi_.assign(np, std::vector<double>(nin, 0.0));
    for (int k = 0; k < nin; ++k)
        for (int p = 0; p < np; ++p) d2odpdi_[p][k] = g[(size_t)k * np + p];
}

void D2ODZDI::grad_rop(const ANN& nn, const std::vector<double>& in, const std::vector<double>& dcdo,
                       const std::vector<double>& v, std::vector<double>& grad) {
    if (nn.nInp() <= 0 || in.size() % nn.nInp()) throw std::invalid_argument("D2ODZDI::grad_rop: input rows expected");
    const size_t n = in.size() / nn.nInp();
    if (v.size() != in.size()) throw std::invalid_argument("D2ODZDI::grad_rop: one direction per input row expected");
    if (!dcdo.empty() && dcdo.size() != n) throw std::invalid_argument("D2ODZDI::grad_rop: dcdo size mismatch");
    ann_mlp* h = nn.handle();
    double* di = up(0, in.data(), in.size());
    double* dv = up(1, v.data(), v.size());
    double* dc = dcdo.empty() ? nullptr : up(2, dcdo.data(), n);
    double* dg = stage.b[3].get(nn.size());
    check(ann_mlp_grad(h, (int)n, di, dc, dv, 0, dg), "D2ODZDI::grad_rop");
    grad.resize(nn.size());
    down(grad.data(), dg, nn.size());
}

}  // namespace NN
}  // namespace ann_gpu
