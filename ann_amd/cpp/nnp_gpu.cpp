// nnp_gpu.cpp -- see nnp_gpu.hpp.  Host code only: every calculation goes
// through the C ABI of libann_amd.so.
#include "nnp_gpu.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "ann_gpu.h"

namespace ann_gpu {

namespace {

void check(int rc, const char* what) {
    if (rc != ANN_OK) throw std::runtime_error(std::string(what) + ": " + ann_last_error());
}
void hipcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

const char* kCut[] = {"NONE", "STEP", "COS", "TANH", "POLY3"};
const char* kRad[] = {"NONE", "GAUSSIAN", "SECH", "LOGISTIC", "TANH", "LOGCOSH", "LOGCOSH2"};
const char* kAng[] = {"NONE", "GAUSS", "GAUSS2", "SECH", "STUDENT3", "STUDENT4", "STUDENT5"};
// NN::Neuron names and numbers (src/ml/nn.hpp:83-107, src/ml/nn.cpp:66-140)
const std::map<std::string, int> kNeuron = {
    {"LINEAR", 1}, {"SIGMOID", 2}, {"TANH", 3},   {"ISRU", 4},    {"ARCTAN", 5},  {"RELU", 6},
    {"ELU", 7},    {"TANHRE", 8},  {"SQRE", 9},   {"SWISH", 10},  {"GELU", 11},   {"MISH", 12},
    {"PFLU", 13},  {"SERF", 14},   {"LOGISH", 15}, {"SOFTPLUS", 16}, {"SQPLUS", 17}, {"ATISH", 18}};

int lookup(const char* const* tab, int n, const std::string& s, const char* what) {
    for (int i = 1; i < n; ++i)
        if (s == tab[i]) return i;
    throw std::invalid_argument(std::string("invalid ") + what + ": " + s);
}

int pair_index(int j, int k) { return j > k ? j * (j + 1) / 2 + k : k * (k + 1) / 2 + j; }

std::vector<std::string> tokens(const std::string& line) {
    std::istringstream is(line);
    std::vector<std::string> t;
    std::string w;
    while (is >> w) t.push_back(w);
    return t;
}

std::string upper(std::string s) {
    for (auto& c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}

// minimal potential for a neighbour-only context (one radial function per pair)
ann_ctx* neighbor_context(int S) {
    static std::map<int, ann_ctx*> cache;
    auto it = cache.find(S);
    if (it != cache.end()) return it->second;
    std::vector<int> rn(S * S, 1), rnm(S * S, 1), roff(S * S), an(S * (S + 1) / 2 * S, 0), anm(an.size(), 1),
        aoff(an.size(), 0), nl(S, 1), noff(S), nodes(S, 1), neu(S, 1), poff(S), ioff(S);
    std::vector<double> rs(S * S, 1.0), eta(S * S, 1.0), ae(1, 1.0), az(1, 1.0), c(S, 1.0), params, iw, ib,
        ow(S, 1.0), ob(S, 0.0), ev(S, 0.0);
    std::vector<int> al(1, 1);
    for (int t = 0; t < S * S; ++t) roff[t] = t;
    for (int a = 0; a < S; ++a) {
        noff[a] = a;
        poff[a] = (int)params.size();
        params.push_back(0.0);                       // bias
        for (int m = 0; m < S; ++m) params.push_back(0.0);
        ioff[a] = (int)iw.size();
        for (int m = 0; m < S; ++m) { iw.push_back(1.0); ib.push_back(0.0); }
    }
    ann_model_desc d{S, 1.0, 1, rnm.data(), rn.data(), roff.data(), rs.data(), eta.data(), anm.data(), an.data(),
                     aoff.data(), ae.data(), az.data(), al.data(), nl.data(), noff.data(), nodes.data(), neu.data(),
                     c.data(), poff.data(), params.data(), ioff.data(), iw.data(), ib.data(), ow.data(), ob.data(),
                     ev.data()};
    ann_ctx* ctx = nullptr;
    check(ann_create(&d, &ctx), "neighbour context");
    cache[S] = ctx;
    return ctx;
}

// device copy of one structure bound to a context
struct DeviceStruc {
    int cap = 0;
    int* off = nullptr;
    double* pos = nullptr;
    double* cell = nullptr;
    int* species = nullptr;
    int hoff[2] = {0, 0};
    void free_all() {
        hipFree(off); hipFree(pos); hipFree(cell); hipFree(species);
        off = nullptr; pos = cell = nullptr; species = nullptr; cap = 0;
    }
    ~DeviceStruc() { free_all(); }
    void upload(ann_ctx* ctx, const Structure& s, double rc) {
        set(ctx, s);
        check(ann_neighbors(ctx, rc), "ann_neighbors");
    }
    // positions only (same atoms, same list): the Verlet-list reuse between builds
    void positions(const Structure& s) {
        const int N = s.nAtoms();
        std::vector<double> p(3 * N);
        for (int i = 0; i < N; ++i)
            for (int x = 0; x < 3; ++x) p[3 * i + x] = s.posn(i)[x];
        hipcheck(hipMemcpy(pos, p.data(), 3 * sizeof(double) * N, hipMemcpyHostToDevice), "hipMemcpy");
    }
    void set(ann_ctx* ctx, const Structure& s) {
        const int N = s.nAtoms();
        if (N < 1) throw std::invalid_argument("structure has no atoms");
        if (N > cap) {
            free_all();
            hipcheck(hipMalloc(&off, 2 * sizeof(int)), "hipMalloc");
            hipcheck(hipMalloc(&pos, 3 * sizeof(double) * N), "hipMalloc");
            hipcheck(hipMalloc(&cell, 9 * sizeof(double)), "hipMalloc");
            hipcheck(hipMalloc(&species, sizeof(int) * N), "hipMalloc");
            cap = N;
        }
        std::vector<double> p(3 * N);
        std::vector<int> sp(N);
        for (int i = 0; i < N; ++i) {
            for (int x = 0; x < 3; ++x) p[3 * i + x] = s.posn(i)[x];
            sp[i] = s.type(i);
        }
        hoff[0] = 0;
        hoff[1] = N;
        hipcheck(hipMemcpy(off, hoff, 2 * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
        hipcheck(hipMemcpy(pos, p.data(), 3 * sizeof(double) * N, hipMemcpyHostToDevice), "hipMemcpy");
        hipcheck(hipMemcpy(cell, s.R().data(), 9 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        hipcheck(hipMemcpy(species, sp.data(), sizeof(int) * N, hipMemcpyHostToDevice), "hipMemcpy");
        ann_batch b{1, hoff, off, pos, cell, species};
        check(ann_set_batch(ctx, &b), "ann_set_batch");
    }
};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    T* get(size_t want) {
        if (want > n) {
            hipFree(p);
            p = nullptr;
            hipcheck(hipMalloc(&p, sizeof(T) * std::max<size_t>(want, 1)), "hipMalloc");
            n = want;
        }
        return p;
    }
    ~DevBuf() { hipFree(p); }
};

int form_from_env() {
    const char* f = std::getenv("ANN_FORM");
    if (f && std::string(f) == "plain") return ANN_FORM_PLAIN;
    if (f && std::string(f) == "triplet") return ANN_FORM_TRIPLET;
    return ANN_FORM_FAST;
}

}  // namespace

// ------------------------------------------------------------------ Structure

void Structure::resize(int n) {
    posn_.assign(n, Vec3{0, 0, 0});
    force_.assign(n, Vec3{0, 0, 0});
    type_.assign(n, 0);
    name_.assign(n, "");
    symm_.assign(n, {});
}

// ------------------------------------------------------------------ NeighborList

namespace {
// the GPU's canonical list of a structure at rc: CSR offsets, partners, images, reverse slots
void gpu_list(const Structure& struc, double rc, std::vector<int>& off, std::vector<int>& idx, std::vector<int>& img,
              std::vector<int>& rev) {
    if (!(rc > 0)) throw std::invalid_argument("NeighborList: invalid cutoff");
    int S = 1;
    for (int i = 0; i < struc.nAtoms(); ++i) S = std::max(S, struc.type(i) + 1);
    ann_ctx* ctx = neighbor_context(S);
    static DeviceStruc ds;
    ds.upload(ctx, struc, rc);
    int64_t P = 0;
    int maxn = 0;
    check(ann_nlist_size(ctx, &P, &maxn), "ann_nlist_size");
    const int N = struc.nAtoms();
    DevBuf<int> doff, didx, dimg, drev;
    check(ann_nlist_copy(ctx, doff.get(N + 1), didx.get(P), dimg.get(3 * P), drev.get(P), nullptr), "ann_nlist_copy");
    off.assign(N + 1, 0);
    idx.assign(P, 0);
    img.assign(3 * P, 0);
    rev.assign(P, 0);
    hipcheck(hipDeviceSynchronize(), "sync");
    hipcheck(hipMemcpy(off.data(), doff.p, sizeof(int) * (N + 1), hipMemcpyDeviceToHost), "hipMemcpy");
    if (P > 0) {
        hipcheck(hipMemcpy(idx.data(), didx.p, sizeof(int) * P, hipMemcpyDeviceToHost), "hipMemcpy");
        hipcheck(hipMemcpy(img.data(), dimg.p, sizeof(int) * 3 * P, hipMemcpyDeviceToHost), "hipMemcpy");
        hipcheck(hipMemcpy(rev.data(), drev.p, sizeof(int) * P, hipMemcpyDeviceToHost), "hipMemcpy");
    }
}
}  // namespace

void NeighborList::build(const Structure& struc, double rc) {
    std::vector<int> idx, img, rev;
    gpu_list(struc, rc, off_, idx, img, rev);
    const int N = struc.nAtoms();
    const int64_t P = off_[N];
    list_.assign(P, Neighbor{});
    const auto& R = struc.R();
    for (int i = 0; i < N; ++i) {
        for (int q = off_[i]; q < off_[i + 1]; ++q) {
            Neighbor& nb = list_[q];
            nb.index_ = idx[q];
            nb.type_ = struc.type(idx[q]);
            nb.cell_ = {img[3 * q], img[3 * q + 1], img[3 * q + 2]};
            nb.rev_ = rev[q] - off_[idx[q]];
            double d2 = 0;
            for (int x = 0; x < 3; ++x) {
                nb.r_[x] = (struc.posn(i)[x] - struc.posn(idx[q])[x]) -
                           ((R[3 * x] * nb.cell_[0] + R[3 * x + 1] * nb.cell_[1]) + R[3 * x + 2] * nb.cell_[2]);
                d2 += nb.r_[x] * nb.r_[x];
            }
            nb.dr_ = std::sqrt(d2);
        }
        // minimum image: shortest entry of each partner (first on ties)
        for (int q = off_[i]; q < off_[i + 1];) {
            int e = q, best = q;
            while (e < off_[i + 1] && list_[e].index_ == list_[q].index_) {
                if (list_[e].dr_ < list_[best].dr_) best = e;
                ++e;
            }
            list_[best].min_ = true;
            q = e;
        }
    }
    rc_ = rc;
}

// ------------------------------------------------------------------ verlet::List

namespace verlet {

void List::build(const Structure& struc) {
    static uint64_t builds = 0;
    std::vector<int> idx, img, rev;
    gpu_list(struc, rc_, off_, idx, img, rev);
    list_.assign(idx.size(), Neighbor{});
    for (size_t q = 0; q < idx.size(); ++q) {
        list_[q].index_ = idx[q];
        list_[q].cell_ = {img[3 * q], img[3 * q + 1], img[3 * q + 2]};
    }
    id_ = ++builds;
}

void List::flat(std::vector<int>& idx, std::vector<int>& img) const {
    idx.resize(list_.size());
    img.resize(3 * list_.size());
    for (size_t q = 0; q < list_.size(); ++q) {
        idx[q] = list_[q].index_;
        for (int x = 0; x < 3; ++x) img[3 * q + x] = list_[q].cell_[x];
    }
}

}  // namespace verlet

// ------------------------------------------------------------------ NNP tables

struct NNP::Tables {
    int cutoff = 0;
    std::vector<int> rad_name, rad_n, rad_off, ang_name, ang_n, ang_off, ang_lambda;
    std::vector<double> rad_rs, rad_eta, ang_eta, ang_zeta;
    std::vector<int> nn_nlayer, nn_nodes_off, nn_nodes, nn_neuron, nn_par_off, nn_inp_off, nin;
    std::vector<double> nn_c, params, inpw, inpb, outw, outb, evac;
    ann_model_desc desc(int S, double rc) {
        return ann_model_desc{S, rc, cutoff, rad_name.data(), rad_n.data(), rad_off.data(), rad_rs.data(),
                              rad_eta.data(), ang_name.data(), ang_n.data(), ang_off.data(), ang_eta.data(),
                              ang_zeta.data(), ang_lambda.data(), nn_nlayer.data(), nn_nodes_off.data(),
                              nn_nodes.data(), nn_neuron.data(), nn_c.data(), nn_par_off.data(), params.data(),
                              nn_inp_off.data(), inpw.data(), inpb.data(), outw.data(), outb.data(), evac.data()};
    }
};

struct NNP::Device {
    DeviceStruc ds;
    DevBuf<float> dE, Ea, P;
    DevBuf<double> G64, F, E;
    uint64_t vlist_id = 0;        // the Verlet list the context's neighbour list was taken from (0: none)
    int natoms = 0;
};

NNP::NNP() : t_(new Tables), dev_(new Device) {}
NNP::~NNP() {
    dev_.reset();
    if (ctx_) ann_destroy(ctx_);
}

int NNP::index(const std::string& name) const {
    for (int i = 0; i < ntypes(); ++i)
        if (types_[i].name == name) return i;
    throw std::invalid_argument("NNP::index: unknown species " + name);
}

int NNP::nInput(int i) const { return t_->nin[i]; }
int NNP::size() const { return (int)t_->params.size(); }
void NNP::params(std::vector<double>& p) const { p = t_->params; }
void NNP::set_params(const std::vector<double>& p) {
    if (p.size() != t_->params.size()) throw std::invalid_argument("NNP::set_params: size mismatch");
    t_->params = p;
    if (ctx_) check(ann_set_params(ctx_, p.data(), (int)p.size()), "ann_set_params");
}

ann_ctx* NNP::ctx() {
    if (!ctx_) {
        ann_model_desc d = t_->desc(ntypes(), rc_);
        check(ann_create(&d, &ctx_), "ann_create");
    }
    return ctx_;
}

void NNP::upload(const Structure& struc) {
    for (int i = 0; i < struc.nAtoms(); ++i)
        if (struc.type(i) < 0 || struc.type(i) >= ntypes()) throw std::invalid_argument("structure species out of range");
    dev_->ds.upload(ctx(), struc, rc_);
    dev_->vlist_id = 0;
    dev_->natoms = struc.nAtoms();
}

// the context's list from a Verlet list: taken over (ann_set_neighbors) once per
// build of the list; later calls with the same build move the positions only
// and refresh the pair geometry from the list's image vectors
void NNP::upload(const Structure& struc, const verlet::List& vlist) {
    if (vlist.nAtoms() != struc.nAtoms() || vlist.id() == 0)
        throw std::invalid_argument("verlet::List: build the list for this structure first");
    if (dev_->vlist_id == vlist.id() && dev_->natoms == struc.nAtoms()) {
        dev_->ds.positions(struc);
        check(ann_refresh_geometry(ctx()), "ann_refresh_geometry");
        return;
    }
    for (int i = 0; i < struc.nAtoms(); ++i)
        if (struc.type(i) < 0 || struc.type(i) >= ntypes()) throw std::invalid_argument("structure species out of range");
    dev_->ds.set(ctx(), struc);
    std::vector<int> idx, img;
    vlist.flat(idx, img);
    check(ann_set_neighbors(ctx(), vlist.off().data(), idx.data(), img.data(), vlist.rc()), "ann_set_neighbors");
    dev_->vlist_id = vlist.id();
    dev_->natoms = struc.nAtoms();
}

// ------------------------------------------------------------------ file format

void NNP::read(const char* file, NNP& nnp) {
    std::ifstream in(file);
    if (!in) throw std::runtime_error(std::string("NNP::read(const char*,NNP&): Could not open nnp file: \"") + file + "\"");
    std::vector<std::vector<std::string>> L;
    std::string line;
    while (std::getline(in, line)) {
        auto t = tokens(line);
        if (!t.empty()) L.push_back(t);
    }
    size_t q = 1;  // skip header
    auto next = [&]() -> const std::vector<std::string>& {
        if (q >= L.size()) throw std::runtime_error("NNP::read: unexpected end of file");
        return L[q++];
    };
    const int S = std::atoi(next().at(1).c_str());
    if (S <= 0) throw std::invalid_argument("NNP::read(FILE*,NNP&): invalid number of types.");
    nnp.types_.assign(S, TypeInfo{});
    for (int a = 0; a < S; ++a) {
        const auto& t = next();
        for (size_t k = 0; k + 1 < t.size(); k += 2) {
            const std::string tag = upper(t[k]);
            if (tag == "NAME") nnp.types_[a].name = t[k + 1];
            else if (tag == "MASS") nnp.types_[a].mass = std::atof(t[k + 1].c_str());
            else if (tag == "ENERGY") nnp.types_[a].energy = std::atof(t[k + 1].c_str());
        }
    }
    nnp.rc_ = std::atof(next().at(1).c_str());
    if (nnp.rc_ <= 0) throw std::invalid_argument("NNP::read(FILE*,NNP&): invalid cutoff.");
    Tables& T = *nnp.t_;
    T = Tables{};
    const int NP = S * (S + 1) / 2;
    T.rad_name.assign(S * S, 0); T.rad_n.assign(S * S, 0); T.rad_off.assign(S * S, 0);
    T.ang_name.assign(S * NP, 0); T.ang_n.assign(S * NP, 0); T.ang_off.assign(S * NP, 0);
    std::vector<std::vector<double>> rrs(S * S), reta(S * S), aeta(S * NP), azeta(S * NP);
    std::vector<std::vector<int>> alam(S * NP);
    for (int c = 0; c < S; ++c) {
        const int a = nnp.index(next().at(1));
        for (int c2 = 0; c2 < S; ++c2) {
            const int b = nnp.index(next().at(1));
            const auto& h = next();
            T.cutoff = lookup(kCut, 5, h.at(1), "cutoff");
            const int t = a * S + b, n = std::atoi(h.at(4).c_str());
            T.rad_name[t] = lookup(kRad, 7, h.at(3), "radial function");
            T.rad_n[t] = n;
            for (int m = 0; m < n; ++m) {
                const auto& v = next();
                rrs[t].push_back(std::atof(v.at(0).c_str()));
                reta[t].push_back(std::atof(v.at(1).c_str()));
            }
        }
        for (int c2 = 0; c2 < NP; ++c2) {
            const auto& pj = next();
            const int p = pair_index(nnp.index(pj.at(1)), nnp.index(pj.at(2)));
            const auto& h = next();
            const int t = a * NP + p, n = std::atoi(h.at(4).c_str());
            T.ang_name[t] = lookup(kAng, 7, h.at(3), "angular function");
            T.ang_n[t] = n;
            for (int m = 0; m < n; ++m) {
                const auto& v = next();
                aeta[t].push_back(std::atof(v.at(0).c_str()));
                azeta[t].push_back(std::atof(v.at(1).c_str()));
                alam[t].push_back(std::atoi(v.at(2).c_str()));
            }
        }
    }
    for (int t = 0; t < S * S; ++t) {
        T.rad_off[t] = (int)T.rad_rs.size();
        T.rad_rs.insert(T.rad_rs.end(), rrs[t].begin(), rrs[t].end());
        T.rad_eta.insert(T.rad_eta.end(), reta[t].begin(), reta[t].end());
    }
    for (int t = 0; t < S * NP; ++t) {
        T.ang_off[t] = (int)T.ang_eta.size();
        T.ang_eta.insert(T.ang_eta.end(), aeta[t].begin(), aeta[t].end());
        T.ang_zeta.insert(T.ang_zeta.end(), azeta[t].begin(), azeta[t].end());
        T.ang_lambda.insert(T.ang_lambda.end(), alam[t].begin(), alam[t].end());
    }
    if (T.rad_rs.empty()) { T.rad_rs.push_back(0); T.rad_eta.push_back(0); }
    if (T.ang_eta.empty()) { T.ang_eta.push_back(0); T.ang_zeta.push_back(0); T.ang_lambda.push_back(0); }
    T.nin.assign(S, 0);
    for (int a = 0; a < S; ++a) {
        for (int b = 0; b < S; ++b) T.nin[a] += T.rad_n[a * S + b];
        for (int p = 0; p < NP; ++p) T.nin[a] += T.ang_n[a * NP + p];
    }
    // networks (ANN::read, src/ml/nn.cpp:1508-1580)
    T.nn_nlayer.assign(S, 0); T.nn_nodes_off.assign(S, 0); T.nn_neuron.assign(S, 0); T.nn_c.assign(S, 1.0);
    T.nn_par_off.assign(S, 0); T.nn_inp_off.assign(S, 0); T.outw.assign(S, 1.0); T.outb.assign(S, 0.0);
    std::vector<std::vector<int>> nodes(S);
    std::vector<std::vector<double>> par(S), iw(S), ib(S);
    for (int c = 0; c < S; ++c) {
        const int a = nnp.index(next().at(1));
        const auto& dims = next();
        const int nin = std::atoi(dims.at(1).c_str());
        if (nin != T.nin[a]) throw std::invalid_argument("NNP::read: network input size does not match the basis");
        for (size_t k = 2; k < dims.size(); ++k) nodes[a].push_back(std::atoi(dims[k].c_str()));
        const auto& neu = next();
        auto it = kNeuron.find(upper(neu.at(1)));
        if (it == kNeuron.end()) throw std::invalid_argument("ANN::read(FILE*,ANN&): Invalid neuron.");
        T.nn_neuron[a] = it->second;
        T.nn_c[a] = std::atof(neu.at(2).c_str());
        auto vec = [&](size_t expect) {
            const auto& v = next();
            if (v.size() != expect + 1) throw std::runtime_error("NNP::read: wrong number of values on line " + v.at(0));
            std::vector<double> o;
            for (size_t k = 1; k < v.size(); ++k) o.push_back(std::atof(v[k].c_str()));
            return o;
        };
        iw[a] = vec(nin);
        T.outw[a] = vec(1)[0];
        ib[a] = vec(nin);
        T.outb[a] = vec(1)[0];
        std::vector<std::vector<double>> bs, ws;
        for (size_t l = 0; l < nodes[a].size(); ++l) bs.push_back(vec(nodes[a][l]));
        int prev = nin;
        for (size_t l = 0; l < nodes[a].size(); ++l) {
            const int no = nodes[a][l];
            auto w = vec((size_t)no * prev);       // row-major (output, input)
            std::vector<double> cm((size_t)no * prev);
            for (int n = 0; n < no; ++n)
                for (int m = 0; m < prev; ++m) cm[(size_t)m * no + n] = w[(size_t)n * prev + m];
            ws.push_back(cm);
            prev = no;
        }
        for (auto& b : bs) par[a].insert(par[a].end(), b.begin(), b.end());
        for (auto& w : ws) par[a].insert(par[a].end(), w.begin(), w.end());
    }
    for (int a = 0; a < S; ++a) {
        T.nn_nlayer[a] = (int)nodes[a].size();
        T.nn_nodes_off[a] = (int)T.nn_nodes.size();
        T.nn_nodes.insert(T.nn_nodes.end(), nodes[a].begin(), nodes[a].end());
        T.nn_par_off[a] = (int)T.params.size();
        T.params.insert(T.params.end(), par[a].begin(), par[a].end());
        T.nn_inp_off[a] = (int)T.inpw.size();
        T.inpw.insert(T.inpw.end(), iw[a].begin(), iw[a].end());
        T.inpb.insert(T.inpb.end(), ib[a].begin(), ib[a].end());
        T.evac.push_back(nnp.types_[a].energy);
    }
    if (nnp.ctx_) { ann_destroy(nnp.ctx_); nnp.ctx_ = nullptr; }
}

// NNP::read_basis (src/nnp/nnp.cpp:96-156): the lines after the "basis NAME"
// line of an nnp file -- "rc", "ntypes", "basis NAME", then per partner type
// "basis_radial J" + BasisR block and per pair j <= k "basis_angular J K" +
// BasisA block.  The cutoff and type count must match the potential.
void NNP::read_basis(const char* file, NNP& nnp, const char* name) {
    std::ifstream in(file);
    if (!in)
        throw std::runtime_error(std::string("NNP::read_basis(const char*,NNP&,const char*): Could not open nnp file: \"") +
                                 file + "\"");
    std::vector<std::vector<std::string>> L;
    std::string line;
    while (std::getline(in, line)) {
        auto t = tokens(line);
        if (!t.empty()) L.push_back(t);
    }
    size_t q = 0;
    auto next = [&]() -> const std::vector<std::string>& {
        if (q >= L.size()) throw std::runtime_error("NNP::read_basis: unexpected end of file");
        return L[q++];
    };
    const int II = nnp.index(name);
    const double rc = std::atof(next().at(1).c_str());
    if (rc != nnp.rc()) throw std::invalid_argument("NNP::read_basis(FILE*,NNP&,const char*): invalid cutoff.");
    const int S = std::atoi(next().at(1).c_str());
    if (S != nnp.ntypes()) throw std::invalid_argument("NNP::read_basis(FILE*,NNP&,const char*): invalid number of types.");
    if (nnp.index(next().at(1)) != II)
        throw std::invalid_argument("NNP::read_basis(FILE*,NNP&,const char*): invalid central types.\n");
    Tables& T = *nnp.t_;
    const int NP = S * (S + 1) / 2;
    // unpack the flat basis tables into one block per (centre, slot)
    std::vector<std::vector<double>> rrs(S * S), reta(S * S), aeta(S * NP), azeta(S * NP);
    std::vector<std::vector<int>> alam(S * NP);
    for (int t = 0; t < S * S; ++t)
        for (int m = 0; m < T.rad_n[t]; ++m) {
            rrs[t].push_back(T.rad_rs[T.rad_off[t] + m]);
            reta[t].push_back(T.rad_eta[T.rad_off[t] + m]);
        }
    for (int t = 0; t < S * NP; ++t)
        for (int m = 0; m < T.ang_n[t]; ++m) {
            aeta[t].push_back(T.ang_eta[T.ang_off[t] + m]);
            azeta[t].push_back(T.ang_zeta[T.ang_off[t] + m]);
            alam[t].push_back(T.ang_lambda[T.ang_off[t] + m]);
        }
    for (int j = 0; j < S; ++j) {
        const int t = II * S + nnp.index(next().at(1));
        const auto& h = next();
        const int cut = lookup(kCut, 5, h.at(1), "cutoff");
        if (cut != T.cutoff) throw std::invalid_argument("NNP::read_basis: the cutoff function differs from the potential's");
        const int n = std::atoi(h.at(4).c_str());
        if (n < 0) throw std::invalid_argument("BasisR::read: invalid number of functions");
        T.rad_name[t] = lookup(kRad, 7, h.at(3), "radial function");
        T.rad_n[t] = n;
        rrs[t].clear(); reta[t].clear();
        for (int m = 0; m < n; ++m) {
            const auto& v = next();
            rrs[t].push_back(std::atof(v.at(0).c_str()));
            reta[t].push_back(std::atof(v.at(1).c_str()));
        }
    }
    for (int j = 0; j < S; ++j)
        for (int k = j; k < S; ++k) {
            const auto& pj = next();
            const int t = II * NP + pair_index(nnp.index(pj.at(1)), nnp.index(pj.at(2)));
            const auto& h = next();
            const int n = std::atoi(h.at(4).c_str());
            if (n < 0) throw std::invalid_argument("BasisA::read: invalid number of functions");
            T.ang_name[t] = lookup(kAng, 7, h.at(3), "angular function");
            T.ang_n[t] = n;
            aeta[t].clear(); azeta[t].clear(); alam[t].clear();
            for (int m = 0; m < n; ++m) {
                const auto& v = next();
                aeta[t].push_back(std::atof(v.at(0).c_str()));
                azeta[t].push_back(std::atof(v.at(1).c_str()));
                alam[t].push_back(std::atoi(v.at(2).c_str()));
            }
        }
    T.rad_rs.clear(); T.rad_eta.clear(); T.ang_eta.clear(); T.ang_zeta.clear(); T.ang_lambda.clear();
    for (int t = 0; t < S * S; ++t) {
        T.rad_off[t] = (int)T.rad_rs.size();
        T.rad_rs.insert(T.rad_rs.end(), rrs[t].begin(), rrs[t].end());
        T.rad_eta.insert(T.rad_eta.end(), reta[t].begin(), reta[t].end());
    }
    for (int t = 0; t < S * NP; ++t) {
        T.ang_off[t] = (int)T.ang_eta.size();
        T.ang_eta.insert(T.ang_eta.end(), aeta[t].begin(), aeta[t].end());
        T.ang_zeta.insert(T.ang_zeta.end(), azeta[t].begin(), azeta[t].end());
        T.ang_lambda.insert(T.ang_lambda.end(), alam[t].begin(), alam[t].end());
    }
    if (T.rad_rs.empty()) { T.rad_rs.push_back(0); T.rad_eta.push_back(0); }
    if (T.ang_eta.empty()) { T.ang_eta.push_back(0); T.ang_zeta.push_back(0); T.ang_lambda.push_back(0); }
    int nin = 0;
    for (int b = 0; b < S; ++b) nin += T.rad_n[II * S + b];
    for (int p = 0; p < NP; ++p) nin += T.ang_n[II * NP + p];
    if (nin != T.nin[II]) {
        // the network's input layer follows the basis (NNH::init_input, src/nnp/nnh.cpp:94-112,
        // then NN::ANN::resize at the training call site, src/nnp/nnptefr2.cpp): hidden
        // layers and neuron are kept, parameters zeroed, input scaling reset to identity
        const int old = T.nin[II];
        const int L = T.nn_nlayer[II];
        int np_old = 0, np_new = 0, prev_o = old, prev_n = nin;
        for (int l = 0; l < L; ++l) {
            const int no = T.nn_nodes[T.nn_nodes_off[II] + l];
            np_old += no + no * prev_o;
            np_new += no + no * prev_n;
            prev_o = prev_n = no;
        }
        std::vector<double> params, inpw, inpb;
        for (int a = 0; a < S; ++a) {
            const int off = (int)params.size(), ioff = (int)inpw.size();
            if (a == II) {
                params.insert(params.end(), np_new, 0.0);
                inpw.insert(inpw.end(), nin, 1.0);
                inpb.insert(inpb.end(), nin, 0.0);
            } else {
                const int cnt = (a + 1 < S ? T.nn_par_off[a + 1] : (int)T.params.size()) - T.nn_par_off[a];
                params.insert(params.end(), T.params.begin() + T.nn_par_off[a], T.params.begin() + T.nn_par_off[a] + cnt);
                inpw.insert(inpw.end(), T.inpw.begin() + T.nn_inp_off[a], T.inpw.begin() + T.nn_inp_off[a] + T.nin[a]);
                inpb.insert(inpb.end(), T.inpb.begin() + T.nn_inp_off[a], T.inpb.begin() + T.nn_inp_off[a] + T.nin[a]);
            }
            T.nn_par_off[a] = off;
            T.nn_inp_off[a] = ioff;
        }
        (void)np_old;
        T.params = params;
        T.inpw = inpw;
        T.inpb = inpb;
        T.nin[II] = nin;
    }
    if (nnp.ctx_) { ann_destroy(nnp.ctx_); nnp.ctx_ = nullptr; }
    nnp.dev_->vlist_id = 0;
}

void NNP::write(const char* file, const NNP& nnp) {
    FILE* w = std::fopen(file, "w");
    if (!w) throw std::runtime_error(std::string("NNP::write(const char*,const NNP&): Could not write to nnh file: \"") + file + "\"");
    const Tables& T = *nnp.t_;
    const int S = nnp.ntypes(), NP = S * (S + 1) / 2;
    std::fprintf(w, "ann\nntypes %i\n", S);
    for (const auto& t : nnp.types_) std::fprintf(w, "name %s mass %f energy %f \n", t.name.c_str(), t.mass, t.energy);
    std::fprintf(w, "rc %f\n", nnp.rc_);
    for (int a = 0; a < S; ++a) {
        std::fprintf(w, "basis %s\n", nnp.types_[a].name.c_str());
        for (int b = 0; b < S; ++b) {
            const int t = a * S + b;
            std::fprintf(w, "basis_radial %s\nBasisR %s %f %s %i\n", nnp.types_[b].name.c_str(), kCut[T.cutoff], nnp.rc_,
                         kRad[T.rad_name[t]], T.rad_n[t]);
            for (int m = 0; m < T.rad_n[t]; ++m)
                std::fprintf(w, "\t%.15f %.15f\n", T.rad_rs[T.rad_off[t] + m], T.rad_eta[T.rad_off[t] + m]);
        }
        for (int j = 0; j < S; ++j)
            for (int k = j; k < S; ++k) {
                const int t = a * NP + pair_index(j, k);
                std::fprintf(w, "basis_angular %s %s\nBasisA %s %f %s %i\n", nnp.types_[j].name.c_str(),
                             nnp.types_[k].name.c_str(), kCut[T.cutoff], nnp.rc_, kAng[T.ang_name[t]], T.ang_n[t]);
                int flag = 0;
                for (int m = 0; m < T.ang_n[t]; ++m) {
                    const int o = T.ang_off[t] + m;
                    if (m > 0 && std::fabs(T.ang_eta[o] - T.ang_eta[o - 1]) >= 1e-6) ++flag;
                    std::fprintf(w, "\t%.15f %.15f %i %i\n", T.ang_eta[o], T.ang_zeta[o], T.ang_lambda[o], flag);
                }
            }
    }
    const char* neuname[19] = {};
    for (const auto& kv : kNeuron) neuname[kv.second] = kv.first.c_str();
    for (int a = 0; a < S; ++a) {
        std::fprintf(w, "nn %s\nnn %i ", nnp.types_[a].name.c_str(), T.nin[a]);
        const int L = T.nn_nlayer[a];
        for (int l = 0; l < L; ++l) std::fprintf(w, "%i ", T.nn_nodes[T.nn_nodes_off[a] + l]);
        std::fprintf(w, "\nneuron %s %f\ninpw ", neuname[T.nn_neuron[a]], T.nn_c[a]);
        for (int m = 0; m < T.nin[a]; ++m) std::fprintf(w, "%.15f ", T.inpw[T.nn_inp_off[a] + m]);
        std::fprintf(w, "\noutw %.15f \ninpb ", T.outw[a]);
        for (int m = 0; m < T.nin[a]; ++m) std::fprintf(w, "%.15f ", T.inpb[T.nn_inp_off[a] + m]);
        std::fprintf(w, "\noutb %.15f \n", T.outb[a]);
        const double* p = T.params.data() + T.nn_par_off[a];
        int cnt = 0;
        for (int l = 0; l < L; ++l) {
            std::fprintf(w, "bias[%i] ", l + 1);
            for (int n = 0; n < T.nn_nodes[T.nn_nodes_off[a] + l]; ++n) std::fprintf(w, "%.15f ", p[cnt++]);
            std::fprintf(w, "\n");
        }
        int prev = T.nin[a];
        for (int l = 0; l < L; ++l) {
            const int no = T.nn_nodes[T.nn_nodes_off[a] + l];
            std::fprintf(w, "weight[%i,%i] ", l, l + 1);
            for (int n = 0; n < no; ++n)
                for (int m = 0; m < prev; ++m) std::fprintf(w, "%.15f ", p[cnt + m * no + n]);
            std::fprintf(w, "\n");
            cnt += no * prev;
            prev = no;
        }
    }
    std::fclose(w);
}

// ------------------------------------------------------------------ calculation

void NNP::init(const NNP& nnp, Structure& struc) {
    for (int i = 0; i < struc.nAtoms(); ++i) struc.symm(i).assign(nnp.nInput(struc.type(i)), 0.0);
}

namespace {

// G of the context's batch in FP64 into struc.symm
void symm_out(NNP& nnp, ann_ctx* ctx, std::vector<double>& h, Structure& struc, DevBuf<double>& G64) {
    int S, nin, ldg, np;
    check(ann_info(ctx, &S, &nin, &ldg, &np), "ann_info");
    const int N = struc.nAtoms();
    double* G = G64.get((size_t)N * ldg);
    check(ann_symm_fwd64(ctx, form_from_env(), G), "ann_symm_fwd64");
    h.resize((size_t)N * ldg);
    hipcheck(hipDeviceSynchronize(), "sync");
    hipcheck(hipMemcpy(h.data(), G, sizeof(double) * h.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    for (int i = 0; i < N; ++i) {
        auto& g = struc.symm(i);
        g.assign(nnp.nInput(struc.type(i)), 0.0);
        for (size_t k = 0; k < g.size(); ++k) g[k] = h[(size_t)i * ldg + k];
    }
}

// struc.symm (FP64, as the reference keeps it) into the device rows
double* symm_in(NNP& nnp, ann_ctx* ctx, const Structure& struc, DevBuf<double>& G64, int& ldg) {
    int S, nin, np;
    check(ann_info(ctx, &S, &nin, &ldg, &np), "ann_info");
    const int N = struc.nAtoms();
    std::vector<double> h((size_t)N * ldg, 0.0);
    for (int i = 0; i < N; ++i) {
        const auto& g = struc.symm(i);
        if ((int)g.size() != nnp.nInput(struc.type(i)))
            throw std::invalid_argument("structure symmetry functions not initialised (NNP::init / NNP::symm)");
        for (size_t k = 0; k < g.size(); ++k) h[(size_t)i * ldg + k] = g[k];
    }
    double* G = G64.get(h.size());
    hipcheck(hipMemcpy(G, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice), "hipMemcpy");
    return G;
}

void forces_out(ann_ctx* ctx, Structure& struc, double* F, double* E) {
    const int N = struc.nAtoms();
    std::vector<double> h(3 * (size_t)N);
    hipcheck(hipDeviceSynchronize(), "sync");
    hipcheck(hipMemcpy(h.data(), F, sizeof(double) * h.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    if (E) hipcheck(hipMemcpy(&struc.pe(), E, sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    for (int i = 0; i < N; ++i) struc.force(i) = Vec3{h[3 * i], h[3 * i + 1], h[3 * i + 2]};
    (void)ctx;
}

}  // namespace

void NNP::symm(NNP& nnp, Structure& struc, const NeighborList& nlist) {
    (void)nlist;  // the list is rebuilt on the device (bit-identical to nlist for the same rc)
    nnp.upload(struc);
    std::vector<double> h;
    symm_out(nnp, nnp.ctx(), h, struc, nnp.dev_->G64);
}

void NNP::symm(NNP& nnp, Structure& struc, const verlet::List& vlist) {
    nnp.upload(struc, vlist);
    std::vector<double> h;
    symm_out(nnp, nnp.ctx(), h, struc, nnp.dev_->G64);
}

double NNP::energy(NNP& nnp, const Structure& struc) {
    nnp.upload(struc);
    const int N = struc.nAtoms();
    int ldg = 0;
    double* G = symm_in(nnp, nnp.ctx(), struc, nnp.dev_->G64, ldg);
    float* Ea = nnp.dev_->Ea.get(N);
    check(ann_net_fwd64(nnp.ctx(), form_from_env(), G, Ea, nullptr), "ann_net_fwd64");
    std::vector<float> e(N);
    hipcheck(hipDeviceSynchronize(), "sync");
    hipcheck(hipMemcpy(e.data(), Ea, sizeof(float) * N, hipMemcpyDeviceToHost), "hipMemcpy");
    double pe = 0;
    for (int i = 0; i < N; ++i) pe += (double)e[i] + nnp.types_[struc.type(i)].energy;
    return pe;
}

// forces from struc.symm through the context's current list (NNP::force,
// src/nnp/nnp.cpp:461-520, 522-587)
static void force_from_symm(NNP& nnp, ann_ctx* ctx, Structure& struc, DevBuf<double>& G64, DevBuf<float>& dEb,
                            DevBuf<float>& Eab, DevBuf<float>& Pb, DevBuf<double>& Fb) {
    const int N = struc.nAtoms();
    int64_t P = 0;
    int maxn = 0;
    check(ann_nlist_size(ctx, &P, &maxn), "ann_nlist_size");
    const int form = form_from_env();
    int ldg = 0;
    double* G = symm_in(nnp, ctx, struc, G64, ldg);
    float* dE = dEb.get((size_t)N * ldg);
    float* Ea = Eab.get(N);
    float* Pd = Pb.get(3 * (size_t)std::max<int64_t>(P, 1));
    double* F = Fb.get(3 * (size_t)N);
    check(ann_net_fwd64(ctx, form, G, Ea, dE), "ann_net_fwd64");
    check(ann_symm_bwd(ctx, form, dE, Pd), "ann_symm_bwd");
    check(ann_force_reduce(ctx, Pd, Ea, 0, F, nullptr, nullptr), "ann_force_reduce");
    forces_out(ctx, struc, F, nullptr);
}

void NNP::force(NNP& nnp, Structure& struc, const NeighborList& nlist) {
    (void)nlist;
    nnp.upload(struc);
    Device& d = *nnp.dev_;
    force_from_symm(nnp, nnp.ctx(), struc, d.G64, d.dE, d.Ea, d.P, d.F);
}

void NNP::force(NNP& nnp, Structure& struc, const verlet::List& vlist) {
    nnp.upload(struc, vlist);
    Device& d = *nnp.dev_;
    force_from_symm(nnp, nnp.ctx(), struc, d.G64, d.dE, d.Ea, d.P, d.F);
}

static void compute_dev(ann_ctx* ctx, Structure& struc, DevBuf<double>& Fb, DevBuf<double>& Eb) {
    const int N = struc.nAtoms();
    double* F = Fb.get(3 * (size_t)N);
    double* E = Eb.get(1);
    check(ann_energy_forces(ctx, form_from_env(), 0, E, F, nullptr), "ann_energy_forces");
    forces_out(ctx, struc, F, E);
}

void NNP::compute(NNP& nnp, Structure& struc, const NeighborList& nlist) {
    (void)nlist;
    nnp.upload(struc);
    compute_dev(nnp.ctx(), struc, nnp.dev_->F, nnp.dev_->E);
}

void NNP::compute(NNP& nnp, Structure& struc, const verlet::List& vlist) {
    nnp.upload(struc, vlist);
    compute_dev(nnp.ctx(), struc, nnp.dev_->F, nnp.dev_->E);
}

}  // namespace ann_gpu
