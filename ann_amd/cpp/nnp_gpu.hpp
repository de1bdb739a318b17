// nnp_gpu.hpp -- C++ host classes that keep the reference's NNP interface
// (src/nnp/nnp.hpp:98-122) and run every calculation through the C ABI of
// libann_amd.so (include/ann_gpu.h) on the GPU.
//
// Names, argument meaning and error behaviour follow the reference:
//   NNP::read / write             src/nnp/nnp.cpp:197-307 (same text format)
//   NNP::init                     src/nnp/nnp.cpp:316-321
//   NNP::symm                     src/nnp/nnp.cpp:333-375
//   NNP::energy                   src/nnp/nnp.cpp:436-447 (adds the vacuum energies)
//   NNP::force                    src/nnp/nnp.cpp:461-520 (uses struc.symm(i))
//   NNP::compute                  src/nnp/nnp.cpp:601-665 (pe without vacuum energies)
//   NeighborList(struc, rc)       call sites src/nnp/nnpce.cpp:282, src/nnp/test_nnp_force.cpp:66
//   Structure accessors           call sites src/nnp/nnp.cpp:316-736
// Errors are reported as std::runtime_error / std::invalid_argument, as in
// the reference.  Vectors are std::array / std::vector instead of Eigen types.
#pragma once
#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

struct ann_ctx;

namespace ann_gpu {

using Vec3 = std::array<double, 3>;

class Structure {
public:
    Structure() = default;
    explicit Structure(int natoms) { resize(natoms); }
    void resize(int natoms);
    int nAtoms() const { return (int)posn_.size(); }
    // lattice vectors are the COLUMNS of R (row-major storage)
    void init(const std::array<double, 9>& lv) { R_ = lv; }
    const std::array<double, 9>& R() const { return R_; }
    Vec3& posn(int i) { return posn_[i]; }
    const Vec3& posn(int i) const { return posn_[i]; }
    Vec3& force(int i) { return force_[i]; }
    const Vec3& force(int i) const { return force_[i]; }
    int& type(int i) { return type_[i]; }
    int type(int i) const { return type_[i]; }
    std::string& name(int i) { return name_[i]; }
    const std::string& name(int i) const { return name_[i]; }
    std::vector<double>& symm(int i) { return symm_[i]; }
    const std::vector<double>& symm(int i) const { return symm_[i]; }
    double& pe() { return pe_; }
    double pe() const { return pe_; }

private:
    std::vector<Vec3> posn_, force_;
    std::vector<int> type_;
    std::vector<std::string> name_;
    std::vector<std::vector<double>> symm_;
    std::array<double, 9> R_{};
    double pe_ = 0;
};

// one neighbour entry: image n of atom index() seen from the centre
class Neighbor {
public:
    int index() const { return index_; }
    int type() const { return type_; }
    const std::array<int, 3>& cell() const { return cell_; }
    const Vec3& r() const { return r_; }          // r_i - r_j - R n
    double dr() const { return dr_; }
    bool min() const { return min_; }             // shortest image of this partner
    int rev() const { return rev_; }              // slot of the reverse entry in the partner's list
    int index_ = 0, type_ = 0, rev_ = 0;
    std::array<int, 3> cell_{};
    Vec3 r_{};
    double dr_ = 0;
    bool min_ = false;
};

// Periodic neighbour list with images, built on the GPU (canonical order:
// partner index ascending, then image vector lexicographic).
class NeighborList {
public:
    NeighborList() = default;
    NeighborList(const Structure& struc, double rc) { build(struc, rc); }
    void build(const Structure& struc, double rc);
    int size(int i) const { return off_[i + 1] - off_[i]; }
    const Neighbor& neigh(int i, int j) const { return list_[off_[i] + j]; }
    double rc() const { return rc_; }

private:
    std::vector<int> off_;
    std::vector<Neighbor> list_;
    double rc_ = 0;
};

namespace verlet {

// one entry of a Verlet list: partner index and image vector (the reference
// reads index() and cell(), src/nnp/nnp.cpp:389-395)
class Neighbor {
public:
    int index() const { return index_; }
    const std::array<int, 3>& cell() const { return cell_; }
    int index_ = 0;
    std::array<int, 3> cell_{};
};

// Verlet list (src/struc absent; contract from src/nnp/nnp.cpp:377-427 and its
// engine src/torch/engine.cpp:64-125: rc() set to the cutoff plus skin, build()
// every stride() steps, entries reused between builds while atoms move).  Built
// on the GPU with the canonical order of NeighborList; the pair vectors are
// r_i - r_j - R n from the structure's current positions.
class List {
public:
    double& rc() { return rc_; }
    double rc() const { return rc_; }
    int& stride() { return stride_; }
    int stride() const { return stride_; }
    void build(const Structure& struc);
    void clear() { off_.assign(1, 0); list_.clear(); id_ = 0; }
    int size(int i) const { return off_[i + 1] - off_[i]; }
    const Neighbor& neigh(int i, int j) const { return list_[off_[i] + j]; }
    int nAtoms() const { return (int)off_.size() - 1; }
    // flat CSR view (off [N+1], idx, img [P][3]) and a build counter, for the GPU context
    const std::vector<int>& off() const { return off_; }
    uint64_t id() const { return id_; }
    void flat(std::vector<int>& idx, std::vector<int>& img) const;

private:
    double rc_ = 0;
    int stride_ = 1;
    uint64_t id_ = 0;
    std::vector<int> off_{0};
    std::vector<Neighbor> list_;
};

}  // namespace verlet

struct TypeInfo {
    std::string name;
    double mass = 0, energy = 0;
};

// Potential: species, global cutoff, per-species bases and networks, all in
// the reference's units and packing, plus the GPU context that evaluates it.
class NNP {
public:
    NNP();
    ~NNP();
    NNP(const NNP&) = delete;
    NNP& operator=(const NNP&) = delete;

    int ntypes() const { return (int)types_.size(); }
    double rc() const { return rc_; }
    int index(const std::string& name) const;
    const TypeInfo& type(int i) const { return types_[i]; }
    int nInput(int i) const;
    int size() const;                                   // number of parameters
    void params(std::vector<double>& p) const;          // all species, reference packing
    void set_params(const std::vector<double>& p);

    static void read(const char* file, NNP& nnp);
    static void write(const char* file, const NNP& nnp);
    // basis of one species from a basis file (src/nnp/nnp.cpp:96-156); the
    // species' network is resized to the new input count with zero parameters
    // when that count changes
    static void read_basis(const char* file, NNP& nnp, const char* name);
    static void init(const NNP& nnp, Structure& struc);
    // symmetry functions in FP64 as the reference stores them (inputs whose
    // scaling needs it are evaluated in FP64 throughout, the rest in FP32)
    static void symm(NNP& nnp, Structure& struc, const NeighborList& nlist);
    static void symm(NNP& nnp, Structure& struc, const verlet::List& vlist);
    static double energy(NNP& nnp, const Structure& struc);
    static void force(NNP& nnp, Structure& struc, const NeighborList& nlist);
    static void force(NNP& nnp, Structure& struc, const verlet::List& vlist);
    static void compute(NNP& nnp, Structure& struc, const NeighborList& nlist);
    static void compute(NNP& nnp, Structure& struc, const verlet::List& vlist);
    ann_ctx* context() { return ctx(); }     // the GPU context evaluating the potential

    // flat tables in the layout of ann_model_desc (public for the C ABI)
    struct Tables;
    std::unique_ptr<Tables> t_;

private:
    ann_ctx* ctx();
    void upload(const Structure& struc);
    void upload(const Structure& struc, const verlet::List& vlist);
    std::vector<TypeInfo> types_;
    double rc_ = 0;
    ann_ctx* ctx_ = nullptr;
    struct Device;
    std::unique_ptr<Device> dev_;
};

}  // namespace ann_gpu
