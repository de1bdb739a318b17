// common.h -- shared definitions of libann_amd (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "ann_gpu.h"

namespace ann {

constexpr int kWave = 64;
constexpr int kMaxSpecies = 8;
constexpr int kMaxNeigh = 512;     // neighbour-list capacity per atom
constexpr int kMaxLayers = 4;      // hidden + output
constexpr int kMaxWidth = 128;     // widest layer supported by the network kernels
constexpr int kVecMaxWidth = 64;   // widest layer of the vector-ALU and fused network paths
constexpr int kMaxFunc = 32;       // functions per basis block
constexpr int kMaxUnionEta = 4;    // distinct angular widths per centre species (fast form)
constexpr int kGridG = 4;          // fast angular tile: widths per block
constexpr int kGridK = 8;          // fast angular tile: (lambda, zeta) combinations per block
constexpr int kMomZ = 5;           // largest integer zeta of the moment form
constexpr int kSumChunk = 4096;    // atoms per block of the per-structure sums
constexpr int kSumQ = 11;          // summed quantities: energy, virial[9], kinetic energy
constexpr int kReadbackWords = 16; // pinned host words of a context (flags and counts for the host)
constexpr int kMaxDevices = 64;    // devices a process may create contexts on
// doubles of the static scratch of the allocation-free force reduction (32 MiB
// per GPU): 9 per atom for the virial rows plus 11 per chunk partial
constexpr int64_t kNoAllocScratch = int64_t(4) << 20;

// Workgroups are dispatched round robin over the 8 XCDs, each with its own L2.
// Map dispatch order b of nb workgroups to a logical index so that XCD x runs
// one contiguous range: kernels whose workgroups read their neighbours' data
// (partner slots, partner positions) then find it in their own XCD's L2.
// A bijection on [0, nb) for every nb.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b % kXcds, q = nb / kXcds, r = nb % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
}

// per (centre, neighbour) radial block
struct RadBlock {
    int name, n, goff, par;        // par: start in rad_rs / rad_eta
};

// per (centre, pair) angular block
struct AngBlock {
    int name, n, goff, par;        // par: start in ang_eta / ang_zeta / ang_lambda
    int grid;                      // 1 if functions = {width g} x {(lambda, zeta) k}, GAUSS family
    int ng, nk;
    int gu[kGridG];                // union width index of group g
    float lamh[kGridK];            // lambda/2 of combination k
    float zeta[kGridK];
    int zint[kGridK];              // integer zeta (>=0) or -1
    int outm[kGridG * kGridK];     // function index of (g, k)
    float mcoef[kGridK][kMomZ + 1];  // moment form: C(zeta,m) 2^(m-zeta) (lambda/2)^m of combination k
};

struct NetDesc {
    int nin, nlayer, neuron;
    int nodes[kMaxLayers];
    int woff[kMaxLayers], boff[kMaxLayers];   // into the FP32 weight array
    int poff;                                  // start of this species in the packed params
    int nparam;
    int inoff;                                 // into inpw / inpb
    int wsf[kMaxLayers], wsb[kMaxLayers];      // hidden layer l: BF16x3 operand tables of W_l and W_l^T
                                               // (16-byte units into the split-weight array)
    int wc0;                                   // start of this species' vector-ALU tables (floats)
    int wcf[kMaxLayers];                       // hidden layer l: rows of W_l (floats into the table array)
    int wcs[kMaxLayers];                       // hidden layer l: row stride
    int wcb[kMaxLayers];                       // hidden layer l: its biases
    int wco;                                   // output neuron: weights (padded to 4), then its bias
    int wcn;                                   // floats of this species' tables (a multiple of 4)
    int wtf[kMaxLayers];                       // hidden layer l: rows of W_l^T (floats into the transposed array)
    int wts[kMaxLayers];                       // hidden layer l: row stride of W_l^T
    float c, outw, outb;
    double evac;
};

// Everything the kernels read about the potential (device resident).
struct DevModel {
    int S, NP, cutoff, ldg;
    float rc, rci, pirc;
    int nin[kMaxSpecies];
    int nunion[kMaxSpecies];
    float uieta2[kMaxSpecies][kMaxUnionEta];
    RadBlock rad[kMaxSpecies * kMaxSpecies];
    AngBlock ang[kMaxSpecies * (kMaxSpecies * (kMaxSpecies + 1) / 2)];
    NetDesc net[kMaxSpecies];
    int all_grid;                               // every angular block is grid-shaped
    int fast_ng, fast_nk;                       // the common grid shape (0 if blocks differ)
    int moment_z;                               // largest zeta if the moment form applies, else 0
    int pre_off[kMaxSpecies + 1];               // FP64-evaluated inputs of species a: items [pre_off[a], pre_off[a+1])
    double uieta2_64[kMaxSpecies][kMaxUnionEta];  // union widths 1/eta^2 in FP64 (FP64-evaluated inputs)
};

// one FP64-evaluated input of a species: its index k, how it is computed
// (0 radial, 1 angular by FP64 moments, 2 angular by FP64 triplets), the
// basis block (neighbour species / pair) and function within it; for the
// moment form also the union width, integer zeta and lambda
struct PreItem {
    int k, mode, blk, fn;
    int u, zint, lam, pad;
};

// device pointers of the variable-size tables
struct DevArrays {
    float* rad_rs = nullptr;
    float* rad_eta = nullptr;
    float* ang_eta = nullptr;
    float* ang_zeta = nullptr;
    float* ang_lam = nullptr;
    float* W = nullptr;          // FP32 copy of every species' parameters (packing order)
    float* inpw = nullptr;
    float* inpb = nullptr;
    double* params = nullptr;    // FP64 master parameters
    // inputs evaluated in FP64 (large |inpw inpb|): items (input, kind, block, function)
    PreItem* pitems = nullptr;
    uint8_t* preflag = nullptr;  // per input (inoff-indexed): 1 if evaluated in FP64
    double* inpb64 = nullptr;    // input bias in FP64
    float* inpb_net = nullptr;   // input bias the network applies to centred rows (0 where FP64-evaluated)
    double* rad_rs64 = nullptr;  // FP64 basis parameters for those inputs
    double* rad_eta64 = nullptr;
    double* ang_eta64 = nullptr;
    double* ang_zeta64 = nullptr;
    double* adam_m = nullptr;
    double* adam_v = nullptr;
};

// neighbour list in CSR form
struct NList {
    int64_t npairs = 0;
    int maxn = 0;
    int* off = nullptr;           // [N+1]
    int* idx = nullptr;           // [P]
    uint32_t* img = nullptr;      // [P] packed image (3 x 10-bit, +512 bias)
    int* rev = nullptr;           // [P]
    float4* geo = nullptr;        // [P] (x, y, z, |r|)
    uint8_t* nsp = nullptr;       // [P] species of the neighbour
    int64_t cap = 0;
};

inline __host__ __device__ uint32_t pack_img(int n0, int n1, int n2) {
    return ((uint32_t)(n0 + 512) << 20) | ((uint32_t)(n1 + 512) << 10) | (uint32_t)(n2 + 512);
}
inline __host__ __device__ void unpack_img(uint32_t p, int& n0, int& n1, int& n2) {
    n0 = (int)((p >> 20) & 1023) - 512;
    n1 = (int)((p >> 10) & 1023) - 512;
    n2 = (int)(p & 1023) - 512;
}

struct Timer {
    std::vector<std::pair<int, int>> ev;   // indices into the context's event pool
    double total_ms = 0;
    int count = 0;
};

struct PairState;

}  // namespace ann

struct ann_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // host copy of the model
    int S = 0, ldg = 0, nparams = 0, max_nin = 0;
    ann::DevModel hm{};
    std::vector<double> params;
    double rc64 = 0;                  // model cutoff in FP64
    std::vector<double> h_inpw, h_inpb;   // input scaling (FP64, inoff-indexed)
    double pre_ratio = 1.0;           // inputs with |inpw inpb| above this are evaluated in FP64 (< 0: all)
    int npre = 0;                     // number of FP64-evaluated inputs over all species
    int nin_total = 0;                // inputs over all species
    uint64_t scaling_gen = 0;         // bumped when the input scaling or its FP64 set changes
    uint64_t batch_gen = 0;           // bumped by every ann_set_batch
    int nb_status = 0;                // ANN_* status of the last neighbour build (capacity, image limit, symmetry)
    // training loss (ann_set_loss / ann_set_struc_weights)
    int loss_type = ANN_LOSS_MSE, loss_norm = ANN_NORM_LINEAR;
    double loss_delta = 1.0;
    const double* d_wEs = nullptr;    // per-structure energy / force weights (caller's device arrays) or null
    const double* d_wFs = nullptr;
    ann::PairState* pair = nullptr;   // device copy of the last ann_pair_compute domain (pair.hip)
    // device model
    ann::DevModel* dm = nullptr;
    ann::DevArrays da;
    // batch
    int nstruc = 0, natoms = 0;
    std::vector<int> atom_off;
    const int* d_atom_off = nullptr;
    const double* d_pos = nullptr;
    const double* d_cells = nullptr;
    const int* d_species = nullptr;
    int* d_struc_of = nullptr;       // structure index per atom
    int struc_cap = 0, atom_cap = 0, e_cap = 0;
    bool G_valid = false;             // G of the current batch is cached (training)
    int* d_rank = nullptr;            // rank of each atom within its species
    int* d_spcount = nullptr;         // atoms per species
    int* d_rkflag = nullptr;          // scratch of the rank scans: flags and prefixes [rank_cap]
    int* d_rkpre = nullptr;
    int64_t rank_cap = 0, perm_cap = 0, tiles_cap = 0;
    bool items_valid = false;         // persistent-dispatch work items of the current batch
    bool rank_valid = false;
    std::vector<int> hm_counts;       // atoms per species (host copy)
    // host copies given with the batch (ann_set_batch_host): the bin grid and
    // the species tiles are then set up without reading the device
    std::vector<double> h_cells;      // [S][9]
    std::vector<int> h_species_count; // atoms per species
    std::vector<int> h_species;       // [N]
    std::vector<int> h_rank;          // [N] rank of each atom within its species (from h_species)
    bool h_cells_valid = false, h_counts_valid = false;
    // pinned staging of the set-up uploads, one slot per upload site (see
    // staged_upload); a slot is rewritten only after its previous copy ended
    struct Staged {
        void* p = nullptr;
        size_t cap = 0;
        hipEvent_t ev = nullptr;
        bool pending = false;
    };
    Staged staged[16];                // kNStaged slots
    // the batch arrays of ann_set_batch_host (context-owned device copies)
    int* d_off_own = nullptr;
    double* d_cells_own = nullptr;
    int* d_species_own = nullptr;
    int own_struc_cap = 0, own_atom_cap = 0;
    int* d_perm = nullptr;            // species-sorted atom order
    int* d_tiles = nullptr;           // (species, start) per 16-atom tile
    int* d_spend = nullptr;           // [2S]: start then end of each species block in d_perm
    int* d_spend_end = nullptr;       // = d_spend + S
    int ntiles = 0;
    bool perm_valid = false;
    std::vector<int> grad_tiles;      // (species, layer, n0, m0) tiles of the parameter gradient
    int* d_grad_tiles = nullptr;
    double* d_gsplit = nullptr;       // per atom-split partial gradients
    int64_t gsplit_cap = 0;
    bool have_batch = false, have_nlist = false;
    double nl_rc = 0;
    ann::NList nl;
    // neighbour-build scratch
    int* d_cellinfo = nullptr;        // per structure grid description
    double* d_cellmat = nullptr;      // per structure R^-1
    int* d_bin_of = nullptr;          // per atom global bin
    int* d_shift = nullptr;           // per atom integer shift [N][3]
    int* d_bin_cnt = nullptr;
    int* d_bin_start = nullptr;
    int* d_bin_struc = nullptr;       // structure of each bin
    int* d_bin_ovf = nullptr;         // bin whose candidate set exceeds the staged search (in d_bin_cnt)
    bool grid_valid = false;          // bin grid of the current batch for grid_rc
    double grid_rc = 0;
    int grid_nbins = 0, grid_nsb = 0;  // bins and staged-search super-bins
    bool pos_ref_valid = false;       // d_pos_ref holds the positions of the last build
    int* d_bin_atoms = nullptr;
    int bin_cap = 0;
    bool nb_dirty = true;             // bin counts or build flags may be non-zero: clear before the next build
    uint64_t* d_keys = nullptr;       // [N][kMaxNeigh] sorted candidate keys
    double4* d_binrec = nullptr;      // [N] atoms in bin order: x, y, z, (atom, packed shift)
    int64_t binrec_cap = 0;
    int* d_cnt = nullptr;             // [N] neighbour counts
    int* d_flag = nullptr;            // error flags
    uint4* d_wsplit = nullptr;        // hidden-layer weights as MFMA operand tables (three BF16 terms)
    int wsplit_n = 0;                 // 16-byte entries
    bool wsplit_valid = false;        // the tables match the FP32 working copy of the weights
    float* d_wchunk = nullptr;        // network parameters in the vector-ALU kernel's layout
    float* d_wtrans = nullptr;        // hidden-layer weights transposed (backward rows of the fused kernel)
    int wtrans_n = 0;                 // floats
    int wchunk_n = 0;                 // floats
    int wchunk_max = 0;               // floats of the largest species
    bool wchunk_valid = false;        // those tables match the FP32 working copy of the weights
    int* h_rb = nullptr;              // pinned host words written by kernels for the host to read after a sync
    int rb_seq = 0;                   // completion marker of the last neighbour-build readback
    int* d_rb = nullptr;              // the same words as seen from the device
    int64_t keys_cap = 0;
    // work buffers
    float* d_G = nullptr;
    float* d_dEdG = nullptr;
    float* d_Eatom = nullptr;
    float* d_P = nullptr;
    float* d_u = nullptr;
    float* d_v = nullptr;
    double* d_F = nullptr;
    double* d_E = nullptr;
    double* d_gE = nullptr;
    double* d_atomvir = nullptr;      // [N][9]
    float* d_scr = nullptr;           // network second-backward per-atom vectors
    int64_t scr_cap = 0;
    double* d_gpart = nullptr;        // partial gradients
    int64_t gpart_cap = 0;
    int64_t work_atoms = 0, work_pairs = 0;
    int* d_adam_t = nullptr;          // Adam step counter (device, so a captured step can advance it)
    // tuning (set by ann_autotune or ann_set_tuning)
    int tune_apb[3] = {4, 4, 1};      // atoms (wavefronts) per workgroup of the moment kernels, per mode
                                      // (the autotuner's usual choice; clamped to 64 KiB of LDS per workgroup)
    int tune_angular = 0;             // 0: moment form where it applies, 1: triplet tiles
    int tune_sb = 0;                  // neighbour search blocks: 0 heuristic (2x2x2 bins where they fit), 1 single bins
    int tune_queue = 0;               // moment kernels: 0 grid dispatch, 1 persistent work queues
    int tune_net = 1;                 // networks (TRIPLET, FAST): 0 vector-ALU tiles where they fit, 1 matrix cores
                                      // (the default: faster on every measured workload)
    int tune_fused = 0;               // FAST E+F in the moment form: 0 separate G / network / adjoint kernels, 1 one fused kernel
    // persistent dispatch: (first atom, count) items per XCD queue, built on first use after ann_set_batch
    int2* d_items = nullptr;
    int* d_qoff = nullptr;
    int* d_queue = nullptr;           // counters, one per queue on its own 128-byte line
    int items_cap = 0;
    int num_cu = 0;
    // training step captured as a HIP graph and replayed
    bool graph_on = false, capturing = false, train_warm = false;
    hipStream_t gstream = nullptr;    // private non-blocking stream the graph is captured on and replayed
    hipEvent_t gev[2] = {nullptr, nullptr};
    hipGraphExec_t gexec = nullptr;
    uint64_t gen = 0;                 // bumped whenever positions, the list or the batch change
    float* d_T = nullptr;             // moment form: per-atom moments T of the last G evaluation
    uint64_t T_gen = ~0ull;           // gen at which d_T was written (~0: never)
    bool ef_graph_on = false, ef_warm = false;   // E+F graph: enabled, one eager call done
    hipGraphExec_t ef_exec = nullptr;
    std::vector<uint64_t> ef_key;               // what the captured E+F kernels depend on
    uint64_t gkey_gen = 0;
    int gkey_form = -1;
    const void* gkey_ptr[4] = {nullptr, nullptr, nullptr, nullptr};
    double gkey_w[3] = {0, 0, 0};
    // per-structure sums: structures split into chunks of at most kSumChunk atoms
    int nchunks = 0, chunk_cap = 0;
    int* d_chunks = nullptr;          // [nchunks][3]: structure, first atom, end atom
    int* d_chunk_off = nullptr;       // [S + 1]: first chunk of each structure
    double* d_part = nullptr;         // [nchunks][kSumQ] partial sums
    // molecular dynamics
    double* d_pos_ref = nullptr;      // positions at the last neighbour build
    double* d_keatom = nullptr;       // per-atom kinetic energy
    int* d_md_flag = nullptr;         // an atom moved more than skin / 2 since the build
    int md_cap = 0;
    // profiling
    bool prof = false;
    std::string prof_only;            // time only this kernel family (empty: all)
    std::map<std::string, ann::Timer> timers;
    std::vector<hipEvent_t> evpool;   // created once, reused between drains
    int evused = 0;
};

// ---- error handling ----
namespace ann {
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
}  // namespace ann

#define ANN_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ann::fail(ANN_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// kernel launchers implemented in the .hip files (return a hipError_t)
namespace ann {
struct Launch {
    ann_ctx* ctx;
    const char* name;
    Launch(ann_ctx* c, const char* n);
    ~Launch();
};
hipError_t neighbors_build(ann_ctx* ctx, double rc);
hipError_t species_perm(ann_ctx* ctx);   // species-sorted atom order and 16-atom tiles of the batch
// rebuild the BF16x3 weight tables of the matrix-core network from the FP32
// weights; tick, if set, is incremented by one thread (the Adam step counter)
hipError_t split_weights(ann_ctx* ctx, int* tick);
// exclusive scan of n ints (out[n-1] = sum of in[0..n-2]); with rb set, also
// copies nflags words of flags, then out[n-1], then seq (a completion marker)
// to the pinned words rb, and clears the first nclear flags once copied
hipError_t exclusive_scan(const int* in, int* out, int n, hipStream_t st, int* flags = nullptr,
                          int nflags = 0, int* rb = nullptr, int seq = 0, int nclear = 0);
// copy n device ints to the pinned words rb (stream ordered; read after a sync)
hipError_t readback(const int* src, int n, int* rb, hipStream_t st);
hipError_t symm_fwd(ann_ctx* ctx, int form, float* G);
// inputs flagged for FP64 evaluation: with Gc set, Gc[i][k] = G_k + inpb_k (FP64
// sum rounded to FP32) for every flagged k (other entries untouched); with G64
// set, G64[i][k] = G_k in FP64 for every flagged k
hipError_t precise_inputs(ann_ctx* ctx, float* Gc, double* G64, float* Tout = nullptr);
// G rows of the context's batch for the network, centred where FP64-evaluated:
// the FP32 basis kernel then the FP64 inputs, or, when every input is
// FP64-evaluated, the FP64 kernel alone (it then also stores the moments the
// force adjoint reads)
hipError_t network_rows(ann_ctx* ctx, int form, float* G);
// G64 = G (widened), then the flagged entries recomputed in FP64
hipError_t widen_rows(ann_ctx* ctx, const float* G, double* G64);
// Gc = G64 rounded to FP32, centred (G + inpb in FP64) where flagged
hipError_t centre_rows(ann_ctx* ctx, const double* G64, float* Gc);
hipError_t symm_bwd(ann_ctx* ctx, int form, const float* dEdG, float* P);
hipError_t symm_tangent(ann_ctx* ctx, int form, const float* u, float* v);
bool fused_supported(const ann_ctx* ctx, int form);
hipError_t symm_fused(ann_ctx* ctx, const float* inpb, float* Eatom, float* P);   // G + network + adjoint
hipError_t chunk_weights(ann_ctx* ctx);
// every weight table from the FP32 weights in one launch (split: also the
// matrix-core operand tables; tick: Adam step counter advanced by the launch)
hipError_t weight_tables(ann_ctx* ctx, int* tick, bool split);
// centred: G rows hold G + inpb at the FP64-evaluated inputs (precise_inputs);
// the network then applies inpb_net instead of inpb
hipError_t net_fwd(ann_ctx* ctx, int form, const float* G, float* Eatom, float* dEdG, bool centred = false);
hipError_t net_bwd2(ann_ctx* ctx, const float* G, const float* v, const double* gE, double* grad, bool centred = false);
hipError_t force_reduce(ann_ctx* ctx, const float* P, const float* Eatom, int with_evac, double* F,
                        double* E, double* V);
hipError_t struc_sums(ann_ctx* ctx, const float* Eatom, int with_evac, const double* avir, const double* keatom,
                      double* E, double* V, double* KE);
// allocation-free force reduction (static scratch): set-up once per device
// (that device current), scratch words a call needs, and the call itself
hipError_t noalloc_init(int device);
int64_t noalloc_need(const ann_ctx* ctx, bool virial);
hipError_t force_reduce_noalloc(ann_ctx* ctx, const float* P, const float* Eatom, int with_evac, double* F,
                                double* E, double* V);
hipError_t refresh_geometry(ann_ctx* ctx);
// the caller's host CSR list (off, idx, img [P][3]) for the current batch (ANN status)
int set_neighbors_host(ann_ctx* ctx, const int* off, const int* idx, const int* img, double rc);
void pair_free(ann_ctx* ctx);
// Upload slots of staged_upload
enum { kStChunks, kStChunkOff, kStCellInfo, kStRinv, kStBinStruc, kStTiles, kStSpend, kStItems, kStQoff,
       kStAtomOff, kStCells, kStSpecies, kStRank, kStSpCount, kStPerm, kNStaged };
static_assert(kNStaged <= 16, "one staging slot per upload site");
// Host-to-device copy of bytes through the context's pinned staging slot, on
// the context's stream, without waiting for the copy: the host data may be
// released on return (it is copied into the slot first; the slot's previous
// copy is waited for before that).
hipError_t staged_upload(ann_ctx* ctx, int slot, void* dst, const void* src, size_t bytes);
hipError_t build_items(ann_ctx* ctx);
hipError_t ensure_items(ann_ctx* ctx);   // build_items unless the current batch has them
hipError_t wrap_positions(ann_ctx* ctx, double* pos);
hipError_t md_kick_drift(ann_ctx* ctx, double dt, const double* mass, const double* F, double* vel, double* pos,
                         double lim2);
hipError_t md_kick(ann_ctx* ctx, double dt, const double* mass, const double* F, double* vel);
}  // namespace ann
