/*
 * ann_gpu.h -- C ABI of the MI355X-native Behler-Parrinello energy / force /
 * weight-gradient path (libann_amd.so, gfx950).
 *
 * Conventions
 *  - Every entry point returns 0 on success or a negative ANN_ERR_* code;
 *    ann_last_error() gives the message of the last failure on the calling
 *    thread.  No exception crosses the ABI.  The reference throws
 *    std::invalid_argument / std::runtime_error from the same places
 *    (src/nnp/cutoff.cpp:51-52, src/nnp/basis_radial.cpp:203, src/ml/nn.cpp:
 *    1022-1023); the C++ shim rethrows the message as std::runtime_error.
 *  - Array arguments named d_* are DEVICE pointers (HBM) on the context's
 *    device; everything else is host memory.  Work is enqueued on the stream
 *    set with ann_set_stream (default: the null stream) and is asynchronous
 *    unless stated otherwise.
 *  - One context per host thread; a context holds its own device buffers.
 *  - Enumerations use the reference numbering: Cutoff::Name
 *    (src/nnp/cutoff.hpp:22-28), BasisR::Name (src/nnp/basis_radial.hpp:27-35),
 *    BasisA::Name (src/nnp/basis_angular.hpp:27-35), NN::Neuron
 *    (src/ml/nn.hpp:83-107).
 *  - Units and formulas are the reference's (radial eta in 1/A, angular eta in
 *    A, cosine with the vectors pointing towards the centre atom).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ANN_OK = 0,
    ANN_ERR_ARG = -1,          /* invalid argument or descriptor */
    ANN_ERR_UNSUPPORTED = -2,  /* basis / neuron / network shape not supported */
    ANN_ERR_CAPACITY = -3,     /* an atom has more neighbours than the list capacity */
    ANN_ERR_HIP = -4,          /* HIP runtime error */
    ANN_ERR_STATE = -5         /* call out of order (e.g. no batch / neighbour list) */
};

/* computation forms */
enum {
    ANN_FORM_PLAIN = 0,  /* FP32, reference-shaped loops: per-triplet function
                            evaluation, LDS-atomic accumulation, FP32 GEMV network */
    ANN_FORM_FAST = 1,   /* angular sums as per-neighbour tensor moments where every
                            angular block is GAUSS with integer zeta <= 5 (exact
                            factorisation, O(n) per atom), otherwise species-blocked
                            triplet tiles (GAUSS: pair-separable exponentials; other
                            families: a distance factor per pair of legs); atomic-free
                            adjoint, network layers on BF16 MFMA with split hi+lo
                            products */
    ANN_FORM_TRIPLET = 2 /* FAST with the triplet tiles everywhere (no moment form) */
};

/* Potential description, flat arrays (host memory, copied by ann_create).
 * S species, NP = S(S+1)/2 neighbour pairs in TensorTri order
 * p(j<=k) = k(k+1)/2 + j (src/mem/tensortri.hpp:196-201).  Radial tables are
 * indexed [centre*S + neighbour], angular tables [centre*NP + p]; *_off give
 * where each basis starts in the value arrays.  Network of species a has
 * nn_nlayer[a] layers (hidden + output, output size 1) of sizes
 * nn_nodes[nn_nodes_off[a] + l]; parameters in the reference packing (all
 * biases layer by layer, then each weight matrix column-major,
 * src/ml/nn.cpp:846-865) at params[nn_par_off[a]]; input scaling at
 * inpw/inpb[nn_inp_off[a]]; outw/outb/evac per species.
 * Limits (ANN_ERR_UNSUPPORTED beyond them): 1..8 species, at most 32
 * functions per radial or angular basis, 1..4 layers per network, layer widths
 * 1..128.
 * Replaces NNP::read + NNH::init_input (src/nnp/nnp.cpp:243-307,
 * src/nnp/nnh.cpp:94-112) as the way a potential reaches the device. */
typedef struct ann_model_desc {
    int nspecies;
    double rc;
    int cutoff;
    const int* rad_name; const int* rad_n; const int* rad_off;
    const double* rad_rs; const double* rad_eta;
    const int* ang_name; const int* ang_n; const int* ang_off;
    const double* ang_eta; const double* ang_zeta; const int* ang_lambda;
    const int* nn_nlayer; const int* nn_nodes_off; const int* nn_nodes;
    const int* nn_neuron; const double* nn_c;
    const int* nn_par_off; const double* params;
    const int* nn_inp_off; const double* inpw; const double* inpb;
    const double* outw; const double* outb; const double* evac;
} ann_model_desc;

/* A batch of periodic structures (Structure, src/struc absent; contract from
 * its call sites src/nnp/nnp.cpp:316-427): atoms of structure s are
 * [atom_off[s], atom_off[s+1]); positions Cartesian (any image), cells
 * row-major 3x3 with the lattice vectors as COLUMNS. */
typedef struct ann_batch {
    int nstruc;
    const int* atom_off;     /* HOST [nstruc+1] */
    const int* d_atom_off;   /* device copy of atom_off */
    const double* d_pos;     /* device [N][3] */
    const double* d_cells;   /* device [nstruc][9] */
    const int* d_species;    /* device [N], species index in [0,S) */
} ann_batch;

typedef struct ann_ctx ann_ctx;

/* ---- context ---- */
const char* ann_last_error(void);
int ann_create(const ann_model_desc* desc, ann_ctx** out);
int ann_destroy(ann_ctx* ctx);
int ann_set_stream(ann_ctx* ctx, void* hip_stream);
/* number of inputs of species a (NNH::nInput) and the G row stride */
int ann_info(ann_ctx* ctx, int* nspecies, int* max_ninput, int* ldg, int* nparams_total);
/* parameters of all species concatenated, reference packing (replaces
 * NN::ANN operator<< , src/ml/nn.cpp:867-885, as used by
 * NNPTEFR::error_cost_o1, src/nnp/nnptefr2.cpp:1293-1303) */
int ann_set_params(ann_ctx* ctx, const double* params, int n);
/* input / output scaling of one species (ANN::inpw/inpb/outw/outb,
 * src/ml/nn.hpp:299-306) */
int ann_set_scaling(ann_ctx* ctx, int species, const double* inpw, const double* inpb,
                    double outw, double outb);

/* ---- neighbour list (NeighborList::build; src/nnp/nnpce.cpp:282) ----
 * Cell-list build with periodic image vectors: every (j, n) with
 * 0 < |r_i - r_j - R n| < rc (rc <= 0: the model cutoff), canonical order
 * j ascending then n lexicographic, plus the reverse-slot map
 * rev: (i->j, n) <-> (j->i, -n).  The pointers of the batch must stay valid
 * until the next ann_set_batch. */
int ann_set_batch(ann_ctx* ctx, const ann_batch* batch);
/* As ann_set_batch, with the atom offsets [nstruc+1], cells [nstruc][9] and
 * species [N] in HOST memory (copied into the context's own device arrays on
 * its stream, without waiting) and only the positions d_pos [N][3] in device
 * memory.  The bin grid of the next build and the species tiles of the
 * networks are then set up from the host copies, so the batch set-up waits
 * for nothing on the device: a training loop that draws a new batch every
 * step (Batch, src/ml/batch.hpp:18-55, at src/nnp/nnptefr2.cpp:1309-1315)
 * queues it behind the previous step.  The host arrays may be released on
 * return; d_pos must stay valid until the next set-up. */
int ann_set_batch_host(ann_ctx* ctx, int nstruc, const int* atom_off, const double* d_pos, const double* cells,
                       const int* species);
int ann_neighbors(ann_ctx* ctx, double rc);
/* A neighbour list supplied by the caller, in place of ann_neighbors (the
 * reference's verlet::List reuse, src/nnp/nnp.cpp:377-427, 522-587, 667-736, and
 * LAMMPS-style callers with their own lists): host CSR off [N+1], partner
 * index idx [P] (batch-global, in the centre's structure) and image vector
 * img [P][3] per entry, r_i - r_j - R n; rows in any order, at most 512
 * entries per atom, images within +-511 cells, no entry twice, no atom with
 * itself at n = 0, and every (i -> j, n) matched by (j -> i, -n) (else
 * ANN_ERR_ARG).  rc is the list's cutoff (entries beyond the model cutoff
 * contribute nothing).  The pair geometry comes from the batch's current
 * positions; synchronous. */
int ann_set_neighbors(ann_ctx* ctx, const int* off, const int* idx, const int* img, double rc);
/* synchronous: total pairs and the largest per-atom count */
int ann_nlist_size(ann_ctx* ctx, int64_t* npairs, int* maxn);
/* device-to-device copies of the list (any pointer may be NULL):
 * off [N+1], idx [P], img [P][3], rev [P], geo [P][4] = (x, y, z, |r|) of
 * r_i - r_j - R n in FP32 */
int ann_nlist_copy(ann_ctx* ctx, int* d_off, int* d_idx, int* d_img, int* d_rev, float* d_geo);

/* ---- forward / backward pairs (shaped for a torch.autograd.Function) ---- */
/* NNP::symm (src/nnp/nnp.cpp:333-375): G [N][ldg] FP32 */
int ann_symm_fwd(ann_ctx* ctx, int form, float* d_G);
/* ANN::fpbp + DODZ::grad (src/ml/nn.cpp:1367-1389, 1801-1830): atom energies
 * (network output, no vacuum energy) and dE/dG [N][ldg] */
int ann_net_fwd(ann_ctx* ctx, int form, const float* d_G, float* d_Eatom, float* d_dEdG);
/* force adjoint of NNP::force (src/nnp/nnp.cpp:461-520) per slot:
 * P[s] = dE_i / d r_ij, [P][3] FP32 */
int ann_symm_bwd(ann_ctx* ctx, int form, const float* d_dEdG, float* d_P);
/* F_a = sum_{s in N(a)} (P[rev s] - P[s]) (FP64 accumulation), per-structure
 * energy E_s = sum_i Eatom_i (+ vacuum energies if with_evac, as
 * NNP::energy), virial W = -sum P (x) r; any output may be NULL */
int ann_force_reduce(ann_ctx* ctx, const float* d_P, const float* d_Eatom, int with_evac,
                     double* d_F, double* d_E, double* d_virial);
/* ann_force_reduce without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP-graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as ann_force_reduce; the per-atom virial rows
 * and the per-chunk partial sums it needs live in a static per-GPU scratch
 * array of 32 MiB (4 Mi doubles: 9 per atom when d_virial is set, plus 11 per
 * chunk of at most 4096 atoms when a structure spans several chunks); larger
 * batches are rejected with ANN_ERR_CAPACITY.  Calls that overlap on one GPU
 * (any context) share that scratch and race.  Kernel timing (ann_profile)
 * does not cover this entry point. */
int ann_force_reduce_noalloc(ann_ctx* ctx, const float* d_P, const float* d_Eatom, int with_evac,
                             double* d_F, double* d_E, double* d_virial);
/* The same forward pair with G in FP64, as the reference stores it
 * (Structure::symm, a VecXd per atom; NNP::symm src/nnp/nnp.cpp:333-375 and
 * NNP::energy :436-447, which applies inpw (G + inpb) in FP64,
 * src/ml/nn.cpp:1346).  Inputs flagged for FP64 evaluation (see
 * ann_set_precise_ratio) are computed in FP64 throughout; the others are the
 * FP32 values of ann_symm_fwd widened.  ann_net_fwd64 forms G + inpb in FP64
 * for the flagged inputs before rounding to the FP32 network.  G [N][ldg]. */
int ann_symm_fwd64(ann_ctx* ctx, int form, double* d_G);
int ann_net_fwd64(ann_ctx* ctx, int form, const double* d_G, float* d_Eatom, float* d_dEdG);
/* Inputs whose scaled value inpw (G + inpb) an FP32 G cannot resolve are
 * evaluated in FP64: every input k with |inpw_k inpb_k| > ratio (default 1;
 * ratio < 0 flags every input).  The reference trainer's preconditioning
 * inpb = -<G>, inpw = 1/sigma(G) (src/nnp/nnptefr2.cpp:776-867) gives
 * |inpw inpb| = <G>/sigma, which reaches ~1e5 for an almost constant input:
 * the ~1e-7 relative rounding of an FP32 G then becomes an O(1e-2) error of
 * the network input (and ratios of 2-20, common, cost the energy bar).  The
 * fused entry points (ann_energy_forces, ann_md_step, ann_train_step) and the
 * FP64 pair above use the FP64 values; ann_symm_fwd / ann_net_fwd stay FP32.
 * The set is recomputed by ann_set_scaling.  No counterpart in the reference
 * (which is FP64 throughout). */
int ann_set_precise_ratio(ann_ctx* ctx, double ratio);
/* number of inputs (over all species) currently evaluated in FP64 */
int ann_precise_count(ann_ctx* ctx, int* count);
/* fused NNP::compute (src/nnp/nnp.cpp:601-665): energies, forces, virials */
int ann_energy_forces(ann_ctx* ctx, int form, int with_evac, double* d_E, double* d_F,
                      double* d_virial);

/* ---- molecular dynamics either side of the E+F call ----
 * A list built with ann_neighbors(ctx, rc + skin) stays valid while no atom
 * moves more than skin/2 (verlet::List reuse, src/nnp/nnp.cpp:377-427 and
 * src/torch/torch.cpp:272-276 rebuilding every `stride` steps); entries
 * beyond the model cutoff contribute nothing. */
/* pair geometry of every slot from the batch's current positions and the
 * list's image vectors */
int ann_refresh_geometry(ann_ctx* ctx);
/* batch positions returned to the cell, x <- R frac(R^-1 x) (WRITES the batch
 * positions; Cell::returnToCell, src/format/xyz_struc.cpp:176-181).  The
 * list's image vectors refer to the old positions: rebuild afterwards. */
int ann_wrap_positions(ann_ctx* ctx);
/* one velocity-Verlet step (Verlet::compute, src/torch/integrator.cpp:177-200)
 * with PotNNPE::compute energies (src/torch/pot_nnpe.cpp:166-287, vacuum
 * energies included): v += dt/2 F/m, x += v dt (the batch positions are
 * WRITTEN, not wrapped), geometry refresh, E / F / virial at the new positions,
 * v += dt/2 F/m, kinetic energy per structure.  d_mass [N], d_vel / d_F
 * [N][3]; d_F holds the forces at the current positions on entry and the new
 * ones on exit.  d_E, d_KE [nstruc], d_virial [nstruc][9] may be NULL.  If
 * `moved` is not NULL the call synchronises and sets *moved = 1 when an atom
 * has moved more than skin/2 since the last build (rebuild before the next
 * step). */
int ann_md_step(ann_ctx* ctx, int form, double dt, const double* d_mass, double* d_vel, double* d_F,
                double* d_E, double* d_KE, double* d_virial, int* moved);

/* ---- second backward (weight gradient of a force loss) ---- */
/* v_i = sum_s dG_i/dr_ij . (u_i - u_j): forward-mode tangent of the basis
 * along the per-atom field u [N][3]; v [N][ldg] */
int ann_symm_tangent(ann_ctx* ctx, int form, const float* d_u, float* d_v);
/* grad = sum_i d/dp [ gE[s(i)] E_i + dE_i/dG . v_i ] for every parameter p
 * (species blocks concatenated, reference packing), FP64 [nparams_total];
 * replaces Cost::grad + D2ODZDI::grad (src/ml/nn.cpp:1712-1776, 2012-2103) */
int ann_net_bwd2(ann_ctx* ctx, const float* d_G, const float* d_v, const double* d_gE,
                 double* d_grad);
/* one training step on the current batch: loss (defaults of ann_set_loss and
 * ann_set_struc_weights)
 * L = (1/B) sum_s [ wE/2 ((E_s-Eref_s)/N_s)^2 + wF/(3N_s) sum |F-Fref|^2 ],
 * gradient, then the Adam update of the FP64 master parameters (ADAM::step,
 * src/opt/algo.cpp:280-296) when gamma > 0.  d_loss [2] = (energy, force) */
int ann_train_step(ann_ctx* ctx, int form, const double* d_Eref, const double* d_Fref, double wE,
                   double wF, double gamma, double* d_loss, double* d_grad);
/* Energy loss of ann_train_step (NNPTEFR::error_cost_o1,
 * src/nnp/nnptefr2.cpp:1336-1372): on dU = (E - Eref) norm(N_s),
 * MSE dU^2/2 (default), MAE |dU|, HUBER delta^2 (sqrt(1 + a^2) - 1) or ASINH
 * delta^2 (1 - sqrt(1 + a^2) + a asinh a), a = dU / delta (opt::Loss numbering,
 * src/opt/loss.hpp:18-24); norm(N) = 1 (IDENTITY), 1/N (LINEAR, default),
 * 1/sqrt N, 1/cbrt N, 1/(ln N + 1) (Norm numbering, src/nnp/nnptefr2.hpp:
 * 117-126, :602-612).  The energy term of the loss becomes
 * (1/B) sum_s wE wE_s loss(dU_s); the force term is unchanged. */
enum { ANN_LOSS_MSE = 1, ANN_LOSS_MAE = 2, ANN_LOSS_HUBER = 3, ANN_LOSS_ASINH = 4 };
enum { ANN_NORM_IDENTITY = 0, ANN_NORM_LINEAR = 1, ANN_NORM_SQRT = 2, ANN_NORM_CBRT = 3, ANN_NORM_LOG = 4 };
int ann_set_loss(ann_ctx* ctx, int loss, double delta, int norm);
/* Per-structure weights of the energy and force terms of ann_train_step
 * (device arrays [nstruc] that must stay valid while used; NULL = all 1): the
 * reference trains Taylor-augmented copies on energies only
 * (src/nnp/nnptefr2.cpp:2466-2473), i.e. with force weight 0. */
int ann_set_struc_weights(ann_ctx* ctx, const double* d_wE, const double* d_wF);
/* enable (1) or disable (0) HIP-graph capture of ann_train_step: after one
 * eager call the step is captured on a private stream and later calls with the
 * same form, buffers and weights replay it (ordered after prior work on the
 * context stream).  A new batch, list or positions re-captures.  No
 * counterpart in the reference (its trainer is host-driven). */
int ann_train_graph(ann_ctx* ctx, int enable);
/* enable (1) or disable (0) HIP-graph capture of ann_energy_forces: after one
 * eager call the E+F kernels (basis, network, adjoint, force gather, sums) are
 * captured on a private stream and later calls replay them while the form,
 * the output buffers, the list's arrays and widest row, and the tuning are
 * unchanged; a neighbour build that keeps them keeps the graph (the kernels
 * read positions and list by pointer).  No counterpart in the reference. */
int ann_ef_graph(ann_ctx* ctx, int enable);
/* copy the FP64 master parameters (device) to host */
int ann_get_params(ann_ctx* ctx, double* params, int n);
/* Adam state for a restart (NNPTEFR restart file, src/nnp/nnptefr2.cpp:73-195,
 * 501-566, holds the optimiser with the potential): first and second moments
 * (n = nparams doubles each, host) and the number of steps taken.  Set after
 * ann_set_params, which restarts Adam from zero moments. */
int ann_get_optimizer_state(ann_ctx* ctx, double* m, double* v, int* step, int n);
int ann_set_optimizer_state(ann_ctx* ctx, const double* m, const double* v, int step, int n);
/* input preconditioning statistics of the current batch (replaces the loops of
 * NNPTEFR::train, src/nnp/nnptefr2.cpp:690-787, over the training structures'
 * symmetry functions): d_G holds N x ldg FP32 rows as written by ann_symm_fwd.
 * Writes count[S] atoms per species and, on the host, stats[S][6][max_ninput]:
 * min, max, harmonic mean N / sum(1/G), mean, deviation sqrt(sum (mean - G)^2 /
 * (N - 1)) and rms sqrt(sum G^2 / N) per input; n = S * 6 * max_ninput.  Sums
 * are FP64 in a fixed order; synchronous. */
int ann_input_stats(ann_ctx* ctx, const float* d_G, int* count, double* stats, int n);

/* ---- tuning ----
 * ann_autotune times the variants on the current batch and keeps the fastest
 * (synchronous): atoms per workgroup (1, 2, 4) of the moment-form kernels
 * for G, adjoint and tangent; moment form against triplet tiles; grid
 * dispatch of the moment kernels against persistent work queues (per XCD,
 * (structure, atom-range) items, largest structure first); neighbour search
 * in 2x2x2-bin or single-bin blocks; the networks (TRIPLET and FAST) as
 * vector-ALU tiles of 16 atoms against matrix-core tiles; FAST E+F in the
 * moment form as one fused kernel (G, network and force adjoint per atom)
 * against separate kernels.  Results do not depend on the choice beyond FP32
 * summation order.  Tuning vector (5 to 8 ints): atoms per workgroup for G,
 * adjoint (also the fused kernel), tangent; angular kernels (0 moment,
 * 1 triplet); search blocks (0 heuristic, 1 single bins); moment-kernel
 * dispatch (0 grid, 1 persistent queues); network kernel (0 vector ALU where
 * the species' weights fit its 64 KiB of LDS, else matrix cores; 1 matrix
 * cores, the default); E+F kernels (0 separate, 1 fused where it applies).
 * Without ann_autotune the defaults are 4, 4, 1, 0, 0, 0, 1, 0.  No counterpart
 * in the reference. */
int ann_autotune(ann_ctx* ctx, int form);
int ann_get_tuning(ann_ctx* ctx, int* tuning, int n);
int ann_set_tuning(ann_ctx* ctx, const int* tuning, int n);

/* ---- kernel timing with HIP events on the context stream ---- */
int ann_profile(ann_ctx* ctx, int enable);
/* accumulated time (ms) and launch count of the kernel family `name`;
 * synchronises the stream */
int ann_profile_get(ann_ctx* ctx, const char* name, double* total_ms, int* count);
/* time only the kernel family `name` (null or empty: every family), so that a
 * timed region carries two event records per step instead of two per kernel */
int ann_profile_only(ann_ctx* ctx, const char* name);

/* ---- LAMMPS-style pair interface ----
 * One periodic domain: nlocal owned atoms and nghost ghost copies (LAMMPS
 * atom->nlocal / nghost), positions x [nlocal + nghost][3] and species
 * [nlocal] on the HOST, the box as the cell (lattice vectors as columns),
 * owner[g] = local index of the atom that ghost nlocal + g images (LAMMPS
 * atom->map of its tag), and a FULL neighbour list of the local atoms,
 * numneigh [nlocal] / firstneigh [nlocal] -> indices in [0, nlocal + nghost)
 * (LAMMPS list->numneigh / firstneigh), with pairs up to rc (the neighbour
 * cutoff, model cutoff + skin).  The reference names the pair style
 * (README.md:411-416) without its source; semantics of its MD caller
 * PotNNPE::compute (src/torch/pot_nnpe.cpp:166-287) with exact forces.
 * Every ghost must be a lattice image of its owner (else ANN_ERR_ARG).  The
 * forces on ghosts are folded back onto their owners: f[i] += F_i for the
 * local atoms (ghost rows of f untouched, as after LAMMPS's reverse
 * communication); *eng += E (vacuum energies included); eatom[i] += E_i and
 * virial[6] += (xx, yy, zz, xy, xz, yz) of -dE/dstrain when not NULL.
 * new_list = 0 reuses the list of the previous call (same atoms; LAMMPS
 * neighbor->ago > 0) and moves only the local positions and the cell.
 * Synchronous. */
int ann_pair_compute(ann_ctx* ctx, int form, int new_list, int nlocal, int nghost, const double* x,
                     const int* species, const double* cell, const int* owner, const int* numneigh,
                     const int* const* firstneigh, double rc, double* f, double* eng, double* eatom,
                     double* virial);

/* ---- FP64 evaluators of the per-function classes ----
 * No context.  npts points per call, device arrays, FP64 throughout,
 * synchronous; output pointers may be null where not wanted.  Names of
 * cutoffs and families as in ann_model_desc. */

/* Cutoff::cutf / cutg / compute (src/nnp/cutoff.hpp:59-61,
 * src/nnp/cutoff.cpp:62-143): f[p], g[p] = cutoff value and d/dr at r[p]. */
int ann_cutoff_eval(int cutoff, double rc, int npts, const double* d_r, double* d_f, double* d_g);

/* One radial basis (BasisR: name, rs, eta; src/nnp/basis_radial.hpp:40-88),
 * at most 64 functions; parameter arrays on the HOST. */
typedef struct ann_radial_desc {
    int cutoff; double rc; int name; int n;
    const double* rs; const double* eta;
} ann_radial_desc;
/* BasisR::symm (src/nnp/basis_radial.cpp:263-309) and BasisR::force
 * (:316-376) per distance r[p]: G [npts][n], dG/dr [npts][n], and
 * amp[p] = -sum_i dEdG[p][i] dG_i/dr (needs dEdG). */
int ann_radial_eval(const ann_radial_desc* basis, int npts, const double* d_r, double* d_G, double* d_dGdr,
                    const double* d_dEdG, double* d_amp);

/* One angular basis (BasisA: name, eta, zeta, lambda;
 * src/nnp/basis_angular.hpp:40-102), at most 64 functions; HOST arrays. */
typedef struct ann_angular_desc {
    int cutoff; double rc; int name; int n;
    const double* eta; const double* zeta; const int* lambda;
} ann_angular_desc;
/* BasisA::symm (src/nnp/basis_angular.cpp:371-420), symmd (:259-364) and
 * force (:430-563) per triplet p with legs r[p][0] = r_ij, r[p][1] = r_ik and
 * cosine cos[p]: G [npts][n]; dG [npts][n][3] = (d/dcos, d/dr_ij, d/dr_ik);
 * phi[p] = -sum_i dEdG dG_i/dcos and eta[p][2] = -sum_i dEdG dG_i/dr (need dEdG). */
int ann_angular_eval(const ann_angular_desc* basis, int npts, const double* d_r, const double* d_cos, double* d_G,
                     double* d_dG, const double* d_dEdG, double* d_phi, double* d_eta);

/* One network (NN::ANN, src/ml/nn.hpp:244-370): nin inputs, nlayer layers of
 * nodes[l] neurons (the last of size 1, linear), one neuron type with
 * sharpness c, parameters in the reference packing (all biases, then each
 * weight matrix column-major), input scaling inpw (G + inpb) (null: 1 and 0),
 * output scaling outw a + outb.  HOST arrays, copied at creation.  Node state
 * lives in 64 KiB of LDS per row: 8 (2 nin + 5 sum(nodes) + 6 max width) bytes. */
typedef struct ann_mlp_desc {
    int nin; int nlayer; const int* nodes; int neuron; double c;
    const double* params; const double* inpw; const double* inpb;
    double outw; double outb;
} ann_mlp_desc;
typedef struct ann_mlp ann_mlp;
int ann_mlp_create(const ann_mlp_desc* desc, ann_mlp** out);
/* operator<< of NN::ANN (src/ml/nn.cpp:867-885): n must equal ann_mlp_size */
int ann_mlp_set_params(ann_mlp* mlp, const double* params, int n);
int ann_mlp_set_scaling(ann_mlp* mlp, const double* inpw, const double* inpb, double outw, double outb);
int ann_mlp_size(const ann_mlp* mlp);
/* ANN::fp / fpbp (src/ml/nn.cpp:1343-1389) per row: out[p], and DODZ::dodi
 * (do/di with the input scaling, src/ml/nn.cpp:1801-1810) in dodi [npts][nin]. */
int ann_mlp_eval(ann_mlp* mlp, int npts, const double* d_in, double* d_out, double* d_dodi);
/* Gradient w.r.t. the parameters of Phi_p = dcdo[p] o_p + sum_k v[p][k] do_p/di_k
 * (dcdo null: 0; v null: 0).  v = 0 is Cost::grad (src/ml/nn.cpp:1712-1762);
 * v = e_k with dcdo = 0 is column k of D2ODZDI::d2odpdi (:2012-2100), here the
 * exact second derivative (the reference's explicit first-layer weight term
 * omits the inpw factor, :2060-2066; the two agree when inpw = 1).
 * per_row: grad [npts][nparam]; else grad [nparam] = sum over rows, in row order. */
int ann_mlp_grad(ann_mlp* mlp, int npts, const double* d_in, const double* d_dcdo, const double* d_v, int per_row,
                 double* d_grad);
void ann_mlp_destroy(ann_mlp* mlp);

#ifdef __cplusplus
}
#endif
