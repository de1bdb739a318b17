// reduce.hip -- atomic-free force assembly and FP64 per-structure sums.
//
// Forces: F_a = -dE/dr_a = sum_{s in N(a)} ( P[rev s] - P[s] ), with
// P[s] = dE_i/dr_ij of slot s = (i -> j, n) and rev the reverse slot
// (j -> i, -n).  This folds the reaction of every periodic image back onto
// its owner, i.e. it is the exact adjoint of the reference force loops
// (src/nnp/nnp.cpp:489-516, which apply reactions only to min() entries).
// Each atom gathers its own sum in FP64 (16 lanes per atom, lanes over
// slots, fixed transposing butterfly; no atomics, deterministic).
// Energy per structure (NNP::energy src/nnp/nnp.cpp:436-447, vacuum energies
// optional as in NNH::energy src/nnp/nnh.cpp:118-121) and the virial
// W = -sum_slots P (x) r are FP64 tree reductions over fixed atom chunks.
// HBM-bound gather: 12 B (own P) + 12 B (reverse P) + 4 B (rev) + 16 B
// (geometry, virial) per slot.
#include <mutex>

#include "common.h"

namespace ann {

// Sum NV per-lane doubles over groups of 16 lanes with a transposing
// butterfly (halving steps over lane bits 8, 4, 2, 1): afterwards lane l of a
// group holds the group total of value index vidx(l) for l < NV rounded up to
// a power of two (16 lanes over NV <= 16 values, then plain steps).
template <int NV>
__device__ __forceinline__ double group16_sum_multi(double (&v)[NV], int& idx) {
    const int lane = threadIdx.x % 16;
    int base = 0;
#pragma unroll
    for (int h = NV / 2, bit = 8; h >= 1; h >>= 1, bit >>= 1) {
        const bool upper = (lane & bit) != 0;
#pragma unroll
        for (int q = 0; q < h; ++q) {
            const double send = upper ? v[q] : v[q + h];
            const double keep = upper ? v[q + h] : v[q];
            v[q] = keep + __shfl_xor(send, bit);
        }
        if (upper) base += h;
    }
    double s = v[0];
    constexpr int levels = NV >= 16 ? 0 : (NV >= 8 ? 1 : (NV >= 4 ? 2 : (NV >= 2 ? 3 : 4)));
    int bit = 1;
#pragma unroll
    for (int q = 0; q < levels; ++q, bit <<= 1) s += __shfl_xor(s, bit);
    idx = base;
    return s;
}

// 16 lanes per atom (four atoms per wavefront): lanes over the atom's slots,
// then one fixed-order transposing FP64 reduction of the 3 force (and 9
// virial) sums
template <bool VIR>
__global__ __launch_bounds__(256) void k_force_atoms(int N, const int* __restrict__ off,
                                                     const int* __restrict__ rev, const float4* __restrict__ geo,
                                                     const float* __restrict__ P, double* __restrict__ F,
                                                     double* __restrict__ avir) {
    constexpr int NV = VIR ? 16 : 4;           // 3 (+ 9) sums padded to a power of two
    const int a = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x / 16) + threadIdx.x / 16;
    const int sub = threadIdx.x % 16;
    // every lane takes part in the shuffles; atoms beyond N sum nothing
    const int s0 = a < N ? off[a] : 0, s1 = a < N ? off[a + 1] : 0;
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.0;
    for (int s = s0 + sub; s < s1; s += 16) {
        const int r = rev[s];
        const double px = P[3 * (size_t)s], py = P[3 * (size_t)s + 1], pz = P[3 * (size_t)s + 2];
        v[0] += (double)P[3 * (size_t)r] - px;
        v[1] += (double)P[3 * (size_t)r + 1] - py;
        v[2] += (double)P[3 * (size_t)r + 2] - pz;
        if (VIR) {
            const float4 g = geo[s];
            const double rr[3] = {g.x, g.y, g.z}, pp[3] = {px, py, pz};
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
                for (int y = 0; y < 3; ++y) v[3 + 3 * x + y] -= pp[x] * rr[y];
        }
    }
    int idx;
    const double t = group16_sum_multi<NV>(v, idx);
    constexpr int kRep = 16 / NV;              // lanes holding the same value index
    if (a < N && sub % kRep == 0) {
        if (idx < 3) {
            if (F) F[3 * (size_t)a + idx] = t;
        } else if (VIR && idx < 12) {
            avir[9 * (size_t)a + idx - 3] = t;
        }
    }
}

// Per-structure sums over chunks of at most kSumChunk atoms (a structure of a
// million atoms is spread over many blocks): quantity q = 0 energy (FP32 atom
// energies + optional vacuum energies), 1..9 virial, 10 kinetic energy.  Each
// block reduces its chunk with a fixed tree; when every structure is a single
// chunk the block writes the result, otherwise k_chunk_combine adds the
// chunk partials of a structure in chunk order (deterministic either way).
__global__ __launch_bounds__(256) void k_chunk_sums(const int* __restrict__ chunks, const float* __restrict__ Eatom,
                                                    const int* __restrict__ species, const DevModel* __restrict__ dm,
                                                    int with_evac, const double* __restrict__ avir,
                                                    const double* __restrict__ keatom, int qmask,
                                                    double* __restrict__ part, double* __restrict__ E,
                                                    double* __restrict__ V, double* __restrict__ KE) {
    __shared__ double red[256 / kWave][kSumQ];
    const int c = blockIdx.x, t = threadIdx.x, lane = t % kWave, w = t / kWave;
    const int s = chunks[3 * c], a0 = chunks[3 * c + 1], a1 = chunks[3 * c + 2];
    // every quantity in one pass over the chunk (thread-strided, fixed order)
    double acc[kSumQ];
#pragma unroll
    for (int q = 0; q < kSumQ; ++q) acc[q] = 0.0;
    for (int a = a0 + t; a < a1; a += blockDim.x) {
        if (qmask & 1) acc[0] += (double)Eatom[a] + (with_evac ? dm->net[species[a]].evac : 0.0);
        if (qmask & 2) {
#pragma unroll
            for (int q = 1; q < 10; ++q) acc[q] += avir[9 * (size_t)a + (q - 1)];
        }
        if (qmask & 4) acc[10] += keatom[a];
    }
    // wavefront butterflies, then the four wavefront partials in order
#pragma unroll
    for (int q = 0; q < kSumQ; ++q) {
        double v = acc[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[w][q] = v;
    }
    __syncthreads();
    if (t < kSumQ) {
        const int q = t;
        const int bit = q == 0 ? 1 : (q < 10 ? 2 : 4);
        if (qmask & bit) {
            double v = red[0][q];
            for (int k = 1; k < 256 / kWave; ++k) v += red[k][q];
            if (part) part[(size_t)kSumQ * c + q] = v;
            else if (q == 0) E[s] = v;
            else if (q < 10) V[9 * (size_t)s + (q - 1)] = v;
            else KE[s] = v;
        }
    }
}

__global__ void k_chunk_combine(int S, const int* __restrict__ chunk_off, const double* __restrict__ part,
                                int qmask, double* __restrict__ E, double* __restrict__ V, double* __restrict__ KE) {
    const int s = blockIdx.x * (blockDim.x / 16) + threadIdx.x / 16, q = threadIdx.x % 16;
    if (s >= S || q >= kSumQ) return;
    const int bit = q == 0 ? 1 : (q < 10 ? 2 : 4);
    if (!(qmask & bit)) return;
    double acc = 0.0;
    for (int c = chunk_off[s]; c < chunk_off[s + 1]; ++c) acc += part[(size_t)kSumQ * c + q];
    if (q == 0) E[s] = acc;
    else if (q < 10) V[9 * (size_t)s + (q - 1)] = acc;
    else KE[s] = acc;
}

// the chunk sums with the chunk partials in `part` (the context's, or scratch)
static hipError_t launch_sums(ann_ctx* ctx, const float* Eatom, int with_evac, const double* avir,
                              const double* keatom, double* part, double* E, double* V, double* KE) {
    const int qmask = (E ? 1 : 0) | (V && avir ? 2 : 0) | (KE && keatom ? 4 : 0);
    if (!qmask || ctx->natoms == 0) return hipSuccess;
    const bool direct = ctx->nchunks == ctx->nstruc;
    k_chunk_sums<<<ctx->nchunks, 256, 0, ctx->stream>>>(ctx->d_chunks, Eatom, ctx->d_species, ctx->dm, with_evac,
                                                       avir, keatom, qmask, direct ? nullptr : part, E, V, KE);
    if (!direct)
        k_chunk_combine<<<(ctx->nstruc + 15) / 16, 256, 0, ctx->stream>>>(ctx->nstruc, ctx->d_chunk_off, part, qmask,
                                                                          E, V, KE);
    return hipGetLastError();
}

// the per-atom gather of forces (and virial rows into avir)
static hipError_t launch_gather(ann_ctx* ctx, const float* P, double* F, double* avir) {
    const int N = ctx->natoms;
    if (avir)
        k_force_atoms<true><<<(N + 15) / 16, 256, 0, ctx->stream>>>(N, ctx->nl.off, ctx->nl.rev, ctx->nl.geo, P, F,
                                                                    avir);
    else
        k_force_atoms<false><<<(N + 15) / 16, 256, 0, ctx->stream>>>(N, ctx->nl.off, ctx->nl.rev, ctx->nl.geo, P, F,
                                                                     nullptr);
    return hipGetLastError();
}

hipError_t struc_sums(ann_ctx* ctx, const float* Eatom, int with_evac, const double* avir, const double* keatom,
                      double* E, double* V, double* KE) {
    if (!(E || (V && avir) || (KE && keatom)) || ctx->natoms == 0) return hipSuccess;
    Launch L(ctx, "struc_sums");
    return launch_sums(ctx, Eatom, with_evac, avir, keatom, ctx->d_part, E, V, KE);
}

hipError_t force_reduce(ann_ctx* ctx, const float* P, const float* Eatom, int with_evac, double* F, double* E,
                        double* V) {
    const int N = ctx->natoms;
    if (N == 0) return hipSuccess;
    hipError_t e;
    {
        Launch L(ctx, "force_gather");
        e = launch_gather(ctx, P, F, V ? ctx->d_atomvir : nullptr);
    }
    if (e != hipSuccess) return e;
    return struc_sums(ctx, Eatom, with_evac, V ? ctx->d_atomvir : nullptr, nullptr, E, V, nullptr);
}

// ---- allocation-free force reduction ----
//
// For callers that cannot allocate device memory (code under HIP-graph
// capture, tight serving loops): the per-atom virial rows and the chunk
// partials live in this file-scope array instead of context buffers.  Each GPU
// holds its own copy for the life of the process; its device address is taken
// once per device at context creation.  Every call on a GPU shares the array,
// so calls that overlap on one GPU race.  Nothing needs clearing: each call
// writes every scratch word it reads before reading it.
static __device__ __align__(256) double scratch[kNoAllocScratch];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

hipError_t noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    double* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

int64_t noalloc_need(const ann_ctx* ctx, bool virial) {
    return (virial ? 9 * (int64_t)ctx->natoms : 0) +
           (ctx->nchunks != ctx->nstruc ? (int64_t)kSumQ * ctx->nchunks : 0);
}

hipError_t force_reduce_noalloc(ann_ctx* ctx, const float* P, const float* Eatom, int with_evac, double* F, double* E,
                                double* V) {
    double* scr = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) scr = g_scratch_ptr[ctx->device];
    }
    if (!scr) return hipErrorNotInitialized;
    if (ctx->natoms == 0) return hipSuccess;
    double* avir = V ? scr : nullptr;
    double* part = scr + (V ? 9 * (size_t)ctx->natoms : 0);
    hipError_t e = launch_gather(ctx, P, F, avir);
    if (e != hipSuccess) return e;
    return launch_sums(ctx, Eatom, with_evac, avir, nullptr, part, E, V, nullptr);
}

}  // namespace ann
