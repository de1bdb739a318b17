// reduce.hip -- atomic-free force assembly and FP64 per-structure sums.
//
// Forces: F_a = -dE/dr_a = sum_{s in N(a)} ( P[rev s] - P[s] ), with
// P[s] = dE_i/dr_ij of slot s = (i -> j, n) and rev the reverse slot
// (j -> i, -n).  This folds the reaction of every periodic image back onto
// its owner, i.e. it is the exact adjoint of the reference force loops
// (src/nnp/nnp.cpp:489-516, which apply reactions only to min() entries).
// Each atom gathers its own sum in FP64 (a wavefront per atom, lanes over
// slots, fixed butterfly; no atomics, deterministic).
// Energy per structure (NNP::energy src/nnp/nnp.cpp:436-447, vacuum energies
// optional as in NNH::energy src/nnp/nnh.cpp:118-121) and the virial
// W = -sum_slots P (x) r are FP64 tree reductions over fixed atom chunks.
// HBM-bound gather: 12 B (own P) + 12 B (reverse P) + 4 B (rev) + 16 B
// (geometry, virial) per slot.
#include "common.h"

namespace ann {

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// one wavefront per atom: lanes over its slots, fixed-order FP64 butterfly
__global__ __launch_bounds__(256) void k_force_atoms(int N, const int* __restrict__ off,
                                                     const int* __restrict__ rev, const float4* __restrict__ geo,
                                                     const float* __restrict__ P, double* __restrict__ F,
                                                     double* __restrict__ avir) {
    const int a = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (a >= N) return;
    double fx = 0, fy = 0, fz = 0;
    double v[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) v[q] = 0.0;
    const int s0 = off[a], s1 = off[a + 1];
    for (int s = s0 + lane; s < s1; s += kWave) {
        const int r = rev[s];
        const double px = P[3 * (size_t)s], py = P[3 * (size_t)s + 1], pz = P[3 * (size_t)s + 2];
        fx += (double)P[3 * (size_t)r] - px;
        fy += (double)P[3 * (size_t)r + 1] - py;
        fz += (double)P[3 * (size_t)r + 2] - pz;
        if (avir) {
            const float4 g = geo[s];
            const double rr[3] = {g.x, g.y, g.z}, pp[3] = {px, py, pz};
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
                for (int y = 0; y < 3; ++y) v[3 * x + y] -= pp[x] * rr[y];
        }
    }
    fx = wave_sum_d(fx);
    fy = wave_sum_d(fy);
    fz = wave_sum_d(fz);
    if (F && lane == 0) {
        F[3 * (size_t)a] = fx;
        F[3 * (size_t)a + 1] = fy;
        F[3 * (size_t)a + 2] = fz;
    }
    if (avir) {
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const double t = wave_sum_d(v[q]);
            if (lane == 0) avir[9 * (size_t)a + q] = t;
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
    __shared__ double red[256];
    const int c = blockIdx.x, t = threadIdx.x;
    const int s = chunks[3 * c], a0 = chunks[3 * c + 1], a1 = chunks[3 * c + 2];
    for (int q = 0; q < kSumQ; ++q) {
        const int bit = q == 0 ? 1 : (q < 10 ? 2 : 4);
        if (!(qmask & bit)) continue;
        double acc = 0.0;
        for (int a = a0 + t; a < a1; a += blockDim.x) {
            if (q == 0) acc += (double)Eatom[a] + (with_evac ? dm->net[species[a]].evac : 0.0);
            else if (q < 10) acc += avir[9 * (size_t)a + (q - 1)];
            else acc += keatom[a];
        }
        red[t] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (t < w) red[t] += red[t + w];
            __syncthreads();
        }
        if (t == 0) {
            if (part) part[(size_t)kSumQ * c + q] = red[0];
            else if (q == 0) E[s] = red[0];
            else if (q < 10) V[9 * (size_t)s + (q - 1)] = red[0];
            else KE[s] = red[0];
        }
        __syncthreads();
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

hipError_t struc_sums(ann_ctx* ctx, const float* Eatom, int with_evac, const double* avir, const double* keatom,
                      double* E, double* V, double* KE) {
    const int qmask = (E ? 1 : 0) | (V && avir ? 2 : 0) | (KE && keatom ? 4 : 0);
    if (!qmask || ctx->natoms == 0) return hipSuccess;
    Launch L(ctx, "struc_sums");
    const bool direct = ctx->nchunks == ctx->nstruc;
    k_chunk_sums<<<ctx->nchunks, 256, 0, ctx->stream>>>(ctx->d_chunks, Eatom, ctx->d_species, ctx->dm, with_evac,
                                                       avir, keatom, qmask, direct ? nullptr : ctx->d_part, E, V,
                                                       KE);
    if (!direct)
        k_chunk_combine<<<(ctx->nstruc + 15) / 16, 256, 0, ctx->stream>>>(ctx->nstruc, ctx->d_chunk_off,
                                                                          ctx->d_part, qmask, E, V, KE);
    return hipGetLastError();
}

hipError_t force_reduce(ann_ctx* ctx, const float* P, const float* Eatom, int with_evac, double* F, double* E,
                        double* V) {
    const int N = ctx->natoms;
    if (N == 0) return hipSuccess;
    {
        Launch L(ctx, "force_gather");
        k_force_atoms<<<(N + 3) / 4, 256, 0, ctx->stream>>>(N, ctx->nl.off, ctx->nl.rev, ctx->nl.geo, P, F,
                                                                 V ? ctx->d_atomvir : nullptr);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return struc_sums(ctx, Eatom, with_evac, V ? ctx->d_atomvir : nullptr, nullptr, E, V, nullptr);
}

}  // namespace ann
