// net_mfma.hip -- per-species networks as batched GEMMs on the matrix cores.
//
// Same mathematics as ANN::fpbp + DODZ::grad (src/ml/nn.cpp:1367-1389,
// 1801-1830).  Atoms are grouped by species (a species-sorted permutation of
// the batch) into tiles of 16 rows; one wavefront owns a tile and runs every
// layer as 16 x N x K products on the matrix cores (BF16 operands split into
// three terms, FP32 accumulation; mm_bf16x3 below):
//   forward:  Z_l = A_{l-1} W_l^T (+ b_l), activation on the accumulator;
//   backward: g_{L-1} = outw W_L^T . sigma'_{L-1}, g_{l-1} = (g_l W_l) . sigma'_{l-1},
//             dE/dG = (g_0 W_0) . inpw.
// Activations move between the accumulator layout and the operand layout
// through LDS; the weights come as operand tables split once per parameter
// update (k_wtables).
#include "net_dev.h"

namespace ann {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

hipError_t species_perm(ann_ctx* ctx);

constexpr int kTile = 16;
constexpr int kMaxK = 128;     // padded input width supported (n_in <= 128)

struct MfmaArgs {
    const DevModel* dm;
    const float* W;
    const float* inpw;
    const float* inpb;
    int ldg;
    const int* perm;       // species-sorted atom order
    const int2* tiles;     // (species, start in perm); count from the species end
    const int* spend;      // end of each species block in perm
    int ntiles;
    const float* G;
    float* Eatom;
    float* dEdG;
    const uint4* wtab;     // BF16x3 operand tables of the hidden-layer weights
};

// acc[ct] += X (16 x K, rows = atoms, from LDS) * B (K x 16 columns of tile ct)
// for every column tile at once on v_mfma_f32_16x16x32_bf16.  BF16 carries 8
// mantissa bits, so both operands are split into three BF16 terms
// x = x0 + x1 + x2 (x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1)) and
// the six products with i + j <= 2 are accumulated in FP32: truncation
// ~2^-24 relative, the FP32 level the 1e-5 force bar needs.  The activation
// operand is split once per k block and reused by every column tile; the
// tiles advance together so one block's weight loads are in flight together.
// Lane l supplies row l % 16 of X and column l % 16 of B for k block
// 8 (l / 16) .. +7 and receives D rows 4 (l / 16) .. +3 of column l % 16.
__device__ __forceinline__ void split3(float x, __bf16& a, __bf16& b, __bf16& c) {
    a = (__bf16)x;
    const float r1 = x - (float)a;
    b = (__bf16)r1;
    c = (__bf16)(r1 - (float)b);
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 u) {
    bf16x8 v;
    __builtin_memcpy(&v, &u, sizeof(v));
    return v;
}

// The right operands come pre-split from a table (k_wtables): for k block kk
// and column tile ct, the three BF16 terms of lane l are the 16-byte entries
// ((kk * nt + ct) * 3 + p) * 64 + l, p = 0, 1, 2.
template <int NT>
__device__ __forceinline__ void mm_bf16x3(int K, int ncol, const float (*X)[kMaxK + 1], const uint4* __restrict__ tab,
                                          f32x4 (&acc)[NT], int wv) {
    const int lane = threadIdx.x % kWave;
    const int row = lane % 16, kb = 8 * (lane / 16);
    const int nt = (ncol + 15) / 16;
    for (int k0 = 0, kk = 0; k0 < K; k0 += 32, ++kk) {
        bf16x8 a0, a1, a2;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = k0 + kb + j;
            __bf16 x0, x1, x2;
            split3(k < K ? X[row][k] : 0.0f, x0, x1, x2);
            a0[j] = x0; a1[j] = x1; a2[j] = x2;
        }
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
            if (ct >= nt) break;
            if ((ct & 1) != wv) continue;           // column tiles alternate between the two wavefronts
            const uint4* e = tab + (size_t)((kk * nt + ct) * 3) * kWave + lane;
            const bf16x8 b0 = as_bf16x8(e[0]), b1 = as_bf16x8(e[kWave]), b2 = as_bf16x8(e[2 * kWave]);
            f32x4 c = acc[ct];
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b0, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b2, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, c, 0, 0, 0);
            acc[ct] = c;
        }
    }
}

// The weight tables, all read from the FP32 weights W and rewritten after
// every parameter update, in one launch (grid row = role and segment):
//  * matrix-core operand tables of every hidden layer, segment (species,
//    layer, direction): direction 0 holds W_l as the right operand of the
//    forward product (K = inputs, columns = neurons), direction 1 holds W_l^T
//    for the backward product (K = neurons, columns = inputs); one thread per
//    16-byte entry; tick (the Adam step counter) is advanced by one thread;
//  * the vector-ALU network's table of each species: the hidden layers' weight
//    rows, their biases and the output neuron's weights and bias, one thread
//    per float including the padding (written as zero);
//  * its transposed weight rows W_l^T[j][k], segment (species, layer).
struct WTabArgs {
    const DevModel* dm;
    const float* W;
    uint4* split;       // null: the operand tables are not rewritten
    float* chunk;
    float* trans;
    int* tick;
    int ys, yc;         // grid rows of the operand tables and of the species tables
};

__device__ __forceinline__ void wsplit_entry(const DevModel* __restrict__ dm, const float* __restrict__ W,
                                             uint4* __restrict__ tab, int seg, int e) {
    const int a = seg / (2 * kMaxLayers), l = (seg / 2) % kMaxLayers, dir = seg % 2;
    if (a >= dm->S) return;
    const NetDesc& D = dm->net[a];
    if (l >= D.nlayer - 1) return;
    const int ni = l == 0 ? D.nin : D.nodes[l - 1], no = D.nodes[l];
    const int K = dir == 0 ? ni : no, ncol = dir == 0 ? no : ni;
    const int nt = (ncol + 15) / 16;
    if (e >= (K + 31) / 32 * nt * 3 * kWave) return;
    const int lane = e % kWave, p = (e / kWave) % 3, ct = (e / (3 * kWave)) % nt, kk = e / (3 * kWave * nt);
    const int col = ct * 16 + lane % 16, kb = kk * 32 + 8 * (lane / 16);
    const float* Wl = W + D.woff[l];
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int k = kb + j;
        const float x = (k < K && col < ncol) ? (dir == 0 ? Wl[k * no + col] : Wl[col * no + k]) : 0.0f;
        __bf16 x0, x1, x2;
        split3(x, x0, x1, x2);
        v[j] = p == 0 ? x0 : (p == 1 ? x1 : x2);
    }
    uint4 u;
    __builtin_memcpy(&u, &v, sizeof(u));
    tab[(dir == 0 ? D.wsf[l] : D.wsb[l]) + e] = u;
}

__device__ __forceinline__ void wchunk_entry(const DevModel* __restrict__ dm, const float* __restrict__ W,
                                             float* __restrict__ tab, int a, int x) {
    if (a >= dm->S) return;
    const NetDesc& D = dm->net[a];
    const int e = D.wc0 + x;
    if (e >= D.wc0 + D.wcn) return;
    const int L = D.nlayer;
    float v = 0.0f;
    if (e >= D.wco) {                                       // output neuron
        const int nlast = L > 1 ? D.nodes[L - 2] : D.nin, m = e - D.wco;
        if (m < nlast) v = W[D.woff[L - 1] + m];
        else if (m == vec_pad4(nlast)) v = W[D.boff[L - 1]];
    } else if (L > 1 && e >= D.wcb[0]) {                    // hidden-layer biases
        int l = 0;
        while (l + 1 < L - 1 && e >= D.wcb[l + 1]) ++l;
        const int j = e - D.wcb[l];
        if (j < D.nodes[l]) v = W[D.boff[l] + j];
    } else {                                                // weight rows
        int l = 0;
        while (l + 1 < L - 1 && e >= D.wcf[l + 1]) ++l;
        const int ni = l == 0 ? D.nin : D.nodes[l - 1], no = D.nodes[l];
        const int r = e - D.wcf[l], k = r / D.wcs[l], j = r % D.wcs[l];
        if (k < ni && j < no) v = W[D.woff[l] + k * no + j];
    }
    tab[e] = v;
}

__device__ __forceinline__ void wtrans_entry(const DevModel* __restrict__ dm, const float* __restrict__ W,
                                             float* __restrict__ tab, int seg, int e) {
    const int a = seg / kMaxLayers, l = seg % kMaxLayers;
    if (a >= dm->S) return;
    const NetDesc& D = dm->net[a];
    if (l >= D.nlayer - 1) return;
    const int ni = l == 0 ? D.nin : D.nodes[l - 1], no = D.nodes[l], ps = D.wts[l];
    if (e >= vec_pad4(no * ps)) return;
    const int j = e / ps, k = e % ps;
    tab[D.wtf[l] + e] = (j < no && k < ni) ? W[D.woff[l] + k * no + j] : 0.0f;
}

__global__ __launch_bounds__(256) void k_wtables(WTabArgs A) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (A.tick && y == 0 && e == 0) *A.tick += 1;
    if (y < A.ys) { wsplit_entry(A.dm, A.W, A.split, y, e); return; }
    y -= A.ys;
    if (y < A.yc) { wchunk_entry(A.dm, A.W, A.chunk, y, e); return; }
    wtrans_entry(A.dm, A.W, A.trans, y - A.yc, e);
}

// All weight tables after a change of the FP32 weights (split: the matrix-core
// operand tables too); tick as in k_wtables.
hipError_t weight_tables(ann_ctx* ctx, int* tick, bool split) {
    const DevModel& M = ctx->hm;
    const int S = std::max(M.S, 1);
    int most = std::max(ctx->wchunk_max, 1);
    for (int a = 0; a < M.S; ++a)
        for (int l = 0; l < M.net[a].nlayer - 1; ++l) {
            const int ni = l == 0 ? M.net[a].nin : M.net[a].nodes[l - 1], no = M.net[a].nodes[l];
            if (split) {
                most = std::max(most, (ni + 31) / 32 * ((no + 15) / 16) * 3 * kWave);
                most = std::max(most, (no + 31) / 32 * ((ni + 15) / 16) * 3 * kWave);
            }
            most = std::max(most, vec_pad4(no * M.net[a].wts[l]));
        }
    WTabArgs A{ctx->dm, ctx->da.W, split ? ctx->d_wsplit : nullptr, ctx->d_wchunk, ctx->d_wtrans, tick,
               split ? S * 2 * kMaxLayers : 0, S};
    dim3 grid((most + 255) / 256, A.ys + A.yc + S * kMaxLayers);
    k_wtables<<<grid, 256, 0, ctx->stream>>>(A);
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        if (split) ctx->wsplit_valid = true;
        ctx->wchunk_valid = true;
    }
    return e;
}

hipError_t split_weights(ann_ctx* ctx, int* tick) {
    // the vector-ALU network's tables follow every change of the weights (a
    // graph that captures a parameter update rebuilds them on every replay)
    return weight_tables(ctx, tick, true);
}

// two wavefronts per 16-atom single-species tile: the column tiles of every
// product alternate between them (the launch has fewer tiles than SIMDs).
// NT: column tiles of 16 of the widest hidden layer (4 up to 64 neurons, 8 up
// to 128), so narrow networks keep their registers and LDS.
template <int NT>
__global__ __launch_bounds__(128) void k_net_mfma(MfmaArgs A) {
    constexpr int W = 16 * NT;                  // widest hidden layer of this instance
    __shared__ float act[kTile][kMaxK + 1];     // activations / gradients, operand side
    __shared__ float sdz[kMaxLayers - 1][kTile][W + 1];
    __shared__ int atoms[kTile];
    const int t = blockIdx.x, tid = threadIdx.x, lane = tid % kWave, wv = tid / kWave;
    constexpr int kThreads = 2 * kWave;
    if (t >= A.ntiles) return;
    const int2 tl = A.tiles[t];
    const int a = tl.x;
    const int nrow = min(kTile, A.spend[a] - tl.y);
    const NetDesc& D = A.dm->net[a];
    const int L = D.nlayer;
    const float* iw = A.inpw + D.inoff;
    const float* ib = A.inpb + D.inoff;
    if (tid < kTile) atoms[tid] = tid < nrow ? A.perm[tl.y + tid] : -1;
    __syncthreads();
    // scaled input s = inpw (G + inpb) into LDS: lanes over the inputs, rows unrolled
    {
        float w2[kMaxK / kWave], b2[kMaxK / kWave];
#pragma unroll
        for (int h = 0; h < kMaxK / kWave; ++h) {
            const int k = lane + h * kWave;
            w2[h] = k < D.nin ? iw[k] : 0.0f;
            b2[h] = k < D.nin ? ib[k] : 0.0f;
        }
#pragma unroll
        for (int r = wv; r < kTile; r += 2) {
            const int ai = atoms[r];
#pragma unroll
            for (int h = 0; h < kMaxK / kWave; ++h) {
                const int k = lane + h * kWave;
                act[r][k] = (ai >= 0 && k < D.nin) ? w2[h] * (A.G[(size_t)ai * A.ldg + k] + b2[h]) : 0.0f;
            }
        }
    }
    __syncthreads();
    const int col = lane % 16, rb = 4 * (lane / 16);
    int ni = D.nin;
    for (int l = 0; l < L - 1; ++l) {
        const int no = D.nodes[l];
        const float* bl = A.W + D.boff[l];
        float out[NT][4];
        f32x4 accs[NT];
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) accs[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
        mm_bf16x3<NT>(ni, no, act, A.wtab + D.wsf[l], accs, wv);
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
            if (ct * 16 >= no) break;
            if ((ct & 1) != wv) continue;
            const f32x4 acc = accs[ct];
            const int n = ct * 16 + col;
            const float bn = n < no ? bl[n] : 0.0f;
            float z[4], av[4], dv[4], d2[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) z[i] = acc[i] + bn;
            act_many<false, 4>(D.neuron, D.c, z, av, dv, d2);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                out[ct][i] = n < no ? av[i] : 0.0f;
                sdz[l][rb + i][n < W ? n : W] = n < no ? dv[i] : 0.0f;
            }
        }
        __syncthreads();
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
            if (ct * 16 >= no) break;
            if ((ct & 1) != wv) continue;
            const int n = ct * 16 + col;
#pragma unroll
            for (int i = 0; i < 4; ++i) act[rb + i][n] = n < no ? out[ct][i] : 0.0f;
        }
        // zero the padding up to the next multiple of 32
        for (int q = tid; q < kTile * 32; q += kThreads) {
            const int r = q / 32, k = no + q % 32;
            if (k < kMaxK) act[r][k] = 0.0f;
        }
        __syncthreads();
        ni = no;
    }
    // linear output layer (1 neuron) on the vector ALU
    const float* WL = A.W + D.woff[L - 1];
    const float bL = A.W[D.boff[L - 1]];
    if (tid < nrow) {
        float z = bL;
        for (int m = 0; m < ni; ++m) z += WL[m] * act[tid][m];
        A.Eatom[atoms[tid]] = D.outb + D.outw * z;
    }
    if (!A.dEdG) return;
    if (L == 1) {
        for (int q = tid; q < kTile * A.ldg; q += kThreads) {
            const int r = q / A.ldg, m = q % A.ldg;
            if (r < nrow) A.dEdG[(size_t)atoms[r] * A.ldg + m] = m < D.nin ? D.outw * WL[m] * iw[m] : 0.0f;
        }
        return;
    }
    __syncthreads();
    // g_{L-1} = outw W_L . sigma'_{L-1}  (operand layout, rows = atoms)
    for (int q = tid; q < kTile * kMaxK; q += kThreads) {
        const int r = q / kMaxK, m = q % kMaxK;
        act[r][m] = m < ni ? D.outw * WL[m] * sdz[L - 2][r][m] : 0.0f;
    }
    __syncthreads();
    for (int l = L - 2; l >= 0; --l) {
        const int no = D.nodes[l];
        const int nprev = l == 0 ? D.nin : D.nodes[l - 1];
        float out[kMaxK / 16][4];
        f32x4 accs[kMaxK / 16];
#pragma unroll
        for (int ct = 0; ct < kMaxK / 16; ++ct) accs[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
        mm_bf16x3<kMaxK / 16>(no, nprev, act, A.wtab + D.wsb[l], accs, wv);
#pragma unroll
        for (int ct = 0; ct < kMaxK / 16; ++ct) {
            if (ct * 16 >= nprev) break;
            if ((ct & 1) != wv) continue;
            const f32x4 acc = accs[ct];
            const int m = ct * 16 + col;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[i];
                if (l > 0) v *= (m < nprev) ? sdz[l - 1][rb + i][m] : 0.0f;
                out[ct][i] = v;
            }
        }
        __syncthreads();
        if (l > 0) {
#pragma unroll
            for (int ct = 0; ct < kMaxK / 16; ++ct) {
                if (ct * 16 >= nprev) break;
                if ((ct & 1) != wv) continue;
                const int m = ct * 16 + col;
#pragma unroll
                for (int i = 0; i < 4; ++i) act[rb + i][m] = m < nprev ? out[ct][i] : 0.0f;
            }
            for (int q = tid; q < kTile * 32; q += kThreads) {
                const int r = q / 32, k = nprev + q % 32;
                if (k < kMaxK) act[r][k] = 0.0f;
            }
        } else {
#pragma unroll
            for (int ct = 0; ct < kMaxK / 16; ++ct) {
                if (ct * 16 >= A.ldg) break;
                if ((ct & 1) != wv) continue;
                const int m = ct * 16 + col;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = rb + i;
                    if (r < nrow && m < A.ldg)
                        A.dEdG[(size_t)atoms[r] * A.ldg + m] = (m < D.nin && ct * 16 < nprev) ? out[ct][i] * iw[m] : 0.0f;
                }
            }
        }
        __syncthreads();
    }
}

hipError_t net_fwd_mfma(ann_ctx* ctx, const float* G, float* Eatom, float* dEdG, const float* inpb) {
    hipError_t e = species_perm(ctx);
    if (e != hipSuccess) return e;
    MfmaArgs A{};
    A.dm = ctx->dm;
    A.W = ctx->da.W;
    A.inpw = ctx->da.inpw;
    A.inpb = inpb;
    A.ldg = ctx->ldg;
    A.perm = ctx->d_perm;
    A.tiles = reinterpret_cast<const int2*>(ctx->d_tiles);
    A.spend = ctx->d_spend_end;
    A.ntiles = ctx->ntiles;
    A.G = G;
    A.Eatom = Eatom;
    A.dEdG = dEdG;
    A.wtab = ctx->d_wsplit;
    if (A.ntiles == 0) return hipSuccess;
    if (!ctx->wsplit_valid && (e = split_weights(ctx, nullptr)) != hipSuccess) return e;
    int wmax = 1;
    for (int a = 0; a < ctx->hm.S; ++a)
        for (int l = 0; l < ctx->hm.net[a].nlayer - 1; ++l) wmax = std::max(wmax, ctx->hm.net[a].nodes[l]);
    Launch L(ctx, "net_mfma");
    if (wmax <= 64) k_net_mfma<4><<<A.ntiles, 128, 0, ctx->stream>>>(A);
    else k_net_mfma<8><<<A.ntiles, 128, 0, ctx->stream>>>(A);
    return hipGetLastError();
}

bool mfma_net_supported(const DevModel& M) {
    for (int a = 0; a < M.S; ++a) {
        if (M.net[a].nin > kMaxK) return false;
        for (int l = 0; l < M.net[a].nlayer; ++l)
            if (M.net[a].nodes[l] > kMaxWidth) return false;
    }
    return true;
}

}  // namespace ann
