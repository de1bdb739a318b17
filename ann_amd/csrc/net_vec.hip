// net_vec.hip -- per-species atomic networks on the vector ALU, FP32.
//
// Same mathematics as ANN::fpbp + DODZ::grad (src/ml/nn.cpp:1367-1389,
// 1801-1830): s = inpw (G + inpb), z_l = b_l + W_l a_{l-1}, a_l = sigma(z_l),
// E = outb + outw (b_L + w_L . a_{L-1}); backward
// g_{L-1} = outw w_L . sigma'_{L-1}, g_{l-1} = (W_l^T g_l) . sigma'_{l-1},
// dE/dG = inpw . (W_0^T g_0).
//
// One wavefront per 16-atom single-species tile of the species-sorted order.
// A layer's outputs go to the lanes: lane (h, j) computes output j for the
// rows h*RP .. h*RP+RP-1 of the tile, with 64 / RP * 16 lanes per row group
// (RP = 16, 8 or 4 rows for up to 64, 32 or 16 outputs).  Each step of the
// reduction reads one weight per lane (a row of W in the forward pass, a
// column in the backward pass: the odd row stride keeps both conflict-free)
// and the RP activations of its rows as broadcast 16-byte LDS reads, and does
// RP fused multiply-adds.  The species' parameters are copied into the
// wavefront's LDS once per tile from tables built once per parameter set
// (layout in net_dev.h); activations and gradients are [neuron][16] rows in
// LDS; no block barriers.
#include "net_dev.h"

namespace ann {

hipError_t species_perm(ann_ctx* ctx);

constexpr int kVecRows = 16;   // atoms per tile (the species tiles of species_perm)

struct VecArgs {
    const DevModel* dm;
    const float* wchunk;
    const float* inpw;
    const float* inpb;
    int ldg;
    const int* perm;
    const int2* tiles;
    const int* spend;
    int ntiles;
    int wmax, maxin, maxw;      // LDS extents: table floats, input rows, hidden width,
    int maxl1;                  // hidden layers
    const float* G;
    float* Eatom;
    float* dEdG;
};

// The species tables (weight rows with their biases and output neuron, and
// the transposed weight rows) are written by k_wtables (net_mfma.hip) beside
// the matrix-core operand tables.
hipError_t chunk_weights(ann_ctx* ctx) { return weight_tables(ctx, nullptr, false); }

// LDS written by one lane of the wavefront is visible to the others after this
__device__ __forceinline__ void vsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// acc[i] += sum_k A[k * sa + o] * B[k * 16 + rows i], k < n: one lane's output o
// (lane-varying A read) over its RP rows (broadcast B reads)
template <int RP>
__device__ __forceinline__ void vec_reduce(const float* __restrict__ A, int sa, int o, const float* __restrict__ B,
                                           int rows0, int n, float (&acc)[RP]) {
    const float* a = A + o;
    const float4* b = reinterpret_cast<const float4*>(B + rows0);
#pragma unroll 4
    for (int k = 0; k < n; ++k) {
        const float w = a[k * sa];
        float x[RP];
#pragma unroll
        for (int q = 0; q < RP / 4; ++q) {
            const float4 v = b[k * (kVecRows / 4) + q];
            x[4 * q] = v.x; x[4 * q + 1] = v.y; x[4 * q + 2] = v.z; x[4 * q + 3] = v.w;
        }
#pragma unroll
        for (int q = 0; q < RP; ++q) acc[q] += w * x[q];
    }
}

// one hidden layer, RP rows per lane: H[j][r] = sigma(z), Ds[j][r] = sigma'(z);
// with last, also the gradient seed Dc[j][r] = outw w_L[j] sigma'(z)
template <int RP>
__device__ __forceinline__ void vec_forward(const NetDesc& D, const float* Wl, int ps, const float* bl, int ni, int no,
                                            const float* Xin, float* Hout, float* Ds, bool last, const float* WL,
                                            float* Dc) {
    constexpr int GW = kWave * RP / kVecRows;              // lanes per row group
    const int lane = threadIdx.x % kWave, j = lane % GW, h = lane / GW;
    const int jj = min(j, no - 1);
    float acc[RP];
#pragma unroll
    for (int q = 0; q < RP; ++q) acc[q] = bl[jj];
    vec_reduce<RP>(Wl, ps, jj, Xin, h * RP, ni, acc);
    float av[RP], dv[RP], d2[RP];
    act_many<false, RP>(D.neuron, D.c, acc, av, dv, d2);
    if (j < no) {
#pragma unroll
        for (int q = 0; q < RP / 4; ++q) {
            reinterpret_cast<float4*>(Hout + j * kVecRows + h * RP)[q] =
                make_float4(av[4 * q], av[4 * q + 1], av[4 * q + 2], av[4 * q + 3]);
            reinterpret_cast<float4*>(Ds + j * kVecRows + h * RP)[q] =
                make_float4(dv[4 * q], dv[4 * q + 1], dv[4 * q + 2], dv[4 * q + 3]);
        }
    }
    if (last && j < no) {
        {
            const float g = D.outw * WL[j];
#pragma unroll
            for (int q = 0; q < RP / 4; ++q)
                reinterpret_cast<float4*>(Dc + j * kVecRows + h * RP)[q] =
                    make_float4(g * dv[4 * q], g * dv[4 * q + 1], g * dv[4 * q + 2], g * dv[4 * q + 3]);
        }
    }
}

// one backward step, lanes over the layer's inputs k0 + k (k < nk), RP rows
// per lane: acc[r] = sum_j W[k][j] g[j][r]; then Dn[k][r] = acc * sigma'_{l-1}
// (inner layer) or dE/dG[atom][k] = acc * inpw[k] (input layer)
template <int RP>
__device__ __forceinline__ void vec_backward(const float* Wl, int ps, int no, int k0, int nk, const float* Dc,
                                             float* Dn, const float* Dsp, const float* IW, const int* rows,
                                             float* dEdG, int ldg) {
    constexpr int GW = kWave * RP / kVecRows;
    const int lane = threadIdx.x % kWave, k = k0 + lane % GW, h = lane / GW;
    const int kk = min(k, k0 + nk - 1);
    float acc[RP];
#pragma unroll
    for (int q = 0; q < RP; ++q) acc[q] = 0.0f;
    vec_reduce<RP>(Wl + (size_t)kk * ps, 1, 0, Dc, h * RP, no, acc);
    if (lane % GW >= nk) return;
    if (Dn) {
#pragma unroll
        for (int q = 0; q < RP / 4; ++q) {
            const float4 d = reinterpret_cast<const float4*>(Dsp + k * kVecRows + h * RP)[q];
            reinterpret_cast<float4*>(Dn + k * kVecRows + h * RP)[q] =
                make_float4(acc[4 * q] * d.x, acc[4 * q + 1] * d.y, acc[4 * q + 2] * d.z, acc[4 * q + 3] * d.w);
        }
    } else {
        const float w = IW[k];
#pragma unroll
        for (int q = 0; q < RP; ++q) {
            const int ar = rows[h * RP + q];
            if (ar >= 0) dEdG[(size_t)ar * ldg + k] = acc[q] * w;
        }
    }
}

__global__ __launch_bounds__(64) void k_net_vec(VecArgs A) {
    extern __shared__ float4 vsm4[];
    __shared__ int rows[kVecRows];                         // atom of each tile row (-1 past the species)
    float* const sm = reinterpret_cast<float*>(vsm4);
    const int t = blockIdx.x, lane = threadIdx.x % kWave;
    if (t >= A.ntiles) return;
    const int2 tl = A.tiles[t];
    const int a = tl.x;
    const int nrow = min(kVecRows, A.spend[a] - tl.y);
    const NetDesc& D = A.dm->net[a];
    const int L = D.nlayer, nin = D.nin;
    float* Wt = sm;                                        // [wmax] this species' tables
    float* X = Wt + A.wmax;                                // [maxin][16] scaled inputs
    float* H0 = X + A.maxin * kVecRows;                    // [maxw][16] activations / gradients
    float* H1 = H0 + A.maxw * kVecRows;
    float* Ds = H1 + A.maxw * kVecRows;                    // [L-1][maxw][16] sigma'
    float* IW = Ds + A.maxl1 * A.maxw * kVecRows;          // [maxin] input scaling
    if (lane < kVecRows) rows[lane] = lane < nrow ? A.perm[tl.y + lane] : -1;
    const float* iw = A.inpw + D.inoff;
    const float* ib = A.inpb + D.inoff;
    // scaled inputs X[k][row] (lanes over the inputs, every row's loads issued
    // before the table copy) and the species' tables, a contiguous copy in
    // batches of eight 16-byte loads per lane
    {
        float v[2][kVecRows];
        const float w0 = lane < nin ? iw[lane] : 0.0f, b0 = lane < nin ? ib[lane] : 0.0f;
        const float w1 = lane + kWave < nin ? iw[lane + kWave] : 0.0f, b1 = lane + kWave < nin ? ib[lane + kWave] : 0.0f;
        vsync();
#pragma unroll
        for (int rr = 0; rr < kVecRows; ++rr) {
            const int ar = rows[rr];
            const float* g = A.G + (size_t)(ar >= 0 ? ar : 0) * A.ldg;
            v[0][rr] = (ar >= 0 && lane < nin) ? w0 * (g[lane] + b0) : 0.0f;
            v[1][rr] = (ar >= 0 && lane + kWave < nin) ? w1 * (g[lane + kWave] + b1) : 0.0f;
        }
        const float4* src = reinterpret_cast<const float4*>(A.wchunk + D.wc0);
        float4* dst = reinterpret_cast<float4*>(Wt);
        const int n4 = D.wcn / 4;
        for (int q0 = 0; q0 < n4; q0 += 8 * kWave) {
            float4 tmp[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int q = q0 + u * kWave + lane;
                tmp[u] = q < n4 ? src[q] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int q = q0 + u * kWave + lane;
                if (q < n4) dst[q] = tmp[u];
            }
        }
        if (lane < nin) IW[lane] = w0;
        if (lane + kWave < nin) IW[lane + kWave] = w1;
#pragma unroll
        for (int rr = 0; rr < kVecRows; ++rr) {
            if (lane < nin) X[lane * kVecRows + rr] = v[0][rr];
            if (lane + kWave < nin) X[(lane + kWave) * kVecRows + rr] = v[1][rr];
        }
    }
    vsync();
    const float* WL = Wt + (D.wco - D.wc0);                // output weights, then (at a multiple of 4) the bias
    // forward
    const float* Xin = X;
    float* Hout = H0;
    float* Dc = H1;                                        // gradient seed of the last hidden layer
    int ni = nin;
#pragma unroll
    for (int l = 0; l < kMaxLayers - 1; ++l) {
        if (l >= L - 1) break;
        const int no = D.nodes[l], ps = D.wcs[l];
        const float* Wl = Wt + (D.wcf[l] - D.wc0);
        const float* bl = Wt + (D.wcb[l] - D.wc0);
        const bool last = l == L - 2;
        float* dc = last ? (Hout == H0 ? H1 : H0) : nullptr;
        float* dsl = Ds + l * A.maxw * kVecRows;
        if (no <= 16) vec_forward<4>(D, Wl, ps, bl, ni, no, Xin, Hout, dsl, last, WL, dc);
        else if (no <= 32) vec_forward<8>(D, Wl, ps, bl, ni, no, Xin, Hout, dsl, last, WL, dc);
        else vec_forward<16>(D, Wl, ps, bl, ni, no, Xin, Hout, dsl, last, WL, dc);
        if (last) Dc = dc;
        vsync();
        Xin = Hout;
        Hout = Hout == H0 ? H1 : H0;
        ni = no;
    }
    // output neuron: E[r] = outb + outw (b_L + sum_j w_L[j] a[j][r]); lane
    // (s, r) sums j = s, s + 4, ... of row r from LDS, two butterflies join them
    {
        const int r = lane % kVecRows, s4 = lane / kVecRows;
        float part = 0.0f;
        for (int m = s4; m < ni; m += 4) part += WL[m] * Xin[m * kVecRows + r];
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        const int ar = rows[r];
        if (s4 == 0 && ar >= 0) A.Eatom[ar] = D.outb + D.outw * (WL[vec_pad4(ni)] + part);
    }
    if (!A.dEdG) return;
    if (L == 1) {
        for (int q = lane; q < kVecRows * A.ldg; q += kWave) {
            const int rr = q / A.ldg, m = q % A.ldg;
            const int ar = rows[rr];
            if (ar >= 0) A.dEdG[(size_t)ar * A.ldg + m] = m < nin ? D.outw * WL[m] * IW[m] : 0.0f;
        }
        return;
    }
    // backward: g_l in Dc, the next gradient into the other activation buffer
    float* Dn = Dc == H0 ? H1 : H0;
#pragma unroll
    for (int l = kMaxLayers - 2; l >= 0; --l) {
        if (l > L - 2) continue;
        const int no = D.nodes[l], nil = l == 0 ? nin : D.nodes[l - 1], ps = D.wcs[l];
        const float* Wl = Wt + (D.wcf[l] - D.wc0);
        float* dn = l > 0 ? Dn : nullptr;
        const float* dsp = l > 0 ? Ds + (l - 1) * A.maxw * kVecRows : nullptr;
        for (int k0 = 0; k0 < nil; k0 += kWave) {
            const int nk = min(kWave, nil - k0);
            if (nk <= 16) vec_backward<4>(Wl, ps, no, k0, nk, Dc, dn, dsp, IW, rows, A.dEdG, A.ldg);
            else if (nk <= 32) vec_backward<8>(Wl, ps, no, k0, nk, Dc, dn, dsp, IW, rows, A.dEdG, A.ldg);
            else vec_backward<16>(Wl, ps, no, k0, nk, Dc, dn, dsp, IW, rows, A.dEdG, A.ldg);
        }
        vsync();
        float* tmp = Dc;
        Dc = Dn;
        Dn = tmp;
    }
    // padding of the dE/dG rows past this species' inputs
    for (int q = lane; q < kVecRows * (A.ldg - nin); q += kWave) {
        const int rr = q / (A.ldg - nin), m = nin + q % (A.ldg - nin);
        const int ar = rows[rr];
        if (ar >= 0) A.dEdG[(size_t)ar * A.ldg + m] = 0.0f;
    }
}

// LDS extents of the kernel for the context's networks; ok is false when a
// network does not fit the kernel's shapes (more than 128 inputs or 64 neurons)
struct VecExtents {
    int maxin = 1, maxw = 1, maxl1 = 1;
    bool ok = true;
};

static VecExtents vec_extents(const DevModel& M) {
    VecExtents x;
    for (int a = 0; a < M.S; ++a) {
        const NetDesc& D = M.net[a];
        if (D.nin > 2 * kWave) x.ok = false;
        x.maxin = std::max(x.maxin, D.nin);
        x.maxl1 = std::max(x.maxl1, D.nlayer - 1);
        for (int l = 0; l < D.nlayer; ++l) {
            if (D.nodes[l] > kVecMaxWidth) x.ok = false;
            if (l < D.nlayer - 1) x.maxw = std::max(x.maxw, D.nodes[l]);
        }
    }
    return x;
}

// LDS bytes of one wavefront (tables, inputs, two activation buffers, sigma'
// of every hidden layer, input scaling), or 0 when the networks do not fit
size_t vec_net_smem(const ann_ctx* ctx) {
    const VecExtents x = vec_extents(ctx->hm);
    if (!x.ok) return 0;
    const size_t f = (size_t)ctx->wchunk_max + (size_t)x.maxin * kVecRows + 2 * (size_t)x.maxw * kVecRows +
                     (size_t)x.maxl1 * x.maxw * kVecRows + (size_t)x.maxin;
    return 4 * f;
}

// the kernel is used when one wavefront's LDS leaves room for two per CU (the
// tiles fill the chip at two or more resident wavefronts per CU)
constexpr size_t kVecSmemMax = 64 * 1024;

bool vec_net_supported(const ann_ctx* ctx) {
    const size_t b = vec_net_smem(ctx);
    return b > 0 && b <= kVecSmemMax;
}

hipError_t net_fwd_vec(ann_ctx* ctx, const float* G, float* Eatom, float* dEdG, const float* inpb) {
    hipError_t e = species_perm(ctx);
    if (e != hipSuccess) return e;
    if (ctx->ntiles == 0) return hipSuccess;
    if (!ctx->wchunk_valid && (e = chunk_weights(ctx)) != hipSuccess) return e;
    const DevModel& M = ctx->hm;
    VecArgs A{};
    A.dm = ctx->dm;
    A.wchunk = ctx->d_wchunk;
    A.inpw = ctx->da.inpw;
    A.inpb = inpb;
    A.ldg = ctx->ldg;
    A.perm = ctx->d_perm;
    A.tiles = reinterpret_cast<const int2*>(ctx->d_tiles);
    A.spend = ctx->d_spend_end;
    A.ntiles = ctx->ntiles;
    const VecExtents x = vec_extents(M);
    A.wmax = ctx->wchunk_max;
    A.maxin = x.maxin;
    A.maxw = x.maxw;
    A.maxl1 = x.maxl1;
    A.G = G;
    A.Eatom = Eatom;
    A.dEdG = dEdG;
    const size_t sm = vec_net_smem(ctx);
    Launch L(ctx, "net_vec");
    k_net_vec<<<A.ntiles, kWave, sm, ctx->stream>>>(A);
    return hipGetLastError();
}

}  // namespace ann
