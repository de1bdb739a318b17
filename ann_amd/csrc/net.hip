// net.hip -- per-species atomic networks.
//
// ANN::fpbp (src/ml/nn.cpp:1367-1389): s = inpw (G + inpb), z_l = b_l + W_l a_{l-1},
// a_l = sigma(z_l), linear output, E = outb + outw a_L; DODZ::grad (:1801-1830):
// dE/dG = inpw . W_0^T diag(sigma'_0) ... W_L^T outw.
//
// PLAIN form: FP32 GEMV chain per atom; 16 lanes cooperate on one atom
// (lane q owns neurons q, q+16, ...), four atoms per wavefront, activations
// shared through LDS.
// The FAST form (BF16 MFMA with split hi+lo products) lives in net_mfma.hip.
#include "net_dev.h"

namespace ann {

struct NetArgs {
    const DevModel* dm;
    const float* W;
    const float* inpw;
    const float* inpb;
    int N, ldg;
    const int* species;
    const float* G;
    float* Eatom;
    float* dEdG;
};

// 4 atoms per wavefront, 16 lanes per atom; lane q owns neurons q, q+16, ...
constexpr int kNetLanes = 16;
constexpr int kNetOut = kMaxWidth / kNetLanes;        // neurons per lane in a hidden layer
constexpr int kNetIn = 128 / kNetLanes;               // inputs per lane for dE/dG (n_in <= 128)

__global__ __launch_bounds__(64) void k_net_plain(NetArgs A) {
    __shared__ float xs[4][kMaxWidth + 1];
    const int sub = threadIdx.x / kNetLanes, q = threadIdx.x % kNetLanes;
    const int i = blockIdx.x * 4 + sub;
    const bool valid = i < A.N;
    const int a = valid ? A.species[i] : 0;
    const NetDesc& D = A.dm->net[a];
    const int L = D.nlayer;
    const float c = D.c;
    const float* iw = A.inpw + D.inoff;
    const float* ib = A.inpb + D.inoff;
    const float* G = A.G + (size_t)(valid ? i : 0) * A.ldg;
    float z[kNetOut];
    float dz[kMaxLayers - 1][kNetOut];
    int ni = D.nin;
    for (int l = 0; l < L; ++l) {
        const int no = D.nodes[l];
        const float* Wl = A.W + D.woff[l];
        const float* bl = A.W + D.boff[l];
        const bool last = l == L - 1;
        if (!last) {
#pragma unroll
            for (int j = 0; j < kNetOut; ++j) {
                const int n = q + kNetLanes * j;
                z[j] = n < no ? bl[n] : 0.0f;
            }
            for (int m = 0; m < ni; ++m) {
                const float x = l == 0 ? iw[m] * (G[m] + ib[m]) : xs[sub][m];
                const float* col = Wl + m * no;
#pragma unroll
                for (int j = 0; j < kNetOut; ++j) {
                    const int n = q + kNetLanes * j;
                    if (n < no) z[j] += col[n] * x;
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kNetOut; ++j) {
                const int n = q + kNetLanes * j;
                float av = 0.0f, dv = 0.0f, d2;
                if (n < no) act_f<false>(D.neuron, c, z[j], av, dv, d2);
                if (n < no) xs[sub][n] = av;
#pragma unroll
                for (int p = 0; p < kMaxLayers - 1; ++p)
                    if (p == l) dz[p][j] = dv;
            }
            __syncthreads();
        } else {
            // linear output neuron: lanes split the inputs, 16-lane reduction
            float part = 0.0f;
            for (int m = q; m < ni; m += kNetLanes) {
                const float x = l == 0 ? iw[m] * (G[m] + ib[m]) : xs[sub][m];
                part += Wl[m] * x;
            }
#pragma unroll
            for (int o = kNetLanes / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, kNetLanes);
            if (valid && q == 0) A.Eatom[i] = D.outb + D.outw * (bl[0] + part);
        }
        ni = no;
    }
    if (!A.dEdG) return;
    __syncthreads();   // the output neuron's reads of xs precede the overwrite below
    // backward: g over the last hidden layer, then through W_l^T . sigma'
    const float* WL = A.W + D.woff[L - 1];
    if (L == 1) {
        for (int m = q; m < A.ldg; m += kNetLanes)
            if (valid) A.dEdG[(size_t)i * A.ldg + m] = m < D.nin ? D.outw * WL[m] * iw[m] : 0.0f;
        return;
    }
    {
        const int nl = D.nodes[L - 2];
#pragma unroll
        for (int j = 0; j < kNetOut; ++j) {
            const int m = q + kNetLanes * j;
            float d = 0.0f;
#pragma unroll
            for (int p = 0; p < kMaxLayers - 1; ++p)
                if (p == L - 2) d = dz[p][j];
            if (m < nl) xs[sub][m] = D.outw * WL[m] * d;
        }
        __syncthreads();
    }
    for (int l = L - 2; l >= 1; --l) {
        const int no = D.nodes[l], np = D.nodes[l - 1];
        const float* Wl = A.W + D.woff[l];
        float h[kNetOut];
#pragma unroll
        for (int j = 0; j < kNetOut; ++j) {
            const int m = q + kNetLanes * j;
            float s = 0.0f;
            if (m < np) {
                const float* col = Wl + m * no;
                for (int n = 0; n < no; ++n) s += col[n] * xs[sub][n];
            }
            float d = 0.0f;
#pragma unroll
            for (int p = 0; p < kMaxLayers - 1; ++p)
                if (p == l - 1) d = dz[p][j];
            h[j] = s * d;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kNetOut; ++j) {
            const int m = q + kNetLanes * j;
            if (m < np) xs[sub][m] = h[j];
        }
        __syncthreads();
    }
    {
        const int no = D.nodes[0];
        const float* W0 = A.W + D.woff[0];
        for (int m = q; m < A.ldg; m += kNetLanes) {
            float s = 0.0f;
            if (m < D.nin) {
                const float* col = W0 + m * no;
                for (int n = 0; n < no; ++n) s += col[n] * xs[sub][n];
                s *= iw[m];
            }
            if (valid) A.dEdG[(size_t)i * A.ldg + m] = s;
        }
    }
}

int max_width(const DevModel& M) {
    int w = 1;
    for (int a = 0; a < M.S; ++a)
        for (int l = 0; l < M.net[a].nlayer; ++l) w = std::max(w, M.net[a].nodes[l]);
    return w;
}

hipError_t net_fwd_plain(ann_ctx* ctx, const float* G, float* Eatom, float* dEdG, const float* inpb) {
    NetArgs A{ctx->dm, ctx->da.W, ctx->da.inpw, inpb, ctx->natoms, ctx->ldg, ctx->d_species,
              G, Eatom, dEdG};
    if (A.N == 0) return hipSuccess;
    Launch L(ctx, "net_plain");
    k_net_plain<<<(A.N + 3) / 4, 64, 0, ctx->stream>>>(A);
    return hipGetLastError();
}

hipError_t net_fwd_mfma(ann_ctx* ctx, const float* G, float* Eatom, float* dEdG, const float* inpb);
bool mfma_net_supported(const DevModel& M);
hipError_t net_fwd_vec(ann_ctx* ctx, const float* G, float* Eatom, float* dEdG, const float* inpb);
bool vec_net_supported(const ann_ctx* ctx);

// TRIPLET and FAST: the vector-ALU tile kernel where the species' weights fit
// its LDS budget (unless the tuning asks for the matrix cores), else the
// matrix-core kernel; PLAIN: one GEMV chain per atom
hipError_t net_fwd(ann_ctx* ctx, int form, const float* G, float* Eatom, float* dEdG, bool centred) {
    const float* inpb = centred ? ctx->da.inpb_net : ctx->da.inpb;
    if (form != ANN_FORM_PLAIN && ctx->tune_net == 0 && vec_net_supported(ctx))
        return net_fwd_vec(ctx, G, Eatom, dEdG, inpb);
    if (form != ANN_FORM_PLAIN && mfma_net_supported(ctx->hm)) return net_fwd_mfma(ctx, G, Eatom, dEdG, inpb);
    return net_fwd_plain(ctx, G, Eatom, dEdG, inpb);
}

}  // namespace ann
