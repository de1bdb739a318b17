// net.hip -- per-species atomic networks.
//
// ANN::fpbp (src/ml/nn.cpp:1367-1389): s = inpw (G + inpb), z_l = b_l + W_l a_{l-1},
// a_l = sigma(z_l), linear output, E = outb + outw a_L; DODZ::grad (:1801-1830):
// dE/dG = inpw . W_0^T diag(sigma'_0) ... W_L^T outw.
//
// PLAIN form: one thread per atom, FP32 GEMV chain with the weights read
// through the scalar/L1 path (every lane of a wave reads the same weight),
// layer widths padded to a compile-time maximum.
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

template <int MW>
__global__ __launch_bounds__(256) void k_net_plain(NetArgs A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.N) return;
    const int a = A.species[i];
    const NetDesc& D = A.dm->net[a];
    const int L = D.nlayer;
    const float c = D.c;
    const float* G = A.G + (size_t)i * A.ldg;
    const float* iw = A.inpw + D.inoff;
    const float* ib = A.inpb + D.inoff;
    float x[MW], z[MW];
    float dz[kMaxLayers - 1][MW];
    // layer 0
    {
        const int no = D.nodes[0];
        const float* W0 = A.W + D.woff[0];
        const float* b0 = A.W + D.boff[0];
#pragma unroll
        for (int n = 0; n < MW; ++n) z[n] = n < no ? b0[n] : 0.0f;
        for (int m = 0; m < D.nin; ++m) {
            const float s = iw[m] * (G[m] + ib[m]);
            const float* col = W0 + m * no;
#pragma unroll
            for (int n = 0; n < MW; ++n)
                if (n < no) z[n] += col[n] * s;
        }
    }
    for (int l = 0; l < L; ++l) {
        const int no = D.nodes[l];
        const bool last = (l == L - 1);
        if (l > 0) {
            const int ni = D.nodes[l - 1];
            const float* Wl = A.W + D.woff[l];
            const float* bl = A.W + D.boff[l];
#pragma unroll
            for (int n = 0; n < MW; ++n) z[n] = n < no ? bl[n] : 0.0f;
#pragma unroll
            for (int m = 0; m < MW; ++m) {
                if (m < ni) {
                    const float xm = x[m];
                    const float* col = Wl + m * no;
#pragma unroll
                    for (int n = 0; n < MW; ++n)
                        if (n < no) z[n] += col[n] * xm;
                }
            }
        }
        if (last) break;
        float d2;
#pragma unroll
        for (int n = 0; n < MW; ++n) {
            float av = 0.0f, dv = 0.0f;
            if (n < no) act_f<false>(D.neuron, c, z[n], av, dv, d2);
            x[n] = av;
#pragma unroll
            for (int q = 0; q < kMaxLayers - 1; ++q)
                if (q == l) dz[q][n] = dv;
        }
    }
    const float e = D.outb + D.outw * z[0];
    A.Eatom[i] = e;
    if (!A.dEdG) return;
    // backward: g = outw at the (linear) output, then W_l^T and sigma'
    float g[MW];
#pragma unroll
    for (int n = 0; n < MW; ++n) g[n] = n == 0 ? D.outw : 0.0f;
    for (int l = L - 1; l >= 1; --l) {
        const int no = D.nodes[l], ni = D.nodes[l - 1];
        const float* Wl = A.W + D.woff[l];
        float h[MW];
#pragma unroll
        for (int m = 0; m < MW; ++m) {
            float s = 0.0f;
            if (m < ni) {
                const float* col = Wl + m * no;
#pragma unroll
                for (int n = 0; n < MW; ++n)
                    if (n < no) s += col[n] * g[n];
            }
            float dd = 0.0f;
#pragma unroll
            for (int q = 0; q < kMaxLayers - 1; ++q)
                if (q == l - 1) dd = dz[q][m];
            h[m] = s * dd;
        }
#pragma unroll
        for (int m = 0; m < MW; ++m) g[m] = h[m];
    }
    {
        const int no = D.nodes[0];
        const float* W0 = A.W + D.woff[0];
        float* out = A.dEdG + (size_t)i * A.ldg;
        for (int m = 0; m < D.nin; ++m) {
            const float* col = W0 + m * no;
            float s = 0.0f;
#pragma unroll
            for (int n = 0; n < MW; ++n)
                if (n < no) s += col[n] * g[n];
            out[m] = s * iw[m];
        }
        for (int m = D.nin; m < A.ldg; ++m) out[m] = 0.0f;
    }
}

int max_width(const DevModel& M) {
    int w = 1;
    for (int a = 0; a < M.S; ++a)
        for (int l = 0; l < M.net[a].nlayer; ++l) w = std::max(w, M.net[a].nodes[l]);
    return w;
}

hipError_t net_fwd_plain(ann_ctx* ctx, const float* G, float* Eatom, float* dEdG) {
    NetArgs A{ctx->dm, ctx->da.W, ctx->da.inpw, ctx->da.inpb, ctx->natoms, ctx->ldg, ctx->d_species,
              G, Eatom, dEdG};
    if (A.N == 0) return hipSuccess;
    const int w = max_width(ctx->hm);
    const int nb = (A.N + 255) / 256;
    Launch L(ctx, "net_plain");
    if (w <= 16) k_net_plain<16><<<nb, 256, 0, ctx->stream>>>(A);
    else if (w <= 32) k_net_plain<32><<<nb, 256, 0, ctx->stream>>>(A);
    else if (w <= 48) k_net_plain<48><<<nb, 256, 0, ctx->stream>>>(A);
    else k_net_plain<64><<<nb, 256, 0, ctx->stream>>>(A);
    return hipGetLastError();
}

hipError_t net_fwd_mfma(ann_ctx* ctx, const float* G, float* Eatom, float* dEdG) {
    return net_fwd_plain(ctx, G, Eatom, dEdG);
}

hipError_t net_fwd(ann_ctx* ctx, int form, const float* G, float* Eatom, float* dEdG) {
    if (form == ANN_FORM_FAST) return net_fwd_mfma(ctx, G, Eatom, dEdG);
    return net_fwd_plain(ctx, G, Eatom, dEdG);
}

}  // namespace ann
