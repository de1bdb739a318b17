// train.hip -- energy + force training step: loss, the second backward
// (weight gradient of the force loss), FP64 reductions and Adam.
//
// Loss of a batch of B structures (energy part as NNPTEFR::error_cost_o1 with
// MSE and Norm::LINEAR, src/nnp/nnptefr2.cpp:1321-1373, plus a force term):
//   L = (1/B) sum_s [ wE/2 ((E_s - Eref_s)/N_s)^2 + wF/(3 N_s) sum_a |F_a - Fref_a|^2 ]
// With u_a = dL/dF_a and F = -(dG/dr)^T dE/dG, the force term reaches the
// parameters only through dE/dG, so dL/d(dE/dG_i) = v_i = -dG_i[u] (a
// forward-mode tangent of the basis, ann_symm_tangent) and
//   dL/dp = sum_i d/dp [ gE_s(i) E_i + dE_i/dG . v_i ].
// The second term is the network's tangent along v back-propagated through
// both the primal and the tangent network (needs sigma''); it replaces the
// reference's full d2O/dp dI tensor (D2ODZDI::grad, src/ml/nn.cpp:2012-2103)
// contracted with dG/dr (src/nnp/nnptef.cpp:1740-1815).
// Parameter gradients are sums over atoms of outer products; they are
// reduced per species with a fixed atom order in FP64.  Adam follows
// ADAM::step (src/opt/algo.cpp:280-296; beta1 0.9, beta2 0.999, eps 1e-16
// by default in the reference).
#include <hipcub/hipcub.hpp>

#include "net_dev.h"

namespace ann {

hipError_t exclusive_scan(const int* in, int* out, int n, hipStream_t st);
hipError_t ensure_work_buffers(ann_ctx* ctx);

// per-atom stride of the second-backward scratch: per layer x, xd (n_in), zbar, zdbar (n_out)
__host__ __device__ inline int bwd2_stride(const NetDesc& D) {
    int s = 0, prev = D.nin;
    for (int l = 0; l < D.nlayer; ++l) { s += 2 * (prev + D.nodes[l]); prev = D.nodes[l]; }
    return s;
}

struct Bwd2Args {
    const DevModel* dm;
    const float* W;
    const float* inpw;
    const float* inpb;
    int N, ldg;
    const int* species;
    const int* sof;
    const int* rank;
    const float* G;
    const float* v;
    const double* gE;
    float* scr;
    int64_t base[kMaxSpecies];   // scratch start of each species
};

template <int MW>
__global__ __launch_bounds__(256) void k_net_bwd2(Bwd2Args A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.N) return;
    const int a = A.species[i];
    const NetDesc& D = A.dm->net[a];
    const int L = D.nlayer;
    const float c = D.c;
    const float* G = A.G + (size_t)i * A.ldg;
    const float* V = A.v ? A.v + (size_t)i * A.ldg : nullptr;
    const float* iw = A.inpw + D.inoff;
    const float* ib = A.inpb + D.inoff;
    float* row = A.scr + A.base[a] + (int64_t)A.rank[i] * bwd2_stride(D);
    const float gE = (float)A.gE[A.sof[i]];
    // scratch offsets per layer
    int loff[kMaxLayers];
    {
        int s = 0, prev = D.nin;
        for (int l = 0; l < L; ++l) { loff[l] = s; s += 2 * (prev + D.nodes[l]); prev = D.nodes[l]; }
    }
    float x[MW], xd[MW];
    float dz[kMaxLayers - 1][MW], q2[kMaxLayers - 1][MW];
    float z[MW], zd[MW];
    // layer 0 from the scaled input s = inpw (G + inpb), tangent sd = inpw v
    {
        const int no = D.nodes[0];
        const float* W0 = A.W + D.woff[0];
        const float* b0 = A.W + D.boff[0];
        float* xr = row + loff[0];
#pragma unroll
        for (int n = 0; n < MW; ++n) { z[n] = n < no ? b0[n] : 0.0f; zd[n] = 0.0f; }
        for (int m = 0; m < D.nin; ++m) {
            const float s = iw[m] * (G[m] + ib[m]);
            const float sd = V ? iw[m] * V[m] : 0.0f;
            xr[m] = s;
            xr[D.nin + m] = sd;
            const float* col = W0 + m * no;
#pragma unroll
            for (int n = 0; n < MW; ++n)
                if (n < no) { z[n] += col[n] * s; zd[n] += col[n] * sd; }
        }
    }
    for (int l = 0; l < L; ++l) {
        const int no = D.nodes[l];
        if (l > 0) {
            const int ni = D.nodes[l - 1];
            const float* Wl = A.W + D.woff[l];
            const float* bl = A.W + D.boff[l];
            float* xr = row + loff[l];
#pragma unroll
            for (int m = 0; m < MW; ++m)
                if (m < ni) { xr[m] = x[m]; xr[ni + m] = xd[m]; }
#pragma unroll
            for (int n = 0; n < MW; ++n) { z[n] = n < no ? bl[n] : 0.0f; zd[n] = 0.0f; }
#pragma unroll
            for (int m = 0; m < MW; ++m) {
                if (m < ni) {
                    const float* col = Wl + m * no;
#pragma unroll
                    for (int n = 0; n < MW; ++n)
                        if (n < no) { z[n] += col[n] * x[m]; zd[n] += col[n] * xd[m]; }
                }
            }
        }
        if (l == L - 1) break;
#pragma unroll
        for (int n = 0; n < MW; ++n) {
            float av = 0.0f, dv = 0.0f, d2 = 0.0f;
            if (n < no) act_f<true>(D.neuron, c, z[n], av, dv, d2);
            x[n] = av;
            xd[n] = dv * zd[n];
#pragma unroll
            for (int q = 0; q < kMaxLayers - 1; ++q)
                if (q == l) { dz[q][n] = dv; q2[q][n] = d2 * zd[n]; }
        }
    }
    // backward: output layer is linear; abar = gE outw, adbar = outw (with v)
    float ab[MW], adb[MW];
#pragma unroll
    for (int n = 0; n < MW; ++n) {
        ab[n] = n == 0 ? gE * D.outw : 0.0f;
        adb[n] = (n == 0 && V) ? D.outw : 0.0f;
    }
    for (int l = L - 1; l >= 0; --l) {
        const int no = D.nodes[l];
        const int ni = l == 0 ? D.nin : D.nodes[l - 1];
        float zb[MW], zdb[MW];
#pragma unroll
        for (int n = 0; n < MW; ++n) {
            float dd = 1.0f, qq = 0.0f;
#pragma unroll
            for (int q = 0; q < kMaxLayers - 1; ++q)
                if (q == l && l < L - 1) { dd = dz[q][n]; qq = q2[q][n]; }
            zdb[n] = adb[n] * dd;
            zb[n] = ab[n] * dd + adb[n] * qq;
        }
        float* br = row + loff[l] + 2 * ni;
#pragma unroll
        for (int n = 0; n < MW; ++n)
            if (n < no) { br[n] = zb[n]; br[no + n] = zdb[n]; }
        if (l == 0) break;
        const float* Wl = A.W + D.woff[l];
#pragma unroll
        for (int m = 0; m < MW; ++m) {
            float s1 = 0.0f, s2 = 0.0f;
            if (m < ni) {
                const float* col = Wl + m * no;
#pragma unroll
                for (int n = 0; n < MW; ++n)
                    if (n < no) { s1 += col[n] * zb[n]; s2 += col[n] * zdb[n]; }
            }
            ab[m] = s1;
            adb[m] = s2;
        }
    }
}

struct GradArgs {
    const DevModel* dm;
    const float* scr;
    const int* spcount;
    int S;
    int64_t base[kMaxSpecies];
    double* grad;
};

// one thread per parameter; sums over the atoms of its species in rank order (FP64)
__global__ __launch_bounds__(256) void k_grad_reduce(GradArgs A, int ptot) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ptot) return;
    int a = 0;
    while (a + 1 < A.S && p >= A.dm->net[a + 1].poff) ++a;
    const NetDesc& D = A.dm->net[a];
    const int loc = p - D.poff;
    const int stride = bwd2_stride(D);
    // locate layer and (n, m)
    int l = 0, cnt = 0, prev = D.nin;
    bool bias = false;
    int n = 0, m = 0;
    for (int q = 0; q < D.nlayer; ++q) {
        if (loc < cnt + D.nodes[q]) { bias = true; l = q; n = loc - cnt; break; }
        cnt += D.nodes[q];
    }
    int lo = 0;
    if (!bias) {
        for (int q = 0; q < D.nlayer; ++q) {
            const int sz = D.nodes[q] * prev;
            if (loc < cnt + sz) { l = q; const int r = loc - cnt; m = r / D.nodes[q]; n = r % D.nodes[q]; break; }
            cnt += sz;
            prev = D.nodes[q];
        }
    }
    int ni = D.nin;
    for (int q = 0; q < l; ++q) { lo += 2 * (ni + D.nodes[q]); ni = D.nodes[q]; }
    const int no = D.nodes[l];
    const float* base = A.scr + A.base[a] + lo;
    const int na = A.spcount[a];
    double acc = 0.0;
    if (bias) {
        for (int r = 0; r < na; ++r) acc += (double)base[(int64_t)r * stride + 2 * ni + n];
    } else {
        for (int r = 0; r < na; ++r) {
            const float* rr = base + (int64_t)r * stride;
            acc += (double)rr[2 * ni + n] * (double)rr[m] + (double)rr[2 * ni + no + n] * (double)rr[ni + m];
        }
    }
    A.grad[p] = acc;
}

__global__ void k_flags(int N, const int* __restrict__ species, int s, int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= N) flag[i] = (i < N && species[i] == s) ? 1 : 0;
}
__global__ void k_take_rank(int N, const int* __restrict__ species, int s, const int* __restrict__ pre,
                            int* __restrict__ rank, int* __restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N && species[i] == s) rank[i] = pre[i];
    if (i == N) count[s] = pre[N];
}

hipError_t species_ranks(ann_ctx* ctx) {
    if (ctx->rank_valid) return hipSuccess;
    const int N = ctx->natoms;
    hipError_t e;
    if (ctx->d_rank) hipFree(ctx->d_rank);
    if (ctx->d_spcount) hipFree(ctx->d_spcount);
    if ((e = hipMalloc(&ctx->d_rank, sizeof(int) * (N + 1))) != hipSuccess) return e;
    if ((e = hipMalloc(&ctx->d_spcount, sizeof(int) * kMaxSpecies)) != hipSuccess) return e;
    int *flag = nullptr, *pre = nullptr;
    if ((e = hipMalloc(&flag, sizeof(int) * (N + 1))) != hipSuccess) return e;
    if ((e = hipMalloc(&pre, sizeof(int) * (N + 1))) != hipSuccess) return e;
    for (int s = 0; s < ctx->S; ++s) {
        k_flags<<<(N + 256) / 256, 256, 0, ctx->stream>>>(N, ctx->d_species, s, flag);
        if ((e = exclusive_scan(flag, pre, N + 1, ctx->stream)) != hipSuccess) return e;
        k_take_rank<<<(N + 256) / 256, 256, 0, ctx->stream>>>(N, ctx->d_species, s, pre, ctx->d_rank, ctx->d_spcount);
    }
    std::vector<int> counts(ctx->S);
    e = hipMemcpyAsync(counts.data(), ctx->d_spcount, sizeof(int) * ctx->S, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(flag);
    hipFree(pre);
    if (e != hipSuccess) return e;
    // scratch layout: species-major
    int64_t tot = 0;
    ctx->rank_valid = true;
    for (int s = 0; s < ctx->S; ++s) tot += (int64_t)counts[s] * bwd2_stride(ctx->hm.net[s]);
    if (tot > ctx->scr_cap) {
        if (ctx->d_scr) hipFree(ctx->d_scr);
        if ((e = hipMalloc(&ctx->d_scr, sizeof(float) * std::max<int64_t>(tot, 1))) != hipSuccess) return e;
        ctx->scr_cap = tot;
    }
    ctx->hm_counts.assign(counts.begin(), counts.end());
    return hipSuccess;
}

int max_width(const DevModel& M);

__global__ void k_perm(int N, const int* __restrict__ species, const int* __restrict__ rank,
                       const int* __restrict__ base, int* __restrict__ perm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) perm[base[species[i]] + rank[i]] = i;
}

// species-sorted permutation of the batch and its 16-atom single-species tiles
hipError_t species_perm(ann_ctx* ctx) {
    if (ctx->perm_valid) return hipSuccess;
    hipError_t e = species_ranks(ctx);
    if (e != hipSuccess) return e;
    const int S = ctx->S, N = ctx->natoms;
    std::vector<int> base(S), end(S), tiles;
    int acc = 0;
    for (int s = 0; s < S; ++s) {
        base[s] = acc;
        for (int t = 0; t < ctx->hm_counts[s]; t += 16) { tiles.push_back(s); tiles.push_back(acc + t); }
        acc += ctx->hm_counts[s];
        end[s] = acc;
    }
    for (int* p : {ctx->d_perm, ctx->d_tiles, ctx->d_spend})
        if (p) hipFree(p);
    ctx->d_perm = ctx->d_tiles = ctx->d_spend = nullptr;
    if ((e = hipMalloc(&ctx->d_perm, sizeof(int) * std::max(N, 1))) != hipSuccess) return e;
    if ((e = hipMalloc(&ctx->d_tiles, sizeof(int) * std::max<size_t>(tiles.size(), 2))) != hipSuccess) return e;
    if ((e = hipMalloc(&ctx->d_spend, sizeof(int) * 2 * S)) != hipSuccess) return e;
    std::vector<int> be(base);
    be.insert(be.end(), end.begin(), end.end());
    if ((e = hipMemcpyAsync(ctx->d_spend, be.data(), sizeof(int) * 2 * S, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess) return e;
    if (!tiles.empty() &&
        (e = hipMemcpyAsync(ctx->d_tiles, tiles.data(), sizeof(int) * tiles.size(), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
        return e;
    k_perm<<<(N + 255) / 256, 256, 0, ctx->stream>>>(N, ctx->d_species, ctx->d_rank, ctx->d_spend, ctx->d_perm);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // spend[] must hold the END of each species block: shift the uploaded (base, end) pair
    ctx->d_spend_end = ctx->d_spend + S;
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
    ctx->ntiles = (int)tiles.size() / 2;
    ctx->perm_valid = true;
    return hipSuccess;
}

hipError_t net_bwd2(ann_ctx* ctx, const float* G, const float* v, const double* gE, double* grad) {
    hipError_t e = species_ranks(ctx);
    if (e != hipSuccess) return e;
    Bwd2Args A{};
    A.dm = ctx->dm;
    A.W = ctx->da.W;
    A.inpw = ctx->da.inpw;
    A.inpb = ctx->da.inpb;
    A.N = ctx->natoms;
    A.ldg = ctx->ldg;
    A.species = ctx->d_species;
    A.sof = ctx->d_struc_of;
    A.rank = ctx->d_rank;
    A.G = G;
    A.v = v;
    A.gE = gE;
    A.scr = ctx->d_scr;
    int64_t b = 0;
    for (int s = 0; s < ctx->S; ++s) {
        A.base[s] = b;
        b += (int64_t)ctx->hm_counts[s] * bwd2_stride(ctx->hm.net[s]);
    }
    const int nb = (A.N + 255) / 256;
    const int w = max_width(ctx->hm);
    {
        Launch L(ctx, "net_bwd2");
        if (w <= 16) k_net_bwd2<16><<<nb, 256, 0, ctx->stream>>>(A);
        else if (w <= 32) k_net_bwd2<32><<<nb, 256, 0, ctx->stream>>>(A);
        else if (w <= 48) k_net_bwd2<48><<<nb, 256, 0, ctx->stream>>>(A);
        else k_net_bwd2<64><<<nb, 256, 0, ctx->stream>>>(A);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    GradArgs R{};
    R.dm = ctx->dm;
    R.scr = ctx->d_scr;
    R.spcount = ctx->d_spcount;
    R.S = ctx->S;
    for (int s = 0; s < kMaxSpecies; ++s) R.base[s] = A.base[s];
    R.grad = grad;
    {
        Launch L(ctx, "grad_reduce");
        k_grad_reduce<<<(ctx->nparams + 255) / 256, 256, 0, ctx->stream>>>(R, ctx->nparams);
    }
    return hipGetLastError();
}

// ---- loss: one block per structure, FP64 ----
__global__ __launch_bounds__(256) void k_loss(const int* __restrict__ aoff, const double* __restrict__ E,
                                              const double* __restrict__ F, const double* __restrict__ Eref,
                                              const double* __restrict__ Fref, double wE, double wF, int B,
                                              double* __restrict__ lossE, double* __restrict__ lossF,
                                              double* __restrict__ gE, float* __restrict__ uneg) {
    __shared__ double red[256];
    const int s = blockIdx.x, t = threadIdx.x;
    const int a0 = aoff[s], a1 = aoff[s + 1], N = a1 - a0;
    const double scale = wF * 2.0 / (3.0 * N) / B;
    double acc = 0.0;
    for (int x = 3 * a0 + t; x < 3 * a1; x += blockDim.x) {
        const double df = F[x] - Fref[x];
        acc += df * df;
        uneg[x] = (float)(-scale * df);
    }
    red[t] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) {
        const double dEn = (E[s] - Eref[s]) / N;
        lossE[s] = 0.5 * wE * dEn * dEn / B;
        lossF[s] = wF * red[0] / (3.0 * N) / B;
        gE[s] = wE * dEn / N / B;
    }
}

// fixed-order sum of the per-structure loss terms
__global__ void k_loss_total(int B, const double* lossE, const double* lossF, double* out) {
    __shared__ double red[2][256];
    const int t = threadIdx.x;
    double e = 0, f = 0;
    for (int s = t; s < B; s += blockDim.x) { e += lossE[s]; f += lossF[s]; }
    red[0][t] = e;
    red[1][t] = f;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) { red[0][t] += red[0][t + w]; red[1][t] += red[1][t + w]; }
        __syncthreads();
    }
    if (t == 0) { out[0] = red[0][0]; out[1] = red[1][0]; }
}

__global__ void k_adam(int n, double* __restrict__ p, double* __restrict__ m1, double* __restrict__ m2,
                       const double* __restrict__ g, double gamma, double b1ii, double b2ii, double eps,
                       float* __restrict__ W) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double b1 = 0.9, b2 = 0.999;
    const double gi = g[i];
    const double a = m1[i] * b1 + (1.0 - b1) * gi;
    const double b = m2[i] * b2 + (1.0 - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const double np = p[i] - gamma * b1ii * a / (sqrt(b2ii * b) + eps);
    p[i] = np;
    W[i] = (float)np;
}

hipError_t train_step_impl(ann_ctx* ctx, int form, const double* Eref, const double* Fref, double wE, double wF,
                           double gamma, double* loss, double* grad) {
    hipError_t e;
    const int B = ctx->nstruc, N = ctx->natoms;
    if (!ctx->G_valid) {
        if ((e = symm_fwd(ctx, form, ctx->d_G)) != hipSuccess) return e;
        ctx->G_valid = true;
    }
    if ((e = net_fwd(ctx, form, ctx->d_G, ctx->d_Eatom, ctx->d_dEdG)) != hipSuccess) return e;
    if ((e = symm_bwd(ctx, form, ctx->d_dEdG, ctx->d_P)) != hipSuccess) return e;
    if ((e = force_reduce(ctx, ctx->d_P, ctx->d_Eatom, 0, ctx->d_F, ctx->d_E, nullptr)) != hipSuccess) return e;
    // per-structure loss terms live after the gradient scratch
    if (ctx->gpart_cap < 2 * (int64_t)B + 2 + ctx->nparams) {
        if (ctx->d_gpart) hipFree(ctx->d_gpart);
        if ((e = hipMalloc(&ctx->d_gpart, sizeof(double) * (2 * (size_t)B + 2 + ctx->nparams))) != hipSuccess) return e;
        ctx->gpart_cap = 2 * (int64_t)B + 2 + ctx->nparams;
    }
    double* lE = ctx->d_gpart;
    double* lF = lE + B;
    double* lt = lF + B;
    double* g = grad ? grad : lt + 2;
    {
        Launch L(ctx, "loss");
        k_loss<<<B, 256, 0, ctx->stream>>>(ctx->d_atom_off, ctx->d_E, ctx->d_F, Eref, Fref, wE, wF, B, lE, lF,
                                           ctx->d_gE, ctx->d_u);
        k_loss_total<<<1, 256, 0, ctx->stream>>>(B, lE, lF, loss ? loss : lt);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    (void)N;
    if ((e = symm_tangent(ctx, form, ctx->d_u, ctx->d_v)) != hipSuccess) return e;
    if ((e = net_bwd2(ctx, ctx->d_G, ctx->d_v, ctx->d_gE, g)) != hipSuccess) return e;
    if (gamma > 0) {
        ctx->adam_t += 1;
        const double b1ii = 1.0 / (1.0 - std::pow(0.9, ctx->adam_t));
        const double b2ii = 1.0 / (1.0 - std::pow(0.999, ctx->adam_t));
        Launch L(ctx, "adam");
        k_adam<<<(ctx->nparams + 255) / 256, 256, 0, ctx->stream>>>(ctx->nparams, ctx->da.params, ctx->da.adam_m,
                                                                    ctx->da.adam_v, g, gamma, b1ii, b2ii, 1e-16,
                                                                    ctx->da.W);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace ann
