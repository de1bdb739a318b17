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

typedef float f32x4 __attribute__((ext_vector_type(4)));
hipError_t species_perm(ann_ctx* ctx);
bool mfma_net_supported(const DevModel& M);

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
    int kp, wp, nlh;             // matrix-core kernel: LDS row strides of the activations and of sigma', sigma'';
                                 // most hidden layers of any species
};

// 16 lanes per atom, 4 atoms per wavefront; lane q owns neurons q, q+16, ...
constexpr int kBLanes = 16;
constexpr int kBOut = kMaxWidth / kBLanes;

__global__ __launch_bounds__(64) void k_net_bwd2(Bwd2Args A) {
    __shared__ float sx[4][kMaxWidth + 1], sxd[4][kMaxWidth + 1];
    const int sub = threadIdx.x / kBLanes, q = threadIdx.x % kBLanes;
    const int i = blockIdx.x * 4 + sub;
    const bool valid = i < A.N;
    const int ii = valid ? i : 0;
    const int a = A.species[ii];
    const NetDesc& D = A.dm->net[a];
    const int L = D.nlayer;
    const float c = D.c;
    const float* G = A.G + (size_t)ii * A.ldg;
    const float* V = A.v ? A.v + (size_t)ii * A.ldg : nullptr;
    const float* iw = A.inpw + D.inoff;
    const float* ib = A.inpb + D.inoff;
    float* row = A.scr + A.base[a] + (int64_t)A.rank[ii] * bwd2_stride(D);
    const float gE = (float)A.gE[A.sof[ii]];
    int loff[kMaxLayers];
    {
        int s0 = 0, prev = D.nin;
        for (int l = 0; l < L; ++l) { loff[l] = s0; s0 += 2 * (prev + D.nodes[l]); prev = D.nodes[l]; }
    }
    // scaled input and its tangent: scratch of layer 0
    for (int m = q; m < D.nin; m += kBLanes)
        if (valid) {
            row[m] = iw[m] * (G[m] + ib[m]);
            row[D.nin + m] = V ? iw[m] * V[m] : 0.0f;
        }
    float dz[kMaxLayers - 1][kBOut], qz[kMaxLayers - 1][kBOut];
    float zo = 0.0f, zdo = 0.0f;   // output neuron (lane 0 after the reduction)
    int ni = D.nin;
    for (int l = 0; l < L; ++l) {
        const int no = D.nodes[l];
        const float* Wl = A.W + D.woff[l];
        const float* bl = A.W + D.boff[l];
        const bool last = l == L - 1;
        if (!last) {
            float z[kBOut], zd[kBOut];
#pragma unroll
            for (int j = 0; j < kBOut; ++j) {
                const int n = q + kBLanes * j;
                z[j] = n < no ? bl[n] : 0.0f;
                zd[j] = 0.0f;
            }
            for (int m = 0; m < ni; ++m) {
                float x, xd;
                if (l == 0) { x = iw[m] * (G[m] + ib[m]); xd = V ? iw[m] * V[m] : 0.0f; }
                else { x = sx[sub][m]; xd = sxd[sub][m]; }
                const float* col = Wl + m * no;
#pragma unroll
                for (int j = 0; j < kBOut; ++j) {
                    const int n = q + kBLanes * j;
                    if (n < no) { z[j] += col[n] * x; zd[j] += col[n] * xd; }
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kBOut; ++j) {
                const int n = q + kBLanes * j;
                float av = 0.0f, dv = 0.0f, d2 = 0.0f;
                if (n < no) act_f<true>(D.neuron, c, z[j], av, dv, d2);
                if (n < no) {
                    sx[sub][n] = av;
                    sxd[sub][n] = dv * zd[j];
                    if (valid) {
                        row[loff[l + 1] + n] = av;
                        row[loff[l + 1] + no + n] = dv * zd[j];
                    }
                }
#pragma unroll
                for (int p = 0; p < kMaxLayers - 1; ++p)
                    if (p == l) { dz[p][j] = dv; qz[p][j] = d2 * zd[j]; }
            }
            __syncthreads();
        } else {
            // linear output neuron: only its pre-activation tangent matters below
            (void)zo; (void)zdo;
        }
        ni = no;
    }
    // backward.  Output layer (linear, one neuron): zbar = gE outw, zdbar = outw (with v)
    const float zbL = gE * D.outw, zdbL = V ? D.outw : 0.0f;
    if (valid && q == 0) {
        const int nprev = L > 1 ? D.nodes[L - 2] : D.nin;
        row[loff[L - 1] + 2 * nprev] = zbL;
        row[loff[L - 1] + 2 * nprev + 1] = zdbL;
    }
    if (L == 1) return;
    const float* WL = A.W + D.woff[L - 1];
    // abar, adbar of the last hidden layer, then zbar/zdbar
    float zb[kBOut], zdb[kBOut];
    {
        const int nl = D.nodes[L - 2];
#pragma unroll
        for (int j = 0; j < kBOut; ++j) {
            const int m = q + kBLanes * j;
            float d = 0.0f, qq = 0.0f;
#pragma unroll
            for (int p = 0; p < kMaxLayers - 1; ++p)
                if (p == L - 2) { d = dz[p][j]; qq = qz[p][j]; }
            const float ab = m < nl ? WL[m] * zbL : 0.0f, adb = m < nl ? WL[m] * zdbL : 0.0f;
            zdb[j] = adb * d;
            zb[j] = ab * d + adb * qq;
        }
    }
    for (int l = L - 2; l >= 0; --l) {
        const int no = D.nodes[l];
        const int nprev = l == 0 ? D.nin : D.nodes[l - 1];
        const int base = loff[l] + 2 * nprev;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kBOut; ++j) {
            const int n = q + kBLanes * j;
            if (n < no) {
                sx[sub][n] = zb[j];
                sxd[sub][n] = zdb[j];
                if (valid) { row[base + n] = zb[j]; row[base + no + n] = zdb[j]; }
            }
        }
        __syncthreads();
        if (l == 0) break;
        const float* Wl = A.W + D.woff[l];
        float nb[kBOut], ndb[kBOut];
#pragma unroll
        for (int j = 0; j < kBOut; ++j) {
            const int m = q + kBLanes * j;
            float s1 = 0.0f, s2 = 0.0f;
            if (m < nprev) {
                const float* col = Wl + m * no;
                for (int n = 0; n < no; ++n) { s1 += col[n] * sx[sub][n]; s2 += col[n] * sxd[sub][n]; }
            }
            float d = 0.0f, qq = 0.0f;
#pragma unroll
            for (int p = 0; p < kMaxLayers - 1; ++p)
                if (p == l - 1) { d = dz[p][j]; qq = qz[p][j]; }
            ndb[j] = s2 * d;
            nb[j] = s1 * d + s2 * qq;
        }
#pragma unroll
        for (int j = 0; j < kBOut; ++j) { zb[j] = nb[j]; zdb[j] = ndb[j]; }
    }
}

// The same second backward on the matrix cores: one wavefront per 16-atom
// single-species tile (species-sorted permutation), every layer product as
// v_mfma_f32_16x16x4_f32 (FP32 operands and accumulation, so no operand
// splitting is needed):
//   forward:   Z = X W^T (+ b), Zd = Xd W^T, a = sigma(Z), ad = sigma'(Z) Zd;
//   backward:  zbar_{l-1} = (zbar_l W_l) sigma' + (zdbar_l W_l) sigma'' zd,
//              zdbar_{l-1} = (zdbar_l W_l) sigma'.
// Lane layout of 16x16x4: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]
// and receives D rows 4 (l / 16) .. +3 of column l % 16.
constexpr int kB2Tile = 16;
constexpr int kB2K = 128;      // widest input supported (same bound as the forward MFMA network)
constexpr int kB2Chunk = 8;    // k steps (of 4) per weight chunk held in registers

// NT: this wavefront's column tiles (2 for hidden layers up to 64 neurons, 4 up to 128)
template <int NT>
__global__ __launch_bounds__(128) void k_bwd2_mfma(Bwd2Args A, const int* __restrict__ perm,
                                                  const int2* __restrict__ tiles, const int* __restrict__ spend) {
    // dynamic LDS sized by the model: activations / gradients [16][kp] (x and its
    // tangent) and sigma', sigma'' zd of every hidden layer [L-1][16][wp]
    extern __shared__ float b2smem[];
    const int KP = A.kp, WP = A.wp;
    float* sx = b2smem;
    float* sxd = sx + kB2Tile * KP;
    float* sdz = sxd + kB2Tile * KP;
    float* sqz = sdz + A.nlh * kB2Tile * WP;
    __shared__ int64_t srow[kB2Tile];
    __shared__ float sgE[kB2Tile];
    // two wavefronts per tile: column tiles ct = 2 h + wv of every product go to wavefront wv
    const int tid = threadIdx.x, lane = tid % kWave, wv = tid / kWave;
    constexpr int kThreads = 2 * kWave;
    const int2 tl = tiles[blockIdx.x];
    const int a = tl.x;
    const int nrow = min(kB2Tile, spend[a] - tl.y);
    const NetDesc& D = A.dm->net[a];
    const int L = D.nlayer;
    const int stride = bwd2_stride(D);
    const float* iw = A.inpw + D.inoff;
    const float* ib = A.inpb + D.inoff;
    if (tid < kB2Tile) {
        const int ai = tid < nrow ? perm[tl.y + tid] : -1;
        srow[tid] = ai >= 0 ? A.base[a] + (int64_t)A.rank[ai] * stride : -1;
        sgE[tid] = ai >= 0 ? (float)A.gE[A.sof[ai]] : 0.0f;
    }
    int loff[kMaxLayers];
    {
        int s0 = 0, prev = D.nin;
        for (int l = 0; l < L; ++l) { loff[l] = s0; s0 += 2 * (prev + D.nodes[l]); prev = D.nodes[l]; }
    }
    // layer-0 input x = inpw (G + inpb), xd = inpw v
#pragma unroll 8
    for (int q = tid; q < kB2Tile * KP; q += kThreads) {
        const int r = q / KP, k = q % KP;
        const int ai = r < nrow ? perm[tl.y + r] : -1;
        float x = 0.0f, xd = 0.0f;
        if (ai >= 0 && k < D.nin) {
            x = iw[k] * (A.G[(size_t)ai * A.ldg + k] + ib[k]);
            xd = A.v ? iw[k] * A.v[(size_t)ai * A.ldg + k] : 0.0f;
        }
        sx[r * KP + k] = x;
        sxd[r * KP + k] = xd;
    }
    __syncthreads();
    __syncthreads();
    for (int q = tid; q < kB2Tile * D.nin; q += kThreads) {
        const int r = q / D.nin, k = q % D.nin;
        if (srow[r] >= 0) {
            float* row = A.scr + srow[r];
            row[k] = sx[r * KP + k];
            row[D.nin + k] = sxd[r * KP + k];
        }
    }
    const int c16 = lane % 16, k4 = lane / 16, rb = 4 * k4;
    int ni = D.nin;
    for (int l = 0; l < L - 1; ++l) {
        const int no = D.nodes[l];
        const float* Wl = A.W + D.woff[l];
        const float* bl = A.W + D.boff[l];
        float oa[NT][4], od[NT][4];
        // all column tiles advance together over k, so the weight loads of one
        // k step (independent across tiles and unrolled steps) are in flight at once
        f32x4 AZ[NT], AD[NT];
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) { AZ[ct] = f32x4{0.f, 0.f, 0.f, 0.f}; AD[ct] = f32x4{0.f, 0.f, 0.f, 0.f}; }
        // this lane's weight operands in chunks of kB2Chunk k steps, the next
        // chunk's loads issued before the current chunk's products (bounded
        // registers, so two wavefronts fit on a SIMD)
        auto wload = [&](float (&w)[kB2Chunk][NT], int c0) {
#pragma unroll
            for (int st = 0; st < kB2Chunk; ++st) {
                const int k = 4 * (c0 + st) + k4;
#pragma unroll
                for (int h = 0; h < NT; ++h) {
                    const int n = (2 * h + wv) * 16 + c16;
                    w[st][h] = (k < ni && n < no) ? Wl[k * no + n] : 0.0f;
                }
            }
        };
        auto wmul = [&](const float (&w)[kB2Chunk][NT], int c0) {
#pragma unroll
            for (int st = 0; st < kB2Chunk; ++st) {
                if (4 * (c0 + st) >= ni) break;
                const int k = 4 * (c0 + st) + k4;
                const bool ok = k < ni;
                const float xa = ok ? sx[c16 * KP + k] : 0.0f, xda = ok ? sxd[c16 * KP + k] : 0.0f;
#pragma unroll
                for (int h = 0; h < NT; ++h) {
                    if ((2 * h + wv) * 16 >= no) break;
                    AZ[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, w[st][h], AZ[h], 0, 0, 0);
                    AD[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(xda, w[st][h], AD[h], 0, 0, 0);
                }
            }
        };
        {
            float wa[kB2Chunk][NT], wb[kB2Chunk][NT];
            const int nst = (ni + 3) / 4;
            wload(wa, 0);
            for (int c0 = 0; c0 < nst; c0 += 2 * kB2Chunk) {
                if (c0 + kB2Chunk < nst) wload(wb, c0 + kB2Chunk);
                wmul(wa, c0);
                if (c0 + kB2Chunk >= nst) break;
                if (c0 + 2 * kB2Chunk < nst) wload(wa, c0 + 2 * kB2Chunk);
                wmul(wb, c0 + kB2Chunk);
            }
        }
#pragma unroll
        for (int h = 0; h < NT; ++h) {
            const int ct = 2 * h + wv;
            if (ct * 16 >= no) break;
            const int n = ct * 16 + c16;
            const f32x4 az = AZ[h], ad = AD[h];
            const float bn = n < no ? bl[n] : 0.0f;
            float z[4], av[4], dv[4], d2[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) z[i] = az[i] + bn;
            act_many<true, 4>(D.neuron, D.c, z, av, dv, d2);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (n >= no) av[i] = dv[i] = d2[i] = 0.0f;
                oa[h][i] = av[i];
                od[h][i] = dv[i] * ad[i];
                if (n < no) {
                    sdz[(l * kB2Tile + rb + i) * WP + n] = dv[i];
                    sqz[(l * kB2Tile + rb + i) * WP + n] = d2[i] * ad[i];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < NT; ++h) {
            const int ct = 2 * h + wv;
            if (ct * 16 >= no) break;
            const int n = ct * 16 + c16;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = rb + i;
                if (n < no) {
                    sx[r * KP + n] = oa[h][i];
                    sxd[r * KP + n] = od[h][i];
                    if (srow[r] >= 0) {
                        float* row = A.scr + srow[r] + loff[l + 1];
                        row[n] = oa[h][i];
                        row[no + n] = od[h][i];
                    }
                }
            }
        }
        __syncthreads();
        ni = no;
    }
    // output layer (linear, one neuron): zbar = gE outw, zdbar = outw
    const float zdbL = A.v ? D.outw : 0.0f;
    const int nlast = L > 1 ? D.nodes[L - 2] : D.nin;
    if (tid < kB2Tile && srow[tid] >= 0) {
        float* row = A.scr + srow[tid] + loff[L - 1] + 2 * nlast;
        row[0] = sgE[tid] * D.outw;
        row[1] = zdbL;
    }
    if (L == 1) return;
    // last hidden layer: abar = W_L^T zbar_L, then zbar / zdbar through sigma', sigma''
    {
        const float* WL = A.W + D.woff[L - 1];
        __syncthreads();
        for (int q = tid; q < kB2Tile * nlast; q += kThreads) {
            const int r = q / nlast, m = q % nlast;
            const float d = sdz[((L - 2) * kB2Tile + r) * WP + m], qq = sqz[((L - 2) * kB2Tile + r) * WP + m];
            const float ab = WL[m] * sgE[r] * D.outw, adb = WL[m] * zdbL;
            sx[r * KP + m] = ab * d + adb * qq;
            sxd[r * KP + m] = adb * d;
        }
        __syncthreads();
    }
    for (int l = L - 2; l >= 0; --l) {
        const int no = D.nodes[l];
        const int nprev = l == 0 ? D.nin : D.nodes[l - 1];
        for (int q = tid; q < kB2Tile * no; q += kThreads) {
            const int r = q / no, n = q % no;
            if (srow[r] >= 0) {
                float* row = A.scr + srow[r] + loff[l] + 2 * nprev;
                row[n] = sx[r * KP + n];
                row[no + n] = sxd[r * KP + n];
            }
        }
        if (l == 0) break;
        const float* Wl = A.W + D.woff[l];
        float nb[NT][4], ndb[NT][4];
        f32x4 A1[NT], A2[NT];
#pragma unroll
        for (int mt = 0; mt < NT; ++mt) { A1[mt] = f32x4{0.f, 0.f, 0.f, 0.f}; A2[mt] = f32x4{0.f, 0.f, 0.f, 0.f}; }
        // W_l^T operands in chunks, double buffered as in the forward
        auto wload = [&](float (&w)[kB2Chunk][NT], int c0) {
#pragma unroll
            for (int st = 0; st < kB2Chunk; ++st) {
                const int k = 4 * (c0 + st) + k4;
#pragma unroll
                for (int h = 0; h < NT; ++h) {
                    const int m = (2 * h + wv) * 16 + c16;
                    w[st][h] = (k < no && m < nprev) ? Wl[m * no + k] : 0.0f;
                }
            }
        };
        auto wmul = [&](const float (&w)[kB2Chunk][NT], int c0) {
#pragma unroll
            for (int st = 0; st < kB2Chunk; ++st) {
                if (4 * (c0 + st) >= no) break;
                const int k = 4 * (c0 + st) + k4;
                const bool ok = k < no;
                const float za = ok ? sx[c16 * KP + k] : 0.0f, zda = ok ? sxd[c16 * KP + k] : 0.0f;
#pragma unroll
                for (int h = 0; h < NT; ++h) {
                    if ((2 * h + wv) * 16 >= nprev) break;
                    A1[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(za, w[st][h], A1[h], 0, 0, 0);
                    A2[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(zda, w[st][h], A2[h], 0, 0, 0);
                }
            }
        };
        {
            float wa[kB2Chunk][NT], wb[kB2Chunk][NT];
            const int nst = (no + 3) / 4;
            wload(wa, 0);
            for (int c0 = 0; c0 < nst; c0 += 2 * kB2Chunk) {
                if (c0 + kB2Chunk < nst) wload(wb, c0 + kB2Chunk);
                wmul(wa, c0);
                if (c0 + kB2Chunk >= nst) break;
                if (c0 + 2 * kB2Chunk < nst) wload(wa, c0 + 2 * kB2Chunk);
                wmul(wb, c0 + kB2Chunk);
            }
        }
#pragma unroll
        for (int h = 0; h < NT; ++h) {
            const int mt = 2 * h + wv;
            if (mt * 16 >= nprev) break;
            const int m = mt * 16 + c16;
            const f32x4 a1 = A1[h], a2 = A2[h];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = rb + i;
                const float d = m < nprev ? sdz[((l - 1) * kB2Tile + r) * WP + m] : 0.0f,
                            qq = m < nprev ? sqz[((l - 1) * kB2Tile + r) * WP + m] : 0.0f;
                nb[h][i] = a1[i] * d + a2[i] * qq;
                ndb[h][i] = a2[i] * d;
            }
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < NT; ++h) {
            const int mt = 2 * h + wv;
            if (mt * 16 >= nprev) break;
            const int m = mt * 16 + c16;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (m < nprev) { sx[(rb + i) * KP + m] = nb[h][i]; sxd[(rb + i) * KP + m] = ndb[h][i]; }
        }
        __syncthreads();
    }
}

struct GradArgs {
    const DevModel* dm;
    const float* scr;
    const int* spcount;
    int64_t base[kMaxSpecies];
    const int4* tiles;        // (species, layer, n0, m0) of 32 x 32 output tiles
    int nsplit;
    double* part;             // [nsplit][nparams]
};

// Parameter gradient of one layer of one species as GEMMs over the atoms on
// the matrix cores (v_mfma_f32_16x16x4_f32: FP32 operands, FP32 accumulate):
//   dL/dW[n][m] = sum_atoms zbar[n] x[m] + zdbar[n] xd[m],   dL/db[n] = sum zbar[n]
// with the bias as the extra column m = ni (x = 1, xd = 0).  One wavefront per
// (32 x 32 output tile, atom split), four atoms per instruction; FP32 over
// 64 atoms, then FP64 (the same rounding scheme as a 64-atom FP32 chunk).
typedef float floatx4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64) void k_grad_mfma(GradArgs A, int nparams) {
    const int4 tl = A.tiles[blockIdx.x];
    const int split = blockIdx.y;
    const int a = tl.x, l = tl.y, n0 = tl.z, m0 = tl.w;
    const NetDesc& D = A.dm->net[a];
    const int stride = bwd2_stride(D);
    int lo = 0, ni = D.nin;
    for (int p = 0; p < l; ++p) { lo += 2 * (ni + D.nodes[p]); ni = D.nodes[p]; }
    const int no = D.nodes[l];
    const int na = A.spcount[a];
    const int per = (na + A.nsplit - 1) / A.nsplit;
    const int r0 = min(na, split * per), r1 = min(na, r0 + per);
    const float* base = A.scr + A.base[a] + lo;
    const int lane = threadIdx.x, c16 = lane % 16, k4 = lane / 16;
    const int nA[2] = {n0 + c16, n0 + 16 + c16};          // A operand rows (outputs n)
    const int mB[2] = {m0 + c16, m0 + 16 + c16};          // B operand columns (inputs m, bias at ni)
    floatx4 acc[2][2];
    double dacc[2][2][4];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            acc[u][v] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < 4; ++q) dacc[u][v][q] = 0.0;
        }
    int steps = 0;
    // four k steps (16 atoms) per trip: every load of the trip is issued before its products
    for (int r = r0; r < r1; r += 16) {
        float z[4][2], zd[4][2], x[4][2], xd[4][2];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int row = r + 4 * t + k4;
            const bool ok = row < r1;
            const float* rr = base + (int64_t)(ok ? row : r0) * stride;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const bool okn = ok && nA[u] < no;
                z[t][u] = okn ? rr[2 * ni + nA[u]] : 0.0f;
                zd[t][u] = okn ? rr[2 * ni + no + nA[u]] : 0.0f;
                const int m = mB[u];
                x[t][u] = ok ? (m < ni ? rr[m] : (m == ni ? 1.0f : 0.0f)) : 0.0f;
                xd[t][u] = ok && m < ni ? rr[ni + m] : 0.0f;
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    acc[u][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(z[t][u], x[t][v], acc[u][v], 0, 0, 0);
                    acc[u][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(zd[t][u], xd[t][v], acc[u][v], 0, 0, 0);
                }
        steps += 4;
        if (steps == 16) {
            steps = 0;
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int v = 0; v < 2; ++v) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) dacc[u][v][q] += (double)acc[u][v][q];
                    acc[u][v] = floatx4{0.f, 0.f, 0.f, 0.f};
                }
        }
    }
    int bo = 0, wo = 0, prev = D.nin;
    for (int q = 0; q < l; ++q) bo += D.nodes[q];
    for (int q = 0; q < D.nlayer; ++q) wo += D.nodes[q];
    for (int q = 0; q < l; ++q) { wo += D.nodes[q] * prev; prev = D.nodes[q]; }
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int n = n0 + 16 * u + 4 * k4 + q, m = m0 + 16 * v + c16;
                if (n >= no || m > ni) continue;
                const int p = m < ni ? D.poff + wo + m * no + n : D.poff + bo + n;
                A.part[(size_t)split * nparams + p] = dacc[u][v][q] + (double)acc[u][v][q];
            }
}

// fixed-order sum over the atom splits
__global__ void k_grad_combine(int nparams, int nsplit, const double* __restrict__ part, double* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nparams) return;
    // the split partials are independent loads: issue them in groups, add in split order
    double s = 0.0;
    int k = 0;
    for (; k + 8 <= nsplit; k += 8) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = part[(size_t)(k + q) * nparams + p];
#pragma unroll
        for (int q = 0; q < 8; ++q) s += v[q];
    }
    for (; k < nsplit; ++k) s += part[(size_t)k * nparams + p];
    grad[p] = s;
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

// Device int arrays that only grow (a new batch of the same or smaller size
// reuses them: no allocation, and no implicit synchronisation of a free).
static hipError_t grow_ints(int*& p, int64_t& cap, int64_t want) {
    if (p && want <= cap) return hipSuccess;
    if (p) {
        const hipError_t e = hipFree(p);
        p = nullptr;
        cap = 0;
        if (e != hipSuccess) return e;
    }
    const int64_t c = std::max<int64_t>(want, 1) + std::max<int64_t>(want, 1) / 4;
    const hipError_t e = hipMalloc(&p, sizeof(int) * c);
    if (e == hipSuccess) cap = c;
    return e;
}

hipError_t species_ranks(ann_ctx* ctx) {
    if (ctx->rank_valid) return hipSuccess;
    const int N = ctx->natoms;
    hipError_t e;
    int64_t c1 = ctx->rank_cap, c2 = ctx->rank_cap, c3 = ctx->rank_cap;
    if ((e = grow_ints(ctx->d_rank, c1, N + 1)) != hipSuccess) return e;
    if ((e = grow_ints(ctx->d_rkflag, c2, N + 1)) != hipSuccess) return e;
    if ((e = grow_ints(ctx->d_rkpre, c3, N + 1)) != hipSuccess) return e;
    ctx->rank_cap = std::min(c1, std::min(c2, c3));
    if (!ctx->d_spcount && (e = hipMalloc(&ctx->d_spcount, sizeof(int) * kMaxSpecies)) != hipSuccess) return e;
    if (ctx->h_counts_valid) {
        // species known on the host (ann_set_batch_host): ranks in atom order,
        // as the device scans give them, and the counts, uploaded without kernels
        ctx->h_rank.resize(N);
        std::vector<int> next(ctx->S, 0);
        for (int i = 0; i < N; ++i) ctx->h_rank[i] = next[ctx->h_species[i]]++;
        if ((e = staged_upload(ctx, kStRank, ctx->d_rank, ctx->h_rank.data(), sizeof(int) * (size_t)N)) != hipSuccess)
            return e;
        if ((e = staged_upload(ctx, kStSpCount, ctx->d_spcount, ctx->h_species_count.data(), sizeof(int) * ctx->S)) !=
            hipSuccess)
            return e;
    }
    int* flag = ctx->d_rkflag;
    int* pre = ctx->d_rkpre;
    for (int s = 0; s < (ctx->h_counts_valid ? 0 : ctx->S); ++s) {
        k_flags<<<(N + 256) / 256, 256, 0, ctx->stream>>>(N, ctx->d_species, s, flag);
        if ((e = exclusive_scan(flag, pre, N + 1, ctx->stream)) != hipSuccess) return e;
        k_take_rank<<<(N + 256) / 256, 256, 0, ctx->stream>>>(N, ctx->d_species, s, pre, ctx->d_rank, ctx->d_spcount);
    }
    // the counts: from the host species given with the batch, else read back
    std::vector<int> counts(ctx->S);
    if (ctx->h_counts_valid) {
        counts = ctx->h_species_count;
    } else {
        e = hipMemcpyAsync(counts.data(), ctx->d_spcount, sizeof(int) * ctx->S, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return e;
    }
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
    if ((e = grow_ints(ctx->d_perm, ctx->perm_cap, N)) != hipSuccess) return e;
    if ((e = grow_ints(ctx->d_tiles, ctx->tiles_cap, (int64_t)tiles.size())) != hipSuccess) return e;
    if (!ctx->d_spend && (e = hipMalloc(&ctx->d_spend, sizeof(int) * 2 * kMaxSpecies)) != hipSuccess) return e;
    std::vector<int> be(base);
    be.insert(be.end(), end.begin(), end.end());
    if ((e = staged_upload(ctx, kStSpend, ctx->d_spend, be.data(), sizeof(int) * 2 * S)) != hipSuccess) return e;
    if ((e = staged_upload(ctx, kStTiles, ctx->d_tiles, tiles.data(), sizeof(int) * tiles.size())) != hipSuccess)
        return e;
    if (ctx->h_counts_valid) {
        std::vector<int> perm(N);
        for (int i = 0; i < N; ++i) perm[base[ctx->h_species[i]] + ctx->h_rank[i]] = i;
        if ((e = staged_upload(ctx, kStPerm, ctx->d_perm, perm.data(), sizeof(int) * (size_t)N)) != hipSuccess) return e;
    } else {
        k_perm<<<(N + 255) / 256, 256, 0, ctx->stream>>>(N, ctx->d_species, ctx->d_rank, ctx->d_spend, ctx->d_perm);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    // spend[] must hold the END of each species block: shift the uploaded (base, end) pair
    ctx->d_spend_end = ctx->d_spend + S;
    ctx->ntiles = (int)tiles.size() / 2;
    ctx->perm_valid = true;
    return hipSuccess;
}

hipError_t net_bwd2(ann_ctx* ctx, const float* G, const float* v, const double* gE, double* grad, bool centred) {
    hipError_t e = species_ranks(ctx);
    if (e != hipSuccess) return e;
    Bwd2Args A{};
    A.dm = ctx->dm;
    A.W = ctx->da.W;
    A.inpw = ctx->da.inpw;
    A.inpb = centred ? ctx->da.inpb_net : ctx->da.inpb;
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
    if (mfma_net_supported(ctx->hm)) {
        if ((e = species_perm(ctx)) != hipSuccess) return e;
        Launch L(ctx, "net_bwd2");
        int kmax = 1, wmax = 1, nlh = 1;
        for (int s = 0; s < ctx->S; ++s) {
            const NetDesc& D = ctx->hm.net[s];
            kmax = std::max(kmax, D.nin);
            nlh = std::max(nlh, D.nlayer - 1);
            for (int l = 0; l < D.nlayer; ++l) { kmax = std::max(kmax, D.nodes[l]); wmax = std::max(wmax, D.nodes[l]); }
        }
        A.kp = (kmax + 1) | 1;
        A.wp = (wmax + 1) | 1;
        A.nlh = nlh;
        const size_t smem = sizeof(float) * (2 * (size_t)kB2Tile * A.kp + 2 * (size_t)nlh * kB2Tile * A.wp);
        int hmax = 1;                                 // widest hidden layer
        for (int s = 0; s < ctx->S; ++s)
            for (int l = 0; l < ctx->hm.net[s].nlayer - 1; ++l) hmax = std::max(hmax, ctx->hm.net[s].nodes[l]);
        if (ctx->ntiles > 0) {
            void (*kb)(Bwd2Args, const int*, const int2*, const int*) = hmax <= 64 ? k_bwd2_mfma<2> : k_bwd2_mfma<4>;
            kb<<<ctx->ntiles, 128, smem, ctx->stream>>>(A, ctx->d_perm, reinterpret_cast<const int2*>(ctx->d_tiles),
                                                       ctx->d_spend_end);
        }
    } else {
        Launch L(ctx, "net_bwd2");
        k_net_bwd2<<<(A.N + 3) / 4, 64, 0, ctx->stream>>>(A);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // 32 x 32 output tiles of every species and layer (bias = column ni)
    if (ctx->grad_tiles.empty()) {
        for (int s = 0; s < ctx->S; ++s) {
            const NetDesc& D = ctx->hm.net[s];
            int ni = D.nin;
            for (int l = 0; l < D.nlayer; ++l) {
                const int no = D.nodes[l];
                for (int n0 = 0; n0 < no; n0 += 32)
                    for (int m0 = 0; m0 <= ni; m0 += 32) ctx->grad_tiles.insert(ctx->grad_tiles.end(), {s, l, n0, m0});
                ni = no;
            }
        }
    }
    const int ntiles = (int)ctx->grad_tiles.size() / 4;
    // atom splits: about 2048 wavefronts in flight, fixed for a given model
    const int nsplit = std::max(8, std::min(64, 2048 / std::max(ntiles, 1)));
    if (!ctx->d_grad_tiles) {
        if ((e = hipMalloc(&ctx->d_grad_tiles, sizeof(int) * ctx->grad_tiles.size())) != hipSuccess) return e;
        if ((e = hipMemcpy(ctx->d_grad_tiles, ctx->grad_tiles.data(), sizeof(int) * ctx->grad_tiles.size(),
                           hipMemcpyHostToDevice)) != hipSuccess)
            return e;
    }
    if (ctx->gsplit_cap < (int64_t)nsplit * ctx->nparams) {
        if (ctx->d_gsplit) hipFree(ctx->d_gsplit);
        if ((e = hipMalloc(&ctx->d_gsplit, sizeof(double) * nsplit * ctx->nparams)) != hipSuccess) return e;
        ctx->gsplit_cap = (int64_t)nsplit * ctx->nparams;
    }
    GradArgs R{};
    R.dm = ctx->dm;
    R.scr = ctx->d_scr;
    R.spcount = ctx->d_spcount;
    for (int s = 0; s < kMaxSpecies; ++s) R.base[s] = A.base[s];
    R.tiles = reinterpret_cast<const int4*>(ctx->d_grad_tiles);
    R.nsplit = nsplit;
    R.part = ctx->d_gsplit;
    {
        Launch L(ctx, "grad_reduce");
        k_grad_mfma<<<dim3(ntiles, nsplit), 64, 0, ctx->stream>>>(R, ctx->nparams);
        k_grad_combine<<<(ctx->nparams + 255) / 256, 256, 0, ctx->stream>>>(ctx->nparams, nsplit, ctx->d_gsplit, grad);
    }
    return hipGetLastError();
}

// ---- loss: one block per structure, FP64 ----
// Energy loss on dU = (E - Eref) norm(N) (NNPTEFR::error_cost_o1,
// src/nnp/nnptefr2.cpp:1336-1372): MSE dU^2/2, MAE |dU|, HUBER
// delta^2 (sqrt(1 + a^2) - 1), ASINH delta^2 (1 - sqrt(1 + a^2) + a asinh a),
// a = dU / delta; norm IDENTITY 1, LINEAR 1/N, SQRT, CBRT, LOG 1/(ln N + 1)
// (:602-612).  Per-structure weights scale both terms (a structure without
// reference forces, as the Taylor-augmented copies of :2466-2473, has force
// weight 0).
struct LossCfg {
    int loss, norm;
    double delta;
    const double* wEs;
    const double* wFs;
};

__device__ inline void energy_loss(const LossCfg& c, double dU, double& l, double& g) {
    switch (c.loss) {
        case ANN_LOSS_MAE: l = fabs(dU); g = (double)((dU > 0) - (dU < 0)); break;
        case ANN_LOSS_HUBER: {
            const double a = dU / c.delta, q = sqrt(1.0 + a * a);
            l = c.delta * c.delta * (q - 1.0);
            g = dU / q;
        } break;
        case ANN_LOSS_ASINH: {
            const double a = dU / c.delta, q = sqrt(1.0 + a * a), as = asinh(a);
            l = c.delta * c.delta * (1.0 - q + a * as);
            g = c.delta * as;
        } break;
        default: l = 0.5 * dU * dU; g = dU; break;
    }
}

__device__ inline double energy_norm(int norm, double n) {
    switch (norm) {
        case ANN_NORM_IDENTITY: return 1.0;
        case ANN_NORM_SQRT: return 1.0 / sqrt(n);
        case ANN_NORM_CBRT: return 1.0 / cbrt(n);
        case ANN_NORM_LOG: return 1.0 / (log(n) + 1.0);
        default: return 1.0 / n;
    }
}

__global__ __launch_bounds__(256) void k_loss(const int* __restrict__ aoff, const double* __restrict__ E,
                                              const double* __restrict__ F, const double* __restrict__ Eref,
                                              const double* __restrict__ Fref, double wE0, double wF0, int B,
                                              double* __restrict__ lossE, double* __restrict__ lossF,
                                              double* __restrict__ gE, float* __restrict__ uneg, LossCfg cfg) {
    __shared__ double red[256];
    const int s = blockIdx.x, t = threadIdx.x;
    const int a0 = aoff[s], a1 = aoff[s + 1], N = a1 - a0;
    const double wE = wE0 * (cfg.wEs ? cfg.wEs[s] : 1.0), wF = wF0 * (cfg.wFs ? cfg.wFs[s] : 1.0);
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
        const double nrm = energy_norm(cfg.norm, (double)N);
        double l, g;
        energy_loss(cfg, (E[s] - Eref[s]) * nrm, l, g);
        lossE[s] = wE * l / B;
        lossF[s] = wF * red[0] / (3.0 * N) / B;
        gE[s] = wE * g * nrm / B;
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

// step t = *tdev + 1 (the counter lives on the device so that a captured
// training step advances it on every replay; the weight-split launch after the
// update increments it)
__global__ void k_adam(int n, double* __restrict__ p, double* __restrict__ m1, double* __restrict__ m2,
                       const double* __restrict__ g, double gamma, const int* __restrict__ tdev, double eps,
                       float* __restrict__ W) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double b1 = 0.9, b2 = 0.999;
    const double t = (double)(*tdev + 1);
    const double b1ii = 1.0 / (1.0 - pow(b1, t)), b2ii = 1.0 / (1.0 - pow(b2, t));
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
    // the network's species tiles first: their host set-up then overlaps the
    // G kernel instead of holding the network kernel back
    if (form != ANN_FORM_PLAIN && (e = species_perm(ctx)) != hipSuccess) return e;
    if (!ctx->G_valid) {
        if ((e = network_rows(ctx, form, ctx->d_G)) != hipSuccess) return e;
        ctx->G_valid = true;
    }
    if ((e = net_fwd(ctx, form, ctx->d_G, ctx->d_Eatom, ctx->d_dEdG, true)) != hipSuccess) return e;
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
        const LossCfg cfg{ctx->loss_type, ctx->loss_norm, ctx->loss_delta, ctx->d_wEs, ctx->d_wFs};
        k_loss<<<B, 256, 0, ctx->stream>>>(ctx->d_atom_off, ctx->d_E, ctx->d_F, Eref, Fref, wE, wF, B, lE, lF,
                                           ctx->d_gE, ctx->d_u, cfg);
        k_loss_total<<<1, 256, 0, ctx->stream>>>(B, lE, lF, loss ? loss : lt);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    (void)N;
    if ((e = symm_tangent(ctx, form, ctx->d_u, ctx->d_v)) != hipSuccess) return e;
    if ((e = net_bwd2(ctx, ctx->d_G, ctx->d_v, ctx->d_gE, g, true)) != hipSuccess) return e;
    if (gamma > 0) {
        Launch L(ctx, "adam");
        k_adam<<<(ctx->nparams + 255) / 256, 256, 0, ctx->stream>>>(ctx->nparams, ctx->da.params, ctx->da.adam_m,
                                                                    ctx->da.adam_v, g, gamma, ctx->d_adam_t, 1e-16,
                                                                    ctx->da.W);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // new FP32 weights: refresh the matrix-core operand tables; the same
        // launch advances the Adam step counter
        if ((e = split_weights(ctx, ctx->d_adam_t)) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace ann
