// symm.hip -- symmetry functions G (NNP::symm), the force adjoint
// P[s] = dE_i/dr_ij (NNP::force / NNP::compute) and the forward-mode basis
// tangent used by the force-loss weight gradient.
//
// Reference loops: src/nnp/nnp.cpp:333-375 (symm), 461-520 and 601-665
// (force/compute; angular assembly :503-516), basis functions
// src/nnp/basis_radial.cpp:263-376 and src/nnp/basis_angular.cpp:371-563.
//
// Two forms, one wavefront (block of 64) per centre atom in both:
//
//  PLAIN (reference-shaped): neighbours in canonical order in LDS; lane j
//    loops k > j exactly like the reference (triangular, unbalanced),
//    evaluates every angular function of the triplet from scratch (cutoffs,
//    exp and pow per function), accumulates G in lane-private LDS columns and
//    scatters the k-side of the force through LDS atomics.
//
//  FAST: neighbours regrouped by species in LDS so that every (species j,
//    species k) block of the angular basis is a rectangle (j, k in different
//    species) or a triangle (same species) of index space; the exponential of
//    the GAUSS family factorises over the pair, exp(-(rij^2+rik^2)/eta^2) =
//    E(rij) E(rik), so E*fc is computed once per neighbour and per distinct
//    width; blocks whose functions form a {width} x {(lambda,zeta)} grid run a
//    kernel specialised on the grid shape with integer powers by repeated
//    squaring and register accumulators.
//      forward: unordered pairs flattened over the 64 lanes (triangle folded
//        into an n x (n-1)/2 rectangle), wave-reduced per block;
//      adjoint / tangent: skewed schedule, lane = owner j, partner
//        k = (j + t) mod n at step t, so that lanes of one instruction touch
//        distinct partners; each unordered triplet is evaluated once and its
//        two slot contributions are accumulated in registers (owner) and LDS
//        (partner) -- no global atomics, deterministic.
//  Both are FP32 VALU-bound (the roofline that prices them is the FP32 peak).
#include "basis_dev.h"

namespace ann {

struct SymmArgs {
    const DevModel* dm;
    const float* rad_rs;
    const float* rad_eta;
    const float* ang_eta;
    const float* ang_zeta;
    const float* ang_lam;
    int N, ldg, maxn;
    const int* species;
    const int* off;
    const float4* geo;
    const uint8_t* nsp;
    const int* nidx;
    float* G;
    const float* dEdG;
    float* P;
    const float* u;
    float* v;
};

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// =========================================================== PLAIN form ====

// mode 0: G, 1: adjoint P, 2: tangent v
template <int MODE>
__global__ __launch_bounds__(64) void k_plain(SymmArgs A) {
    extern __shared__ float smem[];
    const int i = blockIdx.x, lane = threadIdx.x;
    if (i >= A.N) return;
    const DevModel& M = *A.dm;
    const int a = A.species[i], o = A.off[i], n = A.off[i + 1] - o;
    const int S = M.S, NP = M.NP, ldg = A.ldg, maxn = A.maxn;
    float4* sg = reinterpret_cast<float4*>(smem);          // (x, y, z, r) canonical order
    float* sfc = reinterpret_cast<float*>(sg + maxn);
    float* sgc = sfc + maxn;
    int* ssp = reinterpret_cast<int*>(sgc + maxn);
    float* rest = reinterpret_cast<float*>(ssp + maxn);
    constexpr int LP = 65;                                  // padded lane stride
    for (int t = lane; t < n; t += kWave) {
        const float4 gq = A.geo[o + t];
        sg[t] = gq;
        float f, g;
        cutoff_fg(M.cutoff, M.rc, M.rci, M.pirc, gq.w, f, g);
        sfc[t] = f;
        sgc[t] = g;
        ssp[t] = A.nsp[o + t];
    }
    float* acc = rest;                                      // [ldg][LP]   (modes 0, 2)
    float* sd = rest;                                       // [ldg]       (mode 1)
    float* sP = rest + ldg;                                 // [maxn][3]   (mode 1)
    float4* st = reinterpret_cast<float4*>(rest + ldg * LP);  // [maxn] (tx,ty,tz, r^.t) (mode 2)
    if (MODE == 1) {
        for (int m = lane; m < ldg; m += kWave) sd[m] = A.dEdG[(size_t)i * ldg + m];
        for (int t = lane; t < 3 * n; t += kWave) sP[t] = 0.0f;
    } else {
        for (int m = 0; m < ldg; ++m) acc[m * LP + lane] = 0.0f;
    }
    if (MODE == 2) {
        for (int t = lane; t < n; t += kWave) {
            const int j = A.nidx[o + t];
            const float tx = A.u[3 * i] - A.u[3 * j], ty = A.u[3 * i + 1] - A.u[3 * j + 1],
                        tz = A.u[3 * i + 2] - A.u[3 * j + 2];
            const float4 gq = sg[t];
            st[t] = make_float4(tx, ty, tz, (gq.x * tx + gq.y * ty + gq.z * tz) / gq.w);
        }
    }
    __syncthreads();
    // radial
    for (int j = lane; j < n; j += kWave) {
        const RadBlock rb = M.rad[a * S + ssp[j]];
        const float r = sg[j].w;
        float amp = 0.0f;
        for (int m = 0; m < rb.n; ++m) {
            float v, d;
            radial_fd(rb.name, A.rad_eta[rb.par + m], A.rad_rs[rb.par + m], r, sfc[j], sgc[j], v, d);
            if (MODE == 0) acc[(rb.goff + m) * LP + lane] += v;
            if (MODE == 1) amp += sd[rb.goff + m] * d;
            if (MODE == 2) acc[(rb.goff + m) * LP + lane] += d * st[j].w;
        }
        if (MODE == 1) {
            const float4 gq = sg[j];
            atomicAdd(&sP[3 * j], amp * gq.x / r);
            atomicAdd(&sP[3 * j + 1], amp * gq.y / r);
            atomicAdd(&sP[3 * j + 2], amp * gq.z / r);
        }
    }
    // angular: lane j, partners k > j (reference loop order)
    for (int j = lane; j < n; j += kWave) {
        const float4 gj = sg[j];
        const float rj = gj.w;
        float pjx = 0, pjy = 0, pjz = 0;
        for (int k = j + 1; k < n; ++k) {
            const float4 gk = sg[k];
            const float rk = gk.w;
            // cutoffs recomputed per triplet, as BasisA::symm/force do
            float fj, gcj, fk, gck;
            cutoff_fg(M.cutoff, M.rc, M.rci, M.pirc, rj, fj, gcj);
            cutoff_fg(M.cutoff, M.rc, M.rci, M.pirc, rk, fk, gck);
            const float c = (gj.x * gk.x + gj.y * gk.y + gj.z * gk.z) / (rj * rk);
            const AngBlock& ab = M.ang[a * NP + pair_idx(ssp[j], ssp[k])];
            float sA = 0, sBj = 0, sBk = 0;
            for (int m = 0; m < ab.n; ++m) {
                const float eta = A.ang_eta[ab.par + m];
                float G, dc, d0, d1;
                angular_fd(ab.name, 1.0f / (eta * eta), A.ang_zeta[ab.par + m], 0.5f * A.ang_lam[ab.par + m],
                           rj, rk, c, fj, fk, gcj, gck, G, dc, d0, d1);
                if (MODE == 0) acc[(ab.goff + m) * LP + lane] += G;
                if (MODE == 1) {
                    const float de = sd[ab.goff + m];
                    sA += de * dc;
                    sBj += de * d0;
                    sBk += de * d1;
                }
                if (MODE == 2) {
                    const float4 tj = st[j], tk = st[k];
                    const float cdot = ((gk.x * tj.x + gk.y * tj.y + gk.z * tj.z) / rk - c * tj.w) / rj +
                                       ((gj.x * tk.x + gj.y * tk.y + gj.z * tk.z) / rj - c * tk.w) / rk;
                    acc[(ab.goff + m) * LP + lane] += dc * cdot + d0 * tj.w + d1 * tk.w;
                }
            }
            if (MODE == 1) {
                const float hjx = gj.x / rj, hjy = gj.y / rj, hjz = gj.z / rj;
                const float hkx = gk.x / rk, hky = gk.y / rk, hkz = gk.z / rk;
                pjx += sA * (hkx - c * hjx) / rj + sBj * hjx;
                pjy += sA * (hky - c * hjy) / rj + sBj * hjy;
                pjz += sA * (hkz - c * hjz) / rj + sBj * hjz;
                atomicAdd(&sP[3 * k], sA * (hjx - c * hkx) / rk + sBk * hkx);
                atomicAdd(&sP[3 * k + 1], sA * (hjy - c * hky) / rk + sBk * hky);
                atomicAdd(&sP[3 * k + 2], sA * (hjz - c * hkz) / rk + sBk * hkz);
            }
        }
        if (MODE == 1) {
            atomicAdd(&sP[3 * j], pjx);
            atomicAdd(&sP[3 * j + 1], pjy);
            atomicAdd(&sP[3 * j + 2], pjz);
        }
    }
    __syncthreads();
    if (MODE == 1) {
        for (int t = lane; t < 3 * n; t += kWave) A.P[3 * (size_t)o + t] = sP[t];
    } else {
        float* out = (MODE == 0 ? A.G : A.v) + (size_t)i * ldg;
        for (int m = lane; m < ldg; m += kWave) {
            float s = 0.0f;
            for (int l = 0; l < kWave; ++l) s += acc[m * LP + l];
            out[m] = s;
        }
    }
}

// ============================================================ FAST form ====

struct FastSmem {
    float4* g;      // [maxn] (hx, hy, hz, r), species-sorted
    float* ef;      // [U][maxn] exp(-r^2 ieta2_u) fc
    float* efd;     // [U][maxn] d/dr of ef
    float* fc;      // [maxn]
    float* gc;      // [maxn]
    int* canon;     // [maxn] canonical slot of each sorted entry
    int* seg;       // [kMaxSpecies + 1]
    float* extra;   // kernel specific
};

// Regroup the neighbours of atom i by species (stable) and precompute the
// per-neighbour factors.  Returns nothing; seg[] gives the species segments.
__device__ void fast_load(const SymmArgs& A, const DevModel& M, int i, int a, int o, int n, FastSmem& sm,
                          bool need_deriv) {
    const int lane = threadIdx.x, S = M.S, maxn = A.maxn, U = M.nunion[a];
    int cnt[kMaxSpecies];
#pragma unroll
    for (int b = 0; b < kMaxSpecies; ++b) cnt[b] = 0;
    for (int q = 0; q < n; q += kWave) {
        const int t = q + lane;
        const int sp = t < n ? (int)A.nsp[o + t] : -1;
#pragma unroll
        for (int b = 0; b < kMaxSpecies; ++b)
            if (b < S) cnt[b] += __popcll(__ballot(sp == b));
    }
    int base[kMaxSpecies];
    int acc = 0;
#pragma unroll
    for (int b = 0; b < kMaxSpecies; ++b) {
        base[b] = acc;
        if (b < S) acc += cnt[b];
    }
#pragma unroll
    for (int b = 0; b < kMaxSpecies; ++b)
        if (b < S && lane == b) sm.seg[b] = base[b];
    if (lane == S) sm.seg[S] = acc;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int q = 0; q < n; q += kWave) {
        const int t = q + lane;
        const int sp = t < n ? (int)A.nsp[o + t] : -1;
        int pos = 0;
#pragma unroll
        for (int b = 0; b < kMaxSpecies; ++b) {
            if (b < S) {
                const uint64_t mk = __ballot(sp == b);
                if (sp == b) pos = base[b] + __popcll(mk & lt);
                base[b] += __popcll(mk);
            }
        }
        if (t < n) {
            const float4 gq = A.geo[o + t];
            const float r = gq.w, ir = 1.0f / r;
            sm.g[pos] = make_float4(gq.x * ir, gq.y * ir, gq.z * ir, r);
            sm.canon[pos] = t;
            float f, gcut;
            cutoff_fg(M.cutoff, M.rc, M.rci, M.pirc, r, f, gcut);
            sm.fc[pos] = f;
            sm.gc[pos] = gcut;
            for (int uu = 0; uu < U; ++uu) {
                const float ie = M.uieta2[a][uu];
                const float e = __expf(-ie * r * r);
                sm.ef[uu * maxn + pos] = e * f;
                if (need_deriv) sm.efd[uu * maxn + pos] = e * (-2.0f * r * ie * f + gcut);
            }
        }
    }
}

// ---- forward / tangent over one grid block (unordered pairs, flattened) ----
template <int NG, int NK, int MODE>
__device__ void fast_block_flat(const AngBlock& B, const FastSmem& sm, int maxn, int sb, int nb, int sc,
                                int nc, bool tri, const float4* tv, float* out) {
    const int lane = threadIdx.x;
    float lamh[NK], zeta[NK];
    int zi[NK], gu[NG];
#pragma unroll
    for (int k = 0; k < NK; ++k) { lamh[k] = B.lamh[k]; zeta[k] = B.zeta[k]; zi[k] = B.zint[k]; }
#pragma unroll
    for (int g = 0; g < NG; ++g) gu[g] = B.gu[g] * maxn;
    float acc[NG][NK];
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int k = 0; k < NK; ++k) acc[g][k] = 0.0f;
    const int h = tri ? (nb - 1) / 2 : 0;
    const int cols = tri ? h : nc;
    const int main_total = tri ? nb * h : nb * nc;
    const int total = tri ? nb * (nb - 1) / 2 : nb * nc;
    int j = 0, t = 0;
    if (cols > 0) { j = lane / cols; t = lane % cols; }
    const int dj = cols > 0 ? kWave / cols : 0, dt = cols > 0 ? kWave % cols : 0;
    for (int tau = lane; tau < total; tau += kWave) {
        int jj, kk;
        if (tau < main_total) {
            jj = j;
            kk = tri ? j + 1 + t : t;
            if (tri && kk >= nb) kk -= nb;
        } else {
            jj = tau - main_total;
            kk = jj + (nb >> 1);
        }
        jj += sb;
        kk += sc;
        const float4 gj = sm.g[jj], gk = sm.g[kk];
        const float c = gj.x * gk.x + gj.y * gk.y + gj.z * gk.z;
        if (MODE == 0) {
            float p[NK];
#pragma unroll
            for (int k = 0; k < NK; ++k) p[k] = upow(fabsf(0.5f + lamh[k] * c), zi[k], zeta[k]);
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const float D = sm.ef[gu[g] + jj] * sm.ef[gu[g] + kk];
#pragma unroll
                for (int k = 0; k < NK; ++k) acc[g][k] += p[k] * D;
            }
        } else {
            const float4 tj = tv[jj], tk = tv[kk];
            const float cdot = ((gk.x * tj.x + gk.y * tj.y + gk.z * tj.z) - c * tj.w) / gj.w +
                               ((gj.x * tk.x + gj.y * tk.y + gj.z * tk.z) - c * tk.w) / gk.w;
            float p[NK], dp[NK];
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                const float cw = fabsf(0.5f + lamh[k] * c);
                const float w1 = upow(cw, zi[k] >= 1 ? zi[k] - 1 : -1, zeta[k] - 1.0f);
                p[k] = w1 * cw;
                dp[k] = zeta[k] * lamh[k] * w1;
            }
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const float ej = sm.ef[gu[g] + jj], ek = sm.ef[gu[g] + kk];
                const float rdot = sm.efd[gu[g] + jj] * ek * tj.w + ej * sm.efd[gu[g] + kk] * tk.w;
                const float D = ej * ek * cdot;
#pragma unroll
                for (int k = 0; k < NK; ++k) acc[g][k] += dp[k] * D + p[k] * rdot;
            }
        }
        t += dt;
        j += dj;
        if (t >= cols) { t -= cols; j += 1; }
    }
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            const float s = wave_sum(acc[g][k]);
            if (lane == 0) out[B.goff + B.outm[g * NK + k]] = s;
        }
}

// ---- adjoint over one grid block: skewed owner/partner schedule ----
template <int NG, int NK>
__device__ void fast_block_adj(const AngBlock& B, const FastSmem& sm, int maxn, int sb, int nb, int sc,
                               int nc, bool tri, const float* sd, float* sP) {
    const int lane = threadIdx.x;
    float lamh[NK], zeta[NK], dl[NG][NK];
    int zi[NK], gu[NG];
#pragma unroll
    for (int k = 0; k < NK; ++k) { lamh[k] = B.lamh[k]; zeta[k] = B.zeta[k]; zi[k] = B.zint[k]; }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        gu[g] = B.gu[g] * maxn;
#pragma unroll
        for (int k = 0; k < NK; ++k) dl[g][k] = sd[B.goff + B.outm[g * NK + k]];
    }
    // owners: the smaller segment (triangle: all items)
    int so = sb, no = nb, spp = sc, np = nc;
    if (!tri && nb > nc) { so = sc; no = nc; spp = sb; np = nb; }
    const int Q = tri ? 1 : np / no;
    const int items = no * Q;
    const int h = tri ? (no - 1) / 2 : 0;
    for (int r0 = 0; r0 < items; r0 += kWave) {
        const int id = r0 + lane;
        const bool act = id < items;
        const int j = act ? id % no : 0;
        const int q = act ? id / no : 0;
        const int J = so + j;
        const float4 gj = sm.g[J];
        float ej[NG], edj[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) { ej[g] = sm.ef[gu[g] + J]; edj[g] = sm.efd[gu[g] + J]; }
        float px = 0.0f, py = 0.0f, pz = 0.0f;
        int len, koff;
        if (tri) {
            len = h + ((no & 1) == 0 && j < (no >> 1) ? 1 : 0);
            koff = 1;
        } else {
            len = (q < Q - 1) ? no : np - (Q - 1) * no;
            koff = j + q * no;
        }
        int lmax = tri ? h + ((no & 1) == 0 ? 1 : 0) : (np - (Q - 1) * no);
        for (int t = 0; t < lmax; ++t) {
            if (act && t < len) {
                int k = tri ? j + koff + t : koff + t;
                if (k >= np) k -= np;
                const int K = spp + k;
                const float4 gk = sm.g[K];
                const float c = gj.x * gk.x + gj.y * gk.y + gj.z * gk.z;
                float p[NK], dp[NK];
#pragma unroll
                for (int kk = 0; kk < NK; ++kk) {
                    const float cw = fabsf(0.5f + lamh[kk] * c);
                    const float w1 = upow(cw, zi[kk] >= 1 ? zi[kk] - 1 : -1, zeta[kk] - 1.0f);
                    p[kk] = w1 * cw;
                    dp[kk] = zeta[kk] * lamh[kk] * w1;
                }
                float Ac = 0.0f, Bj = 0.0f, Bk = 0.0f;
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    float Sg = 0.0f, Tg = 0.0f;
#pragma unroll
                    for (int kk = 0; kk < NK; ++kk) {
                        Sg += dl[g][kk] * p[kk];
                        Tg += dl[g][kk] * dp[kk];
                    }
                    const float ek = sm.ef[gu[g] + K], edk = sm.efd[gu[g] + K];
                    Ac += Tg * ej[g] * ek;
                    Bj += Sg * edj[g] * ek;
                    Bk += Sg * ej[g] * edk;
                }
                const float aj = Ac / gj.w, ak = Ac / gk.w;
                px += aj * (gk.x - c * gj.x) + Bj * gj.x;
                py += aj * (gk.y - c * gj.y) + Bj * gj.y;
                pz += aj * (gk.z - c * gj.z) + Bj * gj.z;
                atomicAdd(&sP[3 * K], ak * (gj.x - c * gk.x) + Bk * gk.x);
                atomicAdd(&sP[3 * K + 1], ak * (gj.y - c * gk.y) + Bk * gk.y);
                atomicAdd(&sP[3 * K + 2], ak * (gj.z - c * gk.z) + Bk * gk.z);
            }
        }
        if (act) {
            atomicAdd(&sP[3 * J], px);
            atomicAdd(&sP[3 * J + 1], py);
            atomicAdd(&sP[3 * J + 2], pz);
        }
    }
}

template <int MODE>
__device__ void dispatch_flat(const AngBlock& B, const FastSmem& sm, int maxn, int sb, int nb, int sc, int nc,
                              bool tri, const float4* tv, float* out) {
    if (B.ng == 1 && B.nk == 8) fast_block_flat<1, 8, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
    else if (B.ng == 4 && B.nk == 4) fast_block_flat<4, 4, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
    else if (B.ng == 3 && B.nk == 4) fast_block_flat<3, 4, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
    else if (B.ng == 2 && B.nk == 4) fast_block_flat<2, 4, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
    else if (B.ng == 1 && B.nk == 4) fast_block_flat<1, 4, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
    else if (B.ng == 1 && B.nk == 2) fast_block_flat<1, 2, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
    else if (B.ng == 1 && B.nk == 1) fast_block_flat<1, 1, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
    else if (B.ng == 2 && B.nk == 2) fast_block_flat<2, 2, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
    else if (B.ng == 4 && B.nk == 8) fast_block_flat<4, 8, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
    else if (B.ng == 2 && B.nk == 8) fast_block_flat<2, 8, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
    else if (B.ng == 4 && B.nk == 2) fast_block_flat<4, 2, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
    else if (B.ng == 3 && B.nk == 2) fast_block_flat<3, 2, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
    else if (B.ng == 3 && B.nk == 8) fast_block_flat<3, 8, MODE>(B, sm, maxn, sb, nb, sc, nc, tri, tv, out);
}

__device__ void dispatch_adj(const AngBlock& B, const FastSmem& sm, int maxn, int sb, int nb, int sc, int nc,
                             bool tri, const float* sd, float* sP) {
    if (B.ng == 1 && B.nk == 8) fast_block_adj<1, 8>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
    else if (B.ng == 4 && B.nk == 4) fast_block_adj<4, 4>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
    else if (B.ng == 3 && B.nk == 4) fast_block_adj<3, 4>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
    else if (B.ng == 2 && B.nk == 4) fast_block_adj<2, 4>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
    else if (B.ng == 1 && B.nk == 4) fast_block_adj<1, 4>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
    else if (B.ng == 1 && B.nk == 2) fast_block_adj<1, 2>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
    else if (B.ng == 1 && B.nk == 1) fast_block_adj<1, 1>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
    else if (B.ng == 2 && B.nk == 2) fast_block_adj<2, 2>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
    else if (B.ng == 4 && B.nk == 8) fast_block_adj<4, 8>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
    else if (B.ng == 2 && B.nk == 8) fast_block_adj<2, 8>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
    else if (B.ng == 4 && B.nk == 2) fast_block_adj<4, 2>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
    else if (B.ng == 3 && B.nk == 2) fast_block_adj<3, 2>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
    else if (B.ng == 3 && B.nk == 8) fast_block_adj<3, 8>(B, sm, maxn, sb, nb, sc, nc, tri, sd, sP);
}

// shapes handled by the dispatchers above
__host__ __device__ inline bool fast_shape_ok(int ng, int nk) {
    return (ng == 1 && (nk == 1 || nk == 2 || nk == 4 || nk == 8)) ||
           ((ng == 2 || ng == 3 || ng == 4) && (nk == 2 || nk == 4 || nk == 8));
}

// radial part, owner = lane over the neighbours of one species segment
template <int MODE>
__device__ void fast_radial(const SymmArgs& A, const DevModel& M, int a, const FastSmem& sm, const float4* tv,
                            const float* sd, float* sP, float* out) {
    const int lane = threadIdx.x, S = M.S;
    for (int b = 0; b < S; ++b) {
        const RadBlock rb = M.rad[a * S + b];
        if (rb.n == 0) continue;
        const int s0 = sm.seg[b], s1 = sm.seg[b + 1];
        if (MODE == 1) {
            for (int J = s0 + lane; J < s1; J += kWave) {
                const float4 gj = sm.g[J];
                float amp = 0.0f;
                for (int m = 0; m < rb.n; ++m) {
                    float v, d;
                    radial_fd(rb.name, A.rad_eta[rb.par + m], A.rad_rs[rb.par + m], gj.w, sm.fc[J], sm.gc[J], v, d);
                    amp += sd[rb.goff + m] * d;
                }
                sP[3 * J] += amp * gj.x;
                sP[3 * J + 1] += amp * gj.y;
                sP[3 * J + 2] += amp * gj.z;
            }
            continue;
        }
        float acc[kMaxFunc];
#pragma unroll
        for (int m = 0; m < kMaxFunc; ++m) acc[m] = 0.0f;
        for (int J = s0 + lane; J < s1; J += kWave) {
            const float r = sm.g[J].w;
            const float w = MODE == 2 ? tv[J].w : 0.0f;
#pragma unroll
            for (int m = 0; m < kMaxFunc; ++m) {
                if (m < rb.n) {
                    float v, d;
                    radial_fd(rb.name, A.rad_eta[rb.par + m], A.rad_rs[rb.par + m], r, sm.fc[J], sm.gc[J], v, d);
                    acc[m] += MODE == 0 ? v : d * w;
                }
            }
        }
#pragma unroll
        for (int m = 0; m < kMaxFunc; ++m) {
            if (m < rb.n) {
                const float s = wave_sum(acc[m]);
                if (lane == 0) out[rb.goff + m] = s;
            }
        }
    }
}

// mode 0: G, 1: adjoint P, 2: tangent v
template <int MODE>
__global__ __launch_bounds__(64) void k_fast(SymmArgs A) {
    extern __shared__ float smem[];
    const int i = blockIdx.x, lane = threadIdx.x;
    if (i >= A.N) return;
    const DevModel& M = *A.dm;
    const int a = A.species[i], o = A.off[i], n = A.off[i + 1] - o;
    const int S = M.S, NP = M.NP, ldg = A.ldg, maxn = A.maxn, U = M.nunion[a];
    FastSmem sm;
    sm.g = reinterpret_cast<float4*>(smem);
    sm.ef = reinterpret_cast<float*>(sm.g + maxn);
    sm.efd = sm.ef + kMaxUnionEta * maxn;
    sm.fc = sm.efd + (MODE == 0 ? 0 : kMaxUnionEta * maxn);
    sm.gc = sm.fc + maxn;
    sm.canon = reinterpret_cast<int*>(sm.gc + maxn);
    sm.seg = sm.canon + maxn;
    sm.extra = reinterpret_cast<float*>(sm.seg + 16);
    (void)U;
    float* rowbuf = sm.extra;                                   // [ldg] output row (modes 0, 2) / dE/dG (mode 1)
    float* sP = sm.extra + ldg;                                 // [maxn][3] (mode 1)
    float4* tv = reinterpret_cast<float4*>(sm.extra + ldg + (ldg & 3 ? 4 - (ldg & 3) : 0));  // [maxn] (mode 2)
    fast_load(A, M, i, a, o, n, sm, MODE != 0);
    if (MODE == 1) {
        for (int m = lane; m < ldg; m += kWave) rowbuf[m] = A.dEdG[(size_t)i * ldg + m];
    } else {
        for (int m = lane; m < ldg; m += kWave) rowbuf[m] = 0.0f;
    }
    __syncthreads();
    if (MODE == 1) {
        for (int t = lane; t < 3 * n; t += kWave) sP[t] = 0.0f;
    }
    if (MODE == 2) {
        // per neighbour tangent t_j = u_i - u_j and r^_j . t_j, species-sorted
        for (int J = lane; J < n; J += kWave) {
            const int j = A.nidx[o + sm.canon[J]];
            const float tx = A.u[3 * i] - A.u[3 * j], ty = A.u[3 * i + 1] - A.u[3 * j + 1],
                        tz = A.u[3 * i + 2] - A.u[3 * j + 2];
            const float4 gq = sm.g[J];
            tv[J] = make_float4(tx, ty, tz, gq.x * tx + gq.y * ty + gq.z * tz);
        }
    }
    __syncthreads();
    fast_radial<MODE>(A, M, a, sm, tv, rowbuf, sP, rowbuf);
    __syncthreads();
    for (int b = 0; b < S; ++b) {
        for (int c = b; c < S; ++c) {
            const AngBlock& B = M.ang[a * NP + pair_idx(b, c)];
            if (B.n == 0) continue;
            const int sb = sm.seg[b], nb = sm.seg[b + 1] - sb;
            const int sc = sm.seg[c], nc = sm.seg[c + 1] - sc;
            const bool tri = (b == c);
            if (tri ? nb < 2 : (nb == 0 || nc == 0)) continue;
            if (MODE == 1) dispatch_adj(B, sm, maxn, sb, nb, sc, nc, tri, rowbuf, sP);
            else dispatch_flat<MODE == 0 ? 0 : 1>(B, sm, maxn, sb, nb, sc, nc, tri, tv, rowbuf);
            __syncthreads();
        }
    }
    if (MODE == 1) {
        for (int J = lane; J < n; J += kWave) {
            const int t = sm.canon[J];
            A.P[3 * ((size_t)o + t)] = sP[3 * J];
            A.P[3 * ((size_t)o + t) + 1] = sP[3 * J + 1];
            A.P[3 * ((size_t)o + t) + 2] = sP[3 * J + 2];
        }
    } else {
        float* out = (MODE == 0 ? A.G : A.v) + (size_t)i * ldg;
        for (int m = lane; m < ldg; m += kWave) out[m] = rowbuf[m];
    }
}

// ================================================================ host ====

static SymmArgs make_args(ann_ctx* ctx) {
    SymmArgs A{};
    A.dm = ctx->dm;
    A.rad_rs = ctx->da.rad_rs;
    A.rad_eta = ctx->da.rad_eta;
    A.ang_eta = ctx->da.ang_eta;
    A.ang_zeta = ctx->da.ang_zeta;
    A.ang_lam = ctx->da.ang_lam;
    A.N = ctx->natoms;
    A.ldg = ctx->ldg;
    A.maxn = (std::max(ctx->nl.maxn, 1) + 3) / 4 * 4;   // keeps every LDS array 16-byte aligned
    A.species = ctx->d_species;
    A.off = ctx->nl.off;
    A.geo = ctx->nl.geo;
    A.nsp = ctx->nl.nsp;
    A.nidx = ctx->nl.idx;
    return A;
}

static size_t plain_smem(const SymmArgs& A, int mode) {
    size_t b = (size_t)A.maxn * (16 + 4 + 4 + 4);
    if (mode == 1) b += 4 * ((size_t)A.ldg + 3 * A.maxn);
    else b += 4 * (size_t)A.ldg * 65;
    if (mode == 2) b += 16 * (size_t)A.maxn;
    return b;
}

static size_t fast_smem(const SymmArgs& A, int mode) {
    size_t b = 16 * (size_t)A.maxn + 4 * (size_t)kMaxUnionEta * A.maxn * (mode == 0 ? 1 : 2) +
               4 * 3 * (size_t)A.maxn + 4 * 16;
    b += 4 * ((size_t)A.ldg + 4);
    if (mode == 1) b += 12 * (size_t)A.maxn;
    if (mode == 2) b += 16 * (size_t)A.maxn;
    return b;
}

template <int MODE>
static hipError_t run(ann_ctx* ctx, int form, SymmArgs A) {
    if (A.N == 0) return hipSuccess;
    const bool fast = form == ANN_FORM_FAST && ctx->hm.all_grid;
    const char* names[2][3] = {{"symm_fwd_plain", "symm_bwd_plain", "symm_tan_plain"},
                               {"symm_fwd_fast", "symm_bwd_fast", "symm_tan_fast"}};
    Launch L(ctx, names[fast ? 1 : 0][MODE]);
    if (fast) {
        const size_t sm = fast_smem(A, MODE);
        k_fast<MODE><<<A.N, 64, sm, ctx->stream>>>(A);
    } else {
        const size_t sm = plain_smem(A, MODE);
        k_plain<MODE><<<A.N, 64, sm, ctx->stream>>>(A);
    }
    return hipGetLastError();
}

hipError_t symm_fwd(ann_ctx* ctx, int form, float* G) {
    SymmArgs A = make_args(ctx);
    A.G = G;
    return run<0>(ctx, form, A);
}

hipError_t symm_bwd(ann_ctx* ctx, int form, const float* dEdG, float* P) {
    SymmArgs A = make_args(ctx);
    A.dEdG = dEdG;
    A.P = P;
    return run<1>(ctx, form, A);
}

hipError_t symm_tangent(ann_ctx* ctx, int form, const float* u, float* v) {
    SymmArgs A = make_args(ctx);
    A.u = u;
    A.v = v;
    return run<2>(ctx, form, A);
}

bool fast_shape_supported(int ng, int nk) { return fast_shape_ok(ng, nk); }

}  // namespace ann
