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
#include <algorithm>
#include <vector>

#include "basis_dev.h"
#include "net_dev.h"

namespace ann {

struct SymmArgs {
    const DevModel* dm;
    const float* rad_rs;
    const float* rad_eta;
    const float* ang_eta;
    const float* ang_zeta;
    const float* ang_lam;
    int N, ldg, maxn;
    int smem_wave;            // bytes of dynamic LDS per wavefront (moment kernels: one atom per wavefront)
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
    float* Tout;              // moment form, G: store each atom's moments T here (or null)
    const float* Tin;         // moment form, adjoint: moments of the same geometry (or null: recompute)
    // fused E+F (mode 3): the network's parameter tables and inputs, the atom energies
    const float* wchunk;      // forward weight rows, biases, output neuron (net_dev.h layout)
    const float* wtrans;      // transposed weight rows (backward)
    const float* inpw;
    const float* inpb;
    float* Eatom;
    // persistent dispatch of the moment kernels (or null): per XCD a queue of
    // (first atom, atom count) items, its counter at queue[32 q]
    int* queue;
    const int2* items;
    const int* qoff;          // [kXcds + 1] item ranges of the queues
};

// Sum N per-lane values over the 64 lanes with a transposing butterfly: at
// each halving step a lane keeps half of its values and receives the partner's
// copy of that half (N - 1 exchanges instead of 6N), then the lanes that hold
// the same value index finish with plain butterflies.  Afterwards lane l holds
// the total of value index vidx(l); returns it.
template <int N>
__device__ __forceinline__ float wave_sum_multi(float (&v)[N], int& idx) {
    const int lane = threadIdx.x % kWave;
    int base = 0;
#pragma unroll
    for (int h = N / 2, bit = 32; h >= 1; h >>= 1, bit >>= 1) {
        const bool upper = (lane & bit) != 0;
        // bitwise blend of two statically indexed registers: a conditional
        // choice between v[q] and v[q + h] is otherwise rewritten by the
        // compiler into a runtime-indexed register access (a compare/select
        // chain over all N registers)
        const unsigned m = upper ? 0xffffffffu : 0u;
#pragma unroll
        for (int q = 0; q < h; ++q) {
            const unsigned lo = __float_as_uint(v[q]), hi = __float_as_uint(v[q + h]);
            const float send = __uint_as_float((lo & m) | (hi & ~m));
            const float keep = __uint_as_float((hi & m) | (lo & ~m));
            v[q] = keep + __shfl_xor(send, bit);
        }
        if (upper) base += h;
    }
    float s = v[0];
    constexpr int steps = (N >= 64) ? 0 : 6 - __builtin_ctz(N);
    int bit = 1;
#pragma unroll
    for (int q = 0; q < steps; ++q, bit <<= 1) s += __shfl_xor(s, bit);
    idx = base;
    return s;
}

// LDS ordering between the lanes of one wavefront (each wavefront of a
// moment-kernel block owns its own LDS slice, so no block barrier is needed)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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

constexpr int kFastU = 4;   // distinct angular widths per centre species in the fast form

struct FastSmem {
    float4* g;      // [maxn] (hx, hy, hz, 1/r), species-sorted
    float4* ef;     // [maxn] exp(-r^2 ieta2_u) fc for u = 0..3
    float4* efd;    // [maxn] d/dr of ef
    float2* fg;     // [maxn] cutoff and its derivative
    float* r;       // [maxn]
    int* canon;     // [maxn] canonical slot of each sorted entry
    int* seg;       // [kMaxSpecies + 1]
    float* extra;   // kernel specific
};

__device__ __forceinline__ float f4get(const float4& v, int u) {
    return u == 0 ? v.x : (u == 1 ? v.y : (u == 2 ? v.z : v.w));
}

// w^e for a wave-uniform integer e >= 0 (common exponents unrolled)
__device__ __forceinline__ float ipow(float w, int e) {
    switch (e) {
        case 0: return 1.0f;
        case 1: return w;
        case 2: return w * w;
        case 3: return w * w * w;
        case 4: { const float q = w * w; return q * q; }
        case 7: { const float q = w * w, q4 = q * q; return q4 * q * w; }
        case 8: { const float q = w * w, q4 = q * q; return q4 * q4; }
        default: {
            float r = 1.0f, b = w;
            for (int x = e; x; x >>= 1) {
                if (x & 1) r *= b;
                b *= b;
            }
            return r;
        }
    }
}

// w^(zeta-1) for integer (zi >= 1) or general zeta (zi < 0)
__device__ __forceinline__ float pow_m1(float w, int zi, float zeta) {
    if (zi >= 1) return ipow(w, zi - 1);
    return w > 0.0f ? exp2f((zeta - 1.0f) * __log2f(w)) : 0.0f;
}

// Regroup the neighbours of atom i by species (stable) and precompute the
// per-neighbour factors; seg[] gives the species segments.  Entries beyond the
// cutoff (a list built with rc + skin) are dropped here; their adjoint slots
// are written as zero when zero_P is set.  Returns the number kept.
__device__ __forceinline__ int fast_load(const SymmArgs& A, const DevModel& M, int i, int a, int o, int n, FastSmem& sm,
                         bool need_deriv, bool zero_P, float4* tv = nullptr) {
    const int lane = threadIdx.x % kWave, S = M.S, U = M.nunion[a];
    const float rc = M.rc;
    int cnt[kMaxSpecies];
#pragma unroll
    for (int b = 0; b < kMaxSpecies; ++b) cnt[b] = 0;
    // the first two rows of 64 neighbours stay in registers for the second pass
    float4 g0 = make_float4(0.f, 0.f, 0.f, 0.f), g1 = g0;
    int p0 = -1, p1 = -1, j0 = 0, j1 = 0;
    for (int q = 0; q < n; q += kWave) {
        const int t = q + lane;
        float4 gq = make_float4(0.f, 0.f, 0.f, 0.f);
        int raw = -1, jr = 0;
        if (t < n) {
            gq = A.geo[o + t];
            raw = (int)A.nsp[o + t];
            if (tv) jr = A.nidx[o + t];
        }
        if (q == 0) { g0 = gq; p0 = raw; j0 = jr; }
        if (q == kWave) { g1 = gq; p1 = raw; j1 = jr; }
        const int sp = (t < n && gq.w <= rc) ? raw : -1;
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
    float ie[kFastU];
#pragma unroll
    for (int u = 0; u < kFastU; ++u) ie[u] = u < U ? M.uieta2[a][u] : 0.0f;
    const uint64_t lt = (1ull << lane) - 1ull;
    // tangent: the centre's displacement (the same for every lane)
    float uix = 0.f, uiy = 0.f, uiz = 0.f;
    if (tv) {
        uix = A.u[3 * i];
        uiy = A.u[3 * i + 1];
        uiz = A.u[3 * i + 2];
    }
    for (int q = 0; q < n; q += kWave) {
        const int t = q + lane;
        float4 gq;
        int raw, jr;
        if (q == 0) { gq = g0; raw = p0; jr = j0; }
        else if (q == kWave) { gq = g1; raw = p1; jr = j1; }
        else {
            gq = t < n ? A.geo[o + t] : make_float4(0.f, 0.f, 0.f, 0.f);
            raw = t < n ? (int)A.nsp[o + t] : -1;
            jr = (tv && t < n) ? A.nidx[o + t] : 0;
        }
        const bool keep = t < n && gq.w <= rc;
        const int sp = keep ? raw : -1;
        if (zero_P && t < n && !keep) {
            A.P[3 * ((size_t)o + t)] = 0.0f;
            A.P[3 * ((size_t)o + t) + 1] = 0.0f;
            A.P[3 * ((size_t)o + t) + 2] = 0.0f;
        }
        int pos = 0;
#pragma unroll
        for (int b = 0; b < kMaxSpecies; ++b) {
            if (b < S) {
                const uint64_t mk = __ballot(sp == b);
                if (sp == b) pos = base[b] + __popcll(mk & lt);
                base[b] += __popcll(mk);
            }
        }
        if (keep) {
            const float r = gq.w, ir = __builtin_amdgcn_rcpf(r);   // 1 ulp
            sm.g[pos] = make_float4(gq.x * ir, gq.y * ir, gq.z * ir, ir);
            if (tv) {
                // tangent direction t = u_i - u_j and its projection on the unit vector
                const float tx = uix - A.u[3 * jr], ty = uiy - A.u[3 * jr + 1], tz = uiz - A.u[3 * jr + 2];
                tv[pos] = make_float4(tx, ty, tz, (gq.x * ir) * tx + (gq.y * ir) * ty + (gq.z * ir) * tz);
            }
            sm.r[pos] = r;
            if (sm.canon) sm.canon[pos] = t;                     // slot map (kernels that need it)
            float f, gcut;
            cutoff_fg(M.cutoff, M.rc, M.rci, M.pirc, r, f, gcut);
            sm.fg[pos] = make_float2(f, gcut);
            float e[kFastU], d[kFastU];
#pragma unroll
            for (int u = 0; u < kFastU; ++u) {
                const float x = u < U ? __expf(-ie[u] * r * r) : 0.0f;
                e[u] = x * f;
                d[u] = x * (-2.0f * r * ie[u] * f + gcut);
            }
            sm.ef[pos] = make_float4(e[0], e[1], e[2], e[3]);
            if (need_deriv) sm.efd[pos] = make_float4(d[0], d[1], d[2], d[3]);
        }
    }
    return acc;
}

// Distance factor of width g for a pair of legs j, k: GAUSS blocks multiply the
// per-neighbour factors e_g(r) = exp(-r^2/eta^2) fc(r) (and their r-derivatives)
// computed once per neighbour; the other families depend on r_j^2 + r_k^2
// through a factor that does not separate, evaluated here per pair:
// D = fr(r_j^2 + r_k^2) fc_j fc_k, dD/dr_j = (2 r_j fr' fc_j + fr fc'_j) fc_k.
template <bool GS, int NG>
__device__ __forceinline__ void pair_factors(const AngBlock& B, const float* uie, const int (&gu)[NG], const FastSmem& sm,
                                             int jj, int kk, bool deriv, float (&D)[NG], float (&Dj)[NG],
                                             float (&Dk)[NG]) {
    if (GS) {
        const float4 ej4 = sm.ef[jj], ek4 = sm.ef[kk];
        float4 dj4 = make_float4(0.f, 0.f, 0.f, 0.f), dk4 = dj4;
        if (deriv) { dj4 = sm.efd[jj]; dk4 = sm.efd[kk]; }
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const float ej = f4get(ej4, g), ek = f4get(ek4, g);
            D[g] = ej * ek;
            Dj[g] = deriv ? f4get(dj4, g) * ek : 0.0f;
            Dk[g] = deriv ? ej * f4get(dk4, g) : 0.0f;
        }
    } else {
        const float rj = sm.r[jj], rk = sm.r[kk];
        const float2 fj = sm.fg[jj], fk = sm.fg[kk];
        const float r2s = rj * rj + rk * rk, ff = fj.x * fk.x;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            float fr, dfr;
            ang_radial(B.name, uie[gu[g]], r2s, fr, dfr);
            D[g] = fr * ff;
            Dj[g] = deriv ? (2.0f * rj * dfr * fj.x + fr * fj.y) * fk.x : 0.0f;
            Dk[g] = deriv ? (2.0f * rk * dfr * fk.x + fr * fk.y) * fj.x : 0.0f;
        }
    }
}

// ---- forward / tangent over one grid block (unordered pairs, flattened) ----
template <int NG, int NK, int MODE, bool GS>
__device__ __forceinline__ void fast_block_flat(const AngBlock& B, const float* uie, const FastSmem& sm, int sb, int nb,
                                                int sc, int nc, bool tri, const float4* tv, float* out) {
    const int lane = threadIdx.x % kWave;
    float lamh[NK], zeta[NK];
    int zi[NK], gu[NG];
#pragma unroll
    for (int k = 0; k < NK; ++k) { lamh[k] = B.lamh[k]; zeta[k] = B.zeta[k]; zi[k] = B.zint[k]; }
#pragma unroll
    for (int g = 0; g < NG; ++g) gu[g] = B.gu[g];
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
        float Dg[NG], Dj[NG], Dk[NG];
        pair_factors<GS, NG>(B, uie, gu, sm, jj, kk, MODE != 0, Dg, Dj, Dk);
        if (MODE == 0) {
            float p[NK];
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                const float cw = fabsf(0.5f + lamh[k] * c);
                p[k] = zi[k] >= 0 ? ipow(cw, zi[k]) : pow_m1(cw, -1, zeta[k]) * cw;
            }
#pragma unroll
            for (int g = 0; g < NG; ++g) {
#pragma unroll
                for (int k = 0; k < NK; ++k) acc[g][k] += p[k] * Dg[g];
            }
        } else {
            const float4 tj = tv[jj], tk = tv[kk];
            const float cdot = ((gk.x * tj.x + gk.y * tj.y + gk.z * tj.z) - c * tj.w) * gj.w +
                               ((gj.x * tk.x + gj.y * tk.y + gj.z * tk.z) - c * tk.w) * gk.w;
            float p[NK], dp[NK];
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                const float cw = fabsf(0.5f + lamh[k] * c);
                const float w1 = pow_m1(cw, zi[k], zeta[k]);
                p[k] = w1 * cw;
                dp[k] = zeta[k] * lamh[k] * w1;
            }
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const float rdot = Dj[g] * tj.w + Dk[g] * tk.w;
                const float D = Dg[g] * cdot;
#pragma unroll
                for (int k = 0; k < NK; ++k) acc[g][k] += dp[k] * D + p[k] * rdot;
            }
        }
        t += dt;
        j += dj;
        if (t >= cols) { t -= cols; j += 1; }
    }
    float flat[NG * NK];
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int k = 0; k < NK; ++k) flat[g * NK + k] = acc[g][k];
    if constexpr ((NG * NK & (NG * NK - 1)) == 0) {
        int idx;
        const float s = wave_sum_multi<NG * NK>(flat, idx);
        constexpr int lpv = kWave / (NG * NK);   // lanes holding each value
        if ((lane & (lpv - 1)) == 0 && lane < kWave) out[B.goff + B.outm[idx]] = s;
    } else {
#pragma unroll
        for (int q = 0; q < NG * NK; ++q) {
            const float s = wave_sum(flat[q]);
            if (lane == 0) out[B.goff + B.outm[q]] = s;
        }
    }
}

// ---- adjoint over one grid block: skewed owner/partner schedule ----
template <int NG, int NK, bool GS>
__device__ __forceinline__ void fast_block_adj(const AngBlock& B, const float* uie, const FastSmem& sm, int sb, int nb,
                                               int sc, int nc, bool tri, const float* sd, float* sP) {
    const int lane = threadIdx.x % kWave;
    float lamh[NK], zeta[NK], dl[NG][NK];
    int zi[NK], gu[NG];
#pragma unroll
    for (int k = 0; k < NK; ++k) { lamh[k] = B.lamh[k]; zeta[k] = B.zeta[k]; zi[k] = B.zint[k]; }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        gu[g] = B.gu[g];
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
        float ej[NG], edj[NG];                  // GAUSS: the owner's per-neighbour factors, loaded once
        if constexpr (GS) {
            const float4 ej4 = sm.ef[J], dj4 = sm.efd[J];
#pragma unroll
            for (int g = 0; g < NG; ++g) { ej[g] = f4get(ej4, g); edj[g] = f4get(dj4, g); }
        }
        float px = 0.0f, py = 0.0f, pz = 0.0f;
        int len, koff;
        if (tri) {
            len = h + ((no & 1) == 0 && j < (no >> 1) ? 1 : 0);
            koff = j + 1;
        } else {
            len = (q < Q - 1) ? no : np - (Q - 1) * no;
            koff = j + q * no;
        }
        const int lmax = tri ? h + ((no & 1) == 0 ? 1 : 0) : (np - (Q - 1) * no);
        for (int t = 0; t < lmax; ++t) {
            if (act && t < len) {
                int k = koff + t;
                if (k >= np) k -= np;
                const int K = spp + k;
                const float4 gk = sm.g[K];
                const float c = gj.x * gk.x + gj.y * gk.y + gj.z * gk.z;
                float Dg[NG], Dj[NG], Dk[NG];
                if constexpr (GS) {
                    const float4 ek4 = sm.ef[K], dk4 = sm.efd[K];
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        const float ek = f4get(ek4, g);
                        Dg[g] = ej[g] * ek;
                        Dj[g] = edj[g] * ek;
                        Dk[g] = ej[g] * f4get(dk4, g);
                    }
                } else {
                    pair_factors<GS, NG>(B, uie, gu, sm, J, K, true, Dg, Dj, Dk);
                }
                float p[NK], dp[NK];
#pragma unroll
                for (int kk = 0; kk < NK; ++kk) {
                    const float cw = fabsf(0.5f + lamh[kk] * c);
                    const float w1 = pow_m1(cw, zi[kk], zeta[kk]);
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
                    Ac += Tg * Dg[g];
                    Bj += Sg * Dj[g];
                    Bk += Sg * Dk[g];
                }
                const float aj = Ac * gj.w, ak = Ac * gk.w;
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

#define ANN_SHAPES(X) X(1, 8) X(4, 4) X(3, 4) X(2, 4) X(1, 4) X(1, 2) X(1, 1) X(2, 2) X(4, 2) X(3, 2)

// shapes with a compiled fast kernel
__host__ __device__ inline bool fast_shape_ok(int ng, int nk) {
#define X(G, K) if (ng == G && nk == K) return true;
    ANN_SHAPES(X)
#undef X
    return false;
}

// radial part, lane over the neighbours of one species segment, 16 functions at a time
// GAUSSIAN radial functions of one (centre, neighbour species) block, CH per
// pass with widths and centres in registers.  Adjoint: per neighbour,
// d/dr [e fc] = e (gc - 2 eta x fc) weighted by dE/dG, split into two running
// sums; G / tangent: lanes over the neighbours, one transposing reduction.
template <int MODE, int CH>
__device__ __forceinline__ void radial_gauss(const SymmArgs& A, const RadBlock& rb, const FastSmem& sm, int s0, int s1,
                                             const float4* tv, const float* sd, float* sP, float* out) {
    const int lane = threadIdx.x % kWave;
    for (int m0 = 0; m0 < rb.n; m0 += CH) {
        float eta[CH], rs[CH], gs[CH];
#pragma unroll
        for (int m = 0; m < CH; ++m) {
            const bool in = m0 + m < rb.n;
            eta[m] = in ? A.rad_eta[rb.par + m0 + m] : 0.0f;
            rs[m] = in ? A.rad_rs[rb.par + m0 + m] : 0.0f;
            gs[m] = (MODE == 1 && in) ? sd[rb.goff + m0 + m] : 0.0f;
        }
        if (MODE == 1) {
            for (int J = s0 + lane; J < s1; J += kWave) {
                const float r = sm.r[J];
                const float2 fg = sm.fg[J];
                float se = 0.0f, sx = 0.0f;
#pragma unroll
                for (int m = 0; m < CH; ++m) {
                    const float x = eta[m] * (r - rs[m]), ge = gs[m] * __expf(-x * x);
                    se += ge;
                    sx += ge * eta[m] * x;
                }
                const float amp = se * fg.y - 2.0f * fg.x * sx;
                const float4 gj = sm.g[J];
                sP[3 * J] += amp * gj.x;
                sP[3 * J + 1] += amp * gj.y;
                sP[3 * J + 2] += amp * gj.z;
            }
            continue;
        }
        float acc[CH];
#pragma unroll
        for (int m = 0; m < CH; ++m) acc[m] = 0.0f;
        for (int J = s0 + lane; J < s1; J += kWave) {
            const float r = sm.r[J];
            const float w = MODE == 2 ? tv[J].w : 0.0f;
            const float2 fg = sm.fg[J];
            const float fc = fg.x, gc = fg.y;
#pragma unroll
            for (int m = 0; m < CH; ++m) {
                const float x = eta[m] * (r - rs[m]), e = __expf(-x * x);
                acc[m] += MODE == 0 ? e * fc : e * (gc - 2.0f * eta[m] * x * fc) * w;
            }
        }
        int idx;
        const float sum = wave_sum_multi<CH>(acc, idx);
        if ((lane & (kWave / CH - 1)) == 0 && m0 + idx < rb.n) out[rb.goff + m0 + idx] = sum;
    }
}

// GAUSSIAN block of at most 16 functions (G and tangent): lanes over
// (neighbour, function) pairs, W = 8 or 16 functions per lane group and 64 / W
// neighbours per pass, so every lane evaluates one exponential per pass and the
// sum over neighbours is log2(64 / W) butterflies.  Against W exponentials per
// lane on a partly idle wavefront and a W-value transposing reduction it wins
// up to about 32 passes (water: 60 neighbours of one species, 15 passes).
template <int MODE>
__device__ __forceinline__ void radial_gauss_pairs(const SymmArgs& A, const RadBlock& rb, const FastSmem& sm, int s0, int s1,
                                                   int W, const float4* tv, float* out) {
    const int lane = threadIdx.x % kWave;
    const int m = lane & (W - 1), nj = kWave / W;
    const bool in = m < rb.n;
    const float eta = in ? A.rad_eta[rb.par + m] : 0.0f;
    const float rs = in ? A.rad_rs[rb.par + m] : 0.0f;
    float acc = 0.0f;
    for (int J = s0 + lane / W; J < s1; J += nj) {
        const float r = sm.r[J];
        const float2 fg = sm.fg[J];
        const float x = eta * (r - rs), e = __expf(-x * x);
        acc += MODE == 0 ? e * fg.x : e * (fg.y - 2.0f * eta * x * fg.x) * tv[J].w;
    }
    for (int bit = W; bit < kWave; bit <<= 1) acc += __shfl_xor(acc, bit);
    if (lane < W && in) out[rb.goff + m] = acc;
}

// passes of the pair form above which the lane-over-neighbour form is used
constexpr int kPairPasses = 32;

// the moment kernels index block widths by union width: both bounds must agree
static_assert(kMaxUnionEta == kGridG, "union widths and block widths share one index");

template <int MODE>
__device__ __forceinline__ void fast_radial(const SymmArgs& A, const DevModel& M, int a, const FastSmem& sm, const float4* tv,
                            const float* sd, float* sP, float* out) {
    const int lane = threadIdx.x % kWave, S = M.S;
    for (int b = 0; b < S; ++b) {
        const RadBlock rb = M.rad[a * S + b];
        if (rb.n == 0) continue;
        const int s0 = sm.seg[b], s1 = sm.seg[b + 1];
        if (rb.name == RAD_GAUSSIAN && MODE != 1 && rb.n <= 16) {
            const int W = rb.n <= 8 ? 8 : 16;
            if ((s1 - s0) * W <= kPairPasses * kWave) {
                radial_gauss_pairs<MODE>(A, rb, sm, s0, s1, W, tv, out);
                continue;
            }
        }
        if (rb.name == RAD_GAUSSIAN) {
            // eight per pass for blocks of at most eight functions, else sixteen
            // (the tangent keeps sixteen: a second copy raises its register count)
            if (MODE != 2 && rb.n <= 8) radial_gauss<MODE, 8>(A, rb, sm, s0, s1, tv, sd, sP, out);
            else radial_gauss<MODE, 16>(A, rb, sm, s0, s1, tv, sd, sP, out);
            continue;
        }
        if (MODE == 1) {
            for (int J = s0 + lane; J < s1; J += kWave) {
                const float4 gj = sm.g[J];
                const float2 fg = sm.fg[J];
                const float fc = fg.x, gc = fg.y;
                float amp = 0.0f;
                for (int m = 0; m < rb.n; ++m) {
                    float v, d;
                    radial_fd(rb.name, A.rad_eta[rb.par + m], A.rad_rs[rb.par + m], sm.r[J], fc, gc, v, d);
                    amp += sd[rb.goff + m] * d;
                }
                sP[3 * J] += amp * gj.x;
                sP[3 * J + 1] += amp * gj.y;
                sP[3 * J + 2] += amp * gj.z;
            }
            continue;
        }
        // other families: one function at a time, plain wave sums
        for (int m = 0; m < rb.n; ++m) {
            float acc = 0.0f;
            for (int J = s0 + lane; J < s1; J += kWave) {
                const float r = sm.r[J];
                const float2 fg = sm.fg[J];
                const float fc = fg.x, gc = fg.y;
                float v, d;
                radial_fd(rb.name, A.rad_eta[rb.par + m], A.rad_rs[rb.par + m], r, fc, gc, v, d);
                acc += MODE == 0 ? v : d * tv[J].w;
            }
            acc = wave_sum(acc);
            if (lane == 0) out[rb.goff + m] = acc;
        }
    }
}

// mode 0: G, 1: adjoint P, 2: tangent v; every angular block has the grid shape NG x NK
template <int MODE, int NG, int NK>
__global__ __launch_bounds__(64) void k_fast(SymmArgs A) {
    extern __shared__ float4 smem4[];
    const int i = blockIdx.x, lane = threadIdx.x;
    if (i >= A.N) return;
    const DevModel& M = *A.dm;
    const int a = A.species[i], o = A.off[i];
    int n = A.off[i + 1] - o;
    const int S = M.S, NP = M.NP, ldg = A.ldg, maxn = A.maxn;
    FastSmem sm;
    sm.g = smem4;
    sm.ef = sm.g + maxn;
    sm.efd = sm.ef + maxn;
    float4* tv = sm.efd + (MODE == 0 ? 0 : maxn);                // [maxn] (mode 2)
    sm.fg = reinterpret_cast<float2*>(tv + (MODE == 2 ? maxn : 0));
    sm.r = reinterpret_cast<float*>(sm.fg + maxn);
    sm.canon = reinterpret_cast<int*>(sm.r + maxn);
    sm.seg = sm.canon + maxn;
    sm.extra = reinterpret_cast<float*>(sm.seg + 16);
    float* rowbuf = sm.extra;                                   // [ldg] output row (modes 0, 2) / dE/dG (mode 1)
    float* sP = sm.extra + ldg;                                 // [maxn][3] (mode 1)
    n = fast_load(A, M, i, a, o, n, sm, MODE != 0, MODE == 1);   // n: neighbours within rc
    if (MODE == 1) {
        for (int m = lane; m < ldg; m += kWave) rowbuf[m] = A.dEdG[(size_t)i * ldg + m];
        for (int t = lane; t < 3 * n; t += kWave) sP[t] = 0.0f;
    } else {
        for (int m = lane; m < ldg; m += kWave) rowbuf[m] = 0.0f;
    }
    __syncthreads();
    if (MODE == 2) {
        // per neighbour tangent t_j = u_i - u_j and r^_j . t_j, species-sorted
        for (int J = lane; J < n; J += kWave) {
            const int j = A.nidx[o + sm.canon[J]];
            const float tx = A.u[3 * i] - A.u[3 * j], ty = A.u[3 * i + 1] - A.u[3 * j + 1],
                        tz = A.u[3 * i + 2] - A.u[3 * j + 2];
            const float4 gq = sm.g[J];
            tv[J] = make_float4(tx, ty, tz, gq.x * tx + gq.y * ty + gq.z * tz);
        }
        __syncthreads();
    }
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
            const float* uie = M.uieta2[a];
            if (B.name == ANG_GAUSS) {
                if (MODE == 1) fast_block_adj<NG, NK, true>(B, uie, sm, sb, nb, sc, nc, tri, rowbuf, sP);
                else fast_block_flat<NG, NK, MODE == 0 ? 0 : 1, true>(B, uie, sm, sb, nb, sc, nc, tri, tv, rowbuf);
            } else {
                if (MODE == 1) fast_block_adj<NG, NK, false>(B, uie, sm, sb, nb, sc, nc, tri, rowbuf, sP);
                else fast_block_flat<NG, NK, MODE == 0 ? 0 : 1, false>(B, uie, sm, sb, nb, sc, nc, tri, tv, rowbuf);
            }
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

// =========================================================== MOMENT form ====
//
// For GAUSS angular blocks with integer zeta <= 5 the triplet sum factorises
// exactly.  With e_u(r) = exp(-r^2/eta_u^2) fc(r) and c = r^_j . r^_k,
//   (1/2 + lambda c / 2)^zeta = sum_m C(zeta,m) 2^(m-zeta) (lambda/2)^m c^m,
//   c^m = sum_{|alpha|=m} (m!/alpha!) r^_j^alpha r^_k^alpha,
// so a block (species b, species c, width u) needs only the per-species
// moments T_b[u][alpha] = sum_{j in b} e_u(r_j) r^_j^alpha and
// D_b[u] = sum_{j in b} e_u(r_j)^2:
//   A_m = sum_{|alpha|=m} w_alpha T_b T_c                (b != c)
//   A_m = (sum_{|alpha|=m} w_alpha T_b T_b - D_b) / 2    (b == c: pairs j < k)
//   G(u, lambda, zeta) = sum_m C(zeta,m) 2^(m-zeta) (lambda/2)^m A_m.
// Work per atom is O(n M) with M = (Z+1)(Z+2)(Z+3)/6 monomials instead of
// O(n^2) triplets.  The adjoint differentiates the same chain: dE/dT_b from
// the block contractions, then per neighbour the polynomial
// Q(r^) = sum_alpha dE/dT_b[u][alpha] r^^alpha and its gradient.  The
// tangent (force-loss second backward) is the forward-mode derivative.
// FP32 throughout: the alternating binomial sum amplifies rounding by at most
// C(5,2) = 10 for zeta <= 5; larger zeta keeps the triplet kernels.

__host__ __device__ constexpr int nmono(int Z) { return (Z + 1) * (Z + 2) * (Z + 3) / 6; }
// neighbours per chunk of power rows (G: 32 and adjoint: 64, which frees LDS
// for more resident wavefronts; tangent: every neighbour at once, measured
// faster with its two-level rows), and the floats of the region that holds them (it
// also holds the cutoff pairs, two floats per neighbour, before the powers are
// needed; with stored moments only those)
constexpr int kPwRows = 64;
// adjoint: dE/dT entries per lane held in registers while they replace the moments
constexpr int kUbRegs = 8;
__host__ __device__ inline int pw_chunk(int mode, int maxn) {
    return mode == 2 ? maxn : (mode == 0 || mode == 3 ? kPwRows / 2 : kPwRows);
}
__host__ __device__ inline int pw_floats(int mode, int maxn, int pws, bool haveT) {
    const int rows = pw_chunk(mode, maxn);
    if (mode == 3) {                                // fused: also the adjoint vectors and network vectors
        const int m = rows * pws > 3 * maxn ? rows * pws : 3 * maxn;
        return m > 3 * kVecMaxWidth ? m : 3 * kVecMaxWidth;
    }
    return haveT ? 2 * maxn : (rows * pws > 2 * maxn ? rows * pws : 2 * maxn);
}
__host__ __device__ constexpr int mono_start(int m) { return m * (m + 1) * (m + 2) / 6; }
__host__ __device__ constexpr int ifact(int n) { return n <= 1 ? 1 : n * ifact(n - 1); }
// position of x^a y^b z^c in the monomial order below
__host__ __device__ constexpr int mono_index(int a, int b, int c) {
    return mono_start(a + b + c) + (b + c) * (b + c + 1) / 2 + c;
}

// monomials of total degree <= kMomZ ordered by degree, then exponent of x and
// y descending; every smaller degree bound is a prefix
struct MonoTable {
    int a[nmono(kMomZ)], b[nmono(kMomZ)], c[nmono(kMomZ)], m[nmono(kMomZ)];
    float w[nmono(kMomZ)];
    constexpr MonoTable() : a(), b(), c(), m(), w() {
        int q = 0;
        for (int mm = 0; mm <= kMomZ; ++mm)
            for (int aa = mm; aa >= 0; --aa)
                for (int bb = mm - aa; bb >= 0; --bb) {
                    const int cc = mm - aa - bb;
                    a[q] = aa; b[q] = bb; c[q] = cc; m[q] = mm;
                    w[q] = (float)(ifact(mm) / (ifact(aa) * ifact(bb) * ifact(cc)));
                    ++q;
                }
    }
};
__constant__ MonoTable c_mono = MonoTable();

__device__ __forceinline__ int pair_of(int b, int c) { return b <= c ? c * (c + 1) / 2 + b : b * (b + 1) / 2 + c; }

// groups of lanes sharing a reduction of n outputs (n <= 16: four groups of
// 16 lanes, <= 32: two of 32, else one of 64)
__device__ __forceinline__ int lane_groups(int n) { return n <= 16 ? 4 : (n <= 32 ? 2 : 1); }

// The atom's network on its own wavefront (fused E+F; ANN::fpbp + DODZ::grad,
// src/ml/nn.cpp:1367-1389, 1801-1830): rowbuf holds the G row on entry and
// dE/dG on exit (zero past the species' inputs); the atom energy goes to
// A.Eatom.  Lanes over a layer's outputs, its reduction split over up to four
// lane groups joined by butterflies; the layer's inputs are broadcast LDS
// reads, the weights coalesced reads of the row tables (forward: rows of W,
// backward: rows of W^T, net_dev.h layout).  nb: [3][kVecMaxWidth] LDS floats.
__device__ __forceinline__ void fused_net(const SymmArgs& A, const NetDesc& D, int i, float* rowbuf, float* nb) {
    const int lane = threadIdx.x % kWave;
    const int L = D.nlayer, nin = D.nin, ldg = A.ldg;
    const float* iw = A.inpw + D.inoff;
    const float* ib = A.inpb + D.inoff;
    for (int k = lane; k < nin; k += kWave) rowbuf[k] = iw[k] * (rowbuf[k] + ib[k]);
    wave_sync();
    const float* x = rowbuf;
    int ni = nin;
    float sp[kMaxLayers - 1];                               // sigma' of the lane's neuron per hidden layer
#pragma unroll
    for (int l = 0; l < kMaxLayers - 1; ++l) sp[l] = 0.0f;
#pragma unroll
    for (int l = 0; l < kMaxLayers - 1; ++l) {
        if (l >= L - 1) break;
        const int no = D.nodes[l], ps = D.wcs[l];
        const float* W = A.wchunk + D.wcf[l];
        const float* b = A.wchunk + D.wcb[l];
        const int G = lane_groups(no), gw = kWave / G;
        const int j = lane % gw, grp = lane / gw, jj = min(j, no - 1);
        const int kper = (ni + G - 1) / G, k0 = grp * kper, k1 = min(ni, k0 + kper);
        float acc = 0.0f;
#pragma unroll 4
        for (int k = k0; k < k1; ++k) acc += W[k * ps + jj] * x[k];
        for (int m = gw; m < kWave; m <<= 1) acc += __shfl_xor(acc, m);
        float z[1] = {acc + b[jj]}, av[1], dv[1], d2[1];
        act_many<false, 1>(D.neuron, D.c, z, av, dv, d2);
        sp[l] = dv[0];
        float* h = nb + (l & 1) * kVecMaxWidth;
        if (grp == 0 && j < no) h[j] = av[0];
        wave_sync();
        x = h;
        ni = no;
    }
    // linear output neuron
    const float* WL = A.wchunk + D.wco;                     // weights, then (at a multiple of 4) the bias
    float part = 0.0f;
    for (int m = lane; m < ni; m += kWave) part += WL[m] * x[m];
    part = wave_sum(part);
    if (lane == 0) A.Eatom[i] = D.outb + D.outw * (WL[vec_pad4(ni)] + part);
    wave_sync();                                            // every lane is done with x
    if (L == 1) {
        for (int k = lane; k < ldg; k += kWave) rowbuf[k] = k < nin ? D.outw * WL[k] * iw[k] : 0.0f;
        wave_sync();
        return;
    }
    // g of the last hidden layer (its lanes hold sigma' for neuron lane % gw)
    float* dc = nb + 2 * kVecMaxWidth;
    float* dn = nb;
    {
        const int G = lane_groups(ni), gw = kWave / G, j = lane % gw;
        float spl = 0.0f;
#pragma unroll
        for (int q = 0; q < kMaxLayers - 1; ++q)
            if (q == L - 2) spl = sp[q];
        if (lane < gw && j < ni) dc[j] = D.outw * WL[j] * spl;
    }
    wave_sync();
    // hidden layers: g_{l-1}[k] = sigma'_{l-1}[k] sum_j W_l[k][j] g_l[j], lanes over k
#pragma unroll
    for (int l = kMaxLayers - 2; l >= 1; --l) {
        if (l > L - 2) continue;
        const int no = D.nodes[l], nil = D.nodes[l - 1], pst = D.wts[l];
        const float* WT = A.wtrans + D.wtf[l];
        const int G = lane_groups(nil), gw = kWave / G;
        const int k = lane % gw, grp = lane / gw, kk = min(k, nil - 1);
        const int jper = (no + G - 1) / G, j0 = grp * jper, j1 = min(no, j0 + jper);
        float acc = 0.0f;
#pragma unroll 4
        for (int j = j0; j < j1; ++j) acc += WT[j * pst + kk] * dc[j];
        for (int m = gw; m < kWave; m <<= 1) acc += __shfl_xor(acc, m);
        if (grp == 0 && k < nil) dn[k] = acc * sp[l - 1];
        wave_sync();
        float* t = dc;
        dc = dn;
        dn = t;
    }
    // input layer: dE/dG[k] = inpw[k] sum_j W_0[k][j] g_0[j], blocks of up to 64 inputs
    {
        const int no = D.nodes[0], pst = D.wts[0];
        const float* WT = A.wtrans + D.wtf[0];
        for (int kb = 0; kb < nin; kb += kWave) {
            const int nk = min(kWave, nin - kb);
            const int G = lane_groups(nk), gw = kWave / G;
            const int k = kb + lane % gw, grp = lane / gw, kk = min(k, nin - 1);
            const int jper = (no + G - 1) / G, j0 = grp * jper, j1 = min(no, j0 + jper);
            float acc = 0.0f;
#pragma unroll 4
            for (int j = j0; j < j1; ++j) acc += WT[j * pst + kk] * dc[j];
            for (int m = gw; m < kWave; m <<= 1) acc += __shfl_xor(acc, m);
            if (grp == 0 && lane % gw < nk) rowbuf[k] = acc * iw[k];
        }
        for (int k = nin + lane; k < ldg; k += kWave) rowbuf[k] = 0.0f;
        wave_sync();
    }
}

// One wavefront per atom with its own LDS slice (smem4); no block barriers, so
// the waves of a workgroup are independent.
template <int MODE, int Z>
__device__ __forceinline__ void moment_atom(const SymmArgs& A, const int i, float4* smem4) {
    constexpr int NM = nmono(Z);
    constexpr int ZP = Z + 1;
    constexpr int PWS = (3 * ZP * (MODE == 2 ? 2 : 1)) | 1;   // powers per neighbour, odd stride
    constexpr int U4 = kMaxUnionEta;
    constexpr int NG = kWave / NM;                           // lane groups over the neighbours
    constexpr bool FUSED = MODE == 3;                        // G, network and adjoint of the atom
    constexpr bool ADJ = MODE == 1 || FUSED;                 // per-slot adjoint vectors written
    constexpr int GM = FUSED ? 0 : MODE;                     // mode of the basis phase
    const int lane = threadIdx.x % kWave;
    const DevModel& Mo = *A.dm;
    const int a = A.species[i], o = A.off[i];
    int n = A.off[i + 1] - o;
    const int S = Mo.S, NP = Mo.NP, ldg = A.ldg, maxn = A.maxn;
    FastSmem sm;
    sm.g = smem4;
    sm.ef = sm.g + maxn;
    sm.efd = sm.ef + maxn;
    float4* tv = sm.efd + (MODE == 0 ? 0 : maxn);
    sm.r = reinterpret_cast<float*>(tv + (MODE == 2 ? maxn : 0));
    int* const after_r = reinterpret_cast<int*>(sm.r + maxn);
    sm.canon = ADJ ? after_r : nullptr;                         // the adjoint's slot map
    sm.seg = after_r + (ADJ ? maxn : 0);
    sm.extra = reinterpret_cast<float*>(sm.seg + 16);
    float* rowbuf = sm.extra;                                   // [ldg]
    // [maxn][3] adjoint vectors (mode 1; fused: they share the power rows'
    // region, whose rows are dead once the moments are summed)
    float* sP = rowbuf + ldg + (FUSED ? 2 * maxn : 0);
    // adjoint with stored moments: no power rows (the cutoff pairs still use their space)
    const bool haveT = MODE == 1 && A.Tin != nullptr;
    const int TS = S * U4 * NM;                                 // moments per atom
    float* fgr = rowbuf + ldg;                                  // fused: [maxn] cutoff pairs of their own
    float* pw = FUSED ? sP : sP + (MODE == 1 ? 3 * maxn : 0);  // [pw_chunk][PWS] (one chunk of neighbours)
    float* T = pw + pw_floats(MODE, maxn, PWS, haveT);         // [S][U4][NM]
    float* dT = T + S * U4 * NM;                                // [S][U4][NM] (mode 2)
    float* D = dT + (MODE == 2 ? S * U4 * NM : 0);              // [S][U4]
    float* dD = D + S * U4;                                     // [S][U4] (mode 2: dD, mode 1: dE/dD)
    float* Am = dD + S * U4;                                    // [NP][U4][ZP]
    // adjoint: dE/dT overwrites the moments when it fits a few registers per lane
    const bool ub_in_T = S * NM * U4 <= kUbRegs * kWave;
    float* Ub = ub_in_T ? T : Am + NP * U4 * ZP;               // [S][NM][U4] (mode 1)
    float* part = Am + NP * U4 * ZP + (ADJ && !ub_in_T ? S * NM * U4 : 0);   // [NG][NM][U4] x (1 or 2)
    float* wq = part + (NG > 1 ? NG * NM * U4 * (MODE == 2 ? 2 : 1) : 0);   // [NM] monomial weights
    float* nb = FUSED ? pw : wq + NM;                           // fused: [3][kVecMaxWidth] network vectors (power rows' region)
    // cutoffs live in the power rows until the radial part is done (fused: the
    // adjoint's radial part runs after the moments, so they have their own rows)
    sm.fg = reinterpret_cast<float2*>(FUSED ? fgr : pw);

    // adjoint: the dE/dG row and the stored moments are loaded before the neighbours so
    // that their latency overlaps the neighbour load (registers, written to LDS below)
    constexpr int kPreT = 8, kPreD = 2;                         // moments / dE/dG prefetched per lane
    float pre_d[kPreD], pre_t[kPreT];
    if (MODE == 1) {
#pragma unroll
        for (int h = 0; h < kPreD; ++h) {
            const int m = lane + h * kWave;
            pre_d[h] = m < ldg ? A.dEdG[(size_t)i * ldg + m] : 0.0f;
        }
#pragma unroll
        for (int h = 0; h < kPreT; ++h) {
            const int q = lane + h * kWave;
            pre_t[h] = (haveT && q < TS) ? A.Tin[(size_t)i * TS + q] : 0.0f;
        }
    }
    n = fast_load(A, Mo, i, a, o, n, sm, MODE != 0, ADJ, MODE == 2 ? tv : nullptr);
    if (lane < NM) wq[lane] = c_mono.w[lane];
    if (MODE == 1) {
#pragma unroll
        for (int h = 0; h < kPreD; ++h) {
            const int m = lane + h * kWave;
            if (m < ldg) rowbuf[m] = pre_d[h];
        }
        for (int m = kPreD * kWave + lane; m < ldg; m += kWave) rowbuf[m] = A.dEdG[(size_t)i * ldg + m];
        for (int t = lane; t < 3 * n; t += kWave) sP[t] = 0.0f;
    } else {
        for (int m = lane; m < ldg; m += kWave) rowbuf[m] = 0.0f;
    }
    wave_sync();
    fast_radial<GM>(A, Mo, a, sm, tv, rowbuf, sP, rowbuf);
    wave_sync();
    if (haveT) {
#pragma unroll
        for (int h = 0; h < kPreT; ++h) {
            const int q = lane + h * kWave;
            if (q < TS) T[q] = pre_t[h];
        }
        for (int q = kPreT * kWave + lane; q < TS; q += kWave) T[q] = A.Tin[(size_t)i * TS + q];
    }
    // powers of the unit vector (and, for the tangent, their derivatives along
    // the projected direction (t - r^ (r^.t)) / r) of neighbours c0 .. c1-1,
    // at most pw_chunk, into the chunk rows (the moment loop below walks each
    // species segment in such chunks, so for G and the adjoint the rows take a
    // fixed slice of LDS instead of one row per neighbour)
    auto power_rows = [&](int c0, int c1) {
        for (int J = c0 + lane; J < c1; J += kWave) {
            const float4 gq = sm.g[J];
            float* row = pw + (size_t)(J - c0) * PWS;
            const float v3[3] = {gq.x, gq.y, gq.z};
            float d3[3] = {0.f, 0.f, 0.f};
            if (MODE == 2) {
                const float4 t4 = tv[J];
                d3[0] = (t4.x - gq.x * t4.w) * gq.w;
                d3[1] = (t4.y - gq.y * t4.w) * gq.w;
                d3[2] = (t4.z - gq.z * t4.w) * gq.w;
            }
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                float p = 1.0f;
#pragma unroll
                for (int k = 0; k < ZP; ++k) {
                    if (MODE == 2) row[3 * ZP + x * ZP + k] = k == 0 ? 0.0f : (float)k * p * d3[x];
                    row[x * ZP + k] = k == 0 ? 1.0f : (p *= v3[x]);
                }
            }
        }
    };
    wave_sync();
    // moments T_b[u][alpha]: lane = (alpha, group); groups stride the neighbours.
    // G with zeta 4 (35 monomials, one group would leave 29 lanes idle):
    // monomials 1, x, y, z are summed with the diagonal sums below and the other
    // 31 run as two groups of 32 lanes, joined by one cross-half exchange.
    constexpr bool SPLIT = GM == 0 && NM == 35;
    constexpr int NGL = SPLIT ? 2 : NG;                        // neighbour groups of the loop
    if (!haveT) {
        const int al = SPLIT ? min(4 + lane % 32, NM - 1) : lane % NM;
        const int grp = SPLIT ? lane / 32 : lane / NM;
        const bool act = SPLIT ? lane % 32 < 31 : grp < NG;
        const int ea = c_mono.a[al], eb = c_mono.b[al], ec = c_mono.c[al];
        for (int b = 0; b < S; ++b) {
            const int s0 = sm.seg[b], s1 = sm.seg[b + 1];
            float acc[U4], dacc[U4];
#pragma unroll
            for (int u = 0; u < U4; ++u) { acc[u] = 0.0f; dacc[u] = 0.0f; }
            // one neighbour of the segment: T[u][alpha] += e_u(r) r^^alpha
            auto step = [&](const float* row, int J) {
                const float px = row[ea], py = row[ZP + eb], pz = row[2 * ZP + ec];
                const float mono = px * py * pz;
                const float4 e4 = sm.ef[J];
                acc[0] += e4.x * mono; acc[1] += e4.y * mono; acc[2] += e4.z * mono; acc[3] += e4.w * mono;
                if (MODE == 2) {
                    const float dmono = row[3 * ZP + ea] * py * pz + px * row[4 * ZP + eb] * pz +
                                        px * py * row[5 * ZP + ec];
                    const float4 f4 = sm.efd[J];
                    const float wm = tv[J].w * mono;
                    dacc[0] += f4.x * wm + e4.x * dmono;
                    dacc[1] += f4.y * wm + e4.y * dmono;
                    dacc[2] += f4.z * wm + e4.z * dmono;
                    dacc[3] += f4.w * wm + e4.w * dmono;
                }
            };
            const int chunk = pw_chunk(MODE, maxn);
            for (int c0 = s0; c0 < s1; c0 += chunk) {
                const int c1 = min(c0 + chunk, s1);
                wave_sync();                                    // the previous chunk's rows are read
                power_rows(c0, c1);
                wave_sync();
                if (act) {
                    int J = c0 + grp;
                    // four neighbours per trip: their power rows sit at fixed offsets
                    // from one address, so the loads use immediate offsets and overlap
                    for (; J + 3 * NGL < c1; J += 4 * NGL) {
                        const float* row = pw + (size_t)(J - c0) * PWS;
                        step(row, J);
                        step(row + NGL * PWS, J + NGL);
                        step(row + 2 * NGL * PWS, J + 2 * NGL);
                        step(row + 3 * NGL * PWS, J + 3 * NGL);
                    }
                    for (; J < c1; J += NGL) step(pw + (size_t)(J - c0) * PWS, J);
                }
            }
            if (SPLIT) {
#pragma unroll
                for (int u = 0; u < U4; ++u) acc[u] += __shfl_xor(acc[u], 32);
                if (act && grp == 0) {
#pragma unroll
                    for (int u = 0; u < U4; ++u) T[(b * U4 + u) * NM + al] = acc[u];
                }
            } else if (NG == 1) {
                if (act) {
#pragma unroll
                    for (int u = 0; u < U4; ++u) {
                        T[(b * U4 + u) * NM + al] = acc[u];
                        if (MODE == 2) dT[(b * U4 + u) * NM + al] = dacc[u];
                    }
                }
            } else {
                if (act) {
#pragma unroll
                    for (int u = 0; u < U4; ++u) {
                        part[(grp * NM + al) * U4 + u] = acc[u];
                        if (MODE == 2) part[NG * NM * U4 + (grp * NM + al) * U4 + u] = dacc[u];
                    }
                }
                wave_sync();
                for (int q = lane; q < NM * U4; q += kWave) {
                    float s = 0.0f, ds = 0.0f;
                    for (int gq = 0; gq < NG; ++gq) {
                        s += part[gq * NM * U4 + q];
                        if (MODE == 2) ds += part[NG * NM * U4 + gq * NM * U4 + q];
                    }
                    const int al2 = q / U4, u = q % U4;
                    T[(b * U4 + u) * NM + al2] = s;
                    if (MODE == 2) dT[(b * U4 + u) * NM + al2] = ds;
                }
                wave_sync();
            }
        }
        // diagonal sums D_b[u] (mode 2 also their tangent): lanes over the
        // segment, one transposing reduction of the four widths
        if (MODE != 1) {
            for (int b = 0; b < S; ++b) {
                float d4[U4], t4[U4];
#pragma unroll
                for (int u = 0; u < U4; ++u) { d4[u] = 0.0f; t4[u] = 0.0f; }
                float lo[SPLIT ? 4 * U4 : 1];                  // [u][1, x, y, z] (split G)
#pragma unroll
                for (int q = 0; q < (SPLIT ? 4 * U4 : 1); ++q) lo[q] = 0.0f;
                for (int J = sm.seg[b] + lane; J < sm.seg[b + 1]; J += kWave) {
                    const float4 e4 = sm.ef[J];
                    d4[0] += e4.x * e4.x; d4[1] += e4.y * e4.y; d4[2] += e4.z * e4.z; d4[3] += e4.w * e4.w;
                    if (SPLIT) {
                        const float4 gq = sm.g[J];
                        const float ev[U4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
                        for (int u = 0; u < U4; ++u) {
                            lo[4 * u] += ev[u];
                            lo[4 * u + 1] += ev[u] * gq.x;
                            lo[4 * u + 2] += ev[u] * gq.y;
                            lo[4 * u + 3] += ev[u] * gq.z;
                        }
                    }
                    if (MODE == 2) {
                        const float4 f4 = sm.efd[J];
                        const float w2 = 2.0f * tv[J].w;
                        t4[0] += w2 * e4.x * f4.x; t4[1] += w2 * e4.y * f4.y;
                        t4[2] += w2 * e4.z * f4.z; t4[3] += w2 * e4.w * f4.w;
                    }
                }
                int idx;
                const float sd = wave_sum_multi<U4>(d4, idx);
                if ((lane & (kWave / U4 - 1)) == 0) D[b * U4 + idx] = sd;
                if (SPLIT) {
                    // monomial order starts 1, x, y, z (degree 0, then x, y, z)
                    int qi;
                    const float sl = wave_sum_multi<(SPLIT ? 4 * U4 : 1)>(lo, qi);
                    if ((lane & (kWave / (4 * U4) - 1)) == 0) T[(b * U4 + qi / 4) * NM + qi % 4] = sl;
                }
                if (MODE == 2) {
                    const float st = wave_sum_multi<U4>(t4, idx);
                    if ((lane & (kWave / U4 - 1)) == 0) dD[b * U4 + idx] = st;
                }
            }
        }
    }
    wave_sync();
    if (MODE == 0 && A.Tout) {
        for (int q = lane; q < TS; q += kWave) A.Tout[(size_t)i * TS + q] = T[q];
    }
    // block contractions (modes 0, 2, 3) or their adjoint seeds (mode 1; mode 3
    // after its network)
    auto adjoint_seeds = [&]() {
        // lanes over (pair, width, power) of every block at once; each lane loads
        // all of its output indices and coefficients before using any, so the
        // block descriptors cost a few rounds of memory latency, not one per term
        for (int it = lane; it < NP * U4 * ZP; it += kWave) {
            const int p = it / (U4 * ZP), u = (it / ZP) % U4, m = it % ZP;
            const AngBlock& B = Mo.ang[a * NP + p];
            const int nk = min(max(B.nk, 0), kGridK), goff = B.goff;
            const bool use = B.n > 0 && u < B.ng;
            // u * nk + k < kGridG * kGridK for every k < kGridK (u < kGridG, nk <= kGridK):
            // entries past nk are read but not used
            int om[kGridK];
            float mc[kGridK];
#pragma unroll
            for (int k = 0; k < kGridK; ++k) {
                om[k] = B.outm[u * nk + k];
                mc[k] = B.mcoef[k][m];
            }
            float s = 0.0f;
            if (use) {
#pragma unroll
                for (int k = 0; k < kGridK; ++k)
                    if (k < nk) s += rowbuf[goff + om[k]] * mc[k];
            }
            Am[it] = s;
        }
    };
    if (MODE == 1) adjoint_seeds();
    for (int it = lane; MODE != 1 && it < NP * U4 * ZP; it += kWave) {
        const int p = it / (U4 * ZP), u = (it / ZP) % U4, m = it % ZP;
        int c = 0;
        while ((c + 1) * (c + 2) / 2 <= p) ++c;
        const int b = p - c * (c + 1) / 2;
        float s = 0.0f;
        {
            const float* Tb = T + (b * U4 + u) * NM;
            const float* Tc = T + (c * U4 + u) * NM;
            const float* dTb = dT + (b * U4 + u) * NM;
            const float* dTc = dT + (c * U4 + u) * NM;
            // degree m has (m + 1)(m + 2) / 2 monomials; every lane runs the
            // length of the highest degree with the surplus terms weighted 0, so
            // all loads are issued before the first product instead of one
            // dependent round per monomial
            constexpr int QD = (Z + 1) * (Z + 2) / 2;
            const int q0 = mono_start(m), q1 = mono_start(m + 1);
#pragma unroll
            for (int j = 0; j < QD; ++j) {
                const int q = min(q0 + j, NM - 1);
                const float w = q0 + j < q1 ? wq[q] : 0.0f;
                s += GM == 0 ? w * Tb[q] * Tc[q] : w * (dTb[q] * Tc[q] + Tb[q] * dTc[q]);
            }
            if (b == c) s = 0.5f * (s - (GM == 0 ? D[b * U4 + u] : dD[b * U4 + u]));
        }
        Am[it] = s;
    }
    wave_sync();
    if (MODE != 1) {
        // lanes over (pair, width, combination) of every block at once, loads
        // issued before the guards that use them
        for (int it = lane; it < NP * kGridG * kGridK; it += kWave) {
            const int p = it / (kGridG * kGridK), g = (it / kGridK) % kGridG, k = it % kGridK;
            const AngBlock& B = Mo.ang[a * NP + p];
            const int nk = min(max(B.nk, 0), kGridK), goff = B.goff;
            const bool use = B.n > 0 && g < B.ng && k < nk;
            const int om = B.outm[g * nk + k];   // < kGridG * kGridK; used only when use
            float mc[ZP];
#pragma unroll
            for (int m = 0; m < ZP; ++m) mc[m] = B.mcoef[k][m];
            if (use) {
                float v = 0.0f;
#pragma unroll
                for (int m = 0; m < ZP; ++m) v += mc[m] * Am[(p * U4 + g) * ZP + m];
                rowbuf[goff + om] = v;
            }
        }
        wave_sync();
        if (!FUSED) {
            float* out = (MODE == 0 ? A.G : A.v) + (size_t)i * ldg;
            for (int m = lane; m < ldg; m += kWave) out[m] = rowbuf[m];
            return;
        }
    }
    if (FUSED) {
        // the atom's G row in rowbuf: its energy and dE/dG (in place), then the
        // radial adjoint and the angular seeds as the adjoint kernel forms them
        fused_net(A, Mo.net[a], i, rowbuf, nb);
        for (int t = lane; t < 3 * n; t += kWave) sP[t] = 0.0f;
        wave_sync();
        fast_radial<1>(A, Mo, a, sm, tv, rowbuf, sP, rowbuf);
        wave_sync();
        adjoint_seeds();
        wave_sync();
    }
    // adjoint: dE/dT_b[alpha][u] and dE/dD_b[u]
    auto dEdT = [&](int it) {
        const int b = it / (NM * U4), al = (it / U4) % NM, u = it % U4;
        int m = 0;   // degree of monomial al
#pragma unroll
        for (int mm = 1; mm <= Z; ++mm) m += al >= mono_start(mm) ? 1 : 0;
        float s = 0.0f;
        for (int c = 0; c < S; ++c) s += Am[(pair_of(b, c) * U4 + u) * ZP + m] * T[(c * U4 + u) * NM + al];
        return wq[al] * s;
    };
    if (ub_in_T) {
        // every lane's entries in registers, then written over the moments
        float ub[kUbRegs];
#pragma unroll
        for (int h = 0; h < kUbRegs; ++h) {
            const int it = lane + h * kWave;
            ub[h] = it < S * NM * U4 ? dEdT(it) : 0.0f;
        }
        wave_sync();
#pragma unroll
        for (int h = 0; h < kUbRegs; ++h) {
            const int it = lane + h * kWave;
            if (it < S * NM * U4) Ub[it] = ub[h];
        }
    } else {
        for (int it = lane; it < S * NM * U4; it += kWave) Ub[it] = dEdT(it);
    }
    if (lane < S * U4) {
        const int b = lane / U4, u = lane % U4;
        float s = 0.0f;
        for (int m = 0; m < ZP; ++m) s += Am[(pair_of(b, b) * U4 + u) * ZP + m];
        dD[lane] = -0.5f * s;
    }
    wave_sync();
    // per neighbour: Q(r^) = sum_alpha W_alpha r^^alpha and grad of sum V_alpha r^^alpha
    constexpr MonoTable mt{};
    for (int J = lane; J < n; J += kWave) {
        int b = 0;
        while (b + 1 < S && J >= sm.seg[b + 1]) ++b;
        const float4 gq = sm.g[J];
        const float4 e4 = sm.ef[J], f4 = sm.efd[J];
        // every monomial of r^ once (one product each from a lower one); the
        // gradient of x^a y^b z^c is (a, b, c) times monomials of one degree less
        float mono[NM];
        mono[0] = 1.0f;
#pragma unroll
        for (int q = 1; q < NM; ++q) {
            const int ea = mt.a[q], eb = mt.b[q], ec = mt.c[q];
            mono[q] = ea > 0 ? mono[mono_index(ea > 0 ? ea - 1 : 0, eb, ec)] * gq.x
                             : (eb > 0 ? mono[mono_index(ea, eb > 0 ? eb - 1 : 0, ec)] * gq.y
                                       : mono[mono_index(ea, eb, ec > 0 ? ec - 1 : 0)] * gq.z);
        }
        float Qd = 0.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
        const float4* Ubb = reinterpret_cast<const float4*>(Ub + b * NM * U4);
#pragma unroll
        for (int q = 0; q < NM; ++q) {
            const int ea = mt.a[q], eb = mt.b[q], ec = mt.c[q];
            const float4 w4 = Ubb[q];
            const float V = e4.x * w4.x + e4.y * w4.y + e4.z * w4.z + e4.w * w4.w;
            const float W = f4.x * w4.x + f4.y * w4.y + f4.z * w4.z + f4.w * w4.w;
            Qd += W * mono[q];
            if (ea > 0) g0 += (V * (float)ea) * mono[mono_index(ea > 0 ? ea - 1 : 0, eb, ec)];
            if (eb > 0) g1 += (V * (float)eb) * mono[mono_index(ea, eb > 0 ? eb - 1 : 0, ec)];
            if (ec > 0) g2 += (V * (float)ec) * mono[mono_index(ea, eb, ec > 0 ? ec - 1 : 0)];
        }
        const float dd = 2.0f * (dD[b * U4] * e4.x * f4.x + dD[b * U4 + 1] * e4.y * f4.y +
                                 dD[b * U4 + 2] * e4.z * f4.z + dD[b * U4 + 3] * e4.w * f4.w);
        const float rg = gq.x * g0 + gq.y * g1 + gq.z * g2;
        const float radial = Qd + dd;
        sP[3 * J] += radial * gq.x + (g0 - gq.x * rg) * gq.w;
        sP[3 * J + 1] += radial * gq.y + (g1 - gq.y * rg) * gq.w;
        sP[3 * J + 2] += radial * gq.z + (g2 - gq.z * rg) * gq.w;
    }
    wave_sync();
    for (int J = lane; J < n; J += kWave) {
        const int t = sm.canon[J];
        A.P[3 * ((size_t)o + t)] = sP[3 * J];
        A.P[3 * ((size_t)o + t) + 1] = sP[3 * J + 1];
        A.P[3 * ((size_t)o + t) + 2] = sP[3 * J + 2];
    }
}

// Grid dispatch: a block holds blockDim.x / 64 atoms (the "atoms per
// workgroup" tuning).
template <int MODE, int Z>
__global__ __launch_bounds__(256) void k_moment(SymmArgs A) {
    extern __shared__ float4 smem_all[];
    const int wave = threadIdx.x / kWave;
    const int i = blockIdx.x * (blockDim.x / kWave) + wave;
    if (i >= A.N) return;
    moment_atom<MODE, Z>(A, i, smem_all + (size_t)wave * (A.smem_wave / 16));
}

// G: the same with registers held to five wavefronts per SIMD (it fits in 96
// VGPRs without spilling; its LDS per wavefront allows as many)
template <int Z>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_moment_g(SymmArgs A) {
    extern __shared__ float4 smem_all[];
    const int wave = threadIdx.x / kWave;
    const int i = blockIdx.x * (blockDim.x / kWave) + wave;
    if (i >= A.N) return;
    moment_atom<0, Z>(A, i, smem_all + (size_t)wave * (A.smem_wave / 16));
}

// Persistent dispatch (a separate kernel, so that the grid kernel keeps its
// own register allocation): a grid sized to the resident capacity, every
// wavefront pulls (first atom, count) items off the queue of its XCD (blocks
// go to XCD blockIdx.x % 8) until the queue is empty; the items of a batch
// run largest structure first.
template <int MODE, int Z>
__global__ __launch_bounds__(256) void k_moment_queue(SymmArgs A) {
    extern __shared__ float4 smem_all[];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    float4* smem4 = smem_all + (size_t)wave * (A.smem_wave / 16);
    const int q = blockIdx.x % kXcds;
    const int beg = A.qoff[q], cnt = A.qoff[q + 1] - beg;
    int* ctr = A.queue + 32 * q;
    int k = 0;
    if (lane == 0) k = atomicAdd(ctr, 1);
    k = __shfl(k, 0);
    while (k < cnt) {
        // the next item is claimed before this one runs: the atomic's round
        // trip overlaps the item's own loads
        int nk = 0;
        if (lane == 0) nk = atomicAdd(ctr, 1);
        const int2 it = A.items[beg + k];
        for (int a = it.x; a < it.x + it.y; ++a) moment_atom<MODE, Z>(A, a, smem4);
        k = __shfl(nk, 0);
    }
}

__global__ void k_queue_reset(int* __restrict__ queue) {
    if (threadIdx.x < kXcds) queue[32 * threadIdx.x] = 0;
}

// LDS bytes of one wavefront of moment_atom<mode> (the layout at its top)
static size_t moment_smem(const SymmArgs& A, int mode, int Z, int S, int NP) {
    const int NM = nmono(Z), ZP = Z + 1, U4 = kMaxUnionEta;
    const bool adj = mode == 1 || mode == 3;
    const int PWS = (mode == 1 && A.Tin) ? 2 : (3 * ZP * (mode == 2 ? 2 : 1)) | 1;
    const int NG = kWave / NM;
    size_t b = 16 * (size_t)A.maxn * (2 + (mode == 0 ? 0 : 1) + (mode == 2 ? 1 : 0));
    b += 4 * (adj ? 2 : 1) * (size_t)A.maxn + 4 * 16;                              // r, canon (adjoint), seg
    size_t f = (size_t)A.ldg + (mode == 1 ? 3 * (size_t)A.maxn : 0) + (mode == 3 ? 2 * (size_t)A.maxn : 0) +
               (size_t)pw_floats(mode, A.maxn, PWS, mode == 1 && A.Tin);
    f += (size_t)S * U4 * NM * (mode == 2 ? 2 : 1) + 2 * (size_t)S * U4 + (size_t)NP * U4 * ZP;
    f += (adj && S * NM * U4 > kUbRegs * kWave ? (size_t)S * NM * U4 : 0);         // dE/dT beside the moments
    f += (NG > 1 ? (size_t)NG * NM * U4 * (mode == 2 ? 2 : 1) : 0);
    f += NM;
    return b + 4 * f;
}

// ================================================================ host ====

// Work items of the persistent dispatch: every structure cut into ranges of
// up to kItemAtoms atoms (fewer when the batch would leave resident
// wavefronts without items: about four items per wavefront the chip holds),
// structures in decreasing size, items dealt round robin to the XCD queues so
// that each queue gets an equal share of every size.
constexpr int kItemAtoms = 8;

hipError_t build_items(ann_ctx* ctx) {
    const int S = ctx->nstruc;
    hipError_t e;
    if (!ctx->d_queue) {
        if ((e = hipMalloc(&ctx->d_queue, sizeof(int) * 32 * kXcds)) != hipSuccess) return e;
        if ((e = hipMalloc(&ctx->d_qoff, sizeof(int) * (kXcds + 1))) != hipSuccess) return e;
        int dev = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
        if ((e = hipDeviceGetAttribute(&ctx->num_cu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
            return e;
    }
    const int resident = std::max(1, ctx->num_cu) * 32;            // wavefronts per CU the kernels can hold
    const int len = std::max(1, std::min(kItemAtoms, ctx->atom_off[S] / (4 * resident)));
    std::vector<int> order(S);
    for (int s = 0; s < S; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return ctx->atom_off[a + 1] - ctx->atom_off[a] > ctx->atom_off[b + 1] - ctx->atom_off[b];
    });
    std::vector<int2> all;
    for (int s : order)
        for (int a = ctx->atom_off[s]; a < ctx->atom_off[s + 1]; a += len)
            all.push_back(make_int2(a, std::min(len, ctx->atom_off[s + 1] - a)));
    std::vector<int2> items;
    items.reserve(all.size());
    std::vector<int> qoff(kXcds + 1, 0);
    for (int q = 0; q < kXcds; ++q) {
        qoff[q] = (int)items.size();
        for (size_t k = q; k < all.size(); k += kXcds) items.push_back(all[k]);
    }
    qoff[kXcds] = (int)items.size();
    if ((int)items.size() > ctx->items_cap) {
        if (ctx->d_items && (e = hipFree(ctx->d_items)) != hipSuccess) return e;
        ctx->d_items = nullptr;
        ctx->items_cap = 0;
        if ((e = hipMalloc(&ctx->d_items, sizeof(int2) * items.size())) != hipSuccess) return e;
        ctx->items_cap = (int)items.size();
    }
    if ((e = staged_upload(ctx, kStItems, ctx->d_items, items.data(), sizeof(int2) * items.size())) != hipSuccess)
        return e;
    if ((e = staged_upload(ctx, kStQoff, ctx->d_qoff, qoff.data(), sizeof(int) * (kXcds + 1))) != hipSuccess)
        return e;
    ctx->items_valid = true;
    return hipSuccess;
}

// Work items of the current batch, built on the first persistent dispatch
// after ann_set_batch (host set-up and uploads, so callers that capture a
// graph build them before the capture).
hipError_t ensure_items(ann_ctx* ctx) {
    if (ctx->items_valid) return hipSuccess;
    if (ctx->capturing) return hipErrorStreamCaptureUnsupported;
    return build_items(ctx);
}

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
    size_t b = 16 * (size_t)A.maxn * (2 + (mode == 0 ? 0 : 1) + (mode == 2 ? 1 : 0));   // g, ef, efd, tv
    b += 8 * (size_t)A.maxn + 4 * 2 * (size_t)A.maxn + 4 * 16;                         // fg, r, canon, seg
    b += 4 * (size_t)A.ldg;
    if (mode == 1) b += 12 * (size_t)A.maxn;
    return b;
}

template <int MODE>
static hipError_t run(ann_ctx* ctx, int form, SymmArgs A) {
    if (A.N == 0) return hipSuccess;
    const DevModel& M = ctx->hm;
    const bool fast = form != ANN_FORM_PLAIN && M.all_grid && M.fast_ng > 0;
    const bool moment = fast && form == ANN_FORM_FAST && M.moment_z > 0 && ctx->tune_angular == 0;
    const char* names[3][3] = {{"symm_fwd_plain", "symm_bwd_plain", "symm_tan_plain"},
                               {"symm_fwd_fast", "symm_bwd_fast", "symm_tan_fast"},
                               {"symm_fwd_mom", "symm_bwd_mom", "symm_tan_mom"}};
    Launch L(ctx, names[moment ? 2 : (fast ? 1 : 0)][MODE]);
    if (moment) {
        // the moment store holds work_atoms rows: it is used only when the batch
        // fits (a forward call on a grown batch before any fused call skips it)
        const bool tfits = ctx->d_T && ctx->work_atoms >= A.N;
        if (MODE == 0 && tfits) A.Tout = ctx->d_T;
        if (MODE == 1 && tfits && ctx->T_gen == ctx->gen) A.Tin = ctx->d_T;
        const size_t sw = (moment_smem(A, MODE, M.moment_z, M.S, M.NP) + 15) / 16 * 16;
        int apb = ctx->tune_apb[MODE];
        while (apb > 1 && sw * apb > 65536) apb /= 2;          // dynamic LDS of one workgroup
        A.smem_wave = (int)sw;
        unsigned grid = (unsigned)((A.N + apb - 1) / apb), block = 64u * apb;
        const size_t sm = sw * apb;
        const bool queued = ctx->tune_queue == 1;
        if (queued) {
            const hipError_t ei = ensure_items(ctx);
            if (ei != hipSuccess) return ei;
        }
        void (*kern)(SymmArgs) =
            queued ? (M.moment_z == 1 ? k_moment_queue<MODE, 1> : M.moment_z == 2 ? k_moment_queue<MODE, 2>
                      : M.moment_z == 3 ? k_moment_queue<MODE, 3> : M.moment_z == 4 ? k_moment_queue<MODE, 4>
                      : k_moment_queue<MODE, 5>)
                   : MODE == 0 ? (M.moment_z == 1 ? k_moment_g<1> : M.moment_z == 2 ? k_moment_g<2>
                                  : M.moment_z == 3 ? k_moment_g<3> : M.moment_z == 4 ? k_moment_g<4> : k_moment_g<5>)
                   : (M.moment_z == 1 ? k_moment<MODE, 1> : M.moment_z == 2 ? k_moment<MODE, 2>
                      : M.moment_z == 3 ? k_moment<MODE, 3> : M.moment_z == 4 ? k_moment<MODE, 4>
                      : k_moment<MODE, 5>);
        if (queued) {
            // persistent: enough workgroups for the most wavefronts a CU can hold
            // (32); those that find no room start later, find the queues empty
            // and exit.  A multiple of the XCD count, so every queue has workers.
            const unsigned cap = (unsigned)std::max(1, ctx->num_cu) * (32u / (unsigned)apb);
            grid = std::min(grid, cap);
            grid = std::max(kXcds, (int)((grid + kXcds - 1) / kXcds * kXcds));
            A.queue = ctx->d_queue;
            A.items = ctx->d_items;
            A.qoff = ctx->d_qoff;
            k_queue_reset<<<1, 64, 0, ctx->stream>>>(ctx->d_queue);
        }
        kern<<<grid, block, sm, ctx->stream>>>(A);
        if (MODE == 0) ctx->T_gen = A.Tout ? ctx->gen : ~0ull;
        return hipGetLastError();
    }
    if (fast) {
        const size_t sm = fast_smem(A, MODE);
#define X(G, K)                                                                \
        if (M.fast_ng == G && M.fast_nk == K) {                                \
            k_fast<MODE, G, K><<<A.N, 64, sm, ctx->stream>>>(A);               \
            return hipGetLastError();                                          \
        }
        ANN_SHAPES(X)
#undef X
        return hipErrorInvalidValue;
    }
    const size_t sm = plain_smem(A, MODE);
    k_plain<MODE><<<A.N, 64, sm, ctx->stream>>>(A);
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

// Fused E+F: G, the network and the force adjoint of each atom in one
// moment-form kernel (moment_atom<3>), for FAST in the moment form with FP32
// inputs and networks of at most 128 inputs and 64 neurons per layer.
bool fused_supported(const ann_ctx* ctx, int form) {
    const DevModel& M = ctx->hm;
    if (form != ANN_FORM_FAST || !M.all_grid || M.fast_ng <= 0 || M.moment_z <= 0 || ctx->tune_angular != 0 ||
        ctx->npre != 0 || ctx->natoms == 0)
        return false;
    for (int a = 0; a < M.S; ++a) {
        const NetDesc& D = M.net[a];
        if (D.nin > 2 * kWave) return false;
        for (int l = 0; l < D.nlayer - 1; ++l)
            if (D.nodes[l] > kVecMaxWidth) return false;
    }
    const SymmArgs A = make_args(const_cast<ann_ctx*>(ctx));   // reads only
    return moment_smem(A, 3, M.moment_z, M.S, M.NP) <= 65536;
}

hipError_t symm_fused(ann_ctx* ctx, const float* inpb, float* Eatom, float* P) {
    SymmArgs A = make_args(ctx);
    if (A.N == 0) return hipSuccess;
    hipError_t e;
    if (!ctx->wchunk_valid && (e = chunk_weights(ctx)) != hipSuccess) return e;
    A.P = P;
    A.Eatom = Eatom;
    A.wchunk = ctx->d_wchunk;
    A.wtrans = ctx->d_wtrans;
    A.inpw = ctx->da.inpw;
    A.inpb = inpb;
    const DevModel& M = ctx->hm;
    Launch L(ctx, "ef_fused");
    const size_t sw = (moment_smem(A, 3, M.moment_z, M.S, M.NP) + 15) / 16 * 16;
    int apb = ctx->tune_apb[1];                             // the adjoint's atoms per workgroup
    while (apb > 1 && sw * apb > 65536) apb /= 2;
    A.smem_wave = (int)sw;
    const unsigned grid = (unsigned)((A.N + apb - 1) / apb), block = 64u * apb;
    void (*kern)(SymmArgs) = M.moment_z == 1 ? k_moment<3, 1> : M.moment_z == 2 ? k_moment<3, 2>
                             : M.moment_z == 3 ? k_moment<3, 3> : M.moment_z == 4 ? k_moment<3, 4>
                             : k_moment<3, 5>;
    kern<<<grid, block, sw * apb, ctx->stream>>>(A);
    ctx->T_gen = ~0ull;                                     // no stored moments
    return hipGetLastError();
}

hipError_t symm_tangent(ann_ctx* ctx, int form, const float* u, float* v) {
    SymmArgs A = make_args(ctx);
    A.u = u;
    A.v = v;
    return run<2>(ctx, form, A);
}

bool fast_shape_supported(int ng, int nk) { return fast_shape_ok(ng, nk); }

}  // namespace ann
