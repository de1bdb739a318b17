// precise.hip -- network inputs whose FP32 rounding the input scaling would
// amplify, evaluated in FP64.
//
// The network input is s = inpw (G + inpb) (src/ml/nn.cpp:1346).  The
// reference trainer sets inpb = -<G> and inpw = 1/sigma(G) per input
// (src/nnp/nnptefr2.cpp:776-867), and an input that is almost constant over
// the training set (a rigid O-H bond seen by a short-range radial function:
// <G>/sigma ~ 1e5) turns the ~1e-7 relative rounding of an FP32 G into an
// O(1e-2) error of s.  Inputs with |inpw inpb| above a threshold are therefore
// recomputed here from FP64 positions, image vectors and FP64 basis parameters,
// and handed to the network as the FP64 sum G + inpb rounded to FP32 (the
// network then applies inpb = 0 to them).  Everything else stays FP32.
//
// Formulas: Cutoff::cutf (src/nnp/cutoff.cpp:62-113), BasisR::symm
// (src/nnp/basis_radial.cpp:263-309), BasisA::symm (src/nnp/basis_angular.cpp:
// 371-420).  One wavefront per atom; an item is one (input, radial block,
// function) or (input, angular block, function); lanes over neighbours
// (radial) or over first triplet legs with the second leg looped (angular);
// FP64 lane sums and a fixed butterfly, so the result is deterministic.
#include "basis_dev.h"
#include "pairvec_dev.h"

namespace ann {

struct PreciseArgs {
    const DevModel* dm;
    const PreItem* items;
    const double* rad_rs;
    const double* rad_eta;
    const double* ang_eta;
    const double* ang_zeta;
    const float* ang_lam;
    const double* inpb64;
    double rc;
    int N, ldg;
    const int* species;
    const int* sof;
    const double* pos;
    const double* cells;
    const int* off;
    const int* idx;
    const uint32_t* img;
    const uint8_t* nsp;
    float* Gc;                // centred FP32 output (or null)
    double* G64;              // FP64 output (or null)
};

__device__ inline double cut64(int name, double rc, double r) {
    if (name == CUT_STEP) return r <= rc ? 1.0 : 0.0;
    if (!(r < rc)) return 0.0;
    switch (name) {
        case CUT_COS: return 0.5 * (cos(3.141592653589793 * r / rc) + 1.0);
        case CUT_TANH: { const double t = tanh(1.0 - r / rc); return t * t * t; }
        case CUT_POLY3: { const double x = r / rc; return (2.0 * x - 3.0) * x * x + 1.0; }
        default: return 1.0;
    }
}

__device__ inline double sech64(double y) {
    const double e = exp(-fabs(y));
    return 2.0 * e / (1.0 + e * e);
}

// BasisR::symm value of one function at distance r (fc: cutoff value)
__device__ inline double radial64(int name, double eta, double rs, double r, double fc) {
    const double x = r - rs;
    switch (name) {
        case RAD_GAUSSIAN: { const double a = eta * x; return exp(-a * a) * fc; }
        case RAD_SECH: return sech64(eta * x) * fc;
        case RAD_LOGISTIC: { const double s = sech64(eta * x); return s * s * fc; }
        case RAD_TANH: return 0.5 * (tanh(-eta * x) + 1.0) * fc;
        case RAD_LOGCOSH: {   // log(1 + exp(y)) / 2, y = -2 eta x
            const double y = -2.0 * eta * x;
            return 0.5 * (fmax(y, 0.0) + log1p(exp(-fabs(y)))) * fc;
        }
        case RAD_LOGCOSH2: {  // a e / (e - 1), e = exp(2 sqrt2 a); the reference's a == 0 value
            const double k2 = 1.4142135623730951, a = -eta * x;
            if (a == 0.0) return 0.5 * fc / k2;
            const double em1 = expm1(2.0 * k2 * a);
            return a * (em1 + 1.0) / em1 * fc;
        }
        default: return 0.0;
    }
}

// BasisA::symm value of one function of a triplet (legs r0, r1, cosine c)
__device__ inline double angular64(int name, double ieta2, double zeta, double lamh, double r0, double r1, double c,
                                   double f0, double f1) {
    const double r2s = r0 * r0 + r1 * r1, hp = 1.5707963267948966;
    double fr;
    switch (name) {
        case ANG_GAUSS: fr = exp(-ieta2 * r2s); break;
        case ANG_GAUSS2: { const double e = exp(-hp * ieta2 * r2s); fr = 2.0 * e / (1.0 + e); } break;
        case ANG_SECH: { const double e = exp(-hp * ieta2 * r2s); fr = 2.0 * e / (1.0 + e * e); } break;
        default: {
            const int p = name == ANG_STUDENT3 ? 4 : (name == ANG_STUDENT4 ? 5 : 6);
            fr = pow(1.0 + ieta2 * r2s, -0.5 * p);
        } break;
    }
    return pow(fabs(0.5 + lamh * c), zeta) * fr * f0 * f1;
}

__device__ inline double ipow64(double x, int e) {
    double r = 1.0;
    for (int q = 0; q < e; ++q) r *= x;
    return r;
}

// monomials x^a y^b z^c of total degree <= 4 in degree order (as the FP32
// moment kernels number them) and their multinomial weights m!/(a! b! c!)
struct Mono64 {
    int a[35], b[35], c[35], m[35];
    double w[35];
    constexpr Mono64() : a(), b(), c(), m(), w() {
        int q = 0;
        for (int mm = 0; mm <= 4; ++mm)
            for (int aa = mm; aa >= 0; --aa)
                for (int bb = mm - aa; bb >= 0; --bb) {
                    const int cc = mm - aa - bb;
                    int fm = 1, fa = 1, fb = 1, fc = 1;
                    for (int t = 2; t <= mm; ++t) fm *= t;
                    for (int t = 2; t <= aa; ++t) fa *= t;
                    for (int t = 2; t <= bb; ++t) fb *= t;
                    for (int t = 2; t <= cc; ++t) fc *= t;
                    a[q] = aa; b[q] = bb; c[q] = cc; m[q] = mm;
                    w[q] = (double)(fm / (fa * fb * fc));
                    ++q;
                }
    }
};
__constant__ Mono64 c_mono64 = Mono64();

__device__ inline double wave_sum64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// LDS of one atom: neighbours sorted by species (stable), FP64
struct PreSmem {
    double4* v;      // [maxn] (unit vector, r)
    double* fc;      // [maxn]
    double4* ef;     // [maxn] exp(-r^2 / eta_u^2) fc for the union widths u (moment form)
    int* seg;        // [kMaxSpecies + 1]
    double* T;       // [S][4][NM] moments (moment form)
    double* Dg;      // [S][4] diagonal sums
    double* Am;      // [NP][4][Z + 1] degree-m pair sums
};

// One wavefront per atom.  Z: largest zeta of the moment form (0: the model
// has no moment form; angular inputs are then summed over triplets).
template <int Z>
__global__ __launch_bounds__(64) void k_precise(PreciseArgs A, int maxn) {
    extern __shared__ double4 smem64[];
    const int i = blockIdx.x, lane = threadIdx.x;
    if (i >= A.N) return;
    const DevModel& M = *A.dm;
    const int a = A.species[i];
    const int q0 = M.pre_off[a], q1 = M.pre_off[a + 1];
    if (q0 == q1) return;
    const int S = M.S, NP = M.NP;
    constexpr int NM = Z > 0 ? (Z + 1) * (Z + 2) * (Z + 3) / 6 : 1;
    PreSmem sm;
    sm.v = smem64;
    sm.ef = sm.v + maxn;
    sm.fc = reinterpret_cast<double*>(sm.ef + (Z > 0 ? maxn : 0));
    sm.T = sm.fc + maxn;
    sm.Dg = sm.T + (Z > 0 ? S * 4 * NM : 0);
    sm.Am = sm.Dg + (Z > 0 ? S * 4 : 0);
    sm.seg = reinterpret_cast<int*>(sm.Am + (Z > 0 ? NP * 4 * (Z + 1) : 0));
    const int o = A.off[i], n = A.off[i + 1] - o;
    const double* R = A.cells + 9 * A.sof[i];
    const uint64_t lt = (1ull << lane) - 1ull;
    // neighbours within the cutoff, regrouped by species (stable): count, then place
    int cnt[kMaxSpecies];
#pragma unroll
    for (int b = 0; b < kMaxSpecies; ++b) cnt[b] = 0;
    for (int q = 0; q < n; q += kWave) {
        const int t = q + lane;
        int sp = -1;
        if (t < n) {
            int n0, n1, n2;
            unpack_img(A.img[o + t], n0, n1, n2);
            double x, y, z, d2;
            pair_vec(A.pos, i, A.idx[o + t], R, n0, n1, n2, x, y, z, d2);
            if (d2 > 0.0 && sqrt(d2) <= A.rc) sp = A.nsp[o + t];
        }
#pragma unroll
        for (int b = 0; b < kMaxSpecies; ++b)
            if (b < S) cnt[b] += __popcll(__ballot(sp == b));
    }
    int base[kMaxSpecies];
    int tot = 0;
#pragma unroll
    for (int b = 0; b < kMaxSpecies; ++b) {
        base[b] = tot;
        if (b < S) tot += cnt[b];
    }
#pragma unroll
    for (int b = 0; b < kMaxSpecies; ++b)
        if (b < S && lane == b) sm.seg[b] = base[b];
    if (lane == S) sm.seg[S] = tot;
    const int U = M.nunion[a];
    for (int q = 0; q < n; q += kWave) {
        const int t = q + lane;
        int sp = -1;
        double x = 0, y = 0, z = 0, d2 = 0;
        if (t < n) {
            int n0, n1, n2;
            unpack_img(A.img[o + t], n0, n1, n2);
            pair_vec(A.pos, i, A.idx[o + t], R, n0, n1, n2, x, y, z, d2);
            if (d2 > 0.0 && sqrt(d2) <= A.rc) sp = A.nsp[o + t];
        }
        int p = 0;
#pragma unroll
        for (int b = 0; b < kMaxSpecies; ++b) {
            if (b < S) {
                const uint64_t mk = __ballot(sp == b);
                if (sp == b) p = base[b] + __popcll(mk & lt);
                base[b] += __popcll(mk);
            }
        }
        if (sp >= 0) {
            const double r = sqrt(d2), ir = 1.0 / r;
            const double f = cut64(M.cutoff, A.rc, r);
            sm.v[p] = make_double4(x * ir, y * ir, z * ir, r);
            sm.fc[p] = f;
            if (Z > 0) {
                double e[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) e[u] = u < U ? exp(-r * r * M.uieta2_64[a][u]) * f : 0.0;
                sm.ef[p] = make_double4(e[0], e[1], e[2], e[3]);
            }
        }
    }
    __syncthreads();
    const NetDesc& D = M.net[a];
    auto put = [&](int k, double g) {
        const size_t at = (size_t)i * A.ldg + k;
        if (A.Gc) A.Gc[at] = (float)(g + A.inpb64[D.inoff + k]);
        if (A.G64) A.G64[at] = g;
    };
    int q = q0;
    // radial items, one block (neighbour species) at a time: lanes over
    // (neighbour, function) pairs, W lanes per neighbour
    while (q < q1 && A.items[q].mode == 0) {
        const int blk = A.items[q].blk;
        int qe = q;
        while (qe < q1 && A.items[qe].mode == 0 && A.items[qe].blk == blk) ++qe;
        const int nf = qe - q;
        int W = 1;
        while (W < nf) W <<= 1;
        const int f = lane % W, g = lane / W, G = kWave / W;
        const RadBlock& rb = M.rad[a * S + blk];
        double rs = 0.0, eta = 0.0;
        if (f < nf) {
            const int fn = A.items[q + f].fn;
            rs = A.rad_rs[rb.par + fn];
            eta = A.rad_eta[rb.par + fn];
        }
        double acc = 0.0;
        if (f < nf)
            for (int J = sm.seg[blk] + g; J < sm.seg[blk + 1]; J += G) acc += radial64(rb.name, eta, rs, sm.v[J].w, sm.fc[J]);
        for (int bit = W; bit < kWave; bit <<= 1) acc += __shfl_xor(acc, bit);
        if (g == 0 && f < nf) put(A.items[q + f].k, acc);
        q = qe;
    }
    if (q == q1) return;
    if (Z > 0) {
        // moments T_b[u][alpha] = sum_{j in b} e_u(r_j) r^_j^alpha: lanes over alpha
        if (lane < NM) {
            const int ea = c_mono64.a[lane], eb = c_mono64.b[lane], ec = c_mono64.c[lane];
            for (int b = 0; b < S; ++b) {
                double t4[4] = {0.0, 0.0, 0.0, 0.0};
                for (int J = sm.seg[b]; J < sm.seg[b + 1]; ++J) {
                    const double4 v = sm.v[J], e = sm.ef[J];
                    const double mono = ipow64(v.x, ea) * ipow64(v.y, eb) * ipow64(v.z, ec);
                    t4[0] += e.x * mono; t4[1] += e.y * mono; t4[2] += e.z * mono; t4[3] += e.w * mono;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) sm.T[(b * 4 + u) * NM + lane] = t4[u];
            }
        }
        if (lane < S * 4) {
            const int b = lane / 4, u = lane % 4;
            double s = 0.0;
            for (int J = sm.seg[b]; J < sm.seg[b + 1]; ++J) {
                const double4 e = sm.ef[J];
                const double x = u == 0 ? e.x : (u == 1 ? e.y : (u == 2 ? e.z : e.w));
                s += x * x;
            }
            sm.Dg[lane] = s;
        }
        __syncthreads();
        // A_m(p, u) = sum_{|alpha| = m} w_alpha T_b T_c (b != c); (... - D_b) / 2 for b == c
        for (int it = lane; it < NP * 4 * (Z + 1); it += kWave) {
            const int p = it / (4 * (Z + 1)), u = (it / (Z + 1)) % 4, m = it % (Z + 1);
            int c = 0;
            while ((c + 1) * (c + 2) / 2 <= p) ++c;
            const int b = p - c * (c + 1) / 2;
            const int qa = m * (m + 1) * (m + 2) / 6, qb = (m + 1) * (m + 2) * (m + 3) / 6;
            double s = 0.0;
            for (int x = qa; x < qb; ++x) s += c_mono64.w[x] * sm.T[(b * 4 + u) * NM + x] * sm.T[(c * 4 + u) * NM + x];
            if (b == c) s = 0.5 * (s - sm.Dg[b * 4 + u]);
            sm.Am[it] = s;
        }
        __syncthreads();
        // G = sum_m C(zeta, m) 2^-zeta lambda^m A_m: one lane per item
        for (int t = q + lane; t < q1; t += kWave) {
            const PreItem it = A.items[t];
            if (it.mode != 1) continue;
            double g = 0.0, binom = 1.0, lp = 1.0;
            for (int m = 0; m <= it.zint; ++m) {
                g += binom * lp * sm.Am[(it.blk * 4 + it.u) * (Z + 1) + m];
                binom = binom * (it.zint - m) / (m + 1);
                lp *= (double)it.lam;
            }
            put(it.k, g * ldexp(1.0, -it.zint));
        }
    }
    // angular items outside the moment form: lanes over the first leg, the
    // second leg looped (each unordered pair of the block's species once)
    for (; q < q1; ++q) {
        const PreItem it = A.items[q];
        if (it.mode != 2) continue;
        const AngBlock& ab = M.ang[a * NP + it.blk];
        int c = 0;
        while ((c + 1) * (c + 2) / 2 <= it.blk) ++c;
        const int b = it.blk - c * (c + 1) / 2;
        const double eta = A.ang_eta[ab.par + it.fn];
        const double ieta2 = 1.0 / (eta * eta), zeta = A.ang_zeta[ab.par + it.fn];
        const double lamh = 0.5 * (double)A.ang_lam[ab.par + it.fn];
        double acc = 0.0;
        const int kb0 = sm.seg[c], kb1 = sm.seg[c + 1];
        for (int J = sm.seg[b] + lane; J < sm.seg[b + 1]; J += kWave) {
            const double4 vj = sm.v[J];
            for (int K = (b == c ? J + 1 : kb0); K < kb1; ++K) {
                const double4 vk = sm.v[K];
                const double cs = vj.x * vk.x + vj.y * vk.y + vj.z * vk.z;
                acc += angular64(ab.name, ieta2, zeta, lamh, vj.w, vk.w, cs, sm.fc[J], sm.fc[K]);
            }
        }
        acc = wave_sum64(acc);
        if (lane == 0) put(it.k, acc);
    }
}

hipError_t precise_inputs(ann_ctx* ctx, float* Gc, double* G64) {
    if (ctx->npre == 0 || ctx->natoms == 0) return hipSuccess;
    PreciseArgs A{};
    A.dm = ctx->dm;
    A.items = ctx->da.pitems;
    A.rad_rs = ctx->da.rad_rs64;
    A.rad_eta = ctx->da.rad_eta64;
    A.ang_eta = ctx->da.ang_eta64;
    A.ang_zeta = ctx->da.ang_zeta64;
    A.ang_lam = ctx->da.ang_lam;
    A.inpb64 = ctx->da.inpb64;
    A.rc = ctx->rc64;
    A.N = ctx->natoms;
    A.ldg = ctx->ldg;
    A.species = ctx->d_species;
    A.sof = ctx->d_struc_of;
    A.pos = ctx->d_pos;
    A.cells = ctx->d_cells;
    A.off = ctx->nl.off;
    A.idx = ctx->nl.idx;
    A.img = ctx->nl.img;
    A.nsp = ctx->nl.nsp;
    A.Gc = Gc;
    A.G64 = G64;
    const int maxn = std::max(ctx->nl.maxn, 1);
    const int Z = ctx->hm.moment_z, S = ctx->hm.S, NP = ctx->hm.NP;
    const int NM = Z > 0 ? (Z + 1) * (Z + 2) * (Z + 3) / 6 : 0;
    size_t smem = (size_t)maxn * (sizeof(double4) + sizeof(double)) + sizeof(int) * (kMaxSpecies + 1);
    if (Z > 0) smem += (size_t)maxn * sizeof(double4) + sizeof(double) * ((size_t)S * 4 * NM + S * 4 + (size_t)NP * 4 * (Z + 1));
    Launch L(ctx, "precise");
    switch (Z) {
        case 0: k_precise<0><<<ctx->natoms, 64, smem, ctx->stream>>>(A, maxn); break;
        case 1: k_precise<1><<<ctx->natoms, 64, smem, ctx->stream>>>(A, maxn); break;
        case 2: k_precise<2><<<ctx->natoms, 64, smem, ctx->stream>>>(A, maxn); break;
        case 3: k_precise<3><<<ctx->natoms, 64, smem, ctx->stream>>>(A, maxn); break;
        default: k_precise<4><<<ctx->natoms, 64, smem, ctx->stream>>>(A, maxn); break;
    }
    return hipGetLastError();
}

__global__ void k_widen(int64_t n, const float* __restrict__ G, double* __restrict__ G64) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) G64[e] = (double)G[e];
}

// Gc[i][k] = G64 + inpb64 (FP64 sum) where input k of the atom's species is
// FP64-evaluated, else G64 rounded
__global__ void k_centre(int N, int ldg, const DevModel* __restrict__ dm, const int* __restrict__ species,
                         const uint8_t* __restrict__ preflag, const double* __restrict__ inpb64,
                         const double* __restrict__ G64, float* __restrict__ Gc) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)N * ldg) return;
    const int i = (int)(e / ldg), k = (int)(e % ldg);
    const NetDesc& D = dm->net[species[i]];
    const double g = G64[e];
    Gc[e] = (k < D.nin && preflag[D.inoff + k]) ? (float)(g + inpb64[D.inoff + k]) : (float)g;
}

hipError_t widen_rows(ann_ctx* ctx, const float* G, double* G64) {
    const int64_t n = (int64_t)ctx->natoms * ctx->ldg;
    if (n == 0) return hipSuccess;
    k_widen<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(n, G, G64);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return precise_inputs(ctx, nullptr, G64);
}

hipError_t centre_rows(ann_ctx* ctx, const double* G64, float* Gc) {
    const int64_t n = (int64_t)ctx->natoms * ctx->ldg;
    if (n == 0) return hipSuccess;
    k_centre<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(ctx->natoms, ctx->ldg, ctx->dm, ctx->d_species,
                                                                  ctx->da.preflag, ctx->da.inpb64, G64, Gc);
    return hipGetLastError();
}

}  // namespace ann
