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
    const float4* geo;        // FP32 pair geometry (prefilter of the cutoff test)
    float* Tout;              // FP32 moments for the force adjoint (or null)
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

// monomials x^a y^b z^c of total degree <= kMomZ in degree order (as the FP32
// moment kernels number them) and their multinomial weights m!/(a! b! c!)
constexpr int kMonoAll = (kMomZ + 1) * (kMomZ + 2) * (kMomZ + 3) / 6;
struct MonoFixed {
    int a[kMonoAll], b[kMonoAll], c[kMonoAll], m[kMonoAll];
    double w[kMonoAll];
    constexpr MonoFixed() : a(), b(), c(), m(), w() {
        int q = 0;
        for (int mm = 0; mm <= kMomZ; ++mm)
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

__device__ inline double wave_sum64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Sum NV per-lane doubles over the wavefront with a transposing butterfly:
// afterwards lane l holds the total of value index idx (lanes l with
// l % (64 / NV) == 0 hold distinct indices).  NV a power of two <= 64.
template <int NV>
__device__ __forceinline__ double wave_sum_multi64(double (&v)[NV], int& idx) {
    const int lane = threadIdx.x % kWave;
    int base = 0;
#pragma unroll
    for (int h = NV / 2, bit = 32; h >= 1; h >>= 1, bit >>= 1) {
        const bool upper = (lane & bit) != 0;
#pragma unroll
        for (int q = 0; q < h; ++q) {
            const double send = upper ? v[q] : v[q + h];
            const double keep = upper ? v[q + h] : v[q];
            v[q] = keep + __shfl_xor(send, bit);
        }
        if (upper) base += h;
    }
    double s = v[0];
    constexpr int steps = (NV >= 64) ? 0 : 6 - __builtin_ctz(NV);
    int bit = 1;
#pragma unroll
    for (int q = 0; q < steps; ++q, bit <<= 1) s += __shfl_xor(s, bit);
    idx = base;
    return s;
}

// tanh(x) for x >= 0 from expm1, accurate to a few ulp
__device__ inline double tanh64(double x) {
    const double em = expm1(-2.0 * x);
    return -em / (2.0 + em);
}

__device__ inline double cut64f(int name, double rc, double r) {
    if (name == CUT_TANH) {
        if (!(r < rc)) return 0.0;
        const double t = tanh64(1.0 - r / rc);
        return t * t * t;
    }
    return cut64(name, rc, r);
}

// One wavefront per atom.  Z: largest zeta of the moment form (0: the model
// has no moment form; angular inputs are then summed over triplets).  With
// A.Tout set (moment form only) the moments of every atom are also written in
// FP32, in the layout of the FP32 moment kernels, for the force adjoint.
template <int Z>
__global__ __launch_bounds__(64) void k_precise(PreciseArgs A, int maxn) {
    extern __shared__ double4 smem64[];
    const int i = blockIdx.x, lane = threadIdx.x;
    if (i >= A.N) return;
    const DevModel& M = *A.dm;
    const int a = A.species[i];
    const int q0 = M.pre_off[a], q1 = M.pre_off[a + 1];
    const bool wantT = Z > 0 && A.Tout != nullptr;
    if (q0 == q1 && !wantT) return;
    const int S = M.S, NP = M.NP;
    constexpr int ZP = Z + 1;
    constexpr int NM = Z > 0 ? (Z + 1) * (Z + 2) * (Z + 3) / 6 : 1;
    constexpr int PWS = Z > 0 ? (3 * Z) | 1 : 1;          // powers k = 1..Z of the three components
    constexpr int NMH = Z > 1 ? NM - 4 : 1;               // monomials of degree >= 2
    constexpr int GH = Z > 1 ? kWave / NMH : 1;           // neighbour groups of the degree >= 2 sums
    // LDS: (r, fc) and, moment form, widths and powers; triplet form, unit vectors
    double2* rf = reinterpret_cast<double2*>(smem64);
    double4* ef = reinterpret_cast<double4*>(rf + maxn);                   // [maxn] (Z > 0)
    double4* uv = ef;                                                        // [maxn] unit vectors (Z == 0)
    double* pw = reinterpret_cast<double*>(ef + maxn);                      // [maxn][PWS] (Z > 0)
    double* T = pw + (Z > 0 ? (size_t)maxn * PWS : 0);                      // [S][4][NM]
    double* Dg = T + (Z > 0 ? S * 4 * NM : 0);                              // [S][4]
    double* Am = Dg + (Z > 0 ? S * 4 : 0);                                  // [NP][4][ZP]
    double* part = Am + (Z > 0 ? NP * 4 * ZP : 0);                          // [GH][NMH][4]
    int* seg = reinterpret_cast<int*>(part + (Z > 1 ? GH * NMH * 4 : 0));   // [kMaxSpecies + 1]
    const int o = A.off[i], n = A.off[i + 1] - o;
    const double* R = A.cells + 9 * A.sof[i];
    const uint64_t lt = (1ull << lane) - 1ull;
    const double rc = A.rc;
    const float rc_lo = (float)(rc * (1.0 - 1e-5)), rc_hi = (float)(rc * (1.0 + 1e-5));
    // neighbours within the cutoff (the FP32 distance decides except in a thin
    // shell around rc, where the FP64 one does), regrouped by species (stable)
    int cnt[kMaxSpecies];
#pragma unroll
    for (int b = 0; b < kMaxSpecies; ++b) cnt[b] = 0;
    for (int q = 0; q < n; q += kWave) {
        const int t = q + lane;
        int sp = -1;
        if (t < n) {
            const float r32 = A.geo[o + t].w;
            bool keep = r32 < rc_lo;
            if (!keep && r32 <= rc_hi) {
                int n0, n1, n2;
                unpack_img(A.img[o + t], n0, n1, n2);
                double x, y, z, d2;
                pair_vec(A.pos, i, A.idx[o + t], R, n0, n1, n2, x, y, z, d2);
                keep = sqrt(d2) <= rc;
            }
            if (keep) sp = A.nsp[o + t];
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
        if (b < S && lane == b) seg[b] = base[b];
    if (lane == S) seg[S] = tot;
    const int U = M.nunion[a];
    for (int q = 0; q < n; q += kWave) {
        const int t = q + lane;
        int sp = -1;
        double x = 0, y = 0, z = 0, d2 = 0;
        if (t < n) {
            const float r32 = A.geo[o + t].w;
            if (r32 <= rc_hi) {
                int n0, n1, n2;
                unpack_img(A.img[o + t], n0, n1, n2);
                pair_vec(A.pos, i, A.idx[o + t], R, n0, n1, n2, x, y, z, d2);
                if (r32 < rc_lo || sqrt(d2) <= rc) sp = A.nsp[o + t];
            }
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
            const double f = cut64f(M.cutoff, rc, r);
            rf[p] = make_double2(r, f);
            const double h[3] = {x * ir, y * ir, z * ir};
            if (Z > 0) {
                double e[4];
                const double r2 = r * r;
#pragma unroll
                for (int u = 0; u < 4; ++u) e[u] = u < U ? exp(-r2 * M.uieta2_64[a][u]) * f : 0.0;
                ef[p] = make_double4(e[0], e[1], e[2], e[3]);
                double* row = pw + (size_t)p * PWS;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double qv = h[c];
#pragma unroll
                    for (int k = 0; k < Z; ++k) {
                        row[c * Z + k] = qv;
                        qv *= h[c];
                    }
                }
            } else {
                uv[p] = make_double4(h[0], h[1], h[2], 0.0);
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
            for (int J = seg[blk] + g; J < seg[blk + 1]; J += G) {
                const double2 x = rf[J];
                acc += radial64(rb.name, eta, rs, x.x, x.y);
            }
        for (int bit = W; bit < kWave; bit <<= 1) acc += __shfl_xor(acc, bit);
        if (g == 0 && f < nf) put(A.items[q + f].k, acc);
        q = qe;
    }
    if (q == q1 && !wantT) return;
    if constexpr (Z > 0) {
        static_assert(Z <= kMomZ, "the monomial table covers the moment form's degrees");
        constexpr MonoFixed mo{};
        for (int b = 0; b < S; ++b) {
            const int s0 = seg[b], s1 = seg[b + 1];
            // degree <= 1 moments (1, x, y, z) and the diagonal sums: lanes over
            // neighbours, one transposing reduction of 4 x 5 values (padded to 32)
            {
                double v[32];
#pragma unroll
                for (int x = 0; x < 32; ++x) v[x] = 0.0;
                for (int J = s0 + lane; J < s1; J += kWave) {
                    const double4 e4 = ef[J];
                    const double* row = pw + (size_t)J * PWS;
                    const double hx = row[0], hy = row[Z], hz = row[2 * Z];
                    const double ev[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        v[5 * u] += ev[u];
                        v[5 * u + 1] += ev[u] * hx;
                        v[5 * u + 2] += ev[u] * hy;
                        v[5 * u + 3] += ev[u] * hz;
                        v[5 * u + 4] += ev[u] * ev[u];
                    }
                }
                int idx;
                const double sv = wave_sum_multi64<32>(v, idx);
                if ((lane & 1) == 0 && idx < 20) {
                    const int u = idx / 5, c = idx % 5;
                    if (c < 4) T[(b * 4 + u) * NM + c] = sv;
                    else Dg[b * 4 + u] = sv;
                }
            }
            // degree >= 2: lane = (monomial, neighbour group), powers from the rows
            if (Z > 1) {
                const int al = 4 + lane % NMH, g = lane / NMH;
                if (g < GH) {
                    const int ea = mo.a[al], eb = mo.b[al], ec = mo.c[al];
                    double t4[4] = {0.0, 0.0, 0.0, 0.0};
                    auto pv = [&](const double* row, int c, int e) { return e == 0 ? 1.0 : row[c * Z + e - 1]; };
                    for (int J = s0 + g; J < s1; J += GH) {
                        const double* row = pw + (size_t)J * PWS;
                        const double mono = pv(row, 0, ea) * pv(row, 1, eb) * pv(row, 2, ec);
                        const double4 e4 = ef[J];
                        t4[0] += e4.x * mono; t4[1] += e4.y * mono; t4[2] += e4.z * mono; t4[3] += e4.w * mono;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) part[(g * NMH + al - 4) * 4 + u] = t4[u];
                }
                __syncthreads();
                for (int it = lane; it < NMH * 4; it += kWave) {
                    double sum = 0.0;
                    for (int gg = 0; gg < GH; ++gg) sum += part[gg * NMH * 4 + it];
                    T[(b * 4 + it % 4) * NM + 4 + it / 4] = sum;
                }
                __syncthreads();
            }
        }
        __syncthreads();
        if (wantT) {
            const int TS = S * 4 * NM;
            for (int x = lane; x < TS; x += kWave) A.Tout[(size_t)i * TS + x] = (float)T[x];
        }
        // A_m(p, u) = sum_{|alpha| = m} w_alpha T_b T_c (b != c); (... - D_b) / 2 for b == c
        for (int it = lane; it < NP * 4 * ZP; it += kWave) {
            const int p = it / (4 * ZP), u = (it / ZP) % 4, m = it % ZP;
            int c = 0;
            while ((c + 1) * (c + 2) / 2 <= p) ++c;
            const int b = p - c * (c + 1) / 2;
            const int qa = m * (m + 1) * (m + 2) / 6, qb = (m + 1) * (m + 2) * (m + 3) / 6;
            double sum = 0.0;
            for (int x = qa; x < qb; ++x) sum += mo.w[x] * T[(b * 4 + u) * NM + x] * T[(c * 4 + u) * NM + x];
            if (b == c) sum = 0.5 * (sum - Dg[b * 4 + u]);
            Am[it] = sum;
        }
        __syncthreads();
        // G = sum_m C(zeta, m) 2^-zeta lambda^m A_m: one lane per item
        for (int t = q + lane; t < q1; t += kWave) {
            const PreItem it = A.items[t];
            double g = 0.0, binom = 1.0, lp = 1.0;
            for (int m = 0; m <= it.zint; ++m) {
                g += binom * lp * Am[(it.blk * 4 + it.u) * ZP + m];
                binom = binom * (it.zint - m) / (m + 1);
                lp *= (double)it.lam;
            }
            put(it.k, g * ldexp(1.0, -it.zint));
        }
    } else {
        // angular items outside the moment form: lanes over the first leg, the
        // second leg looped (each unordered pair of the block's species once)
        for (; q < q1; ++q) {
            const PreItem it = A.items[q];
            const AngBlock& ab = M.ang[a * NP + it.blk];
            int c = 0;
            while ((c + 1) * (c + 2) / 2 <= it.blk) ++c;
            const int b = it.blk - c * (c + 1) / 2;
            const double eta = A.ang_eta[ab.par + it.fn];
            const double ieta2 = 1.0 / (eta * eta), zeta = A.ang_zeta[ab.par + it.fn];
            const double lamh = 0.5 * (double)A.ang_lam[ab.par + it.fn];
            double acc = 0.0;
            const int kb0 = seg[c], kb1 = seg[c + 1];
            for (int J = seg[b] + lane; J < seg[b + 1]; J += kWave) {
                const double4 vj = uv[J];
                const double2 xj = rf[J];
                for (int K = (b == c ? J + 1 : kb0); K < kb1; ++K) {
                    const double4 vk = uv[K];
                    const double2 xk = rf[K];
                    const double cs = vj.x * vk.x + vj.y * vk.y + vj.z * vk.z;
                    acc += angular64(ab.name, ieta2, zeta, lamh, xj.x, xk.x, cs, xj.y, xk.y);
                }
            }
            acc = wave_sum64(acc);
            if (lane == 0) put(it.k, acc);
        }
    }
}

static size_t precise_smem(int maxn, int Z, int S, int NP) {
    const int NM = Z > 0 ? (Z + 1) * (Z + 2) * (Z + 3) / 6 : 1;
    size_t b = (size_t)maxn * sizeof(double2) + (size_t)maxn * sizeof(double4);
    if (Z > 0) {
        const int PWS = (3 * Z) | 1, NMH = Z > 1 ? NM - 4 : 1, GH = Z > 1 ? kWave / NMH : 1;
        b += sizeof(double) * ((size_t)maxn * PWS + (size_t)S * 4 * NM + S * 4 + (size_t)NP * 4 * (Z + 1) +
                               (Z > 1 ? (size_t)GH * NMH * 4 : 0));
    }
    return b + sizeof(int) * (kMaxSpecies + 1);
}

hipError_t precise_inputs(ann_ctx* ctx, float* Gc, double* G64, float* Tout) {
    if ((ctx->npre == 0 && !Tout) || ctx->natoms == 0) return hipSuccess;
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
    A.geo = ctx->nl.geo;
    A.Gc = Gc;
    A.G64 = G64;
    const int Z = ctx->hm.moment_z;
    A.Tout = Z > 0 ? Tout : nullptr;
    const int maxn = std::max(ctx->nl.maxn, 1);
    const size_t smem = precise_smem(maxn, Z, ctx->hm.S, ctx->hm.NP);
    Launch L(ctx, "precise");
    switch (Z) {
        case 0: k_precise<0><<<ctx->natoms, 64, smem, ctx->stream>>>(A, maxn); break;
        case 1: k_precise<1><<<ctx->natoms, 64, smem, ctx->stream>>>(A, maxn); break;
        case 2: k_precise<2><<<ctx->natoms, 64, smem, ctx->stream>>>(A, maxn); break;
        case 3: k_precise<3><<<ctx->natoms, 64, smem, ctx->stream>>>(A, maxn); break;
        case 4: k_precise<4><<<ctx->natoms, 64, smem, ctx->stream>>>(A, maxn); break;
        default: k_precise<5><<<ctx->natoms, 64, smem, ctx->stream>>>(A, maxn); break;
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

hipError_t network_rows(ann_ctx* ctx, int form, float* G) {
    if (ctx->npre > 0 && ctx->npre == ctx->nin_total) {
        float* T = (ctx->d_T && ctx->work_atoms >= ctx->natoms) ? ctx->d_T : nullptr;
        const hipError_t e = precise_inputs(ctx, G, nullptr, T);
        ctx->T_gen = (e == hipSuccess && T && ctx->hm.moment_z > 0) ? ctx->gen : ~0ull;
        return e;
    }
    const hipError_t e = symm_fwd(ctx, form, G);
    if (e != hipSuccess) return e;
    return precise_inputs(ctx, G, nullptr);
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
