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
    const int4* items;        // (input k, kind 0 radial / 1 angular, block, function)
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

__device__ inline double wave_sum64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ __launch_bounds__(64) void k_precise(PreciseArgs A) {
    extern __shared__ double4 sv[];                  // [maxn] (x, y, z, r) of kept neighbours
    const int i = blockIdx.x, lane = threadIdx.x;
    if (i >= A.N) return;
    const DevModel& M = *A.dm;
    const int a = A.species[i];
    const int q0 = M.pre_off[a], q1 = M.pre_off[a + 1];
    if (q0 == q1) return;
    const int o = A.off[i], n = A.off[i + 1] - o;
    double* sfc = reinterpret_cast<double*>(sv + n);  // [n] cutoff values
    int* ssp = reinterpret_cast<int*>(sfc + n);      // [n] neighbour species
    const double* R = A.cells + 9 * A.sof[i];
    const uint64_t lt = (1ull << lane) - 1ull;
    int kept = 0;
    for (int q = 0; q < n; q += kWave) {
        const int t = q + lane;
        double x = 0, y = 0, z = 0, d2 = 0;
        bool keep = false;
        if (t < n) {
            int n0, n1, n2;
            unpack_img(A.img[o + t], n0, n1, n2);
            pair_vec(A.pos, i, A.idx[o + t], R, n0, n1, n2, x, y, z, d2);
            keep = d2 > 0.0 && sqrt(d2) <= A.rc;
        }
        const uint64_t mk = __ballot(keep);
        if (keep) {
            const int p = kept + __popcll(mk & lt);
            const double r = sqrt(d2);
            sv[p] = make_double4(x, y, z, r);
            sfc[p] = cut64(M.cutoff, A.rc, r);
            ssp[p] = A.nsp[o + t];
        }
        kept += __popcll(mk);
    }
    __syncthreads();
    const NetDesc& D = M.net[a];
    for (int q = q0; q < q1; ++q) {
        const int4 it = A.items[q];
        double acc = 0.0;
        if (it.y == 0) {          // radial: block = neighbour species
            const RadBlock& rb = M.rad[a * M.S + it.z];
            const double rs = A.rad_rs[rb.par + it.w], eta = A.rad_eta[rb.par + it.w];
            for (int J = lane; J < kept; J += kWave)
                if (ssp[J] == it.z) acc += radial64(rb.name, eta, rs, sv[J].w, sfc[J]);
        } else {                  // angular: block = neighbour pair p(b <= c)
            const AngBlock& ab = M.ang[a * M.NP + it.z];
            int c = 0;
            while ((c + 1) * (c + 2) / 2 <= it.z) ++c;
            const int b = it.z - c * (c + 1) / 2;
            const double eta = A.ang_eta[ab.par + it.w];
            const double ieta2 = 1.0 / (eta * eta), zeta = A.ang_zeta[ab.par + it.w];
            const double lamh = 0.5 * (double)A.ang_lam[ab.par + it.w];
            for (int J = lane; J < kept; J += kWave) {
                const int sj = ssp[J];
                if (sj != b && sj != c) continue;
                const int want = sj == b ? c : b;
                const double4 vj = sv[J];
                for (int K = J + 1; K < kept; ++K) {
                    if (ssp[K] != want) continue;
                    const double4 vk = sv[K];
                    const double cs = (vj.x * vk.x + vj.y * vk.y + vj.z * vk.z) / (vj.w * vk.w);
                    acc += angular64(ab.name, ieta2, zeta, lamh, vj.w, vk.w, cs, sfc[J], sfc[K]);
                }
            }
        }
        acc = wave_sum64(acc);
        if (lane == 0) {
            const size_t at = (size_t)i * A.ldg + it.x;
            if (A.Gc) A.Gc[at] = (float)(acc + A.inpb64[D.inoff + it.x]);
            if (A.G64) A.G64[at] = acc;
        }
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
    const size_t maxn = (size_t)std::max(ctx->nl.maxn, 1);
    const size_t smem = maxn * (sizeof(double4) + sizeof(double) + sizeof(int));
    Launch L(ctx, "precise");
    k_precise<<<ctx->natoms, 64, smem, ctx->stream>>>(A);
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
