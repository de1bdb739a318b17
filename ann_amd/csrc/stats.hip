// Input preconditioning statistics of a training batch (host set-up step of
// the trainer, src/nnp/nnptefr2.cpp:690-787): per species and input k over the
// atoms of that species, the count N and
//   min, max                      (:708-723)
//   harmonic mean  N / sum 1/G    (:724-736)
//   mean           sum G / N      (:737-747)
//   deviation      sqrt(sum (mean - G)^2 / (N - 1))  (:748-762, two passes)
//   rms            sqrt(sum G^2 / N)                  (:763-774)
// G rows are FP32 (the library's G); every sum is FP64.  Blocks own a chunk of
// atoms and one species; a second kernel adds the chunk partials in a fixed
// order, so the result does not depend on scheduling.
#include "common.h"

#include <cfloat>

namespace ann {

constexpr int kStatChunk = 128;     // atoms per block
constexpr int kStatThreads = 256;   // threads per block, over the inputs of a row
enum { ST_MIN, ST_MAX, ST_INV, ST_SUM, ST_SQ, ST_NPART };

// pass 1: min, max, sum 1/G, sum G, sum G^2 and the count of one (chunk, species)
__global__ __launch_bounds__(kStatThreads) void k_stats_part(int N, int ldg, int nchunk, const int* species,
                                                             const float* G, const int* nin, double* part,
                                                             int* cnt) {
    const int c = blockIdx.x, a = blockIdx.y;
    const int i0 = c * kStatChunk, i1 = min(N, i0 + kStatChunk);
    const int n_in = nin[a];
    double* out = part + ((size_t)a * nchunk + c) * ST_NPART * ldg;
    for (int k = threadIdx.x; k < n_in; k += kStatThreads) {
        double mn = DBL_MAX, mx = -DBL_MAX, inv = 0.0, s = 0.0, sq = 0.0;
        for (int i = i0; i < i1; ++i) {
            if (species[i] != a) continue;
            const double g = (double)G[(size_t)i * ldg + k];
            mn = g < mn ? g : mn;
            mx = g > mx ? g : mx;
            inv += 1.0 / g;
            s += g;
            sq += g * g;
        }
        out[ST_MIN * ldg + k] = mn;
        out[ST_MAX * ldg + k] = mx;
        out[ST_INV * ldg + k] = inv;
        out[ST_SUM * ldg + k] = s;
        out[ST_SQ * ldg + k] = sq;
    }
    if (threadIdx.x == 0) {
        int m = 0;
        for (int i = i0; i < i1; ++i) m += species[i] == a ? 1 : 0;
        cnt[a * nchunk + c] = m;
    }
}

// pass 2: sum (mean - G)^2 of one (chunk, species)
__global__ __launch_bounds__(kStatThreads) void k_stats_dev(int N, int ldg, int nchunk, const int* species,
                                                            const float* G, const int* nin, const double* mean,
                                                            double* part) {
    const int c = blockIdx.x, a = blockIdx.y;
    const int i0 = c * kStatChunk, i1 = min(N, i0 + kStatChunk);
    const int n_in = nin[a];
    double* out = part + ((size_t)a * nchunk + c) * ldg;
    for (int k = threadIdx.x; k < n_in; k += kStatThreads) {
        const double m = mean[a * ldg + k];
        double d2 = 0.0;
        for (int i = i0; i < i1; ++i) {
            if (species[i] != a) continue;
            const double d = m - (double)G[(size_t)i * ldg + k];
            d2 += d * d;
        }
        out[k] = d2;
    }
}

// chunk partials -> per (species, input) statistics, chunks added in order.
// res[a][s][k] with s: 0 min, 1 max, 2 harmonic mean, 3 mean, 4 deviation, 5 rms
__global__ void k_stats_fin(int S, int ldg, int nchunk, const int* nin, const double* part, const int* cnt,
                            const double* dpart, double* res, int* count, int pass) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= S * ldg) return;
    const int a = t / ldg, k = t % ldg;
    int n = 0;
    for (int c = 0; c < nchunk; ++c) n += cnt[a * nchunk + c];
    if (k == 0 && pass == 0) count[a] = n;
    if (k >= nin[a]) return;
    double* r = res + (size_t)a * 6 * ldg;
    if (pass == 0) {
        double mn = DBL_MAX, mx = -DBL_MAX, inv = 0.0, s = 0.0, sq = 0.0;
        for (int c = 0; c < nchunk; ++c) {
            const double* p = part + ((size_t)a * nchunk + c) * ST_NPART * ldg;
            mn = fmin(mn, p[ST_MIN * ldg + k]);
            mx = fmax(mx, p[ST_MAX * ldg + k]);
            inv += p[ST_INV * ldg + k];
            s += p[ST_SUM * ldg + k];
            sq += p[ST_SQ * ldg + k];
        }
        r[0 * ldg + k] = mn;
        r[1 * ldg + k] = mx;
        r[2 * ldg + k] = (double)n / inv;
        r[3 * ldg + k] = s / (double)n;
        r[5 * ldg + k] = sqrt(sq / (double)n);
    } else {
        double d2 = 0.0;
        for (int c = 0; c < nchunk; ++c) d2 += dpart[((size_t)a * nchunk + c) * ldg + k];
        r[4 * ldg + k] = sqrt(d2 / ((double)n - 1.0));
    }
}

}  // namespace ann

using namespace ann;

// Statistics of the inputs d_G (N x ldg FP32 rows of the current batch, as
// written by ann_symm_fwd).  stats: S x 6 x max_ninput doubles on the host,
// [species][min, max, harmonic mean, mean, deviation, rms][input]; entries past
// a species' input count are zero.  count: S atom counts.  A species with no
// atoms gets the reference's values for N = 0 (+-DBL_MAX extremes, NaN means).
extern "C" int ann_input_stats(ann_ctx* ctx, const float* d_G, int* count, double* stats, int n) {
    if (!ctx) return fail(ANN_ERR_ARG, "null context");
    const int S = ctx->S, ldg = ctx->ldg, N = ctx->natoms, mx = ctx->max_nin;
    if (!d_G || !count || !stats || n != S * 6 * mx) return fail(ANN_ERR_ARG, "input statistics: bad arguments");
    for (int a = 0; a < S; ++a)
        if (ctx->hm.nin[a] < 0 || ctx->hm.nin[a] > ldg || ctx->hm.nin[a] > mx)
            return fail(ANN_ERR_STATE, "input statistics: input count exceeds the G row");
    if (N <= 0 || !ctx->d_species) return fail(ANN_ERR_STATE, "input statistics: no batch");
    const int nchunk = (N + kStatChunk - 1) / kStatChunk;
    const size_t npart = (size_t)S * nchunk * ST_NPART * ldg, ndev = (size_t)S * nchunk * ldg;
    const size_t nres = (size_t)S * 6 * ldg;
    double *part = nullptr, *dpart = nullptr, *res = nullptr;
    int *cnt = nullptr, *dcount = nullptr, *dnin = nullptr;
    hipError_t e = hipMalloc((void**)&part, sizeof(double) * npart);
    if (e == hipSuccess) e = hipMalloc((void**)&dpart, sizeof(double) * ndev);
    if (e == hipSuccess) e = hipMalloc((void**)&res, sizeof(double) * nres);
    if (e == hipSuccess) e = hipMalloc((void**)&cnt, sizeof(int) * S * nchunk);
    if (e == hipSuccess) e = hipMalloc((void**)&dcount, sizeof(int) * S);
    if (e == hipSuccess) e = hipMalloc((void**)&dnin, sizeof(int) * S);
    if (e == hipSuccess) e = hipMemcpyAsync(dnin, ctx->hm.nin, sizeof(int) * S, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(res, 0, sizeof(double) * nres, ctx->stream);
    const dim3 grid(nchunk, S);
    const int fb = (S * ldg + 255) / 256;
    if (e == hipSuccess) {
        k_stats_part<<<grid, kStatThreads, 0, ctx->stream>>>(N, ldg, nchunk, ctx->d_species, d_G, dnin, part, cnt);
        k_stats_fin<<<fb, 256, 0, ctx->stream>>>(S, ldg, nchunk, dnin, part, cnt, dpart, res, dcount, 0);
        // means res[a][3] feed the deviation pass (row stride 6 * ldg per species)
        double* mean = nullptr;
        e = hipMalloc((void**)&mean, sizeof(double) * S * ldg);
        if (e == hipSuccess) {
            for (int a = 0; a < S && e == hipSuccess; ++a)
                e = hipMemcpyAsync(mean + (size_t)a * ldg, res + (size_t)a * 6 * ldg + 3 * ldg, sizeof(double) * ldg,
                                   hipMemcpyDeviceToDevice, ctx->stream);
            if (e == hipSuccess) {
                k_stats_dev<<<grid, kStatThreads, 0, ctx->stream>>>(N, ldg, nchunk, ctx->d_species, d_G, dnin, mean,
                                                                    dpart);
                k_stats_fin<<<fb, 256, 0, ctx->stream>>>(S, ldg, nchunk, dnin, part, cnt, dpart, res, dcount, 1);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            hipFree(mean);
        }
    }
    std::vector<double> h(nres);
    std::vector<int> hc(S);
    if (e == hipSuccess) e = hipMemcpy(h.data(), res, sizeof(double) * nres, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hc.data(), dcount, sizeof(int) * S, hipMemcpyDeviceToHost);
    hipFree(part); hipFree(dpart); hipFree(res); hipFree(cnt); hipFree(dcount); hipFree(dnin);
    if (e != hipSuccess) return fail(ANN_ERR_HIP, hipGetErrorString(e));
    for (int a = 0; a < S; ++a) {
        count[a] = hc[a];
        for (int s = 0; s < 6; ++s)
            for (int k = 0; k < mx; ++k)
                stats[((size_t)a * 6 + s) * mx + k] = k < ctx->hm.nin[a] ? h[((size_t)a * 6 + s) * ldg + k] : 0.0;
    }
    return ANN_OK;
}
