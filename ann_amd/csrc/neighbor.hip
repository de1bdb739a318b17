// neighbor.hip -- periodic cell-list neighbour build with image vectors.
//
// Replaces NeighborList::build / verlet::List::build (src/struc absent; the
// contract comes from the call sites src/nnp/nnpce.cpp:282,
// src/nnp/nnp.cpp:343-348, 389-395, 478-482 and README.md:39): for every atom
// i, every periodic image (j, n) with 0 < |r_i - r_j - R n| < rc, in the
// canonical order j ascending then n lexicographic, plus the reverse-slot map
// (i -> j, n) <-> (j -> i, -n) that the atomic-free force reduction needs.
//
// Layout: positions FP64 [N][3]; per structure R^-1 and a bin grid in
// fractional coordinates (nb_a = floor(w_a / rc) bins along lattice direction
// a, w_a the perpendicular width), so that the search box is +-k_a bins with
// k_a = ceil(rc nb_a / w_a); a wavefront per centre atom walks the flattened
// candidate set of its search box (lanes over candidates, not over bins),
// applies the exact FP64 inclusion test in the same operation order as the
// oracle, compacts accepted keys (j << 30 | packed n) through a ballot into
// LDS and sorts them (ranks in the staged search, a bitonic network in the
// wide one).  Integer work; the roofline is
// HBM/latency (positions in, 16 B + 13 B per pair out).
#include <atomic>
#include <chrono>

#include "common.h"
#include "pairvec_dev.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <tuple>

namespace ann {

constexpr int kMaxImages = 1024;   // bin images searched per atom

// per structure grid: nb[3] bins, search half-width k[3] in bins, bin count
// and offset; the staged search groups sb^3 bins into one block (super-bins:
// nsb[3], offset sboff)
struct CellInfo {
    int nb[3], k[3], nbins, boff;
    int sb, nsb[3], sboff;
};

__global__ void k_struc_of(const int* __restrict__ off, int nstruc, int* __restrict__ sof) {
    const int s = blockIdx.x;
    if (s >= nstruc) return;
    for (int i = off[s] + threadIdx.x; i < off[s + 1]; i += blockDim.x) sof[i] = s;
}

// bin of atom i from its fractional coordinates; writes the atom's integer
// shift (the lattice vector that wraps it into the cell)
// A position whose fractional coordinate is not finite (or beyond 1e8 cells)
// is binned as if at the origin and reported through flag bit 16: the build
// stays in bounds and ann_neighbors returns an argument error.
__device__ __forceinline__ int bin_atom(int i, const double* __restrict__ pos, const int* __restrict__ sof,
                                        const CellInfo* __restrict__ ci, const double* __restrict__ rinv,
                                        int* __restrict__ shift, int* __restrict__ flag) {
    const int s = sof[i];
    const CellInfo c = ci[s];
    const double* Ri = rinv + 9 * s;
    int b[3];
    for (int a = 0; a < 3; ++a) {
        double f = Ri[3 * a] * pos[3 * i] + Ri[3 * a + 1] * pos[3 * i + 1] + Ri[3 * a + 2] * pos[3 * i + 2];
        if (!(fabs(f) < 1e8)) {
            f = 0.0;
            if (flag) atomicOr(flag, 16);
        }
        const double fl = floor(f);
        shift[3 * i + a] = (int)fl;
        int bb = (int)floor((f - fl) * c.nb[a]);
        b[a] = bb < 0 ? 0 : (bb >= c.nb[a] ? c.nb[a] - 1 : bb);
    }
    return c.boff + (b[0] * c.nb[1] + b[1]) * c.nb[2] + b[2];
}

__device__ __forceinline__ double4 bin_record(int i, const double* __restrict__ pos, const int* __restrict__ shift) {
    const uint64_t js = ((uint64_t)(uint32_t)i << 32) | pack_img(shift[3 * i], shift[3 * i + 1], shift[3 * i + 2]);
    return make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], __longlong_as_double((long long)js));
}

// Binning.  The bin counts are zero on entry: k_fill_bins takes every count
// back to zero.
__global__ void k_bin(int N, const double* __restrict__ pos, const int* __restrict__ sof,
                      const CellInfo* __restrict__ ci, const double* __restrict__ rinv,
                      int* __restrict__ bin_of, int* __restrict__ shift, int* __restrict__ bin_cnt,
                      int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int g = bin_atom(i, pos, sof, ci, rinv, shift, flag);
    bin_of[i] = g;
    atomicAdd(&bin_cnt[g], 1);
}

// Atoms in bin order: index, and a 32-byte record (x, y, z, atom | shift)
// that the staged search copies with one load per candidate.  The slot is
// taken by counting the bin's count down, so every count is zero again when
// the kernel ends.
__global__ void k_fill_bins(int N, const int* __restrict__ bin_of, const int* __restrict__ bin_start,
                            int* __restrict__ bin_cnt, int* __restrict__ bin_atoms,
                            const double* __restrict__ pos, const int* __restrict__ shift,
                            double4* __restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int g = bin_of[i];
    const int p = atomicSub(&bin_cnt[g], 1) - 1;
    const int slot = bin_start[g] + p;
    bin_atoms[slot] = i;
    rec[slot] = bin_record(i, pos, shift);
}

__device__ __forceinline__ int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// one wavefront (block of 64) per centre atom
__global__ __launch_bounds__(64) void k_search(
    int N, const double* __restrict__ pos, const int* __restrict__ sof, const CellInfo* __restrict__ ci,
    const double* __restrict__ cells, const int* __restrict__ bin_of, const int* __restrict__ shift,
    const int* __restrict__ bin_start, const int* __restrict__ bin_atoms, double rc2,
    uint64_t* __restrict__ keys_out, int* __restrict__ cnt_out, int* __restrict__ flag,
    const int* __restrict__ bin_only) {
    __shared__ int im_bin[kMaxImages];
    __shared__ int im_m[kMaxImages];     // packed m
    __shared__ int im_pre[kMaxImages + 1];
    __shared__ uint64_t keys[kMaxNeigh];
    __shared__ int s_total, s_cnt;
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    if (i >= N) return;
    if (bin_only && !bin_only[bin_of[i]]) return;     // handled by the staged search
    const int s = sof[i];
    const CellInfo c = ci[s];
    const double* R = cells + 9 * s;
    const int gi = bin_of[i] - c.boff;
    const int b0 = gi / (c.nb[1] * c.nb[2]), b1 = (gi / c.nb[2]) % c.nb[1], b2 = gi % c.nb[2];
    const int e0 = 2 * c.k[0] + 1, e1 = 2 * c.k[1] + 1, e2 = 2 * c.k[2] + 1;
    const int nimg = e0 * e1 * e2;
    if (nimg > kMaxImages) {
        if (lane == 0) { atomicOr(flag, 2); cnt_out[i] = 0; }
        return;
    }
    // bin images of the search box: (wrapped bin, image shift) and candidate counts
    for (int t = lane; t < nimg; t += kWave) {
        const int u0 = b0 - c.k[0] + t / (e1 * e2);
        const int u1 = b1 - c.k[1] + (t / e2) % e1;
        const int u2 = b2 - c.k[2] + t % e2;
        const int m0 = floordiv(u0, c.nb[0]), m1 = floordiv(u1, c.nb[1]), m2 = floordiv(u2, c.nb[2]);
        const int g = c.boff + ((u0 - m0 * c.nb[0]) * c.nb[1] + (u1 - m1 * c.nb[1])) * c.nb[2] + (u2 - m2 * c.nb[2]);
        im_bin[t] = g;
        im_m[t] = (int)pack_img(m0, m1, m2);
        im_pre[t + 1] = bin_start[g + 1] - bin_start[g];
    }
    __syncthreads();
    if (lane == 0) {
        im_pre[0] = 0;
        for (int t = 0; t < nimg; ++t) im_pre[t + 1] += im_pre[t];
        s_total = im_pre[nimg];
        s_cnt = 0;
    }
    __syncthreads();
    const int total = s_total;
    const int si0 = shift[3 * i], si1 = shift[3 * i + 1], si2 = shift[3 * i + 2];
    int count = 0;  // uniform running count
    for (int base = 0; base < total; base += kWave) {
        const int cidx = base + lane;
        bool ok = false;
        uint64_t key = 0;
        if (cidx < total) {
            int lo = 0, hi = nimg;  // largest t with im_pre[t] <= cidx
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (im_pre[mid] <= cidx) lo = mid; else hi = mid;
            }
            const int j = bin_atoms[bin_start[im_bin[lo]] + (cidx - im_pre[lo])];
            int m0, m1, m2;
            unpack_img((uint32_t)im_m[lo], m0, m1, m2);
            const int n0 = si0 - shift[3 * j] + m0;
            const int n1 = si1 - shift[3 * j + 1] + m1;
            const int n2 = si2 - shift[3 * j + 2] + m2;
            double x, y, z, d2;
            pair_vec(pos, i, j, R, n0, n1, n2, x, y, z, d2);
            ok = d2 > 0.0 && d2 < rc2 && !(j == i && n0 == 0 && n1 == 0 && n2 == 0);
            key = ((uint64_t)j << 30) | (uint64_t)pack_img(n0, n1, n2);
        }
        const uint64_t mask = __ballot(ok);
        const int before = __popcll(mask & ((1ull << lane) - 1ull));
        if (ok && count + before < kMaxNeigh) keys[count + before] = key;
        count += __popcll(mask);
    }
    const int n = count < kMaxNeigh ? count : kMaxNeigh;
    if (count > kMaxNeigh && lane == 0) atomicOr(flag, 1);
    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (int t = n + lane; t < np2; t += kWave) keys[t] = ~0ull;
    __syncthreads();
    for (int k = 2; k <= np2; k <<= 1) {
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int t = lane; t < np2; t += kWave) {
                const int p = t ^ jj;
                if (p > t) {
                    const uint64_t a = keys[t], b = keys[p];
                    const bool up = (t & k) == 0;
                    if ((a > b) == up) { keys[t] = b; keys[p] = a; }
                }
            }
            __syncthreads();
        }
    }
    uint64_t* dst = keys_out + (size_t)i * kMaxNeigh;
    for (int t = lane; t < n; t += kWave) dst[t] = keys[t];
    if (lane == 0) { cnt_out[i] = n; atomicMax(flag + 1, count); }
}

// Staged search, one block per bin: the candidates of the bin's search box
// (every atom of the (2k+1)^3 bin images with its image offset) are copied
// to LDS with one 32-byte load each; each wavefront then tests one centre
// atom of the bin against them, compacts the accepted keys (j << 30 | packed
// image) into its own LDS buffer and stores them sorted by rank (no block
// barriers after the staging).  Bins
// whose search box holds more than kStageCap candidates or kStageImages
// images are flagged and left to k_search.  Each block writes the overflow
// mark (0 or 1) of each of its centre bins exactly once per build, so the
// marks need no clearing between builds.
constexpr int kStageCap = 1280;
constexpr int kStageImages = 216;
constexpr int kStageThreads = 512;
constexpr int kWaveKeys = 256;          // accepted keys per centre in the staged search (more: wide path)
static_assert(kWaveKeys == 4 * kWave, "the staged sort covers up to four keys per lane");
constexpr int kStageWaves = kStageThreads / kWave;
constexpr int kStageIt = (kStageCap + kStageThreads - 1) / kStageThreads;

// Sorted store of n distinct keys by rank: every lane counts, for the up to R
// keys it holds, the keys below them (broadcast LDS reads, no dependency
// between iterations), and writes each key at its rank.
template <int R>
__device__ __forceinline__ void rank_and_store(const uint64_t* buf, int n, uint64_t* dst) {
    const int lane = threadIdx.x % kWave;
    uint64_t v[R];
    int rk[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = lane + kWave * r;
        v[r] = e < n ? buf[e] : ~0ull;
        rk[r] = 0;
    }
#pragma unroll 4
    for (int m = 0; m < n; ++m) {
        const uint64_t w = buf[m];
#pragma unroll
        for (int r = 0; r < R; ++r) rk[r] += w < v[r] ? 1 : 0;
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (lane + kWave * r < n) dst[rk[r]] = v[r];
}

__global__ __launch_bounds__(kStageThreads) void k_search_bins(
    const CellInfo* __restrict__ ci, const double* __restrict__ cells, const int* __restrict__ bin_struc,
    const int* __restrict__ bin_start, const double4* __restrict__ rec, const double* __restrict__ pos, double rc2,
    uint64_t* __restrict__ keys_out, int* __restrict__ cnt_out, int* __restrict__ flag, int* __restrict__ bin_ovf) {
    __shared__ int im_first[kStageImages];
    __shared__ int im_m[kStageImages];
    __shared__ int im_pre[kStageImages + 1];
    __shared__ float4 cf[kStageCap];                        // FP32 image position x_j + R (m - shift_j)
    __shared__ uint64_t cjd[kStageCap];                     // j << 32 | packed (image - shift_j)
    __shared__ uint64_t wkeys[kStageWaves][kWaveKeys];
    __shared__ int cen_bin[8], cen_pre[9], cen_img[8];
    __shared__ int s_ovf[8];                                // centre bins left to the wide search
    __shared__ int s_total, s_max, s_xmax;
    const int g = xcd_block(blockIdx.x, gridDim.x), t = threadIdx.x;
    const int lane = t % kWave, wave = t / kWave;
    const int s = bin_struc[g];                             // structure of this super-bin
    const CellInfo c = ci[s];
    const double* R = cells + 9 * s;
    const int gi = g - c.sboff;
    const int B0 = gi / (c.nsb[1] * c.nsb[2]), B1 = (gi / c.nsb[2]) % c.nsb[1], B2 = gi % c.nsb[2];
    const int lo0 = c.sb * B0, lo1 = c.sb * B1, lo2 = c.sb * B2;
    const int n0b = min(c.sb, c.nb[0] - lo0), n1b = min(c.sb, c.nb[1] - lo1), n2b = min(c.sb, c.nb[2] - lo2);
    // centre bins of the super-bin and their atom ranges
    if (t == 0) {
        int acc = 0, q = 0;
        for (int d0 = 0; d0 < n0b; ++d0)
            for (int d1 = 0; d1 < n1b; ++d1)
                for (int d2 = 0; d2 < n2b; ++d2, ++q) {
                    const int bg = c.boff + ((lo0 + d0) * c.nb[1] + (lo1 + d1)) * c.nb[2] + (lo2 + d2);
                    cen_bin[q] = bg;
                    cen_pre[q] = acc;
                    cen_img[q] = ((c.k[0] + d0) * (n1b + 2 * c.k[1]) + (c.k[1] + d1)) * (n2b + 2 * c.k[2]) + (c.k[2] + d2);
                    acc += bin_start[bg + 1] - bin_start[bg];
                }
        cen_pre[q] = acc;
        s_max = 0;
        s_xmax = 0;
    }
    if (t < 8) s_ovf[t] = 0;
    const int ncb = n0b * n1b * n2b;
    const int e0 = n0b + 2 * c.k[0], e1 = n1b + 2 * c.k[1], e2 = n2b + 2 * c.k[2];
    const int nimg = e0 * e1 * e2;
    if (nimg > kStageImages) {
        __syncthreads();
        if (t < ncb) bin_ovf[cen_bin[t]] = 1;
        if (t == 0) atomicOr(flag, 8);
        return;
    }
    for (int q = t; q < nimg; q += kStageThreads) {
        const int u0 = lo0 - c.k[0] + q / (e1 * e2);
        const int u1 = lo1 - c.k[1] + (q / e2) % e1;
        const int u2 = lo2 - c.k[2] + q % e2;
        const int m0 = floordiv(u0, c.nb[0]), m1 = floordiv(u1, c.nb[1]), m2 = floordiv(u2, c.nb[2]);
        const int gg = c.boff + ((u0 - m0 * c.nb[0]) * c.nb[1] + (u1 - m1 * c.nb[1])) * c.nb[2] + (u2 - m2 * c.nb[2]);
        const int f = bin_start[gg];
        im_first[q] = f;
        im_m[q] = (int)pack_img(m0, m1, m2);
        im_pre[q + 1] = bin_start[gg + 1] - f;
    }
    __syncthreads();
    if (t < kWave) {   // inclusive scan of the image counts by the first wavefront
        int carry = 0;
        for (int q0 = 0; q0 < nimg; q0 += kWave) {
            const int q = q0 + lane;
            int v = q < nimg ? im_pre[q + 1] : 0;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int u = __shfl_up(v, o);
                if (lane >= o) v += u;
            }
            if (q < nimg) im_pre[q + 1] = v + carry;
            carry += __shfl(v, kWave - 1);
        }
        if (lane == 0) { im_pre[0] = 0; s_total = carry; }
    }
    __syncthreads();
    const int total = s_total;
    const int ncen = cen_pre[ncb];
    if (ncen == 0) {                                        // no centres in this super-bin
        if (t < ncb) bin_ovf[cen_bin[t]] = 0;
        return;
    }
    if (total > kStageCap) {
        if (t < ncb) bin_ovf[cen_bin[t]] = 1;
        if (t == 0) atomicOr(flag, 8);
        return;
    }
    {
        // every thread resolves all its candidates first, then issues their loads together
        int src[kStageIt], mq[kStageIt];
#pragma unroll
        for (int r = 0; r < kStageIt; ++r) {
            const int q = t + r * kStageThreads;
            src[r] = -1;
            if (q < total) {
                int lo = 0, hi = nimg;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (im_pre[mid] <= q) lo = mid; else hi = mid;
                }
                src[r] = im_first[lo] + (q - im_pre[lo]);
                mq[r] = im_m[lo];
            }
        }
        double4 r4[kStageIt];
#pragma unroll
        for (int r = 0; r < kStageIt; ++r)
            if (src[r] >= 0) r4[r] = rec[src[r]];
        float xm = 0.0f;
#pragma unroll
        for (int r = 0; r < kStageIt; ++r) {
            const int q = t + r * kStageThreads;
            if (src[r] < 0) continue;
            const uint64_t js = (uint64_t)__double_as_longlong(r4[r].w);
            int m0, m1, m2, s0, s1, s2;
            unpack_img((uint32_t)mq[r], m0, m1, m2);
            unpack_img((uint32_t)(js & 0xffffffffu), s0, s1, s2);
            const double d0 = m0 - s0, d1 = m1 - s1, d2 = m2 - s2;
            const float x = (float)(r4[r].x + (R[0] * d0 + R[1] * d1 + R[2] * d2));
            const float y = (float)(r4[r].y + (R[3] * d0 + R[4] * d1 + R[5] * d2));
            const float z = (float)(r4[r].z + (R[6] * d0 + R[7] * d1 + R[8] * d2));
            cf[q] = make_float4(x, y, z, 0.0f);
            xm = fmaxf(xm, fmaxf(fabsf(x), fmaxf(fabsf(y), fabsf(z))));
            cjd[q] = (js & 0xffffffff00000000ull) | (uint64_t)pack_img(m0 - s0, m1 - s1, m2 - s2);
        }
        atomicMax(&s_xmax, __float_as_int(xm));             // non-negative floats order as integers
    }
    __syncthreads();
    // FP32 pre-test.  With u = 2^-24 and every FP32 coordinate within
    // X = max |x| of its FP64 value to u X, the FP32 squared distance of a
    // pair within 1.1 rc is within  e = 2 sqrt3 rc D + 3 D^2 + 4 u rc^2,
    // D = 2 u X + 1.2 u rc, of the FP64 value.  Four times e is the band:
    // below rc^2 - band a pair is in, above rc^2 + band out, and the exact
    // FP64 test of the oracle decides inside the band and near zero.
    const double rc = sqrt(rc2), uu = 5.9604644775390625e-08;
    const double D = 2.0 * uu * (double)__int_as_float(s_xmax) + 1.2 * uu * rc;
    const float band = (float)(4.0 * (3.4641016151377544 * rc * D + 3.0 * D * D + 4.0 * uu * rc2) + 2.0 * uu * rc2);
    const float rc2f = (float)rc2, lo_thr = rc2f - band, hi_thr = rc2f + band;
    uint64_t* wk = wkeys[wave];
    int wmax = 0;          // largest neighbour count of this wavefront's centres
    for (int cq = wave; cq < ncen; cq += kStageWaves) {
        int cb = 0;
        while (cb + 1 < ncb && cq >= cen_pre[cb + 1]) ++cb;
        // the centre is staged too: its own bin's zero image, where (image - shift) = -shift
        const int qc = im_pre[cen_img[cb]] + (cq - cen_pre[cb]);
        const float4 xi = cf[qc];
        const uint64_t jsi = cjd[qc];
        const int i = (int)(jsi >> 32);
        int si0, si1, si2;
        unpack_img((uint32_t)(jsi & 0xffffffffu), si0, si1, si2);
        si0 = -si0;
        si1 = -si1;
        si2 = -si2;
        int count = 0;
        for (int base = 0; base < total; base += kWave) {
            const int q = base + lane;
            bool ok = false, unsure = false;
            if (q < total && q != qc) {
                const float4 xq = cf[q];
                const float dx = xi.x - xq.x, dy = xi.y - xq.y, dz = xi.z - xq.z;
                const float d2f = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                ok = d2f < lo_thr && d2f > band;
                unsure = d2f <= band || (d2f >= lo_thr && d2f <= hi_thr);
            }
            const uint64_t jd = q < total ? cjd[q] : 0ull;
            const int j = (int)(jd >> 32);
            int d0, d1, d2i;
            unpack_img((uint32_t)(jd & 0xffffffffu), d0, d1, d2i);
            const int n0 = si0 + d0, n1 = si1 + d1, n2 = si2 + d2i;
            if (__ballot(unsure)) {
                if (unsure) {
                    double x, y, z, d2;
                    pair_vec(pos, i, j, R, n0, n1, n2, x, y, z, d2);
                    ok = d2 > 0.0 && d2 < rc2;
                }
            }
            const uint64_t key = ((uint64_t)j << 30) | (uint64_t)pack_img(n0, n1, n2);
            const uint64_t mask = __ballot(ok);
            const int before = __popcll(mask & ((1ull << lane) - 1ull));
            if (ok && count + before < kWaveKeys) wk[count + before] = key;
            count += __popcll(mask);
        }
        if (count > kWaveKeys) {
            // more neighbours than the staged buffer holds: the bin goes to the wide search
            if (lane == 0) {
                s_ovf[cb] = 1;
                atomicOr(flag, 8);
            }
            continue;
        }
        const int n = count;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint64_t* dst = keys_out + (size_t)i * kMaxNeigh;
        if (n <= kWave) rank_and_store<1>(wk, n, dst);
        else if (n <= 2 * kWave) rank_and_store<2>(wk, n, dst);
        else rank_and_store<4>(wk, n, dst);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane == 0) cnt_out[i] = n;
        wmax = max(wmax, count);
    }
    // one global atomic per block (a contended atomic per atom serialises the kernel)
    if (lane == 0) atomicMax(&s_max, wmax);
    __syncthreads();
    if (t == 0) atomicMax(flag + 1, s_max);
    if (t < ncb) bin_ovf[cen_bin[t]] = s_ovf[t];
}

// CSR fill: index, packed image, neighbour species and FP32 pair geometry
__global__ void k_compact(int N, const double* __restrict__ pos, const int* __restrict__ sof,
                          const double* __restrict__ cells, const int* __restrict__ species,
                          const uint64_t* __restrict__ keys, const int* __restrict__ off,
                          int* __restrict__ idx, uint32_t* __restrict__ img, int* __restrict__ rev,
                          float4* __restrict__ geo, uint8_t* __restrict__ nsp, const int* __restrict__ order,
                          int cap) {
    // centres in bin order: the blocks of one XCD cover one region of space,
    // so the partners they read stay in that XCD's L2
    const int b = xcd_block(blockIdx.x, gridDim.x);
    if (b >= N || off[N] > cap) return;     // a list that does not fit is compacted again after growing
    const int i = order[b];
    const int o = off[i], n = off[i + 1] - o;
    const double* R = cells + 9 * sof[i];
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const uint64_t key = keys[(size_t)i * kMaxNeigh + t];
        const int j = (int)(key >> 30);
        const uint32_t p = (uint32_t)(key & 0x3fffffffu);
        int n0, n1, n2;
        unpack_img(p, n0, n1, n2);
        double x, y, z, d2;
        pair_vec(pos, i, j, R, n0, n1, n2, x, y, z, d2);
        idx[o + t] = j;
        img[o + t] = p;
        rev[o + t] = o + t;
        geo[o + t] = make_float4((float)x, (float)y, (float)z, (float)sqrt(d2));
        nsp[o + t] = (uint8_t)species[j];
    }
}

// Reverse slots.  rev is an involution, so each pair is resolved once, from its
// upper slot (partner index above the centre's, or the same atom with the
// larger packed image): a binary search in the partner's CSR run, which is
// sorted by (partner index, packed image) like the keys, on the 4-byte partner
// indices, then a scan over the few images of the same partner; the result
// is written to both slots.  k_compact has set every slot to itself before,
// so a slot without a partner (an asymmetric list, impossible for the
// symmetric FP64 test) stays in range; the missing upper match is flagged.
__global__ void k_rev(int N, const int* __restrict__ idx, const uint32_t* __restrict__ img,
                      const int* __restrict__ off, int* __restrict__ rev, int* __restrict__ flag,
                      const int* __restrict__ order, int cap) {
    // one wavefront per centre, the centres of a workgroup consecutive in bin
    // order (see k_compact): their partner rows overlap, so the searches of a
    // workgroup share the lines they bring into its CU's cache
    const int b = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (b >= N || off[N] > cap) return;
    const int i = order[b];
    const int o = off[i], n = off[i + 1] - o;
    for (int t = threadIdx.x % kWave; t < n; t += kWave) {
        const int j = idx[o + t];
        const uint32_t p = img[o + t];
        int n0, n1, n2;
        unpack_img(p, n0, n1, n2);
        const uint32_t want = pack_img(-n0, -n1, -n2);
        if (j < i || (j == i && p < want)) continue;       // lower slot: written by its partner
        const int b0 = off[j], e = off[j + 1];
        int lo = b0, hi = e;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (idx[mid] < i) lo = mid + 1; else hi = mid;
        }
        while (lo < e && idx[lo] == i && img[lo] < want) ++lo;
        if (lo < e && idx[lo] == i && img[lo] == want) {
            rev[o + t] = lo;
            rev[lo] = o + t;
        } else {
            atomicOr(flag, 4);
        }
    }
}

static void inv3(const double* R, double* Ri) {
    const double a = R[0], b = R[1], c = R[2], d = R[3], e = R[4], f = R[5], g = R[6], h = R[7], i = R[8];
    const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const double id = 1.0 / (a * A + b * B + c * C);
    Ri[0] = A * id; Ri[1] = -(b * i - c * h) * id; Ri[2] = (b * f - c * e) * id;
    Ri[3] = B * id; Ri[4] = (a * i - c * g) * id;  Ri[5] = -(a * f - c * d) * id;
    Ri[6] = C * id; Ri[7] = -(a * h - b * g) * id; Ri[8] = (a * e - b * d) * id;
}

template <typename T>
static hipError_t grow(T*& p, int64_t& cap, int64_t want) {
    if (want <= cap && p) return hipSuccess;
    if (p) { hipError_t e = hipFree(p); if (e != hipSuccess) return e; p = nullptr; }
    int64_t c = std::max<int64_t>(want, 1);
    c += c / 4;
    hipError_t e = hipMalloc((void**)&p, sizeof(T) * c);
    if (e == hipSuccess) cap = c;
    return e;
}

// Small exclusive scans in one workgroup: tiles of kScanThreads x 4 values,
// a wavefront scan of the per-thread sums, the wavefront totals scanned by the
// first wavefront, a carry across tiles.  One launch instead of the two of a
// decoupled device scan, and the host-visible words written by the same kernel.
constexpr int kScanThreads = 1024;
constexpr int kScanPre = 4;             // tiles of 4 elements per thread loaded up front
constexpr int kScanMax = 1 << 16;

__device__ __forceinline__ void store_host(int* rb, int q, int v) {
    __hip_atomic_store(rb + q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Exclusive scan of n values by one workgroup of kScanThreads threads; returns
// the sum of all n values to every thread.
__device__ int scan_block(const int* __restrict__ in, int* __restrict__ out, int n) {
    __shared__ int wsum[kScanThreads / kWave];
    __shared__ int s_carry;
    const int t = threadIdx.x, lane = t % kWave, w = t / kWave;
    if (t == 0) s_carry = 0;
    __syncthreads();
    // the first kScanPre tiles are loaded before any is scanned (one memory
    // latency for up to 16384 elements); later tiles load in their own trip
    int pre[kScanPre][4];
#pragma unroll
    for (int p = 0; p < kScanPre; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = p * 4 * kScanThreads + 4 * t + q;
            pre[p][q] = e < n ? in[e] : 0;
        }
    for (int base = 0, tile = 0; base < n; base += 4 * kScanThreads, ++tile) {
        int v[4], sum = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = base + 4 * t + q;
            int x = 0;
#pragma unroll
            for (int p = 0; p < kScanPre; ++p)
                if (p == tile) x = pre[p][q];
            v[q] = tile < kScanPre ? x : (e < n ? in[e] : 0);
            sum += v[q];
        }
        int inc = sum;   // inclusive scan over the wavefront
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int u = __shfl_up(inc, o);
            if (lane >= o) inc += u;
        }
        if (lane == kWave - 1) wsum[w] = inc;
        __syncthreads();
        if (w == 0) {
            int x = lane < kScanThreads / kWave ? wsum[lane] : 0;
#pragma unroll
            for (int o = 1; o < kScanThreads / kWave; o <<= 1) {
                const int u = __shfl_up(x, o);
                if (lane >= o) x += u;
            }
            if (lane < kScanThreads / kWave) wsum[lane] = x;   // inclusive over wavefronts
        }
        __syncthreads();
        int run = s_carry + (w > 0 ? wsum[w - 1] : 0) + inc - sum;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = base + 4 * t + q;
            if (e < n) out[e] = run;
            run += v[q];
        }
        __syncthreads();
        if (t == 0) s_carry += wsum[kScanThreads / kWave - 1];
        __syncthreads();
    }
    return s_carry;
}

// One-workgroup scan; with rb, the flags, the last element's exclusive prefix
// and then seq go to the pinned host words, and the first nclear flags are
// cleared once read (the next build starts from zero without a clear launch).
__global__ __launch_bounds__(kScanThreads) void k_scan_small(const int* __restrict__ in, int* __restrict__ out, int n,
                                                             int* __restrict__ flags, int nflags, int nclear, int* rb,
                                                             int seq) {
    const int tot = scan_block(in, out, n);
    if (rb && threadIdx.x == kScanThreads - 1) {
        for (int q = 0; q < nflags; ++q) store_host(rb, q, flags[q]);
        for (int q = 0; q < nclear; ++q) flags[q] = 0;
        store_host(rb, nflags, n > 0 ? tot - in[n - 1] : 0);
        __threadfence_system();
        store_host(rb, nflags + 1, seq);            // the words above are complete
    }
}

__global__ void k_readback(int* src, int n, int* rb, int seq, int nclear) {
    const int t = threadIdx.x;
    if (t < n) store_host(rb, t, src[t]);
    if (t < nclear) src[t] = 0;                     // cleared once read
    __threadfence_system();
    __syncthreads();
    if (seq && t == 0) {
        store_host(rb, n, seq);
        __threadfence_system();
    }
}

hipError_t readback(const int* src, int n, int* rb, hipStream_t st) {
    k_readback<<<1, 64, 0, st>>>(const_cast<int*>(src), n, rb, 0, 0);
    return hipGetLastError();
}

hipError_t exclusive_scan(const int* in, int* out, int n, hipStream_t st, int* flags, int nflags, int* rb, int seq,
                          int nclear) {
    if (n <= kScanMax) {
        k_scan_small<<<1, kScanThreads, 0, st>>>(in, out, n, flags, nflags, nclear, rb, seq);
        return hipGetLastError();
    }
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, st);
    if (e != hipSuccess) return e;
    static thread_local void* tmp = nullptr;
    static thread_local size_t tmp_cap = 0;
    if (bytes > tmp_cap) {
        if (tmp) hipFree(tmp);
        e = hipMalloc(&tmp, bytes);
        if (e != hipSuccess) { tmp = nullptr; tmp_cap = 0; return e; }
        tmp_cap = bytes;
    }
    if ((e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, st)) != hipSuccess) return e;
    if (rb) {
        if (nflags > 0) {
            k_readback<<<1, 64, 0, st>>>(flags, nflags, rb, 0, nclear);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        k_readback<<<1, 64, 0, st>>>(out + n - 1, 1, rb + nflags, seq, 0);   // total, then the marker
        e = hipGetLastError();
    }
    return e;
}

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return e_; } while (0)

// Bin grid of every structure for cutoff rc: computed on the host from the
// cells once per batch and cutoff (the cells are read at the first build after
// ann_set_batch), then kept on the device.
static hipError_t grid_setup(ann_ctx* ctx, double rc) {
    if (ctx->grid_valid && ctx->grid_rc == rc) return hipSuccess;
    const int S = ctx->nstruc;
    hipStream_t st = ctx->stream;
    // the cells: the host copy given with the batch, else read from the device
    std::vector<double> cells;
    if (ctx->h_cells_valid) {
        cells = ctx->h_cells;
    } else {
        cells.resize(9 * (size_t)S);
        HIPCHK(hipMemcpyAsync(cells.data(), ctx->d_cells, sizeof(double) * 9 * S, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    std::vector<CellInfo> info(S);
    std::vector<double> rinv(9 * (size_t)S);
    std::vector<int> bin_struc;
    int nbins = 0;
    for (int s = 0; s < S; ++s) {
        inv3(&cells[9 * s], &rinv[9 * s]);
        CellInfo& c = info[s];
        for (int a = 0; a < 3; ++a) {
            // bins at least rc/2 wide, so that the search box is +-2 bins
            const double* r = &rinv[9 * s + 3 * a];
            const double w = 1.0 / std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            c.nb[a] = std::max(1, (int)std::floor(2.0 * w / rc * (1.0 - 1e-12)));
            c.nb[a] = std::min(c.nb[a], 256);
            c.k[a] = (int)std::ceil(rc * c.nb[a] / w * (1.0 + 1e-12));
        }
        c.nbins = c.nb[0] * c.nb[1] * c.nb[2];
        c.boff = nbins;
        nbins += c.nbins;
        // 2x2x2 bins per staged block when the enlarged search box still fits
        const double per_bin = (double)(ctx->atom_off[s + 1] - ctx->atom_off[s]) / c.nbins;
        const double box2 = (double)(std::min(2, c.nb[0]) + 2 * c.k[0]) * (std::min(2, c.nb[1]) + 2 * c.k[1]) *
                            (std::min(2, c.nb[2]) + 2 * c.k[2]);
        c.sb = (ctx->tune_sb != 1 && box2 * per_bin <= 0.7 * kStageCap && box2 <= kStageImages) ? 2 : 1;
        for (int a = 0; a < 3; ++a) c.nsb[a] = (c.nb[a] + c.sb - 1) / c.sb;
        const int nsbins = c.nsb[0] * c.nsb[1] * c.nsb[2];
        c.sboff = (int)bin_struc.size();
        bin_struc.insert(bin_struc.end(), nsbins, s);       // structure of each super-bin
    }
    if (S > ctx->struc_cap || !ctx->d_cellinfo) {
        if (ctx->d_cellinfo) { hipFree(ctx->d_cellinfo); hipFree(ctx->d_cellmat); }
        HIPCHK(hipMalloc(&ctx->d_cellinfo, sizeof(CellInfo) * S));
        HIPCHK(hipMalloc(&ctx->d_cellmat, sizeof(double) * 9 * S));
        ctx->struc_cap = S;
    }
    if (nbins + 1 > ctx->bin_cap) {
        if (ctx->d_bin_cnt) { hipFree(ctx->d_bin_cnt); hipFree(ctx->d_bin_start); hipFree(ctx->d_bin_struc); }
        // counts and overflow marks share one allocation; the counts are
        // cleared before the first build and return to zero by the end of
        // every build
        HIPCHK(hipMalloc(&ctx->d_bin_cnt, sizeof(int) * 3 * (nbins + 1)));
        HIPCHK(hipMalloc(&ctx->d_bin_start, sizeof(int) * (nbins + 1)));
        HIPCHK(hipMalloc(&ctx->d_bin_struc, sizeof(int) * (nbins + 1)));
        ctx->bin_cap = nbins + 1;
        ctx->nb_dirty = true;
    }
    ctx->d_bin_ovf = ctx->d_bin_cnt + 2 * (size_t)ctx->bin_cap;
    HIPCHK(staged_upload(ctx, kStCellInfo, ctx->d_cellinfo, info.data(), sizeof(CellInfo) * S));
    HIPCHK(staged_upload(ctx, kStRinv, ctx->d_cellmat, rinv.data(), sizeof(double) * 9 * S));
    HIPCHK(staged_upload(ctx, kStBinStruc, ctx->d_bin_struc, bin_struc.data(), sizeof(int) * bin_struc.size()));
    ctx->grid_valid = true;
    ctx->grid_rc = rc;
    ctx->grid_nbins = nbins;
    ctx->grid_nsb = (int)bin_struc.size();
    return hipSuccess;
}

// Wait until the pinned word slot holds seq (written last by the readback of a
// kernel on the stream), polling instead of synchronising the stream so that
// work queued behind that kernel keeps the device busy meanwhile.
static hipError_t wait_marker(ann_ctx* ctx, int slot, int seq) {
    volatile const int* w = ctx->h_rb + slot;
    // The stream is queried only when the marker is late (2 ms, then every
    // millisecond): queried every 256 polls, the queries while the build's
    // kernels ran left a 5.6 us idle gap before the next kernel on the stream
    // (1.0 us without them).
    auto next = std::chrono::steady_clock::now() + std::chrono::milliseconds(2);
    for (long it = 0;; ++it) {
        if (*w == seq) break;
        if ((it & 255) == 255 && std::chrono::steady_clock::now() >= next) {
            const hipError_t e = hipStreamQuery(ctx->stream);
            if (e == hipSuccess) {                   // stream drained: the marker must be there
                if (*w == seq) break;
                return hipErrorUnknown;
            }
            if (e != hipErrorNotReady) return e;
            next = std::chrono::steady_clock::now() + std::chrono::milliseconds(1);
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return hipSuccess;
}

static void launch_compact(ann_ctx* ctx, int cap) {
    const int N = ctx->natoms;
    hipStream_t st = ctx->stream;
    NList& nl = ctx->nl;
    {
        Launch L(ctx, "nb_compact");
        k_compact<<<N, 128, 0, st>>>(N, ctx->d_pos, ctx->d_struc_of, ctx->d_cells, ctx->d_species, ctx->d_keys,
                                     nl.off, nl.idx, nl.img, nl.rev, nl.geo, nl.nsp, ctx->d_bin_atoms, cap);
    }
    // a missing reverse slot (impossible for the symmetric FP64 test) is
    // reported by the next synchronising call (ann_nlist_size, next build)
    Launch L(ctx, "nb_rev");
    constexpr int kRevRows = 4;                 // centres (wavefronts) per workgroup
    k_rev<<<(N + kRevRows - 1) / kRevRows, kRevRows * kWave, 0, st>>>(N, nl.idx, nl.img, nl.off, nl.rev, ctx->d_flag + 2,
                                                                      ctx->d_bin_atoms, cap);
}

hipError_t neighbors_build(ann_ctx* ctx, double rc) {
    const int N = ctx->natoms;
    hipStream_t st = ctx->stream;
    HIPCHK(grid_setup(ctx, rc));
    const int nbins = ctx->grid_nbins;
    int64_t kc = ctx->keys_cap;
    HIPCHK(grow(ctx->d_keys, kc, (int64_t)N * kMaxNeigh));
    ctx->keys_cap = kc;
    int64_t rcap = ctx->binrec_cap;
    HIPCHK(grow(ctx->d_binrec, rcap, (int64_t)N));
    ctx->binrec_cap = rcap;
    const CellInfo* dci = reinterpret_cast<const CellInfo*>(ctx->d_cellinfo);
    // The bin counts and flag[0..1] are zero at the start of a build without a
    // clear launch: the fill counts every bin back down and the readback scan
    // clears the flags it reports (flag[2], the reverse-slot error, stays set
    // until reported).  New bin arrays, or a build that stopped part-way
    // (nb_dirty), are cleared here first.
    if (ctx->nb_dirty) {
        HIPCHK(hipMemsetAsync(ctx->d_bin_cnt, 0, sizeof(int) * (size_t)ctx->bin_cap, st));
        HIPCHK(hipMemsetAsync(ctx->d_flag, 0, sizeof(int) * 2, st));
    }
    ctx->nb_dirty = true;
    {
        {
            Launch L(ctx, "nb_bin");
            k_bin<<<(N + 255) / 256, 256, 0, st>>>(N, ctx->d_pos, ctx->d_struc_of, dci, ctx->d_cellmat,
                                                  ctx->d_bin_of, ctx->d_shift, ctx->d_bin_cnt, ctx->d_flag);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(exclusive_scan(ctx->d_bin_cnt, ctx->d_bin_start, nbins + 1, st));
        k_fill_bins<<<(N + 255) / 256, 256, 0, st>>>(N, ctx->d_bin_of, ctx->d_bin_start, ctx->d_bin_cnt,
                                                    ctx->d_bin_atoms, ctx->d_pos, ctx->d_shift, ctx->d_binrec);
        HIPCHK(hipGetLastError());
    }
    {
        Launch L(ctx, "nb_search");
        k_search_bins<<<ctx->grid_nsb, kStageThreads, 0, st>>>(dci, ctx->d_cells, ctx->d_bin_struc, ctx->d_bin_start,
                                                      ctx->d_binrec, ctx->d_pos, rc * rc, ctx->d_keys, ctx->d_cnt,
                                                      ctx->d_flag, ctx->d_bin_ovf);
    }
    HIPCHK(hipGetLastError());
    NList& nl = ctx->nl;
    // Offsets, then flags, pair count and a completion marker to pinned host
    // words -- the build's one wait.  While the host waits, the compaction runs
    // speculatively into the current list arrays; it does nothing if the list
    // does not fit, and is repeated after growing them.
    int seq = ++ctx->rb_seq;
    HIPCHK(exclusive_scan(ctx->d_cnt, ctx->nl.off, N + 1, st, ctx->d_flag, 4, ctx->d_rb, seq, 2));
    ctx->nb_dirty = false;
    const int spec_cap = (int)std::min<int64_t>(nl.cap, INT32_MAX);
    if (spec_cap > 0) launch_compact(ctx, spec_cap);
    HIPCHK(hipGetLastError());
    HIPCHK(wait_marker(ctx, 5, seq));
    volatile const int* rb = ctx->h_rb;
    int flags[4] = {rb[0], rb[1], rb[2], rb[3]};
    int total = rb[4];
    if (flags[2]) {
        // reported once: the flag of the list that lacked a reverse slot is cleared
        HIPCHK(hipMemsetAsync(ctx->d_flag + 2, 0, sizeof(int), st));
        ctx->nb_status = ANN_ERR_HIP;
        set_error("the previous neighbour list was not symmetric (reverse slot missing)");
        return hipErrorInvalidValue;
    }
    if (flags[0] & 8) {
        // bins with a search box too large to stage: one wavefront per centre atom
        ctx->nb_dirty = true;
        {
            Launch L(ctx, "nb_search_wide");
            k_search<<<N, 64, 0, st>>>(N, ctx->d_pos, ctx->d_struc_of, dci, ctx->d_cells, ctx->d_bin_of,
                                       ctx->d_shift, ctx->d_bin_start, ctx->d_bin_atoms, rc * rc, ctx->d_keys,
                                       ctx->d_cnt, ctx->d_flag, ctx->d_bin_ovf);
        }
        HIPCHK(hipGetLastError());
        seq = ++ctx->rb_seq;
        HIPCHK(exclusive_scan(ctx->d_cnt, ctx->nl.off, N + 1, st, ctx->d_flag, 4, ctx->d_rb, seq, 2));
        ctx->nb_dirty = false;
        if (spec_cap > 0) launch_compact(ctx, spec_cap);
        HIPCHK(hipGetLastError());
        HIPCHK(wait_marker(ctx, 5, seq));
        // the first pass's flags were cleared when read: merge both passes
        flags[0] |= rb[0];
        flags[1] = std::max(flags[1], (int)rb[1]);
        total = rb[4];
    }
    if (flags[0] & 16) {
        ctx->nb_status = ANN_ERR_ARG;
        set_error("an atom position is not finite (or lies beyond 1e8 cells)");
        return hipErrorInvalidValue;
    }
    if (flags[0] & 2) {
        ctx->nb_status = ANN_ERR_UNSUPPORTED;
        set_error("neighbour search box exceeds the image limit (cell far narrower than rc)");
        return hipErrorInvalidValue;
    }
    if (flags[0] & 1) {
        ctx->nb_status = ANN_ERR_CAPACITY;
        set_error("an atom has " + std::to_string(flags[1]) + " neighbours; capacity is " + std::to_string(kMaxNeigh));
        return hipErrorInvalidValue;
    }
    nl.npairs = total;
    nl.maxn = flags[1];
    if (spec_cap > 0 && total <= spec_cap) return hipSuccess;     // the speculative compaction stands
    if (spec_cap > 0) HIPCHK(hipStreamSynchronize(st));            // its kernels still hold the old arrays
    int64_t c1 = nl.cap, c2 = nl.cap, c3 = nl.cap, c4 = nl.cap, c5 = nl.cap;
    HIPCHK(grow(nl.idx, c1, total));
    HIPCHK(grow(nl.img, c2, total));
    HIPCHK(grow(nl.rev, c3, total));
    HIPCHK(grow(nl.geo, c4, total));
    HIPCHK(grow(nl.nsp, c5, total));
    nl.cap = std::min(std::min(c1, c2), std::min(std::min(c3, c4), c5));
    launch_compact(ctx, (int)std::min<int64_t>(nl.cap, INT32_MAX));
    return hipGetLastError();
}

// ---- a neighbour list supplied by the caller ----

__global__ void k_neigh_species(int64_t P, const int* __restrict__ idx, const int* __restrict__ species,
                                uint8_t* __restrict__ nsp) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < P) nsp[s] = (uint8_t)species[idx[s]];
}

// Host CSR (off [N+1], idx [P], img [P][3]) of the current batch into the
// context's list: packed images, reverse slots found by sorting the entries,
// neighbour species and pair geometry on the device.  Returns an ANN status.
int set_neighbors_host(ann_ctx* ctx, const int* off, const int* idx, const int* img, double rc) {
    const int N = ctx->natoms;
    if (off[0] != 0) return fail(ANN_ERR_ARG, "neighbour list: off[0] must be 0");
    int maxn = 0;
    for (int i = 0; i < N; ++i) {
        if (off[i + 1] < off[i]) return fail(ANN_ERR_ARG, "neighbour list: offsets must not decrease");
        maxn = std::max(maxn, off[i + 1] - off[i]);
    }
    if (maxn > kMaxNeigh)
        return fail(ANN_ERR_CAPACITY, "an atom has " + std::to_string(maxn) + " neighbours; capacity is " +
                                          std::to_string(kMaxNeigh));
    const int64_t P = off[N];
    std::vector<uint32_t> pimg(std::max<int64_t>(P, 1));
    std::vector<int> rev(std::max<int64_t>(P, 1), -1), owner(std::max<int64_t>(P, 1));
    int s = 0;
    for (int i = 0; i < N; ++i) {
        while (ctx->atom_off[s + 1] <= i) ++s;
        for (int q = off[i]; q < off[i + 1]; ++q) {
            const int j = idx[q];
            if (j < ctx->atom_off[s] || j >= ctx->atom_off[s + 1])
                return fail(ANN_ERR_ARG, "neighbour list: partner outside the centre's structure");
            for (int x = 0; x < 3; ++x)
                if (img[3 * (int64_t)q + x] < -511 || img[3 * (int64_t)q + x] > 511)
                    return fail(ANN_ERR_UNSUPPORTED, "neighbour list: image vector beyond +-511 cells");
            pimg[q] = pack_img(img[3 * (int64_t)q], img[3 * (int64_t)q + 1], img[3 * (int64_t)q + 2]);
            owner[q] = i;
            if (j == i && pimg[q] == pack_img(0, 0, 0))
                return fail(ANN_ERR_ARG, "neighbour list: atom " + std::to_string(i) + " lists itself in its own cell");
        }
    }
    // reverse slots: entry (i -> j, n) pairs with (j -> i, -n); sort by (centre, partner, image)
    std::vector<int64_t> order(P);
    for (int64_t q = 0; q < P; ++q) order[q] = q;
    auto key = [&](int64_t q) { return std::make_tuple(owner[q], idx[q], pimg[q]); };
    std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return key(x) < key(y); });
    for (int64_t q = 1; q < P; ++q)
        if (key(order[q]) == key(order[q - 1]))
            return fail(ANN_ERR_ARG, "neighbour list: atom " + std::to_string(owner[order[q]]) + " lists partner " +
                                         std::to_string(idx[order[q]]) + " twice with the same image");
    for (int64_t q = 0; q < P; ++q) {
        int n0, n1, n2;
        unpack_img(pimg[q], n0, n1, n2);
        const auto want = std::make_tuple(idx[q], owner[q], pack_img(-n0, -n1, -n2));
        auto it = std::lower_bound(order.begin(), order.end(), want,
                                   [&](int64_t x, const std::tuple<int, int, uint32_t>& w) { return key(x) < w; });
        if (it == order.end() || key(*it) != want)
            return fail(ANN_ERR_ARG, "neighbour list is not symmetric: entry " + std::to_string(q) + " (atom " +
                                         std::to_string(owner[q]) + " -> " + std::to_string(idx[q]) +
                                         ") has no reverse entry");
        rev[q] = (int)*it;
    }
    NList& nl = ctx->nl;
    hipStream_t st = ctx->stream;
    int64_t c1 = nl.cap, c2 = nl.cap, c3 = nl.cap, c4 = nl.cap, c5 = nl.cap;
    if (P > nl.cap) {
        ANN_HIP(hipStreamSynchronize(st));
        ANN_HIP(grow(nl.idx, c1, P));
        ANN_HIP(grow(nl.img, c2, P));
        ANN_HIP(grow(nl.rev, c3, P));
        ANN_HIP(grow(nl.geo, c4, P));
        ANN_HIP(grow(nl.nsp, c5, P));
        nl.cap = std::min(std::min(c1, c2), std::min(std::min(c3, c4), c5));
    }
    ANN_HIP(hipMemcpyAsync(nl.off, off, sizeof(int) * (N + 1), hipMemcpyHostToDevice, st));
    if (P > 0) {
        ANN_HIP(hipMemcpyAsync(nl.idx, idx, sizeof(int) * P, hipMemcpyHostToDevice, st));
        ANN_HIP(hipMemcpyAsync(nl.img, pimg.data(), sizeof(uint32_t) * P, hipMemcpyHostToDevice, st));
        ANN_HIP(hipMemcpyAsync(nl.rev, rev.data(), sizeof(int) * P, hipMemcpyHostToDevice, st));
        k_neigh_species<<<(unsigned)((P + 255) / 256), 256, 0, st>>>(P, nl.idx, ctx->d_species, nl.nsp);
        ANN_HIP(hipGetLastError());
    }
    nl.npairs = P;
    nl.maxn = maxn;
    ANN_HIP(refresh_geometry(ctx));
    ANN_HIP(hipStreamSynchronize(st));       // the host vectors go out of scope
    (void)rc;
    return ANN_OK;
}

// Pair geometry of every slot from the current positions and the list's image
// vectors (verlet-list reuse between rebuilds, the geometry of
// NNP::symm(..., verlet::List) src/nnp/nnp.cpp:389-395): same FP64 operation
// order as the build, so an unchanged position set reproduces the build's
// geometry bit for bit.  One wavefront per atom, lanes over its slots.
__global__ __launch_bounds__(256) void k_geom(int N, const double* __restrict__ pos, const int* __restrict__ sof,
                                              const double* __restrict__ cells, const int* __restrict__ off,
                                              const int* __restrict__ idx, const uint32_t* __restrict__ img,
                                              float4* __restrict__ geo) {
    const int i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (i >= N) return;
    const double* R = cells + 9 * sof[i];
    for (int s = off[i] + lane; s < off[i + 1]; s += kWave) {
        int n0, n1, n2;
        unpack_img(img[s], n0, n1, n2);
        double x, y, z, d2;
        pair_vec(pos, i, idx[s], R, n0, n1, n2, x, y, z, d2);
        geo[s] = make_float4((float)x, (float)y, (float)z, (float)sqrt(d2));
    }
}

hipError_t refresh_geometry(ann_ctx* ctx) {
    const int N = ctx->natoms;
    if (N == 0) return hipSuccess;
    Launch L(ctx, "nb_geometry");
    k_geom<<<(N + 3) / 4, 256, 0, ctx->stream>>>(N, ctx->d_pos, ctx->d_struc_of, ctx->d_cells, ctx->nl.off,
                                                 ctx->nl.idx, ctx->nl.img, ctx->nl.geo);
    return hipGetLastError();
}

// Positions back into the cell, x <- R (f - floor f) with f = R^-1 x
// (Cell::returnToCell, src/format/xyz_struc.cpp:176-181).
__global__ void k_wrap(int N, double* __restrict__ pos, const int* __restrict__ sof,
                       const double* __restrict__ cells) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double* R = cells + 9 * sof[i];
    const double a = R[0], b = R[1], c = R[2], d = R[3], e = R[4], f = R[5], g = R[6], h = R[7], k = R[8];
    const double A = e * k - f * h, B = -(d * k - f * g), C = d * h - e * g;
    const double id = 1.0 / (a * A + b * B + c * C);
    const double Ri[9] = {A * id, -(b * k - c * h) * id, (b * f - c * e) * id,
                          B * id, (a * k - c * g) * id,  -(a * f - c * d) * id,
                          C * id, -(a * h - b * g) * id, (a * e - b * d) * id};
    const double x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    double fr[3];
    for (int q = 0; q < 3; ++q) {
        const double t = Ri[3 * q] * x[0] + Ri[3 * q + 1] * x[1] + Ri[3 * q + 2] * x[2];
        fr[q] = t - floor(t);
    }
    for (int q = 0; q < 3; ++q) pos[3 * i + q] = R[3 * q] * fr[0] + R[3 * q + 1] * fr[1] + R[3 * q + 2] * fr[2];
}

hipError_t wrap_positions(ann_ctx* ctx, double* pos) {
    const int N = ctx->natoms;
    if (N == 0) return hipSuccess;
    k_wrap<<<(N + 255) / 256, 256, 0, ctx->stream>>>(N, pos, ctx->d_struc_of, ctx->d_cells);
    return hipGetLastError();
}

hipError_t struc_index(ann_ctx* ctx) {
    k_struc_of<<<ctx->nstruc, 256, 0, ctx->stream>>>(ctx->d_atom_off, ctx->nstruc, ctx->d_struc_of);
    return hipGetLastError();
}

}  // namespace ann
