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
// LDS and sorts them with a bitonic network.  Integer work; the roofline is
// HBM/latency (positions in, 16 B + 13 B per pair out).
#include "common.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

namespace ann {

constexpr int kMaxImages = 1024;   // bin images searched per atom

// r_i - r_j - R n in FP64, evaluated exactly as the oracle does (no
// contraction) so that inclusion decisions agree bit for bit.
__device__ __forceinline__ void pair_vec(const double* __restrict__ pos, int i, int j,
                                         const double* __restrict__ R, int n0, int n1, int n2,
                                         double& x, double& y, double& z, double& d2) {
#pragma clang fp contract(off)
    const double m0 = n0, m1 = n1, m2 = n2;
    x = __dsub_rn(__dsub_rn(pos[3 * i], pos[3 * j]),
                  __dadd_rn(__dadd_rn(__dmul_rn(R[0], m0), __dmul_rn(R[1], m1)), __dmul_rn(R[2], m2)));
    y = __dsub_rn(__dsub_rn(pos[3 * i + 1], pos[3 * j + 1]),
                  __dadd_rn(__dadd_rn(__dmul_rn(R[3], m0), __dmul_rn(R[4], m1)), __dmul_rn(R[5], m2)));
    z = __dsub_rn(__dsub_rn(pos[3 * i + 2], pos[3 * j + 2]),
                  __dadd_rn(__dadd_rn(__dmul_rn(R[6], m0), __dmul_rn(R[7], m1)), __dmul_rn(R[8], m2)));
    d2 = __dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z));
}

// the same expression on positions already in registers
__device__ __forceinline__ void pair_vec_v(double pix, double piy, double piz, double pjx, double pjy, double pjz,
                                           const double* __restrict__ R, int n0, int n1, int n2,
                                           double& x, double& y, double& z, double& d2) {
#pragma clang fp contract(off)
    const double m0 = n0, m1 = n1, m2 = n2;
    x = __dsub_rn(__dsub_rn(pix, pjx), __dadd_rn(__dadd_rn(__dmul_rn(R[0], m0), __dmul_rn(R[1], m1)), __dmul_rn(R[2], m2)));
    y = __dsub_rn(__dsub_rn(piy, pjy), __dadd_rn(__dadd_rn(__dmul_rn(R[3], m0), __dmul_rn(R[4], m1)), __dmul_rn(R[5], m2)));
    z = __dsub_rn(__dsub_rn(piz, pjz), __dadd_rn(__dadd_rn(__dmul_rn(R[6], m0), __dmul_rn(R[7], m1)), __dmul_rn(R[8], m2)));
    d2 = __dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z));
}

// per structure grid: nb[3], k[3], nbins, bin offset
struct CellInfo {
    int nb[3], k[3], nbins, boff;
};

__global__ void k_struc_of(const int* __restrict__ off, int nstruc, int* __restrict__ sof) {
    const int s = blockIdx.x;
    if (s >= nstruc) return;
    for (int i = off[s] + threadIdx.x; i < off[s + 1]; i += blockDim.x) sof[i] = s;
}

__global__ void k_bin(int N, const double* __restrict__ pos, const int* __restrict__ sof,
                      const CellInfo* __restrict__ ci, const double* __restrict__ rinv,
                      int* __restrict__ bin_of, int* __restrict__ shift, int* __restrict__ bin_cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int s = sof[i];
    const CellInfo c = ci[s];
    const double* Ri = rinv + 9 * s;
    int b[3];
    for (int a = 0; a < 3; ++a) {
        const double f = Ri[3 * a] * pos[3 * i] + Ri[3 * a + 1] * pos[3 * i + 1] + Ri[3 * a + 2] * pos[3 * i + 2];
        const double fl = floor(f);
        shift[3 * i + a] = (int)fl;
        int bb = (int)floor((f - fl) * c.nb[a]);
        b[a] = bb < 0 ? 0 : (bb >= c.nb[a] ? c.nb[a] - 1 : bb);
    }
    const int g = c.boff + (b[0] * c.nb[1] + b[1]) * c.nb[2] + b[2];
    bin_of[i] = g;
    atomicAdd(&bin_cnt[g], 1);
}

__global__ void k_fill_bins(int N, const int* __restrict__ bin_of, const int* __restrict__ bin_start,
                            int* __restrict__ bin_fill, int* __restrict__ bin_atoms) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int g = bin_of[i];
    const int p = atomicAdd(&bin_fill[g], 1);
    bin_atoms[bin_start[g] + p] = i;
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
// (every atom of the (2k+1)^3 bin images, with its image offset) are copied
// to LDS once and sorted by (j, image - shift_j); each wavefront then tests
// one centre atom of the bin against them in that order, so its accepted
// keys (j, si + image - shift_j) come out already in canonical order and go
// straight to the key scratch without a per-atom sort.  Bins whose search box
// holds more than kStageCap candidates or kStageImages images are flagged and
// left to k_search.
constexpr int kStageCap = 1024;
constexpr int kStageImages = 125;
constexpr int kStageThreads = 256;

__global__ __launch_bounds__(kStageThreads) void k_search_bins(
    const double* __restrict__ pos, const CellInfo* __restrict__ ci, const double* __restrict__ cells,
    const int* __restrict__ bin_struc, const int* __restrict__ shift, const int* __restrict__ bin_start,
    const int* __restrict__ bin_atoms, double rc2, uint64_t* __restrict__ keys_out, int* __restrict__ cnt_out,
    int* __restrict__ flag, int* __restrict__ bin_ovf) {
    __shared__ int im_bin[kStageImages];
    __shared__ int im_m[kStageImages];
    __shared__ int im_pre[kStageImages + 1];
    __shared__ uint64_t ckey[kStageCap];
    __shared__ int cslot[kStageCap];
    __shared__ double cpos[3 * kStageCap];
    __shared__ int s_total;
    const int g = blockIdx.x, t = threadIdx.x;
    const int lane = t % kWave, wave = t / kWave;
    const int a0 = bin_start[g], a1 = bin_start[g + 1];
    if (a0 == a1) return;                                   // empty bin: no centres
    const int s = bin_struc[g];
    const CellInfo c = ci[s];
    const double* R = cells + 9 * s;
    const int gi = g - c.boff;
    const int b0 = gi / (c.nb[1] * c.nb[2]), b1 = (gi / c.nb[2]) % c.nb[1], b2 = gi % c.nb[2];
    const int e0 = 2 * c.k[0] + 1, e1 = 2 * c.k[1] + 1, e2 = 2 * c.k[2] + 1;
    const int nimg = e0 * e1 * e2;
    if (nimg > kStageImages) {
        if (t == 0) { bin_ovf[g] = 1; atomicOr(flag, 8); }
        return;
    }
    for (int q = t; q < nimg; q += kStageThreads) {
        const int u0 = b0 - c.k[0] + q / (e1 * e2);
        const int u1 = b1 - c.k[1] + (q / e2) % e1;
        const int u2 = b2 - c.k[2] + q % e2;
        const int m0 = floordiv(u0, c.nb[0]), m1 = floordiv(u1, c.nb[1]), m2 = floordiv(u2, c.nb[2]);
        const int gg = c.boff + ((u0 - m0 * c.nb[0]) * c.nb[1] + (u1 - m1 * c.nb[1])) * c.nb[2] + (u2 - m2 * c.nb[2]);
        im_bin[q] = gg;
        im_m[q] = (int)pack_img(m0, m1, m2);
        im_pre[q + 1] = bin_start[gg + 1] - bin_start[gg];
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
    if (total > kStageCap) {
        if (t == 0) { bin_ovf[g] = 1; atomicOr(flag, 8); }
        return;
    }
    // stage candidates: key (j, image offset - shift_j), slot, position
    for (int q = t; q < total; q += kStageThreads) {
        int lo = 0, hi = nimg;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (im_pre[mid] <= q) lo = mid; else hi = mid;
        }
        const int j = bin_atoms[bin_start[im_bin[lo]] + (q - im_pre[lo])];
        int m0, m1, m2;
        unpack_img((uint32_t)im_m[lo], m0, m1, m2);
        const uint32_t d = pack_img(m0 - shift[3 * j], m1 - shift[3 * j + 1], m2 - shift[3 * j + 2]);
        ckey[q] = ((uint64_t)j << 30) | (uint64_t)d;
        cslot[q] = q;
        cpos[3 * q] = pos[3 * j];
        cpos[3 * q + 1] = pos[3 * j + 1];
        cpos[3 * q + 2] = pos[3 * j + 2];
    }
    int np2 = 1;
    while (np2 < total) np2 <<= 1;
    for (int q = total + t; q < np2; q += kStageThreads) { ckey[q] = ~0ull; cslot[q] = 0; }
    __syncthreads();
    for (int k = 2; k <= np2; k <<= 1) {
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int q = t; q < np2; q += kStageThreads) {
                const int p = q ^ jj;
                if (p > q) {
                    const uint64_t ka = ckey[q], kb = ckey[p];
                    const bool up = (q & k) == 0;
                    if ((ka > kb) == up) {
                        ckey[q] = kb; ckey[p] = ka;
                        const int sa = cslot[q]; cslot[q] = cslot[p]; cslot[p] = sa;
                    }
                }
            }
            __syncthreads();
        }
    }
    // centres of the bin, one wavefront each
    for (int ca = a0 + wave; ca < a1; ca += kStageThreads / kWave) {
        const int i = bin_atoms[ca];
        const int si0 = shift[3 * i], si1 = shift[3 * i + 1], si2 = shift[3 * i + 2];
        const double pix = pos[3 * i], piy = pos[3 * i + 1], piz = pos[3 * i + 2];
        uint64_t* dst = keys_out + (size_t)i * kMaxNeigh;
        int count = 0;
        for (int base = 0; base < total; base += kWave) {
            const int q = base + lane;
            bool ok = false;
            uint64_t key = 0;
            if (q < total) {
                const uint64_t ck = ckey[q];
                const int sl = cslot[q];
                const int j = (int)(ck >> 30);
                int d0, d1, d2i;
                unpack_img((uint32_t)(ck & 0x3fffffffu), d0, d1, d2i);
                const int n0 = si0 + d0, n1 = si1 + d1, n2 = si2 + d2i;
                double x, y, z, d2;
                pair_vec_v(pix, piy, piz, cpos[3 * sl], cpos[3 * sl + 1], cpos[3 * sl + 2], R, n0, n1, n2, x, y, z, d2);
                ok = d2 > 0.0 && d2 < rc2 && !(j == i && n0 == 0 && n1 == 0 && n2 == 0);
                key = ((uint64_t)j << 30) | (uint64_t)pack_img(n0, n1, n2);
            }
            const uint64_t mask = __ballot(ok);
            const int before = __popcll(mask & ((1ull << lane) - 1ull));
            if (ok && count + before < kMaxNeigh) dst[count + before] = key;
            count += __popcll(mask);
        }
        if (lane == 0) {
            cnt_out[i] = count < kMaxNeigh ? count : kMaxNeigh;
            if (count > kMaxNeigh) atomicOr(flag, 1);
            atomicMax(flag + 1, count);
        }
    }
}

// CSR fill: index, packed image, neighbour species and FP32 pair geometry
__global__ void k_compact(int N, const double* __restrict__ pos, const int* __restrict__ sof,
                          const double* __restrict__ cells, const int* __restrict__ species,
                          const uint64_t* __restrict__ keys, const int* __restrict__ off,
                          int* __restrict__ idx, uint32_t* __restrict__ img, float4* __restrict__ geo,
                          uint8_t* __restrict__ nsp) {
    const int i = blockIdx.x;
    if (i >= N) return;
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
        geo[o + t] = make_float4((float)x, (float)y, (float)z, (float)sqrt(d2));
        nsp[o + t] = (uint8_t)species[j];
    }
}

// reverse slot by binary search in the partner's sorted key list
__global__ void k_rev(int N, const uint64_t* __restrict__ keys, const int* __restrict__ off,
                      int* __restrict__ rev, int* __restrict__ flag) {
    const int i = blockIdx.x;
    if (i >= N) return;
    const int o = off[i], n = off[i + 1] - o;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const uint64_t key = keys[(size_t)i * kMaxNeigh + t];
        const int j = (int)(key >> 30);
        int n0, n1, n2;
        unpack_img((uint32_t)(key & 0x3fffffffu), n0, n1, n2);
        const uint64_t want = ((uint64_t)i << 30) | (uint64_t)pack_img(-n0, -n1, -n2);
        const uint64_t* kj = keys + (size_t)j * kMaxNeigh;
        int lo = 0, hi = off[j + 1] - off[j];
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (kj[mid] < want) lo = mid + 1; else hi = mid;
        }
        if (lo < off[j + 1] - off[j] && kj[lo] == want) rev[o + t] = off[j] + lo;
        else { rev[o + t] = o + t; atomicOr(flag, 4); }
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

hipError_t exclusive_scan(const int* in, int* out, int n, hipStream_t st) {
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
    return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, st);
}

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return e_; } while (0)

// Bin grid of every structure for cutoff rc: computed on the host from the
// cells once per batch and cutoff (the cells are read at the first build after
// ann_set_batch), then kept on the device.
static hipError_t grid_setup(ann_ctx* ctx, double rc) {
    if (ctx->grid_valid && ctx->grid_rc == rc) return hipSuccess;
    const int S = ctx->nstruc;
    hipStream_t st = ctx->stream;
    std::vector<double> cells(9 * (size_t)S);
    HIPCHK(hipMemcpyAsync(cells.data(), ctx->d_cells, sizeof(double) * 9 * S, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
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
        bin_struc.insert(bin_struc.end(), c.nbins, s);
    }
    if (S > ctx->struc_cap || !ctx->d_cellinfo) {
        if (ctx->d_cellinfo) { hipFree(ctx->d_cellinfo); hipFree(ctx->d_cellmat); }
        HIPCHK(hipMalloc(&ctx->d_cellinfo, sizeof(CellInfo) * S));
        HIPCHK(hipMalloc(&ctx->d_cellmat, sizeof(double) * 9 * S));
        ctx->struc_cap = S;
    }
    if (nbins + 1 > ctx->bin_cap) {
        if (ctx->d_bin_cnt) { hipFree(ctx->d_bin_cnt); hipFree(ctx->d_bin_start); hipFree(ctx->d_bin_struc); }
        // counts, fill cursors and overflow marks share one allocation (one clear per build)
        HIPCHK(hipMalloc(&ctx->d_bin_cnt, sizeof(int) * 3 * (nbins + 1)));
        HIPCHK(hipMalloc(&ctx->d_bin_start, sizeof(int) * (nbins + 1)));
        HIPCHK(hipMalloc(&ctx->d_bin_struc, sizeof(int) * (nbins + 1)));
        ctx->bin_cap = nbins + 1;
    }
    ctx->d_bin_ovf = ctx->d_bin_cnt + 2 * (size_t)ctx->bin_cap;
    HIPCHK(hipMemcpyAsync(ctx->d_cellinfo, info.data(), sizeof(CellInfo) * S, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->d_cellmat, rinv.data(), sizeof(double) * 9 * S, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->d_bin_struc, bin_struc.data(), sizeof(int) * nbins, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));      // the host vectors go out of scope
    ctx->grid_valid = true;
    ctx->grid_rc = rc;
    ctx->grid_nbins = nbins;
    return hipSuccess;
}

hipError_t neighbors_build(ann_ctx* ctx, double rc) {
    const int N = ctx->natoms;
    hipStream_t st = ctx->stream;
    HIPCHK(grid_setup(ctx, rc));
    const int nbins = ctx->grid_nbins;
    int64_t kc = ctx->keys_cap;
    HIPCHK(grow(ctx->d_keys, kc, (int64_t)N * kMaxNeigh));
    ctx->keys_cap = kc;
    const CellInfo* dci = reinterpret_cast<const CellInfo*>(ctx->d_cellinfo);
    int* bin_fill = ctx->d_bin_cnt + (nbins + 1);
    HIPCHK(hipMemsetAsync(ctx->d_bin_cnt, 0, sizeof(int) * 3 * (size_t)ctx->bin_cap, st));
    HIPCHK(hipMemsetAsync(ctx->d_flag, 0, sizeof(int) * 2, st));   // flag[2]: sticky reverse-slot error
    {
        Launch L(ctx, "nb_bin");
        k_bin<<<(N + 255) / 256, 256, 0, st>>>(N, ctx->d_pos, ctx->d_struc_of, dci, ctx->d_cellmat,
                                              ctx->d_bin_of, ctx->d_shift, ctx->d_bin_cnt);
    }
    HIPCHK(exclusive_scan(ctx->d_bin_cnt, ctx->d_bin_start, nbins + 1, st));
    k_fill_bins<<<(N + 255) / 256, 256, 0, st>>>(N, ctx->d_bin_of, ctx->d_bin_start, bin_fill,
                                                ctx->d_bin_atoms);
    {
        Launch L(ctx, "nb_search");
        k_search_bins<<<nbins, kStageThreads, 0, st>>>(ctx->d_pos, dci, ctx->d_cells, ctx->d_bin_struc, ctx->d_shift,
                                                      ctx->d_bin_start, ctx->d_bin_atoms, rc * rc, ctx->d_keys,
                                                      ctx->d_cnt, ctx->d_flag, ctx->d_bin_ovf);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(exclusive_scan(ctx->d_cnt, ctx->nl.off, N + 1, st));
    int flags[4] = {0, 0, 0, 0};
    int total = 0;
    HIPCHK(hipMemcpyAsync(flags, ctx->d_flag, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&total, ctx->nl.off + N, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (flags[2]) { set_error("the previous neighbour list was not symmetric (reverse slot missing)"); return hipErrorInvalidValue; }
    if (flags[0] & 8) {
        // bins with a search box too large to stage: one wavefront per centre atom
        {
            Launch L(ctx, "nb_search_wide");
            k_search<<<N, 64, 0, st>>>(N, ctx->d_pos, ctx->d_struc_of, dci, ctx->d_cells, ctx->d_bin_of,
                                       ctx->d_shift, ctx->d_bin_start, ctx->d_bin_atoms, rc * rc, ctx->d_keys,
                                       ctx->d_cnt, ctx->d_flag, ctx->d_bin_ovf);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(exclusive_scan(ctx->d_cnt, ctx->nl.off, N + 1, st));
        HIPCHK(hipMemcpyAsync(flags, ctx->d_flag, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&total, ctx->nl.off + N, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    if (flags[0] & 2) { set_error("neighbour search box exceeds the image limit (cell far narrower than rc)"); return hipErrorInvalidValue; }
    if (flags[0] & 1) {
        set_error("an atom has " + std::to_string(flags[1]) + " neighbours; capacity is " + std::to_string(kMaxNeigh));
        return hipErrorInvalidValue;
    }
    NList& nl = ctx->nl;
    int64_t c1 = nl.cap, c2 = nl.cap, c3 = nl.cap, c4 = nl.cap, c5 = nl.cap;
    HIPCHK(grow(nl.idx, c1, total));
    HIPCHK(grow(nl.img, c2, total));
    HIPCHK(grow(nl.rev, c3, total));
    HIPCHK(grow(nl.geo, c4, total));
    HIPCHK(grow(nl.nsp, c5, total));
    nl.cap = std::min(std::min(c1, c2), std::min(std::min(c3, c4), c5));
    nl.npairs = total;
    nl.maxn = flags[1];
    {
        Launch L(ctx, "nb_compact");
        k_compact<<<N, 128, 0, st>>>(N, ctx->d_pos, ctx->d_struc_of, ctx->d_cells, ctx->d_species, ctx->d_keys,
                                     nl.off, nl.idx, nl.img, nl.geo, nl.nsp);
        // a missing reverse slot (impossible for the symmetric FP64 test) is
        // reported by the next synchronising call (ann_nlist_size, next build)
        k_rev<<<N, 128, 0, st>>>(N, ctx->d_keys, nl.off, nl.rev, ctx->d_flag + 2);
    }
    return hipGetLastError();
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
