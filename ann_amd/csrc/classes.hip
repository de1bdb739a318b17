// classes.hip -- FP64 evaluators behind the reference's per-function classes:
// Cutoff::cutf/cutg/compute (src/nnp/cutoff.hpp:59-61), BasisR::symm/force
// (src/nnp/basis_radial.hpp:87-88), BasisA::symm/force
// (src/nnp/basis_angular.hpp:101-102), NN::ANN::fp/fpbp/fpbp2
// (src/ml/nn.hpp:357-362), DODZ::grad (:485), Cost::grad (:453) and
// D2ODZDI::grad (:588).
//
// The reference evaluates one distance, one triplet or one input vector per
// call; here a call takes npts of them (device arrays) and a launch covers
// them all: one thread per point for the bases, one wavefront per input row
// for the network (node values in LDS, lanes over the nodes of a layer).
// Everything is FP64.  These calls hold no context: parameters travel in the
// kernel arguments (bases) or in an ann_mlp handle (networks).  Synchronous.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "basis_dev.h"
#include "net_dev.h"

namespace ann {

constexpr int kClsMaxFunc = 64;      // functions per basis in one call
constexpr int kMlpMaxLayers = 16;
constexpr int kMlpMaxLds = 64 * 1024;

// ---------------------------------------------------------------- FP64 bases

// Cutoff::compute (src/nnp/cutoff.cpp:115-143): value and d/dr
__device__ inline void cut64_fg(int name, double rc, double r, double& f, double& g) {
    if (name == CUT_STEP) { f = r > rc ? 0.0 : 1.0; g = 0.0; return; }
    if (!(r < rc)) { f = 0.0; g = 0.0; return; }
    const double rci = 1.0 / rc;
    switch (name) {
        case CUT_COS: {
            const double k = 3.141592653589793 * rci, arg = r * k;
            f = 0.5 * (cos(arg) + 1.0);
            g = -0.5 * k * sin(arg);
        } break;
        case CUT_TANH: {
            const double t = tanh(1.0 - r * rci);
            f = t * t * t;
            g = -3.0 * t * t * (1.0 - t * t) * rci;
        } break;
        case CUT_POLY3: {
            const double x = r * rci;
            f = (2.0 * x - 3.0) * x * x + 1.0;
            g = 6.0 * x * (x - 1.0) * rci;
        } break;
        default: f = 1.0; g = 0.0; break;
    }
}

__device__ inline void sech_tanh64(double y, double& s, double& t) {
    const double e = exp(-fabs(y)), e2 = e * e, den = 1.0 / (1.0 + e2);
    s = 2.0 * e * den;
    t = copysign((1.0 - e2) * den, y);
}

// BasisR::symmf / symmd (src/nnp/basis_radial.cpp:175-257): value and d/dr of
// one function (fc, gc: cutoff value and derivative)
__device__ inline void radial64_fd(int name, double eta, double rs, double r, double fc, double gc, double& v,
                                   double& d) {
    const double x = r - rs;
    switch (name) {
        case RAD_GAUSSIAN: {
            const double a = eta * x, e = exp(-a * a);
            v = e * fc;
            d = e * (-2.0 * eta * a * fc + gc);
        } break;
        case RAD_SECH: {
            double s, t;
            sech_tanh64(eta * x, s, t);
            v = s * fc;
            d = s * (-eta * t * fc + gc);
        } break;
        case RAD_LOGISTIC: {
            double s, t;
            sech_tanh64(eta * x, s, t);
            v = s * s * fc;
            d = s * s * (-2.0 * eta * t * fc + gc);
        } break;
        case RAD_TANH: {
            const double t = tanh(-eta * x);
            v = 0.5 * (t + 1.0) * fc;
            d = 0.5 * (-eta * (1.0 - t * t) * fc + (1.0 + t) * gc);
        } break;
        case RAD_LOGCOSH: {
            const double y = -2.0 * eta * x;
            const double lp = fmax(y, 0.0) + log1p(exp(-fabs(y)));
            const double sg = 1.0 / (1.0 + exp(-y));
            v = 0.5 * lp * fc;
            d = -eta * sg * fc + 0.5 * lp * gc;
        } break;
        case RAD_LOGCOSH2: {
            const double k2 = 1.4142135623730951, a = -eta * x;
            if (a == 0.0) {
                // the reference's value at exactly a == 0 (src/nnp/basis_radial.cpp:250,
                // :367), kept as is: it is not the a -> 0 limit of the branch below
                v = 0.5 * fc / k2;
                d = 0.5 * (-eta * gc + fc / k2);
            } else {
                // e - 1 - y by its series where the subtraction would cancel
                const double y = 2.0 * k2 * a;
                const double em1 = expm1(y), e = em1 + 1.0, den = 1.0 / em1;
                const double c = fabs(y) < 1e-2
                                     ? y * y * (0.5 + y * (1.0 / 6.0 + y * (1.0 / 24.0 + y * (1.0 / 120.0 +
                                                          y * (1.0 / 720.0 + y * (1.0 / 5040.0))))))
                                     : em1 - y;
                v = a * e * den * fc;
                d = e * (-eta * c * den * fc + a * gc) * den;
            }
        } break;
        default: v = 0.0; d = 0.0; break;
    }
}

// BasisA::symmf / symmd (src/nnp/basis_angular.cpp:212-364): value and
// d/dcos, d/dr0, d/dr1 of one function of a triplet
__device__ inline void angular64_fd(int name, double ieta2, double zeta, double lamh, double r0, double r1, double c,
                                    double f0, double f1, double g0, double g1, double& G, double& dc, double& d0,
                                    double& d1) {
    const double r2s = r0 * r0 + r1 * r1, hp = 1.5707963267948966;
    double fr, dfr;
    switch (name) {
        case ANG_GAUSS: {
            const double e = exp(-ieta2 * r2s);
            fr = e;
            dfr = -ieta2 * e;
        } break;
        case ANG_GAUSS2: {
            const double e = exp(-hp * ieta2 * r2s), den = 1.0 / (1.0 + e);
            fr = 2.0 * e * den;
            dfr = fr * (-hp * ieta2 * den);
        } break;
        case ANG_SECH: {
            const double e = exp(-hp * ieta2 * r2s), e2 = e * e, den = 1.0 / (1.0 + e2);
            fr = 2.0 * e * den;
            dfr = fr * (-hp * ieta2 * (1.0 - e2) * den);
        } break;
        default: {
            const int p = name == ANG_STUDENT3 ? 4 : (name == ANG_STUDENT4 ? 5 : 6);
            const double den = 1.0 / sqrt(1.0 + ieta2 * r2s);
            double w = 1.0;
            for (int q = 0; q < p; ++q) w *= den;
            fr = w;
            dfr = w * (-0.5 * p * ieta2 * den * den);
        } break;
    }
    const double cw = fabs(0.5 + lamh * c);
    const double wz1 = pow(cw, zeta - 1.0), wz = wz1 * cw;
    G = wz * fr * f0 * f1;
    dc = zeta * lamh * wz1 * fr * f0 * f1;
    d0 = wz * (2.0 * r0 * dfr * f0 + fr * g0) * f1;
    d1 = wz * (2.0 * r1 * dfr * f1 + fr * g1) * f0;
}

struct CutArgs {
    int name;
    double rc;
};

__global__ void k_cutoff(CutArgs A, int npts, const double* __restrict__ r, double* __restrict__ f,
                         double* __restrict__ g) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    double fv, gv;
    cut64_fg(A.name, A.rc, r[p], fv, gv);
    if (f) f[p] = fv;
    if (g) g[p] = gv;
}

struct RadArgs {
    int cut, name, n;
    double rc;
    double rs[kClsMaxFunc], eta[kClsMaxFunc];
};

// one thread per distance: G [npts][n], dG/dr [npts][n], amp = -sum dEdG dG/dr
__global__ void k_radial(RadArgs A, int npts, const double* __restrict__ r, double* __restrict__ G,
                         double* __restrict__ dG, const double* __restrict__ dEdG, double* __restrict__ amp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const double rr = r[p];
    double fc, gc;
    cut64_fg(A.cut, A.rc, rr, fc, gc);
    double a = 0.0;
    for (int i = 0; i < A.n; ++i) {
        double v, d;
        radial64_fd(A.name, A.eta[i], A.rs[i], rr, fc, gc, v, d);
        if (G) G[(size_t)p * A.n + i] = v;
        if (dG) dG[(size_t)p * A.n + i] = d;
        if (dEdG) a -= dEdG[(size_t)p * A.n + i] * d;
    }
    if (amp) amp[p] = a;
}

struct AngArgs {
    int cut, name, n;
    double rc;
    double ieta2[kClsMaxFunc], zeta[kClsMaxFunc], lamh[kClsMaxFunc];
};

// one thread per triplet (legs r[p][0], r[p][1], cosine cs[p]):
// G [npts][n], dG [npts][n][3] = (d/dcos, d/dr0, d/dr1),
// phi = -sum dEdG dG/dcos, eta = -sum dEdG (dG/dr0, dG/dr1)
__global__ void k_angular(AngArgs A, int npts, const double* __restrict__ r, const double* __restrict__ cs,
                          double* __restrict__ G, double* __restrict__ dG, const double* __restrict__ dEdG,
                          double* __restrict__ phi, double* __restrict__ eta) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const double r0 = r[2 * p], r1 = r[2 * p + 1], c = cs[p];
    double f0, g0, f1, g1;
    cut64_fg(A.cut, A.rc, r0, f0, g0);
    cut64_fg(A.cut, A.rc, r1, f1, g1);
    double ph = 0.0, e0 = 0.0, e1 = 0.0;
    for (int i = 0; i < A.n; ++i) {
        double v, dc, d0, d1;
        angular64_fd(A.name, A.ieta2[i], A.zeta[i], A.lamh[i], r0, r1, c, f0, f1, g0, g1, v, dc, d0, d1);
        const size_t q = (size_t)p * A.n + i;
        if (G) G[q] = v;
        if (dG) { dG[3 * q] = dc; dG[3 * q + 1] = d0; dG[3 * q + 2] = d1; }
        if (dEdG) {
            const double w = dEdG[q];
            ph -= w * dc;
            e0 -= w * d0;
            e1 -= w * d1;
        }
    }
    if (phi) phi[p] = ph;
    if (eta) { eta[2 * p] = e0; eta[2 * p + 1] = e1; }
}

// ---------------------------------------------------------------- FP64 network

struct MlpArgs {
    int nin, nl, neuron, nparam;
    double c, outw, outb;
    int nodes[kMlpMaxLayers], boff[kMlpMaxLayers], woff[kMlpMaxLayers], loff[kMlpMaxLayers];
    int maxw, back_off;            // widest layer (inputs included); LDS offset of the backward buffers
    const double* P;               // parameters, reference packing (biases, then column-major weights)
    const double* inpw;
    const double* inpb;
};

// Backward through the layers from (abar, adbar) at the output: bias and
// weight gradients of Phi = gE o + sum_k v_k do/di_k into grow (if set), and
// the input adjoint abar_s into the returned LDS row.  Lanes over nodes.
__device__ double* mlp_backward(const MlpArgs& A, double* lds, double gE, bool tangent, double* __restrict__ grow) {
    const int lane = threadIdx.x;
    const int W = A.maxw;
    double* ab = lds + A.back_off;          // abar of the current layer's outputs
    double* adb = ab + W;                   // adbar
    double* zb = adb + W;                   // zbar
    double* zdb = zb + W;                   // zdbar
    double* s = lds;                        // scaled input
    double* sd = lds + A.nin;               // its tangent
    if (lane == 0) { ab[0] = gE * A.outw; adb[0] = tangent ? A.outw : 0.0; }
    __syncthreads();
    for (int l = A.nl - 1; l >= 0; --l) {
        const int no = A.nodes[l], ni = l ? A.nodes[l - 1] : A.nin;
        const double* L = lds + A.loff[l];       // a, d, d2, zd, ad of layer l
        for (int n = lane; n < no; n += kWave) {
            const double d = L[no + n], d2 = L[2 * no + n], zd = L[3 * no + n];
            zdb[n] = adb[n] * d;
            zb[n] = ab[n] * d + adb[n] * d2 * zd;
        }
        __syncthreads();
        const double* x = l ? lds + A.loff[l - 1] : s;                   // a of the layer below
        const double* xd = l ? lds + A.loff[l - 1] + 4 * ni : sd;        // ad of the layer below
        const double* Wl = A.P + A.woff[l];
        if (grow) {
            for (int n = lane; n < no; n += kWave) grow[A.boff[l] + n] = zb[n];
            for (int q = lane; q < no * ni; q += kWave) {
                const int m = q / no, n = q - m * no;
                grow[A.woff[l] + q] = zb[n] * x[m] + (tangent ? zdb[n] * xd[m] : 0.0);
            }
        }
        // adjoints of the layer below: W^T zbar, W^T zdbar
        double* nab = zdb + W;
        double* nadb = nab + W;
        for (int m = lane; m < ni; m += kWave) {
            const double* col = Wl + (size_t)m * no;
            double s1 = 0.0, s2 = 0.0;
            for (int n = 0; n < no; ++n) { s1 += col[n] * zb[n]; s2 += col[n] * zdb[n]; }
            nab[m] = s1;
            nadb[m] = s2;
        }
        __syncthreads();
        for (int m = lane; m < ni; m += kWave) { ab[m] = nab[m]; adb[m] = nadb[m]; }
        __syncthreads();
    }
    return ab;
}

// One wavefront per input row: ANN::fpbp2 (src/ml/nn.cpp:1391-1412) with the
// tangent along v, the output, DODZ::dodi (src/ml/nn.cpp:1801-1810) and the
// parameter gradient of gE o + sum_k v_k do/di_k (Cost::grad with dc/do = gE,
// :1712-1762; with v the rows of D2ODZDI::d2odpdi, :2012-2100).
__global__ __launch_bounds__(64) void k_mlp(MlpArgs A, int npts, const double* __restrict__ in,
                                             const double* __restrict__ v, const double* __restrict__ dcdo,
                                             double* __restrict__ out, double* __restrict__ dodi,
                                             double* __restrict__ grad) {
    extern __shared__ double lds[];
    const int row = blockIdx.x;
    if (row >= npts) return;
    const int lane = threadIdx.x;
    double* s = lds;
    double* sd = lds + A.nin;
    const bool tangent = v != nullptr;
    for (int k = lane; k < A.nin; k += kWave) {
        const double w = A.inpw[k];
        s[k] = w * (in[(size_t)row * A.nin + k] + A.inpb[k]);
        sd[k] = tangent ? w * v[(size_t)row * A.nin + k] : 0.0;
    }
    __syncthreads();
    const double* x = s;
    const double* xd = sd;
    int ni = A.nin;
    for (int l = 0; l < A.nl; ++l) {
        const int no = A.nodes[l];
        double* L = lds + A.loff[l];
        const double* Wl = A.P + A.woff[l];
        const bool last = l == A.nl - 1;                 // the output layer is linear
        for (int n = lane; n < no; n += kWave) {
            double z = A.P[A.boff[l] + n], zd = 0.0;
            for (int m = 0; m < ni; ++m) {
                const double w = Wl[(size_t)m * no + n];
                z += w * x[m];
                zd += w * xd[m];
            }
            double a, d, d2;
            if (last) { a = z; d = 1.0; d2 = 0.0; }
            else act_f<true>(A.neuron, A.c, z, a, d, d2);
            L[n] = a;
            L[no + n] = d;
            L[2 * no + n] = d2;
            L[3 * no + n] = zd;
            L[4 * no + n] = d * zd;
        }
        __syncthreads();
        x = L;
        xd = L + 4 * no;
        ni = no;
    }
    if (lane == 0 && out) out[row] = A.outb + A.outw * x[0];
    if (dodi) {
        const double* ab = mlp_backward(A, lds, 1.0, false, nullptr);
        for (int k = lane; k < A.nin; k += kWave) dodi[(size_t)row * A.nin + k] = A.inpw[k] * ab[k];
        __syncthreads();
    }
    if (grad) mlp_backward(A, lds, dcdo ? dcdo[row] : 0.0, tangent, grad + (size_t)row * A.nparam);
}

// grad[p] (+)= sum over rows of rows[r][p], rows in order (deterministic)
__global__ void k_rowsum(int nrows, int nparam, const double* __restrict__ rows, double* __restrict__ grad,
                         int accumulate) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nparam) return;
    double s = accumulate ? grad[p] : 0.0;
    for (int r = 0; r < nrows; ++r) s += rows[(size_t)r * nparam + p];
    grad[p] = s;
}

}  // namespace ann

using namespace ann;

struct ann_mlp {
    MlpArgs A{};
    int lds_bytes = 0;
    double* d_par = nullptr;       // params, inpw, inpb (one allocation)
    double* ws = nullptr;          // per-row gradients of a summed call
    size_t ws_rows = 0;
    hipStream_t stream = nullptr;
    int device = 0;
};

namespace {

int check_points(int npts, const void* a) {
    if (npts < 0) return fail(ANN_ERR_ARG, "negative point count");
    if (npts > 0 && !a) return fail(ANN_ERR_ARG, "null input array");
    return ANN_OK;
}

int finish(const char* what) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return fail(ANN_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return ANN_OK;
}

int check_cutoff(int cutoff, double rc) {
    if (cutoff < CUT_STEP || cutoff > CUT_POLY3) return fail(ANN_ERR_ARG, "invalid cutoff function");
    if (!(rc > 0) || std::isinf(rc)) return fail(ANN_ERR_ARG, "invalid cutoff radius");
    return ANN_OK;
}

}  // namespace

extern "C" int ann_cutoff_eval(int cutoff, double rc, int npts, const double* d_r, double* d_f, double* d_g) {
    int rcode = check_cutoff(cutoff, rc);
    if (rcode != ANN_OK) return rcode;
    if ((rcode = check_points(npts, d_r)) != ANN_OK) return rcode;
    if (npts == 0) return ANN_OK;
    k_cutoff<<<(npts + 255) / 256, 256>>>(CutArgs{cutoff, rc}, npts, d_r, d_f, d_g);
    return finish("ann_cutoff_eval");
}

extern "C" int ann_radial_eval(const ann_radial_desc* b, int npts, const double* d_r, double* d_G, double* d_dGdr,
                               const double* d_dEdG, double* d_amp) {
    if (!b) return fail(ANN_ERR_ARG, "null basis");
    int rcode = check_cutoff(b->cutoff, b->rc);
    if (rcode != ANN_OK) return rcode;
    if (b->name < RAD_GAUSSIAN || b->name > RAD_LOGCOSH2) return fail(ANN_ERR_ARG, "invalid radial function");
    if (b->n < 0 || b->n > kClsMaxFunc)
        return fail(ANN_ERR_UNSUPPORTED, "radial basis larger than " + std::to_string(kClsMaxFunc) + " functions");
    if (b->n > 0 && (!b->rs || !b->eta)) return fail(ANN_ERR_ARG, "null basis parameters");
    if ((rcode = check_points(npts, d_r)) != ANN_OK) return rcode;
    if (d_amp && !d_dEdG) return fail(ANN_ERR_ARG, "the force amplitude needs dE/dG");
    RadArgs A{};
    A.cut = b->cutoff;
    A.name = b->name;
    A.n = b->n;
    A.rc = b->rc;
    for (int i = 0; i < b->n; ++i) { A.rs[i] = b->rs[i]; A.eta[i] = b->eta[i]; }
    if (npts == 0) return ANN_OK;
    k_radial<<<(npts + 255) / 256, 256>>>(A, npts, d_r, d_G, d_dGdr, d_dEdG, d_amp);
    return finish("ann_radial_eval");
}

extern "C" int ann_angular_eval(const ann_angular_desc* b, int npts, const double* d_r, const double* d_cos,
                                double* d_G, double* d_dG, const double* d_dEdG, double* d_phi, double* d_eta) {
    if (!b) return fail(ANN_ERR_ARG, "null basis");
    int rcode = check_cutoff(b->cutoff, b->rc);
    if (rcode != ANN_OK) return rcode;
    if (b->name < ANG_GAUSS || b->name > ANG_STUDENT5) return fail(ANN_ERR_ARG, "invalid angular function");
    if (b->n < 0 || b->n > kClsMaxFunc)
        return fail(ANN_ERR_UNSUPPORTED, "angular basis larger than " + std::to_string(kClsMaxFunc) + " functions");
    if (b->n > 0 && (!b->eta || !b->zeta || !b->lambda)) return fail(ANN_ERR_ARG, "null basis parameters");
    if ((rcode = check_points(npts, d_r)) != ANN_OK) return rcode;
    if (npts > 0 && !d_cos) return fail(ANN_ERR_ARG, "null cosine array");
    if ((d_phi || d_eta) && !d_dEdG) return fail(ANN_ERR_ARG, "the force terms need dE/dG");
    AngArgs A{};
    A.cut = b->cutoff;
    A.name = b->name;
    A.n = b->n;
    A.rc = b->rc;
    for (int i = 0; i < b->n; ++i) {
        if (!(b->eta[i] > 0)) return fail(ANN_ERR_ARG, "angular width must be positive");
        A.ieta2[i] = 1.0 / (b->eta[i] * b->eta[i]);
        A.zeta[i] = b->zeta[i];
        A.lamh[i] = 0.5 * b->lambda[i];
    }
    if (npts == 0) return ANN_OK;
    k_angular<<<(npts + 255) / 256, 256>>>(A, npts, d_r, d_cos, d_G, d_dG, d_dEdG, d_phi, d_eta);
    return finish("ann_angular_eval");
}

extern "C" int ann_mlp_create(const ann_mlp_desc* d, ann_mlp** out) {
    if (!d || !out) return fail(ANN_ERR_ARG, "null argument");
    *out = nullptr;
    if (d->nin < 1) return fail(ANN_ERR_ARG, "network needs at least one input");
    if (d->nlayer < 1 || d->nlayer > kMlpMaxLayers)
        return fail(ANN_ERR_UNSUPPORTED, "network layers must be 1.." + std::to_string(kMlpMaxLayers));
    if (!d->nodes) return fail(ANN_ERR_ARG, "null layer sizes");
    if (d->nodes[d->nlayer - 1] != 1) return fail(ANN_ERR_UNSUPPORTED, "only single-output networks");
    if (!neuron_supported(d->neuron)) return fail(ANN_ERR_ARG, "ANN::read(FILE*,ANN&): Invalid neuron.");
    ann_mlp h;
    MlpArgs& A = h.A;
    A.nin = d->nin;
    A.nl = d->nlayer;
    A.neuron = d->neuron;
    A.c = d->c;
    A.outw = d->outw;
    A.outb = d->outb;
    int cnt = 0, prev = d->nin, maxw = d->nin;
    for (int l = 0; l < d->nlayer; ++l) {
        if (d->nodes[l] < 1) return fail(ANN_ERR_ARG, "layer size must be positive");
        A.nodes[l] = d->nodes[l];
        A.boff[l] = cnt;
        cnt += d->nodes[l];
        maxw = std::max(maxw, d->nodes[l]);
    }
    for (int l = 0; l < d->nlayer; ++l) {
        A.woff[l] = cnt;
        cnt += d->nodes[l] * prev;
        prev = d->nodes[l];
    }
    A.nparam = cnt;
    A.maxw = maxw;
    int off = 2 * d->nin;
    for (int l = 0; l < d->nlayer; ++l) { A.loff[l] = off; off += 5 * d->nodes[l]; }
    A.back_off = off;
    off += 6 * maxw;
    h.lds_bytes = off * (int)sizeof(double);
    if (h.lds_bytes > kMlpMaxLds)
        return fail(ANN_ERR_UNSUPPORTED, "network too wide for the FP64 evaluator (" + std::to_string(h.lds_bytes) +
                                             " bytes of node state; limit " + std::to_string(kMlpMaxLds) + ")");
    ANN_HIP(hipGetDevice(&h.device));
    ANN_HIP(hipMalloc(&h.d_par, sizeof(double) * (size_t)(cnt + 2 * d->nin)));
    ANN_HIP(hipStreamCreateWithFlags(&h.stream, hipStreamNonBlocking));
    A.P = h.d_par;
    A.inpw = h.d_par + cnt;
    A.inpb = h.d_par + cnt + d->nin;
    ann_mlp* p = new ann_mlp(h);
    int r = ann_mlp_set_params(p, d->params, cnt);
    if (r == ANN_OK) r = ann_mlp_set_scaling(p, d->inpw, d->inpb, d->outw, d->outb);
    if (r != ANN_OK) { ann_mlp_destroy(p); return r; }
    *out = p;
    return ANN_OK;
}

extern "C" int ann_mlp_set_params(ann_mlp* h, const double* params, int n) {
    if (!h) return fail(ANN_ERR_ARG, "null network");
    if (n != h->A.nparam) return fail(ANN_ERR_ARG, "parameter count mismatch");
    if (!params) return fail(ANN_ERR_ARG, "null parameters");
    ANN_HIP(hipMemcpyAsync(h->d_par, params, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    ANN_HIP(hipStreamSynchronize(h->stream));
    return ANN_OK;
}

extern "C" int ann_mlp_set_scaling(ann_mlp* h, const double* inpw, const double* inpb, double outw, double outb) {
    if (!h) return fail(ANN_ERR_ARG, "null network");
    const int nin = h->A.nin;
    std::vector<double> s(2 * nin);
    for (int k = 0; k < nin; ++k) {
        s[k] = inpw ? inpw[k] : 1.0;
        s[nin + k] = inpb ? inpb[k] : 0.0;
        if (!std::isfinite(s[k]) || !std::isfinite(s[nin + k])) return fail(ANN_ERR_ARG, "non-finite input scaling");
    }
    if (!std::isfinite(outw) || !std::isfinite(outb)) return fail(ANN_ERR_ARG, "non-finite output scaling");
    h->A.outw = outw;
    h->A.outb = outb;
    ANN_HIP(hipMemcpyAsync(h->d_par + h->A.nparam, s.data(), sizeof(double) * 2 * nin, hipMemcpyHostToDevice,
                           h->stream));
    ANN_HIP(hipStreamSynchronize(h->stream));
    return ANN_OK;
}

extern "C" int ann_mlp_size(const ann_mlp* h) { return h ? h->A.nparam : 0; }

extern "C" int ann_mlp_eval(ann_mlp* h, int npts, const double* d_in, double* d_out, double* d_dodi) {
    if (!h) return fail(ANN_ERR_ARG, "null network");
    int r = check_points(npts, d_in);
    if (r != ANN_OK) return r;
    if (npts == 0) return ANN_OK;
    k_mlp<<<npts, kWave, h->lds_bytes, h->stream>>>(h->A, npts, d_in, nullptr, nullptr, d_out, d_dodi, nullptr);
    ANN_HIP(hipGetLastError());
    ANN_HIP(hipStreamSynchronize(h->stream));
    return ANN_OK;
}

extern "C" int ann_mlp_grad(ann_mlp* h, int npts, const double* d_in, const double* d_dcdo, const double* d_v,
                            int per_row, double* d_grad) {
    if (!h) return fail(ANN_ERR_ARG, "null network");
    int r = check_points(npts, d_in);
    if (r != ANN_OK) return r;
    if (!d_grad) return fail(ANN_ERR_ARG, "null gradient");
    const int np = h->A.nparam;
    if (per_row) {
        if (npts == 0) return ANN_OK;
        k_mlp<<<npts, kWave, h->lds_bytes, h->stream>>>(h->A, npts, d_in, d_v, d_dcdo, nullptr, nullptr, d_grad);
        ANN_HIP(hipGetLastError());
        ANN_HIP(hipStreamSynchronize(h->stream));
        return ANN_OK;
    }
    // summed over rows: per-row gradients of a chunk into the workspace, then
    // added in row order (the same sum for any chunking)
    const size_t cap = std::max<size_t>(1, (size_t)(64u << 20) / (sizeof(double) * (size_t)np));
    const size_t chunk = std::min<size_t>(cap, (size_t)std::max(npts, 1));
    if (chunk > h->ws_rows) {
        if (h->ws) ANN_HIP(hipFree(h->ws));
        h->ws = nullptr;
        h->ws_rows = 0;
        ANN_HIP(hipMalloc(&h->ws, sizeof(double) * np * chunk));
        h->ws_rows = chunk;
    }
    if (npts == 0) {
        k_rowsum<<<(np + 255) / 256, 256, 0, h->stream>>>(0, np, h->ws, d_grad, 0);
    }
    for (int r0 = 0; r0 < npts; r0 += (int)chunk) {
        const int nr = std::min<int>((int)chunk, npts - r0);
        k_mlp<<<nr, kWave, h->lds_bytes, h->stream>>>(h->A, nr, d_in + (size_t)r0 * h->A.nin,
                                                      d_v ? d_v + (size_t)r0 * h->A.nin : nullptr,
                                                      d_dcdo ? d_dcdo + r0 : nullptr, nullptr, nullptr, h->ws);
        k_rowsum<<<(np + 255) / 256, 256, 0, h->stream>>>(nr, np, h->ws, d_grad, r0 > 0);
    }
    ANN_HIP(hipGetLastError());
    ANN_HIP(hipStreamSynchronize(h->stream));
    return ANN_OK;
}

extern "C" void ann_mlp_destroy(ann_mlp* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->d_par) (void)hipFree(h->d_par);
    if (h->ws) (void)hipFree(h->ws);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}
