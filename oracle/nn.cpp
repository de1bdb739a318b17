/*
 * nn.cpp -- oracle: per-species atomic network (FP64).  TEST INFRASTRUCTURE
 * ONLY (see oracle.h).
 *
 * Follows:
 *   activations a, da/dz, d2a/dz2   src/ml/nn.cpp:149-237 (linear, sigmoid,
 *                                   tanh), :238-266 (isru), :268-368 (arctan,
 *                                   relu, elu), :369-449 (tanhre, sqre),
 *                                   :451-498 (swish), :500-524 (gelu),
 *                                   :526-586 (mish), :588-616 (pflu),
 *                                   :618-643 (serf), :645-674 (logish),
 *                                   :676-750 (softplus, sqplus), :752-780
 *                                   (atish); softplus and logish use the
 *                                   reference's truncated ln2 constant
 *                                   (src/math/const.hpp:15).  Values and first
 *                                   derivatives are the reference's (serf, logish,
 *                                   atish and sqre ignore the sharpness c there,
 *                                   and so here).  Second derivatives are the
 *                                   true ones: the reference's are zero for
 *                                   gelu, mish, serf, logish and atish, carry an
 *                                   extra factor c for tanhre and have the
 *                                   wrong sign for sqre (:432-449)
 *   ANN::fp / fpbp / fpbp2          src/ml/nn.cpp:1343-1413
 *   DODZ::grad (dE/dG)              src/ml/nn.cpp:1801-1830
 *   Cost::grad (energy backprop)    src/ml/nn.cpp:1712-1776, packing :846-865
 *   second backward                 replaces D2ODZDI::grad (src/ml/nn.cpp:
 *                                   2012-2103) contracted with a G-space
 *                                   vector, as used by src/nnp/nnptef.cpp:
 *                                   1740-1815
 *   ADAM::step                      src/opt/algo.cpp:280-296 (beta1 0.9,
 *                                   beta2 0.999, :260-261)
 *
 * Parity pins (tests/test_oracle_nn.py): activation derivatives by central
 * differences; forward, dE/dG, dE/dp and the second backward against FP64
 * torch.autograd (double backward) and against central differences of
 * dE/dG . v with respect to the parameters; Adam against its closed form
 * after one step (p -= gamma * sign(g) for eps -> 0).
 */
#include "oracle.h"
#include "oracle_impl.h"
#include <cmath>
#include <stdexcept>
#include <vector>

namespace orc {

const double kLN2 = 0.693147180559945;  // truncated as in src/math/const.hpp:15
const double kRadPI = 1.772453850905516; // sqrt(pi), src/math/const.hpp:10

void activation(int neuron, double c, double z, double& a, double& d, double& d2) {
    switch (neuron) {
        case ORC_NEU_LINEAR: a = z; d = 1.0; d2 = 0.0; break;
        case ORC_NEU_SIGMOID: {
            if (z >= 0) {
                const double e = std::exp(-c * z), f = 1.0 / (1.0 + e);
                a = f; d = c * e * f * f; d2 = -c * c * e * (1.0 - e) * f * f * f;
            } else {
                const double e = std::exp(c * z), f = 1.0 / (1.0 + e);
                a = e * f; d = c * e * f * f; d2 = -c * c * e * (e - 1.0) * f * f * f;
            }
        } break;
        case ORC_NEU_TANH: {
            const double t = std::tanh(c * z);
            a = t; d = c * (1.0 - t * t); d2 = c * c * (-2.0 * t * (1.0 - t * t));
        } break;
        case ORC_NEU_ARCTAN: {
            const double zz = c * z * kPI / 2.0, f = 1.0 / (1.0 + zz * zz);
            // the reference's d2 (-2 c zz f^2, src/ml/nn.cpp:289) lacks the chain
            // factor c pi/2; the true second derivative is used here
            a = (2.0 / kPI) * std::atan(zz); d = c * f; d2 = -2.0 * c * zz * f * f * (c * kPI / 2.0);
        } break;
        case ORC_NEU_RELU:
            if (z > 0.0) { a = z; d = 1.0; } else { a = 0.0; d = 0.0; }
            d2 = 0.0;
            break;
        case ORC_NEU_ELU:
            if (z >= 0.0) { a = z; d = 1.0; d2 = 0.0; }
            else { const double e = std::exp(c * z); a = e - 1.0; d = c * e; d2 = c * c * e; }
            break;
        case ORC_NEU_SWISH: {
            const double zz = c * z;
            if (zz > 0.0) {
                const double e = std::exp(-zz), den = 1.0 / (1.0 + e);
                a = z * den; d = (1.0 + e * (1.0 + zz)) * den * den;
                d2 = c * e * (2.0 - zz + e * (2.0 + zz)) * den * den * den;
            } else {
                const double e = std::exp(zz), den = 1.0 / (1.0 + e);
                a = z * e * den; d = e * (1.0 + zz + e) * den * den;
                d2 = c * e * (2.0 + zz + e * (2.0 - zz)) * den * den * den;
            }
        } break;
        case ORC_NEU_SOFTPLUS: {
            const double ci = 1.0 / c;
            if (z >= 0.0) {
                const double e = std::exp(-c * z), den = 1.0 / (e + 1.0);
                a = z + ci * (std::log1p(e) - kLN2); d = den; d2 = c * e * den * den;
            } else {
                const double e = std::exp(c * z), den = 1.0 / (e + 1.0);
                a = ci * (std::log1p(e) - kLN2); d = e * den; d2 = c * e * den * den;
            }
        } break;
        case ORC_NEU_SQPLUS: {
            const double ci = 1.0 / c, zz = c * z, s = std::sqrt(1.0 + zz * zz);
            a = 0.5 * ci * (zz - 1.0 + s); d = 0.5 * (zz / s + 1.0); d2 = 0.5 * c / (s * s * s);
        } break;
        case ORC_NEU_ISRU: {
            const double zz = c * z, isr = 1.0 / std::sqrt(1.0 + zz * zz), isr3 = isr * isr * isr;
            a = zz * isr; d = c * isr3; d2 = -3.0 * c * c * zz * isr3 * isr * isr;
        } break;
        case ORC_NEU_TANHRE:
            if (z >= 0.0) { a = z; d = 1.0; d2 = 0.0; }
            else {
                const double e = std::exp(2.0 * c * z);
                a = (e - 1.0) / (e + 1.0); d = c * (1.0 - a * a); d2 = -2.0 * c * a * d;
            }
            break;
        case ORC_NEU_SQRE:
            if (z >= 0.0) { a = z; d = 1.0; d2 = 0.0; }
            else {
                const double den = 1.0 / std::sqrt(1.0 + z * z), den3 = den * den * den;
                a = z * den; d = den3; d2 = -3.0 * z * den3 * den * den;
            }
            break;
        case ORC_NEU_GELU: {
            const double zz = 0.5 * kRadPI * c * z, g = std::exp(-zz * zz), ep1 = 1.0 + std::erf(zz);
            a = 0.5 * z * ep1; d = 0.5 * (ep1 + c * z * g); d2 = c * g * (1.0 - zz * zz);
        } break;
        case ORC_NEU_MISH: {
            // a = z tanh(softplus(c z)); with f = tanh(softplus(u)), s = sigmoid(u):
            // f' = (1 - f^2) s, f'' = (1 - f^2) s ((1 - s) - 2 f s)
            const double zz = c * z;
            double f, s;
            if (zz > 0.0) {
                const double e = std::exp(-zz), e2 = e * e;
                f = 1.0 - 2.0 * e2 / (1.0 + 2.0 * (e + e2));
                s = 1.0 / (1.0 + e);
            } else {
                const double e = std::exp(zz), ff = e * (e + 2.0);
                f = ff / (ff + 2.0);
                s = e / (1.0 + e);
            }
            const double fp = (1.0 - f * f) * s, fpp = fp * ((1.0 - s) - 2.0 * f * s);
            a = z * f; d = f + zz * fp; d2 = 2.0 * c * fp + c * zz * fpp;
        } break;
        case ORC_NEU_PFLU: {
            const double zz = c * z, z2 = zz * zz, q = 1.0 / std::sqrt(z2 + 1.0), q3 = q * q * q;
            a = 0.5 * z * (1.0 + zz * q); d = 0.5 * (zz * (z2 + 2.0) * q3 + 1.0);
            d2 = -0.5 * c * (z2 - 2.0) * q3 * q * q;
        } break;
        case ORC_NEU_SERF: {
            // a = z erf(softplus(z)); d = erf(sp) + z g with g = 2/sqrt(pi) exp(-sp^2) sigmoid(z)
            const double sp = std::max(z, 0.0) + std::log1p(std::exp(-std::fabs(z)));
            const double s = z >= 0.0 ? 1.0 / (1.0 + std::exp(-z)) : std::exp(z) / (1.0 + std::exp(z));
            const double g = 2.0 / kRadPI * std::exp(-sp * sp) * s;
            a = z * std::erf(sp); d = std::erf(sp) + z * g; d2 = 2.0 * g + z * g * ((1.0 - s) - 2.0 * sp * s);
        } break;
        case ORC_NEU_LOGISH: {
            // a = z log(1 + sigmoid(z)) / ln2; h = s (1 - s) / (1 + s)
            const double s = z >= 0.0 ? 1.0 / (1.0 + std::exp(-z)) : std::exp(z) / (1.0 + std::exp(z));
            const double L = std::log1p(s), h = s * (1.0 - s) / (1.0 + s);
            const double hp = s * (1.0 - s) * ((1.0 - 2.0 * s) / (1.0 + s) - s * (1.0 - s) / ((1.0 + s) * (1.0 + s)));
            a = z * L / kLN2; d = (L + z * h) / kLN2; d2 = (2.0 * h + z * hp) / kLN2;
        } break;
        case ORC_NEU_ATISH: {
            const double zz = 0.5 * kPI * z, q = 1.0 / (1.0 + zz * zz);
            a = z * 0.5 * (1.0 + 2.0 / kPI * std::atan(zz)); d = 0.5 * (1.0 + z * q) + std::atan(zz) / kPI; d2 = q * q;
        } break;
        default: throw std::invalid_argument("orc: unsupported neuron");
    }
}

Net::Net(const orc_model* m, int sp) {
    const int S = m->nspecies;
    if (sp < 0 || sp >= S) throw std::invalid_argument("orc: species out of range");
    nin = ninput(m, sp);
    nl = m->nn_nlayer[sp];
    nodes.assign(m->nn_nodes + m->nn_nodes_off[sp], m->nn_nodes + m->nn_nodes_off[sp] + nl);
    neuron = m->nn_neuron[sp];
    c = m->nn_c[sp];
    const double* p = m->params + m->nn_par_off[sp];
    b.resize(nl);
    W.resize(nl);
    int cnt = 0;
    for (int l = 0; l < nl; ++l) { b[l] = p + cnt; cnt += nodes[l]; }
    for (int l = 0; l < nl; ++l) {
        W[l] = p + cnt;
        cnt += nodes[l] * (l == 0 ? nin : nodes[l - 1]);
    }
    nparam = cnt;
    inpw = m->inpw + m->nn_inp_off[sp];
    inpb = m->inpb + m->nn_inp_off[sp];
    outw = m->outw[sp];
    outb = m->outb[sp];
    if (nodes.back() != 1) throw std::invalid_argument("orc: network output must have size 1");
}

int Net::nprev(int l) const { return l == 0 ? nin : nodes[l - 1]; }

// ANN::fpbp2: scaled input, z, a, a', a'' per layer; last layer linear
double Net::forward(const double* G, State& st) const {
    st.s.resize(nin);
    for (int k = 0; k < nin; ++k) st.s[k] = inpw[k] * (G[k] + inpb[k]);
    st.z.resize(nl); st.a.resize(nl); st.d.resize(nl); st.d2.resize(nl);
    const double* x = st.s.data();
    for (int l = 0; l < nl; ++l) {
        const int no = nodes[l], ni = nprev(l);
        st.z[l].assign(b[l], b[l] + no);
        for (int m = 0; m < ni; ++m) {
            const double xm = x[m];
            const double* col = W[l] + (size_t)m * no;
            for (int n = 0; n < no; ++n) st.z[l][n] += col[n] * xm;
        }
        st.a[l].resize(no); st.d[l].resize(no); st.d2[l].resize(no);
        const int neu = (l == nl - 1) ? ORC_NEU_LINEAR : neuron;
        for (int n = 0; n < no; ++n) activation(neu, c, st.z[l][n], st.a[l][n], st.d[l][n], st.d2[l][n]);
        x = st.a[l].data();
    }
    return outb + outw * st.a[nl - 1][0];
}

// DODZ::grad: dE/dG = inpw . W0^T diag(a0') ... W_{L-1}^T (outw a'_{L-1})
void Net::dedg(const State& st, double* out) const {
    std::vector<double> g(1, outw * st.d[nl - 1][0]);
    for (int l = nl - 1; l >= 0; --l) {
        const int no = nodes[l], ni = nprev(l);
        std::vector<double> h(ni, 0.0);
        for (int m = 0; m < ni; ++m) {
            const double* col = W[l] + (size_t)m * no;
            double acc = 0;
            for (int n = 0; n < no; ++n) acc += col[n] * g[n];
            h[m] = acc;
        }
        if (l > 0) for (int m = 0; m < ni; ++m) h[m] *= st.d[l - 1][m];
        g.swap(h);
    }
    for (int k = 0; k < nin; ++k) out[k] = g[k] * inpw[k];
}

// Second backward.  Phi = gE * E + (dE/dG . v).  The tangent of the network
// along input direction v is propagated forward (zd_l = W_l ad_{l-1},
// ad_l = a'_l zd_l) and Phi is back-propagated through both the primal and
// the tangent network; returns dPhi/dp in the reference packing order.
// With v = 0 this is Cost::grad with dc/do = gE.
void Net::grad_rop(const State& st, const double* v, double gE, double* grad) const {
    std::vector<std::vector<double>> zd(nl), ad(nl);
    std::vector<double> sd(nin, 0.0);
    if (v) for (int k = 0; k < nin; ++k) sd[k] = inpw[k] * v[k];
    const double* xd = sd.data();
    for (int l = 0; l < nl; ++l) {
        const int no = nodes[l], ni = nprev(l);
        zd[l].assign(no, 0.0);
        for (int m = 0; m < ni; ++m) {
            const double* col = W[l] + (size_t)m * no;
            for (int n = 0; n < no; ++n) zd[l][n] += col[n] * xd[m];
        }
        ad[l].resize(no);
        for (int n = 0; n < no; ++n) ad[l][n] = st.d[l][n] * zd[l][n];
        xd = ad[l].data();
    }
    std::vector<double> abar(1, gE * outw), adbar(1, v ? outw : 0.0);
    // offsets of each layer's biases and weights in the packed vector
    std::vector<int> boff(nl), woff(nl);
    int cnt = 0;
    for (int l = 0; l < nl; ++l) { boff[l] = cnt; cnt += nodes[l]; }
    for (int l = 0; l < nl; ++l) { woff[l] = cnt; cnt += nodes[l] * nprev(l); }
    for (int l = nl - 1; l >= 0; --l) {
        const int no = nodes[l], ni = nprev(l);
        std::vector<double> zbar(no), zdbar(no);
        for (int n = 0; n < no; ++n) {
            zdbar[n] = adbar[n] * st.d[l][n];
            zbar[n] = abar[n] * st.d[l][n] + adbar[n] * st.d2[l][n] * zd[l][n];
        }
        const double* xp = (l == 0) ? st.s.data() : st.a[l - 1].data();
        const double* xdp = (l == 0) ? sd.data() : ad[l - 1].data();
        for (int n = 0; n < no; ++n) grad[boff[l] + n] = zbar[n];
        for (int m = 0; m < ni; ++m)
            for (int n = 0; n < no; ++n)
                grad[woff[l] + m * no + n] = zbar[n] * xp[m] + zdbar[n] * xdp[m];
        if (l > 0) {
            std::vector<double> na(ni, 0.0), nad(ni, 0.0);
            for (int m = 0; m < ni; ++m) {
                const double* col = W[l] + (size_t)m * no;
                double s1 = 0, s2 = 0;
                for (int n = 0; n < no; ++n) { s1 += col[n] * zbar[n]; s2 += col[n] * zdbar[n]; }
                na[m] = s1; nad[m] = s2;
            }
            abar.swap(na); adbar.swap(nad);
        }
    }
}

}  // namespace orc

using namespace orc;

extern "C" void orc_activation(int neuron, double c, double z, double* a, double* d, double* d2) {
    activation(neuron, c, z, *a, *d, *d2);
}

extern "C" double orc_nn_forward(const orc_model* m, int sp, const double* G, double* dEdG) {
    Net net(m, sp);
    Net::State st;
    const double e = net.forward(G, st);
    if (dEdG) net.dedg(st, dEdG);
    return e;
}

extern "C" void orc_nn_grad_energy(const orc_model* m, int sp, const double* G, double dcdo,
                                   double* grad) {
    Net net(m, sp);
    Net::State st;
    net.forward(G, st);
    net.grad_rop(st, nullptr, dcdo, grad);
}

extern "C" void orc_nn_grad_rop(const orc_model* m, int sp, const double* G, const double* v,
                                double gE, double* grad) {
    Net net(m, sp);
    Net::State st;
    net.forward(G, st);
    net.grad_rop(st, v, gE, grad);
}

// ADAM::step, src/opt/algo.cpp:280-296; t is the 1-based step index so that
// the bias corrections are 1/(1-beta^t) as with beta1i_/beta2i_.
extern "C" void orc_adam_step(int n, double* p, double* m1, double* m2, const double* g,
                              double gamma, double eps, int t) {
    const double b1 = 0.9, b2 = 0.999;
    const double b1ii = 1.0 / (1.0 - std::pow(b1, t));
    const double b2ii = 1.0 / (1.0 - std::pow(b2, t));
    for (int i = 0; i < n; ++i) {
        m1[i] = m1[i] * b1 + (1.0 - b1) * g[i];
        m2[i] = m2[i] * b2 + (1.0 - b2) * g[i] * g[i];
        p[i] -= gamma * b1ii * m1[i] / (std::sqrt(b2ii * m2[i]) + eps);
    }
}
