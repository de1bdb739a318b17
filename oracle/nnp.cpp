/*
 * nnp.cpp -- oracle: neighbour list, symmetry functions, energy, force
 * adjoint, virial, basis tangent and the energy+force training gradient
 * (FP64).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows:
 *   NNH::init_input (G layout)   src/nnp/nnh.cpp:94-112, TensorTri index
 *                                src/mem/tensortri.hpp:196-201
 *   NeighborList (source absent) contract from src/nnp/nnp.cpp:343-348,
 *                                478-482 and README.md:39: every periodic image
 *                                with 0 < |r_i - r_j - R n| < rc.  The order is
 *                                defined here (j ascending, then n
 *                                lexicographic) because the reference's own
 *                                order cannot be read: parity of ORDER is
 *                                unpinned, parity of the SET is pinned by the
 *                                brute-force / cell-list agreement and by the
 *                                supercell invariance tests.
 *   NNP::symm                    src/nnp/nnp.cpp:333-375
 *   NNP::energy / NNH::energy    src/nnp/nnp.cpp:436-447, src/nnp/nnh.cpp:118-121
 *   NNP::force / NNP::compute    src/nnp/nnp.cpp:461-520, 601-665 (mode 1 keeps
 *                                the reference's min() gating of reactions;
 *                                mode 0 folds every image reaction back, the
 *                                exact adjoint)
 *   thread::Dist block split     src/thread/dist.cpp:29-41 (orc_compute_batch)
 *   energy loss                  src/nnp/nnptefr2.cpp:1321-1373: MSE, MAE, HUBER,
 *                                ASINH on dU = (E - Eref) norm, Norm IDENTITY /
 *                                LINEAR / SQRT / CBRT / LOG (:602-612), per-
 *                                structure weights (energy-only structures of
 *                                the Taylor augmentation, :2466-2473)
 *
 * Parity pins (tests/test_oracle_nnp.py): neighbour sets brute force = cell
 * list, reverse-slot involution; G supercell invariance on the NaCl fixtures
 * of src/nnp/test_nnp_symm.cpp and a random triclinic cell; closed-form G of
 * two- and three-atom clusters; forces = -dE/dr by central differences
 * (src/nnp/test_nnp_force.cpp pattern, eps 1e-6); virial = -dE/dstrain by
 * central differences; sum_i F_i = 0; tangent = directional derivative of G
 * by central differences; training gradient by central differences (h=1e-4),
 * for every loss variant, normalisation and per-structure weights; HUBER and
 * ASINH tend to MSE for large delta.
 */
#include "oracle.h"
#include "oracle_impl.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>
#include <tuple>
#include <vector>

namespace orc {

int ninput(const orc_model* m, int a) {
    const int S = m->nspecies, NP = npairs(S);
    int n = 0;
    for (int b = 0; b < S; ++b) n += m->rad_n[a * S + b];
    for (int p = 0; p < NP; ++p) n += m->ang_n[a * NP + p];
    return n;
}
int offset_r(const orc_model* m, int a, int b) {
    const int S = m->nspecies;
    int o = 0;
    for (int q = 0; q < b; ++q) o += m->rad_n[a * S + q];
    return o;
}
int offset_a(const orc_model* m, int a, int p) {
    const int S = m->nspecies, NP = npairs(S);
    int o = offset_r(m, a, S);
    for (int q = 0; q < p; ++q) o += m->ang_n[a * NP + q];
    return o;
}

static void inv3(const double* R, double* Ri) {
    const double a = R[0], b = R[1], c = R[2], d = R[3], e = R[4], f = R[5], g = R[6], h = R[7],
                 i = R[8];
    const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    if (det == 0.0) throw std::invalid_argument("orc: singular cell");
    const double id = 1.0 / det;
    Ri[0] = A * id; Ri[1] = -(b * i - c * h) * id; Ri[2] = (b * f - c * e) * id;
    Ri[3] = B * id; Ri[4] = (a * i - c * g) * id;  Ri[5] = -(a * f - c * d) * id;
    Ri[6] = C * id; Ri[7] = -(a * h - b * g) * id; Ri[8] = (a * e - b * d) * id;
}

// perpendicular width of the cell along lattice direction a: 1/|row_a(R^-1)|
static void widths(const double* Ri, double* w) {
    for (int a = 0; a < 3; ++a)
        w[a] = 1.0 / std::sqrt(Ri[3 * a] * Ri[3 * a] + Ri[3 * a + 1] * Ri[3 * a + 1] +
                               Ri[3 * a + 2] * Ri[3 * a + 2]);
}

// r_i - r_j - R n with a fixed FP64 operation order; the GPU neighbour kernel
// evaluates the identical expression so that inclusion decisions agree bit for bit.
static inline void pair_vec(const double* pi, const double* pj, const double* R, int n0, int n1,
                            int n2, double* v, double* d2) {
    const double m0 = n0, m1 = n1, m2 = n2;
    v[0] = (pi[0] - pj[0]) - ((R[0] * m0 + R[1] * m1) + R[2] * m2);
    v[1] = (pi[1] - pj[1]) - ((R[3] * m0 + R[4] * m1) + R[5] * m2);
    v[2] = (pi[2] - pj[2]) - ((R[6] * m0 + R[7] * m1) + R[8] * m2);
    *d2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
}

struct Cand { int j, n0, n1, n2; };
static bool cand_less(const Cand& a, const Cand& b) {
    return std::tie(a.j, a.n0, a.n1, a.n2) < std::tie(b.j, b.n0, b.n1, b.n2);
}

static void finish_list(const orc_struc* s, NList& L, std::vector<std::vector<Cand>>& per) {
    const int N = s->natoms;
    L.natoms = N;
    L.off.assign(N + 1, 0);
    for (int i = 0; i < N; ++i) {
        std::sort(per[i].begin(), per[i].end(), cand_less);
        L.off[i + 1] = L.off[i] + (int)per[i].size();
    }
    const int P = L.off[N];
    L.idx.resize(P); L.img.resize(3 * (size_t)P); L.r.resize(3 * (size_t)P); L.d.resize(P);
    L.rev.resize(P); L.minimg.assign(P, 0);
    for (int i = 0; i < N; ++i) {
        for (size_t t = 0; t < per[i].size(); ++t) {
            const int q = L.off[i] + (int)t;
            const Cand& c = per[i][t];
            L.idx[q] = c.j;
            L.img[3 * q] = c.n0; L.img[3 * q + 1] = c.n1; L.img[3 * q + 2] = c.n2;
            double d2;
            pair_vec(s->pos + 3 * i, s->pos + 3 * c.j, s->cell, c.n0, c.n1, c.n2, &L.r[3 * q], &d2);
            L.d[q] = std::sqrt(d2);
        }
        // minimum-image flag: shortest entry per partner j (first on ties)
        for (int q = L.off[i]; q < L.off[i + 1];) {
            int e = q, best = q;
            while (e < L.off[i + 1] && L.idx[e] == L.idx[q]) { if (L.d[e] < L.d[best]) best = e; ++e; }
            L.minimg[best] = 1;
            q = e;
        }
    }
    // reverse slot: (i -> j, n) <-> (j -> i, -n)
    for (int i = 0; i < N; ++i) {
        for (int q = L.off[i]; q < L.off[i + 1]; ++q) {
            const int j = L.idx[q];
            const Cand key{i, -L.img[3 * q], -L.img[3 * q + 1], -L.img[3 * q + 2]};
            int lo = L.off[j], hi = L.off[j + 1];
            while (lo < hi) {
                const int mid = (lo + hi) / 2;
                const Cand cm{L.idx[mid], L.img[3 * mid], L.img[3 * mid + 1], L.img[3 * mid + 2]};
                if (cand_less(cm, key)) lo = mid + 1; else hi = mid;
            }
            if (lo >= L.off[j + 1] || L.idx[lo] != i || L.img[3 * lo] != key.n0 ||
                L.img[3 * lo + 1] != key.n1 || L.img[3 * lo + 2] != key.n2)
                throw std::runtime_error("orc: neighbour list is not symmetric");
            L.rev[q] = lo;
        }
    }
}

static bool accept(double d2, double rc2, int i, int j, int n0, int n1, int n2) {
    if (i == j && n0 == 0 && n1 == 0 && n2 == 0) return false;
    return d2 > 0.0 && d2 < rc2;
}

void build_brute(const orc_struc* s, double rc, NList& L) {
    const int N = s->natoms;
    double Ri[9], w[3];
    inv3(s->cell, Ri);
    widths(Ri, w);
    const double rc2 = rc * rc;
    std::vector<double> frac(3 * (size_t)N);
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < 3; ++a)
            frac[3 * i + a] = Ri[3 * a] * s->pos[3 * i] + Ri[3 * a + 1] * s->pos[3 * i + 1] +
                              Ri[3 * a + 2] * s->pos[3 * i + 2];
    std::vector<std::vector<Cand>> per(N);
    for (int i = 0; i < N; ++i) {
        for (int j = 0; j < N; ++j) {
            int lo[3], hi[3];
            for (int a = 0; a < 3; ++a) {
                const double ds = frac[3 * i + a] - frac[3 * j + a];
                lo[a] = (int)std::floor(ds - rc / w[a]) - 1;
                hi[a] = (int)std::ceil(ds + rc / w[a]) + 1;
            }
            for (int n0 = lo[0]; n0 <= hi[0]; ++n0)
                for (int n1 = lo[1]; n1 <= hi[1]; ++n1)
                    for (int n2 = lo[2]; n2 <= hi[2]; ++n2) {
                        double v[3], d2;
                        pair_vec(s->pos + 3 * i, s->pos + 3 * j, s->cell, n0, n1, n2, v, &d2);
                        if (accept(d2, rc2, i, j, n0, n1, n2)) per[i].push_back({j, n0, n1, n2});
                    }
        }
    }
    finish_list(s, L, per);
}

static inline int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

void build_cells(const orc_struc* s, double rc, NList& L) {
    const int N = s->natoms;
    double Ri[9], w[3];
    inv3(s->cell, Ri);
    widths(Ri, w);
    const double rc2 = rc * rc;
    int nb[3], k[3];
    for (int a = 0; a < 3; ++a) {
        nb[a] = std::max(1, (int)std::floor(w[a] / rc));
        k[a] = (int)std::ceil(rc * nb[a] / w[a]) + 1;
    }
    std::vector<int> shift(3 * (size_t)N), bin(N);
    std::vector<std::vector<int>> members(nb[0] * nb[1] * nb[2]);
    for (int i = 0; i < N; ++i) {
        int b3[3];
        for (int a = 0; a < 3; ++a) {
            const double f = Ri[3 * a] * s->pos[3 * i] + Ri[3 * a + 1] * s->pos[3 * i + 1] +
                             Ri[3 * a + 2] * s->pos[3 * i + 2];
            const double fl = std::floor(f);
            shift[3 * i + a] = (int)fl;
            b3[a] = std::min(nb[a] - 1, std::max(0, (int)std::floor((f - fl) * nb[a])));
        }
        bin[i] = (b3[0] * nb[1] + b3[1]) * nb[2] + b3[2];
        members[bin[i]].push_back(i);
    }
    std::vector<std::vector<Cand>> per(N);
    for (int i = 0; i < N; ++i) {
        const int bi0 = bin[i] / (nb[1] * nb[2]), bi1 = (bin[i] / nb[2]) % nb[1], bi2 = bin[i] % nb[2];
        for (int u0 = bi0 - k[0]; u0 <= bi0 + k[0]; ++u0)
            for (int u1 = bi1 - k[1]; u1 <= bi1 + k[1]; ++u1)
                for (int u2 = bi2 - k[2]; u2 <= bi2 + k[2]; ++u2) {
                    const int m0 = floordiv(u0, nb[0]), m1 = floordiv(u1, nb[1]), m2 = floordiv(u2, nb[2]);
                    const int q = ((u0 - m0 * nb[0]) * nb[1] + (u1 - m1 * nb[1])) * nb[2] + (u2 - m2 * nb[2]);
                    for (int j : members[q]) {
                        const int n0 = shift[3 * i] - shift[3 * j] + m0;
                        const int n1 = shift[3 * i + 1] - shift[3 * j + 1] + m1;
                        const int n2 = shift[3 * i + 2] - shift[3 * j + 2] + m2;
                        double v[3], d2;
                        pair_vec(s->pos + 3 * i, s->pos + 3 * j, s->cell, n0, n1, n2, v, &d2);
                        if (accept(d2, rc2, i, j, n0, n1, n2)) per[i].push_back({j, n0, n1, n2});
                    }
                }
    }
    finish_list(s, L, per);
}

// ---- symmetry functions ----
struct Basis {
    const orc_model* m;
    int S, NP;
    explicit Basis(const orc_model* mm) : m(mm), S(mm->nspecies), NP(npairs(mm->nspecies)) {}
    void radial(int a, int b, double r, double* G, double* dG) const {
        const int t = a * S + b, o = m->rad_off[t];
        orc_radial(m->rad_name[t], m->cutoff, m->rc, m->rad_n[t], m->rad_rs + o, m->rad_eta + o, r, G, dG);
    }
    int nrad(int a, int b) const { return m->rad_n[a * S + b]; }
    void angular(int a, int p, double r0, double r1, double c, double* G, double* dc, double* d0,
                 double* d1) const {
        const int t = a * NP + p, o = m->ang_off[t];
        orc_angular(m->ang_name[t], m->cutoff, m->rc, m->ang_n[t], m->ang_eta + o, m->ang_zeta + o,
                    m->ang_lambda + o, r0, r1, c, G, dc, d0, d1);
    }
    int nang(int a, int p) const { return m->ang_n[a * NP + p]; }
};

static int max_basis(const orc_model* m) {
    const int S = m->nspecies, NP = npairs(S);
    int mx = 1;
    for (int t = 0; t < S * S; ++t) mx = std::max(mx, m->rad_n[t]);
    for (int t = 0; t < S * NP; ++t) mx = std::max(mx, m->ang_n[t]);
    return mx;
}

void symm(const orc_model* m, const orc_struc* s, const NList& L, double* G, int ldg) {
    Basis B(m);
    std::vector<double> buf(max_basis(m));
    for (int i = 0; i < s->natoms; ++i) {
        double* g = G + (size_t)i * ldg;
        std::fill(g, g + ldg, 0.0);
        const int a = s->species[i];
        for (int q = L.off[i]; q < L.off[i + 1]; ++q) {
            const int b = s->species[L.idx[q]];
            B.radial(a, b, L.d[q], buf.data(), nullptr);
            const int oR = offset_r(m, a, b), nR = B.nrad(a, b);
            for (int t = 0; t < nR; ++t) g[oR + t] += buf[t];
            const double* rj = &L.r[3 * q];
            for (int u = q + 1; u < L.off[i + 1]; ++u) {
                const int c = s->species[L.idx[u]];
                const double* rk = &L.r[3 * u];
                const double cs = (rj[0] * rk[0] + rj[1] * rk[1] + rj[2] * rk[2]) / (L.d[q] * L.d[u]);
                const int p = pair_index(b, c);
                B.angular(a, p, L.d[q], L.d[u], cs, buf.data(), nullptr, nullptr, nullptr);
                const int oA = offset_a(m, a, p), nA = B.nang(a, p);
                for (int t = 0; t < nA; ++t) g[oA + t] += buf[t];
            }
        }
    }
}

// P[q] = dE_i/d r_ij for every slot q of centre i (the per-slot force buffer)
void slot_grad(const orc_model* m, const orc_struc* s, const NList& L, const double* dEdG, int ldg,
               std::vector<double>& P) {
    Basis B(m);
    const int mb = max_basis(m);
    std::vector<double> G(mb), dc(mb), d0(mb), d1(mb);
    P.assign(3 * (size_t)L.off[s->natoms], 0.0);
    for (int i = 0; i < s->natoms; ++i) {
        const double* de = dEdG + (size_t)i * ldg;
        const int a = s->species[i];
        for (int q = L.off[i]; q < L.off[i + 1]; ++q) {
            const int b = s->species[L.idx[q]];
            const double dj = L.d[q];
            const double* rj = &L.r[3 * q];
            B.radial(a, b, dj, G.data(), d0.data());
            const int oR = offset_r(m, a, b), nR = B.nrad(a, b);
            double amp = 0;
            for (int t = 0; t < nR; ++t) amp += de[oR + t] * d0[t];
            for (int x = 0; x < 3; ++x) P[3 * q + x] += amp * rj[x] / dj;
            for (int u = q + 1; u < L.off[i + 1]; ++u) {
                const int c = s->species[L.idx[u]];
                const double dk = L.d[u];
                const double* rk = &L.r[3 * u];
                const double cs = (rj[0] * rk[0] + rj[1] * rk[1] + rj[2] * rk[2]) / (dj * dk);
                const int p = pair_index(b, c);
                B.angular(a, p, dj, dk, cs, G.data(), dc.data(), d0.data(), d1.data());
                const int oA = offset_a(m, a, p), nA = B.nang(a, p);
                double A = 0, Bj = 0, Bk = 0;
                for (int t = 0; t < nA; ++t) {
                    A += de[oA + t] * dc[t];
                    Bj += de[oA + t] * d0[t];
                    Bk += de[oA + t] * d1[t];
                }
                for (int x = 0; x < 3; ++x) {
                    const double hj = rj[x] / dj, hk = rk[x] / dk;
                    P[3 * q + x] += A * (hk - cs * hj) / dj + Bj * hj;
                    P[3 * u + x] += A * (hj - cs * hk) / dk + Bk * hk;
                }
            }
        }
    }
}

void force(const orc_model* m, const orc_struc* s, const NList& L, const double* dEdG, int ldg,
           int mode, double* F, double* V, double* Pout) {
    std::vector<double> P;
    slot_grad(m, s, L, dEdG, ldg, P);
    const int N = s->natoms;
    std::fill(F, F + 3 * (size_t)N, 0.0);
    if (V) std::fill(V, V + 9, 0.0);
    for (int i = 0; i < N; ++i)
        for (int q = L.off[i]; q < L.off[i + 1]; ++q) {
            const int j = L.idx[q];
            for (int x = 0; x < 3; ++x) {
                F[3 * i + x] -= P[3 * q + x];
                if (mode == 0 || L.minimg[q]) F[3 * j + x] += P[3 * q + x];
            }
            if (V)
                for (int x = 0; x < 3; ++x)
                    for (int y = 0; y < 3; ++y) V[3 * x + y] -= P[3 * q + x] * L.r[3 * q + y];
        }
    if (Pout) std::copy(P.begin(), P.end(), Pout);
}

// dG_i[u] = sum_slots dG_i/dr_ij . (u_i - u_j): forward-mode tangent of the basis
void tangent(const orc_model* m, const orc_struc* s, const NList& L, const double* u, double* v,
             int ldg) {
    Basis B(m);
    const int mb = max_basis(m);
    std::vector<double> G(mb), dc(mb), d0(mb), d1(mb);
    for (int i = 0; i < s->natoms; ++i) {
        double* g = v + (size_t)i * ldg;
        std::fill(g, g + ldg, 0.0);
        const int a = s->species[i];
        for (int q = L.off[i]; q < L.off[i + 1]; ++q) {
            const int j = L.idx[q], b = s->species[j];
            const double dj = L.d[q];
            const double* rj = &L.r[3 * q];
            double tj[3], hj[3];
            for (int x = 0; x < 3; ++x) { tj[x] = u[3 * i + x] - u[3 * j + x]; hj[x] = rj[x] / dj; }
            const double rdj = hj[0] * tj[0] + hj[1] * tj[1] + hj[2] * tj[2];
            B.radial(a, b, dj, G.data(), d0.data());
            const int oR = offset_r(m, a, b), nR = B.nrad(a, b);
            for (int t = 0; t < nR; ++t) g[oR + t] += d0[t] * rdj;
            for (int w = q + 1; w < L.off[i + 1]; ++w) {
                const int k = L.idx[w], c = s->species[k];
                const double dk = L.d[w];
                const double* rk = &L.r[3 * w];
                double tk[3], hk[3];
                for (int x = 0; x < 3; ++x) { tk[x] = u[3 * i + x] - u[3 * k + x]; hk[x] = rk[x] / dk; }
                const double cs = hj[0] * hk[0] + hj[1] * hk[1] + hj[2] * hk[2];
                const double rdk = hk[0] * tk[0] + hk[1] * tk[1] + hk[2] * tk[2];
                double cdot = 0;
                for (int x = 0; x < 3; ++x)
                    cdot += (hk[x] - cs * hj[x]) / dj * tj[x] + (hj[x] - cs * hk[x]) / dk * tk[x];
                const int p = pair_index(b, c);
                B.angular(a, p, dj, dk, cs, G.data(), dc.data(), d0.data(), d1.data());
                const int oA = offset_a(m, a, p), nA = B.nang(a, p);
                for (int t = 0; t < nA; ++t) g[oA + t] += dc[t] * cdot + d0[t] * rdj + d1[t] * rdk;
            }
        }
    }
}

static int max_ninput(const orc_model* m) {
    int mx = 1;
    for (int a = 0; a < m->nspecies; ++a) mx = std::max(mx, ninput(m, a));
    return mx;
}

double compute(const orc_model* m, const orc_struc* s, int mode, double* Eatom, double* F, double* V) {
    NList L;
    build_cells(s, m->rc, L);
    const int N = s->natoms, ld = max_ninput(m);
    std::vector<double> G((size_t)N * ld), dE((size_t)N * ld, 0.0);
    symm(m, s, L, G.data(), ld);
    double E = 0;
    for (int i = 0; i < N; ++i) {
        const int a = s->species[i];
        const double e = orc_nn_forward(m, a, &G[(size_t)i * ld], &dE[(size_t)i * ld]);
        if (Eatom) Eatom[i] = e;
        E += e + m->evac[a];
    }
    if (F) force(m, s, L, dE.data(), ld, mode, F, V, nullptr);
    return E;
}

}  // namespace orc

using namespace orc;

extern "C" int orc_ninput(const orc_model* m, int a) { return ninput(m, a); }
extern "C" int orc_nparams(const orc_model* m, int a) { return Net(m, a).nparam; }

extern "C" int64_t orc_nlist_build(const orc_struc* s, double rc, int method, void** handle) {
    NList* L = new NList;
    if (method == 0) build_brute(s, rc, *L); else build_cells(s, rc, *L);
    *handle = L;
    return L->off[s->natoms];
}

extern "C" void orc_nlist_get(void* h, int* off, int* idx, int* img, double* r, double* d, int* rev,
                              int* minimg) {
    const NList& L = *static_cast<NList*>(h);
    if (off) std::copy(L.off.begin(), L.off.end(), off);
    if (idx) std::copy(L.idx.begin(), L.idx.end(), idx);
    if (img) std::copy(L.img.begin(), L.img.end(), img);
    if (r) std::copy(L.r.begin(), L.r.end(), r);
    if (d) std::copy(L.d.begin(), L.d.end(), d);
    if (rev) std::copy(L.rev.begin(), L.rev.end(), rev);
    if (minimg) std::copy(L.minimg.begin(), L.minimg.end(), minimg);
}

extern "C" void orc_nlist_free(void* h) { delete static_cast<NList*>(h); }

extern "C" void orc_symm(const orc_model* m, const orc_struc* s, void* nl, double* G, int ldg) {
    symm(m, s, *static_cast<NList*>(nl), G, ldg);
}

extern "C" void orc_net(const orc_model* m, int N, const int* species, const double* G, int ldg,
                        double* Eatom, double* dEdG) {
    for (int i = 0; i < N; ++i) {
        const double e = orc_nn_forward(m, species[i], G + (size_t)i * ldg, dEdG ? dEdG + (size_t)i * ldg : nullptr);
        if (Eatom) Eatom[i] = e;
    }
}

extern "C" void orc_force(const orc_model* m, const orc_struc* s, void* nl, const double* dEdG,
                          int ldg, int mode, double* F, double* V, double* P) {
    force(m, s, *static_cast<NList*>(nl), dEdG, ldg, mode, F, V, P);
}

extern "C" double orc_compute(const orc_model* m, const orc_struc* s, int mode, double* Eatom,
                              double* F, double* V) {
    return compute(m, s, mode, Eatom, F, V);
}

extern "C" void orc_compute_batch(const orc_model* m, int nstruc, const int* off, const double* pos,
                                  const double* cells, const int* species, int nthreads, double* E,
                                  double* F, double* V) {
    if (nthreads < 1) nthreads = 1;
    auto work = [&](int t) {
        // contiguous block split of the structures (thread::Dist semantics)
        const int base = nstruc / nthreads, rem = nstruc % nthreads;
        const int lo = t * base + std::min(t, rem), hi = lo + base + (t < rem ? 1 : 0);
        for (int k = lo; k < hi; ++k) {
            orc_struc s{off[k + 1] - off[k], pos + 3 * (size_t)off[k], cells + 9 * (size_t)k,
                        species + off[k]};
            E[k] = compute(m, &s, 0, nullptr, F ? F + 3 * (size_t)off[k] : nullptr,
                           V ? V + 9 * (size_t)k : nullptr);
        }
    };
    if (nthreads == 1) { work(0); return; }
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
}

extern "C" void orc_symm_tangent(const orc_model* m, const orc_struc* s, void* nl, const double* u,
                                 double* v, int ldg) {
    tangent(m, s, *static_cast<NList*>(nl), u, v, ldg);
}

// Energy+force loss of a batch and its gradient with respect to every
// network parameter (species blocks concatenated in species order):
//   L = (1/B) sum_s [ wE/2 ((E_s - Eref_s)/N_s)^2 + wF/(3 N_s) sum_a |F_a - Fref_a|^2 ]
// E_s is the sum of network outputs (vacuum energies excluded, as in the
// reference trainer which subtracts them from the targets, nnptefr2.cpp:2781-2794).
// Energy loss of one structure on dU = (E - Eref) * norm and its derivative
// d loss / d dU (NNPTEFR::error_cost_o1, src/nnp/nnptefr2.cpp:1336-1372;
// opt::Loss numbering src/opt/loss.hpp:18-24, Norm src/nnp/nnptefr2.hpp:117-126)
static void energy_loss(int loss, double delta, double dU, double& l, double& g) {
    switch (loss) {
        case ORC_LOSS_MAE: l = std::fabs(dU); g = (dU > 0) - (dU < 0); break;
        case ORC_LOSS_HUBER: {
            const double a = dU / delta, q = std::sqrt(1.0 + a * a);
            l = delta * delta * (q - 1.0); g = dU / q;
        } break;
        case ORC_LOSS_ASINH: {
            const double a = dU / delta, q = std::sqrt(1.0 + a * a), as = std::asinh(a);
            l = delta * delta * (1.0 - q + a * as); g = delta * as;
        } break;
        default: l = 0.5 * dU * dU; g = dU; break;
    }
}

static double energy_norm(int norm, double n) {
    switch (norm) {
        case ORC_NORM_IDENTITY: return 1.0;
        case ORC_NORM_SQRT: return 1.0 / std::sqrt(n);
        case ORC_NORM_CBRT: return 1.0 / std::cbrt(n);
        case ORC_NORM_LOG: return 1.0 / (std::log(n) + 1.0);
        default: return 1.0 / n;   // LINEAR
    }
}

extern "C" double orc_train_loss_grad(const orc_model* m, int B, const int* off, const double* pos,
                                      const double* cells, const int* species, const double* Eref,
                                      const double* Fref, double wE, double wF, double* grad,
                                      double* Eout, double* Fout, double* lossparts) {
    return orc_train_loss_grad_ex(m, B, off, pos, cells, species, Eref, Fref, wE, wF, ORC_LOSS_MSE, 1.0,
                                  ORC_NORM_LINEAR, nullptr, nullptr, grad, Eout, Fout, lossparts);
}

// L = (1/B) sum_s [ wE wE_s loss((E_s - Eref_s) norm_s) + wF wF_s / (3 N_s) sum |F - Fref|^2 ]
extern "C" double orc_train_loss_grad_ex(const orc_model* m, int B, const int* off, const double* pos,
                                         const double* cells, const int* species, const double* Eref,
                                         const double* Fref, double wE, double wF, int loss, double delta,
                                         int norm, const double* wEs, const double* wFs, double* grad,
                                         double* Eout, double* Fout, double* lossparts) {
    const int S = m->nspecies;
    std::vector<int> poff(S + 1, 0);
    for (int a = 0; a < S; ++a) poff[a + 1] = poff[a] + orc_nparams(m, a);
    std::fill(grad, grad + poff[S], 0.0);
    const int ld = max_ninput(m);
    double lossE = 0, lossF = 0;
    for (int k = 0; k < B; ++k) {
        const int N = off[k + 1] - off[k];
        orc_struc s{N, pos + 3 * (size_t)off[k], cells + 9 * (size_t)k, species + off[k]};
        NList L;
        build_cells(&s, m->rc, L);
        std::vector<double> G((size_t)N * ld), dE((size_t)N * ld, 0.0), F(3 * (size_t)N), u(3 * (size_t)N),
            v((size_t)N * ld);
        symm(m, &s, L, G.data(), ld);
        double E = 0;
        for (int i = 0; i < N; ++i) E += orc_nn_forward(m, s.species[i], &G[(size_t)i * ld], &dE[(size_t)i * ld]);
        force(m, &s, L, dE.data(), ld, 0, F.data(), nullptr, nullptr);
        const double we = wE * (wEs ? wEs[k] : 1.0), wf = wF * (wFs ? wFs[k] : 1.0);
        const double nrm = energy_norm(norm, (double)N);
        double l, gl;
        energy_loss(loss, delta, (E - Eref[k]) * nrm, l, gl);
        lossE += we * l / B;
        const double gE = we * gl * nrm / B;
        double sf = 0;
        for (int x = 0; x < 3 * N; ++x) {
            const double df = F[x] - Fref[3 * (size_t)off[k] + x];
            sf += df * df;
            u[x] = wf * 2.0 * df / (3.0 * N) / B;
        }
        lossF += wf * sf / (3.0 * N) / B;
        tangent(m, &s, L, u.data(), v.data(), ld);
        std::vector<double> g;
        for (int i = 0; i < N; ++i) {
            const int a = s.species[i];
            g.assign(poff[a + 1] - poff[a], 0.0);
            for (int t = 0; t < ld; ++t) v[(size_t)i * ld + t] = -v[(size_t)i * ld + t];
            orc_nn_grad_rop(m, a, &G[(size_t)i * ld], &v[(size_t)i * ld], gE, g.data());
            for (size_t t = 0; t < g.size(); ++t) grad[poff[a] + t] += g[t];
        }
        if (Eout) Eout[k] = E;
        if (Fout) std::copy(F.begin(), F.end(), Fout + 3 * (size_t)off[k]);
    }
    if (lossparts) { lossparts[0] = lossE; lossparts[1] = lossF; }
    return lossE + lossF;
}
