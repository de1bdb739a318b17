// pair.hip -- LAMMPS-style pair interface: one periodic domain given as owned
// ("local") atoms plus ghost copies, with a full neighbour list of the local
// atoms over local and ghost indices (LAMMPS list->numneigh / firstneigh).
// The reference names the LAMMPS pair style (README.md:411-416,
// examples/Ar/Ar_md_npt.in:24-25) without its source; its in-tree MD caller
// PotNNPE::compute (src/torch/pot_nnpe.cpp:166-287) fixes the semantics:
// energy of every local atom from all its neighbours, forces on every atom the
// pair terms touch.
//
// Each ghost is a lattice image of the local atom that owns it, so a ghost
// partner j of centre i becomes the entry (i -> owner(j), n) with
// R n = x_j - x_owner(j).  The converted list is the context's periodic list
// (ann_set_neighbors), and the force gather over reverse slots then leaves the
// force of every ghost on its owner: the fold-back is exact and needs no
// second pass.  Between list rebuilds only the local positions move; the
// image vectors stay valid (ghosts move with their owners) and the pair
// geometry is refreshed from them.  Host code; the device work is
// ann_energy_forces.
#include <cmath>
#include <string>
#include <vector>

#include "common.h"

namespace ann {

struct PairState {
    int cap = 0;
    double* pos = nullptr;      // local positions [cap][3]
    double* cell = nullptr;     // [9]
    int* species = nullptr;     // [cap]
    int* off = nullptr;         // [2]
    double* out = nullptr;      // E [1], virial [9], F [cap][3]
    int hoff[2] = {0, 0};
    int nlocal = 0;
    bool have_list = false;
    std::vector<double> h_out;
    std::vector<float> h_ea;
};

void pair_free(ann_ctx* ctx) {
    PairState* p = ctx->pair;
    if (!p) return;
    void* ptrs[] = {p->pos, p->cell, p->species, p->off, p->out};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    delete p;
    ctx->pair = nullptr;
}

}  // namespace ann

using namespace ann;

extern "C" int ann_pair_compute(ann_ctx* ctx, int form, int new_list, int nlocal, int nghost, const double* x,
                                const int* species, const double* cell, const int* owner, const int* numneigh,
                                const int* const* firstneigh, double rc, double* f, double* eng, double* eatom,
                                double* virial) {
    if (!ctx) return fail(ANN_ERR_ARG, "null context");
    if (nlocal < 1 || nghost < 0) return fail(ANN_ERR_ARG, "invalid atom counts");
    if (!x || !species || !cell || !f) return fail(ANN_ERR_ARG, "null argument");
    if (new_list && (!numneigh || !firstneigh || (nghost > 0 && !owner)))
        return fail(ANN_ERR_ARG, "null neighbour list");
    const int nall = nlocal + nghost;
    for (int i = 0; i < nlocal; ++i)
        if (species[i] < 0 || species[i] >= ctx->S) return fail(ANN_ERR_ARG, "species index out of range");
    for (int q = 0; q < 9; ++q)
        if (!std::isfinite(cell[q])) return fail(ANN_ERR_ARG, "non-finite cell");
    const double* R = cell;
    const double a = R[0], b = R[1], c = R[2], d = R[3], e = R[4], g = R[5], h = R[6], k = R[7], m = R[8];
    const double A = e * m - g * k, B = -(d * m - g * h), Cc = d * k - e * h;
    const double det = a * A + b * B + c * Cc;
    if (!(std::fabs(det) > 0)) return fail(ANN_ERR_ARG, "singular cell");
    if (!ctx->pair) ctx->pair = new PairState;
    PairState& P = *ctx->pair;
    if (!new_list && (!P.have_list || P.nlocal != nlocal))
        return fail(ANN_ERR_STATE, "no neighbour list for this domain: call with new_list = 1 first");
    if (nlocal > P.cap) {
        hipStream_t st = ctx->stream;
        ANN_HIP(hipStreamSynchronize(st));
        void* ptrs[] = {P.pos, P.cell, P.species, P.off, P.out};
        for (void* q : ptrs)
            if (q) ANN_HIP(hipFree(q));
        P.pos = P.cell = P.out = nullptr;
        P.species = P.off = nullptr;
        P.cap = 0;
        P.have_list = false;
        ANN_HIP(hipMalloc(&P.pos, sizeof(double) * 3 * (size_t)nlocal));
        ANN_HIP(hipMalloc(&P.cell, sizeof(double) * 9));
        ANN_HIP(hipMalloc(&P.species, sizeof(int) * (size_t)nlocal));
        ANN_HIP(hipMalloc(&P.off, sizeof(int) * 2));
        ANN_HIP(hipMalloc(&P.out, sizeof(double) * (10 + 3 * (size_t)nlocal)));
        P.cap = nlocal;
    }
    hipStream_t st = ctx->stream;
    ANN_HIP(hipMemcpyAsync(P.pos, x, sizeof(double) * 3 * (size_t)nlocal, hipMemcpyHostToDevice, st));
    ANN_HIP(hipMemcpyAsync(P.cell, cell, sizeof(double) * 9, hipMemcpyHostToDevice, st));
    if (new_list) {
        P.have_list = false;
        P.hoff[0] = 0;
        P.hoff[1] = nlocal;
        ANN_HIP(hipMemcpyAsync(P.species, species, sizeof(int) * (size_t)nlocal, hipMemcpyHostToDevice, st));
        ANN_HIP(hipMemcpyAsync(P.off, P.hoff, sizeof(int) * 2, hipMemcpyHostToDevice, st));
        ANN_HIP(hipStreamSynchronize(st));     // the host arrays may change after return
        ann_batch bt{1, P.hoff, P.off, P.pos, P.cell, P.species};
        int r = ann_set_batch(ctx, &bt);
        if (r != ANN_OK) return r;
        // ghost partners -> (owner, image vector)
        const double id = 1.0 / det;
        const double Ri[9] = {A * id, -(b * m - c * k) * id, (b * g - c * e) * id,
                              B * id, (a * m - c * h) * id,  -(a * g - c * d) * id,
                              Cc * id, -(a * k - b * h) * id, (a * e - b * d) * id};
        std::vector<int> off(nlocal + 1, 0), idx, img;
        for (int i = 0; i < nlocal; ++i) {
            if (numneigh[i] < 0) return fail(ANN_ERR_ARG, "negative neighbour count");
            if (numneigh[i] > 0 && !firstneigh[i]) return fail(ANN_ERR_ARG, "null neighbour row");
            off[i + 1] = off[i] + numneigh[i];
        }
        idx.resize(off[nlocal]);
        img.resize(3 * (size_t)off[nlocal]);
        for (int i = 0; i < nlocal; ++i)
            for (int q = 0; q < numneigh[i]; ++q) {
                const int j = firstneigh[i][q];
                if (j < 0 || j >= nall) return fail(ANN_ERR_ARG, "neighbour index outside the local and ghost atoms");
                const int o = j < nlocal ? j : owner[j - nlocal];
                if (o < 0 || o >= nlocal)
                    return fail(ANN_ERR_ARG, "ghost " + std::to_string(j) + " has no local owner");
                int n[3] = {0, 0, 0};
                if (j >= nlocal) {
                    const double dx[3] = {x[3 * (size_t)j] - x[3 * (size_t)o], x[3 * (size_t)j + 1] - x[3 * (size_t)o + 1],
                                          x[3 * (size_t)j + 2] - x[3 * (size_t)o + 2]};
                    for (int t = 0; t < 3; ++t) {
                        const double fr = Ri[3 * t] * dx[0] + Ri[3 * t + 1] * dx[1] + Ri[3 * t + 2] * dx[2];
                        const double rn = std::nearbyint(fr);
                        if (!(std::fabs(fr - rn) <= 1e-6 * std::fmax(1.0, std::fabs(fr))))
                            return fail(ANN_ERR_ARG, "ghost " + std::to_string(j) +
                                                         " is not a lattice image of its owner");
                        n[t] = (int)rn;
                    }
                }
                const size_t s = (size_t)off[i] + q;
                idx[s] = o;
                for (int t = 0; t < 3; ++t) img[3 * s + t] = n[t];
            }
        if ((r = ann_set_neighbors(ctx, off.data(), idx.data(), img.data(), rc)) != ANN_OK) return r;
        P.have_list = true;
        P.nlocal = nlocal;
    } else {
        int r = ann_refresh_geometry(ctx);
        if (r != ANN_OK) return r;
    }
    double* dE = P.out;
    double* dV = P.out + 1;
    double* dF = P.out + 10;
    int r = ann_energy_forces(ctx, form, 1, dE, dF, virial ? dV : nullptr);
    if (r != ANN_OK) return r;
    P.h_out.resize(10 + 3 * (size_t)nlocal);
    ANN_HIP(hipMemcpyAsync(P.h_out.data(), P.out, sizeof(double) * P.h_out.size(), hipMemcpyDeviceToHost, st));
    if (eatom) {
        P.h_ea.resize(nlocal);
        ANN_HIP(hipMemcpyAsync(P.h_ea.data(), ctx->d_Eatom, sizeof(float) * (size_t)nlocal, hipMemcpyDeviceToHost, st));
    }
    ANN_HIP(hipStreamSynchronize(st));
    for (size_t q = 0; q < 3 * (size_t)nlocal; ++q) f[q] += P.h_out[10 + q];
    if (eng) *eng += P.h_out[0];
    if (eatom)
        for (int i = 0; i < nlocal; ++i) eatom[i] += (double)P.h_ea[i] + ctx->hm.net[species[i]].evac;
    if (virial) {
        const double* V = P.h_out.data() + 1;
        virial[0] += V[0];
        virial[1] += V[4];
        virial[2] += V[8];
        virial[3] += V[1];
        virial[4] += V[2];
        virial[5] += V[5];
    }
    return ANN_OK;
}
