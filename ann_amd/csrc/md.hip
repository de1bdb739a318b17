// md.hip -- the velocity-Verlet step either side of the energy/force path
// (Verlet::compute, src/torch/integrator.cpp:177-200; the MD loop with its
// list rebuild every `stride` steps, src/torch/torch.cpp:272-276).
//
// Arrays are FP64 [N][3] atom-major (positions, velocities, forces) plus the
// per-atom mass; the update is elementwise and HBM-bound (72 B read, 48 B
// written per atom for the first half, 56 B / 32 B for the second).  The
// first half also measures the displacement since the last neighbour build
// and raises a flag when any atom moved more than half the skin (the list is
// then no longer guaranteed to hold every pair inside rc).
#include "common.h"

namespace ann {

// v += dt/2 F/m; x += v dt; flag if |x - x_ref|^2 > lim2
__global__ void k_kick_drift(int N, double dt, const double* __restrict__ mass, const double* __restrict__ F,
                             double* __restrict__ vel, double* __restrict__ pos, const double* __restrict__ pos_ref,
                             double lim2, int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double h = 0.5 * dt / mass[i];
    double d2 = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double v = vel[3 * i + q] + h * F[3 * i + q];
        const double x = pos[3 * i + q] + v * dt;
        vel[3 * i + q] = v;
        pos[3 * i + q] = x;
        const double dx = x - pos_ref[3 * i + q];
        d2 += dx * dx;
    }
    if (d2 > lim2) atomicOr(flag, 1);
}

// v += dt/2 F/m; per-atom kinetic energy m v^2 / 2
__global__ void k_kick(int N, double dt, const double* __restrict__ mass, const double* __restrict__ F,
                       double* __restrict__ vel, double* __restrict__ keatom) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double m = mass[i], h = 0.5 * dt / m;
    double v2 = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double v = vel[3 * i + q] + h * F[3 * i + q];
        vel[3 * i + q] = v;
        v2 += v * v;
    }
    keatom[i] = 0.5 * m * v2;
}

hipError_t md_kick_drift(ann_ctx* ctx, double dt, const double* mass, const double* F, double* vel, double* pos,
                         double lim2) {
    const int N = ctx->natoms;
    if (N == 0) return hipSuccess;
    Launch L(ctx, "md_kick_drift");
    k_kick_drift<<<(N + 255) / 256, 256, 0, ctx->stream>>>(N, dt, mass, F, vel, pos, ctx->d_pos_ref, lim2,
                                                           ctx->d_md_flag);
    return hipGetLastError();
}

hipError_t md_kick(ann_ctx* ctx, double dt, const double* mass, const double* F, double* vel) {
    const int N = ctx->natoms;
    if (N == 0) return hipSuccess;
    Launch L(ctx, "md_kick");
    k_kick<<<(N + 255) / 256, 256, 0, ctx->stream>>>(N, dt, mass, F, vel, ctx->d_keatom);
    return hipGetLastError();
}

}  // namespace ann
