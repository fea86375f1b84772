// rmpcd_setup.hpp -- set-up kernels shared by the driver programs: Maxwell-Boltzmann
// (or uniform) velocities without net momentum, random placement around obstacles
// (particle_system.f90:250-302) and the initial particle_system_t fields
// (particle_system.f90:81-184).  All are templates, so each translation unit that
// includes this header instantiates its own copies.
#pragma once
#include <math.h>

#include "rmpcd_internal.hpp"

namespace rmpcd {

// ---- set-up kernels --------------------------------------------------------------
// Maxwell-Boltzmann velocities: particle i draws threefry((0, i), (seed, INIT)) and
// threefry((1, i), (seed, INIT)); Box-Muller on the two pairs gives three normals
__device__ inline void normals3(uint64_t seed, uint64_t id, double out[3])
{
    uint64_t a, b, c, d;
    threefry2x64(0, id, seed, STREAM_INIT, a, b);
    threefry2x64(1, id, seed, STREAM_INIT, c, d);
    double u1 = 1.0 - u01(a), u2 = u01(b), u3 = 1.0 - u01(c), u4 = u01(d);
    double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    out[0] = r1 * cos(2 * M_PI * u2);
    out[1] = r1 * sin(2 * M_PI * u2);
    out[2] = r2 * cos(2 * M_PI * u4);
}

// velocities and their per-block sums (summed in block order by k_remove_mean).
// uniform != 0: (u - 1/2) sqrt(12 T) per component from threefry((0, i)) a, b and
// threefry((1, i)) a, the random_number form of setup_single_catalytic_fixed_sphere.f90:143-144
template <class F>
__global__ void k_init_velocities(PView<F> p, uint64_t seed, double sqrtT, double *__restrict__ part,
                                  int uniform = 0)
{
    __shared__ double red[16];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[3] = {0, 0, 0};
    if (i < p.n) {
        if (uniform) {
            uint64_t a, b, c, d;
            threefry2x64(0, (uint64_t)i, seed, STREAM_INIT, a, b);
            threefry2x64(1, (uint64_t)i, seed, STREAM_INIT, c, d);
            v[0] = u01(a) - 0.5;
            v[1] = u01(b) - 0.5;
            v[2] = u01(c) - 0.5;
        } else {
            normals3(seed, (uint64_t)i, v);
        }
        for (int k = 0; k < 3; ++k) {
            p.vel[k * p.n + i] = (typename F::real)(v[k] * sqrtT);
            v[k] = (double)p.vel[k * p.n + i];
        }
    }
    for (int k = 0; k < 3; ++k) {
        double s = block_sum(v[k], red);
        if (threadIdx.x == 0) part[3 * blockIdx.x + k] = s;
    }
}

template <int K = 3>
__global__ void k_sum_blocks3(const double *__restrict__ part, int nb, double *__restrict__ out)
{
    __shared__ double red[16];
    double s[K] = {};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int k = 0; k < K; ++k) s[k] += part[K * b + k];
    for (int k = 0; k < K; ++k) {
        double t = block_sum(s[k], red);
        if (threadIdx.x == 0) out[k] = t;
    }
}

template <class F>
__global__ void k_remove_mean(PView<F> p, const double *__restrict__ sum)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    for (int k = 0; k < 3; ++k)
        p.vel[k * p.n + i] = (typename F::real)((double)p.vel[k * p.n + i] - sum[k] / (double)p.n);
}

// random_placement (particle_system.f90:250-302) with the colloids as obstacles;
// attempt k of particle i uses threefry((k, i), (seed, PLACE)) for x, y and
// threefry((k, i + 2^40), (seed, PLACE)) for z
template <class F>
__global__ void k_place_random(PView<F> p, PView<F64> obst, int n_obst, double L0, double L1, double L2,
                               int n1, double cs0, double cs1, double cs2, double cs3, uint64_t seed,
                               int32_t *__restrict__ fail)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    double cut_sq[4] = {cs0, cs1, cs2, cs3};
    int s1 = (int)p.species[i];
    double x[3] = {0, 0, 0};
    bool ok = false;
    for (uint64_t k = 0; k < 4096 && !ok; ++k) {
        uint64_t a, b, c, d;
        threefry2x64(k, (uint64_t)i, seed, STREAM_PLACE, a, b);
        threefry2x64(k, (uint64_t)i + (1ull << 40), seed, STREAM_PLACE, c, d);
        x[0] = u01(a) * L0;
        x[1] = u01(b) * L1;
        x[2] = u01(c) * L2;
        ok = true;
        for (int j = 0; j < n_obst && ok; ++j) {
            double d0 = min_image(x[0] - obst.pos[j], L0);
            double d1 = min_image(x[1] - obst.pos[obst.n + j], L1);
            double d2 = min_image(x[2] - obst.pos[2 * obst.n + j], L2);
            double rsq = (d0 * d0 + d1 * d1) + d2 * d2;
            int s2 = obst.species[j];
            int q = (s1 - 1) + n1 * (s2 - 1);
            if (q >= 0 && q < 4 && rsq < cut_sq[q]) ok = false;
        }
    }
    if (!ok) atomicOr(fail, 1);
    // positions are stored in the particle precision; a float rounding up to L wraps
    for (int k = 0; k < 3; ++k) {
        typename F::real v = (typename F::real)x[k];
        double e = k == 0 ? L0 : k == 1 ? L1 : L2;
        if ((double)v >= e) v = (typename F::real)0;
        p.pos[k * p.n + i] = v;
    }
}

template <class F>
__global__ void k_fill_state(PView<F> p)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    for (int k = 0; k < 3; ++k) {
        p.force[k * p.n + i] = 0;
        p.force_old[k * p.n + i] = 0;
        p.image[k * p.n + i] = 0;
    }
    p.id[i] = (int32_t)i;
    p.species[i] = (typename F::small)1;
    p.flags[i] = (typename F::small)0;
}


}  // namespace rmpcd
