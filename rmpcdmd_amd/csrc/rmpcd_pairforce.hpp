// rmpcd_pairforce.hpp -- compute_force_n2 (neighbor_list.f90:227-264) for many colloids:
// one workgroup per colloid i, its threads over the partners j, each pair (a, b) with
// a < b evaluated as the reference visits it (cut from cut_sq(s_a, s_b), force and energy
// from epsilon(s_b, s_a)); colloid i receives -f(j, i) for j < i and +f(i, j) for j > i,
// and the energy of its pairs with j > i, summed as a tree over the workgroup and added to
// the force it already holds.  Templates, so that every translation unit may use them.
#pragma once
#include "rmpcd_internal.hpp"

namespace rmpcd {

// the pair sums of colloid i over a workgroup of T threads (every thread returns the tree
// sums of the force and the energy)
template <int T>
__device__ inline void n2_pair_sums(const double *__restrict__ pos, const int32_t *__restrict__ species, int64_t n,
                                    int64_t i, double L0, double L1, double L2, const LJ &lj, double *red, double &f0,
                                    double &f1, double &f2, double &e)
{
    f0 = 0;
    f1 = 0;
    f2 = 0;
    e = 0;
    for (int64_t j = threadIdx.x; j < n; j += T) {
        if (j == i) continue;
        const int64_t a = j < i ? j : i, bb = j < i ? i : j;
        const double d0 = min_image(pos[a] - pos[bb], L0);
        const double d1 = min_image(pos[n + a] - pos[n + bb], L1);
        const double d2 = min_image(pos[2 * n + a] - pos[2 * n + bb], L2);
        const double r_sq = (d0 * d0 + d1 * d1) + d2 * d2;
        const int sa = species[a], sb = species[bb];
        const int pc = (sa - 1) + lj.n1 * (sb - 1), pe = (sb - 1) + lj.n1 * (sa - 1);
        if (r_sq <= lj.cut_sq[pc]) {
            const double fr = lj_force_over_r(r_sq, lj.eps[pe], lj.sigma[pe]);
            if (j < i) {
                f0 = f0 - fr * d0;
                f1 = f1 - fr * d1;
                f2 = f2 - fr * d2;
            } else {
                f0 = f0 + fr * d0;
                f1 = f1 + fr * d1;
                f2 = f2 + fr * d2;
                e = e + lj_energy(r_sq, lj.eps[pe], lj.sigma[pe], lj.cut_energy[pe]);
            }
        }
    }
    f0 = block_sum(f0, red);
    f1 = block_sum(f1, red);
    f2 = block_sum(f2, red);
    e = block_sum(e, red);
}

template <int T>
__global__ void __launch_bounds__(T)
    k_force_n2_blocks(const double *__restrict__ pos, const int32_t *__restrict__ species, int64_t n,
                      const double *__restrict__ force_in, double L0, double L1, double L2, LJ lj,
                      double *__restrict__ fnew, double *__restrict__ e_out)
{
    __shared__ double red[16];
    const int64_t i = blockIdx.x;
    double f0, f1, f2, e;
    n2_pair_sums<T>(pos, species, n, i, L0, L1, L2, lj, red, f0, f1, f2, e);
    if (threadIdx.x == 0) {
        fnew[i] = force_in[i] + f0;
        fnew[n + i] = force_in[n + i] + f1;
        fnew[2 * n + i] = force_in[2 * n + i] + f2;
        e_out[i] = e;
    }
}

// the sum of n doubles by one workgroup (strided partial sums, then a tree)
template <int T>
__global__ void __launch_bounds__(T) k_sum_block(const double *__restrict__ a, int64_t n, double *__restrict__ out)
{
    __shared__ double red[16];
    double s = 0;
    for (int64_t i = threadIdx.x; i < n; i += T) s += a[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[0] = s;
}

}  // namespace rmpcd
