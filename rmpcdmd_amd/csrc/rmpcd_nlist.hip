// rmpcd_nlist.hip -- stencil neighbour list of the colloids, colloid-solvent and
// colloid-colloid Lennard-Jones forces and the catalytic helpers of the drivers
// (src/neighbor_list.f90:76-264, programs/single_dimer_pbc.f90:493-565).
//
// update_list keeps the reference's list order (stencil order, then the cell's sorted
// order) with two passes over (point, stencil cell): a count pass, a per-point scan of
// the counts, and a fill pass that writes every accepted particle at its final offset.
#include "rmpcd_internal.hpp"
#include "rmpcd_pairforce.hpp"

namespace rmpcd {

constexpr int NL_THREADS = 256;

// cell of a colloid and the Hilbert key of stencil cell j (neighbor_list.f90:111-121)
__device__ inline int stencil_key(const Box &b, const double *x, const int32_t *st, int j)
{
    int p0, p1, p2;
    cell_of<double>(b, x[0], x[1], x[2], p0, p1, p2);
    int a0 = (p0 + st[3 * j]) % b.L[0], a1 = (p1 + st[3 * j + 1]) % b.L[1],
        a2 = (p2 + st[3 * j + 2]) % b.L[2];
    a0 += a0 < 0 ? b.L[0] : 0;
    a1 += a1 < 0 ? b.L[1] : 0;
    a2 += a2 < 0 ? b.L[2] : 0;
    return compact_key(a0, a1, a2, b.M[0], b.M[1], b.M[2]);
}

// pass 1 (fill == 0) counts accepted particles per (point, stencil cell) and marks
// is_md / is_reac; pass 2 (fill == 1) writes them at offset[point*ns + j]
template <class F>
__global__ void __launch_bounds__(NL_THREADS)
    k_nl_pass(Box b, PView<F64> pts, PView<F> sol, const int32_t *__restrict__ start,
              const int32_t *__restrict__ count, const int32_t *__restrict__ st, int ns,
              double radiussq, double L0, double L1, double L2, int do_lj, LJ lj,
              int32_t *__restrict__ cnt, uint8_t *__restrict__ is_md, uint8_t *__restrict__ is_reac,
              int fill, int32_t nmax, int32_t *__restrict__ list)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= ns) return;
    double x[3] = {pts.pos[i], pts.pos[pts.n + i], pts.pos[2 * pts.n + i]};
    int si = pts.species[i];
    int key = stencil_key(b, x, st, j);
    int s = start[key], n = count[key];
    int64_t slot = (int64_t)i * ns + j;
    int off = fill ? cnt[slot] : 0;
    int k = 0;
    bool md = false, noreac = false;
    for (int q = s; q < s + n; ++q) {
        int sp = (int)sol.species[q];
        if (sp <= 0) continue;
        double y0 = (double)sol.pos[q], y1 = (double)sol.pos[sol.n + q], y2 = (double)sol.pos[2 * sol.n + q];
        double d0 = min_image(x[0] - y0, L0), d1 = min_image(x[1] - y1, L1), d2 = min_image(x[2] - y2, L2);
        double rsq = (d0 * d0 + d1 * d1) + d2 * d2;
        if (rsq < radiussq) {
            md = true;
            if (do_lj && rsq <= lj.cut_sq[(sp - 1) + lj.n1 * (si - 1)]) noreac = true;
            if (fill) {
                int pos = off + k;
                if (pos < nmax) list[(int64_t)i * nmax + pos] = q;
            }
            ++k;
        }
    }
    if (!fill) {
        cnt[slot] = k;
        if (md) {
            is_md[key] = 1;
            if (noreac) is_reac[key] = 0;
        }
    }
}

// per point: exclusive scan of the stencil counts in place; n[i] = total; flags overflow
__global__ void __launch_bounds__(1024)
    k_nl_scan(int32_t *cnt, int ns, int32_t nmax, int32_t *__restrict__ nout, int32_t *__restrict__ overflow)
{
    __shared__ int lds[16];
    int i = blockIdx.x;
    int32_t *c = cnt + (int64_t)i * ns;
    int carry = 0;
    for (int base = 0; base < ns; base += blockDim.x) {
        int q = base + threadIdx.x;
        int v = q < ns ? c[q] : 0;
        int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
        int inc = wave_inclusive_scan(v);
        if (lane == 63) lds[wid] = inc;
        __syncthreads();
        int off = 0, tot = 0;
        for (int w = 0; w < nw; ++w) {
            int t = lds[w];
            if (w < wid) off += t;
            tot += t;
        }
        __syncthreads();
        if (q < ns) c[q] = carry + off + inc - v;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        // the reference stops here on overflow (neighbor_list.f90:135-137); the count is
        // clamped so that no consumer reads past the list capacity
        nout[i] = carry > nmax ? nmax : carry;
        if (carry > nmax) atomicOr(overflow, 1);
    }
}

__global__ void k_init_cells(int ncells, uint8_t *is_md, uint8_t *is_reac)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    is_md[c] = 0;
    is_reac[c] = 1;
}

// compute_force part 1 (neighbor_list.f90:180-188): PAST_MD <- MD, MD cleared
template <class F>
__global__ void k_roll_md(int64_t n, typename F::small *flags)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int f = (int)flags[i];
    f = (f & MD_MASK) ? (f | PAST_MD_MASK) : (f & ~PAST_MD_MASK);
    flags[i] = (typename F::small)(f & ~MD_MASK);
}

// compute_force part 2 (neighbor_list.f90:191-222): one thread per list entry of the
// colloids c0 .. c0 + gridDim.y - 1.  With one colloid per launch (ordered launches) each
// solvent particle is in the list at most once, so its force is updated without atomics,
// in the reference's colloid order, bit for bit; a launch over several colloids adds
// with atomics (the order of two colloids' contributions to one particle is then not
// fixed).  The colloid force and energy go to per-block partials, summed in block order by
// k_force_nl_sum.
template <class F>
__global__ void __launch_bounds__(256)
    k_force_nl(PView<F64> c, PView<F> sol, const int32_t *__restrict__ list,
               const int32_t *__restrict__ nl, int32_t nmax, double L0, double L1, double L2, LJ lj,
               double *__restrict__ part, int c0, int atomic_add)
{
    using R = typename F::real;
    __shared__ double red[16];
    const int i = c0 + (int)blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int n_i = min(nl[i], nmax);
    double f0 = 0, f1 = 0, f2 = 0, e = 0;
    const double x0 = c.pos[i], x1 = c.pos[c.n + i], x2 = c.pos[2 * c.n + i];
    const int s1 = c.species[i];
    if (k < n_i) {
        const int idx = list[(int64_t)i * nmax + k];
        const int s2 = (int)sol.species[idx];
        if (s2 > 0) {
            const double d0 = min_image(x0 - (double)sol.pos[idx], L0);
            const double d1 = min_image(x1 - (double)sol.pos[sol.n + idx], L1);
            const double d2 = min_image(x2 - (double)sol.pos[2 * sol.n + idx], L2);
            const double r_sq = (d0 * d0 + d1 * d1) + d2 * d2;
            const int pidx = (s2 - 1) + lj.n1 * (s1 - 1);
            if (r_sq <= lj.cut_sq[pidx]) {
                const double fr = lj_force_over_r(r_sq, lj.eps[pidx], lj.sigma[pidx]);
                f0 = fr * d0;
                f1 = fr * d1;
                f2 = fr * d2;
                e = lj_energy(r_sq, lj.eps[pidx], lj.sigma[pidx], lj.cut_energy[pidx]);
                if (atomic_add) {
                    atomicAdd(&sol.force[idx], (R)(-f0));
                    atomicAdd(&sol.force[sol.n + idx], (R)(-f1));
                    atomicAdd(&sol.force[2 * sol.n + idx], (R)(-f2));
                } else {
                    sol.force[idx] = sol.force[idx] - (R)f0;
                    sol.force[sol.n + idx] = sol.force[sol.n + idx] - (R)f1;
                    sol.force[2 * sol.n + idx] = sol.force[2 * sol.n + idx] - (R)f2;
                }
                const int mask = MD_MASK | (s1 == 1 ? CATALYZED_MASK : 0);
                atomic_or_flag(&sol.flags[idx], mask);
            }
        }
    }
    const double b0 = block_sum(f0, red), b1 = block_sum(f1, red), b2 = block_sum(f2, red),
                 be = block_sum(e, red);
    if (threadIdx.x == 0) {
        double *pp = part + ((int64_t)i * gridDim.x + blockIdx.x) * 4;
        pp[0] = b0;
        pp[1] = b1;
        pp[2] = b2;
        pp[3] = be;
    }
}

// the colloid force and energy of compute_force: colloid blockIdx.x's block partials in
// block order, added to the force it holds (one thread per colloid)
__global__ void k_force_nl_sum(PView<F64> c, const double *__restrict__ part, int nb, double *__restrict__ e_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    double a0 = 0, a1 = 0, a2 = 0, ae = 0;
    for (int q = 0; q < nb; ++q) {
        const double *pp = part + ((int64_t)i * nb + q) * 4;
        a0 += pp[0];
        a1 += pp[1];
        a2 += pp[2];
        ae += pp[3];
    }
    c.force[i] = c.force[i] + a0;
    c.force[c.n + i] = c.force[c.n + i] + a1;
    c.force[2 * c.n + i] = c.force[2 * c.n + i] + a2;
    e_out[i] = ae;
}

__global__ void k_copy_f64(const double *__restrict__ a, double *__restrict__ b, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// sum of n doubles in index order by one thread (the small colloid count)
__global__ void k_sum_serial(const double *a, int64_t n, double *out)
{
    if (threadIdx.x || blockIdx.x) return;
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s = s + a[i];
    out[0] = s;
}

// ---- driver helpers -------------------------------------------------------------
// flag_particles_nl (single_dimer_pbc.f90:493-509)
template <class F>
__global__ void k_flag_nl(PView<F> sol, PView<F64> c, const int32_t *__restrict__ list,
                          const int32_t *__restrict__ nl, int32_t nmax, double L0, double L1, double L2,
                          double cut_sq_c)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= min(nl[0], nmax)) return;
    int r = list[k];
    if ((int)sol.species[r] != 1) return;
    double d0 = min_image(c.pos[0] - (double)sol.pos[r], L0);
    double d1 = min_image(c.pos[c.n] - (double)sol.pos[sol.n + r], L1);
    double d2 = min_image(c.pos[2 * c.n] - (double)sol.pos[2 * sol.n + r], L2);
    if ((d0 * d0 + d1 * d1) + d2 * d2 < cut_sq_c)
        sol.flags[r] = (typename F::small)((int)sol.flags[r] | REAC_MASK);
}

// change_species (single_dimer_pbc.f90:511-542)
template <class F>
__global__ void k_change_species(PView<F> sol, PView<F64> c, double L0, double L1, double L2,
                                 double cut_sq_c, double cut_sq_n, double prob,
                                 const double *__restrict__ draws, uint64_t seed, uint64_t step)
{
    int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= sol.n) return;
    int f = (int)sol.flags[m];
    if (!(f & REAC_MASK)) return;
    double y0 = (double)sol.pos[m], y1 = (double)sol.pos[sol.n + m], y2 = (double)sol.pos[2 * sol.n + m];
    double a0 = min_image(c.pos[0] - y0, L0), a1 = min_image(c.pos[c.n] - y1, L1),
           a2 = min_image(c.pos[2 * c.n] - y2, L2);
    double dc = (a0 * a0 + a1 * a1) + a2 * a2;
    double b0 = min_image(c.pos[1] - y0, L0), b1 = min_image(c.pos[c.n + 1] - y1, L1),
           b2 = min_image(c.pos[2 * c.n + 1] - y2, L2);
    double dn = (b0 * b0 + b1 * b1) + b2 * b2;
    if (dc > cut_sq_c && dn > cut_sq_n) {
        double u;
        if (draws) {
            u = draws[m];
        } else {
            uint64_t a, b;
            threefry2x64(step, (uint64_t)sol.id[m], seed, STREAM_CHANGE, a, b);
            u = u01(a);
        }
        if (u <= prob) sol.species[m] = (typename F::small)2;
        sol.flags[m] = (typename F::small)(f & ~REAC_MASK);
    }
}

// flag_particles_nl of the fixed-sphere driver
// (setup_single_catalytic_fixed_sphere.f90:311-329): a species-1 list particle inside
// the cutoff draws u (list slot k: draws[k], or threefry((step, id), (seed, CHANGE)))
// and is flagged when u <= prob
template <class F>
__global__ void k_flag_nl_prob(PView<F> sol, PView<F64> c, const int32_t *__restrict__ list,
                               const int32_t *__restrict__ nl, int32_t nmax, double L0, double L1, double L2,
                               double cut_sq_c, double prob, const double *__restrict__ draws, uint64_t seed,
                               uint64_t step)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= min(nl[0], nmax)) return;
    int r = list[k];
    if ((int)sol.species[r] != 1) return;
    double d0 = min_image(c.pos[0] - (double)sol.pos[r], L0);
    double d1 = min_image(c.pos[c.n] - (double)sol.pos[sol.n + r], L1);
    double d2 = min_image(c.pos[2 * c.n] - (double)sol.pos[2 * sol.n + r], L2);
    if ((d0 * d0 + d1 * d1) + d2 * d2 < cut_sq_c) {
        double u;
        if (draws) {
            u = draws[k];
        } else {
            uint64_t a, b;
            threefry2x64(step, (uint64_t)sol.id[r], seed, STREAM_CHANGE, a, b);
            u = u01(a);
        }
        if (u <= prob) sol.flags[r] = (typename F::small)((int)sol.flags[r] | REAC_MASK);
    }
}

// change_species of the fixed-sphere driver (setup_single_catalytic_fixed_sphere.f90:331-349)
template <class F>
__global__ void k_change_species_sphere(PView<F> sol, PView<F64> c, double L0, double L1, double L2,
                                        double cut_sq_c)
{
    int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= sol.n) return;
    int f = (int)sol.flags[m];
    if (!(f & REAC_MASK)) return;
    double a0 = min_image(c.pos[0] - (double)sol.pos[m], L0);
    double a1 = min_image(c.pos[c.n] - (double)sol.pos[sol.n + m], L1);
    double a2 = min_image(c.pos[2 * c.n] - (double)sol.pos[2 * sol.n + m], L2);
    if ((a0 * a0 + a1 * a1) + a2 * a2 > cut_sq_c) {
        sol.species[m] = (typename F::small)2;
        sol.flags[m] = (typename F::small)(f & ~REAC_MASK);
    }
}

// refuel (single_dimer_pbc.f90:544-565)
template <class F>
__global__ void k_refuel(PView<F> sol, PView<F64> c, double L0, double L1, double L2, double far)
{
    int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= sol.n) return;
    if ((int)sol.species[m] != 2) return;
    double y0 = (double)sol.pos[m], y1 = (double)sol.pos[sol.n + m], y2 = (double)sol.pos[2 * sol.n + m];
    double a0 = min_image(c.pos[0] - y0, L0), a1 = min_image(c.pos[c.n] - y1, L1),
           a2 = min_image(c.pos[2 * c.n] - y2, L2);
    double b0 = min_image(c.pos[1] - y0, L0), b1 = min_image(c.pos[c.n + 1] - y1, L1),
           b2 = min_image(c.pos[2 * c.n + 1] - y2, L2);
    if ((a0 * a0 + a1 * a1) + a2 * a2 > far && (b0 * b0 + b1 * b1) + b2 * b2 > far)
        sol.species[m] = (typename F::small)1;
}

// ---------------------------------------------------------------------------------
template <class F>
int update_list_impl(const rmpcd_box_t *box, rmpcd_cells_t *cells, const rmpcd_particles_t *points,
                     const rmpcd_particles_t *solvent, double radius, const int32_t *stencil, int ns,
                     const rmpcd_lj_t *lj, int32_t nmax, int32_t *list, int32_t *n)
{
    Box b = make_box(box);
    int np = (int)points->n;
    int32_t *cnt = (int32_t *)workspace(WS_NLCNT, sizeof(int32_t) * (size_t)np * ns);
    int32_t *flag = (int32_t *)workspace(WS_FLAG, 64);
    int32_t *host = (int32_t *)pinned(64);
    if (!cnt || !flag || !host) return set_error(RMPCD_ERR_HIP, "update_list: workspace allocation failed");
    hipStream_t s = stream();
    double L0 = b.L[0] * b.a, L1 = b.L[1] * b.a, L2 = b.L[2] * b.a;  // neighbor_list.f90:92
    LJ l{};
    if (lj) l = make_lj(lj);
    k_init_cells<<<blocks_for(b.ncells, 256), 256, 0, s>>>(b.ncells, cells->is_md, cells->is_reac);
    RMPCD_HIP(hipMemsetAsync(flag, 0, sizeof(int32_t), s));
    dim3 grid(blocks_for(ns, NL_THREADS), np);
    k_nl_pass<F><<<grid, NL_THREADS, 0, s>>>(b, view<F64>(points), view<F>(solvent), cells->cell_start,
                                             cells->cell_count, stencil, ns, radius * radius, L0, L1, L2,
                                             lj ? 1 : 0, l, cnt, cells->is_md, cells->is_reac, 0, nmax, list);
    RMPCD_LAUNCHED();
    k_nl_scan<<<np, 1024, 0, s>>>(cnt, ns, nmax, n, flag);
    RMPCD_LAUNCHED();
    k_nl_pass<F><<<grid, NL_THREADS, 0, s>>>(b, view<F64>(points), view<F>(solvent), cells->cell_start,
                                             cells->cell_count, stencil, ns, radius * radius, L0, L1, L2,
                                             lj ? 1 : 0, l, cnt, cells->is_md, cells->is_reac, 1, nmax, list);
    RMPCD_LAUNCHED();
    RMPCD_HIP(hipMemcpyAsync(host, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RMPCD_HIP(hipStreamSynchronize(s));
    if (host[0]) return set_error(RMPCD_ERR_OVERFLOW, "maximum of neighbor list reached");
    return RMPCD_OK;
}

constexpr int FORCE_ORDERED = 64;  // colloids of compute_force taken one launch each

template <class F>
int force_impl(rmpcd_particles_t *ps1, rmpcd_particles_t *ps2, const int32_t *list, const int32_t *nl,
               int32_t nmax, const double *L, const rmpcd_lj_t *lj, double *e)
{
    int np = (int)ps1->n;
    unsigned nb = blocks_for(nmax, 256);
    double *part = (double *)workspace(WS_PART, sizeof(double) * 4 * (size_t)nb * np);
    double *eb = (double *)workspace(WS_RED, sizeof(double) * (np + 1));
    double *host = (double *)pinned(64);
    if (!part || !eb || !host) return set_error(RMPCD_ERR_HIP, "compute_force: workspace allocation failed");
    hipStream_t s = stream();
    PView<F> sol = view<F>(ps2);
    k_roll_md<F><<<blocks_for(sol.n, 256), 256, 0, s>>>(sol.n, sol.flags);
    RMPCD_LAUNCHED();
    // up to FORCE_ORDERED colloids, one launch per colloid: the solvent forces are then
    // updated in the reference's colloid order, without atomics
    const LJ l = make_lj(lj);
    if (np <= FORCE_ORDERED) {
        for (int i = 0; i < np; ++i)
            k_force_nl<F><<<dim3(nb, 1), 256, 0, s>>>(view<F64>(ps1), sol, list, nl, nmax, L[0], L[1], L[2], l, part,
                                                      i, 0);
    } else {
        k_force_nl<F><<<dim3(nb, np), 256, 0, s>>>(view<F64>(ps1), sol, list, nl, nmax, L[0], L[1], L[2], l, part, 0,
                                                   1);
    }
    RMPCD_LAUNCHED();
    k_force_nl_sum<<<blocks_for(np, 64), 64, 0, s>>>(view<F64>(ps1), part, (int)nb, eb);
    RMPCD_LAUNCHED();
    k_sum_serial<<<1, 64, 0, s>>>(eb, np, eb + np);
    RMPCD_LAUNCHED();
    RMPCD_HIP(hipMemcpyAsync(host, eb + np, sizeof(double), hipMemcpyDeviceToHost, s));
    RMPCD_HIP(hipStreamSynchronize(s));
    *e = host[0];
    return RMPCD_OK;
}

}  // namespace rmpcd

using namespace rmpcd;

extern "C" {

int rmpcd_update_list(const rmpcd_box_t *box, rmpcd_cells_t *cells, const rmpcd_particles_t *points,
                      const rmpcd_particles_t *solvent, double radius, const int32_t *stencil,
                      int32_t n_stencil, const rmpcd_lj_t *lj, int32_t nmax, int32_t *list, int32_t *n)
{
    int rc = check_box(box, "update_list");
    if (!rc) rc = check_particles(points, "update_list", false);
    if (!rc) rc = check_particles(solvent, "update_list", false);
    if (rc) return rc;
    RMPCD_CHECK(points->form == RMPCD_FP64, "update_list: points (colloids) must use the FP64 form");
    RMPCD_CHECK(points->n <= 65535, "update_list: at most 65535 points");
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start && cells->is_md && cells->is_reac,
                "update_list: null cell arrays");
    RMPCD_CHECK(stencil && n_stencil > 0 && list && n && nmax > 0, "update_list: bad list arguments");
    return RMPCD_DISPATCH_FORM(solvent, update_list_impl, box, cells, points, solvent, radius, stencil,
                               n_stencil, lj, nmax, list, n);
}

int rmpcd_compute_force(rmpcd_particles_t *ps1, rmpcd_particles_t *ps2, const int32_t *list,
                        const int32_t *n, int32_t nmax, const double L[3], const rmpcd_lj_t *lj, double *e)
{
    int rc = check_particles(ps1, "compute_force", true);
    if (!rc) rc = check_particles(ps2, "compute_force", true);
    if (rc) return rc;
    RMPCD_CHECK(ps1->form == RMPCD_FP64, "compute_force: colloids must use the FP64 form");
    RMPCD_CHECK(ps1->n <= 65535, "compute_force: at most 65535 colloids");
    RMPCD_CHECK(list && n && nmax > 0 && L && lj && e, "compute_force: null argument");
    return RMPCD_DISPATCH_FORM(ps2, force_impl, ps1, ps2, list, n, nmax, L, lj, e);
}

int rmpcd_compute_force_n2(rmpcd_particles_t *ps, const double L[3], const rmpcd_lj_t *lj, double *e)
{
    int rc = check_particles(ps, "compute_force_n2", true);
    if (rc) return rc;
    RMPCD_CHECK(ps->form == RMPCD_FP64, "compute_force_n2: colloids must use the FP64 form");
    RMPCD_CHECK(L && lj && e, "compute_force_n2: null argument");
    int64_t n = ps->n;
    double *fnew = (double *)workspace(WS_AUX, sizeof(double) * 3 * n);
    double *ep = (double *)workspace(WS_PART, sizeof(double) * (n + 1));
    double *host = (double *)pinned(64);
    if (!fnew || !ep || !host) return set_error(RMPCD_ERR_HIP, "compute_force_n2: workspace allocation failed");
    hipStream_t s = stream();
    const PView<F64> c = view<F64>(ps);
    k_force_n2_blocks<256><<<(unsigned)n, 256, 0, s>>>(c.pos, c.species, n, c.force, L[0], L[1], L[2], make_lj(lj), fnew,
                                                       ep);
    RMPCD_LAUNCHED();
    k_copy_f64<<<blocks_for(3 * n, 256), 256, 0, s>>>(fnew, (double *)ps->force, 3 * n);
    RMPCD_LAUNCHED();
    k_sum_block<256><<<1, 256, 0, s>>>(ep, n, ep + n);
    RMPCD_LAUNCHED();
    RMPCD_HIP(hipMemcpyAsync(host, ep + n, sizeof(double), hipMemcpyDeviceToHost, s));
    RMPCD_HIP(hipStreamSynchronize(s));
    *e = host[0];
    return RMPCD_OK;
}

int rmpcd_flag_particles_nl(rmpcd_particles_t *solvent, const rmpcd_particles_t *colloids,
                            const int32_t *list, const int32_t *n, int32_t nmax, const double edges[3],
                            double cut_sq_c)
{
    int rc = check_particles(solvent, "flag_particles_nl", false);
    if (!rc) rc = check_particles(colloids, "flag_particles_nl", false);
    if (rc) return rc;
    RMPCD_CHECK(colloids->form == RMPCD_FP64, "flag_particles_nl: colloids must use the FP64 form");
    RMPCD_CHECK(list && n && edges && nmax > 0, "flag_particles_nl: null argument");
    // the list length is read on the device; launch for the list capacity of colloid 0
    int64_t cap = nmax;
    hipStream_t s = stream();
    if (solvent->form == RMPCD_FP64)
        k_flag_nl<F64><<<blocks_for(cap, 256), 256, 0, s>>>(view<F64>(solvent), view<F64>(colloids), list, n,
                                                            nmax, edges[0], edges[1], edges[2], cut_sq_c);
    else
        k_flag_nl<F32><<<blocks_for(cap, 256), 256, 0, s>>>(view<F32>(solvent), view<F64>(colloids), list, n,
                                                            nmax, edges[0], edges[1], edges[2], cut_sq_c);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

int rmpcd_change_species(rmpcd_particles_t *solvent, const rmpcd_particles_t *colloids, const double edges[3],
                         double cut_sq_c, double cut_sq_n, double prob, const double *draws, uint64_t seed,
                         uint64_t step)
{
    int rc = check_particles(solvent, "change_species", false);
    if (!rc) rc = check_particles(colloids, "change_species", false);
    if (rc) return rc;
    RMPCD_CHECK(colloids->form == RMPCD_FP64 && colloids->n >= 2, "change_species: needs two FP64 colloids");
    RMPCD_CHECK(edges && (draws || solvent->id), "change_species: null argument");
    hipStream_t s = stream();
    if (solvent->form == RMPCD_FP64)
        k_change_species<F64><<<blocks_for(solvent->n, 256), 256, 0, s>>>(
            view<F64>(solvent), view<F64>(colloids), edges[0], edges[1], edges[2], cut_sq_c, cut_sq_n, prob,
            draws, seed, step);
    else
        k_change_species<F32><<<blocks_for(solvent->n, 256), 256, 0, s>>>(
            view<F32>(solvent), view<F64>(colloids), edges[0], edges[1], edges[2], cut_sq_c, cut_sq_n, prob,
            draws, seed, step);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

int rmpcd_flag_particles_nl_prob(rmpcd_particles_t *solvent, const rmpcd_particles_t *colloids,
                                 const int32_t *list, const int32_t *n, int32_t nmax, const double edges[3],
                                 double cut_sq_c, double prob, const double *draws, uint64_t seed, uint64_t step)
{
    int rc = check_particles(solvent, "flag_particles_nl", false);
    if (!rc) rc = check_particles(colloids, "flag_particles_nl", false);
    if (rc) return rc;
    RMPCD_CHECK(colloids->form == RMPCD_FP64, "flag_particles_nl: colloids must use the FP64 form");
    RMPCD_CHECK(list && n && edges && nmax > 0, "flag_particles_nl: null argument");
    RMPCD_CHECK(draws || solvent->id, "flag_particles_nl: null argument");
    hipStream_t s = stream();
    if (solvent->form == RMPCD_FP64)
        k_flag_nl_prob<F64><<<blocks_for(nmax, 256), 256, 0, s>>>(view<F64>(solvent), view<F64>(colloids), list,
                                                                  n, nmax, edges[0], edges[1], edges[2],
                                                                  cut_sq_c, prob, draws, seed, step);
    else
        k_flag_nl_prob<F32><<<blocks_for(nmax, 256), 256, 0, s>>>(view<F32>(solvent), view<F64>(colloids), list,
                                                                  n, nmax, edges[0], edges[1], edges[2],
                                                                  cut_sq_c, prob, draws, seed, step);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

int rmpcd_change_species_sphere(rmpcd_particles_t *solvent, const rmpcd_particles_t *colloids,
                                const double edges[3], double cut_sq_c)
{
    int rc = check_particles(solvent, "change_species", false);
    if (!rc) rc = check_particles(colloids, "change_species", false);
    if (rc) return rc;
    RMPCD_CHECK(colloids->form == RMPCD_FP64 && colloids->n >= 1, "change_species: needs an FP64 colloid");
    RMPCD_CHECK(edges != nullptr, "change_species: null edges");
    hipStream_t s = stream();
    if (solvent->form == RMPCD_FP64)
        k_change_species_sphere<F64><<<blocks_for(solvent->n, 256), 256, 0, s>>>(
            view<F64>(solvent), view<F64>(colloids), edges[0], edges[1], edges[2], cut_sq_c);
    else
        k_change_species_sphere<F32><<<blocks_for(solvent->n, 256), 256, 0, s>>>(
            view<F32>(solvent), view<F64>(colloids), edges[0], edges[1], edges[2], cut_sq_c);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

int rmpcd_refuel(rmpcd_particles_t *solvent, const rmpcd_particles_t *colloids, const double edges[3], int32_t L1)
{
    int rc = check_particles(solvent, "refuel", false);
    if (!rc) rc = check_particles(colloids, "refuel", false);
    if (rc) return rc;
    RMPCD_CHECK(colloids->form == RMPCD_FP64 && colloids->n >= 2, "refuel: needs two FP64 colloids");
    RMPCD_CHECK(edges != nullptr, "refuel: null edges");
    // far = (L(1)*0.45)**2 evaluated in default (single precision) REAL (single_dimer_pbc.f90:551)
    float farf = (float)L1 * 0.45f;
    farf = farf * farf;
    double far = (double)farf;
    hipStream_t s = stream();
    if (solvent->form == RMPCD_FP64)
        k_refuel<F64><<<blocks_for(solvent->n, 256), 256, 0, s>>>(view<F64>(solvent), view<F64>(colloids),
                                                                  edges[0], edges[1], edges[2], far);
    else
        k_refuel<F32><<<blocks_for(solvent->n, 256), 256, 0, s>>>(view<F32>(solvent), view<F64>(colloids),
                                                                  edges[0], edges[1], edges[2], far);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

}  // extern "C"
