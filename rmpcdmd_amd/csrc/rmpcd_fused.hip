// rmpcd_fused.hip -- the single-dimer MD step and collision cycle as fused device-side
// kernels over every replica at once (blockIdx.y = replica), captured as one hipGraph
// per collision cycle.
//
// Per MD step (programs/single_dimer_pbc.f90:306-383):
//   stream    all solvent: x += dt v (+ dt^2/2 f where the MD bit says a force exists)
//   list_a    neighbour-list members: zero the next force buffer and roll the MD bits
//             (each particle once, by an epoch claim), exact displacement of the members;
//             the last block moves the colloids (Verlet + RATTLE) and takes the skin
//             decision on the device
//   exact     only when the bound below is inconclusive: the exact maximum displacement
//   sort      only when the skin test fires: key+count, scan, place, stable gather, list
//   force     colloid-solvent LJ over the list; the last block adds the colloid forces,
//             colloid-colloid LJ, colloid Verlet and RATTLE
//   list_b    list members: velocity Verlet, catalytic flagging and conversion
// Per collision: the same sort with a fresh grid shift, then SRD + refuel (or the bulk
// reaction) per cell.
//
// Why list-only work is exact: forces only reach particles in a colloid's list, the
// list keeps every particle within max_cut + skin, and the skin test rebuilds it before
// any particle that carried a force, an MD/PAST_MD bit or the REAC bit can leave it, so
// the reference's full passes (md_vel, the MD-bit roll, force zeroing, change_species)
// leave every other particle unchanged.  Skin decision without a full pass: between two
// sorts a particle outside the list moves ballistically, so its displacement after k
// steps is at most k (dt |v|max (1 + 1e-6) + rounding slack), with |v|max taken over all
// particles at the last sort or collision; list members are measured exactly.  When
// that bound stays below the threshold the reference's test is false too; otherwise the
// exact pass decides.
#include <math.h>
#include <string.h>

#include <vector>

#include "rmpcd_sim.hpp"

namespace rmpcd {

constexpr int FT = 256;            // threads per block
constexpr int PS = 32768;          // partial-sum slots per replica
constexpr int RED_BLOCKS = 1024;   // blocks per replica for reductions over particles
constexpr int STREAM_BLOCKS = 8192;
constexpr int CT = 128;            // cells per block in the collision kernel
constexpr int CTILE = 2048;        // staged particles per collision block
constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T * SCAN_I;

enum { D_LISTA = 0, D_EXACT, D_KEY, D_FORCE, D_LISTB, D_GATHER, D_COLLIDE };

template <class F>
struct FB {
    typename F::real *pos[2], *vel[2], *force[4], *pos_old;
    int32_t *image[2], *id[2];
    typename F::small *species[2], *flags[2];
    int32_t *id_to_idx, *claim, *key, *slot, *tmp;
    int32_t *count[2], *start;
    uint8_t *is_md, *is_reac;
    int32_t *list, *nlist;
    const int32_t *stencil, *lut;
    double *cpos, *cpos_old, *cpos_rattle, *cvel, *cforce[2], *cmass;
    int32_t *cimage, *cspecies;
    double *part;
    int32_t *scan_sums;
    int64_t scan_stride;
    const RepDev *rep;
    CtlDev *ctl;
};

struct FP {
    double dt, dt_sq, skin, d, prob, bulk_rate, tau, slack;
    double mass[2];
    int nmax, n_stencil, bulk, pad;
    double radius_sq;
    LJ sc, cc;
};

// last-block bookkeeping: every block stores its partial, the last one to arrive sees them
__device__ inline bool arrive_last(uint32_t *counter, unsigned nblocks, bool *s_last)
{
    if (threadIdx.x == 0) {
        __threadfence();
        *s_last = atomicAdd(counter, 1u) == nblocks - 1;
    }
    __syncthreads();
    if (*s_last) __threadfence();
    return *s_last;
}

__device__ inline int64_t gstride() { return (int64_t)gridDim.x * blockDim.x; }
__device__ inline int64_t gindex() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

// ---- colloid helpers (one thread, replica r; the arithmetic of k_colloid_pos,
// k_rattle_pos, k_max_disp, k_force_nl's last block, k_force_n2, k_colloid_vel and
// k_rattle_vel, so that both modes agree bit for bit) --------------------------------
__device__ inline void rattle_pos_2(double *cp, double *cv, const double *pr, const int32_t *sp,
                                    const double *mass, double d, double dt, const double *L)
{
    double a1[3] = {pr[0], pr[2], pr[4]}, a2[3] = {pr[1], pr[3], pr[5]};
    double x1[3] = {cp[0], cp[2], cp[4]}, x2[3] = {cp[1], cp[3], cp[5]};
    double r[3], s[3];
    for (int k = 0; k < 3; ++k) {
        r[k] = min_image(a1[k] - a2[k], L[k]);
        s[k] = min_image(x1[k] - x2[k], L[k]);
    }
    double m1 = mass[(sp[0] - 1) & 1], m2 = mass[(sp[1] - 1) & 1];
    double inv_mass = 1 / m1 + 1 / m2;
    double rsq = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
    double ssq = (s[0] * s[0] + s[1] * s[1]) + s[2] * s[2];
    double rs = (r[0] * s[0] + r[1] * s[1]) + r[2] * s[2];
    double g = rs - sqrt(rs * rs - rsq * (ssq - d * d));
    g = g / ((dt * inv_mass) * rsq);
    for (int k = 0; k < 3; ++k) {
        cp[2 * k] = cp[2 * k] - ((g * dt) * r[k]) / m1;
        cp[2 * k + 1] = cp[2 * k + 1] + ((g * dt) * r[k]) / m2;
        cv[2 * k] = cv[2 * k] - (g * r[k]) / m1;
        cv[2 * k + 1] = cv[2 * k + 1] + (g * r[k]) / m2;
    }
}

__device__ inline void rattle_vel_2(const double *cp, double *cv, const int32_t *sp, const double *mass,
                                    double d, const double *L)
{
    double m1 = mass[(sp[0] - 1) & 1], m2 = mass[(sp[1] - 1) & 1];
    double inv_mass = 1 / m1 + 1 / m2;
    double s[3], dv[3];
    for (int k = 0; k < 3; ++k) {
        s[k] = min_image(cp[2 * k] - cp[2 * k + 1], L[k]);
        dv[k] = cv[2 * k] - cv[2 * k + 1];
    }
    double kk = ((dv[0] * s[0] + dv[1] * s[1]) + dv[2] * s[2]) / ((d * d) * inv_mass);
    for (int k = 0; k < 3; ++k) {
        cv[2 * k] = cv[2 * k] - (kk * s[k]) / m1;
        cv[2 * k + 1] = cv[2 * k + 1] + (kk * s[k]) / m2;
    }
}

// ---- stream ------------------------------------------------------------------------
template <class F>
__global__ void __launch_bounds__(FT) k_fz_stream(FB<F> B, FP P)
{
    using R = typename F::real;
    const int r = blockIdx.y;
    const RepDev &rp = B.rep[r];
    const int cur = B.ctl[r].cur, fp = B.ctl[r].fpar;
    const int64_t n = rp.n;
    R *__restrict__ pos = B.pos[cur] + 3 * rp.p_off;
    const R *__restrict__ vel = B.vel[cur] + 3 * rp.p_off;
    const R *__restrict__ f = B.force[2 * cur + fp] + 3 * rp.p_off;
    const typename F::small *__restrict__ fl = B.flags[cur] + rp.p_off;
    const R rdt = (R)P.dt, dsq = (R)P.dt_sq;
    for (int64_t i = gindex(); i < n; i += gstride()) {
        const bool md = ((int)fl[i] & MD_MASK) != 0;
        R x0 = pos[i] + rdt * vel[i];
        R x1 = pos[n + i] + rdt * vel[n + i];
        R x2 = pos[2 * n + i] + rdt * vel[2 * n + i];
        if (md) {
            x0 = x0 + dsq * f[i];
            x1 = x1 + dsq * f[n + i];
            x2 = x2 + dsq * f[2 * n + i];
        }
        pos[i] = x0;
        pos[n + i] = x1;
        pos[2 * n + i] = x2;
    }
}

// ---- list_a: claims, force zeroing, MD roll, member displacement; colloid step ------
template <class F>
__global__ void __launch_bounds__(FT) k_fz_list_a(FB<F> B, FP P)
{
    using R = typename F::real;
    __shared__ double red[16];
    __shared__ bool s_last;
    const int r = blockIdx.y;
    const RepDev &rp = B.rep[r];
    CtlDev &C = B.ctl[r];
    const int cur = C.cur, fp = C.fpar;
    const int64_t n = rp.n;
    const int e = blockIdx.x * FT + threadIdx.x;
    double r2 = 0;
    if (e < 2 * P.nmax) {
        const int col = e / P.nmax, k = e - col * P.nmax;
        const int nl = min(B.nlist[2 * r + col], P.nmax);
        if (k < nl) {
            const int idx = B.list[(int64_t)(2 * r + col) * P.nmax + k];
            const unsigned epoch = (unsigned)(C.md_step * 4 + 1);
            if (atomicExch((unsigned *)&B.claim[rp.p_off + idx], epoch) != epoch) {
                R *fn = B.force[2 * cur + (1 - fp)] + 3 * rp.p_off;
                fn[idx] = 0;
                fn[n + idx] = 0;
                fn[2 * n + idx] = 0;
                typename F::small *fl = B.flags[cur] + rp.p_off;
                int f = (int)fl[idx];
                f = (f & MD_MASK) ? (f | PAST_MD_MASK) : (f & ~PAST_MD_MASK);
                fl[idx] = (typename F::small)(f & ~MD_MASK);
            }
            const R *pos = B.pos[cur] + 3 * rp.p_off;
            const R *po = B.pos_old + 3 * rp.p_off;
            double d0 = (double)(pos[idx] - po[idx]);
            double d1 = (double)(pos[n + idx] - po[n + idx]);
            double d2 = (double)(pos[2 * n + idx] - po[2 * n + idx]);
            r2 = (d0 * d0 + d1 * d1) + d2 * d2;
        }
    }
    double m = block_max(r2, red);
    if (threadIdx.x == 0) B.part[(int64_t)r * PS + blockIdx.x] = m;
    if (!arrive_last(&C.done[D_LISTA], gridDim.x, &s_last)) return;
    if (threadIdx.x != 0) return;
    double lm = 0;
    for (unsigned q = 0; q < gridDim.x; ++q) lm = fmax(lm, B.part[(int64_t)r * PS + q]);
    C.done[D_LISTA] = 0;
    C.list_max2 = lm;
    // colloid Verlet position step and RATTLE (single_dimer_pbc.f90:311-317)
    double *cp = B.cpos + 6 * r, *cv = B.cvel + 6 * r, *cr = B.cpos_rattle + 6 * r, *co = B.cpos_old + 6 * r;
    const double *cf = B.cforce[C.cfpar] + 6 * r;
    const double *cm = B.cmass + 2 * r;
    const int32_t *cs = B.cspecies + 2 * r;
    const double dt = P.dt;
    for (int q = 0; q < 6; ++q) cr[q] = cp[q];
    for (int i = 0; i < 2; ++i)
        for (int k = 0; k < 3; ++k) {
            int q = 2 * k + i;
            cp[q] = (cp[q] + dt * cv[q]) + ((dt * dt) * cf[q]) / (2 * cm[i]);
        }
    rattle_pos_2(cp, cv, cr, cs, P.mass, P.d, dt, rp.box.edges);
    // maximum displacements and the skin test (single_dimer_pbc.f90:320-323)
    double cmax = 0;
    for (int i = 0; i < 2; ++i) {
        double d0 = cp[i] - co[i], d1 = cp[2 + i] - co[2 + i], d2 = cp[4 + i] - co[4 + i];
        double q2 = (d0 * d0 + d1 * d1) + d2 * d2;
        cmax = q2 > cmax ? q2 : cmax;
    }
    double co_max = sqrt(cmax);
    C.co_max = co_max;
    int k = C.k_since + 1;
    double bound = (double)k * (P.dt * sqrt(C.vmax2) * (1.0 + 1e-6) + P.slack);
    double so_est = fmax(sqrt(lm), bound);
    if (co_max >= P.skin * 0.1) {
        C.need_sort = 1;
        C.need_exact = 0;
    } else if (so_est < P.skin * 0.89) {
        C.need_sort = 0;
        C.need_exact = 0;
    } else {
        C.need_sort = 0;
        C.need_exact = 1;
    }
    // switch(force, force_old) and zero the new colloid force (single_dimer_pbc.f90:341-348)
    C.fpar ^= 1;
    C.cfpar ^= 1;
    double *cfn = B.cforce[C.cfpar] + 6 * r;
    for (int q = 0; q < 6; ++q) cfn[q] = 0;
}

// ---- exact maximum displacement (particle_system.f90:357-378) -----------------------
template <class F>
__global__ void __launch_bounds__(FT) k_fz_exact(FB<F> B, FP P)
{
    using R = typename F::real;
    __shared__ double red[16];
    __shared__ bool s_last;
    const int r = blockIdx.y;
    CtlDev &C = B.ctl[r];
    if (!C.need_exact) return;
    const RepDev &rp = B.rep[r];
    const int64_t n = rp.n;
    const int cur = C.cur;
    const R *pos = B.pos[cur] + 3 * rp.p_off, *po = B.pos_old + 3 * rp.p_off;
    const typename F::small *sp = B.species[cur] + rp.p_off;
    double mx = 0;
    for (int64_t i = gindex(); i < n; i += gstride()) {
        if ((int)sp[i] >= 0) {
            double d0 = (double)(pos[i] - po[i]);
            double d1 = (double)(pos[n + i] - po[n + i]);
            double d2 = (double)(pos[2 * n + i] - po[2 * n + i]);
            double q2 = (d0 * d0 + d1 * d1) + d2 * d2;
            mx = q2 > mx ? q2 : mx;
        }
    }
    double m = block_max(mx, red);
    if (threadIdx.x == 0) B.part[(int64_t)r * PS + blockIdx.x] = m;
    if (!arrive_last(&C.done[D_EXACT], gridDim.x, &s_last)) return;
    if (threadIdx.x != 0) return;
    double all = 0;
    for (unsigned q = 0; q < gridDim.x; ++q) all = fmax(all, B.part[(int64_t)r * PS + q]);
    C.done[D_EXACT] = 0;
    C.need_sort = sqrt(all) >= P.skin * 0.89 ? 1 : 0;
    C.need_exact = 0;
}

// ---- sort ---------------------------------------------------------------------------
// S1: grid shift (collision sort), periodic wrap, key and arrival slot; the last block
// wraps the colloids and records their positions as pos_old
template <class F>
__global__ void __launch_bounds__(FT) k_fz_key(FB<F> B, FP P, int forced)
{
    using R = typename F::real;
    __shared__ bool s_last;
    const int r = blockIdx.y;
    CtlDev &C = B.ctl[r];
    if (!forced && !C.need_sort) return;
    const RepDev &rp = B.rep[r];
    Box b = rp.box;
    if (forced) {
        double u[3];
        uint64_t a0, a1, a2, a3;
        threefry2x64((uint64_t)C.cycle, 0, rp.seed, STREAM_SHIFT, a0, a1);
        threefry2x64((uint64_t)C.cycle, 1, rp.seed, STREAM_SHIFT, a2, a3);
        u[0] = u01(a0);
        u[1] = u01(a1);
        u[2] = u01(a2);
        for (int k = 0; k < 3; ++k) b.origin[k] = u[k] - 1;  // random_shift
    } else {
        for (int k = 0; k < 3; ++k) b.origin[k] = C.origin[k];
    }
    const int64_t n = rp.n;
    const int cur = C.cur;
    const R *pos = B.pos[cur] + 3 * rp.p_off;
    int32_t *cnt = B.count[C.cnt ^ 1] + rp.c_off;
    const int32_t *lut = B.lut + rp.lut_off;
    for (int64_t i = gindex(); i < n; i += gstride()) {
        R x[3] = {pos[i], pos[n + i], pos[2 * n + i]};
        for (int k = 0; k < 3; ++k) {
            R e = (R)b.edges[k];
            int jump = (int)floor((double)(x[k] / e));
            x[k] = x[k] - (R)jump * e;
        }
        int key;
        {
            int p0, p1, p2;
            cell_of<R>(b, x[0], x[1], x[2], p0, p1, p2);
            if ((unsigned)p0 < (unsigned)b.L[0] && (unsigned)p1 < (unsigned)b.L[1] &&
                (unsigned)p2 < (unsigned)b.L[2])
                key = lut[p0 + b.L[0] * (p1 + b.L[1] * p2)];
            else
                key = compact_key(p0, p1, p2, b.M[0], b.M[1], b.M[2]);
        }
        B.key[rp.p_off + i] = key;
        B.slot[rp.p_off + i] = atomicAdd(&cnt[key], 1);
    }
    if (!arrive_last(&C.done[D_KEY], gridDim.x, &s_last)) return;
    if (threadIdx.x != 0) return;
    C.done[D_KEY] = 0;
    for (int k = 0; k < 3; ++k) C.origin[k] = b.origin[k];
    // apply_pbc(colloids) (md.f90:202-217) and colloids%pos_old = colloids%pos
    double *cp = B.cpos + 6 * r, *co = B.cpos_old + 6 * r;
    int32_t *ci = B.cimage + 6 * r;
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 2; ++i) {
            int q = 2 * k + i;
            double x = cp[q];
            int jump = (int)floor(x / b.edges[k]);
            ci[q] += jump;
            cp[q] = x - jump * b.edges[k];
        }
    for (int q = 0; q < 6; ++q) co[q] = cp[q];
}

__device__ inline bool sort_active(const CtlDev &C, int forced) { return forced || C.need_sort; }

__device__ inline int scan_block_exclusive(int v, int *lds, int *total)
{
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
    *total = tot;
    return off + inc - v;
}

template <class F>
__global__ void __launch_bounds__(SCAN_T) k_fz_scan_tiles(FB<F> B, int forced)
{
    __shared__ int lds[SCAN_T / 64];
    const int r = blockIdx.y;
    const CtlDev &C = B.ctl[r];
    if (!sort_active(C, forced)) return;
    const RepDev &rp = B.rep[r];
    const int64_t nc = rp.box.ncells;
    const int64_t tile0 = (int64_t)blockIdx.x * SCAN_TILE;
    if (tile0 >= nc) return;  // uniform per block
    const int32_t *in = B.count[C.cnt ^ 1] + rp.c_off;
    int32_t *out = B.start + rp.c_off;
    int64_t base = tile0 + (int64_t)threadIdx.x * SCAN_I;
    int v[SCAN_I], s = 0;
    for (int k = 0; k < SCAN_I; ++k) {
        v[k] = base + k < nc ? in[base + k] : 0;
        s += v[k];
    }
    int tot;
    int off = scan_block_exclusive(s, lds, &tot);
    for (int k = 0; k < SCAN_I; ++k) {
        if (base + k < nc) out[base + k] = off;
        off += v[k];
    }
    if (threadIdx.x == 0) B.scan_sums[r * B.scan_stride + blockIdx.x] = tot;
}

template <class F>
__global__ void __launch_bounds__(1024) k_fz_scan_top(FB<F> B, int forced)
{
    __shared__ int lds[16];
    const int r = blockIdx.y;
    const CtlDev &C = B.ctl[r];
    if (!sort_active(C, forced)) return;
    const int64_t ntiles = (B.rep[r].box.ncells + SCAN_TILE - 1) / SCAN_TILE;
    int32_t *sums = B.scan_sums + r * B.scan_stride;
    int carry = 0;
    for (int64_t c0 = 0; c0 < ntiles; c0 += blockDim.x) {
        int64_t i = c0 + threadIdx.x;
        int v = i < ntiles ? sums[i] : 0;
        int tot;
        int ex = scan_block_exclusive(v, lds, &tot);
        if (i < ntiles) sums[i] = ex + carry;
        carry += tot;
    }
}

template <class F>
__global__ void __launch_bounds__(SCAN_T) k_fz_scan_add(FB<F> B, int forced)
{
    const int r = blockIdx.y;
    const CtlDev &C = B.ctl[r];
    if (!sort_active(C, forced)) return;
    const RepDev &rp = B.rep[r];
    const int64_t nc = rp.box.ncells;
    const int64_t tile0 = (int64_t)blockIdx.x * SCAN_TILE;
    if (tile0 >= nc) return;
    int32_t *out = B.start + rp.c_off;
    const int add = B.scan_sums[r * B.scan_stride + blockIdx.x];
    for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_T)
        if (tile0 + k < nc) out[tile0 + k] += add;
}

template <class F>
__global__ void __launch_bounds__(FT) k_fz_place(FB<F> B, int forced)
{
    const int r = blockIdx.y;
    const CtlDev &C = B.ctl[r];
    if (!sort_active(C, forced)) return;
    const RepDev &rp = B.rep[r];
    const int64_t n = rp.n;
    const int32_t *start = B.start + rp.c_off;
    const int32_t *key = B.key + rp.p_off, *slot = B.slot + rp.p_off;
    int32_t *tmp = B.tmp + rp.p_off;
    for (int64_t i = gindex(); i < n; i += gstride()) tmp[start[key[i]] + slot[i]] = (int32_t)i;
}

// S4: stable gather into the other buffers (wrap fused), pos_old, |v|max; clears the
// previous cell counts and the is_md/is_reac marks for the list build
template <class F>
__global__ void __launch_bounds__(FT) k_fz_gather(FB<F> B, FP P, int forced)
{
    using R = typename F::real;
    using S = typename F::small;
    __shared__ double red[16];
    __shared__ bool s_last;
    const int r = blockIdx.y;
    CtlDev &C = B.ctl[r];
    if (!sort_active(C, forced)) return;
    const RepDev &rp = B.rep[r];
    const int64_t n = rp.n, nc = rp.box.ncells;
    const int cur = C.cur, dst = 1 - cur, fp = C.fpar;
    const int64_t po3 = 3 * rp.p_off;
    const R *spos = B.pos[cur] + po3, *svel = B.vel[cur] + po3;
    const R *sf = B.force[2 * cur + fp] + po3, *sfo = B.force[2 * cur + 1 - fp] + po3;
    const int32_t *sim = B.image[cur] + po3, *sid = B.id[cur] + rp.p_off;
    const S *ssp = B.species[cur] + rp.p_off, *sfl = B.flags[cur] + rp.p_off;
    R *dpos = B.pos[dst] + po3, *dvel = B.vel[dst] + po3, *dold = B.pos_old + po3;
    R *df = B.force[2 * dst + fp] + po3, *dfo = B.force[2 * dst + 1 - fp] + po3;
    int32_t *dim = B.image[dst] + po3, *did = B.id[dst] + rp.p_off;
    S *dsp = B.species[dst] + rp.p_off, *dfl = B.flags[dst] + rp.p_off;
    const int32_t *tmp = B.tmp + rp.p_off, *key = B.key + rp.p_off;
    const int32_t *start = B.start + rp.c_off, *cnt = B.count[C.cnt ^ 1] + rp.c_off;
    int32_t *old_cnt = B.count[C.cnt] + rp.c_off;
    const int64_t lim = n > nc ? n : nc;
    double vm = 0;
    for (int64_t j = gindex(); j < lim; j += gstride()) {
        if (j < n) {
            const int32_t v = tmp[j];
            const int c = key[v];
            const int s = start[c], m = cnt[c];
            int rank = 0;
            for (int q = s; q < s + m; ++q) rank += tmp[q] < v ? 1 : 0;
            const int64_t d = s + rank;
            const int fl = (int)sfl[v];
            const bool has_f = (fl & (MD_MASK | PAST_MD_MASK)) != 0;
            double v2 = 0;
            for (int k = 0; k < 3; ++k) {
                R x = spos[k * n + v];
                R e = (R)rp.box.edges[k];
                int jump = (int)floor((double)(x / e));
                x = x - (R)jump * e;
                dpos[k * n + d] = x;
                dold[k * n + d] = x;
                dim[k * n + d] = sim[k * n + v] + jump;
                R vv = svel[k * n + v];
                dvel[k * n + d] = vv;
                v2 += (double)vv * (double)vv;
                df[k * n + d] = has_f ? sf[k * n + v] : (R)0;
                dfo[k * n + d] = has_f ? sfo[k * n + v] : (R)0;
            }
            const int32_t id = sid[v];
            did[d] = id;
            dsp[d] = ssp[v];
            dfl[d] = (S)fl;
            if ((uint32_t)id < (uint32_t)n) B.id_to_idx[rp.p_off + id] = (int32_t)d;
            vm = v2 > vm ? v2 : vm;
        }
        if (j < nc) {
            old_cnt[j] = 0;
            B.is_md[rp.c_off + j] = 0;
            B.is_reac[rp.c_off + j] = 1;
        }
    }
    double m = block_max(vm, red);
    if (threadIdx.x == 0) B.part[(int64_t)r * PS + blockIdx.x] = m;
    if (!arrive_last(&C.done[D_GATHER], gridDim.x, &s_last)) return;
    if (threadIdx.x != 0) return;
    double all = 0;
    for (unsigned q = 0; q < gridDim.x; ++q) all = fmax(all, B.part[(int64_t)r * PS + q]);
    C.done[D_GATHER] = 0;
    C.vmax2 = all;
    C.cur = dst;
    C.cnt ^= 1;
    C.k_since = 0;
    if (!forced) C.n_extra += 1;
    // need_sort stays set for the list build; the next list_a pass recomputes it
}

// S5: update_list (neighbor_list.f90:76-157) for colloid blockIdx.x of replica r, one
// workgroup: count per stencil cell, block scan, fill in stencil order
template <class F>
__global__ void __launch_bounds__(1024) k_fz_list_build(FB<F> B, FP P, int forced)
{
    __shared__ int lds[16];
    const int r = blockIdx.y, col = blockIdx.x;
    CtlDev &C = B.ctl[r];
    if (!sort_active(C, forced)) return;
    const RepDev &rp = B.rep[r];
    Box b = rp.box;
    for (int k = 0; k < 3; ++k) b.origin[k] = C.origin[k];
    const int64_t n = rp.n;
    const int cur = C.cur;
    const typename F::real *pos = B.pos[cur] + 3 * rp.p_off;
    const typename F::small *sp = B.species[cur] + rp.p_off;
    const int32_t *cnt = B.count[C.cnt] + rp.c_off, *start = B.start + rp.c_off;
    uint8_t *is_md = B.is_md + rp.c_off, *is_reac = B.is_reac + rp.c_off;
    const double *cp = B.cpos + 6 * r;
    const double x[3] = {cp[col], cp[2 + col], cp[4 + col]};
    const int si = B.cspecies[2 * r + col];
    const double L0 = b.L[0] * b.a, L1 = b.L[1] * b.a, L2 = b.L[2] * b.a;  // neighbor_list.f90:92
    int32_t *lst = B.list + (int64_t)(2 * r + col) * P.nmax;
    int carry = 0;
    for (int base = 0; base < P.n_stencil; base += blockDim.x) {
        const int j = base + threadIdx.x;
        int key = 0, s = 0, m = 0, k = 0;
        bool md = false, noreac = false;
        if (j < P.n_stencil) {
            key = stencil_key_dev(b, x, B.stencil, j);
            s = start[key];
            m = cnt[key];
            for (int q = s; q < s + m; ++q) {
                const int spc = (int)sp[q];
                if (spc <= 0) continue;
                double d0 = min_image(x[0] - (double)pos[q], L0);
                double d1 = min_image(x[1] - (double)pos[n + q], L1);
                double d2 = min_image(x[2] - (double)pos[2 * n + q], L2);
                double rsq = (d0 * d0 + d1 * d1) + d2 * d2;
                if (rsq < P.radius_sq) {
                    md = true;
                    if (rsq <= P.sc.cut_sq[(spc - 1) + P.sc.n1 * (si - 1)]) noreac = true;
                    ++k;
                }
            }
            if (md) {
                is_md[key] = 1;
                if (noreac) is_reac[key] = 0;
            }
        }
        int tot;
        int off = carry + scan_block_exclusive(k, lds, &tot);
        if (j < P.n_stencil && k > 0) {
            int w = 0;
            for (int q = s; q < s + m; ++q) {
                const int spc = (int)sp[q];
                if (spc <= 0) continue;
                double d0 = min_image(x[0] - (double)pos[q], L0);
                double d1 = min_image(x[1] - (double)pos[n + q], L1);
                double d2 = min_image(x[2] - (double)pos[2 * n + q], L2);
                double rsq = (d0 * d0 + d1 * d1) + d2 * d2;
                if (rsq < P.radius_sq) {
                    if (off + w < P.nmax) lst[off + w] = q;
                    ++w;
                }
            }
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        B.nlist[2 * r + col] = carry > P.nmax ? P.nmax : carry;
        if (carry > P.nmax) C.error = RMPCD_ERR_OVERFLOW;
    }
}

// ---- force: colloid-solvent LJ over the lists (neighbor_list.f90:191-222) -------------
template <class F>
__global__ void __launch_bounds__(FT) k_fz_force(FB<F> B, FP P)
{
    using R = typename F::real;
    __shared__ double red[16];
    __shared__ bool s_last;
    const int r = blockIdx.y;
    const RepDev &rp = B.rep[r];
    CtlDev &C = B.ctl[r];
    const int nb = (P.nmax + FT - 1) / FT;
    const int col = blockIdx.x / nb, bb = blockIdx.x - col * nb;
    const int k = bb * FT + threadIdx.x;
    const int64_t n = rp.n;
    const int cur = C.cur, fp = C.fpar;
    const double *cp = B.cpos + 6 * r;
    const double x0 = cp[col], x1 = cp[2 + col], x2 = cp[4 + col];
    const int s1 = B.cspecies[2 * r + col];
    const double *L = rp.box.edges;
    const int nl = min(B.nlist[2 * r + col], P.nmax);
    double f0 = 0, f1 = 0, f2 = 0, e = 0;
    if (k < nl) {
        const int idx = B.list[(int64_t)(2 * r + col) * P.nmax + k];
        const int s2 = (int)B.species[cur][rp.p_off + idx];
        if (s2 > 0) {
            const R *pos = B.pos[cur] + 3 * rp.p_off;
            double d0 = min_image(x0 - (double)pos[idx], L[0]);
            double d1 = min_image(x1 - (double)pos[n + idx], L[1]);
            double d2 = min_image(x2 - (double)pos[2 * n + idx], L[2]);
            double r_sq = (d0 * d0 + d1 * d1) + d2 * d2;
            int pidx = (s2 - 1) + P.sc.n1 * (s1 - 1);
            if (r_sq <= P.sc.cut_sq[pidx]) {
                double fr = lj_force_over_r(r_sq, P.sc.eps[pidx], P.sc.sigma[pidx]);
                f0 = fr * d0;
                f1 = fr * d1;
                f2 = fr * d2;
                e = lj_energy(r_sq, P.sc.eps[pidx], P.sc.sigma[pidx], P.sc.cut_energy[pidx]);
                R *force = B.force[2 * cur + fp] + 3 * rp.p_off;
                atomicAdd(&force[idx], (R)(-f0));
                atomicAdd(&force[n + idx], (R)(-f1));
                atomicAdd(&force[2 * n + idx], (R)(-f2));
                atomic_or_flag(&B.flags[cur][rp.p_off + idx], MD_MASK | (s1 == 1 ? CATALYZED_MASK : 0));
            }
        }
    }
    double b0 = block_sum(f0, red), b1 = block_sum(f1, red), b2 = block_sum(f2, red), be = block_sum(e, red);
    if (threadIdx.x == 0) {
        double *pp = B.part + (int64_t)r * PS + (int64_t)(col * nb + bb) * 4;
        pp[0] = b0;
        pp[1] = b1;
        pp[2] = b2;
        pp[3] = be;
    }
    if (!arrive_last(&C.done[D_FORCE], gridDim.x, &s_last)) return;
    if (threadIdx.x != 0) return;
    C.done[D_FORCE] = 0;
    double *cf = B.cforce[C.cfpar] + 6 * r;
    const double *cfo = B.cforce[1 - C.cfpar] + 6 * r;
    double ecol[2];
    for (int c = 0; c < 2; ++c) {
        double a0 = 0, a1 = 0, a2 = 0, ae = 0;
        for (int q = 0; q < nb; ++q) {
            const double *pp = B.part + (int64_t)r * PS + (int64_t)(c * nb + q) * 4;
            a0 += pp[0];
            a1 += pp[1];
            a2 += pp[2];
            ae += pp[3];
        }
        cf[c] = cf[c] + a0;
        cf[2 + c] = cf[2 + c] + a1;
        cf[4 + c] = cf[4 + c] + a2;
        ecol[c] = ae;
    }
    double e1 = 0;
    e1 = e1 + ecol[0];
    e1 = e1 + ecol[1];
    C.e1 = e1;
    // compute_force_n2 for the pair (neighbor_list.f90:243-262)
    {
        double d0 = min_image(cp[0] - cp[1], L[0]);
        double d1 = min_image(cp[2] - cp[3], L[1]);
        double d2 = min_image(cp[4] - cp[5], L[2]);
        double r_sq = (d0 * d0 + d1 * d1) + d2 * d2;
        const int sa = B.cspecies[2 * r], sb = B.cspecies[2 * r + 1];
        int pc = (sa - 1) + P.cc.n1 * (sb - 1), pe = (sb - 1) + P.cc.n1 * (sa - 1);
        double e2 = 0, e0 = 0;
        if (r_sq <= P.cc.cut_sq[pc]) {
            double fr = lj_force_over_r(r_sq, P.cc.eps[pe], P.cc.sigma[pe]);
            cf[0] = cf[0] + fr * d0;
            cf[2] = cf[2] + fr * d1;
            cf[4] = cf[4] + fr * d2;
            cf[1] = cf[1] - fr * d0;
            cf[3] = cf[3] - fr * d1;
            cf[5] = cf[5] - fr * d2;
            e0 = e0 + lj_energy(r_sq, P.cc.eps[pe], P.cc.sigma[pe], P.cc.cut_energy[pe]);
        }
        e2 = e2 + e0;
        e2 = e2 + 0.0;
        C.e2 = e2;
    }
    // colloid velocity step and RATTLE (single_dimer_pbc.f90:356-361)
    double *cv = B.cvel + 6 * r;
    const double *cm = B.cmass + 2 * r;
    for (int i = 0; i < 2; ++i)
        for (int q3 = 0; q3 < 3; ++q3) {
            int q = 2 * q3 + i;
            cv[q] = cv[q] + (P.dt * (cf[q] + cfo[q])) / (2 * cm[i]);
        }
    rattle_vel_2(cp, cv, B.cspecies + 2 * r, P.mass, P.d, L);
}

// ---- list_b: velocity Verlet of the members, flagging and conversion ------------------
template <class F>
__global__ void __launch_bounds__(FT) k_fz_list_b(FB<F> B, FP P)
{
    using R = typename F::real;
    using S = typename F::small;
    __shared__ bool s_last;
    const int r = blockIdx.y;
    const RepDev &rp = B.rep[r];
    CtlDev &C = B.ctl[r];
    const int cur = C.cur, fp = C.fpar;
    const int64_t n = rp.n;
    const int e = blockIdx.x * FT + threadIdx.x;
    if (e < 2 * P.nmax) {
        const int col = e / P.nmax, k = e - col * P.nmax;
        const int nl = min(B.nlist[2 * r + col], P.nmax);
        if (k < nl) {
            const int idx = B.list[(int64_t)(2 * r + col) * P.nmax + k];
            const int64_t g = rp.p_off + idx;
            const R *pos = B.pos[cur] + 3 * rp.p_off;
            const double *cp = B.cpos + 6 * r;
            const double *L = rp.box.edges;
            const double y0 = (double)pos[idx], y1 = (double)pos[n + idx], y2 = (double)pos[2 * n + idx];
            S *spc = B.species[cur] + rp.p_off;
            S *fl = B.flags[cur] + rp.p_off;
            const double cut_c = P.sc.cut_sq[0], cut_n = P.sc.cut_sq[2];
            const unsigned epoch = (unsigned)(C.md_step * 4 + 2);
            if (atomicExch((unsigned *)&B.claim[g], epoch) != epoch) {
                // md_vel (md.f90:219-237)
                const int f = (int)fl[idx];
                if (!(f & WALL_MASK)) {
                    R *vel = B.vel[cur] + 3 * rp.p_off;
                    const R *fo = B.force[2 * cur + fp] + 3 * rp.p_off;
                    const R *fold = B.force[2 * cur + 1 - fp] + 3 * rp.p_off;
                    const R rdt = (R)P.dt;
                    for (int q = 0; q < 3; ++q)
                        vel[q * n + idx] = vel[q * n + idx] + (rdt * (fo[q * n + idx] + fold[q * n + idx])) / (R)2;
                } else {
                    atomic_and_flag(&fl[idx], ~WALL_MASK);
                }
                // change_species (single_dimer_pbc.f90:511-542), draw threefry((md_step, id), (seed, CHANGE))
                if (f & REAC_MASK) {
                    double a0 = min_image(cp[0] - y0, L[0]), a1 = min_image(cp[2] - y1, L[1]),
                           a2 = min_image(cp[4] - y2, L[2]);
                    double b0 = min_image(cp[1] - y0, L[0]), b1 = min_image(cp[3] - y1, L[1]),
                           b2 = min_image(cp[5] - y2, L[2]);
                    double dc = (a0 * a0 + a1 * a1) + a2 * a2, dn = (b0 * b0 + b1 * b1) + b2 * b2;
                    if (dc > cut_c && dn > cut_n) {
                        uint64_t u0, u1;
                        threefry2x64((uint64_t)C.md_step, (uint64_t)B.id[cur][g], rp.seed, STREAM_CHANGE, u0, u1);
                        if (u01(u0) <= P.prob) spc[idx] = (S)2;
                        atomic_and_flag(&fl[idx], ~REAC_MASK);
                    }
                }
            }
            // flag_particles_nl (single_dimer_pbc.f90:493-509), colloid 1's list
            if (col == 0 && (int)spc[idx] == 1) {
                double a0 = min_image(cp[0] - y0, L[0]), a1 = min_image(cp[2] - y1, L[1]),
                       a2 = min_image(cp[4] - y2, L[2]);
                if ((a0 * a0 + a1 * a1) + a2 * a2 < cut_c) atomic_or_flag(&fl[idx], REAC_MASK);
            }
        }
    }
    if (!arrive_last(&C.done[D_LISTB], gridDim.x, &s_last)) return;
    if (threadIdx.x != 0) return;
    C.done[D_LISTB] = 0;
    C.md_step += 1;
    C.k_since += 1;
}

// ---- collision: SRD (mpcd.f90:120-180) + refuel or bulk reaction, |v|max ---------------
template <class F>
__global__ void __launch_bounds__(CT) k_fz_collide(FB<F> B, FP P)
{
    using R = typename F::real;
    using S = typename F::small;
    __shared__ R lx[CTILE], ly[CTILE], lz[CTILE];
    __shared__ S lsp[CTILE];
    __shared__ double red[16];
    __shared__ int64_t s_range[2];
    __shared__ bool s_last;
    const int r = blockIdx.y;
    const RepDev &rp = B.rep[r];
    CtlDev &C = B.ctl[r];
    const int64_t n = rp.n;
    const int nc = rp.box.ncells;
    const int c0 = blockIdx.x * CT;
    double vm = 0;
    if (c0 < nc) {
        const int c1 = min(c0 + CT, nc);
        const int cur = C.cur;
        const int32_t *cnt = B.count[C.cnt] + rp.c_off, *start = B.start + rp.c_off;
        R *vel = B.vel[cur] + 3 * rp.p_off;
        const R *pos = B.pos[cur] + 3 * rp.p_off;
        S *spc = B.species[cur] + rp.p_off;
        if (threadIdx.x == 0) {
            s_range[0] = start[c0];
            s_range[1] = (int64_t)start[c1 - 1] + cnt[c1 - 1];
        }
        __syncthreads();
        const int64_t base = s_range[0], end = s_range[1];
        const int tcount = (end - base <= CTILE) ? (int)(end - base) : 0;
        for (int q = threadIdx.x; q < tcount; q += CT) {
            lx[q] = vel[base + q];
            ly[q] = vel[n + base + q];
            lz[q] = vel[2 * n + base + q];
            lsp[q] = spc[base + q];
        }
        __syncthreads();
        const int c = c0 + threadIdx.x;
        if (c < c1) {
            const int m = cnt[c];
            const int64_t s = start[c];
            const bool in_lds = tcount > 0;
            const int64_t o = s - base;
            auto V = [&](int k, int64_t i) -> R {
                if (in_lds) return k == 0 ? lx[o + i] : k == 1 ? ly[o + i] : lz[o + i];
                return vel[k * n + s + i];
            };
            auto Wr = [&](int64_t i, R a, R b2, R c2) {
                if (in_lds) {
                    lx[o + i] = a;
                    ly[o + i] = b2;
                    lz[o + i] = c2;
                } else {
                    vel[s + i] = a;
                    vel[n + s + i] = b2;
                    vel[2 * n + s + i] = c2;
                }
            };
            auto Sp = [&](int64_t i) -> int { return in_lds ? (int)lsp[o + i] : (int)spc[s + i]; };
            if (m > 1) {
                R ss[3] = {0, 0, 0}, cc[3] = {0, 0, 0};
                int n_eff = 0;
                for (int i = 0; i < m; ++i) {
                    if (Sp(i) > 0) {
                        for (int k = 0; k < 3; ++k) {
                            R v = V(k, i);
                            if constexpr (sizeof(R) == 4) {
                                R y = v - cc[k];
                                R t = ss[k] + y;
                                cc[k] = (t - ss[k]) - y;
                                ss[k] = t;
                            } else {
                                ss[k] = ss[k] + v;
                            }
                        }
                        ++n_eff;
                    }
                }
                if (n_eff > 1) {
                    R u0 = ss[0] / (R)n_eff, u1 = ss[1] / (R)n_eff, u2 = ss[2] / (R)n_eff;
                    Sphere ax = rand_sphere_threefry(rp.seed, STREAM_SRD, (uint64_t)C.cycle, (uint64_t)c);
                    const double v0 = ax.n[0], v1 = ax.n[1], v2 = ax.n[2];
                    // default alpha: sin = 1, cos = 0 (mpcd.f90:114-117)
                    const double sa = 1, ca = 0, omc = 1 - ca;
                    R o00 = (R)(ca + (v0 * v0) * omc), o10 = (R)((v0 * v1) * omc - v2 * sa),
                      o20 = (R)((v0 * v2) * omc + v1 * sa);
                    R o01 = (R)((v1 * v0) * omc + v2 * sa), o11 = (R)(ca + (v1 * v1) * omc),
                      o21 = (R)((v1 * v2) * omc - v0 * sa);
                    R o02 = (R)((v2 * v0) * omc - v1 * sa), o12 = (R)((v2 * v1) * omc + v0 * sa),
                      o22 = (R)(ca + (v2 * v2) * omc);
                    for (int i = 0; i < m; ++i) {
                        R w0 = V(0, i) - u0, w1 = V(1, i) - u1, w2 = V(2, i) - u2;
                        R r0 = o00 * w0;
                        r0 = r0 + o01 * w1;
                        r0 = r0 + o02 * w2;
                        R r1 = o10 * w0;
                        r1 = r1 + o11 * w1;
                        r1 = r1 + o12 * w2;
                        R r2 = o20 * w0;
                        r2 = r2 + o21 * w1;
                        r2 = r2 + o22 * w2;
                        Wr(i, u0 + r0, u1 + r1, u2 + r2);
                    }
                }
            }
            for (int i = 0; i < m; ++i) {
                double a = V(0, i), b2 = V(1, i), c2 = V(2, i);
                double q2 = (a * a + b2 * b2) + c2 * c2;
                vm = q2 > vm ? q2 : vm;
            }
            if (P.bulk) {
                // bulk_reaction B -> A (mpcd.f90:605-643), one draw threefry((cycle, cell), (seed, REACTION))
                if (m > 1 && B.is_reac[rp.c_off + c]) {
                    double local_rate = 0;
                    int64_t pick = -1;
                    for (int i = 0; i < m; ++i)
                        if (Sp(i) == 2) {
                            local_rate = local_rate + 1;
                            pick = s + i;
                        }
                    local_rate = local_rate * P.bulk_rate;
                    uint64_t a0, a1;
                    threefry2x64((uint64_t)C.cycle, (uint64_t)c, rp.seed, STREAM_REACTION, a0, a1);
                    if (u01(a0) < (1 - exp(-local_rate * P.tau)) && pick >= 0) spc[pick] = (S)1;
                }
            } else {
                // refuel (single_dimer_pbc.f90:544-565): far = (L(1)*0.45)**2 in single precision
                float farf = (float)rp.box.L[0] * 0.45f;
                farf = farf * farf;
                const double far = (double)farf;
                const double *cp = B.cpos + 6 * r;
                const double *L = rp.box.edges;
                for (int i = 0; i < m; ++i) {
                    if (Sp(i) != 2) continue;
                    const double y0 = (double)pos[s + i], y1 = (double)pos[n + s + i], y2 = (double)pos[2 * n + s + i];
                    double a0 = min_image(cp[0] - y0, L[0]), a1 = min_image(cp[2] - y1, L[1]),
                           a2 = min_image(cp[4] - y2, L[2]);
                    double b0 = min_image(cp[1] - y0, L[0]), b1 = min_image(cp[3] - y1, L[1]),
                           b2 = min_image(cp[5] - y2, L[2]);
                    if ((a0 * a0 + a1 * a1) + a2 * a2 > far && (b0 * b0 + b1 * b1) + b2 * b2 > far)
                        spc[s + i] = (S)1;
                }
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < tcount; q += CT) {
            vel[base + q] = lx[q];
            vel[n + base + q] = ly[q];
            vel[2 * n + base + q] = lz[q];
        }
    }
    double mm = block_max(vm, red);
    if (threadIdx.x == 0) B.part[(int64_t)r * PS + blockIdx.x] = mm;
    if (!arrive_last(&C.done[D_COLLIDE], gridDim.x, &s_last)) return;
    if (threadIdx.x != 0) return;
    double all = 0;
    for (unsigned q = 0; q < gridDim.x; ++q) all = fmax(all, B.part[(int64_t)r * PS + q]);
    C.done[D_COLLIDE] = 0;
    C.vmax2 = all;
    C.cycle += 1;
}

// ---------------------------------------------------------------------------------
template <class F>
static FB<F> make_fb(DimerSim &s)
{
    using R = typename F::real;
    using S = typename F::small;
    FB<F> b;
    memset(&b, 0, sizeof(b));
    for (int k = 0; k < 2; ++k) {
        b.pos[k] = (R *)s.sol_pos[k];
        b.vel[k] = (R *)s.sol_vel[k];
        b.image[k] = s.sol_image[k];
        b.id[k] = s.sol_id[k];
        b.species[k] = (S *)s.sol_species[k];
        b.flags[k] = (S *)s.sol_flags[k];
        b.count[k] = s.cell_count[k];
        b.cforce[k] = s.col_force[k];
    }
    for (int k = 0; k < 4; ++k) b.force[k] = (R *)s.sol_force[k];
    b.pos_old = (R *)s.sol_pos_old;
    b.id_to_idx = s.id_to_idx;
    b.claim = s.claim;
    b.key = s.f_key;
    b.slot = s.f_slot;
    b.tmp = s.f_tmp;
    b.start = s.cell_start;
    b.is_md = s.is_md;
    b.is_reac = s.is_reac;
    b.list = s.list;
    b.nlist = s.nlist;
    b.stencil = s.stencil;
    b.lut = s.lut;
    b.cpos = s.col_pos;
    b.cpos_old = s.col_pos_old;
    b.cpos_rattle = s.col_pos_rattle;
    b.cvel = s.col_vel;
    b.cmass = s.col_mass;
    b.cimage = s.col_image;
    b.cspecies = s.col_species;
    b.part = s.part;
    b.scan_sums = s.scan_sums;
    b.scan_stride = (s.ncmax_rep + SCAN_TILE - 1) / SCAN_TILE + 1;
    b.rep = s.rep_dev;
    b.ctl = s.ctl_dev;
    return b;
}

static FP make_fp(const DimerSim &s)
{
    FP p;
    memset(&p, 0, sizeof(p));
    p.dt = s.dt;
    p.dt_sq = (s.dt * s.dt) / 2;
    p.skin = s.skin;
    p.d = s.cfg.d;
    p.prob = s.cfg.prob;
    p.bulk_rate = s.cfg.bulk_rate;
    p.tau = s.cfg.tau;
    p.mass[0] = s.mass[0];
    p.mass[1] = s.mass[1];
    p.nmax = s.nmax;
    p.n_stencil = s.n_stencil;
    p.bulk = s.cfg.bulk_rmpcd ? 1 : 0;
    p.radius_sq = s.radius * s.radius;
    p.sc = make_lj(&s.lj_sc);
    p.cc = make_lj(&s.lj_cc);
    // rounding slack of one ballistic step: a position stays below 2 max(L) between two
    // sorts, so one add rounds by at most half an ulp there per axis
    double lmax = 0;
    for (const Replica &r : s.rep)
        for (int k = 0; k < 3; ++k) lmax = r.box.edges[k] > lmax ? r.box.edges[k] : lmax;
    double ulp = s.cfg.form == RMPCD_FP64 ? ldexp(1.0, -52) : ldexp(1.0, -23);
    p.slack = 2.0 * ulp * (2 * lmax + 4) * sqrt(3.0);
    return p;
}

template <class F>
static int enqueue_md_step(DimerSim &s, const FB<F> &b, const FP &p, hipStream_t st)
{
    const unsigned Rr = (unsigned)s.R;
    const unsigned gs = (unsigned)std::min<int64_t>((s.nmax_rep + FT - 1) / FT, STREAM_BLOCKS);
    const unsigned gr = (unsigned)std::min<int64_t>((s.nmax_rep + FT - 1) / FT, RED_BLOCKS);
    const unsigned gl = (unsigned)((2 * (int64_t)s.nmax + FT - 1) / FT);
    const unsigned gf = (unsigned)(2 * ((s.nmax + FT - 1) / FT));
    const unsigned gt = (unsigned)((s.ncmax_rep + SCAN_TILE - 1) / SCAN_TILE);
    const int64_t lim = std::max(s.nmax_rep, s.ncmax_rep);
    const unsigned gg = (unsigned)std::min<int64_t>((lim + FT - 1) / FT, RED_BLOCKS);
    k_fz_stream<F><<<dim3(gs, Rr), FT, 0, st>>>(b, p);
    k_fz_list_a<F><<<dim3(gl, Rr), FT, 0, st>>>(b, p);
    k_fz_exact<F><<<dim3(gr, Rr), FT, 0, st>>>(b, p);
    k_fz_key<F><<<dim3(gr, Rr), FT, 0, st>>>(b, p, 0);
    k_fz_scan_tiles<F><<<dim3(gt, Rr), SCAN_T, 0, st>>>(b, 0);
    k_fz_scan_top<F><<<dim3(1, Rr), 1024, 0, st>>>(b, 0);
    k_fz_scan_add<F><<<dim3(gt, Rr), SCAN_T, 0, st>>>(b, 0);
    k_fz_place<F><<<dim3(gr, Rr), FT, 0, st>>>(b, 0);
    k_fz_gather<F><<<dim3(gg, Rr), FT, 0, st>>>(b, p, 0);
    k_fz_list_build<F><<<dim3(2, Rr), 1024, 0, st>>>(b, p, 0);
    k_fz_force<F><<<dim3(gf, Rr), FT, 0, st>>>(b, p);
    k_fz_list_b<F><<<dim3(gl, Rr), FT, 0, st>>>(b, p);
    return RMPCD_OK;
}

template <class F>
static void enqueue_collision(DimerSim &s, const FB<F> &b, const FP &p, hipStream_t st)
{
    const unsigned Rr = (unsigned)s.R;
    const unsigned gr = (unsigned)std::min<int64_t>((s.nmax_rep + FT - 1) / FT, RED_BLOCKS);
    const unsigned gt = (unsigned)((s.ncmax_rep + SCAN_TILE - 1) / SCAN_TILE);
    const int64_t lim = std::max(s.nmax_rep, s.ncmax_rep);
    const unsigned gg = (unsigned)std::min<int64_t>((lim + FT - 1) / FT, RED_BLOCKS);
    const unsigned gc = (unsigned)((s.ncmax_rep + CT - 1) / CT);
    k_fz_key<F><<<dim3(gr, Rr), FT, 0, st>>>(b, p, 1);
    k_fz_scan_tiles<F><<<dim3(gt, Rr), SCAN_T, 0, st>>>(b, 1);
    k_fz_scan_top<F><<<dim3(1, Rr), 1024, 0, st>>>(b, 1);
    k_fz_scan_add<F><<<dim3(gt, Rr), SCAN_T, 0, st>>>(b, 1);
    k_fz_place<F><<<dim3(gr, Rr), FT, 0, st>>>(b, 1);
    k_fz_gather<F><<<dim3(gg, Rr), FT, 0, st>>>(b, p, 1);
    k_fz_list_build<F><<<dim3(2, Rr), 1024, 0, st>>>(b, p, 1);
    k_fz_collide<F><<<dim3(gc, Rr), CT, 0, st>>>(b, p);
}

// Event-record nodes around every prof_stride-th launch of the streaming kernel in the
// captured cycle, so that its duration is measured inside the graph replays. The nodes
// are visited in dependency order (Kahn); the streaming kernel is recognised by its
// function handle.
template <class F>
static int add_stream_timers(DimerSim &s, hipGraph_t g)
{
    size_t nn = 0;
    RMPCD_HIP(hipGraphGetNodes(g, nullptr, &nn));
    std::vector<hipGraphNode_t> nodes(nn), order;
    RMPCD_HIP(hipGraphGetNodes(g, nodes.data(), &nn));
    std::vector<size_t> indeg(nn);
    std::vector<std::vector<size_t>> succ(nn);
    auto index_of = [&](hipGraphNode_t x) {
        for (size_t i = 0; i < nn; ++i)
            if (nodes[i] == x) return i;
        return nn;
    };
    for (size_t i = 0; i < nn; ++i) {
        size_t nd = 0;
        RMPCD_HIP(hipGraphNodeGetDependentNodes(nodes[i], nullptr, &nd));
        std::vector<hipGraphNode_t> d(nd);
        if (nd) RMPCD_HIP(hipGraphNodeGetDependentNodes(nodes[i], d.data(), &nd));
        for (hipGraphNode_t x : d) {
            size_t k = index_of(x);
            if (k == nn) return set_error(RMPCD_ERR_HIP, "dimer: inconsistent captured graph");
            succ[i].push_back(k);
            ++indeg[k];
        }
    }
    std::vector<size_t> ready;
    for (size_t i = 0; i < nn; ++i)
        if (!indeg[i]) ready.push_back(i);
    while (!ready.empty()) {
        size_t i = ready.front();
        ready.erase(ready.begin());
        order.push_back(nodes[i]);
        for (size_t k : succ[i])
            if (--indeg[k] == 0) ready.push_back(k);
    }
    const void *stream_fn = reinterpret_cast<const void *>(&k_fz_stream<F>);
    size_t seen = 0, pair = 0;
    for (hipGraphNode_t x : order) {
        hipGraphNodeType t;
        RMPCD_HIP(hipGraphNodeGetType(x, &t));
        if (t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams kp;
        RMPCD_HIP(hipGraphKernelNodeGetParams(x, &kp));
        if (kp.func != stream_fn) continue;
        if (seen++ % (size_t)s.prof_stride || 2 * pair + 1 >= s.prof_ev.size()) continue;
        size_t nd = 0, nn2 = 0;
        RMPCD_HIP(hipGraphNodeGetDependencies(x, nullptr, &nd));
        std::vector<hipGraphNode_t> dep(nd);
        if (nd) RMPCD_HIP(hipGraphNodeGetDependencies(x, dep.data(), &nd));
        RMPCD_HIP(hipGraphNodeGetDependentNodes(x, nullptr, &nn2));
        std::vector<hipGraphNode_t> out(nn2);
        if (nn2) RMPCD_HIP(hipGraphNodeGetDependentNodes(x, out.data(), &nn2));
        hipGraphNode_t e0, e1;
        RMPCD_HIP(hipGraphAddEventRecordNode(&e0, g, dep.data(), nd, s.prof_ev[2 * pair]));
        for (hipGraphNode_t d : dep) RMPCD_HIP(hipGraphRemoveDependencies(g, &d, &x, 1));
        RMPCD_HIP(hipGraphAddDependencies(g, &e0, &x, 1));
        RMPCD_HIP(hipGraphAddEventRecordNode(&e1, g, &x, 1, s.prof_ev[2 * pair + 1]));
        for (hipGraphNode_t d : out) {
            RMPCD_HIP(hipGraphRemoveDependencies(g, &x, &d, 1));
            RMPCD_HIP(hipGraphAddDependencies(g, &e1, &d, 1));
        }
        ++pair;
    }
    if (!pair) return set_error(RMPCD_ERR_HIP, "dimer: streaming kernel not found in the captured graph");
    s.prof_pairs = (int)pair;
    return RMPCD_OK;
}

template <class F>
static int build_graph(DimerSim &s)
{
    FB<F> b = make_fb<F>(s);
    FP p = make_fp(s);
    hipStream_t st = s.fstream;
    if (s.graph_exec) {
        RMPCD_HIP(hipGraphExecDestroy(s.graph_exec));
        s.graph_exec = nullptr;
    }
    if (s.graph) {
        RMPCD_HIP(hipGraphDestroy(s.graph));
        s.graph = nullptr;
    }
    RMPCD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = RMPCD_OK;
    for (int j = 0; j < s.cfg.n_md && rc == RMPCD_OK; ++j) rc = enqueue_md_step<F>(s, b, p, st);
    enqueue_collision<F>(s, b, p, st);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return set_error(RMPCD_ERR_HIP, "dimer: graph capture failed: %s", hipGetErrorString(e));
    s.graph = g;
    s.prof_pairs = 0;
    if (s.prof_stride > 0 && (rc = add_stream_timers<F>(s, g))) return rc;
    RMPCD_HIP(hipGraphInstantiate(&s.graph_exec, g, nullptr, nullptr, 0));
    return RMPCD_OK;
}

int fused_set_profile(DimerSim &s, int stride)
{
    if (!s.fused_ready) return set_error(RMPCD_ERR_ARG, "dimer: profiling needs the fused mode");
    RMPCD_HIP(hipStreamSynchronize(s.fstream));
    for (hipEvent_t e : s.prof_ev) (void)hipEventDestroy(e);
    s.prof_ev.clear();
    s.prof_stride = stride > 0 ? stride : 0;
    if (s.prof_stride) {
        int pairs = (s.cfg.n_md + s.prof_stride - 1) / s.prof_stride;
        s.prof_ev.resize(2 * (size_t)pairs);
        for (hipEvent_t &e : s.prof_ev) RMPCD_HIP(hipEventCreate(&e));
    }
    return s.cfg.form == RMPCD_FP64 ? build_graph<F64>(s) : build_graph<F32>(s);
}

int fused_stream_time(DimerSim &s, double *avg_ms, int32_t *samples)
{
    *avg_ms = 0;
    *samples = 0;
    if (s.prof_pairs == 0) return set_error(RMPCD_ERR_ARG, "dimer: no profiled run yet");
    RMPCD_HIP(hipStreamSynchronize(s.fstream));
    double tot = 0;
    for (int k = 0; k < s.prof_pairs; ++k) {
        float ms = 0;
        RMPCD_HIP(hipEventElapsedTime(&ms, s.prof_ev[2 * k], s.prof_ev[2 * k + 1]));
        tot += ms;
    }
    *avg_ms = tot / s.prof_pairs;
    *samples = s.prof_pairs;
    return RMPCD_OK;
}

int fused_prepare(DimerSim &s)
{
    size_t RS = s.cfg.form == RMPCD_FP64 ? 8 : 4;
    int rc;
    // every per-block partial of a replica must fit its PS slots
    RMPCD_CHECK((s.ncmax_rep + CT - 1) / CT <= PS && 8 * ((int64_t)s.nmax + FT - 1) / FT <= PS &&
                    (2 * (int64_t)s.nmax + FT - 1) / FT <= PS,
                "dimer: box or list too large for the fused partial sums");
    if ((rc = s.alloc(&s.sol_pos_old, 3 * s.ntot * RS))) return rc;
    if ((rc = s.alloc((void **)&s.claim, s.ntot * 4))) return rc;
    if ((rc = s.alloc((void **)&s.f_key, s.ntot * 4))) return rc;
    if ((rc = s.alloc((void **)&s.f_slot, s.ntot * 4))) return rc;
    if ((rc = s.alloc((void **)&s.f_tmp, s.ntot * 4))) return rc;
    if ((rc = s.alloc((void **)&s.lut, s.luttot * 4))) return rc;
    int64_t sstride = (s.ncmax_rep + SCAN_TILE - 1) / SCAN_TILE + 1;
    if ((rc = s.alloc((void **)&s.scan_sums, (size_t)s.R * sstride * 4))) return rc;
    if ((rc = s.alloc((void **)&s.part, (size_t)s.R * PS * 8))) return rc;
    if ((rc = s.alloc((void **)&s.rep_dev, sizeof(RepDev) * s.R))) return rc;
    if ((rc = s.alloc((void **)&s.ctl_dev, sizeof(CtlDev) * s.R))) return rc;
    RMPCD_HIP(hipStreamCreate(&s.fstream));
    std::vector<RepDev> rd(s.R);
    std::vector<CtlDev> cd(s.R);
    for (int r = 0; r < s.R; ++r) {
        Replica &rp = s.rep[r];
        rd[r].box = make_box(&rp.box);
        rd[r].n = rp.n;
        rd[r].p_off = rp.p_off;
        rd[r].c_off = rp.c_off;
        rd[r].lut_off = rp.lut_off;
        rd[r].seed = rp.seed;
        CtlDev &c = cd[r];
        memset(&c, 0, sizeof(c));
        c.cur = rp.cur;
        c.fpar = rp.fpar;
        c.cfpar = rp.cfpar;
        c.cnt = rp.cnt;
        c.md_step = rp.md_steps;
        c.cycle = rp.cycles;
        c.n_extra = rp.n_extra;
        for (int k = 0; k < 3; ++k) c.origin[k] = rp.box.origin[k];
        c.vmax2 = INFINITY;  // the first step compares against the unsorted set-up positions
        c.e1 = rp.e1;
        c.e2 = rp.e2;
        // pos_old starts as the positions before the set-up sort (the reference never
        // copies pos_old before its loop, single_dimer_pbc.f90:270-303)
        rmpcd_particles_t alt = s.solvent_alt(r);
        RMPCD_HIP(hipMemcpy((char *)s.sol_pos_old + 3 * rp.p_off * RS, alt.pos, 3 * rp.n * RS,
                            hipMemcpyDeviceToDevice));
        int32_t *lut_r = s.lut + rp.lut_off;
        if ((rc = rmpcd_hilbert_table(&rp.box, lut_r))) return rc;
    }
    RMPCD_HIP(hipMemcpy(s.rep_dev, rd.data(), sizeof(RepDev) * s.R, hipMemcpyHostToDevice));
    RMPCD_HIP(hipMemcpy(s.ctl_dev, cd.data(), sizeof(CtlDev) * s.R, hipMemcpyHostToDevice));
    RMPCD_HIP(hipDeviceSynchronize());
    rc = s.cfg.form == RMPCD_FP64 ? build_graph<F64>(s) : build_graph<F32>(s);
    if (rc) return rc;
    s.fused_ready = true;
    return RMPCD_OK;
}

int run_fused_cycles(DimerSim &s, int n_cycles)
{
    if (!s.fused_ready) return set_error(RMPCD_ERR_ARG, "dimer: fused mode not prepared");
    for (int c = 0; c < n_cycles; ++c) RMPCD_HIP(hipGraphLaunch(s.graph_exec, s.fstream));
    RMPCD_HIP(hipStreamSynchronize(s.fstream));
    return s.sync_state();
}

int DimerSim::sync_state()
{
    if (cfg.mode != RMPCD_SIM_FUSED || !fused_ready) return RMPCD_OK;
    std::vector<CtlDev> cd(R);
    RMPCD_HIP(hipMemcpy(cd.data(), ctl_dev, sizeof(CtlDev) * R, hipMemcpyDeviceToHost));
    int err = 0;
    for (int r = 0; r < R; ++r) {
        Replica &rp = rep[r];
        const CtlDev &c = cd[r];
        rp.cur = c.cur;
        rp.fpar = c.fpar;
        rp.cfpar = c.cfpar;
        rp.cnt = c.cnt;
        rp.md_steps = c.md_step;
        rp.cycles = c.cycle;
        rp.n_extra = c.n_extra;
        rp.e1 = c.e1;
        rp.e2 = c.e2;
        for (int k = 0; k < 3; ++k) rp.box.origin[k] = c.origin[k];
        if (c.error) err = c.error;
    }
    if (err == RMPCD_ERR_OVERFLOW) return set_error(RMPCD_ERR_OVERFLOW, "maximum of neighbor list reached");
    if (err) return set_error(err, "dimer: device error %d", err);
    return RMPCD_OK;
}

}  // namespace rmpcd
