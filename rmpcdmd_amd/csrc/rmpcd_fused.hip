// rmpcd_fused.hip -- the single-dimer MD step and collision cycle as fused device-side
// kernels over a group of replicas at once (blockIdx.y = replica within the group),
// captured as one hipGraph per collision cycle with one independent branch per group (a
// group's latency-bound kernels run beside another group's streaming).
//
// Per MD step (programs/single_dimer_pbc.f90:306-383), one kernel k_fz_step: the
// neighbour-list members finish the previous step (forces, md_vel, catalytic flagging
// and conversion) and stream; every other particle streams.  Then one block per replica
// (k_fz_colloid) finishes the colloids and takes the skin decision, and, only when the
// skin test fires, an extra sort: it computes the reference's new order without moving
// the particles (key+count, scan, place, rank into perm/inv) or, one extra sort in two,
// defers even that (the key pass keeps the cells, the list build ranks the particles
// near the colloids).  Per collision: a sort with a fresh grid shift that moves the
// particles (the key pass writes a staging record per particle, the gather writes them
// out in cell order), then SRD + refuel (or the bulk reaction) per cell.
//
// Why member-only work is exact: forces only reach particles in a colloid's list, the
// list keeps every particle within max_cut + skin, and the skin test rebuilds it before
// any particle that carried a force, an MD/PAST_MD bit or the REAC bit can leave it, so
// the reference's full passes (md_vel, the MD-bit roll, force zeroing, change_species)
// leave every other particle unchanged.  Skin decision without a full pass: between two
// sorts a particle outside the lists moves ballistically, so its displacement after k
// steps is at most k (dt |v|max (1 + 1e-6) + rounding slack), with |v|max taken over all
// particles at the last sort or collision; list members are measured exactly.  When
// that bound stays below the threshold the reference's test is false too; otherwise the
// stream role measures every particle exactly in the same pass.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "rmpcd_sim.hpp"

namespace rmpcd {

constexpr int FT = 256;            // threads per block
constexpr int COLLOID_T = 256;     // threads per block of the colloid kernel
constexpr int PS = 32768;          // partial-sum slots per replica
constexpr int STREAM_BLOCKS = 16384;  // stream-role blocks per replica, default and maximum
constexpr int LAZY_STREAM_BLOCKS = 1024;  // stream-role blocks per replica in lazy mode
constexpr int SORT_BLOCKS = 4096;     // blocks of the particle passes of a sort (all replicas), default
constexpr int GATHER_GRID_MUL = 4;  // the collision sort's fused gather runs on this many times the sort grid
constexpr int SCAN_BLOCKS = 1024;     // blocks of the cell-count tile scan (all replicas)
constexpr int CT_DEFAULT = 128;  // cells per block in the collision kernel (64, 128 or 256; 12 particles per cell
                                 // staged in LDS, a block with more works from memory)
constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T * SCAN_I;
constexpr int MAX_TILES = 4096;  // cells per replica up to MAX_TILES * SCAN_TILE
constexpr int MAX_STENCIL = 2048;  // stencil cells of the fused list build
constexpr int GPT = 1, GT = FT * GPT;  // destination slots per thread and per unit of the gather
constexpr int GHALO = 32;              // slots staged on either side of a gather or rank unit
constexpr int KPT = 4, KCHUNK = FT * KPT;  // particles per thread and per chunk of the key pass
constexpr int KHASH_BITS = 11, KHASH = 1 << KHASH_BITS;  // LDS hash entries, 2 x KCHUNK
// a replica's error flag: a list longer than its capacity (the reference's error stop),
// or more particles near the colloids than a deferred extra sort can collect; every
// kernel returns at once for a replica with the flag set, so it stops at that step
constexpr int ERR_CANDIDATES = -100;


// The staging record of a sort: a particle's wrapped state, written by the key pass in
// source order and read whole by the gather (one or two cache lines per particle where
// the planes would cost one line per component).
template <class Rl, class S>
struct SortRec;
template <>
struct alignas(16) SortRec<float, uint8_t> {
    float p[3], v[3];
    int32_t im[3], id, key;
    uint32_t sf;  // species | flags << 8
    __device__ int species() const { return (int)(sf & 0xffu); }
    __device__ int flags() const { return (int)(sf >> 8); }
    __device__ void set_sf(int sp, int fl) { sf = (uint32_t)(sp & 0xff) | ((uint32_t)(fl & 0xff) << 8); }
};
template <>
struct alignas(16) SortRec<double, int32_t> {
    double p[3], v[3];
    int32_t im[3], id, key, sp, fl, pad;
    __device__ int species() const { return sp; }
    __device__ int flags() const { return fl; }
    __device__ void set_sf(int s_, int f_) { sp = s_; fl = f_; pad = 0; }
};
static_assert(sizeof(SortRec<float, uint8_t>) == 48, "FP32 sort record");
static_assert(sizeof(SortRec<double, int32_t>) == 80, "FP64 sort record");

template <class F>
struct FB {
    using Rec = SortRec<typename F::real, typename F::small>;
    typename F::real *pos[2], *vel[2], *force[4], *pos_old;
    Rec *rec;
    int32_t *image[2], *id[2];
    typename F::small *species[2], *flags[2];
    int32_t *id_to_idx, *key, *slot;
    // After an extra sort the particles stay where they are: perm[d] is the storage index
    // of the particle at reference position d, inv its inverse, pair the (reference
    // position, storage index) of each slot that the extra sort's place pass fills.
    int32_t *perm, *inv;
    int2 *pair;
    // A deferred extra sort keeps the order of the previous sort in perm/inv and the new
    // cells in dkey (the order it stands for: cell, then previous order); cand lists the
    // particles near the colloids for its list build (cand_cap per replica), stage holds
    // the list build's working entries (nmax per colloid).
    int32_t *dkey, *cand;
    int4 *stage;
    int64_t cand_cap;
    uint8_t *memb;
    int32_t *count[2], *start;
    uint8_t *is_md, *is_reac;
    int32_t *list, *nlist;
    const int32_t *stencil, *lut;
    double *cpos, *cpos_old, *cpos_rattle, *cvel, *cforce[2], *cmass;
    int32_t *cimage, *cspecies;
    double *part;
    int32_t *scan_sums, *slocal;
    int64_t scan_stride;
    int32_t *mslots, *mcount;  // per replica and sort buffer: the member slots of its last lists
    int64_t *plan;             // the replicas a sort chain works on (k_fz_plan)
    double *gpart;             // |v|^2 maximum per gather unit
    int32_t nmax;
    int32_t xcd_order;  // sort passes: units grouped by XCD (1) or a contiguous range per workgroup (0)
    int32_t fuse_srd;   // the collision sort's gather also makes the collision (no bulk reaction)
    // the replica group a launch works on: replicas r0 + blockIdx.y (R of them); plan and
    // gpart are the group's own
    int32_t r0;
    const RepDev *rep;
    CtlDev *ctl;
};

struct FP {
    double dt, dt_sq, skin, d, prob, bulk_rate, tau, slack;
    double mass[2];
    int nmax, n_stencil, bulk, lazy;
    int stencil_m, pad_;  // stencil offsets lie in [-stencil_m, stencil_m]^3
    double radius_sq;
    LJ sc, cc;
};

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

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

// ---- the MD step kernel ---------------------------------------------------------------
// Blocks [0, nbm) take the entries of the two neighbour lists (member role: a particle in
// both lists is taken by its colloid-0 entry, so each member is handled by one thread);
// the other blocks stream every solvent particle that is not a member (stream role).
//   mode 0  first step of a cycle: the members stream with the force of the last step
//   mode 1  the members finish the previous step -- the MD-bit roll and LJ forces of
//           compute_force (neighbor_list.f90:180-222), md_vel (md.f90:219-237),
//           flag_particles_nl and change_species (single_dimer_pbc.f90:493-542) -- and
//           stream for this step (md_pos, md.f90:61-77); the non-members stream
//   mode 2  the members finish the previous step only (the last step before a collision)
// The last block to finish completes the colloids' share of the previous step (force
// sums, compute_force_n2, Verlet velocity, RATTLE), moves them for this step (Verlet
// position, RATTLE) and takes the skin decision (single_dimer_pbc.f90:307-361).
// Member partials: MPV values per member block (colloid-0 force, energy, colloid-1
// force, energy, maximum squared displacement); stream-role maxima follow at MOFF.
constexpr int MPV = 9;
constexpr int MAX_MEMBER_BLOCKS = 256;  // member blocks per replica (lists of up to 32768 entries each)

template <class F>
__global__ void __launch_bounds__(FT) k_fz_step(FB<F> B, FP P, int mode, int nbm)
{
    using R = typename F::real;
    using S = typename F::small;
    __shared__ double red[16];
    const int r = B.r0 + blockIdx.y;
    const RepDev &rp = B.rep[r];
    CtlDev &C = B.ctl[r];
    if (C.error) return;  // the replica stopped at a failed list build
    const int cur = C.cur, fp = C.fpar;
    const bool exact = C.need_exact != 0;
    const int64_t n = rp.n, po3 = 3 * rp.p_off;
    R *__restrict__ pos = B.pos[cur] + po3;
    R *__restrict__ vel = B.vel[cur] + po3;
    const R *__restrict__ po = B.pos_old + po3;
    const uint8_t *__restrict__ memb = B.memb + rp.p_off;
    const R rdt = (R)P.dt, dsq = (R)P.dt_sq;
    double *part = B.part + (int64_t)r * PS;
    const int moff = nbm * MPV;
    if ((int)blockIdx.x < nbm) {
        double c[MPV] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int e = blockIdx.x * FT + threadIdx.x;
        const int col = e / P.nmax, k = e - col * P.nmax;
        if (col < 2 && k < min(B.nlist[2 * r + col], P.nmax)) {
            const int idx = B.list[(int64_t)(2 * r + col) * P.nmax + k];
            const int mb = memb[idx];
            if (col == 0 || !(mb & 1)) {
                R *fnew = B.force[2 * cur + (1 - fp)] + po3;
                const R *fcur = B.force[2 * cur + fp] + po3;
                R x[3] = {pos[idx], pos[n + idx], pos[2 * n + idx]};
                R v[3] = {vel[idx], vel[n + idx], vel[2 * n + idx]};
                R fa[3] = {fcur[idx], fcur[n + idx], fcur[2 * n + idx]};
                if (mode >= 1) {
                    S *spc = B.species[cur] + rp.p_off;
                    S *flg = B.flags[cur] + rp.p_off;
                    const double *cp = B.cpos + 6 * r;
                    const double *L = rp.box.edges;
                    const int s2 = (int)spc[idx];
                    int fl = (int)flg[idx];
                    fl = (fl & MD_MASK) ? (fl | PAST_MD_MASK) : (fl & ~PAST_MD_MASK);
                    fl &= ~MD_MASK;
                    const double y[3] = {(double)x[0], (double)x[1], (double)x[2]};
                    double dsq_c[2] = {0, 0};
                    R fn[3] = {0, 0, 0};
                    for (int cc = 0; cc < 2; ++cc) {
                        const double d0 = min_image(cp[cc] - y[0], L[0]);
                        const double d1 = min_image(cp[2 + cc] - y[1], L[1]);
                        const double d2 = min_image(cp[4 + cc] - y[2], L[2]);
                        const double r_sq = (d0 * d0 + d1 * d1) + d2 * d2;
                        dsq_c[cc] = r_sq;
                        if (!(mb & (1 << cc)) || s2 <= 0) continue;
                        const int s1 = B.cspecies[2 * r + cc];
                        const int pidx = (s2 - 1) + P.sc.n1 * (s1 - 1);
                        if (r_sq <= P.sc.cut_sq[pidx]) {
                            const double fr = lj_force_over_r(r_sq, P.sc.eps[pidx], P.sc.sigma[pidx]);
                            const double g0 = fr * d0, g1 = fr * d1, g2 = fr * d2;
                            fn[0] = fn[0] + (R)(-g0);
                            fn[1] = fn[1] + (R)(-g1);
                            fn[2] = fn[2] + (R)(-g2);
                            fl |= MD_MASK | (s1 == 1 ? CATALYZED_MASK : 0);
                            c[4 * cc] = g0;
                            c[4 * cc + 1] = g1;
                            c[4 * cc + 2] = g2;
                            c[4 * cc + 3] = lj_energy(r_sq, P.sc.eps[pidx], P.sc.sigma[pidx], P.sc.cut_energy[pidx]);
                        }
                    }
                    fnew[idx] = fn[0];
                    fnew[n + idx] = fn[1];
                    fnew[2 * n + idx] = fn[2];
                    if (!(fl & WALL_MASK)) {
                        for (int q = 0; q < 3; ++q) v[q] = v[q] + (rdt * (fn[q] + fa[q])) / (R)2;
                        vel[idx] = v[0];
                        vel[n + idx] = v[1];
                        vel[2 * n + idx] = v[2];
                    } else {
                        fl &= ~WALL_MASK;
                    }
                    const double cut_c = P.sc.cut_sq[0], cut_n = P.sc.cut_sq[2];
                    int snew = s2;
                    if ((mb & 1) && s2 == 1 && dsq_c[0] < cut_c) fl |= REAC_MASK;
                    if ((fl & REAC_MASK) && dsq_c[0] > cut_c && dsq_c[1] > cut_n) {
                        // one draw per conversion candidate: threefry((md_step, id), (seed, CHANGE))
                        uint64_t u0, u1;
                        threefry2x64((uint64_t)C.md_step, (uint64_t)B.id[cur][rp.p_off + idx], rp.seed,
                                     STREAM_CHANGE, u0, u1);
                        if (u01(u0) <= P.prob) snew = 2;
                        fl &= ~REAC_MASK;
                    }
                    if (snew != s2) spc[idx] = (S)snew;
                    flg[idx] = (S)fl;
                    fa[0] = fn[0];
                    fa[1] = fn[1];
                    fa[2] = fn[2];
                }
                if (mode <= 1) {
                    for (int q = 0; q < 3; ++q) x[q] = (x[q] + rdt * v[q]) + dsq * fa[q];
                    pos[idx] = x[0];
                    pos[n + idx] = x[1];
                    pos[2 * n + idx] = x[2];
                    const double d0 = (double)(x[0] - po[idx]);
                    const double d1 = (double)(x[1] - po[n + idx]);
                    const double d2 = (double)(x[2] - po[2 * n + idx]);
                    c[8] = (d0 * d0 + d1 * d1) + d2 * d2;
                }
            }
        }
        // nine reductions at once: wave shuffles, then one exchange through LDS
        __shared__ double wred[FT / 64][MPV];
        for (int off = 32; off > 0; off >>= 1) {
            for (int q = 0; q < 8; ++q) c[q] += __shfl_xor(c[q], off, 64);
            const double o = __shfl_xor(c[8], off, 64);
            c[8] = o > c[8] ? o : c[8];
        }
        if ((threadIdx.x & 63) == 0)
            for (int q = 0; q < MPV; ++q) wred[threadIdx.x >> 6][q] = c[q];
        __syncthreads();
        if (threadIdx.x < MPV) {
            double t = wred[0][threadIdx.x];
            for (int w = 1; w < FT / 64; ++w) {
                const double u = wred[w][threadIdx.x];
                t = threadIdx.x == 8 ? (u > t ? u : t) : t + u;
            }
            part[blockIdx.x * MPV + threadIdx.x] = t;
        }
    } else {
        const int64_t gs = (int64_t)(gridDim.x - nbm) * FT;
        const S *spc = B.species[cur] + rp.p_off;
        double mx = 0;
        if (P.lazy) {
            // non-members keep the position of the last sort (pos == pos_old) until a sort
            // reads them; an exact skin test replays their k additions x = x + dt v
            if (!exact) return;
            const int k = C.k_since + 1;
            for (int64_t i = (int64_t)(blockIdx.x - nbm) * FT + threadIdx.x; i < n; i += gs) {
                if (memb[i] || (int)spc[i] < 0) continue;
                double q2 = 0;
                for (int c = 0; c < 3; ++c) {
                    const R t = rdt * vel[c * n + i];
                    R x = pos[c * n + i];
                    for (int s2 = 0; s2 < k; ++s2) x = x + t;
                    const double dd = (double)(x - po[c * n + i]);
                    q2 = c == 0 ? dd * dd : (c == 1 ? q2 + dd * dd : q2 + dd * dd);
                }
                mx = q2 > mx ? q2 : mx;
            }
            const double m = block_max(mx, red);
            if (threadIdx.x == 0) part[moff + (blockIdx.x - nbm)] = m;
            return;
        }
        if (rp.vec) {
            // 16-byte non-temporal loads and stores: VW particles per thread and plane (the
            // replica's planes and membership bytes are 16-byte aligned).  Nothing read
            // here is read again in this launch, so the lines need not stay in L2.  An
            // exact skin test also reads pos_old and takes the largest displacement.
            constexpr int VW = 16 / sizeof(R);
            using V = typename std::conditional<sizeof(R) == 4, v4f, v2d>::type;
            using M = typename std::conditional<sizeof(R) == 4, uint32_t, uint16_t>::type;
            const int64_t nv = n / VW;
            for (int64_t q = (int64_t)(blockIdx.x - nbm) * FT + threadIdx.x; q < nv; q += gs) {
                const uint32_t mb = (uint32_t)reinterpret_cast<const M *>(memb)[q];
                V pv[3], vv[3];
                for (int k = 0; k < 3; ++k) {
                    pv[k] = __builtin_nontemporal_load(reinterpret_cast<const V *>(pos + k * n) + q);
                    vv[k] = __builtin_nontemporal_load(reinterpret_cast<const V *>(vel + k * n) + q);
                }
                for (int k = 0; k < 3; ++k) {
                    R *pe = reinterpret_cast<R *>(&pv[k]);
                    const R *ve = reinterpret_cast<const R *>(&vv[k]);
                    for (int u = 0; u < VW; ++u)
                        if (!((mb >> (8 * u)) & 0xffu)) pe[u] = pe[u] + rdt * ve[u];
                }
                if (exact) {
                    V ov[3];
                    for (int k = 0; k < 3; ++k)
                        ov[k] = __builtin_nontemporal_load(reinterpret_cast<const V *>(po + k * n) + q);
                    for (int u = 0; u < VW; ++u) {
                        if (((mb >> (8 * u)) & 0xffu) || (int)spc[q * VW + u] < 0) continue;
                        const double d0 = (double)(pv[0][u] - ov[0][u]);
                        const double d1 = (double)(pv[1][u] - ov[1][u]);
                        const double d2 = (double)(pv[2][u] - ov[2][u]);
                        const double q2 = (d0 * d0 + d1 * d1) + d2 * d2;
                        mx = q2 > mx ? q2 : mx;
                    }
                }
                if (!mb) {
                    for (int k = 0; k < 3; ++k)
                        __builtin_nontemporal_store(pv[k], reinterpret_cast<V *>(pos + k * n) + q);
                } else {
                    // a member is streamed by its member thread: write the others only
                    for (int k = 0; k < 3; ++k) {
                        const R *pe = reinterpret_cast<const R *>(&pv[k]);
                        for (int u = 0; u < VW; ++u)
                            if (!((mb >> (8 * u)) & 0xffu)) pos[k * n + q * VW + u] = pe[u];
                    }
                }
            }
            if (exact) {
                const double m = block_max(mx, red);
                if (threadIdx.x == 0) part[moff + (blockIdx.x - nbm)] = m;
            }
            return;
        }
        for (int64_t i = (int64_t)(blockIdx.x - nbm) * FT + threadIdx.x; i < n; i += gs) {
            if (memb[i]) continue;
            const R x0 = pos[i] + rdt * vel[i];
            const R x1 = pos[n + i] + rdt * vel[n + i];
            const R x2 = pos[2 * n + i] + rdt * vel[2 * n + i];
            pos[i] = x0;
            pos[n + i] = x1;
            pos[2 * n + i] = x2;
            if (exact && (int)spc[i] >= 0) {
                const double d0 = (double)(x0 - po[i]);
                const double d1 = (double)(x1 - po[n + i]);
                const double d2 = (double)(x2 - po[2 * n + i]);
                const double q2 = (d0 * d0 + d1 * d1) + d2 * d2;
                mx = q2 > mx ? q2 : mx;
            }
        }
        if (exact) {
            const double m = block_max(mx, red);
            if (threadIdx.x == 0) part[moff + (blockIdx.x - nbm)] = m;
        }
    }
}

// The colloids' share of the step, one block per replica after k_fz_step (COLLOID_T
// threads: most of the work is one thread's, and a small block starts sooner beside
// another replica group's streaming): the member
// partial sums in block order (deterministic), the stream-role maxima, then the
// colloid force, compute_force_n2, Verlet velocity and RATTLE of the previous step and
// the Verlet position step, RATTLE and skin decision of this one.
template <class F>
__global__ void __launch_bounds__(COLLOID_T) k_fz_colloid(FB<F> B, FP P, int mode, int nbm, int nstream)
{
    __shared__ double red[16];
    __shared__ double acc[MPV];
    __shared__ double lp[MAX_MEMBER_BLOCKS * MPV];
    const int r = B.r0 + blockIdx.y;
    const RepDev &rp = B.rep[r];
    CtlDev &C = B.ctl[r];
    if (C.error) return;
    const bool exact = C.need_exact != 0;
    const double *part = B.part + (int64_t)r * PS;
    const int moff = nbm * MPV;
    for (int q = threadIdx.x; q < nbm * MPV; q += blockDim.x) lp[q] = part[q];
    __syncthreads();
    if (threadIdx.x < MPV) {
        double a = 0;
        for (int b = 0; b < nbm; ++b) {
            const double t = lp[b * MPV + threadIdx.x];
            if (threadIdx.x == 8)
                a = t > a ? t : a;
            else
                a = a + t;
        }
        acc[threadIdx.x] = a;
    }
    double smax = 0;
    if (exact && mode <= 1) {
        for (int b = threadIdx.x; b < nstream; b += blockDim.x) {
            const double t = part[moff + b];
            smax = t > smax ? t : smax;
        }
    }
    smax = block_max(smax, red);
    if (threadIdx.x != 0) return;
    // the colloid state in registers: one batch of loads, one batch of stores
    const double *L = rp.box.edges;
    const double dt = P.dt;
    const int cfpar = C.cfpar, fpar = C.fpar, k_since = C.k_since;
    const int64_t md_step = C.md_step;
    const double vmax2 = C.vmax2;
    double cp[6], cv[6], cf[6], cfo[6], co[6], cm[2];
    int32_t cs[2];
    for (int q = 0; q < 6; ++q) {
        cp[q] = B.cpos[6 * r + q];
        cv[q] = B.cvel[6 * r + q];
        cf[q] = B.cforce[cfpar][6 * r + q];
        cfo[q] = B.cforce[cfpar ^ 1][6 * r + q];
        co[q] = B.cpos_old[6 * r + q];
    }
    cm[0] = B.cmass[2 * r];
    cm[1] = B.cmass[2 * r + 1];
    cs[0] = B.cspecies[2 * r];
    cs[1] = B.cspecies[2 * r + 1];
    int ncf = cfpar;
    if (mode >= 1) {
        // switch(force, force_old) of the colloids; the new force is the list sum
        // (single_dimer_pbc.f90:341-352), then compute_force_n2 (neighbor_list.f90:243-262)
        ncf = cfpar ^ 1;
        for (int q = 0; q < 6; ++q) cfo[q] = cf[q];
        for (int cc = 0; cc < 2; ++cc) {
            cf[cc] = 0.0 + acc[4 * cc];
            cf[2 + cc] = 0.0 + acc[4 * cc + 1];
            cf[4 + cc] = 0.0 + acc[4 * cc + 2];
        }
        double e1 = 0;
        e1 = e1 + acc[3];
        e1 = e1 + acc[7];
        C.e1 = e1;
        {
            const double d0 = min_image(cp[0] - cp[1], L[0]);
            const double d1 = min_image(cp[2] - cp[3], L[1]);
            const double d2 = min_image(cp[4] - cp[5], L[2]);
            const double r_sq = (d0 * d0 + d1 * d1) + d2 * d2;
            const int sa = cs[0], sb = cs[1];
            const int pc = (sa - 1) + P.cc.n1 * (sb - 1), pe = (sb - 1) + P.cc.n1 * (sa - 1);
            double e2 = 0, e0 = 0;
            if (r_sq <= P.cc.cut_sq[pc]) {
                const double fr = lj_force_over_r(r_sq, P.cc.eps[pe], P.cc.sigma[pe]);
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
        for (int i = 0; i < 2; ++i)
            for (int q3 = 0; q3 < 3; ++q3) {
                const int q = 2 * q3 + i;
                cv[q] = cv[q] + (dt * (cf[q] + cfo[q])) / (2 * cm[i]);
            }
        rattle_vel_2(cp, cv, cs, P.mass, P.d, L);
        for (int q = 0; q < 6; ++q) B.cforce[ncf][6 * r + q] = cf[q];
        C.cfpar = ncf;
        C.fpar = fpar ^ 1;
        C.md_step = md_step + 1;
    }
    if (mode <= 1) {
        // colloid Verlet position step and RATTLE (single_dimer_pbc.f90:311-317), with the
        // force of the last completed step
        double cr[6];
        for (int q = 0; q < 6; ++q) cr[q] = cp[q];
        for (int i = 0; i < 2; ++i)
            for (int q3 = 0; q3 < 3; ++q3) {
                const int q = 2 * q3 + i;
                cp[q] = (cp[q] + dt * cv[q]) + ((dt * dt) * cf[q]) / (2 * cm[i]);
            }
        rattle_pos_2(cp, cv, cr, cs, P.mass, P.d, dt, L);
        for (int q = 0; q < 6; ++q) B.cpos_rattle[6 * r + q] = cr[q];
        // maximum displacements and the skin test (single_dimer_pbc.f90:320-323)
        double cmax = 0;
        for (int i = 0; i < 2; ++i) {
            const double d0 = cp[i] - co[i], d1 = cp[2 + i] - co[2 + i], d2 = cp[4 + i] - co[4 + i];
            const double q2 = (d0 * d0 + d1 * d1) + d2 * d2;
            cmax = q2 > cmax ? q2 : cmax;
        }
        const double co_max = sqrt(cmax);
        C.co_max = co_max;
        const int k = k_since + 1;
        const double per = P.dt * sqrt(vmax2) * (1.0 + 1e-6) + P.slack;
        const double mem = acc[8];
        C.list_max2 = mem;
        const double so_max = exact ? sqrt(mem > smax ? mem : smax) : fmax(sqrt(mem), (double)k * per);
        C.need_sort = (co_max >= P.skin * 0.1 || so_max >= P.skin * 0.89) ? 1 : 0;
        C.k_since = k;
        // is the ballistic bound conclusive for the next step?
        C.need_exact = (double)(k + 1) * per >= P.skin * 0.89 ? 1 : 0;
    }
    for (int q = 0; q < 6; ++q) {
        B.cpos[6 * r + q] = cp[q];
        B.cvel[6 * r + q] = cv[q];
    }
}

// ---- sort -------------------------------------------------------------------------
// The particle passes of a sort run on a fixed grid over the replicas that sort (a
// graph cannot branch, and a grid of blocks per replica that mostly exit at once costs
// microseconds per launch in a large batch).  k_fz_plan lists those replicas and the
// prefix sums of their work units per pass; a pass's blocks take contiguous unit ranges.
enum { U_KEY = 0, U_SCAN, U_PLACE, U_GATHER, U_KINDS };
constexpr int PT = 4 * FT;  // particles or cells per place unit

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

// S2: exclusive scan of the new cell counts inside tiles of SCAN_TILE cells (tile-local
// offsets into slocal, tile totals into scan_sums)

constexpr int GC = 96, GC_TILE = 12 * GC;  // cells and staged particles per unit of the fused gather
__device__ inline int64_t units_of(const RepDev &rp, int kind, bool fused_gather)
{
    const int64_t n = rp.n, nc = rp.box.ncells;
    if (kind == U_GATHER && fused_gather) return (nc + GC - 1) / GC;
    switch (kind) {
    case U_KEY:
        return (n + KCHUNK - 1) / KCHUNK;
    case U_SCAN:
        return (nc + SCAN_TILE - 1) / SCAN_TILE;
    case U_PLACE:
        return (n + PT - 1) / PT + (nc + PT - 1) / PT;
    default:
        return ((n > nc ? n : nc) + GT - 1) / GT;
    }
}

// plan layout (int64), for a group of R replicas from r0 on: [0] active count, [1 + kind]
// total units, then act[R] (replica of active index a), amap[R] (active index of replica
// r0 + j at j, or -1), pre[kind][R + 1]
__device__ inline const int64_t *plan_act(const int64_t *plan) { return plan + 1 + U_KINDS; }
__device__ inline const int64_t *plan_amap(const int64_t *plan, int R) { return plan + 1 + U_KINDS + R; }
__device__ inline const int64_t *plan_pre(const int64_t *plan, int R, int kind)
{
    return plan + 1 + U_KINDS + 2 * (int64_t)R + (int64_t)kind * (R + 1);
}

template <class F>
__global__ void __launch_bounds__(FT) k_fz_plan(FB<F> B, int forced, int R)
{
    __shared__ int lds[FT / 64];
    int64_t *plan = B.plan;
    int64_t *act = plan + 1 + U_KINDS, *amap = act + R;
    // the common case first: no replica sorts
    bool any = forced != 0;
    for (int r = B.r0 + threadIdx.x; r < B.r0 + R && !any; r += blockDim.x)
        any = B.ctl[r].need_sort != 0 && !B.ctl[r].error;
    if (!__syncthreads_or(any)) {
        if (threadIdx.x <= U_KINDS) plan[threadIdx.x] = 0;
        return;
    }
    int carry = 0;
    int64_t csum[U_KINDS] = {0, 0, 0, 0};
    for (int base = 0; base < R; base += blockDim.x) {
        const int rl = base + threadIdx.x, r = B.r0 + rl;
        // a replica whose list build failed has stopped (the reference's error stop)
        const bool on = rl < R && !B.ctl[r].error && sort_active(B.ctl[r], forced);
        // an extra sort is deferred when the previous one was not and the box allows it:
        // only the key pass and the list build run for it
        const bool dfr = on && !forced && B.ctl[r].def == 0 && B.rep[r].defer != 0;
        if (on) {
            B.ctl[r].kind = dfr ? 1 : 0;
            B.ctl[r].ccount = 0;
        }
        int tot;
        const int off = carry + scan_block_exclusive(on ? 1 : 0, lds, &tot);
        if (rl < R) amap[rl] = on ? off : -1;
        if (on) act[off] = r;
        for (int kind = 0; kind < U_KINDS; ++kind) {
            const int u = on && !(dfr && kind != U_KEY) ? (int)units_of(B.rep[r], kind, forced && B.fuse_srd) : 0;
            int ktot;
            const int kx = scan_block_exclusive(u, lds, &ktot);
            if (on) plan[1 + U_KINDS + 2 * (int64_t)R + (int64_t)kind * (R + 1) + off] = csum[kind] + kx;
            csum[kind] += ktot;
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        plan[0] = carry;
        for (int kind = 0; kind < U_KINDS; ++kind) {
            plan[1 + kind] = csum[kind];
            plan[1 + U_KINDS + 2 * (int64_t)R + (int64_t)kind * (R + 1) + carry] = csum[kind];
        }
    }
}

// the active index whose unit range holds unit u (the last a with pre[a] <= u)
__device__ inline int find_unit(const int64_t *pre, int nact, int64_t u)
{
    int lo = 0, hi = nact - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= u)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// the units u0, u0 + step, ... below u1 of this block for pass `kind`: a contiguous range
// per workgroup, or, XCD-grouped, the workgroups that run on one XCD (consecutive
// workgroups go to the XCDs in turn) share a contiguous eighth of the units, interleaved:
// the units in flight on an XCD are then neighbours, whose particles come from, and go
// to, the same stretch of memory, cached once in that XCD's L2
constexpr int XCDS = 8;
struct UnitRange {
    int64_t u0, u1, step;
    int nact;
    const int64_t *pre, *act;
};
__device__ inline UnitRange unit_range(const int64_t *plan, int R, int kind, int xcd_order)
{
    UnitRange w;
    w.nact = (int)plan[0];
    const int64_t T = plan[1 + kind];
    const int G = (int)gridDim.x;
    if (xcd_order && G % XCDS == 0) {
        const int64_t per = (T + XCDS - 1) / XCDS, lo = (int64_t)(blockIdx.x % XCDS) * per;
        w.u0 = lo + blockIdx.x / XCDS;
        w.u1 = lo + per < T ? lo + per : T;
        w.step = G / XCDS;
    } else {
        const int64_t per = (T + G - 1) / G;
        w.u0 = (int64_t)blockIdx.x * per;
        w.u1 = w.u0 + per < T ? w.u0 + per : T;
        w.step = 1;
    }
    w.pre = plan_pre(plan, R, kind);
    w.act = plan_act(plan);
    return w;
}

// S1: grid shift (collision sort), periodic wrap, key and slot in the cell, and the
// particle's staging record for the gather.  A unit is a chunk of KCHUNK consecutive
// particles: its keys are counted in an LDS hash table,
// then one global atomic per distinct key reserves the chunk's slots in that cell
// (arrival order within a cell is made stable by the gather).  The block with a
// replica's first chunk records the grid shift and wraps the colloids.
template <class F, int kForced>
__device__ inline void key_pass(const FB<F> &B, const FP &P, int R)
{
    constexpr int forced = kForced;
    using Rl = typename F::real;
    // the LDS hash (hkey, hcnt, occ); a collision sort stages each slab of FT records in
    // the same memory before the hash is built
    __shared__ __align__(16) int hbuf[3 * KHASH];
    int *hkey = hbuf, *hcnt = hbuf + KHASH, *occ = hbuf + 2 * KHASH;
    static_assert(sizeof(typename FB<F>::Rec) * FT <= sizeof(int) * 3 * KHASH, "record slab in the hash LDS");
    __shared__ int lds[FT / 64];
    __shared__ double red[16];
    const UnitRange w = unit_range(B.plan, R, U_KEY, B.xcd_order);
    int a = -1;
    for (int64_t u = w.u0; u < w.u1; u += w.step) {
        if (a < 0 || u >= w.pre[a + 1]) a = find_unit(w.pre, w.nact, u);
        const int r = (int)w.act[a];
        const int64_t c0 = (u - w.pre[a]) * KCHUNK;
        CtlDev &C = B.ctl[r];
        const RepDev &rp = B.rep[r];
        Box b = rp.box;
        if (forced) {
            uint64_t a0, a1, a2, a3;
            threefry2x64((uint64_t)C.cycle, 0, rp.seed, STREAM_SHIFT, a0, a1);
            threefry2x64((uint64_t)C.cycle, 1, rp.seed, STREAM_SHIFT, a2, a3);
            b.origin[0] = u01(a0) - 1;  // random_shift
            b.origin[1] = u01(a1) - 1;
            b.origin[2] = u01(a2) - 1;
        } else {
            for (int k = 0; k < 3; ++k) b.origin[k] = C.origin[k];
        }
        const int64_t n = rp.n;
        Rl *pos = B.pos[C.cur] + 3 * rp.p_off, *pold = B.pos_old + 3 * rp.p_off;
        const Rl *vel = B.vel[C.cur] + 3 * rp.p_off;
        int32_t *img = B.image[C.cur] + 3 * rp.p_off;
        const int32_t *ids = B.id[C.cur] + rp.p_off;
        const typename F::small *spc = B.species[C.cur] + rp.p_off, *flg = B.flags[C.cur] + rp.p_off;
        uint8_t *memb = B.memb + rp.p_off;
        typename FB<F>::Rec *rec = B.rec + rp.p_off;
        typename FB<F>::Rec *stage = reinterpret_cast<typename FB<F>::Rec *>(hbuf);
        int32_t *cnt = B.count[C.cnt ^ 1] + rp.c_off;
        const int lazy_k = P.lazy ? C.k_since : 0;
        // before the first sort the bound is not set: take every particle's speed
        const bool all_v = !(C.vmax2 < INFINITY);
        const int32_t *lut = B.lut + rp.lut_off;
        const Rl rdt = (Rl)P.dt;
        int keys[KPT], ent[KPT], loc[KPT];
        double vm = 0;  // extra sort: |v|^2 maximum over the chunk's list members (all_v: all)
        const bool dfr = !forced && C.kind == 1;
        bool candf[KPT];
        double ccol[2][3];
        for (int v = 0; v < KPT; ++v) candf[v] = false;
        for (int cc = 0; cc < 2; ++cc)
            for (int k = 0; k < 3; ++k) ccol[cc][k] = dfr ? B.cpos[6 * r + 2 * k + cc] : 0.0;
        const double cand_r2 = (sqrt(P.radius_sq) + 1e-3) * (sqrt(P.radius_sq) + 1e-3);
        // an extra sort loads the positions and membership bytes of the thread's KPT
        // particles before it stores anything, so that their loads are in flight together
        Rl xs[KPT][3];
        int mbs[KPT];
        if (!forced)
            for (int v = 0; v < KPT; ++v) {
                const int64_t i = c0 + v * FT + threadIdx.x;
                if (i < n) {
                    xs[v][0] = pos[i];
                    xs[v][1] = pos[n + i];
                    xs[v][2] = pos[2 * n + i];
                    mbs[v] = memb[i];
                }
            }
        for (int v = 0; v < KPT; ++v) {
            const int64_t i = c0 + v * FT + threadIdx.x;
            keys[v] = -1;
            ent[v] = 0;
            loc[v] = 0;
            if (i < n) {
                Rl x[3];
                if (forced) {
                    x[0] = pos[i];
                    x[1] = pos[n + i];
                    x[2] = pos[2 * n + i];
                } else {
                    x[0] = xs[v][0];
                    x[1] = xs[v][1];
                    x[2] = xs[v][2];
                }
                int jmp[3];
                if (forced) {
                    const Rl vl[3] = {vel[i], vel[n + i], vel[2 * n + i]};
                    if (lazy_k > 0 && !memb[i]) {
                        // replay the k_since additions x = x + dt v of the lazily moved particle
                        for (int k = 0; k < 3; ++k) {
                            const Rl t = rdt * vl[k];
                            for (int s2 = 0; s2 < lazy_k; ++s2) x[k] = x[k] + t;
                        }
                    }
                    for (int k = 0; k < 3; ++k) {
                        Rl e = (Rl)b.edges[k];
                        int jump = (int)floor((double)(x[k] / e));
                        x[k] = x[k] - (Rl)jump * e;
                        jmp[k] = jump;
                    }
                    // the staging record: wrapped position (apply_pbc, md.f90:202-217), image,
                    // and the rest of the particle's state
                    typename FB<F>::Rec q;
                    for (int k = 0; k < 3; ++k) {
                        q.p[k] = x[k];
                        q.v[k] = vl[k];
                        q.im[k] = img[k * n + i] + jmp[k];
                    }
                    q.id = ids[i];
                    q.set_sf((int)spc[i], (int)flg[i]);
                    int key;
                    {
                        int p0, p1, p2;
                        cell_of<Rl>(b, x[0], x[1], x[2], p0, p1, p2);
                        if ((unsigned)p0 < (unsigned)b.L[0] && (unsigned)p1 < (unsigned)b.L[1] &&
                            (unsigned)p2 < (unsigned)b.L[2])
                            key = lut[p0 + b.L[0] * (p1 + b.L[1] * p2)];
                        else
                            key = compact_key(p0, p1, p2, b.M[0], b.M[1], b.M[2]);
                    }
                    q.key = key;
                    keys[v] = key;
                    stage[threadIdx.x] = q;  // (written out below)
                } else {
                    // extra sort: the particle stays in place; apply_pbc and pos_old = pos
                    // (single_dimer_pbc.f90:325-335) are done here
                    const int mb = mbs[v];
                    bool write = false;
                    if (lazy_k > 0 && !mb) {
                        for (int k = 0; k < 3; ++k) {
                            const Rl t = rdt * vel[k * n + i];
                            for (int s2 = 0; s2 < lazy_k; ++s2) x[k] = x[k] + t;
                        }
                        write = true;
                    }
                    for (int k = 0; k < 3; ++k) {
                        Rl e = (Rl)b.edges[k];
                        int jump = (int)floor((double)(x[k] / e));
                        x[k] = x[k] - (Rl)jump * e;
                        if (jump != 0) {
                            img[k * n + i] += jump;
                            write = true;
                        }
                    }
                    for (int k = 0; k < 3; ++k) {
                        if (write) pos[k * n + i] = x[k];
                        pold[k * n + i] = x[k];
                    }
                    if (mb || all_v) {
                        double v2 = 0;
                        for (int k = 0; k < 3; ++k) {
                            const double vv = (double)vel[k * n + i];
                            v2 += vv * vv;
                        }
                        vm = v2 > vm ? v2 : vm;
                        if (mb) memb[i] = 0;
                    }
                    int key, p0, p1, p2;
                    cell_of<Rl>(b, x[0], x[1], x[2], p0, p1, p2);
                    if ((unsigned)p0 < (unsigned)b.L[0] && (unsigned)p1 < (unsigned)b.L[1] &&
                        (unsigned)p2 < (unsigned)b.L[2])
                        key = lut[p0 + b.L[0] * (p1 + b.L[1] * p2)];
                    else
                        key = compact_key(p0, p1, p2, b.M[0], b.M[1], b.M[2]);
                    keys[v] = key;
                    if (dfr) {
                        // a deferred sort keeps the cell; the particle is a candidate of the
                        // list build when it lies within the list radius of either colloid
                        // (with a margin: the colloids may be read before or after their wrap)
                        for (int cc = 0; cc < 2; ++cc) {
                            const double d0 = min_image(ccol[cc][0] - (double)x[0], b.edges[0]);
                            const double d1 = min_image(ccol[cc][1] - (double)x[1], b.edges[1]);
                            const double d2 = min_image(ccol[cc][2] - (double)x[2], b.edges[2]);
                            candf[v] = candf[v] || (d0 * d0 + d1 * d1) + d2 * d2 < cand_r2;
                        }
                    }
                }
                if (dfr)
                    B.dkey[rp.p_off + i] = keys[v];
                else
                    B.key[rp.p_off + i] = keys[v];
            }
            if (forced) {
                // the records leave through LDS: each store writes one contiguous 16-byte
                // piece per thread, where each thread's own record would take pieces a
                // record apart
                __syncthreads();
                constexpr int RW = (int)(sizeof(typename FB<F>::Rec) / 16);
                const int64_t s0 = c0 + v * FT, nv = n - s0 < FT ? n - s0 : FT;
                uint4 *dst = reinterpret_cast<uint4 *>(rec + s0);
                const uint4 *src = reinterpret_cast<const uint4 *>(hbuf);
                for (int e = threadIdx.x; e < RW * FT; e += FT)
                    if (e < RW * nv) dst[e] = src[e];
                __syncthreads();
            }
        }
        if (!forced) {
            const double m = block_max(vm, red);
            if (threadIdx.x == 0) B.gpart[u] = m;
        }
        if (dfr) {
            // append the chunk's candidates: a block scan, then one atomic per chunk
            int c = 0;
            for (int v = 0; v < KPT; ++v) c += candf[v] ? 1 : 0;
            int tot;
            int o = scan_block_exclusive(c, lds, &tot);
            if (tot) {
                if (threadIdx.x == 0) hcnt[0] = atomicAdd(&C.ccount, tot);
                __syncthreads();
                const int64_t base = hcnt[0];
                int32_t *cand = B.cand + (int64_t)r * B.cand_cap;
                for (int v = 0; v < KPT; ++v)
                    if (candf[v]) {
                        if (base + o < B.cand_cap) cand[base + o] = (int32_t)(c0 + v * FT + threadIdx.x);
                        ++o;
                    }
            }
        }
        if (!dfr) {
            // the LDS hash after all of the thread's loads are issued
            for (int q = threadIdx.x; q < KHASH; q += FT) {
                hkey[q] = -1;
                hcnt[q] = 0;
            }
            __syncthreads();
            for (int v = 0; v < KPT; ++v) {
                const int key = keys[v];
                if (key < 0) continue;
                // open addressing; the table has twice as many entries as a chunk has keys
                unsigned h = ((unsigned)key * 2654435761u) >> (32 - KHASH_BITS);
                for (int probe = 0; probe < KHASH; ++probe) {
                    const int prev = atomicCAS(&hkey[h], -1, key);
                    if (prev == -1 || prev == key) break;
                    h = (h + 1) & (KHASH - 1);
                }
                ent[v] = (int)h;
                loc[v] = atomicAdd(&hcnt[h], 1);
            }
            __syncthreads();
            {
                // the occupied entries compacted first, so that a thread waits for at most
                // one global atomic per round (a chunk has a few hundred distinct keys)
                constexpr int PER = KHASH / FT;
                int c = 0;
                for (int j = 0; j < PER; ++j) c += hkey[threadIdx.x * PER + j] >= 0 ? 1 : 0;
                int tot;
                int o = scan_block_exclusive(c, lds, &tot);
                for (int j = 0; j < PER; ++j)
                    if (hkey[threadIdx.x * PER + j] >= 0) occ[o++] = threadIdx.x * PER + j;
                __syncthreads();
                for (int e = threadIdx.x; e < tot; e += FT) {
                    const int q = occ[e];
                    hcnt[q] = atomicAdd(&cnt[hkey[q]], hcnt[q]);
                }
            }
            __syncthreads();
            for (int v = 0; v < KPT; ++v) {
                const int64_t i = c0 + v * FT + threadIdx.x;
                if (keys[v] < 0) continue;
                B.slot[rp.p_off + i] = hcnt[ent[v]] + loc[v];
            }
        }
        __syncthreads();
        if (c0 == 0 && threadIdx.x == 0) {
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
    }
}

// the collision sort's key pass and the extra sorts' one are separate kernels, so that
// each is given the registers of its own path (the collision sort's fits six waves per SIMD)
template <class F>
__global__ void __launch_bounds__(FT) __attribute__((amdgpu_waves_per_eu(6))) k_fz_key_collision(FB<F> B, FP P, int R)
{
    key_pass<F, 1>(B, P, R);
}
template <class F>
__global__ void __launch_bounds__(FT) k_fz_key(FB<F> B, FP P, int R)
{
    key_pass<F, 0>(B, P, R);
}

// S2: exclusive scan of the new cell counts inside tiles of SCAN_TILE cells (tile-local
// offsets into slocal, tile totals into scan_sums); a unit is a tile
template <class F>
__global__ void __launch_bounds__(SCAN_T) k_fz_scan_tiles(FB<F> B, int R)
{
    __shared__ int lds[SCAN_T / 64];
    const UnitRange w = unit_range(B.plan, R, U_SCAN, B.xcd_order);
    int a = -1;
    for (int64_t u = w.u0; u < w.u1; u += w.step) {
        if (a < 0 || u >= w.pre[a + 1]) a = find_unit(w.pre, w.nact, u);
        const int r = (int)w.act[a];
        const int64_t tile = u - w.pre[a];
        const CtlDev &C = B.ctl[r];
        const RepDev &rp = B.rep[r];
        const int64_t nc = rp.box.ncells;
        const int64_t tile0 = tile * SCAN_TILE;
        const int32_t *in = B.count[C.cnt ^ 1] + rp.c_off;
        int32_t *out = B.slocal + rp.c_off;
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
        if (threadIdx.x == 0) B.scan_sums[r * B.scan_stride + tile] = tot;
    }
}

// S3: a unit is PT particles (pair[start + slot] = (reference position, storage index))
// or PT cells (the final
// cell_start = tile-local offset + tile prefix); the tile prefix of a replica is scanned
// in LDS by every block that works on it.  The block with a replica's first unit marks
// the buffers the sort writes; in a collision sort it also zeroes the destination
// buffers' force slots of the members of the lists saved when those buffers were last in
// use (only those can hold a force; the gather copies forces only where they are
// non-zero).
template <class F>
__global__ void __launch_bounds__(FT) k_fz_place(FB<F> B, int forced, int R)
{
    using Rl = typename F::real;
    __shared__ int pre[MAX_TILES];
    __shared__ int lds[FT / 64];
    const UnitRange w = unit_range(B.plan, R, U_PLACE, B.xcd_order);
    int a = -1, r_pre = -1;
    for (int64_t u = w.u0; u < w.u1; u += w.step) {
        if (a < 0 || u >= w.pre[a + 1]) a = find_unit(w.pre, w.nact, u);
        const int r = (int)w.act[a];
        const int64_t local = u - w.pre[a];
        CtlDev &C = B.ctl[r];
        const RepDev &rp = B.rep[r];
        const int64_t n = rp.n, nc = rp.box.ncells;
        if (r != r_pre) {
            const int nt = (int)((nc + SCAN_TILE - 1) / SCAN_TILE);
            const int32_t *sums = B.scan_sums + r * B.scan_stride;
            const int per = (nt + FT - 1) / FT;
            const int t0 = threadIdx.x * per;
            int loc = 0;
            for (int t = t0; t < t0 + per && t < nt; ++t) loc += sums[t];
            int tot;
            int off = scan_block_exclusive(loc, lds, &tot);
            for (int t = t0; t < t0 + per && t < nt; ++t) {
                pre[t] = off;
                off += sums[t];
            }
            __syncthreads();
            r_pre = r;
        }
        const int32_t *sl = B.slocal + rp.c_off;
        const int64_t np_units = (n + PT - 1) / PT;
        if (local < np_units) {
            // a slot's entry: the particle's reference position (its storage index unless
            // an extra sort left the particles in place) and its storage index
            const int32_t *key = B.key + rp.p_off, *slot = B.slot + rp.p_off, *inv = B.inv + rp.p_off;
            int2 *pair = B.pair + rp.p_off;
            const bool virt_in = C.virt != 0;
            for (int64_t i = local * PT + threadIdx.x; i < n && i < (local + 1) * PT; i += FT) {
                const int c = key[i];
                pair[sl[c] + pre[c / SCAN_TILE] + slot[i]] = make_int2(virt_in ? inv[i] : (int32_t)i, (int32_t)i);
            }
        } else {
            int32_t *start = B.start + rp.c_off;
            const int64_t c0 = (local - np_units) * PT;
            for (int64_t c = c0 + threadIdx.x; c < nc && c < c0 + PT; c += FT) start[c] = sl[c] + pre[c / SCAN_TILE];
        }
        if (local == 0 && !forced) {
            if (threadIdx.x == 0) {
                C.cur_next = C.cur;  // an extra sort leaves the particles in place
                C.cnt_next = C.cnt ^ 1;
            }
        } else if (local == 0) {
            const int dst = 1 - C.cur;
            const int64_t po3 = 3 * rp.p_off;
            Rl *f0 = B.force[2 * dst] + po3, *f1 = B.force[2 * dst + 1] + po3;
            const int32_t *ms = B.mslots + ((int64_t)r * 2 + dst) * 2 * B.nmax;
            const int mc0 = B.mcount[(r * 2 + dst) * 2], mc1 = B.mcount[(r * 2 + dst) * 2 + 1];
            for (int64_t e = threadIdx.x; e < 2 * (int64_t)B.nmax; e += FT) {
                const int col = (int)(e / B.nmax), k = (int)(e - (int64_t)col * B.nmax);
                if (k >= (col ? mc1 : mc0)) continue;
                const int q = ms[e];
                for (int kk = 0; kk < 3; ++kk) {
                    f0[kk * n + q] = 0;
                    f1[kk * n + q] = 0;
                }
            }
            if (threadIdx.x == 0) {
                C.cur_next = 1 - C.cur;
                C.cnt_next = C.cnt ^ 1;
            }
        }
    }
}

// The reference positions of the GHALO slots on either side of a unit's tile, so that
// the rank of a particle in a cell that straddles the tile edge reads LDS only (a cell
// holds far fewer than GHALO particles at the densities of the programs; a larger one
// reads the rest from memory).
__device__ inline void load_halo(int32_t *lv, int32_t *lk, const int2 *pair, const int32_t *dkey, int64_t t0,
                                 int64_t n)
{
    for (int q = threadIdx.x; q < 2 * GHALO; q += FT) {
        const int64_t j = q < GHALO ? t0 - GHALO + q : t0 + GT + (q - GHALO);
        const bool in = j >= 0 && j < n;
        const int2 pr = in ? pair[j] : make_int2(INT32_MAX, 0);
        lv[q < GHALO ? q : GT + q] = pr.x;
        if (dkey) lk[q < GHALO ? q : GT + q] = in ? dkey[pr.y] : INT32_MAX;
    }
}

// rank_in_cell when the previous extra sort was deferred: the reference order is (the
// deferred sort's cell, then the order before it), lexicographic
__device__ inline int rank_in_cell_deferred(const int32_t *lv, const int32_t *lk, const int2 *pair,
                                            const int32_t *dkey, int64_t t0, int s, int m, int32_t ref, int32_t dk)
{
    const int64_t w0 = t0 - GHALO, w1 = t0 + GT + GHALO;
    const int lo = (int)(s > w0 ? s : w0), hi = (int)(s + m < w1 ? s + m : w1);
    int rank = 0;
    for (int e = lo; e < hi; ++e) {
        const int32_t k = lk[e - w0];
        rank += (k < dk || (k == dk && lv[e - w0] < ref)) ? 1 : 0;
    }
    const int a = lo < s + m ? lo : s + m, b = hi > a ? hi : a;
    for (int e = s; e < s + m; ++e) {
        if (e >= a && e < b) continue;
        const int2 pr = pair[e];
        const int32_t k = dkey[pr.y];
        rank += (k < dk || (k == dk && pr.x < ref)) ? 1 : 0;
    }
    return rank;
}

// The number of a cell's arrivals [s, s + m) whose reference position is below ref: the
// window [t0 - GHALO, t0 + GT + GHALO) staged in LDS, anything outside it read from
// memory (two separate loops, so that the common one is LDS only).
__device__ inline int rank_in_cell(const int32_t *lv, const int2 *pair, int64_t t0, int s, int m, int32_t ref)
{
    const int64_t w0 = t0 - GHALO, w1 = t0 + GT + GHALO;
    const int lo = (int)(s > w0 ? s : w0), hi = (int)(s + m < w1 ? s + m : w1);
    int rank = 0;
    const int32_t *l = lv - w0;
    int e = lo;
    for (; e + 4 <= hi; e += 4) {  // four LDS reads in flight per wait
        const int32_t a0 = l[e], a1 = l[e + 1], a2 = l[e + 2], a3 = l[e + 3];
        rank += (a0 < ref ? 1 : 0) + (a1 < ref ? 1 : 0) + (a2 < ref ? 1 : 0) + (a3 < ref ? 1 : 0);
    }
    for (; e < hi; ++e) rank += l[e] < ref ? 1 : 0;
    // the parts before and after the window (empty unless the cell is larger than GHALO)
    const int a = lo < s + m ? lo : s + m, b = hi > a ? hi : a;
    for (int e = s; e < a; ++e) rank += pair[e].x < ref ? 1 : 0;
    for (int e = b; e < s + m; ++e) rank += pair[e].x < ref ? 1 : 0;
    return rank;
}

// S4: stable gather of the staging records into the other buffers, pos_old, |v|max per
// unit; clears the previous cell counts and the is_md/is_reac marks for the list build.  A
// unit is a tile of GT destination slots (GPT per thread, their loads issued together);
// the slots' reference positions are kept in LDS, where the rank of a particle among its
// cell's arrivals reads them (the reference's single-thread sort keeps the order of the
// reference positions within a cell).  Forces are copied only for particles that carry
// one (MD/PAST_MD bits); every other destination force slot is already zero (see
// k_fz_place).
template <class F>
__device__ inline void gather_pass(const FB<F> &B, const FP &P, int R)
{
    using Rl = typename F::real;
    using S = typename F::small;
    __shared__ double red[16];
    __shared__ int32_t ltmp[GT + 2 * GHALO], lkey[GT + 2 * GHALO];
    const UnitRange w = unit_range(B.plan, R, U_GATHER, B.xcd_order);
    int a = -1;
    for (int64_t u = w.u0; u < w.u1; u += w.step) {
        if (a < 0 || u >= w.pre[a + 1]) a = find_unit(w.pre, w.nact, u);
        const int r = (int)w.act[a];
        const int64_t t0 = (u - w.pre[a]) * GT;
        CtlDev &C = B.ctl[r];
        const RepDev &rp = B.rep[r];
        const int64_t n = rp.n, nc = rp.box.ncells;
        const int cur = C.cur, dst = 1 - cur, fp = C.fpar;
        const int64_t po3 = 3 * rp.p_off;
        const typename FB<F>::Rec *rec = B.rec + rp.p_off;
        const Rl *sf = B.force[2 * cur + fp] + po3, *sfo = B.force[2 * cur + 1 - fp] + po3;
        Rl *dpos = B.pos[dst] + po3, *dvel = B.vel[dst] + po3, *dold = B.pos_old + po3;
        Rl *df = B.force[2 * dst + fp] + po3, *dfo = B.force[2 * dst + 1 - fp] + po3;
        int32_t *dim = B.image[dst] + po3, *did = B.id[dst] + rp.p_off;
        S *dsp = B.species[dst] + rp.p_off, *dfl = B.flags[dst] + rp.p_off;
        const int2 *pair = B.pair + rp.p_off;
        const int32_t *start = B.start + rp.c_off, *cnt = B.count[C.cnt ^ 1] + rp.c_off;
        int32_t *old_cnt = B.count[C.cnt] + rp.c_off;
        // GPT destination slots per thread, their loads issued together (unconditionally: a
        // slot past the end reads the last one).  A slot holds the particle's reference
        // position (the rank's order) and its storage index.
        int32_t v[GPT], vs[GPT], s0[GPT], m0[GPT];
        typename FB<F>::Rec q[GPT];
        __syncthreads();
        for (int h = 0; h < GPT; ++h) {
            const int64_t j = t0 + h * FT + threadIdx.x;
            const int2 pr = pair[j < n ? j : n - 1];
            v[h] = j < n ? pr.x : INT32_MAX;
            vs[h] = pr.y;
            ltmp[GHALO + h * FT + threadIdx.x] = v[h];
        }
        for (int h = 0; h < GPT; ++h) q[h] = rec[vs[h]];
        for (int h = 0; h < GPT; ++h) {
            s0[h] = start[q[h].key];
            m0[h] = cnt[q[h].key];
        }
        // after a deferred extra sort the order also compares the cells it kept
        const int32_t *dkey = C.def ? B.dkey + rp.p_off : nullptr;
        int32_t dk[GPT];
        for (int h = 0; h < GPT; ++h) {
            dk[h] = dkey ? dkey[vs[h]] : 0;
            if (dkey) lkey[GHALO + h * FT + threadIdx.x] = v[h] == INT32_MAX ? INT32_MAX : dk[h];
        }
        load_halo(ltmp, lkey, pair, dkey, t0, n);
        __syncthreads();
        double vm = 0;
        for (int h = 0; h < GPT; ++h) {
            if (v[h] == INT32_MAX) continue;
            const int64_t d =
                s0[h] + (dkey ? rank_in_cell_deferred(ltmp, lkey, pair, dkey, t0, s0[h], m0[h], v[h], dk[h])
                              : rank_in_cell(ltmp, pair, t0, s0[h], m0[h], v[h]));
            const int fl = q[h].flags();
            const bool has_f = (fl & (MD_MASK | PAST_MD_MASK)) != 0;
            double v2 = 0;
            for (int k = 0; k < 3; ++k) {
                dpos[k * n + d] = q[h].p[k];
                dold[k * n + d] = q[h].p[k];
                dim[k * n + d] = q[h].im[k];
                const Rl vv = q[h].v[k];
                dvel[k * n + d] = vv;
                v2 += (double)vv * (double)vv;
                if (has_f) {
                    df[k * n + d] = sf[k * n + vs[h]];
                    dfo[k * n + d] = sfo[k * n + vs[h]];
                }
            }
            did[d] = q[h].id;
            dsp[d] = (S)q[h].species();
            dfl[d] = (S)fl;
            B.memb[rp.p_off + d] = 0;
            vm = v2 > vm ? v2 : vm;
        }
        for (int h = 0; h < GPT; ++h) {
            const int64_t jc = t0 + h * FT + threadIdx.x;
            if (jc < nc) {
                old_cnt[jc] = 0;
                B.is_md[rp.c_off + jc] = 0;
                B.is_reac[rp.c_off + jc] = 1;
            }
        }
        const double m = block_max(vm, red);
        if (threadIdx.x == 0) B.gpart[u] = m;
    }
}

// eight waves per SIMD without spills in both forms (one slot per thread)
template <class F>
__global__ void __launch_bounds__(FT) __attribute__((amdgpu_waves_per_eu(8))) k_fz_gather(FB<F> B, FP P, int R)
{
    gather_pass<F>(B, P, R);
}

// S4 of an extra sort: the particles stay where they are and only the order is computed.
// The stable rank of a slot's particle among its cell's arrivals (by reference position,
// read from LDS within the unit) gives its new reference position d: perm[d] = storage
// index, inv[storage index] = d.  Clears the previous cell counts and the is_md/is_reac
// marks like the gather.
template <class F>
__global__ void __launch_bounds__(FT) k_fz_rank(FB<F> B, int R)
{
    __shared__ int32_t lref[GT + 2 * GHALO], lkey[GT + 2 * GHALO];
    const UnitRange w = unit_range(B.plan, R, U_GATHER, B.xcd_order);
    int a = -1;
    for (int64_t u = w.u0; u < w.u1; u += w.step) {
        if (a < 0 || u >= w.pre[a + 1]) a = find_unit(w.pre, w.nact, u);
        const int r = (int)w.act[a];
        const int64_t t0 = (u - w.pre[a]) * GT;
        const CtlDev &C = B.ctl[r];
        const RepDev &rp = B.rep[r];
        const int64_t n = rp.n, nc = rp.box.ncells;
        const int2 *pair = B.pair + rp.p_off;
        const int32_t *key = B.key + rp.p_off;
        int32_t *perm = B.perm + rp.p_off, *inv = B.inv + rp.p_off;
        const int32_t *start = B.start + rp.c_off, *cnt = B.count[C.cnt ^ 1] + rp.c_off;
        int32_t *old_cnt = B.count[C.cnt] + rp.c_off;
        int32_t ref[GPT], sto[GPT], s0[GPT], m0[GPT];
        __syncthreads();
        for (int h = 0; h < GPT; ++h) {
            const int64_t j = t0 + h * FT + threadIdx.x;
            const int2 pr = pair[j < n ? j : n - 1];
            ref[h] = j < n ? pr.x : INT32_MAX;
            sto[h] = pr.y;
            lref[GHALO + h * FT + threadIdx.x] = ref[h];
        }
        const int32_t *dkey = C.def ? B.dkey + rp.p_off : nullptr;
        int32_t dk[GPT];
        for (int h = 0; h < GPT; ++h) {
            dk[h] = dkey ? dkey[sto[h]] : 0;
            if (dkey) lkey[GHALO + h * FT + threadIdx.x] = ref[h] == INT32_MAX ? INT32_MAX : dk[h];
        }
        load_halo(lref, lkey, pair, dkey, t0, n);
        for (int h = 0; h < GPT; ++h) {
            const int c = key[sto[h]];
            s0[h] = start[c];
            m0[h] = cnt[c];
        }
        __syncthreads();
        for (int h = 0; h < GPT; ++h) {
            if (ref[h] == INT32_MAX) continue;
            const int32_t d =
                s0[h] + (dkey ? rank_in_cell_deferred(lref, lkey, pair, dkey, t0, s0[h], m0[h], ref[h], dk[h])
                              : rank_in_cell(lref, pair, t0, s0[h], m0[h], ref[h]));
            perm[d] = sto[h];
            inv[sto[h]] = d;
        }
        for (int h = 0; h < GPT; ++h) {
            const int64_t jc = t0 + h * FT + threadIdx.x;
            if (jc < nc) {
                old_cnt[jc] = 0;
                B.is_md[rp.c_off + jc] = 0;
                B.is_reac[rp.c_off + jc] = 1;
            }
        }
    }
}

// S5 of a deferred extra sort: update_list (neighbor_list.f90:76-157) for colloid col from
// the key pass's candidates, with no cell lists.  A candidate belongs to stencil cell j
// when its cell is the colloid's cell plus stencil offset j (the box holds the stencil
// box, so j is unique); the list keeps the reference's order: stencil order, then, within
// a cell, the order before the sort (a stable sort keeps it).  Counting per stencil cell,
// a scan and a rank within each cell by that order place every entry.  Block 0 commits
// the sort; the particles and the buffers stay as they are.
constexpr int DEFER_BOX = 4096;  // stencil boxes up to (2m + 1)^3 cells
template <class F>
__device__ void deferred_list(FB<F> &B, const FP &P, int R, int r, int col, int *lds, double *red, short *dlut,
                              int *cntj, int *offj, int *cursor, int &nsel)
{
    CtlDev &C = B.ctl[r];
    const RepDev &rp = B.rep[r];
    Box b = rp.box;
    for (int k = 0; k < 3; ++k) b.origin[k] = C.origin[k];
    const int64_t n = rp.n;
    const int cur = C.cur;
    const typename F::real *pos = B.pos[cur] + 3 * rp.p_off;
    const typename F::small *sp = B.species[cur] + rp.p_off;
    const int32_t *inv = B.inv + rp.p_off;
    const bool virt_in = C.virt != 0;
    const double L0 = b.L[0] * b.a, L1 = b.L[1] * b.a, L2 = b.L[2] * b.a;  // neighbor_list.f90:92
    const int m = P.stencil_m, wd = 2 * m + 1, ns = P.n_stencil;
    for (int q = threadIdx.x; q < wd * wd * wd; q += blockDim.x) dlut[q] = -1;
    for (int j = threadIdx.x; j < ns; j += blockDim.x) {
        cntj[j] = 0;
        cursor[j] = 0;
    }
    if (threadIdx.x == 0) nsel = 0;
    __syncthreads();
    for (int j = threadIdx.x; j < ns; j += blockDim.x) {
        const int32_t *o = B.stencil + 3 * j;
        dlut[((o[0] + m) * wd + (o[1] + m)) * wd + (o[2] + m)] = (short)j;
    }
    const double *cp = B.cpos + 6 * r;
    const double x[3] = {cp[col], cp[2 + col], cp[4 + col]};
    int pc[3];
    cell_of<double>(b, x[0], x[1], x[2], pc[0], pc[1], pc[2]);
    int32_t *lst = B.list + (int64_t)(2 * r + col) * P.nmax;
    int32_t *msl = B.mslots + ((int64_t)r * 2 + cur) * 2 * P.nmax + (int64_t)col * P.nmax;
    int4 *stg = B.stage + ((int64_t)r * 2 + col) * 2 * P.nmax, *stg2 = stg + P.nmax;
    const int64_t ncand = C.ccount < B.cand_cap ? C.ccount : B.cand_cap;
    const int32_t *cand = B.cand + (int64_t)r * B.cand_cap;
    __syncthreads();
    constexpr int CU = 4;  // candidates per thread and round, their loads issued together
    for (int64_t e0 = threadIdx.x; e0 < ncand; e0 += CU * blockDim.x) {
        int iv[CU], sv[CU];
        typename F::real xv[CU][3];
        for (int u = 0; u < CU; ++u) {
            const int64_t e = e0 + u * blockDim.x;
            iv[u] = cand[e < ncand ? e : ncand - 1];
        }
        for (int u = 0; u < CU; ++u) {
            sv[u] = (int)sp[iv[u]];
            for (int k = 0; k < 3; ++k) xv[u][k] = pos[k * n + iv[u]];
        }
        for (int u = 0; u < CU; ++u) {
            if (e0 + u * blockDim.x >= ncand) break;
            const int i = iv[u];
            const int spc = sv[u];
            if (spc <= 0) continue;
            const typename F::real xi[3] = {xv[u][0], xv[u][1], xv[u][2]};
            int q[3];
            cell_of<typename F::real>(b, xi[0], xi[1], xi[2], q[0], q[1], q[2]);
            int box = 0;
            bool in = true;
            for (int k = 0; k < 3; ++k) {
                int o = (q[k] - pc[k]) % b.L[k];
                o += o < 0 ? b.L[k] : 0;
                o = o > b.L[k] / 2 ? o - b.L[k] : o;
                in = in && o >= -m && o <= m;
                box = box * wd + (o + m);
            }
            if (!in) continue;
            const int j = (int)dlut[box];
            if (j < 0) continue;
            const double d0 = min_image(x[0] - (double)xi[0], L0);
            const double d1 = min_image(x[1] - (double)xi[1], L1);
            const double d2 = min_image(x[2] - (double)xi[2], L2);
            const double rsq = (d0 * d0 + d1 * d1) + d2 * d2;
            if (!(rsq < P.radius_sq)) continue;
            const int at = atomicAdd(&nsel, 1);
            if (at < P.nmax) {
                stg[at] = make_int4(i, j, virt_in ? inv[i] : i, 0);
                atomicAdd(&cntj[j], 1);
            }
        }
    }
    __syncthreads();
    const int total = nsel;
    const bool over = total > P.nmax || C.ccount > B.cand_cap;
    int carry = 0;
    for (int base = 0; base < ns; base += blockDim.x) {
        const int j = base + threadIdx.x;
        int tot;
        const int off = carry + scan_block_exclusive(j < ns ? cntj[j] : 0, lds, &tot);
        if (j < ns) offj[j] = off;
        carry += tot;
    }
    __syncthreads();
    const int tl = over ? 0 : total;
    for (int e = threadIdx.x; e < tl; e += blockDim.x) {
        const int4 t = stg[e];
        stg2[offj[t.y] + atomicAdd(&cursor[t.y], 1)] = t;
    }
    __threadfence_block();
    __syncthreads();
    for (int e = threadIdx.x; e < tl; e += blockDim.x) {
        const int4 t = stg2[e];
        const int s0 = offj[t.y], c = cntj[t.y];
        int rank = 0;
        for (int q = s0; q < s0 + c; ++q) rank += stg2[q].z < t.z ? 1 : 0;
        lst[s0 + rank] = t.x;
        msl[s0 + rank] = t.x;
        atomic_or_flag(&B.memb[rp.p_off + t.x], 1 << col);
    }
    if (threadIdx.x == 0) {
        B.nlist[2 * r + col] = tl;
        B.mcount[(r * 2 + cur) * 2 + col] = tl;
        if (over) C.error = C.ccount > B.cand_cap ? ERR_CANDIDATES : RMPCD_ERR_OVERFLOW;
    }
    if (col != 0) return;
    // |v|max: the previous bound raised to the list members' speeds (the key pass's units)
    const int a = (int)plan_amap(B.plan, R)[r - B.r0];
    const int64_t *kpre = plan_pre(B.plan, R, U_KEY);
    double all = !(C.vmax2 < INFINITY) ? 0.0 : C.vmax2;
    for (int64_t q = kpre[a] + threadIdx.x; q < kpre[a + 1]; q += blockDim.x) all = fmax(all, B.gpart[q]);
    all = block_max(all, red);
    if (threadIdx.x != 0) return;
    C.vmax2 = all;
    C.def = 1;
    C.k_since = 0;
    C.need_exact = P.dt * sqrt(all) * (1.0 + 1e-6) + P.slack >= P.skin * 0.89 ? 1 : 0;
    C.n_extra += 1;
}

// S5: update_list (neighbor_list.f90:76-157) for colloid blockIdx.x, one workgroup:
// count per stencil cell, block scan, fill in stencil order; block 0 also commits the
// sort (|v|max from the gather partials, buffer flip, counters).  The blocks take the
// new buffer indices from cur_next/cnt_next (set by k_fz_place), so the flip by block 0
// cannot reach the other block.
template <class F>
__global__ void __launch_bounds__(1024) k_fz_commit_list(FB<F> B, FP P, int forced, int R)
{
    __shared__ int lds[16];
    __shared__ double red[16];
    const int r = B.r0 + blockIdx.y;
    CtlDev &C = B.ctl[r];
    if (!sort_active(C, forced)) return;
    // replicas left out of this sort by the plan (stopped at an earlier failed list
    // build; with no replica sorting the plan holds only its zero count); both blocks
    // read the same plan, so they agree even while one of them sets the error flag
    if (B.plan[0] == 0 || plan_amap(B.plan, R)[r - B.r0] < 0) return;
    __shared__ int skey[MAX_STENCIL], sstart[MAX_STENCIL], soff[MAX_STENCIL + 1], sflag[MAX_STENCIL];
    if (!forced && C.kind == 1) {
        __shared__ short dlut[DEFER_BOX];
        __shared__ int nsel;
        deferred_list<F>(B, P, R, r, blockIdx.x, lds, red, dlut, skey, soff, sstart, nsel);
        return;
    }
    const RepDev &rp = B.rep[r];
    const int cur = C.cur_next, cnt_live = C.cnt_next;  // the buffers the gather wrote
    Box b = rp.box;
    for (int k = 0; k < 3; ++k) b.origin[k] = C.origin[k];
    const int64_t n = rp.n;
    const typename F::real *pos = B.pos[cur] + 3 * rp.p_off;
    const typename F::small *sp = B.species[cur] + rp.p_off;
    const int32_t *cnt = B.count[cnt_live] + rp.c_off, *start = B.start + rp.c_off;
    uint8_t *is_md = B.is_md + rp.c_off, *is_reac = B.is_reac + rp.c_off;
    const double L0 = b.L[0] * b.a, L1 = b.L[1] * b.a, L2 = b.L[2] * b.a;  // neighbor_list.f90:92
    // The stencil's cells and their particles are flattened into one index range (cell
    // j owns [soff[j], soff[j+1])), so every thread tests one candidate at a time and the
    // loads of different candidates overlap; a block scan per chunk keeps the reference's
    // order (stencil order, then cell order).
    const int col = blockIdx.x;
    const double *cp = B.cpos + 6 * r;
    const double x[3] = {cp[col], cp[2 + col], cp[4 + col]};
    const int si = B.cspecies[2 * r + col];
    int32_t *lst = B.list + (int64_t)(2 * r + col) * P.nmax;
    int32_t *msl = B.mslots + ((int64_t)r * 2 + cur) * 2 * P.nmax + (int64_t)col * P.nmax;
    const int ns = P.n_stencil;
    int carry = 0;
    for (int base = 0; base < ns; base += blockDim.x) {
        const int j = base + threadIdx.x;
        int m = 0;
        if (j < ns) {
            const int key = stencil_key_dev(b, x, B.stencil, j);
            skey[j] = key;
            sstart[j] = start[key];
            sflag[j] = 0;
            m = cnt[key];
        }
        int tot;
        const int off = carry + scan_block_exclusive(m, lds, &tot);
        if (j < ns) soff[j] = off;
        carry += tot;
    }
    if (threadIdx.x == 0) soff[ns] = carry;
    __syncthreads();
    const int T = carry;
    int total = 0;
    for (int f0 = 0; f0 < T; f0 += blockDim.x) {
        const int f = f0 + threadIdx.x;
        int q = -1, in = 0;
        if (f < T) {
            int lo = 0, hi = ns - 1;  // the last j with soff[j] <= f
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (soff[mid] <= f)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            q = sstart[lo] + (f - soff[lo]);
            if (!forced) q = B.perm[rp.p_off + q];  // an extra sort: the storage index
            const int spc = (int)sp[q];
            if (spc > 0) {
                const double d0 = min_image(x[0] - (double)pos[q], L0);
                const double d1 = min_image(x[1] - (double)pos[n + q], L1);
                const double d2 = min_image(x[2] - (double)pos[2 * n + q], L2);
                const double rsq = (d0 * d0 + d1 * d1) + d2 * d2;
                if (rsq < P.radius_sq) {
                    in = 1;
                    const bool noreac = rsq <= P.sc.cut_sq[(spc - 1) + P.sc.n1 * (si - 1)];
                    atomicOr(&sflag[lo], noreac ? 3 : 1);
                }
            }
        }
        int tot;
        const int off = total + scan_block_exclusive(in, lds, &tot);
        if (in && off < P.nmax) {
            lst[off] = q;
            msl[off] = q;
            atomic_or_flag(&B.memb[rp.p_off + q], 1 << col);
        }
        total += tot;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < ns; j += blockDim.x) {
        const int fl = sflag[j];
        if (fl & 1) is_md[skey[j]] = 1;
        if (fl & 2) is_reac[skey[j]] = 0;
    }
    if (threadIdx.x == 0) {
        B.nlist[2 * r + col] = total > P.nmax ? P.nmax : total;
        B.mcount[(r * 2 + cur) * 2 + col] = total > P.nmax ? P.nmax : total;
        if (total > P.nmax) C.error = RMPCD_ERR_OVERFLOW;
    }
    if (col != 0) return;
    // |v|max over this replica's gather units; after an extra sort, the previous bound
    // raised to the list members' speeds (the key pass's units): only list members have
    // changed their velocity since the bound was taken
    const int a = (int)plan_amap(B.plan, R)[r - B.r0];
    const int64_t *gpre = plan_pre(B.plan, R, forced ? U_GATHER : U_KEY);
    double all = forced || !(C.vmax2 < INFINITY) ? 0.0 : C.vmax2;
    for (int64_t q = gpre[a] + threadIdx.x; q < gpre[a + 1]; q += blockDim.x) all = fmax(all, B.gpart[q]);
    all = block_max(all, red);
    if (threadIdx.x != 0) return;
    C.vmax2 = all;
    C.virt = forced ? 0 : 1;
    C.def = 0;
    C.cur = cur;
    C.cnt = cnt_live;
    C.k_since = 0;
    C.need_exact = P.dt * sqrt(all) * (1.0 + 1e-6) + P.slack >= P.skin * 0.89 ? 1 : 0;
    if (!forced) C.n_extra += 1;
    // need_sort stays set; the colloid kernel of the next step recomputes it
}

// ---- collision: SRD (mpcd.f90:120-180) + refuel or bulk reaction, |v|max ---------------
// refuel's test (single_dimer_pbc.f90:544-565): farther than far = (L(1)*0.45)**2, the
// product in single precision, from both colloids
template <class R>
__device__ inline bool refuel_far(const R *pos, int64_t n, int64_t i, const double *cp, const Box &box)
{
    float farf = (float)box.L[0] * 0.45f;
    farf = farf * farf;
    const double far = (double)farf;
    const double *L = box.edges;
    const double y0 = (double)pos[i], y1 = (double)pos[n + i], y2 = (double)pos[2 * n + i];
    const double a0 = min_image(cp[0] - y0, L[0]), a1 = min_image(cp[2] - y1, L[1]), a2 = min_image(cp[4] - y2, L[2]);
    const double b0 = min_image(cp[1] - y0, L[0]), b1 = min_image(cp[3] - y1, L[1]), b2 = min_image(cp[5] - y2, L[2]);
    return (a0 * a0 + a1 * a1) + a2 * a2 > far && (b0 * b0 + b1 * b1) + b2 * b2 > far;
}

// the SRD rotation of one cell (mpcd.f90:120-180 with the default alpha) on arrays
// vx, vy, vz, sp: the LDS tile or global memory, each addressed directly
template <class R, class S>
__device__ inline void fz_srd_cell(R *vx, R *vy, R *vz, const S *sp, int m, uint64_t seed, uint64_t cycle, int c)
{
    Acc3<R> acc;
    int n_eff = 0;
    for (int i = 0; i < m; ++i) {
        if ((int)sp[i] > 0) {
            acc.add(0, vx[i]);
            acc.add(1, vy[i]);
            acc.add(2, vz[i]);
            ++n_eff;
        }
    }
    if (n_eff <= 1) return;
    const R u0 = acc.sum(0) / (R)n_eff, u1 = acc.sum(1) / (R)n_eff, u2 = acc.sum(2) / (R)n_eff;
    Sphere ax = rand_sphere_threefry(seed, STREAM_SRD, cycle, (uint64_t)c);
    const double v0 = ax.n[0], v1 = ax.n[1], v2 = ax.n[2];
    // default alpha: sin = 1, cos = 0 (mpcd.f90:114-117)
    const double sa = 1, ca = 0, omc = 1 - ca;
    const R o[9] = {(R)(ca + (v0 * v0) * omc), (R)((v1 * v0) * omc + v2 * sa), (R)((v2 * v0) * omc - v1 * sa),
                    (R)((v0 * v1) * omc - v2 * sa), (R)(ca + (v1 * v1) * omc), (R)((v2 * v1) * omc + v0 * sa),
                    (R)((v0 * v2) * omc + v1 * sa), (R)((v1 * v2) * omc - v0 * sa), (R)(ca + (v2 * v2) * omc)};
    srd_rotate(vx, vy, vz, m, o, u0, u1, u2, acc, n_eff == m);
}

// S4 of a collision sort with the collision fused in (no bulk reaction): a unit is a tile
// of GC cells, so that all arrivals of its cells are in the unit.  The slots' reference
// positions (and deferred cells) go to LDS; each slot's particle is ranked among its
// cell's arrivals there and written to its destination, its velocity and species into an
// LDS image of the tile in destination order; then SRD per cell on that image (the
// reference's order within a cell, as k_fz_collide from the sorted arrays), then refuel
// (single_dimer_pbc.f90:544-565), |v|max and the write-back.  A tile of more than
// GC_TILE particles ranks and collides from memory.  Also clears the previous cell counts
// and the is_md/is_reac marks of its cells for the list build.
template <class F>
__global__ void __launch_bounds__(FT) k_fz_gather_srd(FB<F> B, FP P, int R)
{
    using Rl = typename F::real;
    using S = typename F::small;
    __shared__ Rl lx[GC_TILE], ly[GC_TILE], lz[GC_TILE];
    __shared__ S lsp[GC_TILE];
    __shared__ int32_t lref[GC_TILE], lkey[GC_TILE], lsrc[GC_TILE];
    __shared__ int32_t lstart[GC], lcnt[GC];
    __shared__ double red[16];
    const UnitRange w = unit_range(B.plan, R, U_GATHER, B.xcd_order);
    int a = -1;
    double vm = 0;
    for (int64_t u = w.u0; u < w.u1; u += w.step) {
        if (a < 0 || u >= w.pre[a + 1]) a = find_unit(w.pre, w.nact, u);
        const int r = (int)w.act[a];
        const int64_t c0 = (u - w.pre[a]) * GC;
        const CtlDev &C = B.ctl[r];
        const RepDev &rp = B.rep[r];
        const int64_t n = rp.n, nc = rp.box.ncells;
        const int ncl = (int)(c0 + GC < nc ? GC : nc - c0);
        const int cur = C.cur, dst = 1 - cur, fp = C.fpar;
        const int64_t po3 = 3 * rp.p_off;
        const typename FB<F>::Rec *rec = B.rec + rp.p_off;
        const Rl *sf = B.force[2 * cur + fp] + po3, *sfo = B.force[2 * cur + 1 - fp] + po3;
        Rl *dpos = B.pos[dst] + po3, *dvel = B.vel[dst] + po3, *dold = B.pos_old + po3;
        Rl *df = B.force[2 * dst + fp] + po3, *dfo = B.force[2 * dst + 1 - fp] + po3;
        int32_t *dim = B.image[dst] + po3, *did = B.id[dst] + rp.p_off;
        S *dsp = B.species[dst] + rp.p_off, *dfl = B.flags[dst] + rp.p_off;
        const int2 *pair = B.pair + rp.p_off;
        const int32_t *start = B.start + rp.c_off, *cnt = B.count[C.cnt ^ 1] + rp.c_off;
        const int32_t *dkey = C.def ? B.dkey + rp.p_off : nullptr;
        __syncthreads();  // the previous unit's LDS reads are done
        for (int t = threadIdx.x; t < ncl; t += FT) {
            lstart[t] = start[c0 + t];
            lcnt[t] = cnt[c0 + t];
        }
        __syncthreads();
        const int64_t sa = lstart[0], sb = (int64_t)lstart[ncl - 1] + lcnt[ncl - 1];
        const int tcount = (int)(sb - sa);
        const bool fits = tcount <= GC_TILE;
        if (fits) {
            for (int q = threadIdx.x; q < tcount; q += FT) {
                const int2 pr = pair[sa + q];
                lref[q] = pr.x;
                lsrc[q] = pr.y;
                if (dkey) lkey[q] = dkey[pr.y];
            }
            __syncthreads();
        }
        for (int q = threadIdx.x; q < tcount; q += FT) {
            const int2 pr = fits ? make_int2(lref[q], lsrc[q]) : pair[sa + q];
            const int32_t ref = pr.x;
            const typename FB<F>::Rec qr = rec[pr.y];
            const int cl = qr.key - (int)c0;
            const int s0 = lstart[cl], m0 = lcnt[cl];
            int rank = 0;
            if (fits) {
                const int e0 = (int)(s0 - sa);
                if (dkey) {
                    const int32_t dk = lkey[q];
                    for (int e = e0; e < e0 + m0; ++e) {
                        const int32_t k = lkey[e];
                        rank += (k < dk || (k == dk && lref[e] < ref)) ? 1 : 0;
                    }
                } else {
                    for (int e = e0; e < e0 + m0; ++e) rank += lref[e] < ref ? 1 : 0;
                }
            } else if (dkey) {
                const int32_t dk = dkey[pr.y];
                for (int e = s0; e < s0 + m0; ++e) {
                    const int2 pe = pair[e];
                    const int32_t k = dkey[pe.y];
                    rank += (k < dk || (k == dk && pe.x < ref)) ? 1 : 0;
                }
            } else {
                for (int e = s0; e < s0 + m0; ++e) rank += pair[e].x < ref ? 1 : 0;
            }
            const int64_t d = s0 + rank;
            const int fl = qr.flags();
            const bool has_f = (fl & (MD_MASK | PAST_MD_MASK)) != 0;
            for (int k = 0; k < 3; ++k) {
                dpos[k * n + d] = qr.p[k];
                dold[k * n + d] = qr.p[k];
                dim[k * n + d] = qr.im[k];
                if (has_f) {
                    df[k * n + d] = sf[k * n + pr.y];
                    dfo[k * n + d] = sfo[k * n + pr.y];
                }
            }
            did[d] = qr.id;
            dfl[d] = (S)fl;
            B.memb[rp.p_off + d] = 0;
            if (fits) {
                lx[d - sa] = qr.v[0];
                ly[d - sa] = qr.v[1];
                lz[d - sa] = qr.v[2];
                lsp[d - sa] = (S)qr.species();
            } else {
                for (int k = 0; k < 3; ++k) dvel[k * n + d] = qr.v[k];
                dsp[d] = (S)qr.species();
            }
        }
        // the tile's positions and (overflow) velocities in memory are read back below
        __syncthreads();
        for (int t = threadIdx.x; t < ncl; t += FT) {
            const int m = lcnt[t];
            const int64_t sc = lstart[t];
            if (m > 1) {
                if (fits)
                    fz_srd_cell(lx + (sc - sa), ly + (sc - sa), lz + (sc - sa), lsp + (sc - sa), m, rp.seed,
                                (uint64_t)C.cycle, (int)(c0 + t));
                else
                    fz_srd_cell(dvel + sc, dvel + n + sc, dvel + 2 * n + sc, dsp + sc, m, rp.seed, (uint64_t)C.cycle,
                                (int)(c0 + t));
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < tcount; q += FT) {
            const int64_t j = sa + q;
            Rl v0, v1, v2;
            int spc;
            if (fits) {
                v0 = lx[q];
                v1 = ly[q];
                v2 = lz[q];
                spc = (int)lsp[q];
            } else {
                v0 = dvel[j];
                v1 = dvel[n + j];
                v2 = dvel[2 * n + j];
                spc = (int)dsp[j];
            }
            const double q2 = ((double)v0 * (double)v0 + (double)v1 * (double)v1) + (double)v2 * (double)v2;
            vm = q2 > vm ? q2 : vm;
            const bool refuel = spc == 2 && refuel_far(dpos, n, j, B.cpos + 6 * r, rp.box);
            if (fits) {
                dvel[j] = v0;
                dvel[n + j] = v1;
                dvel[2 * n + j] = v2;
                dsp[j] = (S)(refuel ? 1 : spc);
            } else if (refuel) {
                dsp[j] = (S)1;
            }
        }
        int32_t *old_cnt = B.count[C.cnt] + rp.c_off;
        for (int t = threadIdx.x; t < ncl; t += FT) {
            old_cnt[c0 + t] = 0;
            B.is_md[rp.c_off + c0 + t] = 0;
            B.is_reac[rp.c_off + c0 + t] = 1;
        }
        const double m = block_max(vm, red);
        if (threadIdx.x == 0) B.gpart[u] = m;
        vm = 0;
    }
}

// CT cells per workgroup, one thread per cell, their particles staged in LDS (CTILE = 12 CT)
template <class F, int CT>
__global__ void __launch_bounds__(CT) k_fz_collide(FB<F> B, FP P)
{
    using R = typename F::real;
    using S = typename F::small;
    constexpr int CTILE = 12 * CT;
    __shared__ R lx[CTILE], ly[CTILE], lz[CTILE];
    __shared__ S lsp[CTILE];
    __shared__ double red[16];
    __shared__ int64_t s_range[2];
    const int r = B.r0 + blockIdx.y;
    const RepDev &rp = B.rep[r];
    CtlDev &C = B.ctl[r];
    if (C.error) return;
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
        // four slots per thread at a time: their loads are all in flight before the first
        // LDS store waits for one
        int q0 = threadIdx.x;
        for (; q0 + 3 * CT < tcount; q0 += 4 * CT) {
            R x[4], y[4], z[4];
            S sp4[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t g = base + q0 + u * CT;
                x[u] = vel[g];
                y[u] = vel[n + g];
                z[u] = vel[2 * n + g];
                sp4[u] = spc[g];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                lx[q0 + u * CT] = x[u];
                ly[q0 + u * CT] = y[u];
                lz[q0 + u * CT] = z[u];
                lsp[q0 + u * CT] = sp4[u];
            }
        }
        for (; q0 < tcount; q0 += CT) {
            lx[q0] = vel[base + q0];
            ly[q0] = vel[n + base + q0];
            lz[q0] = vel[2 * n + base + q0];
            lsp[q0] = spc[base + q0];
        }
        __syncthreads();
        const int c = c0 + threadIdx.x;
        if (c < c1) {
            const int m = cnt[c];
            const int64_t s = start[c];
            const bool in_lds = tcount > 0;
            const int o = (int)(s - base);
            // the LDS tile and global memory get separate calls, so each is addressed directly
            if (m > 1) {
                if (in_lds)
                    fz_srd_cell(lx + o, ly + o, lz + o, lsp + o, m, rp.seed, (uint64_t)C.cycle, c);
                else
                    fz_srd_cell(vel + s, vel + n + s, vel + 2 * n + s, spc + s, m, rp.seed, (uint64_t)C.cycle, c);
            }
            // with the tile in LDS, |v|max and refuel run over the tile's particles below
            if (!in_lds)
                for (int i = 0; i < m; ++i) {
                    const double a = vel[s + i], b2 = vel[n + s + i], c2 = vel[2 * n + s + i];
                    const double q2 = (a * a + b2 * b2) + c2 * c2;
                    vm = q2 > vm ? q2 : vm;
                }
            if (P.bulk) {
                // bulk_reaction B -> A (mpcd.f90:605-643), one draw threefry((cycle, cell), (seed, REACTION))
                if (m > 1 && B.is_reac[rp.c_off + c]) {
                    double local_rate = 0;
                    int64_t pick = -1;
                    if (in_lds) {
                        for (int i = 0; i < m; ++i)
                            if ((int)lsp[o + i] == 2) {
                                local_rate = local_rate + 1;
                                pick = s + i;
                            }
                    } else {
                        for (int i = 0; i < m; ++i)
                            if ((int)spc[s + i] == 2) {
                                local_rate = local_rate + 1;
                                pick = s + i;
                            }
                    }
                    local_rate = local_rate * P.bulk_rate;
                    uint64_t a0, a1;
                    threefry2x64((uint64_t)C.cycle, (uint64_t)c, rp.seed, STREAM_REACTION, a0, a1);
                    if (u01(a0) < (1 - exp(-local_rate * P.tau)) && pick >= 0) spc[pick] = (S)1;
                }
            } else if (!in_lds) {
                for (int i = 0; i < m; ++i)
                    if ((int)spc[s + i] == 2 && refuel_far(pos, n, s + i, B.cpos + 6 * r, rp.box)) spc[s + i] = (S)1;
            }
        }
        __syncthreads();
        // the tile's particles in parallel: |v|max, refuel, write-back (all coalesced)
        for (int q = threadIdx.x; q < tcount; q += CT) {
            const double a = lx[q], b2 = ly[q], c2 = lz[q];
            const double q2 = (a * a + b2 * b2) + c2 * c2;
            vm = q2 > vm ? q2 : vm;
            if (!P.bulk && lsp[q] == 2 && refuel_far(pos, n, base + q, B.cpos + 6 * r, rp.box)) spc[base + q] = (S)1;
            vel[base + q] = lx[q];
            vel[n + base + q] = ly[q];
            vel[2 * n + base + q] = lz[q];
        }
    }
    double mm = block_max(vm, red);
    if (threadIdx.x == 0) B.part[(int64_t)r * PS + blockIdx.x] = mm;
}

// |v|max after the collision and the cycle counter, one block per replica
// (ncollide < 0: the collision was made by the collision sort's fused gather, whose unit
// maxima the sort's plan locates; R is then the group's replica count)
template <class F>
__global__ void __launch_bounds__(1024) k_fz_collide_commit(FB<F> B, FP P, int ncollide, int R)
{
    __shared__ double red[16];
    const int r = B.r0 + blockIdx.y;
    CtlDev &C = B.ctl[r];
    if (C.error) return;
    double all = 0;
    if (ncollide >= 0) {
        for (int q = threadIdx.x; q < ncollide; q += blockDim.x) all = fmax(all, B.part[(int64_t)r * PS + q]);
    } else {
        const int a = (int)plan_amap(B.plan, R)[r - B.r0];
        const int64_t *gpre = plan_pre(B.plan, R, U_GATHER);
        if (a >= 0)
            for (int64_t q = gpre[a] + threadIdx.x; q < gpre[a + 1]; q += blockDim.x) all = fmax(all, B.gpart[q]);
    }
    all = block_max(all, red);
    if (threadIdx.x != 0) return;
    C.vmax2 = all;
    C.need_exact = P.dt * sqrt(all) * (1.0 + 1e-6) + P.slack >= P.skin * 0.89 ? 1 : 0;
    C.cycle += 1;
}

// list membership bits from the lists built at set-up (memb starts zeroed)
template <class F>
__global__ void k_fz_memb_init(FB<F> B, FP P)
{
    const int r = B.r0 + blockIdx.y;
    const int e = blockIdx.x * FT + threadIdx.x;
    const int col = e / P.nmax, k = e - col * P.nmax;
    if (col >= 2) return;
    const int cur = B.ctl[r].cur, nl = min(B.nlist[2 * r + col], P.nmax);
    if (k == 0) B.mcount[(r * 2 + cur) * 2 + col] = nl;
    if (k < nl) {
        const int q = B.list[(int64_t)(2 * r + col) * P.nmax + k];
        atomic_or_flag(&B.memb[B.rep[r].p_off + q], 1 << col);
        B.mslots[((int64_t)r * 2 + cur) * 2 * P.nmax + (int64_t)col * P.nmax + k] = q;
    }
}

// ---------------------------------------------------------------------------------
constexpr int MAX_GROUPS = 8;
static int64_t plan_len(int R) { return 1 + U_KINDS + 2 * (int64_t)R + (int64_t)U_KINDS * (R + 1); }
static int64_t gather_units(const Replica &rp)
{
    // the units of a gather or rank pass (slot tiles) and of the fused gather (cell tiles)
    return std::max<int64_t>((std::max<int64_t>(rp.n, rp.box.ncells) + GT - 1) / GT, (rp.box.ncells + GC - 1) / GC);
}

// the kernels' view of the buffers, for replica group g (replicas from r0 on; its plan,
// and its partials from gpart_off on)
template <class F>
static FB<F> make_fb(DimerSim &s, int r0 = 0, int g = 0, int64_t gpart_off = 0)
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
    b.memb = s.memb;
    b.key = s.f_key;
    b.perm = s.f_perm;
    b.dkey = s.f_dkey;
    b.cand = s.f_cand;
    b.stage = reinterpret_cast<int4 *>(s.f_stage);
    b.cand_cap = s.cand_cap;
    b.inv = s.f_inv;
    b.pair = reinterpret_cast<int2 *>(s.f_pair);
    b.rec = reinterpret_cast<typename FB<F>::Rec *>(s.f_rec);
    b.slot = s.f_slot;
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
    b.mslots = s.f_mslots;
    b.plan = s.f_plan + g * plan_len(s.R);
    b.gpart = s.f_gpart + gpart_off;
    b.r0 = r0;
    b.mcount = s.f_mcount;
    b.nmax = s.nmax;
    b.xcd_order = s.sort_order == 2 ? 1 : 0;
    b.fuse_srd = s.cfg.bulk_rmpcd ? 0 : 1;
    b.slocal = s.f_slocal;
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
    p.lazy = s.cfg.lazy ? 1 : 0;
    p.mass[0] = s.mass[0];
    p.mass[1] = s.mass[1];
    p.nmax = s.nmax;
    p.n_stencil = s.n_stencil;
    p.stencil_m = s.stencil_m;
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

// grid of the step kernel: the member blocks, then the stream blocks
static int member_blocks(const DimerSim &s) { return (int)((2 * (int64_t)s.nmax + FT - 1) / FT); }
static int stream_blocks(const DimerSim &s)
{
    // lazy mode: the stream role only measures displacements, and only for an exact test
    const int cap = s.cfg.lazy ? std::min(s.stream_blocks, LAZY_STREAM_BLOCKS) : s.stream_blocks;
    // with 16-byte accesses in every replica a thread takes 16 / sizeof(real) particles of
    // each plane, and the grid needs no more blocks than that covers (blocks beyond it
    // would only be dispatched and return: 0.08 ms per cycle of the 16 x L=64 batch)
    bool vec = !s.cfg.lazy;
    for (const Replica &rp : s.rep) vec = vec && rp.n % 4 == 0 && rp.p_off % 4 == 0;
    const int64_t per = (int64_t)FT * (vec ? 16 / (s.cfg.form == RMPCD_FP64 ? 8 : 4) : 1);
    return (int)std::min<int64_t>((s.nmax_rep + per - 1) / per, cap);
}

// (each enqueue_* works on the R replicas of b's group)
template <class F>
static void enqueue_sort(DimerSim &s, const FB<F> &b, const FP &p, hipStream_t st, int R, int forced)
{
    const unsigned Rr = (unsigned)R;
    // the sorts that may follow each MD step get their own grid: when none is due their
    // blocks only read the plan and return, and fewer blocks return sooner
    const int nb = forced ? s.sort_blocks : s.extra_blocks;
    k_fz_plan<F><<<1, FT, 0, st>>>(b, forced, R);
    if (forced)
        k_fz_key_collision<F><<<nb, FT, 0, st>>>(b, p, R);
    else
        k_fz_key<F><<<nb, FT, 0, st>>>(b, p, R);
    k_fz_scan_tiles<F><<<SCAN_BLOCKS, SCAN_T, 0, st>>>(b, R);
    k_fz_place<F><<<nb, FT, 0, st>>>(b, forced, R);
    if (forced && b.fuse_srd)
        // four times the sort grid: more tiles in flight per XCD find more of the records
        // they read in its L2 (measured fetch: 1.17x the used bytes against 1.32x)
        k_fz_gather_srd<F><<<GATHER_GRID_MUL * nb, FT, 0, st>>>(b, p, R);
    else if (forced)
        k_fz_gather<F><<<nb, FT, 0, st>>>(b, p, R);
    else
        k_fz_rank<F><<<nb, FT, 0, st>>>(b, R);
    k_fz_commit_list<F><<<dim3(2, Rr), 1024, 0, st>>>(b, p, forced, R);
}

// an MD step in two parts: the step kernel (the streaming), then the colloids and the
// sort that may follow
template <class F>
static void enqueue_step_kernel(DimerSim &s, const FB<F> &b, const FP &p, hipStream_t st, int R, int mode)
{
    const int nbm = member_blocks(s);
    const unsigned grid = (unsigned)(nbm + (mode <= 1 ? stream_blocks(s) : 0));
    k_fz_step<F><<<dim3(grid, (unsigned)R), FT, 0, st>>>(b, p, mode, nbm);
}
template <class F>
static void enqueue_step_rest(DimerSim &s, const FB<F> &b, const FP &p, hipStream_t st, int R, int mode)
{
    k_fz_colloid<F><<<dim3(1, (unsigned)R), COLLOID_T, 0, st>>>(b, p, mode, member_blocks(s),
                                                            mode <= 1 ? stream_blocks(s) : 0);
    if (mode <= 1) enqueue_sort<F>(s, b, p, st, R, 0);
}

template <class F>
static void enqueue_collision(DimerSim &s, const FB<F> &b, const FP &p, hipStream_t st, int R)
{
    const unsigned Rr = (unsigned)R;
    const int ct = s.collide_cells;
    const unsigned gc = (unsigned)((s.ncmax_rep + ct - 1) / ct);
    enqueue_sort<F>(s, b, p, st, R, 1);
    if (b.fuse_srd) {
        // the collision was made by the sort's gather
        k_fz_collide_commit<F><<<dim3(1, Rr), 1024, 0, st>>>(b, p, -1, R);
        return;
    }
    if (ct == 64)
        k_fz_collide<F, 64><<<dim3(gc, Rr), 64, 0, st>>>(b, p);
    else if (ct == 256)
        k_fz_collide<F, 256><<<dim3(gc, Rr), 256, 0, st>>>(b, p);
    else
        k_fz_collide<F, 128><<<dim3(gc, Rr), 128, 0, st>>>(b, p);
    k_fz_collide_commit<F><<<dim3(1, Rr), 1024, 0, st>>>(b, p, (int)gc, R);
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
    const void *stream_fn = reinterpret_cast<const void *>(&k_fz_step<F>);
    const unsigned full = (unsigned)(member_blocks(s) + stream_blocks(s));
    size_t seen = 0, pair = 0;
    for (hipGraphNode_t x : order) {
        hipGraphNodeType t;
        RMPCD_HIP(hipGraphNodeGetType(x, &t));
        if (t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams kp;
        RMPCD_HIP(hipGraphKernelNodeGetParams(x, &kp));
        if (kp.func != stream_fn || kp.gridDim.x != full) continue;
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

// the first replica of each of G groups (1 <= G <= R) of about equal particle count, and R
// at the end: G + 1 increasing entries
static std::vector<int> group_ranges(const DimerSim &s, int G)
{
    int64_t tot = 0, acc = 0;
    for (const Replica &rp : s.rep) tot += rp.n;
    std::vector<int> r0(1, 0);
    for (int r = 0; r < s.R; ++r) {
        acc += s.rep[r].n;
        const int g = (int)r0.size();  // group g - 1 holds replicas r0[g - 1] .. r
        const int left = s.R - (r + 1);  // replicas for the G - g groups after it
        // close the group once it holds its share, or when each later group needs one of
        // the replicas left; never while a later group would get none
        if (g < G && left >= G - g && (acc * G >= tot * g || left == G - g)) r0.push_back(r + 1);
    }
    r0.push_back(s.R);
    return r0;
}

// One collision cycle per replica group, each group's kernels captured from its own
// stream (forked from and joined to fstream).  Group q starts after group q - 1's first
// step kernel and then runs free, so that the groups stay out of step: the latency-bound
// work after one group's step (its colloids, the sort that may follow) runs while another
// group streams.  (A dependency between streams costs about 10 us, too much to make the
// groups take turns at every step.)  A profiled graph keeps one group, so that every
// timed streaming launch has the device to itself.
template <class F>
static int build_graph(DimerSim &s)
{
    FP p = make_fp(s);
    hipStream_t st = s.fstream;
    const int G = s.prof_stride > 0 ? 1 : s.groups;
    const std::vector<int> r0 = group_ranges(s, G);
    bool ok = (int)r0.size() == G + 1;
    for (int q = 0; ok && q < G; ++q) ok = r0[q] < r0[q + 1];
    RMPCD_CHECK(ok, "dimer: internal error in the replica groups");
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
    hipError_t e = hipSuccess;
    if (G > 1) {
        e = hipEventRecord(s.gevents[0], st);
        for (int q = 1; q < G && e == hipSuccess; ++q) e = hipStreamWaitEvent(s.gstreams[q - 1], s.gevents[0], 0);
    }
    std::vector<FB<F>> bq;
    std::vector<hipStream_t> sq(G);
    int64_t goff = 0;
    for (int q = 0; q < G; ++q) {
        sq[q] = q == 0 ? st : s.gstreams[q - 1];
        bq.push_back(make_fb<F>(s, r0[q], q, goff));
        for (int r = r0[q]; r < r0[q + 1]; ++r) goff += gather_units(s.rep[r]);
    }
    // the first step streams with the last force; each later step first finishes its
    // predecessor; the last step (mode 2) is finished right before the collision
    for (int j = 0; j <= s.cfg.n_md && e == hipSuccess; ++j) {
        const int mode = j == 0 ? 0 : j < s.cfg.n_md ? 1 : 2;
        for (int q = 0; q < G && e == hipSuccess; ++q) {
            const int Rq = r0[q + 1] - r0[q];
            // group q starts after group q - 1's first step kernel
            if (G > 1 && j == 0 && q > 0) e = hipStreamWaitEvent(sq[q], s.gevents[G + q - 1], 0);
            if (e != hipSuccess) break;
            enqueue_step_kernel<F>(s, bq[q], p, sq[q], Rq, mode);
            if (G > 1 && j == 0) e = hipEventRecord(s.gevents[G + q], sq[q]);
            enqueue_step_rest<F>(s, bq[q], p, sq[q], Rq, mode);
        }
    }
    for (int q = 0; q < G && e == hipSuccess; ++q) {
        enqueue_collision<F>(s, bq[q], p, sq[q], r0[q + 1] - r0[q]);
        if (q > 0) {
            e = hipEventRecord(s.gevents[q], sq[q]);
            if (e == hipSuccess) e = hipStreamWaitEvent(st, s.gevents[q], 0);
        }
    }
    hipGraph_t g = nullptr;
    const hipError_t e_end = hipStreamEndCapture(st, &g);
    if (e == hipSuccess) e = e_end;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return set_error(RMPCD_ERR_HIP, "dimer: graph capture failed: %s", hipGetErrorString(e));
    }
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
    s.stream_blocks = s.cfg.stream_blocks > 0 ? s.cfg.stream_blocks : STREAM_BLOCKS;
    s.sort_blocks = s.cfg.sort_blocks > 0 ? s.cfg.sort_blocks : SORT_BLOCKS;
    s.extra_blocks = s.cfg.extra_sort_blocks > 0 ? s.cfg.extra_sort_blocks : s.sort_blocks;
    RMPCD_CHECK(s.cfg.sort_order >= 0 && s.cfg.sort_order <= 2, "dimer: sort_order must be 0, 1 or 2");
    RMPCD_CHECK(s.cfg.groups >= 0 && s.cfg.groups <= MAX_GROUPS, "dimer: groups must be 0 to 8");
    s.sort_order = s.cfg.sort_order > 0 ? s.cfg.sort_order : 2;
    RMPCD_CHECK(s.stream_blocks <= STREAM_BLOCKS && s.sort_blocks <= 65536 && s.extra_blocks <= 65536,
                "dimer: stream_blocks must be at most 16384 and sort_blocks, extra_sort_blocks at most 65536");
    RMPCD_CHECK(s.cfg.collide_cells == 0 || s.cfg.collide_cells == 64 || s.cfg.collide_cells == 128 ||
                    s.cfg.collide_cells == 256,
                "dimer: collide_cells must be 0, 64, 128 or 256");
    s.collide_cells = s.cfg.collide_cells > 0 ? s.cfg.collide_cells : CT_DEFAULT;
    RMPCD_CHECK((s.ncmax_rep + s.collide_cells - 1) / s.collide_cells <= PS && (int64_t)member_blocks(s) * MPV + STREAM_BLOCKS <= PS &&
                    member_blocks(s) <= MAX_MEMBER_BLOCKS && s.n_stencil <= MAX_STENCIL &&
                    (s.ncmax_rep + SCAN_TILE - 1) / SCAN_TILE <= MAX_TILES,
                "dimer: box or list too large for the fused kernels");
    if ((rc = s.alloc(&s.sol_pos_old, 3 * s.ntot * RS))) return rc;
    if ((rc = s.alloc((void **)&s.memb, s.ntot))) return rc;
    if ((rc = s.alloc((void **)&s.f_key, s.ntot * 4))) return rc;
    if ((rc = s.alloc(&s.f_rec, s.ntot * (s.cfg.form == RMPCD_FP32 ? sizeof(SortRec<float, uint8_t>)
                                                                     : sizeof(SortRec<double, int32_t>)))))
        return rc;
    if ((rc = s.alloc((void **)&s.f_slot, s.ntot * 4))) return rc;
    if ((rc = s.alloc((void **)&s.f_perm, s.ntot * 4))) return rc;
    // deferred extra sorts: the stencil's half width, the candidate buffers (the stencil
    // box with one cell of margin, for both colloids, at four times the mean density)
    {
        std::vector<int32_t> st(3 * (size_t)s.n_stencil);
        rmpcd_make_stencil(1.0, s.radius, st.data(), s.n_stencil);
        s.stencil_m = 0;
        for (int32_t o : st) s.stencil_m = std::max(s.stencil_m, std::abs((int)o));
        const int64_t wm = 2 * (int64_t)s.stencil_m + 3;
        s.cand_cap = 2 * wm * wm * wm * 4 * (int64_t)std::max(1, s.cfg.rho);
    }
    if ((rc = s.alloc((void **)&s.f_dkey, s.ntot * 4))) return rc;
    if ((rc = s.alloc((void **)&s.f_cand, (size_t)s.R * s.cand_cap * 4))) return rc;
    if ((rc = s.alloc(&s.f_stage, (size_t)s.R * 4 * s.nmax * 16))) return rc;
    if ((rc = s.alloc((void **)&s.f_inv, s.ntot * 4))) return rc;
    if ((rc = s.alloc((void **)&s.f_pair, s.ntot * 8))) return rc;
    if ((rc = s.alloc((void **)&s.lut, s.luttot * 4))) return rc;
    int64_t sstride = (s.ncmax_rep + SCAN_TILE - 1) / SCAN_TILE + 1;
    if ((rc = s.alloc((void **)&s.scan_sums, (size_t)s.R * sstride * 4))) return rc;
    if ((rc = s.alloc((void **)&s.f_slocal, s.nctot * 4))) return rc;
    if ((rc = s.alloc((void **)&s.f_mslots, (size_t)s.R * 4 * s.nmax * 4))) return rc;
    if ((rc = s.alloc((void **)&s.f_plan, MAX_GROUPS * plan_len(s.R) * 8))) return rc;
    int64_t gunits = 0;
    for (const Replica &rp : s.rep) gunits += gather_units(rp);
    if ((rc = s.alloc((void **)&s.f_gpart, (size_t)gunits * 8))) return rc;
    if ((rc = s.alloc((void **)&s.f_mcount, (size_t)s.R * 4 * 4))) return rc;
    if ((rc = s.alloc((void **)&s.part, (size_t)s.R * PS * 8))) return rc;
    if ((rc = s.alloc((void **)&s.rep_dev, sizeof(RepDev) * s.R))) return rc;
    if ((rc = s.alloc((void **)&s.ctl_dev, sizeof(CtlDev) * s.R))) return rc;
    RMPCD_HIP(hipStreamCreate(&s.fstream));
    // replica groups: streams to capture the groups after the first from, events to fork
    // and join them
    // (events 0 .. groups - 1: the fork, then the joins; groups .. 2 groups - 1: each
    // group's turn at the step kernel)
    s.groups = std::min(s.cfg.groups > 0 ? s.cfg.groups : 2, s.R);
    for (int q = 0; q < 2 * s.groups; ++q) {
        hipEvent_t ev;
        RMPCD_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        s.gevents.push_back(ev);
        if (q == 0 || q >= s.groups) continue;
        hipStream_t sq;
        RMPCD_HIP(hipStreamCreate(&sq));
        s.gstreams.push_back(sq);
    }
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
        rd[r].vec = (rp.n % 4 == 0 && rp.p_off % 4 == 0) ? 1 : 0;
        {
            // a deferred extra sort needs each stencil cell once (the box holds the
            // stencil box) and the stencil box in LDS
            const int wd = 2 * s.stencil_m + 1;
            bool ok = wd * wd * wd <= DEFER_BOX;
            for (int k = 0; k < 3; ++k) ok = ok && rp.box.L[k] >= wd;
            rd[r].defer = ok ? 1 : 0;
        }
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
        c.need_exact = 1;
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
    if (s.cfg.form == RMPCD_FP64)
        k_fz_memb_init<F64><<<dim3(member_blocks(s), s.R), FT, 0, s.fstream>>>(make_fb<F64>(s), make_fp(s));
    else
        k_fz_memb_init<F32><<<dim3(member_blocks(s), s.R), FT, 0, s.fstream>>>(make_fb<F32>(s), make_fp(s));
    RMPCD_LAUNCHED();
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
        rp.error = c.error == ERR_CANDIDATES ? RMPCD_ERR_OVERFLOW : c.error;
        if (c.error) err = c.error;
    }
    if (err == RMPCD_ERR_OVERFLOW) return set_error(RMPCD_ERR_OVERFLOW, "maximum of neighbor list reached");
    if (err == ERR_CANDIDATES)
        return set_error(RMPCD_ERR_OVERFLOW, "dimer: more particles near the colloids than the deferred extra "
                                             "sort's candidate buffer holds (%lld)", (long long)cand_cap);
    if (err) return set_error(err, "dimer: device error %d", err);
    return RMPCD_OK;
}

}  // namespace rmpcd
