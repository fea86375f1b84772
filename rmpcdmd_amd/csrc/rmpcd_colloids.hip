// rmpcd_colloids.hip -- the production loop of programs/n_colloids_pbc.f90 (:286-346) as a
// device-resident program: one collision cycle (N_MD velocity-Verlet steps with lazy
// cell-wise streaming, then the collision) is captured as one hipGraph and replayed,
// with every decision taken on the GPU, so no host scalar is read inside a run.
//
// It continues the state of a ProgramSim (rmpcd_programs.hip: set-up, first sort, list
// and forces) and reproduces the reference call sequence of ProgramSim::colloids_step /
// colloids_collide:
//   * cell_md_pos / cell_md_vel with md_flag true run over the md slot list: the slots of
//     the cells with is_md set by the last list build, collected at that build;
//   * the skin test (cell_maximum_displacement over those cells, the colloids'
//     maximum_displacement) is reduced by one block, which sets need_sort and the catch-up
//     step k dt;
//   * the sort chain (catch-up of the other cells, apply_pbc, key and count, scan, place,
//     stable gather into the other buffer, list build, md slot list, pos_old) is launched
//     every step and returns at once unless need_sort is set (a graph cannot branch); the
//     buffer that holds the solvent is chosen on the device (ctl.cur);
//   * forces: the force parity flips every step as in the reference, so a graph of one
//     cycle fixes it (two graphs, one per starting parity, when N_MD is odd).  The solvent
//     force and the MD-bit roll of compute_force are applied to the md slots only: every
//     particle that carries a force or an MD/PAST_MD bit is within the list radius of a
//     colloid (the skin test rebuilds the list before one can leave it), hence in an is_md
//     cell, and for every other particle the reference's full passes change nothing.
#include "rmpcd_pairforce.hpp"
#include "rmpcd_programs.hpp"

namespace rmpcd {

constexpr int CG_T = 256;        // threads per block
constexpr int CG_GRID = 2048;    // blocks of the grid-stride passes over particles or cells
constexpr int CG_MD_GRID = 4096; // blocks of the md-slot pass (measured against 1024 to 8192)
constexpr int CG_SCAN_I = 8;     // cells per thread of a scan tile
constexpr int CG_TILE = CG_T * CG_SCAN_I;
constexpr int CG_CELL_T = 128;   // cells per block of the collision and temperature passes
constexpr int CG_NL_T = 256;     // threads (and stencil cells at a time) of a colloid's list build
constexpr int CG_DEC_T = 1024;   // threads of the colloid step and skin test (one workgroup)
constexpr int CG_GC = 128, CG_GC_TILE = 12 * CG_GC;  // cells per gather unit, its LDS slots

// device control of the loop
struct ColCtl {
    int32_t cur;        // the solvent buffer holding the state
    int32_t need_sort;  // the skin test of this MD step fired
    int32_t error;      // a list longer than its capacity: the reference stops (error stop)
    int32_t n_md;       // entries of the md slot list
    int64_t md_steps, last_sort, n_extra, cycles;
    double catch_dt;    // k dt of the catch-up of the sort in progress
    double origin[3];   // grid shift of the last collision
    double e1, e2, temperature;
};

template <class F>
struct SolB {
    using R = typename F::real;
    using S = typename F::small;
    int64_t n;
    R *pos[2], *vel[2], *force[4];  // force of buffer b and parity f: force[2b + f]
    int32_t *image[2], *id[2];
    S *species[2], *flags[2];
};

struct ColB {
    int n;
    double *pos, *pos_old, *vel, *force[2], *mass;
    int32_t *image, *species;
};

struct ColloidsGraph {
    ColCtl *ctl = nullptr;
    int32_t *key = nullptr, *slot = nullptr, *tmp = nullptr, *tsum = nullptr, *mdslot = nullptr;
    uint8_t *mdflag = nullptr;
    int32_t *mdoff = nullptr;
    int32_t *lut = nullptr;  // the box's Hilbert table, a copy owned by the program
    void *rec = nullptr, *frec = nullptr;  // staging records of a sort, forces of the members
    double *part = nullptr, *cpart = nullptr, *fpart = nullptr, *npart = nullptr, *epart = nullptr, *tpart = nullptr;
    int ntiles = 0, fblocks = 0;
    hipGraphExec_t exec[2] = {nullptr, nullptr};  // by starting force parity
    std::vector<void *> allocs;
};

__device__ inline bool cg_off(const ColCtl *C, int forced) { return C->error || !(forced || C->need_sort); }

__device__ inline int64_t cg_tid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline int64_t cg_stride() { return (int64_t)gridDim.x * blockDim.x; }

// An XCD-grouped walk over tiles of CG_T items: the workgroups of one XCD (consecutive
// workgroups go to the XCDs in turn, so blockIdx mod 8) take one contiguous eighth of the
// tiles, interleaved, so that the tiles in flight on an XCD are neighbours and share its
// L2 (the sort's place and gather passes, whose scattered accesses stay near the tile)
struct CgWalk {
    int64_t t, t1, step;
};
__device__ inline CgWalk cg_walk_units(int64_t tiles)
{
    const int G = (int)gridDim.x;
    CgWalk w;
    if (G % 8 == 0) {
        const int64_t per = (tiles + 7) / 8, lo = (int64_t)(blockIdx.x % 8) * per;
        w.t = lo + blockIdx.x / 8;
        w.t1 = lo + per < tiles ? lo + per : tiles;
        w.step = G / 8;
    } else {
        w.t = blockIdx.x;
        w.t1 = tiles;
        w.step = G;
    }
    return w;
}
__device__ inline CgWalk cg_walk(int64_t n) { return cg_walk_units((n + CG_T - 1) / CG_T); }

__device__ inline int cg_block_scan(int v, int *lds, int *total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int inc = wave_inclusive_scan(v);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
        const int t = lds[w];
        if (w < wid) off += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// ---- MD step --------------------------------------------------------------------------
// One pass over the md slots per MD step.  With PENDING, first cell_md_vel with md_flag
// true of the step before (md.f90:239-263), which in the reference ends that step; then
// cell_md_pos with md_flag true (md.f90:79-102) and the squared displacement from pos_old
// of each moved particle (cell_maximum_displacement, particle_system.f90:380-410), maximum
// per block; then the solvent part of this step's force reset (switch, zero, MD-bit roll;
// neighbor_list.f90:180-188), which the reference makes after a sort of the same step.
// Taking it before the sort gives the same state: only particles with the MD or PAST_MD
// bit carry a force or change under the roll, each of them sits in an md slot (it is within
// the list radius of a colloid), and the sort carries force planes and flags along.
// The first workgroups also make the colloids' part of the step, one thread per colloid,
// ahead of their share of the md slots: with PENDING the velocity of the step before
// (n_colloids_pbc.f90:312) and the step counter; the velocity-Verlet position
// (n_colloids_pbc.f90:289-291) and its displacement from pos_old (maximum_displacement,
// particle_system.f90:357-378), maximum per workgroup; colloids%force = 0 of this step's
// compute_force (the plane the switch makes the new force; neither the skin test nor a
// sort reads it).
template <class F, bool PENDING>
__global__ void __launch_bounds__(CG_T) k_cg_md(SolB<F> S, ColB c, ColCtl *C, int fp, double dt,
                                                const int32_t *__restrict__ mdslot, double *__restrict__ part,
                                                double *__restrict__ cpart)
{
    using R = typename F::real;
    using Sm = typename F::small;
    __shared__ double red[16];
    double m = 0;
    if ((int64_t)blockIdx.x * CG_T < c.n) {
        double cm = 0;
        const int i = blockIdx.x * CG_T + threadIdx.x;
        if (!C->error && i < c.n) {
            const double *Fc = c.force[fp];
            double *Fz = c.force[1 - fp];
            double r[3];
            for (int k = 0; k < 3; ++k) {
                const int64_t q = (int64_t)k * c.n + i;
                if (PENDING) c.vel[q] = c.vel[q] + (dt * (Fc[q] + Fz[q])) / (2 * c.mass[i]);
                c.pos[q] = (c.pos[q] + dt * c.vel[q]) + ((dt * dt) * Fc[q]) / (2 * c.mass[i]);
                Fz[q] = 0;
                r[k] = c.pos[q] - c.pos_old[q];
            }
            if (c.species[i] >= 0) cm = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
        }
        if (PENDING && i == 0 && !C->error) C->md_steps += 1;
        cm = block_max(cm, red);
        if (threadIdx.x == 0) cpart[blockIdx.x] = cm;
    }
    if (!C->error) {
        const int b = C->cur;
        const int64_t n = S.n;
        R *pos = S.pos[b], *vel = S.vel[b];
        const R *f = S.force[2 * b + fp], *pold = S.pos[1 - b];
        R *fz = S.force[2 * b + 1 - fp];  // force_old of the step before, the new force of this one
        const Sm *sp = S.species[b];
        Sm *flags = S.flags[b];
        const R rdt = (R)dt, dt_sq = (R)((dt * dt) / 2);
        const int nmd = C->n_md;
        for (int64_t k = cg_tid(); k < nmd; k += cg_stride()) {
            const int64_t j = mdslot[k];
            int fl = (int)flags[j];
            R v[3], fc[3], x[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                v[q] = vel[q * n + j];
                fc[q] = f[q * n + j];
            }
            if (PENDING) {
                if (!(fl & WALL_MASK)) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        v[q] = v[q] + (rdt * (fc[q] + fz[q * n + j])) / (R)2;
                        vel[q * n + j] = v[q];
                    }
                } else {
                    fl &= ~WALL_MASK;
                }
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                x[q] = (pos[q * n + j] + rdt * v[q]) + dt_sq * fc[q];
                pos[q * n + j] = x[q];
                fz[q * n + j] = 0;
            }
            fl = (fl & MD_MASK) ? (fl | PAST_MD_MASK) : (fl & ~PAST_MD_MASK);
            flags[j] = (Sm)(fl & ~MD_MASK);
            if ((int)sp[j] >= 0) {
                const double d0 = (double)(x[0] - pold[j]), d1 = (double)(x[1] - pold[n + j]),
                             d2 = (double)(x[2] - pold[2 * n + j]);
                const double r2 = (d0 * d0 + d1 * d1) + d2 * d2;
                m = r2 > m ? r2 : m;
            }
        }
    }
    m = block_max(m, red);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// the skin test (n_colloids_pbc.f90:293-296), one workgroup: so_max is the maximum over the
// md cells (part, from k_cg_md), raised to max_v k dt; co_max over the colloids (cpart);
// k = steps since the last sort
__global__ void __launch_bounds__(CG_DEC_T) k_cg_decide(ColCtl *C, const double *__restrict__ part, int np,
                                                        const double *__restrict__ cpart, int ncp, double dt,
                                                        double skin, double max_v)
{
    __shared__ double red[16];
    double a = 0, bm = 0;
    for (int q = threadIdx.x; q < np; q += blockDim.x) a = fmax(a, part[q]);
    for (int q = threadIdx.x; q < ncp; q += blockDim.x) bm = fmax(bm, cpart[q]);
    a = block_max(a, red);
    bm = block_max(bm, red);
    if (threadIdx.x != 0 || C->error) return;
    const double k1 = (double)(C->md_steps + 1 - C->last_sort);
    double so = sqrt(a);
    const double bnd = max_v * (dt * k1);
    so = so > bnd ? so : bnd;
    const double co = sqrt(bm);
    C->need_sort = (co >= skin * 0.1 || so >= skin * 0.9) ? 1 : 0;
    C->catch_dt = k1 * dt;
}

// ---- sort chain (extra sort when need_sort is set, collision sort when forced) -------
// the grid shift of the collision (cell_system.f90:153-163), threefry((s, 0)) a, b and
// threefry((s, 1)) a of stream SHIFT; the catch-up length since the last sort
__global__ void k_cg_shift(ColCtl *C, uint64_t seed, double dt)
{
    if (threadIdx.x || blockIdx.x || C->error) return;
    uint64_t a, b, c, d;
    const uint64_t s = (uint64_t)C->cycles;
    threefry2x64(s, 0, seed, STREAM_SHIFT, a, b);
    threefry2x64(s, 1, seed, STREAM_SHIFT, c, d);
    C->origin[0] = u01(a) - 1;
    C->origin[1] = u01(b) - 1;
    C->origin[2] = u01(c) - 1;
    C->catch_dt = (double)(C->md_steps - C->last_sort) * dt;
}

// the start of a sort: cell_count = 0, and apply_pbc of the colloids (md.f90:202-217) with
// colloids%pos_old = colloids%pos
__global__ void __launch_bounds__(CG_T) k_cg_sort_begin(ColB c, const ColCtl *C, int forced, double e0, double e1,
                                                        double e2, int32_t *__restrict__ count, int64_t nc)
{
    if (cg_off(C, forced)) return;
    for (int64_t q = cg_tid(); q < nc; q += cg_stride()) count[q] = 0;
    const double e[3] = {e0, e1, e2};
    for (int64_t i = cg_tid(); i < c.n; i += cg_stride())
        for (int k = 0; k < 3; ++k) {
            const int64_t q = (int64_t)k * c.n + i;
            const double x = c.pos[q];
            const int jump = (int)floor(x / e[k]);
            c.image[q] += jump;
            c.pos[q] = x - (double)jump * e[k];
            c.pos_old[q] = c.pos[q];
        }
}

// The staging record of a sort: a particle's caught-up, wrapped state in one piece, written
// by the key pass in storage order and read whole by the gather (one or two cache lines
// where the SoA arrays would cost one line each).  Forces travel apart (only particles
// with the MD or PAST_MD bit carry one; the others' are zero).
template <class R>
struct alignas(16) CgRec {
    R p[3], v[3];
    int32_t im[3], id, key;
    int32_t sf;  // species | flags << 16
};

// The key pass of a sort, a unit being a chunk of CG_KCH particles in storage order:
// cell_md_pos then cell_md_vel with md_flag false and the catch-up k dt for the particles
// outside the md cells (md.f90:79-102, 239-263), apply_pbc (md.f90:202-217), the cell key
// (cell_system.f90:86-129) and the staging record (the forces of the particles with the
// MD or PAST_MD bit in side records, zeroed in place); the chunk's keys are counted in an LDS
// hash, then one global atomic per distinct cell reserves the chunk's slots in that cell
// (the gather makes the order within a cell stable).
constexpr int CG_KPT = 4, CG_KCH = CG_T * CG_KPT;
constexpr int CG_KH_BITS = 11, CG_KH = 1 << CG_KH_BITS;  // hash entries, twice a chunk

template <class F>
__global__ void __launch_bounds__(CG_T) k_cg_key(SolB<F> S, const ColCtl *C, int fp, int forced, Box bx,
                                                 const int32_t *__restrict__ lut, const uint8_t *__restrict__ mdflag,
                                                 int32_t *__restrict__ count, int32_t *__restrict__ key,
                                                 int32_t *__restrict__ slot, CgRec<typename F::real> *__restrict__ rec,
                                                 typename F::real *__restrict__ frec)
{
    using R = typename F::real;
    // the LDS hash (hkey, hcnt, occ); each slab of CG_T records is staged in the same
    // memory before the hash is built, so that the record stores are contiguous
    __shared__ __align__(16) int hbuf[3 * CG_KH];
    int *hkey = hbuf, *hcnt = hbuf + CG_KH, *occ = hbuf + 2 * CG_KH;
    static_assert(sizeof(CgRec<R>) * CG_T <= sizeof(int) * 3 * CG_KH, "record slab in the hash LDS");
    CgRec<R> *stage = reinterpret_cast<CgRec<R> *>(hbuf);
    __shared__ int lds[CG_T / 64];
    if (cg_off(C, forced)) return;
    Box b = bx;
    for (int k = 0; k < 3; ++k) b.origin[k] = C->origin[k];
    const int bf = C->cur;
    const int64_t n = S.n;
    R *pos = S.pos[bf], *vel = S.vel[bf];
    R *f = S.force[2 * bf + fp], *fo = S.force[2 * bf + 1 - fp];
    const int32_t *img = S.image[bf], *ids = S.id[bf];
    const typename F::small *spc = S.species[bf], *flg = S.flags[bf];
    const double cdt = C->catch_dt;
    const R rdt = (R)cdt, dt_sq = (R)((cdt * cdt) / 2);
    const int64_t nch = (n + CG_KCH - 1) / CG_KCH;
    for (int64_t u = blockIdx.x; u < nch; u += gridDim.x) {
        const int64_t c0 = u * CG_KCH;
        int keys[CG_KPT], ent[CG_KPT], loc[CG_KPT];
        for (int v = 0; v < CG_KPT; ++v) {
            const int64_t i = c0 + v * CG_T + threadIdx.x;
            keys[v] = -1;
            ent[v] = 0;
            loc[v] = 0;
            if (i < n) {
                R x[3], vv[3], ff[3] = {0, 0, 0}, fof[3] = {0, 0, 0};
                int fl = (int)flg[i];
                const bool has_f = (fl & (MD_MASK | PAST_MD_MASK)) != 0;
                for (int q = 0; q < 3; ++q) {
                    x[q] = pos[q * n + i];
                    vv[q] = vel[q * n + i];
                    if (has_f) {
                        ff[q] = f[q * n + i];
                        fof[q] = fo[q * n + i];
                    }
                }
                if (!mdflag[i]) {
                    for (int q = 0; q < 3; ++q) x[q] = (x[q] + rdt * vv[q]) + dt_sq * ff[q];
                    if (!(fl & WALL_MASK)) {
                        for (int q = 0; q < 3; ++q) vv[q] = vv[q] + (rdt * (ff[q] + fof[q])) / (R)2;
                    } else {
                        fl &= ~WALL_MASK;
                    }
                }
                CgRec<R> r;
                for (int q = 0; q < 3; ++q) {
                    const R ed = (R)b.edges[q];
                    const int jump = (int)floor((double)(x[q] / ed));
                    x[q] = x[q] - (R)jump * ed;
                    r.p[q] = x[q];
                    r.v[q] = vv[q];
                    r.im[q] = img[q * n + i] + jump;
                }
                int p0, p1, p2, k;
                cell_of<R>(b, x[0], x[1], x[2], p0, p1, p2);
                if ((unsigned)p0 < (unsigned)b.L[0] && (unsigned)p1 < (unsigned)b.L[1] && (unsigned)p2 < (unsigned)b.L[2])
                    k = lut[p0 + b.L[0] * (p1 + b.L[1] * p2)];
                else
                    k = compact_key(p0, p1, p2, b.M[0], b.M[1], b.M[2]);
                r.id = ids[i];
                r.key = k;
                r.sf = (int)spc[i] | (fl << 16);
                stage[threadIdx.x] = r;  // (written out below)
                if (has_f)
                    for (int q = 0; q < 3; ++q) {
                        frec[6 * i + q] = ff[q];
                        frec[6 * i + 3 + q] = fof[q];
                        // the source buffer's forces are zero again once staged, so that
                        // the buffer the next sort writes needs forces for flagged
                        // particles only
                        f[q * n + i] = 0;
                        fo[q * n + i] = 0;
                    }
                keys[v] = k;
            }
            // the slab's records leave through LDS: each store writes one contiguous 16-byte
            // piece per thread, where each thread's own record would take pieces a record apart
            __syncthreads();
            {
                constexpr int RW = (int)(sizeof(CgRec<R>) / 16);
                const int64_t s0 = c0 + v * CG_T, nv = n - s0 < CG_T ? n - s0 : CG_T;
                uint4 *dst = reinterpret_cast<uint4 *>(rec + s0);
                const uint4 *src = reinterpret_cast<const uint4 *>(hbuf);
                for (int e = threadIdx.x; e < RW * CG_T; e += CG_T)
                    if (e < RW * nv) dst[e] = src[e];
            }
            __syncthreads();
        }
        __syncthreads();
        for (int q = threadIdx.x; q < CG_KH; q += CG_T) {
            hkey[q] = -1;
            hcnt[q] = 0;
        }
        __syncthreads();
        for (int v = 0; v < CG_KPT; ++v) {
            const int k = keys[v];
            if (k < 0) continue;
            unsigned h = ((unsigned)k * 2654435761u) >> (32 - CG_KH_BITS);
            for (int probe = 0; probe < CG_KH; ++probe) {
                const int prev = atomicCAS(&hkey[h], -1, k);
                if (prev == -1 || prev == k) break;
                h = (h + 1) & (CG_KH - 1);
            }
            ent[v] = (int)h;
            loc[v] = atomicAdd(&hcnt[h], 1);
        }
        __syncthreads();
        {
            // the occupied entries compacted, so that a thread waits for one global atomic a round
            constexpr int PER = CG_KH / CG_T;
            int c = 0;
            for (int j = 0; j < PER; ++j) c += hkey[threadIdx.x * PER + j] >= 0 ? 1 : 0;
            int tot;
            int o = cg_block_scan(c, lds, &tot);
            for (int j = 0; j < PER; ++j)
                if (hkey[threadIdx.x * PER + j] >= 0) occ[o++] = threadIdx.x * PER + j;
            __syncthreads();
            for (int e = threadIdx.x; e < tot; e += CG_T) {
                const int q = occ[e];
                hcnt[q] = atomicAdd(&count[hkey[q]], hcnt[q]);
            }
        }
        __syncthreads();
        for (int v = 0; v < CG_KPT; ++v) {
            const int64_t i = c0 + v * CG_T + threadIdx.x;
            if (keys[v] < 0) continue;
            key[i] = keys[v];
            slot[i] = hcnt[ent[v]] + loc[v];
        }
        __syncthreads();
    }
}

// the position in the particle's precision, wrapped as apply_pbc does
template <class R>
__device__ inline R cg_wrap(R x, double e, int &jump)
{
    const R ed = (R)e;
    jump = (int)floor((double)(x / ed));
    return x - (R)jump * ed;
}


// exclusive scan of nc int32 in tiles of CG_TILE: tile-local scan and tile totals, one block
// for the totals, then each tile's offset (the total goes to *total when given); with a
// mask, the entries whose mask byte is 0 count as 0
__global__ void __launch_bounds__(CG_T) k_cg_scan_tiles(const int32_t *__restrict__ in, int32_t *__restrict__ out,
                                                        int64_t nc, int32_t *__restrict__ tsum, const ColCtl *C,
                                                        int forced, const uint8_t *__restrict__ mask)
{
    __shared__ int lds[CG_T / 64];
    if (cg_off(C, forced)) return;
    const int64_t ntiles = (nc + CG_TILE - 1) / CG_TILE;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t base = t * CG_TILE + (int64_t)threadIdx.x * CG_SCAN_I;
        int v[CG_SCAN_I], s = 0;
        for (int k = 0; k < CG_SCAN_I; ++k) {
            v[k] = base + k < nc && (!mask || mask[base + k]) ? in[base + k] : 0;
            s += v[k];
        }
        int tot;
        int off = cg_block_scan(s, lds, &tot);
        for (int k = 0; k < CG_SCAN_I; ++k) {
            if (base + k < nc) out[base + k] = off;
            off += v[k];
        }
        if (threadIdx.x == 0) tsum[t] = tot;
    }
}

// the sum of the tile totals before tile t (every thread of the workgroup gets it)
__device__ inline int cg_tile_prefix(const int32_t *__restrict__ tsum, int64_t t, int *lds)
{
    int v = 0;
    for (int64_t u = threadIdx.x; u < t; u += blockDim.x) v += tsum[u];
    int tot;
    (void)cg_block_scan(v, lds, &tot);
    return tot;
}

// the end of a scan: each tile's offsets raised by the totals of the tiles before it (a
// workgroup sums those itself, so there is no pass over the totals alone); the grand
// total goes to *total when given
__global__ void __launch_bounds__(CG_T) k_cg_scan_finish(int32_t *__restrict__ out, int64_t nc,
                                                         const int32_t *__restrict__ tsum, const ColCtl *C, int forced,
                                                         int32_t *total)
{
    __shared__ int lds[CG_T / 64];
    if (cg_off(C, forced)) return;
    const int64_t ntiles = (nc + CG_TILE - 1) / CG_TILE;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int pre = cg_tile_prefix(tsum, t, lds);
        const int64_t c1 = (t + 1) * CG_TILE < nc ? (t + 1) * CG_TILE : nc;
        for (int64_t c = t * CG_TILE + threadIdx.x; c < c1; c += CG_T) out[c] += pre;
        if (total && t == ntiles - 1 && threadIdx.x == 0) *total = pre + tsum[t];
    }
}

// tmp[cell_start[key] + slot] = storage index (particle_system.f90:320-340, a slot per particle)
__global__ void __launch_bounds__(CG_T) k_cg_place(int64_t n, const int32_t *__restrict__ key,
                                                   const int32_t *__restrict__ slot, const int32_t *__restrict__ start,
                                                   int32_t *__restrict__ tmp, const ColCtl *C, int forced)
{
    if (cg_off(C, forced)) return;
    for (CgWalk w = cg_walk(n); w.t < w.t1; w.t += w.step) {
        const int64_t i = w.t * CG_T + threadIdx.x;
        if (i < n) tmp[start[key[i]] + slot[i]] = (int32_t)i;
    }
}

// LDS of a gather unit: the storage indices of its slots and its cells' starts and counts;
// a collision sort's gather adds the image of the velocities and species
template <class F, bool COLL>
struct CgGatherLds {
    int32_t lref[CG_GC_TILE];
    int32_t lstart[CG_GC], lcnt[CG_GC];
};
template <class F>
struct CgGatherLds<F, true> {
    int32_t lref[CG_GC_TILE];
    int32_t lstart[CG_GC], lcnt[CG_GC];
    typename F::real lx[CG_GC_TILE], ly[CG_GC_TILE], lz[CG_GC_TILE];
    typename F::small lsp[CG_GC_TILE];
    double red[16];
};

// The stable gather into the other buffer (particle_system.f90:304-355): the particle of
// each slot goes to cell_start + its rank by storage index among the cell's arrivals,
// which is the reference's single-thread order.  A unit is a tile of CG_GC cells: their
// starts and counts and the storage indices of their slots are loaded into LDS with
// coalesced reads, so that each slot makes one random access (its record) and ranks in
// LDS (tiles of more than CG_GC_TILE particles rank from memory).  Reads only the records
// (never the source arrays), so it also writes pos_old = pos into the source buffer's
// positions (after the sort they are the other buffer's).  Only the particles with the MD
// or PAST_MD bit carry forces (compute_force sets MD on every particle it gives a force,
// the roll keeps PAST_MD for one more step, and the force reset clears both over the md
// slots); their forces come from the side records.  The destination's force planes are
// all zero (the key pass of the previous sort zeroed what it staged, the set-up cleared
// them before the first), so the others' forces are not written.  id_to_idx is not kept
// (nothing in the production loop reads it).
// With COLL (the collision sort) the velocities and species go to an LDS image of the tile,
// and the collision follows on it: simple_mpcd_step with the default angle
// (mpcd.f90:71-183; axes threefry((s, cell + 2^40 k)) of stream SRD) and
// compute_temperature (mpcd.f90:342-407) of the collided cells, per unit the sum of the
// cells' kinetic terms and of the cells counted (k_cg_temp_final sums the units' partials,
// one workgroup).  The list build that follows in the sort chain reads positions only.
template <class F, bool COLL>
__global__ void __launch_bounds__(CG_T) k_cg_gather(SolB<F> S, const ColCtl *C, int fp, int forced, int64_t nc,
                                                    const int32_t *__restrict__ start, const int32_t *__restrict__ count,
                                                    const int32_t *__restrict__ tmp,
                                                    const CgRec<typename F::real> *__restrict__ rec,
                                                    const typename F::real *__restrict__ frec, uint64_t seed,
                                                    double *__restrict__ tpart)
{
    using R = typename F::real;
    using Sm = typename F::small;
    __shared__ CgGatherLds<F, COLL> L;
    if (cg_off(C, forced)) return;
    const int b = C->cur, o = 1 - b;
    const int64_t n = S.n;
    const uint64_t cyc = (uint64_t)C->cycles;
    R *dpos = S.pos[o], *dvel = S.vel[o], *dold = S.pos[b];
    R *df = S.force[2 * o + fp], *dfo = S.force[2 * o + 1 - fp];
    int32_t *dim = S.image[o], *did = S.id[o];
    Sm *dsp = S.species[o], *dfl = S.flags[o];
    for (CgWalk w = cg_walk_units((nc + CG_GC - 1) / CG_GC); w.t < w.t1; w.t += w.step) {
        const int64_t c0 = w.t * CG_GC;
        const int ncl = (int)(c0 + CG_GC < nc ? CG_GC : nc - c0);
        __syncthreads();  // the previous unit's LDS reads are done
        for (int t = threadIdx.x; t < ncl; t += CG_T) {
            L.lstart[t] = start[c0 + t];
            L.lcnt[t] = count[c0 + t];
        }
        __syncthreads();
        const int64_t sa = L.lstart[0], sb = (int64_t)L.lstart[ncl - 1] + L.lcnt[ncl - 1];
        const int tcount = (int)(sb - sa);
        const bool fits = tcount <= CG_GC_TILE;
        if (fits) {
            for (int q = threadIdx.x; q < tcount; q += CG_T) L.lref[q] = tmp[sa + q];
            __syncthreads();
        }
        for (int q = threadIdx.x; q < tcount; q += CG_T) {
            const int32_t v = fits ? L.lref[q] : tmp[sa + q];
            const CgRec<R> r = rec[v];
            const int cl = r.key - (int)c0;
            const int s0 = L.lstart[cl], m0 = L.lcnt[cl];
            int rank = 0;
            if (fits) {
                const int e0 = s0 - (int)sa;
                for (int e = e0; e < e0 + m0; ++e) rank += L.lref[e] < v ? 1 : 0;
            } else {
                for (int e = s0; e < s0 + m0; ++e) rank += tmp[e] < v ? 1 : 0;
            }
            const int64_t d = s0 + rank;
            const int fl = r.sf >> 16;
            const bool has_f = (fl & (MD_MASK | PAST_MD_MASK)) != 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                dpos[k * n + d] = r.p[k];
                dold[k * n + d] = r.p[k];
                dim[k * n + d] = r.im[k];
            }
            if (has_f)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    df[k * n + d] = frec[6 * (int64_t)v + k];
                    dfo[k * n + d] = frec[6 * (int64_t)v + 3 + k];
                }
            did[d] = r.id;
            dfl[d] = (Sm)fl;
            if constexpr (COLL) {
                if (fits) {
                    L.lx[d - sa] = r.v[0];
                    L.ly[d - sa] = r.v[1];
                    L.lz[d - sa] = r.v[2];
                    L.lsp[d - sa] = (Sm)(r.sf & 0xffff);
                    continue;
                }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) dvel[k * n + d] = r.v[k];
            dsp[d] = (Sm)(r.sf & 0xffff);
        }
        if constexpr (COLL) {
            // the collision of the tile's cells, one thread per cell (the tile's image, or
            // memory for a tile that does not fit), then the image written back
            __syncthreads();
            double kin = 0, cnt = 0;
            for (int t = threadIdx.x; t < ncl; t += CG_T) {
                const int m = L.lcnt[t];
                const int64_t sc = L.lstart[t];
                if (m > 1) {
                    const int c = (int)(c0 + t);
                    double k;
                    if (fits) {
                        const int e = (int)(sc - sa);
                        srd_cell(L.lx + e, L.ly + e, L.lz + e, L.lsp + e, m, c, (int)nc, nullptr, seed, cyc, 1.0, 0.0);
                        k = temperature_cell(L.lx + e, L.ly + e, L.lz + e, L.lsp + e, m);
                    } else {
                        srd_cell(dvel + sc, dvel + n + sc, dvel + 2 * n + sc, dsp + sc, m, c, (int)nc, nullptr, seed,
                                 cyc, 1.0, 0.0);
                        k = temperature_cell(dvel + sc, dvel + n + sc, dvel + 2 * n + sc, dsp + sc, m);
                    }
                    cnt += 1;
                    kin += k < 0 ? 0 : k;
                }
            }
            __syncthreads();
            if (fits)
                for (int q = threadIdx.x; q < tcount; q += CG_T) {
                    dvel[sa + q] = L.lx[q];
                    dvel[n + sa + q] = L.ly[q];
                    dvel[2 * n + sa + q] = L.lz[q];
                    dsp[sa + q] = L.lsp[q];
                }
            const double bk = block_sum(kin, L.red);
            const double bc = block_sum(cnt, L.red);
            if (threadIdx.x == 0) {
                tpart[2 * w.t] = bk;
                tpart[2 * w.t + 1] = bc;
            }
        }
    }
}

// The sort is done: the other buffer holds the state (the switch() calls of
// particle_system.f90:346-353), with the md step of the last sort.  Then update_list
// (neighbor_list.f90:76-157) without lj begins: is_md cleared and is_reac set; the lists
// follow (k_cg_nl_build).
__global__ void __launch_bounds__(CG_T) k_cg_list_init(int64_t nc, uint8_t *is_md, uint8_t *is_reac, ColCtl *C,
                                                       int forced)
{
    if (cg_off(C, forced)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        C->cur = 1 - C->cur;
        if (forced) {
            C->last_sort = C->md_steps;
        } else {
            C->last_sort = C->md_steps + 1;
            C->n_extra += 1;
        }
    }
    for (int64_t c = cg_tid(); c < nc; c += cg_stride()) {
        is_md[c] = 0;
        is_reac[c] = 1;
    }
}

// The count and fill of update_list in one pass, a workgroup per colloid: the particle
// ranges of the stencil cells (CG_NL_T cells at a time) are laid end to end by a scan of
// their counts in LDS, the threads walk that sequence (neighbouring threads read
// neighbouring particles of a cell), and an in-order compaction of the accepted particles
// (wave ballots, then the waves' totals) writes the list in the reference's order: stencil
// order, then the cell's sorted order.  A list longer than nmax keeps its first nmax
// entries and raises the overflow (the reference's error stop).
template <class F>
__global__ void __launch_bounds__(CG_NL_T) k_cg_nl_build(SolB<F> S, ColB col, ColCtl *C, int forced, Box bx,
                                                         const int32_t *__restrict__ start,
                                                         const int32_t *__restrict__ count,
                                                         const int32_t *__restrict__ st, int ns, double radius_sq,
                                                         uint8_t *__restrict__ is_md, int32_t nmax,
                                                         int32_t *__restrict__ list, int32_t *__restrict__ nlist)
{
    constexpr int NW = CG_NL_T / 64;
    __shared__ int s_start[CG_NL_T], s_off[CG_NL_T], s_key[CG_NL_T];
    __shared__ int lds[NW], wsum[NW];
    if (cg_off(C, forced)) return;
    const int i = blockIdx.x;
    Box b = bx;
    for (int k = 0; k < 3; ++k) b.origin[k] = C->origin[k];
    const int bf = C->cur;
    const int64_t n = S.n;
    const typename F::real *pos = S.pos[bf];
    const typename F::small *sp = S.species[bf];
    const double x[3] = {col.pos[i], col.pos[col.n + i], col.pos[2 * (int64_t)col.n + i]};
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int total = 0;
    for (int j0 = 0; j0 < ns; j0 += CG_NL_T) {
        const int j = j0 + threadIdx.x;
        int key = 0, s0 = 0, m = 0;
        if (j < ns) {
            key = stencil_key_dev(b, x, st, j);
            s0 = start[key];
            m = count[key];
        }
        int tot;
        const int off = cg_block_scan(m, lds, &tot);
        s_start[threadIdx.x] = s0;
        s_off[threadIdx.x] = off;
        s_key[threadIdx.x] = key;
        __syncthreads();
        for (int p0 = 0; p0 < tot; p0 += CG_NL_T) {
            const int p = p0 + threadIdx.x;
            bool acc = false;
            int q = 0, c = 0;
            if (p < tot) {
                // the cell holding p: the last one whose offset is at most p (the empty cells
                // before it share its offset; the cells past ns have offset tot)
                int lo = 0, hi = CG_NL_T - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= p)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                c = lo;
                q = s_start[c] + (p - s_off[c]);
                if ((int)sp[q] > 0) {
                    const double d0 = min_image(x[0] - (double)pos[q], b.edges[0]);
                    const double d1 = min_image(x[1] - (double)pos[n + q], b.edges[1]);
                    const double d2 = min_image(x[2] - (double)pos[2 * n + q], b.edges[2]);
                    acc = (d0 * d0 + d1 * d1) + d2 * d2 < radius_sq;
                }
            }
            const unsigned long long bal = __ballot(acc);
            if (lane == 0) wsum[w] = __popcll(bal);
            __syncthreads();
            int woff = 0, btot = 0;
            for (int u = 0; u < NW; ++u) {
                const int t = wsum[u];
                woff += u < w ? t : 0;
                btot += t;
            }
            if (acc) {
                const int e = total + woff + __popcll(bal & ((1ull << lane) - 1ull));
                if (e < nmax) list[(int64_t)i * nmax + e] = q;
                is_md[s_key[c]] = 1;
            }
            total += btot;
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        nlist[i] = total > nmax ? nmax : total;
        if (total > nmax) atomicOr(&C->error, RMPCD_ERR_OVERFLOW);
    }
}

// The md slot list: the slots of the is_md cells, in cell order, and a flag per slot.  The
// end of the scan of the md cells' counts (mdoff holds the tile-local offsets from
// k_cg_scan_tiles), as in k_cg_scan_finish, with the list written from the offsets; the
// number of md slots goes to n_md.
__global__ void __launch_bounds__(CG_T) k_cg_md_finish(int64_t nc, const uint8_t *__restrict__ is_md,
                                                       const int32_t *__restrict__ start,
                                                       const int32_t *__restrict__ count,
                                                       const int32_t *__restrict__ mdoff,
                                                       const int32_t *__restrict__ tsum, int32_t *__restrict__ mdslot,
                                                       uint8_t *__restrict__ mdflag, ColCtl *C, int forced)
{
    __shared__ int lds[CG_T / 64];
    if (cg_off(C, forced)) return;
    const int64_t ntiles = (nc + CG_TILE - 1) / CG_TILE;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int pre = cg_tile_prefix(tsum, t, lds);
        const int64_t c1 = (t + 1) * CG_TILE < nc ? (t + 1) * CG_TILE : nc;
        for (int64_t c = t * CG_TILE + threadIdx.x; c < c1; c += CG_T) {
            const int s0 = start[c], m = count[c];
            const bool md = is_md[c] != 0;
            const int o = mdoff[c] + pre;
            for (int k = 0; k < m; ++k) {
                mdflag[s0 + k] = md ? 1 : 0;
                if (md) mdslot[o + k] = s0 + k;
            }
        }
        if (t == ntiles - 1 && threadIdx.x == 0) C->n_md = pre + tsum[t];
    }
}

// ---- forces ---------------------------------------------------------------------------
// compute_force (neighbor_list.f90:191-222): one thread per list entry, solvent forces by
// atomics, the colloid's force and energy as block sums (the partition of the library's
// compute_force, so the same sums).  The grid spans the list capacity; the blocks past a
// colloid's list return at once and k_cg_colloid_forces stops before them (their sums would
// be +0, which leaves a sum that starts at +0 unchanged).  One more workgroup per colloid
// makes its compute_force_n2 sums at the same time.
template <class F>
__global__ void __launch_bounds__(CG_T) k_cg_force_nl(SolB<F> S, ColB col, const ColCtl *C, int fp, int cfp,
                                                      const int32_t *__restrict__ list, const int32_t *__restrict__ nl,
                                                      int32_t nmax, double L0, double L1, double L2, LJ lj,
                                                      double *__restrict__ fpart, LJ lj_cc, double *__restrict__ npart)
{
    using R = typename F::real;
    __shared__ double red[16];
    if (C->error) return;
    const int b = C->cur;
    const int64_t n = S.n;
    const int i = blockIdx.y;
    if (blockIdx.x == gridDim.x - 1) {
        // the last workgroup of a colloid's row: compute_force_n2's sums for it
        // (rmpcd_pairforce.hpp), beside the list blocks
        double g0, g1, g2, ge;
        n2_pair_sums<CG_T>(col.pos, col.species, col.n, i, L0, L1, L2, lj_cc, red, g0, g1, g2, ge);
        if (threadIdx.x == 0) {
            npart[4 * i] = g0;
            npart[4 * i + 1] = g1;
            npart[4 * i + 2] = g2;
            npart[4 * i + 3] = ge;
        }
        return;
    }
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int n_i = min(nl[i], nmax);
    if ((int)(blockIdx.x * blockDim.x) >= n_i) return;
    double f0 = 0, f1 = 0, f2 = 0, e = 0;
    const int64_t cn = col.n;
    const double x0 = col.pos[i], x1 = col.pos[cn + i], x2 = col.pos[2 * cn + i];
    const int s1 = col.species[i];
    if (k < n_i) {
        const int idx = list[(int64_t)i * nmax + k];
        const int s2 = (int)S.species[b][idx];
        if (s2 > 0) {
            const R *pos = S.pos[b];
            const double d0 = min_image(x0 - (double)pos[idx], L0);
            const double d1 = min_image(x1 - (double)pos[n + idx], L1);
            const double d2 = min_image(x2 - (double)pos[2 * n + idx], L2);
            const double r_sq = (d0 * d0 + d1 * d1) + d2 * d2;
            const int pidx = (s2 - 1) + lj.n1 * (s1 - 1);
            if (r_sq <= lj.cut_sq[pidx]) {
                const double fr = lj_force_over_r(r_sq, lj.eps[pidx], lj.sigma[pidx]);
                f0 = fr * d0;
                f1 = fr * d1;
                f2 = fr * d2;
                e = lj_energy(r_sq, lj.eps[pidx], lj.sigma[pidx], lj.cut_energy[pidx]);
                R *f = S.force[2 * b + fp];
                atomicAdd(&f[idx], (R)(-f0));
                atomicAdd(&f[n + idx], (R)(-f1));
                atomicAdd(&f[2 * n + idx], (R)(-f2));
                atomic_or_flag(&S.flags[b][idx], MD_MASK | (s1 == 1 ? CATALYZED_MASK : 0));
            }
        }
    }
    // the four block sums of block_sum (wave sums, then the waves in order) with one barrier
    static_assert(4 * (CG_T / 64) <= 16, "four sums per wave in red");
    const double w4[4] = {wave_sum(f0), wave_sum(f1), wave_sum(f2), wave_sum(e)};
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < 4; ++q) red[4 * wid + q] = w4[q];
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0;
        for (int w = 0; w < CG_T / 64; ++w) t += red[4 * w + threadIdx.x];
        fpart[((int64_t)i * (gridDim.x - 1) + blockIdx.x) * 4 + threadIdx.x] = t;
    }
    (void)cfp;
}

// The colloid forces of one MD step, a thread per colloid: the solvent part, the sum of its
// k_cg_force_nl block sums in block order (over the blocks that cover its list), then the
// compute_force_n2 sums (the last workgroup of its k_cg_force_nl row) added to it.  The
// energies per colloid (solvent, then colloid pairs) go to epart for k_cg_energy.
__global__ void __launch_bounds__(CG_T) k_cg_colloid_forces(ColB col, const ColCtl *C, int cfp,
                                                            const double *__restrict__ fpart, int nb,
                                                            const int32_t *__restrict__ nl, int32_t nmax,
                                                            const double *__restrict__ npart,
                                                            double *__restrict__ epart)
{
    if (C->error) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= col.n) return;
    const int nbi = min(nb, (min(nl[i], nmax) + CG_T - 1) / CG_T);
    double a0 = 0, a1 = 0, a2 = 0, ae = 0;
    for (int q = 0; q < nbi; ++q) {
        const double *pp = fpart + ((int64_t)i * nb + q) * 4;
        a0 += pp[0];
        a1 += pp[1];
        a2 += pp[2];
        ae += pp[3];
    }
    // the force plane holds 0 here; 0 + a is a (a sum that starts at +0 is never -0)
    double *F = col.force[cfp];
    const int64_t cn = col.n;
    F[i] = a0 + npart[4 * i];
    F[cn + i] = a1 + npart[4 * i + 1];
    F[2 * cn + i] = a2 + npart[4 * i + 2];
    epart[i] = ae;
    epart[cn + i] = npart[4 * i + 3];
}

// the energies of the last computed step (tree sums over the colloids): solvent-colloid and
// colloid-colloid.  Run once per cycle, after its last step; it runs after an error stop too,
// when epart still holds the last step that completed.
__global__ void __launch_bounds__(CG_T) k_cg_energy(const ColB col, ColCtl *C, const double *__restrict__ epart)
{
    __shared__ double red[16];
    double s1 = 0, s2 = 0;
    for (int i = threadIdx.x; i < col.n; i += blockDim.x) {
        s1 += epart[i];
        s2 += epart[col.n + i];
    }
    s1 = block_sum(s1, red);
    s2 = block_sum(s2, red);
    if (threadIdx.x == 0) {
        C->e1 = s1;
        C->e2 = s2;
    }
}

// cell_md_vel with md_flag true over the md slots (md.f90:239-263), after the last MD step
// of a cycle (the other steps' velocities are updated in k_cg_md)
template <class F>
__global__ void __launch_bounds__(CG_T) k_cg_vel(SolB<F> S, const ColCtl *C, int fp, double dt,
                                                 const int32_t *__restrict__ mdslot)
{
    using R = typename F::real;
    if (C->error) return;
    const int b = C->cur;
    const int64_t n = S.n;
    R *vel = S.vel[b];
    const R *f = S.force[2 * b + fp], *fo = S.force[2 * b + 1 - fp];
    typename F::small *flags = S.flags[b];
    const R rdt = (R)dt;
    const int nmd = C->n_md;
    for (int64_t k = cg_tid(); k < nmd; k += cg_stride()) {
        const int64_t j = mdslot[k];
        const int fl = (int)flags[j];
        if (!(fl & WALL_MASK)) {
#pragma unroll
            for (int q = 0; q < 3; ++q) vel[q * n + j] = vel[q * n + j] + (rdt * (f[q * n + j] + fo[q * n + j])) / (R)2;
        } else {
            flags[j] = (typename F::small)(fl & ~WALL_MASK);
        }
    }
}

// colloid velocity (n_colloids_pbc.f90:312) and the step counter, after the last MD step of
// a cycle (the other steps': k_cg_md)
__global__ void k_cg_colloid_vel(ColB c, ColCtl *C, int cfp, double dt)
{
    if (C->error) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < c.n) {
        const double *F = c.force[cfp], *Fo = c.force[1 - cfp];
        for (int k = 0; k < 3; ++k) {
            const int64_t q = (int64_t)k * c.n + i;
            c.vel[q] = c.vel[q] + (dt * (F[q] + Fo[q])) / (2 * c.mass[i]);
        }
    }
    if (i == 0) C->md_steps += 1;
}

// ---- collision (in the collision sort's gather) --------------------------------------
__global__ void __launch_bounds__(1024) k_cg_temp_final(ColCtl *C, const double *__restrict__ tpart, int nb)
{
    __shared__ double red[16];
    double a = 0, b = 0;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) {
        a += tpart[2 * q];
        b += tpart[2 * q + 1];
    }
    a = block_sum(a, red);
    b = block_sum(b, red);
    if (threadIdx.x == 0 && !C->error) {
        C->temperature = a / b;
        C->cycles += 1;
    }
}

// ---- host side ------------------------------------------------------------------------
template <class F>
static SolB<F> sol_bufs(const ProgramSim &s)
{
    SolB<F> o;
    o.n = s.n;
    for (int b = 0; b < 2; ++b) {
        o.pos[b] = (typename F::real *)s.sol_pos[b];
        o.vel[b] = (typename F::real *)s.sol_vel[b];
        o.image[b] = s.sol_image[b];
        o.id[b] = s.sol_id[b];
        o.species[b] = (typename F::small *)s.sol_species[b];
        o.flags[b] = (typename F::small *)s.sol_flags[b];
    }
    for (int k = 0; k < 4; ++k) o.force[k] = (typename F::real *)s.sol_force[k];
    return o;
}

static ColB col_bufs(const ProgramSim &s)
{
    ColB c;
    c.n = s.nc;
    c.pos = s.col_pos;
    c.pos_old = s.col_pos_old;
    c.vel = s.col_vel;
    c.force[0] = s.col_force[0];
    c.force[1] = s.col_force[1];
    c.mass = s.col_mass;
    c.image = s.col_image;
    c.species = s.col_species;
    return c;
}

static int cg_alloc(ColloidsGraph &g, void **p, size_t bytes)
{
    bytes = ((bytes ? bytes : 1) + 255) & ~(size_t)255;
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return set_error(RMPCD_ERR_HIP, "colloids graph: hipMalloc of %zu bytes failed", bytes);
    }
    g.allocs.push_back(*p);
    if (hipMemset(*p, 0, bytes) != hipSuccess) return set_error(RMPCD_ERR_HIP, "colloids graph: hipMemset failed");
    return RMPCD_OK;
}

void colloids_graph_free(ColloidsGraph *g)
{
    if (!g) return;
    (void)hipDeviceSynchronize();
    for (int k = 0; k < 2; ++k)
        if (g->exec[k]) (void)hipGraphExecDestroy(g->exec[k]);
    for (void *p : g->allocs) (void)hipFree(p);
    delete g;
}

// the sort chain (ProgramSim::sort, build_list, copy_pos_old): colloid apply_pbc and the
// count reset; the key pass (catch-up, apply_pbc, keys, counts, records); the scan of the
// counts; place; gather (with pos_old); commit and list reset; the lists; the md slots
template <class F>
static int enqueue_sort_chain(ProgramSim &s, ColloidsGraph &g, int fp, int forced, hipStream_t st)
{
    const SolB<F> S = sol_bufs<F>(s);
    const ColB col = col_bufs(s);
    ColCtl *C = g.ctl;
    const Box b = make_box(&s.box);
    const int64_t nc = s.box.ncells, n = s.n;
    const double *E = s.box.edges;
    using R = typename F::real;
    k_cg_sort_begin<<<CG_GRID, CG_T, 0, st>>>(col, C, forced, E[0], E[1], E[2], s.cell_count, nc);
    k_cg_key<F><<<CG_GRID, CG_T, 0, st>>>(S, C, fp, forced, b, g.lut, g.mdflag, s.cell_count, g.key, g.slot,
                                          (CgRec<R> *)g.rec, (R *)g.frec);
    k_cg_scan_tiles<<<(unsigned)std::min<int64_t>(g.ntiles, CG_GRID), CG_T, 0, st>>>(s.cell_count, s.cell_start, nc,
                                                                                      g.tsum, C, forced, nullptr);
    k_cg_scan_finish<<<(unsigned)std::min<int64_t>(g.ntiles, CG_GRID), CG_T, 0, st>>>(s.cell_start, nc, g.tsum, C,
                                                                                       forced, nullptr);
    k_cg_place<<<CG_GRID, CG_T, 0, st>>>(n, g.key, g.slot, s.cell_start, g.tmp, C, forced);
    if (forced)
        k_cg_gather<F, true><<<CG_GRID, CG_T, 0, st>>>(S, C, fp, forced, nc, s.cell_start, s.cell_count, g.tmp,
                                                       (const CgRec<R> *)g.rec, (const R *)g.frec, s.seed, g.tpart);
    else
        k_cg_gather<F, false><<<CG_GRID, CG_T, 0, st>>>(S, C, fp, forced, nc, s.cell_start, s.cell_count, g.tmp,
                                                        (const CgRec<R> *)g.rec, (const R *)g.frec, s.seed, nullptr);
    k_cg_list_init<<<CG_GRID, CG_T, 0, st>>>(nc, s.is_md, s.is_reac, C, forced);
    k_cg_nl_build<F><<<(unsigned)s.nc, CG_NL_T, 0, st>>>(S, col, C, forced, b, s.cell_start, s.cell_count, s.stencil,
                                                         s.n_stencil, s.radius * s.radius, s.is_md, s.nmax, s.list,
                                                         s.nlist);
    k_cg_scan_tiles<<<(unsigned)std::min<int64_t>(g.ntiles, CG_GRID), CG_T, 0, st>>>(s.cell_count, g.mdoff, nc, g.tsum,
                                                                                      C, forced, s.is_md);
    k_cg_md_finish<<<(unsigned)std::min<int64_t>(g.ntiles, CG_GRID), CG_T, 0, st>>>(
        nc, s.is_md, s.cell_start, s.cell_count, g.mdoff, g.tsum, g.mdslot, g.mdflag, C, forced);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

// one MD step (ProgramSim::colloids_step); fp is the force parity before the step.  The
// velocities of a step are updated at the start of the next one (`pending`, the same pass
// over the md slots as the positions) or, after the last step of a cycle, by finish_steps.
template <class F>
static int enqueue_step(ProgramSim &s, ColloidsGraph &g, int fp, bool pending, hipStream_t st)
{
    const SolB<F> S = sol_bufs<F>(s);
    const ColB col = col_bufs(s);
    ColCtl *C = g.ctl;
    const int ncb = (int)blocks_for(s.nc, CG_T);  // the md pass's workgroups that move the colloids
    if (pending)
        k_cg_md<F, true><<<CG_MD_GRID, CG_T, 0, st>>>(S, col, C, fp, s.dt, g.mdslot, g.part, g.cpart);
    else
        k_cg_md<F, false><<<CG_MD_GRID, CG_T, 0, st>>>(S, col, C, fp, s.dt, g.mdslot, g.part, g.cpart);
    k_cg_decide<<<1, CG_DEC_T, 0, st>>>(C, g.part, CG_MD_GRID, g.cpart, ncb, s.dt, s.skin, s.box.max_v);
    int rc = enqueue_sort_chain<F>(s, g, fp, 0, st);
    if (rc) return rc;
    // compute_force(_n2) into the planes zeroed above
    const int np = 1 - fp;
    const double *E = s.box.edges;
    const dim3 fgrid((unsigned)g.fblocks + 1, (unsigned)s.nc);  // list blocks, then the n^2 sums
    k_cg_force_nl<F><<<fgrid, CG_T, 0, st>>>(S, col, C, np, np, s.list, s.nlist, s.nmax, E[0], E[1], E[2],
                                             make_lj(&s.lj_sc), g.fpart, make_lj(&s.lj_cc), g.npart);
    k_cg_colloid_forces<<<blocks_for(s.nc, CG_T), CG_T, 0, st>>>(col, C, np, g.fpart, g.fblocks, s.nlist, s.nmax,
                                                                 g.npart, g.epart);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

// the velocities of the last MD step of a cycle; fp is the force parity after it
template <class F>
static int finish_steps(ProgramSim &s, ColloidsGraph &g, int fp, hipStream_t st)
{
    const SolB<F> S = sol_bufs<F>(s);
    const ColB col = col_bufs(s);
    k_cg_vel<F><<<CG_GRID, CG_T, 0, st>>>(S, g.ctl, fp, s.dt, g.mdslot);
    k_cg_colloid_vel<<<blocks_for(s.nc, CG_T), CG_T, 0, st>>>(col, g.ctl, fp, s.dt);
    k_cg_energy<<<1, CG_T, 0, st>>>(col, g.ctl, g.epart);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

// the collision (ProgramSim::colloids_collide, production loop with SRD)
template <class F>
static int enqueue_collision(ProgramSim &s, ColloidsGraph &g, int fp, hipStream_t st)
{
    ColCtl *C = g.ctl;
    k_cg_shift<<<1, 1, 0, st>>>(C, s.seed, s.dt);
    int rc = enqueue_sort_chain<F>(s, g, fp, 1, st);
    if (rc) return rc;
    // the collision ran in the sort's gather; the temperature of its units' partials
    const int64_t units = (s.box.ncells + CG_GC - 1) / CG_GC;
    k_cg_temp_final<<<1, 1024, 0, st>>>(C, g.tpart, (int)units);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

template <class F>
static int capture_cycle(ProgramSim &s, ColloidsGraph &g, int fp0, hipGraphExec_t *out)
{
    hipStream_t st;
    RMPCD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    RMPCD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = RMPCD_OK, fp = fp0;
    for (int j = 0; j < s.cfg.n_md && !rc; ++j) {
        rc = enqueue_step<F>(s, g, fp, j > 0, st);
        fp = 1 - fp;
    }
    if (!rc && s.cfg.n_md > 0) rc = finish_steps<F>(s, g, fp, st);
    if (!rc) rc = enqueue_collision<F>(s, g, fp, st);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (!rc && e != hipSuccess) rc = set_error(RMPCD_ERR_HIP, "colloids graph: capture failed: %s", hipGetErrorString(e));
    if (!rc) {
        e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = set_error(RMPCD_ERR_HIP, "colloids graph: instantiate failed: %s", hipGetErrorString(e));
    }
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(st);
    return rc;
}

template <class F>
static int graph_setup(ProgramSim &s)
{
    ColloidsGraph *g = new ColloidsGraph();
    s.graph = g;
    const int64_t n = s.n, nc = s.box.ncells;
    int rc;
    g->ntiles = (int)((nc + CG_TILE - 1) / CG_TILE);
    g->fblocks = (int)blocks_for(s.nmax, CG_T);
    RMPCD_CHECK(g->ntiles <= 1 << 20, "colloids graph: too many cells");
    if ((rc = cg_alloc(*g, (void **)&g->ctl, sizeof(ColCtl)))) return rc;
    if ((rc = cg_alloc(*g, (void **)&g->key, 4 * n))) return rc;
    if ((rc = cg_alloc(*g, (void **)&g->slot, 4 * n))) return rc;
    if ((rc = cg_alloc(*g, (void **)&g->tmp, 4 * n))) return rc;
    if ((rc = cg_alloc(*g, (void **)&g->mdslot, 4 * n))) return rc;
    if ((rc = cg_alloc(*g, (void **)&g->mdflag, n))) return rc;
    if ((rc = cg_alloc(*g, (void **)&g->tsum, 4 * (size_t)(g->ntiles + 1)))) return rc;
    if ((rc = cg_alloc(*g, (void **)&g->mdoff, 4 * nc))) return rc;
    if ((rc = cg_alloc(*g, (void **)&g->part, 8 * CG_MD_GRID))) return rc;
    RMPCD_CHECK(blocks_for(s.nc, CG_T) <= CG_MD_GRID, "colloids graph: too many colloids for the md pass");
    if ((rc = cg_alloc(*g, (void **)&g->cpart, 8 * (size_t)blocks_for(s.nc, CG_T)))) return rc;
    if ((rc = cg_alloc(*g, (void **)&g->fpart, 32 * (size_t)g->fblocks * s.nc))) return rc;
    if ((rc = cg_alloc(*g, (void **)&g->epart, 16 * (size_t)s.nc))) return rc;
    if ((rc = cg_alloc(*g, (void **)&g->npart, 32 * (size_t)s.nc))) return rc;
    if ((rc = cg_alloc(*g, (void **)&g->tpart, 16 * (size_t)((nc + CG_GC - 1) / CG_GC)))) return rc;
    {
        using R = typename F::real;
        if ((rc = cg_alloc(*g, &g->rec, sizeof(CgRec<R>) * n))) return rc;
        if ((rc = cg_alloc(*g, &g->frec, 6 * sizeof(R) * n))) return rc;
    }
    const size_t nreal = (size_t)s.box.L[0] * s.box.L[1] * s.box.L[2];
    if ((rc = cg_alloc(*g, (void **)&g->lut, 4 * nreal))) return rc;
    const int32_t *lut = hilbert_lut(make_box(&s.box));
    if (!lut) return set_error(RMPCD_ERR_HIP, "colloids graph: Hilbert table allocation failed");
    RMPCD_HIP(hipMemcpy(g->lut, lut, 4 * nreal, hipMemcpyDeviceToDevice));
    // the host state into the device control, and the md slot list of the current is_md
    ColCtl h{};
    h.cur = s.cur;
    h.md_steps = s.md_steps;
    h.last_sort = s.last_sort;
    h.n_extra = s.n_extra;
    h.cycles = s.cycles;
    for (int k = 0; k < 3; ++k) h.origin[k] = s.box.origin[k];
    h.e1 = s.e1;
    h.e2 = s.e2;
    h.temperature = s.temperature;
    RMPCD_HIP(hipMemcpy(g->ctl, &h, sizeof(h), hipMemcpyHostToDevice));
    hipStream_t st = stream();
    // the buffer the first sort writes starts with zero forces (the gather writes the
    // forces of flagged particles only)
    for (int f = 0; f < 2; ++f)
        RMPCD_HIP(hipMemsetAsync(s.sol_force[2 * (1 - s.cur) + f], 0, 3 * (size_t)n * sizeof(typename F::real), st));
    ColCtl *C = g->ctl;
    k_cg_scan_tiles<<<(unsigned)std::min<int64_t>(g->ntiles, CG_GRID), CG_T, 0, st>>>(s.cell_count, g->mdoff, nc,
                                                                                       g->tsum, C, 1, s.is_md);
    k_cg_md_finish<<<(unsigned)std::min<int64_t>(g->ntiles, CG_GRID), CG_T, 0, st>>>(
        nc, s.is_md, s.cell_start, s.cell_count, g->mdoff, g->tsum, g->mdslot, g->mdflag, C, 1);
    RMPCD_LAUNCHED();
    RMPCD_HIP(hipStreamSynchronize(st));
    return RMPCD_OK;
}

template <class F>
static int graph_run(ProgramSim &s, int n_cycles)
{
    if (!s.graph) {
        int rc = graph_setup<F>(s);
        if (rc) return rc;
    }
    ColloidsGraph &g = *s.graph;
    hipStream_t st = stream();
    for (int cy = 0; cy < n_cycles; ++cy) {
        const int fp = s.fpar;
        if (!g.exec[fp]) {
            int rc = capture_cycle<F>(s, g, fp, &g.exec[fp]);
            if (rc) return rc;
        }
        RMPCD_HIP(hipGraphLaunch(g.exec[fp], st));
        if (s.cfg.n_md % 2) {
            s.fpar = 1 - s.fpar;
            s.cfpar = 1 - s.cfpar;
        }
    }
    ColCtl h;
    RMPCD_HIP(hipMemcpyAsync(&h, g.ctl, sizeof(h), hipMemcpyDeviceToHost, st));
    RMPCD_HIP(hipStreamSynchronize(st));
    s.cur = h.cur;
    s.md_steps = h.md_steps;
    s.last_sort = h.last_sort;
    s.n_extra = h.n_extra;
    s.cycles = h.cycles;
    for (int k = 0; k < 3; ++k) s.box.origin[k] = h.origin[k];
    s.e1 = h.e1;
    s.e2 = h.e2;
    s.temperature = h.temperature;
    if (h.error) return set_error(RMPCD_ERR_OVERFLOW, "maximum of neighbor list reached");
    return RMPCD_OK;
}

int colloids_graph_run(ProgramSim &s, int n_cycles)
{
    RMPCD_CHECK(s.fpar == s.cfpar, "colloids graph: solvent and colloid force parities differ");
    return s.cfg.form == RMPCD_FP64 ? graph_run<F64>(s, n_cycles) : graph_run<F32>(s, n_cycles);
}

}  // namespace rmpcd
