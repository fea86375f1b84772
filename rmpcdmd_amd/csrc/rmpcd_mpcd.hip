// rmpcd_mpcd.hip -- SRD collision and its MPC-AT thermostat branch, the wall collision,
// temperature, streaming (periodic, z walls, forced flow between y/z walls), bulk
// reactions (plain and endothermic) and the cell-wise kinetic-energy rescaling
// (src/mpcd.f90:43-337, 342-407, 440-602, 605-700, 702-745, 748-801).
//
// Cell kernels work on the cell-sorted solvent: a workgroup owns a run of consecutive
// cells, stages their velocities and species from HBM into LDS with coalesced loads,
// each thread then reduces one cell in the reference's summation order (FP32 form: with
// TwoSum compensation), and the rotated velocities go back to HBM coalesced.  Cells whose
// particles do not fit the tile fall back to direct loads.
#include <float.h>

#include <mutex>

#include "rmpcd_internal.hpp"

namespace rmpcd {

constexpr int CELL_THREADS = 128;          // cells per workgroup
constexpr int TILE_CAP = 2048;             // staged particles per workgroup

template <class F>
struct CellTile {
    using R = typename F::real;
    R *vx, *vy, *vz;
    typename F::small *sp;
    int64_t base;  // first staged slot
    int count;     // staged slots (0 when the tile overflowed)
};

// stage the particles of cells [c0, c1) into LDS; returns false (count=0) when the run
// is longer than the tile
template <class F>
__device__ inline void stage_tile(const PView<F> &p, const int32_t *start, const int32_t *count,
                                  int ncells, int c0, int c1, CellTile<F> &t, bool load_vel)
{
    __shared__ int64_t s_range[2];
    if (threadIdx.x == 0) {
        int last = c1 - 1;
        s_range[0] = start[c0];
        s_range[1] = (int64_t)start[last] + count[last];
    }
    __syncthreads();
    int64_t b = s_range[0], e = s_range[1];
    __syncthreads();
    t.base = b;
    t.count = (e - b <= TILE_CAP) ? (int)(e - b) : 0;
    // four slots per thread at a time: their loads are all in flight before the first
    // LDS store waits for one
    const int nt = blockDim.x;
    int q = threadIdx.x;
    for (; q + 3 * nt < t.count; q += 4 * nt) {
        typename F::real x[4], y[4], z[4];
        typename F::small sp[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t g = b + q + u * nt;
            if (load_vel) {
                x[u] = p.vel[g];
                y[u] = p.vel[p.n + g];
                z[u] = p.vel[2 * p.n + g];
            }
            sp[u] = p.species[g];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (load_vel) {
                t.vx[q + u * nt] = x[u];
                t.vy[q + u * nt] = y[u];
                t.vz[q + u * nt] = z[u];
            }
            t.sp[q + u * nt] = sp[u];
        }
    }
    for (; q < t.count; q += nt) {
        if (load_vel) {
            t.vx[q] = p.vel[b + q];
            t.vy[q] = p.vel[p.n + b + q];
            t.vz[q] = p.vel[2 * p.n + b + q];
        }
        t.sp[q] = p.species[b + q];
    }
    __syncthreads();
}

template <class F>
__global__ void __launch_bounds__(CELL_THREADS)
    k_srd(int ncells, const int32_t *__restrict__ start, const int32_t *__restrict__ count,
          PView<F> p, const double *__restrict__ axes, uint64_t seed, uint64_t step,
          double sin_a, double cos_a)
{
    using R = typename F::real;
    __shared__ R lx[TILE_CAP], ly[TILE_CAP], lz[TILE_CAP];
    __shared__ typename F::small lsp[TILE_CAP];
    int c0 = blockIdx.x * CELL_THREADS;
    int c1 = min(c0 + CELL_THREADS, ncells);
    CellTile<F> t{lx, ly, lz, lsp, 0, 0};
    stage_tile<F>(p, start, count, ncells, c0, c1, t, true);
    int c = c0 + threadIdx.x;
    // the LDS tile and global memory get separate calls, so each is addressed directly
    if (c < c1 && count[c] > 1) {
        const int n = count[c];
        const int64_t s = start[c];
        if (t.count > 0) {
            const int o = (int)(s - t.base);
            srd_cell(lx + o, ly + o, lz + o, lsp + o, n, c, ncells, axes, seed, step, sin_a, cos_a);
        } else {
            srd_cell(p.vel + s, p.vel + p.n + s, p.vel + 2 * p.n + s, p.species + s, n, c, ncells, axes, seed, step,
                     sin_a, cos_a);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < t.count; q += blockDim.x) {
        p.vel[t.base + q] = lx[q];
        p.vel[p.n + t.base + q] = ly[q];
        p.vel[2 * p.n + t.base + q] = lz[q];
    }
}

// The MPC-AT branch of simple_mpcd_step (mpcd.f90:144-160): in a cell with more than
// one particle of species > 0, each of them gets sqrt(T) times three normals; with hydro
// the cell's previous mean velocity is restored by adding (mean before - mean after).
// Normals injected per particle slot (3 planes of n) or drawn from Threefry keyed by
// (seed, step, particle id).
constexpr int MPCAT_THREADS = 512;  // a workgroup of k_mpcat: CELL_THREADS cells, the draws over all threads

template <class F>
__global__ void __launch_bounds__(MPCAT_THREADS)
    k_mpcat(int ncells, const int32_t *__restrict__ start, const int32_t *__restrict__ count, PView<F> p,
            const double *__restrict__ normals, uint64_t seed, uint64_t step, double t_factor, int hydro)
{
    using R = typename F::real;
    __shared__ R lx[TILE_CAP], ly[TILE_CAP], lz[TILE_CAP];
    __shared__ typename F::small lsp[TILE_CAP];
    __shared__ uint8_t lcell[TILE_CAP];  // cell (in the workgroup) of each staged slot
    __shared__ R lu[3][CELL_THREADS];    // cell means
    __shared__ int lneff[CELL_THREADS];  // n_eff of a cell that gets new velocities, else 0
    const int tid = threadIdx.x;
    const int c0 = blockIdx.x * CELL_THREADS;
    const int c1 = min(c0 + CELL_THREADS, ncells);
    CellTile<F> t{lx, ly, lz, lsp, 0, 0};
    stage_tile<F>(p, start, count, ncells, c0, c1, t, true);
    const int c = c0 + tid;
    auto draw = [&](int64_t slot, double g[3]) {
        if (normals) {
            for (int k = 0; k < 3; ++k) g[k] = normals[k * p.n + slot];
        } else {
            normal3_threefry(seed, STREAM_THERMO, step, (uint64_t)p.id[slot], g);
        }
    };
    if (t.count > 0) {
        // 1: each cell's mean over species > 0, in slot order (LDS)
        if (tid < CELL_THREADS) {
            lneff[tid] = 0;
            if (c < c1) {
                const int n = count[c];
                const int o = (int)(start[c] - t.base);
                for (int i = 0; i < n; ++i) lcell[o + i] = (uint8_t)tid;
                if (n > 1) {
                    Acc3<R> acc;
                    int n_eff = 0;
                    for (int i = 0; i < n; ++i)
                        if ((int)lsp[o + i] > 0) {
                            acc.add(0, lx[o + i]);
                            acc.add(1, ly[o + i]);
                            acc.add(2, lz[o + i]);
                            ++n_eff;
                        }
                    if (n_eff > 1) {
                        for (int k = 0; k < 3; ++k) lu[k][tid] = acc.sum(k) / (R)n_eff;
                        lneff[tid] = n_eff;
                    }
                }
            }
        }
        __syncthreads();
        // 2: every particle of such a cell with species > 0 draws its three normals (all
        // threads, one slot each); sqrt(T) g replaces its velocity in LDS
        for (int q = tid; q < t.count; q += blockDim.x)
            if (lneff[lcell[q]] > 0 && (int)lsp[q] > 0) {
                double g[3];
                draw(t.base + q, g);
                lx[q] = (R)(g[0] * t_factor);
                ly[q] = (R)(g[1] * t_factor);
                lz[q] = (R)(g[2] * t_factor);
            }
        __syncthreads();
        // 3: with hydro, the cell's mean is put back: v + (mean before - mean of the new)
        if (hydro && tid < CELL_THREADS && c < c1 && lneff[tid] > 0) {
            const int n = count[c];
            const int o = (int)(start[c] - t.base);
            const R n_eff = (R)lneff[tid];
            Acc3<R> vir;
            for (int i = 0; i < n; ++i)
                if ((int)lsp[o + i] > 0) {
                    vir.add(0, lx[o + i]);
                    vir.add(1, ly[o + i]);
                    vir.add(2, lz[o + i]);
                }
            const R d[3] = {lu[0][tid] - vir.sum(0) / n_eff, lu[1][tid] - vir.sum(1) / n_eff,
                            lu[2][tid] - vir.sum(2) / n_eff};
            for (int i = 0; i < n; ++i)
                if ((int)lsp[o + i] > 0) {
                    lx[o + i] = lx[o + i] + d[0];
                    ly[o + i] = ly[o + i] + d[1];
                    lz[o + i] = lz[o + i] + d[2];
                }
        }
        __syncthreads();
        for (int q = tid; q < t.count; q += blockDim.x) {
            p.vel[t.base + q] = lx[q];
            p.vel[p.n + t.base + q] = ly[q];
            p.vel[2 * p.n + t.base + q] = lz[q];
        }
        return;
    }
    // the cells of this workgroup hold more particles than the tile: one thread per cell
    // from global memory, each new velocity drawn again where it is written (so no slot
    // is read after it is written)
    if (tid >= CELL_THREADS || c >= c1) return;
    const int n = count[c];
    if (n <= 1) return;
    const int64_t s = start[c];
    Acc3<R> acc;
    int n_eff = 0;
    for (int i = 0; i < n; ++i)
        if ((int)p.species[s + i] > 0) {
            for (int k = 0; k < 3; ++k) acc.add(k, p.vel[k * p.n + s + i]);
            ++n_eff;
        }
    if (n_eff <= 1) return;
    const R u[3] = {acc.sum(0) / (R)n_eff, acc.sum(1) / (R)n_eff, acc.sum(2) / (R)n_eff};
    R d[3] = {0, 0, 0};
    if (hydro) {
        Acc3<R> vir;
        for (int i = 0; i < n; ++i) {
            if ((int)p.species[s + i] <= 0) continue;
            double g[3];
            draw(s + i, g);
            for (int k = 0; k < 3; ++k) vir.add(k, (R)(g[k] * t_factor));
        }
        for (int k = 0; k < 3; ++k) d[k] = u[k] - vir.sum(k) / (R)n_eff;
    }
    for (int i = 0; i < n; ++i) {
        if ((int)p.species[s + i] <= 0) continue;
        double g[3];
        draw(s + i, g);
        for (int k = 0; k < 3; ++k) {
            const R v = (R)(g[k] * t_factor);
            p.vel[k * p.n + s + i] = hydro ? v + d[k] : v;
        }
    }
}

// rescale_cells (mpcd.f90:748-801): in a cell with more than one particle of species
// > 0, their velocities about the cell mean are scaled so that their kinetic energy is
// 3 (n_eff - 1) T / 2 (sums in FP64 in the reference's order)
// one cell of rescale_cells on arrays vx, vy, vz, sp (the LDS tile or global memory)
template <class R, class S>
__device__ inline void rescale_cell(R *vx, R *vy, R *vz, const S *sp, int n, double T)
{
    double u[3] = {0, 0, 0};
    int n_eff = 0;
    for (int i = 0; i < n; ++i)
        if ((int)sp[i] > 0) {
            u[0] = u[0] + (double)vx[i];
            u[1] = u[1] + (double)vy[i];
            u[2] = u[2] + (double)vz[i];
            ++n_eff;
        }
    if (n_eff <= 1) return;
    for (int k = 0; k < 3; ++k) u[k] = u[k] / n_eff;
    double kin = 0;
    for (int i = 0; i < n; ++i)
        if ((int)sp[i] > 0) {
            const double d0 = vx[i] - u[0], d1 = vy[i] - u[1], d2 = vz[i] - u[2];
            kin = kin + ((d0 * d0 + d1 * d1) + d2 * d2);
        }
    const double f = sqrt((double)(3 * (n_eff - 1)) * T / kin);
    for (int i = 0; i < n; ++i) {
        if ((int)sp[i] <= 0) continue;
        vx[i] = (R)(u[0] + f * ((double)vx[i] - u[0]));
        vy[i] = (R)(u[1] + f * ((double)vy[i] - u[1]));
        vz[i] = (R)(u[2] + f * ((double)vz[i] - u[2]));
    }
}

template <class F>
__global__ void __launch_bounds__(CELL_THREADS)
    k_rescale(int ncells, const int32_t *__restrict__ start, const int32_t *__restrict__ count, PView<F> p,
              double T)
{
    using R = typename F::real;
    __shared__ R lx[TILE_CAP], ly[TILE_CAP], lz[TILE_CAP];
    __shared__ typename F::small lsp[TILE_CAP];
    int c0 = blockIdx.x * CELL_THREADS;
    int c1 = min(c0 + CELL_THREADS, ncells);
    CellTile<F> t{lx, ly, lz, lsp, 0, 0};
    stage_tile<F>(p, start, count, ncells, c0, c1, t, true);
    int c = c0 + threadIdx.x;
    // the LDS tile and global memory get separate calls, so each is addressed directly
    if (c < c1 && count[c] > 1) {
        const int n = count[c];
        const int64_t s = start[c];
        if (t.count > 0) {
            const int o = (int)(s - t.base);
            rescale_cell(lx + o, ly + o, lz + o, lsp + o, n, T);
        } else {
            rescale_cell(p.vel + s, p.vel + p.n + s, p.vel + 2 * p.n + s, p.species + s, n, T);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < t.count; q += blockDim.x) {
        p.vel[t.base + q] = lx[q];
        p.vel[p.n + t.base + q] = ly[q];
        p.vel[2 * p.n + t.base + q] = lz[q];
    }
}

// wall_mpcd_step (mpcd.f90:193-337): the collision of a box with walls at z = 0 and
// z = L.  A wall cell (first or last z layer, when the wall parameters are given) with
// fewer than wall_n particles is completed with wall_n - n virtual particles, whose total
// velocity is sqrt((wall_n - n) T_wall) times three normals; the cell mean counts every
// particle (no species filter).  Then SRD about that mean (axis injected or drawn) or,
// with the thermostat, new velocities sqrt(T) times normals, shifted back to the cell
// mean when keep_cell_v.  Injected draws: axes and vnorm per cell (3 planes of ncells),
// pnorm per particle slot (3 planes of n); otherwise Threefry keyed by the cell (axis:
// SRD stream; virtual normals: WALL stream) or the particle id (THERMO stream).
struct WallParams {
    int present;      // the wall parameters were given
    double temp[2];   // wall temperatures (bottom, top)
    int nw[2];        // particles per wall cell
    int thermostat, keep_cell_v;
    double t_factor;  // sqrt(bulk temperature)
    double sin_a, cos_a;
};

// one cell of wall_mpcd_step on arrays vx, vy, vz (the LDS tile or global memory); s is
// the cell's first slot in the particle arrays (normals and ids are read there).  With
// the thermostat each particle draws once: its new velocity is written, and with
// keep_cell_v the difference of the means is added in a second pass over the written
// values (the same arrays, addressed directly).
template <class R>
__device__ inline void wall_cell(R *vx, R *vy, R *vz, int64_t s, int n, int c, const Box &bx, const WallParams &w,
                                 const double *axes, const double *vnorm, const double *pnorm,
                                 const int32_t *id, int64_t np, uint64_t seed, uint64_t step)
{
    const int ncells = bx.ncells;
    int q[3];
    compact_point(c, bx.M[0], bx.M[1], bx.M[2], q);
    const int wi = !w.present ? -1 : q[2] == 0 ? 0 : q[2] == bx.L[2] - 1 ? 1 : -1;
    int n_virtual = 0;
    double vv[3] = {0, 0, 0};
    if (wi >= 0 && n < w.nw[wi]) {
        n_virtual = w.nw[wi] - n;
        if (vnorm) {
            for (int k = 0; k < 3; ++k) vv[k] = vnorm[k * ncells + c];
        } else {
            normal3_threefry(seed, STREAM_WALL, step, (uint64_t)c, vv);
        }
        const double f = sqrt(n_virtual * w.temp[wi]);
        for (int k = 0; k < 3; ++k) vv[k] = vv[k] * f;
    }
    Acc3<R> acc;
    for (int i = 0; i < n; ++i) {
        acc.add(0, vx[i]);
        acc.add(1, vy[i]);
        acc.add(2, vz[i]);
    }
    R u[3];
    for (int k = 0; k < 3; ++k)
        u[k] = n_virtual > 0 ? (R)(((double)acc.sum(k) + vv[k]) / w.nw[wi]) : acc.sum(k) / (R)n;
    if (w.thermostat) {
        for (int i = 0; i < n; ++i) {
            double g[3];
            if (pnorm) {
                for (int k = 0; k < 3; ++k) g[k] = pnorm[k * np + s + i];
            } else {
                normal3_threefry(seed, STREAM_THERMO, step, (uint64_t)id[s + i], g);
            }
            vx[i] = (R)(g[0] * w.t_factor);
            vy[i] = (R)(g[1] * w.t_factor);
            vz[i] = (R)(g[2] * w.t_factor);
        }
        if (w.keep_cell_v) {
            Acc3<R> sum;
            for (int i = 0; i < n; ++i) {
                sum.add(0, vx[i]);
                sum.add(1, vy[i]);
                sum.add(2, vz[i]);
            }
            const R d[3] = {u[0] - sum.sum(0) / (R)n, u[1] - sum.sum(1) / (R)n, u[2] - sum.sum(2) / (R)n};
            for (int i = 0; i < n; ++i) {
                vx[i] = vx[i] + d[0];
                vy[i] = vy[i] + d[1];
                vz[i] = vz[i] + d[2];
            }
        }
        return;
    }
    double v0, v1, v2;
    if (axes) {
        v0 = axes[c];
        v1 = axes[ncells + c];
        v2 = axes[2 * ncells + c];
    } else {
        Sphere sph = rand_sphere_threefry(seed, STREAM_SRD, step, (uint64_t)c);
        v0 = sph.n[0];
        v1 = sph.n[1];
        v2 = sph.n[2];
    }
    const double sin_a = w.sin_a, cos_a = w.cos_a, omc = 1 - cos_a;
    const R o00 = (R)(cos_a + (v0 * v0) * omc), o10 = (R)((v0 * v1) * omc - v2 * sin_a),
            o20 = (R)((v0 * v2) * omc + v1 * sin_a);
    const R o01 = (R)((v1 * v0) * omc + v2 * sin_a), o11 = (R)(cos_a + (v1 * v1) * omc),
            o21 = (R)((v1 * v2) * omc - v0 * sin_a);
    const R o02 = (R)((v2 * v0) * omc - v1 * sin_a), o12 = (R)((v2 * v1) * omc + v0 * sin_a),
            o22 = (R)(cos_a + (v2 * v2) * omc);
    for (int i = 0; i < n; ++i) {
        const R w0 = vx[i] - u[0], w1 = vy[i] - u[1], w2 = vz[i] - u[2];
        R r0 = o00 * w0;
        r0 = r0 + o01 * w1;
        r0 = r0 + o02 * w2;
        R r1 = o10 * w0;
        r1 = r1 + o11 * w1;
        r1 = r1 + o12 * w2;
        R r2 = o20 * w0;
        r2 = r2 + o21 * w1;
        r2 = r2 + o22 * w2;
        vx[i] = u[0] + r0;
        vy[i] = u[1] + r1;
        vz[i] = u[2] + r2;
    }
}

template <class F>
__global__ void __launch_bounds__(CELL_THREADS)
    k_wall_srd(Box bx, const int32_t *__restrict__ start, const int32_t *__restrict__ count, PView<F> p,
               WallParams w, const double *__restrict__ axes, const double *__restrict__ vnorm,
               const double *__restrict__ pnorm, uint64_t seed, uint64_t step)
{
    using R = typename F::real;
    __shared__ R lx[TILE_CAP], ly[TILE_CAP], lz[TILE_CAP];
    __shared__ typename F::small lsp[TILE_CAP];
    const int ncells = bx.ncells;
    int c0 = blockIdx.x * CELL_THREADS;
    int c1 = min(c0 + CELL_THREADS, ncells);
    CellTile<F> t{lx, ly, lz, lsp, 0, 0};
    stage_tile<F>(p, start, count, ncells, c0, c1, t, true);
    int c = c0 + threadIdx.x;
    // the LDS tile and global memory get separate calls, so each is addressed directly
    if (c < c1 && count[c] > 1) {
        const int n = count[c];
        const int64_t s = start[c];
        if (t.count > 0) {
            const int o = (int)(s - t.base);
            wall_cell(lx + o, ly + o, lz + o, s, n, c, bx, w, axes, vnorm, pnorm, p.id, p.n, seed, step);
        } else {
            wall_cell(p.vel + s, p.vel + p.n + s, p.vel + 2 * p.n + s, s, n, c, bx, w, axes, vnorm, pnorm, p.id, p.n,
                      seed, step);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < t.count; q += blockDim.x) {
        p.vel[t.base + q] = lx[q];
        p.vel[p.n + t.base + q] = ly[q];
        p.vel[2 * p.n + t.base + q] = lz[q];
    }
}

// the optional z profile of compute_temperature (mpcd.f90:354-358, 395-398): a cell's
// kinetic term is added to bin floor((z - xmin) / dx) of data and counted in count, with z
// the height of the cell's centre, origin(3) + (cell(3) + 1/2) a
struct TzProfile {
    double xmin, dx;
    int n;
    double *data;
    int32_t *count;
};

// per workgroup: the sum of the cells' kinetic terms and the number of cells with more
// than one particle (the count goes up before the n_eff test, mpcd.f90:374-386)
template <class F>
__global__ void __launch_bounds__(CELL_THREADS)
    k_temperature(int ncells, const int32_t *__restrict__ start, const int32_t *__restrict__ count,
                  PView<F> p, double *__restrict__ part, Box b, TzProfile tz)
{
    using R = typename F::real;
    __shared__ R lx[TILE_CAP], ly[TILE_CAP], lz[TILE_CAP];
    __shared__ typename F::small lsp[TILE_CAP];
    __shared__ double red[16];
    int c0 = blockIdx.x * CELL_THREADS;
    int c1 = min(c0 + CELL_THREADS, ncells);
    CellTile<F> t{lx, ly, lz, lsp, 0, 0};
    stage_tile<F>(p, start, count, ncells, c0, c1, t, true);
    double kin = 0, cnt = 0;
    int c = c0 + threadIdx.x;
    if (c < c1 && count[c] > 1) {
        const int n = count[c];
        const int64_t s = start[c];
        cnt = 1;
        if (t.count > 0) {
            const int o = (int)(s - t.base);
            kin = temperature_cell(lx + o, ly + o, lz + o, lsp + o, n);
        } else {
            kin = temperature_cell(p.vel + s, p.vel + p.n + s, p.vel + 2 * p.n + s, p.species + s, n);
        }
        if (kin < 0) {
            kin = 0;  // n_eff <= 1
        } else if (tz.data) {
            int q[3];
            compact_point(c, b.M[0], b.M[1], b.M[2], q);
            const double z = b.origin[2] + (q[2] + 0.5) * b.a;
            const double f = floor((z - tz.xmin) / tz.dx);
            if (f >= 0 && f < tz.n) {
                atomicAdd(&tz.data[(int)f], kin);
                atomicAdd(&tz.count[(int)f], 1);
            }
        }
    }
    double bk = block_sum(kin, red);
    double bc = block_sum(cnt, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = bk;
        part[2 * blockIdx.x + 1] = bc;
    }
}

// the workgroups' partial sums, added in a fixed order by one workgroup after the kernel
// boundary (a last-workgroup pattern would need a device-scope fence in every workgroup)
__global__ void __launch_bounds__(CELL_THREADS) k_temperature_final(const double *__restrict__ part, unsigned nb,
                                                                    double *__restrict__ out)
{
    __shared__ double red[16];
    double a = 0, b = 0;
    for (unsigned q = threadIdx.x; q < nb; q += blockDim.x) {
        a += part[2 * q];
        b += part[2 * q + 1];
    }
    a = block_sum(a, red);
    b = block_sum(b, red);
    if (threadIdx.x == 0) out[0] = a / b;
}

// bulk_reaction (mpcd.f90:605-643): one draw per eligible cell, last 'from' converts
template <class F>
__global__ void k_bulk_reaction(int ncells, const int32_t *__restrict__ start,
                                const int32_t *__restrict__ count, const uint8_t *__restrict__ is_reac,
                                PView<F> p, int from, int to, double rate, double tau,
                                const double *__restrict__ draws, uint64_t seed, uint64_t step)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    int n = count[c];
    if (n <= 1 || !is_reac[c]) return;
    int64_t s = start[c];
    double local_rate = 0;
    int64_t pick = -1;
    for (int i = 0; i < n; ++i)
        if ((int)p.species[s + i] == from) {
            local_rate = local_rate + 1;
            pick = s + i;
        }
    local_rate = local_rate * rate;
    double u;
    if (draws) {
        u = draws[c];
    } else {
        uint64_t a, b;
        threefry2x64(step, (uint64_t)c, seed, STREAM_REACTION, a, b);
        u = u01(a);
    }
    if (u < (1 - exp(-local_rate * tau)) && pick >= 0) p.species[pick] = (typename F::small)to;
}

// bulk_reaction_endothermic (mpcd.f90:646-700): in a reactive cell with a 'from'
// particle whose kinetic energy about the cell's centre of mass exceeds delta_u, one draw;
// on success the last 'from' converts and the velocities about the centre of mass are
// scaled by sqrt((kin - delta_u) / kin), which removes delta_u
// one cell of bulk_reaction_endothermic on arrays vx, vy, vz, sp (the LDS tile or global
// memory); returns the converting particle (its index in the cell) or -1
template <class R, class S>
__device__ inline int endothermic_cell(R *vx, R *vy, R *vz, const S *sp, int n, int c, int from, double rate,
                                       double tau, double delta_u, const double *draws, uint64_t seed,
                                       uint64_t step)
{
    double local_rate = 0;
    int pick = -1;
    double com[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        if ((int)sp[i] == from) {
            local_rate = local_rate + 1;
            pick = i;
        }
        com[0] = com[0] + (double)vx[i];
        com[1] = com[1] + (double)vy[i];
        com[2] = com[2] + (double)vz[i];
    }
    for (int k = 0; k < 3; ++k) com[k] = com[k] / n;
    double kin = 0;
    for (int i = 0; i < n; ++i) {
        const double d0 = vx[i] - com[0], d1 = vy[i] - com[1], d2 = vz[i] - com[2];
        kin = kin + ((d0 * d0 + d1 * d1) + d2 * d2);
    }
    kin = kin / 2;
    if (!(local_rate > 0 && kin > delta_u)) return -1;
    local_rate = local_rate * rate;
    double u;
    if (draws) {
        u = draws[c];
    } else {
        uint64_t a, b;
        threefry2x64(step, (uint64_t)c, seed, STREAM_REACTION, a, b);
        u = u01(a);
    }
    if (!(u < (1 - exp(-local_rate * tau)))) return -1;
    const double f = sqrt((kin - delta_u) / kin);
    for (int i = 0; i < n; ++i) {
        vx[i] = (R)(com[0] + f * ((double)vx[i] - com[0]));
        vy[i] = (R)(com[1] + f * ((double)vy[i] - com[1]));
        vz[i] = (R)(com[2] + f * ((double)vz[i] - com[2]));
    }
    return pick;
}

template <class F>
__global__ void __launch_bounds__(CELL_THREADS)
    k_bulk_endothermic(int ncells, const int32_t *__restrict__ start, const int32_t *__restrict__ count,
                       const uint8_t *__restrict__ is_reac, PView<F> p, int from, int to, double rate, double tau,
                       double delta_u, const double *__restrict__ draws, uint64_t seed, uint64_t step)
{
    using R = typename F::real;
    __shared__ R lx[TILE_CAP], ly[TILE_CAP], lz[TILE_CAP];
    __shared__ typename F::small lsp[TILE_CAP];
    int c0 = blockIdx.x * CELL_THREADS;
    int c1 = min(c0 + CELL_THREADS, ncells);
    CellTile<F> t{lx, ly, lz, lsp, 0, 0};
    stage_tile<F>(p, start, count, ncells, c0, c1, t, true);
    int c = c0 + threadIdx.x;
    // the LDS tile and global memory get separate calls, so each is addressed directly
    if (c < c1 && count[c] > 1 && is_reac[c]) {
        const int n = count[c];
        const int64_t s = start[c];
        int pick;
        if (t.count > 0) {
            const int o = (int)(s - t.base);
            pick = endothermic_cell(lx + o, ly + o, lz + o, lsp + o, n, c, from, rate, tau, delta_u, draws, seed,
                                    step);
        } else {
            pick = endothermic_cell(p.vel + s, p.vel + p.n + s, p.vel + 2 * p.n + s, p.species + s, n, c, from, rate,
                                    tau, delta_u, draws, seed, step);
        }
        if (pick >= 0) p.species[s + pick] = (typename F::small)to;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < t.count; q += blockDim.x) {
        p.vel[t.base + q] = lx[q];
        p.vel[p.n + t.base + q] = ly[q];
        p.vel[2 * p.n + t.base + q] = lz[q];
    }
}

// yzwall_collision (mpcd.f90:513-587): the flight of a particle from x0 with velocity v0
// over time t under the acceleration g along x, between the y and z walls, with the
// boundary condition of each axis (bounce-back reverses the velocity, specular the
// normal component, periodic wraps and counts the image, which is added back at the end).
// The alternating y/z wall loop is capped at 64 crossings (the reference's loop only
// ends by exhausting the time; a velocity component of exactly zero at a wall would keep
// it going).
__device__ inline void yzwall_collision(const double x0[3], const double v0[3], double x[3], double v[3],
                                        const double L[3], double t, const int bc[3], double g)
{
    const double gv[3] = {g, 0, 0};
    int im[3] = {0, 0, 0};
    int cd = -1;
    double tcol = DBL_MAX;
    for (int i = 1; i < 3; ++i) {
        const double tt = v0[i] > 0 ? (L[i] - x0[i]) / v0[i] : -x0[i] / v0[i];
        if (tt > 0 && tt < tcol) {
            tcol = tt;
            cd = i;
        }
    }
    if (cd < 0) return;
    double trem = t - tcol;
    for (int k = 0; k < 3; ++k) {
        x[k] = (x0[k] + v0[k] * tcol) + (gv[k] * (tcol * tcol)) / 2;
        v[k] = v0[k] + gv[k] * tcol;
    }
    auto wall = [&](int d) {
        if (bc[d] == RMPCD_BOUNCE_BACK_BC) {
            for (int k = 0; k < 3; ++k) v[k] = -v[k];
        } else if (bc[d] == RMPCD_SPECULAR_BC) {
            v[d] = -v[d];
        } else if (bc[d] == RMPCD_PERIODIC_BC) {
            const int jump = (int)floor(x[d] / L[d]);
            x[d] = x[d] - jump * L[d];
            im[d] = im[d] + jump;
        }
    };
    wall(cd);
    for (int it = 0; it < 64; ++it) {
        cd = cd == 1 ? 2 : 1;
        const double tc = v[cd] > 0 ? (L[cd] - x[cd]) / v[cd] : -x[cd] / v[cd];
        if (tc < 0 || tc > trem) break;
        for (int k = 0; k < 3; ++k) {
            x[k] = (x[k] + v[k] * tc) + (gv[k] * (tc * tc)) / 2;
            v[k] = v[k] + gv[k] * tc;
        }
        trem = trem - tc;
        wall(cd);
    }
    const double tc = trem;
    for (int k = 0; k < 3; ++k) {
        x[k] = ((x[k] + v[k] * tc) + (gv[k] * (tc * tc)) / 2) + im[k] * L[k];
        v[k] = v[k] + gv[k] * tc;
    }
}

// mpcd_stream_xforce_yzwall (mpcd.f90:459-511): x += v dt + g dt^2/2 along x; a particle
// that leaves through a wall (z, and y unless y is periodic) flies through
// yzwall_collision, gets the WALL bit and its new velocity; the others keep theirs (the
// acceleration reaches their velocity elsewhere in the reference's drivers)
template <class F>
__global__ void k_stream_xforce_yzwall(Box b, PView<F> p, double dt, double g)
{
    using R = typename F::real;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    double op[3], ov[3], np[3], nv[3];
    const double gv[3] = {g, 0, 0};
    for (int k = 0; k < 3; ++k) {
        op[k] = (double)p.pos[k * p.n + i];
        ov[k] = (double)p.vel[k * p.n + i];
        np[k] = (op[k] + ov[k] * dt) + (gv[k] * (dt * dt)) / 2;
        nv[k] = ov[k];
    }
    const double *L = b.edges;
    const bool y_out = np[1] < 0 || np[1] > L[1], z_out = np[2] < 0 || np[2] > L[2];
    if ((b.bc[1] == RMPCD_PERIODIC_BC && z_out) || (b.bc[1] != RMPCD_PERIODIC_BC && (y_out || z_out))) {
        yzwall_collision(op, ov, np, nv, L, dt, b.bc, g);
        p.flags[i] = (typename F::small)((int)p.flags[i] | WALL_MASK);
        for (int k = 0; k < 3; ++k) p.vel[k * p.n + i] = (R)nv[k];
    }
    for (int k = 0; k < 3; ++k) p.pos[k * p.n + i] = (R)np[k];
}

// mpcd_stream_periodic (mpcd.f90:440-457)
template <class F>
__global__ void k_stream_periodic(Box b, PView<F> p, double dt)
{
    using R = typename F::real;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        R x = p.pos[k * p.n + i] + p.vel[k * p.n + i] * (R)dt;
        R e = (R)b.edges[k];
        int jump = (int)floor((double)(x / e));
        p.image[k * p.n + i] += jump;
        p.pos[k * p.n + i] = x - (R)jump * e;
    }
}

// mpcd_stream_nogravity_zwall (mpcd.f90:702-745)
template <class F>
__global__ void k_stream_zwall(Box b, PView<F> p, double dt)
{
    using R = typename F::real;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    R op[3], ov[3], np[3], nv[3];
    R rdt = (R)dt, dt2 = (R)(dt * dt);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        op[k] = p.pos[k * p.n + i];
        ov[k] = p.vel[k * p.n + i];
        np[k] = (op[k] + ov[k] * rdt) + (p.force[k * p.n + i] * dt2) / (R)2;
        nv[k] = ov[k];
    }
    R Lz = (R)b.edges[2], t_c = 0;
    bool bounce = false;
    if (np[2] < (R)0) {
        t_c = -op[2] / ov[2];
        bounce = true;
    } else if (np[2] > Lz) {
        t_c = (Lz - op[2]) / ov[2];
        bounce = true;
    }
    if (bounce) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            np[k] = np[k] - (ov[k] * (R)2) * (rdt - t_c);
            nv[k] = -ov[k];
        }
        p.flags[i] = (typename F::small)((int)p.flags[i] | WALL_MASK);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        p.pos[k * p.n + i] = np[k];
        p.vel[k * p.n + i] = nv[k];
    }
}

__global__ void k_rand_sphere(const double *u, int64_t stride, int64_t n_items, double *axes,
                              int32_t *used)
{
    int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n_items) return;
    const double *q = u + it * stride;
    int k = 0;
    double n1 = 0, n2 = 0, s = 2;
    bool ok = false;
    while (k + 2 <= stride) {
        n1 = 2 * q[k] - 1;
        n2 = 2 * q[k + 1] - 1;
        k += 2;
        s = n1 * n1 + n2 * n2;
        if (s < 1.0) {
            ok = true;
            break;
        }
    }
    Sphere sp = marsaglia_from(n1, n2, ok ? s : 0.0);
    axes[it] = sp.n[0];
    axes[n_items + it] = sp.n[1];
    axes[2 * n_items + it] = sp.n[2];
    used[it] = ok ? k : -1;
}

__global__ void k_threefry_uniform(uint64_t seed, uint64_t stream_id, uint64_t step, int64_t n,
                                   double *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t a, b;
    threefry2x64(step, (uint64_t)i, seed, stream_id, a, b);
    out[i] = u01(a);
}

template <class F>
int srd_impl(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *pp,
             const double *axes, uint64_t seed, uint64_t step, double sin_a, double cos_a)
{
    int nc = box->ncells;
    k_srd<F><<<blocks_for(nc, CELL_THREADS), CELL_THREADS, 0, stream()>>>(
        nc, cells->cell_start, cells->cell_count, view<F>(pp), axes, seed, step, sin_a, cos_a);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

template <class F>
int mpcat_impl(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *pp, const double *normals,
               uint64_t seed, uint64_t step, double t_factor, int hydro)
{
    int nc = box->ncells;
    k_mpcat<F><<<blocks_for(nc, CELL_THREADS), MPCAT_THREADS, 0, stream()>>>(
        nc, cells->cell_start, cells->cell_count, view<F>(pp), normals, seed, step, t_factor, hydro);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

template <class F>
int rescale_impl(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *pp, double T)
{
    int nc = box->ncells;
    k_rescale<F><<<blocks_for(nc, CELL_THREADS), CELL_THREADS, 0, stream()>>>(nc, cells->cell_start,
                                                                               cells->cell_count, view<F>(pp), T);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

template <class F>
int bulk_endo_impl(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *pp, int from, int to,
                   double rate, double tau, double delta_u, const double *draws, uint64_t seed, uint64_t step)
{
    int nc = box->ncells;
    k_bulk_endothermic<F><<<blocks_for(nc, CELL_THREADS), CELL_THREADS, 0, stream()>>>(nc, cells->cell_start, cells->cell_count,
                                                                      cells->is_reac, view<F>(pp), from, to, rate,
                                                                      tau, delta_u, draws, seed, step);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

template <class F>
int wall_srd_impl(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *pp, WallParams w,
                  const double *axes, const double *vnorm, const double *pnorm, uint64_t seed, uint64_t step)
{
    int nc = box->ncells;
    k_wall_srd<F><<<blocks_for(nc, CELL_THREADS), CELL_THREADS, 0, stream()>>>(
        make_box(box), cells->cell_start, cells->cell_count, view<F>(pp), w, axes, vnorm, pnorm, seed, step);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

template <class F>
int stream_yzwall_impl(const rmpcd_box_t *box, rmpcd_particles_t *pp, double dt, double g)
{
    PView<F> p = view<F>(pp);
    k_stream_xforce_yzwall<F><<<blocks_for(p.n, 256), 256, 0, stream()>>>(make_box(box), p, dt, g);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

template <class F>
int temperature_impl(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                     const rmpcd_particles_t *pp, double *te, TzProfile tz)
{
    int nc = box->ncells;
    unsigned nb = blocks_for(nc, CELL_THREADS);
    double *part = (double *)workspace(WS_PART, sizeof(double) * 2 * nb);
    double *red = (double *)workspace(WS_RED, 64);
    double *host = (double *)pinned(64);
    if (!part || !red || !host) return set_error(RMPCD_ERR_HIP, "compute_temperature: workspace allocation failed");
    hipStream_t s = stream();
    k_temperature<F><<<nb, CELL_THREADS, 0, s>>>(nc, cells->cell_start, cells->cell_count, view<F>(pp), part,
                                                 make_box(box), tz);
    RMPCD_LAUNCHED();
    k_temperature_final<<<1, CELL_THREADS, 0, s>>>(part, nb, red);
    RMPCD_LAUNCHED();
    RMPCD_HIP(hipMemcpyAsync(host, red, sizeof(double), hipMemcpyDeviceToHost, s));
    RMPCD_HIP(hipStreamSynchronize(s));
    *te = host[0];
    return RMPCD_OK;
}

// ---- compute_temperature without allocation ----------------------------------------
// The workgroups' partial sums live in a fixed device array (one per GPU, held until the
// process ends) instead of the grow-only workspace, and the result goes to a device
// double of the caller, on the library stream, with no synchronisation.  The array's
// device address is looked up once per device (hipGetSymbolAddress) and used wherever
// host code passes it; calls that overlap on one GPU share it and race.
constexpr int NOALLOC_MAX_BLOCKS = 65536;  // cells up to NOALLOC_MAX_BLOCKS * CELL_THREADS = 2^23
static __device__ __align__(256) double scratch[2 * NOALLOC_MAX_BLOCKS];
constexpr int NOALLOC_MAX_DEVICES = 64;
static double *g_scratch_dev[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mutex;

// the current device's address of scratch (set up on the device's first use)
int noalloc_scratch(double **out)
{
    int dev = 0;
    RMPCD_HIP(hipGetDevice(&dev));
    RMPCD_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "compute_temperature_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        RMPCD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = (double *)ptr;
    }
    *out = g_scratch_dev[dev];
    return RMPCD_OK;
}

template <class F>
int temperature_noalloc_impl(const rmpcd_box_t *box, const rmpcd_cells_t *cells, const rmpcd_particles_t *pp,
                             double *te_dev)
{
    const int nc = box->ncells;
    const unsigned nb = blocks_for(nc, CELL_THREADS);
    RMPCD_CHECK(nb <= (unsigned)NOALLOC_MAX_BLOCKS,
                "compute_temperature_noalloc: %d cells, more than the %d of its fixed scratch array", nc,
                NOALLOC_MAX_BLOCKS * CELL_THREADS);
    double *part = nullptr;
    int rc = noalloc_scratch(&part);
    if (rc) return rc;
    hipStream_t s = stream();
    TzProfile tz{0, 0, 0, nullptr, nullptr};
    k_temperature<F><<<nb, CELL_THREADS, 0, s>>>(nc, cells->cell_start, cells->cell_count, view<F>(pp), part,
                                                 make_box(box), tz);
    RMPCD_LAUNCHED();
    k_temperature_final<<<1, CELL_THREADS, 0, s>>>(part, nb, te_dev);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

template <class F>
int stream_periodic_impl(const rmpcd_box_t *box, rmpcd_particles_t *pp, double dt)
{
    PView<F> p = view<F>(pp);
    k_stream_periodic<F><<<blocks_for(p.n, 256), 256, 0, stream()>>>(make_box(box), p, dt);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

template <class F>
int stream_zwall_impl(const rmpcd_box_t *box, rmpcd_particles_t *pp, double dt)
{
    PView<F> p = view<F>(pp);
    k_stream_zwall<F><<<blocks_for(p.n, 256), 256, 0, stream()>>>(make_box(box), p, dt);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

template <class F>
int bulk_impl(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *pp, int from,
              int to, double rate, double tau, const double *draws, uint64_t seed, uint64_t step)
{
    int nc = box->ncells;
    k_bulk_reaction<F><<<blocks_for(nc, 256), 256, 0, stream()>>>(
        nc, cells->cell_start, cells->cell_count, cells->is_reac, view<F>(pp), from, to, rate, tau,
        draws, seed, step);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

}  // namespace rmpcd

using namespace rmpcd;

extern "C" {

int rmpcd_rand_sphere(const double *u, int64_t stride, int64_t n_items, double *axes, int32_t *used)
{
    RMPCD_CHECK(u && axes && used && stride >= 2 && n_items >= 0, "rmpcd_rand_sphere: bad argument");
    if (n_items == 0) return RMPCD_OK;
    k_rand_sphere<<<blocks_for(n_items, 256), 256, 0, stream()>>>(u, stride, n_items, axes, used);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

int rmpcd_simple_mpcd_step(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *p,
                           const double *axes, uint64_t seed, uint64_t step, int32_t alpha_given,
                           double alpha, int32_t thermostat, double T, int32_t hydro)
{
    int rc = check_box(box, "simple_mpcd_step");
    if (!rc) rc = check_particles(p, "simple_mpcd_step", false);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start, "simple_mpcd_step: null cell arrays");
    // mpcd.f90:107-109
    if (!hydro && !thermostat)
        return set_error(RMPCD_ERR_ARG, "requiring .not.do_hydro without thermostat in simple_mpcd_step");
    if (thermostat) {
        // the MPC-AT branch (mpcd.f90:144-160), normals from Threefry
        RMPCD_CHECK(T >= 0, "simple_mpcd_step: negative temperature");
        RMPCD_CHECK(p->id != nullptr, "simple_mpcd_step: the thermostat draws by particle id (null id)");
        return RMPCD_DISPATCH_FORM(p, mpcat_impl, box, cells, p, nullptr, seed, step, sqrt(T), hydro ? 1 : 0);
    }
    double s = 1, c = 0;
    if (alpha_given) {
        s = sin(alpha);
        c = cos(alpha);
    }
    return RMPCD_DISPATCH_FORM(p, srd_impl, box, cells, p, axes, seed, step, s, c);
}

int rmpcd_compute_temperature(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                              const rmpcd_particles_t *p, double *te)
{
    int rc = check_box(box, "compute_temperature");
    if (!rc) rc = check_particles(p, "compute_temperature", false);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start && te,
                "compute_temperature: null argument");
    TzProfile tz{0, 0, 0, nullptr, nullptr};
    return RMPCD_DISPATCH_FORM(p, temperature_impl, box, cells, p, te, tz);
}

int rmpcd_compute_temperature_noalloc(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                                      const rmpcd_particles_t *p, double *te_dev)
{
    int rc = check_box(box, "compute_temperature_noalloc");
    if (!rc) rc = check_particles(p, "compute_temperature_noalloc", false);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start && te_dev,
                "compute_temperature_noalloc: null argument");
    return RMPCD_DISPATCH_FORM(p, temperature_noalloc_impl, box, cells, p, te_dev);
}

int rmpcd_compute_temperature_profile(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                                      const rmpcd_particles_t *p, double *te, double xmin, double dx,
                                      int32_t n_bins, double *tz_data, int32_t *tz_count)
{
    int rc = check_box(box, "compute_temperature");
    if (!rc) rc = check_particles(p, "compute_temperature", false);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start && te,
                "compute_temperature: null argument");
    RMPCD_CHECK(tz_data && tz_count && n_bins > 0, "profile_t: data not allocated");
    RMPCD_CHECK(dx > 0, "negative step in profile_init");
    TzProfile tz{xmin, dx, n_bins, tz_data, tz_count};
    return RMPCD_DISPATCH_FORM(p, temperature_impl, box, cells, p, te, tz);
}

int rmpcd_mpcd_stream_periodic(const rmpcd_box_t *box, rmpcd_particles_t *p, double dt)
{
    int rc = check_box(box, "mpcd_stream_periodic");
    if (!rc) rc = check_particles(p, "mpcd_stream_periodic", false);
    if (rc) return rc;
    RMPCD_CHECK(p->image != nullptr, "mpcd_stream_periodic: null image");
    return RMPCD_DISPATCH_FORM(p, stream_periodic_impl, box, p, dt);
}

int rmpcd_mpcd_stream_nogravity_zwall(const rmpcd_box_t *box, rmpcd_particles_t *p, double dt)
{
    int rc = check_box(box, "mpcd_stream_nogravity_zwall");
    if (!rc) rc = check_particles(p, "mpcd_stream_nogravity_zwall", true);
    if (rc) return rc;
    return RMPCD_DISPATCH_FORM(p, stream_zwall_impl, box, p, dt);
}

int rmpcd_bulk_reaction(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *p,
                        int32_t from, int32_t to, double rate, double tau, const double *draws,
                        uint64_t seed, uint64_t step)
{
    int rc = check_box(box, "bulk_reaction");
    if (!rc) rc = check_particles(p, "bulk_reaction", false);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start && cells->is_reac,
                "bulk_reaction: null cell arrays");
    RMPCD_CHECK(p->form == RMPCD_FP64 || (to >= 0 && to < 256), "bulk_reaction: species out of range");
    return RMPCD_DISPATCH_FORM(p, bulk_impl, box, cells, p, from, to, rate, tau, draws, seed, step);
}

int rmpcd_mpcat_step(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *p,
                     const double *normals, uint64_t seed, uint64_t step, double T, int32_t hydro)
{
    int rc = check_box(box, "mpcat_step");
    if (!rc) rc = check_particles(p, "mpcat_step", false);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start, "mpcat_step: null cell arrays");
    RMPCD_CHECK(T >= 0, "mpcat_step: negative temperature");
    RMPCD_CHECK(normals || p->id, "mpcat_step: the thermostat draws by particle id (null id)");
    return RMPCD_DISPATCH_FORM(p, mpcat_impl, box, cells, p, normals, seed, step, sqrt(T), hydro ? 1 : 0);
}

int rmpcd_rescale_cells(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *p, double T)
{
    int rc = check_box(box, "rescale_cells");
    if (!rc) rc = check_particles(p, "rescale_cells", false);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start, "rescale_cells: null cell arrays");
    return RMPCD_DISPATCH_FORM(p, rescale_impl, box, cells, p, T);
}

int rmpcd_bulk_reaction_endothermic(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *p,
                                    int32_t from, int32_t to, double rate, double tau, double delta_u,
                                    const double *draws, uint64_t seed, uint64_t step)
{
    int rc = check_box(box, "bulk_reaction_endothermic");
    if (!rc) rc = check_particles(p, "bulk_reaction_endothermic", false);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start && cells->is_reac,
                "bulk_reaction_endothermic: null cell arrays");
    RMPCD_CHECK(p->form == RMPCD_FP64 || (to >= 0 && to < 256), "bulk_reaction_endothermic: species out of range");
    return RMPCD_DISPATCH_FORM(p, bulk_endo_impl, box, cells, p, from, to, rate, tau, delta_u, draws, seed, step);
}

int rmpcd_wall_mpcd_step(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *p,
                         int32_t walls_given, const double *wall_temperature, const int32_t *wall_n,
                         int32_t thermostat, double bulk_temperature, int32_t keep_cell_v, int32_t alpha_given,
                         double alpha, const double *axes, const double *vnorm, const double *pnorm, uint64_t seed,
                         uint64_t step)
{
    int rc = check_box(box, "wall_mpcd_step");
    if (!rc) rc = check_particles(p, "wall_mpcd_step", false);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start, "wall_mpcd_step: null cell arrays");
    WallParams w{};
    w.present = walls_given ? 1 : 0;
    if (walls_given) {
        RMPCD_CHECK(wall_temperature && wall_n, "wall parameters must be all present or all absent in wall_mpcd_step");
        for (int k = 0; k < 2; ++k) {
            w.temp[k] = wall_temperature[k];
            w.nw[k] = wall_n[k];
        }
    }
    w.thermostat = thermostat ? 1 : 0;
    // mpcd.f90:247-251: keep_cell_v = .false. without the thermostat is refused
    if (!keep_cell_v && !thermostat)
        return set_error(RMPCD_ERR_ARG, "requiring .not. keep_cell_v and .not. thermostat in wall_mpcd_step");
    w.keep_cell_v = keep_cell_v ? 1 : 0;
    if (thermostat) {
        RMPCD_CHECK(bulk_temperature >= 0, "wall_mpcd_step: negative temperature");
        RMPCD_CHECK(pnorm || p->id, "wall_mpcd_step: the thermostat draws by particle id (null id)");
        w.t_factor = sqrt(bulk_temperature);
    }
    w.sin_a = alpha_given ? sin(alpha) : 1;
    w.cos_a = alpha_given ? cos(alpha) : 0;
    return RMPCD_DISPATCH_FORM(p, wall_srd_impl, box, cells, p, w, axes, vnorm, pnorm, seed, step);
}

int rmpcd_mpcd_stream_xforce_yzwall(const rmpcd_box_t *box, rmpcd_particles_t *p, double dt, double g)
{
    int rc = check_box(box, "mpcd_stream_xforce_yzwall");
    if (!rc) rc = check_particles(p, "mpcd_stream_xforce_yzwall", true);
    if (rc) return rc;
    return RMPCD_DISPATCH_FORM(p, stream_yzwall_impl, box, p, dt, g);
}

int rmpcd_threefry_uniform(uint64_t seed, uint64_t stream_id, uint64_t step, int64_t n, double *out)
{
    RMPCD_CHECK(out && n >= 0, "rmpcd_threefry_uniform: bad argument");
    if (n == 0) return RMPCD_OK;
    k_threefry_uniform<<<blocks_for(n, 256), 256, 0, stream()>>>(seed, stream_id, step, n, out);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

}  // extern "C"
