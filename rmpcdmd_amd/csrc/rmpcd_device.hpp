// rmpcd_device.hpp -- device-side building blocks shared by the kernels: particle
// views per numeric form, the compact Hilbert key, minimum-image distance, LJ
// force/energy, flag bits, Threefry-2x64-20, Marsaglia sphere sampling and wave64
// reductions/scans (DPP).  Compiled for gfx950 only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rmpcd_capi.h"

namespace rmpcd {

// ---------------------------------------------------------------------------------
// numeric forms: FP64 (the reference's arithmetic) and FP32-mixed (float solvent
// state, uint8 species/flags; colloid quantities stay double)
struct F64 {
    using real = double;
    using small = int32_t;
    static constexpr int form = RMPCD_FP64;
};
struct F32 {
    using real = float;
    using small = uint8_t;
    static constexpr int form = RMPCD_FP32;
};

template <class F>
struct PView {
    int64_t n;
    typename F::real *pos, *vel, *force, *force_old;
    int32_t *image, *id;
    typename F::small *species, *flags;
};

template <class F>
inline PView<F> view(const rmpcd_particles_t *p)
{
    PView<F> v;
    v.n = p->n;
    v.pos = (typename F::real *)p->pos;
    v.vel = (typename F::real *)p->vel;
    v.force = (typename F::real *)p->force;
    v.force_old = (typename F::real *)p->force_old;
    v.image = p->image;
    v.id = p->id;
    v.species = (typename F::small *)p->species;
    v.flags = (typename F::small *)p->flags;
    return v;
}

// box parameters passed by value to kernels
struct Box {
    int L[3];
    int M[3];
    int ncells;
    int has_walls;
    double a;
    double edges[3];
    double origin[3];
    int bc[3];
};

inline Box make_box(const rmpcd_box_t *b)
{
    Box o;
    for (int k = 0; k < 3; ++k) {
        o.L[k] = b->L[k];
        o.M[k] = b->M[k];
        o.edges[k] = b->edges[k];
        o.origin[k] = b->origin[k];
        o.bc[k] = b->bc[k];
    }
    o.ncells = b->ncells;
    o.has_walls = b->has_walls;
    o.a = b->a;
    return o;
}

constexpr int REAC_MASK = 1 << RMPCD_REAC_BIT;
constexpr int MD_MASK = 1 << RMPCD_MD_BIT;
constexpr int WALL_MASK = 1 << RMPCD_WALL_BIT;
constexpr int PAST_MD_MASK = 1 << RMPCD_PAST_MD_BIT;
constexpr int CATALYZED_MASK = 1 << RMPCD_CATALYZED_BIT;

// ---------------------------------------------------------------------------------
// compact Hilbert index (Hamilton & Rau-Chaplin, compact Hilbert indices, TR CS-2006-07),
// with the reference's conventions: entry state e=0, d=2, bit j of axis j at level i,
// per-level mask from the axes that still have bits (src/hilbert.f90:241-271).
namespace hilbert {
__host__ __device__ inline int rot_r(int x, int d) { return ((x >> d) | (x << (3 - d))) & 7; }
__host__ __device__ inline int rot_l(int x, int d) { return ((x << d) | (x >> (3 - d))) & 7; }
__host__ __device__ inline int gray(int i) { return i ^ (i >> 1); }
__host__ __device__ inline int gray_inv(int g) { return g ^ (g >> 1) ^ (g >> 2); }
__host__ __device__ inline int entry(int w) { return w == 0 ? 0 : gray(2 * ((w - 1) / 2)); }
__host__ __device__ inline int tz3(int x)  // trailing zeros of a nonzero value below 16
{
    return (x & 1) ? 0 : (x & 2) ? 1 : (x & 4) ? 2 : 3;
}
__host__ __device__ inline int dir(int w)  // intracube direction, w in [0,8)
{
    if (w == 0) return 0;
    int g = (w & 1) ? tz3(gray(w) ^ gray(w + 1)) : tz3(gray(w - 1) ^ gray(w));
    return g % 3;
}
}  // namespace hilbert

__host__ __device__ inline int compact_key(int p0, int p1, int p2, int M0, int M1, int M2)
{
    using namespace hilbert;
    int e = 0, d = 2, h = 0;
    int mmax = M0 > M1 ? M0 : M1;
    mmax = mmax > M2 ? mmax : M2;
    for (int i = mmax - 1; i >= 0; --i) {
        // axes that still own a bit at level i, axis j at bit j
        int mu = (M0 > i ? 1 : 0) | (M1 > i ? 2 : 0) | (M2 > i ? 4 : 0);
        int nbits = (mu & 1) + ((mu >> 1) & 1) + ((mu >> 2) & 1);
        mu = rot_r(mu, d + 1);
        int l = ((p0 >> i) & 1) | (((p1 >> i) & 1) << 1) | (((p2 >> i) & 1) << 2);
        l = rot_r(l ^ e, d + 1);
        int w = gray_inv(l);
        // keep the bits of w selected by mu, most significant first
        int r = 0;
        for (int k = 2; k >= 0; --k)
            if ((mu >> k) & 1) r = (r << 1) | ((w >> k) & 1);
        e = e ^ rot_l(entry(w), d + 1);
        d = (d + dir(w) + 1) % 3;
        h = (h << nbits) | r;
    }
    return h;
}

__host__ __device__ inline void compact_point(int h, int M0, int M1, int M2, int p[3])
{
    // inverse of compact_key (src/hilbert.f90:273-307)
    using namespace hilbert;
    int e = 0, d = 2, k = 0;
    int mmax = M0 > M1 ? M0 : M1;
    mmax = mmax > M2 ? mmax : M2;
    int msum = M0 + M1 + M2;
    p[0] = p[1] = p[2] = 0;
    for (int i = mmax - 1; i >= 0; --i) {
        int mu = (M0 > i ? 1 : 0) | (M1 > i ? 2 : 0) | (M2 > i ? 4 : 0);
        int nbits = (mu & 1) + ((mu >> 1) & 1) + ((mu >> 2) & 1);
        mu = rot_r(mu, d + 1);
        int pi = rot_r(e, d + 1) & (~mu & 7);
        int r = (h >> (msum - k - nbits)) & ((1 << nbits) - 1);
        k += nbits;
        // rebuild w: masked bits from r (msb first), the others from the Gray-code
        // constraint with pi
        int w = 0, j = nbits - 1;
        for (int b = 2; b >= 0; --b) {
            int wb1 = (w >> (b + 1)) & 1;
            if ((mu >> b) & 1) {
                w |= ((r >> j) & 1) << b;
                --j;
            } else {
                // Gray-code bit b is fixed by pi; w_b = g_b xor w_(b+1)
                w |= ((((pi >> b) & 1) + wb1) & 1) << b;
            }
        }
        int l = gray(w);
        // inverse transform
        l = rot_r(l ^ rot_r(e, d + 1), 3 - d - 2 + 1);
        p[0] |= (l & 1) << i;
        p[1] |= ((l >> 1) & 1) << i;
        p[2] |= ((l >> 2) & 1) << i;
        e = e ^ rot_l(entry(w), d + 1);
        d = (d + dir(w) + 1) % 3;
    }
}

// ---------------------------------------------------------------------------------
// cell of a position: floor(x/a - origin), wrapping p == L to 0 (z only without walls)
// (src/cell_system.f90:165-175).  Computed in the precision of the position.
template <class R>
__device__ inline void cell_of(const Box &b, R x, R y, R z, int &p0, int &p1, int &p2)
{
    p0 = (int)floor((double)((R)(x / (R)b.a) - (R)b.origin[0]));
    p1 = (int)floor((double)((R)(y / (R)b.a) - (R)b.origin[1]));
    p2 = (int)floor((double)((R)(z / (R)b.a) - (R)b.origin[2]));
    if (p0 == b.L[0]) p0 = 0;
    if (p1 == b.L[1]) p1 = 0;
    if (!b.has_walls && p2 == b.L[2]) p2 = 0;
}

// minimum image with one shift (src/common.f90:162-178)
__host__ __device__ inline double min_image(double r, double L)
{
    if (r < -0.5 * L) return r + L;
    if (r > 0.5 * L) return r - L;
    return r;
}

// ---------------------------------------------------------------------------------
// Lennard-Jones (src/interaction.f90:83-96, 126-136), powers expanded left to right
__host__ __device__ inline double pow6(double x) { return ((((x * x) * x) * x) * x) * x; }
__host__ __device__ inline double pow3(double x) { return (x * x) * x; }
__host__ __device__ inline double lj_force_over_r(double r_sq, double eps, double sigma)
{
    double s6 = pow6(sigma) / pow3(r_sq);
    return (((24.0 * eps) * s6) / r_sq) * (2.0 * s6 - 1.0);
}
__host__ __device__ inline double lj_energy(double r_sq, double eps, double sigma, double ecut)
{
    double s6 = pow6(sigma) / pow3(r_sq);
    return (4.0 * eps) * (s6 * s6 - s6) - ecut;
}

// LJ parameters passed by value to kernels (lj_params_t without the cut radii)
struct LJ {
    int n1;
    double eps[16], sigma[16], cut_sq[16], cut_energy[16];
};
inline LJ make_lj(const rmpcd_lj_t *l)
{
    LJ o;
    o.n1 = l->n1;
    for (int k = 0; k < 16; ++k) {
        o.eps[k] = l->eps[k];
        o.sigma[k] = l->sigma[k];
        o.cut_sq[k] = l->cut_sq[k];
        o.cut_energy[k] = l->cut_energy[k];
    }
    return o;
}

// ---------------------------------------------------------------------------------
// Threefry-2x64-20 (Salmon, Moraes, Dror, Shaw, SC'11; Random123 rotation constants)
__host__ __device__ inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
__host__ __device__ inline void threefry2x64(uint64_t c0, uint64_t c1, uint64_t k0, uint64_t k1,
                                             uint64_t &o0, uint64_t &o1)
{
    const int R[8] = {16, 42, 12, 31, 16, 32, 24, 21};
    uint64_t ks[3] = {k0, k1, 0x1BD11BDAA9FC1A22ull ^ k0 ^ k1};
    uint64_t x0 = c0 + ks[0], x1 = c1 + ks[1];
#pragma unroll
    for (int r = 0; r < 20; ++r) {
        x0 += x1;
        x1 = rotl64(x1, R[r & 7]);
        x1 ^= x0;
        if ((r & 3) == 3) {
            int s = (r + 1) >> 2;
            x0 += ks[s % 3];
            x1 += ks[(s + 1) % 3] + (uint64_t)s;
        }
    }
    o0 = x0;
    o1 = x1;
}
__host__ __device__ inline double u01(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

// random streams (second key word)
enum : uint64_t {
    STREAM_SRD = 1,
    STREAM_REACTION = 2,
    STREAM_CHANGE = 3,
    STREAM_UNIFORM = 4,
    STREAM_SHIFT = 5,
    STREAM_INIT = 6,
    STREAM_PLACE = 7,
    STREAM_THERMO = 8,
    STREAM_WALL = 9,
};

// three standard normals for item at step: Box-Muller on the four uniforms of the
// counters (2 step, item) and (2 step + 1, item) of stream st
__device__ inline void normal3_threefry(uint64_t seed, uint64_t st, uint64_t step, uint64_t item, double out[3])
{
    uint64_t a, b, c, d;
    threefry2x64(2 * step, item, seed, st, a, b);
    threefry2x64(2 * step + 1, item, seed, st, c, d);
    const double u1 = 1.0 - u01(a), u2 = u01(b), u3 = 1.0 - u01(c), u4 = u01(d);
    const double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    out[0] = r1 * cos(2 * M_PI * u2);
    out[1] = r1 * sin(2 * M_PI * u2);
    out[2] = r2 * cos(2 * M_PI * u4);
}

// Marsaglia's method on a stream of uniform pairs (src/mpcd.f90:43-62)
struct Sphere {
    double n[3];
};
__device__ inline Sphere marsaglia_from(double n1, double n2, double s)
{
    Sphere o;
    double alpha = 2.0 * sqrt(1.0 - s);
    o.n[0] = n1 * alpha;
    o.n[1] = n2 * alpha;
    o.n[2] = 1.0 - 2.0 * s;
    return o;
}
__device__ inline Sphere rand_sphere_threefry(uint64_t seed, uint64_t stream, uint64_t step,
                                              uint64_t item)
{
    // pair k of item comes from counter (step, item + 2^40 k); the loop ends with
    // probability 1 - (1 - pi/4)^k after k pairs; 64 pairs bound it
    double n1 = 0, n2 = 0, s = 0;
    for (uint64_t k = 0; k < 64; ++k) {
        uint64_t a, b;
        threefry2x64(step, item + (k << 40), seed, stream, a, b);
        n1 = 2 * u01(a) - 1;
        n2 = 2 * u01(b) - 1;
        s = n1 * n1 + n2 * n2;
        if (s < 1.0) break;
    }
    return marsaglia_from(n1, n2, s);
}

// ---------------------------------------------------------------------------------
// SRD cell arithmetic shared by the collision kernels

// running sum of three components.  FP32: compensated, each addition split exactly into
// its rounded sum and its rounding error by TwoSum (Knuth), the errors gathered in c, so
// that s + c is as accurate as a sum in twice the precision (Ogita, Rump and Oishi's Sum2:
// error below eps |sum| + (n eps)^2 sum |x|).  FP64: plain, in the reference's order.
template <class R>
struct Acc3 {
    R s[3] = {0, 0, 0};
    R c[3] = {0, 0, 0};
    __device__ inline void add(int k, R v)
    {
        if constexpr (sizeof(R) == 4) {
            const R t = s[k] + v;
            const R bp = t - s[k];
            c[k] = c[k] + ((s[k] - (t - bp)) + (v - bp));
            s[k] = t;
        } else {
            s[k] = s[k] + v;
        }
    }
    __device__ inline R sum(int k) const
    {
        if constexpr (sizeof(R) == 4)
            return s[k] + c[k];
        else
            return s[k];
    }
};

// v <- u + omega (v - u) for the m particles of a cell (mpcd.f90:173-175), omega given
// row-major as o[3 r + k].  In FP32, when every particle of the cell entered the mean
// (then exact arithmetic keeps the cell's momentum; old holds the compensated sums of
// the old velocities), the rounding of the m new velocities is compensated: the
// difference between the compensated sums before and after (s + c each, the s parts
// subtracted first) is added, per component, to the particle whose new component is
// smallest in magnitude, so the cell's momentum is kept to the rounding of that one
// addition.  FP64 is left as the reference computes it.
template <class R>
__device__ inline void srd_rotate(R *vx, R *vy, R *vz, int m, const R *o, R u0, R u1, R u2, const Acc3<R> &old,
                                  bool conserve)
{
    Acc3<R> nw;
    int jmin[3] = {0, 0, 0};
    R amin[3] = {INFINITY, INFINITY, INFINITY};
    for (int i = 0; i < m; ++i) {
        const R w0 = vx[i] - u0, w1 = vy[i] - u1, w2 = vz[i] - u2;
        R r0 = o[0] * w0;
        r0 = r0 + o[1] * w1;
        r0 = r0 + o[2] * w2;
        R r1 = o[3] * w0;
        r1 = r1 + o[4] * w1;
        r1 = r1 + o[5] * w2;
        R r2 = o[6] * w0;
        r2 = r2 + o[7] * w1;
        r2 = r2 + o[8] * w2;
        const R n3[3] = {u0 + r0, u1 + r1, u2 + r2};
        vx[i] = n3[0];
        vy[i] = n3[1];
        vz[i] = n3[2];
        if constexpr (sizeof(R) == 4) {
            for (int k = 0; k < 3; ++k) {
                nw.add(k, n3[k]);
                const R a = fabsf(n3[k]);
                if (a < amin[k]) {
                    amin[k] = a;
                    jmin[k] = i;
                }
            }
        }
    }
    if constexpr (sizeof(R) == 4) {
        if (conserve) {
            R *vk[3] = {vx, vy, vz};
            for (int k = 0; k < 3; ++k) {
                const R d = (old.s[k] - nw.s[k]) + (old.c[k] - nw.c[k]);
                vk[k][jmin[k]] = vk[k][jmin[k]] + d;
            }
        }
    }
}

// SRD rotation (mpcd.f90:120-180, default or given alpha); axes injected per cell or
// drawn from Threefry keyed by (seed, step, cell)
// one cell of the SRD collision on arrays vx, vy, vz, sp (the LDS tile or global memory)
template <class R, class S>
__device__ inline void srd_cell(R *vx, R *vy, R *vz, const S *sp, int n, int c, int ncells,
                                const double *axes, uint64_t seed, uint64_t step, double sin_a, double cos_a)
{
    Acc3<R> acc;
    int n_eff = 0;
    for (int i = 0; i < n; ++i) {
        if ((int)sp[i] > 0) {
            acc.add(0, vx[i]);
            acc.add(1, vy[i]);
            acc.add(2, vz[i]);
            ++n_eff;
        }
    }
    if (n_eff <= 1) return;
    const R u0 = acc.sum(0) / (R)n_eff, u1 = acc.sum(1) / (R)n_eff, u2 = acc.sum(2) / (R)n_eff;
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
    const double omc = 1 - cos_a;
    // omega(r,k) as filled column-major by mpcd.f90:163-171
    const R o[9] = {(R)(cos_a + (v0 * v0) * omc), (R)((v1 * v0) * omc + v2 * sin_a), (R)((v2 * v0) * omc - v1 * sin_a),
                    (R)((v0 * v1) * omc - v2 * sin_a), (R)(cos_a + (v1 * v1) * omc), (R)((v2 * v1) * omc + v0 * sin_a),
                    (R)((v0 * v2) * omc + v1 * sin_a), (R)((v1 * v2) * omc - v0 * sin_a), (R)(cos_a + (v2 * v2) * omc)};
    srd_rotate(vx, vy, vz, n, o, u0, u1, u2, acc, n_eff == n);
}

// one cell of compute_temperature (mpcd.f90:342-407) on arrays vx, vy, vz, sp (the LDS tile or global
// memory): its kinetic term, 0 when n_eff <= 1
template <class R, class S>
__device__ inline double temperature_cell(const R *vx, const R *vy, const R *vz, const S *sp, int n)
{
    double u0 = 0, u1 = 0, u2 = 0;
    int n_eff = 0;
    for (int i = 0; i < n; ++i)
        if ((int)sp[i] > 0) {
            u0 = u0 + vx[i];
            u1 = u1 + vy[i];
            u2 = u2 + vz[i];
            ++n_eff;
        }
    if (n_eff <= 1) return -1;  // the cell counts, but adds nothing (mpcd.f90:385)
    u0 = u0 / n;  // divided by n (mpcd.f90:388)
    u1 = u1 / n;
    u2 = u2 / n;
    double k = 0;
    for (int i = 0; i < n; ++i)
        if ((int)sp[i] > 0) {
            const double d0 = vx[i] - u0, d1 = vy[i] - u1, d2 = vz[i] - u2;
            k = k + ((d0 * d0 + d1 * d1) + d2 * d2);
        }
    return k / (double)(3 * (n_eff - 1));
}

// ---------------------------------------------------------------------------------
// wave64 primitives.  DPP row shifts build scans inside rows of 16 lanes; row
// broadcasts 15 and 31 carry row totals across the wave.
__device__ inline int wave_inclusive_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}
__device__ inline int wave_sum_i(int v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block reductions for blockDim.x a multiple of 64, at most 1024; every thread returns
// the total.  scratch: at least 16 doubles of shared memory.
__device__ inline double block_sum(double v, double *scratch)
{
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    double t = 0;
    for (int w = 0; w < nw; ++w) t += scratch[w];
    return t;
}
__device__ inline double block_max(double v, double *scratch)
{
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_max(v);
    __syncthreads();
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    double t = scratch[0];
    for (int w = 1; w < nw; ++w) t = scratch[w] > t ? scratch[w] : t;
    return t;
}

// atomic max on a non-negative double through its bit pattern (monotone for x >= 0)
__device__ inline void atomic_max_nonneg(double *addr, double v)
{
    atomicMax((unsigned long long *)addr, (unsigned long long)__double_as_longlong(v));
}

// OR / AND into a flag of either width (a uint8 flag is updated through its aligned
// 32-bit word; flag arrays are allocated in multiples of 256 bytes)
__device__ inline void atomic_or_flag(int32_t *f, int m) { atomicOr(f, m); }
__device__ inline void atomic_or_flag(uint8_t *f, int m)
{
    uintptr_t a = (uintptr_t)f;
    unsigned int *w = (unsigned int *)(a & ~(uintptr_t)3);
    atomicOr(w, (unsigned int)(m & 0xff) << (8 * (a & 3)));
}
__device__ inline void atomic_and_flag(int32_t *f, int m) { atomicAnd(f, m); }
__device__ inline void atomic_and_flag(uint8_t *f, int m)
{
    uintptr_t a = (uintptr_t)f;
    unsigned int *w = (unsigned int *)(a & ~(uintptr_t)3);
    unsigned int sh = 8 * (unsigned)(a & 3);
    atomicAnd(w, ~((~(unsigned int)m & 0xffu) << sh));
}

// Hilbert key of stencil cell j around the cell of position x
// (neighbor_list.f90:111-121: modulo(cell + stencil - 1, L) with the 1-based cell)
__device__ inline int stencil_key_dev(const Box &b, const double *x, const int32_t *st, int j)
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

}  // namespace rmpcd
