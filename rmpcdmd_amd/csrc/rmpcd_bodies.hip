// rmpcd_bodies.hip -- colloids held together by links: iterative RATTLE over a link list
// (src/md.f90:321-480), the elastic network (516-544) and the rigid body with
// quaternion velocity Verlet (547-700).
//
// The reference updates the links one after the other, each correction seeing the ones
// before it.  Here the links are grouped into levels on the host: a link's level is one
// more than the highest level of the earlier links that share a particle with it.  Links
// of one level touch disjoint particles, and every earlier link that touches the same
// particle sits in a lower level, so one workgroup updating a level at a time, with a
// barrier between levels, reads and writes every particle in exactly the order of the
// sequential loop: the results are the reference's, bit for bit, with a level's links
// done in parallel.  The body (positions, velocities, inverse masses, the schedule)
// lives in LDS while the sweeps run, or in global memory when it does not fit.
#include <math.h>
#include <string.h>

#include <vector>

#include "rmpcd_internal.hpp"

namespace rmpcd {

constexpr int BT = 256;                // threads of the one workgroup
constexpr size_t BODY_LDS = 64 * 1024; // LDS budget of the staged body
constexpr int RATTLE_MAX = 1000;       // md.f90:341, :432

struct BodyMass {
    double m[16];    // particle_system%mass by species (species s at s - 1)
    double inv[16];  // 1 / mass, as the reference's i_mass
};

__device__ inline double dot3b(const double a[3], const double b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// ---- the link schedule ----------------------------------------------------------------
// Built on the host from the caller's link list; kept on the device while the list is
// unchanged (compared by content on every call).
struct Schedule {
    std::vector<int32_t> links;  // the caller's list, for the comparison
    std::vector<double> dist;
    int64_t n = 0;
    int n_links = 0, n_levels = 0;
    int32_t *d_i1 = nullptr, *d_i2 = nullptr, *d_start = nullptr;  // links by level
    int32_t *d_order = nullptr;  // the caller's index of each scheduled link
    double *d_d = nullptr;
    void *block = nullptr;
};

static Schedule g_sched[4];
static int g_sched_next = 0;

static int get_schedule(const int32_t *links, const double *dist, int n_links, int64_t n, const char *who,
                        const Schedule **out)
{
    for (auto &s : g_sched)
        if (s.block && s.n == n && s.n_links == n_links && n_links > 0 &&
            !memcmp(s.links.data(), links, sizeof(int32_t) * 2 * (size_t)n_links) &&
            !memcmp(s.dist.data(), dist, sizeof(double) * (size_t)n_links)) {
            *out = &s;
            return RMPCD_OK;
        }
    for (int l = 0; l < 2 * n_links; ++l)
        RMPCD_CHECK(links[l] >= 0 && links[l] < n, "%s: link %d names particle %d, outside [0, %lld)", who, l / 2,
                    links[l], (long long)n);
    std::vector<int32_t> last((size_t)n, 0), lvl((size_t)n_links);
    int n_levels = 0;
    for (int l = 0; l < n_links; ++l) {
        const int a = links[2 * l], b = links[2 * l + 1];
        const int v = last[a] > last[b] ? last[a] : last[b];
        lvl[l] = v;
        last[a] = last[b] = v + 1;
        if (v + 1 > n_levels) n_levels = v + 1;
    }
    std::vector<int32_t> start((size_t)n_levels + 1, 0), i1((size_t)n_links), i2((size_t)n_links),
        order((size_t)n_links);
    std::vector<double> dd((size_t)n_links);
    for (int l = 0; l < n_links; ++l) ++start[lvl[l] + 1];
    for (int v = 0; v < n_levels; ++v) start[v + 1] += start[v];
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int l = 0; l < n_links; ++l) {
        const int k = fill[lvl[l]]++;
        i1[k] = links[2 * l];
        i2[k] = links[2 * l + 1];
        dd[k] = dist[l];
        order[k] = l;
    }
    Schedule &s = g_sched[g_sched_next];
    g_sched_next = (g_sched_next + 1) % 4;
    if (s.block) {
        RMPCD_HIP(hipStreamSynchronize(stream()));  // a kernel may still read it
        RMPCD_HIP(hipFree(s.block));
        s.block = nullptr;
    }
    const size_t nl = (size_t)(n_links > 0 ? n_links : 1);
    const size_t bytes = nl * 8 + 3 * nl * 4 + ((size_t)n_levels + 1) * 4 + 64;
    RMPCD_HIP(hipMalloc(&s.block, bytes));
    s.d_d = (double *)s.block;
    s.d_i1 = (int32_t *)(s.d_d + nl);
    s.d_i2 = s.d_i1 + nl;
    s.d_order = s.d_i2 + nl;
    s.d_start = s.d_order + nl;
    if (n_links > 0) {
        RMPCD_HIP(hipMemcpy(s.d_d, dd.data(), sizeof(double) * n_links, hipMemcpyHostToDevice));
        RMPCD_HIP(hipMemcpy(s.d_i1, i1.data(), sizeof(int32_t) * n_links, hipMemcpyHostToDevice));
        RMPCD_HIP(hipMemcpy(s.d_i2, i2.data(), sizeof(int32_t) * n_links, hipMemcpyHostToDevice));
        RMPCD_HIP(hipMemcpy(s.d_order, order.data(), sizeof(int32_t) * n_links, hipMemcpyHostToDevice));
    }
    RMPCD_HIP(hipMemcpy(s.d_start, start.data(), sizeof(int32_t) * (n_levels + 1), hipMemcpyHostToDevice));
    s.links.assign(links, links + 2 * (size_t)n_links);
    s.dist.assign(dist, dist + n_links);
    s.n = n;
    s.n_links = n_links;
    s.n_levels = n_levels;
    *out = &s;
    return RMPCD_OK;
}

struct SchedView {
    const int32_t *i1, *i2, *start, *order;
    const double *d;
    int n_links, n_levels;
};

// bytes of LDS a staged body takes
static size_t body_lds_bytes(int64_t n, int n_links, int n_levels, bool with_pr, bool with_vel)
{
    size_t dbl = (size_t)n * ((with_pr ? 6 : 3) + (with_vel ? 3 : 0) + 1) + n_links;
    size_t ints = 3 * (size_t)n + 2 * (size_t)n_links + n_levels + 1;
    return dbl * 8 + ints * 4 + 64 * 8;  // + the reduction scratch
}

__device__ inline double body_max(double v, double *red)
{
    // every value is >= 0; max is exact in any order
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double r = red[0];
    for (int k = 1; k < BT / 64; ++k) r = fmax(r, red[k]);
    return r;
}

// ---- rattle_body_pos (kPos) / rattle_body_vel ----------------------------------------
// Staging: every particle but the first is unfolded next to particle 0
// (md.f90:343-353, :427-431), and folded back at the end (md.f90:400-402, :474-476).
template <bool kPos, bool kLds>
__global__ void __launch_bounds__(BT) k_rattle_body(PView<F64> p, double *__restrict__ gpr, SchedView sv, BodyMass bm,
                                                    double dt, double e0, double e1, double e2, double precision,
                                                    int32_t *__restrict__ out, double *__restrict__ gminv,
                                                    int32_t *__restrict__ gim)
{
    extern __shared__ double smem[];
    const int64_t n = p.n;
    const int tid = threadIdx.x;
    const double edges[3] = {e0, e1, e2};
    double *pos, *vel, *pr = nullptr, *minv, *d, *red;
    int32_t *im, *i1, *i2, *start;
    int64_t st;
    if (kLds) {
        st = n;
        pos = smem;
        vel = pos + 3 * n;
        double *q = vel + 3 * n;
        if (kPos) {
            pr = q;
            q += 3 * n;
        }
        minv = q;
        d = minv + n;
        red = d + sv.n_links;
        im = (int32_t *)(red + 64);
        i1 = im + 3 * n;
        i2 = i1 + sv.n_links;
        start = i2 + sv.n_links;
        for (int l = tid; l < sv.n_links; l += BT) {
            i1[l] = sv.i1[l];
            i2[l] = sv.i2[l];
            d[l] = sv.d[l];
        }
        for (int l = tid; l <= sv.n_levels; l += BT) start[l] = sv.start[l];
    } else {
        st = n;
        pos = p.pos;
        vel = p.vel;
        pr = gpr;
        minv = gminv;
        d = (double *)sv.d;
        red = smem;
        im = gim;
        i1 = (int32_t *)sv.i1;
        i2 = (int32_t *)sv.i2;
        start = (int32_t *)sv.start;
    }
    for (int64_t i = tid; i < n; i += BT) {
        minv[i] = bm.inv[((int)p.species[i] - 1) & 15];
        for (int k = 0; k < 3; ++k) {
            const double x0 = p.pos[k * n], x = p.pos[k * n + i];
            const int j = i == 0 ? 0 : (int)floor(((x - x0) / edges[k]) + 0.5);
            im[k * st + i] = j;
            pos[k * st + i] = i == 0 ? x : x - j * edges[k];
            if (kLds) vel[k * st + i] = p.vel[k * n + i];
            if (kPos) {
                const double r0 = gpr[k * n], r = gpr[k * n + i];
                const int jr = i == 0 ? 0 : (int)floor(((r - r0) / edges[k]) + 0.5);
                pr[k * st + i] = i == 0 ? r : r - jr * edges[k];
            }
        }
    }
    __syncthreads();
    int it;
    for (it = 1; it <= RATTLE_MAX; ++it) {
        for (int v = 0; v < sv.n_levels; ++v) {
            for (int l = start[v] + tid; l < start[v + 1]; l += BT) {
                const int a = i1[l], b = i2[l];
                const double ma = minv[a], mb = minv[b], dl = d[l];
                double s[3];
                for (int k = 0; k < 3; ++k) s[k] = pos[k * st + a] - pos[k * st + b];
                if (kPos) {
                    double r[3];
                    for (int k = 0; k < 3; ++k) r[k] = pr[k * st + a] - pr[k * st + b];
                    const double rsq = dot3b(r, r), ssq = dot3b(s, s), rs = dot3b(r, s);
                    double g = rs - sqrt(rs * rs - rsq * (ssq - dl * dl));
                    g = g / ((dt * (ma + mb)) * rsq);
                    for (int k = 0; k < 3; ++k) {
                        pos[k * st + a] = pos[k * st + a] - ((g * dt) * r[k]) * ma;
                        pos[k * st + b] = pos[k * st + b] + ((g * dt) * r[k]) * mb;
                        vel[k * st + a] = vel[k * st + a] - (g * r[k]) * ma;
                        vel[k * st + b] = vel[k * st + b] + (g * r[k]) * mb;
                    }
                } else {
                    double dv[3];
                    for (int k = 0; k < 3; ++k) dv[k] = vel[k * st + a] - vel[k * st + b];
                    const double kk = dot3b(dv, s) / ((dl * dl) * (ma + mb));
                    for (int k = 0; k < 3; ++k) {
                        vel[k * st + a] = vel[k * st + a] - (kk * s[k]) * ma;
                        vel[k * st + b] = vel[k * st + b] + (kk * s[k]) * mb;
                    }
                }
            }
            __syncthreads();
        }
        double err = 0;
        for (int l = tid; l < sv.n_links; l += BT) {
            const int a = i1[l], b = i2[l];
            double s[3];
            for (int k = 0; k < 3; ++k) s[k] = pos[k * st + a] - pos[k * st + b];
            double e;
            if (kPos) {
                e = fabs(sqrt(dot3b(s, s)) - d[l]);
            } else {
                double dv[3];
                for (int k = 0; k < 3; ++k) dv[k] = vel[k * st + a] - vel[k * st + b];
                e = fabs(dot3b(dv, s) / ((d[l] * d[l]) * (minv[a] + minv[b])));
            }
            if (e > err) err = e;
        }
        err = body_max(err, red);
        if (err < precision) break;
    }
    __syncthreads();
    for (int64_t i = tid; i < n; i += BT)
        for (int k = 0; k < 3; ++k) {
            p.pos[k * n + i] = i == 0 ? pos[k * st] : pos[k * st + i] + im[k * st + i] * edges[k];
            if (kLds) p.vel[k * n + i] = vel[k * st + i];
            if (kPos && kLds) gpr[k * n + i] = pr[k * st + i];
        }
    if (tid == 0) out[0] = it;
}

// elastic_network (md.f90:516-544): harmonic links, each particle's force receiving the
// link terms in link order (the level schedule), the energy summed in link order
__global__ void __launch_bounds__(BT) k_elastic_network(PView<F64> p, SchedView sv, double kk, double e0, double e1,
                                                        double e2, double *__restrict__ terms,
                                                        double *__restrict__ out)
{
    const int64_t n = p.n;
    const int tid = threadIdx.x;
    const double L[3] = {e0, e1, e2};
    for (int v = 0; v < sv.n_levels; ++v) {
        for (int l = sv.start[v] + tid; l < sv.start[v + 1]; l += BT) {
            const int a = sv.i1[l], b = sv.i2[l];
            const double r0 = sv.d[l];
            double x[3];
            for (int k = 0; k < 3; ++k) x[k] = min_image(p.pos[k * n + a] - p.pos[k * n + b], L[k]);
            const double r = sqrt(dot3b(x, x));
            terms[sv.order[l]] = (kk * ((r - r0) * (r - r0))) / 2;
            for (int k = 0; k < 3; ++k) {
                const double f = ((-kk) * (r - r0)) * x[k] / r;
                p.force[k * n + a] = p.force[k * n + a] + f;
                p.force[k * n + b] = p.force[k * n + b] - f;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        double e = 0;
#pragma unroll 8
        for (int l = 0; l < sv.n_links; ++l) e = e + terms[l];
        out[0] = e;
    }
}

// ---- the rigid body (md.f90:547-700) -------------------------------------------------
// Quaternions scalar first, q = (s, v).  The reference takes them from its quaternion
// module (an absent submodule); these are the textbook definitions: Hamilton product,
// conjugate, rotation v -> q (0, v) q*.
__device__ inline void qmul(const double a[4], const double b[4], double r[4])
{
    const double s = a[0] * b[0] - ((a[1] * b[1] + a[2] * b[2]) + a[3] * b[3]);
    const double x = (a[0] * b[1] + b[0] * a[1]) + (a[2] * b[3] - a[3] * b[2]);
    const double y = (a[0] * b[2] + b[0] * a[2]) + (a[3] * b[1] - a[1] * b[3]);
    const double z = (a[0] * b[3] + b[0] * a[3]) + (a[1] * b[2] - a[2] * b[1]);
    r[0] = s;
    r[1] = x;
    r[2] = y;
    r[3] = z;
}
__device__ inline void qconj(const double q[4], double r[4])
{
    r[0] = q[0];
    r[1] = -q[1];
    r[2] = -q[2];
    r[3] = -q[3];
}
__device__ inline void qrot(const double q[4], const double v[3], double r[3])
{
    const double p[4] = {0, v[0], v[1], v[2]};
    double t[4], c[4], u[4];
    qmul(q, p, t);
    qconj(q, c);
    qmul(t, c, u);
    r[0] = u[1];
    r[1] = u[2];
    r[2] = u[3];
}
__device__ inline double qnorm(const double q[4]) { return sqrt(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]); }
__device__ inline void qnormalize(double q[4])
{
    const double s = qnorm(q);
    for (int k = 0; k < 4; ++k) q[k] = q[k] / s;
}
__device__ inline void cross3(const double a[3], const double b[3], double r[3])
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline bool body_range_ok(const rmpcd_rigid_body_t *b, int64_t n)
{
    return b->i_start >= 0 && b->i_stop >= b->i_start && b->i_stop < n;
}

// Sums over the body's particles keep the reference's order: the terms of a chunk of BT
// particles are made in parallel into LDS and added by thread 0 one after the other.
__global__ void __launch_bounds__(BT) k_rigid_init(rmpcd_rigid_body_t *b, PView<F64> p, BodyMass bm, double e0,
                                                   double e1, double e2)
{
    __shared__ double t[7][BT];
    __shared__ double acc[7];
    const int64_t n = p.n;
    const int tid = threadIdx.x;
    const double edges[3] = {e0, e1, e2};
    if (!body_range_ok(b, n)) return;
    const int i0 = b->i_start, i1 = b->i_stop;
    if (tid == 0)
        for (int k = 0; k < 7; ++k) acc[k] = 0;
    // mass, mass * unfolded position, mass * velocity (md.f90:561-568)
    for (int c0 = i0; c0 <= i1; c0 += BT) {
        const int i = c0 + tid;
        if (i <= i1) {
            const double m = bm.m[((int)p.species[i] - 1) & 15];
            t[0][tid] = m;
            for (int k = 0; k < 3; ++k) {
                t[1 + k][tid] = m * (p.pos[k * n + i] + p.image[k * n + i] * edges[k]);
                t[4 + k][tid] = m * p.vel[k * n + i];
            }
        }
        __syncthreads();
        if (tid == 0) {
            const int m = i1 - c0 + 1 < BT ? i1 - c0 + 1 : BT;
            for (int j = 0; j < m; ++j)
                for (int k = 0; k < 7; ++k) acc[k] = acc[k] + t[k][j];
        }
        __syncthreads();
    }
    const double mass = acc[0];
    const double com[3] = {acc[1] / mass, acc[2] / mass, acc[3] / mass};
    const double cv[3] = {acc[4] / mass, acc[5] / mass, acc[6] / mass};
    __syncthreads();
    if (tid == 0)
        for (int k = 0; k < 3; ++k) acc[k] = 0;
    // body-frame positions and the diagonal of the inertia tensor (md.f90:574-580)
    for (int c0 = i0; c0 <= i1; c0 += BT) {
        const int i = c0 + tid;
        if (i <= i1) {
            const double m = bm.m[((int)p.species[i] - 1) & 15];
            double r[3];
            for (int k = 0; k < 3; ++k) r[k] = (p.pos[k * n + i] + p.image[k * n + i] * edges[k]) - com[k];
            for (int k = 0; k < 3; ++k) b->pos_body[3 * (i - i0) + k] = r[k];
            const double r2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
            for (int k = 0; k < 3; ++k) t[k][tid] = m * (r2 - r[k] * r[k]);
        }
        __syncthreads();
        if (tid == 0) {
            const int m = i1 - c0 + 1 < BT ? i1 - c0 + 1 : BT;
            for (int j = 0; j < m; ++j)
                for (int k = 0; k < 3; ++k) acc[k] = acc[k] + t[k][j];
        }
        __syncthreads();
    }
    if (tid == 0) {
        b->mass = mass;
        for (int k = 0; k < 3; ++k) {
            b->pos[k] = com[k];
            b->vel[k] = cv[k];
            b->I_body[k] = acc[k];
            b->force[k] = 0;
            b->torque[k] = 0;
            b->L[k] = 0;
            b->omega_body[k] = 0;
        }
        b->q[0] = 1;
        b->q[1] = b->q[2] = b->q[3] = 0;
    }
}

// rigid_body_compute_force_torque (md.f90:591-607)
__global__ void __launch_bounds__(BT) k_rigid_force_torque(rmpcd_rigid_body_t *b, PView<F64> p)
{
    __shared__ double t[6][BT];
    __shared__ double acc[6];
    const int64_t n = p.n;
    const int tid = threadIdx.x;
    if (!body_range_ok(b, n)) return;
    const int i0 = b->i_start, i1 = b->i_stop;
    const double q[4] = {b->q[0], b->q[1], b->q[2], b->q[3]};
    if (tid == 0)
        for (int k = 0; k < 6; ++k) acc[k] = 0;
    for (int c0 = i0; c0 <= i1; c0 += BT) {
        const int i = c0 + tid;
        if (i <= i1) {
            const double f[3] = {p.force[i], p.force[n + i], p.force[2 * n + i]};
            double r[3], c[3];
            qrot(q, b->pos_body + 3 * (i - i0), r);
            cross3(r, f, c);
            for (int k = 0; k < 3; ++k) {
                t[k][tid] = f[k];
                t[3 + k][tid] = c[k];
            }
        }
        __syncthreads();
        if (tid == 0) {
            const int m = i1 - c0 + 1 < BT ? i1 - c0 + 1 : BT;
            for (int j = 0; j < m; ++j)
                for (int k = 0; k < 6; ++k) acc[k] = acc[k] + t[k][j];
        }
        __syncthreads();
    }
    if (tid == 0)
        for (int k = 0; k < 3; ++k) {
            b->force[k] = acc[k];
            b->torque[k] = acc[3 + k];
        }
}

// rigid_body_vv1 (md.f90:613-673): thread 0 advances the orientation (the half-step
// quaternion iterated to the threshold, at most 1000 rounds) and the centre of mass; the
// particles are then placed in parallel
__global__ void __launch_bounds__(BT) k_rigid_vv1(rmpcd_rigid_body_t *b, PView<F64> p, double dt, double threshold,
                                                  double e0, double e1, double e2, int32_t *__restrict__ out)
{
    __shared__ double sq[8];
    const int64_t n = p.n;
    const int tid = threadIdx.x;
    const double edges[3] = {e0, e1, e2};
    if (!body_range_ok(b, n)) {
        if (tid == 0) out[0] = -1;
        return;
    }
    if (tid == 0) {
        double I_inv[3], q[4], qc[4], L[3], Lb[3], tb[3], ob[3], Lbd[3], c[3], qd[4], w[4];
        for (int k = 0; k < 3; ++k) I_inv[k] = 1 / b->I_body[k];
        for (int k = 0; k < 4; ++k) q[k] = b->q[k];
        for (int k = 0; k < 3; ++k) L[k] = b->L[k];
        qconj(q, qc);
        qrot(qc, L, Lb);
        qrot(qc, b->torque, tb);
        for (int k = 0; k < 3; ++k) ob[k] = Lb[k] * I_inv[k];
        cross3(ob, Lb, c);
        for (int k = 0; k < 3; ++k) Lbd[k] = tb[k] - c[k];
        for (int k = 0; k < 3; ++k) Lb[k] = Lb[k] + (dt * Lbd[k]) / 2;
        for (int k = 0; k < 3; ++k) ob[k] = Lb[k] * I_inv[k];
        w[0] = 0;
        for (int k = 0; k < 3; ++k) w[k + 1] = ob[k];
        qmul(q, w, qd);
        for (int k = 0; k < 4; ++k) qd[k] = qd[k] / 2;
        for (int k = 0; k < 4; ++k) q[k] = q[k] + (qd[k] * dt) / 2;
        qnormalize(q);
        for (int k = 0; k < 3; ++k) L[k] = b->L[k] + (b->torque[k] * dt) / 2;
        double q_old[4] = {q[0], q[1], q[2], q[3]};
        int it;
        for (it = 1; it <= RATTLE_MAX; ++it) {
            qconj(q, qc);
            qrot(qc, L, Lb);
            for (int k = 0; k < 3; ++k) w[k + 1] = I_inv[k] * Lb[k];
            qmul(q, w, qd);
            for (int k = 0; k < 4; ++k) qd[k] = qd[k] / 2;
            for (int k = 0; k < 4; ++k) q[k] = b->q[k] + (qd[k] * dt) / 2;
            qnormalize(q);
            double dq[4];
            for (int k = 0; k < 4; ++k) dq[k] = q[k] - q_old[k];
            if (qnorm(dq) < threshold) break;
            for (int k = 0; k < 4; ++k) q_old[k] = q[k];
        }
        double qn[4];
        for (int k = 0; k < 4; ++k) qn[k] = b->q[k] + dt * qd[k];
        qnormalize(qn);
        for (int k = 0; k < 4; ++k) b->q[k] = sq[k] = qn[k];
        for (int k = 0; k < 3; ++k) {
            const double x = (b->pos[k] + b->vel[k] * dt) + (b->force[k] * (dt * dt)) / (2 * b->mass);
            b->pos[k] = sq[4 + k] = x;
        }
        out[0] = it;
    }
    __syncthreads();
    const double q[4] = {sq[0], sq[1], sq[2], sq[3]};
    const double com[3] = {sq[4], sq[5], sq[6]};
    const int i0 = b->i_start, i1 = b->i_stop;
    for (int i = i0 + tid; i <= i1; i += BT) {
        double r[3];
        qrot(q, b->pos_body + 3 * (i - i0), r);
        for (int k = 0; k < 3; ++k) {
            const double old = p.pos[k * n + i];
            const double x = com[k] + r[k];
            const double delta = (x - old) + edges[k] / 2;
            p.pos[k * n + i] = x - floor(delta / edges[k]) * edges[k];
        }
    }
    __syncthreads();
    if (tid == 0)
        for (int k = 0; k < 3; ++k) {
            b->L[k] = b->L[k] + (b->torque[k] * dt) / 2;
            b->vel[k] = b->vel[k] + (b->force[k] * dt) / (2 * b->mass);
        }
}

// rigid_body_vv2 (md.f90:679-700)
__global__ void __launch_bounds__(BT) k_rigid_vv2(rmpcd_rigid_body_t *b, PView<F64> p, double dt)
{
    __shared__ double s[10];
    const int64_t n = p.n;
    const int tid = threadIdx.x;
    if (!body_range_ok(b, n)) return;
    if (tid == 0) {
        double qc[4], Lb[3], om[3];
        for (int k = 0; k < 3; ++k) b->L[k] = b->L[k] + (b->torque[k] * dt) / 2;
        qconj(b->q, qc);
        qrot(qc, b->L, Lb);
        for (int k = 0; k < 3; ++k) b->omega_body[k] = Lb[k] / b->I_body[k];
        qrot(b->q, b->omega_body, om);
        for (int k = 0; k < 3; ++k) b->vel[k] = b->vel[k] + (b->force[k] * dt) / (2 * b->mass);
        for (int k = 0; k < 4; ++k) s[k] = b->q[k];
        for (int k = 0; k < 3; ++k) {
            s[4 + k] = om[k];
            s[7 + k] = b->vel[k];
        }
    }
    __syncthreads();
    const double q[4] = {s[0], s[1], s[2], s[3]};
    const double om[3] = {s[4], s[5], s[6]};
    const int i0 = b->i_start, i1 = b->i_stop;
    for (int i = i0 + tid; i <= i1; i += BT) {
        double r[3], c[3];
        qrot(q, b->pos_body + 3 * (i - i0), r);
        cross3(om, r, c);
        for (int k = 0; k < 3; ++k) p.vel[k * n + i] = s[7 + k] + c[k];
    }
}

static int body_mass(const double *mass, int32_t n_mass, BodyMass &m)
{
    if (!mass || n_mass < 1 || n_mass > 16) return set_error(RMPCD_ERR_ARG, "mass table must hold 1..16 species");
    for (int k = 0; k < 16; ++k) {
        m.m[k] = k < n_mass ? mass[k] : 0.0;
        m.inv[k] = k < n_mass ? 1 / mass[k] : 0.0;
    }
    return RMPCD_OK;
}

static int check_body(const rmpcd_particles_t *p, const int32_t *links, const double *dist, int32_t n_links,
                      const char *who, bool need_force)
{
    int rc = check_particles(p, who, need_force);
    if (rc) return rc;
    RMPCD_CHECK(p->form == RMPCD_FP64, "%s: colloids must use the FP64 form", who);
    RMPCD_CHECK(n_links >= 0 && (n_links == 0 || (links && dist)), "%s: bad link arguments", who);
    return RMPCD_OK;
}

// a synchronous int32 result from the device
static int fetch_i32(const int32_t *dev, int32_t *host_out)
{
    int32_t *h = (int32_t *)pinned(64);
    if (!h) return set_error(RMPCD_ERR_HIP, "host staging allocation failed");
    RMPCD_HIP(hipMemcpyAsync(h, dev, sizeof(int32_t), hipMemcpyDeviceToHost, stream()));
    RMPCD_HIP(hipStreamSynchronize(stream()));
    *host_out = h[0];
    return RMPCD_OK;
}

static SchedView sched_view(const Schedule *s)
{
    return SchedView{s->d_i1, s->d_i2, s->d_start, s->d_order, s->d_d, s->n_links, s->n_levels};
}

template <bool kPos>
static int rattle_body(rmpcd_particles_t *p, double *pos_rattle, const int32_t *links, const double *distances,
                       int32_t n_links, const double *mass, int32_t n_mass, double dt, const double edges[3],
                       double precision, int32_t *iterations, const char *who)
{
    int rc = check_body(p, links, distances, n_links, who, false);
    if (rc) return rc;
    RMPCD_CHECK(edges && (!kPos || pos_rattle), "%s: null argument", who);
    BodyMass bm;
    if ((rc = body_mass(mass, n_mass, bm))) return rc;
    const Schedule *s;
    if ((rc = get_schedule(links, distances, n_links, p->n, who, &s))) return rc;
    const int64_t n = p->n;
    const size_t lds = body_lds_bytes(n, n_links, s->n_levels, kPos, true);
    int32_t *out = (int32_t *)workspace(WS_FLAG, 64);
    RMPCD_CHECK(out != nullptr, "%s: workspace allocation failed", who);
    if (lds <= BODY_LDS) {
        k_rattle_body<kPos, true><<<1, BT, lds, stream()>>>(view<F64>(p), pos_rattle, sched_view(s), bm, dt, edges[0],
                                                             edges[1], edges[2], precision, out, nullptr, nullptr);
    } else {
        char *w = (char *)workspace(WS_AUX, (size_t)n * 8 + 3 * (size_t)n * 4 + 256);
        RMPCD_CHECK(w != nullptr, "%s: workspace allocation failed", who);
        k_rattle_body<kPos, false><<<1, BT, 64 * 8, stream()>>>(view<F64>(p), pos_rattle, sched_view(s), bm, dt,
                                                                 edges[0], edges[1], edges[2], precision, out,
                                                                 (double *)w, (int32_t *)(w + (size_t)n * 8));
    }
    RMPCD_LAUNCHED();
    if (iterations) return fetch_i32(out, iterations);
    return RMPCD_OK;
}

}  // namespace rmpcd

using namespace rmpcd;

extern "C" {

int rmpcd_rattle_body_pos(rmpcd_particles_t *p, double *pos_rattle, const int32_t *links, const double *distances,
                          int32_t n_links, const double *mass, int32_t n_mass, double dt, const double edges[3],
                          double precision, int32_t *iterations)
{
    return rattle_body<true>(p, pos_rattle, links, distances, n_links, mass, n_mass, dt, edges, precision, iterations,
                             "rattle_body_pos");
}

int rmpcd_rattle_body_vel(rmpcd_particles_t *p, const int32_t *links, const double *distances, int32_t n_links,
                          const double *mass, int32_t n_mass, double dt, const double edges[3], double precision,
                          int32_t *iterations)
{
    return rattle_body<false>(p, nullptr, links, distances, n_links, mass, n_mass, dt, edges, precision, iterations,
                              "rattle_body_vel");
}

int rmpcd_elastic_network(rmpcd_particles_t *p, const int32_t *links, const double *distances, int32_t n_links,
                          double k, const double edges[3], double *energy)
{
    const char *who = "elastic_network";
    int rc = check_body(p, links, distances, n_links, who, true);
    if (rc) return rc;
    RMPCD_CHECK(edges && energy, "%s: null argument", who);
    const Schedule *s;
    if ((rc = get_schedule(links, distances, n_links, p->n, who, &s))) return rc;
    // the energy terms land at their link's place in the caller's order
    const size_t nl = (size_t)(n_links > 0 ? n_links : 1);
    double *terms = (double *)workspace(WS_AUX, nl * 8 + 256);
    double *h = (double *)pinned(64);
    RMPCD_CHECK(terms && h, "%s: workspace allocation failed", who);
    double *out = terms + nl;
    k_elastic_network<<<1, BT, 0, stream()>>>(view<F64>(p), sched_view(s), k, edges[0], edges[1], edges[2], terms,
                                               out);
    RMPCD_LAUNCHED();
    RMPCD_HIP(hipMemcpyAsync(h, out, sizeof(double), hipMemcpyDeviceToHost, stream()));
    RMPCD_HIP(hipStreamSynchronize(stream()));
    *energy = h[0];
    return RMPCD_OK;
}

int rmpcd_rigid_body_init(rmpcd_rigid_body_t *body, const rmpcd_particles_t *p, const double *mass, int32_t n_mass,
                          const double edges[3], int32_t i_start, int32_t i_stop, double *pos_body)
{
    int rc = check_particles(p, "rigid_body_init", false);
    if (rc) return rc;
    RMPCD_CHECK(p->form == RMPCD_FP64 && p->image, "rigid_body_init: colloids must use the FP64 form, with images");
    RMPCD_CHECK(body && edges && pos_body, "rigid_body_init: null argument");
    RMPCD_CHECK(i_start >= 0 && i_stop >= i_start && i_stop < p->n, "rigid_body_init: particle range out of bounds");
    BodyMass bm;
    if ((rc = body_mass(mass, n_mass, bm))) return rc;
    rmpcd_rigid_body_t h{};
    h.i_start = i_start;
    h.i_stop = i_stop;
    h.pos_body = pos_body;
    RMPCD_HIP(hipMemcpyAsync(body, &h, sizeof(h), hipMemcpyHostToDevice, stream()));
    k_rigid_init<<<1, BT, 0, stream()>>>(body, view<F64>(p), bm, edges[0], edges[1], edges[2]);
    RMPCD_LAUNCHED();
    RMPCD_HIP(hipStreamSynchronize(stream()));
    return RMPCD_OK;
}

int rmpcd_rigid_body_compute_force_torque(rmpcd_rigid_body_t *body, const rmpcd_particles_t *p)
{
    const char *who = "rigid_body_compute_force_torque";
    int rc = check_particles(p, who, true);
    if (rc) return rc;
    RMPCD_CHECK(p->form == RMPCD_FP64, "%s: colloids must use the FP64 form", who);
    RMPCD_CHECK(body != nullptr, "%s: null body", who);
    k_rigid_force_torque<<<1, BT, 0, stream()>>>(body, view<F64>(p));
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

int rmpcd_rigid_body_vv1(rmpcd_rigid_body_t *body, rmpcd_particles_t *p, double dt, double threshold,
                         const double edges[3], int32_t *iterations)
{
    const char *who = "rigid_body_vv1";
    int rc = check_particles(p, who, false);
    if (rc) return rc;
    RMPCD_CHECK(p->form == RMPCD_FP64, "%s: colloids must use the FP64 form", who);
    RMPCD_CHECK(body && edges, "%s: null argument", who);
    int32_t *out = (int32_t *)workspace(WS_FLAG, 64);
    RMPCD_CHECK(out != nullptr, "%s: workspace allocation failed", who);
    k_rigid_vv1<<<1, BT, 0, stream()>>>(body, view<F64>(p), dt, threshold, edges[0], edges[1], edges[2], out);
    RMPCD_LAUNCHED();
    if (!iterations) return RMPCD_OK;
    if ((rc = fetch_i32(out, iterations))) return rc;
    RMPCD_CHECK(*iterations >= 0, "%s: the body's particles are outside [0, n)", who);
    return RMPCD_OK;
}

int rmpcd_rigid_body_vv2(rmpcd_rigid_body_t *body, rmpcd_particles_t *p, double dt)
{
    const char *who = "rigid_body_vv2";
    int rc = check_particles(p, who, false);
    if (rc) return rc;
    RMPCD_CHECK(p->form == RMPCD_FP64, "%s: colloids must use the FP64 form", who);
    RMPCD_CHECK(body != nullptr, "%s: null body", who);
    k_rigid_vv2<<<1, BT, 0, stream()>>>(body, view<F64>(p), dt);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

}  // extern "C"
