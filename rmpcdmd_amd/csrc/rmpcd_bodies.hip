// rmpcd_bodies.hip -- colloids held together by links: iterative RATTLE over a link list
// (src/md.f90:321-480), the elastic network (516-544) and the rigid body with
// quaternion velocity Verlet (547-700).
//
// The reference updates the links one after the other, each correction seeing the ones
// before it, sweep after sweep until the constraints hold.  That order is kept exactly.
// On the host the link list becomes a schedule of levels: an update's level is one more
// than the highest level of the earlier updates that share a particle with it.  Updates
// of one level touch disjoint particles and every earlier update of the same particle
// sits in a lower level, so a workgroup running the levels in turn, with a barrier in
// between, reads and writes every particle in the order of the sequential loops: the
// results are the reference's bit for bit.
//
// Sweeps overlap.  The schedule covers a block of K consecutive sweeps as one unrolled
// sequence, so the first links of sweep s + 1 start while the last ones of sweep s are
// still running (on a 300-bead shell one sweep alone takes 155 levels, each further
// sweep of the block about 16).  The stopping test of a sweep needs the state at its end:
// the update that is a particle's last in a sweep copies the particle to that sweep's
// snapshot, the block's tests then run on the snapshots, and the state of the first sweep
// that passes is put back before the block's later sweeps are forgotten.  Body, links
// and schedule live in LDS (up to 160 KiB on gfx950); bodies too large for it run the
// one-sweep schedule from global memory.
#include <math.h>
#include <string.h>

#include <vector>

#include "rmpcd_internal.hpp"

namespace rmpcd {

constexpr int BT = 256;           // threads of the one workgroup
constexpr int RATTLE_MAX = 1000;  // md.f90:341, :432
constexpr int KMAX = 12;          // sweeps per block at most (4 bits of the entry)

struct BodyMass {
    double m[16];    // particle_system%mass by species (species s at s - 1)
    double inv[16];  // 1 / mass, as the reference's i_mass
};

__device__ inline double dot3b(const double a[3], const double b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// ---- the link schedules -----------------------------------------------------------------
// Built on the host from the caller's link list and kept on the device while the list is
// unchanged (compared by content on every call).
struct BlockSched {
    int K = 0, n_glev = 0;
    uint2 *d_ent = nullptr;   // by level: {a | b << 16, l | s << 24 | last-of-a << 28 | last-of-b << 29}
    int32_t *d_gs = nullptr;  // level starts, n_glev + 1
    void *block = nullptr;
};

struct Schedule {
    std::vector<int32_t> links;  // the caller's list, for the comparison
    std::vector<double> dist;
    int device = -1;             // the device whose memory holds it
    int64_t n = 0;
    int n_links = 0, n_levels = 0;
    // one sweep by level (the global-memory rattle)
    int32_t *d_i1 = nullptr, *d_i2 = nullptr, *d_start = nullptr;
    double *d_d = nullptr;
    // by link: a, b, a | b << 16 (when n < 65536), distance; per particle: the links
    // touching it in link order (bit 31 set where it is the second particle)
    int32_t *d_la = nullptr, *d_lb = nullptr;
    uint32_t *d_lab = nullptr;
    double *d_dist = nullptr;
    int32_t *d_cstart = nullptr, *d_csr = nullptr;
    void *block = nullptr;
    BlockSched blocks[KMAX + 1];  // by K, built on first use
};

static Schedule g_sched[4];
static int g_sched_next = 0;

static void free_schedule(Schedule &s)
{
    for (auto &b : s.blocks) {
        if (b.block) (void)hipFree(b.block);
        b = BlockSched();
    }
    if (s.block) (void)hipFree(s.block);
    s.block = nullptr;
}

template <class T>
static int upload(T *dst, const std::vector<T> &src)
{
    if (!src.empty()) RMPCD_HIP(hipMemcpy(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return RMPCD_OK;
}

static int get_schedule(const int32_t *links, const double *dist, int n_links, int64_t n, const char *who,
                        Schedule **out)
{
    int device = 0;
    RMPCD_HIP(hipGetDevice(&device));
    for (auto &s : g_sched)
        if (s.block && s.device == device && s.n == n && s.n_links == n_links && n_links > 0 &&
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
    const size_t nl = (size_t)n_links;
    std::vector<int32_t> start((size_t)n_levels + 1, 0), i1(nl), i2(nl), la(nl), lb(nl);
    std::vector<uint32_t> lab(nl);
    std::vector<double> dd(nl);
    for (int l = 0; l < n_links; ++l) ++start[lvl[l] + 1];
    for (int v = 0; v < n_levels; ++v) start[v + 1] += start[v];
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int l = 0; l < n_links; ++l) {
        const int k = fill[lvl[l]]++;
        i1[k] = links[2 * l];
        i2[k] = links[2 * l + 1];
        dd[k] = dist[l];
        la[l] = links[2 * l];
        lb[l] = links[2 * l + 1];
        lab[l] = (uint32_t)links[2 * l] | ((uint32_t)links[2 * l + 1] << 16);
    }
    std::vector<int32_t> cstart((size_t)n + 1, 0), csr(2 * nl);
    for (int l = 0; l < n_links; ++l) {
        ++cstart[links[2 * l] + 1];
        ++cstart[links[2 * l + 1] + 1];
    }
    for (int64_t p = 0; p < n; ++p) cstart[p + 1] += cstart[p];
    std::vector<int32_t> cf(cstart.begin(), cstart.end() - 1);
    for (int l = 0; l < n_links; ++l) {
        csr[cf[links[2 * l]]++] = l;
        csr[cf[links[2 * l + 1]]++] = (int32_t)((uint32_t)l | 0x80000000u);
    }
    Schedule &s = g_sched[g_sched_next];
    g_sched_next = (g_sched_next + 1) % 4;
    if (s.block) {
        // a kernel may still read it; its memory belongs to the device it was made on
        int cur = device;
        RMPCD_HIP(hipSetDevice(s.device));
        RMPCD_HIP(hipDeviceSynchronize());
        free_schedule(s);
        RMPCD_HIP(hipSetDevice(cur));
    }
    const size_t nl1 = nl > 0 ? nl : 1;
    const size_t bytes = nl1 * 16 + nl1 * 4 * 5 + ((size_t)n_levels + 1) * 4 + ((size_t)n + 1) * 4 + 2 * nl1 * 4 + 256;
    RMPCD_HIP(hipMalloc(&s.block, bytes));
    s.d_d = (double *)s.block;
    s.d_dist = s.d_d + nl1;
    s.d_i1 = (int32_t *)(s.d_dist + nl1);
    s.d_i2 = s.d_i1 + nl1;
    s.d_la = s.d_i2 + nl1;
    s.d_lb = s.d_la + nl1;
    s.d_lab = (uint32_t *)(s.d_lb + nl1);
    s.d_csr = (int32_t *)(s.d_lab + nl1);
    s.d_cstart = s.d_csr + 2 * nl1;
    s.d_start = s.d_cstart + n + 1;
    std::vector<double> dist_v(dist, dist + n_links);
    int rc;
    if ((rc = upload(s.d_d, dd)) || (rc = upload(s.d_dist, dist_v)) || (rc = upload(s.d_i1, i1)) ||
        (rc = upload(s.d_i2, i2)) || (rc = upload(s.d_la, la)) || (rc = upload(s.d_lb, lb)) ||
        (rc = upload(s.d_lab, lab)) || (rc = upload(s.d_csr, csr)) || (rc = upload(s.d_cstart, cstart)) ||
        (rc = upload(s.d_start, start)))
        return rc;
    s.links.assign(links, links + 2 * nl);
    s.dist = dist_v;
    s.device = device;
    s.n = n;
    s.n_links = n_links;
    s.n_levels = n_levels;
    *out = &s;
    return RMPCD_OK;
}

// the unrolled schedule of K sweeps, built once per list and K
static int get_block(Schedule &s, int K, const BlockSched **out)
{
    BlockSched &b = s.blocks[K];
    if (b.block) {
        *out = &b;
        return RMPCD_OK;
    }
    const int nl = s.n_links;
    const int32_t *links = s.links.data();
    std::vector<int32_t> last((size_t)s.n, 0), lastlink((size_t)s.n, -1), lvl((size_t)K * nl);
    for (int l = 0; l < nl; ++l) lastlink[links[2 * l]] = lastlink[links[2 * l + 1]] = l;
    int n_glev = 0;
    for (int q = 0; q < K; ++q)
        for (int l = 0; l < nl; ++l) {
            const int a = links[2 * l], c = links[2 * l + 1];
            const int v = last[a] > last[c] ? last[a] : last[c];
            lvl[(size_t)q * nl + l] = v;
            last[a] = last[c] = v + 1;
            if (v + 1 > n_glev) n_glev = v + 1;
        }
    std::vector<int32_t> gs((size_t)n_glev + 1, 0);
    for (size_t k = 0; k < lvl.size(); ++k) ++gs[lvl[k] + 1];
    for (int v = 0; v < n_glev; ++v) gs[v + 1] += gs[v];
    std::vector<int32_t> fill(gs.begin(), gs.end() - 1);
    std::vector<uint2> ent(lvl.size());
    for (int q = 0; q < K; ++q)
        for (int l = 0; l < nl; ++l) {
            const int a = links[2 * l], c = links[2 * l + 1];
            uint2 e;
            e.x = (uint32_t)a | ((uint32_t)c << 16);
            e.y = (uint32_t)l | ((uint32_t)q << 24) | ((lastlink[a] == l ? 1u : 0u) << 28) |
                  ((lastlink[c] == l ? 1u : 0u) << 29);
            ent[fill[lvl[(size_t)q * nl + l]]++] = e;
        }
    RMPCD_HIP(hipMalloc(&b.block, ent.size() * sizeof(uint2) + gs.size() * 4 + 256));
    b.d_ent = (uint2 *)b.block;
    b.d_gs = (int32_t *)(b.d_ent + ent.size());
    int rc;
    if ((rc = upload(b.d_ent, ent)) || (rc = upload(b.d_gs, gs))) return rc;
    b.K = K;
    b.n_glev = n_glev;
    *out = &b;
    return RMPCD_OK;
}

// bytes of LDS the pipelined rattle takes: particles (P, V and for positions R, four
// doubles each), links (distance, packed pair), the entries and level starts, flags
static size_t pipe_lds_bytes(int64_t n, int n_links, int K, int n_glev, bool with_pr)
{
    return (size_t)n * 32 * (with_pr ? 3 : 2) + (size_t)n_links * 12 + (size_t)K * n_links * 8 +
           ((size_t)n_glev + 1) * 4 + 16 * 4 + 64;
}

static int lds_limit()
{
    static int limit = 0;
    if (!limit) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 64 * 1024;
        limit = v;
    }
    return limit;
}

// a workgroup barrier that orders LDS only: global stores stay in flight
__device__ inline void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

struct PipeView {
    const uint2 *ent;
    const int32_t *gs;
    const uint32_t *lab;
    const double *dist;
    int n_links, n_ent, n_glev, K;
};

// ---- rattle_body_pos (kPos) / rattle_body_vel, pipelined -------------------------------
// Staging: every particle but the first is unfolded next to particle 0 (md.f90:343-353,
// :427-431) and folded back at the end (md.f90:400-402, :474-476); pos_rattle is left
// unfolded, as the reference leaves it.
template <bool kPos>
__global__ void __launch_bounds__(BT) k_rattle_pipe(PView<F64> p, double *__restrict__ gpr, PipeView pv, BodyMass bm,
                                                    double dt, double e0, double e1, double e2, double precision,
                                                    int32_t *__restrict__ out, int32_t *__restrict__ gim,
                                                    double *__restrict__ snap)
{
    extern __shared__ double smem[];
    const int64_t n = p.n;
    const int tid = threadIdx.x;
    const int nl = pv.n_links, K = pv.K;
    const double edges[3] = {e0, e1, e2};
    double *P = smem;          // {x, y, z, 1/m}
    double *V = P + 4 * n;     // {vx, vy, vz, linked}
    double *R = V + 4 * n;     // {x, y, z, -} (positions only)
    double *D = kPos ? R + 4 * n : R;
    uint2 *E = (uint2 *)(D + nl);
    uint32_t *LAB = (uint32_t *)(E + pv.n_ent);
    int32_t *GS = (int32_t *)(LAB + nl);
    int32_t *NC = GS + pv.n_glev + 1;
    for (int64_t i = tid; i < n; i += BT) {
        for (int k = 0; k < 3; ++k) {
            const double x0 = p.pos[k * n], x = p.pos[k * n + i];
            const int j = i == 0 ? 0 : (int)floor(((x - x0) / edges[k]) + 0.5);
            gim[k * n + i] = j;
            P[4 * i + k] = i == 0 ? x : x - j * edges[k];
            V[4 * i + k] = p.vel[k * n + i];
            if (kPos) {
                const double r0 = gpr[k * n], r = gpr[k * n + i];
                const int jr = i == 0 ? 0 : (int)floor(((r - r0) / edges[k]) + 0.5);
                R[4 * i + k] = i == 0 ? r : r - jr * edges[k];
            }
        }
        P[4 * i + 3] = bm.inv[((int)p.species[i] - 1) & 15];
        V[4 * i + 3] = 0;
    }
    for (int l = tid; l < nl; l += BT) {
        D[l] = pv.dist[l];
        LAB[l] = pv.lab[l];
    }
    for (int x = tid; x < pv.n_ent; x += BT) E[x] = pv.ent[x];
    for (int g = tid; g <= pv.n_glev; g += BT) GS[g] = pv.gs[g];
    __syncthreads();
    for (int l = tid; l < nl; l += BT) {
        V[4 * (LAB[l] & 0xFFFF) + 3] = 1;
        V[4 * (LAB[l] >> 16) + 3] = 1;
    }
    __syncthreads();
    int it = RATTLE_MAX + 1;
    // one link update of the schedule (md.f90:365-383, :435-452)
    auto update = [&](const uint2 e, const int ns) {
        const int q = (e.y >> 24) & 15;
        if (q >= ns) return;
        const int a = e.x & 0xFFFF, b = e.x >> 16, l = e.y & 0xFFFFFF;
        const double pa[3] = {P[4 * a], P[4 * a + 1], P[4 * a + 2]};
        const double pb[3] = {P[4 * b], P[4 * b + 1], P[4 * b + 2]};
        const double va[3] = {V[4 * a], V[4 * a + 1], V[4 * a + 2]};
        const double vb[3] = {V[4 * b], V[4 * b + 1], V[4 * b + 2]};
        const double ma = P[4 * a + 3], mb = P[4 * b + 3], d = D[l];
        double s[3], na[3], nb[3], wa[3], wb[3];
        for (int k = 0; k < 3; ++k) s[k] = pa[k] - pb[k];
        if (kPos) {
            double r[3];
            for (int k = 0; k < 3; ++k) r[k] = R[4 * a + k] - R[4 * b + k];
            const double rsq = dot3b(r, r), ssq = dot3b(s, s), rs = dot3b(r, s);
            double gg = rs - sqrt(rs * rs - rsq * (ssq - d * d));
            gg = gg / ((dt * (ma + mb)) * rsq);
            // a link from a particle to itself reads its own update, as the reference does
            for (int k = 0; k < 3; ++k) {
                na[k] = pa[k] - ((gg * dt) * r[k]) * ma;
                nb[k] = (a == b ? na[k] : pb[k]) + ((gg * dt) * r[k]) * mb;
                wa[k] = va[k] - (gg * r[k]) * ma;
                wb[k] = (a == b ? wa[k] : vb[k]) + (gg * r[k]) * mb;
            }
            for (int k = 0; k < 3; ++k) {
                P[4 * a + k] = na[k];
                V[4 * a + k] = wa[k];
            }
            for (int k = 0; k < 3; ++k) {
                P[4 * b + k] = nb[k];
                V[4 * b + k] = wb[k];
            }
        } else {
            double dv[3];
            for (int k = 0; k < 3; ++k) dv[k] = va[k] - vb[k];
            const double kk = dot3b(dv, s) / ((d * d) * (ma + mb));
            for (int k = 0; k < 3; ++k) {
                na[k] = pa[k];
                nb[k] = pb[k];
                wa[k] = va[k] - (kk * s[k]) * ma;
                wb[k] = (a == b ? wa[k] : vb[k]) + (kk * s[k]) * mb;
            }
            for (int k = 0; k < 3; ++k) V[4 * a + k] = wa[k];
            for (int k = 0; k < 3; ++k) V[4 * b + k] = wb[k];
        }
        // the last update of a particle in a sweep leaves its end-of-sweep state
        if (e.y & (1u << 28)) {
            double *o = snap + ((size_t)q * n + a) * 8;
            for (int k = 0; k < 3; ++k) {
                o[k] = na[k];
                o[3 + k] = wa[k];
            }
        }
        if (e.y & (1u << 29)) {
            double *o = snap + ((size_t)q * n + b) * 8;
            for (int k = 0; k < 3; ++k) {
                o[k] = nb[k];
                o[3 + k] = wb[k];
            }
        }
    };
    const int last_ent = pv.n_ent > 0 ? pv.n_ent - 1 : 0;
    for (int base = 0; base < RATTLE_MAX; base += K) {
        const int ns = K < RATTLE_MAX - base ? K : RATTLE_MAX - base;
        // each level's first entry of this thread is read one level ahead; the barrier
        // between levels orders LDS only (the snapshots are read after the full barrier
        // that ends the block)
        int x0 = GS[0], x1 = pv.n_glev > 0 ? GS[1] : 0;
        uint2 cur = E[min(x0 + tid, last_ent)];
        for (int g = 0; g < pv.n_glev; ++g) {
            const uint2 next = E[min(x1 + tid, last_ent)];
            const int y1 = GS[g + 2 <= pv.n_glev ? g + 2 : pv.n_glev];
            if (x0 + tid < x1) update(cur, ns);
            for (int x = x0 + tid + BT; x < x1; x += BT) update(E[x], ns);
            lds_barrier();
            x0 = x1;
            x1 = y1;
            cur = next;
        }
        __syncthreads();
        // the stopping test of each sweep of the block (md.f90:385-392, :454-465): a sweep
        // passes when no link reaches the precision (a NaN never does).  Items run sweep
        // fastest, four at a time, and skip sweeps already failed.
        if (tid < K) NC[tid] = 0 < precision ? 0 : 1;
        __syncthreads();
        const int n_items = ns * nl;
        for (int x0c = tid; x0c < n_items; x0c += 4 * BT) {
            double err[4];
            int qs[4];
            bool live[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int x = x0c + u * BT;
                const int q = x % ns, l = x / ns;
                qs[u] = q;
                live[u] = x < n_items && !NC[q];
                err[u] = 0;
                if (!live[u]) continue;
                const int a = LAB[l] & 0xFFFF, b = LAB[l] >> 16;
                const double *sa = snap + ((size_t)q * n + a) * 8, *sb = snap + ((size_t)q * n + b) * 8;
                if (kPos) {
                    double r[3];
                    for (int k = 0; k < 3; ++k) r[k] = sa[k] - sb[k];
                    err[u] = fabs(sqrt(dot3b(r, r)) - D[l]);
                } else {
                    double sv[3], dv[3];
                    for (int k = 0; k < 3; ++k) {
                        sv[k] = P[4 * a + k] - P[4 * b + k];
                        dv[k] = sa[3 + k] - sb[3 + k];
                    }
                    err[u] = fabs(dot3b(dv, sv) / ((D[l] * D[l]) * (P[4 * a + 3] + P[4 * b + 3])));
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (live[u] && err[u] >= precision) NC[qs[u]] = 1;
        }
        __syncthreads();
        int conv = -1;
        for (int q = 0; q < ns; ++q)
            if (!NC[q]) {
                conv = q;
                break;
            }
        if (conv >= 0) {
            it = base + conv + 1;
            for (int64_t i = tid; i < n; i += BT)
                if (V[4 * i + 3] != 0) {
                    const double *o = snap + ((size_t)conv * n + i) * 8;
                    for (int k = 0; k < 3; ++k) {
                        P[4 * i + k] = o[k];
                        V[4 * i + k] = o[3 + k];
                    }
                }
            __syncthreads();
            break;
        }
        __syncthreads();
    }
    for (int64_t i = tid; i < n; i += BT)
        for (int k = 0; k < 3; ++k) {
            p.pos[k * n + i] = i == 0 ? P[4 * i + k] : P[4 * i + k] + gim[k * n + i] * edges[k];
            p.vel[k * n + i] = V[4 * i + k];
            if (kPos) gpr[k * n + i] = R[4 * i + k];
        }
    if (tid == 0) out[0] = it;
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

struct SchedView {
    const int32_t *i1, *i2, *start;
    const double *d;
    int n_links, n_levels;
};

// ---- rattle_body_pos (kPos) / rattle_body_vel from global memory ----------------------
// For bodies whose links and schedule do not fit in LDS: one sweep at a time, its levels
// in turn, the particles updated in place.
template <bool kPos>
__global__ void __launch_bounds__(BT) k_rattle_global(PView<F64> p, double *__restrict__ pr, SchedView sv,
                                                      BodyMass bm, double dt, double e0, double e1, double e2,
                                                      double precision, int32_t *__restrict__ out,
                                                      double *__restrict__ minv, int32_t *__restrict__ im)
{
    __shared__ double red[64];
    const int64_t n = p.n;
    const int tid = threadIdx.x;
    const double edges[3] = {e0, e1, e2};
    double *pos = p.pos, *vel = p.vel;
    for (int64_t i = tid; i < n; i += BT) {
        minv[i] = bm.inv[((int)p.species[i] - 1) & 15];
        for (int k = 0; k < 3; ++k) {
            const double x0 = pos[k * n], x = pos[k * n + i];
            const int j = i == 0 ? 0 : (int)floor(((x - x0) / edges[k]) + 0.5);
            im[k * n + i] = j;
            pos[k * n + i] = i == 0 ? x : x - j * edges[k];
            if (kPos) {
                const double r0 = pr[k * n], r = pr[k * n + i];
                const int jr = i == 0 ? 0 : (int)floor(((r - r0) / edges[k]) + 0.5);
                pr[k * n + i] = i == 0 ? r : r - jr * edges[k];
            }
        }
    }
    __syncthreads();
    int it;
    for (it = 1; it <= RATTLE_MAX; ++it) {
        for (int v = 0; v < sv.n_levels; ++v) {
            for (int l = sv.start[v] + tid; l < sv.start[v + 1]; l += BT) {
                const int a = sv.i1[l], b = sv.i2[l];
                const double ma = minv[a], mb = minv[b], dl = sv.d[l];
                double s[3];
                for (int k = 0; k < 3; ++k) s[k] = pos[k * n + a] - pos[k * n + b];
                if (kPos) {
                    double r[3];
                    for (int k = 0; k < 3; ++k) r[k] = pr[k * n + a] - pr[k * n + b];
                    const double rsq = dot3b(r, r), ssq = dot3b(s, s), rs = dot3b(r, s);
                    double g = rs - sqrt(rs * rs - rsq * (ssq - dl * dl));
                    g = g / ((dt * (ma + mb)) * rsq);
                    for (int k = 0; k < 3; ++k) {
                        pos[k * n + a] = pos[k * n + a] - ((g * dt) * r[k]) * ma;
                        pos[k * n + b] = pos[k * n + b] + ((g * dt) * r[k]) * mb;
                        vel[k * n + a] = vel[k * n + a] - (g * r[k]) * ma;
                        vel[k * n + b] = vel[k * n + b] + (g * r[k]) * mb;
                    }
                } else {
                    double dv[3];
                    for (int k = 0; k < 3; ++k) dv[k] = vel[k * n + a] - vel[k * n + b];
                    const double kk = dot3b(dv, s) / ((dl * dl) * (ma + mb));
                    for (int k = 0; k < 3; ++k) {
                        vel[k * n + a] = vel[k * n + a] - (kk * s[k]) * ma;
                        vel[k * n + b] = vel[k * n + b] + (kk * s[k]) * mb;
                    }
                }
            }
            __syncthreads();
        }
        double err = 0;
        for (int l = tid; l < sv.n_links; l += BT) {
            const int a = sv.i1[l], b = sv.i2[l];
            double s[3];
            for (int k = 0; k < 3; ++k) s[k] = pos[k * n + a] - pos[k * n + b];
            double e;
            if (kPos) {
                e = fabs(sqrt(dot3b(s, s)) - sv.d[l]);
            } else {
                double dv[3];
                for (int k = 0; k < 3; ++k) dv[k] = vel[k * n + a] - vel[k * n + b];
                e = fabs(dot3b(dv, s) / ((sv.d[l] * sv.d[l]) * (minv[a] + minv[b])));
            }
            if (e > err) err = e;
        }
        err = body_max(err, red);
        if (err < precision) break;
    }
    __syncthreads();
    for (int64_t i = tid + 1; i < n; i += BT)
        for (int k = 0; k < 3; ++k) pos[k * n + i] = pos[k * n + i] + im[k * n + i] * edges[k];
    if (tid == 0) out[0] = it;
}

// elastic_network (md.f90:516-544).  The link forces are independent: made in parallel
// first.  Each particle then adds the forces of its links in link order (the per-particle
// lists of the schedule), which is the order of the reference's loop, and the energy
// terms are summed in link order.  kLds: the terms and forces of the links in LDS.
template <bool kLds>
__global__ void __launch_bounds__(BT) k_elastic_network(PView<F64> p, const int32_t *__restrict__ la,
                                                        const int32_t *__restrict__ lb, const double *__restrict__ dist,
                                                        const int32_t *__restrict__ cstart,
                                                        const int32_t *__restrict__ csr, int n_links, double kk,
                                                        double e0, double e1, double e2, double *__restrict__ gterms,
                                                        double *__restrict__ out)
{
    extern __shared__ double smem[];
    const int64_t n = p.n;
    const int tid = threadIdx.x;
    const double L[3] = {e0, e1, e2};
    double *T = kLds ? smem : gterms;
    double *F = T + n_links;
    for (int l = tid; l < n_links; l += BT) {
        const int a = la[l], b = lb[l];
        const double r0 = dist[l];
        double x[3];
        for (int k = 0; k < 3; ++k) x[k] = min_image(p.pos[k * n + a] - p.pos[k * n + b], L[k]);
        const double r = sqrt(dot3b(x, x));
        T[l] = (kk * ((r - r0) * (r - r0))) / 2;
        for (int k = 0; k < 3; ++k) F[3 * l + k] = ((-kk) * (r - r0)) * x[k] / r;
    }
    __syncthreads();
    for (int64_t i = tid; i < n; i += BT) {
        const int j0 = cstart[i], j1 = cstart[i + 1];
        if (j0 == j1) continue;
        double f[3] = {p.force[i], p.force[n + i], p.force[2 * n + i]};
        for (int j = j0; j < j1; ++j) {
            const uint32_t c = (uint32_t)csr[j];
            const int l = (int)(c & 0x7FFFFFFFu);
            if (c >> 31)
                for (int k = 0; k < 3; ++k) f[k] = f[k] - F[3 * l + k];
            else
                for (int k = 0; k < 3; ++k) f[k] = f[k] + F[3 * l + k];
        }
        for (int k = 0; k < 3; ++k) p.force[k * n + i] = f[k];
    }
    if (tid == 0) {
        double e = 0;
#pragma unroll 8
        for (int l = 0; l < n_links; ++l) e = e + T[l];
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

// which rattle runs: the pipelined one with the largest K whose body fits in LDS, else
// the global one (K = 0)
static int pick_k(Schedule &s, bool with_pr, const BlockSched **blk, size_t *lds)
{
    *blk = nullptr;
    *lds = 0;
    if (s.n >= 65536 || s.n_links >= (1 << 24) || s.n_links == 0) return RMPCD_OK;
    const size_t limit = (size_t)lds_limit();
    for (int K = KMAX; K >= 1; --K) {
        // a cheap bound first: every entry could be its own level
        if (pipe_lds_bytes(s.n, s.n_links, K, 0, with_pr) > limit) continue;
        const BlockSched *b;
        int rc = get_block(s, K, &b);
        if (rc) return rc;
        const size_t need = pipe_lds_bytes(s.n, s.n_links, K, b->n_glev, with_pr);
        if (need <= limit) {
            *blk = b;
            *lds = need;
            return RMPCD_OK;
        }
    }
    return RMPCD_OK;
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
    const int64_t n = p->n;
    int32_t *out = (int32_t *)workspace(WS_FLAG, 64);
    RMPCD_CHECK(out != nullptr, "%s: workspace allocation failed", who);
    // no links: nothing to correct, the first sweep meets any positive precision, and
    // the positions still go through the reference's unfold and fold
    Schedule *s = nullptr;
    if (n_links > 0 && (rc = get_schedule(links, distances, n_links, n, who, &s))) return rc;
    const BlockSched *blk = nullptr;
    size_t lds = 0;
    if (s && (rc = pick_k(*s, kPos, &blk, &lds))) return rc;
    if (blk) {
        char *w = (char *)workspace(WS_AUX, 3 * (size_t)n * 4 + (size_t)blk->K * n * 64 + 512);
        RMPCD_CHECK(w != nullptr, "%s: workspace allocation failed", who);
        int32_t *gim = (int32_t *)w;
        double *snap = (double *)(w + ((3 * (size_t)n * 4 + 255) & ~(size_t)255));
        if (lds > 64 * 1024)
            RMPCD_HIP(hipFuncSetAttribute((const void *)k_rattle_pipe<kPos>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        PipeView pv{blk->d_ent, blk->d_gs, s->d_lab, s->d_dist, s->n_links, blk->K * s->n_links, blk->n_glev, blk->K};
        k_rattle_pipe<kPos><<<1, BT, lds, stream()>>>(view<F64>(p), pos_rattle, pv, bm, dt, edges[0], edges[1],
                                                       edges[2], precision, out, gim, snap);
    } else {
        char *w = (char *)workspace(WS_AUX, (size_t)n * 8 + 3 * (size_t)n * 4 + 256);
        RMPCD_CHECK(w != nullptr, "%s: workspace allocation failed", who);
        SchedView sv{nullptr, nullptr, nullptr, nullptr, 0, 0};
        if (s) sv = SchedView{s->d_i1, s->d_i2, s->d_start, s->d_d, s->n_links, s->n_levels};
        k_rattle_global<kPos><<<1, BT, 0, stream()>>>(view<F64>(p), pos_rattle, sv, bm, dt, edges[0], edges[1],
                                                       edges[2], precision, out, (double *)w,
                                                       (int32_t *)(w + (size_t)n * 8));
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
    if (n_links == 0) {
        *energy = 0;
        return RMPCD_OK;
    }
    Schedule *s;
    if ((rc = get_schedule(links, distances, n_links, p->n, who, &s))) return rc;
    const size_t tbytes = (size_t)n_links * 32;  // a term and a force per link
    double *w = (double *)workspace(WS_AUX, tbytes + 256);
    double *h = (double *)pinned(64);
    RMPCD_CHECK(w && h, "%s: workspace allocation failed", who);
    double *out = w + 4 * (size_t)n_links;
    if (tbytes <= (size_t)lds_limit()) {
        if (tbytes > 64 * 1024)
            RMPCD_HIP(hipFuncSetAttribute((const void *)k_elastic_network<true>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)tbytes));
        k_elastic_network<true><<<1, BT, tbytes, stream()>>>(view<F64>(p), s->d_la, s->d_lb, s->d_dist, s->d_cstart,
                                                             s->d_csr, n_links, k, edges[0], edges[1], edges[2], w,
                                                             out);
    } else {
        k_elastic_network<false><<<1, BT, 0, stream()>>>(view<F64>(p), s->d_la, s->d_lb, s->d_dist, s->d_cstart,
                                                          s->d_csr, n_links, k, edges[0], edges[1], edges[2], w, out);
    }
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
