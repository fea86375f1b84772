/*
 * rmpcd_oracle.c -- TEST INFRASTRUCTURE ONLY.  CPU (FP64, optional OpenMP)
 * restatement of the RMPCDMD hot path.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker / CPU baseline.
 * Each function cites the reference file:line it follows; see rmpcd_oracle.h for the
 * conventions (0-based indices, (3,N) layout, injected draws) and the parity status.
 *
 * Floating-point expressions keep the reference's evaluation order (Fortran evaluates
 * a + b + c as (a + b) + c, x**2 as x*x, x**3 as (x*x)*x and x**6 as the product of
 * six factors taken left to right, which is how amdflang expands integer powers) and the file is built with
 * -ffp-contract=off so no fused multiply-adds appear.
 */
#include "rmpcd_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* thread count of the parallel loops (1 = the reference's single-thread order, used for
 * parity; more threads only for the CPU baseline, where the loops that the reference
 * parallelises with OpenMP run in parallel too) */
static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }
#define ORC_PRAGMA(x) _Pragma(#x)
#define ORC_PARALLEL_FOR ORC_PRAGMA(omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1))
#define ORC_PARALLEL_FOR_REDUCE(...) \
    ORC_PRAGMA(omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1) reduction(__VA_ARGS__))

/* ======================= hilbert.f90 =========================================== */
/* src/hilbert.f90:14-15 */
#define HDIM 3
#define HMASK 7

static int h_rotate_right(int x, int d) /* hilbert.f90:20-30 */
{
    int r = (int)((unsigned)x >> d);
    int tmp = (int)((unsigned)x << (HDIM - d));
    return (r | tmp) & HMASK;
}
static int h_rotate_left(int x, int d) /* hilbert.f90:32-42 */
{
    int r = (int)((unsigned)x << d);
    int tmp = (int)((unsigned)x >> (HDIM - d));
    return (r | tmp) & HMASK;
}
static int h_gc(int i) { return i ^ (i >> 1); } /* hilbert.f90:63-69 */
static int h_entry_point(int i)                  /* hilbert.f90:71-80 */
{
    return i == 0 ? 0 : h_gc(2 * ((i - 1) / 2));
}
static int h_inverse_gc(int g) /* hilbert.f90:90-102 */
{
    int r = g;
    for (int j = 1; j < HDIM; ++j) r ^= (g >> j);
    return r;
}
static int h_trailz(int x)
{
    if (x == 0) return 32;
    int n = 0;
    while (!(x & 1)) { x >>= 1; ++n; }
    return n;
}
static int h_intercube_g(int i) { return h_trailz(h_gc(i) ^ h_gc(i + 1)); } /* :104-109 */
static int h_mod(int a, int b) { int r = a % b; return r < 0 ? r + b : r; }
static int h_intracube_d(int i) /* hilbert.f90:111-122 */
{
    if (i == 0) return 0;
    if (h_mod(i, 2) == 0) return h_mod(h_intercube_g(i - 1), HDIM);
    return h_mod(h_intercube_g(i), HDIM);
}
static int h_transform(int e, int d, int b) { return h_rotate_right(b ^ e, d + 1); } /* :124-129 */
static int h_inverse_transform(int e, int d, int b)                                   /* :131-137 */
{
    return h_transform(h_rotate_right(e, d + 1), HDIM - d - 2, b);
}
static int h_ibits(int x, int pos, int len) { return (int)(((unsigned)x >> pos) & ((1u << len) - 1u)); }

int orc_p_to_h(const int p[3], int m) /* hilbert.f90:139-162 */
{
    int h = 0, e = 0, d = 2;
    for (int i = m - 1; i >= 0; --i) {
        int l = 0;
        for (int j = 1; j <= HDIM; ++j) l += (1 << (j - 1)) * h_ibits(p[j - 1], i, 1);
        l = h_transform(e, d, l);
        int w = h_inverse_gc(l);
        e = e ^ h_rotate_left(h_entry_point(w), d + 1);
        d = h_mod(d + h_intracube_d(w) + 1, HDIM);
        h = (h << HDIM) | w;
    }
    return h;
}

void orc_h_to_p(int h, int m, int p[3]) /* hilbert.f90:164-186 */
{
    int e = 0, d = 2;
    p[0] = p[1] = p[2] = 0;
    for (int i = m - 1; i >= 0; --i) {
        int w = 0;
        for (int j = 0; j < HDIM; ++j) w += (1 << j) * h_ibits(h, i * HDIM + j, 1);
        int l = h_gc(w);
        l = h_inverse_transform(e, d, l);
        for (int j = 1; j <= HDIM; ++j) p[j - 1] += h_ibits(l, j - 1, 1) << i;
        e = e ^ h_rotate_left(h_entry_point(w), d + 1);
        d = h_mod(d + h_intracube_d(w) + 1, HDIM);
    }
}

static int h_gcr(int i, int mu) /* hilbert.f90:188-200 */
{
    int r = 0;
    for (int k = HDIM - 1; k >= 0; --k)
        if (h_ibits(mu, k, 1) == 1) r = (r << 1) | h_ibits(i, k, 1);
    return r;
}

static int h_inverse_gcr(int r, int mu, int pi) /* hilbert.f90:202-224 */
{
    int i = 0, g = 0, j = -1;
    for (int k = 0; k < HDIM; ++k) j += h_ibits(mu, k, 1);
    for (int k = HDIM - 1; k >= 0; --k) {
        if (h_ibits(mu, k, 1) == 1) {
            i |= h_ibits(r, j, 1) << k;
            g |= h_mod(h_ibits(i, k, 1) + h_ibits(i, k + 1, 1), 2) << k;
            j -= 1;
        } else {
            g |= h_ibits(pi, k, 1) << k;
            i |= h_mod(h_ibits(g, k, 1) + h_ibits(i, k + 1, 1), 2) << k;
        }
    }
    return i;
}

static int h_extract_mask(int i, const int m[3]) /* hilbert.f90:226-239 */
{
    int mu = 0;
    for (int j = HDIM; j >= 1; --j) {
        mu <<= 1;
        if (m[j - 1] > i) mu |= 1;
    }
    return mu;
}

int orc_compact_p_to_h(const int p[3], const int m[3]) /* hilbert.f90:241-271 */
{
    int h = 0, e = 0, d = 2;
    int max_m = m[0];
    if (m[1] > max_m) max_m = m[1];
    if (m[2] > max_m) max_m = m[2];
    for (int i = max_m - 1; i >= 0; --i) {
        int mu = h_extract_mask(i, m);
        int mu_norm = 0;
        for (int j = 0; j < HDIM; ++j) mu_norm += h_ibits(mu, j, 1);
        mu = h_rotate_right(mu, d + 1);
        int l = 0;
        for (int j = 1; j <= HDIM; ++j) l += (1 << (j - 1)) * h_ibits(p[j - 1], i, 1);
        l = h_transform(e, d, l);
        int w = h_inverse_gc(l);
        int r = h_gcr(w, mu);
        e = e ^ h_rotate_left(h_entry_point(w), d + 1);
        d = h_mod(d + h_intracube_d(w) + 1, HDIM);
        h = (h << mu_norm) | r;
    }
    return h;
}

void orc_compact_h_to_p(int h, const int m[3], int p[3]) /* hilbert.f90:273-307 */
{
    int e = 0, d = 2, k = 0;
    int max_m = m[0];
    if (m[1] > max_m) max_m = m[1];
    if (m[2] > max_m) max_m = m[2];
    int sum_m = m[0] + m[1] + m[2];
    p[0] = p[1] = p[2] = 0;
    for (int i = max_m - 1; i >= 0; --i) {
        int mu = h_extract_mask(i, m);
        int mu_norm = 0;
        for (int j = 0; j < HDIM; ++j) mu_norm += h_ibits(mu, j, 1);
        mu = h_rotate_right(mu, d + 1);
        int pi = h_rotate_right(e, d + 1) & (~mu & HMASK);
        int r = 0;
        for (int j = 0; j < mu_norm; ++j) r += (1 << j) * h_ibits(h, sum_m - k - mu_norm + j, 1);
        k += mu_norm;
        int w = h_inverse_gcr(r, mu, pi);
        int l = h_gc(w);
        l = h_inverse_transform(e, d, l);
        for (int j = 1; j <= HDIM; ++j) p[j - 1] += h_ibits(l, j - 1, 1) << i;
        e = e ^ h_rotate_left(h_entry_point(w), d + 1);
        d = h_mod(d + h_intracube_d(w) + 1, HDIM);
    }
}

/* ======================= cell_system.f90 ======================================= */
void orc_cells_init(orc_cells_t *c, const int L[3], double a, int has_walls) /* :49-84 */
{
    memset(c, 0, sizeof(*c));
    for (int k = 0; k < 3; ++k) {
        c->L[k] = L[k];
        c->edges[k] = (double)L[k] * a;
    }
    c->has_walls = has_walls ? 1 : 0;
    if (c->has_walls) c->L[2] = L[2] + 1;
    c->a = a;
    for (int k = 0; k < 3; ++k) {
        c->M[k] = 1;
        /* the loop bound is the argument L, not the padded this%L (cell_system.f90:73) */
        while ((1 << c->M[k]) < L[k]) c->M[k] += 1;
        c->bc[k] = ORC_PERIODIC_BC;
    }
    c->N = (1 << c->M[0]) * (1 << c->M[1]) * (1 << c->M[2]);
}

void orc_cartesian_indices(const orc_cells_t *c, const double x[3], int p[3]) /* :165-175 */
{
    for (int k = 0; k < 3; ++k) p[k] = (int)floor((x[k] / c->a) - c->origin[k]);
    if (p[0] == c->L[0]) p[0] = 0;
    if (p[1] == c->L[1]) p[1] = 0;
    if (!c->has_walls && p[2] == c->L[2]) p[2] = 0;
}

int orc_cell_key(const orc_cells_t *c, const double x[3])
{
    int p[3];
    orc_cartesian_indices(c, x, p);
    return orc_compact_p_to_h(p, c->M);
}

void orc_random_shift(orc_cells_t *c, const double u[3]) /* cell_system.f90:153-163 */
{
    for (int k = 0; k < 3; ++k) c->origin[k] = u[k] - 1;
}

void orc_count_particles(const orc_cells_t *c, int64_t n, const double *pos,
                         int32_t *cell_count, int32_t *cell_start, int nthreads)
{
    /* cell_system.f90:86-129: per-thread histograms, reduction, serial scan */
    int nt = nthreads > 0 ? nthreads : g_threads;
    int32_t *tmp = (int32_t *)calloc((size_t)c->N * nt, sizeof(int32_t));
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int32_t *mine = tmp + (size_t)tid * c->N;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            int idx = orc_cell_key(c, pos + 3 * i);
            mine[idx] += 1;
        }
    }
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int i = 0; i < c->N; ++i) {
        int32_t s = 0;
        for (int t = 0; t < nt; ++t) s += tmp[(size_t)t * c->N + i];
        cell_count[i] = s;
    }
    free(tmp);
    cell_start[0] = 0;
    for (int i = 1; i < c->N; ++i) cell_start[i] = cell_start[i - 1] + cell_count[i - 1];
}

/* ======================= particle_system.f90 ==================================== */
void orc_sort(const orc_cells_t *c, const orc_particles_t *src, orc_particles_t *dst,
              int32_t *id_to_idx, int32_t *cell_count, int32_t *cell_start, int nthreads)
{
    /* particle_system.f90:304-355 (the species==0 test at :325 is a no-op: all
     * particles are sorted) */
    int nt = nthreads > 0 ? nthreads : g_threads;
    orc_count_particles(c, src->n, src->pos, cell_count, cell_start, nt);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = 0; i < src->n; ++i) {
        int idx = orc_cell_key(c, src->pos + 3 * i);
        int32_t start;
#pragma omp atomic capture
        start = cell_start[idx]++;
        for (int k = 0; k < 3; ++k) {
            dst->pos[3 * start + k] = src->pos[3 * i + k];
            dst->image[3 * start + k] = src->image[3 * i + k];
            dst->vel[3 * start + k] = src->vel[3 * i + k];
            dst->force[3 * start + k] = src->force[3 * i + k];
            dst->force_old[3 * start + k] = src->force_old[3 * i + k];
        }
        dst->id[start] = src->id[i];
        id_to_idx[src->id[i]] = start;
        dst->species[start] = src->species[i];
        dst->flags[start] = src->flags[i];
    }
    for (int i = 0; i < c->N; ++i) cell_start[i] -= cell_count[i]; /* :343 */
}

double orc_maximum_displacement(int64_t n, const double *pos, const double *pos_old,
                                const int32_t *species) /* particle_system.f90:357-378 */
{
    double rmax_sq = 0;
    ORC_PARALLEL_FOR_REDUCE(max : rmax_sq)
    for (int64_t i = 0; i < n; ++i) {
        if (species[i] >= 0) {
            double d0 = pos[3 * i] - pos_old[3 * i];
            double d1 = pos[3 * i + 1] - pos_old[3 * i + 1];
            double d2 = pos[3 * i + 2] - pos_old[3 * i + 2];
            double r_sq = (d0 * d0 + d1 * d1) + d2 * d2;
            if (r_sq > rmax_sq) rmax_sq = r_sq;
        }
    }
    return sqrt(rmax_sq);
}

double orc_cell_maximum_displacement(const orc_cells_t *c, const int32_t *cell_count,
                                     const int32_t *cell_start, const uint8_t *is_md,
                                     const double *pos, const double *pos_old,
                                     const int32_t *species, double delta_t)
{
    /* particle_system.f90:380-410 */
    double rmax_sq = 0;
    for (int i = 0; i < c->N; ++i) {
        if (!is_md[i]) continue;
        double local = 0;
        for (int j = cell_start[i]; j < cell_start[i] + cell_count[i]; ++j) {
            if (species[j] >= 0) {
                double d0 = pos[3 * j] - pos_old[3 * j];
                double d1 = pos[3 * j + 1] - pos_old[3 * j + 1];
                double d2 = pos[3 * j + 2] - pos_old[3 * j + 2];
                double r_sq = (d0 * d0 + d1 * d1) + d2 * d2;
                if (r_sq > local) local = r_sq;
            }
        }
        if (local > rmax_sq) rmax_sq = local;
    }
    double r = sqrt(rmax_sq);
    double b = c->max_v * delta_t;
    return r > b ? r : b;
}

/* ======================= mpcd.f90 ============================================== */
int orc_rand_sphere(const double *u, int n_avail, double n_out[3]) /* mpcd.f90:43-62 */
{
    int used = 0;
    double n1, n2, s;
    for (;;) {
        if (used + 2 > n_avail) return -1;
        n1 = u[used];
        n2 = u[used + 1];
        used += 2;
        n1 = 2 * n1 - 1;
        n2 = 2 * n2 - 1;
        s = n1 * n1 + n2 * n2;
        if (s < 1.0) break;
    }
    double alpha = 2.0 * sqrt(1.0 - s);
    n_out[0] = n1 * alpha;
    n_out[1] = n2 * alpha;
    n_out[2] = 1.0 - 2.0 * s;
    return used;
}

int64_t orc_srd_axes_from_stream(const orc_cells_t *c, const int32_t *cell_count,
                                 const int32_t *cell_start, const int32_t *species,
                                 const double *u, int64_t n_u, double *axes)
{
    /* the draw order of simple_mpcd_step at one thread (mpcd.f90:127-162):
     * cells in index order; cells with count<=1 or n_effective<=1 draw nothing */
    int64_t used = 0;
    memset(axes, 0, sizeof(double) * 3 * (size_t)c->N);
    for (int ci = 0; ci < c->N; ++ci) {
        if (cell_count[ci] <= 1) continue;
        int n_eff = 0;
        for (int i = cell_start[ci]; i < cell_start[ci] + cell_count[ci]; ++i)
            if (species[i] > 0) n_eff++;
        if (n_eff <= 1) continue;
        int k = orc_rand_sphere(u + used, (int)((n_u - used) > 1 << 30 ? 1 << 30 : (n_u - used)),
                                axes + 3 * ci);
        if (k < 0) return -1;
        used += k;
    }
    return used;
}

void orc_simple_mpcd_step(const orc_cells_t *c, const int32_t *cell_count,
                          const int32_t *cell_start, double *vel, const int32_t *species,
                          const double *axes, int alpha_given, double alpha)
{
    /* mpcd.f90:71-183, SRD branch (the thermostat branch is orc_mpcat_step) */
    double sin_a, cos_a;
    if (alpha_given) {
        sin_a = sin(alpha);
        cos_a = cos(alpha);
    } else {
        sin_a = 1;
        cos_a = 0;
    }
    ORC_PARALLEL_FOR
    for (int ci = 0; ci < c->N; ++ci) {
        if (cell_count[ci] <= 1) continue;
        int start = cell_start[ci], n = cell_count[ci];
        int n_eff = 0;
        double lv0 = 0, lv1 = 0, lv2 = 0;
        for (int i = start; i < start + n; ++i) {
            if (species[i] > 0) {
                lv0 = lv0 + vel[3 * i];
                lv1 = lv1 + vel[3 * i + 1];
                lv2 = lv2 + vel[3 * i + 2];
                n_eff++;
            }
        }
        if (n_eff <= 1) continue;
        lv0 = lv0 / n_eff;
        lv1 = lv1 / n_eff;
        lv2 = lv2 / n_eff;
        const double *v = axes + 3 * ci;
        double omc = 1 - cos_a;
        /* reshape(...,(/3,3/)) fills column-major (mpcd.f90:163-171): om[r][k] = omega(r+1,k+1) */
        double om[3][3];
        om[0][0] = cos_a + (v[0] * v[0]) * omc;
        om[1][0] = (v[0] * v[1]) * omc - v[2] * sin_a;
        om[2][0] = (v[0] * v[2]) * omc + v[1] * sin_a;
        om[0][1] = (v[1] * v[0]) * omc + v[2] * sin_a;
        om[1][1] = cos_a + (v[1] * v[1]) * omc;
        om[2][1] = (v[1] * v[2]) * omc - v[0] * sin_a;
        om[0][2] = (v[2] * v[0]) * omc - v[1] * sin_a;
        om[1][2] = (v[2] * v[1]) * omc + v[0] * sin_a;
        om[2][2] = cos_a + (v[2] * v[2]) * omc;
        for (int i = start; i < start + n; ++i) {
            double w0 = vel[3 * i] - lv0, w1 = vel[3 * i + 1] - lv1, w2 = vel[3 * i + 2] - lv2;
            for (int r = 0; r < 3; ++r) {
                double t = om[r][0] * w0;
                t = t + om[r][1] * w1;
                t = t + om[r][2] * w2;
                vel[3 * i + r] = (r == 0 ? lv0 : r == 1 ? lv1 : lv2) + t;
            }
        }
    }
}

double orc_compute_temperature(const orc_cells_t *c, const int32_t *cell_count,
                               const int32_t *cell_start, const double *vel,
                               const int32_t *species) /* mpcd.f90:342-407 */
{
    int count = 0;
    double te = 0;
    ORC_PARALLEL_FOR_REDUCE(+ : count, te)
    for (int ci = 0; ci < c->N; ++ci) {
        if (cell_count[ci] <= 1) continue;
        int start = cell_start[ci], n = cell_count[ci];
        count++; /* counted before the n_effective test (mpcd.f90:374) */
        int n_eff = 0;
        double lv0 = 0, lv1 = 0, lv2 = 0, kin = 0;
        for (int i = start; i < start + n; ++i) {
            if (species[i] > 0) {
                lv0 = lv0 + vel[3 * i];
                lv1 = lv1 + vel[3 * i + 1];
                lv2 = lv2 + vel[3 * i + 2];
                n_eff++;
            }
        }
        if (n_eff <= 1) continue;
        lv0 = lv0 / n; /* divides by n, not n_effective (mpcd.f90:388) */
        lv1 = lv1 / n;
        lv2 = lv2 / n;
        for (int i = start; i < start + n; ++i) {
            if (species[i] > 0) {
                double d0 = vel[3 * i] - lv0, d1 = vel[3 * i + 1] - lv1, d2 = vel[3 * i + 2] - lv2;
                kin = kin + ((d0 * d0 + d1 * d1) + d2 * d2);
            }
        }
        kin = kin / (double)(3 * (n_eff - 1));
        te = te + kin;
    }
    return te / count;
}

void orc_temperature_profile(const orc_cells_t *c, const int32_t *cell_count, const int32_t *cell_start,
                             const double *vel, const int32_t *species, double xmin, double dx, int n_bins,
                             double *data, int32_t *count)
{
    /* the tz profile of compute_temperature (mpcd.f90:395-398, profile_t%bin at
       common.f90:196-208): each cell's kinetic term, for cells with n_eff > 1, added at the
       height of the cell's centre origin(3) + (cell(3) + 1/2) a; one thread, cell order */
    for (int ci = 0; ci < c->N; ++ci) {
        if (cell_count[ci] <= 1) continue;
        int start = cell_start[ci], n = cell_count[ci];
        int n_eff = 0;
        double lv0 = 0, lv1 = 0, lv2 = 0, kin = 0;
        for (int i = start; i < start + n; ++i) {
            if (species[i] > 0) {
                lv0 = lv0 + vel[3 * i];
                lv1 = lv1 + vel[3 * i + 1];
                lv2 = lv2 + vel[3 * i + 2];
                n_eff++;
            }
        }
        if (n_eff <= 1) continue;
        lv0 = lv0 / n;
        lv1 = lv1 / n;
        lv2 = lv2 / n;
        for (int i = start; i < start + n; ++i) {
            if (species[i] > 0) {
                double d0 = vel[3 * i] - lv0, d1 = vel[3 * i + 1] - lv1, d2 = vel[3 * i + 2] - lv2;
                kin = kin + ((d0 * d0 + d1 * d1) + d2 * d2);
            }
        }
        kin = kin / (double)(3 * (n_eff - 1));
        int p[3];
        orc_compact_h_to_p(ci, c->M, p);
        double z = c->origin[2] + (p[2] + 0.5) * c->a;
        int idx = (int)floor((z - xmin) / dx);
        if (idx >= 0 && idx < n_bins) {
            data[idx] = data[idx] + kin;
            count[idx] = count[idx] + 1;
        }
    }
}

void orc_mpcd_stream_periodic(const orc_cells_t *c, int64_t n, double *pos,
                              const double *vel, int32_t *image, double dt)
{
    /* mpcd.f90:440-457 */
    ORC_PARALLEL_FOR
    for (int64_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            double x = pos[3 * i + k] + vel[3 * i + k] * dt;
            int jump = (int)floor(x / c->edges[k]);
            image[3 * i + k] += jump;
            pos[3 * i + k] = x - jump * c->edges[k];
        }
    }
}

void orc_bulk_reaction(const orc_cells_t *c, const int32_t *cell_count,
                       const int32_t *cell_start, const uint8_t *is_reac,
                       int32_t *species, int from, int to, double rate, double tau,
                       const double *draws)
{
    /* mpcd.f90:605-643 */
    ORC_PARALLEL_FOR
    for (int ci = 0; ci < c->N; ++ci) {
        if (cell_count[ci] <= 1 || !is_reac[ci]) continue;
        int start = cell_start[ci], n = cell_count[ci];
        double local_rate = 0;
        int pick = -1;
        for (int i = start; i < start + n; ++i) {
            if (species[i] == from) {
                local_rate = local_rate + 1;
                pick = i;
            }
        }
        local_rate = local_rate * rate;
        if (draws[ci] < (1 - exp(-local_rate * tau))) species[pick] = to;
    }
}

void orc_mpcat_step(const orc_cells_t *c, const int32_t *cell_count, const int32_t *cell_start,
                    double *vel, const int32_t *species, const double *normals, double T, int hydro)
{
    /* mpcd.f90:144-160, the thermostat branch of simple_mpcd_step: normals holds the
       three normals drawn for the particle in each slot (read for species > 0) */
    double t_factor = sqrt(T);
    ORC_PARALLEL_FOR
    for (int ci = 0; ci < c->N; ++ci) {
        if (cell_count[ci] <= 1) continue;
        int start = cell_start[ci], n = cell_count[ci];
        int n_eff = 0;
        double lv[3] = {0, 0, 0};
        for (int i = start; i < start + n; ++i) {
            if (species[i] > 0) {
                for (int k = 0; k < 3; ++k) lv[k] = lv[k] + vel[3 * i + k];
                n_eff++;
            }
        }
        if (n_eff <= 1) continue;
        for (int k = 0; k < 3; ++k) lv[k] = lv[k] / n_eff;
        double vv[3] = {0, 0, 0};
        for (int i = start; i < start + n; ++i) {
            if (species[i] <= 0) continue;
            for (int k = 0; k < 3; ++k) {
                vel[3 * i + k] = normals[3 * i + k] * t_factor;
                vv[k] = vv[k] + vel[3 * i + k];
            }
        }
        for (int k = 0; k < 3; ++k) vv[k] = lv[k] - vv[k] / n_eff;
        if (hydro)
            for (int i = start; i < start + n; ++i) {
                if (species[i] <= 0) continue;
                for (int k = 0; k < 3; ++k) vel[3 * i + k] = vel[3 * i + k] + vv[k];
            }
    }
}

void orc_rescale_cells(const orc_cells_t *c, const int32_t *cell_count, const int32_t *cell_start,
                       double *vel, const int32_t *species, double T)
{
    /* mpcd.f90:748-801 */
    ORC_PARALLEL_FOR
    for (int ci = 0; ci < c->N; ++ci) {
        if (cell_count[ci] <= 1) continue;
        int start = cell_start[ci], n = cell_count[ci];
        int n_eff = 0;
        double lv[3] = {0, 0, 0};
        for (int i = start; i < start + n; ++i) {
            if (species[i] > 0) {
                for (int k = 0; k < 3; ++k) lv[k] = lv[k] + vel[3 * i + k];
                n_eff++;
            }
        }
        if (n_eff <= 1) continue;
        for (int k = 0; k < 3; ++k) lv[k] = lv[k] / n_eff;
        double kin = 0;
        for (int i = start; i < start + n; ++i) {
            if (species[i] <= 0) continue;
            double d0 = vel[3 * i] - lv[0], d1 = vel[3 * i + 1] - lv[1], d2 = vel[3 * i + 2] - lv[2];
            kin = kin + ((d0 * d0 + d1 * d1) + d2 * d2);
        }
        double f = sqrt((double)(3 * (n_eff - 1)) * T / kin);
        for (int i = start; i < start + n; ++i) {
            if (species[i] <= 0) continue;
            for (int k = 0; k < 3; ++k) vel[3 * i + k] = lv[k] + f * (vel[3 * i + k] - lv[k]);
        }
    }
}

void orc_bulk_reaction_endothermic(const orc_cells_t *c, const int32_t *cell_count,
                                   const int32_t *cell_start, const uint8_t *is_reac, double *vel,
                                   int32_t *species, int from, int to, double rate, double tau,
                                   double delta_u, const double *draws)
{
    /* mpcd.f90:646-700; draws[cell] is read only by the cells that draw */
    ORC_PARALLEL_FOR
    for (int ci = 0; ci < c->N; ++ci) {
        if (cell_count[ci] <= 1 || !is_reac[ci]) continue;
        int start = cell_start[ci], n = cell_count[ci];
        double local_rate = 0, com[3] = {0, 0, 0};
        int pick = -1;
        for (int i = start; i < start + n; ++i) {
            if (species[i] == from) {
                local_rate = local_rate + 1;
                pick = i;
            }
            for (int k = 0; k < 3; ++k) com[k] = com[k] + vel[3 * i + k];
        }
        for (int k = 0; k < 3; ++k) com[k] = com[k] / n;
        double kin = 0;
        for (int i = start; i < start + n; ++i) {
            double d0 = vel[3 * i] - com[0], d1 = vel[3 * i + 1] - com[1], d2 = vel[3 * i + 2] - com[2];
            kin = kin + ((d0 * d0 + d1 * d1) + d2 * d2);
        }
        kin = kin / 2;
        if (!(local_rate > 0 && kin > delta_u)) continue;
        local_rate = local_rate * rate;
        if (draws[ci] < (1 - exp(-local_rate * tau))) {
            species[pick] = to;
            double f = sqrt((kin - delta_u) / kin);
            for (int i = start; i < start + n; ++i)
                for (int k = 0; k < 3; ++k) vel[3 * i + k] = com[k] + f * (vel[3 * i + k] - com[k]);
        }
    }
}

void orc_wall_mpcd_step(const orc_cells_t *c, const int32_t *cell_count, const int32_t *cell_start,
                        double *vel, const int32_t *id, int walls_given, const double *wall_temperature,
                        const int32_t *wall_n, int thermostat, double bulk_temperature, int keep_cell_v,
                        int alpha_given, double alpha, const double *axes, const double *vnorm,
                        const double *pnorm)
{
    /* mpcd.f90:193-337.  axes and vnorm: 3 per cell; pnorm: 3 per particle slot */
    (void)id;
    double sin_a = alpha_given ? sin(alpha) : 1, cos_a = alpha_given ? cos(alpha) : 0;
    double t_factor = thermostat ? sqrt(bulk_temperature) : 0;
    ORC_PARALLEL_FOR
    for (int ci = 0; ci < c->N; ++ci) {
        if (cell_count[ci] <= 1) continue;
        int start = cell_start[ci], n = cell_count[ci];
        int cell[3];
        orc_compact_h_to_p(ci, c->M, cell);
        int w = -1;
        if (walls_given && cell[2] == 0)
            w = 0;
        else if (walls_given && cell[2] == c->L[2] - 1)
            w = 1;
        int n_virtual = 0;
        double vv[3] = {0, 0, 0};
        if (w >= 0 && n < wall_n[w]) {
            n_virtual = wall_n[w] - n;
            double f = sqrt(n_virtual * wall_temperature[w]);
            for (int k = 0; k < 3; ++k) vv[k] = vnorm[3 * ci + k] * f;
        }
        double lv[3] = {0, 0, 0};
        for (int i = start; i < start + n; ++i)
            for (int k = 0; k < 3; ++k) lv[k] = lv[k] + vel[3 * i + k];
        for (int k = 0; k < 3; ++k) lv[k] = n_virtual > 0 ? (lv[k] + vv[k]) / wall_n[w] : lv[k] / n;
        if (thermostat) {
            double sv[3] = {0, 0, 0};
            for (int i = start; i < start + n; ++i)
                for (int k = 0; k < 3; ++k) {
                    vel[3 * i + k] = pnorm[3 * i + k] * t_factor;
                    sv[k] = sv[k] + vel[3 * i + k];
                }
            if (keep_cell_v) {
                for (int k = 0; k < 3; ++k) sv[k] = lv[k] - sv[k] / (double)n;
                for (int i = start; i < start + n; ++i)
                    for (int k = 0; k < 3; ++k) vel[3 * i + k] = vel[3 * i + k] + sv[k];
            }
        } else {
            const double *v = axes + 3 * ci;
            double omc = 1 - cos_a;
            double om[3][3];
            om[0][0] = cos_a + (v[0] * v[0]) * omc;
            om[1][0] = (v[0] * v[1]) * omc - v[2] * sin_a;
            om[2][0] = (v[0] * v[2]) * omc + v[1] * sin_a;
            om[0][1] = (v[1] * v[0]) * omc + v[2] * sin_a;
            om[1][1] = cos_a + (v[1] * v[1]) * omc;
            om[2][1] = (v[1] * v[2]) * omc - v[0] * sin_a;
            om[0][2] = (v[2] * v[0]) * omc - v[1] * sin_a;
            om[1][2] = (v[2] * v[1]) * omc + v[0] * sin_a;
            om[2][2] = cos_a + (v[2] * v[2]) * omc;
            for (int i = start; i < start + n; ++i) {
                double d[3], r[3];
                for (int k = 0; k < 3; ++k) d[k] = vel[3 * i + k] - lv[k];
                for (int rr = 0; rr < 3; ++rr) r[rr] = (om[rr][0] * d[0] + om[rr][1] * d[1]) + om[rr][2] * d[2];
                for (int k = 0; k < 3; ++k) vel[3 * i + k] = lv[k] + r[k];
            }
        }
    }
}

/* yzwall_collision (mpcd.f90:513-587); the alternating wall loop is capped at 64 */
static void orc_yzwall_collision(const double x0[3], const double v0[3], double x[3], double v[3],
                                 const double L[3], double t, const int bc[3], double g)
{
    double gv[3] = {g, 0, 0};
    int im[3] = {0, 0, 0};
    int cd = -1;
    double tcol = 1.7976931348623157e308;
    for (int i = 1; i < 3; ++i) {
        double tt = v0[i] > 0 ? (L[i] - x0[i]) / v0[i] : -x0[i] / v0[i];
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
    for (int it = -1; it < 64; ++it) {
        if (it >= 0) {
            cd = cd == 1 ? 2 : 1;
            double tc = v[cd] > 0 ? (L[cd] - x[cd]) / v[cd] : -x[cd] / v[cd];
            if (tc < 0 || tc > trem) break;
            for (int k = 0; k < 3; ++k) {
                x[k] = (x[k] + v[k] * tc) + (gv[k] * (tc * tc)) / 2;
                v[k] = v[k] + gv[k] * tc;
            }
            trem = trem - tc;
        }
        if (bc[cd] == ORC_BOUNCE_BACK_BC) {
            for (int k = 0; k < 3; ++k) v[k] = -v[k];
        } else if (bc[cd] == ORC_SPECULAR_BC) {
            v[cd] = -v[cd];
        } else if (bc[cd] == ORC_PERIODIC_BC) {
            int jump = (int)floor(x[cd] / L[cd]);
            x[cd] = x[cd] - jump * L[cd];
            im[cd] = im[cd] + jump;
        }
    }
    double tc = trem;
    for (int k = 0; k < 3; ++k) {
        x[k] = ((x[k] + v[k] * tc) + (gv[k] * (tc * tc)) / 2) + im[k] * L[k];
        v[k] = v[k] + gv[k] * tc;
    }
}

void orc_mpcd_stream_xforce_yzwall(const orc_cells_t *c, int64_t n, double *pos, double *vel,
                                   int32_t *flags, double dt, double g)
{
    /* mpcd.f90:459-511 */
    const double *L = c->edges;
    for (int64_t i = 0; i < n; ++i) {
        double op[3], ov[3], np[3], nv[3], gv[3] = {g, 0, 0};
        for (int k = 0; k < 3; ++k) {
            op[k] = pos[3 * i + k];
            ov[k] = vel[3 * i + k];
            np[k] = (op[k] + ov[k] * dt) + (gv[k] * (dt * dt)) / 2;
            nv[k] = ov[k];
        }
        int y_out = np[1] < 0 || np[1] > L[1], z_out = np[2] < 0 || np[2] > L[2];
        if ((c->bc[1] == ORC_PERIODIC_BC && z_out) || (c->bc[1] != ORC_PERIODIC_BC && (y_out || z_out))) {
            orc_yzwall_collision(op, ov, np, nv, L, dt, c->bc, g);
            flags[i] |= 1 << ORC_WALL_BIT;
            for (int k = 0; k < 3; ++k) vel[3 * i + k] = nv[k];
        }
        for (int k = 0; k < 3; ++k) pos[3 * i + k] = np[k];
    }
}

void orc_mpcd_stream_nogravity_zwall(const orc_cells_t *c, int64_t n, double *pos,
                                     double *vel, const double *force, int32_t *flags,
                                     double dt)
{
    /* mpcd.f90:702-745 */
    const double *L = c->edges;
    double dt2 = dt * dt;
    for (int64_t i = 0; i < n; ++i) {
        double op[3], ov[3], np[3], nv[3];
        for (int k = 0; k < 3; ++k) {
            op[k] = pos[3 * i + k];
            ov[k] = vel[3 * i + k];
            np[k] = (op[k] + ov[k] * dt) + (force[3 * i + k] * dt2) / 2;
            nv[k] = ov[k];
        }
        int bounce = 0;
        double t_c = 0;
        if (np[2] < 0) {
            t_c = -op[2] / ov[2];
            bounce = 1;
        } else if (np[2] > L[2]) {
            t_c = (L[2] - op[2]) / ov[2];
            bounce = 1;
        }
        if (bounce) {
            for (int k = 0; k < 3; ++k) {
                np[k] = np[k] - (ov[k] * 2) * (dt - t_c);
                nv[k] = -ov[k];
            }
            flags[i] |= 1 << ORC_WALL_BIT;
        }
        for (int k = 0; k < 3; ++k) {
            pos[3 * i + k] = np[k];
            vel[3 * i + k] = nv[k];
        }
    }
}

/* ======================= md.f90 ================================================ */
void orc_md_pos(int64_t n, double *pos, const double *vel, const double *force, double dt)
{
    /* md.f90:61-77 */
    double dt_sq = (dt * dt) / 2;
    ORC_PARALLEL_FOR
    for (int64_t i = 0; i < 3 * n; ++i) pos[i] = (pos[i] + dt * vel[i]) + dt_sq * force[i];
}

void orc_md_vel(int64_t n, double *vel, const double *force, const double *force_old,
                int32_t *flags, double dt)
{
    /* md.f90:219-237 */
    ORC_PARALLEL_FOR
    for (int64_t i = 0; i < n; ++i) {
        if (!(flags[i] & (1 << ORC_WALL_BIT))) {
            for (int k = 0; k < 3; ++k)
                vel[3 * i + k] = vel[3 * i + k] + (dt * (force[3 * i + k] + force_old[3 * i + k])) / 2;
        } else {
            flags[i] &= ~(1 << ORC_WALL_BIT);
        }
    }
}

void orc_apply_pbc(int64_t n, double *pos, int32_t *image, const double edges[3])
{
    /* md.f90:202-217 */
    ORC_PARALLEL_FOR
    for (int64_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            int jump = (int)floor(pos[3 * i + k] / edges[k]);
            image[3 * i + k] += jump;
            pos[3 * i + k] = pos[3 * i + k] - jump * edges[k];
        }
    }
}

void orc_cell_md_pos(const orc_cells_t *c, const int32_t *cell_count,
                     const int32_t *cell_start, const uint8_t *is_md, double *pos,
                     const double *vel, const double *force, double dt, int md_flag)
{
    /* md.f90:79-102 */
    double dt_sq = (dt * dt) / 2;
    for (int ci = 0; ci < c->N; ++ci) {
        if ((is_md[ci] != 0) != (md_flag != 0)) continue;
        for (int j = cell_start[ci]; j < cell_start[ci] + cell_count[ci]; ++j)
            for (int k = 0; k < 3; ++k)
                pos[3 * j + k] = (pos[3 * j + k] + dt * vel[3 * j + k]) + dt_sq * force[3 * j + k];
    }
}

void orc_cell_md_vel(const orc_cells_t *c, const int32_t *cell_count,
                     const int32_t *cell_start, const uint8_t *is_md, double *vel,
                     const double *force, const double *force_old, int32_t *flags,
                     double dt, int md_flag)
{
    /* md.f90:239-263 */
    for (int ci = 0; ci < c->N; ++ci) {
        if ((is_md[ci] != 0) != (md_flag != 0)) continue;
        for (int j = cell_start[ci]; j < cell_start[ci] + cell_count[ci]; ++j) {
            if (!(flags[j] & (1 << ORC_WALL_BIT))) {
                for (int k = 0; k < 3; ++k)
                    vel[3 * j + k] = vel[3 * j + k] + (dt * (force[3 * j + k] + force_old[3 * j + k])) / 2;
            } else {
                flags[j] &= ~(1 << ORC_WALL_BIT);
            }
        }
    }
}

void orc_cell_md_pos_wall(const orc_cells_t *c, const int32_t *cell_count,
                          const int32_t *cell_start, const uint8_t *is_md, double *pos,
                          double *vel, const double *force, double dt, int md_flag, int axis)
{
    /* md.f90:104-151 (axis 1, y) and md.f90:153-200 (axis 2, z) */
    double dt_sq = (dt * dt) / 2;
    double Lw = c->edges[axis];
    int bc = c->bc[axis];
    for (int ci = 0; ci < c->N; ++ci) {
        if ((is_md[ci] != 0) != (md_flag != 0)) continue;
        for (int j = cell_start[ci]; j < cell_start[ci] + cell_count[ci]; ++j) {
            double v[3], x[3], overtime = 0;
            for (int k = 0; k < 3; ++k) v[k] = vel[3 * j + k];
            for (int k = 0; k < 3; ++k) x[k] = (pos[3 * j + k] + dt * v[k]) + dt_sq * force[3 * j + k];
            int collide = 0;
            if (x[axis] < 0.0) {
                collide = 1;
                overtime = x[axis] / v[axis];
            } else if (x[axis] > Lw) {
                collide = 1;
                overtime = (x[axis] - Lw) / v[axis];
            }
            if (collide) {
                if (bc == ORC_BOUNCE_BACK_BC) {
                    for (int k = 0; k < 3; ++k) v[k] = -v[k];
                    for (int k = 0; k < 3; ++k) x[k] = x[k] + (2 * overtime) * v[k];
                } else if (bc == ORC_SPECULAR_BC) {
                    v[axis] = -v[axis];
                    x[axis] = x[axis] + (2 * overtime) * v[axis];
                }
                for (int k = 0; k < 3; ++k) vel[3 * j + k] = v[k];
            }
            for (int k = 0; k < 3; ++k) pos[3 * j + k] = x[k];
        }
    }
}

void orc_rel_pos(const double x[3], const double y[3], const double L[3], double r[3])
{
    /* common.f90:162-178 */
    for (int k = 0; k < 3; ++k) {
        r[k] = x[k] - y[k];
        if (r[k] < -0.5 * L[k])
            r[k] = r[k] + L[k];
        else if (r[k] > 0.5 * L[k])
            r[k] = r[k] - L[k];
    }
}

static double dot3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

void orc_rattle_dimer_pos(double *pos, double *vel, const double *pos_rattle,
                          const int32_t *species, const double *mass, double d, double dt,
                          const double edges[3])
{
    /* md.f90:265-296 */
    double r[3], s[3];
    orc_rel_pos(pos_rattle, pos_rattle + 3, edges, r);
    orc_rel_pos(pos, pos + 3, edges, s);
    double m1 = mass[species[0] - 1], m2 = mass[species[1] - 1];
    double inv_mass = 1 / m1 + 1 / m2;
    double rsq = dot3(r, r), ssq = dot3(s, s), rs = dot3(r, s);
    double g = rs - sqrt(rs * rs - rsq * (ssq - d * d));
    g = g / ((dt * inv_mass) * rsq);
    for (int k = 0; k < 3; ++k) {
        pos[k] = pos[k] - ((g * dt) * r[k]) / m1;
        pos[3 + k] = pos[3 + k] + ((g * dt) * r[k]) / m2;
        vel[k] = vel[k] - (g * r[k]) / m1;
        vel[3 + k] = vel[3 + k] + (g * r[k]) / m2;
    }
}

void orc_rattle_dimer_vel(const double *pos, double *vel, const int32_t *species,
                          const double *mass, double d, double dt, const double edges[3])
{
    /* md.f90:298-319 */
    (void)dt;
    double m1 = mass[species[0] - 1], m2 = mass[species[1] - 1];
    double inv_mass = 1 / m1 + 1 / m2;
    double s[3], dv[3];
    orc_rel_pos(pos, pos + 3, edges, s);
    for (int k = 0; k < 3; ++k) dv[k] = vel[k] - vel[3 + k];
    double kk = dot3(dv, s) / ((d * d) * inv_mass);
    for (int k = 0; k < 3; ++k) {
        vel[k] = vel[k] - (kk * s[k]) / m1;
        vel[3 + k] = vel[3 + k] + (kk * s[k]) / m2;
    }
}

void orc_colloid_md_pos(int64_t n, double *pos, const double *vel, const double *force,
                        const double *mass_by_particle, double dt)
{
    /* programs/single_dimer_pbc.f90:312-315 */
    for (int64_t k = 0; k < n; ++k)
        for (int c = 0; c < 3; ++c)
            pos[3 * k + c] = (pos[3 * k + c] + dt * vel[3 * k + c]) +
                             ((dt * dt) * force[3 * k + c]) / (2 * mass_by_particle[k]);
}

void orc_colloid_md_vel(int64_t n, double *vel, const double *force,
                        const double *force_old, const double *mass_by_particle, double dt)
{
    /* programs/single_dimer_pbc.f90:356-359 */
    for (int64_t k = 0; k < n; ++k)
        for (int c = 0; c < 3; ++c)
            vel[3 * k + c] = vel[3 * k + c] +
                             (dt * (force[3 * k + c] + force_old[3 * k + c])) / (2 * mass_by_particle[k]);
}

/* ---- linked bodies (md.f90:321-480, 516-544) ---------------------------------
 * Positions are (n, 3) rows; links are zero-based pairs; mass[s - 1] is the mass of
 * species s.  Each function returns the reference's loop counter at exit (the sweep
 * that met the precision, rattle_max + 1 when none did). */
static void body_unfold(int64_t n, double *x, const double edges[3], int32_t *im)
{
    /* md.f90:343-347: every particle brought next to particle 1 (index 0 here) */
    for (int64_t i = 1; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            int j = (int)floor(((x[3 * i + k] - x[k]) / edges[k]) + 0.5);
            if (im) im[3 * i + k] = j;
            x[3 * i + k] = x[3 * i + k] - j * edges[k];
        }
}

static void body_refold(int64_t n, double *x, const double edges[3], const int32_t *im)
{
    /* md.f90:400-402 */
    for (int64_t i = 1; i < n; ++i)
        for (int k = 0; k < 3; ++k) x[3 * i + k] = x[3 * i + k] + im[3 * i + k] * edges[k];
}

int orc_rattle_body_pos(int64_t n, double *pos, double *vel, double *pos_rattle,
                        const int32_t *species, const double *mass, const int32_t *links,
                        const double *distances, int n_links, double dt, const double edges[3],
                        double precision)
{
    /* md.f90:321-408 */
    int32_t *im = (int32_t *)calloc((size_t)(3 * n), sizeof(int32_t));
    body_unfold(n, pos, edges, im);
    body_unfold(n, pos_rattle, edges, NULL);
    int sweep;
    for (sweep = 1; sweep <= 1000; ++sweep) {
        for (int l = 0; l < n_links; ++l) {
            const int a = links[2 * l], b = links[2 * l + 1];
            const double d = distances[l];
            const double ia = 1 / mass[species[a] - 1], ib = 1 / mass[species[b] - 1];
            double r[3], s[3];
            for (int k = 0; k < 3; ++k) {
                r[k] = pos_rattle[3 * a + k] - pos_rattle[3 * b + k];
                s[k] = pos[3 * a + k] - pos[3 * b + k];
            }
            const double rsq = dot3(r, r), ssq = dot3(s, s), rs = dot3(r, s);
            double g = rs - sqrt(rs * rs - rsq * (ssq - d * d));
            g = g / ((dt * (ia + ib)) * rsq);
            for (int k = 0; k < 3; ++k) {
                pos[3 * a + k] -= ((g * dt) * r[k]) * ia;
                pos[3 * b + k] += ((g * dt) * r[k]) * ib;
                vel[3 * a + k] -= (g * r[k]) * ia;
                vel[3 * b + k] += (g * r[k]) * ib;
            }
        }
        double worst = 0;
        for (int l = 0; l < n_links; ++l) {
            const int a = links[2 * l], b = links[2 * l + 1];
            double r[3];
            for (int k = 0; k < 3; ++k) r[k] = pos[3 * a + k] - pos[3 * b + k];
            const double e = fabs(sqrt(dot3(r, r)) - distances[l]);
            if (e > worst) worst = e;
        }
        if (worst < precision) break;
    }
    body_refold(n, pos, edges, im);
    free(im);
    return sweep;
}

static double link_k(const double *pos, const double *vel, int a, int b, double d, double inv_mass)
{
    double s[3], dv[3];
    for (int k = 0; k < 3; ++k) {
        s[k] = pos[3 * a + k] - pos[3 * b + k];
        dv[k] = vel[3 * a + k] - vel[3 * b + k];
    }
    return dot3(dv, s) / ((d * d) * inv_mass);
}

int orc_rattle_body_vel(int64_t n, double *pos, double *vel, const int32_t *species,
                        const double *mass, const int32_t *links, const double *distances,
                        int n_links, double dt, const double edges[3], double precision)
{
    /* md.f90:410-480 */
    (void)dt;
    int32_t *im = (int32_t *)calloc((size_t)(3 * n), sizeof(int32_t));
    body_unfold(n, pos, edges, im);
    int sweep;
    for (sweep = 1; sweep <= 1000; ++sweep) {
        for (int l = 0; l < n_links; ++l) {
            const int a = links[2 * l], b = links[2 * l + 1];
            const double ia = 1 / mass[species[a] - 1], ib = 1 / mass[species[b] - 1];
            const double kk = link_k(pos, vel, a, b, distances[l], ia + ib);
            for (int k = 0; k < 3; ++k) {
                const double s = pos[3 * a + k] - pos[3 * b + k];
                vel[3 * a + k] -= (kk * s) * ia;
                vel[3 * b + k] += (kk * s) * ib;
            }
        }
        double worst = 0;
        for (int l = 0; l < n_links; ++l) {
            const int a = links[2 * l], b = links[2 * l + 1];
            const double ia = 1 / mass[species[a] - 1], ib = 1 / mass[species[b] - 1];
            const double e = fabs(link_k(pos, vel, a, b, distances[l], ia + ib));
            if (e > worst) worst = e;
        }
        if (worst < precision) break;
    }
    body_refold(n, pos, edges, im);
    free(im);
    return sweep;
}

double orc_elastic_network(const double *pos, double *force, const int32_t *links,
                           const double *distances, int n_links, double k, const double edges[3])
{
    /* md.f90:516-544 */
    double e = 0;
    for (int l = 0; l < n_links; ++l) {
        const int a = links[2 * l], b = links[2 * l + 1];
        double x[3];
        orc_rel_pos(pos + 3 * a, pos + 3 * b, edges, x);
        const double r = sqrt(dot3(x, x)), r0 = distances[l];
        e = e + (k * ((r - r0) * (r - r0))) / 2;
        for (int c = 0; c < 3; ++c) {
            const double f = ((-k) * (r - r0)) * x[c] / r;
            force[3 * a + c] += f;
            force[3 * b + c] -= f;
        }
    }
    return e;
}

/* ---- rigid body (md.f90:547-700) ----------------------------------------------
 * Quaternions (s, x, y, z): the reference takes qmul/qconj/qrot/qnormalize from an
 * external quaternion module; these are the textbook Hamilton product and the
 * rotation v -> q (0, v) q*. */
static void oq_mul(const double a[4], const double b[4], double r[4])
{
    double t[4];
    t[0] = a[0] * b[0] - ((a[1] * b[1] + a[2] * b[2]) + a[3] * b[3]);
    t[1] = (a[0] * b[1] + b[0] * a[1]) + (a[2] * b[3] - a[3] * b[2]);
    t[2] = (a[0] * b[2] + b[0] * a[2]) + (a[3] * b[1] - a[1] * b[3]);
    t[3] = (a[0] * b[3] + b[0] * a[3]) + (a[1] * b[2] - a[2] * b[1]);
    memcpy(r, t, sizeof t);
}
static void oq_conj(const double q[4], double r[4])
{
    r[0] = q[0];
    for (int k = 1; k < 4; ++k) r[k] = -q[k];
}
static void oq_rot(const double q[4], const double v[3], double r[3])
{
    double p[4] = {0, v[0], v[1], v[2]}, c[4], t[4];
    oq_conj(q, c);
    oq_mul(q, p, t);
    oq_mul(t, c, t);
    for (int k = 0; k < 3; ++k) r[k] = t[k + 1];
}
static double oq_norm(const double q[4]) { return sqrt(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]); }
static void oq_normalize(double q[4])
{
    double n = oq_norm(q);
    for (int k = 0; k < 4; ++k) q[k] = q[k] / n;
}
static void ocross(const double a[3], const double b[3], double r[3])
{
    double t[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    memcpy(r, t, sizeof t);
}

void orc_rigid_body_init(orc_body_t *b, const double *pos, const double *vel, const int32_t *image,
                         const int32_t *species, const double *mass, const double edges[3],
                         int i_start, int i_stop)
{
    /* md.f90:547-588 */
    b->i_start = i_start;
    b->i_stop = i_stop;
    b->mass = 0;
    for (int k = 0; k < 3; ++k) b->pos[k] = b->vel[k] = 0;
    for (int i = i_start; i <= i_stop; ++i) {
        const double m = mass[species[i] - 1];
        b->mass = b->mass + m;
        for (int k = 0; k < 3; ++k) {
            b->pos[k] = b->pos[k] + m * (pos[3 * i + k] + image[3 * i + k] * edges[k]);
            b->vel[k] = b->vel[k] + m * vel[3 * i + k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        b->pos[k] = b->pos[k] / b->mass;
        b->vel[k] = b->vel[k] / b->mass;
        b->I_body[k] = 0;
    }
    for (int i = i_start; i <= i_stop; ++i) {
        const double m = mass[species[i] - 1];
        double *x = b->pos_body + 3 * (i - i_start);
        for (int k = 0; k < 3; ++k) x[k] = (pos[3 * i + k] + image[3 * i + k] * edges[k]) - b->pos[k];
        const double x2 = (x[0] * x[0] + x[1] * x[1]) + x[2] * x[2];
        for (int k = 0; k < 3; ++k) b->I_body[k] = b->I_body[k] + m * (x2 - x[k] * x[k]);
    }
    for (int k = 0; k < 3; ++k) b->force[k] = b->torque[k] = b->L[k] = b->omega_body[k] = 0;
    b->q[0] = 1;
    b->q[1] = b->q[2] = b->q[3] = 0;
}

void orc_rigid_body_compute_force_torque(orc_body_t *b, const double *force)
{
    /* md.f90:591-607 */
    for (int k = 0; k < 3; ++k) b->force[k] = b->torque[k] = 0;
    for (int i = b->i_start; i <= b->i_stop; ++i) {
        const double *f = force + 3 * i;
        double x[3], t[3];
        oq_rot(b->q, b->pos_body + 3 * (i - b->i_start), x);
        ocross(x, f, t);
        for (int k = 0; k < 3; ++k) {
            b->force[k] = b->force[k] + f[k];
            b->torque[k] = b->torque[k] + t[k];
        }
    }
}

int orc_rigid_body_vv1(orc_body_t *b, double *pos, double dt, double threshold, const double edges[3])
{
    /* md.f90:613-673 (the iteration capped at 1000 rounds) */
    double I_inv[3], L[3], Lb[3], tb[3], ob[3], Lbd[3], c[3], q[4], qc[4], qd[4], w[4] = {0, 0, 0, 0};
    for (int k = 0; k < 3; ++k) I_inv[k] = 1 / b->I_body[k];
    memcpy(L, b->L, sizeof L);
    memcpy(q, b->q, sizeof q);
    oq_conj(q, qc);
    oq_rot(qc, L, Lb);
    oq_rot(qc, b->torque, tb);
    for (int k = 0; k < 3; ++k) ob[k] = Lb[k] * I_inv[k];
    ocross(ob, Lb, c);
    for (int k = 0; k < 3; ++k) Lbd[k] = tb[k] - c[k];
    for (int k = 0; k < 3; ++k) {
        Lb[k] = Lb[k] + (dt * Lbd[k]) / 2;
        ob[k] = Lb[k] * I_inv[k];
        w[k + 1] = ob[k];
    }
    oq_mul(q, w, qd);
    for (int k = 0; k < 4; ++k) {
        qd[k] = qd[k] / 2;
        q[k] = q[k] + (qd[k] * dt) / 2;
    }
    oq_normalize(q);
    for (int k = 0; k < 3; ++k) L[k] = b->L[k] + (b->torque[k] * dt) / 2;
    double q_old[4];
    memcpy(q_old, q, sizeof q);
    int round;
    for (round = 1; round <= 1000; ++round) {
        oq_conj(q, qc);
        oq_rot(qc, L, Lb);
        for (int k = 0; k < 3; ++k) w[k + 1] = I_inv[k] * Lb[k];
        oq_mul(q, w, qd);
        for (int k = 0; k < 4; ++k) {
            qd[k] = qd[k] / 2;
            q[k] = b->q[k] + (qd[k] * dt) / 2;
        }
        oq_normalize(q);
        double dq[4];
        for (int k = 0; k < 4; ++k) dq[k] = q[k] - q_old[k];
        if (oq_norm(dq) < threshold) break;
        memcpy(q_old, q, sizeof q);
    }
    for (int k = 0; k < 4; ++k) b->q[k] = b->q[k] + dt * qd[k];
    oq_normalize(b->q);
    for (int k = 0; k < 3; ++k)
        b->pos[k] = (b->pos[k] + b->vel[k] * dt) + (b->force[k] * (dt * dt)) / (2 * b->mass);
    for (int i = b->i_start; i <= b->i_stop; ++i) {
        double x[3];
        oq_rot(b->q, b->pos_body + 3 * (i - b->i_start), x);
        for (int k = 0; k < 3; ++k) {
            const double nx = b->pos[k] + x[k];
            const double delta = (nx - pos[3 * i + k]) + edges[k] / 2;
            pos[3 * i + k] = nx - floor(delta / edges[k]) * edges[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        b->L[k] = b->L[k] + (b->torque[k] * dt) / 2;
        b->vel[k] = b->vel[k] + (b->force[k] * dt) / (2 * b->mass);
    }
    return round;
}

void orc_rigid_body_vv2(orc_body_t *b, double *vel, double dt)
{
    /* md.f90:679-700 */
    double qc[4], Lb[3], om[3];
    for (int k = 0; k < 3; ++k) b->L[k] = b->L[k] + (b->torque[k] * dt) / 2;
    oq_conj(b->q, qc);
    oq_rot(qc, b->L, Lb);
    for (int k = 0; k < 3; ++k) b->omega_body[k] = Lb[k] / b->I_body[k];
    oq_rot(b->q, b->omega_body, om);
    for (int k = 0; k < 3; ++k) b->vel[k] = b->vel[k] + (b->force[k] * dt) / (2 * b->mass);
    for (int i = b->i_start; i <= b->i_stop; ++i) {
        double x[3], c[3];
        oq_rot(b->q, b->pos_body + 3 * (i - b->i_start), x);
        ocross(om, x, c);
        for (int k = 0; k < 3; ++k) vel[3 * i + k] = b->vel[k] + c[k];
    }
}

/* ======================= interaction.f90 ======================================= */
/* amdflang evaluates x**6 as ((((x*x)*x)*x)*x)*x: checked bit for bit against the
 * reference's interaction module built from source (tests/golden/interaction_ref.npz) */
static double pow6(double x) { return ((((x * x) * x) * x) * x) * x; }
static double pow3(double x) { return x * (x * x); }

void orc_lj_force(const double d[3], double r_sq, double eps, double sigma, double f[3])
{
    /* interaction.f90:83-96 */
    double sig6_o_r6 = pow6(sigma) / pow3(r_sq);
    double force_over_r = (((24.0 * eps) * sig6_o_r6) / r_sq) * (2.0 * sig6_o_r6 - 1.0);
    for (int k = 0; k < 3; ++k) f[k] = force_over_r * d[k];
}

double orc_lj_energy(double r_sq, double eps, double sigma, double cut_energy)
{
    /* interaction.f90:126-136 */
    double sig6_o_r6 = pow6(sigma) / pow3(r_sq);
    return (4.0 * eps) * (sig6_o_r6 * sig6_o_r6 - sig6_o_r6) - cut_energy;
}

int orc_lj_params_init(orc_lj_t *lj, int n1, int n2, const double *eps,
                       const double *sigma, const double *cut)
{
    /* interaction.f90:38-80; arrays are column-major (n1,n2) */
    if (n1 * n2 > 16 || n1 < 1 || n2 < 1) return -1;
    memset(lj, 0, sizeof(*lj));
    lj->n1 = n1;
    lj->n2 = n2;
    for (int i = 0; i < n1 * n2; ++i) {
        lj->eps[i] = eps[i];
        lj->sigma[i] = sigma[i];
        lj->cut[i] = cut[i];
        lj->cut_sq[i] = cut[i] * cut[i];
        lj->cut_energy[i] = orc_lj_energy(lj->cut_sq[i], eps[i], sigma[i], sigma[i] * 0);
    }
    return 0;
}

/* ======================= neighbor_list.f90 ====================================== */
static double closest(int i, int j, int k) /* neighbor_list.f90:310-322 */
{
    int v[3] = {i, j, k};
    int dist_sq = 0;
    for (int q = 0; q < 3; ++q) {
        int t = abs(v[q]) - 1;
        if (t < 0) t = 0;
        dist_sq += t * t;
    }
    return sqrt((double)dist_sq);
}

int orc_make_stencil(double a, double cut, int32_t *st) /* neighbor_list.f90:271-308 */
{
    int max_i = (int)floor(cut / a) + 1;
    int count = 0;
    for (int i = -max_i; i <= max_i; ++i)
        for (int j = -max_i; j <= max_i; ++j)
            for (int k = -max_i; k <= max_i; ++k)
                if (closest(i, j, k) < cut) {
                    if (st) {
                        st[3 * count] = i;
                        st[3 * count + 1] = j;
                        st[3 * count + 2] = k;
                    }
                    count++;
                }
    return count;
}

int orc_update_list(const orc_cells_t *c, const int32_t *cell_count,
                    const int32_t *cell_start, int64_t n1, const double *pos1,
                    const int32_t *species1, const double *pos2, const int32_t *species2,
                    double radius, const int32_t *stencil, int n_stencil, const orc_lj_t *lj,
                    int32_t nmax, int32_t *list, int32_t *nlist, uint8_t *is_md,
                    uint8_t *is_reac)
{
    /* neighbor_list.f90:76-157 */
    double L[3];
    for (int k = 0; k < 3; ++k) L[k] = (double)c->L[k] * c->a;
    double radiussq = radius * radius;
    for (int i = 0; i < c->N; ++i) {
        is_md[i] = 0;
        is_reac[i] = 1;
    }
    for (int64_t i = 0; i < n1; ++i) {
        const double *x = pos1 + 3 * i;
        int si = species1[i];
        int cell[3];
        orc_cartesian_indices(c, x, cell);
        int list_idx = 0;
        for (int j = 0; j < n_stencil; ++j) {
            int ac[3];
            for (int k = 0; k < 3; ++k) {
                int v = (cell[k] + stencil[3 * j + k]) % c->L[k];
                ac[k] = v < 0 ? v + c->L[k] : v;
            }
            int md = 0, noreac = 0;
            int nidx = orc_compact_p_to_h(ac, c->M);
            int cn = cell_count[nidx], cs = cell_start[nidx];
            for (int ci = cs; ci < cs + cn; ++ci) {
                int s = species2[ci];
                if (s <= 0) continue;
                double d[3];
                orc_rel_pos(x, pos2 + 3 * ci, L, d);
                double rsq = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
                if (rsq < radiussq) {
                    list_idx++;
                    md = 1;
                    if (lj && rsq <= lj->cut_sq[(s - 1) + lj->n1 * (si - 1)]) noreac = 1;
                    if (list_idx > nmax) return -1;
                    list[(int64_t)nmax * i + (list_idx - 1)] = ci;
                }
            }
            if (md) {
                is_md[nidx] = 1;
                if (noreac) is_reac[nidx] = 0;
            }
        }
        nlist[i] = list_idx;
    }
    return 0;
}

double orc_compute_force(int64_t n1, const double *pos1, const int32_t *species1,
                         double *force1, int64_t n2, const double *pos2,
                         const int32_t *species2, double *force2, int32_t *flags2,
                         const int32_t *list, const int32_t *nlist, int32_t nmax,
                         const double L[3], const orc_lj_t *lj)
{
    /* neighbor_list.f90:159-225 */
    const int MD = 1 << ORC_MD_BIT, PAST = 1 << ORC_PAST_MD_BIT, CAT = 1 << ORC_CATALYZED_BIT;
    ORC_PARALLEL_FOR
    for (int64_t i = 0; i < n2; ++i) {
        int f = flags2[i];
        f = (f & MD) ? (f | PAST) : (f & ~PAST);
        flags2[i] = f & ~MD;
    }
    double e = 0;
    for (int64_t i = 0; i < n1; ++i) {
        const double *x = pos1 + 3 * i;
        int s1 = species1[i];
        double f1[3] = {0, 0, 0};
        for (int j = 0; j < nlist[i]; ++j) {
            int idx = list[(int64_t)nmax * i + j];
            int s2 = species2[idx];
            if (s2 <= 0) continue;
            double d[3];
            orc_rel_pos(x, pos2 + 3 * idx, L, d);
            double r_sq = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
            int p = (s2 - 1) + lj->n1 * (s1 - 1);
            if (r_sq <= lj->cut_sq[p]) {
                double f[3];
                orc_lj_force(d, r_sq, lj->eps[p], lj->sigma[p], f);
                e = e + orc_lj_energy(r_sq, lj->eps[p], lj->sigma[p], lj->cut_energy[p]);
                for (int k = 0; k < 3; ++k) {
                    f1[k] = f1[k] + f[k];
                    force2[3 * idx + k] = force2[3 * idx + k] - f[k];
                }
                int mask = MD;
                if (s1 == 1) mask |= CAT;
                flags2[idx] |= mask;
            }
        }
        for (int k = 0; k < 3; ++k) force1[3 * i + k] = force1[3 * i + k] + f1[k];
    }
    return e;
}

double orc_compute_force_n2(int64_t n, const double *pos, const int32_t *species,
                            double *force, const double L[3], const orc_lj_t *lj)
{
    /* neighbor_list.f90:227-264 */
    double e = 0;
    for (int64_t i = 0; i < n; ++i) {
        int s1 = species[i];
        const double *x = pos + 3 * i;
        for (int64_t j = i + 1; j < n; ++j) {
            int s2 = species[j];
            double d[3];
            orc_rel_pos(x, pos + 3 * j, L, d);
            double r_sq = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
            int pc = (s1 - 1) + lj->n1 * (s2 - 1); /* cut_sq(s1,s2) */
            int pe = (s2 - 1) + lj->n1 * (s1 - 1); /* epsilon/sigma(s2,s1) */
            if (r_sq <= lj->cut_sq[pc]) {
                double f[3];
                orc_lj_force(d, r_sq, lj->eps[pe], lj->sigma[pe], f);
                for (int k = 0; k < 3; ++k) {
                    force[3 * i + k] = force[3 * i + k] + f[k];
                    force[3 * j + k] = force[3 * j + k] - f[k];
                }
                e = e + orc_lj_energy(r_sq, lj->eps[pe], lj->sigma[pe], lj->cut_energy[pe]);
            }
        }
    }
    return e;
}

/* ======================= driver helpers ========================================= */
void orc_flag_particles_nl(const int32_t *list1, int32_t n1, const double *cpos,
                           const double *pos, const int32_t *species, int32_t *flags,
                           const double edges[3], double cut_sq_c)
{
    /* programs/single_dimer_pbc.f90:493-509 */
    for (int s = 0; s < n1; ++s) {
        int r = list1[s];
        if (species[r] == 1) {
            double x[3];
            orc_rel_pos(cpos, pos + 3 * r, edges, x);
            if (dot3(x, x) < cut_sq_c) flags[r] |= 1 << ORC_REAC_BIT;
        }
    }
}

void orc_change_species(int64_t n, const double *cpos, const double *pos, int32_t *species,
                        int32_t *flags, const double edges[3], double cut_sq_c,
                        double cut_sq_n, double prob, const double *draws)
{
    /* programs/single_dimer_pbc.f90:511-542 */
    ORC_PARALLEL_FOR
    for (int64_t m = 0; m < n; ++m) {
        if (!(flags[m] & (1 << ORC_REAC_BIT))) continue;
        double x[3];
        orc_rel_pos(cpos, pos + 3 * m, edges, x);
        double dc = dot3(x, x);
        orc_rel_pos(cpos + 3, pos + 3 * m, edges, x);
        double dn = dot3(x, x);
        if (dc > cut_sq_c && dn > cut_sq_n) {
            if (draws[m] <= prob) species[m] = 2;
            flags[m] &= ~(1 << ORC_REAC_BIT);
        }
    }
}

/* splitmix64 (Steele, Lea and Flood 2014): one 64-bit output per counter value */
static uint64_t orc_splitmix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

void orc_change_species_seeded(int64_t n, const double *cpos, const double *pos,
                               int32_t *species, int32_t *flags, const double edges[3],
                               double cut_sq_c, double cut_sq_n, double prob, uint64_t seed,
                               int64_t step, const int32_t *id)
{
    /* programs/single_dimer_pbc.f90:511-542 with the draw made where the reference makes
       it (:533): only for a flagged particle outside both cutoffs, from a counter-based
       stream keyed by (seed, step, particle id), so the loop draws nothing for the others */
    ORC_PARALLEL_FOR
    for (int64_t m = 0; m < n; ++m) {
        if (!(flags[m] & (1 << ORC_REAC_BIT))) continue;
        double x[3];
        orc_rel_pos(cpos, pos + 3 * m, edges, x);
        double dc = dot3(x, x);
        orc_rel_pos(cpos + 3, pos + 3 * m, edges, x);
        double dn = dot3(x, x);
        if (dc > cut_sq_c && dn > cut_sq_n) {
            uint64_t r = orc_splitmix64(seed ^ orc_splitmix64(((uint64_t)step << 32) ^ (uint32_t)id[m]));
            double u = (double)(r >> 11) * 0x1.0p-53;
            if (u <= prob) species[m] = 2;
            flags[m] &= ~(1 << ORC_REAC_BIT);
        }
    }
}

void orc_refuel(int64_t n, const double *cpos, const double *pos, int32_t *species,
                const double edges[3], int L1)
{
    /* programs/single_dimer_pbc.f90:544-565: far = (L(1)*0.45)**2 in default REAL */
    float farf = (float)L1 * 0.45f;
    farf = farf * farf;
    double far = (double)farf;
    ORC_PARALLEL_FOR
    for (int64_t m = 0; m < n; ++m) {
        if (species[m] != 2) continue;
        double x[3];
        orc_rel_pos(cpos, pos + 3 * m, edges, x);
        double dc = dot3(x, x);
        orc_rel_pos(cpos + 3, pos + 3 * m, edges, x);
        double dn = dot3(x, x);
        if (dc > far && dn > far) species[m] = 1;
    }
}

void orc_flag_particles_nl_prob(const int32_t *list1, int32_t n1, const double *cpos,
                                const double *pos, const int32_t *species, int32_t *flags,
                                const double edges[3], double cut_sq_c, double prob,
                                const double *draws)
{
    /* test/setup_single_catalytic_fixed_sphere.f90:311-329: a candidate inside the cutoff
     * consumes one draw (draws[s] for list slot s) and is flagged when it is <= prob */
    for (int s = 0; s < n1; ++s) {
        int r = list1[s];
        if (species[r] == 1) {
            double x[3];
            orc_rel_pos(cpos, pos + 3 * r, edges, x);
            if (dot3(x, x) < cut_sq_c && draws[s] <= prob) flags[r] |= 1 << ORC_REAC_BIT;
        }
    }
}

void orc_change_species_sphere(int64_t n, const double *cpos, const double *pos,
                               int32_t *species, int32_t *flags, const double edges[3],
                               double cut_sq_c)
{
    /* test/setup_single_catalytic_fixed_sphere.f90:331-349 */
    ORC_PARALLEL_FOR
    for (int64_t m = 0; m < n; ++m) {
        if (!(flags[m] & (1 << ORC_REAC_BIT))) continue;
        double x[3];
        orc_rel_pos(cpos, pos + 3 * m, edges, x);
        if (dot3(x, x) > cut_sq_c) {
            species[m] = 2;
            flags[m] &= ~(1 << ORC_REAC_BIT);
        }
    }
}
