// rmpcd_md.hip -- velocity Verlet, periodic wrap, bounce-back walls, RATTLE for the
// dimer, colloid Verlet from the drivers and the maximum displacements
// (src/md.f90:61-319, src/particle_system.f90:357-410,
//  programs/single_dimer_pbc.f90:312-315,356-359).
#include "rmpcd_internal.hpp"

namespace rmpcd {

// md_pos (md.f90:61-77): over the 3n components of the plane layout
template <class F>
__global__ void k_md_pos(int64_t n3, typename F::real *__restrict__ pos,
                         const typename F::real *__restrict__ vel,
                         const typename F::real *__restrict__ force, double dt)
{
    using R = typename F::real;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    R rdt = (R)dt, dt_sq = (R)((dt * dt) / 2);
    pos[i] = (pos[i] + rdt * vel[i]) + dt_sq * force[i];
}

// md_vel (md.f90:219-237)
template <class F>
__global__ void k_md_vel(PView<F> p, double dt)
{
    using R = typename F::real;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    int f = (int)p.flags[i];
    if (!(f & WALL_MASK)) {
        R rdt = (R)dt;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int64_t q = k * p.n + i;
            p.vel[q] = p.vel[q] + (rdt * (p.force[q] + p.force_old[q])) / (R)2;
        }
    } else {
        p.flags[i] = (typename F::small)(f & ~WALL_MASK);
    }
}

// apply_pbc (md.f90:202-217)
template <class F>
__global__ void k_apply_pbc(int64_t n, typename F::real *__restrict__ pos, int32_t *__restrict__ image,
                            double e0, double e1, double e2)
{
    using R = typename F::real;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double e[3] = {e0, e1, e2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        R x = pos[k * n + i], ed = (R)e[k];
        int jump = (int)floor((double)(x / ed));
        image[k * n + i] += jump;
        pos[k * n + i] = x - (R)jump * ed;
    }
}

// cell_md_pos / cell_md_vel / cell_md_pos_{y,z}wall (md.f90:79-263): one thread per cell
template <class F>
__global__ void k_cell_md_pos(Box b, const int32_t *__restrict__ start, const int32_t *__restrict__ count,
                              const uint8_t *__restrict__ is_md, PView<F> p, double dt, int md_flag,
                              int wall_axis)
{
    using R = typename F::real;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b.ncells) return;
    if ((is_md[c] != 0) != (md_flag != 0)) return;
    int64_t s = start[c];
    int n = count[c];
    R rdt = (R)dt, dt_sq = (R)((dt * dt) / 2);
    for (int64_t j = s; j < s + n; ++j) {
        R x[3], v[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[k] = p.vel[k * p.n + j];
            x[k] = (p.pos[k * p.n + j] + rdt * v[k]) + dt_sq * p.force[k * p.n + j];
        }
        if (wall_axis > 0) {
            int ax = wall_axis;
            R Lw = (R)b.edges[ax], overtime = 0;
            bool collide = false;
            if (x[ax] < (R)0) {
                collide = true;
                overtime = x[ax] / v[ax];
            } else if (x[ax] > Lw) {
                collide = true;
                overtime = (x[ax] - Lw) / v[ax];
            }
            if (collide) {
                int bc = b.bc[ax];
                if (bc == RMPCD_BOUNCE_BACK_BC) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) v[k] = -v[k];
#pragma unroll
                    for (int k = 0; k < 3; ++k) x[k] = x[k] + ((R)2 * overtime) * v[k];
                } else if (bc == RMPCD_SPECULAR_BC) {
                    v[ax] = -v[ax];
                    x[ax] = x[ax] + ((R)2 * overtime) * v[ax];
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) p.vel[k * p.n + j] = v[k];
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) p.pos[k * p.n + j] = x[k];
    }
}

template <class F>
__global__ void k_cell_md_vel(int ncells, const int32_t *__restrict__ start,
                              const int32_t *__restrict__ count, const uint8_t *__restrict__ is_md,
                              PView<F> p, double dt, int md_flag)
{
    using R = typename F::real;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    if ((is_md[c] != 0) != (md_flag != 0)) return;
    int64_t s = start[c];
    int n = count[c];
    R rdt = (R)dt;
    for (int64_t j = s; j < s + n; ++j) {
        int f = (int)p.flags[j];
        if (!(f & WALL_MASK)) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                int64_t q = k * p.n + j;
                p.vel[q] = p.vel[q] + (rdt * (p.force[q] + p.force_old[q])) / (R)2;
            }
        } else {
            p.flags[j] = (typename F::small)(f & ~WALL_MASK);
        }
    }
}

// maximum_displacement (particle_system.f90:357-378): max of r^2 as the bit pattern of
// a non-negative double (order independent, so the result is exact)
template <class F>
__global__ void k_max_disp(PView<F> p, const typename F::real *__restrict__ pos_old,
                           double *__restrict__ out)
{
    __shared__ double red[16];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double r_sq = 0;
    if (i < p.n && (int)p.species[i] >= 0) {
        double d0 = (double)(p.pos[i] - pos_old[i]);
        double d1 = (double)(p.pos[p.n + i] - pos_old[p.n + i]);
        double d2 = (double)(p.pos[2 * p.n + i] - pos_old[2 * p.n + i]);
        r_sq = (d0 * d0 + d1 * d1) + d2 * d2;
    }
    double m = block_max(r_sq, red);
    if (threadIdx.x == 0) atomic_max_nonneg(out, m);
}

// cell_maximum_displacement (particle_system.f90:380-410), one thread per cell
template <class F>
__global__ void k_cell_max_disp(int ncells, const int32_t *__restrict__ start,
                                const int32_t *__restrict__ count, const uint8_t *__restrict__ is_md,
                                PView<F> p, const typename F::real *__restrict__ pos_old,
                                double *__restrict__ out)
{
    __shared__ double red[16];
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double local = 0;
    if (c < ncells && is_md[c]) {
        int64_t s = start[c];
        int n = count[c];
        for (int64_t j = s; j < s + n; ++j) {
            if ((int)p.species[j] >= 0) {
                double d0 = (double)(p.pos[j] - pos_old[j]);
                double d1 = (double)(p.pos[p.n + j] - pos_old[p.n + j]);
                double d2 = (double)(p.pos[2 * p.n + j] - pos_old[2 * p.n + j]);
                double r_sq = (d0 * d0 + d1 * d1) + d2 * d2;
                if (r_sq > local) local = r_sq;
            }
        }
    }
    double m = block_max(local, red);
    if (threadIdx.x == 0) atomic_max_nonneg(out, m);
}

__global__ void k_zero_f64(double *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---- colloids (always FP64) -----------------------------------------------------
struct Mass16 {
    double m[16];
};

__device__ inline void rel_pos3(const double *x, const double *y, const double *L, double *r)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = min_image(x[k] - y[k], L[k]);
}
__device__ inline double dot3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// rattle_dimer_pos (md.f90:265-296); closed form for two particles, one thread
__global__ void k_rattle_pos(PView<F64> c, const double *__restrict__ pr, Mass16 mass, double d,
                             double dt, double e0, double e1, double e2)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int64_t n = c.n;
    double L[3] = {e0, e1, e2};
    double a1[3] = {pr[0], pr[n], pr[2 * n]}, a2[3] = {pr[1], pr[n + 1], pr[2 * n + 1]};
    double x1[3] = {c.pos[0], c.pos[n], c.pos[2 * n]}, x2[3] = {c.pos[1], c.pos[n + 1], c.pos[2 * n + 1]};
    double r[3], s[3];
    rel_pos3(a1, a2, L, r);
    rel_pos3(x1, x2, L, s);
    double m1 = mass.m[(c.species[0] - 1) & 15], m2 = mass.m[(c.species[1] - 1) & 15];
    double inv_mass = 1 / m1 + 1 / m2;
    double rsq = dot3(r, r), ssq = dot3(s, s), rs = dot3(r, s);
    double g = rs - sqrt(rs * rs - rsq * (ssq - d * d));
    g = g / ((dt * inv_mass) * rsq);
    for (int k = 0; k < 3; ++k) {
        c.pos[k * n] = c.pos[k * n] - ((g * dt) * r[k]) / m1;
        c.pos[k * n + 1] = c.pos[k * n + 1] + ((g * dt) * r[k]) / m2;
        c.vel[k * n] = c.vel[k * n] - (g * r[k]) / m1;
        c.vel[k * n + 1] = c.vel[k * n + 1] + (g * r[k]) / m2;
    }
}

// rattle_dimer_vel (md.f90:298-319)
__global__ void k_rattle_vel(PView<F64> c, Mass16 mass, double d, double e0, double e1, double e2)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int64_t n = c.n;
    double L[3] = {e0, e1, e2};
    double x1[3] = {c.pos[0], c.pos[n], c.pos[2 * n]}, x2[3] = {c.pos[1], c.pos[n + 1], c.pos[2 * n + 1]};
    double m1 = mass.m[(c.species[0] - 1) & 15], m2 = mass.m[(c.species[1] - 1) & 15];
    double inv_mass = 1 / m1 + 1 / m2;
    double s[3], dv[3];
    rel_pos3(x1, x2, L, s);
    for (int k = 0; k < 3; ++k) dv[k] = c.vel[k * n] - c.vel[k * n + 1];
    double kk = dot3(dv, s) / ((d * d) * inv_mass);
    for (int k = 0; k < 3; ++k) {
        c.vel[k * n] = c.vel[k * n] - (kk * s[k]) / m1;
        c.vel[k * n + 1] = c.vel[k * n + 1] + (kk * s[k]) / m2;
    }
}

// colloid Verlet of the drivers (single_dimer_pbc.f90:312-315, 356-359)
__global__ void k_colloid_pos(PView<F64> c, const double *__restrict__ mass, double dt)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    for (int k = 0; k < 3; ++k) {
        int64_t q = k * c.n + i;
        c.pos[q] = (c.pos[q] + dt * c.vel[q]) + ((dt * dt) * c.force[q]) / (2 * mass[i]);
    }
}
__global__ void k_colloid_vel(PView<F64> c, const double *__restrict__ mass, double dt)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    for (int k = 0; k < 3; ++k) {
        int64_t q = k * c.n + i;
        c.vel[q] = c.vel[q] + (dt * (c.force[q] + c.force_old[q])) / (2 * mass[i]);
    }
}

template <class F>
int md_pos_impl(rmpcd_particles_t *pp, double dt)
{
    PView<F> p = view<F>(pp);
    k_md_pos<F><<<blocks_for(3 * p.n, 256), 256, 0, stream()>>>(3 * p.n, p.pos, p.vel, p.force, dt);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}
template <class F>
int md_vel_impl(rmpcd_particles_t *pp, double dt)
{
    PView<F> p = view<F>(pp);
    k_md_vel<F><<<blocks_for(p.n, 256), 256, 0, stream()>>>(p, dt);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}
template <class F>
int apply_pbc_impl(rmpcd_particles_t *pp, const double *e)
{
    PView<F> p = view<F>(pp);
    k_apply_pbc<F><<<blocks_for(p.n, 256), 256, 0, stream()>>>(p.n, p.pos, p.image, e[0], e[1], e[2]);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}
template <class F>
int cell_md_pos_impl(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *pp,
                     double dt, int md_flag, int axis)
{
    Box b = make_box(box);
    k_cell_md_pos<F><<<blocks_for(b.ncells, 256), 256, 0, stream()>>>(
        b, cells->cell_start, cells->cell_count, cells->is_md, view<F>(pp), dt, md_flag, axis);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}
template <class F>
int cell_md_vel_impl(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *pp,
                     double dt, int md_flag)
{
    k_cell_md_vel<F><<<blocks_for(box->ncells, 256), 256, 0, stream()>>>(
        box->ncells, cells->cell_start, cells->cell_count, cells->is_md, view<F>(pp), dt, md_flag);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

static int fetch_max(double *dev, double *r)
{
    double *host = (double *)pinned(64);
    if (!host) return set_error(RMPCD_ERR_HIP, "pinned staging allocation failed");
    RMPCD_HIP(hipMemcpyAsync(host, dev, sizeof(double), hipMemcpyDeviceToHost, stream()));
    RMPCD_HIP(hipStreamSynchronize(stream()));
    *r = sqrt(host[0]);
    return RMPCD_OK;
}

template <class F>
int max_disp_impl(const rmpcd_particles_t *pp, const void *pos_old, double *r)
{
    PView<F> p = view<F>(pp);
    double *red = (double *)workspace(WS_RED, 64);
    if (!red) return set_error(RMPCD_ERR_HIP, "maximum_displacement: workspace allocation failed");
    k_zero_f64<<<1, 64, 0, stream()>>>(red, 1);
    k_max_disp<F><<<blocks_for(p.n, 256), 256, 0, stream()>>>(p, (const typename F::real *)pos_old, red);
    RMPCD_LAUNCHED();
    return fetch_max(red, r);
}

template <class F>
int cell_max_disp_impl(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                       const rmpcd_particles_t *pp, const void *pos_old, double delta_t, double *r)
{
    PView<F> p = view<F>(pp);
    double *red = (double *)workspace(WS_RED, 64);
    if (!red) return set_error(RMPCD_ERR_HIP, "cell_maximum_displacement: workspace allocation failed");
    k_zero_f64<<<1, 64, 0, stream()>>>(red, 1);
    k_cell_max_disp<F><<<blocks_for(box->ncells, 256), 256, 0, stream()>>>(
        box->ncells, cells->cell_start, cells->cell_count, cells->is_md, p,
        (const typename F::real *)pos_old, red);
    RMPCD_LAUNCHED();
    int rc = fetch_max(red, r);
    if (rc) return rc;
    double bnd = box->max_v * delta_t;
    *r = *r > bnd ? *r : bnd;
    return RMPCD_OK;
}

static int check_colloids(const rmpcd_particles_t *c, const char *who, int need)
{
    int rc = check_particles(c, who, true);
    if (rc) return rc;
    RMPCD_CHECK(c->form == RMPCD_FP64, "%s: colloids must use the FP64 form", who);
    RMPCD_CHECK(c->n >= need, "%s: needs at least %d colloids", who, need);
    return RMPCD_OK;
}

static int mass_table(const double *mass, int32_t n_mass, Mass16 &m)
{
    if (!mass || n_mass < 1 || n_mass > 16)
        return set_error(RMPCD_ERR_ARG, "mass table must hold 1..16 species");
    for (int k = 0; k < 16; ++k) m.m[k] = k < n_mass ? mass[k] : 0.0;
    return RMPCD_OK;
}

}  // namespace rmpcd

using namespace rmpcd;

extern "C" {

int rmpcd_md_pos(rmpcd_particles_t *p, double dt)
{
    int rc = check_particles(p, "md_pos", true);
    if (rc) return rc;
    return RMPCD_DISPATCH_FORM(p, md_pos_impl, p, dt);
}

int rmpcd_md_vel(rmpcd_particles_t *p, double dt)
{
    int rc = check_particles(p, "md_vel", true);
    if (rc) return rc;
    return RMPCD_DISPATCH_FORM(p, md_vel_impl, p, dt);
}

int rmpcd_apply_pbc(rmpcd_particles_t *p, const double edges[3])
{
    int rc = check_particles(p, "apply_pbc", false);
    if (rc) return rc;
    RMPCD_CHECK(edges && p->image, "apply_pbc: null argument");
    return RMPCD_DISPATCH_FORM(p, apply_pbc_impl, p, edges);
}

int rmpcd_cell_md_pos(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *p,
                      double dt, int32_t md_flag)
{
    int rc = check_box(box, "cell_md_pos");
    if (!rc) rc = check_particles(p, "cell_md_pos", true);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start && cells->is_md, "cell_md_pos: null cell arrays");
    return RMPCD_DISPATCH_FORM(p, cell_md_pos_impl, box, cells, p, dt, md_flag, 0);
}

int rmpcd_cell_md_vel(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *p,
                      double dt, int32_t md_flag)
{
    int rc = check_box(box, "cell_md_vel");
    if (!rc) rc = check_particles(p, "cell_md_vel", true);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start && cells->is_md, "cell_md_vel: null cell arrays");
    return RMPCD_DISPATCH_FORM(p, cell_md_vel_impl, box, cells, p, dt, md_flag);
}

int rmpcd_cell_md_pos_wall(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *p,
                           double dt, int32_t md_flag, int32_t axis)
{
    int rc = check_box(box, "cell_md_pos_wall");
    if (!rc) rc = check_particles(p, "cell_md_pos_wall", true);
    if (rc) return rc;
    RMPCD_CHECK(axis == 1 || axis == 2, "cell_md_pos_wall: axis must be 1 (y) or 2 (z)");
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start && cells->is_md, "cell_md_pos_wall: null cell arrays");
    return RMPCD_DISPATCH_FORM(p, cell_md_pos_impl, box, cells, p, dt, md_flag, axis);
}

int rmpcd_maximum_displacement(const rmpcd_particles_t *p, const void *pos_old, double *r)
{
    int rc = check_particles(p, "maximum_displacement", false);
    if (rc) return rc;
    RMPCD_CHECK(pos_old && r, "maximum_displacement: null argument");
    return RMPCD_DISPATCH_FORM(p, max_disp_impl, p, pos_old, r);
}

int rmpcd_cell_maximum_displacement(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                                    const rmpcd_particles_t *p, const void *pos_old, double delta_t,
                                    double *r)
{
    int rc = check_box(box, "cell_maximum_displacement");
    if (!rc) rc = check_particles(p, "cell_maximum_displacement", false);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start && cells->is_md && pos_old && r,
                "cell_maximum_displacement: null argument");
    return RMPCD_DISPATCH_FORM(p, cell_max_disp_impl, box, cells, p, pos_old, delta_t, r);
}

int rmpcd_rattle_dimer_pos(rmpcd_particles_t *colloids, const double *pos_rattle, const double *mass,
                           int32_t n_mass, double d, double dt, const double edges[3])
{
    int rc = check_colloids(colloids, "rattle_dimer_pos", 2);
    if (rc) return rc;
    RMPCD_CHECK(pos_rattle && edges, "rattle_dimer_pos: null argument");
    Mass16 m;
    rc = mass_table(mass, n_mass, m);
    if (rc) return rc;
    k_rattle_pos<<<1, 64, 0, stream()>>>(view<F64>(colloids), pos_rattle, m, d, dt, edges[0], edges[1],
                                         edges[2]);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

int rmpcd_rattle_dimer_vel(rmpcd_particles_t *colloids, const double *mass, int32_t n_mass, double d,
                           double dt, const double edges[3])
{
    (void)dt;
    int rc = check_colloids(colloids, "rattle_dimer_vel", 2);
    if (rc) return rc;
    RMPCD_CHECK(edges != nullptr, "rattle_dimer_vel: null edges");
    Mass16 m;
    rc = mass_table(mass, n_mass, m);
    if (rc) return rc;
    k_rattle_vel<<<1, 64, 0, stream()>>>(view<F64>(colloids), m, d, edges[0], edges[1], edges[2]);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

int rmpcd_colloid_md_pos(rmpcd_particles_t *colloids, const double *mass_by_particle, double dt)
{
    int rc = check_colloids(colloids, "colloid_md_pos", 1);
    if (rc) return rc;
    RMPCD_CHECK(mass_by_particle != nullptr, "colloid_md_pos: null mass");
    k_colloid_pos<<<blocks_for(colloids->n, 256), 256, 0, stream()>>>(view<F64>(colloids),
                                                                       mass_by_particle, dt);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

int rmpcd_colloid_md_vel(rmpcd_particles_t *colloids, const double *mass_by_particle, double dt)
{
    int rc = check_colloids(colloids, "colloid_md_vel", 1);
    if (rc) return rc;
    RMPCD_CHECK(mass_by_particle != nullptr, "colloid_md_vel: null mass");
    k_colloid_vel<<<blocks_for(colloids->n, 256), 256, 0, stream()>>>(view<F64>(colloids),
                                                                       mass_by_particle, dt);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

}  // extern "C"
