/*
 * rmpcd_oracle.h -- CPU restatement (plain C, FP64) of the RMPCDMD solvent/colloid
 * time step, used ONLY as test infrastructure: the parity checker for the HIP path,
 * and the "port" CPU baseline that bench.py times.  Nothing in the product links it.
 *
 * Conventions (differences from the Fortran reference are index bases only):
 *   - all indices are 0-based: particle slots, cell indices (Hilbert key h, the
 *     reference's cell index is h+1), cell_start (reference value - 1), particle ids
 *     (reference id - 1), neighbour-list entries (reference entry - 1);
 *   - particle vectors use the reference's column-major (3,N) layout: x[3*i+k];
 *   - randomness is injected: every routine that draws in the reference takes the
 *     draws as arguments, laid out densely by the item that consumes them (cell or
 *     particle), exactly as a single OpenMP thread would consume them in order.
 *
 * Parity status per function is recorded next to each declaration:
 *   [pinned:ref]   checked against the reference compiled from its own sources
 *                  (oracle/_ref, amdflang) and the committed golden vectors;
 *   [pinned:math]  checked against closed forms / invariants / brute force;
 *   [parity unpinned] nothing independent pins it (stated with the reason).
 */
#ifndef RMPCD_ORACLE_H
#define RMPCD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* threads of the OpenMP loops; 1 (the default) gives the single-thread order */
void orc_set_threads(int n);
int orc_get_threads(void);

/* flag bits, src/common.f90:49-61 */
enum { ORC_REAC_BIT = 1, ORC_MD_BIT = 2, ORC_WALL_BIT = 3, ORC_PAST_MD_BIT = 4,
       ORC_CATALYZED_BIT = 6 };
enum { ORC_PERIODIC_BC = 1, ORC_BOUNCE_BACK_BC = 2, ORC_SPECULAR_BC = 3 };

/* ---- hilbert.f90 ---------------------------------------------------------- */
int  orc_p_to_h(const int p[3], int m);                         /* [pinned:ref] */
void orc_h_to_p(int h, int m, int p[3]);                        /* [pinned:ref] */
int  orc_compact_p_to_h(const int p[3], const int M[3]);        /* [pinned:ref] */
void orc_compact_h_to_p(int h, const int M[3], int p[3]);       /* [pinned:ref] */

/* ---- cell_system.f90 ------------------------------------------------------ */
typedef struct {
    int L[3];          /* cells per axis (z gets +1 with walls) */
    int M[3];          /* bits per axis */
    int N;             /* padded cell count 2^(M1+M2+M3) */
    double a;
    double edges[3];
    double origin[3];
    int has_walls;
    int bc[3];
    double max_v;
} orc_cells_t;

void orc_cells_init(orc_cells_t *c, const int L[3], double a, int has_walls); /* [pinned:math] */
void orc_cartesian_indices(const orc_cells_t *c, const double x[3], int p[3]); /* [pinned:math] */
int  orc_cell_key(const orc_cells_t *c, const double x[3]);                   /* [pinned:math] */
void orc_random_shift(orc_cells_t *c, const double u[3]);                     /* [pinned:math] */
/* count_particles: per-thread histograms + serial scan; cell_start is 0-based */
void orc_count_particles(const orc_cells_t *c, int64_t n, const double *pos,
                         int32_t *cell_count, int32_t *cell_start, int nthreads); /* [pinned:math] */

/* ---- particle_system.f90 -------------------------------------------------- */
typedef struct {
    int64_t n;
    double *pos, *vel, *force, *force_old;  /* (3,n) */
    int32_t *image;                          /* (3,n) */
    int32_t *id, *species, *flags;           /* (n)   */
} orc_particles_t;

/* sort: count + scatter into dst (the reference's *_old / *_store buffers);
 * nthreads==1 gives the reference's deterministic single-thread (stable) order.
 * id_to_idx[id] = new slot.  The caller swaps src/dst afterwards, as the
 * reference's switch() calls do. */
void orc_sort(const orc_cells_t *c, const orc_particles_t *src, orc_particles_t *dst,
              int32_t *id_to_idx, int32_t *cell_count, int32_t *cell_start,
              int nthreads);                                        /* [pinned:math] */
double orc_maximum_displacement(int64_t n, const double *pos, const double *pos_old,
                                const int32_t *species);            /* [pinned:math] */
double orc_cell_maximum_displacement(const orc_cells_t *c, const int32_t *cell_count,
                                     const int32_t *cell_start, const uint8_t *is_md,
                                     const double *pos, const double *pos_old,
                                     const int32_t *species, double delta_t); /* [pinned:math] */

/* ---- mpcd.f90 ------------------------------------------------------------- */
/* Marsaglia: consumes uniforms u[0..]; returns the number of uniforms used. */
int  orc_rand_sphere(const double *u, int n_avail, double n_out[3]);  /* [pinned:math] */
/* build per-cell axes (dense, 3 per cell, zeros for cells that draw nothing)
 * from one uniform stream consumed in single-thread cell order; returns #used or -1 */
int64_t orc_srd_axes_from_stream(const orc_cells_t *c, const int32_t *cell_count,
                                 const int32_t *cell_start, const int32_t *species,
                                 const double *u, int64_t n_u, double *axes);
/* SRD collision, non-thermostat branch; alpha_given=0 -> the default (sin=1,cos=0) */
void orc_simple_mpcd_step(const orc_cells_t *c, const int32_t *cell_count,
                          const int32_t *cell_start, double *vel, const int32_t *species,
                          const double *axes, int alpha_given, double alpha); /* [pinned:math] */
void orc_temperature_profile(const orc_cells_t *c, const int32_t *cell_count, const int32_t *cell_start,
                             const double *vel, const int32_t *species, double xmin, double dx, int n_bins,
                             double *data, int32_t *count);
double orc_compute_temperature(const orc_cells_t *c, const int32_t *cell_count,
                               const int32_t *cell_start, const double *vel,
                               const int32_t *species);              /* [pinned:math] */
void orc_mpcd_stream_periodic(const orc_cells_t *c, int64_t n, double *pos,
                              const double *vel, int32_t *image, double dt); /* [pinned:math] */
/* draws: one per cell (dense); consumed for every eligible cell */
void orc_bulk_reaction(const orc_cells_t *c, const int32_t *cell_count,
                       const int32_t *cell_start, const uint8_t *is_reac,
                       int32_t *species, int from, int to, double rate, double tau,
                       const double *draws);                          /* [pinned:math] */
void orc_mpcat_step(const orc_cells_t *c, const int32_t *cell_count, const int32_t *cell_start,
                    double *vel, const int32_t *species, const double *normals, double T,
                    int hydro);                                              /* [pinned:math] */
void orc_rescale_cells(const orc_cells_t *c, const int32_t *cell_count, const int32_t *cell_start,
                       double *vel, const int32_t *species, double T);       /* [pinned:math] */
void orc_bulk_reaction_endothermic(const orc_cells_t *c, const int32_t *cell_count,
                                   const int32_t *cell_start, const uint8_t *is_reac, double *vel,
                                   int32_t *species, int from, int to, double rate, double tau,
                                   double delta_u, const double *draws);     /* [pinned:math] */
void orc_wall_mpcd_step(const orc_cells_t *c, const int32_t *cell_count, const int32_t *cell_start,
                        double *vel, const int32_t *id, int walls_given, const double *wall_temperature,
                        const int32_t *wall_n, int thermostat, double bulk_temperature, int keep_cell_v,
                        int alpha_given, double alpha, const double *axes, const double *vnorm,
                        const double *pnorm);                               /* [pinned:math] */
void orc_mpcd_stream_xforce_yzwall(const orc_cells_t *c, int64_t n, double *pos, double *vel,
                                   int32_t *flags, double dt, double g);   /* [pinned:math] */
void orc_mpcd_stream_nogravity_zwall(const orc_cells_t *c, int64_t n, double *pos,
                                     double *vel, const double *force, int32_t *flags,
                                     double dt);                      /* [pinned:math] */

/* ---- md.f90 --------------------------------------------------------------- */
void orc_md_pos(int64_t n, double *pos, const double *vel, const double *force, double dt);
void orc_md_vel(int64_t n, double *vel, const double *force, const double *force_old,
                int32_t *flags, double dt);
void orc_apply_pbc(int64_t n, double *pos, int32_t *image, const double edges[3]);
void orc_cell_md_pos(const orc_cells_t *c, const int32_t *cell_count,
                     const int32_t *cell_start, const uint8_t *is_md, double *pos,
                     const double *vel, const double *force, double dt, int md_flag);
void orc_cell_md_vel(const orc_cells_t *c, const int32_t *cell_count,
                     const int32_t *cell_start, const uint8_t *is_md, double *vel,
                     const double *force, const double *force_old, int32_t *flags,
                     double dt, int md_flag);
/* axis: 1 (y, cell_md_pos_ywall) or 2 (z, cell_md_pos_zwall); bc from c->bc[axis] */
void orc_cell_md_pos_wall(const orc_cells_t *c, const int32_t *cell_count,
                          const int32_t *cell_start, const uint8_t *is_md, double *pos,
                          double *vel, const double *force, double dt, int md_flag,
                          int axis);
/* mass[] indexed by (species-1), as p%mass(p%species(k)) */
void orc_rattle_dimer_pos(double *pos, double *vel, const double *pos_rattle,
                          const int32_t *species, const double *mass, double d,
                          double dt, const double edges[3]);
void orc_rattle_dimer_vel(const double *pos, double *vel, const int32_t *species,
                          const double *mass, double d, double dt, const double edges[3]);
/* driver-inline colloid Verlet (single_dimer_pbc.f90:312-315, 356-359):
 * mass indexed by particle k, as the drivers do */
void orc_colloid_md_pos(int64_t n, double *pos, const double *vel, const double *force,
                        const double *mass_by_particle, double dt);
void orc_colloid_md_vel(int64_t n, double *vel, const double *force,
                        const double *force_old, const double *mass_by_particle, double dt);

/* ---- common.f90 / interaction.f90 ------------------------------------------ */
void orc_rel_pos(const double x[3], const double y[3], const double L[3], double r[3]);
/* ---- linked bodies (md.f90:321-480, 516-544), positions (n, 3), links zero-based pairs;
 * the rattles return the sweep counter at exit (1001 when the precision was not met) ---- */
int orc_rattle_body_pos(int64_t n, double *pos, double *vel, double *pos_rattle,
                        const int32_t *species, const double *mass, const int32_t *links,
                        const double *distances, int n_links, double dt, const double edges[3],
                        double precision);                                 /* [pinned:math] */
int orc_rattle_body_vel(int64_t n, double *pos, double *vel, const int32_t *species,
                        const double *mass, const int32_t *links, const double *distances,
                        int n_links, double dt, const double edges[3], double precision); /* [pinned:math] */
double orc_elastic_network(const double *pos, double *force, const int32_t *links,
                           const double *distances, int n_links, double k,
                           const double edges[3]);                         /* [pinned:math] */
/* rigid_body_t (md.f90:39-57), the layout of the library's rmpcd_rigid_body_t;
 * i_start/i_stop zero-based, inclusive; pos_body (n_body, 3) host rows */
typedef struct orc_body {
    int32_t i_start, i_stop;
    double mass;
    double pos[3], vel[3], force[3];
    double q[4];
    double L[3], torque[3];
    double I_body[3], omega_body[3];
    double *pos_body;
} orc_body_t;
void orc_rigid_body_init(orc_body_t *b, const double *pos, const double *vel, const int32_t *image,
                         const int32_t *species, const double *mass, const double edges[3],
                         int i_start, int i_stop);                         /* [pinned:math] */
void orc_rigid_body_compute_force_torque(orc_body_t *b, const double *force); /* [pinned:math] */
int orc_rigid_body_vv1(orc_body_t *b, double *pos, double dt, double threshold,
                       const double edges[3]);                             /* [pinned:math] */
void orc_rigid_body_vv2(orc_body_t *b, double *vel, double dt);            /* [pinned:math] */
typedef struct {
    int n1, n2;                      /* (n1 solvent species, n2 colloid species) */
    double eps[16], sigma[16], cut[16], cut_sq[16], cut_energy[16];  /* [s1-1 + n1*(s2-1)] */
} orc_lj_t;
int  orc_lj_params_init(orc_lj_t *lj, int n1, int n2, const double *eps,
                        const double *sigma, const double *cut);      /* [pinned:ref] */
void orc_lj_force(const double d[3], double r_sq, double eps, double sigma, double f[3]); /* [pinned:ref] */
double orc_lj_energy(double r_sq, double eps, double sigma, double cut_energy);   /* [pinned:ref] */

/* ---- neighbor_list.f90 ---------------------------------------------------- */
/* returns stencil size; fills st[3*k] when st != NULL */
int  orc_make_stencil(double a, double cut, int32_t *st);            /* [pinned:math] */
/* returns 0, or -1 on 'maximum of neighbor list reached'. list is (nmax, npoints). */
int  orc_update_list(const orc_cells_t *c, const int32_t *cell_count,
                     const int32_t *cell_start, int64_t n1, const double *pos1,
                     const int32_t *species1, const double *pos2, const int32_t *species2,
                     double radius, const int32_t *stencil, int n_stencil,
                     const orc_lj_t *lj, int32_t nmax, int32_t *list, int32_t *nlist,
                     uint8_t *is_md, uint8_t *is_reac);                /* [pinned:math] */
double orc_compute_force(int64_t n1, const double *pos1, const int32_t *species1,
                         double *force1, int64_t n2, const double *pos2,
                         const int32_t *species2, double *force2, int32_t *flags2,
                         const int32_t *list, const int32_t *nlist, int32_t nmax,
                         const double L[3], const orc_lj_t *lj);       /* [pinned:math] */
double orc_compute_force_n2(int64_t n, const double *pos, const int32_t *species,
                            double *force, const double L[3], const orc_lj_t *lj);

/* ---- driver helpers (programs/single_dimer_pbc.f90:493-565) ---------------- */
void orc_flag_particles_nl(const int32_t *list1, int32_t n1, const double *cpos,
                           const double *pos, const int32_t *species, int32_t *flags,
                           const double edges[3], double cut_sq_c);
/* draws dense per particle (used only when the particle qualifies) */
void orc_change_species(int64_t n, const double *cpos, const double *pos,
                        int32_t *species, int32_t *flags, const double edges[3],
                        double cut_sq_c, double cut_sq_n, double prob,
                        const double *draws);
/* the same with the draw made inline, only for the particles that convert or not (the
   reference's order of draws is not kept: for timing loops, not for parity) */
void orc_change_species_seeded(int64_t n, const double *cpos, const double *pos,
                               int32_t *species, int32_t *flags, const double edges[3],
                               double cut_sq_c, double cut_sq_n, double prob, uint64_t seed,
                               int64_t step, const int32_t *id);
void orc_refuel(int64_t n, const double *cpos, const double *pos, int32_t *species,
                const double edges[3], int L1);
/* the fixed-sphere driver's helpers (test/setup_single_catalytic_fixed_sphere.f90:311-349) */
void orc_flag_particles_nl_prob(const int32_t *list1, int32_t n1, const double *cpos,
                                const double *pos, const int32_t *species, int32_t *flags,
                                const double edges[3], double cut_sq_c, double prob,
                                const double *draws);
void orc_change_species_sphere(int64_t n, const double *cpos, const double *pos,
                               int32_t *species, int32_t *flags, const double edges[3],
                               double cut_sq_c);

#ifdef __cplusplus
}
#endif
#endif
