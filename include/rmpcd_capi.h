/*
 * rmpcd_capi.h -- C ABI of the MI355X-native RMPCDMD solvent/colloid time step.
 *
 * Every entry point takes plain pointers and sizes.  Particle and cell arrays are
 * DEVICE pointers (HBM) in the library's layout:
 *   - vectors are three contiguous planes of n values: x[0..n), y[n..2n), z[2n..3n)
 *     (the reference's Fortran (3,N) is x/y/z interleaved; the Fortran binding layer
 *     transposes when it moves data between host arrays and the device);
 *   - all indices are 0-based: cell key h (reference cell index h+1), cell_start
 *     (reference value - 1), particle ids and id_to_idx values (reference - 1),
 *     neighbour-list entries (reference - 1);
 *   - form RMPCD_FP64: pos/vel/force are double, species/flags int32;
 *     form RMPCD_FP32 (mixed precision): pos/vel/force float, species/flags uint8;
 *     colloids are always RMPCD_FP64.
 * Scalars returned through host pointers (energies, temperature, displacements) are
 * produced synchronously.  Everything else is ordered on the library stream
 * (rmpcd_set_stream, default: the null stream) and returns without waiting.
 * Every entry point returns RMPCD_OK (0) or a negative status; rmpcd_last_error()
 * gives the message, which for the reference's own failure cases is the reference's
 * error-stop text.
 *
 * Randomness: routines that draw in the reference take either injected draws (a
 * device array laid out by the item that consumes them, the parity interface) or,
 * when that pointer is NULL, draw from the library's counter-based Threefry-2x64-20
 * generator keyed by (seed, stream) with counter (step, item).
 */
#ifndef RMPCD_CAPI_H
#define RMPCD_CAPI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RMPCD_OK 0
#define RMPCD_ERR_ARG -1       /* invalid argument (shape, null pointer, range) */
#define RMPCD_ERR_HIP -2       /* HIP runtime error */
#define RMPCD_ERR_OVERFLOW -3  /* 'maximum of neighbor list reached' */
#define RMPCD_ERR_DOMAIN -4    /* a particle lies outside the box it is keyed in */

#define RMPCD_FP64 0
#define RMPCD_FP32 1

/* flag bits: src/common.f90:49-61 */
#define RMPCD_REAC_BIT 1
#define RMPCD_MD_BIT 2
#define RMPCD_WALL_BIT 3
#define RMPCD_PAST_MD_BIT 4
#define RMPCD_CATALYZED_BIT 6
/* boundary conditions: src/cell_system.f90:43-45 */
#define RMPCD_PERIODIC_BC 1
#define RMPCD_BOUNCE_BACK_BC 2
#define RMPCD_SPECULAR_BC 3

/* cell_system_t scalars (src/cell_system.f90:20-41) */
typedef struct rmpcd_box {
    int32_t L[3];       /* cells per axis (z has L+1 cells with walls) */
    int32_t M[3];       /* Hilbert bits per axis */
    int32_t ncells;     /* padded cell count 2^(M1+M2+M3) (cell_system%N) */
    int32_t has_walls;
    double a;           /* cell size */
    double edges[3];    /* L*a as given to init */
    double origin[3];   /* grid shift, in [-1,0) cells after random_shift */
    int32_t bc[3];      /* PERIODIC/BOUNCE_BACK/SPECULAR per axis */
    double max_v;       /* bound used by cell_maximum_displacement */
} rmpcd_box_t;

/* particle_system_t arrays (src/particle_system.f90:22-77), device pointers */
typedef struct rmpcd_particles {
    int64_t n;          /* particle_system%Nmax */
    int32_t form;       /* RMPCD_FP64 or RMPCD_FP32 */
    int32_t pad_;
    void *pos;          /* 3 planes */
    void *vel;          /* 3 planes */
    void *force;        /* 3 planes */
    void *force_old;    /* 3 planes */
    int32_t *image;     /* 3 planes */
    int32_t *id;        /* n */
    void *species;      /* n, int32 (FP64) / uint8 (FP32) */
    void *flags;        /* n, int32 (FP64) / uint8 (FP32) */
} rmpcd_particles_t;

/* cell arrays, device pointers, each box->ncells long */
typedef struct rmpcd_cells {
    int32_t *cell_count;
    int32_t *cell_start;
    uint8_t *is_md;
    uint8_t *is_reac;
} rmpcd_cells_t;

/* lj_params_t (src/interaction.f90:23-34); arrays column-major (n1, n2), n1*n2 <= 16,
 * index (s1-1) + n1*(s2-1) with s1 the solvent species and s2 the colloid species */
typedef struct rmpcd_lj {
    int32_t n1, n2;
    double eps[16], sigma[16], cut[16], cut_sq[16], cut_energy[16];
} rmpcd_lj_t;

/* ---- runtime -------------------------------------------------------------------- */
const char *rmpcd_last_error(void);
int rmpcd_init(int device);
/* The library orders all its work on ONE stream, and its scratch buffers (sort keys and
 * slots, partial sums, the Hilbert table) are shared by every entry point.  Switching the
 * stream first waits for the work queued on the previous one, so nothing issued on the
 * new stream can overwrite scratch memory that earlier kernels still read.  The library
 * is not thread-safe: one host thread calls it at a time. */
int rmpcd_set_stream(void *hip_stream);
int rmpcd_synchronize(void);
const char *rmpcd_version(void);
/* device memory for callers without another allocator (the Fortran binding layer);
 * rmpcd_memcpy is synchronous, kind 0: host->device, 1: device->host, 2: device->device */
int rmpcd_malloc(void **ptr, size_t bytes);
int rmpcd_free(void *ptr);
int rmpcd_memcpy(void *dst, const void *src, size_t bytes, int32_t kind);

/* ---- hilbert.f90 ------------------------------------------------------------------ */
/* compact_p_to_h (src/hilbert.f90:241-271) and compact_h_to_p (:273-307) on the host */
int32_t rmpcd_compact_p_to_h(const int32_t p[3], const int32_t M[3]);
void rmpcd_compact_h_to_p(int32_t h, const int32_t M[3], int32_t p[3]);
/* device table keys[p0 + L0*(p1 + L1*p2)] for every cell of the box (p in [0,L)) */
int rmpcd_hilbert_table(const rmpcd_box_t *box, int32_t *keys);

/* ---- cell_system.f90 --------------------------------------------------------------- */
/* cell_system%init (src/cell_system.f90:49-84): fills the scalars of *box */
int rmpcd_box_init(rmpcd_box_t *box, const int32_t L[3], double a, int32_t has_walls);
/* cell_system%random_shift (src/cell_system.f90:153-163): origin = u - 1 */
int rmpcd_random_shift(rmpcd_box_t *box, const double u[3]);
/* cell_system%count_particles (src/cell_system.f90:86-129) */
int rmpcd_count_particles(const rmpcd_box_t *box, const rmpcd_particles_t *p,
                          rmpcd_cells_t *cells);
/* cartesian cell of each particle, 3 planes of int32 (cell_system.f90:165-175) */
int rmpcd_cartesian_indices(const rmpcd_box_t *box, const rmpcd_particles_t *p,
                            int32_t *cell_xyz);

/* ---- particle_system.f90 ------------------------------------------------------------ */
/* particle_system%sort (src/particle_system.f90:304-355): counts cells, then writes the
 * stably cell-sorted copy of src into dst (pos, image, vel, force, force_old, id,
 * species, flags) and id_to_idx[id] = new slot.  The caller swaps src and dst after
 * the call, as the reference's switch() calls do.  apply_pbc != 0 fuses md.f90:202-217
 * (applied to the particles before keying, as the drivers call it right before). */
int rmpcd_sort(const rmpcd_box_t *box, const rmpcd_particles_t *src, rmpcd_particles_t *dst,
               int32_t *id_to_idx, rmpcd_cells_t *cells, int32_t apply_pbc);
/* particle_system%maximum_displacement (src/particle_system.f90:357-378) */
int rmpcd_maximum_displacement(const rmpcd_particles_t *p, const void *pos_old, double *r);
/* cell_maximum_displacement (src/particle_system.f90:380-410) */
int rmpcd_cell_maximum_displacement(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                                    const rmpcd_particles_t *p, const void *pos_old,
                                    double delta_t, double *r);

/* ---- mpcd.f90 ----------------------------------------------------------------------- */
/* rand_sphere (src/mpcd.f90:43-62) applied to n_items independent uniform streams
 * u[item*stride .. item*stride+stride) (device); axes (3 planes of n_items) and
 * used[item] (uniforms consumed, -1 when the stream ran out) are device outputs */
int rmpcd_rand_sphere(const double *u, int64_t stride, int64_t n_items, double *axes,
                      int32_t *used);
/* simple_mpcd_step (src/mpcd.f90:71-183).  SRD branch: axes is NULL or a device array
 * of 3 planes x ncells, the unit axis drawn for each cell (read only for cells that
 * draw); otherwise the axis comes from Threefry keyed by (seed, step, cell).
 * alpha_given = 0 gives the reference default (sin = 1, cos = 0).  thermostat != 0 runs
 * the MPC-AT branch (mpcd.f90:144-160) with Threefry normals, as rmpcd_mpcat_step with
 * normals = NULL (axes is not read). */
int rmpcd_simple_mpcd_step(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                           rmpcd_particles_t *p, const double *axes, uint64_t seed,
                           uint64_t step, int32_t alpha_given, double alpha,
                           int32_t thermostat, double T, int32_t hydro);
/* The MPC-AT branch of simple_mpcd_step (src/mpcd.f90:144-160, thermostat = .true.):
 * normals is NULL or a device array of 3 planes x n, the three normals drawn for the
 * particle in each slot (read for species > 0 in cells that draw); otherwise they come
 * from Threefry keyed by (seed, step, particle id).  hydro = 0 leaves the new
 * velocities' cell mean as drawn (mpcd.f90:152-158). */
int rmpcd_mpcat_step(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *p,
                     const double *normals, uint64_t seed, uint64_t step, double T, int32_t hydro);
/* rescale_cells (src/mpcd.f90:748-801) */
int rmpcd_rescale_cells(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *p, double T);
/* bulk_reaction_endothermic (src/mpcd.f90:646-700); draws: NULL or device array, one per
 * cell (read for the cells that draw) */
int rmpcd_bulk_reaction_endothermic(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                                    rmpcd_particles_t *p, int32_t from, int32_t to, double rate,
                                    double tau, double delta_u, const double *draws, uint64_t seed,
                                    uint64_t step);
/* compute_temperature (src/mpcd.f90:342-407) */
int rmpcd_compute_temperature(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                              const rmpcd_particles_t *p, double *te);
/* No counterpart in the reference: compute_temperature (src/mpcd.f90:342-407) for callers
 * that cannot allocate (code under HIP graph capture, tight loops, harnesses that forbid
 * hipMalloc).  It allocates nothing and does not synchronise: the temperature is written to
 * te_dev (a device double of the caller) on the library stream, and the per-workgroup
 * partial sums use a fixed device array of 1 MiB per GPU (held until the process ends), so
 * boxes of at most 8,388,608 cells (2^23; a larger one is refused with an error).  Calls
 * that overlap on one GPU share that array and race.  Same kernels and summation order as
 * rmpcd_compute_temperature, so the same result. */
int rmpcd_compute_temperature_noalloc(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                                      const rmpcd_particles_t *p, double *te_dev);
/* compute_temperature with its optional z profile tz (src/mpcd.f90:342-407, the binning at
 * :395-398 through profile_t%bin, src/common.f90:196-208): each cell's kinetic term is
 * added to tz_data[floor((z - xmin) / dx)] and counted in tz_count (device arrays of n_bins,
 * accumulated, not cleared), z = origin(3) + (cell(3) + 1/2) a */
int rmpcd_compute_temperature_profile(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                                      const rmpcd_particles_t *p, double *te, double xmin, double dx,
                                      int32_t n_bins, double *tz_data, int32_t *tz_count);
/* mpcd_stream_periodic (src/mpcd.f90:440-457) */
int rmpcd_mpcd_stream_periodic(const rmpcd_box_t *box, rmpcd_particles_t *p, double dt);
/* bulk_reaction (src/mpcd.f90:605-643); draws: NULL or device array, one per cell */
int rmpcd_bulk_reaction(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                        rmpcd_particles_t *p, int32_t from, int32_t to, double rate,
                        double tau, const double *draws, uint64_t seed, uint64_t step);
/* wall_mpcd_step (src/mpcd.f90:193-337): the collision of a box with z walls.  walls_given
 * = 1: wall_temperature[2] and wall_n[2] (host arrays) fill the first and last z layer's
 * cells up to wall_n particles with virtual ones.  thermostat = 1: MPC-AT with
 * bulk_temperature (keep_cell_v = 0 drops the cell-mean restoration; it needs the
 * thermostat, as in the reference).  Injected draws, each NULL for Threefry: axes and
 * vnorm (the virtual particles' three normals) as 3 planes x ncells, pnorm (the
 * thermostat's normals) as 3 planes x n. */
int rmpcd_wall_mpcd_step(const rmpcd_box_t *box, const rmpcd_cells_t *cells, rmpcd_particles_t *p,
                         int32_t walls_given, const double *wall_temperature, const int32_t *wall_n,
                         int32_t thermostat, double bulk_temperature, int32_t keep_cell_v,
                         int32_t alpha_given, double alpha, const double *axes, const double *vnorm,
                         const double *pnorm, uint64_t seed, uint64_t step);
/* mpcd_stream_xforce_yzwall (src/mpcd.f90:459-511) with yzwall_collision (513-587): the
 * boundary conditions of y and z are box->bc[1], box->bc[2] */
int rmpcd_mpcd_stream_xforce_yzwall(const rmpcd_box_t *box, rmpcd_particles_t *p, double dt, double g);
/* mpcd_stream_nogravity_zwall (src/mpcd.f90:702-745) */
int rmpcd_mpcd_stream_nogravity_zwall(const rmpcd_box_t *box, rmpcd_particles_t *p,
                                      double dt);

/* ---- md.f90 ------------------------------------------------------------------------- */
int rmpcd_md_pos(rmpcd_particles_t *p, double dt);                 /* md.f90:61-77 */
int rmpcd_md_vel(rmpcd_particles_t *p, double dt);                 /* md.f90:219-237 */
int rmpcd_apply_pbc(rmpcd_particles_t *p, const double edges[3]);  /* md.f90:202-217 */
int rmpcd_cell_md_pos(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                      rmpcd_particles_t *p, double dt, int32_t md_flag);   /* md.f90:79-102 */
int rmpcd_cell_md_vel(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                      rmpcd_particles_t *p, double dt, int32_t md_flag);   /* md.f90:239-263 */
/* cell_md_pos_ywall (axis 1, md.f90:104-151) / cell_md_pos_zwall (axis 2, md.f90:153-200) */
int rmpcd_cell_md_pos_wall(const rmpcd_box_t *box, const rmpcd_cells_t *cells,
                           rmpcd_particles_t *p, double dt, int32_t md_flag, int32_t axis);
/* rattle_dimer_pos (md.f90:265-296) / rattle_dimer_vel (md.f90:298-319) on colloids 0,1;
 * pos_rattle: device, 3 planes of p->n; mass: host, n_mass (<= 16) entries indexed by
 * species-1 (particle_system%mass) */
int rmpcd_rattle_dimer_pos(rmpcd_particles_t *colloids, const double *pos_rattle,
                           const double *mass, int32_t n_mass, double d, double dt,
                           const double edges[3]);
int rmpcd_rattle_dimer_vel(rmpcd_particles_t *colloids, const double *mass, int32_t n_mass,
                           double d, double dt, const double edges[3]);
/* rattle_body_pos (md.f90:321-408) / rattle_body_vel (md.f90:410-480): iterative
 * RATTLE over a link list, links updated one after the other in list order, at most
 * 1000 sweeps.  links: host, 2*n_links zero-based particle indices (pairs); distances:
 * host, n_links (the library keeps a level schedule of the list on the device and
 * rebuilds it when the contents change); pos_rattle: device, 3 planes of p->n, unfolded
 * around particle 0 in place as the reference does; mass: host, by species-1.
 * *iterations (may be NULL, then the call does not wait for the device) receives the
 * sweep count, 1001 when the precision was not reached. */
int rmpcd_rattle_body_pos(rmpcd_particles_t *p, double *pos_rattle, const int32_t *links,
                          const double *distances, int32_t n_links, const double *mass,
                          int32_t n_mass, double dt, const double edges[3], double precision,
                          int32_t *iterations);
int rmpcd_rattle_body_vel(rmpcd_particles_t *p, const int32_t *links, const double *distances,
                          int32_t n_links, const double *mass, int32_t n_mass, double dt,
                          const double edges[3], double precision, int32_t *iterations);
/* elastic_network (md.f90:516-544): harmonic links added to p->force in link order;
 * links/distances: host, as for the rattles; the energy comes back in *energy (host) */
int rmpcd_elastic_network(rmpcd_particles_t *p, const int32_t *links, const double *distances,
                          int32_t n_links, double k, const double edges[3], double *energy);

/* rigid_body_t (md.f90:39-57) in device memory; quaternions scalar first.  pos_body:
 * device, 3*(i_stop-i_start+1), body-frame position of particle i at 3*(i-i_start).
 * i_start/i_stop are zero-based and inclusive. */
typedef struct rmpcd_rigid_body {
    int32_t i_start, i_stop;
    double mass;
    double pos[3], vel[3], force[3];
    double q[4];
    double L[3], torque[3];
    double I_body[3], omega_body[3];
    double *pos_body;
} rmpcd_rigid_body_t;
/* rigid_body_init (md.f90:547-588): body (device) and pos_body (device) are written */
int rmpcd_rigid_body_init(rmpcd_rigid_body_t *body, const rmpcd_particles_t *p, const double *mass,
                          int32_t n_mass, const double edges[3], int32_t i_start, int32_t i_stop,
                          double *pos_body);
/* rigid_body_compute_force_torque (md.f90:591-607) */
int rmpcd_rigid_body_compute_force_torque(rmpcd_rigid_body_t *body, const rmpcd_particles_t *p);
/* rigid_body_vv1 (md.f90:613-673): the half-step quaternion is iterated until it moves
 * less than threshold (at most 1000 rounds, the reference has no cap); *iterations (may
 * be NULL, then the call does not wait for the device) receives the round count, 1001
 * when the threshold was not reached */
int rmpcd_rigid_body_vv1(rmpcd_rigid_body_t *body, rmpcd_particles_t *p, double dt,
                         double threshold, const double edges[3], int32_t *iterations);
/* rigid_body_vv2 (md.f90:679-700) */
int rmpcd_rigid_body_vv2(rmpcd_rigid_body_t *body, rmpcd_particles_t *p, double dt);
/* colloid Verlet written inline in the drivers (programs/single_dimer_pbc.f90:312-315
 * and :356-359); mass_by_particle: device, one per colloid */
int rmpcd_colloid_md_pos(rmpcd_particles_t *colloids, const double *mass_by_particle,
                         double dt);
int rmpcd_colloid_md_vel(rmpcd_particles_t *colloids, const double *mass_by_particle,
                         double dt);

/* ---- interaction.f90 --------------------------------------------------------------- */
/* lj_params_init (src/interaction.f90:38-80), host; eps/sigma/cut column-major (n1,n2) */
int rmpcd_lj_params_init(rmpcd_lj_t *lj, int32_t n1, int32_t n2, const double *eps,
                         const double *sigma, const double *cut);

/* ---- neighbor_list.f90 ------------------------------------------------------------- */
/* make_stencil (src/neighbor_list.f90:271-308), host; returns the stencil size and
 * fills stencil[3*k..3*k+2] when stencil != NULL and cap is large enough */
int32_t rmpcd_make_stencil(double a, double cut, int32_t *stencil, int32_t cap);
/* update_list (src/neighbor_list.f90:76-157).  stencil: device (3*n_stencil);
 * list: device (npoints*nmax), entry k of point i at list[i*nmax+k]; n: device.
 * lj may be NULL (the optional argument absent). */
int rmpcd_update_list(const rmpcd_box_t *box, rmpcd_cells_t *cells,
                      const rmpcd_particles_t *points, const rmpcd_particles_t *solvent,
                      double radius, const int32_t *stencil, int32_t n_stencil,
                      const rmpcd_lj_t *lj, int32_t nmax, int32_t *list, int32_t *n);
/* compute_force (src/neighbor_list.f90:159-225); returns the energy in *e */
int rmpcd_compute_force(rmpcd_particles_t *ps1, rmpcd_particles_t *ps2, const int32_t *list,
                        const int32_t *n, int32_t nmax, const double L[3],
                        const rmpcd_lj_t *lj, double *e);
/* compute_force_n2 (src/neighbor_list.f90:227-264) */
int rmpcd_compute_force_n2(rmpcd_particles_t *ps, const double L[3], const rmpcd_lj_t *lj,
                           double *e);

/* ---- catalytic helpers written in the drivers (single_dimer_pbc.f90:493-565) -------- */
/* flag_particles_nl (:493-509): REAC bit for species-1 particles of colloid 0's list
 * inside cut_sq_c of colloid 0 */
int rmpcd_flag_particles_nl(rmpcd_particles_t *solvent, const rmpcd_particles_t *colloids,
                            const int32_t *list, const int32_t *n, int32_t nmax,
                            const double edges[3], double cut_sq_c);
/* change_species (:511-542); draws: NULL or device array, one per solvent slot */
int rmpcd_change_species(rmpcd_particles_t *solvent, const rmpcd_particles_t *colloids,
                         const double edges[3], double cut_sq_c, double cut_sq_n,
                         double prob, const double *draws, uint64_t seed, uint64_t step);
/* refuel (:544-565) */
int rmpcd_refuel(rmpcd_particles_t *solvent, const rmpcd_particles_t *colloids,
                 const double edges[3], int32_t L1);
/* flag_particles_nl of the fixed-sphere driver
 * (test/setup_single_catalytic_fixed_sphere.f90:311-329): a species-1 particle of colloid
 * 0's list inside cut_sq_c draws u and gets the REAC bit when u <= prob; draws: NULL
 * (Threefry keyed by (seed, step, id)) or device array, one per list slot */
int rmpcd_flag_particles_nl_prob(rmpcd_particles_t *solvent, const rmpcd_particles_t *colloids,
                                 const int32_t *list, const int32_t *n, int32_t nmax,
                                 const double edges[3], double cut_sq_c, double prob,
                                 const double *draws, uint64_t seed, uint64_t step);
/* change_species of the same driver (:331-349): a flagged particle farther than cut_sq_c
 * from colloid 0 becomes species 2 and loses the flag (no draw) */
int rmpcd_change_species_sphere(rmpcd_particles_t *solvent, const rmpcd_particles_t *colloids,
                                const double edges[3], double cut_sq_c);

/* ---- the single-dimer nanomotor program (programs/single_dimer_pbc.f90) -------------- */
/* The time loop of the driver on device-resident state: set-up (:183-296), the MD step
 * (:306-383) and the collision cycle (:386-443), with the catalytic helpers (:493-565).
 * mode RMPCD_SIM_REFERENCE issues the reference's call sequence through this ABI and
 * reads the skin test on the host every step; RMPCD_SIM_FUSED runs the same step as
 * fused device-side kernels, one hipGraph per collision cycle. */
#define RMPCD_SIM_REFERENCE 0
#define RMPCD_SIM_FUSED 1
typedef struct rmpcd_dimer_config {
    int32_t L[3];
    int32_t rho;          /* solvent per cell */
    int32_t n_md;         /* N_MD: MD steps per collision, dt = tau / N_MD */
    int32_t form;         /* RMPCD_FP64 or RMPCD_FP32 for the solvent */
    int32_t mode;         /* RMPCD_SIM_REFERENCE or RMPCD_SIM_FUSED */
    int32_t bulk_rmpcd;   /* bulk B->A reaction instead of refuel */
    double T, tau;
    double sigma_C, sigma_N, d;
    double epsilon_C[2];  /* C sphere with solvent species 1, 2 */
    double epsilon_N[2];  /* N sphere with solvent species 1, 2 */
    double epsilon_CC, epsilon_NC, epsilon_NN;
    double prob;          /* A->B probability on leaving the catalytic sphere */
    double bulk_rate;
    double skin;          /* 0: the driver's value, 2 */
    int32_t nmax_list;    /* 0: max(int(300 sigma^3), 1.3x the expected shell count) */
    int32_t n_replicas;   /* independent systems advanced together (0 or 1: one) */
    uint64_t seed;        /* replica r uses seed + r */
    const int32_t *replica_L;  /* host, 3 per replica; NULL: every replica uses L */
    int32_t stream_blocks;     /* fused mode launch shape, 0: default (16384): stream-role
                                  workgroups per replica of the step kernel */
    int32_t sort_blocks;       /* fused mode launch shape, 0: default (4096): workgroups of
                                  the key, place and gather passes */
    int32_t lazy;              /* fused mode: 1 = particles outside the neighbour lists are
                                  moved only when read (at a sort or an exact skin test), by
                                  replaying their per-step additions; the same trajectory
                                  bit for bit, without the per-step pass over all solvent */
    int32_t extra_sort_blocks; /* fused mode launch shape, 0: sort_blocks: workgroups of the
                                  key, place and rank passes of the sorts that may follow
                                  each MD step (they return at once when none is due) */
    int32_t sort_order;        /* fused mode launch shape: how the sort passes' workgroups
                                  take their units.  0: default (2), 1: a contiguous range
                                  per workgroup, 2: XCD-grouped (the workgroups of one XCD,
                                  blockIdx mod 8, share a contiguous eighth of the units,
                                  interleaved, so that the units in flight on an XCD are
                                  neighbours and share its L2) */
    int32_t groups;            /* fused mode launch shape: the replicas are split into this
                                  many groups of about equal particle count, each group's
                                  collision cycle one branch of the captured graph, so that
                                  one group's latency-bound work (colloids, sorts after MD
                                  steps) runs beside another's streaming.  0: default (2
                                  with two or more replicas), at most 8 and n_replicas */
    int32_t collide_cells;     /* fused mode launch shape: cells (one thread each) per
                                  workgroup of the collision kernel, 64, 128 or 256; 0:
                                  default (128) */
} rmpcd_dimer_config_t;
typedef struct rmpcd_dimer_stats {
    int64_t n_solvent, md_steps, cycles, n_extra_sorting;  /* of the replica asked for */
    int64_t n_replicas, n_solvent_total;
    double e1, e2;        /* last compute_force / compute_force_n2 energies */
    double dt;
    double mass[2];
    int32_t nmax;
    int32_t error;             /* 0, or the status this replica stopped with: a list
                                  longer than nmax stops it at that MD step, as the
                                  reference's error stop does (RMPCD_ERR_OVERFLOW) */
    int32_t *list, *nlist;     /* device: the colloids' neighbour lists */
    void *pos_old;             /* device: solvent positions at the last sort */
    double *colloid_pos_old;   /* device */
} rmpcd_dimer_stats_t;
typedef struct rmpcd_dimer rmpcd_dimer_t;
int rmpcd_dimer_create(const rmpcd_dimer_config_t *cfg, rmpcd_dimer_t **sim);
/* n_cycles collision cycles of N_MD steps each (single_dimer_pbc.f90:303-445) */
int rmpcd_dimer_run(rmpcd_dimer_t *sim, int32_t n_cycles);
/* views of replica r's current device state (valid until the next run) */
int rmpcd_dimer_state(rmpcd_dimer_t *sim, int32_t replica, rmpcd_particles_t *solvent,
                      rmpcd_particles_t *colloids, rmpcd_box_t *box, rmpcd_cells_t *cells,
                      rmpcd_dimer_stats_t *stats);
int rmpcd_dimer_destroy(rmpcd_dimer_t *sim);
/* fused mode: record events around the streaming kernel of every stride-th MD step of
 * the captured cycle (0 switches it off); rmpcd_dimer_stream_time returns the mean
 * duration of those launches in the last cycle run */
int rmpcd_dimer_profile(rmpcd_dimer_t *sim, int32_t stride);
int rmpcd_dimer_stream_time(rmpcd_dimer_t *sim, double *avg_ms, int32_t *samples);

/* ---- the other driver loops on device-resident state --------------------------------
 * RMPCD_PROG_FLUID     test/setup_simple_fluid.f90:89-127: SRD, streaming, grid shift,
 *                      sort; n_md streaming sub-steps of tau/n_md per collision (1 is
 *                      the driver's single mpcd_stream_periodic of tau)
 * RMPCD_PROG_SPHERE    test/setup_single_catalytic_fixed_sphere.f90:89-253: a fixed
 *                      catalytic sphere, solvent Verlet with neighbour-list forces,
 *                      flag/change at the sphere, SRD and the bulk B->A reaction
 * RMPCD_PROG_COLLOIDS  programs/n_colloids_pbc.f90:123-346: many LJ colloids, lazy
 *                      cell-wise streaming (cell_md_pos/vel), colloid-colloid n^2
 *                      forces; the first n_thermo cycles are the thermostat phase
 *                      (:227-280: MPC-AT at T, rescale_cells every 10th), the later
 *                      ones the production loop (:286-346: SRD, or MPC-AT with
 *                      thermostat, at T ramped to T_final over n_loop cycles)
 * Each runs the driver's call sequence through the entry points above. */
#define RMPCD_PROG_FLUID 0
#define RMPCD_PROG_SPHERE 1
#define RMPCD_PROG_COLLOIDS 2
typedef struct rmpcd_program_config {
    int32_t kind;
    int32_t L[3];
    int32_t rho;
    int32_t n_md;             /* MD (or streaming) steps per collision, dt = tau / n_md */
    int32_t form;             /* RMPCD_FP64 or RMPCD_FP32 for the solvent */
    int32_t n_colloids;       /* COLLOIDS */
    int32_t nmax_list;        /* 0: the driver's capacity */
    int32_t n_thermo;         /* COLLOIDS: thermostat-phase cycles (N_thermo_loop) */
    double T, tau;
    double sigma, epsilon;    /* solvent-colloid LJ (SPHERE, COLLOIDS) */
    double epsilon_colloids;  /* COLLOIDS: colloid-colloid LJ, sigma 2 sigma, cut 6 sigma */
    double prob;              /* SPHERE: flag probability */
    double bulk_rate;         /* SPHERE: bulk B->A rate */
    double skin;              /* 0: the driver's value (SPHERE 1.5, COLLOIDS 0.8) */
    uint64_t seed;
    /* COLLOIDS production loop: T = T + (i - 1) (T_final - T) / (n_loop - 1) at loop i
     * (n_loop < 2: T throughout); thermostat: MPC-AT instead of SRD; hydro: the MPC-AT
     * keeps each cell's mean velocity (both phases) */
    double T_final;
    int32_t n_loop;
    int32_t thermostat;
    int32_t hydro;
    /* COLLOIDS: 1 runs each collision cycle as one captured hipGraph with every decision on
     * the device (no host scalar per MD step); production loop with SRD only */
    int32_t graph;
} rmpcd_program_config_t;
typedef struct rmpcd_program_stats {
    int64_t n_solvent, n_colloids, md_steps, cycles, n_extra_sorting;
    double e1, e2;            /* last compute_force / compute_force_n2 energies */
    double temperature;       /* last compute_temperature of the collision cycle */
    double dt, mass;
    int32_t nmax, last_sort;  /* list capacity; COLLOIDS: md step of the last sort */
    int32_t *list, *nlist;    /* device: the colloids' neighbour lists (nmax each) */
    void *pos_old;            /* device: solvent pos_old (the sort's other buffer) */
    double *colloid_pos_old;  /* device */
} rmpcd_program_stats_t;
typedef struct rmpcd_program rmpcd_program_t;
int rmpcd_program_create(const rmpcd_program_config_t *cfg, rmpcd_program_t **prog);
int rmpcd_program_run(rmpcd_program_t *prog, int32_t n_cycles);
int rmpcd_program_state(rmpcd_program_t *prog, rmpcd_particles_t *solvent,
                        rmpcd_particles_t *colloids, rmpcd_box_t *box, rmpcd_cells_t *cells,
                        rmpcd_program_stats_t *stats);
int rmpcd_program_destroy(rmpcd_program_t *prog);

/* ---- counter-based generator used when draws are not injected ------------------------ */
/* uniform doubles in [0,1): out[i] = U(seed, stream, step, i), device output */
int rmpcd_threefry_uniform(uint64_t seed, uint64_t stream, uint64_t step, int64_t n,
                           double *out);
/* Threefry-2x64-20 block function on the host (Random123 spec), for known-answer tests */
void rmpcd_threefry2x64_20(const uint64_t ctr[2], const uint64_t key[2], uint64_t out[2]);

#ifdef __cplusplus
}
#endif
#endif
