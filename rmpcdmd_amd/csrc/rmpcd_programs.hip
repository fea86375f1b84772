// rmpcd_programs.hip -- the driver loops other than the dimer's, on device-resident state,
// in the reference's call order through this library's own entry points:
//   FLUID     test/setup_simple_fluid.f90:45-127
//   SPHERE    test/setup_single_catalytic_fixed_sphere.f90:89-253 and its helpers :311-349
//   COLLOIDS  programs/n_colloids_pbc.f90:123-346: the thermostat phase (:227-280) for
//             the first n_thermo cycles, then the production loop (:286-346) with its
//             temperature ramp
// Random numbers come from Threefry-2x64-20 keyed by (seed, stream), with the counters
// given next to each use, so that a test can reproduce every draw:
//   grid shift s        threefry((s, 0)) a, b and threefry((s, 1)) a, stream SHIFT
//   SRD axes, cycle s   Marsaglia pairs from threefry((s, cell + 2^40 k)), stream SRD
//   bulk reaction       threefry((cycle, cell)) a, stream REACTION
//   sphere flag draw    threefry((md_step, id)) a, stream CHANGE
//   colloid placement   draw k: threefry((k, 2^62)) a, stream PLACE
//   MPC-AT, cycle s     the normals of particle id: Box-Muller on threefry((2s, id)) and
//                       threefry((2s + 1, id)), stream THERMO
#include <math.h>
#include <string.h>

#include <new>
#include <vector>

#include "rmpcd_programs.hpp"

namespace rmpcd {

int ProgramSim::setup(const rmpcd_program_config_t &c)
{
    cfg = c;
    RMPCD_CHECK(cfg.kind == RMPCD_PROG_FLUID || cfg.kind == RMPCD_PROG_SPHERE || cfg.kind == RMPCD_PROG_COLLOIDS,
                "program: unknown kind");
    RMPCD_CHECK(cfg.n_md > 0 && cfg.tau > 0 && cfg.rho > 0, "program: N_MD, tau and rho must be positive");
    RMPCD_CHECK(cfg.form == RMPCD_FP64 || cfg.form == RMPCD_FP32, "program: unknown form");
    RMPCD_CHECK(!cfg.graph || (cfg.kind == RMPCD_PROG_COLLOIDS && cfg.n_thermo == 0 && !cfg.thermostat),
                "program: the graph mode runs the COLLOIDS production loop with SRD (no thermostat)");
    int rc = rmpcd_box_init(&box, cfg.L, 1.0, 0);
    if (rc) return rc;
    seed = cfg.seed;
    dt = cfg.tau / cfg.n_md;
    int64_t vol = (int64_t)cfg.L[0] * cfg.L[1] * cfg.L[2];
    n = (int64_t)cfg.rho * vol;
    double sigma = cfg.sigma, f6 = pow(2.0, 1.0 / 6.0);
    double sigma_cut = sigma * f6;
    if (cfg.kind == RMPCD_PROG_SPHERE) {
        nc = 1;
        // mass = rho * sigma**3 * 4 * 3.14159265/3, the literal in default REAL (:103)
        mass = (((cfg.rho * (sigma * sigma * sigma)) * 4) * (double)3.14159265f) / 3;
        double eps[2] = {cfg.epsilon, cfg.epsilon}, sig[2] = {sigma, sigma}, cut[2] = {sigma_cut, sigma_cut};
        if ((rc = rmpcd_lj_params_init(&lj_sc, 2, 1, eps, sig, cut))) return rc;
        skin = cfg.skin > 0 ? cfg.skin : 1.5;  // :175
        nmax = (int)(300 * (sigma * sigma * sigma));  // neigh%init(colloids%Nmax, int(300*sigma**3)), :173
    } else if (cfg.kind == RMPCD_PROG_COLLOIDS) {
        nc = cfg.n_colloids;
        RMPCD_CHECK(nc > 0 && nc <= 100000, "program: n_colloids must be in [1, 100000]");
        // mass = rho * sigma**3 * 4 * 3.141/3 (n_colloids_pbc.f90:123)
        mass = (((cfg.rho * (sigma * sigma * sigma)) * 4) * (double)3.141f) / 3;
        // N = rho*L(1)*L(2)*L(3) - int(rho*4*3.142/3 * sigma**3*colloids%Nmax) (:147): the
        // left factor is evaluated in default REAL
        float lf = (float)(cfg.rho * 4) * 3.142f;
        lf = lf / 3.0f;
        double excl = ((double)lf * (sigma * sigma * sigma)) * (double)nc;
        n = (int64_t)cfg.rho * vol - (int64_t)excl;
        if (n <= 0) return set_error(RMPCD_ERR_ARG, "Not enough volume available for solvent");
        double eps = cfg.epsilon, e2c = cfg.epsilon_colloids;
        if ((rc = rmpcd_lj_params_init(&lj_sc, 1, 1, &eps, &sigma, &sigma_cut))) return rc;
        double csig = 2 * sigma, ccut = 2 * (3 * sigma);  // colloid_lj (:160-163)
        if ((rc = rmpcd_lj_params_init(&lj_cc, 1, 1, &e2c, &csig, &ccut))) return rc;
        skin = cfg.skin > 0 ? cfg.skin : 0.8;  // :212
        double rs = sigma_cut + skin;
        nmax = (int)(250 * (rs * rs));  // neigh%init(colloids%Nmax, int(250*(sigma_cut+skin)**2)), :213
    } else {
        nc = 0;
    }
    if (cfg.nmax_list > 0) nmax = cfg.nmax_list;
    RMPCD_CHECK(n > 0 && n < ((int64_t)1 << 31), "program: solvent count out of range");
    radius = sigma_cut + skin;

    size_t RS = rsize(cfg.form), SS = ssize(cfg.form);
    for (int b = 0; b < 2; ++b) {
        if ((rc = alloc(&sol_pos[b], 3 * n * RS))) return rc;
        if ((rc = alloc(&sol_vel[b], 3 * n * RS))) return rc;
        if ((rc = alloc((void **)&sol_image[b], 3 * n * 4))) return rc;
        if ((rc = alloc((void **)&sol_id[b], n * 4))) return rc;
        if ((rc = alloc(&sol_species[b], n * SS))) return rc;
        if ((rc = alloc(&sol_flags[b], n * SS))) return rc;
    }
    for (int b = 0; b < 4; ++b)
        if ((rc = alloc(&sol_force[b], 3 * n * RS))) return rc;
    if ((rc = alloc((void **)&id_to_idx, n * 4))) return rc;
    if ((rc = alloc((void **)&cell_count, box.ncells * 4))) return rc;
    if ((rc = alloc((void **)&cell_start, box.ncells * 4))) return rc;
    if ((rc = alloc((void **)&is_md, box.ncells))) return rc;
    if ((rc = alloc((void **)&is_reac, box.ncells))) return rc;
    if ((rc = alloc((void **)&scratch, 1 << 16))) return rc;
    if (nc) {
        size_t c3 = 3 * (size_t)nc * sizeof(double);
        if ((rc = alloc((void **)&col_pos, c3))) return rc;
        if ((rc = alloc((void **)&col_pos_old, c3))) return rc;
        if ((rc = alloc((void **)&col_vel, c3))) return rc;
        if ((rc = alloc((void **)&col_force[0], c3))) return rc;
        if ((rc = alloc((void **)&col_force[1], c3))) return rc;
        if ((rc = alloc((void **)&col_image, 3 * (size_t)nc * 4))) return rc;
        if ((rc = alloc((void **)&col_id, (size_t)nc * 4))) return rc;
        if ((rc = alloc((void **)&col_species, (size_t)nc * 4))) return rc;
        if ((rc = alloc((void **)&col_flags, (size_t)nc * 4))) return rc;
        if ((rc = alloc((void **)&col_mass, (size_t)nc * 8))) return rc;
        if ((rc = alloc((void **)&list, (size_t)nc * nmax * 4))) return rc;
        if ((rc = alloc((void **)&nlist, (size_t)nc * 4))) return rc;
        n_stencil = rmpcd_make_stencil(1.0, radius, nullptr, 0);
        std::vector<int32_t> st(3 * (size_t)n_stencil);
        rmpcd_make_stencil(1.0, radius, st.data(), n_stencil);
        if ((rc = alloc((void **)&stencil, st.size() * 4))) return rc;
        RMPCD_HIP(hipMemcpy(stencil, st.data(), st.size() * 4, hipMemcpyHostToDevice));
        std::vector<int32_t> ids(nc), sp(nc, 1);
        std::vector<double> ms(nc, mass);
        for (int i = 0; i < nc; ++i) ids[i] = i;
        RMPCD_HIP(hipMemcpy(col_id, ids.data(), 4 * (size_t)nc, hipMemcpyHostToDevice));
        RMPCD_HIP(hipMemcpy(col_species, sp.data(), 4 * (size_t)nc, hipMemcpyHostToDevice));
        RMPCD_HIP(hipMemcpy(col_mass, ms.data(), 8 * (size_t)nc, hipMemcpyHostToDevice));
        if ((rc = place_colloids())) return rc;
    }
    if ((rc = init_solvent())) return rc;

    rmpcd_particles_t s, p;
    switch (cfg.kind) {
    case RMPCD_PROG_FLUID:
        // sort, count, first grid shift (setup_simple_fluid.f90:82-86)
        if ((rc = sort(false))) return rc;
        if ((rc = shift(0))) return rc;
        break;
    case RMPCD_PROG_SPHERE:
        // sort, list, first forces (:171-184); pos_old keeps the unsorted positions
        if ((rc = sort(false))) return rc;
        if ((rc = build_list())) return rc;
        s = sv();
        p = cv();
        if ((rc = rmpcd_compute_force(&p, &s, list, nlist, nmax, box.edges, &lj_sc, &e1))) return rc;
        RMPCD_HIP(hipMemcpyAsync(s.force_old, s.force, 3 * n * RS, hipMemcpyDeviceToDevice, stream()));
        RMPCD_HIP(hipMemcpyAsync(p.force_old, p.force, 3 * nc * 8, hipMemcpyDeviceToDevice, stream()));
        break;
    case RMPCD_PROG_COLLOIDS:
        // sort, list, first forces, pos_old = pos (n_colloids_pbc.f90:209-226)
        if ((rc = sort(false))) return rc;
        if ((rc = build_list())) return rc;
        s = sv();
        p = cv();
        if ((rc = rmpcd_compute_force(&p, &s, list, nlist, nmax, box.edges, &lj_sc, &e1))) return rc;
        if ((rc = rmpcd_compute_force_n2(&p, box.edges, &lj_cc, &e2))) return rc;
        RMPCD_HIP(hipMemcpyAsync(s.force_old, s.force, 3 * n * RS, hipMemcpyDeviceToDevice, stream()));
        RMPCD_HIP(hipMemcpyAsync(p.force_old, p.force, 3 * nc * 8, hipMemcpyDeviceToDevice, stream()));
        if ((rc = copy_pos_old())) return rc;
        break;
    }
    RMPCD_HIP(hipStreamSynchronize(stream()));
    return RMPCD_OK;
}

// colloid positions: the box centre (setup_single_catalytic_fixed_sphere.f90:150), or
// uniform draws rejected within colloid_lj sigma of an earlier colloid
// (n_colloids_pbc.f90:189-205)
int ProgramSim::place_colloids()
{
    std::vector<double> pos(3 * (size_t)nc);
    if (cfg.kind == RMPCD_PROG_SPHERE) {
        for (int k = 0; k < 3; ++k) pos[k] = box.edges[k] / 2.0;
    } else {
        double s2 = lj_cc.sigma[0] * lj_cc.sigma[0];
        uint64_t draw = 0;
        auto uni = [&]() {
            uint64_t a, b;
            threefry2x64(draw++, 1ull << 62, seed, STREAM_PLACE, a, b);
            return u01(a);
        };
        int i = 0;
        while (i < nc) {
            RMPCD_CHECK(draw < 3000000000ull, "program: no room for the colloids");
            double x[3];
            for (int k = 0; k < 3; ++k) x[k] = uni() * cfg.L[k];
            bool tooclose = false;
            for (int j = 0; j < i && !tooclose; ++j) {
                double r[3];
                for (int k = 0; k < 3; ++k) r[k] = min_image(x[k] - pos[(size_t)k * nc + j], box.edges[k]);
                if ((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2] < s2) tooclose = true;
            }
            if (!tooclose) {
                for (int k = 0; k < 3; ++k) pos[(size_t)k * nc + i] = x[k];
                ++i;
            }
        }
    }
    RMPCD_HIP(hipMemcpy(col_pos, pos.data(), 8 * pos.size(), hipMemcpyHostToDevice));
    return RMPCD_OK;
}

template <class F>
static int init_solvent_t(ProgramSim &s)
{
    hipStream_t st = stream();
    rmpcd_particles_t sv = s.sv();
    PView<F> p = view<F>(&sv);
    k_fill_state<F><<<blocks_for(s.n, 256), 256, 0, st>>>(p);
    RMPCD_LAUNCHED();
    unsigned nb = blocks_for(s.n, 256);
    double *part = (double *)workspace(WS_PART, sizeof(double) * 3 * (nb + 1));
    if (!part) return set_error(RMPCD_ERR_HIP, "program: workspace allocation failed");
    double *sum = s.scratch;
    // SPHERE: uniform velocities (random_number - 0.5) * sqrt(12 T) (:143-144);
    // FLUID, COLLOIDS: sqrt(T) times normals
    bool uniform = s.cfg.kind == RMPCD_PROG_SPHERE;
    double scale = uniform ? sqrt(12 * s.cfg.T) : sqrt(s.cfg.T);
    k_init_velocities<F><<<nb, 256, 0, st>>>(p, s.seed, scale, part, uniform ? 1 : 0);
    RMPCD_LAUNCHED();
    k_sum_blocks3<3><<<1, 256, 0, st>>>(part, (int)nb, sum);
    RMPCD_LAUNCHED();
    k_remove_mean<F><<<nb, 256, 0, st>>>(p, sum);
    RMPCD_LAUNCHED();
    rmpcd_particles_t cv = s.cv();
    int32_t *fail = (int32_t *)(sum + 4);
    RMPCD_HIP(hipMemsetAsync(fail, 0, 4, st));
    const rmpcd_lj_t &l = s.lj_sc;
    k_place_random<F><<<nb, 256, 0, st>>>(p, view<F64>(&cv), s.nc, s.box.edges[0], s.box.edges[1], s.box.edges[2],
                                          l.n1 > 0 ? l.n1 : 1, l.cut_sq[0], l.cut_sq[1], l.cut_sq[2], l.cut_sq[3],
                                          s.seed, fail);
    RMPCD_LAUNCHED();
    int32_t hf = 0;
    RMPCD_HIP(hipMemcpyAsync(&hf, fail, 4, hipMemcpyDeviceToHost, st));
    RMPCD_HIP(hipStreamSynchronize(st));
    if (hf) return set_error(RMPCD_ERR_ARG, "program: random placement found no free position");
    return RMPCD_OK;
}

int ProgramSim::init_solvent()
{
    return cfg.form == RMPCD_FP64 ? init_solvent_t<F64>(*this) : init_solvent_t<F32>(*this);
}

// FLUID collision cycle c (setup_simple_fluid.f90:108-118): SRD on the cells of the last
// sort, streaming, the next grid shift, sort
int ProgramSim::fluid_cycle()
{
    rmpcd_particles_t s = sv();
    rmpcd_cells_t c = cells();
    int rc = rmpcd_simple_mpcd_step(&box, &c, &s, nullptr, seed, (uint64_t)cycles, 0, 0.0, 0, cfg.T, 1);
    if (rc) return rc;
    for (int k = 0; k < cfg.n_md; ++k)
        if ((rc = rmpcd_mpcd_stream_periodic(&box, &s, dt))) return rc;
    if ((rc = shift((uint64_t)cycles + 1))) return rc;
    if ((rc = sort(false))) return rc;
    md_steps += cfg.n_md;
    ++cycles;
    return RMPCD_OK;
}

// SPHERE MD step (setup_single_catalytic_fixed_sphere.f90:190-228)
int ProgramSim::sphere_step()
{
    rmpcd_particles_t s = sv();
    int rc = rmpcd_md_pos(&s, dt);
    if (rc) return rc;
    double so_max;
    if ((rc = rmpcd_maximum_displacement(&s, alt().pos, &so_max))) return rc;
    if (so_max >= skin) {
        if ((rc = sort(true))) return rc;  // apply_pbc, sort
        if ((rc = build_list())) return rc;
        if ((rc = copy_pos_old())) return rc;
        ++n_extra;
    }
    if ((rc = forces(false))) return rc;  // the driver zeroes the solvent forces only
    s = sv();
    rmpcd_particles_t p = cv();
    if ((rc = rmpcd_md_vel(&s, dt))) return rc;
    if ((rc = rmpcd_flag_particles_nl_prob(&s, &p, list, nlist, nmax, box.edges, lj_sc.cut_sq[0], cfg.prob,
                                           nullptr, seed, (uint64_t)md_steps)))
        return rc;
    if ((rc = rmpcd_change_species_sphere(&s, &p, box.edges, lj_sc.cut_sq[0]))) return rc;
    ++md_steps;
    return RMPCD_OK;
}

// SPHERE collision (:230-253): temperature, grid shift, sort without apply_pbc, list,
// SRD, bulk B->A reaction
int ProgramSim::sphere_collide()
{
    rmpcd_particles_t s = sv();
    rmpcd_cells_t c = cells();
    int rc = rmpcd_compute_temperature(&box, &c, &s, &temperature);
    if (rc) return rc;
    if ((rc = shift((uint64_t)cycles))) return rc;
    if ((rc = sort(false))) return rc;
    if ((rc = build_list())) return rc;
    s = sv();
    if ((rc = rmpcd_simple_mpcd_step(&box, &c, &s, nullptr, seed, (uint64_t)cycles, 0, 0.0, 0, cfg.T, 1))) return rc;
    if ((rc = rmpcd_bulk_reaction(&box, &c, &s, 2, 1, cfg.bulk_rate, cfg.tau, nullptr, seed, (uint64_t)cycles)))
        return rc;
    ++cycles;
    return RMPCD_OK;
}

// COLLOIDS MD step (n_colloids_pbc.f90:288-320): lazy streaming of the cells holding
// list particles; the others catch up (cell_md_pos/vel with md_flag false) at a sort
int ProgramSim::colloids_step()
{
    rmpcd_particles_t s = sv(), p = cv();
    rmpcd_cells_t c = cells();
    int rc = rmpcd_cell_md_pos(&box, &c, &s, dt, 1);
    if (rc) return rc;
    if ((rc = rmpcd_colloid_md_pos(&p, col_mass, dt))) return rc;
    double k1 = (double)(md_steps + 1 - last_sort);
    double so_max, co_max;
    if ((rc = rmpcd_cell_maximum_displacement(&box, &c, &s, alt().pos, dt * k1, &so_max))) return rc;
    if ((rc = rmpcd_maximum_displacement(&p, col_pos_old, &co_max))) return rc;
    if (co_max >= skin * 0.1 || so_max >= skin * 0.9) {
        if ((rc = rmpcd_cell_md_pos(&box, &c, &s, k1 * dt, 0))) return rc;
        if ((rc = rmpcd_cell_md_vel(&box, &c, &s, k1 * dt, 0))) return rc;
        if ((rc = rmpcd_apply_pbc(&p, box.edges))) return rc;
        if ((rc = sort(true))) return rc;
        last_sort = md_steps + 1;
        if ((rc = build_list())) return rc;
        if ((rc = copy_pos_old())) return rc;
        ++n_extra;
    }
    if ((rc = forces(true))) return rc;
    s = sv();
    p = cv();
    if ((rc = rmpcd_cell_md_vel(&box, &c, &s, dt, 1))) return rc;
    if ((rc = rmpcd_colloid_md_vel(&p, col_mass, dt))) return rc;
    ++md_steps;
    return RMPCD_OK;
}

// COLLOIDS collision (:322-344)
int ProgramSim::colloids_collide()
{
    int rc = shift((uint64_t)cycles);
    if (rc) return rc;
    rmpcd_particles_t s = sv(), p = cv();
    rmpcd_cells_t c = cells();
    double k2 = (double)(md_steps - last_sort);
    if ((rc = rmpcd_cell_md_pos(&box, &c, &s, k2 * dt, 0))) return rc;
    if ((rc = rmpcd_cell_md_vel(&box, &c, &s, k2 * dt, 0))) return rc;
    if ((rc = rmpcd_apply_pbc(&p, box.edges))) return rc;
    if ((rc = sort(true))) return rc;
    last_sort = md_steps;
    if ((rc = build_list())) return rc;
    if ((rc = copy_pos_old())) return rc;
    s = sv();
    if (cycles < cfg.n_thermo) {
        // thermostat phase: MPC-AT at T, rescale_cells every 10th loop (:277-279)
        if ((rc = rmpcd_simple_mpcd_step(&box, &c, &s, nullptr, seed, (uint64_t)cycles, 0, 0.0, 1, cfg.T,
                                          cfg.hydro)))
            return rc;
        if ((cycles + 1) % 10 == 0 && (rc = rmpcd_rescale_cells(&box, &c, &s, cfg.T))) return rc;
    } else {
        // production loop i (1-based): T ramped from T to T_final over n_loop loops (:334-336)
        const int64_t i = cycles - cfg.n_thermo + 1;
        const double T = cfg.n_loop > 1 ? cfg.T + (double)(i - 1) * (cfg.T_final - cfg.T) / (double)(cfg.n_loop - 1)
                                        : cfg.T;
        if ((rc = rmpcd_simple_mpcd_step(&box, &c, &s, nullptr, seed, (uint64_t)cycles, 0, 0.0, cfg.thermostat ? 1 : 0,
                                          T, cfg.hydro)))
            return rc;
    }
    if ((rc = rmpcd_compute_temperature(&box, &c, &s, &temperature))) return rc;
    ++cycles;
    return RMPCD_OK;
}

int ProgramSim::run(int n_cycles)
{
    if (cfg.kind == RMPCD_PROG_COLLOIDS && cfg.graph) return colloids_graph_run(*this, n_cycles);
    int rc;
    for (int cy = 0; cy < n_cycles; ++cy) {
        switch (cfg.kind) {
        case RMPCD_PROG_FLUID:
            if ((rc = fluid_cycle())) return rc;
            break;
        case RMPCD_PROG_SPHERE:
            for (int j = 0; j < cfg.n_md; ++j)
                if ((rc = sphere_step())) return rc;
            if ((rc = sphere_collide())) return rc;
            break;
        default:
            for (int j = 0; j < cfg.n_md; ++j)
                if ((rc = colloids_step())) return rc;
            if ((rc = colloids_collide())) return rc;
        }
    }
    RMPCD_HIP(hipStreamSynchronize(stream()));
    return RMPCD_OK;
}

}  // namespace rmpcd

using namespace rmpcd;

extern "C" {

int rmpcd_program_create(const rmpcd_program_config_t *cfg, rmpcd_program_t **out)
{
    RMPCD_CHECK(cfg && out, "rmpcd_program_create: null argument");
    *out = nullptr;
    ProgramSim *s = new (std::nothrow) ProgramSim();
    if (!s) return set_error(RMPCD_ERR_ARG, "rmpcd_program_create: out of host memory");
    int rc = s->setup(*cfg);
    if (rc) {
        delete s;
        return rc;
    }
    *out = (rmpcd_program_t *)s;
    return RMPCD_OK;
}

int rmpcd_program_run(rmpcd_program_t *prog, int32_t n_cycles)
{
    RMPCD_CHECK(prog && n_cycles >= 0, "rmpcd_program_run: bad argument");
    return ((ProgramSim *)prog)->run(n_cycles);
}

int rmpcd_program_state(rmpcd_program_t *prog, rmpcd_particles_t *solvent, rmpcd_particles_t *colloids,
                        rmpcd_box_t *box, rmpcd_cells_t *cells, rmpcd_program_stats_t *stats)
{
    RMPCD_CHECK(prog, "rmpcd_program_state: null program");
    ProgramSim *s = (ProgramSim *)prog;
    RMPCD_HIP(hipStreamSynchronize(stream()));
    if (solvent) *solvent = s->sv();
    if (colloids) *colloids = s->cv();
    if (box) *box = s->box;
    if (cells) *cells = s->cells();
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->n_solvent = s->n;
        stats->n_colloids = s->nc;
        stats->md_steps = s->md_steps;
        stats->cycles = s->cycles;
        stats->n_extra_sorting = s->n_extra;
        stats->e1 = s->e1;
        stats->e2 = s->e2;
        stats->temperature = s->temperature;
        stats->dt = s->dt;
        stats->mass = s->mass;
        stats->nmax = s->nmax;
        stats->last_sort = (int32_t)s->last_sort;
        stats->list = s->list;
        stats->nlist = s->nlist;
        stats->pos_old = s->alt().pos;
        stats->colloid_pos_old = s->col_pos_old;
    }
    return RMPCD_OK;
}

int rmpcd_program_destroy(rmpcd_program_t *prog)
{
    delete (ProgramSim *)prog;
    return RMPCD_OK;
}

}  // extern "C"
