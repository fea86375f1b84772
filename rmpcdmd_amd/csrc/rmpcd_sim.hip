// rmpcd_sim.hip -- the single-dimer nanomotor time loop of
// programs/single_dimer_pbc.f90 (set-up :183-296, MD step :306-383, collision cycle
// :386-443, helpers :493-565) on device-resident state, for one or many independent
// replicas.
//
// Two execution modes share the state:
//   RMPCD_SIM_REFERENCE  the reference's call sequence through the library's own C ABI
//                        entry points, one replica after the other, with the host
//                        reading the skin test every step (the plain form);
//   RMPCD_SIM_FUSED      the fused device-side step of rmpcd_fused.hip: every replica in
//                        each launch, the skin test decided on the device, one hipGraph
//                        per collision cycle.
// Random numbers come from Threefry-2x64-20 keyed by (seed + replica, stream); the
// counters are given next to each use so a test can reproduce every draw.
#include <math.h>
#include <string.h>

#include <new>
#include <vector>

#include "rmpcd_setup.hpp"
#include "rmpcd_sim.hpp"

namespace rmpcd {

// the grid shift of collision s: threefry((s, 0), (seed, SHIFT)) gives u0, u1 and
// threefry((s, 1), ...) gives u2, drawn in the order of cell_system.f90:159-161
__host__ __device__ void shift_uniforms(uint64_t seed, uint64_t s, double u[3])
{
    uint64_t a, b, c, d;
    threefry2x64(s, 0, seed, STREAM_SHIFT, a, b);
    threefry2x64(s, 1, seed, STREAM_SHIFT, c, d);
    u[0] = u01(a);
    u[1] = u01(b);
    u[2] = u01(c);
}
void shift_draws(uint64_t seed, uint64_t s, double u[3]) { shift_uniforms(seed, s, u); }

// ---- host-side state management ---------------------------------------------------
static size_t real_size(int form) { return form == RMPCD_FP64 ? 8 : 4; }
static size_t small_size(int form) { return form == RMPCD_FP64 ? 4 : 1; }

int DimerSim::alloc(void **p, size_t bytes)
{
    bytes = (bytes + 255) & ~(size_t)255;
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return set_error(RMPCD_ERR_HIP, "dimer: hipMalloc of %zu bytes failed", bytes);
    }
    allocs.push_back(*p);
    if (hipMemset(*p, 0, bytes) != hipSuccess) return set_error(RMPCD_ERR_HIP, "dimer: hipMemset failed");
    return RMPCD_OK;
}

DimerSim::~DimerSim()
{
    (void)hipDeviceSynchronize();
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (fstream) (void)hipStreamDestroy(fstream);
    for (hipStream_t st : gstreams) (void)hipStreamDestroy(st);
    for (hipEvent_t e : gevents) (void)hipEventDestroy(e);
    for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
    for (void *p : allocs) (void)hipFree(p);
}

int DimerSim::setup(const rmpcd_dimer_config_t &c, const int32_t *replica_L)
{
    cfg = c;
    R = cfg.n_replicas > 0 ? cfg.n_replicas : 1;
    RMPCD_CHECK(R <= 4096, "dimer: at most 4096 replicas");
    RMPCD_CHECK(cfg.n_md > 0 && cfg.tau > 0 && cfg.rho > 0, "dimer: N_MD, tau and rho must be positive");
    RMPCD_CHECK(cfg.form == RMPCD_FP64 || cfg.form == RMPCD_FP32, "dimer: unknown form");
    RMPCD_CHECK(cfg.mode == RMPCD_SIM_REFERENCE || cfg.mode == RMPCD_SIM_FUSED, "dimer: unknown mode");
    rep.resize(R);
    ntot = nctot = luttot = nmax_rep = ncmax_rep = 0;
    for (int r = 0; r < R; ++r) {
        const int32_t *L = replica_L ? replica_L + 3 * r : cfg.L;
        int rc = rmpcd_box_init(&rep[r].box, L, 1.0, 0);
        if (rc) return rc;
        int64_t n = (int64_t)cfg.rho * L[0] * L[1] * L[2];
        RMPCD_CHECK(n > 0 && n < ((int64_t)1 << 31), "dimer: solvent count out of range");
        rep[r].n = n;
        rep[r].p_off = ntot;
        rep[r].c_off = nctot;
        rep[r].lut_off = luttot;
        rep[r].seed = cfg.seed + (uint64_t)r;
        ntot += n;
        nctot += rep[r].box.ncells;
        luttot += (int64_t)L[0] * L[1] * L[2];
        nmax_rep = n > nmax_rep ? n : nmax_rep;
        ncmax_rep = rep[r].box.ncells > ncmax_rep ? rep[r].box.ncells : ncmax_rep;
    }
    RMPCD_CHECK(ntot < ((int64_t)1 << 31), "dimer: total solvent count exceeds int32 indexing");
    dt = cfg.tau / cfg.n_md;
    skin = cfg.skin > 0 ? cfg.skin : 2.0;  // single_dimer_pbc.f90:274
    // interaction parameters (single_dimer_pbc.f90:153-181)
    double eps[4] = {cfg.epsilon_C[0], cfg.epsilon_C[1], cfg.epsilon_N[0], cfg.epsilon_N[1]};
    double sig[4] = {cfg.sigma_C, cfg.sigma_C, cfg.sigma_N, cfg.sigma_N};
    double cut[4];
    double f6 = pow(2.0, 1.0 / 6.0);
    for (int k = 0; k < 4; ++k) cut[k] = sig[k] * f6;
    max_cut = cut[0];
    for (int k = 1; k < 4; ++k) max_cut = cut[k] > max_cut ? cut[k] : max_cut;
    int rc = rmpcd_lj_params_init(&lj_sc, 2, 2, eps, sig, cut);
    if (rc) return rc;
    double ceps[4] = {cfg.epsilon_CC, cfg.epsilon_NC, cfg.epsilon_NC, cfg.epsilon_NN};
    double csig[4] = {2 * cfg.sigma_C, cfg.sigma_C + cfg.sigma_N, cfg.sigma_C + cfg.sigma_N, 2 * cfg.sigma_N};
    double ccut[4];
    for (int k = 0; k < 4; ++k) ccut[k] = csig[k] * f6;
    if ((rc = rmpcd_lj_params_init(&lj_cc, 2, 2, ceps, csig, ccut))) return rc;
    // default-REAL pi literal 3.14159265 (single_dimer_pbc.f90:180-181)
    const double pi_f = (double)3.14159265f;
    mass[0] = (((cfg.rho * (cfg.sigma_C * cfg.sigma_C * cfg.sigma_C)) * 4) * pi_f) / 3;
    mass[1] = (((cfg.rho * (cfg.sigma_N * cfg.sigma_N * cfg.sigma_N)) * 4) * pi_f) / 3;
    radius = max_cut + skin;
    // list capacity: the reference's int(300*sigma^3) (single_dimer_pbc.f90:272), raised to
    // 1.3x the expected shell population when that would overflow
    double smax = cfg.sigma_C > cfg.sigma_N ? cfg.sigma_C : cfg.sigma_N;
    int nmax_ref = (int)(300 * smax * smax * smax);
    int nmax_est = (int)(1.3 * cfg.rho * 4.0 / 3.0 * M_PI * radius * radius * radius) + 64;
    nmax = cfg.nmax_list > 0 ? cfg.nmax_list : (nmax_ref > nmax_est ? nmax_ref : nmax_est);

    size_t RS = real_size(cfg.form), SS = small_size(cfg.form);
    for (int b = 0; b < 2; ++b) {
        if ((rc = alloc(&sol_pos[b], 3 * ntot * RS))) return rc;
        if ((rc = alloc(&sol_vel[b], 3 * ntot * RS))) return rc;
        if ((rc = alloc((void **)&sol_image[b], 3 * ntot * 4))) return rc;
        if ((rc = alloc((void **)&sol_id[b], ntot * 4))) return rc;
        if ((rc = alloc(&sol_species[b], ntot * SS))) return rc;
        if ((rc = alloc(&sol_flags[b], ntot * SS))) return rc;
        if ((rc = alloc((void **)&cell_count[b], nctot * 4))) return rc;
    }
    for (int b = 0; b < 4; ++b)
        if ((rc = alloc(&sol_force[b], 3 * ntot * RS))) return rc;
    if ((rc = alloc((void **)&id_to_idx, ntot * 4))) return rc;
    if ((rc = alloc((void **)&cell_start, nctot * 4))) return rc;
    if ((rc = alloc((void **)&is_md, nctot))) return rc;
    if ((rc = alloc((void **)&is_reac, nctot))) return rc;
    size_t c6 = 6 * (size_t)R * sizeof(double);
    if ((rc = alloc((void **)&col_pos, c6))) return rc;
    if ((rc = alloc((void **)&col_pos_old, c6))) return rc;
    if ((rc = alloc((void **)&col_pos_rattle, c6))) return rc;
    if ((rc = alloc((void **)&col_vel, c6))) return rc;
    if ((rc = alloc((void **)&col_force[0], c6))) return rc;
    if ((rc = alloc((void **)&col_force[1], c6))) return rc;
    if ((rc = alloc((void **)&col_image, 6 * (size_t)R * 4))) return rc;
    if ((rc = alloc((void **)&col_id, 2 * (size_t)R * 4))) return rc;
    if ((rc = alloc((void **)&col_species, 2 * (size_t)R * 4))) return rc;
    if ((rc = alloc((void **)&col_flags, 2 * (size_t)R * 4))) return rc;
    if ((rc = alloc((void **)&col_mass, 2 * (size_t)R * 8))) return rc;
    if ((rc = alloc((void **)&list, (size_t)2 * R * nmax * 4))) return rc;
    if ((rc = alloc((void **)&nlist, 2 * (size_t)R * 4))) return rc;
    if ((rc = alloc((void **)&scratch, 1 << 16))) return rc;
    // stencil (make_stencil(solvent_cells, max_cut+skin), single_dimer_pbc.f90:277)
    n_stencil = rmpcd_make_stencil(1.0, radius, nullptr, 0);
    std::vector<int32_t> st(3 * (size_t)n_stencil);
    rmpcd_make_stencil(1.0, radius, st.data(), n_stencil);
    if ((rc = alloc((void **)&stencil, st.size() * 4))) return rc;
    RMPCD_HIP(hipMemcpy(stencil, st.data(), st.size() * 4, hipMemcpyHostToDevice));
    for (int r = 0; r < R; ++r)
        if ((rc = initialise(r))) return rc;
    if (cfg.mode == RMPCD_SIM_FUSED) return fused_prepare(*this);
    return RMPCD_OK;
}

rmpcd_cells_t DimerSim::cells_view(int r) const
{
    rmpcd_cells_t c;
    c.cell_count = cell_count[rep[r].cnt] + rep[r].c_off;
    c.cell_start = cell_start + rep[r].c_off;
    c.is_md = is_md + rep[r].c_off;
    c.is_reac = is_reac + rep[r].c_off;
    return c;
}

static rmpcd_particles_t solvent_buffers(const DimerSim &s, int r, int b, int f)
{
    const Replica &p = s.rep[r];
    size_t RS = real_size(s.cfg.form), SS = small_size(s.cfg.form);
    rmpcd_particles_t v;
    memset(&v, 0, sizeof(v));
    v.n = p.n;
    v.form = s.cfg.form;
    v.pos = (char *)s.sol_pos[b] + 3 * p.p_off * RS;
    v.vel = (char *)s.sol_vel[b] + 3 * p.p_off * RS;
    v.force = (char *)s.sol_force[2 * b + f] + 3 * p.p_off * RS;
    v.force_old = (char *)s.sol_force[2 * b + (1 - f)] + 3 * p.p_off * RS;
    v.image = s.sol_image[b] + 3 * p.p_off;
    v.id = s.sol_id[b] + p.p_off;
    v.species = (char *)s.sol_species[b] + p.p_off * SS;
    v.flags = (char *)s.sol_flags[b] + p.p_off * SS;
    return v;
}

rmpcd_particles_t DimerSim::solvent_view(int r) const { return solvent_buffers(*this, r, rep[r].cur, rep[r].fpar); }
rmpcd_particles_t DimerSim::solvent_alt(int r) const { return solvent_buffers(*this, r, 1 - rep[r].cur, rep[r].fpar); }

rmpcd_particles_t DimerSim::colloid_view(int r) const
{
    rmpcd_particles_t p;
    memset(&p, 0, sizeof(p));
    p.n = 2;
    p.form = RMPCD_FP64;
    p.pos = col_pos + 6 * r;
    p.vel = col_vel + 6 * r;
    p.force = col_force[rep[r].cfpar] + 6 * r;
    p.force_old = col_force[1 - rep[r].cfpar] + 6 * r;
    p.image = col_image + 6 * r;
    p.id = col_id + 2 * r;
    p.species = col_species + 2 * r;
    p.flags = col_flags + 2 * r;
    return p;
}

template <class F>
static int init_solvent(DimerSim &s, int r)
{
    hipStream_t st = stream();
    rmpcd_particles_t sv = s.solvent_view(r);
    PView<F> p = view<F>(&sv);
    Replica &rp = s.rep[r];
    k_fill_state<F><<<blocks_for(rp.n, 256), 256, 0, st>>>(p);
    RMPCD_LAUNCHED();
    unsigned nb = blocks_for(rp.n, 256);
    double *part = (double *)workspace(WS_PART, sizeof(double) * 3 * (nb + 1));
    if (!part) return set_error(RMPCD_ERR_HIP, "dimer: workspace allocation failed");
    double *sum = (double *)s.scratch;
    k_init_velocities<F><<<nb, 256, 0, st>>>(p, rp.seed, sqrt(s.cfg.T), part);
    RMPCD_LAUNCHED();
    k_sum_blocks3<3><<<1, 256, 0, st>>>(part, (int)nb, sum);
    RMPCD_LAUNCHED();
    k_remove_mean<F><<<nb, 256, 0, st>>>(p, sum);
    RMPCD_LAUNCHED();
    rmpcd_particles_t cv = s.colloid_view(r);
    int32_t *fail = (int32_t *)(sum + 4);
    RMPCD_HIP(hipMemsetAsync(fail, 0, 4, st));
    k_place_random<F><<<nb, 256, 0, st>>>(p, view<F64>(&cv), 2, rp.box.edges[0], rp.box.edges[1],
                                          rp.box.edges[2], s.lj_sc.n1, s.lj_sc.cut_sq[0], s.lj_sc.cut_sq[1],
                                          s.lj_sc.cut_sq[2], s.lj_sc.cut_sq[3], rp.seed, fail);
    RMPCD_LAUNCHED();
    int32_t hf = 0;
    RMPCD_HIP(hipMemcpyAsync(&hf, fail, 4, hipMemcpyDeviceToHost, st));
    RMPCD_HIP(hipStreamSynchronize(st));
    if (hf) return set_error(RMPCD_ERR_ARG, "dimer: random placement found no free position");
    return RMPCD_OK;
}

int DimerSim::initialise(int r)
{
    hipStream_t st = stream();
    Replica &rp = rep[r];
    // colloids (single_dimer_pbc.f90:197-200, 250-252); pos_old starts at zero
    double cp[6];
    for (int k = 0; k < 3; ++k) {
        cp[2 * k] = rp.box.edges[k] / 2.0;
        cp[2 * k + 1] = rp.box.edges[k] / 2.0;
    }
    cp[1] = cp[1] + cfg.d;  // pos(1,2) += d
    int32_t sp[2] = {1, 2}, ids[2] = {0, 1};
    RMPCD_HIP(hipMemcpy(col_pos + 6 * r, cp, sizeof(cp), hipMemcpyHostToDevice));
    RMPCD_HIP(hipMemcpy(col_species + 2 * r, sp, sizeof(sp), hipMemcpyHostToDevice));
    RMPCD_HIP(hipMemcpy(col_id + 2 * r, ids, sizeof(ids), hipMemcpyHostToDevice));
    // the drivers index colloids%mass by the particle (single_dimer_pbc.f90:314)
    RMPCD_HIP(hipMemcpy(col_mass + 2 * r, mass, sizeof(mass), hipMemcpyHostToDevice));
    int rc = cfg.form == RMPCD_FP64 ? init_solvent<F64>(*this, r) : init_solvent<F32>(*this, r);
    if (rc) return rc;
    // sort, neighbour list, first forces (single_dimer_pbc.f90:270-296)
    if ((rc = sort_solvent(r, false))) return rc;
    if ((rc = build_list(r))) return rc;
    rmpcd_particles_t sv = solvent_view(r), cv = colloid_view(r);
    if ((rc = rmpcd_compute_force(&cv, &sv, list + 2 * (size_t)r * nmax, nlist + 2 * r, nmax, rp.box.edges,
                                  &lj_sc, &rp.e1)))
        return rc;
    if ((rc = rmpcd_compute_force_n2(&cv, rp.box.edges, &lj_cc, &rp.e2))) return rc;
    // force_old = force (single_dimer_pbc.f90:295-296)
    RMPCD_HIP(hipMemcpyAsync(sv.force_old, sv.force, 3 * rp.n * real_size(cfg.form), hipMemcpyDeviceToDevice, st));
    RMPCD_HIP(hipMemcpyAsync(cv.force_old, cv.force, 6 * sizeof(double), hipMemcpyDeviceToDevice, st));
    RMPCD_HIP(hipStreamSynchronize(st));
    return RMPCD_OK;
}

int DimerSim::sort_solvent(int r, bool apply_pbc)
{
    rmpcd_particles_t src = solvent_view(r), dst = solvent_alt(r);
    rmpcd_cells_t c = cells_view(r);
    int rc = rmpcd_sort(&rep[r].box, &src, &dst, id_to_idx + rep[r].p_off, &c, apply_pbc ? 1 : 0);
    if (rc) return rc;
    rep[r].cur = 1 - rep[r].cur;  // the switch() calls of particle_system.f90:346-353
    return RMPCD_OK;
}

int DimerSim::build_list(int r)
{
    rmpcd_particles_t sv = solvent_view(r), cv = colloid_view(r);
    rmpcd_cells_t c = cells_view(r);
    return rmpcd_update_list(&rep[r].box, &c, &cv, &sv, radius, stencil, n_stencil, &lj_sc, nmax,
                             list + 2 * (size_t)r * nmax, nlist + 2 * r);
}

// one MD step in the reference's call sequence (single_dimer_pbc.f90:306-383)
int DimerSim::md_step_reference(int r)
{
    hipStream_t st = stream();
    Replica &rp = rep[r];
    size_t RS = real_size(cfg.form);
    int rc;
    int32_t *lst = list + 2 * (size_t)r * nmax, *nl = nlist + 2 * r;
    double *cpos_old = col_pos_old + 6 * r, *cpos_rattle = col_pos_rattle + 6 * r;
    rmpcd_particles_t sv = solvent_view(r), cv = colloid_view(r);
    if ((rc = rmpcd_md_pos(&sv, dt))) return rc;
    RMPCD_HIP(hipMemcpyAsync(cpos_rattle, cv.pos, 6 * sizeof(double), hipMemcpyDeviceToDevice, st));
    if ((rc = rmpcd_colloid_md_pos(&cv, col_mass + 2 * r, dt))) return rc;
    if ((rc = rmpcd_rattle_dimer_pos(&cv, cpos_rattle, mass, 2, cfg.d, dt, rp.box.edges))) return rc;
    double so_max, co_max;
    if ((rc = rmpcd_maximum_displacement(&sv, solvent_alt(r).pos, &so_max))) return rc;
    if ((rc = rmpcd_maximum_displacement(&cv, cpos_old, &co_max))) return rc;
    if (co_max >= skin * 0.1 || so_max >= skin * 0.89) {
        if ((rc = rmpcd_apply_pbc(&cv, rp.box.edges))) return rc;
        if ((rc = sort_solvent(r, true))) return rc;
        if ((rc = build_list(r))) return rc;
        sv = solvent_view(r);
        RMPCD_HIP(hipMemcpyAsync(solvent_alt(r).pos, sv.pos, 3 * rp.n * RS, hipMemcpyDeviceToDevice, st));
        RMPCD_HIP(hipMemcpyAsync(cpos_old, cv.pos, 6 * sizeof(double), hipMemcpyDeviceToDevice, st));
        ++rp.n_extra;
    }
    // switch(force, force_old) for both systems, then zero the new forces
    rp.fpar = 1 - rp.fpar;
    rp.cfpar = 1 - rp.cfpar;
    sv = solvent_view(r);
    cv = colloid_view(r);
    RMPCD_HIP(hipMemsetAsync(sv.force, 0, 3 * rp.n * RS, st));
    RMPCD_HIP(hipMemsetAsync(cv.force, 0, 6 * sizeof(double), st));
    if ((rc = rmpcd_compute_force(&cv, &sv, lst, nl, nmax, rp.box.edges, &lj_sc, &rp.e1))) return rc;
    if ((rc = rmpcd_compute_force_n2(&cv, rp.box.edges, &lj_cc, &rp.e2))) return rc;
    if ((rc = rmpcd_md_vel(&sv, dt))) return rc;
    if ((rc = rmpcd_colloid_md_vel(&cv, col_mass + 2 * r, dt))) return rc;
    if ((rc = rmpcd_rattle_dimer_vel(&cv, mass, 2, cfg.d, dt, rp.box.edges))) return rc;
    if ((rc = rmpcd_flag_particles_nl(&sv, &cv, lst, nl, nmax, rp.box.edges, lj_sc.cut_sq[0]))) return rc;
    // change_species draws threefry((md_step, id), (seed, CHANGE))
    if ((rc = rmpcd_change_species(&sv, &cv, rp.box.edges, lj_sc.cut_sq[0], lj_sc.cut_sq[2], cfg.prob, nullptr,
                                   rp.seed, (uint64_t)rp.md_steps)))
        return rc;
    ++rp.md_steps;
    return RMPCD_OK;
}

// the collision part of a cycle (single_dimer_pbc.f90:405-428)
int DimerSim::collide_reference(int r)
{
    hipStream_t st = stream();
    Replica &rp = rep[r];
    size_t RS = real_size(cfg.form);
    int rc;
    double u[3];
    shift_draws(rp.seed, (uint64_t)rp.cycles, u);
    if ((rc = rmpcd_random_shift(&rp.box, u))) return rc;
    rmpcd_particles_t cv = colloid_view(r);
    if ((rc = rmpcd_apply_pbc(&cv, rp.box.edges))) return rc;
    if ((rc = sort_solvent(r, true))) return rc;
    if ((rc = build_list(r))) return rc;
    rmpcd_particles_t sv = solvent_view(r);
    rmpcd_cells_t c = cells_view(r);
    RMPCD_HIP(hipMemcpyAsync(solvent_alt(r).pos, sv.pos, 3 * rp.n * RS, hipMemcpyDeviceToDevice, st));
    RMPCD_HIP(hipMemcpyAsync(col_pos_old + 6 * r, cv.pos, 6 * sizeof(double), hipMemcpyDeviceToDevice, st));
    // SRD axes: Marsaglia on threefry((cycle, cell + 2^40 k), (seed, SRD)) pairs
    if ((rc = rmpcd_simple_mpcd_step(&rp.box, &c, &sv, nullptr, rp.seed, (uint64_t)rp.cycles, 0, 0.0, 0, cfg.T, 1)))
        return rc;
    if (cfg.bulk_rmpcd) {
        // one draw per eligible cell: threefry((cycle, cell), (seed, REACTION))
        if ((rc = rmpcd_bulk_reaction(&rp.box, &c, &sv, 2, 1, cfg.bulk_rate, cfg.tau, nullptr, rp.seed,
                                      (uint64_t)rp.cycles)))
            return rc;
    } else {
        if ((rc = rmpcd_refuel(&sv, &cv, rp.box.edges, rp.box.L[0]))) return rc;
    }
    ++rp.cycles;
    return RMPCD_OK;
}

int DimerSim::run(int n_cycles)
{
    if (cfg.mode == RMPCD_SIM_FUSED) return run_fused_cycles(*this, n_cycles);
    for (int c = 0; c < n_cycles; ++c) {
        for (int r = 0; r < R; ++r) {
            int rc;
            for (int j = 0; j < cfg.n_md; ++j)
                if ((rc = md_step_reference(r))) return rc;
            if ((rc = collide_reference(r))) return rc;
        }
    }
    RMPCD_HIP(hipStreamSynchronize(stream()));
    return RMPCD_OK;
}

}  // namespace rmpcd

using namespace rmpcd;

extern "C" {

int rmpcd_dimer_create(const rmpcd_dimer_config_t *cfg, rmpcd_dimer_t **out)
{
    RMPCD_CHECK(cfg && out, "rmpcd_dimer_create: null argument");
    *out = nullptr;
    DimerSim *s = new (std::nothrow) DimerSim();
    if (!s) return set_error(RMPCD_ERR_ARG, "rmpcd_dimer_create: out of host memory");
    int rc = s->setup(*cfg, cfg->replica_L);
    if (rc) {
        delete s;
        return rc;
    }
    *out = (rmpcd_dimer_t *)s;
    return RMPCD_OK;
}

int rmpcd_dimer_run(rmpcd_dimer_t *sim, int32_t n_cycles)
{
    RMPCD_CHECK(sim && n_cycles >= 0, "rmpcd_dimer_run: bad argument");
    return ((DimerSim *)sim)->run(n_cycles);
}

int rmpcd_dimer_state(rmpcd_dimer_t *sim, int32_t replica, rmpcd_particles_t *solvent,
                      rmpcd_particles_t *colloids, rmpcd_box_t *box, rmpcd_cells_t *cells,
                      rmpcd_dimer_stats_t *stats)
{
    RMPCD_CHECK(sim, "rmpcd_dimer_state: null simulation");
    DimerSim *s = (DimerSim *)sim;
    RMPCD_CHECK(replica >= 0 && replica < s->R, "rmpcd_dimer_state: replica out of range");
    // a replica stopped by a failed list build still shows its state; the status it
    // stopped with is in stats->error (rmpcd_dimer_run returned it)
    int rc = s->sync_state();
    if (rc == RMPCD_ERR_HIP || rc == RMPCD_ERR_ARG) return rc;
    int r = replica;
    const Replica &rp = s->rep[r];
    if (solvent) *solvent = s->solvent_view(r);
    if (colloids) *colloids = s->colloid_view(r);
    if (box) *box = rp.box;
    if (cells) *cells = s->cells_view(r);
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->n_solvent = rp.n;
        stats->md_steps = rp.md_steps;
        stats->cycles = rp.cycles;
        stats->n_extra_sorting = rp.n_extra;
        stats->e1 = rp.e1;
        stats->e2 = rp.e2;
        stats->nmax = s->nmax;
        stats->error = rp.error;
        stats->list = s->list + 2 * (size_t)r * s->nmax;
        stats->nlist = s->nlist + 2 * r;
        stats->pos_old = s->cfg.mode == RMPCD_SIM_FUSED
                             ? (void *)((char *)s->sol_pos_old + 3 * rp.p_off * real_size(s->cfg.form))
                             : s->solvent_alt(r).pos;
        stats->colloid_pos_old = s->col_pos_old + 6 * r;
        stats->mass[0] = s->mass[0];
        stats->mass[1] = s->mass[1];
        stats->dt = s->dt;
        stats->n_replicas = s->R;
        stats->n_solvent_total = s->ntot;
    }
    return RMPCD_OK;
}

int rmpcd_dimer_profile(rmpcd_dimer_t *sim, int32_t stride)
{
    RMPCD_CHECK(sim && stride >= 0, "rmpcd_dimer_profile: bad argument");
    return fused_set_profile(*(DimerSim *)sim, stride);
}

int rmpcd_dimer_stream_time(rmpcd_dimer_t *sim, double *avg_ms, int32_t *samples)
{
    RMPCD_CHECK(sim && avg_ms && samples, "rmpcd_dimer_stream_time: null argument");
    return fused_stream_time(*(DimerSim *)sim, avg_ms, samples);
}

int rmpcd_dimer_destroy(rmpcd_dimer_t *sim)
{
    delete (DimerSim *)sim;
    return RMPCD_OK;
}

}  // extern "C"
