// rmpcd_sim.hip -- the single-dimer nanomotor time loop of
// programs/single_dimer_pbc.f90 (set-up :183-296, MD step :306-383, collision cycle
// :386-443, helpers :493-565) on device-resident state.
//
// Two execution modes share the state:
//   RMPCD_SIM_REFERENCE  the reference's call sequence through the library's own C ABI
//                        entry points, with the host reading the skin test every step
//                        (the plain form);
//   RMPCD_SIM_FUSED      the fused device-side step of rmpcd_fused.hip (no host round
//                        trip per step, hipGraph-captured cycles).
// Random numbers come from Threefry-2x64-20 keyed by (seed, stream); the counters are
// listed next to each use so that a test can reproduce every draw on the host.
#include <math.h>
#include <string.h>

#include <new>
#include <vector>

#include "rmpcd_internal.hpp"
#include "rmpcd_sim.hpp"

namespace rmpcd {

// ---- set-up kernels --------------------------------------------------------------
// Maxwell-Boltzmann velocities: particle id draws threefry((0, id), (seed, INIT)) and
// threefry((1, id), (seed, INIT)); Box-Muller on the first two pairs gives 3 normals
__device__ inline void normals3(uint64_t seed, uint64_t id, double out[3])
{
    uint64_t a, b, c, d;
    threefry2x64(0, id, seed, STREAM_INIT, a, b);
    threefry2x64(1, id, seed, STREAM_INIT, c, d);
    double u1 = 1.0 - u01(a), u2 = u01(b), u3 = 1.0 - u01(c), u4 = u01(d);
    double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    out[0] = r1 * cos(2 * M_PI * u2);
    out[1] = r1 * sin(2 * M_PI * u2);
    out[2] = r2 * cos(2 * M_PI * u4);
}

template <class F>
__global__ void k_init_velocities(PView<F> p, uint64_t seed, double sqrtT, double *__restrict__ sum)
{
    __shared__ double red[16];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[3] = {0, 0, 0};
    if (i < p.n) {
        normals3(seed, (uint64_t)i, v);
        for (int k = 0; k < 3; ++k) {
            v[k] *= sqrtT;
            p.vel[k * p.n + i] = (typename F::real)v[k];
            v[k] = (double)p.vel[k * p.n + i];
        }
    }
    for (int k = 0; k < 3; ++k) {
        double s = block_sum(v[k], red);
        if (threadIdx.x == 0) atomicAdd(&sum[k], s);
    }
}

template <class F>
__global__ void k_remove_mean(PView<F> p, const double *__restrict__ sum)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    for (int k = 0; k < 3; ++k)
        p.vel[k * p.n + i] = (typename F::real)((double)p.vel[k * p.n + i] - sum[k] / (double)p.n);
}

// random_placement (particle_system.f90:250-302) with the colloids as obstacles;
// attempt k of particle i uses threefry((k, i), (seed, PLACE)) and threefry((k, i+2^40), ...)
template <class F>
__global__ void k_place(PView<F> p, PView<F64> obst, int n_obst, double L0, double L1, double L2,
                        LJCut cut, uint64_t seed, int32_t *__restrict__ fail)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    int s1 = (int)p.species[i];
    double x[3] = {0, 0, 0};
    bool ok = false;
    for (uint64_t k = 0; k < 4096 && !ok; ++k) {
        uint64_t a, b, c, d;
        threefry2x64(k, (uint64_t)i, seed, STREAM_PLACE, a, b);
        threefry2x64(k, (uint64_t)i + (1ull << 40), seed, STREAM_PLACE, c, d);
        x[0] = u01(a) * L0;
        x[1] = u01(b) * L1;
        x[2] = u01(c) * L2;
        ok = true;
        for (int j = 0; j < n_obst && ok; ++j) {
            double d0 = min_image(x[0] - obst.pos[j], L0);
            double d1 = min_image(x[1] - obst.pos[obst.n + j], L1);
            double d2 = min_image(x[2] - obst.pos[2 * obst.n + j], L2);
            double rsq = (d0 * d0 + d1 * d1) + d2 * d2;
            int s2 = obst.species[j];
            if (rsq < cut.cut_sq[(s1 - 1) + cut.n1 * (s2 - 1)]) ok = false;
        }
    }
    if (!ok) atomicOr(fail, 1);
    // positions are stored in the particle precision; a float rounding up to L wraps
    for (int k = 0; k < 3; ++k) {
        typename F::real v = (typename F::real)x[k];
        double e = k == 0 ? L0 : k == 1 ? L1 : L2;
        if ((double)v >= e) v = (typename F::real)0;
        p.pos[k * p.n + i] = v;
    }
}

template <class F>
__global__ void k_fill_state(PView<F> p)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    for (int k = 0; k < 3; ++k) {
        p.force[k * p.n + i] = 0;
        p.force_old[k * p.n + i] = 0;
        p.image[k * p.n + i] = 0;
    }
    p.id[i] = (int32_t)i;
    p.species[i] = (typename F::small)1;
    p.flags[i] = (typename F::small)0;
}

// three uniforms of the grid shift of collision s: threefry((s, 0), (seed, SHIFT)) gives
// u0, u1 and threefry((s, 1), ...) gives u2 (the reference draws three doubles in order,
// cell_system.f90:159-161)
void shift_draws(uint64_t seed, uint64_t s, double u[3])
{
    uint64_t a, b, c, d;
    threefry2x64(s, 0, seed, STREAM_SHIFT, a, b);
    threefry2x64(s, 1, seed, STREAM_SHIFT, c, d);
    u[0] = u01(a);
    u[1] = u01(b);
    u[2] = u01(c);
}

// ---- host-side state management ---------------------------------------------------
static size_t real_size(int form) { return form == RMPCD_FP64 ? 8 : 4; }
static size_t small_size(int form) { return form == RMPCD_FP64 ? 4 : 1; }

int DimerSim::alloc(void **p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) return set_error(RMPCD_ERR_HIP, "dimer: hipMalloc of %zu bytes failed", bytes);
    allocs.push_back(*p);
    if (hipMemset(*p, 0, bytes) != hipSuccess) return set_error(RMPCD_ERR_HIP, "dimer: hipMemset failed");
    return RMPCD_OK;
}

DimerSim::~DimerSim()
{
    (void)hipDeviceSynchronize();
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    for (void *p : allocs) (void)hipFree(p);
}

int DimerSim::setup(const rmpcd_dimer_config_t &c)
{
    cfg = c;
    int rc = rmpcd_box_init(&box, cfg.L, 1.0, 0);
    if (rc) return rc;
    n = (int64_t)cfg.rho * cfg.L[0] * cfg.L[1] * cfg.L[2];
    RMPCD_CHECK(n > 0 && n < ((int64_t)1 << 31), "dimer: solvent count out of range");
    RMPCD_CHECK(cfg.n_md > 0 && cfg.tau > 0, "dimer: N_MD and tau must be positive");
    RMPCD_CHECK(cfg.form == RMPCD_FP64 || cfg.form == RMPCD_FP32, "dimer: unknown form");
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
    rc = rmpcd_lj_params_init(&lj_sc, 2, 2, eps, sig, cut);
    if (rc) return rc;
    double ceps[4] = {cfg.epsilon_CC, cfg.epsilon_NC, cfg.epsilon_NC, cfg.epsilon_NN};
    double csig[4] = {2 * cfg.sigma_C, cfg.sigma_C + cfg.sigma_N, cfg.sigma_C + cfg.sigma_N, 2 * cfg.sigma_N};
    double ccut[4];
    for (int k = 0; k < 4; ++k) ccut[k] = csig[k] * f6;
    rc = rmpcd_lj_params_init(&lj_cc, 2, 2, ceps, csig, ccut);
    if (rc) return rc;
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

    size_t R = real_size(cfg.form), S = small_size(cfg.form);
    size_t nc = (size_t)box.ncells;
    for (int b = 0; b < 2; ++b) {
        if ((rc = alloc(&sol_pos[b], 3 * n * R))) return rc;
        if ((rc = alloc(&sol_vel[b], 3 * n * R))) return rc;
        if ((rc = alloc((void **)&sol_image[b], 3 * n * 4))) return rc;
        if ((rc = alloc((void **)&sol_id[b], n * 4))) return rc;
        if ((rc = alloc(&sol_species[b], n * S))) return rc;
        if ((rc = alloc(&sol_flags[b], n * S))) return rc;
    }
    for (int b = 0; b < 4; ++b)
        if ((rc = alloc(&sol_force[b], 3 * n * R))) return rc;
    if ((rc = alloc((void **)&id_to_idx, n * 4))) return rc;
    if ((rc = alloc((void **)&claim, n * 4))) return rc;
    if ((rc = alloc((void **)&cells.cell_count, nc * 4))) return rc;
    if ((rc = alloc((void **)&cells.cell_start, nc * 4))) return rc;
    if ((rc = alloc((void **)&cells.is_md, nc))) return rc;
    if ((rc = alloc((void **)&cells.is_reac, nc))) return rc;
    if ((rc = alloc((void **)&col_pos, 3 * 2 * 8))) return rc;
    if ((rc = alloc((void **)&col_pos_old, 3 * 2 * 8))) return rc;
    if ((rc = alloc((void **)&col_pos_rattle, 3 * 2 * 8))) return rc;
    if ((rc = alloc((void **)&col_vel, 3 * 2 * 8))) return rc;
    if ((rc = alloc((void **)&col_force[0], 3 * 2 * 8))) return rc;
    if ((rc = alloc((void **)&col_force[1], 3 * 2 * 8))) return rc;
    if ((rc = alloc((void **)&col_image, 3 * 2 * 4))) return rc;
    if ((rc = alloc((void **)&col_id, 2 * 4))) return rc;
    if ((rc = alloc((void **)&col_species, 2 * 4))) return rc;
    if ((rc = alloc((void **)&col_flags, 2 * 4))) return rc;
    if ((rc = alloc((void **)&col_mass, 2 * 8))) return rc;
    if ((rc = alloc((void **)&list, (size_t)2 * nmax * 4))) return rc;
    if ((rc = alloc((void **)&nlist, 2 * 4))) return rc;
    if ((rc = alloc((void **)&scratch, 4096))) return rc;
    // stencil (make_stencil(solvent_cells, max_cut+skin), single_dimer_pbc.f90:277)
    n_stencil = rmpcd_make_stencil(box.a, radius, nullptr, 0);
    std::vector<int32_t> st(3 * (size_t)n_stencil);
    rmpcd_make_stencil(box.a, radius, st.data(), n_stencil);
    if ((rc = alloc((void **)&stencil, st.size() * 4))) return rc;
    RMPCD_HIP(hipMemcpy(stencil, st.data(), st.size() * 4, hipMemcpyHostToDevice));
    return initialise();
}

rmpcd_particles_t DimerSim::solvent_view() const
{
    rmpcd_particles_t p;
    memset(&p, 0, sizeof(p));
    p.n = n;
    p.form = cfg.form;
    p.pos = sol_pos[cur];
    p.vel = sol_vel[cur];
    p.force = sol_force[2 * cur + fpar];
    p.force_old = sol_force[2 * cur + (1 - fpar)];
    p.image = sol_image[cur];
    p.id = sol_id[cur];
    p.species = sol_species[cur];
    p.flags = sol_flags[cur];
    return p;
}

rmpcd_particles_t DimerSim::solvent_alt() const
{
    rmpcd_particles_t p = solvent_view();
    int o = 1 - cur;
    p.pos = sol_pos[o];
    p.vel = sol_vel[o];
    p.force = sol_force[2 * o + fpar];
    p.force_old = sol_force[2 * o + (1 - fpar)];
    p.image = sol_image[o];
    p.id = sol_id[o];
    p.species = sol_species[o];
    p.flags = sol_flags[o];
    return p;
}

rmpcd_particles_t DimerSim::colloid_view() const
{
    rmpcd_particles_t p;
    memset(&p, 0, sizeof(p));
    p.n = 2;
    p.form = RMPCD_FP64;
    p.pos = col_pos;
    p.vel = col_vel;
    p.force = col_force[cfpar];
    p.force_old = col_force[1 - cfpar];
    p.image = col_image;
    p.id = col_id;
    p.species = col_species;
    p.flags = col_flags;
    return p;
}

template <class F>
static int init_solvent(DimerSim &s)
{
    hipStream_t st = stream();
    rmpcd_particles_t sv = s.solvent_view();
    PView<F> p = view<F>(&sv);
    k_fill_state<F><<<blocks_for(s.n, 256), 256, 0, st>>>(p);
    RMPCD_LAUNCHED();
    double *sum = (double *)s.scratch;
    RMPCD_HIP(hipMemsetAsync(sum, 0, 4 * sizeof(double), st));
    k_init_velocities<F><<<blocks_for(s.n, 256), 256, 0, st>>>(p, s.cfg.seed, sqrt(s.cfg.T), sum);
    RMPCD_LAUNCHED();
    k_remove_mean<F><<<blocks_for(s.n, 256), 256, 0, st>>>(p, sum);
    RMPCD_LAUNCHED();
    rmpcd_particles_t cv = s.colloid_view();
    LJCut cut;
    cut.n1 = s.lj_sc.n1;
    for (int k = 0; k < 16; ++k) cut.cut_sq[k] = s.lj_sc.cut_sq[k];
    int32_t *fail = (int32_t *)(sum + 4);
    RMPCD_HIP(hipMemsetAsync(fail, 0, 4, st));
    k_place<F><<<blocks_for(s.n, 256), 256, 0, st>>>(p, view<F64>(&cv), 2, s.box.edges[0], s.box.edges[1],
                                                    s.box.edges[2], cut, s.cfg.seed, fail);
    RMPCD_LAUNCHED();
    int32_t hf = 0;
    RMPCD_HIP(hipMemcpyAsync(&hf, fail, 4, hipMemcpyDeviceToHost, st));
    RMPCD_HIP(hipStreamSynchronize(st));
    if (hf) return set_error(RMPCD_ERR_ARG, "dimer: random placement found no free position");
    return RMPCD_OK;
}

int DimerSim::initialise()
{
    hipStream_t st = stream();
    // colloids (single_dimer_pbc.f90:197-200, 250-252)
    double cp[6], zero6[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        cp[2 * k] = box.edges[k] / 2.0;
        cp[2 * k + 1] = box.edges[k] / 2.0;
    }
    cp[1] = cp[1] + cfg.d;  // pos(1,2) += d
    int32_t sp[2] = {1, 2}, ids[2] = {0, 1};
    RMPCD_HIP(hipMemcpy(col_pos, cp, sizeof(cp), hipMemcpyHostToDevice));
    RMPCD_HIP(hipMemcpy(col_pos_old, zero6, sizeof(zero6), hipMemcpyHostToDevice));
    RMPCD_HIP(hipMemcpy(col_species, sp, sizeof(sp), hipMemcpyHostToDevice));
    RMPCD_HIP(hipMemcpy(col_id, ids, sizeof(ids), hipMemcpyHostToDevice));
    // the drivers index colloids%mass by the particle (single_dimer_pbc.f90:314)
    RMPCD_HIP(hipMemcpy(col_mass, mass, sizeof(mass), hipMemcpyHostToDevice));
    int rc = cfg.form == RMPCD_FP64 ? init_solvent<F64>(*this) : init_solvent<F32>(*this);
    if (rc) return rc;
    // sort, neighbour list, first forces (single_dimer_pbc.f90:270-296)
    if ((rc = sort_solvent(false))) return rc;
    if ((rc = build_list())) return rc;
    rmpcd_particles_t sv = solvent_view(), cv = colloid_view();
    if ((rc = rmpcd_compute_force(&cv, &sv, list, nlist, nmax, box.edges, &lj_sc, &e1))) return rc;
    if ((rc = rmpcd_compute_force_n2(&cv, box.edges, &lj_cc, &e2))) return rc;
    // force_old = force
    RMPCD_HIP(hipMemcpyAsync(sv.force_old, sv.force, 3 * n * real_size(cfg.form), hipMemcpyDeviceToDevice, st));
    RMPCD_HIP(hipMemcpyAsync(cv.force_old, cv.force, 6 * sizeof(double), hipMemcpyDeviceToDevice, st));
    RMPCD_HIP(hipStreamSynchronize(st));
    return RMPCD_OK;
}

int DimerSim::sort_solvent(bool apply_pbc)
{
    rmpcd_particles_t src = solvent_view(), dst = solvent_alt();
    int rc = rmpcd_sort(&box, &src, &dst, id_to_idx, &cells, apply_pbc ? 1 : 0);
    if (rc) return rc;
    cur = 1 - cur;  // the switch() calls of particle_system.f90:346-353
    return RMPCD_OK;
}

int DimerSim::build_list()
{
    rmpcd_particles_t sv = solvent_view(), cv = colloid_view();
    return rmpcd_update_list(&box, &cells, &cv, &sv, radius, stencil, n_stencil, &lj_sc, nmax, list, nlist);
}

// one MD step in the reference's call sequence (single_dimer_pbc.f90:306-383)
int DimerSim::md_step_reference()
{
    hipStream_t st = stream();
    size_t R = real_size(cfg.form);
    int rc;
    rmpcd_particles_t sv = solvent_view(), cv = colloid_view();
    if ((rc = rmpcd_md_pos(&sv, dt))) return rc;
    RMPCD_HIP(hipMemcpyAsync(col_pos_rattle, col_pos, 6 * sizeof(double), hipMemcpyDeviceToDevice, st));
    if ((rc = rmpcd_colloid_md_pos(&cv, col_mass, dt))) return rc;
    if ((rc = rmpcd_rattle_dimer_pos(&cv, col_pos_rattle, mass, 2, cfg.d, dt, box.edges))) return rc;
    double so_max, co_max;
    if ((rc = rmpcd_maximum_displacement(&sv, sol_pos[1 - cur], &so_max))) return rc;
    if ((rc = rmpcd_maximum_displacement(&cv, col_pos_old, &co_max))) return rc;
    if (co_max >= skin * 0.1 || so_max >= skin * 0.89) {
        if ((rc = rmpcd_apply_pbc(&cv, box.edges))) return rc;
        if ((rc = sort_solvent(true))) return rc;
        if ((rc = build_list())) return rc;
        sv = solvent_view();
        RMPCD_HIP(hipMemcpyAsync(sol_pos[1 - cur], sv.pos, 3 * n * R, hipMemcpyDeviceToDevice, st));
        RMPCD_HIP(hipMemcpyAsync(col_pos_old, col_pos, 6 * sizeof(double), hipMemcpyDeviceToDevice, st));
        ++n_extra_sorting;
    }
    // switch(force, force_old) for both systems, then zero the new forces
    fpar = 1 - fpar;
    cfpar = 1 - cfpar;
    sv = solvent_view();
    cv = colloid_view();
    RMPCD_HIP(hipMemsetAsync(sv.force, 0, 3 * n * R, st));
    RMPCD_HIP(hipMemsetAsync(cv.force, 0, 6 * sizeof(double), st));
    if ((rc = rmpcd_compute_force(&cv, &sv, list, nlist, nmax, box.edges, &lj_sc, &e1))) return rc;
    if ((rc = rmpcd_compute_force_n2(&cv, box.edges, &lj_cc, &e2))) return rc;
    if ((rc = rmpcd_md_vel(&sv, dt))) return rc;
    if ((rc = rmpcd_colloid_md_vel(&cv, col_mass, dt))) return rc;
    if ((rc = rmpcd_rattle_dimer_vel(&cv, mass, 2, cfg.d, dt, box.edges))) return rc;
    if ((rc = rmpcd_flag_particles_nl(&sv, &cv, list, nlist, nmax, box.edges, lj_sc.cut_sq[0]))) return rc;
    if ((rc = rmpcd_change_species(&sv, &cv, box.edges, lj_sc.cut_sq[0], lj_sc.cut_sq[2], cfg.prob, nullptr,
                                   cfg.seed, (uint64_t)md_steps)))
        return rc;
    ++md_steps;
    return RMPCD_OK;
}

// the collision part of a cycle (single_dimer_pbc.f90:405-428)
int DimerSim::collide_reference()
{
    hipStream_t st = stream();
    size_t R = real_size(cfg.form);
    int rc;
    double u[3];
    shift_draws(cfg.seed, (uint64_t)cycles, u);
    if ((rc = rmpcd_random_shift(&box, u))) return rc;
    rmpcd_particles_t cv = colloid_view();
    if ((rc = rmpcd_apply_pbc(&cv, box.edges))) return rc;
    if ((rc = sort_solvent(true))) return rc;
    if ((rc = build_list())) return rc;
    rmpcd_particles_t sv = solvent_view();
    RMPCD_HIP(hipMemcpyAsync(sol_pos[1 - cur], sv.pos, 3 * n * R, hipMemcpyDeviceToDevice, st));
    RMPCD_HIP(hipMemcpyAsync(col_pos_old, col_pos, 6 * sizeof(double), hipMemcpyDeviceToDevice, st));
    if ((rc = rmpcd_simple_mpcd_step(&box, &cells, &sv, nullptr, cfg.seed, (uint64_t)cycles, 0, 0.0, 0, cfg.T, 1)))
        return rc;
    if (cfg.bulk_rmpcd) {
        if ((rc = rmpcd_bulk_reaction(&box, &cells, &sv, 2, 1, cfg.bulk_rate, cfg.tau, nullptr, cfg.seed,
                                      (uint64_t)cycles)))
            return rc;
    } else {
        if ((rc = rmpcd_refuel(&sv, &cv, box.edges, cfg.L[0]))) return rc;
    }
    ++cycles;
    return RMPCD_OK;
}

int DimerSim::run(int n_cycles)
{
    for (int c = 0; c < n_cycles; ++c) {
        int rc;
        if (cfg.mode == RMPCD_SIM_FUSED) {
            rc = run_fused_cycle(*this);
        } else {
            for (int j = 0; j < cfg.n_md; ++j)
                if ((rc = md_step_reference())) return rc;
            rc = collide_reference();
        }
        if (rc) return rc;
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
    int rc = s->setup(*cfg);
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

int rmpcd_dimer_state(rmpcd_dimer_t *sim, rmpcd_particles_t *solvent, rmpcd_particles_t *colloids,
                      rmpcd_box_t *box, rmpcd_cells_t *cells, rmpcd_dimer_stats_t *stats)
{
    RMPCD_CHECK(sim, "rmpcd_dimer_state: null simulation");
    DimerSim *s = (DimerSim *)sim;
    int rc = s->sync_state();
    if (rc) return rc;
    if (solvent) *solvent = s->solvent_view();
    if (colloids) *colloids = s->colloid_view();
    if (box) *box = s->box;
    if (cells) *cells = s->cells;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->n_solvent = s->n;
        stats->md_steps = s->md_steps;
        stats->cycles = s->cycles;
        stats->n_extra_sorting = s->n_extra_sorting;
        stats->e1 = s->e1;
        stats->e2 = s->e2;
        stats->nmax = s->nmax;
        stats->list = s->list;
        stats->nlist = s->nlist;
        stats->pos_old = s->sol_pos[1 - s->cur];
        stats->colloid_pos_old = s->col_pos_old;
        stats->mass[0] = s->mass[0];
        stats->mass[1] = s->mass[1];
        stats->dt = s->dt;
    }
    return RMPCD_OK;
}

int rmpcd_dimer_destroy(rmpcd_dimer_t *sim)
{
    delete (DimerSim *)sim;
    return RMPCD_OK;
}

}  // extern "C"
