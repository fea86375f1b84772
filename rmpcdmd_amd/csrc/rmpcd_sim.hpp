// rmpcd_sim.hpp -- device-resident state of the single-dimer program, shared by the
// reference-sequence mode (rmpcd_sim.hip) and the fused mode (rmpcd_fused.hip).
#pragma once
#include <vector>

#include "rmpcd_internal.hpp"

namespace rmpcd {

struct LJCut {
    int n1;
    double cut_sq[16];
};

void shift_draws(uint64_t seed, uint64_t s, double u[3]);

struct DimerSim {
    rmpcd_dimer_config_t cfg{};
    rmpcd_box_t box{};
    rmpcd_cells_t cells{};
    rmpcd_lj_t lj_sc{}, lj_cc{};
    double mass[2] = {0, 0};
    double dt = 0, skin = 2, max_cut = 0, radius = 0;
    int64_t n = 0;
    int nmax = 0, n_stencil = 0;
    // solvent double buffers (sort) and the four force buffers: force of sort buffer b and
    // force parity f is sol_force[2*b + f], its force_old sol_force[2*b + 1 - f]
    void *sol_pos[2] = {}, *sol_vel[2] = {}, *sol_species[2] = {}, *sol_flags[2] = {};
    int32_t *sol_image[2] = {}, *sol_id[2] = {};
    void *sol_force[4] = {};
    int32_t *id_to_idx = nullptr, *claim = nullptr;
    double *col_pos = nullptr, *col_pos_old = nullptr, *col_pos_rattle = nullptr, *col_vel = nullptr;
    double *col_force[2] = {};
    int32_t *col_image = nullptr, *col_id = nullptr, *col_species = nullptr, *col_flags = nullptr;
    double *col_mass = nullptr;
    int32_t *list = nullptr, *nlist = nullptr, *stencil = nullptr;
    void *scratch = nullptr;
    int cur = 0, fpar = 0, cfpar = 0;
    int64_t md_steps = 0, cycles = 0, n_extra_sorting = 0;
    double e1 = 0, e2 = 0;
    // fused mode
    void *dev_ctl = nullptr;   // device control block
    hipGraphExec_t graph_exec = nullptr;
    bool ctl_ready = false;
    std::vector<void *> allocs;

    ~DimerSim();
    int alloc(void **p, size_t bytes);
    int setup(const rmpcd_dimer_config_t &c);
    int initialise();
    int sort_solvent(bool apply_pbc);
    int build_list();
    int md_step_reference();
    int collide_reference();
    int run(int n_cycles);
    int sync_state();
    rmpcd_particles_t solvent_view() const;
    rmpcd_particles_t solvent_alt() const;
    rmpcd_particles_t colloid_view() const;
};

int run_fused_cycle(DimerSim &s);

}  // namespace rmpcd
