// rmpcd_sim.hpp -- device-resident state of the single-dimer program for R independent
// replicas, shared by the reference-sequence mode (rmpcd_sim.hip) and the fused mode
// (rmpcd_fused.hip).  Replica r owns a contiguous block of every array: solvent planes
// [3][n_r] at 3*p_off_r, scalars at p_off_r, cells at c_off_r, two colloids at 2r,
// two neighbour lists of nmax entries at 2r*nmax.
#pragma once
#include <vector>

#include "rmpcd_internal.hpp"

namespace rmpcd {

void shift_draws(uint64_t seed, uint64_t s, double u[3]);

struct Replica {
    rmpcd_box_t box{};
    int64_t n = 0, p_off = 0, c_off = 0, lut_off = 0;
    uint64_t seed = 0;
    int cur = 0, fpar = 0, cfpar = 0, cnt = 0;  // buffer parities, live cell_count array
    int64_t md_steps = 0, cycles = 0, n_extra = 0;
    double e1 = 0, e2 = 0;
    int error = 0;  // fused mode: the status the replica stopped with (0: running)
};

// per-replica constants and mutable control for the fused kernels (device memory)
struct RepDev {
    Box box;
    int64_t n, p_off, c_off, lut_off;
    uint64_t seed;
    int32_t vec;    // planes and membership bytes allow 16-byte accesses
    int32_t defer;  // an extra sort may be deferred (the box holds the stencil box)
};
struct CtlDev {
    int32_t cur, fpar, cfpar, need_sort;
    int32_t need_exact, k_since, error, cnt;  // cnt: which cell_count array is live
    int32_t cur_next, cnt_next;               // a sort in progress: the buffers it writes
    int32_t virt, def;   // virt: an extra sort since the last collision sort left the
                         // particles in place (the reference order is in perm/inv);
                         // def: the last extra sort was deferred (its cells are in dkey)
    int32_t kind, ccount;  // the sort in progress is deferred; its candidate count
    int64_t md_step, cycle, n_extra;
    double origin[3];
    double vmax2, list_max2, co_max, e1, e2;
    uint32_t done[8];
};

struct DimerSim {
    rmpcd_dimer_config_t cfg{};
    std::vector<Replica> rep;
    int R = 1;
    int64_t ntot = 0, nctot = 0, luttot = 0, nmax_rep = 0, ncmax_rep = 0;
    rmpcd_lj_t lj_sc{}, lj_cc{};
    double mass[2] = {0, 0};
    double dt = 0, skin = 2, max_cut = 0, radius = 0;
    int nmax = 0, n_stencil = 0;
    int stream_blocks = 0, sort_blocks = 0, extra_blocks = 0, sort_order = 2, collide_cells = 128;  // fused-mode launch shape
    // solvent double buffers (sort) and the four force buffers: for sort buffer b and
    // force parity f, force is sol_force[2*b + f] and force_old sol_force[2*b + 1 - f]
    void *sol_pos[2] = {}, *sol_vel[2] = {}, *sol_species[2] = {}, *sol_flags[2] = {};
    int32_t *sol_image[2] = {}, *sol_id[2] = {};
    void *sol_force[4] = {};
    void *sol_pos_old = nullptr;  // fused mode: positions at the last sort
    int32_t *id_to_idx = nullptr;
    uint8_t *memb = nullptr;  // fused mode: bit c set for a member of colloid c's list
    int32_t *cell_count[2] = {}, *cell_start = nullptr;
    uint8_t *is_md = nullptr, *is_reac = nullptr;
    double *col_pos = nullptr, *col_pos_old = nullptr, *col_pos_rattle = nullptr, *col_vel = nullptr;
    double *col_force[2] = {};
    int32_t *col_image = nullptr, *col_id = nullptr, *col_species = nullptr, *col_flags = nullptr;
    double *col_mass = nullptr;
    int32_t *list = nullptr, *nlist = nullptr, *stencil = nullptr;
    int32_t *lut = nullptr;
    void *f_rec = nullptr;  // fused mode: staging records of a sort
    int32_t *f_perm = nullptr, *f_inv = nullptr, *f_pair = nullptr;  // fused mode: extra sorts
    int32_t *f_dkey = nullptr, *f_cand = nullptr;  // fused mode: deferred extra sorts
    void *f_stage = nullptr;
    int64_t cand_cap = 0;
    int stencil_m = 0;
    int32_t *f_key = nullptr, *f_slot = nullptr, *scan_sums = nullptr, *f_slocal = nullptr;
    int32_t *f_mslots = nullptr, *f_mcount = nullptr;
    int64_t *f_plan = nullptr;
    double *f_gpart = nullptr;
    double *part = nullptr;
    void *scratch = nullptr;

    // fused mode
    RepDev *rep_dev = nullptr;
    CtlDev *ctl_dev = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    hipStream_t fstream = nullptr;
    // replica groups, each a branch of the captured cycle; groups after the first are
    // captured from their own streams, forked from and joined to fstream
    int groups = 1;
    std::vector<hipStream_t> gstreams;
    std::vector<hipEvent_t> gevents;
    bool fused_ready = false;
    int prof_stride = 0, prof_pairs = 0;  // event nodes around the streaming kernel
    std::vector<hipEvent_t> prof_ev;
    std::vector<void *> allocs;

    ~DimerSim();
    int alloc(void **p, size_t bytes);
    int setup(const rmpcd_dimer_config_t &c, const int32_t *replica_L);
    int initialise(int r);
    int sort_solvent(int r, bool apply_pbc);
    int build_list(int r);
    int md_step_reference(int r);
    int collide_reference(int r);
    int run(int n_cycles);
    int sync_state();
    rmpcd_cells_t cells_view(int r) const;
    rmpcd_particles_t solvent_view(int r) const;
    rmpcd_particles_t solvent_alt(int r) const;
    rmpcd_particles_t colloid_view(int r) const;
};

int fused_prepare(DimerSim &s);
int run_fused_cycles(DimerSim &s, int n_cycles);
int fused_set_profile(DimerSim &s, int stride);
int fused_stream_time(DimerSim &s, double *avg_ms, int32_t *samples);

}  // namespace rmpcd
