// rmpcd_programs.hpp -- device-resident state of the driver programs other than the
// dimer's (test/setup_simple_fluid.f90, test/setup_single_catalytic_fixed_sphere.f90,
// programs/n_colloids_pbc.f90), shared by the reference call sequence
// (rmpcd_programs.hip) and the graph-captured COLLOIDS loop (rmpcd_colloids.hip).
#pragma once
#include <math.h>
#include <string.h>

#include <vector>

#include "rmpcd_setup.hpp"
#include "rmpcd_sim.hpp"

namespace rmpcd {

struct ColloidsGraph;
struct ProgramSim;
// the graph-captured COLLOIDS loop (rmpcd_colloids.hip): created on the first run
int colloids_graph_run(ProgramSim &s, int n_cycles);
void colloids_graph_free(ColloidsGraph *g);

inline size_t rsize(int form) { return form == RMPCD_FP64 ? 8 : 4; }
inline size_t ssize(int form) { return form == RMPCD_FP64 ? 4 : 1; }

struct ProgramSim {
    rmpcd_program_config_t cfg{};
    rmpcd_box_t box{};
    int64_t n = 0;
    int nc = 0;
    int cur = 0, fpar = 0, cfpar = 0;
    int64_t md_steps = 0, cycles = 0, n_extra = 0, last_sort = 0;
    double e1 = 0, e2 = 0, temperature = 0, dt = 0, skin = 0, mass = 0, radius = 0;
    int nmax = 0, n_stencil = 0;
    uint64_t seed = 0;
    rmpcd_lj_t lj_sc{}, lj_cc{};
    void *sol_pos[2] = {}, *sol_vel[2] = {}, *sol_species[2] = {}, *sol_flags[2] = {};
    int32_t *sol_image[2] = {}, *sol_id[2] = {};
    void *sol_force[4] = {};
    int32_t *id_to_idx = nullptr, *cell_count = nullptr, *cell_start = nullptr;
    uint8_t *is_md = nullptr, *is_reac = nullptr;
    double *col_pos = nullptr, *col_pos_old = nullptr, *col_vel = nullptr, *col_force[2] = {};
    int32_t *col_image = nullptr, *col_id = nullptr, *col_species = nullptr, *col_flags = nullptr;
    double *col_mass = nullptr, *scratch = nullptr;
    int32_t *list = nullptr, *nlist = nullptr, *stencil = nullptr;
    std::vector<void *> allocs;
    ColloidsGraph *graph = nullptr;

    ~ProgramSim()
    {
        colloids_graph_free(graph);
        (void)hipDeviceSynchronize();
        for (void *p : allocs) (void)hipFree(p);
    }
    int alloc(void **p, size_t bytes)
    {
        bytes = ((bytes ? bytes : 1) + 255) & ~(size_t)255;
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            return set_error(RMPCD_ERR_HIP, "program: hipMalloc of %zu bytes failed", bytes);
        }
        allocs.push_back(*p);
        if (hipMemset(*p, 0, bytes) != hipSuccess) return set_error(RMPCD_ERR_HIP, "program: hipMemset failed");
        return RMPCD_OK;
    }

    rmpcd_particles_t solvent(int b, int f) const
    {
        rmpcd_particles_t v;
        memset(&v, 0, sizeof(v));
        v.n = n;
        v.form = cfg.form;
        v.pos = sol_pos[b];
        v.vel = sol_vel[b];
        v.force = sol_force[2 * b + f];
        v.force_old = sol_force[2 * b + 1 - f];
        v.image = sol_image[b];
        v.id = sol_id[b];
        v.species = sol_species[b];
        v.flags = sol_flags[b];
        return v;
    }
    rmpcd_particles_t sv() const { return solvent(cur, fpar); }
    rmpcd_particles_t alt() const { return solvent(1 - cur, fpar); }  // holds pos_old
    rmpcd_particles_t cv() const
    {
        rmpcd_particles_t p;
        memset(&p, 0, sizeof(p));
        p.n = nc;
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
    rmpcd_cells_t cells() const
    {
        rmpcd_cells_t c;
        c.cell_count = cell_count;
        c.cell_start = cell_start;
        c.is_md = is_md;
        c.is_reac = is_reac;
        return c;
    }

    int setup(const rmpcd_program_config_t &c);
    int place_colloids();
    int init_solvent();
    int sort(bool pbc)
    {
        rmpcd_particles_t s = sv(), d = alt();
        rmpcd_cells_t c = cells();
        int rc = rmpcd_sort(&box, &s, &d, id_to_idx, &c, pbc ? 1 : 0);
        if (rc) return rc;
        cur = 1 - cur;  // the switch() calls of particle_system.f90:346-353
        return RMPCD_OK;
    }
    int build_list()
    {
        rmpcd_particles_t s = sv(), p = cv();
        rmpcd_cells_t c = cells();
        // neither driver passes lj to update_list, so every cell stays reactive
        return rmpcd_update_list(&box, &c, &p, &s, radius, stencil, n_stencil, nullptr, nmax, list, nlist);
    }
    int copy_pos_old()  // pos_old = pos for both systems
    {
        hipStream_t st = stream();
        RMPCD_HIP(hipMemcpyAsync(alt().pos, sv().pos, 3 * n * rsize(cfg.form), hipMemcpyDeviceToDevice, st));
        if (nc) RMPCD_HIP(hipMemcpyAsync(col_pos_old, col_pos, 3 * nc * sizeof(double), hipMemcpyDeviceToDevice, st));
        return RMPCD_OK;
    }
    int forces(bool zero_colloid)
    {
        hipStream_t st = stream();
        // switch(force, force_old) for both systems, then zero the new forces
        fpar = 1 - fpar;
        cfpar = 1 - cfpar;
        rmpcd_particles_t s = sv(), p = cv();
        RMPCD_HIP(hipMemsetAsync(s.force, 0, 3 * n * rsize(cfg.form), st));
        if (zero_colloid) RMPCD_HIP(hipMemsetAsync(p.force, 0, 3 * nc * sizeof(double), st));
        int rc = rmpcd_compute_force(&p, &s, list, nlist, nmax, box.edges, &lj_sc, &e1);
        if (!rc && cfg.kind == RMPCD_PROG_COLLOIDS) rc = rmpcd_compute_force_n2(&p, box.edges, &lj_cc, &e2);
        return rc;
    }
    int shift(uint64_t s)
    {
        double u[3];
        shift_draws(seed, s, u);
        return rmpcd_random_shift(&box, u);
    }

    int fluid_cycle();
    int sphere_step();
    int sphere_collide();
    int colloids_step();
    int colloids_collide();
    int run(int n_cycles);
};

}  // namespace rmpcd
