// rmpcd_fused.hip -- fused device-side MD step and collision cycle of the single-dimer
// program (placeholder until the fused kernels land).
#include "rmpcd_sim.hpp"

namespace rmpcd {

int run_fused_cycle(DimerSim &s)
{
    (void)s;
    return set_error(RMPCD_ERR_ARG, "dimer: the fused mode is not available in this build");
}

int DimerSim::sync_state() { return RMPCD_OK; }

}  // namespace rmpcd
