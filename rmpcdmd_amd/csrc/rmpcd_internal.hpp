// rmpcd_internal.hpp -- host-side runtime shared by the translation units of the
// library: error reporting, the launch stream, the grow-only device workspace and
// argument checks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>

#include "rmpcd_capi.h"
#include "rmpcd_device.hpp"

namespace rmpcd {

int set_error(int code, const char *fmt, ...);
hipStream_t stream();

// device workspace: named slots that grow on demand and live until process exit
enum Slot {
    WS_KEY = 0,     // per-particle cell key
    WS_SLOT,        // per-particle arrival slot in its cell
    WS_TMP,         // per-slot particle index before stabilisation
    WS_SCAN,        // scan tile sums
    WS_RED,         // reductions (doubles)
    WS_FLAG,        // status flags (ints)
    WS_LUT,         // Hilbert table of the current box
    WS_NLCNT,       // neighbour-list per-stencil-cell counts
    WS_PART,        // per-block partial sums
    WS_AUX,         // misc
    WS_COUNT
};
void *workspace(Slot s, size_t bytes);
// host-pinned staging for synchronous scalar results
void *pinned(size_t bytes);

// the current device's address of the fixed scratch array of
// rmpcd_compute_temperature_noalloc (looked up on the device's first use)
int noalloc_scratch(double **out);

// Hilbert table of a box (cached by box shape)
const int32_t *hilbert_lut(const Box &b);

#define RMPCD_HIP(call)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::rmpcd::set_error(RMPCD_ERR_HIP, "%s failed: %s", #call,             \
                                      hipGetErrorString(e_));                            \
    } while (0)

#define RMPCD_LAUNCHED()                                                                 \
    do {                                                                                 \
        hipError_t e_ = hipGetLastError();                                               \
        if (e_ != hipSuccess)                                                            \
            return ::rmpcd::set_error(RMPCD_ERR_HIP, "kernel launch failed: %s",         \
                                      hipGetErrorString(e_));                            \
    } while (0)

#define RMPCD_CHECK(cond, ...)                                                           \
    do {                                                                                 \
        if (!(cond)) return ::rmpcd::set_error(RMPCD_ERR_ARG, __VA_ARGS__);              \
    } while (0)

int check_particles(const rmpcd_particles_t *p, const char *who, bool need_force);
int check_box(const rmpcd_box_t *b, const char *who);

inline unsigned blocks_for(int64_t n, int threads)
{
    int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b < 1 ? 1 : b);
}

// dispatch on the numeric form of a particle set
#define RMPCD_DISPATCH_FORM(p, FN, ...)                                                  \
    ((p)->form == RMPCD_FP64 ? FN<::rmpcd::F64>(__VA_ARGS__)                             \
                             : FN<::rmpcd::F32>(__VA_ARGS__))

}  // namespace rmpcd
