// rmpcd_runtime.hip -- runtime of the library (errors, stream, workspace) and the
// host-side entry points that compute scalars: cell_system%init, random_shift,
// lj_params_init, make_stencil, the Hilbert key on the host and its device table.
#include <math.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "rmpcd_internal.hpp"

namespace rmpcd {

static thread_local char g_err[512] = "";
static hipStream_t g_stream = nullptr;

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

hipStream_t stream() { return g_stream; }

struct WsEntry {
    void *ptr = nullptr;
    size_t bytes = 0;
};
static WsEntry g_ws[WS_COUNT];
static void *g_pinned = nullptr;
static size_t g_pinned_bytes = 0;
static std::mutex g_ws_mutex;

void *workspace(Slot s, size_t bytes)
{
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    WsEntry &e = g_ws[s];
    if (bytes <= e.bytes && e.ptr) return e.ptr;
    if (e.ptr) {
        // a grow must not free memory that queued kernels still use
        if (hipStreamSynchronize(g_stream) != hipSuccess) return nullptr;
        (void)hipFree(e.ptr);
        e.ptr = nullptr;
        e.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    want = (want + want / 4 + 255) & ~(size_t)255;
    if (hipMalloc(&e.ptr, want) != hipSuccess) {
        e.ptr = nullptr;
        return nullptr;
    }
    e.bytes = want;
    return e.ptr;
}

void *pinned(size_t bytes)
{
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    if (bytes <= g_pinned_bytes && g_pinned) return g_pinned;
    if (g_pinned) {
        if (hipStreamSynchronize(g_stream) != hipSuccess) return nullptr;
        (void)hipHostFree(g_pinned);
        g_pinned = nullptr;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    if (hipHostMalloc(&g_pinned, want, hipHostMallocDefault) != hipSuccess) {
        g_pinned = nullptr;
        g_pinned_bytes = 0;
        return nullptr;
    }
    g_pinned_bytes = want;
    return g_pinned;
}

int check_particles(const rmpcd_particles_t *p, const char *who, bool need_force)
{
    RMPCD_CHECK(p != nullptr, "%s: null particle set", who);
    RMPCD_CHECK(p->n > 0, "%s: particle count must be positive (got %lld)", who,
                (long long)p->n);
    RMPCD_CHECK(p->n < (int64_t)1 << 31, "%s: particle count exceeds int32 indexing", who);
    RMPCD_CHECK(p->form == RMPCD_FP64 || p->form == RMPCD_FP32, "%s: unknown form %d", who,
                p->form);
    RMPCD_CHECK(p->pos && p->vel && p->species && p->flags, "%s: null particle array", who);
    if (need_force) RMPCD_CHECK(p->force && p->force_old, "%s: null force array", who);
    return RMPCD_OK;
}

int check_box(const rmpcd_box_t *b, const char *who)
{
    RMPCD_CHECK(b != nullptr, "%s: null box", who);
    for (int k = 0; k < 3; ++k) {
        RMPCD_CHECK(b->L[k] >= 1 && b->M[k] >= 1 && b->M[k] <= 10, "%s: bad box axis %d", who, k);
        RMPCD_CHECK((1 << b->M[k]) >= b->L[k] - (k == 2 && b->has_walls ? 1 : 0),
                    "%s: M too small for L on axis %d", who, k);
    }
    RMPCD_CHECK(b->ncells == (1 << (b->M[0] + b->M[1] + b->M[2])), "%s: ncells != 2^sum(M)", who);
    RMPCD_CHECK(b->a > 0, "%s: cell size must be positive", who);
    return RMPCD_OK;
}

// ---------------------------------------------------------------------------------
__global__ void k_hilbert_table(Box b, int32_t *keys)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n = (int64_t)b.L[0] * b.L[1] * b.L[2];
    if (t >= n) return;
    int p0 = (int)(t % b.L[0]);
    int p1 = (int)((t / b.L[0]) % b.L[1]);
    int p2 = (int)(t / ((int64_t)b.L[0] * b.L[1]));
    keys[t] = compact_key(p0, p1, p2, b.M[0], b.M[1], b.M[2]);
}

static int g_lut_shape[6] = {0, 0, 0, 0, 0, 0};

const int32_t *hilbert_lut(const Box &b)
{
    int64_t n = (int64_t)b.L[0] * b.L[1] * b.L[2];
    int32_t *lut = (int32_t *)workspace(WS_LUT, sizeof(int32_t) * n);
    if (!lut) return nullptr;
    int shape[6] = {b.L[0], b.L[1], b.L[2], b.M[0], b.M[1], b.M[2]};
    if (memcmp(shape, g_lut_shape, sizeof(shape)) != 0) {
        k_hilbert_table<<<blocks_for(n, 256), 256, 0, g_stream>>>(b, lut);
        if (hipGetLastError() != hipSuccess) return nullptr;
        memcpy(g_lut_shape, shape, sizeof(shape));
    }
    return lut;
}

}  // namespace rmpcd

using namespace rmpcd;

extern "C" {

const char *rmpcd_last_error(void) { return g_err; }

const char *rmpcd_version(void) { return "rmpcdmd-amd 0.1 (gfx950)"; }

int rmpcd_init(int device)
{
    RMPCD_HIP(hipSetDevice(device));
    RMPCD_HIP(hipFree(nullptr));
    // the device's address of the fixed scratch array of the allocation-free entry point
    double *scratch = nullptr;
    return noalloc_scratch(&scratch);
}

int rmpcd_set_stream(void *s)
{
    // the scratch buffers are shared by all entry points: the kernels queued on the
    // previous stream must be done with them before the new stream can reuse them
    if ((hipStream_t)s != g_stream) RMPCD_HIP(hipStreamSynchronize(g_stream));
    g_stream = (hipStream_t)s;
    return RMPCD_OK;
}

int rmpcd_synchronize(void)
{
    RMPCD_HIP(hipStreamSynchronize(g_stream));
    return RMPCD_OK;
}

int rmpcd_malloc(void **ptr, size_t bytes)
{
    RMPCD_CHECK(ptr != nullptr, "rmpcd_malloc: null argument");
    *ptr = nullptr;
    RMPCD_HIP(hipMalloc(ptr, bytes < 256 ? 256 : (bytes + 255) & ~(size_t)255));
    return RMPCD_OK;
}

int rmpcd_free(void *ptr)
{
    if (ptr) RMPCD_HIP(hipFree(ptr));
    return RMPCD_OK;
}

int rmpcd_memcpy(void *dst, const void *src, size_t bytes, int32_t kind)
{
    RMPCD_CHECK((dst && src) || bytes == 0, "rmpcd_memcpy: null pointer");
    RMPCD_CHECK(kind >= 0 && kind <= 2, "rmpcd_memcpy: kind must be 0, 1 or 2");
    if (bytes == 0) return RMPCD_OK;
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    RMPCD_HIP(hipDeviceSynchronize());
    RMPCD_HIP(hipMemcpy(dst, src, bytes, k));
    return RMPCD_OK;
}

int32_t rmpcd_compact_p_to_h(const int32_t p[3], const int32_t M[3])
{
    return compact_key(p[0], p[1], p[2], M[0], M[1], M[2]);
}

void rmpcd_compact_h_to_p(int32_t h, const int32_t M[3], int32_t p[3])
{
    int q[3];
    compact_point(h, M[0], M[1], M[2], q);
    p[0] = q[0];
    p[1] = q[1];
    p[2] = q[2];
}

int rmpcd_hilbert_table(const rmpcd_box_t *box, int32_t *keys)
{
    int rc = check_box(box, "rmpcd_hilbert_table");
    if (rc) return rc;
    RMPCD_CHECK(keys != nullptr, "rmpcd_hilbert_table: null output");
    Box b = make_box(box);
    int64_t n = (int64_t)b.L[0] * b.L[1] * b.L[2];
    k_hilbert_table<<<blocks_for(n, 256), 256, 0, g_stream>>>(b, keys);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

int rmpcd_box_init(rmpcd_box_t *box, const int32_t L[3], double a, int32_t has_walls)
{
    // cell_system%init (src/cell_system.f90:49-84)
    RMPCD_CHECK(box != nullptr && L != nullptr, "rmpcd_box_init: null argument");
    for (int k = 0; k < 3; ++k)
        RMPCD_CHECK(L[k] >= 1 && L[k] <= 1024, "rmpcd_box_init: L(%d) = %d out of [1,1024]", k + 1,
                    L[k]);
    RMPCD_CHECK(a > 0, "rmpcd_box_init: cell size must be positive");
    memset(box, 0, sizeof(*box));
    for (int k = 0; k < 3; ++k) {
        box->L[k] = L[k];
        box->edges[k] = L[k] * a;
        box->bc[k] = RMPCD_PERIODIC_BC;
        int m = 1;
        while ((1 << m) < L[k]) ++m;  // the bound is the argument L (cell_system.f90:73)
        box->M[k] = m;
    }
    box->has_walls = has_walls ? 1 : 0;
    if (box->has_walls) box->L[2] = L[2] + 1;
    box->a = a;
    RMPCD_CHECK(box->M[0] + box->M[1] + box->M[2] <= 30,
                "rmpcd_box_init: 2^sum(M) exceeds the 32-bit Hilbert key");
    box->ncells = 1 << (box->M[0] + box->M[1] + box->M[2]);
    return RMPCD_OK;
}

int rmpcd_random_shift(rmpcd_box_t *box, const double u[3])
{
    // cell_system%random_shift (src/cell_system.f90:153-163)
    RMPCD_CHECK(box != nullptr && u != nullptr, "rmpcd_random_shift: null argument");
    for (int k = 0; k < 3; ++k) box->origin[k] = u[k] - 1;
    return RMPCD_OK;
}

int rmpcd_lj_params_init(rmpcd_lj_t *lj, int32_t n1, int32_t n2, const double *eps,
                         const double *sigma, const double *cut)
{
    // lj_params_init (src/interaction.f90:38-80)
    RMPCD_CHECK(lj && eps && sigma && cut, "rmpcd_lj_params_init: null argument");
    RMPCD_CHECK(n1 >= 1 && n2 >= 1 && n1 * n2 <= 16,
                "unequal sizes for arguments to lj_params_init");
    memset(lj, 0, sizeof(*lj));
    lj->n1 = n1;
    lj->n2 = n2;
    for (int i = 0; i < n1 * n2; ++i) {
        lj->eps[i] = eps[i];
        lj->sigma[i] = sigma[i];
        lj->cut[i] = cut[i];
        lj->cut_sq[i] = cut[i] * cut[i];
        lj->cut_energy[i] = lj_energy(lj->cut_sq[i], eps[i], sigma[i], sigma[i] * 0);
    }
    return RMPCD_OK;
}

int32_t rmpcd_make_stencil(double a, double cut, int32_t *stencil, int32_t cap)
{
    // make_stencil / closest (src/neighbor_list.f90:271-322): offsets (i,j,k) in
    // [-m,m]^3, m = floor(cut/a)+1, whose nearest cell-to-cell distance is below cut,
    // ordered with i slowest and k fastest
    if (!(a > 0) || !(cut >= 0)) return set_error(RMPCD_ERR_ARG, "rmpcd_make_stencil: bad a/cut");
    int m = (int)floor(cut / a) + 1;
    int32_t count = 0;
    for (int i = -m; i <= m; ++i)
        for (int j = -m; j <= m; ++j)
            for (int k = -m; k <= m; ++k) {
                int di = abs(i) - 1, dj = abs(j) - 1, dk = abs(k) - 1;
                di = di < 0 ? 0 : di;
                dj = dj < 0 ? 0 : dj;
                dk = dk < 0 ? 0 : dk;
                if (sqrt((double)(di * di + dj * dj + dk * dk)) < cut) {
                    if (stencil && count < cap) {
                        stencil[3 * count] = i;
                        stencil[3 * count + 1] = j;
                        stencil[3 * count + 2] = k;
                    }
                    ++count;
                }
            }
    return count;
}

void rmpcd_threefry2x64_20(const uint64_t ctr[2], const uint64_t key[2], uint64_t out[2])
{
    threefry2x64(ctr[0], ctr[1], key[0], key[1], out[0], out[1]);
}

}  // extern "C"
