// rmpcd_sort.hip -- cell keys, cell counts, the cell_start scan and the stable
// counting sort of the solvent (src/cell_system.f90:86-129,
// src/particle_system.f90:304-355, src/md.f90:202-217 fused).
//
// Sort pipeline (all launches on the library stream):
//   k_key_count  one thread per particle: optional periodic wrap, cell key from the
//                box's Hilbert table, arrival slot from an atomic increment of the
//                cell's count (the arrival order is not deterministic)
//   scan         exclusive scan of the counts -> cell_start (wave DPP scan per tile,
//                one block for the tile sums, then the tile offsets)
//   k_place      tmp[cell_start[key] + slot] = particle
//   k_gather     one thread per destination slot: the particle's rank among the
//                cell's entries (by original index) fixes its stable position; it
//                copies the particle there.  The result equals the reference's
//                single-thread order, where each cell keeps increasing original index.
#include "rmpcd_internal.hpp"

namespace rmpcd {

// cell key of a position given in the particle's precision; the periodic wrap of
// apply_pbc (md.f90:209-214) is applied first when wrap != 0
template <class R>
__device__ inline void wrap_pbc(const Box &b, R &x, int &jump, int k)
{
    R e = (R)b.edges[k];
    jump = (int)floor((double)(x / e));
    x = x - (R)jump * e;
}

template <class R>
__device__ inline int key_of(const Box &b, const int32_t *lut, R x, R y, R z)
{
    int p0, p1, p2;
    cell_of<R>(b, x, y, z, p0, p1, p2);
    if ((unsigned)p0 < (unsigned)b.L[0] && (unsigned)p1 < (unsigned)b.L[1] &&
        (unsigned)p2 < (unsigned)b.L[2])
        return lut[p0 + b.L[0] * (p1 + b.L[1] * p2)];
    // outside the box (the reference keys it anyway, through the low bits of p)
    return compact_key(p0, p1, p2, b.M[0], b.M[1], b.M[2]);
}

template <class F>
__global__ void k_key_count(Box b, const int32_t *__restrict__ lut, PView<F> p, int wrap,
                            int32_t *__restrict__ count, int32_t *__restrict__ key,
                            int32_t *__restrict__ slot)
{
    using R = typename F::real;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    R x = p.pos[i], y = p.pos[p.n + i], z = p.pos[2 * p.n + i];
    if (wrap) {
        int j;
        wrap_pbc<R>(b, x, j, 0);
        wrap_pbc<R>(b, y, j, 1);
        wrap_pbc<R>(b, z, j, 2);
    }
    int k = key_of<R>(b, lut, x, y, z);
    int s = atomicAdd(&count[k], 1);
    if (key) {
        key[i] = k;
        slot[i] = s;
    }
}

__global__ void k_fill_i32(int32_t *a, int64_t n, int32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---- exclusive scan of int32 over n (n up to 2^30) ------------------------------
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// block-wide exclusive scan of one value per thread; returns the block total in *total
__device__ inline int block_exclusive_scan(int v, int *lds_waves, int *total)
{
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int inc = wave_inclusive_scan(v);
    if (lane == 63) lds_waves[wid] = inc;
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
        int t = lds_waves[w];
        if (w < wid) off += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// phase 1: per tile exclusive scan (in place into out) and the tile's total
__global__ void __launch_bounds__(SCAN_THREADS)
    k_scan_tiles(const int32_t *__restrict__ in, int32_t *__restrict__ out, int64_t n,
                 int32_t *__restrict__ tile_sums)
{
    __shared__ int lds[SCAN_THREADS / 64];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int tot;
    int off = block_exclusive_scan(s, lds, &tot);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) out[base + k] = off;
        off += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// phase 2: one block scans the tile totals (any count, in chunks of blockDim)
__global__ void __launch_bounds__(1024) k_scan_top(int32_t *tile_sums, int64_t ntiles)
{
    __shared__ int lds[16];
    int carry = 0;
    for (int64_t c = 0; c < ntiles; c += blockDim.x) {
        int64_t i = c + threadIdx.x;
        int v = i < ntiles ? tile_sums[i] : 0;
        int tot;
        int ex = block_exclusive_scan(v, lds, &tot);
        if (i < ntiles) tile_sums[i] = ex + carry;
        carry += tot;
        __syncthreads();
    }
}

// phase 3: add each tile's offset
__global__ void __launch_bounds__(SCAN_THREADS)
    k_scan_add(int32_t *__restrict__ out, int64_t n, const int32_t *__restrict__ tile_sums)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int add = tile_sums[blockIdx.x];
    for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_THREADS)
        if (base + k < n) out[base + k] += add;
}

int exclusive_scan(const int32_t *in, int32_t *out, int64_t n)
{
    int64_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    int32_t *sums = (int32_t *)workspace(WS_SCAN, sizeof(int32_t) * (ntiles + 1));
    if (!sums) return set_error(RMPCD_ERR_HIP, "scan: workspace allocation failed");
    hipStream_t s = stream();
    k_scan_tiles<<<(unsigned)ntiles, SCAN_THREADS, 0, s>>>(in, out, n, sums);
    RMPCD_LAUNCHED();
    k_scan_top<<<1, 1024, 0, s>>>(sums, ntiles);
    RMPCD_LAUNCHED();
    k_scan_add<<<(unsigned)ntiles, SCAN_THREADS, 0, s>>>(out, n, sums);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

// ---- counting sort ---------------------------------------------------------------
__global__ void k_place(int64_t n, const int32_t *__restrict__ key, const int32_t *__restrict__ slot,
                        const int32_t *__restrict__ start, int32_t *__restrict__ tmp)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tmp[start[key[i]] + slot[i]] = (int32_t)i;
}

template <class F>
__global__ void k_gather(Box b, PView<F> src, PView<F> dst, int wrap,
                         const int32_t *__restrict__ key, const int32_t *__restrict__ start,
                         const int32_t *__restrict__ count, const int32_t *__restrict__ tmp,
                         int32_t *__restrict__ id_to_idx)
{
    using R = typename F::real;
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n = src.n;
    if (j >= n) return;
    int32_t v = tmp[j];
    int c = key[v];
    int s = start[c], m = count[c];
    int rank = 0;
    for (int q = s; q < s + m; ++q) rank += tmp[q] < v ? 1 : 0;
    int64_t d = s + rank;
    // copy particle v to slot d (wrap applied to pos/image when requested)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        R x = src.pos[k * n + v];
        int32_t im = src.image[k * n + v];
        if (wrap) {
            int jump;
            wrap_pbc<R>(b, x, jump, k);
            im += jump;
        }
        dst.pos[k * n + d] = x;
        dst.image[k * n + d] = im;
        dst.vel[k * n + d] = src.vel[k * n + v];
        dst.force[k * n + d] = src.force[k * n + v];
        dst.force_old[k * n + d] = src.force_old[k * n + v];
    }
    int32_t id = src.id[v];
    dst.id[d] = id;
    dst.species[d] = src.species[v];
    dst.flags[d] = src.flags[v];
    if ((uint32_t)id < (uint32_t)n) id_to_idx[id] = (int32_t)d;
}

template <class F>
int count_impl(const rmpcd_box_t *box, const rmpcd_particles_t *pp, rmpcd_cells_t *cells,
               int wrap, int32_t *key, int32_t *slot)
{
    Box b = make_box(box);
    PView<F> p = view<F>(pp);
    const int32_t *lut = hilbert_lut(b);
    if (!lut) return set_error(RMPCD_ERR_HIP, "count: Hilbert table allocation failed");
    hipStream_t s = stream();
    k_fill_i32<<<blocks_for(b.ncells, 256), 256, 0, s>>>(cells->cell_count, b.ncells, 0);
    RMPCD_LAUNCHED();
    k_key_count<F><<<blocks_for(p.n, 256), 256, 0, s>>>(b, lut, p, wrap, cells->cell_count,
                                                          key, slot);
    RMPCD_LAUNCHED();
    return exclusive_scan(cells->cell_count, cells->cell_start, b.ncells);
}

template <class F>
int sort_impl(const rmpcd_box_t *box, const rmpcd_particles_t *src, rmpcd_particles_t *dst,
              int32_t *id_to_idx, rmpcd_cells_t *cells, int wrap)
{
    int64_t n = src->n;
    int32_t *key = (int32_t *)workspace(WS_KEY, sizeof(int32_t) * n);
    int32_t *slot = (int32_t *)workspace(WS_SLOT, sizeof(int32_t) * n);
    int32_t *tmp = (int32_t *)workspace(WS_TMP, sizeof(int32_t) * n);
    if (!key || !slot || !tmp) return set_error(RMPCD_ERR_HIP, "rmpcd_sort: workspace allocation failed");
    int rc = count_impl<F>(box, src, cells, wrap, key, slot);
    if (rc) return rc;
    Box b = make_box(box);
    hipStream_t s = stream();
    k_place<<<blocks_for(n, 256), 256, 0, s>>>(n, key, slot, cells->cell_start, tmp);
    RMPCD_LAUNCHED();
    k_gather<F><<<blocks_for(n, 256), 256, 0, s>>>(b, view<F>(src), view<F>(dst), wrap, key,
                                                     cells->cell_start, cells->cell_count, tmp,
                                                     id_to_idx);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

template <class F>
__global__ void k_cartesian(Box b, PView<F> p, int32_t *out)
{
    using R = typename F::real;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    int p0, p1, p2;
    cell_of<R>(b, p.pos[i], p.pos[p.n + i], p.pos[2 * p.n + i], p0, p1, p2);
    out[i] = p0;
    out[p.n + i] = p1;
    out[2 * p.n + i] = p2;
}

template <class F>
int cartesian_impl(const rmpcd_box_t *box, const rmpcd_particles_t *pp, int32_t *out)
{
    PView<F> p = view<F>(pp);
    k_cartesian<F><<<blocks_for(p.n, 256), 256, 0, stream()>>>(make_box(box), p, out);
    RMPCD_LAUNCHED();
    return RMPCD_OK;
}

}  // namespace rmpcd

using namespace rmpcd;

extern "C" {

int rmpcd_count_particles(const rmpcd_box_t *box, const rmpcd_particles_t *p,
                          rmpcd_cells_t *cells)
{
    int rc = check_box(box, "rmpcd_count_particles");
    if (!rc) rc = check_particles(p, "rmpcd_count_particles", false);
    if (rc) return rc;
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start,
                "rmpcd_count_particles: null cell arrays");
    return RMPCD_DISPATCH_FORM(p, count_impl, box, p, cells, 0, nullptr, nullptr);
}

int rmpcd_cartesian_indices(const rmpcd_box_t *box, const rmpcd_particles_t *p, int32_t *xyz)
{
    int rc = check_box(box, "rmpcd_cartesian_indices");
    if (!rc) rc = check_particles(p, "rmpcd_cartesian_indices", false);
    if (rc) return rc;
    RMPCD_CHECK(xyz != nullptr, "rmpcd_cartesian_indices: null output");
    return RMPCD_DISPATCH_FORM(p, cartesian_impl, box, p, xyz);
}

int rmpcd_sort(const rmpcd_box_t *box, const rmpcd_particles_t *src, rmpcd_particles_t *dst,
               int32_t *id_to_idx, rmpcd_cells_t *cells, int32_t apply_pbc)
{
    int rc = check_box(box, "rmpcd_sort");
    if (!rc) rc = check_particles(src, "rmpcd_sort", true);
    if (!rc) rc = check_particles(dst, "rmpcd_sort", true);
    if (rc) return rc;
    RMPCD_CHECK(src->n == dst->n && src->form == dst->form,
                "rmpcd_sort: source and destination differ in size or form");
    RMPCD_CHECK(src->image && dst->image && src->id && dst->id && id_to_idx,
                "rmpcd_sort: null image/id/id_to_idx");
    RMPCD_CHECK(cells && cells->cell_count && cells->cell_start, "rmpcd_sort: null cell arrays");
    // the gather reads every source slot while it writes the destination: the two sets
    // must be different buffers (the reference sorts into its second buffer and swaps)
    RMPCD_CHECK(src->pos != dst->pos && src->vel != dst->vel && src->force != dst->force &&
                    src->force_old != dst->force_old && src->image != dst->image &&
                    src->id != dst->id && src->species != dst->species && src->flags != dst->flags,
                "rmpcd_sort: source and destination share an array (the sort is out of place)");
    return RMPCD_DISPATCH_FORM(src, sort_impl, box, src, dst, id_to_idx, cells, apply_pbc ? 1 : 0);
}

}  // extern "C"
