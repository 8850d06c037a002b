// Counting sort of the macroparticles by mesh cell, so that the lanes of a
// wavefront gather from and deposit to the same few cache lines.  The
// reference keeps its particles in generation order; the permutation is kept
// (orig[j] = reference index of slot j) and undone on download.
#include "puff_common.hpp"
#include "puff_kernels.hpp"

namespace puff {

// Bin of a particle's cell in the slot order (bricks of cells, see
// sort_bins); particles outside the mesh go to the last bin.
template <typename R>
__device__ __forceinline__ int cell_key(const MeshGeom& m, int ncell, R x, R y, double z2) {
    int k = ncell;
    const double zq = floor(z2 * m.inv_dz2);
    int iz = -1;
    if (m.periodic) {
        if (fabs(zq) < 1.0e9) {
            long long r = (long long)zq % m.nzp;
            if (r < 0) r += m.nzp;
            iz = (int)r;
        }
    } else if (zq >= 0.0 && zq <= (double)(m.nzp - 2)) {
        iz = (int)zq;
    }
    if (iz >= 0) {
        if (m.one_d) {
            k = iz;
        } else {
            const R xq = floor((x + (R)m.halfx) * (R)m.inv_dx);
            const R yq = floor((y + (R)m.halfy) * (R)m.inv_dy);
            if (xq >= (R)0 && xq <= (R)(m.nx - 2) && yq >= (R)0 && yq <= (R)(m.ny - 2)) {
                const int ix = (int)xq, iy = (int)yq;
                const int nbx = (m.nx - 1 + BRICK_X - 1) / BRICK_X;
                const int nby = (m.ny - 1 + BRICK_Y - 1) / BRICK_Y;
                // bricks: planes along z2, rows of bricks snaking, so that
                // consecutive bricks are neighbours
                const int bz = iz / BRICK_Z;
                int by = iy / BRICK_Y, bx = ix / BRICK_X;
                if (bz & 1) by = nby - 1 - by;
                const int brow = bz * nby + by;
                if (brow & 1) bx = nbx - 1 - bx;
                int lx = ix % BRICK_X, ly = iy % BRICK_Y, lz = iz % BRICK_Z;
                int cell;
                if (BRICK_XFAST) {
                    // cells of a brick: rows along x in planes, snaking
                    if (lz & 1) ly = BRICK_Y - 1 - ly;
                    const int row = lz * BRICK_Y + ly;
                    if (row & 1) lx = BRICK_X - 1 - lx;
                    cell = row * BRICK_X + lx;
                } else {
                    // cells of a brick: columns along z2, snaking
                    if (ly & 1) lx = BRICK_X - 1 - lx;
                    const int col = ly * BRICK_X + lx;
                    if (col & 1) lz = BRICK_Z - 1 - lz;
                    cell = col * BRICK_Z + lz;
                }
                k = (int)(((int64_t)brow * nbx + bx) * BRICK_CELLS + cell);
            }
        }
    }
    return k;
}

__global__ void k_fill_int(int* p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_iota(int* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int)i;
}

// Runs of equal keys within a wavefront (the input is in the previous sorted
// order, so equal cells are mostly adjacent): the run's first lane does one
// atomic for the whole run.
struct Run {
    int head_lane;
    int len;
    bool head;
};

__device__ __forceinline__ Run run_of(int k, bool valid) {
    const int lane = threadIdx.x & 63;
    const int kk = valid ? k : -1 - lane;
    const int prev = __shfl_up(kk, 1, 64);
    const bool head = (lane == 0) || (prev != kk);
    const unsigned long long heads = __ballot(head);
    const unsigned long long le = (lane == 63) ? heads : (heads & ((2ull << lane) - 1ull));
    const unsigned long long after = (lane == 63) ? 0ull : (heads >> (lane + 1));
    Run r;
    r.head = head;
    r.head_lane = 63 - __builtin_clzll(le);
    const int seg_end = after ? lane + __builtin_ctzll(after) : 63;
    r.len = seg_end - lane + 1;   // meaningful on the head lane
    return r;
}

// Key of every particle, and the histogram of the keys (counts zeroed
// before): one atomic per run of equal keys in a wavefront.
template <typename R>
__global__ __launch_bounds__(256) void k_keys_histogram(const R* x, const R* y, const double* z2,
                                                        int64_t n, MeshGeom m, int ncell, int* key,
                                                        int* counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const int k = valid ? cell_key<R>(m, ncell, x[i], y[i], z2[i]) : 0;
    if (valid) key[i] = k;
    const Run r = run_of(k, valid);
    if (valid && r.head) atomicAdd(&counts[k], r.len);
}

// Exclusive scan in three passes over tiles of SCAN_TILE elements.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ int block_exclusive_scan(int v, int* sh, int* total) {
    // inclusive scan within the wavefront
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int s = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(s, o, 64);
        if (lane >= o) s += t;
    }
    if (lane == 63) sh[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < SCAN_THREADS / 64; ++k) {
            const int t = sh[k];
            sh[k] = acc;
            acc += t;
        }
        sh[SCAN_THREADS / 64] = acc;
    }
    __syncthreads();
    const int excl = s - v + sh[wv];
    *total = sh[SCAN_THREADS / 64];
    __syncthreads();
    return excl;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const int* in, int64_t n, int* tile_sums) {
    __shared__ int sh[SCAN_THREADS / 64 + 1];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int s = 0;
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) s += in[base + k];
    int total;
    block_exclusive_scan(s, sh, &total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Single block: exclusive scan of the tile sums in place.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_sums(int* sums, int64_t n) {
    __shared__ int sh[SCAN_THREADS / 64 + 1];
    int carry = 0;
    for (int64_t off = 0; off < n; off += SCAN_TILE) {
        const int64_t base = off + (int64_t)threadIdx.x * SCAN_ITEMS;
        int v[SCAN_ITEMS];
        int s = 0;
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            v[k] = (base + k < n) ? sums[base + k] : 0;
            s += v[k];
        }
        int total;
        int e = block_exclusive_scan(s, sh, &total) + carry;
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            if (base + k < n) sums[base + k] = e;
            e += v[k];
        }
        carry += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const int* in, int64_t n, const int* tile_sums,
                                                             int* out) {
    __shared__ int sh[SCAN_THREADS / 64 + 1];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int total;
    int e = block_exclusive_scan(s, sh, &total) + tile_sums[blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) out[base + k] = e;
        e += v[k];
    }
}

__global__ __launch_bounds__(256) void k_scatter_perm(const int* key, int64_t n, int* cursor, int* perm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const int k = valid ? key[i] : 0;
    const Run r = run_of(k, valid);
    int base = 0;
    if (valid && r.head) base = atomicAdd(&cursor[k], r.len);
    base = __shfl(base, r.head_lane, 64);
    const int lane = threadIdx.x & 63;
    if (valid) perm[base + (lane - r.head_lane)] = (int)i;
}

// Every per-particle array into slot order: dst[j] = src[perm[j]].
template <typename R>
__global__ __launch_bounds__(256) void k_gather_all(PartPtrs<R> y, const R* chi, const int* orig,
                                                    const int* perm, int64_t n, PartPtrs<R> t,
                                                    R* chi_out, int* orig_out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int p = perm[j];
    t.x[j] = y.x[p];
    t.y[j] = y.y[p];
    t.z2[j] = y.z2[p];
    t.pr[j] = y.pr[p];
    t.pi[j] = y.pi[p];
    t.g[j] = y.g[p];
    chi_out[j] = chi[p];
    orig_out[j] = orig[p];
}

// dst[orig[j]] = (double) src[j]: slot order back to the reference order.
template <typename T>
__global__ __launch_bounds__(256) void k_unpermute(const T* src, const int* orig, int64_t n, double* dst) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) dst[orig[j]] = (double)src[j];
}

// dst[j] = (T) src[orig[j]]: reference order into slot order.
template <typename T>
__global__ __launch_bounds__(256) void k_permute_in(const double* src, const int* orig, int64_t n, T* dst) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) dst[j] = (T)src[orig[j]];
}

static dim3 g1(int64_t n) { return dim3((unsigned)(n > 0 ? (n + 255) / 256 : 1)); }

template <typename R>
void launch_sort(SortArgs<R>& a, hipStream_t s) {
    const int64_t n = a.n;
    const int64_t nb = (int64_t)a.ncell + 1;
    hipLaunchKernelGGL(k_fill_int, g1(nb), dim3(256), 0, s, a.counts, nb, 0);
    hipLaunchKernelGGL((k_keys_histogram<R>), g1(n), dim3(256), 0, s, a.y->x, a.y->y, a.y->z2, n, a.mesh,
                       a.ncell, a.key, a.counts);
    const int64_t ntiles = (nb + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(SCAN_THREADS), 0, s, a.counts, nb, a.tile_sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_THREADS), 0, s, a.tile_sums, ntiles);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)ntiles), dim3(SCAN_THREADS), 0, s, a.counts, nb,
                       a.tile_sums, a.cursor);
    hipLaunchKernelGGL(k_scatter_perm, g1(n), dim3(256), 0, s, a.key, n, a.cursor, a.perm);
    // permute every coordinate into the spare arrays, then swap the pointers
    PartPtrs<R>& y = *a.y;
    PartPtrs<R>& t = *a.spare;
    hipLaunchKernelGGL((k_gather_all<R>), g1(n), dim3(256), 0, s, y, *a.chi, *a.orig, a.perm, n, t,
                       *a.chi_spare, *a.orig_spare);
    std::swap(y.x, t.x);
    std::swap(y.y, t.y);
    std::swap(y.z2, t.z2);
    std::swap(y.pr, t.pr);
    std::swap(y.pi, t.pi);
    std::swap(y.g, t.g);
    std::swap(*a.chi, *a.chi_spare);
    std::swap(*a.orig, *a.orig_spare);
}

void launch_iota(int* p, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_iota, g1(n), dim3(256), 0, s, p, n);
}

template <typename T>
void launch_unpermute(const T* src, const int* orig, int64_t n, double* dst, hipStream_t s) {
    hipLaunchKernelGGL((k_unpermute<T>), g1(n), dim3(256), 0, s, src, orig, n, dst);
}

template <typename T>
void launch_permute_in(const double* src, const int* orig, int64_t n, T* dst, hipStream_t s) {
    hipLaunchKernelGGL((k_permute_in<T>), g1(n), dim3(256), 0, s, src, orig, n, dst);
}

template void launch_sort<double>(SortArgs<double>&, hipStream_t);
template void launch_sort<float>(SortArgs<float>&, hipStream_t);
template void launch_unpermute<double>(const double*, const int*, int64_t, double*, hipStream_t);
template void launch_unpermute<float>(const float*, const int*, int64_t, double*, hipStream_t);
template void launch_permute_in<double>(const double*, const int*, int64_t, double*, hipStream_t);
template void launch_permute_in<float>(const double*, const int*, int64_t, float*, hipStream_t);

}  // namespace puff
