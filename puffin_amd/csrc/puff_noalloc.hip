// Allocation-free slice statistics: getSliceTwiss's per-slice outputs
// (io/av_write.f90:405-731) of macroparticles held in the caller's device
// arrays, on the caller's stream, without a context.
//
// The per-slice sums live in `scratch`, a file-scope device array of
// NOALLOC_MAX_SLICES x NSLICE_SUMS doubles.  Each GPU holds its own copy until
// the process ends; its device address is looked up once per device with
// hipGetSymbolAddress and kept in `g_scratch`.  In host code the array's name
// is only a stand-in for the symbol, never a device address: kernels and
// fills take the looked-up pointer.  Every call on one GPU shares the array:
// calls that overlap on one GPU (different streams, or different host
// threads) race; calls on one stream are ordered.
#include <mutex>

#include "puff_common.hpp"
#include "puff_kernels.hpp"

namespace puff {

constexpr int NOALLOC_MAX_SLICES = PUFF_NOALLOC_MAX_SLICES;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[(size_t)NOALLOC_MAX_SLICES * NSLICE_SUMS];

static std::mutex g_scratch_mutex;
static double* g_scratch[NOALLOC_MAX_DEVICES];

// The device address of `scratch` on the current device (looked up on the
// first call per device).
static int scratch_on_current_device(double** out) {
    int dev = -1;
    PUFF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        set_error("device index " + std::to_string(dev) + " beyond the supported " +
                  std::to_string(NOALLOC_MAX_DEVICES));
        return PUFF_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        PUFF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("hipGetSymbolAddress returned a null address for the slice scratch");
            return PUFF_ERR_HIP;
        }
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return PUFF_OK;
}

// One thread per slice: the getSliceTwiss outputs from the sums, in the
// layout of puff_slice_stats (out[k * nslices + s]); empty slices are zero.
__global__ __launch_bounds__(256) void k_slice_finish(const double* acc, int nslices, double* out) {
    const int is = blockIdx.x * blockDim.x + threadIdx.x;
    if (is >= nslices) return;
    auto A = [&](int k) { return acc[(size_t)k * nslices + is]; };
    auto root = [](double v) { return sqrt(v < 0.0 ? 0.0 : v); };
    const double sd = A(0);
    double o[13] = {sd, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (sd > 0.0) {
        const double ax = A(1) / sd, ay = A(2) / sd, apx = A(3) / sd, apy = A(4) / sd;
        const double ag = A(5) / sd;
        const double vx = A(6) / sd - ax * ax, vy = A(11) / sd - ay * ay;
        const double vpx = A(8) / sd - apx * apx, vpy = A(9) / sd - apy * apy;
        const double vg = A(12) / sd - ag * ag;
        const double cx = A(7) / sd - ax * apx, cy = A(10) / sd - ay * apy;
        o[1] = ax;
        o[2] = ay;
        o[3] = apx;
        o[4] = apy;
        o[5] = root(vx);
        o[6] = root(vy);
        o[7] = root(vpx);
        o[8] = root(vpy);
        o[9] = root(vx * vpx - cx * cx);
        o[10] = root(vy * vpy - cy * cy);
        o[11] = ag;
        o[12] = root(vg);
    }
#pragma unroll
    for (int k = 0; k < 13; ++k) out[(size_t)k * nslices + is] = o[k];
}

}  // namespace puff

using namespace puff;

extern "C" int puff_slice_stats_noalloc(const puff_params* params, int64_t n, const double* x,
                                        const double* y, const double* z2, const double* px,
                                        const double* py, const double* gam, const double* chi,
                                        int32_t nslices, double slice_len, double* out_dev,
                                        void* hip_stream) {
    if (!params || n < 0 || !out_dev || !(slice_len > 0.0)) {
        set_error("puff_slice_stats_noalloc: null params or output, n < 0, or slice_len <= 0");
        return PUFF_ERR_ARG;
    }
    if (nslices < 1 || nslices > NOALLOC_MAX_SLICES) {
        set_error("puff_slice_stats_noalloc: nslices must be in [1, " + std::to_string(NOALLOC_MAX_SLICES) +
                  "] (the size of its static scratch), got " + std::to_string(nslices));
        return PUFF_ERR_ARG;
    }
    if (n > 0 && (!x || !y || !z2 || !px || !py || !gam || !chi)) {
        set_error("puff_slice_stats_noalloc: null particle array");
        return PUFF_ERR_ARG;
    }
    double* acc = nullptr;
    const int st = scratch_on_current_device(&acc);
    if (st != PUFF_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    // the sums start at zero (a fill kernel, so the call may be captured)
    launch_fill_zero(acc, (int64_t)NSLICE_SUMS * nslices, s);
    if (n > 0) {
        PartPtrs<double> p{const_cast<double*>(x), const_cast<double*>(y), const_cast<double*>(z2),
                           const_cast<double*>(px), const_cast<double*>(py), const_cast<double*>(gam)};
        launch_slice_moments<double>(p, chi, n, nslices, slice_len, params->periodic, acc, s);
    }
    hipLaunchKernelGGL(k_slice_finish, dim3((unsigned)((nslices + 255) / 256)), dim3(256), 0, s, acc,
                       (int)nslices, out_dev);
    PUFF_HIP(hipGetLastError());
    return PUFF_OK;
}
