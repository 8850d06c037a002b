// Shared definitions of the gfx950 integrator: error handling, the context,
// complex helpers and wavefront reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/puffin_gpu.h"

namespace puff {

void set_error(const std::string& msg);

#define PUFF_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            ::puff::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return PUFF_ERR_HIP;                                              \
        }                                                                     \
    } while (0)

template <typename R>
struct cplx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) {
    return {a.x + b.x, a.y + b.y};
}
template <typename R>
__device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) {
    return {a.x - b.x, a.y - b.y};
}

// Flag bits raised by the stage kernels (status codes are derived on the host).
constexpr int FLAG_TRANSVERSE = 1;
constexpr int FLAG_BUFFER = 2;
constexpr int FLAG_Z2MESH = 4;
// A source beyond the range the deterministic deposit's fixed point was
// scaled for at the start of the step (never expected: the scale leaves a
// factor 4 of headroom over the sources of the step's start state).
constexpr int FLAG_DET_RANGE = 8;

// Active box of one RK4 step: min/max cell index along x, y, z2.  The stage
// kernels update BOX_COPIES copies (by workgroup) to spread the atomics.
struct Box {
    int lo[3];
    int hi[3];
};
constexpr int BOX_COPIES = 64;

// Workgroup of the RK4 stage kernels (one macroparticle per thread) and the
// node capacity of its LDS deposit tile.  Four-wave workgroups (one wave per
// SIMD each); the stage kernels hold 42-76 VGPRs, so the LDS tiles set how
// many workgroups share a CU: 1280 nodes (2 FP64 each, 20 KiB) in the FP64
// form, 3072 (one packed 8-byte word each, 24 KiB) in the mixed form.
// Measured on the C1 beam: FP64 tiles of 768 and 2048 nodes gave 3.17 and
// 3.07 ms steps against 2.87 ms; mixed FP64 tiles of 1024 and 2048 nodes
// 2.39 and 2.54 ms against 2.28 ms at 1536; the packed mixed tile 2.21 ms
// at 1536 nodes and 2.18 ms at 3072.
#ifndef PUFF_STAGE_THREADS
#define PUFF_STAGE_THREADS 256
#endif
constexpr int STAGE_THREADS = PUFF_STAGE_THREADS;
// (256 or 512: the packed tile's rounding needs STAGE_THREADS >= 256)
constexpr int STAGE_THREADS_LOG2 = STAGE_THREADS == 512 ? 9 : STAGE_THREADS == 256 ? 8 : -1;
static_assert(STAGE_THREADS_LOG2 > 0, "stage workgroups of 256 or 512 threads");
#ifndef PUFF_TILE_NODES_FP64
#define PUFF_TILE_NODES_FP64 1280
#endif
#ifndef PUFF_TILE_NODES_MIXED
#define PUFF_TILE_NODES_MIXED 3072
#endif
// The mixed form's LDS deposit tile: PUFF_MIXED_TILE 0 packed fixed point
// (default), 1 FP64 (the FP64 form's tile, 1536 nodes: the best of 1024,
// 1536 and 2048 for it), 2 FP32 (ds_add_f32; kept to measure it against the
// other two: 6.0 ms C1 steps against 2.06 ms for FP64 and 1.76 ms packed).
#ifndef PUFF_MIXED_TILE
#define PUFF_MIXED_TILE 0
#endif
template <typename R>
constexpr int stage_tile_nodes() {
    return sizeof(R) == 8 ? PUFF_TILE_NODES_FP64 : (PUFF_MIXED_TILE == 1 ? 1536 : PUFF_TILE_NODES_MIXED);
}
// Node capacity of the stage kernels' LDS copy of the field over the
// workgroup's node box (one complex per node; larger boxes gather from
// global memory).
#ifndef PUFF_FIELD_TILE_NODES
#define PUFF_FIELD_TILE_NODES 256
#endif
constexpr int FIELD_TILE_NODES = PUFF_FIELD_TILE_NODES;

// Workgroup and LDS tile of the adjoint stage kernels: more live values per
// particle than the forward, so smaller workgroups with a larger VGPR budget.
constexpr int ADJ_THREADS = 256;
// Deposit tile (two doubles per node): 2000 nodes keep a workgroup's LDS
// (with the field tile) under 40 KiB, four workgroups per CU, as many as the
// kernel's 116 VGPRs allow.  1024 nodes sent more of C4's sparse-beam
// workgroups to global atomics: its adjoint stages took 969 ms per 30 steps
// against 888 ms.
constexpr int ADJ_TILE_NODES = 2000;
constexpr int ADJ_FIELD_TILE_NODES = 256;   // LDS copy of the trial field and source cotangent
constexpr int MAX_BOX_WAVES = (STAGE_THREADS > ADJ_THREADS ? STAGE_THREADS : ADJ_THREADS) / 64;

// Flag words of the step kernels: [0] flag bits of the current step;
// [FLAG_WORD_FALLBACK] the stage workgroups of the step whose node box did not
// fit the LDS tile (global-atomic deposit): the host re-sorts the particles
// when that share grows; [FLAG_WORD_STICKY] the OR of the step flags since
// the host last read them (deferred status).
constexpr int FLAG_WORDS = 3;
constexpr int FLAG_WORD_FALLBACK = 1;
constexpr int FLAG_WORD_STICKY = 2;

// Workgroups of a launch are handed to the XCDs round-robin (workgroup id
// mod NXCD), and each XCD has its own L2.  xcd_block() gives each XCD a
// contiguous range of logical block ids instead, so that neighbouring blocks
// (x-columns sharing cache lines in the z2 FFT pass) meet in one L2.  The
// launch has xcd_grid(nblocks) workgroups; logical ids >= nblocks are idle.
// (The stage kernels keep the round-robin order: measured 3 % faster.)
#ifndef PUFF_XCD_REMAP
#define PUFF_XCD_REMAP 1
#endif
constexpr int NXCD = 8;
inline int64_t xcd_grid(int64_t nblocks) {
    return PUFF_XCD_REMAP ? (nblocks + NXCD - 1) / NXCD * NXCD : nblocks;
}
__device__ __forceinline__ int64_t xcd_block(int64_t nblocks) {
    if (!PUFF_XCD_REMAP) return blockIdx.x;
    const int64_t per = (nblocks + NXCD - 1) / NXCD;
    return (int64_t)(blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
}

// Wavefront (64-lane) reductions.  The integer ones use the ROCm device
// library's wavefront reductions (DPP lane moves, no LDS round trips); all
// 64 lanes must be active.
extern "C" __device__ __attribute__((const)) int __ockl_wfred_min_i32(int);
extern "C" __device__ __attribute__((const)) int __ockl_wfred_max_i32(int);
__device__ __forceinline__ int wave_min(int v) { return __ockl_wfred_min_i32(v); }
__device__ __forceinline__ int wave_max(int v) { return __ockl_wfred_max_i32(v); }
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Per-stage scalars computed on the host: taper strength and the b-field
// coefficients of the stage's zbar (see puff_bfield in puff_push.hip).
struct StageScalars {
    double n2col;
    double c[8];
    int und_type;
};

// 1-D FFT plan of one axis.  Mixed radix: radices (product n) and the
// twiddle table exp(-2 pi i e / n), e < n.  Rader (n a prime with n - 1 a
// power of two, e.g. 17 or 257): the n-point DFT as a cyclic convolution of
// length m = n - 1; radix/npass then describe the m-point transform, tw
// holds exp(-2 pi i e / m) (n entries, the last one unused), and
//   rperm[j]  = g^j mod n, riperm[q] = g^-q mod n (g a primitive root),
//   rb        = cplx<R>[2 m]: DFT_m(b_s) / m for s = -1 then +1, with
//               b_s[j] = exp(s 2 pi i riperm[j] / n).
struct FftAxis {
    int n;
    int npass;
    int radix[16];
    const void* tw;
    int rader;
    const int* rperm;
    const int* riperm;
    const void* rb;
};

}  // namespace puff
