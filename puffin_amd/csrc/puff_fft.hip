// Batched mixed-radix complex FFT along one axis of the field mesh, and the
// diffraction step built from it.
//
// Convention of the reference's FFTW plans (transforms.f90:186-211, 423):
// sign -1 forward, +1 backward, unnormalised; the backward result is divided
// by ffact = NX NY NZ2' afterwards (diffraction.f90:310, 513).
//
// Each workgroup holds B whole lines in LDS and runs a Stockham autosort
// transform (natural order in and out), one radix pass per LDS round trip.
// Radices 2, 3, 4 and 16 have explicit butterflies (16 as 4 x 4 in
// registers, halving the passes and barriers of power-of-two lines); any
// other factor (11, 43, 127, ...) is a direct DFT pass; a prime line length
// with a power-of-two n - 1 (17, 257) uses Rader's algorithm.
//
// DiffractionStep (diffraction.f90:209-348) becomes nine line passes:
//   x fwd | y fwd | z fwd . multiplyexp . z inv | y inv |
//   x inv . /ffact . mask split . x fwd | y fwd | z fwd . damping . z inv |
//   y inv | x inv . /ffact . + interior
// (4096-point z2 lines: each z2 pass as three four-step passes, k_fft_z4).
#include "puff_common.hpp"
#include "puff_kernels.hpp"

namespace puff {

// Threads per z2 workgroup.  512 and 1024 (more loads in flight per line
// at the same LDS) measured 6 % and 30 % slower on the C1 mesh in FP64;
// 512 gained 2 % on the C4 mesh.
#ifndef PUFF_Z_THREADS
#define PUFF_Z_THREADS 256
#endif
// Threads per x/y workgroup on a Rader axis.  Its radix-16 sub-passes keep
// B x 16 of them busy; with the ping-pong transform (7 lines of C4's 257
// points, 112 busy) 128 and 192 threads measured 15 % and 7 % slower C4
// events than 256.
#ifndef PUFF_RADER_THREADS
#define PUFF_RADER_THREADS 256
#endif

// LDS word of flat element e of a line-major buffer (IL false): one spare
// slot after every 16 elements.  A radix-16 pass with p = 1 has each lane
// write 16 consecutive outputs, so lanes write 16 elements apart, all on the
// same LDS banks; with the spare slots they land 17 apart and 16 lanes cover
// all the banks.  Interleaved buffers (IL: y and z2 lines, b fastest) put
// consecutive lanes on consecutive elements and are not padded.
#ifndef PUFF_LDS_PAD_X
#define PUFF_LDS_PAD_X 1
#endif
//
// SW (interleaved buffers of the z2 passes, 0: none) swizzles instead of
// padding: flat index e becomes e ^ ((e >> (SW & 255)) & (SW >> 8)).  With
// B lines interleaved, the first radix-16 pass writes lane i's outputs 16 B
// elements apart (512 bytes for two FP64 or four FP32 lines: one LDS bank);
// XOR-ing the block number 16 B (shift log2(16 B)) into the low bits, within
// the block, puts the lanes on different banks.  A padded buffer would need
// 3 % more LDS, one workgroup per CU fewer at 32 KiB.
template <bool IL, int SW = 0>
__host__ __device__ constexpr int lpad(int e) {
    return (!IL && PUFF_LDS_PAD_X) ? e + (e >> 4) : ((IL && SW) ? e ^ ((e >> (SW & 255)) & (SW >> 8)) : e);
}

// LDS index of element j of line b.  Interleaved layout (b fastest) keeps
// the strided y / z2 loads conflict free; line-major suits contiguous x lines.
template <bool IL, int SW = 0>
__device__ __forceinline__ int lidx(int b, int j, int B, int N) {
    return lpad<IL, SW>(IL ? j * B + b : b * N + j);
}

// Line staging between global memory and LDS with U loads (or stores) per
// thread issued back to back.  A plain element loop (U = 1) writes each
// loaded value to LDS before it issues the next load, so every thread has
// one load in flight and pays the memory latency once per element.  U = 8
// took C4's x and y passes from 4.69 and 3.37 ms to 4.25 and 2.66 ms, and
// (with the XCD-grouped z2 workgroups) C1's FP64 event from 2.47 to 2.40 ms
// through its z2 passes.
constexpr int STAGE_UNROLL = 8;
#ifndef PUFF_Z_UNROLL
#define PUFF_Z_UNROLL 8
#endif

// dst[lpad<IL>(w)] = get(w) for 0 <= w < n.
template <bool IL, int U, int SW = 0, typename R, typename F>
__device__ __forceinline__ void lds_fill(cplx<R>* dst, int n, F get) {
    for (int w0 = threadIdx.x; w0 < n; w0 += U * (int)blockDim.x) {
        cplx<R> v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int w = w0 + u * (int)blockDim.x;
            v[u] = (w < n) ? get(w) : cplx<R>{0, 0};
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int w = w0 + u * (int)blockDim.x;
            if (w < n) dst[lpad<IL, SW>(w)] = v[u];
        }
    }
}

// put(w, src[lpad<IL>(w)]) for 0 <= w < n.  The LDS reads are unconditional
// (element 0 stands in past n), so that v stays in registers: with
// conditional reads the compiler kept v in scratch memory (144 bytes per
// lane in FP64).
template <bool IL, int U, int SW = 0, typename R, typename F>
__device__ __forceinline__ void lds_drain(const cplx<R>* src, int n, F put) {
    for (int w0 = threadIdx.x; w0 < n; w0 += U * (int)blockDim.x) {
        cplx<R> v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int w = w0 + u * (int)blockDim.x;
            v[u] = src[lpad<IL, SW>(w < n ? w : 0)];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int w = w0 + u * (int)blockDim.x;
            if (w < n) put(w, v[u]);
        }
    }
}

// threadIdx.x through an empty asm statement: index arithmetic built on it is
// not merged with the same arithmetic of another call site, so it is
// recomputed where it is used instead of held in registers between the
// passes of a kernel (with two in-place Rader transforms per line, 181
// VGPRs without it).
__device__ __forceinline__ int opaque_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

template <typename R>
__device__ __forceinline__ cplx<R> tw_get(const cplx<R>* tw, int e, int sign) {
    cplx<R> w = tw[e];
    if (sign > 0) w.y = -w.y;
    return w;
}

// 4-point DFT in place (x_k = sum_n x_n W4^(nk), W4 = exp(sign 2 pi i / 4)).
template <typename R>
__device__ __forceinline__ void dft4(cplx<R>& x0, cplx<R>& x1, cplx<R>& x2, cplx<R>& x3, int sign) {
    const cplx<R> a0 = cadd(x0, x2), a1 = csub(x0, x2);
    const cplx<R> b0 = cadd(x1, x3), b1 = csub(x1, x3);
    // W4 b1: sign -1 -> -i b1, sign +1 -> +i b1
    const cplx<R> wb = (sign < 0) ? cplx<R>{b1.y, -b1.x} : cplx<R>{-b1.y, b1.x};
    x0 = cadd(a0, b0);
    x1 = cadd(a1, wb);
    x2 = csub(a0, b0);
    x3 = csub(a1, wb);
}

// 16-point DFT in registers as 4 x 4 (Cooley-Tukey): with j = 4 j1 + j0 and
// q = q0 + 4 q1, X_q = sum_j0 W4^(j0 q1) [w^(j0 q0) sum_j1 v_(4 j1 + j0) W4^(j1 q0)],
// w = exp(sign 2 pi i / 16).  On return v[4 q0 + q1] = X_(q0 + 4 q1).
template <typename R>
__device__ __forceinline__ void dft16(cplx<R>* v, int sign) {
    const R c1 = (R)0.92387953251128675613, s1 = (R)0.38268343236508977173;   // cos, sin of pi/8
    const R c2 = (R)0.70710678118654752440;                                    // cos pi/4
    const R sg = (R)(sign < 0 ? -1 : 1);
#pragma unroll
    for (int j0 = 0; j0 < 4; ++j0) dft4(v[j0], v[4 + j0], v[8 + j0], v[12 + j0], sign);
    // v[4 q0 + j0] = Y[j0][q0]; twiddle by w^(j0 q0): m = 1, 2, 3 | 2, 4, 6 | 3, 6, 9
    const cplx<R> w1 = {c1, sg * s1}, w2 = {c2, sg * c2}, w3 = {s1, sg * c1};
    const cplx<R> w6 = {-c2, sg * c2}, w9 = {-c1, -sg * s1};
    v[5] = cmul(v[5], w1);
    v[9] = cmul(v[9], w2);
    v[13] = cmul(v[13], w3);
    v[6] = cmul(v[6], w2);
    v[10] = (sign < 0) ? cplx<R>{v[10].y, -v[10].x} : cplx<R>{-v[10].y, v[10].x};   // w^4 = W4
    v[14] = cmul(v[14], w6);
    v[7] = cmul(v[7], w3);
    v[11] = cmul(v[11], w6);
    v[15] = cmul(v[15], w9);
#pragma unroll
    for (int q0 = 0; q0 < 4; ++q0) dft4(v[4 * q0], v[4 * q0 + 1], v[4 * q0 + 2], v[4 * q0 + 3], sign);
}

// ---- compile-time line plans --------------------------------------------------
// The line lengths of the workloads get kernels with their length, radix
// passes (factorize()'s order) and lines per workgroup as constants: every
// index division and modulo by B, T, p and N then becomes a multiply and
// shift, and the pass loops unroll.  Other lengths run the same code with
// the runtime plan (NC = 0).
template <int N>
struct CtPlan {
    static constexpr int np = 0;
    static constexpr int rader = 0;
    static constexpr int r[1] = {0};
};
#define PUFF_CT_PLAN(N_, RADER_, ...)                         \
    template <>                                               \
    struct CtPlan<N_> {                                       \
        static constexpr int r[] = {__VA_ARGS__};             \
        static constexpr int np = sizeof(r) / sizeof(int);    \
        static constexpr int rader = RADER_;                  \
    };
PUFF_CT_PLAN(33, 0, 11, 3)
PUFF_CT_PLAN(129, 0, 43, 3)
PUFF_CT_PLAN(257, 1, 16, 16)   // Rader: transforms of M = 256
PUFF_CT_PLAN(128, 0, 16, 4, 2)
PUFF_CT_PLAN(256, 0, 16, 16)
PUFF_CT_PLAN(1024, 0, 16, 16, 4)
PUFF_CT_PLAN(4096, 0, 16, 16, 16)
PUFF_CT_PLAN(64, 0, 16, 4)   // the four-step sub-lines of 4096-point z2 lines
#undef PUFF_CT_PLAN

// Direct DFT pass for an odd prime r (a radix pass of lds_pass), folded:
// with the pre-twiddled inputs y_j = x_j W^(j k twst), u_j = y_j + y_(r-j),
// v_j = y_j - y_(r-j) (j = 1 .. (r-1)/2) and w^(jq) = c + i s
// (w^((r-j)q) = c - i s),
//   X_q     = y_0 + sum u_j c + i sum v_j s
//   X_(r-q) = y_0 + sum u_j c - i sum v_j s,
// four FMAs per input pair for both outputs.  Each work item forms the
// outputs of NQ consecutive values of q (with their mirrors r - q), sharing
// the input loads, the folding and the loop bookkeeping: per input pair it
// reads 2 inputs and NQ rotations from LDS for 4 NQ FMAs.  The pass is bound
// by LDS reads; q >= (r + 1) / 2 slots of the last item read a valid table
// entry (q = 0) and store nothing.
#ifndef PUFF_DFT_WIDE
#define PUFF_DFT_WIDE 4
#endif
#ifndef PUFF_DFT_WIDE_MIN
#define PUFF_DFT_WIDE_MIN 23
#endif
template <typename R, bool IL, int NQ>
__device__ __forceinline__ void direct_pass(const cplx<R>* src, cplx<R>* dst, int B, int LD, int N, int r,
                                            int p, const cplx<R>* tw, int sign) {
    const int T = N / r;
    const int twst = N / (p * r);
    const int nr = N / r;
    const int hq = (r + 1) / 2;
    const int nqq = (hq + NQ - 1) / NQ;
    for (int w = threadIdx.x; w < B * T * nqq; w += blockDim.x) {
        int b, t;
        if (IL) { b = w % B; t = w / B; } else { t = w % (T * nqq); b = w / (T * nqq); }
        const int q0 = NQ * (t / T);
        const int i = t % T;
        const int k = i % p;
        const int dk = k * twst;
        int dq[NQ], f[NQ];
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            dq[m] = (q0 + m < hq) ? (q0 + m) * nr : 0;   // 0: a valid table index, result unused
            f[m] = dq[m];
        }
        const cplx<R> y0 = src[lidx<IL>(b, i, B, LD)];
        int e1 = dk, e2 = (r - 1) * dk;   // e1, e2 < N throughout
        R sr[NQ], si[NQ], tr[NQ], ti[NQ];
#pragma unroll
        for (int m = 0; m < NQ; ++m) sr[m] = si[m] = tr[m] = ti[m] = (R)0;
        constexpr int UJ = NQ >= 4 ? 2 : 4;
#pragma unroll UJ
        for (int j = 1; j < hq; ++j) {
            cplx<R> xa = src[lidx<IL>(b, i + j * T, B, LD)];
            cplx<R> xb = src[lidx<IL>(b, i + (r - j) * T, B, LD)];
            if (dk) {
                xa = cmul(xa, tw_get(tw, e1, sign));
                xb = cmul(xb, tw_get(tw, e2, sign));
            }
            const R ux = xa.x + xb.x, uy = xa.y + xb.y;
            const R vx = xa.x - xb.x, vy = xa.y - xb.y;
#pragma unroll
            for (int m = 0; m < NQ; ++m) {
                const cplx<R> o = tw_get(tw, f[m], sign);
                sr[m] = fma(ux, o.x, sr[m]);
                si[m] = fma(uy, o.x, si[m]);
                tr[m] = fma(vx, o.y, tr[m]);
                ti[m] = fma(vy, o.y, ti[m]);
                f[m] += dq[m];
                if (f[m] >= N) f[m] -= N;
            }
            e1 += dk;
            e2 -= dk;
        }
        const int ob = (i / p) * p * r + k;
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            const int q = q0 + m;
            if (q < hq) {
                dst[lidx<IL>(b, ob + q * p, B, LD)] = {y0.x + sr[m] - ti[m], y0.y + si[m] + tr[m]};
                if (q > 0) dst[lidx<IL>(b, ob + (r - q) * p, B, LD)] = {y0.x + sr[m] + ti[m], y0.y + si[m] - tr[m]};
            }
        }
    }
}

// LDS read of a complex value at a byte offset from base, as one 2-vector
// whose alignment is its size: one ds_read_b128 (FP64) or ds_read_b64
// (FP32) whose 16-bit offset field takes the compile-time part of the
// address.  A cplx<double> read (8-byte aligned) becomes ds_read2_b64,
// whose offsets reach only 2040 bytes.
template <typename R>
struct Vec2;
template <>
struct Vec2<double> {
    typedef double __attribute__((ext_vector_type(2))) T;
};
template <>
struct Vec2<float> {
    typedef float __attribute__((ext_vector_type(2))) T;
};
template <typename R>
__device__ __forceinline__ cplx<R> lds_at(const cplx<R>* base, uint32_t byte_off) {
    const typename Vec2<R>::T v =
        *reinterpret_cast<const typename Vec2<R>::T*>(reinterpret_cast<const char*>(base) + byte_off);
    return {v.x, v.y};
}

// direct_pass for a compile-time radix RR of a compile-time line length N,
// after passes whose radices multiply to P (same arithmetic, bit for bit):
// - the j loop has a compile-time trip count and unrolls in groups;
// - the rotations w^(jq) are read from the forward table and the sign
//   applied to the s sums at the end (X_q takes +-i sum v s; negation is
//   exact and commutes with the FMA rounding), instead of a conditional
//   negate per rotation;
// - each rotation's table offset is kept in bytes and advanced by q nr with
//   one unsigned min for the wrap at N (a + d and a + d - N: the one below N
//   is the smaller as unsigned), instead of add, compare, select and an
//   address computation.
template <typename R, bool IL, int NQ, int N, int RR, int P>
__device__ __forceinline__ void direct_pass_ct(const cplx<R>* src, cplx<R>* dst, int B, const cplx<R>* tw,
                                               int sign) {
    constexpr int T = N / RR;
    constexpr int twst = N / (P * RR);
    constexpr int nr = N / RR;
    constexpr int hq = (RR + 1) / 2;
    constexpr int nqq = (hq + NQ - 1) / NQ;
    constexpr uint32_t ES = sizeof(cplx<R>);
    constexpr uint32_t WRAP = (uint32_t)N * ES;
    for (int w = threadIdx.x; w < B * T * nqq; w += blockDim.x) {
        int b, t;
        if (IL) { b = w % B; t = w / B; } else { t = w % (T * nqq); b = w / (T * nqq); }
        const int q0 = NQ * (t / T);
        const int i = t % T;
        const int k = i % P;
        const int dk = k * twst;
        uint32_t ofs[NQ], stp[NQ], stw[NQ];
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            // q >= hq: offset 0 throughout (a valid table entry, result unused)
            stp[m] = (q0 + m < hq) ? (uint32_t)((q0 + m) * nr) * ES : 0u;
            stw[m] = stp[m] - WRAP;
            ofs[m] = stp[m];
        }
        const uint32_t la = (uint32_t)lidx<IL>(b, i, B, N) * ES;
        const cplx<R> y0 = lds_at(src, la);
        R sr[NQ], si[NQ], tr[NQ], ti[NQ];
#pragma unroll
        for (int m = 0; m < NQ; ++m) sr[m] = si[m] = tr[m] = ti[m] = (R)0;
        // unrolled in groups (a full unroll hoists every LDS read and takes
        // ~490 VGPRs); the groups' reads take their offsets from the
        // instruction
        constexpr int UJ = (hq - 1) <= 6 ? (hq - 1) : ((hq - 1) % 3 == 0 ? 3 : 2);
#pragma unroll UJ
        for (int j = 1; j < hq; ++j) {
            cplx<R> o[NQ];
#pragma unroll
            for (int m = 0; m < NQ; ++m) {
                o[m] = lds_at(tw, ofs[m]);
                ofs[m] = min(ofs[m] + stp[m], ofs[m] + stw[m]);
            }
            // interleaved lines: element i + j T of line b sits j T B
            // elements after element i
            cplx<R> xa = IL ? lds_at(src, la + (uint32_t)(j * T * B) * ES)
                            : lds_at(src, (uint32_t)lidx<IL>(b, i + j * T, B, N) * ES);
            cplx<R> xb = IL ? lds_at(src, la + (uint32_t)((RR - j) * T * B) * ES)
                            : lds_at(src, (uint32_t)lidx<IL>(b, i + (RR - j) * T, B, N) * ES);
            if constexpr (P > 1) {
                if (dk) {   // exponents j dk, (RR - j) dk < N
                    xa = cmul(xa, tw_get(tw, j * dk, sign));
                    xb = cmul(xb, tw_get(tw, (RR - j) * dk, sign));
                }
            }
            const R ux = xa.x + xb.x, uy = xa.y + xb.y;
            const R vx = xa.x - xb.x, vy = xa.y - xb.y;
#pragma unroll
            for (int m = 0; m < NQ; ++m) {
                sr[m] = fma(ux, o[m].x, sr[m]);
                si[m] = fma(uy, o[m].x, si[m]);
                tr[m] = fma(vx, o[m].y, tr[m]);
                ti[m] = fma(vy, o[m].y, ti[m]);
            }
        }
        const int ob = (i / P) * P * RR + k;
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            if (sign > 0) {
                tr[m] = -tr[m];
                ti[m] = -ti[m];
            }
            const int q = q0 + m;
            if (q < hq) {
                dst[lidx<IL>(b, ob + q * P, B, N)] = {y0.x + sr[m] - ti[m], y0.y + si[m] + tr[m]};
                if (q > 0) dst[lidx<IL>(b, ob + (RR - q) * P, B, N)] = {y0.x + sr[m] + ti[m], y0.y + si[m] - tr[m]};
            }
        }
    }
}

// One radix-r pass (after passes whose radices multiply to p) of an in-LDS
// mixed-radix transform of length N over elements 0..N-1 of B lines laid
// out with line length LD (N <= LD; elements N..LD-1 are left alone): src
// to dst.  With constant arguments (ct_passes) every index division folds.
template <typename R, bool IL>
__device__ __forceinline__ void lds_pass(const cplx<R>* src, cplx<R>* dst, int B, int LD, int N, int r,
                                         int p, const cplx<R>* tw, int sign) {
    const int T = N / r;
    // twiddle exponents j k twst (j < r, k < p) stay below N: no reduction
    const int twst = N / (p * r);
    if (r == 16) {
        for (int w = threadIdx.x; w < B * T; w += blockDim.x) {
            int b, i;
            if (IL) { b = w % B; i = w / B; } else { i = w % T; b = w / T; }
            const int k = i % p;
            const int ob = (i / p) * p * 16 + k;
            cplx<R> v[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                v[j] = src[lidx<IL>(b, i + j * T, B, LD)];
                if (j > 0 && k > 0) v[j] = cmul(v[j], tw_get(tw, j * k * twst, sign));
            }
            dft16(v, sign);
#pragma unroll
            for (int q0 = 0; q0 < 4; ++q0)
#pragma unroll
                for (int q1 = 0; q1 < 4; ++q1)
                    dst[lidx<IL>(b, ob + (q0 + 4 * q1) * p, B, LD)] = v[4 * q0 + q1];
        }
    } else if (r == 2 || r == 3 || r == 4) {
        for (int w = threadIdx.x; w < B * T; w += blockDim.x) {
            int b, i;
            if (IL) { b = w % B; i = w / B; } else { i = w % T; b = w / T; }
            const int k = i % p;
            const int ob = (i / p) * p * r + k;
            cplx<R> v[4];
            for (int j = 0; j < r; ++j) {
                v[j] = src[lidx<IL>(b, i + j * T, B, LD)];
                if (j > 0 && k > 0) v[j] = cmul(v[j], tw_get(tw, j * k * twst, sign));
            }
            if (r == 2) {
                dst[lidx<IL>(b, ob, B, LD)] = cadd(v[0], v[1]);
                dst[lidx<IL>(b, ob + p, B, LD)] = csub(v[0], v[1]);
            } else if (r == 3) {
                const R s3 = (R)0.86602540378443864676 * (R)(sign < 0 ? -1 : 1);
                const cplx<R> t1 = cadd(v[1], v[2]);
                const cplx<R> t2 = {v[0].x - (R)0.5 * t1.x, v[0].y - (R)0.5 * t1.y};
                const cplx<R> d = csub(v[1], v[2]);
                const cplx<R> t3 = {-s3 * d.y, s3 * d.x};   // i s s3 (x1 - x2)
                dst[lidx<IL>(b, ob, B, LD)] = cadd(v[0], t1);
                dst[lidx<IL>(b, ob + p, B, LD)] = cadd(t2, t3);
                dst[lidx<IL>(b, ob + 2 * p, B, LD)] = csub(t2, t3);
            } else {
                const cplx<R> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
                const cplx<R> b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
                // W4 b1: sign -1 -> -i b1, sign +1 -> +i b1
                const cplx<R> wb = (sign < 0) ? cplx<R>{b1.y, -b1.x} : cplx<R>{-b1.y, b1.x};
                dst[lidx<IL>(b, ob, B, LD)] = cadd(a0, b0);
                dst[lidx<IL>(b, ob + p, B, LD)] = cadd(a1, wb);
                dst[lidx<IL>(b, ob + 2 * p, B, LD)] = csub(a0, b0);
                dst[lidx<IL>(b, ob + 3 * p, B, LD)] = csub(a1, wb);
            }
        }
    } else if (r >= PUFF_DFT_WIDE_MIN) {
        direct_pass<R, IL, PUFF_DFT_WIDE>(src, dst, B, LD, N, r, p, tw, sign);
    } else {
        direct_pass<R, IL, 2>(src, dst, B, LD, N, r, p, tw, sign);
    }
}

// The passes of a runtime plan.  src holds the input; returns the buffer
// holding the output (src or dst).
template <typename R, bool IL>
__device__ __forceinline__ cplx<R>* lds_passes(cplx<R>* src, cplx<R>* dst, int B, int LD, int N, int npass,
                                               const int* radix, const cplx<R>* tw, int sign) {
    int p = 1;
    for (int ps = 0; ps < npass; ++ps) {
        const int r = radix[ps];
        lds_pass<R, IL>(src, dst, B, LD, N, r, p, tw, sign);
        __syncthreads();
        cplx<R>* t = src;
        src = dst;
        dst = t;
        p *= r;
    }
    return src;
}

// The passes K.. of CtPlan<NC> (P: product of the earlier radices), with
// the radix and its offsets as constants.
template <typename R, bool IL, int NC, int K = 0, int P = 1>
__device__ __forceinline__ cplx<R>* ct_passes(cplx<R>* src, cplx<R>* dst, int B, int LD, int N,
                                              const cplx<R>* tw, int sign) {
    if constexpr (K == CtPlan<NC>::np) {
        return src;
    } else {
        constexpr int r = CtPlan<NC>::r[K];
        constexpr int NN = CtPlan<NC>::rader ? NC - 1 : NC;   // the transform's length
        if constexpr (r != 2 && r != 3 && r != 4 && r != 16 && !CtPlan<NC>::rader) {
            constexpr int nq = r >= PUFF_DFT_WIDE_MIN ? PUFF_DFT_WIDE : 2;
            direct_pass_ct<R, IL, nq, NN, r, P>(src, dst, B, tw, sign);
        } else {
            lds_pass<R, IL>(src, dst, B, LD, N, r, P, tw, sign);
        }
        __syncthreads();
        return ct_passes<R, IL, NC, K + 1, P * r>(dst, src, B, LD, N, tw, sign);
    }
}

// Plan NC's passes when NC > 0, else the runtime plan of ax.
template <typename R, bool IL, int NC>
__device__ __forceinline__ cplx<R>* plan_passes(cplx<R>* src, cplx<R>* dst, int B, int LD, int N,
                                                const FftAxis& ax, const cplx<R>* tw, int sign) {
    if constexpr (NC > 0) return ct_passes<R, IL, NC>(src, dst, B, LD, N, tw, sign);
    else return lds_passes<R, IL>(src, dst, B, LD, N, ax.npass, ax.radix, tw, sign);
}

// Rader's algorithm for a prime N with N - 1 = M a power of two (FftAxis):
// a_j = x[g^j], A = DFT_M(a), C = A . rb_s, c = IDFT_M(C) (unnormalised; rb
// carries the 1/M), X[g^-q] = x_0 + c_q, X_0 = x_0 + A_0.  Element M of each
// line buffer is outside the M-point transforms: it keeps x_0 in dst and
// X_0 in src until the final permutation.
template <typename R, bool IL, int NC>
__device__ __forceinline__ cplx<R>* lds_rader(cplx<R>* src, cplx<R>* dst, int B, const FftAxis& ax,
                                              const cplx<R>* tw, int sign) {
    const int N = NC ? NC : ax.n, M = N - 1;
    const cplx<R>* rb = reinterpret_cast<const cplx<R>*>(ax.rb) + (sign > 0 ? M : 0);
    cplx<R>* P = dst;   // element M: x_0
    cplx<R>* Q = src;   // element M: X_0
    for (int w = threadIdx.x; w < B * N; w += blockDim.x) {
        int b, j;
        if (IL) { b = w % B; j = w / B; } else { j = w % N; b = w / N; }
        P[lidx<IL>(b, j, B, N)] = src[lidx<IL>(b, j < M ? ax.rperm[j] : 0, B, N)];
    }
    __syncthreads();
    cplx<R>* res = plan_passes<R, IL, NC>(P, Q, B, N, M, ax, tw, -1);
    for (int w = threadIdx.x; w < B * M; w += blockDim.x) {
        int b, k;
        if (IL) { b = w % B; k = w / B; } else { k = w % M; b = w / M; }
        const cplx<R> v = res[lidx<IL>(b, k, B, N)];
        if (k == 0) {
            const cplx<R> x0 = P[lidx<IL>(b, M, B, N)];
            Q[lidx<IL>(b, M, B, N)] = cadd(x0, v);
        }
        res[lidx<IL>(b, k, B, N)] = cmul(v, rb[k]);
    }
    __syncthreads();
    cplx<R>* other = (res == P) ? Q : P;
    res = plan_passes<R, IL, NC>(res, other, B, N, M, ax, tw, +1);
    cplx<R>* out = (res == P) ? Q : P;
    // every output element except element M (which holds x_0 or X_0 in out)
    const int qm = M / 2;   // g^-(M/2) = -1 = M (mod N)
    for (int w = threadIdx.x; w < B * M; w += blockDim.x) {
        int b, q;
        if (IL) { b = w % B; q = w / B; } else { q = w % M; b = w / M; }
        const cplx<R> x0 = P[lidx<IL>(b, M, B, N)];
        if (q == qm) {
            out[lidx<IL>(b, 0, B, N)] = Q[lidx<IL>(b, M, B, N)];
        } else {
            out[lidx<IL>(b, ax.riperm[q], B, N)] = cadd(x0, res[lidx<IL>(b, q, B, N)]);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < B; b += blockDim.x) {
        const cplx<R> x0 = P[lidx<IL>(b, M, B, N)];
        out[lidx<IL>(b, M, B, N)] = cadd(x0, res[lidx<IL>(b, qm, B, N)]);
    }
    __syncthreads();
    return out;
}

// In-LDS transform of B lines of length ax.n.  src holds the input; returns
// the buffer holding the output (src or dst).
// Radix-2/3/4 butterfly in place on v[0..r-1] (same arithmetic as the
// ping-pong passes).
template <typename R>
__device__ __forceinline__ void small_radix(cplx<R>* v, int r, int sign) {
    if (r == 2) {
        const cplx<R> t = v[0];
        v[0] = cadd(t, v[1]);
        v[1] = csub(t, v[1]);
    } else if (r == 3) {
        const R s3 = (R)0.86602540378443864676 * (R)(sign < 0 ? -1 : 1);
        const cplx<R> t1 = cadd(v[1], v[2]);
        const cplx<R> t2 = {v[0].x - (R)0.5 * t1.x, v[0].y - (R)0.5 * t1.y};
        const cplx<R> d = csub(v[1], v[2]);
        const cplx<R> t3 = {-s3 * d.y, s3 * d.x};
        v[0] = cadd(v[0], t1);
        v[1] = cadd(t2, t3);
        v[2] = csub(t2, t3);
    } else {
        dft4(v[0], v[1], v[2], v[3], sign);
    }
}

// In-place form of lds_passes for lines whose passes are all radix 16, 4, 3
// or 2, with at most one butterfly per thread in every pass (B N / r <=
// blockDim, checked on the host): each thread reads its butterfly's inputs
// into registers, the workgroup passes a barrier, and the outputs overwrite
// the same buffer.  Half the LDS of the ping-pong form, hence more
// workgroups per CU.
template <typename R, bool IL, int KB = 1, int SW = 0>
__device__ __forceinline__ void inplace_pass(cplx<R>* buf, int B, int N, int r, int p, const cplx<R>* tw,
                                             int sign, int LD) {
    // KB butterflies per thread: all their inputs are read into registers
    // before the barrier, all outputs written after it
    const int T = N / r;
    const int twst = N / (p * r);   // j k twst < N, as in lds_pass
    int b[KB], ob[KB];
    bool act[KB];
    cplx<R> v[KB][16];
    const int tid = opaque_tid();
#pragma unroll
    for (int it = 0; it < KB; ++it) {
        const int w = tid + it * (int)blockDim.x;
        act[it] = w < B * T;
        b[it] = 0;
        ob[it] = 0;
        if (act[it]) {
            int i;
            if (IL) { b[it] = w % B; i = w / B; } else { i = w % T; b[it] = w / T; }
            const int k = i % p;
            ob[it] = (i / p) * p * r + k;
            if (r == 16) {
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    v[it][j] = buf[lidx<IL, SW>(b[it], i + j * T, B, LD)];
                    if (j > 0 && k > 0) v[it][j] = cmul(v[it][j], tw_get(tw, j * k * twst, sign));
                }
                dft16(v[it], sign);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (j < r) {
                        v[it][j] = buf[lidx<IL, SW>(b[it], i + j * T, B, LD)];
                        if (j > 0 && k > 0) v[it][j] = cmul(v[it][j], tw_get(tw, j * k * twst, sign));
                    }
                }
                small_radix(v[it], r, sign);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < KB; ++it) {
        if (act[it]) {
            if (r == 16) {
#pragma unroll
                for (int q0 = 0; q0 < 4; ++q0)
#pragma unroll
                    for (int q1 = 0; q1 < 4; ++q1)
                        buf[lidx<IL, SW>(b[it], ob[it] + (q0 + 4 * q1) * p, B, LD)] = v[it][4 * q0 + q1];
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q < r) buf[lidx<IL, SW>(b[it], ob[it] + q * p, B, LD)] = v[it][q];
            }
        }
    }
}

template <typename R, bool IL>
__device__ __forceinline__ void lds_passes_inplace(cplx<R>* buf, int B, int N, int npass, const int* radix,
                                                   const cplx<R>* tw, int sign) {
    int p = 1;
    for (int ps = 0; ps < npass; ++ps) {
        const int r = radix[ps];
        inplace_pass<R, IL>(buf, B, N, r, p, tw, sign, N);
        __syncthreads();
        p *= r;
    }
}

// Butterflies per thread of an in-place pass of radix r over B lines of N.
constexpr int inplace_kb(int B, int N, int r, int threads) {
    return (B * (N / r) + threads - 1) / threads;
}

// The passes of CtPlan<NC> in place over BC lines stored LD apart, with THR
// threads per workgroup.
template <typename R, bool IL, int NC, int BC, int LD = NC, int THR = PUFF_Z_THREADS, int SW = 0, int K = 0,
          int P = 1>
__device__ __forceinline__ void ct_passes_inplace(cplx<R>* buf, const cplx<R>* tw, int sign) {
    if constexpr (K < CtPlan<NC>::np) {
        constexpr int r = CtPlan<NC>::r[K];
        constexpr int KB = inplace_kb(BC, NC, r, THR);
        inplace_pass<R, IL, KB, SW>(buf, BC, NC, r, P, tw, sign, LD);
        __syncthreads();
        ct_passes_inplace<R, IL, NC, BC, LD, THR, SW, K + 1, P * r>(buf, tw, sign);
    }
}

// The z2 buffers' swizzle (lpad's SW) for B interleaved lines of N points
// of R: B a power of two whose 16 B elements fill whole 256-byte rows of
// LDS banks, so that the first radix-16 pass's writes would land on one
// bank; the block number goes into the low bits.  N a multiple of 16, so
// that the buffer is whole blocks and the swizzle permutes within it.
#ifndef PUFF_Z_SWIZZLE
#define PUFF_Z_SWIZZLE 1
#endif
template <typename R, int B, int N>
constexpr int z_swizzle() {
    if (!PUFF_Z_SWIZZLE || B <= 0 || (B & (B - 1)) != 0 || N % 16 != 0) return 0;
    if ((16 * B * (int)sizeof(cplx<R>)) % 256 != 0) return 0;
    int sh = 0;
    while ((1 << sh) < 16 * B) ++sh;
    return sh | ((16 * B - 1) << 8);
}

template <typename R, bool IL, int NC>
__device__ __forceinline__ cplx<R>* lds_fft(cplx<R>* src, cplx<R>* dst, int B, const FftAxis& ax,
                                            const cplx<R>* tw, int sign) {
    if constexpr (NC > 0) {
        if constexpr (CtPlan<NC>::rader) return lds_rader<R, IL, NC>(src, dst, B, ax, tw, sign);
        else return ct_passes<R, IL, NC>(src, dst, B, NC, NC, tw, sign);
    } else {
        if (ax.rader) return lds_rader<R, IL, 0>(src, dst, B, ax, tw, sign);
        return lds_passes<R, IL>(src, dst, B, ax.n, ax.n, ax.npass, ax.radix, tw, sign);
    }
}

// Rader's algorithm in place, for the compile-time prime lengths (257): one
// LDS buffer of BC lines instead of two, so twice the lines per workgroup at
// the same LDS, and the radix-16 passes of the (N - 1)-point transforms
// keep BC x 16 threads busy.  The two permutations go through registers
// (each thread reads its elements, the workgroup passes a barrier, the
// elements go to their new places); X_0 = x_0 + A_0 waits in x0s[b].  The
// permutation tables are read from LDS (rps: g^j at j < N - 1, then g^-q at
// N + q), staged by the kernel: read from global memory, the compiler kept
// both tables' entries in registers across the kernel (191 VGPRs).
template <typename R, int NC>
constexpr bool rader_ip() { return sizeof(R) == 4 && NC > 0 && CtPlan<NC>::rader; }

// rps[j] = g^j (j < n - 1), rps[n + q] = g^-q (q < n - 1).
__device__ __forceinline__ void stage_rader_perm(int* rps, const FftAxis& ax) {
    for (int w = threadIdx.x; w < ax.n - 1; w += blockDim.x) {
        rps[w] = ax.rperm[w];
        rps[ax.n + w] = ax.riperm[w];
    }
}

// The middle of the in-place Rader transform, between the permutations:
// buf holds a_j = x[g^j] at element j < M and x_0 at element M; on return
// element q < M holds c_q (X[g^-q] = x_0 + c_q, x_0 still at element M) and
// x0s[b] holds X_0.
template <typename R, bool IL, int NC, int BC>
__device__ __forceinline__ void rader_ip_core(cplx<R>* buf, cplx<R>* x0s, const FftAxis& ax, const cplx<R>* tw,
                                              int sign) {
    constexpr int N = NC, M = N - 1;
    constexpr int THR = PUFF_RADER_THREADS;
    const cplx<R>* rb = reinterpret_cast<const cplx<R>*>(ax.rb) + (sign > 0 ? M : 0);
    ct_passes_inplace<R, IL, M, BC, N, THR>(buf, tw, -1);
    // C = A . rb_s (rb carries 1/M); X_0 = x_0 + A_0
    for (int w = threadIdx.x; w < BC * M; w += THR) {
        int b, k;
        if (IL) { b = w % BC; k = w / BC; } else { k = w % M; b = w / M; }
        const cplx<R> v = buf[lidx<IL>(b, k, BC, N)];
        if (k == 0) x0s[b] = cadd(buf[lidx<IL>(b, M, BC, N)], v);
        buf[lidx<IL>(b, k, BC, N)] = cmul(v, rb[k]);
    }
    __syncthreads();
    ct_passes_inplace<R, IL, M, BC, N, THR>(buf, tw, +1);
}

template <typename R, bool IL, int NC, int BC>
__device__ __forceinline__ void lds_rader_ip(cplx<R>* buf, cplx<R>* x0s, const int* rps, const FftAxis& ax,
                                             const cplx<R>* tw, int sign) {
    constexpr int N = NC, M = N - 1;
    constexpr int THR = PUFF_RADER_THREADS;
    constexpr int KP = (BC * N + THR - 1) / THR;   // elements per thread in a permutation
    // a_j = x[g^j] at element j < M; x_0 to element M
    {
        cplx<R> v[KP];
        const int tid = opaque_tid();
#pragma unroll
        for (int u = 0; u < KP; ++u) {
            const int w = tid + u * THR;
            int b, j;
            if (IL) { b = w % BC; j = w / BC; } else { j = w % N; b = w / N; }
            const bool in = w < BC * N;
            v[u] = buf[in ? lidx<IL>(b, j < M ? rps[j] : 0, BC, N) : 0];
        }
        __syncthreads();
        const int tid2 = opaque_tid();
#pragma unroll
        for (int u = 0; u < KP; ++u) {
            const int w = tid2 + u * THR;
            int b, j;
            if (IL) { b = w % BC; j = w / BC; } else { j = w % N; b = w / N; }
            if (w < BC * N) buf[lidx<IL>(b, j < M ? j : M, BC, N)] = v[u];
        }
        __syncthreads();
    }
    rader_ip_core<R, IL, NC, BC>(buf, x0s, ax, tw, sign);
    // X[g^-q] = x_0 + c_q; X_0 to element 0
    {
        cplx<R> v[KP];
        const int tid = opaque_tid();
#pragma unroll
        for (int u = 0; u < KP; ++u) {
            const int w = tid + u * THR;
            int b, q;
            if (IL) { b = w % BC; q = w / BC; } else { q = w % N; b = w / N; }
            const bool in = w < BC * N;
            const int bb = in ? b : 0;
            const cplx<R> c = buf[lidx<IL>(bb, q < M ? q : 0, BC, N)];
            const cplx<R> x0 = buf[lidx<IL>(bb, M, BC, N)];
            v[u] = q < M ? cadd(x0, c) : x0s[bb];
        }
        __syncthreads();
        const int tid2 = opaque_tid();
#pragma unroll
        for (int u = 0; u < KP; ++u) {
            const int w = tid2 + u * THR;
            int b, q;
            if (IL) { b = w % BC; q = w / BC; } else { q = w % N; b = w / N; }
            if (w < BC * N) buf[lidx<IL>(b, q < M ? rps[N + q] : 0, BC, N)] = v[u];
        }
        __syncthreads();
    }
}

// Line transform of the x / y passes: in place for rader_ip lengths (dst
// unused; x0s holds BC values and then the permutation tables), else
// lds_fft.  Returns the buffer holding the result.
template <typename R, bool IL, int NC, int BC>
__device__ __forceinline__ cplx<R>* line_fft(cplx<R>* src, cplx<R>* dst, cplx<R>* x0s, int B, const FftAxis& ax,
                                             const cplx<R>* tw, int sign) {
    if constexpr (rader_ip<R, NC>()) {
        lds_rader_ip<R, IL, NC, BC>(src, x0s, reinterpret_cast<const int*>(x0s + BC), ax, tw, sign);
        return src;
    } else {
        return lds_fft<R, IL, NC>(src, dst, B, ax, tw, sign);
    }
}

// z2-line transform: in place (one LDS buffer) or ping-pong.
template <typename R, bool INPLACE, int NC, int BC>
__device__ __forceinline__ cplx<R>* z_fft(cplx<R>* src, cplx<R>* dst, int B, const FftAxis& ax,
                                          const cplx<R>* tw, int sign) {
    if constexpr (INPLACE) {
        if constexpr (NC > 0) ct_passes_inplace<R, true, NC, BC, NC, PUFF_Z_THREADS, z_swizzle<R, BC, NC>()>(src, tw, sign);
        else lds_passes_inplace<R, true>(src, B, ax.n, ax.npass, ax.radix, tw, sign);
        return src;
    } else {
        return lds_fft<R, true, NC>(src, dst, B, ax, tw, sign);
    }
}

// ---- x lines (contiguous) --------------------------------------------------
// OP (an XOP_ value) is a template parameter: each fused operation gets its
// own register allocation.
template <typename R, int NC, int BC, int OP>
__global__ __launch_bounds__(256) void k_fft_x(DiffArgs<R> a, int Bp, int sign) {
    constexpr int op = OP;
    extern __shared__ unsigned char smem_raw[];
    const int N = NC ? NC : a.nx;
    const int B = BC ? BC : Bp;
    const FftAxis& ax = a.ax;
    cplx<R>* s0 = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* s1 = rader_ip<R, NC>() ? s0 : s0 + lpad<false>(B * N);
    const int64_t nlines = (int64_t)a.ny * a.nzp;
    const int64_t L0 = (int64_t)blockIdx.x * B;
    const int nb = (int)min<int64_t>(B, nlines - L0);
    const cplx<R>* twg = reinterpret_cast<const cplx<R>*>(ax.tw);
    cplx<R>* tw = s1 + lpad<false>(B * N);   // twiddle table staged in LDS
    cplx<R>* x0s = tw + N;            // in-place Rader's X_0 per line and permutations
    for (int w = threadIdx.x; w < N; w += blockDim.x) tw[w] = twg[w];
    if constexpr (rader_ip<R, NC>()) stage_rader_perm(reinterpret_cast<int*>(x0s + B), ax);
    cplx<R>* base = a.a + L0 * N;
    const int64_t ntr = (int64_t)a.nx * a.ny;
    // plane and row of the block's first line; line L0 + b is row iy0 + b
    // counted on from there (b < B)
    const int iy0 = (int)(L0 % a.ny), iz0 = (int)(L0 / a.ny);
    const double rff = 1.0 / a.ffact;
    if constexpr (rader_ip<R, NC>() && (OP == XOP_FORWARD || OP == XOP_INV_SPLIT_FWD || OP == XOP_INV_RECOMBINE)) {
        // Rader's permutations folded into the global reads and writes (as
        // in the y pass; here they scatter within each line's 2 KB): the
        // input permutation into the loads, the output permutation into the
        // stores, and for the split, the inverse transform's output
        // permutation and the forward transform's input permutation into
        // one move within LDS (natural index g^-q = g^(M - q): element q
        // goes to element (M - q) mod M, X_0 to element M)
        constexpr int M = NC - 1;
        constexpr int THR = PUFF_RADER_THREADS;
        constexpr int KP = (BC * NC + THR - 1) / THR;
        const int* rps = reinterpret_cast<const int*>(x0s + BC);
        __syncthreads();   // rps staged
        // the recombination's interior part, read with the lines (read in
        // the final loop, each load waits on the memory latency there)
        cplx<R> anbp[OP == XOP_INV_RECOMBINE ? KP : 1];
        if constexpr (OP == XOP_INV_RECOMBINE) {
            const int tid = opaque_tid();
#pragma unroll
            for (int u = 0; u < KP; ++u) {
                const int w = tid + u * THR;
                const int ww = w < BC * NC ? w : 0;
                const int b = ww / NC, q = ww % NC;
                const int n = q < M ? rps[NC + q] : 0;
                anbp[u] = (w < BC * NC && b < nb) ? a.anb[(L0 + b) * N + n] : cplx<R>{0, 0};
            }
        }
        lds_fill<false, STAGE_UNROLL>(s0, BC * NC, [&](int w) {
            const int b = w / NC, k = w % NC;
            const int j = k < M ? rps[k] : 0;   // element k < M: x[g^k]; element M: x_0
            return b < nb ? base[b * NC + j] : cplx<R>{0, 0};
        });
        __syncthreads();
        rader_ip_core<R, false, NC, BC>(s0, x0s, ax, tw, OP == XOP_FORWARD ? sign : +1);
        if constexpr (OP == XOP_INV_SPLIT_FWD) {
            // /ffact then AbsorptionStep's split (diffraction.f90:310, 454-470)
            cplx<R>* anb = a.anb + L0 * N;
            cplx<R> v[KP];
            int e[KP];
            const int tid = opaque_tid();
#pragma unroll
            for (int u = 0; u < KP; ++u) {
                const int w = tid + u * THR;
                const int ww = w < BC * NC ? w : 0;
                const int b = ww / NC, q = ww % NC;
                const cplx<R> c = s0[lidx<false>(b, q < M ? q : 0, BC, NC)];
                const cplx<R> xz = s0[lidx<false>(b, M, BC, NC)];
                const cplx<R> x = q < M ? cadd(xz, c) : x0s[b];   // X[g^-q] = x_0 + c_q; X_0
                const int n = q < M ? rps[NC + q] : 0;
                e[u] = lidx<false>(b, q < M ? (q ? M - q : 0) : M, BC, NC);
                const int lb = iy0 + b;   // (iy, iz) of line L0 + b
                const int iy = lb % a.ny, iz = iz0 + lb / a.ny;
                const bool st = w < BC * NC && b < nb;
                const double mxy = st ? a.mxy[(int64_t)iy * a.nx + n] : 0.0;
                const double m = mxy + ((st ? a.mz[iz] : 0.0) * (1.0 - mxy));
                const double vr = (double)x.x * rff, vi = (double)x.y * rff;
                if (st) anb[b * NC + n] = {(R)((1.0 - m) * vr), (R)((1.0 - m) * vi)};
                v[u] = {(R)(m * vr), (R)(m * vi)};
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < KP; ++u)
                if (tid + u * THR < BC * NC) s0[e[u]] = v[u];
            __syncthreads();
            rader_ip_core<R, false, NC, BC>(s0, x0s, ax, tw, -1);
        }
        {
            cplx<R> v[KP];
            int n[KP];
            const int tid = opaque_tid();
#pragma unroll
            for (int u = 0; u < KP; ++u) {
                const int w = tid + u * THR;
                const int ww = w < BC * NC ? w : 0;
                const int b = ww / NC, q = ww % NC;
                const cplx<R> c = s0[lidx<false>(b, q < M ? q : 0, BC, NC)];
                const cplx<R> xz = s0[lidx<false>(b, M, BC, NC)];
                v[u] = q < M ? cadd(xz, c) : x0s[b];   // X[g^-q] = x_0 + c_q; X_0
                n[u] = q < M ? rps[NC + q] : 0;
            }
#pragma unroll
            for (int u = 0; u < KP; ++u) {
                const int w = tid + u * THR;
                const int b = w / NC;
                if (w >= BC * NC || b >= nb) continue;
                cplx<R> x = v[u];
                if constexpr (OP == XOP_INV_RECOMBINE) {
                    // /ffact + interior (diffraction.f90:513-517)
                    const cplx<R> qn = anbp[u];
                    x = {(R)((double)x.x * rff) + qn.x, (R)((double)x.y * rff) + qn.y};
                    // keep the aliased plane NZ2 equal to plane 1 (para_field.f90:3197-3244)
                    const int64_t L = L0 + b;
                    if (a.periodic && L < a.ny) a.a[(int64_t)(a.nz2 - 1) * ntr + L * N + n[u]] = x;
                }
                base[b * NC + n[u]] = x;
            }
        }
        return;
    }
    if (op == XOP_ADJ_FOLD_SAVE_FWD) {
        for (int w = threadIdx.x; w < B * N; w += blockDim.x) {
            const int b = w / N;
            cplx<R> v = (b < nb) ? base[w] : cplx<R>{0, 0};
            if (b < nb) {
                const int64_t L = L0 + b;
                if (a.periodic && L < a.ny) {
                    // the output's aliased plane copies plane 1: fold its cotangent
                    cplx<R>* al = a.a + (int64_t)(a.nz2 - 1) * ntr + L * N + (w % N);
                    const cplx<R> t = *al;
                    v.x += t.x;
                    v.y += t.y;
                    *al = cplx<R>{0, 0};
                }
                a.anb[L0 * N + w] = v;
            }
            s0[lpad<false>(w)] = v;
        }
    } else {
        const int nv = nb * N;
        lds_fill<false, STAGE_UNROLL>(s0, B * N, [&](int w) { return w < nv ? base[w] : cplx<R>{0, 0}; });
    }
    __syncthreads();
    const int first_sign = (op == XOP_FORWARD) ? sign : (op == XOP_ADJ_FOLD_SAVE_FWD ? -1 : +1);
    // the backward transform's 1/ffact: a divide in FP64 (the reference's
    // arithmetic), a reciprocal multiply (rff) in the mixed form
    auto unscale = [&](R v) -> double {
        if constexpr (sizeof(R) == 8) return (double)v / a.ffact;
        else return (double)v * rff;
    };
    cplx<R>* res = line_fft<R, false, NC, BC>(s0, s1, x0s, B, ax, tw, first_sign);
    if (op == XOP_ADJ_INV_MIX_FWD) {
        // adjoint of the absorption split: (1 - m) cotangent + m W^H G W cotangent / ffact
        const cplx<R>* lam = a.anb + L0 * N;
        for (int w = threadIdx.x; w < nb * N; w += blockDim.x) {
            const int b = w / N, j = w % N;
            const int lb = iy0 + b;   // (iy, iz) of line L0 + b
            const int iy = lb % a.ny, iz = iz0 + lb / a.ny;
            const double mxy = a.mxy[(int64_t)iy * a.nx + j];
            const double m = mxy + (a.mz[iz] * (1.0 - mxy));
            const cplx<R> v = res[lpad<false>(w)];
            const cplx<R> l = lam[w];
            res[lpad<false>(w)] = {(R)((1.0 - m) * (double)l.x + m * unscale(v.x)),
                            (R)((1.0 - m) * (double)l.y + m * unscale(v.y))};
        }
        __syncthreads();
        cplx<R>* other = (res == s0) ? s1 : s0;
        res = line_fft<R, false, NC, BC>(res, other, x0s, B, ax, tw, -1);
    } else if (op == XOP_ADJ_INV_SCALE) {
        for (int w = threadIdx.x; w < nb * N; w += blockDim.x) {
            const cplx<R> v = res[lpad<false>(w)];
            res[lpad<false>(w)] = {(R)unscale(v.x), (R)unscale(v.y)};
        }
        __syncthreads();
    } else if (op == XOP_INV_SPLIT_FWD) {
        // /ffact then AbsorptionStep's split (diffraction.f90:310, 454-470)
        cplx<R>* anb = a.anb + L0 * N;
        for (int w = threadIdx.x; w < nb * N; w += blockDim.x) {
            const int b = w / N, j = w % N;
            const int lb = iy0 + b;   // (iy, iz) of line L0 + b
            const int iy = lb % a.ny, iz = iz0 + lb / a.ny;
            const double mxy = a.mxy[(int64_t)iy * a.nx + j];
            const double m = mxy + (a.mz[iz] * (1.0 - mxy));
            cplx<R> v = res[lpad<false>(w)];
            const double vr = unscale(v.x), vi = unscale(v.y);
            anb[w] = {(R)((1.0 - m) * vr), (R)((1.0 - m) * vi)};
            res[lpad<false>(w)] = {(R)(m * vr), (R)(m * vi)};
        }
        __syncthreads();
        cplx<R>* other = (res == s0) ? s1 : s0;
        res = line_fft<R, false, NC, BC>(res, other, x0s, B, ax, tw, -1);
    } else if (op == XOP_INV_RECOMBINE) {
        // /ffact + interior (diffraction.f90:513-517)
        const cplx<R>* anb = a.anb + L0 * N;
        for (int w = threadIdx.x; w < nb * N; w += blockDim.x) {
            const cplx<R> v = res[lpad<false>(w)];
            const cplx<R> q = anb[w];
            if (sizeof(R) == 8) {
                res[lpad<false>(w)] = {(R)(unscale(v.x) + (double)q.x), (R)(unscale(v.y) + (double)q.y)};
            } else {
                res[lpad<false>(w)] = {(R)unscale(v.x) + q.x, (R)unscale(v.y) + q.y};
            }
        }
        __syncthreads();
    }
    lds_drain<false, STAGE_UNROLL>(res, nb * N, [&](int w, const cplx<R>& v) { base[w] = v; });
    if (op == XOP_INV_RECOMBINE && a.periodic) {
        // keep the aliased plane NZ2 equal to plane 1 (para_field.f90:3197-3244)
        for (int w = threadIdx.x; w < nb * N; w += blockDim.x) {
            const int64_t L = L0 + w / N;
            if (L < a.ny) a.a[(int64_t)(a.nz2 - 1) * ntr + L * N + (w % N)] = res[lpad<false>(w)];
        }
    }
}

// ---- persistent tile loops -----------------------------------------------------
// With compile-time plans the x and y passes run persistent workgroups (as
// many as fit the device at once): each loops over its tiles of B lines,
// loading tile t + stride into registers while tile t is transformed in LDS
// and stored.  A one-tile workgroup has its loads in flight only between
// its start and its first barrier; here they overlap the transform and the
// stores of the previous tile.  The tiles are dealt to the XCDs in
// contiguous ranges (workgroup g runs on XCD g mod 8, the grid being a
// multiple of 8), so that neighbouring tiles, which share cache lines at
// their edges, meet in one L2.
// Off by default: on the C1 and C4 meshes the persistent y pass measured
// 1-2 % slower events than one tile per workgroup (its prefetch registers
// cost a wave per SIMD of occupancy).
#ifndef PUFF_PERSIST_Y
#define PUFF_PERSIST_Y 0
#endif
struct TileRange {
    int64_t t, end;
    int stride;
};
__device__ __forceinline__ TileRange tile_range(int64_t ntile) {
    const int per_xcd = (int)gridDim.x / NXCD;
    const int xcd = (int)blockIdx.x % NXCD, l = (int)blockIdx.x / NXCD;
    const int64_t per = (ntile + NXCD - 1) / NXCD;
    const int64_t lo = xcd * per;
    return {lo + l, min(lo + per, ntile), per_xcd};
}

// ---- y lines (stride nx), B consecutive x per block -------------------------
template <typename R, int NC, int BC>
__global__ __launch_bounds__(256) void k_fft_y(DiffArgs<R> a, int Bp, int sign) {
    extern __shared__ unsigned char smem_raw[];
    const int N = NC ? NC : a.ny;
    const int B = BC ? BC : Bp;
    const FftAxis& ay = a.ay;
    cplx<R>* s0 = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* s1 = rader_ip<R, NC>() ? s0 : s0 + lpad<true>(B * N);
    // blocks of B x-columns, x fastest; neighbouring blocks share the cache
    // lines at their edges, so they go to one XCD
    const int nxb = (a.nx + B - 1) / B;
    constexpr bool persist = NC > 0 && PUFF_PERSIST_Y;
    const int64_t blk = persist ? 0 : xcd_block((int64_t)nxb * a.nzp);
    if (blk >= (int64_t)nxb * a.nzp) return;
    const int x0 = (int)(blk % nxb) * B;
    const int iz = (int)(blk / nxb);
    const int nb = min(B, a.nx - x0);
    const cplx<R>* twg = reinterpret_cast<const cplx<R>*>(ay.tw);
    cplx<R>* tw = s1 + lpad<true>(B * N);   // twiddle table staged in LDS
    cplx<R>* x0s = tw + N;            // in-place Rader's X_0 per line and permutations
    for (int w = threadIdx.x; w < N; w += blockDim.x) tw[w] = twg[w];
    if constexpr (rader_ip<R, NC>()) stage_rader_perm(reinterpret_cast<int*>(x0s + B), ay);
    if constexpr (persist) {
        // persistent: tiles of B x-columns of one plane, x fastest
        static_assert(PUFF_RADER_THREADS == 256, "the tile loop assumes 256 threads");
        constexpr int U = (BC * NC + 255) / 256;   // elements per thread (256 threads)
        const int64_t ntile = (int64_t)nxb * a.nzp;
        const TileRange tr = tile_range(ntile);
        cplx<R> v[U];
        auto load = [&](int64_t t) {
            const int tx0 = (int)(t % nxb) * BC, tz = (int)(t / nxb);
            const int tnb = min(BC, a.nx - tx0);
            const cplx<R>* tb = a.a + (int64_t)tz * a.nx * a.ny + tx0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int w = (int)threadIdx.x + u * 256;
                const int b = w % BC, j = w / BC;
                v[u] = (w < BC * NC && b < tnb) ? tb[(int64_t)j * a.nx + b] : cplx<R>{0, 0};
            }
        };
        if (tr.t < tr.end) load(tr.t);
        for (int64_t t = tr.t; t < tr.end; t += tr.stride) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int w = (int)threadIdx.x + u * 256;
                if (w < BC * NC) s0[lpad<true>(w)] = v[u];
            }
            __syncthreads();
            if (t + tr.stride < tr.end) load(t + tr.stride);
            cplx<R>* res = line_fft<R, true, NC, BC>(s0, s1, x0s, BC, ay, tw, sign);
            const int tx0 = (int)(t % nxb) * BC, tz = (int)(t / nxb);
            const int tnb = min(BC, a.nx - tx0);
            cplx<R>* tb = a.a + (int64_t)tz * a.nx * a.ny + tx0;
            lds_drain<true, STAGE_UNROLL>(res, BC * NC, [&](int w, const cplx<R>& vv) {
                const int b = w % BC, j = w / BC;
                if (b < tnb) tb[(int64_t)j * a.nx + b] = vv;
            });
            __syncthreads();
        }
        return;
    }
    cplx<R>* base = a.a + (int64_t)iz * a.nx * a.ny + x0;
    if constexpr (rader_ip<R, NC>()) {
        // Rader's permutations folded into the global reads and writes: the
        // permutation moves whole rows of B x-columns, so the accesses stay
        // as contiguous as the plain ones, and the transform saves two LDS
        // round trips and their barriers
        constexpr int M = NC - 1;
        const int* rps = reinterpret_cast<const int*>(x0s + BC);
        __syncthreads();   // rps staged
        lds_fill<true, STAGE_UNROLL>(s0, BC * NC, [&](int w) {
            const int b = w % BC, k = w / BC;
            const int j = k < M ? rps[k] : 0;   // element k < M: x[g^k]; element M: x_0
            return (b < nb) ? base[(int64_t)j * a.nx + b] : cplx<R>{0, 0};
        });
        __syncthreads();
        rader_ip_core<R, true, NC, BC>(s0, x0s, ay, tw, sign);
        for (int w0 = threadIdx.x; w0 < BC * NC; w0 += STAGE_UNROLL * (int)blockDim.x) {
            cplx<R> v[STAGE_UNROLL];
            int j[STAGE_UNROLL];
#pragma unroll
            for (int u = 0; u < STAGE_UNROLL; ++u) {
                const int w = w0 + u * (int)blockDim.x;
                const int ww = w < BC * NC ? w : 0;
                const int b = ww % BC, q = ww / BC;
                const cplx<R> c = s0[lidx<true>(b, q < M ? q : 0, BC, NC)];
                const cplx<R> xz = s0[lidx<true>(b, M, BC, NC)];
                v[u] = q < M ? cadd(xz, c) : x0s[b];   // X[g^-q] = x_0 + c_q; X_0
                j[u] = q < M ? rps[NC + q] : 0;
            }
#pragma unroll
            for (int u = 0; u < STAGE_UNROLL; ++u) {
                const int w = w0 + u * (int)blockDim.x;
                const int b = w % BC;
                if (w < BC * NC && b < nb) base[(int64_t)j[u] * a.nx + b] = v[u];
            }
        }
        return;
    }
    lds_fill<true, STAGE_UNROLL>(s0, B * N, [&](int w) {
        const int b = w % B, j = w / B;
        return (b < nb) ? base[(int64_t)j * a.nx + b] : cplx<R>{0, 0};
    });
    __syncthreads();
    cplx<R>* res = line_fft<R, true, NC, BC>(s0, s1, x0s, B, ay, tw, sign);
    lds_drain<true, STAGE_UNROLL>(res, B * N, [&](int w, const cplx<R>& v) {
        const int b = w % B, j = w / B;
        if (b < nb) base[(int64_t)j * a.nx + b] = v;
    });
}

// The spectral factor between the forward and inverse z2 transforms, at
// wavenumbers (kx, ky, kz): multiplyexp's propagator (with its filter) or
// AbsorptionStep's damping.  FP64 form: in FP64, rounded once.  Mixed form
// (an FP32 field): the propagator phase is formed in FP64 and reduced
// modulo 2 pi there (Cody-Waite, with 2 pi split into a double and its
// remainder), so that the FP32 sine and cosine see |r| <= pi; the damping
// exponent is formed in FP64 and exponentiated in FP32.  FP64 sine, cosine
// and exponential per node cost C4's event 1.2 ms of 22.7.  The mixed form
// also takes rz, 1/(2 kz) (propagator) or 1/sqrt|2 kz| (damping), from a
// per-wavenumber table (kzr) instead of an FP64 divide and square root per
// node; the FP64 form ignores it.
template <typename R>
__device__ __forceinline__ cplx<R> spectral_op(const DiffArgs<R>& a, int op, double kx, double ky, double kz,
                                               double rz, cplx<R> v) {
    if constexpr (sizeof(R) == 4) {
        if (op == SPEC_PROPAGATE) {
            if ((kz > a.cutoff) || (kz < -a.cutoff)) {
                if (kz != 0.0) {
                    constexpr double two_pi_hi = 6.283185307179586232;      // 2 pi rounded to double
                    constexpr double two_pi_lo = 2.4492935982947064e-16;    // 2 pi - two_pi_hi
                    constexpr double inv_two_pi = 0.15915494309189533577;
                    const double th = a.h * (kx * kx + ky * ky) * rz;
                    const double n = rint(th * inv_two_pi);
                    const double r = fma(-n, two_pi_lo, fma(-n, two_pi_hi, th));
                    float sn, cs;
                    sincosf((float)r, &sn, &cs);
                    v = {cs * v.x - sn * v.y, sn * v.x + cs * v.y};
                }
            } else if (a.filter) {
                v = {0.0f, 0.0f};
            }
        } else if (kz != 0.0) {
            const float f = expf((float)(-a.h * a.beta * (fabs(kx) + fabs(ky)) * rz));
            v = {f * v.x, f * v.y};
        }
        return v;
    }
    if (op == SPEC_PROPAGATE) {
        // multiplyexp (diffraction.f90:150-168)
        if ((kz > a.cutoff) || (kz < -a.cutoff)) {
            if (kz != 0.0) {
                const double th = a.h * (kx * kx + ky * ky) / (2.0 * kz);
                double sn, cs;
                sincos(th, &sn, &cs);
                const double vr = (double)v.x, vi = (double)v.y;
                v = {(R)(cs * vr - sn * vi), (R)(sn * vr + cs * vi)};
            }
        } else if (a.filter) {
            v = {(R)0, (R)0};
        }
    } else {
        // AbsorptionStep damping (diffraction.f90:484-497)
        if (kz != 0.0) {
            const double f = exp(-a.h * a.beta * (fabs(kx) + fabs(ky)) / (sqrt(fabs(2.0 * kz))));
            v = {(R)(f * (double)v.x), (R)(f * (double)v.y)};
        }
    }
    return v;
}

// ---- z2 lines (stride nx ny), B consecutive x per block ----------------------
// op < 0: plain transform with `sign`; else forward, spectral op, inverse.
template <typename R, bool INPLACE, int NC, int BC>
__global__ __launch_bounds__(PUFF_Z_THREADS) void k_fft_z(DiffArgs<R> a, int Bp, int op, int sign) {
    extern __shared__ unsigned char smem_raw[];
    const int N = NC ? NC : a.nzp;
    const int B = BC ? BC : Bp;
    const FftAxis& az = a.az;
    cplx<R>* s0 = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* s1 = INPLACE ? s0 : s0 + lpad<true>(B * N);
    // blocks of B x-columns, x fastest; neighbouring blocks share the cache
    // lines of every z2 plane, so they go to one XCD
    const int nxb = (a.nx + B - 1) / B;
    const int64_t blk = xcd_block((int64_t)nxb * a.ny);
    if (blk >= (int64_t)nxb * a.ny) return;
    const int x0 = (int)(blk % nxb) * B;
    const int iy = (int)(blk / nxb);
    const int nb = min(B, a.nx - x0);
    const int64_t ntr = (int64_t)a.nx * a.ny;
    // few lines per workgroup here: the twiddles stay in global memory (L2)
    const cplx<R>* tw = reinterpret_cast<const cplx<R>*>(az.tw);
    cplx<R>* base = a.a + (int64_t)iy * a.nx + x0;
    // the in-place compile-time passes' swizzled buffer (identity otherwise)
    constexpr int SW = (INPLACE && NC > 0) ? z_swizzle<R, BC, NC>() : 0;
    // FP64 compile-time plans: kx and the kz of the thread's elements for the
    // spectral op, loaded with the lines (a load per element in the spectral
    // loop waits on the memory latency each time); the thread's elements
    // share one x-column
    constexpr int KPT = (sizeof(R) == 8 && NC > 0 && BC > 0 && (BC * NC) % PUFF_Z_THREADS == 0 &&
                         PUFF_Z_THREADS % BC == 0)
                            ? (BC * NC) / PUFF_Z_THREADS
                            : 0;
    double kzp[KPT > 0 ? KPT : 1];
    double kxp = 0.0;
    if constexpr (KPT > 0) {
        if (op >= 0) {
            const int b = (int)threadIdx.x % BC;
            kxp = a.kx[x0 + (b < nb ? b : 0)];
#pragma unroll
            for (int it = 0; it < KPT; ++it) kzp[it] = a.kz[((int)threadIdx.x + it * PUFF_Z_THREADS) / BC];
        }
    }
    lds_fill<true, PUFF_Z_UNROLL, SW>(s0, B * N, [&](int w) {
        const int b = w % B, j = w / B;
        return (b < nb) ? base[(int64_t)j * ntr + b] : cplx<R>{0, 0};
    });
    __syncthreads();
    cplx<R>* res;
    if (op < 0) {
        res = z_fft<R, INPLACE, NC, BC>(s0, s1, B, az, tw, sign);
    } else {
        res = z_fft<R, INPLACE, NC, BC>(s0, s1, B, az, tw, -1);
        const double ky = a.ky[iy];
        if constexpr (KPT > 0) {
#pragma unroll
            for (int it = 0; it < KPT; ++it) {
                const int w = (int)threadIdx.x + it * PUFF_Z_THREADS;
                if (w % BC < nb) res[lpad<true, SW>(w)] = spectral_op(a, op, kxp, ky, kzp[it], 0.0, res[lpad<true, SW>(w)]);
            }
        } else {
            for (int w = threadIdx.x; w < B * N; w += blockDim.x) {
                const int b = w % B, q = w / B;
                if (b >= nb) continue;
                const double rz = sizeof(R) == 4 ? a.kzr[2 * q + (op == SPEC_ABSORB ? 1 : 0)] : 0.0;
                res[lpad<true, SW>(w)] = spectral_op(a, op, a.kx[x0 + b], ky, a.kz[q], rz, res[lpad<true, SW>(w)]);
            }
        }
        __syncthreads();
        cplx<R>* other = (res == s0) ? s1 : s0;
        res = z_fft<R, INPLACE, NC, BC>(res, other, B, az, tw, +1);
    }
    lds_drain<true, PUFF_Z_UNROLL, SW>(res, B * N, [&](int w, const cplx<R>& v) {
        const int b = w % B, j = w / B;
        if (b < nb) base[(int64_t)j * ntr + b] = v;
    });
}

// ---- z2 lines of 4096 points as 64 x 64 (four-step) ---------------------------
// A 4096-point line fills 32 KiB of LDS (FP32), so a single-pass workgroup
// holds one or two x-columns and reads 8 or 16 useful bytes of every cache
// line of the 4096 planes it visits (C4's z2 pass fetched 1.4x and wrote
// 3.5x its bytes).  Split as n = 64 n1 + n2, k = k1 + 64 k2:
//   X[k1 + 64 k2] = sum_n2 W64^(n2 k2) W^(n2 k1) sum_n1 x[64 n1 + n2] W64^(n1 k1)
// (W = exp(-2 pi i / 4096)), the transform runs as passes over 64-point
// sub-lines, each workgroup holding BX adjacent x-columns (one 128-byte row
// of every plane it visits) of SETS sub-lines:
//   Z4_FIRST, set n2: transform over n1 (planes 64 n1 + n2), times
//     W^(s n2 k1); result k1 back in plane 64 k1 + n2.
//   Z4_MID, set k1: transform over n2 (planes 64 k1 + n2) -> X[k1 + 64 k2].
//     With a spectral op: the op at q = k1 + 64 k2, the inverse transform
//     over k2 and W^(-n2 k1) (the first steps of the inverse four-step),
//     back in plane 64 k1 + n2.  Without one (plain transform): X written
//     to plane k1 + 64 k2 of `out`.
//   Z4_LAST, set n2: transform over k1 (planes 64 k1 + n2) -> x[64 n1 + n2].
// Each pass moves the field once through HBM in whole rows.
enum { Z4_FIRST = 0, Z4_MID = 1, Z4_LAST = 2 };
constexpr int Z4_M = 64;
template <typename R, int BX, int SETS, int PASS>
__global__ __launch_bounds__(PUFF_Z_THREADS) void k_fft_z4(DiffArgs<R> a, int op, int sign, cplx<R>* out) {
    constexpr int M = Z4_M;
    constexpr int N = M * M;
    constexpr int L = BX * SETS;   // sub-lines per workgroup: b = xi + BX si
    constexpr int NSG = M / SETS;
    extern __shared__ unsigned char smem_raw[];
    cplx<R>* buf = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* tw = buf + lpad<true>(L * M);   // W64^e = W^(64 e), e < 64
    const cplx<R>* twn = reinterpret_cast<const cplx<R>*>(a.az.tw);
    for (int w = threadIdx.x; w < M; w += blockDim.x) tw[w] = twn[w * (N / M)];
    // x-blocks fastest, then set groups, then rows: neighbouring x-blocks
    // share the cache lines at their edges and go to one XCD
    const int nxb = (a.nx + BX - 1) / BX;
    const int64_t nblk = (int64_t)nxb * NSG * a.ny;
    const int64_t blk = xcd_block(nblk);
    if (blk >= nblk) return;
    const int x0 = (int)(blk % nxb) * BX;
    const int sg = (int)((blk / nxb) % NSG) * SETS;
    const int iy = (int)(blk / ((int64_t)nxb * NSG));
    const int nb = min(BX, a.nx - x0);
    const int64_t ntr = (int64_t)a.nx * a.ny;
    const int64_t row = (int64_t)iy * a.nx + x0;
    cplx<R>* base = a.a + row;
    // the spectral op's kz and rz of the workgroup's z2 wavenumbers
    // (sg + si) + 64 k2, entry si + SETS k2, staged with the line loads (a
    // load per element in the spectral loop waits on the memory latency
    // each time)
    double* kzt = reinterpret_cast<double*>(tw + M);
    const bool fused = PASS == Z4_MID && op >= 0;
    // a thread's elements w = tid + 256 u share b = tid mod L (L divides
    // 256), hence its x-column and kx
    static_assert(PUFF_Z_THREADS % L == 0, "one x-column per thread");
    const int xt = ((int)threadIdx.x % L) % BX;
    double kxt = 0.0;
    if (fused) {
        kxt = a.kx[x0 + (xt < nb ? xt : 0)];
        for (int w = threadIdx.x; w < SETS * M; w += blockDim.x) {
            const int iz = (sg + w % SETS) + M * (w / SETS);
            kzt[w] = a.kz[iz];
            kzt[SETS * M + w] = sizeof(R) == 4 ? a.kzr[2 * iz + (op == SPEC_ABSORB ? 1 : 0)] : 0.0;
        }
    }
    // plane of element j of set s
    auto plane = [](int s, int j) { return PASS == Z4_MID ? M * s + j : M * j + s; };
    constexpr int SW = z_swizzle<R, L, M>();
    lds_fill<true, (L * M) / PUFF_Z_THREADS, SW>(buf, L * M, [&](int w) {
        const int b = w % L, j = w / L;
        const int xi = b % BX, si = b / BX;
        return xi < nb ? base[(int64_t)plane(sg + si, j) * ntr + xi] : cplx<R>{0, 0};
    });
    __syncthreads();
    ct_passes_inplace<R, true, M, L, M, PUFF_Z_THREADS, SW>(buf, tw, fused ? -1 : sign);
    if (fused) {
        const double ky = a.ky[iy];
        for (int w = threadIdx.x; w < L * M; w += blockDim.x) {
            const int b = w % L, k2 = w / L;
            const int xi = b % BX, si = b / BX;
            if (xi >= nb) continue;
            const int e = si + SETS * k2;
            buf[lpad<true, SW>(w)] = spectral_op(a, op, kxt, ky, kzt[e], kzt[SETS * M + e], buf[lpad<true, SW>(w)]);
        }
        __syncthreads();
        ct_passes_inplace<R, true, M, L, M, PUFF_Z_THREADS, SW>(buf, tw, +1);
    }
    lds_drain<true, (L * M) / PUFF_Z_THREADS, SW>(buf, L * M, [&](int w, const cplx<R>& v) {
        const int b = w % L, j = w / L;
        const int xi = b % BX, si = b / BX;
        if (xi >= nb) return;
        const int st = sg + si;
        if (PASS == Z4_FIRST) {
            base[(int64_t)plane(st, j) * ntr + xi] = cmul(v, tw_get(twn, st * j, sign));   // W^(s n2 k1)
        } else if (PASS == Z4_MID && !fused) {
            out[(int64_t)(st + M * j) * ntr + row + xi] = v;   // X[k1 + 64 k2]
        } else if (PASS == Z4_MID) {
            base[(int64_t)plane(st, j) * ntr + xi] = cmul(v, tw_get(twn, st * j, +1));     // W^(-n2 k1)
        } else {
            base[(int64_t)plane(st, j) * ntr + xi] = v;
        }
    });
}

// ---- host side -----------------------------------------------------------------


// Dynamic LDS above 64 KiB needs the kernel's opt-in (up to 160 KiB).
static void lds_opt_in(const void* kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

// Lines per workgroup: two LDS line buffers and the twiddle table within
// `budget` bytes.
#ifndef PUFF_Z_LDS_BUDGET
#define PUFF_Z_LDS_BUDGET (32 * 1024)
#endif
#ifndef PUFF_XY_LDS_BUDGET
#define PUFF_XY_LDS_BUDGET (32 * 1024)
#endif
static constexpr int lines_per_block(int n, size_t elem, int cap, size_t budget, bool il) {
    const size_t tw = (size_t)n * elem;
    int b = cap;
    while (b > 1 && (2 * (size_t)(il ? lpad<true>(b * n) : lpad<false>(b * n)) * elem + tw) > budget) --b;
    return b;
}

// z2 lines per workgroup and whether the pass runs in place (one LDS
// buffer): in place when every pass is radix 16, 4, 3 or 2 with at most one
// butterfly per thread, and that leaves at least as many lines per workgroup
// as two buffers do: C1's 1024-point lines in FP64 (616 -> 454 us per pass)
// and C4's 4096-point lines (7.50 -> 6.92 ms).  Otherwise two line buffers,
// no twiddle table (wider workgroups with longer runs along x measured
// slower: occupancy matters more than coalescing here).
struct ZPlan {
    int B;
    bool inplace;
};
static constexpr ZPlan z_plan(int n, size_t elem, bool rader, int npass, const int* radix) {
    bool inplace = !rader;
    int minr = 16;
    for (int k = 0; k < npass; ++k) {
        const int r = radix[k];
        if (r != 2 && r != 3 && r != 4 && r != 16) inplace = false;
        if (r < minr) minr = r;
    }
    int B = (int)((PUFF_Z_LDS_BUDGET) / (2 * (size_t)n * elem));
    B = B < 1 ? 1 : (B > 16 ? 16 : B);   // lines per workgroup with two buffers
    if (inplace) {
        int Bi = (int)((PUFF_Z_LDS_BUDGET) / ((size_t)n * elem));
        Bi = Bi > 16 ? 16 : Bi;
        const int bt = PUFF_Z_THREADS / (n / minr);   // one butterfly per thread
        Bi = Bi > bt ? bt : Bi;
        inplace = Bi >= 1 && Bi >= B;
        if (inplace) B = Bi;
    }
    return ZPlan{B, inplace};
}

// Lines per workgroup of the in-place Rader passes: one line buffer, the
// twiddle table, X_0 and the permutation tables within `budget`; at most 16
// lines (one butterfly per thread in the radix-16 passes).  22 KiB (8 of
// C4's 257-point FP32 lines, 7 workgroups per CU) measured 20.7 ms per C4
// event against 21.7 (18 KiB, 6 lines), 21.6 (26 KiB, 10 lines) and 22.1
// (40 KiB, 16 lines).
#ifndef PUFF_RADER_LDS_BUDGET
#define PUFF_RADER_LDS_BUDGET (22 * 1024)
#endif
static constexpr int lines_per_block_ip(int n, size_t elem, bool il, size_t budget = PUFF_RADER_LDS_BUDGET) {
    int b = 16;
    while (b > 1 && ((size_t)(il ? lpad<true>(b * n) : lpad<false>(b * n)) + n + b) * elem +
                            2 * n * sizeof(int) > budget)
        --b;
    return b;
}
template <typename R, int NC, bool IL>
static constexpr size_t xy_lds_elems(int B) {
    // in-place Rader: line buffer, twiddles, X_0 per line, two int tables of NC
    return rader_ip<R, NC>() ? (size_t)lpad<IL>(B * NC) + NC + B +
                                   (2 * NC * sizeof(int) + sizeof(cplx<R>) - 1) / sizeof(cplx<R>)
                             : 2 * (size_t)lpad<IL>(B * NC) + NC;
}
template <typename R, int NC>
static constexpr int ct_lines_x() {
    return rader_ip<R, NC>() ? lines_per_block_ip(NC, sizeof(cplx<R>), false)
                             : lines_per_block(NC, sizeof(cplx<R>), 64, PUFF_XY_LDS_BUDGET, false);
}
#ifndef PUFF_Y_LDS_BUDGET
#define PUFF_Y_LDS_BUDGET PUFF_XY_LDS_BUDGET
#endif
template <typename R, int NC>
static constexpr int ct_lines_y() {
    return rader_ip<R, NC>() ? lines_per_block_ip(NC, sizeof(cplx<R>), true)
                             : lines_per_block(NC, sizeof(cplx<R>), 16, PUFF_Y_LDS_BUDGET, true);
}
// Compile-time plans may give a thread several butterflies of an in-place
// pass (inplace_pass<KB>), up to Z_MAX_HELD complex values per thread
// (64 VGPRs), so their lines per workgroup follow the LDS budget alone.
template <typename R>
constexpr int z_max_held() { return sizeof(R) == 8 ? 16 : 32; }
template <typename R, int NC>
static constexpr ZPlan ct_zplan() {
    bool inplace = !CtPlan<NC>::rader;
    for (int k = 0; k < CtPlan<NC>::np; ++k) {
        const int r = CtPlan<NC>::r[k];
        if (r != 2 && r != 3 && r != 4 && r != 16) inplace = false;
    }
    const size_t elem = sizeof(cplx<R>);
    int B2 = (int)((PUFF_Z_LDS_BUDGET) / (2 * (size_t)NC * elem));
    B2 = B2 < 1 ? 1 : (B2 > 16 ? 16 : B2);
    if (!inplace) return ZPlan{B2, false};
    int Bi = (int)((PUFF_Z_LDS_BUDGET) / ((size_t)NC * elem));
    Bi = Bi > 16 ? 16 : Bi;
    for (; Bi > 1; --Bi) {
        bool fits = true;
        for (int k = 0; k < CtPlan<NC>::np; ++k) {
            const int r = CtPlan<NC>::r[k];
            if (inplace_kb(Bi, NC, r, PUFF_Z_THREADS) * r > z_max_held<R>()) fits = false;
        }
        if (fits) break;
    }
    if (Bi >= 1 && Bi >= B2) return ZPlan{Bi, true};
    return ZPlan{B2, false};
}

template <typename R, int NC, int BC, int OP>
static void launch_x_op(const DiffArgs<R>& a, int B, int sign, hipStream_t s) {
    const int64_t nlines = (int64_t)a.ny * a.nzp;
    const size_t sh = (NC ? xy_lds_elems<R, NC, false>(B) : 2 * (size_t)lpad<false>(B * a.nx) + a.nx) *
                      sizeof(cplx<R>);
    lds_opt_in(reinterpret_cast<const void*>(&k_fft_x<R, NC, BC, OP>), sh);
    const int thr = a.ax.rader ? PUFF_RADER_THREADS : 256;
    hipLaunchKernelGGL((k_fft_x<R, NC, BC, OP>), dim3((unsigned)((nlines + B - 1) / B)), dim3(thr), sh, s, a, B,
                       sign);
}

template <typename R, int NC, int BC>
static void launch_x(const DiffArgs<R>& a, int B, int op, int sign, hipStream_t s) {
    if (BC) B = BC;   // compile-time plans fix their lines per workgroup
    switch (op) {
        case XOP_FORWARD: return launch_x_op<R, NC, BC, XOP_FORWARD>(a, B, sign, s);
        case XOP_INV_SPLIT_FWD: return launch_x_op<R, NC, BC, XOP_INV_SPLIT_FWD>(a, B, sign, s);
        case XOP_INV_RECOMBINE: return launch_x_op<R, NC, BC, XOP_INV_RECOMBINE>(a, B, sign, s);
        case XOP_ADJ_FOLD_SAVE_FWD: return launch_x_op<R, NC, BC, XOP_ADJ_FOLD_SAVE_FWD>(a, B, sign, s);
        case XOP_ADJ_INV_MIX_FWD: return launch_x_op<R, NC, BC, XOP_ADJ_INV_MIX_FWD>(a, B, sign, s);
        default: return launch_x_op<R, NC, BC, XOP_ADJ_INV_SCALE>(a, B, sign, s);
    }
}

template <typename R>
static void run_x(const DiffArgs<R>& a, int op, int sign, hipStream_t s) {
    const int B = lines_per_block(a.nx, sizeof(cplx<R>), 64, PUFF_XY_LDS_BUDGET, false);
    if (a.nx == 129) return launch_x<R, 129, ct_lines_x<R, 129>()>(a, B, op, sign, s);
    if (a.nx == 257) return launch_x<R, 257, ct_lines_x<R, 257>()>(a, B, op, sign, s);
    if (a.nx == 33) return launch_x<R, 33, ct_lines_x<R, 33>()>(a, B, op, sign, s);
    launch_x<R, 0, 0>(a, B, op, sign, s);
}

// Workgroups of a persistent pass: as many as the device holds at once (the
// kernel's occupancy at `sh` bytes of LDS times the CUs), a multiple of the
// XCD count, and no more than the tiles.
static int64_t persistent_grid(const void* kernel, int threads, size_t sh, int64_t ntile, int* cache) {
    if (*cache <= 0) {
        int dev = 0, cus = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, sh);
        *cache = (per < 1 ? 1 : per) * (cus < 1 ? 1 : cus);
    }
    int64_t g = *cache;
    const int64_t cap = (ntile + NXCD - 1) / NXCD * NXCD;
    if (g > cap) g = cap;
    g = g / NXCD * NXCD;
    return g < NXCD ? NXCD : g;
}

template <typename R, int NC, int BC>
static void launch_y(const DiffArgs<R>& a, int B, int sign, hipStream_t s) {
    if (BC) B = BC;   // compile-time plans fix their lines per workgroup
    const size_t sh = (NC ? xy_lds_elems<R, NC, true>(B) : 2 * (size_t)lpad<true>(B * a.ny) + a.ny) *
                      sizeof(cplx<R>);
    lds_opt_in(reinterpret_cast<const void*>(&k_fft_y<R, NC, BC>), sh);
    const int thr = a.ay.rader ? PUFF_RADER_THREADS : 256;
    const int64_t ntile = (int64_t)((a.nx + B - 1) / B) * a.nzp;
    static int cache = 0;   // resident workgroups (one device)
    const int64_t blocks = (NC > 0 && PUFF_PERSIST_Y) ? persistent_grid(reinterpret_cast<const void*>(&k_fft_y<R, NC, BC>), thr, sh,
                                                    ntile, &cache)
                                  : xcd_grid(ntile);
    hipLaunchKernelGGL((k_fft_y<R, NC, BC>), dim3((unsigned)blocks), dim3(thr), sh, s, a, B, sign);
}

template <typename R>
static void run_y(const DiffArgs<R>& a, int sign, hipStream_t s) {
    const int B = lines_per_block(a.ny, sizeof(cplx<R>), 16, PUFF_Y_LDS_BUDGET, true);
    if (a.ny == 129) return launch_y<R, 129, ct_lines_y<R, 129>()>(a, B, sign, s);
    if (a.ny == 257) return launch_y<R, 257, ct_lines_y<R, 257>()>(a, B, sign, s);
    if (a.ny == 33) return launch_y<R, 33, ct_lines_y<R, 33>()>(a, B, sign, s);
    launch_y<R, 0, 0>(a, B, sign, s);
}

template <typename R, bool IP, int NC, int BC>
static void launch_z(const DiffArgs<R>& a, int B, int op, int sign, hipStream_t s) {
    const size_t sh = (IP ? 1 : 2) * (size_t)lpad<true>(B * a.nzp) * sizeof(cplx<R>);
    const int64_t blocks = xcd_grid((int64_t)((a.nx + B - 1) / B) * a.ny);
    lds_opt_in(reinterpret_cast<const void*>(&k_fft_z<R, IP, NC, BC>), sh);
    hipLaunchKernelGGL((k_fft_z<R, IP, NC, BC>), dim3((unsigned)blocks), dim3(PUFF_Z_THREADS), sh, s, a, B, op,
                       sign);
}

template <typename R, int NC>
static bool launch_z_ct(const DiffArgs<R>& a, int op, int sign, hipStream_t s) {
    if (a.nzp != NC) return false;
    constexpr ZPlan zp = ct_zplan<R, NC>();
    launch_z<R, zp.inplace, NC, zp.B>(a, zp.B, op, sign, s);
    return true;
}

// Four-step z2 transform of 4096-point lines (FP64 always: a whole line and
// its twiddles do not fit one workgroup's LDS; FP32 unless PUFF_Z4 is 0).
#ifndef PUFF_Z4
#define PUFF_Z4 1
#endif
#ifndef PUFF_Z4_SETS
#define PUFF_Z4_SETS 2
#endif
template <typename R>
static bool z4_line(int n) {
    return n == Z4_M * Z4_M && (sizeof(R) == 8 || PUFF_Z4);
}

template <typename R, int PASS>
static void launch_z4(const DiffArgs<R>& a, int op, int sign, cplx<R>* out, hipStream_t s) {
    constexpr int BX = 128 / (int)sizeof(cplx<R>);   // one 128-byte row per plane
    constexpr int SETS = PUFF_Z4_SETS;
    constexpr int L = BX * SETS;
    static_assert((L * Z4_M) % PUFF_Z_THREADS == 0, "whole loads per thread");
    // line buffer, twiddles, the middle pass's kz and rz
    const size_t sh = ((size_t)lpad<true>(L * Z4_M) + Z4_M) * sizeof(cplx<R>) + 2 * SETS * Z4_M * sizeof(double);
    const int nxb = (a.nx + BX - 1) / BX;
    const int64_t blocks = xcd_grid((int64_t)nxb * (Z4_M / SETS) * a.ny);
    lds_opt_in(reinterpret_cast<const void*>(&k_fft_z4<R, BX, SETS, PASS>), sh);
    hipLaunchKernelGGL((k_fft_z4<R, BX, SETS, PASS>), dim3((unsigned)blocks), dim3(PUFF_Z_THREADS), sh, s, a, op,
                       sign, out);
}

// op >= 0: forward, spectral op, inverse, in place.  op < 0: plain transform
// with `sign`, its middle pass writing to tmp (a field-sized buffer), then
// copied back.
template <typename R>
static void run_z4(const DiffArgs<R>& a, int op, int sign, cplx<R>* tmp, hipStream_t s) {
    if (op >= 0) {
        launch_z4<R, Z4_FIRST>(a, op, -1, nullptr, s);
        launch_z4<R, Z4_MID>(a, op, 0, nullptr, s);
        launch_z4<R, Z4_LAST>(a, op, +1, nullptr, s);
    } else {
        launch_z4<R, Z4_FIRST>(a, op, sign, nullptr, s);
        launch_z4<R, Z4_MID>(a, op, sign, tmp, s);
        (void)hipMemcpyAsync(a.a, tmp, sizeof(cplx<R>) * (size_t)a.nx * a.ny * a.nzp, hipMemcpyDeviceToDevice, s);
    }
}

template <typename R>
static void run_z(const DiffArgs<R>& a, int op, int sign, hipStream_t s) {
    if (z4_line<R>(a.nzp)) return run_z4(a, op, sign, a.anb, s);
    if (launch_z_ct<R, 1024>(a, op, sign, s) || launch_z_ct<R, 4096>(a, op, sign, s) ||
        launch_z_ct<R, 128>(a, op, sign, s) || launch_z_ct<R, 256>(a, op, sign, s))
        return;
    const ZPlan zp = z_plan(a.nzp, sizeof(cplx<R>), a.az.rader, a.az.npass, a.az.radix);
    if (zp.inplace) launch_z<R, true, 0, 0>(a, zp.B, op, sign, s);
    else launch_z<R, false, 0, 0>(a, zp.B, op, sign, s);
}

template <typename R>
void launch_diffraction(const DiffArgs<R>& a, hipStream_t s) {
    run_x(a, XOP_FORWARD, -1, s);
    run_y(a, -1, s);
    run_z(a, SPEC_PROPAGATE, 0, s);
    run_y(a, +1, s);
    run_x(a, XOP_INV_SPLIT_FWD, 0, s);
    run_y(a, -1, s);
    run_z(a, SPEC_ABSORB, 0, s);
    run_y(a, +1, s);
    run_x(a, XOP_INV_RECOMBINE, 0, s);
}

// Adjoint of launch_diffraction.  With W the unnormalised forward transform,
// E the multiplyexp factor (with its filter), G the real damping and M the
// real mask, the forward step is D = Q P with P = W^H E W / ffact and
// Q = (1 - M) + W^H G W M / ffact, so D^H = P^H Q^H with
// Q^H = (1 - M) + M W^H G W / ffact and P^H = W^H conj(E) W / ffact:
//   x fwd (fold alias, keep copy) | y fwd | z fwd . damping . z inv | y inv |
//   x inv . /ffact . mix with copy . x fwd | y fwd | z fwd . conj(E) . z inv |
//   y inv | x inv . /ffact
template <typename R>
void launch_diffraction_adjoint(const DiffArgs<R>& a, hipStream_t s) {
    DiffArgs<R> c = a;
    run_x(c, XOP_ADJ_FOLD_SAVE_FWD, 0, s);
    run_y(c, -1, s);
    run_z(c, SPEC_ABSORB, 0, s);
    run_y(c, +1, s);
    run_x(c, XOP_ADJ_INV_MIX_FWD, 0, s);
    run_y(c, -1, s);
    c.h = -a.h;   // conj(E): the propagator phase with -h (the filter is real)
    run_z(c, SPEC_PROPAGATE, 0, s);
    c.h = a.h;
    run_y(c, +1, s);
    run_x(c, XOP_ADJ_INV_SCALE, 0, s);
}

template <typename R>
void launch_fft3d(cplx<R>* data, cplx<R>* tmp, int nx, int ny, int nz, const FftAxis& ax, const FftAxis& ay,
                  const FftAxis& az, int sign, hipStream_t s) {
    DiffArgs<R> a{};
    a.a = data;
    a.anb = tmp;
    a.nx = nx;
    a.ny = ny;
    a.nz2 = nz;
    a.nzp = nz;
    a.ax = ax;
    a.ay = ay;
    a.az = az;
    run_x(a, XOP_FORWARD, sign, s);
    run_y(a, sign, s);
    run_z(a, -1, sign, s);
}

template void launch_diffraction<double>(const DiffArgs<double>&, hipStream_t);
template void launch_diffraction<float>(const DiffArgs<float>&, hipStream_t);
template void launch_diffraction_adjoint<double>(const DiffArgs<double>&, hipStream_t);
template void launch_diffraction_adjoint<float>(const DiffArgs<float>&, hipStream_t);
template void launch_fft3d<double>(cplx<double>*, cplx<double>*, int, int, int, const FftAxis&,
                                   const FftAxis&, const FftAxis&, int, hipStream_t);
template void launch_fft3d<float>(cplx<float>*, cplx<float>*, int, int, int, const FftAxis&, const FftAxis&,
                                  const FftAxis&, int, hipStream_t);

}  // namespace puff
