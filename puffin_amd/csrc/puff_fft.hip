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
#include "puff_common.hpp"
#include "puff_kernels.hpp"

namespace puff {

// Threads per z2 workgroup.  512 and 1024 (more loads in flight per line
// at the same LDS) measured 6 % and 30 % slower on the C1 mesh in FP64;
// 512 gained 2 % on the C4 mesh.
#ifndef PUFF_Z_THREADS
#define PUFF_Z_THREADS 256
#endif

// LDS index of element j of line b.  Interleaved layout (b fastest) keeps
// the strided y / z2 loads conflict free; line-major suits contiguous x lines.
template <bool IL>
__device__ __forceinline__ int lidx(int b, int j, int B, int N) {
    return IL ? j * B + b : b * N + j;
}

// Line staging between global memory and LDS with U loads (or stores) per
// thread issued back to back.  A plain element loop (U = 1) writes each
// loaded value to LDS before it issues the next load, so every thread has
// one load in flight and pays the memory latency once per element.  U = 8
// took C4's x and y passes from 4.69 and 3.37 ms to 4.25 and 2.66 ms; the
// z2 passes, whose elements lie NX NY apart, ran 4-5 % slower with it and
// keep U = 1.
constexpr int STAGE_UNROLL = 8;
#ifndef PUFF_Z_UNROLL
#define PUFF_Z_UNROLL 1
#endif

// dst[w] = get(w) for 0 <= w < n.
template <int U, typename R, typename F>
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
            if (w < n) dst[w] = v[u];
        }
    }
}

// put(w, src[w]) for 0 <= w < n.  The LDS reads are unconditional (element 0
// stands in past n), so that v stays in registers: with conditional reads
// the compiler kept v in scratch memory (144 bytes per lane in FP64).
template <int U, typename R, typename F>
__device__ __forceinline__ void lds_drain(const cplx<R>* src, int n, F put) {
    for (int w0 = threadIdx.x; w0 < n; w0 += U * (int)blockDim.x) {
        cplx<R> v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int w = w0 + u * (int)blockDim.x;
            v[u] = src[w < n ? w : 0];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int w = w0 + u * (int)blockDim.x;
            if (w < n) put(w, v[u]);
        }
    }
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
#undef PUFF_CT_PLAN

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
    } else {
        // direct DFT pass for an odd prime r, folded: with the pre-twiddled
        // inputs y_j = x_j W^(j k twst), u_j = y_j + y_(r-j),
        // v_j = y_j - y_(r-j) (j = 1 .. (r-1)/2) and w^(jq) = c + i s
        // (w^((r-j)q) = c - i s),
        //   X_q     = y_0 + sum u_j c + i sum v_j s
        //   X_(r-q) = y_0 + sum u_j c - i sum v_j s,
        // four FMAs per input pair for both outputs.
        const int nr = N / r;
        const int hq = (r + 1) / 2;
        // each work item forms the outputs of two values of q (qa = 2 qq
        // and qb = qa + 1, with their mirrors r - q), sharing the input
        // loads, the folding and the loop bookkeeping
        const int nqq = (hq + 1) / 2;
        for (int w = threadIdx.x; w < B * T * nqq; w += blockDim.x) {
            int b, t;
            if (IL) { b = w % B; t = w / B; } else { t = w % (T * nqq); b = w / (T * nqq); }
            const int qa = 2 * (t / T);
            const bool hasb = qa + 1 < hq;
            const int qb = hasb ? qa + 1 : 0;   // qb = 0: a valid table index, result unused
            const int i = t % T;
            const int k = i % p;
            const int dk = k * twst;
            const int dqa = qa * nr, dqb = qb * nr;
            const cplx<R> y0 = src[lidx<IL>(b, i, B, LD)];
            int e1 = dk, e2 = (r - 1) * dk, fa = dqa, fb = dqb;   // e1, e2 < N throughout
            R sra = 0, sia = 0, tra = 0, tia = 0;
            R srb = 0, sib = 0, trb = 0, tib = 0;
#pragma unroll 4
            for (int j = 1; j < hq; ++j) {
                cplx<R> xa = src[lidx<IL>(b, i + j * T, B, LD)];
                cplx<R> xb = src[lidx<IL>(b, i + (r - j) * T, B, LD)];
                if (dk) {
                    xa = cmul(xa, tw_get(tw, e1, sign));
                    xb = cmul(xb, tw_get(tw, e2, sign));
                }
                const R ux = xa.x + xb.x, uy = xa.y + xb.y;
                const R vx = xa.x - xb.x, vy = xa.y - xb.y;
                const cplx<R> oa = tw_get(tw, fa, sign);
                const cplx<R> obw = tw_get(tw, fb, sign);
                sra = fma(ux, oa.x, sra);
                sia = fma(uy, oa.x, sia);
                tra = fma(vx, oa.y, tra);
                tia = fma(vy, oa.y, tia);
                srb = fma(ux, obw.x, srb);
                sib = fma(uy, obw.x, sib);
                trb = fma(vx, obw.y, trb);
                tib = fma(vy, obw.y, tib);
                e1 += dk;
                e2 -= dk;
                fa += dqa;
                if (fa >= N) fa -= N;
                fb += dqb;
                if (fb >= N) fb -= N;
            }
            const int ob = (i / p) * p * r + k;
            dst[lidx<IL>(b, ob + qa * p, B, LD)] = {y0.x + sra - tia, y0.y + sia + tra};
            if (qa > 0) dst[lidx<IL>(b, ob + (r - qa) * p, B, LD)] = {y0.x + sra + tia, y0.y + sia - tra};
            if (hasb) {
                dst[lidx<IL>(b, ob + qb * p, B, LD)] = {y0.x + srb - tib, y0.y + sib + trb};
                dst[lidx<IL>(b, ob + (r - qb) * p, B, LD)] = {y0.x + srb + tib, y0.y + sib - trb};
            }
        }
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
        lds_pass<R, IL>(src, dst, B, LD, N, r, P, tw, sign);
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
template <typename R, bool IL, int KB = 1>
__device__ __forceinline__ void inplace_pass(cplx<R>* buf, int B, int N, int r, int p, const cplx<R>* tw,
                                             int sign) {
    // KB butterflies per thread: all their inputs are read into registers
    // before the barrier, all outputs written after it
    const int T = N / r;
    const int twst = N / (p * r);   // j k twst < N, as in lds_pass
    int b[KB], ob[KB];
    bool act[KB];
    cplx<R> v[KB][16];
#pragma unroll
    for (int it = 0; it < KB; ++it) {
        const int w = threadIdx.x + it * (int)blockDim.x;
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
                    v[it][j] = buf[lidx<IL>(b[it], i + j * T, B, N)];
                    if (j > 0 && k > 0) v[it][j] = cmul(v[it][j], tw_get(tw, j * k * twst, sign));
                }
                dft16(v[it], sign);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (j < r) {
                        v[it][j] = buf[lidx<IL>(b[it], i + j * T, B, N)];
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
                        buf[lidx<IL>(b[it], ob[it] + (q0 + 4 * q1) * p, B, N)] = v[it][4 * q0 + q1];
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q < r) buf[lidx<IL>(b[it], ob[it] + q * p, B, N)] = v[it][q];
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
        inplace_pass<R, IL>(buf, B, N, r, p, tw, sign);
        __syncthreads();
        p *= r;
    }
}

// Butterflies per thread of an in-place pass of radix r over B lines of N.
constexpr int inplace_kb(int B, int N, int r, int threads) {
    return (B * (N / r) + threads - 1) / threads;
}

template <typename R, bool IL, int NC, int BC, int K = 0, int P = 1>
__device__ __forceinline__ void ct_passes_inplace(cplx<R>* buf, const cplx<R>* tw, int sign) {
    if constexpr (K < CtPlan<NC>::np) {
        constexpr int r = CtPlan<NC>::r[K];
        constexpr int KB = inplace_kb(BC, NC, r, PUFF_Z_THREADS);
        inplace_pass<R, IL, KB>(buf, BC, NC, r, P, tw, sign);
        __syncthreads();
        ct_passes_inplace<R, IL, NC, BC, K + 1, P * r>(buf, tw, sign);
    }
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

// z2-line transform: in place (one LDS buffer) or ping-pong.
template <typename R, bool INPLACE, int NC, int BC>
__device__ __forceinline__ cplx<R>* z_fft(cplx<R>* src, cplx<R>* dst, int B, const FftAxis& ax,
                                          const cplx<R>* tw, int sign) {
    if constexpr (INPLACE) {
        if constexpr (NC > 0) ct_passes_inplace<R, true, NC, BC>(src, tw, sign);
        else lds_passes_inplace<R, true>(src, B, ax.n, ax.npass, ax.radix, tw, sign);
        return src;
    } else {
        return lds_fft<R, true, NC>(src, dst, B, ax, tw, sign);
    }
}

// ---- x lines (contiguous) --------------------------------------------------
template <typename R, int NC, int BC>
__global__ __launch_bounds__(256) void k_fft_x(DiffArgs<R> a, int Bp, int op, int sign) {
    extern __shared__ unsigned char smem_raw[];
    const int N = NC ? NC : a.nx;
    const int B = BC ? BC : Bp;
    const FftAxis& ax = a.ax;
    cplx<R>* s0 = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* s1 = s0 + B * N;
    const int64_t nlines = (int64_t)a.ny * a.nzp;
    const int64_t L0 = (int64_t)blockIdx.x * B;
    const int nb = (int)min<int64_t>(B, nlines - L0);
    const cplx<R>* twg = reinterpret_cast<const cplx<R>*>(ax.tw);
    cplx<R>* tw = s1 + B * N;   // twiddle table staged in LDS
    for (int w = threadIdx.x; w < N; w += blockDim.x) tw[w] = twg[w];
    cplx<R>* base = a.a + L0 * N;
    const int64_t ntr = (int64_t)a.nx * a.ny;
    // plane and row of the block's first line; line L0 + b is row iy0 + b
    // counted on from there (b < B)
    const int iy0 = (int)(L0 % a.ny), iz0 = (int)(L0 / a.ny);
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
            s0[w] = v;
        }
    } else {
        const int nv = nb * N;
        lds_fill<STAGE_UNROLL>(s0, B * N, [&](int w) { return w < nv ? base[w] : cplx<R>{0, 0}; });
    }
    __syncthreads();
    const int first_sign = (op == XOP_FORWARD) ? sign : (op == XOP_ADJ_FOLD_SAVE_FWD ? -1 : +1);
    cplx<R>* res = lds_fft<R, false, NC>(s0, s1, B, ax, tw, first_sign);
    if (op == XOP_ADJ_INV_MIX_FWD) {
        // adjoint of the absorption split: (1 - m) cotangent + m W^H G W cotangent / ffact
        const cplx<R>* lam = a.anb + L0 * N;
        for (int w = threadIdx.x; w < nb * N; w += blockDim.x) {
            const int b = w / N, j = w % N;
            const int lb = iy0 + b;   // (iy, iz) of line L0 + b
            const int iy = lb % a.ny, iz = iz0 + lb / a.ny;
            const double mxy = a.mxy[(int64_t)iy * a.nx + j];
            const double m = mxy + (a.mz[iz] * (1.0 - mxy));
            const cplx<R> v = res[w];
            const cplx<R> l = lam[w];
            res[w] = {(R)((1.0 - m) * (double)l.x + m * ((double)v.x / a.ffact)),
                      (R)((1.0 - m) * (double)l.y + m * ((double)v.y / a.ffact))};
        }
        __syncthreads();
        cplx<R>* other = (res == s0) ? s1 : s0;
        res = lds_fft<R, false, NC>(res, other, B, ax, tw, -1);
    } else if (op == XOP_ADJ_INV_SCALE) {
        for (int w = threadIdx.x; w < nb * N; w += blockDim.x) {
            const cplx<R> v = res[w];
            res[w] = {(R)((double)v.x / a.ffact), (R)((double)v.y / a.ffact)};
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
            cplx<R> v = res[w];
            const double vr = (double)v.x / a.ffact, vi = (double)v.y / a.ffact;
            anb[w] = {(R)((1.0 - m) * vr), (R)((1.0 - m) * vi)};
            res[w] = {(R)(m * vr), (R)(m * vi)};
        }
        __syncthreads();
        cplx<R>* other = (res == s0) ? s1 : s0;
        res = lds_fft<R, false, NC>(res, other, B, ax, tw, -1);
    } else if (op == XOP_INV_RECOMBINE) {
        // /ffact + interior (diffraction.f90:513-517)
        const cplx<R>* anb = a.anb + L0 * N;
        for (int w = threadIdx.x; w < nb * N; w += blockDim.x) {
            const cplx<R> v = res[w];
            const cplx<R> q = anb[w];
            if (sizeof(R) == 8) {
                res[w] = {(R)((double)v.x / a.ffact + (double)q.x), (R)((double)v.y / a.ffact + (double)q.y)};
            } else {
                res[w] = {(R)((double)v.x / a.ffact) + q.x, (R)((double)v.y / a.ffact) + q.y};
            }
        }
        __syncthreads();
    }
    lds_drain<STAGE_UNROLL>(res, nb * N, [&](int w, const cplx<R>& v) { base[w] = v; });
    if (op == XOP_INV_RECOMBINE && a.periodic) {
        // keep the aliased plane NZ2 equal to plane 1 (para_field.f90:3197-3244)
        for (int w = threadIdx.x; w < nb * N; w += blockDim.x) {
            const int64_t L = L0 + w / N;
            if (L < a.ny) a.a[(int64_t)(a.nz2 - 1) * ntr + L * N + (w % N)] = res[w];
        }
    }
}

// ---- y lines (stride nx), B consecutive x per block -------------------------
template <typename R, int NC, int BC>
__global__ __launch_bounds__(256) void k_fft_y(DiffArgs<R> a, int Bp, int sign) {
    extern __shared__ unsigned char smem_raw[];
    const int N = NC ? NC : a.ny;
    const int B = BC ? BC : Bp;
    const FftAxis& ay = a.ay;
    cplx<R>* s0 = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* s1 = s0 + B * N;
    const int x0 = blockIdx.x * B;
    const int iz = blockIdx.y;
    const int nb = min(B, a.nx - x0);
    const cplx<R>* twg = reinterpret_cast<const cplx<R>*>(ay.tw);
    cplx<R>* tw = s1 + B * N;   // twiddle table staged in LDS
    for (int w = threadIdx.x; w < N; w += blockDim.x) tw[w] = twg[w];
    cplx<R>* base = a.a + (int64_t)iz * a.nx * a.ny + x0;
    lds_fill<STAGE_UNROLL>(s0, B * N, [&](int w) {
        const int b = w % B, j = w / B;
        return (b < nb) ? base[(int64_t)j * a.nx + b] : cplx<R>{0, 0};
    });
    __syncthreads();
    cplx<R>* res = lds_fft<R, true, NC>(s0, s1, B, ay, tw, sign);
    lds_drain<STAGE_UNROLL>(res, B * N, [&](int w, const cplx<R>& v) {
        const int b = w % B, j = w / B;
        if (b < nb) base[(int64_t)j * a.nx + b] = v;
    });
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
    cplx<R>* s1 = INPLACE ? s0 : s0 + B * N;
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
    lds_fill<PUFF_Z_UNROLL>(s0, B * N, [&](int w) {
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
        for (int w = threadIdx.x; w < B * N; w += blockDim.x) {
            const int b = w % B, q = w / B;
            if (b >= nb) continue;
            const double kx = a.kx[x0 + b];
            const double kz = a.kz[q];
            cplx<R> v = res[w];
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
            res[w] = v;
        }
        __syncthreads();
        cplx<R>* other = (res == s0) ? s1 : s0;
        res = z_fft<R, INPLACE, NC, BC>(res, other, B, az, tw, +1);
    }
    lds_drain<PUFF_Z_UNROLL>(res, B * N, [&](int w, const cplx<R>& v) {
        const int b = w % B, j = w / B;
        if (b < nb) base[(int64_t)j * ntr + b] = v;
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
// Threads per x/y workgroup on a Rader axis.  Its radix-16 sub-passes keep
// B x 16 of them busy (112 for C4's 257-point lines), but 128 and 192
// threads measured 15 % and 7 % slower C4 events than 256.
#ifndef PUFF_RADER_THREADS
#define PUFF_RADER_THREADS 256
#endif
static constexpr int lines_per_block(int n, size_t elem, int cap, size_t budget = PUFF_XY_LDS_BUDGET) {
    const size_t tw = (size_t)n * elem;
    int b = budget > tw ? (int)((budget - tw) / (2 * (size_t)n * elem)) : 0;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
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

template <typename R, int NC>
static constexpr int ct_lines_x() { return lines_per_block(NC, sizeof(cplx<R>), 64); }
template <typename R, int NC>
static constexpr int ct_lines_y() { return lines_per_block(NC, sizeof(cplx<R>), 16); }
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

template <typename R, int NC, int BC>
static void launch_x(const DiffArgs<R>& a, int B, int op, int sign, hipStream_t s) {
    const int64_t nlines = (int64_t)a.ny * a.nzp;
    const size_t sh = (2 * (size_t)B + 1) * a.nx * sizeof(cplx<R>);
    lds_opt_in(reinterpret_cast<const void*>(&k_fft_x<R, NC, BC>), sh);
    const int thr = a.ax.rader ? PUFF_RADER_THREADS : 256;
    hipLaunchKernelGGL((k_fft_x<R, NC, BC>), dim3((unsigned)((nlines + B - 1) / B)), dim3(thr), sh, s, a, B,
                       op, sign);
}

template <typename R>
static void run_x(const DiffArgs<R>& a, int op, int sign, hipStream_t s) {
    const int B = lines_per_block(a.nx, sizeof(cplx<R>), 64);
    if (a.nx == 129) return launch_x<R, 129, ct_lines_x<R, 129>()>(a, B, op, sign, s);
    if (a.nx == 257) return launch_x<R, 257, ct_lines_x<R, 257>()>(a, B, op, sign, s);
    if (a.nx == 33) return launch_x<R, 33, ct_lines_x<R, 33>()>(a, B, op, sign, s);
    launch_x<R, 0, 0>(a, B, op, sign, s);
}

template <typename R, int NC, int BC>
static void launch_y(const DiffArgs<R>& a, int B, int sign, hipStream_t s) {
    const size_t sh = (2 * (size_t)B + 1) * a.ny * sizeof(cplx<R>);
    lds_opt_in(reinterpret_cast<const void*>(&k_fft_y<R, NC, BC>), sh);
    const int thr = a.ay.rader ? PUFF_RADER_THREADS : 256;
    hipLaunchKernelGGL((k_fft_y<R, NC, BC>), dim3((unsigned)((a.nx + B - 1) / B), (unsigned)a.nzp), dim3(thr),
                       sh, s, a, B, sign);
}

template <typename R>
static void run_y(const DiffArgs<R>& a, int sign, hipStream_t s) {
    const int B = lines_per_block(a.ny, sizeof(cplx<R>), 16);
    if (a.ny == 129) return launch_y<R, 129, ct_lines_y<R, 129>()>(a, B, sign, s);
    if (a.ny == 257) return launch_y<R, 257, ct_lines_y<R, 257>()>(a, B, sign, s);
    if (a.ny == 33) return launch_y<R, 33, ct_lines_y<R, 33>()>(a, B, sign, s);
    launch_y<R, 0, 0>(a, B, sign, s);
}

template <typename R, bool IP, int NC, int BC>
static void launch_z(const DiffArgs<R>& a, int B, int op, int sign, hipStream_t s) {
    const size_t sh = (IP ? 1 : 2) * (size_t)B * a.nzp * sizeof(cplx<R>);
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

template <typename R>
static void run_z(const DiffArgs<R>& a, int op, int sign, hipStream_t s) {
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
void launch_fft3d(cplx<R>* data, int nx, int ny, int nz, const FftAxis& ax, const FftAxis& ay,
                  const FftAxis& az, int sign, hipStream_t s) {
    DiffArgs<R> a{};
    a.a = data;
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
template void launch_fft3d<double>(cplx<double>*, int, int, int, const FftAxis&, const FftAxis&,
                                   const FftAxis&, int, hipStream_t);
template void launch_fft3d<float>(cplx<float>*, int, int, int, const FftAxis&, const FftAxis&,
                                  const FftAxis&, int, hipStream_t);

}  // namespace puff
