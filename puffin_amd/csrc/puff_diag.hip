// Diagnostics (power, slice bunching) and layout conversion kernels.
//
//   power    gPowerP / fPower_3D / fPower_1D, io/av_write.f90:39-202
//   bunching getSliceTwiss, io/av_write.f90:405-773 (harmonic sums 562-650,
//            magnitudes 725-729)
//   slice moments  getSliceTwiss's per-slice sums (av_write.f90:505-536)
//   current  getCurr's linear binning of the charge (av_write.f90:297-368)
#include "puff_common.hpp"
#include "puff_kernels.hpp"

namespace puff {

// Power of one z2 plane: trapezoidal integral of |A|^2 over x then y
// (m_trapz2D, av_write.f90:217-238) on a uniform axis; |A|^2 in 1-D.
// Accumulated in FP64 (the reference's m_trapz is a default-real function).
template <typename R>
__global__ __launch_bounds__(256) void k_power(const cplx<R>* a, int nx, int ny, double dx,
                                               double dy, double* out) {
    const int iz = blockIdx.x;
    const int64_t ntr = (int64_t)nx * ny;
    const cplx<R>* p = a + (int64_t)iz * ntr;
    __shared__ double red[4];
    double s = 0.0;
    if (nx == 1 && ny == 1) {
        if (threadIdx.x == 0) {
            const double r = (double)p[0].x, i = (double)p[0].y;
            out[iz] = r * r + i * i;
        }
        return;
    }
    for (int64_t t = threadIdx.x; t < ntr; t += blockDim.x) {
        const int ix = (int)(t % nx), iy = (int)(t / nx);
        const double r = (double)p[t].x, i = (double)p[t].y;
        double w = 1.0;
        if (ix == 0 || ix == nx - 1) w *= 0.5;
        if (iy == 0 || iy == ny - 1) w *= 0.5;
        s += w * (r * r + i * i);
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[iz] = (red[0] + red[1] + red[2] + red[3]) * dx * dy;
}

// Per-slice sums sum chi cos(h z2 / 2 rho) w, sum chi sin(...) w and the
// slice charge, with the linear split between slices is and is+1 on a
// temporal mesh (av_write.f90:505-530, 562-572).  acc = [br | bi | sdata].
template <typename R>
__global__ __launch_bounds__(256) void k_bunching(const double* z2, const R* chi, int64_t n,
                                                  double rho, int harmonic, int nslices,
                                                  double slice, int periodic, double* acc) {
    extern __shared__ double sh[];
    for (int t = threadIdx.x; t < 3 * nslices; t += blockDim.x) sh[t] = 0.0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double z = z2[i];
        int is = (int)ceil(z / slice);
        if (periodic) {
            if (is >= nslices) is = is - ((int)floor((double)is / (double)nslices) * nslices) + 1;
            if (is < 1) continue;
        } else {
            if (is > nslices || is < 1) continue;
        }
        const double c = (double)chi[i];
        double sn, cs;
        sincos((double)harmonic * z / (2.0 * rho), &sn, &cs);
        const double w1 = ((is * slice - z) / slice);
        atomicAdd(&sh[is - 1], c * cs * w1);
        atomicAdd(&sh[nslices + is - 1], c * sn * w1);
        atomicAdd(&sh[2 * nslices + is - 1], c);
        if (!periodic && is < nslices) {
            const double w2 = (-((is - 1) * slice - z) / slice);
            atomicAdd(&sh[is], c * cs * w2);
            atomicAdd(&sh[nslices + is], c * sn * w2);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * nslices; t += blockDim.x)
        if (sh[t] != 0.0) atomicAdd(&acc[t], sh[t]);
}

// Per-slice sums of getSliceTwiss for the slices [s0, s0 + nch) of nslices
// (av_write.f90:505-536), in LDS then one global atomic per slice and sum:
// acc[k * nslices + slice] for the NSLICE_SUMS sums, in the order
// chi, chi x, chi y, chi px, -chi py, chi Gamma, chi x^2, chi x px,
// chi px^2, chi py^2, -chi y py, chi y^2, chi Gamma^2 (py is stored as -py,
// hence the signs).  Slice index and periodic wrap as k_bunching.
template <typename R>
__global__ __launch_bounds__(256) void k_slice_moments(PartPtrs<R> y, const R* chi, int64_t n, int nslices,
                                                       double slice, int periodic, int s0, int nch,
                                                       double* acc) {
    extern __shared__ double sh[];
    for (int t = threadIdx.x; t < NSLICE_SUMS * nch; t += blockDim.x) sh[t] = 0.0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double z = y.z2[i];
        int is = (int)ceil(z / slice);
        if (periodic) {
            if (is >= nslices) is = is - ((int)floor((double)is / (double)nslices) * nslices) + 1;
            if (is < 1) continue;
        } else {
            if (is > nslices || is < 1) continue;
        }
        const int l = is - 1 - s0;
        if (l < 0 || l >= nch) continue;
        const double c = (double)chi[i], x = (double)y.x[i], yy = (double)y.y[i];
        const double px = (double)y.pr[i], py = (double)y.pi[i], g = y.g[i];
        const double v[NSLICE_SUMS] = {c, c * x, c * yy, c * px, -(c * py), c * g, c * x * x, c * x * px,
                                       c * px * px, c * py * py, -(c * yy * py), c * yy * yy, c * g * g};
#pragma unroll
        for (int k = 0; k < NSLICE_SUMS; ++k) atomicAdd(&sh[k * nch + l], v[k]);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NSLICE_SUMS * nch; t += blockDim.x)
        if (sh[t] != 0.0) atomicAdd(&acc[(t / nch) * nslices + s0 + (t % nch)], sh[t]);
}

// Charge binned linearly onto npts nodes spaced sam apart (getCurr,
// av_write.f90:297-368): nodes [p0, p0 + nch) of this launch, in LDS.  A
// macroparticle outside the nodes (the reference STOPs) adds 1 to
// acc[npts].
template <typename R>
__global__ __launch_bounds__(256) void k_current(const double* z2, const R* chi, int64_t n, int npts,
                                                 double sam, int periodic, int p0, int nch, double* acc) {
    extern __shared__ double sh[];
    for (int t = threadIdx.x; t < nch; t += blockDim.x) sh[t] = 0.0;
    __syncthreads();
    double bad = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double z = z2[i];
        const int inl = (int)ceil(z / sam);
        const int inu = inl + 1;
        int lo = inl, hi = inu;
        if (periodic) {
            if (inu >= npts) hi = inu - (int)floor((double)inu / (double)npts) * npts + 1;
            if (inl >= npts) lo = inl - (int)floor((double)inl / (double)npts) * npts + 1;
            if (inl <= 0) { bad += 1.0; continue; }
        } else {
            if (inu > npts || inl <= 0) { bad += 1.0; continue; }
        }
        const double li2 = (z - (double)(inl - 1) * sam) / sam;
        const double li1 = 1.0 - li2;
        const double c = (double)chi[i];
        if (lo - 1 - p0 >= 0 && lo - 1 - p0 < nch) atomicAdd(&sh[lo - 1 - p0], li1 * c);
        if (hi - 1 - p0 >= 0 && hi - 1 - p0 < nch) atomicAdd(&sh[hi - 1 - p0], li2 * c);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nch; t += blockDim.x)
        if (sh[t] != 0.0) atomicAdd(&acc[p0 + t], sh[t]);
    bad = wave_sum(bad);
    if (p0 == 0 && (threadIdx.x & 63) == 0 && bad > 0.0) atomicAdd(&acc[npts], bad);
}

template <typename R>
__global__ void k_pack(const double* ar, const double* ai, cplx<R>* a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = {(R)ar[i], (R)ai[i]};
}

template <typename R>
__global__ void k_unpack(const cplx<R>* a, double* ar, double* ai, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        ar[i] = (double)a[i].x;
        ai[i] = (double)a[i].y;
    }
}

template <typename R>
__global__ void k_convert(const double* s, R* d, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (R)s[i];
}

template <typename R>
__global__ void k_convert_back(const R* s, double* d, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (double)s[i];
}

__global__ void k_fill_zero(double* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

static dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1)); }

__global__ void k_store_words(uint32_t* dst, Words64 src, int n) {
    const int t = threadIdx.x;
    if (t < n) dst[t] = src.w[t];
}

__global__ __launch_bounds__(256) void k_copy_sections(CopySections c) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    for (int s = 0; s < c.n; ++s) {
        const int64_t n16 = c.bytes[s] / 16;
        const uint4* sp = reinterpret_cast<const uint4*>(c.src[s]);
        uint4* dp = reinterpret_cast<uint4*>(c.dst[s]);
        for (int64_t i = tid; i < n16; i += nth) dp[i] = sp[i];
        const int64_t rem = (c.bytes[s] - n16 * 16) / 4;
        const uint32_t* s4 = reinterpret_cast<const uint32_t*>(c.src[s] + n16 * 16);
        uint32_t* d4 = reinterpret_cast<uint32_t*>(c.dst[s] + n16 * 16);
        for (int64_t i = tid; i < rem; i += nth) d4[i] = s4[i];
    }
}

void launch_copy_sections(const CopySections& c, hipStream_t s) {
    int64_t total = 0;
    for (int k = 0; k < c.n; ++k) total += c.bytes[k] / 16;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_copy_sections, dim3((unsigned)blocks), dim3(256), 0, s, c);
}

void launch_store_words(uint32_t* dst, const Words64& src, int nwords, hipStream_t s) {
    hipLaunchKernelGGL(k_store_words, dim3(1), dim3(64), 0, s, dst, src, nwords);
}

void launch_fill_zero(double* p, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_zero, grid_for(n), dim3(256), 0, s, p, n);
}

template <typename R>
void launch_power(const cplx<R>* a, int nx, int ny, int nz2, double dx, double dy, double* out,
                  hipStream_t s) {
    hipLaunchKernelGGL((k_power<R>), dim3(nz2), dim3(256), 0, s, a, nx, ny, dx, dy, out);
}

template <typename R>
void launch_bunching(const double* z2, const R* chi, int64_t n, double rho, int harmonic,
                     int nslices, double slice, int periodic, double* acc, hipStream_t s) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_bunching<R>), dim3((unsigned)blocks), dim3(256),
                       3 * (size_t)nslices * sizeof(double), s, z2, chi, n, rho, harmonic, nslices,
                       slice, periodic, acc);
}

template <typename R>
void launch_slice_moments(PartPtrs<R> y, const R* chi, int64_t n, int nslices, double slice, int periodic,
                          double* acc, hipStream_t s) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    for (int s0 = 0; s0 < nslices; s0 += SLICE_CHUNK) {
        const int nch = nslices - s0 < SLICE_CHUNK ? nslices - s0 : SLICE_CHUNK;
        hipLaunchKernelGGL((k_slice_moments<R>), dim3((unsigned)blocks), dim3(256),
                           (size_t)NSLICE_SUMS * nch * sizeof(double), s, y, chi, n, nslices, slice, periodic,
                           s0, nch, acc);
    }
}

template <typename R>
void launch_current(const double* z2, const R* chi, int64_t n, int npts, double sam, int periodic,
                    double* acc, hipStream_t s) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    for (int p0 = 0; p0 < npts; p0 += CURRENT_CHUNK) {
        const int nch = npts - p0 < CURRENT_CHUNK ? npts - p0 : CURRENT_CHUNK;
        hipLaunchKernelGGL((k_current<R>), dim3((unsigned)blocks), dim3(256), (size_t)nch * sizeof(double), s,
                           z2, chi, n, npts, sam, periodic, p0, nch, acc);
    }
}

template <typename R>
void launch_pack_field(const double* ar, const double* ai, cplx<R>* a, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_pack<R>), grid_for(n), dim3(256), 0, s, ar, ai, a, n);
}
template <typename R>
void launch_unpack_field(const cplx<R>* a, double* ar, double* ai, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_unpack<R>), grid_for(n), dim3(256), 0, s, a, ar, ai, n);
}
template <typename R>
void launch_convert(const double* src, R* dst, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<R>), grid_for(n), dim3(256), 0, s, src, dst, n);
}
template <typename R>
void launch_convert_back(const R* src, double* dst, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert_back<R>), grid_for(n), dim3(256), 0, s, src, dst, n);
}

#define PUFF_INST(R)                                                                              \
    template void launch_power<R>(const cplx<R>*, int, int, int, double, double, double*,          \
                                  hipStream_t);                                                   \
    template void launch_bunching<R>(const double*, const R*, int64_t, double, int, int, double,  \
                                     int, double*, hipStream_t);                                  \
    template void launch_slice_moments<R>(PartPtrs<R>, const R*, int64_t, int, double, int,       \
                                          double*, hipStream_t);                                  \
    template void launch_current<R>(const double*, const R*, int64_t, int, double, int, double*,  \
                                    hipStream_t);                                                 \
    template void launch_pack_field<R>(const double*, const double*, cplx<R>*, int64_t,           \
                                       hipStream_t);                                              \
    template void launch_unpack_field<R>(const cplx<R>*, double*, double*, int64_t, hipStream_t); \
    template void launch_convert<R>(const double*, R*, int64_t, hipStream_t);                     \
    template void launch_convert_back<R>(const R*, double*, int64_t, hipStream_t);
PUFF_INST(double)
PUFF_INST(float)
#undef PUFF_INST

}  // namespace puff
