// Lattice elements between undulator modules, applied to the device-resident
// macroparticles (any slot order: every element acts per particle, and the
// transverse corrector's means are order-free sums).
//
// Reference arithmetic restated (UKFELs/Puffin, lib/acc_lattice.f90):
//   driftSection  478-518  (the diffraction over the drift is the caller's)
//   disperse      420-463  (chicane; likewise its diffraction)
//   Quad          531-561
//   bModulation   575-585
//   correctTrans  601-623  (arr_mean_para_weighted, base_functions.f90:552-573)
//   matchIn       743-828, matchOut 638-725, with pxOffset / pyOffset and
//   xOffSet_Array / yOffSet_Array (undulator/init_conds.f90:65-86, 129-152,
//   188-223)
#include "puff_common.hpp"
#include "puff_kernels.hpp"

namespace puff {

static dim3 lat_grid(int64_t n) { return dim3((unsigned)(n > 0 ? (n + 255) / 256 : 1)); }

// getP2 (gamma_to_p2.f90:84-85), as the reference writes it.
__device__ __forceinline__ double lat_p2(const Physics& ph, double g, double pr, double pi) {
    return ((1.0 / sqrt(1.0 - 1.0 / (ph.gamma_r * ph.gamma_r * g * g) *
                                  (1.0 + ph.aw * ph.aw * (pr * pr + pi * pi)))) -
            1.0) / ph.eta;
}

// Free drift over dz in zbar: z2 += dz p2, and (3-D) x, y by their
// transverse velocities 2 rho kappa / sqrt(eta) (1 + eta p2) / Gamma p.
template <typename R>
__global__ __launch_bounds__(256) void k_drift(PartPtrs<R> y, int64_t n, Physics ph, int one_d,
                                               double dz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double pr = (double)y.pr[i], pi = (double)y.pi[i], g = y.g[i];
    const double p2 = lat_p2(ph, g, pr, pi);
    y.z2[i] = y.z2[i] + dz * p2;
    if (!one_d) {
        const double v = 2.0 * ph.rho * ph.kappa / sqrt(ph.eta) * (1.0 + ph.eta * p2) / g;
        y.x[i] = (R)((double)y.x[i] + (v * pr) * dz);
        y.y[i] = (R)((double)y.y[i] - (v * pi) * dz);
    }
}

// Chicane: z2 -= 2 D (Gamma - 1) - slip, then the periodic wrap.
template <typename R>
__global__ __launch_bounds__(256) void k_chicane(PartPtrs<R> y, int64_t n, double disp, double slip,
                                                 double lenz2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double z2 = y.z2[i] - 2.0 * disp * (y.g[i] - 1.0) + slip;
    if (lenz2 > 0.0) z2 = z2 - (floor(z2 / lenz2) * lenz2);
    y.z2[i] = z2;
}

// Thin quadrupole: a kick of p proportional to the offset.
template <typename R>
__global__ __launch_bounds__(256) void k_quad(PartPtrs<R> y, int64_t n, Physics ph, double qfx,
                                              double qfy) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double c = sqrt(ph.eta) / (2.0 * ph.rho * ph.kappa);
    y.pr[i] = (R)((double)y.pr[i] + c * (double)y.x[i] / qfx);
    y.pi[i] = (R)((double)y.pi[i] - c * (double)y.y[i] / qfy);
}

// Energy modulation: Gamma += mag cos(k z2).
template <typename R>
__global__ __launch_bounds__(256) void k_modulation(PartPtrs<R> y, int64_t n, double wavenum, double mag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y.g[i] = y.g[i] + (mag * cos(wavenum * y.z2[i]));
}

// Weighted sums for the transverse corrector: sums[0..4] += (chi, chi px,
// chi py, chi x, chi y), FP64, one atomic per wavefront and sum.
template <typename R>
__global__ __launch_bounds__(256) void k_weighted_sums(PartPtrs<R> y, const R* chi, int64_t n,
                                                       double* sums) {
    double s[5] = {0, 0, 0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double w = (double)chi[i];
        s[0] += w;
        s[1] += w * (double)y.pr[i];
        s[2] += w * (double)y.pi[i];
        s[3] += w * (double)y.x[i];
        s[4] += w * (double)y.y[i];
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double t = wave_sum(s[k]);
        if ((threadIdx.x & 63) == 0) atomicAdd(&sums[k], t);
    }
}

// Subtract the weighted means.
template <typename R>
__global__ __launch_bounds__(256) void k_subtract_means(PartPtrs<R> y, int64_t n, const double* sums) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ws = sums[0];
    y.pr[i] = (R)((double)y.pr[i] - sums[1] / ws);
    y.pi[i] = (R)((double)y.pi[i] - sums[2] / ws);
    y.x[i] = (R)((double)y.x[i] - sums[3] / ws);
    y.y[i] = (R)((double)y.y[i] - sums[4] / ws);
}

__global__ void k_zero_doubles(double* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0;
}

// Undulator entry (sign +1, matchIn) or exit (sign -1, matchOut) offsets at
// module-local zbar sz with the strength n2col: p offsets
//   px0 = pxOffset - (curved) kx^2 x^2/2 + ky^2 y^2/2 ... as below,
//   py0 = -pyOffset (+ curved term),
// and the x, y offsets of xOffSet_Array / yOffSet_Array evaluated with
// (px0, -py0) on entry and (px0, py0) on exit.
template <typename R>
__global__ __launch_bounds__(256) void k_match(PartPtrs<R> y, int64_t n, MatchArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = (double)y.x[i], yy = (double)y.y[i];
    const double c = cos(a.sz / (2.0 * a.rho)), s = sin(a.sz / (2.0 * a.rho));
    const double pxo = -a.uy * a.n2col * c;   // pxOffset
    const double pyo = -a.ux * a.n2col * s;   // pyOffset
    double spx0, spy0;
    if (a.und_type == PUFF_UND_CURVED) {
        spx0 = pxo - 0.5 * a.kx * a.kx * x * x - 0.5 * a.ky * a.ky * yy * yy;
        spy0 = -1.0 * (pyo - a.kx * a.kx * x * yy);
    } else if (a.und_type == PUFF_UND_PLANEPOLE) {
        spx0 = pxo - 0.5 * (a.eta / (4.0 * a.rho * a.rho)) * x * x;
        spy0 = -1.0 * pyo;
    } else {
        spx0 = pxo;
        spy0 = -1.0 * pyo;
    }
    const double pyarg = (a.sign > 0) ? -spy0 : spy0;
    const double gj = a.gamma_r * y.g[i];
    const double root_x = sqrt(gj * gj - (1.0 + a.ncx * (spx0 * spx0 + pyarg * pyarg)));
    const double root_y = sqrt(gj * gj - (1.0 + a.ncy * (spx0 * spx0 + pyarg * pyarg)));
    const double xo = -a.uy * 4.0 * a.kappa * pow(a.rho, 2.0) / sqrt(a.eta) * (a.gamma_r / root_x) *
                      a.n2col * s;
    const double yo = a.ux * 4.0 * a.kappa * pow(a.rho, 2.0) / sqrt(a.eta) * (a.gamma_r / root_y) *
                      a.n2col * c;
    const double sg = (a.sign > 0) ? 1.0 : -1.0;
    y.x[i] = (R)(x + sg * xo);
    y.y[i] = (R)(yy + sg * yo);
    y.pr[i] = (R)((double)y.pr[i] + sg * spx0);
    y.pi[i] = (R)((double)y.pi[i] + sg * spy0);
}

template <typename R>
void launch_lattice(const LatticeOp& op, PartPtrs<R> y, const R* chi, int64_t n, const Physics& ph,
                    const MeshGeom& m, double* sums, hipStream_t s) {
    switch (op.kind) {
        case LAT_DRIFT:
            hipLaunchKernelGGL((k_drift<R>), lat_grid(n), dim3(256), 0, s, y, n, ph, m.one_d, op.a);
            break;
        case LAT_CHICANE:
            hipLaunchKernelGGL((k_chicane<R>), lat_grid(n), dim3(256), 0, s, y, n, op.a, op.b, op.lenz2);
            break;
        case LAT_QUAD:
            if (!m.one_d) hipLaunchKernelGGL((k_quad<R>), lat_grid(n), dim3(256), 0, s, y, n, ph, op.a, op.b);
            break;
        case LAT_MODULATION:
            hipLaunchKernelGGL((k_modulation<R>), lat_grid(n), dim3(256), 0, s, y, n, op.a, op.b);
            break;
        case LAT_CORRECT_TRANS: {
            hipLaunchKernelGGL(k_zero_doubles, dim3(1), dim3(64), 0, s, sums, 5);
            int64_t blocks = (n + 255) / 256;
            if (blocks > 2048) blocks = 2048;
            if (blocks < 1) blocks = 1;
            hipLaunchKernelGGL((k_weighted_sums<R>), dim3((unsigned)blocks), dim3(256), 0, s, y, chi, n, sums);
            hipLaunchKernelGGL((k_subtract_means<R>), lat_grid(n), dim3(256), 0, s, y, n, sums);
            break;
        }
        default:
            hipLaunchKernelGGL((k_match<R>), lat_grid(n), dim3(256), 0, s, y, n, op.match);
            break;
    }
}

template void launch_lattice<double>(const LatticeOp&, PartPtrs<double>, const double*, int64_t,
                                     const Physics&, const MeshGeom&, double*, hipStream_t);
template void launch_lattice<float>(const LatticeOp&, PartPtrs<float>, const float*, int64_t,
                                    const Physics&, const MeshGeom&, double*, hipStream_t);

}  // namespace puff
