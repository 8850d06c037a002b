// RK4 stage kernels of the macroparticle push: trial state, p2, tri-linear
// gather of A_perp, current deposit, undulator fields and the six scaled
// equations, fused into one kernel per RK4 stage; plus the field RK4
// combination.
//
// Reference arithmetic restated (UKFELs/Puffin):
//   rk4par          interaction/puffin_mpi_RK4.f90:38-366
//   getrhs          interaction/rhs.f90:60-330
//   getP2           utilities/gamma_to_p2.f90:84-85
//   getInterps_3D   interaction/system_interpolation.f90:25-98
//   getFFelecs_3D   interaction/system_interpolation.f90:110-142
//   getSource_3D    interaction/system_interpolation.f90:150-248
//   1-D branch      interaction/system_interpolation_1D.f90:24-179
//   equations       interaction/puffin_equations.f90:17-176
//   b-fields        undulator/bfields.f90:36-881 (per-stage coefficients are
//                   computed on the host, see puff_ctx.hip)
#include "puff_common.hpp"
#include "puff_kernels.hpp"

namespace puff {

// Cell of one particle and its interpolation fractions.
template <typename R>
struct Cell {
    int ix, iy, iz0, iz1;
    R fx, fy, fz;   // x_in2, y_in2, z2_in2
};

// Node and fractions of a particle (getInterps_3D / _1D).  Returns flag bits;
// 0 means every node index is inside the mesh.
template <typename R>
__device__ __forceinline__ int locate(const MeshGeom& m, R x, R y, double z2,
                                      Cell<R>& c) {
    int flag = 0;
    const double zq = floor(z2 * m.inv_dz2);
    if (m.periodic) {
        if (!(fabs(zq) < 1.0e9)) return FLAG_Z2MESH;
        long long q = (long long)zq;
        long long r = q % m.nzp;
        if (r < 0) r += m.nzp;
        c.iz0 = (int)r;
        c.iz1 = (c.iz0 + 1 == m.nzp) ? 0 : c.iz0 + 1;
    } else {
        // temporal mesh: z2node >= NZ2 stops the reference
        // (system_interpolation.f90:70-75); z2 < 0 is refused too.
        if (!(zq >= 0.0 && zq <= (double)(m.nzp - 2))) return FLAG_Z2MESH;
        c.iz0 = (int)zq;
        c.iz1 = c.iz0 + 1;
    }
    const double locz2 = z2 - zq * m.dz2;
    c.fz = (R)(locz2 * m.inv_dz2);
    if (m.one_d) {
        c.ix = 0;
        c.iy = 0;
        c.fx = 0;
        c.fy = 0;
        return 0;
    }
    const R tx = (x + (R)m.halfx) * (R)m.inv_dx;
    const R ty = (y + (R)m.halfy) * (R)m.inv_dy;
    const R xq = floor(tx), yq = floor(ty);
    // xnode >= nspinDX .or. xnode < 1 (system_interpolation.f90:60-68)
    if (!(xq >= (R)0 && xq <= (R)(m.nx - 2))) flag |= FLAG_TRANSVERSE;
    if (!(yq >= (R)0 && yq <= (R)(m.ny - 2))) flag |= FLAG_TRANSVERSE;
    if (flag) return flag;
    c.ix = (int)xq;
    c.iy = (int)yq;
    c.fx = (x + (R)m.halfx - xq * (R)m.dx) * (R)m.inv_dx;
    c.fy = (y + (R)m.halfy - yq * (R)m.dy) * (R)m.inv_dy;
    return 0;
}

// Scaled undulator field at one particle from the stage coefficients
// (bfields.f90): see stage_scalars() on the host for the coefficient map.
template <typename R>
__device__ __forceinline__ void bfield(const StageScalars& s, R x, R y, R& bx,
                                       R& by, R& bz) {
    if (s.und_type == PUFF_UND_PLANEPOLE) {
        bx = (R)s.c[3] * y;
        by = (R)s.c[1] + (R)s.c[4] * x;
        bz = (R)s.c[2] * sinh((R)s.c[7] * y);
    } else if (s.und_type == PUFF_UND_CURVED) {
        const R kx = (R)s.c[5], ky = (R)s.c[6];
        bx = (R)s.c[0] * sinh(kx * x) * sinh(ky * y) + (R)s.c[3] * y;
        by = (R)s.c[1] * cosh(kx * x) * cosh(ky * y) + (R)s.c[4] * x;
        bz = (R)s.c[2] * cosh(kx * x) * sinh(ky * y);
    } else {
        bx = (R)s.c[0] + (R)s.c[3] * y;
        by = (R)s.c[1] + (R)s.c[4] * x;
        bz = (R)s.c[5] * x + (R)s.c[6] * y;
    }
}

template <typename R>
__device__ __forceinline__ void atomic_add(R* p, R v) {
    unsafeAtomicAdd(p, v);
}

// One RK4 stage.  STAGE = 1..4 (rk4par's four derivs calls).
template <typename R, int STAGE>
__global__ __launch_bounds__(256) void k_stage(StageArgs<R> a) {
    if (*a.flags) return;   // an earlier kernel of this step raised a flag
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < a.n;
    const MeshGeom& m = a.mesh;
    const Physics& ph = a.phys;

    // ---- trial state y_s = y0 + c k_{s-1} (puffin_mpi_RK4.f90:190-195 etc.)
    R x = 0, y = 0, pr = 0, pi = 0, chi = 0;
    double z2 = 0.0, g = 1.0;
    R x0 = 0, y0 = 0, pr0 = 0, pi0 = 0;
    double z20 = 0.0, g0 = 1.0;
    if (in) {
        x0 = a.y0.x[i];
        y0 = a.y0.y[i];
        z20 = a.y0.z2[i];
        pr0 = a.y0.pr[i];
        pi0 = a.y0.pi[i];
        g0 = a.y0.g[i];
        chi = a.chi[i];
        if (STAGE == 1) {
            x = x0; y = y0; z2 = z20; pr = pr0; pi = pi0; g = g0;
        } else {
            const R c = (R)a.cin;
            x = fma(c, a.kin.x[i], x0);
            y = fma(c, a.kin.y[i], y0);
            z2 = fma(a.cin, a.kin.z2[i], z20);
            pr = fma(c, a.kin.pr[i], pr0);
            pi = fma(c, a.kin.pi[i], pi0);
            g = fma(a.cin, a.kin.g[i], g0);
        }
    }

    // ---- p2 in FP64 (gamma_to_p2.f90:84-85)
    const double prd = (double)pr, pid = (double)pi;
    const double p2 = in ? ((1.0 / sqrt(1.0 - 1.0 / (ph.gamma_r * ph.gamma_r * g * g) *
                                               (1.0 + ph.aw * ph.aw * (prd * prd + pid * pid)))) -
                            1.0) / ph.eta
                         : 0.0;
    const double invg = 1.0 / g;
    const double fac = (1.0 + ph.eta * p2) * invg;   // (1 + eta p2) / Gamma

    // ---- node and weights
    Cell<R> c;
    int flag = in ? locate<R>(m, x, y, z2, c) : 0;
    const bool ok = in && flag == 0;
    if (flag) atomicOr(a.flags, flag);

    R fr = 0, fi = 0;
    if (ok) {
        // active box (nodes touched are cell .. cell+1)
        const cplx<R>* A0 = a.A0;
        const cplx<R>* dA = a.dAin;
        const R cA = (R)a.cin;
        const R sr = (R)((double)chi * ph.inv_dV3 * fac * prd);
        const R si = (R)((double)chi * ph.inv_dV3 * fac * pid);
        if (m.one_d) {
            // lis_GR(1) = 1 - locz2/dz2, lis_GR(2) = 1 - lis_GR(1)
            const R w[2] = {(R)1 - c.fz, (R)1 - ((R)1 - c.fz)};
            const int nd[2] = {c.iz0, c.iz1};
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                cplx<R> v = A0[nd[k]];
                if (STAGE > 1) {
                    const cplx<R> d = dA[nd[k]];
                    v.x = fma(cA, d.x, v.x);
                    v.y = fma(cA, d.y, v.y);
                }
                fr = fma(w[k], v.x, fr);
                fi = fma(w[k], v.y, fi);
            }
            if (ph.field_evolve) {
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    atomic_add(&a.dAout[nd[k]].x, w[k] * sr);
                    atomic_add(&a.dAout[nd[k]].y, w[k] * si);
                }
            }
        } else {
            const R x1 = (R)1 - c.fx, x2 = c.fx;
            const R y1 = (R)1 - c.fy, y2 = c.fy;
            const R z1 = (R)1 - c.fz, z2w = c.fz;
            const R w[8] = {x1 * y1 * z1, x2 * y1 * z1, x1 * y2 * z1, x2 * y2 * z1,
                            x1 * y1 * z2w, x2 * y1 * z2w, x1 * y2 * z2w, x2 * y2 * z2w};
            const int64_t ntr = (int64_t)m.nx * m.ny;
            const int64_t b0 = (int64_t)c.iz0 * ntr + (int64_t)c.iy * m.nx + c.ix;
            const int64_t b1 = (int64_t)c.iz1 * ntr + (int64_t)c.iy * m.nx + c.ix;
            const int64_t nd[8] = {b0, b0 + 1, b0 + m.nx, b0 + m.nx + 1,
                                   b1, b1 + 1, b1 + m.nx, b1 + m.nx + 1};
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                cplx<R> v = A0[nd[k]];
                if (STAGE > 1) {
                    const cplx<R> d = dA[nd[k]];
                    v.x = fma(cA, d.x, v.x);
                    v.y = fma(cA, d.y, v.y);
                }
                fr = fma(w[k], v.x, fr);
                fi = fma(w[k], v.y, fi);
            }
            if (ph.field_evolve) {
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    atomic_add(&a.dAout[nd[k]].x, w[k] * sr);
                    atomic_add(&a.dAout[nd[k]].y, w[k] * si);
                }
            }
        }
    }
    // active box of this step (for the field combination)
    {
        int lx = ok ? c.ix : 0x7fffffff, hx = ok ? c.ix : -1;
        int ly = ok ? c.iy : 0x7fffffff, hy = ok ? c.iy : -1;
        int lz = ok ? c.iz0 : 0x7fffffff, hz = ok ? c.iz0 : -1;
        lx = wave_min(lx); hx = wave_max(hx);
        ly = wave_min(ly); hy = wave_max(hy);
        lz = wave_min(lz); hz = wave_max(hz);
        if ((threadIdx.x & 63) == 0 && hx >= 0) {
            atomicMin(&a.box->lo[0], lx); atomicMax(&a.box->hi[0], hx);
            atomicMin(&a.box->lo[1], ly); atomicMax(&a.box->hi[1], hy);
            atomicMin(&a.box->lo[2], lz); atomicMax(&a.box->hi[2], hz);
        }
    }
    if (!in) return;
    if (!ph.coupling) { fr = 0; fi = 0; }   // rhs.f90:221-224

    // ---- equations (puffin_equations.f90:17-176)
    R kx_ = 0, ky_ = 0, kpr = 0, kpi = 0;
    double kz2 = 0.0, kg = 0.0;
    if (ph.electrons_evolve) {
        R bx, by, bz;
        bfield<R>(a.s, x, y, bx, by, bz);
        const R facr = (R)fac;
        const R n2col = (R)a.s.n2col;
        const R ep2k = (R)(ph.eta * p2 / (ph.kappa * ph.kappa));
        kz2 = p2;
        kx_ = (R)ph.c_dx * facr * pr;
        ky_ = -(R)ph.c_dx * facr * pi;
        kpr = (R)ph.inv2rho * (n2col * by - ep2k * fr) + (R)ph.kappa * pi * facr * n2col * bz;
        kpi = (R)ph.inv2rho * (n2col * bx - ep2k * fi) - (R)ph.kappa * pr * facr * n2col * bz;
        kg = -ph.rho * fac * 2.0 * ((double)pr * (double)fr + (double)pi * (double)fi);
    }

    if (STAGE < 4) {
        a.kout.x[i] = kx_;
        a.kout.y[i] = ky_;
        a.kout.z2[i] = kz2;
        a.kout.pr[i] = kpr;
        a.kout.pi[i] = kpi;
        a.kout.g[i] = kg;
        // next stage's trial state must lie inside the mesh, else the next
        // derivs call would raise the flag (checked here so that no later
        // kernel writes state when it does)
        const R cn = (R)a.cnext;
        Cell<R> cc;
        const int nf = locate<R>(m, fma(cn, kx_, x0), fma(cn, ky_, y0), fma(a.cnext, kz2, z20), cc);
        if (nf) atomicOr(a.flags, nf);
    } else {
        // y += h/6 (k1 + k4 + 2 (k2 + k3))  (puffin_mpi_RK4.f90:274-279, 307-312)
        const R h6 = (R)a.h6;
        a.y0.x[i] = x0 + h6 * (a.k1.x[i] + kx_ + (R)2 * (a.k2.x[i] + a.kin.x[i]));
        a.y0.y[i] = y0 + h6 * (a.k1.y[i] + ky_ + (R)2 * (a.k2.y[i] + a.kin.y[i]));
        a.y0.z2[i] = z20 + a.h6 * (a.k1.z2[i] + kz2 + 2.0 * (a.k2.z2[i] + a.kin.z2[i]));
        a.y0.pr[i] = pr0 + h6 * (a.k1.pr[i] + kpr + (R)2 * (a.k2.pr[i] + a.kin.pr[i]));
        a.y0.pi[i] = pi0 + h6 * (a.k1.pi[i] + kpi + (R)2 * (a.k2.pi[i] + a.kin.pi[i]));
        a.y0.g[i] = g0 + a.h6 * (a.k1.g[i] + kg + 2.0 * (a.k2.g[i] + a.kin.g[i]));
    }
}

// Start of a step: clear flags and the active box.
__global__ void k_step_begin(int* flags, Box* box) {
    if (threadIdx.x == 0) {
        *flags = 0;
        for (int d = 0; d < 3; ++d) {
            box->lo[d] = 0x7fffffff;
            box->hi[d] = -1;
        }
    }
}

// A += h/6 (dA1 + dA4 + 2 (dA2 + dA3)) over the active box
// (puffin_mpi_RK4.f90:281-282, 314-315), then clear the four deposits there.
// The periodic mesh keeps its aliased last plane equal to plane 0.
template <typename R>
__global__ __launch_bounds__(256) void k_field_combine(CombineArgs<R> a) {
    const Box b = *a.box;
    if (b.hi[0] < 0) return;   // no particle deposited
    const int x0 = b.lo[0], x1 = min(b.hi[0] + 1, a.nx - 1);
    const int y0 = a.one_d ? 0 : b.lo[1], y1 = a.one_d ? 0 : min(b.hi[1] + 1, a.ny - 1);
    int z0, z1;
    if (a.periodic) {
        z0 = 0;
        z1 = a.nzp - 1;
    } else {
        z0 = b.lo[2];
        z1 = min(b.hi[2] + 1, a.nzp - 1);
    }
    const int lx = x1 - x0 + 1, ly = y1 - y0 + 1, lz = z1 - z0 + 1;
    const int64_t vol = (int64_t)lx * ly * lz;
    const bool update = (*a.flags == 0);
    const R h6 = (R)a.h6;
    const int64_t ntr = (int64_t)a.nx * a.ny;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < vol;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int ix = x0 + (int)(t % lx);
        const int64_t r = t / lx;
        const int iy = y0 + (int)(r % ly);
        const int iz = z0 + (int)(r / ly);
        const int64_t j = (int64_t)iz * ntr + (int64_t)iy * a.nx + ix;
        const cplx<R> d0 = a.d[0][j], d1 = a.d[1][j], d2 = a.d[2][j], d3 = a.d[3][j];
        if (update) {
            cplx<R> v = a.A[j];
            const R tx = d1.x + d2.x, ty = d1.y + d2.y;
            v.x = v.x + h6 * (d0.x + d3.x + (R)2 * tx);
            v.y = v.y + h6 * (d0.y + d3.y + (R)2 * ty);
            a.A[j] = v;
            if (a.periodic && iz == 0) a.A[(int64_t)(a.nz2 - 1) * ntr + (int64_t)iy * a.nx + ix] = v;
        }
        const cplx<R> zero = {(R)0, (R)0};
        a.d[0][j] = zero;
        a.d[1][j] = zero;
        a.d[2][j] = zero;
        a.d[3][j] = zero;
    }
}

// Periodic z2 wrap (para_field.f90:212-215).
__global__ void k_wrap_z2(double* z2, int64_t n, double lenz2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z2[i] = z2[i] - (floor(z2[i] / lenz2) * lenz2);
}

// ---- host launchers --------------------------------------------------------

template <typename R>
void launch_stage(int stage, const StageArgs<R>& a, hipStream_t s) {
    const int threads = 256;
    const int64_t blocks = (a.n + threads - 1) / threads;
    dim3 grid((unsigned)(blocks > 0 ? blocks : 1));
    switch (stage) {
        case 1: hipLaunchKernelGGL((k_stage<R, 1>), grid, dim3(threads), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_stage<R, 2>), grid, dim3(threads), 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_stage<R, 3>), grid, dim3(threads), 0, s, a); break;
        default: hipLaunchKernelGGL((k_stage<R, 4>), grid, dim3(threads), 0, s, a); break;
    }
}

void launch_step_begin(int* flags, Box* box, hipStream_t s) {
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, s, flags, box);
}

template <typename R>
void launch_combine(const CombineArgs<R>& a, hipStream_t s) {
    hipLaunchKernelGGL((k_field_combine<R>), dim3(2048), dim3(256), 0, s, a);
}

void launch_wrap(double* z2, int64_t n, double lenz2, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_wrap_z2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, z2, n, lenz2);
}

template void launch_stage<double>(int, const StageArgs<double>&, hipStream_t);
template void launch_stage<float>(int, const StageArgs<float>&, hipStream_t);
template void launch_combine<double>(const CombineArgs<double>&, hipStream_t);
template void launch_combine<float>(const CombineArgs<float>&, hipStream_t);

}  // namespace puff
