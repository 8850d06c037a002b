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
__global__ __launch_bounds__(STAGE_THREADS) void k_stage(StageArgs<R> a) {
    // an earlier kernel of this step raised a flag: skip (block-uniform, the
    // box reduction below uses a barrier)
    __shared__ int s_skip;
    if (threadIdx.x == 0) s_skip = *a.flags;
    __syncthreads();
    if (s_skip) return;
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
    Cell<R> c{};
    int flag = in ? locate<R>(m, x, y, z2, c) : 0;
    const bool ok = in && flag == 0;
    if (flag) atomicOr(a.flags, flag);

    // ---- bounding box of this workgroup's cells (wavefront, then workgroup
    // reduction); max values are reduced as min of their negatives
    __shared__ int s_red[6][STAGE_THREADS / 64];
    __shared__ int s_bb[6];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    {
        int v[6] = {ok ? c.ix : 0x7fffffff, ok ? -c.ix : 1,
                    ok ? c.iy : 0x7fffffff, ok ? -c.iy : 1,
                    ok ? c.iz0 : 0x7fffffff, ok ? -c.iz0 : 1};
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            v[d] = wave_min(v[d]);
            if (lane == 0) s_red[d][wv] = v[d];
        }
        __syncthreads();
        if (threadIdx.x < 6) {
            const int d = threadIdx.x;
            int r = s_red[d][0];
            for (int k = 1; k < (int)(blockDim.x >> 6); ++k) r = min(r, s_red[d][k]);
            s_bb[d] = (d & 1) ? -r : r;
            // active box of the step: one atomic per workgroup into one of
            // BOX_COPIES spread copies (same-address atomics would serialise)
            const bool empty = (d & 1) ? (r == 1) : (r == 0x7fffffff);
            if (!empty) {
                Box* b = a.box + (blockIdx.x % BOX_COPIES);
                if (d & 1) atomicMax(&b->hi[d >> 1], -r);
                else atomicMin(&b->lo[d >> 1], r);
            }
        }
        __syncthreads();
    }
    const bool any_ok = s_bb[1] >= 0;
    const int bx0 = s_bb[0], by0 = s_bb[2], bz0 = s_bb[4];
    const int tx = any_ok ? (m.one_d ? 1 : s_bb[1] - bx0 + 2) : 0;
    const int ty = any_ok ? (m.one_d ? 1 : s_bb[3] - by0 + 2) : 0;
    const int tz = any_ok ? s_bb[5] - bz0 + 2 : 0;
    // tile-private deposit when the workgroup's node box fits in LDS and does
    // not wrap around a periodic mesh
    const bool use_tile = ph.field_evolve && any_ok && (int64_t)tx * ty * tz <= TILE_NODES &&
                          !(m.periodic && s_bb[5] + 1 >= m.nzp);

    // ---- nodes and tri-linear weights (zero weights on lanes without a cell)
    R w[8];
    int nd[8];
    int lt[8];    // tile index of each corner
    int nnb, key;
    if (m.one_d) {
        // lis_GR(1) = 1 - locz2/dz2, lis_GR(2) = 1 - lis_GR(1)
        nnb = 2;
        w[0] = ok ? (R)1 - c.fz : (R)0;
        w[1] = ok ? (R)1 - ((R)1 - c.fz) : (R)0;
        nd[0] = ok ? c.iz0 : 0;
        nd[1] = ok ? c.iz1 : 0;
        lt[0] = ok ? c.iz0 - bz0 : 0;
        lt[1] = lt[0] + 1;
        key = ok ? c.iz0 : -1 - lane;
    } else {
        nnb = 8;
        const R x1 = (R)1 - c.fx, x2 = c.fx;
        const R y1 = (R)1 - c.fy, y2 = c.fy;
        const R z1 = (R)1 - c.fz, z2w = c.fz;
        w[0] = x1 * y1 * z1;  w[1] = x2 * y1 * z1;  w[2] = x1 * y2 * z1;  w[3] = x2 * y2 * z1;
        w[4] = x1 * y1 * z2w; w[5] = x2 * y1 * z2w; w[6] = x1 * y2 * z2w; w[7] = x2 * y2 * z2w;
        const int ntr = m.nx * m.ny;
        const int b0 = ok ? c.iz0 * ntr + c.iy * m.nx + c.ix : 0;
        const int b1 = ok ? c.iz1 * ntr + c.iy * m.nx + c.ix : 0;
        nd[0] = b0; nd[1] = b0 + 1; nd[2] = b0 + m.nx; nd[3] = b0 + m.nx + 1;
        nd[4] = b1; nd[5] = b1 + 1; nd[6] = b1 + m.nx; nd[7] = b1 + m.nx + 1;
        const int t0 = ok ? ((c.iz0 - bz0) * ty + (c.iy - by0)) * tx + (c.ix - bx0) : 0;
        const int txy = tx * ty;
        lt[0] = t0; lt[1] = t0 + 1; lt[2] = t0 + tx; lt[3] = t0 + tx + 1;
        lt[4] = t0 + txy; lt[5] = t0 + txy + 1; lt[6] = t0 + txy + tx; lt[7] = t0 + txy + tx + 1;
        key = ok ? b0 : -1 - lane;
        if (!ok) {
#pragma unroll
            for (int k = 0; k < 8; ++k) w[k] = (R)0;
        }
    }

    // ---- gather F = sum_k w_k (A0 + c dA_{s-1})[node_k]  (getFFelecs)
    R fr = 0, fi = 0;
    if (ok) {
        const R cA = (R)a.cin;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k < nnb) {
                cplx<R> v = a.A0[nd[k]];
                if (STAGE > 1) {
                    const cplx<R> d = a.dAin[nd[k]];
                    v.x = fma(cA, d.x, v.x);
                    v.y = fma(cA, d.y, v.y);
                }
                fr = fma(w[k], v.x, fr);
                fi = fma(w[k], v.y, fi);
            }
        }
    }

    // ---- deposit w_k s into dA_s (getSource)
    if (ph.field_evolve) {
        const R sr = ok ? (R)((double)chi * ph.inv_dV3 * fac * prd) : (R)0;
        const R si = ok ? (R)((double)chi * ph.inv_dV3 * fac * pid) : (R)0;
        if (use_tile) {
            // FP64 accumulation in an LDS tile, then one global atomic per
            // touched node and component.  Each lane visits the corners in a
            // rotated order so that lanes of one cell do not hit the same LDS
            // address in the same instruction.
            __shared__ double s_tile[2 * TILE_NODES];
            const int nt = tx * ty * tz;
            for (int t = threadIdx.x; t < 2 * nt; t += blockDim.x) s_tile[t] = 0.0;
            __syncthreads();
            if (ok) {
                const int txy = tx * ty;
                const int t0 = lt[0];
                const double dsr = (double)sr, dsi = (double)si;
                const double wx[2] = {(double)((R)1 - c.fx), (double)c.fx};
                const double wy[2] = {(double)((R)1 - c.fy), (double)c.fy};
                const double wz[2] = {(double)((R)1 - c.fz), (double)c.fz};
                for (int q = 0; q < nnb; ++q) {
                    const int k = (q + lane) & (nnb - 1);
                    double wk;
                    int tk;
                    if (m.one_d) {
                        wk = (double)w[0];
                        if (k) wk = (double)w[1];
                        tk = t0 + k;
                    } else {
                        wk = (double)((R)wx[k & 1] * (R)wy[(k >> 1) & 1] * (R)wz[k >> 2]);
                        tk = t0 + (k & 1) + ((k & 2) ? tx : 0) + ((k & 4) ? txy : 0);
                    }
                    atomicAdd(&s_tile[2 * tk], wk * dsr);
                    atomicAdd(&s_tile[2 * tk + 1], wk * dsi);
                }
            }
            __syncthreads();
            const int ntr = m.nx * m.ny;
            for (int t = threadIdx.x; t < nt; t += blockDim.x) {
                const double tr = s_tile[2 * t], ti = s_tile[2 * t + 1];
                if (tr != 0.0 || ti != 0.0) {
                    const int lx = t % tx, r = t / tx;
                    const int ly = r % ty, lz = r / ty;
                    const int node = (bz0 + lz) * ntr + (by0 + ly) * m.nx + (bx0 + lx);
                    atomic_add(&a.dAout[node].x, (R)tr);
                    atomic_add(&a.dAout[node].y, (R)ti);
                }
            }
        } else {
            // Lanes of a wavefront that share a cell are summed first (runs of
            // equal keys), then one global atomic per run and corner.
            const int prev = __shfl_up(key, 1, 64);
            const bool head = (lane == 0) || (prev != key);
            const unsigned long long heads = __ballot(head);
            const unsigned long long after = (lane == 63) ? 0ull : (heads >> (lane + 1));
            const int seg_end = after ? lane + __builtin_ctzll(after) : 63;
            const int maxlen = wave_max(head ? seg_end - lane + 1 : 1);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (k < nnb) {
                    R vr = w[k] * sr, vi = w[k] * si;
                    for (int o = 1; o < maxlen; o <<= 1) {
                        const R tr = __shfl_down(vr, o, 64);
                        const R ti = __shfl_down(vi, o, 64);
                        if (lane + o <= seg_end) {
                            vr += tr;
                            vi += ti;
                        }
                    }
                    if (ok && head) {
                        atomic_add(&a.dAout[nd[k]].x, vr);
                        atomic_add(&a.dAout[nd[k]].y, vi);
                    }
                }
            }
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
    if (threadIdx.x == 0) *flags = 0;
    for (int k = threadIdx.x; k < BOX_COPIES; k += blockDim.x) {
        for (int d = 0; d < 3; ++d) {
            box[k].lo[d] = 0x7fffffff;
            box[k].hi[d] = -1;
        }
    }
}

// Union of the spread box copies.
__device__ __forceinline__ Box merged_box(const Box* box) {
    Box b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = 0x7fffffff;
        b.hi[d] = -1;
    }
    for (int k = 0; k < BOX_COPIES; ++k)
        for (int d = 0; d < 3; ++d) {
            b.lo[d] = min(b.lo[d], box[k].lo[d]);
            b.hi[d] = max(b.hi[d], box[k].hi[d]);
        }
    return b;
}

// A += h/6 (dA1 + dA4 + 2 (dA2 + dA3)) over the active box
// (puffin_mpi_RK4.f90:281-282, 314-315), then clear the four deposits there.
// The periodic mesh keeps its aliased last plane equal to plane 0.
template <typename R>
__global__ __launch_bounds__(256) void k_field_combine(CombineArgs<R> a) {
    const Box b = merged_box(a.box);
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
    const int threads = STAGE_THREADS;
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
