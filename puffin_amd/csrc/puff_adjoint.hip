// Discrete adjoint of the RK4 step: one kernel per stage (4 .. 1) pulling the
// cotangent of the stage derivative back through the trial state, the
// tri-linear gather, the current deposit, the undulator fields and the six
// scaled equations, plus the field combination and the cotangent of the
// integrated power.
//
// There is no reference counterpart (the reference computes no derivatives).
// The transposed arithmetic is that of the forward stage kernels
// (puff_push.hip), which restate
//   getrhs          interaction/rhs.f90:60-330
//   getP2           utilities/gamma_to_p2.f90:84-85
//   getInterps_3D   interaction/system_interpolation.f90:25-98
//   getFFelecs_3D   interaction/system_interpolation.f90:110-142
//   getSource_3D    interaction/system_interpolation.f90:150-248
//   equations       interaction/puffin_equations.f90:17-176
//
// Per particle, with u = 1 + eta p2 = D^(-1/2), D = 1 - (1 + aw^2 |p|^2) /
// (gamma_r Gamma)^2 and fac = u / Gamma, the stage derivative is
//   dx = c_dx fac pr, dy = -c_dx fac pi, dz2 = (u - 1) / eta,
//   dpr = (n2col by - (u-1)/kappa^2 Fr) / 2rho + kappa pi fac n2col bz,
//   dpi = (n2col bx - (u-1)/kappa^2 Fi) / 2rho - kappa pr fac n2col bz,
//   dGamma = -2 rho fac (pr Fr + pi Fi),
//   dA[node_k] += w_k chi / dV3 fac (pr, pi),
// with F = sum_k w_k A[node_k].  Its transpose applied to the cotangent
// (a_x .. a_Gamma, G at the nodes) is formed below term by term; the weights
// are differentiated inside the particle's cell.
#include "puff_stage.hpp"

namespace puff {

// Trial field of the stage at one node: A0 + c dA_{s-1} (as the forward
// stage forms it).
template <typename R, int STAGE>
__device__ __forceinline__ cplx<R> adj_trial_field(const AdjArgs<R>& a, int node, R cA) {
    cplx<R> v = a.A0[node];
    if (STAGE > 1) {
        const cplx<R> d = a.dAin[node];
        v.x = fma(cA, d.x, v.x);
        v.y = fma(cA, d.y, v.y);
    }
    return v;
}

// Cotangent of the stage's field source at one node: w_s LA + c_{s+1} mA_{s+1}.
template <typename R, int STAGE>
__device__ __forceinline__ double2 adj_source_cotangent(const AdjArgs<R>& a, int node) {
    const cplx<R> l = a.LA[node];
    double gr = a.ws * (double)l.x, gi = a.ws * (double)l.y;
    if (STAGE < 4) {
        const cplx<R> t = a.mAnext[node];
        gr += a.cnext * (double)t.x;
        gi += a.cnext * (double)t.y;
    }
    return double2{gr, gi};
}

template <typename R, int STAGE>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_stage(AdjArgs<R> a) {
    __shared__ int s_skip;
    if (threadIdx.x == 0) s_skip = *a.flags;
    __syncthreads();
    if (s_skip) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < a.n;
    const MeshGeom& m = a.mesh;
    const Physics& ph = a.phys;

    // ---- trial state, formed exactly as the forward stage forms it, and
    // the cotangent of this stage's particle derivative.  All of the
    // particle's loads are issued before any of them is used: split into a
    // block per quantity, the cotangent loads waited for the trial state's.
    R x = 0, y = 0, pr = 0, pi = 0, chi = 0;
    double z2 = 0.0, g = 1.0;
    double ax = 0, ay = 0, az = 0, apr = 0, api = 0, ag = 0;
    if (in) {
        const R x0 = a.y0.x[i], y0 = a.y0.y[i], pr0 = a.y0.pr[i], pi0 = a.y0.pi[i];
        const double z20 = a.y0.z2[i], g0 = a.y0.g[i];
        R kx = 0, ky = 0, kpr = 0, kpi = 0;
        double kz2 = 0.0, kg = 0.0;
        if (STAGE > 1) {
            kx = a.kin.x[i];
            ky = a.kin.y[i];
            kz2 = a.kin.z2[i];
            kpr = a.kin.pr[i];
            kpi = a.kin.pi[i];
            kg = a.kin.g[i];
        }
        chi = a.chi[i];
        const bool ev = ph.electrons_evolve;
        double Lx = 0, Ly = 0, Lz2 = 0, Lpr = 0, Lpi = 0, Lg = 0;
        double nx = 0, ny = 0, nz2 = 0, npr = 0, npi = 0, ng = 0;
        if (ev) {
            Lx = (double)a.L.x[i];
            Ly = (double)a.L.y[i];
            Lz2 = a.L.z2[i];
            Lpr = (double)a.L.pr[i];
            Lpi = (double)a.L.pi[i];
            Lg = a.L.g[i];
            if (STAGE < 4) {
                nx = (double)a.mnext.x[i];
                ny = (double)a.mnext.y[i];
                nz2 = a.mnext.z2[i];
                npr = (double)a.mnext.pr[i];
                npi = (double)a.mnext.pi[i];
                ng = a.mnext.g[i];
            }
        }
        if (STAGE == 1) {
            x = x0; y = y0; z2 = z20; pr = pr0; pi = pi0; g = g0;
        } else {
            const R c = (R)a.cin;
            x = fma(c, kx, x0);
            y = fma(c, ky, y0);
            z2 = fma(a.cin, kz2, z20);
            pr = fma(c, kpr, pr0);
            pi = fma(c, kpi, pi0);
            g = fma(a.cin, kg, g0);
        }
        if (ev) {
            ax = a.ws * Lx;
            ay = a.ws * Ly;
            az = a.ws * Lz2;
            apr = a.ws * Lpr;
            api = a.ws * Lpi;
            ag = a.ws * Lg;
            if (STAGE < 4) {
                ax += a.cnext * nx;
                ay += a.cnext * ny;
                az += a.cnext * nz2;
                apr += a.cnext * npr;
                api += a.cnext * npi;
                ag += a.cnext * ng;
            }
        }
    }

    // ---- p2 family in FP64 (gamma_to_p2.f90:84-85) and its derivatives
    const double prd = (double)pr, pid = (double)pi;
    const double aw2 = ph.aw * ph.aw;
    const double s2 = prd * prd + pid * pid;
    const double invg = 1.0 / g;
    const double u = in ? 1.0 / sqrt(1.0 - ph.inv_gr2 * invg * invg * (1.0 + aw2 * s2)) : 1.0;
    const double p2 = (u - 1.0) * ph.inv_eta;
    const double fac = (1.0 + ph.eta * p2) * invg;
    const double u3 = u * u * u;
    const double q = u3 * ph.inv_gr2 * invg * invg;
    const double du_dpr = q * aw2 * prd;
    const double du_dpi = q * aw2 * pid;
    const double du_dg = -q * (1.0 + aw2 * s2) * invg;

    // ---- cell, box, weights
    Cell<R> c{};
    const int flag = in ? locate<R>(m, x, y, z2, c) : 0;
    const bool ok = in && flag == 0;
    if (flag) atomicOr(a.flags, flag);
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const TileBox tb = workgroup_box<R>(m, ok, c, a.box);
    const bool scat = ph.coupling && ph.electrons_evolve;
    const bool use_tile = scat && tile_fits<ADJ_TILE_NODES>(m, tb);
    const int nnb = m.one_d ? 2 : 8;
    const CellRef cr = cell_ref<R>(m, ok, c, tb, lane);
    // derivatives of the weights inside the cell, formed per corner in the
    // gather below: dw_k/dx = dX[ia] Y[ib] Z[ic] and likewise for y, z2
    // (1-D: dw/dz2 = -+1/dz2)
    const double X[2] = {(double)((R)1 - c.fx), (double)c.fx};
    const double Y[2] = {(double)((R)1 - c.fy), (double)c.fy};
    const double Z[2] = {(double)((R)1 - c.fz), (double)c.fz};

    // ---- undulator fields and their transverse derivatives
    R bxr = 0, byr = 0, bzr = 0;
    if (ok) bfield<R>(a.s, x, y, bxr, byr, bzr);
    const double bx = bxr, by = byr, bz = bzr;
    const double n2col = a.s.n2col;
    const double ep2k = ph.eta * p2 / (ph.kappa * ph.kappa);
    // d/dF of the dp_r, dp_i and dGamma equations
    const double Lfr = ph.coupling ? (-apr * ph.inv2rho * ep2k - ag * 2.0 * ph.rho * fac * prd) : 0.0;
    const double Lfi = ph.coupling ? (-api * ph.inv2rho * ep2k - ag * 2.0 * ph.rho * fac * pid) : 0.0;
    const double dep = ph.field_evolve ? (double)chi * ph.inv_dV3 : 0.0;
    const double depfac = dep * fac;

    // ---- gather of the trial field and of the source cotangent.  When the
    // workgroup's node box is small (the dense part of the beam), both are
    // formed once per node of the box in LDS and the corners read from there
    // (same arithmetic as the global reads).
    const R cA = (R)a.cin;
    const bool lds_field = tile_fits<ADJ_TILE_NODES>(m, tb) && tb.tx * tb.ty * tb.tz <= ADJ_FIELD_TILE_NODES;
    __shared__ cplx<R> s_v[ADJ_FIELD_TILE_NODES];
    __shared__ double2 s_g[ADJ_FIELD_TILE_NODES];
    if (lds_field) {
        static_assert(ADJ_FIELD_TILE_NODES <= ADJ_THREADS, "one field tile node per thread");
        const int q = threadIdx.x;
        if (q < tb.tx * tb.ty * tb.tz) {
            int lx, ly, lz;
            box_coords(box_div(tb), q, lx, ly, lz);
            const int node = (tb.bz0 + lz) * m.nx * m.ny + (tb.by0 + ly) * m.nx + (tb.bx0 + lx);
            s_v[q] = adj_trial_field<R, STAGE>(a, node, cA);
            s_g[q] = adj_source_cotangent<R, STAGE>(a, node);
        }
        __syncthreads();
    }
    double fr = 0, fi = 0, Gr = 0, Gi = 0, sx = 0, sy = 0, sz = 0;
    if (ok) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k < nnb) {
                cplx<R> v;
                double gr, gi;
                if (lds_field) {
                    const int tk = corner_tile(m, tb, cr, k);
                    v = s_v[tk];
                    const double2 gg = s_g[tk];
                    gr = gg.x;
                    gi = gg.y;
                } else {
                    const int node = corner_node(m, cr, k);
                    v = adj_trial_field<R, STAGE>(a, node, cA);
                    const double2 gg = adj_source_cotangent<R, STAGE>(a, node);
                    gr = gg.x;
                    gi = gg.y;
                }
                const double wk = (double)corner_weight<R>(m, c, k);
                fr += wk * (double)v.x;
                fi += wk * (double)v.y;
                Gr += wk * gr;
                Gi += wk * gi;
                const double V = Lfr * (double)v.x + Lfi * (double)v.y + depfac * (prd * gr + pid * gi);
                const int ia = k & 1, ib = (k >> 1) & 1, ic = k >> 2;
                if (m.one_d) {
                    sz += (ia ? m.inv_dz2 : -m.inv_dz2) * V;
                } else {
                    const double dX = ia ? m.inv_dx : -m.inv_dx;
                    const double dY = ib ? m.inv_dy : -m.inv_dy;
                    const double dZ = ic ? m.inv_dz2 : -m.inv_dz2;
                    sx += dX * Y[ib] * Z[ic] * V;
                    sy += X[ia] * dY * Z[ic] * V;
                    sz += X[ia] * Y[ib] * dZ * V;
                }
            }
        }
    }
    if (!ph.coupling) { fr = 0; fi = 0; }   // rhs.f90:221-224

    // ---- field cotangent of this stage: w_k (dL/dFr, dL/dFi) at the nodes
    if (scat) {
        __shared__ double s_tile[2 * ADJ_TILE_NODES];
        scatter_cells<R>(m, tb, use_tile, ok, c, cr, ok ? (R)Lfr : (R)0,
                         ok ? (R)Lfi : (R)0, a.mAout, s_tile);
    }

    // ---- dJ/dn2col of this stage: workgroup sum into a spread accumulator
    {
        double dn = 0.0;
        if (ok)
            dn = apr * (ph.inv2rho * by + ph.kappa * pid * fac * bz) +
                 api * (ph.inv2rho * bx - ph.kappa * prd * fac * bz);
        dn = wave_sum(dn);
        __shared__ double s_dn[ADJ_THREADS / 64];
        if (lane == 0) s_dn[wv] = dn;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += s_dn[k];
            if (t != 0.0) atomicAdd(&a.dn2[blockIdx.x % DN2_COPIES], t);
        }
    }
    if (!in) return;

    // ---- particle cotangent mu_s
    double mx = 0, my = 0, mz2 = 0, mpr = 0, mpi = 0, mg = 0;
    if (ok) {
        const double k2 = ph.kappa * ph.kappa;
        const double nkb = ph.kappa * n2col * bz;
        const double Lu = az / ph.eta - (apr * fr + api * fi) * ph.inv2rho / k2;
        const double Lfac = ax * ph.c_dx * prd - ay * ph.c_dx * pid + nkb * (apr * pid - api * prd) -
                            2.0 * ph.rho * ag * (prd * fr + pid * fi) + dep * (prd * Gr + pid * Gi);
        const double Epr = ax * ph.c_dx * fac - api * nkb * fac - 2.0 * ph.rho * ag * fac * fr + depfac * Gr;
        const double Epi = -ay * ph.c_dx * fac + apr * nkb * fac - 2.0 * ph.rho * ag * fac * fi + depfac * Gi;
        const double dfac_dpr = du_dpr * invg, dfac_dpi = du_dpi * invg;
        const double dfac_dg = du_dg * invg - fac * invg;
        mpr = Epr + Lu * du_dpr + Lfac * dfac_dpr;
        mpi = Epi + Lu * du_dpi + Lfac * dfac_dpi;
        mg = Lu * du_dg + Lfac * dfac_dg;
        double bg[6];   // transverse derivatives of the undulator fields
        bfield_grad(a.s, (double)x, (double)y, bg);
        const double Lbx = api * ph.inv2rho * n2col, Lby = apr * ph.inv2rho * n2col;
        const double Lbz = ph.kappa * fac * n2col * (apr * pid - api * prd);
        mx = sx + Lbx * bg[0] + Lby * bg[2] + Lbz * bg[4];
        my = sy + Lbx * bg[1] + Lby * bg[3] + Lbz * bg[5];
        mz2 = sz;
    }
    if (STAGE > 1) {
        a.mout.x[i] = (R)mx;
        a.mout.y[i] = (R)my;
        a.mout.z2[i] = mz2;
        a.mout.pr[i] = (R)mpr;
        a.mout.pi[i] = (R)mpi;
        a.mout.g[i] = mg;
    } else {
        // L <- L + mu_1 + mu_2 + mu_3 + mu_4
        a.L.x[i] = (R)((double)a.L.x[i] + mx + (double)a.mnext.x[i] + (double)a.m3.x[i] + (double)a.m4.x[i]);
        a.L.y[i] = (R)((double)a.L.y[i] + my + (double)a.mnext.y[i] + (double)a.m3.y[i] + (double)a.m4.y[i]);
        a.L.z2[i] = a.L.z2[i] + mz2 + a.mnext.z2[i] + a.m3.z2[i] + a.m4.z2[i];
        a.L.pr[i] = (R)((double)a.L.pr[i] + mpr + (double)a.mnext.pr[i] + (double)a.m3.pr[i] + (double)a.m4.pr[i]);
        a.L.pi[i] = (R)((double)a.L.pi[i] + mpi + (double)a.mnext.pi[i] + (double)a.m3.pi[i] + (double)a.m4.pi[i]);
        a.L.g[i] = a.L.g[i] + mg + a.mnext.g[i] + a.m3.g[i] + a.m4.g[i];
    }
}

// L_A += mu_A1 + mu_A2 + mu_A3 + mu_A4 over the step's active box, then clear
// the four stage cotangents and the three forward sources there.  The aliased
// last plane of a periodic mesh is never written (its cotangent stays 0).
template <typename R>
__global__ __launch_bounds__(256) void k_adj_combine(AdjCombineArgs<R> a) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.flags[0]) atomicOr(a.flags + FLAG_WORD_STICKY, a.flags[0]);
    const Box b = merged_box(a.box);
    if (b.hi[0] < 0) return;
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
    const int64_t ntr = (int64_t)a.nx * a.ny;
    const cplx<R> zero = {(R)0, (R)0};
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < vol;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int ix = x0 + (int)(t % lx);
        const int64_t r = t / lx;
        const int iy = y0 + (int)(r % ly);
        const int iz = z0 + (int)(r / ly);
        const int64_t j = (int64_t)iz * ntr + (int64_t)iy * a.nx + ix;
        const cplx<R> m0 = a.mA[0][j], m1 = a.mA[1][j], m2 = a.mA[2][j], m3 = a.mA[3][j];
        cplx<R> v = a.LA[j];
        v.x = (R)((double)v.x + (double)m0.x + (double)m1.x + (double)m2.x + (double)m3.x);
        v.y = (R)((double)v.y + (double)m0.y + (double)m1.y + (double)m2.y + (double)m3.y);
        a.LA[j] = v;
        a.mA[0][j] = zero;
        a.mA[1][j] = zero;
        a.mA[2][j] = zero;
        a.mA[3][j] = zero;
        if (a.clear_dA) {
            a.dA[0][j] = zero;
            a.dA[1][j] = zero;
            a.dA[2][j] = zero;
        }
    }
}

// Cotangent of J = sum_iz wz[iz] power(iz), power(iz) = transverse trapezoid
// of |A|^2 on a uniform axis (|A|^2 in 1-D).  The aliased periodic plane's
// share goes to plane 1, which it copies.
template <typename R>
__global__ __launch_bounds__(256) void k_adj_seed_power(cplx<R>* LA, const cplx<R>* A,
                                                        const double* wz, int nx, int ny,
                                                        int nz2, int periodic, double dx,
                                                        double dy, double scale) {
    const int64_t ntr = (int64_t)nx * ny;
    const int nzp = periodic ? nz2 - 1 : nz2;
    const int64_t n = ntr * nzp;
    const bool one = (nx == 1 && ny == 1);
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int iz = (int)(j / ntr);
        const int64_t t = j % ntr;
        const int ix = (int)(t % nx), iy = (int)(t / nx);
        double wxy = 1.0;
        if (!one) {
            wxy = dx * dy;
            if (ix == 0 || ix == nx - 1) wxy *= 0.5;
            if (iy == 0 || iy == ny - 1) wxy *= 0.5;
        }
        const cplx<R> v = A[j];
        double gr = 2.0 * scale * wz[iz] * wxy * (double)v.x;
        double gi = 2.0 * scale * wz[iz] * wxy * (double)v.y;
        if (periodic && iz == 0) {
            const cplx<R> va = A[(int64_t)(nz2 - 1) * ntr + t];
            gr += 2.0 * scale * wz[nz2 - 1] * wxy * (double)va.x;
            gi += 2.0 * scale * wz[nz2 - 1] * wxy * (double)va.y;
        }
        cplx<R> l = LA[j];
        l.x = (R)((double)l.x + gr);
        l.y = (R)((double)l.y + gi);
        LA[j] = l;
    }
}

// Per-stage dJ/dn2col into the per-slot gradient table (one wavefront).
__global__ void k_dn2_accumulate(const double* dn2, Slots4 slots, int ntab, double* table) {
    const int lane = threadIdx.x;
    for (int st = 0; st < 4; ++st) {
        double v = (lane < DN2_COPIES) ? dn2[st * DN2_COPIES + lane] : 0.0;
        v = wave_sum(v);
        if (lane == st) {
            const int k = slots.s[st] < 0 ? ntab : slots.s[st];
            table[k] += v;
        }
    }
}

void launch_dn2_accumulate(const double* dn2, Slots4 slots, int ntab, double* dn2_table, hipStream_t st) {
    hipLaunchKernelGGL(k_dn2_accumulate, dim3(1), dim3(64), 0, st, dn2, slots, ntab, dn2_table);
}

// Periodic mesh: fold the aliased last plane's cotangent into plane 1.
template <typename R>
__global__ __launch_bounds__(256) void k_fold_alias(cplx<R>* LA, int64_t ntr, int nz2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntr) return;
    cplx<R>* al = LA + (int64_t)(nz2 - 1) * ntr + t;
    const cplx<R> a = *al;
    cplx<R> v = LA[t];
    v.x += a.x;
    v.y += a.y;
    LA[t] = v;
    *al = cplx<R>{(R)0, (R)0};
}

__global__ __launch_bounds__(256) void k_fill_zero_words(uint32_t* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// ---- host launchers --------------------------------------------------------

template <typename R>
void launch_fold_alias(cplx<R>* LA, int64_t ntr, int nz2, hipStream_t s) {
    if (ntr <= 0) return;
    hipLaunchKernelGGL((k_fold_alias<R>), dim3((unsigned)((ntr + 255) / 256)), dim3(256), 0, s, LA, ntr, nz2);
}

void launch_fill_zero_words(uint32_t* p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill_zero_words, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
}

template <typename R>
void launch_adj_stage(int stage, const AdjArgs<R>& a, hipStream_t s) {
    const int threads = ADJ_THREADS;
    const int64_t blocks = (a.n + threads - 1) / threads;
    dim3 grid((unsigned)(blocks > 0 ? blocks : 1));
    switch (stage) {
        case 1: hipLaunchKernelGGL((k_adj_stage<R, 1>), grid, dim3(threads), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_adj_stage<R, 2>), grid, dim3(threads), 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_adj_stage<R, 3>), grid, dim3(threads), 0, s, a); break;
        default: hipLaunchKernelGGL((k_adj_stage<R, 4>), grid, dim3(threads), 0, s, a); break;
    }
}

template <typename R>
void launch_adj_combine(const AdjCombineArgs<R>& a, hipStream_t s) {
    hipLaunchKernelGGL((k_adj_combine<R>), dim3(2048), dim3(256), 0, s, a);
}

template <typename R>
void launch_adj_seed_power(cplx<R>* LA, const cplx<R>* A, const double* wz, int nx, int ny,
                           int nz2, int periodic, double dx, double dy, double scale,
                           hipStream_t s) {
    const int64_t n = (int64_t)nx * ny * (periodic ? nz2 - 1 : nz2);
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_adj_seed_power<R>), dim3((unsigned)blocks), dim3(256), 0, s, LA, A, wz,
                       nx, ny, nz2, periodic, dx, dy, scale);
}

template void launch_fold_alias<double>(cplx<double>*, int64_t, int, hipStream_t);
template void launch_fold_alias<float>(cplx<float>*, int64_t, int, hipStream_t);
template void launch_adj_stage<double>(int, const AdjArgs<double>&, hipStream_t);
template void launch_adj_stage<float>(int, const AdjArgs<float>&, hipStream_t);
template void launch_adj_combine<double>(const AdjCombineArgs<double>&, hipStream_t);
template void launch_adj_combine<float>(const AdjCombineArgs<float>&, hipStream_t);
template void launch_adj_seed_power<double>(cplx<double>*, const cplx<double>*, const double*, int,
                                            int, int, int, double, double, double, hipStream_t);
template void launch_adj_seed_power<float>(cplx<float>*, const cplx<float>*, const double*, int,
                                           int, int, int, double, double, double, hipStream_t);

}  // namespace puff
