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
#include "puff_stage.hpp"

namespace puff {

// One RK4 stage.  STAGE = 1..4 (rk4par's four derivs calls).  DET: the
// deterministic fixed-point deposit (FP64 form only; see DetState).
template <typename R, int STAGE, bool DET>
__global__ __launch_bounds__(STAGE_THREADS) void k_stage(StageArgs<R> a) {
    // Flags raised by an earlier kernel of this step do not stop stages 1-3:
    // they write only the step's scratch (k_s, dA_s), and a particle whose
    // trial position is off the mesh takes no part in the gather and deposit
    // (locate raises the flag again).  Stage 4 writes the state, and only
    // when no flag is set: read here, with the particle loads (no kernel of
    // stage 4 raises a flag: stage 3 checked its trial positions).
    const int rejected = (STAGE == 4) ? *a.flags : 0;
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
            if (a.ycap.x) {
                a.ycap.x[i] = x0;
                a.ycap.y[i] = y0;
                a.ycap.z2[i] = z20;
                a.ycap.pr[i] = pr0;
                a.ycap.pi[i] = pi0;
                a.ycap.g[i] = g0;
                a.chicap[i] = chi;
            }
        } else {
            const R c = (R)a.cin;
            const R k3x = a.kin.x[i], k3y = a.kin.y[i], k3pr = a.kin.pr[i], k3pi = a.kin.pi[i];
            const double k3z2 = a.kin.z2[i], k3g = a.kin.g[i];
            x = fma(c, k3x, x0);
            y = fma(c, k3y, y0);
            z2 = fma(a.cin, k3z2, z20);
            pr = fma(c, k3pr, pr0);
            pi = fma(c, k3pi, pi0);
            g = fma(a.cin, k3g, g0);
            if (STAGE == 4) {
                // y0 + h/6 (k1 + 2 (k2 + k3)) formed now, with k1 and k2 loaded
                // beside y0 and k3: the end of the kernel adds h/6 k4 only, and
                // no k array is read after the deposit (puffin_mpi_RK4.f90:274-279)
                const R h6 = (R)a.h6;
                x0 = fma(h6, a.k1.x[i] + (R)2 * (a.k2.x[i] + k3x), x0);
                y0 = fma(h6, a.k1.y[i] + (R)2 * (a.k2.y[i] + k3y), y0);
                z20 = fma(a.h6, a.k1.z2[i] + 2.0 * (a.k2.z2[i] + k3z2), z20);
                pr0 = fma(h6, a.k1.pr[i] + (R)2 * (a.k2.pr[i] + k3pr), pr0);
                pi0 = fma(h6, a.k1.pi[i] + (R)2 * (a.k2.pi[i] + k3pi), pi0);
                g0 = fma(a.h6, a.k1.g[i] + 2.0 * (a.k2.g[i] + k3g), g0);
            }
        }
    }

    // ---- p2 in FP64 (gamma_to_p2.f90:84-85), with 1/eta and 1/gamma_r^2
    // from the host and 1/Gamma formed once
    const double prd = (double)pr, pid = (double)pi;
    const double invg = 1.0 / g;
    const double p2 = in ? ((1.0 / sqrt(1.0 - ph.inv_gr2 * invg * invg *
                                               (1.0 + ph.aw * ph.aw * (prd * prd + pid * pid)))) -
                            1.0) * ph.inv_eta
                         : 0.0;
    const double fac = (1.0 + ph.eta * p2) * invg;   // (1 + eta p2) / Gamma

    // ---- node and weights
    Cell<R> c{};
    int flag = in ? locate<R>(m, x, y, z2, c) : 0;
    const bool ok = in && flag == 0;
    if (flag) atomicOr(a.flags, flag);

    // ---- sources (getSource): s = chi / dV3 (1 + eta p2) p / Gamma
    const R sr = ok ? (R)((double)chi * ph.inv_dV3 * fac * prd) : (R)0;
    const R si = ok ? (R)((double)chi * ph.inv_dV3 * fac * pid) : (R)0;

    // ---- node box of this workgroup's cells and the step's active box (and,
    // for the packed fixed-point tile, the workgroup's largest |s|)
    const int lane = threadIdx.x & 63;
    constexpr bool PACKED = sizeof(R) == 4 && PUFF_MIXED_TILE == 0 && !DET;
    const TileBox tb = workgroup_box<R, PACKED ? 7 : 6>(
        m, ok, c, a.box, PACKED ? __float_as_int(fmaxf(fabsf((float)sr), fabsf((float)si))) : 0);
    const bool use_tile = ph.field_evolve && tile_fits<stage_tile_nodes<R>()>(m, tb);
    if (threadIdx.x == 0 && ph.field_evolve && tb.any && !use_tile) atomicAdd(a.flags + FLAG_WORD_FALLBACK, 1);

    // ---- nodes of the cell (weights are formed per corner where used)
    const int nnb = m.one_d ? 2 : 8;
    const CellRef cr = cell_ref<R>(m, ok, c, tb, lane);

    // ---- gather F = sum_k w_k (A0 + c dA_{s-1})[node_k]  (getFFelecs).
    // When the workgroup's node box is small (the dense part of the beam),
    // A0 + c dA_{s-1} is formed once per node of the box in LDS and the
    // corners are read from there; else from global memory.  Same arithmetic
    // either way.
    // The field tile shares the deposit tile's LDS: it is used only for boxes
    // of at most FIELD_TILE_NODES nodes, whose deposit needs the first
    // WORDS x FIELD_TILE_NODES words, so it sits right after them.
    // Deposit tiles: two doubles per node in the FP64 form; DET_WORDS 8-byte
    // integers in its deterministic mode; one 8-byte packed fixed-point word
    // in the mixed form (scatter_cells_packed: one ds_add_u64 per corner for
    // both components).  Measured on the C1 beam, mixed form: an FP32 tile
    // (ds_add_f32) ran each stage 3.1x slower than the FP64 tile, and a
    // compensated FP32 tile (ds_add_rtn_f32 + TwoSum) 5.7x slower, with the
    // same count of LDS instructions per wave but 6x the LDS-active and 24x
    // the LDS-wait cycles of ds_add_f64; the packed tile cut the LDS-wait
    // cycles per wave from 398 to 6 and the step by 3-4 %.
    constexpr int TILE = stage_tile_nodes<R>();
    using D = typename std::conditional<
        PACKED, unsigned long long,
        typename std::conditional<sizeof(R) == 4 && PUFF_MIXED_TILE == 2, float, double>::type>::type;
    constexpr int WORDS = DET ? DET_WORDS : (PACKED ? 1 : 2);
    static_assert(WORDS * FIELD_TILE_NODES * sizeof(D) + FIELD_TILE_NODES * sizeof(cplx<R>) <=
                      WORDS * TILE * sizeof(D), "field tile inside the deposit tile");
    __shared__ D s_dep[WORDS * TILE];
    cplx<R>* s_fld = reinterpret_cast<cplx<R>*>(s_dep + WORDS * FIELD_TILE_NODES);
    const R cA = (R)a.cin;
    const bool lds_field = use_tile && tb.tx * tb.ty * tb.tz <= FIELD_TILE_NODES;
    if (lds_field) {
        const int nt = tb.tx * tb.ty * tb.tz;
        const int ntr = m.nx * m.ny;
        const BoxDiv bd = box_div(tb);
        for (int q = threadIdx.x; q < nt; q += blockDim.x) {
            int lx, ly, lz;
            box_coords(bd, q, lx, ly, lz);
            const int node = (tb.bz0 + lz) * ntr + (tb.by0 + ly) * m.nx + (tb.bx0 + lx);
            cplx<R> v = a.A0[node];
            if (STAGE > 1) {
                const cplx<R> d = a.dAin[node];
                v.x = fma(cA, d.x, v.x);
                v.y = fma(cA, d.y, v.y);
            }
            s_fld[q] = v;
        }
        __syncthreads();
    }
    R fr = 0, fi = 0;
    if (ok && lds_field && !m.one_d) {
        // getFFelecs' corner order, the weights formed once per lane
        const CornerSet<R> cs = corner_set<R>(c, tb, 0);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const cplx<R> v = s_fld[cr.t0 + cs.oxy[k & 3] + cs.oz[k >> 2]];
            const R wk = cs.xy[k & 3] * cs.z[k >> 2];
            fr = fma(wk, v.x, fr);
            fi = fma(wk, v.y, fi);
        }
    } else if (ok) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k < nnb) {
                cplx<R> v;
                if (lds_field) {
                    v = s_fld[corner_tile(m, tb, cr, k)];
                } else {
                    const int node = corner_node(m, cr, k);
                    v = a.A0[node];
                    if (STAGE > 1) {
                        const cplx<R> d = a.dAin[node];
                        v.x = fma(cA, d.x, v.x);
                        v.y = fma(cA, d.y, v.y);
                    }
                }
                const R wk = corner_weight<R>(m, c, k);
                fr = fma(wk, v.x, fr);
                fi = fma(wk, v.y, fi);
            }
        }
    }

    // ---- deposit w_k s into dA_s (getSource)
    if (ph.field_evolve) {
        if constexpr (DET) {
            // n |s| 2^F <= 2^62 keeps every node's integer sums in range
            const int F = *a.det.shift;
            const double lim = ldexp(1.0, 62 - F) / (double)a.n;
            if (ok && (fabs((double)sr) > lim || fabs((double)si) > lim)) atomicOr(a.flags, FLAG_DET_RANGE);
            scatter_cells_det(m, tb, use_tile, ok, c, cr, (double)sr, (double)si, ldexp(1.0, F), a.det.acc,
                              reinterpret_cast<unsigned long long*>(s_dep));
        } else if constexpr (PACKED) {
            if (use_tile) {
                scatter_cells_packed(m, tb, ok, c, cr, (float)sr, (float)si, packed_scale_log2(tb.extra),
                                     reinterpret_cast<cplx<float>*>(a.dAout), s_dep);
            } else {
                scatter_cells<R, float>(m, tb, false, ok, c, cr, sr, si, a.dAout, (float*)nullptr);
            }
        } else {
            scatter_cells<R, D>(m, tb, use_tile, ok, c, cr, sr, si, a.dAout, s_dep);
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
        // stored in R (FP32 in the mixed form: see KPtrs)
        const R kz2s = (R)kz2;
        a.kout.x[i] = kx_;
        a.kout.y[i] = ky_;
        a.kout.z2[i] = kz2s;
        a.kout.pr[i] = kpr;
        a.kout.pi[i] = kpi;
        a.kout.g[i] = (R)kg;
        // next stage's trial state must lie inside the mesh, else the next
        // derivs call would raise the flag (checked here, on the stored
        // derivative, so that no later kernel writes state when it does)
        const R cn = (R)a.cnext;
        Cell<R> cc;
        const int nf = locate<R>(m, fma(cn, kx_, x0), fma(cn, ky_, y0), fma(a.cnext, (double)kz2s, z20), cc);
        if (nf) atomicOr(a.flags, nf);
    } else {
        // y = (y0 + h/6 (k1 + 2 (k2 + k3))) + h/6 k4  (puffin_mpi_RK4.f90:274-279, 307-312)
        if (rejected) return;   // the step is rejected
        const R h6 = (R)a.h6;
        a.y0.x[i] = fma(h6, kx_, x0);
        a.y0.y[i] = fma(h6, ky_, y0);
        a.y0.z2[i] = fma(a.h6, kz2, z20);
        a.y0.pr[i] = fma(h6, kpr, pr0);
        a.y0.pi[i] = fma(h6, kpi, pi0);
        a.y0.g[i] = fma(a.h6, kg, g0);
    }
}

// Start of a step: clear flags and the active box.
__global__ void k_step_begin(int* flags, Box* box) {
    if (threadIdx.x < FLAG_WORD_STICKY) flags[threadIdx.x] = 0;
    for (int k = threadIdx.x; k < BOX_COPIES; k += blockDim.x) {
        for (int d = 0; d < 3; ++d) {
            box[k].lo[d] = 0x7fffffff;
            box[k].hi[d] = -1;
        }
    }
}

// A += h/6 (dA1 + dA4 + 2 (dA2 + dA3)) over the active box
// (puffin_mpi_RK4.f90:281-282, 314-315), then clear the deposits there (all
// four, or only the last when the first three are kept for the adjoint).
// The periodic mesh keeps its aliased last plane equal to plane 0.
template <typename R>
__global__ __launch_bounds__(256) void k_field_combine(CombineArgs<R> a) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.flags[0]) atomicOr(a.flags + FLAG_WORD_STICKY, a.flags[0]);
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
        for (int k = a.clear_from; k < 4; ++k) a.d[k][j] = zero;
    }
}

// ---- deterministic deposit ---------------------------------------------------

// Largest |source| chi / dV3 (1 + eta p2) |p_r|, |p_i| / Gamma of the state y,
// reduced as the bits of a non-negative double: their integer order is the
// value order, so the integer maximum is exact whatever the atomics' order.
__global__ __launch_bounds__(256) void k_det_bound(PartPtrs<double> y, const double* chi, int64_t n,
                                                   Physics ph, unsigned long long* bound) {
    double v = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double pr = y.pr[i], pi = y.pi[i];
        const double invg = 1.0 / y.g[i];
        const double p2 = ((1.0 / sqrt(1.0 - ph.inv_gr2 * invg * invg * (1.0 + ph.aw * ph.aw * (pr * pr + pi * pi)))) -
                           1.0) * ph.inv_eta;
        const double s = chi[i] * ph.inv_dV3 * (1.0 + ph.eta * p2) * invg;
        v = fmax(v, fmax(fabs(s * pr), fabs(s * pi)));
    }
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0 && v > 0.0) atomicMax(bound, (unsigned long long)__double_as_longlong(v));
}

// F of the step: DET_SAFETY n max|s| < 2^(62 - F), i.e. 62 minus the binary
// exponent e of that bound (frexp: bound < 2^e).  Clears the bound word for
// the next step.
__global__ void k_det_shift(unsigned long long* bound, int* shift, int64_t n) {
    const double m = __longlong_as_double((long long)*bound);
    int F = 0;
    if (m > 0.0) {
        int e = 0;
        frexp((double)DET_SAFETY * (double)n * m, &e);
        F = 62 - e;
        F = F > 1000 ? 1000 : (F < -1000 ? -1000 : F);
    }
    *shift = F;
    *bound = 0ull;
}

// dA = Re hi 2^-F + Re lo 2^-(F + DET_LO_BITS) (and Im) over the step's
// active box so far, clearing the integer sums there.
__global__ __launch_bounds__(256) void k_det_convert(DetState d, cplx<double>* dA, const Box* box, int nx,
                                                     int ny, int nzp, int periodic, int one_d) {
    const Box b = merged_box(box);
    if (b.hi[0] < 0) return;
    const int x0 = b.lo[0], x1 = min(b.hi[0] + 1, nx - 1);
    const int y0 = one_d ? 0 : b.lo[1], y1 = one_d ? 0 : min(b.hi[1] + 1, ny - 1);
    const int z0 = periodic ? 0 : b.lo[2];
    const int z1 = periodic ? nzp - 1 : min(b.hi[2] + 1, nzp - 1);
    const int lx = x1 - x0 + 1, ly = y1 - y0 + 1, lz = z1 - z0 + 1;
    const int64_t vol = (int64_t)lx * ly * lz;
    const int F = *d.shift;
    const double inv_hi = ldexp(1.0, -F), inv_lo = ldexp(1.0, -(F + DET_LO_BITS));
    const int64_t ntr = (int64_t)nx * ny;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < vol;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int ix = x0 + (int)(t % lx);
        const int64_t r = t / lx;
        const int iy = y0 + (int)(r % ly);
        const int iz = z0 + (int)(r / ly);
        const int64_t j = (int64_t)iz * ntr + (int64_t)iy * nx + ix;
        unsigned long long* w = d.acc + DET_WORDS * j;
        const long long rh = (long long)w[0], rl = (long long)w[1], ih = (long long)w[2], il = (long long)w[3];
        dA[j] = {(double)rh * inv_hi + (double)rl * inv_lo, (double)ih * inv_hi + (double)il * inv_lo};
        if (rh | rl | ih | il) {
            w[0] = 0ull;
            w[1] = 0ull;
            w[2] = 0ull;
            w[3] = 0ull;
        }
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
    if constexpr (sizeof(R) == 8) {
        if (a.det.acc) {
            switch (stage) {
                case 1: hipLaunchKernelGGL((k_stage<R, 1, true>), grid, dim3(threads), 0, s, a); break;
                case 2: hipLaunchKernelGGL((k_stage<R, 2, true>), grid, dim3(threads), 0, s, a); break;
                case 3: hipLaunchKernelGGL((k_stage<R, 3, true>), grid, dim3(threads), 0, s, a); break;
                default: hipLaunchKernelGGL((k_stage<R, 4, true>), grid, dim3(threads), 0, s, a); break;
            }
            return;
        }
    }
    switch (stage) {
        case 1: hipLaunchKernelGGL((k_stage<R, 1, false>), grid, dim3(threads), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_stage<R, 2, false>), grid, dim3(threads), 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_stage<R, 3, false>), grid, dim3(threads), 0, s, a); break;
        default: hipLaunchKernelGGL((k_stage<R, 4, false>), grid, dim3(threads), 0, s, a); break;
    }
}

void launch_det_scale(const PartPtrs<double>& y, const double* chi, int64_t n, const Physics& ph,
                      const DetState& d, hipStream_t s) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_det_bound, dim3((unsigned)blocks), dim3(256), 0, s, y, chi, n, ph, d.bound);
    hipLaunchKernelGGL(k_det_shift, dim3(1), dim3(1), 0, s, d.bound, d.shift, n);
}

void launch_det_convert(const DetState& d, cplx<double>* dA, const Box* box, int nx, int ny, int nzp,
                        int periodic, int one_d, hipStream_t s) {
    hipLaunchKernelGGL(k_det_convert, dim3(2048), dim3(256), 0, s, d, dA, box, nx, ny, nzp, periodic, one_d);
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
