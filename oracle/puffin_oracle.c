/*
 * puffin_oracle.c -- CPU FP64 restatement of Puffin's per-step integrator.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP implementation
 * is compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load it.  The product never links or calls it.
 *
 * Every function cites the reference routine (UKFELs/Puffin, path:line) it
 * restates.  Arithmetic is plain FP64 in the reference's order of operations
 * where that order is visible in the source.  Periodic meshes use node indices
 * taken modulo (NZ2-1): the reference keeps aliased tail planes that are
 * summed into the head after every deposit and copied back
 * (para_field.f90:924-1061, 1075-1138), which is the same arithmetic.
 *
 * OpenMP: the particle loops run in parallel; the deposit uses per-thread
 * private grids reduced in thread order, so results do not depend on thread
 * timing (they do depend on the thread count, at the rounding level).
 *
 * How each part is pinned (tests/test_oracle.py, test_oracle_pins.py,
 * test_oracle_ref.py, test_lattice.py, test_adjoint_oracle.py):
 *   the reference's own compiled getP2 / getGamma and constants (oracle/_ref,
 *   built from utilities/gamma_to_p2.f90 and common/puffin_constants.f90):
 *     p2 and Gamma bit for bit, pi, c and q_e;
 *   the reference's own 1-D run (test/expected/f1main_expected_*.h5):
 *     p2, main-section b-fields, 1-D interpolation and deposit, equations,
 *     RK4, taper, matchIn/matchOut and correctTrans (elliptical), power,
 *     bunching, slice statistics and the binned current;
 *   closed forms, invariants and library routines: the FFT and k-vectors
 *   (numpy, brute-force DFT), propagator and absorption, masks, b-fields at
 *   special points, 3-D interpolation/deposit (partition of unity, adjoint
 *   identity, brute force), the lattice elements (drift additivity and its
 *   getGamma inverse, chicane slip and delay, quad inverse, modulation,
 *   corrector means, matching round trips and the plane-pole and curved
 *   entry offsets), slice bounds and the periodic wrap, the adjoint
 *   (central differences, dot-product identity);
 *   physics: the 3-D main-section and ramp b-fields of every undulator type
 *   (curl-free up to the reference's stated departures) and the bz coupling
 *   of the equations (natural-focusing betatron closed form).
 * Parity unpinned against reference outputs (none exist and the path's
 * reference cannot be built here): the 3-D integrator, diffraction and FFT
 * numbers; these rest on the closed forms and invariants above.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/puffin_gpu.h"

#define ORC_FLAG_TRANSVERSE 1
#define ORC_FLAG_BUFFER 2
#define ORC_FLAG_Z2MESH 4

/* ------------------------------------------------------------------------ */
/* p2 from Gamma and p_perp: getP2, gamma_to_p2.f90:84-85                    */
/* ------------------------------------------------------------------------ */
/* gamma0**2 * gamma**2 is (gamma0 gamma0)(gamma gamma), as Fortran forms
   it; the compiled reference routine agrees bit for bit (test_oracle_ref). */
static inline double orc_p2(double gam, double pr, double pi, double eta,
                            double gamma_r, double aw) {
    return ((1.0 / sqrt(1.0 - 1.0 / ((gamma_r * gamma_r) * (gam * gam)) *
                                  (1.0 + aw * aw * (pr * pr + pi * pi)))) -
            1.0) / eta;
}

double orc_get_p2(double gam, double pr, double pi, double eta, double gamma_r,
                  double aw) {
    return orc_p2(gam, pr, pi, eta, gamma_r, aw);
}

/* getGamma, gamma_to_p2.f90:39-41 (inverse of getP2; pinned by a round trip) */
double orc_get_gamma(double p2, double pr, double pi, double eta,
                     double gamma_r, double aw) {
    return sqrt((1.0 + (aw * aw * (pr * pr + pi * pi))) * pow(1.0 + eta * p2, 2.0) /
                (eta * p2 * (eta * p2 + 2.0))) / gamma_r;
}

/* ------------------------------------------------------------------------ */
/* Taper: getAlpha, wiggler_taper.f90:39-45 (with its n2col0 side effect)    */
/* ------------------------------------------------------------------------ */
/* Per-period taper table (no reference counterpart: the reference has only the
   linear taper).  When set, n2col at zbar is values[floor((zbar - z0) /
   period)], clamped to the table, in place of getAlpha. */
static double* g_tt = NULL;
static int g_tt_n = 0;
static double g_tt_z0 = 0.0, g_tt_period = 1.0;

void orc_set_taper_table(const double* values, int nvals, double z0, double period) {
    free(g_tt);
    g_tt = NULL;
    g_tt_n = 0;
    if (nvals > 0 && values) {
        g_tt = (double*)malloc(sizeof(double) * nvals);
        memcpy(g_tt, values, sizeof(double) * nvals);
        g_tt_n = nvals;
        g_tt_z0 = z0;
        g_tt_period = period;
    }
}

/* Table slot of zbar (clamped), or -1 without a table. */
int orc_taper_slot(double sz) {
    if (g_tt_n <= 0) return -1;
    long long k = (long long)floor((sz - g_tt_z0) / g_tt_period);
    if (k < 0) k = 0;
    if (k >= g_tt_n) k = g_tt_n - 1;
    return (int)k;
}

static double orc_alpha(const puff_params* p, double* taper /* n2col, n2col0, undgrad */,
                        double sz) {
    if (g_tt_n > 0) {
        taper[0] = g_tt[orc_taper_slot(sz)];
        return taper[0];
    }
    if (sz >= p->szfs && sz <= p->szfe) {
        taper[0] = taper[1] + taper[2] * (sz - p->szfs);
    } else if (sz > p->szfe) {
        taper[1] = taper[0];
    }
    return taper[0];
}

/* adjUndPlace, puffin_equations.f90:218-254.  0 = start, 1 = main, 2 = end;
   -1 for zbar < 0 (the reference STOPs). */
static int orc_und_place(const puff_params* p, double sz) {
    if (!p->und_ends) return 1;
    if (sz < 0) return -1;
    if (sz <= p->szfs) return 0;
    if (sz >= p->szfe) return 2;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Undulator b-fields at one particle: getBXfield / getBYfield / getBZfield, */
/* bfields.f90:59-309, 313-523, 528-881.                                      */
/* ------------------------------------------------------------------------ */
void orc_bfields(const puff_params* p, int place, double sz, double x, double y,
                 double* bx, double* by, double* bz) {
    const double rho = p->rho, pi = p->pi_ref, seta = sqrt(p->eta);
    double szt = sz / 2.0 / rho;
    double fx = p->fx, fy = p->fy;
    double vx = 0, vy = 0, vz = 0;
    int t = p->und_type;
    if (t == PUFF_UND_HELICAL) { fx = 1.0; fy = 1.0; t = PUFF_UND_ELLIPTICAL; }

    if (t == PUFF_UND_CURVED) {
        const double kx = p->kx_und, ky = p->ky_und;
        if (place == 0) {                                   /* bfields.f90:84-90,334-340,568-575 */
            vx = kx / ky * sinh(kx * x) * sinh(ky * y) * szt / 4.0 / pi * sin(szt);
            vy = cosh(kx * x) * cosh(ky * y) * szt / 4.0 / pi * sin(szt);
            vz = seta / 2.0 / rho / ky * cosh(kx * x) * sinh(ky * y) * (szt / 4.0 / pi) * cos(szt);
        } else if (place == 2) {                            /* :92-101,342-351,577-586 */
            double s = (sz - p->szfe) / 2.0 / rho;
            vx = kx / ky * sinh(kx * x) * sinh(ky * y) * (-s / 4.0 / pi + 1.0) * sin(s);
            vy = cosh(kx * x) * cosh(ky * y) * (-s / 4.0 / pi + 1.0) * sin(s);
            vz = seta / 2.0 / rho / kx * cosh(kx * x) * sinh(ky * y) * (-s / 4.0 / pi + 1.0) * cos(s);
        } else {                                            /* :103-109,353-359,588-595 */
            vx = kx / ky * sinh(kx * x) * sinh(ky * y) * sin(szt);
            vy = cosh(kx * x) * cosh(ky * y) * sin(szt);
            vz = seta / 2.0 / rho / kx * cosh(kx * x) * sinh(ky * y) * cos(szt);
        }
    } else if (t == PUFF_UND_PLANEPOLE) {
        double a = seta / 2.0 / rho * y;
        if (place == 0) {                                   /* :129-133,380-391,613-618 */
            vx = 0.0;
            vy = szt / 4.0 / pi * sin(szt);
            vz = sinh(a) * (szt / 4.0 / pi) * cos(szt);
        } else if (place == 2) {                            /* :135-142,395-407,620-628 */
            double s = (sz - p->szfe) / 2.0 / rho;
            vx = 0.0;
            vy = (-s / 4.0 / pi + 1.0) * sin(s);
            vz = sinh(a) * (-s / 4.0 / pi + 1.0) * cos(s);
        } else {                                            /* :144-148,409-415,630-635 */
            vx = 0.0;
            vy = sin(szt);
            vz = sinh(a) * cos(szt);
        }
    } else { /* elliptical (and helical with fx = fy = 1) */
        if (place == 0) {                                   /* :239-257,481-485,763-787 */
            vx = fx * (sz - pi * rho) / (6.0 * pi * rho) * cos(szt);
            vz = -seta / 2.0 / rho * (sz - pi * rho) / (6.0 * pi * rho) * fx * x * sin(szt);
            if (sz < pi * rho) {
                vx = 0.0;
                vz = 0.0;
            } else if (sz > 7.0 * pi * rho) {
                vx = fx * cos(szt);
                vz = -seta / 2.0 / rho * fx * x * sin(szt);
            }
            vy = fy * szt / 4.0 / pi * sin(szt);
            vz = vz + szt / 4.0 / pi * seta / 2.0 / rho * fy * y * cos(szt);
        } else if (place == 2) {                            /* :259-278,487-494,803-838 */
            double s = sz - p->szfe;
            vx = -fx * (s - 7.0 * pi * rho) / (6.0 * pi * rho) * cos(s / 2.0 / rho);
            vz = -seta / 2.0 / rho * (s - 7.0 * pi * rho) / (6.0 * pi * rho) * fx * x * sin(s / 2.0 / rho);
            if (s < pi * rho) {
                vx = fx * cos(s / 2.0 / rho);
                vz = -seta / 2.0 / rho * fx * x * sin(s / 2.0 / rho);
            } else if (s > 7.0 * pi * rho) {
                vx = 0.0;
                vz = 0.0;
            }
            double s2 = s / 2.0 / rho;
            vy = fy * (-s2 / 4.0 / pi + 1.0) * sin(s2);
            vz = vz + (-s / 8.0 / pi / rho + 1.0) * seta / 2.0 / rho * fy * y * cos(s / 2.0 / rho);
        } else {                                            /* :280-284,496-500,862-867 */
            vx = fx * cos(szt);
            vy = fy * sin(szt);
            vz = seta / 2.0 / rho * (-fx * x * sin(szt) + fy * y * cos(szt));
        }
    }
    if (p->focusing) {                                      /* :299-306, 514-521 */
        vx = seta * p->kbeta_y * p->kbeta_y / p->kappa * y + vx;
        vy = -seta * p->kbeta_x * p->kbeta_x / p->kappa * x + vy;
    }
    if (p->one_d) vz = 0.0;                                 /* :544-551 */
    *bx = vx;
    *by = vy;
    *bz = vz;
}

/* ------------------------------------------------------------------------ */
/* Interpolation weights and node of one particle.                          */
/* 3-D: getInterps_3D, system_interpolation.f90:45-91 with the flat node     */
/*      index of rhs.f90:178-182 (0-based here, whole mesh as the window,    */
/*      halfx = ((NX-1)/2) dx as rhs.f90:395-396 with nspinDX = NX).         */
/* 1-D: getInterps_1D, system_interpolation_1D.f90:41-59; rhs.f90:152.       */
/* Returns flag bits; node[8] and w[8] are filled when no flag is raised.     */
/* ------------------------------------------------------------------------ */
static int orc_interp(const puff_params* p, double x, double y, double z2,
                      int64_t* node, double* w) {
    const int64_t nx = p->nx, ny = p->ny, ntr = (int64_t)p->nx * p->ny;
    const int64_t nzp = p->periodic ? p->nz2 - 1 : p->nz2;
    int flag = 0;
    double zq = floor(z2 / p->dz2);
    double locz2 = z2 - zq * p->dz2;      /* z2node-1 = zq */
    double z2_in2 = locz2 / p->dz2;
    double z2_in1 = 1.0 - z2_in2;
    int64_t iz0, iz1;
    if (p->periodic) {
        if (!(zq == zq) || fabs(zq) > 4.0e15) return ORC_FLAG_Z2MESH;
        int64_t q = (int64_t)zq;
        iz0 = ((q % nzp) + nzp) % nzp;
        iz1 = (iz0 + 1) % nzp;
    } else {
        /* temporal mesh: z2node >= NZ2 STOPs (system_interpolation.f90:70-75);
           z2node < 1 would index before the mesh and is refused as well. */
        if (!(zq >= 0.0 && zq <= (double)(nzp - 2))) return ORC_FLAG_Z2MESH;
        iz0 = (int64_t)zq;
        iz1 = iz0 + 1;
    }
    if (p->one_d) {
        /* 1-D weights: lis_GR(1) = 1 - locz2/dz2, lis_GR(2) = 1 - lis_GR(1) */
        w[0] = 1.0 - locz2 / p->dz2;
        w[1] = 1.0 - w[0];
        node[0] = iz0;
        node[1] = iz1;
        return 0;
    }
    double halfx = ((p->nx - 1) / 2.0) * p->dx;
    double halfy = ((p->ny - 1) / 2.0) * p->dy;
    double xq = floor((x + halfx) / p->dx);
    double yq = floor((y + halfy) / p->dy);
    /* xnode >= nspinDX .or. xnode < 1 raises both flags (:60-68) */
    if (!(xq >= 0.0 && xq <= (double)(nx - 2))) flag |= ORC_FLAG_TRANSVERSE;
    if (!(yq >= 0.0 && yq <= (double)(ny - 2))) flag |= ORC_FLAG_TRANSVERSE;
    if (flag) return flag;
    double locx = x + halfx - xq * p->dx;
    double x_in2 = locx / p->dx, x_in1 = 1.0 - x_in2;
    double locy = y + halfy - yq * p->dy;
    double y_in2 = locy / p->dy, y_in1 = 1.0 - y_in2;
    int64_t ix = (int64_t)xq, iy = (int64_t)yq;
    int64_t b0 = iz0 * ntr + iy * nx + ix;
    int64_t b1 = iz1 * ntr + iy * nx + ix;
    w[0] = x_in1 * y_in1 * z2_in1;
    w[1] = x_in2 * y_in1 * z2_in1;
    w[2] = x_in1 * y_in2 * z2_in1;
    w[3] = x_in2 * y_in2 * z2_in1;
    w[4] = x_in1 * y_in1 * z2_in2;
    w[5] = x_in2 * y_in1 * z2_in2;
    w[6] = x_in1 * y_in2 * z2_in2;
    w[7] = x_in2 * y_in2 * z2_in2;
    node[0] = b0;
    node[1] = b0 + 1;
    node[2] = b0 + nx;
    node[3] = b0 + nx + 1;
    node[4] = b1;
    node[5] = b1 + 1;
    node[6] = b1 + nx;
    node[7] = b1 + nx + 1;
    return 0;
}

/* Exposed for tests: weights / nodes of one particle. */
int orc_interp_one(const puff_params* p, double x, double y, double z2,
                   int64_t* node, double* w) {
    return orc_interp(p, x, y, z2, node, w);
}

/* Deposit mode of the threaded oracle: 0 = per-thread private grids reduced
   in thread order (deterministic, the checker's default); 1 = shared grid
   with atomic adds, as the reference's !$OMP ATOMIC scatter
   (system_interpolation.f90:174-241), used when timing the CPU baseline. */
static int g_deposit_atomic = 0;
void orc_set_deposit_atomic(int on) { g_deposit_atomic = on; }

typedef struct {
    const puff_params* p;
    double* taper;
    int64_t n;
    int nthreads;
    int64_t nmesh;
    /* private deposit grids, nthreads x (2 * nmesh) */
    double* priv;
    int64_t* nodes;   /* n x 8 */
    double* wts;      /* n x 8 */
    double* p2;       /* n */
} orc_work;

/* ------------------------------------------------------------------------ */
/* One derivative evaluation: derivs / getrhs, derivative.f90:36-155,        */
/* rhs.f90:60-330.  Inputs: field (ar, ai), trial particles; outputs: the six */
/* particle derivatives and the field source dA (overwritten, not summed:    */
/* rk4par always passes zeroed dadz arrays, puffin_mpi_RK4.f90:95-100,284).  */
/* Returns flag bits.                                                        */
/* ------------------------------------------------------------------------ */
static int orc_derivs(orc_work* wk, double sz, const double* ar, const double* ai,
                      const double* sx, const double* sy, const double* sz2,
                      const double* spr, const double* spi, const double* sgam,
                      const double* chi, double* dx, double* dy, double* dz2,
                      double* dpr, double* dpi, double* dgam, double* dar,
                      double* dai) {
    const puff_params* p = wk->p;
    const int64_t n = wk->n, nmesh = wk->nmesh;
    const int nnb = p->one_d ? 2 : 8;
    const double dV3 = p->dx * p->dy * p->dz2;        /* rhs.f90:361 */
    const double sInv2rho = 1.0 / (2.0 * p->rho);     /* rhs.f90:366 */
    const double n2col = orc_alpha(p, wk->taper, sz); /* rhs.f90:130 */
    const int place = orc_und_place(p, sz);           /* rhs.f90:131 */
    int flags = 0;

    /* getP2 and getInterps over all particles first (rhs.f90:143,202) */
#pragma omp parallel for reduction(| : flags) num_threads(wk->nthreads) schedule(static)
    for (int64_t i = 0; i < n; i++) {
        wk->p2[i] = orc_p2(sgam[i], spr[i], spi[i], p->eta, p->gamma_r, p->aw);
        flags |= orc_interp(p, sx[i], sy[i], sz2[i], wk->nodes + 8 * i, wk->wts + 8 * i);
    }
    if (place < 0) flags |= ORC_FLAG_Z2MESH;

#pragma omp parallel for num_threads(wk->nthreads) schedule(static)
    for (int64_t j = 0; j < nmesh; j++) {
        dar[j] = 0.0;
        dai[j] = 0.0;
    }
    const int do_fields = (flags == 0);
    const int atomic_dep = g_deposit_atomic && wk->nthreads > 1;
    const int private_dep = wk->nthreads > 1 && !atomic_dep;

    /* getFFelecs + getSource (+ equations) */
#pragma omp parallel num_threads(wk->nthreads)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double* pr_ = wk->priv + (int64_t)t * 2 * nmesh;
        double* pi_ = pr_ + nmesh;
        if (do_fields && private_dep) {
            memset(pr_, 0, sizeof(double) * 2 * nmesh);
        }
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            double fr = 0.0, fi = 0.0;
            const double p2 = wk->p2[i];
            const int64_t* nd = wk->nodes + 8 * i;
            const double* w = wk->wts + 8 * i;
            if (do_fields) {
                /* getFFelecs_3D: F = sum_k w_k A(node_k) (system_interpolation.f90:121-137) */
                for (int k = 0; k < nnb; k++) {
                    fr = w[k] * ar[nd[k]] + fr;
                    fi = w[k] * ai[nd[k]] + fi;
                }
                /* getSource_3D: s = chi/dV3 (1 + eta p2) p / Gamma (:171-172, :208-209) */
                double sr = ((chi[i] / dV3) * (1 + p->eta * p2) * spr[i] / sgam[i]);
                double si = ((chi[i] / dV3) * (1 + p->eta * p2) * spi[i] / sgam[i]);
                if (atomic_dep) {
                    for (int k = 0; k < nnb; k++) {
#pragma omp atomic
                        dar[nd[k]] += w[k] * sr;
#pragma omp atomic
                        dai[nd[k]] += w[k] * si;
                    }
                } else {
                    double* tr = private_dep ? pr_ : dar;
                    double* ti = private_dep ? pi_ : dai;
                    for (int k = 0; k < nnb; k++) {
                        tr[nd[k]] = w[k] * sr + tr[nd[k]];
                        ti[nd[k]] = w[k] * si + ti[nd[k]];
                    }
                }
            }
            if (!p->coupling) { fr = 0.0; fi = 0.0; }     /* rhs.f90:221-224 */
            if (p->electrons_evolve) {
                double bx, by, bz;
                orc_bfields(p, place, sz, sx[i], sy[i], &bx, &by, &bz);
                const double eta = p->eta, kap = p->kappa, g = sgam[i];
                /* puffin_equations.f90:172, 115-117, 144-146, 31-35, 61-65, 86-87 */
                dz2[i] = p2;
                dx[i] = 2 * p->rho * kap / sqrt(eta) * (1 + eta * p2) / g * spr[i];
                dy[i] = -2 * p->rho * kap / sqrt(eta) * (1 + eta * p2) / g * spi[i];
                dpr[i] = sInv2rho * (n2col * by - eta * p2 / (kap * kap) * fr) +
                         kap * spi[i] / g * (1 + eta * p2) * n2col * bz;
                dpi[i] = sInv2rho * (n2col * bx - eta * p2 / (kap * kap) * fi) -
                         kap * spr[i] / g * (1 + eta * p2) * n2col * bz;
                dgam[i] = -p->rho * (1 + eta * p2) / g * 2.0 * (spr[i] * fr + spi[i] * fi);
            } else {                                      /* rhs.f90:301-308 */
                dz2[i] = dx[i] = dy[i] = dpr[i] = dpi[i] = dgam[i] = 0.0;
            }
        }
    }
    if (do_fields && private_dep) {
        /* reduce private grids in thread order (deterministic for a given count) */
#pragma omp parallel for num_threads(wk->nthreads) schedule(static)
        for (int64_t j = 0; j < nmesh; j++) {
            double sr = 0.0, si = 0.0;
            for (int t = 0; t < wk->nthreads; t++) {
                sr += wk->priv[(int64_t)t * 2 * nmesh + j];
                si += wk->priv[(int64_t)t * 2 * nmesh + nmesh + j];
            }
            dar[j] = sr;
            dai[j] = si;
        }
    }
    if (!p->field_evolve) {                               /* rhs.f90:294-297 */
        memset(dar, 0, sizeof(double) * nmesh);
        memset(dai, 0, sizeof(double) * nmesh);
    }
    return flags;
}

/* Periodic mesh: keep the aliased last plane equal to the first
   (redistbackFFT tail copy, para_field.f90:3197-3244). */
static void orc_periodic_alias(const puff_params* p, double* ar, double* ai) {
    if (!p->periodic) return;
    int64_t ntr = (int64_t)p->nx * p->ny;
    memcpy(ar + (int64_t)(p->nz2 - 1) * ntr, ar, sizeof(double) * ntr);
    memcpy(ai + (int64_t)(p->nz2 - 1) * ntr, ai, sizeof(double) * ntr);
}

static int orc_status(int flags) {
    if (flags & ORC_FLAG_Z2MESH) return PUFF_OUT_Z2_MESH;
    if (flags & ORC_FLAG_TRANSVERSE) return PUFF_OUT_TRANSVERSE;
    if (flags & ORC_FLAG_BUFFER) return PUFF_OUT_Z2_BUFFER;
    return PUFF_OK;
}

/* ------------------------------------------------------------------------ */
/* rk4par, puffin_mpi_RK4.f90:38-366.                                        */
/* taper = {n2col, n2col0, undgrad}, updated in place.                       */
/* ------------------------------------------------------------------------ */
int orc_rk4par(const puff_params* p, double* taper, int64_t n, double* x,
               double* y, double* z2, double* pr, double* pi, double* gam,
               const double* chi, double* ar, double* ai, double sz, double h,
               int nthreads) {
    if (nthreads < 1) nthreads = 1;
    const int64_t nmesh = (int64_t)p->nx * p->ny * p->nz2;
    orc_work wk;
    wk.p = p;
    wk.taper = taper;
    wk.n = n;
    wk.nthreads = nthreads;
    wk.nmesh = nmesh;
    wk.priv = (nthreads > 1 && !g_deposit_atomic) ? (double*)malloc(sizeof(double) * 2 * nmesh * nthreads) : NULL;
    wk.nodes = (int64_t*)malloc(sizeof(int64_t) * 8 * (n > 0 ? n : 1));
    wk.wts = (double*)malloc(sizeof(double) * 8 * (n > 0 ? n : 1));
    wk.p2 = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));

    /* work arrays of allact_rk4_arrs (puffin_mpi_RK4.f90:380-412) */
    double* buf = (double*)calloc((size_t)(24 * n + 1), sizeof(double));
    double *dxdx = buf, *dydx = buf + n, *dz2dx = buf + 2 * n, *dpxdx = buf + 3 * n,
           *dpydx = buf + 4 * n, *dpz2dx = buf + 5 * n;
    double *dxt = buf + 6 * n, *dyt = buf + 7 * n, *dz2t = buf + 8 * n,
           *dpxt = buf + 9 * n, *dpyt = buf + 10 * n, *dpz2t = buf + 11 * n;
    double *dxm = buf + 12 * n, *dym = buf + 13 * n, *dz2m = buf + 14 * n,
           *dpxm = buf + 15 * n, *dpym = buf + 16 * n, *dpz2m = buf + 17 * n;
    double *xt = buf + 18 * n, *yt = buf + 19 * n, *z2t = buf + 20 * n,
           *pxt = buf + 21 * n, *pyt = buf + 22 * n, *pz2t = buf + 23 * n;
    double* fbuf = (double*)calloc((size_t)(10 * nmesh), sizeof(double));
    double *dar0 = fbuf, *dai0 = fbuf + nmesh, *dar1 = fbuf + 2 * nmesh,
           *dai1 = fbuf + 3 * nmesh, *dar2 = fbuf + 4 * nmesh, *dai2 = fbuf + 5 * nmesh;
    double *atr = fbuf + 6 * nmesh, *ati = fbuf + 7 * nmesh;
    double *dar3 = fbuf + 8 * nmesh, *dai3 = fbuf + 9 * nmesh;

    const double hh = h * 0.5, h6 = h / 6.0;
    double szh = sz + hh;
    int flags = 0, status = PUFF_OK;

    /* first stage: derivs at (sz, A0, y0) (:174-177) */
    flags = orc_derivs(&wk, sz, ar, ai, x, y, z2, pr, pi, gam, chi, dxdx, dydx,
                       dz2dx, dpxdx, dpydx, dpz2dx, dar0, dai0);
    if (flags) goto done;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; i++) {                  /* :189-198 */
        xt[i] = x[i] + hh * dxdx[i];
        yt[i] = y[i] + hh * dydx[i];
        z2t[i] = z2[i] + hh * dz2dx[i];
        pxt[i] = pr[i] + hh * dpxdx[i];
        pyt[i] = pi[i] + hh * dpydx[i];
        pz2t[i] = gam[i] + hh * dpz2dx[i];
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t j = 0; j < nmesh; j++) {
        atr[j] = ar[j] + hh * dar0[j];
        ati[j] = ai[j] + hh * dai0[j];
    }
    /* second stage (:213-217) */
    flags = orc_derivs(&wk, szh, atr, ati, xt, yt, z2t, pxt, pyt, pz2t, chi, dxt,
                       dyt, dz2t, dpxt, dpyt, dpz2t, dar1, dai1);
    if (flags) goto done;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; i++) {                  /* :226-235 */
        xt[i] = x[i] + hh * dxt[i];
        yt[i] = y[i] + hh * dyt[i];
        z2t[i] = z2[i] + hh * dz2t[i];
        pxt[i] = pr[i] + hh * dpxt[i];
        pyt[i] = pi[i] + hh * dpyt[i];
        pz2t[i] = gam[i] + hh * dpz2t[i];
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t j = 0; j < nmesh; j++) {
        atr[j] = ar[j] + hh * dar1[j];
        ati[j] = ai[j] + hh * dai1[j];
    }
    /* third stage (:249-253) */
    flags = orc_derivs(&wk, szh, atr, ati, xt, yt, z2t, pxt, pyt, pz2t, chi, dxm,
                       dym, dz2m, dpxm, dpym, dpz2m, dar2, dai2);
    if (flags) goto done;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; i++) {                  /* :258-279 */
        xt[i] = x[i] + h * dxm[i];
        yt[i] = y[i] + h * dym[i];
        z2t[i] = z2[i] + h * dz2m[i];
        pxt[i] = pr[i] + h * dpxm[i];
        pyt[i] = pi[i] + h * dpym[i];
        pz2t[i] = gam[i] + h * dpz2m[i];
        dxm[i] = dxt[i] + dxm[i];
        dym[i] = dyt[i] + dym[i];
        dz2m[i] = dz2t[i] + dz2m[i];
        dpxm[i] = dpxt[i] + dpxm[i];
        dpym[i] = dpyt[i] + dpym[i];
        dpz2m[i] = dpz2t[i] + dpz2m[i];
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t j = 0; j < nmesh; j++) {              /* :266-285 */
        atr[j] = ar[j] + h * dar2[j];
        ati[j] = ai[j] + h * dai2[j];
        dar2[j] = dar1[j] + dar2[j];
        dai2[j] = dai1[j] + dai2[j];
    }
    /* fourth stage at sz + h (:291-300) */
    szh = sz + h;
    flags = orc_derivs(&wk, szh, atr, ati, xt, yt, z2t, pxt, pyt, pz2t, chi, dxt,
                       dyt, dz2t, dpxt, dpyt, dpz2t, dar3, dai3);
    if (flags) goto done;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; i++) {                  /* :307-312 */
        x[i] = x[i] + h6 * (dxdx[i] + dxt[i] + 2.0 * dxm[i]);
        y[i] = y[i] + h6 * (dydx[i] + dyt[i] + 2.0 * dym[i]);
        z2[i] = z2[i] + h6 * (dz2dx[i] + dz2t[i] + 2.0 * dz2m[i]);
        pr[i] = pr[i] + h6 * (dpxdx[i] + dpxt[i] + 2.0 * dpxm[i]);
        pi[i] = pi[i] + h6 * (dpydx[i] + dpyt[i] + 2.0 * dpym[i]);
        gam[i] = gam[i] + h6 * (dpz2dx[i] + dpz2t[i] + 2.0 * dpz2m[i]);
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t j = 0; j < nmesh; j++) {              /* :314-315 */
        ar[j] = ar[j] + h6 * (dar0[j] + dar3[j] + 2.0 * dar2[j]);
        ai[j] = ai[j] + h6 * (dai0[j] + dai3[j] + 2.0 * dai2[j]);
    }
    orc_periodic_alias(p, ar, ai);

done:
    status = orc_status(flags);
    free(wk.priv);
    free(wk.nodes);
    free(wk.wts);
    free(wk.p2);
    free(buf);
    free(fbuf);
    return status;
}

/* Source-only evaluation (tests): dA of one getSource_3D call for the given
   particles and the interpolated field F at each particle. */
int orc_interp_deposit(const puff_params* p, int64_t n, const double* x,
                       const double* y, const double* z2, const double* pr,
                       const double* pi, const double* gam, const double* chi,
                       const double* ar, const double* ai, double* fr,
                       double* fi, double* dar, double* dai) {
    const int64_t nmesh = (int64_t)p->nx * p->ny * p->nz2;
    const int nnb = p->one_d ? 2 : 8;
    const double dV3 = p->dx * p->dy * p->dz2;
    int flags = 0;
    memset(dar, 0, sizeof(double) * nmesh);
    memset(dai, 0, sizeof(double) * nmesh);
    for (int64_t i = 0; i < n; i++) {
        int64_t nd[8];
        double w[8];
        int f = orc_interp(p, x[i], y[i], z2[i], nd, w);
        flags |= f;
        fr[i] = fi[i] = 0.0;
        if (f) continue;
        double p2 = orc_p2(gam[i], pr[i], pi[i], p->eta, p->gamma_r, p->aw);
        double sr = ((chi[i] / dV3) * (1 + p->eta * p2) * pr[i] / gam[i]);
        double si = ((chi[i] / dV3) * (1 + p->eta * p2) * pi[i] / gam[i]);
        for (int k = 0; k < nnb; k++) {
            fr[i] = w[k] * ar[nd[k]] + fr[i];
            fi[i] = w[k] * ai[nd[k]] + fi[i];
            dar[nd[k]] += w[k] * sr;
            dai[nd[k]] += w[k] * si;
        }
    }
    return orc_status(flags);
}

/* Periodic wrap of z2 (para_field.f90:212-215). */
void orc_wrap_z2(const puff_params* p, int64_t n, double* z2) {
    if (!p->periodic) return;
    double lenz2 = p->dz2 * (double)(p->nz2 - 1);
    for (int64_t i = 0; i < n; i++) z2[i] = z2[i] - (floor(z2[i] / lenz2) * lenz2);
}

/* ------------------------------------------------------------------------ */
/* Mixed-radix complex DFT (FFTW convention: sign -1 forward, +1 backward,    */
/* unnormalised), standing in for fftw_mpi_plan_dft_3d / execute             */
/* (transforms.f90:186-211, 423).                                            */
/* ------------------------------------------------------------------------ */
static int orc_smallest_factor(int n) {
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int f = 3; (int64_t)f * f <= n; f += 2)
        if (n % f == 0) return f;
    return n;
}

/* Twiddle table of a top-level length N: tw[e] = exp(sign 2 pi i e / N) for
   e = 0..N-1.  A sub-transform of length n = N / twst reads its factor
   exp(sign 2 pi i e / n) at index e * twst: e twst / N and e / n are the same
   rational, so the table holds the very doubles a direct cos / sin of
   sign 2 pi (e / n) gives. */
static double* orc_twiddles(int n, int sign) {
    double* tw = (double*)malloc(sizeof(double) * 2 * n);
    for (int e = 0; e < n; e++) {
        double a = sign * 2.0 * M_PI * (double)e / (double)n;
        tw[2 * e] = cos(a);
        tw[2 * e + 1] = sin(a);
    }
    return tw;
}

/* y[k] = sum_j x[j*xs] exp(sign 2 pi i j k / n); y contiguous.  tw is the
   table of length n * twst (orc_twiddles). */
static void orc_fft_rec(int n, const double* x, int64_t xs, double* y, double* tmp,
                        const double* tw, int64_t twst) {
    if (n == 1) {
        y[0] = x[0];
        y[1] = x[1];
        return;
    }
    int r = orc_smallest_factor(n);
    int m = n / r;
    if (m == 1) {
        for (int k = 0; k < n; k++) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < n; j++) {
                int64_t e = ((int64_t)j * k) % n;
                double c = tw[2 * e * twst], s = tw[2 * e * twst + 1];
                double xr = x[2 * j * xs], xi = x[2 * j * xs + 1];
                sr += xr * c - xi * s;
                si += xr * s + xi * c;
            }
            y[2 * k] = sr;
            y[2 * k + 1] = si;
        }
        return;
    }
    /* decimation in time: r sub-transforms of length m */
    for (int q = 0; q < r; q++)
        orc_fft_rec(m, x + 2 * q * xs, xs * r, y + 2 * q * m, tmp, tw, twst * r);
    for (int k = 0; k < m; k++) {
        for (int s = 0; s < r; s++) {
            int64_t kk = k + (int64_t)s * m;
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < r; q++) {
                int64_t e = ((int64_t)q * kk) % n;
                double c = tw[2 * e * twst], sn = tw[2 * e * twst + 1];
                double yr = y[2 * (q * m + k)], yi = y[2 * (q * m + k) + 1];
                sr += yr * c - yi * sn;
                si += yr * sn + yi * c;
            }
            tmp[2 * kk] = sr;
            tmp[2 * kk + 1] = si;
        }
    }
    memcpy(y, tmp, sizeof(double) * 2 * n);
}

/* 1-D transform of n complex values (interleaved) in place. */
void orc_fft1d(double* data, int n, int sign) {
    double* y = (double*)malloc(sizeof(double) * 2 * n);
    double* in = (double*)malloc(sizeof(double) * 2 * n);
    double* tmp = (double*)malloc(sizeof(double) * 2 * n);
    double* tw = orc_twiddles(n, sign);
    memcpy(in, data, sizeof(double) * 2 * n);
    orc_fft_rec(n, in, 1, y, tmp, tw, 1);
    memcpy(data, y, sizeof(double) * 2 * n);
    free(y);
    free(in);
    free(tmp);
    free(tw);
}

/* 3-D transform of a (n3, n2, n1) row-major complex array (n1 fastest). */
void orc_fft3d(double* a, int n1, int n2, int n3, int sign, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    int nmax = n1 > n2 ? n1 : n2;
    if (n3 > nmax) nmax = n3;
    double* tw1 = orc_twiddles(n1, sign);
    double* tw2 = orc_twiddles(n2, sign);
    double* tw3 = orc_twiddles(n3, sign);
#pragma omp parallel num_threads(nthreads)
    {
        double* line = (double*)malloc(sizeof(double) * 2 * nmax);
        double* y = (double*)malloc(sizeof(double) * 2 * nmax);
        double* tmp = (double*)malloc(sizeof(double) * 2 * nmax);
        /* axis 1 (x) */
#pragma omp for schedule(static)
        for (int64_t l = 0; l < (int64_t)n2 * n3; l++) {
            double* base = a + 2 * l * n1;
            orc_fft_rec(n1, base, 1, y, tmp, tw1, 1);
            memcpy(base, y, sizeof(double) * 2 * n1);
        }
        /* axis 2 (y) */
#pragma omp for schedule(static)
        for (int64_t l = 0; l < (int64_t)n1 * n3; l++) {
            int64_t i1 = l % n1, i3 = l / n1;
            double* base = a + 2 * (i3 * (int64_t)n2 * n1 + i1);
            for (int j = 0; j < n2; j++) {
                line[2 * j] = base[2 * (int64_t)j * n1];
                line[2 * j + 1] = base[2 * (int64_t)j * n1 + 1];
            }
            orc_fft_rec(n2, line, 1, y, tmp, tw2, 1);
            for (int j = 0; j < n2; j++) {
                base[2 * (int64_t)j * n1] = y[2 * j];
                base[2 * (int64_t)j * n1 + 1] = y[2 * j + 1];
            }
        }
        /* axis 3 (z2) */
#pragma omp for schedule(static)
        for (int64_t l = 0; l < (int64_t)n1 * n2; l++) {
            double* base = a + 2 * l;
            int64_t st = (int64_t)n1 * n2;
            for (int j = 0; j < n3; j++) {
                line[2 * j] = base[2 * j * st];
                line[2 * j + 1] = base[2 * j * st + 1];
            }
            orc_fft_rec(n3, line, 1, y, tmp, tw3, 1);
            for (int j = 0; j < n3; j++) {
                base[2 * j * st] = y[2 * j];
                base[2 * j * st + 1] = y[2 * j + 1];
            }
        }
        free(line);
        free(y);
        free(tmp);
    }
    free(tw1);
    free(tw2);
    free(tw3);
}

/* ------------------------------------------------------------------------ */
/* k-vectors: GetKValues, transforms.f90:494-573 (exact pi = 4 atan(1)).     */
/* ------------------------------------------------------------------------ */
void orc_kvalues(const puff_params* p, double* kx, double* ky, double* kz2) {
    const double pi = 4.0 * atan(1.0);
    int nzp = p->periodic ? p->nz2 - 1 : p->nz2;
    double lx = (double)p->nx * p->dx, ly = (double)p->ny * p->dy;
    double lz = (double)nzp * p->dz2;
    int maxx = (int)ceil((double)p->nx / 2.0), maxy = (int)ceil((double)p->ny / 2.0);
    int maxz = (int)ceil((double)nzp / 2.0);
    for (int i = 0; i < p->nx; i++) kx[i] = 2.0 * pi * (double)(i >= maxx ? i - p->nx : i) / lx;
    for (int i = 0; i < p->ny; i++) ky[i] = 2.0 * pi * (double)(i >= maxy ? i - p->ny : i) / ly;
    for (int i = 0; i < nzp; i++) kz2[i] = 2.0 * pi * (double)(i >= maxz ? i - nzp : i) / lz;
}

/* ------------------------------------------------------------------------ */
/* Absorbing masks, masks.f90.                                              */
/* ------------------------------------------------------------------------ */
/* linspace, base_functions.f90:67-97 */
static void orc_linspace(double a, double b, int n, double* out) {
    if (n > 1) {
        double d = (b - a) / (double)(n - 1);
        for (int i = 0; i < n; i++) out[i] = a + i * d;
    } else if (n == 1) {
        out[0] = a;
    }
}

/* Mask1D2, masks.f90:266-302 (1-based Fortran slices restated 0-based) */
static void orc_mask1d2(const double* x, double dx, int nx, int nb, double pi,
                        double* m) {
    double lb = (nb - 1) * dx;
    double x01 = x[nx - 1], x02 = x[0];
    for (int i = nx - nb; i < nx; i++) m[i] = pow(cos(pi * (x[i] - x01) / (2.0 * lb)), 2.0);
    for (int i = 0; i < nb; i++) m[i] = pow(cos(pi * (x[i] - x02) / (2.0 * lb)), 2.0);
    for (int i = nb; i < nx - nb; i++) m[i] = 0.0;
}

/* Maskf4, masks.f90:329-361 */
static void orc_maskf4(const double* x, double dx, int nx, int nb, double pi,
                       double* m) {
    double lb = (nb - 1) * dx;
    double x01 = x[nx - nb], x02 = x[nb - 1];
    for (int i = nx - nb; i < nx; i++) m[i] = pow(cos(pi * (x[i] - x01) / (2.0 * lb)), 2.0);
    for (int i = 0; i < nb; i++) m[i] = pow(cos(pi * (x[i] - x02) / (2.0 * lb)), 2.0);
    for (int i = nb; i < nx - nb; i++) m[i] = 1.0;
}

/* getMask + Mask2D, masks.f90:50-87, 136-189: mask(ix,iy) =
   Maskf3_y(iy) Mask1D2_x(ix) + Maskf4_x(ix) Mask1D2_y(iy), Maskf3 = 1. */
void orc_mask_xy(const puff_params* p, double* mask) {
    int nx = p->nx, ny = p->ny;
    /* a boundary band wider than the mesh would index outside the
       reference's arrays: no mask then (as the HIP side) */
    if (!(nx > 1 && ny > 1 && p->nbx > 1 && p->nby > 1 && nx >= p->nbx && ny >= p->nby)) {
        for (int64_t j = 0; j < (int64_t)nx * ny; j++) mask[j] = 0.0;
        return;
    }
    double* x = (double*)malloc(sizeof(double) * nx);
    double* y = (double*)malloc(sizeof(double) * ny);
    double* mx = (double*)malloc(sizeof(double) * nx);
    double* my = (double*)malloc(sizeof(double) * ny);
    double* mx2 = (double*)malloc(sizeof(double) * nx);
    double lenx = (nx - 1) * p->dx, leny = (ny - 1) * p->dy;
    orc_linspace(-lenx / 2, lenx / 2, nx, x);
    orc_linspace(-leny / 2, leny / 2, ny, y);
    orc_mask1d2(x, p->dx, nx, p->nbx, p->pi_ref, mx);
    orc_mask1d2(y, p->dy, ny, p->nby, p->pi_ref, my);
    orc_maskf4(x, p->dx, nx, p->nbx, p->pi_ref, mx2);
    for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++) mask[iy * nx + ix] = 1.0 * mx[ix] + mx2[ix] * my[iy];
    free(x);
    free(y);
    free(mx);
    free(my);
    free(mx2);
}

/* getZ2Mask, masks.f90:366-430, including its Lf = (nx - 1*dx) expression
   (evaluates to nx - dx); zero on a periodic mesh (diffraction.f90:447). */
void orc_mask_z2(const puff_params* p, double* mz) {
    int nzp = p->periodic ? p->nz2 - 1 : p->nz2;
    if (p->periodic || p->nbz2 <= 1) {   /* nbz2 = 1 would divide by lb = 0 */
        for (int i = 0; i < nzp; i++) mz[i] = 0.0;
        return;
    }
    double dx = p->dz2;
    int nx = p->nz2, nb = p->nbz2, loc_start = 0, loc_nx = nzp;
    double* x = (double*)malloc(sizeof(double) * loc_nx);
    orc_linspace((double)loc_start * dx, (double)(loc_start + loc_nx - 1) * dx, loc_nx, x);
    double lb = (nb - 1) * dx;
    double lf = (nx - 1 * dx);
    double x0 = lf - lb / 2.0;
    int fbn = nx - nb + 1;
    for (int ind = 1; ind <= loc_nx; ind++) {
        mz[ind - 1] = 0.0;
        if (loc_start + ind >= fbn) mz[ind - 1] = pow(cos(p->pi_ref * (x[ind - 1] - x0) / lb), 2.0);
    }
    free(x);
}

/* ------------------------------------------------------------------------ */
/* DiffractionStep, diffraction.f90:209-348, with multiplyexp (:111-184) and */
/* AbsorptionStep (:367-523).  ar / ai: full mesh, nz2 planes.               */
/* ------------------------------------------------------------------------ */
int orc_diffraction_step(const puff_params* p, double* ar, double* ai, double h,
                         int nthreads) {
    const int nx = p->nx, ny = p->ny;
    const int nzp = p->periodic ? p->nz2 - 1 : p->nz2;
    const int64_t ntr = (int64_t)nx * ny, nn = ntr * nzp;
    const double ffact = (double)nx * (double)ny * (double)nzp;   /* setup.f90:334-346 */
    double* a = (double*)malloc(sizeof(double) * 2 * nn);
    double* anb = (double*)malloc(sizeof(double) * 2 * nn);
    double* kx = (double*)malloc(sizeof(double) * nx);
    double* ky = (double*)malloc(sizeof(double) * ny);
    double* kz = (double*)malloc(sizeof(double) * nzp);
    double* mxy = (double*)malloc(sizeof(double) * ntr);
    double* mz = (double*)malloc(sizeof(double) * nzp);
    orc_kvalues(p, kx, ky, kz);
    orc_mask_xy(p, mxy);
    orc_mask_z2(p, mz);

    for (int64_t j = 0; j < nn; j++) {                   /* pack, :255-264 */
        a[2 * j] = ar[j];
        a[2 * j + 1] = ai[j];
    }
    orc_fft3d(a, nx, ny, nzp, -1, nthreads);              /* :272 */
    const double cutoff = p->filt_cutoff;                /* multiplyexp :138-139 */
    for (int iz = 0; iz < nzp; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int64_t j = ((int64_t)iz * ny + iy) * nx + ix;
                if ((kz[iz] > cutoff) || (kz[iz] < -cutoff)) {
                    if (kz[iz] != 0.0) {
                        double th = h * (kx[ix] * kx[ix] + ky[iy] * ky[iy]) / (2.0 * kz[iz]);
                        double c = cos(th), s = sin(th);
                        double re = a[2 * j], im = a[2 * j + 1];
                        a[2 * j] = c * re - s * im;
                        a[2 * j + 1] = s * re + c * im;
                    }
                } else if (p->filter) {
                    a[2 * j] = 0.0;
                    a[2 * j + 1] = 0.0;
                }
            }
    orc_fft3d(a, nx, ny, nzp, +1, nthreads);              /* :297 */
    for (int64_t j = 0; j < 2 * nn; j++) a[j] = a[j] / ffact; /* :310 */

    /* AbsorptionStep: split (:454-470) */
    for (int iz = 0; iz < nzp; iz++)
        for (int64_t t = 0; t < ntr; t++) {
            int64_t j = (int64_t)iz * ntr + t;
            double m = mxy[t] + (mz[iz] * (1.0 - mxy[t]));
            double keep = (1.0 - m);
            anb[2 * j] = keep * a[2 * j];
            anb[2 * j + 1] = keep * a[2 * j + 1];
            a[2 * j] = m * a[2 * j];
            a[2 * j + 1] = m * a[2 * j + 1];
        }
    orc_fft3d(a, nx, ny, nzp, -1, nthreads);              /* :476 */
    for (int iz = 0; iz < nzp; iz++)                     /* :480-501 */
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                if (kz[iz] != 0.0) {
                    int64_t j = ((int64_t)iz * ny + iy) * nx + ix;
                    double f = exp(-h * p->beta * (fabs(kx[ix]) + fabs(ky[iy])) /
                                   (sqrt(fabs(2.0 * kz[iz]))));
                    a[2 * j] *= f;
                    a[2 * j + 1] *= f;
                }
            }
    orc_fft3d(a, nx, ny, nzp, +1, nthreads);              /* :505 */
    for (int64_t j = 0; j < 2 * nn; j++) a[j] = a[j] / ffact + anb[j]; /* :513-517 */

    for (int64_t j = 0; j < nn; j++) {                   /* unpack, :327-336 */
        ar[j] = a[2 * j];
        ai[j] = a[2 * j + 1];
    }
    orc_periodic_alias(p, ar, ai);
    free(a);
    free(anb);
    free(kx);
    free(ky);
    free(kz);
    free(mxy);
    free(mz);
    return PUFF_OK;
}

/* ------------------------------------------------------------------------ */
/* Diagnostics                                                               */
/* ------------------------------------------------------------------------ */
/* Power per z2 node (gPower / fPower_1D / fPower_3D, av_write.f90:108-202):
   trapezoidal integral over x and y on the linspace axes of initPowerCalc
   (:79-98).  f32_accumulate = 1 reproduces m_trapz / m_trapz2D being default
   REAL functions (:217-276): the running sum is rounded to single precision
   after every term. */
void orc_power(const puff_params* p, const double* ar, const double* ai,
               int f32_accumulate, double* power) {
    int nx = p->nx, ny = p->ny;
    int64_t ntr = (int64_t)nx * ny;
    if (nx == 1 && ny == 1) {
        for (int iz = 0; iz < p->nz2; iz++) power[iz] = pow(fabs(ar[iz]), 2.0) + pow(fabs(ai[iz]), 2.0);
        return;
    }
    double* xa = (double*)malloc(sizeof(double) * nx);
    double* ya = (double*)malloc(sizeof(double) * ny);
    double* cul = (double*)malloc(sizeof(double) * ny);
    double lx = p->dx * (nx - 1), ly = p->dy * (ny - 1);
    orc_linspace(-lx / 2.0, lx / 2.0, nx, xa);
    orc_linspace(-ly / 2.0, ly / 2.0, ny, ya);
    for (int iz = 0; iz < p->nz2; iz++) {
        const double* r = ar + iz * ntr;
        const double* im = ai + iz * ntr;
        for (int iy = 0; iy < ny; iy++) {
            double s = 0.0;
            float sf = 0.0f;
            for (int ix = 0; ix < nx - 1; ix++) {
                int64_t a0 = iy * nx + ix, a1 = a0 + 1;
                double f0 = pow(fabs(r[a0]), 2.0) + pow(fabs(im[a0]), 2.0);
                double f1 = pow(fabs(r[a1]), 2.0) + pow(fabs(im[a1]), 2.0);
                if (f32_accumulate)
                    sf = (float)((double)sf + (xa[ix + 1] - xa[ix]) * (f1 + f0) / 2.0);
                else
                    s = s + (xa[ix + 1] - xa[ix]) * (f1 + f0) / 2.0;
            }
            cul[iy] = f32_accumulate ? (double)sf : s;
        }
        double s = 0.0;
        float sf = 0.0f;
        for (int iy = 0; iy < ny - 1; iy++) {
            if (f32_accumulate)
                sf = (float)((double)sf + (ya[iy + 1] - ya[iy]) * (cul[iy + 1] + cul[iy]) / 2.0);
            else
                s = s + (ya[iy + 1] - ya[iy]) * (cul[iy + 1] + cul[iy]) / 2.0;
        }
        power[iz] = f32_accumulate ? (double)sf : s;
    }
    free(xa);
    free(ya);
    free(cul);
}

/* Slice bunching of one harmonic, getSliceTwiss (av_write.f90:498-650,
   725-729).  Linear weights split each particle between slices is and is+1
   on a temporal mesh; the reference writes past the array for is = nslices,
   which is skipped here.  A slice holding no particle of its own reports 0
   (sdata(is) = 0 branch, av_write.f90:732-749). */
int orc_bunching(const puff_params* p, int64_t n, const double* z2,
                 const double* chi, int harmonic, int nslices, double slice,
                 double* b) {
    double* br = (double*)calloc(nslices + 1, sizeof(double));
    double* bi = (double*)calloc(nslices + 1, sizeof(double));
    double* sdata = (double*)calloc(nslices + 1, sizeof(double));
    int status = PUFF_OK;
    for (int64_t i = 0; i < n; i++) {
        int is = (int)ceil(z2[i] / slice);
        if (p->periodic) {
            if (is >= nslices) is = is - ((int)floor((double)is / (double)nslices) * nslices) + 1;
            if (is < 1) { status = PUFF_ERR_ARG; continue; }
        } else {
            if ((is > nslices) || (is < 1)) { status = PUFF_ERR_ARG; continue; }
        }
        sdata[is - 1] += chi[i];
        double ph = (double)harmonic * z2[i] / (2.0 * p->rho);
        double c = cos(ph), s = sin(ph);
        double w1 = ((is * slice - z2[i]) / slice);
        br[is - 1] += chi[i] * c * w1;
        bi[is - 1] += chi[i] * s * w1;
        if (!p->periodic && is < nslices) {
            double w2 = (-((is - 1) * slice - z2[i]) / slice);
            br[is] += chi[i] * c * w2;
            bi[is] += chi[i] * s * w2;
        }
    }
    for (int is = 0; is < nslices; is++)
        b[is] = (sdata[is] > 0.0) ? sqrt(br[is] * br[is] + bi[is] * bi[is]) / slice : 0.0;
    free(br);
    free(bi);
    free(sdata);
    return status;
}

/* Slice statistics of getSliceTwiss (io/av_write.f90:405-731): the charge,
   means, r.m.s. widths, emittances, mean Gamma and energy spread of each
   slice of length `slice`.  Slice index is = ceil(z2 / slice) with the
   periodic wrap of :507-513; out-of-range macroparticles are skipped
   (:514-518).  The p_y sums carry the stored sign of py (-py, :530-536).
   out[k * nslices + s] for k = 0 sdata, 1 aveX, 2 aveY, 3 avePX, 4 avePY,
   5 sdX, 6 sdY, 7 sdpX, 8 sdpY, 9 eX, 10 eY, 11 aveGamma, 12 aveDgamma
   (:664-731, negative variances clipped to zero before the square roots). */
int orc_slice_stats(const puff_params* p, int64_t n, const double* x, const double* y,
                    const double* z2, const double* px, const double* py, const double* gam,
                    const double* chi, int nslices, double slice, double* out) {
    enum { S, MX, MY, MPX, MPY, MG, MXX, MXPX, MPXPX, MPYPY, MYPY, MYY, MGG, NS };
    double* acc = (double*)calloc((size_t)NS * nslices, sizeof(double));
    int status = PUFF_OK;
    for (int64_t i = 0; i < n; i++) {
        int is = (int)ceil(z2[i] / slice);
        if (p->periodic) {
            if (is >= nslices) is = is - ((int)floor((double)is / (double)nslices) * nslices) + 1;
            if (is < 1) { status = PUFF_ERR_ARG; continue; }
        } else {
            if ((is > nslices) || (is < 1)) { status = PUFF_ERR_ARG; continue; }
        }
        double* a = acc + (is - 1);
        const double c = chi[i];
        a[S * nslices] += c;
        a[MX * nslices] += c * x[i];
        a[MY * nslices] += c * y[i];
        a[MPX * nslices] += c * px[i];
        a[MPY * nslices] -= c * py[i];
        a[MG * nslices] += c * gam[i];
        a[MXX * nslices] += c * x[i] * x[i];
        a[MXPX * nslices] += c * x[i] * px[i];
        a[MYY * nslices] += c * y[i] * y[i];
        a[MYPY * nslices] -= c * y[i] * py[i];
        a[MPXPX * nslices] += c * px[i] * px[i];
        a[MPYPY * nslices] += c * py[i] * py[i];
        a[MGG * nslices] += c * gam[i] * gam[i];
    }
    for (int k = 0; k < 13 * nslices; k++) out[k] = 0.0;
    for (int is = 0; is < nslices; is++) {
        double* o = out + is;
        const double sd = acc[S * nslices + is];
        o[0] = sd;
        if (!(sd > 0.0)) continue;
#define A(k) acc[(k) * nslices + is]
        const double ax = A(MX) / sd, ay = A(MY) / sd, apx = A(MPX) / sd, apy = A(MPY) / sd;
        double sdx = (A(MXX) / sd) - pow(A(MX) / sd, 2.0);
        double sdy = (A(MYY) / sd) - pow(A(MY) / sd, 2.0);
        double sdpx = (A(MPXPX) / sd) - pow(A(MPX) / sd, 2.0);
        double sdpy = (A(MPYPY) / sd) - pow(A(MPY) / sd, 2.0);
        double adg = (A(MGG) / sd) - pow(A(MG) / sd, 2.0);
        double ex = sdx * sdpx - pow(A(MXPX) / sd - (ax * apx), 2.0);
        double ey = sdy * sdpy - pow(A(MYPY) / sd - (ay * apy), 2.0);
#undef A
        o[1 * nslices] = ax;
        o[2 * nslices] = ay;
        o[3 * nslices] = apx;
        o[4 * nslices] = apy;
        o[5 * nslices] = sqrt(sdx < 0.0 ? 0.0 : sdx);
        o[6 * nslices] = sqrt(sdy < 0.0 ? 0.0 : sdy);
        o[7 * nslices] = sqrt(sdpx < 0.0 ? 0.0 : sdpx);
        o[8 * nslices] = sqrt(sdpy < 0.0 ? 0.0 : sdpy);
        o[9 * nslices] = sqrt(ex < 0.0 ? 0.0 : ex);
        o[10 * nslices] = sqrt(ey < 0.0 ? 0.0 : ey);
        o[11 * nslices] = acc[MG * nslices + is] / sd;
        o[12 * nslices] = sqrt(adg < 0.0 ? 0.0 : adg);
    }
    free(acc);
    return status;
}

/* Charge binned linearly onto npts nodes spaced sam apart (getCurr,
   io/av_write.f90:297-368, before its unit scalings): node inl =
   ceil(z2 / sam) takes (1 - f) chi and node inl + 1 takes f chi, with f the
   fractional position; the periodic wrap of :317-327.  A macroparticle
   outside the nodes is the reference's STOP: status PUFF_ERR_ARG. */
int orc_current(const puff_params* p, int64_t n, const double* z2, const double* chi, int npts,
                double sam, double* out) {
    int status = PUFF_OK;
    for (int k = 0; k < npts; k++) out[k] = 0.0;
    for (int64_t i = 0; i < n; i++) {
        const int inl = (int)ceil(z2[i] / sam);
        const int inu = inl + 1;
        int lo = inl, hi = inu;
        if (p->periodic) {
            if (inu >= npts) hi = inu - (int)floor((double)inu / (double)npts) * npts + 1;
            if (inl >= npts) lo = inl - (int)floor((double)inl / (double)npts) * npts + 1;
            if (inl <= 0) { status = PUFF_ERR_ARG; continue; }
        } else {
            if ((inu > npts) || (inl <= 0)) { status = PUFF_ERR_ARG; continue; }
        }
        const double locz2 = z2[i] - (double)(inl - 1) * sam;
        const double li2 = locz2 / sam;
        const double li1 = 1.0 - li2;
        out[lo - 1] = li1 * chi[i] + out[lo - 1];
        out[hi - 1] = li2 * chi[i] + out[hi - 1];
    }
    return status;
}

/* ------------------------------------------------------------------------ */
/* Undulator entry / exit matching and the exit corrector, used to rebuild    */
/* the reference's own 1-D test run.                                         */
/* ------------------------------------------------------------------------ */
/* pxOffset / pyOffset (init_conds.f90:188-223) */
static double orc_px_offset(double z, double rho, double uy, double n2col) {
    return -uy * n2col * cos(z / (2.0 * rho));
}
static double orc_py_offset(double z, double rho, double ux, double n2col) {
    return -ux * n2col * sin(z / (2.0 * rho));
}

/* matchIn (sign = +1, acc_lattice.f90:743-828) and matchOut (sign = -1,
   :638-725) with xOffSet_Array and yOffSet_Array (init_conds.f90:65-86,
   129-152).  The p offsets are per particle for the curved-pole and
   plane-pole undulators (2nd-order terms in x, y, :653-685 / :758-790). */
void orc_match(const puff_params* p, double n2col, int64_t n, double* x,
               double* y, double* px, double* py, const double* gam, double sz,
               int sign) {
    double ncx = p->aw * p->aw;
    double ncy = 2.0 * p->aw * p->aw / (p->fx * p->fx + p->fy * p->fy);
    double kx = p->kx_und, ky = p->ky_und;
    for (int64_t i = 0; i < n; i++) {
        double spx0, spy0;
        if (p->und_type == 3) {          /* curved */
            spx0 = orc_px_offset(sz, p->rho, p->fy, n2col) - 0.5 * kx * kx * x[i] * x[i] -
                   0.5 * ky * ky * y[i] * y[i];
            spy0 = -1.0 * (orc_py_offset(sz, p->rho, p->fx, n2col) - kx * kx * x[i] * y[i]);
        } else if (p->und_type == 1) {   /* planepole */
            spx0 = orc_px_offset(sz, p->rho, p->fy, n2col) -
                   0.5 * (p->eta / (4.0 * p->rho * p->rho)) * x[i] * x[i];
            spy0 = -1.0 * orc_py_offset(sz, p->rho, p->fx, n2col);
        } else {
            spx0 = orc_px_offset(sz, p->rho, p->fy, n2col);
            spy0 = -1.0 * orc_py_offset(sz, p->rho, p->fx, n2col);
        }
        /* matchIn passes -spy0_offset as py, matchOut passes spy0_offset */
        double pyarg = (sign > 0) ? -spy0 : spy0;
        double gj = p->gamma_r * gam[i];
        double xo = -p->fy * 4.0 * p->kappa * pow(p->rho, 2.0) / sqrt(p->eta) *
                    (p->gamma_r / sqrt(gj * gj - (1.0 + ncx * (spx0 * spx0 + pyarg * pyarg)))) *
                    n2col * sin(sz / (2.0 * p->rho));
        double yo = p->fx * 4.0 * p->kappa * pow(p->rho, 2.0) / sqrt(p->eta) *
                    (p->gamma_r / sqrt(gj * gj - (1.0 + ncy * (spx0 * spx0 + pyarg * pyarg)))) *
                    n2col * cos(sz / (2.0 * p->rho));
        if (sign > 0) {
            x[i] = x[i] + xo;
            y[i] = y[i] + yo;
            px[i] = px[i] + spx0;
            py[i] = py[i] + spy0;
        } else {
            x[i] = x[i] - xo;
            y[i] = y[i] - yo;
            px[i] = px[i] - spx0;
            py[i] = py[i] - spy0;
        }
    }
}

/* driftSection, acc_lattice.f90:478-518 (particle part): z2 += dz p2 and,
   in 3-D, x += 2 rho kappa/sqrt(eta) (1 + eta p2)/Gamma px dz, y -= ... py dz. */
void orc_drift(const puff_params* p, int64_t n, double* x, double* y, double* z2,
               const double* px, const double* py, const double* gam, double dz) {
    for (int64_t i = 0; i < n; i++) {
        double p2 = orc_p2(gam[i], px[i], py[i], p->eta, p->gamma_r, p->aw);
        z2[i] = z2[i] + dz * p2;
        if (!p->one_d) {
            x[i] = x[i] + (2 * p->rho * p->kappa / sqrt(p->eta) * (1 + p->eta * p2) / gam[i] * px[i]) * dz;
            y[i] = y[i] - (2 * p->rho * p->kappa / sqrt(p->eta) * (1 + p->eta * p2) / gam[i] * py[i]) * dz;
        }
    }
}

/* disperse, acc_lattice.f90:420-463 (particle part), with the periodic wrap. */
void orc_chicane(const puff_params* p, int64_t n, double* z2, const double* gam,
                 double disp, double slip) {
    double lenz2 = p->dz2 * (double)(p->nz2 - 1);
    for (int64_t i = 0; i < n; i++) {
        z2[i] = z2[i] - 2.0 * disp * (gam[i] - 1.0) + slip;
        if (p->periodic) z2[i] = z2[i] - (floor(z2[i] / lenz2) * lenz2);
    }
}

/* Quad, acc_lattice.f90:531-561. */
void orc_quad(const puff_params* p, int64_t n, const double* x, const double* y,
              double* px, double* py, double qfx, double qfy) {
    if (p->one_d) return;
    for (int64_t i = 0; i < n; i++) {
        px[i] = px[i] + sqrt(p->eta) / (2 * p->rho * p->kappa) * x[i] / qfx;
        py[i] = py[i] - sqrt(p->eta) / (2 * p->rho * p->kappa) * y[i] / qfy;
    }
}

/* bModulation, acc_lattice.f90:575-585. */
void orc_modulation(int64_t n, const double* z2, double* gam, double wavenum, double mag) {
    for (int64_t i = 0; i < n; i++) gam[i] = gam[i] + (mag * cos(wavenum * z2[i]));
}

/* correctTrans, acc_lattice.f90:601-623 with arr_mean_para_weighted
   (base_functions.f90:552-573). */
void orc_correct_trans(int64_t n, double* x, double* y, double* px, double* py,
                       const double* chi) {
    double ws = 0.0, spx = 0.0, spy = 0.0, sx = 0.0, sy = 0.0;
    for (int64_t i = 0; i < n; i++) {
        ws += chi[i];
        spx += chi[i] * px[i];
        spy += chi[i] * py[i];
        sx += chi[i] * x[i];
        sy += chi[i] * y[i];
    }
    spx /= ws;
    spy /= ws;
    sx /= ws;
    sy /= ws;
    for (int64_t i = 0; i < n; i++) {
        px[i] -= spx;
        py[i] -= spy;
        x[i] -= sx;
        y[i] -= sy;
    }
}

/* ------------------------------------------------------------------------ */
/* UndSection's split-step schedule, undulator.f90:103-333 (no resume, no    */
/* retries: a raised flag ends the section with its status).                 */
/* write_mask[i] != 0 marks a write after step i+1: power(z2) of the written */
/* field is stored to powers[w * nz2] for the w-th write.                    */
/* taper = {n2col, n2col0, undgrad}; *sz_local is updated.                   */
/* ------------------------------------------------------------------------ */
int orc_und_section(const puff_params* p, double* taper, int64_t n, double* x,
                    double* y, double* z2, double* pr, double* pi, double* gam,
                    const double* chi, double* ar, double* ai, int nsteps,
                    double h, int diffraction, int steps4diff, int redist_every,
                    int64_t* csteps, const uint8_t* write_mask, double* powers,
                    double* sz_local, int nthreads) {
    int status = PUFF_OK;
    double szl = *sz_local;
    double dzds = 0.0, dzdf, dzd;
    int nw = 0;
    orc_wrap_z2(p, n, z2);                                 /* getLocalFieldIndices, :103 */
    if (diffraction) {                                     /* first half step, :156-164 */
        dzds = (double)steps4diff * h / 2.0;
        orc_diffraction_step(p, ar, ai, dzds, nthreads);
    }
    for (int istep = 1; istep <= nsteps; istep++) {
        (*csteps)++;
        status = orc_rk4par(p, taper, n, x, y, z2, pr, pi, gam, chi, ar, ai, szl, h, nthreads);
        if (status != PUFF_OK) break;
        szl = szl + h;                                     /* :222 */
        int wr = write_mask ? write_mask[istep - 1] : 0;
        if (diffraction && ((istep % steps4diff == 0) || (istep == nsteps))) {
            dzdf = dzds;                                   /* :237 */
            if (istep == nsteps) {
                dzds = 0.0;
            } else if ((istep + steps4diff) <= nsteps) {
                dzds = (double)steps4diff * h / 2;
            } else {
                int left = nsteps - istep;
                dzds = (double)left * h / 2;
            }
            if (!wr) {
                dzd = dzdf + dzds;                         /* :269-271 */
                orc_diffraction_step(p, ar, ai, dzd, nthreads);
            } else {
                orc_diffraction_step(p, ar, ai, dzdf, nthreads);  /* :279 */
                if (powers) orc_power(p, ar, ai, 0, powers + (int64_t)nw * p->nz2);
                nw++;
                if (dzds > 0.0) orc_diffraction_step(p, ar, ai, dzds, nthreads);
                wr = 0;                                    /* written already */
            }
        }
        if (wr) {
            if (powers) orc_power(p, ar, ai, 0, powers + (int64_t)nw * p->nz2);
            nw++;
        }
        if (redist_every > 0 && (*csteps % redist_every) == 0) orc_wrap_z2(p, n, z2);
    }
    *sz_local = szl;
    return status;
}

/* ======================================================================== */
/* Discrete adjoint of one rk4par step and of DiffractionStep.               */
/*                                                                          */
/* No reference counterpart: the reference computes no derivatives.  These  */
/* are the exact transposes of the oracle's own FP64 steps above (the        */
/* tri-linear weights are differentiated inside each cell), derived by hand  */
/* from the stage equations.  They are pinned by central finite differences  */
/* of orc_rk4par / orc_diffraction_step and by dot-product identities.       */
/* ======================================================================== */

/* Weights of orc_interp and their derivatives along x, y and z2. */
static int orc_interp_d(const puff_params* p, double x, double y, double z2,
                        int64_t* node, double* w, double* wx, double* wy, double* wz) {
    int f = orc_interp(p, x, y, z2, node, w);
    if (f) return f;
    const double zq = floor(z2 / p->dz2);
    const double fz = (z2 - zq * p->dz2) / p->dz2;
    if (p->one_d) {
        wx[0] = wx[1] = wy[0] = wy[1] = 0.0;
        wz[0] = -1.0 / p->dz2;
        wz[1] = 1.0 / p->dz2;
        return 0;
    }
    const double halfx = ((p->nx - 1) / 2.0) * p->dx, halfy = ((p->ny - 1) / 2.0) * p->dy;
    const double xq = floor((x + halfx) / p->dx), yq = floor((y + halfy) / p->dy);
    const double fx = (x + halfx - xq * p->dx) / p->dx, fy = (y + halfy - yq * p->dy) / p->dy;
    const double X[2] = {1.0 - fx, fx}, Y[2] = {1.0 - fy, fy}, Z[2] = {1.0 - fz, fz};
    const double dX[2] = {-1.0 / p->dx, 1.0 / p->dx};
    const double dY[2] = {-1.0 / p->dy, 1.0 / p->dy};
    const double dZ[2] = {-1.0 / p->dz2, 1.0 / p->dz2};
    for (int k = 0; k < 8; k++) {
        const int a = k & 1, b = (k >> 1) & 1, c = k >> 2;
        wx[k] = dX[a] * Y[b] * Z[c];
        wy[k] = X[a] * dY[b] * Z[c];
        wz[k] = X[a] * Y[b] * dZ[c];
    }
    return 0;
}

/* Transverse derivatives of orc_bfields: g = {dbx/dx, dbx/dy, dby/dx, dby/dy,
   dbz/dx, dbz/dy}. */
void orc_bfields_grad(const puff_params* p, int place, double sz, double x, double y,
                      double* g) {
    const double rho = p->rho, pi = p->pi_ref, seta = sqrt(p->eta);
    const double szt = sz / 2.0 / rho;
    int t = p->und_type;
    for (int k = 0; k < 6; k++) g[k] = 0.0;
    if (t == PUFF_UND_HELICAL) t = PUFF_UND_ELLIPTICAL;
    if (t == PUFF_UND_CURVED) {
        const double kx = p->kx_und, ky = p->ky_und;
        double fs, fc, cz;
        if (place == 0) {
            fs = szt / 4.0 / pi * sin(szt);
            fc = (szt / 4.0 / pi) * cos(szt);
            cz = seta / 2.0 / rho / ky;
        } else if (place == 2) {
            const double s = (sz - p->szfe) / 2.0 / rho;
            fs = (-s / 4.0 / pi + 1.0) * sin(s);
            fc = (-s / 4.0 / pi + 1.0) * cos(s);
            cz = seta / 2.0 / rho / kx;
        } else {
            fs = sin(szt);
            fc = cos(szt);
            cz = seta / 2.0 / rho / kx;
        }
        const double shx = sinh(kx * x), chx = cosh(kx * x), shy = sinh(ky * y), chy = cosh(ky * y);
        g[0] = kx / ky * kx * chx * shy * fs;
        g[1] = kx / ky * shx * ky * chy * fs;
        g[2] = kx * shx * chy * fs;
        g[3] = ky * chx * shy * fs;
        g[4] = cz * kx * shx * shy * fc;
        g[5] = cz * ky * chx * chy * fc;
    } else if (t == PUFF_UND_PLANEPOLE) {
        double fc;
        if (place == 0) {
            fc = (szt / 4.0 / pi) * cos(szt);
        } else if (place == 2) {
            const double s = (sz - p->szfe) / 2.0 / rho;
            fc = (-s / 4.0 / pi + 1.0) * cos(s);
        } else {
            fc = cos(szt);
        }
        g[5] = seta / 2.0 / rho * cosh(seta / 2.0 / rho * y) * fc;
    } else {
        /* elliptical: bz is linear in x and y with no constant term, bx and by
           depend on x, y only through the focusing terms */
        puff_params q = *p;
        double bx, by, bz;
        q.focusing = 0;
        orc_bfields(&q, place, sz, 1.0, 0.0, &bx, &by, &bz);
        g[4] = bz;
        orc_bfields(&q, place, sz, 0.0, 1.0, &bx, &by, &bz);
        g[5] = bz;
    }
    if (p->focusing) {
        g[1] += seta * p->kbeta_y * p->kbeta_y / p->kappa;
        g[2] += -seta * p->kbeta_x * p->kbeta_x / p->kappa;
    }
    if (p->one_d) {
        g[4] = 0.0;
        g[5] = 0.0;
    }
}

/* Adjoint of one orc_derivs evaluation at trial particles s* and trial field
   (ar, ai), with stage taper n2col.  Cotangents: c* of the six particle
   derivatives and (gr, gi) of the field source.  m* (overwritten) receive
   d<c, f>/d(particle); (mr, mi) accumulate d<c, f>/d(field).  Returns
   d<c, f>/d(n2col); *flags_out gets any bounds flag. */
static double orc_derivs_adjoint(const puff_params* p, double n2col, int place, double sz,
                                 int64_t n, const double* ar, const double* ai,
                                 const double* gr, const double* gi, const double* sx,
                                 const double* sy, const double* sz2, const double* spr,
                                 const double* spi, const double* sgam, const double* chi,
                                 const double* cx, const double* cy, const double* cz2,
                                 const double* cpr, const double* cpi, const double* cg,
                                 double* mx, double* my, double* mz2, double* mpr, double* mpi,
                                 double* mg, double* mr, double* mi, int nthreads,
                                 int* flags_out) {
    const int nnb = p->one_d ? 2 : 8;
    const double dV3 = p->dx * p->dy * p->dz2;
    const double inv2rho = 1.0 / (2.0 * p->rho);
    const double eta = p->eta, kap = p->kappa, rho = p->rho;
    const double c_dx = 2 * rho * kap / sqrt(eta);
    const double gr2 = p->gamma_r * p->gamma_r, aw2 = p->aw * p->aw;
    double dn2 = 0.0;
    int flags = 0;
#pragma omp parallel for num_threads(nthreads) reduction(+ : dn2) reduction(| : flags) schedule(static)
    for (int64_t i = 0; i < n; i++) {
        int64_t nd[8];
        double w[8], wxk[8], wyk[8], wzk[8];
        const int f = orc_interp_d(p, sx[i], sy[i], sz2[i], nd, w, wxk, wyk, wzk);
        if (f) {
            flags |= f;
            mx[i] = my[i] = mz2[i] = mpr[i] = mpi[i] = mg[i] = 0.0;
            continue;
        }
        const double qr = spr[i], qi = spi[i], g = sgam[i];
        const double s2 = qr * qr + qi * qi;
        /* u = 1 + eta p2 = D^(-1/2), D = 1 - (1 + aw^2 |p|^2) / (gamma_r Gamma)^2 */
        const double D = 1.0 - 1.0 / (gr2 * g * g) * (1.0 + aw2 * s2);
        const double u = 1.0 / sqrt(D);
        const double u3 = u * u * u;
        const double p2 = orc_p2(g, qr, qi, eta, p->gamma_r, p->aw);
        const double fac = (1 + eta * p2) / g;
        const double du_dpr = u3 * aw2 * qr / (gr2 * g * g);
        const double du_dpi = u3 * aw2 * qi / (gr2 * g * g);
        const double du_dg = -u3 * (1.0 + aw2 * s2) / (gr2 * g * g * g);
        double fr = 0.0, fi = 0.0, Gr = 0.0, Gi = 0.0;
        for (int k = 0; k < nnb; k++) {
            fr += w[k] * ar[nd[k]];
            fi += w[k] * ai[nd[k]];
            Gr += w[k] * gr[nd[k]];
            Gi += w[k] * gi[nd[k]];
        }
        if (!p->coupling) { fr = 0.0; fi = 0.0; }
        const double dep = p->field_evolve ? chi[i] / dV3 : 0.0;
        double ax = 0, ay = 0, az = 0, apr = 0, api = 0, ag = 0;
        if (p->electrons_evolve) {
            ax = cx[i]; ay = cy[i]; az = cz2[i]; apr = cpr[i]; api = cpi[i]; ag = cg[i];
        }
        double bx, by, bz, bg[6];
        orc_bfields(p, place, sz, sx[i], sy[i], &bx, &by, &bz);
        orc_bfields_grad(p, place, sz, sx[i], sy[i], bg);
        const double ep2k = eta * p2 / (kap * kap);
        /* d/dF of the dp_r, dp_i and dGamma equations */
        const double Lfr = p->coupling ? (-apr * inv2rho * ep2k - ag * 2.0 * rho * fac * qr) : 0.0;
        const double Lfi = p->coupling ? (-api * inv2rho * ep2k - ag * 2.0 * rho * fac * qi) : 0.0;
        /* partials through u (p2 = (u-1)/eta) and fac, and the explicit ones */
        const double Lu = az / eta - (apr * fr + api * fi) * inv2rho / (kap * kap);
        const double Lfac = ax * c_dx * qr - ay * c_dx * qi + kap * n2col * bz * (apr * qi - api * qr) -
                            2.0 * rho * ag * (qr * fr + qi * fi) + dep * (qr * Gr + qi * Gi);
        const double Epr = ax * c_dx * fac - api * kap * fac * n2col * bz - 2.0 * rho * ag * fac * fr +
                           dep * fac * Gr;
        const double Epi = -ay * c_dx * fac + apr * kap * fac * n2col * bz - 2.0 * rho * ag * fac * fi +
                           dep * fac * Gi;
        const double dfac_dpr = du_dpr / g, dfac_dpi = du_dpi / g, dfac_dg = du_dg / g - fac / g;
        mpr[i] = Epr + Lu * du_dpr + Lfac * dfac_dpr;
        mpi[i] = Epi + Lu * du_dpi + Lfac * dfac_dpi;
        mg[i] = Lu * du_dg + Lfac * dfac_dg;
        const double Lbx = api * inv2rho * n2col, Lby = apr * inv2rho * n2col;
        const double Lbz = kap * fac * n2col * (apr * qi - api * qr);
        double sxv = 0.0, syv = 0.0, szv = 0.0;
        for (int k = 0; k < nnb; k++) {
            const double V = Lfr * ar[nd[k]] + Lfi * ai[nd[k]] + dep * fac * (qr * gr[nd[k]] + qi * gi[nd[k]]);
            sxv += wxk[k] * V;
            syv += wyk[k] * V;
            szv += wzk[k] * V;
        }
        mx[i] = sxv + Lbx * bg[0] + Lby * bg[2] + Lbz * bg[4];
        my[i] = syv + Lbx * bg[1] + Lby * bg[3] + Lbz * bg[5];
        mz2[i] = szv;
        for (int k = 0; k < nnb; k++) {
#pragma omp atomic
            mr[nd[k]] += w[k] * Lfr;
#pragma omp atomic
            mi[nd[k]] += w[k] * Lfi;
        }
        dn2 += apr * (inv2rho * by + kap * qi * fac * bz) + api * (inv2rho * bx - kap * qr * fac * bz);
    }
    *flags_out |= flags;
    return dn2;
}

/* ------------------------------------------------------------------------ */
/* Adjoint of orc_rk4par.  Input: the state at the start of the step (x ..  */
/* gam, ar, ai; not modified), taper = {n2col, n2col0, undgrad} at the start */
/* of the step (not modified), and the cotangent (lx .. lg, lar, lai) of the */
/* state at the end of the step.  Output: the cotangent of the state at the  */
/* start, in place, and dn2col[s] = dJ/d(n2col of stage s), s = 0..3.        */
/*                                                                          */
/* RK4 transposed: with w = (1, 2, 2, 1) h/6 and the trial coefficients      */
/* c = (h/2, h/2, h), the cotangent of k_s is a_s = w_s L + c_s mu_{s+1}     */
/* (a_4 = w_4 L), mu_s = J_f(Y_s)^T a_s, and L_start = L + sum_s mu_s, for   */
/* the particle and field parts alike.  The aliased last plane of a periodic */
/* mesh is a copy of plane 1: its cotangent is folded into plane 1.          */
/* ------------------------------------------------------------------------ */
int orc_rk4par_adjoint(const puff_params* p, const double* taper_in, int64_t n,
                       const double* x, const double* y, const double* z2,
                       const double* pr, const double* pi, const double* gam,
                       const double* chi, const double* ar, const double* ai,
                       double* lx, double* ly, double* lz2, double* lpr, double* lpi,
                       double* lg, double* lar, double* lai, double sz, double h,
                       double* dn2col, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    const int64_t nmesh = (int64_t)p->nx * p->ny * p->nz2;
    const int64_t ntr = (int64_t)p->nx * p->ny;
    const int64_t nn = n > 0 ? n : 1;
    const double hh = h * 0.5, h6 = h / 6.0;
    const double zs[4] = {sz, sz + hh, sz + hh, sz + h};
    const double cin[4] = {0.0, hh, hh, h};
    const double wts[4] = {1.0, 2.0, 2.0, 1.0};
    double taper[3] = {taper_in[0], taper_in[1], taper_in[2]};
    double n2c[4];
    int flags = 0, status = PUFF_OK;

    orc_work wk;
    wk.p = p;
    wk.taper = taper;
    wk.n = n;
    wk.nthreads = 1;
    wk.nmesh = nmesh;
    wk.priv = NULL;
    wk.nodes = (int64_t*)malloc(sizeof(int64_t) * 8 * nn);
    wk.wts = (double*)malloc(sizeof(double) * 8 * nn);
    wk.p2 = (double*)malloc(sizeof(double) * nn);

    /* particle sets: Y[s] trial states (Y[0] = y0), K[s] derivatives,
       M[s] stage adjoints, C cotangent of the current stage */
    double* pbuf = (double*)calloc((size_t)(6 * 15) * nn, sizeof(double));
    double *Y[4][6], *K[3][6], *M[4][6], *C[6];
    for (int c = 0; c < 6; c++) {
        for (int s = 1; s < 4; s++) Y[s][c] = pbuf + ((s - 1) * 6 + c) * nn;
        for (int s = 0; s < 3; s++) K[s][c] = pbuf + ((18 + s * 6) + c) * nn;
        for (int s = 0; s < 4; s++) M[s][c] = pbuf + ((36 + s * 6) + c) * nn;
        C[c] = pbuf + (60 + c) * nn;
    }
    Y[0][0] = (double*)x; Y[0][1] = (double*)y; Y[0][2] = (double*)z2;
    Y[0][3] = (double*)pr; Y[0][4] = (double*)pi; Y[0][5] = (double*)gam;
    /* fields: At[s] trial fields (At[0] = A0), dA[s] sources, MA[s] stage
       adjoints, CA cotangent of the current stage's source */
    double* fbuf = (double*)calloc((size_t)(2 * 12) * nmesh, sizeof(double));
    double *At[4][2], *dA[3][2], *MA[4][2], *CA[2];
    for (int c = 0; c < 2; c++) {
        for (int s = 1; s < 4; s++) At[s][c] = fbuf + ((s - 1) * 2 + c) * nmesh;
        for (int s = 0; s < 3; s++) dA[s][c] = fbuf + ((6 + s * 2) + c) * nmesh;
        for (int s = 0; s < 4; s++) MA[s][c] = fbuf + ((12 + s * 2) + c) * nmesh;
        CA[c] = fbuf + (20 + c) * nmesh;
    }
    At[0][0] = (double*)ar;
    At[0][1] = (double*)ai;

    /* forward stages 1-3 again, keeping trial states and sources */
    for (int s = 0; s < 3; s++) {
        flags |= orc_derivs(&wk, zs[s], At[s][0], At[s][1], Y[s][0], Y[s][1], Y[s][2], Y[s][3],
                            Y[s][4], Y[s][5], chi, K[s][0], K[s][1], K[s][2], K[s][3], K[s][4],
                            K[s][5], dA[s][0], dA[s][1]);
        n2c[s] = taper[0];
        if (flags) goto done;
        for (int c = 0; c < 6; c++)
            for (int64_t i = 0; i < n; i++) Y[s + 1][c][i] = Y[0][c][i] + cin[s + 1] * K[s][c][i];
        for (int c = 0; c < 2; c++)
            for (int64_t j = 0; j < nmesh; j++) At[s + 1][c][j] = At[0][c][j] + cin[s + 1] * dA[s][c][j];
    }
    n2c[3] = orc_alpha(p, taper, zs[3]);

    if (p->periodic) {
        const int64_t al = (int64_t)(p->nz2 - 1) * ntr;
        for (int64_t t = 0; t < ntr; t++) {
            lar[t] += lar[al + t];
            lai[t] += lai[al + t];
            lar[al + t] = 0.0;
            lai[al + t] = 0.0;
        }
    }
    double* L[6] = {lx, ly, lz2, lpr, lpi, lg};
    double* LA[2] = {lar, lai};
    for (int s = 3; s >= 0; s--) {
        const double ws = wts[s] * h6;
        const double cn = (s < 3) ? cin[s + 1] : 0.0;
        for (int c = 0; c < 6; c++)
            for (int64_t i = 0; i < n; i++) C[c][i] = ws * L[c][i] + ((s < 3) ? cn * M[s + 1][c][i] : 0.0);
        for (int c = 0; c < 2; c++)
            for (int64_t j = 0; j < nmesh; j++) CA[c][j] = ws * LA[c][j] + ((s < 3) ? cn * MA[s + 1][c][j] : 0.0);
        dn2col[s] = orc_derivs_adjoint(p, n2c[s], orc_und_place(p, zs[s]), zs[s], n, At[s][0], At[s][1],
                                       CA[0], CA[1], Y[s][0], Y[s][1], Y[s][2], Y[s][3], Y[s][4],
                                       Y[s][5], chi, C[0], C[1], C[2], C[3], C[4], C[5], M[s][0],
                                       M[s][1], M[s][2], M[s][3], M[s][4], M[s][5], MA[s][0],
                                       MA[s][1], nthreads, &flags);
        if (flags) goto done;
    }
    for (int c = 0; c < 6; c++)
        for (int64_t i = 0; i < n; i++) L[c][i] += M[0][c][i] + M[1][c][i] + M[2][c][i] + M[3][c][i];
    for (int c = 0; c < 2; c++)
        for (int64_t j = 0; j < nmesh; j++) LA[c][j] += MA[0][c][j] + MA[1][c][j] + MA[2][c][j] + MA[3][c][j];

done:
    status = orc_status(flags);
    free(wk.nodes);
    free(wk.wts);
    free(wk.p2);
    free(pbuf);
    free(fbuf);
    return status;
}

/* ------------------------------------------------------------------------ */
/* Adjoint of orc_diffraction_step (in place on the cotangent lr, li).       */
/* With W the unnormalised forward DFT, N = ffact, E the multiplyexp factor  */
/* (with its filter), G the real damping and M the real mask, the forward   */
/* step is D = Q P, P = W^H E W / N, Q = (1 - M) + W^H G W M / N, so        */
/* D^H = P^H Q^H with P^H = W^H conj(E) W / N and Q^H = (1 - M) +            */
/* M W^H G W / N.  The output's aliased periodic plane is folded into plane  */
/* 1 first; the input's aliased plane is not read, so its cotangent is 0.   */
/* ------------------------------------------------------------------------ */
int orc_diffraction_adjoint(const puff_params* p, double* lr, double* li, double h,
                            int nthreads) {
    const int nx = p->nx, ny = p->ny;
    const int nzp = p->periodic ? p->nz2 - 1 : p->nz2;
    const int64_t ntr = (int64_t)nx * ny, nn = ntr * nzp;
    const double ffact = (double)nx * (double)ny * (double)nzp;
    double* a = (double*)malloc(sizeof(double) * 2 * nn);
    double* lam = (double*)malloc(sizeof(double) * 2 * nn);
    double* kx = (double*)malloc(sizeof(double) * nx);
    double* ky = (double*)malloc(sizeof(double) * ny);
    double* kz = (double*)malloc(sizeof(double) * nzp);
    double* mxy = (double*)malloc(sizeof(double) * ntr);
    double* mz = (double*)malloc(sizeof(double) * nzp);
    orc_kvalues(p, kx, ky, kz);
    orc_mask_xy(p, mxy);
    orc_mask_z2(p, mz);
    if (p->periodic) {
        const int64_t al = (int64_t)(p->nz2 - 1) * ntr;
        for (int64_t t = 0; t < ntr; t++) {
            lr[t] += lr[al + t];
            li[t] += li[al + t];
            lr[al + t] = 0.0;
            li[al + t] = 0.0;
        }
    }
    for (int64_t j = 0; j < nn; j++) {
        lam[2 * j] = a[2 * j] = lr[j];
        lam[2 * j + 1] = a[2 * j + 1] = li[j];
    }
    /* Q^H */
    orc_fft3d(a, nx, ny, nzp, -1, nthreads);
    for (int iz = 0; iz < nzp; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                if (kz[iz] != 0.0) {
                    int64_t j = ((int64_t)iz * ny + iy) * nx + ix;
                    double f = exp(-h * p->beta * (fabs(kx[ix]) + fabs(ky[iy])) /
                                   (sqrt(fabs(2.0 * kz[iz]))));
                    a[2 * j] *= f;
                    a[2 * j + 1] *= f;
                }
            }
    orc_fft3d(a, nx, ny, nzp, +1, nthreads);
    for (int iz = 0; iz < nzp; iz++)
        for (int64_t t = 0; t < ntr; t++) {
            int64_t j = (int64_t)iz * ntr + t;
            double m = mxy[t] + (mz[iz] * (1.0 - mxy[t]));
            a[2 * j] = (1.0 - m) * lam[2 * j] + m * (a[2 * j] / ffact);
            a[2 * j + 1] = (1.0 - m) * lam[2 * j + 1] + m * (a[2 * j + 1] / ffact);
        }
    /* P^H */
    orc_fft3d(a, nx, ny, nzp, -1, nthreads);
    const double cutoff = p->filt_cutoff;
    for (int iz = 0; iz < nzp; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int64_t j = ((int64_t)iz * ny + iy) * nx + ix;
                if ((kz[iz] > cutoff) || (kz[iz] < -cutoff)) {
                    if (kz[iz] != 0.0) {
                        double th = -h * (kx[ix] * kx[ix] + ky[iy] * ky[iy]) / (2.0 * kz[iz]);
                        double c = cos(th), s = sin(th);
                        double re = a[2 * j], im = a[2 * j + 1];
                        a[2 * j] = c * re - s * im;
                        a[2 * j + 1] = s * re + c * im;
                    }
                } else if (p->filter) {
                    a[2 * j] = 0.0;
                    a[2 * j + 1] = 0.0;
                }
            }
    orc_fft3d(a, nx, ny, nzp, +1, nthreads);
    for (int64_t j = 0; j < nn; j++) {
        lr[j] = a[2 * j] / ffact;
        li[j] = a[2 * j + 1] / ffact;
    }
    free(a);
    free(lam);
    free(kx);
    free(ky);
    free(kz);
    free(mxy);
    free(mz);
    return PUFF_OK;
}
