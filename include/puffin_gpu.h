/*
 * puffin_gpu.h -- C ABI of the MI355X (gfx950) implementation of Puffin's
 * per-step integrator: the RK4 macroparticle push with tri-linear field
 * interpolation and current deposition (rk4par) and the FFT diffraction step
 * (DiffractionStep / diffractIM).
 *
 * Plain C: pointers, sizes and status codes only.  One context per GPU per
 * process; a context is not re-entrant (calls on one context must not
 * overlap).  All arrays use the reference's data layout:
 *   particles : split SoA, one FP64 array per coordinate of length n
 *               (sElX_G, sElY_G, sElZ2_G, sElPX_G, sElPY_G, sElGam_G and the
 *               charge weights s_chi_bar_G; deriv_globals.f90:94,128-133);
 *               PY holds -p_y (simple_electron_gen.f90:330).
 *   field     : split real / imaginary FP64 arrays of length nx*ny*nz2,
 *               x fastest, then y, then z2 (para_field.f90:1370-1381); the
 *               imaginary part holds -A_y (rhs.f90:42-49).
 *
 * Status codes (the reference raises logical flags instead;
 * system_interpolation.f90:60-80, derivative.f90:96-137):
 *    0  ok
 *    1  a macroparticle is outside the transverse mesh   (qInnerXYOK_G = .false.)
 *    2  a macroparticle is beyond the z2 buffer           (qPArrOK_G   = .false.)
 *    3  a macroparticle is beyond the temporal z2 mesh    (the reference STOPs)
 *   <0  invalid argument / HIP error (see puff_last_error)
 * When rk4par returns 1..3 the particle and field state is left unchanged,
 * as the reference skips every RK4 update once a flag is raised
 * (puffin_mpi_RK4.f90:187,213,225,249,257,296,305).
 */
#ifndef PUFFIN_GPU_H
#define PUFFIN_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum puff_status {
    PUFF_OK = 0,
    PUFF_OUT_TRANSVERSE = 1,
    PUFF_OUT_Z2_BUFFER = 2,
    PUFF_OUT_Z2_MESH = 3,
    PUFF_ERR_ARG = -1,
    PUFF_ERR_HIP = -2,
    PUFF_ERR_ALLOC = -3,
    PUFF_ERR_STATE = -4
};

/* zUndType_G (setup_calcs.f90:757-777; bfields.f90:82,127,165,228) */
enum puff_und_type {
    PUFF_UND_ELLIPTICAL = 0, /* ''         : (fx, fy) free                       */
    PUFF_UND_PLANEPOLE = 1,  /* 'planepole': y-focusing plane pole, sinh in bz   */
    PUFF_UND_HELICAL = 2,    /* 'helical'  : fx = fy = 1                          */
    PUFF_UND_CURVED = 3      /* 'curved'   : curved-pole, sinh/cosh in x and y    */
};

/* Arithmetic of the integrator. */
enum puff_precision {
    PUFF_FP64 = 0,  /* plain FP64, the reference's arithmetic                    */
    PUFF_MIXED = 1  /* FP32 interpolation, deposition, field and FFT; FP64 z2,
                       Gamma and RK4 accumulation of z2 and Gamma                 */
};

enum puff_mem_kind { PUFF_MEM_HOST = 0, PUFF_MEM_DEVICE = 1 };

/*
 * Setup constants of one undulator module: the module globals the reference's
 * rk4par / DiffractionStep read (deriv_globals.f90:25-60,190-225,282-321,
 * 377-415; rhs_vars.f90:23-44).  Filled once per module by the caller,
 * exactly as setup / initUndulator leave them (acc_lattice.f90:848-895).
 */
typedef struct puff_params {
    int32_t nx, ny, nz2;        /* NX_G, NY_G, NZ2_G mesh nodes                  */
    int32_t periodic;           /* fieldMesh == iPeriodic (deriv_globals.f90:60) */
    int32_t one_d;              /* qOneD_G: 1-D interpolation, bz = 0            */
    int32_t und_type;           /* enum puff_und_type                            */
    int32_t und_ends;           /* qUndEnds_G                                    */
    int32_t focusing;           /* qFocussing_G                                  */
    int32_t electrons_evolve;   /* qElectronsEvolve_G                            */
    int32_t field_evolve;       /* qFieldEvolve_G                                */
    int32_t coupling;           /* qElectronFieldCoupling_G                      */
    int32_t filter;             /* qFilter (diffraction.f90:165)                 */
    int32_t nbx, nby, nbz2;     /* absorbing boundary widths NBX_G, NBY_G, NBZ2_G */
    int32_t reserved;
    double dx, dy, dz2;         /* sLengthOfElmX_G, _Y_G, _Z2_G                  */
    double rho, eta, kappa;     /* sRho_G, sEta_G, sKappa_G                      */
    double gamma_r, aw;         /* sGammaR_G, sAw_G                              */
    double fx, fy;              /* fx_G, fy_G (elliptical polarisation)          */
    double szfs, szfe;          /* sZFS, sZFE: module-local zbar of the main section */
    double kbeta_x, kbeta_y;    /* sKBetaXSF_G, sKBetaYSF_G (focusing)           */
    double kx_und, ky_und;      /* kx_und_G, ky_und_G (curved pole)              */
    double beta;                /* sBeta_G absorption coefficient                */
    double filt_cutoff;         /* multiplyexp high-pass cutoff in kz2
                                   (diffraction.f90:138-139)                     */
    double pi_ref;              /* the reference's pi constant, a default-real
                                   literal widened to FP64 (puffin_constants.f90:22) */
} puff_params;

typedef struct puff_ctx puff_ctx;

/* Last error message of this thread (static storage, never NULL). */
const char* puff_last_error(void);

/* Number of visible HIP devices, or <0 on error.  Safe without a GPU. */
int puff_device_count(void);

/*
 * Create a context on the current HIP device for n macroparticles.
 * Allocates all device buffers for one undulator module
 * (replaces allact_rk4_arrs, puffin_mpi_RK4.f90:374-418, and the FFTW plans
 * of getTransformPlans_MultiD, transforms.f90:113-226).
 */
int puff_create(puff_ctx** ctx, const puff_params* params, int64_t n,
                int32_t precision);
int puff_destroy(puff_ctx* ctx);

/* Run on the caller's HIP stream (hipStream_t cast to void*); NULL = own stream. */
int puff_set_stream(puff_ctx* ctx, void* hip_stream);
int puff_synchronize(puff_ctx* ctx);

/* Particle and field exchange in the reference layout (see top of file).
   mem_kind says whether the pointers are host or device memory.  Device
   pointers must hold FP64 arrays of the reference layout as well. */
int puff_upload_particles(puff_ctx* ctx, int32_t mem_kind, const double* x,
                          const double* y, const double* z2, const double* px,
                          const double* py, const double* gam,
                          const double* chi);
int puff_download_particles(puff_ctx* ctx, int32_t mem_kind, double* x,
                            double* y, double* z2, double* px, double* py,
                            double* gam);
int puff_upload_field(puff_ctx* ctx, int32_t mem_kind, const double* ar,
                      const double* ai);
int puff_download_field(puff_ctx* ctx, int32_t mem_kind, double* ar,
                        double* ai);

/* Taper state n2col / n2col0 / undgrad (wiggler_taper.f90:16-53; the
   reference mutates n2col0 once zbar passes sZFE). */
int puff_set_taper(puff_ctx* ctx, double n2col, double n2col0, double undgrad);
int puff_get_taper(puff_ctx* ctx, double* n2col, double* n2col0);

/*
 * Per-stage undulator strength override: when nvals > 0, stage zbar values
 * are mapped to n2col = values[floor((zbar - z0) / period)] (clamped to the
 * vector), replacing getAlpha's linear taper.  This is the per-period a_w
 * model; it has no counterpart in the reference.  nvals = 0 restores the
 * reference's linear taper.
 */
int puff_set_taper_table(puff_ctx* ctx, const double* values, int32_t nvals,
                         double z0, double period);

/*
 * Macroparticles are kept sorted by mesh cell on the device (re-sorted every
 * `steps` RK4 steps, default 30; 0 disables sorting).  The order is internal:
 * downloads always return the reference order of the last upload.
 */
int puff_set_sort_interval(puff_ctx* ctx, int32_t steps);

/*
 * Deterministic deposit (FP64 form only; no reference counterpart).  With
 * on = 1 every contribution to the current deposit is summed as fixed-point
 * integers (a 2 x 64-bit split resolving about 2^-68 of the step's largest
 * source), so puff_rk4par's results do not depend on the order in which
 * atomics land or on the macroparticles' slot order: the same inputs give
 * the same state bit for bit, run after run.  The reference's own deposit
 * (!$OMP ATOMIC, system_interpolation.f90:174-241) has no such mode.  Costs
 * a pass over the macroparticles per step, a conversion pass over the
 * active nodes per stage and 32 bytes of device memory per mesh node.  The
 * adjoint steps keep their atomic deposits.  PUFF_ERR_ARG in the mixed form
 * or above 2^30 macroparticles.
 */
int puff_set_deterministic(puff_ctx* ctx, int32_t on);

/*
 * One RK4 step of particles and field together.
 * Replaces rk4par(sZ, h, qD) (puffin_mpi_RK4.f90:38-366) with its four
 * derivs / getrhs evaluations (derivative.f90:36-155, rhs.f90:60-330).
 * sz is the module-local zbar (undulator.f90:199).  Returns a status code.
 */
int puff_rk4par(puff_ctx* ctx, double sz, double h);

/*
 * Free-space diffraction of the field over distance h.
 * Replaces DiffractionStep(h, sAr, sAi, qOK) (diffraction.f90:209-348):
 * forward 3-D DFT, multiplyexp, backward DFT, /ffact, AbsorptionStep.
 */
int puff_diffraction_step(puff_ctx* ctx, double h);

/* In-place unnormalised 3-D DFT of the field (the first NZ2' planes; sign -1
   forward, +1 backward).  Replaces Transform / Transform_MultiD with the
   plans of getTransformPlans_MultiD (transforms.f90:343-435, 186-211). */
int puff_transform(puff_ctx* ctx, int32_t sign);

/* Periodic z2 wrap applied when the reference recomputes its field layout
   (para_field.f90:212-215).  No-op on a temporal mesh. */
int puff_wrap_z2(puff_ctx* ctx);

/*
 * Lattice elements between undulator modules, on the device macroparticles
 * (acc_lattice.f90; the element chain is puffin_module.f90:58-90).  Each
 * returns after the device work is done.  The diffraction that the reference
 * applies over a drift or a chicane's zbar length (driftSection :512,
 * disperse :447) is the caller's puff_diffraction_step.
 *
 * puff_drift       driftSection (acc_lattice.f90:478-518): z2 += dz p2 and,
 *                  on a 3-D mesh, x += v px dz, y -= v py dz with
 *                  v = 2 rho kappa / sqrt(eta) (1 + eta p2) / Gamma.
 * puff_chicane     disperse (:420-463): z2 -= 2 disp (Gamma - 1) - slip,
 *                  wrapped on a periodic mesh.  disp and slip as the
 *                  reference holds them after readLatt (slip already scaled
 *                  by 4 pi rho, :356-363).
 * puff_quad        Quad (:531-561): px += sqrt(eta)/(2 rho kappa) x / fx,
 *                  py -= sqrt(eta)/(2 rho kappa) y / fy (3-D meshes only).
 * puff_modulation  bModulation (:575-585): Gamma += mag cos(wavenum z2).
 * puff_correct_trans  correctTrans (:601-623): subtract the chi-weighted
 *                  means of x, y, px, py.
 * puff_match       matchIn (sign +1, :743-828) or matchOut (sign -1,
 *                  :638-725) at zbar sz with strength n2col, for the
 *                  context's undulator type (curved and plane-pole
 *                  corrections included; init_conds.f90:65-223).
 * puff_set_undulator  the per-module undulator of initUndulator
 *                  (:848-895): type, polarisation fx/fy, in-undulator
 *                  focusing wavenumbers, and the module-local zbar of the
 *                  main section (sZFS, sZFE).
 */
int puff_drift(puff_ctx* ctx, double dz);
int puff_chicane(puff_ctx* ctx, double disp, double slip);
int puff_quad(puff_ctx* ctx, double fx, double fy);
int puff_modulation(puff_ctx* ctx, double wavenum, double mag);
int puff_correct_trans(puff_ctx* ctx);
int puff_match(puff_ctx* ctx, double sz, double n2col, int32_t sign);
int puff_set_undulator(puff_ctx* ctx, int32_t und_type, double fx, double fy, double kbeta_x,
                       double kbeta_y, double szfs, double szfe);

/* Power per z2 node: trapezoidal integral of |A|^2 over x and y
   (gPowerP / fPower_3D, av_write.f90:39-202), or |A|^2 in 1-D. */
int puff_power(puff_ctx* ctx, double* power_host);

/* Slice bunching |b_h| for harmonic h over nslices slices of width
   slice_len (getSliceTwiss, av_write.f90:405-773); b_host has nslices. */
int puff_bunching(puff_ctx* ctx, int32_t harmonic, int32_t nslices,
                  double slice_len, double* b_host);

/* Slice statistics of getSliceTwiss (av_write.f90:405-731) over nslices
   slices of width slice_len, slice = ceil(z2 / slice_len) with the
   reference's periodic wrap.  out_host has 13 * nslices doubles,
   out[k * nslices + s] for k = 0 slice charge (sum of chi-bar), 1 mean x,
   2 mean y, 3 mean px, 4 mean py (of the stored -py, as the reference),
   5 r.m.s. x, 6 r.m.s. y, 7 r.m.s. px, 8 r.m.s. py, 9 x emittance,
   10 y emittance (scaled, before the 2 kappa / sqrt(eta) / Gamma factor of
   hdf5_puff.f90:330-345), 11 mean Gamma, 12 r.m.s. Gamma.  Empty slices are
   zero. */
int puff_slice_stats(puff_ctx* ctx, int32_t nslices, double slice_len, double* out_host);

/*
 * Allocation-free slice statistics.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The same outputs as puff_slice_stats, in
 * the same layout, for n macroparticles held in the caller's DEVICE arrays
 * (FP64, the reference layout; PY holds -p_y), written to the caller's
 * device array out_dev of 13 * nslices doubles, on the caller's HIP stream
 * (hipStream_t cast to void*; NULL = the null stream) of the current device.
 * No context, no device or host allocation, no wait: it returns once the
 * kernels are queued.  The per-slice sums live in a library-owned static
 * device buffer of PUFF_NOALLOC_MAX_SLICES * 13 doubles on each GPU, which
 * the call clears first; nslices above PUFF_NOALLOC_MAX_SLICES is refused
 * with PUFF_ERR_ARG.  Calls that overlap on one GPU (other streams or host
 * threads) share that buffer and race; calls on one stream are ordered.
 * Only params->periodic is read (the slice wrap).
 */
#define PUFF_NOALLOC_MAX_SLICES 4096
int puff_slice_stats_noalloc(const puff_params* params, int64_t n, const double* x,
                             const double* y, const double* z2, const double* px,
                             const double* py, const double* gam, const double* chi,
                             int32_t nslices, double slice_len, double* out_dev,
                             void* hip_stream);

/* Beam current nodes of getCurr (av_write.f90:297-368): chi-bar binned
   linearly onto npts nodes spaced sam_len apart (node j at (j - 1) sam_len),
   before the reference's unit factors npk_bar (x ata in 1-D) x q_e /
   sam_len x c / Lc.  A macroparticle outside the nodes (the reference
   STOPs) returns PUFF_ERR_STATE. */
int puff_current(puff_ctx* ctx, int32_t npts, double sam_len, double* out_host);

/* Share of the stage workgroups (four stages) whose node box did not fit
   the LDS deposit tile and deposited with global atomics: for the last step
   run without deferred status, and for the first step after the last
   re-sort (the re-sort trigger compares the two). */
int puff_tile_fallback(puff_ctx* ctx, double* last_step, double* after_sort);

/* Number of macroparticles of the context. */
int puff_particle_count(puff_ctx* ctx, int64_t* n);

/*
 * Kernel timing with HIP events on the context's stream, replacing the
 * reference's wall-clock prints per step and per diffraction phase
 * (Get_time, undulator.f90:316-321; diffraction.f90:234-320).
 * Kinds: 0..3 RK4 stage kernels 1..4, 4 field combination, 5 diffraction
 * (all passes), 6 step start, 7 adjoint stage kernels (all four), 8 adjoint
 * field combination, 9 adjoint diffraction, 10 forward stages replayed by
 * the adjoint step, 11 macroparticle re-sort.  ms[k] is the summed duration,
 * counts[k] the number of launches since the last reset;
 * nkinds <= PUFF_PROF_KINDS.
 */
#define PUFF_PROF_KINDS 12
int puff_profile_enable(puff_ctx* ctx, int32_t on);
int puff_profile_read(puff_ctx* ctx, double* ms, int64_t* counts, int32_t nkinds,
                      int32_t reset);

/* Active box of the last step in cells (x0, x1, y0, y1, z0, z1), the region
   whose nodes the deposit touched. */
int puff_active_box(puff_ctx* ctx, int32_t* box6);

/*
 * State snapshots.  No reference counterpart: the reference resumes from
 * HDF5 dumps (io/hdf5_puff.f90), which is out of this library's scope.  A
 * blob holds the particles in the device's slot order, the field and the
 * taper and sort state; it lives in caller-owned device memory of
 * puff_state_bytes bytes (e.g. a torch uint8 CUDA tensor).  Used to take
 * checkpoints for the adjoint's recomputation.  Save and load are ordered on
 * the context's stream and return before the copies finish (puff_synchronize
 * waits for them); load reads the blob's header synchronously.
 */
int puff_state_bytes(puff_ctx* ctx, int64_t* bytes);
int puff_state_save(puff_ctx* ctx, void* device_dst);
int puff_state_load(puff_ctx* ctx, const void* device_src);
/* As above, also writing (save) or taking (load) a host copy of the blob's
   header, PUFF_STATE_HEADER_BYTES bytes: a load given the header of the same
   save does not wait for the device.  host_header may be null. */
#define PUFF_STATE_HEADER_BYTES 256
int puff_state_save_ex(puff_ctx* ctx, void* device_dst, void* host_header);
int puff_state_load_ex(puff_ctx* ctx, const void* device_src, const void* host_header);
/* Sort the macroparticles by mesh cell now (otherwise done every
   puff_set_sort_interval steps inside puff_rk4par).  The order is internal:
   downloads always return the caller's order. */
int puff_sort_now(puff_ctx* ctx);

/*
 * Discrete adjoint of the step.  No reference counterpart (the reference
 * computes no derivatives): the exact transpose of puff_rk4par and
 * puff_diffraction_step as implemented here, for reverse-mode gradients of
 * a scalar output (e.g. the integrated power) through a sequence of steps.
 *
 * The context holds a cotangent of the state (particles and field) beside
 * the state.  puff_adjoint_enable allocates it (zeroed); puff_adjoint_zero
 * clears it; puff_adjoint_seed_power adds scale * d/dA of
 * sum_iz wz[iz] power(iz) at the current field (power as puff_power,
 * wz host array of NZ2 weights).
 *
 * puff_rk4par_adjoint: the current state must be the state at the start of
 * the step (as puff_state_load restores it) and the cotangent that of the
 * state at the end of the step.  It replaces the cotangent by that of the
 * start state, leaves the state unchanged, and returns dn2col[k] =
 * d(output)/d(n2col of RK4 stage k) and slot[k] = the per-period taper
 * table slot of stage k (-1 without a table).  Status as puff_rk4par.
 * puff_diffraction_adjoint applies the transpose of one diffraction step of
 * length h to the field cotangent.
 *
 * puff_adjoint_upload / _download exchange the cotangent in the caller's
 * particle order and split field layout (either field pointer may be null).
 */
/*
 * Deferred status (no reference counterpart).  With puff_set_deferred(ctx,
 * 1), puff_rk4par and puff_rk4par_adjoint return PUFF_OK without waiting for
 * the step: a flagged step still leaves the state unchanged, its flags are
 * kept, and puff_deferred_status waits for the stream and returns the status
 * of all steps since its last call (as puff_rk4par).  In this mode the
 * adjoint step accumulates d(output)/d(n2col) per taper-table slot on the
 * device (slot -1, no table, into entry ntable) instead of returning dn2col;
 * puff_adjoint_taper_gradient copies the first n entries of that table
 * (zeroed by puff_adjoint_zero).  The re-sort between steps then follows only
 * puff_set_sort_interval.
 */
int puff_set_deferred(puff_ctx* ctx, int32_t on);
int puff_deferred_status(puff_ctx* ctx);
int puff_adjoint_taper_gradient(puff_ctx* ctx, double* out, int32_t n);

int puff_adjoint_enable(puff_ctx* ctx);
int puff_adjoint_zero(puff_ctx* ctx);
int puff_adjoint_seed_power(puff_ctx* ctx, const double* wz_host, double scale);
int puff_rk4par_adjoint(puff_ctx* ctx, double sz, double h, double* dn2col, int32_t* slot);
int puff_diffraction_adjoint(puff_ctx* ctx, double h);

/*
 * Step capture: puff_rk4par_capture is puff_rk4par that also keeps, in
 * caller-owned device memory of puff_capture_bytes bytes, the step's start
 * state (after the step's own re-sort) and its stage derivatives and sources,
 * with a host copy of the capture's header (PUFF_STATE_HEADER_BYTES).
 * puff_rk4par_adjoint_captured is puff_rk4par_adjoint for that step from the
 * capture: it neither reads nor changes the context's state and does not
 * recompute the forward stages.
 */
int puff_capture_bytes(puff_ctx* ctx, int64_t* bytes);
int puff_rk4par_capture(puff_ctx* ctx, double sz, double h, void* capture, void* host_header);
int puff_rk4par_adjoint_captured(puff_ctx* ctx, double sz, double h, const void* capture,
                                 const void* host_header, double* dn2col, int32_t* slot);
int puff_adjoint_upload(puff_ctx* ctx, int32_t mem_kind, const double* x, const double* y,
                        const double* z2, const double* px, const double* py,
                        const double* gam, const double* ar, const double* ai);
int puff_adjoint_download(puff_ctx* ctx, int32_t mem_kind, double* x, double* y, double* z2,
                          double* px, double* py, double* gam, double* ar, double* ai);

#ifdef __cplusplus
}
#endif

#endif /* PUFFIN_GPU_H */
