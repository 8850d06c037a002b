// Kernel argument blocks and host launchers shared by the .hip files.
#pragma once

#include "puff_common.hpp"

namespace puff {

// Mesh geometry as seen by the push kernels.  The whole mesh is the
// interpolation window: halfx = ((NX-1)/2) dx (rhs.f90:395-396 with
// nspinDX = NX).  nzp is the z2 period (NZ2-1) on a periodic mesh, else NZ2.
struct MeshGeom {
    int nx, ny, nz2, nzp;
    int periodic, one_d;
    double dx, dy, dz2;
    double inv_dx, inv_dy, inv_dz2;
    double halfx, halfy;
};

struct Physics {
    double eta, rho, kappa, gamma_r, aw;
    double inv2rho;     // 1 / (2 rho)                    (rhs.f90:366)
    double inv_dV3;     // 1 / (dx dy dz2)                (rhs.f90:361)
    double c_dx;        // 2 rho kappa / sqrt(eta)        (puffin_equations.f90:115)
    double inv_eta;     // 1 / eta
    double inv_gr2;     // 1 / gamma_r^2
    int coupling, electrons_evolve, field_evolve;
};

// Six particle coordinates: x, y, p_r, p_i in R; z2 and Gamma in Z (FP64
// for the state).
template <typename R, typename Z = double>
struct PartPtrs {
    R* x;
    R* y;
    Z* z2;
    R* pr;
    R* pi;
    Z* g;
};

// RK4 stage derivatives k_s: all six components in R.  In the mixed form the
// derivatives of z2 and Gamma are stored in FP32, while the trial states
// y0 + c k and the step's accumulation y0 + h/6 (k1 + 2 k2 + 2 k3 + k4) of z2
// and Gamma are formed in FP64 (24 instead of 32 bytes per derivative).
template <typename R>
using KPtrs = PartPtrs<R, R>;

// Deterministic deposit of the FP64 form.  Every contribution v = w_k s of a
// macroparticle to a node is split into two fixed-point integers,
// hi = rint(v 2^F) and lo = rint((v 2^F - hi) 2^DET_LO_BITS), and summed with
// integer atomics: integer sums do not depend on the order in which lanes and
// workgroups add them, so the deposit, and with it the whole step, is the
// same bit for bit from run to run whatever the atomics' order or the
// macroparticles' slot order.  DET_WORDS int64 per node: Re hi, Re lo, Im hi,
// Im lo.  F is set per step from the largest source of the step's start
// state (k_det_bound) so that n (4 max |s|) 2^F < 2^62: no node sum can
// overflow, and the sums resolve 2^-(F + DET_LO_BITS), some 2^-68 of the
// largest source at 2^24 macroparticles.  A conversion pass after each stage
// turns the integers back into FP64 deposits.
constexpr int DET_WORDS = 4;
constexpr int DET_LO_BITS = 32;
constexpr int DET_SAFETY = 4;

struct DetState {
    unsigned long long* acc;   // DET_WORDS per mesh node, zero between stages
    int* shift;                // F of the current step (device word)
    unsigned long long* bound; // bits of the largest |source| (device word)
};

template <typename R>
struct StageArgs {
    int64_t n;
    PartPtrs<R> y0;     // state at the start of the step (written by stage 4)
    KPtrs<R> kin;       // derivative of the previous stage
    KPtrs<R> kout;      // derivative of this stage (stages 1-3)
    KPtrs<R> k1, k2;    // stage 4: first two derivatives (k3 = kin)
    const R* chi;
    double cin;         // trial coefficient of this stage (0, h/2, h/2, h)
    double cnext;       // trial coefficient of the next stage
    double h6;
    const cplx<R>* A0;
    const cplx<R>* dAin;
    cplx<R>* dAout;
    MeshGeom mesh;
    Physics phys;
    StageScalars s;
    int* flags;
    Box* box;
    DetState det;       // acc == nullptr: the atomic FP64 deposit
    // stage 1 of a captured step: the start state and chi also go to the
    // capture (x == nullptr: not captured), written from the registers the
    // stage loads them into instead of by a separate copy
    PartPtrs<R> ycap;
    R* chicap;
};

template <typename R>
struct CombineArgs {
    cplx<R>* A;
    cplx<R>* d[4];
    double h6;
    const Box* box;
    int* flags;
    int nx, ny, nz2, nzp, periodic, one_d;
    int clear_from;        // d[k] with k >= clear_from are cleared after the update
};

template <typename R>
void launch_stage(int stage, const StageArgs<R>& a, hipStream_t s);
void launch_step_begin(int* flags, Box* box, hipStream_t s);
template <typename R>
void launch_combine(const CombineArgs<R>& a, hipStream_t s);
void launch_wrap(double* z2, int64_t n, double lenz2, hipStream_t s);
// Deterministic deposit: F of the step from the sources of the state y
// (particles with a flag contribute nothing later, so all are bounded), and
// the conversion of the integer sums of one stage into dA over the step's
// active box (clearing the sums).
void launch_det_scale(const PartPtrs<double>& y, const double* chi, int64_t n, const Physics& ph,
                      const DetState& d, hipStream_t s);
void launch_det_convert(const DetState& d, cplx<double>* dA, const Box* box, int nx, int ny, int nzp,
                        int periodic, int one_d, hipStream_t s);

// ---- discrete adjoint of one RK4 step -----------------------------------------
// Stage s (4 .. 1) pulls the cotangent a_s = ws L + cnext mu_{s+1} of its
// derivative k_s back through f at the trial state y0 + cin k_{s-1} and trial
// field A0 + cin dA_{s-1}: mu_s = J_f^T a_s (particles into mout, field into
// mAout).  Stage 1 then forms L <- L + mu_1 + mu_2 + mu_3 + mu_4.
constexpr int DN2_COPIES = 64;   // spread accumulators of dJ/dn2col per stage

template <typename R>
struct AdjArgs {
    int64_t n;
    PartPtrs<R> y0;
    KPtrs<R> kin;          // k_{s-1} (stages 2-4)
    const R* chi;
    double cin;
    const cplx<R>* A0;
    const cplx<R>* dAin;   // dA_{s-1} (stages 2-4)
    PartPtrs<R> L;         // cotangent of the step's end state
    const cplx<R>* LA;
    PartPtrs<R> mnext;     // mu_{s+1} (stages 1-3)
    const cplx<R>* mAnext;
    double cnext;
    double ws;             // w_s h / 6, w = (1, 2, 2, 1)
    PartPtrs<R> mout;      // mu_s (stages 2-4)
    PartPtrs<R> m3, m4;    // stage 1: mu_3, mu_4 (mu_2 = mnext)
    cplx<R>* mAout;
    MeshGeom mesh;
    Physics phys;
    StageScalars s;
    int* flags;
    Box* box;
    double* dn2;           // DN2_COPIES accumulators of this stage
};

template <typename R>
struct AdjCombineArgs {
    cplx<R>* LA;
    cplx<R>* mA[4];
    cplx<R>* dA[3];        // forward sources of stages 1-3, cleared here
    const Box* box;
    int* flags;
    int nx, ny, nz2, nzp, periodic, one_d;
    int clear_dA;          // clear dA[0..2] (replayed forward sources)
};

// dn2_table[slot[s]] += sum of the DN2_COPIES accumulators of stage s
// (slot -1: the entry after the table), s = 0..3.
struct Slots4 {
    int s[4];
};
void launch_dn2_accumulate(const double* dn2, Slots4 slots, int ntab, double* dn2_table, hipStream_t st);

template <typename R>
void launch_adj_stage(int stage, const AdjArgs<R>& a, hipStream_t s);
template <typename R>
void launch_adj_combine(const AdjCombineArgs<R>& a, hipStream_t s);
// L_A += scale 2 wz[iz] wxy |A|^2 gradient, i.e. the cotangent of
// sum_iz wz[iz] power(iz) with power the transverse trapezoid of |A|^2.
template <typename R>
void launch_adj_seed_power(cplx<R>* LA, const cplx<R>* A, const double* wz, int nx, int ny,
                           int nz2, int periodic, double dx, double dy, double scale,
                           hipStream_t s);

template <typename R>
void launch_fold_alias(cplx<R>* LA, int64_t ntr, int nz2, hipStream_t s);
void launch_fill_zero_words(uint32_t* p, int64_t n, hipStream_t s);

// ---- particle sort by cell --------------------------------------------------
// Sort order of the macroparticle slots: cells are grouped in bricks of
// BRICK_X x BRICK_Y x BRICK_Z cells, bricks in planes along z2 with snaking
// rows; the cells of a brick run in snaking rows along x (BRICK_XFAST) or in
// snaking columns along z2.  A run of slots (one stage workgroup) then
// covers a compact 3-D block of cells whose node box grows slowly as the
// beam slips along z2 between sorts, and lanes of a wavefront share x-rows of
// nodes.  One bin per cell of the padded brick grid, plus one for particles
// off the mesh.  Bricks of 4 x 2 x 4 with rows along x measured best: 2-5 %
// faster steps than plane/row order on the C1 beam, with 2-6x fewer re-sorts
// once bunched; 1 % faster than 8 x 2 x 4 there and 4 % on the sparser C4
// beam (fewer nodes flushed per macroparticle, fewer tile overflows).
#ifndef PUFF_BRICK_X
#define PUFF_BRICK_X 4
#endif
#ifndef PUFF_BRICK_Y
#define PUFF_BRICK_Y 2
#endif
#ifndef PUFF_BRICK_Z
#define PUFF_BRICK_Z 4
#endif
#ifndef PUFF_BRICK_XFAST
#define PUFF_BRICK_XFAST 1
#endif
constexpr int BRICK_X = PUFF_BRICK_X, BRICK_Y = PUFF_BRICK_Y, BRICK_Z = PUFF_BRICK_Z;
// Re-sort trigger: the share of stage workgroups that overflowed the LDS
// tile in a step exceeds both PUFF_RESORT_MIN_SHARE and PUFF_RESORT_GROWTH
// times its value on the first step after the last sort.
#ifndef PUFF_RESORT_MIN_SHARE
#define PUFF_RESORT_MIN_SHARE 0.02
#endif
#ifndef PUFF_RESORT_GROWTH
#define PUFF_RESORT_GROWTH 2.0
#endif
constexpr bool BRICK_XFAST = PUFF_BRICK_XFAST != 0;
constexpr int BRICK_CELLS = BRICK_X * BRICK_Y * BRICK_Z;

// Number of sort bins for particles on the mesh (the last bin, for particles
// off it, is this value).
__host__ __device__ inline int64_t sort_bins(int nx, int ny, int nzp, bool one_d) {
    if (one_d) return nzp;
    const int64_t bx = (nx - 1 + BRICK_X - 1) / BRICK_X, by = (ny - 1 + BRICK_Y - 1) / BRICK_Y;
    const int64_t bz = (nzp + BRICK_Z - 1) / BRICK_Z;
    return bx * by * bz * BRICK_CELLS;
}

template <typename R>
struct SortArgs {
    int64_t n;
    int ncell;             // bins on the mesh (sort_bins); bin ncell collects particles off it
    MeshGeom mesh;
    PartPtrs<R>* y;        // sorted in place (pointers swapped with spare)
    PartPtrs<R>* spare;
    R** chi;
    R** chi_spare;
    int** orig;            // reference index of each slot
    int** orig_spare;
    int* key;
    int* counts;           // ncell + 1 bins, scanned in place
    int* cursor;
    int* perm;
    int* tile_sums;
};

template <typename R>
void launch_sort(SortArgs<R>& a, hipStream_t s);

// Lattice elements between undulator modules (puff_lattice.hip).
enum LatticeKind { LAT_DRIFT, LAT_CHICANE, LAT_QUAD, LAT_MODULATION, LAT_CORRECT_TRANS, LAT_MATCH };

struct MatchArgs {
    double sz, n2col;             // module-local zbar, undulator strength
    double rho, eta, kappa, gamma_r;
    double ux, uy;                // fx_G, fy_G
    double ncx, ncy;              // aw^2 and 2 aw^2 / (ux^2 + uy^2)
    double kx, ky;                // curved-pole wavenumbers
    int und_type;
    int sign;                     // +1 matchIn, -1 matchOut
};

struct LatticeOp {
    int kind;
    double a, b;                  // drift: dz; chicane: disp, slip; quad: fx, fy;
                                  // modulation: wavenumber, magnitude
    double lenz2;                 // chicane: periodic wrap length (0: none)
    MatchArgs match;
};

template <typename R>
void launch_lattice(const LatticeOp& op, PartPtrs<R> y, const R* chi, int64_t n, const Physics& ph,
                    const MeshGeom& m, double* sums, hipStream_t s);
void launch_iota(int* p, int64_t n, hipStream_t s);
template <typename T>
void launch_unpermute(const T* src, const int* orig, int64_t n, double* dst, hipStream_t s);
template <typename T>
void launch_permute_in(const double* src, const int* orig, int64_t n, T* dst, hipStream_t s);

// ---- diffraction -----------------------------------------------------------

// Spectral operation applied between the forward and inverse z2 transforms.
enum SpecOp { SPEC_PROPAGATE = 0, SPEC_ABSORB = 1 };

// x-line epilogues of the diffraction pipeline.
enum XOp {
    XOP_FORWARD = 0,        // forward transform only
    XOP_INV_SPLIT_FWD = 1,  // inverse, /ffact, split into boundary and interior, forward
    XOP_INV_RECOMBINE = 2,  // inverse, /ffact, + interior
    // adjoint pipeline
    XOP_ADJ_FOLD_SAVE_FWD = 3,  // fold the aliased periodic plane, keep a copy, forward
    XOP_ADJ_INV_MIX_FWD = 4,    // inverse, /ffact, (1 - m) copy + m value, forward
    XOP_ADJ_INV_SCALE = 5       // inverse, /ffact
};

template <typename R>
struct DiffArgs {
    cplx<R>* a;          // field, nz2 planes (only the first nzp take part)
    cplx<R>* anb;        // interior part (AbsorptionStep's sAnb)
    const double* kx;    // GetKValues
    const double* ky;
    const double* kz;
    const double* kzr;   // per kz: 1/(2 kz), 1/sqrt|2 kz| (0 at kz = 0)
    const double* mxy;   // transverse mask
    const double* mz;    // z2 mask
    int nx, ny, nz2, nzp, periodic, filter;
    double h, beta, cutoff, ffact;
    FftAxis ax, ay, az;
};

template <typename R>
void launch_diffraction(const DiffArgs<R>& a, hipStream_t s);
// Adjoint of launch_diffraction applied to the cotangent held in a.a.
template <typename R>
void launch_diffraction_adjoint(const DiffArgs<R>& a, hipStream_t s);

// Plain 3-D transform (tests): sign -1 forward, +1 backward, unnormalised.
// tmp: a buffer of the field's size (the four-step z2 pass's middle output).
template <typename R>
void launch_fft3d(cplx<R>* a, cplx<R>* tmp, int nx, int ny, int nz, const FftAxis& ax, const FftAxis& ay,
                  const FftAxis& az, int sign, hipStream_t s);

// ---- diagnostics -----------------------------------------------------------
template <typename R>
void launch_power(const cplx<R>* a, int nx, int ny, int nz2, double dx, double dy,
                  double* out, hipStream_t s);
template <typename R>
void launch_bunching(const double* z2, const R* chi, int64_t n, double rho, int harmonic,
                     int nslices, double slice, int periodic, double* acc, hipStream_t s);
// getSliceTwiss's per-slice sums: acc holds NSLICE_SUMS x nslices doubles,
// zeroed by the caller; launched in chunks of SLICE_CHUNK slices (LDS).
constexpr int NSLICE_SUMS = 13;
constexpr int SLICE_CHUNK = 448;
template <typename R>
void launch_slice_moments(PartPtrs<R> y, const R* chi, int64_t n, int nslices, double slice, int periodic,
                          double* acc, hipStream_t s);
// getCurr's binned charge: acc holds npts + 1 doubles (the last counts
// macroparticles outside the nodes), zeroed by the caller.
constexpr int CURRENT_CHUNK = 4096;
template <typename R>
void launch_current(const double* z2, const R* chi, int64_t n, int npts, double sam, int periodic,
                    double* acc, hipStream_t s);

// ---- data movement ---------------------------------------------------------
// Up to 64 words written to device memory by a kernel (stream-ordered, no
// host staging buffer to keep alive).
struct Words64 {
    uint32_t w[64];
};
void launch_store_words(uint32_t* dst, const Words64& src, int nwords, hipStream_t s);
// Several device-to-device copies in one launch (byte counts multiple of 4,
// pointers 16-byte aligned).
struct CopySections {
    const uint8_t* src[12];
    uint8_t* dst[12];
    int64_t bytes[12];
    int n;
};
void launch_copy_sections(const CopySections& c, hipStream_t s);
void launch_fill_zero(double* p, int64_t n, hipStream_t s);
template <typename R>
void launch_pack_field(const double* ar, const double* ai, cplx<R>* a, int64_t n, hipStream_t s);
template <typename R>
void launch_unpack_field(const cplx<R>* a, double* ar, double* ai, int64_t n, hipStream_t s);
template <typename R>
void launch_convert(const double* src, R* dst, int64_t n, hipStream_t s);
template <typename R>
void launch_convert_back(const R* src, double* dst, int64_t n, hipStream_t s);

}  // namespace puff
