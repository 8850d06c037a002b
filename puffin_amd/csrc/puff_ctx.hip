// Context, host-side setup and the C ABI (include/puffin_gpu.h).
//
// Host-side restatements used to set up the device:
//   stage b-field coefficients   undulator/bfields.f90:36-881
//   taper                        undulator/wiggler_taper.f90:16-53
//   undulator place              interaction/puffin_equations.f90:218-254
//   k-vectors                    diffraction/transforms.f90:472-590
//   masks                        diffraction/masks.f90:50-430
//   ffact                        setup/setup.f90:334-346
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "puff_common.hpp"
#include "puff_kernels.hpp"

namespace puff {

static thread_local std::string g_err = "";
void set_error(const std::string& m) { g_err = m; }

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

template <typename R>
struct Particles {
    PartPtrs<R> y;
    KPtrs<R> k[3];
    R* chi = nullptr;
    // the sort's spare z2 and Gamma arrays in the mixed form (its spare x, y,
    // p_r, p_i are k[0]'s; the FP64 form's k[0] serves whole)
    double* z2_spare = nullptr;
    double* g_spare = nullptr;
};

constexpr int DN2_TABLE_CAP = 4096;   // taper-table slots of the deferred adjoint

template <typename R>
struct AdjBufs {
    PartPtrs<R> L;
    PartPtrs<R> M[3];
};

struct Ctx {
    puff_params p{};
    int precision = PUFF_FP64;
    int64_t n = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<void*> allocs;
    // particles (typed views; R chosen by precision)
    Particles<double> pd;
    Particles<float> pf;
    // field
    void* A = nullptr;
    void* dA[4] = {nullptr, nullptr, nullptr, nullptr};
    void* anb = nullptr;
    double* kx = nullptr;
    double* ky = nullptr;
    double* kz = nullptr;
    double* kzr = nullptr;   // 1/(2 kz), 1/sqrt|2 kz| per kz (the mixed form's spectral factor)
    double* mxy = nullptr;
    double* mz = nullptr;
    FftAxis ax{}, ay{}, az{};
    int* flags = nullptr;
    Box* box = nullptr;
    int* h_flags = nullptr;   // pinned
    double* stage_buf = nullptr;   // FP64 staging (2 * max(n, nmesh))
    int64_t stage_len = 0;
    double* diag = nullptr;
    int64_t diag_len = 0;
    // taper state (wiggler_taper.f90)
    double n2col = 1.0, n2col0 = 1.0, undgrad = 0.0;
    std::vector<double> taper_table;
    double taper_z0 = 0.0, taper_period = 1.0;
    MeshGeom mesh{};
    Physics phys{};
    // particle order: orig[j] = reference index of slot j
    int* orig = nullptr;
    int* orig_spare = nullptr;
    void* chi_spare = nullptr;
    int* sort_key = nullptr;
    int* sort_counts = nullptr;
    int* sort_perm = nullptr;
    int* sort_tiles = nullptr;
    int ncell = 0;
    bool identity_order = true;
    bool sorted = false;
    int sort_every = 30;
    int64_t steps_since_sort = 0;
    // adaptive re-sort: share of stage workgroups that fell back to global
    // atomics on the first step after the last sort, and the due flag
    double fallback_base = 0.0;
    double fallback_last = 0.0;   // share of the last synchronised step
    bool sort_due = false;
    // kernel timing (puff_profile_*)
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    struct Pend {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Pend> pending;
    double prof_ms[PUFF_PROF_KINDS] = {};
    int64_t prof_n[PUFF_PROF_KINDS] = {};
    // particle order identity: a new id for every sort or full upload
    int64_t order_id = 0;
    int64_t order_counter = 0;
    // discrete adjoint (puff_adjoint_enable): cotangent L of the particles
    // and LA of the field, stage cotangents M (stages 2..4) and mA (1..4)
    bool adj = false;
    AdjBufs<double> ad;
    AdjBufs<float> af;
    void* LA = nullptr;
    void* mA[4] = {nullptr, nullptr, nullptr, nullptr};
    double* dn2 = nullptr;     // 4 x DN2_COPIES
    double* h_dn2 = nullptr;   // pinned
    double* adj_wz = nullptr;  // nz2 power weights
    void* h_hdr = nullptr;     // pinned state-blob header
    int* lam_orig = nullptr;   // slot order of the particle cotangent
    int64_t lam_order_id = 0;
    double* dn2_table = nullptr;   // per-slot dJ/dn2col (deferred adjoint steps)
    // deferred status: steps return without waiting for their flags
    bool deferred = false;
    // deterministic deposit of the FP64 form (puff_set_deterministic)
    bool det = false;
    DetState dets{};

    int64_t nmesh() const { return (int64_t)p.nx * p.ny * p.nz2; }
    int nzp() const { return p.periodic ? p.nz2 - 1 : p.nz2; }
};

template <typename R>
static Particles<R>& parts(Ctx* c);
template <>
Particles<double>& parts<double>(Ctx* c) { return c->pd; }
template <>
Particles<float>& parts<float>(Ctx* c) { return c->pf; }
template <typename R>
static AdjBufs<R>& adjb(Ctx* c);
template <>
AdjBufs<double>& adjb<double>(Ctx* c) { return c->ad; }
template <>
AdjBufs<float>& adjb<float>(Ctx* c) { return c->af; }

static int dalloc(Ctx* c, void** p, size_t bytes) {
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
        return PUFF_ERR_ALLOC;
    }
    c->allocs.push_back(*p);
    e = hipMemset(*p, 0, bytes);
    if (e != hipSuccess) {
        set_error(std::string("hipMemset: ") + hipGetErrorString(e));
        return PUFF_ERR_HIP;
    }
    return PUFF_OK;
}

#define PUFF_TRY(x)                   \
    do {                              \
        int r_ = (x);                 \
        if (r_ != PUFF_OK) return r_; \
    } while (0)

// ---- FFT plans ----------------------------------------------------------------

static bool is_prime(int n) {
    if (n < 2) return false;
    for (int f = 2; (int64_t)f * f <= n; ++f)
        if (n % f == 0) return false;
    return true;
}

static int powmod(int b, int e, int n) {
    int64_t r = 1, x = b % n;
    for (; e > 0; e >>= 1) {
        if (e & 1) r = r * x % n;
        x = x * x % n;
    }
    return (int)r;
}

// Radix passes of an n-point line: prime factors of 5 and more (direct DFT
// passes) first, largest first, so that the costliest pass runs with p = 1
// and needs no input twiddles; then 16s, 4s, 2s and 3s.
static void factorize(int n, FftAxis& ax) {
    ax.n = n;
    ax.npass = 0;
    int m = n;
    int n16 = 0, n4 = 0, n2 = 0, n3 = 0;
    while (m % 16 == 0) { ++n16; m /= 16; }
    while (m % 4 == 0) { ++n4; m /= 4; }
    while (m % 2 == 0) { ++n2; m /= 2; }
    while (m % 3 == 0) { ++n3; m /= 3; }
    int big[32];
    int nbig = 0;
    for (int f = 5; m > 1 && (int64_t)f * f <= m; f += 2)
        while (m % f == 0) { big[nbig++] = f; m /= f; }
    if (m > 1) big[nbig++] = m;
    for (int k = nbig - 1; k >= 0; --k) ax.radix[ax.npass++] = big[k];
    for (int k = 0; k < n16; ++k) ax.radix[ax.npass++] = 16;
    for (int k = 0; k < n4; ++k) ax.radix[ax.npass++] = 4;
    for (int k = 0; k < n2; ++k) ax.radix[ax.npass++] = 2;
    for (int k = 0; k < n3; ++k) ax.radix[ax.npass++] = 3;
}

template <typename R>
static int make_axis(Ctx* c, int n, FftAxis& ax, bool z2_axis = false) {
    // one line, its ping-pong buffer and the twiddle table share a
    // workgroup's LDS (160 KiB on gfx950); 4096-point z2 lines run as
    // 64 x 64 (four-step) in either form
    if (3 * (size_t)n * sizeof(cplx<R>) > 160 * 1024 && !(z2_axis && n == 4096)) {
        set_error("FFT line of " + std::to_string(n) + " points does not fit one workgroup's LDS");
        return PUFF_ERR_ARG;
    }
    const long double two_pi = 2.0L * 3.141592653589793238462643383279502884L;
    const int m = n - 1;
    ax.rader = (n >= 17 && is_prime(n) && (m & (m - 1)) == 0) ? 1 : 0;
    ax.rperm = nullptr;
    ax.riperm = nullptr;
    ax.rb = nullptr;
    const int tn = ax.rader ? m : n;   // length of the radix-pass transform
    factorize(tn, ax);
    ax.n = n;
    std::vector<cplx<R>> tw(n, cplx<R>{0, 0});
    for (int e = 0; e < tn; ++e) {
        // exp(-2 pi i e / tn) in long double, rounded once
        const long double a = -two_pi * (long double)e / (long double)tn;
        tw[e] = {(R)std::cos(a), (R)std::sin(a)};
    }
    void* d = nullptr;
    PUFF_TRY(dalloc(c, &d, sizeof(cplx<R>) * n));
    PUFF_HIP(hipMemcpy(d, tw.data(), sizeof(cplx<R>) * n, hipMemcpyHostToDevice));
    ax.tw = d;
    if (ax.rader) {
        // primitive root: g^((n-1)/2) = -1 (mod n), n - 1 being a power of two
        int g = 2;
        while (powmod(g, m / 2, n) != n - 1) ++g;
        const int ginv = powmod(g, n - 2, n);
        std::vector<int> perm(m), iperm(m);
        for (int j = 0, gj = 1, gi = 1; j < m; ++j) {
            perm[j] = gj;
            iperm[j] = gi;
            gj = (int)((int64_t)gj * g % n);
            gi = (int)((int64_t)gi * ginv % n);
        }
        // rb_s = DFT_m(b_s) / m, b_s[j] = exp(s 2 pi i iperm[j] / n)
        std::vector<cplx<R>> rb(2 * (size_t)m);
        for (int si = 0; si < 2; ++si) {
            const long double sg = si ? 1.0L : -1.0L;
            for (int k = 0; k < m; ++k) {
                long double re = 0.0L, im = 0.0L;
                for (int j = 0; j < m; ++j) {
                    const long double ab = sg * two_pi * (long double)iperm[j] / (long double)n;
                    const long double ak = -two_pi * (long double)(((int64_t)j * k) % m) / (long double)m;
                    re += std::cos(ab + ak);
                    im += std::sin(ab + ak);
                }
                rb[(size_t)si * m + k] = {(R)(re / m), (R)(im / m)};
            }
        }
        void* dp = nullptr;
        void* dq = nullptr;
        void* db = nullptr;
        PUFF_TRY(dalloc(c, &dp, sizeof(int) * m));
        PUFF_TRY(dalloc(c, &dq, sizeof(int) * m));
        PUFF_TRY(dalloc(c, &db, sizeof(cplx<R>) * 2 * m));
        PUFF_HIP(hipMemcpy(dp, perm.data(), sizeof(int) * m, hipMemcpyHostToDevice));
        PUFF_HIP(hipMemcpy(dq, iperm.data(), sizeof(int) * m, hipMemcpyHostToDevice));
        PUFF_HIP(hipMemcpy(db, rb.data(), sizeof(cplx<R>) * 2 * m, hipMemcpyHostToDevice));
        ax.rperm = static_cast<const int*>(dp);
        ax.riperm = static_cast<const int*>(dq);
        ax.rb = db;
    }
    return PUFF_OK;
}

// ---- k-vectors and masks (host restatement) -----------------------------------

static void kvalues(const puff_params& p, std::vector<double>& kx, std::vector<double>& ky,
                    std::vector<double>& kz) {
    const double pi = 4.0 * std::atan(1.0);   // GetKValues' own pi (transforms.f90:494)
    const int nzp = p.periodic ? p.nz2 - 1 : p.nz2;
    const double lx = (double)p.nx * p.dx, ly = (double)p.ny * p.dy, lz = (double)nzp * p.dz2;
    const int mx = (int)std::ceil(p.nx / 2.0), my = (int)std::ceil(p.ny / 2.0), mz = (int)std::ceil(nzp / 2.0);
    kx.resize(p.nx);
    ky.resize(p.ny);
    kz.resize(nzp);
    for (int i = 0; i < p.nx; ++i) kx[i] = 2.0 * pi * (double)(i >= mx ? i - p.nx : i) / lx;
    for (int i = 0; i < p.ny; ++i) ky[i] = 2.0 * pi * (double)(i >= my ? i - p.ny : i) / ly;
    for (int i = 0; i < nzp; ++i) kz[i] = 2.0 * pi * (double)(i >= mz ? i - nzp : i) / lz;
}

static std::vector<double> linspace(double a, double b, int n) {
    std::vector<double> v(n);
    if (n > 1) {
        const double d = (b - a) / (double)(n - 1);
        for (int i = 0; i < n; ++i) v[i] = a + i * d;
    } else if (n == 1) {
        v[0] = a;
    }
    return v;
}

// Mask1D2 (masks.f90:266-302) when edge = true, Maskf4 (:329-361) otherwise.
static std::vector<double> mask_1d(const std::vector<double>& x, double dx, int nb, double pi, bool edge) {
    const int n = (int)x.size();
    std::vector<double> m(n, 0.0);
    const double lb = (nb - 1) * dx;
    const double x01 = edge ? x[n - 1] : x[n - nb];
    const double x02 = edge ? x[0] : x[nb - 1];
    for (int i = n - nb; i < n; ++i) m[i] = std::pow(std::cos(pi * (x[i] - x01) / (2.0 * lb)), 2.0);
    for (int i = 0; i < nb; ++i) m[i] = std::pow(std::cos(pi * (x[i] - x02) / (2.0 * lb)), 2.0);
    for (int i = nb; i < n - nb; ++i) m[i] = edge ? 0.0 : 1.0;
    return m;
}

static void masks(const puff_params& p, std::vector<double>& mxy, std::vector<double>& mz) {
    const int nx = p.nx, ny = p.ny, nzp = p.periodic ? p.nz2 - 1 : p.nz2;
    mxy.assign((size_t)nx * ny, 0.0);
    mz.assign(nzp, 0.0);
    if (nx > 1 && ny > 1 && p.nbx > 1 && p.nby > 1 && nx >= p.nbx && ny >= p.nby) {
        // getMask / Mask2D: Maskf3_y Mask1D2_x + Maskf4_x Mask1D2_y, Maskf3 = 1
        const double lx = (nx - 1) * p.dx, ly = (ny - 1) * p.dy;
        auto x = linspace(-lx / 2, lx / 2, nx);
        auto y = linspace(-ly / 2, ly / 2, ny);
        auto mx = mask_1d(x, p.dx, p.nbx, p.pi_ref, true);
        auto my = mask_1d(y, p.dy, p.nby, p.pi_ref, true);
        auto mx2 = mask_1d(x, p.dx, p.nbx, p.pi_ref, false);
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) mxy[(size_t)iy * nx + ix] = 1.0 * mx[ix] + mx2[ix] * my[iy];
    }
    if (!p.periodic && p.nbz2 > 1) {
        // getZ2Mask with its Lf = (nx - 1*dx) expression (masks.f90:389-422)
        const double dx = p.dz2;
        auto x = linspace(0.0, (double)(nzp - 1) * dx, nzp);
        const double lb = (p.nbz2 - 1) * dx;
        const double lf = (p.nz2 - 1 * dx);
        const double x0 = lf - lb / 2.0;
        const int fbn = p.nz2 - p.nbz2 + 1;
        for (int ind = 1; ind <= nzp; ++ind)
            if (ind >= fbn) mz[ind - 1] = std::pow(std::cos(p.pi_ref * (x[ind - 1] - x0) / lb), 2.0);
    }
}

// ---- per-stage scalars ------------------------------------------------------

// adjUndPlace (puffin_equations.f90:218-254): 0 start, 1 main, 2 end, -1 error
static int und_place(const puff_params& p, double z) {
    if (!p.und_ends) return 1;
    if (z < 0) return -1;
    if (z <= p.szfs) return 0;
    if (z >= p.szfe) return 2;
    return 1;
}

// Coefficient map of the b-fields at zbar z (bfields.f90):
//   elliptical / helical: bx = c0 + c3 y, by = c1 + c4 x, bz = c5 x + c6 y
//   plane pole:           bx = c3 y,      by = c1 + c4 x, bz = c2 sinh(c7 y)
//   curved pole:          bx = c0 sinh(c5 x) sinh(c6 y) + c3 y,
//                         by = c1 cosh(c5 x) cosh(c6 y) + c4 x,
//                         bz = c2 cosh(c5 x) sinh(c6 y)
static int stage_scalars(const puff_params& p, double z, double n2col, StageScalars& s) {
    for (double& v : s.c) v = 0.0;
    s.n2col = n2col;
    const int place = und_place(p, z);
    if (place < 0) {
        set_error("undulator section not recognised, zbar < 0");
        return PUFF_OUT_Z2_MESH;
    }
    const double rho = p.rho, pi = p.pi_ref, se = std::sqrt(p.eta);
    const double szt = z / 2.0 / rho;
    int t = p.und_type;
    double fx = p.fx, fy = p.fy;
    if (t == PUFF_UND_HELICAL) {
        fx = 1.0;
        fy = 1.0;
        t = PUFF_UND_ELLIPTICAL;
    }
    s.und_type = t;
    if (t == PUFF_UND_ELLIPTICAL) {
        if (place == 1) {                                   // bfields.f90:280-284, 496-500, 862-867
            s.c[0] = fx * std::cos(szt);
            s.c[1] = fy * std::sin(szt);
            s.c[5] = -se / 2.0 / rho * fx * std::sin(szt);
            s.c[6] = se / 2.0 / rho * fy * std::cos(szt);
        } else if (place == 0) {                            // :239-257, 481-485, 763-787
            double vx = fx * (z - pi * rho) / (6.0 * pi * rho) * std::cos(szt);
            double czx = -se / 2.0 / rho * (z - pi * rho) / (6.0 * pi * rho) * fx * std::sin(szt);
            if (z < pi * rho) {
                vx = 0.0;
                czx = 0.0;
            } else if (z > 7.0 * pi * rho) {
                vx = fx * std::cos(szt);
                czx = -se / 2.0 / rho * fx * std::sin(szt);
            }
            s.c[0] = vx;
            s.c[5] = czx;
            s.c[1] = fy * szt / 4.0 / pi * std::sin(szt);
            s.c[6] = szt / 4.0 / pi * se / 2.0 / rho * fy * std::cos(szt);
        } else {                                            // :259-278, 487-494, 803-838
            const double sl = z - p.szfe;
            double vx = -fx * (sl - 7.0 * pi * rho) / (6.0 * pi * rho) * std::cos(sl / 2.0 / rho);
            double czx = -se / 2.0 / rho * (sl - 7.0 * pi * rho) / (6.0 * pi * rho) * fx * std::sin(sl / 2.0 / rho);
            if (sl < pi * rho) {
                vx = fx * std::cos(sl / 2.0 / rho);
                czx = -se / 2.0 / rho * fx * std::sin(sl / 2.0 / rho);
            } else if (sl > 7.0 * pi * rho) {
                vx = 0.0;
                czx = 0.0;
            }
            const double s2 = sl / 2.0 / rho;
            s.c[0] = vx;
            s.c[5] = czx;
            s.c[1] = fy * (-s2 / 4.0 / pi + 1.0) * std::sin(s2);
            s.c[6] = (-sl / 8.0 / pi / rho + 1.0) * se / 2.0 / rho * fy * std::cos(sl / 2.0 / rho);
        }
    } else if (t == PUFF_UND_PLANEPOLE) {
        s.c[7] = se / 2.0 / rho;
        if (place == 1) {                                   // :409-415, 630-635
            s.c[1] = std::sin(szt);
            s.c[2] = std::cos(szt);
        } else if (place == 0) {                            // :380-391, 613-618
            s.c[1] = szt / 4.0 / pi * std::sin(szt);
            s.c[2] = (szt / 4.0 / pi) * std::cos(szt);
        } else {                                            // :395-407, 620-628
            const double sl = (z - p.szfe) / 2.0 / rho;
            s.c[1] = (-sl / 4.0 / pi + 1.0) * std::sin(sl);
            s.c[2] = (-sl / 4.0 / pi + 1.0) * std::cos(sl);
        }
    } else {  // curved pole
        const double kx = p.kx_und, ky = p.ky_und;
        s.c[5] = kx;
        s.c[6] = ky;
        if (place == 1) {                                   // :103-109, 353-359, 588-595
            s.c[0] = kx / ky * std::sin(szt);
            s.c[1] = std::sin(szt);
            s.c[2] = se / 2.0 / rho / kx * std::cos(szt);
        } else if (place == 0) {                            // :84-90, 334-340, 568-575
            const double f = szt / 4.0 / pi;
            s.c[0] = kx / ky * f * std::sin(szt);
            s.c[1] = f * std::sin(szt);
            s.c[2] = se / 2.0 / rho / ky * f * std::cos(szt);
        } else {                                            // :92-101, 342-351, 577-586
            const double sl = (z - p.szfe) / 2.0 / rho;
            const double f = (-sl / 4.0 / pi + 1.0);
            s.c[0] = kx / ky * f * std::sin(sl);
            s.c[1] = f * std::sin(sl);
            s.c[2] = se / 2.0 / rho / kx * f * std::cos(sl);
        }
    }
    if (p.focusing) {                                       // :299-306, 514-521
        s.c[3] = se * p.kbeta_y * p.kbeta_y / p.kappa;
        s.c[4] = -se * p.kbeta_x * p.kbeta_x / p.kappa;
    }
    if (p.one_d) {                                          // bz = 0 in 1-D (:544-551)
        if (t == PUFF_UND_ELLIPTICAL) {
            s.c[5] = 0.0;
            s.c[6] = 0.0;
        } else {
            s.c[2] = 0.0;
        }
    }
    return PUFF_OK;
}

// getAlpha with its side effect (wiggler_taper.f90:39-45), or the per-period
// table when one is set.
// Slot of the per-period table at zbar z (clamped), -1 without a table.
static int taper_slot(const Ctx* c, double z) {
    if (c->taper_table.empty()) return -1;
    long long k = (long long)std::floor((z - c->taper_z0) / c->taper_period);
    if (k < 0) k = 0;
    if (k >= (long long)c->taper_table.size()) k = (long long)c->taper_table.size() - 1;
    return (int)k;
}

static double advance_taper(Ctx* c, double z) {
    if (!c->taper_table.empty()) {
        c->n2col = c->taper_table[(size_t)taper_slot(c, z)];
        return c->n2col;
    }
    if (z >= c->p.szfs && z <= c->p.szfe) {
        c->n2col = c->n2col0 + c->undgrad * (z - c->p.szfs);
    } else if (z > c->p.szfe) {
        c->n2col0 = c->n2col;
    }
    return c->n2col;
}

// ---- context creation -------------------------------------------------------

template <typename R>
static int alloc_particles(Ctx* c) {
    Particles<R>& P = parts<R>(c);
    auto setp = [&](PartPtrs<R>& q) -> int {
        PUFF_TRY(dalloc(c, (void**)&q.x, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.y, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.z2, sizeof(double) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.pr, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.pi, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.g, sizeof(double) * c->n));
        return PUFF_OK;
    };
    auto setk = [&](KPtrs<R>& q) -> int {
        PUFF_TRY(dalloc(c, (void**)&q.x, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.y, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.z2, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.pr, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.pi, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.g, sizeof(R) * c->n));
        return PUFF_OK;
    };
    PUFF_TRY(setp(P.y));
    for (int k = 0; k < 3; ++k) PUFF_TRY(setk(P.k[k]));
    if constexpr (sizeof(R) != sizeof(double)) {
        PUFF_TRY(dalloc(c, (void**)&P.z2_spare, sizeof(double) * c->n));
        PUFF_TRY(dalloc(c, (void**)&P.g_spare, sizeof(double) * c->n));
    }
    PUFF_TRY(dalloc(c, (void**)&P.chi, sizeof(R) * c->n));
    PUFF_TRY(dalloc(c, &c->chi_spare, sizeof(R) * c->n));
    PUFF_TRY(dalloc(c, (void**)&c->orig, sizeof(int) * c->n));
    PUFF_TRY(dalloc(c, (void**)&c->orig_spare, sizeof(int) * c->n));
    PUFF_TRY(dalloc(c, (void**)&c->sort_key, sizeof(int) * c->n));
    PUFF_TRY(dalloc(c, (void**)&c->sort_perm, sizeof(int) * c->n));
    const int64_t bins = sort_bins(c->p.nx, c->p.ny, c->nzp(), c->p.one_d != 0);
    if (bins >= (int64_t)INT32_MAX) {
        set_error("mesh too large for 32-bit sort bins");
        return PUFF_ERR_ARG;
    }
    c->ncell = (int)bins;
    PUFF_TRY(dalloc(c, (void**)&c->sort_counts, sizeof(int) * ((size_t)c->ncell + 1)));
    PUFF_TRY(dalloc(c, (void**)&c->sort_tiles, sizeof(int) * ((size_t)c->ncell / 4096 + 2)));
    launch_iota(c->orig, c->n, nullptr);
    PUFF_HIP(hipDeviceSynchronize());
    const int64_t nm = c->nmesh();
    PUFF_TRY(dalloc(c, &c->A, sizeof(cplx<R>) * nm));
    for (int k = 0; k < 4; ++k) PUFF_TRY(dalloc(c, &c->dA[k], sizeof(cplx<R>) * nm));
    if (!c->p.one_d) {
        PUFF_TRY(dalloc(c, &c->anb, sizeof(cplx<R>) * nm));
        PUFF_TRY(make_axis<R>(c, c->p.nx, c->ax));
        PUFF_TRY(make_axis<R>(c, c->p.ny, c->ay));
        PUFF_TRY(make_axis<R>(c, c->nzp(), c->az, true));
    }
    return PUFF_OK;
}

static int validate(const puff_params* p, int64_t n) {
    if (!p) return PUFF_ERR_ARG;
    if (n < 0 || p->nx < 1 || p->ny < 1 || p->nz2 < 3) {
        set_error("invalid sizes");
        return PUFF_ERR_ARG;
    }
    if (p->one_d && (p->nx != 1 || p->ny != 1)) {
        set_error("1-D mesh must have nx = ny = 1");
        return PUFF_ERR_ARG;
    }
    if (!p->one_d && (p->nx < 2 || p->ny < 2)) {
        set_error("3-D mesh needs nx, ny >= 2");
        return PUFF_ERR_ARG;
    }
    if (!(p->dx > 0 && p->dy > 0 && p->dz2 > 0 && p->eta > 0 && p->rho > 0 && p->kappa > 0 &&
          p->gamma_r > 0)) {
        set_error("invalid physical constants");
        return PUFF_ERR_ARG;
    }
    if ((int64_t)p->nx * p->ny * p->nz2 >= (int64_t)1 << 31) {
        set_error("mesh too large for 32-bit node indices");
        return PUFF_ERR_ARG;
    }
    return PUFF_OK;
}

static int create(Ctx* c) {
    const puff_params& p = c->p;
    MeshGeom& m = c->mesh;
    m.nx = p.nx;
    m.ny = p.ny;
    m.nz2 = p.nz2;
    m.nzp = c->nzp();
    m.periodic = p.periodic;
    m.one_d = p.one_d;
    m.dx = p.dx;
    m.dy = p.dy;
    m.dz2 = p.dz2;
    m.inv_dx = 1.0 / p.dx;
    m.inv_dy = 1.0 / p.dy;
    m.inv_dz2 = 1.0 / p.dz2;
    m.halfx = ((p.nx - 1) / 2.0) * p.dx;   // rhs.f90:395-396
    m.halfy = ((p.ny - 1) / 2.0) * p.dy;
    Physics& ph = c->phys;
    ph.eta = p.eta;
    ph.rho = p.rho;
    ph.kappa = p.kappa;
    ph.gamma_r = p.gamma_r;
    ph.aw = p.aw;
    ph.inv2rho = 1.0 / (2.0 * p.rho);
    ph.inv_dV3 = 1.0 / (p.dx * p.dy * p.dz2);
    ph.c_dx = 2 * p.rho * p.kappa / std::sqrt(p.eta);
    ph.inv_eta = 1.0 / p.eta;
    ph.inv_gr2 = 1.0 / (p.gamma_r * p.gamma_r);
    ph.coupling = p.coupling;
    ph.electrons_evolve = p.electrons_evolve;
    ph.field_evolve = p.field_evolve;

    if (c->precision == PUFF_FP64) PUFF_TRY(alloc_particles<double>(c));
    else PUFF_TRY(alloc_particles<float>(c));

    PUFF_TRY(dalloc(c, (void**)&c->flags, sizeof(int) * FLAG_WORDS));
    PUFF_TRY(dalloc(c, (void**)&c->box, sizeof(Box) * BOX_COPIES));
    PUFF_HIP(hipHostMalloc((void**)&c->h_flags, sizeof(int) * FLAG_WORDS, hipHostMallocDefault));
    c->stage_len = 2 * std::max<int64_t>(c->n, c->nmesh());
    PUFF_TRY(dalloc(c, (void**)&c->stage_buf, sizeof(double) * c->stage_len));
    c->diag_len = 3 * 4096 + c->p.nz2;
    PUFF_TRY(dalloc(c, (void**)&c->diag, sizeof(double) * c->diag_len));
    if (!p.one_d) {
        std::vector<double> kx, ky, kz, mxy, mz;
        kvalues(p, kx, ky, kz);
        masks(p, mxy, mz);
        PUFF_TRY(dalloc(c, (void**)&c->kx, sizeof(double) * kx.size()));
        PUFF_TRY(dalloc(c, (void**)&c->ky, sizeof(double) * ky.size()));
        PUFF_TRY(dalloc(c, (void**)&c->kz, sizeof(double) * kz.size()));
        PUFF_TRY(dalloc(c, (void**)&c->mxy, sizeof(double) * mxy.size()));
        PUFF_TRY(dalloc(c, (void**)&c->mz, sizeof(double) * mz.size()));
        PUFF_HIP(hipMemcpy(c->kx, kx.data(), sizeof(double) * kx.size(), hipMemcpyHostToDevice));
        PUFF_HIP(hipMemcpy(c->ky, ky.data(), sizeof(double) * ky.size(), hipMemcpyHostToDevice));
        PUFF_HIP(hipMemcpy(c->kz, kz.data(), sizeof(double) * kz.size(), hipMemcpyHostToDevice));
        std::vector<double> kzr(2 * kz.size(), 0.0);
        for (size_t i = 0; i < kz.size(); ++i) {
            if (kz[i] != 0.0) {
                kzr[2 * i] = 1.0 / (2.0 * kz[i]);
                kzr[2 * i + 1] = 1.0 / std::sqrt(std::fabs(2.0 * kz[i]));
            }
        }
        PUFF_TRY(dalloc(c, (void**)&c->kzr, sizeof(double) * kzr.size()));
        PUFF_HIP(hipMemcpy(c->kzr, kzr.data(), sizeof(double) * kzr.size(), hipMemcpyHostToDevice));
        PUFF_HIP(hipMemcpy(c->mxy, mxy.data(), sizeof(double) * mxy.size(), hipMemcpyHostToDevice));
        PUFF_HIP(hipMemcpy(c->mz, mz.data(), sizeof(double) * mz.size(), hipMemcpyHostToDevice));
    }
    return PUFF_OK;
}

// ---- kernel timing ------------------------------------------------------------

static hipEvent_t take_event(Ctx* c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

// Records the start event of a timed launch; returns its slot or -1.
static int prof_start(Ctx* c, int kind) {
    if (!c->prof) return -1;
    hipEvent_t a = take_event(c), b = take_event(c);
    if (!a || !b) return -1;
    (void)hipEventRecord(a, c->stream);
    c->pending.push_back({kind, a, b});
    return (int)c->pending.size() - 1;
}

static void prof_stop(Ctx* c, int slot) {
    if (slot < 0) return;
    (void)hipEventRecord(c->pending[(size_t)slot].b, c->stream);
}

// After a stream synchronisation: accumulate the recorded durations.
static void prof_collect(Ctx* c) {
    for (const auto& p : c->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->prof_ms[p.kind] += ms;
            c->prof_n[p.kind] += 1;
        }
    }
    c->pending.clear();
    c->ev_used = 0;
}

static void release(Ctx* c) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    c->ev_pool.clear();
    if (c->h_flags) (void)hipHostFree(c->h_flags);
    if (c->h_dn2) (void)hipHostFree(c->h_dn2);
    if (c->h_hdr) (void)hipHostFree(c->h_hdr);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
}

// ---- step capture ---------------------------------------------------------------
// A step's start state and stage data kept in caller-owned device memory for
// the adjoint: 256-byte header, x y z2 pr pi Gamma, chi, orig, A, then k1, k2,
// k3 (six arrays each) and dA1, dA2, dA3, each section 256-byte aligned.

template <typename R>
struct CaptureView {
    PartPtrs<R> y;
    R* chi;
    int* orig;
    cplx<R>* A0;
    KPtrs<R> k[3];
    cplx<R>* dA[3];
};

static size_t align256c(size_t v) { return (v + 255) & ~(size_t)255; }

// Pointers into the capture at base (null base: sizes only); returns its bytes.
template <typename R>
static size_t capture_view(const Ctx* c, char* base, CaptureView<R>* v) {
    size_t o = 256;
    const size_t n = (size_t)c->n, nm = (size_t)c->nmesh();
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + o : nullptr;
        o = align256c(o + bytes);
        return p;
    };
    auto take_parts = [&](PartPtrs<R>& q) {
        q.x = reinterpret_cast<R*>(take(sizeof(R) * n));
        q.y = reinterpret_cast<R*>(take(sizeof(R) * n));
        q.z2 = reinterpret_cast<double*>(take(sizeof(double) * n));
        q.pr = reinterpret_cast<R*>(take(sizeof(R) * n));
        q.pi = reinterpret_cast<R*>(take(sizeof(R) * n));
        q.g = reinterpret_cast<double*>(take(sizeof(double) * n));
    };
    auto take_k = [&](KPtrs<R>& q) {
        q.x = reinterpret_cast<R*>(take(sizeof(R) * n));
        q.y = reinterpret_cast<R*>(take(sizeof(R) * n));
        q.z2 = reinterpret_cast<R*>(take(sizeof(R) * n));
        q.pr = reinterpret_cast<R*>(take(sizeof(R) * n));
        q.pi = reinterpret_cast<R*>(take(sizeof(R) * n));
        q.g = reinterpret_cast<R*>(take(sizeof(R) * n));
    };
    take_parts(v->y);
    v->chi = reinterpret_cast<R*>(take(sizeof(R) * n));
    v->orig = reinterpret_cast<int*>(take(sizeof(int) * n));
    v->A0 = reinterpret_cast<cplx<R>*>(take(sizeof(cplx<R>) * nm));
    for (int k = 0; k < 3; ++k) take_k(v->k[k]);
    for (int k = 0; k < 3; ++k) v->dA[k] = reinterpret_cast<cplx<R>*>(take(sizeof(cplx<R>) * nm));
    return o;
}

// ---- the integrator step ----------------------------------------------------

template <typename R>
static void sort_particles(Ctx* c) {
    const int slot = prof_start(c, 11);
    Particles<R>& P = parts<R>(c);
    SortArgs<R> a{};
    a.n = c->n;
    a.ncell = c->ncell;
    a.mesh = c->mesh;
    a.y = &P.y;
    // spare state arrays: k[0]'s, with z2 and Gamma of the state's type
    KPtrs<R>& k0 = P.k[0];
    PartPtrs<R> spare{k0.x, k0.y, nullptr, k0.pr, k0.pi, nullptr};
    if constexpr (sizeof(R) == sizeof(double)) {
        spare.z2 = reinterpret_cast<double*>(k0.z2);
        spare.g = reinterpret_cast<double*>(k0.g);
    } else {
        spare.z2 = P.z2_spare;
        spare.g = P.g_spare;
    }
    a.spare = &spare;
    a.chi = &P.chi;
    R* spare_chi = reinterpret_cast<R*>(c->chi_spare);
    a.chi_spare = &spare_chi;
    a.orig = &c->orig;
    a.orig_spare = &c->orig_spare;
    a.key = c->sort_key;
    a.counts = c->sort_counts;
    a.cursor = c->sort_counts;
    a.perm = c->sort_perm;
    a.tile_sums = c->sort_tiles;
    launch_sort<R>(a, c->stream);
    // the sort swapped the state and spare pointers: the spares go back
    k0.x = spare.x;
    k0.y = spare.y;
    k0.pr = spare.pr;
    k0.pi = spare.pi;
    if constexpr (sizeof(R) == sizeof(double)) {
        k0.z2 = reinterpret_cast<R*>(spare.z2);
        k0.g = reinterpret_cast<R*>(spare.g);
    } else {
        P.z2_spare = spare.z2;
        P.g_spare = spare.g;
    }
    prof_stop(c, slot);
    c->chi_spare = spare_chi;
    c->sorted = true;
    c->identity_order = false;
    c->steps_since_sort = 0;
    c->order_id = ++c->order_counter;
}

struct BlobHeader;
template <typename R>
static void capture_start(Ctx* c, char* cap, void* cap_hdr, CaptureView<R>& v, const double* taper0);

template <typename R>
static int rk4par(Ctx* c, double sz, double h, char* cap = nullptr, void* cap_hdr = nullptr) {
    Particles<R>& P = parts<R>(c);
    hipStream_t s = c->stream;
    const double hh = h * 0.5, h6 = h / 6.0;   // puffin_mpi_RK4.f90:89-91
    const double zs[4] = {sz, sz + hh, sz + hh, sz + h};
    const double saved[3] = {c->n2col, c->n2col0, c->undgrad};
    StageScalars sc[4];
    for (int k = 0; k < 4; ++k) {
        const double n2 = advance_taper(c, zs[k]);
        const int st = stage_scalars(c->p, zs[k], n2, sc[k]);
        if (st != PUFF_OK) {
            c->n2col = saved[0];
            c->n2col0 = saved[1];
            return st;
        }
    }
    if (c->sort_every > 0 && (!c->sorted || c->sort_due || c->steps_since_sort >= c->sort_every)) {
        sort_particles<R>(c);
    }
    c->steps_since_sort++;
    CaptureView<R> v{};
    if (cap) capture_start<R>(c, cap, cap_hdr, v, saved);
    int slot = prof_start(c, 6);
    launch_step_begin(c->flags, c->box, s);
    DetState det{};
    if constexpr (sizeof(R) == 8) {
        if (c->det) {
            launch_det_scale(P.y, P.chi, c->n, c->phys, c->dets, s);
            det = c->dets;
        }
    }
    prof_stop(c, slot);
    StageArgs<R> a{};
    a.det = det;
    a.n = c->n;
    a.y0 = P.y;
    a.chi = P.chi;
    a.h6 = h6;
    a.A0 = reinterpret_cast<const cplx<R>*>(c->A);
    a.mesh = c->mesh;
    a.phys = c->phys;
    a.flags = c->flags;
    a.box = c->box;
    const double cin[4] = {0.0, hh, hh, h};
    const double cnext[4] = {hh, hh, h, 0.0};
    // stage derivatives and sources: the context's buffers, or the capture's
    KPtrs<R> K[3] = {P.k[0], P.k[1], P.k[2]};
    cplx<R>* D[4];
    for (int k = 0; k < 4; ++k) D[k] = reinterpret_cast<cplx<R>*>(c->dA[k]);
    if (cap) {
        for (int k = 0; k < 3; ++k) {
            K[k] = v.k[k];
            D[k] = v.dA[k];
        }
    }
    for (int st = 1; st <= 4; ++st) {
        a.s = sc[st - 1];
        a.cin = cin[st - 1];
        a.cnext = cnext[st - 1];
        a.kin = (st > 1) ? K[st - 2] : KPtrs<R>{};
        a.kout = (st < 4) ? K[st - 1] : KPtrs<R>{};
        if (st == 4) {
            a.k1 = K[0];
            a.k2 = K[1];
        }
        a.dAin = (st > 1) ? D[st - 2] : nullptr;
        a.dAout = D[st - 1];
        if (cap && st == 1) {
            a.ycap = v.y;
            a.chicap = v.chi;
        } else {
            a.ycap = PartPtrs<R>{};
            a.chicap = nullptr;
        }
        slot = prof_start(c, st - 1);
        launch_stage<R>(st, a, s);
        if constexpr (sizeof(R) == 8) {
            if (det.acc)
                launch_det_convert(det, D[st - 1], c->box, c->p.nx, c->p.ny, c->nzp(), c->p.periodic,
                                   c->p.one_d, s);
        }
        prof_stop(c, slot);
    }
    CombineArgs<R> ca{};
    ca.A = reinterpret_cast<cplx<R>*>(c->A);
    for (int k = 0; k < 4; ++k) ca.d[k] = D[k];
    ca.clear_from = cap ? 3 : 0;   // the capture keeps dA1..dA3
    ca.h6 = h6;
    ca.box = c->box;
    ca.flags = c->flags;
    ca.nx = c->p.nx;
    ca.ny = c->p.ny;
    ca.nz2 = c->p.nz2;
    ca.nzp = c->nzp();
    ca.periodic = c->p.periodic;
    ca.one_d = c->p.one_d;
    slot = prof_start(c, 4);
    launch_combine<R>(ca, s);
    prof_stop(c, slot);
    PUFF_HIP(hipGetLastError());
    if (c->deferred) return PUFF_OK;   // flags collected by puff_deferred_status
    PUFF_HIP(hipMemcpyAsync(c->h_flags, c->flags, sizeof(int) * FLAG_WORDS, hipMemcpyDeviceToHost, s));
    PUFF_HIP(hipStreamSynchronize(s));
    prof_collect(c);
    const int f = c->h_flags[0];
    if (c->sort_every > 0 && c->p.field_evolve) {
        // particles drift out of their sorted runs (faster once the beam's
        // energy spread has grown): re-sort when the share of workgroups that
        // could not use the LDS tile exceeds twice its value after the sort
        // (measured on the bunched C1 beam: sorting at that point beats
        // letting the share grow until the fallback has cost one sort)
        const double groups = 4.0 * (double)((c->n + STAGE_THREADS - 1) / STAGE_THREADS);
        const double share = groups > 0 ? (double)c->h_flags[FLAG_WORD_FALLBACK] / groups : 0.0;
        c->fallback_last = share;
        if (c->steps_since_sort == 1) c->fallback_base = share;
        c->sort_due = share > std::max(PUFF_RESORT_MIN_SHARE, PUFF_RESORT_GROWTH * c->fallback_base);
    }
    if (f) {
        c->n2col = saved[0];
        c->n2col0 = saved[1];
        if (f & FLAG_DET_RANGE) {
            set_error("a source exceeded the range the deterministic deposit was scaled for");
            return PUFF_ERR_STATE;
        }
        if (f & FLAG_Z2MESH) return PUFF_OUT_Z2_MESH;
        if (f & FLAG_TRANSVERSE) return PUFF_OUT_TRANSVERSE;
        return PUFF_OUT_Z2_BUFFER;
    }
    return PUFF_OK;
}

template <typename R>
static int diffraction(Ctx* c, double h) {
    if (c->p.one_d) {
        set_error("no diffraction on a 1-D mesh (read_base_input.f90:389)");
        return PUFF_ERR_STATE;
    }
    DiffArgs<R> a{};
    a.a = reinterpret_cast<cplx<R>*>(c->A);
    a.anb = reinterpret_cast<cplx<R>*>(c->anb);
    a.kx = c->kx;
    a.ky = c->ky;
    a.kz = c->kz;
    a.kzr = c->kzr;
    a.mxy = c->mxy;
    a.mz = c->mz;
    a.nx = c->p.nx;
    a.ny = c->p.ny;
    a.nz2 = c->p.nz2;
    a.nzp = c->nzp();
    a.periodic = c->p.periodic;
    a.filter = c->p.filter;
    a.h = h;
    a.beta = c->p.beta;
    a.cutoff = c->p.filt_cutoff;
    a.ffact = (double)c->p.nx * (double)c->p.ny * (double)a.nzp;   // setup.f90:334-346
    a.ax = c->ax;
    a.ay = c->ay;
    a.az = c->az;
    const int slot = prof_start(c, 5);
    launch_diffraction<R>(a, c->stream);
    prof_stop(c, slot);
    PUFF_HIP(hipGetLastError());
    PUFF_HIP(hipStreamSynchronize(c->stream));
    prof_collect(c);
    return PUFF_OK;
}

// ---- data exchange ------------------------------------------------------------

// One coordinate array from the reference order (FP64) into the slot order
// and storage type T of the device.
template <typename T>
static int upload_array(Ctx* c, int kind, const double* src, T* dst) {
    if (!src) return PUFF_OK;
    const hipMemcpyKind k = kind == PUFF_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (c->identity_order && sizeof(T) == sizeof(double)) {
        PUFF_HIP(hipMemcpyAsync(dst, src, sizeof(double) * c->n, k, c->stream));
    } else {
        PUFF_HIP(hipMemcpyAsync(c->stage_buf, src, sizeof(double) * c->n, k, c->stream));
        if (c->identity_order) launch_convert<T>(c->stage_buf, dst, c->n, c->stream);
        else launch_permute_in<T>(c->stage_buf, c->orig, c->n, dst, c->stream);
        PUFF_HIP(hipGetLastError());
    }
    PUFF_HIP(hipStreamSynchronize(c->stream));
    return PUFF_OK;
}

// One coordinate array from the slot order back to the reference order (FP64).
template <typename T>
static int download_array(Ctx* c, int kind, const T* src, double* dst) {
    if (!dst) return PUFF_OK;
    const hipMemcpyKind k = kind == PUFF_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (c->identity_order && sizeof(T) == sizeof(double)) {
        PUFF_HIP(hipMemcpyAsync(dst, src, sizeof(double) * c->n, k, c->stream));
    } else {
        if (c->identity_order) launch_convert_back<T>(src, c->stage_buf, c->n, c->stream);
        else launch_unpermute<T>(src, c->orig, c->n, c->stage_buf, c->stream);
        PUFF_HIP(hipGetLastError());
        PUFF_HIP(hipMemcpyAsync(dst, c->stage_buf, sizeof(double) * c->n, k, c->stream));
    }
    PUFF_HIP(hipStreamSynchronize(c->stream));
    return PUFF_OK;
}

template <typename R>
static int upload_particles(Ctx* c, int kind, const double* x, const double* y, const double* z2,
                            const double* px, const double* py, const double* g, const double* chi) {
    Particles<R>& P = parts<R>(c);
    if (x && y && z2 && px && py && g && chi) {
        // a complete state: back to the reference order, sorted on the next step
        launch_iota(c->orig, c->n, c->stream);
        PUFF_HIP(hipGetLastError());
        c->identity_order = true;
        c->sorted = false;
        c->order_id = ++c->order_counter;
    }
    PUFF_TRY(upload_array<R>(c, kind, x, P.y.x));
    PUFF_TRY(upload_array<R>(c, kind, y, P.y.y));
    PUFF_TRY(upload_array<double>(c, kind, z2, P.y.z2));
    PUFF_TRY(upload_array<R>(c, kind, px, P.y.pr));
    PUFF_TRY(upload_array<R>(c, kind, py, P.y.pi));
    PUFF_TRY(upload_array<double>(c, kind, g, P.y.g));
    PUFF_TRY(upload_array<R>(c, kind, chi, P.chi));
    return PUFF_OK;
}

template <typename R>
static int download_particles(Ctx* c, int kind, double* x, double* y, double* z2, double* px,
                              double* py, double* g) {
    Particles<R>& P = parts<R>(c);
    PUFF_TRY(download_array<R>(c, kind, P.y.x, x));
    PUFF_TRY(download_array<R>(c, kind, P.y.y, y));
    PUFF_TRY(download_array<double>(c, kind, P.y.z2, z2));
    PUFF_TRY(download_array<R>(c, kind, P.y.pr, px));
    PUFF_TRY(download_array<R>(c, kind, P.y.pi, py));
    PUFF_TRY(download_array<double>(c, kind, P.y.g, g));
    return PUFF_OK;
}

template <typename R>
static int upload_field(Ctx* c, int kind, const double* ar, const double* ai) {
    const int64_t nm = c->nmesh();
    const hipMemcpyKind k = kind == PUFF_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    PUFF_HIP(hipMemcpyAsync(c->stage_buf, ar, sizeof(double) * nm, k, c->stream));
    PUFF_HIP(hipMemcpyAsync(c->stage_buf + nm, ai, sizeof(double) * nm, k, c->stream));
    launch_pack_field<R>(c->stage_buf, c->stage_buf + nm, reinterpret_cast<cplx<R>*>(c->A), nm, c->stream);
    PUFF_HIP(hipGetLastError());
    if (c->p.periodic) {
        // the last plane of a periodic mesh aliases plane 1 (para_field.f90:1075-1138)
        const int64_t ntr = (int64_t)c->p.nx * c->p.ny;
        cplx<R>* A = reinterpret_cast<cplx<R>*>(c->A);
        PUFF_HIP(hipMemcpyAsync(A + (int64_t)(c->p.nz2 - 1) * ntr, A, sizeof(cplx<R>) * ntr,
                                hipMemcpyDeviceToDevice, c->stream));
    }
    PUFF_HIP(hipStreamSynchronize(c->stream));
    return PUFF_OK;
}

template <typename R>
static int download_field(Ctx* c, int kind, double* ar, double* ai) {
    const int64_t nm = c->nmesh();
    const hipMemcpyKind k = kind == PUFF_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    launch_unpack_field<R>(reinterpret_cast<const cplx<R>*>(c->A), c->stage_buf, c->stage_buf + nm, nm, c->stream);
    PUFF_HIP(hipGetLastError());
    if (ar) PUFF_HIP(hipMemcpyAsync(ar, c->stage_buf, sizeof(double) * nm, k, c->stream));
    if (ai) PUFF_HIP(hipMemcpyAsync(ai, c->stage_buf + nm, sizeof(double) * nm, k, c->stream));
    PUFF_HIP(hipStreamSynchronize(c->stream));
    return PUFF_OK;
}

template <typename R>
static int transform(Ctx* c, int sign) {
    if (c->p.one_d) return PUFF_ERR_STATE;
    launch_fft3d<R>(reinterpret_cast<cplx<R>*>(c->A), reinterpret_cast<cplx<R>*>(c->anb), c->p.nx, c->p.ny,
                    c->nzp(), c->ax, c->ay, c->az, sign, c->stream);
    PUFF_HIP(hipGetLastError());
    PUFF_HIP(hipStreamSynchronize(c->stream));
    return PUFF_OK;
}

template <typename R>
static int power(Ctx* c, double* out) {
    launch_power<R>(reinterpret_cast<const cplx<R>*>(c->A), c->p.nx, c->p.ny, c->p.nz2, c->p.dx, c->p.dy,
                    c->diag, c->stream);
    PUFF_HIP(hipGetLastError());
    PUFF_HIP(hipMemcpyAsync(out, c->diag, sizeof(double) * c->p.nz2, hipMemcpyDeviceToHost, c->stream));
    PUFF_HIP(hipStreamSynchronize(c->stream));
    return PUFF_OK;
}

template <typename R>
static int bunching(Ctx* c, int harmonic, int nslices, double slice, double* out) {
    if (nslices < 1 || nslices > 2048) {
        set_error("nslices must be in [1, 2048]");
        return PUFF_ERR_ARG;
    }
    Particles<R>& P = parts<R>(c);
    launch_fill_zero(c->diag, 3 * (int64_t)nslices, c->stream);
    launch_bunching<R>(P.y.z2, P.chi, c->n, c->p.rho, harmonic, nslices, slice, c->p.periodic, c->diag,
                       c->stream);
    PUFF_HIP(hipGetLastError());
    std::vector<double> acc(3 * (size_t)nslices);
    PUFF_HIP(hipMemcpyAsync(acc.data(), c->diag, sizeof(double) * acc.size(), hipMemcpyDeviceToHost, c->stream));
    PUFF_HIP(hipStreamSynchronize(c->stream));
    for (int is = 0; is < nslices; ++is) {
        const double br = acc[is], bi = acc[nslices + is], sd = acc[2 * nslices + is];
        out[is] = (sd > 0.0) ? std::sqrt(br * br + bi * bi) / slice : 0.0;   // av_write.f90:725-749
    }
    return PUFF_OK;
}

// getSliceTwiss's per-slice outputs (av_write.f90:664-731) from the device
// sums: out[k * nslices + s] for k = 0 charge, 1-2 mean x, y, 3-4 mean px,
// py, 5-6 r.m.s. x, y, 7-8 r.m.s. px, py, 9-10 emittances, 11 mean Gamma,
// 12 r.m.s. Gamma; negative variances clipped to zero.
template <typename R>
static int slice_stats(Ctx* c, int nslices, double slice, double* out) {
    if (nslices < 1 || (int64_t)NSLICE_SUMS * nslices > c->diag_len || !(slice > 0.0)) {
        set_error("nslices must be in [1, " + std::to_string(c->diag_len / NSLICE_SUMS) + "] and slice_len > 0");
        return PUFF_ERR_ARG;
    }
    Particles<R>& P = parts<R>(c);
    launch_fill_zero(c->diag, (int64_t)NSLICE_SUMS * nslices, c->stream);
    launch_slice_moments<R>(P.y, P.chi, c->n, nslices, slice, c->p.periodic, c->diag, c->stream);
    PUFF_HIP(hipGetLastError());
    std::vector<double> a((size_t)NSLICE_SUMS * nslices);
    PUFF_HIP(hipMemcpyAsync(a.data(), c->diag, sizeof(double) * a.size(), hipMemcpyDeviceToHost, c->stream));
    PUFF_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 13 * nslices; ++k) out[k] = 0.0;
    auto root = [](double v) { return std::sqrt(v < 0.0 ? 0.0 : v); };
    for (int is = 0; is < nslices; ++is) {
        auto A = [&](int k) { return a[(size_t)k * nslices + is]; };
        const double sd = A(0);
        out[is] = sd;
        if (!(sd > 0.0)) continue;
        const double ax = A(1) / sd, ay = A(2) / sd, apx = A(3) / sd, apy = A(4) / sd;
        const double vx = A(6) / sd - std::pow(A(1) / sd, 2.0), vy = A(11) / sd - std::pow(A(2) / sd, 2.0);
        const double vpx = A(8) / sd - std::pow(A(3) / sd, 2.0), vpy = A(9) / sd - std::pow(A(4) / sd, 2.0);
        const double vg = A(12) / sd - std::pow(A(5) / sd, 2.0);
        const double ex = vx * vpx - std::pow(A(7) / sd - ax * apx, 2.0);
        const double ey = vy * vpy - std::pow(A(10) / sd - ay * apy, 2.0);
        double* o = out + is;
        o[1 * nslices] = ax;
        o[2 * nslices] = ay;
        o[3 * nslices] = apx;
        o[4 * nslices] = apy;
        o[5 * nslices] = root(vx);
        o[6 * nslices] = root(vy);
        o[7 * nslices] = root(vpx);
        o[8 * nslices] = root(vpy);
        o[9 * nslices] = root(ex);
        o[10 * nslices] = root(ey);
        o[11 * nslices] = A(5) / sd;
        o[12 * nslices] = root(vg);
    }
    return PUFF_OK;
}

// getCurr's binned charge (av_write.f90:297-368) before its unit scalings.
template <typename R>
static int current(Ctx* c, int npts, double sam, double* out) {
    if (npts < 2 || npts + 1 > c->diag_len || !(sam > 0.0)) {
        set_error("npts must be in [2, " + std::to_string(c->diag_len - 1) + "] and sam_len > 0");
        return PUFF_ERR_ARG;
    }
    Particles<R>& P = parts<R>(c);
    launch_fill_zero(c->diag, (int64_t)npts + 1, c->stream);
    launch_current<R>(P.y.z2, P.chi, c->n, npts, sam, c->p.periodic, c->diag, c->stream);
    PUFF_HIP(hipGetLastError());
    std::vector<double> a((size_t)npts + 1);
    PUFF_HIP(hipMemcpyAsync(a.data(), c->diag, sizeof(double) * a.size(), hipMemcpyDeviceToHost, c->stream));
    PUFF_HIP(hipStreamSynchronize(c->stream));
    if (a[(size_t)npts] > 0.0) {
        set_error(std::to_string((int64_t)a[(size_t)npts]) + " macroparticles lie outside the current nodes");
        return PUFF_ERR_STATE;
    }
    for (int k = 0; k < npts; ++k) out[k] = a[(size_t)k];
    return PUFF_OK;
}

// ---- discrete adjoint ---------------------------------------------------------

static int need_adj(const Ctx* c) {
    if (!c->adj) {
        set_error("adjoint buffers not allocated (puff_adjoint_enable)");
        return PUFF_ERR_STATE;
    }
    return PUFF_OK;
}

template <typename R>
static int adjoint_zero(Ctx* c);

template <typename R>
static int adjoint_enable(Ctx* c) {
    if (c->adj) return PUFF_OK;
    AdjBufs<R>& B = adjb<R>(c);
    auto setp = [&](PartPtrs<R>& q) -> int {
        PUFF_TRY(dalloc(c, (void**)&q.x, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.y, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.z2, sizeof(double) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.pr, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.pi, sizeof(R) * c->n));
        PUFF_TRY(dalloc(c, (void**)&q.g, sizeof(double) * c->n));
        return PUFF_OK;
    };
    PUFF_TRY(setp(B.L));
    for (int k = 0; k < 3; ++k) PUFF_TRY(setp(B.M[k]));
    const int64_t nm = c->nmesh();
    PUFF_TRY(dalloc(c, &c->LA, sizeof(cplx<R>) * nm));
    for (int k = 0; k < 4; ++k) PUFF_TRY(dalloc(c, &c->mA[k], sizeof(cplx<R>) * nm));
    PUFF_TRY(dalloc(c, (void**)&c->dn2, sizeof(double) * 4 * DN2_COPIES));
    PUFF_TRY(dalloc(c, (void**)&c->adj_wz, sizeof(double) * c->p.nz2));
    PUFF_TRY(dalloc(c, (void**)&c->lam_orig, sizeof(int) * c->n));
    PUFF_TRY(dalloc(c, (void**)&c->dn2_table, sizeof(double) * (DN2_TABLE_CAP + 1)));
    PUFF_HIP(hipHostMalloc((void**)&c->h_dn2, sizeof(double) * 4 * DN2_COPIES, hipHostMallocDefault));
    c->adj = true;
    return adjoint_zero<R>(c);
}

// Zero-fill of a typed array with a fill kernel (no memset nodes).
template <typename T>
static void fill_zero_typed(T* p, int64_t n, hipStream_t s) {
    static_assert(sizeof(T) % 4 == 0, "4-byte multiple");
    launch_fill_zero_words(reinterpret_cast<uint32_t*>(p), (int64_t)(sizeof(T) / 4) * n, s);
}

// Re-index the particle cotangent from slot order orig_from to slot order
// orig_to (reference index of each slot), through the reference order.
template <typename R>
static int repermute_cotangent(Ctx* c, const int* orig_from, const int* orig_to) {
    AdjBufs<R>& B = adjb<R>(c);
    hipStream_t s = c->stream;
    R* rr[4] = {B.L.x, B.L.y, B.L.pr, B.L.pi};
    double* dd[2] = {B.L.z2, B.L.g};
    for (R* a : rr) {
        launch_unpermute<R>(a, orig_from, c->n, c->stage_buf, s);
        launch_permute_in<R>(c->stage_buf, orig_to, c->n, a, s);
    }
    for (double* a : dd) {
        launch_unpermute<double>(a, orig_from, c->n, c->stage_buf, s);
        launch_permute_in<double>(c->stage_buf, orig_to, c->n, a, s);
    }
    PUFF_HIP(hipGetLastError());
    return PUFF_OK;
}

// The particle cotangent keeps its own slot order (lam_orig): a sort or a
// state load changes the state's order only.  Bring it to the state's order.
template <typename R>
static int sync_cotangent_order(Ctx* c) {
    if (c->lam_order_id == c->order_id) return PUFF_OK;
    PUFF_TRY(repermute_cotangent<R>(c, c->lam_orig, c->orig));
    PUFF_HIP(hipMemcpyAsync(c->lam_orig, c->orig, sizeof(int) * c->n, hipMemcpyDeviceToDevice, c->stream));
    c->lam_order_id = c->order_id;
    return PUFF_OK;
}

template <typename R>
static int adjoint_zero(Ctx* c) {
    PUFF_TRY(need_adj(c));
    AdjBufs<R>& B = adjb<R>(c);
    const int64_t n = c->n, nm = c->nmesh();
    fill_zero_typed<R>(B.L.x, n, c->stream);
    fill_zero_typed<R>(B.L.y, n, c->stream);
    fill_zero_typed<R>(B.L.pr, n, c->stream);
    fill_zero_typed<R>(B.L.pi, n, c->stream);
    fill_zero_typed<double>(B.L.z2, n, c->stream);
    fill_zero_typed<double>(B.L.g, n, c->stream);
    fill_zero_typed<cplx<R>>(reinterpret_cast<cplx<R>*>(c->LA), nm, c->stream);
    fill_zero_typed<double>(c->dn2_table, DN2_TABLE_CAP + 1, c->stream);
    PUFF_HIP(hipGetLastError());
    PUFF_HIP(hipMemcpyAsync(c->lam_orig, c->orig, sizeof(int) * n, hipMemcpyDeviceToDevice, c->stream));
    c->lam_order_id = c->order_id;
    PUFF_HIP(hipStreamSynchronize(c->stream));
    return PUFF_OK;
}

template <typename R>
static int adjoint_seed_power(Ctx* c, const double* wz, double scale) {
    PUFF_TRY(need_adj(c));
    PUFF_HIP(hipMemcpyAsync(c->adj_wz, wz, sizeof(double) * c->p.nz2, hipMemcpyHostToDevice, c->stream));
    launch_adj_seed_power<R>(reinterpret_cast<cplx<R>*>(c->LA), reinterpret_cast<const cplx<R>*>(c->A),
                             c->adj_wz, c->p.nx, c->p.ny, c->p.nz2, c->p.periodic, c->p.dx, c->p.dy,
                             scale, c->stream);
    PUFF_HIP(hipGetLastError());
    PUFF_HIP(hipStreamSynchronize(c->stream));
    return PUFF_OK;
}

// The aliased last plane of a periodic mesh copies plane 1: its cotangent is
// folded into plane 1.
template <typename R>
static int fold_alias(Ctx* c) {
    if (!c->p.periodic) return PUFF_OK;
    launch_fold_alias<R>(reinterpret_cast<cplx<R>*>(c->LA), (int64_t)c->p.nx * c->p.ny, c->p.nz2, c->stream);
    PUFF_HIP(hipGetLastError());
    return PUFF_OK;
}

// Adjoint of one rk4par step: the current state is the state at the start of
// the step (left unchanged), L / LA the cotangent of the end state (replaced
// by the cotangent of the start state).  dn2col[k] = dJ/d(n2col of stage k),
// slot[k] its per-period table slot (-1 without a table).
// Forward stages 1-3 of the step starting at the current state again: k_1..k_3
// and dA_1..dA_3 in the context's buffers (the adjoint step without a capture).
template <typename R>
static int replay_stages(Ctx* c, const StageScalars* sc, const double* cin, const double* cnext, double h6) {
    Particles<R>& P = parts<R>(c);
    hipStream_t s = c->stream;
    const int slot_p = prof_start(c, 10);
    StageArgs<R> a{};
    a.n = c->n;
    a.y0 = P.y;
    a.chi = P.chi;
    a.h6 = h6;
    a.A0 = reinterpret_cast<const cplx<R>*>(c->A);
    a.mesh = c->mesh;
    a.phys = c->phys;
    a.flags = c->flags;
    a.box = c->box;
    for (int st = 1; st <= 3; ++st) {
        a.s = sc[st - 1];
        a.cin = cin[st - 1];
        a.cnext = cnext[st - 1];
        a.kin = (st > 1) ? P.k[st - 2] : KPtrs<R>{};
        a.kout = P.k[st - 1];
        a.dAin = (st > 1) ? reinterpret_cast<const cplx<R>*>(c->dA[st - 2]) : nullptr;
        a.dAout = reinterpret_cast<cplx<R>*>(c->dA[st - 1]);
        launch_stage<R>(st, a, s);
    }
    prof_stop(c, slot_p);
    PUFF_HIP(hipGetLastError());
    return PUFF_OK;
}

template <typename R>
static int capture_check(Ctx* c, const void* cap_hdr, double taper[3], int64_t* order_id);

template <typename R>
static int rk4par_adjoint(Ctx* c, double sz, double h, double* dn2col, int32_t* slot,
                          const char* cap = nullptr, const void* cap_hdr = nullptr) {
    PUFF_TRY(need_adj(c));
    Particles<R>& P = parts<R>(c);
    AdjBufs<R>& B = adjb<R>(c);
    hipStream_t s = c->stream;
    const double hh = h * 0.5, h6 = h / 6.0;
    const double zs[4] = {sz, sz + hh, sz + hh, sz + h};
    const double saved[3] = {c->n2col, c->n2col0, c->undgrad};
    CaptureView<R> v{};
    int64_t cap_order = 0;
    if (cap) {
        // the captured step's taper state replaces the context's for the
        // stage coefficients (restored below)
        double tp[3];
        PUFF_TRY(capture_check<R>(c, cap_hdr, tp, &cap_order));
        capture_view<R>(c, const_cast<char*>(cap), &v);
        c->n2col = tp[0];
        c->n2col0 = tp[1];
        c->undgrad = tp[2];
    }
    StageScalars sc[4];
    int st_err = PUFF_OK;
    for (int k = 0; k < 4 && st_err == PUFF_OK; ++k) {
        const double n2 = advance_taper(c, zs[k]);
        st_err = stage_scalars(c->p, zs[k], n2, sc[k]);
        if (slot) slot[k] = taper_slot(c, zs[k]);
    }
    c->n2col = saved[0];
    c->n2col0 = saved[1];
    c->undgrad = saved[2];
    if (st_err != PUFF_OK) return st_err;
    if (c->deferred && c->taper_table.size() > (size_t)DN2_TABLE_CAP) {
        set_error("taper table longer than the deferred adjoint's gradient table");
        return PUFF_ERR_ARG;
    }
    if (cap) {
        // the cotangent follows the captured step's slot order
        if (c->lam_order_id != cap_order) {
            PUFF_TRY(repermute_cotangent<R>(c, c->lam_orig, v.orig));
            PUFF_HIP(hipMemcpyAsync(c->lam_orig, v.orig, sizeof(int) * c->n, hipMemcpyDeviceToDevice, s));
            c->lam_order_id = cap_order;
        }
    } else {
        PUFF_TRY(sync_cotangent_order<R>(c));
    }
    PUFF_TRY(fold_alias<R>(c));
    launch_fill_zero(c->dn2, 4 * DN2_COPIES, s);
    launch_step_begin(c->flags, c->box, s);
    const double cin[4] = {0.0, hh, hh, h};
    const double cnext[4] = {hh, hh, h, 0.0};
    // start state and stage data: the capture's, or the context's state with
    // forward stages 1-3 replayed (k_1..k_3, dA_1..dA_3)
    PartPtrs<R> Y0 = P.y;
    const R* CHI = P.chi;
    const cplx<R>* A0 = reinterpret_cast<const cplx<R>*>(c->A);
    KPtrs<R> K[3] = {P.k[0], P.k[1], P.k[2]};
    const cplx<R>* D[3];
    for (int k = 0; k < 3; ++k) D[k] = reinterpret_cast<const cplx<R>*>(c->dA[k]);
    if (cap) {
        Y0 = v.y;
        CHI = v.chi;
        A0 = v.A0;
        for (int k = 0; k < 3; ++k) {
            K[k] = v.k[k];
            D[k] = v.dA[k];
        }
    } else {
        PUFF_TRY(replay_stages<R>(c, sc, cin, cnext, h6));
    }
    int slot_p;
    // adjoint stages 4 .. 1
    const double wts[4] = {1.0, 2.0, 2.0, 1.0};
    AdjArgs<R> b{};
    b.n = c->n;
    b.y0 = Y0;
    b.chi = CHI;
    b.A0 = A0;
    b.L = B.L;
    b.LA = reinterpret_cast<const cplx<R>*>(c->LA);
    b.mesh = c->mesh;
    b.phys = c->phys;
    b.flags = c->flags;
    b.box = c->box;
    slot_p = prof_start(c, 7);
    for (int st = 4; st >= 1; --st) {
        b.s = sc[st - 1];
        b.cin = cin[st - 1];
        b.kin = (st > 1) ? K[st - 2] : KPtrs<R>{};
        b.dAin = (st > 1) ? D[st - 2] : nullptr;
        b.ws = wts[st - 1] * h6;
        b.cnext = (st < 4) ? cin[st] : 0.0;
        b.mnext = (st < 4) ? B.M[st - 1] : PartPtrs<R>{};
        b.mAnext = (st < 4) ? reinterpret_cast<const cplx<R>*>(c->mA[st]) : nullptr;
        b.mout = (st > 1) ? B.M[st - 2] : PartPtrs<R>{};
        b.m3 = B.M[1];
        b.m4 = B.M[2];
        b.mAout = reinterpret_cast<cplx<R>*>(c->mA[st - 1]);
        b.dn2 = c->dn2 + (st - 1) * DN2_COPIES;
        launch_adj_stage<R>(st, b, s);
    }
    prof_stop(c, slot_p);
    AdjCombineArgs<R> ca{};
    ca.LA = reinterpret_cast<cplx<R>*>(c->LA);
    for (int k = 0; k < 4; ++k) ca.mA[k] = reinterpret_cast<cplx<R>*>(c->mA[k]);
    for (int k = 0; k < 3; ++k) ca.dA[k] = reinterpret_cast<cplx<R>*>(c->dA[k]);
    ca.box = c->box;
    ca.flags = c->flags;
    ca.nx = c->p.nx;
    ca.ny = c->p.ny;
    ca.nz2 = c->p.nz2;
    ca.nzp = c->nzp();
    ca.periodic = c->p.periodic;
    ca.one_d = c->p.one_d;
    ca.clear_dA = cap ? 0 : 1;
    slot_p = prof_start(c, 8);
    launch_adj_combine<R>(ca, s);
    prof_stop(c, slot_p);
    PUFF_HIP(hipGetLastError());
    if (c->deferred) {
        // dJ/dn2col stays on the device (puff_adjoint_taper_gradient)
        Slots4 sl;
        for (int k = 0; k < 4; ++k) sl.s[k] = taper_slot(c, zs[k]);
        launch_dn2_accumulate(c->dn2, sl, (int)c->taper_table.size(), c->dn2_table, s);
        PUFF_HIP(hipGetLastError());
        return PUFF_OK;
    }
    PUFF_HIP(hipMemcpyAsync(c->h_flags, c->flags, sizeof(int) * FLAG_WORDS, hipMemcpyDeviceToHost, s));
    PUFF_HIP(hipMemcpyAsync(c->h_dn2, c->dn2, sizeof(double) * 4 * DN2_COPIES, hipMemcpyDeviceToHost, s));
    PUFF_HIP(hipStreamSynchronize(s));
    prof_collect(c);
    if (dn2col) {
        for (int k = 0; k < 4; ++k) {
            double t = 0.0;
            for (int j = 0; j < DN2_COPIES; ++j) t += c->h_dn2[k * DN2_COPIES + j];
            dn2col[k] = t;
        }
    }
    const int f = c->h_flags[0];
    if (f) {
        set_error("a trial state of the replayed step left the mesh");
        if (f & FLAG_Z2MESH) return PUFF_OUT_Z2_MESH;
        if (f & FLAG_TRANSVERSE) return PUFF_OUT_TRANSVERSE;
        return PUFF_OUT_Z2_BUFFER;
    }
    return PUFF_OK;
}

template <typename R>
static int diffraction_adjoint(Ctx* c, double h) {
    PUFF_TRY(need_adj(c));
    if (c->p.one_d) {
        set_error("no diffraction on a 1-D mesh (read_base_input.f90:389)");
        return PUFF_ERR_STATE;
    }
    DiffArgs<R> a{};
    a.a = reinterpret_cast<cplx<R>*>(c->LA);
    a.anb = reinterpret_cast<cplx<R>*>(c->anb);
    a.kx = c->kx;
    a.ky = c->ky;
    a.kz = c->kz;
    a.kzr = c->kzr;
    a.mxy = c->mxy;
    a.mz = c->mz;
    a.nx = c->p.nx;
    a.ny = c->p.ny;
    a.nz2 = c->p.nz2;
    a.nzp = c->nzp();
    a.periodic = c->p.periodic;
    a.filter = c->p.filter;
    a.h = h;
    a.beta = c->p.beta;
    a.cutoff = c->p.filt_cutoff;
    a.ffact = (double)c->p.nx * (double)c->p.ny * (double)a.nzp;
    a.ax = c->ax;
    a.ay = c->ay;
    a.az = c->az;
    const int slot = prof_start(c, 9);
    launch_diffraction_adjoint<R>(a, c->stream);
    prof_stop(c, slot);
    PUFF_HIP(hipGetLastError());
    PUFF_HIP(hipStreamSynchronize(c->stream));
    prof_collect(c);
    return PUFF_OK;
}

template <typename R>
static int adjoint_upload(Ctx* c, int kind, const double* x, const double* y, const double* z2,
                          const double* px, const double* py, const double* g, const double* ar,
                          const double* ai) {
    PUFF_TRY(need_adj(c));
    PUFF_TRY(sync_cotangent_order<R>(c));
    AdjBufs<R>& B = adjb<R>(c);
    PUFF_TRY(upload_array<R>(c, kind, x, B.L.x));
    PUFF_TRY(upload_array<R>(c, kind, y, B.L.y));
    PUFF_TRY(upload_array<double>(c, kind, z2, B.L.z2));
    PUFF_TRY(upload_array<R>(c, kind, px, B.L.pr));
    PUFF_TRY(upload_array<R>(c, kind, py, B.L.pi));
    PUFF_TRY(upload_array<double>(c, kind, g, B.L.g));
    if (ar && ai) {
        const int64_t nm = c->nmesh();
        const hipMemcpyKind k = kind == PUFF_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        PUFF_HIP(hipMemcpyAsync(c->stage_buf, ar, sizeof(double) * nm, k, c->stream));
        PUFF_HIP(hipMemcpyAsync(c->stage_buf + nm, ai, sizeof(double) * nm, k, c->stream));
        launch_pack_field<R>(c->stage_buf, c->stage_buf + nm, reinterpret_cast<cplx<R>*>(c->LA), nm, c->stream);
        PUFF_HIP(hipGetLastError());
        PUFF_HIP(hipStreamSynchronize(c->stream));
    }
    return PUFF_OK;
}

template <typename R>
static int adjoint_download(Ctx* c, int kind, double* x, double* y, double* z2, double* px, double* py,
                            double* g, double* ar, double* ai) {
    PUFF_TRY(need_adj(c));
    PUFF_TRY(sync_cotangent_order<R>(c));
    AdjBufs<R>& B = adjb<R>(c);
    PUFF_TRY(download_array<R>(c, kind, B.L.x, x));
    PUFF_TRY(download_array<R>(c, kind, B.L.y, y));
    PUFF_TRY(download_array<double>(c, kind, B.L.z2, z2));
    PUFF_TRY(download_array<R>(c, kind, B.L.pr, px));
    PUFF_TRY(download_array<R>(c, kind, B.L.pi, py));
    PUFF_TRY(download_array<double>(c, kind, B.L.g, g));
    if (ar || ai) {
        const int64_t nm = c->nmesh();
        const hipMemcpyKind k = kind == PUFF_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        launch_unpack_field<R>(reinterpret_cast<const cplx<R>*>(c->LA), c->stage_buf, c->stage_buf + nm, nm,
                               c->stream);
        PUFF_HIP(hipGetLastError());
        if (ar) PUFF_HIP(hipMemcpyAsync(ar, c->stage_buf, sizeof(double) * nm, k, c->stream));
        if (ai) PUFF_HIP(hipMemcpyAsync(ai, c->stage_buf + nm, sizeof(double) * nm, k, c->stream));
        PUFF_HIP(hipStreamSynchronize(c->stream));
    }
    return PUFF_OK;
}

// ---- state snapshots ------------------------------------------------------------
// Blob: 256-byte header, then x, y, z2, pr, pi, Gamma, chi, orig (particle slot
// order) and the field, each section 256-byte aligned.

struct BlobHeader {
    uint64_t magic;
    int64_t n, nmesh;
    int32_t precision, sorted, identity_order, sort_due;
    int64_t order_id, steps_since_sort;
    double n2col, n2col0, undgrad;
};
constexpr uint64_t BLOB_MAGIC = 0x31424f4c42465550ull;   // "PUFBLOB1"
static_assert(sizeof(BlobHeader) <= PUFF_STATE_HEADER_BYTES && sizeof(BlobHeader) % 4 == 0, "blob header");

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

template <typename R>
static size_t blob_layout(const Ctx* c, size_t off[9]) {
    const size_t n = (size_t)c->n;
    const size_t sz[9] = {sizeof(R) * n, sizeof(R) * n, sizeof(double) * n, sizeof(R) * n, sizeof(R) * n,
                          sizeof(double) * n, sizeof(R) * n, sizeof(int) * n,
                          sizeof(cplx<R>) * (size_t)c->nmesh()};
    size_t o = 256;
    for (int k = 0; k < 9; ++k) {
        off[k] = o;
        o = align256(o + sz[k]);
    }
    return o;
}

template <typename R>
static int state_bytes(Ctx* c, int64_t* bytes) {
    size_t off[9];
    *bytes = (int64_t)blob_layout<R>(c, off);
    return PUFF_OK;
}

// Sections of a blob and the context arrays they hold.
template <typename R>
static CopySections blob_sections(Ctx* c, char* blob, bool to_blob) {
    Particles<R>& P = parts<R>(c);
    size_t off[9];
    blob_layout<R>(c, off);
    const size_t n = (size_t)c->n;
    void* arr[9] = {P.y.x, P.y.y, P.y.z2, P.y.pr, P.y.pi, P.y.g, P.chi, c->orig, c->A};
    const size_t bytes[9] = {sizeof(R) * n, sizeof(R) * n, sizeof(double) * n, sizeof(R) * n, sizeof(R) * n,
                             sizeof(double) * n, sizeof(R) * n, sizeof(int) * n,
                             sizeof(cplx<R>) * (size_t)c->nmesh()};
    CopySections cs{};
    cs.n = 9;
    for (int k = 0; k < 9; ++k) {
        uint8_t* b = reinterpret_cast<uint8_t*>(blob + off[k]);
        cs.src[k] = to_blob ? reinterpret_cast<const uint8_t*>(arr[k]) : b;
        cs.dst[k] = to_blob ? b : reinterpret_cast<uint8_t*>(arr[k]);
        cs.bytes[k] = (int64_t)bytes[k];
    }
    return cs;
}

template <typename R>
static int state_save(Ctx* c, void* dst, void* host_header) {
    Words64 words{};
    BlobHeader* hd = reinterpret_cast<BlobHeader*>(words.w);
    hd->magic = BLOB_MAGIC;
    hd->n = c->n;
    hd->nmesh = c->nmesh();
    hd->precision = c->precision;
    hd->sorted = c->sorted;
    hd->identity_order = c->identity_order;
    hd->sort_due = c->sort_due;
    hd->order_id = c->order_id;
    hd->steps_since_sort = c->steps_since_sort;
    hd->n2col = c->n2col;
    hd->n2col0 = c->n2col0;
    hd->undgrad = c->undgrad;
    char* d = reinterpret_cast<char*>(dst);
    hipStream_t s = c->stream;
    launch_store_words(reinterpret_cast<uint32_t*>(d), words, (int)(sizeof(BlobHeader) / 4), s);
    launch_copy_sections(blob_sections<R>(c, d, true), s);
    PUFF_HIP(hipGetLastError());
    if (host_header) {
        std::memset(host_header, 0, PUFF_STATE_HEADER_BYTES);
        std::memcpy(host_header, hd, sizeof(BlobHeader));
    }
    return PUFF_OK;
}

template <typename R>
static int state_load(Ctx* c, const void* src, const void* host_header) {
    const char* d = reinterpret_cast<const char*>(src);
    hipStream_t s = c->stream;
    BlobHeader hd;
    if (host_header) {
        std::memcpy(&hd, host_header, sizeof(BlobHeader));
    } else {
        if (!c->h_hdr) PUFF_HIP(hipHostMalloc(&c->h_hdr, 256, hipHostMallocDefault));
        PUFF_HIP(hipMemcpyAsync(c->h_hdr, d, sizeof(BlobHeader), hipMemcpyDeviceToHost, s));
        PUFF_HIP(hipStreamSynchronize(s));
        hd = *reinterpret_cast<const BlobHeader*>(c->h_hdr);
    }
    if (hd.magic != BLOB_MAGIC || hd.n != c->n || hd.nmesh != c->nmesh() || hd.precision != c->precision) {
        set_error("state blob does not belong to this context's shape and precision");
        return PUFF_ERR_ARG;
    }
    launch_copy_sections(blob_sections<R>(c, const_cast<char*>(d), false), s);
    PUFF_HIP(hipGetLastError());
    c->sorted = hd.sorted != 0;
    c->identity_order = hd.identity_order != 0;
    c->sort_due = hd.sort_due != 0;
    c->order_id = hd.order_id;
    c->steps_since_sort = hd.steps_since_sort;
    c->n2col = hd.n2col;
    c->n2col0 = hd.n2col0;
    c->undgrad = hd.undgrad;
    return PUFF_OK;
}

constexpr uint64_t CAPTURE_MAGIC = 0x3150414346465550ull;   // "PUFFCAP1"

// Start of a captured step (after its sort): header, start state and cleared
// stage sources in the capture; a host copy of the header to cap_hdr.
template <typename R>
static void capture_start(Ctx* c, char* cap, void* cap_hdr, CaptureView<R>& v, const double* taper0) {
    Particles<R>& P = parts<R>(c);
    capture_view<R>(c, cap, &v);
    Words64 words{};
    BlobHeader* hd = reinterpret_cast<BlobHeader*>(words.w);
    hd->magic = CAPTURE_MAGIC;
    hd->n = c->n;
    hd->nmesh = c->nmesh();
    hd->precision = c->precision;
    hd->sorted = c->sorted;
    hd->identity_order = c->identity_order;
    hd->sort_due = c->sort_due;
    hd->order_id = c->order_id;
    hd->steps_since_sort = c->steps_since_sort;
    hd->n2col = taper0[0];
    hd->n2col0 = taper0[1];
    hd->undgrad = taper0[2];
    hipStream_t s = c->stream;
    launch_store_words(reinterpret_cast<uint32_t*>(cap), words, (int)(sizeof(BlobHeader) / 4), s);
    const size_t n = (size_t)c->n;
    // the start state and chi are written by stage 1 (StageArgs::ycap); the
    // slot order and the field are copied here
    CopySections cs{};
    const void* src[2] = {c->orig, c->A};
    void* dst[2] = {v.orig, v.A0};
    const size_t bytes[2] = {sizeof(int) * n, sizeof(cplx<R>) * (size_t)c->nmesh()};
    cs.n = 2;
    for (int k = 0; k < 2; ++k) {
        cs.src[k] = reinterpret_cast<const uint8_t*>(src[k]);
        cs.dst[k] = reinterpret_cast<uint8_t*>(dst[k]);
        cs.bytes[k] = (int64_t)bytes[k];
    }
    launch_copy_sections(cs, s);
    for (int k = 0; k < 3; ++k) fill_zero_typed<cplx<R>>(v.dA[k], c->nmesh(), s);
    if (cap_hdr) {
        std::memset(cap_hdr, 0, PUFF_STATE_HEADER_BYTES);
        std::memcpy(cap_hdr, hd, sizeof(BlobHeader));
    }
}

template <typename R>
static int capture_check(Ctx* c, const void* cap_hdr, double taper[3], int64_t* order_id) {
    if (!cap_hdr) {
        set_error("the adjoint of a captured step needs the capture's host header");
        return PUFF_ERR_ARG;
    }
    BlobHeader hd;
    std::memcpy(&hd, cap_hdr, sizeof(BlobHeader));
    if (hd.magic != CAPTURE_MAGIC || hd.n != c->n || hd.nmesh != c->nmesh() || hd.precision != c->precision) {
        set_error("step capture does not belong to this context's shape and precision");
        return PUFF_ERR_ARG;
    }
    taper[0] = hd.n2col;
    taper[1] = hd.n2col0;
    taper[2] = hd.undgrad;
    *order_id = hd.order_id;
    return PUFF_OK;
}

template <typename R>
static int capture_bytes(Ctx* c, int64_t* bytes) {
    CaptureView<R> v{};
    *bytes = (int64_t)capture_view<R>(c, nullptr, &v);
    return PUFF_OK;
}

template <typename R>
static int sort_now(Ctx* c) {
    sort_particles<R>(c);
    PUFF_HIP(hipGetLastError());
    PUFF_HIP(hipStreamSynchronize(c->stream));
    return PUFF_OK;
}

}  // namespace puff

// ============================================================================
// C ABI
// ============================================================================
using puff::Ctx;

struct puff_ctx {
    Ctx c;
};

#define CTX_CHECK(ctx)                                   \
    do {                                                 \
        if (!(ctx)) {                                    \
            puff::set_error("null context");             \
            return PUFF_ERR_ARG;                         \
        }                                                \
        hipError_t e_ = hipSetDevice((ctx)->c.device);   \
        if (e_ != hipSuccess) {                          \
            puff::set_error(hipGetErrorString(e_));      \
            return PUFF_ERR_HIP;                         \
        }                                                \
    } while (0)

namespace puff {

// One lattice element on the device macroparticles (puff_lattice.hip).  The
// corrector's sums use the diagnostics buffer.  Elements that move particles
// leave the slot order stale: the next step re-sorts.
template <typename R>
static int lattice_op(Ctx* c, const LatticeOp& op) {
    Particles<R>& P = parts<R>(c);
    launch_lattice<R>(op, P.y, P.chi, c->n, c->phys, c->mesh, c->diag, c->stream);
    PUFF_HIP(hipGetLastError());
    PUFF_HIP(hipStreamSynchronize(c->stream));
    if (op.kind != LAT_MODULATION && op.kind != LAT_QUAD) c->sort_due = true;
    return PUFF_OK;
}

}  // namespace puff

#define DISPATCH(ctx, fn, ...)                                                      \
    ((ctx)->c.precision == PUFF_FP64 ? puff::fn<double>(&(ctx)->c, __VA_ARGS__)     \
                                     : puff::fn<float>(&(ctx)->c, __VA_ARGS__))

extern "C" {

const char* puff_last_error(void) { return puff::g_err.c_str(); }

int puff_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int puff_create(puff_ctx** out, const puff_params* params, int64_t n, int32_t precision) {
    if (!out) return PUFF_ERR_ARG;
    *out = nullptr;
    int st = puff::validate(params, n);
    if (st != PUFF_OK) return st;
    if (precision != PUFF_FP64 && precision != PUFF_MIXED) {
        puff::set_error("unknown precision");
        return PUFF_ERR_ARG;
    }
    puff_ctx* ctx = new (std::nothrow) puff_ctx();
    if (!ctx) return PUFF_ERR_ALLOC;
    Ctx& c = ctx->c;
    c.p = *params;
    c.n = n;
    c.precision = precision;
    if (hipGetDevice(&c.device) != hipSuccess) {
        puff::set_error("no HIP device");
        delete ctx;
        return PUFF_ERR_HIP;
    }
    if (hipStreamCreate(&c.stream) != hipSuccess) {
        puff::set_error("hipStreamCreate failed");
        delete ctx;
        return PUFF_ERR_HIP;
    }
    c.own_stream = true;
    st = puff::create(&c);
    if (st != PUFF_OK) {
        puff::release(&c);
        delete ctx;
        return st;
    }
    *out = ctx;
    return PUFF_OK;
}

int puff_destroy(puff_ctx* ctx) {
    if (!ctx) return PUFF_OK;
    (void)hipSetDevice(ctx->c.device);
    puff::release(&ctx->c);
    delete ctx;
    return PUFF_OK;
}

int puff_set_stream(puff_ctx* ctx, void* stream) {
    CTX_CHECK(ctx);
    Ctx& c = ctx->c;
    (void)hipStreamSynchronize(c.stream);
    if (c.own_stream && c.stream) (void)hipStreamDestroy(c.stream);
    if (stream) {
        c.stream = (hipStream_t)stream;
        c.own_stream = false;
    } else {
        if (hipStreamCreate(&c.stream) != hipSuccess) return PUFF_ERR_HIP;
        c.own_stream = true;
    }
    return PUFF_OK;
}

int puff_set_deterministic(puff_ctx* ctx, int32_t on) {
    CTX_CHECK(ctx);
    puff::Ctx& c = ctx->c;
    if (on && c.precision != PUFF_FP64) {
        puff::set_error("the deterministic deposit is a mode of the FP64 form");
        return PUFF_ERR_ARG;
    }
    if (on && c.n > (int64_t(1) << 30)) {
        puff::set_error("the deterministic deposit holds at most 2^30 macroparticles");
        return PUFF_ERR_ARG;
    }
    if (on && !c.dets.acc) {
        PUFF_TRY(puff::dalloc(&c, (void**)&c.dets.acc,
                              sizeof(unsigned long long) * puff::DET_WORDS * (size_t)c.nmesh()));
        PUFF_TRY(puff::dalloc(&c, (void**)&c.dets.shift, sizeof(int)));
        PUFF_TRY(puff::dalloc(&c, (void**)&c.dets.bound, sizeof(unsigned long long)));
    }
    c.det = on != 0;
    return PUFF_OK;
}

int puff_synchronize(puff_ctx* ctx) {
    CTX_CHECK(ctx);
    return hipStreamSynchronize(ctx->c.stream) == hipSuccess ? PUFF_OK : PUFF_ERR_HIP;
}

int puff_upload_particles(puff_ctx* ctx, int32_t kind, const double* x, const double* y,
                          const double* z2, const double* px, const double* py, const double* gam,
                          const double* chi) {
    CTX_CHECK(ctx);
    return DISPATCH(ctx, upload_particles, kind, x, y, z2, px, py, gam, chi);
}

int puff_download_particles(puff_ctx* ctx, int32_t kind, double* x, double* y, double* z2,
                            double* px, double* py, double* gam) {
    CTX_CHECK(ctx);
    return DISPATCH(ctx, download_particles, kind, x, y, z2, px, py, gam);
}

int puff_upload_field(puff_ctx* ctx, int32_t kind, const double* ar, const double* ai) {
    CTX_CHECK(ctx);
    if (!ar || !ai) return PUFF_ERR_ARG;
    return DISPATCH(ctx, upload_field, kind, ar, ai);
}

int puff_download_field(puff_ctx* ctx, int32_t kind, double* ar, double* ai) {
    CTX_CHECK(ctx);
    return DISPATCH(ctx, download_field, kind, ar, ai);
}

int puff_set_taper(puff_ctx* ctx, double n2col, double n2col0, double undgrad) {
    if (!ctx) return PUFF_ERR_ARG;
    ctx->c.n2col = n2col;
    ctx->c.n2col0 = n2col0;
    ctx->c.undgrad = undgrad;
    return PUFF_OK;
}

int puff_get_taper(puff_ctx* ctx, double* n2col, double* n2col0) {
    if (!ctx) return PUFF_ERR_ARG;
    if (n2col) *n2col = ctx->c.n2col;
    if (n2col0) *n2col0 = ctx->c.n2col0;
    return PUFF_OK;
}

int puff_set_taper_table(puff_ctx* ctx, const double* values, int32_t nvals, double z0, double period) {
    if (!ctx || nvals < 0 || (nvals > 0 && (!values || !(period > 0)))) return PUFF_ERR_ARG;
    ctx->c.taper_table.assign(values, values + nvals);
    ctx->c.taper_z0 = z0;
    ctx->c.taper_period = period;
    return PUFF_OK;
}

int puff_set_sort_interval(puff_ctx* ctx, int32_t steps) {
    if (!ctx || steps < 0) return PUFF_ERR_ARG;
    ctx->c.sort_every = steps;
    return PUFF_OK;
}

int puff_rk4par(puff_ctx* ctx, double sz, double h) {
    CTX_CHECK(ctx);
    return ctx->c.precision == PUFF_FP64 ? puff::rk4par<double>(&ctx->c, sz, h)
                                         : puff::rk4par<float>(&ctx->c, sz, h);
}

int puff_diffraction_step(puff_ctx* ctx, double h) {
    CTX_CHECK(ctx);
    return ctx->c.precision == PUFF_FP64 ? puff::diffraction<double>(&ctx->c, h)
                                         : puff::diffraction<float>(&ctx->c, h);
}

int puff_transform(puff_ctx* ctx, int32_t sign) {
    CTX_CHECK(ctx);
    if (sign != -1 && sign != 1) return PUFF_ERR_ARG;
    return DISPATCH(ctx, transform, sign);
}

int puff_wrap_z2(puff_ctx* ctx) {
    CTX_CHECK(ctx);
    Ctx& c = ctx->c;
    if (!c.p.periodic) return PUFF_OK;
    double* z2 = c.precision == PUFF_FP64 ? c.pd.y.z2 : c.pf.y.z2;
    puff::launch_wrap(z2, c.n, c.p.dz2 * (double)(c.p.nz2 - 1), c.stream);
    if (hipGetLastError() != hipSuccess) return PUFF_ERR_HIP;
    return hipStreamSynchronize(c.stream) == hipSuccess ? PUFF_OK : PUFF_ERR_HIP;
}

int puff_drift(puff_ctx* ctx, double dz) {
    CTX_CHECK(ctx);
    puff::LatticeOp op{};
    op.kind = puff::LAT_DRIFT;
    op.a = dz;
    return DISPATCH(ctx, lattice_op, op);
}

int puff_chicane(puff_ctx* ctx, double disp, double slip) {
    CTX_CHECK(ctx);
    puff::LatticeOp op{};
    op.kind = puff::LAT_CHICANE;
    op.a = disp;
    op.b = slip;
    op.lenz2 = ctx->c.p.periodic ? ctx->c.p.dz2 * (double)(ctx->c.p.nz2 - 1) : 0.0;
    return DISPATCH(ctx, lattice_op, op);
}

int puff_quad(puff_ctx* ctx, double fx, double fy) {
    CTX_CHECK(ctx);
    if (fx == 0.0 || fy == 0.0) {
        puff::set_error("quad focal strengths must be non-zero");
        return PUFF_ERR_ARG;
    }
    puff::LatticeOp op{};
    op.kind = puff::LAT_QUAD;
    op.a = fx;
    op.b = fy;
    return DISPATCH(ctx, lattice_op, op);
}

int puff_modulation(puff_ctx* ctx, double wavenum, double mag) {
    CTX_CHECK(ctx);
    puff::LatticeOp op{};
    op.kind = puff::LAT_MODULATION;
    op.a = wavenum;
    op.b = mag;
    return DISPATCH(ctx, lattice_op, op);
}

int puff_correct_trans(puff_ctx* ctx) {
    CTX_CHECK(ctx);
    puff::LatticeOp op{};
    op.kind = puff::LAT_CORRECT_TRANS;
    return DISPATCH(ctx, lattice_op, op);
}

int puff_match(puff_ctx* ctx, double sz, double n2col, int32_t sign) {
    CTX_CHECK(ctx);
    if (sign != 1 && sign != -1) {
        puff::set_error("match sign must be +1 (entry) or -1 (exit)");
        return PUFF_ERR_ARG;
    }
    const puff_params& p = ctx->c.p;
    puff::LatticeOp op{};
    op.kind = puff::LAT_MATCH;
    puff::MatchArgs& m = op.match;
    m.sz = sz;
    m.n2col = n2col;
    m.rho = p.rho;
    m.eta = p.eta;
    m.kappa = p.kappa;
    m.gamma_r = p.gamma_r;
    m.ux = p.fx;
    m.uy = p.fy;
    m.ncx = p.aw * p.aw;
    m.ncy = 2.0 * p.aw * p.aw / (p.fx * p.fx + p.fy * p.fy);
    m.kx = p.kx_und;
    m.ky = p.ky_und;
    m.und_type = p.und_type;
    m.sign = sign;
    return DISPATCH(ctx, lattice_op, op);
}

int puff_set_undulator(puff_ctx* ctx, int32_t und_type, double fx, double fy, double kbeta_x,
                       double kbeta_y, double szfs, double szfe) {
    CTX_CHECK(ctx);
    if (und_type < 0 || und_type > PUFF_UND_CURVED || !(fx * fx + fy * fy > 0.0)) {
        puff::set_error("invalid undulator type or polarisation");
        return PUFF_ERR_ARG;
    }
    puff_params& p = ctx->c.p;
    p.und_type = und_type;
    p.fx = fx;
    p.fy = fy;
    p.kbeta_x = kbeta_x;
    p.kbeta_y = kbeta_y;
    p.szfs = szfs;
    p.szfe = szfe;
    return PUFF_OK;
}

int puff_power(puff_ctx* ctx, double* out) {
    CTX_CHECK(ctx);
    if (!out) return PUFF_ERR_ARG;
    return DISPATCH(ctx, power, out);
}

int puff_bunching(puff_ctx* ctx, int32_t harmonic, int32_t nslices, double slice_len, double* out) {
    CTX_CHECK(ctx);
    if (!out || !(slice_len > 0)) return PUFF_ERR_ARG;
    return DISPATCH(ctx, bunching, harmonic, nslices, slice_len, out);
}

int puff_tile_fallback(puff_ctx* ctx, double* last_step, double* after_sort) {
    CTX_CHECK(ctx);
    if (!last_step || !after_sort) return PUFF_ERR_ARG;
    *last_step = ctx->c.fallback_last;
    *after_sort = ctx->c.fallback_base;
    return PUFF_OK;
}

int puff_slice_stats(puff_ctx* ctx, int32_t nslices, double slice_len, double* out) {
    CTX_CHECK(ctx);
    if (!out) return PUFF_ERR_ARG;
    return DISPATCH(ctx, slice_stats, nslices, slice_len, out);
}

int puff_current(puff_ctx* ctx, int32_t npts, double sam_len, double* out) {
    CTX_CHECK(ctx);
    if (!out) return PUFF_ERR_ARG;
    return DISPATCH(ctx, current, npts, sam_len, out);
}

int puff_profile_enable(puff_ctx* ctx, int32_t on) {
    CTX_CHECK(ctx);
    (void)hipStreamSynchronize(ctx->c.stream);
    puff::prof_collect(&ctx->c);
    ctx->c.prof = on != 0;
    return PUFF_OK;
}

int puff_profile_read(puff_ctx* ctx, double* ms, int64_t* counts, int32_t nkinds, int32_t reset) {
    CTX_CHECK(ctx);
    (void)hipStreamSynchronize(ctx->c.stream);
    puff::prof_collect(&ctx->c);
    if (nkinds < 0 || nkinds > PUFF_PROF_KINDS) return PUFF_ERR_ARG;
    for (int k = 0; k < nkinds; ++k) {
        if (ms) ms[k] = ctx->c.prof_ms[k];
        if (counts) counts[k] = ctx->c.prof_n[k];
    }
    if (reset) {
        for (int k = 0; k < PUFF_PROF_KINDS; ++k) {
            ctx->c.prof_ms[k] = 0.0;
            ctx->c.prof_n[k] = 0;
        }
    }
    return PUFF_OK;
}

int puff_active_box(puff_ctx* ctx, int32_t* box6) {
    CTX_CHECK(ctx);
    if (!box6) return PUFF_ERR_ARG;
    puff::Box b[puff::BOX_COPIES];
    if (hipMemcpy(b, ctx->c.box, sizeof(b), hipMemcpyDeviceToHost) != hipSuccess) return PUFF_ERR_HIP;
    for (int d = 0; d < 3; ++d) {
        int lo = 0x7fffffff, hi = -1;
        for (int k = 0; k < puff::BOX_COPIES; ++k) {
            lo = std::min(lo, b[k].lo[d]);
            hi = std::max(hi, b[k].hi[d]);
        }
        box6[2 * d] = lo;
        box6[2 * d + 1] = hi;
    }
    return PUFF_OK;
}

int puff_particle_count(puff_ctx* ctx, int64_t* n) {
    if (!ctx || !n) return PUFF_ERR_ARG;
    *n = ctx->c.n;
    return PUFF_OK;
}

int puff_set_deferred(puff_ctx* ctx, int32_t on) {
    CTX_CHECK(ctx);
    Ctx& c = ctx->c;
    if (hipStreamSynchronize(c.stream) != hipSuccess) return PUFF_ERR_HIP;
    puff::prof_collect(&c);
    c.deferred = on != 0;
    return PUFF_OK;
}

int puff_deferred_status(puff_ctx* ctx) {
    CTX_CHECK(ctx);
    Ctx& c = ctx->c;
    PUFF_HIP(hipMemcpyAsync(c.h_flags, c.flags, sizeof(int) * puff::FLAG_WORDS, hipMemcpyDeviceToHost, c.stream));
    puff::launch_fill_zero_words(reinterpret_cast<uint32_t*>(c.flags + puff::FLAG_WORD_STICKY), 1, c.stream);
    PUFF_HIP(hipStreamSynchronize(c.stream));
    puff::prof_collect(&c);
    const int f = c.h_flags[puff::FLAG_WORD_STICKY];
    if (f & puff::FLAG_DET_RANGE) {
        puff::set_error("a source exceeded the range the deterministic deposit was scaled for");
        return PUFF_ERR_STATE;
    }
    if (f & puff::FLAG_Z2MESH) return PUFF_OUT_Z2_MESH;
    if (f & puff::FLAG_TRANSVERSE) return PUFF_OUT_TRANSVERSE;
    if (f) return PUFF_OUT_Z2_BUFFER;
    return PUFF_OK;
}

int puff_adjoint_taper_gradient(puff_ctx* ctx, double* out, int32_t n) {
    CTX_CHECK(ctx);
    Ctx& c = ctx->c;
    if (!c.adj) {
        puff::set_error("adjoint buffers not allocated (puff_adjoint_enable)");
        return PUFF_ERR_STATE;
    }
    if (!out || n < 0 || n > puff::DN2_TABLE_CAP + 1) return PUFF_ERR_ARG;
    PUFF_HIP(hipMemcpyAsync(out, c.dn2_table, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    PUFF_HIP(hipStreamSynchronize(c.stream));
    return PUFF_OK;
}

int puff_state_bytes(puff_ctx* ctx, int64_t* bytes) {
    if (!ctx || !bytes) return PUFF_ERR_ARG;
    return DISPATCH(ctx, state_bytes, bytes);
}

int puff_state_save(puff_ctx* ctx, void* dst) { return puff_state_save_ex(ctx, dst, nullptr); }

int puff_state_load(puff_ctx* ctx, const void* src) { return puff_state_load_ex(ctx, src, nullptr); }

int puff_state_save_ex(puff_ctx* ctx, void* dst, void* host_header) {
    CTX_CHECK(ctx);
    if (!dst) return PUFF_ERR_ARG;
    return DISPATCH(ctx, state_save, dst, host_header);
}

int puff_state_load_ex(puff_ctx* ctx, const void* src, const void* host_header) {
    CTX_CHECK(ctx);
    if (!src) return PUFF_ERR_ARG;
    return DISPATCH(ctx, state_load, src, host_header);
}

int puff_sort_now(puff_ctx* ctx) {
    CTX_CHECK(ctx);
    return ctx->c.precision == PUFF_FP64 ? puff::sort_now<double>(&ctx->c) : puff::sort_now<float>(&ctx->c);
}

int puff_adjoint_enable(puff_ctx* ctx) {
    CTX_CHECK(ctx);
    return ctx->c.precision == PUFF_FP64 ? puff::adjoint_enable<double>(&ctx->c)
                                         : puff::adjoint_enable<float>(&ctx->c);
}

int puff_adjoint_zero(puff_ctx* ctx) {
    CTX_CHECK(ctx);
    return ctx->c.precision == PUFF_FP64 ? puff::adjoint_zero<double>(&ctx->c)
                                         : puff::adjoint_zero<float>(&ctx->c);
}

int puff_adjoint_seed_power(puff_ctx* ctx, const double* wz, double scale) {
    CTX_CHECK(ctx);
    if (!wz) return PUFF_ERR_ARG;
    return DISPATCH(ctx, adjoint_seed_power, wz, scale);
}

int puff_rk4par_adjoint(puff_ctx* ctx, double sz, double h, double* dn2col, int32_t* slot) {
    CTX_CHECK(ctx);
    return DISPATCH(ctx, rk4par_adjoint, sz, h, dn2col, slot);
}

int puff_capture_bytes(puff_ctx* ctx, int64_t* bytes) {
    if (!ctx || !bytes) return PUFF_ERR_ARG;
    return DISPATCH(ctx, capture_bytes, bytes);
}

int puff_rk4par_capture(puff_ctx* ctx, double sz, double h, void* capture, void* host_header) {
    CTX_CHECK(ctx);
    if (!capture || !host_header) return PUFF_ERR_ARG;
    return ctx->c.precision == PUFF_FP64
               ? puff::rk4par<double>(&ctx->c, sz, h, reinterpret_cast<char*>(capture), host_header)
               : puff::rk4par<float>(&ctx->c, sz, h, reinterpret_cast<char*>(capture), host_header);
}

int puff_rk4par_adjoint_captured(puff_ctx* ctx, double sz, double h, const void* capture,
                                 const void* host_header, double* dn2col, int32_t* slot) {
    CTX_CHECK(ctx);
    if (!capture || !host_header) return PUFF_ERR_ARG;
    const char* cap = reinterpret_cast<const char*>(capture);
    return ctx->c.precision == PUFF_FP64
               ? puff::rk4par_adjoint<double>(&ctx->c, sz, h, dn2col, slot, cap, host_header)
               : puff::rk4par_adjoint<float>(&ctx->c, sz, h, dn2col, slot, cap, host_header);
}

int puff_diffraction_adjoint(puff_ctx* ctx, double h) {
    CTX_CHECK(ctx);
    return ctx->c.precision == PUFF_FP64 ? puff::diffraction_adjoint<double>(&ctx->c, h)
                                         : puff::diffraction_adjoint<float>(&ctx->c, h);
}

int puff_adjoint_upload(puff_ctx* ctx, int32_t kind, const double* x, const double* y, const double* z2,
                        const double* px, const double* py, const double* gam, const double* ar,
                        const double* ai) {
    CTX_CHECK(ctx);
    return DISPATCH(ctx, adjoint_upload, kind, x, y, z2, px, py, gam, ar, ai);
}

int puff_adjoint_download(puff_ctx* ctx, int32_t kind, double* x, double* y, double* z2, double* px,
                          double* py, double* gam, double* ar, double* ai) {
    CTX_CHECK(ctx);
    return DISPATCH(ctx, adjoint_download, kind, x, y, z2, px, py, gam, ar, ai);
}

}  // extern "C"
