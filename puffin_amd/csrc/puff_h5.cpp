// Puffin-format HDF5 exchange over the HDF5 C library (host only).
//
// Layouts and attribute names follow the reference's writers and readers
// (puffin/lib/io/hdf5_puff_coll.f90, hdf5_puff.f90, h5_in.f90): a C-order
// (N, 7) "electrons" table, a component-major "aperp" field, 1-D integrated
// datasets, a "runInfo" group and the position attributes on each dataset.
// Columns are moved with strided hyperslab selections, so no N x 7 staging
// copy is made.
#include "../../include/puffin_h5.h"

#include <hdf5.h>

#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// Scope guard for an HDF5 identifier.
struct Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Id() = default;
    Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    ~Id() {
        if (id >= 0 && close) close(id);
    }
    bool ok() const { return id >= 0; }
    operator hid_t() const { return id; }
};

struct Quiet {
    Quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};

bool put_double(hid_t obj, const char* name, double v) {
    Id sp(H5Screate(H5S_SCALAR), H5Sclose);
    Id a(H5Acreate2(obj, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Awrite(a, H5T_NATIVE_DOUBLE, &v) >= 0;
}

bool put_int(hid_t obj, const char* name, int32_t v) {
    Id sp(H5Screate(H5S_SCALAR), H5Sclose);
    Id a(H5Acreate2(obj, name, H5T_STD_I32LE, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Awrite(a, H5T_NATIVE_INT32, &v) >= 0;
}

// Fixed-length, space-padded ASCII string, as the reference's writers make them.
bool put_string(hid_t obj, const char* name, const std::string& s) {
    Id t(H5Tcopy(H5T_C_S1), H5Tclose);
    if (!t.ok() || H5Tset_size(t, s.empty() ? 1 : s.size()) < 0 || H5Tset_strpad(t, H5T_STR_SPACEPAD) < 0)
        return false;
    Id sp(H5Screate(H5S_SCALAR), H5Sclose);
    Id a(H5Acreate2(obj, name, t, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    std::string buf = s.empty() ? std::string(" ") : s;
    return a.ok() && H5Awrite(a, t, buf.data()) >= 0;
}

bool get_double(hid_t obj, const char* name, double* v) {
    if (H5Aexists(obj, name) <= 0) return false;
    Id a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Aread(a, H5T_NATIVE_DOUBLE, v) >= 0;
}

bool get_int(hid_t obj, const char* name, int32_t* v) {
    if (H5Aexists(obj, name) <= 0) return false;
    Id a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Aread(a, H5T_NATIVE_INT32, v) >= 0;
}

// The position attributes (on every dataset and on runInfo).
bool put_position(hid_t obj, const puff_h5_info& f) {
    return put_double(obj, "time", f.time) && put_double(obj, "zbarTotal", f.zbar_total) &&
           put_double(obj, "zbarLocal", f.zbar_local) && put_double(obj, "zbarInter", f.zbar_inter) &&
           put_double(obj, "zTotal", f.z_total) && put_double(obj, "zLocal", f.z_local) &&
           put_double(obj, "zInter", f.z_inter) && put_int(obj, "istep", f.istep) &&
           put_int(obj, "iCsteps", f.icsteps) && put_int(obj, "iUnd_cr", f.iund_cr) &&
           put_int(obj, "iChic_cr", f.ichic_cr) && put_int(obj, "iDrift_cr", f.idrift_cr) &&
           put_int(obj, "iQuad_cr", f.iquad_cr) && put_int(obj, "iModulation_cr", f.imodulation_cr) &&
           put_int(obj, "iL", f.il) && put_int(obj, "iWrite_cr", f.iwrite_cr);
}

// Missing attributes leave the field as it was (the reference's resume
// reader takes only those it needs, h5_in.f90:504-519).
void get_position(hid_t obj, puff_h5_info* f) {
    get_double(obj, "time", &f->time);
    get_double(obj, "zbarTotal", &f->zbar_total);
    get_double(obj, "zbarLocal", &f->zbar_local);
    get_double(obj, "zbarInter", &f->zbar_inter);
    get_double(obj, "zTotal", &f->z_total);
    get_double(obj, "zLocal", &f->z_local);
    get_double(obj, "zInter", &f->z_inter);
    get_int(obj, "istep", &f->istep);
    get_int(obj, "iCsteps", &f->icsteps);
    get_int(obj, "iUnd_cr", &f->iund_cr);
    get_int(obj, "iChic_cr", &f->ichic_cr);
    get_int(obj, "iDrift_cr", &f->idrift_cr);
    get_int(obj, "iQuad_cr", &f->iquad_cr);
    get_int(obj, "iModulation_cr", &f->imodulation_cr);
    get_int(obj, "iL", &f->il);
    get_int(obj, "iWrite_cr", &f->iwrite_cr);
}

bool put_run_info(hid_t file, const puff_h5_info& f) {
    Id g(H5Gcreate2(file, "runInfo", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return false;
    return put_position(g, f) && put_int(g, "nX", f.nx) && put_int(g, "nY", f.ny) && put_int(g, "nZ2", f.nz2) &&
           put_int(g, "fieldMesh", f.field_mesh) && put_int(g, "nSteps", f.n_steps) &&
           put_int(g, "iScale", f.iscale) && put_double(g, "sLengthOfElmX", f.dx) &&
           put_double(g, "sLengthOfElmY", f.dy) && put_double(g, "sLengthOfElmZ2", f.dz2) &&
           put_double(g, "rho", f.rho) && put_double(g, "eta", f.eta) && put_double(g, "kappa", f.kappa) &&
           put_double(g, "gamma_r", f.gamma_r) && put_double(g, "aw", f.aw) && put_double(g, "Lg", f.lg) &&
           put_double(g, "Lc", f.lc) && put_double(g, "lambda_w", f.lambda_w) &&
           put_double(g, "lambda_r", f.lambda_r) && put_double(g, "sStepSize", f.step_size) &&
           put_double(g, "npk_bar", f.npk_bar) && put_double(g, "transArea", f.trans_area) &&
           put_double(g, "transAreaSI", f.trans_area_si) && put_string(g, "vsType", "runInfo") &&
           put_string(g, "vsSoftware", "PUFFIN");
}

bool get_run_info(hid_t file, puff_h5_info* f) {
    if (H5Lexists(file, "runInfo", H5P_DEFAULT) <= 0) return false;
    Id g(H5Gopen2(file, "runInfo", H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return false;
    get_position(g, f);
    // the field reader's required set (h5_in.f90:775-781)
    bool ok = get_int(g, "nX", &f->nx) && get_int(g, "nY", &f->ny) && get_int(g, "nZ2", &f->nz2) &&
              get_double(g, "sLengthOfElmX", &f->dx) && get_double(g, "sLengthOfElmY", &f->dy) &&
              get_double(g, "sLengthOfElmZ2", &f->dz2);
    get_int(g, "fieldMesh", &f->field_mesh);
    get_int(g, "nSteps", &f->n_steps);
    get_int(g, "iScale", &f->iscale);
    get_double(g, "rho", &f->rho);
    get_double(g, "eta", &f->eta);
    get_double(g, "kappa", &f->kappa);
    get_double(g, "gamma_r", &f->gamma_r);
    get_double(g, "aw", &f->aw);
    get_double(g, "Lg", &f->lg);
    get_double(g, "Lc", &f->lc);
    get_double(g, "lambda_w", &f->lambda_w);
    get_double(g, "lambda_r", &f->lambda_r);
    get_double(g, "sStepSize", &f->step_size);
    get_double(g, "npk_bar", &f->npk_bar);
    get_double(g, "transArea", &f->trans_area);
    get_double(g, "transAreaSI", &f->trans_area_si);
    return ok;
}

bool put_time_group(hid_t file, const puff_h5_info& f) {
    Id g(H5Gcreate2(file, "time", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    return g.ok() && put_string(g, "vsType", "time") && put_int(g, "vsStep", f.icsteps) &&
           put_double(g, "vsTime", f.time);
}

hid_t create_file(const char* path) {
    return H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
}

hid_t open_file(const char* path) { return H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT); }

bool write_1d(hid_t file, const char* name, const double* v, hsize_t n, const puff_h5_info& f) {
    Id sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
    Id d(H5Dcreate2(file, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    return d.ok() && H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0 &&
           put_position(d, f) && put_string(d, "vsType", "variable");
}

// Column c of an (n, ncols) dataset, as a strided hyperslab.
bool column_io(hid_t d, hsize_t n, int c, double* buf, bool write) {
    Id fs(H5Dget_space(d), H5Sclose);
    hsize_t start[2] = {0, (hsize_t)c}, count[2] = {n, 1};
    if (!fs.ok() || H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr) < 0) return false;
    Id ms(H5Screate_simple(1, &n, nullptr), H5Sclose);
    if (!ms.ok()) return false;
    return write ? H5Dwrite(d, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, buf) >= 0
                 : H5Dread(d, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, buf) >= 0;
}

// Component k of a component-major field dataset.
bool component_io(hid_t d, int rank, const hsize_t* dims, int k, double* buf, bool write) {
    Id fs(H5Dget_space(d), H5Sclose);
    hsize_t start[4] = {(hsize_t)k, 0, 0, 0}, count[4] = {1, dims[1], rank > 2 ? dims[2] : 1, rank > 3 ? dims[3] : 1};
    if (!fs.ok() || H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr) < 0) return false;
    hsize_t n = 1;
    for (int r = 1; r < rank; ++r) n *= dims[r];
    Id ms(H5Screate_simple(1, &n, nullptr), H5Sclose);
    if (!ms.ok()) return false;
    return write ? H5Dwrite(d, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, buf) >= 0
                 : H5Dread(d, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, buf) >= 0;
}

bool is_float(hid_t d) {
    Id t(H5Dget_type(d), H5Tclose);
    return t.ok() && H5Tget_class(t) == H5T_FLOAT;
}

}  // namespace

extern "C" {

const char* puff_h5_last_error(void) { return g_err.c_str(); }

int puff_h5_write_electrons(const char* path, const puff_h5_info* info, int64_t n, const double* x,
                            const double* y, const double* z2, const double* px, const double* py,
                            const double* gam, const double* chi) {
    Quiet q;
    if (!path || !info || n < 0 || (n > 0 && (!x || !y || !z2 || !px || !py || !gam || !chi)))
        return fail(PUFF_H5_ERR_ARG, "puff_h5_write_electrons: bad arguments");
    Id f(create_file(path), H5Fclose);
    if (!f.ok()) return fail(PUFF_H5_ERR_IO, std::string("cannot create ") + path);
    hsize_t dims[2] = {(hsize_t)n, 7};
    Id sp(H5Screate_simple(2, dims, nullptr), H5Sclose);
    Id d(H5Dcreate2(f, "electrons", H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return fail(PUFF_H5_ERR_IO, "cannot create dataset electrons");
    const double* cols[7] = {x, y, z2, px, py, gam, chi};
    if (n > 0)
        for (int c = 0; c < 7; ++c)
            if (!column_io(d, (hsize_t)n, c, const_cast<double*>(cols[c]), true))
                return fail(PUFF_H5_ERR_IO, "writing electrons failed");
    const int nd = info->num_spatial_dims;
    bool ok = put_position(d, *info) && put_int(d, "numSpatialDims", nd) &&
              put_int(d, "vsNumSpatialDims", nd) && put_double(d, "gainLength", info->lg) &&
              put_double(d, "cooperationLength", info->lc) && put_string(d, "vsType", "variableWithMesh") &&
              put_string(d, "vsTimeGroup", "time") && put_string(d, "vsAxisLabels", "xbar,ybar,z2bar") &&
              put_string(d, "vsLabels",
                         "electrons_x,electrons_y,electrons_z,electrons_px,electrons_py,"
                         "electrons_gamma,electrons_weight") &&
              put_time_group(f, *info) && put_run_info(f, *info);
    if (!ok) return fail(PUFF_H5_ERR_IO, "writing electrons attributes failed");
    return H5Fflush(f, H5F_SCOPE_LOCAL) >= 0 ? PUFF_H5_OK : fail(PUFF_H5_ERR_IO, "flush failed");
}

int puff_h5_count_electrons(const char* path, int64_t* n) {
    Quiet q;
    if (!path || !n) return fail(PUFF_H5_ERR_ARG, "puff_h5_count_electrons: bad arguments");
    Id f(open_file(path), H5Fclose);
    if (!f.ok()) return fail(PUFF_H5_ERR_IO, std::string("cannot open ") + path);
    if (H5Lexists(f, "electrons", H5P_DEFAULT) <= 0) return fail(PUFF_H5_ERR_FORMAT, "no dataset electrons");
    Id d(H5Dopen2(f, "electrons", H5P_DEFAULT), H5Dclose);
    Id sp(H5Dget_space(d), H5Sclose);
    if (!d.ok() || !sp.ok() || H5Sget_simple_extent_ndims(sp) != 2)
        return fail(PUFF_H5_ERR_FORMAT, "electrons is not a rank-2 dataset");
    hsize_t dims[2];
    H5Sget_simple_extent_dims(sp, dims, nullptr);
    if (dims[1] != 7) return fail(PUFF_H5_ERR_FORMAT, "electrons does not have seven columns");
    *n = (int64_t)dims[0];
    return PUFF_H5_OK;
}

int puff_h5_read_electrons(const char* path, int64_t n, double* x, double* y, double* z2, double* px,
                           double* py, double* gam, double* chi, puff_h5_info* info) {
    Quiet q;
    int64_t nf = 0;
    int st = puff_h5_count_electrons(path, &nf);
    if (st != PUFF_H5_OK) return st;
    if (nf != n) return fail(PUFF_H5_ERR_FORMAT, "electrons holds " + std::to_string(nf) + " rows, not " +
                                                     std::to_string(n));
    if (n > 0 && (!x || !y || !z2 || !px || !py || !gam || !chi))
        return fail(PUFF_H5_ERR_ARG, "puff_h5_read_electrons: null column");
    Id f(open_file(path), H5Fclose);
    Id d(H5Dopen2(f, "electrons", H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return fail(PUFF_H5_ERR_IO, "cannot open electrons");
    if (!is_float(d)) return fail(PUFF_H5_ERR_FORMAT, "electrons is not floating point");
    double* cols[7] = {x, y, z2, px, py, gam, chi};
    if (n > 0)
        for (int c = 0; c < 7; ++c)
            if (!column_io(d, (hsize_t)n, c, cols[c], false)) return fail(PUFF_H5_ERR_IO, "reading electrons failed");
    if (info) {
        get_run_info(f, info);
        get_position(d, info);
        get_int(d, "numSpatialDims", &info->num_spatial_dims);
    }
    return PUFF_H5_OK;
}

int puff_h5_write_field(const char* path, const puff_h5_info* info, const double* ar, const double* ai) {
    Quiet q;
    if (!path || !info || !ar || !ai || info->nz2 < 2 ||
        (info->num_spatial_dims != 1 && (info->nx < 2 || info->ny < 2)))
        return fail(PUFF_H5_ERR_ARG, "puff_h5_write_field: bad arguments");
    Id f(create_file(path), H5Fclose);
    if (!f.ok()) return fail(PUFF_H5_ERR_IO, std::string("cannot create ") + path);
    const bool one_d = info->num_spatial_dims == 1;
    int rank = one_d ? 2 : 4;
    hsize_t dims[4] = {2, (hsize_t)info->nz2, (hsize_t)info->ny, (hsize_t)info->nx};
    Id sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    Id d(H5Dcreate2(f, "aperp", H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return fail(PUFF_H5_ERR_IO, "cannot create dataset aperp");
    if (!component_io(d, rank, dims, 0, const_cast<double*>(ar), true) ||
        !component_io(d, rank, dims, 1, const_cast<double*>(ai), true))
        return fail(PUFF_H5_ERR_IO, "writing aperp failed");
    bool ok = put_position(d, *info) && put_int(d, "numSpatialDims", info->num_spatial_dims) &&
              put_int(d, "vsNumSpatialDims", info->num_spatial_dims) && put_string(d, "vsType", "variable") &&
              put_string(d, "vsIndexOrder", "compMajorC") && put_string(d, "vsCentering", "nodal") &&
              put_string(d, "vsTimeGroup", "time") && put_string(d, "vsLabels", "aperp_real, aperp_imaginary") &&
              put_string(d, "vsAxisLabels", one_d ? "z2" : "x,y,z2") && put_time_group(f, *info) &&
              put_run_info(f, *info);
    if (!ok) return fail(PUFF_H5_ERR_IO, "writing aperp attributes failed");
    return H5Fflush(f, H5F_SCOPE_LOCAL) >= 0 ? PUFF_H5_OK : fail(PUFF_H5_ERR_IO, "flush failed");
}

int puff_h5_read_field_info(const char* path, puff_h5_info* info) {
    Quiet q;
    if (!path || !info) return fail(PUFF_H5_ERR_ARG, "puff_h5_read_field_info: bad arguments");
    Id f(open_file(path), H5Fclose);
    if (!f.ok()) return fail(PUFF_H5_ERR_IO, std::string("cannot open ") + path);
    if (!get_run_info(f, info)) return fail(PUFF_H5_ERR_FORMAT, "runInfo lacks the mesh attributes");
    if (H5Lexists(f, "aperp", H5P_DEFAULT) <= 0) return fail(PUFF_H5_ERR_FORMAT, "no dataset aperp");
    Id d(H5Dopen2(f, "aperp", H5P_DEFAULT), H5Dclose);
    Id sp(H5Dget_space(d), H5Sclose);
    if (!d.ok() || !sp.ok()) return fail(PUFF_H5_ERR_IO, "cannot open aperp");
    get_position(d, info);
    int rank = H5Sget_simple_extent_ndims(sp);
    info->num_spatial_dims = rank == 2 ? 1 : 3;
    return PUFF_H5_OK;
}

int puff_h5_read_field(const char* path, int64_t nmesh, double* ar, double* ai) {
    Quiet q;
    if (!path || !ar || !ai) return fail(PUFF_H5_ERR_ARG, "puff_h5_read_field: bad arguments");
    puff_h5_info info;
    std::memset(&info, 0, sizeof(info));
    int st = puff_h5_read_field_info(path, &info);
    if (st != PUFF_H5_OK) return st;
    Id f(open_file(path), H5Fclose);
    Id d(H5Dopen2(f, "aperp", H5P_DEFAULT), H5Dclose);
    Id sp(H5Dget_space(d), H5Sclose);
    if (!d.ok() || !sp.ok()) return fail(PUFF_H5_ERR_IO, "cannot open aperp");
    if (!is_float(d)) return fail(PUFF_H5_ERR_FORMAT, "aperp is not floating point");
    int rank = H5Sget_simple_extent_ndims(sp);
    if (rank != 2 && rank != 4) return fail(PUFF_H5_ERR_FORMAT, "aperp is neither (2, NZ2) nor (2, NZ2, NY, NX)");
    hsize_t dims[4] = {0, 0, 1, 1};
    H5Sget_simple_extent_dims(sp, dims, nullptr);
    // the shape checks of readH5FieldfileSingleDump (h5_in.f90:815-884)
    if (dims[0] != 2 || dims[1] != (hsize_t)info.nz2 ||
        (rank == 4 && (dims[2] != (hsize_t)info.ny || dims[3] != (hsize_t)info.nx)))
        return fail(PUFF_H5_ERR_FORMAT, "aperp's shape differs from runInfo's mesh");
    const int64_t want = (int64_t)dims[1] * (int64_t)dims[2] * (int64_t)dims[3];
    if (want != nmesh) return fail(PUFF_H5_ERR_FORMAT, "aperp holds " + std::to_string(want) + " nodes, not " +
                                                          std::to_string(nmesh));
    if (!component_io(d, rank, dims, 0, ar, false) || !component_io(d, rank, dims, 1, ai, false))
        return fail(PUFF_H5_ERR_IO, "reading aperp failed");
    return PUFF_H5_OK;
}

int puff_h5_write_integrated(const char* path, const puff_h5_info* info, int32_t nz2, const double* power,
                             int32_t nharm, int32_t nslices, const double* bunching) {
    Quiet q;
    if (!path || !info || nz2 < 1 || !power || nharm < 0 || nharm > 5 || (nharm > 0 && (nslices < 1 || !bunching)))
        return fail(PUFF_H5_ERR_ARG, "puff_h5_write_integrated: bad arguments");
    static const char* names[5] = {"bunchingFundamental", "bunching2ndHarmonic", "bunching3ndHarmonic",
                                   "bunching4thHarmonic", "bunching5thHarmonic"};
    Id f(create_file(path), H5Fclose);
    if (!f.ok()) return fail(PUFF_H5_ERR_IO, std::string("cannot create ") + path);
    bool ok = true;
    if (info->num_spatial_dims == 1) {
        // 1-D: Intensity is |A|^2, power is |A|^2 x transArea (hdf5_puff.f90:203-212)
        std::vector<double> p((size_t)nz2);
        for (int32_t i = 0; i < nz2; ++i) p[(size_t)i] = power[i] * info->trans_area;
        ok = write_1d(f, "Intensity", power, (hsize_t)nz2, *info) && write_1d(f, "power", p.data(), (hsize_t)nz2, *info);
    } else {
        ok = write_1d(f, "power", power, (hsize_t)nz2, *info);
    }
    for (int32_t h = 0; ok && h < nharm; ++h)
        ok = write_1d(f, names[h], bunching + (size_t)h * (size_t)nslices, (hsize_t)nslices, *info);
    ok = ok && put_time_group(f, *info) && put_run_info(f, *info);
    if (!ok) return fail(PUFF_H5_ERR_IO, "writing integrated datasets failed");
    return H5Fflush(f, H5F_SCOPE_LOCAL) >= 0 ? PUFF_H5_OK : fail(PUFF_H5_ERR_IO, "flush failed");
}

int puff_h5_append_1d(const char* path, const puff_h5_info* info, const char* name, int64_t n,
                      const double* v) {
    Quiet q;
    if (!path || !info || !name || n < 1 || !v) return fail(PUFF_H5_ERR_ARG, "puff_h5_append_1d: bad arguments");
    Id f(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(PUFF_H5_ERR_IO, std::string("cannot open ") + path + " for writing");
    if (H5Lexists(f, name, H5P_DEFAULT) > 0) return fail(PUFF_H5_ERR_FORMAT, std::string(name) + " exists");
    if (!write_1d(f, name, v, (hsize_t)n, *info)) return fail(PUFF_H5_ERR_IO, std::string("writing ") + name + " failed");
    return H5Fflush(f, H5F_SCOPE_LOCAL) >= 0 ? PUFF_H5_OK : fail(PUFF_H5_ERR_IO, "flush failed");
}

int puff_h5_read_dataset_1d(const char* path, const char* name, int64_t cap, double* out, int64_t* len) {
    Quiet q;
    if (!path || !name || !len) return fail(PUFF_H5_ERR_ARG, "puff_h5_read_dataset_1d: bad arguments");
    Id f(open_file(path), H5Fclose);
    if (!f.ok()) return fail(PUFF_H5_ERR_IO, std::string("cannot open ") + path);
    if (H5Lexists(f, name, H5P_DEFAULT) <= 0) return fail(PUFF_H5_ERR_FORMAT, std::string("no dataset ") + name);
    Id d(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
    Id sp(H5Dget_space(d), H5Sclose);
    if (!d.ok() || !sp.ok() || H5Sget_simple_extent_ndims(sp) != 1 || !is_float(d))
        return fail(PUFF_H5_ERR_FORMAT, std::string(name) + " is not a 1-D floating-point dataset");
    hsize_t n = 0;
    H5Sget_simple_extent_dims(sp, &n, nullptr);
    *len = (int64_t)n;
    if (!out) return PUFF_H5_OK;
    if (cap < (int64_t)n) return fail(PUFF_H5_ERR_ARG, std::string(name) + ": output too small");
    if (n > 0 && H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
        return fail(PUFF_H5_ERR_IO, std::string("reading ") + name + " failed");
    return PUFF_H5_OK;
}

}  // extern "C"
