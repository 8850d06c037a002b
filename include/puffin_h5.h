/* Puffin-format HDF5 exchange: the particle, field and integrated files the
 * reference writes and reads, so that the same state can be handed to either
 * code and their outputs compared.  Host-only library (libpuffin_h5.so, over
 * the HDF5 C library); no GPU involved.
 *
 * Files and layouts (reference files under puffin/lib/io/):
 *   particles   dataset "electrons", FP64, C shape (N, 7), one row per
 *               macroparticle: x, y, z2, px, py, Gamma, chi-bar
 *               (outputH5BeamFilesSD hdf5_puff_coll.f90:43-665; read by
 *               readH5Beamfile h5_in.f90:416-720, columns 543-548)
 *   field       dataset "aperp", FP64, C shape (2, NZ2, NY, NX) on a 3-D
 *               mesh or (2, NZ2) in 1-D: real part, then imaginary, each x
 *               fastest (outputH5Field3DSD hdf5_puff_coll.f90:672-1072,
 *               outputH5Field1D2CompSD :1094-1473; read by
 *               readH5FieldfileSingleDump h5_in.f90:726-1143)
 *   integrated  1-D FP64 datasets "power" (and "Intensity" in 1-D) over the
 *               z2 nodes, "bunchingFundamental" ... "bunching5thHarmonic"
 *               over the slices (wr_h5 / addH5Field1DFloat hdf5_puff.f90:173-427,
 *               hdf5_puff_coll.f90:1487-1699)
 * Every file carries the group "runInfo" (mesh and scaling attributes) and
 * the position attributes (time, zbar*, z*, istep, iCsteps and the element
 * counters) on its dataset and on runInfo, with the reference's names and
 * HDF5 types (H5T_IEEE_F64LE, H5T_STD_I32LE).
 *
 * All functions return 0 on success and a negative code on failure, with a
 * message in puff_h5_last_error().  Counts and sizes are checked against the
 * file: a shape that does not match is an error, never a partial read.
 */
#ifndef PUFFIN_H5_H
#define PUFFIN_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum puff_h5_status {
    PUFF_H5_OK = 0,
    PUFF_H5_ERR_ARG = -1,      /* null pointer, bad size */
    PUFF_H5_ERR_IO = -2,       /* HDF5 call failed (open, create, read, write) */
    PUFF_H5_ERR_FORMAT = -3    /* file lacks a dataset or attribute, or its shape differs */
};

/* The run and position record of a Puffin file.  Names in comments are the
   reference's attribute names. */
typedef struct puff_h5_info {
    /* position in the machine */
    double time;            /* time (= zbarTotal) */
    double zbar_total;      /* zbarTotal */
    double zbar_local;      /* zbarLocal */
    double zbar_inter;      /* zbarInter */
    double z_total;         /* zTotal (m) */
    double z_local;         /* zLocal (m) */
    double z_inter;         /* zInter (m) */
    int32_t istep;          /* istep */
    int32_t icsteps;        /* iCsteps */
    int32_t iund_cr;        /* iUnd_cr */
    int32_t ichic_cr;       /* iChic_cr */
    int32_t idrift_cr;      /* iDrift_cr */
    int32_t iquad_cr;       /* iQuad_cr */
    int32_t imodulation_cr; /* iModulation_cr */
    int32_t il;             /* iL */
    int32_t iwrite_cr;      /* iWrite_cr */
    /* runInfo */
    int32_t nx, ny, nz2;    /* nX, nY, nZ2 */
    int32_t field_mesh;     /* fieldMesh: 0 temporal, 1 periodic */
    int32_t n_steps;        /* nSteps */
    int32_t num_spatial_dims; /* numSpatialDims: 1 or 3 */
    int32_t iscale;         /* iScale */
    int32_t reserved;
    double dx, dy, dz2;     /* sLengthOfElmX, sLengthOfElmY, sLengthOfElmZ2 */
    double rho, eta, kappa, gamma_r, aw; /* rho, eta, kappa, gamma_r, aw */
    double lg, lc;          /* Lg, Lc: gain and cooperation lengths (m) */
    double lambda_w, lambda_r; /* lambda_w, lambda_r (m) */
    double step_size;       /* sStepSize */
    double npk_bar;         /* npk_bar */
    double trans_area;      /* transArea */
    double trans_area_si;   /* transAreaSI */
} puff_h5_info;

const char* puff_h5_last_error(void);

/* Particles.  puff_h5_count_electrons gives N; puff_h5_read_electrons fills
   the seven length-n columns (n must equal N) and, when info is not null,
   the position attributes of the dataset (the runInfo group when present). */
int puff_h5_write_electrons(const char* path, const puff_h5_info* info, int64_t n, const double* x,
                            const double* y, const double* z2, const double* px, const double* py,
                            const double* gam, const double* chi);
int puff_h5_count_electrons(const char* path, int64_t* n);
int puff_h5_read_electrons(const char* path, int64_t n, double* x, double* y, double* z2, double* px,
                           double* py, double* gam, double* chi, puff_h5_info* info);

/* Field.  The mesh comes from info (nx, ny, nz2, num_spatial_dims) on
   writing and from runInfo on reading: puff_h5_read_field_info fills info,
   puff_h5_read_field the two nmesh-long parts (nmesh = nx ny nz2, 1-D:
   nz2), checking the dataset's shape against runInfo. */
int puff_h5_write_field(const char* path, const puff_h5_info* info, const double* ar, const double* ai);
int puff_h5_read_field_info(const char* path, puff_h5_info* info);
int puff_h5_read_field(const char* path, int64_t nmesh, double* ar, double* ai);

/* Integrated values: power over the nz2 nodes (as the device's puff_power
   gives it: in 1-D written as "Intensity" and, times transArea, as
   "power"), and |b_h| for harmonics 1..nharm (nharm <= 5) over nslices
   slices, bunching[h-1][slice]. */
int puff_h5_write_integrated(const char* path, const puff_h5_info* info, int32_t nz2, const double* power,
                             int32_t nharm, int32_t nslices, const double* bunching);

/* Add a 1-D FP64 dataset `name` (with the position attributes of info) to
   an existing file, such as the slice statistics and beam current of the
   integrated file (meanGamma, sigmaXbar, beamCurrent, ...,
   hdf5_puff.f90:229-416).  An existing dataset of that name is an error. */
int puff_h5_append_1d(const char* path, const puff_h5_info* info, const char* name, int64_t n,
                      const double* v);

/* Any 1-D FP64 dataset by name: *len receives its length; out (capacity
   cap) is filled when not null and cap >= *len. */
int puff_h5_read_dataset_1d(const char* path, const char* name, int64_t cap, double* out, int64_t* len);

#ifdef __cplusplus
}
#endif

#endif
