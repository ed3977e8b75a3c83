/*
 * mlau_io.h -- the reference's output and backup files, written and read by
 * libmlau_hip, so that its readers (the python/ and idl/ scripts and the MPI
 * merge tools) load results of this library unchanged and a run can stop and
 * resume.
 *
 * Fields are passed in solver argument order (1D ro,mx,my,mz,en,by,bz; 2D/3D
 * ro,mx,my,mz,en,bx,by,bz), ghost cells included, x fastest.
 *
 * rank < 0 selects the serial drivers' file names (outdat_CCCCC.dat, x.dat);
 * rank >= 0 the MPI drivers' (outdat_CCCCC_RRRRR.dat, x_XXXXX.dat), which the
 * merge tools read (one rank: rank 0 with every axis index 0).
 */
#ifndef MLAU_IO_H
#define MLAU_IO_H

#include "mlau.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { MLAU_IO_F64 = 0, MLAU_IO_F32 = 1 };
enum { MLAU_NO_BACKUP = 5 };  /* mlau_io_bkup_load: no backup in the directory */

/* Coordinate files, one "%.12f" value per line, ghosts included:
 * x.dat [y.dat] (rank < 0; 2D/SERIAL/OTvortex/init.hpp:10-25,
 * 1D/SERIAL/shock/main.cpp:133-136) or x_/y_/z_%05d.dat (rank >= 0, with the
 * rank's position along each axis; 2D/MPI/OTvortex/dataio.hpp:31-42,
 * 3D/MPI/OTvortex/dataio.hpp:34-52).  y and z may be NULL below dim. */
int mlau_io_write_grid(const char *dir, const mlau_grid *g, const double *x, const double *y, const double *z,
                       int rank, int ix, int iy, int iz);

/* One output record (replaces dataio()):
 *  first != 0: writes params.dat and offsets.dat (1D: params.dat "bx gamma"
 *  and xoff.dat) and starts t.dat; otherwise appends tim to t.dat (rank <= 0
 *  only, as the reference's rank 0).  Then writes the fields as
 *  outdat_<cnt> in FP64 (MLAU_IO_F64, the serial drivers,
 *  2D/SERIAL/OTvortex/dataio.hpp:1-38, 1D/SERIAL/shock/main.cpp:137-159) or
 *  FP32 (MLAU_IO_F32, the MPI drivers' conv_d2f, 2D/MPI/OTvortex/dataio.hpp:
 *  1-76, 3D/MPI/OTvortex/dataio.hpp:1-83). */
int mlau_io_write_output(const char *dir, const mlau_grid *g, int first, int cnt, double tim,
                         const double *const *fields, int fmt, int rank);
/* Read an outdat record back (FP32 records are widened to double). */
int mlau_io_read_output(const char *dir, const mlau_grid *g, int cnt, int fmt, int rank, double *const *fields);

/* Backup (2D/MPI/OTvortex/bkup.hpp:46-76, 3D likewise): bkup_stamp.dat
 * "n cnt", bkup_time.dat {tim, dt, trec} binary, bkup_data_<rank>.dat with
 * the eight fields in the backup order ro,mx,my,mz,bx,by,bz,en.  2D/3D. */
int mlau_io_bkup_save(const char *dir, long long nall, int n, int cnt, double tim, double dt, double trec, int rank,
                      const double *const *fields);
/* Load a backup (bkup.hpp:4-44); MLAU_NO_BACKUP when bkup_stamp.dat is
 * absent (the reference then starts from the initial state). */
int mlau_io_bkup_load(const char *dir, long long nall, int rank, int *n, int *cnt, double *tim, double *dt,
                      double *trec, double *const *fields);

#ifdef __cplusplus
}
#endif
#endif
