/*
 * seiscl_cli.h -- the file-level boundary of SeisCL: the SeisCL_MPI program
 * contract and the HDF5 (.mat) files it exchanges with SeisCL.py.
 *
 * libseiscl_cli.so holds both: the HDF5 helpers (plain C over the image's
 * libhdf5) and seis_cli_main, the body of the SeisCL_MPI executable, which
 * runs the modeling through the C-ABI of seiscl_amd.h.
 */
#ifndef SEISCL_CLI_H
#define SEISCL_CLI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- SeisCL_MPI [-p] <base> [<din>]  (src/SeisCL_MPI.c:47-161) ------------
 * Reads <base>_csts.mat, <base>_model.mat, <din> (default <base>_din.mat)
 * and <base>_res.mat (readhdf5, src/read_hdf5.c:307-787); runs every shot
 * (time_stepping, src/time_stepping.c:692-1012) on one GPU; writes
 * <base>_dout.mat, <base>_gout.mat and <base>_rms.mat (writehdf5,
 * src/writehdf5.c:189-316).  Progress goes to stdout, errors to stderr only
 * on failure (SeisCL.py raises on any stderr output, SeisCL/SeisCL.py:498-
 * 503).  Returns the process exit status: 0 on success. */
int seis_cli_main(int argc, char **argv);

/* ---- HDF5 helpers ---------------------------------------------------------
 * Element types of the helpers below. */
#define SEIS_H5_F32 0
#define SEIS_H5_F64 1
#define SEIS_H5_I32 2
#define SEIS_H5_I64 3
#define SEIS_H5_MAXDIM 8

const char *seis_h5_last_error(void);
/* New file with the MATLAB 7.3 user block (create_file, src/writehdf5.c:
 * 166-188). */
int seis_h5_file_create(const char *path);
/* Write a dataset, dims in file (C) order; floats are chunked and deflated
 * as writetomat does (src/writehdf5.c:24-96). */
int seis_h5_file_write(const char *path, const char *name, int type, const void *buf,
                       int ndim, const long long *dims);
/* Rank and dims of a dataset (getdimmat, src/read_hdf5.c:120-145). */
int seis_h5_file_dims(const char *path, const char *name, int *ndim, long long *dims);
/* Read a whole dataset of n elements, converted to type (readvar, src/
 * read_hdf5.c:95-110). */
int seis_h5_file_read(const char *path, const char *name, int type, void *buf, long long n);
/* Root dataset names separated by '\n'. */
int seis_h5_file_list(const char *path, char *out, size_t cap);

#ifdef __cplusplus
}
#endif

#endif
