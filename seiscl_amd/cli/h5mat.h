/* Handle-level HDF5 helpers shared by the SeisCL_MPI front end (h5mat.c). */
#ifndef SEISCL_H5MAT_H
#define SEISCL_H5MAT_H

#include <hdf5.h>

#include "seiscl_cli.h"

hid_t seis_h5_open(const char *path, int writable);
hid_t seis_h5_create(const char *path);
void seis_h5_close(hid_t f);
int seis_h5_exists(hid_t f, const char *name);
int seis_h5_dims(hid_t f, const char *name, int *ndim, long long *dims);
long long seis_h5_numel(hid_t f, const char *name);
int seis_h5_read(hid_t f, const char *name, int type, void *buf, long long n);
int seis_h5_write(hid_t f, const char *name, int type, const void *buf, int ndim,
                  const long long *dims);

#endif
