/*
 * h5mat.c -- HDF5 files in the layout SeisCL exchanges with its callers.
 *
 * SeisCL's files are HDF5 with a 512-byte MATLAB 7.3 user block; datasets
 * are stored in C order with the Python array's axes reversed (SeisCL.py
 * writes np.transpose(a), SeisCL/SeisCL.py:658-763), floats as F32LE in
 * 8-element chunks deflated at level 6 with a MATLAB_class attribute
 * (src/writehdf5.c:24-96, 166-188).  Readers convert any stored numeric type
 * to the requested one, as the reference does with H5T_NATIVE_* memory types
 * (src/read_hdf5.c:95-110).
 *
 * Plain C on purpose: this library links the image's libhdf5 and nothing
 * C++, so it can carry a run path to that library without pulling another
 * C++ runtime into the process.
 */
#include "h5mat.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

static _Thread_local char h5_err[512];

const char *seis_h5_last_error(void) { return h5_err; }

static int h5_fail(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(h5_err, sizeof(h5_err), fmt, ap);
    va_end(ap);
    return 1;
}

static void quiet(void) {
    static int done = 0;
    if (!done) {
        H5Eset_auto2(H5E_DEFAULT, NULL, NULL);   /* no HDF5 error stack on stderr */
        done = 1;
    }
}

static hid_t mem_type(int type) {
    switch (type) {
        case SEIS_H5_F32: return H5T_NATIVE_FLOAT;
        case SEIS_H5_F64: return H5T_NATIVE_DOUBLE;
        case SEIS_H5_I32: return H5T_NATIVE_INT;
        case SEIS_H5_I64: return H5T_NATIVE_LLONG;
        default: return -1;
    }
}

static hid_t file_type(int type) {
    switch (type) {
        case SEIS_H5_F32: return H5T_IEEE_F32LE;
        case SEIS_H5_F64: return H5T_IEEE_F64LE;
        case SEIS_H5_I32: return H5T_STD_I32LE;
        case SEIS_H5_I64: return H5T_STD_I64LE;
        default: return -1;
    }
}

static const char *matlab_class(int type) {
    switch (type) {
        case SEIS_H5_F32: return "single";
        case SEIS_H5_F64: return "double";
        case SEIS_H5_I32: return "int32";
        default: return "int64";
    }
}

hid_t seis_h5_open(const char *path, int writable) {
    quiet();
    return H5Fopen(path, writable ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
}

void seis_h5_close(hid_t f) {
    if (f >= 0) H5Fclose(f);
}

/* A new file with the 512-byte user block holding a MATLAB 7.3 header. */
hid_t seis_h5_create(const char *path) {
    quiet();
    hid_t fcpl = H5Pcreate(H5P_FILE_CREATE);
    H5Pset_userblock(fcpl, 512);
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, fcpl, H5P_DEFAULT);
    H5Pclose(fcpl);
    if (f < 0) return f;
    H5Fclose(f);
    FILE *fp = fopen(path, "r+b");
    if (!fp) return -1;
    char head[128];
    memset(head, ' ', sizeof(head));
    const char *text = "MATLAB 7.3 MAT-file, Platform: GLNXA64, Created by: seiscl_amd "
                       "HDF5 schema 1.00 .";
    memcpy(head, text, strlen(text));
    /* bytes 116..123: subsystem offset (zero); 124..127: version 0x0200 and
     * the endian indicator "IM" as MATLAB writes them */
    memset(head + 116, 0, 8);
    head[124] = 0x00;
    head[125] = 0x02;
    head[126] = 'I';
    head[127] = 'M';
    fwrite(head, 1, sizeof(head), fp);
    fclose(fp);
    return H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
}

int seis_h5_exists(hid_t f, const char *name) {
    if (f < 0 || !name || !*name) return 0;
    return H5Lexists(f, name, H5P_DEFAULT) > 0;
}

int seis_h5_dims(hid_t f, const char *name, int *ndim, long long *dims) {
    if (!seis_h5_exists(f, name)) return h5_fail("Variable %s is not defined", name);
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (d < 0) return h5_fail("Cannot open variable %s", name);
    hid_t sp = H5Dget_space(d);
    int r = H5Sget_simple_extent_ndims(sp);
    hsize_t h[SEIS_H5_MAXDIM];
    if (r < 0 || r > SEIS_H5_MAXDIM) {
        H5Sclose(sp);
        H5Dclose(d);
        return h5_fail("Variable %s has an unsupported rank", name);
    }
    H5Sget_simple_extent_dims(sp, h, NULL);
    *ndim = r;
    for (int i = 0; i < r; ++i) dims[i] = (long long)h[i];
    H5Sclose(sp);
    H5Dclose(d);
    return 0;
}

long long seis_h5_numel(hid_t f, const char *name) {
    int nd = 0;
    long long dims[SEIS_H5_MAXDIM];
    if (seis_h5_dims(f, name, &nd, dims)) return -1;
    long long n = 1;
    for (int i = 0; i < nd; ++i) n *= dims[i];
    return n;
}

int seis_h5_read(hid_t f, const char *name, int type, void *buf, long long n) {
    long long have = seis_h5_numel(f, name);
    if (have < 0) return 1;
    if (have != n)
        return h5_fail("Variable %s has %lld elements, expected %lld", name, have, n);
    hid_t mt = mem_type(type);
    if (mt < 0) return h5_fail("unknown element type %d", type);
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    herr_t e = H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
    H5Dclose(d);
    if (e < 0) return h5_fail("Cannot read variable %s", name);
    return 0;
}

int seis_h5_write(hid_t f, const char *name, int type, const void *buf, int ndim,
                  const long long *dims) {
    hid_t mt = mem_type(type), ft = file_type(type);
    if (mt < 0) return h5_fail("unknown element type %d", type);
    if (ndim < 0 || ndim > SEIS_H5_MAXDIM) return h5_fail("rank %d out of range", ndim);
    hsize_t h[SEIS_H5_MAXDIM], ch[SEIS_H5_MAXDIM];
    long long n = 1;
    for (int i = 0; i < ndim; ++i) {
        h[i] = (hsize_t)dims[i];
        n *= dims[i];
        ch[i] = h[i] < 8 ? h[i] : 8;
    }
    if (seis_h5_exists(f, name)) {
        /* an existing dataset of the same shape is overwritten in place */
        int nd0;
        long long d0[SEIS_H5_MAXDIM];
        if (seis_h5_dims(f, name, &nd0, d0)) return 1;
        int same = nd0 == ndim;
        for (int i = 0; same && i < ndim; ++i) same = d0[i] == dims[i];
        if (!same) return h5_fail("Variable %s exists with another shape", name);
        hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
        herr_t e = H5Dwrite(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
        H5Dclose(d);
        return e < 0 ? h5_fail("Cannot write variable %s", name) : 0;
    }
    hid_t sp = ndim ? H5Screate_simple(ndim, h, NULL) : H5Screate(H5S_SCALAR);
    hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
    if (ndim && n > 0) {
        ch[0] = h[0];   /* the first axis in one chunk, as src/writehdf5.c:56 */
        H5Pset_chunk(pl, ndim, ch);
        H5Pset_deflate(pl, 6);
    }
    hid_t d = H5Dcreate2(f, name, ft, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
    H5Pclose(pl);
    H5Sclose(sp);
    if (d < 0) return h5_fail("Cannot create variable %s", name);
    herr_t e = n > 0 ? H5Dwrite(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) : 0;
    /* MATLAB_class attribute: a fixed-length string */
    const char *cls = matlab_class(type);
    hid_t st = H5Tcopy(H5T_C_S1);
    H5Tset_size(st, strlen(cls));
    hid_t as = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(d, "MATLAB_class", st, as, H5P_DEFAULT, H5P_DEFAULT);
    if (a >= 0) {
        H5Awrite(a, st, cls);
        H5Aclose(a);
    }
    H5Sclose(as);
    H5Tclose(st);
    H5Dclose(d);
    return e < 0 ? h5_fail("Cannot write variable %s", name) : 0;
}

/* ---- path-level wrappers for foreign callers (ctypes) ------------------- */

int seis_h5_file_create(const char *path) {
    hid_t f = seis_h5_create(path);
    if (f < 0) return h5_fail("Could not create %s", path);
    H5Fclose(f);
    return 0;
}

int seis_h5_file_write(const char *path, const char *name, int type, const void *buf,
                       int ndim, const long long *dims) {
    hid_t f = seis_h5_open(path, 1);
    if (f < 0) return h5_fail("Could not open %s", path);
    int r = seis_h5_write(f, name, type, buf, ndim, dims);
    H5Fclose(f);
    return r;
}

int seis_h5_file_dims(const char *path, const char *name, int *ndim, long long *dims) {
    hid_t f = seis_h5_open(path, 0);
    if (f < 0) return h5_fail("Could not open %s", path);
    int r = seis_h5_dims(f, name, ndim, dims);
    H5Fclose(f);
    return r;
}

int seis_h5_file_read(const char *path, const char *name, int type, void *buf, long long n) {
    hid_t f = seis_h5_open(path, 0);
    if (f < 0) return h5_fail("Could not open %s", path);
    int r = seis_h5_read(f, name, type, buf, n);
    H5Fclose(f);
    return r;
}

/* Names of the datasets at the root of a file, separated by '\n'. */
static herr_t add_name(hid_t g, const char *name, const H5L_info_t *info, void *op) {
    (void)g;
    (void)info;
    char *out = ((char **)op)[0];
    size_t cap = (size_t)((char **)op)[1];
    size_t len = strlen(out);
    if (len + strlen(name) + 2 > cap) return 1;
    strcat(out, name);
    strcat(out, "\n");
    return 0;
}

int seis_h5_file_list(const char *path, char *out, size_t cap) {
    if (!out || cap == 0) return h5_fail("no output buffer");
    out[0] = 0;
    hid_t f = seis_h5_open(path, 0);
    if (f < 0) return h5_fail("Could not open %s", path);
    void *op[2] = {out, (void *)cap};
    herr_t e = H5Literate(f, H5_INDEX_NAME, H5_ITER_INC, NULL, add_name, op);
    H5Fclose(f);
    return e != 0 ? h5_fail("listing %s failed or overflowed", path) : 0;
}
