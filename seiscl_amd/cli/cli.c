/*
 * cli.c -- seis_cli_main: the SeisCL_MPI program on one GPU.
 *
 * Follows the reference's file contract: argument handling and file names
 * (src/SeisCL_MPI.c:47-161), the constants / sources / model / data /
 * residual files (readhdf5, src/read_hdf5.c:307-787), the shot loop with
 * residuals and rms (time_stepping, src/time_stepping.c:692-1012;
 * var_res_raw and res_scale, src/residuals.c) and the output files
 * (writehdf5, src/writehdf5.c:189-316).  The modeling itself runs through
 * the C-ABI of seiscl_amd.h.
 *
 * Messages go to stdout; stderr is written only when the run fails.
 */
#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "h5mat.h"
#include "seiscl_amd.h"

#define NAMELEN 4096

typedef struct {
    char model[NAMELEN], csts[NAMELEN], dout[NAMELEN], gout[NAMELEN], rms[NAMELEN],
        movie[NAMELEN], res[NAMELEN], din[NAMELEN], checkpoint[NAMELEN];
} files_t;

typedef struct {
    seis_csts c;
    int gradout, gradsrcout, resout, rmsout, movout, Hout, inputres, back_prop_type;
    int scalerms, scaleshot, scalermsnorm, restype, halfpar, fp16_file;
    float fmin, fmax;
    /* back_prop_type 2: the gradient frequencies (/gradfreqs, Hz) */
    int nfreqs;
    float *gradfreqs;
    /* sources and receivers, file order (src/read_hdf5.c:307-459) */
    int allns, allng, ns;
    float *src_pos, *src, *rec_pos;
    int *nsrc, *nrec, *src0, *rec0;
    /* model parameters, parameter layout */
    float *par[5];
    /* outputs selected for this run */
    int nout;
    const char *outs[10];
    float *din, *res, *out, *outres;   /* [nout][allng][NT] */
    double rms, rmsnorm;
    float *grad, *hess;                  /* [3][Pi N] */
    /* movie: the output variables that are wavefields, their snapshots
     * [nmov][ns][NT/movout][Pi N] (z fastest) */
    int nmov;
    const char *movs[10];
    float *mov;
    /* checkpoint file of the two-pass inputres flow (src/time_stepping.c:731-745) */
    const char *ckpt_path;
} job_t;

static int err(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "Error: ");
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
    return 1;
}

/* ---- reading ------------------------------------------------------------ */

static int read_scalar(hid_t f, const char *name, int type, void *v) {
    if (!seis_h5_exists(f, name)) return err("Variable %s is not defined", name);
    if (seis_h5_numel(f, name) != 1) return err("Variable %s must be a scalar", name + 1);
    if (seis_h5_read(f, name, type, v, 1)) return err("%s", seis_h5_last_error());
    return 0;
}

static int read_optional(hid_t f, const char *name, int type, void *v) {
    if (!seis_h5_exists(f, name)) return 0;
    if (seis_h5_numel(f, name) != 1) return err("Variable %s must be a scalar", name + 1);
    if (seis_h5_read(f, name, type, v, 1)) return err("%s", seis_h5_last_error());
    return 0;
}

static float *read_floats(hid_t f, const char *name, long long n) {
    float *p = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    if (!p) {
        err("out of host memory reading %s", name);
        return NULL;
    }
    if (seis_h5_read(f, name, SEIS_H5_F32, p, n)) {
        err("%s", seis_h5_last_error());
        free(p);
        return NULL;
    }
    return p;
}

/* read_csts, src/read_hdf5.c:461-552 */
static int read_csts(hid_t f, job_t *j) {
    seis_csts *c = &j->c;
    seis_csts_default(c);
    int s = 0;
    int ignore;
    s = s || read_scalar(f, "/NT", SEIS_H5_I32, &c->NT);
    s = s || read_scalar(f, "/ND", SEIS_H5_I32, &c->ND);
    s = s || read_scalar(f, "/dt", SEIS_H5_F32, &c->dt);
    s = s || read_scalar(f, "/dh", SEIS_H5_F32, &c->dh);
    s = s || read_scalar(f, "/FDORDER", SEIS_H5_I32, &c->FDORDER);
    s = s || read_scalar(f, "/MAXRELERROR", SEIS_H5_I32, &c->MAXRELERROR);
    s = s || read_scalar(f, "/freesurf", SEIS_H5_I32, &c->freesurf);
    s = s || read_scalar(f, "/nab", SEIS_H5_I32, &c->nab);
    s = s || read_scalar(f, "/abs_type", SEIS_H5_I32, &c->abs_type);
    s = s || read_scalar(f, "/L", SEIS_H5_I32, &c->L);
    s = s || read_scalar(f, "/gradout", SEIS_H5_I32, &j->gradout);
    s = s || read_scalar(f, "/gradsrcout", SEIS_H5_I32, &j->gradsrcout);
    s = s || read_scalar(f, "/seisout", SEIS_H5_I32, &c->seisout);
    s = s || read_scalar(f, "/resout", SEIS_H5_I32, &j->resout);
    s = s || read_scalar(f, "/rmsout", SEIS_H5_I32, &j->rmsout);
    /* device selection and MPI layout: one GPU here */
    s = s || read_scalar(f, "/pref_device_type", SEIS_H5_I32, &ignore);
    s = s || read_scalar(f, "/nmax_dev", SEIS_H5_I32, &ignore);
    s = s || read_scalar(f, "/MPI_NPROC_SHOT", SEIS_H5_I32, &ignore);
    s = s || read_scalar(f, "/back_prop_type", SEIS_H5_I32, &j->back_prop_type);
    s = s || read_scalar(f, "/f0", SEIS_H5_F32, &c->f0);
    s = s || read_scalar(f, "/abpc", SEIS_H5_F32, &c->abpc);
    s = s || read_scalar(f, "/VPPML", SEIS_H5_F32, &c->VPPML);
    s = s || read_scalar(f, "/FPML", SEIS_H5_F32, &c->FPML);
    s = s || read_scalar(f, "/NPOWER", SEIS_H5_F32, &c->NPOWER);
    s = s || read_scalar(f, "/K_MAX_CPML", SEIS_H5_F32, &c->K_MAX_CPML);
    if (s) return 1;
    float tmaxf = 0, tminf = 0;
    s = s || read_optional(f, "/fmax", SEIS_H5_F32, &j->fmax);
    s = s || read_optional(f, "/fmin", SEIS_H5_F32, &j->fmin);
    s = s || read_optional(f, "/scalerms", SEIS_H5_I32, &j->scalerms);
    s = s || read_optional(f, "/scaleshot", SEIS_H5_I32, &j->scaleshot);
    s = s || read_optional(f, "/scalermsnorm", SEIS_H5_I32, &j->scalermsnorm);
    s = s || read_optional(f, "/restype", SEIS_H5_I32, &j->restype);
    s = s || read_optional(f, "/Hout", SEIS_H5_I32, &j->Hout);
    s = s || read_optional(f, "/movout", SEIS_H5_I32, &j->movout);
    s = s || read_optional(f, "/param_type", SEIS_H5_I32, &c->param_type);
    s = s || read_optional(f, "/FP16", SEIS_H5_I32, &j->fp16_file);
    s = s || read_optional(f, "/inputres", SEIS_H5_I32, &j->inputres);
    s = s || read_optional(f, "/halfpar", SEIS_H5_I32, &j->halfpar);
    s = s || read_optional(f, "/tmax", SEIS_H5_F32, &tmaxf);
    s = s || read_optional(f, "/tmin", SEIS_H5_F32, &tminf);
    /* gradient corrections (seis_csts.gradfix): absent -> every correction
     * (seis_csts_default); 0 reproduces the reference's gradients */
    s = s || read_optional(f, "/gradfix", SEIS_H5_I32, &c->gradfix);
    if (s) return 1;
    if (c->gradfix < 0 || c->gradfix > 255) return err("gradfix=%d outside 0..255", c->gradfix);
    const int nd = c->ND == 3 ? 3 : 2;
    if (!seis_h5_exists(f, "/N")) return err("Variable /N is not defined");
    if (seis_h5_numel(f, "/N") != nd) return err("N must hold %d sizes for ND=%d", nd, c->ND);
    if (seis_h5_read(f, "/N", SEIS_H5_I32, c->N, nd)) return err("%s", seis_h5_last_error());
    /* FL: one relaxation frequency per Maxwell body (src/assign_modeling_case.c:1185) */
    if (c->L > 0) {
        if (c->L > SEIS_MAXL) return err("L=%d exceeds %d", c->L, SEIS_MAXL);
        if (!seis_h5_exists(f, "/FL")) return err("Variable /FL is not defined");
        if (seis_h5_numel(f, "/FL") != c->L) return err("FL must hold L=%d frequencies", c->L);
        if (seis_h5_read(f, "/FL", SEIS_H5_F32, c->FL, c->L)) return err("%s", seis_h5_last_error());
    }
    /* times in seconds -> steps, float arithmetic (src/read_hdf5.c:523-533) */
    c->tmax = tmaxf > 0 ? (int)(tmaxf / c->dt) : c->NT;
    c->tmin = tminf > 0 ? (int)(tminf / c->dt) : 0;
    if (c->tmin > c->tmax) return err("tmax<tmin");
    if (c->tmax > c->NT) return err("tmax exceeds the NT*dt time window");
    c->gradout = j->gradout;
    /* misfit options of var_res_raw / rtm_res (src/residuals.c:24-450) */
    c->restype = j->restype;
    c->scalerms = j->scalerms;
    c->scaleshot = j->scaleshot;
    c->scalermsnorm = j->scalermsnorm;
    c->fmin = j->fmin;
    c->fmax = j->fmax;
    c->Hout = j->Hout;
    /* precision: the reference's FP16=1 is vectorised FP32 (same results as
     * 0); 2 stores fields in half with FP32 arithmetic; 64 selects FP64.
     * 3 (half storage and half arithmetic) runs as 2: the updates are bound
     * by memory traffic, which half storage already halves, so half
     * arithmetic would save no time and only lose accuracy. */
    switch (j->fp16_file) {
        case 0: case 1: c->FP16 = 0; break;
        case 2: case 3: c->FP16 = 2; break;
        case 64: c->FP16 = 64; break;
        default: return err("FP16=%d is not a precision mode: use 0, 1, 2, 3 or 64",
                            j->fp16_file);
    }
    /* gradients by backpropagation (1) or from the DFT of the wavefields at
     * the frequencies /gradfreqs (2, src/assign_modeling_case.c:1189-1190) */
    if (j->gradout && j->back_prop_type != 1 && j->back_prop_type != 2)
        return err("back_prop_type=%d is not available: 1 (backpropagation) or 2 (DFT)",
                   j->back_prop_type);
    if (j->gradout && j->back_prop_type == 2) {
        if (!seis_h5_exists(f, "/gradfreqs")) return err("Variable /gradfreqs is not defined");
        const long long nf = seis_h5_numel(f, "/gradfreqs");
        if (nf < 1) return err("gradfreqs holds no frequency");
        j->nfreqs = (int)nf;
        j->gradfreqs = (float *)malloc(sizeof(float) * (size_t)nf);
        if (!j->gradfreqs) return err("out of host memory");
        if (seis_h5_read(f, "/gradfreqs", SEIS_H5_F32, j->gradfreqs, nf))
            return err("%s", seis_h5_last_error());
        if (j->inputres) return err("inputres needs back_prop_type=1");
        if (j->movout) return err("movout with gradout needs back_prop_type=1");
    }
    /* features outside this engine fail loudly rather than being ignored */
    if (j->halfpar) return err("half-precision parameters (halfpar) are not available");
    if (j->restype != 0 && j->restype != 1)
        return err("restype=%d is not available (0: L2 residuals, 1: RTM)", j->restype);
    if (j->inputres && j->resout) return err("Cannot input and output residuals simultaneously");
    return 0;
}

/* count the groups of ascending ids in column col of an [n][w] array */
static int count_ids(const float *a, int n, int w, const char *what, int *ngroups) {
    float thisid = -9999;
    int g = 0;
    for (int i = 0; i < n; ++i) {
        const float id = a[3 + i * w];
        if (thisid < id) {
            thisid = id;
            g++;
        } else if (thisid > id) {
            return err("%s ids must be sorted in ascending order", what);
        }
    }
    *ngroups = g;
    return 0;
}

static void group_sizes(const float *a, int n, int w, int *sizes, int *first) {
    int p = 0, cnt = 1;
    first[0] = 0;
    for (int i = 1; i < n; ++i) {
        if (a[3 + i * w] == a[3 + (i - 1) * w]) {
            cnt++;
        } else {
            sizes[p++] = cnt;
            first[p] = i;
            cnt = 1;
        }
    }
    sizes[p] = cnt;
}

/* read_srcrec, src/read_hdf5.c:307-459 */
static int read_srcrec(hid_t f, job_t *j) {
    int nd;
    long long d[SEIS_H5_MAXDIM];
    const char *need[3] = {"/src", "/src_pos", "/rec_pos"};
    for (int i = 0; i < 3; ++i)
        if (!seis_h5_exists(f, need[i])) return err("Variable %s is not defined", need[i]);
    if (seis_h5_dims(f, "/src_pos", &nd, d) || nd != 2) return err("src_pos must have 2 dimensions");
    if (d[1] != 5) return err("src_pos dimension 1 must be 5");
    j->allns = (int)d[0];
    if (seis_h5_dims(f, "/src", &nd, d) || nd != 2 || d[0] != j->allns || d[1] != j->c.NT)
        return err("src must be nt x number of sources");
    if (seis_h5_dims(f, "/rec_pos", &nd, d) || nd != 2) return err("rec_pos must have 2 dimensions");
    if (d[1] != 8) return err("rec_pos dimension 1 must be 8");
    j->allng = (int)d[0];
    if (j->allns < 1 || j->allng < 1) return err("no sources or no receivers");
    if (!(j->src_pos = read_floats(f, "/src_pos", (long long)j->allns * 5))) return 1;
    if (!(j->src = read_floats(f, "/src", (long long)j->allns * j->c.NT))) return 1;
    if (!(j->rec_pos = read_floats(f, "/rec_pos", (long long)j->allng * 8))) return 1;
    int ns, nsg;
    if (count_ids(j->src_pos, j->allns, 5, "Source", &ns)) return 1;
    if (count_ids(j->rec_pos, j->allng, 8, "Src", &nsg)) return 1;
    if (nsg != ns) return err("Number of sources ids in src_pos and rec_pos are not the same");
    j->ns = ns;
    j->nsrc = (int *)calloc(ns, sizeof(int));
    j->nrec = (int *)calloc(ns, sizeof(int));
    j->src0 = (int *)calloc(ns + 1, sizeof(int));
    j->rec0 = (int *)calloc(ns + 1, sizeof(int));
    if (!j->nsrc || !j->nrec || !j->src0 || !j->rec0) return err("out of host memory");
    group_sizes(j->src_pos, j->allns, 5, j->nsrc, j->src0);
    group_sizes(j->rec_pos, j->allng, 8, j->nrec, j->rec0);
    return 0;
}

/* parameter names in the model file (src/assign_modeling_case.c:1204-1240) */
static const char *par_name(int param_type, int k) {
    static const char *names[3][5] = {{"/vp", "/vs", "/rho", "/taup", "/taus"},
                                      {"/M", "/mu", "/rho", "/taup", "/taus"},
                                      {"/Ip", "/Is", "/rho", "/taup", "/taus"}};
    return names[param_type][k];
}

static int par_used(const seis_csts *c, int k) {
    if (k == 1 || k == 4) return c->ND != 22 && (k == 1 || c->L > 0);
    if (k == 3) return c->L > 0;
    return 1;
}

static long long n_cells(const seis_csts *c) {
    long long n = 1;
    for (int d = 0; d < (c->ND == 3 ? 3 : 2); ++d) n *= c->N[d];
    return n;
}

static int read_model(const char *path, job_t *j) {
    const seis_csts *c = &j->c;
    if (c->param_type < 0 || c->param_type > 2)
        return err("param_type=%d is not available (0, 1 or 2)", c->param_type);
    hid_t f = seis_h5_open(path, 0);
    if (f < 0) return err("Could not open the input file model");
    const int nd = c->ND == 3 ? 3 : 2;
    long long want[3];
    for (int i = 0; i < nd; ++i) want[i] = c->N[nd - 1 - i];
    int s = 0;
    for (int k = 0; k < 5 && !s; ++k) {
        if (!par_used(c, k)) continue;
        const char *name = par_name(c->param_type, k);
        int r;
        long long d[SEIS_H5_MAXDIM];
        if (!seis_h5_exists(f, name)) {
            s = err("Variable %s is not defined", name);
            break;
        }
        seis_h5_dims(f, name, &r, d);
        if (r != nd) {
            s = err("Variable %s must have %d dimensions", name, nd);
            break;
        }
        for (int i = 0; i < nd && !s; ++i)
            if (d[i] != want[i])
                s = err("Dimension %d of variable %s should be %lld, got %lld", i, name, want[i], d[i]);
        if (!s && !(j->par[k] = read_floats(f, name, n_cells(c)))) s = 1;
    }
    seis_h5_close(f);
    return s;
}

/* output sets per seisout, in the engine's order */
static int out_set(int ND, int seisout, const char **names) {
    static const char *e3[4][10] = {
        {"vx", "vy", "vz"},
        {"p"},
        {"sxx", "syy", "szz", "sxz", "sxy", "syz"},
        {"vx", "vy", "vz", "sxx", "syy", "szz", "sxz", "sxy", "syz", "p"}};
    static const int n3[4] = {3, 1, 6, 10};
    static const char *e2[4][6] = {{"vx", "vz"}, {"p"}, {"sxx", "szz", "sxz"},
                                   {"vx", "vz", "sxx", "szz", "sxz", "p"}};
    static const int n2[4] = {2, 1, 3, 6};
    static const char *a2[4][3] = {{"vx", "vz"}, {"p"}, {NULL}, {"vx", "vz", "p"}};
    static const int na[4] = {2, 1, 0, 3};
    if (seisout < 1 || seisout > 4) return 0;
    const int s = seisout - 1;
    int n = ND == 3 ? n3[s] : ND == 22 ? na[s] : n2[s];
    for (int i = 0; i < n; ++i) names[i] = ND == 3 ? e3[s][i] : ND == 22 ? a2[s][i] : e2[s][i];
    return n;
}

/* The outputs given by the datasets present in a data or residual file
 * (src/read_hdf5.c:627-758): they must form one of the seisout sets. */
static int outputs_from_file(hid_t f, job_t *j, const char *suffix, const char *what) {
    const char *all[10];
    int nall = out_set(j->c.ND, 4, all);
    int present[10] = {0}, npresent = 0;
    char name[64];
    for (int i = 0; i < nall; ++i) {
        snprintf(name, sizeof(name), "/%s%s", all[i], suffix);
        present[i] = seis_h5_exists(f, name);
        npresent += present[i];
    }
    if (!npresent) {
        if (suffix[0]) return err("Cannot output without the residuals");
        return err("Cannot output gradient, rms or residuals without reference data");
    }
    for (int so = 1; so <= 4; ++so) {
        const char *set[10];
        int n = out_set(j->c.ND, so, set);
        if (n != npresent) continue;
        int ok = 1;
        for (int i = 0; i < nall && ok; ++i) {
            if (!present[i]) continue;
            int found = 0;
            for (int k = 0; k < n; ++k) found |= !strcmp(set[k], all[i]);
            ok = found;
        }
        if (ok) {
            j->c.seisout = so;
            return 0;
        }
    }
    return err("the variables in the %s file must form one output set (vx,vy,vz | p | "
               "the stresses | all)", what);
}

/* per output, rows of [allng][NT]: data selected by receiver id
 * (read_seis, src/read_hdf5.c:61-93), residuals in file order */
static int read_traces(hid_t f, job_t *j, const char *suffix, float *dst, int by_recid) {
    const int NT = j->c.NT;
    char name[64];
    for (int k = 0; k < j->nout; ++k) {
        snprintf(name, sizeof(name), "/%s%s", j->outs[k], suffix);
        int nd;
        long long d[SEIS_H5_MAXDIM];
        if (seis_h5_dims(f, name, &nd, d) || nd != 2 || d[1] < NT || d[0] < (by_recid ? 1 : j->allng))
            return err("%s must be nt x number of recording stations", name + 1);
        float *all = read_floats(f, name, d[0] * d[1]);
        if (!all) return 1;
        for (int g = 0; g < j->allng; ++g) {
            long long row = by_recid ? (long long)j->rec_pos[4 + 8 * g] - 1 : g;
            if (row < 0 || row >= d[0]) {
                free(all);
                return err("receiver id %lld is outside %s", row + 1, name + 1);
            }
            memcpy(dst + ((size_t)k * j->allng + g) * NT, all + row * d[1], sizeof(float) * NT);
        }
        free(all);
    }
    return 0;
}

/* ---- writing ------------------------------------------------------------ */

static int write_f32(hid_t f, const char *name, const float *v, int nd, const long long *dims) {
    if (seis_h5_write(f, name, SEIS_H5_F32, v, nd, dims)) return err("%s", seis_h5_last_error());
    return 0;
}

static int write_outputs(const files_t *fs, job_t *j) {
    const seis_csts *c = &j->c;
    char name[64];
    int s = 0;
    if (j->movout > 0 && j->nmov) {
        /* mov<var>: [ns][NT/movout][NX][(NY)][NZ] (src/writehdf5.c:295-310) */
        hid_t f = seis_h5_create(fs->movie);
        if (f < 0) return err("Could not create %s", fs->movie);
        const int nd = c->ND == 3 ? 3 : 2;
        long long d[5] = {j->ns, c->NT / j->movout, 0, 0, 0};
        for (int i = 0; i < nd; ++i) d[2 + nd - 1 - i] = c->N[i];
        const long long per = (long long)j->ns * d[1] * n_cells(c);
        for (int v = 0; v < j->nmov && !s; ++v) {
            snprintf(name, sizeof(name), "mov%s", j->movs[v]);
            s = write_f32(f, name, j->mov + v * per, nd + 2, d);
        }
        seis_h5_close(f);
        if (s) return s;
    }
    if (c->seisout || j->resout) {
        hid_t f = seis_h5_create(fs->dout);
        if (f < 0) return err("Could not open the input/output file %s", fs->dout);
        long long d[2] = {j->allng, c->NT};
        for (int k = 0; k < j->nout && !s; ++k) {
            snprintf(name, sizeof(name), "%sout", j->outs[k]);
            s = write_f32(f, name, j->out + (size_t)k * j->allng * c->NT, 2, d);
        }
        if (j->resout && !j->inputres)
            for (int k = 0; k < j->nout && !s; ++k) {
                snprintf(name, sizeof(name), "%sres", j->outs[k]);
                s = write_f32(f, name, j->outres + (size_t)k * j->allng * c->NT, 2, d);
            }
        long long ds[2] = {j->allns, 5}, dr[2] = {j->allng, 8};
        if (!s) s = write_f32(f, "/src_pos", j->src_pos, 2, ds);
        if (!s) s = write_f32(f, "/rec_pos", j->rec_pos, 2, dr);
        seis_h5_close(f);
    }
    if (!s && j->rmsout) {
        hid_t f = seis_h5_create(fs->rms);
        if (f < 0) return err("Could not create %s", fs->rms);
        long long d[2] = {1, 1};
        float r = (float)sqrt(j->rms), rn = (float)sqrt(j->rmsnorm);
        s = write_f32(f, "/rms", &r, 2, d);
        if (!s) s = write_f32(f, "/rms_norm", &rn, 2, d);
        seis_h5_close(f);
    }
    if (!s && j->gradout) {
        hid_t f = seis_h5_create(fs->gout);
        if (f < 0) return err("Could not create %s", fs->gout);
        const int nd = c->ND == 3 ? 3 : 2;
        long long d[3];
        for (int i = 0; i < nd; ++i) d[nd - 1 - i] = c->N[i];
        const long long n = n_cells(c);
        const int npar = c->L > 0 && c->param_type <= 2 ? 5 : 3;
        for (int k = 0; k < npar && !s; ++k) {
            if (!par_used(c, k)) continue;
            snprintf(name, sizeof(name), "grad%s", par_name(c->param_type, k) + 1);
            s = write_f32(f, name, j->grad + (size_t)k * n, nd, d);
            if (!s && j->hess && k < 3) {
                /* H<name> (src/writehdf5.c:275-283) */
                snprintf(name, sizeof(name), "H%s", par_name(c->param_type, k) + 1);
                s = write_f32(f, name, j->hess + (size_t)k * n, nd, d);
            }
        }
        if (!s && j->gradsrcout) {
            /* /gradsrc [allns][NT] (src/writehdf5.c:285-288): the reference's
             * source-gradient terms are commented out of its adjoint kernels
             * (src/update_adjv2D.cl:341-357, src/update_adjv3D.cl:500-520,
             * src/update_adjs2D.cl:480-495), so the array it writes holds the
             * zeros of init_gradsrc; this build writes the same */
            long long ds[2] = {j->allns, c->NT};
            float *z = (float *)calloc((size_t)j->allns * c->NT + 1, sizeof(float));
            if (!z) s = err("out of host memory");
            else s = write_f32(f, "/gradsrc", z, 2, ds);
            free(z);
        }
        seis_h5_close(f);
    }
    return s;
}

/* ---- movie (movout, src/time_stepping.c:272-350) ---------------------------- */

/* The output variables that are wavefields (not the elastic pressure, which
 * the reference derives from the stresses and leaves out of the movie). */
static void movie_vars(job_t *j) {
    j->nmov = 0;
    for (int k = 0; k < j->nout; ++k)
        if (strcmp(j->outs[k], "p") != 0 || j->c.ND == 22) j->movs[j->nmov++] = j->outs[k];
}

/* Copy the interior of each movie variable into frame `frame` of shot sh. */
static int movie_frame(seis_model *m, job_t *j, int sh, int frame, float *pad) {
    const seis_csts *c = &j->c;
    const int H = c->FDORDER / 2, nd = c->ND == 3 ? 3 : 2;
    const int NZ = c->N[0], NY = nd == 3 ? c->N[1] : 1, NX = nd == 3 ? c->N[2] : c->N[1];
    const long long n = n_cells(c), nf = c->NT / j->movout;
    const long long pz = NZ + 2 * H, py = nd == 3 ? NY + 2 * H : 1, yo = nd == 3 ? H : 0;
    for (int v = 0; v < j->nmov; ++v) {
        if (seis_read_field(m, j->movs[v], pad)) return err("%s", seis_last_error());
        float *dst = j->mov + (((long long)v * j->ns + sh) * nf + frame) * n;
        for (int x = 0; x < NX; ++x)
            for (int y = 0; y < NY; ++y)
                for (int z = 0; z < NZ; ++z)
                    dst[((long long)x * NY + y) * NZ + z] =
                        pad[((x + H) * py + (y + yo)) * pz + z + H];
    }
    return 0;
}

/* One forward shot with snapshots after every step t with (t+1) % movout == 0. */
static int forward_movie(seis_model *m, job_t *j, int sh, int nsrc, const float *sp,
                         const float *src, int nrec, const float *rp, float *out) {
    const seis_csts *c = &j->c;
    const int H = c->FDORDER / 2, nd = c->ND == 3 ? 3 : 2;
    long long np = 1;
    for (int d = 0; d < nd; ++d) np *= c->N[d] + 2 * H;
    float *pad = (float *)malloc(sizeof(float) * np);
    if (!pad) return err("out of host memory");
    int s = seis_shot_bind(m, nsrc, sp, src, nrec, rp) || seis_launch_init(m);
    const int tmax = c->tmax > 0 ? c->tmax : c->NT;
    int t = 0;
    for (int f = 0; !s && f < c->NT / j->movout; ++f) {
        const int t1 = (f + 1) * j->movout;          /* steps done at frame f */
        if (t1 > tmax) break;
        if (seis_run_steps(m, t, t1 - t)) s = 1;
        t = t1;
        if (!s && movie_frame(m, j, sh, f, pad)) {
            free(pad);
            return 1;
        }
    }
    if (!s && t < tmax && seis_run_steps(m, t, tmax - t)) s = 1;
    if (!s && seis_shot_read(m, out)) s = 1;
    free(pad);
    return s ? err("%s", seis_last_error()) : 0;
}

/* One shot of gradient with the movie of the gradient run (movout with
 * gradout, src/time_stepping.c:926-929): seis_gradient_shot's steps through
 * the public entry points, the adjoint recording the frames. */
static int gradient_movie(seis_model *m, job_t *j, int sh, int nsrc, const float *sp,
                          const float *src, int nrec, const float *rp, const float *din,
                          const float *res_in, float *out, float *res, double *rr) {
    const seis_csts *c = &j->c;
    const long long n = n_cells(c), nf = c->NT / j->movout;
    const size_t nres = (size_t)j->nout * nrec * c->NT;
    float *frames = (float *)malloc(sizeof(float) * (size_t)(j->nmov * nf * n + 1));
    if (!frames) return err("out of host memory");
    int s = seis_forward_save(m, nsrc, sp, src, nrec, rp, out);
    if (!s && res_in) memcpy(res, res_in, sizeof(float) * nres);
    else if (!s) s = seis_residuals(m, nrec, out, din, res, &rr[0], &rr[1]);
    if (!s) s = seis_res_scale(m, nrec, rp, res);
    if (!s) s = seis_adjoint_movie(m, res, j->movout, j->nmov, (const char *const *)j->movs,
                                   frames);
    if (!s)
        for (int v = 0; v < j->nmov; ++v)
            memcpy(j->mov + ((long long)v * j->ns + sh) * nf * n,
                   frames + (long long)v * nf * n, sizeof(float) * nf * n);
    free(frames);
    return s ? err("%s", seis_last_error()) : 0;
}

/* ---- the run -------------------------------------------------------------- */

/* ---- two-pass inputres flow (src/time_stepping.c:92-270, 731-745, 851-859) --- */

/* Pass 1 (inputres=1, gradout=0): the forward run with its saved state (the
 * boundary shell, or the viscoelastic checkpoints) and that state written to
 * the checkpoint file as src<s>_dev0_checkpoint (checkpoint_d2h). */
static int checkpoint_save(seis_model *m, hid_t f, int sh, int nsrc, const float *sp,
                           const float *src, int nrec, const float *rp, float *so) {
    if (seis_forward_save(m, nsrc, sp, src, nrec, rp, so)) return err("%s", seis_last_error());
    size_t bytes = 0;
    if (seis_grad_saved_size(m, &bytes)) return err("%s", seis_last_error());
    float *buf = (float *)malloc(bytes);
    if (!buf) return err("out of host memory for the checkpoint (%zu bytes)", bytes);
    int s = 0;
    if (seis_grad_saved_read(m, buf)) {
        s = err("%s", seis_last_error());
    } else {
        char name[64];
        snprintf(name, sizeof(name), "src%d_dev0_checkpoint", sh);
        long long d[1] = {(long long)(bytes / sizeof(float))};
        s = write_f32(f, name, buf, 1, d);
    }
    free(buf);
    return s;
}

/* Pass 2 (inputres=1, gradout=1): the shot's saved state read back from the
 * checkpoint file instead of a forward run (checkpoint_h2d), then res_scale of
 * the given residuals and the adjoint run.  *found = 0 when the file has no
 * state for this shot. */
static int checkpoint_adjoint(seis_model *m, hid_t f, int sh, int nsrc, const float *sp,
                              const float *src, int nrec, const float *rp, const float *res,
                              float *sres, int nout, int NT, int *found) {
    char name[64];
    snprintf(name, sizeof(name), "src%d_dev0_checkpoint", sh);
    *found = f >= 0 && seis_h5_exists(f, name) > 0;
    if (!*found) return 0;
    const long long n = seis_h5_numel(f, name);
    if (n <= 0) return err("empty checkpoint %s", name);
    float *buf = (float *)malloc(sizeof(float) * (size_t)n);
    if (!buf) return err("out of host memory for the checkpoint");
    int s = 0;
    if (seis_h5_read(f, name, SEIS_H5_F32, buf, n)) s = err("%s", seis_h5_last_error());
    if (!s && (seis_shot_bind(m, nsrc, sp, src, nrec, rp) || seis_grad_saved_write(m, buf)))
        s = err("%s", seis_last_error());
    free(buf);
    if (s) return s;
    memcpy(sres, res, sizeof(float) * (size_t)nout * nrec * NT);
    if (seis_res_scale(m, nrec, rp, sres) || seis_adjoint(m, sres))
        return err("%s", seis_last_error());
    return 0;
}

static int run(job_t *j) {
    seis_ctx *ctx = NULL;
    seis_model *m = NULL;
    const seis_csts *c = &j->c;
    const int NT = c->NT;
    int s = 0;
    float *sd = NULL, *sr = NULL, *so = NULL, *sres = NULL;
    hid_t fck = -1;
    if (seis_ctx_create(0, &ctx)) return err("%s", seis_last_error());
    if (seis_model_create(ctx, c, j->par[0], j->par[1], j->par[2], j->par[3], j->par[4], &m)) {
        s = err("%s", seis_last_error());
        goto done;
    }
    if (seis_out_count(m) != j->nout) {
        s = err("internal: output count mismatch");
        goto done;
    }
    if (j->gradout && j->back_prop_type == 2 && seis_grad_dft(m, j->nfreqs, j->gradfreqs)) {
        s = err("%s", seis_last_error());
        goto done;
    }
    if (j->gradout && seis_grad_zero(m)) {
        s = err("%s", seis_last_error());
        goto done;
    }
    const int need_din = !j->inputres && (j->gradout || j->rmsout || j->resout);
    int ngmax = 0;
    for (int i = 0; i < j->ns; ++i) ngmax = j->nrec[i] > ngmax ? j->nrec[i] : ngmax;
    const size_t shot_n = (size_t)j->nout * ngmax * NT;
    so = (float *)malloc(sizeof(float) * shot_n);
    sd = (float *)malloc(sizeof(float) * shot_n);
    sr = (float *)malloc(sizeof(float) * shot_n);
    sres = (float *)malloc(sizeof(float) * shot_n);
    if (!so || !sd || !sr || !sres) {
        s = err("out of host memory");
        goto done;
    }
    /* the checkpoint file: created by pass 1, read by pass 2 (without it,
     * pass 2 recomputes each shot's forward run) */
    if (j->inputres && !j->gradout && j->ckpt_path) {
        remove(j->ckpt_path);
        if ((fck = seis_h5_create(j->ckpt_path)) < 0) {
            s = err("Could not create the checkpoint file %s", j->ckpt_path);
            goto done;
        }
    } else if (j->inputres && j->gradout && j->ckpt_path) {
        FILE *probe = fopen(j->ckpt_path, "rb");
        if (probe) {
            fclose(probe);
            fck = seis_h5_open(j->ckpt_path, 0);
        }
    }
    for (int sh = 0; sh < j->ns && !s; ++sh) {
        const int nsrc = j->nsrc[sh], nrec = j->nrec[sh], g0 = j->rec0[sh];
        const float *sp = j->src_pos + (size_t)j->src0[sh] * 5;
        const float *src = j->src + (size_t)j->src0[sh] * NT;
        const float *rp = j->rec_pos + (size_t)g0 * 8;
        /* gather this shot's rows of the [nout][allng][NT] arrays */
        for (int k = 0; k < j->nout; ++k)
            for (int g = 0; g < nrec; ++g) {
                const size_t dst = ((size_t)k * nrec + g) * NT;
                const size_t src_row = ((size_t)k * j->allng + g0 + g) * NT;
                if (j->din) memcpy(sd + dst, j->din + src_row, sizeof(float) * NT);
                if (j->res) memcpy(sr + dst, j->res + src_row, sizeof(float) * NT);
            }
        fprintf(stdout, "    shot %d of %d: %d source(s), %d receiver(s)\n", sh + 1, j->ns,
                nsrc, nrec);
        fflush(stdout);
        int from_ckpt = 0;
        if (j->inputres && j->gradout && !(j->movout > 0 && j->nmov) && fck >= 0) {
            if (checkpoint_adjoint(m, fck, sh, nsrc, sp, src, nrec, rp, sr, sres, j->nout, NT,
                                   &from_ckpt)) {
                s = 1;
                break;
            }
            /* no forward run in this pass: its seismograms are not produced */
            if (from_ckpt) memset(so, 0, sizeof(float) * (size_t)j->nout * nrec * NT);
        }
        if (from_ckpt) {
        } else if (j->inputres && !j->gradout && fck >= 0 && !(j->movout > 0)) {
            if (checkpoint_save(m, fck, sh, nsrc, sp, src, nrec, rp, so)) {
                s = 1;
                break;
            }
        } else if (j->gradout && j->movout > 0 && j->nmov) {
            double rr[2] = {0, 0};
            if (gradient_movie(m, j, sh, nsrc, sp, src, nrec, rp, need_din ? sd : NULL,
                               j->inputres ? sr : NULL, so, sres, rr)) {
                s = 1;
                break;
            }
            if (need_din) {
                j->rms += rr[0];
                j->rmsnorm += rr[1];
            }
        } else if (j->gradout) {
            double rr[2] = {0, 0};
            if (seis_gradient_shot(m, nsrc, sp, src, nrec, rp, need_din ? sd : NULL,
                                   j->inputres ? sr : NULL, so, rr)) {
                s = err("%s", seis_last_error());
                break;
            }
            if (need_din) {
                j->rms += rr[0];
                j->rmsnorm += rr[1];
            }
        } else if (j->movout > 0) {
            if (forward_movie(m, j, sh, nsrc, sp, src, nrec, rp, so)) {
                s = 1;
                break;
            }
        } else if (seis_forward_shot(m, nsrc, sp, src, nrec, rp, so)) {
            s = err("%s", seis_last_error());
            break;
        }
        if (!j->gradout && need_din) {
            double r = 0, rn = 0;
            if (seis_residuals(m, nrec, so, sd, sres, &r, &rn)) {
                s = err("%s", seis_last_error());
                break;
            }
            j->rms += r;
            j->rmsnorm += rn;
        }
        if (j->resout) {
            /* the scaled adjoint sources are what the reference writes as
             * <var>res (res_scale runs in place, src/time_stepping.c:845-847) */
            double r = 0, rn = 0;
            if (seis_residuals(m, nrec, so, sd, sres, &r, &rn) ||
                seis_res_scale(m, nrec, rp, sres)) {
                s = err("%s", seis_last_error());
                break;
            }
        }
        for (int k = 0; k < j->nout; ++k)
            for (int g = 0; g < nrec; ++g) {
                const size_t from = ((size_t)k * nrec + g) * NT;
                const size_t to = ((size_t)k * j->allng + g0 + g) * NT;
                memcpy(j->out + to, so + from, sizeof(float) * NT);
                if (j->resout) memcpy(j->outres + to, sres + from, sizeof(float) * NT);
            }
    }
    if (!s && j->gradout) {
        /* grad<name> of every parameter (src/writehdf5.c:261-283): the
         * viscoelastic case adds taup and taus */
        const int ntau = c->L > 0 && c->param_type <= 2 ? 2 : 0;
        j->grad = (float *)calloc((size_t)(3 + ntau) * n_cells(c), sizeof(float));
        if (!j->grad) s = err("out of host memory");
        else if (seis_grad_read(m, j->grad)) s = err("%s", seis_last_error());
        else if (ntau && seis_grad_read_tau(m, j->grad + 3 * (size_t)n_cells(c)))
            s = err("%s", seis_last_error());
        if (!s && j->Hout) {
            j->hess = (float *)malloc(sizeof(float) * 3 * (size_t)n_cells(c));
            if (!j->hess) s = err("out of host memory");
            else if (seis_hess_read(m, j->hess)) s = err("%s", seis_last_error());
        }
    }
done:
    if (fck >= 0) seis_h5_close(fck);
    free(so);
    free(sd);
    free(sr);
    free(sres);
    if (m) seis_model_destroy(m);
    if (ctx) seis_ctx_destroy(ctx);
    return s;
}

static void job_free(job_t *j) {
    free(j->gradfreqs);
    free(j->src_pos);
    free(j->src);
    free(j->rec_pos);
    free(j->nsrc);
    free(j->nrec);
    free(j->src0);
    free(j->rec0);
    for (int k = 0; k < 5; ++k) free(j->par[k]);
    free(j->din);
    free(j->res);
    free(j->out);
    free(j->outres);
    free(j->grad);
    free(j->hess);
    free(j->mov);
}

static int file_names(const char *base, const char *din, files_t *fs) {
    int over = 0;
#define NAME(field, suffix) \
    over |= snprintf(fs->field, NAMELEN, "%s%s", base, suffix) >= NAMELEN
    NAME(model, "_model.mat");
    NAME(csts, "_csts.mat");
    NAME(dout, "_dout.mat");
    NAME(gout, "_gout.mat");
    NAME(rms, "_rms.mat");
    NAME(movie, "_movie.mat");
    NAME(res, "_res.mat");
    /* one process: group id 0 (src/SeisCL_MPI.c:120-121) */
    NAME(checkpoint, "_NP0_checkpoint.mat");
#undef NAME
    if (din) over |= snprintf(fs->din, NAMELEN, "%s", din) >= NAMELEN;
    else over |= snprintf(fs->din, NAMELEN, "%s_din.mat", base) >= NAMELEN;
    return over ? err("file name too long") : 0;
}

int seis_cli_main(int argc, char **argv) {
    const char *base = "SeisCL", *din = NULL;
    int npos = 0;
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        if (a[0] == '-' && a[1]) {
            for (const char *o = a + 1; *o; ++o) {
                if (*o == 'h') {
                    fprintf(stdout, "usage: SeisCL_MPI [-p] <base> [<din>]\n");
                } else if (*o == 'p') {
                    /* the reference prints its generated kernels; the kernels
                     * here are compiled ahead of time */
                    fprintf(stdout, "-p: kernels are precompiled for gfx950\n");
                } else {
                    fprintf(stderr, "Unknown option character `\\x%x'.\n", (unsigned char)*o);
                    return 1;
                }
            }
            continue;
        }
        if (npos == 0) base = a;
        else if (npos == 1) din = a;
        else return err("Too many default argument, expected 2");
        npos++;
    }
    files_t fs;
    if (file_names(base, din, &fs)) return 1;
    fprintf(stdout, "\nInput files for SeisCL: \n    model: %s \n    constants: %s \n"
                    "    output data: %s \n    output gradient: %s \n    output rms: %s \n"
                    "    input data: %s \n\n",
            fs.model, fs.csts, fs.dout, fs.gout, fs.rms, fs.din);
    fflush(stdout);

    job_t j;
    memset(&j, 0, sizeof(j));
    j.ckpt_path = fs.checkpoint;
    int s = 0;
    hid_t f = seis_h5_open(fs.csts, 0);
    if (f < 0) {
        s = err("Could not open the input file csts");
    } else {
        s = read_csts(f, &j);
        if (!s) s = read_srcrec(f, &j);
        seis_h5_close(f);
    }
    if (!s) s = read_model(fs.model, &j);
    /* which outputs: the data or residual file decides when residuals are
     * needed, seisout otherwise (src/assign_modeling_case.c:1243-1326) */
    if (!s && !j.inputres && (j.gradout || j.rmsout || j.resout)) {
        hid_t fd = seis_h5_open(fs.din, 0);
        if (fd < 0) {
            s = j.gradout ? err("Could not open the input file for data_in")
                          : err("Cannot output gradient, rms or residuals without reference data");
        } else {
            s = outputs_from_file(fd, &j, "", "data");
            if (!s) {
                j.nout = out_set(j.c.ND, j.c.seisout, j.outs);
                j.din = (float *)malloc(sizeof(float) * (size_t)j.nout * j.allng * j.c.NT);
                s = j.din ? read_traces(fd, &j, "", j.din, 1) : err("out of host memory");
            }
            seis_h5_close(fd);
        }
    } else if (!s && j.inputres && j.gradout) {
        hid_t fr = seis_h5_open(fs.res, 0);
        if (fr < 0) {
            s = err("Could not open the residuals file");
        } else {
            s = outputs_from_file(fr, &j, "res", "residuals");
            if (!s) {
                j.nout = out_set(j.c.ND, j.c.seisout, j.outs);
                j.res = (float *)malloc(sizeof(float) * (size_t)j.nout * j.allng * j.c.NT);
                s = j.res ? read_traces(fr, &j, "res", j.res, 0) : err("out of host memory");
            }
            seis_h5_close(fr);
        }
    } else if (!s) {
        j.nout = out_set(j.c.ND, j.c.seisout, j.outs);
        if (!j.nout) s = err("seisout=%d has no outputs for ND=%d", j.c.seisout, j.c.ND);
    }
    if (!s && j.movout > 0) {
        movie_vars(&j);
        const size_t n = (size_t)j.nmov * j.ns * (j.c.NT / j.movout) * n_cells(&j.c);
        if (n && !(j.mov = (float *)calloc(n, sizeof(float)))) s = err("out of host memory");
    }
    if (!s) {
        const size_t n = (size_t)j.nout * j.allng * j.c.NT;
        j.out = (float *)calloc(n, sizeof(float));
        if (j.resout) j.outres = (float *)calloc(n, sizeof(float));
        if (!j.out || (j.resout && !j.outres)) s = err("out of host memory");
    }
    if (!s) {
        fprintf(stdout, "Modeling %d shot(s), ND=%d, N=[%d %d%s], NT=%d\n", j.ns, j.c.ND,
                j.c.N[0], j.c.N[1], j.c.ND == 3 ? " ..." : "", j.c.NT);
        fflush(stdout);
        s = run(&j);
    }
    if (!s) s = write_outputs(&fs, &j);
    if (!s) {
        fprintf(stdout, "SeisCL run done\n");
        fflush(stdout);
    }
    job_free(&j);
    return s;
}
