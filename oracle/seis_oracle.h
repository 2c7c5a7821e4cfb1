/*
 * seis_oracle.h -- CPU restatement of SeisCL's time-domain FD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, the smoke entry point and
 * the CPU-baseline leg of bench.py may load it; the product library never
 * links it and never falls back to it.
 *
 * Layout conventions follow the reference (src/F.h, src/update_v3D.cl:128-129):
 *   - N[0]=NZ, N[1]=NX (2D) or N[0]=NZ, N[1]=NY, N[2]=NX (3D), z fastest;
 *   - wavefields are padded by FDOH on every side:
 *       indv = x*NYp*NZp + y*NZp + z  (padded sizes, padded coordinates);
 *   - parameters and gradients are unpadded: indp = x*NY*NZ + y*NZ + z;
 *   - memory variables: r[l*NXp*NYp*NZp + indv];
 *   - CPML strips: src/update_v3D.cl:23-31.
 *
 * Pins: tests/test_oracle_*.py and tests/test_misfit_cpu.py check every
 * function against closed forms, invariants or the reference's own test
 * geometry.  Parity unpinned: the DFT gradient's per-cell coefficients
 * beyond what its Parseval correlation with the time-domain gradient fixes,
 * and the reference's 3D memory-variable index quirks that its restatement
 * keeps (tests/test_oracle_dft.py names them).
 */
#ifndef SEIS_ORACLE_H
#define SEIS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXL 8
#define OR_MAXF 512   /* gradient frequencies of the DFT gradient */
#define OR_FIX_RES_SCALE 1
#define OR_FIX_ADJ_CPML 2
#define OR_FIX_RHO_SCALE 4
#define OR_FIX_SURF_SYY 8
#define OR_FIX_VISCO_F0 16
/* checker only: elastic runs also read the forward state from the stored
 * history (to pin that path against backpropagation) */
#define OR_HISTORY 32
#define OR_FIX_RMSNORM 64
#define OR_FIX_HESS 128

typedef struct or_csts {
    int ND;            /* 2: 2D P-SV elastic, 3: 3D, 22: 2D acoustic           */
    int N[3];          /* see header comment                                     */
    int NT;            /* number of time samples                                 */
    int tmax;          /* forward steps (0 -> NT)                                */
    int tmin;          /* first adjoint step bound (default 0)                   */
    float dt, dh;
    int FDORDER, MAXRELERROR;
    int freesurf, abs_type, nab;
    float abpc, VPPML, NPOWER, FPML, K_MAX_CPML;
    int L;
    float f0;
    float FL[OR_MAXL];
    int seisout;       /* 1 velocities, 2 pressure, 3 stresses, 4 all          */
    int param_type;    /* 0: vp,vs,rho  1: M,mu,rho  2: Ip,Is,rho  3: complex v */
    int nthreads;      /* OpenMP threads (0 -> runtime default)                  */
    /* misfit options (src/residuals.c:24-450)                                    */
    int restype;       /* 1: residuals = data (rtm_res)                          */
    int scalerms, scaleshot, scalermsnorm;
    float fmin, fmax;  /* Butterworth band-pass of order 6, 0: off              */
} or_csts;

/* ---- setup pieces, exported individually so tests can pin them ---------- */
int  or_holberg(int fdorder, int maxrelerror, double hc[7]);
void or_kernel_hc(int fdorder, int maxrelerror, float hcf[7]);
float or_literal9(double v);
int  or_cpml_coeff(float npower, float kmax, float fpml, float vppml, float dh,
                   float dt, int nab, float *K, float *b, float *a,
                   float *K_half, float *b_half, float *a_half);
void or_taper(int nab, float abpc, float *taper);
void or_eta(int L, const float *FL, float dt, float *eta);
/* Parameter transforms of src/assign_modeling_case.c applied in the order of
 * registration (src/assign_modeling_case.c:1001-1040).  Outputs are the 14
 * kernel parameters in reference (unpadded) layout; any pointer may be NULL
 * if the case does not use that parameter. */
int  or_transform(const or_csts *c, const float *p0, const float *p1,
                  const float *rho, const float *taup, const float *taus,
                  float *M, float *mu, float *rhoo, float *rip, float *rjp,
                  float *rkp, float *muipkp, float *muipjp, float *mujpkp,
                  float *taupo, float *tauso, float *tausipkp, float *tausipjp,
                  float *tausjpkp);
int  or_check_stability(const or_csts *c, const float *M, const float *mu,
                        const float *rho, float *dtstable);

/* Residuals of a shot, [nout][nrec][NT] (var_res_raw with its misfit       */
/* options, or rtm_res): rms[0] = sum (in - out)^2, rms[1] = sum in^2.      */
/* gradfix & OR_FIX_RMSNORM: normalised-misfit (scalermsnorm) sources      */
/* weigh the trace by o.d, the derivative of 1/2 |o/|o| - d/|d||^2; the     */
/* reference weighs it by 1 - o.d (src/residuals.c:331-337).                */
void or_residuals(const or_csts *c, int nout, int nrec, const float *out, const float *din,
                  float *res, double *rms, int gradfix);
/* The band-pass alone, in place on ntrace traces data[n*NT + t], t < tmax.  */
void or_butterworth(float *data, float fcl, float fch, float dt, int NT, int tmax,
                    int ntrace, int order);

/* number of output traces per receiver for seisout, and their names */
int  or_nout(const or_csts *c);
const char *or_outname(const or_csts *c, int i);
/* number of wavefield variables (vx..syy/p) and their names */
int  or_nvars(const or_csts *c);
const char *or_varname(const or_csts *c, int i);

/* Forward modeling of one shot (all sources with the same srcid fire
 * together).  out: [nout][nrec][NT].  fields_out (optional): the final
 * wavefields in padded reference layout, [nvars][Pi(N+FDORDER)].
 * Returns 0 on success, non-zero with a message in err otherwise. */
int  or_forward(const or_csts *c, const float *p0, const float *p1,
                const float *rho, const float *taup, const float *taus,
                int nsrc, const float *src_pos, const float *src,
                int nrec, const float *rec_pos, float *out,
                float *fields_out, char *err, int errlen);
/* or_forward started from the padded wavefields fields_in (or_nvars arrays */
/* in or_varname order); memory variables and CPML memory start at zero.    */
int  or_forward_from(const or_csts *c, const float *p0, const float *p1,
                     const float *rho, const float *taup, const float *taus,
                     int nsrc, const float *src_pos, const float *src, int nrec,
                     const float *rec_pos, const float *fields_in, float *out,
                     float *fields_out, char *err, int errlen);

/* wall-clock seconds of the time loop of the last or_forward call */
double or_last_loop_seconds(void);

/* Backpropagation boundary shell: returns the number of inconsistencies
 * between the save enumeration and the inject lookup (0 when they are
 * inverse maps covering the same cells) and the shell size in *nbnd. */
int  or_shell_check(const or_csts *c, long *nbnd);

/* Gradient of one shot (src/time_stepping.c:747-1002, BACK_PROP_TYPE=1),
 * ND 2 or 3.  Elastic runs reconstruct the forward wavefield by
 * backpropagation with boundary saving, as the reference does; viscoelastic
 * runs (L>0) read it from a stored history of post-source states (the
 * reference's viscoelastic backpropagation is unstable) and update the
 * adjoint memory variables with the forward recursion (src/update_adjs2D.cl:
 * 367-408, src/update_adjs3D.cl:549-605, with the memory variables of each
 * l indexed by l and written to the adjoint arrays).  Adjoint sources:
 * res_in [nout][nrec][NT] when given (inputres), else out - din; both are
 * scaled as res_scale does.
 * gradfix selects corrections of reference defects (0 = as the reference):
 *   OR_FIX_RES_SCALE  res_scale reads the buoyancy at the receiver's own
 *                     cell (the reference uses x for all three indices);
 *   OR_FIX_ADJ_CPML   the adjoint run starts from zeroed CPML memory (the
 *                     reference reuses the forward run's);
 *   OR_FIX_RHO_SCALE  the direct density term of the gradient is scaled by
 *                     dh^2/dt^2;
 *   OR_FIX_SURF_SYY   the 3D adjoint free surface reads syy (the reference
 *                     reads sxx twice);
 *   OR_FIX_VISCO_F0   viscoelastic runs: the gradient formulas use the
 *                     moduli at the reference frequency f0 (the reference
 *                     uses the relaxed moduli, which inflates the gradient
 *                     by about (1+s)^(3/2), s the relaxation factor).
 * grad: [3][Pi N] in param_type order (vp|M|Ip, vs|mu|Is, rho), gradients
 * accumulated in double.  res_out (optional): the scaled adjoint sources;
 * rms (optional): [rms, rmsnorm]; fields_out (optional): the backpropagated
 * forward wavefields at the end, padded; hess (optional): the approximate
 * Hessian [3][Pi N] transformed like grad (Hout).  gradtau (optional):
 * [2][Pi N], the gradients of tau_p and tau_s at fixed wave speeds
 * (viscoelastic runs; zero otherwise).  restype 1 accumulates the
 * RTM image of M instead of the gradient.  movie (optional, movout > 0): the
 * movie of the gradient run, [nvars][NT/movout][padded cells]: after the
 * adjoint step t with t % movout == 0, frame (t+1)/movout - 1 holds the
 * backpropagated forward wavefields (src/time_stepping.c:926-929); in
 * viscoelastic runs, which read the forward state from the history, the
 * forward wavefields at the start of step t (what backpropagation would
 * reconstruct there).  Frames no step writes stay as given. */
int  or_gradient(const or_csts *c, const float *p0, const float *p1,
                 const float *rho, const float *taup, const float *taus,
                 int nsrc, const float *src_pos, const float *src,
                 int nrec, const float *rec_pos, const float *din,
                 const float *res_in, int gradfix, float *out,
                 float *res_out, float *grad, double *rms, float *fields_out,
                 float *hess, float *gradtau, int movout, float *movie, char *err,
                 int errlen);

/* DFT-domain gradient of one shot (BACK_PROP_TYPE=2; src/calc_grad.c:
 * 377-968, src/automatic_kernels.c:1030-1128, src/assign_modeling_case.c:
 * 549-570, src/time_stepping.c:658-672, 767-780, 869-965), ND 2 or 3: the
 * forward run accumulates the DFT of every wavefield and memory variable at
 * the nf frequencies freqs (Hz) every DTNYQ steps from tmin, before the
 * sources; the adjoint run propagates the adjoint wavefield alone and
 * accumulates its DFT after the updates; the gradient correlates the two
 * spectra per cell and frequency with the coefficients of param_type.
 * grad: [5][Pi N] (param order, then taup, taus), double, not transformed
 * further (transf_grad applies to BACK_PROP_TYPE=1 only).  rms (optional):
 * [rms, rmsnorm] at the gradient frequencies.  dft_out (optional): the
 * forward spectra [nq][nf][padded cells][2], nq = 5 + 3L (2D: vx vz sxx szz
 * sxz, then rxx rzz rxz by l) or 9 + 6L (3D: vx vy vz sxx syy szz sxy syz
 * sxz, then rxx ryy rzz rxy ryz rxz by l). */
int  or_dft_sampling(const or_csts *c, int nf, const float *freqs, int *dtnyq, int *ntnyq,
                     float *df, float *bins);
void or_dft_coef(double *c, int pt, int visco, float M, float mu, float tp, float ts, float rho,
                 float ND, float L, float al);
int  or_dft_rms(const or_csts *c, int nout, int nrec, const float *out, const float *din,
                int nf, const float *freqs, double *rms);
int  or_gradient_dft(const or_csts *c, const float *p0, const float *p1, const float *rho,
                     const float *taup, const float *taus, int nsrc, const float *src_pos,
                     const float *src, int nrec, const float *rec_pos, const float *din,
                     const float *res_in, int gradfix, int nf, const float *freqs, float *out,
                     double *grad, double *rms, float *dft_out, char *err, int errlen);

#ifdef __cplusplus
}
#endif
#endif
