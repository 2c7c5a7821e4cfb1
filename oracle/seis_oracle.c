/*
 * seis_oracle.c -- CPU restatement of SeisCL's time-domain finite-difference
 * forward path (the checker for the HIP product path).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in seiscl_amd/ links or calls it.
 *
 * Arithmetic semantics follow the reference's OpenCL build:
 *   - HC1..HC6, DT, DT2 and DH are float literals printed with "%9.9f"
 *     (src/clprogram.c:323-348), reproduced by or_literal9();
 *   - unsuffixed literals such as 1.0, 2.0, 0.5 are double, so expressions
 *     that contain them are evaluated in double and rounded to float on
 *     assignment, exactly as written below;
 *   - contraction into FMA is disabled (-ffp-contract=off in oracle/Makefile)
 *     and every sum keeps the reference's left-to-right order.
 *
 * Every function cites the reference file:line it restates.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "seis_oracle.h"

#define PI (3.141592653589793238462643383279502884197169)

/* ------------------------------------------------------------------------ */
/* FD coefficients: src/third_party/SOFI3D/holbergcoeff.c:22-102            */
/* ------------------------------------------------------------------------ */
static const float HOLBERG[5][6][7] = {
    {{23.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f},
     {8.0f, 9.0f / 8.0f, -1.0f / 24.0f, 0.0f, 0.0f, 0.0f, 0.0f},
     {6.0f, 75.0f / 64.0f, -25.0f / 384.0f, 3.0f / 640.0f, 0.0f, 0.0f, 0.0f},
     {5.0f, 1225.0f / 1024.0f, -245.0f / 3072.0f, 49.0f / 5120.0f,
      -5.0f / 7168.0f, 0.0f, 0.0f},
     {5.0f, 19845.0f / 16384.0f, -735.0f / 8192.0f, 567.0f / 40960.0f,
      -405.0f / 229376.0f, 35.0f / 294912.0f, 0.0f},
     {4.0f, 160083.0f / 131072.0f, -12705.0f / 131072.0f,
      22869.0f / 1310720.0f, -5445.0f / 1835008.0f, 847.0f / 2359296.0f,
      -63.0f / 2883584.0f}},
    {{49.7f, 1.0010f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f},
     {8.32f, 1.1382f, -0.046414f, 0.0f, 0.0f, 0.0f, 0.0f},
     {4.77f, 1.1965f, -0.078804f, 0.0081781f, 0.0f, 0.0f, 0.0f},
     {3.69f, 1.2257f, -0.099537f, 0.018063f, -0.0026274f, 0.0f, 0.0f},
     {3.19f, 1.2415f, -0.11231f, 0.026191f, -0.0064682f, 0.001191f, 0.0f},
     {2.91f, 1.2508f, -0.12034f, 0.032131f, -0.010142f, 0.0029857f,
      -0.00066667f}},
    {{22.2f, 1.0050f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f},
     {5.65f, 1.1534f, -0.052806f, 0.0f, 0.0f, 0.0f, 0.0f},
     {3.74f, 1.2111f, -0.088313f, 0.011768f, 0.0f, 0.0f, 0.0f},
     {3.11f, 1.2367f, -0.10815f, 0.023113f, -0.0046905f, 0.0f, 0.0f},
     {2.80f, 1.2496f, -0.11921f, 0.031130f, -0.0093272f, 0.0025161f, 0.0f},
     {2.62f, 1.2568f, -0.12573f, 0.036423f, -0.013132f, 0.0047484f,
      -0.0015979f}},
    {{15.8f, 1.0100f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f},
     {4.80f, 1.1640f, -0.057991f, 0.0f, 0.0f, 0.0f, 0.0f},
     {3.39f, 1.2192f, -0.094070f, 0.014608f, 0.0f, 0.0f, 0.0f},
     {2.90f, 1.2422f, -0.11269f, 0.026140f, -0.0064054f, 0.0f, 0.0f},
     {2.65f, 1.2534f, -0.12257f, 0.033755f, -0.011081f, 0.0036784f, 0.0f},
     {2.51f, 1.2596f, -0.12825f, 0.038550f, -0.014763f, 0.0058619f,
      -0.0024538f}},
    {{9.16f, 1.0300f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f},
     {3.47f, 1.1876f, -0.072518f, 0.0f, 0.0f, 0.0f, 0.0f},
     {2.91f, 1.2341f, -0.10569f, 0.022589f, 0.0f, 0.0f, 0.0f},
     {2.61f, 1.2516f, -0.12085f, 0.032236f, -0.011459f, 0.0f, 0.0f},
     {2.45f, 1.2596f, -0.12829f, 0.038533f, -0.014681f, 0.0072580f, 0.0f},
     {2.36f, 1.2640f, -0.13239f, 0.042217f, -0.017803f, 0.0081959f,
      -0.0051848f}},
};

int or_holberg(int fdorder, int maxrelerror, double hc[7]) {
    if (fdorder != 2 && fdorder != 4 && fdorder != 6 && fdorder != 8 &&
        fdorder != 10 && fdorder != 12)
        return 1;
    if (maxrelerror < 0 || maxrelerror > 4) return 1;
    for (int i = 0; i <= 6; i++)
        hc[i] = (double)HOLBERG[maxrelerror][fdorder / 2 - 1][i];
    return 0;
}

/* A double printed as "%9.9f" and read back as a float literal: the way the
 * reference hands HC, DT, DT2 and DH to its kernels (src/clprogram.c:323-348). */
float or_literal9(double v) {
    char buf[64];
    snprintf(buf, sizeof(buf), "%9.9f", v);
    return strtof(buf, NULL);
}

void or_kernel_hc(int fdorder, int maxrelerror, float hcf[7]) {
    double hc[7] = {0};
    or_holberg(fdorder, maxrelerror, hc);
    hcf[0] = 0.0f;
    for (int i = 1; i <= 6; i++) hcf[i] = or_literal9(hc[i]);
}

/* ------------------------------------------------------------------------ */
/* CPML profiles: src/third_party/SOFI3D/CPML_COEFF.c:26-200                 */
/* Arrays hold 2*nab floats and must be zero on entry (the reference        */
/* allocates them zeroed, src/F.h:59-62).                                    */
/* ------------------------------------------------------------------------ */
int or_cpml_coeff(float NPOWER, float k_max_CPML, float FPML, float VPPML,
                  float dh, float dt, int NAB, float *K_i, float *b_i,
                  float *a_i, float *K_i_half, float *b_i_half,
                  float *a_i_half) {
    int state = 0;
    const float npower = NPOWER;
    const float alpha_max_PML = 2.0 * PI * (FPML / 2.0);
    const float Rcoef = 0.0008;
    const float a = 0.25, b = 0.75, c = 0.0;
    float d0_i, position_norm, position_in_PML;
    float *d_i = calloc(2 * NAB + 1, sizeof(float));
    float *d_i_half = calloc(2 * NAB + 1, sizeof(float));
    float *alpha_prime_i = calloc(2 * NAB + 1, sizeof(float));
    float *alpha_prime_i_half = calloc(2 * NAB + 1, sizeof(float));

    d0_i = -(npower + 1) * VPPML * log(Rcoef) / (2.0 * NAB * dh);
    for (int i = 0; i <= NAB; i++) {
        K_i[i] = 1.0;
        position_in_PML = (NAB - i) * dh;
        position_norm = position_in_PML / (NAB * dh);
        d_i[i] = d0_i * (a * position_norm + b * pow(position_norm, npower) +
                         c * pow(position_norm, (4)));
        K_i[i] = 1.0 + (k_max_CPML - 1.0) * pow(position_norm, npower);
        alpha_prime_i[i] = alpha_max_PML * (1.0 - position_norm);
        if (alpha_prime_i[i] < 0.0) state = 1;
        b_i[i] = exp(-(d_i[i] / K_i[i] + alpha_prime_i[i]) * dt);
        if (fabsf(d_i[i]) > 1.0e-6)
            a_i[i] = d_i[i] * (b_i[i] - 1.0) /
                     (K_i[i] * (d_i[i] + K_i[i] * alpha_prime_i[i]));
        else
            a_i[i] = 0.0;
        if (i <= NAB - 1) {
            position_in_PML = (NAB - i + 0.5 - 1.0) * dh;
            position_norm = position_in_PML / (NAB * dh);
            int i1 = i;
            K_i_half[i1] = 1.0;
            d_i_half[i1] =
                d0_i * (a * position_norm + b * pow(position_norm, npower) +
                        c * pow(position_norm, (4)));
            if (position_in_PML < 0.0) state = 1;
            K_i_half[i1] = 1.0 + (k_max_CPML - 1.0) * pow(position_norm, npower);
            alpha_prime_i_half[i1] = alpha_max_PML * (1.0 - position_norm);
            if (alpha_prime_i_half[i1] < 0.0) state = 1;
            b_i_half[i1] =
                exp(-(d_i_half[i1] / K_i_half[i1] + alpha_prime_i_half[i1]) * dt);
            if (fabs(d_i_half[i1]) > 1.0e-6)
                a_i_half[i1] = d_i_half[i1] * (b_i_half[i1] - 1.0) /
                               (K_i_half[i1] * (d_i_half[i1] +
                                                K_i_half[i1] * alpha_prime_i_half[i1]));
            /* mirror onto the far boundary */
            int l = 2 * NAB - i - 1;
            if (i > 0) {
                K_i[l + 1] = K_i[i];
                b_i[l + 1] = b_i[i];
                if (fabs(d_i[i]) > 1.0e-6) a_i[l + 1] = a_i[i];
            }
            K_i_half[l] = K_i_half[i];
            b_i_half[l] = b_i_half[i];
            if (fabs(d_i[i]) > 1.0e-6) a_i_half[l] = a_i_half[i];
        }
    }
    free(d_i);
    free(d_i_half);
    free(alpha_prime_i);
    free(alpha_prime_i_half);
    return state;
}

/* Cerjan taper: src/assign_modeling_case.c:571-584 */
void or_taper(int nab, float abpc, float *taper) {
    float amp = 1 - abpc / 100;
    float a = sqrt(-log(amp) / ((nab - 1) * (nab - 1)));
    for (int i = 1; i <= nab; i++)
        taper[i - 1] = exp(-(a * a * (nab - i) * (nab - i)));
}

/* Relaxation constants: src/assign_modeling_case.c:536-548 */
void or_eta(int L, const float *FL, float dt, float *eta) {
    for (int l = 0; l < L; l++) eta[l] = (2.0 * PI * FL[l]) * dt;
}

/* ------------------------------------------------------------------------ */
/* Parameter averaging: src/assign_modeling_case.c:50-301                    */
/* ------------------------------------------------------------------------ */
static void dims3(const or_csts *c, int *NZ, int *NY, int *NX) {
    if (c->ND == 3) {
        *NZ = c->N[0];
        *NY = c->N[1];
        *NX = c->N[2];
    } else {
        *NZ = c->N[0];
        *NY = 1;
        *NX = c->N[1];
    }
}

/* ave_arithmetic1 (:50-99): pout = 0.5*(p + p[+dir]); last plane copied */
static void ave_arith1(const float *pin, float *pout, int NZ, int NY, int NX,
                       const int dir[3]) {
    int NX0 = 0, NY0 = 0, NZ0 = 0;
    for (int k = 0; k < NZ - dir[0]; k++)
        for (int j = 0; j < NY - dir[1]; j++)
            for (int i = 0; i < NX - dir[2]; i++) {
                long i1 = (long)i * NY * NZ + (long)j * NZ + k;
                long i2 = (long)(i + dir[2]) * NY * NZ + (long)(j + dir[1]) * NZ +
                          (k + dir[0]);
                pout[i1] = 0.5 * (pin[i1] + pin[i2]);
            }
    if (dir[2] == 1) NX0 = NX - 1;
    else if (dir[1] == 1) NY0 = NY - 1;
    if (dir[0] == 1) NZ0 = NZ - 1;
    for (int k = NZ0; k < NZ; k++)
        for (int j = NY0; j < NY; j++)
            for (int i = NX0; i < NX; i++) {
                long i1 = (long)i * NY * NZ + (long)j * NZ + k;
                pout[i1] = pin[i1];
            }
}

/* Edge handling shared by ave_arithmetic2 (:100-173) and ave_harmonic
 * (:225-301).  The start indices of the second copy are not reset after
 * the first one, so only the corner line of the second edge is copied and
 * the rest of that edge keeps the zero the buffer was allocated with. */
static void ave_edges2(const float *pin, float *pout, int NZ, int NY, int NX,
                       const int dir[2][3]) {
    int NX0 = 0, NY0 = 0, NZ0 = 0;
    for (int e = 0; e < 2; e++) {
        if (dir[e][2] == 1) NX0 = NX - 1;
        else if (dir[e][1] == 1) NY0 = NY - 1;
        if (dir[e][0] == 1) NZ0 = NZ - 1;
        for (int k = NZ0; k < NZ; k++)
            for (int j = NY0; j < NY; j++)
                for (int i = NX0; i < NX; i++) {
                    long i1 = (long)i * NY * NZ + (long)j * NZ + k;
                    pout[i1] = pin[i1];
                }
    }
}

static void ave_arith2(const float *pin, float *pout, int NZ, int NY, int NX,
                       const int dir[2][3]) {
    for (int k = 0; k < NZ - dir[0][0] - dir[1][0]; k++)
        for (int j = 0; j < NY - dir[0][1] - dir[1][1]; j++)
            for (int i = 0; i < NX - dir[0][2] - dir[1][2]; i++) {
                long i1 = (long)i * NY * NZ + (long)j * NZ + k;
                long i2 = (long)(i + dir[0][2]) * NY * NZ +
                          (long)(j + dir[0][1]) * NZ + (k + dir[0][0]);
                long i3 = (long)(i + dir[1][2]) * NY * NZ +
                          (long)(j + dir[1][1]) * NZ + (k + dir[1][0]);
                long i4 = (long)(i + dir[0][2] + dir[1][2]) * NY * NZ +
                          (long)(j + dir[0][1] + dir[1][1]) * NZ +
                          (k + dir[0][0] + dir[1][0]);
                pout[i1] = 0.25 * (pin[i1] + pin[i2] + pin[i3] + pin[i4]);
            }
    ave_edges2(pin, pout, NZ, NY, NX, dir);
}

static void ave_harm(const float *pin, float *pout, int NZ, int NY, int NX,
                     const int dir[2][3]) {
    for (int k = 0; k < NZ - dir[0][0] - dir[1][0]; k++)
        for (int j = 0; j < NY - dir[0][1] - dir[1][1]; j++)
            for (int i = 0; i < NX - dir[0][2] - dir[1][2]; i++) {
                long i1 = (long)i * NY * NZ + (long)j * NZ + k;
                long i2 = (long)(i + dir[0][2]) * NY * NZ +
                          (long)(j + dir[0][1]) * NZ + (k + dir[0][0]);
                long i3 = (long)(i + dir[1][2]) * NY * NZ +
                          (long)(j + dir[1][1]) * NZ + (k + dir[1][0]);
                long i4 = (long)(i + dir[0][2] + dir[1][2]) * NY * NZ +
                          (long)(j + dir[0][1] + dir[1][1]) * NZ +
                          (k + dir[0][0] + dir[1][0]);
                pout[i1] = 4.0 / (1.0 / pin[i1] + 1.0 / pin[i2] +
                                  1.0 / pin[i3] + 1.0 / pin[i4]);
            }
    ave_edges2(pin, pout, NZ, NY, NX, dir);
}

/* powf(x,2) of the reference is emitted by gcc as the correctly rounded x*x */
static inline float sq(float x) { return x * x; }

/* ------------------------------------------------------------------------ */
/* Parameter transforms, in registration order                               */
/* (src/assign_modeling_case.c:1001-1040; transforms :305-535)               */
/* ------------------------------------------------------------------------ */
int or_transform(const or_csts *c, const float *p0, const float *p1,
                 const float *rho_in, const float *taup_in,
                 const float *taus_in, float *M, float *mu, float *rho,
                 float *rip, float *rjp, float *rkp, float *muipkp,
                 float *muipjp, float *mujpkp, float *taup, float *taus,
                 float *tausipkp, float *tausipjp, float *tausjpkp) {
    int NZ, NY, NX;
    dims3(c, &NZ, &NY, &NX);
    long n = (long)NZ * NY * NX;
    int L = c->L;
    int elastic = (c->ND == 2 || c->ND == 3);
    float eta[OR_MAXL];
    or_eta(L, c->FL, c->dt, eta);
    float *tp = taup, *ts = taus;
    float *tmp_tp = NULL, *tmp_ts = NULL;
    if (L > 0) {
        if (!tp) tp = tmp_tp = malloc(n * sizeof(float));
        if (elastic && !ts) ts = tmp_ts = malloc(n * sizeof(float));
        memcpy(tp, taup_in, n * sizeof(float));
        if (elastic) memcpy(ts, taus_in, n * sizeof(float));
    }
    float pts[OR_MAXL];
    for (int l = 0; l < L; l++) pts[l] = c->dt / eta[l];
    float ws = 2.0 * PI * c->f0;

    /* M (:380-449) */
    for (long i = 0; i < n; i++) {
        float v = p0[i];
        if (c->param_type == 0) M[i] = sq(v) * rho_in[i] * c->dt / c->dh;
        if (c->param_type == 1) M[i] = v * c->dt / c->dh;
        else if (c->param_type == 2)
            M[i] = sq(v / rho_in[i]) * rho_in[i] * c->dt / c->dh;
        else if (c->param_type == 3) {
            float thisvp = v - tp[i];
            float thistaup = tp[i] / (v - tp[i]);
            M[i] = sq(thisvp) * rho_in[i] * c->dt / c->dh;
            tp[i] = thistaup;
        }
    }
    if (L > 0) {
        for (long i = 0; i < n; i++) {
            float sumpi = 0.0;
            for (int l = 0; l < L; l++)
                sumpi += ((ws * ws * pts[l] * pts[l] * tp[i]) /
                          (1.0 + ws * ws * pts[l] * pts[l]));
            M[i] = M[i] / (1.0 + sumpi);
        }
    }
    /* mu (:305-379) */
    if (elastic) {
        for (long i = 0; i < n; i++) {
            float v = p1[i];
            if (c->param_type == 0) mu[i] = sq(v) * rho_in[i] * c->dt / c->dh;
            else if (c->param_type == 1) mu[i] = v * c->dt / c->dh;
            else if (c->param_type == 2)
                mu[i] = sq(v / rho_in[i]) * rho_in[i] * c->dt / c->dh;
            else if (c->param_type == 3) {
                float thisvs, thistaus;
                if (v > 0) {
                    thisvs = v - ts[i];
                    thistaus = ts[i] / (v - ts[i]);
                } else {
                    thistaus = 0;
                    thisvs = 0;
                }
                mu[i] = sq(thisvs) * rho_in[i] * c->dt / c->dh;
                ts[i] = thistaus;
            }
        }
        if (L > 0) {
            float sumu = 0.0;
            for (int l = 0; l < L; l++)
                sumu += ((ws * ws * pts[l] * pts[l]) /
                         (1.0 + ws * ws * pts[l] * pts[l]));
            for (long i = 0; i < n; i++) mu[i] = mu[i] / (1.0 + sumu * ts[i]);
        }
    }
    /* rho (:451-462), par_scale = 0 in FP32 */
    for (long i = 0; i < n; i++)
        rho[i] = 1.0 / rho_in[i] * c->dt / c->dh * 1.0; /* powf(2,-par_scale), par_scale=0 */
    /* buoyancy averages (:464-487) */
    {
        int dx[3] = {0, 0, 1}, dy[3] = {0, 1, 0}, dz[3] = {1, 0, 0};
        ave_arith1(rho, rip, NZ, NY, NX, dx);
        if (c->ND == 3) ave_arith1(rho, rjp, NZ, NY, NX, dy);
        ave_arith1(rho, rkp, NZ, NY, NX, dz);
    }
    if (elastic) {
        int dxz[2][3] = {{0, 0, 1}, {1, 0, 0}};
        int dyz[2][3] = {{0, 1, 0}, {1, 0, 0}};
        int dxy[2][3] = {{0, 0, 1}, {0, 1, 0}};
        memset(muipkp, 0, n * sizeof(float));
        ave_harm(mu, muipkp, NZ, NY, NX, dxz);
        if (c->ND == 3) {
            memset(muipjp, 0, n * sizeof(float));
            memset(mujpkp, 0, n * sizeof(float));
            ave_harm(mu, muipjp, NZ, NY, NX, dxy);
            ave_harm(mu, mujpkp, NZ, NY, NX, dyz);
        }
        if (L > 0) {
            memset(tausipkp, 0, n * sizeof(float));
            ave_arith2(ts, tausipkp, NZ, NY, NX, dxz);
            if (c->ND == 3) {
                memset(tausipjp, 0, n * sizeof(float));
                memset(tausjpkp, 0, n * sizeof(float));
                ave_arith2(ts, tausipjp, NZ, NY, NX, dxy);
                ave_arith2(ts, tausjpkp, NZ, NY, NX, dyz);
            }
        }
    }
    free(tmp_tp);
    free(tmp_ts);
    return 0;
}

/* src/assign_modeling_case.c:730-827 */
int or_check_stability(const or_csts *c, const float *M, const float *mu,
                       const float *rho, float *dtstable_out) {
    int NZ, NY, NX;
    dims3(c, &NZ, &NY, &NX);
    long n = (long)NZ * NY * NX;
    float vpmax = 0, vsmax = 0, vpmin = 99999, vsmin = 99999;
    float vmax, g, gamma, dtstable;
    (void)vpmin;
    (void)vsmin;
    if (mu) {
        for (long i = 0; i < n; i++) {
            float thisvs = sqrt(mu[i] * rho[i]) / c->dt * c->dh;
            if (vsmax < thisvs) vsmax = thisvs;
            if (vsmin > thisvs && thisvs > 0.1) vsmin = thisvs;
        }
    }
    if (M) {
        for (long i = 0; i < n; i++) {
            float thisvp = sqrt(M[i] * rho[i]) / c->dt * c->dh;
            if (vpmax < thisvp) vpmax = thisvp;
            if (vpmin > thisvp && thisvp > 0.1) vpmin = thisvp;
        }
    }
    vmax = (vpmax == 0) ? vsmax : vpmax;
    switch (c->FDORDER) {
        case 2: gamma = 1.0; break;
        case 4: gamma = 1.184614; break;
        case 6: gamma = 1.283482; break;
        case 8: gamma = 1.345927; break;
        case 10: gamma = 1.38766; break;
        case 12: gamma = 1.417065; break;
        default: gamma = 1.0; break;
    }
    (void)g;
    if (c->ND == 3) dtstable = c->dh / (gamma * sqrt(3.0) * vmax);
    else dtstable = c->dh / (gamma * sqrt(2.0) * vmax);
    if (dtstable_out) *dtstable_out = dtstable;
    return c->dt > dtstable;
}

/* ------------------------------------------------------------------------ */
/* Output variables (src/assign_modeling_case.c:1244-1326, SeisCL.py        */
/* to_load_names :376-408).                                                  */
/* ------------------------------------------------------------------------ */
static const char *OUT2[4][6] = {{"vx", "vz"},
                                 {"p"},
                                 {"sxx", "szz", "sxz"},
                                 {"vx", "vz", "sxx", "szz", "sxz", "p"}};
static const char *OUT3[4][10] = {
    {"vx", "vy", "vz"},
    {"p"},
    {"sxx", "syy", "szz", "sxz", "sxy", "syz"},
    {"vx", "vy", "vz", "sxx", "syy", "szz", "sxz", "sxy", "syz", "p"}};
static const char *OUTAC[4][3] = {{"vx", "vz"}, {"p"}, {NULL}, {"vx", "vz", "p"}};
static const int NOUT2[4] = {2, 1, 3, 6}, NOUT3[4] = {3, 1, 6, 10},
                 NOUTAC[4] = {2, 1, 0, 3};

int or_nout(const or_csts *c) {
    if (c->seisout < 1 || c->seisout > 4) return 0;
    if (c->ND == 3) return NOUT3[c->seisout - 1];
    if (c->ND == 22) return NOUTAC[c->seisout - 1];
    return NOUT2[c->seisout - 1];
}
const char *or_outname(const or_csts *c, int i) {
    if (i < 0 || i >= or_nout(c)) return NULL;
    if (c->ND == 3) return OUT3[c->seisout - 1][i];
    if (c->ND == 22) return OUTAC[c->seisout - 1][i];
    return OUT2[c->seisout - 1][i];
}

/* Wavefield variables in registration order (src/assign_modeling_case.c:1044-1070) */
static const char *VARS2[] = {"vx", "vz", "sxx", "szz", "sxz"};
static const char *VARS3[] = {"vx", "vy", "vz", "sxx", "szz", "sxz", "sxy", "syz", "syy"};
static const char *VARSAC[] = {"vx", "vz", "p"};
int or_nvars(const or_csts *c) { return c->ND == 3 ? 9 : (c->ND == 22 ? 3 : 5); }
const char *or_varname(const or_csts *c, int i) {
    if (i < 0 || i >= or_nvars(c)) return NULL;
    return c->ND == 3 ? VARS3[i] : (c->ND == 22 ? VARSAC[i] : VARS2[i]);
}

/* ------------------------------------------------------------------------ */
/* Simulation state                                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    const or_csts *c;
    int FDOH, NAB, L, ND;
    int NZ, NY, NX;        /* unpadded */
    int NZp, NYp, NXp;     /* padded (NY=NYp=1 in 2D) */
    long nv, np;           /* padded / unpadded cell counts */
    float hc[7], DT, DT2, DH;
    float eta[OR_MAXL];
    /* wavefields */
    float *vx, *vy, *vz, *sxx, *syy, *szz, *sxy, *syz, *sxz, *p;
    float *rxx, *ryy, *rzz, *rxy, *ryz, *rxz, *rp;
    /* parameters */
    float *M, *mu, *rho, *rip, *rjp, *rkp, *muipkp, *muipjp, *mujpkp;
    float *taup, *taus, *tausipkp, *tausipjp, *tausjpkp;
    /* CPML */
    float K_x[256], a_x[256], b_x[256], K_x_half[256], a_x_half[256],
        b_x_half[256];
    float K_y[256], a_y[256], b_y[256], K_y_half[256], a_y_half[256],
        b_y_half[256];
    float K_z[256], a_z[256], b_z[256], K_z_half[256], a_z_half[256],
        b_z_half[256];
    float taper[256];
    float *psi_sxx_x, *psi_sxy_x, *psi_sxz_x, *psi_sxy_y, *psi_syy_y,
        *psi_syz_y, *psi_sxz_z, *psi_syz_z, *psi_szz_z;
    float *psi_vx_x, *psi_vy_x, *psi_vz_x, *psi_vx_y, *psi_vy_y, *psi_vz_y,
        *psi_vx_z, *psi_vy_z, *psi_vz_z;
    float *psi_p_x, *psi_p_z;
    /* adjoint runs: the increment each update adds before the taper, per
     * cell (2D v: vx vz; 2D s: sxx szz sxz; 3D v: vx vy vz; 3D s: sxx syy
     * szz sxy syz sxz); NULL in forward runs */
    float *inc[6];
    /* viscoelastic adjoint runs with tau gradients: per cell, the change of */
    /* each partial stress increment with tau_p (one value, the same for the */
    /* normal stresses) and with tau_s (per component), both at fixed wave   */
    /* speeds (2D: p xx zz xz; 3D: p xx yy zz xy yz xz); NULL otherwise      */
    float *tinc[7];
    float s0;    /* sum_l w0^2 t_l^2 / (1 + w0^2 t_l^2) of the transform */
} sim;

static float *zalloc(long n) { return n > 0 ? calloc(n, sizeof(float)) : NULL; }

/* stencils of src/header_FD.cl:87-265, summed left to right */
static inline float Dp(const float *v, long i, long s, const float *hc, int fdoh) {
    float acc = hc[1] * (v[i + s] - v[i]);
    for (int k = 2; k <= fdoh; k++)
        acc = acc + hc[k] * (v[i + k * s] - v[i - (k - 1) * s]);
    return acc;
}
static inline float Dm(const float *v, long i, long s, const float *hc, int fdoh) {
    float acc = hc[1] * (v[i] - v[i - s]);
    for (int k = 2; k <= fdoh; k++)
        acc = acc + hc[k] * (v[i + (k - 1) * s] - v[i - k * s]);
    return acc;
}

/* ------------------------------------------------------------------------ */
/* 2D P-SV velocity update: src/update_v2D.cl:32-246                         */
/* ------------------------------------------------------------------------ */
static void update_v2D(sim *S) {
    const int FDOH = S->FDOH, NAB = S->NAB, NZ = S->NZp, NX = S->NXp;
    const int cpml = S->c->abs_type == 1, tap = S->c->abs_type == 2,
              fs = S->c->freesurf;
    const float *hc = S->hc;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++) {
        for (int gidz = FDOH; gidz < NZ - FDOH; gidz++) {
            long indp = (long)(gidx - FDOH) * (NZ - 2 * FDOH) + (gidz - FDOH);
            long indv = (long)gidx * NZ + gidz;
            float sxx_x = Dp(S->sxx, indv, NZ, hc, FDOH);
            float szz_z = Dp(S->szz, indv, 1, hc, FDOH);
            float sxz_z = Dm(S->sxz, indv, 1, hc, FDOH);
            float sxz_x = Dm(S->sxz, indv, NZ, hc, FDOH);
            if (cpml) {
                int i, k, ind;
                if (gidz > NZ - NAB - FDOH - 1) {
                    i = gidx - FDOH;
                    k = gidz - NZ + NAB + FDOH + NAB;
                    ind = 2 * NAB - 1 - k;
                    float *ps = &S->psi_sxz_z[(long)i * (2 * NAB) + k];
                    *ps = S->b_z[ind + 1] * *ps + S->a_z[ind + 1] * sxz_z;
                    sxz_z = sxz_z / S->K_z[ind + 1] + *ps;
                    float *pz = &S->psi_szz_z[(long)i * (2 * NAB) + k];
                    *pz = S->b_z_half[ind] * *pz + S->a_z_half[ind] * szz_z;
                    szz_z = szz_z / S->K_z_half[ind] + *pz;
                } else if (!fs && gidz - FDOH < NAB) {
                    i = gidx - FDOH;
                    k = gidz - FDOH;
                    float *ps = &S->psi_sxz_z[(long)i * (2 * NAB) + k];
                    *ps = S->b_z[k] * *ps + S->a_z[k] * sxz_z;
                    sxz_z = sxz_z / S->K_z[k] + *ps;
                    float *pz = &S->psi_szz_z[(long)i * (2 * NAB) + k];
                    *pz = S->b_z_half[k] * *pz + S->a_z_half[k] * szz_z;
                    szz_z = szz_z / S->K_z_half[k] + *pz;
                }
                if (gidx - FDOH < NAB) {
                    i = gidx - FDOH;
                    k = gidz - FDOH;
                    float *px = &S->psi_sxx_x[(long)i * (NZ - 2 * FDOH) + k];
                    *px = S->b_x_half[i] * *px + S->a_x_half[i] * sxx_x;
                    sxx_x = sxx_x / S->K_x_half[i] + *px;
                    float *pxz = &S->psi_sxz_x[(long)i * (NZ - 2 * FDOH) + k];
                    *pxz = S->b_x[i] * *pxz + S->a_x[i] * sxz_x;
                    sxz_x = sxz_x / S->K_x[i] + *pxz;
                }
                if (gidx > NX - NAB - FDOH - 1) {
                    i = gidx - NX + NAB + FDOH + NAB;
                    k = gidz - FDOH;
                    ind = 2 * NAB - 1 - i;
                    float *px = &S->psi_sxx_x[(long)i * (NZ - 2 * FDOH) + k];
                    *px = S->b_x_half[ind] * *px + S->a_x_half[ind] * sxx_x;
                    sxx_x = sxx_x / S->K_x_half[ind] + *px;
                    float *pxz = &S->psi_sxz_x[(long)i * (NZ - 2 * FDOH) + k];
                    *pxz = S->b_x[ind + 1] * *pxz + S->a_x[ind + 1] * sxz_x;
                    sxz_x = sxz_x / S->K_x[ind + 1] + *pxz;
                }
            }
            float lvx = ((sxx_x + sxz_z) * S->rip[indp]);
            float lvz = ((szz_z + sxz_x) * S->rkp[indp]);
            S->vx[indv] += lvx;
            S->vz[indv] += lvz;
            if (S->inc[0]) {
                S->inc[0][indv] = lvx;
                S->inc[1][indv] = lvz;
            }
            if (tap) {
                if (!fs && gidz - FDOH < NAB) {
                    S->vx[indv] *= S->taper[gidz - FDOH];
                    S->vz[indv] *= S->taper[gidz - FDOH];
                }
                if (gidz > NZ - NAB - FDOH - 1) {
                    S->vx[indv] *= S->taper[NZ - FDOH - gidz - 1];
                    S->vz[indv] *= S->taper[NZ - FDOH - gidz - 1];
                }
                if (gidx - FDOH < NAB) {
                    S->vx[indv] *= S->taper[gidx - FDOH];
                    S->vz[indv] *= S->taper[gidx - FDOH];
                }
                if (gidx > NX - NAB - FDOH - 1) {
                    S->vx[indv] *= S->taper[NX - FDOH - gidx - 1];
                    S->vz[indv] *= S->taper[NX - FDOH - gidx - 1];
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* tau gradients of the GSLS stress update (no counterpart in the reference's */
/* time-domain kernels, which declare gradtaup/gradtaus and never write      */
/* them, src/update_adjs3D.cl:66; its DFT path forms them in                 */
/* src/calc_grad.c:111-152).  The update of a normal stress is               */
/*   inc_ii = g th - f (th - v_ii) + Rm_ii,   g = M(1+L tp), f = 2mu(1+L ts), */
/* with Rm_ii = DT2 sum_l (r_ii,l old + new), the whole memory term of the  */
/* step (the partial increment the M and mu gradients read holds only the  */
/* old level; for tau that halves the memory term, and finite differences  */
/* of the misfit then miss by 60-80%), the memory variables forced by      */
/* e th - d (th - v_ii), e = M tp/DT, d = 2 mu ts/DT, all through the same  */
/* relaxation filters: Rm_ii = e F - d F_ii with F the filtered th, and      */
/* sum_ii Rm_ii = (ND e - (ND-1) d) F.  So F and F_ii follow from the       */
/* memory sums, and the change of inc with tp (at fixed M) is               */
/*   M L th + (M/DT) F,                                                     */
/* with ts (at fixed mu) -2mu L (th - v_ii) - (e F - Rm_ii)/ts, and for a    */
/* shear component (coefficients mu_ij(1+L ts_ij), mu_ij ts_ij/DT)           */
/*   mu_ij L w_ij + Rm_ij/ts_ij.                                            */
/* The wave speeds are held fixed through the relaxation normalisation      */
/* M = M_f0/(1 + s0 tp) (src/assign_modeling_case.c:415-437): that adds     */
/* -s0/(1+s0 tp) times the increment's part linear in M (g th + e F), and   */
/* likewise for mu.  M, mu, tp, ts are recovered from g, f, d, e as the      */
/* device kernels do.  Cells with ts = 0 drop the memory term of ts.        */
static void tau_parts_normal(float theta, const float *vii, const float *Rm, int nd, float g,
                             float f, float d, float e, float L, float DT, float s0, float *Xp,
                             float *Xs) {
    float MR = g - L * e * DT;
    float tp = e * DT / MR;
    float tmu = f - L * d * DT;
    float ts = tmu != 0.0f ? d * DT / tmu : 0.0f;
    float kp = s0 / (1.0f + s0 * tp);
    float ks = s0 / (1.0f + s0 * ts);
    float den = nd == 3 ? 3.0f * e - 2.0f * d : 2.0f * e - d;
    float trm = nd == 3 ? (Rm[0] + Rm[1]) + Rm[2] : Rm[0] + Rm[1];
    float F = den != 0.0f ? trm / den : 0.0f;
    *Xp = (MR * L - kp * g) * theta + (MR / DT - kp * e) * F;
    for (int k = 0; k < nd; k++) {
        float dev = theta - vii[k];
        float q = e * F - Rm[k];
        float Sk = -(f * dev) - q;
        Xs[k] = -(tmu * L * dev) - (ts > 0.0f ? q / ts : 0.0f) - ks * Sk;
    }
}

static float tau_part_shear(float w, float Rm, float fij, float dij, float inc, float L,
                            float DT, float s0) {
    float mua = fij - L * dij * DT;
    float tsa = mua != 0.0f ? dij * DT / mua : 0.0f;
    float ksa = s0 / (1.0f + s0 * tsa);
    return mua * L * w + (tsa > 0.0f ? Rm / tsa : 0.0f) - ksa * inc;
}

/* ------------------------------------------------------------------------ */
/* 2D P-SV stress update: src/update_s2D.cl:30-326                           */
/* ------------------------------------------------------------------------ */
static void update_s2D(sim *S) {
    const int FDOH = S->FDOH, NAB = S->NAB, NZ = S->NZp, NX = S->NXp,
              LVE = S->L;
    const int cpml = S->c->abs_type == 1, tap = S->c->abs_type == 2,
              fs = S->c->freesurf;
    const float *hc = S->hc, DT = S->DT, DT2 = S->DT2;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++) {
        for (int gidz = FDOH; gidz < NZ - FDOH; gidz++) {
            long indp = (long)(gidx - FDOH) * (NZ - 2 * FDOH) + (gidz - FDOH);
            long indv = (long)gidx * NZ + gidz;
            float vxx = Dm(S->vx, indv, NZ, hc, FDOH);
            float vxz = Dp(S->vx, indv, 1, hc, FDOH);
            float vzz = Dm(S->vz, indv, 1, hc, FDOH);
            float vzx = Dp(S->vz, indv, NZ, hc, FDOH);
            if (cpml) {
                int i, k, ind;
                if (gidz > NZ - NAB - FDOH - 1) {
                    i = gidx - FDOH;
                    k = gidz - NZ + NAB + FDOH + NAB;
                    ind = 2 * NAB - 1 - k;
                    float *p1 = &S->psi_vx_z[(long)i * (2 * NAB) + k];
                    *p1 = S->b_z_half[ind] * *p1 + S->a_z_half[ind] * vxz;
                    vxz = vxz / S->K_z_half[ind] + *p1;
                    float *p2 = &S->psi_vz_z[(long)i * (2 * NAB) + k];
                    *p2 = S->b_z[ind + 1] * *p2 + S->a_z[ind + 1] * vzz;
                    vzz = vzz / S->K_z[ind + 1] + *p2;
                } else if (!fs && gidz - FDOH < NAB) {
                    i = gidx - FDOH;
                    k = gidz - FDOH;
                    float *p1 = &S->psi_vx_z[(long)i * (2 * NAB) + k];
                    *p1 = S->b_z_half[k] * *p1 + S->a_z_half[k] * vxz;
                    vxz = vxz / S->K_z_half[k] + *p1;
                    float *p2 = &S->psi_vz_z[(long)i * (2 * NAB) + k];
                    *p2 = S->b_z[k] * *p2 + S->a_z[k] * vzz;
                    vzz = vzz / S->K_z[k] + *p2;
                }
                if (gidx - FDOH < NAB) {
                    i = gidx - FDOH;
                    k = gidz - FDOH;
                    float *p1 = &S->psi_vx_x[(long)i * (NZ - 2 * FDOH) + k];
                    *p1 = S->b_x[i] * *p1 + S->a_x[i] * vxx;
                    vxx = vxx / S->K_x[i] + *p1;
                    float *p2 = &S->psi_vz_x[(long)i * (NZ - 2 * FDOH) + k];
                    *p2 = S->b_x_half[i] * *p2 + S->a_x_half[i] * vzx;
                    vzx = vzx / S->K_x_half[i] + *p2;
                }
                if (gidx > NX - NAB - FDOH - 1) {
                    i = gidx - NX + NAB + FDOH + NAB;
                    k = gidz - FDOH;
                    ind = 2 * NAB - 1 - i;
                    float *p1 = &S->psi_vx_x[(long)i * (NZ - 2 * FDOH) + k];
                    *p1 = S->b_x[ind + 1] * *p1 + S->a_x[ind + 1] * vxx;
                    vxx = vxx / S->K_x[ind + 1] + *p1;
                    float *p2 = &S->psi_vz_x[(long)i * (NZ - 2 * FDOH) + k];
                    *p2 = S->b_x_half[ind] * *p2 + S->a_x_half[ind] * vzx;
                    vzx = vzx / S->K_x_half[ind] + *p2;
                }
            }
            if (LVE == 0) {
                float fipkp = S->muipkp[indp];
                float f = 2.0 * S->mu[indp];
                float g = S->M[indp];
                float lsxz = (fipkp * (vxz + vzx));
                float lsxx = (g * (vxx + vzz)) - (f * vzz);
                float lszz = (g * (vxx + vzz)) - (f * vxx);
                S->sxz[indv] += lsxz;
                S->sxx[indv] += lsxx;
                S->szz[indv] += lszz;
                if (S->inc[0]) {
                    S->inc[0][indv] = lsxx;
                    S->inc[1][indv] = lszz;
                    S->inc[2][indv] = lsxz;
                }
            } else {
                float lM = S->M[indp], lmu = S->mu[indp], lmuipkp = S->muipkp[indp];
                float ltaup = S->taup[indp], ltaus = S->taus[indp],
                      ltausipkp = S->tausipkp[indp];
                float fipkp = lmuipkp * (1.0 + (float)LVE * ltausipkp);
                float g = lM * (1.0 + (float)LVE * ltaup);
                float f = 2.0 * lmu * (1.0 + (float)LVE * ltaus);
                float dipkp = lmuipkp * ltausipkp / DT;
                float d = 2.0 * lmu * ltaus / DT;
                float e = lM * ltaup / DT;
                float sumrxz = 0, sumrxx = 0, sumrzz = 0;
                for (int l = 0; l < LVE; l++) {
                    long indr = (long)l * NX * NZ + indv;
                    sumrxz += S->rxz[indr];
                    sumrxx += S->rxx[indr];
                    sumrzz += S->rzz[indr];
                }
                float lsxz = (fipkp * (vxz + vzx)) + (DT2 * sumrxz);
                float lsxx = ((g * (vxx + vzz)) - (f * vzz)) + (DT2 * sumrxx);
                float lszz = ((g * (vxx + vzz)) - (f * vxx)) + (DT2 * sumrzz);
                if (S->inc[0]) {
                    /* the partial increments, which the reference's viscoelastic */
                    /* gradient reads (src/update_adjs2D.cl:362-364, 441-451)     */
                    S->inc[0][indv] = lsxx;
                    S->inc[1][indv] = lszz;
                    S->inc[2][indv] = lsxz;
                }
                const float oxx = sumrxx, ozz = sumrzz, oxz = sumrxz;
                sumrxz = sumrxx = sumrzz = 0;
                for (int l = 0; l < LVE; l++) {
                    float b = 1.0 / (1.0 + (S->eta[l] * 0.5));
                    float cc = 1.0 - (S->eta[l] * 0.5);
                    long indr = (long)l * NX * NZ + indv;
                    S->rxz[indr] = b * (S->rxz[indr] * cc - S->eta[l] * (dipkp * (vxz + vzx)));
                    S->rxx[indr] = b * (S->rxx[indr] * cc -
                                        S->eta[l] * ((e * (vxx + vzz)) - (d * vzz)));
                    S->rzz[indr] = b * (S->rzz[indr] * cc -
                                        S->eta[l] * ((e * (vxx + vzz)) - (d * vxx)));
                    sumrxz += S->rxz[indr];
                    sumrxx += S->rxx[indr];
                    sumrzz += S->rzz[indr];
                }
                S->sxz[indv] += lsxz + (DT2 * sumrxz);
                S->sxx[indv] += lsxx + (DT2 * sumrxx);
                S->szz[indv] += lszz + (DT2 * sumrzz);
                if (S->tinc[0]) {
                    /* the whole increment: memory of the old and new levels */
                    float vii[2] = {vxx, vzz};
                    float Rm[2] = {DT2 * oxx + DT2 * sumrxx, DT2 * ozz + DT2 * sumrzz};
                    float Xp, Xs[2];
                    tau_parts_normal(vxx + vzz, vii, Rm, 2, g, f, d, e, (float)LVE, DT, S->s0,
                                     &Xp, Xs);
                    S->tinc[0][indv] = Xp;
                    S->tinc[1][indv] = Xs[0];
                    S->tinc[2][indv] = Xs[1];
                    S->tinc[3][indv] =
                        tau_part_shear(vxz + vzx, DT2 * oxz + DT2 * sumrxz, fipkp, dipkp,
                                       lsxz + (DT2 * sumrxz), (float)LVE, DT, S->s0);
                }
            }
            if (tap) {
                float tp = 1.0f;
                int any = 0;
                (void)tp;
                (void)any;
                if (!fs && gidz - FDOH < NAB) {
                    S->sxx[indv] *= S->taper[gidz - FDOH];
                    S->szz[indv] *= S->taper[gidz - FDOH];
                    S->sxz[indv] *= S->taper[gidz - FDOH];
                }
                if (gidz > NZ - NAB - FDOH - 1) {
                    S->sxx[indv] *= S->taper[NZ - FDOH - gidz - 1];
                    S->szz[indv] *= S->taper[NZ - FDOH - gidz - 1];
                    S->sxz[indv] *= S->taper[NZ - FDOH - gidz - 1];
                }
                if (gidx - FDOH < NAB) {
                    S->sxx[indv] *= S->taper[gidx - FDOH];
                    S->szz[indv] *= S->taper[gidx - FDOH];
                    S->sxz[indv] *= S->taper[gidx - FDOH];
                }
                if (gidx > NX - NAB - FDOH - 1) {
                    S->sxx[indv] *= S->taper[NX - FDOH - gidx - 1];
                    S->szz[indv] *= S->taper[NX - FDOH - gidx - 1];
                    S->sxz[indv] *= S->taper[NX - FDOH - gidx - 1];
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* 2D acoustic: src/update_v2D_acc.cl:80-325, src/update_s2D_acc.cl:79-402   */
/* ------------------------------------------------------------------------ */
static void update_v2D_acc(sim *S) {
    const int FDOH = S->FDOH, NAB = S->NAB, NZ = S->NZp, NX = S->NXp;
    const int cpml = S->c->abs_type == 1, tap = S->c->abs_type == 2,
              fs = S->c->freesurf;
    const float *hc = S->hc;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++) {
        for (int gidz = FDOH; gidz < NZ - FDOH; gidz++) {
            long indp = (long)(gidx - FDOH) * (NZ - 2 * FDOH) + (gidz - FDOH);
            long indv = (long)gidx * NZ + gidz;
            float p_x = Dp(S->p, indv, NZ, hc, FDOH);
            float p_z = Dp(S->p, indv, 1, hc, FDOH);
            if (cpml) {
                int i, k, ind;
                if (gidz > NZ - NAB - FDOH - 1) {
                    i = gidx - FDOH;
                    k = gidz - NZ + NAB + FDOH + NAB;
                    ind = 2 * NAB - 1 - k;
                    float *pp = &S->psi_p_z[(long)i * (2 * NAB) + k];
                    *pp = S->b_z_half[ind] * *pp + S->a_z_half[ind] * p_z;
                    p_z = p_z / S->K_z_half[ind] + *pp;
                } else if (!fs && gidz - FDOH < NAB) {
                    i = gidx - FDOH;
                    k = gidz - FDOH;
                    float *pp = &S->psi_p_z[(long)i * (2 * NAB) + k];
                    *pp = S->b_z_half[k] * *pp + S->a_z_half[k] * p_z;
                    p_z = p_z / S->K_z_half[k] + *pp;
                }
                if (gidx - FDOH < NAB) {
                    i = gidx - FDOH;
                    k = gidz - FDOH;
                    float *pp = &S->psi_p_x[(long)i * (NZ - 2 * FDOH) + k];
                    *pp = S->b_x_half[i] * *pp + S->a_x_half[i] * p_x;
                    p_x = p_x / S->K_x_half[i] + *pp;
                }
                if (gidx > NX - NAB - FDOH - 1) {
                    i = gidx - NX + NAB + FDOH + NAB;
                    k = gidz - FDOH;
                    ind = 2 * NAB - 1 - i;
                    float *pp = &S->psi_p_x[(long)i * (NZ - 2 * FDOH) + k];
                    *pp = S->b_x_half[ind] * *pp + S->a_x_half[ind] * p_x;
                    p_x = p_x / S->K_x_half[ind] + *pp;
                }
            }
            S->vx[indv] += (p_x * S->rip[indp]);
            S->vz[indv] += (p_z * S->rkp[indp]);
            if (tap) {
                /* the acoustic kernel tapers the top edge regardless of the
                 * free-surface flag (src/update_v2D_acc.cl:297-300) */
                if (gidz - FDOH < NAB) {
                    S->vx[indv] *= S->taper[gidz - FDOH];
                    S->vz[indv] *= S->taper[gidz - FDOH];
                }
                if (gidz > NZ - NAB - FDOH - 1) {
                    S->vx[indv] *= S->taper[NZ - FDOH - gidz - 1];
                    S->vz[indv] *= S->taper[NZ - FDOH - gidz - 1];
                }
                if (gidx - FDOH < NAB) {
                    S->vx[indv] *= S->taper[gidx - FDOH];
                    S->vz[indv] *= S->taper[gidx - FDOH];
                }
                if (gidx > NX - NAB - FDOH - 1) {
                    S->vx[indv] *= S->taper[NX - FDOH - gidx - 1];
                    S->vz[indv] *= S->taper[NX - FDOH - gidx - 1];
                }
            }
        }
    }
}

static void update_s2D_acc(sim *S) {
    const int FDOH = S->FDOH, NAB = S->NAB, NZ = S->NZp, NX = S->NXp,
              LVE = S->L;
    const int cpml = S->c->abs_type == 1, tap = S->c->abs_type == 2,
              fs = S->c->freesurf;
    const float *hc = S->hc, DT2 = S->DT2;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++) {
        for (int gidz = FDOH; gidz < NZ - FDOH; gidz++) {
            long indp = (long)(gidx - FDOH) * (NZ - 2 * FDOH) + (gidz - FDOH);
            long indv = (long)gidx * NZ + gidz;
            float vxx = Dm(S->vx, indv, NZ, hc, FDOH);
            float vzz = Dm(S->vz, indv, 1, hc, FDOH);
            if (cpml) {
                int i, k, ind;
                if (gidz > NZ - NAB - FDOH - 1) {
                    i = gidx - FDOH;
                    k = gidz - NZ + NAB + FDOH + NAB;
                    ind = 2 * NAB - 1 - k;
                    float *pp = &S->psi_vz_z[(long)i * (2 * NAB) + k];
                    *pp = S->b_z[ind + 1] * *pp + S->a_z[ind + 1] * vzz;
                    vzz = vzz / S->K_z[ind + 1] + *pp;
                } else if (!fs && gidz - FDOH < NAB) {
                    i = gidx - FDOH;
                    k = gidz - FDOH;
                    float *pp = &S->psi_vz_z[(long)i * (2 * NAB) + k];
                    *pp = S->b_z[k] * *pp + S->a_z[k] * vzz;
                    vzz = vzz / S->K_z[k] + *pp;
                }
                if (gidx - FDOH < NAB) {
                    i = gidx - FDOH;
                    k = gidz - FDOH;
                    float *pp = &S->psi_vx_x[(long)i * (NZ - 2 * FDOH) + k];
                    *pp = S->b_x[i] * *pp + S->a_x[i] * vxx;
                    vxx = vxx / S->K_x[i] + *pp;
                }
                if (gidx > NX - NAB - FDOH - 1) {
                    i = gidx - NX + NAB + FDOH + NAB;
                    k = gidz - FDOH;
                    ind = 2 * NAB - 1 - i;
                    float *pp = &S->psi_vx_x[(long)i * (NZ - 2 * FDOH) + k];
                    *pp = S->b_x[ind + 1] * *pp + S->a_x[ind + 1] * vxx;
                    vxx = vxx / S->K_x[ind + 1] + *pp;
                }
            }
            if (LVE == 0) {
                float g = S->M[indp];
                S->p[indv] += (g * (vxx + vzz));
            } else {
                float lM = S->M[indp], ltaup = S->taup[indp];
                float g = lM * (1.0 + (float)LVE * ltaup);
                float e = lM * ltaup; /* no 1/DT here, unlike the elastic kernel */
                float sumrp = 0;
                for (int l = 0; l < LVE; l++) sumrp += S->rp[(long)l * NX * NZ + indv];
                float lp = (g * (vxx + vzz)) + (DT2 * sumrp);
                sumrp = 0;
                for (int l = 0; l < LVE; l++) {
                    float b = 1.0 / (1.0 + (S->eta[l] * 0.5));
                    float cc = 1.0 - (S->eta[l] * 0.5);
                    long indr = (long)l * NX * NZ + indv;
                    S->rp[indr] = b * (S->rp[indr] * cc - S->eta[l] * e * (vxx + vzz));
                    sumrp += S->rp[indr];
                }
                S->p[indv] += lp + (DT2 * sumrp);
            }
            if (tap) {
                if (gidz - FDOH < NAB) S->p[indv] *= S->taper[gidz - FDOH];
                if (gidz > NZ - NAB - FDOH - 1) S->p[indv] *= S->taper[NZ - FDOH - gidz - 1];
                if (gidx - FDOH < NAB) S->p[indv] *= S->taper[gidx - FDOH];
                if (gidx > NX - NAB - FDOH - 1) S->p[indv] *= S->taper[NX - FDOH - gidx - 1];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* 3D velocity update: src/update_v3D.cl:34-369                              */
/* ------------------------------------------------------------------------ */
static void update_v3D(sim *S) {
    const int FDOH = S->FDOH, NAB = S->NAB, NZ = S->NZp, NY = S->NYp, NX = S->NXp;
    const int cpml = S->c->abs_type == 1, tap = S->c->abs_type == 2,
              fs = S->c->freesurf;
    const float *hc = S->hc;
    const long sy = NZ, sx = (long)NY * NZ;
    const long nzi = NZ - 2 * FDOH, nyi = NY - 2 * FDOH;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++) {
        for (int gidy = FDOH; gidy < NY - FDOH; gidy++) {
            for (int gidz = FDOH; gidz < NZ - FDOH; gidz++) {
                long indp = (long)(gidx - FDOH) * nyi * nzi + (long)(gidy - FDOH) * nzi +
                            (gidz - FDOH);
                long indv = (long)gidx * sx + (long)gidy * sy + gidz;
                float szz_z = Dp(S->szz, indv, 1, hc, FDOH);
                float sxx_x = Dp(S->sxx, indv, sx, hc, FDOH);
                float sxz_x = Dm(S->sxz, indv, sx, hc, FDOH);
                float sxz_z = Dm(S->sxz, indv, 1, hc, FDOH);
                float syz_y = Dm(S->syz, indv, sy, hc, FDOH);
                float syz_z = Dm(S->syz, indv, 1, hc, FDOH);
                float syy_y = Dp(S->syy, indv, sy, hc, FDOH);
                float sxy_x = Dm(S->sxy, indv, sx, hc, FDOH);
                float sxy_y = Dm(S->sxy, indv, sy, hc, FDOH);
                if (cpml) {
                    int i, j, k, indm;
#define PZ(a) a[(long)i * nyi * (2 * NAB) + (long)j * (2 * NAB) + k]
#define PY(a) a[(long)i * (2 * NAB) * nzi + (long)j * nzi + k]
#define PX(a) a[(long)i * nyi * nzi + (long)j * nzi + k]
                    if (gidz > NZ - NAB - FDOH - 1) {
                        i = gidx - FDOH;
                        j = gidy - FDOH;
                        k = gidz - NZ + NAB + FDOH + NAB;
                        indm = 2 * NAB - 1 - k;
                        PZ(S->psi_sxz_z) = S->b_z[indm + 1] * PZ(S->psi_sxz_z) + S->a_z[indm + 1] * sxz_z;
                        sxz_z = sxz_z / S->K_z[indm + 1] + PZ(S->psi_sxz_z);
                        PZ(S->psi_syz_z) = S->b_z[indm + 1] * PZ(S->psi_syz_z) + S->a_z[indm + 1] * syz_z;
                        syz_z = syz_z / S->K_z[indm + 1] + PZ(S->psi_syz_z);
                        PZ(S->psi_szz_z) = S->b_z_half[indm] * PZ(S->psi_szz_z) + S->a_z_half[indm] * szz_z;
                        szz_z = szz_z / S->K_z_half[indm] + PZ(S->psi_szz_z);
                    } else if (!fs && gidz - FDOH < NAB) {
                        i = gidx - FDOH;
                        j = gidy - FDOH;
                        k = gidz - FDOH;
                        PZ(S->psi_sxz_z) = S->b_z[k] * PZ(S->psi_sxz_z) + S->a_z[k] * sxz_z;
                        sxz_z = sxz_z / S->K_z[k] + PZ(S->psi_sxz_z);
                        PZ(S->psi_syz_z) = S->b_z[k] * PZ(S->psi_syz_z) + S->a_z[k] * syz_z;
                        syz_z = syz_z / S->K_z[k] + PZ(S->psi_syz_z);
                        PZ(S->psi_szz_z) = S->b_z_half[k] * PZ(S->psi_szz_z) + S->a_z_half[k] * szz_z;
                        szz_z = szz_z / S->K_z_half[k] + PZ(S->psi_szz_z);
                    }
                    if (gidy - FDOH < NAB) {
                        i = gidx - FDOH;
                        j = gidy - FDOH;
                        k = gidz - FDOH;
                        PY(S->psi_sxy_y) = S->b_y[j] * PY(S->psi_sxy_y) + S->a_y[j] * sxy_y;
                        sxy_y = sxy_y / S->K_y[j] + PY(S->psi_sxy_y);
                        PY(S->psi_syy_y) = S->b_y_half[j] * PY(S->psi_syy_y) + S->a_y_half[j] * syy_y;
                        syy_y = syy_y / S->K_y_half[j] + PY(S->psi_syy_y);
                        PY(S->psi_syz_y) = S->b_y[j] * PY(S->psi_syz_y) + S->a_y[j] * syz_y;
                        syz_y = syz_y / S->K_y[j] + PY(S->psi_syz_y);
                    } else if (gidy > NY - NAB - FDOH - 1) {
                        i = gidx - FDOH;
                        j = gidy - NY + NAB + FDOH + NAB;
                        k = gidz - FDOH;
                        indm = 2 * NAB - 1 - j;
                        PY(S->psi_sxy_y) = S->b_y[indm + 1] * PY(S->psi_sxy_y) + S->a_y[indm + 1] * sxy_y;
                        sxy_y = sxy_y / S->K_y[indm + 1] + PY(S->psi_sxy_y);
                        PY(S->psi_syy_y) = S->b_y_half[indm] * PY(S->psi_syy_y) + S->a_y_half[indm] * syy_y;
                        syy_y = syy_y / S->K_y_half[indm] + PY(S->psi_syy_y);
                        PY(S->psi_syz_y) = S->b_y[indm + 1] * PY(S->psi_syz_y) + S->a_y[indm + 1] * syz_y;
                        syz_y = syz_y / S->K_y[indm + 1] + PY(S->psi_syz_y);
                    }
                    if (gidx - FDOH < NAB) {
                        i = gidx - FDOH;
                        j = gidy - FDOH;
                        k = gidz - FDOH;
                        PX(S->psi_sxx_x) = S->b_x_half[i] * PX(S->psi_sxx_x) + S->a_x_half[i] * sxx_x;
                        sxx_x = sxx_x / S->K_x_half[i] + PX(S->psi_sxx_x);
                        PX(S->psi_sxy_x) = S->b_x[i] * PX(S->psi_sxy_x) + S->a_x[i] * sxy_x;
                        sxy_x = sxy_x / S->K_x[i] + PX(S->psi_sxy_x);
                        PX(S->psi_sxz_x) = S->b_x[i] * PX(S->psi_sxz_x) + S->a_x[i] * sxz_x;
                        sxz_x = sxz_x / S->K_x[i] + PX(S->psi_sxz_x);
                    }
                    if (gidx > NX - NAB - FDOH - 1) {
                        i = gidx - NX + NAB + FDOH + NAB;
                        j = gidy - FDOH;
                        k = gidz - FDOH;
                        indm = 2 * NAB - 1 - i;
                        PX(S->psi_sxx_x) = S->b_x_half[indm] * PX(S->psi_sxx_x) + S->a_x_half[indm] * sxx_x;
                        sxx_x = sxx_x / S->K_x_half[indm] + PX(S->psi_sxx_x);
                        PX(S->psi_sxy_x) = S->b_x[indm + 1] * PX(S->psi_sxy_x) + S->a_x[indm + 1] * sxy_x;
                        sxy_x = sxy_x / S->K_x[indm + 1] + PX(S->psi_sxy_x);
                        PX(S->psi_sxz_x) = S->b_x[indm + 1] * PX(S->psi_sxz_x) + S->a_x[indm + 1] * sxz_x;
                        sxz_x = sxz_x / S->K_x[indm + 1] + PX(S->psi_sxz_x);
                    }
                }
                float lvx = ((sxx_x + sxy_y + sxz_z) * S->rip[indp]);
                float lvy = ((syy_y + sxy_x + syz_z) * S->rjp[indp]);
                float lvz = ((szz_z + sxz_x + syz_y) * S->rkp[indp]);
                S->vx[indv] += lvx;
                S->vy[indv] += lvy;
                S->vz[indv] += lvz;
                if (S->inc[0]) {
                    S->inc[0][indv] = lvx;
                    S->inc[1][indv] = lvy;
                    S->inc[2][indv] = lvz;
                }
                if (tap) {
                    float *V[3] = {&S->vx[indv], &S->vy[indv], &S->vz[indv]};
                    for (int q = 0; q < 3; q++) {
                        if (!fs && gidz - FDOH < NAB) *V[q] *= S->taper[gidz - FDOH];
                        if (gidz > NZ - NAB - FDOH - 1) *V[q] *= S->taper[NZ - FDOH - gidz - 1];
                        if (gidy - FDOH < NAB) *V[q] *= S->taper[gidy - FDOH];
                        if (gidy > NY - NAB - FDOH - 1) *V[q] *= S->taper[NY - FDOH - gidy - 1];
                        if (gidx - FDOH < NAB) *V[q] *= S->taper[gidx - FDOH];
                        if (gidx > NX - NAB - FDOH - 1) *V[q] *= S->taper[NX - FDOH - gidx - 1];
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* 3D stress update: src/update_s3D.cl:35-489                                */
/* ------------------------------------------------------------------------ */
static void update_s3D(sim *S) {
    const int FDOH = S->FDOH, NAB = S->NAB, NZ = S->NZp, NY = S->NYp, NX = S->NXp,
              LVE = S->L;
    const int cpml = S->c->abs_type == 1, tap = S->c->abs_type == 2,
              fs = S->c->freesurf;
    const float *hc = S->hc, DT = S->DT, DT2 = S->DT2;
    const long sy = NZ, sx = (long)NY * NZ;
    const long nzi = NZ - 2 * FDOH, nyi = NY - 2 * FDOH;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++) {
        for (int gidy = FDOH; gidy < NY - FDOH; gidy++) {
            for (int gidz = FDOH; gidz < NZ - FDOH; gidz++) {
                long indp = (long)(gidx - FDOH) * nyi * nzi + (long)(gidy - FDOH) * nzi +
                            (gidz - FDOH);
                long indv = (long)gidx * sx + (long)gidy * sy + gidz;
                float vzx = Dp(S->vz, indv, sx, hc, FDOH);
                float vzy = Dp(S->vz, indv, sy, hc, FDOH);
                float vzz = Dm(S->vz, indv, 1, hc, FDOH);
                float vyx = Dp(S->vy, indv, sx, hc, FDOH);
                float vyy = Dm(S->vy, indv, sy, hc, FDOH);
                float vyz = Dp(S->vy, indv, 1, hc, FDOH);
                float vxx = Dm(S->vx, indv, sx, hc, FDOH);
                float vxy = Dp(S->vx, indv, sy, hc, FDOH);
                float vxz = Dp(S->vx, indv, 1, hc, FDOH);
                if (cpml) {
                    int i, j, k, indm;
                    if (gidz > NZ - NAB - FDOH - 1) {
                        i = gidx - FDOH;
                        j = gidy - FDOH;
                        k = gidz - NZ + NAB + FDOH + NAB;
                        indm = 2 * NAB - 1 - k;
                        PZ(S->psi_vx_z) = S->b_z_half[indm] * PZ(S->psi_vx_z) + S->a_z_half[indm] * vxz;
                        vxz = vxz / S->K_z_half[indm] + PZ(S->psi_vx_z);
                        PZ(S->psi_vy_z) = S->b_z_half[indm] * PZ(S->psi_vy_z) + S->a_z_half[indm] * vyz;
                        vyz = vyz / S->K_z_half[indm] + PZ(S->psi_vy_z);
                        PZ(S->psi_vz_z) = S->b_z[indm + 1] * PZ(S->psi_vz_z) + S->a_z[indm + 1] * vzz;
                        vzz = vzz / S->K_z[indm + 1] + PZ(S->psi_vz_z);
                    }
                    if (!fs && gidz - FDOH < NAB) {
                        i = gidx - FDOH;
                        j = gidy - FDOH;
                        k = gidz - FDOH;
                        PZ(S->psi_vx_z) = S->b_z_half[k] * PZ(S->psi_vx_z) + S->a_z_half[k] * vxz;
                        vxz = vxz / S->K_z_half[k] + PZ(S->psi_vx_z);
                        PZ(S->psi_vy_z) = S->b_z_half[k] * PZ(S->psi_vy_z) + S->a_z_half[k] * vyz;
                        vyz = vyz / S->K_z_half[k] + PZ(S->psi_vy_z);
                        PZ(S->psi_vz_z) = S->b_z[k] * PZ(S->psi_vz_z) + S->a_z[k] * vzz;
                        vzz = vzz / S->K_z[k] + PZ(S->psi_vz_z);
                    }
                    if (gidy - FDOH < NAB) {
                        i = gidx - FDOH;
                        j = gidy - FDOH;
                        k = gidz - FDOH;
                        PY(S->psi_vx_y) = S->b_y_half[j] * PY(S->psi_vx_y) + S->a_y_half[j] * vxy;
                        vxy = vxy / S->K_y_half[j] + PY(S->psi_vx_y);
                        PY(S->psi_vy_y) = S->b_y[j] * PY(S->psi_vy_y) + S->a_y[j] * vyy;
                        vyy = vyy / S->K_y[j] + PY(S->psi_vy_y);
                        PY(S->psi_vz_y) = S->b_y_half[j] * PY(S->psi_vz_y) + S->a_y_half[j] * vzy;
                        vzy = vzy / S->K_y_half[j] + PY(S->psi_vz_y);
                    }
                    if (gidy > NY - NAB - FDOH - 1) {
                        i = gidx - FDOH;
                        j = gidy - NY + NAB + FDOH + NAB;
                        k = gidz - FDOH;
                        indm = 2 * NAB - 1 - j;
                        PY(S->psi_vx_y) = S->b_y_half[indm] * PY(S->psi_vx_y) + S->a_y_half[indm] * vxy;
                        vxy = vxy / S->K_y_half[indm] + PY(S->psi_vx_y);
                        PY(S->psi_vy_y) = S->b_y[indm + 1] * PY(S->psi_vy_y) + S->a_y[indm + 1] * vyy;
                        vyy = vyy / S->K_y[indm + 1] + PY(S->psi_vy_y);
                        PY(S->psi_vz_y) = S->b_y_half[indm] * PY(S->psi_vz_y) + S->a_y_half[indm] * vzy;
                        vzy = vzy / S->K_y_half[indm] + PY(S->psi_vz_y);
                    }
                    if (gidx - FDOH < NAB) {
                        i = gidx - FDOH;
                        j = gidy - FDOH;
                        k = gidz - FDOH;
                        PX(S->psi_vx_x) = S->b_x[i] * PX(S->psi_vx_x) + S->a_x[i] * vxx;
                        vxx = vxx / S->K_x[i] + PX(S->psi_vx_x);
                        PX(S->psi_vy_x) = S->b_x_half[i] * PX(S->psi_vy_x) + S->a_x_half[i] * vyx;
                        vyx = vyx / S->K_x_half[i] + PX(S->psi_vy_x);
                        PX(S->psi_vz_x) = S->b_x_half[i] * PX(S->psi_vz_x) + S->a_x_half[i] * vzx;
                        vzx = vzx / S->K_x_half[i] + PX(S->psi_vz_x);
                    }
                    if (gidx > NX - NAB - FDOH - 1) {
                        i = gidx - NX + NAB + FDOH + NAB;
                        j = gidy - FDOH;
                        k = gidz - FDOH;
                        indm = 2 * NAB - 1 - i;
                        PX(S->psi_vx_x) = S->b_x[indm + 1] * PX(S->psi_vx_x) + S->a_x[indm + 1] * vxx;
                        vxx = vxx / S->K_x[indm + 1] + PX(S->psi_vx_x);
                        PX(S->psi_vy_x) = S->b_x_half[indm] * PX(S->psi_vy_x) + S->a_x_half[indm] * vyx;
                        vyx = vyx / S->K_x_half[indm] + PX(S->psi_vy_x);
                        PX(S->psi_vz_x) = S->b_x_half[indm] * PX(S->psi_vz_x) + S->a_x_half[indm] * vzx;
                        vzx = vzx / S->K_x_half[indm] + PX(S->psi_vz_x);
                    }
                }
                float vxyyx = vxy + vyx;
                float vyzzy = vyz + vzy;
                float vxzzx = vxz + vzx;
                float vxxyyzz = vxx + vyy + vzz;
                float vyyzz = vyy + vzz;
                float vxxzz = vxx + vzz;
                float vxxyy = vxx + vyy;
                if (LVE == 0) {
                    float fipjp = S->muipjp[indp];
                    float fjpkp = S->mujpkp[indp];
                    float fipkp = S->muipkp[indp];
                    float g = S->M[indp];
                    float f = 2.0 * S->mu[indp];
                    float lsxy = (fipjp * vxyyx), lsyz = (fjpkp * vyzzy), lsxz = (fipkp * vxzzx);
                    float lsxx = ((g * vxxyyzz) - (f * vyyzz));
                    float lsyy = ((g * vxxyyzz) - (f * vxxzz));
                    float lszz = ((g * vxxyyzz) - (f * vxxyy));
                    S->sxy[indv] += lsxy;
                    S->syz[indv] += lsyz;
                    S->sxz[indv] += lsxz;
                    S->sxx[indv] += lsxx;
                    S->syy[indv] += lsyy;
                    S->szz[indv] += lszz;
                    if (S->inc[0]) {
                        S->inc[0][indv] = lsxx;
                        S->inc[1][indv] = lsyy;
                        S->inc[2][indv] = lszz;
                        S->inc[3][indv] = lsxy;
                        S->inc[4][indv] = lsyz;
                        S->inc[5][indv] = lsxz;
                    }
                } else {
                    float lM = S->M[indp], lmu = S->mu[indp];
                    float lmuipkp = S->muipkp[indp], lmuipjp = S->muipjp[indp],
                          lmujpkp = S->mujpkp[indp];
                    float ltaup = S->taup[indp], ltaus = S->taus[indp];
                    float ltausipkp = S->tausipkp[indp], ltausipjp = S->tausipjp[indp],
                          ltausjpkp = S->tausjpkp[indp];
                    float fipjp = lmuipjp * (1.0 + (float)LVE * ltausipjp);
                    float fjpkp = lmujpkp * (1.0 + (float)LVE * ltausjpkp);
                    float fipkp = lmuipkp * (1.0 + (float)LVE * ltausipkp);
                    float g = lM * (1.0 + (float)LVE * ltaup);
                    float f = 2.0 * lmu * (1.0 + (float)LVE * ltaus);
                    float dipjp = lmuipjp * ltausipjp / DT;
                    float djpkp = lmujpkp * ltausjpkp / DT;
                    float dipkp = lmuipkp * ltausipkp / DT;
                    float d = 2.0 * lmu * ltaus / DT;
                    float e = lM * ltaup / DT;
                    float sumrxy = 0, sumryz = 0, sumrxz = 0, sumrxx = 0, sumryy = 0,
                          sumrzz = 0;
                    const long NV = (long)NX * NY * NZ;
                    for (int l = 0; l < LVE; l++) {
                        long indr = l * NV + indv;
                        sumrxy += S->rxy[indr];
                        sumryz += S->ryz[indr];
                        sumrxz += S->rxz[indr];
                        sumrxx += S->rxx[indr];
                        sumryy += S->ryy[indr];
                        sumrzz += S->rzz[indr];
                    }
                    float lsxy = (fipjp * vxyyx) + (DT2 * sumrxy);
                    float lsyz = (fjpkp * vyzzy) + (DT2 * sumryz);
                    float lsxz = (fipkp * vxzzx) + (DT2 * sumrxz);
                    float lsxx = ((g * vxxyyzz) - (f * vyyzz)) + (DT2 * sumrxx);
                    float lsyy = ((g * vxxyyzz) - (f * vxxzz)) + (DT2 * sumryy);
                    float lszz = ((g * vxxyyzz) - (f * vxxyy)) + (DT2 * sumrzz);
                    if (S->inc[0]) {
                        /* partial increments (src/update_adjs3D.cl:570-575, 679-689) */
                        S->inc[0][indv] = lsxx;
                        S->inc[1][indv] = lsyy;
                        S->inc[2][indv] = lszz;
                        S->inc[3][indv] = lsxy;
                        S->inc[4][indv] = lsyz;
                        S->inc[5][indv] = lsxz;
                    }
                    const float oxy = sumrxy, oyz = sumryz, oxz = sumrxz, oxx = sumrxx,
                                oyy = sumryy, ozz = sumrzz;
                    sumrxy = sumryz = sumrxz = sumrxx = sumryy = sumrzz = 0;
                    for (int l = 0; l < LVE; l++) {
                        float b = 1.0 / (1.0 + (S->eta[l] * 0.5));
                        float cc = 1.0 - (S->eta[l] * 0.5);
                        long indr = l * NV + indv;
                        S->rxy[indr] = b * (S->rxy[indr] * cc - S->eta[l] * (dipjp * vxyyx));
                        S->ryz[indr] = b * (S->ryz[indr] * cc - S->eta[l] * (djpkp * vyzzy));
                        S->rxz[indr] = b * (S->rxz[indr] * cc - S->eta[l] * (dipkp * vxzzx));
                        S->rxx[indr] = b * (S->rxx[indr] * cc -
                                            S->eta[l] * ((e * vxxyyzz) - (d * vyyzz)));
                        S->ryy[indr] = b * (S->ryy[indr] * cc -
                                            S->eta[l] * ((e * vxxyyzz) - (d * vxxzz)));
                        S->rzz[indr] = b * (S->rzz[indr] * cc -
                                            S->eta[l] * ((e * vxxyyzz) - (d * vxxyy)));
                        sumrxy += S->rxy[indr];
                        sumryz += S->ryz[indr];
                        sumrxz += S->rxz[indr];
                        sumrxx += S->rxx[indr];
                        sumryy += S->ryy[indr];
                        sumrzz += S->rzz[indr];
                    }
                    S->sxy[indv] += lsxy + (DT2 * sumrxy);
                    S->syz[indv] += lsyz + (DT2 * sumryz);
                    S->sxz[indv] += lsxz + (DT2 * sumrxz);
                    S->sxx[indv] += lsxx + (DT2 * sumrxx);
                    S->syy[indv] += lsyy + (DT2 * sumryy);
                    S->szz[indv] += lszz + (DT2 * sumrzz);
                    if (S->tinc[0]) {
                        /* the whole increment: memory of the old and new levels */
                        float vii[3] = {vxx, vyy, vzz};
                        float Rm[3] = {DT2 * oxx + DT2 * sumrxx, DT2 * oyy + DT2 * sumryy,
                                       DT2 * ozz + DT2 * sumrzz};
                        float Xp, Xs[3];
                        const float Lf = (float)LVE;
                        tau_parts_normal(vxxyyzz, vii, Rm, 3, g, f, d, e, Lf, DT, S->s0, &Xp, Xs);
                        S->tinc[0][indv] = Xp;
                        S->tinc[1][indv] = Xs[0];
                        S->tinc[2][indv] = Xs[1];
                        S->tinc[3][indv] = Xs[2];
                        S->tinc[4][indv] =
                            tau_part_shear(vxyyx, DT2 * oxy + DT2 * sumrxy, fipjp, dipjp,
                                           lsxy + (DT2 * sumrxy), Lf, DT, S->s0);
                        S->tinc[5][indv] =
                            tau_part_shear(vyzzy, DT2 * oyz + DT2 * sumryz, fjpkp, djpkp,
                                           lsyz + (DT2 * sumryz), Lf, DT, S->s0);
                        S->tinc[6][indv] =
                            tau_part_shear(vxzzx, DT2 * oxz + DT2 * sumrxz, fipkp, dipkp,
                                           lsxz + (DT2 * sumrxz), Lf, DT, S->s0);
                    }
                }
                if (tap) {
                    float *V[6] = {&S->sxy[indv], &S->syz[indv], &S->sxz[indv],
                                   &S->sxx[indv], &S->syy[indv], &S->szz[indv]};
                    for (int q = 0; q < 6; q++) {
                        if (!fs && gidz - FDOH < NAB) *V[q] *= S->taper[gidz - FDOH];
                        if (gidz > NZ - NAB - FDOH - 1) *V[q] *= S->taper[NZ - FDOH - gidz - 1];
                        if (gidy - FDOH < NAB) *V[q] *= S->taper[gidy - FDOH];
                        if (gidy > NY - NAB - FDOH - 1) *V[q] *= S->taper[NY - FDOH - gidy - 1];
                        if (gidx - FDOH < NAB) *V[q] *= S->taper[gidx - FDOH];
                        if (gidx > NX - NAB - FDOH - 1) *V[q] *= S->taper[NX - FDOH - gidx - 1];
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Free surface, image method: src/surface2D.cl:30-193                       */
/* ------------------------------------------------------------------------ */
static void surface2D(sim *S, int pdir) {
    const int FDOH = S->FDOH, NAB = S->NAB, NZ = S->NZp, NX = S->NXp, LVE = S->L;
    const float *hc = S->hc, DT = S->DT, DT2 = S->DT2;
    const int cpml = S->c->abs_type == 1;
    const int gidz = FDOH;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++) {
        long indp = (long)(gidx - FDOH) * (NZ - 2 * FDOH) + (gidz - FDOH);
        long iv = (long)gidx * NZ + gidz;
        S->szz[iv] = 0.0;
        for (int l = 0; l < LVE; l++) S->rzz[(long)l * NX * NZ + iv] = 0.0;
        for (int m = 1; m <= FDOH; m++) {
            S->szz[iv - m] = -S->szz[iv + m];
            S->sxz[iv - m] = -S->sxz[iv + m - 1];
        }
        float vxx = Dm(S->vx, iv, NZ, hc, FDOH);
        float vzz = Dm(S->vz, iv, 1, hc, FDOH);
        if (cpml) {
            int i, k, ind;
            if (gidx - FDOH < NAB) {
                i = gidx - FDOH;
                k = gidz - FDOH;
                vxx = vxx / S->K_x[i] + S->psi_vx_x[(long)i * (NZ - 2 * FDOH) + k];
            }
            if (gidx > NX - NAB - FDOH - 1) {
                i = gidx - NX + NAB + FDOH + NAB;
                k = gidz - FDOH;
                ind = 2 * NAB - 1 - i;
                vxx = vxx / S->K_x[ind + 1] + S->psi_vx_x[(long)i * (NZ - 2 * FDOH) + k];
            }
        }
        if (LVE == 0) {
            float f = S->mu[indp] * 2.0;
            float g = S->M[indp];
            float h = -((g - f) * (g - f) * (vxx) / g) - ((g - f) * vzz);
            S->sxx[iv] += pdir * h;
        } else {
            float f = S->mu[indp] * 2.0 * (1.0 + LVE * S->taus[indp]);
            float g = S->M[indp] * (1.0 + LVE * S->taup[indp]);
            float h = -((g - f) * (g - f) * (vxx) / g) - ((g - f) * vzz);
            float sump = 0;
            for (int l = 0; l < LVE; l++) sump += S->rxx[(long)l * NX * NZ + iv];
            h += -DT2 * sump;
            S->sxx[iv] += pdir * h;
            float d = 2.0 * S->mu[indp] * S->taus[indp] / DT;
            float e = S->M[indp] * S->taup[indp] / DT;
            sump = 0;
            for (int l = 0; l < LVE; l++) {
                float b = S->eta[l] / (1.0 + (S->eta[l] * 0.5));
                h = b * (((d - e) * ((f / g) - 1.0) * vxx) - ((d - e) * vzz));
                S->rxx[(long)l * NX * NZ + iv] += pdir * h;
                sump += S->rxx[(long)l * NX * NZ + iv];
            }
            S->sxx[iv] += pdir * (DT2 * sump);
        }
    }
}

/* Free surface 3D: src/surface3D.cl:39-271 */
static void surface3D(sim *S, int pdir) {
    const int FDOH = S->FDOH, NAB = S->NAB, NZ = S->NZp, NY = S->NYp, NX = S->NXp,
              LVE = S->L;
    const float *hc = S->hc, DT = S->DT, DT2 = S->DT2;
    const int cpml = S->c->abs_type == 1;
    const long sy = NZ, sx = (long)NY * NZ, NV = (long)NX * NY * NZ;
    const long nzi = NZ - 2 * FDOH, nyi = NY - 2 * FDOH;
    const int gidz = FDOH;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++) {
        for (int gidy = FDOH; gidy < NY - FDOH; gidy++) {
            long indp = (long)(gidx - FDOH) * nyi * nzi + (long)(gidy - FDOH) * nzi +
                        (gidz - FDOH);
            long iv = (long)gidx * sx + (long)gidy * sy + gidz;
            S->szz[iv] = 0.0;
            for (int l = 0; l < LVE; l++) S->rzz[l * NV + iv] = 0.0;
            for (int m = 1; m <= FDOH; m++) {
                S->szz[iv - m] = -S->szz[iv + m];
                S->sxz[iv - m] = -S->sxz[iv + m - 1];
                S->syz[iv - m] = -S->syz[iv + m - 1];
            }
            float vxx = Dm(S->vx, iv, sx, hc, FDOH);
            float vyy = Dm(S->vy, iv, sy, hc, FDOH);
            float vzz = Dm(S->vz, iv, 1, hc, FDOH);
            if (cpml) {
                int i, j, k, ind;
                if (gidy - FDOH < NAB) {
                    i = gidx - FDOH;
                    j = gidy - FDOH;
                    k = gidz - FDOH;
                    vyy = vyy / S->K_y[j] + PY(S->psi_vy_y);
                } else if (gidy > NY - NAB - FDOH - 1) {
                    i = gidx - FDOH;
                    j = gidy - NY + NAB + FDOH + NAB;
                    k = gidz - FDOH;
                    ind = 2 * NAB - 1 - j;
                    vyy = vyy / S->K_y[ind + 1] + PY(S->psi_vy_y);
                }
                if (gidx - FDOH < NAB) {
                    i = gidx - FDOH;
                    j = gidy - FDOH;
                    k = gidz - FDOH;
                    vxx = vxx / S->K_x[i] + PX(S->psi_vx_x);
                }
                if (gidx > NX - NAB - FDOH - 1) {
                    i = gidx - NX + NAB + FDOH + NAB;
                    j = gidy - FDOH;
                    k = gidz - FDOH;
                    ind = 2 * NAB - 1 - i;
                    vxx = vxx / S->K_x[ind + 1] + PX(S->psi_vx_x);
                }
            }
            if (LVE == 0) {
                float f = S->mu[indp] * 2.0f;
                float g = S->M[indp];
                float h = -((g - f) * (g - f) * (vxx + vyy) / g) - ((g - f) * vzz);
                S->sxx[iv] += h;
                S->syy[iv] += h;
            } else {
                float f = S->mu[indp] * 2.0 * (1.0 + LVE * S->taus[indp]);
                float g = S->M[indp] * (1.0 + LVE * S->taup[indp]);
                float h = -((g - f) * (g - f) * (vxx + vyy) / g) - ((g - f) * vzz);
                float sumx = 0, sumy = 0;
                for (int l = 0; l < LVE; l++) {
                    sumx += S->rxx[l * NV + iv];
                    sumy += S->ryy[l * NV + iv];
                }
                S->sxx[iv] += pdir * (h - DT2 * sumx);
                S->syy[iv] += pdir * (h - DT2 * sumy);
                float d = 2.0 * S->mu[indp] * S->taus[indp] / DT;
                float e = S->M[indp] * S->taup[indp] / DT;
                sumx = 0;
                sumy = 0;
                for (int l = 0; l < LVE; l++) {
                    float b = S->eta[l] / (1.0 + (S->eta[l] * 0.5));
                    h = b * (((d - e) * ((f / g) - 1.0) * (vxx + vyy)) - ((d - e) * vzz));
                    S->rxx[l * NV + iv] += h;
                    S->ryy[l * NV + iv] += h;
                    sumx += S->rxx[l * NV + iv];
                    sumy += S->ryy[l * NV + iv];
                }
                S->sxx[iv] += (DT2 * sumx);
                S->syy[iv] += (DT2 * sumy);
            }
        }
    }
}
#undef PX
#undef PY
#undef PZ

/* ------------------------------------------------------------------------ */
/* Sources: generated kernel of src/automatic_kernels.c:394-629.             */
/* Each source is dispatched on its own type code (the reference compares    */
/* against the variable index when several types are mixed, :562-585).      */
/* ------------------------------------------------------------------------ */
static long cell_of(const sim *S, const float *pos, int *i_out, int *k_out) {
    int i = (int)(pos[0] / S->DH) + S->FDOH;
    int j = (int)(pos[1] / S->DH) + S->FDOH;
    int k = (int)(pos[2] / S->DH) + S->FDOH;
    *i_out = i;
    *k_out = k;
    if (S->ND == 3) return (long)i * S->NYp * S->NZp + (long)j * S->NZp + k;
    return (long)i * S->NZp + k;
}

static int check_positions(const sim *S, int n, const float *pos, int stride,
                           char *err, int errlen) {
    for (int s = 0; s < n; s++) {
        int i = (int)(pos[s * stride + 0] / S->DH) + S->FDOH;
        int j = (int)(pos[s * stride + 1] / S->DH) + S->FDOH;
        int k = (int)(pos[s * stride + 2] / S->DH) + S->FDOH;
        int bad = k < 0 || k >= S->NZp;
        if (S->ND == 3) bad |= j < 0 || j >= S->NYp;
        if (bad) {
            snprintf(err, errlen, "position %d lies outside the padded grid", s);
            return 1;
        }
        (void)i;
    }
    return 0;
}

static void sources(sim *S, int nt, int pdir, int nsrc, const float *src_pos,
                    const float *src) {
    const int NT = S->c->NT;
    for (int s = 0; s < nsrc; s++) {
        int i, k;
        long pos = cell_of(S, &src_pos[5 * s], &i, &k);
        if (i < S->FDOH || i > S->NXp - S->FDOH - 1) continue;
        if (S->c->freesurf == 1 && pdir == -1 && k < 2 * S->FDOH) continue;
        int type = (int)src_pos[4 + 5 * s];
        float amp = (float)pdir * (S->DT * src[s * NT + nt]);
        if (type == 0 && S->vx) S->vx[pos] += amp;
        else if (type == 1 && S->vy) S->vy[pos] += amp;
        else if (type == 2 && S->vz) S->vz[pos] += amp;
        else if ((type == 3 || type == 100) && S->ND == 22) S->p[pos] += amp;
        else if (type == 3 || type == 100) {
            /* generated text: "+=amp/%f" with (float)ND, a double literal */
            double nd = (double)(S->ND == 3 ? 3 : 2);
            S->sxx[pos] = (float)((double)S->sxx[pos] + (double)amp / nd);
            S->szz[pos] = (float)((double)S->szz[pos] + (double)amp / nd);
            if (S->ND == 3) S->syy[pos] = (float)((double)S->syy[pos] + (double)amp / nd);
        }
    }
}

/* Seismogram sampling: generated kernel of src/automatic_kernels.c:22-219 */
static float *named_var(sim *S, const char *n) {
    if (!strcmp(n, "vx")) return S->vx;
    if (!strcmp(n, "vy")) return S->vy;
    if (!strcmp(n, "vz")) return S->vz;
    if (!strcmp(n, "sxx")) return S->sxx;
    if (!strcmp(n, "syy")) return S->syy;
    if (!strcmp(n, "szz")) return S->szz;
    if (!strcmp(n, "sxy")) return S->sxy;
    if (!strcmp(n, "syz")) return S->syz;
    if (!strcmp(n, "sxz")) return S->sxz;
    if (!strcmp(n, "p")) return S->p;
    return NULL;
}

static void varsout(sim *S, int nt, int nrec, const float *rec_pos, float *out) {
    const int NT = S->c->NT, nout = or_nout(S->c);
    for (int r = 0; r < nrec; r++) {
        int i, k;
        long pos = cell_of(S, &rec_pos[8 * r], &i, &k);
        if (i < S->FDOH || i > S->NXp - S->FDOH - 1) continue;
        for (int o = 0; o < nout; o++) {
            const char *name = or_outname(S->c, o);
            float v;
            if (!strcmp(name, "p") && S->ND != 22) {
                /* (sxx+szz[+syy])/ND with "%f"-printed ND, a double literal */
                float s = S->sxx[pos] + S->szz[pos];
                if (S->ND == 3) s = s + S->syy[pos];
                v = (float)((double)s / (double)(S->ND == 3 ? 3 : 2));
            } else {
                v = named_var(S, name)[pos];
            }
            out[((long)o * nrec + r) * NT + nt] = v;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Forward driver: src/time_stepping.c:747-835 (one shot)                    */
/* ------------------------------------------------------------------------ */
static int sim_init(sim *S, const or_csts *c, char *err, int errlen) {
    memset(S, 0, sizeof(*S));
    S->c = c;
    S->ND = c->ND;
    S->FDOH = c->FDORDER / 2;
    S->NAB = c->nab;
    S->L = c->L;
    if (c->ND != 2 && c->ND != 3 && c->ND != 22) {
        snprintf(err, errlen, "ND=%d is not supported", c->ND);
        return 1;
    }
    if (c->L < 0 || c->L > OR_MAXL) {
        snprintf(err, errlen, "L must be in [0, %d]", OR_MAXL);
        return 1;
    }
    if (c->nab > 120) {
        snprintf(err, errlen, "nab too large");
        return 1;
    }
    int NZ, NY, NX;
    dims3(c, &NZ, &NY, &NX);
    S->NZ = NZ;
    S->NY = NY;
    S->NX = NX;
    S->NZp = NZ + 2 * S->FDOH;
    S->NXp = NX + 2 * S->FDOH;
    S->NYp = c->ND == 3 ? NY + 2 * S->FDOH : 1;
    S->nv = (long)S->NZp * S->NYp * S->NXp;
    S->np = (long)NZ * NY * NX;
    double hc[7];
    if (or_holberg(c->FDORDER, c->MAXRELERROR, hc)) {
        snprintf(err, errlen, "invalid FDORDER/MAXRELERROR");
        return 1;
    }
    or_kernel_hc(c->FDORDER, c->MAXRELERROR, S->hc);
    S->DT = or_literal9((double)c->dt);
    S->DT2 = or_literal9((double)c->dt / 2.0);
    S->DH = or_literal9((double)c->dh);
    or_eta(c->L, c->FL, c->dt, S->eta);
    if (c->abs_type == 1) {
        or_cpml_coeff(c->NPOWER, c->K_MAX_CPML, c->FPML, c->VPPML, c->dh, c->dt,
                      c->nab, S->K_x, S->b_x, S->a_x, S->K_x_half, S->b_x_half,
                      S->a_x_half);
        memcpy(S->K_y, S->K_x, sizeof(S->K_x));
        memcpy(S->a_y, S->a_x, sizeof(S->a_x));
        memcpy(S->b_y, S->b_x, sizeof(S->b_x));
        memcpy(S->K_y_half, S->K_x_half, sizeof(S->K_x));
        memcpy(S->a_y_half, S->a_x_half, sizeof(S->a_x));
        memcpy(S->b_y_half, S->b_x_half, sizeof(S->b_x));
        memcpy(S->K_z, S->K_x, sizeof(S->K_x));
        memcpy(S->a_z, S->a_x, sizeof(S->a_x));
        memcpy(S->b_z, S->b_x, sizeof(S->b_x));
        memcpy(S->K_z_half, S->K_x_half, sizeof(S->K_x));
        memcpy(S->a_z_half, S->a_x_half, sizeof(S->a_x));
        memcpy(S->b_z_half, S->b_x_half, sizeof(S->b_x));
    } else if (c->abs_type == 2) {
        or_taper(c->nab, c->abpc, S->taper);
    }
    long nv = S->nv, L = c->L;
    S->vx = zalloc(nv);
    S->vz = zalloc(nv);
    if (c->ND == 3) S->vy = zalloc(nv);
    if (c->ND == 22) {
        S->p = zalloc(nv);
        if (L) S->rp = zalloc(nv * L);
    } else {
        S->sxx = zalloc(nv);
        S->szz = zalloc(nv);
        S->sxz = zalloc(nv);
        if (L) {
            S->rxx = zalloc(nv * L);
            S->rzz = zalloc(nv * L);
            S->rxz = zalloc(nv * L);
        }
        if (c->ND == 3) {
            S->syy = zalloc(nv);
            S->sxy = zalloc(nv);
            S->syz = zalloc(nv);
            if (L) {
                S->ryy = zalloc(nv * L);
                S->rxy = zalloc(nv * L);
                S->ryz = zalloc(nv * L);
            }
        }
    }
    long np = S->np;
    S->M = zalloc(np);
    S->rho = zalloc(np);
    S->rip = zalloc(np);
    S->rkp = zalloc(np);
    if (c->ND == 3) S->rjp = zalloc(np);
    if (c->ND != 22) {
        S->mu = zalloc(np);
        S->muipkp = zalloc(np);
        if (c->ND == 3) {
            S->muipjp = zalloc(np);
            S->mujpkp = zalloc(np);
        }
    }
    if (L) {
        S->taup = zalloc(np);
        if (c->ND != 22) {
            S->taus = zalloc(np);
            S->tausipkp = zalloc(np);
            if (c->ND == 3) {
                S->tausipjp = zalloc(np);
                S->tausjpkp = zalloc(np);
            }
        }
    }
    if (c->abs_type == 1) {
        long nab2 = 2 * c->nab;
        long nzi = NZ, nyi = (c->ND == 3) ? NY : 1, nxi = NX;
        long sxs = nab2 * nyi * nzi, sys = nxi * nab2 * nzi, szs = nxi * nyi * nab2;
        if (c->ND == 22) {
            S->psi_p_x = zalloc(sxs);
            S->psi_p_z = zalloc(szs);
            S->psi_vx_x = zalloc(sxs);
            S->psi_vz_z = zalloc(szs);
        } else {
            S->psi_sxx_x = zalloc(sxs);
            S->psi_sxz_x = zalloc(sxs);
            S->psi_szz_z = zalloc(szs);
            S->psi_sxz_z = zalloc(szs);
            S->psi_vx_x = zalloc(sxs);
            S->psi_vz_x = zalloc(sxs);
            S->psi_vx_z = zalloc(szs);
            S->psi_vz_z = zalloc(szs);
            if (c->ND == 3) {
                S->psi_sxy_x = zalloc(sxs);
                S->psi_syz_z = zalloc(szs);
                S->psi_vy_x = zalloc(sxs);
                S->psi_vy_z = zalloc(szs);
                S->psi_syy_y = zalloc(sys);
                S->psi_sxy_y = zalloc(sys);
                S->psi_syz_y = zalloc(sys);
                S->psi_vx_y = zalloc(sys);
                S->psi_vy_y = zalloc(sys);
                S->psi_vz_y = zalloc(sys);
            }
        }
    }
    return 0;
}

static void sim_free(sim *S) {
    float **ptrs[] = {&S->vx, &S->vy, &S->vz, &S->sxx, &S->syy, &S->szz, &S->sxy,
                      &S->syz, &S->sxz, &S->p, &S->rxx, &S->ryy, &S->rzz, &S->rxy,
                      &S->ryz, &S->rxz, &S->rp, &S->M, &S->mu, &S->rho, &S->rip,
                      &S->rjp, &S->rkp, &S->muipkp, &S->muipjp, &S->mujpkp,
                      &S->taup, &S->taus, &S->tausipkp, &S->tausipjp, &S->tausjpkp,
                      &S->psi_sxx_x, &S->psi_sxy_x, &S->psi_sxz_x, &S->psi_sxy_y,
                      &S->psi_syy_y, &S->psi_syz_y, &S->psi_sxz_z, &S->psi_syz_z,
                      &S->psi_szz_z, &S->psi_vx_x, &S->psi_vy_x, &S->psi_vz_x,
                      &S->psi_vx_y, &S->psi_vy_y, &S->psi_vz_y, &S->psi_vx_z,
                      &S->psi_vy_z, &S->psi_vz_z, &S->psi_p_x, &S->psi_p_z};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) {
        free(*ptrs[i]);
        *ptrs[i] = NULL;
    }
}

static double g_loop_seconds = 0.0;
double or_last_loop_seconds(void) { return g_loop_seconds; }

static double wall_seconds(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

int or_forward(const or_csts *c, const float *p0, const float *p1,
               const float *rho, const float *taup, const float *taus,
               int nsrc, const float *src_pos, const float *src, int nrec,
               const float *rec_pos, float *out, float *fields_out, char *err,
               int errlen) {
    return or_forward_from(c, p0, p1, rho, taup, taus, nsrc, src_pos, src, nrec, rec_pos,
                           NULL, out, fields_out, err, errlen);
}

/* The same loop started from given wavefields (or_nvars padded fields in   */
/* or_varname order) instead of zeros; memory variables and CPML memory     */
/* start at zero.  Test use: a 3D run on fields copied along y from a 2D    */
/* state must reproduce the 2D run away from the y edges.                  */
int or_forward_from(const or_csts *c, const float *p0, const float *p1,
                    const float *rho, const float *taup, const float *taus,
                    int nsrc, const float *src_pos, const float *src, int nrec,
                    const float *rec_pos, const float *fields_in, float *out,
                    float *fields_out, char *err, int errlen) {
    sim S;
    char dummy[8];
    if (!err) {
        err = dummy;
        errlen = sizeof(dummy);
    }
    if (sim_init(&S, c, err, errlen)) {
        sim_free(&S);
        return 1;
    }
#ifdef _OPENMP
    if (c->nthreads > 0) omp_set_num_threads(c->nthreads);
#endif
    if (or_nout(c) == 0) {
        snprintf(err, errlen, "seisout=%d has no outputs for ND=%d", c->seisout, c->ND);
        sim_free(&S);
        return 1;
    }
    if (c->ND == 22 && c->freesurf) {
        snprintf(err, errlen, "the free surface is not available for ND=22");
        sim_free(&S);
        return 1;
    }
    or_transform(c, p0, p1, rho, taup, taus, S.M, S.mu, S.rho, S.rip, S.rjp, S.rkp,
                 S.muipkp, S.muipjp, S.mujpkp, S.taup, S.taus, S.tausipkp,
                 S.tausipjp, S.tausjpkp);
    float dts;
    if (or_check_stability(c, S.M, S.mu, S.rho, &dts)) {
        snprintf(err, errlen, "Time step too large, to be stable, set dt<%f", dts);
        sim_free(&S);
        return 1;
    }
    if (check_positions(&S, nsrc, src_pos, 5, err, errlen) ||
        check_positions(&S, nrec, rec_pos, 8, err, errlen)) {
        sim_free(&S);
        return 1;
    }
    int tmax = c->tmax > 0 ? c->tmax : c->NT;
    if (tmax > c->NT || c->tmin < 0 || c->tmin >= tmax) {
        snprintf(err, errlen, "time window [tmin=%d, tmax=%d) outside [0, NT=%d]", c->tmin,
                 tmax, c->NT);
        sim_free(&S);
        return 1;
    }
    memset(out, 0, sizeof(float) * (size_t)or_nout(c) * nrec * c->NT);
    if (fields_in) {
        for (int v = 0; v < or_nvars(c); v++)
            memcpy(named_var(&S, or_varname(c, v)), &fields_in[v * S.nv],
                   S.nv * sizeof(float));
    }
    double t_start = wall_seconds();
    for (int t = 0; t < tmax; t++) {
        sources(&S, t, 1, nsrc, src_pos, src);
        if (c->ND == 3) {
            update_v3D(&S);
            update_s3D(&S);
        } else if (c->ND == 2) {
            update_v2D(&S);
            update_s2D(&S);
        } else {
            update_v2D_acc(&S);
            update_s2D_acc(&S);
        }
        if (c->freesurf) {
            if (c->ND == 3) surface3D(&S, 1);
            else surface2D(&S, 1);
        }
        varsout(&S, t, nrec, rec_pos, out);
    }
    g_loop_seconds = wall_seconds() - t_start;
    if (fields_out) {
        for (int v = 0; v < or_nvars(c); v++)
            memcpy(&fields_out[v * S.nv], named_var(&S, or_varname(c, v)),
                   S.nv * sizeof(float));
    }
    sim_free(&S);
    return 0;
}

/* ======================================================================== */
/* Gradient by backpropagation (BACK_PROP_TYPE=1), elastic 2D and 3D.        */
/* ======================================================================== */

/* Boundary shell saved during the forward run and re-injected while the    */
/* forward wavefield is run backwards: an FDOH-thick shell on the inner face */
/* of the absorbing layer.  One device, FP32 storage.                        */
/* savebnd: src/savebnd2D.cl:43-223, src/savebnd3D.cl:38-284;               */
/* inject_ind: src/header_injectbnd.cl:25-232 (3D), :238-398 (2D);          */
/* NBND: src/Init_OpenCL.c:988-1038.  Coordinates are padded.               */
typedef struct {
    int NXbnd, NYbnd, NZbnd, lbnd, lbnds;
    long n;
} shell_t;

static void shell_init(const sim *S, shell_t *B) {
    const int F = S->FDOH, NAB = S->NAB;
    B->NXbnd = S->NXp - 2 * F - 2 * NAB;
    B->NYbnd = S->ND == 3 ? S->NYp - 2 * F - 2 * NAB : 0;
    B->NZbnd = S->c->freesurf ? S->NZp - 2 * F - NAB : S->NZp - 2 * F - 2 * NAB;
    B->lbnd = F + NAB;
    B->lbnds = S->c->freesurf ? F : F + NAB;
    if (S->ND == 3)
        B->n = (long)B->NYbnd * B->NZbnd * 2 * F + (long)B->NZbnd * (B->NXbnd - 2 * F) * 2 * F +
               (long)(B->NYbnd - 2 * F) * (B->NXbnd - 2 * F) * 2 * F;
    else
        B->n = (long)B->NZbnd * F * 2 + (long)(B->NXbnd - 2 * F) * 2 * F;
}

/* padded cell of shell entry gid (the savebnd enumeration) */
static long shell_cell(const sim *S, const shell_t *B, long gid) {
    const int F = S->FDOH;
    const long NXb = B->NXbnd, NYb = B->NYbnd, NZb = B->NZbnd;
    long i, j = 0, k, g;
    if (S->ND == 3) {
        const long fb = NYb * NZb * F, lr = NZb * (NXb - 2 * F) * F,
                   du = (NYb - 2 * F) * (NXb - 2 * F) * F;
        if (gid < fb) {                                  /* front */
            g = gid;
            i = g / (NYb * NZb) + B->lbnd;
            j = (g / NZb) % NYb + B->lbnd;
            k = g % NZb + B->lbnds;
        } else if (gid < 2 * fb) {                       /* back */
            g = gid - fb;
            i = g / (NYb * NZb) + NXb + S->NAB;
            j = (g / NZb) % NYb + B->lbnd;
            k = g % NZb + B->lbnds;
        } else if (gid < 2 * fb + lr) {                  /* left */
            g = gid - 2 * fb;
            i = g / (NZb * F) + B->lbnd + F;
            j = (g / NZb) % F + B->lbnd;
            k = g % NZb + B->lbnds;
        } else if (gid < 2 * fb + 2 * lr) {              /* right */
            g = gid - 2 * fb - lr;
            i = g / (NZb * F) + B->lbnd + F;
            j = (g / NZb) % F + NYb + S->NAB;
            k = g % NZb + B->lbnds;
        } else if (gid < 2 * fb + 2 * lr + du) {         /* down */
            g = gid - 2 * fb - 2 * lr;
            i = g / (F * (NYb - 2 * F)) + B->lbnd + F;
            j = (g / F) % (NYb - 2 * F) + B->lbnd + F;
            k = g % F + NZb + B->lbnds - F;
        } else {                                         /* up */
            g = gid - 2 * fb - 2 * lr - du;
            i = g / (F * (NYb - 2 * F)) + B->lbnd + F;
            j = (g / F) % (NYb - 2 * F) + B->lbnd + F;
            k = g % F + B->lbnds;
        }
        return i * S->NYp * S->NZp + j * S->NZp + k;
    }
    const long fb = NZb * F, du = (NXb - 2 * F) * F;
    if (gid < fb) {                                      /* front */
        i = gid / NZb + B->lbnd;
        k = gid % NZb + B->lbnds;
    } else if (gid < 2 * fb) {                           /* back */
        g = gid - fb;
        i = g / NZb + NXb + S->NAB;
        k = g % NZb + B->lbnds;
    } else if (gid < 2 * fb + du) {                      /* bottom */
        g = gid - 2 * fb;
        i = g % (NXb - 2 * F) + B->lbnd + F;
        k = g / (NXb - 2 * F) + NZb + B->lbnds - F;
    } else {                                             /* up */
        g = gid - 2 * fb - du;
        i = g % (NXb - 2 * F) + B->lbnd + F;
        k = g / (NXb - 2 * F) + B->lbnds;
    }
    return i * S->NZp + k;
}

/* shell entry of padded cell (k, j, i), or -1 (inject_ind) */
static long inject_ind(const sim *S, const shell_t *B, int k, int j, int i) {
    const int F = S->FDOH;
    const long NXb = B->NXbnd, NYb = B->NYbnd, NZb = B->NZbnd;
    i -= B->lbnd;
    k -= B->lbnds;
    if (S->ND == 3) {
        j -= B->lbnd;
        if ((k > F - 1 && k < NZb - F) && (j > F - 1 && j < NYb - F) && (i > F - 1 && i < NXb - F))
            return -1;
        if (k < 0 || k > NZb - 1 || j < 0 || j > NYb - 1 || i < 0 || i > NXb - 1) return -1;
        if (i < F) return (long)i * NYb * NZb + (long)j * NZb + k;
        if (i > NXb - 1 - F) {
            i = i - NXb + F;
            return NYb * NZb * F + (long)i * NYb * NZb + (long)j * NZb + k;
        }
        if (j < F) {
            i = i - F;
            return NYb * NZb * F * 2 + (long)i * F * NZb + (long)j * NZb + k;
        }
        if (j > NYb - 1 - F) {
            i = i - F;
            j = j - NYb + F;
            return NYb * NZb * F * 2 + (NXb - 2 * F) * NZb * F + (long)i * F * NZb + (long)j * NZb + k;
        }
        if (k < F) {
            i = i - F;
            j = j - F;
            return NYb * NZb * F * 2 + (NXb - 2 * F) * NZb * F * 2 + (NXb - 2 * F) * (NYb - 2 * F) * F +
                   (long)i * (NYb - 2 * F) * F + (long)j * F + k;
        }
        i = i - F;
        j = j - F;
        k = k - NZb + F;
        return NYb * NZb * F * 2 + (NXb - 2 * F) * NZb * F * 2 + (long)i * (NYb - 2 * F) * F +
               (long)j * F + k;
    }
    if ((k > F - 1 && k < NZb - F) && (i > F - 1 && i < NXb - F)) return -1;
    if (k < 0 || k > NZb - 1 || i < 0 || i > NXb - 1) return -1;
    if (i < F) return (long)i * NZb + k;
    if (i > NXb - 1 - F) {
        i = i - NXb + F;
        return NZb * F + (long)i * NZb + k;
    }
    if (k < F) {
        i = i - F;
        return NZb * F * 2 + (NXb - 2 * F) * F + i + (long)k * (NXb - 2 * F);
    }
    i = i - F;
    k = k - NZb + F;
    return NZb * F * 2 + i + (long)k * (NXb - 2 * F);
}

/* forward variables kept on the shell, in the order of the boundary buffer */
static int shell_vars(const sim *S, float **v) {
    if (S->ND == 3) {
        float *w[9] = {S->vx, S->vy, S->vz, S->sxx, S->syy, S->szz, S->sxy, S->syz, S->sxz};
        memcpy(v, w, sizeof(w));
        return 9;
    }
    float *w[5] = {S->vx, S->vz, S->sxx, S->szz, S->sxz};
    memcpy(v, w, sizeof(w));
    return 5;
}

int or_shell_check(const or_csts *c, long *nbnd) {
    sim S;
    char err[128];
    if (sim_init(&S, c, err, sizeof(err))) {
        sim_free(&S);
        return -1;
    }
    shell_t B;
    shell_init(&S, &B);
    *nbnd = B.n;
    long bad = 0, inside = 0;
    char *seen = calloc(B.n > 0 ? B.n : 1, 1);
    for (long g = 0; g < B.n; g++) {
        long cell = shell_cell(&S, &B, g);
        int i = (int)(cell / ((long)S.NYp * S.NZp));
        int j = (int)((cell / S.NZp) % S.NYp);
        int k = (int)(cell % S.NZp);
        if (inject_ind(&S, &B, k, j, i) != g) bad++;
    }
    for (int i = 0; i < S.NXp; i++)
        for (int j = 0; j < S.NYp; j++)
            for (int k = 0; k < S.NZp; k++) {
                long m = inject_ind(&S, &B, k, j, i);
                if (m < 0) continue;
                inside++;
                if (m >= B.n || seen[m]) bad++;
                else seen[m] = 1;
            }
    free(seen);
    sim_free(&S);
    if (inside != B.n) bad++;
    return (int)bad;
}

static void save_shell(sim *S, const shell_t *B, float *dst) {
    float *v[9];
    const int nv = shell_vars(S, v);
    for (int q = 0; q < nv; q++)
        for (long g = 0; g < B->n; g++) dst[q * B->n + g] = v[q][shell_cell(S, B, g)];
}

static void inject_shell(sim *S, const shell_t *B, const float *src, int first, int count) {
    float *v[9];
    shell_vars(S, v);
    for (int q = first; q < first + count; q++)
        for (long g = 0; g < B->n; g++) v[q][shell_cell(S, B, g)] = src[q * B->n + g];
}

/* Backpropagation of the forward stresses and velocities: the forward       */
/* increment without CPML, subtracted (src/update_adjs2D.cl:205-250,        */
/* src/update_adjv2D.cl:196-210, src/update_adjs3D.cl:309-324,               */
/* src/update_adjv3D.cl:296-305).  Elastic only.                             */
static void backprop_s(sim *S) {
    const int FDOH = S->FDOH, NZ = S->NZp, NY = S->NYp, NX = S->NXp;
    const float *hc = S->hc;
    const long sy = NZ, sx = (long)NY * NZ;
    const long nzi = NZ - 2 * FDOH, nyi = NY - 2 * FDOH;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++)
        for (int gidy = (S->ND == 3 ? FDOH : 0); gidy < (S->ND == 3 ? NY - FDOH : 1); gidy++)
            for (int gidz = FDOH; gidz < NZ - FDOH; gidz++) {
                long indp = (long)(gidx - FDOH) * nyi * nzi + (long)(gidy - (S->ND == 3 ? FDOH : 0)) * nzi + (gidz - FDOH);
                long indv = (long)gidx * sx + (long)gidy * sy + gidz;
                if (S->ND == 2) {
                    indp = (long)(gidx - FDOH) * nzi + (gidz - FDOH);
                    indv = (long)gidx * NZ + gidz;
                    float vxx = Dm(S->vx, indv, NZ, hc, FDOH);
                    float vxz = Dp(S->vx, indv, 1, hc, FDOH);
                    float vzz = Dm(S->vz, indv, 1, hc, FDOH);
                    float vzx = Dp(S->vz, indv, NZ, hc, FDOH);
                    float fipkp = S->muipkp[indp];
                    float f = 2.0 * S->mu[indp];
                    float g = S->M[indp];
                    S->sxz[indv] -= (fipkp * (vxz + vzx));
                    S->sxx[indv] -= (g * (vxx + vzz)) - (f * vzz);
                    S->szz[indv] -= (g * (vxx + vzz)) - (f * vxx);
                    continue;
                }
                float vxx = Dm(S->vx, indv, sx, hc, FDOH);
                float vxy = Dp(S->vx, indv, sy, hc, FDOH);
                float vxz = Dp(S->vx, indv, 1, hc, FDOH);
                float vyx = Dp(S->vy, indv, sx, hc, FDOH);
                float vyy = Dm(S->vy, indv, sy, hc, FDOH);
                float vyz = Dp(S->vy, indv, 1, hc, FDOH);
                float vzx = Dp(S->vz, indv, sx, hc, FDOH);
                float vzy = Dp(S->vz, indv, sy, hc, FDOH);
                float vzz = Dm(S->vz, indv, 1, hc, FDOH);
                float fipjp = S->muipjp[indp], fjpkp = S->mujpkp[indp], fipkp = S->muipkp[indp];
                float g = S->M[indp];
                float f = 2.0 * S->mu[indp];
                float vxyyx = vxy + vyx, vyzzy = vyz + vzy, vxzzx = vxz + vzx;
                float vxxyyzz = vxx + vyy + vzz, vyyzz = vyy + vzz, vxxzz = vxx + vzz,
                      vxxyy = vxx + vyy;
                S->sxy[indv] -= (fipjp * vxyyx);
                S->syz[indv] -= (fjpkp * vyzzy);
                S->sxz[indv] -= (fipkp * vxzzx);
                S->sxx[indv] -= ((g * vxxyyzz) - (f * vyyzz));
                S->syy[indv] -= ((g * vxxyyzz) - (f * vxxzz));
                S->szz[indv] -= ((g * vxxyyzz) - (f * vxxyy));
            }
}

static void backprop_v(sim *S) {
    const int FDOH = S->FDOH, NZ = S->NZp, NY = S->NYp, NX = S->NXp;
    const float *hc = S->hc;
    const long sy = NZ, sx = (long)NY * NZ;
    const long nzi = NZ - 2 * FDOH, nyi = NY - 2 * FDOH;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++)
        for (int gidy = (S->ND == 3 ? FDOH : 0); gidy < (S->ND == 3 ? NY - FDOH : 1); gidy++)
            for (int gidz = FDOH; gidz < NZ - FDOH; gidz++) {
                if (S->ND == 2) {
                    long indp = (long)(gidx - FDOH) * nzi + (gidz - FDOH);
                    long indv = (long)gidx * NZ + gidz;
                    float sxx_x = Dp(S->sxx, indv, NZ, hc, FDOH);
                    float szz_z = Dp(S->szz, indv, 1, hc, FDOH);
                    float sxz_z = Dm(S->sxz, indv, 1, hc, FDOH);
                    float sxz_x = Dm(S->sxz, indv, NZ, hc, FDOH);
                    float lvx = ((sxx_x + sxz_z) * S->rip[indp]);
                    float lvz = ((szz_z + sxz_x) * S->rkp[indp]);
                    S->vx[indv] -= lvx;
                    S->vz[indv] -= lvz;
                    continue;
                }
                long indp = (long)(gidx - FDOH) * nyi * nzi + (long)(gidy - FDOH) * nzi + (gidz - FDOH);
                long indv = (long)gidx * sx + (long)gidy * sy + gidz;
                float szz_z = Dp(S->szz, indv, 1, hc, FDOH);
                float sxx_x = Dp(S->sxx, indv, sx, hc, FDOH);
                float sxz_x = Dm(S->sxz, indv, sx, hc, FDOH);
                float sxz_z = Dm(S->sxz, indv, 1, hc, FDOH);
                float syz_y = Dm(S->syz, indv, sy, hc, FDOH);
                float syz_z = Dm(S->syz, indv, 1, hc, FDOH);
                float syy_y = Dp(S->syy, indv, sy, hc, FDOH);
                float sxy_x = Dm(S->sxy, indv, sx, hc, FDOH);
                float sxy_y = Dm(S->sxy, indv, sy, hc, FDOH);
                float lvx = ((sxx_x + sxy_y + sxz_z) * S->rip[indp]);
                float lvy = ((syy_y + sxy_x + syz_z) * S->rjp[indp]);
                float lvz = ((szz_z + sxz_x + syz_y) * S->rkp[indp]);
                S->vx[indv] -= lvx;
                S->vy[indv] -= lvy;
                S->vz[indv] -= lvz;
            }
}

/* Gradient increments of one adjoint step, accumulated in double.          */
/* rho: src/update_adjv2D.cl:326, src/update_adjv3D.cl:481.                  */
/* M, mu: src/update_adjs2D.cl:441-451, src/update_adjs3D.cl:679-689.        */
static void grad_rho(const sim *S, const sim *A, double *grho) {
    const int FDOH = S->FDOH, NZ = S->NZp, NY = S->NYp, NX = S->NXp;
    const long nzi = NZ - 2 * FDOH, nyi = S->ND == 3 ? NY - 2 * FDOH : 1;
    const int y0 = S->ND == 3 ? FDOH : 0, y1 = S->ND == 3 ? NY - FDOH : 1;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++)
        for (int gidy = y0; gidy < y1; gidy++)
            for (int gidz = FDOH; gidz < NZ - FDOH; gidz++) {
                long indp = (long)(gidx - FDOH) * nyi * nzi + (long)(gidy - y0) * nzi + (gidz - FDOH);
                long indv = (long)gidx * NY * NZ + (long)gidy * NZ + gidz;
                float d;
                if (S->ND == 3)
                    d = -S->vx[indv] * A->inc[0][indv] - S->vy[indv] * A->inc[1][indv] -
                        S->vz[indv] * A->inc[2][indv];
                else
                    d = -S->vx[indv] * A->inc[0][indv] - S->vz[indv] * A->inc[1][indv];
                grho[indp] += d;
            }
}

/* lMs, lmus: the moduli of the stress-to-strain conversion (the relaxed   */
/* ones in the reference; those at f0 for viscoelastic runs with           */
/* OR_FIX_VISCO_F0).                                                       */
/* M and mu gradients from the forward stresses S and the adjoint stress  */
/* increments A->inc (src/update_adjs2D.cl:441-468, src/update_adjs3D.cl:   */
/* 678-710).  restype 1 (RTM): the M image without the moduli, no mu term. */
/* hM/hmu (Hout, may be NULL): the approximate Hessian, whose mu part       */
/* subtracts the Hessian's own M term with hfix (the 2D reference           */
/* subtracts the gradient's, :456; the 3D reference's expression does not   */
/* compile, :694-697, and c5 is taken to weigh all three squares).          */
static void grad_Mmu(const sim *S, const sim *A, const float *lMs, const float *lmus,
                     double *gM, double *gmu, double *hM, double *hmu, int hfix) {
    const int FDOH = S->FDOH, NZ = S->NZp, NY = S->NYp, NX = S->NXp;
    const long nzi = NZ - 2 * FDOH, nyi = S->ND == 3 ? NY - 2 * FDOH : 1;
    const int y0 = S->ND == 3 ? FDOH : 0, y1 = S->ND == 3 ? NY - FDOH : 1;
    const int rtm = S->c->restype == 1;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++)
        for (int gidy = y0; gidy < y1; gidy++)
            for (int gidz = FDOH; gidz < NZ - FDOH; gidz++) {
                long indp = (long)(gidx - FDOH) * nyi * nzi + (long)(gidy - y0) * nzi + (gidz - FDOH);
                long indv = (long)gidx * NY * NZ + (long)gidy * NZ + gidz;
                float lM = lMs[indp], lmu = lmus[indp];
                if (S->ND == 2) {
                    float lsxx = A->inc[0][indv], lszz = A->inc[1][indv], lsxz = A->inc[2][indv];
                    float sxx = S->sxx[indv], szz = S->szz[indv], sxz = S->sxz[indv];
                    if (rtm) {
                        float dM = (sxx + szz) * (lsxx + lszz);
                        gM[indp] += -dM;
                        if (hM) hM[indp] += (float)((sxx + szz) * (sxx + szz));
                        continue;
                    }
                    float c1 = 1.0 / ((2.0 * lM - 2.0 * lmu) * (2.0 * lM - 2.0 * lmu));
                    float c3 = 1.0 / (lmu * lmu);
                    float c5 = 0.25 * c3;
                    float dM = c1 * (sxx + szz) * (lsxx + lszz);
                    gM[indp] += -dM;
                    gmu[indp] += -c3 * (sxz * lsxz) + dM - c5 * ((sxx - szz) * (lsxx - lszz));
                    if (hM) {
                        float dMH = c1 * (sxx + szz) * (sxx + szz);
                        hM[indp] += dMH;
                        hmu[indp] += (float)(c3 * sxz * sxz - (hfix ? dMH : dM) +
                                             c5 * (sxx - szz) * (sxx - szz));
                    }
                    continue;
                }
                float lsxx = A->inc[0][indv], lsyy = A->inc[1][indv], lszz = A->inc[2][indv];
                float lsxy = A->inc[3][indv], lsyz = A->inc[4][indv], lsxz = A->inc[5][indv];
                float sxx = S->sxx[indv], syy = S->syy[indv], szz = S->szz[indv];
                float sxy = S->sxy[indv], syz = S->syz[indv], sxz = S->sxz[indv];
                if (rtm) {
                    float dM = (sxx + syy + szz) * (lsxx + lsyy + lszz);
                    gM[indp] += -dM;
                    if (hM) hM[indp] += (float)((sxx + syy + szz) * (sxx + syy + szz));
                    continue;
                }
                float c1 = 1.0 / (3.0 * lM - 4.0 * lmu) / (3.0 * lM - 4.0 * lmu);
                float c3 = 1.0 / lmu / lmu;
                float c5 = 1.0 / 6.0 * c3;
                float dM = c1 * (sxx + syy + szz) * (lsxx + lsyy + lszz);
                gM[indp] += -dM;
                gmu[indp] += -c3 * (sxz * lsxz + sxy * lsxy + syz * lsyz) + 4.0 / 3 * dM -
                             c5 * (lsxx * (2.0 * sxx - syy - szz) + lsyy * (2.0 * syy - sxx - szz) +
                                   lszz * (2.0 * szz - sxx - syy));
                if (hM) {
                    float dMH = c1 * (sxx + syy + szz) * (sxx + syy + szz);
                    hM[indp] += dMH;
                    hmu[indp] += c3 * (sxz * sxz + sxy * sxy + syz * syz) - 4.0 / 3 * dMH +
                                 c5 * ((2.0 * sxx - syy - szz) * (2.0 * sxx - syy - szz) +
                                       (2.0 * syy - sxx - szz) * (2.0 * syy - sxx - szz) +
                                       (2.0 * szz - sxx - syy) * (2.0 * szz - sxx - syy));
                }
            }
}

/* tau gradients from the forward stresses S and the parts A->tinc: minus  */
/* the forward strain (the stresses through the compliance of lMs, lmus,    */
/* as for M and mu) times the change of each increment.                      */
static void grad_tau(const sim *S, const sim *A, const float *lMs, const float *lmus,
                     double *gtp, double *gts) {
    const int FDOH = S->FDOH, NZ = S->NZp, NY = S->NYp, NX = S->NXp;
    const long nzi = NZ - 2 * FDOH, nyi = S->ND == 3 ? NY - 2 * FDOH : 1;
    const int y0 = S->ND == 3 ? FDOH : 0, y1 = S->ND == 3 ? NY - FDOH : 1;
#pragma omp parallel for schedule(static)
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++)
        for (int gidy = y0; gidy < y1; gidy++)
            for (int gidz = FDOH; gidz < NZ - FDOH; gidz++) {
                long indp = (long)(gidx - FDOH) * nyi * nzi + (long)(gidy - y0) * nzi + (gidz - FDOH);
                long indv = (long)gidx * NY * NZ + (long)gidy * NZ + gidz;
                float lM = lMs[indp], lmu = lmus[indp];
                float lam = lM - 2.0f * lmu;
                if (S->ND == 2) {
                    float sxx = S->sxx[indv], szz = S->szz[indv], sxz = S->sxz[indv];
                    float th = (sxx + szz) / (2.0f * lM - 2.0f * lmu);
                    gtp[indp] += (double)(-(th * A->tinc[0][indv]));
                    if (lmu == 0.0f) continue;
                    float exx = (sxx - lam * th) / (2.0f * lmu);
                    float ezz = (szz - lam * th) / (2.0f * lmu);
                    float gxz = sxz / lmu;
                    gts[indp] += (double)(-((exx * A->tinc[1][indv] + ezz * A->tinc[2][indv]) +
                                            gxz * A->tinc[3][indv]));
                    continue;
                }
                float sxx = S->sxx[indv], syy = S->syy[indv], szz = S->szz[indv];
                float sxy = S->sxy[indv], syz = S->syz[indv], sxz = S->sxz[indv];
                float th = ((sxx + syy) + szz) / (3.0f * lM - 4.0f * lmu);
                gtp[indp] += (double)(-(th * A->tinc[0][indv]));
                if (lmu == 0.0f) continue;
                float exx = (sxx - lam * th) / (2.0f * lmu);
                float eyy = (syy - lam * th) / (2.0f * lmu);
                float ezz = (szz - lam * th) / (2.0f * lmu);
                float n3 = (exx * A->tinc[1][indv] + eyy * A->tinc[2][indv]) + ezz * A->tinc[3][indv];
                float s3 = ((sxy / lmu) * A->tinc[4][indv] + (syz / lmu) * A->tinc[5][indv]) +
                           (sxz / lmu) * A->tinc[6][indv];
                gts[indp] += (double)(-(n3 + s3));
            }
}

/* Adjoint of the image-method free surface, applied to the adjoint fields */
/* at the start of each adjoint step (src/surface2D_adj.cl:35-148,          */
/* src/surface3D_adj.cl:29-180), elastic.  The reference scatters into the  */
/* neighbouring columns' velocities; run column by column in order, as     */
/* here, that is deterministic.  The 3D kernel reads sxxr where syyr is     */
/* meant (src/surface3D_adj.cl:69); syy_fix reads syyr.                      */
static void surface_adj2D(sim *A) {
    const int FDOH = A->FDOH, NZ = A->NZp, NX = A->NXp;
    const float *hc = &A->hc[1];          /* {HC1..HC6} */
    const int gidz = FDOH;
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++) {
        long indp = (long)(gidx - FDOH) * (NZ - 2 * FDOH) + (gidz - FDOH);
        long iv = (long)gidx * NZ + gidz;
        float lszz = A->szz[iv], lsxx = A->sxx[iv];
        float lmu = A->mu[indp], lM = A->M[indp];
        {
            float temp1 = lsxx, temp2 = lszz;
            float a = (float)1.0 / ((float)4.0 * lM * lmu - (float)4.0 * lmu * lmu);
            lsxx = -a * (lM * temp1 + (-lM + (float)2.0 * lmu) * temp2);
            lszz = -a * (lM * temp2 + (-lM + (float)2.0 * lmu) * temp1);
        }
        lszz = 0.0;
        float hx = -((lM - lmu * (float)2.0) * (lM - lmu * (float)2.0) / lM);
        float hz = -(lM - lmu * (float)2.0);
        for (int m = 0; m < FDOH; m++) {
            if ((gidx - FDOH + m) < (NX - 2 * FDOH)) {
                long ip = (long)(gidx + m - FDOH) * (NZ - 2 * FDOH) + (gidz - FDOH);
                A->vx[iv + (long)m * NZ] += (hc[m] * hx * lsxx * A->rip[ip]);
            }
            if ((gidx - (m + 1)) > (FDOH - 1)) {
                long ip = (long)(gidx - (m + 1) - FDOH) * (NZ - 2 * FDOH) + (gidz - FDOH);
                A->vx[iv - (long)(m + 1) * NZ] += (-hc[m] * hx * lsxx * A->rip[ip]);
            }
            A->vz[iv + m] += (hc[m] * hz * lsxx * A->rkp[indp + m]);
        }
        A->szz[iv] = -(lM * lszz + (lM - (float)2.0 * lmu) * lsxx);
        A->sxx[iv] = -(lM * lsxx + (lM - (float)2.0 * lmu) * lszz);
    }
}

static void surface_adj3D(sim *A, int syy_fix) {
    const int FDOH = A->FDOH, NZ = A->NZp, NY = A->NYp, NX = A->NXp;
    const float *hc = &A->hc[1];
    const long nzi = NZ - 2 * FDOH, nyi = NY - 2 * FDOH;
    const int gidz = FDOH;
    for (int gidx = FDOH; gidx < NX - FDOH; gidx++)
        for (int gidy = FDOH; gidy < NY - FDOH; gidy++) {
#define IV(z, y, x) ((long)(x) * NY * NZ + (long)(y) * NZ + (z))
#define IP(z, y, x) ((long)((x) - FDOH) * nyi * nzi + (long)((y) - FDOH) * nzi + ((z) - FDOH))
            long iv = IV(gidz, gidy, gidx);
            float lszz = -A->szz[iv];
            float lsxx = -A->sxx[iv];
            float lsyy = syy_fix ? -A->syy[iv] : -A->sxx[iv];
            float lmu = A->mu[IP(gidz, gidy, gidx)], lM = A->M[IP(gidz, gidy, gidx)];
            {
                float temp1 = lsxx, temp2 = lsyy, temp3 = lszz;
                float a = (float)1.0 / ((float)6.0 * lM * lmu - (float)8.0 * lmu * lmu);
                lsxx = a * ((float)2.0 * (lM - lmu) * temp1 + (-lM + (float)2.0 * lmu) * (temp2 + temp3));
                lsyy = a * ((float)2.0 * (lM - lmu) * temp2 + (-lM + (float)2.0 * lmu) * (temp1 + temp3));
                lszz = a * ((float)2.0 * (lM - lmu) * temp3 + (-lM + (float)2.0 * lmu) * (temp1 + temp2));
            }
            lszz = 0.0;
            float h = -((lM - lmu * (float)2.0) * (lM - lmu * (float)2.0) / lM);
            float hz = -(lM - lmu * (float)2.0);
            for (int m = 0; m < FDOH; m++) {
                if ((gidx - FDOH + m) < (NX - 2 * FDOH))
                    A->vx[IV(gidz, gidy, gidx + m)] +=
                        (hc[m] * h * (lsxx + lsyy) * A->rip[IP(gidz, gidy, gidx + m)]);
                if ((gidx - (m + 1)) > (FDOH - 1))
                    A->vx[IV(gidz, gidy, gidx - (m + 1))] +=
                        (-hc[m] * h * (lsxx + lsyy) * A->rip[IP(gidz, gidy, gidx - (m + 1))]);
                if ((gidy - FDOH + m) < (NY - 2 * FDOH))
                    A->vy[IV(gidz, gidy + m, gidx)] +=
                        (hc[m] * h * (lsxx + lsyy) * A->rjp[IP(gidz, gidy + m, gidx)]);
                if ((gidy - (m + 1)) > (FDOH - 1))
                    A->vy[IV(gidz, gidy - (m + 1), gidx)] +=
                        (-hc[m] * h * (lsxx + lsyy) * A->rjp[IP(gidz, gidy - (m + 1), gidx)]);
                A->vz[IV(gidz + m, gidy, gidx)] +=
                    (hc[m] * hz * (lsxx + lsyy) * A->rkp[IP(gidz + m, gidy, gidx)]);
            }
            A->sxx[iv] = -(lM * lsxx + (lM - (float)2.0 * lmu) * (lsyy + lszz));
            A->syy[iv] = -(lM * lsyy + (lM - (float)2.0 * lmu) * (lsxx + lszz));
            A->szz[iv] = -(lM * lszz + (lM - (float)2.0 * lmu) * (lsxx + lsyy));
#undef IV
#undef IP
        }
}

/* Adjoint source injection: generated kernel of                           */
/* src/automatic_kernels.c:631-836; p spreads res/ND over the normal        */
/* stresses (integer literal ND, float division).                           */
static void residuals_inject(sim *A, int nt, int nrec, const float *rec_pos, const float *res) {
    const int NT = A->c->NT, nout = or_nout(A->c);
    for (int r = 0; r < nrec; r++) {
        int i, k;
        long pos = cell_of(A, &rec_pos[8 * r], &i, &k);
        if (i < A->FDOH || i > A->NXp - A->FDOH - 1) continue;
        for (int o = 0; o < nout; o++) {
            const char *name = or_outname(A->c, o);
            float v = res[((long)o * nrec + r) * NT + nt];
            if (!strcmp(name, "p") && A->ND != 22) {
                float q = v / (A->ND == 3 ? 3 : 2);
                A->sxx[pos] += q;
                A->szz[pos] += q;
                if (A->ND == 3) A->syy[pos] += q;
            } else {
                named_var(A, name)[pos] += v;
            }
        }
    }
}

/* Butterworth band-pass (src/butterworth.c:24-86): each trace (t < tmax)  */
/* is extended to nfft = the smallest 2,3,5-smooth length >= tmax by        */
/* repeating its start, transformed, multiplied by                           */
/* H(f) = (1 - sqrt(1/(1+(f/fcl)^2n))) sqrt(1/(1+(f/fch)^2n)) and            */
/* transformed back.  Restated with a direct DFT in double (O(nfft^2)).     */
void or_butterworth(float *data, float fcl, float fch, float dt, int NT, int tmax, int ntrace,
                    int order) {
    int nfft = tmax < 1 ? 1 : tmax;
    for (;; nfft++) {
        int k = nfft;
        while (k % 2 == 0) k /= 2;
        while (k % 3 == 0) k /= 3;
        while (k % 5 == 0) k /= 5;
        if (k == 1) break;
    }
    float *H = malloc(sizeof(float) * (nfft / 2 + 1));
    for (int t = 0; t < nfft / 2 + 1; t++) {
        float freq = t / (dt * nfft);
        H[t] = 1.;
        if (fcl > 0) H[t] *= 1. - sqrt(1. / (1. + pow(freq / fcl, 2 * order)));
        if (fch > 0) H[t] *= sqrt(1. / (1. + pow(freq / fch, 2 * order)));
    }
    double *cs = malloc(sizeof(double) * nfft), *sn = malloc(sizeof(double) * nfft);
    for (int k = 0; k < nfft; k++) {
        cs[k] = cos(2.0 * M_PI * k / nfft);
        sn[k] = sin(2.0 * M_PI * k / nfft);
    }
    double *x = malloc(sizeof(double) * nfft), *Xr = malloc(sizeof(double) * nfft),
           *Xi = malloc(sizeof(double) * nfft);
    for (int n = 0; n < ntrace; n++) {
        float *d = data + (long)n * NT;
        for (int t = 0; t < nfft; t++) x[t] = t < tmax ? d[t] : d[t - tmax];
        for (int k = 0; k < nfft; k++) {
            double re = 0, im = 0;
            for (int t = 0; t < nfft; t++) {
                int q = (int)(((long)k * t) % nfft);
                re += x[t] * cs[q];
                im -= x[t] * sn[q];
            }
            double h = H[k <= nfft / 2 ? k : nfft - k];
            Xr[k] = re * h;
            Xi[k] = im * h;
        }
        for (int t = 0; t < tmax; t++) {
            double re = 0;
            for (int k = 0; k < nfft; k++) {
                int q = (int)(((long)k * t) % nfft);
                re += Xr[k] * cs[q] - Xi[k] * sn[q];
            }
            d[t] = (float)(re / nfft);
        }
    }
    free(H); free(cs); free(sn); free(x); free(Xr); free(Xi);
}

/* var_res_raw (src/residuals.c:58-450, back_prop_type 1) and rtm_res       */
/* (:24-55): band-pass (fmin/fmax), per-trace rms scalings over [tmin,     */
/* tmax) in the reference's float accumulators, then residuals and rms for */
/* t < tmax.  Arrays [nout][nrec][NT].                                       */
void or_residuals(const or_csts *c, int nout, int nrec, const float *out, const float *din,
                  float *res, double *rms, int gradfix) {
    const int NT = c->NT, tmax = c->tmax > 0 ? c->tmax : c->NT, tmin = c->tmin;
    const long n = (long)nout * nrec * NT;
    if (c->restype == 1) {
        for (long q = 0; q < n; q++)
            if (q % NT < tmax) res[q] = din[q];
        return;
    }
    float *vo = malloc(sizeof(float) * (n ? n : 1)), *vi = malloc(sizeof(float) * (n ? n : 1));
    memcpy(vo, out, sizeof(float) * n);
    memcpy(vi, din, sizeof(float) * n);
    if (c->fmax > 0 || c->fmin > 0)
        for (int o = 0; o < nout; o++) {
            or_butterworth(vo + (long)o * nrec * NT, c->fmin, c->fmax, c->dt, NT, tmax, nrec, 6);
            or_butterworth(vi + (long)o * nrec * NT, c->fmin, c->fmax, c->dt, NT, tmax, nrec, 6);
        }
#define AT(v, o, g, t) v[((long)(o) * nrec + (g)) * NT + (t)]
    const int scaled = c->scalerms || c->scaleshot || c->scalermsnorm;
    const int norm = c->scalerms || c->scalermsnorm;
    float *rs = malloc(sizeof(float) * (nrec + 1)), *rs0 = malloc(sizeof(float) * (nrec + 1)),
          *rn = malloc(sizeof(float) * (nrec + 1));
    float ws = 0;
    for (int g = 0; g < nrec; g++) rs[g] = rs0[g] = 1.0f, rn[g] = 0.0f;
    if (scaled) {
        for (int g = 0; g < nrec; g++) {
            rs[g] = 0;
            rs0[g] = 0;
            for (int o = 0; o < nout; o++)
                for (int t = tmin; t < tmax; t++) {
                    rs[g] += pow(AT(vo, o, g, t), 2);
                    rs0[g] += pow(AT(vi, o, g, t), 2);
                }
            rs[g] = sqrt(1. / rs[g]);
            rs0[g] = sqrt(1. / rs0[g]);
        }
        if (c->scaleshot) {
            ws = 0;
            for (int g = 0; g < nrec; g++) ws += 1 / pow(rs0[g], 2);
            ws = 1 / sqrt(ws);
            for (int g = 0; g < nrec; g++) rs[g] = rs0[g] = ws;
        }
        if (c->scalerms) {
            ws = 0;
            for (int g = 0; g < nrec; g++) ws += 1 / pow(rs0[g], 2);
            ws = 1 / sqrt(ws);
            for (int g = 0; g < nrec; g++) {
                rs[g] *= ws / rs0[g];
                rs0[g] = ws;
            }
        }
        for (int g = 0; g < nrec; g++)
            for (int o = 0; o < nout; o++)
                for (int t = tmin; t < tmax; t++) {
                    AT(vo, o, g, t) *= rs[g];
                    AT(vi, o, g, t) *= rs0[g];
                }
    }
    if (norm)
        for (int g = 0; g < nrec; g++) {
            rn[g] = 0;
            for (int o = 0; o < nout; o++)
                for (int t = tmin; t < tmax; t++)
                    rn[g] += AT(vo, o, g, t) * (AT(vo, o, g, t) - AT(vi, o, g, t));
            if (c->scalerms) rn[g] *= pow(ws / rs0[g] / ws, 2);
            rn[g] = 1.0 - rn[g];
        }
    const int fixnorm = (gradfix & OR_FIX_RMSNORM) && c->scalermsnorm && !c->scalerms;
    for (int g = 0; g < nrec; g++)
        for (int t = 0; t < tmax; t++)
            for (int o = 0; o < nout; o++) {
                const float a = AT(vo, o, g, t), b = AT(vi, o, g, t);
                rms[0] += pow(-a + b, 2);
                rms[1] += pow(b, 2);
                float *r = &res[((long)o * nrec + g) * NT + t];
                if (fixnorm) *r = rs[g] * (rn[g] * a - b);
                else *r = norm ? rs[g] * ((1.0 - rn[g]) * a - b) : rs[g] * (a - b);
            }
#undef AT
    free(vo); free(vi); free(rs); free(rs0); free(rn);
}

/* res_scale (src/residuals.c:452-570), FP32 storage.  With at_receiver=0  */
/* the velocity scaling reads the buoyancy at (z=x, y=x, x=x) of the       */
/* receiver's x index, as the reference does (:487-491); at_receiver=1     */
/* reads it at the receiver's own cell.                                      */
static void res_scale(const sim *S, int nrec, const float *rec_pos, float *res, int at_receiver) {
    const or_csts *c = S->c;
    const int NT = c->NT, nout = or_nout(c), tmax = c->tmax > 0 ? c->tmax : c->NT;
    for (int o = 0; o < nout; o++) {
        const char *name = or_outname(c, o);
        const float *par = NULL;
        int isv = 0;
        if (!strcmp(name, "vx")) { par = S->rip; isv = 1; }
        else if (!strcmp(name, "vy")) { par = S->rjp; isv = 1; }
        else if (!strcmp(name, "vz")) { par = S->rkp; isv = 1; }
        else if (!strcmp(name, "p")) { par = S->M; }
        else continue;
        if (!isv && S->ND == 22) continue;
        for (int g = 0; g < nrec; g++) {
            int x, y, z;
            if (isv && !at_receiver) {
                x = y = z = (int)(rec_pos[0 + 8 * g] / c->dh);
            } else {
                x = (int)(rec_pos[0 + 8 * g] / c->dh);
                y = (int)(rec_pos[1 + 8 * g] / c->dh);
                z = (int)(rec_pos[2 + 8 * g] / c->dh);
            }
            long pos = S->ND == 3 ? (long)x * S->NZ * S->NY + (long)y * S->NZ + z
                                  : (long)x * S->NZ + z;
            if (pos < 0 || pos >= S->np) continue;
            float *r = &res[((long)o * nrec + g) * NT];
            if (isv) {
                float parscal = 1.0 / par[pos] * c->dh / c->dt * powf(2, 0);
                for (int t = 0; t < tmax; t++) r[t] *= 1.0 / parscal * c->dt;
            } else {
                float parscal = -2.0 * (par[pos] - S->mu[pos]) * c->dh / c->dt * powf(2, -0);
                for (int t = 0; t < tmax; t++) r[t] *= parscal * c->dt;
            }
        }
    }
}

/* transf_grad (src/calc_grad.c:975-1110), FP32 storage, par_scale 0, in  */
/* double; with rho_fix the direct density term is scaled by dh^2/dt^2     */
/* (finite differences of the misfit confirm the scaled form).              */
/* double from the double accumulators.  Outputs follow param_type: 0      */
/* vp, vs, rho; 1 M, mu, rho; 2 Ip, Is, rho.                                 */
static void transf_grad(const sim *S, const float *Ms, const float *mus, const double *gM,
                        const double *gmu, const double *grho, float *out, int rho_fix) {
    const or_csts *c = S->c;
    const long n = S->np;
    for (long i = 0; i < n; i++) {
        double rho = 1.0 / S->rho[i] * c->dt / c->dh;
        double M = (double)Ms[i] * c->dh / c->dt;
        double mu = (double)mus[i] * c->dh / c->dt;
        double gr = grho[i] / c->dt, gm = gM[i] / c->dt, gu = gmu[i] / c->dt;
        /* the density kernel accumulates v . dv_adjoint, whose scale differs */
        /* from dJ/drho by dh^2/dt^2; the reference omits that factor         */
        if (rho_fix) gr *= (double)c->dh * c->dh / ((double)c->dt * c->dt);
        if (c->param_type == 0 || c->param_type == 2) {
            gr = gr + M / rho * gm;
            if (mu > 0) gr = gr + mu / rho * gu;
            if (c->param_type == 0) {
                gm = 2.0 * sqrt(rho * M) * gm;
                gu = 2.0 * sqrt(rho * mu) * gu;
            } else {
                gm = 2.0 * sqrt(M / rho) * gm;
                gu = 2.0 * sqrt(mu / rho) * gu;
            }
        }
        out[i] = (float)gm;
        out[n + i] = (float)gu;
        out[2 * n + i] = (float)gr;
    }
}

/* The Hessian through the same transform (src/calc_grad.c:1028-1094), not */
/* divided by dt; its density part holds only the chain-rule terms.         */
static void transf_hess(const sim *S, const float *Ms, const float *mus, const double *hM,
                        const double *hmu, float *out) {
    const or_csts *c = S->c;
    const long n = S->np;
    for (long i = 0; i < n; i++) {
        double rho = 1.0 / S->rho[i] * c->dt / c->dh;
        double M = (double)Ms[i] * c->dh / c->dt;
        double mu = (double)mus[i] * c->dh / c->dt;
        double hr = 0, hm = hM[i], hu = hmu[i];
        if (c->param_type == 0 || c->param_type == 2) {
            hr = hr + M / rho * hm;
            if (mu > 0) hr = hr + mu / rho * hu;
            if (c->param_type == 0) {
                hm = 2.0 * sqrt(rho * M) * hm;
                hu = 2.0 * sqrt(rho * mu) * hu;
            } else {
                hm = 2.0 * sqrt(M / rho) * hm;
                hu = 2.0 * sqrt(mu / rho) * hu;
            }
        }
        out[i] = (float)hm;
        out[n + i] = (float)hu;
        out[2 * n + i] = (float)hr;
    }
}

static int adj_alloc(sim *A, const sim *S, char *err, int errlen) {
    /* the adjoint run shares the model and the absorbing profiles; it has   */
    /* its own wavefields and its own zeroed CPML memory                      */
    if (sim_init(A, S->c, err, errlen)) return 1;
    free(A->M); free(A->mu); free(A->rho); free(A->rip); free(A->rjp); free(A->rkp);
    free(A->muipkp); free(A->muipjp); free(A->mujpkp);
    free(A->taup); free(A->taus); free(A->tausipkp); free(A->tausipjp); free(A->tausjpkp);
    A->M = S->M; A->mu = S->mu; A->rho = S->rho; A->rip = S->rip; A->rjp = S->rjp;
    A->rkp = S->rkp; A->muipkp = S->muipkp; A->muipjp = S->muipjp; A->mujpkp = S->mujpkp;
    A->taup = S->taup; A->taus = S->taus; A->tausipkp = S->tausipkp;
    A->tausipjp = S->tausipjp; A->tausjpkp = S->tausjpkp;
    int ninc = S->ND == 3 ? 6 : 3;
    for (int q = 0; q < ninc; q++) A->inc[q] = zalloc(S->nv);
    if (S->L > 0) {
        /* the tau parts and the relaxation sum of the mu transform */
        for (int q = 0; q < (S->ND == 3 ? 7 : 4); q++) A->tinc[q] = zalloc(S->nv);
        float pts[OR_MAXL];
        for (int l = 0; l < S->L; l++) pts[l] = S->c->dt / S->eta[l];
        float ws = 2.0 * PI * S->c->f0;
        float sumu = 0.0;
        for (int l = 0; l < S->L; l++)
            sumu += ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l]));
        A->s0 = sumu;
    }
    return 0;
}

static void adj_free(sim *A) {
    A->M = A->mu = A->rho = A->rip = A->rjp = A->rkp = A->muipkp = A->muipjp = A->mujpkp = NULL;
    A->taup = A->taus = A->tausipkp = A->tausipjp = A->tausjpkp = NULL;
    for (int q = 0; q < 6; q++) {
        free(A->inc[q]);
        A->inc[q] = NULL;
    }
    for (int q = 0; q < 7; q++) {
        free(A->tinc[q]);
        A->tinc[q] = NULL;
    }
    sim_free(A);
}

/* The reference's adjoint kernels update the forward CPML memory, which   */
/* still holds the end of the forward run: the adjoint starts from a copy.  */
static void adj_share_cpml(sim *A, const sim *S) {
    float *const *fw[] = {&S->psi_sxx_x, &S->psi_sxy_x, &S->psi_sxz_x, &S->psi_sxy_y,
                          &S->psi_syy_y, &S->psi_syz_y, &S->psi_sxz_z, &S->psi_syz_z,
                          &S->psi_szz_z, &S->psi_vx_x,  &S->psi_vy_x,  &S->psi_vz_x,
                          &S->psi_vx_y,  &S->psi_vy_y,  &S->psi_vz_y,  &S->psi_vx_z,
                          &S->psi_vy_z,  &S->psi_vz_z};
    float **aw[] = {&A->psi_sxx_x, &A->psi_sxy_x, &A->psi_sxz_x, &A->psi_sxy_y, &A->psi_syy_y,
                    &A->psi_syz_y, &A->psi_sxz_z, &A->psi_syz_z, &A->psi_szz_z, &A->psi_vx_x,
                    &A->psi_vy_x,  &A->psi_vz_x,  &A->psi_vx_y,  &A->psi_vy_y,  &A->psi_vz_y,
                    &A->psi_vx_z,  &A->psi_vy_z,  &A->psi_vz_z};
    long nab2 = 2 * S->NAB, nyi = S->ND == 3 ? S->NY : 1;
    long sizes[3] = {nab2 * nyi * S->NZ, (long)S->NX * nab2 * S->NZ, (long)S->NX * nyi * nab2};
    const int axis[18] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 0, 0, 0, 1, 1, 1, 2, 2, 2};
    for (int q = 0; q < 18; q++)
        if (*fw[q] && *aw[q]) memcpy(*aw[q], *fw[q], sizeof(float) * sizes[axis[q]]);
}

/* Movie frame of a gradient run (src/time_stepping.c:926-929, frame index
 * of :317-320): every variable, padded, into frame (t+1)/movout - 1 when
 * t % movout == 0. */
static void movie_snap(const or_csts *c, sim *S, int movout, float *movie, int t) {
    if (!movie || movout <= 0 || t % movout) return;
    const int nfr = c->NT / movout, f = (t + 1) / movout - 1;
    if (f < 0 || f >= nfr) return;
    for (int v = 0; v < or_nvars(c); v++)
        memcpy(movie + ((size_t)v * nfr + f) * S->nv, named_var(S, or_varname(c, v)),
               sizeof(float) * S->nv);
}

int or_gradient(const or_csts *c, const float *p0, const float *p1, const float *rho,
                const float *taup, const float *taus, int nsrc, const float *src_pos,
                const float *src, int nrec, const float *rec_pos, const float *din,
                const float *res_in, int gradfix, float *out, float *res_out,
                float *grad, double *rms, float *fields_out, float *hess, float *gradtau,
                int movout,
                float *movie, char *err, int errlen) {
    sim S, A;
    char dummy[8];
    if (!err) {
        err = dummy;
        errlen = sizeof(dummy);
    }
    if (c->ND == 22) {
        snprintf(err, errlen, "gradients need the elastic or viscoelastic case (ND 2 or 3)");
        return 1;
    }
    if (sim_init(&S, c, err, errlen)) {
        sim_free(&S);
        return 1;
    }
#ifdef _OPENMP
    if (c->nthreads > 0) omp_set_num_threads(c->nthreads);
#endif
    or_transform(c, p0, p1, rho, taup, taus, S.M, S.mu, S.rho, S.rip, S.rjp, S.rkp,
                 S.muipkp, S.muipjp, S.mujpkp, S.taup, S.taus, S.tausipkp,
                 S.tausipjp, S.tausjpkp);
    float dts;
    if (or_check_stability(c, S.M, S.mu, S.rho, &dts)) {
        snprintf(err, errlen, "Time step too large, to be stable, set dt<%f", dts);
        sim_free(&S);
        return 1;
    }
    if (check_positions(&S, nsrc, src_pos, 5, err, errlen) ||
        check_positions(&S, nrec, rec_pos, 8, err, errlen)) {
        sim_free(&S);
        return 1;
    }
    shell_t B;
    shell_init(&S, &B);
    const int tmax = c->tmax > 0 ? c->tmax : c->NT, tmin = c->tmin;
    if (tmax > c->NT || tmin < 0 || tmin >= tmax) {
        snprintf(err, errlen, "time window [tmin=%d, tmax=%d) outside [0, NT=%d]", tmin, tmax,
                 c->NT);
        sim_free(&S);
        return 1;
    }
    const int nvb = S.ND == 3 ? 9 : 5, nout = or_nout(c);
    /* Viscoelastic runs cannot be reconstructed backwards (the memory-     */
    /* variable recursion decays forwards and grows backwards, which makes  */
    /* the reference's backpropagation unstable): the adjoint reads the     */
    /* forward state each step needs from a stored history instead, here    */
    /* the post-source velocities and stresses of every step t in           */
    /* (tmin, tmax).  What the elastic backpropagation reconstructs at step  */
    /* t is exactly this state: the stresses and velocities after sources(t) */
    /* and before update_v(t).                                                */
    const int visco = c->L > 0 || (gradfix & OR_HISTORY);
    float *hist = NULL;
    const size_t nh = (size_t)(tmax - tmin - 1) * nvb;
    float *fld[9];
    int nfld = 0;
    if (S.ND == 3) {
        float *f3[9] = {S.vx, S.vy, S.vz, S.sxx, S.syy, S.szz, S.sxy, S.syz, S.sxz};
        for (int q = 0; q < 9; q++) fld[nfld++] = f3[q];
    } else {
        float *f2[5] = {S.vx, S.vz, S.sxx, S.szz, S.sxz};
        for (int q = 0; q < 5; q++) fld[nfld++] = f2[q];
    }
    float *bnd = visco ? NULL : malloc(sizeof(float) * (size_t)tmax * nvb * B.n);
    if (visco) {
        hist = malloc(sizeof(float) * (nh ? nh : 1) * S.nv);
        if (!hist) {
            snprintf(err, errlen, "out of host memory for the forward history");
            sim_free(&S);
            return 1;
        }
    }
    memset(out, 0, sizeof(float) * (size_t)nout * nrec * c->NT);
    /* forward with boundary saving: src/time_stepping.c:755-835 */
    for (int t = 0; t < tmax; t++) {
        /* movie of a viscoelastic gradient run: the forward state at the */
        /* start of step t, what backpropagation would reconstruct after  */
        /* the adjoint step t                                             */
        if (visco && t > tmin) movie_snap(c, &S, movout, movie, t);
        sources(&S, t, 1, nsrc, src_pos, src);
        if (visco && t > tmin)
            for (int q = 0; q < nfld; q++)
                memcpy(hist + ((size_t)(t - tmin - 1) * nvb + q) * S.nv, fld[q],
                       sizeof(float) * S.nv);
        if (S.ND == 3) {
            update_v3D(&S);
            update_s3D(&S);
        } else {
            update_v2D(&S);
            update_s2D(&S);
        }
        if (!visco) save_shell(&S, &B, bnd + (size_t)t * nvb * B.n);
        if (c->freesurf) {
            if (S.ND == 3) surface3D(&S, 1);
            else surface2D(&S, 1);
        }
        varsout(&S, t, nrec, rec_pos, out);
    }
    /* residuals and their scaling: src/time_stepping.c:838-848 */
    size_t nres = (size_t)nout * nrec * c->NT;
    float *res = calloc(nres, sizeof(float));
    double rr[2] = {0, 0};
    if (res_in) memcpy(res, res_in, nres * sizeof(float));
    else or_residuals(c, nout, nrec, out, din, res, rr, gradfix);
    res_scale(&S, nrec, rec_pos, res, gradfix & OR_FIX_RES_SCALE);
    if (res_out) memcpy(res_out, res, nres * sizeof(float));
    if (rms) {
        rms[0] = rr[0];
        rms[1] = rr[1];
    }
    /* adjoint: src/time_stepping.c:869-940 */
    if (adj_alloc(&A, &S, err, errlen)) {
        free(bnd);
        free(hist);
        free(res);
        sim_free(&S);
        return 1;
    }
    if (!(gradfix & OR_FIX_ADJ_CPML)) adj_share_cpml(&A, &S);
    /* Moduli of the gradient formulas.  The reference uses the relaxed     */
    /* moduli M, mu for the stress-to-strain conversion and the chain rule   */
    /* (src/update_adjs2D.cl:441-451, src/calc_grad.c:1020-1056); for a       */
    /* viscoelastic medium the stresses relate to the strains through the    */
    /* moduli at the reference frequency f0, larger by the relaxation factor */
    /* 1+s, which inflates the reference's gradient by about (1+s)^(3/2).    */
    /* With OR_FIX_VISCO_F0 the formulas use the moduli at f0: what the      */
    /* transform gives without the relaxation division (that is, with L=0).  */
    const float *Mg = S.M, *mug = S.mu;
    float *Mf = NULL, *muf = NULL;
    if (c->L > 0 && (gradfix & OR_FIX_VISCO_F0)) {
        or_csts c0 = *c;
        c0.L = 0;
        float *tmp[14];
        for (int q = 0; q < 14; q++) tmp[q] = zalloc(S.np);
        Mf = tmp[0];
        muf = tmp[1];
        or_transform(&c0, p0, p1, rho, taup, taus, tmp[0], tmp[1], tmp[2], tmp[3], tmp[4],
                     tmp[5], tmp[6], tmp[7], tmp[8], tmp[9], tmp[10], tmp[11], tmp[12], tmp[13]);
        for (int q = 2; q < 14; q++) free(tmp[q]);
        Mg = Mf;
        mug = muf;
    }
    double *gM = calloc(S.np, sizeof(double)), *gmu = calloc(S.np, sizeof(double)),
           *grho = calloc(S.np, sizeof(double));
    double *hM = hess ? calloc(S.np, sizeof(double)) : NULL,
           *hmu = hess ? calloc(S.np, sizeof(double)) : NULL;
    const int want_tau = c->L > 0 && gradtau && c->restype != 1 && c->param_type <= 2;
    double *gtp = want_tau ? calloc(S.np, sizeof(double)) : NULL,
           *gts = want_tau ? calloc(S.np, sizeof(double)) : NULL;
    for (int t = tmax - 1; t > tmin; t--) {
        if (c->freesurf) {
            if (S.ND == 3) surface_adj3D(&A, gradfix & OR_FIX_SURF_SYY);
            else surface_adj2D(&A);
        }
        residuals_inject(&A, t, nrec, rec_pos, res);
        if (visco) {
            /* the forward state of step t from the history; the adjoint      */
            /* stress update with memory variables is the forward one run on  */
            /* the adjoint fields (src/update_adjs2D.cl:367-408,               */
            /* src/update_adjs3D.cl:549-605)                                   */
            for (int q = 0; q < nfld; q++)
                memcpy(fld[q], hist + ((size_t)(t - tmin - 1) * nvb + q) * S.nv,
                       sizeof(float) * S.nv);
            if (S.ND == 3) update_s3D(&A);
            else update_s2D(&A);
            grad_Mmu(&S, &A, Mg, mug, gM, gmu, hM, hmu, gradfix & OR_FIX_HESS);
            if (gtp) grad_tau(&S, &A, Mg, mug, gtp, gts);
            if (S.ND == 3) update_v3D(&A);
            else update_v2D(&A);
            grad_rho(&S, &A, grho);
            continue;
        }
        const float *bprev = bnd + (size_t)(t - 1) * nvb * B.n;
        /* update_adjs: backpropagate the stresses, re-inject their shell,  */
        /* update the adjoint stresses, accumulate the M and mu gradients    */
        backprop_s(&S);
        inject_shell(&S, &B, bprev, S.ND == 3 ? 3 : 2, S.ND == 3 ? 6 : 3);
        if (S.ND == 3) update_s3D(&A);
        else update_s2D(&A);
        grad_Mmu(&S, &A, Mg, mug, gM, gmu, hM, hmu, gradfix & OR_FIX_HESS);
        /* update_adjv: the same for the velocities and the density */
        backprop_v(&S);
        inject_shell(&S, &B, bprev, 0, S.ND == 3 ? 3 : 2);
        if (S.ND == 3) update_v3D(&A);
        else update_v2D(&A);
        grad_rho(&S, &A, grho);
        sources(&S, t, -1, nsrc, src_pos, src);
        movie_snap(c, &S, movout, movie, t);
    }
    transf_grad(&S, Mg, mug, gM, gmu, grho, grad, gradfix & OR_FIX_RHO_SCALE);
    if (hess) transf_hess(&S, Mg, mug, hM, hmu, hess);
    if (gradtau) {
        /* the tau gradients at fixed wave speeds: the accumulators hold     */
        /* dt^2/dh times dJ/dtau (as the moduli's hold dt^2/dh times the     */
        /* derivative by the scaled moduli, transf_grad); zero for elastic   */
        const double k = (double)c->dh / ((double)c->dt * c->dt);
        for (long i = 0; i < S.np; i++) {
            gradtau[i] = gtp ? (float)(gtp[i] * k) : 0.0f;
            gradtau[S.np + i] = gts ? (float)(gts[i] * k) : 0.0f;
        }
    }
    free(gtp);
    free(gts);
    free(hM);
    free(hmu);
    if (fields_out) {
        for (int v = 0; v < or_nvars(c); v++)
            memcpy(&fields_out[v * S.nv], named_var(&S, or_varname(c, v)),
                   S.nv * sizeof(float));
    }
    free(gM);
    free(gmu);
    free(grho);
    free(Mf);
    free(muf);
    free(bnd);
    free(hist);
    free(res);
    adj_free(&A);
    sim_free(&S);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* DFT-domain gradient (BACK_PROP_TYPE=2).                                   */
/*   sampling: src/assign_modeling_case.c:549-570 (gradfreqsn);               */
/*   forward: the DFT of every wavefield and memory variable before the      */
/*     sources of each sampled step (src/time_stepping.c:767-780, savefreqs  */
/*     src/automatic_kernels.c:1030-1128);                                    */
/*   adjoint: the adjoint wavefield alone (no backpropagation), its DFT after */
/*     the updates of each sampled step (src/time_stepping.c:869-925);        */
/*   gradient: the correlation of the two spectra per frequency with the     */
/*     coefficients of the parameterisation (src/calc_grad.c:111-376,        */
/*     377-968); no transf_grad (:982-1002 runs for BACK_PROP_TYPE=1 only);  */
/*   misfit: rms at the gradient frequencies (src/residuals.c:364-403).       */
/* ------------------------------------------------------------------------ */
int or_dft_sampling(const or_csts *c, int nf, const float *freqs, int *dtnyq, int *ntnyq,
                    float *df, float *bins) {
    float fmax = 0;
    for (int f = 0; f < nf; f++)
        if (freqs[f] > fmax) fmax = freqs[f];
    if (nf <= 0 || fmax <= 0) return 1;
    const int tmax = c->tmax > 0 ? c->tmax : c->NT;
    *dtnyq = (int)ceil(0.0156 / fmax / c->dt);
    *ntnyq = (tmax - c->tmin) / *dtnyq + 1;
    *df = (float)(1.0 / *ntnyq / c->dt / *dtnyq);
    for (int f = 0; f < nf; f++) {
        bins[f] = (float)floor(freqs[f] / *df);
        /* bin 0 (below df) would divide the memory-variable terms by 0 */
        if (bins[f] < 1) return 2;
    }
    return 0;
}

/* factor of sample thist at every frequency: dtnyq dt e^{-i pi 2 bin thist / ntnyq} */
static void dft_fact(const or_csts *c, int nf, const float *bins, int dtnyq, int ntnyq,
                     int thist, float *fact) {
    const float s = (float)dtnyq * c->dt;
    for (int f = 0; f < nf; f++) {
        const double a = 2.0 * bins[f] * thist / ntnyq;
        fact[2 * f] = (float)(s * cos(PI * a));
        fact[2 * f + 1] = (float)(-s * sin(PI * a));
    }
}

/* the arrays a DFT covers, in order: velocities and stresses, then each  */
/* memory-variable component's L slabs                                      */
static int dft_arrays(sim *S, float **q) {
    int n = 0;
    if (S->ND == 3) {
        float *f[9] = {S->vx, S->vy, S->vz, S->sxx, S->syy, S->szz, S->sxy, S->syz, S->sxz};
        for (int k = 0; k < 9; k++) q[n++] = f[k];
        float *r[6] = {S->rxx, S->ryy, S->rzz, S->rxy, S->ryz, S->rxz};
        for (int k = 0; k < 6 && S->L > 0; k++)
            for (int l = 0; l < S->L; l++) q[n++] = r[k] + (long)l * S->nv;
    } else {
        float *f[5] = {S->vx, S->vz, S->sxx, S->szz, S->sxz};
        for (int k = 0; k < 5; k++) q[n++] = f[k];
        float *r[3] = {S->rxx, S->rzz, S->rxz};
        for (int k = 0; k < 3 && S->L > 0; k++)
            for (int l = 0; l < S->L; l++) q[n++] = r[k] + (long)l * S->nv;
    }
    return n;
}

/* F[q][f][cell] (complex) += fact[f] u_q[cell] over the padded arrays */
static void dft_acc(sim *S, float *F, int nf, const float *fact) {
    float *q[9 + 6 * OR_MAXL];
    const int nq = dft_arrays(S, q);
    const long nv = S->nv;
    for (int a = 0; a < nq; a++)
        for (int f = 0; f < nf; f++) {
            float *Fq = F + ((long)a * nf + f) * nv * 2;
            const float fx = fact[2 * f], fy = fact[2 * f + 1];
#pragma omp parallel for schedule(static)
            for (long i = 0; i < nv; i++) {
                const float u = q[a][i];
                Fq[2 * i] += fx * u;
                Fq[2 * i + 1] += fy * u;
            }
        }
}

typedef struct {
    float x, y;
} cfl;
static inline cfl c_add(cfl a, cfl b) { return (cfl){a.x + b.x, a.y + b.y}; }
static inline cfl c_sub(cfl a, cfl b) { return (cfl){a.x - b.x, a.y - b.y}; }
static inline cfl c_add3(cfl a, cfl b, cfl c) { return (cfl){a.x + b.x + c.x, a.y + b.y + c.y}; }
static inline cfl c_sub3(cfl a, cfl b, cfl c) { return (cfl){a.x - b.x - c.x, a.y - b.y - c.y}; }
/* the spectrum of the time integral: a / (i w) */
static inline cfl c_integ(cfl a, float w) { return (cfl){a.y / w, -a.x / w}; }
/* Im(conj(b) a) */
static inline float c_itreal(cfl a, cfl b) { return a.y * b.x - a.x * b.y; }
/* the memory-variable correlation: ts Re(conj(b) a) + Im(conj(a) b) / w */
static inline float c_rm(cfl a, cfl b, float ts, float w) {
    return ts * (a.x * b.x + a.y * b.y) + (a.x * b.y - a.y * b.x) / w;
}

/* The 24 coefficients of the scalar products for one cell
 * (src/calc_grad.c:111-376: grad_coefvisc_0..3, grad_coefelast_0..2), in
 * double.  M, mu, rho are the transformed parameters (rho the buoyancy
 * dt/(dh rho)), ND and L as floats, al = sum_l (w0 tau_l)^2/(1+(w0 tau_l)^2).
 * The square roots are the reference's float ones. */
void or_dft_coef(double *c, int pt, int visco, float M, float mu, float tp, float ts, float rho,
                 float ND, float L, float al) {
    for (int k = 0; k < 24; k++) c[k] = 0;
    const double Md = M, mud = mu, rd = rho, n3 = (ND + 1.0) / 3.0, Lal = L - al;
    const double kp = 1.0 + L * tp, ks = 1.0 + L * ts, ap = 1.0 + al * tp, as = 1.0 + al * ts;
    const double mu2 = mud * mud, as2 = as * as, ap2 = ap * ap;
    if (!visco) {
        const double F = (ND * Md - 2.0 * (ND - 1.0) * mud) * (ND * Md - 2.0 * (ND - 1.0) * mud);
        if (pt == 1) {
            c[0] = 1.0 / F;
            c[2] = 1.0 / mu2;
            c[3] = n3 / F;
            c[4] = 1.0 / (2 * ND * mu2);
            return;
        }
        const double A = pt == 0 ? 2.0 * sqrtf(rho * M) : 2.0 * sqrtf(M / rho);
        const double B = pt == 0 ? 2.0 * sqrtf(rho * mu) : 2.0 * sqrtf(mu / rho);
        c[0] = A / F;
        c[2] = B / mu2;
        c[3] = B * n3 / F;
        c[4] = B / (2 * ND * mu2);
        if (pt == 0) {
            c[16] = Md / rd / F;
            c[18] = mud / rd / mu2;
            c[19] = mud / rd * n3 / F;
            c[20] = mud / rd / (2 * ND * mu2);
        } else {
            c[16] = -Md / rd * kp * ap * as2;
            c[18] = -mud / rd / mu2;
            c[19] = -mud / rd * n3 / F;
            c[20] = -mud / rd / (2 * ND * mu2);
        }
        return;
    }
    const double g1 = ND * Md * kp * as - 2.0 * (ND - 1.0) * mud * ks * ap;
    const double g2 = ND * Md * tp * as - 2.0 * (ND - 1.0) * mud * ts * ap;
    const double f1 = g1 * g1, f2 = g2 * g2;
    /* the tau parts (c8..c15) and the unweighted moduli parts */
    const double t8 = Md * Lal * as2 / f1, t9 = Md * as2 / f2;
    const double t10 = Lal / (mud * ks * ks), t11 = n3 * mud * Lal * ap2 / f1;
    const double t12 = Lal / (2 * ND * mud * ks * ks), t13 = 1.0 / (mud * ts * ts);
    const double t14 = n3 * mud * ap2 / f2, t15 = 1.0 / (2 * ND * mud * ts * ts);
    const double u[8] = {kp * ap * as2 / f1,        tp * ap * as2 / f2,
                         as / (mu2 * ks),           n3 * ks * as * ap2 / f1,
                         as / (2.0 * ND * mu2 * ks), as / (mu2 * ts),
                         n3 * ts * as * ap2 / f2,   as / (2.0 * ND * mu2 * ts)};
    double A = 1, B = 1;
    if (pt == 0 || pt == 3) {
        A = 2.0 * sqrtf(rho * M);
        B = 2.0 * sqrtf(rho * mu);
    } else if (pt == 2) {
        A = 2.0 * sqrtf(M / rho);
        B = 2.0 * sqrtf(mu / rho);
    }
    for (int k = 0; k < 8; k++) c[k] = (k < 2 ? A : B) * u[k];
    const double t[8] = {t8, t9, t10, t11, t12, t13, t14, t15};
    for (int k = 0; k < 8; k++) c[8 + k] = t[k];
    if (pt == 3) {
        /* taup is the Q parameter: its chain through vp = sqrt(M/rho)(1+taup) */
        const double sm = sqrtf(M / rho), R = tp / sm, Rp = (1.0 + tp) / sm;
        const double w[8] = {Md * Lal * as2, Md * as2, Lal / (mud * ks * ks),
                             n3 * mud * Lal * ap2, Lal / (2 * ND * mud * ks * ks),
                             1.0 / (mud * ts * ts), n3 * mud * ap2, 1.0 / (2 * ND * mud * ts * ts)};
        const double dv[8] = {f1, f2, 1, f1, 1, 1, f2, 1};
        const double P[8] = {A * kp * ap * as2, A * tp * ap * as2, B * as / (mu2 * ks),
                             B * n3 * ks * as * ap2, B * as / (2.0 * ND * mu2 * ks),
                             B * as / (mu2 * ts), B * n3 * ts * as * ap2,
                             B * as / (2.0 * ND * mu2 * ts)};
        for (int k = 0; k < 8; k++) {
            c[k] = (P[k] - R * w[k]) / dv[k];
            c[8 + k] = (-P[k] + Rp * w[k]) / dv[k];
        }
    }
    if (pt == 1) return;
    const double s = pt == 2 ? -1.0 : 1.0, m_r = Md / rd, u_r = mud / rd;
    c[16] = s * m_r * u[0];
    c[17] = s * m_r * u[1];
    for (int k = 2; k < 8; k++) c[16 + k] = s * u_r * u[k];
}

/* Gradient increments of one shot from the forward (Ff) and adjoint (Fa)
 * spectra (src/calc_grad.c:608-968): per interior cell and frequency, the
 * memory-variable spectra are integrated out of the stress spectra, the
 * scalar products dot[0..8] formed, and grad[5][np] (param order, then
 * taup, taus) incremented in double.  The reference's 3D memory-variable
 * terms are kept as written: ryy_mxxzz and rzz_mxxyy repeat rxx_myyzz, and
 * the adjoint syz spectrum is corrected with the forward ryz. */
static void dft_grad(sim *S, const float *Ff, const float *Fa, int nf, const float *bins,
                     float df, int ntnyq, double *grad) {
    const or_csts *c = S->c;
    const int ND3 = S->ND == 3, L = S->L, nq = ND3 ? 9 + 6 * L : 5 + 3 * L;
    const long nv = S->nv, np = S->np;
    const float NDf = (float)S->ND, Lf = (float)L;
    float tausig[OR_MAXL];
    float al = 0;
    const float w0 = 2.0 * PI * c->f0;
    for (int l = 0; l < L; l++) {
        tausig[l] = 1.0 / (2.0 * PI * c->FL[l]);
        const double r = w0 / (2.0 * PI * c->FL[l]);
        al += r * r / (1.0 + r * r);
    }
    const int pt = c->param_type;
    const int H = S->FDOH;
#pragma omp parallel for schedule(static)
    for (long m = 0; m < np; m++) {
        int i, j, k;
        if (ND3) {
            i = (int)(m / ((long)S->NY * S->NZ));
            j = (int)(m / S->NZ % S->NY);
            k = (int)(m % S->NZ);
        } else {
            i = (int)(m / S->NZ);
            j = 0;
            k = (int)(m % S->NZ);
        }
        const long iv = ND3 ? ((long)(i + H) * S->NYp + (j + H)) * S->NZp + (k + H)
                            : (long)(i + H) * S->NZp + (k + H);
        double cf[24];
        or_dft_coef(cf, pt, L > 0, S->M[m], S->mu[m], L > 0 ? S->taup[m] : 0.f,
                    L > 0 ? S->taus[m] : 0.f, S->rho[m], NDf, Lf, al);
        if (S->mu[m] < 1) {
            for (int n = 2; n < 8; n++) cf[n] = 0;
            for (int n = 10; n < 16; n++) cf[n] = 0;
            for (int n = 18; n < 24; n++) cf[n] = 0;
        }
        for (int f = 0; f < nf; f++) {
            const float w = (float)(2.0 * PI * df * bins[f]);
            cfl A[9 + 6 * OR_MAXL], B[9 + 6 * OR_MAXL];   /* forward, adjoint */
            for (int q = 0; q < nq; q++) {
                const long o = (((long)q * nf + f) * nv + iv) * 2;
                A[q] = (cfl){Ff[o], Ff[o + 1]};
                B[q] = (cfl){Fa[o], Fa[o + 1]};
            }
            double dot[9] = {0};
            if (ND3) {
                /* A/B: vx vy vz sxx syy szz sxy syz sxz, rxx.. ryy.. rzz.. rxy.. ryz.. rxz.. */
                cfl *sxx = &A[3], *syy = &A[4], *szz = &A[5], *sxy = &A[6], *syz = &A[7],
                    *sxz = &A[8];
                cfl *sxxr = &B[3], *syyr = &B[4], *szzr = &B[5], *sxyr = &B[6], *syzr = &B[7],
                    *sxzr = &B[8];
                for (int l = 0; l < L; l++) {
                    const cfl rxx = A[9 + l], ryy = A[9 + L + l], rzz = A[9 + 2 * L + l],
                              rxy = A[9 + 3 * L + l], ryz = A[9 + 4 * L + l],
                              rxz = A[9 + 5 * L + l];
                    const cfl rxxr = B[9 + l], ryyr = B[9 + L + l], rzzr = B[9 + 2 * L + l],
                              rxyr = B[9 + 3 * L + l], ryzr = B[9 + 4 * L + l],
                              rxzr = B[9 + 5 * L + l];
                    *sxx = c_sub(*sxx, c_integ(rxx, w));
                    *szz = c_sub(*szz, c_integ(rzz, w));
                    *syy = c_sub(*syy, c_integ(ryy, w));
                    *sxz = c_sub(*sxz, c_integ(rxz, w));
                    *sxy = c_sub(*sxy, c_integ(rxy, w));
                    *syz = c_sub(*syz, c_integ(ryz, w));
                    *sxxr = c_sub(*sxxr, c_integ(rxxr, w));
                    *szzr = c_sub(*szzr, c_integ(rzzr, w));
                    *syyr = c_sub(*syyr, c_integ(ryyr, w));
                    *sxzr = c_sub(*sxzr, c_integ(rxzr, w));
                    *sxyr = c_sub(*sxyr, c_integ(rxyr, w));
                    *syzr = c_sub(*syzr, c_integ(ryz, w));
                    const cfl rsum = c_add3(rxx, ryy, rzz), rsumr = c_add3(rxxr, ryyr, rzzr);
                    const cfl rdev = c_sub3(rxx, ryy, rzz);
                    dot[1] += c_rm(rsumr, rsum, tausig[l], w) / ntnyq;
                    dot[5] += (c_rm(rxyr, rxy, tausig[l], w) + c_rm(rxzr, rxz, tausig[l], w) +
                               c_rm(ryzr, ryz, tausig[l], w)) /
                              ntnyq;
                    dot[6] = dot[1];
                    dot[7] += (c_rm(rxxr, rdev, tausig[l], w) + c_rm(ryyr, rdev, tausig[l], w) +
                               c_rm(rzzr, rdev, tausig[l], w)) /
                              ntnyq;
                }
                const cfl ssum = c_add3(*sxx, *syy, *szz), ssumr = c_add3(*sxxr, *syyr, *szzr);
                const cfl dxx = c_sub3(*sxx, *syy, *szz), dyy = c_sub3(*syy, *sxx, *szz),
                          dzz = c_sub3(*szz, *sxx, *syy);
                dot[0] = w * c_itreal(ssumr, ssum) / ntnyq;
                dot[2] = w * (c_itreal(*sxyr, *sxy) + c_itreal(*sxzr, *sxz) +
                              c_itreal(*syzr, *syz)) / ntnyq;
                dot[3] = dot[0];
                dot[4] = w * (c_itreal(*sxxr, dxx) + c_itreal(*syyr, dyy) +
                              c_itreal(*szzr, dzz)) / ntnyq;
                dot[8] = w * (c_itreal(B[0], A[0]) + c_itreal(B[1], A[1]) +
                              c_itreal(B[2], A[2])) / ntnyq;
            } else {
                /* A/B: vx vz sxx szz sxz, rxx.. rzz.. rxz.. */
                cfl *sxx = &A[2], *szz = &A[3], *sxz = &A[4];
                cfl *sxxr = &B[2], *szzr = &B[3], *sxzr = &B[4];
                for (int l = 0; l < L; l++) {
                    const cfl rxx = A[5 + l], rzz = A[5 + L + l], rxz = A[5 + 2 * L + l];
                    const cfl rxxr = B[5 + l], rzzr = B[5 + L + l], rxzr = B[5 + 2 * L + l];
                    *sxx = c_sub(*sxx, c_integ(rxx, w));
                    *szz = c_sub(*szz, c_integ(rzz, w));
                    *sxz = c_sub(*sxz, c_integ(rxz, w));
                    *sxxr = c_sub(*sxxr, c_integ(rxxr, w));
                    *szzr = c_sub(*szzr, c_integ(rzzr, w));
                    *sxzr = c_sub(*sxzr, c_integ(rxzr, w));
                    const cfl rsum = c_add(rxx, rzz), rsumr = c_add(rxxr, rzzr);
                    const cfl dxx = c_sub(rxx, rzz), dzz = c_sub(rzz, rxx);
                    dot[1] += c_rm(rsumr, rsum, tausig[l], w) / ntnyq;
                    dot[5] += (c_rm(rxzr, rxz, tausig[l], w)) / ntnyq;
                    dot[6] = dot[1];
                    dot[7] += (c_rm(rxxr, dxx, tausig[l], w) + c_rm(rzzr, dzz, tausig[l], w)) /
                              ntnyq;
                }
                const cfl ssum = c_add(*sxx, *szz), ssumr = c_add(*sxxr, *szzr);
                const cfl dxx = c_sub(*sxx, *szz), dzz = c_sub(*szz, *sxx);
                dot[0] = w * c_itreal(ssumr, ssum) / ntnyq;
                dot[2] = w * (c_itreal(*sxzr, *sxz)) / ntnyq;
                dot[3] = dot[0];
                dot[4] = w * (c_itreal(*sxxr, dxx) + c_itreal(*szzr, dzz)) / ntnyq;
                dot[8] = w * (c_itreal(B[0], A[0]) + c_itreal(B[1], A[1])) / ntnyq;
            }
            grad[m] += -cf[0] * dot[0] + cf[1] * dot[1];
            grad[np + m] += -cf[2] * dot[2] + cf[3] * dot[3] - cf[4] * dot[4] + cf[5] * dot[5] -
                            cf[6] * dot[6] + cf[7] * dot[7];
            if (L > 0) {
                grad[3 * np + m] += -cf[8] * dot[0] + cf[9] * dot[1];
                grad[4 * np + m] += -cf[10] * dot[2] + cf[11] * dot[3] - cf[12] * dot[4] +
                                    cf[13] * dot[5] - cf[14] * dot[6] + cf[15] * dot[7];
            }
            grad[2 * np + m] += -dot[8] + cf[16] * dot[0] - cf[17] * dot[1] + cf[18] * dot[2] -
                                cf[19] * dot[3] + cf[20] * dot[4] - cf[21] * dot[5] +
                                cf[22] * dot[6] - cf[23] * dot[7];
        }
    }
}

/* rms and rmsnorm at the gradient frequencies (src/residuals.c:364-403):   */
/* the traces -out + din and din over t < tmax, band-passed as for the     */
/* residuals, zero-padded to nfft (the smallest 2,3,5-smooth length >=     */
/* tmax); rms sums |X(k)|^2 at k = (int)(freq nfft dt + 1) of every output  */
/* and receiver, with the DFT bin evaluated directly in double.             */
int or_dft_rms(const or_csts *c, int nout, int nrec, const float *out, const float *din,
               int nf, const float *freqs, double *rms) {
    const int NT = c->NT, tmax = c->tmax > 0 ? c->tmax : c->NT;
    int nfft = tmax < 1 ? 1 : tmax;
    for (;; nfft++) {
        int k = nfft;
        while (k % 2 == 0) k /= 2;
        while (k % 3 == 0) k /= 3;
        while (k % 5 == 0) k /= 5;
        if (k == 1) break;
    }
    int kb[OR_MAXF];
    if (nf > OR_MAXF) return 1;
    for (int f = 0; f < nf; f++) {
        kb[f] = (int)(freqs[f] * nfft * c->dt + 1);
        if (kb[f] > nfft / 2) return 1;
    }
    const long n = (long)nout * nrec * NT;
    float *vo = malloc(sizeof(float) * (n ? n : 1)), *vi = malloc(sizeof(float) * (n ? n : 1));
    memcpy(vo, out, sizeof(float) * n);
    memcpy(vi, din, sizeof(float) * n);
    if (c->fmax > 0 || c->fmin > 0)
        for (int o = 0; o < nout; o++) {
            or_butterworth(vo + (long)o * nrec * NT, c->fmin, c->fmax, c->dt, NT, tmax, nrec, 6);
            or_butterworth(vi + (long)o * nrec * NT, c->fmin, c->fmax, c->dt, NT, tmax, nrec, 6);
        }
    rms[0] = rms[1] = 0;
    for (int g = 0; g < nrec; g++)
        for (int o = 0; o < nout; o++) {
            const float *a = vo + ((long)o * nrec + g) * NT, *b = vi + ((long)o * nrec + g) * NT;
            for (int f = 0; f < nf; f++) {
                double rr = 0, ri = 0, nr = 0, ni = 0;
                for (int t = 0; t < tmax; t++) {
                    const double ph = 2.0 * PI * (double)(((long)kb[f] * t) % nfft) / nfft;
                    const double cs = cos(ph), sn = sin(ph);
                    const float r = -a[t] + b[t];
                    rr += r * cs;
                    ri -= r * sn;
                    nr += b[t] * cs;
                    ni -= b[t] * sn;
                }
                rms[0] += rr * rr + ri * ri;
                rms[1] += nr * nr + ni * ni;
            }
        }
    free(vo);
    free(vi);
    return 0;
}

int or_gradient_dft(const or_csts *c, const float *p0, const float *p1, const float *rho,
                    const float *taup, const float *taus, int nsrc, const float *src_pos,
                    const float *src, int nrec, const float *rec_pos, const float *din,
                    const float *res_in, int gradfix, int nf, const float *freqs, float *out,
                    double *grad, double *rms, float *dft_out, char *err, int errlen) {
    sim S, A;
    char dummy[8];
    if (!err) {
        err = dummy;
        errlen = sizeof(dummy);
    }
    if (c->ND != 2 && c->ND != 3) {
        snprintf(err, errlen, "the DFT gradient needs the elastic or viscoelastic case (ND 2 or 3)");
        return 1;
    }
    if (c->param_type == 3 && c->L == 0) {
        snprintf(err, errlen, "param_type 3 needs a viscoelastic model (L > 0)");
        return 1;
    }
    int dtnyq, ntnyq;
    float df, bins[OR_MAXF];
    const int smp = nf > 0 && nf <= OR_MAXF
                        ? or_dft_sampling(c, nf, freqs, &dtnyq, &ntnyq, &df, bins) : 1;
    if (smp == 1) {
        snprintf(err, errlen, "gradfreqs: 1 to %d positive frequencies", OR_MAXF);
        return 1;
    }
    if (smp == 2) {
        snprintf(err, errlen, "gradfreqs below the frequency step of the record (bin 0)");
        return 1;
    }
    if (sim_init(&S, c, err, errlen)) {
        sim_free(&S);
        return 1;
    }
#ifdef _OPENMP
    if (c->nthreads > 0) omp_set_num_threads(c->nthreads);
#endif
    or_transform(c, p0, p1, rho, taup, taus, S.M, S.mu, S.rho, S.rip, S.rjp, S.rkp,
                 S.muipkp, S.muipjp, S.mujpkp, S.taup, S.taus, S.tausipkp,
                 S.tausipjp, S.tausjpkp);
    float dts;
    if (or_check_stability(c, S.M, S.mu, S.rho, &dts)) {
        snprintf(err, errlen, "Time step too large, to be stable, set dt<%f", dts);
        sim_free(&S);
        return 1;
    }
    if (check_positions(&S, nsrc, src_pos, 5, err, errlen) ||
        check_positions(&S, nrec, rec_pos, 8, err, errlen)) {
        sim_free(&S);
        return 1;
    }
    const int tmax = c->tmax > 0 ? c->tmax : c->NT, tmin = c->tmin, nout = or_nout(c);
    const int nq = S.ND == 3 ? 9 + 6 * S.L : 5 + 3 * S.L;
    const size_t nF = (size_t)nq * nf * S.nv * 2;
    float *Ff = calloc(nF, sizeof(float)), *Fa = calloc(nF, sizeof(float));
    float fact[2 * OR_MAXF];
    memset(out, 0, sizeof(float) * (size_t)nout * nrec * c->NT);
    /* forward: the DFT of the state before the sources of each sampled step */
    for (int t = 0; t < tmax; t++) {
        if (t >= tmin && (t - tmin) % dtnyq == 0) {
            dft_fact(c, nf, bins, dtnyq, ntnyq, (t - tmin) / dtnyq, fact);
            dft_acc(&S, Ff, nf, fact);
        }
        sources(&S, t, 1, nsrc, src_pos, src);
        if (S.ND == 3) {
            update_v3D(&S);
            update_s3D(&S);
            if (c->freesurf) surface3D(&S, 1);
        } else {
            update_v2D(&S);
            update_s2D(&S);
            if (c->freesurf) surface2D(&S, 1);
        }
        varsout(&S, t, nrec, rec_pos, out);
    }
    if (dft_out) memcpy(dft_out, Ff, nF * sizeof(float));
    size_t nres = (size_t)nout * nrec * c->NT;
    float *res = calloc(nres ? nres : 1, sizeof(float));
    double rr[2] = {0, 0};
    if (res_in) memcpy(res, res_in, nres * sizeof(float));
    else {
        or_residuals(c, nout, nrec, out, din, res, rr, gradfix);
        if (or_dft_rms(c, nout, nrec, out, din, nf, freqs, rr)) {
            snprintf(err, errlen, "gradfreqs above the Nyquist frequency of the record");
            free(Ff);
            free(Fa);
            free(res);
            sim_free(&S);
            return 1;
        }
    }
    res_scale(&S, nrec, rec_pos, res, gradfix & OR_FIX_RES_SCALE);
    if (rms) {
        rms[0] = rr[0];
        rms[1] = rr[1];
    }
    if (adj_alloc(&A, &S, err, errlen)) {
        free(Ff);
        free(Fa);
        free(res);
        sim_free(&S);
        return 1;
    }
    if (!(gradfix & OR_FIX_ADJ_CPML)) adj_share_cpml(&A, &S);
    /* adjoint: the adjoint wavefield alone, its DFT after the updates */
    for (int t = tmax - 1; t > tmin; t--) {
        if (c->freesurf) {
            if (S.ND == 3) surface_adj3D(&A, gradfix & OR_FIX_SURF_SYY);
            else surface_adj2D(&A);
        }
        residuals_inject(&A, t, nrec, rec_pos, res);
        if (S.ND == 3) {
            update_s3D(&A);
            update_v3D(&A);
        } else {
            update_s2D(&A);
            update_v2D(&A);
        }
        if ((t - tmin) % dtnyq == 0) {
            dft_fact(c, nf, bins, dtnyq, ntnyq, (t - tmin) / dtnyq, fact);
            dft_acc(&A, Fa, nf, fact);
        }
    }
    memset(grad, 0, sizeof(double) * 5 * S.np);
    dft_grad(&S, Ff, Fa, nf, bins, df, ntnyq, grad);
    free(Ff);
    free(Fa);
    free(res);
    adj_free(&A);
    sim_free(&S);
    return 0;
}
