// DFT-domain gradient (back_prop_type 2): spectra of the wavefields and
// their correlation.
//
// The forward run, and later the adjoint run (the adjoint wavefield alone,
// no backpropagation), add every DTNYQ-th state into running DFTs at the
// gradient frequencies (savefreqs, src/automatic_kernels.c:1030-1128):
//   F[q][f][cell] += fact[thist][f] * u_q[cell],
//   fact = DTNYQ dt e^{-i pi 2 bin_f thist / NTNYQ} (a host-built table).
// One lane per interior cell, all arrays q and frequencies f in the lane,
// each spectrum a float2 in the padded layout; the kernel streams the
// fields once and each spectrum once in and out (HBM-bound, run every
// DTNYQ steps).
//
// The gradient kernel then correlates the forward and adjoint spectra cell
// by cell (src/calc_grad.c:608-968): the memory-variable spectra are
// integrated out of the stress spectra, the scalar products dot[0..8]
// formed in float as the reference does, and weighted by the 24
// coefficients of the parameterisation (src/calc_grad.c:111-376, here in
// double) into the double accumulators.
#include <hip/hip_runtime.h>

#include "fd_common.hpp"
#include "seis_internal.hpp"

namespace seis {
namespace {

__device__ __forceinline__ bool dft_cell(const Layout &g, int nbz, unsigned nb, int &z, int &y,
                                         int &x) {
    const unsigned tile = blockIdx.x;
    if (tile >= nb) return false;
    const int bz = tile % nbz;
    const unsigned row = (tile / nbz) * blockDim.y + threadIdx.y;   // y + NY x
    z = bz * BZ + threadIdx.x;
    y = (int)(row % (unsigned)g.NY);
    x = (int)(row / (unsigned)g.NY);
    return z < g.NZ && x < g.NX;
}

__global__ __launch_bounds__(256) void k_dft_acc(DftArrays q, float2 *F, const float2 *fact,
                                                 int nf, Layout g, int nbz, unsigned nb) {
    int z, y, x;
    if (!dft_cell(g, nbz, nb, z, y, x)) return;
    const long i = g.idx(z, y, x), nv = g.nfield;
    for (int a = 0; a < q.n; ++a) {
        const float u = q.p[a][i];
        float2 *Fa = F + (long)a * nf * nv + i;
        for (int f = 0; f < nf; ++f) {
            const float2 c = fact[f];
            float2 s = Fa[(long)f * nv];
            s.x += c.x * u;
            s.y += c.y * u;
            Fa[(long)f * nv] = s;
        }
    }
}

struct cf2 {
    float x, y;
};
__device__ __forceinline__ cf2 cadd(cf2 a, cf2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf2 csub(cf2 a, cf2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cf2 cadd3(cf2 a, cf2 b, cf2 c) {
    return {a.x + b.x + c.x, a.y + b.y + c.y};
}
__device__ __forceinline__ cf2 csub3(cf2 a, cf2 b, cf2 c) {
    return {a.x - b.x - c.x, a.y - b.y - c.y};
}
// spectrum of the time integral, a / (i w)
__device__ __forceinline__ cf2 cinteg(cf2 a, float w) { return {a.y / w, -a.x / w}; }
// Im(conj(b) a)
__device__ __forceinline__ float citreal(cf2 a, cf2 b) { return a.y * b.x - a.x * b.y; }
// memory-variable correlation: ts Re(conj(b) a) + Im(conj(a) b) / w
__device__ __forceinline__ float crm(cf2 a, cf2 b, float ts, float w) {
    return ts * (a.x * b.x + a.y * b.y) + (a.x * b.y - a.y * b.x) / w;
}

// The 24 coefficients of one cell in double (grad_coefvisc_0..3,
// grad_coefelast_0..2 of src/calc_grad.c:111-376).  M, mu, rho are the
// transformed parameters (rho the buoyancy dt/(dh rho)); the square roots
// are the float ones of the reference.
__device__ void dft_coef(double *c, int pt, bool visco, float M, float mu, float tp, float ts,
                         float rho, float ND, float L, float al) {
    for (int k = 0; k < 24; ++k) c[k] = 0;
    const double Md = M, mud = mu, rd = rho, n3 = (ND + 1.0) / 3.0, Lal = L - al;
    const double kp = 1.0 + L * tp, ks = 1.0 + L * ts, ap = 1.0 + al * tp, as = 1.0 + al * ts;
    const double mu2 = mud * mud, as2 = as * as, ap2 = ap * ap;
    if (!visco) {
        const double e = ND * Md - 2.0 * (ND - 1.0) * mud, F = e * e;
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
        const double s = pt == 0 ? 1.0 : -1.0;
        c[16] = pt == 0 ? Md / rd / F : -Md / rd * kp * ap * as2;
        c[18] = s * (mud / rd / mu2);
        c[19] = s * (mud / rd * n3 / F);
        c[20] = s * (mud / rd / (2 * ND * mu2));
        return;
    }
    const double g1 = ND * Md * kp * as - 2.0 * (ND - 1.0) * mud * ks * ap;
    const double g2 = ND * Md * tp * as - 2.0 * (ND - 1.0) * mud * ts * ap;
    const double f1 = g1 * g1, f2 = g2 * g2;
    const double u[8] = {kp * ap * as2 / f1,         tp * ap * as2 / f2,
                         as / (mu2 * ks),            n3 * ks * as * ap2 / f1,
                         as / (2.0 * ND * mu2 * ks), as / (mu2 * ts),
                         n3 * ts * as * ap2 / f2,    as / (2.0 * ND * mu2 * ts)};
    const double t[8] = {Md * Lal * as2 / f1,
                         Md * as2 / f2,
                         Lal / (mud * ks * ks),
                         n3 * mud * Lal * ap2 / f1,
                         Lal / (2 * ND * mud * ks * ks),
                         1.0 / (mud * ts * ts),
                         n3 * mud * ap2 / f2,
                         1.0 / (2 * ND * mud * ts * ts)};
    double A = 1, B = 1;
    if (pt == 0 || pt == 3) {
        A = 2.0 * sqrtf(rho * M);
        B = 2.0 * sqrtf(rho * mu);
    } else if (pt == 2) {
        A = 2.0 * sqrtf(M / rho);
        B = 2.0 * sqrtf(mu / rho);
    }
    for (int k = 0; k < 8; ++k) c[k] = (k < 2 ? A : B) * u[k];
    for (int k = 0; k < 8; ++k) c[8 + k] = t[k];
    if (pt == 3) {
        // taup is the Q parameter: its chain through vp = sqrt(M/rho)(1+taup)
        const double sm = sqrtf(M / rho), R = tp / sm, Rp = (1.0 + tp) / sm;
        const double w[8] = {Md * Lal * as2, Md * as2, Lal / (mud * ks * ks),
                             n3 * mud * Lal * ap2, Lal / (2 * ND * mud * ks * ks),
                             1.0 / (mud * ts * ts), n3 * mud * ap2, 1.0 / (2 * ND * mud * ts * ts)};
        const double dv[8] = {f1, f2, 1, f1, 1, 1, f2, 1};
        const double P[8] = {A * kp * ap * as2, A * tp * ap * as2, B * as / (mu2 * ks),
                             B * n3 * ks * as * ap2, B * as / (2.0 * ND * mu2 * ks),
                             B * as / (mu2 * ts), B * n3 * ts * as * ap2,
                             B * as / (2.0 * ND * mu2 * ts)};
        for (int k = 0; k < 8; ++k) {
            c[k] = (P[k] - R * w[k]) / dv[k];
            c[8 + k] = (-P[k] + Rp * w[k]) / dv[k];
        }
    }
    if (pt == 1) return;
    const double s = pt == 2 ? -1.0 : 1.0, m_r = Md / rd, u_r = mud / rd;
    c[16] = s * m_r * u[0];
    c[17] = s * m_r * u[1];
    for (int k = 2; k < 8; ++k) c[16 + k] = s * u_r * u[k];
}

__global__ __launch_bounds__(256) void k_dft_grad(DftGrad d, Layout g, int nbz, unsigned nb) {
    int z, y, x;
    if (!dft_cell(g, nbz, nb, z, y, x)) return;
    const long i = g.idx(z, y, x), nv = g.nfield;
    const long m = ((long)x * g.NY + y) * g.NZ + z;     // unpadded, z fastest
    const long np = (long)g.NX * g.NY * g.NZ;
    const bool nd3 = g.nd == 3;
    const int L = d.L, nf = d.nf;
    const float M = d.par[m], mu = d.par[np + m];
    const float tp = L > 0 ? d.par[2 * np + m] : 0.f, ts = L > 0 ? d.par[3 * np + m] : 0.f;
    const float rho = d.par[4 * np + m];
    double cf[24];
    dft_coef(cf, d.pt, L > 0, M, mu, tp, ts, rho, nd3 ? 3.f : 2.f, (float)L, d.al);
    if (mu < 1) {
        for (int n = 2; n < 8; ++n) cf[n] = 0;
        for (int n = 10; n < 16; ++n) cf[n] = 0;
        for (int n = 18; n < 24; ++n) cf[n] = 0;
    }
    double gM = 0, gmu = 0, grho = 0, gtp = 0, gts = 0;
    const int ntnyq = d.ntnyq;
    for (int f = 0; f < nf; ++f) {
        const float w = d.w[f];
        auto A = [&](int q) -> cf2 {
            const float2 v = ((const float2 *)d.Ff)[((long)q * nf + f) * nv + i];
            return {v.x, v.y};
        };
        auto B = [&](int q) -> cf2 {
            const float2 v = ((const float2 *)d.Fa)[((long)q * nf + f) * nv + i];
            return {v.x, v.y};
        };
        double dot[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (nd3) {
            cf2 sxx = A(3), syy = A(4), szz = A(5), sxy = A(6), syz = A(7), sxz = A(8);
            cf2 sxxr = B(3), syyr = B(4), szzr = B(5), sxyr = B(6), syzr = B(7), sxzr = B(8);
            for (int l = 0; l < L; ++l) {
                const cf2 rxx = A(9 + l), ryy = A(9 + L + l), rzz = A(9 + 2 * L + l),
                          rxy = A(9 + 3 * L + l), ryz = A(9 + 4 * L + l), rxz = A(9 + 5 * L + l);
                const cf2 rxxr = B(9 + l), ryyr = B(9 + L + l), rzzr = B(9 + 2 * L + l),
                          rxyr = B(9 + 3 * L + l), ryzr = B(9 + 4 * L + l),
                          rxzr = B(9 + 5 * L + l);
                const float tsl = d.tausig[l];
                sxx = csub(sxx, cinteg(rxx, w));
                szz = csub(szz, cinteg(rzz, w));
                syy = csub(syy, cinteg(ryy, w));
                sxz = csub(sxz, cinteg(rxz, w));
                sxy = csub(sxy, cinteg(rxy, w));
                syz = csub(syz, cinteg(ryz, w));
                sxxr = csub(sxxr, cinteg(rxxr, w));
                szzr = csub(szzr, cinteg(rzzr, w));
                syyr = csub(syyr, cinteg(ryyr, w));
                sxzr = csub(sxzr, cinteg(rxzr, w));
                sxyr = csub(sxyr, cinteg(rxyr, w));
                syzr = csub(syzr, cinteg(ryz, w));   // the forward ryz, as the reference
                const cf2 rsum = cadd3(rxx, ryy, rzz), rsumr = cadd3(rxxr, ryyr, rzzr);
                const cf2 rdev = csub3(rxx, ryy, rzz);   // all three deviators, as the reference
                dot[1] += crm(rsumr, rsum, tsl, w) / ntnyq;
                dot[5] += (crm(rxyr, rxy, tsl, w) + crm(rxzr, rxz, tsl, w) +
                           crm(ryzr, ryz, tsl, w)) /
                          ntnyq;
                dot[6] = dot[1];
                dot[7] += (crm(rxxr, rdev, tsl, w) + crm(ryyr, rdev, tsl, w) +
                           crm(rzzr, rdev, tsl, w)) /
                          ntnyq;
            }
            const cf2 ssum = cadd3(sxx, syy, szz), ssumr = cadd3(sxxr, syyr, szzr);
            const cf2 dxx = csub3(sxx, syy, szz), dyy = csub3(syy, sxx, szz),
                      dzz = csub3(szz, sxx, syy);
            dot[0] = w * citreal(ssumr, ssum) / ntnyq;
            dot[2] = w * (citreal(sxyr, sxy) + citreal(sxzr, sxz) + citreal(syzr, syz)) / ntnyq;
            dot[3] = dot[0];
            dot[4] = w * (citreal(sxxr, dxx) + citreal(syyr, dyy) + citreal(szzr, dzz)) / ntnyq;
            dot[8] = w * (citreal(B(0), A(0)) + citreal(B(1), A(1)) + citreal(B(2), A(2))) / ntnyq;
        } else {
            cf2 sxx = A(2), szz = A(3), sxz = A(4);
            cf2 sxxr = B(2), szzr = B(3), sxzr = B(4);
            for (int l = 0; l < L; ++l) {
                const cf2 rxx = A(5 + l), rzz = A(5 + L + l), rxz = A(5 + 2 * L + l);
                const cf2 rxxr = B(5 + l), rzzr = B(5 + L + l), rxzr = B(5 + 2 * L + l);
                const float tsl = d.tausig[l];
                sxx = csub(sxx, cinteg(rxx, w));
                szz = csub(szz, cinteg(rzz, w));
                sxz = csub(sxz, cinteg(rxz, w));
                sxxr = csub(sxxr, cinteg(rxxr, w));
                szzr = csub(szzr, cinteg(rzzr, w));
                sxzr = csub(sxzr, cinteg(rxzr, w));
                const cf2 rsum = cadd(rxx, rzz), rsumr = cadd(rxxr, rzzr);
                const cf2 dxx = csub(rxx, rzz), dzz = csub(rzz, rxx);
                dot[1] += crm(rsumr, rsum, tsl, w) / ntnyq;
                dot[5] += (crm(rxzr, rxz, tsl, w)) / ntnyq;
                dot[6] = dot[1];
                dot[7] += (crm(rxxr, dxx, tsl, w) + crm(rzzr, dzz, tsl, w)) / ntnyq;
            }
            const cf2 ssum = cadd(sxx, szz), ssumr = cadd(sxxr, szzr);
            const cf2 dxx = csub(sxx, szz), dzz = csub(szz, sxx);
            dot[0] = w * citreal(ssumr, ssum) / ntnyq;
            dot[2] = w * (citreal(sxzr, sxz)) / ntnyq;
            dot[3] = dot[0];
            dot[4] = w * (citreal(sxxr, dxx) + citreal(szzr, dzz)) / ntnyq;
            dot[8] = w * (citreal(B(0), A(0)) + citreal(B(1), A(1))) / ntnyq;
        }
        gM += -cf[0] * dot[0] + cf[1] * dot[1];
        gmu += -cf[2] * dot[2] + cf[3] * dot[3] - cf[4] * dot[4] + cf[5] * dot[5] -
               cf[6] * dot[6] + cf[7] * dot[7];
        if (L > 0) {
            gtp += -cf[8] * dot[0] + cf[9] * dot[1];
            gts += -cf[10] * dot[2] + cf[11] * dot[3] - cf[12] * dot[4] + cf[13] * dot[5] -
                   cf[14] * dot[6] + cf[15] * dot[7];
        }
        grho += -dot[8] + cf[16] * dot[0] - cf[17] * dot[1] + cf[18] * dot[2] -
                cf[19] * dot[3] + cf[20] * dot[4] - cf[21] * dot[5] + cf[22] * dot[6] -
                cf[23] * dot[7];
    }
    d.gM[i] += gM;
    d.gmu[i] += gmu;
    d.grho[i] += grho;
    if (L > 0 && d.gtp) {
        d.gtp[i] += gtp;
        d.gts[i] += gts;
    }
}

void dft_grid(const Layout &g, int &nbz, unsigned &nb) {
    nbz = (g.NZ + BZ - 1) / BZ;
    const unsigned rows = (unsigned)g.NY * (unsigned)g.NX;
    nb = (unsigned)nbz * ((rows + 3) / 4);
}

}  // namespace

void launch_dft_acc(const DftArrays &q, float *F, const float *fact, int nf, const Layout &g,
                    void *stream) {
    int nbz;
    unsigned nb;
    dft_grid(g, nbz, nb);
    hipLaunchKernelGGL(k_dft_acc, dim3(nb), dim3(BZ, 4), 0, (hipStream_t)stream, q, (float2 *)F,
                       (const float2 *)fact, nf, g, nbz, nb);
}

void launch_dft_grad(const DftGrad &d, const Layout &g, void *stream) {
    int nbz;
    unsigned nb;
    dft_grid(g, nbz, nb);
    hipLaunchKernelGGL(k_dft_grad, dim3(nb), dim3(BZ, 4), 0, (hipStream_t)stream, d, g, nbz, nb);
}

}  // namespace seis
