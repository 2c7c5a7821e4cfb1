// Host side of the gradient path: the boundary shell geometry, L2
// residuals, their scaling and the final parameter transform of the
// gradient.  Arithmetic follows the reference's C host (mixed float/double
// exactly as written there).
#include <cmath>
#include <cstring>

#include "seis_internal.hpp"

namespace seis {

// NBND geometry of src/Init_OpenCL.c:988-1038 and src/savebnd2D.cl:43-80
// for one device with FP32 storage.  The shell holds exactly the cells that
// savebnd enumerates (with a free surface the reference allocates FDOH*2*FDOH
// more entries than it fills).
Shell make_shell(const Layout &g, int nab, int freesurf) {
    Shell s;
    s.nd = g.nd;
    s.F = g.H;
    s.NAB = nab;
    const int F = g.H;
    s.NXbnd = g.NX - 2 * nab;
    s.NYbnd = g.nd == 3 ? g.NY - 2 * nab : 0;
    s.NZbnd = freesurf ? g.NZ - nab : g.NZ - 2 * nab;
    s.lbnd = F + nab;
    s.lbnds = freesurf ? F : F + nab;
    if (g.nd == 3)
        s.n = (long)s.NYbnd * s.NZbnd * 2 * F + (long)s.NZbnd * (s.NXbnd - 2 * F) * 2 * F +
              (long)(s.NYbnd - 2 * F) * (s.NXbnd - 2 * F) * 2 * F;
    else
        s.n = (long)s.NZbnd * F * 2 + (long)(s.NXbnd - 2 * F) * 2 * F;
    return s;
}

// var_res_raw, L2 without filtering or rms scaling (src/residuals.c:58-450):
// res = 1.0f * (out - in) for t < tmax; returns rms = sum (in - out)^2 and
// rmsnorm = sum in^2 (pow in double).  Arrays are [nout][nrec][NT].
double residuals_l2(const seis_csts &c, int nout, int nrec, const float *out, const float *din,
                    float *res, double *rmsnorm) {
    const int NT = c.NT, tmax = c.tmax > 0 ? c.tmax : c.NT;
    double rms = 0, rn = 0;
    for (int g = 0; g < nrec; ++g)
        for (int t = 0; t < tmax; ++t)
            for (int o = 0; o < nout; ++o) {
                const size_t q = ((size_t)o * nrec + g) * NT + t;
                rms += std::pow((double)(-out[q] + din[q]), 2);
                rn += std::pow((double)din[q], 2);
                res[q] = 1.0f * (out[q] - din[q]);
            }
    if (rmsnorm) *rmsnorm = rn;
    return rms;
}

// res_scale (src/residuals.c:452-570), FP32 storage (par_scale 0).
// Velocity outputs are scaled by dt / (dh / (dt * buoyancy)), with the
// buoyancy read at (z=x, y=x, x=x) of the receiver's x index as the reference
// does, or at the receiver's own cell with at_receiver; pressure by
// -2 (M - mu) dh/dt * dt.
void res_scale(const seis_csts &c, const RefParams &p, const std::vector<std::string> &outn,
               int nrec, const float *rec_pos, float *res, int at_receiver) {
    const int NT = c.NT, tmax = c.tmax > 0 ? c.tmax : c.NT;
    const int NZ = c.N[0], NY = c.ND == 3 ? c.N[1] : 1, NX = c.ND == 3 ? c.N[2] : c.N[1];
    const long np = (long)NZ * NY * NX;
    for (size_t o = 0; o < outn.size(); ++o) {
        const std::string &name = outn[o];
        const std::vector<float> *par = nullptr;
        bool isv = false;
        if (name == "vx") par = &p.rip, isv = true;
        else if (name == "vy") par = &p.rjp, isv = true;
        else if (name == "vz") par = &p.rkp, isv = true;
        else if (name == "p" && c.ND != 22) par = &p.M;
        else continue;
        for (int g = 0; g < nrec; ++g) {
            int x, y, z;
            if (isv && !at_receiver) {
                x = y = z = (int)(rec_pos[0 + 8 * g] / c.dh);
            } else {
                x = (int)(rec_pos[0 + 8 * g] / c.dh);
                y = (int)(rec_pos[1 + 8 * g] / c.dh);
                z = (int)(rec_pos[2 + 8 * g] / c.dh);
            }
            const long pos = c.ND == 3 ? (long)x * NZ * NY + (long)y * NZ + z : (long)x * NZ + z;
            if (pos < 0 || pos >= np) continue;
            float *r = &res[((size_t)o * nrec + g) * NT];
            if (isv) {
                const float parscal = 1.0 / (*par)[pos] * c.dh / c.dt * powf(2, 0);
                for (int t = 0; t < tmax; ++t) r[t] *= 1.0 / parscal * c.dt;
            } else {
                const float parscal = -2.0 * ((*par)[pos] - p.mu[pos]) * c.dh / c.dt * powf(2, -0);
                for (int t = 0; t < tmax; ++t) r[t] *= parscal * c.dt;
            }
        }
    }
}

// transf_grad (src/calc_grad.c:975-1110) in double from the double
// accumulators (unpadded, reference layout).  Output [3][N] in param_type
// order: vp|M|Ip, vs|mu|Is, rho.  With rho_fix the direct density term is
// scaled by dh^2/dt^2, the factor between the density kernel's
// v . dv_adjoint and the derivative of the misfit.
// Ms, mus: the moduli of the chain rule (the relaxed ones, or those at f0 for
// viscoelastic runs with FIX_VISCO_F0).
void transf_grad(const seis_csts &c, const RefParams &p, const std::vector<float> &Ms,
                 const std::vector<float> &mus, const double *gM, const double *gmu,
                 const double *grho, float *out, int rho_fix) {
    const size_t n = Ms.size();
    const bool elastic = !mus.empty();
    for (size_t i = 0; i < n; ++i) {
        const double rho = 1.0 / p.rho[i] * c.dt / c.dh;
        const double M = (double)Ms[i] * c.dh / c.dt;
        const double mu = elastic ? (double)mus[i] * c.dh / c.dt : 0.0;
        double gr = grho[i] / c.dt, gm = gM[i] / c.dt, gu = elastic ? gmu[i] / c.dt : 0.0;
        if (rho_fix) gr *= (double)c.dh * c.dh / ((double)c.dt * c.dt);
        if (c.param_type == 0 || c.param_type == 2) {
            gr = gr + M / rho * gm;
            if (mu > 0) gr = gr + mu / rho * gu;
            if (c.param_type == 0) {
                gm = 2.0 * std::sqrt(rho * M) * gm;
                gu = 2.0 * std::sqrt(rho * mu) * gu;
            } else {
                gm = 2.0 * std::sqrt(M / rho) * gm;
                gu = 2.0 * std::sqrt(mu / rho) * gu;
            }
        }
        out[i] = (float)gm;
        out[n + i] = (float)gu;
        out[2 * n + i] = (float)gr;
    }
}

}  // namespace seis
