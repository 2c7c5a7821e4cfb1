// Host side of the gradient path: the boundary shell geometry, L2
// residuals, their scaling and the final parameter transform of the
// gradient.  Arithmetic follows the reference's C host (mixed float/double
// exactly as written there).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

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

// ---- misfit (var_res_raw and rtm_res, src/residuals.c:24-450) ------------

// Smallest n >= m whose prime factors are 2, 3 and 5 (the FFT length the
// reference's filter uses).
int fft_fast_size(int m) {
    for (int n = std::max(m, 1);; ++n) {
        int k = n;
        for (int p : {2, 3, 5})
            while (k % p == 0) k /= p;
        if (k == 1) return n;
    }
}

namespace {

using cplx = std::complex<double>;

// Mixed-radix decimation-in-time DFT of x[0], x[s], ..., x[(n-1)s] into
// out[0..n-1], for n with factors 2, 3, 5.  w holds exp(sign 2 pi i k / N)
// for the top-level length N; this level's length n = N / ws.
void fft_rec(const cplx *x, long s, int n, cplx *out, const std::vector<cplx> &w, long ws,
             std::vector<cplx> &tmp) {
    if (n == 1) {
        out[0] = x[0];
        return;
    }
    const int p = n % 2 == 0 ? 2 : n % 3 == 0 ? 3 : 5;
    const int m = n / p;
    const long N = (long)w.size();
    for (int r = 0; r < p; ++r) fft_rec(x + r * s, s * p, m, out + r * m, w, ws * p, tmp);
    // X[k + q m] = sum_r W_n^{r k} W_p^{r q} Y_r[k]
    cplx y[5];
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) y[r] = out[r * m + k] * w[(long)r * k * ws % N];
        for (int q = 0; q < p; ++q) {
            cplx acc = 0;
            for (int r = 0; r < p; ++r) acc += y[r] * w[(long)r * q * m * ws % N];
            tmp[q * m + k] = acc;
        }
    }
    for (int k = 0; k < n; ++k) out[k] = tmp[k];
}

void fft(std::vector<cplx> &a, int sign) {
    const int n = (int)a.size();
    std::vector<cplx> w(n), out(n), tmp(n);
    for (int k = 0; k < n; ++k) w[k] = std::polar(1.0, sign * 2.0 * M_PI * k / n);
    fft_rec(a.data(), 1, n, out.data(), w, 1, tmp);
    a.swap(out);
}

}  // namespace

// butterworth (src/butterworth.c:24-86): zero-phase band-pass of order
// `order` applied in the frequency domain to each of ntrace traces of
// data[n*NT + t], t < tmax, over nfft = fft_fast_size(tmax) samples whose
// tail repeats the start of the trace.
void butterworth(float *data, float fcl, float fch, float dt, int NT, int tmax, int ntrace,
                 int order) {
    const int nfft = fft_fast_size(tmax);
    std::vector<float> H(nfft / 2 + 1, 1.0f);
    for (int t = 0; t < nfft / 2 + 1; ++t) {
        const float freq = t / (dt * nfft);
        if (fcl > 0) H[t] *= 1. - std::sqrt(1. / (1. + std::pow(freq / fcl, 2 * order)));
        if (fch > 0) H[t] *= std::sqrt(1. / (1. + std::pow(freq / fch, 2 * order)));
    }
    std::vector<cplx> a(nfft);
    for (int n = 0; n < ntrace; ++n) {
        float *d = data + (size_t)n * NT;
        for (int t = 0; t < nfft; ++t) a[t] = cplx(t < tmax ? d[t] : d[t - tmax], 0.0);
        fft(a, -1);
        for (int k = 0; k < nfft; ++k) a[k] *= (double)H[k <= nfft / 2 ? k : nfft - k];
        fft(a, 1);
        for (int t = 0; t < tmax; ++t) d[t] = (float)(a[t].real() / nfft);
    }
}

// var_res_raw (src/residuals.c:58-450) with back_prop_type 1, or rtm_res
// (:24-55) with restype 1.  out and din are [nout][nrec][NT]; the data are
// first band-passed (fmin/fmax), then scaled by trace rms (scalerms,
// scaleshot, scalermsnorm) over [tmin, tmax); res and the returned
// rms = sum (in - out)^2, rmsnorm = sum in^2 follow for t < tmax, with the
// reference's float accumulators and double pow.
double residuals(const seis_csts &c, int nout, int nrec, const float *out, const float *din,
                 float *res, double *rmsnorm) {
    const int NT = c.NT, tmax = c.tmax > 0 ? c.tmax : c.NT, tmin = c.tmin;
    const size_t n = (size_t)nout * nrec * NT;
    if (rmsnorm) *rmsnorm = 0;
    if (c.restype == 1) {
        for (int o = 0; o < nout; ++o)
            for (int g = 0; g < nrec; ++g)
                for (int t = 0; t < tmax; ++t) {
                    const size_t q = ((size_t)o * nrec + g) * NT + t;
                    res[q] = din[q];
                }
        return 0;
    }
    std::vector<float> vo(out, out + n), vi(din, din + n);
    if (c.fmax > 0 || c.fmin > 0)
        for (int o = 0; o < nout; ++o) {
            butterworth(vo.data() + (size_t)o * nrec * NT, c.fmin, c.fmax, c.dt, NT, tmax, nrec, 6);
            butterworth(vi.data() + (size_t)o * nrec * NT, c.fmin, c.fmax, c.dt, NT, tmax, nrec, 6);
        }
    auto at = [&](std::vector<float> &v, int o, int g, int t) -> float & {
        return v[((size_t)o * nrec + g) * NT + t];
    };
    const bool scaled = c.scalerms || c.scaleshot || c.scalermsnorm;
    const bool norm = c.scalerms || c.scalermsnorm;
    std::vector<float> rs(nrec, 1.0f), rs0(nrec, 1.0f), rn(nrec, 0.0f);
    float ws = 0;
    if (scaled) {
        for (int g = 0; g < nrec; ++g) {
            float a = 0, b = 0;
            for (int o = 0; o < nout; ++o)
                for (int t = tmin; t < tmax; ++t) {
                    a += std::pow(at(vo, o, g, t), 2);
                    b += std::pow(at(vi, o, g, t), 2);
                }
            rs[g] = std::sqrt(1. / a);
            rs0[g] = std::sqrt(1. / b);
        }
        if (c.scaleshot) {
            ws = 0;
            for (int g = 0; g < nrec; ++g) ws += 1 / std::pow(rs0[g], 2);
            ws = 1 / std::sqrt((double)ws);   // C's double sqrt
            for (int g = 0; g < nrec; ++g) rs[g] = rs0[g] = ws;
        }
        if (c.scalerms) {
            ws = 0;
            for (int g = 0; g < nrec; ++g) ws += 1 / std::pow(rs0[g], 2);
            ws = 1 / std::sqrt((double)ws);   // C's double sqrt
            for (int g = 0; g < nrec; ++g) {
                rs[g] *= ws / rs0[g];
                rs0[g] = ws;
            }
        }
        for (int g = 0; g < nrec; ++g)
            for (int o = 0; o < nout; ++o)
                for (int t = tmin; t < tmax; ++t) {
                    at(vo, o, g, t) *= rs[g];
                    at(vi, o, g, t) *= rs0[g];
                }
    }
    if (norm)
        for (int g = 0; g < nrec; ++g) {
            float a = 0;
            for (int o = 0; o < nout; ++o)
                for (int t = tmin; t < tmax; ++t)
                    a += at(vo, o, g, t) * (at(vo, o, g, t) - at(vi, o, g, t));
            if (c.scalerms) a *= std::pow(ws / rs0[g] / ws, 2);
            rn[g] = 1.0 - a;
        }
    // normalised misfit: the reference weighs the trace by 1 - o.d
    // (rn = o.d for unit traces); the derivative of 1/2 |o/|o| - d/|d||^2
    // weighs it by o.d (FIX_RMSNORM)
    const bool fixnorm = (c.gradfix & FIX_RMSNORM) && c.scalermsnorm && !c.scalerms;
    double rms = 0, rnorm = 0;
    for (int g = 0; g < nrec; ++g)
        for (int t = 0; t < tmax; ++t)
            for (int o = 0; o < nout; ++o) {
                const float vout = at(vo, o, g, t), vin = at(vi, o, g, t);
                rms += std::pow(-vout + vin, 2);
                rnorm += std::pow(vin, 2);
                float &r = res[((size_t)o * nrec + g) * NT + t];
                if (fixnorm) r = rs[g] * (rn[g] * vout - vin);
                else r = norm ? rs[g] * ((1.0 - rn[g]) * vout - vin) : rs[g] * (vout - vin);
            }
    if (rmsnorm) *rmsnorm = rnorm;
    return rms;
}

double dft_rms(const seis_csts &c, int nout, int nrec, const float *out, const float *din,
               const std::vector<float> &freqs, double *rmsnorm) {
    const int NT = c.NT, tmax = c.tmax > 0 ? c.tmax : c.NT, nfft = fft_fast_size(tmax);
    const size_t n = (size_t)nout * nrec * NT;
    std::vector<float> vo(out, out + n), vi(din, din + n);
    if (c.fmax > 0 || c.fmin > 0)
        for (int o = 0; o < nout; ++o) {
            butterworth(vo.data() + (size_t)o * nrec * NT, c.fmin, c.fmax, c.dt, NT, tmax, nrec, 6);
            butterworth(vi.data() + (size_t)o * nrec * NT, c.fmin, c.fmax, c.dt, NT, tmax, nrec, 6);
        }
    std::vector<int> kb;
    for (float f : freqs) kb.push_back((int)(f * nfft * c.dt + 1));
    const double pi = 3.141592653589793238462643383279502884197169;
    double rms = 0, rnorm = 0;
    for (int g = 0; g < nrec; ++g)
        for (int o = 0; o < nout; ++o) {
            const float *a = vo.data() + ((size_t)o * nrec + g) * NT;
            const float *b = vi.data() + ((size_t)o * nrec + g) * NT;
            for (int k : kb) {
                double rr = 0, ri = 0, nr = 0, ni = 0;
                for (int t = 0; t < tmax; ++t) {
                    const double ph = 2.0 * pi * (double)(((long)k * t) % nfft) / nfft;
                    const double cs = std::cos(ph), sn = std::sin(ph);
                    const float r = -a[t] + b[t];
                    rr += r * cs;
                    ri -= r * sn;
                    nr += b[t] * cs;
                    ni -= b[t] * sn;
                }
                rms += rr * rr + ri * ri;
                rnorm += nr * nr + ni * ni;
            }
        }
    if (rmsnorm) *rmsnorm = rnorm;
    return rms;
}

// res_scale (src/residuals.c:452-570), FP32 storage (par_scale 0).
// Velocity outputs are scaled by dt / (dh / (dt * buoyancy)), with the
// buoyancy read at (z=x, y=x, x=x) of the receiver's x index as the reference
// does, or at the receiver's own cell with at_receiver; pressure by
// -2 (M - mu) dh/dt * dt.
// The factor of every (output, receiver) trace; 1 where nothing is scaled.
// Applying it as (float)((double)r * f) equals both of the reference's
// updates: r *= 1.0 / parscal * dt (double factor) for velocities and
// r *= parscal * dt (a float factor, whose product with r is exact in
// double) for pressure.
void res_scale_factors(const seis_csts &c, const RefParams &p, const std::vector<std::string> &outn,
                       int nrec, const float *rec_pos, double *fac, int at_receiver) {
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
        for (int g = 0; g < nrec; ++g) {
            double &f = fac[o * nrec + g];
            f = 1.0;
            if (!par) continue;
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
            if (isv) {
                const float parscal = 1.0 / (*par)[pos] * c.dh / c.dt * powf(2, 0);
                f = 1.0 / parscal * c.dt;
            } else {
                const float parscal = -2.0 * ((*par)[pos] - p.mu[pos]) * c.dh / c.dt * powf(2, -0);
                f = (double)(parscal * c.dt);
            }
        }
    }
}

void res_scale(const seis_csts &c, const RefParams &p, const std::vector<std::string> &outn,
               int nrec, const float *rec_pos, float *res, int at_receiver) {
    const int NT = c.NT, tmax = c.tmax > 0 ? c.tmax : c.NT;
    std::vector<double> fac(outn.size() * (size_t)nrec);
    res_scale_factors(c, p, outn, nrec, rec_pos, fac.data(), at_receiver);
    for (size_t k = 0; k < fac.size(); ++k) {
        if (fac[k] == 1.0) continue;
        float *r = &res[k * NT];
        for (int t = 0; t < tmax; ++t) r[t] = (float)((double)r[t] * fac[k]);
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

// The approximate Hessian through the same parameter transform
// (src/calc_grad.c:1028-1094): not divided by dt; the density term holds
// only the chain-rule contributions of M and mu (the reference's kernels
// leave it at zero).
void transf_hess(const seis_csts &c, const RefParams &p, const std::vector<float> &Ms,
                 const std::vector<float> &mus, const double *hM, const double *hmu,
                 float *out) {
    const size_t n = Ms.size();
    const bool elastic = !mus.empty();
    for (size_t i = 0; i < n; ++i) {
        const double rho = 1.0 / p.rho[i] * c.dt / c.dh;
        const double M = (double)Ms[i] * c.dh / c.dt;
        const double mu = elastic ? (double)mus[i] * c.dh / c.dt : 0.0;
        double hr = 0, hm = hM[i], hu = elastic ? hmu[i] : 0.0;
        if (c.param_type == 0 || c.param_type == 2) {
            hr = hr + M / rho * hm;
            if (mu > 0) hr = hr + mu / rho * hu;
            if (c.param_type == 0) {
                hm = 2.0 * std::sqrt(rho * M) * hm;
                hu = 2.0 * std::sqrt(rho * mu) * hu;
            } else {
                hm = 2.0 * std::sqrt(M / rho) * hm;
                hu = 2.0 * std::sqrt(mu / rho) * hu;
            }
        }
        out[i] = (float)hm;
        out[n + i] = (float)hu;
        out[2 * n + i] = (float)hr;
    }
}

}  // namespace seis
