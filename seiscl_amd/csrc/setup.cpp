// Host-side set-up of a modeling case: FD coefficients, CPML profiles,
// relaxation constants and the material-parameter transforms of the
// reference.  Arithmetic reproduces the reference's mixed float/double
// expressions: unsuffixed literals are doubles, so every expression that
// contains one is evaluated in double and rounded once to float.  C math
// functions are called on doubles explicitly: the reference is C, where
// log/exp/pow/sqrt of a float argument run in double, while the C++
// std:: overloads would pick the float versions.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "seis_internal.hpp"

namespace seis {

namespace {
constexpr double kPi = 3.141592653589793238462643383279502884197169;

// Holberg / Taylor coefficient table, src/third_party/SOFI3D/holbergcoeff.c:41-83.
// Rows: MAXRELERROR 0..4; columns: FDORDER/2 = 1..6; entries: points per
// wavelength then HC1..HC6.  The reference stores it as float.
const float kHolberg[5][6][7] = {
    {{23.0f, 1.0f, 0, 0, 0, 0, 0},
     {8.0f, 9.0f / 8.0f, -1.0f / 24.0f, 0, 0, 0, 0},
     {6.0f, 75.0f / 64.0f, -25.0f / 384.0f, 3.0f / 640.0f, 0, 0, 0},
     {5.0f, 1225.0f / 1024.0f, -245.0f / 3072.0f, 49.0f / 5120.0f, -5.0f / 7168.0f, 0, 0},
     {5.0f, 19845.0f / 16384.0f, -735.0f / 8192.0f, 567.0f / 40960.0f,
      -405.0f / 229376.0f, 35.0f / 294912.0f, 0},
     {4.0f, 160083.0f / 131072.0f, -12705.0f / 131072.0f, 22869.0f / 1310720.0f,
      -5445.0f / 1835008.0f, 847.0f / 2359296.0f, -63.0f / 2883584.0f}},
    {{49.7f, 1.0010f, 0, 0, 0, 0, 0},
     {8.32f, 1.1382f, -0.046414f, 0, 0, 0, 0},
     {4.77f, 1.1965f, -0.078804f, 0.0081781f, 0, 0, 0},
     {3.69f, 1.2257f, -0.099537f, 0.018063f, -0.0026274f, 0, 0},
     {3.19f, 1.2415f, -0.11231f, 0.026191f, -0.0064682f, 0.001191f, 0},
     {2.91f, 1.2508f, -0.12034f, 0.032131f, -0.010142f, 0.0029857f, -0.00066667f}},
    {{22.2f, 1.0050f, 0, 0, 0, 0, 0},
     {5.65f, 1.1534f, -0.052806f, 0, 0, 0, 0},
     {3.74f, 1.2111f, -0.088313f, 0.011768f, 0, 0, 0},
     {3.11f, 1.2367f, -0.10815f, 0.023113f, -0.0046905f, 0, 0},
     {2.80f, 1.2496f, -0.11921f, 0.031130f, -0.0093272f, 0.0025161f, 0},
     {2.62f, 1.2568f, -0.12573f, 0.036423f, -0.013132f, 0.0047484f, -0.0015979f}},
    {{15.8f, 1.0100f, 0, 0, 0, 0, 0},
     {4.80f, 1.1640f, -0.057991f, 0, 0, 0, 0},
     {3.39f, 1.2192f, -0.094070f, 0.014608f, 0, 0, 0},
     {2.90f, 1.2422f, -0.11269f, 0.026140f, -0.0064054f, 0, 0},
     {2.65f, 1.2534f, -0.12257f, 0.033755f, -0.011081f, 0.0036784f, 0},
     {2.51f, 1.2596f, -0.12825f, 0.038550f, -0.014763f, 0.0058619f, -0.0024538f}},
    {{9.16f, 1.0300f, 0, 0, 0, 0, 0},
     {3.47f, 1.1876f, -0.072518f, 0, 0, 0, 0},
     {2.91f, 1.2341f, -0.10569f, 0.022589f, 0, 0, 0},
     {2.61f, 1.2516f, -0.12085f, 0.032236f, -0.011459f, 0, 0},
     {2.45f, 1.2596f, -0.12829f, 0.038533f, -0.014681f, 0.0072580f, 0},
     {2.36f, 1.2640f, -0.13239f, 0.042217f, -0.017803f, 0.0081959f, -0.0051848f}},
};

// The reference passes HC1..6, DT, DT2 and DH to its kernels as text,
// "%9.9f" followed by an f suffix (src/clprogram.c:323-348).  Reading that
// text back as a float gives the exact constant the kernels use.
float as_kernel_literal(double v) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%9.9f", v);
    return std::strtof(buf, nullptr);
}

// One axis of CPML profiles; src/third_party/SOFI3D/CPML_COEFF.c:26-200.
// Output arrays have 2*nab entries and start zeroed.
int cpml_profile(float NPOWER, float kmax, float FPML, float VPPML, float dh,
                 float dt, int nab, float *K, float *a, float *b, float *Kh,
                 float *ah, float *bh) {
    int bad = 0;
    const float npower = NPOWER;
    const float alpha_max = 2.0 * kPi * (FPML / 2.0);
    const float R = 0.0008;
    const float ca = 0.25, cb = 0.75, cc = 0.0;
    std::vector<float> d(2 * nab + 1, 0.f), dh_(2 * nab + 1, 0.f);
    const float d0 = -(npower + 1) * VPPML * std::log((double)R) / (2.0 * nab * dh);
    for (int i = 0; i <= nab; ++i) {
        float pos = (nab - i) * dh;
        float xn = pos / (nab * dh);
        d[i] = d0 * (ca * xn + cb * std::pow((double)xn, (double)npower) +
                     cc * std::pow((double)xn, 4.0));
        K[i] = 1.0 + (kmax - 1.0) * std::pow((double)xn, (double)npower);
        float alpha = alpha_max * (1.0 - xn);
        bad |= alpha < 0.0;
        b[i] = std::exp((double)(-(d[i] / K[i] + alpha) * dt));
        a[i] = std::fabs(d[i]) > 1.0e-6
                   ? (float)(d[i] * (b[i] - 1.0) / (K[i] * (d[i] + K[i] * alpha)))
                   : 0.0f;
        if (i < nab) {
            float posh = (nab - i + 0.5 - 1.0) * dh;
            float xh = posh / (nab * dh);
            dh_[i] = d0 * (ca * xh + cb * std::pow((double)xh, (double)npower) +
                           cc * std::pow((double)xh, 4.0));
            bad |= posh < 0.0;
            Kh[i] = 1.0 + (kmax - 1.0) * std::pow((double)xh, (double)npower);
            float alphah = alpha_max * (1.0 - xh);
            bad |= alphah < 0.0;
            bh[i] = std::exp((double)(-(dh_[i] / Kh[i] + alphah) * dt));
            if (std::fabs(dh_[i]) > 1.0e-6)
                ah[i] = dh_[i] * (bh[i] - 1.0) / (Kh[i] * (dh_[i] + Kh[i] * alphah));
            // far boundary is the mirror image; the half-grid a is mirrored
            // under the integer-grid test, as in the reference
            int l = 2 * nab - i - 1;
            if (i > 0) {
                K[l + 1] = K[i];
                b[l + 1] = b[i];
                if (std::fabs(d[i]) > 1.0e-6) a[l + 1] = a[i];
            }
            Kh[l] = Kh[i];
            bh[l] = bh[i];
            if (std::fabs(d[i]) > 1.0e-6) ah[l] = ah[i];
        }
    }
    return bad;
}

void dims(const seis_csts &c, int &NZ, int &NY, int &NX) {
    NZ = c.N[0];
    if (c.ND == 3) {
        NY = c.N[1];
        NX = c.N[2];
    } else {
        NY = 1;
        NX = c.N[1];
    }
}

// Host loops use threads only for grids of at least this many cells: the
// 2D jobs of a serving batch are set up on a thread beside the loop that
// feeds the GPU, and idle OpenMP workers spinning after a parallel region
// would take cores from it.
constexpr long kParCells = 1L << 21;

// The reference's powf(x, 2): compilers emit the correctly rounded x*x.
inline float sq(float x) { return x * x; }

inline long lin(int z, int y, int x, int NZ, int NY) {
    return (long)x * NY * NZ + (long)y * NZ + z;
}

// 2-point arithmetic average along one axis; last plane copied
// (ave_arithmetic1, src/assign_modeling_case.c:50-99).
void avg2(const std::vector<float> &in, std::vector<float> &out, int NZ, int NY,
          int NX, int dz, int dy, int dx) {
    out.assign(in.size(), 0.f);
#pragma omp parallel for schedule(static) if ((long)in.size() >= kParCells)
    for (int x = 0; x < NX - dx; ++x)
        for (int y = 0; y < NY - dy; ++y)
            for (int z = 0; z < NZ - dz; ++z)
                out[lin(z, y, x, NZ, NY)] =
                    0.5 * (in[lin(z, y, x, NZ, NY)] + in[lin(z + dz, y + dy, x + dx, NZ, NY)]);
    int x0 = 0, y0 = 0, z0 = 0;
    if (dx) x0 = NX - 1;
    else if (dy) y0 = NY - 1;
    if (dz) z0 = NZ - 1;
    for (int x = x0; x < NX; ++x)
        for (int y = y0; y < NY; ++y)
            for (int z = z0; z < NZ; ++z) out[lin(z, y, x, NZ, NY)] = in[lin(z, y, x, NZ, NY)];
}

// 4-point averages in a plane (ave_arithmetic2 :100-173, ave_harmonic
// :225-301).  The reference does not reset the start of the second edge copy,
// so the second edge keeps zeros except for its corner line; that is kept.
template <bool Harmonic>
void avg4(const std::vector<float> &in, std::vector<float> &out, int NZ, int NY,
          int NX, const int d[2][3]) {
    out.assign(in.size(), 0.f);
    const int ez = d[0][0] + d[1][0], ey = d[0][1] + d[1][1], ex = d[0][2] + d[1][2];
#pragma omp parallel for schedule(static) if ((long)in.size() >= kParCells)
    for (int x = 0; x < NX - ex; ++x)
        for (int y = 0; y < NY - ey; ++y)
            for (int z = 0; z < NZ - ez; ++z) {
                float p1 = in[lin(z, y, x, NZ, NY)];
                float p2 = in[lin(z + d[0][0], y + d[0][1], x + d[0][2], NZ, NY)];
                float p3 = in[lin(z + d[1][0], y + d[1][1], x + d[1][2], NZ, NY)];
                float p4 = in[lin(z + ez, y + ey, x + ex, NZ, NY)];
                if (Harmonic)
                    out[lin(z, y, x, NZ, NY)] =
                        4.0 / (1.0 / p1 + 1.0 / p2 + 1.0 / p3 + 1.0 / p4);
                else
                    out[lin(z, y, x, NZ, NY)] = 0.25 * (p1 + p2 + p3 + p4);
            }
    int x0 = 0, y0 = 0, z0 = 0;
    for (int e = 0; e < 2; ++e) {
        if (d[e][2]) x0 = NX - 1;
        else if (d[e][1]) y0 = NY - 1;
        if (d[e][0]) z0 = NZ - 1;
        for (int x = x0; x < NX; ++x)
            for (int y = y0; y < NY; ++y)
                for (int z = z0; z < NZ; ++z)
                    out[lin(z, y, x, NZ, NY)] = in[lin(z, y, x, NZ, NY)];
    }
}
}  // namespace

Layout make_layout(int nd, int NZ, int NY, int NX, int H, int elem) {
    Layout g;
    g.nd = nd;
    g.NZ = NZ;
    g.NY = nd == 3 ? NY : 1;
    g.NX = NX;
    g.H = H;
    g.elem = elem;
    const int line = 128 / elem;          // elements per 128-byte line
    g.ZO = line;                          // interior z=0 on a line boundary
    g.ldz = ((g.ZO + NZ + H + line - 1) / line) * line;
    g.sy = nd == 3 ? g.ldz : 0;
    g.sx = nd == 3 ? g.ldz * (NY + 2 * H) : g.ldz;
    g.nfield = g.sx * (NX + 2 * H);
    return g;
}

int build_setup(const seis_csts &c, Setup &s, std::string &err) {
    if (c.FDORDER < 2 || c.FDORDER > 12 || c.FDORDER % 2) {
        err = "FDORDER must be 2, 4, 6, 8, 10 or 12";
        return 1;
    }
    if (c.MAXRELERROR < 0 || c.MAXRELERROR > 4) {
        err = "MAXRELERROR must be in [0, 4]";
        return 1;
    }
    for (int k = 1; k <= 6; ++k)
        s.hc[k] = as_kernel_literal((double)kHolberg[c.MAXRELERROR][c.FDORDER / 2 - 1][k]);
    s.DT = as_kernel_literal((double)c.dt);
    s.DT2 = as_kernel_literal((double)c.dt / 2.0);
    s.DH = as_kernel_literal((double)c.dh);
    for (int l = 0; l < c.L; ++l) {
        s.eta[l] = (2.0 * kPi * c.FL[l]) * c.dt;             // :536-548
        s.bl[l] = 1.0 / (1.0 + (s.eta[l] * 0.5));            // src/update_s3D.cl:406
        s.cl[l] = 1.0 - (s.eta[l] * 0.5);                    // src/update_s3D.cl:407
        s.bs[l] = s.eta[l] / (1.0 + (s.eta[l] * 0.5));       // src/surface3D.cl:239
    }
    const int nab = c.nab;
    if (c.abs_type == 1) {
        s.cpml.assign(3 * 6 * 2 * nab, 0.f);
        float *x = s.cpml.data();
        if (cpml_profile(c.NPOWER, c.K_MAX_CPML, c.FPML, c.VPPML, c.dh, c.dt, nab,
                         x, x + 2 * nab, x + 4 * nab, x + 6 * nab, x + 8 * nab,
                         x + 10 * nab)) {
            err = "CPML profile: negative damping coordinate";
            return 1;
        }
        // the three axes share one profile (CPML_x/y/z, src/assign_modeling_case.c:585-660)
        std::memcpy(x + 12 * nab, x, 12 * nab * sizeof(float));
        std::memcpy(x + 24 * nab, x, 12 * nab * sizeof(float));
    } else if (c.abs_type == 2) {
        // Cerjan taper, src/assign_modeling_case.c:571-584
        s.taper.assign(nab, 0.f);
        float amp = 1 - c.abpc / 100;
        float a = std::sqrt(-std::log((double)amp) / ((nab - 1) * (nab - 1)));
        for (int i = 1; i <= nab; ++i) s.taper[i - 1] = std::exp((double)(-(a * a * (nab - i) * (nab - i))));
    }
    return 0;
}

float relax_sum(const seis_csts &c) {
    float s = 0.0;
    const float ws = 2.0 * kPi * c.f0;
    for (int l = 0; l < c.L; ++l) {
        const float eta = (2.0 * kPi * c.FL[l]) * c.dt;
        const float pts = c.dt / eta;
        s += ((ws * ws * pts * pts) / (1.0 + ws * ws * pts * pts));
    }
    return s;
}

int transform_params(const seis_csts &c, const float *p0, const float *p1,
                     const float *rho_in, const float *taup_in, const float *taus_in,
                     RefParams &o, std::string &err) {
    int NZ, NY, NX;
    dims(c, NZ, NY, NX);
    const long n = (long)NZ * NY * NX;
    const bool elastic = c.ND == 2 || c.ND == 3;
    const int L = c.L;
    if (!p0 || !rho_in || (elastic && !p1) || (L > 0 && !taup_in) ||
        (L > 0 && elastic && !taus_in)) {
        err = "a material parameter required by this case is missing";
        return 1;
    }
    float eta[kMaxL], pts[kMaxL];
    for (int l = 0; l < L; ++l) {
        eta[l] = (2.0 * kPi * c.FL[l]) * c.dt;
        pts[l] = c.dt / eta[l];
    }
    const float ws = 2.0 * kPi * c.f0;
    if (L > 0) {
        o.taup.assign(taup_in, taup_in + n);
        if (elastic) o.taus.assign(taus_in, taus_in + n);
    }
    if (c.param_type < 0 || c.param_type > 3) {
        err = "param_type must be 0, 1, 2 or 3";
        return 1;
    }
    // Every loop below computes each cell on its own, so the threads change
    // nothing in the results.
    // M (src/assign_modeling_case.c:380-449)
    o.M.resize(n);
#pragma omp parallel for schedule(static) if (n >= kParCells)
    for (long i = 0; i < n; ++i) {
        const float v = p0[i], r = rho_in[i];
        switch (c.param_type) {
            case 0: o.M[i] = sq(v) * r * c.dt / c.dh; break;
            case 1: o.M[i] = v * c.dt / c.dh; break;
            case 2: o.M[i] = sq(v / r) * r * c.dt / c.dh; break;
            case 3: {
                float vr = v - o.taup[i];
                float tp = o.taup[i] / (v - o.taup[i]);
                o.M[i] = sq(vr) * r * c.dt / c.dh;
                o.taup[i] = tp;
            } break;
        }
    }
    if (L > 0) o.Mf0 = o.M;
    if (L > 0)
#pragma omp parallel for schedule(static) if (n >= kParCells)
        for (long i = 0; i < n; ++i) {
            float s = 0.0;
            for (int l = 0; l < L; ++l)
                s += ((ws * ws * pts[l] * pts[l] * o.taup[i]) / (1.0 + ws * ws * pts[l] * pts[l]));
            o.M[i] = o.M[i] / (1.0 + s);
        }
    // mu (src/assign_modeling_case.c:305-379)
    if (elastic) {
        o.mu.resize(n);
#pragma omp parallel for schedule(static) if (n >= kParCells)
        for (long i = 0; i < n; ++i) {
            const float v = p1[i], r = rho_in[i];
            switch (c.param_type) {
                case 0: o.mu[i] = sq(v) * r * c.dt / c.dh; break;
                case 1: o.mu[i] = v * c.dt / c.dh; break;
                case 2: o.mu[i] = sq(v / r) * r * c.dt / c.dh; break;
                case 3: {
                    float vs = 0, ts = 0;
                    if (v > 0) {
                        vs = v - o.taus[i];
                        ts = o.taus[i] / (v - o.taus[i]);
                    }
                    o.mu[i] = sq(vs) * r * c.dt / c.dh;
                    o.taus[i] = ts;
                } break;
            }
        }
        if (L > 0) {
            o.muf0 = o.mu;
            const float s = relax_sum(c);
#pragma omp parallel for schedule(static) if (n >= kParCells)
            for (long i = 0; i < n; ++i) o.mu[i] = o.mu[i] / (1.0 + s * o.taus[i]);
        }
    }
    // buoyancy dt/(dh rho) (src/assign_modeling_case.c:451-462)
    o.rho.resize(n);
#pragma omp parallel for schedule(static) if (n >= kParCells)
    for (long i = 0; i < n; ++i) o.rho[i] = 1.0 / rho_in[i] * c.dt / c.dh * 1.0;
    avg2(o.rho, o.rip, NZ, NY, NX, 0, 0, 1);
    if (c.ND == 3) avg2(o.rho, o.rjp, NZ, NY, NX, 0, 1, 0);
    avg2(o.rho, o.rkp, NZ, NY, NX, 1, 0, 0);
    if (elastic) {
        const int dxz[2][3] = {{0, 0, 1}, {1, 0, 0}};
        const int dyz[2][3] = {{0, 1, 0}, {1, 0, 0}};
        const int dxy[2][3] = {{0, 0, 1}, {0, 1, 0}};
        avg4<true>(o.mu, o.muipkp, NZ, NY, NX, dxz);
        if (c.ND == 3) {
            avg4<true>(o.mu, o.muipjp, NZ, NY, NX, dxy);
            avg4<true>(o.mu, o.mujpkp, NZ, NY, NX, dyz);
        }
        if (L > 0) {
            avg4<false>(o.taus, o.tausipkp, NZ, NY, NX, dxz);
            if (c.ND == 3) {
                avg4<false>(o.taus, o.tausipjp, NZ, NY, NX, dxy);
                avg4<false>(o.taus, o.tausjpkp, NZ, NY, NX, dyz);
            }
        }
    }
    return 0;
}

int check_stability(const seis_csts &c, const RefParams &p, float *dtstable) {
    // maxima per thread with the serial loop's comparison (NaNs never win),
    // then over threads: the same value as one pass
    auto vmax_of = [&](const std::vector<float> &m) {
        float best = 0;
#pragma omp parallel if ((long)m.size() >= kParCells)
        {
            float b = 0;
#pragma omp for schedule(static) nowait
            for (long i = 0; i < (long)m.size(); ++i) {
                float v = std::sqrt((double)(m[i] * p.rho[i])) / c.dt * c.dh;
                if (b < v) b = v;
            }
#pragma omp critical
            if (best < b) best = b;
        }
        return best;
    };
    const float vsmax = vmax_of(p.mu), vpmax = vmax_of(p.M);
    const float vmax = vpmax == 0 ? vsmax : vpmax;
    float gamma;
    switch (c.FDORDER) {
        case 4: gamma = 1.184614; break;
        case 6: gamma = 1.283482; break;
        case 8: gamma = 1.345927; break;
        case 10: gamma = 1.38766; break;
        case 12: gamma = 1.417065; break;
        default: gamma = 1.0; break;
    }
    float dts = c.ND == 3 ? c.dh / (gamma * std::sqrt(3.0) * vmax)
                          : c.dh / (gamma * std::sqrt(2.0) * vmax);
    if (dtstable) *dtstable = dts;
    return c.dt > dts;
}

template <typename T>
void to_padded(const Layout &g, const float *ref, float mul, std::vector<T> &out) {
    out.resize((size_t)g.nfield);
    T *o = out.data();
    const long nf = g.nfield;
    const bool par = nf >= kParCells;
#pragma omp parallel for schedule(static) if (par)
    for (long i = 0; i < nf; ++i) o[i] = 0;
#pragma omp parallel for schedule(static) if (par)
    for (int x = 0; x < g.NX; ++x)
        for (int y = 0; y < g.NY; ++y) {
            const float *src = ref + ((size_t)x * g.NY + y) * g.NZ;
            T *dst = o + g.idx(0, y, x);
            if (mul == 1)
                for (int z = 0; z < g.NZ; ++z) dst[z] = (T)src[z];
            else
                for (int z = 0; z < g.NZ; ++z) dst[z] = (T)src[z] * (T)mul;
        }
}
template void to_padded<float>(const Layout &, const float *, float, std::vector<float> &);
template void to_padded<double>(const Layout &, const float *, float, std::vector<double> &);

void visco_coeffs(const seis_csts &c, const RefParams &p, float DT, std::vector<float> &g,
                  std::vector<float> &f, std::vector<float> &d, std::vector<float> &e) {
    const long n = (long)p.M.size();
    const int L = c.L;
    const bool elastic = c.ND != 22;
    g.resize(n);
    e.resize(n);
    if (elastic) {
        f.resize(n);
        d.resize(n);
    }
#pragma omp parallel for schedule(static) if (n >= kParCells)
    for (long i = 0; i < n; ++i) {
        g[i] = p.M[i] * (1.0 + (float)L * p.taup[i]);
        if (elastic) {
            f[i] = 2.0 * p.mu[i] * (1.0 + (float)L * p.taus[i]);
            d[i] = 2.0 * p.mu[i] * p.taus[i] / DT;
            e[i] = p.M[i] * p.taup[i] / DT;
        } else {
            e[i] = p.M[i] * p.taup[i];
        }
    }
}

void visco_pair(const std::vector<float> &mu, const std::vector<float> &ts, int L, float DT,
                std::vector<float> &ff, std::vector<float> &dd) {
    const long n = (long)mu.size();
    ff.resize(n);
    dd.resize(n);
#pragma omp parallel for schedule(static) if (n >= kParCells)
    for (long i = 0; i < n; ++i) {
        ff[i] = mu[i] * (1.0 + (float)L * ts[i]);
        dd[i] = mu[i] * ts[i] / DT;
    }
}

std::vector<std::string> out_names(const seis_csts &c) {
    const int so = c.seisout;
    if (c.ND == 3) {
        switch (so) {
            case 1: return {"vx", "vy", "vz"};
            case 2: return {"p"};
            case 3: return {"sxx", "syy", "szz", "sxz", "sxy", "syz"};
            case 4: return {"vx", "vy", "vz", "sxx", "syy", "szz", "sxz", "sxy", "syz", "p"};
        }
    } else if (c.ND == 22) {
        switch (so) {
            case 1: return {"vx", "vz"};
            case 2: return {"p"};
            case 4: return {"vx", "vz", "p"};
        }
    } else {
        switch (so) {
            case 1: return {"vx", "vz"};
            case 2: return {"p"};
            case 3: return {"sxx", "szz", "sxz"};
            case 4: return {"vx", "vz", "sxx", "szz", "sxz", "p"};
        }
    }
    return {};
}


// ---------------------------------------------------------------------------
// Gradients through the staggered averages (gradfix FIX_AVG): transposes of
// avg2 and avg4, dJ/dp_c = sum over the staggered nodes s that cell c enters
// of dJ/dp_s dp_s/dp_c, with the node classes of those functions: averaged
// nodes, nodes copied from their own cell, and (avg4's second edge) nodes
// left at zero, which depend on no cell.
// ---------------------------------------------------------------------------
namespace {
// 1 where avg4 copies the node from its cell (its edge loops)
std::vector<char> avg4_copied(int NZ, int NY, int NX, const int d[2][3]) {
    std::vector<char> cp((size_t)NZ * NY * NX, 0);
    int x0 = 0, y0 = 0, z0 = 0;
    for (int e = 0; e < 2; ++e) {
        if (d[e][2]) x0 = NX - 1;
        else if (d[e][1]) y0 = NY - 1;
        if (d[e][0]) z0 = NZ - 1;
        for (int x = x0; x < NX; ++x)
            for (int y = y0; y < NY; ++y)
                for (int z = z0; z < NZ; ++z) cp[lin(z, y, x, NZ, NY)] = 1;
    }
    return cp;
}
}  // namespace

void chain_density(const std::vector<double> &G, const std::vector<float> &b,
                   const std::vector<float> &bs, int NZ, int NY, int NX, int dz, int dy, int dx,
                   double *out) {
    // bs = avg2(b) along (dz, dy, dx); rho_s = k / bs, so
    // drho_s/drho_c = (1/2) (b_c / bs_s)^2 for an averaged node, 1 for a copy
    for (int x = 0; x < NX; ++x)
        for (int y = 0; y < NY; ++y)
            for (int z = 0; z < NZ; ++z) {
                const long s = lin(z, y, x, NZ, NY);
                const bool avg = x < NX - dx && y < NY - dy && z < NZ - dz;
                if (!avg) {
                    out[s] += G[s];
                    continue;
                }
                const long s2 = lin(z + dz, y + dy, x + dx, NZ, NY);
                const double q = (double)bs[s];
                if (q == 0) continue;
                const double r1 = (double)b[s] / q, r2 = (double)b[s2] / q;
                out[s] += 0.5 * r1 * r1 * G[s];
                out[s2] += 0.5 * r2 * r2 * G[s];
            }
}

void chain_shear(const std::vector<double> &P, const std::vector<float> &mus, int NZ, int NY,
                 int NX, const int d[2][3], double *out) {
    // P_s: the shear stress times its adjoint increment, summed over time;
    // dJ/dmu_s = -P_s / mu_s^2 and the harmonic mean gives dmu_s/dmu_c =
    // mu_s^2 / (4 mu_c^2): -P_s / (4 mu_c^2) per cell of an averaged node,
    // -P_s / mu_c^2 for a copied one (the gradient formulas' own moduli mus)
    const std::vector<char> cp = avg4_copied(NZ, NY, NX, d);
    const int ez = d[0][0] + d[1][0], ey = d[0][1] + d[1][1], ex = d[0][2] + d[1][2];
    auto add = [&](long c, double v) {
        const double m = mus[c];
        if (m != 0) out[c] += -v / (m * m);
    };
    for (int x = 0; x < NX; ++x)
        for (int y = 0; y < NY; ++y)
            for (int z = 0; z < NZ; ++z) {
                const long s = lin(z, y, x, NZ, NY);
                if (cp[s]) {
                    add(s, P[s]);
                    continue;
                }
                if (!(x < NX - ex && y < NY - ey && z < NZ - ez)) continue;   // left at zero
                const double q = 0.25 * P[s];
                add(s, q);
                add(lin(z + d[0][0], y + d[0][1], x + d[0][2], NZ, NY), q);
                add(lin(z + d[1][0], y + d[1][1], x + d[1][2], NZ, NY), q);
                add(lin(z + ez, y + ey, x + ex, NZ, NY), q);
            }
}

void chain_staggered(const seis_csts &c, const RefParams &p, const std::vector<float> &mus,
                     const std::vector<double> *sh, const std::vector<double> *rh, double *gmu,
                     double *grho) {
    const int NZ = c.N[0], NY = c.ND == 3 ? c.N[1] : 1, NX = c.ND == 3 ? c.N[2] : c.N[1];
    const size_t n = (size_t)NZ * NY * NX;
    // density: rh[0] vx (x nodes), rh[1] vy (y), rh[2] vz (z); grho is
    // rebuilt from them
    std::fill(grho, grho + n, 0.0);
    chain_density(rh[0], p.rho, p.rip, NZ, NY, NX, 0, 0, 1, grho);
    if (c.ND == 3) chain_density(rh[1], p.rho, p.rjp, NZ, NY, NX, 0, 1, 0, grho);
    chain_density(rh[2], p.rho, p.rkp, NZ, NY, NX, 1, 0, 0, grho);
    // shear: sh[0] xz, sh[1] xy, sh[2] yz, added to the normal-stress part
    const int dxz[2][3] = {{0, 0, 1}, {1, 0, 0}};
    const int dxy[2][3] = {{0, 0, 1}, {0, 1, 0}};
    const int dyz[2][3] = {{0, 1, 0}, {1, 0, 0}};
    chain_shear(sh[0], mus, NZ, NY, NX, dxz, gmu);
    if (c.ND == 3) {
        chain_shear(sh[1], mus, NZ, NY, NX, dxy, gmu);
        chain_shear(sh[2], mus, NZ, NY, NX, dyz, gmu);
    }
}

}  // namespace seis
