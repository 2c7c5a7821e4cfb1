"""Closed-form Green's functions of a homogeneous (visco)elastic full space
for a vertical point force, used to check the oracle.

2D: Carcione (2014), Wave Fields in Real Media, eq. for a line force
(Hankel functions of the second kind, e^{i w t} convention).
3D: Gosselin-Cliche & Giroux (2014), Geophysics 79(3) T169-T188.
Attenuation: generalized standard linear solid (Blanch et al. 1995), with the
modulus normalised so that the phase velocity at w0 equals the given
velocity (the convention of SeisCL/analytical/viscoelastic.py:14-56).
"""
import numpy as np
from scipy.special import hankel2


def gsls_moduli(vp, vs, rho, taup, taus, FL, f0, omega):
    """Complex P and S moduli of the GSLS at angular frequencies omega."""
    omega = np.asarray(omega, np.complex128)
    wl = 2 * np.pi * np.atleast_1d(np.asarray(FL, np.float64)) if len(np.atleast_1d(FL)) else []
    w0 = 2 * np.pi * f0
    si = np.zeros_like(omega)
    sr = 0.0
    for w in wl:
        si = si + 1j * omega / w / (1.0 + 1j * omega / w)
        sr = sr + w0 ** 2 / w ** 2 / (1.0 + w0 ** 2 / w ** 2)
    M = vp ** 2 * rho * (1.0 + si * taup) / (1.0 + sr * taup)
    mu = vs ** 2 * rho * (1.0 + si * taus) / (1.0 + sr * taus)
    return M, mu


def _to_time(spec, nt):
    full = np.zeros((nt,) + spec.shape[1:], np.complex128)
    full[:spec.shape[0]] = spec
    for ii in range(1, nt // 2):
        full[-ii] = np.conj(full[ii])
    v = nt * np.fft.ifft(full, axis=0)
    return np.real(v)


def force_z_2d(vp, vs, rho, taup, taus, f0, FL, dt, offsets, src):
    """(vx, vz) at receivers offsets[:, (x, z)] from a z line force in 2D."""
    nt = src.shape[0]
    F = np.fft.fft(src)
    omega = 2 * np.pi * np.fft.fftfreq(nt, dt)
    kk = np.arange(1, nt // 2)
    w = omega[kk]
    M, mu = gsls_moduli(vp, vs, rho, taup, taus, FL, f0, w)
    a, b = np.sqrt(M / rho), np.sqrt(mu / rho)         # complex P, S velocities
    x = offsets[:, 0][None, :]
    z = offsets[:, 1][None, :]
    r = np.sqrt(x ** 2 + z ** 2)
    w_, a_, b_ = w[:, None], a[:, None], b[:, None]
    kp, ks = w_ / a_, w_ / b_
    h0p, h0s = hankel2(0, kp * r), hankel2(0, ks * r)
    h1p, h1s = hankel2(1, kp * r), hankel2(1, ks * r)
    G1 = -1j * np.pi / 2 * (h0p / a_ ** 2 + h1s / (w_ * r * b_) - h1p / (w_ * r * a_))
    G2 = 1j * np.pi / 2 * (h0s / b_ ** 2 - h1s / (w_ * r * b_) + h1p / (w_ * r * a_))
    pre = F[kk][:, None] / (2 * np.pi * rho) * 1j * w_
    spec_x = np.zeros((nt // 2,) + x.shape[1:], np.complex128)
    spec_z = np.zeros_like(spec_x)
    spec_x[1:] = pre * (x * z / r ** 2) * (G1 + G2)
    spec_z[1:] = pre / r ** 2 * (z ** 2 * G1 - x ** 2 * G2)
    return _to_time(spec_x, nt), _to_time(spec_z, nt)


def force_z_3d(vp, vs, rho, taup, taus, f0, FL, dt, offsets, src):
    """(vx, vy, vz) at receivers offsets[:, (x, y, z)] from a z point force."""
    nt = src.shape[0]
    F = np.fft.fft(src)
    omega = 2 * np.pi * np.fft.fftfreq(nt, dt)
    kk = np.arange(1, nt // 2)
    w = omega[kk][:, None]
    M, mu = gsls_moduli(vp, vs, rho, taup, taus, FL, f0, omega[kk])
    kp = w / np.sqrt(M / rho)[:, None]
    ks = w / np.sqrt(mu / rho)[:, None]
    x, y, z = (offsets[:, i][None, :] for i in range(3))
    R = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    amp = F[kk][:, None] / (4.0 * np.pi * rho * R ** 5 * w ** 2) * 1j * w
    ep, es = np.exp(-1j * kp * R), np.exp(-1j * ks * R)
    lat = (R ** 2 * kp ** 2 - 3.0 - 3.0j * R * kp) * ep + (3.0 + 3.0j * R * ks - R ** 2 * ks ** 2) * es
    q = x ** 2 + y ** 2 - 2.0 * z ** 2
    vzs = (q * (ep - es) + (z ** 2 * R ** 2 * kp ** 2 + 1j * q * R * kp) * ep
           + ((x ** 2 + y ** 2) * R ** 2 * ks ** 2 - 1j * q * R * ks) * es)
    out = []
    for s in (amp * x * z * lat, amp * y * z * lat, amp * vzs):
        spec = np.zeros((nt // 2, s.shape[1]), np.complex128)
        spec[1:] = s
        out.append(_to_time(spec, nt))
    return out


# ------------------------------------------------------------------ sources
# The finite-difference runs add dt*s(t_n) to the stresses (explosion: s/ND
# to each normal stress; 2D acoustic: s to p) at step n before the velocity
# update, and record v^{n+1/2} and sigma^{n+1} at step n.  A cell holds the
# source over its volume dh^ND, so the continuous source is a moment-rate
# density dh^ND s(t)/ND * delta(x) (explosion) or a pressure-rate density
# dh^2 s(t) delta(x) (acoustic).  In continuous time the sample s_n acts at
# t_n - dt/2; receivers therefore see the continuous response at t_n + dt
# (velocities) and t_n + 1.5 dt (stresses and pressure).  V_SHIFT, P_SHIFT
# are those offsets in steps.
V_SHIFT, P_SHIFT = 1.0, 1.5


def ricker_t(f0, t):
    """Continuous Ricker of SeisCL.ricker_wavelet (t0 = -1.5/f0 at sample 0)."""
    tt = t - 1.5 / f0
    a = (np.pi * f0) ** 2
    return (1.0 - 2.0 * a * tt ** 2) * np.exp(-a * tt ** 2)


def ricker_dot(f0, t):
    """Time derivative of ricker_t."""
    tt = t - 1.5 / f0
    a = (np.pi * f0) ** 2
    return (-6.0 * a * tt + 4.0 * a * a * tt ** 3) * np.exp(-a * tt ** 2)


def _response(kernel, f0, dt, nt, shift, pad=8):
    """Samples n = 0..nt-1 of the response whose spectrum is kernel(omega)
    times that of the Ricker, evaluated at t_n + shift*dt (e^{i w t})."""
    n = pad * nt
    s = ricker_t(f0, (np.arange(n) + shift) * dt)
    w = 2 * np.pi * np.fft.rfftfreq(n, dt)
    K = np.zeros((w.size,) + np.shape(kernel(w[1:2]))[1:], np.complex128)
    K[1:] = kernel(w[1:])
    S = np.fft.rfft(s)
    return np.fft.irfft(S.reshape((-1,) + (1,) * (K.ndim - 1)) * K, n, axis=0)[:nt]


def explosion_3d(vp, vs, rho, dh, f0, dt, nt, r_p, r_v):
    """Homogeneous full space, one-cell explosion of the FD scheme (P waves
    only).  From u = grad(phi), phi = dh^3 S1(t - r/a)/(12 pi rho a^2 r),
    S1 the time integral of s:
      p   = (sxx+syy+szz)/3 = (a^2 - 4 b^2/3) dh^3 s'(t - r/a) / (12 pi a^4 r),
      v_r = -dh^3/(12 pi rho a^2) [s'(t - r/a)/(a r) + s(t - r/a)/r^2].
    r_p: distances of the pressure receivers, r_v: of the radial velocity.
    Returns (p (nt, len r_p), v_r (nt, len r_v))."""
    a, b = vp, vs

    def kp(w):
        w = w[:, None]
        return (a * a - 4 * b * b / 3) * dh ** 3 * 1j * w * np.exp(-1j * w * r_p / a) / (
            12 * np.pi * a ** 4 * r_p)

    def kv(w):
        w = w[:, None]
        return -dh ** 3 / (12 * np.pi * rho * a * a) * (1j * w / (a * r_v) + 1 / r_v ** 2) * \
            np.exp(-1j * w * r_v / a)

    return (_response(kp, f0, dt, nt, P_SHIFT), _response(kv, f0, dt, nt, V_SHIFT))


def explosion_2d(vp, vs, rho, dh, f0, dt, nt, r_p, r_v, acoustic=False):
    """Homogeneous 2D (plane strain) full space, one-cell explosion (P waves
    only), e^{i w t} convention, k = w/a:
      elastic:  p = (sxx+szz)/2, phi = -(dh^2/2) S1/(rho a^2) (i/4) H0(kr),
                p = (lam+mu) w dh^2 s H0(kr)/(8 rho a^4),
                v_r = i (dh^2/2) s k H1(kr)/(4 rho a^2);
      acoustic: dp/dt = M div v + s delta, p = w dh^2 s H0(kr)/(4 a^2),
                v_r = dp/dr/(i w rho) = i dh^2 s k H1(kr)/(4 rho a^2)
    (H = Hankel functions of the second kind)."""
    a = vp
    lam_mu = rho * (vp * vp - vs * vs)

    def kp(w):
        w = w[:, None]
        if acoustic:
            amp = w * dh ** 2 / (4 * a * a)
        else:
            amp = lam_mu * w * dh ** 2 / (8 * rho * a ** 4)
        return amp * hankel2(0, w / a * r_p)

    def kv(w):
        w = w[:, None]
        amp = (1.0 if acoustic else 0.5) * 1j * dh ** 2 * (w / a) / (4 * rho * a * a)
        return amp * hankel2(1, w / a * r_v)

    return (_response(kp, f0, dt, nt, P_SHIFT), _response(kv, f0, dt, nt, V_SHIFT))


# ------------------------------------------------------------------ Garvin
def _t_ps(x, z, h, vp, vs):
    """Arrival of the P-to-S surface reflection (Fermat: least time over the
    reflection point)."""
    from scipy.optimize import minimize_scalar
    if z == 0:
        return float(np.hypot(x, h) / vp)
    T = lambda xs: np.hypot(xs, h) / vp + np.hypot(x - xs, z) / vs   # noqa: E731
    return float(minimize_scalar(T, bounds=(0.0, x), method="bounded",
                                 options=dict(xatol=1e-12 * x)).fun)


def ricker_ddot(f0, t):
    """Second time derivative of ricker_t."""
    tt = t - 1.5 / f0
    a = (np.pi * f0) ** 2
    return (-6.0 * a + 24.0 * a * a * tt ** 2 - 8.0 * a ** 3 * tt ** 4) * np.exp(-a * tt ** 2)


def garvin_parts(x, z, h, vp, vs):
    """Displacement (u along +x, w downwards) at range x > 0 and depth z of
    a line blast with Heaviside time function at depth h in a homogeneous
    half space with a free surface at z = 0, per unit shear modulus (up to a
    common factor).  Garvin (1956) in the algebraic form of Alterman &
    Loewenthal (1969) and Sanchez-Sesma, Iturraran-Viveros & Kausel (2013):
    the direct P wave from tP, its PP image from tPP, and the PS reflection
    from tPS, whose complex slowness q3 is the root (Re >= 0, Im >= 0, least
    Im) of A q^4 - 4i B q^3 - 2 C q^2 + 4i D q + E = 0.

    Returns the arrival times (tP, tPP, tPS), the integral of the direct
    wave from tP (closed form) and the PP and PS terms as functions of t
    (each valid after its own arrival; the direct and PP terms have
    integrable 1/sqrt singularities at their arrivals)."""
    a2 = (vs / vp) ** 2
    a = np.sqrt(a2)
    r1, r2 = np.hypot(x, z - h), np.hypot(x, z + h)
    s1, c1, s2, c2 = x / r1, (z - h) / r1, x / r2, (z + h) / r2
    tP, tPP, tPS = r1 / vp, r2 / vp, _t_ps(x, z, h, vp, vs)

    def direct_int(t):
        g = 0.5 / r1 * np.sqrt(np.maximum(t * t - tP * tP, 0.0)) / np.pi
        return g * s1, g * c1

    def pp(t):
        T2 = np.sqrt(t * t - tPP * tPP)
        q = (c2 * T2 + 1j * s2 * t) * vs / r2
        dq = c2 * t / T2 + 1j * s2
        Q = q * q
        R = (1 + 2 * Q) ** 2
        D = R - 4 * Q * np.sqrt(Q + 1) * np.sqrt(Q + a2)
        f2 = 0.5 / r2 * t / T2
        u = -f2 * s2 - 4 / r2 * np.imag(q ** 3 * np.sqrt(Q + 1) * dq / D)
        w = f2 * c2 - 1 / r2 * np.real(R * dq / D)
        return u / np.pi, w / np.pi

    def ps(t):
        tau = t * vs / r2
        X, Z, Hh = x / r2, z / r2, h / r2
        A = ((Hh + Z) ** 2 + X ** 2) * ((Hh - Z) ** 2 + X ** 2)
        B = tau * X * (X ** 2 + Hh ** 2 + Z ** 2)
        C = tau ** 2 * (3 * X ** 2 + Hh ** 2 + Z ** 2) - (
            X ** 2 * (a2 * Hh ** 2 + Z ** 2) + (Hh ** 2 - Z ** 2) * (a2 * Hh ** 2 - Z ** 2))
        Dd = tau * (tau ** 2 * X - X * (a2 * Hh ** 2 + Z ** 2))
        E = (tau ** 2 - (a * Hh + Z) ** 2) * (tau ** 2 - (a * Hh - Z) ** 2)
        comp = np.zeros((tau.size, 4, 4), np.complex128)   # companion of the monic quartic
        comp[:, 0, 0], comp[:, 0, 1] = 4j * B / A, 2 * C / A
        comp[:, 0, 2], comp[:, 0, 3] = -4j * Dd / A, -E / A
        comp[:, 1, 0] = comp[:, 2, 1] = comp[:, 3, 2] = 1
        roots = np.linalg.eigvals(comp)
        key = np.where((roots.real >= 0) & (roots.imag >= 0), roots.imag, np.inf)
        q = roots[np.arange(tau.size), np.argmin(key, axis=1)]
        Q = q * q
        QS, QP = np.sqrt(Q + 1), np.sqrt(Q + a2)
        S = 1 + 2 * Q
        D = S * S - 4 * Q * QS * QP
        dq = 1 / ((h / r2 / QP + z / r2 / QS) * q - 1j * x / r2)
        return (2 / r2 * np.imag(q * S * QS * dq / D) / np.pi,
                2 / r2 * np.real(Q * S * dq / D) / np.pi)

    return (tP, tPP, tPS), direct_int, pp, ps


def _integral_from(f, t0, t, n=6000):
    """Integral of f from t0 to each t (0 before t0), with t = t0 + xi^2 so
    that a 1/sqrt(t - t0) singularity becomes smooth (midpoint rule in xi);
    f returns a pair of arrays."""
    out = [np.zeros_like(t), np.zeros_like(t)]
    xmax = np.sqrt(max(t[-1] - t0, 0.0))
    if xmax == 0:
        return out
    dxi = xmax / n
    xi = (np.arange(n) + 0.5) * dxi
    vals = f(t0 + xi * xi)
    m = t > t0
    for k in range(2):
        cum = np.concatenate([[0.0], np.cumsum(vals[k] * 2 * xi * dxi)])
        out[k][m] = np.interp(np.sqrt(t[m] - t0), np.arange(n + 1) * dxi, cum)
    return out


def garvin_shot(offsets, zsrc, zrec, vp, vs, f0, dt, nt, comp="x", resamp=20):
    """vx (comp 'x') or vz seismograms (nt, len offsets) of the FD explosion
    (a positive s adds tension: an implosion) at depth zsrc recorded at depth
    zrec, up to a common positive factor: -(u_step * s') = -(U * s'') with U
    the time integral of the step response, sampled at t_n + V_SHIFT dt on a
    grid resamp times finer."""
    dtf = dt / resamp
    nf = nt * resamp
    tf = np.arange(nf + 1) * dtf
    sdd = ricker_ddot(f0, np.arange(nf) * dtf + V_SHIFT * dt)
    k = 0 if comp == "x" else 1
    out = np.zeros((nt, len(offsets)))
    for i, x in enumerate(offsets):
        (tP, tPP, tPS), direct_int, pp, ps = garvin_parts(float(x), zrec, zsrc, vp, vs)
        U = direct_int(tf)[k] + _integral_from(pp, tPP, tf)[k] + _integral_from(ps, tPS, tf)[k]
        v = -np.fft.irfft(np.fft.rfft(U[:nf], 2 * nf) * np.fft.rfft(sdd, 2 * nf), 2 * nf)[:nf] * dtf
        out[:, i] = v[::resamp]
    return out


def gsls_inverse_q(tau, FL, freqs):
    """1/Q(f) = Im M / Re M of the GSLS modulus M_R (1 + tau sum_l
    i w t_l / (1 + i w t_l)), t_l = 1/(2 pi FL_l)."""
    w = 2 * np.pi * np.asarray(freqs, np.float64)
    tl = 1.0 / (2 * np.pi * np.atleast_1d(np.asarray(FL, np.float64)))
    Mw = 1.0 + tau * sum(1j * w * t / (1.0 + 1j * w * t) for t in tl)
    return Mw.imag / Mw.real


# ------------------------------------------------------------------ Lamb 3D
def lamb_roots(pois):
    """Roots x1 <= x2 <= x3 of the Rayleigh cubic 16(1-a2)x^3 - 8(3-2a2)x^2
    + 8x - 1 = 0, a2 = (Cs/Cp)^2; x3 = (Cs/Cr)^2 (Kausel 2012, Proc. R. Soc.
    A 468, 'Lamb's problem at its simplest', eq. for the Rayleigh function)."""
    a2 = (1 - 2 * pois) / (2 - 2 * pois)
    x = np.roots([16 * (1 - a2), -8 * (3 - 2 * a2), 8.0, -1.0])
    x = np.sort_complex(x)
    if np.any(np.abs(x.imag) > 1e-12):
        raise ValueError("complex Rayleigh roots (Poisson's ratio above 0.2631) not handled")
    return a2, np.sort(x.real)


def lamb_step_integral(tau, pois, comp="r"):
    """Time integral, in dimensionless time tau = t Cs / r, of the surface
    displacement under a horizontal step point load on the surface of a
    half space (Kausel 2012): comp 'r' the radial displacement along the
    load (varies as cos(theta)), 't' the tangential one (as -sin(theta)),
    both times 2 pi mu r.  The step responses are, with S_j = |tau^2 - x_j|^(1/2)
    and C_j = (1 - x_j) |a2 - x_j|^(1/2) / prod_{k != j} (x_j - x_k):
      radial:     (1-nu)/2 tau^2 (C1/S1 + C2/S2 + C3/S3)   tp < tau < 1,
                  1 + (1-nu) C3 tau^2 / S3                 1 < tau < tr,
                  1                                         tau > tr;
      tangential: (1-nu)/2 (1 - C1 S1 - C2 S2 + C3 S3)     tp < tau < 1,
                  (1-nu) (1 + C3 S3)                        1 < tau < tr,
                  1 - nu                                    tau > tr,
    integrated in closed form (the radial one has an integrable singularity
    at the Rayleigh arrival tr = x3^(1/2))."""
    a2, (x1, x2, x3) = lamb_roots(pois)
    xs = (x1, x2, x3)
    C = [(1 - xj) * np.sqrt(abs(a2 - xj)) / np.prod([xj - xk for xk in xs if xk != xj])
         for xj in xs]
    tp, tr = np.sqrt(a2), np.sqrt(x3)
    nu1 = 1 - pois
    tau = np.asarray(tau, np.float64)

    def Ij(t, xj):       # int t^2 / sqrt(t^2 - xj) dt
        s = np.sqrt(np.maximum(t * t - xj, 0.0))
        return 0.5 * t * s + 0.5 * xj * np.log(t + s)

    def Jj(t, xj):       # int sqrt(t^2 - xj) dt
        s = np.sqrt(np.maximum(t * t - xj, 0.0))
        return 0.5 * t * s - 0.5 * xj * np.log(t + s)

    def K3(t):           # int t^2 / sqrt(x3 - t^2) dt
        t = np.minimum(t, tr)
        return 0.5 * x3 * np.arcsin(t / tr) - 0.5 * t * np.sqrt(np.maximum(x3 - t * t, 0.0))

    def L3(t):           # int sqrt(x3 - t^2) dt
        t = np.minimum(t, tr)
        return 0.5 * t * np.sqrt(np.maximum(x3 - t * t, 0.0)) + 0.5 * x3 * np.arcsin(t / tr)

    def partA(t):        # integral from tp to t, t in [tp, 1]
        if comp == "r":
            return 0.5 * nu1 * (C[0] * (Ij(t, x1) - Ij(tp, x1)) + C[1] * (Ij(t, x2) - Ij(tp, x2))
                                + C[2] * (K3(t) - K3(tp)))
        return 0.5 * nu1 * ((t - tp) - C[0] * (Jj(t, x1) - Jj(tp, x1))
                            - C[1] * (Jj(t, x2) - Jj(tp, x2)) + C[2] * (L3(t) - L3(tp)))

    def partB(t):        # integral from 1 to t, t in [1, tr]
        if comp == "r":
            return (t - 1.0) + nu1 * C[2] * (K3(t) - K3(1.0))
        return nu1 * ((t - 1.0) + C[2] * (L3(t) - L3(1.0)))

    out = np.zeros_like(tau)
    a = (tau > tp) & (tau <= 1.0)
    out[a] = partA(tau[a])
    A1 = partA(np.array(1.0))
    b = (tau > 1.0) & (tau <= tr)
    out[b] = A1 + partB(tau[b])
    B1 = partB(np.array(tr))
    c = tau > tr
    last = 1.0 if comp == "r" else nu1
    out[c] = A1 + B1 + last * (tau[c] - tr)
    return out


def ricker_dddot(f0, t):
    """Third time derivative of ricker_t."""
    tt = t - 1.5 / f0
    a = (np.pi * f0) ** 2
    e = np.exp(-a * tt ** 2)
    # d/dt of ricker_ddot's (-6a + 24a^2 tt^2 - 8a^3 tt^4) e
    return (48.0 * a * a * tt - 32.0 * a ** 3 * tt ** 3) * e \
        - 2.0 * a * tt * (-6.0 * a + 24.0 * a * a * tt ** 2 - 8.0 * a ** 3 * tt ** 4) * e


def lamb_shot(offsets, vp, vs, rho, f0, dt, nt, comp="r", shift=V_SHIFT, resamp=20):
    """Surface particle velocity (nt, len offsets) from a horizontal point
    force with the Ricker time function on the free surface of a
    homogeneous half space, up to a common positive factor: v = s''' * W,
    W the time integral of the step response (lamb_step_integral), sampled
    at t_n + shift dt on a grid resamp times finer."""
    pois = 0.5 * (vp * vp - 2 * vs * vs) / (vp * vp - vs * vs)
    mu = rho * vs * vs
    dtf = dt / resamp
    nf = nt * resamp
    tf = np.arange(nf + 1) * dtf
    s3 = ricker_dddot(f0, np.arange(nf) * dtf + shift * dt)
    out = np.zeros((nt, len(offsets)))
    for i, r in enumerate(offsets):
        W = lamb_step_integral(tf * vs / r, pois, comp) * (r / vs) / (2 * np.pi * mu * r)
        v = np.fft.irfft(np.fft.rfft(W[:nf], 2 * nf) * np.fft.rfft(s3, 2 * nf), 2 * nf)[:nf] * dtf
        out[:, i] = v[::resamp]
    return out
