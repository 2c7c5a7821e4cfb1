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
