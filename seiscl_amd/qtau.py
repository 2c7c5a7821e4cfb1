"""Quality factor <-> tau of the generalized standard linear solid.

Host-side calibration used to set the taup/taus parameters for a target Q
over a frequency band, with the interface of SeisCL's QTAU
(SeisCL/Q_tau.py:25-147): Qprofile, calibrate_Qm1, tau_optimize, get_FL,
Q2tau, tau2Qi.

For relaxation times t_l = 1/(2 pi FL_l) and attenuation level tau, the
GSLS quality factor is (SeisCL/Q_tau.py:52-66)

    Q(w) = (1/tau + sum_l w^2 t_l^2/(1 + w^2 t_l^2)) / sum_l w t_l/(1 + w^2 t_l^2),

i.e. Re M / Im M of M(w) = M_R (1 + tau sum_l i w t_l/(1 + i w t_l)).  The
calibration tabulates the band average of 1/Q on 40 log-spaced tau between
1/(2 Qmax) and 2/Qmin and interpolates log tau against log <1/Q> linearly.
tau_optimize moves the relaxation times so that <1/Q> is as nearly
proportional to tau as possible (log-log slope 1) and flat over the band.
"""
from __future__ import annotations

import numpy as np


class QTAU:
    def __init__(self, Qmin: float, Qmax: float, fmin: float, fmax: float, FL):
        self.Qmin, self.Qmax = float(Qmin), float(Qmax)
        self.fmin, self.fmax = float(fmin), float(fmax)
        self.taus = [1.0 / (2.0 * np.pi * float(f)) for f in np.atleast_1d(FL)]
        self._table()

    # 1/Q(w) over the band and the table behind Q2tau
    def _omega(self):
        return 2.0 * np.pi * np.arange(self.fmin, self.fmax, 0.1)

    def Qprofile(self, tau: float, taus=None):
        """(omega, Q(omega)) over [fmin, fmax) in steps of 0.1 Hz."""
        t = np.asarray(self.taus if taus is None else taus, np.float64)
        w = self._omega()
        wt = w[:, None] * t[None, :]
        num = 1.0 / tau + np.sum(wt ** 2 / (1.0 + wt ** 2), axis=1)
        den = np.sum(wt / (1.0 + wt ** 2), axis=1)
        return w, num / den

    def calibrate_Qm1(self, taus=None):
        """(tau grid, band-averaged 1/Q, spread, spread measure, log-log
        slope) over 40 log-spaced tau, as QTAU.calibrate_Qm1 returns them:
        the spread is the standard deviation of 1/Q over the band at the
        largest tau of the grid, the measure that spread times the sum of
        1/<1/Q> over the grid, the slope the least-squares slope of
        log <1/Q> against log tau."""
        lt = np.linspace(np.log(1.0 / self.Qmax / 2.0), np.log(1.0 / self.Qmin * 2.0), 40)
        qi = np.zeros(lt.size)
        inv = None
        for k, l in enumerate(lt):
            inv = 1.0 / self.Qprofile(np.exp(l), taus)[1]
            qi[k] = np.mean(inv)
        spread = float(np.std(inv))
        slope = np.polyfit(lt, np.log(qi), 1)[0]
        return np.exp(lt), qi, spread, float(np.sum(spread / qi)), float(slope)

    def _table(self):
        tau, qi, _, _, _ = self.calibrate_Qm1()
        self._ltau, self._lqi = np.log(tau), np.log(qi)

    def tau_optimize(self):
        """Relaxation times making <1/Q> proportional to tau and flat over
        the band: minimise (slope - 1)^2 + summed relative spread, times
        bounded to [1e-6, 0.3] s."""
        from scipy.optimize import minimize

        def loss(t):
            _, _, _, spread, slope = self.calibrate_Qm1(taus=t)
            return (slope - 1.0) ** 2 + spread

        res = minimize(loss, np.asarray(self.taus), bounds=[(1e-6, 0.3)] * len(self.taus))
        self.taus = [float(v) for v in res.x]
        self._table()

    def get_FL(self):
        return [1.0 / (2.0 * np.pi * t) for t in self.taus]

    def Q2tau(self, Q):
        """tau whose band-averaged 1/Q is 1/Q (interpolated in log-log)."""
        return np.exp(np.interp(np.log(1.0 / np.asarray(Q, np.float64)), self._lqi, self._ltau))

    def tau2Qi(self, logtau, taus=None):
        """Band-averaged 1/Q of each tau, given as log tau (QTAU.tau2Qi takes
        the logarithm, SeisCL/Q_tau.py:133-147)."""
        lt = np.atleast_1d(np.asarray(logtau, np.float64))
        return np.array([np.mean(1.0 / self.Qprofile(np.exp(t), taus)[1]) for t in lt])
