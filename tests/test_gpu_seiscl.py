"""The SeisCL-compatible front end on the GPU: a two-shot FWI round
(forward, observed data, gradient) through SeisCL.set_forward / execute /
read_data / read_grad / read_rms gives what the engine and the oracle give
shot by shot."""
import numpy as np
import pytest

from cases import rel_l2

pytestmark = pytest.mark.gpu


def test_fwi_round(oracle_lib):
    from seiscl_amd.seiscl import SeisCL
    from seiscl_amd.workloads import layered_model
    N = np.array([60, 90])
    seis = SeisCL(N=N, ND=2, dh=10.0, dt=1e-3, NT=220, FDORDER=4, nab=10, seisout=1,
                  f0=15.0)
    seis.surface_acquisition_2d(dg=3, ds=25)
    rng = np.random.default_rng(2)
    model = layered_model(tuple(N), rng, nlayers=4)
    true = {k: (v * (1 + 0.03 * rng.standard_normal(v.shape))).astype(np.float32)
            for k, v in model.items()}
    ids = [0, 1]
    seis.set_forward(ids, true, withgrad=False)
    seis.execute()
    dobs = seis.read_data()
    assert [d.shape for d in dobs] == [(220, seis.rec_pos.shape[1])] * 2
    # the oracle, shot by shot
    csts = dict(N=N, ND=2, dh=10.0, dt=1e-3, NT=220, FDORDER=4, nab=10, seisout=1, f0=15.0)
    refs, grads = [], []
    for sid in ids:
        sc = seis.src_pos[3].astype(int) == sid
        rc = seis.rec_pos[3].astype(int) == sid
        shot = (seis.src_pos[:, sc], seis.src[:, sc], seis.rec_pos[:, rc])
        ref = oracle_lib.forward(true, *shot, **csts)
        for d, r in zip(dobs, ref):
            assert rel_l2(d[:, rc], r) <= 1e-6
        din = ref
        g = oracle_lib.gradient(model, *shot, din=din, **csts)["grad"]
        grads.append(g)
    seis.write_data({"vx": dobs[0], "vz": dobs[1]})
    seis.set_forward(ids, model, withgrad=True)
    seis.execute()
    gout = seis.read_grad()
    rms, norm = seis.read_rms()
    assert rms > 0 and norm > 0
    nab = seis.nab
    for name, g in zip(seis.params, gout):
        tot = grads[0][name].astype(np.float64) + grads[1][name]
        tot[:nab, :] = 0
        tot[-nab:, :] = 0
        tot[:, :nab] = 0
        tot[:, -nab:] = 0
        assert rel_l2(g, tot) <= 1e-5, name
    seis.close()


@pytest.mark.parametrize("misfit", [dict(scalermsnorm=1), dict(scaleshot=1, fmin=4.0, fmax=30.0),
                                    dict(restype=1)],
                         ids=lambda o: "-".join(f"{k}{v}" for k, v in o.items()))
def test_fwi_round_misfit_options(oracle_lib, misfit):
    """The misfit options (src/residuals.c:24-450) through the in-process
    front end: gradients and written residuals equal the oracle's."""
    from seiscl_amd.seiscl import SeisCL
    from seiscl_amd.workloads import layered_model
    N = np.array([60, 90])
    seis = SeisCL(N=N, ND=2, dh=10.0, dt=1e-3, NT=220, FDORDER=4, nab=10, seisout=1,
                  f0=15.0, **misfit)
    seis.surface_acquisition_2d(dg=3, ds=25)
    rng = np.random.default_rng(2)
    model = layered_model(tuple(N), rng, nlayers=4)
    true = {k: (v * (1 + 0.03 * rng.standard_normal(v.shape))).astype(np.float32)
            for k, v in model.items()}
    csts = dict(N=N, ND=2, dh=10.0, dt=1e-3, NT=220, FDORDER=4, nab=10, seisout=1, f0=15.0,
                **misfit)
    ids = [0, 1]
    seis.set_forward(ids, true, withgrad=False)
    seis.execute()
    dobs = seis.read_data()
    seis.write_data({"vx": dobs[0], "vz": dobs[1]})
    seis.resout = 1
    seis.set_forward(ids, model, withgrad=True)
    seis.execute()
    gout = seis.read_grad()
    res = seis.read_data(residuals=True)
    tot = None
    for sid in ids:
        sc = seis.src_pos[3].astype(int) == sid
        rc = seis.rec_pos[3].astype(int) == sid
        shot = (seis.src_pos[:, sc], seis.src[:, sc], seis.rec_pos[:, rc])
        r = oracle_lib.gradient(model, *shot, din=[d[:, rc] for d in dobs], **csts)
        for a, b in zip(res, r["res"]):
            assert np.abs(b).max() > 0
            assert rel_l2(a[:, rc], b) <= 1e-6
        g = {k: v.astype(np.float64) for k, v in r["grad"].items()}
        tot = g if tot is None else {k: tot[k] + g[k] for k in g}
    nab = seis.nab
    for name, g in zip(seis.params, gout):
        t = tot[name]
        t[:nab, :] = 0
        t[-nab:, :] = 0
        t[:, :nab] = 0
        t[:, -nab:] = 0
        assert rel_l2(g, t) <= 1e-5, name
    seis.close()
