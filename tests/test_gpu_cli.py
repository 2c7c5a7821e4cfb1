"""The SeisCL_MPI file boundary on the GPU (its body runs in this process
through seis_cli_main, so no program is started from a process holding the
GPU).  Files are written and read with SeisCL.py's conventions; results are
checked shot by shot against the oracle:

- forward seismograms (<var>out), 2D and 3D;
- gradient files (grad<name>) and the rms file for a two-shot FWI round with
  observed data;
- the inputres mode (adjoint sources from the residual file);
- resout: the scaled adjoint sources the reference writes as <var>res;
- movies (movout) of forward and of gradient runs.
"""
import os

import numpy as np
import pytest

from cases import rel_l2, small_case

pytestmark = pytest.mark.gpu


def _fwi_case(tmp_path, **kw):
    from seiscl_amd.seiscl import SeisCL
    from seiscl_amd.workloads import layered_model
    N = np.array([60, 90])
    args = dict(N=N, ND=2, dh=10.0, dt=1e-3, NT=220, FDORDER=4, nab=10, seisout=1, f0=15.0,
                files=True, spawn=False, workdir=str(tmp_path))
    args.update(kw)
    seis = SeisCL(**args)
    seis.surface_acquisition_2d(dg=3, ds=25)
    rng = np.random.default_rng(2)
    model = layered_model(tuple(N), rng, nlayers=4)
    true = {k: (v * (1 + 0.03 * rng.standard_normal(v.shape))).astype(np.float32)
            for k, v in model.items()}
    csts = {k: args[k] for k in ("N", "ND", "dh", "dt", "NT", "FDORDER", "nab", "seisout", "f0")}
    return seis, model, true, csts


def _shots(seis):
    for sid in sorted(set(seis.src_pos[3].astype(int))):
        sc = seis.src_pos[3].astype(int) == sid
        rc = seis.rec_pos[3].astype(int) == sid
        yield rc, (seis.src_pos[:, sc], seis.src[:, sc], seis.rec_pos[:, rc])


def _cropped(g, nab):
    g = g.astype(np.float64)
    g[:nab, :] = 0
    g[-nab:, :] = 0
    g[:, :nab] = 0
    g[:, -nab:] = 0
    return g


def test_forward_and_gradient_files(oracle_lib, tmp_path):
    seis, model, true, csts = _fwi_case(tmp_path)
    print("[cli] forward files", flush=True)
    seis.set_forward([0, 1], true, withgrad=False)
    seis.execute()
    dobs = seis.read_data()
    assert [d.shape for d in dobs] == [(220, seis.rec_pos.shape[1])] * 2
    refs = []
    for rc, shot in _shots(seis):
        ref = oracle_lib.forward(true, *shot, **csts)
        for d, r in zip(dobs, ref):
            assert rel_l2(d[:, rc], r) <= 1e-6
        refs.append((rc, shot, ref))
    # data file with every trace, rows by receiver id (read_seis)
    seis.write_data({"vx": dobs[0], "vz": dobs[1]})
    print("[cli] gradient files", flush=True)
    seis.rmsout = 1
    seis.Hout = 1
    seis.set_forward([0, 1], model, withgrad=True)
    seis.execute()
    gout = seis.read_grad()
    hout = seis.read_Hessian()
    rms, norm = seis.read_rms()
    tot = {k: 0 for k in seis.params}
    htot = {k: 0 for k in seis.params}
    r_sum = rn_sum = 0.0
    for rc, shot, ref in refs:
        g = oracle_lib.gradient(model, *shot, din=ref, Hout=1, **csts)
        for k in tot:
            tot[k] = tot[k] + g["grad"][k].astype(np.float64)
            htot[k] = htot[k] + g["hess"][k].astype(np.float64)
        r_sum += g["rms"][0]
        rn_sum += g["rms"][1]
    for name, g in zip(seis.params, gout):
        assert rel_l2(g, _cropped(tot[name], seis.nab)) <= 1e-5, name
    # the approximate Hessian, H<name> of the gradient file
    for name, h in zip(seis.params, hout):
        assert np.abs(htot[name]).max() > 0
        assert rel_l2(h, _cropped(htot[name], seis.nab)) <= 1e-5, name
    # the rms file holds sqrt of the sums (src/writehdf5.c:236-245)
    assert abs(norm - np.sqrt(rn_sum)) <= 1e-5 * np.sqrt(rn_sum)
    assert abs(rms - np.sqrt(r_sum) / np.sqrt(rn_sum)) <= 1e-5 * rms
    # dout carries src_pos and rec_pos as given
    dout = os.path.join(str(tmp_path), "SeisCL_dout.mat")
    from seiscl_amd import h5mat
    np.testing.assert_array_equal(h5mat.read(dout, "rec_pos"), seis.rec_pos.T.astype(np.float32))


def test_inputres_files(oracle_lib, tmp_path):
    seis, model, _, csts = _fwi_case(tmp_path)
    seis.prepare_data([0, 1])
    rng = np.random.default_rng(5)
    res = [rng.standard_normal((220, seis.rec_pos.shape[1])).astype(np.float32) for _ in range(2)]
    print("[cli] inputres files", flush=True)
    seis.set_forward([0, 1], model, withgrad=True)
    seis.set_backward(res)
    seis.execute()
    gout = seis.read_grad()
    tot = {k: 0 for k in seis.params}
    for rc, shot in _shots(seis):
        g = oracle_lib.gradient(model, *shot, res=[r[:, rc] for r in res], **csts)
        for k in tot:
            tot[k] = tot[k] + g["grad"][k].astype(np.float64)
    for name, g in zip(seis.params, gout):
        assert rel_l2(g, _cropped(tot[name], seis.nab)) <= 1e-5, name


@pytest.mark.parametrize("gradfix", [0, 223])
def test_gradfix_files(oracle_lib, tmp_path, gradfix):
    """/gradfix in the constants file selects the gradient corrections at the
    file boundary: 0 gives the reference's gradients (the oracle's reference
    mode), 223 the corrected ones; the two differ."""
    seis, model, true, csts = _fwi_case(tmp_path, gradfix=gradfix)
    print(f"[cli] gradfix={gradfix} files", flush=True)
    seis.set_forward([0, 1], true, withgrad=False)
    seis.execute()
    dobs = seis.read_data()
    seis.write_data({"vx": dobs[0], "vz": dobs[1]})
    seis.set_forward([0, 1], model, withgrad=True)
    seis.execute()
    gout = seis.read_grad()
    tot = {k: 0 for k in seis.params}
    other = {k: 0 for k in seis.params}
    for rc, shot in _shots(seis):
        din = [d[:, rc] for d in dobs]
        g = oracle_lib.gradient(model, *shot, din=din, gradfix=gradfix, **csts)
        o = oracle_lib.gradient(model, *shot, din=din, gradfix=223 - gradfix, **csts)
        for k in tot:
            tot[k] = tot[k] + g["grad"][k].astype(np.float64)
            other[k] = other[k] + o["grad"][k].astype(np.float64)
    for name, g in zip(seis.params, gout):
        ref = _cropped(tot[name], seis.nab)
        assert rel_l2(g, ref) <= 1e-5, name
        assert rel_l2(_cropped(other[name], seis.nab), ref) > 1e-3, name


@pytest.mark.parametrize("files", [True, False])
def test_gradient_movie(oracle_lib, tmp_path, files):
    """movout with gradout (src/time_stepping.c:926-929): the movie holds the
    backpropagated forward wavefields of the gradient run, frame
    (t+1)/movout - 1 after the adjoint step t; equal to the oracle's frames
    bit for bit, shot by shot, through the files (SeisCL_MPI) and in
    process; the gradient is unchanged by the recording."""
    seis, model, _, csts = _fwi_case(tmp_path, movout=50, files=files)
    seis.prepare_data([0, 1])
    rng = np.random.default_rng(5)
    res = [rng.standard_normal((220, seis.rec_pos.shape[1])).astype(np.float32) for _ in range(2)]
    print(f"[cli] gradient movie files={files}", flush=True)
    seis.set_forward([0, 1], model, withgrad=True)
    seis.set_backward(res)
    seis.execute()
    gout = seis.read_grad()
    movie = seis.read_movie()
    assert [m.shape for m in movie] == [(60, 90, 220 // 50, 2)] * 2
    tot = {k: 0 for k in seis.params}
    for k, (rc, shot) in enumerate(_shots(seis)):
        g = oracle_lib.gradient(model, *shot, res=[r[:, rc] for r in res], movout=50, **csts)
        for name, mov in zip(("vx", "vz"), movie):
            ref = g["movie"][name]
            assert np.abs(ref).max() > 0
            for f in range(220 // 50):
                np.testing.assert_array_equal(mov[..., f, k], ref[f], err_msg=f"{name} {f}")
        for p in tot:
            tot[p] = tot[p] + g["grad"][p].astype(np.float64)
    for name, gr in zip(seis.params, gout):
        assert rel_l2(gr, _cropped(tot[name], seis.nab)) <= 1e-5, name
    seis.close()


MISFITS = [dict(), dict(scalermsnorm=1), dict(scaleshot=1, fmin=4.0, fmax=30.0),
           dict(scalerms=1), dict(restype=1)]


@pytest.mark.parametrize("misfit", MISFITS,
                         ids=lambda o: "-".join(f"{k}{v}" for k, v in o.items()) or "L2")
def test_resout_files(oracle_lib, tmp_path, misfit):
    seis, model, true, csts = _fwi_case(tmp_path, resout=1, rmsout=1, **misfit)
    csts.update(misfit)
    seis.prepare_data([0, 1])
    dobs = []
    for rc, shot in _shots(seis):
        dobs.append((rc, oracle_lib.forward(true, *shot, **csts)))
    full = [np.zeros((220, seis.rec_pos.shape[1]), np.float32) for _ in range(2)]
    for rc, d in dobs:
        for f, x in zip(full, d):
            f[:, rc] = x
    print(f"[cli] resout files {misfit}", flush=True)
    seis.set_forward([0, 1], model, withgrad=False)
    seis.write_data({"vx": full[0], "vz": full[1]})
    seis.execute()
    res = seis.read_data(residuals=True)
    for (rc, d), (_, shot) in zip(dobs, _shots(seis)):
        ref = oracle_lib.gradient(model, *shot, din=d, **csts)["res"]
        for a, r in zip(res, ref):
            assert np.abs(r).max() > 0
            assert rel_l2(a[:, rc], r) <= 1e-6


@pytest.mark.parametrize("seisout", [1, 2])
def test_forward_files_3d(oracle_lib, tmp_path, seisout):
    from seiscl_amd.seiscl import SeisCL
    csts, params, (sp, s, rp) = small_case(ND=3, FDORDER=4, L=2, freesurf=1, NT=80,
                                           seisout=seisout)
    keys = ("N", "ND", "dh", "dt", "NT", "FDORDER", "MAXRELERROR", "abs_type", "nab",
            "freesurf", "L", "FL", "f0", "seisout", "abpc")
    seis = SeisCL(files=True, spawn=False, workdir=str(tmp_path),
                  **{k: csts[k] for k in keys})
    seis.src_pos_all, seis.src_all, seis.rec_pos_all = sp, s, rp.copy()
    print(f"[cli] 3D forward seisout={seisout}", flush=True)
    seis.set_forward([0], params, withgrad=False)
    seis.execute()
    out = seis.read_data()
    ref = oracle_lib.forward(params, sp, s, seis.rec_pos, **csts)
    assert len(out) == len(ref)
    for o, r in zip(out, ref):
        assert rel_l2(o, r) <= 1e-6


@pytest.mark.parametrize("files", [True, False], ids=["files", "in_process"])
def test_movie(oracle_lib, tmp_path, files):
    """movout (src/time_stepping.c:272-350, src/writehdf5.c:295-310): the
    wavefields every movout steps, [NZ, NX, NT/movout, shots] after
    read_movie's transpose; frames equal the oracle's fields after as many
    steps, bit for bit."""
    from seiscl_amd.seiscl import SeisCL
    from seiscl_amd.workloads import layered_model
    N, NT, mv = np.array([60, 90]), 200, 50
    args = dict(N=N, ND=2, dh=10.0, dt=1e-3, NT=NT, FDORDER=4, nab=10, seisout=1, f0=15.0,
                movout=mv)
    if files:
        args.update(files=True, spawn=False, workdir=str(tmp_path))
    seis = SeisCL(**args)
    seis.surface_acquisition_2d(dg=3, ds=25)
    model = layered_model(tuple(N), np.random.default_rng(2), nlayers=3)
    print(f"[cli] movie files={files}", flush=True)
    seis.set_forward([0, 1], model, withgrad=False)
    seis.execute()
    movie = seis.read_movie()
    assert [m.shape for m in movie] == [(60, 90, NT // mv, 2)] * 2
    csts = dict(N=N, ND=2, dh=10.0, dt=1e-3, FDORDER=4, nab=10, seisout=1, f0=15.0)
    for k, sid in enumerate((0, 1)):
        sc = seis.src_pos[3].astype(int) == sid
        rc = seis.rec_pos[3].astype(int) == sid
        for f in (0, NT // mv - 1):
            nt = (f + 1) * mv
            _, fl = oracle_lib.forward(model, seis.src_pos[:, sc], seis.src[:nt, sc],
                                       seis.rec_pos[:, rc], fields=True, NT=nt, **csts)
            for name, mov in zip(("vx", "vz"), movie):
                assert np.abs(fl[name]).max() > 0
                np.testing.assert_array_equal(mov[..., f, k], fl[name][2:-2, 2:-2])
    seis.close()


def test_movie_tmax(tmp_path):
    """With tmax < NT the forward movie stops at tmax: the in-process front
    end and SeisCL_MPI return the same seismograms and frames (frames past
    tmax stay zero)."""
    from seiscl_amd.seiscl import SeisCL
    from seiscl_amd.workloads import layered_model
    N, NT, mv, dt = np.array([60, 90]), 200, 40, 1e-3
    model = layered_model(tuple(N), np.random.default_rng(3), nlayers=3)
    res = {}
    for files in (True, False):
        args = dict(N=N, ND=2, dh=10.0, dt=dt, NT=NT, FDORDER=4, nab=10, seisout=1, f0=15.0,
                    movout=mv, tmax=130 * dt)
        if files:
            args.update(files=True, spawn=False, workdir=str(tmp_path))
        seis = SeisCL(**args)
        seis.surface_acquisition_2d(dg=3, ds=25)
        print(f"[cli] movie tmax files={files}", flush=True)
        seis.set_forward([0], model, withgrad=False)
        seis.execute()
        res[files] = (seis.read_data(), seis.read_movie())
        seis.close()
    (d1, m1), (d0, m0) = res[True], res[False]
    for a, b in zip(d1, d0):
        assert np.abs(a).max() > 0
        np.testing.assert_array_equal(a, b)
    for a, b in zip(m1, m0):
        np.testing.assert_array_equal(a, b)
        assert np.abs(a[..., 130 // mv - 1, :]).max() > 0
        assert np.all(a[..., 130 // mv:, :] == 0)


def test_visco_gradient_files(oracle_lib, tmp_path):
    """Viscoelastic gradient through SeisCL_MPI: the gradient file holds
    grad<name> of every parameter, taup and taus included (the reference
    writes all five, src/writehdf5.c:261-283, and SeisCL.py reads them back
    for L > 0); each equals the oracle's, shot by shot."""
    from seiscl_amd.workloads import q_to_tau
    seis, model, true, csts = _fwi_case(tmp_path, L=2, FL=np.array([8.0, 30.0]))
    tau = q_to_tau(50.0, 5.0, 40.0, [8.0, 30.0])
    for p in (model, true):
        p["taup"] = np.full(p["vp"].shape, tau, np.float32)
        p["taus"] = np.full(p["vp"].shape, tau, np.float32)
    csts = dict(csts, L=2, FL=np.array([8.0, 30.0]))
    print("[cli] visco gradient files", flush=True)
    seis.set_forward([0, 1], true, withgrad=False)
    seis.execute()
    dobs = seis.read_data()
    seis.write_data({"vx": dobs[0], "vz": dobs[1]})
    seis.set_forward([0, 1], model, withgrad=True)
    seis.execute()
    gout = seis.read_grad()
    assert seis.params == ["vp", "vs", "rho", "taup", "taus"]
    tot = {k: 0 for k in seis.params}
    for rc, shot in _shots(seis):
        g = oracle_lib.gradient(model, *shot, din=[d[:, rc] for d in dobs], **csts)
        for k in tot:
            tot[k] = tot[k] + g["grad"][k].astype(np.float64)
    for name, g in zip(seis.params, gout):
        ref = _cropped(tot[name], seis.nab)
        assert np.abs(ref).max() > 0, name
        assert rel_l2(g, ref) <= 1e-5, name


@pytest.mark.parametrize("files,gradfix,L", [(True, 223, 0), (True, 0, 0), (False, 223, 0),
                                              (True, 223, 2)])
def test_inputres_two_pass(oracle_lib, tmp_path, files, gradfix, L):
    """The reference's two-pass inputres flow (src/time_stepping.c:731-745,
    850-859): pass 1 (inputres=1, gradout=0) models the shots and writes
    each shot's saved forward run to <base>_NP0_checkpoint.mat (in process:
    keeps it); the residuals are made outside; pass 2 (set_backward) reads
    the checkpoint instead of modeling again and runs the adjoint.  Pass 1's
    seismograms and pass 2's gradients equal the oracle's, in the reference's
    gradient mode (gradfix 0) and the corrected one."""
    kw = dict(inputres=1, gradfix=gradfix, files=files)
    if L:
        kw.update(L=L, FL=np.array([8.0, 30.0]))
    seis, model, _, csts = _fwi_case(tmp_path, **kw)
    if L:
        from seiscl_amd.workloads import q_to_tau
        tau = q_to_tau(50.0, 5.0, 40.0, [8.0, 30.0])
        model["taup"] = np.full(model["vp"].shape, tau, np.float32)
        model["taus"] = np.full(model["vp"].shape, tau, np.float32)
        csts = dict(csts, L=L, FL=np.array([8.0, 30.0]))
    print(f"[cli] inputres two-pass files={files} gradfix={gradfix} L={L}", flush=True)
    seis.set_forward([0, 1], model, withgrad=False)
    seis.execute()
    ckpt = os.path.join(str(tmp_path), "SeisCL_NP0_checkpoint.mat")
    assert os.path.exists(ckpt) == files
    dout = seis.read_data()
    refs = []
    for rc, shot in _shots(seis):
        ref = oracle_lib.forward(model, *shot, **csts)
        for d, r in zip(dout, ref):
            assert rel_l2(d[:, rc], r) <= 1e-6
        refs.append((rc, shot))
    # adjoint sources made outside the program: the modeled data scaled
    res = [(0.5 * d).astype(np.float32) for d in dout]
    seis.set_backward(res)
    seis.execute()
    gout = seis.read_grad()
    tot = {k: 0 for k in seis.params}
    for rc, shot in refs:
        g = oracle_lib.gradient(model, *shot, res=[r[:, rc] for r in res], gradfix=gradfix,
                                **csts)
        for k in tot:
            tot[k] = tot[k] + g["grad"][k].astype(np.float64)
    for name, g in zip(seis.params, gout):
        ref = _cropped(tot[name], seis.nab)
        assert np.abs(ref).max() > 0, name
        assert rel_l2(g, ref) <= 1e-5, (name, rel_l2(g, ref))


@pytest.mark.parametrize("files", [True, False])
def test_dft_gradient_files(oracle_lib, tmp_path, files):
    """back_prop_type=2 with /gradfreqs (src/calc_grad.c:377-968): the DFT
    gradient through the files (SeisCL_MPI) and in process, two shots,
    against the oracle's DFT gradient summed over the shots."""
    freqs = np.array([6.0, 12.0, 20.0], np.float32)
    seis, model, true, csts = _fwi_case(tmp_path, files=files, back_prop_type=2,
                                        gradfreqs=freqs)
    print(f"[cli] DFT gradient files={files}", flush=True)
    seis.set_forward([0, 1], true, withgrad=False)
    seis.execute()
    dobs = seis.read_data()
    seis.write_data({"vx": dobs[0], "vz": dobs[1]})
    seis.set_forward([0, 1], model, withgrad=True)
    seis.execute()
    gout = seis.read_grad()
    tot = {k: 0 for k in seis.params}
    for rc, shot in _shots(seis):
        din = [d[:, rc] for d in dobs]
        g = oracle_lib.gradient_dft(model, *shot, freqs, din=din, **csts)
        for k in tot:
            tot[k] = tot[k] + g["grad"][k]
    for name, g in zip(seis.params, gout):
        ref = _cropped(tot[name], seis.nab)
        assert np.abs(ref).max() > 0, name
        assert rel_l2(g, ref) <= 1e-5, name


def test_gradsrc_file(tmp_path):
    """gradsrcout=1 writes /gradsrc [allns][NT] into the gradient file
    (src/writehdf5.c:285-288).  The reference's source-gradient terms are
    commented out of its adjoint kernels, so its array holds the zeros of
    init_gradsrc; the file here holds the same, next to the usual
    gradients."""
    from seiscl_amd import h5mat
    seis, model, true, csts = _fwi_case(tmp_path, gradsrcout=1)
    seis.set_forward([0, 1], true, withgrad=False)
    seis.execute()
    dobs = seis.read_data()
    seis.write_data({"vx": dobs[0], "vz": dobs[1]})
    seis.set_forward([0, 1], model, withgrad=True)
    seis.execute()
    path = os.path.join(seis.workdir, seis.file_gout)
    g = h5mat.loadmat(path)
    assert "gradsrc" in g and "gradvp" in g
    nsrc = int(np.sum(np.isin(seis.src_pos[3].astype(int), [0, 1])))
    assert g["gradsrc"].size == nsrc * int(csts["NT"])
    assert not np.any(g["gradsrc"])
    assert np.abs(g["gradvp"]).max() > 0
