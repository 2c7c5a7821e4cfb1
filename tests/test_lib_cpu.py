"""The C-ABI library loads and exports every entry point the header
declares; host-side plumbing works without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from seiscl_amd import _lib
from seiscl_amd.engine import from_c_layout, make_csts, to_c_layout

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(header="seiscl_amd.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(seis_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_cli_library_exports_every_declared_symbol():
    from seiscl_amd import h5mat
    lib = h5mat.load()
    names = header_functions("seiscl_cli.h")
    assert "seis_cli_main" in names and len(names) >= 7
    for n in names:
        assert hasattr(lib, n), n
    assert os.access(h5mat.BIN_PATH, os.X_OK)


def test_csts_struct_matches_c():
    lib = _lib.load()
    assert lib.seis_csts_size() == ctypes.sizeof(_lib.SeisCsts)
    c = _lib.SeisCsts()
    lib.seis_csts_default(ctypes.byref(c))
    # SeisCL.py constructor defaults (SeisCL/SeisCL.py:36-83)
    assert (c.ND, c.NT, c.FDORDER, c.nab, c.seisout) == (2, 875, 8, 16, 2)
    assert np.isclose(c.dt, 0.0008) and c.dh == 10 and c.K_MAX_CPML == 2


def test_make_csts_time_window():
    # tmax in seconds becomes a step count by float division and truncation,
    # as in read_csts (src/read_hdf5.c:523-526): 0.5f / 0.001f = 499.99997
    c = make_csts(N=[100, 200], NT=1000, dt=1e-3, tmax=0.5)
    assert c.tmax == 499 and c.N[0] == 100 and c.N[1] == 200


def test_layout_roundtrip():
    a = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)   # [NZ, NY, NX]
    flat = to_c_layout(a)
    # x slowest, z fastest: element (z=1, y=2, x=3) at x*NY*NZ + y*NZ + z
    assert flat[3 * 3 * 2 + 2 * 2 + 1] == a[1, 2, 3]
    np.testing.assert_array_equal(from_c_layout(flat, a.shape), a)


def test_context_fails_cleanly_without_device():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.seis_ctx_create(0, ctypes.byref(h)) != 0
    assert lib.seis_last_error()


def test_boundary_shell_maps_host(tmp_path):
    """The product's boundary shell (save enumeration / inject lookup) is a
    pair of inverse maps over in-grid cells, checked on the host."""
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    csrc = os.path.join(here, "..", "seiscl_amd", "csrc")
    exe = str(tmp_path / "shell_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-fopenmp", "-I", csrc, "-o", exe,
                    os.path.join(here, "shell_check.cpp"), os.path.join(csrc, "grad_host.cpp"),
                    os.path.join(csrc, "setup.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_autograd_module_needs_the_gpu():
    """The autograd wrapper imports without a GPU and fails loudly on use."""
    from seiscl_amd import Context
    from seiscl_amd.autograd import Propagator, seismic  # noqa: F401
    with pytest.raises(Exception):
        Context(0)


def test_seiscl_front_end_acquisition():
    """The SeisCL-compatible class selects shots, receivers and output names
    as SeisCL.py does (no GPU needed before execute)."""
    from seiscl_amd.seiscl import SeisCL
    seis = SeisCL(N=np.array([60, 120]), ND=2, nab=10, seisout=1, NT=300, dt=1e-3)
    seis.surface_acquisition_2d(dg=4, ds=20)
    assert seis.params == ["vp", "vs", "rho"]
    assert seis.to_load_names == ["vx", "vz"]
    nsrc = seis.src_pos_all.shape[1]
    assert nsrc == len(range(12, 120 - 12, 20))
    seis.prepare_data([1, 2])
    assert set(seis.src_pos[3].astype(int)) == {1, 2}
    assert set(seis.rec_pos[3].astype(int)) == {1, 2}
    ids = np.where(np.isin(seis.rec_pos_all[3].astype(int), [1, 2]))[0]
    np.testing.assert_array_equal(seis.rec_pos[4], ids + 1)
    assert seis.src.shape == (300, 2)
    seis.param_type, seis.L = 0, 1
    assert seis.params == ["vp", "vs", "rho", "taup", "taus"]
