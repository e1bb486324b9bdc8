"""A Fortran program drives the library through sunhip_bridge, as the reference's control code
would (src/dqmc_ft_ctrl.f90:37-49, :68-104): allocate_core -> input_conf ->
dqmc_ft_sweep_start_0b -> dqmc_ft_sweep_b0 / 0b -> the HS field, G, diagnostics and bins.

CPU: sunhubbardmy_amd/fortran/ft_driver.f90 compiles with the image's flang and links against
libsunhip.so.  GPU: the program's output after two sweeps equals the oracle run of the same
model and seed (HS field bit-identical, G within 1e-10 relative, acceptance counts equal,
observable bin sums within 1e-9).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from sunhubbardmy_amd import Model, chain_seed, make_tables
from sunhubbardmy_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(HERE, "..", "sunhubbardmy_amd", "fortran")
FLANG = shutil.which("flang") or ("/opt/rocm/lib/llvm/bin/flang" if os.path.exists("/opt/rocm/lib/llvm/bin/flang") else None)

KW = dict(lattice="square", la=4, lb=4, beta=2.0, dtau=0.1, rhub=4.0, kmode="checkerboard", update="delay",
          nublock=4, nwrap=10)
NSWEEP = 2


def _build(tmp_path):
    lib = os.path.abspath(_lib.LIB_PATH)
    subprocess.run([FLANG, "-c", os.path.abspath(os.path.join(FDIR, "sunhip_bridge.f90")), "-o", "bridge.o"],
                   cwd=tmp_path, check=True, capture_output=True)
    subprocess.run([FLANG, os.path.abspath(os.path.join(FDIR, "ft_driver.f90")), "bridge.o", lib,
                    "-Wl,-rpath," + os.path.dirname(lib), "-o", "ft_driver"], cwd=tmp_path, check=True,
                   capture_output=True)
    return tmp_path / "ft_driver"


def _write_tables(path, m, nsweep):
    t = make_tables(m)
    lat, hs = t["lat"], t["hs"]
    n = t["ndim"]
    hdr = np.array([n, t["lq"], t["ltrot"], t["nwrap"], t["nst"], t["lwrapT"], t["lwrapu"], t["kmode"],
                    lat["nn_nf"], lat["nn_lf"], hs["lcomp"], m.nflr, t["phase_from_conf"], t["update"],
                    t["nublock"], int(np.int32(m.seed)), nsweep], np.int32)
    cz = lambda a: np.ascontiguousarray(a, np.complex128)
    parts = [hdr, np.array([lat["kin_weight"]], np.float64), np.array(t["fft_dims"], np.int32),
             np.asarray(t["iwrap_nt"], np.int32), np.asarray(t["wrap_step"], np.int32).reshape(-1),
             np.asarray(lat["bonds"], np.int32).reshape(-1), cz(t["urt"]), cz(t["urtm1"]), cz(hs["bmat"]),
             cz(hs["bmat_inv"]), np.asarray(hs["phase_arg"], np.float64),
             cz(hs["delta_bmat"].reshape(-1, order="F")), cz(hs["phase_ratio"].reshape(-1, order="F")),
             np.asarray(hs["gaml"], np.float64), np.asarray(lat["stag"], np.float64)]
    with open(path, "wb") as f:
        for a in parts:
            f.write(np.ascontiguousarray(a).tobytes())
    return t


@pytest.mark.skipif(FLANG is None, reason="no flang in this image")
def test_fortran_driver_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    assert exe.exists()


@pytest.mark.gpu
@pytest.mark.skipif(FLANG is None, reason="no flang in this image")
def test_fortran_driver_matches_the_oracle(tmp_path):
    exe = _build(tmp_path)
    seed = chain_seed(0)
    m = Model(seed=seed, **KW)
    t = _write_tables(tmp_path / "tables.bin", m, NSWEEP)
    print("fortran driver: 4x4, 2 sweeps, seed", seed, flush=True)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "sweeps done" in r.stdout
    n, ltrot = t["ndim"], t["ltrot"]
    raw = (tmp_path / "out.bin").read_bytes()
    off = 0
    conf = np.frombuffer(raw, np.int32, n * ltrot, off).reshape(ltrot, n)
    off += 4 * n * ltrot
    g = np.frombuffer(raw, np.complex128, n * n, off).reshape(n, n, order="F")
    off += 16 * n * n
    diag = np.frombuffer(raw, np.float64, 10, off)
    obs = np.frombuffer(raw, np.float64, 11, off + 80)
    o = oracle.OracleChain(seed=seed, **KW)
    o.init_conf()
    o.sweep_start()
    for _ in range(NSWEEP):
        o.sweep_b0(True, True)
        o.sweep_0b(True, True)
    assert np.array_equal(conf, o.conf)
    assert np.abs(g - o.G).max() / np.abs(o.G).max() < 1e-10
    do = o.diag()
    assert diag[6] == do["accepted"] and diag[7] == do["tried"]
    assert diag[0] < 1e-8
    oo = o.obs()
    assert obs[0] == oo["nobs"]
    z = obs[1::2] + 1j * obs[2::2]
    for k, name in enumerate(("sign", "density", "ekin", "dblocc", "s_stag")):
        assert abs(z[k] - oo[name]) <= 1e-9 * max(1.0, abs(oo[name])), name
