"""CPU tests of the host side: model tables, the C ABI library's exports and its argument checks
(no compute calls; there is no GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from sunhubbardmy_amd import Model, SweepError, chain_seed, make_tables
from sunhubbardmy_amd import _lib
from sunhubbardmy_amd.model import default_nublock, schedule

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(sh[pl]?_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = (_declared("sunhip.h") + _declared("sunhip_testing.h") + _declared("sunhip_proj.h") +
             _declared("sunhip_learn.h"))
    assert len(names) >= 60
    for n in names:
        assert hasattr(L, n), n
    assert set(_declared("sunhip.h")) == set(_lib.EXPORTS)
    assert set(_declared("sunhip_proj.h")) == set(_lib.EXPORTS_PROJ)
    assert set(_declared("sunhip_learn.h")) == set(_lib.EXPORTS_LEARN)


@pytest.mark.parametrize("lattice,la,kmode", [("square", 4, "checkerboard"), ("square", 6, "dense"),
                                              ("honeycomb", 3, "checkerboard"), ("honeycomb", 4, "dense")])
def test_tables_match_oracle(lattice, la, kmode):
    kw = dict(lattice=lattice, la=la, lb=la, kmode=kmode, rhub=4.0, beta=2.3, nwrap=7)
    t = make_tables(Model(**kw))
    o = oracle.OracleChain(**kw)
    assert np.abs(t["urt"] - o.table(4)).max() < 1e-14
    assert np.abs(t["urtm1"] - o.table(5)).max() < 1e-14
    assert np.array_equal(t["lat"]["bonds"].reshape(-1), o.table(6).astype(int))
    assert np.array_equal(t["iwrap_nt"], o.table(7).astype(int))
    assert np.array_equal(t["wrap_step"].reshape(-1), o.table(8).astype(int))
    assert t["nst"] == o.nst and t["ltrot"] == o.ltrot


@pytest.mark.parametrize("nflr,exact", [(2, True), (4, True), (4, False), (6, True)])
def test_hs_tables_match_oracle(nflr, exact):
    hs = make_tables(Model(nflr=nflr, exact_hs=exact, rhub=6.0))["hs"]
    o = oracle.OracleChain(nflr=nflr, exact_hs=exact, rhub=6.0)
    t0 = o.table(0).reshape(-1, 3)
    assert np.abs(hs["bmat"] - t0[:, 0]).max() == 0
    assert np.abs(hs["bmat_inv"] - t0[:, 1]).max() == 0
    assert np.abs(np.exp(1j * hs["phase_arg"]) - t0[:, 2]).max() < 1e-15
    assert np.abs(hs["delta_bmat"].reshape(-1, order="F") - o.table(1)).max() == 0
    assert np.abs(hs["phase_ratio"].reshape(-1, order="F") - o.table(2)).max() == 0


def test_schedule_matches_reference_rules():
    # ltrot multiple of nwrap
    nst, nw, iw, ws = schedule(80, 10)
    assert nst == 8 and nw == 10 and list(np.nonzero(iw)[0]) == list(range(10, 81, 10))
    assert ws[0].tolist() == [1, 10] and ws[-1].tolist() == [71, 80]
    # remainder slice block, and nwrap > ltrot clamps to ltrot
    nst, nw, iw, ws = schedule(23, 10)
    assert nst == 3 and iw[23] == 3 and ws[-1].tolist() == [21, 23]
    nst, nw, iw, ws = schedule(8, 10)
    assert nst == 1 and nw == 8 and ws[0].tolist() == [1, 8]


def test_default_nublock_tables():
    assert default_nublock(16, "delay") == 4
    assert default_nublock(256, "delay") == 16     # 256/5 = 51 < 64
    assert default_nublock(288, "delay") == 16
    assert default_nublock(400, "delay") == 32
    assert default_nublock(1400, "delay") == 64
    assert default_nublock(256, "submatrix") == 32
    # the reference's non-monotone entry (ndim/20 in [64,128) -> 16)
    assert default_nublock(1500, "submatrix") == 16


def test_chain_seed_matches_oracle():
    for c in range(50):
        assert chain_seed(c) == oracle.chain_seed(c)


def _model_struct(t, **over):
    from sunhubbardmy_amd.walkers import FTWalkers
    w = FTWalkers.__new__(FTWalkers)
    w._keep = []
    s = FTWalkers._shmodel(w, t)
    for k, v in over.items():
        setattr(s, k, v)
    return s, w._keep


def _create(structs):
    arr = (_lib.ShModel * len(structs))(*structs)
    ctx = ctypes.c_void_p()
    rc = _lib.lib().sh_create(ctypes.byref(ctx), len(structs), arr, 0)
    return rc, ctx


def test_create_rejects_bad_arguments_before_touching_the_device():
    t = make_tables(Model(la=4, lb=4, beta=2.0))
    t2 = make_tables(Model(la=4, lb=4, beta=3.0))
    s1, k1 = _model_struct(t)
    s2, k2 = _model_struct(t2)
    assert _create([s1, s2])[0] == -1          # chains disagree on ltrot
    s3, k3 = _model_struct(t, ndim=2048)
    assert _create([s3])[0] == -1              # above the 1024 supported
    s4, k4 = _model_struct(t, nublock=65)
    assert _create([s4])[0] == -1
    bad = np.array(t["lat"]["bonds"]).reshape(-1).copy()
    bad[3] = 16
    s5, k5 = _model_struct(t)
    s5.bonds = bad.ctypes.data_as(_lib.ip)
    assert _create([s5])[0] == -1              # bond index outside the lattice
    s6, k6 = _model_struct(t, lcomp=5)
    assert _create([s6])[0] == -1
    assert _lib.lib().sh_destroy(None) == 0
    assert _lib.lib().sh_sweep_b0(None, 1, 1) == -1
    assert _lib.lib().sh_last_error(None) == b"null context"


def test_create_rejects_batches_the_pass_loops_cannot_share():
    """The pass loops take the bond-family count and the on-site switch of the whole batch: a
    batch whose bond-field chains differ in nn_nf (square 4 families, chain 2), or whose chains
    differ in lwrapu (U = 0 next to U != 0), is refused with SH_E_ARG before any device call."""
    tsq = make_tables(Model(lattice="square", la=4, lb=4, beta=1.0, rv=1.0))
    tch = make_tables(Model(lattice="chain", la=16, lb=1, beta=1.0, rv=1.0))
    assert tsq["v"] is not None and tch["v"] is not None
    a, ka = _model_struct(tsq)
    b, kb = _model_struct(tch)
    assert _create([a, b])[0] == -1
    u0 = make_tables(Model(la=4, lb=4, beta=1.0, rhub=0.0))
    u4 = make_tables(Model(la=4, lb=4, beta=1.0, rhub=4.0))
    c, kc = _model_struct(u0)
    d, kd = _model_struct(u4)
    assert c.lwrapu == 0 and d.lwrapu == 1
    assert _create([c, d])[0] == -1
    assert _create([d, c])[0] == -1


def test_walkers_raise_on_mixed_batches():
    with pytest.raises(SweepError):
        from sunhubbardmy_amd import FTWalkers
        FTWalkers([Model(beta=2.0), Model(beta=3.0)])


# ---- projective (t-V) host side --------------------------------------------------------------
from sunhubbardmy_amd.proj import ProjModel, ShpModel, make_proj_tables, trial_hamiltonian  # noqa: E402


@pytest.mark.parametrize("lattice,la,kmode", [("honeycomb", 3, "dense"), ("honeycomb", 2, "checkerboard"),
                                              ("square", 4, "checkerboard")])
def test_proj_tables_match_oracle(lattice, la, kmode):
    m = ProjModel(lattice=lattice, la=la, lb=la, kmode=kmode, beta=1.3, nwrap=4, rv=1.7, nustock=6)
    t = make_proj_tables(m)
    o = oracle.OracleProjChain(lattice=lattice, la=la, lb=la, kmode=kmode, beta=1.3, nwrap=4, rv=1.7, nustock=6)
    assert (t["ndim"], t["ne"], t["ltrot"], t["nst"]) == (o.ndim, o.ne, o.ltrot, o.nst)
    assert np.abs(t["urt"] - o.table(4)).max() < 1e-14
    assert np.abs(t["urtm1"] - o.table(5)).max() < 1e-14
    assert np.array_equal(t["lat"]["bonds"].reshape(-1), o.table(6).astype(int))
    assert np.array_equal(t["iwrap_nt"], o.table(7).astype(int))
    assert np.abs(t["v"]["vt"].reshape(-1) - o.table(100)).max() < 1e-15
    assert np.abs(t["v"]["vd"].reshape(-1) - o.table(101)).max() < 1e-15


def test_proj_trial_hamiltonian_matches_oracle():
    seed = chain_seed(4)
    m = ProjModel(lattice="honeycomb", la=3, lb=3, beta=1.0, rndness=0.25, mu=0.1, seed=seed)
    t = make_proj_tables(m)
    o = oracle.OracleProjChain(lattice="honeycomb", la=3, lb=3, beta=1.0, rndness=0.25, mu=0.1, seed=seed)
    o.init_conf()
    o.set_trial()
    p_o, w_o = o.proj
    lat = t["lat"]
    nconf = t["ndim"] * lat["nn_nf"] * t["ltrot"]
    u = oracle.sfmt_stream(seed, nconf + lat["nn_lf"] * lat["nn_nf"])[nconf:]
    w, v = np.linalg.eigh(trial_hamiltonian(m, lat, u))
    assert np.abs(w - w_o).max() < 1e-12
    p = v[:, :t["ne"]]
    assert np.abs(p @ p.T - p_o @ p_o.conj().T).max() < 1e-11


def test_proj_default_stock_size():
    assert make_proj_tables(ProjModel(lattice="honeycomb", la=18, lb=18))["nustock"] == 64   # N = 648
    assert make_proj_tables(ProjModel(lattice="honeycomb", la=2, lb=2))["nustock"] == 32


def _proj_create(tables, **override):
    from sunhubbardmy_amd.proj import ProjWalkers
    w = ProjWalkers.__new__(ProjWalkers)
    w._keep = []
    arr = (ShpModel * len(tables))()
    for k, t in enumerate(tables):
        arr[k] = w._shpmodel(t)
        for key, val in override.items():
            setattr(arr[k], key, val)
    ctx = ctypes.c_void_p()
    return _lib.lib().shp_create(ctypes.byref(ctx), len(tables), ctypes.cast(arr, ctypes.c_void_p), 0)


def test_proj_create_rejects_bad_arguments():
    """Argument errors come back as SH_E_ARG (-1) before any device call."""
    t1 = make_proj_tables(ProjModel(beta=1.0))
    t2 = make_proj_tables(ProjModel(beta=2.0))
    assert _proj_create([t1, t2]) == -1            # chains disagree on ltrot
    assert _proj_create([t1], nustock=66) == -1    # stock above 64 rows
    assert _proj_create([t1], nustock=7) == -1     # odd stock
    assert _proj_create([t1], ne=0) == -1
    assert _proj_create([t1], ndim=2048) == -1
    assert _lib.lib().shp_destroy(None) == 0
    assert _lib.lib().shp_sweep_b0(None, 1, 1) == -1
    assert _lib.lib().shp_last_error(None) == b"null context"


def test_proj_onsite_update_variants_are_checked():
    """The G-based on-site updates need the on-site field without a bond field and a depth
    1..64; an unknown algorithm is refused; all before any device call."""
    from sunhubbardmy_amd.proj import PUPDATES
    hub = make_proj_tables(ProjModel(lattice="square", la=4, lb=4, beta=1.0, rv=0.0, rhub=4.0, nflr=2,
                                     update="delay", gblock=8))
    assert hub["update"] == PUPDATES["delay"] and hub["gblock"] == 8
    assert _proj_create([hub], gblock=0) == -1
    assert _proj_create([hub], gblock=65) == -1
    assert _proj_create([hub], update=4) == -1
    tuv = make_proj_tables(ProjModel(lattice="square", la=4, lb=4, beta=1.0, rv=1.0, rhub=4.0, nflr=2))
    assert _proj_create([tuv], update=2, gblock=8) == -1      # bond field present
    with pytest.raises(ValueError):
        make_proj_tables(ProjModel(lattice="square", la=4, lb=4, rv=1.0, rhub=4.0, update="submatrix_g"))
    with pytest.raises(ValueError):
        make_proj_tables(ProjModel(update="unknown"))


# ---- the DFT form of the kinetic term ---------------------------------------------------------
@pytest.mark.parametrize("kw", [dict(lattice="square", la=4, lb=6), dict(lattice="square", la=12, lb=12, mu=0.3),
                                dict(lattice="chain", la=10, lb=1), dict(lattice="cubic", la=4, lb=6, lc=4)])
def test_fft_tables_reproduce_the_dense_exponential(kw):
    """F^H diag(exph0k) F, applied as the kernels do (unnormalised transforms along each lattice
    direction of the site order, the diagonal divided by N), equals expm(-dtau H / 2) built from
    the bond list; the same for the inverse.  A transposed direction order or a sign error in the
    exponent fails on the non-square shapes."""
    import scipy.linalg
    from sunhubbardmy_amd.model import h0_matrix
    m = Model(kmode="fft", dtau=0.1, **kw)
    t = make_tables(m)
    n = t["ndim"]
    dims = t["fft_dims"]
    assert int(np.prod(dims)) == n
    H = h0_matrix(m, t["lat"])
    shp = tuple(dims[::-1])

    def apply(e, X, side):
        # left: columns transformed forward, scaled, transformed back; right: rows conjugate first
        Y = X if side == "left" else X.T
        A = Y.reshape(shp + (n,))
        fwd, bwd = (np.fft.fftn, np.fft.ifftn) if side == "left" else (np.fft.ifftn, np.fft.fftn)
        A = fwd(A, axes=(0, 1, 2), norm="ortho").reshape(n, n) * e[:, None]
        A = bwd(A.reshape(shp + (n,)), axes=(0, 1, 2), norm="ortho").reshape(n, n)
        return A if side == "left" else A.T
    eye = np.eye(n, dtype=complex)
    for e, s in ((t["urt"], -0.05), (t["urtm1"], 0.05)):
        ref = scipy.linalg.expm(s * H)
        assert np.abs(apply(e, eye, "left") - ref).max() < 1e-13
        assert np.abs(apply(e, eye, "right") - ref).max() < 1e-13
    assert np.abs(t["urt"] * t["urtm1"] - 1.0).max() < 1e-14


def test_fft_tables_refuse_lattices_they_cannot_diagonalise():
    with pytest.raises(ValueError):
        make_tables(Model(lattice="honeycomb", la=3, lb=3, kmode="fft"))
    with pytest.raises(ValueError):
        make_tables(Model(lattice="square", la=17, lb=4, kmode="fft"))
    from sunhubbardmy_amd.proj import ProjModel, make_proj_tables
    with pytest.raises(ValueError):
        make_proj_tables(ProjModel(kmode="fft"))


def test_create_rejects_bad_fft_dimensions():
    t = make_tables(Model(lattice="square", la=4, lb=6, beta=1.0, kmode="fft"))
    s, k = _model_struct(t)
    assert list(s.fft_dims) == [6, 4, 1]
    s.fft_dims[:] = [6, 5, 1]                  # product != ndim
    assert _create([s])[0] == -1
    s.fft_dims[:] = [17, 1, 1]                 # a length the transform does not support
    assert _create([s])[0] == -1
    s2, k2 = _model_struct(t)
    s2.kmode = 3                               # unknown kinetic form
    assert _create([s2])[0] == -1
