"""Pins of the oracle's time-displaced path (ltau): green_tau (src/dqmc_ft_core/green_tau.f90),
the 0b pass without updates that carries G(t,0) and G(0,t) (dqmc_ft_sweep_0b.f90:36-49,
:160-204, dqmc_ft_core/dyn.f90) and the znn_tau bins (src/dqmc_measure/dyn_measure.f90,
dyn_measure1d.f90, dyn_measure3d.f90) on the square, honeycomb, chain and cubic lattices.

For one fixed HS field every Green's function has a closed form in the dense slice matrices
B(t) (built here with numpy from the bond tables, independently of the oracle's propagation):
G(0,0) = (1 + B(beta,0))^-1, G(t,0) = B(t,0) G(0,0), G(0,t) = -(1 - G(0,0)) B(t,0)^-1,
G(t,t) = B(t,0) G(0,0) B(t,0)^-1; at t = beta they reduce to the antiperiodicity relations
G(beta,0) = 1 - G(0,0) and G(0,beta) = -G(0,0).  Every B(t) conserves the particle number, so
the correlation summed over all pairs is the same at every t for one field."""
import numpy as np
import pytest

import oracle
from sunhubbardmy_amd import Model, chain_seed
from sunhubbardmy_amd.model import make_tables

CASES = {
    "square4_u4": dict(lattice="square", la=4, lb=4, beta=1.0, rhub=4.0, kmode="checkerboard", nwrap=4),
    "honeycomb2_u3": dict(lattice="honeycomb", la=2, lb=2, beta=1.2, rhub=3.0, kmode="checkerboard", nwrap=3),
    "chain8_u4": dict(lattice="chain", la=8, lb=1, beta=1.0, rhub=4.0, kmode="checkerboard", nwrap=4),
    "cubic4_u4": dict(lattice="cubic", la=4, lb=4, lc=4, beta=0.6, rhub=4.0, kmode="checkerboard", nwrap=3),
    "square4_tv": dict(lattice="square", la=4, lb=4, beta=1.0, rhub=0.0, rv=1.5, nflr=1, kmode="checkerboard",
                       nwrap=4),
}


def _cells(t):
    """1-based cell coordinates (x, y, z), orbital of every site from the site numbering, and the
    cell grid (l1, l2, l3)"""
    m, n, lq = t["model"], t["ndim"], t["lq"]
    if m.lattice == "chain":
        return [(s + 1, 1, 1, 0) for s in range(n)], (m.la, 1, 1)
    if m.lattice == "cubic":
        a = m.la * m.lb
        return [((s % a) // m.lb + 1, s % m.lb + 1, s // a + 1, 0) for s in range(n)], (m.la, m.lb, m.lc)
    nsub = n // lq
    return [((s // nsub) // m.lb + 1, (s // nsub) % m.lb + 1, 1, s % nsub) for s in range(n)], (m.la, m.lb, 1)


def _vfield(t, confv, nt):
    """the t-V bond factors of slice nt, V_nn ... V_1 with V_f = prod over the bonds of u diag(bp, bm) u"""
    lat, n, m = t["lat"], t["ndim"], t["model"]
    al = np.arccosh(np.exp(m.dtau * m.rv * 0.5))
    u = np.array([[1.0, 1.0], [1.0, -1.0]]) / np.sqrt(2.0)
    V = np.eye(n)
    for nf in range(lat["nn_nf"]):
        for q in range(lat["nn_lf"]):
            i1, i2 = lat["bonds"][nf, q]
            e = (-1.0, 1.0)[confv[nt, nf, q] - 1]
            E = np.eye(n)
            E[np.ix_([i1, i2], [i1, i2])] = u @ np.diag([np.exp(al * e), np.exp(-al * e)]) @ u
            V = E @ V
    return V


def _slices(t, conf, bm, confv=None):
    """[B(1), ..., B(ltrot)] with B(t) = A2 exp(V(t)) [bond factors] A1 from the checkerboard bond
    matrices"""
    lat, n = t["lat"], t["ndim"]
    fams = []
    for nf in range(lat["nn_nf"]):
        F = np.eye(n, dtype=complex)
        for b in range(lat["nn_lf"]):
            ist = nf * lat["nn_lf"] + b
            i1, i2 = lat["bonds"][nf, b]
            E = np.eye(n, dtype=complex)
            E[np.ix_([i1, i2], [i1, i2])] = t["urt"][4 * ist:4 * ist + 4].reshape(2, 2, order="F")
            F = E @ F
        fams.append(F)
    A1 = np.eye(n, dtype=complex)
    for F in fams:
        A1 = F @ A1
    A2 = np.eye(n, dtype=complex)
    for F in fams[::-1]:
        A2 = F @ A2
    if confv is None:
        return [A2 @ np.diag(bm[conf[nt] - 1]) @ A1 for nt in range(t["ltrot"])]
    return [A2 @ np.diag(bm[conf[nt] - 1]) @ _vfield(t, confv, nt) @ A1 for nt in range(t["ltrot"])]


def _run(name, ntdm=None):
    kw = CASES[name]
    seed = chain_seed(1)
    o = oracle.OracleChain(seed=seed, **kw)
    o.init_conf()
    o.sweep_start()
    o.sweep(lupdate=True, lmeasure=False)
    o.sweep_b0(lupdate=False, lmeasure=False)     # the left chain the dynamic 0b pass needs
    o.set_dyn(ntdm)
    o.set_dyn_trace(o.ltrot + 1)
    o.sweep_0b_dyn()
    return o, make_tables(Model(seed=seed, **kw))


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", CASES)
def test_dyn_pass_matches_dense_products(name):
    o, t = _run(name)
    n = t["ndim"]
    Bs = _slices(t, o.conf, t["hs"]["bmat"], o.confv if CASES[name].get("rv") else None)
    Bb = np.eye(n)
    for B in Bs:
        Bb = B @ Bb
    G00 = np.linalg.inv(np.eye(n) + Bb)
    tr = o.dyn_trace()
    assert tr.shape[0] == o.ltrot + 1
    Bt = np.eye(n, dtype=complex)
    for tau in range(o.ltrot + 1):
        if tau:
            Bt = Bs[tau - 1] @ Bt
        Bi = np.linalg.inv(Bt)
        gt0, g0t, gtt, g00 = tr[tau]
        assert _rel(gt0, Bt @ G00) < 1e-9, tau
        assert _rel(g0t, -(np.eye(n) - G00) @ Bi) < 1e-9, tau
        assert _rel(gtt, Bt @ G00 @ Bi) < 1e-9, tau
        assert _rel(g00, G00) < 1e-9, tau
    gt0, g0t, g00 = o.gtau()
    assert np.abs(gt0 - (np.eye(n) - G00)).max() < 1e-9
    assert np.abs(g0t + G00).max() < 1e-9
    # the propagated G(t,0), G(0,t) agree with green_tau's at every stabilization (-DDYNERROR)
    e = o.dyn_error()
    assert e["count"] == 2 * (o.nst) and e["xmax_dyn"] < 1e-10


@pytest.mark.parametrize("name", CASES)
def test_znn_bins(name):
    o, t = _run(name)
    z = o.znn()                                     # [t][orbital j][orbital i][imj]
    n, nflr = t["ndim"], t["model"].nflr
    cells, (l1, l2, l3) = _cells(t)
    phase = o.diag()["phase"]
    zphi = np.sign(phase.real) / phase.real * phase
    ref = np.zeros_like(z)
    for tau, (gt0, g0t, gtt, g00) in enumerate(o.dyn_trace()):
        v = np.outer(0.5 - gtt.diagonal().real, 0.5 - g00.diagonal().real) - g0t.T * gt0 / nflr
        for i, (xi, yi, zi, oi) in enumerate(cells):
            for j, (xj, yj, zj, oj) in enumerate(cells):
                dx = (xi - xj - 1) % l1 + 1
                dy = (yi - yj - 1) % l2 + 1
                dz = (zi - zj - 1) % l3 + 1
                ref[tau, oj, oi, (dz - 1) * l1 * l2 + dy + (dx - 1) * l2 - 1] += v[i, j] * zphi
    assert np.abs(z - ref).max() < 1e-10
    # number conservation: the total over all pairs does not depend on t for one field
    tot = z.sum(axis=(1, 2, 3))
    assert np.abs(tot - tot[0]).max() < 1e-9 * max(abs(tot[0]), 1.0)


def test_ntdm_shorter_than_ltrot():
    o, t = _run("square4_u4", ntdm=3)
    z = o.znn()
    assert z.shape[0] == 4 and np.abs(z).max() > 0
    assert o.dyn_trace().shape[0] == o.ltrot + 1
