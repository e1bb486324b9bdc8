"""The plaquette field of the finite-temperature sweep on the host tables and the oracle
(src/dqmc_config_plqu/*.f90 in dqmc_ft_sweep_0b.f90:82-93, dqmc_ft_sweep_b0.f90:91-100, the B
products of src/dqmc_ft_core.f90:91-96, :158-163, set_phase.f90:9-15).

Pins:
* tables: V Hermitian with the ring of the lattice's plaquette (consecutive sites of every cycle
  are nearest neighbours, every link carries alpha), bmat = expm(i a eta V) (scipy), bmat bmat_inv
  = 1, delta = bmat(isp) bmat_inv(is) - 1, phase_ratio = phase(isp) / phase(is), and the
  four-component identity sum_l gamma_l exp(i a eta_l V) / 4 = exp(-a^2 V^2) + O((a |V|)^8);
  the plaquette order puts the classes one after the other and covers every site 4 (square) or 3
  (honeycomb) times;
* after every pass the oracle's G equals (1 + B(beta, 0))^-1 with B built densely by numpy from the
  host tables and the final fields, B(t) = A2 exp(V_u) P_lq ... P_1 A1 (a wrong plaquette order or
  factor fails it);
* on a 2x2 square lattice with 2 slices (4^8 plaquette fields) the Monte Carlo averages of density,
  kinetic energy and sum n_i^2 agree with the exact sums sum_C W_C O_C / sum_C W_C, W_C =
  prod gamma phase det(1 + B_C)^nflr, within 4 standard errors."""
import itertools

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from sunhubbardmy_amd.model import PLQ_ETA, PLQ_GAML, Model, make_tables


def _tables(**kw):
    return make_tables(Model(**kw))


@pytest.mark.parametrize("lattice,l,theta", [("square", 4, 0.0), ("square", 6, 0.3), ("honeycomb", 3, 0.25)])
def test_plq_tables(lattice, l, theta):
    t = _tables(lattice=lattice, la=l, lb=l, beta=0.4, dtau=0.1, rhub=0.0, rhub_plq=1.7, alpha=0.35, theta=theta)
    p, lat = t["plq"], t["lat"]
    z, v, a = p["z"], p["v"], p["a"]
    assert z == (4 if lattice == "square" else 6) and p["lqp"] == t["lq"]
    assert np.abs(v - v.conj().T).max() == 0.0
    assert np.allclose(np.diag(v), 1.0 / lat["nn_nf"])
    for i in range(z):
        j = (i + 1) % z
        assert abs(abs(v[j, i]) - 0.35) < 1e-15          # ring links
        for k in range(z):
            if k not in (i, j, (i - 1) % z):
                assert v[k, i] == 0
    nb = {(int(s), int(x)) for s in range(t["ndim"]) for x in lat["nnlist"][s]}
    cnt = np.zeros(t["ndim"], int)
    for cyc in p["cord"]:
        assert len(set(cyc)) == z
        for i in range(z):
            assert (int(cyc[i]), int(cyc[(i + 1) % z])) in nb      # a closed walk over nn bonds
        cnt[cyc] += 1
    assert np.all(cnt == (4 if lattice == "square" else 3))
    for is_ in range(4):
        assert np.abs(p["b"][is_] - sl.expm(1j * a * PLQ_ETA[is_] * v)).max() < 1e-13
        assert np.abs(p["b"][is_] @ p["bi"][is_] - np.eye(z)).max() < 1e-13
        assert np.abs(p["bi"][is_] - p["b"][is_].conj().T).max() < 1e-14
        for fl in range(1, 4):
            isp = (is_ + fl) % 4
            assert np.abs(p["delta"][is_, fl - 1] - (p["b"][isp] @ p["bi"][is_] - np.eye(z))).max() < 1e-13
            assert abs(p["pratio"][is_, fl - 1] - np.exp(1j * (p["parg"][isp] - p["parg"][is_]))) < 1e-14
    # the four-component decoupling at a small step: a |V| ~ 0.1
    ts = _tables(lattice=lattice, la=l, lb=l, beta=0.4, dtau=0.004, rhub=0.0, rhub_plq=1.7, alpha=0.35, theta=theta)
    ps = ts["plq"]
    lhs = sum(PLQ_GAML[k] * ps["b"][k] for k in range(4)) / 4.0
    rhs = sl.expm(-ps["a"] ** 2 * ps["v"] @ ps["v"])
    assert np.abs(lhs - rhs).max() < 1e-7


def test_plaquette_order_square():
    """The first plaquettes of the 4x4 square lattice: class A cells (1,1), (1,3), (3,1), (3,3),
    then class B from (2,1) (src/lattice.f90:919-958)."""
    c = _tables(lattice="square", la=4, lb=4, rhub=0.0, rhub_plq=1.0)["plq"]["cord"]
    site = lambda x, y: ((x - 1) % 4) * 4 + (y - 1) % 4
    assert c[0].tolist() == [site(1, 1), site(2, 1), site(2, 2), site(1, 2)]
    assert c[1].tolist() == [site(1, 3), site(2, 3), site(2, 4), site(1, 4)]
    assert c[3].tolist() == [site(3, 3), site(4, 3), site(4, 4), site(3, 4)]
    assert c[4].tolist() == [site(2, 1), site(3, 1), site(3, 2), site(2, 2)]


def _kin(t):
    n = t["ndim"]
    if t["kmode"] == 1:
        K = t["urt"].reshape(n, n, order="F")
        return K, K
    lat = t["lat"]
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
    return A1, A2


def _slice(t, confp_nt, conf_nt=None):
    """A2 exp(V_u) P_lq ... P_1 A1 of one slice"""
    n, p = t["ndim"], t["plq"]
    A1, A2 = _kin(t)
    P = np.eye(n, dtype=complex)
    for q, cyc in enumerate(p["cord"]):
        E = np.eye(n, dtype=complex)
        E[np.ix_(cyc, cyc)] = p["b"][confp_nt[q] - 1]
        P = E @ P
    U = np.eye(n) if conf_nt is None or t["model"].rhub == 0 else np.diag(t["hs"]["bmat"][conf_nt - 1])
    return A2 @ U @ P @ A1


def _b_beta(t, confp, conf=None):
    B = np.eye(t["ndim"], dtype=complex)
    for nt in range(t["ltrot"]):
        B = _slice(t, confp[nt], None if conf is None else conf[nt]) @ B
    return B


@pytest.mark.parametrize("kw", [
    dict(lattice="square", la=4, lb=4, rhub=3.0, nflr=2, kmode="checkerboard", theta=0.0),
    dict(lattice="honeycomb", la=3, lb=3, rhub=0.0, nflr=2, kmode="dense", theta=0.3),
    dict(lattice="square", la=4, lb=4, rhub=2.0, nflr=4, exact_hs=False, kmode="dense", theta=0.5),
])
def test_g_equals_dense_product_after_every_pass(kw):
    kw = dict(kw, beta=0.6, dtau=0.1, nwrap=3, rhub_plq=1.5, alpha=0.3)
    t = _tables(**kw)
    o = oracle.OracleChain(seed=oracle.chain_seed(3), **kw)
    o.init_conf()
    c0 = o.confp
    assert c0.shape == (t["ltrot"], t["lq"]) and set(np.unique(c0)) == {1, 2, 3, 4}
    o.sweep_start()
    n = o.ndim
    for p in ("b0", "0b"):
        getattr(o, "sweep_" + p)(lupdate=True, lmeasure=False)
        G = np.linalg.inv(np.eye(n) + _b_beta(t, o.confp, o.conf))
        assert np.abs(o.G - G).max() < 1e-9, p
    acc = o.pacc()
    assert 0 < acc[0] < acc[1] and acc[1] == 2 * o.ltrot * t["lq"]
    assert not np.array_equal(o.confp, c0)
    assert o.diag()["max_wrap_error"] < 1e-10


def test_exact_enumeration_square2():
    kw = dict(lattice="square", la=2, lb=2, beta=0.2, dtau=0.1, rhub=0.0, nflr=2, rt=1.0, kmode="dense", nwrap=1,
              rhub_plq=2.0, alpha=0.4, theta=0.0)
    t = _tables(**kw)
    n, lq, p = t["ndim"], t["lq"], t["plq"]
    assert t["ltrot"] == 2 and lq == 4
    # the slice matrices of all 4^4 plaquette fields of one slice, then every pair of slices
    fields = np.array(list(itertools.product((1, 2, 3, 4), repeat=lq)), np.int32)
    S = np.array([_slice(t, f) for f in fields])
    wslice = np.array([np.prod(PLQ_GAML[f - 1] * np.exp(1j * p["parg"][f - 1])) for f in fields])
    B = np.einsum("aij,bjk->abik", S, S).reshape(-1, n, n)       # slice 2 (a) after slice 1 (b)
    w = np.outer(wslice, wslice).reshape(-1)
    M = np.eye(n) + B
    W = w * np.linalg.det(M) ** kw["nflr"]
    G = np.linalg.inv(M)
    gc = np.eye(n)[None] - np.transpose(G, (0, 2, 1))
    lat = t["lat"]
    ek = sum(-2.0 * gc[:, i0, i1].real for nf in range(lat["nn_nf"]) for i0, i1 in lat["bonds"][nf])
    dens = np.trace(gc, axis1=1, axis2=2).real / n
    d2 = np.sum(np.diagonal(gc, axis1=1, axis2=2) ** 2, axis=1).real
    exact = np.array([np.sum(W * o) / np.sum(W) for o in (dens, ek, d2)]).real
    o = oracle.OracleChain(seed=oracle.chain_seed(4), **kw)
    o.init_conf()
    o.sweep_start()
    for _ in range(200):
        o.sweep(lupdate=True, lmeasure=False)
    vals = []
    for _ in range(60):
        o.reset_obs()
        for _ in range(60):
            o.sweep(lupdate=True, lmeasure=True)
        ob = o.obs()
        s = ob["sign"]
        vals.append([(ob["density"] / s).real, (ob["ekin"] / s).real, (ob["dblocc"] / s).real])
    vals = np.array(vals)
    mean = vals.mean(0)
    err = vals.std(0, ddof=1) / np.sqrt(len(vals))
    assert np.all(np.abs(mean - exact) < 4 * err + 1e-12), (mean, exact, err)
    assert np.all(err < 0.05 * np.abs(exact) + 1e-3)
    acc = o.pacc()
    assert 0.05 < acc[0] / acc[1] < 0.95
