"""The bond field of the finite-temperature t-V model on the oracle (src/dqmc_config_v/*.f90 in the
FT sweep: dqmc_ft_sweep_b0.f90:78-89, dqmc_ft_sweep_0b.f90:95-106, the B products of
src/dqmc_ft_core.f90:98-104, :150-156).

Pins: after every pass G equals (1 + B(beta, 0))^-1 with B built densely by numpy from the bond
matrices of the final fields (B(t) = A2 exp(V_u) V_nn ... V_1 A1, V_f the product over the bonds
of family f of u diag(bp, bm) u); and on a 4-site chain with 2 slices (2^8 bond fields) the Monte
Carlo averages of density, kinetic energy and sum n_i^2 agree with the exact weighted sums over
all fields, sum_C det(1 + B_C) O_C / sum_C det(1 + B_C), within 4 standard errors."""
import itertools

import numpy as np
import pytest

import oracle
from sunhubbardmy_amd.model import Model, make_tables


def _kin_fams(t):
    lat, n = t["lat"], t["ndim"]
    if t["kmode"] == 1:                                  # dense exp(-dtau K / 2) on both sides
        K = t["urt"].reshape(n, n, order="F")
        return K, K
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


def _b_beta(t, confv, rv, dtau, conf=None):
    lat, n = t["lat"], t["ndim"]
    al = np.arccosh(np.exp(dtau * rv * 0.5))
    eta = np.array([-1.0, 1.0])
    u = np.array([[1.0, 1.0], [1.0, -1.0]]) / np.sqrt(2.0)
    A1, A2 = _kin_fams(t)
    B = np.eye(n, dtype=complex)
    for nt in range(t["ltrot"]):
        V = np.eye(n)
        for nf in range(lat["nn_nf"]):
            Vf = np.eye(n)
            for q in range(lat["nn_lf"]):
                i1, i2 = lat["bonds"][nf, q]
                e = eta[confv[nt, nf, q] - 1]
                E = np.eye(n)
                E[np.ix_([i1, i2], [i1, i2])] = u @ np.diag([np.exp(al * e), np.exp(-al * e)]) @ u
                Vf = E @ Vf
            V = Vf @ V
        U = np.eye(n) if conf is None or t["model"].rhub == 0 else np.diag(t["hs"]["bmat"][conf[nt] - 1])
        B = A2 @ U @ V @ A1 @ B
    return B


@pytest.mark.parametrize("rhub,nflr,kmode", [(0.0, 1, "checkerboard"), (3.0, 2, "checkerboard"), (3.0, 2, "dense")])
def test_g_equals_dense_product_after_every_pass(rhub, nflr, kmode):
    kw = dict(lattice="square", la=4, lb=4, beta=1.0, dtau=0.1, rhub=rhub, nflr=nflr, kmode=kmode, nwrap=5)
    rv = 1.5
    o = oracle.OracleChain(seed=oracle.chain_seed(2), rv=rv, **kw)
    t = make_tables(Model(**kw))
    o.init_conf()
    c0 = o.confv
    assert set(np.unique(c0)) <= {1, 2} and len(np.unique(c0)) == 2
    o.sweep_start()
    n = o.ndim
    for p in ("b0", "0b"):
        getattr(o, "sweep_" + p)(lupdate=True, lmeasure=False)
        G = np.linalg.inv(np.eye(n) + _b_beta(t, o.confv, rv, 0.1, o.conf))
        assert np.abs(o.G - G).max() < 1e-9, p
    acc = o.vacc()
    assert 0 < acc[0] < acc[1] and acc[1] == 2 * o.ltrot * o.nn_nf * o.nn_lf
    assert not np.array_equal(o.confv[:, :, :o.nn_lf], c0[:, :, :o.nn_lf])
    assert o.diag()["max_wrap_error"] < 1e-10


def test_exact_enumeration_chain4():
    kw = dict(lattice="chain", la=4, lb=1, beta=0.2, dtau=0.1, rhub=0.0, nflr=1, mu=0.3, kmode="checkerboard",
              nwrap=1)
    rv = 2.0
    t = make_tables(Model(**kw))
    lat, n = t["lat"], t["ndim"]
    nb = lat["nn_nf"] * lat["nn_lf"]
    num, den = np.zeros(3), 0.0
    for bits in itertools.product((1, 2), repeat=t["ltrot"] * nb):
        confv = np.ones((t["ltrot"], lat["nn_nf"], n), np.int32)
        confv[:, :, :lat["nn_lf"]] = np.array(bits).reshape(t["ltrot"], lat["nn_nf"], lat["nn_lf"])
        M = np.eye(n) + _b_beta(t, confv, rv, 0.1)
        w = np.linalg.det(M).real
        G = np.linalg.inv(M)
        gc = np.eye(n) - G.T
        ek = sum(-2.0 * gc[i0, i1].real for nf in range(lat["nn_nf"]) for i0, i1 in lat["bonds"][nf])
        num += w * np.array([np.trace(gc).real / n, ek, np.sum(np.diag(gc) ** 2).real])
        den += w
    exact = num / den
    o = oracle.OracleChain(seed=oracle.chain_seed(5), rv=rv, **kw)
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
        s = ob["sign"].real
        vals.append([ob["density"].real / s, ob["ekin"].real / s, ob["dblocc"].real / s])
    vals = np.array(vals)
    mean = vals.mean(0)
    err = vals.std(0, ddof=1) / np.sqrt(len(vals))
    assert np.all(np.abs(mean - exact) < 4 * err + 1e-12), (mean, exact, err)
    assert np.all(err < 0.05 * np.abs(exact) + 1e-3)
