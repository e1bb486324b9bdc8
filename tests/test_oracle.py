"""Pins of the CPU oracle (oracle/dqmc_oracle.c) against facts it does not itself compute.

The reference ships no tests or golden outputs (its .gitignore drops *.out/*.bin) and cannot be
built here without Intel MKL and MPI, so each oracle function is pinned by an independent fact:

  SFMT19937            published known-answer vector of init_gen_rand(1234) (SFMT.19937.out.txt)
  chain seeds          the int32 product of src/dqmc_util.f90:70-72 evaluated by hand
  HS tables            the HS identities they exist to satisfy (exact for SU(2)/SU(4), O(dtau^4) general)
  kinetic tables       scipy.linalg.expm; FFT operator of the reference = dense exp(-dtau K/2)
  QR with pivoting     LAPACK zgeqp3 via scipy.linalg.qr(pivoting=True): same pivots, same R
  LU inverse/log det   numpy.linalg.inv / slogdet
  propagation + UDV    U = 0 closed form G = (1 + B(beta,0))^-1 with B built by numpy
  whole Markov chain   exact enumeration of every HS field of a 4-site, 4-slice system
  update algorithms    fast, delay and submatrix give the same HS-field trajectory
"""
import numpy as np
import pytest
import scipy.linalg

import oracle
from sunhubbardmy_amd.model import Model, make_tables

# first outputs of SFMT19937 after init_gen_rand(1234), from the SFMT reference distribution
SFMT_KAT_1234 = [3440181298, 1564997079, 1510669302, 2930277156, 1452439940, 3796268453, 423124208, 2143818589]


def test_sfmt_known_answer():
    assert list(oracle.sfmt_u32(1234, 8)) == SFMT_KAT_1234


def test_sfmt_stream_conversion():
    u = oracle.sfmt_u32(4357, 3000)
    d = oracle.sfmt_stream(4357, 3000)
    assert np.array_equal(d, (u.astype(np.float64) + 0.5) / 4294967296.0)
    assert d.min() > 0.0 and d.max() < 1.0
    # refills past the first 624-word block keep the stream going (no repetition of block 1)
    assert not np.array_equal(u[:624], u[624:1248])


def test_chain_seeds():
    # (irank*1981 + 2008)*951049 wraps in int32 for irank >= 1
    assert oracle.chain_seed(0) == 2008 * 951049 == 1909706392
    v = 3989 * 951049 - 2 ** 32
    assert oracle.chain_seed(1) == abs(v) == 501232835


def _hs(nflr, exact, u=6.0, dtau=0.1):
    return make_tables(Model(nflr=nflr, exact_hs=exact, rhub=u, dtau=dtau))["hs"]


def test_hs_su2_identity():
    # sum_eta gamma exp(i alpha eta x) / 2 = exp(-dtau U x^2 / 2) for x = n - 1 in {-1, 0, 1}
    for u in (2.0, 4.0, 8.0):
        hs = _hs(2, True, u)
        o = oracle.OracleChain(nflr=2, exact_hs=True, rhub=u)
        for x in (-1, 0, 1):
            s = np.sum(hs["gaml"] * np.exp(1j * hs["alpha"] * hs["etal"] * x)) / 2
            assert abs(s - np.exp(-0.1 * u * x * x / 2)) < 1e-14
        t0 = o.table(0).reshape(-1, 3)
        assert np.allclose(t0[:, 0], np.exp(1j * hs["alpha"] * hs["etal"]), atol=1e-15)


def test_hs_su4_exact_identity():
    # exact SU(4) HS: sum_l gamma_l exp(i eta_l x) is proportional to exp(-dtau U x^2 / 2), x = n - 2
    u, dtau = 6.0, 0.1
    hs = _hs(4, True, u, dtau)
    vals = [np.sum(hs["gaml"] * np.exp(1j * hs["alpha"] * hs["etal"] * x)) for x in range(-2, 3)]
    ratio = [v / vals[2] for v in vals]
    for x, r in zip(range(-2, 3), ratio):
        assert abs(r - np.exp(-dtau * u * x * x / 2)) < 1e-12


def test_hs_general_four_component():
    # general 4-component HS: sum gamma exp(i alpha eta x) = 4 exp(-alpha^2 x^2) + O(alpha^8 x^8)
    for dtau in (0.05, 0.1):
        hs = _hs(4, False, 6.0, dtau)
        a = hs["alpha"]
        for x in (-2, -1, 0, 1, 2):
            s = np.sum(hs["gaml"] * np.exp(1j * a * hs["etal"] * x))
            assert abs(s - 4 * np.exp(-a * a * x * x)) < 5 * (a * x) ** 8 + 1e-14


def test_hs_ratio_tables_consistent():
    for nflr, ex in ((2, True), (4, True), (4, False), (6, True)):
        o = oracle.OracleChain(nflr=nflr, exact_hs=ex, rhub=6.0)
        L = o.lcomp
        t0 = o.table(0).reshape(-1, 3)
        bm, ph = t0[:, 0], t0[:, 2]
        dl = o.table(1).reshape(L, L - 1).T
        pr = o.table(2).reshape(L, L - 1).T
        for is_ in range(L):
            for fl in range(1, L):
                isp = (is_ + fl) % L
                assert abs(dl[fl - 1, is_] - (bm[isp] / bm[is_] - 1)) < 1e-14
                assert abs(pr[fl - 1, is_] - ph[isp] / ph[is_]) < 1e-14


def test_kinetic_checkerboard_bond_exponentials():
    o = oracle.OracleChain(kmode="checkerboard", mu=0.3, dtau=0.1)
    urt = o.table(4).reshape(-1, 2, 2).transpose(0, 2, 1)
    h = np.array([[-0.3 / 4, -1.0], [-1.0, -0.3 / 4]])
    ref = scipy.linalg.expm(-0.05 * h)
    assert np.abs(urt - ref).max() < 1e-15
    urtm1 = o.table(5).reshape(-1, 2, 2).transpose(0, 2, 1)
    assert np.abs(urt[0] @ urtm1[0] - np.eye(2)).max() < 1e-15


@pytest.mark.parametrize("lattice,la", [("square", 4), ("square", 6), ("honeycomb", 3)])
def test_kinetic_dense_and_fft(lattice, la):
    m = Model(lattice=lattice, la=la, lb=la, kmode="dense", mu=0.2)
    t = make_tables(m)
    o = oracle.OracleChain(lattice=lattice, la=la, lb=la, kmode="dense", mu=0.2)
    from sunhubbardmy_amd.model import h0_matrix
    H = h0_matrix(m, t["lat"])
    ref = scipy.linalg.expm(-0.05 * H)
    n = t["ndim"]
    urt = o.table(4).reshape(n, n).T
    assert np.abs(urt - ref).max() < 1e-13
    assert np.abs(t["urt"].reshape(n, n, order="F") - ref).max() < 1e-13
    if lattice == "square":
        # the FFT build applies F^-1 diag(exp(-dtau eps_k/2)) F (dqmc_config_h0/dqmc_left_forward_prop.f90:95-134)
        L = la
        eps = np.zeros((L, L))
        for kx in range(L):
            for ky in range(L):
                eps[kx, ky] = -2.0 * (np.cos(2 * np.pi * kx / L) + np.cos(2 * np.pi * ky / L)) - 0.2
        F = np.kron(np.fft.fft(np.eye(L)) / np.sqrt(L), np.fft.fft(np.eye(L)) / np.sqrt(L))
        op = F.conj().T @ np.diag(np.exp(-0.05 * eps.reshape(-1))) @ F
        assert np.abs(op - ref).max() < 1e-13


@pytest.mark.parametrize("n", [5, 16, 33])
def test_qr_pivoted_matches_lapack(n):
    rng = np.random.default_rng(n)
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) @ np.diag(np.exp(rng.uniform(-6, 6, n)))
    q, r, jp, ld = oracle.qr(a, pivot=1)
    Q, R, P = scipy.linalg.qr(a, pivoting=True)
    assert np.array_equal(jp - 1, P)
    assert np.abs(r - R).max() / np.abs(R).max() < 1e-12
    assert np.abs(q - Q).max() < 1e-10
    assert np.allclose(np.imag(np.diag(r)), 0.0)
    assert np.abs(q @ r - a[:, jp - 1]).max() / np.abs(a).max() < 1e-13
    d = np.abs(np.diag(r))
    assert np.all(d[:-1] >= d[1:] * (1 - 1e-12))
    assert abs(np.exp(ld) - np.linalg.det(q)) < 1e-10


def test_qr_unpivoted_matches_lapack():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((20, 20)) + 1j * rng.standard_normal((20, 20))
    q, r, jp, ld = oracle.qr(a, pivot=0)
    Q, R = scipy.linalg.qr(a)
    assert np.array_equal(jp, np.arange(1, 21))
    assert np.abs(r - R).max() < 1e-12 and np.abs(q - Q).max() < 1e-12
    assert abs(np.exp(ld) - np.linalg.det(q)) < 1e-12


def test_inv_logdet():
    rng = np.random.default_rng(9)
    for n in (1, 7, 30):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        inv, ld, info = oracle.inv_logdet(a)
        assert info == 0
        assert np.abs(inv - np.linalg.inv(a)).max() < 1e-11
        s, la = np.linalg.slogdet(a)
        assert abs(ld.real - la) < 1e-11 and abs(np.exp(1j * ld.imag) - s) < 1e-11
        assert -np.pi <= ld.imag <= np.pi
    inv, ld, info = oracle.inv_logdet(np.zeros((3, 3)))
    assert info == 1


def _b_product_checkerboard(t, conf=None, bm=None):
    """B(beta, 0) from numpy: slices A2 exp(V) A1 with the bond matrices of the tables."""
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
    B = np.eye(n, dtype=complex)
    for nt in range(t["ltrot"]):
        V = np.eye(n) if conf is None else np.diag(bm[conf[nt] - 1])
        B = A2 @ V @ A1 @ B
    return B


@pytest.mark.parametrize("kmode", ["checkerboard", "dense"])
def test_u0_closed_form(kmode):
    kw = dict(lattice="square", la=4, lb=4, beta=3.0, rhub=0.0, kmode=kmode, nwrap=5)
    o = oracle.OracleChain(**kw)
    o.sweep_start()
    t = make_tables(Model(**kw))
    if kmode == "dense":
        from sunhubbardmy_amd.model import h0_matrix
        H = h0_matrix(t["model"], t["lat"])
        ref = np.linalg.inv(np.eye(16) + scipy.linalg.expm(-3.0 * H))
    else:
        ref = np.linalg.inv(np.eye(16) + _b_product_checkerboard(t))
    # the direct float64 inverse of 1 + B(beta,0) at beta = 3 is itself good to ~1e-11
    assert np.abs(o.G - ref).max() < 1e-10
    o.sweep(lupdate=False, lmeasure=True)
    assert o.diag()["max_wrap_error"] < 1e-12
    assert np.abs(o.G - ref).max() < 1e-10
    ob = o.obs()
    # half filling, U = 0: density exactly 1/2 per flavour, no double-occupancy correlation
    assert abs(ob["density"] / ob["nobs"] - 0.5) < 1e-12


def _exact_enumeration(t):
    """<O> = sum_s w_s O_s / sum_s w_s over every HS field of a tiny system."""
    n, lt = t["ndim"], t["ltrot"]
    hs = t["hs"]
    nflr = t["model"].nflr
    num = np.zeros(3)
    den = 0.0
    lat = t["lat"]
    for code in range(2 ** (n * lt)):
        bits = [(code >> k) & 1 for k in range(n * lt)]
        conf = np.array(bits, np.int32).reshape(lt, n) + 1
        B = _b_product_checkerboard(t, conf, hs["bmat"])
        M = np.eye(n) + B
        G = np.linalg.inv(M)
        w = np.linalg.det(M) ** nflr * np.prod(hs["phase"][conf - 1])
        gc = np.eye(n) - G.T
        dbl = np.sum(np.diag(gc) ** 2)
        ek = 0.0
        for nf in range(lat["nn_nf"]):
            for b in range(lat["nn_lf"]):
                i0, i1 = lat["bonds"][nf, b]
                ek += -2.0 * np.real(gc[i0, i1])
        stag = lat["stag"]
        ss = np.sum(np.outer(stag, stag) * gc * G) * (1 - 1 / nflr ** 2) / t["lq"]
        num += np.real(w) * np.array([dbl.real, ek, ss.real])
        den += np.real(w)
    return num / den


def test_exact_enumeration_small_system():
    # 2x2 square lattice (N=4), 4 slices: 2^16 HS fields enumerated exactly
    kw = dict(lattice="square", la=2, lb=2, beta=0.8, dtau=0.2, rhub=4.0, kmode="checkerboard", nwrap=2,
              update="delay", nublock=2)
    t = make_tables(Model(**kw))
    exact = _exact_enumeration(t)
    o = oracle.OracleChain(seed=oracle.chain_seed(3), **kw)
    o.init_conf()
    o.sweep_start()
    for _ in range(200):
        o.sweep(lupdate=True, lmeasure=False)
    vals = []
    for _ in range(60):
        o.reset_obs()
        for _ in range(50):
            o.sweep(lupdate=True, lmeasure=True)
        ob = o.obs()
        vals.append([ob["dblocc"].real / ob["nobs"], ob["ekin"].real / ob["nobs"], ob["s_stag"].real / ob["nobs"]])
    vals = np.array(vals)
    mean = vals.mean(0)
    err = vals.std(0, ddof=1) / np.sqrt(len(vals))
    assert np.all(np.abs(mean - exact) < 4 * err + 1e-12), (mean, exact, err)


def test_update_algorithms_agree():
    for kmode in ("checkerboard", "dense"):
        out = []
        for upd in ("fast", "delay", "submatrix"):
            o = oracle.OracleChain(la=4, lb=4, beta=2.0, rhub=4.0, update=upd, kmode=kmode, nwrap=5, nublock=3,
                                   seed=oracle.chain_seed(1))
            o.init_conf()
            o.sweep_start()
            for _ in range(2):
                o.sweep()
            out.append((o.conf, o.diag()))
        for conf, d in out[1:]:
            assert np.array_equal(conf, out[0][0])
            assert d["accepted"] == out[0][1]["accepted"]
        for _, d in out:
            assert d["max_wrap_error"] < 1e-12 and d["max_phasediff"] < 1e-12


def test_reference_probe_values_4x4():
    # values measured with a build of the reference at this shape (4x4 square, U = 4, beta = 4, fast
    # update): acceptance 0.736, E_kin/N per flavour -0.656, <n_up n_dn> 0.1284 over 1200 sweeps;
    # loose statistical match
    o = oracle.OracleChain(la=4, lb=4, beta=4.0, rhub=4.0, update="fast", kmode="checkerboard", nwrap=10,
                           seed=oracle.chain_seed(0))
    o.init_conf()
    o.sweep_start()
    for _ in range(300):
        o.sweep(lmeasure=True)
    d, ob = o.diag(), o.obs()
    assert abs(d["accepted"] / d["tried"] - 0.736) < 0.02
    assert abs(ob["ekin"].real / ob["nobs"] / 16 - (-0.656)) < 0.03
    assert abs(ob["dblocc"].real / ob["nobs"] / 16 - 0.1284) < 0.02
