"""Pins of the projective (t-V, submatrix-LR) part of the CPU oracle.

Each test checks the oracle against a fact that does not come from the oracle itself: the
Hubbard-Stratonovich identity of the bond field, LAPACK's QR and numpy's eigensolver, the
non-interacting closed form, a direct dense-matrix construction of G from the fields, and an
exact enumeration of every field configuration of a 4-site system against the Monte Carlo
average.  References: src/dqmc_config_v/dqmc_set_conf.f90, src/dqmc_config_v/*_prop.f90,
src/dqmc_config_v/dqmc_proj_update_submatrix_LR.f90, src/dqmc_proj_core/*.f90,
src/dqmc_config_h0/set_trial_h0.f90, lib/s_matrix.f90:1822-1996.
"""
import itertools

import numpy as np
import pytest
import scipy.linalg

import oracle

SX = np.array([[0.0, 1.0], [1.0, 0.0]])


def bonds_by_family(o):
    b = o.table(6).astype(int).reshape(o.nn_nf, o.nn_lf, 2)
    return b


def test_bond_hs_identity():
    """lambda sum_eta exp(alpha eta K_b) = exp(-dtau V (n_i - 1/2)(n_j - 1/2)) on the 2-site
    Fock space, and the bond matrix u diag(bp, bm) ut = expm(alpha eta sigma_x)."""
    dtau, v = 0.1, 1.2
    o = oracle.OracleProjChain(lattice="honeycomb", la=1, lb=1, beta=0.2, dtau=dtau, rv=v, nwrap=1, nustock=2)
    t = o.table(100).reshape(2, 4)       # per field value: bp, bm, bpi, bmi
    d = o.table(101).reshape(2, 2)       # per field value: dp, dm
    alpha = np.arccosh(np.exp(dtau * v / 2))
    eta = np.array([-1.0, 1.0])
    h = np.array([[1.0, 1.0], [1.0, -1.0]]) / np.sqrt(2.0)
    for k in range(2):
        bp, bm, bpi, bmi = t[k]
        np.testing.assert_allclose([bp, bm, bpi, bmi],
                                   np.exp([alpha * eta[k], -alpha * eta[k], -alpha * eta[k], alpha * eta[k]]),
                                   rtol=1e-15)
        blk = h @ np.diag([bp.real, bm.real]) @ h
        np.testing.assert_allclose(blk, scipy.linalg.expm(alpha * eta[k] * SX), rtol=1e-14, atol=1e-15)
        kp = 1 - k
        np.testing.assert_allclose(d[k], [t[kp][0] / t[k][0] - 1, t[kp][1] / t[k][1] - 1], rtol=1e-13)
    # Fock space of two spinless sites: basis |n_i n_j>, hopping K_b = c_i^+ c_j + h.c.
    ci = np.array([[0, 1], [0, 0]], float)
    z = np.diag([1.0, -1.0])
    c1 = np.kron(ci, np.eye(2))
    c2 = np.kron(z, ci)          # Jordan-Wigner sign
    n1, n2 = c1.T @ c1, c2.T @ c2
    kb = c1.T @ c2 + c2.T @ c1
    lam = 0.5 * np.exp(-dtau * v / 4)
    lhs = lam * sum(scipy.linalg.expm(alpha * e * kb) for e in eta)
    rhs = scipy.linalg.expm(-dtau * v * (n1 - 0.5 * np.eye(4)) @ (n2 - 0.5 * np.eye(4)))
    np.testing.assert_allclose(lhs, rhs, atol=1e-14)


@pytest.mark.parametrize("lattice,la", [("honeycomb", 2), ("square", 4), ("honeycomb", 3)])
def test_trial_wavefunction(lattice, la):
    """set_trial_h0 draws one uniform per (A site, family) after the field's N nn_nf ltrot
    draws; the projector on the lowest ne eigenvectors matches numpy's eigh."""
    seed = oracle.chain_seed(3)
    o = oracle.OracleProjChain(lattice=lattice, la=la, lb=la, beta=1.0, rndness=0.3, mu=0.05, seed=seed)
    o.init_conf()
    gap = o.set_trial()
    p, w = o.proj
    n_conf = o.ndim * o.nn_nf * o.ltrot
    u = oracle.sfmt_stream(seed, n_conf + o.nn_lf * o.nn_nf)[n_conf:]
    b = bonds_by_family(o)
    H = np.zeros((o.ndim, o.ndim))
    k = 0
    for i in range(o.nn_lf):
        for f in range(o.nn_nf):
            i0, i1 = b[f, i]
            z1 = -1.0 + 0.3 * (u[k] - 1.0)
            k += 1
            H[i0, i1] += z1
            H[i1, i0] += z1
    H -= 0.05 * np.eye(o.ndim)
    ev, vec = np.linalg.eigh(H)
    np.testing.assert_allclose(w, ev, atol=1e-12)
    assert gap == pytest.approx(ev[o.ne] - ev[o.ne - 1], abs=1e-12)
    P = vec[:, :o.ne]
    np.testing.assert_allclose(p @ p.conj().T, P @ P.T, atol=1e-11)
    np.testing.assert_allclose(p.conj().T @ p, np.eye(o.ne), atol=1e-12)


@pytest.mark.parametrize("shape", [(24, 10), (10, 24), (16, 16)])
def test_qr_pqmc(shape):
    """s_zgeQR_pqmc: zgeqrf Householder Q (LAPACK via scipy) with the first column negated for
    an odd number of negative R diagonal entries; log|det R|; LQ of wide matrices by plain
    transpose."""
    rng = np.random.default_rng(7)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    q, rl = oracle.qr_pqmc(a)
    wide = shape[0] < shape[1]
    at = a.T if wide else a
    qs, rs = scipy.linalg.qr(at, mode="economic")
    d = np.real(np.diag(rs))
    if np.sum(d < 0) % 2 == 1:
        qs[:, 0] *= -1
    expect = qs.T if wide else qs
    np.testing.assert_allclose(q, expect, atol=1e-12)
    assert rl == pytest.approx(np.sum(np.log(np.abs(d))), rel=1e-12)
    # the discarded R has a positive determinant: A = Q R (tall) or A = R^T Q (wide)
    r = a @ q.conj().T if wide else q.conj().T @ a
    np.testing.assert_allclose(np.imag(np.diag(r)), 0, atol=1e-12)
    assert np.prod(np.sign(np.real(np.diag(r)))) > 0


def test_free_fermion_closed_form():
    """V = 0 and a trial state of eigenvectors of K: G = 1 - P P^H at every time slice."""
    o = oracle.OracleProjChain(lattice="honeycomb", la=2, lb=2, beta=2.0, rv=0.0, rndness=0.0, kmode="dense",
                               nwrap=5, nustock=8)
    o.init_conf()
    o.set_trial()
    p, w = o.proj
    assert w[o.ne] - w[o.ne - 1] > 0.1
    o.sweep_start()
    o.set_trace(64)
    o.sweep()
    expect = np.eye(o.ndim) - p @ p.conj().T
    tr = o.trace()
    assert len(tr) == 2 * o.nst
    for g in tr:
        np.testing.assert_allclose(g, expect, atol=1e-10)
    np.testing.assert_allclose(o.G, expect, atol=1e-10)


def slice_matrix(o, conf_t, dtau, rv, kin):
    """B(tau) of one slice built densely: K-half (families 1..nf), bond fields (families
    1..nf), K-half (families nf..1), per src/dqmc_proj_core.f90:68-116."""
    N = o.ndim
    alpha = np.arccosh(np.exp(dtau * rv / 2))
    b = bonds_by_family(o)
    M = np.eye(N, dtype=complex)
    for f in kin[0]:
        M = f @ M
    for f in range(o.nn_nf):
        V = np.eye(N, dtype=complex)
        for q in range(o.nn_lf):
            i1, i2 = b[f, q]
            eta = -1.0 if conf_t[f, q] == 1 else 1.0
            blk = scipy.linalg.expm(alpha * eta * SX)
            V[np.ix_([i1, i2], [i1, i2])] = blk
        M = V @ M
    for f in kin[1]:
        M = f @ M
    return M


def kinetic_factors(o, kmode, dtau):
    """exp(-dtau K / 2) as per-family bond products (checkerboard) or one dense factor."""
    N = o.ndim
    b = bonds_by_family(o)
    if kmode == "dense":
        K = np.zeros((N, N))
        for f in range(o.nn_nf):
            for q in range(o.nn_lf):
                i1, i2 = b[f, q]
                K[i1, i2] -= 1.0
                K[i2, i1] -= 1.0
        e = scipy.linalg.expm(-0.5 * dtau * K)
        return [e], [e]
    fams = []
    for f in range(o.nn_nf):
        T = np.eye(N, dtype=complex)
        for q in range(o.nn_lf):
            i1, i2 = b[f, q]
            T[np.ix_([i1, i2], [i1, i2])] = scipy.linalg.expm(0.5 * dtau * SX)
        fams.append(T)
    return fams, fams[::-1]


@pytest.mark.parametrize("kmode,lattice,la", [("dense", "honeycomb", 2), ("checkerboard", "honeycomb", 2),
                                              ("checkerboard", "square", 4), ("checkerboard", "chain", 10),
                                              ("dense", "chain", 12)])
def test_direct_product(kmode, lattice, la):
    """After a 0b pass, G = 1 - R (P^H R)^-1 P^H with R = B(theta, 0) P built densely from
    the final fields."""
    dtau, rv = 0.1, 1.2
    o = oracle.OracleProjChain(lattice=lattice, la=la, lb=la, beta=1.0, dtau=dtau, rv=rv, rndness=0.2, kmode=kmode,
                               nwrap=5, nustock=6, seed=oracle.chain_seed(5))
    o.init_conf()
    o.set_trial()
    o.sweep_start()
    o.sweep()
    conf = o.conf[:, :, :o.nn_lf]
    p, _ = o.proj
    kin = kinetic_factors(o, kmode, dtau)
    B = np.eye(o.ndim, dtype=complex)
    for t in range(o.ltrot):
        B = slice_matrix(o, conf[t], dtau, rv, kin) @ B
    R = B @ p
    G = np.eye(o.ndim) - R @ np.linalg.solve(p.conj().T @ R, p.conj().T)
    np.testing.assert_allclose(o.G, G, atol=1e-10)
    assert o.diag()["max_wrap_error"] < 1e-10


def test_stock_size_is_invisible():
    """Submatrix-LR with one bond per stock (Gamma^-1 is one 2x2 block) and with 4 bonds per
    stock (Gamma^-1 grown by bordering) accept the same bonds and give the same G."""
    kw = dict(lattice="honeycomb", la=3, lb=3, beta=1.0, rv=1.2, rndness=0.2, nwrap=5, seed=oracle.chain_seed(2))
    a = oracle.OracleProjChain(nustock=2, **kw)
    b = oracle.OracleProjChain(nustock=8, **kw)
    for o in (a, b):
        o.init_conf()
        o.set_trial()
        o.sweep_start()
    for _ in range(2):
        a.sweep()
        b.sweep()
        assert np.array_equal(a.conf, b.conf)
        np.testing.assert_allclose(a.G, b.G, atol=1e-11)
    da, db = a.diag(), b.diag()
    assert da["accepted"] == db["accepted"] and 0 < da["accepted"] < da["tried"]


def test_exact_enumeration_square_2x2():
    """4 sites, 2 particles, 2 slices, 8 bond fields per slice: all 2^16 field configurations
    enumerated; the Monte Carlo estimates of the kinetic energy and sum_i n_i^2 at the middle
    slice agree with the exact weighted averages within 4 standard errors."""
    dtau, rv = 0.1, 2.0
    kw = dict(lattice="square", la=2, lb=2, beta=0.2, dtau=dtau, rv=rv, rndness=0.4, kmode="dense", nwrap=1,
              nustock=4, mid_len=1)
    o = oracle.OracleProjChain(seed=oracle.chain_seed(1), **kw)
    o.init_conf()
    o.set_trial()
    p, _ = o.proj
    assert o.ltrot == 2 and o.ndim == 4 and o.ne == 2
    kin = kinetic_factors(o, "dense", dtau)
    b = bonds_by_family(o)
    nb = o.nn_nf * o.nn_lf
    # per-slice matrices for all 2^8 field settings of one slice
    sl = []
    for bits in itertools.product([1, 2], repeat=nb):
        ct = np.array(bits).reshape(o.nn_nf, o.nn_lf)
        sl.append(slice_matrix(o, ct, dtau, rv, kin))
    sl = np.array(sl)
    B1 = sl[:, None]           # slice 1 varies along axis 0
    B2 = sl[None, :]           # slice 2 along axis 1
    R = np.einsum("abij,jk->abik", np.broadcast_to(B1, (len(sl), len(sl), 4, 4)), p)
    L = np.einsum("ji,abjk->abik", p.conj(), np.broadcast_to(B2, (len(sl), len(sl), 4, 4)))
    LR = L @ R
    w = np.linalg.det(LR).real
    G = np.eye(4) - R @ np.linalg.solve(LR, L)
    ekin = np.zeros(w.shape)
    for f in range(o.nn_nf):
        for q in range(o.nn_lf):
            i0, i1 = b[f, q]
            ekin += -2.0 * (-(G[..., i1, i0].real))
    dbl = np.sum((1.0 - np.real(np.diagonal(G, axis1=-2, axis2=-1))) ** 2, axis=-1)
    exact_ekin = np.sum(w * ekin) / np.sum(w)
    exact_dbl = np.sum(w * dbl) / np.sum(w)
    # Monte Carlo: per-sweep ratios of the weighted sums, binned
    o.sweep_start()
    for _ in range(200):
        o.sweep(lmeasure=False)
    nsw, nbin = 24000, 40
    ek, db, sg = np.zeros(nsw), np.zeros(nsw), np.zeros(nsw)
    for s in range(nsw):
        o.reset_obs()
        o.sweep()
        ob = o.obs()
        ek[s], db[s], sg[s] = ob["ekin"].real, ob["dblocc"].real, ob["sign"].real
    for vals, exact in ((ek, exact_ekin), (db, exact_dbl)):
        vb = vals.reshape(nbin, -1).sum(1)
        sb = sg.reshape(nbin, -1).sum(1)
        est = vb.sum() / sb.sum()
        jk = np.array([(vb.sum() - vb[i]) / (sb.sum() - sb[i]) for i in range(nbin)])
        err = np.sqrt((nbin - 1) * np.mean((jk - jk.mean()) ** 2))
        assert abs(est - exact) < 4 * err + 1e-12, (est, exact, err)


# ---- the on-site (Hubbard) field in the projective sweep: src/dqmc_config_u/*_prop.f90,
# ---- src/dqmc_config_u/dqmc_proj_update_submatrix_LR.f90, src/dqmc_proj_core.f90:68-116

def _u_tables(dtau, rhub, nflr=2):
    from sunhubbardmy_amd.model import Model, hs_tables
    return hs_tables(Model(dtau=dtau, rhub=rhub, nflr=nflr, exact_hs=True))


def slice_matrix_tuv(o, conf_v_t, conf_u_t, dtau, rv, hs, kin):
    """B(tau) with the bond field (when rv != 0) and the on-site field: K-half, bond fields,
    exp(V_U) = diag(bmat(conf_u)), K-half (src/dqmc_proj_core.f90:68-116)."""
    N = o.ndim
    M = np.eye(N, dtype=complex)
    for f in kin[0]:
        M = f @ M
    if rv != 0.0:
        alpha = np.arccosh(np.exp(dtau * rv / 2))
        b = bonds_by_family(o)
        for f in range(o.nn_nf):
            V = np.eye(N, dtype=complex)
            for q in range(o.nn_lf):
                i1, i2 = b[f, q]
                eta = -1.0 if conf_v_t[f, q] == 1 else 1.0
                V[np.ix_([i1, i2], [i1, i2])] = scipy.linalg.expm(alpha * eta * SX)
            M = V @ M
    M = np.diag(hs["bmat"][conf_u_t - 1]) @ M
    for f in kin[1]:
        M = f @ M
    return M


@pytest.mark.parametrize("kmode,rv", [("dense", 0.0), ("checkerboard", 0.0), ("checkerboard", 1.0)])
def test_hubbard_direct_product(kmode, rv):
    """Hubbard (and t-U-V) projective chain: after a sweep, G = 1 - R (P^H R)^-1 P^H with
    R = B(theta, 0) P built densely from the final on-site and bond fields."""
    dtau, rhub = 0.1, 4.0
    o = oracle.OracleProjChain(lattice="square", la=4, lb=4, beta=1.0, dtau=dtau, rv=rv, rhub=rhub, nflr=2,
                               rndness=0.2, kmode=kmode, nwrap=5, nustock=6, seed=oracle.chain_seed(4))
    o.init_conf()
    o.set_trial()
    o.sweep_start()
    u0 = o.uconf.copy()
    assert set(np.unique(u0)) <= {1, 2} and len(np.unique(u0)) == 2
    o.sweep()
    assert not np.array_equal(o.uconf, u0)                # the on-site field was updated
    hs = _u_tables(dtau, rhub)
    p, _ = o.proj
    kin = kinetic_factors(o, kmode, dtau)
    B = np.eye(o.ndim, dtype=complex)
    conf_v = o.conf[:, :, :o.nn_lf]
    for t in range(o.ltrot):
        B = slice_matrix_tuv(o, conf_v[t], o.uconf[t], dtau, rv, hs, kin) @ B
    R = B @ p
    G = np.eye(o.ndim) - R @ np.linalg.solve(p.conj().T @ R, p.conj().T)
    np.testing.assert_allclose(o.G, G, atol=1e-10)
    d = o.diag()
    assert d["max_wrap_error"] < 1e-10 and 0 < d["accepted"] < d["tried"]


def test_hubbard_stock_size_is_invisible():
    """The on-site submatrix-LR update with 2 and with 8 sites per stock accepts the same flips."""
    kw = dict(lattice="square", la=4, lb=4, beta=1.0, rv=0.0, rhub=4.0, nflr=2, rndness=0.2, nwrap=5,
              kmode="checkerboard", seed=oracle.chain_seed(6))
    a = oracle.OracleProjChain(nustock=2, **kw)
    b = oracle.OracleProjChain(nustock=8, **kw)
    for o in (a, b):
        o.init_conf()
        o.set_trial()
        o.sweep_start()
    for _ in range(2):
        a.sweep()
        b.sweep()
        assert np.array_equal(a.uconf, b.uconf)
        np.testing.assert_allclose(a.G, b.G, atol=1e-11)
    assert a.diag()["accepted"] == b.diag()["accepted"]


def test_hubbard_exact_enumeration_square_2x2():
    """4 sites, 2 particles per flavour, SU(2), 2 slices: all 2^8 on-site fields enumerated
    with weight Re[det(P^H B P)^2 prod phase(conf)]; the Monte Carlo kinetic energy and
    sum_i (1 - G_ii)^2 at the middle slice agree with the exact averages within 4 errors."""
    dtau, rhub = 0.25, 4.0
    kw = dict(lattice="square", la=2, lb=2, beta=0.5, dtau=dtau, rv=0.0, rhub=rhub, nflr=2, rndness=0.4,
              kmode="dense", nwrap=1, nustock=4, mid_len=1)
    o = oracle.OracleProjChain(seed=oracle.chain_seed(3), **kw)
    o.init_conf()
    o.set_trial()
    p, _ = o.proj
    assert o.ltrot == 2 and o.ndim == 4 and o.ne == 2
    hs = _u_tables(dtau, rhub)
    kin = kinetic_factors(o, "dense", dtau)
    b = bonds_by_family(o)
    sl, ph = [], []
    for bits in itertools.product([1, 2], repeat=o.ndim):
        ct = np.array(bits)
        sl.append(slice_matrix_tuv(o, None, ct, dtau, 0.0, hs, kin))
        ph.append(np.prod(hs["phase"][ct - 1]))
    sl, ph = np.array(sl), np.array(ph)
    n1 = len(sl)
    B1 = np.broadcast_to(sl[:, None], (n1, n1, 4, 4))
    B2 = np.broadcast_to(sl[None, :], (n1, n1, 4, 4))
    R = B1 @ p
    L = p.conj().T @ B2
    LR = L @ R
    # complex weights; the chain samples |Re W| and measures O W / |Re W| (the phase-weighted
    # bins of equaltime_measure), so the exact value of a bin ratio is sum Re(W O) / sum Re(W)
    wc = np.linalg.det(LR) ** 2 * ph[:, None] * ph[None, :]
    assert wc.real.min() > 0
    G = np.eye(4) - R @ np.linalg.solve(LR, L)
    ekin = np.zeros(wc.shape)
    for f in range(o.nn_nf):
        for q in range(o.nn_lf):
            i0, i1 = b[f, q]
            ekin += 2.0 * G[..., i1, i0].real
    dbl = np.sum((1.0 - np.diagonal(G, axis1=-2, axis2=-1)) ** 2, axis=-1)
    exact_ekin = np.sum((wc * ekin).real) / np.sum(wc.real)
    exact_dbl = np.sum((wc * dbl).real) / np.sum(wc.real)
    o.sweep_start()
    for _ in range(200):
        o.sweep(lmeasure=False)
    nsw, nbin = 16000, 40
    ek, db, sg = np.zeros(nsw), np.zeros(nsw), np.zeros(nsw)
    for s in range(nsw):
        o.reset_obs()
        o.sweep()
        ob = o.obs()
        ek[s], db[s], sg[s] = ob["ekin"].real, ob["dblocc"].real, ob["sign"].real
    for vals, exact in ((ek, exact_ekin), (db, exact_dbl)):
        vb = vals.reshape(nbin, -1).sum(1)
        sb = sg.reshape(nbin, -1).sum(1)
        est = vb.sum() / sb.sum()
        jk = np.array([(vb.sum() - vb[i]) / (sb.sum() - sb[i]) for i in range(nbin)])
        err = np.sqrt((nbin - 1) * np.mean((jk - jk.mean()) ** 2))
        assert abs(est - exact) < 4 * err + 1e-12, (est, exact, err)
