"""Kernel-level GPU tests: every kernel group of libsunhip.so against a float64 numpy reference
(and the oracle's QR/LU where the pivot order is part of the contract).  Uses the test hooks
of include/sunhip_testing.h."""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from sunhubbardmy_amd import FTWalkers, Model
from sunhubbardmy_amd import _lib
from sunhubbardmy_amd.model import make_tables

pytestmark = pytest.mark.gpu

SIZES = [(4, 4), (6, 8), (8, 8), (10, 10), (9, 13)]
SEL = {"G": 0, "Gs": 1, "W1": 2, "W2": 3, "W3": 4, "W4": 5, "Us": 6, "Vs": 7, "U": 8, "V": 9}


def say(*a):
    print(*a, flush=True)


class Box:
    def __init__(self, la, lb, **kw):
        self.m = Model(la=la, lb=lb, beta=1.0, nwrap=5, **kw)
        self.w = FTWalkers([self.m])
        self.n = self.w.tables[0]["ndim"]
        self.L = _lib.testing_lib()

    def put(self, name, a, pt=0):
        a = np.asfortranarray(a, dtype=np.complex128)
        assert self.L.sh_dbg_put(self.w._ctx, 0, SEL[name], pt, a.ctypes.data_as(_lib.dp)) == 0

    def get(self, name, pt=0):
        a = np.zeros((self.n, self.n), np.complex128, order="F")
        assert self.L.sh_dbg_get(self.w._ctx, 0, SEL[name], pt, a.ctypes.data_as(_lib.dp)) == 0
        return a

    def put_d(self, d, pt):
        d = np.ascontiguousarray(d, dtype=np.float64)
        assert self.L.sh_dbg_put(self.w._ctx, 0, 100, pt, d.ctypes.data_as(_lib.dp)) == 0

    def get_vec(self, sel, pt=0, size=None):
        out = np.zeros(size or 2 * self.n)
        assert self.L.sh_dbg_get(self.w._ctx, 0, sel, pt, out.ctypes.data_as(_lib.dp)) == 0
        return out

    def d(self, pt):
        return self.get_vec(100, pt, self.n)

    def ldq(self, pt):
        v = self.get_vec(103, pt, 2)
        return complex(v[0], v[1])

    def sc(self):
        return self.get_vec(104, 0, 32)

    def op(self, code, *args):
        a = np.zeros(8, np.int32)
        a[:len(args)] = args
        rc = self.L.sh_dbg_op(self.w._ctx, code, a.ctypes.data_as(_lib.ip))
        assert rc == 0, rc


def rnd(n, rng, scale=1.0):
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * scale


@pytest.mark.parametrize("la,lb", SIZES)
def test_gemm(la, lb):
    b = Box(la, lb)
    rng = np.random.default_rng(la * 100 + lb)
    A, B = rnd(b.n, rng), rnd(b.n, rng)
    b.put("W1", A)
    b.put("W2", B)
    say("gemm", b.n)
    for code, ref in ((0, A @ B), (1, A.conj().T @ B), (2, A @ B.conj().T)):
        b.op(code)
        C = b.get("W3")
        err = np.abs(C - ref).max() / np.abs(ref).max()
        assert err < 1e-14, (code, err)


@pytest.mark.parametrize("la,lb", SIZES)
def test_qr_pivoted(la, lb):
    b = Box(la, lb)
    rng = np.random.default_rng(7 + la)
    # graded columns, as in a propagated B product
    A = rnd(b.n, rng) @ np.diag(np.exp(rng.uniform(-8, 8, b.n)))
    b.put("W1", A)
    say("qrcp", b.n)
    b.op(3, 1)
    U, V, D = b.get("U", 1), b.get("V", 1), b.d(1)
    assert np.abs(U.conj().T @ U - np.eye(b.n)).max() < 1e-13
    rec = U @ np.diag(D) @ V
    assert np.abs(rec - A).max() / np.abs(A).max() < 1e-13
    q, r, jp, ld = oracle.qr(A, pivot=1)
    assert np.array_equal(b.get_vec(101, 0, b.n).astype(int), jp)
    assert np.allclose(D, np.real(np.diag(r)), rtol=1e-12, atol=0)
    assert abs(np.exp(1j * b.ldq(1).imag) - np.linalg.det(U)) < 1e-10
    assert abs(b.ldq(1) - ld) < 1e-10


@pytest.mark.parametrize("la,lb", SIZES)
def test_qr_prepivoted(la, lb):
    b = Box(la, lb)
    rng = np.random.default_rng(11 + lb)
    Bd = rnd(b.n, rng)
    Dp = np.exp(rng.uniform(-10, 10, b.n)) * rng.choice([-1.0, 1.0], b.n)
    Vp = np.triu(rnd(b.n, rng)) + 3 * np.eye(b.n)
    b.put("W1", Bd)
    b.put_d(Dp, 0)
    b.put("V", Vp, 0)
    say("qr prepivot", b.n)
    b.op(4, 1, 0)
    U, V, D = b.get("U", 1), b.get("V", 1), b.d(1)
    assert np.abs(U.conj().T @ U - np.eye(b.n)).max() < 1e-13
    ref = Bd @ np.diag(Dp) @ Vp
    rec = U @ np.diag(D) @ V
    # graded scales: compare column by column relative to each column's size
    err = np.abs(rec - ref).max(axis=0) / np.abs(ref).max(axis=0)
    assert err.max() < 1e-11
    assert abs(np.exp(1j * b.ldq(1).imag) - np.linalg.det(U)) < 1e-10


@pytest.mark.parametrize("la,lb", SIZES)
def test_lu_solve(la, lb):
    b = Box(la, lb)
    rng = np.random.default_rng(3 + la + lb)
    M, X = rnd(b.n, rng) + 2 * np.eye(b.n), rnd(b.n, rng)
    b.put("W3", M)
    b.put("W4", X)
    say("lu", b.n)
    b.op(5)
    Y = b.get("W4")
    ref = np.linalg.solve(M, X)
    assert np.abs(Y - ref).max() / np.abs(ref).max() < 1e-11
    sc = b.sc()
    sign, logabs = np.linalg.slogdet(M)
    assert abs(sc[6] - logabs) < 1e-10
    assert abs(np.exp(1j * sc[7]) - sign) < 1e-10
    assert -np.pi <= sc[7] <= np.pi and sc[14] == 0
    _, ld, _ = oracle.inv_logdet(M)
    assert abs(complex(sc[6], sc[7]) - ld) < 1e-9


def _bond_ops(t, ksel):
    lat = t["lat"]
    n = t["ndim"]
    K = t["urt"] if ksel == 0 else t["urtm1"]
    mats = []
    for nf in range(lat["nn_nf"]):
        F = np.eye(n, dtype=complex)
        for b_ in range(lat["nn_lf"]):
            ist = nf * lat["nn_lf"] + b_
            i1, i2 = lat["bonds"][nf, b_]
            blk = K[4 * ist:4 * ist + 4].reshape(2, 2, order="F")
            E = np.eye(n, dtype=complex)
            E[np.ix_([i1, i2], [i1, i2])] = blk
            F = E @ F
        mats.append(F)
    return mats


@pytest.mark.parametrize("la,lb", [(4, 4), (6, 8), (10, 10)])
def test_checkerboard_passes(la, lb):
    b = Box(la, lb)
    t = b.w.tables[0]
    rng = np.random.default_rng(5)
    conf = rng.integers(1, 3, size=(t["ltrot"], b.n)).astype(np.int32)
    b.w.set_conf(conf)
    hs = t["hs"]
    A = rnd(b.n, rng)
    say("checkerboard", b.n)
    fam = _bond_ops(t, 0)
    Kf = np.eye(b.n, dtype=complex)
    for F in fam:
        Kf = F @ Kf
    Kr = np.eye(b.n, dtype=complex)
    for F in fam[::-1]:
        Kr = F @ Kr
    # full B slices 2..4 from the left (Bmat_left_forward)
    b.put("W1", A)
    b.op(6, 1 | 2 | 4, 0, 0, 2, 4, 1)
    ref = A.copy()
    for nt in (2, 3, 4):
        ref = Kr @ (np.diag(hs["bmat"][conf[nt - 1] - 1]) @ (Kf @ ref))
    assert np.abs(b.get("W1") - ref).max() / np.abs(ref).max() < 1e-13
    # conj(bmat), slices 4 down to 2 (Bmat_left_forward_hc)
    b.put("W1", A)
    b.op(6, 1 | 2 | 4, 0, 2, 4, 2, -1)
    ref = A.copy()
    for nt in (4, 3, 2):
        ref = Kr @ (np.diag(hs["bmat"][conf[nt - 1] - 1].conj()) @ (Kf @ ref))
    assert np.abs(b.get("W1") - ref).max() / np.abs(ref).max() < 1e-13
    # right pass on G: V (bmat_inv) before the reverse families of urtm1
    famm = _bond_ops(t, 1)
    R = np.eye(b.n, dtype=complex)
    for F in famm[::-1]:
        R = R @ F          # families nn_nf..1 applied from the right in that order
    b.put("G", A)
    b.op(7, 8 | 4, 1, 1, 3)
    ref = A @ np.diag(hs["bmat_inv"][conf[2] - 1]) @ R
    assert np.abs(b.get("G") - ref).max() / np.abs(ref).max() < 1e-13


@pytest.mark.parametrize("la,lb", [(4, 4), (8, 8), (9, 13)])
def test_green(la, lb):
    b = Box(la, lb)
    rng = np.random.default_rng(19)
    n = b.n

    def udv(scale):
        U = np.linalg.qr(rnd(n, rng))[0]
        D = np.exp(rng.uniform(-scale, scale, n)) * rng.choice([-1.0, 1.0], n)
        V = np.triu(rnd(n, rng) * 0.3) + np.eye(n)
        return U, D, V

    # scales kept inside what a direct float64 inverse of I + B_R B_L resolves
    UR, DR, VR = udv(3)
    UL, DL, VL = udv(3)
    b.put("U", UR, 1); b.put_d(DR, 1); b.put("V", VR, 1)
    b.put("Us", UL); b.put("Vs", VL)
    # saved D goes through the saved point: write it into point 0 then save via D of point 0?  The
    # saved D has its own slot; use point 2 as the left point instead
    b.put("U", UL, 2); b.put_d(DL, 2); b.put("V", VL, 2)
    say("green", n)
    b.op(8, 0, 1, 0, 2)
    G = b.get("Gs")
    BR = UR @ np.diag(DR) @ VR
    BL = VL.conj().T @ np.diag(DL) @ UL.conj().T
    ref = np.linalg.inv(np.eye(n) + BR @ BL)
    assert np.abs(G - ref).max() < 1e-10
    b.op(9, 1)
    Gbb = b.get("Gs")
    ref_bb = np.linalg.inv(np.eye(n) + BR)
    assert np.abs(Gbb - ref_bb).max() < 1e-10
