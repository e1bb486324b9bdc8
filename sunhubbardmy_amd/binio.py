"""Equal-time bins in the reference's layout and its per-bin output files.

The library returns the correlation bins of one chain as 20 lq + 8 complex numbers (each summed
with the sign weight zphi over the measured slices, like the reference's *_bin arrays):
zcpcm, zspsm, znn, pair_onsite as (imj, no_i, no_j) at imj + lq (no_i + 2 no_j), then zjj, zbb,
pair_nn, pair_sn (lq each), then zn(1), zn(2), zj, zb, zqsq, ztsq, ztq, zint_v
(src/dqmc_measure/equaltime_measure.f90:44-68; zqsq without its 1/9 and zint_v without rv, both
applied when energy.bin is written).

`write_bin` appends one bin to the files equaltime_output writes on the square and honeycomb
lattices (src/dqmc_measure/equaltime_output.f90:1-180): energy.bin (16e16.8), n.bin, j.bin,
b.bin, nnnn.bin, Crmax.bin and the *_r.bin / *_k.bin pairs of fourier_trans_eqt and
fourier_trans_eqt_22, with the MPI reduction replaced by the sum over chains and rnorm =
(number of chains) * nobs.
"""
import os

import numpy as np

KEYS22 = ("cpcm", "spsm", "nn", "pair_onsite")
KEYS1 = ("jj", "bb", "pair_nn", "pair_sn")
SCALARS = ("zn1", "zn2", "zj", "zb", "zqsq", "ztsq", "ztq", "zint_v")


def eqt_size(lq):
    return 20 * lq + 8


def unpack_eqt(flat, lq):
    """flat complex bins -> dict of arrays: (lq, 2, 2) [imj, no_i, no_j] for KEYS22, (lq,) for KEYS1,
    complex scalars for SCALARS"""
    flat = np.asarray(flat)
    out = {}
    for q, k in enumerate(KEYS22):
        out[k] = flat[4 * lq * q:4 * lq * (q + 1)].reshape(2, 2, lq).transpose(2, 1, 0)
    for q, k in enumerate(KEYS1):
        out[k] = flat[lq * (16 + q):lq * (17 + q)]
    for q, k in enumerate(SCALARS):
        out[k] = flat[20 * lq + q]
    return out


def fortran_e(v, d=8, w=16):
    """Fortran Ew.d edit descriptor (0.ddddddddE+xx), right-justified in w columns"""
    v = float(v)
    if v == 0.0 or not np.isfinite(v):
        body = ("-" if np.signbit(v) else "") + ("0." + "0" * d + "E+00" if v == 0.0 else str(v))
        return body.rjust(w)
    m, e = ("%.*e" % (d - 1, abs(v))).split("e")
    digits = m.replace(".", "")
    exp = int(e) + 1
    es = "E%+03d" % exp if abs(exp) < 100 else "%+04d" % exp
    return (("-" if v < 0 else "") + "0." + digits + es).rjust(w)


def _record(f, values):
    f.write("".join(fortran_e(x) for x in values) + "\n")


def _reals(zs):
    out = []
    for z in np.ravel(zs):
        out += [z.real, z.imag]
    return out


def k_list(l1, l2):
    """latt%listk of setup_latt2d_sub (src/lattice.f90:463-470): m outer, n inner"""
    return [(m, n) for m in range(-((l1 + 1) // 2) + 1, l1 // 2 + 1) for n in range(-((l2 + 1) // 2) + 1, l2 // 2 + 1)]


def _phase_table(l1, l2):
    """zexpiqr(nk, imj) = exp(i k . a_imj), k . a = 2 pi (m x / l1 + n y / l2) for cell imj at
    (x, y), imj = y + (x - 1) l2 (lattice.f90:481-489)"""
    ks = np.array(k_list(l1, l2), float)
    x = np.repeat(np.arange(1, l1 + 1), l2)
    y = np.tile(np.arange(1, l2 + 1), l1)
    return np.exp(2j * np.pi * (np.outer(ks[:, 0], x) / l1 + np.outer(ks[:, 1], y) / l2)), ks


def _fourier(gr, l1, l2, tag, path):
    """fourier_trans_eqt (gr (lq,)) or fourier_trans_eqt_22 (gr (lq, nsub, nsub))"""
    lq = l1 * l2
    z, ks = _phase_table(l1, l2)
    with open(os.path.join(path, tag + "_r.bin"), "a") as f:
        for imj in range(lq):
            _record(f, _reals(gr[imj].T if gr.ndim == 3 else gr[imj]))
    gk = np.tensordot(1.0 / z, gr, axes=(1, 0)) / lq
    with open(os.path.join(path, tag + "_k.bin"), "a") as f:
        for nk in range(lq):
            _record(f, [ks[nk, 0] / l1, ks[nk, 1] / l2] + _reals(gk[nk].T if gr.ndim == 3 else gk[nk]))


def write_bin(path, lattice, l1, l2, nobs, obs, eqt, rv=0.0):
    """Append one bin: obs = per chain the complex sums (sign, density, zkint, zint_u) of the
    library's equal-time measurement, eqt = per chain the flat correlation bins, nobs = measured
    slices per chain in the bin."""
    if lattice not in ("square", "honeycomb"):
        raise ValueError("the *.bin writers cover the square and honeycomb lattices")
    nsub = 2 if lattice == "honeycomb" else 1
    lq = l1 * l2
    rnorm = float(len(eqt) * nobs)
    tot = unpack_eqt(np.sum(np.asarray(eqt), axis=0), lq)
    e4 = np.sum(np.asarray(obs), axis=0)
    energy = np.array([e4[0], e4[1], e4[2], e4[3], tot["zqsq"] / 9.0, tot["ztsq"], tot["ztq"],
                       tot["zint_v"] * rv]) / rnorm
    os.makedirs(path, exist_ok=True)
    for tag in KEYS22:
        _fourier(tot[tag][:, :nsub, :nsub] / rnorm, l1, l2, tag, path)
    for tag in KEYS1:
        _fourier(tot[tag] / rnorm, l1, l2, tag, path)
    with open(os.path.join(path, "energy.bin"), "a") as f:
        _record(f, _reals(energy))
    with open(os.path.join(path, "n.bin"), "a") as f:
        _record(f, _reals(np.array([tot["zn1"], tot["zn2"]]) / rnorm))
    for name, v in (("j.bin", tot["zj"]), ("b.bin", tot["zb"]), ("nnnn.bin", 0j), ("Crmax.bin", 0j)):
        # znnnn and zCrmax are never accumulated in the reference (their calls are commented out)
        with open(os.path.join(path, name), "a") as f:
            _record(f, _reals(np.array([v]) / rnorm))
