"""Equal-time bins in the reference's layout and its per-bin output files.

The library returns the correlation bins of one chain as 20 lq + 8 complex numbers (each summed
with the sign weight zphi over the measured slices, like the reference's *_bin arrays):
zcpcm, zspsm, znn, pair_onsite as (imj, no_i, no_j) at imj + lq (no_i + 2 no_j), then zjj, zbb,
pair_nn, pair_sn (lq each), then zn(1), zn(2), zj, zb, zqsq, ztsq, ztq, zint_v
(src/dqmc_measure/equaltime_measure.f90:44-68; zqsq without its 1/9 and zint_v without rv, both
applied when energy.bin is written).

`write_bin` appends one bin to the files equaltime_output writes, with the MPI reduction
replaced by the sum over chains and rnorm = (number of chains) * nobs:
* square and honeycomb (src/dqmc_measure/equaltime_output.f90:1-180): energy.bin (16e16.8),
  n.bin, j.bin, b.bin, nnnn.bin, Crmax.bin and the *_r.bin / *_k.bin pairs of the eight
  correlation functions;
* chain (equaltime_output1d.f90): energy.bin (20e16.8 of energy_bin(1:10), zint_v in column 5),
  n.bin, j.bin, b.bin and the same eight pairs with one k coordinate;
* cube (equaltime_output3d.f90): energy.bin (8e16.8), n.bin (2e16.8), b.bin and the pairs of
  cpcm, spsm, nn, bb with (3f16.8, 2e16.8) k lines.
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


def _krange(l):
    return range(-((l + 1) // 2) + 1, l // 2 + 1)


def k_list(l1, l2, l3=None):
    """latt%listk: 2D m outer, n inner (src/lattice.f90:463-470); 3D l (z) outermost, then m, n
    (:1116-1125); the chain's list is m alone (:105-109, l2 = l3 = None is not used for it)"""
    if l3 is None:
        return [(m, n) for m in _krange(l1) for n in _krange(l2)]
    return [(m, n, l) for l in _krange(l3) for m in _krange(l1) for n in _krange(l2)]


def _phase_table(dims):
    """zexpiqr(nk, imj) = exp(i k . a_imj) with k . a = 2 pi sum_d k_d x_d / l_d for the cell imj at
    x (imj = y + (x - 1) l2 in 2D, + (z - 1) l1 l2 in 3D, x on the chain: lattice.f90:481-489,
    :1147-1154)"""
    if len(dims) == 1:
        ks = np.array([(m,) for m in _krange(dims[0])], float)
        xs = np.arange(1, dims[0] + 1)[:, None]
    elif len(dims) == 2:
        ks = np.array(k_list(*dims), float)
        xs = np.array([(x, y) for x in range(1, dims[0] + 1) for y in range(1, dims[1] + 1)], float)
    else:
        l1, l2, l3 = dims
        ks = np.array(k_list(l1, l2, l3), float)
        xs = np.array([(x, y, z) for z in range(1, l3 + 1) for x in range(1, l1 + 1) for y in range(1, l2 + 1)], float)
    arg = sum(np.outer(ks[:, d], xs[:, d]) / dims[d] for d in range(len(dims)))
    return np.exp(2j * np.pi * arg), ks


def _fourier(gr, dims, tag, path):
    """fourier_trans_eqt (gr (lq,)) or fourier_trans_eqt_22 (gr (lq, nsub, nsub))"""
    lq = int(np.prod(dims))
    z, ks = _phase_table(dims)
    with open(os.path.join(path, tag + "_r.bin"), "a") as f:
        for imj in range(lq):
            _record(f, _reals(gr[imj].T if gr.ndim == 3 else gr[imj]))
    gk = np.tensordot(1.0 / z, gr, axes=(1, 0)) / lq
    with open(os.path.join(path, tag + "_k.bin"), "a") as f:
        for nk in range(lq):
            kc = [ks[nk, d] / dims[d] for d in range(len(dims))]
            vals = _reals(gk[nk].T if gr.ndim == 3 else gk[nk])
            if len(dims) == 3:
                f.write("".join("%16.8f" % x for x in kc) + "".join(fortran_e(x) for x in vals) + "\n")
            else:
                _record(f, kc + vals)


def write_bin(path, lattice, l1, l2, nobs, obs, eqt, rv=0.0, l3=1):
    """Append one bin: obs = per chain the complex sums (sign, density, zkint, zint_u) of the
    library's equal-time measurement, eqt = per chain the flat correlation bins, nobs = measured
    slices per chain in the bin; (l1, l2, l3) the cell grid (l2 ignored on the chain)."""
    if lattice not in ("square", "honeycomb", "chain", "cubic"):
        raise ValueError("unknown lattice %r" % lattice)
    nsub = 2 if lattice == "honeycomb" else 1
    dims = {"chain": (l1,), "cubic": (l1, l2, l3)}.get(lattice, (l1, l2))
    lq = int(np.prod(dims))
    rnorm = float(len(eqt) * nobs)
    tot = unpack_eqt(np.sum(np.asarray(eqt), axis=0), lq)
    e4 = np.sum(np.asarray(obs), axis=0)
    os.makedirs(path, exist_ok=True)
    if lattice == "cubic":
        for tag in ("cpcm", "spsm", "nn"):
            _fourier(tot[tag][:, 0, 0] / rnorm, dims, tag, path)
        _fourier(tot["bb"] / rnorm, dims, "bb", path)
        with open(os.path.join(path, "energy.bin"), "a") as f:
            _record(f, _reals(np.array(e4[:4]) / rnorm))
        for name, v in (("n.bin", tot["zn1"]), ("b.bin", tot["zb"])):
            with open(os.path.join(path, name), "a") as f:
                _record(f, _reals(np.array([v]) / rnorm))
        return
    for tag in KEYS22:
        _fourier(tot[tag][:, :nsub, :nsub] / rnorm, dims, tag, path)
    for tag in KEYS1:
        _fourier(tot[tag] / rnorm, dims, tag, path)
    if lattice == "chain":
        energy = np.zeros(10, complex)
        energy[:4] = e4[:4]
        energy[4] = tot["zint_v"]
    else:
        energy = np.array([e4[0], e4[1], e4[2], e4[3], tot["zqsq"] / 9.0, tot["ztsq"], tot["ztq"],
                           tot["zint_v"] * rv])
    with open(os.path.join(path, "energy.bin"), "a") as f:
        _record(f, _reals(energy / rnorm))
    with open(os.path.join(path, "n.bin"), "a") as f:
        _record(f, _reals(np.array([tot["zn1"], tot["zn2"]]) / rnorm))
    extra = (("j.bin", tot["zj"]), ("b.bin", tot["zb"]))
    if lattice != "chain":
        extra += (("nnnn.bin", 0j), ("Crmax.bin", 0j))
    for name, v in extra:
        # znnnn and zCrmax are never accumulated in the reference (their calls are commented out)
        with open(os.path.join(path, name), "a") as f:
            _record(f, _reals(np.array([v]) / rnorm))


def write_dyn_bin(path, lattice, l1, l2, nsweep, znn, l3=1):
    """Append one bin of dyn_output: znn = per chain the library's znn_tau bins
    [ntdm+1][no_j][no_i][lq], summed over nsweep dynamic passes, normalised by
    1 / (nsweep * chains) and written as nn_rtau.bin / nn_ktau.bin in the layout of the lattice's
    dyn_output:
      square, honeycomb (src/dqmc_measure/dyn_output.f90) and chain (dyn_output1d.f90): per imj,
        per t = 0 .. ntdm the 8e16.8 record of gr(imj, 1:nsub, 1:nsub, t); per k its coordinates
        (2f16.8 in 2D; the one coordinate k/l1 on the chain), then per t the 8e16.8 components;
      cube (dyn_output3d.f90): per imj, per t = 1 .. ntdm (the reference skips t = 0) the 2e16.8
        record of gr(imj, t); per k its 3f16.8 coordinates, then per t the 2e16.8 components.
    (dyn_output3d also writes spsm_tau and gf_tau files; those bins are not measured here.)"""
    if lattice not in ("square", "honeycomb", "chain", "cubic"):
        raise ValueError("unknown lattice %r" % lattice)
    nsub = 2 if lattice == "honeycomb" else 1
    dims = {"chain": (l1,), "cubic": (l1, l2, l3)}.get(lattice, (l1, l2))
    z = np.sum(np.asarray(znn), axis=0) / float(nsweep * len(znn))     # [t][no_j][no_i][imj]
    gr = z.transpose(3, 2, 1, 0)[:, :nsub, :nsub, :]                   # [imj][no_i][no_j][t]
    lq, nt = gr.shape[0], gr.shape[3]
    ph, ks = _phase_table(dims)
    gk = np.tensordot(1.0 / ph, gr, axes=(1, 0)) / lq
    os.makedirs(path, exist_ok=True)
    t0 = 1 if lattice == "cubic" else 0
    with open(os.path.join(path, "nn_rtau.bin"), "a") as f:
        for imj in range(lq):
            for t in range(t0, nt):
                _record(f, _reals(gr[imj, :, :, t].T))
    with open(os.path.join(path, "nn_ktau.bin"), "a") as f:
        for nk in range(lq):
            f.write("".join("%16.8f" % (ks[nk, d] / dims[d]) for d in range(len(dims))) + "\n")
            for t in range(t0, nt):
                _record(f, _reals(gk[nk, :, :, t].T))
