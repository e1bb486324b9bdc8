"""HS-field files in the reference's text formats, and generator state for bit-exact resume.

conf_u.out (src/dqmc_config_u/dqmc_output_conf.f90:27-41): a first record `1` written
list-directed (gfortran: an integer right-justified in 12 columns), then one `(i4)` record per
value, conf(i, nt) with nt outer and i inner, rank 0's field first and then every other rank's
in rank order.  conf_v.out (src/dqmc_config_v/dqmc_output_conf.f90:27-44) is the same with
conf(i, nf, nt), nf innermost; conf_plqu.out (src/dqmc_config_plqu/dqmc_output_plqconf.f90:
15-39) with conf_plqu(i, nt) over the plaquettes.  Here chain c plays the role of MPI rank c.

Reading follows dqmc_input_conf (src/dqmc_config_u/dqmc_input_conf.f90:21-99): a missing file or
a first value of 0 means "start from scratch" (fresh fields drawn from the chains' own streams,
warm-up needed); otherwise every READ takes one value per record (list-directed input of one
item consumes one record, so only a record's first value counts and blank records are
skipped), rank 0 first.  When the file runs out while a later chain is being read, that chain
and all after it copy chain 0's field (:85-95).  A file too short for chain 0 itself leaves the
reference with an uninitialised field; here it is an error.

The reference reseeds its generator on every start (src/dqmc_util.f90:68-75), so a restart
from conf_u.in continues with different random numbers.  `save_state` / `load_state` keep the
fields together with each chain's generator state (the library's sh_get_rng_state /
shp_get_rng_state), which resumes a run bit-exactly.
"""
import os

import numpy as np


def _write_records(path, head, values):
    with open(path, "w") as f:
        f.write("%12d\n" % head)
        f.write("".join("%4d\n" % v for v in values))


def _records(path):
    """First value of every non-blank record, as list-directed READs of one item see them."""
    with open(path) as f:
        for line in f:
            tok = line.replace(",", " ").split()
            if tok:
                yield int(tok[0])


def write_conf_u(path, confs):
    """confs: per chain an int array [ltrot][nsites] (values 1..lcomp)."""
    vals = [v for c in confs for v in np.asarray(c, np.int64).reshape(-1)]
    _write_records(path, 1, vals)


def write_conf_v(path, confs):
    """confs: per chain an int array [ltrot][nn_nf][nsites]; written nf innermost."""
    vals = [v for c in confs for v in np.asarray(c, np.int64).transpose(0, 2, 1).reshape(-1)]
    _write_records(path, 1, vals)


def write_conf_plqu(path, confs):
    """confs: per chain an int array [ltrot][lqp] (values 1..4)."""
    vals = [v for c in confs for v in np.asarray(c, np.int64).reshape(-1)]
    _write_records(path, 1, vals)


def _read(path, nchains, count):
    """(lwarmup, list of flat per-chain arrays or None)"""
    if not os.path.exists(path):
        return True, None
    it = _records(path)
    iseed0 = next(it, 0)
    if iseed0 == 0:
        return True, None
    out = []
    for c in range(nchains):
        vals = []
        for v in it:
            vals.append(v)
            if len(vals) == count:
                break
        if len(vals) < count:
            if c == 0:
                raise ValueError("%s holds %d of the %d values of the first chain" % (path, len(vals), count))
            out.extend(out[0].copy() for _ in range(c, nchains))
            break
        out.append(np.array(vals, np.int32))
    return False, out


def read_conf_u(path, nchains, nsites, ltrot):
    """(lwarmup, [conf [ltrot][nsites]] per chain, or None to start from scratch)"""
    lw, flat = _read(path, nchains, nsites * ltrot)
    return lw, None if flat is None else [f.reshape(ltrot, nsites) for f in flat]


def read_conf_v(path, nchains, nsites, nn_nf, ltrot):
    """(lwarmup, [conf [ltrot][nn_nf][nsites]] per chain, or None to start from scratch)"""
    lw, flat = _read(path, nchains, nsites * nn_nf * ltrot)
    if flat is None:
        return lw, None
    return lw, [np.ascontiguousarray(f.reshape(ltrot, nsites, nn_nf).transpose(0, 2, 1)) for f in flat]


def read_conf_plqu(path, nchains, lqp, ltrot):
    """dqmc_input_plqconf (src/dqmc_config_plqu/dqmc_input_plqconf.f90:21-99): (lwarmup,
    [conf_plqu [ltrot][lqp]] per chain, or None to start from scratch)"""
    lw, flat = _read(path, nchains, lqp * ltrot)
    return lw, None if flat is None else [f.reshape(ltrot, lqp) for f in flat]


def save_state(path, walkers):
    """Fields and generator states of every chain of FTWalkers or ProjWalkers to an .npz file."""
    d = {"kind": np.array(type(walkers).__name__), "rng": np.stack([walkers.rng_state(c) for c in range(walkers.nchains)])}
    d["conf"] = np.stack([walkers.get_conf(c) for c in range(walkers.nchains)])
    if getattr(walkers, "has_uconf", False):
        d["uconf"] = np.stack([walkers.get_uconf(c) for c in range(walkers.nchains)])
    if type(walkers).__name__ == "FTWalkers" and walkers.has_vconf:
        d["confv"] = np.stack([walkers.get_confv(c) for c in range(walkers.nchains)])
    if type(walkers).__name__ == "FTWalkers" and walkers.has_pconf:
        d["confp"] = np.stack([walkers.get_confp(c) for c in range(walkers.nchains)])
    np.savez(path, **d)


def load_state(path, walkers):
    """Restore what save_state wrote into walkers of the same shape (call sweep_start_0b after)."""
    with np.load(path, allow_pickle=False) as z:
        if str(z["kind"]) != type(walkers).__name__ or z["rng"].shape[0] != walkers.nchains:
            raise ValueError("%s holds %s chains of another shape" % (path, str(z["kind"])))
        for c in range(walkers.nchains):
            if type(walkers).__name__ == "FTWalkers":
                walkers.set_conf(z["conf"][c], c)
            else:
                walkers.set_conf(c, z["conf"][c])
            if "uconf" in z:
                walkers.set_uconf(c, z["uconf"][c])
            if "confv" in z:
                walkers.set_confv(z["confv"][c], c)
            if "confp" in z:
                walkers.set_confp(z["confp"][c], c)
            walkers.set_rng_state(c, z["rng"][c])
