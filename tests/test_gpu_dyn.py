"""GPU parity of the time-displaced pass (ltau) through the C ABI: the chain is swept with
updates, then a b0 pass builds the left chain, and the dynamic 0b pass (no updates, green_tau at
every stabilization point, znn_tau bins) runs on the device and in the CPU oracle from the same
field.  Bar: G(t,0), G(0,t), G(0,0) and G(t,t) at the end within 1e-9 relative max-norm of the
oracle's, the znn_tau bins within 1e-9 relative, the propagated-versus-recomputed difference
(-DDYNERROR) below 1e-8 on both, and push_stage / pop_stage leaving the chain as it was."""
import sys

import numpy as np
import pytest

import oracle
from sunhubbardmy_amd import FTWalkers, Model, chain_seed

pytestmark = pytest.mark.gpu

CASES = {
    "square4_delay": dict(lattice="square", la=4, lb=4, beta=1.0, rhub=4.0, kmode="checkerboard", nwrap=4,
                          update="delay", nublock=4),
    "honeycomb3_dense": dict(lattice="honeycomb", la=3, lb=3, beta=1.2, rhub=3.0, kmode="dense", nwrap=3,
                             update="fast"),
    "square6_ntdm5": dict(lattice="square", la=6, lb=6, beta=1.5, rhub=5.0, kmode="checkerboard", nwrap=5,
                          update="delay", nublock=8, ntdm=5),
    "chain16_delay": dict(lattice="chain", la=16, lb=1, beta=1.0, rhub=4.0, kmode="checkerboard", nwrap=4,
                          update="delay", nublock=4),
    "cubic4_dense": dict(lattice="cubic", la=4, lb=4, lc=4, beta=0.8, rhub=4.0, kmode="dense", nwrap=4,
                         update="fast"),
    "square4_tv": dict(lattice="square", la=4, lb=4, beta=1.0, rhub=0.0, rv=1.5, nflr=1, kmode="checkerboard",
                       nwrap=4, update="delay", nublock=4),
    "honeycomb3_tuv_dense": dict(lattice="honeycomb", la=3, lb=3, beta=1.0, rhub=3.0, rv=1.0, nflr=2, kmode="dense",
                                 nwrap=5, update="delay", nublock=4),
    # the plaquette field (hconf): its factors in G(t,0), G(0,t) and the B products
    "square4_plq_delay": dict(lattice="square", la=4, lb=4, beta=1.0, rhub=3.0, rhub_plq=1.5, alpha=0.3,
                              kmode="checkerboard", nwrap=4, update="delay", nublock=4),
    "honeycomb3_plq_dense": dict(lattice="honeycomb", la=3, lb=3, beta=0.8, rhub=0.0, rhub_plq=1.2, alpha=0.3,
                                 theta=0.25, kmode="dense", nwrap=4, update="fast"),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("name", CASES)
def test_dyn_parity(name):
    kw = dict(CASES[name])
    ntdm = kw.pop("ntdm", None)
    seed = chain_seed(2)
    say("case", name, "seed", seed)
    w = FTWalkers([Model(seed=seed, **kw)])
    o = oracle.OracleChain(seed=seed, **kw)
    w.init_conf()
    o.init_conf()
    assert np.array_equal(w.get_conf(), o.conf)
    w.sweep_start_0b()
    o.sweep_start()
    w.sweep(lupdate=True, lmeasure=False)
    o.sweep(lupdate=True, lmeasure=False)
    assert np.array_equal(w.get_conf(), o.conf), "field differs after the updating sweep"
    if kw.get("rv"):
        assert np.array_equal(w.get_confv(), o.confv), "bond field differs after the updating sweep"
    if kw.get("rhub_plq"):
        assert np.array_equal(w.get_confp(), o.confp), "plaquette field differs after the updating sweep"
    w.sweep_b0(False, False)
    o.sweep_b0(lupdate=False, lmeasure=False)
    g_before = w.G()
    w.set_dyn(ntdm)
    o.set_dyn(ntdm)
    w.push_stage()
    w.sweep_0b(False, False, True)
    o.sweep_0b_dyn()
    gt0, g0t, g00 = w.gtau()
    ogt0, og0t, og00 = o.gtau()
    say("gt0", rel(gt0, ogt0), "g0t", rel(g0t, og0t), "g00", rel(g00, og00), "gtt", rel(w.G(), o.G))
    assert rel(gt0, ogt0) < 1e-9 and rel(g0t, og0t) < 1e-9 and rel(g00, og00) < 1e-9
    assert rel(w.G(), o.G) < 1e-9
    z, err = w.dyn_obs()
    oz = o.znn()
    say("znn", rel(z, oz), "xmax_dyn", err["xmax_dyn"], o.dyn_error()["xmax_dyn"])
    assert z.shape == oz.shape and rel(z, oz) < 1e-9
    assert err["count"] == o.dyn_error()["count"]
    assert err["xmax_dyn"] < 1e-8 and o.dyn_error()["xmax_dyn"] < 1e-8
    w.pop_stage()
    assert np.array_equal(w.G(), g_before)
    # a second dynamic pass adds the same bins again (the field is unchanged)
    w.push_stage()
    w.sweep_0b(False, False, True)
    w.pop_stage()
    z2, _ = w.dyn_obs()
    assert rel(z2, 2 * z) < 1e-9


def test_dyn_requires_setup_and_no_updates():
    kw = CASES["square4_delay"]
    w = FTWalkers([Model(seed=chain_seed(0), **kw)])
    w.init_conf()
    w.sweep_start_0b()
    with pytest.raises(Exception):
        w.sweep_0b(False, False, True)          # before set_dyn
    w.set_dyn()
    with pytest.raises(Exception):
        w.sweep_0b(True, False, True)           # updates and the dynamic pass together
