"""Bit-exact resume on the GPU: fields plus generator states saved after some sweeps and loaded
into a fresh context (then sweep_start_0b, as a restart does) continue the chains exactly as the
uninterrupted run does; also with learned block proposals and for the projective sweep."""
import sys

import numpy as np
import pytest

from sunhubbardmy_amd import FTWalkers, Model, chain_seed, confio
from sunhubbardmy_amd.learned import EffectiveAction
from sunhubbardmy_amd.proj import ProjModel, ProjWalkers

pytestmark = pytest.mark.gpu

FT = dict(lattice="square", la=4, lb=4, beta=2.0, dtau=0.1, rt=1.0, mu=0.0, rhub=4.0, nu=0.0, nflr=2,
          exact_hs=True, kmode="checkerboard", update="delay", nwrap=5, nublock=8)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sweeps(w, n):
    confs = []
    for _ in range(n):
        w.sweep_b0(True, True)
        w.sweep_0b(True, True)
        confs.append([w.get_conf(c).copy() for c in range(w.nchains)])
    return confs


@pytest.mark.parametrize("blocks", [False, True])
def test_ft_resume(tmp_path, blocks):
    say("ft resume, block proposals", blocks)
    models = [Model(seed=chain_seed(c), **FT) for c in range(3)]

    def setup(w):
        if blocks:
            w.set_neighbours()
            net = EffectiveAction(capacity=16, seed=7)
            net.set_norm(np.array([0.0, 0.0, 0.0, 0.0, 0.3, 0.3, 0.3, 0.3, 0.0, 0.05]))
            w.set_proposal(net, nprop=8, local_updates=True)
            w._net = net

    a = FTWalkers(models)
    setup(a)
    a.init_conf()
    a.sweep_start_0b()
    sweeps(a, 2)
    path = tmp_path / "state.npz"
    confio.save_state(path, a)
    ref = sweeps(a, 3)
    b = FTWalkers(models)
    setup(b)
    confio.load_state(path, b)
    b.sweep_start_0b()
    got = sweeps(b, 3)
    for s in range(3):
        for c in range(3):
            assert np.array_equal(ref[s][c], got[s][c]), "sweep %d chain %d differs" % (s, c)
    for c in range(3):
        assert np.array_equal(a.rng_state(c), b.rng_state(c))
        np.testing.assert_allclose(b.G(c), a.G(c), rtol=0, atol=1e-9)


def test_pqmc_resume(tmp_path):
    kw = dict(lattice="honeycomb", la=3, lb=3, beta=1.0, dtau=0.1, rt=1.0, mu=0.0, rv=1.0, rhub=3.0, nu=0.0,
              rndness=0.2, nflr=2, kmode="checkerboard", nwrap=5, nustock=8, mid_len=2)
    say("pqmc resume")
    models = [ProjModel(seed=chain_seed(c), **kw) for c in range(2)]
    a = ProjWalkers(models)
    a.init_conf()
    a.set_trial()
    a.sweep_start_0b()
    for _ in range(2):
        a.sweep_b0(True, True)
        a.sweep_0b(True, True)
    path = tmp_path / "pstate.npz"
    confio.save_state(path, a)
    ref = []
    for _ in range(2):
        a.sweep_b0(True, True)
        a.sweep_0b(True, True)
        ref.append([(a.get_conf(c).copy(), a.get_uconf(c).copy()) for c in range(2)])
    b = ProjWalkers(models)
    for c in range(2):
        b.set_proj(c, a.projs[c])
    confio.load_state(path, b)
    b.sweep_start_0b()
    for s in range(2):
        b.sweep_b0(True, True)
        b.sweep_0b(True, True)
        for c in range(2):
            assert np.array_equal(ref[s][c][0], b.get_conf(c)), "bond field, sweep %d chain %d" % (s, c)
            assert np.array_equal(ref[s][c][1], b.get_uconf(c)), "on-site field, sweep %d chain %d" % (s, c)


def test_conf_files_round_trip_through_the_library(tmp_path):
    models = [Model(seed=chain_seed(c), **FT) for c in range(2)]
    a = FTWalkers(models)
    a.init_conf()
    a.output_conf(tmp_path / "conf_u.out")
    b = FTWalkers(models)
    assert b.input_conf(tmp_path / "conf_u.out") is False
    for c in range(2):
        assert np.array_equal(a.get_conf(c), b.get_conf(c))
    # no file: a fresh start that draws the same fields as init_conf
    assert b.input_conf(tmp_path / "absent.in") is True
    for c in range(2):
        assert np.array_equal(a.get_conf(c), b.get_conf(c))
