"""The HS-field text formats of conf_u / conf_v (src/dqmc_config_u/dqmc_output_conf.f90:27-41,
src/dqmc_config_v/dqmc_output_conf.f90:27-44, src/dqmc_config_u/dqmc_input_conf.f90:21-99):
record layout, value order, the list-directed reading rules and the short-file rule."""
import numpy as np
import pytest

from sunhubbardmy_amd import confio


def test_conf_u_text_layout(tmp_path):
    # two chains, 3 sites, 2 slices: a list-directed `1` (12 columns), then (i4) records with
    # i inner, nt outer, chain 0 first
    c0 = np.array([[1, 2, 2], [2, 1, 1]])
    c1 = np.array([[2, 2, 2], [1, 1, 2]])
    p = tmp_path / "conf_u.out"
    confio.write_conf_u(p, [c0, c1])
    text = p.read_text()
    assert text == "           1\n" + "".join("%4d\n" % v for v in [1, 2, 2, 2, 1, 1, 2, 2, 2, 1, 1, 2])


def test_conf_v_nf_innermost(tmp_path):
    # conf(i, nf, nt) with nf innermost in the file; the library's layout is [nt][nf][i]
    conf = np.arange(2 * 3 * 4).reshape(2, 3, 4) % 2 + 1     # ltrot 2, nn_nf 3, nsites 4
    p = tmp_path / "conf_v.out"
    confio.write_conf_v(p, [conf])
    vals = [int(x) for x in p.read_text().split()[1:]]
    expect = [conf[nt, nf, i] for nt in range(2) for i in range(4) for nf in range(3)]
    assert vals == expect
    lw, back = confio.read_conf_v(p, 1, 4, 3, 2)
    assert lw is False and np.array_equal(back[0], conf)


def test_conf_plqu_layout_and_short_file(tmp_path):
    # conf_plqu(i, nt): plaquette inner, nt outer (dqmc_output_plqconf.f90:23-38), same records as
    # conf_u; a file short for chain 1 copies chain 0 (dqmc_input_plqconf.f90:85-95)
    c0 = np.array([[1, 4, 2], [3, 3, 1]])
    p = tmp_path / "conf_plqu.out"
    confio.write_conf_plqu(p, [c0])
    assert p.read_text() == "           1\n" + "".join("%4d\n" % v for v in [1, 4, 2, 3, 3, 1])
    lw, back = confio.read_conf_plqu(p, 3, 3, 2)
    assert lw is False and len(back) == 3 and all(np.array_equal(b, c0) for b in back)
    (tmp_path / "zero").write_text("0\n")
    assert confio.read_conf_plqu(tmp_path / "zero", 2, 3, 2) == (True, None)


def test_round_trip_many_chains(tmp_path):
    rng = np.random.default_rng(5)
    confs = [rng.integers(1, 5, size=(7, 9)) for _ in range(4)]
    p = tmp_path / "c.out"
    confio.write_conf_u(p, confs)
    lw, back = confio.read_conf_u(p, 4, 9, 7)
    assert lw is False
    assert all(np.array_equal(a, b) for a, b in zip(confs, back))


def test_fresh_start_when_missing_or_zero(tmp_path):
    assert confio.read_conf_u(tmp_path / "none.in", 2, 3, 2) == (True, None)
    p = tmp_path / "zero.in"
    p.write_text("0\n1\n2\n")
    assert confio.read_conf_u(p, 1, 1, 2) == (True, None)


def test_one_value_per_record(tmp_path):
    # each READ consumes one record: values after the first on a line are skipped, blank records
    # are passed over
    p = tmp_path / "rec.in"
    p.write_text("  5  9\n\n 1 2 2\n2\n\n 1\n2 1\n")
    lw, back = confio.read_conf_u(p, 1, 2, 2)
    assert lw is False
    assert np.array_equal(back[0], [[1, 2], [1, 2]])


def test_short_file_copies_the_first_chain(tmp_path):
    # chain 1 runs out: chains 1 and 2 take chain 0's field (dqmc_input_conf.f90:85-95)
    p = tmp_path / "short.in"
    p.write_text("1\n" + "".join("%d\n" % v for v in [1, 2, 2, 1] + [2, 2]))
    lw, back = confio.read_conf_u(p, 3, 2, 2)
    assert lw is False
    assert [b.tolist() for b in back] == [[[1, 2], [2, 1]]] * 3
    # too short for chain 0: the reference would run on an uninitialised field
    p.write_text("1\n1\n2\n")
    with pytest.raises(ValueError):
        confio.read_conf_u(p, 1, 2, 2)
