"""Throughput of the DQMC sweep on one MI355X.

Workload: SU(2) Hubbard on the 16x16 periodic square lattice, t = 1, U = 6, half filling,
beta = 8, dtau = 0.1 (Ltrot = 80), nwrap = 10, exact SU(2) HS field, checkerboard kinetic
term, delay update; a batch of independent chains (seeds abs((c*1981+2008)*951049)), each
what one MPI rank of the reference runs.  Synthetic: random initial HS fields from each
chain's SFMT stream.  A step is one sweep (b0 pass + 0b pass, measurement on every slice) of
every chain.  Metric: tau_int-corrected samples/s = chain sweeps/s / (2 tau_int), tau_int the
largest of double occupancy, kinetic energy and S(pi,pi) from the per-sweep bins: the
reference's auto_cor rho(t) with Sokal's window, on the deviations from the batch mean of each
sweep (sunhubbardmy_amd/autocorr.py tau_int_batch), so that the relaxation every chain of the
batch still shares after the warm-up is not read as autocorrelation.  The line also carries the
estimate with per-chain means only (tau_int_per_chain_means).

Sequence: W warm-up sweeps (at least --therm, 20) of the plain delay update from random fields, then 40
more plain sweeps whose configurations after every pass train the effective-action network
(40 sweeps x 2 passes x 256 chains = 20480 configurations, whatever W is; the library's FP32
kernels train it).  The same chains are then timed for exactly K sweeps with the plain sweep,
and K sweeps with block proposals after every slice's local update.  Each phase continues
untimed to --tau-sweeps sweeps in all, so that tau_int comes from a series long enough to
resolve it (the window closes before a quarter of the series); the timed sweeps are the first
K of that series.  value is the learned form's tau_int-corrected samples/s; the line carries
both phases, every tau with its Sokal error, window and a resolved flag, and learned_over_plain
with its error.  A tau whose window did not close makes value null.

CPU baseline: the oracle (plain C restatement of the same sweep) runs one chain per core on up
to 16 cores of the host for one sweep each; the line states the core count and the CPU model,
and a one-core sweep at U = 4 relates the oracle to the reference's single-core figure at that
shape (0.366 sweeps/s, measured on a 2.7 GHz Xeon when the reference was surveyed).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--chains C] [--nublock k] [--proposal both]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "τ_int-corrected samples/s, SU(2) 16×16 β=8, learned vs plain, 1 GPU, FP64 G"
FP64_MFMA_PEAK = 78.6e12   # MI355X dense FP64 matrix, FLOP/s
HBM_PEAK = 8.0e12          # MI355X HBM3E, B/s


def model_kw(args):
    # nublock 0: start from the reference's heuristic depth, then the autotuner picks it
    return dict(lattice="square", la=16, lb=16, beta=8.0, dtau=0.1, rt=1.0, mu=0.0, rhub=6.0, nu=0.0, nflr=2,
                exact_hs=True, kmode="checkerboard", update="delay", nwrap=10, nublock=args.nublock)


def graph_demo(sweeps=40):
    """Host launches against captured passes on the CPU-runnable configuration (one chain of the
    4x4 square lattice, U = 4, beta = 4, fast update, checkerboard K), where the sweep is
    launch-bound: chain sweeps/s with every kernel launched from the host, then with every pass
    replayed from its hipGraph (same results, bit for bit, tests/test_gpu_graphs_tune.py)."""
    from sunhubbardmy_amd import FTWalkers, Model, chain_seed
    kw = dict(lattice="square", la=4, lb=4, beta=4.0, dtau=0.1, rhub=4.0, kmode="checkerboard", update="fast",
              nwrap=10)
    out = {"workload": "SU(2) square 4x4, U=4, beta=4 (Ltrot=40), fast update, checkerboard K, 1 chain"}
    for mode in ("direct", "graphs"):
        w = FTWalkers([Model(seed=chain_seed(0), **kw)])
        w.set_graphs(mode == "graphs")
        w.init_conf()
        w.sweep_start_0b()
        w.sweep(True, True)
        w.synchronize()
        t0 = time.perf_counter()
        for _ in range(sweeps):
            w.sweep(True, True)
        w.synchronize()
        out[mode + "_sweeps_per_s"] = sweeps / (time.perf_counter() - t0)
        w.close()
    out["speedup"] = out["graphs_sweeps_per_s"] / out["direct_sweeps_per_s"]
    return out


def algorithmic_work(group, n, nchains, ltrot, nwrap, acc, k):
    """(kind, amount per launch) of the kernel group, all chains of the launch."""
    nn = float(n) * n
    if group == "wrap":
        return "bytes", nchains * 2 * 16 * nn                       # read + write G
    if group == "measure":
        return "bytes", nchains * 16 * nn                           # read G
    if group == "bprod":
        return "bytes", nchains * 2 * 16 * nn                       # one pass over the propagated U
    if group == "update":
        return "flops", nchains * 8.0 * acc * n * nn * (1.0 + k / n)  # rank-k flushes + column builds
    if group in ("qr", "formq"):
        return "flops", nchains * 16.0 / 3.0 * n * nn
    if group == "gemm":
        return "flops", nchains * 8.0 * n * nn
    if group == "lu":
        return "flops", nchains * 8.0 / 3.0 * n * nn
    if group == "solve":
        return "flops", nchains * 8.0 * n * nn
    return "bytes", nchains * 2 * 16 * nn


def _oracle_sweep(args):
    """One measured sweep of one oracle chain (run in a worker process); returns its seconds."""
    kw, chain = args
    import oracle
    o = oracle.OracleChain(seed=oracle.chain_seed(chain), **kw)
    o.init_conf()
    o.sweep_start()
    t0 = time.perf_counter()
    o.sweep(lupdate=True, lmeasure=True)
    return time.perf_counter() - t0


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(kw):
    """The CPU oracle on a bounded sample: one measured sweep of one chain per core, all cores
    at once (one process each, the reference's one chain per MPI rank), plus one sweep at U = 4
    on one core for the calibration against the reference's surveyed single-core rate.  Runs
    before this process touches the GPU; the tau_int division happens after the GPU run."""
    import multiprocessing as mp
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    cores = max(1, min(16, avail))
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        per = pool.map(_oracle_sweep, [(kw, c) for c in range(cores)])
    wall = time.perf_counter() - t0
    agg = cores / max(per)                       # sweeps per second of the whole host
    kw4 = dict(kw)
    kw4["rhub"] = 4.0
    t_u4 = _oracle_sweep((kw4, 0))
    return {"value": None, "unit": "samples/s", "cores": cores, "kind": "port",
            "cpu_model": _cpu_model(),
            "sample": "one measured sweep (b0+0b) of one chain per core, %d chains on %d cores in parallel "
                      "(slowest %.1f s, wall %.1f s incl. start-up); sweeps/s divided by 2 tau_int of the GPU's plain "
                      "sweep" % (cores, cores, max(per), wall),
            "sweeps_per_s": agg, "sweeps_per_s_one_core": 1.0 / (sum(per) / len(per)),
            "calibration_u4": {"oracle_sweeps_per_s_one_core": 1.0 / t_u4,
                               "reference_sweeps_per_s_one_core": 0.366,
                               "note": "same shape at U = 4; the reference figure was measured on another CPU "
                                       "(2.7 GHz Xeon, AVX-512, MKL) when the reference was surveyed"}}


def copy_calibration():
    """Three copies of 512 MiB between two device buffers, for counter calibration."""
    import torch
    a = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": a.numel() * 8, "launches": 3}), flush=True)


def bench_pqmc(args):
    """Projective t-V sweep: honeycomb 18x18 (N=648, ne=324), t=1, V=1.2, theta=20 (Ltrot=200),
    dtau=0.1, rndness 1e-4, nwrap=10, dense kinetic term, submatrix-LR with the reference's
    default stock (64 rows).  A step is one sweep (b0 + 0b, measurement in the middle slice)."""
    from sunhubbardmy_amd import chain_seed
    from sunhubbardmy_amd.proj import ProjModel, ProjWalkers

    chains = args.chains if args.chains != 256 else 16
    steps = args.steps if args.steps != 60 else 2
    warmup = args.warmup if args.warmup != 60 else 1
    kw = dict(lattice="honeycomb", la=18, lb=18, beta=20.0, dtau=0.1, rt=1.0, rv=1.2, rndness=1e-4, kmode="dense",
              nwrap=10, nustock=0, mid_len=1, wrap_check=False)
    w = ProjWalkers([ProjModel(seed=chain_seed(c), **kw) for c in range(chains)])
    print("# pqmc %d chains, N=%d, ne=%d, Ltrot=%d" % (chains, w.ndim, w.ne, w.ltrot), file=sys.stderr, flush=True)
    w.init_conf()
    w.set_trial()
    w.sweep_start_0b()
    for s in range(warmup):
        w.sweep(lupdate=True, lmeasure=False)
        print("# warmup %d" % s, file=sys.stderr, flush=True)
    w.synchronize()
    w.timers(enable=args.timers)
    t0 = time.perf_counter()
    for s in range(steps):
        w.sweep(lupdate=True, lmeasure=True)
        print("# step %d" % s, file=sys.stderr, flush=True)
    w.synchronize()
    dt = time.perf_counter() - t0
    tm = w.timers(enable=False)
    d = [w.diag(c) for c in range(min(chains, 4))]
    line = {"metric": "projective t-V chain sweeps/s", "value": chains * steps / dt, "unit": "sweeps/s",
            "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": 1000.0 * dt / steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "spinless t-V honeycomb 18x18 (N=648, 324 particles), V=1.2, theta=20, dtau=0.1, "
                                   "nwrap=10, dense K, submatrix-LR (64-row stocks)", "chains": chains},
            "acceptance": float(np.mean([x["accepted"] / max(x["tried"], 1) for x in d]))}
    if args.timers:
        line["kernel_ms"] = {g: round(v[0], 3) for g, v in tm.items()}
        line["kernel_launches"] = {g: v[1] for g, v in tm.items()}
    print(json.dumps(line), flush=True)


def bench_c3(args):
    """The batched submatrix configuration: SU(4) Hubbard, general 4-component HS field, U = 6,
    half filling, beta = 8, dtau = 0.1 (Ltrot = 80), nwrap = 10, FFT kinetic term, submatrix
    update; 32 chains at L = 8 (N = 64), 24 at L = 12 (N = 144) and 16 at L = 16 (N = 256) in
    one batch, one launch per kernel for all of them; the depth and the DFT tiles picked by the
    autotuner.  A step is one sweep of every chain; value = chain sweeps/s of the batch."""
    from sunhubbardmy_amd import FTWalkers, Model, chain_seed
    steps = args.steps if args.steps != 60 else 5
    warmup = args.warmup if args.warmup != 60 else 2
    kw = dict(lattice="square", beta=8.0, dtau=0.1, rt=1.0, mu=0.0, rhub=6.0, nflr=4, exact_hs=False, kmode="fft",
              update="submatrix", nwrap=10, nublock=args.nublock if args.nublock > 0 else 0)
    models, sizes = [], []
    for l, cnt in ((8, 32), (12, 24), (16, 16)):
        for _ in range(cnt):
            models.append(Model(seed=chain_seed(len(models)), la=l, lb=l, **kw))
            sizes.append(l * l)
    w = FTWalkers(models)
    w.init_conf()
    w.sweep_start_0b()
    tuned = w.autotune() if args.nublock == 0 else None
    for s in range(warmup):
        w.sweep(True, False)
    w.reset_diag()
    w.synchronize()
    w.timers(enable=args.timers)
    t0 = time.perf_counter()
    for s in range(steps):
        w.sweep(True, True)
        print("# c3 step %d" % s, file=sys.stderr, flush=True)
    w.synchronize()
    dt = time.perf_counter() - t0
    tm = w.timers(enable=False)
    acc = [w.diag(c)["accepted"] / max(w.diag(c)["tried"], 1) for c in range(len(models))]
    line = {"metric": "chain sweeps/s, SU(4) batched submatrix (C3)", "value": len(models) * steps / dt,
            "unit": "sweeps/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": 1000.0 * dt / steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "SU(4) Hubbard, general 4-component HS, U=6, beta=8, dtau=0.1, FFT K, submatrix; "
                                   "32 x L=8 + 24 x L=12 + 16 x L=16 square chains in one batch",
                       "chains": len(models), "autotuned": tuned,
                       "nublock": w.tables[0]["nublock"]},
            "acceptance_by_lattice": {str(l): float(np.mean([a for a, n in zip(acc, sizes) if n == l * l]))
                                      for l in (8, 12, 16)}}
    if args.timers:
        best = max(tm, key=lambda g: tm[g][0])
        ms, cnt = tm[best]
        a = float(np.mean(acc))
        k = w.tables[0]["nublock"]
        if best == "update":
            amount = sum(8.0 * a * n * n * (n + k) for n in sizes)
            kind = "flops"
        else:
            amount = sum(2 * 16.0 * n * n for n in sizes)
            kind = "bytes"
        avg_s = ms / 1000.0 / max(cnt, 1)
        if kind == "flops":
            line["roofline"] = {"bound": "mfma", "achieved": amount / avg_s / 1e12, "peak": FP64_MFMA_PEAK / 1e12,
                                "unit": "TFLOP/s", "frac": amount / avg_s / FP64_MFMA_PEAK, "traffic": None}
        else:
            line["roofline"] = {"bound": "hbm", "achieved": amount / avg_s / 1e9, "peak": HBM_PEAK / 1e9,
                                "unit": "GB/s", "frac": amount / avg_s / HBM_PEAK, "traffic": None}
        line["roofline"].update({"kernel_group": best, "avg_launch_ms": ms / max(cnt, 1)})
        line["kernel_ms"] = {g: round(v[0], 3) for g, v in tm.items()}
        line["kernel_launches"] = {g: v[1] for g, v in tm.items()}
    print(json.dumps(line), flush=True)


def timed_phase(w, steps, tau_sweeps, timers, label):
    """steps sweeps with measurement, bracketed by device synchronisation, then untimed sweeps up
    to tau_sweeps in all; per-sweep bins of the whole series give tau_int."""
    from sunhubbardmy_amd.autocorr import tau_estimate
    w.reset_obs()
    w.reset_diag()
    w.synchronize()
    w.timers(enable=timers)
    series = []
    t0 = time.perf_counter()
    for s in range(steps):
        w.sweep(lupdate=True, lmeasure=True)
        z, nobs = w.obs_all(reset=True)
        series.append(np.real(z) / nobs[:, None])
        if s % 10 == 0:
            print("# %s step %d" % (label, s), file=sys.stderr, flush=True)
    w.synchronize()
    dt = time.perf_counter() - t0
    tm = w.timers(enable=False)
    nd = min(w.nchains, 8)
    acc = np.mean([w.diag(c)["accepted"] / max(w.diag(c)["tried"], 1) for c in range(nd)])
    for s in range(steps, max(steps, tau_sweeps)):
        w.sweep(lupdate=True, lmeasure=True)
        z, nobs = w.obs_all(reset=True)
        series.append(np.real(z) / nobs[:, None])
        if s % 20 == 0:
            print("# %s tau series sweep %d" % (label, s), file=sys.stderr, flush=True)
    series = np.array(series)            # [sweeps, chains, 5]
    taus, taus_chain = {}, {}
    for name, col in (("dblocc", 3), ("ekin", 2), ("s_pipi", 4)):
        r = tau_estimate(series[:, :, col].T, batch=True)
        taus[name] = {"tau": round(r["tau"], 4), "sigma": round(r["sigma"], 4), "window": r["window"],
                      "max_lag": r["max_lag"], "resolved": r["resolved"]}
        rc = tau_estimate(series[:, :, col].T, batch=False)
        taus_chain[name] = {"tau": round(rc["tau"], 4), "window": rc["window"], "resolved": rc["resolved"]}
    slow = max(taus, key=lambda k: taus[k]["tau"])
    tau, sig = taus[slow]["tau"], taus[slow]["sigma"]
    resolved = all(v["resolved"] for v in taus.values())
    sweeps_per_s = w.nchains * steps / dt
    means = series.mean(axis=(0, 1))
    return {"dt": dt, "tm": tm, "tau": tau, "tau_sigma": sig, "tau_obs": slow, "resolved": resolved, "taus": taus,
            "taus_chain": taus_chain, "sweeps_per_s": sweeps_per_s, "samples_per_s": sweeps_per_s / (2.0 * tau),
            "acc": float(acc), "means": means, "series": series, "series_len": int(series.shape[0])}


def read_traffic(chains, n, nublock):
    """HBM bytes per launch of the update kernel, replayed from the committed counter summary
    (two rocprofv3 --pmc runs of this workload), when the run is the workload it was measured on."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_ft_hubbard16x16_256chains.txt")
    if not (chains == 256 and n == 256 and nublock == 32) or not os.path.exists(path):
        return None
    kv = dict(l.strip().split("=", 1) for l in open(path) if "=" in l and not l.startswith("#"))
    return float(kv["fetch_bytes"]) + float(kv["write_bytes"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--chains", type=int, default=256)
    ap.add_argument("--nublock", type=int, default=0,
                    help="delay depth; 0: the autotuner picks it on the warmed-up chains")
    ap.add_argument("--nprop", type=int, default=16, help="inner Metropolis attempts per block proposal (<= 64)")
    ap.add_argument("--tau-sweeps", type=int, default=100,
                    help="length of each phase's per-sweep series for tau_int (the first --steps are timed)")
    ap.add_argument("--therm", type=int, default=20,
                    help="least number of plain warm-up sweeps from the random start, whatever --warmup is")
    ap.add_argument("--train-sweeps", type=int, default=40,
                    help="plain sweeps after the warm-up whose configurations train the network")
    ap.add_argument("--epochs", type=int, default=30, help="training epochs of the effective action")
    ap.add_argument("--proposal", choices=("both", "plain", "learned"), default="both",
                    help="time the plain delay sweep, the learned-proposal sweep, or both (value = learned)")
    ap.add_argument("--series-out", default=None,
                    help="write the per-sweep observable series of the timed phases to this .npz file")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph-demo", action="store_true",
                    help="skip the direct-launch vs captured-pass comparison on the 4x4 configuration")
    ap.add_argument("--no-timers", dest="timers", action="store_false",
                    help="skip the per-kernel hipEvent timing of the timed steps")
    ap.add_argument("--workload", choices=("ft", "pqmc", "c3"), default="ft",
                    help="ft: the headline finite-temperature workload; pqmc: the projective t-V workload "
                         "(honeycomb L=18, N=648, 324 particles, V=1.2, theta=20, submatrix-LR); c3: the SU(4) "
                         "batched submatrix workload (32 x L=8 + 24 x L=12 + 16 x L=16, FFT K, beta=8)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 512 MiB device-to-device copies (the known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; chains are independent, so N GPUs would run N "
              "independent batches (no collective on the sweep path). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "pqmc":
        return bench_pqmc(args)
    if args.workload == "c3":
        return bench_c3(args)

    from sunhubbardmy_amd import FTWalkers, Model, chain_seed
    from sunhubbardmy_amd.learned import EffectiveAction

    kw = model_kw(args)
    cpu = None if args.no_cpu else cpu_baseline(kw)
    models = [Model(seed=chain_seed(c), **kw) for c in range(args.chains)]
    w = FTWalkers(models)
    t = w.tables[0]
    n, ltrot = t["ndim"], t["ltrot"]
    print("# %d chains, N=%d, Ltrot=%d, nublock=%d" % (args.chains, n, ltrot, t["nublock"]), file=sys.stderr,
          flush=True)
    w.set_neighbours()
    w.init_conf()
    w.sweep_start_0b()
    # warm-up with the plain sweep, then train_sweeps more whose configurations after every pass
    # train the effective action
    learn = args.proposal != "plain"
    nwarm = max(args.warmup, args.therm)
    net = EffectiveAction(capacity=max(2 * args.train_sweeps * args.chains, 1), seed=chain_seed(0)) if learn else None
    for s in range(nwarm + args.train_sweeps):
        collect = learn and s >= nwarm
        w.sweep_b0(True, False)
        if collect:
            net.add_samples(w)
        w.sweep_0b(True, False)
        if collect:
            net.add_samples(w)
        if s % 10 == 0:
            print("# warmup %d" % s, file=sys.stderr, flush=True)
    tuned = None
    if args.nublock == 0:
        tuned = w.autotune()
        print("# autotuned: %s" % tuned, file=sys.stderr, flush=True)
    train = None
    if learn:
        t0 = time.perf_counter()
        loss = net.train(epochs=args.epochs, batch=512, lr=1e-3)
        train = {"samples": net.num_samples(), "epochs": args.epochs, "seconds": round(time.perf_counter() - t0, 3),
                 "loss_first_epoch": float(loss[0]), "loss_last_epoch": float(loss[-1])}
        print("# trained on %d configurations: loss %.4g -> %.4g" % (net.num_samples(), loss[0], loss[-1]),
              file=sys.stderr, flush=True)
    plain = learned = None
    if args.proposal in ("both", "plain"):
        plain = timed_phase(w, args.steps, args.tau_sweeps, args.timers, "plain")
    if learn:
        w.set_proposal(net, nprop=args.nprop, local_updates=True)
        w.reset_proposal_stats()
        learned = timed_phase(w, args.steps, args.tau_sweeps, args.timers, "learned")
        pst = [w.proposal_stats(c) for c in range(min(args.chains, 16))]
        learned["block"] = {
            "acceptance": sum(x["accepted"] for x in pst) / max(sum(x["moves"] for x in pst), 1),
            "spins_per_accepted_block": sum(x["spins_accepted"] for x in pst) / max(sum(x["accepted"] for x in pst), 1),
            "spins_proposed_per_block": sum(x["spins_proposed"] for x in pst) / max(sum(x["moves"] for x in pst), 1)}
    if args.series_out:
        np.savez(args.series_out, **{k: r["series"] for k, r in (("plain", plain), ("learned", learned))
                                     if r is not None})
    head = learned if learned is not None else plain
    wl = ("SU(2) Hubbard square 16x16 (N=256), U=6, t=1, half filling, beta=8, dtau=0.1 (Ltrot=80), nwrap=10, "
          "checkerboard K, delay update k=%d" % t["nublock"])

    def summary(r):
        return {"samples_per_s": r["samples_per_s"] if r["resolved"] else None, "sweeps_per_s": r["sweeps_per_s"],
                "tau_int": r["taus"], "tau_int_headline": {"observable": r["tau_obs"], "tau": r["tau"],
                                                           "sigma": r["tau_sigma"], "resolved": r["resolved"],
                                                           "series_sweeps": r["series_len"]},
                "tau_int_per_chain_means": r["taus_chain"],
                "acceptance": round(r["acc"], 4), "ms_per_step": 1000.0 * r["dt"] / args.steps,
                "observables": {"dblocc_per_site": r["means"][3] / n, "ekin_per_site_flavour": r["means"][2] / n,
                                "s_pipi": r["means"][4]}}

    line = {
        "metric": METRIC, "value": head["samples_per_s"] if head["resolved"] else None, "unit": "samples/s",
        "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000.0 * head["dt"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl, "chains": args.chains, "nublock": t["nublock"],
                   "nublock_from": "autotuner (sh_autotune)" if tuned is not None else "--nublock",
                   "proposal": "learned (block moves of %d attempts after each slice's local update)" % args.nprop
                   if learned is not None else "plain local delay update"},
    }
    if plain is not None:
        line["plain"] = summary(plain)
    if learned is not None:
        line["learned"] = summary(learned)
        line["learned"]["block"] = learned["block"]
        line["training"] = train
    if plain is not None and learned is not None:
        ratio = learned["samples_per_s"] / plain["samples_per_s"]
        rel = np.hypot(plain["tau_sigma"] / plain["tau"], learned["tau_sigma"] / learned["tau"])
        line["learned_over_plain"] = ratio
        line["learned_over_plain_sigma"] = ratio * rel
        line["learned_over_plain_sweeps"] = learned["sweeps_per_s"] / plain["sweeps_per_s"]
    if args.timers:
        tm = head["tm"]
        a = head["acc"]
        best = max(tm, key=lambda g: tm[g][0])
        ms, cnt = tm[best]
        kind, amount = algorithmic_work(best, n, args.chains, ltrot, 10, a, t["nublock"])
        avg_s = ms / 1000.0 / max(cnt, 1)
        achieved = amount / avg_s
        traffic = read_traffic(args.chains, n, t["nublock"]) if best == "update" else None
        traffic_src = "profiles/hbm_traffic_ft_hubbard16x16_256chains.txt (replayed, not measured in this run)" \
            if traffic is not None else None
        if kind == "flops":
            roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": FP64_MFMA_PEAK / 1e12, "unit": "TFLOP/s",
                    "frac": achieved / FP64_MFMA_PEAK, "traffic": traffic}
        else:
            roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK, "traffic": traffic}
        roof["traffic_source"] = traffic_src
        roof["kernel_group"] = best
        roof["avg_launch_ms"] = ms / max(cnt, 1)
        line["roofline"] = roof
        line["kernel_ms"] = {g: round(v[0], 3) for g, v in tm.items()}
        line["kernel_launches"] = {g: v[1] for g, v in tm.items()}
    if cpu is not None:
        cpu["value"] = cpu["sweeps_per_s"] / (2.0 * (plain or head)["tau"])
        line["cpu_baseline"] = cpu
    if not args.no_graph_demo:
        line["graph_capture"] = graph_demo()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
