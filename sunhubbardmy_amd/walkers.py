"""FTWalkers: a batch of independent finite-temperature DQMC chains advanced on one GPU.

Host mirror of the reference's sweep interface (src/dqmc_ft_ctrl.f90:9-126 calls these):
``allocate_core`` -> ``FTWalkers(models)``, ``u0conf%input_conf`` -> ``init_conf``,
``dqmc_ft_sweep_start_0b`` -> ``sweep_start_0b``, ``dqmc_ft_sweep_b0(lupdate, lmeasure)`` ->
``sweep_b0``, ``dqmc_ft_sweep_0b(lupdate, lmeasure, ldyn)`` -> ``sweep_0b``, ``push_stage`` /
``pop_stage``.  Each chain is what one MPI rank of the reference runs; the batch replaces the
ranks.  All compute goes through libsunhip.so; nothing here falls back to the CPU.
"""
import ctypes

import numpy as np

from . import _lib
from .model import Model, make_tables


TIMER_GROUPS = ["wrap", "update", "bprod", "qr", "formq", "gemm", "lu", "solve", "measure", "other", "proposal"]


class SweepError(RuntimeError):
    pass


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class FTWalkers:
    def __init__(self, models, device=0):
        if isinstance(models, Model):
            models = [models]
        self.models = list(models)
        self.tables = [make_tables(m) for m in self.models]
        self._keep = []
        arr = (_lib.ShModel * len(self.tables))()
        for k, t in enumerate(self.tables):
            arr[k] = self._shmodel(t)
        L = _lib.lib()
        self._ctx = ctypes.c_void_p()
        rc = L.sh_create(ctypes.byref(self._ctx), len(self.tables), arr, int(device))
        self._keep = []
        if rc != 0:
            raise SweepError("sh_create failed (%d)" % rc)
        self.nchains = len(self.tables)

    def _shmodel(self, t):
        hs, lat = t["hs"], t["lat"]

        def keep(a, dtype):
            a = _c(a, dtype)
            self._keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32 if dtype == np.int32 else ctypes.c_double))

        cz = lambda z: np.stack([np.real(z), np.imag(z)], axis=-1).reshape(-1)
        m = t["model"]
        s = _lib.ShModel()
        s.ndim, s.lq, s.ltrot, s.nwrap, s.nst = t["ndim"], t["lq"], t["ltrot"], t["nwrap"], t["nst"]
        s.iwrap_nt = keep(t["iwrap_nt"], np.int32)
        s.wrap_step = keep(t["wrap_step"].reshape(-1), np.int32)
        s.lwrapT, s.lwrapu, s.kmode = t["lwrapT"], t["lwrapu"], t["kmode"]
        s.nn_nf, s.nn_lf = lat["nn_nf"], lat["nn_lf"]
        s.bonds = keep(lat["bonds"].reshape(-1), np.int32)
        s.urt = keep(cz(t["urt"]), np.float64)
        s.urtm1 = keep(cz(t["urtm1"]), np.float64)
        s.lcomp, s.nflr = hs["lcomp"], m.nflr
        s.bmat = keep(cz(hs["bmat"]), np.float64)
        s.bmat_inv = keep(cz(hs["bmat_inv"]), np.float64)
        s.phase_arg = keep(hs["phase_arg"], np.float64)
        s.delta_bmat = keep(cz(hs["delta_bmat"].reshape(-1, order="F")), np.float64)
        s.phase_ratio = keep(cz(hs["phase_ratio"].reshape(-1, order="F")), np.float64)
        s.gaml = keep(hs["gaml"], np.float64)
        s.phase_from_conf = t["phase_from_conf"]
        s.stag = keep(lat["stag"], np.float64)
        s.update, s.nublock, s.seed = t["update"], t["nublock"], int(np.int32(m.seed))
        s.kin_weight = lat["kin_weight"]
        s.fft_dims[:] = [int(d) for d in t["fft_dims"]]
        if t["v"] is not None:
            s.vtab = keep(cz(t["v"]["vt"].reshape(-1)), np.float64)
            s.vdelta = keep(cz(t["v"]["vd"].reshape(-1)), np.float64)
        if t["plq"] is not None:
            pq = t["plq"]
            s.zplq, s.lqp = pq["z"], pq["lqp"]
            s.plq_cord = keep(pq["cord"].reshape(-1), np.int32)
            # z x z blocks column-major: the transpose of each block in C order
            s.plq_b = keep(cz(pq["b"].transpose(0, 2, 1).reshape(-1)), np.float64)
            s.plq_bi = keep(cz(pq["bi"].transpose(0, 2, 1).reshape(-1)), np.float64)
            s.plq_delta = keep(cz(pq["delta"].transpose(0, 1, 3, 2).reshape(-1)), np.float64)
            s.plq_pratio = keep(cz(pq["pratio"].reshape(-1)), np.float64)
            s.plq_parg = keep(pq["parg"], np.float64)
            s.plq_gaml = keep(pq["gaml"], np.float64)
        return s

    # -- lifecycle ---------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None) and self._ctx.value:
            _lib.lib().sh_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = _lib.lib().sh_last_error(self._ctx)
            raise SweepError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    # -- configuration -----------------------------------------------------------------------
    def init_conf(self):
        self._check(_lib.lib().sh_init_conf(self._ctx), "init_conf")

    def get_conf(self, chain=0):
        t = self.tables[chain]
        c = np.zeros((t["ltrot"], t["ndim"]), np.int32)
        self._check(_lib.lib().sh_get_conf(self._ctx, chain, c.ctypes.data_as(_lib.ip)), "get_conf")
        return c

    def set_conf(self, conf, chain=0):
        t = self.tables[chain]
        c = _c(conf, np.int32).reshape(t["ltrot"], t["ndim"])
        self._check(_lib.lib().sh_set_conf(self._ctx, chain, c.ctypes.data_as(_lib.ip)), "set_conf")

    @property
    def has_vconf(self):
        return self.tables[0]["v"] is not None

    @property
    def has_pconf(self):
        return self.tables[0]["plq"] is not None

    def get_confv(self, chain=0):
        """the t-V bond field [ltrot][nn_nf][ndim] (entries i < nn_lf are the bonds)"""
        t = self.tables[chain]
        c = np.zeros((t["ltrot"], t["lat"]["nn_nf"], t["ndim"]), np.int32)
        self._check(_lib.lib().sh_get_confv(self._ctx, chain, c.ctypes.data_as(_lib.ip)), "get_confv")
        return c

    def set_confv(self, conf, chain=0):
        t = self.tables[chain]
        c = _c(conf, np.int32).reshape(t["ltrot"], t["lat"]["nn_nf"], t["ndim"])
        self._check(_lib.lib().sh_set_confv(self._ctx, chain, c.ctypes.data_as(_lib.ip)), "set_confv")

    def vacc(self, chain=0):
        """(accepted, tried) bond-field flips"""
        a = np.zeros(2)
        self._check(_lib.lib().sh_get_vacc(self._ctx, chain, a.ctypes.data_as(_lib.dp)), "get_vacc")
        return a

    def get_confp(self, chain=0):
        """the plaquette field [ltrot][lqp], values 1..4"""
        t = self.tables[chain]
        c = np.zeros((t["ltrot"], t["plq"]["lqp"]), np.int32)
        self._check(_lib.lib().sh_get_confp(self._ctx, chain, c.ctypes.data_as(_lib.ip)), "get_confp")
        return c

    def set_confp(self, conf, chain=0):
        t = self.tables[chain]
        c = _c(conf, np.int32).reshape(t["ltrot"], t["plq"]["lqp"])
        self._check(_lib.lib().sh_set_confp(self._ctx, chain, c.ctypes.data_as(_lib.ip)), "set_confp")

    def pacc(self, chain=0):
        """(accepted, tried) plaquette-field flips"""
        a = np.zeros(2)
        self._check(_lib.lib().sh_get_pacc(self._ctx, chain, a.ctypes.data_as(_lib.dp)), "get_pacc")
        return a

    def rng_state(self, chain=0):
        """The chain's generator state (SH_RNG_WORDS uint32: SFMT19937 words, read index, block
        proposal state)."""
        s = np.zeros(_lib.SH_RNG_WORDS, np.uint32)
        self._check(_lib.lib().sh_get_rng_state(self._ctx, chain, s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
                    "get_rng_state")
        return s

    def set_rng_state(self, chain, state):
        s = _c(state, np.uint32)
        if s.size != _lib.SH_RNG_WORDS:
            raise ValueError("a generator state has %d words" % _lib.SH_RNG_WORDS)
        self._check(_lib.lib().sh_set_rng_state(self._ctx, chain, s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
                    "set_rng_state")

    def output_conf(self, path="conf_u.out"):
        """u0conf%output_conf: every chain's field to `path` in the reference's text format."""
        from .confio import write_conf_u
        write_conf_u(path, [self.get_conf(c) for c in range(self.nchains)])

    def input_conf(self, path="conf_u.in"):
        """u0conf%input_conf: the fields from `path`, or fresh ones when it is missing or starts
        with 0.  Returns lwarmup (True for a fresh start)."""
        from .confio import read_conf_u
        t = self.tables[0]
        lw, confs = read_conf_u(path, self.nchains, t["ndim"], t["ltrot"])
        if confs is None:
            self.init_conf()
        else:
            for c, conf in enumerate(confs):
                self.set_conf(conf, c)
        return lw

    def output_plqconf(self, path="conf_plqu.out"):
        """hconf%output_plqconf: every chain's plaquette field to `path` in the reference's format."""
        from .confio import write_conf_plqu
        write_conf_plqu(path, [self.get_confp(c) for c in range(self.nchains)])

    def input_plqconf(self, path="conf_plqu.in"):
        """hconf%input_plqconf: the plaquette fields from `path`; returns lwarmup.  A missing file
        or a leading 0 keeps the fields sh_init_conf drew (the library draws every field at once,
        the plaquette field first, as dqmc_ft_ctrl.f90:19-35 does)."""
        from .confio import read_conf_plqu
        t = self.tables[0]
        lw, confs = read_conf_plqu(path, self.nchains, t["plq"]["lqp"], t["ltrot"])
        if confs is not None:
            for c, conf in enumerate(confs):
                self.set_confp(conf, c)
        return lw

    # -- sweeps ------------------------------------------------------------------------------
    def sweep_start_0b(self):
        self._check(_lib.lib().sh_sweep_start_0b(self._ctx), "sweep_start_0b")

    def sweep_b0(self, lupdate=True, lmeasure_equaltime=True):
        self._check(_lib.lib().sh_sweep_b0(self._ctx, int(lupdate), int(lmeasure_equaltime)), "sweep_b0")

    def sweep_0b(self, lupdate=True, lmeasure_equaltime=True, lmeasure_dyn=False):
        self._check(_lib.lib().sh_sweep_0b(self._ctx, int(lupdate), int(lmeasure_equaltime), int(lmeasure_dyn)),
                    "sweep_0b")

    def sweep(self, lupdate=True, lmeasure=True):
        """One reference sweep: a b0 pass then a 0b pass (src/dqmc_ft_ctrl.f90:90-115)."""
        self.sweep_b0(lupdate, lmeasure)
        self.sweep_0b(lupdate, lmeasure)

    # -- time-displaced measurement (ltau) ---------------------------------------------------
    def set_dyn(self, ntdm=None):
        """Switch on G(t,0), G(0,t) and the znn_tau bins for t = 0 .. ntdm slices (default ltrot);
        sweep_0b(lupdate=False, lmeasure_dyn=True) then runs the dynamic pass."""
        for c, t in enumerate(self.tables):
            l1, l2, l3 = t["lat"]["dims"]
            sites = _c(t["lat"]["cells"], np.int32)
            nd = t["ltrot"] if ntdm is None else int(ntdm)
            self._check(_lib.lib().sh_set_dyn(self._ctx, c, nd, l1, l2, l3, sites.ctypes.data_as(_lib.ip)), "set_dyn")
        self.ntdm = self.tables[0]["ltrot"] if ntdm is None else int(ntdm)

    # -- correlation bins of the equal-time measurement --------------------------------------
    def set_eqt(self, theta=0.0):
        """Switch on zcpcm, zspsm, znn, pair_onsite, the bond correlations and the plaquette and
        zint_v energy terms at every measured slice (theta in radians: model_para's theta * pi)."""
        for c, t in enumerate(self.tables):
            lat = t["lat"]
            l1, l2, l3 = lat["dims"]
            sites, cells, plq = _c(lat["cells"], np.int32), _c(lat["eqt_cells"], np.int32), _c(lat["plq"], np.int32)
            self._check(_lib.lib().sh_set_eqt(self._ctx, c, l1, l2, l3, sites.ctypes.data_as(_lib.ip),
                                              cells.ctypes.data_as(_lib.ip), plq.shape[1],
                                              plq.ctypes.data_as(_lib.ip), float(theta)), "set_eqt")

    def eqt(self, chain=0):
        """the flat correlation bins of a chain, 20 lq + 8 complex (binio.unpack_eqt splits them)"""
        lq = self.tables[chain]["lq"]
        b = np.zeros(20 * lq + 8, np.complex128)
        self._check(_lib.lib().sh_get_eqt(self._ctx, chain, b.ctypes.data_as(_lib.dp)), "get_eqt")
        return b

    def dyn_obs(self, chain=0):
        """(znn_tau bins [ntdm+1][orbital j][orbital i][lq], {xmax_dyn, count})"""
        lq = self.tables[chain]["lq"]
        z = np.zeros((self.ntdm + 1, 2, 2, lq), np.complex128)
        e = np.zeros(2)
        self._check(_lib.lib().sh_get_dyn_obs(self._ctx, chain, z.ctypes.data_as(_lib.dp), e.ctypes.data_as(_lib.dp)),
                    "get_dyn_obs")
        return z, {"xmax_dyn": e[0], "count": e[1]}

    def reset_dyn_obs(self):
        self._check(_lib.lib().sh_reset_dyn_obs(self._ctx), "reset_dyn_obs")

    def gtau(self, chain=0):
        """(G(t,0), G(0,t), G(0,0)) after the last dynamic pass"""
        n = self.tables[chain]["ndim"]
        m = [np.zeros((n, n), np.complex128, order="F") for _ in range(3)]
        self._check(_lib.lib().sh_get_gtau(self._ctx, chain, *[x.ctypes.data_as(_lib.dp) for x in m]), "get_gtau")
        return tuple(m)

    def push_stage(self):
        self._check(_lib.lib().sh_push_stage(self._ctx), "push_stage")

    def pop_stage(self):
        self._check(_lib.lib().sh_pop_stage(self._ctx), "pop_stage")

    def synchronize(self):
        self._check(_lib.lib().sh_synchronize(self._ctx), "synchronize")

    def set_graphs(self, enable=True):
        """Replay every kind of pass from a captured hipGraph (one host launch per pass)."""
        self._check(_lib.lib().sh_set_graphs(self._ctx, int(bool(enable))), "set_graphs")

    def autotune(self):
        """Pick the delay depth / submatrix block (and the DFT kernels' tile) by timing them on
        the chains' current state, which is left unchanged; call after sweep_start_0b."""
        out = np.zeros(2, np.int32)
        self._check(_lib.lib().sh_autotune(self._ctx, out.ctypes.data_as(_lib.ip)), "autotune")
        for t in self.tables:
            t["nublock"] = int(out[0])
        return {"nublock": int(out[0]), "fft_vectors_per_workgroup": int(out[1])}

    # -- state readback ------------------------------------------------------------------------
    def G(self, chain=0):
        n = self.tables[chain]["ndim"]
        g = np.zeros((n, n), np.complex128, order="F")
        self._check(_lib.lib().sh_get_g(self._ctx, chain, g.ctypes.data_as(_lib.dp)), "get_g")
        return g

    def diag(self, chain=0):
        d = np.zeros(10)
        self._check(_lib.lib().sh_get_diag(self._ctx, chain, d.ctypes.data_as(_lib.dp)), "get_diag")
        return {"max_wrap_error": d[0], "max_phasediff": d[1], "phase": complex(d[2], d[3]),
                "logweightf": complex(d[4], d[5]), "accepted": d[6], "tried": d[7], "nchecks": d[8],
                "sumlog10err": d[9]}

    def reset_diag(self):
        self._check(_lib.lib().sh_reset_diag(self._ctx), "reset_diag")

    def obs(self, chain=0):
        o = np.zeros(11)
        self._check(_lib.lib().sh_get_obs(self._ctx, chain, o.ctypes.data_as(_lib.dp)), "get_obs")
        z = o[1::2] + 1j * o[2::2]
        return {"nobs": o[0], "sign": z[0], "density": z[1], "ekin": z[2], "dblocc": z[3], "s_stag": z[4]}

    def obs_all(self, reset=True):
        """Observable bin sums of every chain (array [nchains, 5] complex) and nobs; resets them."""
        o = np.zeros((self.nchains, 11))
        self._check(_lib.lib().sh_get_obs_all(self._ctx, o.ctypes.data_as(_lib.dp), int(reset)), "get_obs_all")
        return o[:, 1::2] + 1j * o[:, 2::2], o[:, 0]

    def reset_obs(self):
        self._check(_lib.lib().sh_reset_obs(self._ctx), "reset_obs")

    def set_trace(self, cap, chain=0):
        self._check(_lib.lib().sh_set_trace(self._ctx, chain, int(cap)), "set_trace")
        self._trace_cap = cap

    def trace(self, chain=0):
        n = self.tables[chain]["ndim"]
        cnt = ctypes.c_int32(0)
        buf = np.zeros((max(self._trace_cap, 1), n, n), np.complex128)
        self._check(_lib.lib().sh_get_trace(self._ctx, chain, buf.ctypes.data_as(_lib.dp), ctypes.byref(cnt)),
                    "get_trace")
        # snapshots are column-major matrices
        return np.array([b.reshape(n, n).T for b in buf[:cnt.value]])

    def timers(self, enable=True):
        """Device ms per kernel class since the last call, and launch counts."""
        t = np.zeros(2 * len(TIMER_GROUPS))
        self._check(_lib.lib().sh_timers(self._ctx, int(enable), t.ctypes.data_as(_lib.dp)), "timers")
        g = len(TIMER_GROUPS)
        return {n: (t[k], int(t[g + k])) for k, n in enumerate(TIMER_GROUPS)}

    # -- learned block proposals (include/sunhip_learn.h) ----------------------------------
    def set_neighbours(self):
        """Hand the nearest-neighbour lists of every chain's lattice to the library (features
        and block proposals need them)."""
        z = {t["lat"]["nnlist"].shape[1] for t in self.tables}
        if len(z) != 1:
            raise SweepError("all chains need the same coordination number")
        nbr = _c(np.concatenate([t["lat"]["nnlist"].reshape(-1) for t in self.tables]), np.int32)
        self._check(_lib.lib().sh_set_neighbours(self._ctx, z.pop(), nbr.ctypes.data_as(_lib.ip)), "set_neighbours")

    def features(self):
        """[nchains, 5]: the four field features f0..f3 and E/Ns of each chain's current field."""
        out = np.zeros((self.nchains, 5))
        self._check(_lib.lib().sh_features(self._ctx, out.ctypes.data_as(_lib.dp)), "features")
        return out

    def set_proposal(self, net, nprop=64, local_updates=True):
        """Block proposals from the effective action `net` (None disables them)."""
        h = None if net is None else net._net
        self._check(_lib.lib().sh_set_proposal(self._ctx, h, int(nprop), int(local_updates)), "set_proposal")

    def proposal_stats(self, chain=0):
        s = np.zeros(5)
        self._check(_lib.lib().sh_get_proposal_stats(self._ctx, chain, s.ctypes.data_as(_lib.dp)), "proposal_stats")
        return {"moves": s[0], "accepted": s[1], "spins_proposed": s[2], "spins_accepted": s[3],
                "inner_accepted": s[4]}

    def reset_proposal_stats(self):
        self._check(_lib.lib().sh_reset_proposal_stats(self._ctx), "reset_proposal_stats")
