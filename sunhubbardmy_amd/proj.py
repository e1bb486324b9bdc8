"""ProjWalkers: a batch of projective (ground-state) DQMC chains on one GPU: the t-V model's
bond field (rv != 0) and/or the Hubbard model's on-site field (rhub != 0).

Host mirror of the reference's projective interface (src/dqmc_proj_ctrl.f90:8-129 calls
these): ``allocate_core`` -> ``ProjWalkers(models)``, ``v0conf%input_conf`` -> ``init_conf``,
``dqmc_set_proj`` -> ``set_trial`` (or ``set_proj``), ``dqmc_proj_sweep_start_0b`` ->
``sweep_start_0b``, ``dqmc_proj_sweep_b0 / _0b`` -> ``sweep_b0 / sweep_0b``.  Tables restate
``v0conf%set_conf`` (src/dqmc_config_v/dqmc_set_conf.f90:14-77, lprojv = .false.),
``u0conf%set_conf`` (src/dqmc_config_u/dqmc_set_conf.f90, lproju = .false., via
``model.hs_tables``), ``set_trial_h0`` (src/dqmc_config_h0/set_trial_h0.f90:1-29, expar = 1) and ``make_tables``
(src/model_para.f90:376-453).  All compute goes through libsunhip.so.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .model import KMODES, Model, default_nublock, hs_tables, kinetic_tables, lattice_tables, schedule, vconf_tables
from .walkers import TIMER_GROUPS, SweepError


@dataclass
class ProjModel:
    lattice: str = "honeycomb"
    la: int = 2
    lb: int = 2
    beta: float = 2.0          # projection length (the reference's beta), ltrot = nint(beta/dtau)
    dtau: float = 0.1
    rt: float = 1.0
    mu: float = 0.0
    rv: float = 1.2
    nu: float = 0.0
    rndness: float = 1e-4
    nflr: int = 1
    kmode: str = "dense"
    nwrap: int = 10
    nustock: int = 0           # 0: the reference's submatrix default (nustock = nublock)
    mid_len: int = 1           # obs_eqt_mid_len
    wrap_check: bool = True
    seed: int = 1909706392
    rhub: float = 0.0          # on-site U (the uconf field when != 0)
    exact_hs: bool = True      # EXACTHS tables of the on-site field
    update: str = "submatrix_lr"  # update of both fields: submatrix_lr, fast, delay, submatrix_g
    gblock: int = 0            # depth of the G-based forms (0: the reference's nublock default)


PUPDATES = {"submatrix_lr": 0, "fast": 1, "delay": 2, "submatrix_g": 3}


def make_proj_tables(m):
    if KMODES.get(m.kmode) not in (0, 1):
        raise ValueError("the projective sweep applies the kinetic term as 'checkerboard' or 'dense'")
    if m.update not in PUPDATES:
        raise ValueError("update must be one of %s" % sorted(PUPDATES))
    if PUPDATES[m.update] >= 2 and (m.rhub != 0.0) == (m.rv != 0.0):
        raise ValueError("the G-based updates (delay, submatrix_g) need one field alone (rhub or rv)")
    lat = lattice_tables(m)
    n = lat["nsites"]
    ltrot = int(round(m.beta / m.dtau))
    nst, nw, iw, ws = schedule(ltrot, m.nwrap)
    urt, urtm1 = kinetic_tables(m, lat)
    ns = m.nustock if m.nustock > 0 else default_nublock(n, "submatrix")
    ns = min(ns, 64)
    ns -= ns % 2
    u = None
    if m.rhub != 0.0:
        u = hs_tables(Model(dtau=m.dtau, rhub=m.rhub, nu=m.nu, nflr=m.nflr, exact_hs=m.exact_hs))
    return {"model": m, "lat": lat, "ndim": n, "ne": n // 2 + int(m.nu * n / 8.0), "lq": lat["lq"],
            "ltrot": ltrot, "nst": nst, "nwrap": nw, "iwrap_nt": iw, "wrap_step": ws, "urt": urt,
            "urtm1": urtm1, "v": vconf_tables(m) if m.rv != 0.0 else None, "u": u, "nustock": max(ns, 2),
            "kmode": KMODES[m.kmode], "update": PUPDATES[m.update],
            "gblock": min(64, m.gblock if m.gblock > 0 else default_nublock(n, "delay" if m.update == "delay" else "submatrix"))}


def trial_hamiltonian(m, lat, u):
    """set_trial_h0: hopping -rt + rndness (u - 1) per (A site, family), -mu on the diagonal."""
    n = lat["nsites"]
    h = np.zeros((n, n))
    k = 0
    for i in range(lat["nn_lf"]):
        for f in range(lat["nn_nf"]):
            i0, i1 = lat["bonds"][f, i]
            z1 = -m.rt + m.rndness * (u[k] - 1.0)
            k += 1
            h[i0, i1] += z1
            h[i1, i0] += z1
    h -= m.mu * np.eye(n)
    return h


class ShpModel(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int), ("ne", ctypes.c_int),
        ("ltrot", ctypes.c_int), ("nwrap", ctypes.c_int), ("nst", ctypes.c_int),
        ("iwrap_nt", _lib.ip), ("wrap_step", _lib.ip),
        ("kmode", ctypes.c_int), ("nn_nf", ctypes.c_int), ("nn_lf", ctypes.c_int), ("bonds", _lib.ip),
        ("urt", _lib.dp), ("urtm1", _lib.dp), ("vtab", _lib.dp), ("vdelta", _lib.dp), ("stag", _lib.dp),
        ("lq", ctypes.c_int), ("nflr", ctypes.c_int), ("nustock", ctypes.c_int), ("mid_len", ctypes.c_int),
        ("wrap_check", ctypes.c_int), ("seed", ctypes.c_int32),
        ("ulcomp", ctypes.c_int), ("ubmat", _lib.dp), ("ubmat_inv", _lib.dp), ("udelta", _lib.dp),
        ("uratio", _lib.dp), ("ugaml", _lib.dp), ("uphase_arg", _lib.dp), ("uphase_from_conf", ctypes.c_int),
        ("kin_weight", ctypes.c_double), ("update", ctypes.c_int), ("gblock", ctypes.c_int),
    ]


class ProjWalkers:
    def __init__(self, models, device=0):
        if isinstance(models, ProjModel):
            models = [models]
        self.models = list(models)
        self.tables = [make_proj_tables(m) for m in self.models]
        self._keep = []
        arr = (ShpModel * len(self.tables))()
        for k, t in enumerate(self.tables):
            arr[k] = self._shpmodel(t)
        L = _lib.lib()
        self._ctx = ctypes.c_void_p()
        rc = L.shp_create(ctypes.byref(self._ctx), len(self.tables), ctypes.cast(arr, ctypes.c_void_p), int(device))
        self._keep = []
        if rc != 0:
            raise SweepError("shp_create failed (%d)" % rc)
        self.nchains = len(self.tables)
        t = self.tables[0]
        self.ndim, self.ne, self.ltrot, self.nst = t["ndim"], t["ne"], t["ltrot"], t["nst"]
        self.nn_nf, self.nn_lf = t["lat"]["nn_nf"], t["lat"]["nn_lf"]

    def _shpmodel(self, t):
        lat = t["lat"]

        def keep(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32 if dtype == np.int32 else ctypes.c_double))

        cz = lambda z: np.stack([np.real(z), np.imag(z)], axis=-1).reshape(-1)
        m = t["model"]
        s = ShpModel()
        s.ndim, s.ne, s.ltrot, s.nwrap, s.nst = t["ndim"], t["ne"], t["ltrot"], t["nwrap"], t["nst"]
        s.iwrap_nt = keep(t["iwrap_nt"], np.int32)
        s.wrap_step = keep(t["wrap_step"].reshape(-1), np.int32)
        s.kmode, s.nn_nf, s.nn_lf = t["kmode"], lat["nn_nf"], lat["nn_lf"]
        s.bonds = keep(lat["bonds"].reshape(-1), np.int32)
        s.urt = keep(cz(t["urt"]), np.float64)
        s.urtm1 = keep(cz(t["urtm1"]), np.float64)
        if t["v"] is not None:
            s.vtab = keep(cz(t["v"]["vt"].reshape(-1)), np.float64)
            s.vdelta = keep(cz(t["v"]["vd"].reshape(-1)), np.float64)
        hs = t["u"]
        if hs is not None:
            s.ulcomp = hs["lcomp"]
            s.ubmat = keep(cz(hs["bmat"]), np.float64)
            s.ubmat_inv = keep(cz(hs["bmat_inv"]), np.float64)
            s.udelta = keep(cz(hs["delta_bmat"].reshape(-1, order="F")), np.float64)
            s.uratio = keep(cz(hs["phase_ratio"].reshape(-1, order="F")), np.float64)
            s.ugaml = keep(hs["gaml"], np.float64)
            s.uphase_arg = keep(hs["phase_arg"], np.float64)
            s.uphase_from_conf = int(m.rhub > 0.0)
        s.stag = keep(lat["stag"], np.float64)
        s.lq, s.nflr, s.nustock, s.mid_len = t["lq"], m.nflr, t["nustock"], m.mid_len
        s.wrap_check, s.seed = int(bool(m.wrap_check)), int(np.int32(m.seed))
        s.kin_weight = lat["kin_weight"]
        s.update, s.gblock = t["update"], t["gblock"]
        return s

    def close(self):
        if getattr(self, "_ctx", None) and self._ctx.value:
            _lib.lib().shp_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = _lib.lib().shp_last_error(self._ctx)
            raise SweepError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    # -- fields and trial state ---------------------------------------------------------------
    def init_conf(self):
        self._check(_lib.lib().shp_init_conf(self._ctx), "init_conf")

    def get_conf(self, chain=0):
        c = np.zeros((self.ltrot, self.nn_nf, self.ndim), np.int32)
        self._check(_lib.lib().shp_get_conf(self._ctx, chain, c.ctypes.data_as(_lib.ip)), "get_conf")
        return c

    def set_conf(self, chain, conf):
        c = np.ascontiguousarray(conf, dtype=np.int32).reshape(self.ltrot, self.nn_nf, self.ndim)
        self._check(_lib.lib().shp_set_conf(self._ctx, chain, c.ctypes.data_as(_lib.ip)), "set_conf")

    def get_uconf(self, chain=0):
        """the on-site field [ltrot, ndim] (models with rhub != 0)"""
        c = np.zeros((self.ltrot, self.ndim), np.int32)
        self._check(_lib.lib().shp_get_uconf(self._ctx, chain, c.ctypes.data_as(_lib.ip)), "get_uconf")
        return c

    def set_uconf(self, chain, conf):
        c = np.ascontiguousarray(conf, dtype=np.int32).reshape(self.ltrot, self.ndim)
        self._check(_lib.lib().shp_set_uconf(self._ctx, chain, c.ctypes.data_as(_lib.ip)), "set_uconf")

    @property
    def has_uconf(self):
        return self.tables[0].get("u") is not None

    @property
    def has_vconf(self):
        return self.tables[0].get("v") is not None

    def set_eqt(self, theta=0.0):
        """correlation bins of the equal-time measurement at the measured slices (see FTWalkers.set_eqt)"""
        for c, t in enumerate(self.tables):
            lat = t["lat"]
            l1, l2, l3 = lat["dims"]
            sites = np.ascontiguousarray(lat["cells"], np.int32)
            cells = np.ascontiguousarray(lat["eqt_cells"], np.int32)
            plq = np.ascontiguousarray(lat["plq"], np.int32)
            self._check(_lib.lib().shp_set_eqt(self._ctx, c, l1, l2, l3, sites.ctypes.data_as(_lib.ip),
                                               cells.ctypes.data_as(_lib.ip), plq.shape[1],
                                               plq.ctypes.data_as(_lib.ip), float(theta)), "set_eqt")

    def eqt(self, chain=0):
        """the flat correlation bins of a chain, 20 lq + 8 complex (binio.unpack_eqt splits them)"""
        b = np.zeros(20 * self.tables[chain]["lq"] + 8, np.complex128)
        self._check(_lib.lib().shp_get_eqt(self._ctx, chain, b.ctypes.data_as(_lib.dp)), "get_eqt")
        return b

    def rng_state(self, chain=0):
        """The chain's SFMT19937 state (SHP_RNG_WORDS uint32: words, read index)."""
        s = np.zeros(_lib.SHP_RNG_WORDS, np.uint32)
        self._check(_lib.lib().shp_get_rng_state(self._ctx, chain, s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
                    "get_rng_state")
        return s

    def set_rng_state(self, chain, state):
        s = np.ascontiguousarray(state, dtype=np.uint32)
        if s.size != _lib.SHP_RNG_WORDS:
            raise ValueError("a generator state has %d words" % _lib.SHP_RNG_WORDS)
        self._check(_lib.lib().shp_set_rng_state(self._ctx, chain, s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
                    "set_rng_state")

    def output_conf(self, path_v="conf_v.out", path_u="conf_u.out"):
        """v0conf%output_conf and u0conf%output_conf for the fields the model has."""
        from .confio import write_conf_u, write_conf_v
        if self.has_vconf:
            write_conf_v(path_v, [self.get_conf(c) for c in range(self.nchains)])
        if self.has_uconf:
            write_conf_u(path_u, [self.get_uconf(c) for c in range(self.nchains)])

    def input_conf(self, path_v="conf_v.in", path_u="conf_u.in"):
        """u0conf%input_conf then v0conf%input_conf: fresh fields (drawn by the library in that
        order) unless every field the model has is in its file.  Returns lwarmup."""
        from .confio import read_conf_u, read_conf_v
        lw_u, cu = read_conf_u(path_u, self.nchains, self.ndim, self.ltrot) if self.has_uconf else (False, None)
        lw_v, cv = read_conf_v(path_v, self.nchains, self.ndim, self.nn_nf, self.ltrot) if self.has_vconf else (False, None)
        if (self.has_uconf and cu is None) or (self.has_vconf and cv is None):
            self.init_conf()
        for c in range(self.nchains):
            if cu is not None:
                self.set_uconf(c, cu[c])
            if cv is not None:
                self.set_conf(c, cv[c])
        return lw_u or lw_v

    def draw_uniforms(self, chain, count):
        out = np.zeros(count)
        self._check(_lib.lib().shp_draw_uniforms(self._ctx, chain, count, out.ctypes.data_as(_lib.dp)),
                    "draw_uniforms")
        return out

    def set_proj(self, chain, proj):
        p = np.asfortranarray(proj, dtype=np.complex128)
        assert p.shape == (self.ndim, self.ne)
        self._check(_lib.lib().shp_set_proj(self._ctx, chain, p.ctypes.data_as(_lib.dp)), "set_proj")

    def set_trial(self):
        """dqmc_set_proj for every chain: the trial Hamiltonian from the chain's next
        nn_lf nn_nf draws, its lowest ne eigenvectors as projL = projR.  Returns the gaps."""
        gaps = []
        self.projs = []
        for c, t in enumerate(self.tables):
            lat = t["lat"]
            u = self.draw_uniforms(c, lat["nn_lf"] * lat["nn_nf"])
            h = trial_hamiltonian(t["model"], lat, u)
            w, v = np.linalg.eigh(h)
            p = v[:, :t["ne"]].astype(np.complex128)
            self.set_proj(c, p)
            self.projs.append(p)
            gaps.append(w[t["ne"]] - w[t["ne"] - 1] if t["ne"] < t["ndim"] else 0.0)
        return gaps

    # -- sweeps ---------------------------------------------------------------------------------
    def sweep_start_0b(self):
        self._check(_lib.lib().shp_sweep_start_0b(self._ctx), "sweep_start_0b")

    def sweep_b0(self, lupdate=True, lmeasure=True):
        self._check(_lib.lib().shp_sweep_b0(self._ctx, int(lupdate), int(lmeasure)), "sweep_b0")

    def sweep_0b(self, lupdate=True, lmeasure=True):
        self._check(_lib.lib().shp_sweep_0b(self._ctx, int(lupdate), int(lmeasure), 0), "sweep_0b")

    def sweep(self, lupdate=True, lmeasure=True):
        self.sweep_b0(lupdate, lmeasure)
        self.sweep_0b(lupdate, lmeasure)

    def synchronize(self):
        self._check(_lib.lib().shp_synchronize(self._ctx), "synchronize")

    # -- state readback --------------------------------------------------------------------------
    def G(self, chain=0):
        g = np.zeros((self.ndim, self.ndim), np.complex128, order="F")
        self._check(_lib.lib().shp_get_g(self._ctx, chain, g.ctypes.data_as(_lib.dp)), "get_g")
        return g

    def state(self, chain=0):
        ul = np.zeros((self.ne, self.ndim), np.complex128, order="F")
        ur = np.zeros((self.ndim, self.ne), np.complex128, order="F")
        inv = np.zeros((self.ne, self.ne), np.complex128, order="F")
        self._check(_lib.lib().shp_get_state(self._ctx, chain, ul.ctypes.data_as(_lib.dp),
                                             ur.ctypes.data_as(_lib.dp), inv.ctypes.data_as(_lib.dp)), "get_state")
        return ul, ur, inv

    def diag(self, chain=0):
        d = np.zeros(10)
        self._check(_lib.lib().shp_get_diag(self._ctx, chain, d.ctypes.data_as(_lib.dp)), "get_diag")
        return {"max_wrap_error": d[0], "max_phasediff": d[1], "phase": complex(d[2], d[3]),
                "logweightf": complex(d[4], d[5]), "accepted": d[6], "tried": d[7], "nchecks": d[8],
                "sumlog10err": d[9]}

    def reset_diag(self):
        self._check(_lib.lib().shp_reset_diag(self._ctx), "reset_diag")

    def obs(self, chain=0):
        o = np.zeros(11)
        self._check(_lib.lib().shp_get_obs(self._ctx, chain, o.ctypes.data_as(_lib.dp)), "get_obs")
        z = o[1::2] + 1j * o[2::2]
        return {"nobs": o[0], "sign": z[0], "density": z[1], "ekin": z[2], "dblocc": z[3], "s_stag": z[4]}

    def reset_obs(self):
        self._check(_lib.lib().shp_reset_obs(self._ctx), "reset_obs")

    def set_trace(self, cap):
        self._check(_lib.lib().shp_set_trace(self._ctx, int(cap > 0), int(cap)), "set_trace")
        self._trace_cap = cap

    def trace(self, chain=0):
        n = self.ndim
        buf = np.zeros((max(self._trace_cap, 1), n, n), np.complex128)
        cnt = ctypes.c_int32()
        self._check(_lib.lib().shp_get_trace(self._ctx, chain, buf.ctypes.data_as(_lib.dp), ctypes.byref(cnt)),
                    "get_trace")
        return np.array([b.T for b in buf[:cnt.value]])

    def timers(self, enable=True):
        t = np.zeros(2 * len(TIMER_GROUPS))
        self._check(_lib.lib().shp_timers(self._ctx, int(enable), t.ctypes.data_as(_lib.dp)), "timers")
        g = len(TIMER_GROUPS)
        return {n: (t[k], int(t[g + k])) for k, n in enumerate(TIMER_GROUPS)}
