"""ctypes binding of libsunhip.so (include/sunhip.h).  Loading fails loudly when the library
has not been built; there is no CPU fallback."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsunhip.so")
_LIB = None

ip = ctypes.POINTER(ctypes.c_int32)
dp = ctypes.POINTER(ctypes.c_double)


class ShModel(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int), ("lq", ctypes.c_int),
        ("ltrot", ctypes.c_int), ("nwrap", ctypes.c_int), ("nst", ctypes.c_int),
        ("iwrap_nt", ip), ("wrap_step", ip),
        ("lwrapT", ctypes.c_int), ("lwrapu", ctypes.c_int), ("kmode", ctypes.c_int),
        ("nn_nf", ctypes.c_int), ("nn_lf", ctypes.c_int), ("bonds", ip),
        ("urt", dp), ("urtm1", dp),
        ("lcomp", ctypes.c_int), ("nflr", ctypes.c_int),
        ("bmat", dp), ("bmat_inv", dp), ("phase_arg", dp), ("delta_bmat", dp), ("phase_ratio", dp),
        ("gaml", dp), ("phase_from_conf", ctypes.c_int), ("stag", dp),
        ("update", ctypes.c_int), ("nublock", ctypes.c_int), ("seed", ctypes.c_int32),
        ("kin_weight", ctypes.c_double), ("vtab", dp), ("vdelta", dp),
        ("fft_dims", ctypes.c_int * 3),
        ("zplq", ctypes.c_int), ("lqp", ctypes.c_int), ("plq_cord", ip),
        ("plq_b", dp), ("plq_bi", dp), ("plq_delta", dp), ("plq_pratio", dp), ("plq_parg", dp), ("plq_gaml", dp),
    ]


EXPORTS = [
    "sh_create", "sh_destroy", "sh_last_error", "sh_init_conf", "sh_set_conf", "sh_get_conf",
    "sh_sweep_start_0b", "sh_sweep_b0", "sh_sweep_0b", "sh_push_stage", "sh_pop_stage",
    "sh_get_g", "sh_get_diag", "sh_reset_diag", "sh_get_obs", "sh_get_obs_all", "sh_reset_obs", "sh_set_trace",
    "sh_get_trace", "sh_timers", "sh_synchronize", "sh_set_graphs", "sh_autotune", "sh_get_rng_state", "sh_set_rng_state", "sh_set_dyn", "sh_set_eqt", "sh_get_eqt", "sh_set_confv", "sh_get_confv", "sh_get_vacc",
    "sh_set_confp", "sh_get_confp", "sh_get_pacc",
    "sh_get_dyn_obs", "sh_reset_dyn_obs", "sh_get_gtau",
]

SH_RNG_WORDS = 627     # include/sunhip.h
SHP_RNG_WORDS = 625    # include/sunhip_proj.h

# include/sunhip_proj.h
EXPORTS_PROJ = [
    "shp_create", "shp_destroy", "shp_last_error", "shp_init_conf", "shp_set_conf", "shp_get_conf",
    "shp_draw_uniforms", "shp_set_proj", "shp_sweep_start_0b", "shp_sweep_b0", "shp_sweep_0b", "shp_get_g",
    "shp_get_state", "shp_get_diag", "shp_reset_diag", "shp_get_obs", "shp_reset_obs", "shp_set_trace",
    "shp_get_trace", "shp_timers", "shp_synchronize", "shp_set_uconf", "shp_get_uconf", "shp_get_rng_state",
    "shp_set_rng_state", "shp_set_eqt", "shp_get_eqt",
]


# include/sunhip_learn.h
EXPORTS_LEARN = [
    "shl_create", "shl_destroy", "shl_last_error", "shl_get_weights", "shl_set_weights", "shl_set_norm",
    "shl_get_norm", "shl_add_samples", "shl_push_samples", "shl_num_samples", "shl_get_samples",
    "shl_clear_samples", "shl_train", "shl_loss_grad", "shl_predict", "shl_loss_grad_noalloc", "sh_set_neighbours",
    "sh_features",
    "sh_set_proposal", "sh_get_proposal_stats", "sh_reset_proposal_stats",
]


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError("libsunhip.so is not built: run `python -m sunhubbardmy_amd.build` (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    L.sh_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.POINTER(ShModel), ctypes.c_int]
    L.sh_destroy.argtypes = [P]
    L.sh_last_error.argtypes = [P]
    L.sh_last_error.restype = ctypes.c_char_p
    L.sh_init_conf.argtypes = [P]
    L.sh_set_conf.argtypes = [P, ctypes.c_int, ip]
    L.sh_get_conf.argtypes = [P, ctypes.c_int, ip]
    L.sh_sweep_start_0b.argtypes = [P]
    L.sh_sweep_b0.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.sh_sweep_0b.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.sh_push_stage.argtypes = [P]
    L.sh_pop_stage.argtypes = [P]
    L.sh_get_g.argtypes = [P, ctypes.c_int, dp]
    L.sh_get_diag.argtypes = [P, ctypes.c_int, dp]
    L.sh_reset_diag.argtypes = [P]
    L.sh_get_obs.argtypes = [P, ctypes.c_int, dp]
    L.sh_get_obs_all.argtypes = [P, dp, ctypes.c_int]
    L.sh_reset_obs.argtypes = [P]
    L.sh_set_trace.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.sh_get_trace.argtypes = [P, ctypes.c_int, dp, ip]
    L.sh_timers.argtypes = [P, ctypes.c_int, dp]
    L.sh_synchronize.argtypes = [P]
    L.sh_set_graphs.argtypes = [P, ctypes.c_int]
    L.sh_autotune.argtypes = [P, ip]
    u32p = ctypes.POINTER(ctypes.c_uint32)
    L.sh_get_rng_state.argtypes = [P, ctypes.c_int, u32p]
    L.sh_set_rng_state.argtypes = [P, ctypes.c_int, u32p]
    L.sh_set_dyn.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip]
    L.sh_set_eqt.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ip, ctypes.c_int, ip,
                             ctypes.c_double]
    L.sh_get_eqt.argtypes = [P, ctypes.c_int, dp]
    L.sh_set_confv.argtypes = [P, ctypes.c_int, ip]
    L.sh_get_confv.argtypes = [P, ctypes.c_int, ip]
    L.sh_get_vacc.argtypes = [P, ctypes.c_int, dp]
    L.sh_set_confp.argtypes = [P, ctypes.c_int, ip]
    L.sh_get_confp.argtypes = [P, ctypes.c_int, ip]
    L.sh_get_pacc.argtypes = [P, ctypes.c_int, dp]
    L.sh_get_dyn_obs.argtypes = [P, ctypes.c_int, dp, dp]
    L.sh_reset_dyn_obs.argtypes = [P]
    L.sh_get_gtau.argtypes = [P, ctypes.c_int, dp, dp, dp]
    for name in EXPORTS:
        if name != "sh_last_error":
            getattr(L, name).restype = ctypes.c_int
    L.shp_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, P, ctypes.c_int]
    L.shp_destroy.argtypes = [P]
    L.shp_last_error.argtypes = [P]
    L.shp_last_error.restype = ctypes.c_char_p
    L.shp_init_conf.argtypes = [P]
    L.shp_set_conf.argtypes = [P, ctypes.c_int, ip]
    L.shp_set_uconf.argtypes = [P, ctypes.c_int, ip]
    L.shp_get_uconf.argtypes = [P, ctypes.c_int, ip]
    L.shp_get_conf.argtypes = [P, ctypes.c_int, ip]
    L.shp_draw_uniforms.argtypes = [P, ctypes.c_int, ctypes.c_int, dp]
    L.shp_set_proj.argtypes = [P, ctypes.c_int, dp]
    L.shp_sweep_start_0b.argtypes = [P]
    L.shp_sweep_b0.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.shp_sweep_0b.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.shp_get_g.argtypes = [P, ctypes.c_int, dp]
    L.shp_get_state.argtypes = [P, ctypes.c_int, dp, dp, dp]
    L.shp_get_diag.argtypes = [P, ctypes.c_int, dp]
    L.shp_reset_diag.argtypes = [P]
    L.shp_get_obs.argtypes = [P, ctypes.c_int, dp]
    L.shp_reset_obs.argtypes = [P]
    L.shp_set_trace.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.shp_get_trace.argtypes = [P, ctypes.c_int, dp, ip]
    L.shp_timers.argtypes = [P, ctypes.c_int, dp]
    L.shp_synchronize.argtypes = [P]
    L.shp_get_rng_state.argtypes = [P, ctypes.c_int, u32p]
    L.shp_set_eqt.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ip, ctypes.c_int, ip,
                              ctypes.c_double]
    L.shp_get_eqt.argtypes = [P, ctypes.c_int, dp]
    L.shp_set_rng_state.argtypes = [P, ctypes.c_int, u32p]
    for name in EXPORTS_PROJ:
        if name != "shp_last_error":
            getattr(L, name).restype = ctypes.c_int
    fp = ctypes.POINTER(ctypes.c_float)
    L.shl_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.c_int32, ctypes.c_int]
    L.shl_destroy.argtypes = [P]
    L.shl_last_error.argtypes = [P]
    L.shl_last_error.restype = ctypes.c_char_p
    L.shl_get_weights.argtypes = [P, fp]
    L.shl_set_weights.argtypes = [P, fp]
    L.shl_set_norm.argtypes = [P, dp]
    L.shl_get_norm.argtypes = [P, dp]
    L.shl_add_samples.argtypes = [P, P]
    L.shl_push_samples.argtypes = [P, dp, ctypes.c_int]
    L.shl_num_samples.argtypes = [P]
    L.shl_get_samples.argtypes = [P, dp, ctypes.c_int, ctypes.c_int]
    L.shl_clear_samples.argtypes = [P]
    L.shl_train.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp]
    L.shl_loss_grad.argtypes = [P, dp, ctypes.c_int, dp, fp]
    L.shl_predict.argtypes = [P, dp, ctypes.c_int, dp]
    # device pointers (x, loss, grad) as integers
    L.shl_loss_grad_noalloc.argtypes = [P, P, ctypes.c_int, P, P]
    L.sh_set_neighbours.argtypes = [P, ctypes.c_int, ip]
    L.sh_features.argtypes = [P, dp]
    L.sh_set_proposal.argtypes = [P, P, ctypes.c_int, ctypes.c_int]
    L.sh_get_proposal_stats.argtypes = [P, ctypes.c_int, dp]
    L.sh_reset_proposal_stats.argtypes = [P]
    for name in EXPORTS_LEARN:
        if name != "shl_last_error":
            getattr(L, name).restype = ctypes.c_int
    _LIB = L
    return L


def testing_lib():
    """libsunhip.so with the test hooks of include/sunhip_testing.h bound."""
    L = lib()
    if not hasattr(L, "_dbg_bound"):
        P = ctypes.c_void_p
        L.sh_dbg_put.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.sh_dbg_get.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.sh_dbg_op.argtypes = [P, ctypes.c_int, ip]
        for n in ("sh_dbg_put", "sh_dbg_get", "sh_dbg_op"):
            getattr(L, n).restype = ctypes.c_int
        L._dbg_bound = True
    return L
