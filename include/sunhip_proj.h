/*
 * sunhip_proj.h -- C ABI of libsunhip.so for the projective (ground-state) DQMC sweep: the
 * bond Hubbard-Stratonovich field of the t-V model and/or the on-site field of the Hubbard
 * model, submatrix-LR local updates, QR of the rectangular propagators every nwrap slices.
 *
 * Drop-in for the module procedures of src/dqmc_proj_core.f90 that the reference's
 * src/dqmc_proj_ctrl.f90 calls, and for the vconf and uconf operator plugins they drive
 * (src/dqmc_config_v.f90, src/dqmc_config_u.f90).  Each entry point below names the reference
 * routine it replaces.
 * Conventions as in sunhip.h: int return codes (0 = success, SH_E_*), column-major complex128
 * (re, im interleaved) host buffers owned by the caller and read or written only during the
 * call, one context per host thread, all chains of a context advanced together.
 */
#ifndef SUNHIP_PROJ_H
#define SUNHIP_PROJ_H
#include <stdint.h>
#include "sunhip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct shp_ctx shp_ctx;

/* One chain's model: what make_tables, h0c%set_h0conf, v0conf%set_conf and u0conf%set_conf
 * leave in module state (src/model_para.f90:376-453, src/dqmc_config_h0/dqmc_set_h0conf.f90,
 * src/dqmc_config_v/dqmc_set_conf.f90:1-79, src/dqmc_config_u/dqmc_set_conf.f90:1-176).
 * vtab == NULL: no bond field (rv = 0); ubmat == NULL: no on-site field (rhub = 0). */
typedef struct shp_model {
    int ndim, ne;                 /* sites, particles (ne = ndim/2 + int(nu*ndim/8)) */
    int ltrot, nwrap, nst;        /* slices, stabilization interval and count */
    const int32_t *iwrap_nt;      /* [ltrot+1] */
    const int32_t *wrap_step;     /* [2*nst] (1-based slices) */
    int kmode;                    /* SH_K_CHECKERBOARD or SH_K_DENSE */
    int nn_nf, nn_lf;             /* bond families and bonds per family */
    const int32_t *bonds;         /* [nn_nf][nn_lf][2] 0-based (nnlf_list(i), nnlist(.,nf)) */
    const double *urt, *urtm1;    /* exp(-+dtau K/2): per bond 2x2 column-major, or N x N */
    const double *vtab;           /* 8 complex: (bmat_p, bmat_m, bmat_p_inv, bmat_m_inv) for field 1, 2 */
    const double *vdelta;         /* 4 complex: (delta_bmat_p, delta_bmat_m)(1, is) for field 1, 2 */
    const double *stag;           /* [ndim] +-1 staggering for the S(pi,pi) bin */
    int lq, nflr;
    int nustock;                  /* rows per stock of the submatrix-LR update (even, <= 64) */
    int mid_len;                  /* obs_eqt_mid_len: measured slices around ltrot/2 */
    int wrap_check;               /* 1: the WRAPERROR comparison at every stabilization */
    int32_t seed;                 /* SFMT19937 seed of the chain (src/dqmc_util.f90:68-75) */
    /* the on-site field (u0conf, lproju = .false.) */
    int ulcomp;                   /* components 2 or 4 */
    const double *ubmat;          /* [ulcomp] complex bmat */
    const double *ubmat_inv;      /* [ulcomp] complex */
    const double *udelta;         /* [(ulcomp-1)*ulcomp] complex delta_bmat(iflip, is), column-major */
    const double *uratio;         /* [(ulcomp-1)*ulcomp] complex phase_ratio(iflip, is) */
    const double *ugaml;          /* [ulcomp] gaml */
    const double *uphase_arg;     /* [ulcomp] arg of phase(is), for set_phase */
    int uphase_from_conf;         /* 1 when set_phase multiplies the field's phases (rhub > 0) */
    double kin_weight;            /* factor of the zkint bin: 1 (0 is read as 1), nflr * rt for the
                                     chain (src/dqmc_measure/equaltime_measure1d.f90:76) */
    /* the update algorithm of both fields (src/dqmc_config_u/dqmc_proj_update*.f90,
     * src/dqmc_config_v/dqmc_proj_update*.f90; the reference selects it for both by its build):
     * SHP_U_LR submatrix-LR in stocks of nustock rows (default 0);
     * SHP_U_FAST one site (bond) at a time: rank-1 (rank-2) updates of UR and (UL UR)^-1;
     * SHP_U_DELAY / SHP_U_SUBMATRIX_G updates of G = 1 - UR (UL UR)^-1 UL (the delay form with
     * depth gblock 1..64, or G updated at every accepted site or bond), then UR's flipped rows
     * and a fresh (UL UR)^-1.  The four take the same decisions from the same draws.  The
     * G-based forms need one field alone (on-site or bond).  All chains of a batch share update
     * and gblock. */
    int update, gblock;
} shp_model;

#define SHP_U_LR 0
#define SHP_U_FAST 1
#define SHP_U_DELAY 2
#define SHP_U_SUBMATRIX_G 3

/* allocate_core (src/dqmc_proj_core.f90:23-49) for a batch of chains sharing ndim, ne, ltrot,
 * nwrap, kmode and nustock */
int shp_create(shp_ctx **ctx, int nchains, const shp_model *models, int device);
/* the correlation bins of equaltime_measure at the measured slices, as sh_set_eqt / sh_get_eqt
 * (the projective sweep calls the same equaltime_measure: src/dqmc_proj_core/dqmc_proj_sweep_0b.f90) */
int shp_set_eqt(shp_ctx *ctx, int chain, int l1, int l2, int l3, const int32_t *sites, const int32_t *cells, int zplq,
                const int32_t *plq, double theta);
int shp_get_eqt(shp_ctx *ctx, int chain, double *bins);
int shp_destroy(shp_ctx *ctx);                         /* deallocate_core */
const char *shp_last_error(const shp_ctx *ctx);

/* v0conf%input_conf from scratch (src/dqmc_config_v/dqmc_input_conf.f90:55-66): every chain's
 * field from its own stream, nt outer, site i, family nf innermost */
int shp_init_conf(shp_ctx *ctx);
int shp_set_conf(shp_ctx *ctx, int chain, const int32_t *conf);   /* [ltrot][nn_nf][ndim] */
int shp_get_conf(shp_ctx *ctx, int chain, int32_t *conf);
/* the on-site field u0conf%conf(1:ndim, 1:ltrot) as int32 [ltrot][ndim], values 1..ulcomp;
 * shp_init_conf draws it first (u0conf%input_conf precedes v0conf%input_conf in dqmc_start,
 * src/dqmc_proj_ctrl.f90:27-36) */
int shp_set_uconf(shp_ctx *ctx, int chain, const int32_t *conf);
int shp_get_uconf(shp_ctx *ctx, int chain, int32_t *conf);

/* the chain's SFMT19937 state (lib/m_spring.f90:66-71): 624 words, then the read index 0..624;
 * with the fields it resumes a chain bit-exactly (no reference counterpart: the reference
 * restarts from conf_v.in / conf_u.in with a fresh seed) */
#define SHP_RNG_WORDS 625
int shp_get_rng_state(shp_ctx *ctx, int chain, uint32_t *state);
int shp_set_rng_state(shp_ctx *ctx, int chain, const uint32_t *state);

/* the next n numbers of the chain's stream (set_trial_h0 draws them after the field:
 * src/dqmc_config_h0/set_trial_h0.f90:20) */
int shp_draw_uniforms(shp_ctx *ctx, int chain, int n, double *out);
/* dqmc_set_proj (src/dqmc_proj_core/dqmc_set_proj.f90:58-59): projR = projL = the trial
 * Slater determinant, ndim x ne complex */
int shp_set_proj(shp_ctx *ctx, int chain, const double *proj);

int shp_sweep_start_0b(shp_ctx *ctx);                               /* dqmc_proj_sweep_start_0b */
int shp_sweep_b0(shp_ctx *ctx, int lupdate, int lmeasure);          /* dqmc_proj_sweep_b0 */
int shp_sweep_0b(shp_ctx *ctx, int lupdate, int lmeasure, int ldyn);/* dqmc_proj_sweep_0b; ldyn must be 0 */

/* G = 1 - UR (UL UR)^-1 UL of the current propagators (green_equaltime of
 * src/dqmc_proj_core/green_equaltime.f90 with a fresh inverse), ndim x ndim */
int shp_get_g(shp_ctx *ctx, int chain, double *g);
/* UL (ne x ndim), UR (ndim x ne), ULRINV (ne x ne); any pointer may be NULL */
int shp_get_state(shp_ctx *ctx, int chain, double *ul, double *ur, double *ulrinv);
/* [0]=max_wrap_error [1]=max_phasediff [2,3]=phase [4,5]=logweightf [6]=accepted [7]=tried
 * (main_obs(4)) [8]=wrap checks [9]=sum log10 wrap error */
int shp_get_diag(shp_ctx *ctx, int chain, double *d);
int shp_reset_diag(shp_ctx *ctx);
/* [0]=nobs, then complex bin sums of sign, density/lq, zkint, zint_u, S_stag as sh_get_obs */
int shp_get_obs(shp_ctx *ctx, int chain, double *obs);
int shp_reset_obs(shp_ctx *ctx);
/* keep G after every stabilization, up to cap snapshots per chain */
int shp_set_trace(shp_ctx *ctx, int enable, int cap);
int shp_get_trace(shp_ctx *ctx, int chain, double *buf, int32_t *count);
/* device ms per kernel group since the last call (groups as sh_timers) */
int shp_timers(shp_ctx *ctx, int enable, double *t);
int shp_synchronize(shp_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
