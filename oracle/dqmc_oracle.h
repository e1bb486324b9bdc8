/*
 * dqmc_oracle.h -- CPU restatement of the reference's finite-temperature DQMC sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle for the HIP sweep in
 * sunhubbardmy_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker (or the timed CPU baseline), never as the
 * thing measured or shipped.  The product path never calls into it.
 *
 * Every routine in dqmc_oracle.c cites the reference file:line it restates
 * (paths relative to the YuanyueChen/sunhubbardmy tree).  Storage is column-major
 * complex128 (re, im interleaved), the layout of the reference's complex(dp) arrays.
 */
#ifndef DQMC_ORACLE_H
#define DQMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int lattice;    /* 0 = square (src/lattice.f90:792), 1 = honeycomb (:611), 2 = chain (:240), 3 = cubic (:1278) */
    int la, lb;
    double beta, dtau, rt, mu, rhub, nu;
    int nflr;
    int exact_hs;   /* 1 = EXACTHS tables, 0 = general 4-component (dqmc_set_conf.f90:99-113) */
    int kmode;      /* 0 = checkerboard (BREAKUP_T), 1 = dense exp(-dtau K/2) (also what FFT computes) */
    int update;     /* 0 = fast, 1 = delay, 2 = submatrix */
    int nwrap, nublock;
    int seed;       /* SFMT19937 seed, int32 */
    int lc;         /* third linear size (cubic) */
    double rv;      /* nearest-neighbour V: the bond field of the finite-temperature t-V model (0: off) */
    double rhub_plq;        /* plaquette interaction: the plaquette field (0: off; square, honeycomb) */
    double alpha, theta;    /* its assisting hopping: amplitude, phase in units of pi */
} orc_params;

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const orc_params *p);
void orc_destroy(orc_ctx *c);

/* sizes: out[0]=ndim, [1]=lq, [2]=ltrot, [3]=nst, [4]=lcomp, [5]=nn_nf, [6]=nn_lf, [7]=nublock */
void orc_sizes(const orc_ctx *c, int *out);

/* random initial HS field from the chain's own SFMT stream (dqmc_input_conf.f90:55-63) */
void orc_init_conf(orc_ctx *c);
void orc_get_conf(const orc_ctx *c, int32_t *conf);       /* [ltrot][nsites], values 1..lcomp */
void orc_set_conf(orc_ctx *c, const int32_t *conf);

int orc_sweep_start(orc_ctx *c);                           /* dqmc_ft_sweep_start_0b */
int orc_sweep_b0(orc_ctx *c, int lupdate, int lmeasure);   /* dqmc_ft_sweep_b0 */
int orc_sweep_0b(orc_ctx *c, int lupdate, int lmeasure);   /* dqmc_ft_sweep_0b */

void orc_get_g(const orc_ctx *c, double *g);               /* N*N complex, column-major */
/* diag: [0]=max_wrap_error [1]=max_phasediff [2,3]=phase [4,5]=logweightf
 *       [6]=accepted [7]=tried [8]=nwrap_checks [9]=sum log10 wrap error */
void orc_get_diag(const orc_ctx *c, double *d);
void orc_reset_diag(orc_ctx *c);
/* obs: [0]=nobs, then complex sums (re,im) of: sign, density/lq, zkint, zint_u, S_stag */
void orc_get_obs(const orc_ctx *c, double *o);
void orc_reset_obs(orc_ctx *c);

/* Learned block proposals (the HIP library's block move restated; the reference has none):
 * after the local update of every slice (skipped when local == 0), nprop Metropolis attempts
 * on E_eff = Ns (ymu + ysd g(x)) at uniform sites of the slice (splitmix64 stream seeded from
 * the chain seed), then the flipped set is accepted with the exact determinant ratio.
 * nbr[ndim][z] neighbour lists, w the float network weights, norm[10] = xmu, xsd, ymu, ysd.
 * w == NULL switches proposals off.  stats[5]: moves, accepted, spins proposed, spins accepted,
 * inner attempts accepted. */
void orc_set_proposal(orc_ctx *c, int z, const int32_t *nbr, int nprop, int local, const float *w, const double *norm);
void orc_get_proposal_stats(const orc_ctx *c, double *stats);

/* per-stabilization trace of G (after replacement by the scratch G): cap snapshots */
void orc_set_trace(orc_ctx *c, double *buf, int cap);
/* time-displaced measurement (ltau): green_tau, the 0b pass without updates, znn_tau bins */
int orc_set_dyn(orc_ctx *c, int ntdm);
/* correlation bins of equaltime_measure (zcpcm, zspsm, znn, pair_onsite, zjj, zbb, pair_nn,
 * pair_sn, zn, zj, zb, zqsq, ztsq, ztq, zint_v), accumulated at every measured slice from now on;
 * theta in radians (model_para theta * pi).  Layout of orc_get_eqt: 20 lq + 8 complex, see the
 * source. */
int orc_set_eqt(orc_ctx *c, double theta);
/* the bond field conf(i, nf, nt) of the finite-temperature t-V model, [ltrot][nn_nf][nsites]
 * (entries i < nn_lf are the bonds); acceptance sums (accepted, tried) of its updates */
void orc_get_confv(const orc_ctx *c, int32_t *conf);
void orc_set_confv(orc_ctx *c, const int32_t *conf);
void orc_get_vacc(const orc_ctx *c, double *acc);
/* the plaquette field [ltrot][lqp] (values 1..4), its (accepted, tried) counts and size lqp */
void orc_get_confp(const orc_ctx *c, int32_t *conf);
void orc_set_confp(orc_ctx *c, const int32_t *conf);
void orc_get_pacc(const orc_ctx *c, double *acc);
int orc_lqp(const orc_ctx *c);
void orc_get_eqt(const orc_ctx *c, double *out);
int orc_sweep_0b_dyn(orc_ctx *c, int lmeasure);
void orc_get_gtau(const orc_ctx *c, double *gt0, double *g0t, double *g00);
void orc_get_znn(const orc_ctx *c, double *znn);
void orc_reset_znn(orc_ctx *c);
void orc_get_dyn_error(const orc_ctx *c, double *out);
void orc_set_dyn_trace(orc_ctx *c, double *buf, int cap);
int orc_dyn_trace_count(const orc_ctx *c);
int orc_trace_count(const orc_ctx *c);

/* tables, for cross-checking the product's host-side setup:
 * kind 0: hs (lcomp x [bmat, bmat_inv, phase]) complex -> 3*lcomp complex
 * kind 1: delta_bmat (lcomp-1)*lcomp complex; kind 2: phase_ratio (lcomp-1)*lcomp complex
 * kind 3: gaml (lcomp doubles);  kind 4: K urt (dense N*N or nn_nf*nn_lf*4 complex)
 * kind 5: K urtm1; kind 6: bonds (int32 pairs as doubles, nn_nf*nn_lf*2)
 * kind 7: iwrap_nt (ltrot+1 as doubles); kind 8: wrap_step (2*nst as doubles) */
int orc_get_table(const orc_ctx *c, int kind, double *out);

/* RNG restatement of lib/m_spring.f90 */
void orc_sfmt_u32(int32_t seed, int n, uint32_t *out);
void orc_sfmt_stream(int32_t seed, int n, double *out);
int32_t orc_chain_seed(int32_t chain);    /* abs((c*1981+2008)*951049) in int32 */

/* linear-algebra primitives, exposed for invariant tests */
/* A (m x n) -> Q (m x m), R (m x n), jpvt (1-based), logdetQ (re,im); pivot=1: column pivoting */
void orc_qr(int m, int n, const double *a, double *q, double *r, int32_t *jpvt, double *logdetq, int pivot);
/* inverse with log det via LU (s_inv_logdet_lu_z); returns 0 or the singular column */
int orc_inv_logdet(int n, double *a, double *logdet);


/* ---- projective DQMC: spinless t-V model, bond HS field, submatrix-LR update ---------- */
typedef struct orc_pparams {
    int lattice, la, lb;
    double beta, dtau, rt, mu, rv, nu, rndness;
    int nflr;
    int kmode;      /* 0 = checkerboard, 1 = dense */
    int nwrap, nustock, mid_len;
    int wrap_check; /* 1: the WRAPERROR comparison at every stabilization */
    int seed;
    double rhub;    /* on-site U: the uconf field (src/dqmc_config_u) joins the bond field when != 0 */
    int exact_hs;   /* uconf tables: 1 = EXACTHS, 0 = general 4-component */
} orc_pparams;

typedef struct orc_pctx orc_pctx;
/* the on-site field of a projective chain with rhub != 0: [ltrot][N], values 1..lcomp */
void orc_pget_uconf(const orc_pctx *c, int32_t *conf);
void orc_pset_uconf(orc_pctx *c, const int32_t *conf);
orc_pctx *orc_pcreate(const orc_pparams *p);
void orc_pdestroy(orc_pctx *c);
/* sizes: [0]=N [1]=ne [2]=ltrot [3]=nst [4]=nn_nf [5]=nn_lf [6]=nustock [7]=lq */
void orc_psizes(const orc_pctx *c, int *out);
void orc_pinit_conf(orc_pctx *c);                          /* N*nn_nf*ltrot draws */
void orc_pget_conf(const orc_pctx *c, int32_t *conf);      /* [ltrot][nn_nf][N] */
void orc_pset_conf(orc_pctx *c, const int32_t *conf);
double orc_pset_trial(orc_pctx *c);                        /* nn_lf*nn_nf draws; returns the gap */
void orc_pset_proj(orc_pctx *c, const double *projR);      /* N x ne complex */
void orc_pget_proj(const orc_pctx *c, double *projR, double *w);
int orc_psweep_start(orc_pctx *c);
int orc_psweep_b0(orc_pctx *c, int lupdate, int lmeasure);
int orc_psweep_0b(orc_pctx *c, int lupdate, int lmeasure);
void orc_pget_g(orc_pctx *c, double *g);
void orc_pget_state(const orc_pctx *c, double *ul, double *ur, double *ulrinv);
void orc_pget_diag(const orc_pctx *c, double *d);
void orc_preset_diag(orc_pctx *c);
void orc_pget_obs(const orc_pctx *c, double *o);
/* the correlation bins of the measured slices (layout of orc_get_eqt) */
int orc_pset_eqt(orc_pctx *c, double theta);
void orc_pget_eqt(const orc_pctx *c, double *out);
void orc_preset_obs(orc_pctx *c);
void orc_pset_trace(orc_pctx *c, double *buf, int cap);
int orc_ptrace_count(const orc_pctx *c);
int orc_pget_table(const orc_pctx *c, int kind, double *out);
/* s_zgeQR_pqmc: a (rdim x cdim) -> Q in place (LQ by plain transpose when rdim < cdim); returns log|det R| */
double orc_qr_pqmc(int rdim, int cdim, double *a);

#ifdef __cplusplus
}
#endif
#endif
