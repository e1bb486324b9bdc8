/*
 * sunhip.h -- C ABI of the MI355X-native DQMC sweep (libsunhip.so).
 *
 * Drop-in boundary for the reference's sweep routines (YuanyueChen/sunhubbardmy):
 * a Fortran control layer keeps dqmc.in, the driver and the measurement output, and calls
 * these entry points through ISO_C_BINDING in place of the module procedures of
 * src/dqmc_ft_core.f90.  Plain pointers and sizes only; every host buffer is owned by the
 * caller and read or written only during the call.  Complex numbers are complex(dp) /
 * double _Complex: two doubles (re, im); matrices are column-major.
 *
 * Return codes: 0 = success; > 0 = numerical breakdown (zero pivot in the LU of
 * green_equaltime, index + 1); < 0 = usage or runtime error (SH_E_*).  The library never
 * aborts; sh_last_error() gives a message.
 */
#ifndef SUNHIP_H
#define SUNHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SH_E_ARG (-1)      /* bad argument */
#define SH_E_HIP (-2)      /* HIP runtime failure (allocation, launch) */
#define SH_E_STATE (-3)    /* call out of order (e.g. sweep before sweep_start) */
#define SH_E_NONFINITE (-4) /* non-finite Green's function detected */

#define SH_K_CHECKERBOARD 0 /* BREAKUP_T: 2x2 bond exponentials, nn_nf families */
#define SH_K_DENSE 1        /* dense exp(-dtau K/2), N x N */
#define SH_K_FFT 2          /* FFT build: exp(-dtau K/2) = F^H diag(exph0k) F over the lattice cells
                               (src/dqmc_config_h0/dqmc_left_forward_prop.f90:95-134,
                               dqmc_set_h0conf.f90:225-291); one site per cell */

#define SH_U_DELAY 0        /* dqmc_update_delay.f90 (nublock = 1 is the fast update) */
#define SH_U_SUBMATRIX 1    /* dqmc_update_submatrix.f90 */

/*
 * One chain's model tables: what model_para%make_tables, h0c%set_h0conf and
 * u0conf%set_conf leave in the reference's module state
 * (src/model_para.f90:421-453, src/dqmc_config_h0/dqmc_set_h0conf.f90,
 *  src/dqmc_config_u/dqmc_set_conf.f90:122-175).  All index tables are 0-based.
 */
typedef struct sh_model {
    int ndim;               /* matrix dimension N (= nsites) */
    int lq;                 /* number of unit cells (normalises S_stag, density) */
    int ltrot, nwrap, nst;  /* slices, stabilization interval, number of stabilization points */
    const int32_t *iwrap_nt;   /* [ltrot+1] iwrap_nt(0:ltrot) */
    const int32_t *wrap_step;  /* [2*nst] wrap_step(1:2, 1:nst), 1-based slice numbers */
    int lwrapT, lwrapu;     /* kinetic / Hubbard terms present (rt != 0, rhub != 0) */
    int kmode;              /* SH_K_* */
    int nn_nf, nn_lf;       /* bond families, bonds per family */
    const int32_t *bonds;   /* [nn_nf*nn_lf*2] (A site, B site) per bond, family-major */
    const double *urt;      /* checkerboard: [nn_nf*nn_lf] 2x2 complex col-major; dense: N*N complex;
                               FFT: [N] complex exph0k in the order of the wave-vector index */
    const double *urtm1;    /* inverse propagator, same layout */
    int lcomp, nflr;
    const double *bmat;        /* [lcomp] complex  exp(i alpha eta_l) */
    const double *bmat_inv;    /* [lcomp] complex */
    const double *phase_arg;   /* [lcomp] real: arg of u0conf%phase(l) (set_phase product) */
    const double *delta_bmat;  /* [(lcomp-1)*lcomp] complex, (iflip, is) column-major */
    const double *phase_ratio; /* [(lcomp-1)*lcomp] complex */
    const double *gaml;        /* [lcomp] */
    int phase_from_conf;    /* 1 when set_phase multiplies u0conf%phase (rhub > 0) */
    const double *stag;     /* [ndim] +-1 staggering used for S_stag (S(pi,pi) on the square lattice) */
    int update;             /* SH_U_* */
    int nublock;            /* delay depth / submatrix block */
    int32_t seed;           /* SFMT19937 seed of this chain (spring_sfmt_init) */
    double kin_weight;      /* factor of the zkint bin: 1 (0 is read as 1) in 2D and 3D, nflr * rt for
                               the chain (src/dqmc_measure/equaltime_measure1d.f90:76) */
    /* the bond field of the t-V model (v0conf, rv != 0; NULL: none): vtab 8 complex (bmat_p,
     * bmat_m, bmat_p_inv, bmat_m_inv of component 1, then 2), vdelta 4 complex (delta_bmat_p,
     * delta_bmat_m)(1, is) for is = 1, 2 (src/dqmc_config_v/dqmc_set_conf.f90:29-68); the
     * bonds are the (nn_nf, nn_lf) table above, also for the dense kinetic term */
    const double *vtab;
    const double *vdelta;
    /* SH_K_FFT: the lattice directions in site order, site s = i0 + n0 (i1 + n1 i2) with
     * fft_dims = {n0, n1, n2}, n0 n1 n2 = ndim, each 1..16, 18, 20 or 24 (the reference's
     * DFTI dimensions dimm); the wave-vector index of exph0k is ordered the same way.  Ignored
     * for the other kinetic forms. */
    int fft_dims[3];
    /* the plaquette field (hconf with rhub_plq != 0 and lprojplqu = .false.; zplq = 0: none):
     * zplq sites per plaquette (4 square, 6 honeycomb), lqp plaquettes in the reference's order,
     * plq_cord [lqp][zplq] their 0-based sites (latt%plq_cord, src/lattice.f90:735-773, 898-958);
     * per component is = 1..4 the zplq x zplq complex column-major factors bmat_plqu_blk1(:,:,is)
     * (plq_b) and bmat_plqu_blk1_inv(:,:,is) (plq_bi), per (is, iflip) delta_bmat_plqu_blk1(:,:,
     * iflip, is) at block (is-1)*3 + iflip-1 (plq_delta); phase_ratio(iflip, is) complex at
     * (iflip-1) + 3 (is-1); plq_parg[4] the arg of hconf%phase(is); plq_gaml[4]
     * (src/dqmc_config_plqu/dqmc_set_plqconf.f90:57-199, dqmc_config_plqu.f90:42-51) */
    int zplq, lqp;
    const int32_t *plq_cord;
    const double *plq_b, *plq_bi, *plq_delta, *plq_pratio, *plq_parg, *plq_gaml;
} sh_model;

typedef struct sh_ctx sh_ctx;

/* allocate_core for a batch of independent chains (uneven ndim allowed; all chains share
 * ltrot, nwrap and nst).  Replaces allocate_core (src/dqmc_ft_core.f90:22-50) per MPI rank. */
int sh_create(sh_ctx **ctx, int nchains, const sh_model *models, int device);
/* deallocate_core (src/dqmc_ft_core.f90:52-58) */
int sh_destroy(sh_ctx *ctx);
const char *sh_last_error(const sh_ctx *ctx);

/* u0conf%input_conf, fresh start: each chain draws conf(i,nt) = ceiling(u*lcomp) from its own
 * SFMT stream, nt outer, i inner (src/dqmc_config_u/dqmc_input_conf.f90:55-63) */
int sh_init_conf(sh_ctx *ctx);
/* conf(1:nsites, 1:ltrot) as int32 [ltrot][ndim], values 1..lcomp (conf_u.in restart path) */
int sh_set_conf(sh_ctx *ctx, int chain, const int32_t *conf);
int sh_get_conf(sh_ctx *ctx, int chain, int32_t *conf);

/* Time-displaced measurement (ltau, dyntau = ntdm * dtau: src/model_para.f90:457).  sh_set_dyn
 * gives a chain the unit-cell coordinates of its sites, [ndim][4] int32: x in 1..l1, y in 1..l2,
 * z in 1..l3, orbital 0 or 1 (latt%list(i,1:2) with latt%ulist, src/lattice.f90:67-135, 386-491,
 * 1033-1157; l2 = l3 = 1 for the chain, l3 = 1 in 2D) and switches on G(t,0),
 * G(0,t), G(0,0) and the znn_tau bins; every chain of a batch takes the same ntdm (<= ltrot).
 * sh_sweep_0b(ctx, 0, lmeasure, 1) is then the reference's dynamic pass (dqmc_ft_ctrl.f90:99-101
 * between sh_push_stage and sh_pop_stage): no updates, green_tau at every stabilization point
 * (src/dqmc_ft_core/green_tau.f90:1-122), znn_tau(imj, no_i, no_j, t) += ... for t = 0 .. ntdm
 * (src/dqmc_measure/dyn_measure.f90:51-64; the reference writes past its array when ntdm < ltrot,
 * here the slices past ntdm are not measured). */
int sh_set_dyn(sh_ctx *ctx, int chain, int ntdm, int l1, int l2, int l3, const int32_t *sites);
/* znn_tau_bin as complex [ntdm+1][2][2][lq] (the Fortran znn_tau(lq, 2, 2, ntdm+1)); dynerr (may be
 * NULL): the largest |propagated - green_tau| of G(t,0) and G(0,t) and the number of such
 * comparisons (xmax_dyn, logdyncount of -DDYNERROR) */
int sh_get_dyn_obs(sh_ctx *ctx, int chain, double *znn, double *dynerr);
int sh_reset_dyn_obs(sh_ctx *ctx);
/* G(t,0), G(0,t), G(0,0) (N x N complex, column-major) as the last dynamic pass left them */
int sh_get_gtau(sh_ctx *ctx, int chain, double *gt0, double *g0t, double *g00);

/* The correlation bins of equaltime_measure (src/dqmc_measure/equaltime_measure.f90:44-68), on top
 * of the sh_get_obs sums at every measured slice from this call on (off by default).  sites as
 * for sh_set_dyn; cells [lq][4]: the A site of each unit cell, its nnlist(., 1), its snlist(., 1)
 * and its current partner (snlist on the honeycomb, nnlist elsewhere: equaltime_measure.f90:
 * 378-394), 0-based and in latt%imj's cell order; plq [lq][zplq] the plaquette site cycles
 * (latt%plq_cord, src/lattice.f90:954-957, :767-772; zplq = 0: no plaquette terms); theta in
 * radians.  Needs 16 lq + ndim + 544 <= 10240. */
int sh_set_eqt(sh_ctx *ctx, int chain, int l1, int l2, int l3, const int32_t *sites, const int32_t *cells, int zplq,
               const int32_t *plq, double theta);
/* 20 lq + 8 complex, each summed with zphi: zcpcm, zspsm, znn, pair_onsite as (imj, no_i, no_j) at
 * q 4 lq + imj + lq (no_i + 2 no_j); zjj, zbb, pair_nn, pair_sn at (16 + q) lq + imj; then zn(1),
 * zn(2), zj, zb, zqsq (without its 1/9), ztsq, ztq, zint_v (without rv).  Cleared by sh_reset_obs. */
int sh_get_eqt(sh_ctx *ctx, int chain, double *bins);

/* The bond field of the t-V model (v0conf%conf(i, nf, nt), src/dqmc_config_v): int32
 * [ltrot][nn_nf][ndim], values 1..2, entries i < nn_lf are the bonds (conf_v.in / conf_v.out);
 * sh_init_conf draws it after the on-site field.  sh_get_vacc: accepted and tried bond flips
 * (main_obs(4) of dqmc_update.f90:97). */
int sh_set_confv(sh_ctx *ctx, int chain, const int32_t *conf);
int sh_get_confv(sh_ctx *ctx, int chain, int32_t *conf);
int sh_get_vacc(sh_ctx *ctx, int chain, double *acc);
/* The plaquette field (hconf%conf_plqu(1:lqp, 1:ltrot), values 1..4) as int32 [ltrot][lqp], and
 * its acceptance sums (accepted, tried): main_obs(2) of dqmc_update_plqu.f90:85.  The field is
 * drawn first by sh_init_conf (hconf%input_plqconf before u0conf%input_conf, src/dqmc_ft_ctrl.f90:
 * 19-35), ceiling(u*4), nt outer, plaquette inner (dqmc_input_plqconf.f90:55-62). */
int sh_set_confp(sh_ctx *ctx, int chain, const int32_t *conf);
int sh_get_confp(sh_ctx *ctx, int chain, int32_t *conf);
int sh_get_pacc(sh_ctx *ctx, int chain, double *acc);

/* The chain's random-number state, SH_RNG_WORDS uint32: the SFMT19937 state that spring_sfmt
 * keeps in module storage (lib/m_spring.f90:66-71: 624 words, then the read index 0..624) and,
 * for chains with learned block proposals, the proposal generator's 64-bit state (low word
 * first; zeros when proposals were never set up, and ignored by the setter then).  Together with
 * the HS field it resumes a chain bit-exactly; the reference itself restarts from conf_u.in with
 * a fresh seed (src/dqmc_util.f90:68-75), so this pair has no reference counterpart. */
#define SH_RNG_WORDS 627
int sh_get_rng_state(sh_ctx *ctx, int chain, uint32_t *state);
int sh_set_rng_state(sh_ctx *ctx, int chain, const uint32_t *state);

/* dqmc_ft_sweep_start_0b (src/dqmc_ft_core/dqmc_ft_sweep_start_0b.f90:1-71) */
int sh_sweep_start_0b(sh_ctx *ctx);
/* dqmc_ft_sweep_b0(lupdate, lmeasure_equaltime) (src/dqmc_ft_core/dqmc_ft_sweep_b0.f90:1-216) */
int sh_sweep_b0(sh_ctx *ctx, int lupdate, int lmeasure);
/* dqmc_ft_sweep_0b(lupdate, lmeasure_equaltime, lmeasure_dyn)
 * (src/dqmc_ft_core/dqmc_ft_sweep_0b.f90:1-288); lmeasure_dyn = 1 needs sh_set_dyn and lupdate = 0 */
int sh_sweep_0b(sh_ctx *ctx, int lupdate, int lmeasure, int lmeasure_dyn);
/* push_stage / pop_stage (src/dqmc_ft_core.f90:290-312) */
int sh_push_stage(sh_ctx *ctx);
int sh_pop_stage(sh_ctx *ctx);

/* gf%blk1 of one chain, N*N complex column-major */
int sh_get_g(sh_ctx *ctx, int chain, double *g);
/* diag[10]: max_wrap_error, max_phasediff, phase(re,im), logweightf(re,im),
 * accepted, tried (main_obs(3)), number of wrap checks, sum log10 wrap error */
int sh_get_diag(sh_ctx *ctx, int chain, double *diag);
/* clears the wrap-error and acceptance accumulators and the non-finite flag */
int sh_reset_diag(sh_ctx *ctx);
/* obs[11]: nobs, then complex bin sums of sign, density/lq (energy_bin(1:2)), zkint
 * (energy_bin(3)), zint_u (energy_bin(4)), S_stag (zspsm Fourier component at the
 * staggering wave vector), all weighted by zphi as in equaltime_measure */
int sh_get_obs(sh_ctx *ctx, int chain, double *obs);
/* obs[11*nchains] of every chain in one readback, then (reset != 0) zero the sums: the
 * per-sweep bins of a batch (equaltime_output with nsweep = 1) */
int sh_get_obs_all(sh_ctx *ctx, double *obs, int reset);
int sh_reset_obs(sh_ctx *ctx);

/* Record G after every stabilization (nwrap) of chain `chain`, up to cap snapshots; buf is
 * device-owned and read back with sh_get_trace (N*N complex per snapshot). cap = 0 stops. */
int sh_set_trace(sh_ctx *ctx, int chain, int cap);
int sh_get_trace(sh_ctx *ctx, int chain, double *buf, int *count);

/* per-kernel-group device time since the last call, from hipEvent pairs on the library's
 * stream: t[0..10] ms of wrap (K and V on G), update, stabilization B products, QR (prepivot +
 * Householder), form Q, ZGEMM, LU, LU solve, measure, other, block proposals; t[11..21] launch counts of the
 * same groups.  enable = 0 stops recording (events cost a little launch overhead). */
int sh_timers(sh_ctx *ctx, int enable, double *t);
int sh_synchronize(sh_ctx *ctx);

/* Graph-captured passes: with enable = 1 (and the timers off) each kind of sh_sweep_b0 /
 * sh_sweep_0b call (pass, lupdate, lmeasure) is captured once into a hipGraph and replayed, one
 * host launch per pass instead of one per kernel; results are those of the direct launches.
 * Setters that change what a pass launches retire the captured graphs.  The reference has no
 * counterpart (its passes are host loops, src/dqmc_ft_core/dqmc_ft_sweep_b0.f90:36-207). */
int sh_set_graphs(sh_ctx *ctx, int enable);

/* Autotuner of the delay depth / submatrix block and of the DFT kernels' vectors per workgroup,
 * in place of the heuristic tables of src/model_para.f90:382-415: after sh_sweep_start_0b, the
 * update of one slice is timed on the chains' own G at every candidate depth (delay 4, 8, 16, 32,
 * and 64 where the delayed vectors do not fit in registers; submatrix 8 .. 64) and the fastest is
 * kept for every chain; with the FFT kinetic form the DFT kernels are timed at 2 .. 16 vectors
 * per workgroup.  The chains' state is left bit for bit as it was (the next updating pass draws
 * its own uniforms).  tuned[0] = chosen nublock, tuned[1] = vectors per workgroup (0: not FFT).
 * The delay depth changes rounding only, not the algorithm (dqmc_update_delay.f90:39-128). */
int sh_autotune(sh_ctx *ctx, int *tuned);

#ifdef __cplusplus
}
#endif
#endif
