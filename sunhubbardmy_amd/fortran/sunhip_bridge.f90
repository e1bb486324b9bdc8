! sunhip_bridge.f90 -- ISO_C_BINDING interface to libsunhip.so for the reference's Fortran
! control layer.  A maintainer adds this module next to src/dqmc_ft_core.f90 and calls the
! sh_* entry points where src/dqmc_ft_ctrl.f90 calls allocate_core, dqmc_ft_sweep_start_0b,
! dqmc_ft_sweep_b0/0b and push_stage/pop_stage, and the shp_* entry points where
! src/dqmc_proj_ctrl.f90 calls allocate_core, dqmc_set_proj, dqmc_proj_sweep_start_0b and
! dqmc_proj_sweep_b0/0b.  Nonzero return codes map onto the reference's `stop`
! (dqmc_ft_sweep_b0.f90:188-194, dqmc_ft_sweep_start_0b.f90:44-47).
module sunhip_bridge
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: SH_K_CHECKERBOARD = 0, SH_K_DENSE = 1, SH_K_FFT = 2
  integer(c_int), parameter :: SH_U_DELAY = 0, SH_U_SUBMATRIX = 1
  integer(c_int), parameter :: SHP_U_LR = 0, SHP_U_FAST = 1, SHP_U_DELAY = 2, SHP_U_SUBMATRIX_G = 3

  ! mirrors struct sh_model of include/sunhip.h
  type, bind(C) :: sh_model
     integer(c_int) :: ndim, lq, ltrot, nwrap, nst
     type(c_ptr) :: iwrap_nt, wrap_step
     integer(c_int) :: lwrapT, lwrapu, kmode, nn_nf, nn_lf
     type(c_ptr) :: bonds, urt, urtm1
     integer(c_int) :: lcomp, nflr
     type(c_ptr) :: bmat, bmat_inv, phase_arg, delta_bmat, phase_ratio, gaml
     integer(c_int) :: phase_from_conf
     type(c_ptr) :: stag
     integer(c_int) :: update, nublock
     integer(c_int32_t) :: seed
     real(c_double) :: kin_weight
     ! the t-V bond field (c_null_ptr: none)
     type(c_ptr) :: vtab, vdelta
     ! SH_K_FFT: lattice directions in site order (n0 fastest); urt / urtm1 = exph0k / exph0kinv
     integer(c_int) :: fft_dims(3)
     ! the plaquette field (zplq = 0: none): sites per plaquette, plaquettes, plq_cord(zplq, lqp)
     ! (0-based), the factor tables bmat / bmat_inv (zplq, zplq, 4), delta (zplq, zplq, 3, 4),
     ! phase_ratio (3, 4) complex, the phase args (4) and gaml (4)
     integer(c_int) :: zplq, lqp
     type(c_ptr) :: plq_cord, plq_b, plq_bi, plq_delta, plq_pratio, plq_parg, plq_gaml
  end type sh_model

  ! mirrors struct shp_model of include/sunhip_proj.h
  type, bind(C) :: shp_model
     integer(c_int) :: ndim, ne, ltrot, nwrap, nst
     type(c_ptr) :: iwrap_nt, wrap_step
     integer(c_int) :: kmode, nn_nf, nn_lf
     type(c_ptr) :: bonds, urt, urtm1, vtab, vdelta, stag
     integer(c_int) :: lq, nflr, nustock, mid_len, wrap_check
     integer(c_int32_t) :: seed
     ! the on-site field (c_null_ptr ubmat: none); vtab = c_null_ptr: no bond field
     integer(c_int) :: ulcomp
     type(c_ptr) :: ubmat, ubmat_inv, udelta, uratio, ugaml, uphase_arg
     integer(c_int) :: uphase_from_conf
     real(c_double) :: kin_weight
     ! the on-site update: SHP_U_LR, SHP_U_FAST, SHP_U_DELAY, SHP_U_SUBMATRIX_G, and its depth
     integer(c_int) :: update, gblock
  end type shp_model

  interface
     integer(c_int) function sh_create(ctx, nchains, models, device) bind(C, name="sh_create")
       import :: c_ptr, c_int, sh_model
       type(c_ptr), intent(out) :: ctx
       integer(c_int), value :: nchains, device
       type(sh_model), intent(in) :: models(*)
     end function sh_create
     integer(c_int) function sh_destroy(ctx) bind(C, name="sh_destroy")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function sh_destroy
     integer(c_int) function sh_init_conf(ctx) bind(C, name="sh_init_conf")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function sh_init_conf
     integer(c_int) function sh_set_conf(ctx, chain, conf) bind(C, name="sh_set_conf")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(in) :: conf(*)
     end function sh_set_conf
     integer(c_int) function sh_get_conf(ctx, chain, conf) bind(C, name="sh_get_conf")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(out) :: conf(*)
     end function sh_get_conf
     ! the t-V bond field conf(ndim, nn_nf, ltrot), values 1..2, and its acceptance sums
     integer(c_int) function sh_set_confv(ctx, chain, conf) bind(C, name="sh_set_confv")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(in) :: conf(*)
     end function sh_set_confv
     integer(c_int) function sh_get_confv(ctx, chain, conf) bind(C, name="sh_get_confv")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(out) :: conf(*)
     end function sh_get_confv
     integer(c_int) function sh_get_vacc(ctx, chain, acc) bind(C, name="sh_get_vacc")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       real(c_double), intent(out) :: acc(2)
     end function sh_get_vacc
     ! the plaquette field conf_plqu(lqp, ltrot) and its (accepted, tried) counts
     integer(c_int) function sh_set_confp(ctx, chain, conf) bind(C, name="sh_set_confp")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(in) :: conf(*)
     end function sh_set_confp
     integer(c_int) function sh_get_confp(ctx, chain, conf) bind(C, name="sh_get_confp")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(out) :: conf(*)
     end function sh_get_confp
     integer(c_int) function sh_get_pacc(ctx, chain, acc) bind(C, name="sh_get_pacc")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       real(c_double), intent(out) :: acc(2)
     end function sh_get_pacc
     ! correlation bins of equaltime_measure: cells(4, lq) = A site, nnlist(., 1), snlist(., 1),
     ! current partner (0-based); plq(zplq, lq) plaquette cycles; bins(2, 20 lq + 8)
     integer(c_int) function sh_set_eqt(ctx, chain, l1, l2, l3, sites, cells, zplq, plq, theta) &
          bind(C, name="sh_set_eqt")
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain, l1, l2, l3, zplq
       integer(c_int32_t), intent(in) :: sites(*), cells(*), plq(*)
       real(c_double), value :: theta
     end function sh_set_eqt
     integer(c_int) function sh_get_eqt(ctx, chain, bins) bind(C, name="sh_get_eqt")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       real(c_double), intent(out) :: bins(*)
     end function sh_get_eqt
     ! ltau: sites(4, ndim) = cell x (1..l1), y (1..l2), z (1..l3), orbital (0, 1) per site
     integer(c_int) function sh_set_dyn(ctx, chain, ntdm, l1, l2, l3, sites) bind(C, name="sh_set_dyn")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain, ntdm, l1, l2, l3
       integer(c_int32_t), intent(in) :: sites(*)
     end function sh_set_dyn
     ! znn_tau_bin(lq, 2, 2, ntdm+1) as complex(c_double_complex); dynerr(2) = xmax_dyn, count
     integer(c_int) function sh_get_dyn_obs(ctx, chain, znn, dynerr) bind(C, name="sh_get_dyn_obs")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       real(c_double), intent(out) :: znn(*), dynerr(2)
     end function sh_get_dyn_obs
     integer(c_int) function sh_reset_dyn_obs(ctx) bind(C, name="sh_reset_dyn_obs")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function sh_reset_dyn_obs
     integer(c_int) function sh_get_gtau(ctx, chain, gt0, g0t, g00) bind(C, name="sh_get_gtau")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       real(c_double), intent(out) :: gt0(*), g0t(*), g00(*)
     end function sh_get_gtau
     ! the chain's generator state, SH_RNG_WORDS = 627 words (uint32 as int32): bit-exact resume
     integer(c_int) function sh_get_rng_state(ctx, chain, state) bind(C, name="sh_get_rng_state")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(out) :: state(*)
     end function sh_get_rng_state
     integer(c_int) function sh_set_rng_state(ctx, chain, state) bind(C, name="sh_set_rng_state")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(in) :: state(*)
     end function sh_set_rng_state
     integer(c_int) function sh_sweep_start_0b(ctx) bind(C, name="sh_sweep_start_0b")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function sh_sweep_start_0b
     integer(c_int) function sh_sweep_b0(ctx, lupdate, lmeasure) bind(C, name="sh_sweep_b0")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: lupdate, lmeasure
     end function sh_sweep_b0
     integer(c_int) function sh_sweep_0b(ctx, lupdate, lmeasure, ldyn) bind(C, name="sh_sweep_0b")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: lupdate, lmeasure, ldyn
     end function sh_sweep_0b
     integer(c_int) function sh_push_stage(ctx) bind(C, name="sh_push_stage")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function sh_push_stage
     integer(c_int) function sh_pop_stage(ctx) bind(C, name="sh_pop_stage")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function sh_pop_stage
     integer(c_int) function sh_get_g(ctx, chain, g) bind(C, name="sh_get_g")
       import :: c_ptr, c_int, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       complex(c_double_complex), intent(out) :: g(*)
     end function sh_get_g
     integer(c_int) function sh_get_diag(ctx, chain, d) bind(C, name="sh_get_diag")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       real(c_double), intent(out) :: d(10)
     end function sh_get_diag
     integer(c_int) function sh_get_obs(ctx, chain, o) bind(C, name="sh_get_obs")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       real(c_double), intent(out) :: o(11)
     end function sh_get_obs
     integer(c_int) function sh_reset_obs(ctx) bind(C, name="sh_reset_obs")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function sh_reset_obs

     ! projective sweep (include/sunhip_proj.h)
     integer(c_int) function shp_create(ctx, nchains, models, device) bind(C, name="shp_create")
       import :: c_ptr, c_int, shp_model
       type(c_ptr), intent(out) :: ctx
       integer(c_int), value :: nchains, device
       type(shp_model), intent(in) :: models(*)
     end function shp_create
     integer(c_int) function shp_destroy(ctx) bind(C, name="shp_destroy")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function shp_destroy
     integer(c_int) function shp_init_conf(ctx) bind(C, name="shp_init_conf")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function shp_init_conf
     integer(c_int) function shp_set_conf(ctx, chain, conf) bind(C, name="shp_set_conf")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(in) :: conf(*)
     end function shp_set_conf
     integer(c_int) function shp_get_conf(ctx, chain, conf) bind(C, name="shp_get_conf")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(out) :: conf(*)
     end function shp_get_conf
     integer(c_int) function shp_set_uconf(ctx, chain, conf) bind(C, name="shp_set_uconf")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(in) :: conf(*)
     end function shp_set_uconf
     integer(c_int) function shp_get_uconf(ctx, chain, conf) bind(C, name="shp_get_uconf")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(out) :: conf(*)
     end function shp_get_uconf
     ! SHP_RNG_WORDS = 625 words
     integer(c_int) function shp_get_rng_state(ctx, chain, state) bind(C, name="shp_get_rng_state")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(out) :: state(*)
     end function shp_get_rng_state
     integer(c_int) function shp_set_rng_state(ctx, chain, state) bind(C, name="shp_set_rng_state")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       integer(c_int32_t), intent(in) :: state(*)
     end function shp_set_rng_state
     integer(c_int) function shp_draw_uniforms(ctx, chain, n, u) bind(C, name="shp_draw_uniforms")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain, n
       real(c_double), intent(out) :: u(*)
     end function shp_draw_uniforms
     integer(c_int) function shp_set_proj(ctx, chain, proj) bind(C, name="shp_set_proj")
       import :: c_ptr, c_int, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       complex(c_double_complex), intent(in) :: proj(*)
     end function shp_set_proj
     integer(c_int) function shp_sweep_start_0b(ctx) bind(C, name="shp_sweep_start_0b")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function shp_sweep_start_0b
     integer(c_int) function shp_sweep_b0(ctx, lupdate, lmeasure) bind(C, name="shp_sweep_b0")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: lupdate, lmeasure
     end function shp_sweep_b0
     integer(c_int) function shp_sweep_0b(ctx, lupdate, lmeasure, ldyn) bind(C, name="shp_sweep_0b")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: lupdate, lmeasure, ldyn
     end function shp_sweep_0b
     integer(c_int) function shp_get_g(ctx, chain, g) bind(C, name="shp_get_g")
       import :: c_ptr, c_int, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       complex(c_double_complex), intent(out) :: g(*)
     end function shp_get_g
     integer(c_int) function shp_get_diag(ctx, chain, d) bind(C, name="shp_get_diag")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       real(c_double), intent(out) :: d(10)
     end function shp_get_diag
     integer(c_int) function shp_get_obs(ctx, chain, o) bind(C, name="shp_get_obs")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       real(c_double), intent(out) :: o(11)
     end function shp_get_obs
     integer(c_int) function shp_reset_obs(ctx) bind(C, name="shp_reset_obs")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function shp_reset_obs
     ! learned block proposals (include/sunhip_learn.h)
     integer(c_int) function shl_create(net, capacity, seed, device) bind(C, name="shl_create")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr) :: net
       integer(c_int), value :: capacity, device
       integer(c_int32_t), value :: seed
     end function shl_create
     integer(c_int) function shl_destroy(net) bind(C, name="shl_destroy")
       import :: c_ptr, c_int
       type(c_ptr), value :: net
     end function shl_destroy
     integer(c_int) function shl_add_samples(net, ctx) bind(C, name="shl_add_samples")
       import :: c_ptr, c_int
       type(c_ptr), value :: net, ctx
     end function shl_add_samples
     integer(c_int) function shl_train(net, epochs, batch, lr, loss) bind(C, name="shl_train")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: net
       integer(c_int), value :: epochs, batch
       real(c_double), value :: lr
       real(c_double) :: loss(*)
     end function shl_train
     integer(c_int) function sh_set_neighbours(ctx, z, nbr) bind(C, name="sh_set_neighbours")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: z
       integer(c_int32_t) :: nbr(*)
     end function sh_set_neighbours
     integer(c_int) function sh_set_proposal(ctx, net, nprop, local_updates) bind(C, name="sh_set_proposal")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, net
       integer(c_int), value :: nprop, local_updates
     end function sh_set_proposal
     integer(c_int) function sh_get_proposal_stats(ctx, chain, stats) bind(C, name="sh_get_proposal_stats")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: chain
       real(c_double) :: stats(*)
     end function sh_get_proposal_stats
  end interface

contains

  ! the reference's failure convention: print and stop on a nonzero code
  subroutine sh_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc /= 0) then
       write(*, '(a,a,i6)') ' Error in ', where, rc
       stop
    end if
  end subroutine sh_check

end module sunhip_bridge
