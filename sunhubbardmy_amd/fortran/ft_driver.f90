! ft_driver.f90 -- a Fortran caller of libsunhip.so through sunhip_bridge, in the shape of the
! reference's finite-temperature control code (src/dqmc_ft_ctrl.f90:37-49 allocate_core and
! dqmc_ft_sweep_start_0b, :68-104 the b0 / 0b sweeps, deallocate_core).  It reads one chain's
! model tables (what make_tables, h0c%set_h0conf and u0conf%set_conf leave in module state) from
! tables.bin, runs nsweep sweeps with updates and measurement, and writes the HS field, G, the
! diagnostics and the observable bin sums to out.bin (all stream files, native byte order).
!
! tables.bin: int32 hdr(17) = ndim, lq, ltrot, nwrap, nst, lwrapT, lwrapu, kmode, nn_nf, nn_lf,
!   lcomp, nflr, phase_from_conf, update, nublock, seed, nsweep; then float64 kin_weight; int32
!   fft_dims(3); int32 iwrap_nt(0:ltrot), wrap_step(2, nst), bonds(2, nn_lf, nn_nf); complex
!   urt(nk), urtm1(nk) (nk = 4 nn_nf nn_lf checkerboard, ndim**2 dense, ndim FFT); complex
!   bmat(lcomp), bmat_inv(lcomp); float64 phase_arg(lcomp); complex delta_bmat(lcomp-1, lcomp),
!   phase_ratio(lcomp-1, lcomp); float64 gaml(lcomp), stag(ndim).
! out.bin: int32 conf(ndim, ltrot); complex g(ndim, ndim); float64 diag(10), obs(11).
program ft_driver
  use sunhip_bridge
  use, intrinsic :: iso_c_binding
  implicit none
  type(sh_model), target :: m(1)
  type(c_ptr) :: ctx
  integer(c_int32_t) :: hdr(17), fdims(3)
  integer(c_int32_t), allocatable, target :: iwrap(:), wstep(:), bonds(:), conf(:)
  complex(c_double_complex), allocatable, target :: urt(:), urtm1(:), bmat(:), bmat_inv(:), delta(:), pratio(:)
  real(c_double), allocatable, target :: parg(:), gaml(:), stag(:)
  complex(c_double_complex), allocatable :: g(:)
  real(c_double) :: kw, diag(10), obs(11)
  integer :: u, n, ltrot, nst, nnf, nlf, lcomp, nk, nsweep, isw

  open(newunit=u, file='tables.bin', access='stream', form='unformatted', status='old')
  read(u) hdr
  read(u) kw
  read(u) fdims
  n = hdr(1); ltrot = hdr(3); nst = hdr(5); nnf = hdr(9); nlf = hdr(10); lcomp = hdr(11); nsweep = hdr(17)
  select case (hdr(8))
  case (SH_K_CHECKERBOARD)
     nk = 4 * nnf * nlf
  case (SH_K_DENSE)
     nk = n * n
  case default
     nk = n
  end select
  allocate(iwrap(ltrot + 1), wstep(2 * nst), bonds(2 * nnf * nlf), urt(nk), urtm1(nk))
  allocate(bmat(lcomp), bmat_inv(lcomp), parg(lcomp), delta((lcomp - 1) * lcomp), pratio((lcomp - 1) * lcomp))
  allocate(gaml(lcomp), stag(n), conf(n * ltrot), g(n * n))
  read(u) iwrap, wstep, bonds, urt, urtm1, bmat, bmat_inv, parg, delta, pratio, gaml, stag
  close(u)

  m(1)%ndim = hdr(1); m(1)%lq = hdr(2); m(1)%ltrot = hdr(3); m(1)%nwrap = hdr(4); m(1)%nst = hdr(5)
  m(1)%iwrap_nt = c_loc(iwrap); m(1)%wrap_step = c_loc(wstep)
  m(1)%lwrapT = hdr(6); m(1)%lwrapu = hdr(7); m(1)%kmode = hdr(8); m(1)%nn_nf = hdr(9); m(1)%nn_lf = hdr(10)
  m(1)%bonds = c_loc(bonds); m(1)%urt = c_loc(urt); m(1)%urtm1 = c_loc(urtm1)
  m(1)%lcomp = hdr(11); m(1)%nflr = hdr(12)
  m(1)%bmat = c_loc(bmat); m(1)%bmat_inv = c_loc(bmat_inv); m(1)%phase_arg = c_loc(parg)
  m(1)%delta_bmat = c_loc(delta); m(1)%phase_ratio = c_loc(pratio); m(1)%gaml = c_loc(gaml)
  m(1)%phase_from_conf = hdr(13); m(1)%stag = c_loc(stag)
  m(1)%update = hdr(14); m(1)%nublock = hdr(15); m(1)%seed = hdr(16)
  m(1)%kin_weight = kw
  m(1)%vtab = c_null_ptr; m(1)%vdelta = c_null_ptr
  m(1)%fft_dims = fdims
  m(1)%zplq = 0; m(1)%lqp = 0
  m(1)%plq_cord = c_null_ptr; m(1)%plq_b = c_null_ptr; m(1)%plq_bi = c_null_ptr; m(1)%plq_delta = c_null_ptr
  m(1)%plq_pratio = c_null_ptr; m(1)%plq_parg = c_null_ptr; m(1)%plq_gaml = c_null_ptr

  call sh_check(sh_create(ctx, 1_c_int, m, 0_c_int), 'allocate_core')
  call sh_check(sh_init_conf(ctx), 'input_conf')
  call sh_check(sh_sweep_start_0b(ctx), 'dqmc_ft_sweep_start_0b')
  do isw = 1, nsweep
     call sh_check(sh_sweep_b0(ctx, 1_c_int, 1_c_int), 'dqmc_ft_sweep_b0')
     call sh_check(sh_sweep_0b(ctx, 1_c_int, 1_c_int, 0_c_int), 'dqmc_ft_sweep_0b')
  end do
  call sh_check(sh_get_conf(ctx, 0_c_int, conf), 'get_conf')
  call sh_check(sh_get_g(ctx, 0_c_int, g), 'get_g')
  call sh_check(sh_get_diag(ctx, 0_c_int, diag), 'get_diag')
  call sh_check(sh_get_obs(ctx, 0_c_int, obs), 'get_obs')
  call sh_check(sh_destroy(ctx), 'deallocate_core')

  open(newunit=u, file='out.bin', access='stream', form='unformatted', status='replace')
  write(u) conf, g, diag, obs
  close(u)
  write(*, '(a, i0, a)') ' ft_driver: ', nsweep, ' sweeps done'
end program ft_driver
