!> Fortran binding of the MI355X engine (include/champ_gpu.h) for CHAMP's
!> drivers.  The interfaces are iso_c_binding declarations in the style of
!> CHAMP's own binding of its xoshiro C file (src/vmc/random.f90:3-22); the
!> wrappers keep the names of the routines they replace (hpsi, psie,
!> compute_determinante_grad, metrop1, rotqua) with a leading engine context
!> that holds the whole walker population on the GPU.
module champ_gpu_mod
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: CHAMP_PREC_FP64 = 0
  integer, parameter, public :: CHAMP_MPS_L = 5, CHAMP_MGAUSS = 8, CHAMP_NCOEF = 5

  !> champ_system_t: read-only tables of one system (pointers to caller arrays)
  type, bind(c), public :: champ_system_t
    integer(c_int32_t) :: nelec, nup, ncent, nctype
    type(c_ptr) :: cent, iwctype, znuc
    integer(c_int32_t) :: nr
    real(c_double) :: grid_r0, grid_arg
    integer(c_int32_t) :: nrbas_max
    type(c_ptr) :: nrbas, rwf, d2rwf, ce, ae, rmaxwf
    integer(c_int32_t) :: nbastyp_max
    type(c_ptr) :: nbastyp, iwrwf, iwlbas
    integer(c_int32_t) :: nbasis
    integer(c_int32_t) :: norb
    type(c_ptr) :: coef
    integer(c_int32_t) :: ndet
    type(c_ptr) :: iworbd, cdet
    integer(c_int32_t) :: norda, nordb, nordc, nspin2b, nocuspb
    real(c_double) :: scalek
    type(c_ptr) :: a4, b, c
    integer(c_int32_t) :: nloc
    type(c_ptr) :: lpot, necp_term, necp_power, ecp_coef, ecp_exponent
    integer(c_int32_t) :: nquad
    type(c_ptr) :: xq0, yq0, zq0, wq
  end type

  type, bind(c), public :: champ_sweep_stats
    real(c_double) :: accepted, proposed, esum, esum2
    integer(c_int64_t) :: nflagged
  end type

  type, bind(c), public :: champ_vmc_result
    real(c_double) :: efin, eerr, sigma, tcorr, accept
    integer(c_int64_t) :: nsamples, nflagged
  end type

  type, bind(c), public :: champ_dmc_step_stats
    integer(c_int32_t) :: nwalk, pad
    real(c_double) :: wsum, esum, wgsum, egsum, eest, eigv, fprod, accept
    integer(c_int64_t) :: nodecr, nbranch
  end type

  type, bind(c), public :: champ_sr_info
    real(c_double) :: adiag, dparm_norm
    integer(c_int32_t) :: cg_iters, retries
  end type

  type, bind(c), public :: champ_dmc_result
    real(c_double) :: egave, egerr, eest, accept, tau_eff
    integer(c_int32_t) :: nwalk, nwalk_max
    integer(c_int64_t) :: nbranch, nodecr
  end type

  interface
    integer(c_int) function champ_gpu_init(ctx, sys, nwalk, precision) bind(c, name="champ_gpu_init")
      import :: c_int, c_ptr, c_int32_t, champ_system_t
      type(c_ptr), intent(out) :: ctx
      type(champ_system_t), intent(in) :: sys
      integer(c_int32_t), value :: nwalk, precision
    end function
    integer(c_int) function champ_gpu_free(ctx) bind(c, name="champ_gpu_free")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function champ_gpu_last_error() bind(c, name="champ_gpu_last_error")
      import :: c_ptr
    end function
    integer(c_int) function champ_gpu_set_walkers(ctx, x) bind(c, name="champ_gpu_set_walkers")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: x(*)
    end function
    integer(c_int) function champ_gpu_get_walkers(ctx, x) bind(c, name="champ_gpu_get_walkers")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: x(*)
    end function
    integer(c_int) function champ_gpu_seed_rng(ctx, seed) bind(c, name="champ_gpu_seed_rng")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: seed(8)
    end function
    integer(c_int) function champ_gpu_rotqua(ctx) bind(c, name="champ_gpu_rotqua")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function champ_gpu_hpsi(ctx, psid, psij, ekin, energy) bind(c, name="champ_gpu_hpsi")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: psid(*), psij(*), ekin(*), energy(*)
    end function
    integer(c_int) function champ_gpu_compute_determinante_grad(ctx, iel, iflag_move, vd) &
        bind(c, name="champ_gpu_compute_determinante_grad")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: iel, iflag_move
      real(c_double), intent(out) :: vd(*)
    end function
    integer(c_int) function champ_gpu_psie(ctx, iel, xnew, psidn, psijn, iflag) bind(c, name="champ_gpu_psie")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: iel, iflag
      real(c_double), intent(in) :: xnew(*)
      real(c_double), intent(out) :: psidn(*), psijn(*)
    end function
    integer(c_int) function champ_gpu_accept(ctx, iel, accept) bind(c, name="champ_gpu_accept")
      import :: c_int, c_ptr, c_int32_t, c_int8_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: iel
      integer(c_int8_t), intent(in) :: accept(*)
    end function
    integer(c_int) function champ_gpu_metrop1(ctx, nsweep, tau, stats) bind(c, name="champ_gpu_metrop1")
      import :: c_int, c_ptr, c_int32_t, c_double, champ_sweep_stats
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsweep
      real(c_double), value :: tau
      type(champ_sweep_stats), intent(out) :: stats
    end function
    integer(c_int) function champ_gpu_metrop1_batch(ctxs, n, nsweep, tau, stats) &
        bind(c, name="champ_gpu_metrop1_batch")
      import :: c_int, c_ptr, c_int32_t, c_double, champ_sweep_stats
      type(c_ptr), intent(in) :: ctxs(*)
      integer(c_int32_t), value :: n, nsweep
      real(c_double), value :: tau
      type(champ_sweep_stats), intent(out) :: stats(*)
    end function
    integer(c_int) function champ_gpu_autotune(ctx, choices) bind(c, name="champ_gpu_autotune")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: choices(4)
    end function
    integer(c_int) function champ_gpu_set_graphs(ctx, on) bind(c, name="champ_gpu_set_graphs")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function champ_gpu_get_walker_scalars(ctx, psido, psijo, eold, psi2o) &
        bind(c, name="champ_gpu_get_walker_scalars")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: psido(*), psijo(*), eold(*), psi2o(*)
    end function
    integer(c_int) function champ_gpu_metrop6(ctx, nsweep, stats) bind(c, name="champ_gpu_metrop6")
      import :: c_int, c_ptr, c_int32_t, champ_sweep_stats
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsweep
      type(champ_sweep_stats), intent(out) :: stats
    end function
    integer(c_int) function champ_gpu_set_metrop6(ctx, deltar, deltat, fbias) bind(c, name="champ_gpu_set_metrop6")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltar, deltat, fbias
    end function
    integer(c_int) function champ_gpu_vmc(ctx, nblk, nstep, nblkeq, tau, imetro, enow, res) &
        bind(c, name="champ_gpu_vmc")
      import :: c_int, c_ptr, c_int32_t, c_double, champ_vmc_result
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nblk, nstep, nblkeq
      real(c_double), value :: tau
      integer(c_int32_t), value :: imetro
      real(c_double), intent(out) :: enow(*)
      type(champ_vmc_result), intent(out) :: res
    end function
    integer(c_int) function champ_gpu_dmc_init(ctx, nconf, tau, etrial, nfprod, pop_state) &
        bind(c, name="champ_gpu_dmc_init")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nconf
      real(c_double), value :: tau, etrial
      integer(c_int32_t), value :: nfprod
      integer(c_int64_t), intent(in) :: pop_state(4)
    end function
    integer(c_int) function champ_gpu_dmc_step(ctx, stats) bind(c, name="champ_gpu_dmc_step")
      import :: c_int, c_ptr, champ_dmc_step_stats
      type(c_ptr), value :: ctx
      type(champ_dmc_step_stats), intent(out) :: stats
    end function
    integer(c_int) function champ_gpu_dmc(ctx, nblk, nstep, nblkeq, egnow, res) bind(c, name="champ_gpu_dmc")
      import :: c_int, c_ptr, c_int32_t, c_double, champ_dmc_result
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nblk, nstep, nblkeq
      real(c_double), intent(out) :: egnow(*)
      type(champ_dmc_result), intent(out) :: res
    end function
    integer(c_int) function champ_gpu_dmc_set_icasula(ctx, icasula) bind(c, name="champ_gpu_dmc_set_icasula")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: icasula
    end function
    integer(c_int) function champ_gpu_dmc_weights(ctx, wt, nwalk) bind(c, name="champ_gpu_dmc_weights")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: wt(*)
      integer(c_int32_t), intent(out) :: nwalk
    end function
    integer(c_int) function champ_gpu_deriv_count(ctx, kinds, nparm) bind(c, name="champ_gpu_deriv_count")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kinds
      integer(c_int32_t), intent(out) :: nparm
    end function
    integer(c_int) function champ_gpu_deriv_psi(ctx, kinds, o) bind(c, name="champ_gpu_deriv_psi")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kinds
      real(c_double), intent(out) :: o(*)        ! (nparm, nwalk)
    end function
    integer(c_int) function champ_gpu_sr_sums(ctx, kinds, s, osum, eosum, esum) bind(c, name="champ_gpu_sr_sums")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kinds
      real(c_double), intent(out) :: s(*), osum(*), eosum(*)
      real(c_double), intent(out) :: esum
    end function
    integer(c_int) function champ_gpu_set_deriv_map(ctx, kinds, nout, map) bind(c, name="champ_gpu_set_deriv_map")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kinds, nout
      real(c_double), intent(in) :: map(*)      ! (nout, nraw)
    end function
    integer(c_int) function champ_gpu_set_orb_terms(ctx, nterm, ideriv) bind(c, name="champ_gpu_set_orb_terms")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nterm
      integer(c_int32_t), intent(in) :: ideriv(2, *)   ! (io, jo), 0-based
    end function
    integer(c_int) function champ_gpu_set_orbitals(ctx, coef) bind(c, name="champ_gpu_set_orbitals")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: coef(*)          ! (nbasis, norb)
    end function
    integer(c_int) function champ_gpu_set_params(ctx, a4, b, c, cdet) bind(c, name="champ_gpu_set_params")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: a4, b, c, cdet       ! c_loc of the arrays, or c_null_ptr to keep
    end function
    integer(c_int) function champ_sr_solve(nparm, wsum, s, osum, eosum, esum, sr_tau, sr_eps, sr_adiag, &
        dparm_norm_min, dparm, info) bind(c, name="champ_sr_solve")
      import :: c_int, c_int32_t, c_double, champ_sr_info
      integer(c_int32_t), value :: nparm
      real(c_double), value :: wsum
      real(c_double), intent(in) :: s(*), osum(*), eosum(*)
      real(c_double), value :: esum, sr_tau, sr_eps, sr_adiag, dparm_norm_min
      real(c_double), intent(out) :: dparm(*)
      type(champ_sr_info), intent(out) :: info
    end function
    integer(c_int) function champ_gpu_sr_accumulate(ctx, kinds, reset, wsum, esum) &
        bind(c, name="champ_gpu_sr_accumulate")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kinds, reset
      real(c_double), intent(out) :: wsum, esum
    end function
    integer(c_int) function champ_gpu_sr_solve(ctx, sr_tau, sr_eps, sr_adiag, dparm_norm_min, dparm, info) &
        bind(c, name="champ_gpu_sr_solve")
      import :: c_int, c_ptr, c_double, champ_sr_info
      type(c_ptr), value :: ctx
      real(c_double), value :: sr_tau, sr_eps, sr_adiag, dparm_norm_min
      real(c_double), intent(out) :: dparm(*)
      type(champ_sr_info), intent(out) :: info
    end function
  end interface

  public :: champ_gpu_init, champ_gpu_free, champ_gpu_set_walkers, champ_gpu_get_walkers, champ_gpu_seed_rng
  public :: champ_gpu_rotqua, champ_gpu_hpsi, champ_gpu_compute_determinante_grad, champ_gpu_psie
  public :: champ_gpu_accept, champ_gpu_metrop1, champ_gpu_get_walker_scalars, champ_gpu_autotune
  public :: champ_gpu_set_graphs, champ_gpu_metrop1_batch, champ_gpu_metrop6, champ_gpu_set_metrop6
  public :: champ_gpu_vmc, champ_gpu_dmc_init, champ_gpu_dmc_step, champ_gpu_dmc, champ_gpu_dmc_weights
  public :: champ_gpu_dmc_set_icasula
  public :: champ_gpu_deriv_count, champ_gpu_deriv_psi, champ_gpu_sr_sums, champ_gpu_set_deriv_map
  public :: champ_gpu_set_params, champ_sr_solve, champ_gpu_set_orb_terms, champ_gpu_set_orbitals
  public :: champ_gpu_sr_accumulate, champ_gpu_sr_solve
  integer(c_int32_t), parameter, public :: CHAMP_DERIV_JAS_EN = 1, CHAMP_DERIV_JAS_EE = 2
  integer(c_int32_t), parameter, public :: CHAMP_DERIV_JAS_EEN = 4, CHAMP_DERIV_CI = 8, CHAMP_DERIV_MAPPED = 16
  integer(c_int32_t), parameter, public :: CHAMP_DERIV_ORB = 32
  public :: gpu_check, gpu_hpsi, gpu_psie, gpu_compute_determinante_grad, gpu_metrop1

contains

  !> Stop with the engine's message on a non-zero return code (CHAMP aborts
  !> through fatal_error on input/linear-algebra failures, m_error.f90).
  subroutine gpu_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(champ_gpu_last_error(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n+1) == c_null_char) exit
      n = n + 1
    end do
    write(*,'(a,a,i0,a,512a)') what, ' failed, rc=', rc, ': ', msg(1:n)
    error stop 1
  end subroutine

  !> hpsi for every walker (hpsi.f90:3-273), energies per walker
  subroutine gpu_hpsi(ctx, psid, psij, ekin, energy)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out) :: psid(:), psij(:), ekin(:), energy(:)
    call gpu_check(champ_gpu_hpsi(ctx, psid, psij, ekin, energy), 'hpsi')
  end subroutine

  !> psie for electron iel (1-based as in CHAMP) of every walker (hpsie.f90:3-79)
  subroutine gpu_psie(ctx, iel, xnew, psidn, psijn)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: iel
    real(c_double), intent(in) :: xnew(:,:)     ! (3, nwalk)
    real(c_double), intent(out) :: psidn(:), psijn(:)
    call gpu_check(champ_gpu_psie(ctx, int(iel-1, c_int32_t), xnew, psidn, psijn, 0_c_int32_t), 'psie')
  end subroutine

  !> compute_determinante_grad (determinante.f90:69-283), 1-based iel
  subroutine gpu_compute_determinante_grad(ctx, iel, iflag_move, vd)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: iel, iflag_move
    real(c_double), intent(out) :: vd(:,:)      ! (3, nwalk)
    call gpu_check(champ_gpu_compute_determinante_grad(ctx, int(iel-1, c_int32_t), int(iflag_move, c_int32_t), vd), &
                   'compute_determinante_grad')
  end subroutine

  !> nsweep metrop1 sweeps of the population (metrop_mov1_driftdif.f90:4-363)
  subroutine gpu_metrop1(ctx, nsweep, tau, stats)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nsweep
    real(c_double), intent(in) :: tau
    type(champ_sweep_stats), intent(out) :: stats
    call gpu_check(champ_gpu_metrop1(ctx, int(nsweep, c_int32_t), tau, stats), 'metrop1')
  end subroutine

end module champ_gpu_mod
