!> Population VMC driver in Fortran on top of the GPU engine: the block/step
!> loop of vmc.f90:103-134 with metrop1 moves (metrop_mov1_driftdif.f90) or
!> metrop6 moves (metrop_mov1_slat.f90, CHAMP's default mover, tau unused)
!> executed for all walkers at once through champ_gpu_mod.
!>
!> usage: vmc_gpu <system.bin> <nblk> <nstep> <tau> [seed32 [imetro]]
!> The system file holds the tables of champ_system_t and the initial walker
!> positions (written by champ_amd.system.System.write_binary).
program vmc_gpu
  use iso_c_binding
  use champ_gpu_mod
  implicit none

  integer(c_int32_t) :: hdr(20)
  real(c_double) :: dh(3)
  type(champ_system_t) :: sys
  type(c_ptr) :: ctx
  type(champ_sweep_stats) :: st
  real(c_double), allocatable, target :: cent(:), znuc(:), rwf(:), d2rwf(:), ce(:), ae(:), rmaxwf(:), coef(:), &
       cdet(:), a4(:), b(:), c(:), ecp_coef(:), ecp_exponent(:), xq0(:), yq0(:), zq0(:), wq(:), x0(:)
  integer(c_int32_t), allocatable, target :: iwctype(:), nrbas(:), nbastyp(:), iwrwf(:), iwlbas(:), iworbd(:), &
       lpot(:), necp_term(:), necp_power(:)
  real(c_double), allocatable :: psid(:), psij(:), ekin(:), energy(:)
  character(len=512) :: path, arg
  character(len=32) :: cseed
  integer(c_int32_t) :: irn(8)
  integer :: u, nelec, ncent, nctype, nr, nrbm, nbm, nbasis, norb, ndet, norda, nordb, nordc, nspin2b, nquad
  integer :: nterms, nwalk, nblk, nstep, iblk, imetro
  real(c_double) :: tau, eblk, esum, esum2, acc, t0, t1, moves
  integer(c_int64_t) :: c0, c1, crate

  call get_command_argument(1, path)
  call get_command_argument(2, arg); read(arg, *) nblk
  call get_command_argument(3, arg); read(arg, *) nstep
  call get_command_argument(4, arg); read(arg, *) tau
  cseed = '18123437465549275475255231234865'
  if (command_argument_count() >= 5) call get_command_argument(5, cseed)
  imetro = 1
  if (command_argument_count() >= 6) then
    call get_command_argument(6, arg); read(arg, *) imetro
  end if
  if (imetro /= 1 .and. imetro /= 6) then
    write(*,'(a,i0)') ' vmc_gpu: imetro must be 1 or 6, got ', imetro
    error stop 1
  end if

  open(newunit=u, file=trim(path), access='stream', form='unformatted', status='old')
  read(u) hdr
  read(u) dh
  nelec = hdr(1); ncent = hdr(3); nctype = hdr(4); nr = hdr(5); nrbm = hdr(6); nbm = hdr(7); nbasis = hdr(8)
  norb = hdr(9); ndet = hdr(10); norda = hdr(11); nordb = hdr(12); nordc = hdr(13); nspin2b = hdr(14)
  nquad = hdr(17); nterms = hdr(18); nwalk = hdr(19)
  allocate(cent(3*ncent), iwctype(ncent), znuc(nctype), nrbas(nctype), rwf(nctype*nrbm*nr), d2rwf(nctype*nrbm*nr), &
           ce(nctype*nrbm*CHAMP_NCOEF), ae(nctype*nrbm*2), rmaxwf(nctype*nrbm), nbastyp(nctype), iwrwf(nctype*nbm), &
           iwlbas(nctype*nbm), coef(norb*nbasis), iworbd(ndet*nelec), cdet(ndet), a4(nctype*(norda+1)), &
           b(nspin2b*(nordb+1)), c(max(1, nctype*nterms)), lpot(nctype), necp_term(nctype*CHAMP_MPS_L), &
           necp_power(nctype*CHAMP_MPS_L*CHAMP_MGAUSS), ecp_coef(nctype*CHAMP_MPS_L*CHAMP_MGAUSS), &
           ecp_exponent(nctype*CHAMP_MPS_L*CHAMP_MGAUSS), xq0(max(1, nquad)), yq0(max(1, nquad)), zq0(max(1, nquad)), &
           wq(max(1, nquad)), x0(3*nelec*nwalk))
  read(u) cent, iwctype, znuc, nrbas, rwf, d2rwf, ce, ae, rmaxwf, nbastyp, iwrwf, iwlbas, coef, iworbd, cdet
  read(u) a4, b
  if (nterms > 0) read(u) c
  read(u) lpot, necp_term, necp_power, ecp_coef, ecp_exponent
  if (nquad > 0) read(u) xq0(1:nquad), yq0(1:nquad), zq0(1:nquad), wq(1:nquad)
  read(u) x0
  close(u)

  sys%nelec = hdr(1); sys%nup = hdr(2); sys%ncent = hdr(3); sys%nctype = hdr(4)
  sys%cent = c_loc(cent); sys%iwctype = c_loc(iwctype); sys%znuc = c_loc(znuc)
  sys%nr = hdr(5); sys%grid_r0 = dh(1); sys%grid_arg = dh(2)
  sys%nrbas_max = hdr(6); sys%nrbas = c_loc(nrbas); sys%rwf = c_loc(rwf); sys%d2rwf = c_loc(d2rwf)
  sys%ce = c_loc(ce); sys%ae = c_loc(ae); sys%rmaxwf = c_loc(rmaxwf)
  sys%nbastyp_max = hdr(7); sys%nbastyp = c_loc(nbastyp); sys%iwrwf = c_loc(iwrwf); sys%iwlbas = c_loc(iwlbas)
  sys%nbasis = hdr(8); sys%norb = hdr(9); sys%coef = c_loc(coef)
  sys%ndet = hdr(10); sys%iworbd = c_loc(iworbd); sys%cdet = c_loc(cdet)
  sys%norda = hdr(11); sys%nordb = hdr(12); sys%nordc = hdr(13); sys%nspin2b = hdr(14); sys%nocuspb = hdr(15)
  sys%scalek = dh(3); sys%a4 = c_loc(a4); sys%b = c_loc(b); sys%c = c_loc(c)
  sys%nloc = hdr(16); sys%lpot = c_loc(lpot); sys%necp_term = c_loc(necp_term); sys%necp_power = c_loc(necp_power)
  sys%ecp_coef = c_loc(ecp_coef); sys%ecp_exponent = c_loc(ecp_exponent)
  sys%nquad = hdr(17); sys%xq0 = c_loc(xq0); sys%yq0 = c_loc(yq0); sys%zq0 = c_loc(zq0); sys%wq = c_loc(wq)

  call gpu_check(champ_gpu_init(ctx, sys, int(nwalk, c_int32_t), int(CHAMP_PREC_FP64, c_int32_t)), 'init')
  call gpu_check(champ_gpu_set_walkers(ctx, x0), 'set_walkers')
  ! seed as parser.f90:651-661: 8 x i4 from the 32-character string, jump per walker
  read(cseed, '(8i4)') irn
  call gpu_check(champ_gpu_seed_rng(ctx, irn), 'seed_rng')
  allocate(psid(nwalk), psij(nwalk), ekin(nwalk), energy(nwalk))
  call gpu_hpsi(ctx, psid, psij, ekin, energy)
  write(*,'(a,i0,a,i0,a,i0)') ' vmc_gpu: nelec=', nelec, ' nbasis=', nbasis, ' nwalk=', nwalk
  write(*,'(a,f20.10)') ' initial <E_L> =', sum(energy) / nwalk

  esum = 0; esum2 = 0; acc = 0
  call system_clock(c0, crate)
  do iblk = 1, nblk
    if (imetro == 6) then
      call gpu_check(champ_gpu_metrop6(ctx, int(nstep, c_int32_t), st), 'metrop6')
    else
      call gpu_metrop1(ctx, nstep, tau, st)
    end if
    eblk = st%esum / (nwalk * nstep)
    esum = esum + eblk
    esum2 = esum2 + eblk * eblk
    acc = acc + st%accepted / st%proposed
    write(*,'(a,i6,a,es25.17,a,f8.5,a,i0)') ' block', iblk, '  E_L', eblk, '  acc', st%accepted / st%proposed, &
         '  flagged ', st%nflagged
  end do
  call system_clock(c1)
  t1 = real(c1 - c0, c_double) / real(crate, c_double)
  t0 = esum / nblk
  moves = real(nwalk, c_double) * nelec * nstep * nblk
  write(*,'(a,f22.12,a,es12.4)') ' total E =', t0, ' +- ', sqrt(max(0d0, esum2 / nblk - t0 * t0) / max(1, nblk - 1))
  write(*,'(a,es14.6)') ' walker-moves/s =', moves / t1
  call gpu_check(champ_gpu_free(ctx), 'free')
end program vmc_gpu
