!> Prints the byte offset of every component of the bind(c) types of
!> champ_gpu_mod and their sizes, one "name offset" line each, for comparison
!> with offsetof() of the C declarations in include/champ_gpu.h.
program layout_check
  use iso_c_binding
  use champ_gpu_mod
  implicit none
  type(champ_system_t), target :: s
  type(champ_sweep_stats), target :: st
  type(champ_vmc_result), target :: v
  type(champ_dmc_step_stats), target :: d
  type(champ_dmc_result), target :: r
  type(champ_sr_info), target :: q
  integer(c_intptr_t) :: b

  b = transfer(c_loc(s), b)
  call p('nelec', c_loc(s%nelec)); call p('nup', c_loc(s%nup)); call p('ncent', c_loc(s%ncent))
  call p('nctype', c_loc(s%nctype)); call p('cent', c_loc(s%cent)); call p('iwctype', c_loc(s%iwctype))
  call p('znuc', c_loc(s%znuc)); call p('nr', c_loc(s%nr)); call p('grid_r0', c_loc(s%grid_r0))
  call p('grid_arg', c_loc(s%grid_arg)); call p('nrbas_max', c_loc(s%nrbas_max)); call p('nrbas', c_loc(s%nrbas))
  call p('rwf', c_loc(s%rwf)); call p('d2rwf', c_loc(s%d2rwf)); call p('ce', c_loc(s%ce)); call p('ae', c_loc(s%ae))
  call p('rmaxwf', c_loc(s%rmaxwf)); call p('nbastyp_max', c_loc(s%nbastyp_max)); call p('nbastyp', c_loc(s%nbastyp))
  call p('iwrwf', c_loc(s%iwrwf)); call p('iwlbas', c_loc(s%iwlbas)); call p('nbasis', c_loc(s%nbasis))
  call p('norb', c_loc(s%norb)); call p('coef', c_loc(s%coef)); call p('ndet', c_loc(s%ndet))
  call p('iworbd', c_loc(s%iworbd)); call p('cdet', c_loc(s%cdet)); call p('norda', c_loc(s%norda))
  call p('nordb', c_loc(s%nordb)); call p('nordc', c_loc(s%nordc)); call p('nspin2b', c_loc(s%nspin2b))
  call p('nocuspb', c_loc(s%nocuspb)); call p('scalek', c_loc(s%scalek)); call p('a4', c_loc(s%a4))
  call p('b', c_loc(s%b)); call p('c', c_loc(s%c)); call p('nloc', c_loc(s%nloc)); call p('lpot', c_loc(s%lpot))
  call p('necp_term', c_loc(s%necp_term)); call p('necp_power', c_loc(s%necp_power))
  call p('ecp_coef', c_loc(s%ecp_coef)); call p('ecp_exponent', c_loc(s%ecp_exponent))
  call p('nquad', c_loc(s%nquad)); call p('xq0', c_loc(s%xq0)); call p('yq0', c_loc(s%yq0))
  call p('zq0', c_loc(s%zq0)); call p('wq', c_loc(s%wq))
  write(*,'(a,1x,i0)') 'sizeof_system', c_sizeof(s)

  b = transfer(c_loc(st), b)
  call p('stats.accepted', c_loc(st%accepted)); call p('stats.proposed', c_loc(st%proposed))
  call p('stats.esum', c_loc(st%esum)); call p('stats.esum2', c_loc(st%esum2))
  call p('stats.nflagged', c_loc(st%nflagged))
  write(*,'(a,1x,i0)') 'sizeof_stats', c_sizeof(st)

  b = transfer(c_loc(v), b)
  call p('vmc.efin', c_loc(v%efin)); call p('vmc.eerr', c_loc(v%eerr)); call p('vmc.sigma', c_loc(v%sigma))
  call p('vmc.tcorr', c_loc(v%tcorr)); call p('vmc.accept', c_loc(v%accept))
  call p('vmc.nsamples', c_loc(v%nsamples)); call p('vmc.nflagged', c_loc(v%nflagged))
  write(*,'(a,1x,i0)') 'sizeof_vmc', c_sizeof(v)

  b = transfer(c_loc(d), b)
  call p('dmcstep.nwalk', c_loc(d%nwalk)); call p('dmcstep.pad', c_loc(d%pad)); call p('dmcstep.wsum', c_loc(d%wsum))
  call p('dmcstep.esum', c_loc(d%esum)); call p('dmcstep.wgsum', c_loc(d%wgsum))
  call p('dmcstep.egsum', c_loc(d%egsum)); call p('dmcstep.eest', c_loc(d%eest))
  call p('dmcstep.eigv', c_loc(d%eigv)); call p('dmcstep.fprod', c_loc(d%fprod))
  call p('dmcstep.accept', c_loc(d%accept)); call p('dmcstep.nodecr', c_loc(d%nodecr))
  call p('dmcstep.nbranch', c_loc(d%nbranch))
  write(*,'(a,1x,i0)') 'sizeof_dmcstep', c_sizeof(d)

  b = transfer(c_loc(r), b)
  call p('dmcres.egave', c_loc(r%egave)); call p('dmcres.egerr', c_loc(r%egerr)); call p('dmcres.eest', c_loc(r%eest))
  call p('dmcres.accept', c_loc(r%accept)); call p('dmcres.tau_eff', c_loc(r%tau_eff))
  call p('dmcres.nwalk', c_loc(r%nwalk)); call p('dmcres.nwalk_max', c_loc(r%nwalk_max))
  call p('dmcres.nbranch', c_loc(r%nbranch)); call p('dmcres.nodecr', c_loc(r%nodecr))
  write(*,'(a,1x,i0)') 'sizeof_dmcres', c_sizeof(r)

  b = transfer(c_loc(q), b)
  call p('sr.adiag', c_loc(q%adiag)); call p('sr.dparm_norm', c_loc(q%dparm_norm))
  call p('sr.cg_iters', c_loc(q%cg_iters)); call p('sr.retries', c_loc(q%retries))
  write(*,'(a,1x,i0)') 'sizeof_sr', c_sizeof(q)

contains

  subroutine p(name, a)
    character(len=*), intent(in) :: name
    type(c_ptr), intent(in) :: a
    write(*,'(a,1x,i0)') name, transfer(a, b) - b
  end subroutine

end program layout_check
