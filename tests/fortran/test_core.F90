! test_core.F90 -- runs OpenQG's time loop shape through the drop-in `module
! core` (openqg_amd/fortran/core.F90): step_qg on a qg_type filled from an
! input file, with suppress_comp_mode every `supp` steps, the way
! driver.F90:73-81 and :107 call them.
!
! Usage: test_core <input> <output>.  The input holds the basin (sizes,
! parameters, init_qg's precompute), its state and the forcing, as raw
! native-endian records (layout in tests/test_gpu_fortran_core.py); the output
! holds q, qm, p, pm and the constraint scalars after the run.  With
! OQG_RESIDENT=1 the device state is authoritative: suppression runs on the
! device (qg_suppress_device) and the state is fetched once at the end
! (qg_to_host); otherwise the host-side suppress_comp_mode of driver.F90:174-189
! runs between steps on qg's own arrays.
program test_core

  use qg, only: qg_type
  use core, only: step_qg, qg_to_host, qg_suppress_device

  implicit none

  type(qg_type), target :: g
  integer :: hdr(9), nxp, nyp, nl, nsteps, supp, nt, m
  double precision :: par(9), tdt, txis, txin
  double precision, allocatable, target :: ent(:,:,:), wekp(:,:), coef(:)
  character(len=512) :: inpath, outpath
  character(len=16) :: env
  integer :: envlen, envstat
  logical :: resident

  call get_command_argument(1, inpath)
  call get_command_argument(2, outpath)
  call get_environment_variable('OQG_RESIDENT', env, envlen, envstat)
  resident = (envstat == 0 .and. envlen > 0 .and. env(1:1) == '1')

  open(10, file=trim(inpath), access='stream', form='unformatted', status='old')
  read(10) hdr
  nxp = hdr(1); nyp = hdr(2); nl = hdr(3)
  nsteps = hdr(5); supp = hdr(6)
  read(10) par
  tdt = par(7); txis = par(8); txin = par(9)

  g%b%nxp = nxp; g%b%nyp = nyp; g%b%nl = nl
  g%b%nxt = nxp - 1; g%b%nyt = nyp - 1
  g%b%cyclic = (hdr(4) /= 0)
  g%topo%k_topo = hdr(7)
  g%tau_sign = hdr(8)
  g%b%dz_sign = hdr(9)
  g%b%dx = par(1); g%b%dy = par(2); g%b%fnot = par(3); g%b%beta = par(4)
  g%bcco = par(5); g%delek = par(6)
  allocate(g%b%h(nl), g%gp(nl-1), g%ah2(nl), g%ah4(nl))
  read(10) g%b%h, g%gp, g%ah2, g%ah4
  allocate(g%topo%ddyn(nxp,nyp))
  read(10) g%topo%ddyn
  allocate(g%mod%amat(nl,nl), g%mod%ctl2m(nl,nl), g%mod%ctm2l(nl,nl), g%mod%rdm2(nl))
  read(10) g%mod%amat, g%mod%ctl2m, g%mod%ctm2l, g%mod%rdm2
  if (g%b%cyclic) then
     allocate(g%hom%cyc%hom_sol_bt(nyp), g%hom%cyc%hom_sol_bc1(nyp,2:nl), g%hom%cyc%hom_sol_bc2(nyp,2:nl))
     allocate(g%hom%cyc%hc1s(2:nl), g%hom%cyc%hc2s(2:nl), g%hom%cyc%hc1n(2:nl), g%hom%cyc%hc2n(2:nl))
     allocate(coef(4*(nl-1)+1))
     read(10) g%hom%cyc%hom_sol_bt, g%hom%cyc%hom_sol_bc1, g%hom%cyc%hom_sol_bc2, coef
     do m = 2, nl
        g%hom%cyc%hc1s(m) = coef(4*(m-2)+1)
        g%hom%cyc%hc2s(m) = coef(4*(m-2)+2)
        g%hom%cyc%hc1n(m) = coef(4*(m-2)+3)
        g%hom%cyc%hc2n(m) = coef(4*(m-2)+4)
     enddo
     g%hom%cyc%hbsi = coef(4*(nl-1)+1)
  else
     allocate(g%hom%box%hom_sol_bc(nxp,nyp,2:nl))
     read(10) g%hom%box%hom_sol_bc
  endif
  allocate(g%q(nxp,nyp,nl), g%qm(nxp,nyp,nl), g%p(nxp,nyp,nl), g%pm(nxp,nyp,nl))
  read(10) g%q, g%qm, g%p, g%pm
  allocate(g%con%dpi(nl-1), g%con%dpip(nl-1))
  read(10) g%con%dpi, g%con%dpip
  allocate(g%constr%cs(nl), g%constr%cn(nl), g%constr%csp(nl), g%constr%cnp(nl))
  if (g%b%cyclic) read(10) g%constr%cs, g%constr%cn, g%constr%csp, g%constr%cnp
  allocate(ent(nxp,nyp,nl), wekp(nxp,nyp))
  read(10) ent, wekp
  close(10)
  g%active = .true.

  do nt = 1, nsteps
     call step_qg(tdt, ent, wekp, txis, txin, g)
     if (supp > 0 .and. mod(nt - 1, supp) == 0) then
        if (resident) then
           call qg_suppress_device(g)
        else
           call suppress_comp_mode(g)
        endif
     endif
  enddo
  if (resident) call qg_to_host(g)

  open(11, file=trim(outpath), access='stream', form='unformatted', status='replace')
  write(11) g%q, g%qm, g%p, g%pm, g%con%dpi, g%con%dpip
  if (g%b%cyclic) write(11) g%constr%cs, g%constr%cn, g%constr%csp, g%constr%cnp
  close(11)

contains

  ! the host-side averaging of time levels (driver.F90:174-189)
  subroutine suppress_comp_mode(s)
    type(qg_type), intent(inout) :: s
    s%q = 0.5d0*(s%q + s%qm)
    s%p = 0.5d0*(s%p + s%pm)
    s%con%dpi = 0.5d0*(s%con%dpi + s%con%dpip)
    if (s%b%cyclic) then
       s%constr%cs = 0.5d0*(s%constr%cs + s%constr%csp)
       s%constr%cn = 0.5d0*(s%constr%cn + s%constr%cnp)
    endif
  end subroutine suppress_comp_mode

end program test_core
