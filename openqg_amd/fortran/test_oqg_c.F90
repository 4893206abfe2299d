! test_oqg_c.F90 -- drives liboqg.so from Fortran through oqg_c (ISO_C_BINDING).
!
! Builds a 65x65, 3-layer double-gyre box, sets p = pm to a smooth analytic
! field, derives q (init_pv), steps 10 times with the leapfrog suppression at
! step 1, and writes the inputs p, wekp and then q, qm, p, pm (raw FP64,
! Fortran order) to the file named on the command line.
! tests/test_gpu_fortran.py runs the same case through the Python binding and
! checks the two agree bitwise.
program test_oqg_c

  use, intrinsic :: iso_c_binding
  use oqg_c

  implicit none

  integer, parameter :: nxp = 65, nyp = 65, nl = 3
  double precision, parameter :: dx = 30.0d3, pi = 3.14159265358979324d0
  type(oqg_desc) :: d
  type(c_ptr) :: ctx
  double precision, target :: q(nxp,nyp,nl), qm(nxp,nyp,nl), p(nxp,nyp,nl), pm(nxp,nyp,nl)
  double precision, target :: ent(nxp,nyp,nl-1), wekp(nxp,nyp)
  integer :: i, j, k, nt, rc
  character(len=512) :: outpath

  call get_command_argument(1, outpath)

  d%nxp = nxp
  d%nyp = nyp
  d%nl = nl
  d%cyclic = 0
  d%dx = dx
  d%dy = dx
  d%fnot = 1.0d-4
  d%beta = 2.0d-11
  d%h(1:3) = (/300.0d0, 1100.0d0, 2600.0d0/)
  d%gp(1:2) = (/0.05d0, 0.025d0/)
  d%ah2(1:3) = 0.0d0
  d%ah4(1:3) = 2.0d10
  d%bcco = 0.5d0
  d%delek = 2.0d0
  d%dz_sign = -1
  d%tau_sign = 1

  rc = oqg_create(d, ctx)
  if (rc /= 0) then
     print *, 'oqg_create failed: ', trim(oqg_error_message(c_null_ptr))
     stop 1
  endif

  do k = 1, nl
     do j = 1, nyp
        do i = 1, nxp
           p(i,j,k) = 0.1d0*dble(4-k)*sin(pi*dble(i-1)/dble(nxp-1))*sin(2.0d0*pi*dble(j-1)/dble(nyp-1))
        enddo
     enddo
  enddo
  pm = p
  q = 0.0d0
  qm = 0.0d0
  ent = 0.0d0
  do j = 1, nyp
     do i = 1, nxp
        wekp(i,j) = 1.0d-6*sin(2.0d0*pi*dble(j-1)/dble(nyp-1))
     enddo
  enddo

  open(unit=10, file=trim(outpath), access='stream', form='unformatted', status='replace')
  write(10) p, wekp

  call ok(oqg_set_state(ctx, c_loc(q), c_loc(qm), c_loc(p), c_loc(pm), OQG_HOST))
  call ok(oqg_init_state(ctx))
  do nt = 1, 10
     call ok(oqg_step(ctx, 7200.0d0, c_loc(ent), int(nl-1, c_int), c_loc(wekp), 0.0d0, 0.0d0))
     if (nt == 1) call ok(oqg_suppress(ctx))
  enddo
  call ok(oqg_get_state(ctx, c_loc(q), c_loc(qm), c_loc(p), c_loc(pm), OQG_HOST))
  call oqg_destroy(ctx)

  write(10) q, qm, p, pm
  close(10)
  print '(a,es24.16)', 'fortran oqg: max|p| = ', maxval(abs(p))

contains

  subroutine ok(code)
    integer(c_int), intent(in) :: code
    if (code /= 0) then
       print *, 'oqg error ', code, ': ', trim(oqg_error_message(ctx))
       stop 1
    endif
  end subroutine ok

end program test_oqg_c
