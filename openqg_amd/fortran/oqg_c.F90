! oqg_c.F90 -- ISO_C_BINDING interface to liboqg.so (include/oqg.h).
!
! Fortran callers (OpenQG's driver through the replacement `core` module in
! core.F90, or any other program) reach the HIP QG step through these
! bindings.  All fields are Fortran (nxp,nyp,nl) double precision arrays,
! passed by address.
module oqg_c

  use, intrinsic :: iso_c_binding

  implicit none

  private

  integer(c_int), parameter, public :: OQG_MAX_LAYERS = 8
  integer(c_int), parameter, public :: OQG_HOST = 0, OQG_DEVICE = 1
  integer(c_int), parameter, public :: OQG_FIELD_Q = 0, OQG_FIELD_QM = 1, OQG_FIELD_P = 2, OQG_FIELD_PM = 3
  integer(c_int), parameter, public :: OQG_FLAG_YSOLVE_THOMAS = 1

  ! Mirror of oqg_desc
  type, bind(C), public :: oqg_desc
     integer(c_int) :: nxp = 0, nyp = 0, nl = 0, cyclic = 0
     real(c_double) :: dx = 0.0d0, dy = 0.0d0, fnot = 0.0d0, beta = 0.0d0
     real(c_double) :: h(OQG_MAX_LAYERS) = 0.0d0
     real(c_double) :: gp(OQG_MAX_LAYERS) = 0.0d0
     real(c_double) :: ah2(OQG_MAX_LAYERS) = 0.0d0
     real(c_double) :: ah4(OQG_MAX_LAYERS) = 0.0d0
     real(c_double) :: bcco = 0.0d0, delek = 0.0d0
     integer(c_int) :: dz_sign = 0, tau_sign = 0, k_topo = 0, flags = 0
     type(c_ptr) :: ddyn = c_null_ptr
     type(c_ptr) :: amat = c_null_ptr, ctl2m = c_null_ptr, ctm2l = c_null_ptr, rdm2 = c_null_ptr
     type(c_ptr) :: hom_sol_bc = c_null_ptr, hom_bt = c_null_ptr, hom_bc1 = c_null_ptr
     type(c_ptr) :: hom_bc2 = c_null_ptr, hom_coef = c_null_ptr
  end type oqg_desc

  public oqg_create, oqg_destroy, oqg_last_error, oqg_set_state, oqg_get_state, oqg_get_layer
  public oqg_constr_len, oqg_set_constr, oqg_get_constr, oqg_init_state, oqg_set_forcing
  public oqg_step, oqg_suppress, oqg_run, oqg_sync, oqg_solve_inhomog, oqg_get_modes
  public oqg_error_message

  interface
     integer(c_int) function oqg_create(desc, ctx) bind(C, name="oqg_create")
       import :: c_int, c_ptr, oqg_desc
       type(oqg_desc), intent(in) :: desc
       type(c_ptr), intent(out) :: ctx
     end function oqg_create

     subroutine oqg_destroy(ctx) bind(C, name="oqg_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine oqg_destroy

     type(c_ptr) function oqg_last_error(ctx) bind(C, name="oqg_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function oqg_last_error

     integer(c_int) function oqg_set_state(ctx, q, qm, p, pm, kind) bind(C, name="oqg_set_state")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, q, qm, p, pm
       integer(c_int), value :: kind
     end function oqg_set_state

     integer(c_int) function oqg_get_state(ctx, q, qm, p, pm, kind) bind(C, name="oqg_get_state")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, q, qm, p, pm
       integer(c_int), value :: kind
     end function oqg_get_state

     integer(c_int) function oqg_get_layer(ctx, which, k, out, kind) bind(C, name="oqg_get_layer")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, out
       integer(c_int), value :: which, k, kind
     end function oqg_get_layer

     integer(c_int) function oqg_constr_len(ctx) bind(C, name="oqg_constr_len")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function oqg_constr_len

     integer(c_int) function oqg_set_constr(ctx, v) bind(C, name="oqg_set_constr")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: v(*)
     end function oqg_set_constr

     integer(c_int) function oqg_get_constr(ctx, v) bind(C, name="oqg_get_constr")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: v(*)
     end function oqg_get_constr

     integer(c_int) function oqg_init_state(ctx) bind(C, name="oqg_init_state")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function oqg_init_state

     integer(c_int) function oqg_set_forcing(ctx, ent, n_ent, wekp, txis, txin, kind) &
          bind(C, name="oqg_set_forcing")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, ent, wekp
       integer(c_int), value :: n_ent, kind
       real(c_double), value :: txis, txin
     end function oqg_set_forcing

     integer(c_int) function oqg_step(ctx, tdt, ent, n_ent, wekp, txis, txin) bind(C, name="oqg_step")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, ent, wekp
       real(c_double), value :: tdt, txis, txin
       integer(c_int), value :: n_ent
     end function oqg_step

     integer(c_int) function oqg_suppress(ctx) bind(C, name="oqg_suppress")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function oqg_suppress

     integer(c_int) function oqg_run(ctx, tdt, nsteps, step0, supp_every) bind(C, name="oqg_run")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: tdt
       integer(c_int), value :: nsteps, step0, supp_every
     end function oqg_run

     integer(c_int) function oqg_sync(ctx) bind(C, name="oqg_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function oqg_sync

     integer(c_int) function oqg_solve_inhomog(ctx, m, rhs, out, kind) bind(C, name="oqg_solve_inhomog")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, rhs, out
       integer(c_int), value :: m, kind
     end function oqg_solve_inhomog

     integer(c_int) function oqg_get_modes(ctx, amat, ctl2m, ctm2l, rdm2) bind(C, name="oqg_get_modes")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, amat, ctl2m, ctm2l, rdm2
     end function oqg_get_modes
  end interface

contains

  ! Message of the last failure as a Fortran string.
  function oqg_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: chars(:)
    type(c_ptr) :: cmsg
    integer :: i
    msg = ' '
    cmsg = oqg_last_error(ctx)
    if (.not. c_associated(cmsg)) return
    call c_f_pointer(cmsg, chars, [256])
    do i = 1, 256
       if (chars(i) == c_null_char) exit
       msg(i:i) = chars(i)
    enddo
  end function oqg_error_message

end module oqg_c
