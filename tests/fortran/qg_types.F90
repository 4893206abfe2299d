! qg_types.F90 -- test-only stand-in for OpenQG's `module qg` and the types
! qg_type nests, so that openqg_amd/fortran/core.F90 (the drop-in
! `module core`) compiles and runs outside the reference tree.
!
! Component names, kinds and ranks follow the reference's declarations, which
! core.F90 reads: qg_type (src/qg/qg.F90:17-53), box_type
! (src/glam/box.F90:11-41), topog_type (src/glam/topog.F90:16-28), modes_type
! (src/qg/modes.F90:11-23), homog_type and its box/cyclic parts
! (src/qg/homog.F90:15-61), constraint_type and mass_constr_type
! (src/qg/constraint.F90:22-67).  The inhomogeneous solver's FFT plans are not
! read by core.F90 and are left out.
module qg

  implicit none

  private

  type box_type
     integer :: nxt = 0, nyt = 0, nl = 0
     double precision :: dx = 0.0d0, dy = 0.0d0
     double precision :: x0, y0
     double precision, allocatable :: h(:)
     double precision :: hm
     integer :: dz_sign
     logical :: cyclic
     double precision, allocatable :: yp(:), yprel(:), yt(:), ytrel(:), xp(:), xt(:)
     double precision :: fnot, beta
     integer :: nxp = 0, nyp = 0
     double precision :: xl, yl, dxm2, rdxf0
  end type box_type

  type topog_type
     double precision, allocatable :: dtop(:,:)
     double precision :: davg
     double precision, allocatable :: ddyn(:,:)
     integer :: k_topo
  end type topog_type

  type modes_type
     double precision, allocatable :: amat(:,:), c_phase(:), rdef(:), rdm2(:), ctl2m(:,:), ctm2l(:,:)
  end type modes_type

  type homog_box_type
     double precision, allocatable :: hom_sol_bc(:,:,:)
     double precision, allocatable :: cdhoc(:,:), LU(:,:)
     integer, allocatable :: ipiv(:)
  end type homog_box_type

  type homog_cyclic_type
     double precision, allocatable :: hom_sol_bc1(:,:), hom_sol_bc2(:,:)
     double precision, allocatable :: hom_sol_bt(:)
     double precision, allocatable :: hc1s(:), hc2s(:), hc1n(:), hc2n(:)
     double precision :: hbsi
  end type homog_cyclic_type

  type homog_type
     type(homog_cyclic_type) :: cyc
     type(homog_box_type) :: box
  end type homog_type

  type constraint_type
     double precision, allocatable :: cs(:), cn(:), csp(:), cnp(:)
     double precision, allocatable :: ajis(:), ajin(:)
  end type constraint_type

  type mass_constr_type
     double precision, allocatable :: dpi(:), dpip(:)
  end type mass_constr_type

  type qg_type
     logical :: active = .false.
     double precision :: bcco
     double precision :: delek
     double precision, allocatable :: ah2(:), ah4(:)
     double precision, allocatable :: gp(:)
     double precision :: rho
     integer :: tau_sign
     type(box_type) :: b
     type(topog_type) :: topo
     type(modes_type) :: mod
     type(homog_type) :: hom
     double precision, allocatable :: q(:,:,:), qm(:,:,:)
     double precision, allocatable :: p(:,:,:), pm(:,:,:)
     type(constraint_type) :: constr
     type(mass_constr_type) :: con
  end type qg_type

  public qg_type

end module qg
