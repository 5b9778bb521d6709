! iris_c_api -- Fortran view of the engine's C ABI (include/iris.h).
!
! This is the iso_c_binding layer the Fortran host uses in place of the
! reference's `call driver(e,h,eps_inv,...)` (reference code/main.f90:179,
! interface code/driver.f90:25-58).  The derived type mirrors `iris_params`
! field for field; C arrays [a][b][c] appear with reversed Fortran extents.
module iris_c_api
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: IRIS_ABI_VERSION = 1
  ! iris_get_plane selectors: base + component (0: x, 1: y, 2: z)
  integer(c_int32_t), parameter :: IRIS_PLANE_E = 0, IRIS_PLANE_H_PREV = 3, IRIS_PLANE_H = 6

  type, bind(C) :: iris_params_t
    integer(c_int32_t) :: abi_version
    integer(c_int32_t) :: precision
    integer(c_int32_t) :: nx, ny, nz
    integer(c_int32_t) :: pmlzi, pmlzf
    integer(c_int32_t) :: hi, hf
    integer(c_int32_t) :: i_pol
    integer(c_int32_t) :: em_source
    integer(c_int32_t) :: plane_wave
    integer(c_int32_t) :: cosine
    integer(c_int32_t) :: iz_min, iz_max
    integer(c_int32_t) :: ixo
    integer(c_int32_t) :: nt
    integer(c_int32_t) :: to_fourier
    integer(c_int32_t) :: sample_fourier
    integer(c_int32_t) :: no_downloads
    integer(c_int32_t) :: nof
    integer(c_int32_t) :: nbox
    integer(c_int32_t) :: box(6, 2)          ! C box[2][6]
    real(c_double) :: q, dt, c0, eps0, mu0
    real(c_double) :: pi, f_abohr
    real(c_double) :: eps(3)
    real(c_double) :: eps_in(3)
    real(c_double) :: epsa_inv(3, 3, 2)      ! C epsa_inv[region][row][col] = (col,row,region)
    real(c_double) :: eamp, wlpump, gpump, zopump
    real(c_double) :: ratio_angle, ux, uz
    real(c_double) :: win
    real(c_double) :: ak(2, 3)               ! C ak[axis][re/im]
  end type iris_params_t

  interface
    integer(c_int) function iris_create(p, eps_inv, fpzmn, fmzmn, fpzmx, fmzmx, fw, h) bind(C, name="iris_create")
      import :: c_int, c_ptr, iris_params_t
      type(iris_params_t), intent(in) :: p
      type(c_ptr), value :: eps_inv, fpzmn, fmzmn, fpzmx, fmzmx, fw
      type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function iris_step(h, it_first, it_last) bind(C, name="iris_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: it_first, it_last
    end function

    integer(c_int) function iris_sync(h) bind(C, name="iris_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function iris_flux(h, nobox, jabs) bind(C, name="iris_flux")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nobox
      type(c_ptr), value :: jabs
    end function

    ! one field component on one plane (field snapshots, reference code/sdl.f90)
    integer(c_int) function iris_get_plane(h, what, axis, index, range, out) bind(C, name="iris_get_plane")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: what, axis, index
      integer(c_int32_t), intent(in) :: range(4)
      type(c_ptr), value :: out
    end function

    ! energy and divergence monitors (reference code/energy.f90, uncalled there)
    integer(c_int) function iris_diagnostics(h, zi, zf, out) bind(C, name="iris_diagnostics")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: zi, zf
      type(c_ptr), value :: out
    end function

    ! gradient of a flux objective (no counterpart in the reference)
    integer(c_int) function iris_checkpoint(h, segment, record_bytes) bind(C, name="iris_checkpoint")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: segment
      integer(c_int64_t), value :: record_bytes
    end function

    integer(c_int) function iris_autotune(h, steps, info) bind(C, name="iris_autotune")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: steps
      type(c_ptr), value :: info
    end function

    integer(c_int) function iris_graph_info(h, info) bind(C, name="iris_graph_info")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      type(c_ptr), value :: info
    end function

    integer(c_int) function iris_checkpoint_info(h, info) bind(C, name="iris_checkpoint_info")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      type(c_ptr), value :: info
    end function

    integer(c_int) function iris_adjoint(h, nobox, weights, grad) bind(C, name="iris_adjoint")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nobox
      type(c_ptr), value :: weights, grad
    end function

    integer(c_int) function iris_destroy(h) bind(C, name="iris_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    type(c_ptr) function iris_last_error(h) bind(C, name="iris_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

end module iris_c_api
