! irisfdtd -- drop-in program for IrisFDTD-Academic running the time loop on
! the MI355X engine.
!
! Same program flow and output files as the reference
! (reference code/program.f90:25-61, code/main.f90:43-213,
! code/driver.f90:338-500):
!   inputFDTD.dat in the working directory, optional <<LOOP>> first line;
!   log.dat; dielectric_layer_spectra.dat;
!   <box>_scatt<j>.dat (j=1..6) and <box>_scatt.dat appended with '#',it
!   blocks of nof values at the output cadence.
! The loop body is replaced by iris_step on the device; at each output step
! the host fetches Jabs with iris_flux and writes the files with the same
! list-directed formatting as scattering_fourier_out
! (reference code/dataprocess.f90:279-299).
!
! At the same output steps the field snapshots of the enabled SDL planes
! (reference code/driver.f90:468-495, code/sdl.f90) are written by the
! iris_sdl writer from planes fetched with iris_get_plane.  With
! IRIS_MONITOR=1 in the environment the program also appends, per output
! step, the energy and divergence monitors of iris_diagnostics over the
! region between the UPML slabs to monitor.dat (it, W_E, W_H, W_E+W_H,
! max |div D|, max |div B|); the reference has these routines
! (code/energy.f90) but never calls them.
program irisfdtd
  use iso_c_binding
  use iris_c_api
  use iris_setup
  use iris_sdl
  implicit none
  character(len=100) :: first
  integer :: u, ios, io_main, if_main, i, status, prec_default
  real(8) :: dxl, dyl, dzl
  character(len=16) :: envp
  integer, parameter :: SNAPSHOT_IO_ERROR = -20
  logical :: monitor

  prec_default = 8
  call get_environment_variable('IRIS_PRECISION', envp, status=status)
  if (status == 0) then
    read(envp, *, iostat=ios) prec_default
    if (ios /= 0 .or. (prec_default /= 4 .and. prec_default /= 8)) prec_default = 8
  end if

  monitor = .false.
  call get_environment_variable('IRIS_MONITOR', envp, status=status)
  if (status == 0) monitor = envp(1:1) == '1'
  if (monitor) then
    ! one file per program invocation; the runs of a <<LOOP>> append to it
    open(newunit=u, file='monitor.dat', status='replace', action='write')
    close(u)
  end if

  open(newunit=u, file='inputFDTD.dat', status='old', action='read', iostat=ios)
  if (ios /= 0) then
    write(*, '(A)') 'irisfdtd: cannot open inputFDTD.dat'
    error stop 1
  end if
  read(u, '(A100)', iostat=ios) first
  if (first /= '<<LOOP>>') then
    close(u)
    call run_once(.true., status)
  else
    ! <<LOOP>>: repeat the whole run for i = io_main..if_main
    ! (reference code/program.f90:40-56); the delta values are read and unused
    read(u, *) io_main
    read(u, *) if_main
    read(u, *) dxl
    read(u, *) dyl
    read(u, *) dzl
    close(u)
    status = 0
    do i = io_main, if_main
      call run_once(.false., status)
      if (status /= 0) exit
    end do
  end if
  if (status /= 0) error stop 2

contains

  ! Convert the library's NUL-terminated error message to a Fortran string.
  function iris_error_text(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=:), allocatable :: msg
    type(c_ptr) :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    s = iris_last_error(h)
    n = int(c_strlen(s))
    call c_f_pointer(s, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function iris_error_text

  subroutine run_once(with_description, status)
    logical, intent(in) :: with_description
    integer, intent(out) :: status
    type(iris_run), target :: run
    type(c_ptr) :: h
    character(len=400) :: errmsg
    integer :: ierr, lg, it, stop_at, cand, nos, nb, ot, rc, um
    real(8), target :: diag(4)
    real(8), allocatable, target :: jabs(:,:)
    integer(8) :: c0, c1, crate
    real(8) :: secs

    status = 0
    open(newunit=lg, file='log.dat', status='replace', action='write')
    write(lg, *) 'irisfdtd (MI355X engine for the IrisFDTD-Academic time loop)'
    call iris_load_input('inputFDTD.dat', run, ierr, errmsg, prec_default)
    if (.not. with_description) run%description = ''
    write(lg, *) 'Description     : ', trim(run%description)
    if (ierr /= 0) then
      write(lg, *) '...Error -> ', trim(errmsg)
      write(*, '(A)') 'irisfdtd: '//trim(errmsg)
      close(lg)
      status = ierr
      return
    end if
    call iris_log_parameters(run, lg)
    call write_spectra(run)

    rc = iris_create(run%p, c_loc(run%eps_inv), c_loc(run%fpzmn), c_loc(run%fmzmn), &
                     c_loc(run%fpzmx), c_loc(run%fmzmx), c_loc(run%fw), h)
    if (rc /= 0) then
      write(lg, *) '...Error -> iris_create: ', iris_error_text(c_null_ptr)
      write(*, '(A)') 'irisfdtd: iris_create failed: '//iris_error_text(c_null_ptr)
      close(lg)
      status = rc
      return
    end if
    write(lg, *) '...Starting main calculation'
    allocate(jabs(6, run%p%nof))

    call system_clock(c0, crate)
    ot = max(run%output_time_step, 1)
    nos = 1
    it = 1
    do while (it <= run%p%nt)
      cand = run%p%to_fourier + run%p%no_downloads*nos
      if (cand >= it) then
        stop_at = min(cand, run%p%nt)
      else
        stop_at = run%p%nt
      end if
      ! progress lines at multiples of OUTPUT_TIME_STEP (reference code/driver.f90:340-353)
      stop_at = min(stop_at, (it/ot + 1)*ot)
      rc = iris_step(h, it, stop_at)
      if (rc /= 0) exit
      if (mod(stop_at, ot) == 0) then
        rc = iris_sync(h)
        if (rc /= 0) exit
        call system_clock(c1)
        secs = real(c1 - c0, 8)/real(crate, 8)
        print *, '  Iteration ', stop_at, ' of ', run%p%nt, '  (', stop_at/max(secs, 1d-9), ' it/s)'
      end if
      if (stop_at >= run%p%to_fourier .and. (stop_at == cand .or. stop_at == run%p%nt)) then
        do nb = 1, run%p%nbox
          rc = iris_flux(h, nb, c_loc(jabs))
          if (rc /= 0) exit
          call write_flux(nb, stop_at, jabs)
        end do
        if (rc /= 0) exit
        if (sdl_enabled(run%sdl)) then
          call write_snapshots(h, run, nos + 1, rc, errmsg)
          if (rc /= 0) exit
        end if
        if (monitor) then
          rc = iris_diagnostics(h, run%p%pmlzi + 1, run%p%pmlzf - 1, c_loc(diag))
          if (rc /= 0) exit
          open(newunit=um, file='monitor.dat', status='unknown', action='write', position='append')
          write(um, '(I10,5(1X,ES24.16E3))') stop_at, diag(1), diag(2), diag(1) + diag(2), diag(3), diag(4)
          close(um)
        end if
        nos = nos + 1
      end if
      it = stop_at + 1
    end do
    if (rc /= 0) then
      if (rc == SNAPSHOT_IO_ERROR) then
        write(lg, *) '...Error -> ', trim(errmsg)
        write(*, '(A)') 'irisfdtd: '//trim(errmsg)
      else
        write(lg, *) '...Error -> ', iris_error_text(h)
        write(*, '(A)') 'irisfdtd: '//iris_error_text(h)
      end if
      status = rc
    end if
    call system_clock(c1)
    secs = real(c1 - c0, 8)/real(crate, 8)
    write(lg, *) '...Main calculation end; loop seconds = ', secs
    write(lg, *) '...Gcell-updates/s = ', &
      real(run%p%nx, 8)*real(run%p%ny, 8)*real(run%p%nz, 8)*real(run%p%nt, 8)/max(secs, 1d-9)/1d9
    rc = iris_destroy(h)
    close(lg)
  end subroutine run_once

  ! Snapshot files of every enabled plane and layer for output number ino
  ! (reference code/driver.f90:468-495): E after the step and H of the step
  ! before, fetched plane by plane.
  subroutine write_snapshots(h, run, ino, rc, errmsg)
    type(c_ptr), intent(in) :: h
    type(iris_run), intent(in) :: run
    integer, intent(in) :: ino
    integer, intent(out) :: rc
    character(len=*), intent(out) :: errmsg
    complex(8), allocatable, target :: e(:,:,:), hp(:,:,:)
    integer :: layer, il, c, fx, a0, a1, b0, b1, ierr
    integer(c_int32_t) :: rng(4)
    logical :: need_e, need_h
    rc = 0
    errmsg = ''
    need_e = run%sdl%module_e == 1 .or. any(run%sdl%field(1:3) >= 1)
    need_h = run%sdl%module_e == 2 .or. any(run%sdl%field(4:6) >= 1)
    do layer = 1, 3
      if (run%sdl%plane_on(layer) /= 1) cycle
      do il = 1, run%sdl%nlayers(layer)
        call sdl_geometry(run%sdl, layer, il, fx, a0, a1, b0, b1)
        allocate(e(a0:a1, b0:b1, 3), hp(a0:a1, b0:b1, 3))
        e = (0.0d0, 0.0d0)
        hp = (0.0d0, 0.0d0)
        rng = [a0, a1, b0, b1]
        do c = 1, 3
          if (need_e) then
            rc = iris_get_plane(h, IRIS_PLANE_E + c - 1, layer, fx, rng, c_loc(e(a0, b0, c)))
            if (rc /= 0) return
          end if
          if (need_h) then
            rc = iris_get_plane(h, IRIS_PLANE_H_PREV + c - 1, layer, fx, rng, c_loc(hp(a0, b0, c)))
            if (rc /= 0) return
          end if
        end do
        call sdl_write_plane(run%sdl, layer, fx, a0, b0, ino, e, hp, run%p%q*run%p%eamp, ierr, errmsg)
        deallocate(e, hp)
        if (ierr /= 0) then
          rc = SNAPSHOT_IO_ERROR
          return
        end if
      end do
    end do
  end subroutine write_snapshots

  ! <box>_scatt<j>.dat and <box>_scatt.dat (reference code/dataprocess.f90:279-299)
  subroutine write_flux(nobox, it, jabs)
    integer, intent(in) :: nobox, it
    real(8), intent(in) :: jabs(:,:)
    integer :: j, iw, uo
    character(len=32) :: fname
    do j = 1, 6
      write(fname, '(I0,A,I0,A)') nobox, '_scatt', j, '.dat'
      open(newunit=uo, file=trim(fname), status='unknown', action='write', position='append')
      write(uo, *) '#', it
      do iw = 1, size(jabs, 2)
        write(uo, *) jabs(j, iw)
      end do
      close(uo)
    end do
    write(fname, '(I0,A)') nobox, '_scatt.dat'
    open(newunit=uo, file=trim(fname), status='unknown', action='write', position='append')
    write(uo, *) '#', it
    do iw = 1, size(jabs, 2)
      write(uo, *) jabs(1, iw) + jabs(2, iw) + jabs(3, iw) + jabs(4, iw) + jabs(5, iw) + jabs(6, iw)
    end do
    close(uo)
  end subroutine write_flux

  ! dielectric_layer_spectra.dat (reference code/common.f90:114-130)
  subroutine write_spectra(run)
    type(iris_run), intent(in) :: run
    integer :: us, iw
    open(newunit=us, file='dielectric_layer_spectra.dat', status='unknown', action='write', position='rewind')
    do iw = 1, size(run%spectra, 2)
      write(us, *) run%spectra(1, iw), run%spectra(2, iw), run%spectra(3, iw)
    end do
    close(us)
  end subroutine write_spectra

end program irisfdtd
