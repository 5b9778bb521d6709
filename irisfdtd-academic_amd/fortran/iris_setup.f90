! iris_setup -- host-side set-up for the MI355X engine.
!
! Reads an IrisFDTD-Academic `inputFDTD.dat` and derives everything the time
! loop needs, with the reference's arithmetic so that the FP64 values match
! bit for bit:
!   parameters and units     reference code/files.f90:39-344 (setparam)
!   slit limits              reference code/files.f90:385-409 (struct_limit)
!   eps_inv, epsa_inv        reference code/mesh.f90:75-109 (set_emconstants)
!   slit                     reference code/mesh.f90:57-69
!   UPML coefficient tables  reference code/boundary_II.f90:224-276 (set_sz)
!   TF/SF ranges, dtcq2      reference code/driver.f90:310-335
!   analytic film spectrum   reference code/common.f90:104-131
!   snapshot options         reference code/sdl.f90:868-897 (setparam_SDL),
!                            code/files.f90:291-319 (extra layers)
!
! Differences from the reference, by design: errors are returned as a code
! plus message instead of STOP (reference code/files.f90:352-381), and an
! optional "Engine options" section selects the field precision (8 or 4);
! the reference skips sections it does not know.
!
! The module also exports a small bind(C) interface (iris_host_*) so that
! non-Fortran callers (the Python binding, the tests) derive parameters from
! the same code the Fortran program uses.
module iris_setup
  use iso_c_binding
  use iris_c_api, only: iris_params_t, IRIS_ABI_VERSION
  use iris_sdl, only: sdl_config, sdl_geometry, sdl_write_plane
  implicit none
  private

  integer, parameter :: dp = 8

  ! constants of the reference (code/physconst.f90:31-41): most literals are
  ! default REAL, so they carry float32 rounding into FP64
  real(dp), parameter, public :: pi_r = 3.14159265358979323846
  real(dp), parameter, public :: c0_r = 137.0360
  real(dp), parameter, public :: eps0_r = 1.0/(4.0*pi_r)
  real(dp), parameter, public :: mu0_r = 1.0/(eps0_r*(c0_r**2))
  real(dp), parameter, public :: f_abohr_r = 10/0.529177249
  real(dp), parameter, public :: uat_r = 24.188843265E-3
  real(dp), parameter :: f_e_lit = 514.2208281e9
  real(dp), parameter :: pmlalfa = 1.8

  integer, parameter, public :: IRIS_HOST_OK = 0
  integer, parameter, public :: IRIS_HOST_ENOFILE = -10
  integer, parameter, public :: IRIS_HOST_EBADINPUT = -11

  ! everything a run needs, one record per loaded input
  type, public :: iris_run
    type(iris_params_t) :: p
    real(dp) :: q_nm, cfl, wzpml_in, texit_fs, dlamda, lno, lnf
    real(dp) :: pump_dt_fs, def_pulse, tincr_pw, tdecay_pw, epump_kvcm
    real(dp) :: wzpml
    complex(dp) :: eps_bbox
    integer :: pml_taflove, output_time_step, no_downloads_in
    integer :: dh, dx, aix, afx, tfsf_zi
    integer :: sdl_plane_on(3)
    type(sdl_config) :: sdl                     ! snapshot options (read when a plane is on)
    integer :: ngrid_eps                        ! elements of eps_inv
    character(len=100) :: description
    complex(dp), allocatable :: eps_inv(:,:,:)  ! (0:nx+1, 0:ny+1, hi-1:hf+1)
    real(dp), allocatable :: fpzmn(:,:), fmzmn(:,:), fpzmx(:,:), fmzmx(:,:)
    real(dp), allocatable :: fw(:)
    real(dp), allocatable :: spectra(:,:)       ! (3, nof): lambda, R, T
  end type iris_run

  public :: iris_load_input, iris_log_parameters

  ! state behind the bind(C) interface
  type(iris_run), save, target :: c_run
  character(len=256), save :: c_errmsg = ''

contains

  ! Scan from the top of the file for a line whose first list-directed token
  ! equals `header` (the reference's rewind/read loops, code/files.f90:48-51).
  subroutine find_section(u, header, found)
    integer, intent(in) :: u
    character(len=*), intent(in) :: header
    logical, intent(out) :: found
    character(len=100) :: tok
    integer :: ios
    found = .false.
    rewind(u)
    do
      tok = ''
      read(u, *, iostat=ios) tok
      if (ios < 0) return
      if (ios > 0) cycle
      if (tok == header) then
        found = .true.
        return
      end if
    end do
  end subroutine find_section

  ! struct_limit (reference code/files.f90:385-409): centred slit of width a
  ! on a period of N cells
  subroutine slit_limits(n, a, ai, af)
    integer, intent(in) :: n, a
    integer, intent(out) :: ai, af
    if (mod(a, 2) == 0) then
      if (mod(n, 2) == 0) then
        ai = n/2 - (a/2 - 1); af = n/2 + a/2
      else
        ai = (n + 1)/2 - a/2; af = (n + 1)/2 + a/2 - 1
      end if
    else
      if (mod(n, 2) == 0) then
        ai = n/2 - ((a + 1)/2 - 1); af = n/2 + (a - 1)/2
      else
        ai = (n + 1)/2 - (a - 1)/2; af = (n + 1)/2 + (a - 1)/2
      end if
    end if
  end subroutine slit_limits

  ! Parse `path` and derive the run.  ierr = 0 on success; otherwise errmsg
  ! says what was wrong.  precision_default applies when the input has no
  ! "Engine options" section.
  subroutine iris_load_input(path, run, ierr, errmsg, precision_default)
    character(len=*), intent(in) :: path
    type(iris_run), intent(out) :: run
    integer, intent(out) :: ierr
    character(len=*), intent(out) :: errmsg
    integer, intent(in), optional :: precision_default
    integer :: u, ios, i, ix, iy, texit, nof, prec
    integer :: sxi, sxf, syi, syf, szi, szf, mxi, mxf, myi, myf, mzi, mzf
    integer :: to_fourier, sample_fourier, plane_wave, cosine, i_pol
    integer :: nx, ny, nz, pmlzi, pmlzf, hi, hf
    real(dp) :: q, dt, cfl, wz, lno_aux, epump, wlpump, pump_dt, def_pulse, gpump, zopump
    real(dp) :: eps(3), eps_in(3), delta, eps_bbox_re
    real(dp) :: sg, kk
    character(len=100) :: tok, line1
    character(len=400) :: ebuf
    logical :: found
    complex(dp) :: eps_bbox

    ierr = IRIS_HOST_OK
    errmsg = ''
    prec = 8
    if (present(precision_default)) prec = precision_default

    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      ierr = IRIS_HOST_ENOFILE
      errmsg = 'cannot open input file: '//trim(path)
      return
    end if

    ! description: the first line (reference code/main.f90:111)
    read(u, '(A100)', iostat=ios) line1
    run%description = line1

    ! ---- "Mesh options" (code/files.f90:48-111)
    call find_section(u, 'Mesh options', found)
    if (.not. found) then; call fail('section "Mesh options" not found'); return; end if
    read(u, *, iostat=ios) i_pol;            if (bad(ios, 'i_pol')) return
    if (i_pol < 1 .or. i_pol > 4) then; call fail('i_pol must be 1..4'); return; end if
    read(u, *, iostat=ios) run%q_nm;         if (bad(ios, 'q')) return
    if (run%q_nm <= 0.0) then; call fail('q must be > 0'); return; end if
    read(u, *, iostat=ios) nx;               if (bad(ios, 'ixmax')) return
    read(u, *, iostat=ios) ny;               if (bad(ios, 'iymax')) return
    read(u, *, iostat=ios) nz;               if (bad(ios, 'izmax')) return
    if (nx < 1 .or. ny < 1 .or. nz < 1) then; call fail('ixmax, iymax, izmax must be >= 1'); return; end if
    q = f_abohr_r*run%q_nm
    read(u, *, iostat=ios) cfl;              if (bad(ios, 'CFL')) return
    if (cfl <= 0.0) then; call fail('CFL must be > 0'); return; end if
    run%cfl = cfl
    dt = cfl/(c0_r*dsqrt(3.0/q**2))
    read(u, *, iostat=ios) wz;               if (bad(ios, 'wzPML')) return
    if (wz < 0.0) then; call fail('wzPML must be >= 0'); return; end if
    run%wzpml_in = wz
    run%wzpml = f_abohr_r*(wz/q)
    read(u, *, iostat=ios) run%pml_taflove;  if (bad(ios, 'nPML')) return
    read(u, *, iostat=ios) pmlzi;            if (bad(ios, 'pmlzi')) return
    read(u, *, iostat=ios) pmlzf;            if (bad(ios, 'pmlzf')) return

    ! ---- "Job options" (code/files.f90:115-154)
    call find_section(u, 'Job options', found)
    if (.not. found) then; call fail('section "Job options" not found'); return; end if
    read(u, *, iostat=ios) run%no_downloads_in, run%output_time_step; if (bad(ios, 'no_downloads')) return
    read(u, *, iostat=ios) run%texit_fs;     if (bad(ios, 'itmax')) return
    texit = int(run%texit_fs/(dt*uat_r))
    if (run%no_downloads_in < 1) then; call fail('no_downloads must be >= 1'); return; end if
    read(u, *, iostat=ios) hi;               if (bad(ios, 'hi')) return
    read(u, *, iostat=ios) run%sdl_plane_on; if (bad(ios, 'SDL planes')) return

    ! ---- "Dielectric constants" (code/files.f90:155-170); eps_bbox may be
    ! written as a real or as a complex literal
    call find_section(u, 'Dielectric constants', found)
    if (.not. found) then; call fail('section "Dielectric constants" not found'); return; end if
    read(u, *, iostat=ios) eps(1);           if (bad(ios, 'eps(1)')) return
    read(u, '(A)', iostat=ios) ebuf;         if (bad(ios, 'eps_bbox')) return
    read(ebuf, *, iostat=ios) eps_bbox
    if (ios /= 0) then
      read(ebuf, *, iostat=ios) eps_bbox_re; if (bad(ios, 'eps_bbox')) return
      eps_bbox = eps_bbox_re
    end if
    read(u, *, iostat=ios) eps(3);           if (bad(ios, 'eps(3)')) return
    eps(2) = real(eps_bbox, dp)
    eps_in(1) = 1.0/eps(1)
    eps_in(2) = real(1.0/eps_bbox, dp)
    eps_in(3) = 1.0/eps(3)
    run%eps_bbox = eps_bbox

    ! ---- "Fourier transformation options" (code/files.f90:172-254)
    call find_section(u, 'Fourier transformation options', found)
    if (.not. found) then; call fail('section "Fourier transformation options" not found'); return; end if
    read(u, *, iostat=ios) sxi, mxi;         if (bad(ios, 'sxi')) return
    read(u, *, iostat=ios) sxf, mxf;         if (bad(ios, 'sxf')) return
    read(u, *, iostat=ios) syi, myi;         if (bad(ios, 'syi')) return
    read(u, *, iostat=ios) syf, myf;         if (bad(ios, 'syf')) return
    read(u, *, iostat=ios) szi, mzi;         if (bad(ios, 'szi')) return
    read(u, *, iostat=ios) szf, mzf;         if (bad(ios, 'szf')) return
    read(u, *, iostat=ios) nof;              if (bad(ios, 'nof')) return
    if (nof < 1) then; call fail('nof must be >= 1'); return; end if
    read(u, *, iostat=ios) run%dlamda;       if (bad(ios, 'dlambda')) return
    read(u, *, iostat=ios) run%lno;          if (bad(ios, 'lno')) return
    allocate(run%fw(nof))
    lno_aux = run%lno
    do i = 1, nof
      run%fw(i) = 2.0*pi_r*c0_r*dt/(f_abohr_r*lno_aux)
      lno_aux = lno_aux + run%dlamda
    end do
    run%lnf = lno_aux
    read(u, *, iostat=ios) to_fourier, sample_fourier; if (bad(ios, 'to_Fourier')) return
    if (sample_fourier < 1) then; call fail('sample_Fourier must be >= 1'); return; end if

    ! ---- "Source options" (code/files.f90:256-290)
    call find_section(u, 'Source options', found)
    if (.not. found) then; call fail('section "Source options" not found (the TF/SF source is required)'); return; end if
    read(u, *, iostat=ios) plane_wave;       if (bad(ios, 'plane_wave')) return
    read(u, *, iostat=ios) run%tincr_pw, run%tdecay_pw; if (bad(ios, 'tincr_PW')) return
    read(u, *, iostat=ios) epump, wlpump, pump_dt, def_pulse; if (bad(ios, 'Epump')) return
    read(u, *, iostat=ios) run%tfsf_zi;      if (bad(ios, 'tfsf_o')) return
    read(u, *, iostat=ios) cosine;           if (bad(ios, 'cosine')) return
    if (plane_wave /= 1) then
      call fail('plane_wave must be 1 (Gaussian pulse); the CW source of the reference is broken (GAP unset)')
      return
    end if
    run%epump_kvcm = epump
    run%pump_dt_fs = pump_dt
    run%def_pulse = def_pulse
    epump = epump/(f_e_lit*1e-5)
    pump_dt = pump_dt/uat_r
    gpump = c0_r*pump_dt/(2.d0*q*sqrt(eps(1)*def_pulse))
    zopump = -sqrt(def_pulse)*gpump - run%tfsf_zi

    ! ---- optional engine section (not in the reference; it skips it)
    call find_section(u, 'Engine options', found)
    if (found) then
      read(u, *, iostat=ios) prec;           if (bad(ios, 'precision')) return
      if (prec /= 4 .and. prec /= 8) then; call fail('precision must be 4 or 8'); return; end if
    end if

    ! ---- "Slit Array" (code/files.f90:321-342)
    call find_section(u, 'Slit Array', found)
    if (.not. found) then; call fail('section "Slit Array" not found'); return; end if
    read(u, *, iostat=ios) run%dh;           if (bad(ios, 'Dh')) return
    hf = (hi + run%dh) - 1
    read(u, *, iostat=ios) run%dx;           if (bad(ios, 'dx')) return
    call slit_limits(nx, run%dx, run%aix, run%afx)

    ! ---- snapshot options, read when any plane is on (code/files.f90:151-154,
    ! code/driver.f90:289-292 -> setparam_SDL code/sdl.f90:868-897)
    run%sdl%plane_on = run%sdl_plane_on
    if (sum(run%sdl_plane_on) /= 0) then
      call find_section(u, 'Time evolution options', found)
      if (.not. found) then
        call fail('SDL planes are on but section "Time evolution options" is missing'); return
      end if
      read(u, *, iostat=ios) run%sdl%copy_flag, run%sdl%copy_no; if (bad(ios, 'SDL_copy_flag')) return
      read(u, *, iostat=ios) run%sdl%xi;        if (bad(ios, 'SDL sxi')) return
      read(u, *, iostat=ios) run%sdl%xf;        if (bad(ios, 'SDL sxf')) return
      read(u, *, iostat=ios) run%sdl%yi;        if (bad(ios, 'SDL syi')) return
      read(u, *, iostat=ios) run%sdl%yf;        if (bad(ios, 'SDL syf')) return
      read(u, *, iostat=ios) run%sdl%zi;        if (bad(ios, 'SDL szi')) return
      read(u, *, iostat=ios) run%sdl%zf;        if (bad(ios, 'SDL szf')) return
      read(u, *, iostat=ios) run%sdl%module_e;  if (bad(ios, 'SDL E')) return
      do i = 1, 6
        read(u, *, iostat=ios) run%sdl%field(i); if (bad(ios, 'SDL field flags')) return
      end do
      if (run%sdl%copy_flag /= 0 .and. run%sdl%copy_no < 1) then
        call fail('SDL_copy_no must be >= 1'); return
      end if
      ! optional extra layers (code/files.f90:291-319)
      run%sdl%nlayers = 1
      call find_section(u, 'Extra parameters-SDLextended', found)
      if (found) then
        read(u, *, iostat=ios) run%sdl%nlayers; if (bad(ios, 'SDL extra layer counts')) return
        if (any(run%sdl%nlayers < 1)) then; call fail('SDL extra layer counts must be >= 1'); return; end if
      end if
      allocate(run%sdl%delta(maxval(run%sdl%nlayers), 3))
      run%sdl%delta = 0
      if (found) then
        do iy = 1, 3
          do i = 1, run%sdl%nlayers(iy)
            read(u, *, iostat=ios) run%sdl%delta(i, iy); if (bad(ios, 'SDL layer offset')) return
          end do
        end do
      end if
      ! every enabled plane must lie inside the grid (the reference would read
      ! ghost or out-of-range cells)
      do iy = 1, 3
        if (run%sdl%plane_on(iy) /= 1) cycle
        do i = 1, run%sdl%nlayers(iy)
          if (.not. plane_ok(iy, i)) then
            write(ebuf, '(A,I0,A,I0,A)') 'SDL plane ', iy, ' layer ', i, ' lies outside the grid'
            call fail(trim(ebuf)); return
          end if
        end do
      end do
    end if
    close(u)

    ! ---- consistency checks the engine relies on (the reference would run
    ! out of bounds instead)
    if (pmlzi < 2 .or. pmlzf <= pmlzi .or. pmlzf >= nz) then
      call fail('need 2 <= pmlzi < pmlzf < izmax'); return
    end if
    if (hi < 1 .or. hf < hi .or. hf > nz) then
      call fail('structure [hi, hi+Dh-1] must lie in [1, izmax]'); return
    end if
    if (run%tfsf_zi < 2 .or. run%tfsf_zi > nz) then
      call fail('tfsf_o must lie in [2, izmax]'); return
    end if
    if (.not. box_ok(sxi, sxf, syi, syf, szi, szf) .or. .not. box_ok(mxi, mxf, myi, myf, mzi, mzf)) then
      call fail('Fourier box planes must satisfy 1 <= i <= f <= N on each axis'); return
    end if
    if (i_pol /= 1 .and. i_pol /= 2) then
      call fail('i_pol 3 and 4 have no source in the reference (code/tfsf.f90:299-350)'); return
    end if

    ! ---- eps_inv (code/mesh.f90:97-106) and the slit (code/mesh.f90:65)
    allocate(run%eps_inv(0:nx + 1, 0:ny + 1, hi - 1:hf + 1))
    run%eps_inv = 1.0/eps_bbox
    run%eps_inv(:, :, hi - 1) = 1.0/eps(1)
    run%eps_inv(:, :, hf + 1) = 1.0/eps(3)
    if (run%afx >= run%aix) run%eps_inv(run%aix:run%afx, :, hi:hf) = 1.0d0
    run%ngrid_eps = size(run%eps_inv)

    ! ---- UPML tables (code/boundary_II.f90:224-276); only the z column is graded
    allocate(run%fpzmn(3, pmlzi), run%fmzmn(3, pmlzi))
    allocate(run%fpzmx(3, pmlzf:nz), run%fmzmx(3, pmlzf:nz))
    do i = 1, pmlzi
      do ix = 1, 3
        sg = 0.0; kk = 1.0
        if (ix == 3) then
          sg = sqrt(eps(1))*run%wzpml*((pmlzi - i)**2/(dble(pmlzi - 1))**2)
          kk = 1 + sqrt(eps(1))*pmlalfa*((pmlzi - i)**2/(dble(pmlzi - 1))**2)
        end if
        run%fpzmn(ix, i) = 2*eps0_r*eps(1)*kk + dt*sg
        run%fmzmn(ix, i) = 2*eps0_r*eps(1)*kk - dt*sg
      end do
    end do
    do i = pmlzf, nz
      do ix = 1, 3
        sg = 0.0; kk = 1.0
        if (ix == 3) then
          sg = sqrt(eps(3))*run%wzpml*((i - pmlzf))**2/(dble(nz - pmlzf)**2)
          kk = 1 + sqrt(eps(3))*pmlalfa*((i - pmlzf))**2/(dble(nz - pmlzf)**2)
        end if
        run%fpzmx(ix, i) = 2*eps0_r*eps(3)*kk + dt*sg
        run%fmzmx(ix, i) = 2*eps0_r*eps(3)*kk - dt*sg
      end do
    end do

    ! ---- analytic lossless-film spectrum (code/common.f90:104-131)
    call film_spectrum(run, eps, q, nof)

    ! ---- the parameter block (include/iris.h)
    run%p%abi_version = IRIS_ABI_VERSION
    run%p%precision = prec
    run%p%nx = nx; run%p%ny = ny; run%p%nz = nz
    run%p%pmlzi = pmlzi; run%p%pmlzf = pmlzf
    run%p%hi = hi; run%p%hf = hf
    run%p%i_pol = i_pol
    run%p%em_source = 1
    run%p%plane_wave = plane_wave
    run%p%cosine = cosine
    run%p%iz_min = run%tfsf_zi
    run%p%iz_max = run%tfsf_zi + 5
    run%p%ixo = pmlzi - 50
    run%p%nt = texit
    run%p%to_fourier = to_fourier
    run%p%sample_fourier = sample_fourier
    run%p%no_downloads = int(texit/run%no_downloads_in)
    run%p%nof = nof
    run%p%box(:, 1) = [sxi, sxf, syi, syf, szi, szf]
    run%p%box(:, 2) = [mxi, mxf, myi, myf, mzi, mzf]
    if (all(run%p%box(:, 1) == run%p%box(:, 2))) then
      run%p%nbox = 1
    else
      run%p%nbox = 2
    end if
    run%p%q = q; run%p%dt = dt
    run%p%c0 = c0_r; run%p%eps0 = eps0_r; run%p%mu0 = mu0_r
    run%p%pi = pi_r; run%p%f_abohr = f_abohr_r
    run%p%eps = eps
    run%p%eps_in = eps_in
    run%p%epsa_inv = 0.0_dp
    do i = 1, 3
      run%p%epsa_inv(i, i, 1) = 1.0d0/eps(1)
      run%p%epsa_inv(i, i, 2) = 1.0d0/eps(3)
    end do
    run%p%eamp = epump
    run%p%wlpump = wlpump
    run%p%gpump = gpump
    run%p%zopump = zopump
    run%p%ratio_angle = 1.0
    delta = run%p%ratio_angle*pi_r/2.0
    run%p%ux = dcos(delta)
    run%p%uz = dsin(delta)
    run%p%win = 1.0
    run%p%ak = 0.0_dp
    iy = 0

  contains

    subroutine fail(msg)
      character(len=*), intent(in) :: msg
      ierr = IRIS_HOST_EBADINPUT
      errmsg = msg
      close(u)
    end subroutine fail

    logical function bad(ios_in, what)
      integer, intent(in) :: ios_in
      character(len=*), intent(in) :: what
      bad = (ios_in /= 0)
      if (bad) call fail('cannot read '//what)
    end function bad

    logical function plane_ok(layer, ilayer)
      integer, intent(in) :: layer, ilayer
      integer :: fx, a0, a1, b0, b1, n(3), na, nb
      n = [nx, ny, nz]
      call sdl_geometry(run%sdl, layer, ilayer, fx, a0, a1, b0, b1)
      select case (layer)
      case (1); na = ny; nb = nz
      case (2); na = nx; nb = nz
      case default; na = nx; nb = ny
      end select
      plane_ok = fx >= 1 .and. fx <= n(layer) .and. a0 >= 1 .and. a0 <= a1 .and. a1 <= na &
                 .and. b0 >= 1 .and. b0 <= b1 .and. b1 <= nb
    end function plane_ok

    logical function box_ok(a, b, c, d, e, f)
      integer, intent(in) :: a, b, c, d, e, f
      box_ok = a >= 1 .and. a <= b .and. b <= nx .and. c >= 1 .and. c <= d .and. d <= ny &
               .and. e >= 1 .and. e <= f .and. f <= nz
    end function box_ok
  end subroutine iris_load_input

  ! R and T of a lossless film of thickness Dh*q (reference code/common.f90:104-131)
  subroutine film_spectrum(run, eps, q, nof)
    type(iris_run), intent(inout) :: run
    real(dp), intent(in) :: eps(3), q
    integer, intent(in) :: nof
    integer :: iw
    real(dp) :: lp, del, r
    complex(dp) :: n2, n3, cn
    complex(dp), parameter :: ci = (0.0, 1.0)
    allocate(run%spectra(3, nof))
    do iw = 1, nof
      lp = run%lno + run%dlamda*(iw - 1.0)
      n2 = sqrt(run%eps_bbox)
      n3 = sqrt(eps(3))
      del = 2*pi_r*n2*run%dh*q/(lp*f_abohr_r)
      cn = (dcos(del) + ci*(n3/n2)*dsin(del) - n3*dcos(del) - ci*n2*dsin(del))
      cn = cn/(dcos(del) + ci*(n3/n2)*dsin(del) + n3*dcos(del) + ci*n2*dsin(del))
      r = cdabs(dconjg(cn)*cn)
      run%spectra(:, iw) = [lp, r, 1.0d0 - r]
    end do
  end subroutine film_spectrum

  ! log.dat parameter lines (same quantities as reference code/files.f90:60-342)
  subroutine iris_log_parameters(run, u)
    type(iris_run), intent(in) :: run
    integer, intent(in) :: u
    select case (run%p%i_pol)
    case (1); write(u, '(A1,A20)') '#', 'Polarization = E//y'
    case (2); write(u, '(A1,A20)') '#', 'Polarization = E//x'
    end select
    write(u, '(A1,A12,F15.5,A12,F15.5)') '#', 'q(nm)=', run%q_nm, 'q(u.a)=', run%p%q
    write(u, '(A1,A12,I15,A12,F15.5)') '#', 'ixmax=', run%p%nx, 'Lx(nm)=', run%p%nx*run%q_nm
    write(u, '(A1,A12,I15,A12,F15.5)') '#', 'iymax=', run%p%ny, 'Ly(nm)=', run%p%ny*run%q_nm
    write(u, '(A1,A12,I15,A12,F15.5)') '#', 'izmax=', run%p%nz, 'Lz(nm)=', run%p%nz*run%q_nm
    write(u, '(A1,A12,F15.5)') '#', 'CFL =', run%cfl
    write(u, '(A1,A15,F15.5)') '#', 'PML strength =', run%wzpml
    write(u, '(A1,A15,I10)') '#', 'PML exponent =', 2
    write(u, '(A1,A12,I10)') '#', 'pmlzi=', run%p%pmlzi
    write(u, '(A1,A12,I10)') '#', 'pmlzf=', run%p%pmlzf
    write(u, '(A1,A12,F15.8)') '#', 'dt(fs)=', run%p%dt*uat_r
    write(u, '(A1,A26,I10,A26,F15.5,A7)') '#', 'Iteration max. number =', run%p%nt, ' Simulation time =', &
      uat_r*run%p%dt*run%p%nt, ' (fs)'
    write(u, '(A1,A12,I10)') '#', 'hi=', run%p%hi
    write(u, '(A4,A20,F10.5)') '#  -', 'Eps region I  =', run%p%eps(1)
    write(u, '(A4,A20,F10.5,F10.5)') '   -', 'Eps region II =', run%eps_bbox
    write(u, '(A4,A20,F10.5)') '   -', 'Eps region III=', run%p%eps(3)
    write(u, '(A1,A12,(I2))') '#', 'to_Fourier=', run%p%to_fourier
    write(u, '(A1,A22,I1)') '#', 'Plane-Wave=', run%p%plane_wave
    write(u, *) run%p%gpump, run%p%zopump
    write(u, '(A1,A12,I10)') '#', 'System starts at z=hi=', run%p%hi
    write(u, '(A1,A12,I10)') '#', 'System ends   at z=hf=', run%p%hf
    write(u, '(A1,A12,I10)') '#', 'aix=', run%aix
    write(u, '(A1,A12,I10)') '#', 'afx=', run%afx
    write(u, '(A1,A12,I2)') '#', 'precision=', run%p%precision
  end subroutine iris_log_parameters

  ! ------------------------------------------------------------------
  ! bind(C) interface for non-Fortran hosts
  ! ------------------------------------------------------------------

  ! Load an input file into the module's run record.  Returns 0 or an error
  ! code; iris_host_error() gives the message.
  integer(c_int) function iris_host_load(path, path_len, precision_default, p) bind(C, name="iris_host_load")
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int32_t), value :: path_len, precision_default
    type(iris_params_t), intent(out) :: p
    character(len=path_len) :: fpath
    integer :: i, ierr
    do i = 1, path_len
      fpath(i:i) = path(i)
    end do
    if (allocated(c_run%eps_inv)) deallocate(c_run%eps_inv)
    call iris_load_input(fpath, c_run, ierr, c_errmsg, int(precision_default))
    iris_host_load = ierr
    if (ierr == 0) p = c_run%p
  end function iris_host_load

  ! Copy the arrays of the last loaded run into caller buffers sized from the
  ! parameter block: eps_inv complex (nx+2)(ny+2)(hf-hi+3); fpzmn/fmzmn
  ! 3*pmlzi; fpzmx/fmzmx 3*(nz-pmlzf+1); fw nof; spectra 3*nof.
  integer(c_int) function iris_host_arrays(eps_inv, fpzmn, fmzmn, fpzmx, fmzmx, fw, spectra) &
      bind(C, name="iris_host_arrays")
    type(c_ptr), value :: eps_inv, fpzmn, fmzmn, fpzmx, fmzmx, fw, spectra
    complex(c_double_complex), pointer :: pe(:)
    real(c_double), pointer :: pr(:)
    if (.not. allocated(c_run%eps_inv)) then
      iris_host_arrays = -1
      return
    end if
    call c_f_pointer(eps_inv, pe, [c_run%ngrid_eps]);  pe = reshape(c_run%eps_inv, [c_run%ngrid_eps])
    call c_f_pointer(fpzmn, pr, [size(c_run%fpzmn)]);  pr = reshape(c_run%fpzmn, [size(c_run%fpzmn)])
    call c_f_pointer(fmzmn, pr, [size(c_run%fmzmn)]);  pr = reshape(c_run%fmzmn, [size(c_run%fmzmn)])
    call c_f_pointer(fpzmx, pr, [size(c_run%fpzmx)]);  pr = reshape(c_run%fpzmx, [size(c_run%fpzmx)])
    call c_f_pointer(fmzmx, pr, [size(c_run%fmzmx)]);  pr = reshape(c_run%fmzmx, [size(c_run%fmzmx)])
    call c_f_pointer(fw, pr, [size(c_run%fw)]);        pr = c_run%fw
    call c_f_pointer(spectra, pr, [size(c_run%spectra)]); pr = reshape(c_run%spectra, [size(c_run%spectra)])
    iris_host_arrays = 0
  end function iris_host_arrays

  ! Input-file quantities that are not part of the parameter block:
  ! out(1:8) = q_nm, texit_fs, dlamda, lno, lnf, pump_DT(fs), def_pulse, wzPML(a.u.)
  integer(c_int) function iris_host_extra(out, iout) bind(C, name="iris_host_extra")
    real(c_double), intent(out) :: out(8)
    integer(c_int32_t), intent(out) :: iout(8)
    out = [c_run%q_nm, c_run%texit_fs, c_run%dlamda, c_run%lno, c_run%lnf, c_run%pump_dt_fs, &
           c_run%def_pulse, c_run%wzpml]
    ! iout = no_downloads (input), OUTPUT_TIME_STEP, Dh, dx, aix, afx, SDL on, tfsf_zi
    iout = [c_run%no_downloads_in, c_run%output_time_step, c_run%dh, c_run%dx, c_run%aix, c_run%afx, &
            sum(c_run%sdl_plane_on), c_run%tfsf_zi]
    iris_host_extra = 0
  end function iris_host_extra

  ! Snapshot options of the last loaded run: iout = plane_on(3), copy_flag,
  ! copy_no, module_e, field(6), nlayers(3), then 1 when any plane is on.
  integer(c_int) function iris_host_sdl_config(iout) bind(C, name="iris_host_sdl_config")
    integer(c_int32_t), intent(out) :: iout(16)
    iout = [c_run%sdl%plane_on, c_run%sdl%copy_flag, c_run%sdl%copy_no, c_run%sdl%module_e, &
            c_run%sdl%field, c_run%sdl%nlayers, merge(1, 0, any(c_run%sdl%plane_on == 1))]
    iris_host_sdl_config = 0
  end function iris_host_sdl_config

  ! Plane of layer `ilayer` of snapshot plane `layer`: geo = fixed
  ! coordinate, a0, a1, b0, b1 (free ranges).
  integer(c_int) function iris_host_sdl_geometry(layer, ilayer, geo) bind(C, name="iris_host_sdl_geometry")
    integer(c_int32_t), value :: layer, ilayer
    integer(c_int32_t), intent(out) :: geo(5)
    integer :: fx, a0, a1, b0, b1
    if (layer < 1 .or. layer > 3 .or. ilayer < 1 .or. ilayer > c_run%sdl%nlayers(max(1, min(3, layer)))) then
      iris_host_sdl_geometry = -1
      return
    end if
    call sdl_geometry(c_run%sdl, int(layer), int(ilayer), fx, a0, a1, b0, b1)
    geo = [fx, a0, a1, b0, b1]
    iris_host_sdl_geometry = 0
  end function iris_host_sdl_geometry

  ! Write the snapshot files of one plane of the last loaded run into the
  ! current directory, from caller-supplied fields on that plane: e and hp
  ! are complex (na, nb, 3) (E after the step, H of the step before), ino the
  ! output number.  Lets a caller without the engine (the tests, with the
  ! oracle's fields) drive the same writer as the program.
  integer(c_int) function iris_host_sdl_write(layer, ilayer, ino, e, hp) bind(C, name="iris_host_sdl_write")
    integer(c_int32_t), value :: layer, ilayer, ino
    type(c_ptr), value :: e, hp
    complex(c_double_complex), pointer :: pe(:, :, :), ph(:, :, :)
    integer :: fx, a0, a1, b0, b1, ierr
    character(len=200) :: msg
    if (layer < 1 .or. layer > 3 .or. ilayer < 1 .or. ilayer > c_run%sdl%nlayers(max(1, min(3, layer)))) then
      iris_host_sdl_write = -1
      return
    end if
    call sdl_geometry(c_run%sdl, int(layer), int(ilayer), fx, a0, a1, b0, b1)
    call c_f_pointer(e, pe, [a1 - a0 + 1, b1 - b0 + 1, 3])
    call c_f_pointer(hp, ph, [a1 - a0 + 1, b1 - b0 + 1, 3])
    call sdl_write_plane(c_run%sdl, int(layer), fx, a0, b0, int(ino), pe, ph, c_run%p%q*c_run%p%eamp, ierr, msg)
    if (ierr /= 0) c_errmsg = msg
    iris_host_sdl_write = ierr
  end function iris_host_sdl_write

  ! Message of the last failed iris_host_load, NUL-terminated into buf.
  integer(c_int) function iris_host_error(buf, buflen) bind(C, name="iris_host_error")
    character(kind=c_char), intent(out) :: buf(*)
    integer(c_int32_t), value :: buflen
    integer :: i, n
    n = min(len_trim(c_errmsg), int(buflen) - 1)
    do i = 1, n
      buf(i) = c_errmsg(i:i)
    end do
    buf(n + 1) = c_null_char
    iris_host_error = n
  end function iris_host_error

end module iris_setup
