! iris_sdl -- field snapshots ("SDL" planes) of the time loop.
!
! At every output step the reference writes, for each enabled plane
! (SDL_plane_on(k) == 1) and each of its extra layers, the files of
! setSDL_FILE (reference code/sdl.f90:923-967) through SDL_output
! (code/sdl.f90:68-517), with the fields e_cur and h_prev
! (code/driver.f90:468-495): E after the step and H of the step before.
! This module restates that writer on the host; the plane values come from
! the engine (iris_get_plane) or, in the tests, from the oracle.
!
!   plane            X: x = xi(1)+deltaX(i), free (y, z) over [yi(1),yf(1)] x [zi(1),zf(1)]
!                    Y: y = yi(2)+deltaY(i), free (x, z) over [xi(2),xf(2)] x [zi(2),zf(2)]
!                    Z: z = zi(3)+deltaZ(i), free (x, y) over [xi(3),xf(3)] x [yi(3),yf(3)]
!   SDLModuleE 1/2   sqrt(sum_c conjg(F_c)*F_c) of E or H, file ".e", tag "_E"/"_H"
!                    (code/sdl.f90:288-316)
!   SDLField(c)>=1   E component c; SDLField(3+c)>=1: H component c.  Both use
!                    the mode SDLField(c) (1 Re, 2 Im, 3 |F|, 4 |Re|, 5 |Im|;
!                    code/sdl.f90:142-153, :210-285): the reference indexes the
!                    mode by the component number for either field.
!   value            real(v) (float32) / (q*Epump) in FP64 (code/sdl.f90:340-346)
!   copies           SDL_copy_flag: the plane repeated SDL_copy_no times along
!                    the first free axis (both axes for Z planes), code/sdl.f90:374-405
!   record           one unformatted record of float32, first free axis fastest
!                    (code/sdl.f90:474-481)
!   file name        "./X_SDL" // tag // n // ext with n = output number
!                    (inttoASCII, code/common.f90:40-101); with extra layers
!                    "./X_SDL" // coordinate // "layer" // tag // n // ext
!                    (code/sdl.f90:175-217)
!
! Where the reference writes memory it never set (an H component whose E
! mode SDLField(c) is 0 or above 5), this writer writes zeros.
module iris_sdl
  implicit none
  private

  integer, parameter :: dp = 8

  type, public :: sdl_config
    integer :: plane_on(3) = 0
    integer :: copy_flag = 0, copy_no = 1
    integer :: xi(3) = 0, xf(3) = 0, yi(3) = 0, yf(3) = 0, zi(3) = 0, zf(3) = 0
    integer :: module_e = 0
    integer :: field(6) = 0
    integer :: nlayers(3) = 1
    integer, allocatable :: delta(:, :)  ! (max layers, 3)
  end type sdl_config

  public :: sdl_enabled, sdl_geometry, sdl_write_plane, sdl_int_ascii, sdl_file_name

contains

  logical function sdl_enabled(cfg)
    type(sdl_config), intent(in) :: cfg
    sdl_enabled = any(cfg%plane_on == 1)
  end function sdl_enabled

  ! Fixed coordinate and free ranges of layer `ilayer` of plane `layer`
  ! (code/driver.f90:470-493 shifts xi(1), yi(2) or zi(3) by the layer's delta).
  subroutine sdl_geometry(cfg, layer, ilayer, fixed, a0, a1, b0, b1)
    type(sdl_config), intent(in) :: cfg
    integer, intent(in) :: layer, ilayer
    integer, intent(out) :: fixed, a0, a1, b0, b1
    integer :: d
    d = 0
    if (allocated(cfg%delta)) d = cfg%delta(ilayer, layer)
    select case (layer)
    case (1)
      fixed = cfg%xi(1) + d; a0 = cfg%yi(1); a1 = cfg%yf(1); b0 = cfg%zi(1); b1 = cfg%zf(1)
    case (2)
      fixed = cfg%yi(2) + d; a0 = cfg%xi(2); a1 = cfg%xf(2); b0 = cfg%zi(2); b1 = cfg%zf(2)
    case default
      fixed = cfg%zi(3) + d; a0 = cfg%xi(3); a1 = cfg%xf(3); b0 = cfg%yi(3); b1 = cfg%yf(3)
    end select
  end subroutine sdl_geometry

  ! Decimal digits of ino as inttoASCII writes them (code/common.f90:40-101),
  ! including its digit arithmetic for five-digit numbers.
  function sdl_int_ascii(ino) result(s)
    integer, intent(in) :: ino
    character(len=5) :: s
    integer :: u, d, c, m, dm, r
    s = ''
    if (ino <= 9) then
      s = achar(ino + 48)
    else if (ino <= 99) then
      s = achar(ino/10 + 48)//achar(mod(ino, 10) + 48)
    else if (ino <= 999) then
      c = ino/100
      u = mod(ino, 10)
      d = (ino - (100*c + u))/10
      s = achar(c + 48)//achar(d + 48)//achar(u + 48)
    else if (ino <= 9999) then
      m = ino/1000
      r = ino - 1000*m
      c = r/100
      u = mod(r, 10)
      d = (r - (100*c + u))/10
      s = achar(m + 48)//achar(c + 48)//achar(d + 48)//achar(u + 48)
    else
      dm = ino/10000
      r = ino - 10000*dm
      m = r/1000
      r = ino - 10000*dm - 1000*m
      c = r/100
      u = mod(r, 10)
      d = (r - (100*c + u))/10
      s = achar(dm + 48)//achar(m + 48)//achar(c + 48)//achar(d + 48)//achar(u + 48)
    end if
  end function sdl_int_ascii

  ! File name of one output (em = 0 modulus, 1 E component, 2 H component).
  function sdl_file_name(cfg, layer, fixed, ino, em, comp) result(name)
    type(sdl_config), intent(in) :: cfg
    integer, intent(in) :: layer, fixed, ino, em, comp
    character(len=100) :: name
    character(len=4) :: ext, tag
    character(len=1), parameter :: axis_letter(3) = ['X', 'Y', 'Z']
    if (em == 0) then
      ext = '.e'
      tag = '_H'
      if (cfg%module_e == 1) tag = '_E'
    else
      if (em == 1) then
        ext = '.e'//achar(48 + comp)
      else
        ext = '.h'//achar(48 + comp)
      end if
      select case (cfg%field(comp))
      case (1); tag = '_R'
      case (2); tag = '_I'
      case (3); tag = '_MC'
      case (4); tag = '_MR'
      case (5); tag = '_MI'
      case default; tag = ''
      end select
    end if
    if (all(cfg%nlayers == 1)) then
      name = './'//axis_letter(layer)//'_SDL'//trim(tag)//trim(sdl_int_ascii(ino))//trim(ext)
    else
      name = './'//axis_letter(layer)//'_SDL'//trim(sdl_int_ascii(fixed))//'layer'//trim(tag)// &
             trim(sdl_int_ascii(ino))//trim(ext)
    end if
  end function sdl_file_name

  ! Every file setSDL_FILE writes for one plane.  e and hp hold E and H_prev
  ! on the plane, (a0:a1, b0:b1, 3) with a0, b0 the first free indices;
  ! norm = q*Epump.
  subroutine sdl_write_plane(cfg, layer, fixed, a0, b0, ino, e, hp, norm, ierr, errmsg)
    type(sdl_config), intent(in) :: cfg
    integer, intent(in) :: layer, fixed, a0, b0, ino
    complex(dp), intent(in) :: e(:, :, :), hp(:, :, :)
    real(dp), intent(in) :: norm
    integer, intent(out) :: ierr
    character(len=*), intent(out) :: errmsg
    integer :: c
    ierr = 0
    errmsg = ''
    if (cfg%module_e == 1) call one_output(0, 0, e)
    if (cfg%module_e == 2) call one_output(0, 0, hp)
    do c = 1, 3
      if (ierr == 0 .and. cfg%field(c) >= 1) call one_output(1, c, e)
    end do
    do c = 1, 3
      if (ierr == 0 .and. cfg%field(3 + c) >= 1) call one_output(2, c, hp)
    end do

  contains

    subroutine one_output(em, comp, f)
      integer, intent(in) :: em, comp
      complex(dp), intent(in) :: f(:, :, :)
      real(dp), allocatable :: output(:, :), aux(:)
      real(dp) :: glog, sdlnorm
      integer :: na, nb, i, j, ic, jc, n, lo1, up1, lo2, up2, ext1, ext2, ti, tj, off, u, ios
      character(len=100) :: name
      na = size(f, 1)
      nb = size(f, 2)
      allocate(output(na, nb))
      output = 0.0_dp
      if (em == 0) then
        ! code/sdl.f90:288-316 (the Y plane takes abs() before sqrt; the sum is >= 0)
        do j = 1, nb
          do i = 1, na
            glog = conjg(f(i, j, 1))*f(i, j, 1)
            glog = glog + conjg(f(i, j, 2))*f(i, j, 2)
            glog = glog + conjg(f(i, j, 3))*f(i, j, 3)
            if (layer == 2) then
              glog = sqrt(abs(glog))
            else
              glog = sqrt(glog)
            end if
            output(i, j) = glog
          end do
        end do
      else
        ! code/sdl.f90:210-285
        do j = 1, nb
          do i = 1, na
            select case (cfg%field(comp))
            case (1)
              output(i, j) = dble(f(i, j, comp))
            case (2)
              output(i, j) = aimag(f(i, j, comp))
            case (3)
              output(i, j) = sqrt(conjg(f(i, j, comp))*f(i, j, comp))
            case (4)
              output(i, j) = sqrt(dble(f(i, j, comp))*dble(f(i, j, comp)))
            case (5)
              output(i, j) = sqrt(aimag(f(i, j, comp))*aimag(f(i, j, comp)))
            end select
          end do
        end do
      end if
      sdlnorm = norm
      if (sdlnorm == 0.0) sdlnorm = 1.0
      output = real(output)/sdlnorm

      name = sdl_file_name(cfg, layer, fixed, ino, em, comp)
      open(newunit=u, file=trim(name), status='replace', form='unformatted', position='rewind', &
           action='write', iostat=ios)
      if (ios /= 0) then
        ierr = -1
        errmsg = 'cannot open snapshot file '//trim(name)
        return
      end if
      if (cfg%copy_flag /= 0) then
        ! output_aux(a0:n*na, b0:b1) -- (a0:n*na, b0:n*nb) for Z planes --
        ! filled with the copies in the reference's loop order; a target
        ! whose first index passes the upper bound lands, in column-major
        ! storage, in the next column, and one past the end of the array is
        ! not part of the record
        n = cfg%copy_no
        lo1 = a0
        up1 = n*na
        lo2 = b0
        up2 = b0 + nb - 1
        if (layer == 3) up2 = n*nb
        ext1 = max(0, up1 - lo1 + 1)
        ext2 = max(0, up2 - lo2 + 1)
        allocate(aux(0:max(ext1*ext2, 1) - 1))
        aux = 0.0_dp
        do ic = 0, n - 1
          do jc = 0, merge(n - 1, 0, layer == 3)
            do i = 1, na
              do j = 1, nb
                ti = a0 + (i - 1) + ic*na
                tj = b0 + (j - 1) + jc*nb
                off = (ti - lo1) + (tj - lo2)*ext1
                if (off >= 0 .and. off < ext1*ext2) aux(off) = output(i, j)
              end do
            end do
          end do
        end do
        write(u) real(aux(0:ext1*ext2 - 1))
      else
        write(u) real(output)
      end if
      close(u)
      print '(1X,A)', trim(name)
    end subroutine one_output
  end subroutine sdl_write_plane

end module iris_sdl
