! Opt-in dump of the reference's on-the-fly DFT accumulators.
!
! TEST INFRASTRUCTURE ONLY: linked into the oracle binary oracle/_ref/irisfdtd_ref
! by oracle/patch_ref.py, never into the product.  The reference writes only the
! face fluxes (code/dataprocess.f90:279-299); the accumulators emg(j)%fdata
! themselves are needed to check DFT fields.  When the environment variable
! IRIS_REF_DUMP is set, the final output call (it == nt) writes
! emg<box>_<face-pair>.bin as a raw stream: int32 n1, int32 n2, then n1*n2
! complex(8) values in Fortran order (points fastest, wavelengths slowest).
subroutine iris_ref_dump(fdata, n1, n2, nobox, j, it, nt)
  implicit none
  integer, intent(in) :: n1, n2, nobox, j, it, nt
  complex(8), intent(in) :: fdata(n1, n2)
  character(len=64) :: fname
  character(len=16) :: flag
  integer :: stat, u

  if (it /= nt) return
  call get_environment_variable('IRIS_REF_DUMP', flag, status=stat)
  if (stat /= 0) return
  write(fname, '(A,I0,A,I0,A)') 'emg', nobox, '_', j, '.bin'
  open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
  write(u) n1, n2
  write(u) fdata
  close(u)
end subroutine iris_ref_dump
