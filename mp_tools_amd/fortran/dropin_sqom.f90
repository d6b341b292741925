! dropin_sqom.f90 -- Fortran caller of libmpsq with the reference's call shape.
!
! Reads nj points (x, y, strengths) from a raw FP64 file and computes the
! type-1 transform twice:
!   1. through the implicit-interface call the reference makes, with the
!      same argument kinds (integer(8) counts, real(8) coordinates,
!      complex(8) strengths and modes, default integer iflag/ier, real(8) eps
!      and an unallocated integer(8) allocatable passed for the options):
!        call finufft2d1(n_fft,xf,yf,cf,iflag,eps_fft,n_qx8,n_qy8,ampl_tot,nul_opt,ier)
!      (code/src/mp_sqom56.f90:94-106, 195-196, 1148);
!   2. through the explicit plan interface of module mpsq, reusing one plan
!      for two strength vectors (the per-species loop of
!      code/src/mp_sqom56.f90:1101-1156).
! Writes both mode arrays (CMCL order, k1 fastest) to a raw FP64 file.
!
! usage: dropin_sqom <in.bin> <out.bin>
!   in.bin : int64 nj, int64 ms, int64 mt, real(8) eps, x(nj), y(nj), c(2*nj)
!   out.bin: ier, rc_plan (int32), f_simple(2*ms*mt), f_plan(2*ms*mt), f_plan2(2*ms*mt)
program dropin_sqom
  use mpsq
  implicit none
  integer(8) :: n_fft, n_qx8, n_qy8
  integer(8), allocatable :: nul_opt
  real(8), allocatable :: xf(:), yf(:)
  complex(8), allocatable :: cf(:), ampl_tot(:), f_plan(:), f_plan2(:)
  real(8) :: eps_fft
  integer :: iflag, ier, u, rc
  character(len=512) :: fin, fout
  type(c_ptr) :: plan
  integer(c_int64_t) :: nm(3)
  real(8) :: dummy(1)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n_fft, n_qx8, n_qy8, eps_fft
  allocate(xf(n_fft), yf(n_fft), cf(n_fft))
  read(u) xf, yf, cf
  close(u)
  allocate(ampl_tot(n_qx8*n_qy8), f_plan(n_qx8*n_qy8), f_plan2(n_qx8*n_qy8))
  iflag = -1
  ier = -99

  call finufft2d1(n_fft,xf,yf,cf,iflag,eps_fft,n_qx8,n_qy8,ampl_tot,nul_opt,ier)

  nm = [n_qx8, n_qy8, 1_8]
  rc = mpsq_makeplan(1, 2, nm, iflag, 1, eps_fft, plan, c_null_ptr)
  if (rc < 2) rc = max(rc, mpsq_setpts(plan, n_fft, xf, yf, dummy))
  if (rc < 2) rc = max(rc, mpsq_execute(plan, cf, f_plan))
  if (rc < 2) rc = max(rc, mpsq_execute(plan, 2.0d0 * cf, f_plan2))
  if (c_associated(plan)) rc = max(rc, mpsq_destroy(plan))

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) int(ier, 4), int(rc, 4), ampl_tot, f_plan, f_plan2
  close(u)
end program dropin_sqom
