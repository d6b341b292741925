! mpsq_mod.f90 -- ISO_C_BINDING declarations of libmpsq.so for Fortran callers.
!
! The reference tools call FINUFFT through an implicit interface,
!   call finufft2d1(n_fft,xf,yf,cf,iflag,eps_fft,n_qx8,n_qy8,ampl_tot,nul_opt,ier)
! (code/src/mp_sqom56.f90:1148, code/src/mp_sql56.f90:1129); libmpsq exports
! finufft1d1_/finufft2d1_/finufft3d1_ with exactly that by-reference argument
! list, so those call sites relink unchanged.  This module adds explicit
! interfaces to the plan API (makeplan / setpts / execute / destroy, FP64) for
! code that wants to reuse one device plan across frames and species, and to
! the entry points that replace mp_sql's integration loops
! (code/src/mp_sql56.f90:1146-1254), mp_dos's FFT and species sums
! (code/src/mp_dos56.f90:501-525) and mp_sqom's speckle filter
! (code/src/mp_sqom56.f90:1530-1570).  type(c_ptr) arguments named d_* are
! device addresses; a_cell_inv is row-major (element (i, j) at 3 (i-1) + j),
! i.e. the transpose of a Fortran a_cell_inv(3,3).
! Return codes: 0 ok, 1 warning (eps below the floor), >= 2 error with the
! outputs zero-filled.
module mpsq
  use, intrinsic :: iso_c_binding
  implicit none

  type, bind(C) :: mpsq_opts
     integer(c_int) :: device = 0
     type(c_ptr) :: stream = c_null_ptr
     integer(c_int) :: tile_by = 0
     integer(c_int) :: tile_bz = 0
     integer(c_int) :: sort_chunk = 0
     integer(c_int) :: debug = 0
  end type mpsq_opts

  interface
     integer(c_int) function mpsq_makeplan(ntype, dim, n_modes, iflag, ntrans, eps, plan, opts) &
          bind(C, name="mpsq_makeplan")
       import :: c_int, c_int64_t, c_double, c_ptr
       integer(c_int), value :: ntype, dim, iflag, ntrans
       integer(c_int64_t), intent(in) :: n_modes(3)
       real(c_double), value :: eps
       type(c_ptr), intent(out) :: plan
       type(c_ptr), value :: opts
     end function mpsq_makeplan

     integer(c_int) function mpsq_setpts(plan, m, x, y, z) bind(C, name="mpsq_setpts")
       import :: c_int, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: plan
       integer(c_int64_t), value :: m
       real(c_double), intent(in) :: x(*), y(*), z(*)
     end function mpsq_setpts

     integer(c_int) function mpsq_execute(plan, c, f) bind(C, name="mpsq_execute")
       import :: c_int, c_double_complex, c_ptr
       type(c_ptr), value :: plan
       complex(c_double_complex), intent(in) :: c(*)
       complex(c_double_complex), intent(out) :: f(*)
     end function mpsq_execute

     integer(c_int) function mpsq_destroy(plan) bind(C, name="mpsq_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: plan
     end function mpsq_destroy

     integer(c_int) function mpsq_nufft2d1(nj, xj, yj, cj, iflag, eps, ms, mt, fk, opts) &
          bind(C, name="mpsq_nufft2d1")
       import :: c_int, c_int64_t, c_double, c_double_complex, c_ptr
       integer(c_int64_t), value :: nj, ms, mt
       real(c_double), intent(in) :: xj(*), yj(*)
       complex(c_double_complex), intent(in) :: cj(*)
       integer(c_int), value :: iflag
       real(c_double), value :: eps
       complex(c_double_complex), intent(out) :: fk(*)
       type(c_ptr), value :: opts
     end function mpsq_nufft2d1

     integer(c_int) function mpsq_nufft3d1(nj, xj, yj, zj, cj, iflag, eps, ms, mt, mu, fk, opts) &
          bind(C, name="mpsq_nufft3d1")
       import :: c_int, c_int64_t, c_double, c_double_complex, c_ptr
       integer(c_int64_t), value :: nj, ms, mt, mu
       real(c_double), intent(in) :: xj(*), yj(*), zj(*)
       complex(c_double_complex), intent(in) :: cj(*)
       integer(c_int), value :: iflag
       real(c_double), value :: eps
       complex(c_double_complex), intent(out) :: fk(*)
       type(c_ptr), value :: opts
     end function mpsq_nufft3d1

     integer(c_int) function mpsq_sql_hist(d_amp, nsp, precision, n_modes, dq, a_cell_inv, q_range, q_step, &
                                           n_hist, d_hist, stream) bind(C, name="mpsq_sql_hist")
       import :: c_int, c_int64_t, c_float, c_ptr
       type(c_ptr), value :: d_amp, d_hist, stream
       integer(c_int), value :: nsp, precision, n_hist
       integer(c_int64_t), intent(in) :: n_modes(3)
       real(c_float), intent(in) :: dq(3), a_cell_inv(9)
       real(c_float), value :: q_range, q_step
     end function mpsq_sql_hist

     ! seed: the 64 bits of the generator's key; tag: its 32-bit stream number
     integer(c_int) function mpsq_sql_hist_mc(d_amp, nsp, precision, n_modes, dq, a_cell_inv, q_range, q_step, &
                                              n_hist, n_h, n_mc, seed, tag, d_hist, d_skip, stream) &
                                              bind(C, name="mpsq_sql_hist_mc")
       import :: c_int, c_int32_t, c_int64_t, c_float, c_ptr
       type(c_ptr), value :: d_amp, d_hist, d_skip, stream
       integer(c_int), value :: nsp, precision, n_hist
       integer(c_int64_t), intent(in) :: n_modes(3)
       real(c_float), intent(in) :: dq(3), a_cell_inv(9)
       real(c_float), value :: q_range, q_step
       integer(c_int64_t), value :: n_h, n_mc, seed
       integer(c_int32_t), value :: tag
     end function mpsq_sql_hist_mc

     integer(c_int) function mpsq_dos(d_vel, nfile, n_tot, d_at_ind, nsp, n_freq, precision, d_dos, stream) &
                                      bind(C, name="mpsq_dos")
       import :: c_int, c_int64_t, c_ptr
       type(c_ptr), value :: d_vel, d_at_ind, d_dos, stream
       integer(c_int64_t), value :: nfile, n_tot, n_freq
       integer(c_int), value :: nsp, precision
     end function mpsq_dos

     ! host array; column_walk 0 walks the first index fastest
     integer(c_int) function mpsq_speckle_filter(map, n_rows, n_cols, column_walk, s_trig) &
                                                 bind(C, name="mpsq_speckle_filter")
       import :: c_int, c_int64_t, c_double
       real(c_double), intent(inout) :: map(*)
       integer(c_int64_t), value :: n_rows, n_cols
       integer(c_int), value :: column_walk
       real(c_double), value :: s_trig
     end function mpsq_speckle_filter
  end interface
end module mpsq
