! run_gs N hw R [ML MS] -- the ODE ground-state decoupling (gs_IMSRG) on the GPU.
!
! Same command line and per-output line as the reference's gs_IMSRG
! (src/gs_IMSRG.f90:35-56: N, R, ML, MS read with (I5), hw with (f5.2);
! per-output line '(I5,3(e14.6))' at :117).  H(0) is the Hartree-Fock
! Hamiltonian built from the analytic Coulomb elements; dH/ds = dGam(H)
! (White generator, calc_cc-filled CC inputs) is integrated by the
! Shampine-Gordon Adams code in output intervals of 0.1 until |dE/E| over one
! interval is below 1e-7.  ML and MS select the reference's excited-state
! targets, which this driver does not compute (no EOM): they are read and
! ignored.  Last line: the reference's TRAD summary '(I4,I5,3(f15.9),f15.4)'
! with the log of the wall time (src/gs_IMSRG.f90:136-139).
program run_gs
  use iso_c_binding
  use qdimsrg_c
  implicit none
  character(5) :: nstr, hwstr, emaxstr
  integer :: n, emax, m
  real(8) :: hw, e2nd, t0, t1
  type(c_ptr) :: ctx, h0, hs
  type(qd_info) :: info
  type(qd_gs_params) :: p
  type(qd_gs_result) :: res
  real(c_double), allocatable :: vec(:), eig(:)
  real(c_double) :: trace(4), e2
  integer(c_int) :: it

  call getarg(1, nstr)
  call getarg(2, hwstr)
  call getarg(3, emaxstr)
  read(nstr, '(I5)') n
  read(hwstr, '(f5.2)') hw
  read(emaxstr, '(I5)') emax
  m = emax * (emax + 1)

  call cpu_time(t0)
  call qd_check(qd_context_create(int(n, c_int), int(emax, c_int), ctx), 'context')
  call qd_check(qd_context_info(ctx, info), 'info')
  allocate(vec(info%neq), eig(m))
  call qd_check(qd_hf_hamiltonian(int(n, c_int), int(emax, c_int), real(hw, c_double), vec, info%neq, eig, it), &
       'Hartree-Fock')
  print *, 'Hartree-Fock Energy: ', vec(1)
  flush(6)
  call qd_check(qd_op_create(ctx, 1_c_int, h0), 'op')
  call qd_check(qd_op_create(ctx, 1_c_int, hs), 'op')
  call qd_check(qd_op_upload(h0, vec, info%neq), 'upload')
  call qd_check(qd_mbpt2(ctx, h0, e2), 'MBPT2')
  e2nd = vec(1) + e2

  call qd_gs_params_default(p)
  p%verbose = 1
  call qd_check(qd_gs_flow(ctx, h0, hs, p, res, trace, 0_c_int), 'gs flow')
  call cpu_time(t1)
  write(*, '(I4,I5,3(f15.9),f15.4)') emax, m, vec(1), e2nd, res%e0, log(t1 - t0)

  call qd_check(qd_op_destroy(h0), 'destroy')
  call qd_check(qd_op_destroy(hs), 'destroy')
  call qd_check(qd_context_destroy(ctx), 'destroy')
end program run_gs
