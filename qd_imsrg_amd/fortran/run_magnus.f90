! run_magnus N hw R [ML MS cut] [pandya] -- the Magnus IM-SRG(2) ground-state
! flow of the reference's magnus_IMSRG program, on the GPU.
!
! Same command line and stdout contract as the reference (src/magnus_IMSRG.f90:
! 32-56: N read with (I5), hw with (f5.2) from a CHARACTER(5), R with (I5);
! per-step line '(I5,5(f18.9))' and ' faillll' on a rejected step at :218-
! 221).  The reference stops right after the flow (:144), so no summary line
! follows.  H(0) is the Hartree-Fock Hamiltonian built from the analytic
! Coulomb elements; the flow runs as the reference ships it (first-order
! Magnus, zero CC inputs).  An argument "pandya" after the others adds
! calc_cc-filled CC inputs.
program run_magnus
  use iso_c_binding
  use qdimsrg_c
  implicit none
  character(5) :: nstr, hwstr, emaxstr
  character(16) :: arg
  integer :: n, emax, m, k
  logical :: pandya
  real(8) :: hw
  type(c_ptr) :: ctx, h0, hs, om
  type(qd_info) :: info
  type(qd_flow_params) :: p
  type(qd_flow_result) :: res
  real(c_double), allocatable :: vec(:), eig(:)
  real(c_double) :: trace(6)
  integer(c_int) :: it

  call getarg(1, nstr)
  call getarg(2, hwstr)
  call getarg(3, emaxstr)
  pandya = .false.
  do k = 4, command_argument_count()
     call getarg(k, arg)
     if (trim(arg) == 'pandya') pandya = .true.
  end do
  read(nstr, '(I5)') n
  read(hwstr, '(f5.2)') hw
  read(emaxstr, '(I5)') emax
  m = emax * (emax + 1)

  call qd_check(qd_context_create(int(n, c_int), int(emax, c_int), ctx), 'context')
  call qd_check(qd_context_info(ctx, info), 'info')
  allocate(vec(info%neq), eig(m))
  call qd_check(qd_hf_hamiltonian(int(n, c_int), int(emax, c_int), real(hw, c_double), vec, info%neq, eig, it), &
       'Hartree-Fock')
  call qd_check(qd_op_create(ctx, 1_c_int, h0), 'op')
  call qd_check(qd_op_create(ctx, 1_c_int, hs), 'op')
  call qd_check(qd_op_create(ctx, -1_c_int, om), 'op')
  call qd_check(qd_op_upload(h0, vec, info%neq), 'upload')

  call qd_flow_params_default(p)
  p%verbose = 1
  if (pandya) p%pandya = 1
  call qd_check(qd_magnus_flow(ctx, h0, hs, om, p, res, trace, 0_c_int), 'magnus flow')

  call qd_check(qd_op_destroy(h0), 'destroy')
  call qd_check(qd_op_destroy(hs), 'destroy')
  call qd_check(qd_op_destroy(om), 'destroy')
  call qd_check(qd_context_destroy(ctx), 'destroy')
end program run_magnus
