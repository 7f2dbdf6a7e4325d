! run_magnus N hw R [ML MS cut] [pandya] -- the Magnus IM-SRG(2) ground-state
! flow of the reference's magnus_IMSRG program, on the GPU.
!
! Same command line and outputs as the reference (src/magnus_IMSRG.f90):
!   * N read with (I5), hw with (f5.2) from a CHARACTER(5), R with (I5); ML, MS
!     and the state cut with (I5) when all three are given (:32-56);
!   * the per-step line '(I5,5(f18.9))' and ' faillll' on a rejected step
!     (:218-221);
!   * spectrum_magnus.dat: s, E0 and E0 + the TDA eigenvalues of the (ML, MS)
!     block below the cut, at the start and twice per accepted step
!     (:130-136, 165-189, 232-243);
!   * after the flow the EOM-IM-SRG block of calculate_excited_states for
!     (ML, MS) with 10 states (:142; src/EOM_IMSRG.f90:10-47); the reference
!     stops there (:144), so no summary line follows.
! H(0) is the Hartree-Fock Hamiltonian built from the analytic Coulomb
! elements; the flow runs as the reference ships it (first-order Magnus, zero
! CC inputs).  An argument "pandya" after the others adds calc_cc-filled CC
! inputs.  Without ML MS cut the TDA block is empty and the EOM targets are 0.
program run_magnus
  use iso_c_binding
  use qdimsrg_c
  implicit none
  character(5) :: nstr, hwstr, emaxstr, mlstr, msstr, cutstr
  character(16) :: arg
  character(kind=c_char, len=32), target :: specfile
  integer :: n, emax, m, k, ml, ms, cut, ios
  logical :: pandya
  real(8) :: hw, e0
  type(c_ptr) :: ctx, h0, hs, om
  type(qd_info) :: info
  type(qd_flow_params) :: p
  type(qd_flow_result) :: res
  real(c_double), allocatable :: eig(:)
  real(c_double) :: trace(6), de(10), w(10)
  integer(c_int) :: it

  call getarg(1, nstr)
  call getarg(2, hwstr)
  call getarg(3, emaxstr)
  call getarg(4, mlstr)
  call getarg(5, msstr)
  call getarg(6, cutstr)
  pandya = .false.
  do k = 4, command_argument_count()
     call getarg(k, arg)
     if (trim(arg) == 'pandya') pandya = .true.
  end do
  ml = 0
  ms = 0
  cut = 0
  if (len_trim(adjustl(mlstr)) * len_trim(adjustl(msstr)) * len_trim(adjustl(cutstr)) > 0) then
     read(mlstr, '(I5)', iostat=ios) ml
     if (ios == 0) read(msstr, '(I5)', iostat=ios) ms
     if (ios == 0) read(cutstr, '(I5)', iostat=ios) cut
     if (ios /= 0) then  ! the fourth argument was "pandya"
        ml = 0
        ms = 0
        cut = 0
     end if
  end if
  read(nstr, '(I5)') n
  read(hwstr, '(f5.2)') hw
  read(emaxstr, '(I5)') emax
  m = emax * (emax + 1)

  call qd_check(qd_context_create(int(n, c_int), int(emax, c_int), ctx), 'context')
  call qd_check(qd_context_info(ctx, info), 'info')
  allocate(eig(m))
  call qd_check(qd_op_create(ctx, 1_c_int, h0), 'op')
  call qd_check(qd_op_create(ctx, 1_c_int, hs), 'op')
  call qd_check(qd_op_create(ctx, -1_c_int, om), 'op')
  call qd_check(qd_hf_hamiltonian_device(ctx, real(hw, c_double), h0, eig, it), 'Hartree-Fock')

  call qd_flow_params_default(p)
  p%verbose = 1
  if (pandya) p%pandya = 1
  p%tda_ml = ml
  p%tda_ms = ms
  p%tda_cut = cut
  specfile = 'spectrum_magnus.dat'//c_null_char
  p%spectrum_path = c_loc(specfile)
  call qd_check(qd_magnus_flow(ctx, h0, hs, om, p, res, trace, 0_c_int), 'magnus flow')
  e0 = res%e0

  ! calculate_excited_states(HS%Mltarg, HS%Mstarg, 10, HS), src/EOM_IMSRG.f90:10-47
  print *
  write(*, '((A55),(I2),(A4),(I2))') 'EXECUTING EOM CALCULATION'//' FOR EXCITED STATES: Ml=', ml, ' Ms=', ms
  print *
  call qd_check(qd_eom_excited_states(ctx, hs, int(ml, c_int), int(ms, c_int), 10_c_int, de, w, it), 'EOM')
  write(6, *)
  print *
  write(*, '((A21),(f16.9))') 'Ground State Energy: ', e0
  print *
  print *, 'EXCITED STATE ENERGIES:'
  print *, '=================================='
  print *, '      dE             E_0 + dE'
  print *, '=================================='
  do k = 1, 10
     write(*, '(3(f16.9))') de(k), de(k) + e0, w(k)
  end do

  call qd_check(qd_op_destroy(h0), 'destroy')
  call qd_check(qd_op_destroy(hs), 'destroy')
  call qd_check(qd_op_destroy(om), 'destroy')
  call qd_check(qd_context_destroy(ctx), 'destroy')
end program run_magnus
