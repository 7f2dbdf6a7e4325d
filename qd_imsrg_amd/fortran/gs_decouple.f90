! gs_decouple N hw R ML MS [cut] -- the ODE ground-state decoupling of the
! reference's gs_IMSRG program, on the GPU.
!
! Same program name, command line and outputs as the reference (src/makefile:6
! builds gs_IMSRG.f90 as gs_decouple; build_job_script.py:223 runs
! "./gs_decouple N hwd0 R 0 1 20"):
!   * N, R, ML, MS read with (I5), hw with (f5.2) from a CHARACTER(5)
!     (src/gs_IMSRG.f90:35-56);
!   * 'Hartree-Fock Energy: ' line, then the per-output line '(I5,3(e14.6))'
!     (src/gs_IMSRG.f90:66,117);
!   * ML and MS are read (the reference's EOM targets, src/gs_IMSRG.f90:
!     122-129);
!   * '(I4,I5,3(f15.9),f15.4)' emax, M, E_HF, E_HF + MBPT2, E_gs, log(wall
!     seconds) appended to ../output/TRAD_<N>_<hw>.dat (src/gs_IMSRG.f90:
!     134-139);
!   * H(s) written by export_hamiltonian to ../hamiltonians/ham_<N>_<hw>_<R>
!     (src/gs_IMSRG.f90:141-142).
! H(0) is the Hartree-Fock Hamiltonian from the analytic Coulomb elements;
! dH/ds = dGam(H) (White generator, calc_cc-filled CC inputs) is integrated by
! the Shampine-Gordon Adams code in output intervals of 0.1 until |dE/E| over
! one interval is below 1e-7.  A file the reference would fail to open
! (missing ../output or ../hamiltonians) is reported on stderr and skipped.
program gs_decouple
  use iso_c_binding
  use qdimsrg_c
  implicit none
  character(5) :: nstr, hwstr, emaxstr, mlstr, msstr
  integer :: n, emax, m, ml, ms, u, ios
  integer(8) :: c0, c1, rate
  real(8) :: hw, e2nd, wall
  type(c_ptr) :: ctx, h0, hs
  type(qd_info) :: info
  type(qd_gs_params) :: p
  type(qd_gs_result) :: res
  real(c_double), allocatable :: vec(:), eig(:)
  real(c_double) :: trace(4), e2
  integer(c_int) :: it

  call system_clock(c0, rate)
  call getarg(1, nstr)
  call getarg(2, hwstr)
  call getarg(3, emaxstr)
  call getarg(4, mlstr)
  call getarg(5, msstr)
  read(nstr, '(I5)') n
  read(hwstr, '(f5.2)') hw
  read(emaxstr, '(I5)') emax
  ml = 0
  ms = 0
  if (len_trim(mlstr) > 0) read(mlstr, '(I5)') ml
  if (len_trim(msstr) > 0) read(msstr, '(I5)') ms
  m = emax * (emax + 1)
  write(hwstr, '(f5.2)') hw
  nstr = adjustl(nstr)
  hwstr = adjustl(hwstr)
  emaxstr = adjustl(emaxstr)

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

  call system_clock(c1)
  wall = real(c1 - c0, 8) / real(rate, 8)
  open(newunit=u, file='../output/TRAD_'//trim(nstr)//'_'//trim(hwstr)//'.dat', position='append', &
       action='write', iostat=ios)
  if (ios == 0) then
     write(u, '(I4,I5,3(f15.9),f15.4)') emax, m, vec(1), e2nd, res%e0, log(wall)
     close(u)
  else
     write(0, *) 'gs_decouple: cannot open ../output/TRAD_'//trim(nstr)//'_'//trim(hwstr)//'.dat'
  end if
  if (qd_export_hamiltonian(hs, '../hamiltonians/ham_'//trim(nstr)//'_'//trim(hwstr)//'_'//trim(emaxstr)// &
       c_null_char) /= QD_OK) then
     write(0, *) 'gs_decouple: cannot write ../hamiltonians/ham_'//trim(nstr)//'_'//trim(hwstr)//'_'// &
          trim(emaxstr)
  end if

  call qd_check(qd_op_destroy(h0), 'destroy')
  call qd_check(qd_op_destroy(hs), 'destroy')
  call qd_check(qd_context_destroy(ctx), 'destroy')
end program gs_decouple
