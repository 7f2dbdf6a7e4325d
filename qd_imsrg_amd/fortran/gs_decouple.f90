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
!   * EOM-IM-SRG on the decoupled H(s) with 5 states (src/gs_IMSRG.f90:122-
!     129): for MS = 0 the excited states of (ML, MS), otherwise one particle
!     attached and one removed, each printing the reference's block and
!     appending '(4(I5),f8.2,4(f17.9))' to ../output/EOM_particle_attached.dat
!     / EOM_particle_removed.dat (src/EOM_IMSRG.f90:10-128);
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
  call qd_check(qd_op_create(ctx, 1_c_int, h0), 'op')
  call qd_check(qd_op_create(ctx, 1_c_int, hs), 'op')
  call qd_check(qd_hf_hamiltonian_device(ctx, real(hw, c_double), h0, eig, it), 'Hartree-Fock')
  call qd_check(qd_op_download(h0, vec, info%neq), 'download')
  print *, 'Hartree-Fock Energy: ', vec(1)
  flush(6)
  call qd_check(qd_mbpt2(ctx, h0, e2), 'MBPT2')
  e2nd = vec(1) + e2

  call qd_gs_params_default(p)
  p%verbose = 1
  call qd_check(qd_gs_flow(ctx, h0, hs, p, res, trace, 0_c_int), 'gs flow')
  if (ms == 0) then
     call excited_states(ctx, hs, ml, ms, res%e0)
  else
     call attached_removed(ctx, hs, 1, ml, ms, res%e0, emax, n, hw)
     call attached_removed(ctx, hs, 2, ml, ms, res%e0, emax, n, hw)
  end if

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
contains

  ! calculate_excited_states(ML, MS, 5, HS), src/EOM_IMSRG.f90:10-47
  subroutine excited_states(ctx, hs, ml, ms, e0)
    type(c_ptr), intent(in) :: ctx, hs
    integer, intent(in) :: ml, ms
    real(8), intent(in) :: e0
    real(c_double) :: de(5), w(5)
    integer(c_int) :: it
    integer :: k
    print *
    write(*, '((A55),(I2),(A4),(I2))') 'EXECUTING EOM CALCULATION'//' FOR EXCITED STATES: Ml=', ml, ' Ms=', ms
    print *
    call qd_check(qd_eom_excited_states(ctx, hs, int(ml, c_int), int(ms, c_int), 5_c_int, de, w, it), 'EOM')
    write(6, *)
    print *
    write(*, '((A21),(f16.9))') 'Ground State Energy: ', e0
    print *
    print *, 'EXCITED STATE ENERGIES:'
    print *, '=================================='
    print *, '      dE             E_0 + dE'
    print *, '=================================='
    do k = 1, 5
       write(*, '(3(f16.9))') de(k), de(k) + e0, w(k)
    end do
  end subroutine excited_states

  ! calculate_1p_attached (mode 1) / calculate_1h_removed (mode 2) with 5
  ! states, src/EOM_IMSRG.f90:50-128
  subroutine attached_removed(ctx, hs, mode, ml, ms, e0, emax, n, hw)
    type(c_ptr), intent(in) :: ctx, hs
    integer, intent(in) :: mode, ml, ms, emax, n
    real(8), intent(in) :: e0, hw
    real(c_double) :: de(5), w(5)
    integer(c_int) :: it
    integer :: k, u, ios, nshell, nfill
    character(40) :: fname
    nshell = emax
    nfill = nint((sqrt(4.d0 * n + 1.d0) - 1) / 2.d0)
    print *
    if (mode == 1) then
       write(*, '((A65),(I2),(A4),(I2))') 'EXECUTING EOM CALCULATION'//' FOR PARTICLE ATTACHED STATES: Ml=', &
            ml, ' Ms=', ms
    else
       write(*, '((A64),(I2),(A4),(I2))') 'EXECUTING EOM CALCULATION'//' FOR PARTICLE REMOVED STATES: Ml=', &
            ml, ' Ms=', ms
    end if
    print *
    call qd_check(qd_eom_attached_removed(ctx, hs, int(mode, c_int), int(ml, c_int), int(ms, c_int), 5_c_int, &
         de, w, it), 'EOM')
    write(6, *)
    print *
    write(*, '((A21),(f16.9))') 'Ground State Energy: ', e0
    print *
    if (mode == 1) then
       print *, '1p ATTACHED ENERGIES:'
    else
       print *, '1p REMOVED ENERGIES:'
    end if
    print *, '=================================='
    print *, '      dE             E_0 + dE'
    print *, '=================================='
    do k = 1, 5
       write(*, '(3(f16.9))') de(k), de(k) + e0, w(k)
    end do
    if (mode == 1) then
       fname = '../output/EOM_particle_attached.dat'
    else
       fname = '../output/EOM_particle_removed.dat'
    end if
    open(newunit=u, file=trim(fname), position='append', action='write', iostat=ios)
    if (ios == 0) then
       write(u, '(4(I5),f8.2,4(f17.9))') nshell, nfill, ml, ms, hw, e0, de(1), de(1) + e0, w(1)
       close(u)
    else
       write(0, *) 'gs_decouple: cannot open ', trim(fname)
    end if
  end subroutine attached_removed

end program gs_decouple
