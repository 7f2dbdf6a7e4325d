! iso_c_binding interface to libqdimsrg.so (include/qdimsrg.h).
!
! A Fortran caller of the reference's operator API (src/commutators.f90,
! src/magnus_IMSRG.f90, src/gs_IMSRG.f90) uses these procedures on opaque
! device handles; full_ham data crosses as the reference's vectorize vector
! (src/ME_general.f90:1871-1989) through qd_op_upload / qd_op_download.
module qdimsrg_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: QD_OK = 0

  type, bind(C) :: qd_info
     integer(c_int64_t) :: nbody, nshell, msp, nblock, ncc, neq, cc_size, n2b
  end type qd_info

  type, bind(C) :: qd_flow_params
     real(c_double) :: stp, conv
     integer(c_int) :: max_iter, pandya, verbose, warm_start, diis, stop_rule
     real(c_double) :: conv_deds
     integer(c_int) :: tda_ml, tda_ms, tda_cut
     type(c_ptr) :: spectrum_path
  end type qd_flow_params

  type, bind(C) :: qd_gs_params
     real(c_double) :: relerr, abserr, stp, conv
     integer(c_int) :: max_outputs, verbose, maxnum
  end type qd_gs_params

  type, bind(C) :: qd_gs_result
     real(c_double) :: s, e0, crit
     integer(c_int) :: outputs, nsteps
     integer(c_int64_t) :: nfe
     integer(c_int) :: iflag, converged
  end type qd_gs_result

  type, bind(C) :: qd_flow_result
     real(c_double) :: s, e0, crit
     integer(c_int) :: iterations, rejected
     integer(c_int64_t) :: bch_terms
     integer(c_int) :: converged, trace_rows
     real(c_double) :: mbpt2, deds
  end type qd_flow_result

  interface
     function qd_context_create(nbody, nshell, ctx) bind(C, name="qd_context_create")
       import :: c_int, c_ptr
       integer(c_int), value :: nbody, nshell
       type(c_ptr) :: ctx
       integer(c_int) :: qd_context_create
     end function
     function qd_context_destroy(ctx) bind(C, name="qd_context_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: qd_context_destroy
     end function
     function qd_context_info(ctx, info) bind(C, name="qd_context_info")
       import :: c_int, c_ptr, qd_info
       type(c_ptr), value :: ctx
       type(qd_info) :: info
       integer(c_int) :: qd_context_info
     end function
     function qd_op_create(ctx, herm, op) bind(C, name="qd_op_create")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: herm
       type(c_ptr) :: op
       integer(c_int) :: qd_op_create
     end function
     function qd_op_destroy(op) bind(C, name="qd_op_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: op
       integer(c_int) :: qd_op_destroy
     end function
     function qd_op_set_herm(op, herm) bind(C, name="qd_op_set_herm")
       import :: c_int, c_ptr
       type(c_ptr), value :: op
       integer(c_int), value :: herm
       integer(c_int) :: qd_op_set_herm
     end function
     function qd_op_upload(op, vec, neq) bind(C, name="qd_op_upload")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: op
       real(c_double) :: vec(*)
       integer(c_int64_t), value :: neq
       integer(c_int) :: qd_op_upload
     end function
     function qd_op_download(op, vec, neq) bind(C, name="qd_op_download")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: op
       real(c_double) :: vec(*)
       integer(c_int64_t), value :: neq
       integer(c_int) :: qd_op_download
     end function
     function qd_cc_create(ctx, herm, cc) bind(C, name="qd_cc_create")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: herm
       type(c_ptr) :: cc
       integer(c_int) :: qd_cc_create
     end function
     function qd_cc_destroy(cc) bind(C, name="qd_cc_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: cc
       integer(c_int) :: qd_cc_destroy
     end function
     ! ---- commutators (src/commutators.f90) ----
     function qd_commutator_110(ctx, r1, r2, c) bind(C, name="qd_commutator_110")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, r1, r2
       real(c_double) :: c
       integer(c_int) :: qd_commutator_110
     end function
     function qd_commutator_220(ctx, r1, r2, c) bind(C, name="qd_commutator_220")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, r1, r2
       real(c_double) :: c
       integer(c_int) :: qd_commutator_220
     end function
     function qd_commutator_111(ctx, r1, r2, r3) bind(C, name="qd_commutator_111")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, r1, r2, r3
       integer(c_int) :: qd_commutator_111
     end function
     function qd_commutator_121(ctx, r1, r2, r3) bind(C, name="qd_commutator_121")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, r1, r2, r3
       integer(c_int) :: qd_commutator_121
     end function
     function qd_commutator_221(ctx, r1, r2, r3) bind(C, name="qd_commutator_221")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, r1, r2, r3
       integer(c_int) :: qd_commutator_221
     end function
     function qd_commutator_122(ctx, r1, r2, r3) bind(C, name="qd_commutator_122")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, r1, r2, r3
       integer(c_int) :: qd_commutator_122
     end function
     function qd_commutator_222(ctx, r1, r2, r3, hcc, etacc) bind(C, name="qd_commutator_222")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, r1, r2, r3, hcc, etacc
       integer(c_int) :: qd_commutator_222
     end function
     function qd_commutator(ctx, r1, r2, r3, hcc, etacc, e0) bind(C, name="qd_commutator")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, r1, r2, r3, hcc, etacc
       real(c_double), intent(out) :: e0
       integer(c_int) :: qd_commutator
     end function
     function qd_commutator_cc(ctx, r1, r2, r3, hcc, etacc, e0) bind(C, name="qd_commutator_cc")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, r1, r2, r3, hcc, etacc
       real(c_double), intent(out) :: e0
       integer(c_int) :: qd_commutator_cc
     end function
     function qd_calc_cc(ctx, h, hcc) bind(C, name="qd_calc_cc")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, h, hcc
       integer(c_int) :: qd_calc_cc
     end function
     function qd_build_white(ctx, gen, rec) bind(C, name="qd_build_white")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, gen, rec
       integer(c_int) :: qd_build_white
     end function
     function qd_add_arrays(ctx, r1, s1, r2, s2, r3) bind(C, name="qd_add_arrays")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, r1, r2, r3
       real(c_double), value :: s1, s2
       integer(c_int) :: qd_add_arrays
     end function
     function qd_copy_arrays(ctx, r1, r2) bind(C, name="qd_copy_arrays")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, r1, r2
       integer(c_int) :: qd_copy_arrays
     end function
     function qd_mat_2_norm(ctx, r, x) bind(C, name="qd_mat_2_norm")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, r
       real(c_double) :: x
       integer(c_int) :: qd_mat_2_norm
     end function
     function qd_mbpt2(ctx, h, x) bind(C, name="qd_mbpt2")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, h
       real(c_double) :: x
       integer(c_int) :: qd_mbpt2
     end function
     ! ---- flow level (src/magnus_IMSRG.f90, src/gs_IMSRG.f90) ----
     function qd_bch_expand(ctx, hs, g, h, hcc, occ, nterms) bind(C, name="qd_bch_expand")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, hs, g, h, hcc, occ
       integer(c_int) :: nterms
       integer(c_int) :: qd_bch_expand
     end function
     function qd_euler_step(ctx, g, dg, s, stp) bind(C, name="qd_euler_step")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, g, dg
       real(c_double) :: s
       real(c_double), value :: stp
       integer(c_int) :: qd_euler_step
     end function
     function qd_dgam(ctx, h, der, hcc, etacc) bind(C, name="qd_dgam")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, h, der, hcc, etacc
       integer(c_int) :: qd_dgam
     end function
     subroutine qd_flow_params_default(p) bind(C, name="qd_flow_params_default")
       import :: qd_flow_params
       type(qd_flow_params) :: p
     end subroutine
     function qd_magnus_flow(ctx, h0, hs, omega, p, res, trace, trace_len) bind(C, name="qd_magnus_flow")
       import :: c_int, c_ptr, c_double, qd_flow_params, qd_flow_result
       type(c_ptr), value :: ctx, h0, hs, omega
       type(qd_flow_params) :: p
       type(qd_flow_result) :: res
       real(c_double) :: trace(*)
       integer(c_int), value :: trace_len
       integer(c_int) :: qd_magnus_flow
     end function
     ! ---- the ODE path (src/adams_ode.f90:7, src/gs_IMSRG.f90:100-120) ----
     function qd_ode_create(ctx, ode) bind(C, name="qd_ode_create")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr) :: ode
       integer(c_int) :: qd_ode_create
     end function
     function qd_ode_destroy(ode) bind(C, name="qd_ode_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ode
       integer(c_int) :: qd_ode_destroy
     end function
     function qd_ode_integrate(ode, y, t, tout, relerr, abserr, iflag) bind(C, name="qd_ode_integrate")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ode, y
       real(c_double) :: t, relerr, abserr
       real(c_double), value :: tout
       integer(c_int) :: iflag
       integer(c_int) :: qd_ode_integrate
     end function
     subroutine qd_gs_params_default(p) bind(C, name="qd_gs_params_default")
       import :: qd_gs_params
       type(qd_gs_params) :: p
     end subroutine
     function qd_gs_flow(ctx, h0, hs, p, res, trace, trace_len) bind(C, name="qd_gs_flow")
       import :: c_int, c_ptr, c_double, qd_gs_params, qd_gs_result
       type(c_ptr), value :: ctx, h0, hs
       type(qd_gs_params) :: p
       type(qd_gs_result) :: res
       real(c_double) :: trace(*)
       integer(c_int), value :: trace_len
       integer(c_int) :: qd_gs_flow
     end function
     ! ---- batched and precision forms ----
     function qd_set_graphs(ctx, on) bind(C, name="qd_set_graphs")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: on
       integer(c_int) :: qd_set_graphs
     end function
     function qd_autotune(ctx, reps, ms_before, ms_after) bind(C, name="qd_autotune")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: reps
       real(c_double) :: ms_before, ms_after
       integer(c_int) :: qd_autotune
     end function
     function qd_set_precision(ctx, prec) bind(C, name="qd_set_precision")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: prec
       integer(c_int) :: qd_set_precision
     end function
     ! ---- input side ----
     function qd_hf_hamiltonian(nbody, nshell, hw, h0, neq, eig, iters) bind(C, name="qd_hf_hamiltonian")
       import :: c_int, c_double, c_int64_t
       integer(c_int), value :: nbody, nshell
       real(c_double), value :: hw
       real(c_double) :: h0(*), eig(*)
       integer(c_int64_t), value :: neq
       integer(c_int) :: iters
       integer(c_int) :: qd_hf_hamiltonian
     end function
     function qd_hf_hamiltonian_device(ctx, hw, h0, eig, iters) bind(C, name="qd_hf_hamiltonian_device")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx, h0
       real(c_double), value :: hw
       real(c_double) :: eig(*)
       integer(c_int) :: iters
       integer(c_int) :: qd_hf_hamiltonian_device
     end function
     ! ---- consumers of H(s) (src/operators.f90:7-124, src/EOM_IMSRG.f90:10-47) ----
     function qd_eom_excited_states(ctx, h, ml, ms, nev, eig, w1p1h, iters) bind(C, name="qd_eom_excited_states")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, h
       integer(c_int), value :: ml, ms, nev
       real(c_double) :: eig(*), w1p1h(*)
       integer(c_int) :: iters
       integer(c_int) :: qd_eom_excited_states
     end function
     function qd_eom_attached_removed(ctx, h, mode, ml, ms, nev, eig, w1, iters) &
          bind(C, name="qd_eom_attached_removed")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, h
       integer(c_int), value :: mode, ml, ms, nev
       real(c_double) :: eig(*), w1(*)
       integer(c_int) :: iters
       integer(c_int) :: qd_eom_attached_removed
     end function
     function qd_tda_spectrum(ctx, h, ml, ms, cut, eig, cap, dim) bind(C, name="qd_tda_spectrum")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, h
       integer(c_int), value :: ml, ms, cut, cap
       real(c_double) :: eig(*)
       integer(c_int) :: dim
       integer(c_int) :: qd_tda_spectrum
     end function
     ! ---- on-disk format (src/ME_general.f90:207-361) ----
     function qd_export_hamiltonian(op, path) bind(C, name="qd_export_hamiltonian")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: op
       character(kind=c_char), dimension(*) :: path
       integer(c_int) :: qd_export_hamiltonian
     end function
     function qd_import_hamiltonian(op, path) bind(C, name="qd_import_hamiltonian")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: op
       character(kind=c_char), dimension(*) :: path
       integer(c_int) :: qd_import_hamiltonian
     end function
     function qd_last_error() bind(C, name="qd_last_error")
       import :: c_ptr
       type(c_ptr) :: qd_last_error
     end function
  end interface

contains

  subroutine qd_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc /= QD_OK) then
       print *, 'qdimsrg error in ', what, ': code ', rc
       stop 1
    end if
  end subroutine qd_check

end module qdimsrg_c
