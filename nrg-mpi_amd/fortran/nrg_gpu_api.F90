! iso_c_binding interfaces of libnrg_gpu.so (include/nrg_gpu.h).
module nrg_gpu_api
  use iso_c_binding
  implicit none
  interface
     integer(c_int) function nrg_gpu_init(device) bind(C, name="nrg_gpu_init")
       import :: c_int
       integer(c_int), value :: device
     end function
     integer(c_int) function nrg_gpu_finalize() bind(C, name="nrg_gpu_finalize")
       import :: c_int
     end function
     type(c_ptr) function nrg_gpu_last_error() bind(C, name="nrg_gpu_last_error")
       import :: c_ptr
     end function
     integer(c_int) function nrg_gpu_set_model(nv, nops, nchain, ncm, ncoef, operation, coef_diag_type, &
          coef_diag, coef_inv_type, coef_inv, conservation_block, nmat, coef_spec, cons_diff_spec) &
          bind(C, name="nrg_gpu_set_model")
       import :: c_int, c_double, c_ptr
       integer(c_int), value :: nv, nops, nchain, ncm, ncoef, nmat
       integer(c_int), intent(in) :: operation(*), coef_diag_type(*), coef_inv_type(*), conservation_block(*)
       real(c_double), intent(in) :: coef_diag(*), coef_inv(*)
       type(c_ptr), value :: coef_spec, cons_diff_spec
     end function
     integer(c_int) function nrg_gpu_set_bcs(nbcs, coef_bcs, nrow, chain_bcs) bind(C, name="nrg_gpu_set_bcs")
       import :: c_int, c_double
       integer(c_int), value :: nbcs, nrow
       integer(c_int), intent(in) :: coef_bcs(*)
       real(c_double), intent(in) :: chain_bcs(*)
     end function
     integer(c_int) function nrg_gpu_set_invariant(nold, inv) bind(C, name="nrg_gpu_set_invariant")
       import :: c_int, c_double
       integer(c_int), value :: nold
       real(c_double), intent(in) :: inv(*)
     end function
     integer(c_int) function nrg_gpu_set_invariant_spectrum(nold, inv) bind(C, name="nrg_gpu_set_invariant_spectrum")
       import :: c_int, c_double
       integer(c_int), value :: nold
       real(c_double), intent(in) :: inv(*)
     end function
     integer(c_int) function nrg_gpu_diagonalize(iteration, nsub, subspace_info, nfull, basis_input, nold, &
          eigenvalue_old, chain_diag, chain_nondiag, sqrt_lambda, eigenvalue, eigenvector, eigenvector_size) &
          bind(C, name="nrg_gpu_diagonalize")
       import :: c_int, c_double, c_ptr, c_int64_t
       integer(c_int), value :: iteration, nsub, nfull, nold
       integer(c_int), intent(in) :: subspace_info(*), basis_input(*)
       real(c_double), intent(in) :: eigenvalue_old(*), chain_diag(*), chain_nondiag(*)
       real(c_double), value :: sqrt_lambda
       real(c_double), intent(out) :: eigenvalue(*)
       type(c_ptr), value :: eigenvector
       integer(c_int64_t), value :: eigenvector_size
     end function
     integer(c_int) function nrg_gpu_invariant(kept, out) bind(C, name="nrg_gpu_invariant")
       import :: c_int, c_ptr
       integer(c_int), intent(in) :: kept(*)
       type(c_ptr), value :: out
     end function
     integer(c_int) function nrg_gpu_invariant_spectrum(kept, out) bind(C, name="nrg_gpu_invariant_spectrum")
       import :: c_int, c_ptr
       integer(c_int), intent(in) :: kept(*)
       type(c_ptr), value :: out
     end function
     integer(c_int) function nrg_gpu_save_restart(path, iteration, nstate, state) &
          bind(C, name="nrg_gpu_save_restart")
       import :: c_int, c_char
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), value :: iteration, nstate
       integer(c_int), intent(in) :: state(*)
     end function
     integer(c_int) function nrg_gpu_load_restart(path, iteration, nstate, state) &
          bind(C, name="nrg_gpu_load_restart")
       import :: c_int, c_char
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), intent(out) :: iteration
       integer(c_int), value :: nstate
       integer(c_int), intent(out) :: state(*)
     end function
     integer(c_int) function nrg_gpu_zero_temperature_weights(gs_block, type, positive, w, cap, excited, n) &
          bind(C, name="nrg_gpu_zero_temperature_weights")
       import :: c_int, c_double
       integer(c_int), value :: gs_block, type, positive, cap
       real(c_double), intent(out) :: w(*)
       integer(c_int), intent(out) :: excited, n
     end function
     integer(c_int) function nrg_gpu_set_spectrum_operators(nmat, pos_op, neg_op) &
          bind(C, name="nrg_gpu_set_spectrum_operators")
       import :: c_int
       integer(c_int), value :: nmat
       integer(c_int), intent(in) :: pos_op(*), neg_op(*)
     end function
     integer(c_int) function nrg_gpu_stock(iteration, kept, last) bind(C, name="nrg_gpu_stock")
       import :: c_int
       integer(c_int), value :: iteration, last
       integer(c_int), intent(in) :: kept(*)
     end function
     integer(c_int) function nrg_gpu_rho_init(last_iteration) bind(C, name="nrg_gpu_rho_init")
       import :: c_int
       integer(c_int), value :: last_iteration
     end function
     integer(c_int) function nrg_gpu_rho_shell(iteration, trace) bind(C, name="nrg_gpu_rho_shell")
       import :: c_int, c_double
       integer(c_int), value :: iteration
       real(c_double), intent(out) :: trace
     end function
     integer(c_int) function nrg_gpu_spectrum_setup(nbin) bind(C, name="nrg_gpu_spectrum_setup")
       import :: c_int
       integer(c_int), value :: nbin
     end function
     integer(c_int) function nrg_gpu_spike_cfs(iteration, scale) bind(C, name="nrg_gpu_spike_cfs")
       import :: c_int, c_double
       integer(c_int), value :: iteration
       real(c_double), value :: scale
     end function
     integer(c_int) function nrg_gpu_spike_zero_temperature(iteration, scale) &
          bind(C, name="nrg_gpu_spike_zero_temperature")
       import :: c_int, c_double
       integer(c_int), value :: iteration
       real(c_double), value :: scale
     end function
     integer(c_int) function nrg_gpu_spectrum_bins(negative, positive, n) bind(C, name="nrg_gpu_spectrum_bins")
       import :: c_int, c_double, c_int64_t
       real(c_double), intent(out) :: negative(*), positive(*)
       integer(c_int64_t), value :: n
     end function
     integer(c_int) function nrg_gpu_history_clear() bind(C, name="nrg_gpu_history_clear")
       import :: c_int
     end function
     integer(c_int) function nrg_gpu_stage_times(t, n) bind(C, name="nrg_gpu_stage_times")
       import :: c_int, c_double
       real(c_double), intent(out) :: t(*)
       integer(c_int), value :: n
     end function
     ! options with no counterpart in the reference
     integer(c_int) function nrg_gpu_set_gemm_split(slices) bind(C, name="nrg_gpu_set_gemm_split")
       import :: c_int
       integer(c_int), value :: slices
     end function
     integer(c_int) function nrg_gpu_set_gemm_split_families(invariant_slices, spectral_slices, min_batch_flops) &
         bind(C, name="nrg_gpu_set_gemm_split_families")
       import :: c_int, c_double
       integer(c_int), value :: invariant_slices, spectral_slices
       real(c_double), value :: min_batch_flops
     end function
     integer(c_int) function nrg_gpu_set_batch_limit(max_blocks) bind(C, name="nrg_gpu_set_batch_limit")
       import :: c_int
       integer(c_int), value :: max_blocks
     end function
     integer(c_int) function nrg_gpu_set_tridiag_schedule(align_end) bind(C, name="nrg_gpu_set_tridiag_schedule")
       import :: c_int
       integer(c_int), value :: align_end
     end function
     integer(c_int) function nrg_gpu_set_tridiag_method(two_stage) bind(C, name="nrg_gpu_set_tridiag_method")
       import :: c_int
       integer(c_int), value :: two_stage
     end function
     integer(c_int) function nrg_gpu_set_tridiag_groups(groups) bind(C, name="nrg_gpu_set_tridiag_groups")
       import :: c_int
       integer(c_int), value :: groups
     end function
     integer(c_int) function nrg_gpu_set_tridiag_finish(rows) bind(C, name="nrg_gpu_set_tridiag_finish")
       import :: c_int
       integer(c_int), value :: rows
     end function
     integer(c_int) function nrg_gpu_set_tridiag_symv(kind) bind(C, name="nrg_gpu_set_tridiag_symv")
       import :: c_int
       integer(c_int), value :: kind
     end function
     integer(c_int) function nrg_gpu_set_graphs(on) bind(C, name="nrg_gpu_set_graphs")
       import :: c_int
       integer(c_int), value :: on
     end function
     integer(c_int) function nrg_gpu_set_tuning(mode, cache_path) bind(C, name="nrg_gpu_set_tuning")
       import :: c_int, c_char
       integer(c_int), value :: mode
       character(kind=c_char), intent(in) :: cache_path(*)
     end function
     integer(c_int) function nrg_gpu_profile(enable) bind(C, name="nrg_gpu_profile")
       import :: c_int
       integer(c_int), value :: enable
     end function
     integer(c_int) function nrg_gpu_profile_sample(every) bind(C, name="nrg_gpu_profile_sample")
       import :: c_int
       integer(c_int), value :: every
     end function
     integer(c_int) function nrg_gpu_profile_exclusive(on) bind(C, name="nrg_gpu_profile_exclusive")
       import :: c_int
       integer(c_int), value :: on
     end function
  end interface

contains

  ! Message of the last library error as a Fortran string.
  function gpu_error_text() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: cs(:)
    integer :: n
    p = nrg_gpu_last_error()
    msg = ""
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [4096])
    n = 0
    do while (n < 4096)
       if (cs(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(cs(1:n), msg)
  end function gpu_error_text

end module nrg_gpu_api
