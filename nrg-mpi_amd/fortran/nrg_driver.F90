! bind(C) entry points of libnrg_host.so: the driver surface used by the
! Python tests and the benchmark (and by any C caller).  They wrap the
! Fortran host's NRG loop; the program nrg_amd uses the same routines.
module nrg_driver
  use iso_c_binding
  use nrg_gpu_api
  use nrg_host
  implicit none
contains

  function c_string(p, n) result(s)
    type(c_ptr), value :: p
    integer(c_int), value :: n
    character(len=n) :: s
    character(kind=c_char), pointer :: cs(:)
    integer :: i
    call c_f_pointer(p, cs, [n])
    do i = 1, n
       s(i:i) = cs(i)
    end do
  end function c_string

  ! Read the three input files (no device needed).
  integer(c_int) function nrg_host_read(dir, n) bind(C, name="nrg_host_read")
    type(c_ptr), value :: dir
    integer(c_int), value :: n
    integer :: ierr
    call read_inputs(c_string(dir, n), ierr)
    nrg_host_read = ierr
  end function nrg_host_read

  ! Read the inputs, initialise the device and upload the model.
  !   sort_mode 0: exact replay of the reference's exchange sort, 1: canonical
  integer(c_int) function nrg_host_open(dir, n, device, smode, keepvec) bind(C, name="nrg_host_open")
    type(c_ptr), value :: dir
    integer(c_int), value :: n, device, smode, keepvec
    integer :: ierr
    sort_mode = smode
    keep_vectors = (keepvec /= 0)
    call setup(c_string(dir, n), int(device), ierr)
    nrg_host_open = ierr
  end function nrg_host_open

  ! One NRG iteration (nrgIteration.F90:1-31).
  integer(c_int) function nrg_host_iterate(it) bind(C, name="nrg_host_iterate")
    integer(c_int), value :: it
    integer :: ierr
    call nrg_iteration(int(it), ierr)
    nrg_host_iterate = ierr
  end function nrg_host_iterate

  ! prepareBasis alone on the current basis_output (CPU only, for tests).
  integer(c_int) function nrg_host_prepare(smode) bind(C, name="nrg_host_prepare")
    integer(c_int), value :: smode
    integer :: ierr
    sort_mode = smode
    call prepare_basis(ierr)
    nrg_host_prepare = ierr
  end function nrg_host_prepare

  ! Replace the current basis_output / eigenvalues (CPU tests of the host
  ! bookkeeping): basis(nv+3, n), e(n), nold = n.
  integer(c_int) function nrg_host_set_basis_output(n, basis, e) bind(C, name="nrg_host_set_basis_output")
    integer(c_int), value :: n
    integer(c_int), intent(in) :: basis(*)
    real(c_double), intent(in) :: e(*)
    integer :: i
    if (allocated(basis_out)) deallocate(basis_out)
    if (allocated(eigval)) deallocate(eigval)
    allocate(basis_out(nv + 3, n), eigval(n))
    bo_partial = .false.
    ev_partial = .false.
    do i = 1, n
       basis_out(:, i) = basis((i - 1) * (nv + 3) + 1:i * (nv + 3))
       eigval(i) = e(i)
    end do
    nold = n
    nrg_host_set_basis_output = 0
  end function nrg_host_set_basis_output

  ! Truncation and reordering alone (CPU tests): eigenvalues e(nfull) in
  ! block order of the current subspace table; returns the kept count.
  integer(c_int) function nrg_host_truncate(it, n, e) bind(C, name="nrg_host_truncate")
    integer(c_int), value :: it, n
    real(c_double), intent(in) :: e(*)
    integer :: b, j, s0
    if (allocated(eigval)) deallocate(eigval)
    allocate(eigval(n))
    ev_partial = .false.
    eigval = e(1:n)
    nfull = n
    if (allocated(basis_out)) deallocate(basis_out)
    allocate(basis_out(nv + 3, nfull))
    do b = 1, nsub
       s0 = subspace(nv + 3, b)
       do j = 1, subspace(nv + 1, b)
          basis_out(1:nv, s0 + j - 1) = subspace(1:nv, b)
          basis_out(nv + 1, s0 + j - 1) = j
          basis_out(nv + 2:nv + 3, s0 + j - 1) = 0
       end do
    end do
    call calc_thermo()
    call truncate(int(it))
    call kept_per_block()
    nrg_host_truncate = nkept
  end function nrg_host_truncate

  ! The eigenvalue sort of the truncation on caller values (CPU tests of the
  ! tie order): perm(n) receives the 1-based source position of each sorted
  ! element, val is sorted in place.
  integer(c_int) function nrg_host_sort_values(n, val, perm) bind(C, name="nrg_host_sort_values")
    integer(c_int), value :: n
    real(c_double), intent(inout) :: val(*)
    integer(c_int), intent(out) :: perm(*)
    double precision, allocatable :: v(:)
    integer, allocatable :: p(:)
    integer :: i
    allocate(v(n), p(n))
    v = val(1:n)
    do i = 1, n
       p(i) = i
    end do
    call run_merge_sort_r8(int(n), v, p)
    val(1:n) = v
    perm(1:n) = p
    nrg_host_sort_values = 0
  end function nrg_host_sort_values

  ! The basis sort on caller keys (CPU tests of the exact tie order):
  ! mode 0 the O(N log N) replay, 1 the swap-by-swap exchange sort; perm(n)
  ! receives the 1-based source position of each sorted element.
  integer(c_int) function nrg_host_sort_keys(n, key, perm, mode) bind(C, name="nrg_host_sort_keys")
    integer(c_int), value :: n, mode
    integer(c_int64_t), intent(inout) :: key(*)
    integer(c_int), intent(out) :: perm(*)
    integer(8), allocatable :: k(:)
    integer, allocatable :: p(:)
    integer :: i
    allocate(k(n), p(n))
    k = key(1:n)
    do i = 1, n
       p(i) = i
    end do
    if (mode == 0) then
       call exchange_sort_lazy(int(n), k, p)
    else
       call exchange_sort(int(n), k, p)
    end if
    key(1:n) = k
    perm(1:n) = p
    nrg_host_sort_keys = 0
  end function nrg_host_sort_keys

  ! info(1:16): iteration, nsub, nfull, nkept, start_trunc, nold, i_first,
  ! n_iter, flag_spectrum, nv, nops, ncm, nmat, nkeep, flag_trunc, nchain
  integer(c_int) function nrg_host_info(info) bind(C, name="nrg_host_info")
    integer(c_int), intent(out) :: info(16)
    info = 0
    info(1) = cur_iteration; info(2) = nsub; info(3) = nfull; info(4) = nkept
    info(5) = start_trunc; info(6) = nold; info(7) = i_first; info(8) = n_iter
    info(9) = merge(1, 0, flag_spectrum); info(10) = nv; info(11) = nops; info(12) = ncm
    info(13) = nmat; info(14) = nkeep; info(15) = merge(1, 0, flag_trunc); info(16) = nchain
    nrg_host_info = 0
  end function nrg_host_info

  integer(c_int) function nrg_host_get_eigenvalues(out, n) bind(C, name="nrg_host_get_eigenvalues")
    real(c_double), intent(out) :: out(*)
    integer(c_int), value :: n
    integer :: m
    call eigval_complete()
    m = min(int(n), size(eigval))
    out(1:m) = eigval(1:m)
    nrg_host_get_eigenvalues = m
  end function nrg_host_get_eigenvalues

  integer(c_int) function nrg_host_get_ints(which, out, n) bind(C, name="nrg_host_get_ints")
    ! which: 1 subspace, 2 basis_input, 3 basis_output, 4 kept per block
    integer(c_int), value :: which, n
    integer(c_int), intent(out) :: out(*)
    integer :: m
    m = 0
    select case (which)
    case (1)
       m = min(int(n), size(subspace))
       out(1:m) = reshape(subspace, [size(subspace)])
    case (2)
       m = min(int(n), size(basis_in))
       out(1:m) = reshape(basis_in, [size(basis_in)])
    case (3)
       call basis_out_complete()
       m = min(int(n), size(basis_out))
       out(1:m) = reshape(basis_out, [size(basis_out)])
    case (4)
       m = min(int(n), size(kept_blk))
       out(1:m) = kept_blk(1:m)
    end select
    nrg_host_get_ints = m
  end function nrg_host_get_ints

  integer(c_int) function nrg_host_get_eigenvectors(out, n) bind(C, name="nrg_host_get_eigenvectors")
    real(c_double), intent(out) :: out(*)
    integer(c_int64_t), value :: n
    integer(8) :: m
    m = 0
    if (allocated(eigvec)) then
       m = min(n, int(size(eigvec, kind=8), c_int64_t))
       out(1:m) = eigvec(1:m)
    end if
    nrg_host_get_eigenvectors = int(min(m, 2147483647_8), c_int)
  end function nrg_host_get_eigenvectors

  ! t(1:8): prepare, diagonalization, postprocess, invariant, iteration (s),
  ! thermo X, S, C of the last iteration; t(9) truncation margin
  integer(c_int) function nrg_host_get_times(t) bind(C, name="nrg_host_get_times")
    real(c_double), intent(out) :: t(9)
    t(1) = t_prepare; t(2) = t_diag; t(3) = t_post; t(4) = t_invariant; t(5) = t_iter
    t(6) = thermo_x; t(7) = thermo_s; t(8) = thermo_c; t(9) = trunc_margin
    nrg_host_get_times = 0
  end function nrg_host_get_times

  ! CFS-NRG spectrum phase after the last iteration (main.F90:185-224):
  ! reduced density sweep, CFS spikes, bins copied to the host.
  integer(c_int) function nrg_host_spectrum(keep_history) bind(C, name="nrg_host_spectrum")
    integer(c_int), value :: keep_history
    integer :: ierr
    call spectrum_phase(ierr, keep_history /= 0)
    nrg_host_spectrum = ierr
  end function nrg_host_spectrum

  ! Bin weights of binnedSpikeNegative/Positive (nbin*nmat each); returns
  ! numberOfBin (0 before the spectrum phase).
  integer(c_int) function nrg_host_get_bins(negative, positive, n) bind(C, name="nrg_host_get_bins")
    integer(c_int), value :: n
    real(c_double), intent(out) :: negative(*), positive(*)
    integer :: m
    nrg_host_get_bins = 0
    if (.not. allocated(binned_pos)) return
    m = min(int(n), size(binned_pos))
    negative(1:m) = binned_neg(1:m)
    positive(1:m) = binned_pos(1:m)
    nrg_host_get_bins = number_of_bin
  end function nrg_host_get_bins

  ! Tr rho of the shells i_first..n_iter (zero outside the swept range).
  integer(c_int) function nrg_host_get_rho_traces(out, n) bind(C, name="nrg_host_get_rho_traces")
    integer(c_int), value :: n
    real(c_double), intent(out) :: out(*)
    integer :: m
    nrg_host_get_rho_traces = 0
    if (.not. allocated(rho_trace)) return
    m = min(int(n), size(rho_trace))
    out(1:m) = rho_trace(i_first:i_first + m - 1)
    nrg_host_get_rho_traces = m
  end function nrg_host_get_rho_traces

  ! t(1:3): stock of the last iteration, rho sweep, CFS spikes (s)
  integer(c_int) function nrg_host_get_spectrum_times(t) bind(C, name="nrg_host_get_spectrum_times")
    real(c_double), intent(out) :: t(3)
    t(1) = t_stock; t(2) = t_rho; t(3) = t_cfs
    nrg_host_get_spectrum_times = 0
  end function nrg_host_get_spectrum_times

  ! Write the reference's output files into dir (n > 0), or stop writing
  ! them (n = 0): Xn, Sn, Cn, eigenodd.txt, eigeneven.txt, keepingState.txt
  ! and error.log per iteration, the CFS-spectrum-* files after the
  ! spectrum phase.
  integer(c_int) function nrg_host_set_output(dir, n) bind(C, name="nrg_host_set_output")
    type(c_ptr), value :: dir
    integer(c_int), value :: n
    integer :: ierr
    ierr = 0
    if (n > 0) then
       call open_outputs(c_string(dir, n), ierr)
    else
       call close_outputs()
    end if
    nrg_host_set_output = ierr
  end function nrg_host_set_output

  ! Restart of the CFS phase: save (before every CFS step) and load
  ! (resume at the saved step, skipping the forward pass and the reduced-
  ! density sweep), overriding saveRestart / loadRestart of parameter.txt;
  ! a file name of length n > 0 replaces restart.dat beside the inputs;
  ! stop_at >= 0 ends the spectrum phase right after the save at that
  ! iteration (an interrupted run).
  integer(c_int) function nrg_host_set_restart(save, load, stop_at, path, n) bind(C, name="nrg_host_set_restart")
    integer(c_int), value :: save, load, stop_at, n
    type(c_ptr), value :: path
    save_restart = (save /= 0)
    load_restart = (load /= 0)
    restart_stop = huge(restart_stop)
    if (stop_at >= 0) restart_stop = stop_at
    if (n > 0) restart_path = c_string(path, n)
    nrg_host_set_restart = 0
  end function nrg_host_set_restart

  integer(c_int) function nrg_host_close() bind(C, name="nrg_host_close")
    call close_outputs()
    nrg_host_close = nrg_gpu_finalize()
  end function nrg_host_close

end module nrg_driver
