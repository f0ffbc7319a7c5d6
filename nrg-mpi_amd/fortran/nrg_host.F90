! Fortran host of the MI355X NRG hot path.
!
! Keeps the reference's call surface and bookkeeping on the host:
!   nrgIteration(it) = prepareBasis -> diagonalization -> postprocess
! (nrgIteration.F90:18-28) with the numerical bodies of diagonalization,
! invariantMatrixDGEMM and invariantMatrixForSpectrumDGEMM on the GPU
! through the C ABI of libnrg_gpu.so.  MPI and loadbalance.F90 have no
! counterpart: one device owns every block.
!
! Data layout mirrors the reference's packed records (module.F90:122-604):
!   subspace(nv+3, nsub)  labels, dimension, count_eigenvector, start_input
!   basis_in(nv+3, nfull) labels, reference, variation, operation
!   basis_out(nv+3, nfull)
! so that the arrays cross the C ABI without conversion.
module nrg_host
  use iso_c_binding
  use nrg_gpu_api
  implicit none

  ! thresholds (module.F90:5-7, 629-630, 652)
  double precision, parameter :: numerical_threshold = 1.0d-12
  double precision, parameter :: degeneracy_threshold = 1.0d-8
  double precision, parameter :: truncation_threshold = 1.0d-6
  double precision, parameter :: truncation_limit_scale = 1.2d0
  double precision, parameter :: betabar = real(0.8, kind(1.0)) * 1.0d0

  ! ---- model (namelists + initial.dat)
  integer :: n_iter, i_first, nchain, nv, nmat, ncm, nb0, ncoef, nops
  integer :: cons_block(2, 2)
  logical :: flag_spectrum = .false.
  double precision :: band_width, ztrick, lambda
  integer :: scaling_type, nkeep
  integer, allocatable :: operation(:, :)
  integer, allocatable :: coef_inv_type(:, :), coef_diag_type(:, :)
  double precision, allocatable :: coef_inv(:), coef_diag(:)
  double precision, allocatable :: chain_diag(:, :), chain_nondiag(:, :)
  double precision, allocatable, target :: coef_spec(:, :), inv0(:, :, :), spec0(:, :, :)
  integer, allocatable, target :: cons_diff_spec(:, :), pos_op(:, :), neg_op(:, :)
  ! superconducting (BCS) chain: coefficient_BCS(nbcs,3) and the raw
  ! chain_BCS(0:n_iter+1, nchain) (read_initial_basis.F90:185-201)
  logical :: flag_bcs = .false.
  integer :: nbcs = 0
  integer, allocatable, target :: coef_bcs(:, :)
  double precision, allocatable, target :: chain_bcs(:, :)

  ! ---- iteration state
  integer :: nold = 0, nsub = 0, nfull = 0, nkept = 0
  integer :: start_trunc, cur_iteration
  logical :: flag_trunc = .false.
  integer, allocatable, target :: basis_out(:, :), basis_in(:, :), subspace(:, :), kept_blk(:)
  double precision, allocatable, target :: eigval(:), eigval_old(:)
  double precision :: trunc_margin = -1.0d0
  double precision :: thermo_x, thermo_s, thermo_c
  ! After a truncation eigval and basis_out hold the kept prefix; the
  ! discarded states follow in ascending order of energy, formed when a
  ! reader asks for the whole arrays (eigval_complete, basis_out_complete)
  ! from the eigenvalues before the truncation (ev_pre), the kept count
  ! (tr_nb) and the block table they index (bo_sub).
  double precision, allocatable :: ev_pre(:)
  integer, allocatable :: bo_sub(:, :)
  integer :: tr_nb = 0
  logical :: bo_partial = .false., ev_partial = .false.
  ! k-way merge of ascending runs (kw_*): run starts, current heads, heap
  integer, allocatable :: kw_rs(:), kw_hd(:), kw_hp(:)
  integer :: kw_nr = 0, kw_nh = 0

  ! ---- options
  integer :: sort_mode = 0          ! 0: exact replay of the reference sort, 1: canonical
  logical :: keep_vectors = .false. ! copy eigenvectors to the host (spectra / tests)
  double precision, allocatable, target :: eigvec(:)

  ! ---- CFS spectra (main.F90:141-156): bins of types 1..nmat, weight column
  integer, parameter :: mesh_of_bin = 1000
  double precision, parameter :: log10_emin = -15.0d0
  integer :: number_of_bin = 0
  double precision, allocatable, target :: binned_pos(:), binned_neg(:)
  double precision, allocatable :: rho_trace(:)   ! Tr rho per shell, indexed by iteration

  ! ---- the reference's output files, written when write_files is on
  ! (open_outputs): Xn/Sn/Cn (postprocess.F90:60-67), eigenodd.txt /
  ! eigeneven.txt (:128-137), keepingState.txt (:163-167), error.log, and
  ! after the spectrum phase CFS-spectrum-{positive,negative}-typeNNN
  ! (main.F90:228-252) and the zero-temperature spikes step_ 02_itr_NNN.*
  ! (calculateSpectrumDGEMM.F90:80-90,157-159,203-205).  The records are
  ! written by the same list-directed statements as the reference's.
  logical :: write_files = .false.
  ! ---- restart of the CFS phase (main.F90:165-221): loadRestart /
  ! saveRestart of parameter.txt; the file sits beside the inputs (the
  ! reference writes restart.dat in its working directory); restart_stop
  ! ends the spectrum phase right after the save at that iteration (an
  ! interrupted run, for tests)
  logical :: load_restart = .false., save_restart = .false.
  character(len=:), allocatable :: restart_path
  integer :: restart_stop
  ! open units (newunit= numbers are negative; 0 marks a closed file)
  integer :: u_xn = 0, u_sn = 0, u_cn = 0, u_odd = 0, u_even = 0, u_info = 0, u_err = 0
  character(len=:), allocatable :: out_dir

  ! ---- timing (seconds)
  double precision :: t_prepare = 0, t_diag = 0, t_post = 0, t_invariant = 0, t_iter = 0
  double precision :: t_stock = 0, t_rho = 0, t_cfs = 0

contains

  ! =====================================================================
  ! input (module.F90:656-688, 826-944; read_initial_basis.F90:146-302)
  ! =====================================================================
  ! namelist /hamiltonianInfo/ (module.F90:846-861); the locals carry the
  ! namelist names and hide the module variables of the same name
  subroutine read_hamiltonian_nml(path, ierr, o_iter, o_nchain, o_nv, o_cons, o_nmat, o_ncm, o_nb, &
       o_ncoef, o_nops, o_i0, o_spectrum, o_bcs, o_nbcs)
    character(len=*), intent(in) :: path
    integer, intent(out) :: ierr, o_iter, o_nchain, o_nv, o_cons(2, 2), o_nmat, o_ncm, o_nb, o_ncoef, o_nops, o_i0
    integer, intent(out) :: o_nbcs
    logical, intent(out) :: o_spectrum, o_bcs
    integer :: numberOfIteration, numberOfChain, numberOfVariation, conservationBlock(2, 2)
    integer :: numberOfMatrix, numberOfConductionMatrix, numberOfBasis, numberOfCoefficient
    integer :: numberOfCoefficientImpurity, numberOfOperation, numberOfBCSCoefficient, indexOfFirstIteration
    logical :: flag_spectrum, flag_local, flag_BCS
    integer :: u
    namelist /hamiltonianInfo/ numberOfIteration, numberOfChain, numberOfVariation, conservationBlock, &
         numberOfMatrix, numberOfConductionMatrix, numberOfBasis, numberOfCoefficient, numberOfBCSCoefficient, &
         numberOfCoefficientImpurity, numberOfOperation, indexOfFirstIteration, flag_spectrum, flag_local, flag_BCS
    numberOfCoefficientImpurity = 0
    numberOfBCSCoefficient = 0
    flag_BCS = .false.
    flag_local = .false.
    flag_spectrum = .false.
    open(newunit=u, file=path, status="old", action="read", iostat=ierr)
    if (ierr /= 0) return
    read(u, nml=hamiltonianInfo, iostat=ierr)
    close(u)
    if (ierr /= 0) return
    o_iter = numberOfIteration; o_nchain = numberOfChain; o_nv = numberOfVariation; o_cons = conservationBlock
    o_nmat = numberOfMatrix; o_ncm = numberOfConductionMatrix; o_nb = numberOfBasis; o_ncoef = numberOfCoefficient
    o_nops = numberOfOperation; o_i0 = indexOfFirstIteration; o_spectrum = flag_spectrum; o_bcs = flag_BCS
    o_nbcs = numberOfBCSCoefficient
  end subroutine read_hamiltonian_nml

  ! namelist /nrg/ (module.F90:665-675)
  subroutine read_parameter_nml(path, ierr, o_band, o_z, o_lambda, o_scaling, o_nkeep, o_load, o_save)
    character(len=*), intent(in) :: path
    integer, intent(out) :: ierr, o_scaling, o_nkeep
    double precision, intent(out) :: o_band, o_z, o_lambda
    logical, intent(out) :: o_load, o_save
    double precision :: bandWidth, ztrick, Lambda
    integer :: scalingType, truncation, u
    logical :: cfsnrg, noimpurity, loadRestart, saveRestart
    namelist /nrg/ bandWidth, ztrick, Lambda, scalingType, truncation, cfsnrg, noimpurity, loadRestart, saveRestart
    cfsnrg = .false.; noimpurity = .false.; loadRestart = .false.; saveRestart = .false.
    open(newunit=u, file=path, status="old", action="read", iostat=ierr)
    if (ierr /= 0) return
    read(u, nml=nrg, iostat=ierr)
    close(u)
    if (ierr /= 0) return
    o_band = bandWidth; o_z = ztrick; o_lambda = Lambda; o_scaling = scalingType; o_nkeep = truncation
    o_load = loadRestart; o_save = saveRestart
  end subroutine read_parameter_nml

  ! =====================================================================
  ! input (module.F90:656-688, 826-944; read_initial_basis.F90:146-302)
  ! =====================================================================
  subroutine read_inputs(dir, ierr)
    character(len=*), intent(in) :: dir
    integer, intent(out) :: ierr
    integer :: u, i, j, k, c
    character(len=64) :: hdr
    double precision :: f
    logical :: bcs

    call read_hamiltonian_nml(dir // "/hamiltonianInfo.dat", ierr, n_iter, nchain, nv, cons_block, nmat, ncm, &
         nb0, ncoef, nops, i_first, flag_spectrum, bcs, nbcs)
    if (ierr /= 0) return
    call read_parameter_nml(dir // "/parameter.txt", ierr, band_width, ztrick, lambda, scaling_type, nkeep, &
         load_restart, save_restart)
    restart_path = dir // "/restart.dat"
    restart_stop = huge(restart_stop)
    if (ierr /= 0) return
    flag_bcs = bcs
    if (.not. flag_bcs) nbcs = 0

    if (allocated(operation)) deallocate(operation, coef_inv_type, coef_diag_type, coef_inv, coef_diag, &
         chain_diag, chain_nondiag, inv0)
    if (allocated(basis_out)) deallocate(basis_out)
    if (allocated(eigval)) deallocate(eigval)
    allocate(operation(nops, nv + nchain), coef_inv_type(ncoef, 3), coef_diag_type(ncoef, 4))
    allocate(coef_inv(ncoef), coef_diag(ncoef))
    allocate(chain_diag(0:n_iter + 1, nchain), chain_nondiag(i_first:n_iter + 1, nchain))
    allocate(inv0(nb0, nb0, ncm), basis_out(nv + 3, nb0), eigval(nb0))

    open(newunit=u, file=dir // "/initial.dat", status="old", action="read", iostat=ierr)
    if (ierr /= 0) return
    read(u, *) hdr
    do i = 1, nb0
       read(u, *) basis_out(1:nv, i), basis_out(nv + 1, i), eigval(i)
       basis_out(nv + 2:nv + 3, i) = 0
    end do
    bo_partial = .false.
    ev_partial = .false.
    read(u, *) hdr
    do c = 1, ncm
       do j = 1, nb0
          do k = 1, nb0
             read(u, *) inv0(k, j, c)
          end do
       end do
    end do
    read(u, *) hdr
    do i = 1, ncoef
       read(u, *) coef_inv_type(i, 1:3), coef_inv(i)
    end do
    read(u, *) hdr
    do i = 1, ncoef
       read(u, *) coef_diag_type(i, 1:4), coef_diag(i)
    end do
    read(u, *) hdr
    do i = 1, nops
       read(u, *) operation(i, 1:nv + nchain)
    end do
    read(u, *) hdr
    do c = 1, nchain
       read(u, *) hdr
       do j = 0, n_iter + 1
          read(u, *) chain_diag(j, c)
       end do
    end do
    do c = 1, nchain
       read(u, *) hdr
       do j = i_first, n_iter + 1
          read(u, *) chain_nondiag(j, c)
       end do
    end do
    ! rescaling of indices >= 0 (read_initial_basis.F90:222-261)
    do c = 1, nchain
       do j = 0, n_iter + 1
          select case (scaling_type)
          case (0)
             chain_diag(j, c) = chain_diag(j, c) * lambda**(dble(j - 1) * 0.5d0)
             chain_nondiag(j, c) = chain_nondiag(j, c) * lambda**(dble(j) * 0.5d0)
          case (2)
             f = log(lambda) / band_width / (lambda - 1.0d0)
             chain_diag(j, c) = chain_diag(j, c) * lambda**(dble(j - 1) * 0.5d0 + ztrick) * f
             chain_nondiag(j, c) = chain_nondiag(j, c) * lambda**(dble(j) * 0.5d0 + ztrick) * f
          case default
             f = 1.0d0 / band_width * 2.0d0 / (1.0d0 + 1.0d0 / lambda)
             chain_diag(j, c) = chain_diag(j, c) * lambda**(dble(j - 1) * 0.5d0 + ztrick - 1.0d0) * f
             chain_nondiag(j, c) = chain_nondiag(j, c) * lambda**(dble(j) * 0.5d0 + ztrick - 1.0d0) * f
          end select
       end do
    end do
    if (flag_spectrum) then
       if (allocated(coef_spec)) deallocate(coef_spec, spec0, cons_diff_spec, pos_op, neg_op)
       allocate(coef_spec(nmat, nops), spec0(nb0, nb0, nmat), cons_diff_spec(nmat, nv))
       allocate(pos_op(nmat, nv), neg_op(nmat, nv))
       read(u, *) hdr
       do i = 1, nmat
          do j = 1, nops
             read(u, *) coef_spec(i, j)
          end do
       end do
       read(u, *) hdr
       do i = 1, nmat
          do j = 1, nb0
             do k = 1, nb0
                read(u, *) spec0(k, j, i)
             end do
          end do
       end do
       read(u, *) hdr
       do i = 1, nmat
          read(u, *) cons_diff_spec(i, 1:nv)
       end do
       read(u, *) hdr
       do i = 1, nmat
          read(u, *) pos_op(i, 1:nv)
       end do
       read(u, *) hdr
       do i = 1, nmat
          read(u, *) neg_op(i, 1:nv)
       end do
    end if
    if (flag_bcs) then
       ! chain_BCS per chain, not rescaled, then the pair rows
       ! (read_initial_basis.F90:185-201)
       if (allocated(coef_bcs)) deallocate(coef_bcs, chain_bcs)
       allocate(chain_bcs(0:n_iter + 1, nchain), coef_bcs(max(nbcs, 1), 3))
       coef_bcs = 0
       do c = 1, nchain
          read(u, *) hdr
          do j = 0, n_iter + 1
             read(u, *) chain_bcs(j, c)
          end do
       end do
       read(u, *) hdr
       do i = 1, nbcs
          read(u, *) coef_bcs(i, 1:3)
       end do
    end if
    close(u)
    nold = nb0
    start_trunc = n_iter + 1
    flag_trunc = .false.
  end subroutine read_inputs

  double precision function energy_scale(n)
    ! module.F90:690-711
    integer, intent(in) :: n
    select case (scaling_type)
    case (0)
       energy_scale = lambda**(-dble(n) * 0.5d0)
    case (2)
       energy_scale = lambda**(-dble(n) * 0.5d0 - ztrick) * band_width * (lambda - 1.0d0) / log(lambda)
    case default
       energy_scale = band_width * 0.5d0 * (1.0d0 + 1.0d0 / lambda) * lambda**(-dble(n) * 0.5d0 - ztrick + 1.0d0)
    end select
  end function energy_scale

  ! =====================================================================
  ! sorts with the reference's permutations
  ! =====================================================================
  ! Replay of the exchange sort of module.F90:537-543: for n, for m > n, if
  ! key(n) > key(m) swap.  Ties (equal labels and reference) keep the order
  ! that the swap history gives them, which this loop reproduces exactly.
  subroutine exchange_sort(n, key, perm)
    integer, intent(in) :: n
    integer(8), intent(inout) :: key(n)
    integer, intent(inout) :: perm(n)
    integer :: i, m, pi, pm
    integer(8) :: cur, km
    do i = 1, n - 1
       cur = key(i)
       pi = perm(i)
       do m = i + 1, n
          km = key(m)
          if (km < cur) then
             key(m) = cur
             pm = perm(m)
             perm(m) = pi
             cur = km
             pi = pm
          end if
       end do
       key(i) = cur
       perm(i) = pi
    end do
  end subroutine exchange_sort

  ! The same exchange sort in O(N log N).  Fuse all steps: element m passes
  ! the holders of steps 1..m-1 in order (swapping where it is smaller) and
  ! becomes the holder of step m.  The holders stay sorted, so a passing
  ! element x of key v lands right after the holders of key v, and each swap
  ! hands the first holder of the next key group to the slot of the first
  ! holder of the group after it: every group of keys above v rotates its
  ! first member to its end, and the carried element (first member of the
  ! top group) becomes the new last holder, which rotates the top group too.
  ! Groups are kept per distinct key (rank); the rotations of all groups
  ! above v are one range update on a Fenwick tree over ranks, applied to a
  ! group only when it is next touched.  The final order is the groups in
  ! key order, each in its rotated order: identical to the swap-by-swap
  ! loop above, tie order included.
  subroutine exchange_sort_lazy(n, key, perm)
    integer, intent(in) :: n
    integer(8), intent(inout) :: key(n)
    integer, intent(inout) :: perm(n)
    integer(8), allocatable :: sk(:), wk(:), k2(:)
    integer, allocatable :: idx(:), wp(:), rnk(:), gofs(:), glen(:), gbase(:), fw(:), gel(:), order(:), p2(:)
    integer :: i, r, nr, top, m, rot, L, j, p, q, t
    if (n <= 1) return
    allocate(sk(n), wk(n), idx(n), wp(n), rnk(n))
    sk = key
    do i = 1, n
       idx(i) = i
    end do
    call merge_sort_i8(n, sk, idx, wk, wp)
    nr = 0
    do i = 1, n
       if (i == 1) then
          nr = 1
       else if (sk(i) /= sk(i - 1)) then
          nr = nr + 1
       end if
       rnk(idx(i)) = nr
    end do
    deallocate(sk, wk, idx, wp)
    ! group storage: members of rank r at gel(gofs(r)+1 : gofs(r)+glen(r))
    allocate(gofs(nr + 1), glen(nr), gbase(nr), fw(nr), gel(n), order(n))
    gofs = 0
    do i = 1, n
       gofs(rnk(i) + 1) = gofs(rnk(i) + 1) + 1
    end do
    do r = 2, nr + 1
       gofs(r) = gofs(r) + gofs(r - 1)
    end do
    glen = 0
    gbase = 0
    fw = 0
    top = 0
    do m = 1, n
       r = rnk(m)
       q = fen_query(r)
       L = glen(r)
       if (L > 1) then
          rot = mod(q - gbase(r), L)
          if (rot /= 0) call rotate_left(gel(gofs(r) + 1:gofs(r) + L), L, rot)
       end if
       gbase(r) = q
       glen(r) = L + 1
       gel(gofs(r) + L + 1) = m
       if (r < top) call fen_add(r + 1)
       if (r > top) top = r
    end do
    p = 0
    do r = 1, nr
       L = glen(r)
       rot = mod(fen_query(r) - gbase(r), L)
       do j = 0, L - 1
          p = p + 1
          order(p) = gel(gofs(r) + 1 + mod(j + rot, L))
       end do
    end do
    allocate(k2(n), p2(n))
    do i = 1, n
       k2(i) = key(order(i))
       p2(i) = perm(order(i))
    end do
    key = k2
    perm = p2
  contains
    ! +1 on every rank >= i (difference form)
    subroutine fen_add(i0)
      integer, intent(in) :: i0
      integer :: ii
      ii = i0
      do while (ii <= nr)
         fw(ii) = fw(ii) + 1
         ii = ii + iand(ii, -ii)
      end do
    end subroutine fen_add
    ! rotations applied so far to rank i0
    integer function fen_query(i0)
      integer, intent(in) :: i0
      integer :: ii
      fen_query = 0
      ii = i0
      do while (ii > 0)
         fen_query = fen_query + fw(ii)
         ii = ii - iand(ii, -ii)
      end do
    end function fen_query
    subroutine rotate_left(a, la, k)
      integer, intent(in) :: la, k
      integer, intent(inout) :: a(la)
      integer :: tmp(la)
      tmp = a
      a(1:la - k) = tmp(k + 1:la)
      a(la - k + 1:la) = tmp(1:k)
    end subroutine rotate_left
  end subroutine exchange_sort_lazy

  ! Stable merge sort of (key, perm) by key (the permutation of an insertion
  ! sort with strict comparisons, module.F90:16-44).
  recursive subroutine merge_sort_i8(n, key, perm, wk, wp)
    integer, intent(in) :: n
    integer(8), intent(inout) :: key(n)
    integer, intent(inout) :: perm(n)
    integer(8), intent(inout) :: wk(n)
    integer, intent(inout) :: wp(n)
    integer :: h, i, j, k
    if (n <= 16) then
       call insertion_i8(n, key, perm)
       return
    end if
    h = n / 2
    call merge_sort_i8(h, key(1:h), perm(1:h), wk(1:h), wp(1:h))
    call merge_sort_i8(n - h, key(h + 1:n), perm(h + 1:n), wk(h + 1:n), wp(h + 1:n))
    i = 1; j = h + 1; k = 1
    do while (i <= h .and. j <= n)
       if (key(j) < key(i)) then
          wk(k) = key(j); wp(k) = perm(j); j = j + 1
       else
          wk(k) = key(i); wp(k) = perm(i); i = i + 1
       end if
       k = k + 1
    end do
    do while (i <= h)
       wk(k) = key(i); wp(k) = perm(i); i = i + 1; k = k + 1
    end do
    do while (j <= n)
       wk(k) = key(j); wp(k) = perm(j); j = j + 1; k = k + 1
    end do
    key = wk
    perm = wp
  end subroutine merge_sort_i8

  subroutine insertion_i8(n, key, perm)
    integer, intent(in) :: n
    integer(8), intent(inout) :: key(n)
    integer, intent(inout) :: perm(n)
    integer :: i, j, tp
    integer(8) :: tk
    do i = 2, n
       tk = key(i); tp = perm(i)
       j = i - 1
       do while (j >= 1)
          if (key(j) <= tk) exit
          key(j + 1) = key(j); perm(j + 1) = perm(j)
          j = j - 1
       end do
       key(j + 1) = tk; perm(j + 1) = tp
    end do
  end subroutine insertion_i8

  ! Stable ascending sort of (val, perm) by val that merges the ascending
  ! runs the input already holds (the eigenvalues arrive as one ascending
  ! run per block): ceil(log2 r) passes for r runs.  Ties keep their input
  ! order, so the permutation is that of the reference's insertion sort
  ! with strict comparisons (module.F90:16-44).
  subroutine run_merge_sort_r8(n, val, perm)
    integer, intent(in) :: n
    double precision, intent(inout) :: val(n)
    integer, intent(inout) :: perm(n)
    double precision, allocatable :: vb(:)
    integer, allocatable :: pb(:), rs(:)
    integer :: nr, i
    logical :: inbuf
    if (n < 2) return
    allocate(rs(n + 1))
    nr = 1
    rs(1) = 1
    do i = 2, n
       if (val(i) < val(i - 1)) then
          nr = nr + 1
          rs(nr) = i
       end if
    end do
    rs(nr + 1) = n + 1
    if (nr == 1) return
    allocate(vb(n), pb(n))
    inbuf = .false.
    do while (nr > 1)
       if (inbuf) then
          call merge_run_pairs(n, vb, pb, val, perm, nr, rs)
       else
          call merge_run_pairs(n, val, perm, vb, pb, nr, rs)
       end if
       inbuf = .not. inbuf
    end do
    if (inbuf) then
       val = vb
       perm = pb
    end if
  end subroutine run_merge_sort_r8

  ! One pass: runs 2i-1 and 2i of (sv, sp) merged into (dv, dp), the right
  ! element taken only when strictly smaller; an odd last run is copied.
  ! rs(1:nr+1) holds the run starts and n + 1; on return the merged ones.
  subroutine merge_run_pairs(n, sv, sp, dv, dp, nr, rs)
    integer, intent(in) :: n
    double precision, intent(in) :: sv(n)
    integer, intent(in) :: sp(n)
    double precision, intent(out) :: dv(n)
    integer, intent(out) :: dp(n)
    integer, intent(inout) :: nr, rs(:)
    integer :: r, i, j, k, m, e, nr2
    nr2 = 0
    r = 1
    do while (r <= nr)
       nr2 = nr2 + 1
       i = rs(r)
       rs(nr2) = i
       if (r == nr) then
          dv(i:n) = sv(i:n)
          dp(i:n) = sp(i:n)
          exit
       end if
       m = rs(r + 1)
       e = rs(r + 2)
       j = m
       k = i
       do while (i < m .and. j < e)
          if (sv(j) < sv(i)) then
             dv(k) = sv(j); dp(k) = sp(j); j = j + 1
          else
             dv(k) = sv(i); dp(k) = sp(i); i = i + 1
          end if
          k = k + 1
       end do
       do while (i < m)
          dv(k) = sv(i); dp(k) = sp(i); i = i + 1; k = k + 1
       end do
       do while (j < e)
          dv(k) = sv(j); dp(k) = sp(j); j = j + 1; k = k + 1
       end do
       r = r + 2
    end do
    rs(nr2 + 1) = n + 1
    nr = nr2
  end subroutine merge_run_pairs

  ! =====================================================================
  ! prepareBasis (prepareBasis.F90:41-114)
  ! =====================================================================
  subroutine prepare_basis(ierr)
    integer, intent(out) :: ierr
    integer :: i, j, p, c, b, cnt
    integer(8), allocatable :: key(:), wk(:)
    integer, allocatable :: perm(:), wp(:), tmp(:, :)
    integer(8) :: kk
    integer :: lab
    ierr = 0
    nfull = nold * nops
    if (allocated(basis_in)) deallocate(basis_in)
    allocate(basis_in(nv + 3, nfull), key(nfull), perm(nfull))
    p = 0
    do i = 1, nold
       do j = 1, nops
          p = p + 1
          basis_in(1:nv, p) = basis_out(1:nv, i) - operation(j, 1:nv)
          basis_in(nv + 1, p) = i
          basis_in(nv + 2, p) = basis_out(nv + 1, i)
          basis_in(nv + 3, p) = j
          kk = 0
          do c = 1, nv
             lab = basis_in(c, p)
             if (abs(lab) >= 512) then
                ierr = 3
                return
             end if
             kk = kk * 1024_8 + int(lab + 512, 8)
          end do
          key(p) = kk * 16777216_8 + int(i, 8)
          perm(p) = p
       end do
    end do
    if (nold >= 16777216) then
       ierr = 3
       return
    end if
    if (sort_mode == 0) then
       call exchange_sort_lazy(nfull, key, perm)
    else
       allocate(wk(nfull), wp(nfull))
       call merge_sort_i8(nfull, key, perm, wk, wp)
       deallocate(wk, wp)
    end if
    allocate(tmp(nv + 3, nfull))
    tmp = basis_in(:, perm)
    call move_alloc(tmp, basis_in)
    ! block table
    cnt = 1
    do i = 2, nfull
       if (any(basis_in(1:nv, i) /= basis_in(1:nv, i - 1))) cnt = cnt + 1
    end do
    nsub = cnt
    if (allocated(subspace)) deallocate(subspace)
    allocate(subspace(nv + 3, nsub))
    b = 0
    do i = 1, nfull
       if (i == 1) then
          b = 1
          subspace(1:nv, b) = basis_in(1:nv, i)
          subspace(nv + 1, b) = 1
          subspace(nv + 3, b) = i
       else if (any(basis_in(1:nv, i) /= basis_in(1:nv, i - 1))) then
          b = b + 1
          subspace(1:nv, b) = basis_in(1:nv, i)
          subspace(nv + 1, b) = 1
          subspace(nv + 3, b) = i
       else
          subspace(nv + 1, b) = subspace(nv + 1, b) + 1
       end if
    end do
    ! count_eigenvector: running sum of dimension^2 (int32, as the reference)
    kk = 1
    do b = 1, nsub
       subspace(nv + 2, b) = int(kk)          ! may wrap beyond 2^31 like the reference
       kk = kk + int(subspace(nv + 1, b), 8)**2
    end do
    deallocate(key, perm)
  end subroutine prepare_basis

  ! =====================================================================
  ! diagonalization (diagonalization.F90:1-410) on the GPU
  ! =====================================================================
  subroutine diagonalization(it, ierr)
    integer, intent(in) :: it
    integer, intent(out) :: ierr
    double precision :: cd(nchain), cnd(nchain)
    integer(8) :: n2
    integer :: b
    type(c_ptr) :: pvec
    do b = 1, nchain
       cd(b) = chain_diag(it + 1, b)
       cnd(b) = chain_nondiag(it, b)
    end do
    n2 = 0
    do b = 1, nsub
       n2 = n2 + int(subspace(nv + 1, b), 8)**2
    end do
    if (allocated(eigval_old)) deallocate(eigval_old)
    allocate(eigval_old(nold))
    eigval_old = eigval(1:nold)
    if (allocated(eigval)) deallocate(eigval)
    allocate(eigval(nfull))
    ev_partial = .false.
    pvec = c_null_ptr
    if (keep_vectors) then
       if (allocated(eigvec)) deallocate(eigvec)
       allocate(eigvec(n2))
       pvec = c_loc(eigvec)
    end if
    ierr = nrg_gpu_diagonalize(int(it, c_int), int(nsub, c_int), subspace, int(nfull, c_int), basis_in, &
         int(nold, c_int), eigval_old, cd, cnd, sqrt(lambda), eigval, pvec, int(n2, c_int64_t))
    ! basis_output <- (block labels, column index) (diagonalization.F90:395-407)
    ! is formed in its post-truncation order by truncate: state s0 + j - 1 of
    ! block b is (subspace(1:nv, b), j, 0, 0)
  end subroutine diagonalization

  ! =====================================================================
  ! thermodynamics (module.F90:1931-1982), before truncation
  ! =====================================================================
  subroutine calc_thermo()
    integer :: i, j, b, jj
    double precision :: be, ex, zn, en, sz2, sz, cn, rs
    integer :: szt
    double precision, allocatable :: exv(:)
    allocate(exv(nfull))
    zn = 0; en = 0; sz2 = 0; sz = 0; cn = 0
    ! states in index order: block by block (the blocks are contiguous and
    ! ascending), the labels of state i those of its block
    i = 0
    do b = 1, nsub
       do jj = 1, subspace(nv + 1, b)
          i = i + 1
          be = betabar * eigval(i)
          ex = exp(-be)
          exv(i) = ex
          zn = zn + ex
          en = en + be * ex
          szt = 0
          do j = cons_block(2, 1), cons_block(2, 2)
             szt = szt + state_label(subspace, j, b, jj)
          end do
          rs = dble(szt) * 0.5d0
          sz2 = sz2 + ex * rs * rs
          sz = sz + ex * rs
       end do
    end do
    do i = 1, nfull
       be = betabar * eigval(i)
       cn = cn + (be - en / zn) * (be - en / zn) * exv(i)
    end do
    thermo_c = cn / zn
    thermo_s = en / zn + log(zn)
    thermo_x = sz2 / zn - (sz / zn) * (sz / zn)
  end subroutine calc_thermo

  ! Entry c of basis_output for state jj of block b of the table sub
  ! (labels 1..nv, the column index, then two zeros).
  pure integer function state_label(sub, c, b, jj)
    integer, intent(in) :: sub(:, :), c, b, jj
    if (c <= nv) then
       state_label = sub(c, b)
    else if (c == nv + 1) then
       state_label = jj
    else
       state_label = 0
    end if
  end function state_label

  subroutine set_state(i, sub, b, jj)
    integer, intent(in) :: i, sub(:, :), b, jj
    integer :: c
    do c = 1, nv
       basis_out(c, i) = sub(c, b)
    end do
    basis_out(nv + 1, i) = jj
    basis_out(nv + 2, i) = 0
    basis_out(nv + 3, i) = 0
  end subroutine set_state

  ! Indices (before the truncation) of the discarded states in ascending
  ! order of energy, ties in index order: the stable sort of ev_pre past
  ! its first tr_nb entries.
  subroutine discarded_order(rest)
    integer, allocatable, intent(out) :: rest(:)
    integer, allocatable :: idx(:)
    double precision, allocatable :: es(:)
    integer :: i, n
    n = size(ev_pre)
    allocate(idx(n), es(n))
    do i = 1, n
       idx(i) = i
    end do
    es = ev_pre
    call run_merge_sort_r8(n, es, idx)
    rest = idx(tr_nb + 1:n)
  end subroutine discarded_order

  ! The discarded part of basis_output, for readers of the whole array.
  subroutine basis_out_complete()
    integer, allocatable :: blk(:), rest(:)
    integer :: b, j, i, s
    if (.not. bo_partial) return
    call discarded_order(rest)
    allocate(blk(size(ev_pre)))
    do b = 1, size(bo_sub, 2)
       do j = 1, bo_sub(nv + 1, b)
          blk(bo_sub(nv + 3, b) + j - 1) = b
       end do
    end do
    do i = 1, size(rest)
       s = rest(i)
       b = blk(s)
       call set_state(tr_nb + i, bo_sub, b, s - bo_sub(nv + 3, b) + 1)
    end do
    bo_partial = .false.
  end subroutine basis_out_complete

  ! The discarded part of eigval, for readers of the whole array.
  subroutine eigval_complete()
    integer, allocatable :: rest(:)
    if (.not. ev_partial) return
    call discarded_order(rest)
    eigval(tr_nb + 1:tr_nb + size(rest)) = ev_pre(rest)
    ev_partial = .false.
  end subroutine eigval_complete

  ! k-way merge of the ascending runs of v(1:n) (a new run wherever a value
  ! is below its predecessor; the eigenvalues arrive as one run per block):
  ! a binary heap of run heads ordered by (value, run), so equal values
  ! leave in index order and kw_next yields the stable ascending order,
  ! one element per call, without sorting what is never asked for.
  subroutine kw_init(n, v)
    integer, intent(in) :: n
    double precision, intent(in) :: v(n)
    integer :: i
    if (allocated(kw_rs)) deallocate(kw_rs, kw_hd, kw_hp)
    kw_nr = 1
    do i = 2, n
       if (v(i) < v(i - 1)) kw_nr = kw_nr + 1
    end do
    allocate(kw_rs(kw_nr + 1))
    kw_nr = 1
    kw_rs(1) = 1
    do i = 2, n
       if (v(i) < v(i - 1)) then
          kw_nr = kw_nr + 1
          kw_rs(kw_nr) = i
       end if
    end do
    kw_rs(kw_nr + 1) = n + 1
    allocate(kw_hd(kw_nr), kw_hp(kw_nr))
    kw_nh = 0
    do i = 1, kw_nr
       kw_hd(i) = kw_rs(i)
       if (kw_hd(i) < kw_rs(i + 1)) then
          kw_nh = kw_nh + 1
          kw_hp(kw_nh) = i
       end if
    end do
    do i = kw_nh / 2, 1, -1
       call kw_down(v, i)
    end do
  end subroutine kw_init

  logical function kw_less(v, a, b)
    double precision, intent(in) :: v(:)
    integer, intent(in) :: a, b
    kw_less = v(kw_hd(a)) < v(kw_hd(b)) .or. (v(kw_hd(a)) == v(kw_hd(b)) .and. a < b)
  end function kw_less

  subroutine kw_down(v, i0)
    double precision, intent(in) :: v(:)
    integer, intent(in) :: i0
    integer :: i, c, t
    i = i0
    do
       c = 2 * i
       if (c > kw_nh) exit
       if (c < kw_nh) then
          if (kw_less(v, kw_hp(c + 1), kw_hp(c))) c = c + 1
       end if
       if (.not. kw_less(v, kw_hp(c), kw_hp(i))) exit
       t = kw_hp(i); kw_hp(i) = kw_hp(c); kw_hp(c) = t
       i = c
    end do
  end subroutine kw_down

  ! next element of the merge: its value and index; .false. when exhausted
  logical function kw_next(v, val, pos)
    double precision, intent(in) :: v(:)
    double precision, intent(out) :: val
    integer, intent(out) :: pos
    integer :: r
    kw_next = kw_nh > 0
    if (.not. kw_next) return
    r = kw_hp(1)
    pos = kw_hd(r)
    val = v(pos)
    kw_hd(r) = pos + 1
    if (kw_hd(r) >= kw_rs(r + 1)) then
       kw_hp(1) = kw_hp(kw_nh)
       kw_nh = kw_nh - 1
    end if
    if (kw_nh > 0) call kw_down(v, 1)
  end function kw_next

  ! =====================================================================
  ! postprocess (postprocess.F90:71-207): sort, truncate, reorder
  ! =====================================================================
  subroutine truncate(it)
    integer, intent(in) :: it
    integer, allocatable :: idx(:), wp(:), kidx(:)
    double precision, allocatable :: es(:)
    integer(8), allocatable :: kk(:), wk(:)
    integer :: i, limit, cutoff, nb, b, ne, pos
    double precision :: ecut, val
    logical :: more
    ! the stable ascending order of the eigenvalues (postprocess.F90:71-77),
    ! es(1:ne) with their indices idx(1:ne): as far as the cutoff needs it
    ! when truncating (k-way merge of the per-block runs), whole when every
    ! state is kept and the kept energies are written, not at all otherwise
    allocate(idx(nfull), es(nfull))
    ne = 0
    nb = nfull
    trunc_margin = -1.0d0
    if (it /= i_first .and. it /= n_iter .and. nfull > nkeep) then
       if (start_trunc > it) then
          start_trunc = it
          flag_trunc = .true.
       end if
       call kw_init(nfull, eigval)
       do while (ne < nkeep)
          more = kw_next(eigval, val, pos)
          ne = ne + 1
          es(ne) = val
          idx(ne) = pos
       end do
       limit = int(nkeep * truncation_limit_scale)
       ecut = es(nkeep) + truncation_threshold
       cutoff = nkeep
       do while (ne < nfull)
          more = kw_next(eigval, val, pos)
          ne = ne + 1
          es(ne) = val
          idx(ne) = pos
          if (val <= ecut) then
             cutoff = cutoff + 1
          else
             exit
          end if
       end do
       do while (cutoff > limit)
          ecut = ecut - 0.001d0
          do i = cutoff, 1, -1
             if (es(i) > ecut) then
                cutoff = cutoff - 1
             else
                exit
             end if
          end do
       end do
       nb = cutoff
       ! the same set of values as the sorted array: the same minimum
       trunc_margin = minval(abs(eigval - ecut))
    else if (write_files) then
       do i = 1, nfull
          idx(i) = i
       end do
       es = eigval
       call run_merge_sort_r8(nfull, es, idx)
       ne = nfull
    end if
    if (write_files) then
       ! kept energies in ascending order (postprocess.F90:128-137)
       do i = 1, nb
          if (mod(it + 1, 2) == 1) then
             write(u_odd, *) it + 1, real(es(i))
          else
             write(u_even, *) it + 1, real(es(i))
          end if
       end do
    end if
    ! kept prefix back to index order (sortByIndex on 1..nb); every state
    ! kept: index order as it stands
    allocate(kidx(nb))
    if (nb < nfull) then
       allocate(kk(nb), wk(nb), wp(nb))
       do i = 1, nb
          kk(i) = idx(i)
          kidx(i) = idx(i)
       end do
       if (nb > 0) call merge_sort_i8(nb, kk, kidx, wk, wp)
    else
       do i = 1, nb
          kidx(i) = i
       end do
    end if
    ! eigval and basis_output in the new order: the kept prefix now, the
    ! discarded states (ascending energy) when a reader asks
    call move_alloc(eigval, ev_pre)
    allocate(eigval(nfull))
    eigval(1:nb) = ev_pre(kidx)
    if (allocated(basis_out)) deallocate(basis_out)
    allocate(basis_out(nv + 3, nfull))
    b = 1
    do i = 1, nb
       do while (kidx(i) >= subspace(nv + 3, b) + subspace(nv + 1, b))
          b = b + 1
       end do
       call set_state(i, subspace, b, kidx(i) - subspace(nv + 3, b) + 1)
    end do
    bo_sub = subspace(:, 1:nsub)
    tr_nb = nb
    bo_partial = nb < nfull
    ev_partial = nb < nfull
    nkept = nb
    if (write_files) then
       ! kept states in index order with their labels (postprocess.F90:163-167)
       do i = 1, nb
          write(u_info, *) it + 1, real(eigval(i)), basis_out(1:nv, i)
       end do
    end if
  end subroutine truncate

  ! kept states per block (invariantMatrixDGEMM.F90:84-100): blocks are
  ! matched by labels against the kept prefix of basis_output
  subroutine kept_per_block()
    integer :: b, i
    if (allocated(kept_blk)) deallocate(kept_blk)
    allocate(kept_blk(nsub))
    kept_blk = 0
    ! the kept prefix is block-ordered: walk both lists once
    b = 1
    do i = 1, nkept
       do while (b <= nsub)
          if (all(subspace(1:nv, b) == basis_out(1:nv, i))) exit
          b = b + 1
       end do
       if (b > nsub) then
          ! fall back to a full search (never expected)
          do b = 1, nsub
             if (all(subspace(1:nv, b) == basis_out(1:nv, i))) exit
          end do
       end if
       kept_blk(b) = kept_blk(b) + 1
    end do
  end subroutine kept_per_block

  subroutine postprocess(it, ierr)
    integer, intent(in) :: it
    integer, intent(out) :: ierr
    double precision :: t0, t1
    ierr = 0
    call calc_thermo()
    if (write_files) then
       write(u_xn, *) thermo_x
       write(u_sn, *) thermo_s
       write(u_cn, *) thermo_c
    end if
    call truncate(it)
    call kept_per_block()
    ! stock of the truncated iterations (postprocess.F90:213-374); the
    ! library forms this iteration's spectral products now, while the
    ! previous spectral operator is still resident
    if (flag_spectrum .and. it >= start_trunc) then
       call cpu_time_wall(t0)
       ierr = nrg_gpu_stock(int(it, c_int), kept_blk, merge(1_c_int, 0_c_int, it == n_iter))
       if (ierr /= 0) return
       call cpu_time_wall(t1)
       t_stock = t_stock + (t1 - t0)   ! summed over the truncated iterations
    end if
    if (it < n_iter) then
       call cpu_time_wall(t0)
       ierr = nrg_gpu_invariant(kept_blk, c_null_ptr)
       if (ierr /= 0) return
       if (flag_spectrum) then
          ierr = nrg_gpu_invariant_spectrum(kept_blk, c_null_ptr)
          if (ierr /= 0) return
       end if
       call cpu_time_wall(t1)
       t_invariant = t1 - t0
    end if
    nold = nkept
  end subroutine postprocess

  subroutine cpu_time_wall(t)
    double precision, intent(out) :: t
    integer(8) :: c, r
    call system_clock(c, r)
    t = dble(c) / dble(r)
  end subroutine cpu_time_wall

  ! =====================================================================
  ! nrgIteration (nrgIteration.F90:1-31)
  ! =====================================================================
  subroutine nrg_iteration(it, ierr)
    integer, intent(in) :: it
    integer, intent(out) :: ierr
    double precision :: t0, t1, t2, t3
    cur_iteration = it
    call cpu_time_wall(t0)
    call prepare_basis(ierr)
    if (ierr /= 0) return
    call cpu_time_wall(t1)
    call diagonalization(it, ierr)
    if (ierr /= 0) return
    call cpu_time_wall(t2)
    call postprocess(it, ierr)
    if (ierr /= 0) return
    call cpu_time_wall(t3)
    t_prepare = t1 - t0
    t_diag = t2 - t1
    t_post = t3 - t2
    t_iter = t3 - t0
  end subroutine nrg_iteration

  ! =====================================================================
  ! CFS-NRG spectrum phase (main.F90:185-224), after the forward loop
  ! =====================================================================
  ! initializeReducedDensity (initializeReducedDensity.F90:8-81)
  subroutine initialize_reduced_density(ierr)
    integer, intent(out) :: ierr
    ierr = nrg_gpu_rho_init(int(n_iter, c_int))
  end subroutine initialize_reduced_density

  ! reducedDensityMatrixCalculation (reducedDensityMatrixDGEMM.F90:2-377):
  ! shells n_iter-1 down to startTrancation
  subroutine reduced_density_matrix_calculation(ierr)
    integer, intent(out) :: ierr
    integer :: it
    real(c_double) :: tr
    ierr = 0
    if (allocated(rho_trace)) deallocate(rho_trace)
    allocate(rho_trace(i_first:n_iter))
    rho_trace = 0.0d0
    if (flag_trunc) rho_trace(n_iter) = 1.0d0
    do it = n_iter - 1, start_trunc, -1
       ierr = nrg_gpu_rho_shell(int(it, c_int), tr)
       if (ierr /= 0) return
       rho_trace(it) = tr
    end do
  end subroutine reduced_density_matrix_calculation

  ! cfsSpectrumCalculation (module.F90:1683-1706)
  subroutine cfs_spectrum_calculation(it, ierr)
    integer, intent(in) :: it
    integer, intent(out) :: ierr
    if (it == n_iter) then
       ierr = nrg_gpu_spike_zero_temperature(int(it, c_int), energy_scale(it))
    else
       ierr = nrg_gpu_spike_cfs(int(it, c_int), energy_scale(it))
    end if
  end subroutine cfs_spectrum_calculation

  ! keep_history: leave the stocked history (rho, products) on the device
  ! for inspection instead of freeing it
  subroutine spectrum_phase(ierr, keep_history)
    integer, intent(out) :: ierr
    logical, intent(in) :: keep_history
    integer :: it, it_restart
    integer(c_int) :: st(2), itc
    double precision :: t0, t1, t2
    ierr = 0
    if (.not. flag_spectrum) return
    call cpu_time_wall(t0)
    if (load_restart) then
       ! resume: the history, reduced density and bins come from the file
       ! (the NRG loop and the sweep are skipped, main.F90:171-199)
       ierr = nrg_gpu_load_restart(restart_path // c_null_char, itc, 2_c_int, st)
       if (ierr /= 0) return
       it_restart = itc
       start_trunc = st(1)
       flag_trunc = (st(2) /= 0)
       if (allocated(rho_trace)) deallocate(rho_trace)
       allocate(rho_trace(i_first:n_iter))
       rho_trace = 0.0d0   ! the traces of the swept shells are not in the file
    else
       if (flag_trunc) then
          call initialize_reduced_density(ierr)
          if (ierr /= 0) return
       end if
       call reduced_density_matrix_calculation(ierr)
       if (ierr /= 0) return
       it_restart = start_trunc
    end if
    call cpu_time_wall(t1)
    st(1) = start_trunc
    st(2) = merge(1, 0, flag_trunc)
    do it = start_trunc, n_iter
       if (load_restart .and. it < it_restart) cycle
       ! the file is rewritten before every step, after the restart step
       ! when resuming (main.F90:204-218)
       if (save_restart .and. (.not. load_restart .or. it > it_restart)) then
          ierr = nrg_gpu_save_restart(restart_path // c_null_char, int(it, c_int), 2_c_int, st)
          if (ierr /= 0) return
          if (it == restart_stop) return
       end if
       call cfs_spectrum_calculation(it, ierr)
       if (ierr /= 0) return
    end do
    if (allocated(binned_pos)) deallocate(binned_pos, binned_neg)
    allocate(binned_pos(number_of_bin * nmat), binned_neg(number_of_bin * nmat))
    ierr = nrg_gpu_spectrum_bins(binned_neg, binned_pos, int(number_of_bin * nmat, c_int64_t))
    if (ierr /= 0) return
    if (write_files) call write_spectrum_files()
    if (.not. keep_history) ierr = nrg_gpu_history_clear()
    call cpu_time_wall(t2)
    t_rho = t1 - t0
    t_cfs = t2 - t1
  end subroutine spectrum_phase

  ! =====================================================================
  ! output files
  ! =====================================================================
  subroutine open_outputs(dir, ierr)
    character(len=*), intent(in) :: dir
    integer, intent(out) :: ierr
    call close_outputs()
    out_dir = dir
    open(newunit=u_err, file=dir // "/error.log", status="replace", iostat=ierr)
    if (ierr == 0) open(newunit=u_even, file=dir // "/eigeneven.txt", status="replace", iostat=ierr)
    if (ierr == 0) open(newunit=u_odd, file=dir // "/eigenodd.txt", status="replace", iostat=ierr)
    if (ierr == 0) open(newunit=u_xn, file=dir // "/Xn", status="replace", iostat=ierr)
    if (ierr == 0) open(newunit=u_sn, file=dir // "/Sn", status="replace", iostat=ierr)
    if (ierr == 0) open(newunit=u_cn, file=dir // "/Cn", status="replace", iostat=ierr)
    if (ierr == 0) open(newunit=u_info, file=dir // "/keepingState.txt", status="replace", iostat=ierr)
    write_files = (ierr == 0)
  end subroutine open_outputs

  subroutine close_outputs()
    if (u_err /= 0) close(u_err)
    if (u_even /= 0) close(u_even)
    if (u_odd /= 0) close(u_odd)
    if (u_xn /= 0) close(u_xn)
    if (u_sn /= 0) close(u_sn)
    if (u_cn /= 0) close(u_cn)
    if (u_info /= 0) close(u_info)
    u_err = 0; u_even = 0; u_odd = 0; u_xn = 0; u_sn = 0; u_cn = 0; u_info = 0
    write_files = .false.
  end subroutine close_outputs

  ! binned spectra of every type with |weight| > 1e-16 scale(N), frequency
  ! and weight per record (main.F90:228-252); the frequency grid is the
  ! one main.F90:145-150 builds, 1000 bins per decade from 1e-15
  subroutine write_spectrum_files()
    character(len=64) :: fname
    integer :: u, imat, i, side
    double precision :: thr, rec(2)
    thr = 1.0d-16 * energy_scale(n_iter)
    do side = 1, 2
       do imat = 1, nmat
          if (side == 1) then
             write(fname, '("CFS-spectrum-positive-type",i3.3)') imat
          else
             write(fname, '("CFS-spectrum-negative-type",i3.3)') imat
          end if
          open(newunit=u, file=out_dir // "/" // trim(fname), status="replace")
          do i = 1, number_of_bin
             if (side == 1) then
                rec(1) = +10.0d0**(log10_emin + dble(i - 1) / dble(mesh_of_bin))
                rec(2) = binned_pos(i + (imat - 1) * number_of_bin)
             else
                rec(1) = -10.0d0**(log10_emin + dble(i - 1) / dble(mesh_of_bin))
                rec(2) = binned_neg(i + (imat - 1) * number_of_bin)
             end if
             if (abs(rec(2)) > thr) write(u, *) rec(1:2)
          end do
          close(u)
       end do
    end do
    ! the spike files need the last iteration's block table and energies,
    ! which a run resumed from a restart file does not have on the host
    if (nsub > 0 .and. cur_iteration == n_iter) call write_step_files()
  end subroutine write_spectrum_files

  ! zero-temperature spikes of the last iteration, one record per ground
  ! state and state of the excited block, ground states in index order
  ! (calculateSpectrumDGEMM.F90:80-90,114-210): frequency E_k scale(N)
  ! (negated on the negative side) and weight w(k)^2 / g, for |frequency|
  ! > 1e-16; the file names keep the reference's i3.2 fields
  subroutine write_step_files()
    integer :: ip, side, i, b, k, kb, n, g, ierr
    integer, allocatable :: uf(:, :)
    double precision, allocatable :: w(:)
    double precision :: rec(2), sc
    character(len=64) :: fname
    allocate(uf(nmat, 2), w(max(1, maxval(subspace(nv + 1, 1:nsub)))))
    do ip = 1, nmat
       write(fname, '("step_",i3.2,"_itr_",i3.2,".positive_type",i3.3)') 2, n_iter, ip
       open(newunit=uf(ip, 1), file=out_dir // "/" // trim(fname), status="replace")
       write(fname, '("step_",i3.2,"_itr_",i3.2,".negative_type",i3.3)') 2, n_iter, ip
       open(newunit=uf(ip, 2), file=out_dir // "/" // trim(fname), status="replace")
    end do
    g = count(eigval(1:nfull) < degeneracy_threshold)
    sc = energy_scale(n_iter)
    b = 1
    do i = 1, nfull
       if (eigval(i) >= degeneracy_threshold) cycle
       do while (i >= subspace(nv + 3, b) + subspace(nv + 1, b))
          b = b + 1
       end do
       do side = 1, 2
          do ip = 1, nmat
             ierr = nrg_gpu_zero_temperature_weights(int(b, c_int), int(ip, c_int), &
                  merge(1_c_int, 0_c_int, side == 1), w, int(size(w), c_int), kb, n)
             if (ierr /= 0) exit
             do k = 1, n
                rec(1) = eigval(subspace(nv + 3, kb) + k - 1) * sc
                if (side == 2) rec(1) = -rec(1)
                rec(2) = w(k) * w(k) / dble(g)
                if (abs(rec(1)) > 1.0d-16) write(uf(ip, side), *) rec(1:2)
             end do
          end do
       end do
    end do
    do ip = 1, nmat
       close(uf(ip, 1))
       close(uf(ip, 2))
    end do
  end subroutine write_step_files

  ! =====================================================================
  ! setup: inputs + device
  ! =====================================================================
  subroutine setup(dir, device, ierr)
    character(len=*), intent(in) :: dir
    integer, intent(in) :: device
    integer, intent(out) :: ierr
    double precision :: emax, log10emax
    double precision :: lambda_dummy(1)
    lambda_dummy = 0
    t_stock = 0; t_rho = 0; t_cfs = 0
    call read_inputs(dir, ierr)
    if (ierr /= 0) then
       ierr = 100 + ierr
       return
    end if
    ierr = nrg_gpu_init(int(device, c_int))
    if (ierr /= 0) return
    if (flag_spectrum) then
       ierr = nrg_gpu_set_model(int(nv, c_int), int(nops, c_int), int(nchain, c_int), int(ncm, c_int), &
            int(ncoef, c_int), operation, coef_diag_type, coef_diag, coef_inv_type, coef_inv, cons_block, &
            int(nmat, c_int), c_loc(coef_spec), c_loc(cons_diff_spec))
    else
       ierr = nrg_gpu_set_model(int(nv, c_int), int(nops, c_int), int(nchain, c_int), int(ncm, c_int), &
            int(ncoef, c_int), operation, coef_diag_type, coef_diag, coef_inv_type, coef_inv, cons_block, &
            0_c_int, c_null_ptr, c_null_ptr)
    end if
    if (ierr /= 0) return
    if (flag_bcs) then
       ierr = nrg_gpu_set_bcs(int(nbcs, c_int), coef_bcs, int(n_iter + 2, c_int), chain_bcs)
    else
       ierr = nrg_gpu_set_bcs(0_c_int, cons_block, 0_c_int, lambda_dummy)
    end if
    if (ierr /= 0) return
    ierr = nrg_gpu_set_invariant(int(nb0, c_int), inv0)
    if (ierr /= 0) return
    if (flag_spectrum) then
       ierr = nrg_gpu_set_invariant_spectrum(int(nb0, c_int), spec0)
       if (ierr /= 0) return
       ierr = nrg_gpu_set_spectrum_operators(int(nmat, c_int), pos_op, neg_op)
       if (ierr /= 0) return
       ! bins (main.F90:141-156): Emax = scale(-1), 1000 bins per decade from 1e-15
       emax = energy_scale(-1)
       log10emax = dble(int(log10(emax))) + 1.0d0
       number_of_bin = int(log10emax - log10_emin) * mesh_of_bin + 1
       ierr = nrg_gpu_spectrum_setup(int(number_of_bin, c_int))
       if (ierr /= 0) return
    end if
  end subroutine setup

end module nrg_host
