! Test-infrastructure driver around the reference Fortran sources.
!
! This file is NOT part of the product.  It is compiled together with the
! reference's own routines (from the read-only reference tree, by
! oracle/Makefile) into oracle/_ref/nrg_ref, replacing only the reference's
! `program main` (main.F90) so that the module state can be dumped in full
! precision after every NRG iteration and after the CFS phase.  Those dumps
! become the golden fixtures under tests/golden/.
!
! The sequence of calls is the one main.F90 performs (main.F90:22-224):
! setupHamiltonian -> setSpectrumInfo -> read_parameter -> read_initial_basis
! -> stockInitialState -> bin grid -> nrgIteration(i0..N) ->
! initializeReducedDensity -> reducedDensityMatrixCalculation ->
! cfsSpectrumCalculation(startTrancation..N) -> showCount.
!
! Usage (inside a directory holding the three input files):
!   nrg_ref [full_dump_upto [stop_after]]
! full_dump_upto (default -100): iterations <= this value also dump the
! eigenvector array and the invariant matrices.
! stop_after (default: the deck's last iteration): the run ends after this
! iteration, skipping the spectrum phase.  Iterations up to it are exactly
! those of the complete run (the last-iteration rule of postprocess.F90:79-81
! looks at the deck's numberOfIteration, not at stop_after).
program ref_driver
  use parallel
  use system
  use io_param
  use hamiltonian
  use spectrum
  implicit none
  include 'mpif.h'

  integer :: ierr, it, imat, ib, full_upto, nargs, stop_after
  character(64) :: arg

  call MPI_INIT(ierr)
  call MPI_COMM_SIZE(MPI_COMM_WORLD, numberOfProcess, ierr)
  call MPI_COMM_RANK(MPI_COMM_WORLD, my_rank, ierr)
  call initCount
  call startCount("main")

  full_upto = -100
  nargs = command_argument_count()
  if (nargs >= 1) then
     call get_command_argument(1, arg)
     read(arg, *) full_upto
  end if
  stop_after = huge(stop_after)
  if (nargs >= 2) then
     call get_command_argument(2, arg)
     read(arg, *) stop_after
  end if

  call setupHamiltonian
  if (hami%flag_spectrum) call setSpectrumInfo
  allocate(chain_diagonal(0:hami%numberOfIteration+1, hami%numberOfChain))
  allocate(chain_nondiagonal(hami%indexOfFirstIteration:hami%numberOfIteration+1, hami%numberOfChain))
  if (hami%flag_BCS) allocate(chain_BCS(0:hami%numberOfIteration+1, hami%numberOfChain))
  call read_parameter
  open(errorlog, file=errorfile)
  open(eigeneven, file=fileEigenEven)
  open(eigenodd, file=fileEigenOdd)
  open(thermoXn, file=fileXn)
  open(thermoSn, file=fileSn)
  open(thermoCn, file=fileCn)
  open(stateInfomation, file=fileInfo)
  call read_initial_basis
  call stockInitialState

  ! logarithmic bin grid, 1000 bins per decade from 1e-15 up to the first scale
  Emax = scale(-1)
  log10Emax = int(log10(Emax)) + 1.0d0
  numberOfBin = int(log10Emax - log10Emin) * meshOfBin + 1
  allocate(binnedSpikePositive(numberOfBin*hami%numberOfMatrix, 2))
  allocate(binnedSpikeNegative(numberOfBin*hami%numberOfMatrix, 2))
  binnedSpikePositive = 0.0d0
  binnedSpikeNegative = 0.0d0
  do imat = 1, hami%numberOfMatrix
     do ib = 1, numberOfBin
        binnedSpikePositive(ib+(imat-1)*numberOfBin, 1) = +10.0d0**(log10Emin+dble(ib-1)/dble(meshOfBin))
        binnedSpikeNegative(ib+(imat-1)*numberOfBin, 1) = -10.0d0**(log10Emin+dble(ib-1)/dble(meshOfBin))
     end do
  end do
  numberOfBasis_old = hami%numberOfBasis

  call startCount("main:nrg")
  do it = hami%indexOfFirstIteration, hami%numberOfIteration
     call nrgIteration(it)
     call dump_iteration(it, it <= full_upto)
     if (it >= stop_after) exit
  end do
  call stopCount

  if (hami%flag_spectrum .and. stop_after >= hami%numberOfIteration) then
     call startCount("main:spec")
     if (hami%flag_truncation) call initializeReducedDensity
     elementNumFullBasisInput  = pointFullBasisInput - 1
     elementNumFullBasisOutput = pointFullBasisOutput - 1
     elementNumFullSubspace    = pointFullSubspace - 1
     elementNumReducedDensity  = pointReducedDensity - 1
     elementNumFullEigenVector = int(pointFullEigenvector - 1)
     call reducedDensityMatrixCalculation
     do it = hami%startTrancation, hami%numberOfIteration
        call cfsSpectrumCalculation(it)
     end do
     call stopCount
     call dump_spectrum()
     call write_spectrum_files()
  end if

  call stopCount
  call showCount
  call MPI_FINALIZE(ierr)

contains

  subroutine dump_iteration(it, full)
    integer, intent(in) :: it
    logical, intent(in) :: full
    character(64) :: fname
    integer :: u, nv
    if (my_rank /= 0) return
    nv = hami%numberOfVariation
    write(fname, '("dump_it",i4.4,".bin")') it + 1000
    open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
    write(u) int(it, 4), int(nv, 4), int(numberOfSubspace, 4), int(numberOfBasis_full, 4), &
         int(hami%numberOfBasis, 4), int(hami%startTrancation, 4), int(hami%numberOfConductionMatrix, 4)
    write(u) subspaceInfo(1)%memory(1:numberOfSubspace*(nv+3))
    write(u) basis_input(1)%memory(1:numberOfBasis_full*(nv+3))
    write(u) basis_output(1)%memory(1:numberOfBasis_full*(nv+3))
    write(u) eigenvalue(1:numberOfBasis_full)
    if (full) then
       write(u) int(1, 4), int(size(eigenvector), 8)
       write(u) eigenvector
       if (it < hami%numberOfIteration) then
          write(u) int(size(invariant_matrix, 1), 4)
          write(u) invariant_matrix
          if (hami%flag_spectrum) then
             write(u) int(size(invariant_matrix_spectrum, 1), 4)
             write(u) invariant_matrix_spectrum
          else
             write(u) int(0, 4)
          end if
       else
          write(u) int(0, 4), int(0, 4)
       end if
    else
       write(u) int(0, 4)
    end if
    close(u)
  end subroutine dump_iteration

  ! the binned-spectrum files main.F90:228-252 writes: per type and side,
  ! the (frequency, weight) records of the bins above 1e-16 scale(N)
  subroutine write_spectrum_files()
    character(64) :: fname
    integer :: u, imat, i
    if (my_rank /= 0) return
    do imat = 1, hami%numberOfMatrix
       write(fname, '("CFS-spectrum-positive-type",i3.3)') imat
       open(newunit=u, file=trim(fname), status='replace')
       do i = 1, numberOfBin
          if (abs(binnedSpikePositive(i+(imat-1)*numberOfBin, 2)) > 1.0d-16*scale(hami%numberOfIteration)) &
               write(u, *) binnedSpikePositive(i+(imat-1)*numberOfBin, 1:2)
       end do
       close(u)
    end do
    do imat = 1, hami%numberOfMatrix
       write(fname, '("CFS-spectrum-negative-type",i3.3)') imat
       open(newunit=u, file=trim(fname), status='replace')
       do i = 1, numberOfBin
          if (abs(binnedSpikeNegative(i+(imat-1)*numberOfBin, 2)) > 1.0d-16*scale(hami%numberOfIteration)) &
               write(u, *) binnedSpikeNegative(i+(imat-1)*numberOfBin, 1:2)
       end do
       close(u)
    end do
  end subroutine write_spectrum_files

  subroutine dump_spectrum()
    integer :: u, nv
    if (my_rank /= 0) return
    nv = hami%numberOfVariation
    open(newunit=u, file='dump_spectrum.bin', access='stream', form='unformatted', status='replace')
    write(u) int(numberOfBin, 4), int(hami%numberOfMatrix, 4), int(hami%startTrancation, 4), &
         int(elementNumFullSubspace, 4), int(sizeOfReducedDensity, 4), int(nv, 4)
    write(u) binnedSpikePositive(:, 2)
    write(u) binnedSpikeNegative(:, 2)
    write(u) mapFullSubspaceInfo(hami%startTrancation:hami%numberOfIteration)
    write(u) fullSubspaceInfo(1)%memory(1:elementNumFullSubspace*(nv+6))
    write(u) reducedDensityMatrix(1:sizeOfReducedDensity)
    close(u)
  end subroutine dump_spectrum

end program ref_driver
