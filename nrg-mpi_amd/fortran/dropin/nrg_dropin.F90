! Drop-in bodies of the reference's hot-path routines (eminamitani/nrg-mpi)
! on libnrg_gpu.so.  A maintainer replaces the reference's files with these
! subroutines (same names and arguments, same module globals in and out) and
! links -lnrg_gpu; nrg_gpu_api.F90 supplies the interfaces.  The file is not
! part of the library build: it compiles against the reference's own
! modules (hamiltonian, system, spectrum), so that the binding it shows is
! checked by a compiler.
!
! Replaced routines and the call sites that stay as they are:
!   diagonalization(iteration, ismysubspace)   diagonalization.F90:1-410, nrgIteration.F90:18-28
!   invariantMatrixDGEMM(iteration)            invariantMatrixDGEMM.F90:1-394, postprocess.F90:194
!   invariantMatrixForSpectrumDGEMM(iteration) invariantMatrixForSpectrumDGEMM.F90:1-392, postprocess.F90:197
!   initializeReducedDensity                   initializeReducedDensity.F90:8-81, main.F90:189
!   reducedDensityMatrixCalculation            reducedDensityMatrixDGEMM.F90:2-377, main.F90:200
!   spikeZeroTemperature(iteration)            calculateSpectrumDGEMM.F90:1-215, module.F90:1700
!   spikeCFS(iteration, numberOfBasisFull)     calculateSpectrumDGEMM.F90:217-646, module.F90:1702
! and three additions the caller makes:
!   gpuSetup          after read_initial_basis and the bin container set-up (main.F90:120-156)
!   gpuStock(iteration)  in postprocess, after the truncation and before the invariant calls
!                     (in place of the restore bookkeeping, postprocess.F90:213-374)
!   gpuSpectrumBins   after the CFS loop, before the spectrum files are written (main.F90:224-228)
! cfsSpectrumCalculation (module.F90:1683-1706) keeps its body except the
! getRestore call: the history stays in device memory.

module nrg_dropin_util
  use iso_c_binding
  use nrg_gpu_api
  implicit none
contains
  subroutine gpu_check(ierr, where)
    integer(c_int), intent(in) :: ierr
    character(*), intent(in) :: where
    if (ierr /= 0) then
       print *, "libnrg_gpu: ", where, ": ", gpu_error_text()
       stop 1
    end if
  end subroutine gpu_check

  ! kept states per block after the truncation: the count of basis_output
  ! records with the block's labels (invariantMatrixDGEMM.F90:84-100)
  subroutine kept_per_block(kept)
    use hamiltonian, only: hami, subspaceInfo, basis_output, numberOfSubspace
    integer(c_int), intent(out) :: kept(:)
    integer :: isub, j
    do isub = 1, numberOfSubspace
       kept(isub) = 0
       do j = 1, hami%numberOfBasis
          if (all(subspaceInfo(isub)%basis(:) == basis_output(j)%basis(:))) kept(isub) = kept(isub) + 1
       end do
    end do
  end subroutine kept_per_block
end module nrg_dropin_util

subroutine gpuSetup
  use iso_c_binding
  use nrg_gpu_api
  use nrg_dropin_util
  use hamiltonian
  use spectrum, only: numberOfBin, positiveSpectrumOperation, negativeSpectrumOperation
  implicit none
  type(c_ptr) :: pspec, pdiff
  ! the module arrays are not TARGETs: the optional tables go through
  ! copies (the library copies them during the call)
  double precision, allocatable, target :: spec(:, :)
  integer(c_int), allocatable, target :: diff(:, :)
  call gpu_check(nrg_gpu_init(0_c_int), "init")
  pspec = c_null_ptr
  pdiff = c_null_ptr
  if (hami%flag_spectrum) then
     spec = coefficient_invariant_matrix_spectrum
     diff = conservation_difference_spectrum
     pspec = c_loc(spec)
     pdiff = c_loc(diff)
  end if
  call gpu_check(nrg_gpu_set_model(int(hami%numberOfVariation, c_int), int(hami%numberOfOperation, c_int), &
       int(hami%numberOfChain, c_int), int(hami%numberOfConductionMatrix, c_int), &
       int(hami%numberOfCoefficient, c_int), operation, coefficient_diagonalization_type, &
       coefficient_diagonalization, coefficient_invariant_matrix_type, coefficient_invariant_matrix, &
       hami%conservationBlock, int(hami%numberOfMatrix, c_int), pspec, pdiff), "set_model")
  if (hami%flag_BCS) &
       call gpu_check(nrg_gpu_set_bcs(int(hami%numberOfBCSCoefficient, c_int), coefficient_BCS, &
       int(size(chain_BCS, 1), c_int), chain_BCS), "set_bcs")
  call gpu_check(nrg_gpu_set_invariant(int(hami%numberOfBasis, c_int), invariant_matrix), "set_invariant")
  if (hami%flag_spectrum) then
     call gpu_check(nrg_gpu_set_invariant_spectrum(int(hami%numberOfBasis, c_int), invariant_matrix_spectrum), &
          "set_invariant_spectrum")
     call gpu_check(nrg_gpu_set_spectrum_operators(int(hami%numberOfMatrix, c_int), &
          positiveSpectrumOperation, negativeSpectrumOperation), "set_spectrum_operators")
     call gpu_check(nrg_gpu_spectrum_setup(int(numberOfBin, c_int)), "spectrum_setup")
  end if
end subroutine gpuSetup

subroutine diagonalization(iteration, ismysubspace)
  use iso_c_binding
  use nrg_gpu_api
  use nrg_dropin_util
  use parallel, only: numberOfProcess
  use hamiltonian, only: hami, subspaceInfo, basis_input, basis_output, eigenvalue, numberOfSubspace, &
       groundStateEnergy, sizeOfEigenvector, chain_diagonal, chain_nondiagonal, numberOfBasis_old, allocateType
  use system, only: param
  implicit none
  integer, intent(in) :: iteration
  ! one device owns every block: the rank map is not read
  logical, intent(in) :: ismysubspace(numberOfSubspace, 0:numberOfProcess - 1)
  double precision, allocatable :: eigenvalue_old(:)
  integer :: isub, j, start_output
  integer(c_int64_t) :: n2

  allocate(eigenvalue_old(max(numberOfBasis_old, 1)))
  eigenvalue_old(1:numberOfBasis_old) = eigenvalue(1:numberOfBasis_old)
  n2 = 0
  do isub = 1, numberOfSubspace
     n2 = n2 + int(subspaceInfo(isub)%dimension, c_int64_t)**2
  end do
  if (allocated(eigenvalue)) deallocate(eigenvalue)
  allocate(eigenvalue(hami%numberOfBasis))
  ! the eigenvectors stay in device memory for the invariant and CFS calls
  ! of this iteration (eigenvector = c_null_ptr); the library applies the
  ! ground-state shift and the 1e-12 chop of diagonalization.F90:348-356
  call gpu_check(nrg_gpu_diagonalize(int(iteration, c_int), int(numberOfSubspace, c_int), &
       subspaceInfo(1)%memory, int(hami%numberOfBasis, c_int), basis_input(1)%memory, &
       int(numberOfBasis_old, c_int), eigenvalue_old, chain_diagonal(iteration + 1, :), &
       chain_nondiagonal(iteration, :), sqrt(param%Lambda), eigenvalue, c_null_ptr, n2), "diagonalize")
  sizeOfEigenvector = int(n2)
  ! written by the reference (diagonalization.F90:351) and read nowhere; the
  ! library shifts the eigenvalues on the device and does not return it
  groundStateEnergy = 0.0d0
  call allocateType(basis_output, hami%numberOfBasis, hami%numberOfVariation)
  do isub = 1, numberOfSubspace
     start_output = subspaceInfo(isub)%start_input
     do j = 1, subspaceInfo(isub)%dimension
        basis_output(start_output + j - 1)%basis(:) = subspaceInfo(isub)%basis(:)
        basis_output(start_output + j - 1)%reference = j
        basis_output(start_output + j - 1)%variation = 0
        basis_output(start_output + j - 1)%operation = 0
     end do
  end do
  if (.false.) print *, ismysubspace(1, 0)
end subroutine diagonalization

subroutine invariantMatrixDGEMM(iteration)
  use iso_c_binding
  use nrg_gpu_api
  use nrg_dropin_util
  use hamiltonian, only: numberOfSubspace
  implicit none
  integer :: iteration
  integer(c_int), allocatable :: kept(:)
  allocate(kept(numberOfSubspace))
  call kept_per_block(kept)
  ! the invariant matrix stays in device memory for the next assembly
  call gpu_check(nrg_gpu_invariant(kept, c_null_ptr), "invariant")
  if (.false.) print *, iteration
end subroutine invariantMatrixDGEMM

subroutine invariantMatrixForSpectrumDGEMM(iteration)
  use iso_c_binding
  use nrg_gpu_api
  use nrg_dropin_util
  use hamiltonian, only: numberOfSubspace
  implicit none
  integer :: iteration
  integer(c_int), allocatable :: kept(:)
  allocate(kept(numberOfSubspace))
  call kept_per_block(kept)
  call gpu_check(nrg_gpu_invariant_spectrum(kept, c_null_ptr), "invariant_spectrum")
  if (.false.) print *, iteration
end subroutine invariantMatrixForSpectrumDGEMM

subroutine gpuStock(iteration)
  use iso_c_binding
  use nrg_gpu_api
  use nrg_dropin_util
  use hamiltonian, only: hami, numberOfSubspace
  implicit none
  integer, intent(in) :: iteration
  integer(c_int), allocatable :: kept(:)
  if (.not. hami%flag_spectrum .or. iteration < hami%startTrancation) return
  allocate(kept(numberOfSubspace))
  call kept_per_block(kept)
  call gpu_check(nrg_gpu_stock(int(iteration, c_int), kept, &
       merge(1_c_int, 0_c_int, iteration == hami%numberOfIteration)), "stock")
end subroutine gpuStock

subroutine initializeReducedDensity
  use iso_c_binding
  use nrg_gpu_api
  use nrg_dropin_util
  use hamiltonian, only: hami
  implicit none
  call gpu_check(nrg_gpu_rho_init(int(hami%numberOfIteration, c_int)), "rho_init")
end subroutine initializeReducedDensity

subroutine reducedDensityMatrixCalculation
  use iso_c_binding
  use nrg_gpu_api
  use nrg_dropin_util
  use hamiltonian, only: hami
  implicit none
  integer :: iteration
  real(c_double) :: trace
  do iteration = hami%numberOfIteration - 1, hami%startTrancation, -1
     call gpu_check(nrg_gpu_rho_shell(int(iteration, c_int), trace), "rho_shell")
  end do
end subroutine reducedDensityMatrixCalculation

subroutine spikeZeroTemperature(iteration)
  use iso_c_binding
  use nrg_gpu_api
  use nrg_dropin_util
  use system, only: scale
  implicit none
  integer :: iteration
  call gpu_check(nrg_gpu_spike_zero_temperature(int(iteration, c_int), scale(iteration)), "spike_zero_temperature")
end subroutine spikeZeroTemperature

subroutine spikeCFS(iteration, numberOfBasisFull)
  use iso_c_binding
  use nrg_gpu_api
  use nrg_dropin_util
  use system, only: scale
  implicit none
  integer :: iteration, numberOfBasisFull
  call gpu_check(nrg_gpu_spike_cfs(int(iteration, c_int), scale(iteration)), "spike_cfs")
  if (.false.) print *, numberOfBasisFull
end subroutine spikeCFS

subroutine gpuSpectrumBins
  use iso_c_binding
  use nrg_gpu_api
  use nrg_dropin_util
  use hamiltonian, only: hami
  use spectrum, only: numberOfBin, binnedSpikePositive, binnedSpikeNegative
  implicit none
  ! the weight columns; the omega columns are the reference's grid
  call gpu_check(nrg_gpu_spectrum_bins(binnedSpikeNegative(:, 2), binnedSpikePositive(:, 2), &
       int(numberOfBin, c_int64_t) * int(hami%numberOfMatrix, c_int64_t)), "spectrum_bins")
end subroutine gpuSpectrumBins
