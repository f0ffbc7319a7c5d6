/*
 * nrg_gpu.h -- C ABI of the MI355X (gfx950) NRG hot-path library libnrg_gpu.so.
 *
 * The reference program (eminamitani/nrg-mpi, Fortran 90 + MPI + MKL) runs
 * its hot path as argument-less subroutines that read and write module
 * globals.  Each entry point below replaces the numerical body of one of
 * them; the caller (a Fortran host through iso_c_binding, or any C caller)
 * keeps the NRG loop, the (Q,Sz) block bookkeeping and the truncation, and
 * passes the module arrays as plain pointers.
 *
 * Conventions
 *   - integers are int32_t (Fortran default integer), offsets into the packed
 *     eigenvector array are int64_t (the reference's int32 offsets overflow
 *     for the large workloads);
 *   - arrays are Fortran column-major; the packed records are the reference's
 *     contiguous `memory(:)` arrays: SubspaceInfo (labels(nv), dimension,
 *     count_eigenvector, start_input) and Basis (labels(nv), reference,
 *     variation, operation), all 1-based values;
 *   - every call returns 0 on success, a nonzero status otherwise, with a
 *     message from nrg_gpu_last_error();
 *   - one device, one host thread, not reentrant.  Heavy arrays
 *     (eigenvectors, invariant matrices, history) stay resident in HBM
 *     between calls; host copies are made only when an output pointer is
 *     non-NULL.
 */
#ifndef NRG_GPU_H
#define NRG_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- lifecycle -------------------------------------------------------- */
/* Select the device and create the stream/context.  Replaces the MPI setup
 * of main.F90:18-20 (one device instead of ranks). */
int nrg_gpu_init(int device);
int nrg_gpu_finalize(void);
const char* nrg_gpu_last_error(void);

/* ---- model constants --------------------------------------------------- */
/* Static tables read by setupHamiltonian / read_initial_basis
 * (module.F90:826-944, read_initial_basis.F90:160-302).
 *   operation(nops, nv+nchain), coefficient_diagonalization_type(ncoef,4),
 *   coefficient_diagonalization(ncoef), coefficient_invariant_matrix_type
 *   (ncoef,3), coefficient_invariant_matrix(ncoef), conservationBlock(2,2),
 *   coefficient_invariant_matrix_spectrum(nmat,nops) and
 *   conservation_difference_spectrum(nmat,nv) (both may be NULL when the
 *   spectrum is off). */
int nrg_gpu_set_model(int nv, int nops, int nchain, int ncm, int ncoef,
                      const int32_t* operation, const int32_t* coef_diag_type,
                      const double* coef_diag, const int32_t* coef_inv_type,
                      const double* coef_inv, const int32_t* conservation_block,
                      int nmat, const double* coef_spec, const int32_t* cons_diff_spec);

/* Superconducting (BCS) chain of a flag_BCS model (read_initial_basis.F90:
 * 185-201; used at diagonalization.F90:215-245): coefficient_BCS(nbcs,3)
 * rows (op_i, op_j, channel) and the raw chain_BCS(0:nrow-1, 1:nchain),
 * column-major (nrow = numberOfIteration + 2).  During the H assembly of
 * iteration N, a product-state pair i < j of one block whose operations
 * match a row in this order and whose references are equal gets
 * H(i,j) = H(j,i) = chain_BCS(N+1, channel) of the last matching row, in
 * place of the hopping term.  nbcs = 0 removes the BCS terms (the default
 * after nrg_gpu_set_model). */
int nrg_gpu_set_bcs(int nbcs, const int32_t* coef_bcs, int nrow, const double* chain_bcs);

/* Initial invariant matrices of the impurity basis:
 * invariant_matrix(nold,nold,ncm) (read_initial_basis.F90:163-173) and
 * invariant_matrix_spectrum(nold,nold,nmat) (read_initial_basis.F90:271-282). */
int nrg_gpu_set_invariant(int nold, const double* invariant_matrix);
int nrg_gpu_set_invariant_spectrum(int nold, const double* invariant_matrix_spectrum);

/* ---- diagonalization (diagonalization.F90:1-410) ------------------------
 * Builds every block Hamiltonian of the iteration from the previous kept
 * eigenvalues and the device-resident invariant matrix (:153-263), solves all
 * blocks in one batched eigensolver (:266-269), chops |U| < 1e-14
 * (:273-280), shifts the eigenvalues by their global minimum and zeroes
 * those <= 1e-12 (:351-356).  `ismysubspace` has no counterpart: one device
 * owns every block.
 *   subspace_info  nsub*(nv+3)  (prepareBasis.F90:93-114)
 *   basis_input    nfull*(nv+3) (prepareBasis.F90:41-59)
 *   eigenvalue_old nold          eigenvalues of the previous kept basis
 *   chain_diag     chain_diagonal(iteration+1, 1:nchain) after rescaling
 *   chain_nondiag  chain_nondiagonal(iteration, 1:nchain) after rescaling
 *   eigenvalue     out, nfull, block order (start_input offsets)
 *   eigenvector    out (may be NULL), packed at count_eigenvector offsets,
 *                  eigenvector_size = sum of dimension^2 */
int nrg_gpu_diagonalize(int iteration, int nsub, const int32_t* subspace_info, int nfull,
                        const int32_t* basis_input, int nold, const double* eigenvalue_old,
                        const double* chain_diag, const double* chain_nondiag,
                        double sqrt_lambda, double* eigenvalue, double* eigenvector,
                        int64_t eigenvector_size);

/* ---- invariant matrices ------------------------------------------------
 * invariantMatrixDGEMM (invariantMatrixDGEMM.F90:1-394): new conduction
 * invariant matrices of the kept basis from the eigenvectors of the last
 * nrg_gpu_diagonalize call.  kept_per_block(nsub) is keeped_basis_number
 * (:84-100).  The result stays on the device as the input of the next
 * diagonalization; invariant_matrix_out (k,k,ncm) may be NULL, and then the
 * call returns as soon as the work is queued on the library's stream (every
 * later call of the library is ordered after it). */
int nrg_gpu_invariant(const int32_t* kept_per_block, double* invariant_matrix_out);

/* Copy the device-resident invariant matrix of the current kept basis (the
 * reference's module global invariant_matrix(k,k,numberOfConductionMatrix),
 * module.F90:809, column-major) to the host: the last nrg_gpu_invariant
 * result, or the nrg_gpu_set_invariant input before the first iteration.
 * size = k*k*ncm (checked). */
int nrg_gpu_invariant_get(double* invariant_matrix, int64_t size);

/* invariantMatrixForSpectrumDGEMM (invariantMatrixForSpectrumDGEMM.F90:1-392). */
int nrg_gpu_invariant_spectrum(const int32_t* kept_per_block, double* invariant_matrix_spectrum_out);

/* ---- CFS spectra: history, reduced density, spikes -----------------------
 * Label shifts of the spectral operators, positive/negative_spectrum_operation
 * (nmat,nv) read with the model (read_initial_basis.F90:168-182). */
int nrg_gpu_set_spectrum_operators(int nmat, const int32_t* pos_op, const int32_t* neg_op);

/* Stock of one iteration >= startTrancation (postprocess.F90:221-374 and
 * appendRestore, module.F90:1345-1570): the library keeps the kept columns of
 * every block, the block table, the product-basis records and the energies
 * of the last nrg_gpu_diagonalize call in HBM, and forms the spectral
 * products that spikeCFS (or, with last != 0, spikeZeroTemperature) of this
 * iteration will weight, while the previous spectral operator is resident.
 * Call after the truncation and before nrg_gpu_invariant_spectrum.  A stock
 * of an iteration not above the last stocked one starts a new history. */
int nrg_gpu_stock(int iteration, const int32_t* kept_per_block, int last);

/* initializeReducedDensity (initializeReducedDensity.F90:8-81): rho of the
 * last iteration = 1/g on its g ground states (E < 1e-8). */
int nrg_gpu_rho_init(int last_iteration);

/* One shell of reducedDensityMatrixCalculation (reducedDensityMatrixDGEMM.F90:
 * 62-321): rho(iteration) from rho(iteration+1); *trace receives Tr rho
 * (printed by the reference at :364-367). */
int nrg_gpu_rho_shell(int iteration, double* trace);

/* Copy rho(iteration) to the host in the reference's reducedDensityMatrix
 * layout: one k_b x k_b row-major block per stocked block in block order
 * (:308-314).  size = sum k_b^2 (checked). */
int nrg_gpu_rho_get(int iteration, double* rho, int64_t size);

/* Bins of binnedSpikeNegative/Positive (main.F90:142-156): clear them for
 * numberOfBin bins per type. */
int nrg_gpu_spectrum_setup(int nbin);

/* spikeCFS (calculateSpectrumDGEMM.F90:217-646) for iteration < last and
 * spikeZeroTemperature (:1-215) for the last iteration, both as
 * cfsSpectrumCalculation calls them (module.F90:1683-1706).  scale =
 * scale(iteration) (module.F90:690-711); spikes are binned on the device. */
int nrg_gpu_spike_cfs(int iteration, double scale);
int nrg_gpu_spike_zero_temperature(int iteration, double scale);

/* The spectral products behind spikeZeroTemperature's spikes
 * (calculateSpectrumDGEMM.F90:114-210) after nrg_gpu_stock of the last
 * iteration, for the step_* spike files the reference writes there
 * (:80-90,157-159,203-205): for the ground-state block gs_block (1-based,
 * last iteration) and spectral type `type` (1-based), on the positive
 * (positive != 0) or negative side, *excited_block receives the 1-based
 * block whose labels are the ground-state labels plus the type's
 * positive/negative operation (0 when there is none; then *n = 0), *n its
 * dimension, and w(k) the product of its state k with column 1 of the
 * ground-state block (zeros when the two blocks share no operation).  A
 * ground state's spike of state k has weight w(k)^2 / g (g ground states).
 * cap >= the excited block's dimension (checked). */
int nrg_gpu_zero_temperature_weights(int gs_block, int type, int positive, double* w, int cap,
                                     int* excited_block, int* n);

/* Copy the bin weights: negative and positive, each (nbin*nmat) with bin i
 * of type t at i + t*nbin (the weight column of binnedSpike*, main.F90:146). */
int nrg_gpu_spectrum_bins(double* negative, double* positive, int64_t size);

/* Restart file of the CFS phase (saveRestart / loadRestart, module.F90:
 * 1723-1890, used at main.F90:165-221).  Save before the CFS step of
 * `iteration`: the bins so far, the device-resident history of the
 * iterations >= `iteration` (kept columns, spectral products, reduced
 * density) and nstate integers of the caller's own state, into one stream
 * file led by the reference's mode word MODE_RESTART_CFS (1002) and the
 * iteration; the file is replaced only once complete.  Load restores the
 * history and bins after nrg_gpu_init / set_model / spectrum set-up and
 * returns the iteration at which the CFS steps resume and the caller's
 * state; the forward pass and the reduced-density sweep are not needed.
 * Errors: a missing file, a mode word other than 1002 ("broken restart
 * file"), another model or a different nstate. */
int nrg_gpu_save_restart(const char* path, int iteration, int nstate, const int32_t* state);
int nrg_gpu_load_restart(const char* path, int* iteration, int nstate, int32_t* state);

/* Free the stocked history (no counterpart; the reference frees at exit). */
int nrg_gpu_history_clear(void);

/* ---- per-stage timing of the last calls (milliseconds) -----------------
 * Never waits for the device: t[4] (its deferred part) and t[5] time work
 * that returns before the device has done it, and report the last such
 * span that has completed.
 * t[0] H assembly, t[1] eigensolver total, t[2] tridiagonalisation,
 * t[3] divide and conquer, t[4] back-transform, t[5] invariant,
 * t[6] spectral invariant, t[7] rho sweep, t[8] CFS.  Also returns the
 * algorithmic FP64 flops of the last eigensolve (sum 10/3 n^3) in t[9] and,
 * for n >= 12, the tridiagonalisation graphs captured (t[10]) and replayed
 * (t[11]) since init (nrg_gpu_set_graphs); for n >= 13, the host time spent
 * issuing the last tridiagonalisation's launches (t[12], <= t[2]: equal when
 * the host, not the device, sets its pace). */
int nrg_gpu_stage_times(double* t, int n);

/* ---- batched eigensolver on host arrays (kernel-level interface) -------
 * Solves nblk packed symmetric blocks like one dsyev('V','U') each
 * (diagonalization.F90:267): H packed column-major at offsets sum n^2,
 * w at offsets sum n.  U receives the eigenvectors (no chop, no shift). */
int nrg_gpu_syev_batched(int nblk, const int32_t* n, const double* H, double* w, double* U);

/* ---- batched FP64 GEMM on host arrays (kernel-level interface) ---------
 * C = alpha op(A) op(B) + beta C, column-major, op = N (0) or T (1): the
 * dgemm call sites of the reference (e.g. invariantMatrixDGEMM.F90:300). */
int nrg_gpu_dgemm(int transa, int transb, int m, int n, int k, double alpha,
                  const double* A, int lda, const double* B, int ldb, double beta,
                  double* C, int ldc);

/* ---- split-precision contraction GEMMs ------------------------------------
 * Form with no counterpart in the reference: the contraction products of the
 * invariant (invariantMatrixDGEMM.F90:300,316), spectral invariant
 * (invariantMatrixForSpectrumDGEMM.F90:296,312), CFS weight
 * (module.F90:1100,1113) and reduced-density (reducedDensityMatrixDGEMM.F90:
 * 280,296) call sites run as FP64 products recovered from int8 matrix-core
 * products of `slices` signed 7-bit digits per element (Ozaki slicing, exact
 * int32 accumulation flushed to FP64; 7*slices - 1 bits of each operand).
 * 0 (the default) keeps FP64 matrix cores; 3..8 slices.  The eigensolver
 * always stays FP64. */
int nrg_gpu_set_gemm_split(int slices);
/* The same option per product family (no counterpart in the reference):
 * invariant_slices for the invariant-matrix products, whose results feed the
 * next iteration's Hamiltonian and so its eigenvalues (1e-12 relative: 8
 * slices), spectral_slices for the spectral invariant, CFS weight and
 * reduced-density products, which feed only the binned spectrum (1e-8 of
 * the peak bin: 6 slices, 41 bits per operand, are enough); a batch below
 * min_batch_flops FP64-equivalent flops (2mnk summed over its products)
 * stays on FP64 matrix cores.  nrg_gpu_set_gemm_split(s) is
 * (s, s, 0). */
int nrg_gpu_set_gemm_split_families(int invariant_slices, int spectral_slices, double min_batch_flops);

/* Batched GEMM on host arrays (kernel-level interface for the synthetic
 * block batch): task i computes C_i = op(A_i) op(B_i), m x n x k = mnk[3i..],
 * packed column-major at the given element offsets (ld = stored rows).
 * slices 0: FP64 matrix cores, 3..8: split precision.  With repeat > 0 the
 * product is run `repeat` more times on the device and *ms receives the mean
 * milliseconds per batch (HIP events). */
int nrg_gpu_dgemm_batched(int ntask, const int32_t* mnk, int transa, int transb, const double* A,
                          const int64_t* offA, const double* B, const int64_t* offB, double* C,
                          const int64_t* offC, int slices, int repeat, double* ms);

/* No counterpart in the reference: the FP64 form of nrg_gpu_dgemm_batched
 * for callers that cannot allocate device memory (code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc).  A, B and C are
 * device arrays owned by the caller (same task layout and offsets as
 * nrg_gpu_dgemm_batched, C_i = op(A_i) op(B_i)); the task tables and split-K
 * partial products go to a statically allocated per-device scratch array of
 * 128 MiB, so the call allocates nothing.  At most 4096 products per call,
 * and their tables and partials must fit the scratch (both checked, status
 * 2 with a message).  Returns after the products are done.  Every call on a
 * device shares that device's scratch: calls that overlap on one device
 * race. */
int nrg_gpu_dgemm_batched_noalloc(int ntask, const int32_t* mnk, int transa, int transb, const double* A,
                                  const int64_t* offA, const double* B, const int64_t* offB, double* C,
                                  const int64_t* offC);

/* Per-kernel-family HIP-event timing (no counterpart in the reference; the
 * reference's count.f90 timers measure host wall clock only).  enable = 0
 * stops, -1 times every family, any other value is a bitmask of the
 * families to time (bit f = family f); each call clears the accumulators.
 * A timed launch carries its start/stop events in the dispatch itself
 * (hipExtLaunchKernel); timing does not make any call synchronous.
 * nrg_gpu_kernel_stats waits for the device and writes, for family
 * f < nfam, out[4f..4f+3] = {milliseconds, launches, algorithmic bytes,
 * algorithmic flops}.  Families: 0 H assembly, 1 tridiagonal column update,
 * 2 tridiagonal SYMV, 3 tridiagonal trailing GEMM, 4 D&C leaves, 5 D&C merge
 * set-up, 6 D&C secular equations, 7 D&C vectors, 8 D&C GEMM, 9 D&C
 * reorder, 10 back-transform small kernels, 11 back-transform GEMM,
 * 12 invariant GEMM, 13 spectral invariant GEMM, 14 CFS spectral products,
 * 15 CFS rho weighting GEMM, 16 CFS binning, 17 rho GEMM, 18 rho row
 * gather, 19 other. */
int nrg_gpu_profile(int enable);
int nrg_gpu_kernel_stats(double* out, int nfam);
/* Time only every `every`-th launch of each timed family (default 1: all);
 * the statistics then cover the sampled launches. */
int nrg_gpu_profile_sample(int every);
/* Measurement option with no counterpart in the reference: on != 0 (the
 * default) fences a timed SYMV launch of the multi-stream tridiagonalisation
 * against the other stream groups, so that its event interval is its
 * execution with the device to itself (each fence waits for the work the
 * host has already queued on the other streams, which slows the schedule);
 * 0 times it in place, sharing the device with the other groups' kernels as
 * an untimed launch does. */
int nrg_gpu_profile_exclusive(int on);
/* Launch log for matching profiler counters to algorithmic work: with
 * fam_mask != 0 every launch of a family in the mask is recorded (the call
 * clears the log; 0 stops).  nrg_gpu_launch_log_get writes up to cap records
 * out[7i..7i+6] = {family, threads per workgroup, kernel id, ordinal of the
 * launch among all launches of the same kernel function since this call,
 * algorithmic bytes, algorithmic flops, 1 if the launch carried timing
 * events (nrg_gpu_profile) else 0} and the record count to *count;
 * nrg_gpu_launch_log_kernel gives the (mangled) name of a kernel id. */
int nrg_gpu_launch_log(int fam_mask);
/* Work issued per kernel family (no counterpart in the reference), counted
 * for every launch whether timed or not: out[3f..3f+2] = {launches,
 * algorithmic bytes, algorithmic flops} since the last reset, for family
 * f < nfam (families as for nrg_gpu_kernel_stats; GEMM flops are 2mnk per
 * product, halved for lower-triangle-only products); reset != 0 clears the
 * counters after reading. */
int nrg_gpu_work_counters(double* out, int nfam, int reset);
int nrg_gpu_launch_log_get(double* out, int64_t cap, int64_t* count);
/* Measurement helper with no counterpart in the reference: copies nbytes
 * (a multiple of 16) device to device reps times with the library's own
 * streaming kernel (k_stream_copy: every byte read once and written once),
 * the known-size kernel that scales the rocprofv3 FETCH_SIZE / WRITE_SIZE
 * counters of the NRG kernels to bytes; *ok = 1 when the copy arrived. */
int nrg_gpu_copy_calibration(int64_t nbytes, int reps, int* ok);
int nrg_gpu_launch_log_kernel(int kid, char* name, int cap);

/* Algorithm option with no counterpart in the reference (it replaces the
 * tridiagonalisation inside dsyev, diagonalization.F90:266-269):
 * two_stage = 0 reduces each block to tridiagonal form in one stage
 * (Householder columns with a trailing SYMV, the default); 1 reduces it to a
 * band of width 32 by panel QRs and matrix-core updates, then the band to
 * tridiagonal form by bulge chasing, and back-transforms the requested
 * eigenvector columns through both stages. */
int nrg_gpu_set_tridiag_method(int two_stage);
/* Launch option with no counterpart in the reference: the number of stream
 * groups (1..4) the blocks of the tridiagonalisation are split over (each
 * block, in size order, to the group with the least sum n^3 so far), in
 * place of the tuner's choice; 0 returns the choice to the tuner (the
 * default, 4 groups, with tuning off). */
int nrg_gpu_set_tridiag_groups(int groups);
/* Kernel option with no counterpart in the reference: the one-stage
 * tridiagonalisation stops its column steps at the panel boundary from which
 * every block has at most `rows` rows left (0..192, default 192) and reduces
 * those last columns of all blocks in one launch, one workgroup per block,
 * with the trailing block in LDS; 0 runs the column steps to the end. */
int nrg_gpu_set_tridiag_finish(int rows);
/* Kernel option with no counterpart in the reference: the SYMV of the
 * one-stage tridiagonalisation adds the row parts of its columns with LDS
 * atomics (kind 0, the default) or sums them in registers and 256-row LDS
 * chunks in a fixed order (kind 1, bit-reproducible). */
int nrg_gpu_set_tridiag_symv(int kind);

/* Scheduling option with no counterpart in the reference: solve the blocks
 * of an iteration in groups of at most max_blocks (0 = all in one batched
 * pass, the default; 1 = one block at a time, the per-block schedule of the
 * reference's dsyev loop, used as the plain-FP64 comparison). */
int nrg_gpu_set_batch_limit(int max_blocks);

/* Launch option with no counterpart in the reference: on != 0 captures the
 * launch sequence of each batched tridiagonalisation (the column kernels,
 * SYMVs and trailing updates of all blocks on all stream groups) as a HIP
 * graph, keyed on the block set (sizes, offsets, schedule, buffers); a
 * later iteration with the same block set -- the NRG steady state repeats
 * its block tables -- replays the graph instead of issuing the launches
 * (at most 4 graphs are kept, least recently used replaced).  Off by
 * default; kernel timing (nrg_gpu_profile) and the launch log disable it. */
int nrg_gpu_set_graphs(int on);

/* Tuning option with no counterpart in the reference (the role the
 * reference's n^3 cost model plays in loadbalance.F90:102-186): the launch
 * configuration of the batched tridiagonalisation -- the number of stream
 * groups sharing the device and the launch size from which the SYMV uses
 * 32-column panels -- chosen per block-size histogram (block count, largest
 * block, sum n^3).  mode 0: defaults (4 groups, 512 workgroups); 1: time the
 * candidate configurations on successive iterations with the same
 * histogram (three times each), then keep the one with the lowest mean
 * time, updating the cache file; 2: use the cache file's choice only.  cache_path: a text file of measurements (NULL
 * or "" keeps them in memory). */
int nrg_gpu_set_tuning(int mode, const char* cache_path);

/* Scheduling option with no counterpart in the reference: the column order
 * of the batched tridiagonalisation.  align_end != 0 (the default) starts
 * the blocks of an iteration at staggered steps so that they finish
 * together; 0 starts them together. */
int nrg_gpu_set_tridiag_schedule(int align_end);

/* Test hook with no counterpart in the reference: make the next
 * nrg_gpu_diagonalize of `iteration` fail with status 5 (error-path tests). */
int nrg_gpu_inject_fault(int iteration);

#ifdef __cplusplus
}
#endif
#endif /* NRG_GPU_H */
