/*
 * dcaspt2_amd.h -- C ABI of the MI355X (gfx950) CASCI/CASPT2 hot path.
 *
 * Plain C: fixed-width integers, double pointers.  Complex arrays are
 * interleaved (re, im) doubles, i.e. the memory of Fortran complex*16 or C
 * double _Complex, column-major (Fortran order), 0-based in C / 1-based
 * labels where the reference uses them.  Every function returns 0 on success
 * or a non-zero status (1 bad argument/state, 2 HIP error, 3 rocSOLVER error,
 * >= 100: 100 + LAPACK-style info of a failed factorisation); the message is
 * available from dc2_last_error().  A Fortran host turns a non-zero status into
 * stop_with_errorcode(status) (reference: module_error.F90:15-39).
 *
 * The entry points map onto the reference's Fortran subroutines (paths relative
 * to the reference repository):
 *
 *   dc2_set_spaces / dc2_set_symmetry / dc2_set_ras / dc2_set_one_electron /
 *   dc2_set_integral_block   -- the module_global_variables state the reference's
 *       readers fill (module_global_variables.F90:55-183: ninact/nact/nsec/nelec,
 *       irpamo, MULTB_*, ras*_list, one_elec_int_r/i, inttwr/inttwi,
 *       int2r/i_f1, int2r/i_f2) and the subspace integral files A1int..Hint
 *       (divide_2_elec_integral_into_subspaces.F90:133-384)
 *   dc2_casci        -- casci (casci.F90:3-72): search_cas_configuration
 *                       (search_cas_configuration.F90:3-181) + casmat_complex
 *                       (casmat.F90:3-222) + cdiag (diag.F90:66-128)
 *   dc2_caspt2       -- r4dcaspt2_tra (r4dcaspt2_tra.F90:4-482): fockcasci_complex
 *                       (fockcasci.F90:4-90), fockdiag (fockdiag.F90:4-78) with
 *                       tracic (trac.F90:76-156), calce0 (calce0.F90:4-50),
 *                       intra_1/2/3 (module_intra.F90:19-51), tramo1
 *                       (tramo.F90:4-68) and solve_A..H_subspace
 *                       (solve_[A-H]_subspace.F90, with cutoff.F90:37-70 and
 *                       ulambda_s_half.F90:47-75)
 *   dc2_casmat       -- casmat_complex alone (CI matrix to host memory)
 *   dc2_cdiag        -- cdiag (diag.F90:66-128) on a host matrix, on the GPU
 *   dc2_zgemm        -- gemm (module_blas.F90:32-57) on device memory
 */
#ifndef DCASPT2_AMD_H
#define DCASPT2_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dc2_context dc2_context;

/* subspace indices used by the result arrays: A=0 ... H=7 */
#define DC2_NSUBSPACE 8

/* ---- context -------------------------------------------------------------- */
int dc2_context_create(int32_t device, dc2_context** out);
int dc2_context_destroy(dc2_context* ctx);
const char* dc2_last_error(void);
/* version string of the library build */
const char* dc2_version(void);

/* ---- problem definition (module_global_variables.F90) ---------------------- */
/* spinor spaces: ninact, nact, nsec, nelec (module_global_variables.F90:21-24) */
int dc2_set_spaces(dc2_context* ctx, int32_t ninact, int32_t nact, int32_t nsec, int32_t nelec);
/* double-group symmetry: nsymrpa, MULTB (2n x 2n, column-major, 1-based values),
 * irpamo (nmo entries, CAS order) and the target total symmetry totsym
 * (read_mrconee.F90:383-453).  Without this call the problem is C1. */
int dc2_set_symmetry(dc2_context* ctx, int32_t nsymrpa, const int64_t* multb, const int64_t* irpamo, int32_t totsym);
/* RAS restriction (read_input_module.F90:150-172; ras_det_check.F90) */
int dc2_set_ras(dc2_context* ctx, int32_t ras1_size, int32_t ras2_size, int32_t ras3_size, int32_t ras1_max_hole,
                int32_t ras3_max_elec, int32_t min_hole_ras1, int32_t min_elec_ras3);
/* one-electron integrals h (nmo x nmo complex, CAS order) and the core energy */
int dc2_set_one_electron(dc2_context* ctx, const double* h, double ecore);
/* Two-electron integral blocks in the original spinor basis, column-major over
 * (i,j,k,l) = (ij|kl).  name / index spaces (i inactive, a active, s secondary,
 * o = inactive+active):
 *   "OOOO" (o,o|o,o)  = inttwr/inttwi      (module_2integrals.F90:395-434)
 *   "F1"   (s,s|o,o)  = int2r/i_f1, "F2" (s,o|o,s) = int2r/i_f2
 *   "A1" aiaa "A2" aiii "B" aiai (optional: taken from OOOO when absent)
 *   "C1" saaa "C2" saii "C3" siia "D1" siaa "D2" saai "D3" siii
 *   "E" siai "F" sasa "G" sisa "H" sisi      (the A1int..Hint subspace files)
 * where: 0 = host memory (copied), 1 = device memory (copied),
 *        2 = device memory borrowed (must stay valid until the context is
 *            destroyed or the block replaced). */
int dc2_set_integral_block(dc2_context* ctx, const char* name, const double* data, int64_t count, int32_t where);

/* ---- CASCI (casci.F90:31-72) --------------------------------------------------
 * nroot energies eigen(k) = ecas(k) + ecore are written to eigen[0..nroot-1]
 * (fewer if ndet < nroot; *ndet_out receives ndet).  At least selectroot roots must be
 * requested before dc2_caspt2 (the Davidson solver computes only the requested ones). */
int dc2_casci(dc2_context* ctx, int32_t nroot, double* eigen, int64_t* ndet_out);
/* determinant bit strings (ndet int64) and the CI vector of a root (ndet complex) */
int dc2_get_dets(dc2_context* ctx, int64_t* dets);
int dc2_get_ci_vector(dc2_context* ctx, int32_t root, double* c);

/* ---- CASPT2 (r4dcaspt2_tra.F90:113-442) ------------------------------------------
 * selectroot is 1-based.  Outputs: e2[8], sumc2[8] per subspace A..H (0 when
 * skipped), scalars[6] = { e0, e2all, sumc2_all, weight0,
 * total second-order energy (e2all - eshift*sumc2), total energy }. */
int dc2_caspt2(dc2_context* ctx, int32_t selectroot, double eshift, double* e2, double* sumc2, double* scalars);
/* Per subspace diagnostics of the last dc2_caspt2: info[4*k..4*k+3] =
 * {dimn, dimm, npert, ran}, margin[k] = min |w_S - 1e-8| (distance of the
 * closest S eigenvalue to the linear-dependency threshold). */
int dc2_get_subspace_info(dc2_context* ctx, int64_t* info, double* margin);
/* orbital energies after fockdiag (nmo doubles) and the rotation U (nmo x nmo complex) */
int dc2_get_fock_rotation(dc2_context* ctx, double* eps, double* U);
/* rotated CI vector of the last dc2_caspt2 (ndet complex) */
int dc2_get_rotated_ci(dc2_context* ctx, double* c);
/* density tensors of the rotated state, 0:<E_ab> (na^2) 1:<E_ab F> 2:<E_ab E_cd> (na^4)
 * 3:<E_ab E_cd F>; layouts documented in dirac_caspt2_amd/csrc/dc2_pt2.hip */
int dc2_get_density(dc2_context* ctx, int32_t which, double* out, int64_t count);
/* a transformed subspace block of the last dc2_caspt2 (same names as above) */
int dc2_get_transformed_block(dc2_context* ctx, const char* name, double* out, int64_t count);
/* wall-clock seconds of the stages of the last dc2_casci/dc2_caspt2 call,
 * measured with HIP events on the library stream:
 * t[0] casci, t[1] fock+fockdiag+tracic, t[2] densities, t[3] transforms,
 * t[4..11] subspaces A..H (assembly + solves, transforms excluded) */
int dc2_get_timings(dc2_context* ctx, double* t, int32_t n);

/* Progress and checkpoints (no counterpart in the reference, whose restart files
 * module_restart_file.F90:55-172 reads are written by hand): cb(user, k, e2, sumc2) is called
 * from the stream thread that finished subspace k (0 = A .. 7 = H) of a dc2_caspt2 on this
 * context, as soon as its energy is known (subspaces run concurrently, so in any order; the
 * callback must be thread-safe and must not call into this context).  NULL removes it. */
int dc2_set_subspace_callback(dc2_context* ctx, void (*cb)(void* user, int32_t subspace, double e2, double sumc2),
                              void* user);
/* Writes the caspt2_restart file read_and_validate_restart_file reads (module_restart_file.F90:
 * 55-172): "totsym: N", "selectroot: N", "Next subspace: X" (X = 'A' + next; 'I' when next = 8)
 * and one line "Subspace: Y Sumc2: v Energy: v" for each subspace before X, values printed to
 * 17 significant digits; written to path.tmp, then renamed.  No GPU needed. */
int dc2_write_restart(const char* path, int32_t totsym, int32_t selectroot, int32_t next, const double* e2,
                      const double* sumc2);

/* ---- stand-alone pieces --------------------------------------------------------- */
/* CI matrix <i|H|j> (ndet x ndet complex, host) after dc2_casci (casmat.F90:3-222) */
int dc2_casmat(dc2_context* ctx, double* mat);
/* cdiag (diag.F90:66-128): Hermitian eigendecomposition of the host matrix c
 * (n x n complex, upper triangle read) on the GPU; c is overwritten by the
 * eigenvectors, w by the ascending eigenvalues, *dimm = count(w >= threshold)
 * (or n when threshold == 0) */
int dc2_cdiag(dc2_context* ctx, double* c, int32_t n, int32_t* dimm, double* w, double threshold);
/* C = alpha op(A) op(B) + beta C on device memory, batched by strides;
 * opA in 'N','T','C'; opB in 'N','R' (conjugate, no transpose),'T','C' */
int dc2_zgemm(dc2_context* ctx, char opA, char opB, int64_t M, int64_t N, int64_t K, const double* alpha,
              const double* A, int64_t lda, int64_t strideA, const double* B, int64_t ldb, int64_t strideB,
              const double* beta, double* C, int64_t ldc, int64_t strideC, int64_t batch);
/* C (m x m, ldc) = A A^H for A (m x k, lda) on device memory, the whole Hermitian matrix
 * (no counterpart in the reference, which forms the A/C metrics element by element: the
 * Gram matrix of the determinant-space metric route).  split: 0 chosen by size, 1 unsplit,
 * 2..64 that many chunks of k summed in a fixed order. */
int dc2_zherk(dc2_context* ctx, int64_t m, int64_t k, const double* A, int64_t lda, double* C, int64_t ldc,
              int32_t split);
/* wait for all work queued on the library stream */
int dc2_synchronize(dc2_context* ctx);
/* GEMM launch accounting with HIP events on the library stream (resets on each call):
 * dc2_get_gemm_stats(tag) returns out[4] = {launches, seconds, algorithmic flops
 * (8*M*N*K per complex GEMM), largest single launch flops} of the launches with that
 * tag (-1: all): 1 integral transforms, 2 density contractions, 3 V contractions,
 * 4 S^-1/2 projections, 5 integral blocks from factors, 0 other */
int dc2_set_profiling(dc2_context* ctx, int32_t on);
int dc2_get_gemm_stats(dc2_context* ctx, int32_t tag, double* out);
/* GEMM autotuning (no counterpart in the reference: module_blas.F90 calls the vendor BLAS).
 * dc2_get_gemm_shapes: the distinct shapes of the MFMA-kernel launches recorded while
 *   profiling was on, rows of 6 int64 {opA, opB, M, N, K, batch} (opA/opB as character
 *   codes), at most cap rows written, *count = all of them.
 * dc2_gemm_tune: times every candidate for one shape on scratch operands on the context's
 *   stream (kernel variant 1..4: K tile 16/32 x LDS padding 1/16; split-K 1, 2, 4 .. 32 for
 *   single products with few tiles), installs the fastest in the process-wide table when it
 *   beats the built-in choice; out[4] = {variant, split (0, 0: built-in kept), its seconds,
 *   the built-in choice's seconds}.  reps: timed repetitions per candidate.
 * dc2_gemm_set_choice: installs a choice (a table saved from an earlier tuning run).
 * dc2_gemm_clear_choices: empties the table (*count_before may be NULL).
 * Choices apply to products of exactly that shape; results change only by the order of
 * the split-K summation (still fixed, so repeatable). */
int dc2_get_gemm_shapes(dc2_context* ctx, int64_t* out, int64_t cap, int64_t* count);
/* Captured graphs of the block Davidson (no counterpart in the reference): the launches
 * between two host round trips of an iteration are captured once per size and buffer set
 * and replayed (option "graphs", default 1).  out[4] = {captures, replays, graphs held,
 * graphs still enabled (a failed capture turns them off for the context)} */
int dc2_get_graph_stats(dc2_context* ctx, int64_t* out);
int dc2_gemm_tune(dc2_context* ctx, char opA, char opB, int64_t M, int64_t N, int64_t K, int64_t batch, int32_t reps,
                  double* out);
int dc2_gemm_set_choice(char opA, char opB, int64_t M, int64_t N, int64_t K, int64_t batch, int32_t variant,
                        int32_t split);
int dc2_gemm_clear_choices(int64_t* count_before);
/* the same figures for the launches of one MFMA kernel instantiation k_zgemm<opA, opB>
 * (products routed to the kernels for skinny shapes are counted by dc2_get_gemm_stats only) */
int dc2_get_gemm_op_stats(dc2_context* ctx, char opA, char opB, double* out);
/* Numerical options (no counterpart in the reference):
 *   "eigensolver": -1 automatic (default: matrices up to 32, and Fock blocks up to 64, on the
 *                  host, 4 for 33 <= n <= native_max outside the batched server, rocSOLVER
 *                  divide and conquer otherwise, with Jacobi leaves for rank-deficient S
 *                  matrices); any other value puts every eigenproblem on the device: 0 tridiagonal
 *                  QR (the ZHEEV algorithm), 1 divide and conquer, 2 D&C with Jacobi leaves,
 *                  3 Jacobi, 4 this library's Householder tridiagonalisation (one persistent
 *                  kernel) + its tridiagonal divide and conquer (or rocSOLVER dstedc, option
 *                  tridiag_solver) + its back-transformation (n <= 2048; D&C otherwise);
 *   "ac_metric":   0 automatic (default: the smallest exact basis; for A the tuples t < u
 *                  through the determinant-space factor (4) when they are fewer than the
 *                  (nelec+1)-electron strings), 1 the RDM-formula S/B matrices of size
 *                  nact^3, 2 the determinant-space factorisation of the A and C metrics, 3 for
 *                  A the formula matrices on the tuples t < u (a+_t a+_u is antisymmetric, so
 *                  these span the same space and S's eigenvalues halve), 4 for A the same
 *                  reduced matrices as products of the determinant-space factor (S = Phi^H
 *                  Phi, B = Phi^H F Phi on the columns t < u; no F-weighted 3-RDM), for C the
 *                  determinant-space factorisation;
 *   "lanes":       1..4 concurrent streams (each with a host thread) for the subspace
 *                  solves (default 4: A, C, D and the rest; 3: A, C and the rest);
 *   "lane_plan":   with 4 lanes, -1 (default) B and F on D's lane from 256 secondary spinors,
 *                  0 never, 1 always;
 *   "early_v":     1 (default) with 4 lanes, V of A and C formed on D's lane as soon as their
 *                  transformed blocks exist (the A and C lanes use it after their eigensolves),
 *                  0 on their own lanes;
 *   "c_thin":      1 (default) the determinant-space C metric through its thin factor (an
 *                  eigenproblem of nact^2 (nact-1)/2 instead of C(nact, nelec-1)) when that is
 *                  smaller and more than 32 directions are cut anyway, 0 never;
 *   "graphs":      1 (default) block-Davidson iterations replayed from captured graphs
 *                  (dc2_get_graph_stats), 0 every launch issued directly;
 *   "schur":       1 (default) E2 of the determinant-space route through the compressed
 *                  resolvent when at most 32 directions are cut, 0 always through the
 *                  eigendecomposition of B~;
 *   "split_slices": 0 (default) FP64 integral transforms, 3..8 split-precision passes
 *                  (dc2_zgemm_split) with this many slices;
 *   "split_min_k": smallest contracted dimension of a split-precision pass (default 32);
 *   "davidson_maxsub": vectors of the block Davidson's subspace before a thick restart
 *                  (0: automatic, max(6 b, 24) for a block of b vectors, 12 b from 1500
 *                  determinants);
 *   "davidson_block": vectors per block (0: automatic, max(nroot + 3, 4); at 3003
 *                  determinants 4 with 48 subspace vectors was fastest, 13.1 ms against 14.5
 *                  to 26 ms with blocks of 6 to 12);
 *   "native_max":  largest n this library's eigensolver takes (default 2048, at most 2048);
 *                  rocSOLVER's divide and conquer above;
 *   "native_min":  n >= this (and n within this library's eigensolver, 3..2048) uses this
 *                  library's eigensolver even where the eigensolves are batched across served
 *                  requests (default 0: batched rocSOLVER for every size there);
 *   "tridiag_solver": the tridiagonal stage of this library's eigensolver: 0 (default) its
 *                  own divide and conquer on the device, 1 rocSOLVER dstedc;
 *   "casci_solver": 0 automatic (default: block Davidson for ndet >= 300 and nroot <= 16,
 *                  the dense eigensolver otherwise), 1 dense (all roots, as cdiag), 2 Davidson
 *                  (the nroot lowest roots; falls back to dense when it does not converge). */
int dc2_set_option(dc2_context* ctx, const char* key, int64_t value);
/* Split-precision complex GEMM (Ozaki scheme on INT8 matrix cores; no counterpart in
 * the reference): C[m + ldc*j] = sum_k A[k + lda*m] * op(B)[k + ldb*j] on device memory,
 * op(B) = B ('N') or conj(B) ('R'), every row of A and column of B cut into `slices`
 * 7-bit digits (3..8) after power-of-two scaling; K <= 4096.  The integral transforms
 * use it for passes with K >= "split_min_k" when the option "split_slices" is non-zero. */
int dc2_zgemm_split(dc2_context* ctx, int32_t slices, char opB, int64_t M, int64_t N, int64_t K, const double* A,
                    int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc);
/* The same product as dc2_zgemm_split, allocating nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  The INT8 slices and exponents live in a fixed 64 MiB device array of the library
 * (held on every device it runs on until the process ends); a product needing more,
 * slices * (M + 2N) * 2 * ceil32(K) + 4 (M + N) bytes rounded up to 256, is refused with
 * status 1.  Asynchronous on the context's stream (dc2_synchronize before reading C), no
 * hipMalloc/hipFree and no host synchronisation.  Every call on a device shares the array:
 * calls on one context are ordered by its stream; calls that overlap on one device race. */
int dc2_zgemm_split_noalloc(dc2_context* ctx, int32_t slices, char opB, int64_t M, int64_t N, int64_t K,
                            const double* A, int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc);
/* transform flops of the last dc2_caspt2: out[0] in split precision, out[1] in FP64 */
int dc2_get_transform_flops(dc2_context* ctx, double* out);
/* Measurement, no counterpart in the reference: out[2] = {calls, summed device seconds} of
 * this library's tridiagonalisation kernel for n x n matrices in this process (HIP events
 * around each launch); n = 0 resets the counts. */
int dc2_get_hetrd_stats(int32_t n, double* out);
/* out[3] = {ndet, roots held by the last dc2_casci, Davidson iterations (0: dense solver)} */
int dc2_get_casci_info(dc2_context* ctx, int64_t* out);
/* Hermitian eigendecomposition on the host (the Davidson subspace solver): a (n x n
 * complex, column-major) is overwritten by the eigenvectors, w by the ascending eigenvalues.
 * method 0: Householder tridiagonalisation + implicit QL, Jacobi if QL does not converge;
 * 1: cyclic Jacobi; 2: QL only (status 3 if it does not converge); 3: the sequence of
 * LAPACK's ZHEEV('V','U') (zhetd2, zungtr, zsteqr), whose eigenvector signs the reference's
 * IVO coefficients carry (diag.F90:66-128), only the upper triangle read; 4: every eigenvalue
 * and the eigenvectors of the min(n, 4) lowest only, the other columns of a set to zero (what
 * the Davidson solver uses on its Rayleigh matrices: the lower triangle reduced, QL, the
 * rotations applied to the selected unit vectors alone).  No GPU needed. */
int dc2_host_heev(int32_t n, double* a, double* w, int32_t method);
/* fockivo (fockivo.F90:40-230, driven by r4divo.F90:60-80): the IVO Fock matrix of the
 * secondary spinors, F_ab = delta_ab eps_a + sum_k [(ak|kb) - (ab|kk)] / numh over the numh
 * highest occupied spinors k of the inactive+active space (a factor 1/2 on the last two for odd
 * nelec), built on the device from the resident F1/F2 blocks, then diagonalised per irreducible
 * representation irreps[0..nirr-1] (values of irpamo; every spinor is irrep 1 without
 * dc2_set_symmetry) over the secondary spinors of that irrep, by the sequence of LAPACK's ZHEEV
 * (the reference's cdiag) on the host: these blocks are small and the IVO coefficients carry
 * ZHEEV's eigenvector signs.  numh <= 0: the number of spinors within 0.1 Eh of the HOMO energy
 * mo_energy[ninact + nelec - 1].  mo_energy: nmo orbital energies (CAS order, caspt2_mo_energy).
 * Outputs: fock (nsec x nsec complex, NULL to skip), w (nsec: the eigenvalue of eigenvector j of
 * an irrep's block at the position of that block's j-th spinor), U (nsec x nsec complex, block
 * diagonal: U[mos[i] + nsec * mos[j]] = component i of eigenvector j, mos the irrep's spinors
 * in order; zero outside the blocks of the irreps listed). */
int dc2_fockivo(dc2_context* ctx, int32_t numh, const double* mo_energy, int32_t nirr, const int32_t* irreps,
                double* fock, double* w, double* U);
/* number of eigensolves that fell back to the tridiagonal QR algorithm */
int dc2_get_eig_fallbacks(dc2_context* ctx, int64_t* n);

/* ---- device memory for hosts without their own GPU runtime (e.g. Fortran) ----------
 * kind: 1 host->device, 2 device->host, 3 device->device (synchronous) */
int dc2_malloc(dc2_context* ctx, int64_t bytes, void** ptr);
int dc2_free(dc2_context* ctx, void* ptr);
int dc2_memcpy(dc2_context* ctx, void* dst, const void* src, int64_t bytes, int32_t kind);
/* Input preparation for factorised integrals (pq|rs) = sum_P B[P,p,q] B[P,r,s]:
 * B is (naux, nmo, nmo) complex, column-major (P fastest); every block the path
 * needs is formed on the device (no counterpart in the reference, which reads
 * the four-index MDCINT files; this feeds the same blocks without a host round
 * trip).  where: 0 host, 1 device. */
int dc2_set_integral_factors(dc2_context* ctx, const double* B, int64_t naux, int32_t where);

/* CASCI result computed elsewhere, e.g. read from a CIDATA_sym<totsym> file (casci.F90:117-160,
 * read_cidata.F90): ndet determinants (bit strings of occupied active spinors, the values of
 * dict_cas_idx, which must equal this problem's determinants in order), nroot CI vectors
 * (complex, ndet x nroot, column-major) and the CASCI energies ecas without ecore.  Replaces
 * dc2_casci before dc2_caspt2. */
int dc2_set_casci_result(dc2_context* ctx, int64_t ndet, int32_t nroot, const int64_t* dets, const double* coeffs,
                         const double* ecas);
/* dc2_caspt2 resumed from a caspt2_restart_<totsym>_<root> file (module_restart_file.F90:55-172):
 * subspaces before `first` (0 = A .. 7 = H, 8 = none computed) are skipped, e2_file[8] and
 * sumc2_file[8] hold the file's values (zero where it lists nothing).  As the reference
 * (r4dcaspt2_tra.F90:168, solve_X_subspace), e2all and sumc2 start from the sums of the file's
 * values and every computed subspace adds its own; e2[k] and sumc2[k] of a skipped subspace
 * are the file's. */
int dc2_caspt2_restart(dc2_context* ctx, int32_t selectroot, double eshift, int32_t first, const double* e2_file,
                       const double* sumc2_file, double* e2_subspace, double* sumc2_subspace, double* scalars);
/* Integral ingest on the GPU: n MDCINTNEW-equivalent records (create_newmdcint.F90:173-176),
 * ijkl[4n] = 1-based CAS-order spinor labels (i, j, k, l) of (ij|kl), values[2n] = (re, im)
 * as written there.  Every record's 16 symmetry images (module_2integrals.F90:395-434) are
 * scattered on the device into every block of dc2_set_integral_block (OOOO, F1, F2, A1 ... H)
 * whose space pattern they match; where two records reach the same element the later one
 * wins.  This replaces the dense fill and the subspace files of
 * divide_2_elec_integral_into_subspaces.F90.  where: 0 host arrays, 1 device arrays. */
int dc2_set_integral_records(dc2_context* ctx, int64_t n, const int32_t* ijkl, const double* values, int32_t where);
/* DIRAC's unique integral set -> MDCINTNEW-equivalent records on the GPU (create_newmdcint.F90:
 * 86-177, the step before dc2_set_integral_records; the reference writes the records to the
 * MDCINTNEW file, here they stay on the device).  Input as create_newmdcint reads it
 * (read_mdcint_header, read_one_line_of_mdcint, create_newmdcint.F90:209-280): the nrec
 * records (ikr, jkr, nz) of MDCINT, MDCINXXXX1, ... in file order (the ikr = 0 record that
 * ends a file excluded), their entries concatenated: indk[], indl[], rklr[] and rkli[] (NULL:
 * real integrals, realonly) of sum(nz) elements; kr[2 nkr + 1] the header's kr(-nkr:nkr)
 * (kr[nkr + m] = DIRAC spinor of Kramers label m); d2c[nd2c] = indmo_dirac_to_cas (d2c[k] the
 * 1-based CAS index of DIRAC spinor k, read_mrconee.F90:500-541); scheme the MDCINT SCHEME
 * of active.inp (6: no amplification).  Every record and its Kramers image are expanded
 * (and particle-exchanged when the SCHEME amplifies), relabelled, and kept under the
 * reference's cutoff (|v| > 0.25e-12) and canonical-order conditions, in the order the
 * reference writes them.  *nout = number of records; when ijkl/values are not NULL (both or
 * neither) and cap >= *nout, the records (1-based CAS labels, (re, im) as written to
 * MDCINTNEW) are copied there (cap too small: status 1, *nout set); ingest != 0: the records
 * also replace every integral block as dc2_set_integral_records does.  Labels outside kr /
 * d2c: status 1. */
int dc2_mdcint_records(dc2_context* ctx, int64_t nrec, const int64_t* ikr, const int64_t* jkr, const int64_t* nz,
                       const int64_t* indk, const int64_t* indl, const double* rklr, const double* rkli, int64_t nkr,
                       const int64_t* kr, int64_t nd2c, const int64_t* d2c, int32_t scheme, int64_t cap,
                       int64_t* nout, int32_t* ijkl, double* values, int32_t ingest);
/* device pointer and element count of a resident integral block, e.g. to lend it to
 * another context with dc2_set_integral_block(..., where = 2) */
int dc2_get_integral_block_device(dc2_context* ctx, const char* name, const double** ptr, int64_t* count);

/* ---- request server: continuous batching on one GPU (no counterpart in the reference,
 * whose driver runs one problem per process, dcaspt2_run_subprograms.F90:57-62) ------
 * nslots (1..256) slots, each with its own context and host thread (lanes_per_slot, 1..4,
 * concurrent subspace streams per request).  Molecules are registered once (factorised integrals,
 * as dc2_set_integral_factors; the blocks stay resident in HBM and are shared read-only
 * by all requests on the molecule).  A request is one CASCI + CASPT2 energy for a
 * one-electron matrix h (nmo x nmo complex, copied at submit), root and eshift, with
 * an optional RAS restriction ras[7] (the arguments of dc2_set_ras, or NULL).  A slot's
 * context keeps its scratch between requests, so each slot holds the largest footprint
 * (dc2_request_footprint) it has run; a free slot admits the oldest pending request once
 * the memory all slots hold plus the growth the request needs in this slot fits mem_budget
 * bytes, and when nothing runs and it still does not fit, idle slots release their memory.
 * A request larger than the budget is refused at submit.  flags: 0 FP64 integral transforms, 3..8 split-
 * precision transforms with that many slices (the option "split_slices"). */
typedef struct dc2_server dc2_server;
int dc2_request_footprint(int32_t ninact, int32_t nact, int32_t nsec, int32_t nelec, int64_t* bytes);
int dc2_server_create(int32_t device, int32_t nslots, int64_t mem_budget, int32_t lanes_per_slot, dc2_server** out);
/* the same with cross-request eigensolve batching: the Hermitian eigenproblems of the
 * running requests are grouped by size and solved eig_batch at a time by one batched
 * rocSOLVER call (eig_batch <= 256); a group is flushed when every running request waits
 * for it, when it is full, or eig_window_s seconds after its oldest problem (eig_batch <= 1:
 * off) */
int dc2_server_create_batched(int32_t device, int32_t nslots, int64_t mem_budget, int32_t lanes_per_slot,
                              int32_t eig_batch, double eig_window_s, dc2_server** out);
/* finishes every pending request, then stops the slots and frees the molecules */
int dc2_server_destroy(dc2_server* srv);
int dc2_server_add_molecule(dc2_server* srv, int32_t ninact, int32_t nact, int32_t nsec, int32_t nelec,
                            const double* B, int64_t naux, int32_t* mol_id);
int dc2_server_submit(dc2_server* srv, int32_t mol_id, const double* h, double ecore, int32_t selectroot,
                      double eshift, const int32_t* ras, int32_t flags, int64_t* req_id);
/* next completed request, waiting at most timeout_s seconds (*req_id = 0 on timeout):
 * out[0..7] e2, out[8..15] sumc2, out[16..21] the scalars of dc2_caspt2,
 * out[22..24] submit/start/end times (s since server creation), out[25] slot;
 * *req_status is the request's status (its message via dc2_last_error) */
int dc2_server_wait(dc2_server* srv, double timeout_s, int64_t* req_id, int32_t* req_status, double* out);
/* Per-subspace checkpoints of the served requests: with a directory set, every request writes
 * <dir>/<request id>/caspt2_restart_1_<selectroot> (dc2_write_restart) whenever the subspaces
 * finished so far extend the completed prefix A, B, ... (the reference resumes from a prefix), so
 * an interrupted request can be resubmitted with dc2_server_submit_restart.  NULL or "": off. */
int dc2_server_set_checkpoint_dir(dc2_server* srv, const char* dir);
/* dc2_server_submit resumed from a restart: subspaces before `first` (0 = A .. 8 = none) are taken
 * from e2_file[8] / sumc2_file[8], as dc2_caspt2_restart does. */
int dc2_server_submit_restart(dc2_server* srv, int32_t mol_id, const double* h, double ecore, int32_t selectroot,
                              double eshift, const int32_t* ras, int32_t flags, int32_t first, const double* e2_file,
                              const double* sumc2_file, int64_t* req_id);
/* dc2_set_option on every slot's context, now and for contexts the slots create later; call it
 * before the first dc2_server_submit */
int dc2_server_set_option(dc2_server* srv, const char* key, int64_t value);
/* out[7] = {pending, running, completed, footprint bytes the slots hold, its peak,
 *           eigensolve batches, eigenproblems solved in them} */
int dc2_server_stats(dc2_server* srv, int64_t* out);

#ifdef __cplusplus
}
#endif

#endif /* DCASPT2_AMD_H */
