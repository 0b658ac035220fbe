// Launchers of the CASCI / CASPT2 kernels (defined in dc2_ci.hip and dc2_pt2.hip).
#pragma once

#include "dc2_common.hpp"

#include <rocblas/rocblas.h>

namespace dc2 {

// ---------------------------------------------------------------- dc2_ci.hip
struct DetParams {
  int nact, nelec, ninact;
  int ras1, ras2, ras3, ras1_max_hole, ras3_max_elec, min_hole_ras1, min_elec_ras3;
  int nsym, totsym;
  const int* irp_act;   // device, nact entries (1-based fermion irreps of the active spinors)
  const int* multb_d;   // device, nsym x nsym (MULTB_D, column-major)
  const int* multb_ds;  // device, nsym x nsym (MULTB_DS, column-major)
};

// Enumerate all nelec-in-nact strings (full) and the CAS/RAS/symmetry-selected
// subset, both in ascending bit-string order (the Gosper order of
// search_cas_configuration.F90).  Returns counts through the host pointers.
void enumerate_dets(hipStream_t s, const DetParams& p, int* flag_full, int* flag_cas, int* rank_full,
                    int* rank_cas, int64_t* full_dets, int64_t* cas_dets, int* cas2full, int64_t* n_full,
                    int64_t* n_cas, int* counts_dev);
// tab[(p + na*q)*nf + j] = (rank(E_pq |j>) << 1) | odd_phase, or -1 when E_pq|j> = 0
void build_exc_table(hipStream_t s, int na, int64_t nf, const int64_t* full_dets, const int* rank_full, int* tab);
// Y[j + nf*((p + na*q) + na*na*x)] = (E_pq X_x)_j for x < nx
void apply_E(hipStream_t s, int na, int64_t nf, int64_t nx, const int* tab, const cplx* X, cplx* Y);
// c_full = embed(c_cas) (zeros elsewhere), cF = f * c_full with f_j = sum_w eps_w n_w(j)
void embed_ci(hipStream_t s, int na, int64_t nf, int64_t ncas, const int* cas2full, const cplx* c_cas,
              const int64_t* full_dets, const double* eps_act, cplx* c_full, cplx* cF);
// CASCI matrix (casmat.F90:3-222)
void casmat(hipStream_t s, int ni, int na, int ne, int64_t ndet, const int64_t* cas_dets, const int* rank_cas,
            const cplx* O, int nmoc, const cplx* h, int ldh, cplx* mat);
// Fock matrix, occupied and secondary diagonal blocks (fockcasci.F90:4-90); upper
// triangles written as computed, lower triangles as conjugates
// fockivo.F90:46-78: IVO secondary-space Fock matrix (ns x ns, both triangles) from F1, F2 and
// the secondary orbital energies; the numh highest occupied spinors of the nmoc are removed
void fockivo_build(hipStream_t s, int nmoc, int ns, int numh, bool odd_nelec, const double* eps_sec, const cplx* F1,
                   const cplx* F2, cplx* fock);
void fock_build(hipStream_t s, int ni, int na, int ns, const cplx* h, int ldh, const cplx* O, const cplx* F1,
                const cplx* F2, const cplx* D1, cplx* fock);
// gather a (space, irrep) block: out(i,j) = fock(ind_i, ind_j) (upper values)
void gather_block(hipStream_t s, int n, const int* ind, const cplx* fock, int ldf, cplx* out);
void scatter_block(hipStream_t s, int n, const int* ind, const cplx* V, const double* w, cplx* U, int ldu,
                   double* eps);
// canonical eigenvector gauge (see canonical_gauge in the oracle); scratch >= n*n complex,
// starts >= n + 2 ints
void canonical_gauge(hipStream_t s, int n, const double* w, cplx* V, cplx* scratch, int* starts);
// ds(k,kt) = det(Uact[occ_k, occ_kt]) with partial pivoting (trac.F90:105-131, sdet.F90:87-108)
void tracic_ds(hipStream_t s, int na, int ne, int64_t ndet, const int64_t* cas_dets, const cplx* Uact, cplx* ds);
// e0 = sum_w eps_w Re D1(w,w) (calce0.F90)
void calce0(hipStream_t s, int na, const double* eps_act, const cplx* D1, double* e0);
// out[perm] generic 6-index permutation of a column-major tensor
void permute6(hipStream_t s, const int* dims_out, const int* src_stride_of_out_axis, const cplx* in, cplx* out,
              int64_t total);
// copy a 4-index sub-block of a column-major tensor
void extract4(hipStream_t s, const cplx* in, const int64_t* ld_in, const int* off, const int* n, cplx* out);
void fill_zero(hipStream_t s, cplx* p, int64_t n);
// true when none of the n doubles at p is NaN or infinite (a device reduction into flag_dev,
// then a 4-byte copy to flag_host, page-locked host memory); synchronises the stream
bool all_finite_dev(hipStream_t s, const double* p, int64_t n, int* flag_dev, int* flag_host);
// uc[:, c] = vs[:, n-dimm+c] * ws[n-dimm+c]^(-1/2)  (or ^-1 with inverse)
void scale_cols_tail(hipStream_t s, int64_t n, int64_t dimm, const cplx* vs, const double* ws, cplx* uc,
                     bool inverse = false);
int64_t binom_host(int n, int k);
// Determinant-space factor of the A (create) / C metric: the (ne+-1)-electron strings
// (ascending), Phi[r + m*(t*na*na + u*na + v)] = <r|a+_t E_uv|0> or <r|a_t E_uv|0>, so that
// S = Phi^H Phi and B = Phi^H diag(fdiag) Phi with fdiag[r] = sum_w eps_w n_w(r)
void build_phi(hipStream_t s, int na, int ne, bool create, int64_t nf, const int* rank_full, const cplx* Y1,
               const double* eps_act, int64_t m, int64_t* strings, cplx* Phi, double* fdiag);
// Thin factor of the C metric G = Phi Phi^H (Phi of build_phi, m x na^3, ne electrons in the
// CASCI state): PsiH (nred x m, nred = na*na*(na-1)/2) with PsiH^H PsiH = G; Av: m x na scratch
void build_thin_c(hipStream_t s, int na, int ne, int64_t m, const cplx* Phi, cplx* Av, cplx* PsiH);
// W = (M M^T)^(-1/2) M V (nred x npert) for V (na^3 x npert) with Phi = Psi M of build_thin_c:
// with T t = l t (T = PsiH PsiH^H), u = M^T (M M^T)^(-1/2) t are the orthonormal eigenvectors
// of the tuple-space metric Phi^H Phi, so (U L^-1/2)^H V = L^-1/2 t^H W.  MV: nred x npert scratch
void thin_c_w(hipStream_t s, int na, int ne, int64_t npert, const cplx* V, cplx* MV, cplx* W);
// The antisymmetric reduction of the A factor in adjoint layout: the (ne+1)-electron strings,
// Psi (nred x m) with Psi[k + nred*r] = conj(<r|a+_t E_uv|0>) for the tuple tup[k] = t*na*na +
// u*na + v (t < u, nred = na*na*(na-1)/2 of them) and Psif = Psi diag(f), so that the reduced
// S = Psi Psi^H and B = Psif Psi^H
void build_phi_adj(hipStream_t s, int na, int ne, int64_t nf, const int* rank_full, const cplx* Y1,
                   const double* eps_act, int64_t m, int64_t* strings, int64_t nred, const int* tup, cplx* Psi,
                   cplx* Psif);
void scale_rows(hipStream_t s, int64_t m, int64_t ncol, const double* f, const cplx* in, cplx* out);
// block Davidson helpers: diagonal of H, unit start vectors e_rows[j], Ritz residuals
// R = AX - X diag(theta), diagonal preconditioner T[:, c] = R[:, sel_c] / (theta_sel_c - d)
void davidson_diag(hipStream_t s, int64_t n, const cplx* H, double* d);
void davidson_unit(hipStream_t s, int64_t n, int ncol, const int* rows, cplx* V);
void davidson_residual(hipStream_t s, int64_t n, int ncol, const cplx* AX, const cplx* X, const double* theta,
                       cplx* R);
void davidson_precond(hipStream_t s, int64_t n, int ncol, const int* sel, const cplx* R, const double* theta,
                      const double* d, cplx* T);
void davidson_normalize(hipStream_t s, int64_t n, int ncol, cplx* T);
// the same four steps in one launch for a subspace of mm <= 64 vectors: X = V Y, AX = W Y,
// R = AX - X diag(theta), T[:, c] = R[:, c] / (theta_c - d) (Y = Sy, mm x ncol)
void davidson_ritz(hipStream_t s, int64_t n, int ncol, int mm, const cplx* V, const cplx* W, const cplx* Sy,
                   const double* theta, const double* d, cplx* X, cplx* AX, cplx* R, cplx* T);
// orthonormal combinations Sc (4 x 4, kept columns first, then zeros) of a correction block of
// four columns from its Gram matrix G (4 x 4); the kept count in nk[0]
void davidson_block_orth(hipStream_t s, const cplx* G, cplx* Sc, int* nk);

// ---------------------------------------------------------------- dc2_ozaki.hip
// Split-precision (Ozaki, INT8 MFMA) complex GEMM C = A^T-view * op(B) of the transform
// passes: A[m][k] = A[k + lda*m], op(B) = B ('N') or conj(B) ('R'); slices in 3..8, K <= 4096.
// work: zgemm_ozaki_work_bytes() bytes (16-byte aligned), expo: M + N ints.
int64_t zgemm_ozaki_work_bytes(int slices, int64_t M, int64_t N, int64_t K);
void zgemm_ozaki(hipStream_t s, int slices, char opB, int64_t M, int64_t N, int64_t K, const cplx* A, int64_t lda,
                 const cplx* B, int64_t ldb, cplx* C, int64_t ldc, int8_t* work, int* expo);

// ---------------------------------------------------------------- dc2_ingest.hip
// Expand MDCINTNEW-equivalent records (1-based CAS-order i, j, k, l and (re, im)) into the
// dense blocks: spaces[b][4] in {0 inactive, 1 active, 2 secondary, 3 inactive+active},
// dims[b][4] the block extents, owner[b] a scratch key array of counts[b] elements.
// MDCINT records -> MDCINTNEW-equivalent records (create_newmdcint.F90:86-177) on the device:
// nrec records (ikr, jkr, nz) with entries at eoff[r] .. eoff[r] + nz - 1 of indk/indl/rr/ri
// (ri null: real), coff[r] the record's first candidate (coff[nrec] = ncand, candidates per
// record 2 nz, 4 nz when the SCHEME amplifies), cas[nkr + m] the CAS index of Kramers label m
// (indmo_dirac_to_cas of the header's kr(m)).  Writes the kept records to ijkl[4 .] / val[2 .]
// (room for ncand) in the reference's order and their count to scratch[cdiv(ncand, 256) * 2];
// scratch: mdcint_convert_scratch(ncand) bytes.  Labels must be within cas.
int64_t mdcint_convert_scratch(int64_t ncand);
void mdcint_convert(hipStream_t s, int64_t nrec, const int64_t* ikr, const int64_t* jkr, const int64_t* nz,
                    const int64_t* eoff, const int64_t* coff, const int* indk, const int* indl,
                    const double* rr, const double* ri, int64_t nkr, const int* cas, int scheme, int64_t ncand,
                    int64_t* scratch, int* ijkl, double* val);
void ingest_records(hipStream_t s, int ni, int na, int ns, int nblocks, const int* const* spaces, const int* const* dims,
                    cplx* const* data, unsigned long long* const* owner, const int64_t* counts, int64_t nrec,
                    const int* ijkl, const double* val);

// ---------------------------------------------------------------- dc2_heev.hip
// Hermitian eigendecomposition of A (n x n, 3 <= n <= 2048, the matrix its upper triangle
// defines): eigenvectors over A, ascending eigenvalues in w; the tridiagonal solver's info in
// info_dev (own_tridiag: this library's divide and conquer, else rocSOLVER dstedc).  false:
// nothing usable was computed (the caller falls back to rocSOLVER).  work: at least
// heev_native_work_bytes(n) bytes of device memory; flags_host: two ints of page-locked host
// memory for the reduction's status.  narrow: fewer workgroups for the reduction (n in
// 1025..1280) at some speed, for calls that run beside other work.
int64_t heev_native_work_bytes(int n);
bool heev_native(hipStream_t s, rocblas_handle h, int n, cplx* A, int64_t lda, double* w, void* work, int* info_dev,
                 int* flags_host, bool own_tridiag = true, bool report = false, bool narrow = false);

// ---------------------------------------------------------------- dc2_tridc.hip
// Eigenpairs of the real symmetric tridiagonal (d in w, off-diagonal e[0 .. n-2]), 1 <= n <= 2048,
// by divide and conquer on the device: ascending eigenvalues over w, eigenvectors in Z
// (n x n, leading dimension n), info_dev[0] = 0, or 1 / 2 if a leaf's QL iteration / a secular
// root did not converge (info_dev[1]: roots that did not; report: print the first few).
// Asynchronous on s; work: tri_dc_work_bytes(n) bytes of device memory.
int64_t tri_dc_work_bytes(int n);
void tri_dc(hipStream_t s, int n, double* w, const double* e, double* Z, void* work, int* info_dev,
            bool report = false);
// {calls, summed device seconds} of the tridiagonalisation kernel for size n in this process
void hetrd_stats(int n, double* out);
void hetrd_stats_reset();

// ---------------------------------------------------------------- dc2_pt2.hip
struct Dens {
  int na;
  const cplx *G1, *G1F, *G2, *G2F, *G3, *G3F;  // see dc2_pt2.hip for the index conventions
};
// map (optional, increasing): build the rows/columns map[0..n) of the formula matrices only
void build_SB(hipStream_t s, char X, const Dens& d, const double* eps_act, const double* e0_dev, int64_t n,
              const int* pairs, cplx* S, cplx* B, const int* map = nullptr);
// out[i + nr*c] = in[map[i] + ldin*c]
void gather_rows(hipStream_t s, int64_t nr, int64_t ncol, const int* map, const cplx* in, int64_t ldin, cplx* out);
// effective one-electron terms of A, C, D (tramo1 + integral sums)
void effh_A(hipStream_t s, int ni, int na, const cplx* hp, int ldh, const cplx* A2, cplx* effh);
void effh_C(hipStream_t s, int ni, int na, int ns, const cplx* hp, int ldh, const cplx* C1, const cplx* C2,
            const cplx* C3, cplx* effh);
void effh_D(hipStream_t s, int ni, int na, int ns, const cplx* hp, int ldh, const cplx* D3b, cplx* effh);
// D3 reshapes used by the A and C V-matrix GEMMs
// the 3-RDM part of V_A / V_C from the E_qr|0> vectors (no 3-RDM): Wt[(x + nx t) + nx na I] =
// sum_s <I|E_ts|Z_{x,s}> (sum_annihilation, V_C) or sum_s <I|E_st|Z_{x,s}> (V_A), Z_{x,s} at
// Zt[x xs + s ss + ld J]; V[(v + na u + na2 t) + na3 x] = scale Out[(u + na v) + na2 (x + nx t)]
void v3_esum(hipStream_t s, bool sum_annihilation, int na, int64_t nf, int nx, int64_t xs, int64_t ss, int64_t ld,
             const int* tab, const cplx* Zt, cplx* Wt);
void v3_permute(hipStream_t s, int na, int nx, double scale, const cplx* Out, cplx* V);
void d3_perm_A(hipStream_t s, int na, const cplx* G3, cplx* Dm);
void d3_perm_C(hipStream_t s, int na, const cplx* G3, cplx* Dm);
void a1_perm(hipStream_t s, int na, int ni, const cplx* A1, cplx* Am);
void V_A_rest(hipStream_t s, int na, int ni, const Dens& d, const cplx* A1, const cplx* effh, cplx* V);
void V_C_rest(hipStream_t s, int na, int ns, const Dens& d, const cplx* effh, cplx* V);
void V_B(hipStream_t s, int na, int ni, const Dens& d, const cplx* Bb, const int* tpairs, const int* ipairs,
         cplx* V);
void V_D(hipStream_t s, int na, int ni, int ns, const Dens& d, const cplx* D1b, const cplx* D2b, const cplx* effh,
         cplx* V);
void V_E(hipStream_t s, int na, int ni, int ns, const Dens& d, const cplx* Eb, const int* ipairs, cplx* V);
// scratch of V_F_scratch(na, ns) complex elements (the permuted block, the density matrix
// and the all-pairs product)
int64_t V_F_scratch(int na, int ns);
void V_F(hipStream_t s, int na, int ns, const Dens& d, const cplx* Fb, const int* tpairs, const int* spairs, cplx* V,
         cplx* scratch);
void V_G(hipStream_t s, int na, int ni, int ns, const Dens& d, const cplx* Gb, const int* spairs, cplx* V);
// alpha per perturber of subspace X (solve_X_subspace.F90 'alpha =' lines)
void alphas(hipStream_t s, char X, int ni, int na, int ns, const double* eps, const double* e0_dev, double eshift,
            const int* ipairs, const int* spairs, int64_t npert, double* alpha);
// e2 = -sum |X_jk|^2/(alpha_k + wb_j), sumc2 = sum |X_jk|^2/(alpha_k + wb_j)^2 ; partial sums per block,
// then a single-block final pass (deterministic order)
void e2_reduce(hipStream_t s, int64_t dimm, int64_t npert, const cplx* X, const double* wb, const double* alpha,
               double* partial, int64_t npartial_max, double* out2);
void e2_H(hipStream_t s, int ni, int ns, const cplx* Hb, const double* eps, double eshift, const int* ipairs,
          const int* spairs, double* partial, int64_t npartial_max, double* out2);
int64_t e2_partials_needed(int64_t nwork);
// E2/sumc2 of the determinant-space route without the B~ eigensolve: per perturber k the
// compressed resolvent of diag(f + alpha_k) on the retained directions Gk (Gd = the r
// dropped eigenvectors of G, Z = Gk Lambda^-1 Gk^H Phi V); partial needs 2*npert doubles
int e2_schur_max_dropped();
void e2_schur(hipStream_t s, int64_t m, int r, int64_t npert, const cplx* Gd, const cplx* Z, const double* f,
              const double* alpha, double* partial, double* out2);

}  // namespace dc2
