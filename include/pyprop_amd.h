/*
 * pyprop_amd.h -- C ABI of the MI355X tensor-potential Krylov path.
 *
 * Every entry point returns an int status and never throws:
 *   0  ok
 *  -1  invalid argument (null pointer, shape / storage mismatch)
 *  -2  combination not implemented on this path
 *  -3  HIP runtime error
 *  -4  a grid barrier of a persistent launch timed out: the launch's
 *      results are invalid (reported by the *_diagnostics calls)
 *  >0  numerical failure (band Cholesky: LAPACK info; see each function)
 * Device pointers are complex128 (interleaved re, im; the layout of
 * std::complex<double> / blitz::Array<cplx, R>), C-ordered with the largest
 * stride first.  `stream` is a hipStream_t (NULL = default stream).
 * Handles are not re-entrant: one stream per handle.
 */
#ifndef PYPROP_AMD_H
#define PYPROP_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pp_overlap pp_overlap;
typedef struct pp_pamp pp_pamp;
typedef struct pp_batch_graph pp_batch_graph;
typedef struct pp_pamp3 pp_pamp3;
typedef void *pp_stream;  /* hipStream_t */

/* storage ids (core/tensorpotential/tensorpotentialmultiply_generator.py:1336) */
enum { PP_STORAGE_IDENT = 0, PP_STORAGE_DIAG = 1, PP_STORAGE_SIMP = 2,
       PP_STORAGE_HERM = 3, PP_STORAGE_BANDNH = 4, PP_STORAGE_BAND = 5 };

/* overlap operations (core/representation/combinedrepresentation.cpp:186-238) */
enum { PP_OV_SOLVE_SQRT = 0,       /* x := U^-1 x   SolveSqrtOverlap(false)    */
       PP_OV_SOLVE_SQRT_CONJ = 1,  /* x := U^-H x   SolveSqrtOverlap(true)     */
       PP_OV_MUL_SQRT = 2,         /* x := U x      MultiplySqrtOverlap(false) */
       PP_OV_MUL_SQRT_CONJ = 3,    /* x := U^H x    MultiplySqrtOverlap(true)  */
       PP_OV_SOLVE = 4,            /* x := S^-1 x   SolveOverlap               */
       PP_OV_MUL = 5 };            /* x := S x      MultiplyOverlap            */

/* zpbtrf 'U' on host band storage ab[j*(kd+1) + kd+i-j] (complex interleaved).
 * Replaces blitz::linalg::LAPACK::CalculateCholeskyFactorizationPositiveDefiniteBanded
 * (core/utility/blitzlapack.h:843-856) as called by OverlapMatrix::Setup
 * (core/representation/overlapmatrix.cpp:35-37).  Returns LAPACK info. */
int pp_band_cholesky_upper(int N, int kd, double *ab);

/* OverlapMatrix::Setup (overlapmatrix.cpp:9-104): S in upper band storage on
 * the host, factorised and uploaded with both solve plans.  *info = zpbtrf
 * info (the reference leaves it unchecked; here info>0 returns 1).
 * Supported: 2 <= kd <= 8 (B-spline orders 3 to 9); N up to 64 chunks of the
 * solver (N <= 16384 with the default chunk choice). */
int pp_overlap_create(int N, int kd, const double *S_upper, pp_overlap **out, int *info);
/* Same with an explicit solver chunk length B (power of two >= max(kd, 8),
 * at most 64 chunks; 0 = default).  No reference counterpart: the reference's
 * solve is a sequential ztbsv; the batch autotuner picks B per shape. */
int pp_overlap_create_chunked(int N, int kd, const double *S_upper, int B, pp_overlap **out, int *info);
/* Copy the real Cholesky factor U (N*(kd+1) doubles, upper band storage) to the host. */
int pp_overlap_factor(const pp_overlap *ov, double *U_upper_host);
int pp_overlap_destroy(pp_overlap *ov);
/* Apply `op` along rank `rank` of a rank-R tensor x (device), every fiber of
 * MapToRank3 (core/utility/blitztricks.h:63-107).  Replaces
 * OverlapMatrix::{Solve,Multiply}{Sqrt,}OverlapTensor (overlapmatrix.cpp:174-229). */
int pp_overlap_apply(const pp_overlap *ov, int op, int rank, int R, const int *shape, void *x,
                     pp_stream stream);

/* core.TensorPotentialMultiply_<ids>(potential, scaling, source, dest, pairs...)
 * (generator wrapper tensorpotentialmultiply_generator.py:1428-1521):
 * dest += scaling * (x_r V_r) source.  storage[R] ids, potShape[R] extents of
 * the potential (device), shape[R] extents of source/dest (device),
 * pairs[r] host int32 (potShape[r] x 2) for Simp/Herm ranks, else NULL.
 * R = 1..3. */
int pp_tp_multiply(int R, const int *storage, const int *potShape, const void *potential,
                   double scaling, const int *shape, const void *source, void *dest,
                   const int *const *pairs, pp_stream stream);

/* pAMP propagator (core.krylov_PampWrapper_R, core/krylov/pamp/pampwrapper.h:38-40,
 * pampwrapper.cpp:51-108, pamp.h:613-669) with the balanced-overlap matvec of
 * BasisPropagator.MultiplyHamiltonianBalancedOverlap (BasisPropagator.py:131-159)
 * fused in.  State: (L angular, N B-spline) complex128; m = krylov_basis_size;
 * reorth = krylov_double_orthogonalization. */
int pp_pamp_create(int L, int N, int m, int reorth, const pp_overlap *ov, pp_pamp **out);
/* One consolidated tensor potential: angular storage (Ident/Diag/Simp/Herm,
 * P0 compressed indices, host pairs for Simp/Herm) x radial BandNH of total
 * band width bw; `potential` device (P0, N*bw) in the reference layout, as
 * TensorPotential.PotentialData (pyprop/tensorpotential/TensorPotential.py:33-77).
 * The library factors it into real radial bands x complex angular
 * coefficients and returns 1 if that reconstruction misses V by more than
 * 1e-12 max|V| (the assembled tensor itself is low rank only to ~1e-14). */
int pp_pamp_add_term(pp_pamp *p, int angular_storage, int P0, const int *pairs, int radial_storage,
                     int bw, const void *potential);
/* Same term already factored: V[p,:] = sum_r coef[p,r] bands[r,:] with host
 * coef (P0 x rank complex) and real host bands (rank x N*bw, BandNH layout).
 * No reference counterpart: the generator of the reference always stores
 * the full tensor (TensorPotentialGenerator.py:72-256). */
int pp_pamp_add_factored_term(pp_pamp *p, int angular_storage, int P0, const int *pairs,
                              int radial_storage, int bw, int rank, const double *coef,
                              const double *bands);
/* Host helper behind pp_pamp_add_term: V (host, P0 x M complex) ~= coef
 * (P0 x rank complex) x bands (rank x M real rows of Re/Im V, pivoted).
 * No reference counterpart.  *relerr = max|V - recon| / max|V|. */
int pp_lowrank_factor(int P0, long M, const double *V, int maxrank, int *rank_out, double *coef,
                      double *bands, double *relerr);
/* nsteps consecutive PamPropagator.AdvanceStep(t_n, dt) (PamPropagator.py:50-59)
 * in one launch; scal (host, nsteps x nterms) = each term's time factor at t_n. */
int pp_pamp_propagate(pp_pamp *p, void *psi, int nsteps, double dt_re, double dt_im,
                      const double *scal, pp_stream stream);
/* Per step of the last launch (host, nsteps x 4): |last Krylov coefficient|
 * (pamp.h:666; -1 if the Pade denominator was singular, pamp.h:544-550),
 * basis size used (< m after happy breakdown), |U psi|, re-orthogonalisation passes.
 * |U psi| = 0 marks a zero wavefunction (the step leaves it zero; the
 * reference prints "Wavefunction is zero" and divides by the zero norm).
 * Returns -4 if a grid barrier of the launch timed out.  The Python layer
 * raises RuntimeError on each of these, as the reference throws. */
int pp_pamp_diagnostics(pp_pamp *p, int nsteps, double *diag, pp_stream stream);
/* nsteps consecutive ExpokitPropagator.AdvanceStep(t_n, dt)
 * (pyprop/propagator/KrylovPropagator.py:41-45, core.krylov_ExpokitPropagator_R,
 * core/krylov/expokit/expokitpropagator.h:29-31, expokitpropagator.cpp:36-103)
 * on the same engine: psi~ = U psi, w = exp(|dt| A) psi~ by zgexpv
 * (zgexpv.cpp:32-586, adaptive substeps with error control), psi = U^-1 w,
 * with A = -i H^ (real dt) or -H^ (dt_re == 0, imaginary time).  tol =
 * krylov_tolerance (<= eps means sqrt(eps)), anorm = krylov_norm.  Complex
 * dt and reorth != 0 return -2.  The reference passes dt^2 as the Krylov time
 * for real dt (expokitpropagator.cpp:61); here it is |dt|. */
int pp_expokit_propagate(pp_pamp *p, void *psi, int nsteps, double dt_re, double dt_im,
                         const double *scal, double tol, double anorm, pp_stream stream);
/* Per step of the last pp_expokit_propagate (host, nsteps x 6): matvecs,
 * exponentials, substeps, rejected substeps, basis size of the last substep,
 * zgexpv iflag (0 ok, 1 more than 500 substeps, -9 Pade failure, -8 more than
 * 100 rejections of one substep, a bound zgexpv does not have, -10 zero
 * wavefunction: no substep, psi stays zero); the reference throws on
 * iflag != 0 (expokitpropagator.cpp:96-100).  Returns -4 if a grid barrier
 * of the launch timed out. */
int pp_expokit_diagnostics(pp_pamp *p, int nsteps, double *stats, pp_stream stream);
/* Batched form (no reference counterpart; the reference runs one problem per
 * process): nsteps pAMP steps of each of ninst independent engines (own
 * operator, overlap, psi[i]; dt: ninst x 2 re/im; scal[i]: host nsteps x
 * nterms_i), packed in order into cooperative launches whose workgroups (one
 * per fiber) are all co-resident; larger batches take several launches on the
 * stream.  Engines must share the B-spline order and the precision form. */
int pp_pamp_batch_propagate(int ninst, pp_pamp *const *inst, void *const *psi, int nsteps, const double *dt,
                            const double *const *scal, pp_stream stream);
/* The same batch recorded once as a HIP graph (no reference counterpart):
 * per pack a kernel node that clears the barrier words and the batch kernel
 * as a cooperative kernel node, chained in order.  The graph keeps psi[i] and
 * the engines' buffers; each pp_pamp_batch_graph_launch uploads the replay's
 * time factors (scal[i]: host nsteps x nterms_i) and replays all packs on the
 * stream, advancing every psi[i] by nsteps steps.  A launch returns PP_EARG
 * if an engine changed since recording (terms, precision, Pade scaling, or a
 * launch that reallocated its buffers); destroy and record again then. */
int pp_pamp_batch_graph_create(int ninst, pp_pamp *const *inst, void *const *psi, int nsteps, const double *dt,
                               pp_batch_graph **out);
int pp_pamp_batch_graph_launch(pp_batch_graph *g, const double *const *scal, pp_stream stream);
int pp_pamp_batch_graph_destroy(pp_batch_graph *g);
/* LDS bytes per workgroup a pAMP launch of an (L, N) problem with Krylov
 * dimension m, nslots coupled fibers per row, kd super-diagonals and solver
 * chunk length B needs (B = 0: the default chunk of pp_overlap_create_chunked;
 * -1: invalid B).  No reference counterpart: the batch autotuner uses it to
 * keep two workgroups per CU, and the Expokit host side to pick a chunk whose
 * kernel fits 160 KB of LDS. */
long pp_pamp_lds_estimate(int L, int N, int m, int nslots, int kd, int B);
/* Pade scaling of later pp_pamp_propagate launches (no reference
 * counterpart): 0 (default) scales A = -i dt H_m by its largest element as
 * pAMP does (pamp.h:442-445); 1 scales until the 1-norm is at most 1/2, which
 * keeps the Pade(6) error at rounding level and makes a step a smooth
 * function of its inputs (used by the adjoint gradients). */
int pp_pamp_set_pade_scaling(pp_pamp *p, int mode);
/* Exponential of the Krylov matrix in later launches: the reference's
 * krylov_exponentiation_method (core/krylov/pamp/pampwrapper.cpp:56-61):
 * 0 scaling + Pade(6) (pamp.h:430-564, the default), 1 the eigenvector
 * exponential R exp(Lambda) R^H e1 (pAMP::EigenvectorExponential,
 * pamp.h:567-607; Jacobi on the Hermitian part of the Krylov matrix). */
int pp_pamp_set_exponentiation(pp_pamp *p, int method);
/* Mixed-precision form (no reference counterpart; the reference is FP64
 * throughout): later launches apply the terms whose bit is set in term_mask
 * (bit t = t-th added term) in complex FP32 (FP32 radial bands, FP32 band
 * sums); Gram-Schmidt, the overlap solves, the exchanged U^-1 w and the Pade
 * exponential stay FP64.  0 restores the all-FP64 drop-in. */
int pp_pamp_set_mixed(pp_pamp *p, unsigned term_mask);
/* Launch shape of pp_pamp_propagate (no reference counterpart; the
 * reference runs one fiber loop per MPI rank): q workgroups share each
 * angular fiber, each owning a contiguous run of the overlap solver's
 * chunks.  0 (default) chooses q so that L*q workgroups fill the CUs (at
 * most 8, at least 2 chunks each; 1 with re-orthogonalisation), 1 keeps one
 * workgroup per fiber, q >= 2 forces q (PP_EUNSUP above the chunk count).
 * Results agree with q = 1 to rounding (partial sums are grouped
 * differently); batch and Expokit launches are not affected. */
int pp_pamp_set_fiber_split(pp_pamp *p, int q);
/* Profiling hook (no reference counterpart; the reference keeps per-phase
 * Timers in pAMP, pamp.h:41,125,295-349): record wall-clock stamps (100 MHz)
 * of workgroup 0 around every grid barrier of step 0 of later launches. */
int pp_pamp_set_timing(pp_pamp *p, int enable);
int pp_pamp_timings(pp_pamp *p, long long *out, int max, int *count);
int pp_pamp_destroy(pp_pamp *p);

/* ---- two-electron states: coupled channels x two B-spline radial ranks ----
 * Replaces core.krylov_PampWrapper_3 driven by PamPropagator.AdvanceStep
 * (pyprop/propagator/PamPropagator.py:50-59, pampwrapper.cpp:51-108) with
 * the balanced rank-3 matvec BasisPropagator.MultiplyHamiltonianBalancedOverlap
 * (BasisPropagator.py:131-159) over the potentials of
 * examples/tensor/helium_stabilization/config.ini (TensorPotentialMultiply_
 * Simp_Simp_Simp and _Simp_Simp_BandNH).  psi is (C, N, N) complex128, the
 * second radial rank contiguous; both radial ranks use the overlap ov (order
 * 7 B-splines).  reorth != 0 returns -2. */
int pp_pamp3_create(int C, int N, int m, int reorth, const pp_overlap *ov, pp_pamp3 **out);
/* A factored potential term sum_p coef[p] |pairs[p,0]><pairs[p,1]| (x) A (x) B
 * (coef: P complex as 2P doubles; A, B real N x 13 row-major bands, element
 * (i, j) at i*13 + 6 + j - i; slot: column of the per-step time scalings,
 * -1 = static).  One-electron terms and the velocity-gauge dipole of the
 * reference's helium potentials are sums of such terms
 * (coupledspherical.cpp:120-205, potential.cpp:7-77, coupledvelocity.cpp). */
int pp_pamp3_add_kron_term(pp_pamp3 *p, int P, const int *pairs, const double *coef, const double *A,
                           const double *B, int slot);
/* A multipole term sum_p coef[p] |pairs[p,0]><pairs[p,1]| (x) R with the real
 * 4-index radial tensor R[i1][d1][i2][d2] (N x 13 x N x 13, j = i + d - 6):
 * one l of CustomPotentialEvaluationR12 (coupledspherical.cpp:16-114).
 * Pband, Qband (optional, both or neither): <B|r^l|B> and <B|r^-(l+1)|B> as
 * row-major bands; with them and the radial supports set, the engine applies
 * R as P (x) Q / Q (x) P wherever the supports of B_i1 and B_i2 are disjoint
 * (exact for product quadrature) and the full R only on the near band. */
int pp_pamp3_add_multipole_term(pp_pamp3 *p, int P, const int *pairs, const double *coef, const double *R,
                                const double *Pband, const double *Qband, int slot);
/* Support of every B-spline as an interval range [lo[i], hi[i]] of the
 * breakpoint intervals (both non-decreasing in i). */
int pp_pamp3_set_radial_supports(pp_pamp3 *p, const int *lo, const int *hi);
/* dst = H src, or U^-H H U^-1 src with balanced != 0 (src and dst device,
 * scal: nslots time scalings on the host).  Synchronises the stream. */
int pp_pamp3_apply(pp_pamp3 *p, const void *src, void *dst, const double *scal, int balanced, pp_stream stream);
/* nsteps PamPropagator.AdvanceStep's of psi (device): psi~ = (U x U) psi,
 * pAMP on the balanced operator with scal[n*nslots + slot] at step n,
 * psi = (U x U)^-1 psi~.  Synchronises the stream. */
int pp_pamp3_propagate(pp_pamp3 *p, void *psi, int nsteps, double dt_re, double dt_im, const double *scal,
                       pp_stream stream);
/* Per step of the last pp_pamp3_propagate (host, nsteps x 4): |last Krylov
 * coefficient|, basis size used, |psi~|, 0. */
int pp_pamp3_diagnostics(pp_pamp3 *p, int nsteps, double *diag, pp_stream stream);
/* Mixed form of the rank-3 engine (no reference counterpart: the reference
 * is complex128 throughout, helium_stabilization/config.ini): with mixed != 0
 * the multipole terms (the near/far split stage) run in FP32 with complex64
 * intermediates and FP32 radial tables, each output added once into the
 * FP64 result; the one-electron and laser terms, Gram-Schmidt, the overlap
 * solves and the Pade exponential stay FP64.  Needs the split (radial
 * supports and P_l, Q_l of every multipole term): later calls return
 * PP_EUNSUP otherwise. */
int pp_pamp3_set_mixed(pp_pamp3 *p, int mixed);
/* krylov_exponentiation_method for the two-electron engine, as
 * pp_pamp_set_exponentiation (pampwrapper.cpp:56-61). */
int pp_pamp3_set_exponentiation(pp_pamp3 *p, int method);
/* Measurement hooks (no reference counterpart): with timing enabled, every
 * multipole contraction launch of later calls is bracketed by HIP events on
 * its stream; pp_pamp3_kernel_time returns their summed duration (ms) and
 * count since timing was enabled; pp_pamp3_multipole_flops the algorithmic
 * flops of one such launch. */
int pp_pamp3_set_timing(pp_pamp3 *p, int enable);
int pp_pamp3_kernel_time(pp_pamp3 *p, double *ms_total, long *launches);
int pp_pamp3_multipole_flops(pp_pamp3 *p, double *flops);
int pp_pamp3_destroy(pp_pamp3 *p);

/* Krylov-basis vector work of the eigenvalue solver (PiramSolver), whose
 * reference does it with BLAS on the host inside pIRAM
 * (core/krylov/piram/piram/piram.h:320-445 Arnoldi step and
 * orthogonalisation, :540-640 restart).  Basis V: k vectors of n complex128,
 * vector j at V + j*n.
 * pp_gram_schmidt: classical Gram-Schmidt of w against V_0..V_{k-1}, passes
 * times (1-3; deterministic reductions), h (host, 2k doubles re/im) = the
 * summed coefficients, *wnorm = |w| afterwards; synchronises the stream.
 * pp_basis_combine: out[p] = sum_j V_j Q[j, p], p < np, Q host column-major
 * m x np complex128; out must not overlap V. */
int pp_gram_schmidt(const void *V, long n, int k, void *w, int passes, double *h, double *wnorm, pp_stream stream);
int pp_basis_combine(const void *V, long n, int m, const double *Q, int np, void *out, pp_stream stream);
/* out (host, 2 doubles re/im) = sum_i conj(a_i) b_i over n complex128
 * (device): the zdotc of CombinedRepresentation::InnerProduct
 * (core/representation/combinedrepresentation.cpp:423-474) after the overlap
 * multiply, and of Wavefunction::GetNorm (core/wavefunction.cpp:194-198);
 * fixed-order reduction; synchronises the stream. */
int pp_inner_product(const void *a, const void *b, long n, double *out, pp_stream stream);
/* pp_gram_schmidt without any allocation.  No reference counterpart: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch is a fixed device array of the library, one per GPU, sized for
 * k <= 128 basis vectors (PP_EARG beyond); every call on a GPU shares that
 * GPU's array, so calls that overlap on one GPU race.  Same kernels, same
 * results as pp_gram_schmidt; synchronises the stream. */
int pp_gram_schmidt_noalloc(const void *V, long n, int k, void *w, int passes, double *h, double *wnorm,
                            pp_stream stream);
/* Per-fiber banded LU solve of the shift-invert and Cayley preconditioner (the device
 * side of pyprop's BSplinePreconditioner.Solve,
 * examples/tensor/helium_stabilization_sae/example.py:275-318): x_a := A_a^-1
 * x_a for nfib fibers of length N, A_a factored by LAPACK zgbtrf (lub: nfib
 * x N x (2kl+ku+1) complex128 column bands, piv: nfib x N 0-based). */
int pp_band_lu_solve(int nfib, int N, int kl, int ku, const void *lub, const int *piv, void *x, pp_stream stream);

/* Two-electron radial preconditioner (the device side of pyprop's
 * RadialTwoElectronPreconditioner Setup / Solve,
 * examples/tensor/helium_stabilization/preconditioner.py:77-244): on a
 * (C, N, N) state, x_c := M_c^-1 x_c with M_c = b S (x) S + a (H1_c (x) S +
 * S (x) H2_c), by fast diagonalisation.  create copies, for nmat distinct
 * real symmetric radial matrices H, the S-orthonormal generalised
 * eigenvectors V (nmat x N x N row-major, column n the n-th eigenvector,
 * V^T S V = I) and eigenvalues lam (nmat x N); idx1[c], idx2[c] select
 * channel c's matrix on radial rank 1 and 2 (host arrays).  solve runs
 * V1 [(V1^T X V2) ./ (b + a (lam1_i + lam2_k))] V2^T per channel,
 * asynchronously on `stream`; solves that overlap on one handle race (one
 * intermediate buffer per handle). */
/* The B-spline rank's projection step of tensor-potential generation
 * (TensorPotentialGenerator.GeneratePotential, pyprop/tensorpotential/
 * TensorPotentialGenerator.py:72-256, with basis_bspline.h's
 * <B_i| f d^d |B_j>): out[r][e] = sum_q MT[e][q] f[r][q] for rows x G
 * weighted quadrature samples f (device complex128), MT the sparse map to
 * the band pairs as host CSR (ptr nout+1, idx/val nnz, copied at create).
 * Selected by the `generate_on_device = True` key of a potential section
 * (no reference counterpart: the reference projects on the host). */
typedef struct pp_projector pp_projector;
int pp_projector_create(int nout, int G, const int *ptr, const int *idx, const double *val, pp_projector **out);
int pp_projector_apply(const pp_projector *p, long rows, const void *f, void *out, pp_stream stream);
int pp_projector_destroy(pp_projector *p);

typedef struct pp_radial_prec pp_radial_prec;
int pp_radial_prec_create(int C, int N, int nmat, const int *idx1, const int *idx2, const double *V,
                          const double *lam, pp_radial_prec **out);
int pp_radial_prec_solve(pp_radial_prec *p, void *psi, double a_re, double a_im, double b_re, double b_im,
                         pp_stream stream);
int pp_radial_prec_destroy(pp_radial_prec *p);

#ifdef __cplusplus
}
#endif
#endif
