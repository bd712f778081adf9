// pAMP Krylov propagation on MI355X: one persistent cooperative launch runs
// whole time steps of
//
//   psi~ = U psi;  fixed-m Arnoldi on  H^ = U^-H (sum_t s_t(t_n) V_t) U^-1
//   with classical Gram-Schmidt;  exp(-i dt H_m) by scaling + Pade(6);
//   psi~ <- |psi~| V exp(-i dt H_m) e1;  psi = U^-1 psi~
//
// Reference: pyprop/propagator/PamPropagator.py:50-75 (balanced overlap
// wrapping), pyprop/tensorpotential/BasisPropagator.py:131-159 (the matvec),
// core/krylov/pamp/pamp/pamp.h:271-350 (Arnoldi/CGS), :353-400 (optional
// re-orthogonalisation, 0.717 criterion, <= 3 passes), :430-564 (scaling and
// Pade), :613-669 (PropagateVector); Problem.py:158-173 (H held at t_n).
//
// Work split: the wavefunction is (L, N), angular rank outer, B-spline rank
// contiguous; workgroup a owns fiber a for the whole launch.  Its Arnoldi
// vectors (fiber-major, L2-resident), residual (LDS) and the chunked solver
// plan (LDS) never leave it.  Only U^-1 w (needed by the neighbours through
// the angular couplings), the per-fiber partial inner products and the
// Krylov coefficients cross workgroups; those go through agent-scope atomic
// loads/stores, so the grid barrier needs no cache write-back/invalidate and
// the private data stays cached.  Each Arnoldi iteration costs two barriers;
// every workgroup reduces the partial sums itself in one fixed order, and
// breakdown / re-orthogonalisation only predicate work, so all workgroups
// execute the same barrier sequence whatever the data.
//
// Tensor potentials arrive factored as sum_r c_{p,r} B_r with real radial
// bands B_r (diagonal-major, shared by all workgroups and L2-resident) and
// complex angular coefficients c (scalar per workgroup).
#include "pp_common.h"
#include <hip/hip_cooperative_groups.h>
#include <cstring>
#include <cstdio>
#include <cmath>
#include <vector>
#include <algorithm>

#define PP_MAXTERMS 8
#define PP_MAXM 64
#define PP_MAXSLOTS 8    // distinct source fibers per row (all terms)
#define PP_MAXENT 4      // couplings per row of one term
#define PP_MAXCOMBO 16
#ifndef PP_THREADS
#define PP_THREADS 256   // threads per workgroup (one workgroup per fiber)
#endif
#ifndef PP_VUNROLL
#define PP_VUNROLL 4     // Krylov-vector loads in flight per thread in the dot and projection loops
#endif
#ifndef PP_SOLVE_THREADS
#define PP_SOLVE_THREADS (PP_THREADS < 256 ? PP_THREADS : 256)   // threads in the band solves
#endif
#define PP_TLOG_MAX 512
#include "pp_pade.h"
// LDS doubles of the per-row combo table (coefficients, slots, counts, terms)
#define PP_CMB_DOUBLES (PP_MAXCOMBO * PP_MAXENT * 5 + 2 * PP_MAXCOMBO)

struct pp_fterm {
  const int *rowptr;   // L+1
  const int *ent;      // (slot, pidx, conj) per entry
  const zd *coef;      // P0 x rank
  const double *bands; // rank x bw x N, element (i, i+d) at (r*bw + d+s)*N + i
  const float *bands32; // the same in FP32 (mixed-precision form)
  int fp32;             // apply this term in complex FP32
  int rank;
};

struct pp_barrier {
  unsigned count, gen, err, pad;
};

struct pp_pamp_args {
  int L, N, m, s, bw, nterms, nsteps, reorth, padN, nslots, ncombo, scrDoubles, nlevEven;
  int normscale;         // Pade scaling by the 1-norm instead of the largest element
  double dt_re, dt_im;
  pp_fterm term[PP_MAXTERMS];
  const int *slotcols;   // L x nslots source fibers staged per row (-1 unused)
  const double *scal;    // nsteps x nterms
  const double *Uband;   // real U band for psi~ = U psi
  pp_band_plan plan;     // chunk solver plan of U (scan matrices copied to LDS)
  zd *psi;               // L x N, in/out (fiber a private to workgroup a)
  zd *V;                 // L x (m+1) x N (fiber-major, private)
  zd *y;                 // L x N, U^-1 w (exchanged)
  double *part;          // L x pstr partial sums (exchanged)
  int pstr;
  zd *Hm;                // m x m column-major, private to workgroup 0
  zd *cvec;              // m, exp(-i dt H) e1 (exchanged)
  double *diag;          // nsteps x 4
  pp_barrier *bar;
  long long *tlog;
  int Q;                 // split kernel: workgroups per fiber
  zd *sb;                // split kernel: L x 2 x P*KD chunk boundary values (exchanged)
  unsigned *subw;        // split kernel: L x Q arrival words of the per-fiber barriers
  int vcache;            // split kernel: Krylov vectors whose own rows are also kept in LDS
  const double *srows;   // split kernel: per part q (stride sstride doubles) the chunk-boundary rows
  int sstride;           //   of the band solves (pamp_split_rows), copied to LDS in place of the scan matrices
};

struct pp_pamp {
  int L, N, m, reorth, s, bw;
  const pp_overlap *ov;
  int nterms, nslots;
  pp_fterm term[PP_MAXTERMS];
  std::vector<std::vector<int>> rowcols;   // per row: distinct source fibers
  std::vector<std::vector<int>> termrows;  // per term: rowptr (L+1) + (col, pidx, conj) entries
  void *termbuf[PP_MAXTERMS];
  int *slotcols;
  bool dirty;
  zd *V, *y, *Hm, *cvec;
  unsigned mixed;                          // bit t: term t applied in complex FP32
  int normscale;                           // Pade scaling by the 1-norm
  int expmethod;                           // 0 Pade, 1 eigenvector exponential
  double *part, *diag, *scal;
  pp_barrier *bar;
  int diag_cap, scal_cap, pstr;
  long long *tlog;
  int timing;
  int fsplit;                              // workgroups per fiber: 0 automatic, 1 none, >= 2 fixed
  size_t bar_words, part_cap, sb_cap;      // allocated arrival words, partial-sum doubles, boundary values
  zd *sb;
  double *srows;                           // split kernel: chunk-boundary rows per part (pamp_split_rows)
  int srows_q, sstride;                    // the split factor they were built for, doubles per part
  unsigned gen;                            // bumped whenever a buffer or setting a recorded launch uses changes
};

// Barrier over n co-resident workgroups (member `me`; a whole instance, or
// the workgroups sharing one fiber in the split kernel): every workgroup
// publishes the number of barriers it has reached in its own arrival word
// (word[me]) and wavefront 0 polls the n words until each has reached the
// same count, so a barrier costs one store plus the polling round trips.  Data crossing it is written/read with agent-scope atomics
// (above); the workgroup-scope fences order this workgroup's accesses around
// the arrival.  A bounded spin turns a lost arrival into an error flag
// instead of a hang (*err).  `ep` counts the barriers of this launch
// (identical in every member, since all execute the same barrier sequence).
__device__ inline void group_barrier(unsigned *word, unsigned *err, unsigned &ep, int n, int me)
{
  const int tid = threadIdx.x;
  ep++;
  // every wave waits until its own stores (the exchanged values, written
  // with agent-scope atomics) are acknowledged before the workgroup barrier:
  // the arrival word below must not become visible before them
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid < 64) {
    if (tid == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __hip_atomic_store(word + me, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    long spins = 0;
    while (true) {
      bool ok = true;
      for (int f = tid; f < n; f += 64)
        ok = ok && (int)(__hip_atomic_load(word + f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - ep) >= 0;
      if (__all(ok)) break;
      __builtin_amdgcn_s_sleep(1);
      if (++spins > (1l << 24)) {
        if (tid == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  __syncthreads();
}

__device__ inline void grid_barrier(pp_barrier *b, unsigned &ep, int n, int me)
{
  group_barrier((unsigned *)(b + 1), &b->err, ep, n, me);
}

// ------------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// out[c] = sum_{a=0}^{L-1} part[a*pstr + off + c], c < ncomp, in one fixed
// order (identical in every workgroup).  The L*ncomp partials are staged in
// LDS (scratch) with independent loads first, then summed per component.
__device__ inline void grid_partial_sum(const double *part, int pstr, int off, int ncomp, int L,
                                        double *out_lds, double *scratch)
{
  const int tid = threadIdx.x;
  const int tot = L * ncomp;
#pragma unroll 4
  for (int f = tid; f < tot; f += PP_THREADS) {
    const int a = f / ncomp, c = f - a * ncomp;
    scratch[f] = xld(part + (size_t)a * pstr + off + c);
  }
  __syncthreads();
  if (tid < ncomp) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int a = 0;
    for (; a + 3 < L; a += 4) {
      s0 += scratch[a * ncomp + tid]; s1 += scratch[(a + 1) * ncomp + tid];
      s2 += scratch[(a + 2) * ncomp + tid]; s3 += scratch[(a + 3) * ncomp + tid];
    }
    for (; a < L; a++) s0 += scratch[a * ncomp + tid];
    out_lds[tid] = (s0 + s1) + (s2 + s3);
  }
  __syncthreads();
}

__device__ inline double block_sum(double v, double *wave_lds)
{
  const int tid = threadIdx.x;
  v = wave_sum(v);
  if ((tid & 63) == 0) wave_lds[tid >> 6] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < PP_THREADS / 64; w++) r += wave_lds[w];
  __syncthreads();
  return r;
}

// out[2k], out[2k+1] = <V_k, W> over n rows for k < nk (V_k at V + k*ldv);
// with extra_norm also out[2nk] = |W|^2.  Wave w owns k = w, w+4, ...; each
// lane streams a strided slice so the loads of one k are independent and
// pipeline.
__device__ inline void fiber_dots(const zd *V, int ldv, const zd *W, int nk, int N, bool extra_norm, double *out)
{
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int ntot = nk + (extra_norm ? 1 : 0);
  for (int k = wave; k < ntot; k += PP_THREADS / 64) {
    double re = 0.0, im = 0.0;
    if (k < nk) {
      const zd *vk = V + (size_t)k * ldv;
#pragma unroll PP_VUNROLL
      for (int i = lane; i < N; i += 64) {
        zd v = vk[i], w = W[i];
        re = fma(v.x, w.x, re); re = fma(v.y, w.y, re);
        im = fma(v.x, w.y, im); im = fma(-v.y, w.x, im);
      }
    } else {
#pragma unroll PP_VUNROLL
      for (int i = lane; i < N; i += 64) { zd w = W[i]; re = fma(w.x, w.x, re); re = fma(w.y, w.y, re); }
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) { out[2 * k] = re; if (k < nk) out[2 * k + 1] = im; }
  }
  __syncthreads();
}

// fiber_dots over one workgroup's rows of the Krylov vectors with the first
// kc vectors read from their LDS copy (stride ldl) and the rest from global
// memory (stride ldg): the same values in the same order as fiber_dots.
__device__ __attribute__((always_inline)) inline void split_dots(const zd *Vl, int ldl, int kc, const zd *Vg, int ldg,
                                                                 const zd *W, int nk, int N, double *out)
{
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (int k = wave; k < nk; k += PP_THREADS / 64) {
    double re = 0.0, im = 0.0;
    if (k < kc) {
      const zd *vk = Vl + (size_t)k * ldl;
#pragma unroll PP_VUNROLL
      for (int i = lane; i < N; i += 64) {
        zd v = vk[i], w = W[i];
        re = fma(v.x, w.x, re); re = fma(v.y, w.y, re);
        im = fma(v.x, w.y, im); im = fma(-v.y, w.x, im);
      }
    } else {
      const zd *vk = Vg + (size_t)k * ldg;
#pragma unroll PP_VUNROLL
      for (int i = lane; i < N; i += 64) {
        zd v = vk[i], w = W[i];
        re = fma(v.x, w.x, re); re = fma(v.y, w.y, re);
        im = fma(v.x, w.y, im); im = fma(-v.y, w.x, im);
      }
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) { out[2 * k] = re; out[2 * k + 1] = im; }
  }
  __syncthreads();
}

// ------------------------------------------------------------------ the step kernel
// ------------------------------------------------------------------ LDS map and the operator apply
// LDS carve-up shared by the step kernels.  Persistent for a launch: the
// solver's scan matrices and the combo table; the rest is per step (and
// reused by the Pade exponential).
struct pp_lds_map {
  double *nl;
  zd *cmb_c; int *cmb_q, *cmb_n, *cmb_t;   // per row: coefficients, slots, counts, terms
  zd *cur_c; int *cur_q, *cur_n;           // rewritten before every matvec
  zd *W, *X, *tb, *Ys, *Z;                 // residual, solve in/out, solver scratch, staged fibers
  double *scr, *wsum, *hred, *gscr, *hacc;
  zd *cv;
};

template <int KD>
__device__ __attribute__((always_inline)) inline pp_lds_map pp_map_lds(const pp_pamp_args &A, double *lds_d,
                                                                      int extra = 0)
{
  pp_lds_map M;
  const int padN = A.padN, m = A.m;
  M.nl = lds_d;
  M.cmb_c = (zd *)(lds_d + A.nlevEven);
  M.cmb_q = (int *)(M.cmb_c + PP_MAXCOMBO * PP_MAXENT);
  M.cmb_n = M.cmb_q + PP_MAXCOMBO * PP_MAXENT;
  M.cmb_t = M.cmb_n + PP_MAXCOMBO;
  M.cur_c = (zd *)(M.cmb_t + PP_MAXCOMBO);
  M.cur_q = (int *)(M.cur_c + PP_MAXCOMBO * PP_MAXENT);
  M.cur_n = M.cur_q + PP_MAXCOMBO * PP_MAXENT;
  M.W = (zd *)(lds_d + A.nlevEven + PP_CMB_DOUBLES + extra);   // extra: kernel-specific persistent doubles
  M.X = M.W + padN;
  M.tb = M.X + padN;
  M.Ys = M.tb + A.plan.P * KD;
  M.Z = M.Ys;
  M.scr = (double *)M.Ys;
  M.wsum = (double *)M.Ys + A.scrDoubles;
  M.hred = M.wsum + 16;
  M.gscr = M.hred + (2 * m + 4);
  M.hacc = M.gscr + PP_THREADS;
  M.cv = (zd *)(M.hacc + 2 * m);
  return M;
}

// Scan matrices to LDS and the combo table of row a: one entry per
// (term, rank) with the angular coefficients of its couplings (time factors
// applied per matvec).
__device__ __attribute__((always_inline)) inline void pp_load_persistent(const pp_pamp_args &A, const pp_lds_map &M, int a,
                                                                         const double *nsrc = nullptr, int nlen = -1)
{
  const int tid = threadIdx.x;
  // the solver's persistent matrices: the scan matrices, or those given (nsrc, nlen)
  if (!nsrc) { nsrc = A.plan.nlevs; nlen = A.plan.nlev; }
  for (int i = tid; i < nlen; i += PP_THREADS) M.nl[i] = nsrc[i];
  if (tid == 0) {
    int k = 0;
    for (int t = 0; t < A.nterms; t++) {
      const pp_fterm T = A.term[t];
      const int e0 = T.rowptr[a], e1 = T.rowptr[a + 1];
      for (int r = 0; r < T.rank; r++, k++) {
        int ne = 0;
        for (int e = e0; e < e1; e++, ne++) {
          zd c = T.coef[(size_t)T.ent[3 * e + 1] * T.rank + r];
          if (T.ent[3 * e + 2]) c = zconj(c);
          M.cmb_c[k * PP_MAXENT + ne] = c;
          M.cmb_q[k * PP_MAXENT + ne] = T.ent[3 * e];
        }
        M.cmb_n[k] = ne;
        M.cmb_t[k] = t;
      }
    }
  }
  __syncthreads();
}

// Coefficients of the combo table times the terms' time factors at t_step
// and the factor fac (rewritten before every operator apply).
__device__ inline void pp_step_coeffs(const pp_pamp_args &A, const pp_lds_map &M, int step, zd fac)
{
  for (int e = threadIdx.x; e < A.ncombo * PP_MAXENT; e += PP_THREADS) {
    M.cur_c[e] = zmul(M.cmb_c[e], zscale(fac, A.scal[(size_t)step * A.nterms + M.cmb_t[e / PP_MAXENT]]));
    M.cur_q[e] = M.cmb_q[e];
    if (e % PP_MAXENT == 0) M.cur_n[e / PP_MAXENT] = M.cmb_n[e / PP_MAXENT];
  }
}

// X[i - r0] = sum_k B_k (sum_q c_kq Y_q) at rows r0 <= i < r0 + nloc (zero
// for rows >= N), with row j of staged source fiber q at
// Ys[q * ystr + j - ybase] for every |j - i| <= KD inside [0, N).  Bands
// diagonal-major, all loads independent; the combo table and coefficients
// are in LDS (pp_step_coeffs).
template <int KD, bool MIXED>
__device__ __attribute__((always_inline)) inline void pp_band_rows(const pp_pamp_args &A, const pp_lds_map &M, const zd *Ys,
                                                                   int ystr, int ybase, int r0, int nloc, zd *X)
{
  const int N = A.N, tid = threadIdx.x;
  const zd *cur_c = M.cur_c;
  const int *cur_q = M.cur_q, *cur_n = M.cur_n;
  if (MIXED) for (int i = r0 + tid; i < r0 + nloc; i += PP_THREADS) {
    zd acc = zmk(0.0, 0.0);
    if (i < N) {
      const bool interior = i >= KD && i + KD < N;
      int k = 0;
      for (int t = 0; t < A.nterms; t++) {
        const bool f32 = A.term[t].fp32 != 0;
        for (int r = 0; r < A.term[t].rank; r++, k++) {
          const int ne = cur_n[k];
          if (ne == 0) continue;
          if (f32) {
            const float *Br = A.term[t].bands32 + (size_t)r * (2 * KD + 1) * N;
            float bv[2 * KD + 1];
#pragma unroll
            for (int d = -KD; d <= KD; d++)
              bv[d + KD] = (interior || (i + d >= 0 && i + d < N)) ? Br[(size_t)(d + KD) * N + i] : 0.f;
#pragma unroll
            for (int q = 0; q < PP_MAXENT; q++) {
              if (q < ne) {
                const zd *Yq = Ys + cur_q[k * PP_MAXENT + q] * ystr;
                float zr = 0.f, zi = 0.f;
#pragma unroll
                for (int d = -KD; d <= KD; d++) {
                  const int j = i + d;
                  const zd yv = (interior || (j >= 0 && j < N)) ? Yq[j - ybase] : zmk(0.0, 0.0);
                  zr = fmaf(bv[d + KD], (float)yv.x, zr);
                  zi = fmaf(bv[d + KD], (float)yv.y, zi);
                }
                zfma(acc, cur_c[k * PP_MAXENT + q], zmk((double)zr, (double)zi));
              }
            }
          } else {
            const double *Br = A.term[t].bands + (size_t)r * (2 * KD + 1) * N;
            double bv[2 * KD + 1];
#pragma unroll
            for (int d = -KD; d <= KD; d++)
              bv[d + KD] = (interior || (i + d >= 0 && i + d < N)) ? Br[(size_t)(d + KD) * N + i] : 0.0;
#pragma unroll
            for (int q = 0; q < PP_MAXENT; q++) {
              if (q < ne) {
                const zd *Yq = Ys + cur_q[k * PP_MAXENT + q] * ystr;
                zd z = zmk(0.0, 0.0);
#pragma unroll
                for (int d = -KD; d <= KD; d++) {
                  const int j = i + d;
                  const zd yv = (interior || (j >= 0 && j < N)) ? Yq[j - ybase] : zmk(0.0, 0.0);
                  zfmar(z, bv[d + KD], yv);
                }
                zfma(acc, cur_c[k * PP_MAXENT + q], z);
              }
            }
          }
        }
      }
    }
    X[i - r0] = acc;
  }
  else for (int i = r0 + tid; i < r0 + nloc; i += PP_THREADS) {
    zd acc = zmk(0.0, 0.0);
    if (i < N) {
      const bool interior = i >= KD && i + KD < N;
      int k = 0;
      for (int t = 0; t < A.nterms; t++) {
        const double *Bt = A.term[t].bands;
        for (int r = 0; r < A.term[t].rank; r++, k++) {
          const int ne = cur_n[k];
          if (ne == 0) continue;
          const double *Br = Bt + (size_t)r * (2 * KD + 1) * N;
          double bv[2 * KD + 1];
#pragma unroll
          for (int d = -KD; d <= KD; d++)
            bv[d + KD] = (interior || (i + d >= 0 && i + d < N)) ? Br[(size_t)(d + KD) * N + i] : 0.0;
#pragma unroll
          for (int q = 0; q < PP_MAXENT; q++) {
            if (q < ne) {
              const zd cq = cur_c[k * PP_MAXENT + q];
              const zd *Yq = Ys + cur_q[k * PP_MAXENT + q] * ystr;
              zd z = zmk(0.0, 0.0);
#pragma unroll
              for (int d = -KD; d <= KD; d++) {
                const int j = i + d;
                const zd yv = (interior || (j >= 0 && j < N)) ? Yq[j - ybase] : zmk(0.0, 0.0);
                zfmar(z, bv[d + KD], yv);
              }
              zfma(acc, cq, z);
            }
          }
        }
      }
    }
    X[i - r0] = acc;
  }
}

// W := fac * U^-H (sum_t s_t(t_step) V_t) y for fiber a, with y
// the exchanged U^-1 v of every fiber (A.y) staged for this row's couplings.
// The balanced-overlap matvec of BasisPropagator.py:131-159.  `stamp` is a
// profiling hook (after the apply and inside the solve).
template <int KD, bool MIXED, class Stamp>
__device__ __attribute__((always_inline)) inline void pp_matvec(const pp_pamp_args &A, const pp_lds_map &M, int a, int step, zd fac,
                                                                Stamp stamp)
{
  const int N = A.N, tid = threadIdx.x;
  // stage the source fibers U^-1 w of this row's couplings
  for (int q = 0; q < A.nslots; q++) {
    const int col = A.slotcols[a * A.nslots + q];
    if (col >= 0)
      for (int i = tid; i < N; i += PP_THREADS) M.Ys[q * N + i] = zld(A.y + (size_t)col * N + i);
  }
  pp_step_coeffs(A, M, step, fac);
  __syncthreads();
  pp_band_rows<KD, MIXED>(A, M, M.Ys, N, 0, 0, A.padN, M.X);
  __syncthreads();
  stamp();
  pp_band_solve<KD, false, PP_SOLVE_THREADS>(A.plan, M.nl, M.X, M.W, M.tb, stamp);
}

// MIXED: terms flagged fp32 are applied in complex FP32 (FP32 radial bands,
// FP32 band sums); Gram-Schmidt, the overlap solves, the exchanged U^-1 w and
// the Pade exponential stay FP64.
template <int KD, bool MIXED>
__device__ __attribute__((always_inline)) inline void pp_pamp_body(const pp_pamp_args &A, int a)
{
  extern __shared__ double lds_d[];
  const int N = A.N, L = A.L, m = A.m, tid = threadIdx.x;
  const int padN = A.padN;
  const pp_band_plan &pl = A.plan;
  const pp_lds_map M = pp_map_lds<KD>(A, lds_d);
  double *nl = M.nl;
  zd *W = M.W, *X = M.X, *tb = M.tb, *Ys = M.Ys, *Z = M.Z, *cv = M.cv;
  double *scr = M.scr, *wsum = M.wsum, *hred = M.hred, *hacc = M.hacc;
  const double tol = 1.4901161193847656e-08;           // sqrt(DBL_EPSILON), pamp.h:75
  zd *Va = A.V + (size_t)a * (m + 1) * N;
  double *mypart = A.part + (size_t)a * A.pstr;
  const int NRM = 2 * m, ONRM = 2 * m + 1;
  unsigned ep = 0;   // grid barriers passed in this launch
  int tcount = 0;
  auto stamp = [&](int step) {
    if (A.tlog && step == 0 && a == 0 && tid == 0 && tcount < PP_TLOG_MAX) A.tlog[tcount++] = (long long)wall_clock64();
  };
  pp_load_persistent(A, M, a);

  for (int step = 0; step < A.nsteps; step++) {
    // ---- psi~ = U psi (MultiplySqrtOverlap, PamPropagator.py:53) ; y = U^-1 psi~
    for (int i = tid; i < N; i += PP_THREADS) Ys[i] = A.psi[(size_t)a * N + i];
    __syncthreads();
    double nrm = 0.0;
    for (int i = tid; i < N; i += PP_THREADS) {
      zd acc = zmk(0.0, 0.0);
      const int j1 = i + KD < N - 1 ? i + KD : N - 1;
      for (int j = i; j <= j1; j++) zfmar(acc, A.Uband[(size_t)j * (KD + 1) + KD + i - j], Ys[j]);
      W[i] = acc;
      nrm += zabs2(acc);
    }
    nrm = block_sum(nrm, wsum);
    if (tid == 0) xst(mypart + NRM, nrm);
    for (int i = tid; i < padN; i += PP_THREADS) X[i] = i < N ? W[i] : zmk(0.0, 0.0);
    __syncthreads();
    pp_band_solve<KD, true, PP_SOLVE_THREADS>(pl, nl, X, Z, tb);
    for (int i = tid; i < N; i += PP_THREADS) zst(A.y + (size_t)a * N + i, Z[i]);
    stamp(step);
    grid_barrier(A.bar, ep, L, a);
    stamp(step);

    double beta0 = 0.0, beta = 0.0;
    int mEff = m;
    bool broken = false;
    int reorth_passes = 0;
    for (int j = 0; j < m; j++) {
      // ---------------- phase A: beta_j, v_j = w/beta, w' = U^-H H U^-1 v_j, <v_k, w'>
      const int nk = j + 1, ncomp = 2 * nk;
      grid_partial_sum(A.part, A.pstr, NRM, 1, L, hred, scr);
      beta = sqrt(hred[0]);
      if (j == 0) {
        beta0 = beta;
        // zero (or non-finite) start vector: no Arnoldi step; exp(A) 0 = 0,
        // and the host reports "Wavefunction is zero" from |psi~| = 0
        if (!(beta > 0.0 && beta <= 1.79e308)) { broken = true; mEff = 0; }
      }
      else if (!broken) {
        if (fabs(beta) < tol) { broken = true; mEff = j; }   // happy breakdown (pamp.h:305-312)
        else if (a == 0 && tid == 0) A.Hm[j + (j - 1) * m] = zmk(beta, 0.0);
      }
      if (!broken) {
        const double inv = 1.0 / beta;
        for (int i = tid; i < N; i += PP_THREADS) {
          zd v = zscale(W[i], inv);
          W[i] = v;
          Va[(size_t)j * N + i] = v;
        }
        pp_matvec<KD, MIXED>(A, M, a, step, zmk(inv, 0.0), [&]() { stamp(step); });
        stamp(step);
        fiber_dots(Va, N, W, nk, N, A.reorth != 0, hred);
        for (int c = tid; c < ncomp; c += PP_THREADS) xst(mypart + c, hred[c]);
        if (A.reorth && tid == 0) xst(mypart + ONRM, hred[ncomp]);
      }
      stamp(step);
      grid_barrier(A.bar, ep, L, a);
      stamp(step);

      // ---------------- phase B: h = sum_a partials ; w' -= V h ; |w'|^2 ; y = U^-1 w'
      if (!broken) {
        grid_partial_sum(A.part, A.pstr, 0, ncomp, L, hred, scr);
        if (a == 0 && tid <= j) A.Hm[tid + j * m] = zmk(hred[2 * tid], hred[2 * tid + 1]);
      }
      if (j == m - 1) break;   // the last residual is not used (pamp.h:419-426)
      if (!broken) {
        for (int i = tid; i < N; i += PP_THREADS) {
          zd w = W[i];
#pragma unroll PP_VUNROLL
          for (int k = 0; k <= j; k++) w = zsub(w, zmul(Va[(size_t)k * N + i], zmk(hred[2 * k], hred[2 * k + 1])));
          W[i] = w;
        }
        __syncthreads();
      }
      if (A.reorth) {
        // PerformOrthogonalization (pamp.h:353-400): repeat CGS while the
        // residual shrank below 0.717 of the previous norm, at most 3 times.
        double orig = 0.0;
        bool stop = broken;
        if (!broken) {
          for (int c = tid; c < ncomp; c += PP_THREADS) hacc[c] = hred[c];
          grid_partial_sum(A.part, A.pstr, ONRM, 1, L, hred, scr);
          orig = sqrt(hred[0]);
          __syncthreads();
        }
        for (int pass = 0; pass < 4; pass++) {
          if (!stop) {
            double v = 0.0;
            for (int i = tid; i < N; i += PP_THREADS) v += zabs2(W[i]);
            v = block_sum(v, wsum);
            if (tid == 0) xst(mypart + NRM, v);
          }
          grid_barrier(A.bar, ep, L, a);
          bool doit = false;
          double res = 0.0;
          if (!stop) {
            grid_partial_sum(A.part, A.pstr, NRM, 1, L, hred, scr);
            res = sqrt(hred[0]);
            __syncthreads();
            if (fabs(res) < 0.717 * fabs(orig) && pass < 3) doit = true;
            else stop = true;
          }
          if (doit) {
            fiber_dots(Va, N, W, nk, N, false, hred);
            for (int c = tid; c < ncomp; c += PP_THREADS) xst(mypart + c, hred[c]);
          }
          grid_barrier(A.bar, ep, L, a);
          if (doit) {
            grid_partial_sum(A.part, A.pstr, 0, ncomp, L, hred, scr);
            for (int i = tid; i < N; i += PP_THREADS) {
              zd w = W[i];
              for (int k = 0; k <= j; k++) w = zsub(w, zmul(Va[(size_t)k * N + i], zmk(hred[2 * k], hred[2 * k + 1])));
              W[i] = w;
            }
            __syncthreads();
            for (int c = tid; c < ncomp; c += PP_THREADS) hacc[c] += hred[c];
            __syncthreads();
            if (a == 0 && tid <= j) A.Hm[tid + j * m] = zmk(hacc[2 * tid], hacc[2 * tid + 1]);
            orig = res;
            reorth_passes++;
            __syncthreads();
          }
        }
        grid_barrier(A.bar, ep, L, a);   // everyone has read NRM before it is rewritten below
      }
      if (!broken) {
        double v = 0.0;
        for (int i = tid; i < N; i += PP_THREADS) v += zabs2(W[i]);
        v = block_sum(v, wsum);
        if (tid == 0) xst(mypart + NRM, v);
        for (int i = tid; i < padN; i += PP_THREADS) X[i] = i < N ? W[i] : zmk(0.0, 0.0);
        __syncthreads();
        stamp(step);
        pp_band_solve<KD, true, PP_SOLVE_THREADS>(pl, nl, X, Z, tb);
        for (int i = tid; i < N; i += PP_THREADS) zst(A.y + (size_t)a * N + i, Z[i]);
      }
      stamp(step);
      grid_barrier(A.bar, ep, L, a);
      stamp(step);
    }

    // ---------------- exponentiate on workgroup 0, broadcast exp(-i dt H) e1
    if (a == 0) {
      pade_block(m, A.Hm, A.dt_re, A.dt_im, A.nlevEven + PP_CMB_DOUBLES, A.cvec, A.diag + (size_t)step * 4,
                 A.normscale);
      if (tid == 0) {
        A.diag[(size_t)step * 4 + 1] = mEff;
        A.diag[(size_t)step * 4 + 2] = beta0;
        A.diag[(size_t)step * 4 + 3] = reorth_passes;
      }
      __syncthreads();
      for (int e = tid; e < m * m; e += PP_THREADS) A.Hm[e] = zmk(0.0, 0.0);
    }
    stamp(step);
    grid_barrier(A.bar, ep, L, a);
    stamp(step);
    // ---------------- psi~ = |psi~| V c ; psi = U^-1 psi~ (SolveSqrtOverlap)
    for (int k = tid; k < m; k += PP_THREADS) cv[k] = zld(A.cvec + k);
    __syncthreads();
    for (int i = tid; i < padN; i += PP_THREADS) {
      zd acc = zmk(0.0, 0.0);
      if (i < N)
        for (int k = 0; k < mEff; k++) zfma(acc, Va[(size_t)k * N + i], cv[k]);
      X[i] = zscale(acc, beta0);
    }
    __syncthreads();
    pp_band_solve<KD, true, PP_SOLVE_THREADS>(pl, nl, X, Z, tb);
    for (int i = tid; i < N; i += PP_THREADS) A.psi[(size_t)a * N + i] = Z[i];
    __syncthreads();
    stamp(step);
    // cvec is rewritten by the next step's Pade only after further barriers
  }
  if (A.tlog && a == 0 && tid == 0) A.tlog[PP_TLOG_MAX] = tcount;
}

// One instance: workgroup a owns fiber a.
template <int KD, bool MIXED>
__global__ void __launch_bounds__(PP_THREADS) pp_pamp_kernel(pp_pamp_args A)
{
  pp_pamp_body<KD, MIXED>(A, blockIdx.x);
}

// Many independent instances in one cooperative grid: workgroup w runs fiber
// map[w].y of instance map[w].x; every instance has its own arguments
// (sizes, operator, solver plan, buffers, barrier words) and synchronises
// only its own workgroups.
template <int KD, bool MIXED>
__global__ void __launch_bounds__(PP_THREADS) pp_pamp_batch_kernel(const pp_pamp_args *args, const int2 *map)
{
  const int2 w = map[blockIdx.x];
  const pp_pamp_args A = args[w.x];
  pp_pamp_body<KD, MIXED>(A, w.y);
}

// ------------------------------------------------------------------ fiber split
// Split form of the step kernel for problems with fewer fibers than CUs:
// Q workgroups share each fiber, workgroup (a, q) owning the solver chunks
// c0 = q*P/Q .. (q+1)*P/Q - 1 of fiber a (rows r0 = c0*B onwards), so every
// CU streams its share of the Krylov basis.  A band solve becomes: stage 1
// on the own chunks; their boundary values to the fiber's exchange slot
// (sb, double-buffered by solve parity); a barrier among the fiber's Q
// workgroups; the log-depth scan over all P chunks, repeated by each of
// them; stage 3 on the own chunks.  The operator apply stages rows
// r0-KD .. r1+KD-1 of the coupled fibers; inner products and norms are
// partial over own rows and summed over all L*Q workgroups in one fixed
// order.  Same algorithm and barrier discipline as pp_pamp_body, without
// re-orthogonalisation (the launcher keeps that form on pp_pamp_kernel).

// Workgroup b -> (fiber a, part q).  Workgroups are dispatched round-robin
// over the 8 XCDs, so when 8 divides the grid the parts of one fiber (and
// neighbouring fibers) are given blocks b = x + 8s with one x: they share
// one XCD's L2 for the per-fiber barrier and the boundary exchange.
__device__ inline void pp_split_map(int b, int L, int Q, int &a, int &q)
{
  const int nw = L * Q;
  const int idx = (nw & 7) == 0 ? (b & 7) * (nw >> 3) + (b >> 3) : b;
  a = idx / Q;
  q = idx - a * Q;
}

// out[c] = sum_{w < nw} part[w*pstr + off + c] for c < ncomp, in one fixed
// order (identical in every workgroup): G = a power of two <= 64 lanes per
// component each sum the partials w = g, g+G, ... in order, then a
// butterfly over the G lanes.
__device__ inline void grid_partial_sum_wide(const double *part, int pstr, int off, int ncomp, int nw,
                                             double *out_lds)
{
  const int tid = threadIdx.x;
  int G = 64;
  while (G > 1 && G * ncomp > PP_THREADS) G >>= 1;
  const int c = tid / G, g = tid - c * G;
  double s = 0.0;
  if (c < ncomp) {
#pragma unroll 8
    for (int w = g; w < nw; w += G) s += xld(part + (size_t)w * pstr + off + c);
  }
  for (int o = G >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (c < ncomp && g == 0) out_lds[c] = s;
  __syncthreads();
}

// x = U^-1 b (UPPER) or U^-T b over the own chunks c0 .. c0+nch-1 of one
// fiber (in/out: those chunks' rows, LDS; tb: nch*KD complex; bb: P*KD
// complex).  All Q workgroups of the fiber call it in the same sequence.
// Stage 1 on the own chunks; their boundary values to the fiber's exchange
// slot and a barrier among the fiber's workgroups; then, instead of the
// log-depth scan, the boundary term every own chunk needs from its
// neighbour (x_{c+1} for U^-1, x_{c-1} for U^-T) as one dot product with a
// precomputed row over the other chunks' boundary values: the scan's
// recurrence x_c = b_c + N_c x_{c+1} (x_c = b_c + N'_c x_{c-1}) has the
// solution x_c = sum_{c' >= c} N_c N_{c+1} .. N_{c'-1} b_{c'}, and the
// workgroup keeps the rows of those products for its own chunks (srows,
// pamp_split_rows).  Each row's terms are added by SEG lanes, lane s
// taking columns s, s+SEG, ..., then a fixed butterfly over the SEG lanes.
// Stage 3 on the own chunks.  Ends with a __syncthreads.
template <int KD, bool UPPER>
__device__ __attribute__((always_inline)) inline void pp_split_solve(const pp_pamp_args &A, const double *rows,
                                                                     const zd *in, zd *out, zd *tb, zd *bb, int c0,
                                                                     int nch, zd *sbuf, unsigned *fw, unsigned &sep,
                                                                     int q)
{
  const pp_band_plan &pl = A.plan;
  const int tid = threadIdx.x, lb = pl.lb, off = UPPER ? 0 : pl.B - KD, P = pl.P;
  if (tid < PP_SOLVE_THREADS) pp_solve_blocks<UPPER, PP_SOLVE_THREADS>(pl, in, out, c0, nch);
  __syncthreads();
  for (int e = tid; e < nch * KD; e += PP_THREADS) {
    const int cl = e / KD, k = e - cl * KD;
    zst(sbuf + (size_t)(c0 + cl) * KD + k, out[(cl << lb) + off + k]);
  }
  group_barrier(fw, &A.bar->err, sep, A.Q, q);
  for (int e = tid; e < P * KD; e += PP_THREADS) {
    const int c = e / KD, k = e - c * KD;
    bb[e] = (c >= c0 && c < c0 + nch) ? out[((c - c0) << lb) + off + k] : zld(sbuf + e);
  }
  __syncthreads();
  const int nr = nch * KD;
  int SEG = 64;
  while (SEG > 1 && SEG * nr > PP_THREADS) SEG >>= 1;
  const int r = tid / SEG, sg = tid - r * SEG;
  zd acc = zmk(0.0, 0.0);
  if (r < nr) {
    const int cl = r / KD, i = r - cl * KD, c = c0 + cl;
    // this part's rows: the U^-1 block (row length (P-c-1)*KD, columns from
    // chunk c+1 on), then the U^-T block (row length c*KD, columns from chunk 0)
    const int kk = KD * KD;
    const int ubase = kk * (cl * (P - c0 - 1) - cl * (cl - 1) / 2);
    const int lbase = kk * (nch * (P - c0 - 1) - nch * (nch - 1) / 2) + kk * (cl * c0 + cl * (cl - 1) / 2);
    const int len = UPPER ? (P - c - 1) * KD : c * KD;
    const double *row = rows + (UPPER ? ubase : lbase) + i * len;
    const zd *b = UPPER ? bb + (c + 1) * KD : bb;
#pragma unroll 8
    for (int j = sg; j < len; j += SEG) zfmar(acc, row[j], b[j]);
  }
  for (int o = SEG >> 1; o > 0; o >>= 1) {
    acc.x += __shfl_xor(acc.x, o, 64);
    acc.y += __shfl_xor(acc.y, o, 64);
  }
  if (r < nr && sg == 0) tb[r] = acc;
  __syncthreads();
  if (tid < PP_SOLVE_THREADS) pp_solve_fixup<KD, UPPER, PP_SOLVE_THREADS>(pl, tb, out, c0, nch);
  __syncthreads();
}

// LDS doubles of the split kernel's per-workgroup buffers above the
// persistent chunk-boundary rows and combo table (maxch = chunks per
// workgroup, rounded up; P chunks per fiber): W, X (maxch*B each), tb, all
// chunks' boundary values (P*kd), the staged coupled rows, the reduction
// scratch and exp(-i dt H) e1.
__host__ __device__ inline int pp_split_doubles(int nslots, int m, int kd, int B, int maxch, int P)
{
  const int R = maxch * B;
  return 2 * (2 * R + maxch * kd + P * kd + nslots * (R + 2 * kd)) + 16 + (2 * m + 4) + 2 * m;
}

// Doubles of one part's chunk-boundary rows (kd^2 (P-1) per chunk), even.
__host__ __device__ inline int pp_split_rows_doubles(int kd, int P, int maxch)
{
  const int n = kd * kd * (P - 1) * maxch;
  return n + (n & 1);
}

// Hand-over of U^-1 v between the split kernels' workgroups without a grid
// barrier: every workgroup counts its productions of U^-1 v (ep) and posts
// the count in its word of `flag` after the values; a consumer waits until
// the producers of the rows it stages (at most nslots * Q of them, one per
// lane of wavefront 0) have posted the same count.  A producer overwrites
// its rows only after a grid barrier that every consumer reaches after
// staging them.
struct pp_yx {
  unsigned *flag;
  unsigned ep;
  int prod, wid;
};

// rows r0 .. r0+nloc-1 of fiber a staged with a halo of KD rows
__device__ inline pp_yx pp_yx_init(const pp_pamp_args &A, int a, int wid, int nwg, int r0, int nloc, int KD)
{
  const int N = A.N, Q = A.Q, P = A.plan.P, lb = A.plan.lb, tid = threadIdx.x;
  pp_yx y;
  y.flag = A.subw + (size_t)nwg;
  y.ep = 0;
  y.prod = -1;
  y.wid = wid;
  const int ylo = r0 - KD > 0 ? r0 - KD : 0, yhi = r0 + nloc + KD < N ? r0 + nloc + KD : N;
  int cnt = 0;
  for (int sl = 0; sl < A.nslots; sl++) {
    const int col = A.slotcols[a * A.nslots + sl];
    if (col < 0) continue;
    for (int q2 = 0; q2 < Q; q2++) {
      const int rlo = (q2 * P / Q) << lb, rhi0 = ((q2 + 1) * P / Q) << lb, rhi = rhi0 < N ? rhi0 : N;
      if (rlo < yhi && rhi > ylo) {
        if (cnt == tid) y.prod = col * Q + q2;
        cnt++;
      }
    }
  }
  return y;
}

// after this workgroup's stores of U^-1 v (all threads call it)
__device__ inline void pp_yx_publish(pp_yx &y)
{
  y.ep++;
  __builtin_amdgcn_s_waitcnt(0);   // this wave's stores are acknowledged before the count
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __hip_atomic_store(y.flag + y.wid, y.ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// until the producers have posted this workgroup's own count (all threads call it)
__device__ inline void pp_yx_wait(const pp_yx &y, unsigned *err)
{
  if (threadIdx.x < 64) {
    long spins = 0;
    while (true) {
      const bool ok =
          y.prod < 0 || (int)(__hip_atomic_load(y.flag + y.prod, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - y.ep) >= 0;
      if (__all(ok)) break;
      __builtin_amdgcn_s_sleep(1);
      if (++spins > (1l << 24)) {
        if (threadIdx.x == 0) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  __syncthreads();
}

template <int KD, bool MIXED>
__global__ void __launch_bounds__(PP_THREADS) pp_pamp_split_kernel(pp_pamp_args A)
{
  extern __shared__ double lds_d[];
  const int N = A.N, L = A.L, m = A.m, Q = A.Q, tid = threadIdx.x;
  const pp_band_plan &pl = A.plan;
  const int P = pl.P, lb = pl.lb;
  int a, q;
  pp_split_map(blockIdx.x, L, Q, a, q);
  const int c0 = q * P / Q, nch = (q + 1) * P / Q - c0, maxch = (P + Q - 1) / Q;
  const int r0 = c0 << lb, nloc = nch << lb, R = maxch << lb;
  const int nown = (r0 + nloc < N ? r0 + nloc : N) - r0;   // own rows below N (>= 1: only the last chunk pads)
  const int ystr = R + 2 * KD;
  pp_lds_map M = pp_map_lds<KD>(A, lds_d);
  M.W = (zd *)(lds_d + A.nlevEven + PP_CMB_DOUBLES);
  M.X = M.W + R;
  M.tb = M.X + R;
  zd *bb = M.tb + maxch * KD;   // all chunks' boundary values of a split solve
  M.Ys = bb + P * KD;
  M.Z = M.Ys;
  M.wsum = (double *)(M.Ys + A.nslots * ystr);
  M.hred = M.wsum + 16;
  M.cv = (zd *)(M.hred + (2 * m + 4));
  zd *W = M.W, *X = M.X, *tb = M.tb, *Ys = M.Ys, *Z = M.Z, *cv = M.cv;
  // own rows of the first KC Krylov vectors, kept in LDS as well (stride R)
  // for the inner products and the projection; the Pade exponential of the
  // lead workgroup reuses this space, so the final combination reads them
  // back from global memory
  zd *Vl = cv + m;
  const int KC = A.vcache;
  double *wsum = M.wsum, *hred = M.hred;
  const double tol = 1.4901161193847656e-08;           // sqrt(DBL_EPSILON), pamp.h:75
  zd *Va = A.V + (size_t)a * (m + 1) * N;
  const int nwg = L * Q, wid = a * Q + q;
  unsigned *fw = A.subw + (size_t)a * Q;
  zd *sbase = A.sb + (size_t)a * 2 * P * KD;
  const bool lead = a == 0 && q == 0;
  unsigned ep = 0, sep = 0;   // grid and per-fiber barriers passed
  int nsolve = 0;             // split solves done (selects the boundary buffer)
  auto sbuf = [&]() { return sbase + (size_t)((nsolve++) & 1) * P * KD; };
  int tcount = 0;
  auto stamp = [&](int step) {
    if (A.tlog && step == 0 && lead && tid == 0 && tcount < PP_TLOG_MAX) A.tlog[tcount++] = (long long)wall_clock64();
  };
  pp_load_persistent(A, M, a, A.srows + (size_t)q * A.sstride, A.sstride);

  // U^-1 u of the coupled fibers arrives through production counts (pp_yx)
  pp_yx yx = pp_yx_init(A, a, wid, nwg, r0, nloc, KD);
  const int HB = 2 * m + 3;   // partial sums per iteration; two buffers alternate by iteration parity

  for (int step = 0; step < A.nsteps; step++) {
    // ---- u_0 = psi~ = U psi on the own rows (psi rows up to r0+nloc+KD-1,
    // written by the neighbouring part at the end of the previous step) ;
    // y_0 = U^-1 u_0
    if (step > 0) group_barrier(fw, &A.bar->err, sep, Q, q);
    const int phi = r0 + nloc + KD < N ? r0 + nloc + KD : N;
    for (int j = r0 + tid; j < phi; j += PP_THREADS) Ys[j - r0] = zld(A.psi + (size_t)a * N + j);
    __syncthreads();
    double nrm = 0.0;
    for (int i = tid; i < nloc; i += PP_THREADS) {
      const int gi = r0 + i;
      zd acc = zmk(0.0, 0.0);
      if (gi < N) {
        const int j1 = gi + KD < N - 1 ? gi + KD : N - 1;
        for (int j = gi; j <= j1; j++) zfmar(acc, A.Uband[(size_t)j * (KD + 1) + KD + gi - j], Ys[j - r0]);
      }
      W[i] = acc;
      X[i] = acc;
      nrm += zabs2(acc);
    }
    double unrm = block_sum(nrm, wsum);   // this workgroup's part of |u_j|^2
    pp_split_solve<KD, true>(A, M.nl, X, Z, tb, bb, c0, nch, sbuf(), fw, sep, q);
    for (int i = tid; i < nown; i += PP_THREADS) zst(A.y + (size_t)a * N + r0 + i, Z[i]);
    pp_yx_publish(yx);
    stamp(step);

    double beta0 = 0.0;
    int mEff = m;
    pp_step_coeffs(A, M, step, zmk(1.0, 0.0));   // the apply's coefficients at t_step (read after a barrier below)
    for (int j = 0; j < m; j++) {
      // ---- Arnoldi iteration j with the normalisation deferred by one
      // reduction (pamp.h:293-350: v_j = w / |w|, w = H^ v_j, h = V^H w,
      // w -= V h).  W holds u_j, the unnormalised residual of iteration j-1
      // (u_0 = psi~), and the coupled fibers' U^-1 u_j have been posted.
      // z = H^ u_j = U^-H H U^-1 u_j; one grid reduction of <v_k, z>
      // (k < j), <u_j, z> and |u_j|^2 gives beta_j = |u_j|, v_j = u_j /
      // beta_j, h_kj = <v_k, z> / beta_j, h_jj = <u_j, z> / beta_j^2 and
      // u_{j+1} = z / beta_j - sum_k h_kj v_k: the same quantities with H^
      // applied to u_j instead of v_j (moves results by rounding only), and
      // one grid barrier per iteration.
      const int ncomp = 2 * j + 3;
      double *mypart = A.part + (size_t)wid * A.pstr + (j & 1) * HB;
      pp_yx_wait(yx, &A.bar->err);
      const int ybase = r0 - KD, ylo = ybase > 0 ? ybase : 0;
      for (int sl = 0; sl < A.nslots; sl++) {
        const int col = A.slotcols[a * A.nslots + sl];
        if (col >= 0)
          for (int jj = ylo + tid; jj < phi; jj += PP_THREADS) Ys[sl * ystr + jj - ybase] = zld(A.y + (size_t)col * N + jj);
      }
      __syncthreads();
      pp_band_rows<KD, MIXED>(A, M, Ys, ystr, ybase, r0, nloc, X);
      __syncthreads();
      stamp(step);
      pp_split_solve<KD, false>(A, M.nl, X, Z, tb, bb, c0, nch, sbuf(), fw, sep, q);   // z in Z (over Ys)
      stamp(step);
      if (j > 0) split_dots(Vl, R, KC, Va + r0, N, Z, j, nown, hred);
      double ure = 0.0, uim = 0.0;
      for (int i = tid; i < nown; i += PP_THREADS) {
        const zd u = W[i], zz = Z[i];
        ure = fma(u.x, zz.x, ure); ure = fma(u.y, zz.y, ure);
        uim = fma(u.x, zz.y, uim); uim = fma(-u.y, zz.x, uim);
      }
      ure = block_sum(ure, wsum);
      uim = block_sum(uim, wsum);
      if (tid == 0) { hred[2 * j] = ure; hred[2 * j + 1] = uim; hred[2 * j + 2] = unrm; }
      __syncthreads();
      for (int c = tid; c < ncomp; c += PP_THREADS) xst(mypart + c, hred[c]);
      stamp(step);
      grid_barrier(A.bar, ep, nwg, wid);
      stamp(step);
      grid_partial_sum_wide(A.part + (j & 1) * HB, A.pstr, 0, ncomp, nwg, hred);
      const double bj = sqrt(hred[2 * j + 2]);
      if (j == 0) {
        beta0 = bj;
        // zero (or non-finite) start vector: no Arnoldi step; exp(A) 0 = 0,
        // and the host reports "Wavefunction is zero" from |psi~| = 0
        if (!(bj > 0.0 && bj <= 1.79e308)) { mEff = 0; break; }
      } else {
        if (fabs(bj) < tol) { mEff = j; break; }   // happy breakdown (pamp.h:305-312)
        if (lead && tid == 0) A.Hm[j + (j - 1) * m] = zmk(bj, 0.0);
      }
      const double inv = 1.0 / bj, inv2 = inv * inv;
      if (lead && tid <= j) {
        const double f = tid < j ? inv : inv2;
        A.Hm[tid + j * m] = zmk(hred[2 * tid] * f, hred[2 * tid + 1] * f);
      }
      for (int i = tid; i < nown; i += PP_THREADS) {
        const zd v = zscale(W[i], inv);
        Va[(size_t)j * N + r0 + i] = v;
        if (j < KC) Vl[(size_t)j * R + i] = v;
      }
      if (j == m - 1) break;   // the last residual is not used (pamp.h:419-426)
      double v2 = 0.0;
      for (int i = tid; i < nloc; i += PP_THREADS) {
        zd w = zmk(0.0, 0.0);
        if (i < nown) {
          w = zscale(Z[i], inv);
          const int kl = j + 1 < KC ? j + 1 : KC;   // LDS copies first, same k order
#pragma unroll PP_VUNROLL
          for (int k = 0; k < kl; k++) {
            const double f = k < j ? inv : inv2;
            w = zsub(w, zmul(Vl[(size_t)k * R + i], zmk(hred[2 * k] * f, hred[2 * k + 1] * f)));
          }
#pragma unroll PP_VUNROLL
          for (int k = kl; k <= j; k++) {
            const double f = k < j ? inv : inv2;
            w = zsub(w, zmul(Va[(size_t)k * N + r0 + i], zmk(hred[2 * k] * f, hred[2 * k + 1] * f)));
          }
        }
        W[i] = w;
        X[i] = w;
        v2 += zabs2(w);
      }
      unrm = block_sum(v2, wsum);
      stamp(step);
      pp_split_solve<KD, true>(A, M.nl, X, Z, tb, bb, c0, nch, sbuf(), fw, sep, q);
      for (int i = tid; i < nown; i += PP_THREADS) zst(A.y + (size_t)a * N + r0 + i, Z[i]);
      pp_yx_publish(yx);
      stamp(step);
    }

    // ---------------- exponentiate on the lead workgroup, broadcast exp(-i dt H) e1
    if (lead) {
      pade_block(m, A.Hm, A.dt_re, A.dt_im, A.nlevEven + PP_CMB_DOUBLES, A.cvec, A.diag + (size_t)step * 4,
                 A.normscale);
      if (tid == 0) {
        A.diag[(size_t)step * 4 + 1] = mEff;
        A.diag[(size_t)step * 4 + 2] = beta0;
        A.diag[(size_t)step * 4 + 3] = 0;
      }
      __syncthreads();
      for (int e = tid; e < m * m; e += PP_THREADS) A.Hm[e] = zmk(0.0, 0.0);
    }
    stamp(step);
    grid_barrier(A.bar, ep, nwg, wid);
    stamp(step);
    // ---------------- psi~ = |psi~| V c ; psi = U^-1 psi~ on the own rows
    for (int k = tid; k < m; k += PP_THREADS) cv[k] = zld(A.cvec + k);
    __syncthreads();
    for (int i = tid; i < nloc; i += PP_THREADS) {
      zd acc = zmk(0.0, 0.0);
      if (i < nown)
        for (int k = 0; k < mEff; k++) zfma(acc, Va[(size_t)k * N + r0 + i], cv[k]);
      X[i] = zscale(acc, beta0);
    }
    __syncthreads();
    pp_split_solve<KD, true>(A, M.nl, X, Z, tb, bb, c0, nch, sbuf(), fw, sep, q);
    for (int i = tid; i < nown; i += PP_THREADS) zst(A.psi + (size_t)a * N + r0 + i, Z[i]);
    __syncthreads();
    stamp(step);
  }
  if (A.tlog && lead && tid == 0) A.tlog[PP_TLOG_MAX] = tcount;
}

// ------------------------------------------------------------------ Expokit
// zgpadm (Expokit, Sidje 1998; core/krylov/expokit/f2c/zgpadm.cpp:28-285):
// exp(t H) of the leading mx x mx block of Hm (column-major, leading
// dimension mh, LDS), irreducible Pade(6) with ns = max(0, int(log2 ||tH||_inf)
// + 2) squarings.  Writes the first column to ecol (LDS) and returns ns, or
// -1 for a null H or a singular denominator.  Scratch: 5 mx^2 + 5 mx complex
// and PP_THREADS/64 doubles at double offset `off` of the dynamic LDS.  The
// solve of zgesv is an in-place Gauss-Jordan with the same partial pivoting.
__device__ __attribute__((noinline)) int pp_zgpadm_block(int mx, int mh, const zd *Hm, double t, int off, zd *ecol)
{
  extern __shared__ double pp_lds[];
  const int tid = threadIdx.x, mm = mx * mx;
  zd *Hs = (zd *)(pp_lds + off), *P = Hs + mm, *Q = P + mm, *H2 = Q + mm, *F = H2 + mm;
  zd *G = H2;                                   // [Q - P | P], mx x 2mx, after H2 and F are dead
  zd *rk = F + mm, *rowk = rk + 2 * mx, *cm = rowk + 2 * mx;
  double *red = (double *)(cm + mx);
  const int i0 = tid % mx, j0 = tid / mx, di = PP_THREADS % mx, dj = PP_THREADS / mx;
  for (int e = tid; e < mm; e += PP_THREADS) Hs[e] = Hm[(e % mx) + (e / mx) * mh];
  __syncthreads();
  // infinity norm
  double rs = 0.0;
  if (tid < mx)
    for (int j = 0; j < mx; j++) rs += sqrt(zabs2(Hs[tid + j * mx]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) rs = fmax(rs, __shfl_xor(rs, o, 64));
  if ((tid & 63) == 0) red[tid >> 6] = rs;
  __syncthreads();
  double hnorm = red[0];
  for (int w = 1; w < PP_THREADS / 64; w++) hnorm = fmax(hnorm, red[w]);
  hnorm = fabs(t * hnorm);
  if (hnorm == 0.0) return -1;
  int ns = (int)(log(hnorm) / log(2.0)) + 2;
  if (ns < 0) ns = 0;
  const double scale = t / pow(2.0, (double)ns), scale2 = scale * scale;
  const int ideg = 6;
  double coef[ideg + 1];
  coef[0] = 1.0;
  for (int k = 1; k <= ideg; k++) coef[k] = coef[k - 1] * (double)(ideg + 1 - k) / (double)(k * (2 * ideg + 1 - k));
  auto matmul = [&](const zd *X, const zd *Y, zd *Z, double sc) {   // Z = sc X Y
    int i = i0, j = j0;
    for (int e = tid; e < mm; e += PP_THREADS) {
      const zd *xr = X + i, *yc = Y + j * mx;
      zd a0 = zmk(0.0, 0.0), a1 = zmk(0.0, 0.0), a2 = zmk(0.0, 0.0), a3 = zmk(0.0, 0.0);
      int k = 0;
      for (; k + 3 < mx; k += 4) {
        const zd x0 = xr[k * mx], x1 = xr[(k + 1) * mx], x2 = xr[(k + 2) * mx], x3 = xr[(k + 3) * mx];
        const zd y0 = yc[k], y1 = yc[k + 1], y2 = yc[k + 2], y3 = yc[k + 3];
        zfma(a0, x0, y0);
        zfma(a1, x1, y1);
        zfma(a2, x2, y2);
        zfma(a3, x3, y3);
      }
      for (; k < mx; k++) zfma(a0, xr[k * mx], yc[k]);
      Z[e] = zscale(zadd(zadd(a0, a1), zadd(a2, a3)), sc);
      i += di; j += dj;
      if (i >= mx) { i -= mx; j++; }
    }
    __syncthreads();
  };
  matmul(Hs, Hs, H2, scale2);
  {
    int i = i0, j = j0;
    for (int e = tid; e < mm; e += PP_THREADS) {
      P[e] = zmk(i == j ? coef[ideg - 1] : 0.0, 0.0);
      Q[e] = zmk(i == j ? coef[ideg] : 0.0, 0.0);
      i += di; j += dj;
      if (i >= mx) { i -= mx; j++; }
    }
  }
  __syncthreads();
  int iodd = 1;
  for (int k = ideg - 1; k > 0; k--) {
    zd *U = iodd ? Q : P;
    matmul(U, H2, F, 1.0);
    int i = i0, j = j0;
    for (int e = tid; e < mm; e += PP_THREADS) {
      zd v = F[e];
      if (i == j) v.x += coef[k - 1];
      U[e] = v;
      i += di; j += dj;
      if (i >= mx) { i -= mx; j++; }
    }
    __syncthreads();
    iodd = 1 - iodd;
  }
  zd *Uo = iodd ? Q : P;
  matmul(Uo, Hs, F, scale);
  for (int e = tid; e < mm; e += PP_THREADS) Uo[e] = F[e];
  __syncthreads();
  for (int e = tid; e < mm; e += PP_THREADS) {
    G[e] = zsub(Q[e], P[e]);
    G[mm + e] = P[e];
  }
  __syncthreads();
  // in-place Gauss-Jordan on G: the pivot row, the old row k and the
  // multipliers go to small buffers first, so every element update reads
  // only its own entry and the buffers
  const int w2 = 2 * mm;
  for (int k = 0; k < mx; k++) {
    int p = mx;
    double best = -1.0;
    for (int i = k + (tid & 63); i < mx; i += 64) {
      const double v = fabs(G[i + k * mx].x) + fabs(G[i + k * mx].y);
      if (v > best) { best = v; p = i; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int op = __shfl_xor(p, o, 64);
      if (ob > best || (ob == best && op < p)) { best = ob; p = op; }
    }
    if (best == 0.0) return -1;
    const zd piv = G[p + k * mx];
    const double dn = zabs2(piv);
    const zd pinv = zmk(piv.x / dn, -piv.y / dn);
    if (tid < 2 * mx) {
      rowk[tid] = G[k + tid * mx];
      rk[tid] = zmul(G[p + tid * mx], pinv);
    }
    if (tid < mx) {
      const int si = (tid == p) ? k : tid;
      cm[tid] = (tid == k) ? zmk(0.0, 0.0) : G[si + k * mx];
    }
    __syncthreads();
    int i = i0, j = j0;
    for (int e = tid; e < w2; e += PP_THREADS) {
      if (j >= k) {
        if (i == k) G[e] = rk[j];
        else {
          const zd src = (i == p) ? rowk[j] : G[e];
          G[e] = zsub(src, zmul(cm[i], rk[j]));
        }
      }
      i += di; j += dj;
      if (i >= mx) { i -= mx; j++; }
    }
    __syncthreads();
  }
  // E = 2X + I (X = right half of G), sign / squarings as zgpadm
  if (ns == 0) {
    if (tid < mx) {
      zd v = zscale(G[mm + tid], 2.0);
      if (tid == 0) v.x += 1.0;
      if (iodd) v = zmk(-v.x, -v.y);
      ecol[tid] = v;
    }
    __syncthreads();
    return 0;
  }
  {
    int i = i0, j = j0;
    for (int e = tid; e < mm; e += PP_THREADS) {
      zd v = zscale(G[mm + e], 2.0);
      if (i == j) v.x += 1.0;
      P[e] = v;
      i += di; j += dj;
      if (i >= mx) { i -= mx; j++; }
    }
  }
  __syncthreads();
  for (int k = 0; k < ns; k++) {
    matmul(P, P, Q, 1.0);
    for (int e = tid; e < mm; e += PP_THREADS) P[e] = Q[e];
    __syncthreads();
  }
  if (tid < mx) ecol[tid] = P[tid];
  __syncthreads();
  return ns;
}

// x_roundstep of zgexpv (zgexpv.cpp:297-310): keep two significant digits.
__device__ inline double pp_roundstep(double t)
{
  const double e = log10(t) - sqrt(0.1);
  const int ie = (int)(e >= 0 ? floor(e + 0.5) : -floor(0.5 - e)) - 1;
  const double p1 = pow(10.0, (double)ie);
  return trunc(t / p1 + 0.55) * p1;
}

struct pp_expo_args {
  double tol, anorm;   // krylov_tolerance (<= eps: sqrt(eps)), krylov_norm
  double t_out;        // |dt|
  int imag;            // imaginary time: A = -H^ (else A = -i H^)
  double *stats;       // nsteps x 6: nmult, nexph, nstep, nreject, mbrkdwn, iflag
};

// nsteps consecutive ExpokitPropagator.AdvanceStep's (expokitpropagator.cpp:
// 36-103 inside KrylovPropagator.py:41-45): psi~ = U psi; w = exp(t A) psi~ by
// zgexpv (zgexpv.cpp:32-586: adaptive substeps, m+1 matvecs each, 2-corrected
// error estimate, step rejection, at most 500 substeps); psi = U^-1 w.  The
// Arnoldi basis is orthogonalised by classical Gram-Schmidt with one
// re-orthogonalisation pass (two grid reductions) in place of zgexpv's
// modified Gram-Schmidt (one reduction per basis vector); the two agree in
// exact arithmetic.  Every workgroup holds the Hessenberg matrix and runs the
// exponential itself on identical, identically ordered data, so the step
// control takes the same branches in every workgroup without a broadcast.
template <int KD, bool MIXED>
__global__ void __launch_bounds__(PP_THREADS) pp_expokit_kernel(pp_pamp_args A, pp_expo_args E)
{
  extern __shared__ double lds_d[];
  const int N = A.N, L = A.L, m = A.m, a = blockIdx.x, tid = threadIdx.x, padN = A.padN, mh = m + 2;
  const pp_band_plan &pl = A.plan;
  const int hm_doubles = 2 * (mh * mh + mh);
  const pp_lds_map M = pp_map_lds<KD>(A, lds_d, hm_doubles);
  zd *Hm = (zd *)(lds_d + A.nlevEven + PP_CMB_DOUBLES);   // (m+2)^2, column-major
  zd *ecol = Hm + mh * mh;                                 // first column of exp(t H)
  double *nl = M.nl;
  zd *W = M.W, *X = M.X, *tb = M.tb, *Z = M.Z;
  double *scr = M.scr, *wsum = M.wsum, *hred = M.hred, *hacc = M.hacc;
  zd *Va = A.V + (size_t)a * (m + 1) * N;
  double *mypart = A.part + (size_t)a * A.pstr;
  const int D1 = 0, D2 = 2 * m, NRM = 4 * m + 2, ANRM = 4 * m + 3, P0NRM = 4 * m + 4;
  const int pade_off = A.nlevEven + PP_CMB_DOUBLES + hm_doubles;
  unsigned ep = 0;
  pp_load_persistent(A, M, a);
  const zd fac = E.imag ? zmk(-1.0, 0.0) : zmk(0.0, -1.0);
  const double eps = 2.220446049250313e-16;
  const double tol = E.tol <= eps ? sqrt(eps) : E.tol;
  const double rndoff = eps * E.anorm, break_tol = 1e-7, t_out = E.t_out;
  auto reduce1 = [&](int slot) {
    grid_partial_sum(A.part, A.pstr, slot, 1, L, hred, scr);
    return hred[0];
  };
  auto exchange_y = [&]() {   // X holds v (padded); y = U^-1 v to the other fibers
    __syncthreads();
    pp_band_solve<KD, true, PP_SOLVE_THREADS>(pl, nl, X, Z, tb);
    for (int i = tid; i < N; i += PP_THREADS) zst(A.y + (size_t)a * N + i, Z[i]);
  };
  for (int step = 0; step < A.nsteps; step++) {
    // ---- psi~ = U psi into W, |psi~|^2
    for (int i = tid; i < N; i += PP_THREADS) Z[i] = A.psi[(size_t)a * N + i];
    __syncthreads();
    double nrm = 0.0;
    for (int i = tid; i < N; i += PP_THREADS) {
      zd acc = zmk(0.0, 0.0);
      const int j1 = i + KD < N - 1 ? i + KD : N - 1;
      for (int j = i; j <= j1; j++) zfmar(acc, A.Uband[(size_t)j * (KD + 1) + KD + i - j], Z[j]);
      W[i] = acc;
      nrm += zabs2(acc);
    }
    nrm = block_sum(nrm, wsum);
    if (tid == 0) xst(mypart + P0NRM, nrm);
    grid_barrier(A.bar, ep, L, a);
    double beta = sqrt(reduce1(P0NRM));
    int nmult = 0, nexph = 0, nstep = 0, nreject = 0, mbrkdwn = m, k1 = 2, iflag = 0;
    // zero (or non-finite) psi: no substep (w = 0 stays 0), iflag -10; every
    // workgroup holds the same beta, so all of them skip alike
    if (!(beta > 0.0 && beta <= 1.79e308)) iflag = -10;
    double t_now = 0.0, err_loc = 0.0, t_step = 0.0, avnorm = 0.0, xmloc = 1.0 / m;
    const double p2 = tol * pow((m + 1) / 2.72, (double)(m + 1)) * sqrt((m + 1) * 6.28);
    double t_new = pp_roundstep((1.0 / E.anorm) * pow(p2 / (4.0 * beta * E.anorm), 1.0 / m));
    while (!iflag && t_now < t_out) {
      nstep++;
      t_step = fmin(t_out - t_now, t_new);
      const double inv = 1.0 / beta;
      for (int i = tid; i < padN; i += PP_THREADS) {
        const zd v = i < N ? zscale(W[i], inv) : zmk(0.0, 0.0);
        if (i < N) Va[i] = v;
        X[i] = v;
      }
      for (int e = tid; e < mh * mh; e += PP_THREADS) Hm[e] = zmk(0.0, 0.0);
      exchange_y();
      grid_barrier(A.bar, ep, L, a);
      k1 = 2;
      mbrkdwn = m;
      for (int j = 1; j <= m; j++) {
        nmult++;
        pp_matvec<KD, MIXED>(A, M, a, step, fac, pp_nostamp());
        // first pass: h1 = V^H z
        fiber_dots(Va, N, W, j, N, false, hred);
        for (int c = tid; c < 2 * j; c += PP_THREADS) xst(mypart + D1 + c, hred[c]);
        grid_barrier(A.bar, ep, L, a);
        grid_partial_sum(A.part, A.pstr, D1, 2 * j, L, hred, scr);
        for (int c = tid; c < 2 * j; c += PP_THREADS) hacc[c] = hred[c];
        for (int i = tid; i < N; i += PP_THREADS) {
          zd w = W[i];
          for (int k = 0; k < j; k++) w = zsub(w, zmul(Va[(size_t)k * N + i], zmk(hred[2 * k], hred[2 * k + 1])));
          W[i] = w;
        }
        __syncthreads();
        // second pass: h2 = V^H z', |z'|^2
        fiber_dots(Va, N, W, j, N, true, hred);
        for (int c = tid; c <= 2 * j; c += PP_THREADS) xst(mypart + D2 + c, hred[c]);
        grid_barrier(A.bar, ep, L, a);
        grid_partial_sum(A.part, A.pstr, D2, 2 * j + 1, L, hred, scr);
        double h2n = 0.0;
        for (int k = 0; k < j; k++) h2n += hred[2 * k] * hred[2 * k] + hred[2 * k + 1] * hred[2 * k + 1];
        const double hj1j = sqrt(fmax(hred[2 * j] - h2n, 0.0));
        for (int i = tid; i < N; i += PP_THREADS) {
          zd w = W[i];
          for (int k = 0; k < j; k++) w = zsub(w, zmul(Va[(size_t)k * N + i], zmk(hred[2 * k], hred[2 * k + 1])));
          W[i] = w;
        }
        if (tid < j) Hm[tid + (j - 1) * mh] = zmk(hacc[2 * tid] + hred[2 * tid], hacc[2 * tid + 1] + hred[2 * tid + 1]);
        __syncthreads();
        if (hj1j <= break_tol) {   // happy breakdown (zgexpv.cpp:366-372)
          k1 = 0;
          mbrkdwn = j;
          t_step = t_out - t_now;
          break;
        }
        if (tid == 0) Hm[j + (j - 1) * mh] = zmk(hj1j, 0.0);
        const double hinv = 1.0 / hj1j;
        for (int i = tid; i < padN; i += PP_THREADS) {
          const zd v = i < N ? zscale(W[i], hinv) : zmk(0.0, 0.0);
          if (i < N) Va[(size_t)j * N + i] = v;
          X[i] = v;
        }
        exchange_y();
        grid_barrier(A.bar, ep, L, a);
      }
      if (k1 != 0) {   // the extra matvec of the 2-corrected error estimate
        nmult++;
        pp_matvec<KD, MIXED>(A, M, a, step, fac, pp_nostamp());
        double v = 0.0;
        for (int i = tid; i < N; i += PP_THREADS) v += zabs2(W[i]);
        v = block_sum(v, wsum);
        if (tid == 0) xst(mypart + ANRM, v);
        grid_barrier(A.bar, ep, L, a);
        avnorm = sqrt(reduce1(ANRM));
      }
      if (tid == 0) Hm[(m + 1) + m * mh] = zmk(1.0, 0.0);
      __syncthreads();
      // exponential with step rejection (zgexpv.cpp:397-470)
      int ireject = 0;
      for (;;) {
        nexph++;
        const int mx = mbrkdwn + k1;
        const int ns = pp_zgpadm_block(mx, mh, Hm, t_step, pade_off, ecol);
        if (ns < 0) { iflag = -9; break; }
        if (k1 == 0) err_loc = tol;
        else {
          const double q1 = sqrt(zabs2(ecol[m])) * beta, q2 = sqrt(zabs2(ecol[m + 1])) * beta * avnorm;
          if (q1 > 10.0 * q2) { err_loc = q2; xmloc = 1.0 / m; }
          else if (q1 > q2) { err_loc = q1 * q2 / (q1 - q2); xmloc = 1.0 / m; }
          else { err_loc = q1; xmloc = 1.0 / (m - 1); }
        }
        if (k1 != 0 && err_loc > 1.2 * t_step * tol) {
          t_step = pp_roundstep(0.9 * t_step * pow(t_step * tol / err_loc, xmloc));
          nreject++;
          if (++ireject > 100) { iflag = -8; break; }   // bound (not in zgexpv): no accepted step
          continue;
        }
        break;
      }
      if (iflag) break;
      // w = beta V exp(t H) e1, |w|
      const int mx = mbrkdwn + (k1 - 1 > 0 ? k1 - 1 : 0);
      double nw = 0.0;
      for (int i = tid; i < N; i += PP_THREADS) {
        zd acc = zmk(0.0, 0.0);
        for (int k = 0; k < mx; k++) zfma(acc, Va[(size_t)k * N + i], ecol[k]);
        acc = zscale(acc, beta);
        W[i] = acc;
        nw += zabs2(acc);
      }
      nw = block_sum(nw, wsum);
      if (tid == 0) xst(mypart + NRM, nw);
      grid_barrier(A.bar, ep, L, a);
      beta = sqrt(reduce1(NRM));
      t_new = pp_roundstep(0.9 * t_step * pow(t_step * tol / err_loc, xmloc));
      err_loc = fmax(err_loc, rndoff);
      t_now += t_step;
      if (nstep >= 500 && t_now < t_out) { iflag = 1; break; }
    }
    // ---- psi = U^-1 w (SolveSqrtOverlap)
    for (int i = tid; i < padN; i += PP_THREADS) X[i] = i < N ? W[i] : zmk(0.0, 0.0);
    __syncthreads();
    pp_band_solve<KD, true, PP_SOLVE_THREADS>(pl, nl, X, Z, tb);
    for (int i = tid; i < N; i += PP_THREADS) A.psi[(size_t)a * N + i] = Z[i];
    if (a == 0 && tid == 0) {
      double *st = E.stats + (size_t)step * 6;
      st[0] = nmult; st[1] = nexph; st[2] = nstep; st[3] = nreject; st[4] = mbrkdwn; st[5] = iflag;
    }
    __syncthreads();
  }
}

// Fiber-split form of pp_expokit_kernel for problems with fewer fibers than
// CUs (the layout of pp_pamp_split_kernel: Q workgroups per fiber, own chunk
// ranges, split band solves, coupled rows staged with a halo, U^-1 v handed
// over through production counts).  The same zgexpv: classical
// Gram-Schmidt applied twice, with the inner products and norms partial over
// each workgroup's own rows and summed over all L*Q workgroups in one fixed
// order; every workgroup holds the Hessenberg matrix and runs the
// exponential on identical data, so the step control takes the same
// branches everywhere.
template <int KD, bool MIXED>
__global__ void __launch_bounds__(PP_THREADS) pp_expokit_split_kernel(pp_pamp_args A, pp_expo_args E)
{
  extern __shared__ double lds_d[];
  const int N = A.N, L = A.L, m = A.m, Q = A.Q, tid = threadIdx.x, mh = m + 2;
  const pp_band_plan &pl = A.plan;
  const int P = pl.P, lb = pl.lb;
  int a, q;
  pp_split_map(blockIdx.x, L, Q, a, q);
  const int c0 = q * P / Q, nch = (q + 1) * P / Q - c0, maxch = (P + Q - 1) / Q;
  const int r0 = c0 << lb, nloc = nch << lb, R = maxch << lb;
  const int nown = (r0 + nloc < N ? r0 + nloc : N) - r0;
  const int ystr = R + 2 * KD;
  pp_lds_map M = pp_map_lds<KD>(A, lds_d);
  M.W = (zd *)(lds_d + A.nlevEven + PP_CMB_DOUBLES);
  M.X = M.W + R;
  M.tb = M.X + R;
  zd *bb = M.tb + maxch * KD;
  M.Ys = bb + P * KD;
  M.Z = M.Ys;
  M.wsum = (double *)(M.Ys + A.nslots * ystr);
  M.hred = M.wsum + 16;
  double *hacc = M.hred + (2 * m + 4);             // 2m doubles (the cv slot of pp_split_doubles)
  zd *Hm = (zd *)(hacc + 2 * m);                   // mh x mh, column-major
  zd *ecol = Hm + mh * mh;                         // first column of exp(t H)
  const int pade_off = (int)((double *)(ecol + mh) - lds_d);
  zd *W = M.W, *X = M.X, *tb = M.tb, *Ys = M.Ys, *Z = M.Z;
  double *wsum = M.wsum, *hred = M.hred;
  zd *Va = A.V + (size_t)a * (m + 1) * N;
  const int nwg = L * Q, wid = a * Q + q;
  double *mypart = A.part + (size_t)wid * A.pstr;
  unsigned *fw = A.subw + (size_t)a * Q;
  zd *sbase = A.sb + (size_t)a * 2 * P * KD;
  const bool lead = a == 0 && q == 0;
  const int D1 = 0, D2 = 2 * m, NRM = 4 * m + 2, ANRM = 4 * m + 3, P0NRM = 4 * m + 4;
  unsigned ep = 0, sep = 0;
  int nsolve = 0;
  auto sbuf = [&]() { return sbase + (size_t)((nsolve++) & 1) * P * KD; };
  pp_load_persistent(A, M, a, A.srows + (size_t)q * A.sstride, A.sstride);
  pp_yx yx = pp_yx_init(A, a, wid, nwg, r0, nloc, KD);
  const zd fac = E.imag ? zmk(-1.0, 0.0) : zmk(0.0, -1.0);
  const double eps = 2.220446049250313e-16;
  const double tol = E.tol <= eps ? sqrt(eps) : E.tol;
  const double rndoff = eps * E.anorm, break_tol = 1e-7, t_out = E.t_out;
  auto reduce = [&](int slot, int n) { grid_partial_sum_wide(A.part, A.pstr, slot, n, nwg, hred); };
  // y = U^-1 X (own rows) to the coupled fibers
  auto publish_y = [&]() {
    pp_split_solve<KD, true>(A, M.nl, X, Z, tb, bb, c0, nch, sbuf(), fw, sep, q);
    for (int i = tid; i < nown; i += PP_THREADS) zst(A.y + (size_t)a * N + r0 + i, Z[i]);
    pp_yx_publish(yx);
  };
  // W = fac U^-H H U^-1 v on the own rows, from the coupled fibers' posted U^-1 v
  auto matvec = [&](int step) {
    pp_yx_wait(yx, &A.bar->err);
    const int ybase = r0 - KD, ylo = ybase > 0 ? ybase : 0;
    const int phi = r0 + nloc + KD < N ? r0 + nloc + KD : N;
    for (int sl = 0; sl < A.nslots; sl++) {
      const int col = A.slotcols[a * A.nslots + sl];
      if (col >= 0)
        for (int jj = ylo + tid; jj < phi; jj += PP_THREADS) Ys[sl * ystr + jj - ybase] = zld(A.y + (size_t)col * N + jj);
    }
    __syncthreads();
    pp_band_rows<KD, MIXED>(A, M, Ys, ystr, ybase, r0, nloc, X);
    __syncthreads();
    pp_split_solve<KD, false>(A, M.nl, X, W, tb, bb, c0, nch, sbuf(), fw, sep, q);
  };
  for (int step = 0; step < A.nsteps; step++) {
    // ---- psi~ = U psi on the own rows (psi rows up to r0+nloc+KD-1 were
    // written by the neighbouring part at the end of the previous step)
    if (step > 0) group_barrier(fw, &A.bar->err, sep, Q, q);
    pp_step_coeffs(A, M, step, fac);   // the matvec's coefficients at t_step (read after barriers below)
    const int phi = r0 + nloc + KD < N ? r0 + nloc + KD : N;
    for (int j = r0 + tid; j < phi; j += PP_THREADS) Ys[j - r0] = zld(A.psi + (size_t)a * N + j);
    __syncthreads();
    double nrm = 0.0;
    for (int i = tid; i < nloc; i += PP_THREADS) {
      const int gi = r0 + i;
      zd acc = zmk(0.0, 0.0);
      if (gi < N) {
        const int j1 = gi + KD < N - 1 ? gi + KD : N - 1;
        for (int j = gi; j <= j1; j++) zfmar(acc, A.Uband[(size_t)j * (KD + 1) + KD + gi - j], Ys[j - r0]);
      }
      W[i] = acc;
      nrm += zabs2(acc);
    }
    nrm = block_sum(nrm, wsum);
    if (tid == 0) xst(mypart + P0NRM, nrm);
    grid_barrier(A.bar, ep, nwg, wid);
    reduce(P0NRM, 1);
    double beta = sqrt(hred[0]);
    int nmult = 0, nexph = 0, nstep = 0, nreject = 0, mbrkdwn = m, k1 = 2, iflag = 0;
    if (!(beta > 0.0 && beta <= 1.79e308)) iflag = -10;
    double t_now = 0.0, err_loc = 0.0, t_step = 0.0, avnorm = 0.0, xmloc = 1.0 / m;
    const double p2 = tol * pow((m + 1) / 2.72, (double)(m + 1)) * sqrt((m + 1) * 6.28);
    double t_new = pp_roundstep((1.0 / E.anorm) * pow(p2 / (4.0 * beta * E.anorm), 1.0 / m));
    while (!iflag && t_now < t_out) {
      nstep++;
      t_step = fmin(t_out - t_now, t_new);
      const double inv = 1.0 / beta;
      for (int i = tid; i < nloc; i += PP_THREADS) {
        const zd v = i < nown ? zscale(W[i], inv) : zmk(0.0, 0.0);
        if (i < nown) Va[r0 + i] = v;
        X[i] = v;
      }
      for (int e = tid; e < mh * mh; e += PP_THREADS) Hm[e] = zmk(0.0, 0.0);
      __syncthreads();
      publish_y();
      k1 = 2;
      mbrkdwn = m;
      for (int j = 1; j <= m; j++) {
        nmult++;
        matvec(step);
        // first pass: h1 = V^H z
        split_dots(nullptr, R, 0, Va + r0, N, W, j, nown, hred);
        for (int c = tid; c < 2 * j; c += PP_THREADS) xst(mypart + D1 + c, hred[c]);
        grid_barrier(A.bar, ep, nwg, wid);
        reduce(D1, 2 * j);
        for (int c = tid; c < 2 * j; c += PP_THREADS) hacc[c] = hred[c];
        for (int i = tid; i < nown; i += PP_THREADS) {
          zd w = W[i];
          for (int k = 0; k < j; k++) w = zsub(w, zmul(Va[(size_t)k * N + r0 + i], zmk(hred[2 * k], hred[2 * k + 1])));
          W[i] = w;
        }
        __syncthreads();
        // second pass: h2 = V^H z', |z'|^2
        split_dots(nullptr, R, 0, Va + r0, N, W, j, nown, hred);
        double v2 = 0.0;
        for (int i = tid; i < nown; i += PP_THREADS) v2 += zabs2(W[i]);
        v2 = block_sum(v2, wsum);
        if (tid == 0) hred[2 * j] = v2;
        __syncthreads();
        for (int c = tid; c <= 2 * j; c += PP_THREADS) xst(mypart + D2 + c, hred[c]);
        grid_barrier(A.bar, ep, nwg, wid);
        reduce(D2, 2 * j + 1);
        double h2n = 0.0;
        for (int k = 0; k < j; k++) h2n += hred[2 * k] * hred[2 * k] + hred[2 * k + 1] * hred[2 * k + 1];
        const double hj1j = sqrt(fmax(hred[2 * j] - h2n, 0.0));
        for (int i = tid; i < nown; i += PP_THREADS) {
          zd w = W[i];
          for (int k = 0; k < j; k++) w = zsub(w, zmul(Va[(size_t)k * N + r0 + i], zmk(hred[2 * k], hred[2 * k + 1])));
          W[i] = w;
        }
        if (tid < j) Hm[tid + (j - 1) * mh] = zmk(hacc[2 * tid] + hred[2 * tid], hacc[2 * tid + 1] + hred[2 * tid + 1]);
        __syncthreads();
        if (hj1j <= break_tol) {   // happy breakdown (zgexpv.cpp:366-372)
          k1 = 0;
          mbrkdwn = j;
          t_step = t_out - t_now;
          break;
        }
        if (tid == 0) Hm[j + (j - 1) * mh] = zmk(hj1j, 0.0);
        const double hinv = 1.0 / hj1j;
        for (int i = tid; i < nloc; i += PP_THREADS) {
          const zd v = i < nown ? zscale(W[i], hinv) : zmk(0.0, 0.0);
          if (i < nown) Va[(size_t)j * N + r0 + i] = v;
          X[i] = v;
        }
        __syncthreads();
        publish_y();
      }
      if (k1 != 0) {   // the extra matvec of the 2-corrected error estimate
        nmult++;
        matvec(step);
        double v = 0.0;
        for (int i = tid; i < nown; i += PP_THREADS) v += zabs2(W[i]);
        v = block_sum(v, wsum);
        if (tid == 0) xst(mypart + ANRM, v);
        grid_barrier(A.bar, ep, nwg, wid);
        reduce(ANRM, 1);
        avnorm = sqrt(hred[0]);
      }
      if (tid == 0) Hm[(m + 1) + m * mh] = zmk(1.0, 0.0);
      __syncthreads();
      // exponential with step rejection (zgexpv.cpp:397-470)
      int ireject = 0;
      for (;;) {
        nexph++;
        const int mx = mbrkdwn + k1;
        const int ns = pp_zgpadm_block(mx, mh, Hm, t_step, pade_off, ecol);
        if (ns < 0) { iflag = -9; break; }
        if (k1 == 0) err_loc = tol;
        else {
          const double q1 = sqrt(zabs2(ecol[m])) * beta, q2 = sqrt(zabs2(ecol[m + 1])) * beta * avnorm;
          if (q1 > 10.0 * q2) { err_loc = q2; xmloc = 1.0 / m; }
          else if (q1 > q2) { err_loc = q1 * q2 / (q1 - q2); xmloc = 1.0 / m; }
          else { err_loc = q1; xmloc = 1.0 / (m - 1); }
        }
        if (k1 != 0 && err_loc > 1.2 * t_step * tol) {
          t_step = pp_roundstep(0.9 * t_step * pow(t_step * tol / err_loc, xmloc));
          nreject++;
          if (++ireject > 100) { iflag = -8; break; }   // bound (not in zgexpv): no accepted step
          continue;
        }
        break;
      }
      if (iflag) break;
      // w = beta V exp(t H) e1, |w|
      const int mx = mbrkdwn + (k1 - 1 > 0 ? k1 - 1 : 0);
      double nw = 0.0;
      for (int i = tid; i < nloc; i += PP_THREADS) {
        zd acc = zmk(0.0, 0.0);
        if (i < nown)
          for (int k = 0; k < mx; k++) zfma(acc, Va[(size_t)k * N + r0 + i], ecol[k]);
        acc = zscale(acc, beta);
        W[i] = acc;
        nw += zabs2(acc);
      }
      nw = block_sum(nw, wsum);
      if (tid == 0) xst(mypart + NRM, nw);
      grid_barrier(A.bar, ep, nwg, wid);
      reduce(NRM, 1);
      beta = sqrt(hred[0]);
      t_new = pp_roundstep(0.9 * t_step * pow(t_step * tol / err_loc, xmloc));
      err_loc = fmax(err_loc, rndoff);
      t_now += t_step;
      if (nstep >= 500 && t_now < t_out) { iflag = 1; break; }
    }
    // ---- psi = U^-1 w (SolveSqrtOverlap) on the own rows
    for (int i = tid; i < nloc; i += PP_THREADS) X[i] = i < nown ? W[i] : zmk(0.0, 0.0);
    __syncthreads();
    pp_split_solve<KD, true>(A, M.nl, X, Z, tb, bb, c0, nch, sbuf(), fw, sep, q);
    for (int i = tid; i < nown; i += PP_THREADS) zst(A.psi + (size_t)a * N + r0 + i, Z[i]);
    if (lead && tid == 0) {
      double *st = E.stats + (size_t)step * 6;
      st[0] = nmult; st[1] = nexph; st[2] = nstep; st[3] = nreject; st[4] = mbrkdwn; st[5] = iflag;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ host API
extern "C" int pp_pamp_create(int L, int N, int m, int reorth, const pp_overlap *ov, pp_pamp **out)
{
  if (!out || !ov || L < 1 || N < 1 || m < 1) return PP_EARG;
  if (ov->N != N) return PP_EARG;
  if (m > PP_MAXM) return PP_EUNSUP;
  pp_pamp *p = new pp_pamp();
  p->L = L; p->N = N; p->m = m; p->reorth = reorth ? 1 : 0; p->ov = ov;
  p->pstr = 4 * m + 6;
  p->s = -1; p->bw = 0; p->nterms = 0; p->nslots = 0;
  p->rowcols.assign(L, std::vector<int>());
  p->slotcols = nullptr; p->dirty = true;
  p->V = p->y = p->Hm = p->cvec = nullptr;
  p->mixed = 0;
  p->normscale = 0;
  p->expmethod = 0;
  p->gen = 0;
  p->part = p->diag = p->scal = nullptr;
  p->bar = nullptr; p->tlog = nullptr; p->timing = 0;
  p->fsplit = 0; p->sb = nullptr; p->sb_cap = 0; p->srows = nullptr; p->srows_q = 0; p->sstride = 0;
  p->bar_words = (size_t)L; p->part_cap = (size_t)L * p->pstr;
  p->diag_cap = p->scal_cap = 0;
  for (int t = 0; t < PP_MAXTERMS; t++) p->termbuf[t] = nullptr;
  PP_HIP(hipMalloc(&p->V, sizeof(zd) * (size_t)L * (m + 1) * N));
  PP_HIP(hipMalloc(&p->y, sizeof(zd) * (size_t)L * N));
  PP_HIP(hipMalloc(&p->Hm, sizeof(zd) * (size_t)m * m));
  PP_HIP(hipMalloc(&p->cvec, sizeof(zd) * (size_t)m));
  PP_HIP(hipMalloc(&p->part, sizeof(double) * (size_t)L * p->pstr));
  PP_HIP(hipMalloc(&p->bar, sizeof(pp_barrier) + sizeof(unsigned) * (size_t)L));
  PP_HIP(hipMemset(p->V, 0, sizeof(zd) * (size_t)L * (m + 1) * N));
  PP_HIP(hipMemset(p->Hm, 0, sizeof(zd) * (size_t)m * m));
  PP_HIP(hipMemset(p->bar, 0, sizeof(pp_barrier) + sizeof(unsigned) * (size_t)L));
  *out = p;
  return PP_OK;
}

// Register one factored tensor potential  V[p, :] = sum_r coef[p, r] bands[r, :]
// with real radial bands in the reference BandNH layout (rank x N*bw) and
// complex angular coefficients (P0 x rank), angular storage with P0 indices.
static int pamp_add_factored(pp_pamp *p, int angular_storage, int P0, const int *pairs, int bw, int rank,
                             const double *coef, const double *bandsNH)
{
  if (p->nterms >= PP_MAXTERMS) return PP_EUNSUP;
  if (bw < 1 || (bw % 2) == 0 || rank < 1) return PP_EARG;
  if (p->s >= 0 && bw != p->bw) return PP_EUNSUP;
  pp_rank_csr csr;
  int rc = pp_build_csr(angular_storage, P0, p->L, pairs, csr);
  if (rc) return rc;
  const int N = p->N, s = (bw - 1) / 2;
  // slots: distinct source fibers per row
  std::vector<int> ent;
  for (int a = 0; a < p->L; a++)
    for (int e = csr.rowptr[a]; e < csr.rowptr[a + 1]; e++) {
      int col = csr.ent[3 * e];
      auto &rc_ = p->rowcols[a];
      auto it = std::find(rc_.begin(), rc_.end(), col);
      if (it == rc_.end()) { if ((int)rc_.size() >= PP_MAXSLOTS) return PP_EUNSUP; rc_.push_back(col); }
    }
  for (int a = 0; a < p->L; a++) {
    int cnt = csr.rowptr[a + 1] - csr.rowptr[a];
    if (cnt > PP_MAXENT) return PP_EUNSUP;
  }
  for (size_t e = 0; e < csr.ent.size() / 3; e++) {
    // find row of entry e
    (void)e;
  }
  // device term buffer: rowptr | ent(slot,pidx,conj) | coef | bands
  std::vector<int> ints(csr.rowptr);
  for (int a = 0; a < p->L; a++)
    for (int e = csr.rowptr[a]; e < csr.rowptr[a + 1]; e++) {
      int col = csr.ent[3 * e];
      int slot = (int)(std::find(p->rowcols[a].begin(), p->rowcols[a].end(), col) - p->rowcols[a].begin());
      ints.push_back(slot);
      ints.push_back(csr.ent[3 * e + 1]);
      ints.push_back(csr.ent[3 * e + 2]);
    }
  size_t nint = ints.size();
  size_t off_coef = ((nint * sizeof(int) + 15) / 16) * 16;
  size_t off_band = off_coef + sizeof(zd) * (size_t)P0 * rank;
  size_t off_band32 = off_band + sizeof(double) * (size_t)rank * bw * N;
  size_t total = off_band32 + sizeof(float) * (size_t)rank * bw * N;
  std::vector<char> host(total, 0);
  memcpy(host.data(), ints.data(), nint * sizeof(int));
  memcpy(host.data() + off_coef, coef, sizeof(zd) * (size_t)P0 * rank);
  double *bd = (double *)(host.data() + off_band);
  for (int r = 0; r < rank; r++)
    for (int j = 0; j < N; j++)
      for (int i = j - s; i <= j + s; i++) {
        if (i < 0 || i >= N) continue;
        int d = j - i;   // element (i, i+d)
        bd[((size_t)r * bw + d + s) * N + i] = bandsNH[(size_t)r * N * bw + (size_t)j * bw + s + i - j];
      }
  float *bf = (float *)(host.data() + off_band32);
  for (size_t e = 0; e < (size_t)rank * bw * N; e++) bf[e] = (float)bd[e];
  void *dev = nullptr;
  PP_HIP(hipMalloc(&dev, total));
  PP_HIP(hipMemcpy(dev, host.data(), total, hipMemcpyHostToDevice));
  int t = p->nterms++;
  p->termbuf[t] = dev;
  p->term[t].rowptr = (const int *)dev;
  p->term[t].ent = (const int *)dev + csr.rowptr.size();
  p->term[t].coef = (const zd *)((char *)dev + off_coef);
  p->term[t].bands = (const double *)((char *)dev + off_band);
  p->term[t].bands32 = (const float *)((char *)dev + off_band32);
  p->term[t].rank = rank;
  p->bw = bw;
  p->s = s;
  p->dirty = true;
  p->gen++;
  return PP_OK;
}

extern "C" int pp_pamp_add_factored_term(pp_pamp *p, int angular_storage, int P0, const int *pairs,
                                         int radial_storage, int bw, int rank, const double *coef,
                                         const double *bands)
{
  if (!p || !coef || !bands || P0 < 1) return PP_EARG;
  if (radial_storage != PP_BANDNH) return PP_EUNSUP;
  return pamp_add_factored(p, angular_storage, P0, pairs, bw, rank, coef, bands);
}

// Low-rank factorisation of a tensor potential (host, complex interleaved,
// P0 x M): V[p,:] = sum_r coef[p,r] bands[r,:] with real bands.  Bands are
// rows of the real and imaginary parts chosen by pivoted Gram-Schmidt (an
// interpolative decomposition, so coefficients stay O(1)); coef solves the
// triangular system in the orthonormal basis.  *rank_out = rank, *relerr =
// max |V - recon| / max |V|.  coef: P0 x maxrank complex, bands: maxrank x M.
extern "C" int pp_lowrank_factor(int P0, long M, const double *V, int maxrank, int *rank_out, double *coef,
                                 double *bands, double *relerr)
{
  if (P0 < 1 || M < 1 || !V || !rank_out || !coef || !bands || maxrank < 1) return PP_EARG;
  const int R2 = 2 * P0;
  std::vector<double> rows((size_t)R2 * M);
  for (int p0 = 0; p0 < P0; p0++)
    for (long e = 0; e < M; e++) {
      rows[(size_t)p0 * M + e] = V[2 * ((size_t)p0 * M + e)];
      rows[(size_t)(P0 + p0) * M + e] = V[2 * ((size_t)p0 * M + e) + 1];
    }
  double maxnorm = 0.0, maxval = 0.0;
  for (int r = 0; r < R2; r++) {
    double n2 = 0.0;
    for (long e = 0; e < M; e++) { double v = rows[(size_t)r * M + e]; n2 += v * v; maxval = std::max(maxval, std::fabs(v)); }
    maxnorm = std::max(maxnorm, std::sqrt(n2));
  }
  std::vector<double> Q, res(rows);
  std::vector<int> piv;
  int rank = 0;
  while (rank < std::min(R2, maxrank) && maxnorm > 0.0) {
    int best = -1;
    double bn = 0.0;
    for (int r = 0; r < R2; r++) {
      double n2 = 0.0;
      for (long e = 0; e < M; e++) n2 += res[(size_t)r * M + e] * res[(size_t)r * M + e];
      if (std::sqrt(n2) > bn) { bn = std::sqrt(n2); best = r; }
    }
    if (bn <= 1e-13 * maxnorm) break;   // below the assembly rounding of the rows
    std::vector<double> q(res.begin() + (size_t)best * M, res.begin() + (size_t)(best + 1) * M);
    for (int pass = 0; pass < 2; pass++)
      for (int k = 0; k < rank; k++) {
        double d = 0.0;
        for (long e = 0; e < M; e++) d += Q[(size_t)k * M + e] * q[e];
        for (long e = 0; e < M; e++) q[e] -= d * Q[(size_t)k * M + e];
      }
    double n2 = 0.0;
    for (long e = 0; e < M; e++) n2 += q[e] * q[e];
    const double nq = std::sqrt(n2);
    for (long e = 0; e < M; e++) q[e] /= nq;
    Q.insert(Q.end(), q.begin(), q.end());
    piv.push_back(best);
    rank++;
    // project the new direction out of every residual row twice (one pass
    // leaves ~eps*sqrt(M) of the row behind, which would pass for rank);
    // the pivot row lies in the span exactly
    for (int r = 0; r < R2; r++) {
      double *rr = res.data() + (size_t)r * M;
      if (r == best) { std::fill(rr, rr + M, 0.0); continue; }
      for (int pass = 0; pass < 2; pass++) {
        double d = 0.0;
        for (long e = 0; e < M; e++) d += q[e] * rr[e];
        for (long e = 0; e < M; e++) rr[e] -= d * q[e];
      }
    }
  }
  if (rank == 0) {   // zero potential
    *rank_out = 1;
    for (long e = 0; e < M; e++) bands[e] = 0.0;
    for (int p0 = 0; p0 < P0; p0++) coef[2 * p0] = coef[2 * p0 + 1] = 0.0;
    if (relerr) *relerr = 0.0;
    return PP_OK;
  }
  // pivot rows B_j = sum_k R[k][j] q_k (R upper triangular); coef solves R c = Q^T row
  std::vector<double> R((size_t)rank * rank, 0.0);
  for (int j = 0; j < rank; j++)
    for (int k = 0; k <= j; k++) {
      double d = 0.0;
      for (long e = 0; e < M; e++) d += Q[(size_t)k * M + e] * rows[(size_t)piv[j] * M + e];
      R[(size_t)k * rank + j] = d;
    }
  for (int j = 0; j < rank; j++)
    for (long e = 0; e < M; e++) bands[(size_t)j * M + e] = rows[(size_t)piv[j] * M + e];
  std::vector<double> yv(rank);
  double maxerr = 0.0;
  for (int p0 = 0; p0 < P0; p0++)
    for (int part = 0; part < 2; part++) {
      const double *row = rows.data() + (size_t)(part * P0 + p0) * M;
      for (int k = 0; k < rank; k++) {
        double d = 0.0;
        for (long e = 0; e < M; e++) d += Q[(size_t)k * M + e] * row[e];
        yv[k] = d;
      }
      for (int k = rank - 1; k >= 0; k--) {
        double v = yv[k];
        for (int j = k + 1; j < rank; j++) v -= R[(size_t)k * rank + j] * yv[j];
        yv[k] = v / R[(size_t)k * rank + k];
      }
      for (int k = 0; k < rank; k++) coef[2 * ((size_t)p0 * rank + k) + part] = yv[k];
      for (long e = 0; e < M; e++) {
        double v = 0.0;
        for (int k = 0; k < rank; k++) v += yv[k] * bands[(size_t)k * M + e];
        maxerr = std::max(maxerr, std::fabs(v - row[e]));
      }
    }
  *rank_out = rank;
  if (relerr) *relerr = maxval > 0.0 ? maxerr / maxval : 0.0;
  return PP_OK;
}

// One consolidated tensor potential in the reference layout (device,
// P0 x N*bw complex).  The library factors it on the host
// (pp_lowrank_factor) into real radial bands and complex angular
// coefficients and rejects factorisations that miss V by more than
// 1e-12 max|V| (the assembled tensor is itself low rank only up to the
// rounding of its separately projected parts, a few 1e-14).
extern "C" int pp_pamp_add_term(pp_pamp *p, int angular_storage, int P0, const int *pairs, int radial_storage,
                                int bw, const void *pot)
{
  if (!p || !pot || P0 < 1) return PP_EARG;
  if (radial_storage != PP_BANDNH) return PP_EUNSUP;
  if (bw < 1 || (bw % 2) == 0) return PP_EARG;
  const long M = (long)p->N * bw;
  std::vector<double> T((size_t)P0 * M * 2);
  PP_HIP(hipMemcpy(T.data(), pot, T.size() * sizeof(double), hipMemcpyDeviceToHost));
  const int maxrank = std::min(2 * P0, 16);
  std::vector<double> coef((size_t)P0 * maxrank * 2), bands((size_t)maxrank * M);
  int rank = 0;
  double relerr = 0.0;
  int rc = pp_lowrank_factor(P0, M, T.data(), maxrank, &rank, coef.data(), bands.data(), &relerr);
  if (rc) return rc;
  if (relerr > 1e-12) return PP_ENUM;
  return pamp_add_factored(p, angular_storage, P0, pairs, bw, rank, coef.data(), bands.data());
}

static int pamp_scr_doubles(int L, int N, int m, int nslots, int padN)
{
  return std::max(std::max(2 * nslots * N, L * (2 * m + 2)), 2 * padN);
}

static int pamp_scr_doubles(const pp_pamp *p, int padN)
{
  return pamp_scr_doubles(p->L, p->N, p->m, p->nslots, padN);
}

// LDS bytes of one pAMP step workgroup: per-fiber buffers, or the Pade
// scratch on workgroup 0, above the persistent scan matrices and combo table.
static size_t pamp_lds_bytes(int L, int N, int m, int nslots, int kd, int P, int B, int nlev)
{
  const int padN = P * B;
  const size_t nle = (size_t)((nlev + 1) & ~1) + PP_CMB_DOUBLES;
  size_t fib = sizeof(double) * nle + sizeof(zd) * (size_t)(2 * padN + P * kd) +
               sizeof(double) * (size_t)pamp_scr_doubles(L, N, m, nslots, padN) +
               sizeof(double) * (size_t)(16 + (2 * m + 4) + PP_THREADS + 2 * m) + sizeof(zd) * (size_t)m + 64;
  size_t pade = sizeof(double) * (nle + (size_t)pp_pade_lds_doubles(m)) + 64;
  return fib > pade ? fib : pade;
}

static size_t pamp_lds_bytes(const pp_pamp *p, int padN)
{
  const pp_band_plan &pl = p->ov->plan;
  (void)padN;
  return pamp_lds_bytes(p->L, p->N, p->m, p->nslots, p->ov->kd, pl.P, pl.B, pl.nlev);
}

// LDS bytes a pAMP launch of an (L angular, N radial) problem with Krylov
// dimension m, nslots coupled fibers per row and solver chunk length B needs
// per workgroup (-1 if B is not a valid chunk length).  No reference
// counterpart: the batch autotuner uses it to keep two workgroups per CU.
extern "C" long pp_pamp_lds_estimate(int L, int N, int m, int nslots, int kd, int B)
{
  if (B == 0) B = pp_plan_chunk(N, kd);
  if (L < 1 || N < 1 || m < 1 || nslots < 1 || kd < 1 || B < 8 || B < kd || (B & (B - 1))) return -1;
  const int P = (N + B - 1) / B;
  if (P > 64) return -1;
  int levels = 0;
  while ((1 << levels) < P) levels++;
  return (long)pamp_lds_bytes(L, N, m, nslots, kd, P, B, 2 * levels * kd * kd * P);
}

static int pamp_upload_slots(pp_pamp *p)
{
  if (!p->dirty) return PP_OK;
  int ns = 0;
  for (auto &rc : p->rowcols) ns = std::max(ns, (int)rc.size());
  if (ns < 1) ns = 1;
  std::vector<int> sc((size_t)p->L * ns, -1);
  for (int a = 0; a < p->L; a++)
    for (size_t q = 0; q < p->rowcols[a].size(); q++) sc[(size_t)a * ns + q] = p->rowcols[a][q];
  if (p->slotcols) (void)hipFree(p->slotcols);
  p->gen++;
  PP_HIP(hipMalloc(&p->slotcols, sc.size() * sizeof(int)));
  PP_HIP(hipMemcpy(p->slotcols, sc.data(), sc.size() * sizeof(int), hipMemcpyHostToDevice));
  p->nslots = ns;
  p->dirty = false;
  return PP_OK;
}

// Launch set-up of the step kernels without stream work: buffers for
// nsteps steps and the kernel arguments.
static int pamp_args(pp_pamp *p, void *psi, int nsteps, double dt_re, double dt_im, pp_pamp_args &A)
{
  int rc = pamp_upload_slots(p);
  if (rc) return rc;
  const pp_overlap *ov = p->ov;
  if (p->s != ov->kd) return PP_EUNSUP;   // radial bands must have the overlap's bandwidth
  if (nsteps * p->nterms > p->scal_cap) {
    if (p->scal) (void)hipFree(p->scal);
    p->gen++;
    p->scal_cap = nsteps * p->nterms;
    PP_HIP(hipMalloc(&p->scal, sizeof(double) * p->scal_cap));
  }
  if (nsteps > p->diag_cap) {
    if (p->diag) (void)hipFree(p->diag);
    p->gen++;
    p->diag_cap = nsteps;
    PP_HIP(hipMalloc(&p->diag, sizeof(double) * 6 * (size_t)nsteps));
  }
  memset(&A, 0, sizeof(A));
  A.L = p->L; A.N = p->N; A.m = p->m; A.s = p->s; A.bw = p->bw; A.nterms = p->nterms;
  A.nsteps = nsteps; A.reorth = p->reorth; A.nslots = p->nslots; A.normscale = p->expmethod ? 2 : p->normscale;
  A.ncombo = 0;
  for (int t = 0; t < p->nterms; t++) A.ncombo += p->term[t].rank;
  if (A.ncombo > PP_MAXCOMBO) return PP_EUNSUP;
  const pp_band_plan &pl = ov->plan;
  A.padN = pl.P * pl.B;
  A.scrDoubles = pamp_scr_doubles(p, A.padN);
  A.nlevEven = (pl.nlev + 1) & ~1;
  A.dt_re = dt_re; A.dt_im = dt_im;
  for (int t = 0; t < p->nterms; t++) {
    A.term[t] = p->term[t];
    A.term[t].fp32 = (p->mixed >> t) & 1u;
  }
  A.slotcols = p->slotcols;
  A.scal = p->scal;
  A.Uband = ov->Uband;
  A.plan = pl;
  A.psi = (zd *)psi; A.V = p->V; A.y = p->y; A.part = p->part; A.pstr = p->pstr;
  A.Hm = p->Hm; A.cvec = p->cvec; A.diag = p->diag; A.bar = p->bar;
  A.tlog = p->timing ? p->tlog : nullptr;
  return PP_OK;
}

// Common launch set-up of the step kernels: pamp_args, then the time
// factors uploaded and the barrier words cleared on the stream.
static int pamp_prepare(pp_pamp *p, void *psi, int nsteps, double dt_re, double dt_im, const double *scal,
                        hipStream_t stream, pp_pamp_args &A)
{
  int rc = pamp_args(p, psi, nsteps, dt_re, dt_im, A);
  if (rc) return rc;
  PP_HIP(hipMemcpyAsync(p->scal, scal, sizeof(double) * nsteps * p->nterms, hipMemcpyHostToDevice, stream));
  PP_HIP(hipMemsetAsync(p->bar, 0, sizeof(pp_barrier) + sizeof(unsigned) * p->bar_words, stream));
  return PP_OK;
}

static int pamp_launch(const void *fn, int L, void **args, size_t lds, hipStream_t stream)
{
  if (lds > 160 * 1024) return PP_EUNSUP;
  hipError_t err = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err == hipSuccess) err = hipLaunchCooperativeKernel(fn, dim3(L), dim3(PP_THREADS), args, (unsigned)lds, stream);
  return err == hipSuccess ? PP_OK : PP_EHIP;
}

// Workgroups per fiber for pp_pamp_propagate: as many as keep one workgroup
// per CU (at most 8) and at least two solver chunks per workgroup; 1 (the
// one-workgroup-per-fiber kernel) with re-orthogonalisation or when that
// leaves fewer than 2.  A fixed request must leave every workgroup a chunk.
static int pamp_split_factor(const pp_pamp *p, int &Q)
{
  const int P = p->ov->plan.P;
  Q = 1;
  if (p->reorth || p->fsplit == 1) return PP_OK;
  // the split kernel waits on the producers of its coupled rows with one
  // lane each: nslots * Q of them at most
  const int qmax = 64 / (p->nslots > 0 ? p->nslots : 1);
  if (p->fsplit >= 2) {
    if (p->fsplit > P || p->fsplit > qmax) return PP_EUNSUP;
    Q = p->fsplit;
    return PP_OK;
  }
  int dev = 0, cus = 0;
  PP_HIP(hipGetDevice(&dev));
  PP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int q = cus / p->L;
  if (q > 8) q = 8;
  if (q > P / 2) q = P / 2;
  if (q > qmax) q = qmax;
  Q = q >= 2 ? q : 1;
  return PP_OK;
}

// LDS bytes of one split-kernel workgroup (the lead one also runs the Pade
// exponential in the same space).
static size_t pamp_split_lds_bytes(const pp_pamp *p, int Q)
{
  const pp_band_plan &pl = p->ov->plan;
  const int maxch = (pl.P + Q - 1) / Q;
  const size_t nle = (size_t)pp_split_rows_doubles(p->ov->kd, pl.P, maxch) + PP_CMB_DOUBLES;
  const size_t fib = sizeof(double) * (nle + (size_t)pp_split_doubles(p->nslots, p->m, p->ov->kd, pl.B, maxch, pl.P)) + 64;
  const size_t pade = sizeof(double) * (nle + (size_t)pp_pade_lds_doubles(p->m)) + 64;
  return fib > pade ? fib : pade;
}

// Krylov vectors whose own rows the split kernel also keeps in LDS: as many
// as fit beside its per-workgroup buffers within 160 KB (the LDS the lead's
// Pade exponential needs is allocated anyway); the LDS bytes to allocate.
static int pamp_split_vcache(const pp_pamp *p, int Q, size_t &lds)
{
  const pp_band_plan &pl = p->ov->plan;
  const int maxch = (pl.P + Q - 1) / Q;
  const size_t nle = (size_t)pp_split_rows_doubles(p->ov->kd, pl.P, maxch) + PP_CMB_DOUBLES;
  const size_t R = (size_t)maxch << pl.lb;
  const size_t fib = sizeof(double) * (nle + (size_t)pp_split_doubles(p->nslots, p->m, p->ov->kd, pl.B, maxch, pl.P)) + 64;
  lds = pamp_split_lds_bytes(p, Q);
  const size_t cap = 160 * 1024, per = sizeof(zd) * R;
  long kc = fib < cap ? (long)((cap - fib) / per) : 0;
  if (kc > p->m) kc = p->m;
  if (fib + per * (size_t)kc > lds) lds = fib + per * (size_t)kc;
  return (int)kc;
}

// Chunk-boundary rows of the split band solves for split factor Q: the
// scan of the band solver solves x_c = b_c + N_c x_{c+1} (U^-1) and x_c =
// b_c + N'_c x_{c-1} (U^-T) over the chunks' boundary values, with N_c, N'_c
// its level-0 matrices (the first kd*kd*P doubles of each half of the plan's
// scan matrices, [(i*kd + k)*P + c]).  Part q (chunks c0 .. c0+nch-1) keeps,
// per own chunk c and component i, the row of x_{c+1} = sum_{c' > c}
// (N_{c+1} .. N_{c'-1}) b_{c'} over (c', k) for c' = c+1 .. P-1, then the
// rows of x_{c-1} = sum_{c' < c} (N'_{c-1} .. N'_{c'+1}) b_{c'} over
// c' = 0 .. c-1 (pp_split_solve reads them in this order); FP64 products on
// the host, once per split factor.
static int pamp_split_rows(pp_pamp *p, int Q)
{
  if (p->srows && p->srows_q == Q) return PP_OK;
  const pp_band_plan &pl = p->ov->plan;
  const int P = pl.P, kd = p->ov->kd, kk = kd * kd, maxch = (P + Q - 1) / Q;
  const size_t nL = (size_t)pl.levels * kk * P;
  std::vector<double> n0((size_t)kk * P), n0t((size_t)kk * P);
  PP_HIP(hipMemcpy(n0.data(), pl.nlevs, sizeof(double) * n0.size(), hipMemcpyDeviceToHost));
  PP_HIP(hipMemcpy(n0t.data(), pl.nlevs + nL, sizeof(double) * n0t.size(), hipMemcpyDeviceToHost));
  const int stride = pp_split_rows_doubles(kd, P, maxch);
  std::vector<double> rows((size_t)stride * Q, 0.0), T(kk), T2(kk);
  auto times = [&](const std::vector<double> &Nm, int c) {   // T := T N_c
    for (int i = 0; i < kd; i++)
      for (int k = 0; k < kd; k++) {
        double a = 0.0;
        for (int u = 0; u < kd; u++) a += T[(size_t)i * kd + u] * Nm[((size_t)u * kd + k) * P + c];
        T2[(size_t)i * kd + k] = a;
      }
    T.swap(T2);
  };
  auto identity = [&]() { for (int e = 0; e < kk; e++) T[e] = (e % (kd + 1) == 0) ? 1.0 : 0.0; };
  for (int q = 0; q < Q; q++) {
    const int c0 = q * P / Q, nch = (q + 1) * P / Q - c0;
    double *out = rows.data() + (size_t)q * stride;
    size_t o = 0;
    for (int cl = 0; cl < nch; cl++) {   // x_{c+1}: columns c' = c+1 .. P-1
      const int c = c0 + cl, len = (P - c - 1) * kd;
      identity();
      for (int cp = c + 1; cp < P; cp++) {
        for (int i = 0; i < kd; i++)
          for (int k = 0; k < kd; k++) out[o + (size_t)i * len + (size_t)(cp - c - 1) * kd + k] = T[(size_t)i * kd + k];
        if (cp + 1 < P) times(n0, cp);
      }
      o += (size_t)kd * len;
    }
    for (int cl = 0; cl < nch; cl++) {   // x_{c-1}: columns c' = 0 .. c-1
      const int c = c0 + cl, len = c * kd;
      identity();
      for (int cp = c - 1; cp >= 0; cp--) {
        for (int i = 0; i < kd; i++)
          for (int k = 0; k < kd; k++) out[o + (size_t)i * len + (size_t)cp * kd + k] = T[(size_t)i * kd + k];
        if (cp > 0) times(n0t, cp);
      }
      o += (size_t)kd * len;
    }
    if (o > (size_t)stride) return PP_EUNSUP;   // cannot happen: kd^2 (P-1) per chunk
  }
  if (p->srows) (void)hipFree(p->srows);
  p->srows = nullptr;
  p->gen++;
  PP_HIP(hipMalloc(&p->srows, sizeof(double) * rows.size()));
  PP_HIP(hipMemcpy(p->srows, rows.data(), sizeof(double) * rows.size(), hipMemcpyHostToDevice));
  p->srows_q = Q;
  p->sstride = stride;
  return PP_OK;
}

// Buffers of the split kernel: partial sums of L*Q workgroups, grid and
// per-fiber arrival words, the chunk boundary exchange.
static int pamp_split_buffers(pp_pamp *p, int Q)
{
  const size_t nw = (size_t)p->L * Q;
  if (nw * p->pstr > p->part_cap) {
    p->gen++;
    (void)hipFree(p->part);
    p->part = nullptr;
    PP_HIP(hipMalloc(&p->part, sizeof(double) * nw * p->pstr));
    p->part_cap = nw * p->pstr;
  }
  if (3 * nw > p->bar_words) {   // grid arrival words, per-fiber arrival words, U^-1 u production counts
    p->gen++;
    (void)hipFree(p->bar);
    p->bar = nullptr;
    PP_HIP(hipMalloc(&p->bar, sizeof(pp_barrier) + sizeof(unsigned) * 3 * nw));
    p->bar_words = 3 * nw;
  }
  const size_t nsb = (size_t)p->L * 2 * p->ov->plan.P * p->ov->kd;
  if (nsb > p->sb_cap) {
    if (p->sb) (void)hipFree(p->sb);
    p->sb = nullptr;
    PP_HIP(hipMalloc(&p->sb, sizeof(zd) * nsb));
    p->sb_cap = nsb;
  }
  return PP_OK;
}

// Propagate psi (device, L x N complex128) through nsteps steps of length dt.
// scal: host array nsteps x nterms with each term's time factor at t_n.
extern "C" int pp_pamp_propagate(pp_pamp *p, void *psi, int nsteps, double dt_re, double dt_im,
                                 const double *scal, hipStream_t stream)
{
  if (!p || !psi || nsteps < 0 || (!scal && p->nterms > 0)) return PP_EARG;
  if (nsteps == 0) return PP_OK;
  if (p->nterms == 0) return PP_EARG;
  int rc = pamp_upload_slots(p);   // nslots bounds the split factor
  if (rc) return rc;
  int Q = 1;
  rc = pamp_split_factor(p, Q);
  if (rc) return rc;
  if (Q > 1 && ((rc = pamp_split_buffers(p, Q)) || (rc = pamp_split_rows(p, Q)))) return rc;
  pp_pamp_args A;
  rc = pamp_prepare(p, psi, nsteps, dt_re, dt_im, scal, stream, A);
  if (rc) return rc;
  void *args[] = {&A};
  rc = PP_EUNSUP;
  if (Q > 1) {
    A.Q = Q;
    A.sb = p->sb;
    A.srows = p->srows;
    A.sstride = p->sstride;
    A.nlevEven = p->sstride;   // the chunk-boundary rows take the scan matrices' place in LDS
    A.subw = (unsigned *)(p->bar + 1) + (size_t)p->L * Q;
    size_t lds = 0;
    A.vcache = pamp_split_vcache(p, Q, lds);
    PP_KD_SWITCH(p->ov->kd, {
      rc = pamp_launch(p->mixed ? (const void *)pp_pamp_split_kernel<KD_, true>
                                : (const void *)pp_pamp_split_kernel<KD_, false>,
                       p->L * Q, args, lds, stream); })
    return rc;
  }
  A.Q = 1;
  const size_t lds = pamp_lds_bytes(p, A.padN);
  PP_KD_SWITCH(p->ov->kd, {
    rc = pamp_launch(p->mixed ? (const void *)pp_pamp_kernel<KD_, true> : (const void *)pp_pamp_kernel<KD_, false>,
                     p->L, args, lds, stream); })
  return rc;
}

// Workgroups per fiber of pp_pamp_propagate: 0 automatic, 1 one workgroup
// per fiber, q >= 2 that many (at most the solver's chunk count).  No
// reference counterpart: a launch-shape control of this implementation.
extern "C" int pp_pamp_set_fiber_split(pp_pamp *p, int q)
{
  if (!p || q < 0 || q > 64) return PP_EARG;
  p->fsplit = q;
  return PP_OK;
}

// Many independent instances (each a pp_pamp with its own operator, overlap
// and psi) advanced nsteps steps each.  Instances are packed in order into
// cooperative launches whose workgroups (one per fiber) are all co-resident;
// a batch larger than the GPU holds at once becomes several launches on the
// stream.  dt: ninst x 2 (re, im); scal[i]: host nsteps x nterms_i.  All
// instances need the same B-spline order and precision form.
struct pp_batch_pack {
  int i0, i1, wgs;         // instances [i0, i1), workgroups
  size_t lds;              // dynamic LDS of the launch (the group's largest)
};

// Arguments of every instance and the packing of the batch (no stream work).
static int batch_plan(int ninst, pp_pamp *const *inst, void *const *psi, int nsteps, const double *dt,
                      std::vector<pp_pamp_args> &args, const void *&fn, std::vector<pp_batch_pack> &packs)
{
  if (ninst <= 0 || !inst || !psi || !dt || nsteps <= 0) return PP_EARG;
  const int kd = inst[0] ? inst[0]->ov->kd : 0;
  const bool mixed = inst[0] && inst[0]->mixed != 0;
  for (int i = 0; i < ninst; i++) {
    if (!inst[i] || !psi[i] || inst[i]->nterms == 0) return PP_EARG;
    if (inst[i]->ov->kd != kd || (inst[i]->mixed != 0) != mixed) return PP_EUNSUP;
  }
  args.assign(ninst, pp_pamp_args());
  std::vector<size_t> lds(ninst);
  for (int i = 0; i < ninst; i++) {
    int rc = pamp_args(inst[i], psi[i], nsteps, dt[2 * i], dt[2 * i + 1], args[i]);
    if (rc) return rc;
    lds[i] = pamp_lds_bytes(inst[i], args[i].padN);
    if (lds[i] > 160 * 1024) return PP_EUNSUP;
  }
  fn = nullptr;
  PP_KD_SWITCH(kd, { fn = mixed ? (const void *)pp_pamp_batch_kernel<KD_, true> : (const void *)pp_pamp_batch_kernel<KD_, false>; })
  if (!fn) return PP_EUNSUP;
  int dev = 0, ncu = 0, reg_per_cu = 0;
  PP_HIP(hipGetDevice(&dev));
  PP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  // workgroups per CU: the register limit from the runtime, the LDS limit
  // (160 KB per CU) from the group's largest allocation
  PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&reg_per_cu, fn, PP_THREADS, 0));
  packs.clear();
  int i0 = 0;
  while (i0 < ninst) {
    // grow the group while all its workgroups fit on the GPU at once
    size_t lmax = 0;
    int wgs = 0, i1 = i0;
    while (i1 < ninst) {
      const size_t l2 = std::max(lmax, lds[i1]);
      const int per_cu = std::min(reg_per_cu, (int)((160 * 1024) / l2));
      if (wgs + inst[i1]->L > per_cu * ncu) break;
      lmax = l2;
      wgs += inst[i1]->L;
      i1++;
    }
    if (i1 == i0) return PP_EUNSUP;   // one instance alone does not fit
    packs.push_back({i0, i1, wgs, lmax});
    i0 = i1;
  }
  return PP_OK;
}

// Workgroup -> (instance within the pack, fiber) map of one pack.
static std::vector<int2> batch_map(pp_pamp *const *inst, const pp_batch_pack &pk)
{
  std::vector<int2> map;
  map.reserve(pk.wgs);
  for (int i = pk.i0; i < pk.i1; i++)
    for (int a = 0; a < inst[i]->L; a++) map.push_back(make_int2(i - pk.i0, a));
  return map;
}

extern "C" int pp_pamp_batch_propagate(int ninst, pp_pamp *const *inst, void *const *psi, int nsteps,
                                       const double *dt, const double *const *scal, hipStream_t stream)
{
  if (ninst < 0 || (ninst > 0 && (!inst || !psi || !dt || !scal)) || nsteps < 0) return PP_EARG;
  if (ninst == 0 || nsteps == 0) return PP_OK;
  for (int i = 0; i < ninst; i++)
    if (!scal[i]) return PP_EARG;
  std::vector<pp_pamp_args> args;
  std::vector<pp_batch_pack> packs;
  const void *fn = nullptr;
  int rc = batch_plan(ninst, inst, psi, nsteps, dt, args, fn, packs);
  if (rc) return rc;
  for (int i = 0; i < ninst; i++) {
    pp_pamp *p = inst[i];
    PP_HIP(hipMemcpyAsync(p->scal, scal[i], sizeof(double) * nsteps * p->nterms, hipMemcpyHostToDevice, stream));
    PP_HIP(hipMemsetAsync(p->bar, 0, sizeof(pp_barrier) + sizeof(unsigned) * p->bar_words, stream));
  }
  for (const pp_batch_pack &pk : packs) {
    const std::vector<int2> map = batch_map(inst, pk);
    pp_pamp_args *dargs = nullptr;
    int2 *dmap = nullptr;
    PP_HIP(hipMallocAsync((void **)&dargs, sizeof(pp_pamp_args) * (pk.i1 - pk.i0), stream));
    PP_HIP(hipMallocAsync((void **)&dmap, sizeof(int2) * map.size(), stream));
    PP_HIP(hipMemcpyAsync(dargs, args.data() + pk.i0, sizeof(pp_pamp_args) * (pk.i1 - pk.i0), hipMemcpyHostToDevice,
                          stream));
    PP_HIP(hipMemcpyAsync(dmap, map.data(), sizeof(int2) * map.size(), hipMemcpyHostToDevice, stream));
    PP_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pk.lds));
    void *kargs[] = {&dargs, &dmap};
    PP_HIP(hipLaunchCooperativeKernel(fn, dim3(pk.wgs), dim3(PP_THREADS), kargs, (unsigned)pk.lds, stream));
    PP_HIP(hipFreeAsync(dargs, stream));
    PP_HIP(hipFreeAsync(dmap, stream));
  }
  return PP_OK;
}

// ------------------------------------------------------------------ batch graph
// The packed launches of pp_pamp_batch_propagate recorded once as a HIP
// graph: per pack, a kernel node that clears the pack's barrier words (a
// kernel, not a memset node) and the batch kernel as a cooperative kernel
// node, the packs chained in order.  The graph holds the instances' buffers
// and psi pointers; a replay reads the time factors from the instances'
// device buffers, which pp_pamp_batch_graph_launch refreshes first.
__global__ void __launch_bounds__(256) pp_batch_bar_reset(const pp_pamp_args *args, const int *nwords, int ninst)
{
  for (int i = blockIdx.x; i < ninst; i += gridDim.x) {
    unsigned *w = (unsigned *)args[i].bar;
    for (int e = threadIdx.x; e < nwords[i]; e += 256) w[e] = 0u;
  }
}

struct pp_batch_graph {
  int ninst, nsteps;
  std::vector<pp_pamp *> inst;
  std::vector<unsigned> gen;   // each instance's generation when recorded
  std::vector<void *> dbuf;    // device argument blocks, maps and word counts
  const void *fn;              // the batch kernel and its largest dynamic LDS
  size_t lds;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" int pp_pamp_batch_graph_destroy(pp_batch_graph *g)
{
  if (!g) return PP_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  for (void *b : g->dbuf)
    if (b) (void)hipFree(b);
  delete g;
  return PP_OK;
}

extern "C" int pp_pamp_batch_graph_create(int ninst, pp_pamp *const *inst, void *const *psi, int nsteps,
                                          const double *dt, pp_batch_graph **out)
{
  if (!out) return PP_EARG;
  *out = nullptr;
  std::vector<pp_pamp_args> args;
  std::vector<pp_batch_pack> packs;
  const void *fn = nullptr;
  int rc = batch_plan(ninst, inst, psi, nsteps, dt, args, fn, packs);
  if (rc) return rc;
  pp_batch_graph *g = new pp_batch_graph();
  g->ninst = ninst;
  g->nsteps = nsteps;
  g->graph = nullptr;
  g->exec = nullptr;
  g->fn = fn;
  g->lds = 0;
  for (const pp_batch_pack &pk : packs) g->lds = std::max(g->lds, pk.lds);
  g->inst.assign(inst, inst + ninst);
  for (int i = 0; i < ninst; i++) g->gen.push_back(inst[i]->gen);
  auto fail = [&](int code) { pp_pamp_batch_graph_destroy(g); return code; };
  auto upload = [&](const void *src, size_t bytes, void **dst) {
    if (hipMalloc(dst, bytes) != hipSuccess) return false;
    g->dbuf.push_back(*dst);
    return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  if (hipGraphCreate(&g->graph, 0) != hipSuccess) return fail(PP_EHIP);
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)g->lds) != hipSuccess)
    return fail(PP_EHIP);
  hipGraphNode_t prev = nullptr;
  for (const pp_batch_pack &pk : packs) {
    const int np = pk.i1 - pk.i0;
    const std::vector<int2> map = batch_map(inst, pk);
    std::vector<int> nwords(np);
    for (int i = 0; i < np; i++)
      nwords[i] = (int)((sizeof(pp_barrier) + sizeof(unsigned) * inst[pk.i0 + i]->bar_words) / sizeof(unsigned));
    void *dargs = nullptr, *dmap = nullptr, *dnw = nullptr;
    if (!upload(args.data() + pk.i0, sizeof(pp_pamp_args) * np, &dargs) ||
        !upload(map.data(), sizeof(int2) * map.size(), &dmap) || !upload(nwords.data(), sizeof(int) * np, &dnw))
      return fail(PP_EHIP);
    hipKernelNodeParams rp = {};
    void *rargs[] = {&dargs, &dnw, (void *)&np};
    rp.func = (void *)pp_batch_bar_reset;
    rp.gridDim = dim3(np < 1024 ? np : 1024);
    rp.blockDim = dim3(256);
    rp.sharedMemBytes = 0;
    rp.kernelParams = rargs;
    hipGraphNode_t rn = nullptr, kn = nullptr;
    if (hipGraphAddKernelNode(&rn, g->graph, prev ? &prev : nullptr, prev ? 1 : 0, &rp) != hipSuccess)
      return fail(PP_EHIP);
    hipKernelNodeParams kp = {};
    void *kargs[] = {&dargs, &dmap};
    kp.func = (void *)fn;
    kp.gridDim = dim3(pk.wgs);
    kp.blockDim = dim3(PP_THREADS);
    kp.sharedMemBytes = (unsigned)pk.lds;
    kp.kernelParams = kargs;
    if (hipGraphAddKernelNode(&kn, g->graph, &rn, 1, &kp) != hipSuccess) return fail(PP_EHIP);
    hipKernelNodeAttrValue coop = {};
    coop.cooperative = 1;
    if (hipGraphKernelNodeSetAttribute(kn, hipKernelNodeAttributeCooperative, &coop) != hipSuccess)
      return fail(PP_EUNSUP);
    prev = kn;
  }
  if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) return fail(PP_EHIP);
  *out = g;
  return PP_OK;
}

// One replay: every instance advances the recorded nsteps steps from its
// current psi.  scal[i]: host nsteps x nterms_i time factors of the replay.
// PP_EARG if an instance changed since the graph was recorded (new terms,
// precision or scaling mode, or a launch that grew its buffers).
extern "C" int pp_pamp_batch_graph_launch(pp_batch_graph *g, const double *const *scal, hipStream_t stream)
{
  if (!g || !g->exec || !scal) return PP_EARG;
  for (int i = 0; i < g->ninst; i++)
    if (!scal[i] || g->inst[i]->gen != g->gen[i] || g->inst[i]->dirty) return PP_EARG;
  for (int i = 0; i < g->ninst; i++) {
    pp_pamp *p = g->inst[i];
    PP_HIP(hipMemcpyAsync(p->scal, scal[i], sizeof(double) * g->nsteps * p->nterms, hipMemcpyHostToDevice, stream));
  }
  // the stream form sets this attribute per launch; restore the graph's largest
  PP_HIP(hipFuncSetAttribute(g->fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)g->lds));
  PP_HIP(hipGraphLaunch(g->exec, stream));
  return PP_OK;
}

// Expokit form of the same engine: nsteps ExpokitPropagator.AdvanceStep's
// with zgexpv (tol <= eps means sqrt(eps), anorm = krylov_norm).  Real dt:
// A = -i H^ over |dt|; imaginary dt (dt_re == 0): A = -H^ over |dt_im|.
extern "C" int pp_expokit_propagate(pp_pamp *p, void *psi, int nsteps, double dt_re, double dt_im,
                                    const double *scal, double tol, double anorm, hipStream_t stream)
{
  if (!p || !psi || nsteps < 0 || (!scal && p->nterms > 0) || !(anorm > 0.0)) return PP_EARG;
  if (nsteps == 0) return PP_OK;
  if (p->nterms == 0 || p->m < 2) return PP_EARG;
  if (dt_re != 0.0 && dt_im != 0.0) return PP_EUNSUP;   // complex time steps
  if (p->reorth) return PP_EUNSUP;                      // pAMP-only option
  int rc = pamp_upload_slots(p);   // nslots sizes the split form's LDS
  if (rc) return rc;
  const int mh = p->m + 2;
  const size_t hm = sizeof(double) * 2 * (size_t)(mh * mh + mh);
  const size_t zgp = sizeof(zd) * (size_t)(5 * mh * mh + 5 * mh) + sizeof(double) * (PP_THREADS / 64) + 64;
  // the fiber-split form where it applies and its LDS (stage-2 rows, combo
  // table, split buffers, Hessenberg matrix, zgpadm scratch) fits
  int Q = 1;
  rc = pamp_split_factor(p, Q);
  if (rc) return rc;
  size_t slds = 0;
  if (Q > 1) {
    const pp_band_plan &pl = p->ov->plan;
    const int maxch = (pl.P + Q - 1) / Q;
    slds = sizeof(double) * ((size_t)pp_split_rows_doubles(p->ov->kd, pl.P, maxch) + PP_CMB_DOUBLES +
                             (size_t)pp_split_doubles(p->nslots, p->m, p->ov->kd, pl.B, maxch, pl.P)) + hm + zgp;
    if (slds > 160 * 1024) Q = 1;
  }
  if (Q > 1 && ((rc = pamp_split_buffers(p, Q)) || (rc = pamp_split_rows(p, Q)))) return rc;
  pp_pamp_args A;
  rc = pamp_prepare(p, psi, nsteps, dt_re, dt_im, scal, stream, A);
  if (rc) return rc;
  pp_expo_args E;
  E.tol = tol;
  E.anorm = anorm;
  E.imag = dt_re == 0.0 ? 1 : 0;
  E.t_out = E.imag ? fabs(dt_im) : fabs(dt_re);
  E.stats = p->diag;
  void *args[] = {&A, &E};
  rc = PP_EUNSUP;
  if (Q > 1) {
    A.Q = Q;
    A.sb = p->sb;
    A.subw = (unsigned *)(p->bar + 1) + (size_t)p->L * Q;
    A.srows = p->srows;
    A.sstride = p->sstride;
    A.nlevEven = p->sstride;
    PP_KD_SWITCH(p->ov->kd, {
      rc = pamp_launch(p->mixed ? (const void *)pp_expokit_split_kernel<KD_, true>
                                : (const void *)pp_expokit_split_kernel<KD_, false>,
                       p->L * Q, args, slds, stream); })
    return rc;
  }
  const size_t nle = sizeof(double) * ((size_t)((p->ov->plan.nlev + 1) & ~1) + PP_CMB_DOUBLES);
  size_t fib = pamp_lds_bytes(p, A.padN) + hm;
  size_t pade = nle + hm + zgp;
  const size_t lds = fib > pade ? fib : pade;
  PP_KD_SWITCH(p->ov->kd, {
    rc = pamp_launch(p->mixed ? (const void *)pp_expokit_kernel<KD_, true> : (const void *)pp_expokit_kernel<KD_, false>,
                     p->L, args, lds, stream); })
  return rc;
}

// stats: host nsteps x 6 of the last pp_expokit_propagate call (matvecs,
// exponentials, substeps, rejected substeps, basis size of the last substep,
// zgexpv iflag: 0 ok, 1 more than 500 substeps, -9 Pade failure, -8 more
// than 100 rejections of one substep);
// synchronises the stream.  Returns a HIP error status if a grid barrier
// timed out.
extern "C" int pp_expokit_diagnostics(pp_pamp *p, int nsteps, double *stats, hipStream_t stream)
{
  if (!p || !stats || nsteps > p->diag_cap) return PP_EARG;
  pp_barrier hb;
  PP_HIP(hipMemcpyAsync(stats, p->diag, sizeof(double) * 6 * (size_t)nsteps, hipMemcpyDeviceToHost, stream));
  PP_HIP(hipMemcpyAsync(&hb, p->bar, sizeof(hb), hipMemcpyDeviceToHost, stream));
  PP_HIP(hipStreamSynchronize(stream));
  if (hb.err) return PP_ETIMEOUT;
  return PP_OK;
}

// diag: host nsteps x 4 (error estimate |c_{m-1}|, basis used, |psi~|, re-orth passes)
// of the last pp_pamp_propagate call; synchronises the stream.  Returns a HIP
// error status if a grid barrier timed out.
extern "C" int pp_pamp_diagnostics(pp_pamp *p, int nsteps, double *diag, hipStream_t stream)
{
  if (!p || !diag || nsteps > p->diag_cap) return PP_EARG;
  pp_barrier hb;
  PP_HIP(hipMemcpyAsync(diag, p->diag, sizeof(double) * 4 * (size_t)nsteps, hipMemcpyDeviceToHost, stream));
  PP_HIP(hipMemcpyAsync(&hb, p->bar, sizeof(hb), hipMemcpyDeviceToHost, stream));
  PP_HIP(hipStreamSynchronize(stream));
  if (hb.err) return PP_ETIMEOUT;
  return PP_OK;
}

// Pade scaling of later launches: 0 (default) the largest |A_ij| as pAMP
// (pamp.h:442-445); 1 the 1-norm with one extra halving, so the exponential
// is accurate to rounding and smooth in the inputs (needed by derivatives).
extern "C" int pp_pamp_set_pade_scaling(pp_pamp *p, int mode)
{
  if (!p || mode < 0 || mode > 1) return PP_EARG;
  p->normscale = mode;
  p->gen++;
  return PP_OK;
}

// Exponential of the Krylov matrix in later launches (the reference's
// krylov_exponentiation_method, pampwrapper.cpp:56-61): 0 scaling + Pade(6)
// (pamp.h:430-564), 1 the eigenvector exponential (pamp.h:567-607).
extern "C" int pp_pamp_set_exponentiation(pp_pamp *p, int method)
{
  if (!p || method < 0 || method > 1) return PP_EARG;
  p->expmethod = method;
  p->gen++;
  return PP_OK;
}

// Mixed-precision form: later launches apply the terms whose bit is set in
// term_mask in complex FP32 (FP32 radial bands and band sums); 0 restores the
// all-FP64 drop-in.
extern "C" int pp_pamp_set_mixed(pp_pamp *p, unsigned term_mask)
{
  if (!p) return PP_EARG;
  if (term_mask >> PP_MAXTERMS) return PP_EARG;
  p->mixed = term_mask;
  p->gen++;
  return PP_OK;
}

// Profiling hook: when enabled, later launches record wall-clock stamps
// (100 MHz counter) of workgroup 0 around the phases of step 0.
extern "C" int pp_pamp_set_timing(pp_pamp *p, int enable)
{
  if (!p) return PP_EARG;
  if (enable && !p->tlog) {
    PP_HIP(hipMalloc(&p->tlog, sizeof(long long) * (PP_TLOG_MAX + 1)));
    PP_HIP(hipMemset(p->tlog, 0, sizeof(long long) * (PP_TLOG_MAX + 1)));
  }
  p->timing = enable ? 1 : 0;
  return PP_OK;
}

extern "C" int pp_pamp_timings(pp_pamp *p, long long *out, int max, int *count)
{
  if (!p || !out || !count || !p->tlog) return PP_EARG;
  long long buf[PP_TLOG_MAX + 1];
  PP_HIP(hipDeviceSynchronize());
  PP_HIP(hipMemcpy(buf, p->tlog, sizeof(buf), hipMemcpyDeviceToHost));
  int n = (int)buf[PP_TLOG_MAX];
  if (n > max) n = max;
  for (int i = 0; i < n; i++) out[i] = buf[i];
  *count = n;
  return PP_OK;
}

extern "C" int pp_pamp_destroy(pp_pamp *p)
{
  if (!p) return PP_OK;
  for (int t = 0; t < p->nterms; t++) if (p->termbuf[t]) (void)hipFree(p->termbuf[t]);
  if (p->slotcols) (void)hipFree(p->slotcols);
  if (p->tlog) (void)hipFree(p->tlog);
  if (p->V) (void)hipFree(p->V);
  if (p->y) (void)hipFree(p->y);
  if (p->Hm) (void)hipFree(p->Hm);
  if (p->cvec) (void)hipFree(p->cvec);
  if (p->part) (void)hipFree(p->part);
  if (p->diag) (void)hipFree(p->diag);
  if (p->scal) (void)hipFree(p->scal);
  if (p->bar) (void)hipFree(p->bar);
  if (p->sb) (void)hipFree(p->sb);
  if (p->srows) (void)hipFree(p->srows);
  delete p;
  return PP_OK;
}
