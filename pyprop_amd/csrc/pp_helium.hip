// pAMP propagation of two-electron states on MI355X: coupled angular
// channels x two B-spline radial ranks, psi(c, i1, i2) with i2 contiguous.
//
//   psi~ = (U (x) U) psi;  fixed-m Arnoldi with classical Gram-Schmidt on
//   H^ = (U (x) U)^-H (sum_t s_t(t_n) V_t) (U (x) U)^-1;  exp(-i dt H_m) e1
//   by scaling + Pade(6);  psi = (U (x) U)^-1 |psi~| V c.
//
// Reference: pyprop/propagator/PamPropagator.py:50-75 and
// core/krylov/pamp/pamp/pamp.h:271-350, :430-564, :613-669 (the step, as
// for one radial rank), with the balanced matvec of
// pyprop/tensorpotential/BasisPropagator.py:131-159 over the rank-3
// potentials of examples/tensor/helium_stabilization/config.ini
// (Simp_Simp_Simp one-electron terms, Simp_Simp_BandNH electron-electron
// term, generated by TensorPotentialMultiply_* of
// core/tensorpotential/tensorpotentialmultiply_generator.py).
//
// The state is 0.95 M amplitudes at the helium workload, far more than one
// workgroup per fiber can keep resident, so this engine is a stream of
// ordinary launches on one stream; the host issues a whole step without
// synchronising (scalars stay on the device and later kernels read them).
//
// Operator, factored (no reference-layout tensor is formed):
//  * Kronecker terms  sum_p a_p |c_p><c'_p| (x) A (x) B  with real banded
//    A, B (the one-electron terms, the velocity-gauge dipole).  They are
//    applied in two banded passes along i2 then i1 through intermediates:
//    for a term whose B is shared by other terms the i2 pass is plain
//    (X = y B^T per channel) and the angular mixing happens in the i1 pass;
//    otherwise the mixing happens in the i2 pass and the i1 pass is plain.
//  * Multipole terms  sum_p a_p |c_p><c'_p| (x) R  with a real 4-index
//    radial tensor R[i1][d1][i2][d2] (the r<^l / r>^(l+1) term of one l).
//    The angular mixing is applied first (one "mixed input" per term and
//    16-channel chunk, channel-innermost), then one workgroup per (chunk,
//    i1, half of i2) runs the 169-tap contraction of every term for 16
//    channels at once: R slices and mixed input rows staged in LDS, each
//    thread accumulating 2 channels x 4 consecutive i2 in registers so one
//    R value feeds two channels and one input value feeds up to four
//    outputs.
#include "pp_common.h"
#include <cstring>
#include <cstdio>
#include <cmath>
#include <vector>
#include <algorithm>

#define PP_THREADS 256
#include "pp_pade.h"

#define P3_CHUNK 16        // channels per multipole chunk
#ifndef P3_RED_BLOCKS
#define P3_RED_BLOCKS 480  // blocks of the Gram-Schmidt reductions
#endif

struct p3_ent {            // one coupling of a banded pass or the mixing
  int src, col, mat, slot; // source buffer, source channel, radial matrix, time slot (-1: 1.0)
  double cre, cim;
};
struct p3_grp {            // entries of one i2-pass row sharing a band matrix and time slot
  int m0, m1, mat, slot;   // their channel mix: p3_ent m0 .. m1-1 (col, cre, cim)
};

// Working precision of the split multipole stage: the FP64 form keeps the
// mixed input, the far-field intermediates and the radial tables in double;
// the mixed form (pp_pamp3_set_mixed) stores them as complex64 / float and
// accumulates in FP32, adding each output into the FP64 result once.
template <class T> struct p3_cx;
template <> struct p3_cx<double> { typedef double2 type; };
template <> struct p3_cx<float> { typedef float2 type; };
template <class Z> __device__ inline Z p3_czero();
template <> __device__ inline double2 p3_czero<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ inline float2 p3_czero<float2>() { return make_float2(0.f, 0.f); }
template <class Z> __device__ inline Z p3_cfrom(zd v);
template <> __device__ inline double2 p3_cfrom<double2>(zd v) { return v; }
template <> __device__ inline float2 p3_cfrom<float2>(zd v) { return make_float2((float)v.x, (float)v.y); }
__device__ inline zd p3_ctod(double2 v) { return v; }
__device__ inline zd p3_ctod(float2 v) { return zmk((double)v.x, (double)v.y); }
__device__ inline double2 p3_cadd(double2 a, double2 b) { return zadd(a, b); }
__device__ inline float2 p3_cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline void p3_cfmar(double2 &a, double s, double2 c) { zfmar(a, s, c); }
__device__ inline void p3_cfmar(float2 &a, float s, float2 c)
{
  a.x = fmaf(s, c.x, a.x);
  a.y = fmaf(s, c.y, a.y);
}

struct pp_pamp3 {
  int C, N, m, s, bw, reorth;
  long n;
  const pp_overlap *ov;
  // host description
  std::vector<std::vector<double>> mats;
  struct kterm { std::vector<int> pr; std::vector<double> cf; int A, B, slot; };
  std::vector<kterm> kterms;
  struct mterm {
    std::vector<int> pr; std::vector<double> cf; double *R; int slot;
    std::vector<double> Rh, Pl, Ql;   // host copies for the near/far split (Pl, Ql empty: no split)
    double *dRn, *dP, *dQ;
    float *fRn, *fP, *fQ;             // the same in FP32 (mixed form)
  };
  std::vector<int> sup_lo, sup_hi;     // quadrature-interval support of each B-spline
  bool split;
  int2 *nr;
  std::vector<int2> nrh;               // near band per i1 (host copy, for the flop count)
  const double **dRn, **dPt, **dQt;
  const float **fRnt, **fPt, **fQt;    // per-term FP32 tables of the mixed form
  int mixed;                           // multipole stage in FP32 (complex64 intermediates)
  int expmethod;                       // 0 Pade, 1 eigenvector exponential
  zd *Tq, *Tp;
  std::vector<mterm> mterms;
  int nslots;
  bool dirty;
  // device plan
  double *dmats;
  int K;                   // intermediates of the banded passes
  int *rrow, *lrow, *mrow; // CSR row pointers (right: K*C rows of groups, left: C, mix: nblk*16)
  p3_grp *rgrp;            // right pass: (matrix, slot) groups, each a channel mix in rent
  p3_ent *rent, *lent, *ment;
  int nblk;                // multipole (term, chunk) blocks
  int nchunk;
  int *cb_ptr, *cb_blk;    // per chunk: its blocks
  int *blk_term;
  const double **dR;       // per multipole term: device R
  // buffers
  zd *Y, *X, *Z, *P, *V, *Ym;
  double *part, *scal, *diag, *scal_dt;
  double *urd;             // reciprocals of U's diagonal (the fiber solves)
  double *tinv;            // U^-1, U^-T and U as MFMA blocks (p3_tgemm; null: substitution tiles)
  int tnp;                 // their padded size (0: substitution tiles)
  zd *Hm, *cvec, *hcoef;
  double *beta;
  int *state;              // [0] broken, [1] mEff
  int scal_cap, diag_cap;
  // optional HIP-event timing of the multipole contraction launches
  int timing;
  std::vector<hipEvent_t> ev;   // pairs (before, after) of the current propagate
  double ee_ms;
  long ee_count;
};

// ------------------------------------------------------------------ banded passes
// Logical block of dispatched block b of nb: blocks go round-robin over the
// 8 XCDs (b % 8), so logical blocks L = 0, 1, 2, ... are given to one XCD in
// order and neighbouring logical blocks share that XCD's L2 (a bijection for
// any nb).
__device__ inline int p3_xcd_order(int b, int nb)
{
  const int x = b & 7, slot = b >> 3, q = nb >> 3, rem = nb & 7;
  return x * q + (x < rem ? x : rem) + slot;
}

#define P3_RT 8            // rows per workgroup of the banded passes
// X[k][c][i1][i2] = sum_g sum_d B_g[i2][d] * m_g[i1][i2+d-s],
// m_g = sum_{e in g} coef_e * s_e * y[c'_e]: the row's couplings grouped by
// band matrix B_g and time slot, each group's channel mix formed first
// (rows i1a .. i1a+P3_RT-1 and the i2 window of the chunk with its halo, in
// LDS) and its 13-tap stencil applied once.  workgroup: (row r = k*C + c,
// P3_RT rows i1, P3_RC columns i2), thread: one i2.  The mixing loads of
// four window elements are issued together.  Logical blocks run tile-major
// (every row r of one i1 tile, then the next tile), so the workgroups that
// read the same rows of y share an XCD.
#define P3_RC 128          // i2 columns per workgroup of the i2 pass (= its threads)
template <int S>
__global__ void __launch_bounds__(P3_RC) p3_right(const double *mats, const int *grow, const p3_grp *grp,
                                                   const p3_ent *mix, const zd *y, zd *X, int C, int N,
                                                   const double *scal, const int *state)
{
  if (state[0]) return;
  constexpr int s = S, bw = 2 * S + 1, WM = P3_RC + 2 * S, NE = P3_RT * WM;
  __shared__ zd m[P3_RT * WM];
  const int ntile = (N + P3_RT - 1) / P3_RT, nch = (N + P3_RC - 1) / P3_RC, nr = gridDim.x / (ntile * nch);
  const int L = p3_xcd_order(blockIdx.x, gridDim.x);
  const int r = L % nr, tc = L / nr, i1a = (tc / nch) * P3_RT, i2c = (tc % nch) * P3_RC;   // r = k*C + c
  const long NN = (long)N * N;
  const int tid = threadIdx.x, i2 = i2c + tid;
  zd acc[P3_RT];
#pragma unroll
  for (int rr = 0; rr < P3_RT; rr++) acc[rr] = zmk(0.0, 0.0);
  for (int gi = grow[r]; gi < grow[r + 1]; gi++) {
    const p3_grp G = grp[gi];
    const double sc = G.slot < 0 ? 1.0 : scal[G.slot];
    for (int base = 0; base < NE; base += 4 * P3_RC) {
      long off[4];
      bool ok[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int idx = base + tid + u * P3_RC, rr = idx / WM, j = i2c - s + (idx - rr * WM), i1 = i1a + rr;
        ok[u] = idx < NE && i1 < N && j >= 0 && j < N;
        off[u] = ok[u] ? (long)i1 * N + j : 0;   // clamped: the loads need no branches
      }
      zd v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = zmk(0.0, 0.0);
      for (int e = G.m0; e < G.m1; e++) {
        const p3_ent E = mix[e];
        const zd cf = zmk(E.cre * sc, E.cim * sc);
        const zd *src = y + E.col * NN;
        zd t[4];
#pragma unroll
        for (int u = 0; u < 4; u++) t[u] = src[off[u]];
#pragma unroll
        for (int u = 0; u < 4; u++) zfma(v[u], cf, t[u]);
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int idx = base + tid + u * P3_RC;
        if (idx < NE) m[idx] = ok[u] ? v[u] : zmk(0.0, 0.0);
      }
    }
    __syncthreads();
    if (i2 < N) {
      const double *Bm = mats + ((size_t)G.mat * N + i2) * bw;
      double bv[bw];
#pragma unroll
      for (int d = 0; d < bw; d++) bv[d] = Bm[d];
      const int d0 = s - i2 > 0 ? s - i2 : 0, d1 = N - 1 - i2 + s < bw - 1 ? N - 1 - i2 + s : bw - 1;
#pragma unroll
      for (int rr = 0; rr < P3_RT; rr++) {
        zd t = zmk(0.0, 0.0);
#pragma unroll
        for (int d = 0; d < bw; d++)
          if (d >= d0 && d <= d1) zfmar(t, bv[d], m[rr * WM + tid + d]);
        acc[rr] = zadd(acc[rr], t);
      }
    }
    __syncthreads();
  }
  if (i2 < N)
#pragma unroll
    for (int rr = 0; rr < P3_RT; rr++)
      if (i1a + rr < N) X[(long)r * NN + (long)(i1a + rr) * N + i2] = acc[rr];
}

// Z[c][i1][i2] = sum_e coef_e * s_e * sum_d A_e[i1][d] * X[k_e][c'_e][i1+d-s][i2]
// workgroup: (channel c, P3_LT rows i1); thread: i2.  Per entry the column
// window X[.][i1a-s .. i1a+P3_LT-1+s][i2] is loaded once into registers.
// Logical blocks run tile-major, so neighbouring tiles (overlapping windows)
// and the channels that read the same sources share an XCD.
#ifndef P3_LT
#define P3_LT 4            // rows per workgroup of the i1 pass
#endif
template <int S>
__global__ void __launch_bounds__(128) p3_left(const double *mats, const int *row, const p3_ent *ent, const zd *X,
                                                zd *Z, int C, int N, const double *scal, const int *state)
{
  if (state[0]) return;
  constexpr int s = S, bw = 2 * S + 1, WIN = P3_LT + 2 * S;
  const int L = p3_xcd_order(blockIdx.x, gridDim.x);
  const int c = L % C, i1a = (L / C) * P3_LT;
  const long NN = (long)N * N;
  for (int i2 = threadIdx.x; i2 < N; i2 += blockDim.x) {
    zd acc[P3_LT];
#pragma unroll
    for (int rr = 0; rr < P3_LT; rr++) acc[rr] = zmk(0.0, 0.0);
    for (int e = row[c]; e < row[c + 1]; e++) {
      const p3_ent E = ent[e];
      const double sc = E.slot < 0 ? 1.0 : scal[E.slot];
      const zd cf = zmk(E.cre * sc, E.cim * sc);
      const zd *src = X + ((long)E.src * C + E.col) * NN + i2;
      zd win[WIN];
#pragma unroll
      for (int w = 0; w < WIN; w++) {
        const int j1 = i1a - s + w;
        win[w] = (j1 >= 0 && j1 < N) ? src[(long)j1 * N] : zmk(0.0, 0.0);
      }
#pragma unroll
      for (int rr = 0; rr < P3_LT; rr++) {
        const int i1 = i1a + rr;
        if (i1 >= N) break;
        const double *Am = mats + ((size_t)E.mat * N + i1) * bw;   // uniform over the workgroup
        zd t = zmk(0.0, 0.0);
#pragma unroll
        for (int d = 0; d < bw; d++) zfmar(t, Am[d], win[rr + d]);   // A is zero outside the matrix
        zfma(acc[rr], cf, t);
      }
    }
#pragma unroll
    for (int rr = 0; rr < P3_LT; rr++)
      if (i1a + rr < N) Z[(long)c * NN + (long)(i1a + rr) * N + i2] = acc[rr];
  }
}

// ------------------------------------------------------------------ multipole terms
// Ym[b][j1][j2][cl] = sum_e coef_e * s_e * y[c'_e][j1][j2], row b*16 + cl
__global__ void __launch_bounds__(256) p3_mix(const int *row, const p3_ent *ent, const zd *y, zd *Ym, int N,
                                               const double *scal, const int *state)
{
  if (state[0]) return;
  const int b = blockIdx.x / N, j1 = blockIdx.x % N;
  const long NN = (long)N * N;
  const int cl = threadIdx.x & (P3_CHUNK - 1);
  const int r = b * P3_CHUNK + cl;
  const int e0 = row[r], e1 = row[r + 1];
  zd *out = Ym + ((long)b * NN + (long)j1 * N) * P3_CHUNK;
  for (int j2 = threadIdx.x / P3_CHUNK; j2 < N; j2 += 256 / P3_CHUNK) {
    zd acc = zmk(0.0, 0.0);
    for (int e = e0; e < e1; e++) {
      const p3_ent E = ent[e];
      const double sc = E.slot < 0 ? 1.0 : scal[E.slot];
      zfma(acc, zmk(E.cre * sc, E.cim * sc), y[E.col * NN + (long)j1 * N + j2]);
    }
    out[(long)j2 * P3_CHUNK + cl] = acc;
  }
}

#define P3_HALF 64         // i2 per multipole workgroup
#define P3_EE_THREADS 128  // (8 channel pairs) x (16 i2 quads)

// Z[c][i1][i2] += sum_blocks sum_{d1,d2} R_t[i1][d1][i2][d2] Ym[b][i1+d1-s][i2+d2-s][cl]
// workgroup: (chunk, i1, i2 half); thread: channel pair cp (8) x i2 quad q (16).
// Many small workgroups (up to six per CU) keep enough loads in flight: the
// stages are short (one R slice and one mixed input row), and a register-
// blocked variant with three i1 rows per workgroup and the next stage
// prefetched into registers measured slower (fewer resident waves).
template <int S>
__global__ void __launch_bounds__(P3_EE_THREADS) p3_ee(const double *const *Rt, const int *blk_term, const int *cb_ptr,
                                                       const int *cb_blk, const zd *Ym, zd *Z, int C, int N,
                                                       int nhalf, int nchunk, const int *state)
{
  if (state[0]) return;
  constexpr int BW = 2 * S + 1;
  constexpr int WIN = P3_HALF + 2 * S;                  // staged input positions
  __shared__ double Rs[P3_HALF * BW];                   // [i2 local][d2]
  __shared__ zd Ys[WIN * P3_CHUNK];                     // [position][cl]
  // blockIdx -> (group of 8 WGs on one XCD) so the chunks of one (i1, half)
  // share the XCD's L2 for their R slices
  const int g = blockIdx.x, xcd = g & 7, w = g >> 3;
  const int pairIdx = (w / nchunk) * 8 + xcd;           // (i1, half) pair
  const int chunk = w % nchunk;
  if (pairIdx >= N * nhalf) return;
  const int i1 = pairIdx / nhalf, half = pairIdx % nhalf;
  const int i2base = half * P3_HALF;
  const int nloc = N - i2base < P3_HALF ? N - i2base : P3_HALF;
  const long NN = (long)N * N;
  const int tid = threadIdx.x, cp = tid >> 4, q = tid & 15;
  zd acc[2][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int t = 0; t < 4; t++) acc[a][t] = zmk(0.0, 0.0);
  for (int bi = cb_ptr[chunk]; bi < cb_ptr[chunk + 1]; bi++) {
    const int b = cb_blk[bi];
    const double *R = Rt[blk_term[b]];
    const zd *Yb = Ym + (long)b * NN * P3_CHUNK;
    for (int d1 = 0; d1 < BW; d1++) {
      const int j1 = i1 + d1 - S;
      if (j1 < 0 || j1 >= N) continue;      // uniform over the workgroup
      __syncthreads();
      const double *Rrow = R + (((size_t)i1 * BW + d1) * N + i2base) * BW;
      for (int e = tid; e < nloc * BW; e += P3_EE_THREADS) Rs[e] = Rrow[e];
      for (int e = nloc * BW + tid; e < P3_HALF * BW; e += P3_EE_THREADS) Rs[e] = 0.0;
      const zd *Yrow = Yb + (long)j1 * N * P3_CHUNK;
      for (int e = tid; e < WIN * P3_CHUNK; e += P3_EE_THREADS) {
        const int p = i2base - S + e / P3_CHUNK;
        Ys[e] = (p >= 0 && p < N) ? Yrow[(long)p * P3_CHUNK + (e & (P3_CHUNK - 1))] : zmk(0.0, 0.0);
      }
      __syncthreads();
      // outputs i2 = 4q + t; input position of tap d2 is 4q + t + d2 (window base i2base - S)
      const double *Rq = Rs + (4 * q) * BW;
      const zd *Yq = Ys + (4 * q) * P3_CHUNK + 2 * cp;
      zd w0[4], w1[4];
#pragma unroll
      for (int t = 0; t < 3; t++) { w0[t + 1] = Yq[t * P3_CHUNK]; w1[t + 1] = Yq[t * P3_CHUNK + 1]; }
#pragma unroll
      for (int d2 = 0; d2 < BW; d2++) {
        // slide: window holds positions d2 .. d2+3
#pragma unroll
        for (int t = 0; t < 3; t++) { w0[t] = w0[t + 1]; w1[t] = w1[t + 1]; }
        w0[3] = Yq[(d2 + 3) * P3_CHUNK];
        w1[3] = Yq[(d2 + 3) * P3_CHUNK + 1];
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const double r = Rq[t * BW + d2];
          zfmar(acc[0][t], r, w0[t]);
          zfmar(acc[1][t], r, w1[t]);
        }
      }
    }
  }
  const int cbase = chunk * P3_CHUNK + 2 * cp;
#pragma unroll
  for (int a = 0; a < 2; a++) {
    const int c = cbase + a;
    if (c >= C) continue;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int i2l = 4 * q + t;
      if (i2l >= nloc) continue;
      zd *z = Z + (long)c * NN + (long)i1 * N + i2base + i2l;
      *z = zadd(*z, acc[a][t]);
    }
  }
}

// ------------------------------------------------------------------ multipole terms, near/far split
// Where the supports of B_i1 and B_i2 do not overlap (i2 > nhi[i1]: every
// quadrature point of electron 1 lies below every point of electron 2;
// i2 < nlo[i1]: above), r<^l / r>^(l+1) is a product on the whole product-
// quadrature support and R_l[i1, j1, i2, j2] = P_l[i1, j1] Q_l[i2, j2] (resp.
// Q_l[i1, j1] P_l[i2, j2]) exactly, with P_l = <B|r^l|B>, Q_l = <B|r^-(l+1)|B>.
// Those outputs take two banded passes over the mixed input; only the near
// band nlo[i1] <= i2 <= nhi[i1] (at most 2k-1 wide) needs the 169-tap
// contraction, with R stored for that band only.
#define P3_NW 16           // near-band width handled per i1 (>= 2k-1)
#define P3_MAXB 32         // multipole blocks per chunk (terms l <= 31)

// Mixing fused with the far field's i2 pass: workgroup (block b, row j1)
// forms the mixed input row Ym[b][j1][.][.] in LDS (written out for the
// near band), then Tq/Tp of that row from LDS.
#define P3_MIXE 12         // mixing entries per row held in registers
template <int S, class T>
__global__ void __launch_bounds__(256) p3_mixT(const int *row, const p3_ent *ent, const zd *y,
                                               typename p3_cx<T>::type *Ym, const T *const *Pt, const T *const *Qt,
                                               const int *blk_term, const int2 *nr, typename p3_cx<T>::type *Tq,
                                               typename p3_cx<T>::type *Tp, int N, const double *scal,
                                               const int *state)
{
  typedef typename p3_cx<T>::type Z;
  if (state[0]) return;
  constexpr int BW = 2 * S + 1;
  extern __shared__ double p3_lds[];
  Z *rowb = (Z *)p3_lds;                                 // positions -S .. N+S-1, 16 channels each
  const int b = blockIdx.x / N, j1 = blockIdx.x % N;
  const long NN = (long)N * N;
  const int cl = threadIdx.x & (P3_CHUNK - 1);
  const int r = b * P3_CHUNK + cl;
  const int e0 = row[r], ne = row[r + 1] - e0;
  // this lane's couplings: source row offset and scaled coefficient
  long src[P3_MIXE];
  zd cf[P3_MIXE];
#pragma unroll
  for (int q = 0; q < P3_MIXE; q++) {
    src[q] = 0;
    cf[q] = zmk(0.0, 0.0);
    if (q < ne) {
      const p3_ent E = ent[e0 + q];
      const double sc = E.slot < 0 ? 1.0 : scal[E.slot];
      src[q] = E.col * NN + (long)j1 * N;
      cf[q] = zmk(E.cre * sc, E.cim * sc);
    }
  }
  // the near band reads this row only at positions nlo(j1-S)-S .. nhi(j1+S)+S
  const int wlo = nr[j1 - S > 0 ? j1 - S : 0].x - S, whi = nr[j1 + S < N - 1 ? j1 + S : N - 1].y + S;
  Z *out = Ym + ((long)b * NN + (long)j1 * N) * P3_CHUNK;
  for (int e = threadIdx.x; e < (N + 2 * S) * P3_CHUNK; e += 256) {
    const int pos = e / P3_CHUNK - S;
    zd acc = zmk(0.0, 0.0);
    if (pos >= 0 && pos < N) {
#pragma unroll
      for (int q = 0; q < P3_MIXE; q++)
        if (q < ne) zfma(acc, cf[q], y[src[q] + pos]);
      for (int q = P3_MIXE; q < ne; q++) {               // rows with more couplings
        const p3_ent E = ent[e0 + q];
        const double sc = E.slot < 0 ? 1.0 : scal[E.slot];
        zfma(acc, zmk(E.cre * sc, E.cim * sc), y[E.col * NN + (long)j1 * N + pos]);
      }
      if (pos >= wlo && pos <= whi) out[(long)pos * P3_CHUNK + cl] = p3_cfrom<Z>(acc);
    }
    rowb[e] = p3_cfrom<Z>(acc);
  }
  __syncthreads();
  const T *P = Pt[blk_term[b]], *Q = Qt[blk_term[b]];
  Z *oq = Tq + ((long)b * NN + (long)j1 * N) * P3_CHUNK, *op = Tp + ((long)b * NN + (long)j1 * N) * P3_CHUNK;
  // p3_farL reads row j1 of Tq only for outputs i1 <= j1 + S with
  // i2 > nhi[i1] (so i2 > nhi[j1 - S], nhi non-decreasing) and of Tp only for
  // i1 >= j1 - S with i2 < nlo[i1] (so i2 < nlo[j1 + S]): the rest of each row
  // is neither computed nor stored.
  const int qlo = nr[j1 - S > 0 ? j1 - S : 0].y, phi = nr[j1 + S < N - 1 ? j1 + S : N - 1].x;
  for (int e = threadIdx.x; e < N * P3_CHUNK; e += 256) {
    const int i2 = e / P3_CHUNK, c2 = e & (P3_CHUNK - 1);
    if (i2 > qlo) {
      const T *qr = Q + (size_t)i2 * BW;
      Z tq = p3_czero<Z>();
#pragma unroll
      for (int d = 0; d < BW; d++) p3_cfmar(tq, qr[d], rowb[(i2 + d) * P3_CHUNK + c2]);
      oq[e] = tq;
    }
    if (i2 < phi) {
      const T *pr = P + (size_t)i2 * BW;
      Z tp = p3_czero<Z>();
#pragma unroll
      for (int d = 0; d < BW; d++) p3_cfmar(tp, pr[d], rowb[(i2 + d) * P3_CHUNK + c2]);
      op[e] = tp;
    }
  }
}

// Z[c][i1][i2] += sum_blocks [ i2 > nhi[i1]:  sum_d1 P[i1][d1] Tq[b][i1+d1-S][i2][cl]
//                              i2 < nlo[i1]:  sum_d1 Q[i1][d1] Tp[b][i1+d1-S][i2][cl] ]
// workgroup: (chunk, P3_RT rows i1, 16 columns i2); thread: (i2, channel lane).
#ifndef P3_FRT
#define P3_FRT 12          // output rows i1 per far-field workgroup (its 24-row window: 2x re-read, 8 rows: 2.5x)
#endif
template <int S, class T>
__global__ void __launch_bounds__(256) p3_farL(const T *const *Pt, const T *const *Qt, const int *blk_term,
                                               const int *cb_ptr, const int *cb_blk, const int2 *nr,
                                               const typename p3_cx<T>::type *Tq, const typename p3_cx<T>::type *Tp,
                                               zd *Z, int C, int N, int nchunk, const int *state)
{
  typedef typename p3_cx<T>::type ZT;
  if (state[0]) return;
  constexpr int BW = 2 * S + 1, WIN = P3_FRT + 2 * S;
  const int ntile = (N + P3_FRT - 1) / P3_FRT, ngrp = (N + 15) / 16;
  int g = blockIdx.x;
  const int chunk = g % nchunk; g /= nchunk;
  const int i1a = (g % ntile) * P3_FRT; g /= ntile;
  const int i2 = g * 16 + (threadIdx.x >> 4), cl = threadIdx.x & (P3_CHUNK - 1);
  (void)ngrp;
  const long NN = (long)N * N;
  const bool col_ok = i2 < N;
  bool below[P3_FRT], above[P3_FRT];
#pragma unroll
  for (int rr = 0; rr < P3_FRT; rr++) {
    const int i1 = i1a + rr;
    const int2 r = i1 < N ? nr[i1] : make_int2(N, -1);
    below[rr] = col_ok && i1 < N && i2 > r.y;
    above[rr] = col_ok && i1 < N && i2 < r.x;
  }
  bool anyb = false, anya = false;
#pragma unroll
  for (int rr = 0; rr < P3_FRT; rr++) { anyb = anyb || below[rr]; anya = anya || above[rr]; }
  __shared__ const T *sP[P3_MAXB];
  __shared__ const T *sQ[P3_MAXB];
  __shared__ long sT[P3_MAXB];
  const int b0 = cb_ptr[chunk], nb = cb_ptr[chunk + 1] - b0;
  for (int k = threadIdx.x; k < nb; k += 256) {
    const int b = cb_blk[b0 + k];
    sP[k] = Pt[blk_term[b]];
    sQ[k] = Qt[blk_term[b]];
    sT[k] = (long)b * NN * P3_CHUNK;
  }
  __syncthreads();
  ZT acc[P3_FRT];
#pragma unroll
  for (int rr = 0; rr < P3_FRT; rr++) acc[rr] = p3_czero<ZT>();
  for (int k = 0; k < nb; k++) {
    const T *P = sP[k], *Q = sQ[k];
    const long tb = sT[k];
    ZT win[WIN];
    if (anyb) {
#pragma unroll
      for (int w = 0; w < WIN; w++) {
        const int j1 = i1a - S + w;
        win[w] = (j1 >= 0 && j1 < N && col_ok) ? Tq[tb + ((long)j1 * N + i2) * P3_CHUNK + cl] : p3_czero<ZT>();
      }
#pragma unroll
      for (int rr = 0; rr < P3_FRT; rr++) {
        if (!below[rr]) continue;
        const T *pr = P + (size_t)(i1a + rr) * BW;
        ZT t = p3_czero<ZT>();
#pragma unroll
        for (int d = 0; d < BW; d++) p3_cfmar(t, pr[d], win[rr + d]);
        acc[rr] = p3_cadd(acc[rr], t);
      }
    }
    if (anya) {
#pragma unroll
      for (int w = 0; w < WIN; w++) {
        const int j1 = i1a - S + w;
        win[w] = (j1 >= 0 && j1 < N && col_ok) ? Tp[tb + ((long)j1 * N + i2) * P3_CHUNK + cl] : p3_czero<ZT>();
      }
#pragma unroll
      for (int rr = 0; rr < P3_FRT; rr++) {
        if (!above[rr]) continue;
        const T *qr = Q + (size_t)(i1a + rr) * BW;
        ZT t = p3_czero<ZT>();
#pragma unroll
        for (int d = 0; d < BW; d++) p3_cfmar(t, qr[d], win[rr + d]);
        acc[rr] = p3_cadd(acc[rr], t);
      }
    }
  }
  const int c = chunk * P3_CHUNK + cl;
  if (c >= C || !col_ok) return;
#pragma unroll
  for (int rr = 0; rr < P3_FRT; rr++) {
    if (!(below[rr] || above[rr])) continue;
    const long o = (long)c * NN + (long)(i1a + rr) * N + i2;
    Z[o] = zadd(Z[o], p3_ctod(acc[rr]));
  }
}

// Z[c][i1][i2] += sum_blocks sum_{d1,d2} Rn_t[i1][d1][w][d2] Ym[b][i1+d1-S][i2+d2-S][cl]
// for the near band i2 = nlo[i1] + w <= nhi[i1]; workgroup: (chunk, i1),
// P3_NG wavefronts = P3_NG groups x (4 w-quads x 16 channel lanes).  A lane
// computes the 4 outputs w = 4 wq .. 4 wq + 3 of its channel, so each input
// row is loaded once per lane as 16 consecutive positions shared by the 4
// outputs' 13-tap windows (a quarter of the loads of one output per lane).
// The (block, tap row d1) pairs of the workgroup are dealt round-robin to the
// groups (pair index g, g + P3_NG, ...: equal shares), which keeps P3_NG
// short dependent load chains in flight per workgroup; the groups' partial
// sums are added in LDS in group order, so the result is deterministic.
// Lanes whose outputs fall past the band accumulate values that are never
// stored.
#ifndef P3_NG
#define P3_NG 16
#endif
template <int S, class T>
__global__ void __launch_bounds__(64 * P3_NG) p3_near(const T *const *Rn, const int *blk_term, const int *cb_ptr,
                                                      const int *cb_blk, const int2 *nr,
                                                      const typename p3_cx<T>::type *Ym, zd *Z, int C, int N,
                                                      int nchunk, const int *state)
{
  typedef typename p3_cx<T>::type ZT;
  if (state[0]) return;
  constexpr int BW = 2 * S + 1, WIN = 4 + 2 * S;
  __shared__ const T *sR[P3_MAXB];
  __shared__ long sY[P3_MAXB];
  __shared__ ZT part[P3_NG][64 * 4];
  const int chunk = blockIdx.x % nchunk, i1 = blockIdx.x / nchunk;
  const long NN = (long)N * N;
  const int b0 = cb_ptr[chunk], nb = cb_ptr[chunk + 1] - b0;
  for (int k = threadIdx.x; k < nb; k += 64 * P3_NG) {
    const int b = cb_blk[b0 + k];
    sR[k] = Rn[blk_term[b]] + (size_t)i1 * BW * P3_NW * BW;
    sY[k] = (long)b * NN * P3_CHUNK;
  }
  __syncthreads();
  const int grp = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wq = lane >> 4, cl = lane & (P3_CHUNK - 1);
  const int2 r = nr[i1];
  const int w0 = 4 * wq, p0 = r.x + w0 - S;     // first input position of this lane's window
  const bool any = r.x + w0 <= r.y;
  // tap rows with j1 = i1 + d1 - S inside the grid
  const int d1lo = S - i1 > 0 ? S - i1 : 0, d1hi = N - 1 - i1 + S < BW - 1 ? N - 1 - i1 + S : BW - 1;
  const int nd = d1hi - d1lo + 1;
  ZT acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = p3_czero<ZT>();
  // this wavefront's copy of R[d1][.][.] for the current (block, d1): the
  // coefficients are shared by the 16 channel lanes, so one coalesced load
  // per lane fills it and the taps read LDS instead of issuing 52 loads each
  __shared__ T rs[P3_NG][P3_NW * BW];
  T *rw = rs[grp];
  {
    bool pv[WIN];
#pragma unroll
    for (int p = 0; p < WIN; p++) pv[p] = any && p0 + p >= 0 && p0 + p < N;
    for (int it = grp; it < nb * nd; it += P3_NG) {
      const int k = it / nd, d1 = d1lo + it % nd;
      const int j1 = i1 + d1 - S;
      const T *Rd = sR[k] + (size_t)d1 * P3_NW * BW;
      for (int e = lane; e < P3_NW * BW; e += 64) rw[e] = Rd[e];
      const ZT *Yr = Ym + sY[k] + cl + (long)p0 * P3_CHUNK + (long)j1 * N * P3_CHUNK;
      ZT y[WIN];
#pragma unroll
      for (int p = 0; p < WIN; p++) y[p] = pv[p] ? Yr[(long)p * P3_CHUNK] : p3_czero<ZT>();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const T *Rr = rw + (size_t)w0 * BW;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        ZT a = p3_czero<ZT>();
#pragma unroll
        for (int d2 = 0; d2 < BW; d2++) p3_cfmar(a, Rr[t * BW + d2], y[t + d2]);
        acc[t] = p3_cadd(acc[t], a);
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
#pragma unroll
  for (int t = 0; t < 4; t++) part[grp][lane * 4 + t] = acc[t];
  __syncthreads();
  if (grp != 0) return;
  const int c = chunk * P3_CHUNK + cl;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int i2 = r.x + w0 + t;
    if (c >= C || i2 > r.y) continue;
    ZT v = part[0][lane * 4 + t];
#pragma unroll
    for (int g = 1; g < P3_NG; g++) v = p3_cadd(v, part[g][lane * 4 + t]);
    const long o = (long)c * NN + (long)i1 * N + i2;
    Z[o] = zadd(Z[o], p3_ctod(v));
  }
}

// ------------------------------------------------------------------ overlap solves
// Fiber tiles in LDS for both radial ranks.  A workgroup stages P3_FT
// fibers (rank 2: P3_FT consecutive rows; rank 1: P3_FT consecutive columns
// of one channel) with coalesced loads, the band of U and the reciprocals of
// its diagonal, then one thread per fiber runs the substitution in LDS and
// the tile is stored back coalesced.  OP 0: U x = b (back substitution,
// ztbsv 'U','N'), OP 1: U^T x = b (forward; U is real so this is U^H, ztbsv
// 'U','C'), OP 2: x := U x (ztbmv).  U(i, j) at Ub[j*(KD+1) + KD + i - j]
// (overlapmatrix.cpp:22-32).
#ifndef P3_FT
#define P3_FT 32           // fibers per workgroup
#endif
#ifndef P3_FTT
#define P3_FTT 256         // threads per workgroup
#endif
#define P3_SB 8            // rows per group of the substitution's interior
template <int KD, int OP>
__global__ void __launch_bounds__(P3_FTT) p3_ftile(const double *Ub, const double *rdg, zd *x, int C, int N, int rank)
{
  extern __shared__ double p3_lds[];
  double *uL = p3_lds;                                   // N*(KD+1) band
  double *rd = uL + (size_t)N * (KD + 1);                // N reciprocals of the diagonal
  const int toff = N * (KD + 1) + N + ((N * (KD + 1) + N) & 1);   // doubles before the tile (16-byte aligned)
  zd *tile = (zd *)(p3_lds + toff);                      // fiber f, element i at f*(N+1) + i
  const int ld = N + 1;
  const long NN = (long)N * N;
  for (int e = threadIdx.x; e < N * (KD + 1); e += P3_FTT) uL[e] = Ub[e];
  for (int i = threadIdx.x; i < N; i += P3_FTT) rd[i] = rdg[i];
  int nf;
  zd *g;
  long gstr, fstr;   // global element i of fiber f at g[f*fstr + i*gstr]
  if (rank == 2) {
    const long r0 = (long)blockIdx.x * P3_FT, nrow = (long)C * N;
    nf = nrow - r0 < P3_FT ? (int)(nrow - r0) : P3_FT;
    g = x + r0 * N; fstr = N; gstr = 1;
    const int tot = nf * N;
    for (int e0 = 0; e0 < tot; e0 += 8 * P3_FTT) {   // 8 loads in flight per thread (clamped: no branches)
      zd v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) { const int e = e0 + threadIdx.x + u * P3_FTT; v[u] = g[e < tot ? e : tot - 1]; }
#pragma unroll
      for (int u = 0; u < 8; u++) { const int e = e0 + threadIdx.x + u * P3_FTT; if (e < tot) tile[(e / N) * ld + e % N] = v[u]; }
    }
  } else {
    const int ncol = (N + P3_FT - 1) / P3_FT;
    const int c = blockIdx.x / ncol, i2a = (blockIdx.x % ncol) * P3_FT;
    nf = N - i2a < P3_FT ? N - i2a : P3_FT;
    g = x + (long)c * NN + i2a; fstr = 1; gstr = N;
    const int tot = nf * N;
    for (int e0 = 0; e0 < tot; e0 += 8 * P3_FTT) {   // consecutive threads: consecutive columns
      zd v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int e = e0 + threadIdx.x + u * P3_FTT, ec = e < tot ? e : tot - 1;
        v[u] = g[(long)(ec / nf) * N + ec % nf];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int e = e0 + threadIdx.x + u * P3_FTT, i = e / nf, f = e % nf;
        if (e < tot) tile[f * ld + i] = v[u];
      }
    }
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < nf) {
    zd *b = tile + t * ld;
    auto U = [&](int i, int j) { return uL[(size_t)j * (KD + 1) + KD + i - j]; };
    zd win[KD + 1];
    if (OP == 0) {
#pragma unroll
      for (int d = 0; d < KD; d++) win[d] = zmk(0.0, 0.0);
      // the last KD rows see the end of the band; the rest take all KD taps
      // with no bounds tests (same terms in the same order)
      int i = N - 1;
      for (; i >= 0 && i + KD >= N; i--) {
        zd s = b[i];
#pragma unroll
        for (int d = KD; d >= 1; d--)
          if (i + d < N) { const double u = U(i, i + d); s.x -= u * win[d - 1].x; s.y -= u * win[d - 1].y; }
        const zd xi = zscale(s, rd[i]);
        b[i] = xi;
#pragma unroll
        for (int d = KD - 1; d >= 1; d--) win[d] = win[d - 1];
        win[0] = xi;
      }
      // interior rows in groups of P3_SB: the group's right-hand sides are
      // loaded together, so the recurrence waits on LDS once per group
      for (; i >= P3_SB - 1; i -= P3_SB) {
        zd bv[P3_SB];
#pragma unroll
        for (int r = 0; r < P3_SB; r++) bv[r] = b[i - r];
#pragma unroll
        for (int r = 0; r < P3_SB; r++) {
          const int ii = i - r;
          zd s = bv[r];
#pragma unroll
          for (int d = KD; d >= 1; d--) { const double u = U(ii, ii + d); s.x -= u * win[d - 1].x; s.y -= u * win[d - 1].y; }
          const zd xi = zscale(s, rd[ii]);
          b[ii] = xi;
#pragma unroll
          for (int d = KD - 1; d >= 1; d--) win[d] = win[d - 1];
          win[0] = xi;
        }
      }
      for (; i >= 0; i--) {
        zd s = b[i];
#pragma unroll
        for (int d = KD; d >= 1; d--) { const double u = U(i, i + d); s.x -= u * win[d - 1].x; s.y -= u * win[d - 1].y; }
        const zd xi = zscale(s, rd[i]);
        b[i] = xi;
#pragma unroll
        for (int d = KD - 1; d >= 1; d--) win[d] = win[d - 1];
        win[0] = xi;
      }
    } else if (OP == 1) {
#pragma unroll
      for (int d = 0; d < KD; d++) win[d] = zmk(0.0, 0.0);
      int i = 0;
      for (; i < N && i < KD; i++) {   // the first KD rows see the start of the band
        zd s = b[i];
#pragma unroll
        for (int d = KD; d >= 1; d--)
          if (i - d >= 0) { const double u = U(i - d, i); s.x -= u * win[d - 1].x; s.y -= u * win[d - 1].y; }
        const zd xi = zscale(s, rd[i]);
        b[i] = xi;
#pragma unroll
        for (int d = KD - 1; d >= 1; d--) win[d] = win[d - 1];
        win[0] = xi;
      }
      for (; i + P3_SB <= N; i += P3_SB) {   // interior rows in groups, as above
        zd bv[P3_SB];
#pragma unroll
        for (int r = 0; r < P3_SB; r++) bv[r] = b[i + r];
#pragma unroll
        for (int r = 0; r < P3_SB; r++) {
          const int ii = i + r;
          zd s = bv[r];
#pragma unroll
          for (int d = KD; d >= 1; d--) { const double u = U(ii - d, ii); s.x -= u * win[d - 1].x; s.y -= u * win[d - 1].y; }
          const zd xi = zscale(s, rd[ii]);
          b[ii] = xi;
#pragma unroll
          for (int d = KD - 1; d >= 1; d--) win[d] = win[d - 1];
          win[0] = xi;
        }
      }
      for (; i < N; i++) {
        zd s = b[i];
#pragma unroll
        for (int d = KD; d >= 1; d--) { const double u = U(i - d, i); s.x -= u * win[d - 1].x; s.y -= u * win[d - 1].y; }
        const zd xi = zscale(s, rd[i]);
        b[i] = xi;
#pragma unroll
        for (int d = KD - 1; d >= 1; d--) win[d] = win[d - 1];
        win[0] = xi;
      }
    } else {
      for (int i = 0; i < N; i++) {
        zd s = zmk(0.0, 0.0);
#pragma unroll
        for (int d = 0; d <= KD; d++)
          if (i + d < N) { const double u = U(i, i + d); s.x += u * b[i + d].x; s.y += u * b[i + d].y; }
        b[i] = s;                                        // b[i] is not read again
      }
    }
  }
  __syncthreads();
  if (rank == 2) {
    for (int e = threadIdx.x; e < nf * N; e += P3_FTT) g[e] = tile[(e / N) * ld + e % N];
  } else {
    for (int e = threadIdx.x; e < nf * N; e += P3_FTT) {
      const int i = e / nf, f = e % nf;
      g[(long)i * N + f] = tile[f * ld + i];
    }
  }
  (void)fstr; (void)gstr;
}

// Dense triangular products for the overlap solves at N <= 128: y = T x
// per fiber with T = U^-1 (OP 0, upper triangular) or U^-T (OP 1, lower;
// U is real, so this is U^-H), padded with zeros to NP = 16 * NB rows and
// columns, on v_mfma_f64_16x16x4f64.  The substitution over a fiber is a
// chain of N dependent steps on one lane; the product has no chain.  T is
// kept as its nonzero 16 x 4 blocks (row block rb, column step ks), each 64
// doubles in MFMA operand order (lane i + 16k: T[16rb + i][4ks + k]); it
// stays in L2 and every wavefront streams the blocks it needs.  One
// wavefront per workgroup owns 16 real columns (fiber f, part ri at column
// 2f + ri) and computes all NP rows of them: rank 1 (fibers along i1,
// 8 consecutive i2 of one channel) D = T X with T as the A operand; rank 2
// (fibers along i2, 8 consecutive rows c*N + i1) D = X^T T^T with the same
// T blocks as the B operand.  The wave loads all of its columns first (so
// the product is in place), T one column step ahead of the MFMAs; terms are
// added in k order, four per MFMA, over the blocks inside the triangle.
#ifndef P3_TPROD
#define P3_TPROD 1         // overlap solves as products with the dense inverses where N <= 128
#endif
__host__ __device__ inline int p3_tblk_first(int OP, int rb) { return OP == 0 ? 4 * rb : 0; }
__host__ __device__ inline int p3_tblk_end(int OP, int rb, int KS) { return OP == 0 ? KS : 4 * rb + 4; }
__host__ __device__ inline int p3_tblk_base(int OP, int rb, int KS)   // blocks of row blocks < rb
{
  return OP == 0 ? rb * KS - 2 * rb * (rb - 1) : 2 * rb * (rb + 1);
}

template <int NB, int OP, int RANK>
__global__ void __launch_bounds__(64) p3_tgemm(const double *__restrict__ Tb, double *x, int C, int N)
{
  constexpr int NP = 16 * NB, KS = NP / 4;
  const int lane = threadIdx.x, li = lane & 15, lk = lane >> 4;
  double *g;
  int ncols;        // valid real columns of this wave (<= 16)
  long kstr;        // doubles from element k to k + 1 of a column
  if (RANK == 1) {
    const int nb = (2 * N + 15) / 16, c = blockIdx.x / nb, j = blockIdx.x - c * nb;
    g = x + (long)c * 2 * N * N + 16 * j;
    ncols = 2 * N - 16 * j < 16 ? 2 * N - 16 * j : 16;
    kstr = 2L * N;
  } else {
    const long f0 = (long)blockIdx.x * 8, nrow = (long)C * N;
    g = x + f0 * 2 * N;
    ncols = nrow - f0 < 8 ? 2 * (int)(nrow - f0) : 16;
    kstr = 2;
  }
  // the lane's operand column li (B[k][j] for rank 1, A[c][k] for rank 2; k = lk)
  const bool colok = li < ncols;
  const double *src = RANK == 1 ? g + li : g + (long)(li >> 1) * 2 * N + (li & 1);
  double xa[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ks++) {
    const int k = 4 * ks + lk;
    xa[ks] = (colok && k < N) ? src[(long)k * kstr] : 0.0;
  }
  auto tblk = [&](int rb, int ks) { return Tb[(p3_tblk_base(OP, rb, KS) + ks - p3_tblk_first(OP, rb)) * 64 + lane]; };
  auto inside = [](int rb, int ks) { return ks >= p3_tblk_first(OP, rb) && ks < p3_tblk_end(OP, rb, KS); };
  double tn[NB];
#pragma unroll
  for (int rb = 0; rb < NB; rb++) tn[rb] = inside(rb, 0) ? tblk(rb, 0) : 0.0;
  pp_d4 acc[NB];
#pragma unroll
  for (int rb = 0; rb < NB; rb++) acc[rb] = pp_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < KS; ks++) {
    double tc[NB];
#pragma unroll
    for (int rb = 0; rb < NB; rb++) {
      tc[rb] = tn[rb];
      if (ks + 1 < KS && inside(rb, ks + 1)) tn[rb] = tblk(rb, ks + 1);
    }
#pragma unroll
    for (int rb = 0; rb < NB; rb++) {
      if (!inside(rb, ks)) continue;
      if (RANK == 1) acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(tc[rb], xa[ks], acc[rb], 0, 0, 0);
      else acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[ks], tc[rb], acc[rb], 0, 0, 0);
    }
  }
  // D[4r + lane/16][lane%16] in register r of each block
#pragma unroll
  for (int rb = 0; rb < NB; rb++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (RANK == 1) {   // row i = 16rb + 4r + lk, column li
        const int i = 16 * rb + 4 * r + lk;
        if (i < N && colok) g[(long)i * 2 * N + li] = acc[rb][r];
      } else {           // column 4r + lk, row i = 16rb + li
        const int col = 4 * r + lk, i = 16 * rb + li;
        if (i < N && col < ncols) g[(long)(col >> 1) * 2 * N + 2 * i + (col & 1)] = acc[rb][r];
      }
    }
}

// Columns of U^-1 by back substitution (one thread per column j: U t = e_j,
// t_i = 0 for i > j), into T0 (row-major U^-1) and T1 (row-major U^-T);
// both NP x NP and zeroed beforehand.
__global__ void p3_tinv_build(const double *Ub, double *T0, double *T1, int N, int NP, int kd)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  auto U = [&](int i, int k) { return Ub[(size_t)k * (kd + 1) + kd + i - k]; };
  for (int i = j; i >= 0; i--) {
    double s = i == j ? 1.0 : 0.0;
    for (int d = 1; d <= kd && i + d <= j; d++) s -= U(i, i + d) * T0[(size_t)(i + d) * NP + j];
    const double t = s / U(i, i);
    T0[(size_t)i * NP + j] = t;
    T1[(size_t)j * NP + i] = t;
  }
}

// ------------------------------------------------------------------ Gram-Schmidt
// The thread's partial <V_u, w> (u < KN, V_u at vk + u*n) over i = lo+tid,
// lo+tid+256, ... < hi, accumulated in that order; the loads of P3_DUR
// consecutive elements are issued before their multiply-adds.
#ifndef P3_DUR
#define P3_DUR 4
#endif
template <int KN>
__device__ __attribute__((always_inline)) inline void p3_dots_group(const zd *vk, const zd *w, long n, long lo, long hi,
                                                                    double *re, double *im)
{
  for (long i = lo + threadIdx.x; i < hi; i += 256 * P3_DUR) {
    long ii[P3_DUR];
    zd x[P3_DUR], v[P3_DUR][KN];
#pragma unroll
    for (int r = 0; r < P3_DUR; r++) ii[r] = i + 256 * r < hi ? i + 256 * r : i;   // clamped: no branches on the loads
#pragma unroll
    for (int r = 0; r < P3_DUR; r++) {
      x[r] = w[ii[r]];
#pragma unroll
      for (int u = 0; u < KN; u++) v[r][u] = vk[(long)u * n + ii[r]];
    }
#pragma unroll
    for (int r = 0; r < P3_DUR; r++)
      if (r == 0 || i + 256 * r < hi)
#pragma unroll
        for (int u = 0; u < KN; u++) {
          re[u] = fma(v[r][u].x, x[r].x, re[u]); re[u] = fma(v[r][u].y, x[r].y, re[u]);
          im[u] = fma(v[r][u].x, x[r].y, im[u]); im[u] = fma(-v[r][u].y, x[r].x, im[u]);
        }
  }
}

// part[blk*pstr + 2k..] = partial <V_k, w> over this block's range, k < nk;
// nk == 0: part[blk*pstr] = partial |w|^2.
__global__ void __launch_bounds__(256) p3_dots(const zd *V, const zd *w, long n, int nk, double *part, int pstr,
                                                const int *state, int skip_if_broken)
{
  if (skip_if_broken && state[0]) return;
  __shared__ double red[8];
  const long per = (n + gridDim.x - 1) / gridDim.x;
  const long lo = blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (nk == 0) {
    double a = 0.0;
    for (long i = lo + tid; i < hi; i += 256) { const zd v = w[i]; a = fma(v.x, v.x, fma(v.y, v.y, a)); }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if (lane == 0) red[wv] = a;
    __syncthreads();
    if (tid == 0) part[(long)blockIdx.x * pstr] = (red[0] + red[1]) + (red[2] + red[3]);
    return;
  }
  // groups of 8 vectors: each thread reads w[i] once per group and carries
  // 8 partial inner products; one block reduction per group, summed over the
  // waves in a fixed order
  __shared__ double wred[4][16];
  for (int k0 = 0; k0 < nk; k0 += 8) {
    const int kn = nk - k0 < 8 ? nk - k0 : 8;
    double re[8], im[8];
#pragma unroll
    for (int u = 0; u < 8; u++) { re[u] = 0.0; im[u] = 0.0; }
    const zd *vk = V + (long)k0 * n;
    switch (kn) {
      case 1: p3_dots_group<1>(vk, w, n, lo, hi, re, im); break;
      case 2: p3_dots_group<2>(vk, w, n, lo, hi, re, im); break;
      case 3: p3_dots_group<3>(vk, w, n, lo, hi, re, im); break;
      case 4: p3_dots_group<4>(vk, w, n, lo, hi, re, im); break;
      case 5: p3_dots_group<5>(vk, w, n, lo, hi, re, im); break;
      case 6: p3_dots_group<6>(vk, w, n, lo, hi, re, im); break;
      case 7: p3_dots_group<7>(vk, w, n, lo, hi, re, im); break;
      default: p3_dots_group<8>(vk, w, n, lo, hi, re, im); break;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) { re[u] += __shfl_xor(re[u], o, 64); im[u] += __shfl_xor(im[u], o, 64); }
    }
    if (lane == 0)
#pragma unroll
      for (int u = 0; u < 8; u++) { wred[wv][2 * u] = re[u]; wred[wv][2 * u + 1] = im[u]; }
    __syncthreads();
    if (tid < 2 * kn)
      part[(long)blockIdx.x * pstr + 2 * k0 + tid] = (wred[0][tid] + wred[1][tid]) + (wred[2][tid] + wred[3][tid]);
    __syncthreads();
  }
}

// part[l*pstr + c] + part[(l+64)*pstr + c] + ... (blocks below nblk), added
// in that order, the loads of eight terms issued together
__device__ __attribute__((always_inline)) inline double p3_lane_sum(const double *part, int pstr, int nblk, int c,
                                                                    int lane)
{
  double a = 0.0;
  for (int b0 = lane; b0 < nblk; b0 += 64 * 8) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = part[(long)min(b0 + 64 * u, nblk - 1) * pstr + c];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (b0 + 64 * u < nblk) a += x[u];
  }
  return a;
}

// h_k = <V_k, w> for k <= j from the partials, written to hcoef and column j
// of Hm: summed in one fixed order, one wavefront per component, lane l
// summing blocks l, l+64, ... and a fixed shuffle tree.  (|w| and |psi~| are
// summed the same way by p3_scale, which uses them.)
__global__ void __launch_bounds__(1024) p3_reduce(const double *part, int pstr, int nblk, int j, int m, zd *hcoef,
                                                   zd *Hm, const int *state)
{
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (state[0]) return;
  const int ncomp = 2 * (j + 1);
  for (int c = wv; c < ncomp; c += 16) {
    double a = p3_lane_sum(part, pstr, nblk, c, lane);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if (lane == 0) ((double *)hcoef)[c] = a;   // h_k = (v[2k], v[2k+1])
  }
  __syncthreads();
  for (int k = tid; k <= j; k += 1024) Hm[k + j * m] = hcoef[k];
}

// w -= sum_{k<nk} V_k h_k;  partial |w|^2 per block
__global__ void __launch_bounds__(256) p3_project(const zd *V, zd *w, long n, int nk, const zd *hcoef, double *part,
                                                   int pstr, const int *state)
{
  if (state[0]) return;
  __shared__ zd h[64];
  __shared__ double red[4];
  const int tid = threadIdx.x;
  for (int k = tid; k < nk; k += 256) h[k] = hcoef[k];
  __syncthreads();
  const long per = (n + gridDim.x - 1) / gridDim.x;
  const long lo = blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  double a = 0.0;
  for (long i = lo + tid; i < hi; i += 256) {
    zd x = w[i];
#pragma unroll 4
    for (int k = 0; k < nk; k++) x = zsub(x, zmul(V[(long)k * n + i], h[k]));
    w[i] = x;
    a = fma(x.x, x.x, fma(x.y, x.y, a));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
  if ((tid & 63) == 0) red[tid >> 6] = a;
  __syncthreads();
  if (tid == 0) part[(long)blockIdx.x * pstr] = (red[0] + red[1]) + (red[2] + red[3]);
}

// dst1 = dst2 = src / b, b = sqrt of the sum of the nblk partials
// part[blk*pstr], summed by every block in p3_reduce's fixed order (lane l of
// wavefront 0 sums blocks l, l+64, ..., then the same shuffle tree), so that
// no reducing launch sits between the partials and their use.  mode 1: b =
// beta_bi with the breakdown test (pamp.h:305-312), recorded by block 0;
// mode 2: b = beta_0 = |psi~| at the start of a step (block 0 resets the
// breakdown state).
__global__ void __launch_bounds__(256) p3_scale(const zd *src, zd *dst1, zd *dst2, long n, const double *part, int pstr,
                                                 int nblk, int mode, int bi, int m, double *beta, zd *Hm, int *state)
{
  if (mode == 1 && state[0]) return;
  __shared__ double sb;
  if (threadIdx.x < 64) {
    double a = p3_lane_sum(part, pstr, nblk, 0, threadIdx.x);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if (threadIdx.x == 0) sb = sqrt(a);
  }
  __syncthreads();
  const double bt = sb;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (mode == 1) {
    const double tol = 1.4901161193847656e-08;   // sqrt(DBL_EPSILON), pamp.h:75
    if (fabs(bt) < tol) {
      if (lead) { state[0] = 1; state[1] = bi; }
      return;
    }
    if (lead) { beta[bi] = bt; Hm[bi + (bi - 1) * m] = zmk(bt, 0.0); }
  } else if (lead) {
    beta[0] = bt;
    state[0] = 0;
    state[1] = m;
  }
  const double inv = 1.0 / bt;
  for (long i = blockIdx.x * 256l + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const zd v = zscale(src[i], inv);
    dst1[i] = v;
    dst2[i] = v;
  }
}

__global__ void __launch_bounds__(256) p3_zero(zd *a, long cnt)
{
  for (long i = blockIdx.x * 256l + threadIdx.x; i < cnt; i += (long)gridDim.x * 256) a[i] = zmk(0.0, 0.0);
}

__global__ void __launch_bounds__(PP_THREADS) p3_pade(int m, zd *Hm, double dt_re, double dt_im, zd *cvec, double *diag,
                                                     const int *state, const double *beta, int normscale)
{
  pade_block(m, Hm, dt_re, dt_im, 0, cvec, diag, normscale);
  if (threadIdx.x == 0) {
    diag[1] = state[1];
    diag[2] = beta[0];
    diag[3] = 0.0;
  }
}

// out = beta_0 sum_{k < mEff} c_k V_k
__global__ void __launch_bounds__(256) p3_combine(const zd *V, const zd *cvec, const double *beta, const int *state,
                                                   zd *out, long n)
{
  __shared__ zd c[64];
  const int mEff = state[1];
  for (int k = threadIdx.x; k < mEff; k += 256) c[k] = zld(cvec + k);
  __syncthreads();
  const double b0 = beta[0];
  for (long i = blockIdx.x * 256l + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    zd acc = zmk(0.0, 0.0);
    for (int k = 0; k < mEff; k++) zfma(acc, V[(long)k * n + i], c[k]);
    out[i] = zscale(acc, b0);
  }
}

// ------------------------------------------------------------------ host side
static void p3_free_plan(pp_pamp3 *p)
{
  void *bufs[] = {p->dmats, p->rrow, p->rgrp, p->lrow, p->mrow, p->rent, p->lent, p->ment, p->cb_ptr, p->cb_blk,
                  p->blk_term, (void *)p->dR, p->X, p->Ym, p->nr, (void *)p->dRn, (void *)p->dPt, (void *)p->dQt,
                  (void *)p->fRnt, (void *)p->fPt, (void *)p->fQt, p->Tq, p->Tp};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (auto &t : p->mterms) {
    void *tb[] = {t.dRn, t.dP, t.dQ, t.fRn, t.fP, t.fQ};
    for (void *b : tb)
      if (b) (void)hipFree(b);
    t.dRn = t.dP = t.dQ = nullptr;
    t.fRn = t.fP = t.fQ = nullptr;
  }
  p->nr = nullptr; p->dRn = p->dPt = p->dQt = nullptr; p->Tq = p->Tp = nullptr; p->split = false;
  p->fRnt = p->fPt = p->fQt = nullptr;
  p->dmats = nullptr; p->rrow = p->lrow = p->mrow = nullptr; p->rgrp = nullptr; p->rent = p->lent = p->ment = nullptr;
  p->cb_ptr = p->cb_blk = p->blk_term = nullptr; p->dR = nullptr; p->X = nullptr; p->Ym = nullptr;
}

static void p3_collect(pp_pamp3 *p);

// rd[i] = 1 / U(i, i) of the overlap factor (the fiber solves' diagonal)
__global__ void p3_recip_diag(const double *Ub, double *rd, int N, int kd)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
    rd[i] = 1.0 / Ub[(size_t)i * (kd + 1) + kd];
}

// Dense U^-1 and U^-T on the device, then their nonzero 16 x 4 blocks in
// MFMA operand order (p3_tgemm), both triangles one after the other.
static int p3_tprod_build(pp_pamp3 *p)
{
  const int N = p->N, NP = p->tnp, NB = NP / 16, KS = NP / 4, nblk = 2 * NB * (NB + 1);
  const size_t dn = (size_t)NP * NP;
  double *dT = nullptr;
  if (hipMalloc(&dT, sizeof(double) * 2 * dn) != hipSuccess) return PP_EHIP;
  std::vector<double> T(2 * dn), B(2 * (size_t)nblk * 64, 0.0);
  bool ok = hipMemset(dT, 0, sizeof(double) * 2 * dn) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(p3_tinv_build, dim3((N + 63) / 64), dim3(64), 0, 0, p->ov->Uband, dT, dT + dn, N, NP, p->ov->kd);
    ok = hipGetLastError() == hipSuccess &&
         hipMemcpy(T.data(), dT, sizeof(double) * 2 * dn, hipMemcpyDeviceToHost) == hipSuccess;
  }
  (void)hipFree(dT);
  if (!ok) return PP_EHIP;
  // U itself, dense and upper (U(i, j) at Uband[j*(KD+1) + KD + i - j]), for x := U x
  const int KD = p->ov->kd;
  std::vector<double> Ub((size_t)N * (KD + 1));
  if (hipMemcpy(Ub.data(), p->ov->Uband, sizeof(double) * Ub.size(), hipMemcpyDeviceToHost) != hipSuccess)
    return PP_EHIP;
  T.resize(3 * dn, 0.0);
  for (int j = 0; j < N; j++)
    for (int i = j - KD > 0 ? j - KD : 0; i <= j; i++) T[2 * dn + (size_t)i * NP + j] = Ub[(size_t)j * (KD + 1) + KD + i - j];
  B.assign(3 * (size_t)nblk * 64, 0.0);
  for (int set = 0; set < 3; set++) {   // U^-1, U^-T, U: sets 0 and 2 upper, 1 lower
    const int op = set == 1 ? 1 : 0;
    for (int rb = 0; rb < NB; rb++)
      for (int ks = p3_tblk_first(op, rb); ks < p3_tblk_end(op, rb, KS); ks++) {
        double *blk = B.data() + ((size_t)set * nblk + p3_tblk_base(op, rb, KS) + ks - p3_tblk_first(op, rb)) * 64;
        for (int l = 0; l < 64; l++) blk[l] = T[(size_t)set * dn + (size_t)(16 * rb + (l & 15)) * NP + 4 * ks + (l >> 4)];
      }
  }
  if (hipMalloc(&p->tinv, sizeof(double) * B.size()) != hipSuccess) return PP_EHIP;
  return hipMemcpy(p->tinv, B.data(), sizeof(double) * B.size(), hipMemcpyHostToDevice) == hipSuccess ? PP_OK : PP_EHIP;
}

extern "C" int pp_pamp3_destroy(pp_pamp3 *p);

extern "C" int pp_pamp3_create(int C, int N, int m, int reorth, const pp_overlap *ov, pp_pamp3 **out)
{
  if (!out || !ov || C <= 0 || N <= 0 || m < 1 || m > 48) return PP_EARG;
  if (reorth) return PP_EUNSUP;
  if (ov->N != N || ov->kd != 6) return PP_EUNSUP;   // order-7 B-splines (13 diagonals)
  pp_pamp3 *p = new pp_pamp3();
  p->C = C; p->N = N; p->m = m; p->reorth = 0; p->s = ov->kd; p->bw = 2 * ov->kd + 1;
  p->n = (long)C * N * N; p->ov = ov;
  p->nslots = 0; p->dirty = true;
  p->dmats = nullptr; p->rrow = p->lrow = p->mrow = nullptr; p->rgrp = nullptr; p->rent = p->lent = p->ment = nullptr;
  p->cb_ptr = p->cb_blk = p->blk_term = nullptr; p->dR = nullptr; p->X = nullptr; p->Ym = nullptr;
  p->nr = nullptr; p->dRn = p->dPt = p->dQt = nullptr; p->Tq = p->Tp = nullptr; p->split = false;
  p->fRnt = p->fPt = p->fQt = nullptr; p->mixed = 0; p->expmethod = 0;
  p->K = 0; p->nblk = 0; p->nchunk = (C + P3_CHUNK - 1) / P3_CHUNK;
  p->scal = p->diag = nullptr; p->scal_cap = p->diag_cap = 0;
  p->timing = 0; p->ee_ms = 0.0; p->ee_count = 0;
  p->tinv = nullptr; p->tnp = 0;
  const size_t vb = sizeof(zd) * (size_t)p->n;
  bool ok = hipMalloc(&p->Y, vb) == hipSuccess && hipMalloc(&p->Z, vb) == hipSuccess &&
            hipMalloc(&p->P, vb) == hipSuccess && hipMalloc(&p->V, vb * (size_t)(m + 1)) == hipSuccess &&
            hipMalloc(&p->part, sizeof(double) * P3_RED_BLOCKS * 128) == hipSuccess &&
            hipMalloc(&p->Hm, sizeof(zd) * m * m) == hipSuccess && hipMalloc(&p->cvec, sizeof(zd) * m) == hipSuccess &&
            hipMalloc(&p->hcoef, sizeof(zd) * 64) == hipSuccess && hipMalloc(&p->beta, sizeof(double) * 64) == hipSuccess &&
            hipMalloc(&p->state, sizeof(int) * 4) == hipSuccess &&
            hipMalloc(&p->urd, sizeof(double) * (size_t)N) == hipSuccess;
  if (!ok) { pp_pamp3_destroy(p); return PP_EHIP; }
  hipLaunchKernelGGL(p3_recip_diag, dim3(1), dim3(256), 0, 0, ov->Uband, p->urd, N, ov->kd);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) { pp_pamp3_destroy(p); return PP_EHIP; }
  // U^-1, U^-T and U as MFMA blocks for the product form of the overlap solves and U x
  p->tnp = P3_TPROD ? (N <= 64 ? 64 : N <= 128 ? 128 : 0) : 0;
  if (p->tnp && p3_tprod_build(p) != PP_OK) { pp_pamp3_destroy(p); return PP_EHIP; }
  *out = p;
  return PP_OK;
}

extern "C" int pp_pamp3_destroy(pp_pamp3 *p)
{
  if (!p) return PP_OK;
  p3_collect(p);
  p3_free_plan(p);
  void *bufs[] = {p->Y, p->Z, p->P, p->V, p->part, p->Hm, p->cvec, p->hcoef, p->beta, p->state, p->scal, p->diag,
                  p->urd, p->tinv};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (auto &t : p->mterms)
    if (t.R) (void)hipFree(t.R);
  delete p;
  return PP_OK;
}

static int p3_mat_id(pp_pamp3 *p, const double *M)
{
  const size_t cnt = (size_t)p->N * p->bw;
  for (size_t i = 0; i < p->mats.size(); i++)
    if (std::memcmp(p->mats[i].data(), M, cnt * sizeof(double)) == 0) return (int)i;
  p->mats.emplace_back(M, M + cnt);
  return (int)p->mats.size() - 1;
}

static int p3_check_pairs(const pp_pamp3 *p, int P, const int *pairs, const double *coef, int slot)
{
  if (P < 0 || (P > 0 && (!pairs || !coef)) || slot < -1 || slot >= 32) return PP_EARG;
  for (int i = 0; i < 2 * P; i++)
    if (pairs[i] < 0 || pairs[i] >= p->C) return PP_EARG;
  return PP_OK;
}

extern "C" int pp_pamp3_set_radial_supports(pp_pamp3 *p, const int *lo, const int *hi)
{
  if (!p || !lo || !hi) return PP_EARG;
  for (int i = 0; i < p->N; i++)
    if (lo[i] > hi[i] || (i > 0 && (lo[i] < lo[i - 1] || hi[i] < hi[i - 1]))) return PP_EARG;
  p->sup_lo.assign(lo, lo + p->N);
  p->sup_hi.assign(hi, hi + p->N);
  p->dirty = true;
  return PP_OK;
}

extern "C" int pp_pamp3_add_kron_term(pp_pamp3 *p, int P, const int *pairs, const double *coef, const double *A,
                                      const double *B, int slot)
{
  if (!p || !A || !B) return PP_EARG;
  int rc = p3_check_pairs(p, P, pairs, coef, slot);
  if (rc) return rc;
  pp_pamp3::kterm t;
  t.pr.assign(pairs, pairs + 2 * P);
  t.cf.assign(coef, coef + 2 * P);
  t.A = p3_mat_id(p, A);
  t.B = p3_mat_id(p, B);
  t.slot = slot;
  p->kterms.push_back(t);
  if (slot + 1 > p->nslots) p->nslots = slot + 1;
  p->dirty = true;
  return PP_OK;
}

extern "C" int pp_pamp3_add_multipole_term(pp_pamp3 *p, int P, const int *pairs, const double *coef, const double *R,
                                           const double *Pband, const double *Qband, int slot)
{
  if (!p || !R || (!Pband) != (!Qband)) return PP_EARG;
  int rc = p3_check_pairs(p, P, pairs, coef, slot);
  if (rc) return rc;
  const size_t cnt = (size_t)p->N * p->bw * p->N * p->bw;
  pp_pamp3::mterm t;
  t.pr.assign(pairs, pairs + 2 * P);
  t.cf.assign(coef, coef + 2 * P);
  t.slot = slot;
  t.R = nullptr;
  t.dRn = t.dP = t.dQ = nullptr;
  t.fRn = t.fP = t.fQ = nullptr;
  if (Pband) {
    t.Rh.assign(R, R + cnt);
    t.Pl.assign(Pband, Pband + (size_t)p->N * p->bw);
    t.Ql.assign(Qband, Qband + (size_t)p->N * p->bw);
  }
  if (hipMalloc(&t.R, cnt * sizeof(double)) != hipSuccess) return PP_EHIP;
  if (hipMemcpy(t.R, R, cnt * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) { (void)hipFree(t.R); return PP_EHIP; }
  p->mterms.push_back(t);
  if (slot + 1 > p->nslots) p->nslots = slot + 1;
  p->dirty = true;
  return PP_OK;
}

template <class T>
static int p3_upload(T **dst, const std::vector<T> &v)
{
  const size_t bytes = sizeof(T) * (v.empty() ? 1 : v.size());
  if (hipMalloc(dst, bytes) != hipSuccess) return PP_EHIP;
  if (!v.empty() && hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess) return PP_EHIP;
  return PP_OK;
}

// Build the pass tables (see the file header): every Kronecker term goes to
// the side whose matrix it shares with more terms.
static int p3_build(pp_pamp3 *p)
{
  p3_free_plan(p);
  const int C = p->C;
  std::vector<int> cnt(p->mats.size(), 0);
  for (auto &t : p->kterms) { cnt[t.A]++; cnt[t.B]++; }
  // intermediates: ('L', B) plain i2 pass with matrix B, mixing in the i1 pass;
  //                ('R', A) mixing in the i2 pass, plain i1 pass with matrix A
  std::vector<std::pair<char, int>> inter;
  auto find = [&](char kind, int mat) {
    for (size_t i = 0; i < inter.size(); i++)
      if (inter[i].first == kind && inter[i].second == mat) return (int)i;
    inter.push_back({kind, mat});
    return (int)inter.size() - 1;
  };
  std::vector<std::vector<p3_ent>> rrows, lrows(C);
  for (auto &t : p->kterms) {
    const bool left_mix = cnt[t.B] >= cnt[t.A];
    const int k = find(left_mix ? 'L' : 'R', left_mix ? t.B : t.A);
    if ((int)rrows.size() < (k + 1) * C) rrows.resize((size_t)(k + 1) * C);
    const int P = (int)t.pr.size() / 2;
    for (int q = 0; q < P; q++) {
      const int row = t.pr[2 * q], col = t.pr[2 * q + 1];
      if (left_mix) lrows[row].push_back({k, col, t.A, t.slot, t.cf[2 * q], t.cf[2 * q + 1]});
      else rrows[(size_t)k * C + row].push_back({0, col, t.B, t.slot, t.cf[2 * q], t.cf[2 * q + 1]});
    }
  }
  p->K = (int)inter.size();
  rrows.resize((size_t)p->K * C);
  for (int k = 0; k < p->K; k++) {
    if (inter[k].first == 'L') {
      for (int c = 0; c < C; c++) rrows[(size_t)k * C + c].push_back({0, c, inter[k].second, -1, 1.0, 0.0});
    } else {
      for (int c = 0; c < C; c++) lrows[c].push_back({k, c, inter[k].second, -1, 1.0, 0.0});
    }
  }
  // Entries of one row with the same source, channel and time slot whose
  // coefficients are all real (or all imaginary) become one entry with the
  // summed band matrix (e.g. the kinetic, centrifugal and Coulomb bands of a
  // channel, or the 1/r and d/dr dipole bands of a channel pair).
  auto merge_row = [&](std::vector<p3_ent> &r) {
    std::vector<p3_ent> out;
    std::vector<char> used(r.size(), 0);
    const size_t cnt = (size_t)p->N * p->bw;
    for (size_t a = 0; a < r.size(); a++) {
      if (used[a]) continue;
      used[a] = 1;
      const bool re = r[a].cim == 0.0, im = r[a].cre == 0.0;
      std::vector<size_t> grp(1, a);
      for (size_t b = a + 1; b < r.size() && (re || im); b++) {
        if (used[b] || r[b].src != r[a].src || r[b].col != r[a].col || r[b].slot != r[a].slot) continue;
        if ((re && r[b].cim == 0.0) || (!re && r[b].cre == 0.0)) { grp.push_back(b); used[b] = 1; }
      }
      if (grp.size() == 1) { out.push_back(r[a]); continue; }
      std::vector<double> M(cnt, 0.0);
      for (size_t gi : grp) {
        const double f = re ? r[gi].cre : r[gi].cim;
        const double *src = p->mats[r[gi].mat].data();
        for (size_t i = 0; i < cnt; i++) M[i] += f * src[i];
      }
      p3_ent e = r[a];
      e.mat = p3_mat_id(p, M.data());
      e.cre = re ? 1.0 : 0.0;
      e.cim = re ? 0.0 : 1.0;
      out.push_back(e);
    }
    r.swap(out);
  };
  for (auto &r : lrows) merge_row(r);
  // i2-pass rows: entries grouped by (band matrix, time slot); within a
  // group the coefficients of one source channel are summed
  std::vector<int> rgp(1, 0);
  std::vector<p3_grp> rg;
  std::vector<p3_ent> rm;
  for (auto &r : rrows) {
    std::vector<char> used(r.size(), 0);
    for (size_t a = 0; a < r.size(); a++) {
      if (used[a]) continue;
      p3_grp G{(int)rm.size(), 0, r[a].mat, r[a].slot};
      for (size_t b = a; b < r.size(); b++) {
        if (used[b] || r[b].mat != r[a].mat || r[b].slot != r[a].slot) continue;
        used[b] = 1;
        bool found = false;
        for (int q = G.m0; q < (int)rm.size() && !found; q++)
          if (rm[q].col == r[b].col) { rm[q].cre += r[b].cre; rm[q].cim += r[b].cim; found = true; }
        if (!found) rm.push_back({0, r[b].col, r[b].mat, r[b].slot, r[b].cre, r[b].cim});
      }
      G.m1 = (int)rm.size();
      rg.push_back(G);
    }
    rgp.push_back((int)rg.size());
  }
  auto flatten = [](const std::vector<std::vector<p3_ent>> &rows, std::vector<int> &ptr, std::vector<p3_ent> &ent) {
    ptr.assign(1, 0);
    for (auto &r : rows) { ent.insert(ent.end(), r.begin(), r.end()); ptr.push_back((int)ent.size()); }
  };
  std::vector<int> lp, mp;
  std::vector<p3_ent> le, me;
  flatten(lrows, lp, le);
  // multipole blocks: (term, chunk) with any row in the chunk
  std::vector<std::vector<p3_ent>> mrows;
  std::vector<int> blk_term;
  std::vector<std::vector<int>> chunk_blocks(p->nchunk);
  for (size_t t = 0; t < p->mterms.size(); t++) {
    auto &mt = p->mterms[t];
    std::vector<std::vector<p3_ent>> rows(C);
    const int P = (int)mt.pr.size() / 2;
    for (int q = 0; q < P; q++)
      rows[mt.pr[2 * q]].push_back({0, mt.pr[2 * q + 1], 0, mt.slot, mt.cf[2 * q], mt.cf[2 * q + 1]});
    for (int ch = 0; ch < p->nchunk; ch++) {
      bool any = false;
      for (int cl = 0; cl < P3_CHUNK && ch * P3_CHUNK + cl < C; cl++) any = any || !rows[ch * P3_CHUNK + cl].empty();
      if (!any) continue;
      const int b = (int)blk_term.size();
      blk_term.push_back((int)t);
      chunk_blocks[ch].push_back(b);
      for (int cl = 0; cl < P3_CHUNK; cl++) {
        const int c = ch * P3_CHUNK + cl;
        mrows.push_back(c < C ? rows[c] : std::vector<p3_ent>());
      }
    }
  }
  p->nblk = (int)blk_term.size();
  flatten(mrows, mp, me);
  std::vector<int> cbp(1, 0), cbb;
  for (auto &v : chunk_blocks) { cbb.insert(cbb.end(), v.begin(), v.end()); cbp.push_back((int)cbb.size()); }
  // band rows are uploaded with the entries outside the matrix (j < 0 or
  // j >= N) set to 0: the i1 pass reads them against its zero-padded window
  // instead of testing every tap, whatever bands the caller handed over
  std::vector<double> mats;
  for (auto &mm : p->mats) {
    const size_t base = mats.size();
    mats.insert(mats.end(), mm.begin(), mm.end());
    for (int i = 0; i < p->N; i++)
      for (int d = 0; d < p->bw; d++) {
        const int j = i + d - p->s;
        if (j < 0 || j >= p->N) mats[base + (size_t)i * p->bw + d] = 0.0;
      }
  }
  std::vector<const double *> rptr;
  for (auto &mt : p->mterms) rptr.push_back(mt.R);
  int rc;
  if ((rc = p3_upload(&p->dmats, mats)) || (rc = p3_upload(&p->rrow, rgp)) || (rc = p3_upload(&p->rgrp, rg)) ||
      (rc = p3_upload(&p->rent, rm)) ||
      (rc = p3_upload(&p->lrow, lp)) || (rc = p3_upload(&p->lent, le)) || (rc = p3_upload(&p->mrow, mp)) ||
      (rc = p3_upload(&p->ment, me)) || (rc = p3_upload(&p->cb_ptr, cbp)) || (rc = p3_upload(&p->cb_blk, cbb)) ||
      (rc = p3_upload(&p->blk_term, blk_term)) || (rc = p3_upload(&p->dR, rptr)))
    return rc;
  const size_t vb = sizeof(zd) * (size_t)p->n;
  if (p->K > 0 && hipMalloc(&p->X, vb * (size_t)p->K) != hipSuccess) return PP_EHIP;
  if (p->nblk > 0 && hipMalloc(&p->Ym, sizeof(zd) * (size_t)p->nblk * p->N * p->N * P3_CHUNK) != hipSuccess)
    return PP_EHIP;
  // near/far split when every multipole term brings P_l, Q_l and the supports are known
  bool split = !p->mterms.empty() && (int)p->sup_lo.size() == p->N;
  for (auto &t : p->mterms) split = split && !t.Pl.empty();
  if (split) {
    const int N = p->N, bw = p->bw;
    std::vector<int2> nrh(N);
    for (int i1 = 0; i1 < N; i1++) {
      int lo = N, hi = -1;
      for (int i2 = 0; i2 < N; i2++) {
        const bool below = p->sup_hi[i1] < p->sup_lo[i2], above = p->sup_hi[i2] < p->sup_lo[i1];
        if (!below && !above) { lo = std::min(lo, i2); hi = std::max(hi, i2); }
      }
      if (hi - lo + 1 > P3_NW || hi < lo) return PP_EUNSUP;
      for (int i2 = lo; i2 <= hi; i2++)   // the near set must be contiguous
        if (p->sup_hi[i1] < p->sup_lo[i2] || p->sup_hi[i2] < p->sup_lo[i1]) return PP_EUNSUP;
      nrh[i1] = make_int2(lo, hi);
    }
    std::vector<const double *> rn, pp, qq;
    std::vector<const float *> frn, fpp, fqq;
    auto tof = [](const std::vector<double> &v) { return std::vector<float>(v.begin(), v.end()); };
    for (auto &t : p->mterms) {
      std::vector<double> Rn((size_t)N * bw * P3_NW * bw, 0.0);
      for (int i1 = 0; i1 < N; i1++)
        for (int d1 = 0; d1 < bw; d1++)
          for (int w = 0; w <= nrh[i1].y - nrh[i1].x; w++)
            for (int d2 = 0; d2 < bw; d2++)
              Rn[(((size_t)i1 * bw + d1) * P3_NW + w) * bw + d2] =
                  t.Rh[(((size_t)i1 * bw + d1) * N + nrh[i1].x + w) * bw + d2];
      if ((rc = p3_upload(&t.dRn, Rn)) || (rc = p3_upload(&t.dP, t.Pl)) || (rc = p3_upload(&t.dQ, t.Ql)) ||
          (rc = p3_upload(&t.fRn, tof(Rn))) || (rc = p3_upload(&t.fP, tof(t.Pl))) || (rc = p3_upload(&t.fQ, tof(t.Ql))))
        return rc;
      rn.push_back(t.dRn); pp.push_back(t.dP); qq.push_back(t.dQ);
      frn.push_back(t.fRn); fpp.push_back(t.fP); fqq.push_back(t.fQ);
    }
    for (int ch = 0; ch < p->nchunk; ch++)
      if (cbp[ch + 1] - cbp[ch] > P3_MAXB) return PP_EUNSUP;
    p->nrh = nrh;
    if ((rc = p3_upload(&p->nr, nrh)) || (rc = p3_upload(&p->dRn, rn)) || (rc = p3_upload(&p->dPt, pp)) ||
        (rc = p3_upload(&p->dQt, qq)) || (rc = p3_upload(&p->fRnt, frn)) || (rc = p3_upload(&p->fPt, fpp)) ||
        (rc = p3_upload(&p->fQt, fqq)))
      return rc;
    const size_t tb = sizeof(zd) * (size_t)p->nblk * N * N * P3_CHUNK;
    if (hipMalloc(&p->Tq, tb) != hipSuccess || hipMalloc(&p->Tp, tb) != hipSuccess) return PP_EHIP;
    p->split = true;
  }
  p->dirty = false;
  return PP_OK;
}

// z = H y (not balanced) with the time scalings at `scal` (device, nslots)
static int p3_apply_raw(pp_pamp3 *p, const zd *y, zd *z, const double *scal, hipStream_t st)
{
  const int C = p->C, N = p->N;
  if (p->K > 0) {
    const int ntile = (N + P3_RT - 1) / P3_RT, nch = (N + P3_RC - 1) / P3_RC;
    hipLaunchKernelGGL(p3_right<6>, dim3(p->K * C * ntile * nch), dim3(P3_RC), 0, st, p->dmats, p->rrow, p->rgrp,
                       p->rent, y, p->X, C, N, scal, p->state);
    const int ltile = (N + P3_LT - 1) / P3_LT;
    hipLaunchKernelGGL(p3_left<6>, dim3(C * ltile), dim3(128), 0, st, p->dmats, p->lrow, p->lent, p->X, z, C, N, scal,
                       p->state);
  } else {
    hipLaunchKernelGGL(p3_zero, dim3(1024), dim3(256), 0, st, z, p->n);
  }
  if (p->nblk > 0) {
    const int nhalf = (N + P3_HALF - 1) / P3_HALF;
    const int pairs = N * nhalf;
    const int groups = (pairs + 7) / 8;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (p->timing) {
      if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return PP_EHIP;
      (void)hipEventRecord(e0, st);
    }
    const int ntile = (N + P3_FRT - 1) / P3_FRT, ngrp = (N + 15) / 16;
    if (p->split && p->mixed) {
      // FP32 stage: the complex64 intermediates use the first half of the FP64 buffers
      float2 *Ym = (float2 *)p->Ym, *Tq = (float2 *)p->Tq, *Tp = (float2 *)p->Tp;
      const size_t lds = sizeof(float2) * (size_t)(N + 12) * P3_CHUNK;
      hipLaunchKernelGGL((p3_mixT<6, float>), dim3(p->nblk * N), dim3(256), lds, st, p->mrow, p->ment, y, Ym, p->fPt,
                         p->fQt, p->blk_term, p->nr, Tq, Tp, N, scal, p->state);
      hipLaunchKernelGGL((p3_near<6, float>), dim3(p->nchunk * N), dim3(64 * P3_NG), 0, st, p->fRnt, p->blk_term,
                         p->cb_ptr, p->cb_blk, p->nr, Ym, z, C, N, p->nchunk, p->state);
      hipLaunchKernelGGL((p3_farL<6, float>), dim3(p->nchunk * ntile * ngrp), dim3(256), 0, st, p->fPt, p->fQt,
                         p->blk_term, p->cb_ptr, p->cb_blk, p->nr, Tq, Tp, z, C, N, p->nchunk, p->state);
    } else if (p->split) {
      const size_t lds = sizeof(zd) * (size_t)(N + 12) * P3_CHUNK;
      hipLaunchKernelGGL((p3_mixT<6, double>), dim3(p->nblk * N), dim3(256), lds, st, p->mrow, p->ment, y, p->Ym,
                         p->dPt, p->dQt, p->blk_term, p->nr, p->Tq, p->Tp, N, scal, p->state);
      hipLaunchKernelGGL((p3_near<6, double>), dim3(p->nchunk * N), dim3(64 * P3_NG), 0, st, p->dRn, p->blk_term,
                         p->cb_ptr, p->cb_blk, p->nr, p->Ym, z, C, N, p->nchunk, p->state);
      hipLaunchKernelGGL((p3_farL<6, double>), dim3(p->nchunk * ntile * ngrp), dim3(256), 0, st, p->dPt, p->dQt,
                         p->blk_term, p->cb_ptr, p->cb_blk, p->nr, p->Tq, p->Tp, z, C, N, p->nchunk, p->state);
    } else {
      hipLaunchKernelGGL(p3_mix, dim3(p->nblk * N), dim3(256), 0, st, p->mrow, p->ment, y, p->Ym, N, scal, p->state);
      hipLaunchKernelGGL(p3_ee<6>, dim3(groups * 8 * p->nchunk), dim3(P3_EE_THREADS), 0, st, p->dR, p->blk_term,
                         p->cb_ptr, p->cb_blk, p->Ym, z, C, N, nhalf, p->nchunk, p->state);
    }
    if (p->timing) {
      (void)hipEventRecord(e1, st);
      p->ev.push_back(e0);
      p->ev.push_back(e1);
    }

  }
  return hipGetLastError() == hipSuccess ? PP_OK : PP_EHIP;
}

// op 0: (U x U)^-1, 1: (U x U)^-H, 2: (U x U) on both radial ranks.
template <int NB>
static int p3_tsolve2(pp_pamp3 *p, int op, zd *x, hipStream_t st)
{
  const int C = p->C, N = p->N, nblk = 2 * NB * (NB + 1);
  const double *T = p->tinv + (size_t)op * nblk * 64;   // op 0: U^-1, 1: U^-T, 2: U
  const dim3 g2((unsigned)(((long)C * N + 7) / 8)), g1((unsigned)(C * ((2 * N + 15) / 16)));
  double *xd = (double *)x;
  if (op != 1) {
    hipLaunchKernelGGL((p3_tgemm<NB, 0, 2>), g2, dim3(64), 0, st, T, xd, C, N);
    hipLaunchKernelGGL((p3_tgemm<NB, 0, 1>), g1, dim3(64), 0, st, T, xd, C, N);
  } else {
    hipLaunchKernelGGL((p3_tgemm<NB, 1, 2>), g2, dim3(64), 0, st, T, xd, C, N);
    hipLaunchKernelGGL((p3_tgemm<NB, 1, 1>), g1, dim3(64), 0, st, T, xd, C, N);
  }
  return hipGetLastError() == hipSuccess ? PP_OK : PP_EHIP;
}

static int p3_solve2(pp_pamp3 *p, int op, zd *x, hipStream_t st)
{
  if (p->tnp == 64) return p3_tsolve2<4>(p, op, x, st);
  if (p->tnp == 128) return p3_tsolve2<8>(p, op, x, st);
  const int C = p->C, N = p->N;
  const size_t lds = sizeof(double) * ((size_t)N * 7 + N + ((N * 7 + N) & 1)) + sizeof(zd) * (size_t)P3_FT * (N + 1);
  const dim3 g2((unsigned)(((long)C * N + P3_FT - 1) / P3_FT)), g1((unsigned)(C * ((N + P3_FT - 1) / P3_FT)));
  const void *fn[3] = {(const void *)p3_ftile<6, 0>, (const void *)p3_ftile<6, 1>, (const void *)p3_ftile<6, 2>};
  if (hipFuncSetAttribute(fn[op], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) return PP_EHIP;
  for (int rank = 2; rank >= 1; rank--) {
    const dim3 grid = rank == 2 ? g2 : g1;
    if (op == 0) hipLaunchKernelGGL((p3_ftile<6, 0>), grid, dim3(P3_FTT), lds, st, p->ov->Uband, p->urd, x, C, N, rank);
    else if (op == 1) hipLaunchKernelGGL((p3_ftile<6, 1>), grid, dim3(P3_FTT), lds, st, p->ov->Uband, p->urd, x, C, N, rank);
    else hipLaunchKernelGGL((p3_ftile<6, 2>), grid, dim3(P3_FTT), lds, st, p->ov->Uband, p->urd, x, C, N, rank);
  }
  return hipGetLastError() == hipSuccess ? PP_OK : PP_EHIP;
}

static int p3_ensure(pp_pamp3 *p, int nsteps, const double *scal)
{
  if (p->dirty) {
    int rc = p3_build(p);
    if (rc) return rc;
  }
  // the mixed form needs the near/far split (radial supports and P_l, Q_l of every multipole term)
  if (p->mixed && p->nblk > 0 && !p->split) return PP_EUNSUP;
  const int ns = p->nslots > 0 ? p->nslots : 1;
  if (nsteps * ns > p->scal_cap) {
    if (p->scal) (void)hipFree(p->scal);
    p->scal_cap = nsteps * ns;
    if (hipMalloc(&p->scal, sizeof(double) * p->scal_cap) != hipSuccess) { p->scal = nullptr; p->scal_cap = 0; return PP_EHIP; }
  }
  if (nsteps * 4 > p->diag_cap) {
    if (p->diag) (void)hipFree(p->diag);
    p->diag_cap = nsteps * 4;
    if (hipMalloc(&p->diag, sizeof(double) * p->diag_cap) != hipSuccess) { p->diag = nullptr; p->diag_cap = 0; return PP_EHIP; }
  }
  if (p->nslots > 0) {
    if (!scal) return PP_EARG;
    if (hipMemcpy(p->scal, scal, sizeof(double) * nsteps * p->nslots, hipMemcpyHostToDevice) != hipSuccess) return PP_EHIP;
  }
  return PP_OK;
}

extern "C" int pp_pamp3_apply(pp_pamp3 *p, const void *src, void *dst, const double *scal, int balanced,
                              hipStream_t stream)
{
  if (!p || !src || !dst) return PP_EARG;
  int rc = p3_ensure(p, 1, scal);
  if (rc) return rc;
  int zero_state[4] = {0, p->m, 0, 0};
  if (hipMemcpyAsync(p->state, zero_state, sizeof(zero_state), hipMemcpyHostToDevice, stream) != hipSuccess) return PP_EHIP;
  if (hipMemcpyAsync(p->Y, src, sizeof(zd) * (size_t)p->n, hipMemcpyDeviceToDevice, stream) != hipSuccess) return PP_EHIP;
  if (balanced && (rc = p3_solve2(p, 0, p->Y, stream))) return rc;
  if ((rc = p3_apply_raw(p, p->Y, (zd *)dst, p->scal, stream))) return rc;
  if (balanced && (rc = p3_solve2(p, 1, (zd *)dst, stream))) return rc;
  return hipStreamSynchronize(stream) == hipSuccess ? PP_OK : PP_EHIP;
}

extern "C" int pp_pamp3_propagate(pp_pamp3 *p, void *psi, int nsteps, double dt_re, double dt_im, const double *scal,
                                  hipStream_t stream)
{
  if (!p || !psi || nsteps < 0) return PP_EARG;
  if (nsteps == 0) return PP_OK;
  int rc = p3_ensure(p, nsteps, scal);
  if (rc) return rc;
  const long n = p->n;
  const int m = p->m, G = P3_RED_BLOCKS, pstr = 128;
  const size_t vb = sizeof(zd) * (size_t)n;
  zd *ps = (zd *)psi;
  const int gridv = 1024;
  const size_t lds = sizeof(double) * (size_t)pp_pade_lds_doubles(m) + 64 * sizeof(double);
  if (hipFuncSetAttribute((const void *)p3_pade, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    return PP_EHIP;
  for (int step = 0; step < nsteps; step++) {
    const double *sc = p->scal + (size_t)step * (p->nslots > 0 ? p->nslots : 1);
    hipLaunchKernelGGL(p3_zero, dim3(1), dim3(256), 0, stream, p->Hm, (long)m * m);
    // psi~ = (U x U) psi
    if (hipMemcpyAsync(p->P, ps, vb, hipMemcpyDeviceToDevice, stream) != hipSuccess) return PP_EHIP;
    if ((rc = p3_solve2(p, 2 /* U x */, p->P, stream))) return rc;
    hipLaunchKernelGGL(p3_dots, dim3(G), dim3(256), 0, stream, p->V, p->P, n, 0, p->part, pstr, p->state, 0);
    hipLaunchKernelGGL(p3_scale, dim3(gridv), dim3(256), 0, stream, p->P, p->V, p->Y, n, p->part, pstr, G, 2, 0, m,
                       p->beta, p->Hm, p->state);
    for (int j = 0; j < m; j++) {
      if ((rc = p3_solve2(p, 0, p->Y, stream))) return rc;
      if ((rc = p3_apply_raw(p, p->Y, p->Z, sc, stream))) return rc;
      if ((rc = p3_solve2(p, 1, p->Z, stream))) return rc;
      hipLaunchKernelGGL(p3_dots, dim3(G), dim3(256), 0, stream, p->V, p->Z, n, j + 1, p->part, pstr, p->state, 1);
      hipLaunchKernelGGL(p3_reduce, dim3(1), dim3(1024), 0, stream, p->part, pstr, G, j, m, p->hcoef, p->Hm,
                         p->state);
      if (j == m - 1) break;   // the last residual is not used (pamp.h:419-426)
      hipLaunchKernelGGL(p3_project, dim3(G), dim3(256), 0, stream, p->V, p->Z, n, j + 1, p->hcoef, p->part, pstr,
                         p->state);
      hipLaunchKernelGGL(p3_scale, dim3(gridv), dim3(256), 0, stream, p->Z, p->V + (size_t)(j + 1) * n, p->Y, n,
                         p->part, pstr, G, 1, j + 1, m, p->beta, p->Hm, p->state);
    }
    hipLaunchKernelGGL(p3_pade, dim3(1), dim3(PP_THREADS), lds, stream, m, p->Hm, dt_re, dt_im, p->cvec,
                       p->diag + (size_t)step * 4, p->state, p->beta, p->expmethod ? 2 : 0);
    hipLaunchKernelGGL(p3_combine, dim3(gridv), dim3(256), 0, stream, p->V, p->cvec, p->beta, p->state, ps, n);
    if ((rc = p3_solve2(p, 0, ps, stream))) return rc;
    if (hipGetLastError() != hipSuccess) return PP_EHIP;
  }
  return hipStreamSynchronize(stream) == hipSuccess ? PP_OK : PP_EHIP;
}

// Collect the event pairs recorded since the last call (the stream has been
// synchronised by the caller).
static void p3_collect(pp_pamp3 *p)
{
  for (size_t i = 0; i + 1 < p->ev.size(); i += 2) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p->ev[i], p->ev[i + 1]) == hipSuccess) {
      p->ee_ms += ms;
      p->ee_count++;
    }
    (void)hipEventDestroy(p->ev[i]);
    (void)hipEventDestroy(p->ev[i + 1]);
  }
  p->ev.clear();
}

// Exponential of the Krylov matrix (krylov_exponentiation_method): 0 Pade,
// 1 the eigenvector exponential, as pp_pamp_set_exponentiation.
extern "C" int pp_pamp3_set_exponentiation(pp_pamp3 *p, int method)
{
  if (!p || method < 0 || method > 1) return PP_EARG;
  p->expmethod = method;
  return PP_OK;
}

extern "C" int pp_pamp3_set_mixed(pp_pamp3 *p, int mixed)
{
  if (!p) return PP_EARG;
  p->mixed = mixed ? 1 : 0;
  return PP_OK;
}

extern "C" int pp_pamp3_set_timing(pp_pamp3 *p, int enable)
{
  if (!p) return PP_EARG;
  p3_collect(p);
  p->timing = enable != 0;
  p->ee_ms = 0.0;
  p->ee_count = 0;
  return PP_OK;
}

extern "C" int pp_pamp3_kernel_time(pp_pamp3 *p, double *ms_total, long *launches)
{
  if (!p || !ms_total || !launches) return PP_EARG;
  if (hipDeviceSynchronize() != hipSuccess) return PP_EHIP;
  p3_collect(p);
  *ms_total = p->ee_ms;
  *launches = p->ee_count;
  return PP_OK;
}

// Algorithmic flops of one multipole contraction launch: per term, every
// output channel with a coupling times every (i1, j1, i2, j2) inside both
// bands, 4 flops per real x complex multiply-add.
extern "C" int pp_pamp3_multipole_flops(pp_pamp3 *p, double *flops)
{
  if (!p || !flops) return PP_EARG;
  const long N = p->N, s = p->s;
  long bandpairs = 0;
  for (long i = 0; i < N; i++) bandpairs += std::min(N - 1, i + s) - std::max(0l, i - s) + 1;
  if (p->dirty) {
    int rc = p3_build(p);
    if (rc) return rc;
  }
  const double NN = (double)N * N;
  double f = 0.0;
  for (auto &t : p->mterms) {
    std::vector<char> rowset(p->C, 0);
    for (size_t q = 0; q < t.pr.size(); q += 2) rowset[t.pr[q]] = 1;
    long rows = 0;
    for (char r : rowset) rows += r;
    f += 8.0 * (double)(t.pr.size() / 2) * NN;   // channel mixing, complex x complex
    if (!p->split) {
      f += 4.0 * (double)rows * (double)bandpairs * (double)bandpairs;
      continue;
    }
    // near band: every (j1, j2) inside both bands; far field: two i2 passes
    // (Q and P) over every row and one i1 pass per far output
    double near = 0.0, far = 0.0;
    for (long i1 = 0; i1 < N; i1++) {
      const long t1 = std::min(N - 1, i1 + s) - std::max(0l, i1 - s) + 1;
      for (long i2 = p->nrh[i1].x; i2 <= p->nrh[i1].y; i2++)
        near += (double)t1 * (double)(std::min(N - 1, i2 + s) - std::max(0l, i2 - s) + 1);
      far += (double)(N - (p->nrh[i1].y - p->nrh[i1].x + 1)) * (double)t1;
    }
    f += 4.0 * (double)rows * (near + 2.0 * NN * (2 * s + 1) + far);
  }
  *flops = f;
  return PP_OK;
}

extern "C" int pp_pamp3_diagnostics(pp_pamp3 *p, int nsteps, double *diag, hipStream_t stream)
{
  if (!p || !diag || nsteps < 0 || nsteps * 4 > p->diag_cap) return PP_EARG;
  if (hipMemcpyAsync(diag, p->diag, sizeof(double) * 4 * nsteps, hipMemcpyDeviceToHost, stream) != hipSuccess)
    return PP_EHIP;
  return hipStreamSynchronize(stream) == hipSuccess ? PP_OK : PP_EHIP;
}
