// Shared device/host helpers for the pyprop_amd HIP library (gfx950).
// Complex amplitudes are stored interleaved (re, im) exactly like
// std::complex<double> / numpy complex128, so device buffers are
// layout-compatible with the reference's blitz::Array<cplx, R> data.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <vector>

#define PP_OK 0
#define PP_EARG (-1)       // bad argument (shape, storage id, null pointer)
#define PP_EUNSUP (-2)     // combination not implemented on this path
#define PP_EHIP (-3)       // HIP runtime error
#define PP_ETIMEOUT (-4)   // a grid barrier of a persistent launch timed out (its results are invalid)
#define PP_ENUM 1          // numerical failure (Cholesky info, singular Pade denominator)

#ifdef PP_DEBUG_ERRORS   // a diagnostic build: name the failing call
#include <cstdio>
#define PP_HIP(call)                                                                            \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      fprintf(stderr, "pyprop_amd: %s:%d: %s: %s\n", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return PP_EHIP;                                                                           \
    }                                                                                           \
  } while (0)
#else
#define PP_HIP(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return PP_EHIP;              \
  } while (0)
#endif

typedef double2 zd;

__host__ __device__ inline zd zmk(double a, double b) { return make_double2(a, b); }
__host__ __device__ inline zd zadd(zd a, zd b) { return zmk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline zd zsub(zd a, zd b) { return zmk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline zd zmul(zd a, zd b) { return zmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline zd zconj(zd a) { return zmk(a.x, -a.y); }
__host__ __device__ inline zd zscale(zd a, double s) { return zmk(a.x * s, a.y * s); }
// a += b * c
__host__ __device__ inline void zfma(zd &a, zd b, zd c)
{
  a.x = fma(b.x, c.x, a.x);
  a.x = fma(-b.y, c.y, a.x);
  a.y = fma(b.x, c.y, a.y);
  a.y = fma(b.y, c.x, a.y);
}
// a += conj(b) * c
__host__ __device__ inline void zfmac(zd &a, zd b, zd c)
{
  a.x = fma(b.x, c.x, a.x);
  a.x = fma(b.y, c.y, a.x);
  a.y = fma(b.x, c.y, a.y);
  a.y = fma(-b.y, c.x, a.y);
}
// a += s * c (s real)
__host__ __device__ inline void zfmar(zd &a, double s, zd c)
{
  a.x = fma(s, c.x, a.x);
  a.y = fma(s, c.y, a.y);
}
__host__ __device__ inline double zabs2(zd a) { return a.x * a.x + a.y * a.y; }

// Storage ids of the reference generator (tensorpotentialmultiply_generator.py:1336).
enum pp_storage { PP_IDENT = 0, PP_DIAG = 1, PP_SIMP = 2, PP_HERM = 3, PP_BANDNH = 4, PP_BAND = 5 };

// Chunked parallel solver for one real upper-banded triangular factor U
// (kd super-diagonals), for U x = b and U^T x = b (= U^H x = b, U real).
// Rows are padded to P*B (chunk length B a power of two >= kd, P <= 64
// chunks; padded rows are identity rows).
//   1. y_c = U_cc^-1 b_c for every chunk at once: a block-diagonal matvec
//      with the precomputed triangular inverse blocks (all threads; the
//      blocks are read-only and stay in L2).
//   2. The chunk couplings give t_c = y_c[0:kd] - M_c t_{c+1} (U) or
//      t_c = y_c[B-kd:B] - M'_c t_{c-1} (U^T), solved by a log-depth scan
//      on one wavefront (lane = chunk) whose matrix products are
//      precomputed (LDS).
//   3. x_c = y_c - G_c t_{c+1} (resp. G'_c t_{c-1}) for all rows at once.
struct pp_band_plan {
  int N, kd, B, lb, P, levels;
  int nlev;          // doubles of NlevN + NlevC (copied to LDS by the kernels)
  double *packed;    // device allocation holding the arrays below
  double *nlevs;     // [NlevN | NlevC], [((l*kd+i)*kd+k)*P + c]
  double *binv;      // [(c*B + j)*B + i] = (U_cc^-1)(i, j)
  double *binvT;     // [(c*B + j)*B + i] = (U_cc^-1)(j, i)
  double *G;         // [k*P*B + c*B + i] = (U_cc^-1 C_c)(i, k)
  double *GT;        // [k*P*B + c*B + i] = (U_cc^-T C'_c)(i, k)
};

// Overlap factor on the device: S = U^H U for one B-spline rank.
struct pp_overlap {
  int N, kd;
  double *Uband;       // device, [j*(kd+1) + kd+i-j] = U(i,j), real
  pp_band_plan plan;
};

struct pp_nostamp {
  __device__ void operator()() const {}
};

// Stage 1: out_c = U_cc^-1 in_c (UPPER) or U_cc^-T in_c for the nch chunks
// c0 <= c < c0+nch (all chunks by default); in/out hold those chunks' rows
// only (row c*B+i at (c-c0)*B+i).  8-column blocks that are structurally
// zero for a row are skipped (upper: j < i, lower: j > i).  The blocks are
// read-only and stream from L2.
template <bool UPPER, int NT>
__device__ inline void pp_solve_blocks(const pp_band_plan &pl, const zd *in, zd *out, int c0 = 0, int nch = -1)
{
  const int tid = threadIdx.x, B = pl.B, lb = pl.lb, nloc = (nch < 0 ? pl.P : nch) << lb;
  const double *Bm = UPPER ? pl.binv : pl.binvT;
  for (int row = tid; row < nloc; row += NT) {
    const int c = c0 + (row >> lb), i = row & (B - 1);
    const double *col = Bm + ((size_t)c << (2 * lb)) + i;
    const zd *bb = in + ((row >> lb) << lb);
    zd a0 = zmk(0.0, 0.0), a1 = zmk(0.0, 0.0);
    for (int j0 = 0; j0 < B; j0 += 8) {
      if (UPPER ? (j0 + 7 < i) : (j0 > i)) continue;
      const double *cj = col + ((size_t)j0 << lb);
#pragma unroll
      for (int jj = 0; jj < 8; jj += 2) {
        zfmar(a0, cj[(size_t)jj << lb], bb[j0 + jj]);
        zfmar(a1, cj[(size_t)(jj + 1) << lb], bb[j0 + jj + 1]);
      }
    }
    out[row] = zadd(a0, a1);
  }
}

// Stage 2 (wavefront 0 only): boundary unknowns of every chunk from the
// chunk couplings, t_c = y_c[0:kd] - M_c t_{c+1} (UPPER) or
// t_c = y_c[B-kd:B] - M'_c t_{c-1}, by a log-depth scan (lane = chunk).
// `src(c, k)` returns chunk c's boundary value k (y_c[k], resp. y_c[B-kd+k]);
// tb[(c-c0)*KD + k] receives the neighbour's t (t_{c+1} resp. t_{c-1}) for
// c0 <= c < c1.
template <int KD, bool UPPER, class Src>
__device__ inline void pp_solve_scan_src(const pp_band_plan &pl, const double *nl, Src src, zd *tb, int c0, int c1)
{
  const int P = pl.P;
  const int c = threadIdx.x;
  const bool act = c < P;
  const double *Nl = UPPER ? nl : nl + pl.levels * KD * KD * P;
  zd a[KD];
#pragma unroll
  for (int k = 0; k < KD; k++) a[k] = act ? src(c, k) : zmk(0.0, 0.0);
  for (int l = 0; l < pl.levels; l++) {
    const int s = 1 << l;
    zd an[KD];
#pragma unroll
    for (int k = 0; k < KD; k++) {
      if (UPPER) { an[k].x = __shfl_down(a[k].x, s, 64); an[k].y = __shfl_down(a[k].y, s, 64); }
      else { an[k].x = __shfl_up(a[k].x, s, 64); an[k].y = __shfl_up(a[k].y, s, 64); }
    }
    const bool has = UPPER ? (c + s < P) : (c - s >= 0);
    if (act && has) {
      const double *Nc = Nl + (size_t)l * KD * KD * P + c;
#pragma unroll
      for (int i = 0; i < KD; i++) {
        zd acc = a[i];
#pragma unroll
        for (int k = 0; k < KD; k++) zfmar(acc, Nc[(i * KD + k) * P], an[k]);
        a[i] = acc;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < KD; k++) {
    zd t;
    if (UPPER) { t.x = __shfl_down(a[k].x, 1, 64); t.y = __shfl_down(a[k].y, 1, 64); if (c + 1 >= P) t = zmk(0.0, 0.0); }
    else { t.x = __shfl_up(a[k].x, 1, 64); t.y = __shfl_up(a[k].y, 1, 64); if (c == 0) t = zmk(0.0, 0.0); }
    if (c >= c0 && c < c1) tb[(c - c0) * KD + k] = t;
  }
}

// Stage 2 over a whole fiber whose stage-1 output y (all P*B rows) is in LDS.
template <int KD, bool UPPER>
__device__ inline void pp_solve_scan(const pp_band_plan &pl, const double *nl, const zd *y, zd *tb)
{
  const int lb = pl.lb, off = UPPER ? 0 : pl.B - KD;
  pp_solve_scan_src<KD, UPPER>(pl, nl, [&](int c, int k) { return y[(c << lb) + off + k]; }, tb, 0, pl.P);
}

// Stage 3: x_c = y_c - G_c t_{c+1} (UPPER) or y_c - G'_c t_{c-1}, in place,
// for the nch chunks from c0 (all by default; out and tb hold those only).
template <int KD, bool UPPER, int NT>
__device__ inline void pp_solve_fixup(const pp_band_plan &pl, const zd *tb, zd *out, int c0 = 0, int nch = -1)
{
  const int npad = pl.P << pl.lb, lb = pl.lb, nloc = (nch < 0 ? pl.P : nch) << lb, r0 = c0 << lb;
  const double *Gm = UPPER ? pl.G : pl.GT;
  for (int row = threadIdx.x; row < nloc; row += NT) {
    const zd *t = tb + (row >> lb) * KD;
    zd v = out[row];
#pragma unroll
    for (int k = 0; k < KD; k++) zfmar(v, -Gm[(size_t)k * npad + r0 + row], t[k]);
    out[row] = v;
  }
}

// Solve U x = b (UPPER) or U^T x = b (!UPPER) for one fiber with the first
// NT threads of the workgroup (all threads must call it).  in: b (LDS, P*B, zero beyond N); out: x (LDS,
// P*B, distinct from in; zero beyond N on return); nl: LDS copy of
// pl.nlevs; tb: LDS scratch of P*KD complex.  Ends with a __syncthreads.
// `stamp` is called after stages 1 and 2 (profiling hook).
template <int KD, bool UPPER, int NT, class Stamp = pp_nostamp>
__device__ __attribute__((always_inline)) inline void pp_band_solve(const pp_band_plan &pl, const double *nl, const zd *in, zd *out, zd *tb,
                                     Stamp stamp = Stamp())
{
  if (threadIdx.x < NT) pp_solve_blocks<UPPER, NT>(pl, in, out);
  __syncthreads();
  stamp();
  if (threadIdx.x < 64) pp_solve_scan<KD, UPPER>(pl, nl, out, tb);
  __syncthreads();
  stamp();
  if (threadIdx.x < NT) pp_solve_fixup<KD, UPPER, NT>(pl, tb, out);
  __syncthreads();
}

// Chunk length for N rows: the smallest power of two B >= max(kd, 16) with
// at most 64 chunks whose scan matrices (both directions) fit in 72 KB of LDS.
inline int pp_plan_chunk(int N, int kd)
{
  for (int B = 16; B <= 256; B *= 2) {
    if (B < kd) continue;
    const int P = (N + B - 1) / B;
    int levels = 0;
    while ((1 << levels) < P) levels++;
    if (P <= 64 && 2.0 * levels * kd * kd * P * sizeof(double) <= 72.0 * 1024) return B;
  }
  return 0;
}

// Dispatch helper over the supported super-diagonal counts kd = order - 1:
// B-spline orders 3 to 9 (every order the reference's example configs use,
// and cubic splines).
#define PP_KD_MIN 2
#define PP_KD_MAX 8
#define PP_KD_SWITCH(kd, ...) \
  switch (kd) {                      \
    case 2: { constexpr int KD_ = 2; __VA_ARGS__; } break; \
    case 3: { constexpr int KD_ = 3; __VA_ARGS__; } break; \
    case 4: { constexpr int KD_ = 4; __VA_ARGS__; } break; \
    case 5: { constexpr int KD_ = 5; __VA_ARGS__; } break; \
    case 6: { constexpr int KD_ = 6; __VA_ARGS__; } break; \
    case 7: { constexpr int KD_ = 7; __VA_ARGS__; } break; \
    case 8: { constexpr int KD_ = 8; __VA_ARGS__; } break; \
    default: break; }

int pp_plan_build(pp_band_plan *pl, int N, int kd, const double *Uhost, int B);
void pp_plan_free(pp_band_plan *pl);

// Per-destination-row entry lists for one rank's storage id (host side).
struct pp_rank_csr {
  std::vector<int> rowptr;   // extent+1
  std::vector<int> ent;      // 3 ints per entry: source col, potential index, conjugate flag
};
// entry flags: conjugate V; the entry's own conjugation replaces the outer
// ranks' (innermost Band, zhbmv); only the real part of V (zhbmv's diagonal)
enum { PP_ENT_CONJ = 1, PP_ENT_OWN = 2, PP_ENT_REAL = 4 };
int pp_build_csr(int storage, int potExt, int srcExt, const int *pairs, pp_rank_csr &csr, bool innermost = false);
