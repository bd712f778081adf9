// Krylov-basis vector operations on the device for the eigenvalue solvers
// (PiramSolver): classical Gram-Schmidt of one vector against k basis
// vectors, repeated `passes` times, with deterministic reductions, and the
// recombination V Q of a basis by a small matrix (the implicit restart).
//
// Reference: the Arnoldi step and restart of pIRAM,
// core/krylov/piram/piram/piram.h:320-445 (PerformArnoldiStep, orthogonalisation
// with the 0.717 re-orthogonalisation rule) and :540-640 (PerformRestartStep),
// whose vector work runs through BLAS zgemv/zaxpy on the host there.
//
// Basis layout: V[j * n + i] (vector j contiguous), complex128 interleaved.
#include "pp_common.h"
#include <cstdio>
#include <mutex>
#include <vector>

#define PV_THREADS 256
#define PV_BLOCKS 240   // row blocks of the dot kernel

// partial[(c * nblk) + b] for c = 2j (re), 2j+1 (im): sum over rows of block b
// of conj(V_j[i]) w[i]; grid (nblk, k).
__global__ void __launch_bounds__(PV_THREADS) pv_dots(const zd *V, long n, const zd *w, double *partial)
{
  const int j = blockIdx.y, nblk = gridDim.x;
  const zd *v = V + (long)j * n;
  zd acc = zmk(0.0, 0.0);
  for (long i = (long)blockIdx.x * PV_THREADS + threadIdx.x; i < n; i += (long)nblk * PV_THREADS)
    zfmac(acc, v[i], w[i]);
  __shared__ double sr[PV_THREADS / 64], si[PV_THREADS / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    acc.x += __shfl_xor(acc.x, o, 64);
    acc.y += __shfl_xor(acc.y, o, 64);
  }
  if ((threadIdx.x & 63) == 0) { sr[threadIdx.x >> 6] = acc.x; si[threadIdx.x >> 6] = acc.y; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double re = 0.0, im = 0.0;
    for (int q = 0; q < PV_THREADS / 64; q++) { re += sr[q]; im += si[q]; }
    partial[(long)(2 * j) * nblk + blockIdx.x] = re;
    partial[(long)(2 * j + 1) * nblk + blockIdx.x] = im;
  }
}

// s_c = sum_b partial[c * nblk + b] in block order, for c < ncomp:
// cur[c] = s_c, and acc[c] = s_c (first pass) or acc[c] + s_c (acc may be null)
__global__ void __launch_bounds__(PV_THREADS) pv_reduce(const double *partial, int nblk, int ncomp, double *cur,
                                                        double *acc, int accumulate)
{
  for (int c = threadIdx.x; c < ncomp; c += PV_THREADS) {
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += partial[(long)c * nblk + b];
    cur[c] = s;
    if (acc) acc[c] = accumulate ? acc[c] + s : s;
  }
}

// w[i] -= sum_j V_j[i] h_j, with h_j = (hp[2j], hp[2j+1])
__global__ void __launch_bounds__(PV_THREADS) pv_project(const zd *V, long n, int k, const double *hp, zd *w)
{
  for (long i = (long)blockIdx.x * PV_THREADS + threadIdx.x; i < n; i += (long)gridDim.x * PV_THREADS) {
    zd acc = w[i];
    for (int j = 0; j < k; j++) acc = zsub(acc, zmul(V[(long)j * n + i], zmk(hp[2 * j], hp[2 * j + 1])));
    w[i] = acc;
  }
}

// out[p][i] = sum_j V[j][i] Q[j + p * m]   (Q column-major m x np, complex)
__global__ void __launch_bounds__(PV_THREADS) pv_combine(const zd *V, long n, int m, const zd *Q, int np, zd *out)
{
  const int p = blockIdx.y;
  for (long i = (long)blockIdx.x * PV_THREADS + threadIdx.x; i < n; i += (long)gridDim.x * PV_THREADS) {
    zd acc = zmk(0.0, 0.0);
    for (int j = 0; j < m; j++) zfma(acc, V[(long)j * n + i], Q[j + (long)p * m]);
    out[(long)p * n + i] = acc;
  }
}

static unsigned pv_grid(long n)
{
  const long g = (n + PV_THREADS - 1) / PV_THREADS;
  return (unsigned)(g < 1024 ? (g > 0 ? g : 1) : 1024);
}

// Device scratch of one Gram-Schmidt call: the dot kernel's partial sums,
// 2 max(k, 1) rows of PV_BLOCKS doubles (the norm pass uses rows 0 and 1 even
// when k = 0), then [h accumulated over passes | h of this pass | |w|^2].
static size_t pv_gs_partial_doubles(int k) { return 2 * (size_t)(k > 0 ? k : 1) * PV_BLOCKS; }
static size_t pv_gs_scalar_doubles(int k) { return 4 * (size_t)k + 1; }

static bool pv_gs_args_ok(const void *V, long n, int k, const void *w, int passes, const double *h,
                          const double *wnorm)
{
  return !(n <= 0 || k < 0 || k > 4096 || (k > 0 && (!V || !h)) || !w || !wnorm || passes < 1 || passes > 3);
}

// The launches and the read-back of one Gram-Schmidt call on caller-given
// device scratch: part (pv_gs_partial_doubles(k)) and hd
// (pv_gs_scalar_doubles(k)).  Every scratch value is written before it is
// read, so the scratch needs no clearing.  Synchronises the stream.
static int pv_gram_schmidt_on(const zd *Vz, long n, int k, zd *wz, int passes, double *h, double *wnorm,
                              double *part, double *hd, hipStream_t stream)
{
  const int nblk = PV_BLOCKS;
  const size_t nc = 2 * (size_t)k;
  double *hacc = hd, *hcur = hd + nc, *nrm = hd + 2 * nc;
  for (int ps = 0; k > 0 && ps < passes; ps++) {
    hipLaunchKernelGGL(pv_dots, dim3(nblk, k), dim3(PV_THREADS), 0, stream, Vz, n, (const zd *)wz, part);
    hipLaunchKernelGGL(pv_reduce, dim3(1), dim3(PV_THREADS), 0, stream, (const double *)part, nblk, (int)nc, hcur,
                       hacc, ps > 0 ? 1 : 0);
    hipLaunchKernelGGL(pv_project, dim3(pv_grid(n)), dim3(PV_THREADS), 0, stream, Vz, n, k, (const double *)hcur, wz);
  }
  hipLaunchKernelGGL(pv_dots, dim3(nblk, 1), dim3(PV_THREADS), 0, stream, (const zd *)wz, n, (const zd *)wz, part);
  hipLaunchKernelGGL(pv_reduce, dim3(1), dim3(PV_THREADS), 0, stream, (const double *)part, nblk, 1, nrm,
                     (double *)nullptr, 0);
  PP_HIP(hipGetLastError());
  double nn = 0.0;
  if (nc > 0) PP_HIP(hipMemcpyAsync(h, hacc, sizeof(double) * nc, hipMemcpyDeviceToHost, stream));
  PP_HIP(hipMemcpyAsync(&nn, nrm, sizeof(double), hipMemcpyDeviceToHost, stream));
  PP_HIP(hipStreamSynchronize(stream));
  *wnorm = sqrt(nn > 0.0 ? nn : 0.0);
  return PP_OK;
}

// Classical Gram-Schmidt of w (n) against V_0 .. V_{k-1}, `passes` times
// (1 or 2): h = sum over passes of V^H w, w -= V h each pass; then |w|.
// h (host, 2k doubles re/im) and wnorm (host) are returned; synchronises the
// stream.  k = 0 computes only the norm.
extern "C" int pp_gram_schmidt(const void *V, long n, int k, void *w, int passes, double *h, double *wnorm,
                               hipStream_t stream)
{
  if (!pv_gs_args_ok(V, n, k, w, passes, h, wnorm)) return PP_EARG;
  double *part = nullptr, *hd = nullptr;
  PP_HIP(hipMallocAsync((void **)&part, sizeof(double) * pv_gs_partial_doubles(k), stream));
  PP_HIP(hipMallocAsync((void **)&hd, sizeof(double) * pv_gs_scalar_doubles(k), stream));
  const int rc = pv_gram_schmidt_on((const zd *)V, n, k, (zd *)w, passes, h, wnorm, part, hd, stream);
  const hipError_t e1 = hipFreeAsync(part, stream), e2 = hipFreeAsync(hd, stream);
  if (rc != PP_OK) return rc;
  if (e1 != hipSuccess || e2 != hipSuccess) return PP_EHIP;
  PP_HIP(hipStreamSynchronize(stream));
  return PP_OK;
}

// out (host, 2 doubles) = sum_i conj(a_i) b_i: the zdotc after the overlap
// multiply of CombinedRepresentation::InnerProduct (combinedrepresentation.cpp:
// 423-474), as 240 row-block partial sums added in block order (the same
// kernels as the Gram-Schmidt norms, so the result is reproducible).
// Synchronises the stream.
extern "C" int pp_inner_product(const void *a, const void *b, long n, double *out, hipStream_t stream)
{
  if (!a || !b || !out || n <= 0) return PP_EARG;
  double *part = nullptr;
  PP_HIP(hipMallocAsync((void **)&part, sizeof(double) * (2 * PV_BLOCKS + 2), stream));
  hipLaunchKernelGGL(pv_dots, dim3(PV_BLOCKS, 1), dim3(PV_THREADS), 0, stream, (const zd *)a, n, (const zd *)b, part);
  hipLaunchKernelGGL(pv_reduce, dim3(1), dim3(PV_THREADS), 0, stream, (const double *)part, PV_BLOCKS, 2,
                     part + 2 * PV_BLOCKS, (double *)nullptr, 0);
  const hipError_t le = hipGetLastError();
  const hipError_t ce = le == hipSuccess ? hipMemcpyAsync(out, part + 2 * PV_BLOCKS, 2 * sizeof(double),
                                                          hipMemcpyDeviceToHost, stream) : le;
  const hipError_t fe = hipFreeAsync(part, stream);
  if (ce != hipSuccess || fe != hipSuccess) return PP_EHIP;
  PP_HIP(hipStreamSynchronize(stream));
  return PP_OK;
}

// ---- the same call without allocation ----
// One file-scope device array per GPU holds the scratch of
// pp_gram_schmidt_noalloc: the partial sums of PV_GS_NOALLOC_MAXK basis
// vectors and the scalars after them.  Host code never passes the array's
// name to a kernel (on the host it is only a stand-in for the symbol): the
// device address is looked up once per device with hipGetSymbolAddress and
// kept in pv_gs_scratch_ptr.
#define PV_GS_NOALLOC_MAXK 128
#define PV_GS_NOALLOC_DOUBLES (2 * PV_GS_NOALLOC_MAXK * PV_BLOCKS + 4 * PV_GS_NOALLOC_MAXK + 1)
__device__ __align__(256) double scratch[PV_GS_NOALLOC_DOUBLES];

#define PV_MAX_DEVICES 64
static double *pv_gs_scratch_ptr[PV_MAX_DEVICES];
static std::mutex pv_gs_scratch_mutex;

static int pv_gs_scratch(double **out)
{
  int dev = 0;
  PP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= PV_MAX_DEVICES) return PP_EUNSUP;
  std::lock_guard<std::mutex> lock(pv_gs_scratch_mutex);
  if (!pv_gs_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return PP_EHIP;
    pv_gs_scratch_ptr[dev] = (double *)ptr;
  }
  *out = pv_gs_scratch_ptr[dev];
  return PP_OK;
}

// pp_gram_schmidt with its scratch in the device array above instead of
// stream-ordered allocations: k <= PV_GS_NOALLOC_MAXK (PP_EARG beyond).
// Calls on one GPU share that GPU's array, so calls that overlap on one GPU
// (other streams or threads) race; calls on one stream are ordered.
extern "C" int pp_gram_schmidt_noalloc(const void *V, long n, int k, void *w, int passes, double *h,
                                       double *wnorm, hipStream_t stream)
{
  if (!pv_gs_args_ok(V, n, k, w, passes, h, wnorm) || k > PV_GS_NOALLOC_MAXK) return PP_EARG;
  double *base = nullptr;
  const int rc = pv_gs_scratch(&base);
  if (rc != PP_OK) return rc;
  return pv_gram_schmidt_on((const zd *)V, n, k, (zd *)w, passes, h, wnorm, base,
                            base + pv_gs_partial_doubles(PV_GS_NOALLOC_MAXK), stream);
}

// out[p] = sum_j V_j Q[j, p] for p < np (Q host, column-major m x np,
// complex interleaved); out must not overlap V.  Asynchronous on the stream.
extern "C" int pp_basis_combine(const void *V, long n, int m, const double *Q, int np, void *out, hipStream_t stream)
{
  if (!V || !Q || !out || n <= 0 || m <= 0 || np <= 0 || np > 65535) return PP_EARG;
  zd *Qd = nullptr;
  PP_HIP(hipMallocAsync((void **)&Qd, sizeof(zd) * (size_t)m * np, stream));
  PP_HIP(hipMemcpyAsync(Qd, Q, sizeof(zd) * (size_t)m * np, hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(pv_combine, dim3(pv_grid(n), np), dim3(PV_THREADS), 0, stream, (const zd *)V, n, m, Qd, np,
                     (zd *)out);
  PP_HIP(hipGetLastError());
  PP_HIP(hipFreeAsync(Qd, stream));
  return PP_OK;
}

// ------------------------------------------------------------------ banded LU solve per fiber
// x_a := A_a^-1 x_a for fibers a < nfib, with A_a = P L U factored by LAPACK
// zgbtrf (band storage, column-major, ldab = 2 kl + ku + 1, U's diagonal in
// row kl + ku, multipliers below it; piv 0-based).  The substitutions follow
// zgbtrs 'N': row interchanges with the forward elimination, then the upper
// band solve by columns.  One thread per fiber (the recurrences are
// sequential); used by the shift-invert preconditioner, not on the
// propagation path.
__global__ void __launch_bounds__(64) pv_gbtrs(int nfib, int N, int kl, int ku, const zd *lub, const int *piv, zd *x)
{
  const int a = blockIdx.x * 64 + threadIdx.x;
  if (a >= nfib) return;
  const int ldab = 2 * kl + ku + 1, kd = kl + ku;
  const zd *A = lub + (long)a * ldab * N;
  const int *p = piv + (long)a * N;
  zd *b = x + (long)a * N;
  for (int j = 0; j < N - 1; j++) {
    const int lm = kl < N - 1 - j ? kl : N - 1 - j;
    const int l = p[j];
    zd bj = b[l];
    if (l != j) { b[l] = b[j]; b[j] = bj; }
    for (int i = 1; i <= lm; i++) b[j + i] = zsub(b[j + i], zmul(A[(long)j * ldab + kd + i], bj));
  }
  for (int j = N - 1; j >= 0; j--) {
    const zd d = A[(long)j * ldab + kd];
    const double dn = zabs2(d);
    const zd bj = zmul(b[j], zmk(d.x / dn, -d.y / dn));
    b[j] = bj;
    const int i0 = j - kd > 0 ? j - kd : 0;
    for (int i = i0; i < j; i++) b[i] = zsub(b[i], zmul(A[(long)j * ldab + kd + i - j], bj));
  }
}

// The same substitutions with one wavefront per fiber: the fiber's vector
// lives in LDS and the band factor is staged through LDS 64 rows at a time
// (coalesced loads), so each step of the recurrence waits on LDS rather than
// on a global-memory round trip.  Per step the pivot swap and the kl (forward)
// or kd (backward) updates run on separate lanes.
#define PV_GB_ROWS 64
__global__ void __launch_bounds__(64) pv_gbtrs_lds(int N, int kl, int ku, const zd *lub, const int *piv, zd *x)
{
  extern __shared__ zd pv_lds[];
  const int a = blockIdx.x, lane = threadIdx.x;
  const int ldab = 2 * kl + ku + 1, kd = kl + ku;
  zd *bs = pv_lds;                              // N
  zd *As = bs + N;                              // PV_GB_ROWS columns of ldab
  int *ps = (int *)(As + PV_GB_ROWS * ldab);    // PV_GB_ROWS pivots
  const zd *A = lub + (long)a * ldab * N;
  const int *p = piv + (long)a * N;
  zd *b = x + (long)a * N;
  for (int i = lane; i < N; i += 64) bs[i] = b[i];
  for (int j0 = 0; j0 < N - 1; j0 += PV_GB_ROWS) {
    const int nr = min(PV_GB_ROWS, N - 1 - j0);
    __syncthreads();
    for (int e = lane; e < nr * ldab; e += 64) As[e] = A[(long)j0 * ldab + e];
    for (int e = lane; e < nr; e += 64) ps[e] = p[j0 + e];
    __syncthreads();
    for (int r = 0; r < nr; r++) {
      const int j = j0 + r;
      const int lm = min(kl, N - 1 - j);
      const int l = ps[r];
      const zd bl = bs[l], bj = bs[j];
      zd v = zmk(0.0, 0.0);
      if (lane >= 1 && lane <= lm) v = (j + lane == l) ? bj : bs[j + lane];
      __syncthreads();
      if (lane == 0) bs[j] = bl;
      else if (lane <= lm) bs[j + lane] = zsub(v, zmul(As[r * ldab + kd + lane], bl));
      __syncthreads();
    }
  }
  for (int jh = N - 1; jh >= 0; jh -= PV_GB_ROWS) {
    const int jl = jh - PV_GB_ROWS + 1 > 0 ? jh - PV_GB_ROWS + 1 : 0;
    __syncthreads();
    for (int e = lane; e < (jh - jl + 1) * ldab; e += 64) As[e] = A[(long)jl * ldab + e];
    __syncthreads();
    for (int j = jh; j >= jl; j--) {
      const zd *col = As + (j - jl) * ldab;
      const zd d = col[kd];
      const double dn = zabs2(d);
      const zd bj = zmul(bs[j], zmk(d.x / dn, -d.y / dn));
      const int i = j - lane;
      zd v = zmk(0.0, 0.0);
      if (lane >= 1 && lane <= kd && i >= 0) v = bs[i];
      __syncthreads();
      if (lane == 0) bs[j] = bj;
      else if (lane <= kd && i >= 0) bs[i] = zsub(v, zmul(col[kd - lane], bj));
      __syncthreads();
    }
  }
  __syncthreads();
  for (int i = lane; i < N; i += 64) b[i] = bs[i];
}

// Solve A_a x_a = b_a in place for nfib fibers of length N (x: nfib x N
// complex128, device); lub: nfib band LU factors (device, nfib x N x ldab
// complex128), piv: nfib x N 0-based pivots (device).  Asynchronous.
extern "C" int pp_band_lu_solve(int nfib, int N, int kl, int ku, const void *lub, const int *piv, void *x,
                                hipStream_t stream)
{
  if (nfib <= 0 || N <= 0 || kl < 0 || ku < 0 || !lub || !piv || !x) return PP_EARG;
  const int ldab = 2 * kl + ku + 1;
  const size_t lds = sizeof(zd) * ((size_t)N + (size_t)PV_GB_ROWS * ldab) + sizeof(int) * PV_GB_ROWS;
  if (kl + ku < 64 && lds <= 64 * 1024)
    hipLaunchKernelGGL(pv_gbtrs_lds, dim3(nfib), dim3(64), lds, stream, N, kl, ku, (const zd *)lub, piv, (zd *)x);
  else
    hipLaunchKernelGGL(pv_gbtrs, dim3((nfib + 63) / 64), dim3(64), 0, stream, nfib, N, kl, ku, (const zd *)lub, piv,
                       (zd *)x);
  return hipGetLastError() == hipSuccess ? PP_OK : PP_EHIP;
}
