// Tensor-potential multiply: dest += scaling * (x_r V_r) src, the contract of
// the generated TensorPotentialMultiply_<ids> routines
// (core/tensorpotential/tensorpotentialmultiply_generator.py:302-763,
// wrapper signature :1428-1521).  dest is accumulated, scaling is real, the
// potential is stored with one compressed index per rank in the reference
// layout (C order, largest stride first).
//
// Each rank's storage id is expanded on the host into a per-destination-row
// list of (source col, potential index, conjugate) entries; one device
// thread then owns one destination element and walks the product of its
// rows' lists.  That keeps the kernel free of atomics and deterministic for
// every storage combination the generator emits on a single process.
#include "pp_common.h"
#include <cstring>


// Expand one rank's storage into destination-row lists of (source col,
// potential index, flags).  Ident :436-462, Diag :403-433, Simp :302-363,
// Herm :370-398, BandNH :622-763 (general band, pidx = col*bandCount +
// s+row-col), Band :467-618: slot j*k + d (d = 0..k-1) belongs to the pair
// (j, j+d); on the innermost rank the generator calls zhbmv 'L' (:562-575),
// so A[j+d][j] = V, A[j][j+d] = conj V, the diagonal's real part only and
// no conjugation from an outer Herm mirror (PP_ENT_OWN); on an outer rank
// its loops (:513-551) apply A[j][j+d] = V and A[j+d][j] = conj V with the
// complex diagonal.
int pp_build_csr(int storage, int potExt, int srcExt, const int *pairs, pp_rank_csr &csr, bool innermost)
{
  std::vector<std::vector<int>> rows((size_t)srcExt);
  auto add = [&](int row, int col, int pidx, int flags) -> bool {
    if (row < 0 || row >= srcExt || col < 0 || col >= srcExt) return false;
    rows[row].push_back(col); rows[row].push_back(pidx); rows[row].push_back(flags);
    return true;
  };
  switch (storage) {
    case PP_IDENT:
      for (int r = 0; r < srcExt; r++) add(r, r, 0, 0);
      break;
    case PP_DIAG:
      if (potExt > srcExt) return PP_EARG;
      for (int i = 0; i < potExt; i++) add(i, i, i, 0);
      break;
    case PP_SIMP:
      if (!pairs) return PP_EARG;
      for (int i = 0; i < potExt; i++)
        if (!add(pairs[2 * i], pairs[2 * i + 1], i, 0)) return PP_EARG;
      break;
    case PP_HERM:
      if (!pairs) return PP_EARG;
      for (int i = 0; i < potExt; i++) {
        if (!add(pairs[2 * i], pairs[2 * i + 1], i, 0)) return PP_EARG;
        if (!add(pairs[2 * i + 1], pairs[2 * i], i, PP_ENT_CONJ)) return PP_EARG;
      }
      break;
    case PP_BANDNH: {
      int N = srcExt, bandCount = potExt / N, s = (bandCount - 1) / 2;
      for (int col = 0; col < N; col++)
        for (int row = col - s; row <= col + s; row++)
          if (row >= 0 && row < N) add(row, col, col * bandCount + s + row - col, 0);
      break;
    }
    case PP_BAND: {
      int N = srcExt, k = potExt / N;
      for (int j = 0; j < N; j++)
        for (int d = 0; d < k && j + d < N; d++) {
          const int slot = j * k + d;
          if (innermost) {
            if (d == 0) { add(j, j, slot, PP_ENT_OWN | PP_ENT_REAL); continue; }
            add(j + d, j, slot, PP_ENT_OWN);
            add(j, j + d, slot, PP_ENT_OWN | PP_ENT_CONJ);
          } else {
            add(j, j + d, slot, 0);
            if (d > 0) add(j + d, j, slot, PP_ENT_CONJ);
          }
        }
      break;
    }
    default:
      return PP_EARG;
  }
  csr.rowptr.assign((size_t)srcExt + 1, 0);
  csr.ent.clear();
  for (int r = 0; r < srcExt; r++) {
    csr.rowptr[r] = (int)(csr.ent.size() / 3);
    csr.ent.insert(csr.ent.end(), rows[r].begin(), rows[r].end());
  }
  csr.rowptr[srcExt] = (int)(csr.ent.size() / 3);
  return PP_OK;
}

struct pp_csr_dev {
  const int *rowptr[3];
  const int *ent[3];
  int shape[3];
  long potStride[3];
  long srcStride[3];
};

template <int R>
__global__ void __launch_bounds__(256) pp_tp_kernel(pp_csr_dev C, const zd *__restrict__ pot, double scaling,
                                                    const zd *__restrict__ src, zd *__restrict__ dst, long n)
{
  long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  int row[3] = {0, 0, 0};
  long rem = id;
  for (int r = R - 1; r >= 0; r--) { row[r] = (int)(rem % C.shape[r]); rem /= C.shape[r]; }
  zd acc = zmk(0.0, 0.0);
  // innermost entry with flags f, outer ranks' conjugation cj
  auto term = [&](zd v, zd x, int cj, int f) {
    if (f & PP_ENT_OWN) cj = 0;
    if (f & PP_ENT_REAL) v.y = 0.0;
    if ((cj ^ f) & PP_ENT_CONJ) zfmac(acc, v, x); else zfma(acc, v, x);
  };
  for (int e0 = C.rowptr[0][row[0]]; e0 < C.rowptr[0][row[0] + 1]; e0++) {
    const int *E0 = C.ent[0] + 3 * e0;
    long s0 = (long)E0[0] * C.srcStride[0], p0 = (long)E0[1] * C.potStride[0];
    int c0 = E0[2] & PP_ENT_CONJ;
    if (R == 1) {
      term(pot[p0], src[s0], 0, E0[2]);
      continue;
    }
    for (int e1 = C.rowptr[1][row[1]]; e1 < C.rowptr[1][row[1] + 1]; e1++) {
      const int *E1 = C.ent[1] + 3 * e1;
      long s1 = s0 + (long)E1[0] * C.srcStride[1], p1 = p0 + (long)E1[1] * C.potStride[1];
      if (R == 2) {
        term(pot[p1], src[s1], c0, E1[2]);
        continue;
      }
      const int c1 = c0 ^ (E1[2] & PP_ENT_CONJ);
      for (int e2 = C.rowptr[2][row[2]]; e2 < C.rowptr[2][row[2] + 1]; e2++) {
        const int *E2 = C.ent[2] + 3 * e2;
        term(pot[p1 + (long)E2[1] * C.potStride[2]], src[s1 + (long)E2[0] * C.srcStride[2]], c1, E2[2]);
      }
    }
  }
  zd d = dst[id];
  d.x = fma(scaling, acc.x, d.x);
  d.y = fma(scaling, acc.y, d.y);
  dst[id] = d;
}

// storage[R], potShape[R] (extents of the potential), shape[R] (extents of
// src/dst), pairs[r] host int32 (potShape[r] x 2) for Simp/Herm ranks or
// NULL.  potential/src/dst are device pointers to complex128.
extern "C" int pp_tp_multiply(int R, const int *storage, const int *potShape, const void *potential,
                              double scaling, const int *shape, const void *src, void *dst,
                              const int *const *pairs, hipStream_t stream)
{
  if (R < 1 || R > 3 || !storage || !potShape || !shape || !potential || !src || !dst) return PP_EARG;
  if (src == dst) return PP_EARG;   // the generated code assumes distinct buffers
  pp_rank_csr csr[3];
  pp_csr_dev C;
  memset(&C, 0, sizeof(C));
  long n = 1;
  for (int r = 0; r < R; r++) {
    int rc = pp_build_csr(storage[r], potShape[r], shape[r], pairs ? pairs[r] : nullptr, csr[r], r == R - 1);
    if (rc) return rc;
    n *= shape[r];
    C.shape[r] = shape[r];
  }
  C.potStride[R - 1] = 1; C.srcStride[R - 1] = 1;
  for (int r = R - 2; r >= 0; r--) {
    C.potStride[r] = C.potStride[r + 1] * potShape[r + 1];
    C.srcStride[r] = C.srcStride[r + 1] * shape[r + 1];
  }
  if (n == 0) return PP_OK;
  size_t bytes = 0;
  for (int r = 0; r < R; r++) bytes += (csr[r].rowptr.size() + csr[r].ent.size()) * sizeof(int);
  std::vector<int> host;
  host.reserve(bytes / sizeof(int));
  size_t off[3][2];
  for (int r = 0; r < R; r++) {
    off[r][0] = host.size(); host.insert(host.end(), csr[r].rowptr.begin(), csr[r].rowptr.end());
    off[r][1] = host.size(); host.insert(host.end(), csr[r].ent.begin(), csr[r].ent.end());
  }
  int *dev = nullptr;
  PP_HIP(hipMallocAsync((void **)&dev, host.size() * sizeof(int) + 4, stream));
  PP_HIP(hipMemcpyAsync(dev, host.data(), host.size() * sizeof(int), hipMemcpyHostToDevice, stream));
  PP_HIP(hipStreamSynchronize(stream));  // host vector goes out of scope below
  for (int r = 0; r < R; r++) { C.rowptr[r] = dev + off[r][0]; C.ent[r] = dev + off[r][1]; }
  dim3 grid((unsigned)((n + 255) / 256));
  if (R == 1) hipLaunchKernelGGL(pp_tp_kernel<1>, grid, dim3(256), 0, stream, C, (const zd *)potential, scaling, (const zd *)src, (zd *)dst, n);
  else if (R == 2) hipLaunchKernelGGL(pp_tp_kernel<2>, grid, dim3(256), 0, stream, C, (const zd *)potential, scaling, (const zd *)src, (zd *)dst, n);
  else hipLaunchKernelGGL(pp_tp_kernel<3>, grid, dim3(256), 0, stream, C, (const zd *)potential, scaling, (const zd *)src, (zd *)dst, n);
  PP_HIP(hipGetLastError());
  PP_HIP(hipFreeAsync(dev, stream));
  return PP_OK;
}

// ---------------------------------------------------------------- potential projection
// The B-spline rank's step of TensorPotentialGenerator.GeneratePotential
// (TensorPotentialGenerator.py:72-256 with basis_bspline.h's projection):
// out[r][e] = sum_q M^T[e][q] f[r][q] for every row r of the grid-evaluated
// potential, with M^T the sparse map from quadrature samples to the band
// pairs (<B_i| f d^d |B_j> = sum_q w_q B_i(q) f(q) d^d B_j(q) over the
// quadrature points the two supports share).  CSR rows of M^T in ascending
// q, so a pair sums its points in the same order on every run.
struct pp_projector {
  int nout, G;
  int *ptr, *idx;    // device CSR of M^T (nout + 1, nnz)
  double *val;       // device, nnz
};

// one thread per (row, pair); consecutive threads take consecutive pairs of a
// row, whose point ranges overlap, so the row's samples come from L1 / L2
__global__ void __launch_bounds__(256) pp_project_kernel(pp_projector pj, long rows, const zd *__restrict__ f,
                                                         zd *__restrict__ out)
{
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= rows * pj.nout) return;
  const long r = t / pj.nout;
  const int e = (int)(t % pj.nout);
  const zd *fr = f + r * pj.G;
  zd acc = zmk(0.0, 0.0);
  for (int k = pj.ptr[e]; k < pj.ptr[e + 1]; k++) zfmar(acc, pj.val[k], fr[pj.idx[k]]);
  out[t] = acc;
}

// CSR of M^T (host arrays: ptr nout+1, idx/val nnz; idx in [0, G)); copied.
extern "C" int pp_projector_create(int nout, int G, const int *ptr, const int *idx, const double *val,
                                   pp_projector **out)
{
  if (!out) return PP_EARG;
  *out = nullptr;
  if (nout < 1 || G < 1 || !ptr || !idx || !val || ptr[0] != 0) return PP_EARG;
  const int nnz = ptr[nout];
  for (int e = 0; e < nout; e++)
    if (ptr[e + 1] < ptr[e]) return PP_EARG;
  for (int k = 0; k < nnz; k++)
    if (idx[k] < 0 || idx[k] >= G) return PP_EARG;
  pp_projector *p = new pp_projector();
  p->nout = nout; p->G = G;
  p->ptr = p->idx = nullptr;
  p->val = nullptr;
  hipError_t e = hipMalloc((void **)&p->ptr, sizeof(int) * (nout + 1));
  if (e == hipSuccess) e = hipMalloc((void **)&p->idx, sizeof(int) * (nnz > 0 ? nnz : 1));
  if (e == hipSuccess) e = hipMalloc((void **)&p->val, sizeof(double) * (nnz > 0 ? nnz : 1));
  if (e == hipSuccess) e = hipMemcpy(p->ptr, ptr, sizeof(int) * (nout + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz > 0) e = hipMemcpy(p->idx, idx, sizeof(int) * nnz, hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz > 0) e = hipMemcpy(p->val, val, sizeof(double) * nnz, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(p->ptr); (void)hipFree(p->idx); (void)hipFree(p->val);
    delete p;
    return PP_EHIP;
  }
  *out = p;
  return PP_OK;
}

// out (rows x nout complex128, device) = f (rows x G complex128, device,
// the samples already multiplied by the quadrature weights) projected.
extern "C" int pp_projector_apply(const pp_projector *p, long rows, const void *f, void *out, hipStream_t stream)
{
  if (!p || rows < 0 || (!f && rows > 0) || (!out && rows > 0)) return PP_EARG;
  const long n = rows * p->nout;
  if (n == 0) return PP_OK;
  if ((n + 255) / 256 > 0x7fffffffL) return PP_EARG;
  hipLaunchKernelGGL(pp_project_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, *p, rows,
                     (const zd *)f, (zd *)out);
  return hipGetLastError() == hipSuccess ? PP_OK : PP_EHIP;
}

extern "C" int pp_projector_destroy(pp_projector *p)
{
  if (!p) return PP_OK;
  (void)hipFree(p->ptr);
  (void)hipFree(p->idx);
  (void)hipFree(p->val);
  delete p;
  return PP_OK;
}
