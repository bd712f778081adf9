// Two-electron radial preconditioner: the exact solve of one Kronecker-sum
// matrix per coupled channel c on a (C, N, N) state,
//
//     M_c = b S (x) S + a (H1_c (x) S + S (x) H2_c),
//
// where H1_c, H2_c are real symmetric N x N matrices (the one-electron
// radial kinetic, centrifugal and nuclear terms of the channel's l1 and l2)
// and S is the B-spline overlap.  With the S-orthonormal generalised
// eigenvectors H V = S V diag(lam), V^T S V = I of each distinct matrix,
//
//     M_c^-1 = (V1 (x) V2) diag(1 / (b + a (lam1_i + lam2_k))) (V1^T (x) V2^T),
//
// which on the channel's N x N block X is V1 [(V1^T X V2) ./ D] V2^T: four
// dense real-by-complex N x N products per channel and one division, fused
// into the second product.  This replaces the per-channel sparse
// factorisation of RadialTwoElectronPreconditioner
// (examples/tensor/helium_stabilization/preconditioner.py:77-200) for the
// separable part of its matrix.
//
// Kernel: one workgroup per TI x TK output tile of one channel; the real
// matrix and the complex operand stream through LDS in KC-deep slices; each
// thread accumulates a 4 x 4 block of complex outputs (rows ty + TY r,
// columns tx + TX q, so that LDS reads are broadcasts or consecutive).
// Compute-bound in FP64 (4 N^3 flops per channel and pass against
// 32 N^2 bytes of HBM traffic).
#include "pp_common.h"
#include <type_traits>

#ifndef PR_TI
#define PR_TI 64
#endif
#ifndef PR_TK
#define PR_TK 64
#endif
#define PR_KC 16
#define PR_TY (PR_TI / 4)
#define PR_TX (PR_TK / 4)
#define PR_THREADS (PR_TX * PR_TY)

struct pr_prec {
  int C, N, nmat;
  int *idx1, *idx2;  // device, matrix index of each channel on radial rank 1 / 2
  double *V, *VT;    // device, nmat x N x N row-major eigenvector matrices and their transposes
  double *lam;       // device, nmat x N eigenvalues
  zd *tmp;           // device, C x N x N intermediate
};

__device__ inline void pr_mac(zd &acc, double r, zd z)
{
  acc.x = fma(r, z.x, acc.x);
  acc.y = fma(r, z.y, acc.y);
}
__device__ inline void pr_mac(zd &acc, zd z, double r)
{
  acc.x = fma(r, z.x, acc.x);
  acc.y = fma(r, z.y, acc.y);
}
template <typename T> __device__ inline T pr_zero();
template <> __device__ inline double pr_zero<double>() { return 0.0; }
template <> __device__ inline zd pr_zero<zd>() { return zmk(0.0, 0.0); }

// Y[c] = R_{ridx[c]} Z[c] (RIGHT false) or Z[c] R_{ridx[c]} (RIGHT true), R
// real N x N row-major, Z and Y complex N x N per channel.  SCALE divides
// output (i, k) by b + a (lam[idx1[c]][i] + lam[idx2[c]][k]).
template <bool RIGHT, bool SCALE>
__global__ __launch_bounds__(PR_THREADS) void pr_mm(const zd *__restrict__ Z, zd *__restrict__ Y,
                                                    const double *__restrict__ R, const int *__restrict__ ridx,
                                                    int N, int tilesI, int tilesK, const double *__restrict__ lam,
                                                    const int *__restrict__ idx1, const int *__restrict__ idx2,
                                                    zd a, zd b)
{
  typedef typename std::conditional<RIGHT, zd, double>::type RowT;   // operand indexed by output row
  typedef typename std::conditional<RIGHT, double, zd>::type ColT;   // operand indexed by output column
  __shared__ RowT srow[PR_KC][PR_TI + 1];
  __shared__ ColT scol[PR_KC][PR_TK];

  const int tile = blockIdx.x;
  const int tk = tile % tilesK;
  const int ti = (tile / tilesK) % tilesI;
  const int c = tile / (tilesK * tilesI);
  const int i0 = ti * PR_TI, k0 = tk * PR_TK;
  const int tid = threadIdx.x;
  const int tx = tid % PR_TX, ty = tid / PR_TX;
  const size_t NN = (size_t)N * N;
  const zd *Zc = Z + (size_t)c * NN;
  const double *Rc = R + (size_t)ridx[c] * NN;

  zd acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[r][q] = zmk(0.0, 0.0);

  for (int j0 = 0; j0 < N; j0 += PR_KC) {
    // row operand slice: rows i0.., columns j0.. (R for the left product, Z for the right)
    for (int t = tid; t < PR_TI * PR_KC; t += PR_THREADS) {
      const int jj = t % PR_KC, i = t / PR_KC;
      const int gi = i0 + i, gj = j0 + jj;
      RowT v = pr_zero<RowT>();
      if (gi < N && gj < N) {
        if constexpr (RIGHT) v = Zc[(size_t)gi * N + gj];
        else v = Rc[(size_t)gi * N + gj];
      }
      srow[jj][i] = v;
    }
    // column operand slice: rows j0.., columns k0..
    for (int t = tid; t < PR_KC * PR_TK; t += PR_THREADS) {
      const int k = t % PR_TK, jj = t / PR_TK;
      const int gj = j0 + jj, gk = k0 + k;
      ColT v = pr_zero<ColT>();
      if (gj < N && gk < N) {
        if constexpr (RIGHT) v = Rc[(size_t)gj * N + gk];
        else v = Zc[(size_t)gj * N + gk];
      }
      scol[jj][k] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int jj = 0; jj < PR_KC; ++jj) {
      RowT rv[4];
      ColT cv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) rv[r] = srow[jj][ty + PR_TY * r];
#pragma unroll
      for (int q = 0; q < 4; ++q) cv[q] = scol[jj][tx + PR_TX * q];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) pr_mac(acc[r][q], rv[r], cv[q]);
    }
    __syncthreads();
  }

  zd *Yc = Y + (size_t)c * NN;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + ty + PR_TY * r;
    if (i >= N) continue;
    double l1 = 0.0;
    if (SCALE) l1 = lam[(size_t)idx1[c] * N + i];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + tx + PR_TX * q;
      if (k >= N) continue;
      zd v = acc[r][q];
      if (SCALE) {
        const double s = l1 + lam[(size_t)idx2[c] * N + k];
        const zd d = zmk(b.x + a.x * s, b.y + a.y * s);
        const double inv = 1.0 / zabs2(d);
        v = zscale(zmul(v, zconj(d)), inv);
      }
      Yc[(size_t)i * N + k] = v;
    }
  }
}

static void pr_free(pr_prec *p)
{
  if (!p) return;
  (void)hipFree(p->idx1);
  (void)hipFree(p->idx2);
  (void)hipFree(p->V);
  (void)hipFree(p->VT);
  (void)hipFree(p->lam);
  (void)hipFree(p->tmp);
  delete p;
}

// C channels of N x N blocks; nmat distinct eigen-decompositions: V (nmat x
// N x N row-major, column n the n-th eigenvector, V^T S V = I) and lam (nmat x
// N); idx1[c], idx2[c] in [0, nmat) select the decomposition of channel c on
// radial rank 1 and 2.  Host pointers; the data are copied.
extern "C" int pp_radial_prec_create(int C, int N, int nmat, const int *idx1, const int *idx2, const double *V,
                                     const double *lam, void **out)
{
  if (!out) return PP_EARG;
  *out = nullptr;
  if (C <= 0 || N <= 0 || nmat <= 0 || !idx1 || !idx2 || !V || !lam) return PP_EARG;
  for (int c = 0; c < C; ++c)
    if (idx1[c] < 0 || idx1[c] >= nmat || idx2[c] < 0 || idx2[c] >= nmat) return PP_EARG;
  const size_t NN = (size_t)N * N;
  std::vector<double> VT(nmat * NN);
  for (int m = 0; m < nmat; ++m)
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) VT[m * NN + (size_t)j * N + i] = V[m * NN + (size_t)i * N + j];
  pr_prec *p = new pr_prec();
  p->C = C;
  p->N = N;
  p->nmat = nmat;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc((void **)&p->idx1, sizeof(int) * C);
  if (e == hipSuccess) e = hipMalloc((void **)&p->idx2, sizeof(int) * C);
  if (e == hipSuccess) e = hipMalloc((void **)&p->V, sizeof(double) * nmat * NN);
  if (e == hipSuccess) e = hipMalloc((void **)&p->VT, sizeof(double) * nmat * NN);
  if (e == hipSuccess) e = hipMalloc((void **)&p->lam, sizeof(double) * nmat * N);
  if (e == hipSuccess) e = hipMalloc((void **)&p->tmp, sizeof(zd) * C * NN);
  if (e == hipSuccess) e = hipMemcpy(p->idx1, idx1, sizeof(int) * C, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->idx2, idx2, sizeof(int) * C, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->V, V, sizeof(double) * nmat * NN, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->VT, VT.data(), sizeof(double) * nmat * NN, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->lam, lam, sizeof(double) * nmat * N, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    pr_free(p);
    return PP_EHIP;
  }
  *out = p;
  return PP_OK;
}

// psi (C x N x N complex128, device) <- M^-1 psi with M_c = b S(x)S +
// a (H1_c (x) S + S (x) H2_c).  Asynchronous on `stream`; the handle's
// intermediate buffer makes overlapping solves on one handle race.
extern "C" int pp_radial_prec_solve(void *h, void *psi, double a_re, double a_im, double b_re, double b_im,
                                    hipStream_t stream)
{
  pr_prec *p = (pr_prec *)h;
  if (!p || !psi) return PP_EARG;
  const int N = p->N;
  const int tilesI = (N + PR_TI - 1) / PR_TI, tilesK = (N + PR_TK - 1) / PR_TK;
  const long blocks = (long)p->C * tilesI * tilesK;
  if (blocks > 0x7fffffffL) return PP_EARG;
  const dim3 grid((unsigned)blocks), block(PR_THREADS);
  const zd a = zmk(a_re, a_im), b = zmk(b_re, b_im);
  zd *x = (zd *)psi;
  // T = V1^T X;  X = (T V2) ./ D;  T = V1 X;  X = T V2^T
  hipLaunchKernelGGL((pr_mm<false, false>), grid, block, 0, stream, x, p->tmp, p->VT, p->idx1, N, tilesI, tilesK,
                     p->lam, p->idx1, p->idx2, a, b);
  hipLaunchKernelGGL((pr_mm<true, true>), grid, block, 0, stream, p->tmp, x, p->V, p->idx2, N, tilesI, tilesK,
                     p->lam, p->idx1, p->idx2, a, b);
  hipLaunchKernelGGL((pr_mm<false, false>), grid, block, 0, stream, x, p->tmp, p->V, p->idx1, N, tilesI, tilesK,
                     p->lam, p->idx1, p->idx2, a, b);
  hipLaunchKernelGGL((pr_mm<true, false>), grid, block, 0, stream, p->tmp, x, p->VT, p->idx2, N, tilesI, tilesK,
                     p->lam, p->idx1, p->idx2, a, b);
  return hipGetLastError() == hipSuccess ? PP_OK : PP_EHIP;
}

extern "C" int pp_radial_prec_destroy(void *h)
{
  pr_free((pr_prec *)h);
  return PP_OK;
}
