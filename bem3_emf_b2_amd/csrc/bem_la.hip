// bem_la.hip -- dense complex linear algebra for the BIE solve (FP64):
//   * ZGEMM C = alpha A B + beta C on v_mfma_f64_16x16x4f64 (4 real MFMAs per
//     complex product), 64x64 tiles per 256-thread workgroup, 16-deep K
//     slices staged in LDS as split real/imaginary planes;
//   * blocked right-looking LU with partial pivoting (row-major): per-column
//     pivot/swap + panel update kernels, row interchanges on the rest of the
//     matrix, unit-lower triangular solve for the U12 block row, and the
//     rank-nb trailing update on the ZGEMM above;
//   * forward/backward substitution with the factors for any number of
//     right-hand sides.
// Replaces LAPACKE_zgetrf/zgetri (solve_bieq2.c:389-399) and zgesv
// (mo_solve.c:751) of the reference.
#include "bem_dev.h"
#include "bem_prof.h"

typedef double d4 __attribute__((ext_vector_type(4)));

#define GB_M 64
#define GB_N 64
#define GB_K 16

// C[m x n] = alpha * A[m x k] B[k x n] + beta * C   (row-major, complex)
__global__ void __launch_bounds__(256) k_zgemm(int m, int n, int k, cd alpha, const cd *__restrict__ A, size_t lda,
                                               const cd *__restrict__ B, size_t ldb, cd beta, cd *__restrict__ C,
                                               size_t ldc)
{
  __shared__ double sAr[GB_K][GB_M + 1], sAi[GB_K][GB_M + 1];
  __shared__ double sBr[GB_K][GB_N + 1], sBi[GB_K][GB_N + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = blockIdx.y * GB_M, j0 = blockIdx.x * GB_N;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;   // wave sub-tile origin
  d4 cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) { cr[a][b] = (d4){0, 0, 0, 0}; ci[a][b] = (d4){0, 0, 0, 0}; }
  for (int k0 = 0; k0 < k; k0 += GB_K) {
    // stage A (64 x 16) and B (16 x 64): 4 complex values per thread each
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int idx = tid + q * 256;
      int ar = idx >> 4, ac = idx & 15;          // A tile row, k column
      int gi = i0 + ar, gk = k0 + ac;
      cd va = (gi < m && gk < k) ? A[(size_t)gi * lda + gk] : mk(0, 0);
      sAr[ac][ar] = va.x; sAi[ac][ar] = va.y;
      int br = idx >> 6, bc = idx & 63;          // B tile k row, column
      int gk2 = k0 + br, gj = j0 + bc;
      cd vb = (gk2 < k && gj < n) ? B[(size_t)gk2 * ldb + gj] : mk(0, 0);
      sBr[br][bc] = vb.x; sBi[br][bc] = vb.y;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GB_K; kk += 4) {
      int kr = kk + (lane >> 4), cc = lane & 15;
      double ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        ar[a] = sAr[kr][wr + 16 * a + cc];
        ai[a] = sAi[kr][wr + 16 * a + cc];
        nai[a] = -ai[a];
        br[a] = sBr[kr][wc + 16 * a + cc];
        bi[a] = sBi[kr][wc + 16 * a + cc];
      }
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], cr[a][b], 0, 0, 0);
          cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[a], bi[b], cr[a][b], 0, 0, 0);
          ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], ci[a][b], 0, 0, 0);
          ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], ci[a][b], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  // D layout of the 16x16 f64 MFMA (measured): lane l, register r -> row l/16 + 4r, col l%16
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int gi = i0 + wr + 16 * a + (lane >> 4) + 4 * r;
        int gj = j0 + wc + 16 * b + (lane & 15);
        if (gi < m && gj < n) {
          cd ab = mk(cr[a][b][r], ci[a][b][r]);
          cd *c = C + (size_t)gi * ldc + gj;
          cd v = alpha * ab;
          if (beta.x != 0.0 || beta.y != 0.0) v = v + beta * (*c);
          *c = v;
        }
      }
}

// ---- LU helpers ----
// pivot search in column j (rows j..n-1), swap rows j and p inside the panel
// columns [c0, c1), record piv[j] = p and keep the pivot row segment.
__global__ void __launch_bounds__(1024) k_lu_pivot(cd *A, size_t lda, int n, int j, int c0, int c1, int *piv,
                                                   cd *prow, int *info)
{
  __shared__ double sv[1024];
  __shared__ int si[1024];
  const int tid = threadIdx.x;
  double best = -1.0;
  int bi = j;
  for (int i = j + tid; i < n; i += 1024) {
    cd v = A[(size_t)i * lda + j];
    double m = fabs(v.x) + fabs(v.y);   // |re|+|im| as LAPACK izamax
    if (m > best) { best = m; bi = i; }
  }
  sv[tid] = best; si[tid] = bi;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (tid < s) {
      double o = sv[tid + s];
      int oi = si[tid + s];
      if (o > sv[tid] || (o == sv[tid] && oi < si[tid])) { sv[tid] = o; si[tid] = oi; }
    }
    __syncthreads();
  }
  int p = si[0];
  if (tid == 0) {
    piv[j] = p;
    if (sv[0] == 0.0 && *info == 0) *info = j + 1;
  }
  for (int c = c0 + tid; c < c1; c += 1024) {
    cd a = A[(size_t)j * lda + c], b = A[(size_t)p * lda + c];
    A[(size_t)j * lda + c] = b;
    A[(size_t)p * lda + c] = a;
    prow[c - c0] = b;
  }
}

// scale column j below the diagonal and update panel columns j+1..c1-1
__global__ void __launch_bounds__(256) k_lu_update(cd *A, size_t lda, int n, int j, int c0, int c1, const cd *prow)
{
  int i = j + 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cd p = prow[j - c0];
  if (p.x == 0.0 && p.y == 0.0) return;
  cd *ri = A + (size_t)i * lda;
  cd l = cdiv(ri[j], p);
  ri[j] = l;
  for (int c = j + 1; c < c1; c++) ri[c] = ri[c] - l * prow[c - c0];
}

// apply interchanges piv[k0..k1) to columns [c0, c1) of A (or of a RHS block)
__global__ void __launch_bounds__(256) k_laswp(cd *A, size_t lda, int c0, int c1, int k0, int k1, const int *piv)
{
  int c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= c1) return;
  for (int j = k0; j < k1; j++) {
    int p = piv[j];
    if (p != j) {
      cd a = A[(size_t)j * lda + c];
      A[(size_t)j * lda + c] = A[(size_t)p * lda + c];
      A[(size_t)p * lda + c] = a;
    }
  }
}

// X := L^{-1} X with L the unit-lower nb x nb block at A(k0,k0), X the
// nb x ncol block at B (row stride ldb); 64 columns per workgroup
__global__ void __launch_bounds__(64) k_trsm_lunit(const cd *__restrict__ L, size_t lda, int nb, cd *X, size_t ldb,
                                                   int ncol)
{
  extern __shared__ double sm[];
  cd *sL = (cd *)sm;                  // nb x nb
  cd *sX = sL + nb * nb;              // nb x 64
  const int c = threadIdx.x, gc = blockIdx.x * 64 + c;
  for (int idx = c; idx < nb * nb; idx += 64) sL[idx] = L[(size_t)(idx / nb) * lda + idx % nb];
  for (int r = 0; r < nb; r++) sX[r * 64 + c] = gc < ncol ? X[(size_t)r * ldb + gc] : mk(0, 0);
  __syncthreads();
  for (int r = 1; r < nb; r++) {
    cd acc = sX[r * 64 + c];
    for (int q = 0; q < r; q++) acc = acc - sL[r * nb + q] * sX[q * 64 + c];
    sX[r * 64 + c] = acc;
  }
  if (gc < ncol)
    for (int r = 0; r < nb; r++) X[(size_t)r * ldb + gc] = sX[r * 64 + c];
}

// X := U^{-1} X with U the upper nb x nb block (non-unit diagonal)
__global__ void __launch_bounds__(64) k_trsm_upper(const cd *__restrict__ U, size_t lda, int nb, cd *X, size_t ldb,
                                                   int ncol)
{
  extern __shared__ double sm[];
  cd *sU = (cd *)sm;
  cd *sX = sU + nb * nb;
  const int c = threadIdx.x, gc = blockIdx.x * 64 + c;
  for (int idx = c; idx < nb * nb; idx += 64) sU[idx] = U[(size_t)(idx / nb) * lda + idx % nb];
  for (int r = 0; r < nb; r++) sX[r * 64 + c] = gc < ncol ? X[(size_t)r * ldb + gc] : mk(0, 0);
  __syncthreads();
  for (int r = nb - 1; r >= 0; r--) {
    cd acc = sX[r * 64 + c];
    for (int q = r + 1; q < nb; q++) acc = acc - sU[r * nb + q] * sX[q * 64 + c];
    sX[r * 64 + c] = cdiv(acc, sU[r * nb + r]);
  }
  if (gc < ncol)
    for (int r = 0; r < nb; r++) X[(size_t)r * ldb + gc] = sX[r * 64 + c];
}

extern "C" {

// the triangular-solve kernels stage a 64x64 factor block and a 64x64 RHS
// block (128 KiB) in LDS, above the default dynamic-LDS limit
int bem_la_init(void)
{
  const int bytes = 2 * 64 * 64 * (int)sizeof(cd);
  if (hipFuncSetAttribute((const void *)k_trsm_lunit, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess) return -1;
  if (hipFuncSetAttribute((const void *)k_trsm_upper, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess) return -1;
  return 0;
}

void bem_zgemm(int m, int n, int k, double ar, double ai, const void *A, size_t lda, const void *B, size_t ldb,
               double br, double bi, void *C, size_t ldc, hipStream_t st)
{
  if (m <= 0 || n <= 0) return;
  dim3 grid((n + GB_N - 1) / GB_N, (m + GB_M - 1) / GB_M);
  bem_prof_begin(BH_K_ZGEMM, 8.0 * m * n * k, st);
  hipLaunchKernelGGL(k_zgemm, grid, dim3(256), 0, st, m, n, k, mk(ar, ai), (const cd *)A, lda, (const cd *)B, ldb,
                     mk(br, bi), (cd *)C, ldc);
  bem_prof_end(BH_K_ZGEMM, st);
}

// LU factorisation in place.  piv (device, n ints) receives 0-based pivot
// rows; work (device) holds at least nb complex values; info (device int)
// receives 0 or the 1-based index of the first zero pivot.
void bem_zgetrf(int n, void *Av, size_t lda, int *piv, void *work, int *info, int nb, hipStream_t st)
{
  cd *A = (cd *)Av;
  cd *prow = (cd *)work;
  (void)hipMemsetAsync(info, 0, sizeof(int), st);
  for (int k0 = 0; k0 < n; k0 += nb) {
    int kb = n - k0 < nb ? n - k0 : nb, k1 = k0 + kb;
    // panel: about 8 (n - k0) kb^2 / 2 real flops of rank-1 updates
    bem_prof_begin(BH_K_LU_PANEL, 4.0 * (double)(n - k0) * kb * kb, st);
    for (int j = k0; j < k1; j++) {
      hipLaunchKernelGGL(k_lu_pivot, dim3(1), dim3(1024), 0, st, A, lda, n, j, k0, k1, piv, prow, info);
      int rows = n - j - 1;
      if (rows > 0) hipLaunchKernelGGL(k_lu_update, dim3((rows + 255) / 256), dim3(256), 0, st, A, lda, n, j, k0, k1, prow);
    }
    bem_prof_end(BH_K_LU_PANEL, st);
    // interchanges on the columns left and right of the panel
    if (k0 > 0) hipLaunchKernelGGL(k_laswp, dim3((k0 + 255) / 256), dim3(256), 0, st, A, lda, 0, k0, k0, k1, piv);
    if (k1 < n) {
      int nr = n - k1;
      hipLaunchKernelGGL(k_laswp, dim3((nr + 255) / 256), dim3(256), 0, st, A, lda, k1, n, k0, k1, piv);
      size_t shm = (size_t)(kb * kb + kb * 64) * sizeof(cd);
      bem_prof_begin(BH_K_TRSM, 4.0 * kb * kb * (double)nr, st);
      hipLaunchKernelGGL(k_trsm_lunit, dim3((nr + 63) / 64), dim3(64), shm, st, A + (size_t)k0 * lda + k0, lda, kb,
                         A + (size_t)k0 * lda + k1, lda, nr);
      bem_prof_end(BH_K_TRSM, st);
      bem_zgemm(nr, nr, kb, -1.0, 0.0, A + (size_t)k1 * lda + k0, lda, A + (size_t)k0 * lda + k1, lda, 1.0, 0.0,
                A + (size_t)k1 * lda + k1, lda, st);
    }
  }
}

// solve A X = B with the factors of bem_zgetrf; B is n x nrhs (row stride ldb)
void bem_zgetrs(int n, int nrhs, const void *Av, size_t lda, const int *piv, void *Bv, size_t ldb, int nb, hipStream_t st)
{
  const cd *A = (const cd *)Av;
  cd *B = (cd *)Bv;
  hipLaunchKernelGGL(k_laswp, dim3((nrhs + 255) / 256), dim3(256), 0, st, B, ldb, 0, nrhs, 0, n, piv);
  // forward substitution with the unit lower factor
  for (int k0 = 0; k0 < n; k0 += nb) {
    int kb = n - k0 < nb ? n - k0 : nb, k1 = k0 + kb;
    size_t shm = (size_t)(kb * kb + kb * 64) * sizeof(cd);
    hipLaunchKernelGGL(k_trsm_lunit, dim3((nrhs + 63) / 64), dim3(64), shm, st, A + (size_t)k0 * lda + k0, lda, kb,
                       B + (size_t)k0 * ldb, ldb, nrhs);
    if (k1 < n)
      bem_zgemm(n - k1, nrhs, kb, -1.0, 0.0, A + (size_t)k1 * lda + k0, lda, B + (size_t)k0 * ldb, ldb, 1.0, 0.0,
                B + (size_t)k1 * ldb, ldb, st);
  }
  // backward substitution with the upper factor
  int last = ((n - 1) / nb) * nb;
  for (int k0 = last; k0 >= 0; k0 -= nb) {
    int kb = n - k0 < nb ? n - k0 : nb;
    size_t shm = (size_t)(kb * kb + kb * 64) * sizeof(cd);
    hipLaunchKernelGGL(k_trsm_upper, dim3((nrhs + 63) / 64), dim3(64), shm, st, A + (size_t)k0 * lda + k0, lda, kb,
                       B + (size_t)k0 * ldb, ldb, nrhs);
    if (k0 > 0)
      bem_zgemm(k0, nrhs, kb, -1.0, 0.0, A + k0, lda, B + (size_t)k0 * ldb, ldb, 1.0, 0.0, B, ldb, st);
  }
}
}
