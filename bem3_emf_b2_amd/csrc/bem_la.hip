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

// ---- large-tile ZGEMM ----
// 128x128 tile per 256-thread workgroup (2x2 waves of 64x64, i.e. 4x4
// MFMA tiles of 16x16 per wave, real and imaginary accumulators), K slices
// of 8 complex values double-buffered in LDS as split real/imaginary planes
// (row padded by 4 doubles against bank conflicts), the next slice
// prefetched into registers while the current one feeds the MFMAs.  Tiles
// are assigned so that each XCD (workgroup id mod 8) works on a contiguous
// band of C, grouped 8 tile-rows high, for L2 reuse of the A and B panels.
#define ZT 128
#define ZK 8
#define ZP (ZT + 4)
#define ZLDS (2 * 4 * ZK * ZP * (int)sizeof(double))

__global__ void __launch_bounds__(256) k_zgemm128(int m, int n, int k, cd alpha, const cd *__restrict__ A, size_t lda,
                                                  const cd *__restrict__ B, size_t ldb, cd beta, cd *__restrict__ C,
                                                  size_t ldc, int tiles_x, int tiles_y)
{
  extern __shared__ double zsm[];
  const int T = tiles_x * tiles_y, per = (T + 7) / 8;
  const int L = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (L >= T) return;
  const int GM = 8, gw = GM * tiles_x;
  const int first = (L / gw) * GM, gsz = min(tiles_y - first, GM);
  const int ty = first + (L % gw) % gsz, tx = (L % gw) / gsz;
  const int i0 = ty * ZT, j0 = tx * ZT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  const int ra = tid >> 1, ca = (tid & 1) * 4;     // A slice: row, first k column
  const int rb = tid >> 5, cb = (tid & 31) * 4;    // B slice: k row, first column
  d4 cr[4][4], ci[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) { cr[a][b] = (d4){0, 0, 0, 0}; ci[a][b] = (d4){0, 0, 0, 0}; }
  cd pa[4], pb[4];
  auto gload = [&](int k0) {
    const int gi = i0 + ra, gk = k0 + rb;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      pa[q] = (gi < m && k0 + ca + q < k) ? A[(size_t)gi * lda + k0 + ca + q] : mk(0, 0);
      pb[q] = (gk < k && j0 + cb + q < n) ? B[(size_t)gk * ldb + j0 + cb + q] : mk(0, 0);
    }
  };
  auto lstore = [&](int s) {
    double *Ar = zsm + s * 4 * ZK * ZP, *Ai = Ar + ZK * ZP, *Br = Ai + ZK * ZP, *Bi = Br + ZK * ZP;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      Ar[(ca + q) * ZP + ra] = pa[q].x;
      Ai[(ca + q) * ZP + ra] = pa[q].y;
      Br[rb * ZP + cb + q] = pb[q].x;
      Bi[rb * ZP + cb + q] = pb[q].y;
    }
  };
  gload(0);
  lstore(0);
  __syncthreads();
  int s = 0;
  for (int k0 = 0; k0 < k; k0 += ZK) {
    const bool more = k0 + ZK < k;
    if (more) gload(k0 + ZK);
    const double *Ar = zsm + s * 4 * ZK * ZP, *Ai = Ar + ZK * ZP, *Br = Ai + ZK * ZP, *Bi = Br + ZK * ZP;
#pragma unroll
    for (int ks = 0; ks < ZK; ks += 4) {
      const int kr = ks + (lane >> 4), cc = lane & 15;
      double ar[4], ai[4], nai[4], br[4], bi[4];
#pragma unroll
      for (int a = 0; a < 4; a++) {
        ar[a] = Ar[kr * ZP + wr + 16 * a + cc];
        ai[a] = Ai[kr * ZP + wr + 16 * a + cc];
        nai[a] = -ai[a];
        br[a] = Br[kr * ZP + wc + 16 * a + cc];
        bi[a] = Bi[kr * ZP + wc + 16 * a + cc];
      }
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
          cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], cr[a][b], 0, 0, 0);
          cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[a], bi[b], cr[a][b], 0, 0, 0);
          ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], ci[a][b], 0, 0, 0);
          ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], ci[a][b], 0, 0, 0);
        }
    }
    if (more) lstore(s ^ 1);
    __syncthreads();
    s ^= 1;
  }
  // D layout of the 16x16 f64 MFMA (measured): lane l, register r -> row l/16 + 4r, col l%16
  const bool bz = beta.x == 0.0 && beta.y == 0.0;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gi = i0 + wr + 16 * a + (lane >> 4) + 4 * r;
        const int gj = j0 + wc + 16 * b + (lane & 15);
        if (gi < m && gj < n) {
          cd *c = C + (size_t)gi * ldc + gj;
          cd v = alpha * mk(cr[a][b][r], ci[a][b][r]);
          if (!bz) v = v + beta * (*c);
          *c = v;
        }
      }
}

// ---- LU panel ----
// The nb-column panel (rows k0..n-1) is factorised in a column-major copy P
// (P[c * ldp + r], r = row - k0) without moving rows: the pivot of column j is
// searched among the rows not yet used as pivots, and used[r] records the
// column at which row r became a pivot (INT_MAX while unused).  One launch
// per column: it applies the elimination of column j-1 to the thread's own
// row, then finds the per-workgroup maximum of column j; the last workgroup
// to finish (atomic ticket) reduces the partial maxima and publishes the
// pivot.  Rows are interchanged afterwards, once per panel, in LAPACK's
// sequential-swap form.
#define LU_NT 256

__global__ void __launch_bounds__(256) k_panel_load(const cd *__restrict__ A, size_t lda, int k0, int kb, int m,
                                                    cd *__restrict__ P, size_t ldp, int *__restrict__ used)
{
  __shared__ cd t[64][33];
  const int r0 = blockIdx.x * 32, tid = threadIdx.x;
  for (int idx = tid; idx < 32 * 64; idx += 256) {
    int rr = idx >> 6, c = idx & 63;
    if (r0 + rr < m && c < kb) t[c][rr] = A[(size_t)(k0 + r0 + rr) * lda + k0 + c];
  }
  __syncthreads();
  for (int idx = tid; idx < 32 * 64; idx += 256) {
    int c = idx >> 5, rr = idx & 31;
    if (r0 + rr < m && c < kb) P[(size_t)c * ldp + r0 + rr] = t[c][rr];
  }
  if (tid < 32 && r0 + tid < m) used[r0 + tid] = 0x7fffffff;
}

__global__ void __launch_bounds__(256) k_panel_store(cd *__restrict__ A, size_t lda, int k0, int kb, int m,
                                                     const cd *__restrict__ P, size_t ldp)
{
  __shared__ cd t[64][33];
  const int r0 = blockIdx.x * 32, tid = threadIdx.x;
  for (int idx = tid; idx < 32 * 64; idx += 256) {
    int c = idx >> 5, rr = idx & 31;
    if (r0 + rr < m && c < kb) t[c][rr] = P[(size_t)c * ldp + r0 + rr];
  }
  __syncthreads();
  for (int idx = tid; idx < 32 * 64; idx += 256) {
    int rr = idx >> 6, c = idx & 63;
    if (r0 + rr < m && c < kb) A[(size_t)(k0 + r0 + rr) * lda + k0 + c] = t[c][rr];
  }
}

__global__ void __launch_bounds__(LU_NT) k_panel_col(cd *__restrict__ P, size_t ldp, int m, int kb, int j,
                                                     int *__restrict__ used, int *__restrict__ pv,
                                                     unsigned long long *__restrict__ partV, int *__restrict__ partI,
                                                     unsigned int *__restrict__ ticket, int *__restrict__ info, int k0)
{
  __shared__ double sv[LU_NT];
  __shared__ int si[LU_NT];
  __shared__ bool last;
  const int tid = threadIdx.x, r = blockIdx.x * LU_NT + tid;
  // elimination of column j-1 on this row
  if (j > 0 && r < m && used[r] > j - 1) {
    const int p = pv[j - 1];
    // p is a row of the panel (published by the previous launch); the check
    // only keeps a corrupted workspace from turning into a wild access
    const bool ok = p >= 0 && p < m;
    const cd a = ok ? P[(size_t)(j - 1) * ldp + p] : mk(0, 0);
    if (a.x != 0.0 || a.y != 0.0) {
      cd l = cdiv(P[(size_t)(j - 1) * ldp + r], a);
      P[(size_t)(j - 1) * ldp + r] = l;
      for (int c = j; c < kb; c++) P[(size_t)c * ldp + r] = P[(size_t)c * ldp + r] - l * P[(size_t)c * ldp + p];
    }
  }
  if (j >= kb) return;
  // candidates for the pivot of column j
  double best = -1.0;
  int bidx = 0x7fffffff;
  if (r < m && used[r] > j - 1) {
    cd v = P[(size_t)j * ldp + r];
    best = fabs(v.x) + fabs(v.y);   // |re|+|im| as LAPACK izamax
    if (!(best >= 0.0)) best = INFINITY;   // a NaN entry still wins over non-candidates
    bidx = r;
  }
  sv[tid] = best;
  si[tid] = bidx;
  __syncthreads();
  for (int s = LU_NT / 2; s > 0; s >>= 1) {
    if (tid < s) {
      double o = sv[tid + s];
      int oi = si[tid + s];
      if (o > sv[tid] || (o == sv[tid] && oi < si[tid])) { sv[tid] = o; si[tid] = oi; }
    }
    __syncthreads();
  }
  if (tid == 0) {
    __hip_atomic_store(&partV[blockIdx.x], (unsigned long long)__double_as_longlong(sv[0]), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&partI[blockIdx.x], si[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned int t = atomicAdd(ticket, 1u);
    last = t == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double b = -1.0;
  int bi = 0x7fffffff;
  for (int q = tid; q < (int)gridDim.x; q += LU_NT) {
    double o = __longlong_as_double(
        (long long)__hip_atomic_load(&partV[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    int oi = __hip_atomic_load(&partI[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (o > b || (o == b && oi < bi)) { b = o; bi = oi; }
  }
  sv[tid] = b;
  si[tid] = bi;
  __syncthreads();
  for (int s = LU_NT / 2; s > 0; s >>= 1) {
    if (tid < s) {
      double o = sv[tid + s];
      int oi = si[tid + s];
      if (o > sv[tid] || (o == sv[tid] && oi < si[tid])) { sv[tid] = o; si[tid] = oi; }
    }
    __syncthreads();
  }
  if (tid == 0) {
    int p = si[0];   // a candidate row: m - j >= 1 rows are unused at column j
    if (p >= m) { p = m - 1; if (*info == 0) *info = k0 + j + 1; }
    pv[j] = p;
    used[p] = j;
    if (!(sv[0] > 0.0) && *info == 0) *info = k0 + j + 1;
    *ticket = 0u;
  }
}

// pivot rows of the panel -> LAPACK interchanges piv[k0+j] (absolute rows)
__global__ void k_panel_ipiv(const int *__restrict__ pv, int kb, int k0, int *__restrict__ piv)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int ip[64];
  for (int j = 0; j < kb; j++) {
    int cur = pv[j];
    for (int q = 0; q < j; q++) {
      if (cur == q) cur = ip[q];
      else if (cur == ip[q]) cur = q;
    }
    ip[j] = cur;
    piv[k0 + j] = k0 + cur;
  }
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
static bool g_big_tiles = false;   // set once the 128x128 kernel's LDS limit is raised

int bem_la_init(void)
{
  const int bytes = 2 * 64 * 64 * (int)sizeof(cd);
  if (hipFuncSetAttribute((const void *)k_trsm_lunit, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess) return -1;
  if (hipFuncSetAttribute((const void *)k_trsm_upper, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess) return -1;
  if (hipFuncSetAttribute((const void *)k_zgemm128, hipFuncAttributeMaxDynamicSharedMemorySize, ZLDS) != hipSuccess) return -1;
  g_big_tiles = true;
  return 0;
}

void bem_zgemm(int m, int n, int k, double ar, double ai, const void *A, size_t lda, const void *B, size_t ldb,
               double br, double bi, void *C, size_t ldc, hipStream_t st)
{
  if (m <= 0 || n <= 0) return;
  bem_prof_begin(BH_K_ZGEMM, 8.0 * m * n * k, st);
  if (g_big_tiles && (size_t)m * n >= (size_t)512 * ZT * ZT) {
    // enough 128x128 tiles to fill the chip twice over
    int tx = (n + ZT - 1) / ZT, ty = (m + ZT - 1) / ZT, T = tx * ty;
    hipLaunchKernelGGL(k_zgemm128, dim3(8 * ((T + 7) / 8)), dim3(256), ZLDS, st, m, n, k, mk(ar, ai), (const cd *)A,
                       lda, (const cd *)B, ldb, mk(br, bi), (cd *)C, ldc, tx, ty);
  } else {
    dim3 grid((n + GB_N - 1) / GB_N, (m + GB_M - 1) / GB_M);
    hipLaunchKernelGGL(k_zgemm, grid, dim3(256), 0, st, m, n, k, mk(ar, ai), (const cd *)A, lda, (const cd *)B, ldb,
                       mk(br, bi), (cd *)C, ldc);
  }
  bem_prof_end(BH_K_ZGEMM, st);
}

// grow-only scratch of the panel factorisation (device memory kept for the
// life of the process; calls are serialised by the C-ABI)
static void *g_lu_ws = nullptr;
static size_t g_lu_ws_bytes = 0;

// LU factorisation in place (nb <= 64).  piv (device, n ints) receives
// 0-based pivot rows; info (device int) receives 0 or the 1-based index of
// the first zero pivot.  Returns 0, or -1 if the scratch cannot be allocated.
int bem_zgetrf(int n, void *Av, size_t lda, int *piv, void *work, int *info, int nb, hipStream_t st)
{
  (void)work;
  cd *A = (cd *)Av;
  if (nb > 64) nb = 64;
  size_t ldp = ((size_t)n + 63) / 64 * 64;
  int nblk = (n + LU_NT - 1) / LU_NT;
  size_t need = 256 + (size_t)nb * ldp * sizeof(cd) + ldp * sizeof(int) + 64 * sizeof(int) +
                (size_t)nblk * (sizeof(unsigned long long) + sizeof(int));
  if (need > g_lu_ws_bytes) {
    if (g_lu_ws) {
      (void)hipStreamSynchronize(st);   // earlier factorisations may still use it
      (void)hipFree(g_lu_ws);
    }
    g_lu_ws = nullptr;
    g_lu_ws_bytes = 0;
    if (hipMalloc(&g_lu_ws, need) != hipSuccess) return -1;
    g_lu_ws_bytes = need;
  }
  // layout for this n (the buffer may be larger, from an earlier call):
  // ticket at a fixed offset, cleared on every call
  char *w = (char *)g_lu_ws;
  unsigned int *ticket = (unsigned int *)w; w += 256;
  cd *P = (cd *)w;                       w += (size_t)nb * ldp * sizeof(cd);
  int *used = (int *)w;                  w += ldp * sizeof(int);
  int *pv = (int *)w;                    w += 64 * sizeof(int);
  unsigned long long *partV = (unsigned long long *)w; w += (size_t)nblk * sizeof(unsigned long long);
  int *partI = (int *)w;
  (void)hipMemsetAsync(ticket, 0, sizeof(unsigned int), st);
  (void)hipMemsetAsync(info, 0, sizeof(int), st);
  for (int k0 = 0; k0 < n; k0 += nb) {
    int kb = n - k0 < nb ? n - k0 : nb, k1 = k0 + kb, m = n - k0;
    int gb = (m + LU_NT - 1) / LU_NT;
    // panel: about 8 m kb^2 / 2 real flops of rank-1 updates
    bem_prof_begin(BH_K_LU_PANEL, 4.0 * (double)m * kb * kb, st);
    hipLaunchKernelGGL(k_panel_load, dim3((m + 31) / 32), dim3(256), 0, st, A, lda, k0, kb, m, P, ldp, used);
    for (int j = 0; j <= kb; j++)
      hipLaunchKernelGGL(k_panel_col, dim3(gb), dim3(LU_NT), 0, st, P, ldp, m, kb, j, used, pv, partV, partI, ticket,
                         info, k0);
    hipLaunchKernelGGL(k_panel_store, dim3((m + 31) / 32), dim3(256), 0, st, A, lda, k0, kb, m, P, ldp);
    hipLaunchKernelGGL(k_panel_ipiv, dim3(1), dim3(64), 0, st, pv, kb, k0, piv);
    bem_prof_end(BH_K_LU_PANEL, st);
    // interchanges on every column (the panel was factorised in place order)
    hipLaunchKernelGGL(k_laswp, dim3((n + 255) / 256), dim3(256), 0, st, A, lda, 0, n, k0, k1, piv);
    if (k1 < n) {
      int nr = n - k1;
      size_t shm = (size_t)(kb * kb + kb * 64) * sizeof(cd);
      bem_prof_begin(BH_K_TRSM, 4.0 * kb * kb * (double)nr, st);
      hipLaunchKernelGGL(k_trsm_lunit, dim3((nr + 63) / 64), dim3(64), shm, st, A + (size_t)k0 * lda + k0, lda, kb,
                         A + (size_t)k0 * lda + k1, lda, nr);
      bem_prof_end(BH_K_TRSM, st);
      bem_zgemm(nr, nr, kb, -1.0, 0.0, A + (size_t)k1 * lda + k0, lda, A + (size_t)k0 * lda + k1, lda, 1.0, 0.0,
                A + (size_t)k1 * lda + k1, lda, st);
    }
  }
  return 0;
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
