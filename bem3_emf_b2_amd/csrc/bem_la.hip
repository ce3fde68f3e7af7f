// bem_la.hip -- dense complex linear algebra for the BIE solve, in complex
// FP64 (cd) and complex FP32 (cf) from the same templates:
//   * GEMM C = alpha A B + beta C on the 16x16x4 MFMAs (v_mfma_f64_16x16x4f64,
//     v_mfma_f32_16x16x4f32): 64x64 tiles (four real products per complex
//     product) for small shapes, 128x128 tiles in the 3M form for large ones,
//     a streaming kernel for a few right-hand sides;
//   * recursive LU with partial pivoting (row-major) whose panels are
//     factorised in one cooperative launch, LAPACK's pivot choice and
//     interchange vector;
//   * recursive triangular solves with the factors for any number of
//     right-hand sides.
// Replaces LAPACKE_zgetrf/zgetri (solve_bieq2.c:389-399) and zgesv
// (mo_solve.c:751) of the reference.  The FP32 instances serve the
// mixed-precision form (a complex-FP32 factorisation refined in FP64).
#include "bem_dev.h"
#include <mutex>
#include "bem_prof.h"

namespace bem {
int tune_value(int key);   // the tuning table (bem_tune.hip)
}
using bem::tune_value;

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// per-precision types, LDS row padding of the 128x128 kernel (double rows
// are read 16 lanes at a time: 132 keeps them conflict-free; float rows 32
// lanes at a time: 144 = 16 mod 32 puts the two k rows of a half-wave on
// disjoint banks) and the kernel-time registry ids
template <class R> struct Cx;
template <> struct Cx<double> {
  typedef cd Z;
  typedef d4 V;
  static constexpr int ZP = 132;
  static constexpr int KG = BH_K_ZGEMM, KGS = BH_K_ZGEMM_SMALL, KGV = BH_K_ZGEMV, KP = BH_K_LU_PANEL, KT = BH_K_TRSM;
};
template <> struct Cx<float> {
  typedef cf Z;
  typedef f4 V;
  static constexpr int ZP = 144;
  static constexpr int KG = BH_K_CGEMM, KGS = BH_K_CGEMM_SMALL, KGV = BH_K_CGEMV, KP = BH_K_LU_PANEL_C,
                       KT = BH_K_TRSM_C;
};

template <class Z> __host__ __device__ __forceinline__ Z zmk(double a, double b)
{
  Z r;
  r.x = a;
  r.y = b;
  return r;
}

__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
// row (within the 16x16 block) of accumulator register r of lane l; the
// column is l % 16.  f64: l/16 + 4r (measured); f32: 4(l/16) + r.
__device__ __forceinline__ int drow(double, int l, int r) { return (l >> 4) + 4 * r; }
__device__ __forceinline__ int drow(float, int l, int r) { return 4 * (l >> 4) + r; }

// |re| + |im| (LAPACK's izamax measure), compared in double
__device__ __forceinline__ double zabs1(cd x) { return fabs(x.x) + fabs(x.y); }
__device__ __forceinline__ double zabs1(cf x) { return (double)(fabsf(x.x) + fabsf(x.y)); }

#define GB_M 64
#define GB_N 64
#define GB_K 16

// C[m x n] = alpha * A[m x k] B[k x n] + beta * C   (row-major, complex)
template <class R>
__global__ void __launch_bounds__(256) k_gemm64(int m, int n, int k, typename Cx<R>::Z alpha,
                                                const typename Cx<R>::Z *__restrict__ A, size_t lda,
                                                const typename Cx<R>::Z *__restrict__ B, size_t ldb,
                                                typename Cx<R>::Z beta, typename Cx<R>::Z *__restrict__ C, size_t ldc)
{
  typedef typename Cx<R>::Z Z;
  typedef typename Cx<R>::V V;
  __shared__ R sAr[GB_K][GB_M + 1], sAi[GB_K][GB_M + 1];
  __shared__ R sBr[GB_K][GB_N + 1], sBi[GB_K][GB_N + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = blockIdx.y * GB_M, j0 = blockIdx.x * GB_N;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;   // wave sub-tile origin
  V cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) { cr[a][b] = (V){0, 0, 0, 0}; ci[a][b] = (V){0, 0, 0, 0}; }
  for (int k0 = 0; k0 < k; k0 += GB_K) {
    // stage A (64 x 16) and B (16 x 64): 4 complex values per thread each
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int idx = tid + q * 256;
      int ar = idx >> 4, ac = idx & 15;          // A tile row, k column
      int gi = i0 + ar, gk = k0 + ac;
      Z va = (gi < m && gk < k) ? A[(size_t)gi * lda + gk] : zmk<Z>(0, 0);
      sAr[ac][ar] = va.x; sAi[ac][ar] = va.y;
      int br = idx >> 6, bc = idx & 63;          // B tile k row, column
      int gk2 = k0 + br, gj = j0 + bc;
      Z vb = (gk2 < k && gj < n) ? B[(size_t)gk2 * ldb + gj] : zmk<Z>(0, 0);
      sBr[br][bc] = vb.x; sBi[br][bc] = vb.y;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GB_K; kk += 4) {
      int kr = kk + (lane >> 4), cc = lane & 15;
      R ar[2], ai[2], nai[2], br[2], bi[2];
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
          cr[a][b] = mfma4(ar[a], br[b], cr[a][b]);
          cr[a][b] = mfma4(nai[a], bi[b], cr[a][b]);
          ci[a][b] = mfma4(ar[a], bi[b], ci[a][b]);
          ci[a][b] = mfma4(ai[a], br[b], ci[a][b]);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int gi = i0 + wr + 16 * a + drow(R(0), lane, r);
        int gj = j0 + wc + 16 * b + (lane & 15);
        if (gi < m && gj < n) {
          Z ab = zmk<Z>(cr[a][b][r], ci[a][b][r]);
          Z *c = C + (size_t)gi * ldc + gj;
          Z v = alpha * ab;
          if (beta.x != 0 || beta.y != 0) v = v + beta * (*c);
          *c = v;
        }
      }
}

// ---- skinny GEMM (n <= 8 right-hand sides) ----
// Memory-bound products such as G1 N Y in the block solve: each wavefront
// streams 4 rows of A along k (lanes over k, coalesced), the matching rows of
// B (NC contiguous values per k) are reused across the 4 rows, and the
// 4 x NC partial sums are reduced across the wave at the end.
// rows row0 .. row0+3 of C = alpha A B + beta C, one wave
template <class R, int NC>
__device__ __forceinline__ void skinny_rows(int row0, int m, int n, int k, typename Cx<R>::Z alpha,
                                            const typename Cx<R>::Z *__restrict__ A, size_t lda,
                                            const typename Cx<R>::Z *__restrict__ B, size_t ldb,
                                            typename Cx<R>::Z beta, typename Cx<R>::Z *__restrict__ C, size_t ldc)
{
  typedef typename Cx<R>::Z Z;
  const int lane = threadIdx.x & 63;
  if (row0 >= m) return;
  Z acc[4][NC];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int j = 0; j < NC; j++) acc[r][j] = zmk<Z>(0, 0);
  const Z *a0 = A + (size_t)min(row0 + 0, m - 1) * lda, *a1 = A + (size_t)min(row0 + 1, m - 1) * lda;
  const Z *a2 = A + (size_t)min(row0 + 2, m - 1) * lda, *a3 = A + (size_t)min(row0 + 3, m - 1) * lda;
  for (int kk = lane; kk < k; kk += 64) {
    Z b[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) b[j] = j < n ? B[(size_t)kk * ldb + j] : zmk<Z>(0, 0);
    const Z v[4] = {a0[kk], a1[kk], a2[kk], a3[kk]};
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int j = 0; j < NC; j++) acc[r][j] = acc[r][j] + v[r] * b[j];
  }
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int j = 0; j < NC; j++) acc[r][j] = wsum(acc[r][j]);
  const bool bz = beta.x == 0 && beta.y == 0;
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int j = 0; j < NC; j++)
      if (lane == r * NC + j && row0 + r < m && j < n) {
        Z *c = C + (size_t)(row0 + r) * ldc + j;
        Z v = alpha * acc[r][j];
        if (!bz) v = v + beta * (*c);
        *c = v;
      }
}

template <class R, int NC>
__global__ void __launch_bounds__(256) k_gemm_skinny(int m, int n, int k, typename Cx<R>::Z alpha,
                                                     const typename Cx<R>::Z *__restrict__ A, size_t lda,
                                                     const typename Cx<R>::Z *__restrict__ B, size_t ldb,
                                                     typename Cx<R>::Z beta, typename Cx<R>::Z *__restrict__ C,
                                                     size_t ldc)
{
  skinny_rows<R, NC>((blockIdx.x * 4 + (threadIdx.x >> 6)) * 4, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// many complex-FP64 products with 4 right-hand sides in one launch (the
// refinement residuals of all objects and domains): C = A B + beta C with
// B, C row-major k x 4 / m x 4; workgroup b belongs to the job whose range
// of the prefix table `first` holds it (16 rows per workgroup), and each
// job's rows are those of k_gemm_skinny on it alone
struct SkJob {
  const cd *A;
  size_t lda;
  const cd *B;
  cd *C;
  int m, k;
  double beta;
};
__global__ void __launch_bounds__(256) k_gemm_skinny4_batch(const SkJob *__restrict__ jobs, const int *__restrict__ first,
                                                            int nj)
{
  const int j = job_of(first, nj, blockIdx.x);
  const SkJob J = jobs[j];
  skinny_rows<double, 4>(((blockIdx.x - first[j]) * 4 + (threadIdx.x >> 6)) * 4, J.m, 4, J.k, mk(1.0, 0.0), J.A, J.lda,
                         J.B, 4, mk(J.beta, 0.0), J.C, 4);
}

// ---- large-tile GEMM ----
// Complex product in the 3M (Gauss) form: T1 = Ar Br, T2 = Ai Bi,
// T3 = (Ar + Ai)(Br + Bi), Cr = T1 - T2, Ci = T3 - T1 - T2: three real
// 16x16x4 MFMAs per step instead of four (normwise stable, the error bound
// of a real product of the summed operands).
// 128x128 tile per 512-thread workgroup: 8 waves in a 2x4 grid, each wave a
// 64x32 block = 4x2 MFMA tiles of 16x16 (three accumulators each, all in
// VGPRs, two waves per SIMD).  K slices of 8 complex values are
// double-buffered in LDS as six planes (Ar, Ai, Ar+Ai, Br, Bi, Br+Bi; rows
// padded against bank conflicts); the next slice is loaded into registers
// from clamped, always-valid addresses before the current slice's MFMAs and
// masked to zero outside the matrix only when written to LDS, so the load
// latency overlaps the MFMAs.  Tiles are assigned so that each XCD
// (workgroup id mod 8) works on a contiguous band of C, grouped 8 tile-rows
// high, for L2 reuse of the A and B panels.
#define ZT 128
#define ZK 8
template <class R> constexpr int zlds_bytes() { return 2 * 6 * ZK * Cx<R>::ZP * (int)sizeof(R); }

template <class R>
__global__ void __launch_bounds__(512) k_gemm128(int m, int n, int k, typename Cx<R>::Z alpha,
                                                 const typename Cx<R>::Z *__restrict__ A, size_t lda,
                                                 const typename Cx<R>::Z *__restrict__ B, size_t ldb,
                                                 typename Cx<R>::Z beta, typename Cx<R>::Z *__restrict__ C, size_t ldc,
                                                 int tiles_x, int tiles_y)
{
  typedef typename Cx<R>::Z Z;
  typedef typename Cx<R>::V V;
  constexpr int ZP = Cx<R>::ZP;
  extern __shared__ double dyn_lds[];
  R *zsm = (R *)dyn_lds;
  const int T = tiles_x * tiles_y, per = (T + 7) / 8;
  const int L = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (L >= T) return;
  const int GM = 8, gw = GM * tiles_x;
  const int first = (L / gw) * GM, gsz = min(tiles_y - first, GM);
  const int ty = first + (L % gw) % gsz, tx = (L % gw) / gsz;
  const int i0 = ty * ZT, j0 = tx * ZT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 2) * 64, wc = (wave & 3) * 32;
  const int ra = tid >> 2, ca = (tid & 3) * 2;
  const int rb = tid >> 6, cb = (tid & 63) * 2;
  V t1[4][2], t2[4][2], t3[4][2];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) { t1[a][b] = (V){0, 0, 0, 0}; t2[a][b] = (V){0, 0, 0, 0}; t3[a][b] = (V){0, 0, 0, 0}; }
  Z pa[2], pb[2];
  bool oka[2], okb[2];
  const int gi = min(i0 + ra, m - 1);
  const bool rowok = i0 + ra < m;
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 2; q++) {
      int ka = k0 + ca + q, kb2 = k0 + rb, jb = j0 + cb + q;
      pa[q] = A[(size_t)gi * lda + min(ka, k - 1)];
      pb[q] = B[(size_t)min(kb2, k - 1) * ldb + min(jb, n - 1)];
      oka[q] = rowok && ka < k;
      okb[q] = kb2 < k && jb < n;
    }
  };
  const int PL = ZK * ZP;   // one LDS plane
  auto lstore = [&](int s) {
    R *Ar = zsm + s * 6 * PL, *Ai = Ar + PL, *As = Ai + PL, *Br = As + PL, *Bi = Br + PL, *Bs = Bi + PL;
#pragma unroll
    for (int q = 0; q < 2; q++) {
      R ax = oka[q] ? pa[q].x : (R)0, ay = oka[q] ? pa[q].y : (R)0;
      R bx = okb[q] ? pb[q].x : (R)0, by = okb[q] ? pb[q].y : (R)0;
      Ar[(ca + q) * ZP + ra] = ax;
      Ai[(ca + q) * ZP + ra] = ay;
      As[(ca + q) * ZP + ra] = ax + ay;
      Br[rb * ZP + cb + q] = bx;
      Bi[rb * ZP + cb + q] = by;
      Bs[rb * ZP + cb + q] = bx + by;
    }
  };
  gload(0);
  lstore(0);
  __syncthreads();
  int s = 0;
  for (int k0 = 0; k0 < k; k0 += ZK) {
    const bool more = k0 + ZK < k;
    if (more) gload(k0 + ZK);
    const R *Ar = zsm + s * 6 * PL, *Ai = Ar + PL, *As = Ai + PL, *Br = As + PL, *Bi = Br + PL, *Bs = Bi + PL;
#pragma unroll
    for (int ks = 0; ks < ZK; ks += 4) {
      const int kr = ks + (lane >> 4), cc = lane & 15;
      R ar[4], ai[4], as[4], br[2], bi[2], bs[2];
#pragma unroll
      for (int a = 0; a < 4; a++) {
        ar[a] = Ar[kr * ZP + wr + 16 * a + cc];
        ai[a] = Ai[kr * ZP + wr + 16 * a + cc];
        as[a] = As[kr * ZP + wr + 16 * a + cc];
      }
#pragma unroll
      for (int b = 0; b < 2; b++) {
        br[b] = Br[kr * ZP + wc + 16 * b + cc];
        bi[b] = Bi[kr * ZP + wc + 16 * b + cc];
        bs[b] = Bs[kr * ZP + wc + 16 * b + cc];
      }
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          t1[a][b] = mfma4(ar[a], br[b], t1[a][b]);
          t2[a][b] = mfma4(ai[a], bi[b], t2[a][b]);
          t3[a][b] = mfma4(as[a], bs[b], t3[a][b]);
        }
    }
    if (more) lstore(s ^ 1);
    __syncthreads();
    s ^= 1;
  }
  const bool bz = beta.x == 0 && beta.y == 0;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gi2 = i0 + wr + 16 * a + drow(R(0), lane, r);
        const int gj = j0 + wc + 16 * b + (lane & 15);
        if (gi2 < m && gj < n) {
          Z *c = C + (size_t)gi2 * ldc + gj;
          Z v = alpha * zmk<Z>(t1[a][b][r] - t2[a][b][r], t3[a][b][r] - t1[a][b][r] - t2[a][b][r]);
          if (!bz) v = v + beta * (*c);
          *c = v;
        }
      }
}

// ---- large-tile complex-FP32 GEMM on the bf16 MFMAs ----
// The 3M products T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi), each real
// FP32 operand split into three bf16 parts x = x1 + x2 + x3 (x1 = bf16(x),
// x2 = bf16(x - x1), x3 = bf16(x - x1 - x2): 24 significant bits, all of
// FP32's) and each real product formed from the six part products whose
// orders add to at most 4, x1 y1 + x1 y2 + x2 y1 + x2 y2 + x1 y3 + x3 y1,
// accumulated in FP32 by v_mfma_f32_32x32x16_bf16 (the part products are
// exact in FP32; the three dropped ones are below 2^-24 of the product).
// The result carries FP32 rounding error like the FP32-MFMA kernel's, at
// bf16 matrix rates (16x the FP32 MFMA rate; six products per real product).
// 128x128 tile per 512-thread workgroup: 8 waves in a 2x4 grid, each a
// 64x32 block of 2x1 MFMA tiles of 32x32 with three accumulators (96
// accumulator registers: two waves per SIMD, so that one wave's operand
// splitting and waits overlap the other's MFMAs).  A K step of 16 stages 18
// bf16 planes in LDS (A and B, x {r, i, r+i} x 3 parts; 128 rows or columns
// x 16 k each, the two 8-value halves of a row swapped on every fourth row
// so that the 16-byte fragment reads of eight lanes hit distinct banks),
// double-buffered; the next step's operands are loaded into registers
// before the current step's MFMAs and split into the other buffer after
// them.  Tiles are ordered per XCD as in k_gemm128.
// MFMA operand layout (checked on gfx950): lane l holds A[l % 32][8 (l / 32)
// + j] and B[8 (l / 32) + j][l % 32], j < 8; accumulator r of lane l is
// C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
#define BF_T 128
#define BF_K 16
#define BF_PLANE (BF_T * BF_K)                                  // bf16 values per plane
#define BF_LDS_BYTES (2 * 18 * BF_PLANE * (int)sizeof(__bf16))  // two stages of 18 planes

// element offset of (row, k) in a plane: 8-value half h = k / 8 stored at
// half h ^ ((row / 4) & 1)
__device__ __forceinline__ int bf_off(int row, int k)
{
  return row * BF_K + ((((k >> 3) ^ (row >> 2)) & 1) << 3) + (k & 7);
}
// x = x1 + x2 + x3 in bf16 parts, into element j of v[q], v[q+1], v[q+2]
__device__ __forceinline__ void bf_split(float x, bf16x4 *v, int q, int j)
{
  const __bf16 h = (__bf16)x;
  const float r = x - (float)h;
  const __bf16 m = (__bf16)r;
  v[q][j] = h;
  v[q + 1][j] = m;
  v[q + 2][j] = (__bf16)(r - (float)m);
}

__global__ void __launch_bounds__(512, 1) k_cgemm_bf6(int m, int n, int k, cf alpha, const cf *__restrict__ A,
                                                      size_t lda, const cf *__restrict__ B, size_t ldb, cf beta,
                                                      cf *__restrict__ C, size_t ldc, int tiles_x, int tiles_y)
{
  extern __shared__ double dyn_lds[];
  __bf16 *lds = (__bf16 *)dyn_lds;
  const int T = tiles_x * tiles_y, per = (T + 7) / 8;
  const int L = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (L >= T) return;
  const int GM = 8, gw = GM * tiles_x;
  const int first = (L / gw) * GM, gsz = min(tiles_y - first, GM);
  const int ty = first + (L % gw) % gsz, tx = (L % gw) / gsz;
  const int i0 = ty * BF_T, j0 = tx * BF_T;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 2) * 64, wc = (wave & 3) * 32;
  // staging: thread t holds row (A) / column (B) t % 128 and the k quarter t / 128
  const int sr = tid & 127, sq = tid >> 7;
  const int ga = min(i0 + sr, m - 1), gb = min(j0 + sr, n - 1);
  const bool rowok = i0 + sr < m, colok = j0 + sr < n;
  const int so = bf_off(sr, 4 * sq);
  f32x16 acc[3][2];
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[p][a][r] = 0.0f;
  cf pa[4], pb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int kk = min(k0 + 4 * sq + j, k - 1);
      pa[j] = A[(size_t)ga * lda + kk];
      pb[j] = B[(size_t)kk * ldb + gb];
    }
  };
  // the staged operands split into the 18 bf16 planes of stage st
  auto lstore = [&](int st, int k0) {
    __bf16 *base = lds + st * 18 * BF_PLANE + so;
    bf16x4 v[9];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const bool ok = rowok && k0 + 4 * sq + j < k;
      const float x = ok ? pa[j].x : 0.0f, y = ok ? pa[j].y : 0.0f;
      bf_split(x, v, 0, j);
      bf_split(y, v, 3, j);
      bf_split(x + y, v, 6, j);
    }
#pragma unroll
    for (int q = 0; q < 9; q++) *(bf16x4 *)(base + q * BF_PLANE) = v[q];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const bool ok = colok && k0 + 4 * sq + j < k;
      const float x = ok ? pb[j].x : 0.0f, y = ok ? pb[j].y : 0.0f;
      bf_split(x, v, 0, j);
      bf_split(y, v, 3, j);
      bf_split(x + y, v, 6, j);
    }
#pragma unroll
    for (int q = 0; q < 9; q++) *(bf16x4 *)(base + (9 + q) * BF_PLANE) = v[q];
  };
  gload(0);
  lstore(0, 0);
  __syncthreads();
  int st = 0;
  const int fr = lane & 31, fk = 8 * (lane >> 5);
  const int oa0 = bf_off(wr + fr, fk), oa1 = bf_off(wr + 32 + fr, fk), ob = bf_off(wc + fr, fk);
  for (int k0 = 0; k0 < k; k0 += BF_K) {
    if (k0 + BF_K < k) gload(k0 + BF_K);
    const __bf16 *base = lds + st * 18 * BF_PLANE;
#pragma unroll
    for (int p = 0; p < 3; p++) {
      bf16x8 fa[2][3], fb[3];
#pragma unroll
      for (int q = 0; q < 3; q++) {
        fa[0][q] = *(const bf16x8 *)(base + (3 * p + q) * BF_PLANE + oa0);
        fa[1][q] = *(const bf16x8 *)(base + (3 * p + q) * BF_PLANE + oa1);
        fb[q] = *(const bf16x8 *)(base + (9 + 3 * p + q) * BF_PLANE + ob);
      }
      // part pairs of order <= 4, the small ones first
#pragma unroll
      for (int e = 0; e < 6; e++) {
        const int qa = e == 0 ? 0 : e == 1 ? 2 : e == 2 ? 1 : e == 3 ? 0 : e == 4 ? 1 : 0;
        const int qb = e == 0 ? 2 : e == 1 ? 0 : e == 2 ? 1 : e == 3 ? 1 : e == 4 ? 0 : 0;
#pragma unroll
        for (int a = 0; a < 2; a++)
          acc[p][a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a][qa], fb[qb], acc[p][a], 0, 0, 0);
      }
    }
    // the next step's operands into the other stage (not read in this
    // step); after the last step they fill the unused stage
    lstore(st ^ 1, k0 + BF_K);
    __syncthreads();
    st ^= 1;
  }
  const bool bz = beta.x == 0.0f && beta.y == 0.0f;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int gi = i0 + wr + 32 * a + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      const int gj = j0 + wc + fr;
      if (gi < m && gj < n) {
        cf *c = C + (size_t)gi * ldc + gj;
        const float t1 = acc[0][a][r], t2 = acc[1][a][r], t3 = acc[2][a][r];
        cf v = alpha * mkf(t1 - t2, t3 - t1 - t2);
        if (!bz) v = v + beta * (*c);
        *c = v;
      }
    }
}

// ---- LU panel, launch-per-column form ----
// Used where a cooperative launch is not available.  The panel (at most 64
// columns, rows k0..n-1) is factorised in a column-major copy P
// (P[c * ldp + r], r = row - k0) without moving rows: the pivot of column j
// is searched among the rows not yet used as pivots, and used[r] records
// the column at which row r became a pivot (INT_MAX while unused).  One
// launch per column (see k_panel_col).  Rows are interchanged afterwards,
// once per panel, in LAPACK's sequential-swap form.
#define LU_NT 256
#define PERM_LDS_MAX (2 * 16384 * (int)sizeof(int))   // position maps of panels up to 16384 rows

template <class R>
__global__ void __launch_bounds__(256) k_panel_load(const typename Cx<R>::Z *__restrict__ A, size_t lda, int k0, int kb,
                                                    int m, typename Cx<R>::Z *__restrict__ P, size_t ldp,
                                                    int *__restrict__ used)
{
  typedef typename Cx<R>::Z Z;
  __shared__ Z t[64][33];
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

// panel back into A with the panel's row interchanges applied: position t
// of the panel receives physical panel row rowAt[t]
template <class R>
__global__ void __launch_bounds__(256) k_panel_store(typename Cx<R>::Z *__restrict__ A, size_t lda, int k0, int kb, int m,
                                                     const typename Cx<R>::Z *__restrict__ P, size_t ldp,
                                                     const int *__restrict__ rowAt)
{
  typedef typename Cx<R>::Z Z;
  __shared__ Z t[64][33];
  const int r0 = blockIdx.x * 32, tid = threadIdx.x;
  for (int idx = tid; idx < 32 * 64; idx += 256) {
    int c = idx >> 5, rr = idx & 31;
    if (r0 + rr < m && c < kb) t[c][rr] = P[(size_t)c * ldp + rowAt[r0 + rr]];
  }
  __syncthreads();
  for (int idx = tid; idx < 32 * 64; idx += 256) {
    int rr = idx >> 6, c = idx & 63;
    if (r0 + rr < m && c < kb) A[(size_t)(k0 + r0 + rr) * lda + k0 + c] = t[c][rr];
  }
}

// (value, index) maximum with the smallest index on ties, over one wave
__device__ __forceinline__ void wave_argmax(double &v, int &i)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double ov = __shfl_xor(v, o, BEM_WAVE);
    int oi = __shfl_xor(i, o, BEM_WAVE);
    if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
  }
}

// Launch j of the panel: (1) every workgroup reduces the per-workgroup
// maxima that launch j-1 left for column j-1 (the same deterministic result
// everywhere: no atomics, no fences; the kernel boundary orders the data),
// records the pivot and stages the pivot row in LDS; (2) eliminates column
// j-1 on its own rows; (3) leaves its maximum of column j for launch j+1.
// Maxima of consecutive columns alternate between two halves of partV/partI.
template <class R>
__global__ void __launch_bounds__(LU_NT) k_panel_col(typename Cx<R>::Z *__restrict__ P, size_t ldp, int m, int kb, int j,
                                                     int *__restrict__ used, int *__restrict__ pv,
                                                     double *__restrict__ partV, int *__restrict__ partI, int nbmax,
                                                     int *__restrict__ info, int k0)
{
  typedef typename Cx<R>::Z Z;
  __shared__ Z prow[64];
  __shared__ double sv[LU_NT / 64];
  __shared__ int si[LU_NT / 64];
  __shared__ int sp;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, r = blockIdx.x * LU_NT + tid;
  int p = -1;
  if (j > 0) {
    const double *pV = partV + ((j - 1) & 1) * nbmax;
    const int *pI = partI + ((j - 1) & 1) * nbmax;
    double b = -1.0;
    int bi = 0x7fffffff;
    for (int q = tid; q < (int)gridDim.x; q += LU_NT)
      if (pV[q] > b || (pV[q] == b && pI[q] < bi)) { b = pV[q]; bi = pI[q]; }
    wave_argmax(b, bi);
    if (lane == 0) { sv[wv] = b; si[wv] = bi; }
    __syncthreads();
    if (tid == 0) {
      for (int q = 1; q < LU_NT / 64; q++)
        if (sv[q] > b || (sv[q] == b && si[q] < bi)) { b = sv[q]; bi = si[q]; }
      bool bad = bi < 0 || bi >= m;   // never: m - j + 1 >= 1 rows were candidates
      if (bad) bi = m - 1;
      sp = bi;
      if (blockIdx.x == 0) {
        pv[j - 1] = bi;
        if ((bad || !(b > 0.0)) && *info == 0) *info = k0 + j;
      }
    }
    __syncthreads();
    p = sp;
    if (r == p) used[r] = j - 1;   // the owner of the pivot row marks it
    if (tid < kb - j + 1) prow[tid] = P[(size_t)(j - 1 + tid) * ldp + p];
    __syncthreads();
  }
  // rows not yet used as pivots: eliminated for column j-1, candidates for column j
  const bool act = r < m && r != p && used[r] > j - 1;
  Z xj = zmk<Z>(0, 0);   // this row's value in column j after the elimination
  if (act && j < kb) xj = P[(size_t)j * ldp + r];
  if (act && j > 0) {
    const Z a = prow[0];
    if (a.x != 0 || a.y != 0) {
      const Z l = cdiv(P[(size_t)(j - 1) * ldp + r], a);
      P[(size_t)(j - 1) * ldp + r] = l;
      // columns j .. kb-1 in batches of 16 independent loads
      for (int c0 = j; c0 < kb; c0 += 16) {
        Z x[16];
#pragma unroll
        for (int q = 0; q < 16; q++)
          if (c0 + q < kb) x[q] = P[(size_t)(c0 + q) * ldp + r];
#pragma unroll
        for (int q = 0; q < 16; q++)
          if (c0 + q < kb) {
            x[q] = x[q] - l * prow[c0 + q - j + 1];
            P[(size_t)(c0 + q) * ldp + r] = x[q];
          }
        if (c0 == j) xj = x[0];
      }
    }
  }
  if (j >= kb) return;
  // this workgroup's pivot candidate for column j
  double best = -1.0;
  int bidx = 0x7fffffff;
  if (act) {
    best = zabs1(xj);                       // |re|+|im| as LAPACK izamax
    if (!(best >= 0.0)) best = INFINITY;   // a NaN entry still wins over non-candidates
    bidx = r;
  }
  wave_argmax(best, bidx);
  if (lane == 0) { sv[wv] = best; si[wv] = bidx; }
  __syncthreads();
  if (tid == 0) {
    for (int q = 1; q < LU_NT / 64; q++)
      if (sv[q] > best || (sv[q] == best && si[q] < bidx)) { best = sv[q]; bidx = si[q]; }
    partV[(j & 1) * nbmax + blockIdx.x] = best;
    partI[(j & 1) * nbmax + blockIdx.x] = bidx;
  }
}

// Relaxed agent-scope atomic loads of values exchanged through uncached
// memory between workgroups of one launch: no wait is inserted between them
// (a volatile load waits for each one), so a batch costs one round trip.
__device__ inline double ld_rel(const double *p)
{
  return __longlong_as_double((long long)__hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT));
}
__device__ inline int ld_rel(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
// one exchanged complex value
__device__ inline cd ld_x(const cd *p)
{
  const unsigned long long *q = (const unsigned long long *)p;
  const unsigned long long re = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long im = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return zmk<cd>(__longlong_as_double((long long)re), __longlong_as_double((long long)im));
}
__device__ inline cf ld_x(const cf *p)
{
  const unsigned long long u = __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return zmk<cf>(__int_as_float((int)(unsigned int)u), __int_as_float((int)(unsigned int)(u >> 32)));
}

// ---- one-launch panel (cooperative) ----
// The whole panel in one cooperative launch: one 64-lane workgroup per 64
// panel rows, the rows held in LDS (column-major), a grid barrier per
// column.  Column j: every workgroup publishes its best candidate (maximum
// |re|+|im| among its rows not yet used as pivots, smallest row on ties)
// together with that row's entries j..kb-1; after the barrier every
// workgroup reduces the candidates to the same pivot, takes the pivot row
// from the published copy and eliminates column j on its rows.  Each lane
// tracks the position of its row under LAPACK's sequential interchanges
// (interchange j is published by the pivot row's owner and applied by all
// after the next barrier), and the rows are stored once, at their final
// positions, after the last column.  The barrier spins with a bounded count:
// on timeout the error flag is raised and every wave still runs to the end.
#define COOP_SPIN_LIMIT (1 << 24)
#define COOP_GS 16         // workgroups per arrival group of the grid barrier
#define COOP_CSTRIDE 64    // spacing of the arrival counters (unsigned ints: 256 B)
struct CoopWs {
  double *candV;         // [2][nbmax]
  int *candI;            // [2][nbmax]
  void *candRow;         // [2][nbmax][64] complex values of the precision in use
  int *ip;               // [64] interchange of each column (panel positions)
  unsigned int *cnt;     // top counter, then one counter per arrival group
  int *err;
  int nbmax;
};

// Two-level grid barrier on monotonically increasing arrival counts (all
// cleared before the launch).  Barrier e (1, 2, ...) of the launch: each
// workgroup adds 1 to its group's counter; the group's last arriver (the
// one that reads e * groupsize - 1 back) adds 1 to the top counter; the
// barrier completes when the top counter reaches e * ngroups.  Groups of 16
// keep the serialised atomics on any one counter short.  The counters and
// everything exchanged across workgroups live in uncached device memory, so
// visibility needs no cache write-back or invalidation: a workgroup waits
// for its own stores to complete before it arrives, and compiler barriers
// keep the exchange loads after the poll.
__device__ __forceinline__ void coop_barrier(unsigned int *cnt, int *err, unsigned int &epoch, int blk, unsigned int nb)
{
  epoch += 1;
  const unsigned int ng = (nb + COOP_GS - 1) / COOP_GS, g = (unsigned int)blk / COOP_GS;
  const unsigned int gsz = min((unsigned int)COOP_GS, nb - g * COOP_GS);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);   // this wave's stores have completed
  if ((threadIdx.x & 63) == 0) {
    const unsigned int old = __hip_atomic_fetch_add(cnt + COOP_CSTRIDE * (1 + g), 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
    if (old == epoch * gsz - 1) __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int target = epoch * ng;
    int spins = 0;
    for (;;) {   // both loads of a poll issued together
      const unsigned int c = __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int e = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (c >= target || e != 0) break;
      if (++spins > COOP_SPIN_LIMIT) { atomicExch(err, 1); break; }   // after one timeout nobody waits
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// S[c * 64 + lane] -= l sp[c] for c in [c0, kb): eight columns at a time,
// their loads issued before the first update (one LDS round trip per eight
// columns instead of one per column)
template <class Z>
__device__ __forceinline__ void panel_update(Z *S, const Z *sp, Z l, int c0, int kb, int lane)
{
  int c = c0;
  for (; c + 8 <= kb; c += 8) {
    Z a[8], p[8];
#pragma unroll
    for (int u = 0; u < 8; u++) { a[u] = S[(c + u) * 64 + lane]; p[u] = sp[c + u]; }
#pragma unroll
    for (int u = 0; u < 8; u++) S[(c + u) * 64 + lane] = a[u] - l * p[u];
  }
  for (; c < kb; c++) S[c * 64 + lane] = S[c * 64 + lane] - l * sp[c];
}

template <class R>
__global__ void __launch_bounds__(64) k_panel_coop(typename Cx<R>::Z *__restrict__ A, size_t lda, int k0, int kb, int m,
                                                   int *__restrict__ piv, int *__restrict__ info, CoopWs w)
{
  typedef typename Cx<R>::Z Z;
  __shared__ Z S[64 * 64];   // S[c * 64 + lane]: column c of this workgroup's 64 rows
  __shared__ Z sp[64];       // pivot row of the current column
  Z *candRow = (Z *)w.candRow;
  const int lane = threadIdx.x, blk = blockIdx.x, nb = gridDim.x;
  const int r = blk * 64 + lane;
  const bool live = r < m;
  unsigned int epoch = 0;
  for (int c = 0; c < kb; c++) S[c * 64 + lane] = live ? A[(size_t)(k0 + r) * lda + k0 + c] : zmk<Z>(0, 0);
  bool used = false;
  int pos = r;
  for (int j = 0; j < kb; j++) {
    // this workgroup's candidate for column j
    double v = -1.0;
    int vi = 0x7fffffff;
    if (live && !used) {
      v = zabs1(S[j * 64 + lane]);        // |re|+|im| as LAPACK izamax
      if (!(v >= 0.0)) v = INFINITY;      // a NaN entry still wins over non-candidates
      vi = r;
    }
    wave_argmax(v, vi);
    const int buf = j & 1;
    if (lane == 0) { w.candV[buf * w.nbmax + blk] = v; w.candI[buf * w.nbmax + blk] = vi; }
    // the candidate row's entries j..kb-1, one per lane (vi is the same in
    // every lane after the reduction)
    if (v >= 0.0 && lane >= j && lane < kb)
      candRow[((size_t)buf * w.nbmax + blk) * 64 + lane] = S[lane * 64 + (vi - blk * 64)];
    coop_barrier(w.cnt, w.err, epoch, blk, nb);
    // the interchange of column j-1 (published before this barrier) and the
    // pivot: maximum over the workgroups' candidates, loads issued together
    const int cj = j > 0 ? ld_rel(w.ip + j - 1) : 0;
    double b = -1.0;
    int bi = 0x7fffffff, bw = 0;
    for (int q0 = 0; q0 < nb; q0 += 256) {
      double o[4];
      int oi[4];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int q = q0 + 64 * t + lane;
        o[t] = q < nb ? ld_rel(w.candV + buf * w.nbmax + q) : -1.0;
        oi[t] = q < nb ? ld_rel(w.candI + buf * w.nbmax + q) : 0x7fffffff;
      }
#pragma unroll
      for (int t = 0; t < 4; t++)
        if (o[t] > b || (o[t] == b && oi[t] < bi)) { b = o[t]; bi = oi[t]; }
    }
    if (j > 0) {
      if (pos == j - 1) pos = cj; else if (pos == cj) pos = j - 1;
    }
    wave_argmax(b, bi);
    bw = bi / 64;
    if (bi < 0 || bi >= m) { bi = m - 1; bw = bi / 64; if (blk == 0 && lane == 0 && *info == 0) *info = k0 + j + 1; }
    if (blk == 0 && lane == 0 && !(b > 0.0) && *info == 0) *info = k0 + j + 1;
    if (lane >= j && lane < kb) sp[lane] = ld_x(candRow + ((size_t)buf * w.nbmax + bw) * 64 + lane);
    __syncthreads();
    if (r == bi) {
      used = true;
      w.ip[j] = pos;   // LAPACK interchange: column j's position <-> the pivot row's current position
    }
    const Z a = sp[j];
    if (live && !used && (a.x != 0 || a.y != 0)) {
      const Z l = cdiv(S[j * 64 + lane], a);
      S[j * 64 + lane] = l;
      panel_update(S, sp, l, j + 1, kb, lane);
    }
    __syncthreads();
  }
  coop_barrier(w.cnt, w.err, epoch, blk, nb);
  {
    const int cj = ld_rel(w.ip + kb - 1);
    if (pos == kb - 1) pos = cj; else if (pos == cj) pos = kb - 1;
  }
  if (blk == 0 && lane < kb) piv[k0 + lane] = k0 + ld_rel(w.ip + lane);
  // every row to its final position (a permutation of 0..m-1; all reads of
  // the panel happened before the first barrier)
  if (live)
    for (int c = 0; c < kb; c++) A[(size_t)(k0 + pos) * lda + k0 + c] = S[c * 64 + lane];
}

// The complex-FP32 panel without a grid barrier: the same steps as
// k_panel_coop<float> (the same pivots), but each workgroup publishes its
// candidate for column j as one 8-byte word -- |re| + |im| as a float (the
// exact value the FP32 entries give), the row (24 bits) and the tag j + 1
// (8 bits) -- in slot [j & 1][workgroup] (cleared before the launch) after
// its candidate row's stores have completed, and every workgroup reads the
// slots of column j until they carry its tag.  The candidates' arrival is
// then the barrier: a column costs the store of the candidate row, the
// poll of the slots and the load of the pivot row, where the counter
// barrier adds two atomic round trips and a separate load of the
// candidates.  A workgroup runs at most one column ahead of any other (it
// needs everyone's slot of the column before), so two slot and row buffers
// suffice.  Step kb (tag kb + 1) only waits for every workgroup: after it
// nobody reads the panel, and the rows are stored at their final
// positions.  Every wait is bounded (error flag on timeout).
__global__ void __launch_bounds__(64) k_panel_slot(cf *__restrict__ A, size_t lda, int k0, int kb, int m,
                                                   int *__restrict__ piv, int *__restrict__ info, CoopWs w)
{
  typedef cf Z;
  __shared__ Z S[64 * 64];   // S[c * 64 + lane]: column c of this workgroup's 64 rows
  __shared__ Z sp[64];       // pivot row of the current column
  Z *candRow = (Z *)w.candRow;
  unsigned long long *slot = (unsigned long long *)w.candV;   // [2][nbmax]
  const int lane = threadIdx.x, blk = blockIdx.x, nb = gridDim.x;
  const int r = blk * 64 + lane;
  const bool live = r < m;
  for (int c = 0; c < kb; c++) S[c * 64 + lane] = live ? A[(size_t)(k0 + r) * lda + k0 + c] : zmk<Z>(0, 0);
  bool used = false;
  int pos = r;
  for (int j = 0; j <= kb; j++) {
    const int buf = j & 1;
    double v = -1.0;
    int vi = 0xffffff;
    if (j < kb && live && !used) {
      v = zabs1(S[j * 64 + lane]);        // |re|+|im| as LAPACK izamax
      if (!(v >= 0.0)) v = INFINITY;      // a NaN entry still wins over non-candidates
      vi = r;
    }
    wave_argmax(v, vi);
    if (j < kb && v >= 0.0 && lane >= j && lane < kb)
      candRow[((size_t)buf * w.nbmax + blk) * 64 + lane] = S[lane * 64 + (vi - blk * 64)];
    __builtin_amdgcn_s_waitcnt(0);   // the candidate row (and last column's interchange) stores have completed
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (lane == 0) {
      const unsigned int fv = (unsigned int)__float_as_int((float)v);
      const unsigned long long word =
          (unsigned long long)fv | ((unsigned long long)(((unsigned int)vi & 0xffffffu) | ((unsigned int)(j + 1) << 24)) << 32);
      __hip_atomic_store(slot + (size_t)buf * w.nbmax + blk, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // every workgroup's candidate for this column: lane q reads slots q, q + 64, ...
    double b = -1.0;
    int bi = 0x7fffffff;
    for (int q0 = 0; q0 < nb; q0 += 64) {
      const int q = q0 + lane;
      if (q < nb) {
        unsigned long long word = 0;
        int spins = 0;
        for (;;) {
          word = __hip_atomic_load(slot + (size_t)buf * w.nbmax + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if ((int)(word >> 56) == j + 1) break;
          if (++spins > COOP_SPIN_LIMIT || __hip_atomic_load(w.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            atomicExch(w.err, 1);
            word = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        const double fv = (double)__int_as_float((int)(unsigned int)word);
        const int fi = (int)((word >> 32) & 0xffffffu);
        if ((int)(word >> 56) == j + 1 && (fv > b || (fv == b && fi < bi))) { b = fv; bi = fi; }
      }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (j == kb) break;   // the last step only waited for every workgroup
    wave_argmax(b, bi);
    // the interchange of column j-1, published before its owner's slot of this column
    const int cj = j > 0 ? ld_rel(w.ip + j - 1) : 0;
    if (j > 0) {
      if (pos == j - 1) pos = cj; else if (pos == cj) pos = j - 1;
    }
    int bw;
    if (bi < 0 || bi >= m) { bi = m - 1; if (blk == 0 && lane == 0 && *info == 0) *info = k0 + j + 1; }
    bw = bi / 64;
    if (blk == 0 && lane == 0 && !(b > 0.0) && *info == 0) *info = k0 + j + 1;
    if (lane >= j && lane < kb) sp[lane] = ld_x(candRow + ((size_t)buf * w.nbmax + bw) * 64 + lane);
    __syncthreads();
    if (r == bi) {
      used = true;
      w.ip[j] = pos;   // LAPACK interchange: column j's position <-> the pivot row's current position
    }
    const Z a = sp[j];
    if (live && !used && (a.x != 0 || a.y != 0)) {
      const Z l = cdiv(S[j * 64 + lane], a);
      S[j * 64 + lane] = l;
      panel_update(S, sp, l, j + 1, kb, lane);
    }
    __syncthreads();
  }
  {
    const int cj = ld_rel(w.ip + kb - 1);
    if (pos == kb - 1) pos = cj; else if (pos == cj) pos = kb - 1;
  }
  if (blk == 0 && lane < kb) piv[k0 + lane] = k0 + ld_rel(w.ip + lane);
  // every row to its final position (a permutation of 0..m-1; all reads of
  // the panel happened before step kb)
  if (live)
    for (int c = 0; c < kb; c++) A[(size_t)(k0 + pos) * lda + k0 + c] = S[c * 64 + lane];
}

// pivot rows of the panel -> LAPACK interchanges piv[k0+j] (absolute rows)
// and the panel's final row order rowAt[t] (physical panel row at position
// t).  LAPACK swaps position j with the current position of pivot row pv[j];
// the position map and its inverse live in LDS (dynamic, 2 m ints) when
// they fit, else in the global scratch arrays.  One thread performs the kb
// swaps, O(1) each.
__global__ void __launch_bounds__(256) k_panel_perm(const int *__restrict__ pv, int kb, int k0, int m, int *__restrict__ piv,
                                                    int *__restrict__ rowAtG, int *__restrict__ posG, int in_lds)
{
  extern __shared__ int pm[];
  int *rowAt = in_lds ? pm : rowAtG;
  int *pos = in_lds ? pm + m : posG;
  for (int t = threadIdx.x; t < m; t += blockDim.x) { rowAt[t] = t; pos[t] = t; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < kb; j++) {
      const int x = pv[j], cur = pos[x], a = rowAt[j];
      piv[k0 + j] = k0 + cur;
      rowAt[j] = x; pos[x] = j;
      rowAt[cur] = a; pos[a] = cur;
    }
  }
  __syncthreads();
  if (in_lds)
    for (int t = threadIdx.x; t < m; t += blockDim.x) rowAtG[t] = rowAt[t];
}

// apply interchanges piv[k0..k1) to columns [c0, c1) of A (or of a RHS
// block), one column per thread.  The interchanges are read into LDS in
// chunks first, those that change a row compacted to the front of the
// chunk, so the dependent loop over them holds no global loads of piv and
// no iterations without a swap.
#define LASWP_CHUNK 2048
// the interchanges of rows k0..k1-1 in column c (< c1: this thread's
// column), every thread of the 256-thread workgroup taking part
template <class Z>
__device__ __forceinline__ void laswp_body(Z *A, size_t lda, int c, int c1, int k0, int k1, const int *piv)
{
  __shared__ int sj[LASWP_CHUNK], sp[LASWP_CHUNK];
  __shared__ int cnt, wbase[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j0 = k0; j0 < k1; j0 += LASWP_CHUNK) {
    const int len = min(LASWP_CHUNK, k1 - j0);
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    // compaction in order: each wave ballots its swaps, wave results are
    // placed in wave order by a running count
    for (int t0 = 0; t0 < len; t0 += 256) {
      const int t = t0 + threadIdx.x;
      const int p = t < len ? piv[j0 + t] : 0;
      const bool sw = t < len && p != j0 + t;
      const unsigned long long m = __ballot(sw);
      if (lane == 0) wbase[wv] = __popcll(m);
      __syncthreads();
      int base = cnt;
      for (int u = 0; u < wv; u++) base += wbase[u];
      if (sw) {
        const int slot = base + __popcll(m & ((1ull << lane) - 1));
        sj[slot] = j0 + t;
        sp[slot] = p;
      }
      __syncthreads();
      if (threadIdx.x == 0) cnt += wbase[0] + wbase[1] + wbase[2] + wbase[3];
      __syncthreads();
    }
    const int ns = cnt;
    if (c < c1)
      for (int q = 0; q < ns; q++) {
        const int j = sj[q], p = sp[q];
        const Z a = A[(size_t)j * lda + c];
        A[(size_t)j * lda + c] = A[(size_t)p * lda + c];
        A[(size_t)p * lda + c] = a;
      }
    __syncthreads();   // the chunk's tables are reused
  }
}
template <class Z>
__global__ void __launch_bounds__(256) k_laswp(Z *A, size_t lda, int c0, int c1, int k0, int k1, const int *piv)
{
  laswp_body<Z>(A, lda, c0 + blockIdx.x * blockDim.x + threadIdx.x, c1, k0, k1, piv);
}

// Triangular solves with one nb x nb (nb <= 64) diagonal block for ncol
// right-hand-side columns, 64 columns per 256-thread workgroup.  The block
// sits in LDS; a column's rows are spread over 4 neighbouring lanes (lane g
// of the quad owns rows 4i + g, i < 16, in registers), so step q of the
// substitution broadcasts x_q inside the quad with a shuffle and every lane
// updates its rows at once: nb dependent steps of 16 independent complex
// FMAs.  The interleaved ownership makes the four rows a quad reads together
// consecutive rows of the LDS block (distinct banks with the padded stride).
#define TRS_LD 65
template <class R>
__global__ void __launch_bounds__(256) k_trsm_lunit(const typename Cx<R>::Z *__restrict__ L, size_t lda, int nb,
                                                    typename Cx<R>::Z *X, size_t ldb, int ncol)
{
  typedef typename Cx<R>::Z Z;
  extern __shared__ double dyn_lds[];
  Z *sL = (Z *)dyn_lds;   // nb x TRS_LD
  const int tid = threadIdx.x, lane = tid & 63, g = lane & 3;
  const int col = blockIdx.x * 64 + (tid >> 6) * 16 + (lane >> 2);
  for (int idx = tid; idx < nb * nb; idx += 256) sL[(idx / nb) * TRS_LD + idx % nb] = L[(size_t)(idx / nb) * lda + idx % nb];
  Z x[16];
  const bool cok = col < ncol;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = 4 * i + g;
    x[i] = (cok && r < nb) ? X[(size_t)r * ldb + col] : zmk<Z>(0, 0);
  }
  __syncthreads();
  const int qsrc = lane & ~3;
  // branch-free steps: the 16 factor entries of a step are read together
  // (rows clamped into the block) and masked by a select
#pragma unroll
  for (int q = 0; q < 64; q++) {
    if (q < nb) {
      Z lv[16];
#pragma unroll
      for (int i = 0; i < 16; i++) lv[i] = sL[min(4 * i + g, nb - 1) * TRS_LD + q];
      const R vx = __shfl(x[q >> 2].x, qsrc | (q & 3), BEM_WAVE);
      const R vy = __shfl(x[q >> 2].y, qsrc | (q & 3), BEM_WAVE);
      const Z xq = zmk<Z>(vx, vy);
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int r = 4 * i + g;
        const bool on = r > q && r < nb;
        const Z c = zmk<Z>(on ? lv[i].x : (R)0, on ? lv[i].y : (R)0);
        x[i] = x[i] - c * xq;
      }
    }
  }
  if (cok)
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int r = 4 * i + g;
      if (r < nb) X[(size_t)r * ldb + col] = x[i];
    }
}

// one step q of the backward substitution, then the steps q-1 .. 0
template <int Q, class R>
__device__ __forceinline__ void trsm_upper_steps(typename Cx<R>::Z (&x)[16], const typename Cx<R>::Z *sU,
                                                 const typename Cx<R>::Z *dinv, int nb, int g, int qsrc)
{
  typedef typename Cx<R>::Z Z;
  if (Q < nb) {
    Z uv[16];
#pragma unroll
    for (int i = 0; i < 16; i++) uv[i] = sU[min(4 * i + g, nb - 1) * TRS_LD + Q];
    const Z xo = x[Q >> 2] * dinv[Q];   // used by the owner lanes only
    if (g == (Q & 3)) x[Q >> 2] = xo;
    const R vx = __shfl(x[Q >> 2].x, qsrc | (Q & 3), BEM_WAVE);
    const R vy = __shfl(x[Q >> 2].y, qsrc | (Q & 3), BEM_WAVE);
    const Z xq = zmk<Z>(vx, vy);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int r = 4 * i + g;
      const bool on = r < Q;
      const Z c = zmk<Z>(on ? uv[i].x : (R)0, on ? uv[i].y : (R)0);
      x[i] = x[i] - c * xq;
    }
  }
  if constexpr (Q > 0) trsm_upper_steps<Q - 1, R>(x, sU, dinv, nb, g, qsrc);
}

// X := U^{-1} X with U the upper nb x nb block (non-unit diagonal)
template <class R>
__global__ void __launch_bounds__(256) k_trsm_upper(const typename Cx<R>::Z *__restrict__ U, size_t lda, int nb,
                                                    typename Cx<R>::Z *X, size_t ldb, int ncol)
{
  typedef typename Cx<R>::Z Z;
  extern __shared__ double dyn_lds[];
  Z *sU = (Z *)dyn_lds;
  __shared__ Z dinv[64];   // reciprocals of the diagonal
  const int tid = threadIdx.x, lane = tid & 63, g = lane & 3;
  const int col = blockIdx.x * 64 + (tid >> 6) * 16 + (lane >> 2);
  for (int idx = tid; idx < nb * nb; idx += 256) sU[(idx / nb) * TRS_LD + idx % nb] = U[(size_t)(idx / nb) * lda + idx % nb];
  if (tid < nb) dinv[tid] = cdiv(zmk<Z>(1, 0), U[(size_t)tid * lda + tid]);
  Z x[16];
  const bool cok = col < ncol;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = 4 * i + g;
    x[i] = (cok && r < nb) ? X[(size_t)r * ldb + col] : zmk<Z>(0, 0);
  }
  __syncthreads();
  const int qsrc = lane & ~3;
  // the 64 steps are written out through a template so that every x index
  // is a compile-time constant (a runtime index would put x in scratch)
  trsm_upper_steps<63, R>(x, sU, dinv, nb, g, qsrc);
  if (cok)
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int r = 4 * i + g;
      if (r < nb) X[(size_t)r * ldb + col] = x[i];
    }
}

// ---- triangular solves with a few right-hand sides, one launch ----
// X := L^{-1} X (unit lower) or U^{-1} X (upper) for nrhs <= TRSV_MAXR
// columns.  Workgroup b owns one 64-row block (block i = b for L, nb-1-b for
// U, so a block depends only on workgroups launched before it).  With D the
// block's diagonal block, N the block that finishes last (the neighbour) and
// r the right-hand side minus the other finished blocks' terms,
//   x_i = D^{-1} r - W x_nbr,   W = D^{-1} N.
// Before any wait the block inverts D in LDS (wave 0, one lane per column)
// and forms W (four waves, one lane per row).  Its four waves then fold in
// the other finished blocks j (acc -= A_ij x_j) as their flags appear, wave
// w taking every fourth j; wave 0 forms y = D^{-1} r while the neighbour is
// still running, so after the neighbour's flag only y - W x_nbr, one 64-term
// product, lies on the chain from block to block.  The block then publishes
// x, raises its flag and only then writes its rows of B.  x and the flags
// live in uncached memory and are read with relaxed agent-scope atomic loads
// (no wait between them; a volatile load would wait for each one); flag =
// launch epoch, so flags never need clearing; a block waits for its own x
// stores to complete before it raises its flag.  (Polling x itself against
// a sentinel would save the flag's round trips, but a complex FP64 value
// then needs two single-copy-atomic 8-byte stores, and the doubled store
// transactions cost more than the flag.)  Every wait is bounded: on timeout
// the error flag is raised and the launch runs to its end.
#define TRSV_NBMAX 4096    // 64-row blocks (n <= 262144)
#define TRSV_XMAX 65536    // rows of the exchange copy of x
#define TRSV_MAXR 4        // right-hand sides of the one-launch path
struct TrsvWs {
  void *x;                 // [TRSV_XMAX][TRSV_MAXR] complex values of the precision in use
  int *flag;               // [TRSV_NBMAX]
  int *err;
};
// false: the one-launch triangular solves are unavailable (LDS limit), the
// recursive path runs instead
static bool g_trsv_ok = true;

// Per-stream state of the factorisation and the one-launch solves: work on
// different streams (the coupling loop factorises and solves objects on
// several streams at once) must not share exchange areas, flags or epochs.
// A lane is bound to a stream on first use; its uncached areas are
// allocated once and kept until the process ends (memory a kernel has read
// uncached is not returned to the allocator).
#define LA_NLANES 16
struct LaLane {
  hipStream_t st;
  bool used;
  void *lu_ws;           // grow-only work area of the panel factorisation
  size_t lu_bytes;
  void *coop_buf;        // exchange area of the one-launch panel (uncached)
  void *trsv_buf;        // exchange area of the one-launch triangular solves (uncached)
  int trsv_epoch;
  void *trsvb_buf;       // exchange area of the batched one-launch solves (uncached)
  int trsvb_epoch;
  unsigned trsvb_ticket; // the batch area's ticket counter after the launches enqueued so far
};
static LaLane g_lanes[LA_NLANES];
static std::mutex g_lanes_mu;
// the lane of a stream (null if every lane is bound to another stream)
static LaLane *la_lane(hipStream_t st)
{
  std::lock_guard<std::mutex> lk(g_lanes_mu);
  for (int i = 0; i < LA_NLANES; i++)
    if (g_lanes[i].used && g_lanes[i].st == st) return &g_lanes[i];
  for (int i = 0; i < LA_NLANES; i++)
    if (!g_lanes[i].used) {
      g_lanes[i] = LaLane{st, true, nullptr, 0, nullptr, nullptr, 0};
      return &g_lanes[i];
    }
  return nullptr;
}

// LDS of one block: D^{-1} and W (rows padded to 65), an x block per wave
template <class R> constexpr int trsv_lds_bytes()
{
  return (2 * 64 * 65 + 4 * 64 * TRSV_MAXR) * (int)sizeof(typename Cx<R>::Z);
}

// lane 0 waits for flag >= epoch (bounded)
__device__ inline void trsv_wait(const TrsvWs &w, int j, int epoch, int lane)
{
  if (lane == 0) {
    int spins = 0;
    for (;;) {   // both loads of a poll issued together
      const int f = __hip_atomic_load(w.flag + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int e = __hip_atomic_load(w.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (f >= epoch || e != 0) break;
      if (++spins > COOP_SPIN_LIMIT) { atomicExch(w.err, 1); break; }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// the x block of a finished block into LDS: all TRSV_MAXR slots of a row
// are in the exchange area, so every load is issued before the one wait
template <class Z>
__device__ inline void trsv_fetch_x(Z *xw, const Z *xs, int c0, int cols, int lane)
{
  Z xv[TRSV_MAXR];
  if (lane < cols) {
    const Z *src = xs + (size_t)(c0 + lane) * TRSV_MAXR;
#pragma unroll
    for (int c = 0; c < TRSV_MAXR; c++) xv[c] = ld_x(src + c);
  } else {
#pragma unroll
    for (int c = 0; c < TRSV_MAXR; c++) xv[c] = zmk<Z>(0, 0);
  }
#pragma unroll
  for (int c = 0; c < TRSV_MAXR; c++) xw[lane * TRSV_MAXR + c] = xv[c];
}

// acc_c += sum_k a_k x_kc over cols <= 64 columns, the row of A read 16
// values at a time (the loads of a chunk are issued before the first use)
template <class Z, int NR>
__device__ inline void trsv_rowdot(Z *acc, const Z *arow, const Z *xw, int cols)
{
  for (int k0 = 0; k0 < cols; k0 += 16) {
    Z a[16];
    if (k0 + 16 <= cols) {
#pragma unroll
      for (int u = 0; u < 16; u++) a[u] = arow[k0 + u];
    } else {
#pragma unroll
      for (int u = 0; u < 16; u++) a[u] = k0 + u < cols ? arow[k0 + u] : zmk<Z>(0, 0);
    }
#pragma unroll
    for (int u = 0; u < 16; u++)
#pragma unroll
      for (int c = 0; c < NR; c++) acc[c] = acc[c] + a[u] * xw[(k0 + u) * TRSV_MAXR + c];
  }
}

// D (64x64, row stride 65, in LDS) := D^{-1}, one wave, lane = column c:
// row k of X = D^{-1} solves sum_m D_km X_mc = delta_kc from the rows of X
// already formed (above k for L, unit diagonal; below k for U), then
// overwrites row k of D, whose entries that row's sum has consumed.  The
// wave's LDS operations run in program order, so every lane's reads of
// row k precede the writes.  The sum runs in two halves to shorten the
// dependent chain of multiply-adds.
template <class Z, bool UPPER>
__device__ inline void trsv_invert(Z *D, int lane)
{
  for (int t = 0; t < 64; t++) {
    const int k = UPPER ? 63 - t : t;
    const int m0 = UPPER ? k + 1 : 0, m1 = UPPER ? 64 : k;
    Z s0 = zmk<Z>(lane == k ? 1 : 0, 0), s1 = zmk<Z>(0, 0);
    const Z dk = UPPER ? cdiv(zmk<Z>(1, 0), D[k * 65 + k]) : zmk<Z>(1, 0);
    int m = m0;
    for (; m + 1 < m1; m += 2) {
      s0 = s0 - D[k * 65 + m] * D[m * 65 + lane];
      s1 = s1 - D[k * 65 + m + 1] * D[(m + 1) * 65 + lane];
    }
    if (m < m1) s0 = s0 - D[k * 65 + m] * D[m * 65 + lane];
    D[k * 65 + lane] = UPPER ? (s0 + s1) * dk : s0 + s1;
  }
}

// block b of the one-launch solve of one system (w: its exchange rows and
// flags)
template <class R, bool UPPER, int NR>
__device__ __forceinline__ void trsv_body(const typename Cx<R>::Z *__restrict__ A, size_t lda, int n,
                                          typename Cx<R>::Z *B, size_t ldb, TrsvWs w, int epoch, int b)
{
  typedef typename Cx<R>::Z Z;
  extern __shared__ double dyn_lds[];
  Z *Dv = (Z *)dyn_lds;                 // [64][65] the diagonal block, then its inverse
  Z *Wn = Dv + 64 * 65;                 // [64][65] the neighbour block, then W = D^{-1} N
  Z *Xw = Wn + 64 * 65;                 // [4][64][TRSV_MAXR] x block, then partial sums, per wave
  const int nb = (n + 63) / 64;
  const int i = UPPER ? nb - 1 - b : b;
  const int r0 = 64 * i, rows = min(64, n - r0);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, tid = threadIdx.x;
  const bool live = lane < rows;
  Z *xs = (Z *)w.x;
  const int ndep = UPPER ? nb - 1 - i : i;
  const int jlast = UPPER ? i + 1 : i - 1;
  // before any wait: this block's right-hand-side rows (wave 0), and the
  // diagonal and neighbour blocks staged in LDS (coalesced rows; the rows
  // of a partial block padded with identity rows of D and zero rows of N)
  Z bpre[NR];
#pragma unroll
  for (int c = 0; c < NR; c++) bpre[c] = (wv == 0 && live) ? B[(size_t)(r0 + lane) * ldb + c] : zmk<Z>(0, 0);
  {
    const int c0 = 64 * jlast, cols = ndep > 0 ? min(64, n - c0) : 0;
    for (int idx = tid; idx < 64 * 64; idx += 256) {
      const int rr = idx >> 6, kk = idx & 63;
      const bool rok = rr < rows;
      Dv[rr * 65 + kk] = (rok && kk < rows) ? A[(size_t)(r0 + rr) * lda + r0 + kk] : zmk<Z>(rr == kk ? 1 : 0, 0);
      if (ndep > 0) Wn[rr * 65 + kk] = (rok && kk < cols) ? A[(size_t)(r0 + rr) * lda + c0 + kk] : zmk<Z>(0, 0);
    }
  }
  __syncthreads();
  if (wv == 0) trsv_invert<Z, UPPER>(Dv, lane);
  __syncthreads();
  // W = D^{-1} N in place: wave w owns columns 16w..16w+15, lane = row; a
  // wave reads a column group whole before it writes it
  if (ndep > 0) {
    for (int cb = 16 * wv; cb < 16 * wv + 16; cb += 4) {
      Z s[4] = {zmk<Z>(0, 0), zmk<Z>(0, 0), zmk<Z>(0, 0), zmk<Z>(0, 0)};
#pragma unroll 4
      for (int m = 0; m < 64; m++) {
        const Z d = Dv[lane * 65 + m];
#pragma unroll
        for (int u = 0; u < 4; u++) s[u] = s[u] + d * Wn[m * 65 + cb + u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) Wn[lane * 65 + cb + u] = s[u];
    }
  }
  Z acc[NR];
#pragma unroll
  for (int c = 0; c < NR; c++) acc[c] = zmk<Z>(0, 0);
  const Z *arow = A + (size_t)(r0 + (live ? lane : 0)) * lda;
  Z *xw = Xw + wv * 64 * TRSV_MAXR;
  // the other dependencies in the order they finish, wave w every fourth
  for (int q = wv; q < ndep - 1; q += 4) {
    const int j = UPPER ? nb - 1 - q : q;
    trsv_wait(w, j, epoch, lane);
    const int c0 = 64 * j, cols = min(64, n - c0);   // U: the first block j = nb-1 may be partial
    trsv_fetch_x(xw, xs, c0, cols, lane);
    trsv_rowdot<Z, NR>(acc, arow + c0, xw, cols);
  }
  if (wv > 0)
#pragma unroll
    for (int c = 0; c < NR; c++) xw[lane * TRSV_MAXR + c] = acc[c];
  __syncthreads();   // also completes W
  if (wv != 0) return;
  // r = b - (the other blocks' terms), through LDS for the product with D^{-1}
#pragma unroll
  for (int c = 0; c < NR; c++) {
    const Z v = acc[c] + Xw[(64 + lane) * TRSV_MAXR + c] + Xw[(128 + lane) * TRSV_MAXR + c] +
                Xw[(192 + lane) * TRSV_MAXR + c];
    xw[lane * TRSV_MAXR + c] = live ? bpre[c] - v : zmk<Z>(0, 0);
  }
  Z y[NR];
#pragma unroll
  for (int c = 0; c < NR; c++) y[c] = zmk<Z>(0, 0);
#pragma unroll 8
  for (int k = 0; k < 64; k++) {
    const Z d = Dv[lane * 65 + k];
#pragma unroll
    for (int c = 0; c < NR; c++) y[c] = y[c] + d * xw[k * TRSV_MAXR + c];
  }
  // the neighbour: the wait on the critical path
  if (ndep > 0) {
    trsv_wait(w, jlast, epoch, lane);
    const int c0 = 64 * jlast, cols = min(64, n - c0);
    trsv_fetch_x(xw, xs, c0, cols, lane);
#pragma unroll 8
    for (int k = 0; k < 64; k++) {
      const Z a = Wn[lane * 65 + k];
#pragma unroll
      for (int c = 0; c < NR; c++) y[c] = y[c] - a * xw[k * TRSV_MAXR + c];
    }
  }
  if (live) {
#pragma unroll
    for (int c = 0; c < NR; c++) xs[(size_t)(r0 + lane) * TRSV_MAXR + c] = y[c];
  }
  __builtin_amdgcn_s_waitcnt(0);   // this wave's x stores have completed
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  if (lane == 0) __hip_atomic_store(w.flag + i, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // this block's rows of B after the flag: no other block reads them
  if (live) {
#pragma unroll
    for (int c = 0; c < NR; c++) B[(size_t)(r0 + lane) * ldb + c] = y[c];
  }
}

template <class R, bool UPPER, int NR>
__global__ void __launch_bounds__(256) k_trsv_sf(const typename Cx<R>::Z *__restrict__ A, size_t lda, int n,
                                                 typename Cx<R>::Z *B, size_t ldb, TrsvWs w, int epoch)
{
  trsv_body<R, UPPER, NR>(A, lda, n, B, ldb, w, epoch, blockIdx.x);
}

// ---- the one-launch solves of several systems together ----
// N systems (at most TRSVB_MAXSYS, sorted by block count, largest first)
// in one launch: the blocks of all systems interleaved round by round
// (round r: block r of every system that has more than r blocks), so every
// block comes after the blocks it depends on and all chains advance at
// once.  A workgroup takes its position from a ticket counter when it
// starts, not from its workgroup id: a block then waits only for blocks
// whose workgroups are already running or done, whatever order the
// hardware dispatches workgroups in, so the launch needs no co-residency.
// The host passes the counter's value before the launch (base).  System s
// exchanges its x rows at xoff and its flags at foff of the batch area.
#define TRSVB_MAXSYS 32
#define TRSVB_XMAX 262144      // exchange rows of all systems of a launch
#define TRSVB_NBMAX 8192       // blocks of all systems of a launch
template <class Z> struct TrsvbSys {
  const Z *A;
  Z *B;
  size_t lda, ldb;
  int n, nb, xoff, foff;
};
template <class Z> struct TrsvbArgs {
  TrsvbSys<Z> s[TRSVB_MAXSYS];
  int nsys;
  unsigned base;
};

template <class R, bool UPPER, int NR>
__global__ void __launch_bounds__(256) k_trsv_batch(TrsvbArgs<typename Cx<R>::Z> a, void *xbase, int *flags, int *err,
                                                    unsigned *ticket, int epoch)
{
  typedef typename Cx<R>::Z Z;
  __shared__ int s_pos;
  if (threadIdx.x == 0) s_pos = (int)(atomicAdd(ticket, 1u) - a.base);
  __syncthreads();
  const int p = s_pos;
  // the round r of position p: P(r) <= p < P(r + 1), P(r) = sum_s min(nb_s, r)
  // (the argument table is read at constant indices only)
  int lo = 0, hi = a.s[0].nb;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    int P = 0;
#pragma unroll
    for (int q = 0; q < TRSVB_MAXSYS; q++)
      if (q < a.nsys) P += min(a.s[q].nb, mid);
    if (P <= p) lo = mid; else hi = mid;
  }
  int P = 0, cnt = 0;
#pragma unroll
  for (int q = 0; q < TRSVB_MAXSYS; q++)
    if (q < a.nsys) {
      P += min(a.s[q].nb, lo);
      cnt += a.s[q].nb > lo ? 1 : 0;
    }
  const int sy = p - P;   // the systems with more than lo blocks are the first ones
  if (p < 0 || sy >= cnt) {   // a ticket outside this launch: the counter is off
    if (threadIdx.x == 0) atomicExch(err, 1);
    return;
  }
  TrsvbSys<Z> y = a.s[0];
#pragma unroll
  for (int q = 1; q < TRSVB_MAXSYS; q++)
    if (q == sy) y = a.s[q];
  TrsvWs w;
  w.x = (void *)((Z *)xbase + (size_t)y.xoff * TRSV_MAXR);
  w.flag = flags + y.foff;
  w.err = err;
  trsv_body<R, UPPER, NR>(y.A, y.lda, y.n, y.B, y.ldb, w, epoch, lo);
}

// row interchanges piv[0..n) of the first nc columns of N right-hand sides
// (workgroup s: system s), in order, one thread per column
template <class Z> struct LaswpSys {
  Z *B;
  const int *piv;
  size_t ldb;
  int n, pad;
};
template <class Z> struct LaswpArgs {
  LaswpSys<Z> s[TRSVB_MAXSYS];
};
template <class Z> __global__ void __launch_bounds__(256) k_laswp_batch(LaswpArgs<Z> a, int nc)
{
  LaswpSys<Z> q = a.s[0];
#pragma unroll
  for (int k = 1; k < TRSVB_MAXSYS; k++)
    if ((int)blockIdx.x == k) q = a.s[k];
  laswp_body<Z>(q.B, q.ldb, threadIdx.x, nc, 0, q.n, q.piv);
}

__global__ void k_zero_bytes(unsigned int *p, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ---- host side ----
static bool g_big_tiles = false;   // set once the 128x128 kernels' LDS limits are raised
static bool g_coop = false;        // the device supports cooperative launches
// exchange area of the one-launch panel: up to COOP_NBMAX workgroups
// (panels of up to 65536 rows)
#define COOP_NBMAX 1024
#define COOP_CNT_BYTES ((size_t)COOP_CSTRIDE * sizeof(unsigned int) * (1 + COOP_NBMAX / COOP_GS))
#define COOP_BYTES (COOP_CNT_BYTES + 256 + 2 * COOP_NBMAX * (sizeof(double) + sizeof(int)) + 64 * sizeof(int) + \
                    2 * (size_t)COOP_NBMAX * 64 * sizeof(cd))
template <class R> static int la_init_t()
{
  // the triangular-solve kernels stage a 64x64 factor block in LDS, the
  // 128x128 GEMM its double-buffered planes: above the default dynamic limit
  const int bytes = 64 * TRS_LD * (int)sizeof(typename Cx<R>::Z);
  if (hipFuncSetAttribute((const void *)k_trsm_lunit<R>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess) return -1;
  if (hipFuncSetAttribute((const void *)k_trsm_upper<R>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess) return -1;
  if (hipFuncSetAttribute((const void *)k_gemm128<R>, hipFuncAttributeMaxDynamicSharedMemorySize, zlds_bytes<R>()) != hipSuccess)
    return -1;
  const void *sf[8] = {(const void *)k_trsv_sf<R, false, 1>, (const void *)k_trsv_sf<R, false, 2>,
                       (const void *)k_trsv_sf<R, false, 3>, (const void *)k_trsv_sf<R, false, 4>,
                       (const void *)k_trsv_sf<R, true, 1>,  (const void *)k_trsv_sf<R, true, 2>,
                       (const void *)k_trsv_sf<R, true, 3>,  (const void *)k_trsv_sf<R, true, 4>};
  const void *sb[8] = {(const void *)k_trsv_batch<R, false, 1>, (const void *)k_trsv_batch<R, false, 2>,
                       (const void *)k_trsv_batch<R, false, 3>, (const void *)k_trsv_batch<R, false, 4>,
                       (const void *)k_trsv_batch<R, true, 1>,  (const void *)k_trsv_batch<R, true, 2>,
                       (const void *)k_trsv_batch<R, true, 3>,  (const void *)k_trsv_batch<R, true, 4>};
  for (int q = 0; q < 8; q++)
    if (hipFuncSetAttribute(sf[q], hipFuncAttributeMaxDynamicSharedMemorySize, trsv_lds_bytes<R>()) != hipSuccess ||
        hipFuncSetAttribute(sb[q], hipFuncAttributeMaxDynamicSharedMemorySize, trsv_lds_bytes<R>()) != hipSuccess) {
      (void)hipGetLastError();
      g_trsv_ok = false;
    }
  return 0;
}

// the bf16-split kernel's LDS limit is raised (the tuning table's gemm_bf6
// entry chooses it for complex-FP32 products on 128x128 tiles)
static bool g_gemm_bf6 = true;
// the kernel bh_gemm_time forces on the calling thread (-1: the shape and
// the tuning table decide; 0 64x64 tiles, 1 128x128 tiles, 2 bf16-split)
static thread_local int g_gemm_force = -1;

template <class R>
static void gemm_t(int m, int n, int k, typename Cx<R>::Z alpha, const typename Cx<R>::Z *A, size_t lda,
                   const typename Cx<R>::Z *B, size_t ldb, typename Cx<R>::Z beta, typename Cx<R>::Z *C, size_t ldc,
                   hipStream_t st)
{
  if (m <= 0 || n <= 0) return;
  const int tx = (n + ZT - 1) / ZT, ty = (m + ZT - 1) / ZT, T = tx * ty;
  // the 128x128 kernel takes products of at least the tuning table's tile
  // count (defaults measured on gfx950: about 160 tiles in both
  // precisions), and with a long k (the products of many right-hand sides,
  // m x 256 x m) already from fewer (64), where its double-buffered 3M
  // slices outrun the 64x64 kernel even with part of the CUs idle
  const int force = g_gemm_force;
  const bool skinny = force < 0 && n <= 8 && k >= 256;
  const int tiles = tune_value(sizeof(R) == sizeof(double) ? BH_TUNE_GEMM_T_D : BH_TUNE_GEMM_T_F);
  const bool big = g_big_tiles && !skinny &&
                   (force >= 0 ? force >= 1
                               : ((T >= tiles && k >= 32) ||
                                  (T >= tune_value(BH_TUNE_GEMM_LONGK_T) && k >= tune_value(BH_TUNE_GEMM_LONGK_K))));
  const bool bf6 = g_gemm_bf6 && (force >= 0 ? force == 2 : tune_value(BH_TUNE_GEMM_BF6) != 0);
  const int kid = big ? Cx<R>::KG : skinny ? Cx<R>::KGV : Cx<R>::KGS;
  const bool bz = beta.x == 0 && beta.y == 0;
  bem_prof_begin_b(kid, 8.0 * m * n * k,
                   ((double)m * k + (double)k * n + (double)m * n * (bz ? 1.0 : 2.0)) * sizeof(typename Cx<R>::Z), st);
  if (skinny) {
    // a few right-hand sides against a long k: stream A once
    dim3 grid((m + 15) / 16);
    if (n == 1)
      hipLaunchKernelGGL((k_gemm_skinny<R, 1>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (n == 2)
      hipLaunchKernelGGL((k_gemm_skinny<R, 2>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (n <= 4)
      hipLaunchKernelGGL((k_gemm_skinny<R, 4>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else
      hipLaunchKernelGGL((k_gemm_skinny<R, 8>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  } else if (big && sizeof(R) == sizeof(float) && bf6) {
    hipLaunchKernelGGL(k_cgemm_bf6, dim3(8 * ((T + 7) / 8)), dim3(512), BF_LDS_BYTES, st, m, n, k, *(const cf *)&alpha,
                       (const cf *)A, lda, (const cf *)B, ldb, *(const cf *)&beta, (cf *)C, ldc, tx, ty);
  } else if (big) {
    // at least one 128x128 tile per CU
    hipLaunchKernelGGL(k_gemm128<R>, dim3(8 * ((T + 7) / 8)), dim3(512), zlds_bytes<R>(), st, m, n, k, alpha, A, lda, B,
                       ldb, beta, C, ldc, tx, ty);
  } else {
    dim3 grid((n + GB_N - 1) / GB_N, (m + GB_M - 1) / GB_M);
    hipLaunchKernelGGL(k_gemm64<R>, grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  }
  bem_prof_end(kid, st);
}


// workspace of the panel factorisation for one matrix order
struct LuWs {
  void *P;   // column-major panel copy (complex values of the precision in use)
  size_t ldp;
  int *used, *pv, *partI, *info, *rowAt, *pos;
  double *partV;
  int nbmax;
  CoopWs coop;
  bool coop_ok;
  bool ordinary;   // the panels of this factorisation take ordinary launches
};

// split point of a recursive step: about half, a multiple of the 64-wide base
static inline int half64(int n) { return std::max(64, (n / 2) / 64 * 64); }

// columns [k0, k0+kb) (kb <= 64) of rows [k0, n): pivots into piv[k0..k0+kb)
// (absolute rows), rows interchanged inside these columns only
// Resident workgroups of k_panel_coop<float> on the device (0: unknown),
// a k_panel_coop<double> workgroup's share of them in units of the float
// one, and the units held by factorisations running now.  A factorisation
// reserves the residency of its largest panel and takes ordinary launches
// for its panels; while the factorisations running on other streams hold
// too much of the device it waits for them (each releases its share when
// it returns, without waiting for anyone).  Only a factorisation whose
// panel alone exceeds the device takes cooperative launches, which run
// alone on the device (BEM3_HIP_PANEL_COOP=1: always).
static int g_panel_cap = 0, g_panel_unit_d = 2;
static std::atomic<long> g_panel_held{0};
static bool g_panel_ordinary = true;
static bool g_panel_slot = true;   // BEM3_HIP_PANEL_SLOT=0: complex-FP32 panels by the counter barrier
template <class R> static long panel_units(int nbc) { return (long)nbc * (sizeof(R) == sizeof(float) ? 1 : g_panel_unit_d); }
template <class R> static bool panel_reserve(int nbc)
{
  if (!g_panel_ordinary || g_panel_cap <= 0) return false;
  const long u = panel_units<R>(nbc);
  if (u > (long)g_panel_cap) return false;
  for (int k = 0;; k++) {
    long cur = g_panel_held.load();
    while (cur + u <= (long)g_panel_cap)
      if (g_panel_held.compare_exchange_weak(cur, cur + u)) return true;
    std::this_thread::sleep_for(std::chrono::microseconds(k < 200 ? 20 : 200));
  }
}
template <class R> static void panel_release(int nbc) { g_panel_held -= panel_units<R>(nbc); }

template <class R>
static void lu_panel(typename Cx<R>::Z *A, size_t lda, int n, int k0, int kb, int *piv, const LuWs &w, hipStream_t st)
{
  typedef typename Cx<R>::Z Z;
  const int m = n - k0, gb = (m + LU_NT - 1) / LU_NT;
  // about 8 m kb^2 / 2 real flops of rank-1 updates
  bem_prof_begin(Cx<R>::KP, 4.0 * (double)m * kb * kb, st);
  const int nbc = (m + 63) / 64;
  if (w.coop_ok && nbc <= w.coop.nbmax) {
    CoopWs cw = w.coop;
    // arrival counters of this launch (the counter barrier): the top one
    // and one per group
    if (!(sizeof(R) == sizeof(float) && g_panel_slot))
      (void)hipMemsetAsync(cw.cnt, 0, (size_t)COOP_CSTRIDE * sizeof(unsigned int) * (1 + (nbc + COOP_GS - 1) / COOP_GS), st);
    // An ordinary launch lets the work of other streams (other objects'
    // products and panels) run beside the panel; a cooperative launch runs
    // alone on the device.  The grid barrier needs every workgroup of the
    // panel resident: that holds once the other kernels finish if the panels
    // of all streams that factorise at once fit the device together
    // (lanes x workgroups <= resident capacity), the only other waiters
    // being the triangular solves' blocks, which always progress.
    // complex FP32: the slot form (k_panel_slot), its slots cleared first
    const void *fn = (const void *)k_panel_coop<R>;
    if (sizeof(R) == sizeof(float) && g_panel_slot) {
      fn = (const void *)k_panel_slot;
      (void)hipMemsetAsync(cw.candV, 0, 2 * (size_t)cw.nbmax * sizeof(double), st);
    }
    if (w.ordinary) {
      if (fn == (const void *)k_panel_slot)
        hipLaunchKernelGGL(k_panel_slot, dim3(nbc), dim3(64), 0, st, (cf *)A, lda, k0, kb, m, piv, w.info, cw);
      else
        hipLaunchKernelGGL(k_panel_coop<R>, dim3(nbc), dim3(64), 0, st, A, lda, k0, kb, m, piv, w.info, cw);
      if (hipGetLastError() == hipSuccess) {
        bem_prof_end(Cx<R>::KP, st);
        return;
      }
    }
    void *args[] = {&A, &lda, (void *)&k0, (void *)&kb, (void *)&m, &piv, (void *)&w.info, &cw};
    if (hipLaunchCooperativeKernel(fn, dim3(nbc), dim3(64), args, 0, st) == hipSuccess) {
      bem_prof_end(Cx<R>::KP, st);
      return;
    }
    (void)hipGetLastError();   // not co-resident: the launch-per-column path below
  }
  Z *P = (Z *)w.P;
  hipLaunchKernelGGL(k_panel_load<R>, dim3((m + 31) / 32), dim3(256), 0, st, A, lda, k0, kb, m, P, w.ldp, w.used);
  for (int j = 0; j <= kb; j++)
    hipLaunchKernelGGL(k_panel_col<R>, dim3(gb), dim3(LU_NT), 0, st, P, w.ldp, m, kb, j, w.used, w.pv, w.partV, w.partI,
                       w.nbmax, w.info, k0);
  const int in_lds = 2 * (size_t)m * sizeof(int) <= (size_t)PERM_LDS_MAX;
  hipLaunchKernelGGL(k_panel_perm, dim3(1), dim3(256), in_lds ? 2 * (size_t)m * sizeof(int) : 0, st, w.pv, kb, k0, m, piv,
                     w.rowAt, w.pos, in_lds);
  hipLaunchKernelGGL(k_panel_store<R>, dim3((m + 31) / 32), dim3(256), 0, st, A, lda, k0, kb, m, P, w.ldp, w.rowAt);
  bem_prof_end(Cx<R>::KP, st);
}

// X := L^{-1} X with L the nr x nr unit-lower block at L (stride lda) and X
// nr x ncol at B (stride ldb); recursive halving so that the off-diagonal
// work is large-K GEMM, 64x64 diagonal blocks solved in LDS
template <class R>
static void trsm_lunit_rec(const typename Cx<R>::Z *L, size_t lda, int nr, typename Cx<R>::Z *B, size_t ldb, int ncol,
                           hipStream_t st)
{
  typedef typename Cx<R>::Z Z;
  if (nr <= 64) {
    size_t shm = (size_t)nr * TRS_LD * sizeof(Z);
    bem_prof_begin(Cx<R>::KT, 4.0 * nr * nr * (double)ncol, st);
    hipLaunchKernelGGL(k_trsm_lunit<R>, dim3((ncol + 63) / 64), dim3(256), shm, st, L, lda, nr, B, ldb, ncol);
    bem_prof_end(Cx<R>::KT, st);
    return;
  }
  const int h = half64(nr);
  trsm_lunit_rec<R>(L, lda, h, B, ldb, ncol, st);
  gemm_t<R>(nr - h, ncol, h, zmk<Z>(-1, 0), L + (size_t)h * lda, lda, B, ldb, zmk<Z>(1, 0), B + (size_t)h * ldb, ldb, st);
  trsm_lunit_rec<R>(L + (size_t)h * lda + h, lda, nr - h, B + (size_t)h * ldb, ldb, ncol, st);
}

// X := U^{-1} X with U the nr x nr upper block (non-unit diagonal)
template <class R>
static void trsm_upper_rec(const typename Cx<R>::Z *U, size_t lda, int nr, typename Cx<R>::Z *B, size_t ldb, int ncol,
                           hipStream_t st)
{
  typedef typename Cx<R>::Z Z;
  if (nr <= 64) {
    size_t shm = (size_t)nr * TRS_LD * sizeof(Z);
    bem_prof_begin(Cx<R>::KT, 4.0 * nr * nr * (double)ncol, st);
    hipLaunchKernelGGL(k_trsm_upper<R>, dim3((ncol + 63) / 64), dim3(256), shm, st, U, lda, nr, B, ldb, ncol);
    bem_prof_end(Cx<R>::KT, st);
    return;
  }
  const int h = half64(nr);
  trsm_upper_rec<R>(U + (size_t)h * lda + h, lda, nr - h, B + (size_t)h * ldb, ldb, ncol, st);
  gemm_t<R>(h, ncol, nr - h, zmk<Z>(-1, 0), U + h, lda, B + (size_t)h * ldb, ldb, zmk<Z>(1, 0), B, ldb, st);
  trsm_upper_rec<R>(U, lda, h, B, ldb, ncol, st);
}

// recursive LU (left half, then the right half after its update) of columns
// [k0, k0+nc) over rows [k0, n); every interchange found below reaches every
// column of the block through the two laswp calls
template <class R>
static void lu_rec(typename Cx<R>::Z *A, size_t lda, int n, int k0, int nc, int *piv, const LuWs &w, hipStream_t st)
{
  typedef typename Cx<R>::Z Z;
  if (nc <= 64) {
    lu_panel<R>(A, lda, n, k0, nc, piv, w, st);
    return;
  }
  const int h = half64(nc);
  lu_rec<R>(A, lda, n, k0, h, piv, w, st);
  hipLaunchKernelGGL(k_laswp<Z>, dim3((nc - h + 255) / 256), dim3(256), 0, st, A, lda, k0 + h, k0 + nc, k0, k0 + h, piv);
  trsm_lunit_rec<R>(A + (size_t)k0 * lda + k0, lda, h, A + (size_t)k0 * lda + k0 + h, lda, nc - h, st);
  gemm_t<R>(n - k0 - h, nc - h, h, zmk<Z>(-1, 0), A + (size_t)(k0 + h) * lda + k0, lda, A + (size_t)k0 * lda + k0 + h, lda,
            zmk<Z>(1, 0), A + (size_t)(k0 + h) * lda + k0 + h, lda, st);
  lu_rec<R>(A, lda, n, k0 + h, nc - h, piv, w, st);
  hipLaunchKernelGGL(k_laswp<Z>, dim3((h + 255) / 256), dim3(256), 0, st, A, lda, k0, k0 + h, k0 + h, k0 + nc, piv);
}

// LU factorisation in place, LAPACK's pivot choice and interchange vector.
// piv (device, n ints) receives 0-based pivot rows; info (device int)
// receives 0 or the 1-based index of the first zero pivot.  Returns 0, -1 if
// the scratch cannot be allocated, -2 if a panel barrier timed out.
template <class R> static int getrf_t(int n, typename Cx<R>::Z *A, size_t lda, int *piv, int *info, hipStream_t st)
{
  const int kbmax = 64;
  size_t ldp = ((size_t)n + 63) / 64 * 64;
  int nblk = (n + LU_NT - 1) / LU_NT;
  const int nbc = (n + 63) / 64;   // workgroups of the one-launch panel
  size_t need = (size_t)kbmax * ldp * sizeof(cd) + 3 * ldp * sizeof(int) + 64 * sizeof(int) +
                2 * (size_t)nblk * (sizeof(double) + sizeof(int));
  LaLane *L = la_lane(st);
  if (!L) return -1;
  if (need > L->lu_bytes) {
    if (L->lu_ws) {
      (void)hipStreamSynchronize(st);   // earlier factorisations on this stream may still use it
      (void)hipFree(L->lu_ws);
    }
    L->lu_ws = nullptr;
    L->lu_bytes = 0;
    if (hipMalloc(&L->lu_ws, need) != hipSuccess) return -1;
    L->lu_bytes = need;
  }
  // layout for this n (the buffer may be larger, from an earlier call);
  // nothing in it needs clearing: the panel loader initialises used[], every
  // launch writes the maxima the next one reads
  char *p = (char *)L->lu_ws;
  LuWs w;
  w.P = (void *)p;                       p += (size_t)kbmax * ldp * sizeof(cd);
  w.ldp = ldp;
  w.used = (int *)p;                     p += ldp * sizeof(int);
  w.rowAt = (int *)p;                    p += ldp * sizeof(int);
  w.pos = (int *)p;                      p += ldp * sizeof(int);
  w.pv = (int *)p;                       p += 64 * sizeof(int);
  w.partV = (double *)p;                 p += 2 * (size_t)nblk * sizeof(double);
  w.partI = (int *)p;                    p += 2 * (size_t)nblk * sizeof(int);
  w.nbmax = nblk;
  w.info = info;
  // one-launch panel: exchange area in uncached memory (allocated once for
  // the life of the process, never freed: memory a kernel has read uncached
  // is not returned to the allocator)
  w.coop_ok = false;
  if (g_coop && nbc <= COOP_NBMAX) {
    if (!L->coop_buf && hipExtMallocWithFlags(&L->coop_buf, COOP_BYTES, hipDeviceMallocUncached) != hipSuccess)
      L->coop_buf = nullptr;   // this stream takes the launch-per-column path
    if (L->coop_buf) {
      char *q = (char *)L->coop_buf;
      w.coop.cnt = (unsigned int *)q;        q += COOP_CNT_BYTES;
      w.coop.err = (int *)q;                 q += 256;
      w.coop.candV = (double *)q;            q += 2 * COOP_NBMAX * sizeof(double);
      w.coop.candI = (int *)q;               q += 2 * COOP_NBMAX * sizeof(int);
      w.coop.ip = (int *)q;                  q += 64 * sizeof(int);
      w.coop.candRow = (void *)q;
      w.coop.nbmax = COOP_NBMAX;
      w.coop_ok = true;
    }
  }
  (void)hipMemsetAsync(info, 0, sizeof(int), st);
  if (w.coop_ok) (void)hipMemsetAsync(w.coop.err, 0, sizeof(int), st);   // error flag of all panels
  // the largest panel is the first one (nbc workgroups): reserved until the
  // factorisation has finished (the stream is waited for below)
  w.ordinary = w.coop_ok && panel_reserve<R>(nbc);
  if (n > 0) lu_rec<R>(A, lda, n, 0, n, piv, w, st);
  if (w.coop_ok) {
    int herr = 0;
    const bool ok = hipMemcpyAsync(&herr, w.coop.err, sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
                    hipStreamSynchronize(st) == hipSuccess;
    if (w.ordinary) panel_release<R>(nbc);
    if (!ok || herr != 0) return -2;   // a panel barrier timed out
  }
  return 0;
}

// solve A X = B with the factors of getrf_t; B is n x nrhs (row stride ldb)
#define TRSV_XBYTES ((size_t)TRSV_XMAX * TRSV_MAXR * sizeof(cd))
#define TRSV_BYTES (TRSV_XBYTES + TRSV_NBMAX * sizeof(int) + 256)

// the lane's exchange area of the one-launch solves (its error flag is
// shared with the batched solves); false if it cannot be allocated
static bool trsv_area(LaLane *L, hipStream_t st)
{
  if (L->trsv_buf) return true;
  if (hipExtMallocWithFlags(&L->trsv_buf, TRSV_BYTES, hipDeviceMallocUncached) != hipSuccess) {
    L->trsv_buf = nullptr;
    return false;
  }
  // flags start below every epoch (a kernel clears them: no memset node)
  hipLaunchKernelGGL(k_zero_bytes, dim3((TRSV_BYTES / 4 + 255) / 256), dim3(256), 0, st, (unsigned int *)L->trsv_buf,
                     TRSV_BYTES / 4);
  return true;
}

// one-launch solve with the factors for nrhs <= TRSV_MAXR; false if not possible
// (then the recursive path runs)
template <class R, bool UPPER>
static bool trsv_batch(int N, const int *n, int nrhs, const typename Cx<R>::Z *const *A, const size_t *lda,
                       typename Cx<R>::Z *const *B, const size_t *ldb, hipStream_t st);
static bool trsv_batch_fits(int N, const int *n, int nrhs);

// one-launch solve with the factors for nrhs <= TRSV_MAXR (the batched
// launch with one system: ticket-ordered blocks, an ordinary launch that
// work on other streams can overlap); false if not possible (then the
// recursive path runs)
template <class R, bool UPPER>
static bool trsv_sf(const typename Cx<R>::Z *A, size_t lda, int n, typename Cx<R>::Z *B, size_t ldb, int nrhs,
                    hipStream_t st)
{
  if (!trsv_batch_fits(1, &n, nrhs)) return false;
  return trsv_batch<R, UPPER>(1, &n, nrhs, &A, &lda, &B, &ldb, st);
}

#define TRSVB_XBYTES ((size_t)TRSVB_XMAX * TRSV_MAXR * sizeof(cd))
#define TRSVB_BYTES (TRSVB_XBYTES + TRSVB_NBMAX * sizeof(int) + 256)

// whether N systems of orders n[] fit one batched launch
static bool trsv_batch_fits(int N, const int *n, int nrhs)
{
  if (!g_trsv_ok || N <= 0 || N > TRSVB_MAXSYS || nrhs < 1 || nrhs > TRSV_MAXR) return false;
  size_t rows = 0, blocks = 0;
  for (int q = 0; q < N; q++) {
    if (n[q] <= 0) return false;
    rows += (size_t)n[q];
    blocks += (size_t)((n[q] + 63) / 64);
  }
  return rows <= TRSVB_XMAX && blocks <= TRSVB_NBMAX;
}

// X_s := L_s^{-1} X_s (unit lower) or U_s^{-1} X_s of N systems in one
// launch (k_trsv_batch); the systems must fit (trsv_batch_fits).  False if
// the exchange area or the launch fails.
template <class R, bool UPPER>
static bool trsv_batch(int N, const int *n, int nrhs, const typename Cx<R>::Z *const *A, const size_t *lda,
                       typename Cx<R>::Z *const *B, const size_t *ldb, hipStream_t st)
{
  typedef typename Cx<R>::Z Z;
  LaLane *L = la_lane(st);
  if (!L || !trsv_area(L, st)) return false;
  if (!L->trsvb_buf) {
    if (hipExtMallocWithFlags(&L->trsvb_buf, TRSVB_BYTES, hipDeviceMallocUncached) != hipSuccess) {
      L->trsvb_buf = nullptr;
      return false;
    }
    hipLaunchKernelGGL(k_zero_bytes, dim3((TRSVB_BYTES / 4 + 255) / 256), dim3(256), 0, st, (unsigned int *)L->trsvb_buf,
                       TRSVB_BYTES / 4);
    L->trsvb_ticket = 0;
  }
  std::vector<int> order(N);
  for (int q = 0; q < N; q++) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
  TrsvbArgs<Z> a{};
  int xoff = 0, foff = 0;
  double work = 0.0;
  for (int q = 0; q < N; q++) {
    const int k = order[q];
    TrsvbSys<Z> &y = a.s[q];
    y.A = A[k]; y.B = B[k]; y.lda = lda[k]; y.ldb = ldb[k];
    y.n = n[k]; y.nb = (n[k] + 63) / 64; y.xoff = xoff; y.foff = foff;
    xoff += y.n;
    foff += y.nb;
    work += 4.0 * (double)n[k] * n[k] * nrhs;
  }
  a.nsys = N;
  a.base = L->trsvb_ticket;
  char *p = (char *)L->trsvb_buf;
  void *xb = (void *)p;
  int *flags = (int *)(p + TRSVB_XBYTES);
  unsigned *ticket = (unsigned *)(p + TRSVB_XBYTES + TRSVB_NBMAX * sizeof(int));
  int *err = (int *)((char *)L->trsv_buf + TRSV_XBYTES + TRSV_NBMAX * sizeof(int));
  const int epoch = ++L->trsvb_epoch;
  bem_prof_begin(Cx<R>::KT, work, st);
  switch (nrhs) {
  case 1: hipLaunchKernelGGL((k_trsv_batch<R, UPPER, 1>), dim3(foff), dim3(256), trsv_lds_bytes<R>(), st, a, xb, flags, err, ticket, epoch); break;
  case 2: hipLaunchKernelGGL((k_trsv_batch<R, UPPER, 2>), dim3(foff), dim3(256), trsv_lds_bytes<R>(), st, a, xb, flags, err, ticket, epoch); break;
  case 3: hipLaunchKernelGGL((k_trsv_batch<R, UPPER, 3>), dim3(foff), dim3(256), trsv_lds_bytes<R>(), st, a, xb, flags, err, ticket, epoch); break;
  default: hipLaunchKernelGGL((k_trsv_batch<R, UPPER, 4>), dim3(foff), dim3(256), trsv_lds_bytes<R>(), st, a, xb, flags, err, ticket, epoch); break;
  }
  bem_prof_end(Cx<R>::KT, st);
  if (hipGetLastError() != hipSuccess) return false;
  L->trsvb_ticket += (unsigned)foff;   // every workgroup of the launch takes one ticket
  return true;
}

template <class R>
static void getrs_t(int n, int nrhs, const typename Cx<R>::Z *A, size_t lda, const int *piv, typename Cx<R>::Z *B,
                    size_t ldb, hipStream_t st);

// solves with the factors of N systems (getrs_t of each): the row
// interchanges of all systems in one launch, then the unit-lower and the
// upper one-launch solves of all systems together; systems that do not fit
// one launch are solved one after another.  Returns 0, or -1 if a batched
// launch failed after the interchanges.
template <class R>
static int getrs_batch_t(int N, const int *n, int nrhs, const typename Cx<R>::Z *const *A, const size_t *lda,
                         const int *const *piv, typename Cx<R>::Z *const *B, const size_t *ldb, hipStream_t st)
{
  typedef typename Cx<R>::Z Z;
  if (N <= 0 || nrhs <= 0) return 0;
  if (!trsv_batch_fits(N, n, nrhs)) {
    for (int q = 0; q < N; q++) getrs_t<R>(n[q], nrhs, A[q], lda[q], piv[q], B[q], ldb[q], st);
    return 0;
  }
  LaswpArgs<Z> la{};
  for (int q = 0; q < N; q++) la.s[q] = LaswpSys<Z>{B[q], piv[q], ldb[q], n[q], 0};
  hipLaunchKernelGGL(k_laswp_batch<Z>, dim3(N), dim3(256), 0, st, la, nrhs);
  if (!trsv_batch<R, false>(N, n, nrhs, A, lda, B, ldb, st)) return -1;
  if (!trsv_batch<R, true>(N, n, nrhs, A, lda, B, ldb, st)) return -1;
  return 0;
}

template <class R>
static void getrs_t(int n, int nrhs, const typename Cx<R>::Z *A, size_t lda, const int *piv, typename Cx<R>::Z *B,
                    size_t ldb, hipStream_t st)
{
  typedef typename Cx<R>::Z Z;
  if (n <= 0 || nrhs <= 0) return;
  hipLaunchKernelGGL(k_laswp<Z>, dim3((nrhs + 255) / 256), dim3(256), 0, st, B, ldb, 0, nrhs, 0, n, piv);
  // up to two one-launch solves of TRSV_MAXR columns each, else (and for
  // any group the one-launch path cannot take) the recursive solves
  const int ng = nrhs <= 2 * TRSV_MAXR ? (nrhs + TRSV_MAXR - 1) / TRSV_MAXR : 0;
  for (int q = 0; q < 2; q++) {
    const bool upper = q == 1;
    int c0 = 0;
    for (int g = 0; g < ng; g++, c0 += TRSV_MAXR) {
      const int nc = min(TRSV_MAXR, nrhs - c0);
      const bool ok = upper ? trsv_sf<R, true>(A, lda, n, B + c0, ldb, nc, st)
                            : trsv_sf<R, false>(A, lda, n, B + c0, ldb, nc, st);
      if (!ok) break;
    }
    if (c0 < nrhs) {
      if (upper) trsm_upper_rec<R>(A, lda, n, B + c0, ldb, nrhs - c0, st);
      else trsm_lunit_rec<R>(A, lda, n, B + c0, ldb, nrhs - c0, st);
    }
  }
}


// ---- the allocation-free solve ----
// The work area of the factorisation of bem_zgesv_noalloc lives in a fixed
// device array of the library, sized for n <= NOALLOC_NMAX: the column-major
// panel copy (64 columns), three row tables, the panel's pivots and the
// per-block pivot candidates of the launch-per-column panel.  Its device
// address is taken once per device at set-up (bem_la_noalloc_init, with the
// device current); host code uses that address, never the array's name.
// Every call on a GPU shares the array: calls that overlap on one GPU race.
#define NOALLOC_NMAX 8192
#define NOALLOC_NBLK ((NOALLOC_NMAX + LU_NT - 1) / LU_NT)
#define NOALLOC_BYTES ((size_t)64 * NOALLOC_NMAX * sizeof(cd) + 3 * (size_t)NOALLOC_NMAX * sizeof(int) + \
                       64 * sizeof(int) + 2 * (size_t)NOALLOC_NBLK * (sizeof(double) + sizeof(int)))
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
#define NOALLOC_NDEV 16
static void *g_scratch_ptr[NOALLOC_NDEV];

// the factorisation of getrf_t with its work area at ws and the
// launch-per-column panel (no exchange areas, no memset: info is cleared by
// a kernel), then the row interchanges and the recursive triangular solves
// of B; enqueued on st, nothing waited for
template <class R>
static void gesv_noalloc_t(int n, int nrhs, typename Cx<R>::Z *A, typename Cx<R>::Z *B, int *piv, int *info, char *ws,
                           hipStream_t st)
{
  const size_t ldp = ((size_t)n + 63) / 64 * 64;
  const int nblk = (n + LU_NT - 1) / LU_NT;
  LuWs w;
  char *p = ws;
  w.P = (void *)p;                       p += (size_t)64 * ldp * sizeof(cd);
  w.ldp = ldp;
  w.used = (int *)p;                     p += ldp * sizeof(int);
  w.rowAt = (int *)p;                    p += ldp * sizeof(int);
  w.pos = (int *)p;                      p += ldp * sizeof(int);
  w.pv = (int *)p;                       p += 64 * sizeof(int);
  w.partV = (double *)p;                 p += 2 * (size_t)nblk * sizeof(double);
  w.partI = (int *)p;
  w.nbmax = nblk;
  w.info = info;
  w.coop_ok = false;
  w.ordinary = false;
  hipLaunchKernelGGL(k_zero_bytes, dim3(1), dim3(64), 0, st, (unsigned int *)info, (size_t)1);
  lu_rec<R>(A, (size_t)n, n, 0, n, piv, w, st);
  typedef typename Cx<R>::Z Z;
  hipLaunchKernelGGL(k_laswp<Z>, dim3((nrhs + 255) / 256), dim3(256), 0, st, B, (size_t)nrhs, 0, nrhs, 0, n, piv);
  trsm_lunit_rec<R>(A, (size_t)n, n, B, (size_t)nrhs, nrhs, st);
  trsm_upper_rec<R>(A, (size_t)n, n, B, (size_t)nrhs, nrhs, st);
}

// the error flag of the one-launch triangular solves (a wait timed out);
// clears it
static int trsv_error(hipStream_t st)
{
  LaLane *L = la_lane(st);
  if (!L || !L->trsv_buf) return 0;
  int *err = (int *)((char *)L->trsv_buf + TRSV_XBYTES + TRSV_NBMAX * sizeof(int));
  int h = 0;
  if (hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    return -1;
  if (h) (void)hipMemsetAsync(err, 0, sizeof(int), st);
  return h;
}

extern "C" {

int bem_la_init(void)
{
  if (la_init_t<double>() || la_init_t<float>()) return -1;
  if (hipFuncSetAttribute((const void *)k_panel_perm, hipFuncAttributeMaxDynamicSharedMemorySize, PERM_LDS_MAX) != hipSuccess)
    return -1;
  g_big_tiles = true;
  int dev = 0, coop = 0, ncu = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess)
    g_coop = coop != 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0) {
    int bd = 0, bf = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bd, (const void *)k_panel_coop<double>, 64, 0) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&bf, (const void *)k_panel_coop<float>, 64, 0) == hipSuccess &&
        bd > 0 && bf > 0) {
      g_panel_cap = bf * ncu;
      g_panel_unit_d = (bf + bd - 1) / bd;
    }
  }
  (void)hipGetLastError();
  const char *e = getenv("BEM3_HIP_PANEL_COOP");
  g_panel_ordinary = !(e && strcmp(e, "1") == 0);
  const char *es = getenv("BEM3_HIP_PANEL_SLOT");
  g_panel_slot = !(es && strcmp(es, "0") == 0);
  g_gemm_bf6 = true;
  if (g_gemm_bf6 &&
      hipFuncSetAttribute((const void *)k_cgemm_bf6, hipFuncAttributeMaxDynamicSharedMemorySize, BF_LDS_BYTES) != hipSuccess) {
    (void)hipGetLastError();
    g_gemm_bf6 = false;
  }
  return 0;
}

// the device address of the allocation-free solve's work area on device dev
// (the current device); 0 or -1
int bem_la_noalloc_init(int dev)
{
  if (dev < 0 || dev >= NOALLOC_NDEV) return -1;
  if (g_scratch_ptr[dev]) return 0;
  void *ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return -1;
  g_scratch_ptr[dev] = ptr;
  return 0;
}

// A X = B on device-resident row-major A (n x n) and B (n x nrhs) with the
// fixed work area of device dev: 0, -1 (n beyond the work area, or no work
// area on dev), enqueued on st only
int bem_zgesv_noalloc(int dev, int n, int nrhs, void *A, void *B, int *piv, int *info, hipStream_t st)
{
  if (dev < 0 || dev >= NOALLOC_NDEV || !g_scratch_ptr[dev] || n < 1 || n > NOALLOC_NMAX || nrhs < 1) return -1;
  gesv_noalloc_t<double>(n, nrhs, (cd *)A, (cd *)B, piv, info, (char *)g_scratch_ptr[dev], st);
  return 0;
}
int bem_noalloc_nmax(void) { return NOALLOC_NMAX; }

// nonzero if a one-launch triangular solve timed out since the last call
int bem_trsv_error(hipStream_t st) { return trsv_error(st); }

// frees the cached work area of the factorisation; the uncached exchange
// areas of the one-launch panel and triangular solves stay allocated until
// the process ends (memory freed after a kernel read it uncached must not be
// handed out again)
void bem_la_release(void)
{
  std::lock_guard<std::mutex> lk(g_lanes_mu);
  for (int i = 0; i < LA_NLANES; i++) {
    if (g_lanes[i].lu_ws) (void)hipFree(g_lanes[i].lu_ws);
    g_lanes[i].lu_ws = nullptr;
    g_lanes[i].lu_bytes = 0;
  }
}

// nj products of k_gemm_skinny4_batch: jobs (host, nj entries of
// {A, lda, B, C, m, k, beta}) copied with their prefix table into dev (a
// device area of at least bem_skinny_batch_bytes(nj) bytes), one launch
size_t bem_skinny_batch_bytes(int nj) { return ((size_t)nj * sizeof(SkJob) + 255) / 256 * 256 + (size_t)(nj + 1) * sizeof(int); }
int bem_zgemm_skinny4_batch(const void *jobs, int nj, void *dev, void *host_tab, hipStream_t st)
{
  if (nj <= 0) return 0;
  const SkJob *J = (const SkJob *)jobs;
  const size_t jb = ((size_t)nj * sizeof(SkJob) + 255) / 256 * 256;
  char *h = (char *)host_tab;   // bem_skinny_batch_bytes(nj) bytes, alive until the copy completes
  memcpy(h, J, (size_t)nj * sizeof(SkJob));
  int *f = (int *)(h + jb);
  long tot = 0;
  double flops = 0.0, bytes = 0.0;
  for (int i = 0; i < nj; i++) {
    f[i] = (int)tot;
    tot += (J[i].m + 15) / 16;
    flops += 8.0 * J[i].m * 4.0 * J[i].k;
    bytes += ((double)J[i].m * J[i].k + 4.0 * J[i].k + 4.0 * J[i].m * (J[i].beta != 0.0 ? 2.0 : 1.0)) * sizeof(cd);
  }
  if (tot > 0x7fffffffL) return -1;
  f[nj] = (int)tot;
  if (hipMemcpyAsync(dev, h, bem_skinny_batch_bytes(nj), hipMemcpyHostToDevice, st) != hipSuccess) return -1;
  bem_prof_begin_b(Cx<double>::KGV, flops, bytes, st);
  hipLaunchKernelGGL(k_gemm_skinny4_batch, dim3((unsigned)tot), dim3(256), 0, st, (const SkJob *)dev,
                     (const int *)((char *)dev + jb), nj);
  bem_prof_end(Cx<double>::KGV, st);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// complex FP64 (interleaved double pairs)
void bem_zgemm(int m, int n, int k, double ar, double ai, const void *A, size_t lda, const void *B, size_t ldb,
               double br, double bi, void *C, size_t ldc, hipStream_t st)
{
  gemm_t<double>(m, n, k, mk(ar, ai), (const cd *)A, lda, (const cd *)B, ldb, mk(br, bi), (cd *)C, ldc, st);
}

int bem_zgetrf(int n, void *A, size_t lda, int *piv, void *work, int *info, int nb, hipStream_t st)
{
  (void)work;
  (void)nb;
  return getrf_t<double>(n, (cd *)A, lda, piv, info, st);
}

void bem_zgetrs(int n, int nrhs, const void *A, size_t lda, const int *piv, void *B, size_t ldb, int nb, hipStream_t st)
{
  (void)nb;
  getrs_t<double>(n, nrhs, (const cd *)A, lda, piv, (cd *)B, ldb, st);
}

// mean time (ms) of reps launches of one product kernel (kernel 0 64x64
// tiles, 1 128x128 tiles, 2 bf16-split) on operands of its own, after one
// warm-up launch: 0, -1 on a HIP error, -2 if the kernel is unavailable
int bem_gemm_time(int prec, int kernel, int m, int n, int k, int reps, double *ms, hipStream_t st)
{
  if (kernel >= 1 && !g_big_tiles) return -2;
  if (kernel == 2 && (prec != 1 || !g_gemm_bf6)) return -2;
  const size_t es = prec == 0 ? sizeof(cd) : sizeof(cf);
  void *A = nullptr, *B = nullptr, *C = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -1;
  if (hipMalloc(&A, es * m * k) == hipSuccess && hipMalloc(&B, es * k * n) == hipSuccess &&
      hipMalloc(&C, es * m * n) == hipSuccess && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
      hipMemsetAsync(A, 0x3c, es * m * k, st) == hipSuccess && hipMemsetAsync(B, 0x3b, es * k * n, st) == hipSuccess) {
    g_gemm_force = kernel;
    for (int r = 0; r <= reps; r++) {
      if (r == 1) (void)hipEventRecord(e0, st);
      if (prec == 0)
        gemm_t<double>(m, n, k, mk(1, 0), (const cd *)A, k, (const cd *)B, n, mk(0, 0), (cd *)C, n, st);
      else
        gemm_t<float>(m, n, k, mkf(1, 0), (const cf *)A, k, (const cf *)B, n, mkf(0, 0), (cf *)C, n, st);
    }
    g_gemm_force = -1;
    float t = 0.f;
    if (hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
        hipEventElapsedTime(&t, e0, e1) == hipSuccess && hipGetLastError() == hipSuccess) {
      *ms = (double)t / reps;
      rc = 0;
    }
  }
  (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (A) (void)hipFree(A);
  if (B) (void)hipFree(B);
  if (C) (void)hipFree(C);
  return rc;
}

// complex FP32 (interleaved float pairs)
void bem_cgemm(int m, int n, int k, double ar, double ai, const void *A, size_t lda, const void *B, size_t ldb,
               double br, double bi, void *C, size_t ldc, hipStream_t st)
{
  gemm_t<float>(m, n, k, mkf((float)ar, (float)ai), (const cf *)A, lda, (const cf *)B, ldb, mkf((float)br, (float)bi),
                (cf *)C, ldc, st);
}

// solves with the factors of N systems together (nrhs columns each, row
// stride ldb[s]); 0, or -1 if a batched launch failed
int bem_cgetrs_batch(int N, const int *n, int nrhs, const void *const *A, const size_t *lda, const int *const *piv,
                     void *const *B, const size_t *ldb, hipStream_t st)
{
  return getrs_batch_t<float>(N, n, nrhs, (const cf *const *)A, lda, piv, (cf *const *)B, ldb, st);
}
int bem_zgetrs_batch(int N, const int *n, int nrhs, const void *const *A, const size_t *lda, const int *const *piv,
                     void *const *B, const size_t *ldb, hipStream_t st)
{
  return getrs_batch_t<double>(N, n, nrhs, (const cd *const *)A, lda, piv, (cd *const *)B, ldb, st);
}

int bem_cgetrf(int n, void *A, size_t lda, int *piv, int *info, hipStream_t st)
{
  return getrf_t<float>(n, (cf *)A, lda, piv, info, st);
}

void bem_cgetrs(int n, int nrhs, const void *A, size_t lda, const int *piv, void *B, size_t ldb, hipStream_t st)
{
  getrs_t<float>(n, nrhs, (const cf *)A, lda, piv, (cf *)B, ldb, st);
}
}
