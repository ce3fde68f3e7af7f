// bem_schur.hip -- one-medium block elimination of the boundary-integral
// system (the exact equivalent of factorising the reference's dense 8NN
// matrix of create_Amatrix_csr_dac, solve_bieq2.c:209-341, for an object
// with one interior medium) in complex FP64 or FP32, and the FP64 residual
// of the full system that the mixed-precision form refines against.
#include "bem_schur.h"
#include "../../include/bem3_hip.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

extern "C" {
void bem_zgemm(int, int, int, double, double, const void *, size_t, const void *, size_t, double, double, void *, size_t,
               hipStream_t);
void bem_cgemm(int, int, int, double, double, const void *, size_t, const void *, size_t, double, double, void *, size_t,
               hipStream_t);
int bem_cgetrf(int, void *, size_t, int *, int *, hipStream_t);
int bem_cgetrs_batch(int, const int *, int, const void *const *, const size_t *, const int *const *, void *const *,
                     const size_t *, hipStream_t);
int bem_zgetrf(int, void *, size_t, int *, void *, int *, int, hipStream_t);
void bem_zgetrs(int, int, const void *, size_t, const int *, void *, size_t, int, hipStream_t);
void bem_cgetrs(int, int, const void *, size_t, const int *, void *, size_t, hipStream_t);
}

static inline dim3 grid1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

static void xgemm(int m, int n, int k, cd a, const cd *A, size_t lda, const cd *B, size_t ldb, cd b, cd *C, size_t ldc,
                  hipStream_t st)
{
  bem_zgemm(m, n, k, a.x, a.y, A, lda, B, ldb, b.x, b.y, C, ldc, st);
}
static void xgemm(int m, int n, int k, cf a, const cf *A, size_t lda, const cf *B, size_t ldb, cf b, cf *C, size_t ldc,
                  hipStream_t st)
{
  bem_cgemm(m, n, k, a.x, a.y, A, lda, B, ldb, b.x, b.y, C, ldc, st);
}
static void xgetrs(int n, int nrhs, const cd *A, size_t lda, const int *piv, cd *B, size_t ldb, hipStream_t st)
{
  bem_zgetrs(n, nrhs, A, lda, piv, B, ldb, 64, st);
}
static void xgetrs(int n, int nrhs, const cf *A, size_t lda, const int *piv, cf *B, size_t ldb, hipStream_t st)
{
  bem_cgetrs(n, nrhs, A, lda, piv, B, ldb, st);
}
static int xgetrf(int n, cd *A, size_t lda, int *piv, int *info, hipStream_t st)
{
  return bem_zgetrf(n, A, lda, piv, nullptr, info, 64, st);
}
static int xgetrf(int n, cf *A, size_t lda, int *piv, int *info, hipStream_t st) { return bem_cgetrf(n, A, lda, piv, info, st); }
template <class Zt> static Zt one() { Zt r; r.x = 1; r.y = 0; return r; }
template <class Zt> static Zt zero() { Zt r; r.x = 0; r.y = 0; return r; }

// dst[r*ncol + c] = src[c*sstride + r]
template <class Zt>
__global__ void k_gather_cols(const Zt *__restrict__ src, size_t sstride, int ncol, Zt *__restrict__ dst, int n)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * ncol) return;
  int r = i / ncol, c = i % ncol;
  dst[i] = src[(size_t)c * sstride + r];
}
// v[r] = sum_c n_c(r) Y[r][c]  (3 columns)
template <class Zt> __global__ void k_ndot(const Zt *__restrict__ Y, const ColInfo *__restrict__ col, int n, Zt *__restrict__ v)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double *nn = col[r].n;
  v[r] = Y[3 * r] * nn[0] + Y[3 * r + 1] * nn[1] + Y[3 * r + 2] * nn[2];
}
// Y[r][c] = n_c(r) u[r]  (3 columns)
template <class Zt> __global__ void k_nscale(const Zt *__restrict__ u, const ColInfo *__restrict__ col, int n, Zt *__restrict__ Y)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int c = 0; c < 3; c++) Y[3 * r + c] = u[r] * col[r].n[c];
}
// domain-0 right-hand side carried into the domain-1 rows:
// Y[r][c] += T[r][c] (c < 3), b3[r] += beta T[r][3]
template <class Zt> __global__ void k_rhs_q0(Zt *__restrict__ Y, Zt *__restrict__ b3, const Zt *__restrict__ T, Zt beta, int n)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int c = 0; c < 3; c++) Y[3 * r + c] = Y[3 * r + c] + T[4 * r + c];
  b3[r] = b3[r] + beta * T[4 * r + 3];
}
// U4[r][c] = Y[r][c] (c<3), u3[r] (c=3)
template <class Zt> __global__ void k_pack_u4(const Zt *__restrict__ Y, const Zt *__restrict__ u3, int n, Zt *__restrict__ U4)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int c = 0; c < 3; c++) U4[4 * r + c] = Y[3 * r + c];
  U4[4 * r + 3] = u3[r];
}
// X[c n + r] = U4[r][c], X[(c+4) n + r] = Q4[r][c] - q0[r][c] (q0 optional)
template <class Zt>
__global__ void k_fill_x(const Zt *__restrict__ U4, const Zt *__restrict__ Q4, const Zt *__restrict__ q0, int n,
                         Zt *__restrict__ X)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int c = 0; c < 4; c++) {
    X[(size_t)c * n + r] = U4[4 * r + c];
    X[(size_t)(c + 4) * n + r] = q0 ? Q4[4 * r + c] - q0[4 * r + c] : Q4[4 * r + c];
  }
}

// ---- S systems without S: GMRES on the operator ----
__device__ __forceinline__ void zget(const cd &a, double &x, double &y) { x = a.x; y = a.y; }
__device__ __forceinline__ void zget(const cf &a, double &x, double &y) { x = a.x; y = a.y; }

// out[2i], out[2i+1] = sum_r conj(V_i[r]) w[r] for i < k (one workgroup per
// i, double accumulation in a fixed order)
template <class Zt>
__global__ void __launch_bounds__(256) k_mdot(const Zt *__restrict__ V, size_t vstride, int k, const Zt *__restrict__ w,
                                              int n, double *__restrict__ out)
{
  __shared__ double sx[4], sy[4];
  const int i = blockIdx.x;
  if (i >= k) return;
  const Zt *v = V + (size_t)i * vstride;
  double ax = 0.0, ay = 0.0;
  for (int r = threadIdx.x; r < n; r += 256) {
    double vx, vy, wx, wy;
    zget(v[r], vx, vy);
    zget(w[r], wx, wy);
    ax += vx * wx + vy * wy;
    ay += vx * wy - vy * wx;
  }
  ax = wsum(ax);
  ay = wsum(ay);
  if ((threadIdx.x & 63) == 0) { sx[threadIdx.x >> 6] = ax; sy[threadIdx.x >> 6] = ay; }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[2 * i] = (sx[0] + sx[1]) + (sx[2] + sx[3]);
    out[2 * i + 1] = (sy[0] + sy[1]) + (sy[2] + sy[3]);
  }
}
// w[r] += sgn * sum_{i<k} h_i V_i[r]   (h: k complex doubles)
template <class Zt>
__global__ void k_maxpy(Zt *__restrict__ w, const Zt *__restrict__ V, size_t vstride, int k, const double *__restrict__ h,
                        double sgn, int n)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double ax, ay;
  zget(w[r], ax, ay);
  for (int i = 0; i < k; i++) {
    double vx, vy;
    zget(V[(size_t)i * vstride + r], vx, vy);
    const double hx = sgn * h[2 * i], hy = sgn * h[2 * i + 1];
    ax += hx * vx - hy * vy;
    ay += hx * vy + hy * vx;
  }
  w[r].x = ax;
  w[r].y = ay;
}
template <class Zt> __global__ void k_vscale(Zt *__restrict__ w, double s, int n)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) { w[r].x = w[r].x * s; w[r].y = w[r].y * s; }
}
// w = b - w
template <class Zt> __global__ void k_bminus(const Zt *__restrict__ b, Zt *__restrict__ w, int n)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) w[r] = b[r] - w[r];
}

// y = S x = Sp x - sigma delta G1 sum_c N_c M^{-1} G1 N_c x   (t: 7 n)
template <class Zt> static void s_apply(const SchurF<Zt> &f, const Zt *x, Zt *y, Zt *t, hipStream_t st)
{
  const int n = f.n;
  const size_t N = (size_t)n;
  Zt *U3 = t, *T3 = t + 3 * N, *v = t + 6 * N;
  hipLaunchKernelGGL(k_nscale<Zt>, grid1(N), dim3(256), 0, st, x, f.col1, n, U3);
  xgemm(n, 3, n, one<Zt>(), f.G1, f.ld, U3, 3, zero<Zt>(), T3, 3, st);
  xgetrs(n, 3, f.Mf, f.ld, f.pivM, T3, 3, st);
  hipLaunchKernelGGL(k_ndot<Zt>, grid1(N), dim3(256), 0, st, T3, f.col1, n, v);
  xgemm(n, 1, n, one<Zt>(), f.Sf, f.ld, x, 1, zero<Zt>(), y, 1, st);
  Zt sd = f.sigm * f.delt;
  xgemm(n, 1, n, -sd, f.G1, f.ld, v, 1, one<Zt>(), y, 1, st);
}

static int copy_dots(const double *d, double *h, int k, hipStream_t st)
{
  if (hipMemcpyAsync(h, d, 2 * (size_t)k * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
  return hipStreamSynchronize(st) == hipSuccess ? 0 : -1;
}

// S x = b by restarted GMRES (classical Gram-Schmidt with one
// re-orthogonalisation, Givens rotations on the host); bx: in b, out x.
// gw: (SCHUR_GM_M + 1) n Krylov vectors, 7 n operator scratch, 2 n (b, x),
// 512 values for the dot products.  Returns 0 or BH_ECONV.
template <class Zt> static int gmres_s(const SchurF<Zt> &f, Zt *bx, Zt *gw, hipStream_t st, int *steps)
{
  const int n = f.n, m = SCHUR_GM_M;
  const size_t N = (size_t)n;
  Zt *V = gw, *t = gw + (m + 1) * N, *b = t + 7 * N, *x = b + N;
  double *dd = (double *)(x + N);
  (void)hipMemcpyAsync(b, bx, N * sizeof(Zt), hipMemcpyDeviceToDevice, st);
  (void)hipMemsetAsync(x, 0, N * sizeof(Zt), st);
  std::vector<double> H((size_t)(m + 1) * m * 2, 0.0), h1(2 * (m + 1)), h2(2 * (m + 1));
  std::vector<std::complex<double>> g(m + 1), cs(m), sn(m);
  double bnorm = -1.0;
  int total = 0;
  for (int cycle = 0; cycle < 20; cycle++) {
    // r = b - S x into V_0
    if (cycle == 0) {
      (void)hipMemcpyAsync(V, b, N * sizeof(Zt), hipMemcpyDeviceToDevice, st);
    } else {
      s_apply(f, x, V, t, st);
      hipLaunchKernelGGL(k_bminus<Zt>, grid1(N), dim3(256), 0, st, b, V, n);
    }
    hipLaunchKernelGGL(k_mdot<Zt>, dim3(1), dim3(256), 0, st, V, N, 1, V, n, dd);
    if (copy_dots(dd, h1.data(), 1, st)) return BH_EHIP;
    const double beta = sqrt(h1[0]);
    if (bnorm < 0.0) bnorm = beta;
    if (bnorm == 0.0 || beta <= f.gm_tol * bnorm) {
      if (steps) *steps = total;
      (void)hipMemcpyAsync(bx, x, N * sizeof(Zt), hipMemcpyDeviceToDevice, st);
      return 0;
    }
    hipLaunchKernelGGL(k_vscale<Zt>, grid1(N), dim3(256), 0, st, V, 1.0 / beta, n);
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    bool conv = false;
    for (int j = 0; j < m; j++) {
      Zt *w = V + (size_t)(j + 1) * N;
      s_apply(f, V + (size_t)j * N, w, t, st);
      // two passes of classical Gram-Schmidt against V_0..V_j
      std::complex<double> hc[SCHUR_GM_M + 1];
      for (int i = 0; i <= j; i++) hc[i] = 0.0;
      for (int pass = 0; pass < 2; pass++) {
        hipLaunchKernelGGL(k_mdot<Zt>, dim3(j + 1), dim3(256), 0, st, V, N, j + 1, w, n, dd);
        hipLaunchKernelGGL(k_maxpy<Zt>, grid1(N), dim3(256), 0, st, w, V, N, j + 1, dd, -1.0, n);
        if (copy_dots(dd, h1.data(), j + 1, st)) return BH_EHIP;
        for (int i = 0; i <= j; i++) hc[i] += std::complex<double>(h1[2 * i], h1[2 * i + 1]);
      }
      hipLaunchKernelGGL(k_mdot<Zt>, dim3(1), dim3(256), 0, st, w, N, 1, w, n, dd);
      if (copy_dots(dd, h2.data(), 1, st)) return BH_EHIP;
      const double hn = sqrt(h2[0]);
      if (hn > 0.0) hipLaunchKernelGGL(k_vscale<Zt>, grid1(N), dim3(256), 0, st, w, 1.0 / hn, n);
      // previous rotations on the new column, then this step's rotation
      std::complex<double> col[SCHUR_GM_M + 2];
      for (int i = 0; i <= j; i++) col[i] = hc[i];
      col[j + 1] = hn;
      for (int i = 0; i < j; i++) {
        const std::complex<double> a = col[i], c2 = col[i + 1];
        col[i] = std::conj(cs[i]) * a + std::conj(sn[i]) * c2;
        col[i + 1] = -sn[i] * a + cs[i] * c2;
      }
      const double na = std::abs(col[j]), nb = std::abs(col[j + 1]), den = std::sqrt(na * na + nb * nb);
      if (den == 0.0) { cs[j] = 1.0; sn[j] = 0.0; }
      else { cs[j] = col[j] / den; sn[j] = col[j + 1] / den; }
      col[j] = std::conj(cs[j]) * col[j] + std::conj(sn[j]) * col[j + 1];
      col[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = std::conj(cs[j]) * g[j];
      for (int i = 0; i <= j; i++) { H[2 * ((size_t)i * m + j)] = col[i].real(); H[2 * ((size_t)i * m + j) + 1] = col[i].imag(); }
      k = j + 1;
      total++;
      if (std::abs(g[j + 1]) <= f.gm_tol * bnorm || hn == 0.0) { conv = true; break; }
    }
    // y = R^{-1} g, x += V y
    std::vector<std::complex<double>> y(k);
    for (int i = k - 1; i >= 0; i--) {
      std::complex<double> a = g[i];
      for (int l = i + 1; l < k; l++)
        a -= std::complex<double>(H[2 * ((size_t)i * m + l)], H[2 * ((size_t)i * m + l) + 1]) * y[l];
      y[i] = a / std::complex<double>(H[2 * ((size_t)i * m + i)], H[2 * ((size_t)i * m + i) + 1]);
    }
    for (int i = 0; i < k; i++) { h2[2 * i] = y[i].real(); h2[2 * i + 1] = y[i].imag(); }
    if (hipMemcpyAsync(dd, h2.data(), 2 * (size_t)k * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
      return BH_EHIP;
    hipLaunchKernelGGL(k_maxpy<Zt>, grid1(N), dim3(256), 0, st, x, V, N, k, dd, 1.0, n);
    if (hipStreamSynchronize(st) != hipSuccess) return BH_EHIP;   // dd is rewritten by the next cycle
    if (conv) {
      if (steps) *steps = total;
      (void)hipMemcpyAsync(bx, x, N * sizeof(Zt), hipMemcpyDeviceToDevice, st);
      return 0;
    }
  }
  return BH_ECONV;
}

template <class Zt>
static int schur_apply_t(const SchurF<Zt> &f, const Zt *rhs, bool r0nz, Zt *X, Zt *w, hipStream_t st, int *gm_steps)
{
  const int n = f.n;
  const size_t N = (size_t)n;
  Zt *Y = w, *Bc = w + 3 * N, *v = w + 6 * N, *u3 = w + 7 * N, *Qn = w + 8 * N, *q0 = w + 11 * N, *T = w + 15 * N,
     *U4 = w + 19 * N, *Q4 = w + 23 * N;
  const Zt *b1 = rhs + 4 * N;   // domain-1 rows
  hipLaunchKernelGGL(k_gather_cols<Zt>, grid1(3 * N), dim3(256), 0, st, b1, N, 3, Bc, n);
  (void)hipMemcpyAsync(u3, b1 + 3 * N, N * sizeof(Zt), hipMemcpyDeviceToDevice, st);
  if (r0nz) {
    // Q_c = Z U_c - q0_c with q0 = G0^{-1} r0: the domain-1 rows gain G1 q0
    hipLaunchKernelGGL(k_gather_cols<Zt>, grid1(4 * N), dim3(256), 0, st, rhs, N, 4, q0, n);
    xgetrs(n, 4, f.G0f, f.ld, f.piv0, q0, 4, st);
    xgemm(n, 4, n, one<Zt>(), f.G1, f.ld, q0, 4, zero<Zt>(), T, 4, st);
    hipLaunchKernelGGL(k_rhs_q0<Zt>, grid1(N), dim3(256), 0, st, Bc, u3, T, f.beps, n);
  }
  // Y = M^{-1} [b0 b1 b2]
  (void)hipMemcpyAsync(Y, Bc, 3 * N * sizeof(Zt), hipMemcpyDeviceToDevice, st);
  xgetrs(n, 3, f.Mf, f.ld, f.pivM, Y, 3, st);
  // u3 = S^{-1} (b3 - sigma G1 sum_c N_c Y_c)
  hipLaunchKernelGGL(k_ndot<Zt>, grid1(N), dim3(256), 0, st, Y, f.col1, n, v);
  xgemm(n, 1, n, -f.sigm, f.G1, f.ld, v, 1, one<Zt>(), u3, 1, st);
  if (gm_steps) *gm_steps = 0;
  if (f.s_lu) {
    xgetrs(n, 1, f.Sf, f.ld, f.pivS, u3, 1, st);
  } else {
    int rc = gmres_s(f, u3, w + 27 * N, st, gm_steps);
    if (rc) return rc;
  }
  // u_c = M^{-1} (b_c - delta G1 N_c u3)
  hipLaunchKernelGGL(k_nscale<Zt>, grid1(N), dim3(256), 0, st, u3, f.col1, n, Qn);
  xgemm(n, 3, n, -f.delt, f.G1, f.ld, Qn, 3, one<Zt>(), Bc, 3, st);
  xgetrs(n, 3, f.Mf, f.ld, f.pivM, Bc, 3, st);
  // U4 = [u0 u1 u2 u3], Q4 = Z U4 (- q0), X = [U4^T ; Q4^T]
  hipLaunchKernelGGL(k_pack_u4<Zt>, grid1(N), dim3(256), 0, st, Bc, u3, n, U4);
  xgemm(n, 4, n, one<Zt>(), f.Zm, f.ld, U4, 4, zero<Zt>(), Q4, 4, st);
  hipLaunchKernelGGL(k_fill_x<Zt>, grid1(N), dim3(256), 0, st, U4, Q4, r0nz ? (const Zt *)q0 : nullptr, n, X);
  return 0;
}

template <class Zt> __global__ void k_copy_v(Zt *__restrict__ dst, const Zt *__restrict__ src, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ---- several objects at once ----
// The steps of schur_apply_t (r0nz, S formed) for N objects in lock step:
// each solve with LU(G0), LU(M) or LU(S) is one batched launch per
// triangle over every object (the chains of all objects advance together),
// the products and elementwise steps run per object in order on st.
int schur_apply_batch(int N, const SchurF<cf> *f, const cf *const *rhs, cf *const *X, cf *const *w, hipStream_t st)
{
  if (N <= 0) return 0;
  for (int o = 0; o < N; o++)
    if (!f[o].s_lu) return BH_ESTATE;
  std::vector<int> n(N);
  std::vector<size_t> ld(N), l3(N, 3), l4(N, 4), l1(N, 1);
  std::vector<const void *> G0f(N), Mf(N), Sf(N);
  std::vector<const int *> p0(N), pM(N), pS(N);
  std::vector<void *> q0v(N), Yv(N), u3v(N), Bcv(N);
  for (int o = 0; o < N; o++) {
    n[o] = f[o].n;
    ld[o] = f[o].ld;
    G0f[o] = f[o].G0f; Mf[o] = f[o].Mf; Sf[o] = f[o].Sf;
    p0[o] = f[o].piv0; pM[o] = f[o].pivM; pS[o] = f[o].pivS;
    const size_t M = (size_t)n[o];
    cf *ww = w[o];
    Yv[o] = ww; Bcv[o] = ww + 3 * M; u3v[o] = ww + 7 * M; q0v[o] = ww + 11 * M;
  }
  auto P = [&](int o, int k) -> cf * { return w[o] + (size_t)k * n[o]; };   // Y 0, Bc 3, v 6, u3 7, Qn 8, q0 11, T 15, U4 19, Q4 23
  // q0 = G0^{-1} r0 (4 columns), Bc, u3 from the domain-1 rows
  for (int o = 0; o < N; o++) {
    const size_t M = (size_t)n[o];
    const cf *b1 = rhs[o] + 4 * M;
    hipLaunchKernelGGL(k_gather_cols<cf>, grid1(3 * M), dim3(256), 0, st, b1, M, 3, P(o, 3), n[o]);
    hipLaunchKernelGGL(k_copy_v<cf>, grid1(M), dim3(256), 0, st, P(o, 7), b1 + 3 * M, M);
    hipLaunchKernelGGL(k_gather_cols<cf>, grid1(4 * M), dim3(256), 0, st, rhs[o], M, 4, P(o, 11), n[o]);
  }
  if (bem_cgetrs_batch(N, n.data(), 4, G0f.data(), ld.data(), p0.data(), q0v.data(), l4.data(), st)) return BH_EHIP;
  // the domain-1 rows gain G1 q0; Y = M^{-1} [b0 b1 b2]
  for (int o = 0; o < N; o++) {
    const size_t M = (size_t)n[o];
    xgemm(n[o], 4, n[o], one<cf>(), f[o].G1, f[o].ld, P(o, 11), 4, zero<cf>(), P(o, 15), 4, st);
    hipLaunchKernelGGL(k_rhs_q0<cf>, grid1(M), dim3(256), 0, st, P(o, 3), P(o, 7), P(o, 15), f[o].beps, n[o]);
    hipLaunchKernelGGL(k_copy_v<cf>, grid1(3 * M), dim3(256), 0, st, P(o, 0), P(o, 3), 3 * M);
  }
  if (bem_cgetrs_batch(N, n.data(), 3, Mf.data(), ld.data(), pM.data(), Yv.data(), l3.data(), st)) return BH_EHIP;
  // u3 = S^{-1} (b3 - sigma G1 sum_c N_c Y_c)
  for (int o = 0; o < N; o++) {
    hipLaunchKernelGGL(k_ndot<cf>, grid1((size_t)n[o]), dim3(256), 0, st, P(o, 0), f[o].col1, n[o], P(o, 6));
    xgemm(n[o], 1, n[o], -f[o].sigm, f[o].G1, f[o].ld, P(o, 6), 1, one<cf>(), P(o, 7), 1, st);
  }
  if (bem_cgetrs_batch(N, n.data(), 1, Sf.data(), ld.data(), pS.data(), u3v.data(), l1.data(), st)) return BH_EHIP;
  // u_c = M^{-1} (b_c - delta G1 N_c u3)
  for (int o = 0; o < N; o++) {
    hipLaunchKernelGGL(k_nscale<cf>, grid1((size_t)n[o]), dim3(256), 0, st, P(o, 7), f[o].col1, n[o], P(o, 8));
    xgemm(n[o], 3, n[o], -f[o].delt, f[o].G1, f[o].ld, P(o, 8), 3, one<cf>(), P(o, 3), 3, st);
  }
  if (bem_cgetrs_batch(N, n.data(), 3, Mf.data(), ld.data(), pM.data(), Bcv.data(), l3.data(), st)) return BH_EHIP;
  // U4 = [u0 u1 u2 u3], Q4 = Z U4 - q0, X = [U4^T ; Q4^T]
  for (int o = 0; o < N; o++) {
    hipLaunchKernelGGL(k_pack_u4<cf>, grid1((size_t)n[o]), dim3(256), 0, st, P(o, 3), P(o, 7), n[o], P(o, 19));
    xgemm(n[o], 4, n[o], one<cf>(), f[o].Zm, f[o].ld, P(o, 19), 4, zero<cf>(), P(o, 23), 4, st);
    hipLaunchKernelGGL(k_fill_x<cf>, grid1((size_t)n[o]), dim3(256), 0, st, P(o, 19), P(o, 23), (const cf *)P(o, 11), n[o],
                       X[o]);
  }
  return 0;
}

// ---- many right-hand sides at once (nb beams) ----
// Layouts: row-major n x (K nb) with the column of component c of beam b at
// b K + c; right-hand sides and solutions in the 8NN layout, beam b at
// rhs + b 8N.
template <class Zt>
__global__ void k_mgather(const Zt *__restrict__ rhs, size_t N, int nb, int K, int c0, Zt *__restrict__ dst, int n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * nb * K) return;
  const int r = (int)(i / ((size_t)nb * K)), q = (int)(i % ((size_t)nb * K)), b = q / K, c = q % K;
  dst[i] = rhs[(size_t)b * 8 * N + (size_t)(c0 + c) * N + r];
}
template <class Zt>
__global__ void k_mndot(const Zt *__restrict__ Y, const ColInfo *__restrict__ col, int n, int nb, Zt *__restrict__ v)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * nb) return;
  const int r = (int)(i / nb), b = (int)(i % nb);
  const double *nn = col[r].n;
  const Zt *y = Y + (size_t)r * 3 * nb + 3 * b;
  v[i] = y[0] * nn[0] + y[1] * nn[1] + y[2] * nn[2];
}
template <class Zt>
__global__ void k_mnscale(const Zt *__restrict__ u, const ColInfo *__restrict__ col, int n, int nb, Zt *__restrict__ Y)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * nb) return;
  const int r = (int)(i / nb), b = (int)(i % nb);
  for (int c = 0; c < 3; c++) Y[(size_t)r * 3 * nb + 3 * b + c] = u[i] * col[r].n[c];
}
template <class Zt>
__global__ void k_mrhs_q0(Zt *__restrict__ Y, Zt *__restrict__ b3, const Zt *__restrict__ T, Zt beta, int n, int nb)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * nb) return;
  const int r = (int)(i / nb), b = (int)(i % nb);
  const Zt *t = T + (size_t)r * 4 * nb + 4 * b;
  for (int c = 0; c < 3; c++) Y[(size_t)r * 3 * nb + 3 * b + c] = Y[(size_t)r * 3 * nb + 3 * b + c] + t[c];
  b3[i] = b3[i] + beta * t[3];
}
template <class Zt>
__global__ void k_mpack_u4(const Zt *__restrict__ Y, const Zt *__restrict__ u3, int n, int nb, Zt *__restrict__ U4)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * nb) return;
  const int r = (int)(i / nb), b = (int)(i % nb);
  for (int c = 0; c < 3; c++) U4[(size_t)r * 4 * nb + 4 * b + c] = Y[(size_t)r * 3 * nb + 3 * b + c];
  U4[(size_t)r * 4 * nb + 4 * b + 3] = u3[i];
}
template <class Zt>
__global__ void k_mfill_x(const Zt *__restrict__ U4, const Zt *__restrict__ Q4, const Zt *__restrict__ q0, int n, int nb,
                          Zt *__restrict__ X)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * nb) return;
  const int r = (int)(i / nb), b = (int)(i % nb);
  Zt *x = X + (size_t)b * 8 * n;
  for (int c = 0; c < 4; c++) {
    const size_t o = (size_t)r * 4 * nb + 4 * b + c;
    x[(size_t)c * n + r] = U4[o];
    x[(size_t)(c + 4) * n + r] = q0 ? Q4[o] - q0[o] : Q4[o];
  }
}

// X_b = A^{-1} rhs_b for nb right-hand sides with S formed (the steps of
// schur_apply_t, every product and triangular solve on all beams at once);
// work holds schur_work_multi(n, nb) values
template <class Zt>
static int schur_apply_multi_t(const SchurF<Zt> &f, const Zt *rhs, bool r0nz, Zt *X, Zt *w, int nb, hipStream_t st,
                               const SchurInv<Zt> *inv)
{
  if (!f.s_lu) return BH_ESTATE;
  const int n = f.n;
  const size_t N = (size_t)n, M = (size_t)nb;
  // 27 N M values: Y 3, Bc 3, v 1, u3 1, Qn 3, q0 4, T 4, U4 4, Q4 4 (with
  // the inverses a product writes into a free area of the same size and
  // the two names swap)
  Zt *Y = w, *Bc = Y + 3 * N * M, *v = Bc + 3 * N * M, *u3 = v + N * M, *Qn = u3 + N * M, *q0 = Qn + 3 * N * M,
     *T = q0 + 4 * N * M, *U4 = T + 4 * N * M, *Q4 = U4 + 4 * N * M;
  hipLaunchKernelGGL(k_mgather<Zt>, grid1(3 * N * M), dim3(256), 0, st, rhs, N, nb, 3, 4, Bc, n);
  hipLaunchKernelGGL(k_mgather<Zt>, grid1(N * M), dim3(256), 0, st, rhs, N, nb, 1, 7, u3, n);
  if (r0nz) {
    hipLaunchKernelGGL(k_mgather<Zt>, grid1(4 * N * M), dim3(256), 0, st, rhs, N, nb, 4, 0, q0, n);
    if (inv) {
      xgemm(n, 4 * nb, n, one<Zt>(), inv->G0i, N, q0, 4 * M, zero<Zt>(), U4, 4 * M, st);
      std::swap(q0, U4);
    } else {
      xgetrs(n, 4 * nb, f.G0f, f.ld, f.piv0, q0, 4 * M, st);
    }
    xgemm(n, 4 * nb, n, one<Zt>(), f.G1, f.ld, q0, 4 * M, zero<Zt>(), T, 4 * M, st);
    hipLaunchKernelGGL(k_mrhs_q0<Zt>, grid1(N * M), dim3(256), 0, st, Bc, u3, T, f.beps, n, nb);
  }
  if (inv) {
    xgemm(n, 3 * nb, n, one<Zt>(), inv->Mi, N, Bc, 3 * M, zero<Zt>(), Y, 3 * M, st);
  } else {
    (void)hipMemcpyAsync(Y, Bc, 3 * N * M * sizeof(Zt), hipMemcpyDeviceToDevice, st);
    xgetrs(n, 3 * nb, f.Mf, f.ld, f.pivM, Y, 3 * M, st);
  }
  hipLaunchKernelGGL(k_mndot<Zt>, grid1(N * M), dim3(256), 0, st, Y, f.col1, n, nb, v);
  xgemm(n, nb, n, -f.sigm, f.G1, f.ld, v, M, one<Zt>(), u3, M, st);
  if (inv) {
    xgemm(n, nb, n, one<Zt>(), inv->Si, N, u3, M, zero<Zt>(), v, M, st);
    std::swap(u3, v);
  } else {
    xgetrs(n, nb, f.Sf, f.ld, f.pivS, u3, M, st);
  }
  hipLaunchKernelGGL(k_mnscale<Zt>, grid1(N * M), dim3(256), 0, st, u3, f.col1, n, nb, Qn);
  xgemm(n, 3 * nb, n, -f.delt, f.G1, f.ld, Qn, 3 * M, one<Zt>(), Bc, 3 * M, st);
  if (inv) {
    xgemm(n, 3 * nb, n, one<Zt>(), inv->Mi, N, Bc, 3 * M, zero<Zt>(), Y, 3 * M, st);
    std::swap(Bc, Y);
  } else {
    xgetrs(n, 3 * nb, f.Mf, f.ld, f.pivM, Bc, 3 * M, st);
  }
  hipLaunchKernelGGL(k_mpack_u4<Zt>, grid1(N * M), dim3(256), 0, st, Bc, u3, n, nb, U4);
  xgemm(n, 4 * nb, n, one<Zt>(), f.Zm, f.ld, U4, 4 * M, zero<Zt>(), Q4, 4 * M, st);
  hipLaunchKernelGGL(k_mfill_x<Zt>, grid1(N * M), dim3(256), 0, st, U4, Q4, r0nz ? (const Zt *)q0 : nullptr, n, nb, X);
  return 0;
}

int schur_apply_multi(const SchurF<cf> &f, const cf *rhs, bool r0nz, cf *X, cf *work, int nb, hipStream_t st,
                      const SchurInv<cf> *inv)
{
  return schur_apply_multi_t<cf>(f, rhs, r0nz, X, work, nb, st, inv);
}
int schur_apply_multi(const SchurF<cd> &f, const cd *rhs, bool r0nz, cd *X, cd *work, int nb, hipStream_t st,
                      const SchurInv<cd> *inv)
{
  return schur_apply_multi_t<cd>(f, rhs, r0nz, X, work, nb, st, inv);
}

template <class Zt> __global__ void k_eye(Zt *__restrict__ A, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * n) return;
  Zt r;
  r.x = (i / n == i % n) ? 1 : 0;
  r.y = 0;
  A[i] = r;
}

int schur_inverses(const SchurF<cf> &f, cf *G0i, cf *Mi, cf *Si, hipStream_t st)
{
  if (!f.s_lu) return BH_ESTATE;
  const size_t N = (size_t)f.n;
  const cf *lu[3] = {f.G0f, f.Mf, f.Sf};
  const int *pv[3] = {f.piv0, f.pivM, f.pivS};
  cf *out[3] = {G0i, Mi, Si};
  for (int q = 0; q < 3; q++) {
    hipLaunchKernelGGL(k_eye<cf>, grid1(N * N), dim3(256), 0, st, out[q], N);
    xgetrs(f.n, f.n, lu[q], f.ld, pv[q], out[q], N, st);
  }
  return hipGetLastError() == hipSuccess ? 0 : BH_EHIP;
}

// R_b = B_b - A X_b over the rows of one domain for nb beams (8NN layout,
// beam b at + b 8NN): the domain's column vectors of every beam, then two
// products with 4 nb columns
__global__ void k_mcolvec(const cd *__restrict__ X, const ColInfo *__restrict__ col, int nd, size_t NN, int nb,
                          cd *__restrict__ a, cd *__restrict__ b)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nd * nb) return;
  const int j = (int)(i / nb), q = (int)(i % nb);
  const ColInfo c = col[j];
  const cd *x = X + (size_t)q * 8 * NN;
  const size_t cn = (size_t)c.node, o = ((size_t)j * nb + q) * 4;
  cd U[4], Q[4];
  for (int k = 0; k < 4; k++) { U[k] = x[NN * k + cn]; Q[k] = x[NN * (k + 4) + cn]; }
  for (int k = 0; k < 4; k++) a[o + k] = U[k];
  if (c.main) {
    for (int k = 0; k < 4; k++) b[o + k] = -Q[k];
  } else {
    for (int k = 0; k < 3; k++) b[o + k] = c.delt * (U[3] * c.n[k]) + Q[k];
    b[o + 3] = c.sigm * (U[0] * c.n[0] + U[1] * c.n[1] + U[2] * c.n[2]) + c.beps * Q[3];
  }
}
__global__ void k_mresid(const cd *__restrict__ B, const cd *__restrict__ Y, int nd, size_t NN, size_t rowBase, int nb,
                         cd *__restrict__ R)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)4 * nd * nb) return;
  const int q = (int)(i / (4 * (size_t)nd)), rem = (int)(i % (4 * (size_t)nd)), c = rem / nd, r = rem % nd;
  const size_t o = (size_t)q * 8 * NN + rowBase + (size_t)c * nd + r;
  R[o] = B[o] - Y[((size_t)r * nb + q) * 4 + c];
}

void bem_resid_domain_multi(const cd *G, const cd *H, int nd, size_t ld, const ColInfo *col, const cd *X, size_t NN,
                            size_t rowBase, const cd *B, cd *R, cd *w, int nb, hipStream_t st)
{
  const size_t N = (size_t)nd, M = (size_t)nb;
  cd *a = w, *b = w + 4 * N * M, *Y = w + 8 * N * M;
  hipLaunchKernelGGL(k_mcolvec, grid1(N * M), dim3(256), 0, st, X, col, nd, NN, nb, a, b);
  bem_zgemm(nd, 4 * nb, nd, 1.0, 0.0, H, ld, a, 4 * M, 0.0, 0.0, Y, 4 * M, st);
  bem_zgemm(nd, 4 * nb, nd, 1.0, 0.0, G, ld, b, 4 * M, 1.0, 0.0, Y, 4 * M, st);
  hipLaunchKernelGGL(k_mresid, grid1(4 * N * M), dim3(256), 0, st, B, Y, nd, NN, rowBase, nb, R);
}

__global__ void k_mbcol(const ColInfo *__restrict__ col, int nd, int nb, const cd *__restrict__ Ui,
                        const cd *__restrict__ dUi, size_t istride, cd *__restrict__ a, cd *__restrict__ c)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nd * nb) return;
  const int j = (int)(i / nb), q = (int)(i % nb);
  const ColInfo cc = col[j];
  const size_t o = ((size_t)j * nb + q) * 4;
  if (cc.main || cc.mdid != 0) {
    for (int k = 0; k < 4; k++) { a[o + k] = mk(0, 0); c[o + k] = mk(0, 0); }
    return;
  }
  const size_t e = (size_t)q * istride + ((size_t)cc.elem * 4 + cc.slot) * 4;
  const cd *ui = Ui + e, *dui = dUi + e;
  for (int k = 0; k < 4; k++) a[o + k] = ui[k];
  for (int k = 0; k < 3; k++) c[o + k] = dui[k] + ui[3] * (cc.delt * cc.n[k]);
  const cd nu = ui[0] * cc.n[0] + ui[1] * cc.n[1] + ui[2] * cc.n[2];
  c[o + 3] = cc.sigm * nu + cc.beps * dui[3];
}
__global__ void k_mB(const cd *__restrict__ Y, int nd, size_t NN, size_t rowBase, int nb, cd *__restrict__ B)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)4 * nd * nb) return;
  const int q = (int)(i / (4 * (size_t)nd)), rem = (int)(i % (4 * (size_t)nd)), cid = rem / nd, r = rem % nd;
  const cd y = Y[((size_t)r * nb + q) * 4 + cid];
  B[(size_t)q * 8 * NN + rowBase + (size_t)cid * nd + r] = mk(-y.x, -y.y);
}

void bem_build_B_multi(const cd *G, const cd *H, int nd, size_t ld, const ColInfo *col, const cd *Ui, const cd *dUi,
                       size_t istride, size_t NN, size_t rowBase, cd *B, cd *w, int nb, hipStream_t st)
{
  const size_t N = (size_t)nd, M = (size_t)nb;
  cd *a = w, *c = w + 4 * N * M, *Y = w + 8 * N * M;
  hipLaunchKernelGGL(k_mbcol, grid1(N * M), dim3(256), 0, st, col, nd, nb, Ui, dUi, istride, a, c);
  bem_zgemm(nd, 4 * nb, nd, 1.0, 0.0, H, ld, a, 4 * M, 0.0, 0.0, Y, 4 * M, st);
  bem_zgemm(nd, 4 * nb, nd, 1.0, 0.0, G, ld, c, 4 * M, 1.0, 0.0, Y, 4 * M, st);
  hipLaunchKernelGGL(k_mB, grid1(4 * N * M), dim3(256), 0, st, Y, nd, NN, rowBase, nb, B);
}

// |x_b|^2 for nb vectors of n values (x_b at x + b n): out[b] (one
// workgroup per vector, a fixed order)
__global__ void __launch_bounds__(256) k_mnrm2(const cd *__restrict__ x, size_t n, double *__restrict__ out)
{
  __shared__ double s[4];
  const cd *v = x + (size_t)blockIdx.x * n;
  double acc = 0.0;
  for (size_t i = threadIdx.x; i < n; i += 256) acc += abs2(v[i]);
  acc = wsum(acc);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}
void bem_mnrm2(const cd *x, size_t n, int nb, double *out, hipStream_t st)
{
  hipLaunchKernelGGL(k_mnrm2, dim3(nb), dim3(256), 0, st, x, n, out);
}

int schur_apply(const SchurF<cd> &f, const cd *rhs, bool r0nz, cd *X, cd *work, hipStream_t st, int *gm_steps)
{
  return schur_apply_t<cd>(f, rhs, r0nz, X, work, st, gm_steps);
}
int schur_apply(const SchurF<cf> &f, const cf *rhs, bool r0nz, cf *X, cf *work, hipStream_t st, int *gm_steps)
{
  return schur_apply_t<cf>(f, rhs, r0nz, X, work, st, gm_steps);
}

// ---- complex-FP32 factors ----
__global__ void k_cvt_z2c(const cd *__restrict__ s, cf *__restrict__ d, size_t n)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = to_cf(s[i]);
}
__global__ void k_acc_c2z(cd *__restrict__ X, const cf *__restrict__ d, size_t n)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) X[i] = X[i] + to_cd(d[i]);
}
// out = (cf) x + a y
__global__ void k_caxpby(const cd *__restrict__ x, cf a, const cf *__restrict__ y, cf *__restrict__ out, size_t n)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = to_cf(x[i]) + a * y[i];
}
// W[i][j] *= n(i) . n(j)
template <class Zt> __global__ void k_hadamard_nn(Zt *__restrict__ W, int nd, size_t ld, const ColInfo *__restrict__ col)
{
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nd * nd) return;
  int i = (int)(idx / nd), j = (int)(idx % nd);
  const double *a = col[i].n, *b = col[j].n;
  Zt &w = W[(size_t)i * ld + j];
  const double s = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  w.x = w.x * s;
  w.y = w.y * s;
}

void bem_cvt_z2c(const cd *src, cf *dst, size_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_cvt_z2c, grid1(n), dim3(256), 0, st, src, dst, n);
}
void bem_acc_c2z(cd *X, const cf *d, size_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_acc_c2z, grid1(n), dim3(256), 0, st, X, d, n);
}

int schur_factor_c(int n, size_t ld, const cd *G0, const cd *H0, const cd *G1, const cd *H1, cd beps, cf *G0f, cf *Zm,
                   cf *Mf, cf *Sf, cf *G1c, int *piv0, int *pivM, int *info, hipStream_t st)
{
  const size_t n2 = (size_t)n * ld;
  int rc;
  // Z = G0^{-1} H0
  bem_cvt_z2c(G0, G0f, n2, st);
  bem_cvt_z2c(H0, Zm, n2, st);
  if ((rc = bem_cgetrf(n, G0f, ld, piv0, info, st))) return rc;
  bem_cgetrs(n, n, G0f, ld, piv0, Zm, ld, st);
  // P = G1 Z (in Sf), M = H1 + P, Sp = H1 + beta P
  bem_cvt_z2c(G1, G1c, n2, st);
  bem_cgemm(n, n, n, 1.0, 0.0, G1c, ld, Zm, ld, 0.0, 0.0, Sf, ld, st);
  hipLaunchKernelGGL(k_caxpby, grid1(n2), dim3(256), 0, st, H1, mkf(1.0f, 0.0f), (const cf *)Sf, Mf, n2);
  hipLaunchKernelGGL(k_caxpby, grid1(n2), dim3(256), 0, st, H1, to_cf(beps), (const cf *)Sf, Sf, n2);
  return bem_cgetrf(n, Mf, ld, pivM, info, st);
}

// y = x + a y in complex FP32
__global__ void k_cxpay(const cf *__restrict__ x, cf a, cf *__restrict__ y, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] + a * y[i];
}

int schur_factor_c_inplace(int n, size_t ld, cf *G0, cf *H0, cf *G1, cf *H1, cd beps, cf *P, int *piv0, int *pivM,
                           int *info, hipStream_t st)
{
  const size_t n2 = (size_t)n * ld;
  int rc;
  if ((rc = bem_cgetrf(n, G0, ld, piv0, info, st))) return rc;   // G0 -> LU(G0)
  bem_cgetrs(n, n, G0, ld, piv0, H0, ld, st);                      // H0 -> Z
  bem_cgemm(n, n, n, 1.0, 0.0, G1, ld, H0, ld, 0.0, 0.0, P, ld, st);   // P = G1 Z
  hipLaunchKernelGGL(k_cxpay, grid1(n2), dim3(256), 0, st, (const cf *)P, mkf(1.0f, 0.0f), H1, n2);   // H1 -> M
  // P -> Sp = M + (beta - 1) P = H1 + beta P
  hipLaunchKernelGGL(k_cxpay, grid1(n2), dim3(256), 0, st, (const cf *)H1, mkf((float)beps.x - 1.0f, (float)beps.y), P,
                     n2);
  return bem_cgetrf(n, H1, ld, pivM, info, st);                    // M -> LU(M)
}

template <class Zt> static int form_s_t(const SchurF<Zt> &f, Zt *Sf, int *pivS, Zt *Wt, int *info, hipStream_t st)
{
  const int n = f.n;
  const size_t n2 = (size_t)n * f.ld;
  // W = M^{-1} G1, T = W o (n n^T), S = Sp - sigma delta G1 T
  (void)hipMemcpyAsync(Wt, f.G1, n2 * sizeof(Zt), hipMemcpyDeviceToDevice, st);
  xgetrs(n, n, f.Mf, f.ld, f.pivM, Wt, f.ld, st);
  hipLaunchKernelGGL(k_hadamard_nn<Zt>, grid1((size_t)n * n), dim3(256), 0, st, Wt, n, f.ld, f.col1);
  const Zt sd = -(f.sigm * f.delt);
  xgemm(n, n, n, sd, f.G1, f.ld, Wt, f.ld, one<Zt>(), Sf, f.ld, st);
  return xgetrf(n, Sf, f.ld, pivS, info, st);
}
int schur_form_s(const SchurF<cd> &f, cd *Sf, int *pivS, cd *Wt, int *info, hipStream_t st)
{
  return form_s_t<cd>(f, Sf, pivS, Wt, info, st);
}
int schur_form_s(const SchurF<cf> &f, cf *Sf, int *pivS, cf *Wt, int *info, hipStream_t st)
{
  return form_s_t<cf>(f, Sf, pivS, Wt, info, st);
}

// ---- FP64 residual of the full system ----
// Row r of domain d: sum_j H[r][j] a_c(j) + G[r][j] b_c(j) with, per column
// j (node cn): main side a_c = U_c, b_c = -Q_c; sub side a_c = U_c,
// b_c = delta n_c U_3 + Q_c (c < 3), b_3 = sigma n.U + beta Q_3 -- the
// entries k_build_A places in the row (create_Amatrix_csr_dac).
__global__ void k_colvec(const cd *__restrict__ X, const ColInfo *__restrict__ col, int nd, size_t NN, cd *__restrict__ a,
                         cd *__restrict__ b)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nd) return;
  const ColInfo c = col[j];
  const size_t cn = (size_t)c.node;
  cd U[4], Q[4];
  for (int i = 0; i < 4; i++) { U[i] = X[NN * i + cn]; Q[i] = X[NN * (i + 4) + cn]; }
  for (int i = 0; i < 4; i++) a[4 * j + i] = U[i];
  if (c.main) {
    for (int i = 0; i < 4; i++) b[4 * j + i] = -Q[i];
  } else {
    for (int i = 0; i < 3; i++) b[4 * j + i] = c.delt * (U[3] * c.n[i]) + Q[i];
    b[4 * j + 3] = c.sigm * (U[0] * c.n[0] + U[1] * c.n[1] + U[2] * c.n[2]) + c.beps * Q[3];
  }
}
// R[c nd + r] = B[c nd + r] - Y[r][c]
__global__ void k_resid(const cd *__restrict__ B, const cd *__restrict__ Y, int nd, cd *__restrict__ R)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * nd) return;
  int c = i / nd, r = i % nd;
  R[i] = B[i] - Y[4 * r + c];
}

// rows r0 .. r0+nrows of the domain: R[c nd + r] = B[c nd + r] - Y[r - r0][c]
__global__ void k_resid_rows(const cd *__restrict__ B, const cd *__restrict__ Y, int nd, int r0, int nrows,
                             cd *__restrict__ R)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * nrows) return;
  int c = i / nrows, r = i % nrows;
  const size_t o = (size_t)c * nd + r0 + r;
  R[o] = B[o] - Y[4 * r + c];
}

void bem_resid_colvec(const ColInfo *col, int nd, const cd *X, size_t NN, cd *w, hipStream_t st)
{
  const size_t N = (size_t)nd;
  hipLaunchKernelGGL(k_colvec, grid1(N), dim3(256), 0, st, X, col, nd, NN, w, w + 4 * N);
}

void bem_resid_rows(const cd *G, const cd *H, int nrows, int nd, size_t ld, int r0, size_t rowBase, const cd *B, cd *R,
                    cd *w, hipStream_t st)
{
  const size_t N = (size_t)nd;
  cd *a = w, *b = w + 4 * N, *Y = w + 8 * N;
  bem_zgemm(nrows, 4, nd, 1.0, 0.0, H, ld, a, 4, 0.0, 0.0, Y, 4, st);
  bem_zgemm(nrows, 4, nd, 1.0, 0.0, G, ld, b, 4, 1.0, 0.0, Y, 4, st);
  hipLaunchKernelGGL(k_resid_rows, grid1(4 * (size_t)nrows), dim3(256), 0, st, B + rowBase, Y, nd, r0, nrows,
                     R + rowBase);
}

// R = B - Y on a domain's nd rows (rows rowBase.. of the 8NN system), Y the
// products of bem_resid_rows formed elsewhere (the batched residuals)
void bem_resid_fin(const cd *B, const cd *Y, int nd, size_t rowBase, cd *R, hipStream_t st)
{
  hipLaunchKernelGGL(k_resid_rows, grid1(4 * (size_t)nd), dim3(256), 0, st, B + rowBase, Y, nd, 0, nd, R + rowBase);
}

void bem_resid_domain(const cd *G, const cd *H, int nd, size_t ld, const ColInfo *col, const cd *X, size_t NN,
                      size_t rowBase, const cd *B, cd *R, cd *w, hipStream_t st)
{
  bem_resid_colvec(col, nd, X, NN, w, st);
  bem_resid_rows(G, H, nd, nd, ld, 0, rowBase, B, R, w, st);
}

// sum of |x|^2: 256 workgroup partials in a fixed order, then one workgroup
__global__ void __launch_bounds__(256) k_nrm2_part(const cd *__restrict__ x, size_t n, double *__restrict__ part)
{
  __shared__ double s[4];
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) acc += abs2(x[i]);
  acc = wsum(acc);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}
__global__ void __launch_bounds__(256) k_nrm2_fin(double *part)
{
  __shared__ double s[4];
  double v = wsum(part[threadIdx.x]);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[0] = (s[0] + s[1]) + (s[2] + s[3]);
}

__global__ void __launch_bounds__(256) k_nrm2_fin_to(const double *part, double *out)
{
  __shared__ double s[4];
  double v = wsum(part[threadIdx.x]);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) *out = (s[0] + s[1]) + (s[2] + s[3]);
}

void bem_nrm2sq_to(const cd *x, size_t n, double *tmp, double *out, hipStream_t st)
{
  hipLaunchKernelGGL(k_nrm2_part, dim3(256), dim3(256), 0, st, x, n, tmp);
  hipLaunchKernelGGL(k_nrm2_fin_to, dim3(1), dim3(256), 0, st, tmp, out);
}

double bem_nrm2(const cd *x, size_t n, double *tmp, hipStream_t st)
{
  hipLaunchKernelGGL(k_nrm2_part, dim3(256), dim3(256), 0, st, x, n, tmp);
  hipLaunchKernelGGL(k_nrm2_fin, dim3(1), dim3(256), 0, st, tmp);
  double h = -1.0;
  if (hipMemcpyAsync(&h, tmp, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return -1.0;
  return sqrt(h);
}
