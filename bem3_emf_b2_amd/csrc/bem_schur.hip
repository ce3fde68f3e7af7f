// bem_schur.hip -- one-medium block elimination of the boundary-integral
// system (the exact equivalent of factorising the reference's dense 8NN
// matrix of create_Amatrix_csr_dac, solve_bieq2.c:209-341, for an object
// with one interior medium) in complex FP64 or FP32, and the FP64 residual
// of the full system that the mixed-precision form refines against.
#include "bem_schur.h"

extern "C" {
void bem_zgemm(int, int, int, double, double, const void *, size_t, const void *, size_t, double, double, void *, size_t,
               hipStream_t);
void bem_cgemm(int, int, int, double, double, const void *, size_t, const void *, size_t, double, double, void *, size_t,
               hipStream_t);
int bem_cgetrf(int, void *, size_t, int *, int *, hipStream_t);
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

template <class Zt>
static void schur_apply_t(const SchurF<Zt> &f, const Zt *rhs, bool r0nz, Zt *X, Zt *w, hipStream_t st)
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
  xgetrs(n, 1, f.Sf, f.ld, f.pivS, u3, 1, st);
  // u_c = M^{-1} (b_c - delta G1 N_c u3)
  hipLaunchKernelGGL(k_nscale<Zt>, grid1(N), dim3(256), 0, st, u3, f.col1, n, Qn);
  xgemm(n, 3, n, -f.delt, f.G1, f.ld, Qn, 3, one<Zt>(), Bc, 3, st);
  xgetrs(n, 3, f.Mf, f.ld, f.pivM, Bc, 3, st);
  // U4 = [u0 u1 u2 u3], Q4 = Z U4 (- q0), X = [U4^T ; Q4^T]
  hipLaunchKernelGGL(k_pack_u4<Zt>, grid1(N), dim3(256), 0, st, Bc, u3, n, U4);
  xgemm(n, 4, n, one<Zt>(), f.Zm, f.ld, U4, 4, zero<Zt>(), Q4, 4, st);
  hipLaunchKernelGGL(k_fill_x<Zt>, grid1(N), dim3(256), 0, st, U4, Q4, r0nz ? (const Zt *)q0 : nullptr, n, X);
}

void schur_apply(const SchurF<cd> &f, const cd *rhs, bool r0nz, cd *X, cd *work, hipStream_t st)
{
  schur_apply_t<cd>(f, rhs, r0nz, X, work, st);
}
void schur_apply(const SchurF<cf> &f, const cf *rhs, bool r0nz, cf *X, cf *work, hipStream_t st)
{
  schur_apply_t<cf>(f, rhs, r0nz, X, work, st);
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
__global__ void k_hadamard_nn_c(cf *__restrict__ W, int nd, size_t ld, const ColInfo *__restrict__ col)
{
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nd * nd) return;
  int i = (int)(idx / nd), j = (int)(idx % nd);
  const double *a = col[i].n, *b = col[j].n;
  W[(size_t)i * ld + j] = W[(size_t)i * ld + j] * (float)(a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
}

void bem_cvt_z2c(const cd *src, cf *dst, size_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_cvt_z2c, grid1(n), dim3(256), 0, st, src, dst, n);
}
void bem_acc_c2z(cd *X, const cf *d, size_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_acc_c2z, grid1(n), dim3(256), 0, st, X, d, n);
}

int schur_factor_c(int n, size_t ld, const cd *G0, const cd *H0, const cd *G1, const cd *H1, const ColInfo *col1,
                   cd beps, cd sigm, cd delt, cf *G0f, cf *Zm, cf *Mf, cf *Sf, cf *G1c, cf *Wt, int *piv0, int *pivM,
                   int *pivS, int *info, hipStream_t st)
{
  const size_t n2 = (size_t)n * ld;
  int rc;
  // Z = G0^{-1} H0
  bem_cvt_z2c(G0, G0f, n2, st);
  bem_cvt_z2c(H0, Zm, n2, st);
  if ((rc = bem_cgetrf(n, G0f, ld, piv0, info, st))) return rc;
  bem_cgetrs(n, n, G0f, ld, piv0, Zm, ld, st);
  // P = G1 Z (in Sf), M = H1 + P, S starts as H1 + beta P
  bem_cvt_z2c(G1, G1c, n2, st);
  bem_cgemm(n, n, n, 1.0, 0.0, G1c, ld, Zm, ld, 0.0, 0.0, Sf, ld, st);
  hipLaunchKernelGGL(k_caxpby, grid1(n2), dim3(256), 0, st, H1, mkf(1.0f, 0.0f), (const cf *)Sf, Mf, n2);
  hipLaunchKernelGGL(k_caxpby, grid1(n2), dim3(256), 0, st, H1, to_cf(beps), (const cf *)Sf, Sf, n2);
  if ((rc = bem_cgetrf(n, Mf, ld, pivM, info, st))) return rc;
  // W = M^{-1} G1, S -= sigma delta G1 (W o n n^T)
  (void)hipMemcpyAsync(Wt, G1c, n2 * sizeof(cf), hipMemcpyDeviceToDevice, st);
  bem_cgetrs(n, n, Mf, ld, pivM, Wt, ld, st);
  hipLaunchKernelGGL(k_hadamard_nn_c, grid1((size_t)n * n), dim3(256), 0, st, Wt, n, ld, col1);
  const cd sd = -(sigm * delt);
  bem_cgemm(n, n, n, sd.x, sd.y, G1c, ld, Wt, ld, 1.0, 0.0, Sf, ld, st);
  return bem_cgetrf(n, Sf, ld, pivS, info, st);
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

void bem_resid_domain(const cd *G, const cd *H, int nd, size_t ld, const ColInfo *col, const cd *X, size_t NN,
                      size_t rowBase, const cd *B, cd *R, cd *w, hipStream_t st)
{
  const size_t N = (size_t)nd;
  cd *a = w, *b = w + 4 * N, *Y = w + 8 * N;
  hipLaunchKernelGGL(k_colvec, grid1(N), dim3(256), 0, st, X, col, nd, NN, a, b);
  bem_zgemm(nd, 4, nd, 1.0, 0.0, H, ld, a, 4, 0.0, 0.0, Y, 4, st);
  bem_zgemm(nd, 4, nd, 1.0, 0.0, G, ld, b, 4, 1.0, 0.0, Y, 4, st);
  hipLaunchKernelGGL(k_resid, grid1(4 * N), dim3(256), 0, st, B + rowBase, Y, nd, R + rowBase);
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
