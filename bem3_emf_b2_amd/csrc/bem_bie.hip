// bem_bie.hip -- boundary-integral system of one object on the device.
//
// Domain d has nd node rows (3 per triangle, in the domain's element order)
// and the same nd shape-function columns.  Column j of domain d maps to the
// global element-node index colNode[j] (the 8NN unknown numbering of
// solve_bieq2.c:199-331), to the element slot (colElem, colSlot) and carries
// the boundary-condition class of its element: main side (md == d) or sub
// side with the coefficients delta = k1^2-k2^2, beta = k1^2/k2^2,
// sigma = 1-beta of its (main, sub) media.
//
//   k_build_A     dense 8NN matrix rows of create_Amatrix_csr_dac
//                 (solve_bieq2.c:209-341)
//   k_build_B     right-hand side of create_Bmatrix_dac (mo_solve.c:137-244)
//   k_store_X     tmatrix_bd_store (mo_solve.c:457-545)
//   k_eh_rows     solve_eh_bv + dudt_fcm + inv_matrix_33 (mo_solve.c:547-661)
#include "bem_dev.h"


// A(row, col) entries from G, H of domain d.  rowBase = first A row of the
// domain; rows of component cid are rowBase + cid*nd + r.
__global__ void k_build_A(const cd *__restrict__ G, const cd *__restrict__ H, int nd, size_t ld,
                          const ColInfo *__restrict__ col, size_t rowBase, size_t NN, cd *__restrict__ A, size_t na)
{
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nd * nd) return;
  int r = (int)(idx / nd), j = (int)(idx % nd);
  cd g = G[(size_t)r * ld + j], h = H[(size_t)r * ld + j];
  const ColInfo c = col[j];
  size_t cn = (size_t)c.node;
  for (int cid = 0; cid < 4; cid++) {
    cd *ar = A + (rowBase + (size_t)cid * nd + r) * na;
    if (c.main) {
      ar[NN * cid + cn] = h;
      ar[NN * (cid + 4) + cn] = -g;
    } else if (cid < 3) {
      ar[NN * cid + cn] = h;
      ar[NN * 3 + cn] = c.delt * (c.n[cid] * g);
      ar[NN * (cid + 4) + cn] = g;
    } else {
      ar[NN * 0 + cn] = c.sigm * (c.n[0] * g);
      ar[NN * 1 + cn] = c.sigm * (c.n[1] * g);
      ar[NN * 2 + cn] = c.sigm * (c.n[2] * g);
      ar[NN * 3 + cn] = h;
      ar[NN * 7 + cn] = c.beps * g;
    }
  }
}

// B rows of domain d: only columns whose element has this domain as its
// sub side and the open domain as its main side carry incident data
// (nrows rows of the domain's nd, from the row at rowBase: a row block of
// the matrices)
__global__ void __launch_bounds__(256) k_build_B(const cd *__restrict__ G, const cd *__restrict__ H, int nrows, int nd,
                                                 size_t ld, const ColInfo *__restrict__ col, const cd *__restrict__ Ui,
                                                 const cd *__restrict__ dUi, size_t rowBase, cd *__restrict__ B)
{
  const int r = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (r >= nrows) return;
  cd b[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  for (int j = lane; j < nd; j += 64) {
    const ColInfo c = col[j];
    if (c.main || c.mdid != 0) continue;
    cd g = G[(size_t)r * ld + j], h = H[(size_t)r * ld + j];
    const cd *ui = Ui + ((size_t)c.elem * 4 + c.slot) * 4, *dui = dUi + ((size_t)c.elem * 4 + c.slot) * 4;
    for (int cid = 0; cid < 3; cid++) b[cid] = b[cid] - h * ui[cid] - g * dui[cid] - g * ui[3] * (c.delt * c.n[cid]);
    cd nu = ui[0] * c.n[0] + ui[1] * c.n[1] + ui[2] * c.n[2];
    b[3] = b[3] - h * ui[3] - c.sigm * (g * nu) - c.beps * (g * dui[3]);
  }
  for (int cid = 0; cid < 4; cid++) {
    cd v = wsum(b[cid]);
    if (lane == 0) B[rowBase + (size_t)cid * nd + r] = v;
  }
}

// X -> per-domain boundary values U, dU (slot layout [(s*4+slot)*4+c])
__global__ void k_store_X(const cd *__restrict__ X, int nd, const ColInfo *__restrict__ col, size_t NN,
                          const int *__restrict__ colS, const cd *__restrict__ Ui, const cd *__restrict__ dUi,
                          cd *__restrict__ U, cd *__restrict__ dU)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nd) return;
  const ColInfo c = col[j];
  size_t cn = (size_t)c.node;
  size_t o = ((size_t)colS[j] * 4 + c.slot) * 4;
  if (c.main) {
    for (int i = 0; i < 4; i++) { U[o + i] = X[NN * i + cn]; dU[o + i] = X[NN * (i + 4) + cn]; }
    return;
  }
  cd tU[4], tUi[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
  tU[3] = X[NN * 3 + cn];
  const cd *ui = Ui + ((size_t)c.elem * 4 + c.slot) * 4, *dui = dUi + ((size_t)c.elem * 4 + c.slot) * 4;
  for (int i = 0; i < 3; i++) {
    tU[i] = X[NN * i + cn];
    cd u = tU[i], du = -(X[NN * (i + 4) + cn] + c.delt * tU[3] * c.n[i]);
    if (c.mdid == 0) { tUi[i] = ui[i]; u = u + tUi[i]; du = du - dui[i]; }
    U[o + i] = u;
    dU[o + i] = du;
  }
  U[o + 3] = tU[3];
  cd d3 = -(c.beps * X[NN * 7 + cn] + c.sigm * (tU[0] * c.n[0] + tU[1] * c.n[1] + tU[2] * c.n[2]));
  if (c.mdid == 0) d3 = d3 - c.sigm * (tUi[0] * c.n[0] + tUi[1] * c.n[1] + tUi[2] * c.n[2]);
  dU[o + 3] = d3;
}


// (nrows rows of a row block: dG, dH, dF and rows start at the block's
// first row; nd columns)
__global__ void __launch_bounds__(256) k_eh_rows(const cd *__restrict__ dG, const cd *__restrict__ dH, int nrows, int nd,
                                                 size_t ld, const double *__restrict__ dF,
                                                 const RowEH *__restrict__ rows, const int *__restrict__ colS,
                                                 const ColInfo *__restrict__ col, const cd *__restrict__ U,
                                                 const cd *__restrict__ dU, cd ch, cd *__restrict__ E,
                                                 cd *__restrict__ H)
{
  const int r = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (r >= nrows) return;
  cd dz[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)}, de[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  const cd *gz = dG + (size_t)r * 2 * ld, *ge = gz + ld, *hz = dH + (size_t)r * 2 * ld, *he = hz + ld;
  for (int j = lane; j < nd; j += 64) {
    size_t o = ((size_t)colS[j] * 4 + col[j].slot) * 4;
    cd a = gz[j], b = hz[j], c = ge[j], d = he[j];
    for (int i = 0; i < 4; i++) {
      dz[i] = dz[i] + a * dU[o + i] - b * U[o + i];
      de[i] = de[i] + c * dU[o + i] - d * U[o + i];
    }
  }
  for (int i = 0; i < 4; i++) { dz[i] = wsum(dz[i]); de[i] = wsum(de[i]); }
  if (lane) return;
  const RowEH rw = rows[r];
  size_t ot = ((size_t)rw.s * 4 + rw.slot) * 4;
  const double f0 = dF[3 * r], f1 = dF[3 * r + 1], f2 = dF[3 * r + 2];
  cd Ut[4], dUt[4];
  for (int i = 0; i < 4; i++) {
    Ut[i] = U[ot + i]; dUt[i] = dU[ot + i];
    dz[i] = (dz[i] - Ut[i] * f1) * (1.0 / f0);
    de[i] = (de[i] - Ut[i] * f2) * (1.0 / f0);
  }
  // rows of the inverse of [n; t_z; t_e]
  const double *n = rw.vn, *a = rw.vtz, *b = rw.vte;
  double det = n[0] * a[1] * b[2] + n[1] * a[2] * b[0] + n[2] * a[0] * b[1] - (n[2] * a[1] * b[0] + n[1] * a[0] * b[2] + n[0] * a[2] * b[1]);
  double id = 1.0 / det, m[9];
  m[0] = a[1] * b[2] - a[2] * b[1]; m[1] = -n[1] * b[2] + n[2] * b[1]; m[2] = n[1] * a[2] - n[2] * a[1];
  m[3] = -a[0] * b[2] + a[2] * b[0]; m[4] = n[0] * b[2] - n[2] * b[0]; m[5] = -n[0] * a[2] + n[2] * a[0];
  m[6] = a[0] * b[1] - a[1] * b[0]; m[7] = -n[0] * b[1] + n[1] * b[0]; m[8] = n[0] * a[1] - n[1] * a[0];
  for (int i = 0; i < 9; i++) m[i] *= id;
  cd grd[4][3];   // gradient of component c along axis i
  for (int c = 0; c < 4; c++)
    for (int i = 0; i < 3; i++) grd[c][i] = dUt[c] * m[i * 3] + dz[c] * m[i * 3 + 1] + de[c] * m[i * 3 + 2];
  size_t oe = ((size_t)rw.s * 4 + rw.slot) * 3;
  for (int i = 0; i < 3; i++) E[oe + i] = Ut[i] - grd[3][i];
  H[oe + 0] = ch * (grd[2][1] - grd[1][2]);
  H[oe + 1] = ch * (grd[0][2] - grd[2][0]);
  H[oe + 2] = ch * (grd[1][0] - grd[0][1]);
}

// pack [E|H] node values of domain d in column order (nd x 6) and back
__global__ void k_pack_eh(int nd, const ColInfo *__restrict__ col, const int *__restrict__ colS,
                          const cd *__restrict__ E, const cd *__restrict__ H, cd *__restrict__ P)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nd) return;
  size_t o = ((size_t)colS[j] * 4 + col[j].slot) * 3;
  for (int i = 0; i < 3; i++) { P[(size_t)j * 6 + i] = E[o + i]; P[(size_t)j * 6 + 3 + i] = H[o + i]; }
}

__global__ void k_unpack_eh(int nd, const ColInfo *__restrict__ col, const int *__restrict__ colS,
                            const cd *__restrict__ P, cd *__restrict__ dE, cd *__restrict__ dH)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nd) return;
  size_t o = ((size_t)colS[j] * 4 + col[j].slot) * 3;
  for (int i = 0; i < 3; i++) { dE[o + i] = P[(size_t)j * 6 + i]; dH[o + i] = P[(size_t)j * 6 + 3 + i]; }
}

// ---- Schur-complement helpers for the one-medium object ----
// T(i,j) = W(i,j) * (n_i . n_j) ; S = Mb - sd * G1 T is done with ZGEMM
__global__ void k_hadamard_nn(cd *__restrict__ W, int nd, size_t ld, const ColInfo *__restrict__ col)
{
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nd * nd) return;
  int i = (int)(idx / nd), j = (int)(idx % nd);
  const double *a = col[i].n, *b = col[j].n;
  W[(size_t)i * ld + j] = W[(size_t)i * ld + j] * (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
}

// C = A + alpha * B (elementwise, complex alpha)
__global__ void k_axpby(const cd *__restrict__ A, cd alpha, const cd *__restrict__ B, cd *__restrict__ C, size_t n)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) C[i] = A[i] + alpha * B[i];
}

// gather / scatter of the Schur solve vectors between the 8NN layout
// X[c*NN + node] and the compact nd-order vectors (one medium: node = j)
__global__ void k_scale_rows(cd *__restrict__ X, const ColInfo *__restrict__ col, int nd, int comp, cd alpha)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nd) X[j] = X[j] * (alpha * col[j].n[comp]);
}


// ---- many beams: the boundary E, H steps on all beams at once ----
// PU[j][4b+i], PdU[j][4b+i]: U, dU of beam b at the node of column j (U_b
// at U + b ustride)
__global__ void k_meh_gather(int nd, int nb, const ColInfo *__restrict__ col, const int *__restrict__ colS,
                             const cd *__restrict__ U, const cd *__restrict__ dU, size_t ustride, cd *__restrict__ PU,
                             cd *__restrict__ PdU)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nd * nb) return;
  const int j = (int)(i / nb), b = (int)(i % nb);
  const size_t o = ((size_t)colS[j] * 4 + col[j].slot) * 4 + (size_t)b * ustride;
  for (int k = 0; k < 4; k++) { PU[i * 4 + k] = U[o + k]; PdU[i * 4 + k] = dU[o + k]; }
}

// the finish of k_eh_rows for row r and beam b from the tangential
// products Z[2r][4b+i] (zeta) and Z[2r+1][4b+i] (eta)
__global__ void k_meh_rows(int nd, int nb, const cd *__restrict__ Z, const double *__restrict__ dF,
                           const RowEH *__restrict__ rows, const cd *__restrict__ U, const cd *__restrict__ dU,
                           size_t ustride, cd ch, cd *__restrict__ E, cd *__restrict__ H, size_t estride)
{
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nd * nb) return;
  const int r = (int)(t / nb), b = (int)(t % nb);
  const RowEH rw = rows[r];
  const size_t ot = ((size_t)rw.s * 4 + rw.slot) * 4 + (size_t)b * ustride;
  const double f0 = dF[3 * r], f1 = dF[3 * r + 1], f2 = dF[3 * r + 2];
  const cd *zz = Z + (size_t)(2 * r) * 4 * nb + 4 * b, *ze = Z + (size_t)(2 * r + 1) * 4 * nb + 4 * b;
  cd Ut[4], dUt[4], dz[4], de[4];
  for (int i = 0; i < 4; i++) {
    Ut[i] = U[ot + i]; dUt[i] = dU[ot + i];
    dz[i] = (zz[i] - Ut[i] * f1) * (1.0 / f0);
    de[i] = (ze[i] - Ut[i] * f2) * (1.0 / f0);
  }
  const double *n = rw.vn, *a = rw.vtz, *bb = rw.vte;
  double det = n[0] * a[1] * bb[2] + n[1] * a[2] * bb[0] + n[2] * a[0] * bb[1] -
               (n[2] * a[1] * bb[0] + n[1] * a[0] * bb[2] + n[0] * a[2] * bb[1]);
  double id = 1.0 / det, m[9];
  m[0] = a[1] * bb[2] - a[2] * bb[1]; m[1] = -n[1] * bb[2] + n[2] * bb[1]; m[2] = n[1] * a[2] - n[2] * a[1];
  m[3] = -a[0] * bb[2] + a[2] * bb[0]; m[4] = n[0] * bb[2] - n[2] * bb[0]; m[5] = -n[0] * a[2] + n[2] * a[0];
  m[6] = a[0] * bb[1] - a[1] * bb[0]; m[7] = -n[0] * bb[1] + n[1] * bb[0]; m[8] = n[0] * a[1] - n[1] * a[0];
  for (int i = 0; i < 9; i++) m[i] *= id;
  cd grd[4][3];
  for (int c = 0; c < 4; c++)
    for (int i = 0; i < 3; i++) grd[c][i] = dUt[c] * m[i * 3] + dz[c] * m[i * 3 + 1] + de[c] * m[i * 3 + 2];
  const size_t oe = ((size_t)rw.s * 4 + rw.slot) * 3 + (size_t)b * estride;
  for (int i = 0; i < 3; i++) E[oe + i] = Ut[i] - grd[3][i];
  H[oe + 0] = ch * (grd[2][1] - grd[1][2]);
  H[oe + 1] = ch * (grd[0][2] - grd[2][0]);
  H[oe + 2] = ch * (grd[1][0] - grd[0][1]);
}

// [E|H] of every beam in column order, P[j][6b+i], and back (dE, dH)
__global__ void k_mpack_eh(int nd, int nb, const ColInfo *__restrict__ col, const int *__restrict__ colS,
                           const cd *__restrict__ E, const cd *__restrict__ H, size_t estride, cd *__restrict__ P)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nd * nb) return;
  const int j = (int)(i / nb), b = (int)(i % nb);
  const size_t o = ((size_t)colS[j] * 4 + col[j].slot) * 3 + (size_t)b * estride;
  for (int k = 0; k < 3; k++) { P[i * 6 + k] = E[o + k]; P[i * 6 + 3 + k] = H[o + k]; }
}
__global__ void k_munpack_eh(int nd, int nb, const ColInfo *__restrict__ col, const int *__restrict__ colS,
                             const cd *__restrict__ P, cd *__restrict__ dE, cd *__restrict__ dH, size_t estride)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nd * nb) return;
  const int j = (int)(i / nb), b = (int)(i % nb);
  const size_t o = ((size_t)colS[j] * 4 + col[j].slot) * 3 + (size_t)b * estride;
  for (int k = 0; k < 3; k++) { dE[o + k] = P[i * 6 + k]; dH[o + k] = P[i * 6 + 3 + k]; }
}

extern "C" {
void bem_launch_build_A(const void *G, const void *H, int nd, size_t ld, const void *col, size_t rowBase, size_t NN,
                        void *A, size_t na, hipStream_t st)
{
  size_t n = (size_t)nd * nd;
  hipLaunchKernelGGL(k_build_A, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (const cd *)G, (const cd *)H, nd, ld,
                     (const ColInfo *)col, rowBase, NN, (cd *)A, na);
}
void bem_launch_build_B(const void *G, const void *H, int nrows, int nd, size_t ld, const void *col, const void *Ui,
                        const void *dUi, size_t rowBase, void *B, hipStream_t st)
{
  hipLaunchKernelGGL(k_build_B, dim3((nrows + 3) / 4), dim3(256), 0, st, (const cd *)G, (const cd *)H, nrows, nd, ld,
                     (const ColInfo *)col, (const cd *)Ui, (const cd *)dUi, rowBase, (cd *)B);
}
void bem_launch_store_X(const void *X, int nd, const void *col, size_t NN, const int *colS, const void *Ui,
                        const void *dUi, void *U, void *dU, hipStream_t st)
{
  hipLaunchKernelGGL(k_store_X, dim3((nd + 255) / 256), dim3(256), 0, st, (const cd *)X, nd, (const ColInfo *)col, NN,
                     colS, (const cd *)Ui, (const cd *)dUi, (cd *)U, (cd *)dU);
}
void bem_launch_eh_rows(const void *dG, const void *dH, int nrows, int nd, size_t ld, const double *dF, const void *rows,
                        const int *colS, const void *col, const void *U, const void *dU, double chr, double chi,
                        void *E, void *H, hipStream_t st)
{
  hipLaunchKernelGGL(k_eh_rows, dim3((nrows + 3) / 4), dim3(256), 0, st, (const cd *)dG, (const cd *)dH, nrows, nd, ld, dF,
                     (const RowEH *)rows, colS, (const ColInfo *)col, (const cd *)U, (const cd *)dU, mk(chr, chi),
                     (cd *)E, (cd *)H);
}
void bem_launch_pack_eh(int nd, const void *col, const int *colS, const void *E, const void *H, void *P, hipStream_t st)
{
  hipLaunchKernelGGL(k_pack_eh, dim3((nd + 255) / 256), dim3(256), 0, st, nd, (const ColInfo *)col, colS, (const cd *)E,
                     (const cd *)H, (cd *)P);
}
void bem_launch_unpack_eh(int nd, const void *col, const int *colS, const void *P, void *dE, void *dH, hipStream_t st)
{
  hipLaunchKernelGGL(k_unpack_eh, dim3((nd + 255) / 256), dim3(256), 0, st, nd, (const ColInfo *)col, colS,
                     (const cd *)P, (cd *)dE, (cd *)dH);
}
void bem_launch_hadamard_nn(void *W, int nd, size_t ld, const void *col, hipStream_t st)
{
  size_t n = (size_t)nd * nd;
  hipLaunchKernelGGL(k_hadamard_nn, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (cd *)W, nd, ld,
                     (const ColInfo *)col);
}
void bem_launch_axpby(const void *A, double ar, double ai, const void *B, void *C, size_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (const cd *)A, mk(ar, ai),
                     (const cd *)B, (cd *)C, n);
}
void bem_launch_scale_rows(void *X, const void *col, int nd, int comp, double ar, double ai, hipStream_t st)
{
  hipLaunchKernelGGL(k_scale_rows, dim3((nd + 255) / 256), dim3(256), 0, st, (cd *)X, (const ColInfo *)col, nd, comp,
                     mk(ar, ai));
}
void bem_launch_meh_gather(int nd, int nb, const void *col, const int *colS, const void *U, const void *dU, size_t ustride,
                           void *PU, void *PdU, hipStream_t st)
{
  const size_t n = (size_t)nd * nb;
  hipLaunchKernelGGL(k_meh_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nd, nb, (const ColInfo *)col, colS,
                     (const cd *)U, (const cd *)dU, ustride, (cd *)PU, (cd *)PdU);
}
void bem_launch_meh_rows(int nd, int nb, const void *Z, const double *dF, const void *rows, const void *U, const void *dU,
                         size_t ustride, double chr, double chi, void *E, void *H, size_t estride, hipStream_t st)
{
  const size_t n = (size_t)nd * nb;
  hipLaunchKernelGGL(k_meh_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nd, nb, (const cd *)Z, dF,
                     (const RowEH *)rows, (const cd *)U, (const cd *)dU, ustride, mk(chr, chi), (cd *)E, (cd *)H, estride);
}
void bem_launch_mpack_eh(int nd, int nb, const void *col, const int *colS, const void *E, const void *H, size_t estride,
                         void *P, hipStream_t st)
{
  const size_t n = (size_t)nd * nb;
  hipLaunchKernelGGL(k_mpack_eh, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nd, nb, (const ColInfo *)col, colS,
                     (const cd *)E, (const cd *)H, estride, (cd *)P);
}
void bem_launch_munpack_eh(int nd, int nb, const void *col, const int *colS, const void *P, void *dE, void *dH,
                           size_t estride, hipStream_t st)
{
  const size_t n = (size_t)nd * nb;
  hipLaunchKernelGGL(k_munpack_eh, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nd, nb, (const ColInfo *)col, colS,
                     (const cd *)P, (cd *)dE, (cd *)dH, estride);
}
}
