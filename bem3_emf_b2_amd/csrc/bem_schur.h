// bem_schur.h -- the one-medium block elimination of the boundary-integral
// system in either precision, and the FP64 residual of the full system
// (bem_schur.hip).
#pragma once
#include "bem_dev.h"

// Factors of the block elimination of one object with one interior medium:
// domain-0 rows H0 U_c - G0 Q_c = r0_c, domain-1 rows (sub side of every
// element) H1 U_c + delta G1 N_c U_3 + G1 Q_c = r1_c (c < 3) and
// sigma G1 sum_c N_c U_c + H1 U_3 + beta G1 Q_3 = r1_3 (N_c: the normal's
// component c at each column).  With Z = G0^{-1} H0, M = H1 + G1 Z,
// W = M^{-1} G1 and S = H1 + beta G1 Z - sigma delta G1 (W o n n^T).
// S is formed only on demand (s_lu): until then Sf holds Sp = H1 + beta G1 Z
// and the S systems are solved by GMRES with the operator
// S x = Sp x - sigma delta G1 sum_c N_c M^{-1} G1 N_c x (well conditioned:
// cond(S) is about 2 on sphere meshes of 960 and 3840 nodes, where GMRES
// converges in 15 steps at either size).
template <class Zt> struct SchurF {
  int n;                                   // NN (nodes = columns of each domain)
  size_t ld;                               // row stride of the matrices
  const Zt *G0f, *Zm, *Mf, *Sf, *G1;       // LU(G0), Z, LU(M), LU(S) or Sp, G1
  const int *piv0, *pivM, *pivS;
  const ColInfo *col1;                     // domain-1 columns (unit normals)
  Zt beps, sigm, delt;
  bool s_lu;                               // Sf holds LU(S) (else Sp, GMRES)
  double gm_tol;                           // GMRES relative residual target
};

// Krylov dimension of one GMRES cycle and the work (in complex values of
// the precision in use) of schur_apply for n = NN
#define SCHUR_GM_M 60
static inline size_t schur_work(size_t n) { return (27 + SCHUR_GM_M + 1 + 9) * n + 512; }

// X = A^{-1} rhs through the factors.  rhs and X use the 8NN layout of the
// system (rows: domain 0 components 0..3, then domain 1; unknowns U_0..U_3,
// Q_0..Q_3, NN each).  r0nz = false asserts that the domain-0 rows of rhs
// are zero (the case of every incident-field right-hand side).  work holds
// schur_work(NN) values.  Returns 0, or BH_ECONV if GMRES on S does not
// reach gm_tol within 20 cycles.  gm_steps (optional) receives the GMRES
// steps taken.
int schur_apply(const SchurF<cd> &f, const cd *rhs, bool r0nz, cd *X, cd *work, hipStream_t st, int *gm_steps);
int schur_apply(const SchurF<cf> &f, const cf *rhs, bool r0nz, cf *X, cf *work, hipStream_t st, int *gm_steps);

// complex-FP32 factors from the FP64 matrices of both domains: LU(G0), Z,
// LU(M), G1 and Sp = H1 + beta G1 Z (S itself by schur_form_s).  Returns 0,
// or the code of the LU (-1 scratch, -2 barrier).
int schur_factor_c(int n, size_t ld, const cd *G0, const cd *H0, const cd *G1, const cd *H1, cd beps, cf *G0f, cf *Zm,
                   cf *Mf, cf *Sf, cf *G1c, int *piv0, int *pivM, int *info, hipStream_t st);

// the same factors for the low-memory form, in place over the complex-FP32
// matrices: G0 -> LU(G0), H0 -> Z, H1 -> LU(M); P (NN^2 values) receives
// Sp; G1 is kept.  Returns 0 or the code of the LU.
int schur_factor_c_inplace(int n, size_t ld, cf *G0, cf *H0, cf *G1, cf *H1, cd beps, cf *P, int *piv0, int *pivM,
                           int *info, hipStream_t st);

// S from Sp in place and its LU: W = M^{-1} G1 (Wt: NN^2 scratch),
// S = Sp - sigma delta G1 (W o n n^T), then LU(S) into Sf / pivS
int schur_form_s(const SchurF<cd> &f, cd *Sf, int *pivS, cd *Wt, int *info, hipStream_t st);
int schur_form_s(const SchurF<cf> &f, cf *Sf, int *pivS, cf *Wt, int *info, hipStream_t st);

// R = B - A X over the rows of one domain (rows rowBase + c nd + r), from
// the domain's FP64 G, H and its column classes; work holds 12 nd values
void bem_resid_domain(const cd *G, const cd *H, int nd, size_t ld, const ColInfo *col, const cd *X, size_t NN,
                      size_t rowBase, const cd *B, cd *R, cd *work, hipStream_t st);
// the same in two parts, for row blocks of G, H: the column vectors of X
// (work[0, 8 nd)), then rows r0 .. r0+nrows from the block (work[8 nd,
// 8 nd + 4 nrows))
void bem_resid_colvec(const ColInfo *col, int nd, const cd *X, size_t NN, cd *work, hipStream_t st);
void bem_resid_rows(const cd *G, const cd *H, int nrows, int nd, size_t ld, int r0, size_t rowBase, const cd *B, cd *R,
                    cd *work, hipStream_t st);
void bem_resid_fin(const cd *B, const cd *Y, int nd, size_t rowBase, cd *R, hipStream_t st);
extern "C" size_t bem_skinny_batch_bytes(int nj);
extern "C" int bem_zgemm_skinny4_batch(const void *jobs, int nj, void *dev, void *host_tab, hipStream_t st);

// Euclidean norm of n complex values (synchronises the stream); tmp holds
// 256 doubles of device memory
double bem_nrm2(const cd *x, size_t n, double *tmp, hipStream_t st);

// the same, enqueued only: |x|^2 (not its root) into the device double *out
void bem_nrm2sq_to(const cd *x, size_t n, double *tmp, double *out, hipStream_t st);

// schur_apply (r0nz, S formed) of N objects together: every solve with a
// factor is one batched launch per triangle over all objects (f[o], rhs[o],
// X[o], work w[o] of schur_work(n) values per object).  Returns 0,
// BH_ESTATE if an object's S is not formed, or BH_EHIP.
int schur_apply_batch(int N, const SchurF<cf> *f, const cf *const *rhs, cf *const *X, cf *const *w, hipStream_t st);

// elementwise conversions: dst = src rounded to FP32; X += d widened to FP64
void bem_cvt_z2c(const cd *src, cf *dst, size_t n, hipStream_t st);
void bem_acc_c2z(cd *X, const cf *d, size_t n, hipStream_t st);

// many right-hand sides (nb, beam b at + b 8NN) with S formed: work holds
// schur_work_multi(NN, nb) values; BH_ESTATE if S is not formed
static inline size_t schur_work_multi(size_t n, size_t nb) { return 27 * n * nb + 512; }
// explicit inverses of the factors (n x n, row stride n): with many
// right-hand sides each solve with LU(G0), LU(M), LU(S) becomes one product
// (one launch in place of a chain of blocked triangular solves); null
// pointers: solve with the factors
template <class Zt> struct SchurInv {
  const Zt *G0i, *Mi, *Si;
};
int schur_apply_multi(const SchurF<cf> &f, const cf *rhs, bool r0nz, cf *X, cf *work, int nb, hipStream_t st,
                      const SchurInv<cf> *inv = nullptr);
int schur_apply_multi(const SchurF<cd> &f, const cd *rhs, bool r0nz, cd *X, cd *work, int nb, hipStream_t st,
                      const SchurInv<cd> *inv = nullptr);
// the inverses from the factors (S formed): each LU^{-1} I by the blocked
// triangular solves with n right-hand sides
int schur_inverses(const SchurF<cf> &f, cf *G0i, cf *Mi, cf *Si, hipStream_t st);
// R_b = B_b - A X_b over the rows of one domain for nb beams; work holds
// 12 nd nb values
// B rows of one domain for nb beams (k_build_B of every beam at once): the
// incident data of the open-domain sub-side columns (beam b's at Ui + b
// istride) as two operands a, c (nd x 4 nb), B = -(H a + G c); work holds
// 12 nd nb values
void bem_build_B_multi(const cd *G, const cd *H, int nd, size_t ld, const ColInfo *col, const cd *Ui, const cd *dUi,
                       size_t istride, size_t NN, size_t rowBase, cd *B, cd *w, int nb, hipStream_t st);
void bem_resid_domain_multi(const cd *G, const cd *H, int nd, size_t ld, const ColInfo *col, const cd *X, size_t NN,
                            size_t rowBase, const cd *B, cd *R, cd *work, int nb, hipStream_t st);
// |x_b|^2 of nb vectors of n values into out[nb] (device)
void bem_mnrm2(const cd *x, size_t n, int nb, double *out, hipStream_t st);
