/*
 * orc_bie.c -- oracle: single-object boundary-integral system.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 *   orc_cmatrix_domain  create_cmatrix_domain, solve_bieq2.c:98-184
 *   orc_build_A         create_Amatrix_csr_dac, solve_bieq2.c:209-341 (dense)
 *   orc_zgetrf/zgetrs   LAPACK zgetrf + zgetrs semantics (the reference
 *                       calls MKL zgetrf+zgetri, solve_bieq2.c:389-399, and
 *                       multiplies by the inverse, mo_solve.c:246-274; a
 *                       solve with the LU factors is the same linear map)
 *   orc_build_B         create_Bmatrix_dac, mo_solve.c:137-244
 *   orc_store_X         tmatrix_bd_store, mo_solve.c:457-545
 *   orc_solve_EH        solve_eh_bv + dudt_fcm + inv_matrix_33, mo_solve.c:547-661
 *   orc_solve_dEH       solve_deh_bv (zgesv, 6 right-hand sides), mo_solve.c:663-789
 *
 * Parity: the assembled G/H satisfy the v1.dat fixture's stored
 * G dE = H E and H U = G dU identities (both domains, both objects), the
 * dG/dH/dF rows of triangle nodes reproduce the fixture's stored boundary
 * E and H, and the dense solve is checked against numpy's LAPACK.
 */
#include "orc.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void *orc_xcalloc(size_t n, size_t sz)
{
  void *p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "oracle: out of memory (%zu x %zu)\n", n, sz); exit(1); }
  return p;
}

/* Assemble the padded N x N (N = 4 Ne_d) matrices of domain did.  Row
 * (t,tn) = node slot tn of the t-th element seen by the domain, column
 * (s,i) = shape function i of the s-th element.  Diagonal H gets 1+F on the
 * open domain and F inside (solid-angle term). */
int orc_cmatrix_domain(const orc_obj *o, int did, int type, zc *G, zc *H, zc *dG, zc *dH, double *dF, int *gh_count)
{
  return orc_cmatrix_rows(o, did, type, 1, o->dNe[did] + 1, G, H, dG, dH, dF, gh_count);
}

/* The rows of elements t0 <= t < t1 only (row (t - t0) * 4 + tn of the
 * outputs); the full matrix is t0 = 1, t1 = Ne + 1.  Used to time a bounded
 * sample of a large assembly. */
int orc_cmatrix_rows(const orc_obj *o, int did, int type, int t0, int t1, zc *G, zc *H, zc *dG, zc *dH, double *dF,
                     int *gh_count)
{
  size_t Ne = (size_t)o->dNe[did], N = 4 * Ne;
  zc kc = o->kn[did];
  int ngh = 0, nbad = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : ngh, nbad)
  for (size_t t = (size_t)t0; t < (size_t)t1; t++) {
    int td = o->sid[did][t];
    int tl = orc_is_tri(o, td) ? 3 : 4;
    for (int tn = 0; tn < tl; tn++) {
      double vtz[3], vte[3];
      orc_tz_te_node(vtz, vte, o, td, tn);
      double F = 0.0, dFz = 0.0, dFe = 0.0;
      size_t row = (t - (size_t)t0) * 4 + tn, drow = (t - 1) * 4 + tn;
      for (size_t s = 1; s <= Ne; s++) {
        int sd = o->sid[did][s];
        zc CC[9], dCz[9], dCe[9];
        int rc = orc_dcoef_bd_node_t2(CC, dCz, dCe, td, tn, vtz, vte, sd, kc, type, o);
        if (rc == 1) ngh++;
        if (rc < 0) nbad++;
        for (int i = 0; i < 4; i++) {
          size_t col = (s - 1) * 4 + i;
          if (G) G[row * N + col] = CC[i];
          if (H) H[row * N + col] = CC[i + 4];
          if (dG) { dG[(row * 2 + 0) * N + col] = dCz[i]; dG[(row * 2 + 1) * N + col] = dCe[i]; }
          if (dH) { dH[(row * 2 + 0) * N + col] = dCz[i + 4]; dH[(row * 2 + 1) * N + col] = dCe[i + 4]; }
        }
        F += creal(CC[8]); dFz += creal(dCz[8]); dFe += creal(dCe[8]);
      }
      double diag = did == 0 ? 1.0 + F : F;
      if (H) H[row * N + drow] += diag;
      if (dF) { dF[row * 3 + 0] = diag; dF[row * 3 + 1] = dFz; dF[row * 3 + 2] = dFe; }
    }
  }
  if (gh_count) *gh_count = ngh;
  return nbad;
}

/* ---- dense complex LU with partial pivoting (row-major) ---- */
int orc_zgetrf(int n, zc *A, int lda, int *piv)
{
  const int nb = 48;
  int info = 0;
  double *a = (double *)A;   /* interleaved re/im */
  for (int k0 = 0; k0 < n; k0 += nb) {
    int kb = n - k0 < nb ? n - k0 : nb;
    /* panel: columns k0 .. k0+kb-1, all rows below */
    for (int k = k0; k < k0 + kb; k++) {
      int p = k; double best = -1.0;
      for (int i = k; i < n; i++) { double m = cabs(A[(size_t)i * lda + k]); if (m > best) { best = m; p = i; } }
      piv[k] = p;
      if (best == 0.0) { if (!info) info = k + 1; continue; }
      if (p != k) for (int j = 0; j < n; j++) { zc t = A[(size_t)k * lda + j]; A[(size_t)k * lda + j] = A[(size_t)p * lda + j]; A[(size_t)p * lda + j] = t; }
      zc inv = 1.0 / A[(size_t)k * lda + k];
      for (int i = k + 1; i < n; i++) A[(size_t)i * lda + k] *= inv;
      /* rank-1 update restricted to the panel columns */
#pragma omp parallel for schedule(static) if (n - k > 256)
      for (int i = k + 1; i < n; i++) {
        zc l = A[(size_t)i * lda + k];
        if (l == 0.0) continue;
        for (int j = k + 1; j < k0 + kb; j++) A[(size_t)i * lda + j] -= l * A[(size_t)k * lda + j];
      }
    }
    int c1 = k0 + kb;
    if (c1 >= n) break;
    /* U12 = L11^{-1} A12 (unit lower triangular forward substitution) */
    for (int k = k0; k < c1; k++)
      for (int i = k + 1; i < c1; i++) {
        zc l = A[(size_t)i * lda + k];
        if (l == 0.0) continue;
        for (int j = c1; j < n; j++) A[(size_t)i * lda + j] -= l * A[(size_t)k * lda + j];
      }
    /* A22 -= L21 U12, written on split re/im to avoid library complex calls */
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = c1; i < n; i++) {
      double *ri = a + 2 * ((size_t)i * lda);
      for (int k = k0; k < c1; k++) {
        double lr = ri[2 * k], li = ri[2 * k + 1];
        if (lr == 0.0 && li == 0.0) continue;
        const double *rk = a + 2 * ((size_t)k * lda);
        for (int j = c1; j < n; j++) {
          double ur = rk[2 * j], ui = rk[2 * j + 1];
          ri[2 * j] -= lr * ur - li * ui;
          ri[2 * j + 1] -= lr * ui + li * ur;
        }
      }
    }
  }
  return info;
}

void orc_zgetrs(int n, int nrhs, const zc *A, int lda, const int *piv, zc *B, int ldb)
{
  for (int k = 0; k < n; k++)
    if (piv[k] != k) for (int j = 0; j < nrhs; j++) { zc t = B[(size_t)k * ldb + j]; B[(size_t)k * ldb + j] = B[(size_t)piv[k] * ldb + j]; B[(size_t)piv[k] * ldb + j] = t; }
  for (int i = 0; i < n; i++)
    for (int k = 0; k < i; k++) {
      zc l = A[(size_t)i * lda + k];
      if (l != 0.0) for (int j = 0; j < nrhs; j++) B[(size_t)i * ldb + j] -= l * B[(size_t)k * ldb + j];
    }
  for (int i = n - 1; i >= 0; i--) {
    for (int k = i + 1; k < n; k++) {
      zc u = A[(size_t)i * lda + k];
      if (u != 0.0) for (int j = 0; j < nrhs; j++) B[(size_t)i * ldb + j] -= u * B[(size_t)k * ldb + j];
    }
    zc inv = 1.0 / A[(size_t)i * lda + i];
    for (int j = 0; j < nrhs; j++) B[(size_t)i * ldb + j] *= inv;
  }
}

/* unit node normal of a slot, d3b1_elem.c:38-44 */
static void orc_node_normal(double *vn, const orc_obj *o, int s, int l)
{
  for (int d = 0; d < 3; d++) vn[d] = ORC_WEN(o, s, l, d);
  orc_unit(vn);
}

/* boundary-condition coefficients between main domain mdid and sub domain sdid */
static void orc_bc(const orc_obj *o, int mdid, int sdid, zc *delt, zc *beps, zc *sigm)
{
  zc k1 = o->kn[mdid], k2 = o->kn[sdid];
  *delt = k1 * k1 - k2 * k2;
  *beps = k1 * k1 / (k2 * k2);
  *sigm = 1.0 - *beps;
}

/* Dense 8NN matrix, row-major.  Row order: domain, component (Ux,Uy,Uz,phi),
 * element, node slot; column blocks: U (0-3), dU/dn (4-7), each NN long. */
void orc_build_A(const orc_obj *o, const orc_sys *sys, zc *A)
{
  size_t nn = (size_t)o->NN, na = 8 * nn, row = 0;
  memset(A, 0, sizeof(zc) * na * na);
  for (int did = 0; did <= o->MN; did++) {
    size_t Ne = (size_t)o->dNe[did], N = 4 * Ne;
    for (int cid = 0; cid < 4; cid++)
      for (size_t t = 1; t <= Ne; t++) {
        int td = o->sid[did][t];
        for (int tn = 0; tn < 4; tn++) {
          if (tn == 3 && orc_is_tri(o, td)) continue;
          const zc *gr = sys->G[did] + ((t - 1) * 4 + tn) * N, *hr = sys->H[did] + ((t - 1) * 4 + tn) * N;
          zc *ar = A + row * na;
          for (size_t s = 1; s <= Ne; s++) {
            int sd = o->sid[did][s], asd = sd < 0 ? -sd : sd;
            int mdid = o->md[asd], sdid = o->sd[asd], nv = orc_is_tri(o, sd) ? 3 : 4;
            for (int l = 0; l < nv; l++) {
              size_t c = (size_t)ORC_ENI(o, asd, l);
              zc g = gr[(s - 1) * 4 + l], h = hr[(s - 1) * 4 + l];
              if (did == mdid) { ar[nn * cid + c] = h; ar[nn * (cid + 4) + c] = -g; }
              else {
                zc delt, beps, sigm; double vn[3];
                orc_bc(o, mdid, sdid, &delt, &beps, &sigm);
                orc_node_normal(vn, o, sd < 0 ? -sd : sd, l);
                if (cid < 3) {
                  ar[nn * cid + c] = h;
                  ar[nn * 3 + c] = delt * vn[cid] * g;
                  ar[nn * (cid + 4) + c] = g;
                } else {
                  for (int q = 0; q < 3; q++) ar[nn * q + c] = sigm * vn[q] * g;
                  ar[nn * 3 + c] = h;
                  ar[nn * 7 + c] = beps * g;
                }
              }
            }
          }
          row++;
        }
      }
  }
}

/* right-hand side from the incident field on the sub-domain rows */
void orc_build_B(const orc_obj *o, const orc_sys *sys, zc *B)
{
  size_t row = 0;
  for (int did = 0; did <= o->MN; did++) {
    size_t Ne = (size_t)o->dNe[did], N = 4 * Ne;
    for (int cid = 0; cid < 4; cid++)
      for (size_t t = 1; t <= Ne; t++) {
        int td = o->sid[did][t];
        for (int tn = 0; tn < 4; tn++) {
          if (tn == 3 && orc_is_tri(o, td)) continue;
          const zc *gr = sys->G[did] + ((t - 1) * 4 + tn) * N, *hr = sys->H[did] + ((t - 1) * 4 + tn) * N;
          zc b = 0.0;
          for (size_t s = 1; s <= Ne; s++) {
            int sd = o->sid[did][s], asd = sd < 0 ? -sd : sd;
            int mdid = o->md[asd], sdid = o->sd[asd], nv = orc_is_tri(o, sd) ? 3 : 4;
            if (did == mdid || mdid != 0) continue;
            zc delt, beps, sigm;
            orc_bc(o, mdid, sdid, &delt, &beps, &sigm);
            for (int l = 0; l < nv; l++) {
              double vn[3];
              orc_node_normal(vn, o, asd, l);
              zc g = gr[(s - 1) * 4 + l], h = hr[(s - 1) * 4 + l];
              const zc *ui = &o->Ui[((size_t)asd * 4 + l) * 4], *dui = &o->dUi[((size_t)asd * 4 + l) * 4];
              if (cid < 3) b += -h * ui[cid] - g * dui[cid] - g * ui[3] * delt * vn[cid];
              else b += -h * ui[3] - sigm * g * (vn[0] * ui[0] + vn[1] * ui[1] + vn[2] * ui[2]) - beps * g * dui[3];
            }
          }
          B[row++] = b;
        }
      }
  }
}

/* solution vector -> per-domain boundary values (U, dU/dn on each side) */
void orc_store_X(orc_obj *o, const zc *X)
{
  size_t nn = (size_t)o->NN;
  for (int d = 0; d <= o->MN; d++)
    for (int s = 1; s <= o->dNe[d]; s++) {
      int sd = o->sid[d][s], asd = sd < 0 ? -sd : sd;
      int mdid = o->md[asd], sdid = o->sd[asd], nv = orc_is_tri(o, sd) ? 3 : 4;
      zc *U = o->U[d] + (size_t)s * 16, *dU = o->dU[d] + (size_t)s * 16;
      for (int l = 0; l < nv; l++) {
        size_t c = (size_t)ORC_ENI(o, asd, l);
        if (d == mdid) {
          for (int i = 0; i < 4; i++) { U[l * 4 + i] = X[nn * i + c]; dU[l * 4 + i] = X[nn * (i + 4) + c]; }
        } else {
          zc delt, beps, sigm, tU[4], tUi[3] = {0, 0, 0};
          double n[3];
          orc_bc(o, mdid, sdid, &delt, &beps, &sigm);
          orc_node_normal(n, o, asd, l);
          tU[3] = X[nn * 3 + c];
          for (int i = 0; i < 3; i++) {
            tU[i] = X[nn * i + c];
            U[l * 4 + i] = tU[i];
            dU[l * 4 + i] = -(X[nn * (i + 4) + c] + delt * tU[3] * n[i]);
            if (mdid == 0) {
              tUi[i] = o->Ui[((size_t)asd * 4 + l) * 4 + i];
              U[l * 4 + i] += tUi[i];
              dU[l * 4 + i] += -o->dUi[((size_t)asd * 4 + l) * 4 + i];
            }
          }
          U[l * 4 + 3] = tU[3];
          dU[l * 4 + 3] = -(beps * X[nn * 7 + c] + sigm * (n[0] * tU[0] + n[1] * tU[1] + n[2] * tU[2]));
          if (mdid == 0) dU[l * 4 + 3] += -sigm * (n[0] * tUi[0] + n[1] * tUi[1] + n[2] * tUi[2]);
        }
      }
    }
}

int orc_create_matrix(orc_obj *o, orc_sys *sys, int type)
{
  memset(sys, 0, sizeof *sys);
  if (o->MN > 1) { fprintf(stderr, "oracle: create_matrix restated for MN = 1 only\n"); return -1; }
  int bad = 0;
  for (int d = 0; d <= o->MN; d++) {
    size_t N = 4 * (size_t)o->dNe[d];
    sys->G[d] = orc_xcalloc(N * N, sizeof(zc)); sys->H[d] = orc_xcalloc(N * N, sizeof(zc));
    sys->dG[d] = orc_xcalloc(2 * N * N, sizeof(zc)); sys->dH[d] = orc_xcalloc(2 * N * N, sizeof(zc));
    sys->dF[d] = orc_xcalloc(3 * N, sizeof(double));
    bad += orc_cmatrix_domain(o, d, type, sys->G[d], sys->H[d], sys->dG[d], sys->dH[d], sys->dF[d], NULL);
  }
  sys->nn = o->NN; sys->na = 8 * o->NN;
  size_t na = (size_t)sys->na;
  sys->lu = orc_xcalloc(na * na, sizeof(zc));
  sys->piv = orc_xcalloc(na, sizeof(int));
  orc_build_A(o, sys, sys->lu);
  int info = orc_zgetrf(sys->na, sys->lu, sys->na, sys->piv);
  if (info) { fprintf(stderr, "oracle: singular A (info %d)\n", info); return -2; }
  return bad;
}

void orc_sys_free(orc_sys *sys)
{
  free(sys->lu); free(sys->piv);
  for (int d = 0; d < 2; d++) { free(sys->G[d]); free(sys->H[d]); free(sys->dG[d]); free(sys->dH[d]); free(sys->dF[d]); }
  memset(sys, 0, sizeof *sys);
}

void orc_solve_mEMP(orc_obj *o, orc_sys *sys)
{
  zc *B = orc_xcalloc((size_t)sys->na, sizeof(zc));
  orc_build_B(o, sys, B);
  orc_zgetrs(sys->na, 1, sys->lu, sys->na, sys->piv, B, 1);
  orc_store_X(o, B);
  free(B);
}

/* inverse of the 3x3 matrix with rows (n, t_zeta, t_eta), mo_solve.c:618-639 */
static void orc_inv33(double *m, const double *n, const double *a, const double *b)
{
  double det = n[0] * a[1] * b[2] + n[1] * a[2] * b[0] + n[2] * a[0] * b[1] - (n[2] * a[1] * b[0] + n[1] * a[0] * b[2] + n[0] * a[2] * b[1]);
  double id = 1.0 / det;
  m[0] = a[1] * b[2] - a[2] * b[1]; m[1] = -n[1] * b[2] + n[2] * b[1]; m[2] = n[1] * a[2] - n[2] * a[1];
  m[3] = -a[0] * b[2] + a[2] * b[0]; m[4] = n[0] * b[2] - n[2] * b[0]; m[5] = -n[0] * a[2] + n[2] * a[0];
  m[6] = a[0] * b[1] - a[1] * b[0]; m[7] = -n[0] * b[1] + n[1] * b[0]; m[8] = n[0] * a[1] - n[1] * a[0];
  for (int i = 0; i < 9; i++) m[i] *= id;
}

/* boundary E, H from U, dU/dn and the tangential derivatives of U obtained
 * from the derivative BIE rows; E = -grad(phi) + A, H = lambda0/(2 pi i) curl A */
void orc_solve_EH(orc_obj *o, orc_sys *sys)
{
  zc ch = o->bm.lambda0 / (2.0 * M_PI * I);
  for (int d = 0; d <= o->MN; d++) {
    size_t Ne = (size_t)o->dNe[d], N = 4 * Ne;
#pragma omp parallel for schedule(dynamic)
    for (size_t t = 1; t <= Ne; t++) {
      int at = o->sid[d][t], atd = at < 0 ? -at : at;
      double sig = at < 0 ? -1.0 : 1.0;
      for (int tn = 0; tn < 4; tn++) {
        if (tn == 3 && orc_is_tri(o, atd)) continue;
        size_t row = (t - 1) * 4 + tn;
        const zc *gz = sys->dG[d] + (row * 2) * N, *ge = gz + N, *hz = sys->dH[d] + (row * 2) * N, *he = hz + N;
        const double *f = sys->dF[d] + row * 3;
        zc dz[4] = {0, 0, 0, 0}, de[4] = {0, 0, 0, 0};
        for (size_t s = 1; s <= Ne; s++)
          for (int sn = 0; sn < 4; sn++) {
            const zc *Us = o->U[d] + s * 16 + sn * 4, *dUs = o->dU[d] + s * 16 + sn * 4;
            size_t c = (s - 1) * 4 + sn;
            for (int i = 0; i < 4; i++) { dz[i] += gz[c] * dUs[i] - hz[c] * Us[i]; de[i] += ge[c] * dUs[i] - he[c] * Us[i]; }
          }
        const zc *Ut = o->U[d] + t * 16 + tn * 4, *dUt = o->dU[d] + t * 16 + tn * 4;
        for (int i = 0; i < 4; i++) { dz[i] = (dz[i] - Ut[i] * f[1]) / f[0]; de[i] = (de[i] - Ut[i] * f[2]) / f[0]; }
        double vtz[3], vte[3], vn[3], a[9];
        orc_tz_te_node(vtz, vte, o, atd, tn);
        for (int q = 0; q < 3; q++) vn[q] = sig * ORC_WEN(o, atd, tn, q);
        orc_unit(vn);
        orc_inv33(a, vn, vtz, vte);
        zc *E = o->E[d] + t * 12 + tn * 3, *H = o->H[d] + t * 12 + tn * 3;
        /* gradient of component c: a * (dU/dn, dU/dtz, dU/dte) */
#define GRD(c, i) (dUt[c] * a[(i) * 3 + 0] + dz[c] * a[(i) * 3 + 1] + de[c] * a[(i) * 3 + 2])
        for (int i = 0; i < 3; i++) E[i] = -GRD(3, i) + Ut[i];
        H[0] = ch * (GRD(2, 1) - GRD(1, 2));
        H[1] = ch * (GRD(0, 2) - GRD(2, 0));
        H[2] = ch * (GRD(1, 0) - GRD(0, 1));
#undef GRD
      }
    }
  }
}

/* normal derivatives of E and H from G [dE|dH] = H [E|H] per domain */
int orc_solve_dEH(orc_obj *o, orc_sys *sys)
{
  for (int d = 0; d <= o->MN; d++) {
    size_t Ne = (size_t)o->dNe[d], N = 4 * Ne;
    if (!Ne) continue;
    size_t nn = 0;
    for (size_t s = 1; s <= Ne; s++) nn += orc_is_tri(o, o->sid[d][s]) ? 3 : 4;
    size_t *map = orc_xcalloc(nn, sizeof(size_t));   /* compact index -> padded (s-1)*4+n */
    size_t c = 0;
    for (size_t s = 1; s <= Ne; s++) { int nv = orc_is_tri(o, o->sid[d][s]) ? 3 : 4; for (int n = 0; n < nv; n++) map[c++] = (s - 1) * 4 + n; }
    zc *A = orc_xcalloc(nn * nn, sizeof(zc)), *B = orc_xcalloc(nn * 6, sizeof(zc));
    int *piv = orc_xcalloc(nn, sizeof(int));
    for (size_t r = 0; r < nn; r++) {
      const zc *gr = sys->G[d] + map[r] * N, *hr = sys->H[d] + map[r] * N;
      for (size_t q = 0; q < nn; q++) {
        size_t pc = map[q], s = pc / 4 + 1, sn = pc % 4;
        A[r * nn + q] = gr[pc];
        for (int i = 0; i < 3; i++) {
          B[r * 6 + i] += hr[pc] * o->E[d][s * 12 + sn * 3 + i];
          B[r * 6 + 3 + i] += hr[pc] * o->H[d][s * 12 + sn * 3 + i];
        }
      }
    }
    int info = orc_zgetrf((int)nn, A, (int)nn, piv);
    if (info) { free(A); free(B); free(piv); free(map); return info; }
    orc_zgetrs((int)nn, 6, A, (int)nn, piv, B, 6);
    for (size_t q = 0; q < nn; q++) {
      size_t pc = map[q], s = pc / 4 + 1, sn = pc % 4;
      for (int i = 0; i < 3; i++) { o->dE[d][s * 12 + sn * 3 + i] = B[q * 6 + i]; o->dH[d][s * 12 + sn * 3 + i] = B[q * 6 + 3 + i]; }
    }
    free(A); free(B); free(piv); free(map);
  }
  return 0;
}
