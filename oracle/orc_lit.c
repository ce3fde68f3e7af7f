/*
 * orc_lit.c -- oracle: linear-triangle element integrals (G, H, F and their
 * directional derivatives).  TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Reference element: equilateral triangle with vertices (-1/2,-sqrt3/2),
 * (1,0), (-1/2,sqrt3/2); node slots at the 4-point rule abscissae
 * (d3b1_elem.c:688-711, setup2.c:735-737).
 *
 * Parity: pinned by tests against (a) the v1.dat fixture (28 triangles of
 * cone_m1_p5, 2 objects): fields/forces of example1.log and the boundary
 * E/H that the reference derived from these derivative coefficients, (b) a
 * finite-difference check of every d/dv term against the undifferentiated
 * kernel, (c) partition of unity / closed forms (sum of G over shape
 * functions vs brute-force integration, solid-angle values of F).
 */
#include "orc.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#define SQ3 1.732050807568877293528
#define I_EP 3.978873577297383394222e-2     /* 1/(8 pi)            */
#define I_SPSQ3 3.062938307898844719507e-2  /* 1/(6 pi sqrt(3))    */

/* interpolation functions at the node slots, d3b1_elem.c:700-711 */
double orc_lit_Nn(int n, double z, double e)
{
  const double ia = 2.5;  /* 0.5 / P34_N0 */
  switch (n) {
  case 0: return (1.0 - ia * (z + SQ3 * e)) / 3.0;
  case 1: return (1.0 + 2.0 * ia * z) / 3.0;
  default: return (1.0 - ia * (z - SQ3 * e)) / 3.0;
  }
}

/* d N_n / d r along (cos th, sin th), lit_dcoef.c:844-855 */
static double orc_lit_dNn(int n, double c, double s)
{
  const double ia = 2.5;
  switch (n) {
  case 0: return -2.0 / 3.0 * ia * (c + SQ3 * s);
  case 1: return 4.0 / 3.0 * ia * c;
  default: return -2.0 / 3.0 * ia * (c - SQ3 * s);
  }
}

static inline zc orc_eikr(zc k, double r)
{
  if (cimag(k) == 0.0) { double a = creal(k) * r; return cos(a) + I * sin(a); }
  return cexp(I * k * r);
}

/* integrand values at one source point.  vR = r_source - r_target,
 * W = unnormalised surface normal.  Kernel forms: lit_dcoef_v2.c:381-413 */
typedef struct { zc g, h, dg[2], dh[2]; double f, df[2]; } kv_t;

static inline void orc_kernel(kv_t *kv, const double *vR, const double *W, zc k, const double *const *v, int nd)
{
  double R = orc_norm(vR), iR = 1.0 / R;
  zc kR = k * R, cb = I * kR - 1.0;
  zc ce = orc_eikr(k, R);
  double rdW = orc_dot(vR, W) * iR;
  kv->g = ce * iR;
  kv->h = cb * ce * rdW * iR * iR;
  kv->f = rdW * iR * iR;
  for (int d = 0; d < nd; d++) {
    double rdv = orc_dot(vR, v[d]) * iR, Wdv = orc_dot(W, v[d]);
    kv->dg[d] = cb * ce * iR * iR * rdv;
    kv->dh[d] = ce * iR * iR * iR * ((kR * kR + 3.0 * cb) * rdv * rdW - cb * Wdv);
    kv->df[d] = iR * iR * iR * (3.0 * rdv * rdW - Wdv);
  }
}

/* 3-part Gauss-Legendre rule: the triangle is cut into three quadrilaterals
 * mapped from [-1,1]^2 (lit_coef.c:144-179).  Tables are built once. */
typedef struct { int np; double *z, *e, *w; } orc_trule;
static orc_trule TR_GL, TR_GH;

static void orc_trule_build(orc_trule *t, int n, const double *x, const double *w)
{
  t->np = 3 * n * n;
  t->z = malloc(sizeof(double) * t->np); t->e = malloc(sizeof(double) * t->np); t->w = malloc(sizeof(double) * t->np);
  int p = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double a = x[i], b = x[j], wt = (1.0 + 0.25 * a + 0.25 * b) * w[i] * w[j];
      double zz[3] = {0.125 * (3.0 + 2.0 * a + 2.0 * b + a * b), 0.0625 * (-3.0 + a - 5.0 * b - a * b), 0.0625 * (-3.0 - 5.0 * a + b - a * b)};
      double ee[3] = {0.125 * SQ3 * (-a + b), 0.0625 * SQ3 * (3.0 + 3.0 * a + b + a * b), 0.0625 * SQ3 * (-3.0 - a - 3.0 * b - a * b)};
      for (int q = 0; q < 3; q++) { t->z[p] = zz[q]; t->e[p] = ee[q]; t->w[p] = wt; p++; }
    }
}

static const orc_trule *orc_lit_rule(int high)
{
  if (!TR_GL.np) {
#pragma omp critical(orc_trule_init)
    {
      if (!TR_GL.np) {
        orc_trule_build(&TR_GH, ORC_GHN, ORCQ.xhi, ORCQ.whi);
        orc_trule_build(&TR_GL, ORC_GLN, ORCQ.xli, ORCQ.wli);
      }
    }
  }
  return high ? &TR_GH : &TR_GL;
}

/* sign convention of the sub-domain side: H, F and their derivatives change
 * sign, G does not (d3b1_elem.c:728-746) */
static void orc_lit_flip(zc *C) { C[4] = -C[4]; C[5] = -C[5]; C[6] = -C[6]; C[8] = -C[8]; }

static void orc_elem(const orc_obj *o, int s, double cr[3][3], double W[3], double *aW)
{
  for (int d = 0; d < 3; d++) { cr[d][0] = ORC_CR(o, s, d, 0); cr[d][1] = ORC_CR(o, s, d, 1); cr[d][2] = ORC_CR(o, s, d, 2); W[d] = ORC_CW(o, s, d, 0); }
  *aW = orc_norm(W);
}

/* general driver: accumulate over a point set.  mode 0: 4-point node rule
 * (nodes carry their own shape function exactly, centroid shared 1/3 each),
 * 1: 7-point rule, 2: 3-part GL (GLN), 3: 3-part GL (GHN).  Outputs are the
 * 9-slot arrays of the reference (G 0-2, H 4-6, F 8). */
static void orc_lit_run(zc *CC, zc *dC0, zc *dC1, const double *rt, const double *v0, const double *v1,
                        int seid, zc k, const orc_obj *o, int mode)
{
  int s = seid < 0 ? -seid : seid;
  double cr[3][3], W[3], aW;
  orc_elem(o, s, cr, W, &aW);
  const double *vv[2] = {v0, v1};
  int nd = v1 ? 2 : (v0 ? 1 : 0);
  zc acc[3][5];           /* per shape fn: g, h, dg0, dh0, dg1/dh1 packed below */
  zc accd[3][4];
  double f = 0.0, df[2] = {0.0, 0.0};
  for (int a = 0; a < 3; a++) { for (int b = 0; b < 5; b++) acc[a][b] = 0.0; for (int b = 0; b < 4; b++) accd[a][b] = 0.0; }
  double scale;
  int np;
  const orc_trule *tr = NULL;
  if (mode == 0) { np = 4; scale = I_EP; }
  else if (mode == 1) { np = 7; scale = I_EP; }
  else { tr = orc_lit_rule(mode == 3); np = tr->np; scale = 0.0625 * SQ3 * I_SPSQ3; }
  for (int p = 0; p < np; p++) {
    double z, e, w, N[3], vR[3];
    if (mode == 0) {
      z = ORCQ.zt34[p]; e = ORCQ.et34[p]; w = ORCQ.wt34[p];
      for (int d = 0; d < 3; d++) vR[d] = ORC_REN(o, s, p, d) - rt[d];
      if (p < 3) { N[0] = N[1] = N[2] = 0.0; N[p] = 1.0; }
      else for (int a = 0; a < 3; a++) N[a] = orc_lit_Nn(a, z, e);
    } else {
      if (mode == 1) { z = ORCQ.zt37[p]; e = ORCQ.et37[p]; w = ORCQ.wt37[p]; }
      else { z = tr->z[p]; e = tr->e[p]; w = tr->w[p]; }
      for (int d = 0; d < 3; d++) vR[d] = cr[d][0] + cr[d][1] * z + cr[d][2] * e - rt[d];
      for (int a = 0; a < 3; a++) N[a] = orc_lit_Nn(a, z, e);
    }
    kv_t kv;
    orc_kernel(&kv, vR, W, k, vv, nd);
    f += w * kv.f;
    for (int d = 0; d < nd; d++) df[d] += w * kv.df[d];
    for (int a = 0; a < 3; a++) {
      double wn = w * N[a];
      acc[a][0] += wn * kv.g; acc[a][1] += wn * kv.h;
      for (int d = 0; d < nd; d++) { accd[a][2 * d] += wn * kv.dg[d]; accd[a][2 * d + 1] += wn * kv.dh[d]; }
    }
  }
  for (int i = 0; i < 9; i++) { CC[i] = 0.0; if (dC0) dC0[i] = 0.0; if (dC1) dC1[i] = 0.0; }
  for (int a = 0; a < 3; a++) {
    CC[a] = acc[a][0] * scale * aW;
    CC[a + 4] = acc[a][1] * scale;
    if (nd > 0) { dC0[a] = -accd[a][0] * scale * aW; dC0[a + 4] = accd[a][1] * scale; }
    if (nd > 1) { dC1[a] = -accd[a][2] * scale * aW; dC1[a + 4] = accd[a][3] * scale; }
  }
  CC[8] = f * scale;
  if (nd > 0) dC0[8] = df[0] * scale;
  if (nd > 1) dC1[8] = df[1] * scale;
  if (seid < 0) { orc_lit_flip(CC); if (nd > 0) orc_lit_flip(dC0); if (nd > 1) orc_lit_flip(dC1); }
}

/* lit_coef.c:26-78 */
void orc_lit_coef_4p(zc *CC, const double *rt, int s, zc k, const orc_obj *o) { orc_lit_run(CC, NULL, NULL, rt, NULL, NULL, s, k, o, 0); }
/* lit_coef.c:80-122 */
void orc_lit_coef_7p(zc *CC, const double *rt, int s, zc k, const orc_obj *o) { orc_lit_run(CC, NULL, NULL, rt, NULL, NULL, s, k, o, 1); }
/* lit_coef.c:124-242 */
void orc_lit_coef_GLH(zc *CC, const double *rt, int s, zc k, const orc_obj *o, int high) { orc_lit_run(CC, NULL, NULL, rt, NULL, NULL, s, k, o, high ? 3 : 2); }
/* lit_dcoef.c:29-330 */
void orc_lit_dcoef_4p(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, const orc_obj *o) { orc_lit_run(CC, dCC, NULL, rt, v, NULL, s, k, o, 0); }
void orc_lit_dcoef_7p(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, const orc_obj *o) { orc_lit_run(CC, dCC, NULL, rt, v, NULL, s, k, o, 1); }
void orc_lit_dcoef_GLH(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, const orc_obj *o, int high) { orc_lit_run(CC, dCC, NULL, rt, v, NULL, s, k, o, high ? 3 : 2); }
/* lit_dcoef_v2.c:13-357 */
void orc_lit_dcoef2_4p(zc *CC, zc *d0, zc *d1, const double *rt, const double *v0, const double *v1, int s, zc k, const orc_obj *o) { orc_lit_run(CC, d0, d1, rt, v0, v1, s, k, o, 0); }
void orc_lit_dcoef2_7p(zc *CC, zc *d0, zc *d1, const double *rt, const double *v0, const double *v1, int s, zc k, const orc_obj *o) { orc_lit_run(CC, d0, d1, rt, v0, v1, s, k, o, 1); }
void orc_lit_dcoef2_GLH(zc *CC, zc *d0, zc *d1, const double *rt, const double *v0, const double *v1, int s, zc k, const orc_obj *o, int high) { orc_lit_run(CC, d0, d1, rt, v0, v1, s, k, o, high ? 3 : 2); }

/* polar set-up about (zt,et): distances and angles of the 3 vertices,
 * d3b1_elem.c:802-815 */
static void orc_lit_polar(double *r, double *th, double zt, double et)
{
  double zp = 1.0 - zt, zm = -0.5 - zt, ep = 0.5 * SQ3 - et, em = -0.5 * SQ3 - et;
  r[0] = sqrt(zm * zm + em * em); r[1] = sqrt(zp * zp + et * et); r[2] = sqrt(zm * zm + ep * ep); r[3] = r[0];
  th[0] = atan2(em, zm); th[1] = atan2(-et, zp); th[2] = atan2(ep, zm); th[3] = th[0] + 2.0 * M_PI;
}

/* distance from the pole to the edge (r0,th0)-(r1,th1) along direction th */
static double orc_edge_r(double r0, double r1, double th0, double th1, double th)
{
  double rm = r0 * r1 * sin(th1 - th0) / (r0 * sin(th - th0) - r1 * sin(th - th1));
  if (rm < 0.0) { fprintf(stderr, "oracle: integral region error rm=%g\n", rm); exit(1); }
  return rm;
}

/* self-element G (singular 1/R) in polar coordinates: per edge two
 * GHN-point angular halves, GLN radial points; lit_coef.c:290-348, 543-578 */
void orc_lit_coef_bd(zc *CC, double zt, double et, int seid, zc k, const orc_obj *o)
{
  int s = seid < 0 ? -seid : seid;
  double cr[3][3], W[3], aW, r[4], th[4];
  orc_elem(o, s, cr, W, &aW);
  orc_lit_polar(r, th, zt, et);
  for (int i = 0; i < 9; i++) CC[i] = 0.0;
  for (int a = 0; a < 3; a++) {
    zc sum = 0.0;
    for (int j = 0; j < 3; j++) {
      double dth = th[j + 1] - th[j];
      zc tc = 0.0;
      for (int m = 0; m < ORC_GHN; m++) {
        double tq[2] = {0.25 * dth * ORCQ.xhi[m] + 0.25 * (th[j + 1] + 3.0 * th[j]),
                        0.25 * dth * ORCQ.xhi[m] + 0.25 * (3.0 * th[j + 1] + th[j])};
        zc part = 0.0;
        for (int q = 0; q < 2; q++) {
          double t = tq[q], rm = orc_edge_r(r[j], r[j + 1], th[j], th[j + 1], t);
          double c = cos(t), sn = sin(t), K[3];
          for (int d = 0; d < 3; d++) K[d] = cr[d][1] * c + cr[d][2] * sn;
          double aK = orc_norm(K);
          zc rad = 0.0;
          for (int ii = 0; ii < ORC_GLN; ii++) {
            double rr = 0.5 * rm * (ORCQ.xli[ii] + 1.0);
            rad += orc_eikr(k, aK * rr) / aK * orc_lit_Nn(a, rr * c + zt, rr * sn + et) * ORCQ.wli[ii];
          }
          part += rad * 0.5 * rm;
        }
        tc += part * ORCQ.whi[m];
      }
      sum += tc * 0.25 * dth;
    }
    CC[a] = sum * I_SPSQ3 * aW;
  }
  if (seid < 0) orc_lit_flip(CC);
}

/* Principal-value tangential derivative of the self-element G along
 * t_zeta (dir 0) or t_eta (dir 1): the element is split into the disk D_eps
 * of radius DEPS (closed-form radial factor) and its complement (radial GHN
 * rule from DEPS to the edge); lit_dcoef.c:460-635, 857-951 */
void orc_lit_dcoef_bd_PV(zc *dCC, double zt, double et, int seid, zc k, const orc_obj *o, int dir)
{
  int s = seid < 0 ? -seid : seid;
  double cr[3][3], W[3], aW, r[4], th[4], vt[3];
  orc_elem(o, s, cr, W, &aW);
  orc_lit_polar(r, th, zt, et);
  for (int d = 0; d < 3; d++) vt[d] = cr[d][1 + dir];
  double iaT = 1.0 / orc_norm(vt);
  const double eps = ORC_DEPS;
  for (int i = 0; i < 9; i++) dCC[i] = 0.0;
  for (int a = 0; a < 3; a++) {
    zc sum = 0.0;
    /* complement of D_eps */
    for (int j = 0; j < 3; j++) {
      double dth = th[j + 1] - th[j];
      zc tc = 0.0;
      for (int m = 0; m < ORC_GHN; m++) {
        double tq[2] = {0.25 * dth * ORCQ.xhi[m] + 0.25 * (th[j + 1] + 3.0 * th[j]),
                        0.25 * dth * ORCQ.xhi[m] + 0.25 * (3.0 * th[j + 1] + th[j])};
        zc part = 0.0;
        for (int q = 0; q < 2; q++) {
          double t = tq[q], rm = orc_edge_r(r[j], r[j + 1], th[j], th[j + 1], t);
          double c = cos(t), sn = sin(t), K[3];
          for (int d = 0; d < 3; d++) K[d] = cr[d][1] * c + cr[d][2] * sn;
          double aK = orc_norm(K), iaK = 1.0 / aK, Kdv = orc_dot(K, vt) * iaK;
          zc rad = 0.0;
          for (int ii = 0; ii < ORC_GHN; ii++) {
            double rr = 0.5 * (rm - eps) * ORCQ.xhi[ii] + 0.5 * (rm + eps);
            zc kR = k * aK * rr, cb = I * kR - 1.0;
            rad += cb * orc_eikr(k, aK * rr) / rr * iaK * iaK * Kdv * orc_lit_Nn(a, rr * c + zt, rr * sn + et) * ORCQ.whi[ii];
          }
          part += rad * 0.5 * (rm - eps);
        }
        tc += part * ORCQ.whi[m];
      }
      sum += tc * 0.25 * dth;
    }
    /* D_eps: radial integral in closed form, full circle in 4 quarters */
    zc tde = 0.0, tdp = 0.0;
    for (int m = 0; m < ORC_GHN; m++) {
      zc pe = 0.0, pp = 0.0;
      for (int q = 0; q < 4; q++) {
        double t = 0.25 * M_PI * ORCQ.xhi[m] + (-0.75 + 0.5 * q) * M_PI;
        double c = cos(t), sn = sin(t), K[3];
        for (int d = 0; d < 3; d++) K[d] = cr[d][1] * c + cr[d][2] * sn;
        double aK = orc_norm(K), iaK = 1.0 / aK, Kdv = orc_dot(K, vt) * iaK;
        zc ke = k * eps * aK, cs, cc;
        if (cimag(k) == 0.0) { cc = cos(creal(ke)); cs = sin(creal(ke)); } else { cc = ccos(ke); cs = csin(ke); }
        pe += orc_lit_dNn(a, c, sn) * iaK * iaK * iaK * Kdv * ((ke + I) * cs + (2.0 - I * ke) * cc - 2.0);
        if (q >= 2) {   /* principal-value part over the half circle (0, pi) */
          zc sk = cimag(k) == 0.0 ? sin(creal(k) * aK * eps) : csin(k * aK * eps);
          pp += orc_lit_dNn(a, c, sn) * sk * iaK * iaK * iaK * Kdv;
        }
      }
      tde += pe * ORCQ.whi[m];
      tdp += pp * ORCQ.whi[m];
    }
    sum += I * (tde * 0.25 * M_PI) / k;
    sum -= (tdp * 0.25 * M_PI) / k;
    dCC[a] = sum * (-I_SPSQ3 * aW * iaT);
  }
  if (seid < 0) orc_lit_flip(dCC);
}
