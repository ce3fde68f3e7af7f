/*
 * orc_bil.c -- oracle: bi-linear quadrangle element integrals.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Reference element [-1,1]^2, node slots at the 2x2 Gauss points
 * (d3b1_elem.c:543-568).
 *
 * Parity: pinned by the v1.dat fixture replays (553 quads per object):
 * fields and forces of example1.log, the stored G dE = H E identity, and
 * the stored boundary E, H of every quad node, which the reference computed
 * from the tangential-derivative rows, i.e. from the self-element
 * principal-value integrals of orc_bil_dcoef_bd_PV.
 */
#include "orc.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#define I_FP 7.957747154594766788444e-2   /* 1/(4 pi) */

/* bil_Nn, d3b1_elem.c:556-568 */
double orc_bil_Nn(int n, double z, double e)
{
  const double ia = 1.0 / 0.5773502691896257645091;
  switch (n) {
  case 0: return 0.25 * (1.0 - z * ia) * (1.0 - e * ia);
  case 1: return 0.25 * (1.0 + z * ia) * (1.0 - e * ia);
  case 2: return 0.25 * (1.0 + z * ia) * (1.0 + e * ia);
  default: return 0.25 * (1.0 - z * ia) * (1.0 + e * ia);
  }
}

static inline zc orc_eikr(zc k, double r)
{
  if (cimag(k) == 0.0) { double a = creal(k) * r; return cos(a) + I * sin(a); }
  return cexp(I * k * r);
}

/* H, F (and derivatives) change sign on the sub-domain side, d3b1_elem.c:580-600 */
static void orc_bil_flip(zc *C) { for (int i = 4; i < 9; i++) C[i] = -C[i]; }

static void orc_bil_rw(const orc_obj *o, int s, double z, double e, double *r, double *w)
{
  for (int d = 0; d < 3; d++) {
    r[d] = ORC_CR(o, s, d, 0) + ORC_CR(o, s, d, 1) * z + ORC_CR(o, s, d, 2) * e + ORC_CR(o, s, d, 3) * z * e;
    w[d] = ORC_CW(o, s, d, 0) + ORC_CW(o, s, d, 1) * z + ORC_CW(o, s, d, 2) * e;
  }
}

/* Common driver.  mode 0: 4 node points (each node carries only its own
 * shape function, weight 1 for G/H), 1: 9-point rule, 2/3: GLN/GHN tensor
 * GL.  Kernel forms of bil_coef.c:25-202, bil_dcoef.c:50-320,
 * bil_dcoef_v2.c:13-326. */
static void orc_bil_run(zc *CC, zc *d0, zc *d1, const double *rt, const double *v0, const double *v1,
                        int seid, zc k, const orc_obj *o, int mode)
{
  int s = seid < 0 ? -seid : seid;
  const double *vv[2] = {v0, v1};
  int nd = v1 ? 2 : (v0 ? 1 : 0);
  zc ag[4] = {0}, ah[4] = {0}, adg[2][4] = {{0}}, adh[2][4] = {{0}};
  double f = 0.0, df[2] = {0.0, 0.0};
  int n1 = mode == 0 ? 4 : mode == 1 ? 9 : mode == 2 ? ORC_GLN : ORC_GHN;
  int n2 = mode <= 1 ? 1 : n1;
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++) {
      double z, e, w, vR[3], W[3], N[4];
      if (mode == 0) {
        z = ORCQ.zt44[i]; e = ORCQ.et44[i]; w = ORCQ.wt44[i];
        for (int d = 0; d < 3; d++) { vR[d] = ORC_REN(o, s, i, d) - rt[d]; W[d] = ORC_WEN(o, s, i, d); }
        N[0] = N[1] = N[2] = N[3] = 0.0; N[i] = 1.0;
      } else {
        if (mode == 1) { z = ORCQ.zt49[i]; e = ORCQ.et49[i]; w = ORCQ.wt49[i]; }
        else if (mode == 2) { z = ORCQ.xli[i]; e = ORCQ.xli[j]; w = ORCQ.wli[i] * ORCQ.wli[j]; }
        else { z = ORCQ.xhi[i]; e = ORCQ.xhi[j]; w = ORCQ.whi[i] * ORCQ.whi[j]; }
        double r[3];
        orc_bil_rw(o, s, z, e, r, W);
        for (int d = 0; d < 3; d++) vR[d] = r[d] - rt[d];
        for (int a = 0; a < 4; a++) N[a] = orc_bil_Nn(a, z, e);
      }
      double R = orc_norm(vR), iR = 1.0 / R, aW = orc_norm(W);
      zc kR = k * R, cb = I * kR - 1.0, ce = orc_eikr(k, R);
      double rdW = orc_dot(vR, W) * iR, tD = rdW * iR * iR;
      zc g = ce * iR * aW, h = cb * ce * tD;
      /* the 4-node rule weights G and H by 1 and only F by wt44 */
      double wg = mode == 0 ? 1.0 : w;
      f += w * tD;
      zc dg[2] = {0, 0}, dh[2] = {0, 0};
      for (int d = 0; d < nd; d++) {
        double rdv = orc_dot(vR, vv[d]) * iR, Wdv = orc_dot(W, vv[d]);
        dg[d] = cb * ce * iR * iR * rdv * aW;
        dh[d] = ce * iR * iR * iR * ((k * k * R * R + 3.0 * cb) * rdv * rdW - cb * Wdv);
        df[d] += w * iR * iR * iR * (3.0 * rdv * rdW - Wdv);
      }
      for (int a = 0; a < 4; a++) {
        double wn = wg * N[a];
        if (wn == 0.0) continue;
        ag[a] += wn * g; ah[a] += wn * h;
        for (int d = 0; d < nd; d++) { adg[d][a] += wn * dg[d]; adh[d][a] += wn * dh[d]; }
      }
    }
  zc *dd[2] = {d0, d1};
  for (int i = 0; i < 9; i++) { CC[i] = 0.0; for (int d = 0; d < nd; d++) dd[d][i] = 0.0; }
  for (int a = 0; a < 4; a++) {
    CC[a] = I_FP * ag[a]; CC[a + 4] = I_FP * ah[a];
    for (int d = 0; d < nd; d++) { dd[d][a] = -I_FP * adg[d][a]; dd[d][a + 4] = I_FP * adh[d][a]; }
  }
  CC[8] = I_FP * f;
  for (int d = 0; d < nd; d++) dd[d][8] = I_FP * df[d];
  if (seid < 0) { orc_bil_flip(CC); for (int d = 0; d < nd; d++) orc_bil_flip(dd[d]); }
}

void orc_bil_coef_4p(zc *CC, const double *rt, int s, zc k, const orc_obj *o) { orc_bil_run(CC, NULL, NULL, rt, NULL, NULL, s, k, o, 0); }
void orc_bil_coef_9p(zc *CC, const double *rt, int s, zc k, const orc_obj *o) { orc_bil_run(CC, NULL, NULL, rt, NULL, NULL, s, k, o, 1); }
void orc_bil_coef_GLH(zc *CC, const double *rt, int s, zc k, const orc_obj *o, int high) { orc_bil_run(CC, NULL, NULL, rt, NULL, NULL, s, k, o, high ? 3 : 2); }
void orc_bil_dcoef_4p(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, const orc_obj *o) { orc_bil_run(CC, dCC, NULL, rt, v, NULL, s, k, o, 0); }
void orc_bil_dcoef_9p(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, const orc_obj *o) { orc_bil_run(CC, dCC, NULL, rt, v, NULL, s, k, o, 1); }
void orc_bil_dcoef_GLH(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, const orc_obj *o, int high) { orc_bil_run(CC, dCC, NULL, rt, v, NULL, s, k, o, high ? 3 : 2); }
void orc_bil_dcoef2_GLH(zc *CC, zc *d0, zc *d1, const double *rt, const double *v0, const double *v1, int s, zc k, const orc_obj *o, int high) { orc_bil_run(CC, d0, d1, rt, v0, v1, s, k, o, high ? 3 : 2); }

/* flatness test of bil_check_plane, d3b1_elem.c:640-649 (kept with its
 * 0/0 behaviour for parallelograms, where the test reports "not plane") */
static int orc_bil_plane(const orc_obj *o, int s)
{
  double b[3], c[3];
  for (int d = 0; d < 3; d++) { b[d] = ORC_CR(o, s, d, 1); c[d] = ORC_CW(o, s, d, 2); }
  double q = orc_dot(b, c) / (orc_norm(b) * orc_norm(c));
  return fabs(q) < ORC_GEPS;
}

/* polar set-up about (zt,et) for the square, d3b1_elem.c:667-681 */
static void orc_bil_polar(double *r, double *th, double zt, double et)
{
  double zp = 1.0 - zt, zm = -1.0 - zt, ep = 1.0 - et, em = -1.0 - et;
  r[0] = sqrt(zm * zm + em * em); r[1] = sqrt(zp * zp + em * em);
  r[2] = sqrt(zp * zp + ep * ep); r[3] = sqrt(zm * zm + ep * ep); r[4] = r[0];
  th[0] = atan2(em, zm); th[1] = atan2(em, zp); th[2] = atan2(ep, zp); th[3] = atan2(ep, zm); th[4] = th[0] + 2.0 * M_PI;
}

/* Self-element G (and H, F on a warped quad) in polar coordinates about the
 * point, bil_coef.c:244-342, 426-536.  which: 0 -> G slot a, 1 -> H slot a,
 * 2 -> F.  Returns the unscaled angular-radial integral. */
static zc orc_bil_bd_int(const orc_obj *o, int s, zc k, double zt, double et, int a, int which)
{
  double r[5], th[5], cr[3][4], cw[3][3];
  for (int d = 0; d < 3; d++) { for (int j = 0; j < 4; j++) cr[d][j] = ORC_CR(o, s, d, j); for (int j = 0; j < 3; j++) cw[d][j] = ORC_CW(o, s, d, j); }
  orc_bil_polar(r, th, zt, et);
  zc sum = 0.0;
  for (int j = 0; j < 4; j++) {
    double dth = th[j + 1] - th[j];
    zc tc = 0.0;
    for (int m = 0; m < ORC_GHN; m++) {
      double tq[2] = {0.25 * dth * ORCQ.xhi[m] + 0.25 * (th[j + 1] + 3.0 * th[j]),
                      0.25 * dth * ORCQ.xhi[m] + 0.25 * (3.0 * th[j + 1] + th[j])};
      zc part = 0.0;
      for (int q = 0; q < 2; q++) {
        double t = tq[q];
        double rm = r[j] * r[j + 1] * sin(th[j + 1] - th[j]) / (r[j] * sin(t - th[j]) - r[j + 1] * sin(t - th[j + 1]));
        if (rm < 0.0) { fprintf(stderr, "oracle: integral region error\n"); exit(1); }
        double c = cos(t), sn = sin(t);
        zc rad = 0.0;
        for (int ii = 0; ii < ORC_GLN; ii++) {
          double rr = 0.5 * rm * (ORCQ.xli[ii] + 1.0), z = rr * c + zt, e = rr * sn + et, K[3], W[3];
          for (int d = 0; d < 3; d++) {
            K[d] = cr[d][1] * c + cr[d][2] * sn + cr[d][3] * (rr * c * sn + zt * sn + et * c);
            W[d] = cw[d][0] + cw[d][1] * z + cw[d][2] * e;
          }
          double aK = orc_norm(K);
          zc val;
          if (which == 0) val = orc_eikr(k, aK * rr) / aK * orc_norm(W) * orc_bil_Nn(a, z, e);
          else if (which == 1) { zc kR = k * rr * aK; val = (I * kR - 1.0) * orc_eikr(k, rr * aK) / (aK * aK * aK) * c * sn * orc_bil_Nn(a, z, e); }
          else val = 1.0 / (aK * aK * aK) * c * sn;
          rad += val * ORCQ.wli[ii];
        }
        part += rad * 0.5 * rm;
      }
      tc += part * ORCQ.whi[m];
    }
    sum += tc * 0.25 * dth;
  }
  return sum;
}

void orc_bil_coef_bd(zc *CC, double zt, double et, int seid, zc k, const orc_obj *o)
{
  int s = seid < 0 ? -seid : seid;
  for (int i = 0; i < 9; i++) CC[i] = 0.0;
  for (int a = 0; a < 4; a++) CC[a] = I_FP * orc_bil_bd_int(o, s, k, zt, et, a, 0);
  if (!orc_bil_plane(o, s)) {
    double bdC = 0.0;
    for (int d = 0; d < 3; d++) bdC += ORC_CR(o, s, d, 1) * ORC_CW(o, s, d, 2);
    for (int a = 0; a < 4; a++) CC[a + 4] = I_FP * bdC * orc_bil_bd_int(o, s, k, zt, et, a, 1);
    CC[8] = I_FP * bdC * orc_bil_bd_int(o, s, k, zt, et, 0, 2);
  }
  if (seid < 0) orc_bil_flip(CC);
}

/* ---- self-element principal-value tangential derivatives ----
 * bil_dcoef_bd_tz0_PV / bil_dcoef_bd_te0_PV, bil_dcoef.c:369-740, with the
 * integrands of bil_dcoef.c:916-1631.  The element is split about the pole
 * (zt,et) into the disk D_eps (radius DEPS) and its complement.  Complement:
 * per edge two GHN-point angular halves, GHN radial points on (DEPS, rm).
 * D_eps: four GHN-point quarters of (-pi, pi), GLN radial points on
 * (0, DEPS); the principal-value part: two quarters of (0, pi).  dir 0 is
 * the t_zeta derivative, 1 the t_eta one; t is the unnormalised tangent at
 * the pole, the result is divided by |t|. */
typedef struct {
  double cr[3][4], cw[3][3], vt[3], zt, et;
  zc k;
  int dir;
} orc_bpv;

/* K(r, theta) = dr/dr_polar of the bilinear map, and W at the point */
static void orc_bpv_KW(const orc_bpv *b, double rr, double c, double sn, double *K, double *W, double *z, double *e)
{
  *z = rr * c + b->zt;
  *e = rr * sn + b->et;
  for (int d = 0; d < 3; d++) {
    K[d] = b->cr[d][1] * c + b->cr[d][2] * sn + b->cr[d][3] * (rr * c * sn + b->zt * sn + b->et * c);
    if (W) W[d] = b->cw[d][0] + b->cw[d][1] * *z + b->cw[d][2] * *e;
  }
}

/* which 0: G slot a, 1: H slot a, 2: F.  part 0: complement of D_eps
 * (radial integrand at rr), 1: D_eps, 2: principal-value pair at +rr / -rr */
static zc orc_bpv_f(const orc_bpv *b, int which, int a, int part, double rr, double c, double sn)
{
  const double ang = b->dir == 0 ? c : sn;   /* K.t enters with cos (zeta) or sin (eta) */
  double K[3], W[3], z, e;
  if (part == 2) {
    zc v[2];
    for (int q = 0; q < 2; q++) {
      double sr = q == 0 ? rr : -rr;
      orc_bpv_KW(b, sr, c, sn, K, W, &z, &e);
      double aK = orc_norm(K), KdT = orc_dot(K, b->vt);
      zc cc = cimag(b->k) == 0.0 ? cos(creal(b->k) * aK * sr) : ccos(b->k * aK * sr);
      if (which == 0) v[q] = cc / (aK * aK) * (KdT / aK) * orc_bil_Nn(a, z, e) * orc_norm(W);
      else if (which == 1) v[q] = cc / pow(aK, 5) * (3.0 * KdT * ang - aK * aK) * orc_bil_Nn(a, z, e);
      else v[q] = 1.0 / pow(aK, 5) * (3.0 * KdT * ang - aK * aK);
    }
    return (v[0] - v[1]) / rr;
  }
  orc_bpv_KW(b, rr, c, sn, K, W, &z, &e);
  double aK = orc_norm(K), KdT = orc_dot(K, b->vt);
  zc kR = b->k * aK * rr;
  if (part == 0) {
    zc ce = orc_eikr(b->k, aK * rr), cb = I * kR - 1.0;
    if (which == 0) return cb * ce / (aK * aK) / rr * (KdT / aK) * orc_bil_Nn(a, z, e) * orc_norm(W);
    if (which == 1) return ce / (rr * pow(aK, 5)) * ((kR * kR + 3.0 * cb) * KdT * ang - cb * aK * aK) * orc_bil_Nn(a, z, e);
    return 1.0 / (rr * pow(aK, 5)) * (3.0 * KdT * ang - aK * aK);
  }
  /* part 1 */
  zc cs = cimag(b->k) == 0.0 ? sin(creal(kR)) : csin(kR);
  if (which == 0) {
    zc cc = cimag(b->k) == 0.0 ? cos(creal(kR)) : ccos(kR), cb = I * kR - 1.0;
    return (cb * cs + kR * cc) / (aK * aK) / rr * (KdT / aK) * orc_norm(W) * orc_bil_Nn(a, z, e);
  }
  /* H: (De0 - i De1) of bil_sdHz_De_theta */
  zc ce = orc_eikr(b->k, aK * rr), ik = I * b->k;
  zc d0 = ce / pow(aK, 4) * ((b->k * kR + 3.0 * ik) * KdT * ang - ik * aK * aK) * orc_bil_Nn(a, z, e);
  zc d1 = cs / (pow(aK, 5) * rr) * (3.0 * KdT * ang - aK * aK) * orc_bil_Nn(a, z, e);
  return d0 - I * d1;
}

/* the three integrals of one quantity: complement, D_eps, PV half disk,
 * each including the angular factor (1 for G, sin/cos theta for H and F) */
static void orc_bpv_int(const orc_bpv *b, int which, int a, zc *comp, zc *disk, zc *pv)
{
  double r[5], th[5];
  const double eps = ORC_DEPS;
  orc_bil_polar(r, th, b->zt, b->et);
  *comp = *disk = *pv = 0.0;
  for (int j = 0; j < 4; j++) {
    double dth = th[j + 1] - th[j];
    zc tc = 0.0;
    for (int m = 0; m < ORC_GHN; m++) {
      double tq[2] = {0.25 * dth * ORCQ.xhi[m] + 0.25 * (th[j + 1] + 3.0 * th[j]),
                      0.25 * dth * ORCQ.xhi[m] + 0.25 * (3.0 * th[j + 1] + th[j])};
      zc part = 0.0;
      for (int q = 0; q < 2; q++) {
        double t = tq[q];
        double rm = r[j] * r[j + 1] * sin(th[j + 1] - th[j]) / (r[j] * sin(t - th[j]) - r[j + 1] * sin(t - th[j + 1]));
        if (rm < 0.0) { fprintf(stderr, "oracle: integral region error\n"); exit(1); }
        double c = cos(t), sn = sin(t), fa = which == 0 ? 1.0 : (b->dir == 0 ? sn : c);
        zc rad = 0.0;
        for (int ii = 0; ii < ORC_GHN; ii++)
          rad += orc_bpv_f(b, which, a, 0, 0.5 * (rm - eps) * ORCQ.xhi[ii] + 0.5 * (rm + eps), c, sn) * ORCQ.whi[ii];
        part += fa * rad * 0.5 * (rm - eps);
      }
      tc += part * ORCQ.whi[m];
    }
    *comp += tc * 0.25 * dth;
  }
  zc td = 0.0, tp = 0.0;
  for (int m = 0; m < ORC_GHN; m++) {
    for (int q = 0; q < 4; q++) {
      double t = 0.25 * M_PI * ORCQ.xhi[m] + (-0.75 + 0.5 * q) * M_PI;
      double c = cos(t), sn = sin(t), fa = which == 0 ? 1.0 : (b->dir == 0 ? sn : c);
      zc rd = 0.0, rp = 0.0;
      for (int ii = 0; ii < ORC_GLN; ii++) {
        double rr = 0.5 * eps * (ORCQ.xli[ii] + 1.0);
        if (which != 2) rd += orc_bpv_f(b, which, a, 1, rr, c, sn) * ORCQ.wli[ii];
        if (q >= 2) rp += orc_bpv_f(b, which, a, 2, rr, c, sn) * ORCQ.wli[ii];
      }
      td += fa * rd * 0.5 * eps * ORCQ.whi[m];
      tp += fa * rp * 0.5 * eps * ORCQ.whi[m];
    }
  }
  *disk = td * 0.25 * M_PI;
  *pv = tp * 0.25 * M_PI;
}

void orc_bil_dcoef_bd_PV(zc *dCC, double zt, double et, int seid, zc k, const orc_obj *o, int dir)
{
  int s = seid < 0 ? -seid : seid;
  orc_bpv b;
  for (int d = 0; d < 3; d++) {
    for (int j = 0; j < 4; j++) b.cr[d][j] = ORC_CR(o, s, d, j);
    for (int j = 0; j < 3; j++) b.cw[d][j] = ORC_CW(o, s, d, j);
    b.vt[d] = dir == 0 ? b.cr[d][1] + b.cr[d][3] * et : b.cr[d][2] + b.cr[d][3] * zt;   /* bil_t_zeta / bil_t_eta */
  }
  b.zt = zt; b.et = et; b.k = k; b.dir = dir;
  double iaT = 1.0 / orc_norm(b.vt);
  zc comp, disk, pv;
  for (int i = 0; i < 9; i++) dCC[i] = 0.0;
  for (int a = 0; a < 4; a++) {
    orc_bpv_int(&b, 0, a, &comp, &disk, &pv);
    dCC[a] = (comp + I * disk - pv) * (-I_FP * iaT);
  }
  if (!orc_bil_plane(o, s)) {
    double bdC = 0.0;
    for (int d = 0; d < 3; d++) bdC += b.cr[d][1] * b.cw[d][2];
    for (int a = 0; a < 4; a++) {
      orc_bpv_int(&b, 1, a, &comp, &disk, &pv);
      dCC[a + 4] = (comp + disk - pv) * (bdC * I_FP * iaT);
    }
    orc_bpv_int(&b, 2, 0, &comp, &disk, &pv);
    dCC[8] = (comp + pv) * (bdC * I_FP * iaT);
  }
  if (seid < 0) orc_bil_flip(dCC);
}
