/*
 * orc_beam.c -- oracle: the beams of multi_fbeam other than the plane wave.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 *   type 1  focused plane wave            fpw.c:3-77 (rotation, translation),
 *                                         :115-157 (set-up), :172-382 (field)
 *   type 2  focused, spiral phase         lgb.c:116-164, phase_lgb :177-187,
 *                                         ac_eh_lgb :260-272
 *   type 3  Bessel beam                   bsb.c:112-148, :157-236
 *   type 4  Bessel beam, spiral phase     bslgb.c:115-156, :166-256
 *   type 5  radially / azimuthally        rab.c:111-158, ac_eh_rab :242-254
 *           polarised focused beam
 *   sum over all beams, type order        multi_fbeam.c:764-876
 *   common medium and wavelength          multi_fbeam.c:878-962
 *
 * Every beam is an angular spectrum of plane waves: a Gauss-Legendre rule
 * over sin(theta) in (0, NA/n) (types 1, 2, 5) or the single cone angle of a
 * Bessel beam (3, 4), and in phi the reference's adaptive composite
 * trapezoid: step halving from one point up to 2^13 intervals on [-pi, pi),
 * stopped when the change of sum(eps |E|^2) + sum(|H|^2) falls below
 * 1e-15 sqrt(points) times its largest value, after at least two halvings
 * (mfb_const.h:4-7).  The beam frame is rotated by (theta, phi) and
 * translated by (fx, fy, fz).
 *
 * Parity: the focused plane wave by the reference's analysis_sample3 fixture
 * (ex30.dat with example1_ex30.log: the total field at a point is the stored
 * scattered field plus this incident field); every type by Maxwell's
 * equations (curl E = i k0 H and curl H = -i k0 n^2 E from the directional
 * derivatives, div E = 0), by closed forms on the beam axis checked against
 * an independent quadrature, and by the spiral-phase types with mode 0
 * reducing to their plain counterparts.
 */
#include "orc.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_TRAP_MNUM 12
#define ORC_TRAP_LNUM 1
#define ORC_TRAP_EPS 1.0e-15
#define ORC_NPHI ((2 << ORC_TRAP_MNUM) + 1)

/* phi points of the trapezoid in the order they are added: -pi, then the
 * midpoints of each halving (fpw.c:140-156) */
static double PCP[ORC_NPHI], PSP[ORC_NPHI], PANG[ORC_NPHI];
static int phi_ready = 0;

static void orc_phi_table(void)
{
  if (phi_ready) return;
  double h = 2.0 * M_PI;
  PCP[0] = -1.0; PSP[0] = 0.0; PANG[0] = -M_PI;
  int j = 1, st = 1;
  for (int i = 0; i <= ORC_TRAP_MNUM; i++) {
    h *= 0.5;
    for (int s = 1; s <= st; s++) {
      double a = -M_PI + (double)(2 * s - 1) * h;
      PCP[j] = cos(a); PSP[j] = sin(a); PANG[j] = a;
      j++;
    }
    st *= 2;
  }
  phi_ready = 1;
}

/* exp(i m phi) with phi taken into [0, 2 pi), lgb.c:177-187 */
static zc orc_spiral(int m, double phi)
{
  if (m == 0) return 1.0;
  double pp = phi < 0.0 ? phi + 2.0 * M_PI : phi;
  double t = fmod((double)m * pp, 2.0 * M_PI);
  return cos(t) + I * sin(t);
}

/* one plane-wave component of the spectrum at (ct, st) x (cp, sp) */
static void orc_fb_ac(zc *eh, zc *deh, const orc_fb *b, double ct, double st, int j, const double *x, const double *v)
{
  const double cp = PCP[j], sp = PSP[j];
  double a = b->ki * (st * cp * x[0] + st * sp * x[1] + ct * x[2]);
  zc ce = cos(a) + I * sin(a);
  if (b->type == 2 || b->type == 4) ce *= orc_spiral(b->lg_m, PANG[j]);
  if (b->type == 5) {
    eh[0] = (b->ex * ct * cp - b->ey * sp) * ce;
    eh[1] = (b->ex * ct * sp + b->ey * cp) * ce;
    eh[2] = (-b->ex * st) * ce;
    eh[3] = (b->hx * ct * cp - b->hy * sp) * ce;
    eh[4] = (b->hx * ct * sp + b->hy * cp) * ce;
    eh[5] = (-b->hx * st) * ce;
  } else {
    double m11 = sp * sp + ct * cp * cp, m12 = sp * cp * (ct - 1.0), m22 = ct * sp * sp + cp * cp;
    eh[0] = (b->ex * m11 + b->ey * m12) * ce;
    eh[1] = (b->ex * m12 + b->ey * m22) * ce;
    eh[2] = (-b->ex * st * cp - b->ey * st * sp) * ce;
    eh[3] = (b->hx * m11 + b->hy * m12) * ce;
    eh[4] = (b->hx * m12 + b->hy * m22) * ce;
    eh[5] = (-b->hx * st * cp - b->hy * st * sp) * ce;
  }
  if (deh) {
    zc g = I * b->ki * (st * cp * v[0] + st * sp * v[1] + ct * v[2]);
    for (int i = 0; i < 6; i++) deh[i] = eh[i] * g;
  }
}

/* adaptive composite trapezoid over phi (fpw.c:188-248, 310-366) */
static void orc_fb_phi(zc *It, zc *dIt, const orc_fb *b, double ct, double st, const double *x, const double *v)
{
  const double cep = b->ni * b->ni;
  zc Ia[6], dIa[6];
  double h = 2.0 * M_PI;
  orc_fb_ac(Ia, dIt ? dIa : NULL, b, ct, st, 0, x, v);
  double abst = 0.0;
  for (int i = 0; i < 6; i++) {
    It[i] = h * Ia[i];
    if (dIt) dIt[i] = h * dIa[i];
    abst += (i < 3 ? cep : 1.0) * creal(It[i] * conj(It[i]));
  }
  double f_order = abst;
  int j = 1, sc = 1, cc = 0;
  while (cc <= ORC_TRAP_MNUM) {
    h *= 0.5;
    zc Is[6] = {0}, dIs[6] = {0};
    for (int s = 1; s <= sc; s++, j++) {
      orc_fb_ac(Ia, dIt ? dIa : NULL, b, ct, st, j, x, v);
      for (int i = 0; i < 6; i++) { Is[i] += Ia[i]; if (dIt) dIs[i] += dIa[i]; }
    }
    double abss = 0.0;
    for (int i = 0; i < 6; i++) {
      It[i] = 0.5 * It[i] + Is[i] * h;
      if (dIt) dIt[i] = 0.5 * dIt[i] + dIs[i] * h;
      abss += (i < 3 ? cep : 1.0) * creal(It[i] * conj(It[i]));
    }
    double crc = sqrt((double)sc) * f_order;
    if ((cc > ORC_TRAP_LNUM && fabs(abst - abss) < ORC_TRAP_EPS * crc) || (abss == 0.0 && abst == 0.0)) break;
    abst = abss;
    if (abst > f_order) f_order = abst;
    sc *= 2;
    cc++;
  }
  if (cc >= ORC_TRAP_MNUM) { fprintf(stderr, "oracle: phi integration limit over\n"); exit(1); }
}

/* rotation of the beam frame, fpw.c:13-27 (forward) and the transpose back */
static void orc_fb_rot(double *o, const double *a, const orc_fb *b)
{
  const double ct = b->cos_t, st = b->sin_t, cp = b->cos_p, sp = b->sin_p;
  o[0] = a[0] * (ct * cp * cp + sp * sp) + a[1] * sp * cp * (ct - 1.0) - a[2] * st * cp;
  o[1] = a[0] * sp * cp * (ct - 1.0) + a[1] * (ct * sp * sp + cp * cp) - a[2] * st * sp;
  o[2] = a[0] * st * cp + a[1] * st * sp + a[2] * ct;
}
static void orc_fb_back(zc *o, const zc *a, const orc_fb *b)
{
  const double ct = b->cos_t, st = b->sin_t, cp = b->cos_p, sp = b->sin_p;
  o[0] = a[0] * (ct * cp * cp + sp * sp) + a[1] * sp * cp * (ct - 1.0) + a[2] * st * cp;
  o[1] = a[0] * sp * cp * (ct - 1.0) + a[1] * (ct * sp * sp + cp * cp) + a[2] * st * sp;
  o[2] = -a[0] * st * cp - a[1] * st * sp + a[2] * ct;
}

/* E, H (and their derivatives along v when de != NULL) of one beam */
void orc_fb_EH(zc *e, zc *h, zc *de, zc *dh, const double *x, const double *v, const orc_fb *b)
{
  orc_phi_table();
  double xb[3] = {x[0] - b->fx, x[1] - b->fy, x[2] - b->fz}, xn[3], vn[3] = {0, 0, 0};
  orc_fb_rot(xn, xb, b);
  if (de) orc_fb_rot(vn, v, b);
  zc s[6] = {0}, ds[6] = {0}, It[6], dIt[6];
  if (b->type == 3 || b->type == 4) {
    orc_fb_phi(It, de ? dIt : NULL, b, b->ct, b->st, xn, vn);
    for (int i = 0; i < 6; i++) { s[i] = It[i] * b->st; if (de) ds[i] = dIt[i] * b->st; }
  } else {
    for (int j = 0; j < b->nt; j++) {
      orc_fb_phi(It, de ? dIt : NULL, b, b->tc[j], b->ts[j], xn, vn);
      for (int i = 0; i < 6; i++) { s[i] += It[i] * b->tw[j]; if (de) ds[i] += dIt[i] * b->tw[j]; }
    }
  }
  orc_fb_back(e, s, b);
  orc_fb_back(h, s + 3, b);
  for (int i = 0; i < 3; i++) { e[i] *= b->E0; h[i] *= b->E0; }
  if (de) {
    orc_fb_back(de, ds, b);
    orc_fb_back(dh, ds + 3, b);
    for (int i = 0; i < 3; i++) { de[i] *= b->E0; dh[i] *= b->E0; }
  }
}

/* derived data of one beam (the set-up functions of each type) */
static void orc_fb_setup(orc_fb *b)
{
  b->ki = 2.0 * M_PI * b->ni / b->lambda0;
  b->cos_t = cos(b->theta); b->sin_t = sin(b->theta);
  b->cos_p = cos(b->phi); b->sin_p = sin(b->phi);
  double a0 = sqrt(pow(cabs(b->e0x), 2) + pow(cabs(b->e0y), 2));
  if (b->type == 3 || b->type == 4) {
    b->st = sin(b->na); b->ct = cos(b->na);
    b->ex = b->e0x / a0; b->ey = b->e0y / a0;
    b->E0 = sqrt(2.0 * b->power / b->ni);
    b->nt = 0;
  } else {
    double kr = b->na / b->ni, ce = sqrt(1.0 / M_PI) / (kr * a0);
    b->ex = b->e0x * ce; b->ey = b->e0y * ce;
    b->E0 = sqrt(2.0 * b->power * b->ni) / b->lambda0;
    b->nt = b->nn;
    b->tc = calloc((size_t)b->nt, sizeof(double));
    b->ts = calloc((size_t)b->nt, sizeof(double));
    b->tw = calloc((size_t)b->nt, sizeof(double));
    if (!b->tc || !b->ts || !b->tw) { fprintf(stderr, "oracle: out of memory\n"); exit(1); }
    orc_gauleg_ab(0.0, kr, b->ts, b->tw, b->nt);
    for (int i = 0; i < b->nt; i++) { b->tc[i] = sqrt(1.0 - b->ts[i] * b->ts[i]); b->tw[i] *= b->ts[i]; }
  }
  /* H polarisation (for type 5: radial / azimuthal components) */
  b->hx = -b->ey * b->ni;
  b->hy = b->ex * b->ni;
}

void orc_fb_free(orc_beams *bm)
{
  for (int i = 0; i < bm->nf; i++) { free(bm->f[i].tc); free(bm->f[i].ts); free(bm->f[i].tw); }
  free(bm->f);
  bm->f = NULL;
  bm->nf = 0;
}

/* Every beam of a beam set as 16 doubles: type (0 plane wave, 1-5 as
 * above), lambda0, ni, NA or deflection angle, power, re/im e0x (e0r),
 * re/im e0y (e0a), fx, fy, fz, theta, phi, spiral mode, Gauss-Legendre
 * points.  Plane waves go to the plane-wave list; all beams share the
 * medium index and wavelength (multi_fbeam.c:878-962). */
void orc_set_mfb(orc_obj *o, int nb, const double *par)
{
  int nw = 0, nf = 0;
  for (int i = 0; i < nb; i++) { if ((int)par[16 * i] == 0) nw++; else nf++; }
  double *pw = calloc((size_t)(nw ? nw : 1) * 12, sizeof(double));
  int k = 0;
  for (int i = 0; i < nb; i++) {
    const double *p = par + 16 * i;
    if ((int)p[0] != 0) continue;
    double w[12] = {p[1], p[2], p[4], p[5], p[6], p[7], p[8], p[9], p[10], p[11], p[12], p[13]};
    memcpy(pw + 12 * k++, w, sizeof w);
  }
  orc_fb_free(&o->bm);
  if (nw) orc_set_beams(o, nw, pw);
  else { free(o->bm.w); o->bm.w = NULL; o->bm.n = 0; }
  free(pw);
  o->bm.f = calloc((size_t)(nf ? nf : 1), sizeof(orc_fb));
  o->bm.nf = nf;
  k = 0;
  for (int i = 0; i < nb; i++) {
    const double *p = par + 16 * i;
    if ((int)p[0] == 0) continue;
    orc_fb *b = &o->bm.f[k++];
    b->type = (int)p[0];
    b->lambda0 = p[1]; b->ni = p[2]; b->na = p[3]; b->power = p[4];
    b->e0x = p[5] + I * p[6]; b->e0y = p[7] + I * p[8];
    b->fx = p[9]; b->fy = p[10]; b->fz = p[11]; b->theta = p[12]; b->phi = p[13];
    b->lg_m = (int)p[14]; b->nn = (int)p[15];
    orc_fb_setup(b);
  }
  /* the common medium index and wavelength: the first beam's */
  double n0 = par[2], lam = par[1];
  for (int i = 1; i < nb; i++)
    if (par[16 * i + 2] != n0 || par[16 * i + 1] != lam) { fprintf(stderr, "oracle: beam data mismatch\n"); exit(1); }
  o->bm.n0 = n0;
  o->bm.lambda0 = lam;
  o->bm.omega = 2.0 * M_PI / lam;
}
