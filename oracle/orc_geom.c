/*
 * orc_geom.c -- oracle: quadrature tables, mesh input, element constants,
 * incident plane waves and the solid-angle domain test.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Parity: pinned through the v1.dat fixture replays in tests/ (fields and
 * forces of example1.log, stored boundary state), the GL node/weight check
 * against numpy.polynomial.legendre.leggauss, the node-interpolation identity
 * N_i(node_j) = delta_ij, and an analytic inside/outside test of domain_id on
 * a sphere.
 */
#include "orc.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

orc_quad ORCQ;

double orc_norm(const double *a) { return sqrt(a[0]*a[0]+a[1]*a[1]+a[2]*a[2]); }

/* my_utils.c:15-23 : normalise, -1 when the vector is zero */
int orc_unit(double *a)
{
  double m = orc_norm(a);
  if (m == 0.0) return -1;
  a[0] /= m; a[1] /= m; a[2] /= m;
  return 0;
}

static void *orc_calloc(size_t n, size_t sz)
{
  void *p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); exit(1); }
  return p;
}

/* Gauss-Legendre nodes on (a,b) by Newton iteration on P_n, gauleg.c:3-36 */
static void orc_gauleg(double a, double b, double *x, double *w, int n)
{
  int half = (n + 1) / 2;
  double mid = 0.5 * (b + a), hl = 0.5 * (b - a);
  for (int i = 1; i <= half; i++) {
    double z = cos(M_PI * ((double)i - 0.25) / ((double)n + 0.5));
    double pn, pd;
    for (;;) {
      double p0 = 1.0, p1 = 0.0;   /* P_j, P_{j-1} */
      for (int j = 1; j <= n; j++) {
        double pm = p1;
        p1 = p0;
        p0 = ((double)(2 * j - 1) * z * p1 - (double)(j - 1) * pm) / (double)j;
      }
      pn = p0;
      pd = (double)n * (z * p0 - p1) / (z * z - 1.0);
      double zo = z;
      z = zo - pn / pd;
      if (fabs(z - zo) <= 3.0e-16) break;
    }
    x[i - 1] = mid - hl * z;
    x[n - i] = mid + hl * z;
    w[i - 1] = 2.0 * hl / ((1.0 - z * z) * pd * pd);
    w[n - i] = w[i - 1];
  }
  if (a == -b && (n % 2) == 1) x[n / 2] = 0.0;
}

/* fixed rules of init_elem_const, setup2.c:719-748 and d3b1_const.h:50-67 */
void orc_quad_init(void)
{
  const double p44 = 0.5773502691896257645091;            /* sqrt(1/3)              */
  const double p49 = 0.7745966692414833770359;            /* sqrt(3/5)              */
  const double w490 = 0.3086419753086419753086, w491 = 0.4938271604938271604938, w492 = 0.7901234567901234567901;
  const double p340 = 0.2, p341 = 0.3464101615137754587055;
  const double w340 = 0.5208333333333333333333, w341 = 0.5625;
  const double p370 = 0.3480702390148154917985, p371 = 0.2052130961576726346557;
  const double p372 = 0.6028753385763033130539, p373 = 0.3554395089236065568357;
  const double w370 = 0.1259391805448271525957, w371 = 0.1323941527885061807376, w372 = 0.225;
  orc_quad *q = &ORCQ;
  /* 4-point quad rule: counter-clockwise corners of the Gauss square */
  const double z44[4] = {-p44, p44, p44, -p44}, e44[4] = {-p44, -p44, p44, p44};
  for (int i = 0; i < 4; i++) { q->zt44[i] = z44[i]; q->et44[i] = e44[i]; q->wt44[i] = 1.0; }
  const double z49[9] = {-p49, p49, p49, -p49, 0.0, p49, 0.0, -p49, 0.0};
  const double e49[9] = {-p49, -p49, p49, p49, -p49, 0.0, p49, 0.0, 0.0};
  const double w49[9] = {w490, w490, w490, w490, w491, w491, w491, w491, w492};
  memcpy(q->zt49, z49, sizeof z49); memcpy(q->et49, e49, sizeof e49); memcpy(q->wt49, w49, sizeof w49);
  /* triangle: three node points plus the centroid with a negative weight */
  const double z34[4] = {-p340, 2.0 * p340, -p340, 0.0}, e34[4] = {-p341, 0.0, p341, 0.0};
  const double w34[4] = {w340, w340, w340, -w341};
  memcpy(q->zt34, z34, sizeof z34); memcpy(q->et34, e34, sizeof e34); memcpy(q->wt34, w34, sizeof w34);
  const double z37[7] = {-p370, 2.0 * p370, -p370, p371, -2.0 * p371, p371, 0.0};
  const double e37[7] = {-p372, 0.0, p372, p373, 0.0, -p373, 0.0};
  const double w37[7] = {w370, w370, w370, w371, w371, w371, w372};
  memcpy(q->zt37, z37, sizeof z37); memcpy(q->et37, e37, sizeof e37); memcpy(q->wt37, w37, sizeof w37);
  orc_gauleg(-1.0, 1.0, q->xli, q->wli, ORC_GLN);
  orc_gauleg(-1.0, 1.0, q->xhi, q->whi, ORC_GHN);
}

/* allocate the per-element arrays of an object */
static orc_obj *orc_alloc(int Nn, int Ne, int MN)
{
  orc_obj *o = orc_calloc(1, sizeof(orc_obj));
  o->Nn = Nn; o->Ne = Ne; o->MN = MN;
  o->rn = orc_calloc((size_t)(Nn + 1) * 3, sizeof(double));
  o->ed = orc_calloc((size_t)(Ne + 1) * 4, sizeof(int));
  o->eni = orc_calloc((size_t)(Ne + 1) * 4, sizeof(int));
  o->md = orc_calloc(Ne + 1, sizeof(int));
  o->sd = orc_calloc(Ne + 1, sizeof(int));
  o->gd = orc_calloc(Ne + 1, sizeof(int));
  o->cr = orc_calloc((size_t)(Ne + 1) * 12, sizeof(double));
  o->cw = orc_calloc((size_t)(Ne + 1) * 9, sizeof(double));
  o->ren = orc_calloc((size_t)(Ne + 1) * 12, sizeof(double));
  o->wen = orc_calloc((size_t)(Ne + 1) * 12, sizeof(double));
  o->n = orc_calloc(MN + 1, sizeof(zc));
  o->kn = orc_calloc(MN + 1, sizeof(zc));
  size_t ns = (size_t)(Ne + 1) * 16;
  o->Ui = orc_calloc(ns, sizeof(zc)); o->dUi = orc_calloc(ns, sizeof(zc));
  o->Ui0 = orc_calloc(ns, sizeof(zc)); o->dUi0 = orc_calloc(ns, sizeof(zc));
  o->Uip = orc_calloc(ns, sizeof(zc));
  return o;
}

/* element-node numbering: sequential per element, 3 or 4 slots (setup2.c:371-383) */
static void orc_number_nodes(orc_obj *o)
{
  int nc = 0;
  for (int i = 1; i <= o->Ne; i++) {
    int nv = ORC_ED(o, i, 3) == 0 ? 3 : 4;
    for (int j = 0; j < 4; j++) ORC_ENI(o, i, j) = j < nv ? nc++ : -1;
  }
  o->NN = nc;
}

/* domain membership lists, setup2.c:751-835 */
static void orc_subdomains(orc_obj *o)
{
  o->dNe = orc_calloc(o->MN + 1, sizeof(int));
  o->sid = orc_calloc(o->MN + 1, sizeof(int *));
  for (int i = 1; i <= o->Ne; i++) { o->dNe[o->md[i]]++; o->dNe[o->sd[i]]++; }
  for (int d = 0; d <= o->MN; d++) {
    o->sid[d] = orc_calloc(o->dNe[d] + 1, sizeof(int));
    int c = 1;
    for (int i = 1; i <= o->Ne; i++) {
      if (o->md[i] == d) o->sid[d][c++] = i;
      else if (o->sd[i] == d) o->sid[d][c++] = -i;
    }
  }
}

void orc_alloc_state(orc_obj *o)
{
  o->U = orc_calloc(o->MN + 1, sizeof(zc *)); o->dU = orc_calloc(o->MN + 1, sizeof(zc *));
  o->E = orc_calloc(o->MN + 1, sizeof(zc *)); o->dE = orc_calloc(o->MN + 1, sizeof(zc *));
  o->H = orc_calloc(o->MN + 1, sizeof(zc *)); o->dH = orc_calloc(o->MN + 1, sizeof(zc *));
  for (int d = 0; d <= o->MN; d++) {
    size_t n4 = (size_t)(o->dNe[d] + 1) * 16, n3 = (size_t)(o->dNe[d] + 1) * 12;
    o->U[d] = orc_calloc(n4, sizeof(zc)); o->dU[d] = orc_calloc(n4, sizeof(zc));
    o->E[d] = orc_calloc(n3, sizeof(zc)); o->dE[d] = orc_calloc(n3, sizeof(zc));
    o->H[d] = orc_calloc(n3, sizeof(zc)); o->dH[d] = orc_calloc(n3, sizeof(zc));
  }
}

/* Build an object from already paired boundary data (the open-region side
 * already swapped to the main side, as setup2.c:484-498 leaves it). */
orc_obj *orc_obj_new(int Nn, const double *rn, int Ne, const int *ed, const int *md, const int *sd,
                     const int *gd, int MN, const zc *n_med)
{
  orc_obj *o = orc_alloc(Nn, Ne, MN);
  memcpy(o->rn, rn, sizeof(double) * 3 * (size_t)(Nn + 1));
  memcpy(o->ed, ed, sizeof(int) * 4 * (size_t)(Ne + 1));
  memcpy(o->md, md, sizeof(int) * (size_t)(Ne + 1));
  memcpy(o->sd, sd, sizeof(int) * (size_t)(Ne + 1));
  if (gd) memcpy(o->gd, gd, sizeof(int) * (size_t)(Ne + 1));
  for (int i = 1; i <= MN; i++) o->n[i] = n_med[i];
  orc_number_nodes(o);
  orc_subdomains(o);
  orc_alloc_state(o);
  return o;
}

static int orc_next_int(FILE *fp, int *v) { return fscanf(fp, "%d", v) == 1; }

/* Gmsh 2.2 ASCII reader with the paired-face convention, setup2.c:205-503.
 * Every face is written twice: element 2i-1 (physical = domain id, 99 = open)
 * and element 2i with the reversed node order; the open side becomes the
 * main side with the node order flipped.  Returns NULL on malformed input. */
orc_obj *orc_obj_from_msh(const char *path, int MN, const zc *n_med)
{
  FILE *fp = fopen(path, "r");
  if (!fp) return NULL;
  char line[512];
  double ver; int ftype, prec, Nn, Nel;
  if (!fgets(line, sizeof line, fp)) goto bad;
  if (fscanf(fp, "%lf %d %d", &ver, &ftype, &prec) != 3) goto bad;
  if (ver != 2.2 || ftype != 0 || prec != 8) goto bad;
  if (fscanf(fp, " %511s", line) != 1) goto bad;           /* $EndMeshFormat */
  if (fscanf(fp, " %511s", line) != 1) goto bad;           /* $Nodes */
  if (!orc_next_int(fp, &Nn)) goto bad;
  double *rn = orc_calloc((size_t)(Nn + 1) * 3, sizeof(double));
  for (int i = 1; i <= Nn; i++) {
    int id;
    if (fscanf(fp, "%d %lf %lf %lf", &id, &rn[3 * i], &rn[3 * i + 1], &rn[3 * i + 2]) != 4) { free(rn); goto bad; }
  }
  if (fscanf(fp, " %511s", line) != 1 || fscanf(fp, " %511s", line) != 1) { free(rn); goto bad; }
  if (!orc_next_int(fp, &Nel)) { free(rn); goto bad; }
  int Ne = Nel / 2;
  int *ed = orc_calloc((size_t)(Ne + 1) * 4, sizeof(int));
  int *mdv = orc_calloc(Ne + 1, sizeof(int)), *sdv = orc_calloc(Ne + 1, sizeof(int)), *gdv = orc_calloc(Ne + 1, sizeof(int));
  int ok = 1;
  for (int i = 1; i <= Ne && ok; i++) {
    int id, et, nt, tag, nv;
    ok = orc_next_int(fp, &id) && id == 2 * i - 1 && orc_next_int(fp, &et) && (et == 2 || et == 3) && orc_next_int(fp, &nt);
    if (!ok) break;
    for (int j = 0; j < nt && ok; j++) {
      ok = orc_next_int(fp, &tag);
      if (j == 0) { if (tag == 99) tag = 0; if (tag > MN) ok = 0; mdv[i] = tag; }
      else if (j == 1) gdv[i] = tag;
    }
    nv = et == 2 ? 3 : 4;
    for (int j = 0; j < nv && ok; j++) ok = orc_next_int(fp, &ed[4 * i + j]);
    if (nv == 3) ed[4 * i + 3] = 0;
    int id2, et2, nt2;
    ok = ok && orc_next_int(fp, &id2) && id2 == 2 * i && orc_next_int(fp, &et2) && orc_next_int(fp, &nt2);
    for (int j = 0; j < nt2 && ok; j++) {
      ok = orc_next_int(fp, &tag);
      if (j == 0) { if (tag == 99) tag = 0; if (tag > MN) ok = 0; sdv[i] = tag; }
    }
    /* the partner lists the same nodes in reversed order: 0, n-1, ..., 1 */
    for (int j = 0; j < nv && ok; j++) {
      int v; ok = orc_next_int(fp, &v);
      int want = j == 0 ? ed[4 * i] : ed[4 * i + (nv - j)];
      if (ok && v != want) ok = 0;
    }
    if (ok && sdv[i] == 0) {         /* open region becomes the main side */
      sdv[i] = mdv[i]; mdv[i] = 0;
      if (nv == 3) { int t = ed[4 * i + 1]; ed[4 * i + 1] = ed[4 * i + 2]; ed[4 * i + 2] = t; }
      else { int t = ed[4 * i + 3]; ed[4 * i + 3] = ed[4 * i + 1]; ed[4 * i + 1] = t; }
    }
  }
  fclose(fp);
  orc_obj *o = NULL;
  if (ok) o = orc_obj_new(Nn, rn, Ne, ed, mdv, sdv, gdv, MN, n_med);
  free(rn); free(ed); free(mdv); free(sdv); free(gdv);
  return o;
bad:
  fclose(fp);
  return NULL;
}

void orc_obj_free(orc_obj *o)
{
  if (!o) return;
  free(o->rn); free(o->ed); free(o->eni); free(o->md); free(o->sd); free(o->gd);
  free(o->cr); free(o->cw); free(o->ren); free(o->wen); free(o->n); free(o->kn);
  free(o->Ui); free(o->dUi); free(o->Ui0); free(o->dUi0); free(o->Uip);
  if (o->sid) for (int d = 0; d <= o->MN; d++) free(o->sid[d]);
  if (o->U) for (int d = 0; d <= o->MN; d++) { free(o->U[d]); free(o->dU[d]); free(o->E[d]); free(o->dE[d]); free(o->H[d]); free(o->dH[d]); }
  free(o->sid); free(o->dNe); free(o->U); free(o->dU); free(o->E); free(o->dE); free(o->H); free(o->dH);
  orc_fb_free(&o->bm);
  free(o->bm.w);
  free(o);
}

/* Rodrigues rotation about rv by th, then translation by tv, setup2.c:637-664 */
void orc_obj_translate(orc_obj *o, const double *rv, double th, const double *tv)
{
  double a[3] = {rv[0], rv[1], rv[2]};
  orc_unit(a);
  double c = cos(th), s = sin(th), oc = 1.0 - c;
  double M[9] = {c + a[0] * a[0] * oc, a[0] * a[1] * oc - a[2] * s, a[2] * a[0] * oc + a[1] * s,
                 a[0] * a[1] * oc + a[2] * s, c + a[1] * a[1] * oc, a[1] * a[2] * oc - a[0] * s,
                 a[2] * a[0] * oc - a[1] * s, a[1] * a[2] * oc + a[0] * s, c + a[2] * a[2] * oc};
  for (int i = 1; i <= o->Nn; i++) {
    double *p = &o->rn[3 * i], r[3];
    for (int d = 0; d < 3; d++) r[d] = M[3 * d] * p[0] + M[3 * d + 1] * p[1] + M[3 * d + 2] * p[2] + tv[d];
    p[0] = r[0]; p[1] = r[1]; p[2] = r[2];
  }
}

/* geometric constants: r(z,e) = c0 + c1 z + c2 e (+ c3 z e) and the surface
 * normal W = dr/dz x dr/de (constant on a triangle, linear on a quad),
 * setup2.c:666-717 */
void orc_elem_const(orc_obj *o)
{
  for (int i = 1; i <= o->Ne; i++) {
    double p[4][3];
    for (int j = 0; j < 4; j++) for (int d = 0; d < 3; d++) p[j][d] = o->rn[3 * ORC_ED(o, i, j) + d];
    if (ORC_ED(o, i, 3) != 0) {
      for (int d = 0; d < 3; d++) {
        ORC_CR(o, i, d, 0) = 0.25 * (p[0][d] + p[1][d] + p[2][d] + p[3][d]);
        ORC_CR(o, i, d, 1) = 0.25 * (-p[0][d] + p[1][d] + p[2][d] - p[3][d]);
        ORC_CR(o, i, d, 2) = 0.25 * (-p[0][d] - p[1][d] + p[2][d] + p[3][d]);
        ORC_CR(o, i, d, 3) = 0.25 * (p[0][d] - p[1][d] + p[2][d] - p[3][d]);
      }
      /* W(z,e) = W0 + z W1 + e W2 with W0 = (d13 x d02)/8, W1 = (d12 x d03)/8,
         W2 = (d01 x d23)/8 written through node differences d_ab = p_a - p_b */
      for (int d = 0; d < 3; d++) {
        int a = (d + 1) % 3, b = (d + 2) % 3;
        ORC_CW(o, i, d, 0) = 0.125 * ((p[0][a] - p[2][a]) * (p[1][b] - p[3][b]) - (p[1][a] - p[3][a]) * (p[0][b] - p[2][b]));
        ORC_CW(o, i, d, 1) = 0.125 * ((p[2][a] - p[3][a]) * (p[0][b] - p[1][b]) - (p[0][a] - p[1][a]) * (p[2][b] - p[3][b]));
        ORC_CW(o, i, d, 2) = 0.125 * ((p[1][a] - p[2][a]) * (p[0][b] - p[3][b]) - (p[0][a] - p[3][a]) * (p[1][b] - p[2][b]));
      }
    } else {
      const double i3 = 1.0 / 3.0, isq3 = 1.0 / sqrt(3.0);
      for (int d = 0; d < 3; d++) {
        ORC_CR(o, i, d, 0) = i3 * (p[0][d] + p[1][d] + p[2][d]);
        ORC_CR(o, i, d, 1) = i3 * (-p[0][d] + 2.0 * p[1][d] - p[2][d]);
        ORC_CR(o, i, d, 2) = isq3 * (p[2][d] - p[0][d]);
        ORC_CR(o, i, d, 3) = 0.0;
      }
      double u[3], v[3];
      for (int d = 0; d < 3; d++) { u[d] = p[1][d] - p[0][d]; v[d] = p[2][d] - p[0][d]; }
      for (int d = 0; d < 3; d++) {
        int a = (d + 1) % 3, b = (d + 2) % 3;
        ORC_CW(o, i, d, 0) = u[a] * v[b] - v[a] * u[b];
        ORC_CW(o, i, d, 1) = 0.0;
        ORC_CW(o, i, d, 2) = 0.0;
      }
    }
  }
}

/* node points and (unnormalised) normals of the 4 slots, setup2.c:838-869 */
void orc_node_data(orc_obj *o)
{
  for (int i = 1; i <= o->Ne; i++) {
    int tri = ORC_ED(o, i, 3) == 0;
    for (int j = 0; j < 4; j++) {
      double z = tri ? ORCQ.zt34[j] : ORCQ.zt44[j], e = tri ? ORCQ.et34[j] : ORCQ.et44[j];
      for (int d = 0; d < 3; d++) {
        ORC_REN(o, i, j, d) = ORC_CR(o, i, d, 0) + ORC_CR(o, i, d, 1) * z + ORC_CR(o, i, d, 2) * e + (tri ? 0.0 : ORC_CR(o, i, d, 3) * z * e);
        ORC_WEN(o, i, j, d) = ORC_CW(o, i, d, 0) + (tri ? 0.0 : ORC_CW(o, i, d, 1) * z + ORC_CW(o, i, d, 2) * e);
      }
    }
  }
}

/* point on an element, d3b1_elem.c:18-29 */
void orc_r_bd(double *r, const orc_obj *o, int seid, double z, double e)
{
  int s = seid < 0 ? -seid : seid;
  int tri = ORC_ED(o, s, 3) == 0;
  for (int d = 0; d < 3; d++)
    r[d] = ORC_CR(o, s, d, 0) + ORC_CR(o, s, d, 1) * z + ORC_CR(o, s, d, 2) * e + (tri ? 0.0 : ORC_CR(o, s, d, 3) * z * e);
}

/* unit tangents at a node slot, d3b1_elem.c:122-143 */
void orc_tz_te_node(double *vtz, double *vte, const orc_obj *o, int seid, int nid)
{
  int s = seid < 0 ? -seid : seid;
  int tri = ORC_ED(o, s, 3) == 0;
  for (int d = 0; d < 3; d++) {
    vtz[d] = ORC_CR(o, s, d, 1) + (tri ? 0.0 : ORC_CR(o, s, d, 3) * ORCQ.et44[nid]);
    vte[d] = ORC_CR(o, s, d, 2) + (tri ? 0.0 : ORC_CR(o, s, d, 3) * ORCQ.zt44[nid]);
  }
  orc_unit(vtz);
  orc_unit(vte);
}

/* ---- incident plane waves: ipw.c:3-33, 65-89; multi_fbeam.c:764-876, 878-962 ---- */
void orc_set_beams(orc_obj *o, int nb, const double *par)
{
  free(o->bm.w);
  o->bm.n = nb;
  o->bm.w = orc_calloc(nb, sizeof(orc_ipw));
  for (int i = 0; i < nb; i++) {
    const double *p = par + 12 * i;
    orc_ipw *w = &o->bm.w[i];
    w->lambda0 = p[0]; w->ni = p[1]; w->power = p[2];
    w->e0x = p[3] + I * p[4]; w->e0y = p[5] + I * p[6];
    w->fx = p[7]; w->fy = p[8]; w->fz = p[9]; w->theta = p[10]; w->phi = p[11];
    double nrm = 1.0 / sqrt(pow(cabs(w->e0x), 2) + pow(cabs(w->e0y), 2));
    zc ex = w->e0x * nrm, ey = w->e0y * nrm;
    zc hx = -ey * w->ni, hy = ex * w->ni;
    w->E0 = sqrt(w->power * 2.0 / w->ni);
    w->ki = 2.0 * M_PI * w->ni / w->lambda0;
    double st = sin(w->theta), ct = cos(w->theta), sp = sin(w->phi), cp = cos(w->phi);
    w->sin_t = st; w->cos_t = ct; w->sin_p = sp; w->cos_p = cp;
    /* rotate the x/y polarisation of a +z wave to the direction (theta, phi) */
    double a11 = sp * sp + cp * cp * ct, a12 = sp * cp * ct - sp * cp, a22 = sp * sp * ct + cp * cp;
    w->ex = ex * a11 + ey * a12;  w->ey = ex * a12 + ey * a22;  w->ez = -(ex * cp + ey * sp) * st;
    w->hx = hx * a11 + hy * a12;  w->hy = hx * a12 + hy * a22;  w->hz = -(hx * cp + hy * sp) * st;
    if (i == 0) { o->bm.n0 = w->ni; o->bm.lambda0 = w->lambda0; }
    else if (w->ni != o->bm.n0 || w->lambda0 != o->bm.lambda0) { fprintf(stderr, "oracle: beam data mismatch\n"); exit(1); }
  }
  o->bm.omega = 2.0 * M_PI / o->bm.lambda0;
}

void orc_gauleg_ab(double a, double b, double *x, double *w, int n) { orc_gauleg(a, b, x, w, n); }

/* medium wave numbers, setup2.c:111-113 */
void orc_set_k(orc_obj *o)
{
  o->n[0] = o->bm.n0;
  for (int i = 0; i <= o->MN; i++) o->kn[i] = 2.0 * M_PI / o->bm.lambda0 * o->n[i];
}

void orc_beam_EH(zc *e, zc *h, const double *x, const orc_beams *bm)
{
  for (int j = 0; j < 3; j++) { e[j] = 0.0; h[j] = 0.0; }
  for (int i = 0; i < bm->n; i++) {
    const orc_ipw *w = &bm->w[i];
    double ph = w->ki * ((x[0] - w->fx) * w->sin_t * w->cos_p + (x[1] - w->fy) * w->sin_t * w->sin_p + (x[2] - w->fz) * w->cos_t);
    zc a = (cos(ph) + I * sin(ph)) * w->E0;
    e[0] += w->ex * a; e[1] += w->ey * a; e[2] += w->ez * a;
    h[0] += w->hx * a; h[1] += w->hy * a; h[2] += w->hz * a;
  }
  for (int i = 0; i < bm->nf; i++) {   /* the other beam types after the plane waves */
    zc te[3], th[3];
    orc_fb_EH(te, th, NULL, NULL, x, NULL, &bm->f[i]);
    for (int j = 0; j < 3; j++) { e[j] += te[j]; h[j] += th[j]; }
  }
}

void orc_beam_EH_dv(zc *e, zc *h, zc *de, zc *dh, const double *x, const double *v, const orc_beams *bm)
{
  for (int j = 0; j < 3; j++) { e[j] = 0.0; h[j] = 0.0; de[j] = 0.0; dh[j] = 0.0; }
  for (int i = 0; i < bm->n; i++) {
    const orc_ipw *w = &bm->w[i];
    double ph = w->ki * ((x[0] - w->fx) * w->sin_t * w->cos_p + (x[1] - w->fy) * w->sin_t * w->sin_p + (x[2] - w->fz) * w->cos_t);
    zc a = (cos(ph) + I * sin(ph)) * w->E0;
    zc g = I * w->ki * (w->sin_t * w->cos_p * v[0] + w->sin_t * w->sin_p * v[1] + w->cos_t * v[2]);
    zc f[6] = {w->ex * a, w->ey * a, w->ez * a, w->hx * a, w->hy * a, w->hz * a};
    for (int j = 0; j < 3; j++) { e[j] += f[j]; h[j] += f[3 + j]; de[j] += f[j] * g; dh[j] += f[3 + j] * g; }
  }
  for (int i = 0; i < bm->nf; i++) {
    zc te[3], th[3], tde[3], tdh[3];
    orc_fb_EH(te, th, tde, tdh, x, v, &bm->f[i]);
    for (int j = 0; j < 3; j++) { e[j] += te[j]; h[j] += th[j]; de[j] += tde[j]; dh[j] += tdh[j]; }
  }
}

/* incident potential and its normal derivative at the domain-0 node slots,
 * mo_setup.c:704-757 (phi slot set to zero) */
void orc_incident_bd(orc_obj *o)
{
  for (int i = 1; i <= o->Ne; i++) {
    if (o->md[i] != 0) continue;
    for (int j = 0; j < 4; j++) {
      double r[3], w[3];
      for (int d = 0; d < 3; d++) { r[d] = ORC_REN(o, i, j, d); w[d] = ORC_WEN(o, i, j, d); }
      orc_unit(w);
      zc e[3], h[3], de[3], dh[3];
      orc_beam_EH_dv(e, h, de, dh, r, w, &o->bm);
      zc *u = &o->Ui[(i * 4 + j) * 4], *du = &o->dUi[(i * 4 + j) * 4];
      for (int c = 0; c < 3; c++) { u[c] = e[c]; du[c] = de[c]; }
      u[3] = 0.0; du[3] = 0.0;
      for (int c = 0; c < 4; c++) {
        o->Ui0[(i * 4 + j) * 4 + c] = u[c]; o->dUi0[(i * 4 + j) * 4 + c] = du[c]; o->Uip[(i * 4 + j) * 4 + c] = u[c];
      }
    }
  }
}

static void orc_cross(double *r, const double *a, const double *b)
{
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

/* signed solid angle of a flat triangle (and a second one for the quad's
 * other half), setup2.c:871-942.  flg = -1 when the point lies on it. */
static double orc_solid_angle(int quad, double r[4][3], int *flg)
{
  double om = 0.0, n01[3], n12[3], n20[3], t[3];
  *flg = 0;
  int bad = 0;
  for (int i = 0; i < (quad ? 4 : 3); i++) bad += orc_unit(r[i]);
  if (bad < 0) { *flg = -1; return 0.0; }
  for (int half = 0; half < (quad ? 2 : 1); half++) {
    const double *a = r[0], *b = half ? r[2] : r[1], *c = half ? r[3] : r[2];
    if (half == 0) { orc_cross(n01, a, b); orc_cross(n12, b, c); orc_cross(n20, c, a); }
    else { for (int d = 0; d < 3; d++) n01[d] = -n20[d]; orc_cross(n12, b, c); orc_cross(n20, c, a); }
    orc_cross(t, n01, n20); double A0 = atan2(orc_norm(t), -orc_dot(n01, n20));
    orc_cross(t, n12, n01); double A1 = atan2(orc_norm(t), -orc_dot(n12, n01));
    orc_cross(t, n20, n12); double A2 = atan2(orc_norm(t), -orc_dot(n20, n12));
    double D = orc_dot(a, n12);
    if (D > 0.0) om += A0 + A1 + A2 - M_PI;
    else if (D < 0.0) om -= A0 + A1 + A2 - M_PI;
    else { *flg = -1; return 0.0; }
  }
  return om;
}

/* domain of a point from the solid angles of the main-side faces,
 * setup2.c:944-978.  Returns -1 if no domain is closed around it. */
int orc_domain_id(const double *rt, const orc_obj *o)
{
  double *Og = orc_calloc(o->MN + 1, sizeof(double));
  int res = -1;
  for (int d = 0; d <= o->MN && res < 0; d++) {
    for (int i = 1; i <= o->dNe[d]; i++) {
      int s = o->sid[d][i];
      if (s <= 0) continue;
      double rv[4][3];
      for (int j = 0; j < 4; j++) for (int k = 0; k < 3; k++) rv[j][k] = o->rn[3 * ORC_ED(o, s, j) + k] - rt[k];
      int flg;
      double om = orc_solid_angle(ORC_ED(o, s, 3) != 0, rv, &flg);
      if (flg < 0) { res = d; break; }
      Og[d] += om;
      Og[o->sd[s]] -= om;
    }
    if (res >= 0) break;
    if (d == 0 && fabs(Og[d]) < 2.0 * M_PI) res = d;
    else if (Og[d] > 2.0 * M_PI) res = d;
  }
  free(Og);
  return res;
}
