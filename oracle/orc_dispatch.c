/*
 * orc_dispatch.c -- oracle: integration-rule dispatch per element type,
 * d3b1_elem.c:148-539.  TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * type 0: 4-point, 1: 7-point (triangle) / 9-point (quad), 2: GLN GL,
 * 3: GHN GL.  type 4 (double-exponential, a test-only path of the
 * reference) is not restated.  Derivative rules switch to GHN when
 * |Re dF/dv| > CBD_SDF, exactly as the reference decides it.
 * Parity: see orc_lit.c / orc_bil.c.
 */
#include "orc.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static void orc_bad_type(int type)
{
  fprintf(stderr, "oracle: integration type %d is not available\n", type);
  exit(1);
}

/* d3b1_elem.c:148-185 */
void orc_coef_rt(zc *CC, const double *rt, int s, zc k, int type, const orc_obj *o)
{
  if (!orc_is_tri(o, s)) {
    switch (type) {
    case 0: orc_bil_coef_4p(CC, rt, s, k, o); break;
    case 1: orc_bil_coef_9p(CC, rt, s, k, o); break;
    case 2: case 3: orc_bil_coef_GLH(CC, rt, s, k, o, type == 3); break;
    default: orc_bad_type(type);
    }
  } else {
    switch (type) {
    case 0: orc_lit_coef_4p(CC, rt, s, k, o); break;
    case 1: orc_lit_coef_7p(CC, rt, s, k, o); break;
    case 2: case 3: orc_lit_coef_GLH(CC, rt, s, k, o, type == 3); break;
    default: orc_bad_type(type);
    }
  }
}

/* d3b1_elem.c:187-230 : the self element uses the polar rule */
void orc_coef_bd(zc *CC, const double *rt, int t, double zt, double et, int s, zc k, int type, const orc_obj *o)
{
  if (t != s) { orc_coef_rt(CC, rt, s, k, type, o); return; }
  if (orc_is_tri(o, s)) orc_lit_coef_bd(CC, zt, et, s, k, o);
  else orc_bil_coef_bd(CC, zt, et, s, k, o);
}

/* d3b1_elem.c:277-316 */
void orc_dcoef_rt(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, int type, const orc_obj *o)
{
  int tri = orc_is_tri(o, s);
  switch (type) {
  case 0: if (tri) orc_lit_dcoef_4p(CC, dCC, rt, v, s, k, o); else orc_bil_dcoef_4p(CC, dCC, rt, v, s, k, o); break;
  case 1: if (tri) orc_lit_dcoef_7p(CC, dCC, rt, v, s, k, o); else orc_bil_dcoef_9p(CC, dCC, rt, v, s, k, o); break;
  case 2: case 3:
    if (tri) orc_lit_dcoef_GLH(CC, dCC, rt, v, s, k, o, type == 3);
    else orc_bil_dcoef_GLH(CC, dCC, rt, v, s, k, o, type == 3);
    break;
  default: orc_bad_type(type);
  }
  if (type < 3 && fabs(creal(dCC[8])) > ORC_CBD_SDF) {
    if (tri) orc_lit_dcoef_GLH(CC, dCC, rt, v, s, k, o, 1);
    else orc_bil_dcoef_GLH(CC, dCC, rt, v, s, k, o, 1);
  }
}

/* Coefficients of one source element at a boundary point with two
 * tangential derivatives, d3b1_elem.c:418-480 (dcoef_bd_t2): t != s the
 * point rt with the rule of type (GHN again when |Re dF/dv| > CBD_SDF along
 * either direction); t == s the self element at (zt, et), polar rule and
 * principal-value derivatives.  Returns 1 when the GHN fallback was taken,
 * 2 for the self element, 0 otherwise. */
int orc_dcoef_bd_t2(zc *CC, zc *dCz, zc *dCe, const double *rt, const double *vtz, const double *vte, int t,
                    double zt, double et, int s, zc k, int type, const orc_obj *o)
{
  int tri = orc_is_tri(o, s);
  if (t != s) {
    if (tri) {
      if (type == 0) orc_lit_dcoef2_4p(CC, dCz, dCe, rt, vtz, vte, s, k, o);
      else if (type == 1) orc_lit_dcoef2_7p(CC, dCz, dCe, rt, vtz, vte, s, k, o);
      else if (type == 2 || type == 3) orc_lit_dcoef2_GLH(CC, dCz, dCe, rt, vtz, vte, s, k, o, type == 3);
      else orc_bad_type(type);
    } else {
      if (type == 2 || type == 3) orc_bil_dcoef2_GLH(CC, dCz, dCe, rt, vtz, vte, s, k, o, type == 3);
      else orc_bad_type(type);
    }
    if (type < 3 && (fabs(creal(dCz[8])) > ORC_CBD_SDF || fabs(creal(dCe[8])) > ORC_CBD_SDF)) {
      if (tri) orc_lit_dcoef2_GLH(CC, dCz, dCe, rt, vtz, vte, s, k, o, 1);
      else orc_bil_dcoef2_GLH(CC, dCz, dCe, rt, vtz, vte, s, k, o, 1);
      return 1;
    }
    return 0;
  }
  if (tri) {
    orc_lit_coef_bd(CC, zt, et, s, k, o);
    orc_lit_dcoef_bd_PV(dCz, zt, et, s, k, o, 0);
    orc_lit_dcoef_bd_PV(dCe, zt, et, s, k, o, 1);
    return 2;
  }
  orc_bil_coef_bd(CC, zt, et, s, k, o);
  orc_bil_dcoef_bd_PV(dCz, zt, et, s, k, o, 0);
  orc_bil_dcoef_bd_PV(dCe, zt, et, s, k, o, 1);
  return 2;
}

/* Coefficients of one (row node, source element) pair with two tangential
 * derivatives, d3b1_elem.c:482-539: dcoef_bd_t2 at node nid of element t. */
int orc_dcoef_bd_node_t2(zc *CC, zc *dCz, zc *dCe, int t, int nid, const double *vtz, const double *vte,
                         int s, zc k, int type, const orc_obj *o)
{
  int at = t < 0 ? -t : t;
  double rt[3];
  for (int d = 0; d < 3; d++) rt[d] = ORC_REN(o, at, nid, d);
  const int tri = orc_is_tri(o, s);
  const double zt = tri ? ORCQ.zt34[nid] : ORCQ.zt44[nid], et = tri ? ORCQ.et34[nid] : ORCQ.et44[nid];
  return orc_dcoef_bd_t2(CC, dCz, dCe, rt, vtz, vte, t, zt, et, s, k, type, o);
}
