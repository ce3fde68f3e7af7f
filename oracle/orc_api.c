/*
 * orc_api.c -- flat-array entry points of the oracle for the Python test
 * harness (ctypes).  TEST INFRASTRUCTURE ONLY (see orc.h).
 * Complex arrays cross this boundary as interleaved (re, im) doubles.
 */
#include "orc.h"
#include <stdlib.h>
#include <string.h>

void orc_init(void) { orc_quad_init(); }

void orc_api_quad(double *xl, double *wl, double *xh, double *wh)
{
  memcpy(xl, ORCQ.xli, sizeof ORCQ.xli); memcpy(wl, ORCQ.wli, sizeof ORCQ.wli);
  memcpy(xh, ORCQ.xhi, sizeof ORCQ.xhi); memcpy(wh, ORCQ.whi, sizeof ORCQ.whi);
}

orc_obj *orc_api_from_msh(const char *path, int MN, const double *nmed)
{
  return orc_obj_from_msh(path, MN, (const zc *)nmed);
}

orc_obj *orc_api_from_arrays(int Nn, const double *rn, int Ne, const int *ed, const int *md, const int *sd,
                             const int *gd, int MN, const double *nmed)
{
  return orc_obj_new(Nn, rn, Ne, ed, md, sd, gd, MN, (const zc *)nmed);
}

void orc_api_free(orc_obj *o) { orc_obj_free(o); }

/* translate (optional), element constants, node data, beams, wave numbers,
 * incident boundary data */
void orc_api_setup(orc_obj *o, int nb, const double *beams, const double *tv, int translate)
{
  if (translate) { double rv[3] = {1.0, 0.0, 0.0}; orc_obj_translate(o, rv, 0.0, tv); }
  orc_elem_const(o);
  orc_node_data(o);
  orc_set_beams(o, nb, beams);
  orc_set_k(o);
  orc_incident_bd(o);
}

/* the same with a beam set of any multi_fbeam types, 16 doubles per beam
 * (orc_set_mfb) */
void orc_api_setup_mfb(orc_obj *o, int nb, const double *beams, const double *tv, int translate)
{
  if (translate) { double rv[3] = {1.0, 0.0, 0.0}; orc_obj_translate(o, rv, 0.0, tv); }
  orc_elem_const(o);
  orc_node_data(o);
  orc_set_mfb(o, nb, beams);
  orc_set_k(o);
  orc_incident_bd(o);
}

/* incident E, H (and their derivatives along v, when v != NULL) of the
 * object's beams at np points: out rows of 6 (12 with v) */
void orc_api_beam_EH(const orc_obj *o, int np, const double *x, const double *v, double *out)
{
  zc *z = (zc *)out;
  for (int p = 0; p < np; p++) {
    if (v) orc_beam_EH_dv(z + 12 * p, z + 12 * p + 3, z + 12 * p + 6, z + 12 * p + 9, x + 3 * p, v + 3 * p, &o->bm);
    else orc_beam_EH(z + 6 * p, z + 6 * p + 3, x + 3 * p, &o->bm);
  }
}

/* replace node points / normals and the wave numbers by stored values */
void orc_api_override(orc_obj *o, const double *ren, const double *wen, const double *kn)
{
  if (ren) memcpy(o->ren, ren, sizeof(double) * 12 * (size_t)(o->Ne + 1));
  if (wen) memcpy(o->wen, wen, sizeof(double) * 12 * (size_t)(o->Ne + 1));
  if (kn) memcpy(o->kn, kn, sizeof(zc) * (size_t)(o->MN + 1));
}

void orc_api_refresh_incident(orc_obj *o) { orc_incident_bd(o); }

int orc_api_info(const orc_obj *o, int *out)
{
  out[0] = o->Nn; out[1] = o->Ne; out[2] = o->NN; out[3] = o->MN;
  for (int d = 0; d <= o->MN && d < 8; d++) out[4 + d] = o->dNe[d];
  return 0;
}

void orc_api_geometry(const orc_obj *o, double *rn, int *ed, int *eni, int *md, int *sd, double *cr, double *cw, double *ren, double *wen)
{
  size_t E1 = (size_t)o->Ne + 1;
  if (rn) memcpy(rn, o->rn, sizeof(double) * 3 * ((size_t)o->Nn + 1));
  if (ed) memcpy(ed, o->ed, sizeof(int) * 4 * E1);
  if (eni) memcpy(eni, o->eni, sizeof(int) * 4 * E1);
  if (md) memcpy(md, o->md, sizeof(int) * E1);
  if (sd) memcpy(sd, o->sd, sizeof(int) * E1);
  if (cr) memcpy(cr, o->cr, sizeof(double) * 12 * E1);
  if (cw) memcpy(cw, o->cw, sizeof(double) * 9 * E1);
  if (ren) memcpy(ren, o->ren, sizeof(double) * 12 * E1);
  if (wen) memcpy(wen, o->wen, sizeof(double) * 12 * E1);
}

void orc_api_sid(const orc_obj *o, int d, int *sid) { memcpy(sid, o->sid[d], sizeof(int) * ((size_t)o->dNe[d] + 1)); }

/* which: 0 U, 1 dU, 2 E, 3 dE, 4 H, 5 dH; set = 1 writes into the object */
void orc_api_state(orc_obj *o, int d, int which, double *buf, int set)
{
  zc **tab[6] = {o->U, o->dU, o->E, o->dE, o->H, o->dH};
  size_t n = ((size_t)o->dNe[d] + 1) * (which < 2 ? 16 : 12);
  if (set) memcpy(tab[which][d], buf, n * sizeof(zc));
  else memcpy(buf, tab[which][d], n * sizeof(zc));
}

/* which: 0 Ui, 1 dUi, 2 Ui0, 3 dUi0 */
void orc_api_incident(orc_obj *o, int which, double *buf, int set)
{
  zc *tab[4] = {o->Ui, o->dUi, o->Ui0, o->dUi0};
  size_t n = ((size_t)o->Ne + 1) * 16;
  if (set) memcpy(tab[which], buf, n * sizeof(zc));
  else memcpy(buf, tab[which], n * sizeof(zc));
}

void orc_api_coef_rt(const orc_obj *o, const double *rt, int s, const double *k, int type, double *CC)
{
  orc_coef_rt((zc *)CC, rt, s, k[0] + I * k[1], type, o);
}

void orc_api_coef_bd(const orc_obj *o, const double *rt, int t, double zt, double et, int s, const double *k, int type, double *CC)
{
  orc_coef_bd((zc *)CC, rt, t, zt, et, s, k[0] + I * k[1], type, o);
}

void orc_api_dcoef_rt(const orc_obj *o, const double *rt, const double *v, int s, const double *k, int type, double *CC, double *dCC)
{
  orc_dcoef_rt((zc *)CC, (zc *)dCC, rt, v, s, k[0] + I * k[1], type, o);
}

int orc_api_dcoef_node_t2(const orc_obj *o, int t, int nid, int s, const double *k, int type, double *CC, double *dCz, double *dCe)
{
  double vtz[3], vte[3];
  orc_tz_te_node(vtz, vte, o, t, nid);
  return orc_dcoef_bd_node_t2((zc *)CC, (zc *)dCz, (zc *)dCe, t, nid, vtz, vte, s, k[0] + I * k[1], type, o);
}

int orc_api_dcoef_bd_t2(const orc_obj *o, const double *rt, const double *vtz, const double *vte, int t, double zt,
                        double et, int s, const double *k, int type, double *CC, double *dCz, double *dCe)
{
  return orc_dcoef_bd_t2((zc *)CC, (zc *)dCz, (zc *)dCe, rt, vtz, vte, t, zt, et, s, k[0] + I * k[1], type, o);
}

/* padded domain matrices: G,H (N x N), dG,dH (2N x N), dF (3N), N = 4 dNe */
int orc_api_cmatrix(const orc_obj *o, int did, int type, double *G, double *H, double *dG, double *dH, double *dF, int *ngh)
{
  return orc_cmatrix_domain(o, did, type, (zc *)G, (zc *)H, (zc *)dG, (zc *)dH, dF, ngh);
}

/* rows of elements t0 <= t < t1 of the padded domain matrices */
int orc_api_cmatrix_rows(const orc_obj *o, int did, int type, int t0, int t1, double *G, double *H, double *dG, double *dH,
                         double *dF, int *ngh)
{
  return orc_cmatrix_rows(o, did, type, t0, t1, (zc *)G, (zc *)H, (zc *)dG, (zc *)dH, dF, ngh);
}

orc_sys *orc_api_create_matrix(orc_obj *o, int type, int *status)
{
  orc_sys *sys = calloc(1, sizeof(orc_sys));
  *status = orc_create_matrix(o, sys, type);
  return sys;
}

void orc_api_sys_free(orc_sys *sys) { orc_sys_free(sys); free(sys); }

/* coefficient matrices only (no 8NN factorisation): enough for the
 * boundary-field steps orc_solve_EH / orc_solve_dEH */
orc_sys *orc_api_assemble(orc_obj *o, int type, int *status)
{
  orc_sys *sys = calloc(1, sizeof(orc_sys));
  int bad = 0;
  for (int d = 0; d <= o->MN && d < 2; d++) {
    size_t N = 4 * (size_t)o->dNe[d];
    sys->G[d] = calloc(N * N, sizeof(zc)); sys->H[d] = calloc(N * N, sizeof(zc));
    sys->dG[d] = calloc(2 * N * N, sizeof(zc)); sys->dH[d] = calloc(2 * N * N, sizeof(zc));
    sys->dF[d] = calloc(3 * N, sizeof(double));
    bad += orc_cmatrix_domain(o, d, type, sys->G[d], sys->H[d], sys->dG[d], sys->dH[d], sys->dF[d], NULL);
  }
  sys->nn = o->NN; sys->na = 8 * o->NN;
  *status = bad;
  return sys;
}

/* the dense 8NN matrix of an object (row-major, na x na) */
void orc_api_build_A(const orc_obj *o, const orc_sys *sys, double *A) { orc_build_A(o, sys, (zc *)A); }
void orc_api_build_B(const orc_obj *o, const orc_sys *sys, double *B) { orc_build_B(o, sys, (zc *)B); }
void orc_api_store_X(orc_obj *o, const double *X) { orc_store_X(o, (const zc *)X); }
void orc_api_solve_EH(orc_obj *o, orc_sys *sys) { orc_solve_EH(o, sys); }
int orc_api_solve_dEH(orc_obj *o, orc_sys *sys) { return orc_solve_dEH(o, sys); }

int orc_api_mo_solve(orc_obj **mo, orc_sys **sys, int N, double ccd, int iter_max, double *cc_trace)
{
  orc_sys *s = calloc((size_t)N, sizeof(orc_sys));
  for (int m = 0; m < N; m++) s[m] = *sys[m];
  int it = orc_mo_solve(mo, s, N, ccd, iter_max, cc_trace);
  free(s);
  return it;
}

int orc_api_mEMP2_s(const orc_obj *o, const double *rt, const double *v, int type, double *U, double *dU)
{
  return orc_mEMP2_s((zc *)U, (zc *)dU, rt, v, type, o);
}

void orc_api_add_scattered(orc_obj *dst, const orc_obj *src) { orc_add_scattered(dst, src); }
void orc_api_reset_incident(orc_obj *o) { orc_reset_incident(o); }
double orc_api_ccd(orc_obj **mo, int N) { return orc_ccd(mo, N); }

int orc_api_domain_id(const orc_obj *o, const double *rt) { return orc_domain_id(rt, o); }

/* fields at np points; what: 0 EH (E,H), 1 mEMP (U), 2 EH via potentials;
 * mode 0 s, 1 t, 2 i.  Output rows of 6 (EH) or 4 (U) complex values. */
void orc_api_mo_field(orc_obj **mo, int N, int what, int mode, int type, int np, const double *pts, double *out, int *oid)
{
#pragma omp parallel for schedule(dynamic)
  for (int p = 0; p < np; p++) {
    zc *o = (zc *)out + (size_t)p * (what == 1 ? 4 : 6);
    if (what == 0) oid[p] = orc_mo_EH(o, o + 3, pts + 3 * p, type, mo, N, mode);
    else if (what == 1) oid[p] = orc_mo_mEMP(o, pts + 3 * p, type, mo, N, mode);
    else oid[p] = orc_mo_EH_mEMP(o, o + 3, pts + 3 * p, type, mo, N, mode);
  }
}

void orc_api_mo_EH_bd(orc_obj **mo, int N, int oid, int did, int t, double zt, double et, int type, int mode, double *out)
{
  orc_mo_EH_bd((zc *)out, (zc *)out + 3, oid, did, t, zt, et, type, mo, N, mode);
}

int orc_api_mo_force(orc_obj **mo, int N, int oid, int type, const double *rc, double *F, double *Nq)
{
  return orc_mo_force(F, Nq, rc, oid, type, mo, N);
}

int orc_api_mo_absorb(orc_obj **mo, int N, int oid, int type, double *P)
{
  return orc_mo_absorb(P, oid, type, mo, N);
}

int orc_api_zgesv(int n, int nrhs, double *A, double *B)
{
  int *piv = malloc(sizeof(int) * (size_t)n);
  int info = orc_zgetrf(n, (zc *)A, n, piv);
  if (!info) orc_zgetrs(n, nrhs, (zc *)A, n, piv, (zc *)B, nrhs);
  free(piv);
  return info;
}
