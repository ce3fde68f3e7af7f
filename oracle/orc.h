/*
 * orc.h -- CPU oracle for the bem3_emf_b2 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C (C99 complex, OpenMP)
 * restatement of the reference algorithms, used by tests/, by the smoke
 * check and by the cpu_baseline leg of bench.py as the checker.  Nothing in
 * the product (the HIP library and its host layer) links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference tree).  Data are held in flat arrays instead of the
 * reference's pointer-to-pointer tables; indices keep the reference meaning
 * (1-based element ids, 4 node slots per element, signed ids per domain).
 *
 * Parity status of each part is recorded in the header of the file that
 * implements it.
 */
#ifndef ORC_H
#define ORC_H

#include <complex.h>
#include <stddef.h>

typedef double complex zc;

/* quadrature orders, d3b1_const.h:70-71 */
#define ORC_GLN 12
#define ORC_GHN 64
/* GH fallback threshold on |dF/dv|, d3b1_const.h:79 */
#define ORC_CBD_SDF 0.5
/* field-derivative cut-off, d3b1_const.h:78 */
#define ORC_CBD_CDF 100.0
/* D_eps radius of the principal-value split, d3b1_const.h:30 */
#define ORC_DEPS 0.1
/* geometric tolerance, d3b1_const.h:28 */
#define ORC_GEPS 1.0e-15
/* run constants, bem3_emf_b2.h:15-18 */
#define ORC_ITER_MIN 10
#define ORC_ITER_MAX 100
#define ORC_PREC_DEF 2

/* one plane wave, ipw.h:79-99 (user parameters + derived data) */
typedef struct {
  double lambda0, ni, power;
  zc e0x, e0y;
  double fx, fy, fz, theta, phi;
  /* derived, ipw.c:65-89 */
  double E0, ki, cos_t, sin_t, cos_p, sin_p;
  zc ex, ey, ez, hx, hy, hz;
} orc_ipw;

/* one beam of the other multi_fbeam types (orc_beam.c): user parameters
 * and derived data; type 1 focused plane wave, 2 with spiral phase, 3
 * Bessel, 4 Bessel with spiral phase, 5 radially / azimuthally polarised
 * (e0x, e0y, ex, ey, hx, hy are then the radial / azimuthal components) */
typedef struct {
  int type;
  double lambda0, ni, na, power;   /* na: NA, or the Bessel deflection angle */
  zc e0x, e0y;
  double fx, fy, fz, theta, phi;
  int lg_m, nn;
  double ki, E0, cos_t, sin_t, cos_p, sin_p, ct, st;
  zc ex, ey, hx, hy;
  int nt;
  double *tc, *ts, *tw;            /* spectral angles: cos, sin, weight x sin */
} orc_fb;

typedef struct {
  int n;
  orc_ipw *w;
  int nf;
  orc_fb *f;
  double n0, lambda0, omega;
} orc_beams;

typedef struct {
  int Nn, Ne, NN, MN;
  double *rn;              /* [Nn+1][3]                          */
  int *ed, *eni;           /* [Ne+1][4]                          */
  int *md, *sd, *gd;       /* [Ne+1]                             */
  double *cr;              /* [Ne+1][3][4]                       */
  double *cw;              /* [Ne+1][3][3]                       */
  double *ren, *wen;       /* [Ne+1][4][3]                       */
  zc *n, *kn;              /* [MN+1]                             */
  int *dNe;                /* [MN+1] elements seen by domain d   */
  int **sid;               /* [MN+1][dNe+1] signed element ids   */
  /* boundary values per domain, slot layout [(s*4+node)*4+c] (U,dU) and
     [(s*4+node)*3+c] (E,dE,H,dH), s = 0..dNe                        */
  zc **U, **dU, **E, **dE, **H, **dH;
  /* incident boundary data per element, [(i*4+node)*4+c]          */
  zc *Ui, *dUi, *Ui0, *dUi0, *Uip;
  orc_beams bm;
} orc_obj;

/* quadrature tables shared by every object (setup2.c:719-748) */
typedef struct {
  double zt44[4], et44[4], wt44[4];
  double zt49[9], et49[9], wt49[9];
  double zt34[4], et34[4], wt34[4];
  double zt37[7], et37[7], wt37[7];
  double xli[ORC_GLN], wli[ORC_GLN];
  double xhi[ORC_GHN], whi[ORC_GHN];
} orc_quad;

extern orc_quad ORCQ;
void orc_quad_init(void);

/* ---- small vector helpers (my_utils.c) ---- */
static inline double orc_dot(const double *a, const double *b) { return a[0]*b[0]+a[1]*b[1]+a[2]*b[2]; }
double orc_norm(const double *a);
int orc_unit(double *a);

#define ORC_CR(o,s,d,j) ((o)->cr[((size_t)(s)*3+(d))*4+(j)])
#define ORC_CW(o,s,d,j) ((o)->cw[((size_t)(s)*3+(d))*3+(j)])
#define ORC_REN(o,s,n,d) ((o)->ren[((size_t)(s)*4+(n))*3+(d)])
#define ORC_WEN(o,s,n,d) ((o)->wen[((size_t)(s)*4+(n))*3+(d)])
#define ORC_ED(o,s,j) ((o)->ed[(size_t)(s)*4+(j)])
#define ORC_ENI(o,s,j) ((o)->eni[(size_t)(s)*4+(j)])

static inline int orc_is_tri(const orc_obj *o, int seid) { int a = seid < 0 ? -seid : seid; return ORC_ED(o,a,3) == 0; }

/* ---- geometry / setup (orc_geom.c) ---- */
orc_obj *orc_obj_new(int Nn, const double *rn, int Ne, const int *ed, const int *md, const int *sd,
                     const int *gd, int MN, const zc *n_med);
orc_obj *orc_obj_from_msh(const char *path, int MN, const zc *n_med);
void orc_obj_free(orc_obj *o);
void orc_obj_translate(orc_obj *o, const double *rv, double th, const double *tv);
void orc_elem_const(orc_obj *o);
void orc_node_data(orc_obj *o);
void orc_set_beams(orc_obj *o, int nb, const double *par);
void orc_set_k(orc_obj *o);
void orc_incident_bd(orc_obj *o);
void orc_alloc_state(orc_obj *o);
int orc_domain_id(const double *rt, const orc_obj *o);
void orc_beam_EH(zc *e, zc *h, const double *x, const orc_beams *bm);
void orc_beam_EH_dv(zc *e, zc *h, zc *de, zc *dh, const double *x, const double *v, const orc_beams *bm);
void orc_gauleg_ab(double a, double b, double *x, double *w, int n);
/* ---- other beams (orc_beam.c) ---- */
void orc_set_mfb(orc_obj *o, int nb, const double *par);
void orc_fb_EH(zc *e, zc *h, zc *de, zc *dh, const double *x, const double *v, const orc_fb *b);
void orc_fb_free(orc_beams *bm);
void orc_r_bd(double *r, const orc_obj *o, int seid, double zeta, double eta);
void orc_tz_te_node(double *vtz, double *vte, const orc_obj *o, int seid, int nid);

/* ---- element integrals: triangles (orc_lit.c) ---- */
double orc_lit_Nn(int n, double zeta, double eta);
void orc_lit_coef_4p(zc *CC, const double *rt, int s, zc k, const orc_obj *o);
void orc_lit_coef_7p(zc *CC, const double *rt, int s, zc k, const orc_obj *o);
void orc_lit_coef_GLH(zc *CC, const double *rt, int s, zc k, const orc_obj *o, int high);
void orc_lit_coef_bd(zc *CC, double zt, double et, int s, zc k, const orc_obj *o);
void orc_lit_dcoef_4p(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, const orc_obj *o);
void orc_lit_dcoef_7p(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, const orc_obj *o);
void orc_lit_dcoef_GLH(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, const orc_obj *o, int high);
void orc_lit_dcoef2_4p(zc *CC, zc *d0, zc *d1, const double *rt, const double *v0, const double *v1, int s, zc k, const orc_obj *o);
void orc_lit_dcoef2_7p(zc *CC, zc *d0, zc *d1, const double *rt, const double *v0, const double *v1, int s, zc k, const orc_obj *o);
void orc_lit_dcoef2_GLH(zc *CC, zc *d0, zc *d1, const double *rt, const double *v0, const double *v1, int s, zc k, const orc_obj *o, int high);
void orc_lit_dcoef_bd_PV(zc *dCC, double zt, double et, int s, zc k, const orc_obj *o, int dir);

/* ---- element integrals: bi-linear quads (orc_bil.c) ---- */
double orc_bil_Nn(int n, double zeta, double eta);
void orc_bil_coef_4p(zc *CC, const double *rt, int s, zc k, const orc_obj *o);
void orc_bil_coef_9p(zc *CC, const double *rt, int s, zc k, const orc_obj *o);
void orc_bil_coef_GLH(zc *CC, const double *rt, int s, zc k, const orc_obj *o, int high);
void orc_bil_coef_bd(zc *CC, double zt, double et, int s, zc k, const orc_obj *o);
void orc_bil_dcoef_bd_PV(zc *dCC, double zt, double et, int s, zc k, const orc_obj *o, int dir);
void orc_bil_dcoef_4p(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, const orc_obj *o);
void orc_bil_dcoef_9p(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, const orc_obj *o);
void orc_bil_dcoef_GLH(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, const orc_obj *o, int high);
void orc_bil_dcoef2_GLH(zc *CC, zc *d0, zc *d1, const double *rt, const double *v0, const double *v1, int s, zc k, const orc_obj *o, int high);

/* ---- dispatch (orc_dispatch.c), d3b1_elem.c:148-539 ---- */
void orc_coef_rt(zc *CC, const double *rt, int s, zc k, int type, const orc_obj *o);
void orc_coef_bd(zc *CC, const double *rt, int t, double zt, double et, int s, zc k, int type, const orc_obj *o);
void orc_dcoef_rt(zc *CC, zc *dCC, const double *rt, const double *v, int s, zc k, int type, const orc_obj *o);
int orc_dcoef_bd_t2(zc *CC, zc *dCz, zc *dCe, const double *rt, const double *vtz, const double *vte, int t,
                    double zt, double et, int s, zc k, int type, const orc_obj *o);
int orc_dcoef_bd_node_t2(zc *CC, zc *dCz, zc *dCe, int t, int nid, const double *vtz, const double *vte,
                         int s, zc k, int type, const orc_obj *o);

/* ---- single-object BIE (orc_bie.c) ---- */
int orc_cmatrix_domain(const orc_obj *o, int did, int type, zc *G, zc *H, zc *dG, zc *dH, double *dF, int *gh_count);
int orc_cmatrix_rows(const orc_obj *o, int did, int type, int t0, int t1, zc *G, zc *H, zc *dG, zc *dH, double *dF,
                     int *gh_count);
int orc_zgetrf(int n, zc *A, int lda, int *piv);
void orc_zgetrs(int n, int nrhs, const zc *A, int lda, const int *piv, zc *B, int ldb);
typedef struct {
  int na, nn;
  zc *lu;   int *piv;            /* factors of the 8NN matrix A */
  zc *G[2], *H[2];               /* padded N x N per domain (N = 4 Ne_d) */
  zc *dG[2], *dH[2];             /* padded 2N x N */
  double *dF[2];                 /* 3N */
} orc_sys;
int orc_create_matrix(orc_obj *o, orc_sys *sys, int type);
void orc_sys_free(orc_sys *sys);
void orc_build_A(const orc_obj *o, const orc_sys *sys, zc *A);
void orc_build_B(const orc_obj *o, const orc_sys *sys, zc *B);
void orc_store_X(orc_obj *o, const zc *X);
void orc_solve_mEMP(orc_obj *o, orc_sys *sys);
void orc_solve_EH(orc_obj *o, orc_sys *sys);
int orc_solve_dEH(orc_obj *o, orc_sys *sys);

/* ---- multi-object layer (orc_mo.c) ---- */
int orc_mEMP2_s(zc *U, zc *dU, const double *rt, const double *v, int type, const orc_obj *o);
void orc_add_scattered(orc_obj *dst, const orc_obj *src);
void orc_reset_incident(orc_obj *o);
double orc_ccd(orc_obj **mo, int N);
int orc_mo_solve(orc_obj **mo, orc_sys *sys, int N, double ccd, int iter_max, double *cc_trace);
int orc_EH_s(zc *E, zc *H, const double *rt, int type, const orc_obj *o);
int orc_mEMP_s(zc *U, const double *rt, int type, const orc_obj *o);
int orc_EH_mEMP_s(zc *E, zc *H, const double *rt, int type, const orc_obj *o);
void orc_EH_s_bd(zc *E, zc *H, int did, int t, double zt, double et, int type, const orc_obj *o);
int orc_mo_domain(int *did, const double *rt, orc_obj **mo, int N);
int orc_mo_EH(zc *E, zc *H, const double *rt, int type, orc_obj **mo, int N, int mode);
int orc_mo_mEMP(zc *U, const double *rt, int type, orc_obj **mo, int N, int mode);
int orc_mo_EH_mEMP(zc *E, zc *H, const double *rt, int type, orc_obj **mo, int N, int mode);
void orc_mo_EH_bd(zc *E, zc *H, int oid, int did, int t, double zt, double et, int type, orc_obj **mo, int N, int mode);
int orc_mo_force(double *F, double *Nq, const double *rc, int oid, int type, orc_obj **mo, int N);
int orc_mo_absorb(double *P, int oid, int type, orc_obj **mo, int N);

#endif
