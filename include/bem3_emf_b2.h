/*
 * bem3_emf_b2.h -- drop-in C API of the MI355X build of bem3_emf_b2.
 *
 * Same functions, argument meanings, return codes and data types as the
 * reference's d3b2_src/bem3_emf_b2.h:29-122 with d3b1_elem.h:13-97 and
 * multi_fbeam.h:19-41 / ipw.h:12-32 (struct layouts kept member for member,
 * because callers read them, e.g. mo.md[0].mw.lambda_0, and because Bobj and
 * Ipw are written raw into the .obj / .dat files).  DOMD gains one trailing
 * member, the handle of the object's device copy.
 *
 * Host code is plain C; every computation of the hot path (coefficient
 * matrices, the linear solve, the coupling loop, fields, force, absorbed
 * power, the domain test) runs through the C-ABI of bem3_hip.h on the GPU.
 * Errors print a message and exit(1) as in the reference.
 */
#ifndef BEM3_EMF_B2_H_
#define BEM3_EMF_B2_H_

#include <complex.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

/* ---- constants (bem3_emf_b2.h:14-18, d3b1_const.h:70-71) ---- */
#define FNLGTH 256
#define ITER_MIN 10
#define ITER_MAX 100
#define PREC_DEF 2
#define CCD 1.0e-6
#define GLN 12
#define GHN 64

/* ---- incident beams (ipw.h:12-32, multi_fbeam.h:19-41) ---- */
typedef struct incident_planewave_data {
  double E0;
  double ki;
  double cos_t, sin_t;
  double cos_p, sin_p;
  double complex ex, ey, ez;
  double complex hx, hy, hz;
} IpwD;

typedef struct incident_planewave {
  double lambda0;
  double ni;
  double power;
  double complex e0x;
  double complex e0y;
  double fx, fy, fz;
  double theta;
  double phi;
  IpwD data;
} Ipw;

/* The five other beams of multi_fbeam (fpw.h:14-39, lgb.h:14-41,
 * bsb.h:13-37, bslgb.h:13-39, rab.h:13-44).  Each record is written raw into
 * .obj / .dat files, so the member order and types are the reference's; the
 * pointers of the derived-data part are meaningless in a file and are
 * rebuilt by the set-up functions. */
typedef struct fpw_data {
  double ki, E0;
  double complex ex, ey;
  double complex hx, hy;
  double cos_t, sin_t;
  double cos_p, sin_p;
  int nt;
  double *wt, *ct, *st;
  double *cp, *sp;
} FpwD;

typedef struct fpw {              /* focused plane wave */
  double lambda0;
  double ni;
  double NA;
  double power;
  double complex e0x;
  double complex e0y;
  double fx, fy, fz;
  double theta, phi;
  int nn;                         /* Gauss-Legendre points over sin(theta) */
  FpwD data;
} Fpw;

typedef struct lgb_data {
  double ki, E0;
  double complex ex, ey;
  double complex hx, hy;
  double cos_t, sin_t;
  double cos_p, sin_p;
  int nt;
  double *wt, *ct, *st;
  double *cp, *sp;
  double complex *ph_p;
} LGbD;

typedef struct lgb {              /* focused beam with spiral phase modulation */
  double lambda0;
  double ni;
  double NA;
  double power;
  double complex e0x;
  double complex e0y;
  double fx, fy, fz;
  double theta, phi;
  int lg_m;                       /* mode number of the spiral phase */
  int nn;
  LGbD data;
} LGb;

typedef struct bsb_data {
  double ki, E0;
  double complex ex, ey;
  double complex hx, hy;
  double cos_t, sin_t;
  double cos_p, sin_p;
  double ct, st;                  /* cone (deflection) angle */
  double *cp;
  double *sp;
} BsbD;

typedef struct bsb {              /* Bessel beam */
  double lambda0;
  double ni;
  double d_angle;
  double power;
  double complex e0x;
  double complex e0y;
  double fx, fy, fz;
  double theta, phi;
  BsbD data;
} Bsb;

typedef struct bslgb_data {
  double ki, E0;
  double complex ex, ey;
  double complex hx, hy;
  double cos_t, sin_t;
  double cos_p, sin_p;
  double ct, st;
  double *cp, *sp;
  double complex *ph_p;
} BsLGbD;

typedef struct bslgb {            /* Bessel beam with spiral phase modulation */
  double lambda0;
  double ni;
  double d_angle;
  double power;
  double complex e0x;
  double complex e0y;
  double fx, fy, fz;
  double theta, phi;
  int lg_m;
  BsLGbD data;
} BsLGb;

typedef struct rab_data {
  double ki, E0;
  double complex er, ea;
  double complex hr, ha;
  double cos_t, sin_t;
  double cos_p, sin_p;
  int nt;
  double *wt;
  double *ct;
  double *st;
  double *cp;
  double *sp;
} RAbD;

typedef struct rab {              /* focused radially / azimuthally polarised beam */
  double lambda0;
  double ni;
  double NA;
  double power;
  double complex e0r;
  double complex e0a;
  double fx, fy, fz;
  double theta, phi;
  int nn;
  RAbD data;
} RAb;

typedef struct beam_data {
  Ipw *ipw;
  Fpw *fpw;
  LGb *lgb;
  Bsb *bsb;
  BsLGb *blg;
  RAb *rab;
} Bdata;

typedef struct beam_object {
  int n_ipw; char fname_ipw[28];
  int n_fpw; char fname_fpw[28];
  int n_lgb; char fname_lgb[28];
  int n_bsb; char fname_bsb[28];
  int n_blg; char fname_blg[28];
  int n_rab; char fname_rab[28];
  double n_0;
  double lambda_0;
  double omega;
  Bdata bd;
} Bobj;

/* ---- object data (d3b1_elem.h:13-97) ---- */
typedef struct coef_matrix {
  int type;
  int MN;
  char **tgfn;
  char **thfn;
  char **tdgfn;
  char **tdhfn;
  char **tdffn;
  char *lupfn;
  size_t nn;
  size_t na;
  size_t nnz;
  char *aval;
  char *aptr;
  char *aidx;
  char *b;
} CMD;

typedef struct sub_domain_data {
  int Ne;
  int *sid;
  double complex ***U, ***dU;
  double complex ***E, ***dE;
  double complex ***H, ***dH;
} SUBD;

typedef struct boundary_data {
  int Nn;
  int Ne;
  double **rn;
  int **ed;
  int NN;
  int **eni;
  int *md;
  int *sd;
  int *gd;
  double ***cr, ***cw;
  double ***ren, ***wen;
  double complex ***Ui, ***dUi;
  double complex ***Ui0, ***dUi0;
  double complex ***Uip;
  double zt_44[4], et_44[4], wt_44[4], zt_49[9], et_49[9], wt_49[9];
  double zt_34[4], et_34[4], wt_34[4], zt_37[7], et_37[7], wt_37[7];
  double xli[GLN], wli[GLN];
  double xhi[GHN], whi[GHN];
  SUBD *sb;
} BOUD;

typedef struct domain_data {
  char med_fn[128], msh_fn[128];
  double rv[3], th;
  double tv[3];
  int MN;
  double complex *n;
  double complex *kn;
  Bobj mw;
  BOUD bd;
  CMD cm;
  void *dev;          /* device copy (bh_object *), owned by the library */
} DOMD;

typedef struct multi_object_data {
  int N;
  char **mod_fn;
  double **rs;
  DOMD *md;
} MOBJ;

/* ---- create_matrix side (setup2.c, solve_bieq2.c) ---- */
void read_domd2(int argc, char **argv, DOMD *md);
void print_domd2(DOMD *md);
void print_domd2_mksa(DOMD *md);
void initialize_domd2(DOMD *md);
void finalize_domd2(DOMD *md);
int domain_id(double *rt, DOMD *md);
void create_matrix(DOMD *md, char *ofn);

/* ---- bv_solver side (mo_setup.c, mo_solve.c) ---- */
void mo_initialize(int argc, char **argv, MOBJ *mo);
void mo_print_data(MOBJ *mo);
void mo_print_data_mksa(MOBJ *mo);
void mo_finalize(MOBJ *mo);
void mo_dat_write(char *fname, MOBJ *mo);
void mo_dat_read(char *fname, MOBJ *mo);
void mo_output_node_particles(char *fname, MOBJ *mo);
void mo_solve_bv(MOBJ *mo);

/* ---- analysis (mo_field.c, mo_force.c, mo_absorb.c) ---- */
void mo_object_domain_id(int *oid, int *did, double *rt, MOBJ *mo);
int mo_mEMP_s(double complex *U, double *rt, int type, MOBJ *mo);
int mo_mEMP_t(double complex *U, double *rt, int type, MOBJ *mo);
int mo_mEMP_i(double complex *U, double *rt, int type, MOBJ *mo);
int mo_EH_s(double complex *E, double complex *H, double *rt, int type, MOBJ *mo);
int mo_EH_t(double complex *E, double complex *H, double *rt, int type, MOBJ *mo);
int mo_EH_i(double complex *E, double complex *H, double *rt, int type, MOBJ *mo);
int mo_EH_mEMP_s(double complex *E, double complex *H, double *rt, int type, MOBJ *mo);
int mo_EH_mEMP_t(double complex *E, double complex *H, double *rt, int type, MOBJ *mo);
int mo_EH_mEMP_i(double complex *E, double complex *H, double *rt, int type, MOBJ *mo);
void mo_EH_s_bd(double complex *E, double complex *H, int oid, int did, int t, double zeta_t, double eta_t, int type, MOBJ *mo);
void mo_EH_t_bd(double complex *E, double complex *H, int oid, int did, int t, double zeta_t, double eta_t, int type, MOBJ *mo);
void mo_EH_i_bd(double complex *E, double complex *H, int oid, int did, int t, double zeta_t, double eta_t, int type, MOBJ *mo);
void mo_EH_a_bd(double complex *E, double complex *H, int oid, int did, int t, double zeta_t, double eta_t, int type, MOBJ *mo);
int mo_force_FN(double *F, double *N, double *rc, int oid, int type, MOBJ *mo);
int mo_absorb_P(double *P, int oid, int type, MOBJ *mo);

/* ---- element layer (d3b1_elem.h:101-124; the constants d3b1_const.h:44-45) ---- */
#define ELT3 2
#define ELT4 3
int check_element_type(int seid, BOUD *bd);
void r_bd(double *r, int seid, double zeta, double eta, BOUD *bd);
void r_bd_node(double *r, int seid, int nid, BOUD *bd);
void n_bd_node(double *n, int seid, int nid, BOUD *bd);
void r_tz_bd(double *r, double *vtz, int seid, double zeta, double eta, BOUD *bd);
void r_te_bd(double *r, double *vte, int seid, double zeta, double eta, BOUD *bd);
void Tz_bd(double *vTz, int seid, double zeta, double eta, BOUD *bd);
void Te_bd(double *vTe, int seid, double zeta, double eta, BOUD *bd);
void tz_te_bd_node(double *vtz, double *vte, int seid, int nid, BOUD *bd);
void coef_rt(double complex *CC, double *rt, int s, double complex k, int type, BOUD *bd);
void coef_bd(double complex *CC, double *rt, int t, double zeta_t, double eta_t, int s, double complex k, int type, BOUD *bd);
void coef_bd_node(double complex *CC, int t, int nid, int s, double complex k, int type, BOUD *bd);
void dcoef_rt(double complex *CC, double complex *dCC, double *rt, double *vt, int s, double complex k, int type, BOUD *bd);
void dcoef_bd_tz(double complex *CC, double complex *dCC, double *rt, double *vtz, int t, double zeta_t, double eta_t,
                 int s, double complex k, int type, BOUD *bd);
void dcoef_bd_te(double complex *CC, double complex *dCC, double *rt, double *vte, int t, double zeta_t, double eta_t,
                 int s, double complex k, int type, BOUD *bd);
void dcoef_bd_t2(double complex *CC, double complex *dCdtz, double complex *dCdte, double *rt, double *vtz,
                 double *vte, int t, double zeta_t, double eta_t, int s, double complex k, int type, BOUD *bd);
void dcoef_bd_node_t2(double complex *CC, double complex *dCdtz, double complex *dCdte, int t, int nid, double *vtz,
                      double *vte, int s, double complex k, int type, BOUD *bd);

/* ---- utility the reference's examples call (com_src/my_utils.h:32):
 *      calloc that prints txt and exits on failure ---- */
void *m_alloc2(size_t num, size_t size, char *txt);

/* ---- batched forms (no counterpart in the reference; one device call
 *      for many points instead of one call per point) ---- */
int mo_EH_t_batch(double complex *E, double complex *H, const double *rt, int np, int type, MOBJ *mo, int *oid);
int mo_EH_s_batch(double complex *E, double complex *H, const double *rt, int np, int type, MOBJ *mo, int *oid);

#endif /* BEM3_EMF_B2_H_ */
