/*
 * bem_host.c -- host side of the drop-in bem3_emf_b2 API (include/bem3_emf_b2.h).
 *
 * Reads the reference's input files (medium, Gmsh 2.2 mesh, the six beam files,
 * object settings), keeps the reference's in-memory tables (so callers that
 * read MOBJ / DOMD keep working) and the reference's binary files (.obj of
 * create_matrix, .dat of the solver), and runs every computation of the
 * hot path through the device C-ABI (include/bem3_hip.h).
 *
 * Differences from the reference, by design:
 *  - create_matrix keeps the factorised system in HBM; the .obj file is
 *    written in the reference layout, the coefficient-matrix files only when
 *    BEM3_WRITE_CMAT=1 (mo_initialize rebuilds the matrices on the GPU from
 *    the .obj geometry, which is faster than reading them back);
 *  - the beam records' derived-data pointers (quadrature tables) are left
 *    NULL: the device builds its own tables.
 */
#include "../../../include/bem3_emf_b2.h"
#include "../../../include/bem3_hip.h"

#include <stddef.h>
#include <time.h>


void *m_alloc2(size_t num, size_t size, char *txt)
{
  void *p = calloc(num ? num : 1, size);
  if (!p) { printf("memory allocation error!\n%s. Exit..\n", txt); exit(1); }
  return p;
}

static void *xalloc(size_t n, size_t sz, const char *what) { return m_alloc2(n, sz, (char *)what); }

static void die(const char *msg)
{
  printf("%s Exit...\n", msg);
  exit(1);
}

static void bh_or_die(int rc, const char *where)
{
  if (rc != 0) { printf("%s: device error %d (%s). Exit...\n", where, rc, bh_last_error()); exit(1); }
}

static double vabs3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static void vunit3(double *v) { double a = vabs3(v); if (a != 0.0) { v[0] /= a; v[1] /= a; v[2] /= a; } }

/* ------------------------------------------------------------ beams */
/* The six beam files of multi_fbeam (multi_fbeam.h:10-16, read by
 * read_data_mfb_*, multi_fbeam.c:145-762).  Every record starts with the
 * same doubles -- lambda0, ni, [NA or deflection angle,] power, e0x, e0y,
 * fx, fy, fz, theta, phi -- then the type's int parameters; the reader fills
 * those members in that order, as the reference's readers do. */
enum { BT_IPW, BT_FPW, BT_LGB, BT_BSB, BT_BLG, BT_RAB, BT_N };
typedef struct {
  const char *file, *title;
  size_t size;
  int nd;               /* leading doubles (complex values as two) */
  int nint;             /* trailing int members */
  size_t oint[2];       /* their offsets */
} btype_t;
static const btype_t BT[BT_N] = {
  {"ipw.txt", "plane wave", sizeof(Ipw), 12, 0, {0, 0}},
  {"fpw.txt", "focused plane wave", sizeof(Fpw), 13, 1, {offsetof(Fpw, nn), 0}},
  {"lgb.txt", "focused plane wave with spiral phase modulation", sizeof(LGb), 13, 2, {offsetof(LGb, lg_m), offsetof(LGb, nn)}},
  {"bsb.txt", "Bessel beam", sizeof(Bsb), 13, 0, {0, 0}},
  {"blg.txt", "Bessel beam with spiral phase modulation", sizeof(BsLGb), 13, 1, {offsetof(BsLGb, lg_m), 0}},
  {"rab.txt", "focused radially and azimuthally polarized beam", sizeof(RAb), 13, 1, {offsetof(RAb, nn), 0}},
};
_Static_assert(offsetof(Fpw, phi) == 96 && offsetof(LGb, phi) == 96 && offsetof(Bsb, phi) == 96 &&
               offsetof(BsLGb, phi) == 96 && offsetof(RAb, phi) == 96 && offsetof(Ipw, phi) == 88,
               "beam records start with their doubles");

static int *bt_count(Bobj *b, int t)
{
  int *c[BT_N] = {&b->n_ipw, &b->n_fpw, &b->n_lgb, &b->n_bsb, &b->n_blg, &b->n_rab};
  return c[t];
}
static char *bt_fname(Bobj *b, int t)
{
  char *c[BT_N] = {b->fname_ipw, b->fname_fpw, b->fname_lgb, b->fname_bsb, b->fname_blg, b->fname_rab};
  return c[t];
}
static void **bt_data(Bobj *b, int t)
{
  void **c[BT_N] = {(void **)&b->bd.ipw, (void **)&b->bd.fpw, (void **)&b->bd.lgb, (void **)&b->bd.bsb,
                    (void **)&b->bd.blg, (void **)&b->bd.rab};
  return c[t];
}
static char *bt_rec(Bobj *b, int t, int i) { return (char *)*bt_data(b, t) + BT[t].size * (size_t)i; }

static void beams_init(Bobj *b)
{
  memset(b, 0, sizeof *b);
  for (int t = 0; t < BT_N; t++) {
    strcpy(bt_fname(b, t), "null");
    *bt_data(b, t) = xalloc(1, BT[t].size, "beams");
  }
}

static void beams_free(Bobj *b)
{
  for (int t = 0; t < BT_N; t++) { free(*bt_data(b, t)); *bt_data(b, t) = NULL; *bt_count(b, t) = 0; }
}

/* one beam file of the working directory; the number of beams it defines */
static int beams_read_file(Bobj *b, int t)
{
  FILE *fp = fopen(BT[t].file, "r");
  if (!fp) return 0;
  char buf[256], msg[160];
  int nn;
  snprintf(msg, sizeof msg, "multi_fbeam.c, read_data_mfb(), %s: failed to read the header.", BT[t].file);
  if (!fgets(buf, sizeof buf, fp) || !fgets(buf, sizeof buf, fp) || fscanf(fp, "%d\n", &nn) != 1 || !fgets(buf, sizeof buf, fp))
    die(msg);
  snprintf(bt_fname(b, t), 28, "%s", BT[t].file);
  if (nn <= 0) { fclose(fp); return 0; }
  free(*bt_data(b, t));
  *bt_data(b, t) = xalloc(nn, BT[t].size, "beams");
  *bt_count(b, t) = nn;
  snprintf(msg, sizeof msg, "multi_fbeam.c, read_data_mfb(), %s: failed to read a beam parameter.", BT[t].file);
  for (int i = 0; i < nn; i++) {
    double *v = (double *)bt_rec(b, t, i);
    for (int k = 0; k < BT[t].nd; k++)
      if (fscanf(fp, "%lf", &v[k]) != 1) die(msg);
    for (int k = 0; k < BT[t].nint; k++)
      if (fscanf(fp, "%d", (int *)(bt_rec(b, t, i) + BT[t].oint[k])) != 1) die(msg);
  }
  fclose(fp);
  return nn;
}

/* read_data_mfb (multi_fbeam.c:20-36): every beam file present */
static void beams_read(Bobj *b)
{
  int wn = 0;
  for (int t = 0; t < BT_N; t++) wn += beams_read_file(b, t);
  if (wn == 0) {
    printf("read_data_mfb() No beam defined, check beam data file : %s, %s, %s, %s, %s, %s.\n", BT[0].file, BT[1].file,
           BT[2].file, BT[3].file, BT[4].file, BT[5].file);
    exit(1);
  }
}

/* the members common to every record */
typedef struct {
  double lambda0, ni, na, power;
  double complex e0x, e0y;
  double fx, fy, fz, theta, phi;
  int lg_m, nn;
} bcommon_t;
static bcommon_t bt_common(Bobj *b, int t, int i)
{
  bcommon_t c;
  memset(&c, 0, sizeof c);
  const double *v = (const double *)bt_rec(b, t, i);
  const int k = t == BT_IPW ? 0 : 1;   /* the plane wave has no third double */
  c.lambda0 = v[0]; c.ni = v[1]; c.na = k ? v[2] : 0.0; c.power = v[2 + k];
  c.e0x = v[3 + k] + I * v[4 + k]; c.e0y = v[5 + k] + I * v[6 + k];
  c.fx = v[7 + k]; c.fy = v[8 + k]; c.fz = v[9 + k]; c.theta = v[10 + k]; c.phi = v[11 + k];
  if (t == BT_LGB || t == BT_BLG) c.lg_m = *(const int *)(bt_rec(b, t, i) + BT[t].oint[0]);
  if (t == BT_FPW || t == BT_RAB) c.nn = *(const int *)(bt_rec(b, t, i) + BT[t].oint[0]);
  if (t == BT_LGB) c.nn = *(const int *)(bt_rec(b, t, i) + BT[t].oint[1]);
  return c;
}

/* check_data_mfb (multi_fbeam.c:878-962), then the derived data of the
 * set-up functions that live in the records (ipw.c:65-89, fpw.c:115-133,
 * lgb.c:116-134, bsb.c:112-127, bslgb.c:115-130, rab.c:111-130; the
 * quadrature tables are the device's) */
static void beams_setup(Bobj *b)
{
  static const char *tag[BT_N] = {"ipw", "fpw", "lgb", "bsb", "blg", "rab"};
  int st = 0, tot = 0;
  b->n_0 = 0.0;
  b->lambda_0 = 0.0;
  for (int t = 0; t < BT_N; t++)
    for (int i = 0; i < *bt_count(b, t); i++, tot++) {
      bcommon_t c = bt_common(b, t, i);
      if (st == 0) { b->n_0 = c.ni; b->lambda_0 = c.lambda0; st = 1; }
      else {
        if (b->n_0 != c.ni) { printf("data error %s ni, exit\n", tag[t]); exit(1); }
        if (b->lambda_0 != c.lambda0) { printf("data error %s lambda0, exit\n", tag[t]); exit(1); }
      }
    }
  if (tot == 0) { printf("no beams defined. check datafile name. Exit...\n"); exit(0); }
  for (int i = 0; i < b->n_ipw; i++) {
    Ipw *w = &b->bd.ipw[i];
    double cs = 1.0 / sqrt(pow(cabs(w->e0x), 2) + pow(cabs(w->e0y), 2));
    double complex ex = w->e0x * cs, ey = w->e0y * cs, hx = -ey * w->ni, hy = ex * w->ni;
    w->data.E0 = sqrt(w->power * 2.0 / w->ni);
    w->data.ki = 2.0 * M_PI * w->ni / w->lambda0;
    double sn = sin(w->theta), ct = cos(w->theta), sp = sin(w->phi), cp = cos(w->phi);
    w->data.sin_t = sn; w->data.cos_t = ct; w->data.sin_p = sp; w->data.cos_p = cp;
    double a11 = sp * sp + cp * cp * ct, a12 = sp * cp * ct - sp * cp, a22 = sp * sp * ct + cp * cp;
    w->data.ex = ex * a11 + ey * a12; w->data.ey = ex * a12 + ey * a22; w->data.ez = -(ex * cp + ey * sp) * sn;
    w->data.hx = hx * a11 + hy * a12; w->data.hy = hx * a12 + hy * a22; w->data.hz = -(hx * cp + hy * sp) * sn;
  }
#define FOCUSED_SETUP(p)                                                                                   \
  do {                                                                                                   \
    double kr_ = (p)->NA / (p)->ni, ce_ = sqrt(1.0 / M_PI) / (kr_ * sqrt(pow(cabs((p)->e0x), 2) + pow(cabs((p)->e0y), 2))); \
    (p)->data.ki = 2.0 * M_PI * (p)->ni / (p)->lambda0;                                                  \
    (p)->data.ex = (p)->e0x * ce_; (p)->data.ey = (p)->e0y * ce_;                                        \
    (p)->data.hx = -(p)->data.ey * (p)->ni; (p)->data.hy = (p)->data.ex * (p)->ni;                       \
    (p)->data.E0 = sqrt(2.0 * (p)->power * (p)->ni) / (p)->lambda0;                                      \
    (p)->data.cos_t = cos((p)->theta); (p)->data.sin_t = sin((p)->theta);                                \
    (p)->data.cos_p = cos((p)->phi); (p)->data.sin_p = sin((p)->phi);                                    \
    (p)->data.nt = (p)->nn;                                                                              \
  } while (0)
#define BESSEL_SETUP(p)                                                                                    \
  do {                                                                                                   \
    double ce_ = 1.0 / sqrt(pow(cabs((p)->e0x), 2) + pow(cabs((p)->e0y), 2));                             \
    (p)->data.st = sin((p)->d_angle); (p)->data.ct = cos((p)->d_angle);                                  \
    (p)->data.ki = 2.0 * M_PI * (p)->ni / (p)->lambda0;                                                  \
    (p)->data.ex = (p)->e0x * ce_; (p)->data.ey = (p)->e0y * ce_;                                        \
    (p)->data.hx = -(p)->data.ey * (p)->ni; (p)->data.hy = (p)->data.ex * (p)->ni;                       \
    (p)->data.E0 = sqrt(2.0 * (p)->power / (p)->ni);                                                     \
    (p)->data.cos_t = cos((p)->theta); (p)->data.sin_t = sin((p)->theta);                                \
    (p)->data.cos_p = cos((p)->phi); (p)->data.sin_p = sin((p)->phi);                                    \
  } while (0)
  for (int i = 0; i < b->n_fpw; i++) FOCUSED_SETUP(&b->bd.fpw[i]);
  for (int i = 0; i < b->n_lgb; i++) FOCUSED_SETUP(&b->bd.lgb[i]);
  for (int i = 0; i < b->n_bsb; i++) BESSEL_SETUP(&b->bd.bsb[i]);
  for (int i = 0; i < b->n_blg; i++) BESSEL_SETUP(&b->bd.blg[i]);
  for (int i = 0; i < b->n_rab; i++) {
    RAb *r = &b->bd.rab[i];
    double kr = r->NA / r->ni, ce = sqrt(1.0 / M_PI) / (kr * sqrt(pow(cabs(r->e0r), 2) + pow(cabs(r->e0a), 2)));
    r->data.ki = 2.0 * M_PI * r->ni / r->lambda0;
    r->data.er = r->e0r * ce; r->data.ea = r->e0a * ce;
    r->data.hr = -r->data.ea * r->ni; r->data.ha = r->data.er * r->ni;
    r->data.E0 = sqrt(2.0 * r->power * r->ni) / r->lambda0;
    r->data.cos_t = cos(r->theta); r->data.sin_t = sin(r->theta);
    r->data.cos_p = cos(r->phi); r->data.sin_p = sin(r->phi);
    r->data.nt = r->nn;
  }
#undef FOCUSED_SETUP
#undef BESSEL_SETUP
  /* pointers of the derived data are not used on the host (and stale in a file) */
  for (int i = 0; i < b->n_fpw; i++) { FpwD *d = &b->bd.fpw[i].data; d->wt = d->ct = d->st = d->cp = d->sp = NULL; }
  for (int i = 0; i < b->n_lgb; i++) { LGbD *d = &b->bd.lgb[i].data; d->wt = d->ct = d->st = d->cp = d->sp = NULL; d->ph_p = NULL; }
  for (int i = 0; i < b->n_bsb; i++) { BsbD *d = &b->bd.bsb[i].data; d->cp = d->sp = NULL; }
  for (int i = 0; i < b->n_blg; i++) { BsLGbD *d = &b->bd.blg[i].data; d->cp = d->sp = NULL; d->ph_p = NULL; }
  for (int i = 0; i < b->n_rab; i++) { RAbD *d = &b->bd.rab[i].data; d->wt = d->ct = d->st = d->cp = d->sp = NULL; }
  b->omega = 2.0 * M_PI / b->lambda_0;
}

static int beams_total(Bobj *b)
{
  int n = 0;
  for (int t = 0; t < BT_N; t++) n += *bt_count(b, t);
  return n;
}

/* the device's beam records (16 doubles each, bh_object_set_mfb), in the
 * type order in which calc_mfb_EH sums them */
static void beam_params(Bobj *b, double *par)
{
  int k = 0;
  for (int t = 0; t < BT_N; t++)
    for (int i = 0; i < *bt_count(b, t); i++, k++) {
      bcommon_t c = bt_common(b, t, i);
      double v[16] = {(double)t, c.lambda0, c.ni, c.na, c.power, creal(c.e0x), cimag(c.e0x), creal(c.e0y), cimag(c.e0y),
                      c.fx, c.fy, c.fz, c.theta, c.phi, (double)c.lg_m, (double)c.nn};
      memcpy(par + 16 * k, v, sizeof v);
    }
}

/* print_data_mfb (multi_fbeam.c:38-67) with each type's print function */
static void print_beams(Bobj *b, int mksa)
{
  for (int t = 0; t < BT_N; t++)
    for (int i = 0; i < *bt_count(b, t); i++) {
      bcommon_t c = bt_common(b, t, i);
      printf(" \"%s\" No.%02d -- %s%s --\n", bt_fname(b, t), i, BT[t].title, mksa ? ", MKSA system" : "");
      printf("wave length of incident beam in vacuum      : %15.14g\n", c.lambda0);
      printf("refractive index of surrounding             : %15.14g\n", c.ni);
      if (t == BT_FPW || t == BT_LGB || t == BT_RAB) printf("numerical aperture                          : %15.14g\n", c.na);
      if (t == BT_BSB || t == BT_BLG) printf("deflection angle                       [rad]: %15.14g\n", c.na);
      if (t == BT_IPW || t == BT_BSB || t == BT_BLG) printf("incident beam power per unit area           : %15.14g\n", c.power);
      else printf("incident beam power                         : %15.14g\n", c.power);
      if (t == BT_RAB) {
        printf("radial-component of polarization coefficient:%7.6g+%7.6gi\n", creal(c.e0x), cimag(c.e0x));
        printf("azimuthal-component of polarization coef    :%7.6g+%7.6gi\n", creal(c.e0y), cimag(c.e0y));
      } else {
        printf("x-component of polarization coefficient     :%7.6g+%7.6gi\n", creal(c.e0x), cimag(c.e0x));
        printf("y-component of polarization coefficient     :%7.6g+%7.6gi\n", creal(c.e0y), cimag(c.e0y));
      }
      printf("x-component of translation vector           : %15.14g\n", c.fx);
      printf("y-component of translation vector           : %15.14g\n", c.fy);
      printf("z-component of translation vector           : %15.14g\n", c.fz);
      printf("rotation parameter theta               [rad]: %15.14g\n", c.theta);
      printf("rotation parameter phi                 [rad]: %15.14g\n", c.phi);
      if (t == BT_LGB || t == BT_BLG) printf("mode number for spiral phase modulation     : %15d\n", c.lg_m);
      if (t == BT_FPW || t == BT_LGB || t == BT_RAB) printf("gauss-legendre integration sampling number  : %15d\n", c.nn);
    }
  printf("\n");
}

/* the beam records of a .obj / .dat file after the raw Bobj
 * (solve_bieq2.c:459-465, mo_setup.c:170-176) */
static void beams_write(FILE *fp, Bobj *b)
{
  fwrite(b, sizeof(Bobj), 1, fp);
  for (int t = 0; t < BT_N; t++) fwrite(*bt_data(b, t), BT[t].size, *bt_count(b, t), fp);
}

/* ------------------------------------------------------------ tables */
static void alloc_nodes(BOUD *bd)
{
  bd->rn = xalloc(bd->Nn + 1, sizeof(double *), "rn");
  for (int i = 0; i <= bd->Nn; i++) bd->rn[i] = xalloc(3, sizeof(double), "rn[i]");
}

static double ***alloc3d(int n1, int n2, int n3)
{
  double ***p = xalloc(n1, sizeof(double **), "table");
  for (int i = 0; i < n1; i++) {
    p[i] = xalloc(n2, sizeof(double *), "table");
    for (int j = 0; j < n2; j++) p[i][j] = xalloc(n3, sizeof(double), "table");
  }
  return p;
}

static double complex ***alloc3z(int n1, int n2, int n3)
{
  double complex ***p = xalloc(n1, sizeof(double complex **), "table");
  for (int i = 0; i < n1; i++) {
    p[i] = xalloc(n2, sizeof(double complex *), "table");
    for (int j = 0; j < n2; j++) p[i][j] = xalloc(n3, sizeof(double complex), "table");
  }
  return p;
}

static void free3(void ***p, int n1, int n2)
{
  if (!p) return;
  for (int i = 0; i < n1; i++) { for (int j = 0; j < n2; j++) free(p[i][j]); free(p[i]); }
  free(p);
}

static void alloc_elems(BOUD *bd)
{
  int n = bd->Ne + 1;
  bd->ed = xalloc(n, sizeof(int *), "ed");
  bd->eni = xalloc(n, sizeof(int *), "eni");
  for (int i = 0; i < n; i++) { bd->ed[i] = xalloc(4, sizeof(int), "ed[i]"); bd->eni[i] = xalloc(4, sizeof(int), "eni[i]"); }
  bd->md = xalloc(n, sizeof(int), "md"); bd->sd = xalloc(n, sizeof(int), "sd"); bd->gd = xalloc(n, sizeof(int), "gd");
  bd->cr = alloc3d(n, 3, 4); bd->cw = alloc3d(n, 3, 3);
  bd->ren = alloc3d(n, 4, 3); bd->wen = alloc3d(n, 4, 3);
  bd->Ui = alloc3z(n, 4, 4); bd->dUi = alloc3z(n, 4, 4); bd->Ui0 = alloc3z(n, 4, 4); bd->dUi0 = alloc3z(n, 4, 4);
  bd->Uip = alloc3z(n, 4, 4);
}

static void free_elems(BOUD *bd)
{
  int n = bd->Ne + 1;
  for (int i = 0; i < n; i++) { free(bd->ed[i]); free(bd->eni[i]); }
  free(bd->ed); free(bd->eni); free(bd->md); free(bd->sd); free(bd->gd);
  free3((void ***)bd->cr, n, 3); free3((void ***)bd->cw, n, 3);
  free3((void ***)bd->ren, n, 4); free3((void ***)bd->wen, n, 4);
  free3((void ***)bd->Ui, n, 4); free3((void ***)bd->dUi, n, 4); free3((void ***)bd->Ui0, n, 4);
  free3((void ***)bd->dUi0, n, 4); free3((void ***)bd->Uip, n, 4);
}

static void alloc_subdomains(DOMD *md)
{
  BOUD *bd = &md->bd;
  int *cnt = xalloc(md->MN + 1, sizeof(int), "count");
  for (int i = 1; i <= bd->Ne; i++) { cnt[bd->md[i]]++; cnt[bd->sd[i]]++; }
  bd->sb = xalloc(md->MN + 1, sizeof(SUBD), "sb");
  for (int d = 0; d <= md->MN; d++) {
    SUBD *s = &bd->sb[d];
    s->Ne = cnt[d];
    s->sid = xalloc(cnt[d] + 1, sizeof(int), "sid");
    s->U = alloc3z(cnt[d] + 1, 4, 4); s->dU = alloc3z(cnt[d] + 1, 4, 4);
    s->E = alloc3z(cnt[d] + 1, 4, 3); s->dE = alloc3z(cnt[d] + 1, 4, 3);
    s->H = alloc3z(cnt[d] + 1, 4, 3); s->dH = alloc3z(cnt[d] + 1, 4, 3);
  }
  free(cnt);
}

static void free_subdomains(DOMD *md)
{
  for (int d = 0; d <= md->MN; d++) {
    SUBD *s = &md->bd.sb[d];
    int n = s->Ne + 1;
    free(s->sid);
    free3((void ***)s->U, n, 4); free3((void ***)s->dU, n, 4); free3((void ***)s->E, n, 4);
    free3((void ***)s->dE, n, 4); free3((void ***)s->H, n, 4); free3((void ***)s->dH, n, 4);
  }
  free(md->bd.sb);
}

/* signed element lists per domain (setup2.c:817-835) */
static void fill_sid(DOMD *md)
{
  for (int d = 0; d <= md->MN; d++) {
    int c = 1;
    for (int i = 1; i <= md->bd.Ne; i++) {
      if (md->bd.md[i] == d) md->bd.sb[d].sid[c++] = i;
      else if (md->bd.sd[i] == d) md->bd.sb[d].sid[c++] = -i;
    }
  }
}

/* rule tables of BOUD (setup2.c:719-748; quadrature nodes from the device
 * library's tables are not needed on the host, Gauss-Legendre nodes are
 * kept for callers that read them) */
static void gl_host(int n, double *x, double *w)
{
  for (int i = 1; i <= (n + 1) / 2; i++) {
    double z = cos(M_PI * (i - 0.25) / (n + 0.5)), pd;
    for (;;) {
      double p0 = 1.0, p1 = 0.0;
      for (int j = 1; j <= n; j++) { double pm = p1; p1 = p0; p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * pm) / j; }
      pd = n * (z * p0 - p1) / (z * z - 1.0);
      double zo = z;
      z = zo - p0 / pd;
      if (fabs(z - zo) <= 3.0e-16) break;
    }
    x[i - 1] = -z; x[n - i] = z;
    w[i - 1] = 2.0 / ((1.0 - z * z) * pd * pd); w[n - i] = w[i - 1];
  }
  if (n % 2) x[n / 2] = 0.0;
}

static void rule_tables(BOUD *bd)
{
  const double p44 = 0.5773502691896257645091, p49 = 0.7745966692414833770359;
  const double z44[4] = {-p44, p44, p44, -p44}, e44[4] = {-p44, -p44, p44, p44};
  for (int i = 0; i < 4; i++) { bd->zt_44[i] = z44[i]; bd->et_44[i] = e44[i]; bd->wt_44[i] = 1.0; }
  const double z49[9] = {-p49, p49, p49, -p49, 0.0, p49, 0.0, -p49, 0.0}, e49[9] = {-p49, -p49, p49, p49, -p49, 0.0, p49, 0.0, 0.0};
  const double w49[9] = {0.3086419753086419753086, 0.3086419753086419753086, 0.3086419753086419753086, 0.3086419753086419753086,
                         0.4938271604938271604938, 0.4938271604938271604938, 0.4938271604938271604938, 0.4938271604938271604938,
                         0.7901234567901234567901};
  memcpy(bd->zt_49, z49, sizeof z49); memcpy(bd->et_49, e49, sizeof e49); memcpy(bd->wt_49, w49, sizeof w49);
  const double z34[4] = {-0.2, 0.4, -0.2, 0.0}, e34[4] = {-0.3464101615137754587055, 0.0, 0.3464101615137754587055, 0.0};
  const double w34[4] = {0.5208333333333333333333, 0.5208333333333333333333, 0.5208333333333333333333, -0.5625};
  memcpy(bd->zt_34, z34, sizeof z34); memcpy(bd->et_34, e34, sizeof e34); memcpy(bd->wt_34, w34, sizeof w34);
  const double a0 = 0.3480702390148154917985, a1 = 0.2052130961576726346557, a2 = 0.6028753385763033130539, a3 = 0.3554395089236065568357;
  const double z37[7] = {-a0, 2 * a0, -a0, a1, -2 * a1, a1, 0.0}, e37[7] = {-a2, 0.0, a2, a3, 0.0, -a3, 0.0};
  const double w37[7] = {0.1259391805448271525957, 0.1259391805448271525957, 0.1259391805448271525957, 0.1323941527885061807376,
                         0.1323941527885061807376, 0.1323941527885061807376, 0.225};
  memcpy(bd->zt_37, z37, sizeof z37); memcpy(bd->et_37, e37, sizeof e37); memcpy(bd->wt_37, w37, sizeof w37);
  gl_host(GLN, bd->xli, bd->wli);
  gl_host(GHN, bd->xhi, bd->whi);
}

/* element constants and node data of triangles (setup2.c:666-749, 838-869) */
/* element constants (setup2.c:666-749, 838-869): triangles on the
 * equilateral reference triangle, quadrangles r = c0 + c1 z + c2 e + c3 z e
 * on [-1,1]^2 with W(z,e) = dr/dz x dr/de = W0 + W1 z + W2 e and node slots
 * at the 2x2 Gauss points */
static void elem_const_quad(BOUD *bd, int i)
{
  double p[4][3];
  for (int j = 0; j < 4; j++) for (int d = 0; d < 3; d++) p[j][d] = bd->rn[bd->ed[i][j]][d];
  for (int d = 0; d < 3; d++) {
    bd->cr[i][d][0] = 0.25 * (p[0][d] + p[1][d] + p[2][d] + p[3][d]);
    bd->cr[i][d][1] = 0.25 * (p[1][d] + p[2][d] - p[0][d] - p[3][d]);
    bd->cr[i][d][2] = 0.25 * (p[2][d] + p[3][d] - p[0][d] - p[1][d]);
    bd->cr[i][d][3] = 0.25 * (p[0][d] + p[2][d] - p[1][d] - p[3][d]);
  }
  for (int d = 0; d < 3; d++) {
    int a = (d + 1) % 3, b = (d + 2) % 3;
    bd->cw[i][d][0] = bd->cr[i][a][1] * bd->cr[i][b][2] - bd->cr[i][b][1] * bd->cr[i][a][2];
    bd->cw[i][d][1] = bd->cr[i][a][1] * bd->cr[i][b][3] - bd->cr[i][b][1] * bd->cr[i][a][3];
    bd->cw[i][d][2] = bd->cr[i][a][3] * bd->cr[i][b][2] - bd->cr[i][b][3] * bd->cr[i][a][2];
  }
  for (int j = 0; j < 4; j++) {
    double z = bd->zt_44[j], e = bd->et_44[j];
    for (int d = 0; d < 3; d++) {
      bd->ren[i][j][d] = bd->cr[i][d][0] + bd->cr[i][d][1] * z + bd->cr[i][d][2] * e + bd->cr[i][d][3] * z * e;
      bd->wen[i][j][d] = bd->cw[i][d][0] + bd->cw[i][d][1] * z + bd->cw[i][d][2] * e;
    }
  }
}

static void elem_const(BOUD *bd)
{
  for (int i = 1; i <= bd->Ne; i++) {
    if (bd->ed[i][3] != 0) { elem_const_quad(bd, i); continue; }
    double p[3][3];
    for (int j = 0; j < 3; j++) for (int d = 0; d < 3; d++) p[j][d] = bd->rn[bd->ed[i][j]][d];
    for (int d = 0; d < 3; d++) {
      bd->cr[i][d][0] = (p[0][d] + p[1][d] + p[2][d]) / 3.0;
      bd->cr[i][d][1] = (-p[0][d] + 2.0 * p[1][d] - p[2][d]) / 3.0;
      bd->cr[i][d][2] = (p[2][d] - p[0][d]) / sqrt(3.0);
      bd->cr[i][d][3] = 0.0;
    }
    double u[3], v[3];
    for (int d = 0; d < 3; d++) { u[d] = p[1][d] - p[0][d]; v[d] = p[2][d] - p[0][d]; }
    for (int d = 0; d < 3; d++) {
      int a = (d + 1) % 3, b = (d + 2) % 3;
      bd->cw[i][d][0] = u[a] * v[b] - v[a] * u[b];
      bd->cw[i][d][1] = bd->cw[i][d][2] = 0.0;
    }
    for (int j = 0; j < 4; j++)
      for (int d = 0; d < 3; d++) {
        bd->ren[i][j][d] = bd->cr[i][d][0] + bd->cr[i][d][1] * bd->zt_34[j] + bd->cr[i][d][2] * bd->et_34[j];
        bd->wen[i][j][d] = bd->cw[i][d][0];
      }
  }
}

static void translate_obj(DOMD *md, const double *rv, double th, const double *tv)
{
  double nv[3] = {rv[0], rv[1], rv[2]};
  vunit3(nv);
  double st = sin(th), ct = cos(th), oc = 1.0 - ct;
  double M[9] = {ct + nv[0] * nv[0] * oc, nv[0] * nv[1] * oc - nv[2] * st, nv[2] * nv[0] * oc + nv[1] * st,
                 nv[0] * nv[1] * oc + nv[2] * st, ct + nv[1] * nv[1] * oc, nv[1] * nv[2] * oc - nv[0] * st,
                 nv[2] * nv[0] * oc - nv[1] * st, nv[1] * nv[2] * oc + nv[0] * st, ct + nv[2] * nv[2] * oc};
  for (int s = 1; s <= md->bd.Nn; s++) {
    double *r = md->bd.rn[s], q[3];
    for (int i = 0; i < 3; i++) q[i] = M[3 * i] * r[0] + M[3 * i + 1] * r[1] + M[3 * i + 2] * r[2] + tv[i];
    r[0] = q[0]; r[1] = q[1]; r[2] = q[2];
  }
}

/* device copy of the object (geometry, media, beams) */
static bh_object *dev(DOMD *md)
{
  if (md->dev) return (bh_object *)md->dev;
  BOUD *bd = &md->bd;
  double *rn = xalloc((size_t)(bd->Nn + 1) * 3, sizeof(double), "rn");
  int *ed = xalloc((size_t)(bd->Ne + 1) * 4, sizeof(int), "ed");
  for (int i = 0; i <= bd->Nn; i++) for (int d = 0; d < 3; d++) rn[3 * i + d] = bd->rn[i][d];
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) ed[4 * i + j] = bd->ed[i][j];
  double *nm = xalloc((size_t)(md->MN + 1) * 2, sizeof(double), "n");
  for (int i = 0; i <= md->MN; i++) { nm[2 * i] = creal(md->n[i]); nm[2 * i + 1] = cimag(md->n[i]); }
  bh_object *o = bh_object_create(bd->Nn, rn, bd->Ne, ed, bd->md, bd->sd, md->MN, nm, md->mw.lambda_0);
  free(rn); free(ed); free(nm);
  if (!o) { printf("device object: %s. Exit...\n", bh_last_error()); exit(1); }
  const int nb = beams_total(&md->mw);
  double *par = xalloc((size_t)nb * 16, sizeof(double), "beams");
  beam_params(&md->mw, par);
  bh_or_die(bh_object_set_mfb(o, nb, par), "bh_object_set_mfb");
  free(par);
  md->dev = o;
  return o;
}

/* ------------------------------------------------------------ create_matrix side */
void read_domd2(int argc, char **argv, DOMD *md)
{
  if (argc != 4 && argc != 11) {
    printf("This program needs command line arguments as follows.\n");
    printf("%s medium_datafile_name mesh_datafile_name output_datafile_name [rv_x rv_y rv_z theta tr_x tr_y tr_z ](optional)\n", argv[0]);
    printf("rv : vector defining rotation axis, theta : rotation angle ( using Rodrigues' rotation formula ), tr : translation vector\n");
    printf("Exit...\n");
    exit(0);
  }
  memset(md, 0, sizeof *md);
  beams_init(&md->mw);
  beams_read(&md->mw);
  /* medium file (setup2.c:153-193) */
  FILE *fp = fopen(argv[1], "r");
  if (!fp) { printf("Can not open the '%s' file. Exit...\n", argv[1]); exit(1); }
  snprintf(md->med_fn, sizeof md->med_fn, "%s", argv[1]);
  char buf[256];
  int MN;
  if (!fgets(buf, sizeof buf, fp) || !fgets(buf, sizeof buf, fp) || fscanf(fp, "%d\n", &MN) != 1 || !fgets(buf, sizeof buf, fp))
    die("read_medium_data(): failed to read the header.");
  md->MN = MN;
  md->n = xalloc(MN + 1, sizeof(double complex), "n");
  md->kn = xalloc(MN + 1, sizeof(double complex), "kn");
  for (int i = 1; i <= MN; i++) {
    double a, b;
    if (fscanf(fp, "%lf %lf", &a, &b) != 2) die("read_medium_data(): failed to read the refractive index.");
    md->n[i] = a + I * b;
  }
  fclose(fp);
  /* Gmsh 2.2 ASCII mesh, paired faces (setup2.c:205-503) */
  fp = fopen(argv[2], "r");
  if (!fp) { printf("Can not open the '%s' file. Exit...\n", argv[2]); exit(1); }
  snprintf(md->msh_fn, sizeof md->msh_fn, "%s", argv[2]);
  double ver;
  int ft, pr, Nn, Nel;
  if (!fgets(buf, sizeof buf, fp) || fscanf(fp, "%lf %d %d", &ver, &ft, &pr) != 3) die("read_mesh_data(): bad header.");
  if (ver != 2.2) { printf("This program supports mesh file version 2.2. Reading file version is %g. Exit...\n", ver); exit(1); }
  if (ft != 0) die("This program supports 'ASCII' data format mesh file.");
  if (pr != 8) die("This program supports double precision mesh data.");
  if (fscanf(fp, " %255s %255s %d", buf, buf, &Nn) != 3) die("read_mesh_data(): failed to read the node count.");
  BOUD *bd = &md->bd;
  bd->Nn = Nn;
  alloc_nodes(bd);
  for (int i = 1; i <= Nn; i++) {
    int id;
    if (fscanf(fp, "%d %lf %lf %lf", &id, &bd->rn[i][0], &bd->rn[i][1], &bd->rn[i][2]) != 4) die("read_mesh_data(): failed to read a node.");
    if (id != i) printf("bad id %d\n", id);
  }
  if (fscanf(fp, " %255s %255s %d", buf, buf, &Nel) != 3) die("read_mesh_data(): failed to read the element count.");
  bd->Ne = Nel / 2;
  alloc_elems(bd);
  int nc = 0;
  for (int i = 1; i <= bd->Ne; i++) {
    int id, et, nt, tag;
    if (fscanf(fp, "%d %d %d", &id, &et, &nt) != 3) die("read_mesh_data(): failed to read an element.");
    if (id != 2 * i - 1) { printf("bad id :%d. Exit...\n", id); exit(1); }
    if (et != 2 && et != 3) die("bad element type. element type must be 2 or 3.");
    for (int j = 0; j < nt; j++) {
      if (fscanf(fp, "%d", &tag) != 1) die("read_mesh_data(): failed to read a tag.");
      if (j == 0) {
        if (tag == 99) tag = 0;
        if (tag > MN) { printf("domain id %d is not defined medium data. check domain and medium data. exit..\n", tag); exit(1); }
        bd->md[i] = tag;
      } else if (j == 1) bd->gd[i] = tag;
    }
    int nv = et == 2 ? 3 : 4;
    for (int j = 0; j < nv; j++) if (fscanf(fp, "%d", &bd->ed[i][j]) != 1) die("read_mesh_data(): failed to read a node id.");
    if (nv == 3) bd->ed[i][3] = 0;
    for (int j = 0; j < 4; j++) bd->eni[i][j] = j < nv ? nc++ : -1;
    int id2, et2, nt2;
    if (fscanf(fp, "%d %d %d", &id2, &et2, &nt2) != 3) die("read_mesh_data(): failed to read an element.");
    if (id2 != 2 * i) { printf("bad id :%d. Exit...\n", id2); exit(1); }
    for (int j = 0; j < nt2; j++) {
      if (fscanf(fp, "%d", &tag) != 1) die("read_mesh_data(): failed to read a tag.");
      if (j == 0) {
        if (tag == 99) tag = 0;
        if (tag > MN) { printf("domain id %d is not defined medium data. check domain and medium data! exit..\n", tag); exit(1); }
        bd->sd[i] = tag;
      }
    }
    for (int j = 0; j < nv; j++) {
      int v, want = j == 0 ? bd->ed[i][0] : bd->ed[i][nv - j];
      if (fscanf(fp, "%d", &v) != 1) die("read_mesh_data(): failed to read a node id.");
      if (v != want) { printf("node id miss matched. check element id %d. Exit...\n", 2 * i); exit(1); }
    }
    if (bd->sd[i] == 0) {
      bd->sd[i] = bd->md[i];
      bd->md[i] = 0;
      if (nv == 3) { int t = bd->ed[i][1]; bd->ed[i][1] = bd->ed[i][2]; bd->ed[i][2] = t; }
      else { int t = bd->ed[i][3]; bd->ed[i][3] = bd->ed[i][1]; bd->ed[i][1] = t; }
    }
  }
  fclose(fp);
  bd->NN = nc;
  if (argc == 11) {
    md->rv[0] = atof(argv[4]); md->rv[1] = atof(argv[5]); md->rv[2] = atof(argv[6]); md->th = atof(argv[7]);
    md->tv[0] = atof(argv[8]); md->tv[1] = atof(argv[9]); md->tv[2] = atof(argv[10]);
  } else {
    md->rv[0] = 1.0;
  }
}

static void print_medium_mesh(DOMD *md)
{
  printf("-- medium data --\n");
  printf("medium data file name                  : %s\n", md->med_fn);
  for (int i = 1; i <= md->MN; i++) printf("medium (domain) id %2d refractive index :%8.7g + %8.7gI\n", i, creal(md->n[i]), cimag(md->n[i]));
  printf("\n-- mesh data --\n");
  printf("mesh data file name    : %s\n", md->msh_fn);
  printf("node number            : %8d\n", md->bd.Nn);
  printf("defined element number : %8d\n\n", md->bd.Ne * 2);
}

void print_domd2(DOMD *md)
{
  printf("-- multi_fbeam data --\n");
  print_beams(&md->mw, 0);
  print_medium_mesh(md);
}

void print_domd2_mksa(DOMD *md) { print_domd2(md); }

void initialize_domd2(DOMD *md)
{
  beams_setup(&md->mw);
  md->n[0] = md->mw.n_0;
  for (int i = 0; i <= md->MN; i++) md->kn[i] = 2.0 * M_PI / md->mw.lambda_0 * md->n[i];
  if (md->th != 0.0 || vabs3(md->tv) != 0.0) translate_obj(md, md->rv, md->th, md->tv);
  rule_tables(&md->bd);
  elem_const(&md->bd);
  alloc_subdomains(md);
  fill_sid(md);
}

static void free_names(CMD *cm)
{
  if (!cm->tgfn) return;
  for (int i = 0; i <= cm->MN; i++) { free(cm->tgfn[i]); free(cm->thfn[i]); free(cm->tdgfn[i]); free(cm->tdhfn[i]); free(cm->tdffn[i]); }
  free(cm->tgfn); free(cm->thfn); free(cm->tdgfn); free(cm->tdhfn); free(cm->tdffn); free(cm->lupfn);
  free(cm->aval); free(cm->aptr); free(cm->aidx); free(cm->b);
  cm->tgfn = NULL;
}

void finalize_domd2(DOMD *md)
{
  if (md->dev) { bh_object_destroy((bh_object *)md->dev); md->dev = NULL; }
  free_subdomains(md);
  free_elems(&md->bd);
  for (int i = 0; i <= md->bd.Nn; i++) free(md->bd.rn[i]);
  free(md->bd.rn);
  free(md->n); free(md->kn);
  beams_free(&md->mw);
  free_names(&md->cm);
}

int domain_id(double *rt, DOMD *md)
{
  int did = -1;
  bh_or_die(bh_object_domain(dev(md), 1, rt, &did), "domain_id");
  return did;
}

static void init_names(CMD *cm, int MN)
{
  cm->MN = MN;
  cm->tgfn = xalloc(MN + 1, sizeof(char *), "names"); cm->thfn = xalloc(MN + 1, sizeof(char *), "names");
  cm->tdgfn = xalloc(MN + 1, sizeof(char *), "names"); cm->tdhfn = xalloc(MN + 1, sizeof(char *), "names");
  cm->tdffn = xalloc(MN + 1, sizeof(char *), "names");
  for (int i = 0; i <= MN; i++) {
    cm->tgfn[i] = xalloc(128, 1, "name"); cm->thfn[i] = xalloc(128, 1, "name"); cm->tdgfn[i] = xalloc(128, 1, "name");
    cm->tdhfn[i] = xalloc(128, 1, "name"); cm->tdffn[i] = xalloc(128, 1, "name");
    snprintf(cm->tgfn[i], 128, "tmpG_%05d.cmat", i); snprintf(cm->thfn[i], 128, "tmpH_%05d.cmat", i);
    snprintf(cm->tdgfn[i], 128, "tmpdG_%05d.cmat", i); snprintf(cm->tdhfn[i], 128, "tmpdH_%05d.cmat", i);
    snprintf(cm->tdffn[i], 128, "tmpdF_%05d.cmat", i);
  }
  cm->lupfn = xalloc(128, 1, "name");
  snprintf(cm->lupfn, 128, "tmpinvA.cmat");
  cm->aval = xalloc(16, 1, "name"); cm->aptr = xalloc(16, 1, "name"); cm->aidx = xalloc(16, 1, "name"); cm->b = xalloc(16, 1, "name");
  snprintf(cm->aval, 16, "tmpAval.dat"); snprintf(cm->aptr, 16, "tmpAprt.dat"); snprintf(cm->aidx, 16, "tmpAidx.dat");
  snprintf(cm->b, 16, "tmpB.dat");
}

/* coefficient matrices of domain d in the reference's padded layout
 * (solve_bieq2.c:168-183), written only on request */
static void write_cmat(DOMD *md, int d, const char *gfn, const char *hfn, const char *dgfn, const char *dhfn, const char *dffn)
{
  bh_object *o = dev(md);
  int info[20];
  bh_object_info(o, info);
  size_t nd = (size_t)info[12 + d], Ne = (size_t)md->bd.sb[d].Ne, N = 4 * Ne;
  double complex *c = xalloc(2 * nd * nd, sizeof(double complex), "matrix");
  double complex *pad = xalloc(2 * N * N, sizeof(double complex), "matrix");
  double *f = xalloc(3 * nd, sizeof(double), "dF"), *fp3 = xalloc(3 * N, sizeof(double), "dF");
  const char *names[4] = {gfn, hfn, dgfn, dhfn};
  for (int w = 0; w < 4; w++) {
    bh_or_die(bh_object_matrix(o, d, w, (double *)c), "bh_object_matrix");
    size_t nr = w < 2 ? 1 : 2;
    memset(pad, 0, nr * N * N * sizeof(double complex));
    for (size_t r = 0; r < nd; r++)
      for (size_t q = 0; q < nr; q++)
        for (size_t j = 0; j < nd; j++)
          pad[((r / 3 * 4 + r % 3) * nr + q) * N + j / 3 * 4 + j % 3] = c[(r * nr + q) * nd + j];
    FILE *fo = fopen(names[w], "wb");
    if (!fo) { printf("Failed to create %s file.\n", names[w]); exit(1); }
    fwrite(pad, sizeof(double complex), nr * N * N, fo);
    fclose(fo);
  }
  bh_or_die(bh_object_matrix(o, d, 4, f), "bh_object_matrix");
  for (size_t r = 0; r < nd; r++) for (int q = 0; q < 3; q++) fp3[(r / 3 * 4 + r % 3) * 3 + q] = f[3 * r + q];
  FILE *fo = fopen(dffn, "wb");
  if (!fo) { printf("Failed to create %s file.\n", dffn); exit(1); }
  fwrite(fp3, sizeof(double), 3 * N, fo);
  fclose(fo);
  free(c); free(pad); free(f); free(fp3);
}

/* .obj file of create_matrix (dat_write2, solve_bieq2.c:439-502) */
static void write_obj(const char *fname, DOMD *md)
{
  FILE *fp = fopen(fname, "wb");
  if (!fp) { printf("dat_write2(), Failed to create the %s file.\n", fname); exit(1); }
  BOUD *bd = &md->bd;
  fwrite(md->med_fn, 1, 128, fp); fwrite(md->msh_fn, 1, 128, fp);
  fwrite(md->rv, sizeof(double), 3, fp); fwrite(&md->th, sizeof(double), 1, fp); fwrite(md->tv, sizeof(double), 3, fp);
  fwrite(&md->MN, sizeof(int), 1, fp);
  fwrite(md->n, sizeof(double complex), md->MN + 1, fp); fwrite(md->kn, sizeof(double complex), md->MN + 1, fp);
  beams_write(fp, &md->mw);
  fwrite(&bd->Nn, sizeof(int), 1, fp); fwrite(&bd->Ne, sizeof(int), 1, fp);
  for (int i = 0; i <= bd->Nn; i++) fwrite(bd->rn[i], sizeof(double), 3, fp);
  for (int i = 0; i <= bd->Ne; i++) fwrite(bd->ed[i], sizeof(int), 4, fp);
  fwrite(&bd->NN, sizeof(int), 1, fp);
  for (int i = 0; i <= bd->Ne; i++) fwrite(bd->eni[i], sizeof(int), 4, fp);
  fwrite(bd->md, sizeof(int), bd->Ne + 1, fp); fwrite(bd->sd, sizeof(int), bd->Ne + 1, fp); fwrite(bd->gd, sizeof(int), bd->Ne + 1, fp);
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) fwrite(bd->ren[i][j], sizeof(double), 3, fp);
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) fwrite(bd->wen[i][j], sizeof(double), 3, fp);
  for (int d = 0; d <= md->MN; d++) { fwrite(&bd->sb[d].Ne, sizeof(int), 1, fp); fwrite(bd->sb[d].sid, sizeof(int), bd->sb[d].Ne + 1, fp); }
  /* stored names: "<prefix>_G_00000.cmat" ..., prefix = fname up to its
   * first '.' (solve_bieq2.c:486-499); fields are 128 bytes, zero padded */
  char pfn[100] = {0};
  for (int i = 0; i < 99 && fname[i] && fname[i] != '.'; i++) pfn[i] = fname[i];
  int wcm = getenv("BEM3_WRITE_CMAT") && atoi(getenv("BEM3_WRITE_CMAT")) == 1;
  for (int d = 0; d <= md->MN; d++) {
    char *nm[5] = {md->cm.tgfn[d], md->cm.thfn[d], md->cm.tdgfn[d], md->cm.tdhfn[d], md->cm.tdffn[d]};
    char full[5][128];
    for (int k = 0; k < 5; k++) {
      memset(full[k], 0, 128);
      snprintf(full[k], 128, "%s_%.24s", pfn, nm[k] + 3);
    }
    if (wcm) write_cmat(md, d, full[0], full[1], full[2], full[3], full[4]);
    for (int k = 0; k < 5; k++) fwrite(full[k], 1, 128, fp);
  }
  char lu[128] = {0};
  snprintf(lu, 128, "%s_%.24s", pfn, md->cm.lupfn + 3);
  fwrite(lu, 1, 128, fp);
  fclose(fp);
}

void create_matrix(DOMD *md, char *ofn)
{
  time_t t0, t1, ms, me;
  printf("\ncreate matrix data \n");
  time(&t0);
  printf("  coefficient matrix          "); fflush(stdout);
  time(&ms);
  md->cm.type = PREC_DEF;
  init_names(&md->cm, md->MN);
  bh_object *o = dev(md);
  bh_or_die(bh_object_assemble(o, PREC_DEF), "bh_object_assemble");
  time(&me);
  printf("finished. Elapsed time : %5g (sec)\n", difftime(me, ms));
  printf("  inverse matrix              "); fflush(stdout);
  time(&ms);
  md->cm.nn = (size_t)md->bd.NN;
  md->cm.na = md->cm.nn * 8;
  bh_or_die(bh_object_factor(o, 0), "bh_object_factor");
  time(&me);
  printf("finished. Elapsed time : %5g (sec)\n", difftime(me, ms));
  write_obj(ofn, md);
  time(&t1);
  printf("Total elapsed time : %g (sec)\n", difftime(t1, t0));
}

/* ------------------------------------------------------------ bv_solver side */
static void mo_alloc(MOBJ *mo)
{
  mo->md = xalloc(mo->N, sizeof(DOMD), "md");
  mo->mod_fn = xalloc(mo->N, sizeof(char *), "mod_fn");
  mo->rs = xalloc(mo->N, sizeof(double *), "rs");
  for (int i = 0; i < mo->N; i++) { mo->mod_fn[i] = xalloc(FNLGTH, 1, "mod_fn"); mo->rs[i] = xalloc(3, sizeof(double), "rs"); }
}

static void rd(void *p, size_t sz, size_t n, FILE *fp, const char *what)
{
  if (fread(p, sz, n, fp) != n) { printf("failed to read the %s. exit...\n", what); exit(1); }
}

/* geometry part shared by .obj and .dat records */
static void read_geometry(FILE *fp, DOMD *md)
{
  rd(md->med_fn, 1, 128, fp, "med_fn"); rd(md->msh_fn, 1, 128, fp, "msh_fn");
  rd(md->rv, sizeof(double), 3, fp, "rv"); rd(&md->th, sizeof(double), 1, fp, "th"); rd(md->tv, sizeof(double), 3, fp, "tv");
  rd(&md->MN, sizeof(int), 1, fp, "MN");
  md->n = xalloc(md->MN + 1, sizeof(double complex), "n"); md->kn = xalloc(md->MN + 1, sizeof(double complex), "kn");
  rd(md->n, sizeof(double complex), md->MN + 1, fp, "n"); rd(md->kn, sizeof(double complex), md->MN + 1, fp, "kn");
  rd(&md->mw, sizeof(Bobj), 1, fp, "Bobj");
  for (int t = 0; t < BT_N; t++) {   /* the stored pointers are stale */
    *bt_data(&md->mw, t) = xalloc(*bt_count(&md->mw, t), BT[t].size, "beams");
    rd(*bt_data(&md->mw, t), BT[t].size, *bt_count(&md->mw, t), fp, BT[t].file);
  }
  beams_setup(&md->mw);
  BOUD *bd = &md->bd;
  rd(&bd->Nn, sizeof(int), 1, fp, "Nn"); rd(&bd->Ne, sizeof(int), 1, fp, "Ne");
  alloc_nodes(bd);
  alloc_elems(bd);
  for (int i = 0; i <= bd->Nn; i++) rd(bd->rn[i], sizeof(double), 3, fp, "rn");
  for (int i = 0; i <= bd->Ne; i++) rd(bd->ed[i], sizeof(int), 4, fp, "ed");
  rd(&bd->NN, sizeof(int), 1, fp, "NN");
  for (int i = 0; i <= bd->Ne; i++) rd(bd->eni[i], sizeof(int), 4, fp, "eni");
  rd(bd->md, sizeof(int), bd->Ne + 1, fp, "md"); rd(bd->sd, sizeof(int), bd->Ne + 1, fp, "sd"); rd(bd->gd, sizeof(int), bd->Ne + 1, fp, "gd");
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) rd(bd->ren[i][j], sizeof(double), 3, fp, "ren");
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) rd(bd->wen[i][j], sizeof(double), 3, fp, "wen");
  rule_tables(bd);
  alloc_subdomains(md);
}

static void read_obj(const char *fname, DOMD *md)
{
  FILE *fp = fopen(fname, "rb");
  if (!fp) { printf("dat_read2(), Failed to open the %s file.\n", fname); exit(1); }
  memset(md, 0, sizeof *md);
  read_geometry(fp, md);
  for (int d = 0; d <= md->MN; d++) {
    int t;
    rd(&t, sizeof(int), 1, fp, "sub-domain Ne");
    rd(md->bd.sb[d].sid, sizeof(int), md->bd.sb[d].Ne + 1, fp, "sid");
  }
  init_names(&md->cm, md->MN);
  for (int d = 0; d <= md->MN; d++) {
    rd(md->cm.tgfn[d], 1, 128, fp, "G name"); rd(md->cm.thfn[d], 1, 128, fp, "H name");
    rd(md->cm.tdgfn[d], 1, 128, fp, "dG name"); rd(md->cm.tdhfn[d], 1, 128, fp, "dH name");
    rd(md->cm.tdffn[d], 1, 128, fp, "dF name");
  }
  rd(md->cm.lupfn, 1, 128, fp, "LU name");
  md->cm.nn = (size_t)md->bd.NN;
  md->cm.na = 8 * md->cm.nn;
  fclose(fp);
}

void mo_initialize(int argc, char **argv, MOBJ *mo)
{
  if (argc != 3) {
    printf("This program needs command line argument as follows.\n");
    printf("%s object_setting_file output_data_file_name\n", argv[0]);
    exit(0);
  }
  FILE *fp = fopen(argv[1], "r");
  if (!fp) { printf("mo_read_data(),Can not open the %s file.\n", argv[1]); exit(1); }
  char tmp[FNLGTH];
  int n, flg;
  if (!fgets(tmp, FNLGTH, fp) || fscanf(fp, "%d %d\n", &n, &flg) != 2 || !fgets(tmp, FNLGTH, fp))
    die("mo_initialize(): failed to read the object settings.");
  mo->N = n;
  mo_alloc(mo);
  for (int i = 0; i < n; i++) {
    if (fscanf(fp, "%255s %lf %lf %lf", mo->mod_fn[i], &mo->rs[i][0], &mo->rs[i][1], &mo->rs[i][2]) != 4)
      die("mo_initialize(): failed to read the mod_fn,rs.");
    DOMD *md = &mo->md[i];
    read_obj(mo->mod_fn[i], md);
    if (flg == 1) {  /* renew beams from the working directory's beam files */
      double lam = md->mw.lambda_0, n0 = md->mw.n_0;
      beams_free(&md->mw);
      beams_init(&md->mw);
      beams_read(&md->mw);
      beams_setup(&md->mw);
      if (lam != md->mw.lambda_0 || n0 != md->mw.n_0) die("renew_beam_data(), wave number check error!");
    }
    translate_obj(md, md->rv, 0.0, mo->rs[i]);
    elem_const(&md->bd);
    bh_object *o = dev(md);
    bh_or_die(bh_object_assemble(o, PREC_DEF), "bh_object_assemble");
    bh_or_die(bh_object_factor(o, 0), "bh_object_factor");
  }
  fclose(fp);
}

void mo_print_data(MOBJ *mo)
{
  printf("-- multi_fbeam --\n");
  print_beams(&mo->md[0].mw, 0);
  for (int i = 0; i < mo->N; i++) {
    printf("********** object %d **********\n", i);
    printf("data file name                : %s\n", mo->mod_fn[i]);
    printf("additional translation vector : (% 8.7g,% 8.7g,% 8.7g)\n", mo->rs[i][0], mo->rs[i][1], mo->rs[i][2]);
    print_medium_mesh(&mo->md[i]);
  }
  printf("\n");
}

void mo_print_data_mksa(MOBJ *mo) { mo_print_data(mo); }

void mo_finalize(MOBJ *mo)
{
  for (int i = 0; i < mo->N; i++) { finalize_domd2(&mo->md[i]); free(mo->mod_fn[i]); free(mo->rs[i]); }
  free(mo->md); free(mo->mod_fn); free(mo->rs);
}

/* device boundary values -> the caller-visible tables */
static void download_state(DOMD *md)
{
  bh_object *o = dev(md);
  for (int d = 0; d <= md->MN; d++) {
    SUBD *s = &md->bd.sb[d];
    size_t n1 = (size_t)s->Ne + 1;
    double complex *buf = xalloc(n1 * 16, sizeof(double complex), "state");
    double complex ****tab[6] = {&s->U, &s->dU, &s->E, &s->dE, &s->H, &s->dH};
    for (int w = 0; w < 6; w++) {
      int nc = w < 2 ? 4 : 3;
      bh_or_die(bh_object_state(o, d, w, (double *)buf, 0), "bh_object_state");
      for (size_t i = 0; i < n1; i++) for (int j = 0; j < 4; j++) for (int c = 0; c < nc; c++) (*tab[w])[i][j][c] = buf[(i * 4 + j) * nc + c];
    }
    free(buf);
  }
  double complex *ui = xalloc((size_t)(md->bd.Ne + 1) * 16, sizeof(double complex), "incident");
  double complex ****itab[4] = {&md->bd.Ui, &md->bd.dUi, &md->bd.Ui0, &md->bd.dUi0};
  for (int w = 0; w < 4; w++) {
    bh_or_die(bh_object_incident(o, w, (double *)ui, 0), "bh_object_incident");
    for (int i = 0; i <= md->bd.Ne; i++) for (int j = 0; j < 4; j++) for (int c = 0; c < 4; c++) (*itab[w])[i][j][c] = ui[(i * 4 + j) * 4 + c];
  }
  free(ui);
}

static void upload_state(DOMD *md)
{
  bh_object *o = dev(md);
  for (int d = 0; d <= md->MN; d++) {
    SUBD *s = &md->bd.sb[d];
    size_t n1 = (size_t)s->Ne + 1;
    double complex *buf = xalloc(n1 * 16, sizeof(double complex), "state");
    double complex ***tab[6] = {s->U, s->dU, s->E, s->dE, s->H, s->dH};
    for (int w = 0; w < 6; w++) {
      int nc = w < 2 ? 4 : 3;
      for (size_t i = 0; i < n1; i++) for (int j = 0; j < 4; j++) for (int c = 0; c < nc; c++) buf[(i * 4 + j) * nc + c] = tab[w][i][j][c];
      bh_or_die(bh_object_state(o, d, w, (double *)buf, 1), "bh_object_state");
    }
    free(buf);
  }
}

static bh_object **handles(MOBJ *mo)
{
  bh_object **h = xalloc(mo->N, sizeof(bh_object *), "handles");
  for (int m = 0; m < mo->N; m++) h[m] = dev(&mo->md[m]);
  return h;
}

void mo_solve_bv(MOBJ *mo)
{
  time_t ms, me;
  bh_object **h = handles(mo);
  double *trace = xalloc(ITER_MAX + 1, sizeof(double), "trace");
  int iters = 0;
  printf(mo->N == 1 ? "\nsingle object\n" : "\ninitialize boundary value\n");
  time(&ms);
  bh_or_die(bh_mo_solve(h, mo->N, CCD, ITER_MIN, ITER_MAX, trace, &iters), "bh_mo_solve");
  time(&me);
  if (mo->N > 1) {
    printf("iterative operation start. convergence criterion cc < %g and count > %d\n", CCD, ITER_MIN);
    for (int it = 2; it <= iters && it < ITER_MAX; it++) printf("count =% 3d, cc = %g\n", it, trace[it]);
    if (iters == ITER_MAX) printf("The maximum number of iterations has been reached (The result has not converged).\n");
  }
  printf("finished. Elapsed time : %5g (sec)\n", difftime(me, ms));
  for (int m = 0; m < mo->N; m++) download_state(&mo->md[m]);
  free(trace);
  free(h);
}

void mo_dat_write(char *fname, MOBJ *mo)
{
  FILE *fp = fopen(fname, "wb");
  if (!fp) { printf("mo_dat_write(), Failed to create the %s file.\n", fname); exit(1); }
  fwrite(&mo->N, sizeof(int), 1, fp);
  for (int m = 0; m < mo->N; m++) {
    DOMD *md = &mo->md[m];
    BOUD *bd = &md->bd;
    fwrite(mo->mod_fn[m], 1, FNLGTH, fp);
    fwrite(mo->rs[m], sizeof(double), 3, fp);
    fwrite(md->med_fn, 1, 128, fp); fwrite(md->msh_fn, 1, 128, fp);
    fwrite(md->rv, sizeof(double), 3, fp); fwrite(&md->th, sizeof(double), 1, fp); fwrite(md->tv, sizeof(double), 3, fp);
    fwrite(&md->MN, sizeof(int), 1, fp);
    fwrite(md->n, sizeof(double complex), md->MN + 1, fp); fwrite(md->kn, sizeof(double complex), md->MN + 1, fp);
    beams_write(fp, &md->mw);
    fwrite(&bd->Nn, sizeof(int), 1, fp); fwrite(&bd->Ne, sizeof(int), 1, fp);
    for (int i = 0; i <= bd->Nn; i++) fwrite(bd->rn[i], sizeof(double), 3, fp);
    for (int i = 0; i <= bd->Ne; i++) fwrite(bd->ed[i], sizeof(int), 4, fp);
    fwrite(&bd->NN, sizeof(int), 1, fp);
    for (int i = 0; i <= bd->Ne; i++) fwrite(bd->eni[i], sizeof(int), 4, fp);
    fwrite(bd->md, sizeof(int), bd->Ne + 1, fp); fwrite(bd->sd, sizeof(int), bd->Ne + 1, fp); fwrite(bd->gd, sizeof(int), bd->Ne + 1, fp);
    for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) fwrite(bd->ren[i][j], sizeof(double), 3, fp);
    for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) fwrite(bd->wen[i][j], sizeof(double), 3, fp);
    for (int d = 0; d <= md->MN; d++) {
      SUBD *s = &bd->sb[d];
      fwrite(&s->Ne, sizeof(int), 1, fp);
      fwrite(s->sid, sizeof(int), s->Ne + 1, fp);
      double complex ***tab[6] = {s->U, s->dU, s->E, s->dE, s->H, s->dH};
      for (int w = 0; w < 6; w++)
        for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) fwrite(tab[w][i][j], sizeof(double complex), w < 2 ? 4 : 3, fp);
    }
  }
  fclose(fp);
}

void mo_dat_read(char *fname, MOBJ *mo)
{
  FILE *fp = fopen(fname, "rb");
  if (!fp) { printf("mo_dat_read(), Failed to open the %s file.\n", fname); exit(1); }
  rd(&mo->N, sizeof(int), 1, fp, "N");
  mo_alloc(mo);
  for (int m = 0; m < mo->N; m++) {
    DOMD *md = &mo->md[m];
    rd(mo->mod_fn[m], 1, FNLGTH, fp, "mod_fn");
    rd(mo->rs[m], sizeof(double), 3, fp, "rs");
    memset(md, 0, sizeof *md);
    read_geometry(fp, md);
    elem_const(&md->bd);
    for (int d = 0; d <= md->MN; d++) {
      SUBD *s = &md->bd.sb[d];
      int t;
      rd(&t, sizeof(int), 1, fp, "sub-domain Ne");
      rd(s->sid, sizeof(int), s->Ne + 1, fp, "sid");
      double complex ***tab[6] = {s->U, s->dU, s->E, s->dE, s->H, s->dH};
      for (int w = 0; w < 6; w++)
        for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) rd(tab[w][i][j], sizeof(double complex), w < 2 ? 4 : 3, fp, "state");
    }
    init_names(&md->cm, md->MN);
    md->cm.nn = (size_t)md->bd.NN;
    md->cm.na = 8 * md->cm.nn;
    upload_state(md);
  }
  fclose(fp);
}

void mo_output_node_particles(char *fname, MOBJ *mo)
{
  char fo[256] = {0}, tf[200] = {0};
  size_t s1 = strlen(fname);
  if (s1 > 200) die("mo_output_node_particles(), file name is too long.");
  snprintf(fo, sizeof fo, "%s", fname);
  char *sd = strrchr(fo, '.');
  if (sd) { strncpy(tf, fname, s1 - strlen(sd)); snprintf(fo, sizeof fo, "%s.particles", tf); }
  FILE *fp = fopen(fo, "w");
  if (!fp) { printf("Failed to create the %s file.\n", fo); exit(1); }
  fprintf(fp, "# x y z object_id\n");
  for (int o = 0; o < mo->N; o++)
    for (int i = 1; i <= mo->md[o].bd.Ne; i++)
      for (int j = 0; j < 4; j++)
        fprintf(fp, "%15.14e %15.14e %15.14e %d\n", mo->md[o].bd.ren[i][j][0], mo->md[o].bd.ren[i][j][1], mo->md[o].bd.ren[i][j][2], o);
  fclose(fp);
}

/* ------------------------------------------------------------ analysis */
void mo_object_domain_id(int *oid, int *did, double *rt, MOBJ *mo)
{
  for (int o = 0; o < mo->N; o++) {
    bh_or_die(bh_object_domain(dev(&mo->md[o]), 1, rt, did), "domain_id");
    if (*did != 0) { *oid = o; return; }
  }
  *oid = -1;
}

static int field1(double complex *out, double *rt, int type, MOBJ *mo, int what, int mode)
{
  bh_object **h = handles(mo);
  int oid;
  bh_or_die(bh_mo_field(h, mo->N, what, mode, type, 1, rt, (double *)out, &oid), "bh_mo_field");
  free(h);
  return oid;
}

int mo_mEMP_s(double complex *U, double *rt, int type, MOBJ *mo) { return field1(U, rt, type, mo, 1, 0); }
int mo_mEMP_t(double complex *U, double *rt, int type, MOBJ *mo) { return field1(U, rt, type, mo, 1, 1); }
int mo_mEMP_i(double complex *U, double *rt, int type, MOBJ *mo) { return field1(U, rt, type, mo, 1, 2); }

static int eh1(double complex *E, double complex *H, double *rt, int type, MOBJ *mo, int mode, int what)
{
  double complex o[6];
  int oid = field1(o, rt, type, mo, what, mode);
  for (int i = 0; i < 3; i++) { E[i] = o[i]; H[i] = o[3 + i]; }
  return oid;
}

int mo_EH_s(double complex *E, double complex *H, double *rt, int type, MOBJ *mo) { return eh1(E, H, rt, type, mo, 0, 0); }
int mo_EH_t(double complex *E, double complex *H, double *rt, int type, MOBJ *mo) { return eh1(E, H, rt, type, mo, 1, 0); }
int mo_EH_i(double complex *E, double complex *H, double *rt, int type, MOBJ *mo) { return eh1(E, H, rt, type, mo, 2, 0); }

static int eh_batch(double complex *E, double complex *H, const double *rt, int np, int type, MOBJ *mo, int *oid, int mode)
{
  bh_object **h = handles(mo);
  double complex *o = xalloc((size_t)np * 6, sizeof(double complex), "fields");
  int *id = oid ? oid : xalloc(np, sizeof(int), "oid");
  int rc = bh_mo_field(h, mo->N, 0, mode, type, np, rt, (double *)o, id);
  if (rc == 0)
    for (int p = 0; p < np; p++) for (int i = 0; i < 3; i++) { E[3 * p + i] = o[6 * p + i]; H[3 * p + i] = o[6 * p + 3 + i]; }
  free(o); free(h);
  if (!oid) free(id);
  return rc;
}

int mo_EH_t_batch(double complex *E, double complex *H, const double *rt, int np, int type, MOBJ *mo, int *oid) { return eh_batch(E, H, rt, np, type, mo, oid, 1); }
int mo_EH_s_batch(double complex *E, double complex *H, const double *rt, int np, int type, MOBJ *mo, int *oid) { return eh_batch(E, H, rt, np, type, mo, oid, 0); }

/* E, H from the potentials' gradients (the far-field form of the reference) */
int mo_EH_mEMP_s(double complex *E, double complex *H, double *rt, int type, MOBJ *mo) { return eh1(E, H, rt, type, mo, 0, 2); }
int mo_EH_mEMP_t(double complex *E, double complex *H, double *rt, int type, MOBJ *mo) { return eh1(E, H, rt, type, mo, 1, 2); }
int mo_EH_mEMP_i(double complex *E, double complex *H, double *rt, int type, MOBJ *mo) { return eh1(E, H, rt, type, mo, 2, 2); }

static void eh_bd(double complex *E, double complex *H, int oid, int did, int t, double zt, double et, int type, MOBJ *mo, int mode)
{
  bh_object **h = handles(mo);
  double ze[2] = {zt, et};
  double complex o[6];
  (void)type;
  bh_or_die(bh_mo_EH_bd(h, mo->N, oid, did, 1, &t, ze, PREC_DEF, mode, (double *)o), "bh_mo_EH_bd");
  for (int i = 0; i < 3; i++) { E[i] = o[i]; H[i] = o[3 + i]; }
  free(h);
}

void mo_EH_s_bd(double complex *E, double complex *H, int oid, int did, int t, double zeta_t, double eta_t, int type, MOBJ *mo) { eh_bd(E, H, oid, did, t, zeta_t, eta_t, type, mo, 0); }
void mo_EH_t_bd(double complex *E, double complex *H, int oid, int did, int t, double zeta_t, double eta_t, int type, MOBJ *mo) { eh_bd(E, H, oid, did, t, zeta_t, eta_t, type, mo, 1); }
void mo_EH_i_bd(double complex *E, double complex *H, int oid, int did, int t, double zeta_t, double eta_t, int type, MOBJ *mo) { eh_bd(E, H, oid, did, t, zeta_t, eta_t, type, mo, 2); }
void mo_EH_a_bd(double complex *E, double complex *H, int oid, int did, int t, double zeta_t, double eta_t, int type, MOBJ *mo) { eh_bd(E, H, oid, did, t, zeta_t, eta_t, type, mo, 3); }

int mo_force_FN(double *F, double *N, double *rc, int oid, int type, MOBJ *mo)
{
  bh_object **h = handles(mo);
  int r = bh_mo_force(h, mo->N, oid, type, rc, F, N);
  free(h);
  if (r != 0 && r != -1) bh_or_die(r, "bh_mo_force");
  return r;
}

int mo_absorb_P(double *P, int oid, int type, MOBJ *mo)
{
  bh_object **h = handles(mo);
  int r = bh_mo_absorb(h, mo->N, oid, type, P);
  free(h);
  if (r != 0 && r != -1 && r != -2) bh_or_die(r, "bh_mo_absorb");
  return r;
}

/* ------------------------------------------------------------ element layer
 * The public element functions of d3b1_elem.h:101-124.  The geometry
 * functions read the element constants the host keeps (d3b1_elem.c:10-146);
 * the coefficient dispatchers (d3b1_elem.c:148-539) run on the device
 * through bh_elem_coef, one request per call (bh_elem_coef itself takes
 * any number).  Rule type 4 (the double-exponential rules) is not
 * provided. */
int check_element_type(int seid, BOUD *bd) { return bd->ed[abs(seid)][3] == 0 ? ELT3 : ELT4; }

void r_bd(double *r, int seid, double zeta, double eta, BOUD *bd)
{
  const int s = abs(seid);
  const double q = check_element_type(seid, bd) == ELT4 ? zeta * eta : 0.0;
  for (int i = 0; i < 3; i++) r[i] = bd->cr[s][i][0] + bd->cr[s][i][1] * zeta + bd->cr[s][i][2] * eta + bd->cr[s][i][3] * q;
}

void r_bd_node(double *r, int seid, int nid, BOUD *bd)
{
  for (int i = 0; i < 3; i++) r[i] = bd->ren[abs(seid)][nid][i];
}

void n_bd_node(double *n, int seid, int nid, BOUD *bd)
{
  for (int i = 0; i < 3; i++) n[i] = bd->wen[abs(seid)][nid][i];
  vunit3(n);
}

void Tz_bd(double *vTz, int seid, double zeta, double eta, BOUD *bd)
{
  const int s = abs(seid), q = check_element_type(seid, bd) == ELT4;
  (void)zeta;
  for (int i = 0; i < 3; i++) vTz[i] = bd->cr[s][i][1] + (q ? bd->cr[s][i][3] * eta : 0.0);
}

void Te_bd(double *vTe, int seid, double zeta, double eta, BOUD *bd)
{
  const int s = abs(seid), q = check_element_type(seid, bd) == ELT4;
  (void)eta;
  for (int i = 0; i < 3; i++) vTe[i] = bd->cr[s][i][2] + (q ? bd->cr[s][i][3] * zeta : 0.0);
}

void r_tz_bd(double *r, double *vtz, int seid, double zeta, double eta, BOUD *bd)
{
  r_bd(r, seid, zeta, eta, bd);
  Tz_bd(vtz, seid, zeta, eta, bd);
  vunit3(vtz);
}

void r_te_bd(double *r, double *vte, int seid, double zeta, double eta, BOUD *bd)
{
  r_bd(r, seid, zeta, eta, bd);
  Te_bd(vte, seid, zeta, eta, bd);
  vunit3(vte);
}

void tz_te_bd_node(double *vtz, double *vte, int seid, int nid, BOUD *bd)
{
  const int q = check_element_type(seid, bd) == ELT4;
  Tz_bd(vtz, seid, q ? bd->zt_44[nid] : 0.0, q ? bd->et_44[nid] : 0.0, bd);
  Te_bd(vte, seid, q ? bd->zt_44[nid] : 0.0, q ? bd->et_44[nid] : 0.0, bd);
  vunit3(vtz);
  vunit3(vte);
}

/* one request of bh_elem_coef on the object whose boundary data bd is
 * (BOUD sits inside its DOMD); o27: CC, dC/dv0, dC/dv1 */
static void elem_request(BOUD *bd, const char *fn, int nd, int type, double complex k, int s, int t, double zt,
                         double et, const double *rt, const double *v0, const double *v1, double complex *o27)
{
  if (type < 0 || type > 3) {
    printf("b_elem.c, %s(), type number error! type=%d is not defined.Exit...\n", fn, type);
    exit(1);
  }
  DOMD *md = (DOMD *)((char *)bd - offsetof(DOMD, bd));
  const double zero[3] = {0.0, 0.0, 0.0}, ze[2] = {zt, et};
  bh_or_die(bh_elem_coef(dev(md), 1, nd, type, creal(k), cimag(k), &s, &t, ze, rt ? rt : zero, v0 ? v0 : zero,
                         v1 ? v1 : zero, (double *)o27),
            "bh_elem_coef");
}

/* local coordinates of node nid of the element that source element s's kind
 * uses for its self terms (zt_44 / zt_34, d3b1_elem.c:239,262) */
static void node_ze(BOUD *bd, int s, int nid, double *zt, double *et)
{
  const int q = check_element_type(s, bd) == ELT4;
  *zt = q ? bd->zt_44[nid] : bd->zt_34[nid];
  *et = q ? bd->et_44[nid] : bd->et_34[nid];
}

void coef_rt(double complex *CC, double *rt, int s, double complex k, int type, BOUD *bd)
{
  double complex o[27];
  elem_request(bd, "coef_rt", 0, type, k, s, 0, 0.0, 0.0, rt, NULL, NULL, o);
  memcpy(CC, o, 9 * sizeof(double complex));
}

void coef_bd(double complex *CC, double *rt, int t, double zeta_t, double eta_t, int s, double complex k, int type,
             BOUD *bd)
{
  double complex o[27];
  elem_request(bd, "coef_bd", 0, type, k, s, t, zeta_t, eta_t, rt, NULL, NULL, o);
  memcpy(CC, o, 9 * sizeof(double complex));
}

void coef_bd_node(double complex *CC, int t, int nid, int s, double complex k, int type, BOUD *bd)
{
  double complex o[27];
  double zt, et;
  node_ze(bd, s, nid, &zt, &et);
  elem_request(bd, "coef_bd_node", 0, type, k, s, t, zt, et, bd->ren[abs(t)][nid], NULL, NULL, o);
  memcpy(CC, o, 9 * sizeof(double complex));
}

void dcoef_rt(double complex *CC, double complex *dCC, double *rt, double *vt, int s, double complex k, int type,
              BOUD *bd)
{
  double complex o[27];
  elem_request(bd, "dcoef_rt", 1, type, k, s, 0, 0.0, 0.0, rt, vt, NULL, o);
  memcpy(CC, o, 9 * sizeof(double complex));
  memcpy(dCC, o + 9, 9 * sizeof(double complex));
}

void dcoef_bd_tz(double complex *CC, double complex *dCC, double *rt, double *vtz, int t, double zeta_t, double eta_t,
                 int s, double complex k, int type, BOUD *bd)
{
  double complex o[27];
  elem_request(bd, "dcoef_bd_tz", 1, type, k, s, t, zeta_t, eta_t, rt, vtz, NULL, o);
  memcpy(CC, o, 9 * sizeof(double complex));
  memcpy(dCC, o + 9, 9 * sizeof(double complex));
}

void dcoef_bd_te(double complex *CC, double complex *dCC, double *rt, double *vte, int t, double zeta_t, double eta_t,
                 int s, double complex k, int type, BOUD *bd)
{
  double complex o[27];
  elem_request(bd, "dcoef_bd_te", 1, type, k, s, t, zeta_t, eta_t, rt, vte, NULL, o);
  memcpy(CC, o, 9 * sizeof(double complex));
  /* the self element's derivatives come along both tangents: t_eta second */
  memcpy(dCC, o + (t == s ? 18 : 9), 9 * sizeof(double complex));
}

void dcoef_bd_t2(double complex *CC, double complex *dCdtz, double complex *dCdte, double *rt, double *vtz,
                 double *vte, int t, double zeta_t, double eta_t, int s, double complex k, int type, BOUD *bd)
{
  double complex o[27];
  elem_request(bd, "dcoef_bd_t2", 2, type, k, s, t, zeta_t, eta_t, rt, vtz, vte, o);
  memcpy(CC, o, 9 * sizeof(double complex));
  memcpy(dCdtz, o + 9, 9 * sizeof(double complex));
  memcpy(dCdte, o + 18, 9 * sizeof(double complex));
}

void dcoef_bd_node_t2(double complex *CC, double complex *dCdtz, double complex *dCdte, int t, int nid, double *vtz,
                      double *vte, int s, double complex k, int type, BOUD *bd)
{
  double complex o[27];
  double zt, et;
  node_ze(bd, s, nid, &zt, &et);
  elem_request(bd, "dcoef_bd_t2", 2, type, k, s, t, zt, et, bd->ren[abs(t)][nid], vtz, vte, o);
  memcpy(CC, o, 9 * sizeof(double complex));
  memcpy(dCdtz, o + 9, 9 * sizeof(double complex));
  memcpy(dCdte, o + 18, 9 * sizeof(double complex));
}
