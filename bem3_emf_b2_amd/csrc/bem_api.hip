// bem_api.hip -- C-ABI of the MI355X bem3_emf_b2 hot path (include/bem3_hip.h).
//
// Objects live in HBM for their whole life: element records, domain lists,
// the compact coefficient matrices of every domain, the factorised system,
// incident and boundary values.  Host code only moves inputs in and results
// out.  Every entry point returns 0 on success and a negative code on error
// (the message is kept for bh_last_error()).
#include "bem_obj.h"

#include <climits>

using namespace bem;

extern "C" {
void bem_launch_asm_p1(int, const void *, const int *, int, int, double, const void *, int, const void *, int,
                       hipStream_t);
void bem_launch_asm_p2(const void *, const int *, int, int, const void *, int, const void *, int, const void *, int,
                       const void *, int, hipStream_t);
void bem_launch_asm_p34(const void *, const int *, int, int, const double *, const double *, const double *,
                        const double *, hipStream_t);
int bem_rule32(const void *, int, void **);
int bem_gemm_time(int, int, int, int, int, int, double *, hipStream_t);
void bem_zgemm(int, int, int, double, double, const void *, size_t, const void *, size_t, double, double, void *, size_t,
               hipStream_t);
int bem_zgetrf(int, void *, size_t, int *, void *, int *, int, hipStream_t);
void bem_zgetrs(int, int, const void *, size_t, const int *, void *, size_t, int, hipStream_t);
int bem_la_init(void);
int bem_trsv_error(hipStream_t);
void bem_la_release(void);
int bem_la_noalloc_init(int);
int bem_zgesv_noalloc(int, int, int, void *, void *, int *, int *, hipStream_t);
int bem_noalloc_nmax(void);
void bem_prof_release(void);
void bem_mo_release(void);
void bem_cgemm(int, int, int, double, double, const void *, size_t, const void *, size_t, double, double, void *, size_t,
               hipStream_t);
int bem_cgetrf(int, void *, size_t, int *, int *, hipStream_t);
void bem_cgetrs(int, int, const void *, size_t, const int *, void *, size_t, hipStream_t);
}

static_assert(sizeof(cd) == 16, "complex layout");

// ------------------------------------------------------------------ errors, context
static thread_local std::string g_err;
namespace bem {
int fail(int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
std::recursive_mutex g_mu;
Ctx g;
thread_local hipStream_t t_cur_st = nullptr;
// a one-launch triangular solve whose wait timed out since the last check
// (reads and clears the flag)
int trsv_failed(hipStream_t st)
{
  int e = bem_trsv_error(st);
  return e ? fail(BH_EHIP, "triangular solve: a block wait timed out") : 0;
}
}  // namespace bem
static int lu_failed(void) { return fail(BH_EHIP, "LU factorisation: scratch allocation failed or a panel barrier timed out"); }

extern "C" const char *bh_last_error(void) { return g_err.c_str(); }

// Gauss-Legendre nodes on (a, b) by Newton iteration on P_n (same fixed
// point as the reference's gauleg.c:3-36: start cos(pi(i-1/4)/(n+1/2)),
// stop |dz|<=3e-16)
static void gl_nodes_ab(double a, double b, int n, double *x, double *w)
{
  const double mid = 0.5 * (b + a), hl = 0.5 * (b - a);
  int half = (n + 1) / 2;
  for (int i = 1; i <= half; i++) {
    double z = cos(M_PI * (i - 0.25) / (n + 0.5)), pd;
    for (;;) {
      double p0 = 1.0, p1 = 0.0;
      for (int j = 1; j <= n; j++) {
        double pm = p1;
        p1 = p0;
        p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * pm) / j;
      }
      pd = n * (z * p0 - p1) / (z * z - 1.0);
      double zo = z;
      z = zo - p0 / pd;
      if (fabs(z - zo) <= 3.0e-16) break;
    }
    x[i - 1] = mid - hl * z; x[n - i] = mid + hl * z;
    w[i - 1] = 2.0 * hl / ((1.0 - z * z) * pd * pd); w[n - i] = w[i - 1];
  }
  if (a == -b && (n % 2)) x[n / 2] = 0.0;
}
static void gl_nodes(int n, double *x, double *w) { gl_nodes_ab(-1.0, 1.0, n, x, w); }

// 3-part map of the equilateral triangle onto three [-1,1]^2 squares
static std::vector<QPt> tri_rule(int n, const double *x, const double *w)
{
  std::vector<QPt> r;
  const double S3 = BEM_SQ3;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double a = x[i], b = x[j], wt = (1.0 + 0.25 * a + 0.25 * b) * w[i] * w[j];
      QPt p;
      p = {0.125 * (3.0 + 2.0 * a + 2.0 * b + a * b), 0.125 * S3 * (-a + b), wt}; r.push_back(p);
      p = {0.0625 * (-3.0 + a - 5.0 * b - a * b), 0.0625 * S3 * (3.0 + 3.0 * a + b + a * b), wt}; r.push_back(p);
      p = {0.0625 * (-3.0 - 5.0 * a + b - a * b), 0.0625 * S3 * (-3.0 - a - 3.0 * b - a * b), wt}; r.push_back(p);
    }
  return r;
}


static int ensure_init()
{
  if (g.dev >= 0) return 0;
  return bh_init(0);
}

extern "C" int bh_init(int device)
{
  API_LOCK;
  if (g.dev >= 0) return 0;
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (n <= device) return fail(BH_ENODEV, "no HIP device %d (found %d)", device, n);
  {
    int rc = tune_init();
    if (rc) return rc;
  }
  HIPCHK(hipSetDevice(device));
  // the library's stream lives as long as the process (bh_finalize keeps it)
  static hipStream_t s_st = nullptr;
  if (!s_st) HIPCHK(hipStreamCreate(&s_st));
  g.st = s_st;
  HIPCHK(hipEventCreate(&g.ev[0]));
  HIPCHK(hipEventCreate(&g.ev[1]));
  if (bem_la_init()) fprintf(stderr, "bem3_hip: could not raise the dynamic LDS limit (%s)\n", hipGetErrorString(hipGetLastError()));
  // the allocation-free solve's work area: its device address, once per device
  if (bem_la_noalloc_init(device)) return fail(BH_EHIP, "device address of the allocation-free work area");
  gl_nodes(12, g.xl, g.wl);
  gl_nodes(64, g.xh, g.wh);
  std::vector<QPt> a = tri_rule(12, g.xl, g.wl), b = tri_rule(64, g.xh, g.wh);
  g.n_gl = (int)a.size(); g.n_gh = (int)b.size();
  if (upload(&g.d_gl, a.data(), a.size()) || upload(&g.d_gh, b.data(), b.size()) || upload(&g.d_xl, g.xl, 12) ||
      upload(&g.d_wl, g.wl, 12) || upload(&g.d_xh, g.xh, 64) || upload(&g.d_wh, g.wh, 64))
    return BH_EHIP;
  HIPCHK(hipStreamSynchronize(cur_st()));
  {
    // far-pair table of the mixed form: 12x12, 6x6, 5x5, 4x4
    double x6[6], w6[6], x5[5], w5[5], x4[4], w4[4];
    gl_nodes(6, x6, w6);
    gl_nodes(5, x5, w5);
    gl_nodes(4, x4, w4);
    std::vector<QPt> r32 = a, r6 = tri_rule(6, x6, w6), r5 = tri_rule(5, x5, w5), r4 = tri_rule(4, x4, w4);
    r32.insert(r32.end(), r6.begin(), r6.end());
    r32.insert(r32.end(), r5.begin(), r5.end());
    r32.insert(r32.end(), r4.begin(), r4.end());
    if ((int)a.size() != BEM_FAR_OFF6 || (int)r32.size() != BEM_FAR_NPT || bem_rule32(r32.data(), BEM_FAR_NPT, &g.d_gl32))
      return fail(BH_EHIP, "FP32 rule table");
    if (upload(&g.d_glr, r32.data(), r32.size())) return BH_EHIP;
    HIPCHK(hipStreamSynchronize(cur_st()));
  }
  g.dev = device;
  return 0;
}

extern "C" int bh_gemm_time(int prec, int kernel, int m, int n, int k, int reps, double *ms)
{
  API_LOCK;
  if (ensure_init()) return BH_ENODEV;
  if (prec < 0 || prec > 1 || kernel < 0 || kernel > 2 || m < 1 || n < 1 || k < 1 || reps < 1 || !ms)
    return fail(BH_EARG, "bh_gemm_time: prec %d kernel %d shape %d x %d x %d reps %d", prec, kernel, m, n, k, reps);
  const int rc = bem_gemm_time(prec, kernel, m, n, k, reps, ms, cur_st());
  if (rc == -2) return fail(BH_EUNSUP, "bh_gemm_time: kernel %d is not available for precision %d", kernel, prec);
  if (rc) return fail(BH_EHIP, "bh_gemm_time: %s", hipGetErrorString(hipGetLastError()));
  return 0;
}

extern "C" int bh_sync(void)
{
  API_LOCK;
  if (g.dev < 0) return 0;
  HIPCHK(hipStreamSynchronize(cur_st()));
  return 0;
}

extern "C" void *bh_stream(void) { return (void *)g.st; }

// ------------------------------------------------------------------ objects

// element constants and node data (reference setup2.c:666-749, 838-869):
// triangles r = c0 + c1 z + c2 e on the equilateral reference triangle,
// node slots at the 4-point rule's abscissae (centroid in slot 3); quads
// r = c0 + c1 z + c2 e + c3 z e on [-1,1]^2 with W(z,e) = W0 + W1 z + W2 e,
// node slots at the 2x2 Gauss points
static const double Q44 = 0.5773502691896257645091;   // 1/sqrt(3)
static const double ZT44[4] = {-Q44, Q44, Q44, -Q44}, ET44[4] = {-Q44, -Q44, Q44, Q44};

static void elem_const(bh_object *o)
{
  const double P340 = 0.2, P341 = 0.3464101615137754587055;
  const double zt[4] = {-P340, 2.0 * P340, -P340, 0.0}, et[4] = {-P341, 0.0, P341, 0.0};
  o->cr.assign((size_t)(o->Ne + 1) * 12, 0.0);
  o->cw.assign((size_t)(o->Ne + 1) * 9, 0.0);
  o->ren.assign((size_t)(o->Ne + 1) * 12, 0.0);
  o->wen.assign((size_t)(o->Ne + 1) * 12, 0.0);
  for (int i = 1; i <= o->Ne; i++) {
    double p[4][3];
    const bool quad = o->ed[4 * i + 3] != 0;
    for (int j = 0; j < (quad ? 4 : 3); j++)
      for (int d = 0; d < 3; d++) p[j][d] = o->rn[3 * (size_t)o->ed[4 * i + j] + d];
    double *cr = &o->cr[(size_t)i * 12], *cw = &o->cw[(size_t)i * 9];
    if (quad) {
      for (int d = 0; d < 3; d++) {
        cr[d * 4 + 0] = 0.25 * (p[0][d] + p[1][d] + p[2][d] + p[3][d]);
        cr[d * 4 + 1] = 0.25 * (p[1][d] + p[2][d] - p[0][d] - p[3][d]);
        cr[d * 4 + 2] = 0.25 * (p[2][d] + p[3][d] - p[0][d] - p[1][d]);
        cr[d * 4 + 3] = 0.25 * (p[0][d] + p[2][d] - p[1][d] - p[3][d]);
      }
      // dr/dz x dr/de = (c1 + c3 e) x (c2 + c3 z) = c1 x c2 + z (c1 x c3) + e (c3 x c2)
      for (int d = 0; d < 3; d++) {
        int a = (d + 1) % 3, b = (d + 2) % 3;
        cw[d * 3 + 0] = cr[a * 4 + 1] * cr[b * 4 + 2] - cr[b * 4 + 1] * cr[a * 4 + 2];
        cw[d * 3 + 1] = cr[a * 4 + 1] * cr[b * 4 + 3] - cr[b * 4 + 1] * cr[a * 4 + 3];
        cw[d * 3 + 2] = cr[a * 4 + 3] * cr[b * 4 + 2] - cr[b * 4 + 3] * cr[a * 4 + 2];
      }
      for (int j = 0; j < 4; j++)
        for (int d = 0; d < 3; d++) {
          const double z = ZT44[j], e = ET44[j];
          o->ren[(size_t)i * 12 + j * 3 + d] = cr[d * 4] + cr[d * 4 + 1] * z + cr[d * 4 + 2] * e + cr[d * 4 + 3] * z * e;
          o->wen[(size_t)i * 12 + j * 3 + d] = cw[d * 3] + cw[d * 3 + 1] * z + cw[d * 3 + 2] * e;
        }
      continue;
    }
    for (int d = 0; d < 3; d++) {
      cr[d * 4 + 0] = (p[0][d] + p[1][d] + p[2][d]) / 3.0;
      cr[d * 4 + 1] = (-p[0][d] + 2.0 * p[1][d] - p[2][d]) / 3.0;
      cr[d * 4 + 2] = (p[2][d] - p[0][d]) / sqrt(3.0);
      cr[d * 4 + 3] = 0.0;
    }
    double u[3], v[3];
    for (int d = 0; d < 3; d++) { u[d] = p[1][d] - p[0][d]; v[d] = p[2][d] - p[0][d]; }
    for (int d = 0; d < 3; d++) {
      int a = (d + 1) % 3, b = (d + 2) % 3;
      cw[d * 3 + 0] = u[a] * v[b] - v[a] * u[b];
    }
    for (int j = 0; j < 4; j++)
      for (int d = 0; d < 3; d++) {
        o->ren[(size_t)i * 12 + j * 3 + d] = cr[d * 4] + cr[d * 4 + 1] * zt[j] + cr[d * 4 + 2] * et[j];
        o->wen[(size_t)i * 12 + j * 3 + d] = cw[d * 3];
      }
  }
}

// unit tangents along z and e at node slot j of element a
static void node_tangents(const bh_object *o, int a, int j, double *vtz, double *vte)
{
  const double *cr = &o->cr[(size_t)a * 12];
  const bool quad = o->ed[4 * a + 3] != 0;
  for (int d = 0; d < 3; d++) {
    vtz[d] = cr[d * 4 + 1] + (quad ? cr[d * 4 + 3] * ET44[j] : 0.0);
    vte[d] = cr[d * 4 + 2] + (quad ? cr[d * 4 + 3] * ZT44[j] : 0.0);
  }
}

static void unit3(double *v)
{
  double m = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (m != 0.0) { v[0] /= m; v[1] /= m; v[2] /= m; }
}

static int build_domains(bh_object *o)
{
  o->dom.assign(o->MN + 1, Dom());
  for (int d = 0; d <= o->MN; d++) {
    Dom &D = o->dom[d];
    D.sid.assign(1, 0);
    for (int i = 1; i <= o->Ne; i++) {
      if (o->md[i] == d) D.sid.push_back(i);
      else if (o->sd[i] == d) D.sid.push_back(-i);
    }
    D.dNe = (int)D.sid.size() - 1;
    D.colOff.assign(D.dNe + 1, 0);
    int off = 0;
    for (int s = 1; s <= D.dNe; s++) {
      int a = std::abs(D.sid[s]);
      D.colOff[s] = off;
      off += o->ed[4 * a + 3] == 0 ? 3 : 4;
    }
    D.nd = off;
    D.rows.clear();
    for (int t = 1; t <= D.dNe; t++) {
      int a = std::abs(D.sid[t]);
      const int ns = o->ed[4 * a + 3] == 0 ? 3 : 4;
      for (int tn = 0; tn < ns; tn++) {
        double vtz[3], vte[3];
        node_tangents(o, a, tn, vtz, vte);
        unit3(vtz);
        unit3(vte);
        RowPt r;
        memset(&r, 0, sizeof r);
        for (int q = 0; q < 3; q++) { r.rt[q] = o->ren[(size_t)a * 12 + tn * 3 + q]; r.v0[q] = vtz[q]; r.v1[q] = vte[q]; }
        r.t = t;
        r.slot = tn;
        D.rows.push_back(r);
      }
    }
    if (upload(&D.d_sid, D.sid.data(), D.sid.size()) || upload(&D.d_colOff, D.colOff.data(), D.colOff.size()) ||
        upload(&D.d_rows, D.rows.data(), D.rows.size()))
      return BH_EHIP;
  }
  return 0;
}

static int upload_geo(bh_object *o)
{
  std::vector<TriGeo> geo(o->Ne + 1);
  memset(geo.data(), 0, geo.size() * sizeof(TriGeo));
  for (int i = 1; i <= o->Ne; i++) {
    const double *cr = &o->cr[(size_t)i * 12], *cw = &o->cw[(size_t)i * 9];
    for (int d = 0; d < 3; d++) {
      geo[i].c0[d] = cr[d * 4]; geo[i].c1[d] = cr[d * 4 + 1]; geo[i].c2[d] = cr[d * 4 + 2]; geo[i].W[d] = cw[d * 3];
    }
    geo[i].aW = sqrt(cw[0] * cw[0] + cw[3] * cw[3] + cw[6] * cw[6]);
    geo[i].quad = o->ed[4 * i + 3] != 0 ? 1.0 : 0.0;
    if (o->ed[4 * i + 3] != 0) o->nquad++;
  }
  std::vector<QuadExt> gq(o->Ne + 1);
  memset(gq.data(), 0, gq.size() * sizeof(QuadExt));
  for (int i = 1; i <= o->Ne; i++) {
    if (o->ed[4 * i + 3] == 0) continue;
    const double *cr = &o->cr[(size_t)i * 12], *cw = &o->cw[(size_t)i * 9];
    for (int d = 0; d < 3; d++) { gq[i].c3[d] = cr[d * 4 + 3]; gq[i].W1[d] = cw[d * 3 + 1]; gq[i].W2[d] = cw[d * 3 + 2]; }
  }
  return upload(&o->d_geo, geo.data(), geo.size()) || upload(&o->d_gq, gq.data(), gq.size());
}

extern "C" bh_object *bh_object_create(int Nn, const double *rn, int Ne, const int *ed, const int *md, const int *sd,
                                       int MN, const double *n_med, double lambda0)
{
  API_LOCK;
  if (ensure_init()) return nullptr;
  if (Ne <= 0 || Nn <= 0 || MN < 1) { fail(BH_EARG, "empty object"); return nullptr; }
  bh_object *o = new bh_object();
  o->Nn = Nn; o->Ne = Ne; o->MN = MN;
  o->rn.assign(rn, rn + 3 * (size_t)(Nn + 1));
  o->ed.assign(ed, ed + 4 * (size_t)(Ne + 1));
  o->md.assign(md, md + Ne + 1);
  o->sd.assign(sd, sd + Ne + 1);
  for (int i = 1; i <= Ne; i++) {
    if (o->ed[4 * i + 3] != 0) o->all_tri = false;
    if (o->md[i] < 0 || o->md[i] > MN || o->sd[i] < 0 || o->sd[i] > MN) {
      fail(BH_EARG, "element %d: domain id out of range", i);
      delete o;
      return nullptr;
    }
  }

  // node numbers sequential per element: 3 for a triangle, 4 for a
  // quadrangle (setup2.c:372-383)
  o->eni.assign(4 * (size_t)(Ne + 1), -1);
  int nc = 0;
  for (int i = 1; i <= Ne; i++)
    for (int j = 0; j < (o->ed[4 * i + 3] != 0 ? 4 : 3); j++) o->eni[4 * i + j] = nc++;
  o->NN = nc;
  o->n.assign(MN + 1, 0.0);
  o->kn.assign(MN + 1, 0.0);
  for (int i = 0; i <= MN; i++) {
    o->n[i] = zd(n_med[2 * i], n_med[2 * i + 1]);
    o->kn[i] = 2.0 * M_PI / lambda0 * o->n[i];
  }
  elem_const(o);
  if (build_domains(o) || upload_geo(o)) { delete o; return nullptr; }
  if (hipStreamSynchronize(cur_st()) != hipSuccess) { fail(BH_EHIP, "upload failed"); delete o; return nullptr; }
  return o;
}

extern "C" void bh_object_destroy(bh_object *o)
{
  API_LOCK;
  if (!o) return;
  (void)hipStreamSynchronize(cur_st());   // no kernel still uses the object's memory
  for (Dom &D : o->dom) {
    dfree(D.d_sid); dfree(D.d_colOff); dfree(D.d_rows);
    dfree(D.d_G); dfree(D.d_H); dfree(D.d_dG); dfree(D.d_dH); dfree(D.d_dF);
    dfree(D.d_col); dfree(D.d_colS); dfree(D.d_rowEH);
    dfree(D.d_U); dfree(D.d_dU); dfree(D.d_E); dfree(D.d_dE); dfree(D.d_Hf); dfree(D.d_dHf);
    dfree(D.d_Gf); dfree(D.d_Gpiv); dfree(D.c_Gf);
  }
  dfree(o->d_geo); dfree(o->d_gq);
  dfree(o->d_Ui); dfree(o->d_dUi); dfree(o->d_Ui0); dfree(o->d_dUi0); dfree(o->d_Uip);
  dfree(o->d_slotPt); dfree(o->d_slotN); dfree(o->d_slotIdx); dfree(o->d_fv); dfree(o->d_fsd); dfree(o->d_fq); dfree(o->d_seg);
  dfree(o->d_beams); dfree(o->d_fb); dfree(o->d_fbt);
  dfree(o->d_A); dfree(o->d_work); dfree(o->d_piv); dfree(o->d_info);
  dfree(o->d_G0f); dfree(o->d_Z); dfree(o->d_Mf); dfree(o->d_Sf);
  dfree(o->d_piv0); dfree(o->d_pivM); dfree(o->d_pivS);
  dfree(o->d_X); dfree(o->d_B); dfree(o->d_tmp); dfree(o->d_w);
  dfree(o->c_G0f); dfree(o->c_Z); dfree(o->c_Mf); dfree(o->c_Sf); dfree(o->c_G1); dfree(o->c_w);
  dfree(o->d_R); dfree(o->d_nrm); dfree(o->d_ginfo);
  delete o;
}

extern "C" int bh_object_info(const bh_object *o, int *out)
{
  API_LOCK;
  out[0] = o->Nn; out[1] = o->Ne; out[2] = o->NN; out[3] = o->MN;
  for (int d = 0; d <= o->MN && d < 8; d++) { out[4 + d] = o->dom[d].dNe; out[12 + d] = o->dom[d].nd; }
  return 0;
}

extern "C" int bh_object_geometry(const bh_object *o, double *cr, double *cw, double *ren, double *wen, int *eni)
{
  API_LOCK;
  if (cr) memcpy(cr, o->cr.data(), o->cr.size() * sizeof(double));
  if (cw) memcpy(cw, o->cw.data(), o->cw.size() * sizeof(double));
  if (ren) memcpy(ren, o->ren.data(), o->ren.size() * sizeof(double));
  if (wen) memcpy(wen, o->wen.data(), o->wen.size() * sizeof(double));
  if (eni) memcpy(eni, o->eni.data(), o->eni.size() * sizeof(int));
  return 0;
}

// rule tables (QPtN on the triangle, QPtQ on the square), built at first use
namespace bem {
QPtN *g_rule[4] = {nullptr, nullptr, nullptr, nullptr};
int g_nrule[4] = {0, 0, 0, 0};
double g_rscale[4];
QPtQ *g_qrule[4] = {nullptr, nullptr, nullptr, nullptr};   // the same rules on the square (quads)
int g_nqrule[4] = {0, 0, 0, 0};
}  // namespace bem

// ------------------------------------------------------------------ assembly
template <class T> static int dalloc(T **p, size_t n)
{
  HIPCHK(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
  HIPCHK(hipMemsetAsync(*p, 0, std::max<size_t>(n, 1) * sizeof(T), cur_st()));
  return 0;
}

// one block of rows of one domain in a batched assembly: rows r0 ..
// r0+nrows of domain d's matrices into G, H, dG, dH (row stride = the
// domain's node count, the block's first row first) and dF (3 per row);
// *ngh receives the near pairs redone with the 64-point rule
struct AsmSpec {
  bh_object *o;
  int d, r0, nrows;
  cd *G, *H, *dG, *dH;
  double *dF;
  int *ngh;
};

// Passes 1-4 of the assembly for every spec of the batch at once: one launch
// per pass over all of them (the objects and domains of bh_mo_prepare, or
// the domains of one object), one read-back of the near-pair flags, one
// upload of the pair lists.  The specs of one object carve its grow-only
// work areas between them.  All objects of a batch take one precision form.
static int assemble_batch(const std::vector<AsmSpec> &sp)
{
  const int nj = (int)sp.size();
  if (nj == 0) return 0;
  const bool mixed = sp[0].o->prec != BH_PREC_FP64;
  for (const AsmSpec &x : sp)
    if ((x.o->prec != BH_PREC_FP64) != mixed) return fail(BH_EARG, "batched assembly: objects of both precision forms");
  // objects of the batch and each spec's offsets into its object's areas
  std::vector<bh_object *> obs;
  std::vector<int> oi(nj);
  for (int j = 0; j < nj; j++) {
    int k = (int)(std::find(obs.begin(), obs.end(), sp[j].o) - obs.begin());
    if (k == (int)obs.size()) obs.push_back(sp[j].o);
    oi[j] = k;
  }
  const int no = (int)obs.size();
  std::vector<int> nblk(nj), nwv(nj);
  std::vector<size_t> offM(nj), offS(nj), offT(nj), totM(no, 0), totS(no, 0), totT(no, 0);
  std::vector<int> f1(nj + 1, 0), f2(nj + 1, 0), f3(nj + 1, 0);
  size_t b1 = 0, r3 = 0;
  double work = 0.0;
  for (int j = 0; j < nj; j++) {
    const Dom &D = sp[j].o->dom[sp[j].d];
    const size_t nr = (size_t)sp[j].nrows;
    nblk[j] = (D.dNe + 255) / 256;
    nwv[j] = nblk[j] * 4;
    offM[j] = totM[oi[j]]; totM[oi[j]] += nr * nwv[j];   // masks; partial F sums 3 per mask
    offS[j] = totS[oi[j]]; totS[oi[j]] += nr * 3;        // self F terms
    offT[j] = totT[oi[j]]; totT[oi[j]] += nr + 1;        // row starts of the pair lists
    f1[j] = (int)b1; b1 += nr * nblk[j];
    f3[j] = (int)r3; r3 += nr;
    work += (double)nr * D.dNe;
  }
  if (b1 > (size_t)INT_MAX || r3 > (size_t)INT_MAX / 64) return fail(BH_EARG, "batched assembly: %zu workgroups", b1);
  f1[nj] = (int)b1; f3[nj] = (int)r3;
  std::vector<double *> part(no), self(no);
  std::vector<unsigned long long *> mask(no);
  std::vector<int *> start(no);
  for (int k = 0; k < no; k++) {
    part[k] = obs[k]->ws_part.get<double>(totM[k] * 3);
    mask[k] = obs[k]->ws_mask.get<unsigned long long>(totM[k]);
    self[k] = obs[k]->ws_self.get<double>(totS[k]);
    start[k] = obs[k]->ws_start.get<int>(totT[k]);
    if (!part[k] || !mask[k] || !self[k] || !start[k]) return fail(BH_EHIP, "assembly: work areas of %zu rows", totS[k] / 3);
  }
  std::vector<AsmJob> jobs(nj);
  for (int j = 0; j < nj; j++) {
    const AsmSpec &x = sp[j];
    const Dom &D = x.o->dom[x.d];
    AsmJob &J = jobs[j];
    J.geo = (const TriGeo *)x.o->d_geo; J.gq = (const QuadExt *)x.o->d_gq; J.sid = D.d_sid; J.rows = D.d_rows + x.r0;
    J.colOff = D.d_colOff;
    J.G = x.G; J.H = x.H; J.dG = x.dG; J.dH = x.dH; J.dF = x.dF;
    J.dfPart = part[oi[j]] + offM[j] * 3; J.ghMask = mask[oi[j]] + offM[j];
    J.pairs = nullptr; J.ghStart = start[oi[j]] + offT[j]; J.ghDF = nullptr;
    J.selfDF = self[oi[j]] + offS[j];
    J.kr = x.o->kn[x.d].real(); J.ki = x.o->kn[x.d].imag();
    J.ld = (size_t)D.nd; J.dNe = D.dNe; J.nrows = x.nrows; J.nblk = nblk[j]; J.r0 = x.r0; J.open = x.d == 0; J.pad = 0;
  }
  // the device job table: jobs, then the prefix tables of passes 1, 2, 3
  const size_t jb = ((size_t)nj * sizeof(AsmJob) + 255) / 256 * 256, tb = jb + 3 * (size_t)(nj + 1) * sizeof(int);
  char *tab = sp[0].o->ws_jobs.get<char>(tb);
  if (!tab) return fail(BH_EHIP, "assembly: job table");
  std::vector<char> h1(tb, 0);
  memcpy(h1.data(), jobs.data(), (size_t)nj * sizeof(AsmJob));
  int *hf = (int *)(h1.data() + jb);
  for (int j = 0; j <= nj; j++) { hf[j] = f1[j]; hf[nj + 1 + j] = 0; hf[2 * (nj + 1) + j] = f3[j]; }
  const int *d_f1 = (const int *)(tab + jb), *d_f2 = d_f1 + (nj + 1), *d_f3 = d_f1 + 2 * (nj + 1);
  HIPCHK(hipMemcpyAsync(tab, h1.data(), tb, hipMemcpyHostToDevice, cur_st()));
  if (mixed)
    bem_launch_asm_p1(1, tab, d_f1, nj, (int)b1, work, g.d_glr, 0, g_qrule[2], g_nqrule[2], cur_st());
  else
    bem_launch_asm_p1(0, tab, d_f1, nj, (int)b1, work, g.d_gl, g.n_gl, g_qrule[2], g_nqrule[2], cur_st());
  HIPCHK(hipGetLastError());
  // near pairs of every job in row-major order (deterministic)
  std::vector<std::vector<unsigned long long>> hm(nj);
  for (int j = 0; j < nj; j++) {
    hm[j].resize((size_t)sp[j].nrows * nwv[j]);
    HIPCHK(hipMemcpyAsync(hm[j].data(), jobs[j].ghMask, hm[j].size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          cur_st()));
  }
  HIPCHK(hipStreamSynchronize(cur_st()));
  std::vector<std::vector<GhPair>> pairs(nj);
  std::vector<std::vector<int>> st(nj);
  std::vector<size_t> offP(nj), totP(no, 0);
  size_t p2 = 0;
  for (int j = 0; j < nj; j++) {
    const size_t nr = (size_t)sp[j].nrows;
    st[j].assign(nr + 1, 0);
    for (size_t r = 0; r < nr; r++) {
      st[j][r] = (int)pairs[j].size();
      for (int w = 0; w < nwv[j]; w++) {
        unsigned long long m = hm[j][r * nwv[j] + w];
        while (m) {
          const int bt = __builtin_ctzll(m);
          m &= m - 1;
          pairs[j].push_back({(int)r, w * 64 + bt + 1});
        }
      }
    }
    st[j][nr] = (int)pairs[j].size();
    *sp[j].ngh = (int)pairs[j].size();
    offP[j] = totP[oi[j]]; totP[oi[j]] += pairs[j].size();
    f2[j] = (int)p2; p2 += pairs[j].size();
  }
  if (p2 > (size_t)INT_MAX / 64) return fail(BH_EARG, "batched assembly: %zu near pairs", p2);
  f2[nj] = (int)p2;
  std::vector<GhPair *> dpr(no);
  std::vector<double *> dgh(no);
  for (int k = 0; k < no; k++) {
    dpr[k] = obs[k]->ws_pairs.get<GhPair>(std::max<size_t>(totP[k], 1));
    dgh[k] = obs[k]->ws_ghdf.get<double>(std::max<size_t>(totP[k], 1) * 3);
    if (!dpr[k] || !dgh[k]) return fail(BH_EHIP, "assembly: work area of %zu near pairs", totP[k]);
  }
  for (int j = 0; j < nj; j++) {
    jobs[j].pairs = dpr[oi[j]] + offP[j];
    jobs[j].ghDF = dgh[oi[j]] + offP[j] * 3;
    if (!pairs[j].empty())
      HIPCHK(hipMemcpyAsync((void *)jobs[j].pairs, pairs[j].data(), pairs[j].size() * sizeof(GhPair),
                            hipMemcpyHostToDevice, cur_st()));
    HIPCHK(hipMemcpyAsync((void *)jobs[j].ghStart, st[j].data(), st[j].size() * sizeof(int), hipMemcpyHostToDevice,
                          cur_st()));
  }
  std::vector<char> h2(tb, 0);
  memcpy(h2.data(), jobs.data(), (size_t)nj * sizeof(AsmJob));
  hf = (int *)(h2.data() + jb);
  for (int j = 0; j <= nj; j++) { hf[j] = f1[j]; hf[nj + 1 + j] = f2[j]; hf[2 * (nj + 1) + j] = f3[j]; }
  HIPCHK(hipMemcpyAsync(tab, h2.data(), tb, hipMemcpyHostToDevice, cur_st()));
  bem_launch_asm_p2(tab, d_f2, nj, (int)p2, g.d_gh, g.n_gh, mixed ? (const void *)g.d_gl : nullptr, g.n_gl, g_qrule[3],
                    g_nqrule[3], g_qrule[2], g_nqrule[2], cur_st());
  bem_launch_asm_p34(tab, d_f3, nj, (int)r3, g.d_xl, g.d_wl, g.d_xh, g.d_wh, cur_st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(cur_st()));   // the host tables above stay alive until here
  return 0;
}

// rows r0 .. r0+nrows of domain d (a batch of one)
static int assemble_rows(bh_object *o, int d, int r0, int nrows, cd *G, cd *H, cd *dG, cd *dH, double *dF, int *ngh)
{
  return assemble_batch({AsmSpec{o, d, r0, nrows, G, H, dG, dH, dF, ngh}});
}

// the FP64 matrices of domain d (allocated at the first assembly) as one
// spec of a batch
static int domain_spec(bh_object *o, int d, AsmSpec &x)
{
  Dom &D = o->dom[d];
  size_t nd = (size_t)D.nd;
  if (!D.d_G) {
    HIPCHK(hipMalloc((void **)&D.d_G, nd * nd * sizeof(cd)));
    HIPCHK(hipMalloc((void **)&D.d_H, nd * nd * sizeof(cd)));
    HIPCHK(hipMalloc((void **)&D.d_dG, 2 * nd * nd * sizeof(cd)));
    HIPCHK(hipMalloc((void **)&D.d_dH, 2 * nd * nd * sizeof(cd)));
    HIPCHK(hipMalloc((void **)&D.d_dF, 3 * nd * sizeof(double)));
  }
  x = AsmSpec{o, d, 0, (int)nd, D.d_G, D.d_H, D.d_dG, D.d_dH, D.d_dF, &D.n_gh};
  return 0;
}

// Domain d's FP64 matrices a block of rows at a time: fn(r0, nrows, G, H,
// dG, dH, dF) with pointers at the block's first row (row stride nd).  The
// FP64 forms hold them whole (one call); the low-memory form assembles each
// block again (rows per block within about 16 GB of FP64 rows).
template <class F> static int domain_blocks(bh_object *o, int d, F fn)
{
  Dom &D = o->dom[d];
  const size_t nd = (size_t)D.nd;
  if (o->prec != BH_PREC_MIXED_LM) return fn(0, (int)nd, D.d_G, D.d_H, D.d_dG, D.d_dH, D.d_dF);
  size_t rb = o->lm_rows > 0 ? (size_t)o->lm_rows : ((size_t)16 << 30) / (6 * nd * sizeof(cd));
  rb = std::max<size_t>(64, std::min(rb, nd));
  cd *G = o->lm_blk[0].get<cd>(rb * nd), *H = o->lm_blk[1].get<cd>(rb * nd);
  cd *dG = o->lm_blk[2].get<cd>(2 * rb * nd), *dH = o->lm_blk[3].get<cd>(2 * rb * nd);
  double *dF = o->lm_blk[4].get<double>(3 * rb);
  if (!G || !H || !dG || !dH || !dF) return fail(BH_EHIP, "low-memory form: row blocks of %zu rows", rb);
  int ngh = 0;
  for (size_t r0 = 0; r0 < nd; r0 += rb) {
    const int n = (int)std::min(rb, nd - r0);
    int k = 0, rc = assemble_rows(o, d, (int)r0, n, G, H, dG, dH, dF, &k);
    if (rc) return rc;
    ngh += k;
    if ((rc = fn((int)r0, n, G, H, dG, dH, dF))) return rc;
  }
  D.n_gh = ngh;
  return 0;
}

// the low-memory form's assembly: every block of G, H rounded into the
// complex-FP32 storage the factorisation works in (c_G0f: G0, c_Z: H0,
// c_G1: G1, c_Mf: H1)
static int assemble_lm(bh_object *o)
{
  if (o->MN != 1) return fail(BH_EUNSUP, "the low-memory form needs exactly one interior medium");
  const size_t NN = (size_t)o->NN, n2 = NN * NN;
  for (int d = 0; d <= 1; d++)
    if ((size_t)o->dom[d].nd != NN) return fail(BH_EUNSUP, "low-memory form: unexpected domain sizes");
  if (!o->c_G0f && (dalloc(&o->c_G0f, n2) || dalloc(&o->c_Z, n2) || dalloc(&o->c_Mf, n2) || dalloc(&o->c_G1, n2)))
    return BH_EHIP;
  cf *Gs[2] = {o->c_G0f, o->c_G1}, *Hs[2] = {o->c_Z, o->c_Mf};
  for (int d = 0; d <= 1; d++) {
    int rc = domain_blocks(o, d, [&](int r0, int n, const cd *G, const cd *H, const cd *, const cd *, const double *) {
      bem_cvt_z2c(G, Gs[d] + (size_t)r0 * NN, (size_t)n * NN, cur_st());
      bem_cvt_z2c(H, Hs[d] + (size_t)r0 * NN, (size_t)n * NN, cur_st());
      HIPCHK(hipStreamSynchronize(cur_st()));   // the block is assembled again for the next rows
      return 0;
    });
    if (rc) return rc;
  }
  o->lm_raw = true;
  return 0;
}

// every factorisation of earlier matrices is stale from the assembly on
static void asm_reset(bh_object *o)
{
  o->fmode = 0;
  o->x_warm = false;
  o->r_valid = false;
  o->mb_xwarm = false;
  for (Dom &D : o->dom) D.gf_valid = D.gfc_valid = false;
  o->asm_prec = -1;
  o->lm_raw = false;
}

// Assembly of N objects (create_cmatrix_domain of every domain,
// solve_bieq2.c:98-184): the FP64 and mixed forms of all objects and
// domains as one batch per form, one launch per assembly pass; the
// low-memory form object by object (its row blocks)
int bem::objects_assemble(bh_object **objs, int N, int type)
{
  if (type != 2) return fail(BH_EUNSUP, "assembly rule type %d: only 2 (GLN with GHN fallback) is implemented", type);
  {
    int rc = ensure_rules();   // the square's rules for quadrangles
    if (rc) return fail(rc, "rule tables");
  }
  std::vector<AsmSpec> sp[2];   // FP64 form, mixed form
  for (int i = 0; i < N; i++) {
    bh_object *o = objs[i];
    asm_reset(o);
    if (o->prec == BH_PREC_MIXED_LM) continue;
    for (int d = 0; d <= o->MN; d++) {
      AsmSpec x;
      int rc = domain_spec(o, d, x);
      if (rc) return rc;
      sp[o->prec != BH_PREC_FP64].push_back(x);
    }
  }
  for (int f = 0; f < 2; f++) {
    int rc = assemble_batch(sp[f]);
    if (rc) return rc;
  }
  for (int i = 0; i < N; i++) {
    if (objs[i]->prec == BH_PREC_MIXED_LM) {
      int rc = assemble_lm(objs[i]);
      if (rc) return rc;
    }
    objs[i]->asm_prec = objs[i]->prec;
  }
  return 0;
}

int bem::object_assemble(bh_object *o, int type)
{
  return objects_assemble(&o, 1, type);
}

extern "C" int bh_object_assemble(bh_object *o, int type)
{
  API_LOCK;
  return object_assemble(o, type);
}

extern "C" int bh_object_set_precision(bh_object *o, int prec)
{
  API_LOCK;
  if (prec != BH_PREC_FP64 && prec != BH_PREC_MIXED && prec != BH_PREC_MIXED_LM) return fail(BH_EARG, "precision %d", prec);
  if (prec == o->prec) return 0;
  // matrices, factors and warm starts of the other precision are stale:
  // the object has to be assembled (and factorised) again
  o->prec = prec;
  o->fmode = 0;
  o->x_warm = false;
  o->r_valid = false;
  o->mb_xwarm = false;
  o->s_formed = false;
  for (Dom &D : o->dom) D.gf_valid = D.gfc_valid = false;
  return 0;
}

extern "C" int bh_object_gh_count(const bh_object *o, int did) { return o->dom[did].n_gh; }

// rows per block of the low-memory form's on-demand assembly (0: as many as
// fit about 16 GB of FP64 rows)
extern "C" int bh_object_set_block_rows(bh_object *o, int rows)
{
  API_LOCK;
  if (rows < 0) return fail(BH_EARG, "block rows %d", rows);
  o->lm_rows = rows;
  return 0;
}

// which: 0 G, 1 H (nd x nd), 2 dG, 3 dH (nd x 2 x nd), 4 dF (nd x 3 doubles)
extern "C" int bh_object_matrix(const bh_object *o, int did, int which, double *out)
{
  if (o->prec == BH_PREC_MIXED_LM) return fail(BH_EUNSUP, "the low-memory form keeps no FP64 matrices");
  API_LOCK;
  if (did < 0 || did > o->MN) return fail(BH_EARG, "bad domain %d", did);
  const Dom &D = o->dom[did];
  size_t nd = (size_t)D.nd;
  const void *src = which == 0 ? (const void *)D.d_G : which == 1 ? (const void *)D.d_H : which == 2 ? (const void *)D.d_dG
                  : which == 3 ? (const void *)D.d_dH : (const void *)D.d_dF;
  size_t bytes = which < 2 ? nd * nd * sizeof(cd) : which < 4 ? 2 * nd * nd * sizeof(cd) : 3 * nd * sizeof(double);
  if (!src) return fail(BH_ESTATE, "matrices not assembled");
  HIPCHK(hipMemcpy(out, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}

// nrows rows of one matrix (rows[i] in 0..nd-1): out receives them one after
// another, each as bh_object_matrix lays a row out
extern "C" int bh_object_matrix_rows(const bh_object *o, int did, int which, int nrows, const int *rows, double *out)
{
  if (o->prec == BH_PREC_MIXED_LM) return fail(BH_EUNSUP, "the low-memory form keeps no FP64 matrices");
  API_LOCK;
  if (did < 0 || did > o->MN || which < 0 || which > 4 || nrows < 0) return fail(BH_EARG, "bad matrix selector");
  const Dom &D = o->dom[did];
  const size_t nd = (size_t)D.nd;
  const char *src = (const char *)(which == 0 ? (const void *)D.d_G : which == 1 ? (const void *)D.d_H
                                   : which == 2 ? (const void *)D.d_dG : which == 3 ? (const void *)D.d_dH
                                                                       : (const void *)D.d_dF);
  if (!src) return fail(BH_ESTATE, "matrices not assembled");
  const size_t rb = which < 2 ? nd * sizeof(cd) : which < 4 ? 2 * nd * sizeof(cd) : 3 * sizeof(double);
  for (int i = 0; i < nrows; i++) {
    if (rows[i] < 0 || (size_t)rows[i] >= nd) return fail(BH_EARG, "row %d out of range", rows[i]);
    HIPCHK(hipMemcpyAsync((char *)out + i * rb, src + (size_t)rows[i] * rb, rb, hipMemcpyDeviceToHost, cur_st()));
  }
  HIPCHK(hipStreamSynchronize(cur_st()));
  return 0;
}

// ------------------------------------------------------------------ dense LA entry points
extern "C" int bh_zgesv(int n, int nrhs, double *A, double *B)
{
  API_LOCK;
  if (ensure_init()) return BH_EHIP;
  cd *dA, *dB, *work;
  int *piv, *info;
  HIPCHK(hipMalloc((void **)&dA, (size_t)n * n * sizeof(cd)));
  HIPCHK(hipMalloc((void **)&dB, (size_t)n * nrhs * sizeof(cd)));
  HIPCHK(hipMalloc((void **)&work, 256 * sizeof(cd)));
  HIPCHK(hipMalloc((void **)&piv, (size_t)n * sizeof(int)));
  HIPCHK(hipMalloc((void **)&info, sizeof(int)));
  HIPCHK(hipMemcpy(dA, A, (size_t)n * n * sizeof(cd), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dB, B, (size_t)n * nrhs * sizeof(cd), hipMemcpyHostToDevice));
  if (bem_zgetrf(n, dA, n, piv, work, info, 64, cur_st())) return lu_failed();
  bem_zgetrs(n, nrhs, dA, n, piv, dB, nrhs, 64, cur_st());
  HIPCHK(hipGetLastError());
  int hinfo = 0;
  HIPCHK(hipMemcpy(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(B, dB, (size_t)n * nrhs * sizeof(cd), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(A, dA, (size_t)n * n * sizeof(cd), hipMemcpyDeviceToHost));
  dfree(dA); dfree(dB); dfree(work); dfree(piv); dfree(info);
  if (hinfo) return fail(BH_ESING, "singular matrix (zero pivot at %d)", hinfo);
  return trsv_failed(cur_st());
}

// A X = B on device buffers with no allocation (the caller passes A, B, piv
// and info; the factorisation's work area is the library's fixed device
// array, set up by bh_init); enqueue only
extern "C" int bh_zgesv_noalloc(int n, int nrhs, void *A, void *B, int *piv, int *info, void *stream)
{
  API_LOCK;
  if (g.dev < 0) return fail(BH_ESTATE, "bh_zgesv_noalloc: call bh_init first (it sets up the work area)");
  if (n < 1 || nrhs < 1 || !A || !B || !piv || !info) return fail(BH_EARG, "bh_zgesv_noalloc: bad arguments");
  if (n > bem_noalloc_nmax())
    return fail(BH_EARG, "bh_zgesv_noalloc: n = %d exceeds the fixed work area (n <= %d)", n, bem_noalloc_nmax());
  if (bem_zgesv_noalloc(g.dev, n, nrhs, A, B, piv, info, stream ? (hipStream_t)stream : g.st))
    return fail(BH_EHIP, "bh_zgesv_noalloc: no work area on device %d", g.dev);
  return 0;
}

extern "C" int bh_zgemm_host(int m, int n, int k, const double *alpha, const double *A, const double *B,
                             const double *beta, double *C)
{
  API_LOCK;
  if (ensure_init()) return BH_EHIP;
  cd *dA, *dB, *dC;
  HIPCHK(hipMalloc((void **)&dA, (size_t)m * k * sizeof(cd)));
  HIPCHK(hipMalloc((void **)&dB, (size_t)k * n * sizeof(cd)));
  HIPCHK(hipMalloc((void **)&dC, (size_t)m * n * sizeof(cd)));
  HIPCHK(hipMemcpy(dA, A, (size_t)m * k * sizeof(cd), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dB, B, (size_t)k * n * sizeof(cd), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dC, C, (size_t)m * n * sizeof(cd), hipMemcpyHostToDevice));
  bem_zgemm(m, n, k, alpha[0], alpha[1], dA, k, dB, n, beta[0], beta[1], dC, n, cur_st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(C, dC, (size_t)m * n * sizeof(cd), hipMemcpyDeviceToHost));
  dfree(dA); dfree(dB); dfree(dC);
  return 0;
}

// complex FP32 counterparts (the factorisation of the mixed-precision form)
extern "C" int bh_cgesv(int n, int nrhs, float *A, float *B)
{
  API_LOCK;
  if (ensure_init()) return BH_EHIP;
  cf *dA, *dB;
  int *piv, *info;
  HIPCHK(hipMalloc((void **)&dA, (size_t)n * n * sizeof(cf)));
  HIPCHK(hipMalloc((void **)&dB, (size_t)n * nrhs * sizeof(cf)));
  HIPCHK(hipMalloc((void **)&piv, (size_t)n * sizeof(int)));
  HIPCHK(hipMalloc((void **)&info, sizeof(int)));
  HIPCHK(hipMemcpy(dA, A, (size_t)n * n * sizeof(cf), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dB, B, (size_t)n * nrhs * sizeof(cf), hipMemcpyHostToDevice));
  if (bem_cgetrf(n, dA, n, piv, info, cur_st())) return lu_failed();
  bem_cgetrs(n, nrhs, dA, n, piv, dB, nrhs, cur_st());
  HIPCHK(hipGetLastError());
  int hinfo = 0;
  HIPCHK(hipMemcpy(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(B, dB, (size_t)n * nrhs * sizeof(cf), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(A, dA, (size_t)n * n * sizeof(cf), hipMemcpyDeviceToHost));
  dfree(dA); dfree(dB); dfree(piv); dfree(info);
  if (hinfo) return fail(BH_ESING, "singular matrix (zero pivot at %d)", hinfo);
  return trsv_failed(cur_st());
}

extern "C" int bh_cgemm_host(int m, int n, int k, const float *alpha, const float *A, const float *B, const float *beta,
                             float *C)
{
  API_LOCK;
  if (ensure_init()) return BH_EHIP;
  cf *dA, *dB, *dC;
  HIPCHK(hipMalloc((void **)&dA, (size_t)m * k * sizeof(cf)));
  HIPCHK(hipMalloc((void **)&dB, (size_t)k * n * sizeof(cf)));
  HIPCHK(hipMalloc((void **)&dC, (size_t)m * n * sizeof(cf)));
  HIPCHK(hipMemcpy(dA, A, (size_t)m * k * sizeof(cf), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dB, B, (size_t)k * n * sizeof(cf), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dC, C, (size_t)m * n * sizeof(cf), hipMemcpyHostToDevice));
  bem_cgemm(m, n, k, alpha[0], alpha[1], dA, k, dB, n, beta[0], beta[1], dC, n, cur_st());
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(C, dC, (size_t)m * n * sizeof(cf), hipMemcpyDeviceToHost));
  dfree(dA); dfree(dB); dfree(dC);
  return 0;
}

// ================================================================== solve side
extern "C" {
void bem_launch_build_A(const void *, const void *, int, size_t, const void *, size_t, size_t, void *, size_t,
                        hipStream_t);
void bem_launch_build_B(const void *, const void *, int, int, size_t, const void *, const void *, const void *, size_t,
                        void *, hipStream_t);
void bem_launch_store_X(const void *, int, const void *, size_t, const int *, const void *, const void *, void *, void *,
                        hipStream_t);
void bem_launch_eh_rows(const void *, const void *, int, int, size_t, const double *, const void *, const int *,
                        const void *, const void *, const void *, double, double, void *, void *, hipStream_t);
void bem_launch_pack_eh(int, const void *, const int *, const void *, const void *, void *, hipStream_t);
void bem_launch_unpack_eh(int, const void *, const int *, const void *, void *, void *, hipStream_t);
void bem_launch_hadamard_nn(void *, int, size_t, const void *, hipStream_t);
void bem_launch_axpby(const void *, double, double, const void *, void *, size_t, hipStream_t);
void bem_launch_scale_rows(void *, const void *, int, int, double, double, hipStream_t);
void bem_launch_domain(const double *, int, const double *, const int *, const int *, const int *, int, int *,
                       hipStream_t);
void bem_launch_field(const double *, const int *, int, int, const void *, const int *, int, double, double, const void *,
                      int, const void *, double, const void *, const void *, int, int, const void *, const void *,
                      const void *, const void *, void *, int, hipStream_t);
void bem_launch_field_memp(const double *, const int *, int, int, const void *, const void *, const int *, int, double,
                           double, const void *, int, double, int, const void *, int, double, const void *, int,
                           const void *, int, const void *, const void *, double, void *, hipStream_t);
void bem_launch_scatter(const double *, const double *, const int *, int, const void *, const void *, const int *, int,
                        double, double, const void *, int, const void *, int, double, const void *, int, const void *,
                        int, const void *, const void *, const void *, void *, void *, int, double *, hipStream_t);
void bem_launch_field_bd(const int *, const double *, int, int, const void *, const void *, const int *, int, double,
                         double, const void *, const void *, int, double, const void *, int, const double *,
                         const double *, const double *, const double *, const void *, const void *, const void *,
                         const void *, void *, double *, hipStream_t);
void bem_launch_incident_bd(const double *, const double *, const int *, int, const void *, int, void *, void *, void *,
                            void *, void *, hipStream_t);
void bem_launch_add_incident(const double *, const int *, int, const void *, int, void *, int, hipStream_t);
void bem_launch_fbeam(const double *, const double *, const int *, int, const void *, int, const double *,
                      const double *, int *, int, void *, const int *, void *, void *, void *, void *, void *,
                      hipStream_t);
void bem_launch_stress(const void *, const double *, const double *, const double *, int, double, const double *,
                       double *, hipStream_t);
void bem_launch_poynting(const void *, const double *, const double *, int, double *, hipStream_t);
void bem_launch_field_rows(const double *, const int *, const double *, int, int, const void *, const void *,
                           const int *, int, double, double, const void *, int, const void *, double, const void *, int,
                           const double *, const double *, const double *, const double *, size_t, void *, double *,
                           hipStream_t);
void bem_launch_pack_ehbv(const void *, const void *, const void *, const void *, size_t, int, int, void *, hipStream_t);
void bem_launch_elem_coef(int, const void *, const void *, int, const int *, const double *, const double *,
                          const double *, const int *, double, double, const void *, int, double, const void *, int,
                          const void *, int, double, const void *, int, int, void *, hipStream_t);
void bem_launch_elem_self(const void *, const void *, const int *, const double *, const int *, int, double, double,
                          const double *, const double *, const double *, const double *, void *, double *, void *,
                          hipStream_t);
void bem_launch_rows_add(const void *, int, int, const double *, int, const int *, size_t, void *, hipStream_t);
void bem_launch_meh_gather(int, int, const void *, const int *, const void *, const void *, size_t, void *, void *,
                           hipStream_t);
void bem_launch_meh_rows(int, int, const void *, const double *, const void *, const void *, const void *, size_t, double,
                         double, void *, void *, size_t, hipStream_t);
void bem_launch_mpack_eh(int, int, const void *, const int *, const void *, const void *, size_t, void *, hipStream_t);
void bem_launch_munpack_eh(int, int, const void *, const int *, const void *, void *, void *, size_t, hipStream_t);
}


int bem::ensure_rules()
{
  if (g_rule[0]) return 0;
  const double S3 = BEM_SQ3;
  auto N = [&](double z, double e, double *n) {
    n[0] = (1.0 - 2.5 * (z + S3 * e)) / 3.0;
    n[1] = (1.0 + 5.0 * z) / 3.0;
    n[2] = (1.0 - 2.5 * (z - S3 * e)) / 3.0;
  };
  std::vector<QPtN> r4, r7, rgl, rgh;
  // 4-point rule (lit_coef_4p, lit_coef.c:26-78): each node carries only its
  // own shape function, the centroid (weight -27/48) is shared
  const double P0 = 0.2, P1 = 0.3464101615137754587055, W0 = 0.5208333333333333333333, W1 = 0.5625;
  r4.push_back({-P0, -P1, W0, 1, 0, 0});
  r4.push_back({2 * P0, 0.0, W0, 0, 1, 0});
  r4.push_back({-P0, P1, W0, 0, 0, 1});
  { double n[3]; N(0, 0, n); r4.push_back({0.0, 0.0, -W1, n[0], n[1], n[2]}); }
  // 7-point rule (d3b1_const.h:60-67)
  const double a0 = 0.3480702390148154917985, a1 = 0.2052130961576726346557, a2 = 0.6028753385763033130539,
               a3 = 0.3554395089236065568357, w0 = 0.1259391805448271525957, w1 = 0.1323941527885061807376,
               w2 = 0.225;
  const double z7[7] = {-a0, 2 * a0, -a0, a1, -2 * a1, a1, 0.0}, e7[7] = {-a2, 0.0, a2, a3, 0.0, -a3, 0.0};
  const double ww7[7] = {w0, w0, w0, w1, w1, w1, w2};
  for (int i = 0; i < 7; i++) { double n[3]; N(z7[i], e7[i], n); r7.push_back({z7[i], e7[i], ww7[i], n[0], n[1], n[2]}); }
  for (const QPt &q : tri_rule(12, g.xl, g.wl)) { double n[3]; N(q.z, q.e, n); rgl.push_back({q.z, q.e, q.w, n[0], n[1], n[2]}); }
  for (const QPt &q : tri_rule(64, g.xh, g.wh)) { double n[3]; N(q.z, q.e, n); rgh.push_back({q.z, q.e, q.w, n[0], n[1], n[2]}); }
  std::vector<QPtN> *all[4] = {&r4, &r7, &rgl, &rgh};
  for (int i = 0; i < 4; i++) {
    if (upload(&g_rule[i], all[i]->data(), all[i]->size())) return BH_EHIP;
    g_nrule[i] = (int)all[i]->size();
  }
  g_rscale[0] = g_rscale[1] = BEM_I_EP;
  g_rscale[2] = g_rscale[3] = 0.0625 * BEM_SQ3 * BEM_I_SPSQ3;
  // the square's rules (bil_coef_4p/9p/GL/GH; constants d3b1_const.h:50-59):
  // node points with N = delta and weight 1, the 3x3 Gauss rule, and the
  // GLN x GLN and GHN x GHN tensor rules
  auto NQ = [](double z, double e, double *n) {
    const double ia = 1.0 / 0.5773502691896257645091;
    n[0] = 0.25 * (1.0 - z * ia) * (1.0 - e * ia);
    n[1] = 0.25 * (1.0 + z * ia) * (1.0 - e * ia);
    n[2] = 0.25 * (1.0 + z * ia) * (1.0 + e * ia);
    n[3] = 0.25 * (1.0 - z * ia) * (1.0 + e * ia);
  };
  std::vector<QPtQ> q4, q9, qgl, qgh;
  for (int i = 0; i < 4; i++) {
    QPtQ q = {ZT44[i], ET44[i], 1.0, {0.0, 0.0, 0.0, 0.0}};
    q.n[i] = 1.0;
    q4.push_back(q);
  }
  const double p9 = 0.7745966692414833770359, v0 = 0.3086419753086419753086, v1 = 0.4938271604938271604938,
               v2 = 0.7901234567901234567901;
  const double z9[9] = {-p9, p9, p9, -p9, 0.0, p9, 0.0, -p9, 0.0}, e9[9] = {-p9, -p9, p9, p9, -p9, 0.0, p9, 0.0, 0.0};
  const double w9[9] = {v0, v0, v0, v0, v1, v1, v1, v1, v2};
  for (int i = 0; i < 9; i++) { QPtQ q = {z9[i], e9[i], w9[i], {}}; NQ(q.z, q.e, q.n); q9.push_back(q); }
  for (int i = 0; i < 12; i++)
    for (int j = 0; j < 12; j++) { QPtQ q = {g.xl[i], g.xl[j], g.wl[i] * g.wl[j], {}}; NQ(q.z, q.e, q.n); qgl.push_back(q); }
  for (int i = 0; i < 64; i++)
    for (int j = 0; j < 64; j++) { QPtQ q = {g.xh[i], g.xh[j], g.wh[i] * g.wh[j], {}}; NQ(q.z, q.e, q.n); qgh.push_back(q); }
  std::vector<QPtQ> *qall[4] = {&q4, &q9, &qgl, &qgh};
  for (int i = 0; i < 4; i++) {
    if (upload(&g_qrule[i], qall[i]->data(), qall[i]->size())) return BH_EHIP;
    g_nqrule[i] = (int)qall[i]->size();
  }
  return 0;
}


// per-domain column tables, state buffers, slot and face tables
int bem::prepare_solve(bh_object *o)
{
  if (o->d_Ui) return 0;
  size_t E1 = (size_t)o->Ne + 1;
  if (dalloc(&o->d_Ui, E1 * 16) || dalloc(&o->d_dUi, E1 * 16) || dalloc(&o->d_Ui0, E1 * 16) ||
      dalloc(&o->d_dUi0, E1 * 16) || dalloc(&o->d_Uip, E1 * 16))
    return BH_EHIP;
  for (int d = 0; d <= o->MN; d++) {
    Dom &D = o->dom[d];
    std::vector<ColInfo> col(D.nd);
    std::vector<int> colS(D.nd);
    std::vector<RowEH> rows(D.nd);
    for (int s = 1; s <= D.dNe; s++) {
      int sd = D.sid[s], a = std::abs(sd);
      int mdid = o->md[a], sdid = o->sd[a];
      zd k1 = o->kn[mdid], k2 = o->kn[sdid], delt = k1 * k1 - k2 * k2, beps = k1 * k1 / (k2 * k2), sigm = 1.0 - beps;
      const int ns = o->ed[4 * a + 3] == 0 ? 3 : 4;
      for (int l = 0; l < ns; l++) {
        double vtz[3], vte[3];
        node_tangents(o, a, l, vtz, vte);
        unit3(vtz);
        unit3(vte);
        int j = D.colOff[s] + l;
        ColInfo &c = col[j];
        memset(&c, 0, sizeof c);
        c.node = o->eni[4 * a + l];
        c.elem = a;
        c.slot = l;
        c.main = mdid == d;
        c.mdid = mdid;
        for (int q = 0; q < 3; q++) c.n[q] = o->wen[(size_t)a * 12 + l * 3 + q];
        unit3(c.n);
        c.delt = mk(delt.real(), delt.imag());
        c.beps = mk(beps.real(), beps.imag());
        c.sigm = mk(sigm.real(), sigm.imag());
        colS[j] = s;
        RowEH &r = rows[j];
        double sg = sd < 0 ? -1.0 : 1.0;
        for (int q = 0; q < 3; q++) { r.vn[q] = sg * c.n[q]; r.vtz[q] = vtz[q]; r.vte[q] = vte[q]; }
        r.s = s;
        r.slot = l;
      }
    }
    size_t n1 = (size_t)(D.dNe + 1);
    D.has_inc = false;
    for (const ColInfo &c : col) D.has_inc = D.has_inc || (!c.main && c.mdid == 0);
    if (upload(&D.d_col, col.data(), col.size()) || upload(&D.d_colS, colS.data(), colS.size()) ||
        upload(&D.d_rowEH, rows.data(), rows.size()) || dalloc(&D.d_U, n1 * 16) || dalloc(&D.d_dU, n1 * 16) ||
        dalloc(&D.d_E, n1 * 12) || dalloc(&D.d_dE, n1 * 12) || dalloc(&D.d_Hf, n1 * 12) || dalloc(&D.d_dHf, n1 * 12))
      return BH_EHIP;
  }
  // open-domain node slots (4 per element, centroid included)
  std::vector<double> pt, nm;
  std::vector<int> idx;
  for (int i = 1; i <= o->Ne; i++) {
    if (o->md[i] != 0) continue;
    for (int j = 0; j < 4; j++) {
      double w[3];
      for (int q = 0; q < 3; q++) { pt.push_back(o->ren[(size_t)i * 12 + j * 3 + q]); w[q] = o->wen[(size_t)i * 12 + j * 3 + q]; }
      unit3(w);
      nm.insert(nm.end(), w, w + 3);
      idx.push_back(i * 4 + j);
    }
  }
  o->nslot = (int)idx.size();
  // faces by main domain for the solid-angle test
  // (4 vertices per face; a triangle repeats its first vertex and is
  // flagged by fq = 0)
  std::vector<double> fv;
  std::vector<int> fsd, fq, seg(o->MN + 2, 0);
  for (int d = 0; d <= o->MN; d++) {
    seg[d] = (int)fsd.size();
    for (int i = 1; i <= o->Ne; i++) {
      if (o->md[i] != d) continue;
      const bool quad = o->ed[4 * i + 3] != 0;
      for (int j = 0; j < 4; j++)
        for (int q = 0; q < 3; q++) fv.push_back(o->rn[3 * (size_t)o->ed[4 * i + (j < 3 || quad ? j : 0)] + q]);
      fsd.push_back(o->sd[i]);
      fq.push_back(quad ? 1 : 0);
    }
  }
  seg[o->MN + 1] = (int)fsd.size();
  if (upload(&o->d_slotPt, pt.data(), pt.size()) || upload(&o->d_slotN, nm.data(), nm.size()) ||
      upload(&o->d_slotIdx, idx.data(), idx.size()) || upload(&o->d_fv, fv.data(), fv.size()) ||
      upload(&o->d_fsd, fsd.data(), fsd.size()) || upload(&o->d_fq, fq.data(), fq.size()) ||
      upload(&o->d_seg, seg.data(), seg.size()))
    return BH_EHIP;
  HIPCHK(hipStreamSynchronize(cur_st()));
  return 0;
}

// one plane wave from 12 doubles (lambda0, ni, power, re/im e0x, re/im e0y,
// fx, fy, fz, theta, phi); derived data as setup_ipw (ipw.c:65-89)
static Beam plane_wave(const double *p)
{
  zd e0x(p[3], p[4]), e0y(p[5], p[6]);
  double nrm = 1.0 / sqrt(std::norm(e0x) + std::norm(e0y));
  zd ex = e0x * nrm, ey = e0y * nrm, hx = -ey * p[1], hy = ex * p[1];
  double st = sin(p[10]), ct = cos(p[10]), sp = sin(p[11]), cp = cos(p[11]);
  double a11 = sp * sp + cp * cp * ct, a12 = sp * cp * ct - sp * cp, a22 = sp * sp * ct + cp * cp;
  zd vex = ex * a11 + ey * a12, vey = ex * a12 + ey * a22, vez = -(ex * cp + ey * sp) * st;
  zd vhx = hx * a11 + hy * a12, vhy = hx * a12 + hy * a22, vhz = -(hx * cp + hy * sp) * st;
  Beam b;
  b.ki = 2.0 * M_PI * p[1] / p[0];
  b.E0 = sqrt(p[2] * 2.0 / p[1]);
  b.st = st; b.ct = ct; b.sp = sp; b.cp = cp; b.fx = p[7]; b.fy = p[8]; b.fz = p[9]; b.pad = 0;
  b.ex = mk(vex.real(), vex.imag()); b.ey = mk(vey.real(), vey.imag()); b.ez = mk(vez.real(), vez.imag());
  b.hx = mk(vhx.real(), vhx.imag()); b.hy = mk(vhy.real(), vhy.imag()); b.hz = mk(vhz.real(), vhz.imag());
  return b;
}

// phi points of the trapezoid in the order they are added (-pi, then the
// midpoints of each halving, fpw.c:140-156): cos, sin, angle; and the
// error flag of the phi integrals
static double *g_phit = nullptr;
static int *g_fberr = nullptr;
static int ensure_phi_table()
{
  if (g_phit) return 0;
  std::vector<double> t(3 * (size_t)BEM_NPHI, 0.0);
  double h = 2.0 * M_PI;
  t[0] = -1.0; t[1] = 0.0; t[2] = -M_PI;
  int j = 1, st = 1;
  for (int i = 0; i <= BEM_TRAP_MNUM; i++) {
    h *= 0.5;
    for (int q = 1; q <= st; q++, j++) {
      const double a = -M_PI + (double)(2 * q - 1) * h;
      t[3 * j] = cos(a); t[3 * j + 1] = sin(a); t[3 * j + 2] = a;
    }
    st *= 2;
  }
  if (upload(&g_phit, t.data(), t.size())) return BH_EHIP;
  HIPCHK(hipMalloc((void **)&g_fberr, sizeof(int)));
  HIPCHK(hipMemset(g_fberr, 0, sizeof(int)));
  return 0;
}

// the phi-integral error flag (read and cleared, synchronises the stream)
static int fbeam_failed()
{
  if (!g_fberr) return 0;
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, g_fberr, sizeof(int), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  if (!h) return 0;
  HIPCHK(hipMemsetAsync(g_fberr, 0, sizeof(int), cur_st()));
  return fail(BH_ECONV, "phi integration limit over (a beam's phi integral did not converge in 2^13 intervals)");
}

// incident fields of an object's beams added to np points (mask: points
// that take it; what 0: E, H rows of 6, 1: E into rows of 4)
static void add_incident(const bh_object *o, const double *d_pts, const int *d_mask, int np, cd *d_out, int what)
{
  bem_launch_add_incident(d_pts, d_mask, np, o->d_beams, (int)o->beams.size(), d_out, what, cur_st());
  if (!o->fbeams.empty())
    bem_launch_fbeam(d_pts, nullptr, d_mask, np, o->d_fb, (int)o->fbeams.size(), o->d_fbt, g_phit, g_fberr, what,
                     d_out, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, cur_st());
}

// a beam set in place: device copies, then the incident boundary data of the
// open-domain node slots (plane waves, then the other beams added)
static int install_beams(bh_object *o, const std::vector<Beam> &bm, const std::vector<FBeam> &fb,
                         const std::vector<double> &tht, double lam)
{
  if (!fb.empty() && ensure_phi_table()) return BH_EHIP;
  o->beams = bm;
  o->fbeams = fb;
  o->lambda0 = lam;
  dfree(o->d_beams); dfree(o->d_fb); dfree(o->d_fbt);
  o->d_beams = nullptr; o->d_fb = nullptr; o->d_fbt = nullptr;
  const Beam none = {};
  if (bm.empty() ? upload(&o->d_beams, &none, 1) : upload(&o->d_beams, bm.data(), bm.size())) return BH_EHIP;
  if (!fb.empty() && (upload(&o->d_fb, fb.data(), fb.size()) || upload(&o->d_fbt, tht.data(), tht.size())))
    return BH_EHIP;
  if (prepare_solve(o)) return BH_EHIP;
  bem_launch_incident_bd(o->d_slotPt, o->d_slotN, o->d_slotIdx, o->nslot, o->d_beams, (int)bm.size(), o->d_Ui, o->d_dUi,
                         o->d_Ui0, o->d_dUi0, o->d_Uip, cur_st());
  if (!fb.empty())
    bem_launch_fbeam(o->d_slotPt, o->d_slotN, nullptr, o->nslot, o->d_fb, (int)fb.size(), o->d_fbt, g_phit, g_fberr, 2,
                     nullptr, o->d_slotIdx, o->d_Ui, o->d_dUi, o->d_Ui0, o->d_dUi0, o->d_Uip, cur_st());
  HIPCHK(hipGetLastError());
  return fbeam_failed();
}

// plane waves only: 12 doubles per beam (ipw.txt columns)
extern "C" int bh_object_set_beams(bh_object *o, int nb, const double *par)
{
  API_LOCK;
  if (nb <= 0) return fail(BH_EARG, "no beams defined");
  std::vector<Beam> bm(nb);
  double n0 = par[1], lam = par[0];
  for (int i = 0; i < nb; i++) {
    const double *p = par + 12 * i;
    if (p[0] != lam || p[1] != n0) return fail(BH_EARG, "beam %d: wavelength / medium differ from beam 0", i);
    bm[i] = plane_wave(p);
  }
  return install_beams(o, bm, std::vector<FBeam>(), std::vector<double>(), lam);
}

// a beam set of any multi_fbeam types, 16 doubles per beam: type (0 plane
// wave, 1 focused plane wave, 2 focused with spiral phase, 3 Bessel, 4
// Bessel with spiral phase, 5 radially / azimuthally polarised focused),
// lambda0, ni, NA (types 1, 2, 5) or deflection angle (3, 4), power, re/im
// e0x (e0r), re/im e0y (e0a), fx, fy, fz, theta, phi, spiral mode (2, 4),
// Gauss-Legendre points over the aperture (1, 2, 5).  Derived data as
// setup_ipw, setup_Fpw (fpw.c:115-157), setup_LGb (lgb.c:116-164),
// setup_Bsb (bsb.c:112-148), setup_BsLGb (bslgb.c:115-156), setup_rab
// (rab.c:111-158); all beams share the medium index and wavelength
// (check_data_mfb, multi_fbeam.c:878-962).
extern "C" int bh_object_set_mfb(bh_object *o, int nb, const double *par)
{
  API_LOCK;
  if (nb <= 0) return fail(BH_EARG, "no beams defined");
  const double lam = par[1], n0 = par[2];
  std::vector<Beam> bm;
  std::vector<FBeam> fb;
  std::vector<double> tht;
  for (int i = 0; i < nb; i++) {
    const double *p = par + 16 * i;
    const int t = (int)p[0];
    if (p[1] != lam || p[2] != n0) return fail(BH_EARG, "beam %d: wavelength / medium differ from beam 0", i);
    if (t == 0) {
      const double w[12] = {p[1], p[2], p[4], p[5], p[6], p[7], p[8], p[9], p[10], p[11], p[12], p[13]};
      bm.push_back(plane_wave(w));
      continue;
    }
    if (t < 0 || t > 5) return fail(BH_EARG, "beam %d: unknown beam type %d", i, t);
    FBeam b;
    memset(&b, 0, sizeof b);
    b.type = t;
    b.ni = p[2];
    b.ki = 2.0 * M_PI * p[2] / p[1];
    b.cos_t = cos(p[12]); b.sin_t = sin(p[12]); b.cos_p = cos(p[13]); b.sin_p = sin(p[13]);
    b.fx = p[9]; b.fy = p[10]; b.fz = p[11];
    b.lg_m = (t == 2 || t == 4) ? (int)p[14] : 0;
    const zd e0x(p[5], p[6]), e0y(p[7], p[8]);
    const double a0 = sqrt(std::norm(e0x) + std::norm(e0y));
    zd ex, ey;
    if (t == 3 || t == 4) {
      b.st = sin(p[3]); b.ct = cos(p[3]);
      const double ce = 1.0 / a0;
      ex = e0x * ce; ey = e0y * ce;
      b.E0 = sqrt(2.0 * p[4] / p[2]);
    } else {
      const int nn = (int)p[15];
      if (nn < 1 || nn > 4096) return fail(BH_EARG, "beam %d: %d Gauss-Legendre points", i, nn);
      const double kr = p[3] / p[2], ce = sqrt(1.0 / M_PI) / (kr * a0);
      ex = e0x * ce; ey = e0y * ce;
      b.E0 = sqrt(2.0 * p[4] * p[2]) / p[1];
      b.nt = nn;
      b.toff = (int)(tht.size() / 3);
      std::vector<double> x(nn), w(nn);
      gl_nodes_ab(0.0, kr, nn, x.data(), w.data());
      for (int q = 0; q < nn; q++) {
        tht.push_back(sqrt(1.0 - x[q] * x[q]));
        tht.push_back(x[q]);
        tht.push_back(w[q] * x[q]);
      }
    }
    const zd hx = -ey * p[2], hy = ex * p[2];
    b.ex = mk(ex.real(), ex.imag()); b.ey = mk(ey.real(), ey.imag());
    b.hx = mk(hx.real(), hx.imag()); b.hy = mk(hy.real(), hy.imag());
    fb.push_back(b);
  }
  if (tht.empty()) tht.assign(3, 0.0);   // Bessel beams only: no aperture rule
  return install_beams(o, bm, fb, tht, lam);
}

static size_t row_base(const bh_object *o, int d)
{
  size_t b = 0;
  for (int q = 0; q < d; q++) b += 4 * (size_t)o->dom[q].nd;
  return b;
}

// the FP64 / FP32 factors of the block elimination as plain pointers
static SchurF<cd> schur64(const bh_object *o)
{
  const Dom &D1 = o->dom[1];
  return SchurF<cd>{o->NN, (size_t)o->NN, o->d_G0f, o->d_Z, o->d_Mf, o->d_Sf, D1.d_G, o->d_piv0, o->d_pivM, o->d_pivS,
                    D1.d_col, o->sch_beps, o->sch_sigm, o->sch_delt, o->s_formed, 1e-13};
}
static SchurF<cf> schur32(const bh_object *o)
{
  const Dom &D1 = o->dom[1];
  return SchurF<cf>{o->NN, (size_t)o->NN, o->c_G0f, o->c_Z, o->c_Mf, o->c_Sf, o->c_G1, o->d_piv0, o->d_pivM, o->d_pivS,
                    D1.d_col, to_cf(o->sch_beps), to_cf(o->sch_sigm), to_cf(o->sch_delt), o->s_formed, 1e-6};
}

// S = Sp - sigma delta G1 (M^{-1} G1 o n n^T) and its LU, for objects
// solved many times (the coupling loop); a single solve uses GMRES instead
int bem::form_s(bh_object *o)
{
  if (o->s_formed || (o->fmode != 2 && o->fmode != 3)) return 0;
  size_t n2 = (size_t)o->NN * o->NN;
  int rc;
  if (o->fmode == 2) {
    cd *Wt = o->ws_wt.get<cd>(n2);
    if (!Wt) return fail(BH_EHIP, "Schur complement: work area of %zu entries", n2);
    rc = schur_form_s(schur64(o), o->d_Sf, o->d_pivS, Wt, o->d_info, cur_st());
    (void)hipStreamSynchronize(cur_st());
  } else {
    cf *Wt = o->ws_wt.get<cf>(n2);
    if (!Wt) return fail(BH_EHIP, "Schur complement: work area of %zu entries", n2);
    rc = schur_form_s(schur32(o), o->c_Sf, o->d_pivS, Wt, o->d_info, cur_st());
    (void)hipStreamSynchronize(cur_st());
  }
  if (rc) return lu_failed();
  int info = 0;
  HIPCHK(hipMemcpyAsync(&info, o->d_info, sizeof(int), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  if (info) return fail(BH_ESING, "singular Schur complement (zero pivot %d)", info);
  o->s_formed = true;
  o->fact_gen++;
  return 0;
}

// mode 0: auto (Schur for one medium, dense otherwise), 1: dense 8NN,
// 2: Schur, 3: Schur in complex FP32 refined in FP64 (mixed precision)
int bem::object_factor(bh_object *o, int mode)
{
  const bool lm = o->prec == BH_PREC_MIXED_LM;
  if ((!lm && !o->dom[0].d_G) || o->asm_prec < 0) return fail(BH_ESTATE, "assemble before factor");
  if (o->asm_prec != o->prec) return fail(BH_ESTATE, "the precision changed since the assembly: assemble again");
  if (mode < 0 || mode > 3) return fail(BH_EARG, "factor mode %d", mode);
  if (lm && mode != 0 && mode != 3) return fail(BH_EUNSUP, "the low-memory form factorises in mode 3 only");
  if (lm && !o->lm_raw) return fail(BH_ESTATE, "low-memory form: the matrices were factorised in place: assemble again");
  if (prepare_solve(o)) return BH_EHIP;
  size_t NN = (size_t)o->NN, na = 8 * NN;
  if (mode == 0) mode = o->MN == 1 ? (o->prec != BH_PREC_FP64 ? 3 : 2) : 1;
  if (mode >= 2 && o->MN != 1) return fail(BH_EUNSUP, "the block elimination needs exactly one interior medium");
  o->fmode = 0;
  o->x_warm = false;
  o->r_valid = false;
  if (!o->d_info) { HIPCHK(hipMalloc((void **)&o->d_info, sizeof(int))); HIPCHK(hipMalloc((void **)&o->d_work, 256 * sizeof(cd))); }
  if (!o->d_X) { if (dalloc(&o->d_X, na) || dalloc(&o->d_B, na) || dalloc(&o->d_tmp, 8 * NN)) return BH_EHIP; }
  int info = 0;
  if (mode == 1) {
    if (!o->d_A) { HIPCHK(hipMalloc((void **)&o->d_A, na * na * sizeof(cd))); HIPCHK(hipMalloc((void **)&o->d_piv, na * sizeof(int))); }
    HIPCHK(hipMemsetAsync(o->d_A, 0, na * na * sizeof(cd), cur_st()));
    for (int d = 0; d <= o->MN; d++) {
      Dom &D = o->dom[d];
      bem_launch_build_A(D.d_G, D.d_H, D.nd, D.nd, D.d_col, row_base(o, d), NN, o->d_A, na, cur_st());
    }
    if (bem_zgetrf((int)na, o->d_A, na, o->d_piv, o->d_work, o->d_info, 64, cur_st())) return lu_failed();
  } else {
    // one medium: domain 0 rows H0 u - G0 q = 0, domain 1 rows couple through phi
    Dom &D0 = o->dom[0], &D1 = o->dom[1];
    if (D0.nd != (int)NN || D1.nd != (int)NN) return fail(BH_EUNSUP, "block elimination: unexpected domain sizes");
    size_t n2 = NN * NN;
    if (!o->d_piv0 && (dalloc(&o->d_piv0, NN) || dalloc(&o->d_pivM, NN) || dalloc(&o->d_pivS, NN) ||
                       dalloc(&o->d_w, schur_work(NN)) || dalloc(&o->d_R, na) || dalloc(&o->d_nrm, 256)))
      return BH_EHIP;
    o->s_formed = false;
    zd k1 = o->kn[0], k2 = o->kn[1], delt = k1 * k1 - k2 * k2, beps = k1 * k1 / (k2 * k2), sigm = 1.0 - beps;
    o->sch_delt = mk(delt.real(), delt.imag());
    o->sch_beps = mk(beps.real(), beps.imag());
    o->sch_sigm = mk(sigm.real(), sigm.imag());
    int n = (int)NN;
    if (mode == 3 && lm) {
      // the assembled FP32 matrices become the factors in place
      if ((!o->c_Sf && dalloc(&o->c_Sf, n2)) || (!o->c_w && dalloc(&o->c_w, schur_work(NN) + 16 * NN))) return BH_EHIP;
      o->lm_raw = false;
      if (schur_factor_c_inplace(n, NN, o->c_G0f, o->c_Z, o->c_G1, o->c_Mf, o->sch_beps, o->c_Sf, o->d_piv0, o->d_pivM,
                                 o->d_info, cur_st()))
        return lu_failed();
    } else if (mode == 3) {
      if (!o->c_G0f && (dalloc(&o->c_G0f, n2) || dalloc(&o->c_Z, n2) || dalloc(&o->c_Mf, n2) || dalloc(&o->c_Sf, n2) ||
                        dalloc(&o->c_G1, n2) || dalloc(&o->c_w, schur_work(NN) + 16 * NN)))
        return BH_EHIP;
      if (schur_factor_c(n, NN, D0.d_G, D0.d_H, D1.d_G, D1.d_H, o->sch_beps, o->c_G0f, o->c_Z, o->c_Mf, o->c_Sf,
                         o->c_G1, o->d_piv0, o->d_pivM, o->d_info, cur_st()))
        return lu_failed();
    } else {
      if (!o->d_G0f && (dalloc(&o->d_G0f, n2) || dalloc(&o->d_Z, n2) || dalloc(&o->d_Mf, n2) || dalloc(&o->d_Sf, n2)))
        return BH_EHIP;
      // Z = G0^{-1} H0
      HIPCHK(hipMemcpyAsync(o->d_G0f, D0.d_G, n2 * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
      HIPCHK(hipMemcpyAsync(o->d_Z, D0.d_H, n2 * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
      if (bem_zgetrf(n, o->d_G0f, NN, o->d_piv0, o->d_work, o->d_info, 64, cur_st())) return lu_failed();
      bem_zgetrs(n, n, o->d_G0f, NN, o->d_piv0, o->d_Z, NN, 64, cur_st());
      // M = H1 + G1 Z  (Mf), P = G1 Z kept in Sf for the beta term
      HIPCHK(hipMemcpyAsync(o->d_Mf, D1.d_H, n2 * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
      bem_zgemm(n, n, n, 1.0, 0.0, D1.d_G, NN, o->d_Z, NN, 0.0, 0.0, o->d_Sf, NN, cur_st());
      bem_launch_axpby(o->d_Mf, 1.0, 0.0, o->d_Sf, o->d_Mf, n2, cur_st());
      // S starts as H1 + beta P
      bem_launch_axpby(D1.d_H, beps.real(), beps.imag(), o->d_Sf, o->d_Sf, n2, cur_st());
      if (bem_zgetrf(n, o->d_Mf, NN, o->d_pivM, o->d_work, o->d_info, 64, cur_st())) return lu_failed();
      // S itself is formed on demand (form_s): a single solve takes GMRES on Sp
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&info, o->d_info, sizeof(int), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  if (info) return fail(BH_ESING, "singular system (zero pivot %d)", info);
  o->fmode = mode;
  o->fact_gen++;
  o->mb_xwarm = false;
  return 0;
}

extern "C" int bh_object_factor(bh_object *o, int mode)
{
  API_LOCK;
  return object_factor(o, mode);
}


// iterative refinement of the mixed-precision form: relative residual
// target ||b - A x|| / ||b|| and step limit
#define IR_TOL 1e-10
#define IR_MAX 30

// B from the incident data, X = A^{-1} B, boundary values (solve_coefficient_mEMP)
// B rows of domain d (create_Bmatrix_dac) and the FP64 residual rows of
// domain d, over the domain's matrices a block of rows at a time
static int dom_build_B(bh_object *o, int d)
{
  Dom &D = o->dom[d];
  if (!D.has_inc) return 0;   // no incident data in these rows: they stay zero
  const size_t rb = row_base(o, d);
  if (o->prec != BH_PREC_MIXED_LM) {
    // the incident data as two column operands a, c (nd x 4): B = -(H a + G c),
    // two streaming products over the stored matrices (work: 12 nd values)
    cd *w = o->ws_b.get<cd>(12 * (size_t)D.nd);
    if (!w) return fail(BH_EHIP, "right-hand side: work area");
    bem_build_B_multi(D.d_G, D.d_H, D.nd, D.nd, D.d_col, o->d_Ui, o->d_dUi, ((size_t)o->Ne + 1) * 16, (size_t)o->NN, rb,
                      o->d_B, w, 1, cur_st());
    return 0;
  }
  return domain_blocks(o, d, [&](int r0, int n, const cd *G, const cd *H, const cd *, const cd *, const double *) {
    bem_launch_build_B(G, H, n, D.nd, D.nd, D.d_col, o->d_Ui, o->d_dUi, rb + r0, o->d_B, cur_st());
    return 0;
  });
}

static int dom_resid(bh_object *o, int d)
{
  Dom &D = o->dom[d];
  const size_t rb = row_base(o, d);
  bem_resid_colvec(D.d_col, D.nd, o->d_X, (size_t)o->NN, o->d_w, cur_st());
  return domain_blocks(o, d, [&](int r0, int n, const cd *G, const cd *H, const cd *, const cd *, const double *) {
    bem_resid_rows(G, H, n, D.nd, D.nd, r0, rb, o->d_B, o->d_R, o->d_w, cur_st());
    return 0;
  });
}

static int object_solve_core(bh_object *o)
{
  if (!o->fmode) return fail(BH_ESTATE, "factor before solve");
  if (!o->d_beams) return fail(BH_ESTATE, "set beams before solve");
  size_t NN = (size_t)o->NN, na = 8 * NN;
  o->r_valid = false;   // B changes here; R is not kept by this path
  HIPCHK(hipMemsetAsync(o->d_B, 0, na * sizeof(cd), cur_st()));
  for (int d = 0; d <= o->MN; d++) {
    int rc = dom_build_B(o, d);
    if (rc) return rc;
  }
  o->ir_iters = 0;
  o->ir_res = 0.0;
  if (o->fmode == 1) {
    HIPCHK(hipMemcpyAsync(o->d_X, o->d_B, na * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
    bem_zgetrs((int)na, 1, o->d_A, na, o->d_piv, o->d_X, 1, 64, cur_st());
  } else if (o->fmode == 2) {
    // the incident data enter the domain-1 rows only
    int rc = schur_apply(schur64(o), o->d_B, false, o->d_X, o->d_w, cur_st(), &o->gm_steps);
    if (rc) return fail(rc, "GMRES on the Schur complement did not converge");
  } else {
    // complex-FP32 solves of the residual equation, residuals in FP64; a
    // repeated solve (the coupling loop) starts from the previous solution,
    // whose residual for the new B is usually a small fraction of B
    const double nb = bem_nrm2(o->d_B, na, o->d_nrm, cur_st());
    if (nb < 0.0) return fail(BH_EHIP, "norm of the right-hand side");
    double res = nb > 0.0 ? 1.0 : 0.0;
    if (o->x_warm && nb > 0.0) {
      for (int d = 0; d <= o->MN; d++) {
        int rc = dom_resid(o, d);
        if (rc) return rc;
      }
      const double nr = bem_nrm2(o->d_R, na, o->d_nrm, cur_st());
      if (nr < 0.0) return fail(BH_EHIP, "norm of the residual");
      res = nr / nb;
    } else {
      HIPCHK(hipMemsetAsync(o->d_X, 0, na * sizeof(cd), cur_st()));
      HIPCHK(hipMemcpyAsync(o->d_R, o->d_B, na * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
    }
    o->x_warm = false;
    o->r_valid = false;
    cf *cw = o->c_w, *crhs = cw + schur_work(NN), *cdx = crhs + 8 * NN;
    o->gm_steps = 0;
    const SchurF<cf> f = schur32(o);
    int it = 0;
    while (res > IR_TOL && it < IR_MAX) {
      bem_cvt_z2c(o->d_R, crhs, na, cur_st());
      int gs = 0;
      if (schur_apply(f, crhs, true, cdx, cw, cur_st(), &gs)) return fail(BH_ECONV, "GMRES on the FP32 Schur complement did not converge");
      o->gm_steps += gs;
      bem_acc_c2z(o->d_X, cdx, na, cur_st());
      for (int d = 0; d <= o->MN; d++) {
        int rc = dom_resid(o, d);
        if (rc) return rc;
      }
      const double nr = bem_nrm2(o->d_R, na, o->d_nrm, cur_st());
      if (nr < 0.0) return fail(BH_EHIP, "norm of the residual");
      res = nr / nb;
      it++;
    }
    o->ir_iters = it;
    o->ir_res = res;
    if (!(res <= IR_TOL))
      return fail(BH_ECONV, "mixed-precision refinement: relative residual %.3g after %d steps", res, it);
    o->x_warm = true;
  }
  for (int d = 0; d <= o->MN; d++) {
    Dom &D = o->dom[d];
    bem_launch_store_X(o->d_X, D.nd, D.d_col, NN, D.d_colS, o->d_Ui, o->d_dUi, D.d_U, D.d_dU, cur_st());
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(cur_st()));
  return 0;
}

int bem::object_solve(bh_object *o)
{
  const int rc = object_solve_core(o);
  // the one-launch triangular solves' timeout flag is read and cleared on
  // every exit: a timed-out wait is the cause of whatever failed after it
  const int tr = trsv_failed(cur_st());
  return tr ? tr : rc;
}

extern "C" int bh_object_solve(bh_object *o)
{
  API_LOCK;
  return object_solve(o);
}

extern "C" int bh_object_solve_stats(const bh_object *o, int *iters, double *resid, int *gmres_steps)
{
  API_LOCK;
  if (iters) *iters = o->ir_iters;
  if (resid) *resid = o->ir_res;
  if (gmres_steps) *gmres_steps = o->gm_steps;
  return 0;
}

// G x = b for 6 right-hand sides (nd x 6, b in R, x into R) with the
// complex-FP32 LU of G, refined against the FP64 G
template <class F> static int domain_blocks(bh_object *o, int d, F fn);
static int solve_g_mixed(bh_object *o, int d, const cf *Gc, const int *piv, cd *R, int ncol = 6)
{
  const Dom &D = o->dom[d];
  const size_t nd = (size_t)D.nd, m = (size_t)ncol * nd;
  cd *X = o->ws_eh[2].get<cd>(m), *Rr = o->ws_eh[3].get<cd>(m);
  cf *C = o->ws_eh[4].get<cf>(m);
  if (!X || !Rr || !C) return fail(BH_EHIP, "refinement of the boundary fields: work area");
  HIPCHK(hipMemsetAsync(X, 0, m * sizeof(cd), cur_st()));
  HIPCHK(hipMemcpyAsync(Rr, R, m * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
  const double nb = bem_nrm2(R, m, o->d_nrm, cur_st());
  double res = nb > 0.0 ? 1.0 : 0.0;
  int it = 0, rc = 0;
  while (res > IR_TOL && it < IR_MAX) {
    bem_cvt_z2c(Rr, C, m, cur_st());
    bem_cgetrs(D.nd, ncol, Gc, nd, piv, C, ncol, cur_st());
    bem_acc_c2z(X, C, m, cur_st());
    HIPCHK(hipMemcpyAsync(Rr, R, m * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
    rc = domain_blocks(o, d, [&](int r0, int n, const cd *G, const cd *, const cd *, const cd *, const double *) {
      bem_zgemm(n, ncol, D.nd, -1.0, 0.0, G, nd, X, ncol, 1.0, 0.0, Rr + (size_t)r0 * ncol, ncol, cur_st());
      return 0;
    });
    if (rc) break;
    const double nr = bem_nrm2(Rr, m, o->d_nrm, cur_st());
    if (nr < 0.0) { rc = fail(BH_EHIP, "norm of the residual"); break; }
    res = nr / nb;
    it++;
  }
  if (!rc && !(res <= IR_TOL)) rc = fail(BH_ECONV, "mixed-precision refinement (G): relative residual %.3g after %d steps", res, it);
  if (!rc) HIPCHK(hipMemcpyAsync(R, X, m * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  return rc;
}

__global__ void k_zero_int(int *__restrict__ a, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// the complex-FP32 LU of G_d (d >= 1) the mixed form's boundary fields
// solve with (solve_deh_bv, mo_solve.c:663-789), once per assembly; info:
// a device int for the LU's zero-pivot report
static int factor_g_mixed(bh_object *o, int d, int *info)
{
  Dom &D = o->dom[d];
  const size_t nd = (size_t)D.nd;
  if (!D.c_Gf && dalloc(&D.c_Gf, nd * nd)) return BH_EHIP;
  if (!D.d_Gpiv && dalloc(&D.d_Gpiv, nd)) return BH_EHIP;
  if (o->prec == BH_PREC_MIXED_LM)   // G1 in complex FP32, kept by the factorisation
    HIPCHK(hipMemcpyAsync(D.c_Gf, o->c_G1, nd * nd * sizeof(cf), hipMemcpyDeviceToDevice, cur_st()));
  else
    bem_cvt_z2c(D.d_G, D.c_Gf, nd * nd, cur_st());
  if (bem_cgetrf(D.nd, D.c_Gf, nd, D.d_Gpiv, info, cur_st())) return lu_failed();
  D.gfc_valid = true;
  D.gf_valid = false;   // the shared pivot vector now belongs to the FP32 factors
  return 0;
}

// factor_g_mixed of every domain d >= 1 of a mixed-form object ahead of its
// boundary fields (the coupling loop runs it beside the sweeps); its own
// zero-pivot check, on the thread's current stream
int bem::object_factor_g(bh_object *o)
{
  if (o->fmode != 3) return 0;
  if (!o->d_ginfo && dalloc(&o->d_ginfo, 1)) return BH_EHIP;
  for (int d = 1; d <= o->MN; d++) {
    if (o->dom[d].gfc_valid) continue;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, cur_st(), o->d_ginfo, 1);
    int rc = factor_g_mixed(o, d, o->d_ginfo);
    if (rc) return rc;
    int info = 0;
    HIPCHK(hipMemcpyAsync(&info, o->d_ginfo, sizeof(int), hipMemcpyDeviceToHost, cur_st()));
    HIPCHK(hipStreamSynchronize(cur_st()));
    if (info) {
      o->dom[d].gfc_valid = false;
      return fail(BH_ESING, "singular G of domain %d (zero pivot %d)", d, info);
    }
  }
  return 0;
}

// boundary E, H from the derivative rows, then dE/dn, dH/dn from
// G [dE|dH] = H [E|H] per domain (solve_eh_bv + solve_deh_bv)
int bem::object_solve_EH(bh_object *o)
{
  if (!o->dom[0].d_U) return fail(BH_ESTATE, "solve before the boundary fields");
  zd ch = o->lambda0 / (2.0 * M_PI * zd(0.0, 1.0));
  if (!o->d_info) { HIPCHK(hipMalloc((void **)&o->d_info, sizeof(int))); HIPCHK(hipMalloc((void **)&o->d_work, 256 * sizeof(cd))); }
  HIPCHK(hipMemsetAsync(o->d_info, 0, sizeof(int), cur_st()));
  const bool mixed = o->fmode == 3;
  if (mixed && !o->d_nrm && dalloc(&o->d_nrm, 256)) return BH_EHIP;
  for (int d = 0; d <= o->MN; d++) {
    Dom &D = o->dom[d];
    size_t nd = (size_t)D.nd;
    int rc = domain_blocks(o, d, [&](int r0, int n, const cd *, const cd *, const cd *dG, const cd *dH, const double *dF) {
      bem_launch_eh_rows(dG, dH, n, D.nd, nd, dF, D.d_rowEH + r0, D.d_colS, D.d_col, D.d_U, D.d_dU, ch.real(),
                         ch.imag(), D.d_E, D.d_Hf, cur_st());
      return 0;
    });
    if (rc) return rc;
    // LU of G_d: domain 0 of the block form already factorised G0 for this
    // assembly (in the precision of the form); otherwise factorise G_d once
    // per assembly
    const cd *Gf = D.d_Gf;
    const cf *Gc = D.c_Gf;
    const int *Gpiv = D.d_Gpiv;
    if (d == 0 && (o->fmode == 2 || mixed)) {
      Gf = o->d_G0f;
      Gc = o->c_G0f;
      Gpiv = o->d_piv0;
    } else if (mixed) {
      if (!D.gfc_valid) {
        rc = factor_g_mixed(o, d, o->d_info);
        if (rc) return rc;
      }
      Gc = D.c_Gf;
      Gpiv = D.d_Gpiv;
    } else if (!D.gf_valid) {
      if (!D.d_Gf && dalloc(&D.d_Gf, nd * nd)) return BH_EHIP;
      if (!D.d_Gpiv && dalloc(&D.d_Gpiv, nd)) return BH_EHIP;
      HIPCHK(hipMemcpyAsync(D.d_Gf, D.d_G, nd * nd * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
      if (bem_zgetrf(D.nd, D.d_Gf, nd, D.d_Gpiv, o->d_work, o->d_info, 64, cur_st())) return lu_failed();
      D.gf_valid = true;
      D.gfc_valid = false;
      Gf = D.d_Gf;
      Gpiv = D.d_Gpiv;
    }
    cd *P = o->ws_eh[0].get<cd>(nd * 6), *R = o->ws_eh[1].get<cd>(nd * 6);
    if (!P || !R) return fail(BH_EHIP, "boundary fields: work area");
    bem_launch_pack_eh(D.nd, D.d_col, D.d_colS, D.d_E, D.d_Hf, P, cur_st());
    rc = domain_blocks(o, d, [&](int r0, int n, const cd *, const cd *H, const cd *, const cd *, const double *) {
      bem_zgemm(n, 6, D.nd, 1.0, 0.0, H, nd, P, 6, 0.0, 0.0, R + (size_t)r0 * 6, 6, cur_st());
      return 0;
    });
    if (rc) return rc;
    if (mixed)
      rc = solve_g_mixed(o, d, Gc, Gpiv, R);
    else
      bem_zgetrs(D.nd, 6, Gf, nd, Gpiv, R, 6, 64, cur_st());
    if (!rc) bem_launch_unpack_eh(D.nd, D.d_col, D.d_colS, R, D.d_dE, D.d_dHf, cur_st());
    HIPCHK(hipStreamSynchronize(cur_st()));
    if (rc) return rc;
  }
  HIPCHK(hipGetLastError());
  int info = 0;
  HIPCHK(hipMemcpyAsync(&info, o->d_info, sizeof(int), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  if (info) return fail(BH_ESING, "singular G (zero pivot %d)", info);
  return trsv_failed(cur_st());
}

extern "C" int bh_object_solve_EH(bh_object *o)
{
  API_LOCK;
  return object_solve_EH(o);
}


// ------------------------------------------------------------------ many beams as right-hand sides
__global__ void k_zero_cd_api(cd *__restrict__ a, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = mk(0.0, 0.0);
}
static void zero_cd(cd *a, size_t n) { hipLaunchKernelGGL(k_zero_cd_api, g1(n), dim3(256), 0, cur_st(), a, n); }

// R += B - Bp (the residual of a warm start from the last one)
__global__ void k_mresid_update(cd *__restrict__ R, const cd *__restrict__ B, const cd *__restrict__ Bp, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) R[i] = R[i] + (B[i] - Bp[i]);
}

cd *bem::mb_ptr(bh_object *o, int which) { return (cd *)o->mb_inc[which].p; }
cd *bem::mb_dom_ptr(bh_object *o, int d, int which) { return (cd *)o->mb_dom[(size_t)d * 6 + which].p; }

// nb plane waves as separate right-hand sides: device records and the
// incident data of every beam at the open-domain slots (init_boundary_data_u2,
// mo_setup.c:704-757, one beam at a time)
int bem::object_beams_set(bh_object *o, int nb, const double *par)
{
  if (nb <= 0) return fail(BH_EARG, "no beams");
  if (o->prec == BH_PREC_MIXED_LM) return fail(BH_EUNSUP, "many beams: not in the low-memory form");
  if (prepare_solve(o)) return BH_EHIP;
  const size_t E1 = (size_t)o->Ne + 1;
  for (int q = 0; q < 5; q++)
    if (!o->mb_inc[q].get<cd>(E1 * 16 * nb)) return fail(BH_EHIP, "beams: incident data of %d beams", nb);
  o->mb_dom.resize((size_t)(o->MN + 1) * 6);
  for (int d = 0; d <= o->MN; d++)
    for (int q = 0; q < 6; q++)
      if (!o->mb_dom[(size_t)d * 6 + q].get<cd>((size_t)(o->dom[d].dNe + 1) * (q < 2 ? 16 : 12) * nb))
        return fail(BH_EHIP, "beams: boundary values of %d beams", nb);
  const size_t na = 8 * (size_t)o->NN;
  for (int q = 0; q < 4; q++)
    if (!o->mb_sys[q].get<cd>(na * nb)) return fail(BH_EHIP, "beams: systems of %d beams", nb);
  std::vector<Beam> bm(nb);
  for (int b = 0; b < nb; b++) {
    const double *p = par + 12 * (size_t)b;
    if (p[0] != o->lambda0 || p[1] != o->n[0].real() || o->n[0].imag() != 0.0)
      return fail(BH_EARG, "beam %d: wavelength / medium differ from the object's", b);
    bm[b] = plane_wave(p);
  }
  Beam *db = o->mb_beams.get<Beam>(nb);
  if (!db) return fail(BH_EHIP, "beams: records");
  HIPCHK(hipMemcpyAsync(db, bm.data(), nb * sizeof(Beam), hipMemcpyHostToDevice, cur_st()));
  o->mb_hbeams = bm;
  for (int q = 0; q < 5; q++) zero_cd(mb_ptr(o, q), E1 * 16 * nb);
  // the boundary values start at zero like the one-beam arrays (dalloc):
  // the solves write only the slots the elements use, and the areas may
  // hold an earlier allocation's data
  for (int d = 0; d <= o->MN; d++)
    for (int q = 0; q < 6; q++)
      zero_cd(mb_dom_ptr(o, d, q), (size_t)(o->dom[d].dNe + 1) * (q < 2 ? 16 : 12) * nb);
  for (int b = 0; b < nb; b++) {
    const size_t off = (size_t)b * E1 * 16;
    bem_launch_incident_bd(o->d_slotPt, o->d_slotN, o->d_slotIdx, o->nslot, db + b, 1, mb_ptr(o, 2) + off,
                           mb_ptr(o, 3) + off, mb_ptr(o, 0) + off, mb_ptr(o, 1) + off, mb_ptr(o, 4) + off, cur_st());
  }
  o->nbm = nb;
  o->mb_xwarm = false;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(cur_st()));   // the host records stay alive until here
  return 0;
}

// B, X = A^{-1} B and the boundary values of every beam at once: the FP32
// block solve of the residual equations on all beams (its products and
// triangular solves with 3 nb to 4 nb columns), the FP64 residuals of all
// beams as two products per domain with 4 nb columns, until every beam's
// relative residual is at most IR_TOL
int bem::object_solve_multi(bh_object *o)
{
  if (o->fmode != 3 || !o->s_formed) return fail(BH_EUNSUP, "many beams: the mixed-precision form with S formed");
  const int nb = o->nbm;
  if (nb <= 0) return fail(BH_ESTATE, "many beams: no beams set");
  const size_t NN = (size_t)o->NN, na = 8 * NN, E1 = (size_t)o->Ne + 1;
  cd *B = (cd *)o->mb_sys[0].p, *X = (cd *)o->mb_sys[1].p, *R = (cd *)o->mb_sys[2].p, *Bp = (cd *)o->mb_sys[3].p;
  cf *crhs = o->mb_ws[0].get<cf>(na * nb), *cdx = o->mb_ws[1].get<cf>(na * nb);
  cf *cw = o->mb_ws[2].get<cf>(schur_work_multi(NN, nb));
  size_t ndmax = 0;
  for (int d = 0; d <= o->MN; d++) ndmax = std::max(ndmax, (size_t)o->dom[d].nd);
  cd *rw = o->mb_ws[3].get<cd>(12 * ndmax * nb);
  double *nrm = o->mb_ws[4].get<double>(2 * (size_t)nb);
  if (!crhs || !cdx || !cw || !rw || !nrm) return fail(BH_EHIP, "many beams: work areas");
  // B of every beam: per domain two products with 4 nb columns
  zero_cd(B, na * nb);
  for (int d = 0; d <= o->MN; d++) {
    Dom &D = o->dom[d];
    if (D.has_inc)
      bem_build_B_multi(D.d_G, D.d_H, D.nd, D.nd, D.d_col, mb_ptr(o, 2), mb_ptr(o, 3), E1 * 16, NN, row_base(o, d), B,
                        rw, nb, cur_st());
  }
  auto resid = [&]() {
    for (int d = 0; d <= o->MN; d++) {
      Dom &D = o->dom[d];
      bem_resid_domain_multi(D.d_G, D.d_H, D.nd, D.nd, D.d_col, X, NN, row_base(o, d), B, R, rw, nb, cur_st());
    }
  };
  std::vector<double> hb(nb), hr(nb);
  bem_mnrm2(B, na, nb, nrm, cur_st());
  HIPCHK(hipMemcpyAsync(hb.data(), nrm, nb * sizeof(double), hipMemcpyDeviceToHost, cur_st()));
  // warm start: R holds Bp - A X of the last solve, so B - A X = R + (B - Bp)
  if (o->mb_xwarm) hipLaunchKernelGGL(k_mresid_update, g1(na * nb), dim3(256), 0, cur_st(), R, B, Bp, na * nb);
  else {
    zero_cd(X, na * nb);
    HIPCHK(hipMemcpyAsync(R, B, na * nb * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
  }
  o->mb_xwarm = false;
  const SchurF<cf> f = schur32(o);
  // explicit FP32 inverses of the factors, formed once per factorisation:
  // every solve of the refinement is then a few products over all beams
  const size_t n2 = NN * NN;
  cf *G0i = o->mb_inv[0].get<cf>(n2), *Mi = o->mb_inv[1].get<cf>(n2), *Si = o->mb_inv[2].get<cf>(n2);
  if (!G0i || !Mi || !Si) return fail(BH_EHIP, "many beams: inverses of order %zu", NN);
  if (o->mb_inv_gen != o->fact_gen) {
    if (schur_inverses(f, G0i, Mi, Si, cur_st())) return fail(BH_EHIP, "many beams: inverses");
    o->mb_inv_gen = o->fact_gen;
  }
  const SchurInv<cf> inv{G0i, Mi, Si};
  int it = 0;
  double worst = 0.0;
  for (;;) {
    bem_mnrm2(R, na, nb, nrm + nb, cur_st());
    HIPCHK(hipMemcpyAsync(hr.data(), nrm + nb, nb * sizeof(double), hipMemcpyDeviceToHost, cur_st()));
    HIPCHK(hipStreamSynchronize(cur_st()));
    worst = 0.0;
    for (int b = 0; b < nb; b++) worst = std::max(worst, hb[b] > 0.0 ? sqrt(hr[b] / hb[b]) : 0.0);
    if (worst <= IR_TOL || it >= IR_MAX) break;
    bem_cvt_z2c(R, crhs, na * nb, cur_st());
    if (schur_apply_multi(f, crhs, true, cdx, cw, nb, cur_st(), &inv)) return fail(BH_ESTATE, "many beams: S is not formed");
    bem_acc_c2z(X, cdx, na * nb, cur_st());
    resid();
    it++;
  }
  o->mb_iters = it;
  o->mb_res = worst;
  if (!(worst <= IR_TOL))
    return fail(BH_ECONV, "many beams: relative residual %.3g after %d refinement steps", worst, it);
  o->mb_xwarm = true;
  HIPCHK(hipMemcpyAsync(Bp, B, na * nb * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
  for (int b = 0; b < nb; b++)
    for (int d = 0; d <= o->MN; d++) {
      Dom &D = o->dom[d];
      const size_t dn = (size_t)(D.dNe + 1) * 16;
      bem_launch_store_X(X + (size_t)b * na, D.nd, D.d_col, NN, D.d_colS, mb_ptr(o, 2) + (size_t)b * E1 * 16,
                         mb_ptr(o, 3) + (size_t)b * E1 * 16, mb_dom_ptr(o, d, 0) + b * dn, mb_dom_ptr(o, d, 1) + b * dn,
                         cur_st());
    }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(cur_st()));
  return 0;
}

// boundary E, H and their normal derivatives of every beam (solve_eh_bv +
// solve_deh_bv, mo_solve.c:547-789): the tangential derivative rows as two
// products dG [dU], dH [U] with 4 nb columns, then G [dE|dH] = H [E|H] with
// 6 nb columns (the FP32 LU of G refined in FP64)
int bem::object_solve_EH_multi(bh_object *o)
{
  if (o->fmode != 3) return fail(BH_EUNSUP, "many beams: the mixed-precision form");
  const int nb = o->nbm;
  zd ch = o->lambda0 / (2.0 * M_PI * zd(0.0, 1.0));
  HIPCHK(hipMemsetAsync(o->d_info, 0, sizeof(int), cur_st()));
  for (int d = 0; d <= o->MN; d++) {
    Dom &D = o->dom[d];
    const size_t nd = (size_t)D.nd, us = (size_t)(D.dNe + 1) * 16, es = (size_t)(D.dNe + 1) * 12;
    cd *PU = o->mb_ws[5].get<cd>(2 * nd * 6 * nb), *PdU = PU ? PU + 4 * nd * nb : nullptr;
    cd *Z = o->mb_ws[6].get<cd>(2 * nd * 4 * nb);
    if (!PU || !Z) return fail(BH_EHIP, "many beams: boundary-field work areas");
    bem_launch_meh_gather(D.nd, nb, D.d_col, D.d_colS, mb_dom_ptr(o, d, 0), mb_dom_ptr(o, d, 1), us, PU, PdU, cur_st());
    bem_zgemm(2 * D.nd, 4 * nb, D.nd, 1.0, 0.0, D.d_dG, nd, PdU, 4 * nb, 0.0, 0.0, Z, 4 * nb, cur_st());
    bem_zgemm(2 * D.nd, 4 * nb, D.nd, -1.0, 0.0, D.d_dH, nd, PU, 4 * nb, 1.0, 0.0, Z, 4 * nb, cur_st());
    bem_launch_meh_rows(D.nd, nb, Z, D.d_dF, D.d_rowEH, mb_dom_ptr(o, d, 0), mb_dom_ptr(o, d, 1), us, ch.real(),
                        ch.imag(), mb_dom_ptr(o, d, 2), mb_dom_ptr(o, d, 4), es, cur_st());
    // G [dE|dH] = H [E|H], 6 nb columns
    const cf *Gc = D.c_Gf;
    const int *Gpiv = D.d_Gpiv;
    if (d == 0) { Gc = o->c_G0f; Gpiv = o->d_piv0; }
    else if (!D.gfc_valid) {
      if (!D.c_Gf && dalloc(&D.c_Gf, nd * nd)) return BH_EHIP;
      if (!D.d_Gpiv && dalloc(&D.d_Gpiv, nd)) return BH_EHIP;
      bem_cvt_z2c(D.d_G, D.c_Gf, nd * nd, cur_st());
      if (bem_cgetrf(D.nd, D.c_Gf, nd, D.d_Gpiv, o->d_info, cur_st())) return lu_failed();
      D.gfc_valid = true;
      D.gf_valid = false;
      Gc = D.c_Gf;
      Gpiv = D.d_Gpiv;
    }
    cd *P = PU, *Rm = Z;   // reused: [nd x 6nb] each
    bem_launch_mpack_eh(D.nd, nb, D.d_col, D.d_colS, mb_dom_ptr(o, d, 2), mb_dom_ptr(o, d, 4), es, P, cur_st());
    bem_zgemm(D.nd, 6 * nb, D.nd, 1.0, 0.0, D.d_H, nd, P, 6 * nb, 0.0, 0.0, Rm, 6 * nb, cur_st());
    int rc = solve_g_mixed(o, d, Gc, Gpiv, Rm, 6 * nb);
    if (rc) return rc;
    bem_launch_munpack_eh(D.nd, nb, D.d_col, D.d_colS, Rm, mb_dom_ptr(o, d, 3), mb_dom_ptr(o, d, 5), es, cur_st());
  }
  HIPCHK(hipGetLastError());
  int info = 0;
  HIPCHK(hipMemcpyAsync(&info, o->d_info, sizeof(int), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  if (info) return fail(BH_ESING, "singular G (zero pivot %d)", info);
  return 0;
}

// one beam's boundary values of one domain (which as bh_object_state) or
// incident data (which 6 + 0..3: Ui, dUi, Ui0, dUi0) after bh_mo_solve_beams
extern "C" int bh_object_beam_state(bh_object *o, int b, int did, int which, double *buf)
{
  API_LOCK;
  if (b < 0 || b >= o->nbm) return fail(BH_EARG, "beam %d of %d", b, o->nbm);
  if (which >= 6 && which < 10) {
    const int map[4] = {2, 3, 0, 1};
    const size_t n = ((size_t)o->Ne + 1) * 16;
    HIPCHK(hipMemcpy(buf, mb_ptr(o, map[which - 6]) + b * n, n * sizeof(cd), hipMemcpyDeviceToHost));
    return 0;
  }
  if (did < 0 || did > o->MN || which < 0 || which > 5) return fail(BH_EARG, "bad state selector");
  const size_t n = (size_t)(o->dom[did].dNe + 1) * (which < 2 ? 16 : 12);
  HIPCHK(hipMemcpy(buf, mb_dom_ptr(o, did, which) + b * n, n * sizeof(cd), hipMemcpyDeviceToHost));
  return 0;
}

// beam b of the many-beam solve as the object's one-beam state: its
// boundary values and incident data copied into the object's own arrays and
// its plane wave made the object's beam set, so that the field, force and
// power entry points report beam b (what one reference run with that beam
// alone would hold after mo_solve_iter, mo_solve.c:10-88)
int bem::object_select_beam(bh_object *o, int b)
{
  if (b < 0 || b >= o->nbm || b >= (int)o->mb_hbeams.size()) return fail(BH_EARG, "beam %d of %d", b, o->nbm);
  if (prepare_solve(o)) return BH_EHIP;
  const size_t E1 = (size_t)o->Ne + 1, ni = E1 * 16;
  cd *inc[5] = {o->d_Ui0, o->d_dUi0, o->d_Ui, o->d_dUi, o->d_Uip};
  for (int q = 0; q < 5; q++)
    HIPCHK(hipMemcpyAsync(inc[q], mb_ptr(o, q) + b * ni, ni * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
  for (int d = 0; d <= o->MN; d++) {
    Dom &D = o->dom[d];
    cd *tab[6] = {D.d_U, D.d_dU, D.d_E, D.d_dE, D.d_Hf, D.d_dHf};
    for (int q = 0; q < 6; q++) {
      const size_t n = (size_t)(D.dNe + 1) * (q < 2 ? 16 : 12);
      HIPCHK(hipMemcpyAsync(tab[q], mb_dom_ptr(o, d, q) + b * n, n * sizeof(cd), hipMemcpyDeviceToDevice, cur_st()));
    }
  }
  o->beams.assign(1, o->mb_hbeams[b]);
  o->fbeams.clear();
  dfree(o->d_fb); dfree(o->d_fbt);
  o->d_fb = nullptr; o->d_fbt = nullptr;
  if (!o->d_beams && upload(&o->d_beams, &o->beams[0], 1)) return BH_EHIP;
  HIPCHK(hipMemcpyAsync(o->d_beams, &o->beams[0], sizeof(Beam), hipMemcpyHostToDevice, cur_st()));
  o->x_warm = false;   // the one-beam solve's X belongs to another right-hand side
  o->r_valid = false;
  HIPCHK(hipStreamSynchronize(cur_st()));   // the host record stays alive until here
  return 0;
}

extern "C" int bh_object_select_beam(bh_object *o, int b)
{
  API_LOCK;
  return object_select_beam(o, b);
}

// which: 0 U, 1 dU ([dNe+1][4][4]), 2 E, 3 dE, 4 H, 5 dH ([dNe+1][4][3]);
// set != 0 uploads (e.g. a stored solution), else downloads
extern "C" int bh_object_state(bh_object *o, int did, int which, double *buf, int set)
{
  API_LOCK;
  if (did < 0 || did > o->MN || which < 0 || which > 5) return fail(BH_EARG, "bad state selector");
  if (prepare_solve(o)) return BH_EHIP;
  Dom &D = o->dom[did];
  cd *tab[6] = {D.d_U, D.d_dU, D.d_E, D.d_dE, D.d_Hf, D.d_dHf};
  size_t n = (size_t)(D.dNe + 1) * (which < 2 ? 16 : 12) * sizeof(cd);
  if (set) HIPCHK(hipMemcpy(tab[which], buf, n, hipMemcpyHostToDevice));
  else HIPCHK(hipMemcpy(buf, tab[which], n, hipMemcpyDeviceToHost));
  return 0;
}

// which: 0 Ui, 1 dUi, 2 Ui0, 3 dUi0 ([Ne+1][4][4])
extern "C" int bh_object_incident(bh_object *o, int which, double *buf, int set)
{
  API_LOCK;
  if (prepare_solve(o)) return BH_EHIP;
  cd *tab[4] = {o->d_Ui, o->d_dUi, o->d_Ui0, o->d_dUi0};
  if (which < 0 || which > 3) return fail(BH_EARG, "bad incident selector");
  size_t n = (size_t)(o->Ne + 1) * 16 * sizeof(cd);
  if (set) HIPCHK(hipMemcpy(tab[which], buf, n, hipMemcpyHostToDevice));
  else HIPCHK(hipMemcpy(buf, tab[which], n, hipMemcpyDeviceToHost));
  return 0;
}

// ------------------------------------------------------------------ multi-object layer

int bem::domain_ids(bh_object *o, const double *d_pts, int np, int *d_did)
{
  if (prepare_solve(o)) return BH_EHIP;
  bem_launch_domain(d_pts, np, o->d_fv, o->d_fsd, o->d_fq, o->d_seg, o->MN, d_did, cur_st());
  HIPCHK(hipGetLastError());
  return 0;
}

// grow-only device work areas of the field entry points (calls are
// serialised; bh_finalize releases them).  Slots 0-15: bh_mo_field,
// bh_mo_EH_bd, bh_object_domain and the surface points of the force and
// power use separate slots where they nest.
#define BEM_NCWS 24
static Scratch *g_cws = nullptr;
template <class T> static T *cws(int i, size_t n)
{
  if (!g_cws) g_cws = new Scratch[BEM_NCWS];
  return g_cws[i].get<T>(std::max<size_t>(n, 1));
}
// a grow-only work area holding n host values (copied on the library's stream)
template <class T> static int cput(int i, T **d, const T *h, size_t n)
{
  *d = cws<T>(i, n);
  if (!*d) return fail(BH_EHIP, "device work area of %zu bytes", n * sizeof(T));
  if (n) HIPCHK(hipMemcpyAsync(*d, h, n * sizeof(T), hipMemcpyHostToDevice, cur_st()));
  return 0;
}
#define CWS_CHECK(p) do { if (!(p)) return fail(BH_EHIP, "device work area"); } while (0)

// The coupling loop's solves of N objects of the mixed-precision form (S
// formed, matrices stored) in lock step on the library's stream: the
// right-hand sides of all objects, then refinement steps of every object
// whose relative residual is above IR_TOL, each step one schur_apply_batch
// (the triangular solves of all objects together, one launch per factor
// and triangle) and the objects' FP64 residuals, with one read-back of all
// norms per step.  A repeated solve starts from the previous solution with
// the residual updated by the change of the right-hand side (R = B - A X
// becomes R - B_old + B_new: no matrix pass for the warm start).  Returns
// BH_EUNSUP (nothing done) if an object does not qualify.
// one product of the batched residuals (layout of SkJob, bem_la.hip):
// C (m x 4) = A (m x k) B (k x 4) + beta C
struct SkJobH {
  const cd *A;
  size_t lda;
  const cd *B;
  cd *C;
  int m, k;
  double beta;
};
// device job tables of the batched residuals (entry points are serialised;
// released by bh_finalize, never by a destructor at process exit)
static Scratch *g_skjob = nullptr;

int bem::objects_solve_batch(bh_object **objs, int N)
{
  if (N <= 0) return 0;
  for (int m = 0; m < N; m++) {
    const bh_object *o = objs[m];
    if (o->fmode != 3 || !o->s_formed || o->prec == BH_PREC_MIXED_LM || o->MN != 1 || !o->d_beams)
      return BH_EUNSUP;
  }
  hipStream_t st = cur_st();
  double *dn = cws<double>(22, 2 * (size_t)N);
  CWS_CHECK(dn);
  std::vector<double> hn(2 * (size_t)N), bn(N), res(N);
  // right-hand sides, and the residuals of the starting solutions
  for (int m = 0; m < N; m++) {
    bh_object *o = objs[m];
    const size_t na = 8 * (size_t)o->NN;
    if (o->r_valid) bem_launch_axpby(o->d_R, -1.0, 0.0, o->d_B, o->d_R, na, st);   // R = -A X
    HIPCHK(hipMemsetAsync(o->d_B, 0, na * sizeof(cd), st));
    for (int d = 0; d <= o->MN; d++) {
      int rc = dom_build_B(o, d);
      if (rc) return rc;
    }
    if (o->r_valid) {
      bem_launch_axpby(o->d_R, 1.0, 0.0, o->d_B, o->d_R, na, st);
    } else {
      HIPCHK(hipMemsetAsync(o->d_X, 0, na * sizeof(cd), st));
      HIPCHK(hipMemcpyAsync(o->d_R, o->d_B, na * sizeof(cd), hipMemcpyDeviceToDevice, st));
    }
    o->r_valid = false;
    o->x_warm = false;
    o->ir_iters = 0;
    o->gm_steps = 0;
    bem_nrm2sq_to(o->d_B, na, o->d_nrm, dn + 2 * m, st);
    bem_nrm2sq_to(o->d_R, na, o->d_nrm, dn + 2 * m + 1, st);
  }
  HIPCHK(hipMemcpyAsync(hn.data(), dn, hn.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<int> act;
  for (int m = 0; m < N; m++) {
    bn[m] = sqrt(hn[2 * m]);
    res[m] = bn[m] > 0.0 ? sqrt(hn[2 * m + 1]) / bn[m] : 0.0;
    if (res[m] > IR_TOL) act.push_back(m);
  }
  std::vector<SchurF<cf>> f;
  std::vector<const cf *> rhs;
  std::vector<cf *> dx, w;
  for (int it = 0; !act.empty() && it < IR_MAX; it++) {
    const int na_ = (int)act.size();
    f.resize(na_); rhs.resize(na_); dx.resize(na_); w.resize(na_);
    for (int q = 0; q < na_; q++) {
      bh_object *o = objs[act[q]];
      const size_t NN = (size_t)o->NN;
      cf *cw = o->c_w, *crhs = cw + schur_work(NN), *cdx = crhs + 8 * NN;
      bem_cvt_z2c(o->d_R, crhs, 8 * NN, st);
      f[q] = schur32(o);
      rhs[q] = crhs; dx[q] = cdx; w[q] = cw;
    }
    int rc = schur_apply_batch(na_, f.data(), rhs.data(), dx.data(), w.data(), st);
    if (rc) return fail(rc, "batched block solve of %d objects", na_);
    // R = B - A X of every object and domain: the column vectors of X, then
    // the H and the G products of all of them as one launch each
    std::vector<SkJobH> jh, jg;
    std::vector<cd *> Ys;
    for (int q = 0; q < na_; q++) {
      bh_object *o = objs[act[q]];
      bem_acc_c2z(o->d_X, dx[q], 8 * (size_t)o->NN, st);
      size_t rows = 0;
      for (int d = 0; d <= o->MN; d++) rows += (size_t)o->dom[d].nd;
      cd *wr = o->ws_res.get<cd>(12 * rows);
      if (!wr) return fail(BH_EHIP, "refinement residual: work area of %zu rows", rows);
      size_t off = 0;
      for (int d = 0; d <= o->MN; d++) {
        Dom &D = o->dom[d];
        cd *a = wr + off, *c = a + 4 * (size_t)D.nd, *Y = a + 8 * (size_t)D.nd;
        off += 12 * (size_t)D.nd;
        bem_resid_colvec(D.d_col, D.nd, o->d_X, (size_t)o->NN, a, st);
        jh.push_back({D.d_H, (size_t)D.nd, a, Y, D.nd, D.nd, 0.0});
        jg.push_back({D.d_G, (size_t)D.nd, c, Y, D.nd, D.nd, 1.0});
        Ys.push_back(Y);
      }
    }
    {
      const size_t tb = (bem_skinny_batch_bytes((int)jh.size()) + 255) / 256 * 256;
      if (!g_skjob) g_skjob = new Scratch;
      char *dtab = g_skjob->get<char>(2 * tb);
      std::vector<char> h1(tb), h2(tb);
      if (!dtab || bem_zgemm_skinny4_batch(jh.data(), (int)jh.size(), dtab, h1.data(), st) ||
          bem_zgemm_skinny4_batch(jg.data(), (int)jg.size(), dtab + tb, h2.data(), st))
        return fail(BH_EHIP, "batched refinement residuals of %d objects", na_);
      size_t k = 0;
      for (int q = 0; q < na_; q++) {
        const int m = act[q];
        bh_object *o = objs[m];
        for (int d = 0; d <= o->MN; d++, k++) bem_resid_fin(o->d_B, Ys[k], o->dom[d].nd, row_base(o, d), o->d_R, st);
        bem_nrm2sq_to(o->d_R, 8 * (size_t)o->NN, o->d_nrm, dn + 2 * m + 1, st);
        o->ir_iters++;
      }
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(st));   // the host job tables stay alive until here
    }
    HIPCHK(hipMemcpyAsync(hn.data(), dn, hn.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int> next;
    for (int m : act) {
      res[m] = sqrt(hn[2 * m + 1]) / bn[m];
      if (res[m] > IR_TOL) next.push_back(m);
    }
    act.swap(next);
  }
  for (int m = 0; m < N; m++) {
    bh_object *o = objs[m];
    o->ir_res = res[m];
    if (!(res[m] <= IR_TOL))
      return fail(BH_ECONV, "mixed-precision refinement (object %d of %d): relative residual %.3g after %d steps", m, N,
                  res[m], o->ir_iters);
    const size_t NN = (size_t)o->NN;
    for (int d = 0; d <= o->MN; d++) {
      Dom &D = o->dom[d];
      bem_launch_store_X(o->d_X, D.nd, D.d_col, NN, D.d_colS, o->d_Ui, o->d_dUi, D.d_U, D.d_dU, st);
    }
    o->x_warm = true;
    o->r_valid = true;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return trsv_failed(st);
}

// fields at points (mo_EH_s/t/i, mo_mEMP_s/t/i; mo_field.c:11-282):
// what 0 E,H (6 complex per point), 1 U (4); mode 0 s, 1 t, 2 i;
// type 0/1/2/3 integration rule.  oid[p] = object containing the point or -1.
extern "C" int bh_mo_field(bh_object **objs, int N, int what, int mode, int type, int np, const double *pts, double *out,
                           int *oid)
{
  API_LOCK;
  if (N <= 0 || np < 0 || type < 0 || type > 3 || what < 0 || what > 2) return fail(BH_EARG, "bad field request");
  if (np == 0) return 0;
  if (ensure_rules()) return BH_EHIP;
  for (int m = 0; m < N; m++) {
    if (!objs[m]->d_beams) return fail(BH_ESTATE, "object %d has no beams", m);
    if (mode != 2 && !objs[m]->dom[0].d_U) return fail(BH_ESTATE, "object %d is not solved", m);
  }
  int nc = what == 1 ? 4 : 6;
  double *d_pts;
  cd *d_out;
  int *d_did, *d_sel, *d_mask;
  if (cput(0, &d_pts, pts, 3 * (size_t)np)) return BH_EHIP;
  d_out = cws<cd>(1, (size_t)np * nc);
  d_did = cws<int>(2, (size_t)np * N);
  d_sel = cws<int>(3, np);
  d_mask = cws<int>(4, np);
  CWS_CHECK(d_out && d_did && d_sel && d_mask);
  HIPCHK(hipMemsetAsync(d_out, 0, (size_t)np * nc * sizeof(cd), cur_st()));
  for (int m = 0; m < N; m++)
    if (domain_ids(objs[m], d_pts, np, d_did + (size_t)m * np)) return BH_EHIP;
  std::vector<int> did((size_t)np * N), sel(np), mask(np);
  HIPCHK(hipMemcpyAsync(did.data(), d_did, did.size() * sizeof(int), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  // object containing each point: the first one whose domain is not open
  std::vector<int> pin(np, -1), pdom(np, 0);
  for (int p = 0; p < np; p++)
    for (int m = 0; m < N; m++)
      if (did[(size_t)m * np + p] != 0) { pin[p] = m; pdom[p] = did[(size_t)m * np + p]; break; }
  if (mode != 2) {
    for (int m = 0; m < N; m++) {
      bh_object *o = objs[m];
      for (int p = 0; p < np; p++) sel[p] = pin[p] < 0 ? 0 : (pin[p] == m ? pdom[p] : -2);
      HIPCHK(hipMemcpyAsync(d_sel, sel.data(), np * sizeof(int), hipMemcpyHostToDevice, cur_st()));
      for (int d = 0; d <= o->MN; d++) {
        Dom &D = o->dom[d];
        if (what == 2) {
          // E, H from the potentials' gradients (mo_EH_mEMP_s/t, mo_field.c:154-207)
          bem_launch_field_memp(d_pts, d_sel, np, d, o->d_geo, o->d_gq, D.d_sid, D.dNe, o->kn[d].real(),
                                o->kn[d].imag(), g_rule[type], g_nrule[type], g_rscale[type], type < 3, g_rule[3],
                                g_nrule[3], g_rscale[3], g_qrule[type], g_nqrule[type], g_qrule[3], g_nqrule[3], D.d_U,
                                D.d_dU, o->lambda0, d_out, cur_st());
          continue;
        }
        const void *A0 = what == 0 ? (const void *)D.d_E : (const void *)D.d_U;
        const void *dA0 = what == 0 ? (const void *)D.d_dE : (const void *)D.d_dU;
        bem_launch_field(d_pts, d_sel, np, d, o->d_geo, D.d_sid, D.dNe, o->kn[d].real(), o->kn[d].imag(), g_rule[type],
                         g_nrule[type], nullptr, g_rscale[type], o->d_gq, g_qrule[type], g_nqrule[type], what, A0, dA0, D.d_Hf,
                         D.d_dHf, d_out, 1, cur_st());
      }
      HIPCHK(hipStreamSynchronize(cur_st()));
    }
  }
  // incident field: everywhere for mode i, open-region points for mode t
  if (mode != 0) {
    for (int p = 0; p < np; p++) mask[p] = mode == 2 ? 1 : pin[p] < 0;
    HIPCHK(hipMemcpyAsync(d_mask, mask.data(), np * sizeof(int), hipMemcpyHostToDevice, cur_st()));
    add_incident(objs[0], d_pts, d_mask, np, d_out, what == 1);
  }
  HIPCHK(hipGetLastError());
  if (mode != 0 && fbeam_failed()) return BH_ECONV;
  HIPCHK(hipMemcpyAsync(out, d_out, (size_t)np * nc * sizeof(cd), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  if (oid) memcpy(oid, pin.data(), np * sizeof(int));
  return 0;
}

// surface points of the force / power rules of object oid (domain-0 faces):
// triangles type 0: 3 node slots (stored fields) + centroid (boundary
// integral), type 1: 7 interior points (boundary integral); quadrangles
// type 0: the 4 node slots (stored), type 1: the 3x3 Gauss points
// (mo_force.c:145-240, mo_absorb.c:9-162).  A quadrangle's per-element
// factor is twice the triangle's (force -1 vs -1/2, power -1/2 vs -1/4):
// its weights carry the 2.
struct SurfPts {
  std::vector<int> t, stored;      // element (domain-local), stored slot or -1
  std::vector<double> ze, r, w, wt;
};

static void surface_points(const bh_object *o, int type, SurfPts &sp)
{
  const Dom &D = o->dom[0];
  const double P0 = 0.2, P1 = 0.3464101615137754587055, W0 = 0.5208333333333333333333, W1 = 0.5625;
  const double a0 = 0.3480702390148154917985, a1 = 0.2052130961576726346557, a2 = 0.6028753385763033130539,
               a3 = 0.3554395089236065568357, w0 = 0.1259391805448271525957, w1 = 0.1323941527885061807376,
               w2 = 0.225;
  const double z4[4] = {-P0, 2 * P0, -P0, 0.0}, e4[4] = {-P1, 0.0, P1, 0.0}, ww4[4] = {W0, W0, W0, -W1};
  const double z7[7] = {-a0, 2 * a0, -a0, a1, -2 * a1, a1, 0.0}, e7[7] = {-a2, 0.0, a2, a3, 0.0, -a3, 0.0};
  const double ww7[7] = {w0, w0, w0, w1, w1, w1, w2};
  const double p9 = 0.7745966692414833770359, v0 = 0.3086419753086419753086, v1 = 0.4938271604938271604938,
               v2 = 0.7901234567901234567901;
  const double z9[9] = {-p9, p9, p9, -p9, 0.0, p9, 0.0, -p9, 0.0}, e9[9] = {-p9, -p9, p9, p9, -p9, 0.0, p9, 0.0, 0.0};
  const double ww9[9] = {v0, v0, v0, v0, v1, v1, v1, v1, v2};
  for (int t = 1; t <= D.dNe; t++) {
    int a = std::abs(D.sid[t]);
    const double *cr = &o->cr[(size_t)a * 12], *cw = &o->cw[(size_t)a * 9];
    if (o->ed[4 * a + 3] != 0) {
      const int np = type == 0 ? 4 : 9;
      for (int i = 0; i < np; i++) {
        const double z = type == 0 ? ZT44[i] : z9[i], e = type == 0 ? ET44[i] : e9[i];
        const bool st = type == 0;
        sp.t.push_back(t);
        sp.stored.push_back(st ? i : -1);
        sp.ze.push_back(z); sp.ze.push_back(e);
        for (int d = 0; d < 3; d++) {
          sp.r.push_back(st ? o->ren[(size_t)a * 12 + i * 3 + d]
                            : cr[d * 4] + cr[d * 4 + 1] * z + cr[d * 4 + 2] * e + cr[d * 4 + 3] * z * e);
          sp.w.push_back(st ? o->wen[(size_t)a * 12 + i * 3 + d] : cw[d * 3] + cw[d * 3 + 1] * z + cw[d * 3 + 2] * e);
        }
        sp.wt.push_back(2.0 * (type == 0 ? 1.0 : ww9[i]));
      }
      continue;
    }
    int np = type == 0 ? 4 : 7;
    for (int i = 0; i < np; i++) {
      double z = type == 0 ? z4[i] : z7[i], e = type == 0 ? e4[i] : e7[i];
      bool st = type == 0 && i < 3;
      sp.t.push_back(t);
      sp.stored.push_back(st ? i : -1);
      sp.ze.push_back(z); sp.ze.push_back(e);
      for (int d = 0; d < 3; d++) {
        sp.r.push_back(st ? o->ren[(size_t)a * 12 + i * 3 + d] : cr[d * 4] + cr[d * 4 + 1] * z + cr[d * 4 + 2] * e);
        sp.w.push_back(st ? o->wen[(size_t)a * 12 + i * 3 + d] : cw[d * 3]);
      }
      sp.wt.push_back(type == 0 ? ww4[i] : ww7[i]);
    }
  }
}

// out[idx[i]] = in[i] (6 complex values per point)
__global__ void k_scatter6(const cd *__restrict__ in, const int *__restrict__ idx, int n, cd *__restrict__ out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < 6; c++) out[6 * (size_t)idx[i] + c] = in[6 * (size_t)i + c];
}

__global__ void k_gather_stored(const int *__restrict__ t, const int *__restrict__ stored, int np,
                                const cd *__restrict__ E, const cd *__restrict__ H, cd *__restrict__ out)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np || stored[p] < 0) return;
  size_t o = ((size_t)t[p] * 4 + stored[p]) * 3;
  for (int i = 0; i < 3; i++) { out[6 * (size_t)p + i] = E[o + i]; out[6 * (size_t)p + 3 + i] = H[o + i]; }
}

// total field on the surface points: self boundary integral (or stored node
// value) + other objects + incident (mo_EH_t_bd / mo_EH_i_bd)
static int surface_fields(bh_object **objs, int N, int oid, const SurfPts &sp, cd **d_EH_out, double **d_r_out,
                          double **d_w_out, double **d_wt_out)
{
  bh_object *o = objs[oid];
  Dom &D = o->dom[0];
  int np = (int)sp.t.size();
  int *d_t, *d_st, *d_did, *d_sel, *d_mask;
  double *d_ze, *d_r, *d_w, *d_wt, *d_rt;
  cd *d_EH, *d_tmp;
  if (cput(0, &d_t, sp.t.data(), np) || cput(1, &d_st, sp.stored.data(), np) || cput(2, &d_ze, sp.ze.data(), 2 * (size_t)np) ||
      cput(3, &d_r, sp.r.data(), 3 * (size_t)np) || cput(4, &d_w, sp.w.data(), 3 * (size_t)np) ||
      cput(5, &d_wt, sp.wt.data(), np))
    return BH_EHIP;
  d_EH = cws<cd>(6, 6 * (size_t)np);
  d_tmp = cws<cd>(7, 6 * (size_t)np);
  d_rt = cws<double>(8, 3 * (size_t)np);
  d_did = cws<int>(9, np);
  d_sel = cws<int>(10, np);
  d_mask = cws<int>(11, np);
  CWS_CHECK(d_EH && d_tmp && d_rt && d_did && d_sel && d_mask);
  HIPCHK(hipMemsetAsync(d_EH, 0, 6 * (size_t)np * sizeof(cd), cur_st()));
  // self object: the boundary integral (EH_s_bd) only at the points that are
  // not node slots, the node slots from the stored boundary values
  {
    std::vector<int> ci, ct;
    std::vector<double> cze;
    for (int p = 0; p < np; p++)
      if (sp.stored[p] < 0) {
        ci.push_back(p);
        ct.push_back(sp.t[p]);
        cze.push_back(sp.ze[2 * p]);
        cze.push_back(sp.ze[2 * p + 1]);
      }
    int nc = (int)ci.size();
    if (nc > 0) {
      int *d_ci, *d_ct;
      double *d_cze;
      if (cput(12, &d_ci, ci.data(), nc) || cput(13, &d_ct, ct.data(), nc) || cput(14, &d_cze, cze.data(), cze.size()))
        return BH_EHIP;
      bem_launch_field_bd(d_ct, d_cze, nc, 0, o->d_geo, o->d_gq, D.d_sid, D.dNe, o->kn[0].real(), o->kn[0].imag(),
                          g_rule[2], o->prec != BH_PREC_FP64 ? g.d_gl32 : nullptr, g_nrule[2], g_rscale[2], g_qrule[2],
                          g_nqrule[2], g.d_xl, g.d_wl, g.d_xh, g.d_wh, D.d_E, D.d_dE, D.d_Hf, D.d_dHf,
                          d_tmp, d_rt, cur_st());
      hipLaunchKernelGGL(k_scatter6, g1(nc), dim3(256), 0, cur_st(), d_tmp, d_ci, nc, d_EH);
      HIPCHK(hipStreamSynchronize(cur_st()));   // the host tables above stay alive until here
    }
  }
  hipLaunchKernelGGL(k_gather_stored, g1(np), dim3(256), 0, cur_st(), d_t, d_st, np, D.d_E, D.d_Hf, d_EH);
  // other objects at the point r (their open domain, checked by solid angle)
  for (int m = 0; m < N; m++) {
    if (m == oid) continue;
    bh_object *q = objs[m];
    if (domain_ids(q, d_r, np, d_did)) return BH_EHIP;
    for (int d = 0; d <= q->MN; d++) {
      Dom &Q = q->dom[d];
      // the mixed form: far triangles in FP32 with the reduced far rules, as
      // in the object's own boundary integral
      bem_launch_field(d_r, d_did, np, d, q->d_geo, Q.d_sid, Q.dNe, q->kn[d].real(), q->kn[d].imag(), g_rule[2],
                       g_nrule[2], q->prec != BH_PREC_FP64 ? g.d_gl32 : nullptr, g_rscale[2], q->d_gq, g_qrule[2],
                       g_nqrule[2], 0, Q.d_E, Q.d_dE, Q.d_Hf, Q.d_dHf, d_EH, 1, cur_st());
    }
  }
  {
    std::vector<int> ones(np, 1);
    HIPCHK(hipMemcpyAsync(d_mask, ones.data(), np * sizeof(int), hipMemcpyHostToDevice, cur_st()));
  }
  add_incident(objs[0], d_r, d_mask, np, d_EH, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(cur_st()));
  if (fbeam_failed()) return BH_ECONV;
  *d_EH_out = d_EH; *d_r_out = d_r; *d_w_out = d_w; *d_wt_out = d_wt;
  return 0;
}

static int check_ready(bh_object **objs, int N, int oid)
{
  if (N <= 0 || oid < 0 || oid >= N) return fail(BH_EARG, "bad object id %d", oid);
  if (ensure_rules()) return BH_EHIP;
  for (int m = 0; m < N; m++)
    if (!objs[m]->dom[0].d_U || !objs[m]->d_beams) return fail(BH_ESTATE, "object %d is not solved", m);
  return 0;
}

// radiation force and torque about rc (mo_force_FN, mo_force.c:10-240);
// returns 0, or -1 when the object has no open-domain faces
static int mo_force(bh_object **objs, int N, int oid, int type, const double *rc, double *F, double *Nq)
{
  int rc0 = check_ready(objs, N, oid);
  if (rc0) return rc0;
  bh_object *o = objs[oid];
  SurfPts sp;
  surface_points(o, type, sp);
  int np = (int)sp.t.size();
  for (int j = 0; j < 3; j++) { F[j] = 0.0; Nq[j] = 0.0; }
  if (np == 0) return -1;
  cd *d_EH;
  double *d_r, *d_w, *d_wt, *d_part, *d_rc;
  int rc1 = surface_fields(objs, N, oid, sp, &d_EH, &d_r, &d_w, &d_wt);
  if (rc1) return rc1;
  d_part = cws<double>(15, 6 * (size_t)np);
  CWS_CHECK(d_part);
  if (cput(16, &d_rc, rc, 3)) return BH_EHIP;
  double epd = o->n[0].real() * o->n[0].real();
  bem_launch_stress(d_EH, d_r, d_w, d_wt, np, epd, d_rc, d_part, cur_st());
  std::vector<double> part(6 * (size_t)np);
  HIPCHK(hipMemcpyAsync(part.data(), d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  for (int p = 0; p < np; p++)
    for (int j = 0; j < 3; j++) { F[j] += -0.5 * part[6 * p + j]; Nq[j] += -0.5 * part[6 * p + 3 + j]; }
  return 0;
}

extern "C" int bh_mo_force(bh_object **objs, int N, int oid, int type, const double *rc, double *F, double *Nq)
{
  API_LOCK;
  return mo_force(objs, N, oid, type, rc, F, Nq);
}

// E, H of nb right-hand sides at n points from one object's domain-0
// boundary values of the many-beam solve, added into EH[b][p][6] (beam b
// at + b ldb; point p = idx[i], or i for idx null): the rows of the E, H
// integral at the points (k_field_rows; on the object itself when d_t is
// given) times the packed boundary values of every beam, in chunks of
// points whose rows fit in about 4 GB
static int rows_field(bh_object *src, const double *d_pts, const int *d_t, const double *d_ze, int n, const int *d_idx,
                      int nb, cd *EH, size_t ldb)
{
  Dom &D = src->dom[0];
  const size_t K = 8 * ((size_t)D.dNe + 1), ncol = 6 * (size_t)nb;
  const int pc = (int)std::max<size_t>(4, std::min<size_t>((size_t)n, ((size_t)256 << 20) / K));
  cd *X = cws<cd>(18, K * ncol), *R = cws<cd>(19, (size_t)pc * K), *Cm = cws<cd>(20, (size_t)pc * ncol);
  double *Fv = cws<double>(21, pc);
  CWS_CHECK(X && R && Cm && Fv);
  const size_t es = (size_t)(D.dNe + 1) * 12;
  bem_launch_pack_ehbv(mb_dom_ptr(src, 0, 2), mb_dom_ptr(src, 0, 3), mb_dom_ptr(src, 0, 4), mb_dom_ptr(src, 0, 5), es,
                       D.dNe, nb, X, cur_st());
  for (int p0 = 0; p0 < n; p0 += pc) {
    const int m = std::min(pc, n - p0);
    bem_launch_field_rows(d_pts ? d_pts + 3 * (size_t)p0 : nullptr, d_t ? d_t + p0 : nullptr,
                       d_ze ? d_ze + 2 * (size_t)p0 : nullptr, m, 0, src->d_geo, src->d_gq, D.d_sid, D.dNe,
                       src->kn[0].real(), src->kn[0].imag(), g_rule[2], g_nrule[2], g.d_glr, g_rscale[2], g_qrule[2],
                       g_nqrule[2], g.d_xl, g.d_wl, g.d_xh, g.d_wh, K, R, Fv, cur_st());
    bem_zgemm(m, (int)ncol, (int)K, 1.0, 0.0, R, K, X, ncol, 0.0, 0.0, Cm, ncol, cur_st());
    bem_launch_rows_add(Cm, m, nb, Fv, 1, d_idx ? d_idx + p0 : nullptr, ldb, d_idx ? EH : EH + 6 * (size_t)p0, cur_st());
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// force and torque of object oid under all nb beams (mo_force_FN with the
// total field of mo_EH_t_bd, mo_force.c:10-240, mo_field.c:227-254, for
// every beam at once): the stored node values of every beam (rule 0), the
// self boundary integral at the other surface points and the fields of the
// other objects as products of integral rows with the boundary values of
// all beams, each beam's incident wave, then the stress sums per beam
static int force_beams_one(bh_object **objs, int N, int oid, int type, const double *rc, double *F, double *Nq,
                           int *flag)
{
  bh_object *o = objs[oid];
  const int nb = o->nbm;
  Dom &D = o->dom[0];
  SurfPts sp;
  surface_points(o, type, sp);
  const int np = (int)sp.t.size();
  for (int b = 0; b < nb; b++) {
    for (int j = 0; j < 3; j++) { F[3 * (size_t)b * N + j] = 0.0; Nq[3 * (size_t)b * N + j] = 0.0; }
    if (flag) flag[(size_t)b * N] = np ? 0 : -1;
  }
  if (np == 0) return 0;
  int *d_t, *d_st, *d_did;
  double *d_ze, *d_r, *d_w, *d_wt, *d_rc;
  if (cput(0, &d_t, sp.t.data(), np) || cput(1, &d_st, sp.stored.data(), np) ||
      cput(2, &d_ze, sp.ze.data(), 2 * (size_t)np) || cput(3, &d_r, sp.r.data(), 3 * (size_t)np) ||
      cput(4, &d_w, sp.w.data(), 3 * (size_t)np) || cput(5, &d_wt, sp.wt.data(), np) || cput(16, &d_rc, rc, 3))
    return BH_EHIP;
  const size_t ldb = (size_t)np * 6;
  cd *EH = cws<cd>(17, ldb * nb);
  d_did = cws<int>(9, np);
  double *part = cws<double>(15, 6 * (size_t)np * nb);
  CWS_CHECK(EH && d_did && part);
  HIPCHK(hipMemsetAsync(EH, 0, ldb * nb * sizeof(cd), cur_st()));
  // stored node values of every beam
  const size_t es = (size_t)(D.dNe + 1) * 12;
  for (int b = 0; b < nb; b++)
    hipLaunchKernelGGL(k_gather_stored, g1(np), dim3(256), 0, cur_st(), d_t, d_st, np,
                       (const cd *)(mb_dom_ptr(o, 0, 2) + b * es), (const cd *)(mb_dom_ptr(o, 0, 4) + b * es), EH + b * ldb);
  // self boundary integral at the other points
  std::vector<int> ci, ct;
  std::vector<double> cze;
  for (int p = 0; p < np; p++)
    if (sp.stored[p] < 0) {
      ci.push_back(p);
      ct.push_back(sp.t[p]);
      cze.push_back(sp.ze[2 * p]);
      cze.push_back(sp.ze[2 * p + 1]);
    }
  const int nc = (int)ci.size();
  if (nc > 0) {
    int *d_ci, *d_ct;
    double *d_cze;
    if (cput(12, &d_ci, ci.data(), nc) || cput(13, &d_ct, ct.data(), nc) || cput(14, &d_cze, cze.data(), cze.size()))
      return BH_EHIP;
    int r = rows_field(o, nullptr, d_ct, d_cze, nc, d_ci, nb, EH, ldb);
    if (r) return r;
  }
  // the other objects (every point in their open domain)
  std::vector<int> did(np);
  for (int m = 0; m < N; m++) {
    if (m == oid) continue;
    bh_object *q = objs[m];
    if (domain_ids(q, d_r, np, d_did)) return BH_EHIP;
    HIPCHK(hipMemcpyAsync(did.data(), d_did, np * sizeof(int), hipMemcpyDeviceToHost, cur_st()));
    HIPCHK(hipStreamSynchronize(cur_st()));
    for (int p = 0; p < np; p++)
      if (did[p] != 0) return fail(BH_EDOM, "force: a surface point of object %d lies inside object %d", oid, m);
    int r = rows_field(q, d_r, nullptr, nullptr, np, nullptr, nb, EH, ldb);
    if (r) return r;
  }
  // incident wave of each beam, stress sums
  const Beam *db = (const Beam *)objs[0]->mb_beams.p;
  const double epd = o->n[0].real() * o->n[0].real();
  int *d_ones;
  if (cput(11, &d_ones, std::vector<int>(np, 1).data(), np)) return BH_EHIP;
  for (int b = 0; b < nb; b++) {
    bem_launch_add_incident(d_r, d_ones, np, db + b, 1, EH + b * ldb, 0, cur_st());
    bem_launch_stress(EH + b * ldb, d_r, d_w, d_wt, np, epd, d_rc, part + 6 * (size_t)np * b, cur_st());
  }
  std::vector<double> hp(6 * (size_t)np * nb);
  HIPCHK(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  for (int b = 0; b < nb; b++) {
    double *Fb = F + 3 * (size_t)b * N, *Nb = Nq + 3 * (size_t)b * N;
    const double *pb = hp.data() + 6 * (size_t)np * b;
    for (int p = 0; p < np; p++)
      for (int j = 0; j < 3; j++) { Fb[j] += -0.5 * pb[6 * p + j]; Nb[j] += -0.5 * pb[6 * p + 3 + j]; }
  }
  return 0;
}

// force and torque of every object under every beam of the many-beam solve
// (mo_force_FN of each object in one reference run per beam): rc[3 m] the
// torque centre of object m (null: the origin); F, Nq: [nb][N][3];
// flag[b N + m] = -1 when object m has no open-domain faces, else 0
extern "C" int bh_mo_force_beams(bh_object **objs, int N, int type, const double *rc, double *F, double *Nq, int *flag)
{
  API_LOCK;
  if (N <= 0 || type < 0 || type > 1) return fail(BH_EARG, "bad force request");
  if (ensure_rules()) return BH_EHIP;
  const int nb = objs[0]->nbm;
  for (int m = 0; m < N; m++)
    if (objs[m]->nbm != nb || nb <= 0 || objs[m]->mb_dom.empty())
      return fail(BH_ESTATE, "object %d: no many-beam solve of %d beams", m, nb);
  const double zero[3] = {0.0, 0.0, 0.0};
  for (int m = 0; m < N; m++) {
    int r = force_beams_one(objs, N, m, type, rc ? rc + 3 * (size_t)m : zero, F + 3 * (size_t)m, Nq + 3 * (size_t)m,
                            flag ? flag + m : nullptr);
    if (r) return r;
  }
  return 0;
}

// element coefficients of the reference's element layer (coef_rt,
// coef_bd, coef_bd_node, dcoef_rt, dcoef_bd_tz/te, dcoef_bd_t2,
// dcoef_bd_node_t2; d3b1_elem.c:148-539) for n requests on object o's
// elements: nd 0 CC only, 1 with dC along v0, 2 along v0 and v1; rule type
// 0..3; k = kr + i ki.  Request i: signed source element s[i]; t (may be
// null) the signed target element: t[i] == s[i] takes the self element at
// local coordinates ze[2i], ze[2i+1] (polar rules, principal-value
// derivatives along the element's own tangents), otherwise the point
// rt[3i].  out: 27 complex per request (CC, dC0, dC1; 9 each: G 0-3, H
// 4-7, F 8).
extern "C" int bh_elem_coef(bh_object *o, int n, int nd, int type, double kr, double ki, const int *s, const int *t,
                            const double *ze, const double *rt, const double *v0, const double *v1, double *out)
{
  API_LOCK;
  if (n < 0 || nd < 0 || nd > 2 || type < 0 || type > 3) return fail(BH_EARG, "bad element-coefficient request");
  if (n == 0) return 0;
  if (ensure_rules()) return BH_EHIP;
  std::vector<int> so, io, ss, is;
  std::vector<double> ro, a0, a1, zs;
  for (int i = 0; i < n; i++) {
    const int as = std::abs(s[i]);
    if (as < 1 || as > o->Ne) return fail(BH_EARG, "element %d of %d", s[i], o->Ne);
    if (t && t[i] == s[i]) {
      ss.push_back(s[i]);
      is.push_back(i);
      zs.push_back(ze[2 * i]);
      zs.push_back(ze[2 * i + 1]);
      continue;
    }
    so.push_back(s[i]);
    io.push_back(i);
    for (int c = 0; c < 3; c++) {
      ro.push_back(rt[3 * i + c]);
      a0.push_back(nd > 0 ? v0[3 * i + c] : 0.0);
      a1.push_back(nd > 1 ? v1[3 * i + c] : 0.0);
    }
  }
  cd *d_out = cws<cd>(17, 27 * (size_t)n);
  CWS_CHECK(d_out);
  const int no = (int)so.size(), ns = (int)ss.size();
  if (no > 0) {
    int *d_s, *d_i;
    double *d_r, *d_a0, *d_a1;
    if (cput(0, &d_s, so.data(), no) || cput(1, &d_i, io.data(), no) || cput(2, &d_r, ro.data(), ro.size()) ||
        cput(3, &d_a0, a0.data(), a0.size()) || cput(4, &d_a1, a1.data(), a1.size()))
      return BH_EHIP;
    bem_launch_elem_coef(nd, o->d_geo, o->d_gq, no, d_s, d_r, d_a0, d_a1, d_i, kr, ki, g_rule[type], g_nrule[type],
                         g_rscale[type], g_qrule[type], g_nqrule[type], g_rule[3], g_nrule[3], g_rscale[3], g_qrule[3],
                         g_nqrule[3], type < 3, d_out, cur_st());
  }
  if (ns > 0) {
    int *d_s, *d_i;
    double *d_z, *d_tf;
    cd *d_t;
    if (cput(5, &d_s, ss.data(), ns) || cput(6, &d_i, is.data(), ns) || cput(7, &d_z, zs.data(), zs.size()))
      return BH_EHIP;
    d_t = cws<cd>(8, 24 * (size_t)ns);
    d_tf = cws<double>(9, 3 * (size_t)ns);
    CWS_CHECK(d_t && d_tf);
    bem_launch_elem_self(o->d_geo, o->d_gq, d_s, d_z, d_i, ns, kr, ki, g.d_xl, g.d_wl, g.d_xh, g.d_wh, d_t, d_tf, d_out,
                         cur_st());
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, d_out, 27 * (size_t)n * sizeof(cd), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  return 0;
}

// absorbed power (mo_absorb_P, mo_absorb.c:9-44): 0, -1 (no faces) or -2
// (negative flux, magnitude returned)
extern "C" int bh_mo_absorb(bh_object **objs, int N, int oid, int type, double *P)
{
  API_LOCK;
  int rc0 = check_ready(objs, N, oid);
  if (rc0) return rc0;
  SurfPts sp;
  surface_points(objs[oid], type, sp);
  int np = (int)sp.t.size();
  *P = 0.0;
  if (np == 0) return -1;
  cd *d_EH;
  double *d_r, *d_w, *d_wt, *d_part;
  int rc1 = surface_fields(objs, N, oid, sp, &d_EH, &d_r, &d_w, &d_wt);
  if (rc1) return rc1;
  d_part = cws<double>(15, np);
  CWS_CHECK(d_part);
  bem_launch_poynting(d_EH, d_w, d_wt, np, d_part, cur_st());
  std::vector<double> part(np);
  HIPCHK(hipMemcpyAsync(part.data(), d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  double p = 0.0;
  for (int i = 0; i < np; i++) p += -0.25 * part[i];
  *P = -p;
  if (*P < 0.0) { *P = -*P; return -2; }
  return 0;
}

// domain of np points w.r.t. one object (domain_id, setup2.c:944-978)
extern "C" int bh_object_domain(bh_object *o, int np, const double *pts, int *did)
{
  API_LOCK;
  if (np <= 0) return 0;
  double *d_pts;
  if (cput(0, &d_pts, pts, 3 * (size_t)np)) return BH_EHIP;
  int *d_did = cws<int>(1, np);
  CWS_CHECK(d_did);
  if (domain_ids(o, d_pts, np, d_did)) return BH_EHIP;
  HIPCHK(hipMemcpyAsync(did, d_did, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  return 0;
}

// fields at np points on the surface of object oid (element t of domain
// did, local coordinates zeta/eta): mode 0 s (own boundary integral),
// 1 t (+ other objects + incident), 2 i (incident + other objects),
// 3 a (own + other objects); mo_EH_*_bd, mo_field.c:632-718
extern "C" int bh_mo_EH_bd(bh_object **objs, int N, int oid, int did, int np, const int *t, const double *ze, int type,
                           int mode, double *out)
{
  API_LOCK;
  if (N <= 0 || oid < 0 || oid >= N || np <= 0) return fail(BH_EARG, "bad boundary-field request");
  if (type != 2) return fail(BH_EUNSUP, "boundary fields use the GLN rule (type 2)");
  if (ensure_rules()) return BH_EHIP;
  bh_object *o = objs[oid];
  if (did < 0 || did > o->MN) return fail(BH_EARG, "bad domain %d", did);
  Dom &D = o->dom[did];
  for (int p = 0; p < np; p++)
    if (t[p] < 1 || t[p] > D.dNe) return fail(BH_EARG, "element %d out of range", t[p]);
  if (prepare_solve(o)) return BH_EHIP;
  if (mode != 2 && !o->fmode && !D.d_E) return fail(BH_ESTATE, "object not solved");
  for (int m = 0; m < N; m++)
    if (!objs[m]->d_beams) return fail(BH_ESTATE, "object %d has no beams", m);
  int *d_t, *d_did, *d_mask;
  double *d_ze, *d_rt;
  cd *d_out;
  if (cput(0, &d_t, t, np) || cput(1, &d_ze, ze, 2 * (size_t)np)) return BH_EHIP;
  d_rt = cws<double>(2, 3 * (size_t)np);
  d_out = cws<cd>(3, 6 * (size_t)np);
  d_did = cws<int>(4, np);
  d_mask = cws<int>(5, np);
  double *d_rt2 = cws<double>(6, 3 * (size_t)np);
  CWS_CHECK(d_rt && d_out && d_did && d_mask && d_rt2);
  HIPCHK(hipMemsetAsync(d_out, 0, 6 * (size_t)np * sizeof(cd), cur_st()));
  // positions on the element (r = c0 + c1 zeta + c2 eta)
  {
    std::vector<double> rt(3 * (size_t)np);
    for (int p = 0; p < np; p++) {
      const double *cr = &o->cr[(size_t)std::abs(D.sid[t[p]]) * 12];
      for (int q = 0; q < 3; q++) rt[3 * p + q] = cr[q * 4] + cr[q * 4 + 1] * ze[2 * p] + cr[q * 4 + 2] * ze[2 * p + 1];
    }
    HIPCHK(hipMemcpyAsync(d_rt, rt.data(), rt.size() * sizeof(double), hipMemcpyHostToDevice, cur_st()));
  }
  // own boundary integral (needs the solved boundary values)
  if (mode != 2) {
    bem_launch_field_bd(d_t, d_ze, np, did, o->d_geo, o->d_gq, D.d_sid, D.dNe, o->kn[did].real(), o->kn[did].imag(),
                        g_rule[2], o->prec != BH_PREC_FP64 ? g.d_gl32 : nullptr, g_nrule[2], g_rscale[2], g_qrule[2],
                        g_nqrule[2], g.d_xl, g.d_wl, g.d_xh, g.d_wh, D.d_E, D.d_dE, D.d_Hf, D.d_dHf, d_out,
                        d_rt2, cur_st());
  }
  if (did == 0 && mode != 0) {
    for (int m = 0; m < N; m++) {
      if (m == oid) continue;
      bh_object *q = objs[m];
      if (domain_ids(q, d_rt, np, d_did)) return BH_EHIP;
      for (int d = 0; d <= q->MN; d++) {
        Dom &Q = q->dom[d];
        bem_launch_field(d_rt, d_did, np, d, q->d_geo, Q.d_sid, Q.dNe, q->kn[d].real(), q->kn[d].imag(), g_rule[2],
                         g_nrule[2], q->prec != BH_PREC_FP64 ? g.d_gl32 : nullptr, g_rscale[2], q->d_gq, g_qrule[2],
                         g_nqrule[2], 0, Q.d_E, Q.d_dE, Q.d_Hf, Q.d_dHf, d_out, 1, cur_st());
      }
    }
    if (mode == 1 || mode == 2) {
      std::vector<int> ones(np, 1);
      HIPCHK(hipMemcpyAsync(d_mask, ones.data(), np * sizeof(int), hipMemcpyHostToDevice, cur_st()));
      add_incident(objs[0], d_rt, d_mask, np, d_out, 0);
      if (fbeam_failed()) return BH_ECONV;
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, d_out, 6 * (size_t)np * sizeof(cd), hipMemcpyDeviceToHost, cur_st()));
  HIPCHK(hipStreamSynchronize(cur_st()));
  return 0;
}

// device state of the library released (every object destroyed first):
// rule tables, work areas, timing events.  bh_init (or any entry point) sets
// the library up again.  The uncached exchange areas of the one-launch LU
// panel and triangular solves and the device streams stay until the
// process ends.
extern "C" int bh_finalize(void)
{
  API_LOCK;
  if (g.dev < 0) return 0;
  (void)hipStreamSynchronize(cur_st());
  bem_mo_release();
  if (g_cws) { delete[] g_cws; g_cws = nullptr; }
  if (g_skjob) { delete g_skjob; g_skjob = nullptr; }
  for (int i = 0; i < 4; i++) {
    dfree(g_rule[i]); g_rule[i] = nullptr; g_nrule[i] = 0;
    dfree(g_qrule[i]); g_qrule[i] = nullptr; g_nqrule[i] = 0;
  }
  dfree(g_phit); g_phit = nullptr;
  dfree(g_fberr); g_fberr = nullptr;
  dfree(g.d_gl); dfree(g.d_gh); dfree(g.d_gl32); dfree(g.d_glr);
  dfree(g.d_xl); dfree(g.d_wl); dfree(g.d_xh); dfree(g.d_wh);
  bem_la_release();
  bem_prof_release();
  (void)hipEventDestroy(g.ev[0]);
  (void)hipEventDestroy(g.ev[1]);
  // The streams (this one, the lanes', the coupling loop's G1 stream) stay
  // for the next bh_init: a process whose streams had carried cooperative
  // launches and were destroyed crashed in the runtime's teardown at exit.
  g = Ctx();
  return 0;
}
