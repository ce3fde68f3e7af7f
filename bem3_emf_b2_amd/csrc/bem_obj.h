// bem_obj.h -- internals shared by the host-side sources of the HIP library:
// the device context, the object record that lives in HBM for the object's
// whole life, grow-only work areas and the error / locking conventions of
// the C-ABI (include/bem3_hip.h).  Everything but the object record sits in
// namespace bem, so the shared object exports only the bh_* entry points
// under C names.
#pragma once
#include "bem_dev.h"
#include "bem_schur.h"
#include "../../include/bem3_hip.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

typedef std::complex<double> zd;

namespace bem {

// the message of the last failure on the calling thread; returns code
int fail(int code, const char *fmt, ...);
inline void dfree(void *p) { (void)hipFree(p); }

// Entry points are serialised: callers may call the scalar field functions
// from their own thread pools (e.g. OpenMP loops over grid points); all work
// goes to one device stream in call order.
extern std::recursive_mutex g_mu;

struct Ctx {
  int dev = -1;
  hipStream_t st = nullptr;
  QPt *d_gl = nullptr, *d_gh = nullptr;   // 3-part triangle rules
  int n_gl = 0, n_gh = 0;
  void *d_gl32 = nullptr;                 // FP32 far-pair rules, 12x12 / 6x6 / 5x5 / 4x4 (QPt32)
  QPt *d_glr = nullptr;                   // the same rules in FP64 (QPt): far pairs of the mixed-form assembly
  double *d_xl = nullptr, *d_wl = nullptr, *d_xh = nullptr, *d_wh = nullptr;
  double xl[12], wl[12], xh[64], wh[64];
  hipEvent_t ev[2];
};
extern Ctx g;

// the stream a thread's entry-point work goes to: the library's stream,
// unless the thread is a worker of the coupling loop bound to a stream of
// its own (set_cur_st)
extern thread_local hipStream_t t_cur_st;
inline hipStream_t cur_st() { return t_cur_st ? t_cur_st : g.st; }
inline void set_cur_st(hipStream_t s) { t_cur_st = s; }

// a grow-only device buffer
struct Scratch {
  void *p = nullptr;
  size_t n = 0;
  Scratch() = default;
  Scratch(const Scratch &) = delete;
  Scratch(Scratch &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  Scratch &operator=(const Scratch &) = delete;
  ~Scratch() { dfree(p); }
  void release() { dfree(p); p = nullptr; n = 0; }
  template <class T> T *get(size_t cnt)
  {
    size_t b = cnt * sizeof(T);
    if (b > n) { dfree(p); p = nullptr; if (hipMalloc(&p, b) != hipSuccess) { n = 0; return nullptr; } n = b; }
    return (T *)p;
  }
};


inline dim3 g1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// rules of the field kernels: 0 4-point, 1 7-point, 2 GLN, 3 GHN
extern QPtN *g_rule[4];
extern int g_nrule[4];
extern double g_rscale[4];
extern QPtQ *g_qrule[4];
extern int g_nqrule[4];
int ensure_rules();

struct Dom {
  int dNe = 0, nd = 0;                 // elements seen, compact node count
  std::vector<int> sid, colOff;        // [dNe+1]
  std::vector<RowPt> rows;             // [nd]
  int *d_sid = nullptr, *d_colOff = nullptr;
  RowPt *d_rows = nullptr;
  cd *d_G = nullptr, *d_H = nullptr, *d_dG = nullptr, *d_dH = nullptr;
  double *d_dF = nullptr;
  int n_gh = 0;
  // solve-side tables and boundary values ([dNe+1][4][4] U/dU, [dNe+1][4][3] E/H)
  ColInfo *d_col = nullptr;
  int *d_colS = nullptr;
  RowEH *d_rowEH = nullptr;
  cd *d_U = nullptr, *d_dU = nullptr, *d_E = nullptr, *d_dE = nullptr, *d_Hf = nullptr, *d_dHf = nullptr;
  cd *d_Gf = nullptr;                  // LU factors of G (normal derivatives of E, H)
  int *d_Gpiv = nullptr;
  bool gf_valid = false;               // d_Gf factors the current G (cleared by assembly)
  cf *c_Gf = nullptr;                  // complex-FP32 LU factors of G (mixed-precision form)
  bool gfc_valid = false;
  bool has_inc = false;                // some column takes incident data (open-domain sub side)
};


}  // namespace bem

#define API_LOCK std::lock_guard<std::recursive_mutex> api_lock_(bem::g_mu)
#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return bem::fail(BH_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

namespace bem {
// a device copy of n host values (on the library's stream)
template <class T> int upload(T **d, const T *h, size_t n)
{
  HIPCHK(hipMalloc((void **)d, std::max<size_t>(n, 1) * sizeof(T)));
  if (n) HIPCHK(hipMemcpyAsync(*d, h, n * sizeof(T), hipMemcpyHostToDevice, cur_st()));
  return 0;
}
}  // namespace bem

struct bh_object {
  int Nn, Ne, MN, NN;
  int nquad = 0;                       // quadrangle elements
  std::vector<double> rn;              // [Nn+1][3]
  std::vector<int> ed, md, sd, eni;    // [Ne+1][4], [Ne+1]
  std::vector<double> cr, cw, ren, wen;// [Ne+1][3][4], [Ne+1][3][3], [Ne+1][4][3] x2
  std::vector<zd> n, kn;
  std::vector<bem::Dom> dom;
  TriGeo *d_geo = nullptr;
  QuadExt *d_gq = nullptr;             // bi-linear terms of quadrangles (zero rows for triangles)
  bool all_tri = true;
  // incident data at every element slot [Ne+1][4][4]
  cd *d_Ui = nullptr, *d_dUi = nullptr, *d_Ui0 = nullptr, *d_dUi0 = nullptr, *d_Uip = nullptr;
  // open-domain node slots (targets of the incident and coupling fields)
  int nslot = 0;
  double *d_slotPt = nullptr, *d_slotN = nullptr;
  int *d_slotIdx = nullptr;
  // faces grouped by main domain for the domain test
  double *d_fv = nullptr;
  int *d_fsd = nullptr, *d_fq = nullptr, *d_seg = nullptr;
  // beams
  std::vector<Beam> beams;             // plane waves
  std::vector<FBeam> fbeams;           // the other multi_fbeam types
  FBeam *d_fb = nullptr;
  double *d_fbt = nullptr;             // their spectral angles (cos, sin, weight x sin)
  Beam *d_beams = nullptr;
  double lambda0 = 0.0;
  int prec = BH_PREC_FP64;             // arithmetic of assembly, field sums and factor mode 0
  // factorisation: 0 none, 1 dense 8NN, 2 Schur (one medium), 3 Schur in FP32 refined in FP64
  int fmode = 0;                       // factorisation of the current matrices (cleared by assembly)
  cd *d_A = nullptr, *d_work = nullptr;
  int *d_piv = nullptr, *d_info = nullptr;
  cd *d_G0f = nullptr, *d_Z = nullptr, *d_Mf = nullptr, *d_Sf = nullptr;
  int *d_piv0 = nullptr, *d_pivM = nullptr, *d_pivS = nullptr;
  cd sch_delt, sch_beps, sch_sigm;
  bool s_formed = false;               // Sf holds LU(S) (else Sp = H1 + beta G1 Z, S solved by GMRES)
  bool x_warm = false;                 // d_X solves the current factors' system for an earlier B
                                       // (the refinement of the mixed form starts from it)
  bool r_valid = false;                // d_R = d_B - A d_X for the current factors' system (kept by
                                       // objects_solve_batch for its warm start)
  int gm_steps = 0;                    // GMRES steps of the last S solve(s)
  cd *d_X = nullptr, *d_B = nullptr, *d_tmp = nullptr;
  cd *d_w = nullptr;                   // work of the block solve (schur_work(NN))
  // mixed-precision form (fmode 3): complex-FP32 factors of the block
  // elimination, refined against the FP64 matrices
  cf *c_G0f = nullptr, *c_Z = nullptr, *c_Mf = nullptr, *c_Sf = nullptr, *c_G1 = nullptr;
  cf *c_w = nullptr;                   // schur_work(NN) work, 8 NN right-hand side, 8 NN correction
  cd *d_R = nullptr;                   // 8 NN residual
  double *d_nrm = nullptr;             // 256 partial sums of the norm
  int *d_ginfo = nullptr;              // zero-pivot report of object_factor_g
  int ir_iters = 0;                    // refinement steps of the last solve
  double ir_res = 0.0;                 // its final relative residual
  int asm_prec = -1;                   // precision the current matrices were assembled in (-1: none)
  // grow-only device work areas of the entry points (freed with the object)
  bem::Scratch ws_part, ws_mask, ws_pairs, ws_start, ws_ghdf, ws_self;   // assembly
  bem::Scratch ws_jobs;                // job table of a batched assembly this object heads
  bem::Scratch ws_res;                 // batched refinement residuals: per domain a, c, Y (12 nd)
  bem::Scratch ws_wt;                  // forming S
  bem::Scratch ws_b;                   // right-hand side: the incident data as column operands
  // low-memory mixed form: FP64 row blocks of G, H, dG, dH, dF assembled on
  // demand; lm_raw: the FP32 storage (c_G0f: G0, c_Z: H0, c_G1: G1, c_Mf: H1)
  // holds the assembled matrices, not yet the factors
  bem::Scratch lm_blk[5];
  bool lm_raw = false;
  int lm_rows = 0;                     // rows per block (0: by the memory budget)
  bem::Scratch ws_eh[5];               // boundary E, H: packed rows, products, refinement
  // many beams as right-hand sides (bh_mo_solve_beams): nbm plane waves;
  // beam b of each array at + b x (its one-beam size)
  int nbm = 0;
  bem::Scratch mb_inc[5];              // Ui0, dUi0, Ui, dUi, Uip: [nb][Ne+1][16]
  std::vector<bem::Scratch> mb_dom;    // per domain: U, dU ([nb][dNe+1][16]), E, dE, H, dH ([nb][dNe+1][12])
  bem::Scratch mb_sys[4];              // B, X, R, the B of R: [nb][8NN]
  bem::Scratch mb_inv[3];              // FP32 inverses of LU(G0), LU(M), LU(S) ([NN][NN])
  unsigned fact_gen = 0;               // factorisations so far (the inverses' validity)
  unsigned mb_inv_gen = ~0u;           // fact_gen the inverses were formed for
  bem::Scratch mb_ws[8];               // solve and boundary-field work areas
  bem::Scratch mb_beams;               // the beams (Beam records)
  std::vector<Beam> mb_hbeams;         // and their host copies
  bool mb_xwarm = false;               // X solves the current factors' system for an earlier B
  int mb_iters = 0;                    // refinement steps of the last many-beam solve
  double mb_res = 0.0;                 // its largest final relative residual
};

namespace bem {
// the entry points' work without the entry lock, on the thread's current
// stream (the coupling loop's workers call these)
int object_assemble(bh_object *o, int type);
int objects_assemble(bh_object **objs, int N, int type);
// the tuning table (bem_tune.hip)
int tune_value(int key);
int tune_init();
int object_factor(bh_object *o, int mode);
int object_solve(bh_object *o);
// the mixed-form solves of N objects in lock step (BH_EUNSUP: an object
// does not qualify, nothing done)
int objects_solve_batch(bh_object **objs, int N);
int object_solve_EH(bh_object *o);
// the FP32 LU of G_d (d >= 1) of a mixed-form object that object_solve_EH
// needs, ahead of it (once per assembly)
int object_factor_g(bh_object *o);
// many beams (bh_mo_solve_beams): nb plane waves (12 doubles each) as the
// object's right-hand sides and their incident data; the solve and the
// boundary E, H of all of them (the mixed-precision form with S formed)
int object_beams_set(bh_object *o, int nb, const double *par);
int object_solve_multi(bh_object *o);
int object_solve_EH_multi(bh_object *o);
int object_select_beam(bh_object *o, int b);
cd *mb_ptr(bh_object *o, int which);   // 0 Ui0, 1 dUi0, 2 Ui, 3 dUi, 4 Uip
cd *mb_dom_ptr(bh_object *o, int d, int which);   // 0 U, 1 dU, 2 E, 3 dE, 4 H, 5 dH
int prepare_solve(bh_object *o);
int form_s(bh_object *o);
int domain_ids(bh_object *o, const double *d_pts, int np, int *d_did);
int trsv_failed(hipStream_t st);
}  // namespace bem
