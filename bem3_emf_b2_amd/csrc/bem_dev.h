// bem_dev.h -- device-side helpers shared by the HIP kernels of the
// bem3_emf_b2 hot path (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BEM_WAVE 64

struct cd {
  double x, y;
};

__host__ __device__ __forceinline__ cd mk(double a, double b) { cd r; r.x = a; r.y = b; return r; }
__host__ __device__ __forceinline__ cd operator+(cd a, cd b) { return mk(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ cd operator-(cd a, cd b) { return mk(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ cd operator-(cd a) { return mk(-a.x, -a.y); }
__host__ __device__ __forceinline__ cd operator*(cd a, cd b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ __forceinline__ cd operator*(cd a, double s) { return mk(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ cd operator*(double s, cd a) { return mk(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ cd cdiv(cd a, cd b)
{
  double d = b.x * b.x + b.y * b.y;
  return mk((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__host__ __device__ __forceinline__ cd conjd(cd a) { return mk(a.x, -a.y); }
__host__ __device__ __forceinline__ double abs2(cd a) { return a.x * a.x + a.y * a.y; }
// a += b * s  (s real)
__device__ __forceinline__ void fma_cr(cd &a, cd b, double s) { a.x = fma(b.x, s, a.x); a.y = fma(b.y, s, a.y); }

// complex single precision (the mixed-precision factorisation)
struct cf {
  float x, y;
};
__host__ __device__ __forceinline__ cf mkf(float a, float b) { cf r; r.x = a; r.y = b; return r; }
__host__ __device__ __forceinline__ cf operator+(cf a, cf b) { return mkf(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ cf operator-(cf a, cf b) { return mkf(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ cf operator-(cf a) { return mkf(-a.x, -a.y); }
__host__ __device__ __forceinline__ cf operator*(cf a, cf b) { return mkf(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ __forceinline__ cf operator*(cf a, float s) { return mkf(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ cf cdiv(cf a, cf b)
{
  float d = b.x * b.x + b.y * b.y;
  return mkf((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__host__ __device__ __forceinline__ cf to_cf(cd a) { return mkf((float)a.x, (float)a.y); }
__host__ __device__ __forceinline__ cd to_cd(cf a) { return mk((double)a.x, (double)a.y); }

// e^{i k R} for complex k = kr + i ki
__device__ __forceinline__ cd eikr(double kr, double ki, double R)
{
  double s, c;
  sincos(kr * R, &s, &c);
  if (ki != 0.0) {
    double a = exp(-ki * R);
    s *= a; c *= a;
  }
  return mk(c, s);
}

// ---- wave64 reductions (DPP/shuffle via __shfl_xor) ----
__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, BEM_WAVE);
  return v;
}
__device__ __forceinline__ cd wsum(cd v) { return mk(wsum(v.x), wsum(v.y)); }
__device__ __forceinline__ float wsum(float v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, BEM_WAVE);
  return v;
}
__device__ __forceinline__ cf wsum(cf v) { return mkf(wsum(v.x), wsum(v.y)); }

// element record: r(z,e) = c0 + c1 z + c2 e (+ c3 z e on a quad), W =
// unnormalised normal (constant on a triangle; W0 of W(z,e) = W0 + W1 z +
// W2 e on a quad), aW = |W|, quad = 1 for a bi-linear quadrangle
struct TriGeo {
  double c0[3], c1[3], c2[3], W[3], aW, quad, pad[2];
};
// the bi-linear terms of a quadrangle (zero on triangles)
struct QuadExt {
  double c3[3], W1[3], W2[3], pad[3];
};

// quadrature-point record of a fixed rule: position in the reference
// element and the full weight (rule weight times mapping factor)
struct QPt {
  double z, e, w;
};

// interpolation functions at the node slots of the triangle rule
// (reference: d3b1_elem.c:700-711), 1/(2*P34_N0) = 2.5
__device__ __forceinline__ void tri_N(double z, double e, double &n0, double &n1, double &n2)
{
  const double SQ3 = 1.732050807568877293528;
  n0 = (1.0 - 2.5 * (z + SQ3 * e)) * (1.0 / 3.0);
  n1 = (1.0 + 5.0 * z) * (1.0 / 3.0);
  n2 = (1.0 - 2.5 * (z - SQ3 * e)) * (1.0 / 3.0);
}

#define BEM_SQ3 1.732050807568877293528
#define BEM_I_EP 3.978873577297383394222e-2    // 1/(8 pi)
#define BEM_I_FP 7.957747154594766788444e-2    // 1/(4 pi)
#define BEM_I_SPSQ3 3.062938307898844719507e-2 // 1/(6 pi sqrt 3)
#define BEM_DEPS 0.1
#define BEM_GEPS 1.0e-15   // geometric tolerance of bil_check_plane (d3b1_const.h:28)
#define BEM_CBD_SDF 0.5
#define BEM_CBD_CDF 100.0   // field-derivative cut-off of EH_mEMP_s (d3b1_const.h:78)

// one row of a domain's coefficient matrices: node slot `slot` of the
// domain's t-th element (1-based), its position and unit tangents
struct RowPt {
  double rt[3], v0[3], v1[3];
  int t, slot;
  int pad[2];
};

// a near (row, source element) pair re-integrated with the 64x64 rule
struct GhPair {
  int row, s;
};

// One job of a batched assembly: rows [r0, r0 + nrows) of one domain of one
// object.  Every assembly pass is one launch over the jobs of a batch (all
// objects and domains of bh_mo_prepare, or the domains of one object); a
// workgroup finds its job from the launch's prefix tables (job_of).
struct AsmJob {
  const TriGeo *geo;
  const QuadExt *gq;
  const int *sid;              // signed element of each source slot (1-based)
  const RowPt *rows;           // the job's rows
  const int *colOff;           // first column of each source element
  cd *G, *H, *dG, *dH;         // the job's first row; row stride ld
  double *dF;                  // 3 per row
  double *dfPart;              // pass 1: [nrows][nwv][3] per-wave F sums
  unsigned long long *ghMask;  // pass 1: [nrows][nwv] near-pair flags
  const GhPair *pairs;         // pass 2: the job's near pairs
  const int *ghStart;          // [nrows + 1] first pair of each row
  double *ghDF;                // [npairs][3] their F sums
  double *selfDF;              // pass 3: [nrows][3] the self element's F terms
  double kr, ki;               // wavenumber of the domain
  size_t ld;
  int dNe;                     // source elements of the domain
  int nrows, nblk;             // rows; 256-lane workgroups per row in pass 1 (nwv = 4 nblk)
  int r0;                      // the job's first row in the domain (its diagonal column offset)
  int open;                    // 1: the open domain 0 (diagonal 1 + F)
  int pad;
};

// Prefix tables of a batch of nj jobs: first[j] is job j's first global
// index (workgroup of pass 1, pair of pass 2, row of passes 3 and 4),
// first[nj] the total.  The job holding global index i (0 <= i < first[nj]):
// the last j with first[j] <= i (empty jobs are skipped that way).
__host__ __device__ inline int job_of(const int *first, int nj, int i)
{
  int lo = 0, hi = nj;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (first[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// column of a domain matrix in the 8NN unknown numbering and its boundary
// condition class (main side of its element or sub side with delta, beta,
// sigma of the element's media)
struct ColInfo {
  int node;      // global element-node index (0..NN-1)
  int elem;      // element id (1-based)
  int slot;      // node slot 0..2
  int main;      // 1: the element's main side is this domain
  int mdid;      // main domain of the element
  int pad;
  double n[3];   // unit normal at the node slot
  cd delt, beps, sigm;
};

// node row of a domain for the boundary E, H step
struct RowEH {
  double vn[3], vtz[3], vte[3];   // unit normal (domain side), unit tangents
  int s, slot;                    // domain-local element, node slot
};

// quadrature point with its shape-function weights
struct QPtN {
  double z, e, w, n0, n1, n2;
};

// shape functions at a rule point: the triangle's interpolation functions
// (a QPt table: the far rules), or the values a QPtN table holds (the
// 4-point node rule carries its own)
__device__ __forceinline__ void rule_N(const QPt &q, double &n0, double &n1, double &n2) { tri_N(q.z, q.e, n0, n1, n2); }
__device__ __forceinline__ void rule_N(const QPtN &q, double &n0, double &n1, double &n2)
{
  n0 = q.n0; n1 = q.n1; n2 = q.n2;
}

// point of a rule on the square [-1,1]^2 with its four shape functions
// (the 4-point node rule carries N = delta and weight 1)
struct QPtQ {
  double z, e, w, n[4];
};

// FP32 copy of a triangle rule point: position, weight, weight x shape functions
struct QPt32 {
  float z, e, w, wn0, wn1, wn2;
};

// Far-pair rules of the mixed-precision form, one table (an FP64 copy for
// its assembly, an FP32 copy for its boundary-field and coupling sums): the
// 3-part triangle map with 12x12 (432 points), 6x6 (108), 5x5 (75) and 4x4
// (48) Gauss-Legendre squares.  The element radius is the largest distance
// from the centre c0 to a vertex (elem_rad2).  A pair farther than 4 radii
// takes 5x5 while |k| times the radius is at most 1.2, 6x6 up to 3, else
// 12x12; farther than 8 radii with |k| x radius at most 0.5 it takes 4x4.
// Against a 20x20 rule the relative error of every G, H, dG/dv, dH/dv group
// stays below 3e-8 there, for general (also skewed) triangles and real and
// lossy k (largest value seen over thousands of random pairs per tier: 2e-8).
#define BEM_FAR_OFF6 432
#define BEM_FAR_OFF5 540
#define BEM_FAR_OFF4 615
#define BEM_FAR_NPT 663
#define BEM_FAR_RADII 4.0
// squared element radius: the largest squared distance from the centre c0
// of a triangle (the reference triangle's origin) to its vertices c0 + c1,
// c0 - c1/2 -+ (sqrt3/2) c2
__host__ __device__ __forceinline__ double elem_rad2(const TriGeo &g)
{
  double a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double h = -0.5 * g.c1[d], q = 0.8660254037844386467637 * g.c2[d];
    a += g.c1[d] * g.c1[d];
    b += (h - q) * (h - q);
    c += (h + q) * (h + q);
  }
  return fmax(a, fmax(b, c));
}
// Phase of an FP32 far-pair point.  e^{ikR} is split as e^{ik Rc} e^{ik dR}:
// Rc, the distance from the row point to the element's centre, and its
// factor are formed once per pair in FP64 (far32_centre); a point adds only
// dR = R - Rc = (2 d0.dl + |dl|^2) / (R + Rc), with d0 the centre offset and
// dl the point's offset inside the element, so the FP32 rounding of d0
// (relative 6e-8 of R, a phase error of 6e-8 k R) stays out of the phase.
__device__ __forceinline__ cf far32_centre(double dx, double dy, double dz, double kr, double ki, float &Rc)
{
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  double s, c;
  sincos(kr * r, &s, &c);
  const double a = ki != 0.0 ? exp(-ki * r) : 1.0;
  Rc = (float)r;
  return mkf((float)(c * a), (float)(s * a));
}
__device__ __forceinline__ cf far32_phase(cf ec, float dR, float kr, float ki)
{
  float sn, cs;
  sincosf(kr * dR, &sn, &cs);
  if (ki != 0.0f) {
    const float at = expf(-ki * dR);
    sn *= at;
    cs *= at;
  }
  return ec * mkf(cs, sn);
}

// wave-uniform rule of the far lanes: the most points any of them needs
// (call from every lane that may be far; kR = |k| x element radius, d2 =
// squared distance to the element's centre over the squared radius)
__device__ __forceinline__ void far_rule(bool far, double kR, double d2, int &off, int &nq)
{
  const bool need12 = far && kR > 3.0, need6 = far && kR > 1.2;
  const bool need5 = far && (kR > 0.5 || d2 <= 64.0);
  if (__ballot(need12)) { off = 0; nq = 432; }
  else if (__ballot(need6)) { off = BEM_FAR_OFF6; nq = 108; }
  else if (__ballot(need5)) { off = BEM_FAR_OFF5; nq = 75; }
  else { off = BEM_FAR_OFF4; nq = 48; }
}

// one incident plane wave (ipw.h:79-99 derived data)
struct Beam {
  double ki, E0, st, ct, sp, cp, fx, fy, fz, pad;
  cd ex, ey, ez, hx, hy, hz;
};

// one beam of the other multi_fbeam types: 1 focused plane wave, 2 focused
// with spiral phase, 3 Bessel, 4 Bessel with spiral phase, 5 radially /
// azimuthally polarised focused (ex, ey, hx, hy then hold the radial and
// azimuthal components); derived data as the set-up functions of fpw.c,
// lgb.c, bsb.c, bslgb.c, rab.c.  toff: offset of the beam's spectral angles
// (cos, sin, weight x sin per point, nt points) in the angle table.
struct FBeam {
  int type, nt, lg_m, toff;
  double ki, E0, ni, pad;
  double cos_t, sin_t, cos_p, sin_p, fx, fy, fz;
  double ct, st;                  // cone angle of a Bessel beam
  cd ex, ey, hx, hy;
};
#define BEM_TRAP_MNUM 12                         // halvings of the phi trapezoid (mfb_const.h:4)
#define BEM_NPHI ((2 << BEM_TRAP_MNUM) + 1)      // its points
