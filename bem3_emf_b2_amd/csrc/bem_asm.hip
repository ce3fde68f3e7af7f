// bem_asm.hip -- coefficient-matrix assembly for meshes of linear triangles
// and bi-linear quadrangles.
//
// For one domain d (element list sid_d, wave number k_d) and every pair
// (row node slot, source element) the kernels produce the 18 coefficients of
// the reference's dcoef_bd_node_t2 (d3b1_elem.c:482-539):
//   G, H (3 shape functions), dG/dt_z, dG/dt_e, dH/dt_z, dH/dt_e and the
//   per-row sums F, dF/dt_z, dF/dt_e (solve_bieq2.c:124-166).
//
// Pass 1 (k_asm_gl)   : 3-part 12x12 Gauss-Legendre rule (432 points, the
//                       lit_dcoef_v2_GL rule) for every pair, one lane per
//                       source element, 64 consecutive elements per
//                       wavefront, rule table read through scalar loads.
//                       Pairs with |dF/dt| > 0.5 (the reference's GHN switch)
//                       are flagged in a ballot bitmap and left for pass 2.
// Pass 2 (k_asm_gh)   : 3-part 64x64 rule (12288 points) for flagged pairs,
//                       one wavefront per pair, lanes over points.
// Pass 3 (k_asm_self) : self element: polar-coordinate G (lit_coef_bd) and
//                       the principal-value tangential derivatives
//                       (lit_dcoef_bd_tz0/te0_PV), one wavefront per row.
// Pass 4 (k_asm_fin)  : per-row F sums in a fixed order, the 1+F / F
//                       diagonal term, dF rows.
//
// A quadrangle source element takes the square's rules instead (12x12 and
// 64x64 tensor Gauss-Legendre, bil_dcoef_v2_GL/GH, bil_dcoef_v2.c:142-326)
// with four shape functions, and a quadrangle self element the polar rule
// about the node with the principal-value derivatives of bil_dcoef.c:369-740.
//
// Matrices are compact: row = node slot index in the domain's element order,
// column = shape-function index in the same order (row-major, leading
// dimension ld, complex128 interleaved).  dG/dH rows are [row][dir][ld].
#include "bem_dev.h"
#include "bem_prof.h"

#include <cstdlib>

struct AsmAcc {
  cd G[3], H[3], DG0[3], DG1[3], DH0[3], DH1[3];
  double F, DF0, DF1;
};

__device__ __forceinline__ void acc_zero(AsmAcc &a)
{
#pragma unroll
  for (int i = 0; i < 3; i++) {
    a.G[i] = mk(0, 0); a.H[i] = mk(0, 0); a.DG0[i] = mk(0, 0); a.DG1[i] = mk(0, 0); a.DH0[i] = mk(0, 0); a.DH1[i] = mk(0, 0);
  }
  a.F = a.DF0 = a.DF1 = 0.0;
}

// integrand of lit_sdGL_v2_zeta_eta (lit_dcoef_v2.c:381-413) at one point,
// all three shape functions at once
__device__ __forceinline__ void asm_point(AsmAcc &a, const TriGeo &g, const RowPt &rp, double Wdv0, double Wdv1,
                                          double kr, double ki, double z, double e, double w)
{
  double n0, n1, n2;
  tri_N(z, e, n0, n1, n2);
  double vx = fma(g.c2[0], e, fma(g.c1[0], z, g.c0[0])) - rp.rt[0];
  double vy = fma(g.c2[1], e, fma(g.c1[1], z, g.c0[1])) - rp.rt[1];
  double vz = fma(g.c2[2], e, fma(g.c1[2], z, g.c0[2])) - rp.rt[2];
  double R = sqrt(vx * vx + vy * vy + vz * vz);
  double iR = 1.0 / R, iR2 = iR * iR, iR3 = iR2 * iR;
  double rdW = (vx * g.W[0] + vy * g.W[1] + vz * g.W[2]) * iR;
  double rd0 = (vx * rp.v0[0] + vy * rp.v0[1] + vz * rp.v0[2]) * iR;
  double rd1 = (vx * rp.v1[0] + vy * rp.v1[1] + vz * rp.v1[2]) * iR;
  cd ce = eikr(kr, ki, R);
  cd cb = mk(-1.0 - ki * R, kr * R);            // i k R - 1
  cd cbce = cb * ce;
  cd gk = ce * iR;
  cd hk = cbce * (rdW * iR2);
  cd tk = cbce * iR2;
  cd kR = mk(kr * R, ki * R);
  cd q = kR * kR + cb * 3.0;
  cd ce3 = ce * iR3;
  cd dh0 = ce3 * (q * (rd0 * rdW) - cb * Wdv0);
  cd dh1 = ce3 * (q * (rd1 * rdW) - cb * Wdv1);
  double wn[3] = {w * n0, w * n1, w * n2};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    fma_cr(a.G[i], gk, wn[i]);
    fma_cr(a.H[i], hk, wn[i]);
    fma_cr(a.DG0[i], tk, wn[i] * rd0);
    fma_cr(a.DG1[i], tk, wn[i] * rd1);
    fma_cr(a.DH0[i], dh0, wn[i]);
    fma_cr(a.DH1[i], dh1, wn[i]);
  }
  a.F = fma(w, rdW * iR2, a.F);
  a.DF0 = fma(w, iR3 * (3.0 * rd0 * rdW - Wdv0), a.DF0);
  a.DF1 = fma(w, iR3 * (3.0 * rd1 * rdW - Wdv1), a.DF1);
}

// final scale of the 3-part rule (lit_dcoef_v2.c:252-262) and the sign
// change of H, F and their derivatives on the sub-domain side
__device__ __forceinline__ void acc_finish(AsmAcc &a, double aW, bool flip)
{
  const double S = 0.0625 * BEM_SQ3 * BEM_I_SPSQ3;
  double sh = flip ? -S : S;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    a.G[i] = a.G[i] * (S * aW);
    a.H[i] = a.H[i] * sh;
    a.DG0[i] = a.DG0[i] * (-S * aW);
    a.DG1[i] = a.DG1[i] * (-S * aW);
    a.DH0[i] = a.DH0[i] * sh;
    a.DH1[i] = a.DH1[i] * sh;
  }
  a.F *= sh; a.DF0 *= sh; a.DF1 *= sh;
}

__device__ __forceinline__ void acc_store(const AsmAcc &a, size_t row, size_t col, size_t ld, cd *G, cd *H, cd *dG, cd *dH)
{
#pragma unroll
  for (int i = 0; i < 3; i++) {
    G[row * ld + col + i] = a.G[i];
    H[row * ld + col + i] = a.H[i];
    dG[(row * 2 + 0) * ld + col + i] = a.DG0[i];
    dG[(row * 2 + 1) * ld + col + i] = a.DG1[i];
    dH[(row * 2 + 0) * ld + col + i] = a.DH0[i];
    dH[(row * 2 + 1) * ld + col + i] = a.DH1[i];
  }
}

// ---- quadrangle source elements ----
struct AsmAccQ {
  cd G[4], H[4], DG0[4], DG1[4], DH0[4], DH1[4];
  double F, DF0, DF1;
};

__device__ __forceinline__ void accq_zero(AsmAccQ &a)
{
#pragma unroll
  for (int i = 0; i < 4; i++) {
    a.G[i] = mk(0, 0); a.H[i] = mk(0, 0); a.DG0[i] = mk(0, 0); a.DG1[i] = mk(0, 0); a.DH0[i] = mk(0, 0); a.DH1[i] = mk(0, 0);
  }
  a.F = a.DF0 = a.DF1 = 0.0;
}

// integrand of bil_dcoef_v2_GL/GH (bil_dcoef_v2.c:142-326) at one point of
// the square: r and the unnormalised normal W vary over the element, |W| is
// the area factor of G and dG
__device__ __forceinline__ void asmq_point(AsmAccQ &a, const TriGeo &g, const QuadExt &qx, const RowPt &rp, double kr,
                                           double ki, const QPtQ &q)
{
  double vx, vy, vz, W[3];
  {
    const double z = q.z, e = q.e, ze = z * e;
    vx = g.c0[0] + g.c1[0] * z + g.c2[0] * e + qx.c3[0] * ze - rp.rt[0];
    vy = g.c0[1] + g.c1[1] * z + g.c2[1] * e + qx.c3[1] * ze - rp.rt[1];
    vz = g.c0[2] + g.c1[2] * z + g.c2[2] * e + qx.c3[2] * ze - rp.rt[2];
#pragma unroll
    for (int d = 0; d < 3; d++) W[d] = g.W[d] + qx.W1[d] * z + qx.W2[d] * e;
  }
  const double R = sqrt(vx * vx + vy * vy + vz * vz);
  const double iR = 1.0 / R, iR2 = iR * iR, iR3 = iR2 * iR;
  const double aW = sqrt(W[0] * W[0] + W[1] * W[1] + W[2] * W[2]);
  const double rdW = (vx * W[0] + vy * W[1] + vz * W[2]) * iR;
  const double rd0 = (vx * rp.v0[0] + vy * rp.v0[1] + vz * rp.v0[2]) * iR;
  const double rd1 = (vx * rp.v1[0] + vy * rp.v1[1] + vz * rp.v1[2]) * iR;
  const double Wdv0 = W[0] * rp.v0[0] + W[1] * rp.v0[1] + W[2] * rp.v0[2];
  const double Wdv1 = W[0] * rp.v1[0] + W[1] * rp.v1[1] + W[2] * rp.v1[2];
  const cd ce = eikr(kr, ki, R);
  const cd cb = mk(-1.0 - ki * R, kr * R);   // i k R - 1
  const cd cbce = cb * ce;
  const cd gk = ce * (iR * aW);
  const cd hk = cbce * (rdW * iR2);
  const cd tk = cbce * (iR2 * aW);
  const cd kR = mk(kr * R, ki * R);
  const cd qq = kR * kR + cb * 3.0;
  const cd ce3 = ce * iR3;
  const cd dh0 = ce3 * (qq * (rd0 * rdW) - cb * Wdv0);
  const cd dh1 = ce3 * (qq * (rd1 * rdW) - cb * Wdv1);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const double wn = q.w * q.n[i];
    fma_cr(a.G[i], gk, wn);
    fma_cr(a.H[i], hk, wn);
    fma_cr(a.DG0[i], tk, wn * rd0);
    fma_cr(a.DG1[i], tk, wn * rd1);
    fma_cr(a.DH0[i], dh0, wn);
    fma_cr(a.DH1[i], dh1, wn);
  }
  a.F = fma(q.w, rdW * iR2, a.F);
  a.DF0 = fma(q.w, iR3 * (3.0 * rd0 * rdW - Wdv0), a.DF0);
  a.DF1 = fma(q.w, iR3 * (3.0 * rd1 * rdW - Wdv1), a.DF1);
}

// scale 1/(4 pi), -1/(4 pi) for dG, and the sign change of H, F and their
// derivatives on the sub-domain side (bil_convert_CC / bil_convert_dCC)
__device__ __forceinline__ void accq_finish(AsmAccQ &a, bool flip)
{
  const double sh = flip ? -BEM_I_FP : BEM_I_FP;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    a.G[i] = a.G[i] * BEM_I_FP;
    a.H[i] = a.H[i] * sh;
    a.DG0[i] = a.DG0[i] * (-BEM_I_FP);
    a.DG1[i] = a.DG1[i] * (-BEM_I_FP);
    a.DH0[i] = a.DH0[i] * sh;
    a.DH1[i] = a.DH1[i] * sh;
  }
  a.F *= sh; a.DF0 *= sh; a.DF1 *= sh;
}

__device__ __forceinline__ void accq_store(const AsmAccQ &a, size_t row, size_t col, size_t ld, cd *G, cd *H, cd *dG, cd *dH)
{
#pragma unroll
  for (int i = 0; i < 4; i++) {
    G[row * ld + col + i] = a.G[i];
    H[row * ld + col + i] = a.H[i];
    dG[(row * 2 + 0) * ld + col + i] = a.DG0[i];
    dG[(row * 2 + 1) * ld + col + i] = a.DG1[i];
    dH[(row * 2 + 0) * ld + col + i] = a.DH0[i];
    dH[(row * 2 + 1) * ld + col + i] = a.DH1[i];
  }
}

__device__ __forceinline__ void accq_wsum(AsmAccQ &a)
{
#pragma unroll
  for (int i = 0; i < 4; i++) {
    a.G[i] = wsum(a.G[i]); a.H[i] = wsum(a.H[i]); a.DG0[i] = wsum(a.DG0[i]); a.DG1[i] = wsum(a.DG1[i]);
    a.DH0[i] = wsum(a.DH0[i]); a.DH1[i] = wsum(a.DH1[i]);
  }
  a.F = wsum(a.F); a.DF0 = wsum(a.DF0); a.DF1 = wsum(a.DF1);
}

// one quadrangle pair of pass 1 (the lane's whole rule): the GHN test and,
// when the pair stays, its store and its F sums
__device__ __forceinline__ bool asm_gl_quad(const TriGeo &g, const QuadExt &qx, const RowPt &rp, int sd, double kr,
                                            double ki, const QPtQ *__restrict__ qrule, int nqq, double lim, size_t row,
                                            size_t col, size_t ld, cd *G, cd *H, cd *dG, cd *dH, double &f, double &df0,
                                            double &df1)
{
  AsmAccQ a;
  accq_zero(a);
  for (int p = 0; p < nqq; p++) asmq_point(a, g, qx, rp, kr, ki, qrule[p]);
  accq_finish(a, sd < 0);
  if (fabs(a.DF0) > lim || fabs(a.DF1) > lim) return true;
  accq_store(a, row, col, ld, G, H, dG, dH);
  f = a.F; df0 = a.DF0; df1 = a.DF1;
  return false;
}

// ---- pass 1 ----
// The fields of a batch job that passes 1-4 read, in registers.  Pass 1:
// workgroup b of the launch is workgroup xb of row `row` of its job.
struct AsmView {
  const TriGeo *__restrict__ geo;
  const QuadExt *__restrict__ gq;
  const int *__restrict__ sid;
  const int *__restrict__ colOff;
  cd *__restrict__ G, *__restrict__ H, *__restrict__ dG, *__restrict__ dH;
  size_t ld;
  double kr, ki;
};

__device__ __forceinline__ AsmView asm_view(const AsmJob &J)
{
  return {J.geo, J.gq, J.sid, J.colOff, J.G, J.H, J.dG, J.dH, J.ld, J.kr, J.ki};
}

__global__ void __launch_bounds__(256) k_asm_gl(const AsmJob *__restrict__ jobs, const int *__restrict__ first, int nj,
                                                const QPt *__restrict__ rule, int nq, const QPtQ *__restrict__ qrule,
                                                int nqq)
{
  const int j = job_of(first, nj, blockIdx.x);
  const AsmJob &J = jobs[j];
  const int loc = blockIdx.x - first[j], row = loc / J.nblk, xb = loc - row * J.nblk;
  const AsmView V = asm_view(J);
  const TriGeo *__restrict__ geo = V.geo;
  const QuadExt *__restrict__ gq = V.gq;
  const int *__restrict__ sid = V.sid;
  const int *__restrict__ colOff = V.colOff;
  cd *__restrict__ G = V.G, *__restrict__ H = V.H, *__restrict__ dG = V.dG, *__restrict__ dH = V.dH;
  const size_t ld = V.ld;
  const double kr = V.kr, ki = V.ki;
  const int dNe = J.dNe;
  double *__restrict__ dfPart = J.dfPart;
  unsigned long long *__restrict__ ghMask = J.ghMask;
  const RowPt rp = J.rows[row];
  const int s = xb * blockDim.x + threadIdx.x + 1;
  const int lane = threadIdx.x & 63;
  const int wv = (xb * blockDim.x + threadIdx.x) >> 6;
  const int nwv = J.nblk * (blockDim.x >> 6);
  bool active = s <= dNe && s != rp.t;
  bool flagged = false;
  double f = 0.0, df0 = 0.0, df1 = 0.0;
  if (active) {
    int sd = sid[s];
    int as = sd < 0 ? -sd : sd;
    const TriGeo g = geo[as];
    if (g.quad != 0.0) {
      flagged = asm_gl_quad(g, gq[as], rp, sd, kr, ki, qrule, nqq, BEM_CBD_SDF, row, colOff[s], ld, G, H, dG, dH, f,
                            df0, df1);
    } else {
      double Wdv0 = g.W[0] * rp.v0[0] + g.W[1] * rp.v0[1] + g.W[2] * rp.v0[2];
      double Wdv1 = g.W[0] * rp.v1[0] + g.W[1] * rp.v1[1] + g.W[2] * rp.v1[2];
      AsmAcc a;
      acc_zero(a);
      for (int p = 0; p < nq; p++) {
        const QPt qp = rule[p];
        asm_point(a, g, rp, Wdv0, Wdv1, kr, ki, qp.z, qp.e, qp.w);
      }
      acc_finish(a, g.aW, sd < 0);
      flagged = fabs(a.DF0) > BEM_CBD_SDF || fabs(a.DF1) > BEM_CBD_SDF;
      if (!flagged) {
        acc_store(a, row, colOff[s], ld, G, H, dG, dH);
        f = a.F; df0 = a.DF0; df1 = a.DF1;
      }
    }
  }
  unsigned long long m = __ballot(flagged);
  f = wsum(f); df0 = wsum(df0); df1 = wsum(df1);
  if (lane == 0) {
    size_t o = (size_t)row * nwv + wv;
    ghMask[o] = m;
    dfPart[o * 3 + 0] = f; dfPart[o * 3 + 1] = df0; dfPart[o * 3 + 2] = df1;
  }
}

// ---- pass 1, mixed-precision form ----
// Far pairs (row point farther than 4 element radii from the element's
// centre) take the reduced far rule the wave's far lanes need (far_rule:
// 4x4, 5x5, 6x6 or 12x12 squares of the 3-part map, by distance and |k| x
// element radius), with a rule error below 3e-8 there.  Nearer pairs take
// the reference's 12x12 rule.
// The arithmetic stays FP64: FP32 would put 1e-7 relative noise on the
// matrix entries, which the solved state carries on the total field, and
// the force on a weakly scattering object (a few per cent index contrast),
// which comes from the scattered part, then sees it magnified past 1e-6.
// Non-packed FP32 runs at the FP64 rate on CDNA4, so the reduced rule is
// where the time goes.  Pairs whose |dF/dt| comes within a relative 1e-4
// of the GHN switch level are flagged as in the FP32 form of the boundary
// sums: pass 2 redoes them and takes the reference's decision there.
__global__ void __launch_bounds__(256) k_asm_glr(const AsmJob *__restrict__ jobs, const int *__restrict__ first, int nj,
                                                 const QPt *__restrict__ rule, const QPtQ *__restrict__ qrule, int nqq)
{
  const int j = job_of(first, nj, blockIdx.x);
  const AsmJob &J = jobs[j];
  const int loc = blockIdx.x - first[j], row = loc / J.nblk, xb = loc - row * J.nblk;
  const AsmView V = asm_view(J);
  const TriGeo *__restrict__ geo = V.geo;
  const QuadExt *__restrict__ gq = V.gq;
  const int *__restrict__ sid = V.sid;
  const int *__restrict__ colOff = V.colOff;
  cd *__restrict__ G = V.G, *__restrict__ H = V.H, *__restrict__ dG = V.dG, *__restrict__ dH = V.dH;
  const size_t ld = V.ld;
  const double kr = V.kr, ki = V.ki;
  const int dNe = J.dNe;
  double *__restrict__ dfPart = J.dfPart;
  unsigned long long *__restrict__ ghMask = J.ghMask;
  const RowPt rp = J.rows[row];
  const int s = xb * blockDim.x + threadIdx.x + 1;
  const int lane = threadIdx.x & 63;
  const int wv = (xb * blockDim.x + threadIdx.x) >> 6;
  const int nwv = J.nblk * (blockDim.x >> 6);
  bool active = s <= dNe && s != rp.t;
  bool flagged = false;
  double f = 0.0, df0 = 0.0, df1 = 0.0;
  if (active && geo[sid[s] < 0 ? -sid[s] : sid[s]].quad != 0.0) {
    // quadrangles: the square's 12x12 rule
    const int sd = sid[s], as = sd < 0 ? -sd : sd;
    flagged = asm_gl_quad(geo[as], gq[as], rp, sd, kr, ki, qrule, nqq, BEM_CBD_SDF * (1.0 - 1e-4), row, colOff[s], ld, G,
                          H, dG, dH, f, df0, df1);
  } else if (active) {
    int sd = sid[s];
    int as = sd < 0 ? -sd : sd;
    const TriGeo g = geo[as];
    const double Wdv0 = g.W[0] * rp.v0[0] + g.W[1] * rp.v0[1] + g.W[2] * rp.v0[2];
    const double Wdv1 = g.W[0] * rp.v1[0] + g.W[1] * rp.v1[1] + g.W[2] * rp.v1[2];
    const double dx = g.c0[0] - rp.rt[0], dy = g.c0[1] - rp.rt[1], dz = g.c0[2] - rp.rt[2];
    const double rad2 = elem_rad2(g);
    const bool far = dx * dx + dy * dy + dz * dz > BEM_FAR_RADII * BEM_FAR_RADII * rad2;
    int qoff, nqf;
    far_rule(far, sqrt((kr * kr + ki * ki) * rad2), (dx * dx + dy * dy + dz * dz) / rad2, qoff, nqf);
    if (!far) { qoff = 0; nqf = BEM_FAR_OFF6; }   // the 12x12 rule heads the table
    AsmAcc a;
    acc_zero(a);
    for (int p = qoff; p < qoff + nqf; p++) {
      const QPt qp = rule[p];
      asm_point(a, g, rp, Wdv0, Wdv1, kr, ki, qp.z, qp.e, qp.w);
    }
    acc_finish(a, g.aW, sd < 0);
    const double lim = BEM_CBD_SDF * (1.0 - 1e-4);
    flagged = fabs(a.DF0) > lim || fabs(a.DF1) > lim;
    if (!flagged) {
      acc_store(a, row, colOff[s], ld, G, H, dG, dH);
      f = a.F; df0 = a.DF0; df1 = a.DF1;
    }
  }
  unsigned long long m = __ballot(flagged);
  f = wsum(f); df0 = wsum(df0); df1 = wsum(df1);
  if (lane == 0) {
    size_t o = (size_t)row * nwv + wv;
    ghMask[o] = m;
    dfPart[o * 3 + 0] = f; dfPart[o * 3 + 1] = df0; dfPart[o * 3 + 2] = df1;
  }
}

// ---- pass 2: flagged pairs with the 64x64 rule ----
// a flagged quadrangle pair: the FP64 12x12 rule first when the pair comes
// from the FP32 pass (qgl != nullptr), then the 64x64 rule
__device__ __forceinline__ void asm_gh_quad(const TriGeo &g, const QuadExt &qx, const RowPt &rp, int sd, double kr,
                                            double ki, const QPtQ *__restrict__ qgh, int nqgh,
                                            const QPtQ *__restrict__ qgl, int nqgl, int lane, size_t row, size_t col,
                                            size_t ld, cd *G, cd *H, cd *dG, cd *dH, double *df)
{
  AsmAccQ a;
  if (qgl) {
    accq_zero(a);
    for (int p = lane; p < nqgl; p += 64) asmq_point(a, g, qx, rp, kr, ki, qgl[p]);
    accq_wsum(a);
    accq_finish(a, sd < 0);
    if (!(fabs(a.DF0) > BEM_CBD_SDF || fabs(a.DF1) > BEM_CBD_SDF)) {
      if (lane == 0) {
        accq_store(a, row, col, ld, G, H, dG, dH);
        df[0] = a.F; df[1] = a.DF0; df[2] = a.DF1;
      }
      return;
    }
  }
  accq_zero(a);
  for (int p = lane; p < nqgh; p += 64) asmq_point(a, g, qx, rp, kr, ki, qgh[p]);
  accq_wsum(a);
  accq_finish(a, sd < 0);
  if (lane == 0) {
    accq_store(a, row, col, ld, G, H, dG, dH);
    df[0] = a.F; df[1] = a.DF0; df[2] = a.DF1;
  }
}

__global__ void __launch_bounds__(256) k_asm_gh(const AsmJob *__restrict__ jobs, const int *__restrict__ first, int nj,
                                                const QPt *__restrict__ rule, int nq, const QPt *__restrict__ gl, int ngl,
                                                const QPtQ *__restrict__ qgh, int nqgh, const QPtQ *__restrict__ qgl,
                                                int nqgl)
{
  // one wave per pair; the waves of a workgroup may belong to different jobs
  const int gp = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (gp >= first[nj]) return;
  const int j = job_of(first, nj, gp);
  const AsmJob &J = jobs[j];
  const int pid = gp - first[j];
  const AsmView V = asm_view(J);
  const TriGeo *__restrict__ geo = V.geo;
  const QuadExt *__restrict__ gq = V.gq;
  const int *__restrict__ sid = V.sid;
  const int *__restrict__ colOff = V.colOff;
  cd *__restrict__ G = V.G, *__restrict__ H = V.H, *__restrict__ dG = V.dG, *__restrict__ dH = V.dH;
  const size_t ld = V.ld;
  const double kr = V.kr, ki = V.ki;
  const GhPair *__restrict__ pairs = J.pairs;
  const RowPt *__restrict__ rows = J.rows;
  double *__restrict__ ghDF = J.ghDF;
  const GhPair pr = pairs[pid];
  const RowPt rp = rows[pr.row];
  int sd = sid[pr.s];
  int as = sd < 0 ? -sd : sd;
  const TriGeo g = geo[as];
  if (g.quad != 0.0) {
    asm_gh_quad(g, gq[as], rp, sd, kr, ki, qgh, nqgh, gl ? qgl : nullptr, nqgl, lane, pr.row, colOff[pr.s], ld, G, H, dG,
                dH, ghDF + (size_t)pid * 3);
    return;
  }
  double Wdv0 = g.W[0] * rp.v0[0] + g.W[1] * rp.v0[1] + g.W[2] * rp.v0[2];
  double Wdv1 = g.W[0] * rp.v1[0] + g.W[1] * rp.v1[1] + g.W[2] * rp.v1[2];
  AsmAcc a;
  if (gl) {
    // pairs flagged near the switch level by the FP32 pass: the FP64
    // 12x12 rule first, and the 64x64 rule only where it says so
    acc_zero(a);
    for (int p = lane; p < ngl; p += 64) {
      const QPt qp = gl[p];
      asm_point(a, g, rp, Wdv0, Wdv1, kr, ki, qp.z, qp.e, qp.w);
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
      a.G[i] = wsum(a.G[i]); a.H[i] = wsum(a.H[i]); a.DG0[i] = wsum(a.DG0[i]); a.DG1[i] = wsum(a.DG1[i]);
      a.DH0[i] = wsum(a.DH0[i]); a.DH1[i] = wsum(a.DH1[i]);
    }
    a.F = wsum(a.F); a.DF0 = wsum(a.DF0); a.DF1 = wsum(a.DF1);
    acc_finish(a, g.aW, sd < 0);
    if (!(fabs(a.DF0) > BEM_CBD_SDF || fabs(a.DF1) > BEM_CBD_SDF)) {   // the same value in every lane
      if (lane == 0) {
        acc_store(a, pr.row, colOff[pr.s], ld, G, H, dG, dH);
        ghDF[pid * 3 + 0] = a.F; ghDF[pid * 3 + 1] = a.DF0; ghDF[pid * 3 + 2] = a.DF1;
      }
      return;
    }
  }
  acc_zero(a);
  for (int p = lane; p < nq; p += 64) {
    const QPt qp = rule[p];
    asm_point(a, g, rp, Wdv0, Wdv1, kr, ki, qp.z, qp.e, qp.w);
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    a.G[i] = wsum(a.G[i]); a.H[i] = wsum(a.H[i]); a.DG0[i] = wsum(a.DG0[i]); a.DG1[i] = wsum(a.DG1[i]);
    a.DH0[i] = wsum(a.DH0[i]); a.DH1[i] = wsum(a.DH1[i]);
  }
  a.F = wsum(a.F); a.DF0 = wsum(a.DF0); a.DF1 = wsum(a.DF1);
  acc_finish(a, g.aW, sd < 0);
  if (lane == 0) {
    acc_store(a, pr.row, colOff[pr.s], ld, G, H, dG, dH);
    ghDF[pid * 3 + 0] = a.F; ghDF[pid * 3 + 1] = a.DF0; ghDF[pid * 3 + 2] = a.DF1;
  }
}

// ---- pass 3: self element ----
// node slot positions of the triangle rule (setup2.c:735-737)
__constant__ double c_tri_zt[4] = {-0.2, 0.4, -0.2, 0.0};
__constant__ double c_tri_et[4] = {-0.3464101615137754587055, 0.0, 0.3464101615137754587055, 0.0};

__device__ __forceinline__ void tri_dN(double c, double s, double &d0, double &d1, double &d2)
{
  // derivative of the interpolation functions along (cos, sin), lit_dcoef.c:844-855
  d0 = -2.0 / 3.0 * 2.5 * (c + BEM_SQ3 * s);
  d1 = 4.0 / 3.0 * 2.5 * c;
  d2 = -2.0 / 3.0 * 2.5 * (c - BEM_SQ3 * s);
}

__device__ __forceinline__ void csincos(cd a, cd &sn, cd &cs)
{
  double s, c;
  sincos(a.x, &s, &c);
  if (a.y == 0.0) { sn = mk(s, 0.0); cs = mk(c, 0.0); return; }
  double ch = cosh(a.y), sh = sinh(a.y);
  sn = mk(s * ch, c * sh);
  cs = mk(c * ch, -s * sh);
}

// ---- self quadrangle ----
// node slot positions on the square (d3b1_const.h:50, the 2x2 Gauss points)
__constant__ double c_quad_zt[4] = {-0.5773502691896257645091, 0.5773502691896257645091, 0.5773502691896257645091,
                                    -0.5773502691896257645091};
__constant__ double c_quad_et[4] = {-0.5773502691896257645091, -0.5773502691896257645091, 0.5773502691896257645091,
                                    0.5773502691896257645091};

// the map at polar (rr, c, sn) about (zt, et): K = dr/drr, W, |K|, |W|,
// K.t0, K.t1 and the four shape functions
struct QPol {
  double aK, aW, Kt0, Kt1, n[4];
};
__device__ __forceinline__ QPol qpol(const TriGeo &g, const QuadExt &qx, const double *t0, const double *t1, double zt,
                                     double et, double rr, double c, double sn)
{
  QPol p;
  const double z = rr * c + zt, e = rr * sn + et, cs = rr * c * sn + zt * sn + et * c;
  double K[3], W[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    K[d] = g.c1[d] * c + g.c2[d] * sn + qx.c3[d] * cs;
    W[d] = g.W[d] + qx.W1[d] * z + qx.W2[d] * e;
  }
  p.aK = sqrt(K[0] * K[0] + K[1] * K[1] + K[2] * K[2]);
  p.aW = sqrt(W[0] * W[0] + W[1] * W[1] + W[2] * W[2]);
  p.Kt0 = K[0] * t0[0] + K[1] * t0[1] + K[2] * t0[2];
  p.Kt1 = K[0] * t1[0] + K[1] * t1[1] + K[2] * t1[2];
  const double ia = 1.7320508075688772935;   // 1 / (1/sqrt(3))
  p.n[0] = 0.25 * (1.0 - z * ia) * (1.0 - e * ia);
  p.n[1] = 0.25 * (1.0 + z * ia) * (1.0 - e * ia);
  p.n[2] = 0.25 * (1.0 + z * ia) * (1.0 + e * ia);
  p.n[3] = 0.25 * (1.0 - z * ia) * (1.0 + e * ia);
  return p;
}

// Self quadrangle of one row, one wavefront: G (and H, F on a warped
// quadrangle) by the polar rule of bil_coef_bd (bil_coef.c:244-342), and the
// tangential derivatives of bil_dcoef_bd_tz0_PV / te0_PV (bil_dcoef.c:369-740
// with the integrands of :916-1631): the complement of the disk D_eps (4
// edges x 2 x 64 angles, 64 radial points from DEPS to the edge), the disk
// (4 quarters x 64 angles, 12 radial points) and its principal-value half
// (pairs at +r and -r over the upper two quarters).  Direction 0 is t_zeta,
// 1 is t_eta (unnormalised tangents at the node; results divided by their
// lengths).  H, F and their derivatives carry the factor c1 . W2 and vanish
// on a plane quadrangle (bil_check_plane).
__device__ void asm_self_quad(const TriGeo &g, const QuadExt &qx, double zt, double et, int td, int row, size_t col,
                              double kr, double ki, const double *__restrict__ xl, const double *__restrict__ wl,
                              const double *__restrict__ xh, const double *__restrict__ wh, int lane, cd *G, cd *H,
                              cd *dG, cd *dH, size_t ld, double *selfDF)
{
  double zp = 1.0 - zt, zm = -1.0 - zt, ep = 1.0 - et, em = -1.0 - et;
  double r[5] = {sqrt(zm * zm + em * em), sqrt(zp * zp + em * em), sqrt(zp * zp + ep * ep), sqrt(zm * zm + ep * ep), 0.0};
  double th[5] = {atan2(em, zm), atan2(em, zp), atan2(ep, zp), atan2(ep, zm), 0.0};
  r[4] = r[0];
  th[4] = th[0] + 2.0 * M_PI;
  double t0[3], t1[3];
#pragma unroll
  for (int d = 0; d < 3; d++) { t0[d] = g.c1[d] + qx.c3[d] * et; t1[d] = g.c2[d] + qx.c3[d] * zt; }
  const double iaT0 = 1.0 / sqrt(t0[0] * t0[0] + t0[1] * t0[1] + t0[2] * t0[2]);
  const double iaT1 = 1.0 / sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
  const double bdC = g.c1[0] * qx.W2[0] + g.c1[1] * qx.W2[1] + g.c1[2] * qx.W2[2];
  const double ab = sqrt(g.c1[0] * g.c1[0] + g.c1[1] * g.c1[1] + g.c1[2] * g.c1[2]);
  const double aC = sqrt(qx.W2[0] * qx.W2[0] + qx.W2[1] * qx.W2[1] + qx.W2[2] * qx.W2[2]);
  const bool warped = !(fabs(bdC / (ab * aC)) < BEM_GEPS);   // bil_check_plane (0/0 counts as warped)
  const double eps = BEM_DEPS;
  const cd kk = mk(kr, ki), ik = mk(-ki, kr);
  cd sg[4], sh[4], g0[4], g1[4], h0[4], h1[4];
#pragma unroll
  for (int a = 0; a < 4; a++) { sg[a] = sh[a] = g0[a] = g1[a] = h0[a] = h1[a] = mk(0, 0); }
  double sf = 0.0, f0 = 0.0, f1 = 0.0;
  // complement of D_eps, and the G (H, F) integrals over the whole element
  for (int ia = lane; ia < 512; ia += 64) {
    const int j = ia >> 7, m = (ia & 127) >> 1, half = ia & 1;
    const double dth = th[j + 1] - th[j];
    const double t = 0.25 * dth * xh[m] + (half ? 0.25 * (3.0 * th[j + 1] + th[j]) : 0.25 * (th[j + 1] + 3.0 * th[j]));
    const double rm = r[j] * r[j + 1] * sin(th[j + 1] - th[j]) / (r[j] * sin(t - th[j]) - r[j + 1] * sin(t - th[j + 1]));
    double c, sn;
    sincos(t, &sn, &c);
    const double wa = wh[m] * 0.25 * dth;
    // coefficients: radial GLN rule on (0, rm)
    for (int i = 0; i < 12; i++) {
      const double rr = 0.5 * rm * (xl[i] + 1.0);
      const QPol p = qpol(g, qx, t0, t1, zt, et, rr, c, sn);
      const double w = wl[i] * wa * 0.5 * rm, R = p.aK * rr;
      const cd ce = eikr(kr, ki, R);
      const cd vg = ce * (p.aW / p.aK * w);
#pragma unroll
      for (int a = 0; a < 4; a++) sg[a] = sg[a] + vg * p.n[a];
      if (warped) {
        const double w3 = w * c * sn / (p.aK * p.aK * p.aK);
        const cd vh = (mk(-1.0 - ki * R, kr * R) * ce) * w3;
#pragma unroll
        for (int a = 0; a < 4; a++) sh[a] = sh[a] + vh * p.n[a];
        sf += w3;
      }
    }
    // derivatives: radial GHN rule on (eps, rm)
    for (int i = 0; i < 64; i++) {
      const double rr = 0.5 * (rm - eps) * xh[i] + 0.5 * (rm + eps);
      const QPol p = qpol(g, qx, t0, t1, zt, et, rr, c, sn);
      const double w = wh[i] * wa * 0.5 * (rm - eps), R = p.aK * rr;
      const double aK2 = p.aK * p.aK, iK5r = 1.0 / (aK2 * aK2 * p.aK * rr);
      const cd ce = eikr(kr, ki, R);
      const cd cb = mk(-1.0 - ki * R, kr * R);
      const cd vg = (cb * ce) * (p.aW / (aK2 * p.aK * rr) * w);
#pragma unroll
      for (int a = 0; a < 4; a++) { g0[a] = g0[a] + vg * (p.Kt0 * p.n[a]); g1[a] = g1[a] + vg * (p.Kt1 * p.n[a]); }
      if (warped) {
        const cd kR = mk(kr * R, ki * R);
        const cd q3 = kR * kR + cb * 3.0;
        const cd e0 = (ce * (iK5r * w * sn)) * (q3 * (p.Kt0 * c) - cb * aK2);
        const cd e1 = (ce * (iK5r * w * c)) * (q3 * (p.Kt1 * sn) - cb * aK2);
#pragma unroll
        for (int a = 0; a < 4; a++) { h0[a] = h0[a] + e0 * p.n[a]; h1[a] = h1[a] + e1 * p.n[a]; }
        f0 += iK5r * w * sn * (3.0 * p.Kt0 * c - aK2);
        f1 += iK5r * w * c * (3.0 * p.Kt1 * sn - aK2);
      }
    }
  }
  // D_eps (x i for G) and its principal-value half (subtracted from G and
  // H, added to F)
  for (int ia = lane; ia < 256; ia += 64) {
    const int q = ia >> 6, m = ia & 63;
    const double t = 0.25 * M_PI * xh[m] + (-0.75 + 0.5 * q) * M_PI;
    double c, sn;
    sincos(t, &sn, &c);
    const double wa = wh[m] * 0.25 * M_PI * 0.5 * eps;
    for (int i = 0; i < 12; i++) {
      const double rr = 0.5 * eps * (xl[i] + 1.0), w = wl[i] * wa;
      const QPol p = qpol(g, qx, t0, t1, zt, et, rr, c, sn);
      const double R = p.aK * rr, aK2 = p.aK * p.aK;
      const cd kR = mk(kr * R, ki * R);
      cd skr, ckr;
      csincos(kR, skr, ckr);
      const cd cb = mk(-1.0 - ki * R, kr * R);
      const cd vg = mk(0, 1) * ((cb * skr + kR * ckr) * (p.aW / (aK2 * p.aK * rr) * w));
#pragma unroll
      for (int a = 0; a < 4; a++) { g0[a] = g0[a] + vg * (p.Kt0 * p.n[a]); g1[a] = g1[a] + vg * (p.Kt1 * p.n[a]); }
      if (warped) {
        // sdH_De: (De0 - i De1) with De0 = e^{ikR}/aK^4 ((k kR + 3ik) K.t ang - ik aK^2),
        // De1 = sin(kR)/(aK^5 r) (3 K.t ang - aK^2)
        const cd ce = eikr(kr, ki, R);
        const cd kkR = kk * kR + ik * 3.0;
        const double iK4 = 1.0 / (aK2 * aK2), iK5r = iK4 / (p.aK * rr);
        const cd e0 = (ce * (iK4 * w * sn)) * (kkR * (p.Kt0 * c) - ik * aK2) -
                      mk(0, 1) * (skr * (iK5r * w * sn * (3.0 * p.Kt0 * c - aK2)));
        const cd e1 = (ce * (iK4 * w * c)) * (kkR * (p.Kt1 * sn) - ik * aK2) -
                      mk(0, 1) * (skr * (iK5r * w * c * (3.0 * p.Kt1 * sn - aK2)));
#pragma unroll
        for (int a = 0; a < 4; a++) { h0[a] = h0[a] + e0 * p.n[a]; h1[a] = h1[a] + e1 * p.n[a]; }
      }
      if (q >= 2) {
        const QPol pm = qpol(g, qx, t0, t1, zt, et, -rr, c, sn);
        cd sp, cp, smv, cmv;
        csincos(mk(kr * p.aK * rr, ki * p.aK * rr), sp, cp);
        csincos(mk(kr * pm.aK * rr, ki * pm.aK * rr), smv, cmv);
        const double ir = w / rr;
        const cd up = cp * (p.aW / (aK2 * p.aK) * ir), um = cmv * (pm.aW / (pm.aK * pm.aK * pm.aK) * ir);
#pragma unroll
        for (int a = 0; a < 4; a++) {
          g0[a] = g0[a] - (up * (p.Kt0 * p.n[a]) - um * (pm.Kt0 * pm.n[a]));
          g1[a] = g1[a] - (up * (p.Kt1 * p.n[a]) - um * (pm.Kt1 * pm.n[a]));
        }
        if (warped) {
          const double pk2 = pm.aK * pm.aK, iP5 = 1.0 / (aK2 * aK2 * p.aK), iM5 = 1.0 / (pk2 * pk2 * pm.aK);
          const double ap0 = iP5 * (3.0 * p.Kt0 * c - aK2), am0 = iM5 * (3.0 * pm.Kt0 * c - pk2);
          const double ap1 = iP5 * (3.0 * p.Kt1 * sn - aK2), am1 = iM5 * (3.0 * pm.Kt1 * sn - pk2);
#pragma unroll
          for (int a = 0; a < 4; a++) {
            h0[a] = h0[a] - (cp * (ap0 * p.n[a]) - cmv * (am0 * pm.n[a])) * (ir * sn);
            h1[a] = h1[a] - (cp * (ap1 * p.n[a]) - cmv * (am1 * pm.n[a])) * (ir * c);
          }
          f0 += (ap0 - am0) * ir * sn;
          f1 += (ap1 - am1) * ir * c;
        }
      }
    }
  }
  const bool flip = td < 0;
  const double sH = (flip ? -BEM_I_FP : BEM_I_FP) * bdC;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const cd ga = wsum(sg[a]), d0 = wsum(g0[a]), d1 = wsum(g1[a]);
    const cd ha = wsum(sh[a]), e0 = wsum(h0[a]), e1 = wsum(h1[a]);
    if (lane == 0) {
      G[(size_t)row * ld + col + a] = ga * BEM_I_FP;
      dG[((size_t)row * 2 + 0) * ld + col + a] = d0 * (-BEM_I_FP * iaT0);
      dG[((size_t)row * 2 + 1) * ld + col + a] = d1 * (-BEM_I_FP * iaT1);
      H[(size_t)row * ld + col + a] = warped ? ha * sH : mk(0, 0);
      dH[((size_t)row * 2 + 0) * ld + col + a] = warped ? e0 * (sH * iaT0) : mk(0, 0);
      dH[((size_t)row * 2 + 1) * ld + col + a] = warped ? e1 * (sH * iaT1) : mk(0, 0);
    }
  }
  sf = wsum(sf); f0 = wsum(f0); f1 = wsum(f1);
  if (lane == 0) {
    selfDF[(size_t)row * 3 + 0] = warped ? sf * sH : 0.0;
    selfDF[(size_t)row * 3 + 1] = warped ? f0 * sH * iaT0 : 0.0;
    selfDF[(size_t)row * 3 + 2] = warped ? f1 * sH * iaT1 : 0.0;
  }
}

// Self triangle at (zt, et), one wavefront: G by the polar rule of
// lit_coef_bd (lit_coef.c:290-348) and the principal-value tangential
// derivatives dG/dt_zeta, dG/dt_eta of lit_dcoef_bd_tz0_PV / te0_PV
// (lit_dcoef.c:460-635): the complement of the disk D_eps by the polar
// rule, the disk by its closed-form radial factor.  H, F and their
// derivatives vanish on a flat triangle.  Every lane returns the wave's
// scaled sums.
__device__ void self_tri(const TriGeo &g, double zt, double et, double kr, double ki, const double *__restrict__ xl,
                         const double *__restrict__ wl, const double *__restrict__ xh, const double *__restrict__ wh,
                         int lane, cd *Gs, cd *D0s, cd *D1s)
{
  // polar set-up about (zt, et): vertices (-1/2,-sqrt3/2), (1,0), (-1/2,sqrt3/2)
  double zp = 1.0 - zt, zm = -0.5 - zt, ep = 0.5 * BEM_SQ3 - et, em = -0.5 * BEM_SQ3 - et;
  double r[4] = {sqrt(zm * zm + em * em), sqrt(zp * zp + et * et), sqrt(zm * zm + ep * ep), 0.0};
  double th[4] = {atan2(em, zm), atan2(-et, zp), atan2(ep, zm), 0.0};
  r[3] = r[0];
  th[3] = th[0] + 2.0 * M_PI;
  const double eps = BEM_DEPS;
  double iaT0 = 1.0 / sqrt(g.c1[0] * g.c1[0] + g.c1[1] * g.c1[1] + g.c1[2] * g.c1[2]);
  double iaT1 = 1.0 / sqrt(g.c2[0] * g.c2[0] + g.c2[1] * g.c2[1] + g.c2[2] * g.c2[2]);
  cd Ga[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
  cd D0[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, D1[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
  // complement of D_eps and the plain G integral share the angular points:
  // 3 edges x 64 GH nodes x 2 halves = 384 angles, 6 per lane
  for (int ia = lane; ia < 384; ia += 64) {
    int j = ia / 128, m = (ia % 128) >> 1, half = ia & 1;
    double dth = th[j + 1] - th[j];
    double t = 0.25 * dth * xh[m] + (half ? 0.25 * (3.0 * th[j + 1] + th[j]) : 0.25 * (th[j + 1] + 3.0 * th[j]));
    double rm = r[j] * r[j + 1] * sin(th[j + 1] - th[j]) / (r[j] * sin(t - th[j]) - r[j + 1] * sin(t - th[j + 1]));
    double c, sn;
    sincos(t, &sn, &c);
    double K[3] = {g.c1[0] * c + g.c2[0] * sn, g.c1[1] * c + g.c2[1] * sn, g.c1[2] * c + g.c2[2] * sn};
    double aK = sqrt(K[0] * K[0] + K[1] * K[1] + K[2] * K[2]), iaK = 1.0 / aK;
    double Kd0 = (K[0] * g.c1[0] + K[1] * g.c1[1] + K[2] * g.c1[2]) * iaK;
    double Kd1 = (K[0] * g.c2[0] + K[1] * g.c2[1] + K[2] * g.c2[2]) * iaK;
    double wang = wh[m] * 0.25 * dth;
    // G: radial GLN rule on (0, rm)
    cd sg[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    for (int i = 0; i < 12; i++) {
      double rr = 0.5 * rm * (xl[i] + 1.0);
      double n0, n1, n2;
      tri_N(rr * c + zt, rr * sn + et, n0, n1, n2);
      cd v = eikr(kr, ki, aK * rr) * (iaK * wl[i]);
      fma_cr(sg[0], v, n0); fma_cr(sg[1], v, n1); fma_cr(sg[2], v, n2);
    }
    // dG, complement of D_eps: radial GHN rule on (eps, rm)
    cd sd3[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    for (int i = 0; i < 64; i++) {
      double rr = 0.5 * (rm - eps) * xh[i] + 0.5 * (rm + eps);
      double n0, n1, n2;
      tri_N(rr * c + zt, rr * sn + et, n0, n1, n2);
      double R = aK * rr;
      cd cb = mk(-1.0 - ki * R, kr * R);
      cd v = (cb * eikr(kr, ki, R)) * (iaK * iaK / rr * wh[i]);
      fma_cr(sd3[0], v, n0); fma_cr(sd3[1], v, n1); fma_cr(sd3[2], v, n2);
    }
    double fg = wang * 0.5 * rm, fd = wang * 0.5 * (rm - eps);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      fma_cr(Ga[a], sg[a], fg);
      fma_cr(D0[a], sd3[a], fd * Kd0);
      fma_cr(D1[a], sd3[a], fd * Kd1);
    }
  }
  // D_eps: closed-form radial factor, 4 quarters x 64 nodes = 256 angles
  cd kk = mk(kr, ki);
  cd pe0[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, pe1[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
  cd pp0[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, pp1[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
  for (int ia = lane; ia < 256; ia += 64) {
    int q = ia >> 6, m = ia & 63;
    double t = 0.25 * M_PI * xh[m] + (-0.75 + 0.5 * q) * M_PI;
    double c, sn;
    sincos(t, &sn, &c);
    double K[3] = {g.c1[0] * c + g.c2[0] * sn, g.c1[1] * c + g.c2[1] * sn, g.c1[2] * c + g.c2[2] * sn};
    double aK = sqrt(K[0] * K[0] + K[1] * K[1] + K[2] * K[2]), iaK = 1.0 / aK;
    double Kd0 = (K[0] * g.c1[0] + K[1] * g.c1[1] + K[2] * g.c1[2]) * iaK;
    double Kd1 = (K[0] * g.c2[0] + K[1] * g.c2[1] + K[2] * g.c2[2]) * iaK;
    double d[3];
    tri_dN(c, sn, d[0], d[1], d[2]);
    cd ke = kk * (eps * aK), sk, ck;
    csincos(ke, sk, ck);
    cd fe = (ke + mk(0, 1)) * sk + (mk(2, 0) - mk(0, 1) * ke) * ck - mk(2, 0);
    double ia3 = iaK * iaK * iaK * wh[m];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      fma_cr(pe0[a], fe, d[a] * ia3 * Kd0);
      fma_cr(pe1[a], fe, d[a] * ia3 * Kd1);
      if (q >= 2) { fma_cr(pp0[a], sk, d[a] * ia3 * Kd0); fma_cr(pp1[a], sk, d[a] * ia3 * Kd1); }
    }
  }
  const double sG = BEM_I_SPSQ3 * g.aW;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    cd ga = wsum(Ga[a]);
    cd d0 = wsum(D0[a]), d1 = wsum(D1[a]);
    cd e0 = wsum(pe0[a]), e1 = wsum(pe1[a]), p0 = wsum(pp0[a]), p1 = wsum(pp1[a]);
    // + i tde (pi/4) / k - tdp (pi/4) / k
    d0 = d0 + cdiv(mk(0, 1) * e0 * (0.25 * M_PI) - p0 * (0.25 * M_PI), kk);
    d1 = d1 + cdiv(mk(0, 1) * e1 * (0.25 * M_PI) - p1 * (0.25 * M_PI), kk);
    Gs[a] = ga * sG;
    D0s[a] = d0 * (-sG * iaT0);
    D1s[a] = d1 * (-sG * iaT1);
  }
}

__global__ void __launch_bounds__(256) k_asm_self(const AsmJob *__restrict__ jobs, const int *__restrict__ first, int nj,
                                                  const double *__restrict__ xl, const double *__restrict__ wl,
                                                  const double *__restrict__ xh, const double *__restrict__ wh)
{
  // one wave per row; the waves of a workgroup may belong to different jobs
  const int gr = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (gr >= first[nj]) return;
  const int j = job_of(first, nj, gr);
  const AsmJob &J = jobs[j];
  const int row = gr - first[j];
  const AsmView V = asm_view(J);
  const TriGeo *__restrict__ geo = V.geo;
  const QuadExt *__restrict__ gq = V.gq;
  const int *__restrict__ sid = V.sid;
  const int *__restrict__ colOff = V.colOff;
  cd *__restrict__ G = V.G, *__restrict__ H = V.H, *__restrict__ dG = V.dG, *__restrict__ dH = V.dH;
  const size_t ld = V.ld;
  const double kr = V.kr, ki = V.ki;
  const RowPt *__restrict__ rows = J.rows;
  double *__restrict__ selfDF = J.selfDF;
  const RowPt rp = rows[row];
  int td = sid[rp.t];
  int at = td < 0 ? -td : td;
  const TriGeo g = geo[at];
  if (g.quad != 0.0) {
    asm_self_quad(g, gq[at], c_quad_zt[rp.slot], c_quad_et[rp.slot], td, row, colOff[rp.t], kr, ki, xl, wl, xh, wh,
                  lane, G, H, dG, dH, ld, selfDF);
    return;
  }
  if (lane == 0) {   // a flat triangle has no self F terms
    selfDF[(size_t)row * 3 + 0] = 0.0; selfDF[(size_t)row * 3 + 1] = 0.0; selfDF[(size_t)row * 3 + 2] = 0.0;
  }
  cd Gs[3], D0s[3], D1s[3];
  self_tri(g, c_tri_zt[rp.slot], c_tri_et[rp.slot], kr, ki, xl, wl, xh, wh, lane, Gs, D0s, D1s);
  size_t col = colOff[rp.t];
  if (lane == 0) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      G[(size_t)row * ld + col + a] = Gs[a];
      dG[((size_t)row * 2 + 0) * ld + col + a] = D0s[a];
      dG[((size_t)row * 2 + 1) * ld + col + a] = D1s[a];
      H[(size_t)row * ld + col + a] = mk(0, 0);
      dH[((size_t)row * 2 + 0) * ld + col + a] = mk(0, 0);
      dH[((size_t)row * 2 + 1) * ld + col + a] = mk(0, 0);
    }
  }
}

// ---- pass 4: per-row sums and the diagonal term ----
__global__ void __launch_bounds__(256) k_asm_fin(const AsmJob *__restrict__ jobs, const int *__restrict__ first, int nj)
{
  const int gr = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (gr >= first[nj]) return;
  const int j = job_of(first, nj, gr);
  const AsmJob &J = jobs[j];
  const int row = gr - first[j];
  const int nwv = J.nblk * 4, open_domain = J.open, r0 = J.r0;
  const size_t ld = J.ld;
  const double *__restrict__ dfPart = J.dfPart;
  const double *__restrict__ ghDF = J.ghDF;
  const int *__restrict__ ghStart = J.ghStart;
  const double *__restrict__ selfDF = J.selfDF;
  cd *__restrict__ H = J.H;
  double *__restrict__ dF = J.dF;
  double f = 0.0, f0 = 0.0, f1 = 0.0;
  for (int w = lane; w < nwv; w += 64) {
    const double *p = dfPart + ((size_t)row * nwv + w) * 3;
    f += p[0]; f0 += p[1]; f1 += p[2];
  }
  for (int k = ghStart[row] + lane; k < ghStart[row + 1]; k += 64) {
    f += ghDF[k * 3]; f0 += ghDF[k * 3 + 1]; f1 += ghDF[k * 3 + 2];
  }
  f = wsum(f); f0 = wsum(f0); f1 = wsum(f1);
  if (lane == 0) {
    // the self element last (a warped quadrangle's F terms)
    f += selfDF[(size_t)row * 3]; f0 += selfDF[(size_t)row * 3 + 1]; f1 += selfDF[(size_t)row * 3 + 2];
    double diag = open_domain ? 1.0 + f : f;
    cd *h = H + (size_t)row * ld + row + r0;   // the row's own column (rows r0.. of the domain)
    h->x += diag;
    dF[row * 3 + 0] = diag; dF[row * 3 + 1] = f0; dF[row * 3 + 2] = f1;
  }
}

// ---- element layer: the self element at a point of its own ----
// Request i: signed element s[i] at local coordinates (ze[2i], ze[2i+1]);
// out[27 idx[i] + ...]: CC (G 0-3, H 4-7, F 8), dC/dt_zeta (9-17) and
// dC/dt_eta (18-26) of lit_coef_bd + lit_dcoef_bd_vt2_PV, or bil_coef_bd +
// bil_dcoef_bd_vt2_PV for a quadrangle (d3b1_elem.c:187-230, 400-480).  tmp
// holds 24 complex values and tmpF 3 doubles per request (the quadrangle
// routine's outputs).
__global__ void __launch_bounds__(256) k_elem_self(const TriGeo *__restrict__ geo, const QuadExt *__restrict__ gq,
                                                   const int *__restrict__ s, const double *__restrict__ ze,
                                                   const int *__restrict__ idx, int n, double kr, double ki,
                                                   const double *__restrict__ xl, const double *__restrict__ wl,
                                                   const double *__restrict__ xh, const double *__restrict__ wh,
                                                   cd *__restrict__ tmp, double *__restrict__ tmpF, cd *__restrict__ out)
{
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int sd = s[i], as = sd < 0 ? -sd : sd;
  const TriGeo g = geo[as];
  const double zt = ze[2 * i], et = ze[2 * i + 1];
  cd *o = out + 27 * (size_t)idx[i];
  if (g.quad != 0.0) {
    cd *T = tmp + 24 * (size_t)i;   // G 0-3, H 4-7, dG 8-15, dH 16-23 (t_zeta then t_eta)
    double *Fq = tmpF + 3 * (size_t)i;
    asm_self_quad(g, gq[as], zt, et, sd, 0, 0, kr, ki, xl, wl, xh, wh, lane, T, T + 4, T + 8, T + 16, 4, Fq);
    if (lane == 0) {
      for (int a = 0; a < 4; a++) {
        o[a] = T[a];
        o[4 + a] = T[4 + a];
        o[9 + a] = T[8 + a];
        o[13 + a] = T[16 + a];
        o[18 + a] = T[12 + a];
        o[22 + a] = T[20 + a];
      }
      o[8] = mk(Fq[0], 0.0);
      o[17] = mk(Fq[1], 0.0);
      o[26] = mk(Fq[2], 0.0);
    }
    return;
  }
  cd Gs[3], D0s[3], D1s[3];
  self_tri(g, zt, et, kr, ki, xl, wl, xh, wh, lane, Gs, D0s, D1s);
  if (lane == 0) {
    for (int j = 0; j < 27; j++) o[j] = mk(0, 0);
    for (int a = 0; a < 3; a++) {
      o[a] = Gs[a];
      o[9 + a] = D0s[a];
      o[18 + a] = D1s[a];
    }
  }
}

extern "C" {
// launchers used by the C-ABI layer (bem_api.hip)

// Passes of a batched assembly over nj jobs (AsmJob, device) with their
// prefix tables (device, nj + 1 ints each).  Pass 1: first1 over the jobs'
// workgroups (nrows x nblk each), nblocks = first1[nj]; pairs = rows x
// source elements of the batch (the registry's work).  mixed: the reduced
// far rules (rule: the far-pair table, qrule: the square's 12x12 rule).
void bem_launch_asm_p1(int mixed, const void *jobs, const int *first1, int nj, int nblocks, double pairs,
                       const void *rule, int nq, const void *qrule, int nqq, hipStream_t st)
{
  if (nblocks <= 0) return;
  const int kid = mixed ? BH_K_ASM_GLR : BH_K_ASM_GL;
  bem_prof_begin(kid, pairs, st);
  if (mixed)
    hipLaunchKernelGGL(k_asm_glr, dim3(nblocks), dim3(256), 0, st, (const AsmJob *)jobs, first1, nj, (const QPt *)rule,
                       (const QPtQ *)qrule, nqq);
  else
    hipLaunchKernelGGL(k_asm_gl, dim3(nblocks), dim3(256), 0, st, (const AsmJob *)jobs, first1, nj, (const QPt *)rule, nq,
                       (const QPtQ *)qrule, nqq);
  bem_prof_end(kid, st);
}

// Pass 2: the near pairs of every job (first2 over the pairs, npairs =
// first2[nj]) with the 64x64 rule (rule, nq; qgh, nqgh on quadrangles).
// gl != nullptr: the pairs come from the mixed form's pass and are first
// redone with the FP64 12x12 rule (gl, ngl; qgl, nqgl on quadrangles).
void bem_launch_asm_p2(const void *jobs, const int *first2, int nj, int npairs, const void *rule, int nq, const void *gl,
                       int ngl, const void *qgh, int nqgh, const void *qgl, int nqgl, hipStream_t st)
{
  if (npairs <= 0) return;
  bem_prof_begin(BH_K_ASM_GH, (double)npairs, st);
  hipLaunchKernelGGL(k_asm_gh, dim3((npairs + 3) / 4), dim3(256), 0, st, (const AsmJob *)jobs, first2, nj,
                     (const QPt *)rule, nq, (const QPt *)gl, ngl, (const QPtQ *)qgh, nqgh, (const QPtQ *)qgl, nqgl);
  bem_prof_end(BH_K_ASM_GH, st);
}

// FP32 copy (device) of a 3-part rule given on the host: point, weight and
// weight x shape functions (the far-pair table: the 12x12, 6x6 and 5x5 rules
// back to back)
int bem_rule32(const void *rule, int nq, void **out)
{
  const QPt *r = (const QPt *)rule;
  QPt32 *h = (QPt32 *)malloc(sizeof(QPt32) * (size_t)nq);
  if (!h) return -1;
  for (int p = 0; p < nq; p++) {
    const double S3 = BEM_SQ3, z = r[p].z, e = r[p].e, w = r[p].w;
    const double n0 = (1.0 - 2.5 * (z + S3 * e)) / 3.0, n1 = (1.0 + 5.0 * z) / 3.0, n2 = (1.0 - 2.5 * (z - S3 * e)) / 3.0;
    h[p] = {(float)z, (float)e, (float)w, (float)(w * n0), (float)(w * n1), (float)(w * n2)};
  }
  int rc = 0;
  if (hipMalloc(out, sizeof(QPt32) * (size_t)nq) != hipSuccess ||
      hipMemcpy(*out, h, sizeof(QPt32) * (size_t)nq, hipMemcpyHostToDevice) != hipSuccess)
    rc = -1;
  free(h);
  return rc;
}

// Passes 3 and 4: the self element of every row (its F terms into the
// job's selfDF), then the row sums of F and the diagonal term; first3 over
// the rows of the jobs, nrows = first3[nj]
void bem_launch_asm_p34(const void *jobs, const int *first3, int nj, int nrows, const double *xl, const double *wl,
                        const double *xh, const double *wh, hipStream_t st)
{
  if (nrows <= 0) return;
  const int nb = (nrows + 3) / 4;
  bem_prof_begin(BH_K_ASM_SELF, (double)nrows, st);
  hipLaunchKernelGGL(k_asm_self, dim3(nb), dim3(256), 0, st, (const AsmJob *)jobs, first3, nj, xl, wl, xh, wh);
  bem_prof_end(BH_K_ASM_SELF, st);
  hipLaunchKernelGGL(k_asm_fin, dim3(nb), dim3(256), 0, st, (const AsmJob *)jobs, first3, nj);
}

void bem_launch_elem_self(const void *geo, const void *gq, const int *s, const double *ze, const int *idx, int n,
                          double kr, double ki, const double *xl, const double *wl, const double *xh, const double *wh,
                          void *tmp, double *tmpF, void *out, hipStream_t st)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_elem_self, dim3((n + 3) / 4), dim3(256), 0, st, (const TriGeo *)geo, (const QuadExt *)gq, s, ze,
                     idx, n, kr, ki, xl, wl, xh, wh, (cd *)tmp, tmpF, (cd *)out);
}
}
