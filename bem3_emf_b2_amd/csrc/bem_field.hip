// bem_field.hip -- surface-integral field evaluation on linear triangles:
//   * k_domain      : domain of target points by solid angles
//                     (domain_id + fid_calc_solid_angle, setup2.c:871-978)
//   * k_field       : E, H (or U) at points from the boundary values of one
//                     object (EH_s / mEMP_s, d3b1_field.c:12-209) with the
//                     4-point, 7-point, GLN or GHN rule (coef_rt)
//   * k_scatter     : potentials and their normal derivative at the node
//                     slots of another object (mEMP2_s + add_scattered_mEMP,
//                     mo_solve.c:296-396), GLN with the per-pair GHN switch
//   * k_field_bd    : E, H at points on the object's own surface (EH_s_bd,
//                     d3b1_field.c:238-279): GLN for other elements, the polar
//                     rule (lit_coef_bd) for the element holding the point
//   * k_incident    : plane-wave boundary data (init_boundary_data_u2,
//                     mo_setup.c:704-757; calc_ipw_EH_dv, ipw.c:15-33)
//   * k_stress      : Maxwell-stress quadrature (mo_force.c:145-240)
// One wavefront per target point, lanes over source elements, wave
// reductions at the end.
#include "bem_dev.h"
#include "bem_prof.h"



__device__ __forceinline__ void tri_pos(const TriGeo &g, double z, double e, double *r)
{
#pragma unroll
  for (int d = 0; d < 3; d++) r[d] = fma(g.c2[d], e, fma(g.c1[d], z, g.c0[d]));
}

// G[3], H[3], F at rt from one triangle with a tabulated rule; optional
// derivative along v (dG, dH, dF).  scale = rule scale (aW applied to G).
template <bool DER, class RT = QPtN>
__device__ __forceinline__ void tri_coef(const TriGeo &g, const double *rt, const double *v, double kr, double ki,
                                         const RT *__restrict__ rule, int nq, double scale, bool flip, cd *G, cd *H,
                                         double &F, cd *dG, cd *dH, double &dF)
{
  cd aG[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, aH[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
  cd aDG[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, aDH[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
  double f = 0.0, df = 0.0;
  double Wdv = DER ? g.W[0] * v[0] + g.W[1] * v[1] + g.W[2] * v[2] : 0.0;
  for (int p = 0; p < nq; p++) {
    const RT q = rule[p];
    double r[3];
    tri_pos(g, q.z, q.e, r);
    double vx = r[0] - rt[0], vy = r[1] - rt[1], vz = r[2] - rt[2];
    double R = sqrt(vx * vx + vy * vy + vz * vz), iR = 1.0 / R, iR2 = iR * iR;
    double rdW = (vx * g.W[0] + vy * g.W[1] + vz * g.W[2]) * iR;
    cd ce = eikr(kr, ki, R);
    cd cb = mk(-1.0 - ki * R, kr * R);
    cd cbce = cb * ce;
    cd gk = ce * iR, hk = cbce * (rdW * iR2);
    double n0, n1, n2;
    rule_N(q, n0, n1, n2);
    double wn[3] = {q.w * n0, q.w * n1, q.w * n2};
#pragma unroll
    for (int i = 0; i < 3; i++) { fma_cr(aG[i], gk, wn[i]); fma_cr(aH[i], hk, wn[i]); }
    f = fma(q.w, rdW * iR2, f);
    if (DER) {
      double rd = (vx * v[0] + vy * v[1] + vz * v[2]) * iR, iR3 = iR2 * iR;
      cd tk = cbce * iR2;
      cd kR = mk(kr * R, ki * R);
      cd qq = kR * kR + cb * 3.0;
      cd dh = (ce * iR3) * (qq * (rd * rdW) - cb * Wdv);
#pragma unroll
      for (int i = 0; i < 3; i++) { fma_cr(aDG[i], tk, wn[i] * rd); fma_cr(aDH[i], dh, wn[i]); }
      df = fma(q.w, iR3 * (3.0 * rd * rdW - Wdv), df);
    }
  }
  double sh = flip ? -scale : scale;
#pragma unroll
  for (int i = 0; i < 3; i++) { G[i] = aG[i] * (scale * g.aW); H[i] = aH[i] * sh; }
  F = f * sh;
  if (DER) {
#pragma unroll
    for (int i = 0; i < 3; i++) { dG[i] = aDG[i] * (-scale * g.aW); dH[i] = aDH[i] * sh; }
    dF = df * sh;
  }
}

// G[3], H[3], F at rt from one triangle with the FP32 copy of a rule
// (mixed-precision form): position relative to rt formed in FP64, the
// integrand in FP32, sums over blocks of 12 points added in FP64 (PAIR32:
// the pair's sum in FP32 throughout, fewer registers, for the coupling sums
// whose error enters only through the other objects' scattered field);
// optional derivative along v (dG, dH, dF) as tri_coef
__device__ __forceinline__ void fma_cf(cf &a, cf b, float s) { a.x = fmaf(b.x, s, a.x); a.y = fmaf(b.y, s, a.y); }
template <bool DER, bool PAIR32 = false>
__device__ __forceinline__ void tri_coef32(const TriGeo &g, const double *rt, const double *v, double krd, double kid,
                                           const QPt32 *__restrict__ rule, int nq, double scale, bool flip, cd *G, cd *H,
                                           double &F, cd *dG, cd *dH, double &dF)
{
  const float kr = (float)krd, ki = (float)kid;
  const float d0x = (float)(g.c0[0] - rt[0]), d0y = (float)(g.c0[1] - rt[1]), d0z = (float)(g.c0[2] - rt[2]);
  const float c1x = (float)g.c1[0], c1y = (float)g.c1[1], c1z = (float)g.c1[2];
  const float c2x = (float)g.c2[0], c2y = (float)g.c2[1], c2z = (float)g.c2[2];
  const float vx0 = DER ? (float)v[0] : 0.0f, vy0 = DER ? (float)v[1] : 0.0f, vz0 = DER ? (float)v[2] : 0.0f;
  const float Wdv = DER ? (float)(g.W[0] * v[0] + g.W[1] * v[1] + g.W[2] * v[2]) : 0.0f;
  cd aG[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, aH[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
  cd aDG[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, aDH[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
  double f = 0.0, df = 0.0;
  float Rc;
  const cf ec = far32_centre(g.c0[0] - rt[0], g.c0[1] - rt[1], g.c0[2] - rt[2], krd, kid, Rc);
  // (r - rt).W is the same at every point of a flat triangle: formed once in
  // FP64, so that coplanar pairs keep H = F = 0 exactly
  const float dW = (float)((g.c0[0] - rt[0]) * g.W[0] + (g.c0[1] - rt[1]) * g.W[1] + (g.c0[2] - rt[2]) * g.W[2]);
  const int blk = PAIR32 ? nq : 12;
  for (int p0 = 0; p0 < nq; p0 += blk) {
    cf bG[3] = {mkf(0, 0), mkf(0, 0), mkf(0, 0)}, bH[3] = {mkf(0, 0), mkf(0, 0), mkf(0, 0)};
    cf bDG[3] = {mkf(0, 0), mkf(0, 0), mkf(0, 0)}, bDH[3] = {mkf(0, 0), mkf(0, 0), mkf(0, 0)};
    float bf = 0.0f, bdf = 0.0f;
    const int p1 = min(p0 + blk, nq);
#pragma unroll 2
    for (int p = p0; p < p1; p++) {
      const QPt32 q = rule[p];
      const float lx = fmaf(c2x, q.e, c1x * q.z), ly = fmaf(c2y, q.e, c1y * q.z), lz = fmaf(c2z, q.e, c1z * q.z);
      const float vx = d0x + lx, vy = d0y + ly, vz = d0z + lz;
      const float R2 = fmaf(vx, vx, fmaf(vy, vy, vz * vz));
      const float iR = rsqrtf(R2), R = R2 * iR, iR2 = iR * iR;
      const float rdW = dW * iR;
      const float dR = fmaf(2.0f, fmaf(d0x, lx, fmaf(d0y, ly, d0z * lz)), fmaf(lx, lx, fmaf(ly, ly, lz * lz))) / (R + Rc);
      const cf ce = far32_phase(ec, dR, kr, ki);
      const cf cb = mkf(-1.0f - ki * R, kr * R);
      const cf cbce = cb * ce;
      const cf gk = ce * iR, hk = cbce * (rdW * iR2);
      const float wn[3] = {q.wn0, q.wn1, q.wn2};
#pragma unroll
      for (int i = 0; i < 3; i++) { fma_cf(bG[i], gk, wn[i]); fma_cf(bH[i], hk, wn[i]); }
      bf = fmaf(q.w, rdW * iR2, bf);
      if (DER) {
        const float rd = fmaf(vx, vx0, fmaf(vy, vy0, vz * vz0)) * iR, iR3 = iR2 * iR;
        const cf tk = cbce * iR2;
        const cf kR = mkf(kr * R, ki * R);
        const cf qq = kR * kR + cb * 3.0f;
        const cf dh = (ce * iR3) * (qq * (rd * rdW) - cb * Wdv);
#pragma unroll
        for (int i = 0; i < 3; i++) { fma_cf(bDG[i], tk, wn[i] * rd); fma_cf(bDH[i], dh, wn[i]); }
        bdf = fmaf(q.w, iR3 * fmaf(3.0f * rd, rdW, -Wdv), bdf);
      }
    }
    if constexpr (PAIR32) {   // one block: the pair's FP32 sums, converted once
#pragma unroll
      for (int i = 0; i < 3; i++) { aG[i] = to_cd(bG[i]); aH[i] = to_cd(bH[i]); }
      f = (double)bf;
      if (DER) {
#pragma unroll
        for (int i = 0; i < 3; i++) { aDG[i] = to_cd(bDG[i]); aDH[i] = to_cd(bDH[i]); }
        df = (double)bdf;
      }
      break;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) { aG[i] = aG[i] + to_cd(bG[i]); aH[i] = aH[i] + to_cd(bH[i]); }
    f += (double)bf;
    if (DER) {
#pragma unroll
      for (int i = 0; i < 3; i++) { aDG[i] = aDG[i] + to_cd(bDG[i]); aDH[i] = aDH[i] + to_cd(bDH[i]); }
      df += (double)bdf;
    }
  }
  double sh = flip ? -scale : scale;
#pragma unroll
  for (int i = 0; i < 3; i++) { G[i] = aG[i] * (scale * g.aW); H[i] = aH[i] * sh; }
  F = f * sh;
  if (DER) {
#pragma unroll
    for (int i = 0; i < 3; i++) { dG[i] = aDG[i] * (-scale * g.aW); dH[i] = aDH[i] * sh; }
    dF = df * sh;
  }
}

// ---------------------------------------------------------------- quadrangles
__device__ __forceinline__ void quad_geom(const TriGeo &g, const QuadExt &q, double z, double e, double *r, double *W)
{
#pragma unroll
  for (int d = 0; d < 3; d++) {
    r[d] = g.c0[d] + g.c1[d] * z + g.c2[d] * e + q.c3[d] * z * e;
    W[d] = g.W[d] + q.W1[d] * z + q.W2[d] * e;
  }
}

// G[4], H[4], F at rt from one bi-linear quadrangle with a tabulated rule
// (bil_coef_4p / bil_coef_9p / bil_coef_GL / bil_coef_GH, bil_coef.c:25-202):
// G_a = 1/(4 pi) sum w N_a e^{ikR}/R |W|, H_a = 1/(4 pi) sum w N_a (ikR-1)
// e^{ikR} (R.W)/R^3, F = 1/(4 pi) sum w (R.W)/R^3, H and F negated on the
// sub-domain side
__device__ __forceinline__ void quad_coef(const TriGeo &g, const QuadExt &qx, const double *rt, double kr, double ki,
                                          const QPtQ *__restrict__ rule, int nq, bool flip, cd *G, cd *H, double &F)
{
  cd aG[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)}, aH[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  double f = 0.0;
  for (int p = 0; p < nq; p++) {
    const QPtQ q = rule[p];
    double r[3], W[3];
    quad_geom(g, qx, q.z, q.e, r, W);
    const double vx = r[0] - rt[0], vy = r[1] - rt[1], vz = r[2] - rt[2];
    const double R = sqrt(vx * vx + vy * vy + vz * vz), iR = 1.0 / R;
    const double aW = sqrt(W[0] * W[0] + W[1] * W[1] + W[2] * W[2]);
    const double tD = (vx * W[0] + vy * W[1] + vz * W[2]) * iR * iR * iR;
    const cd ce = eikr(kr, ki, R);
    const cd cb = mk(-1.0 - ki * R, kr * R);
    const cd gk = ce * (iR * aW), hk = (cb * ce) * tD;
#pragma unroll
    for (int a = 0; a < 4; a++) { fma_cr(aG[a], gk, q.w * q.n[a]); fma_cr(aH[a], hk, q.w * q.n[a]); }
    f = fma(q.w, tD, f);
  }
  const double sh = flip ? -BEM_I_FP : BEM_I_FP;
#pragma unroll
  for (int a = 0; a < 4; a++) { G[a] = aG[a] * BEM_I_FP; H[a] = aH[a] * sh; }
  F = f * sh;
}

// bi-linear shape functions on the square (node slots at the 2x2 Gauss
// points, d3b1_elem.c:556-568)
__device__ __forceinline__ void quad_N(double z, double e, double *n)
{
  const double ia = 1.7320508075688772935;   // 1 / (1/sqrt(3))
  n[0] = 0.25 * (1.0 - z * ia) * (1.0 - e * ia);
  n[1] = 0.25 * (1.0 + z * ia) * (1.0 - e * ia);
  n[2] = 0.25 * (1.0 + z * ia) * (1.0 + e * ia);
  n[3] = 0.25 * (1.0 - z * ia) * (1.0 + e * ia);
}

// G[4], H[4], F and their derivatives along ND directions v[d] (unit
// vectors) at rt from one bi-linear quadrangle with a tabulated rule: the
// bil_dcoef_4p/9p/GL/GH kernels (bil_dcoef.c:50-320) with the ND dcoef_rt
// calls of one point sharing one evaluation of the kernel per rule point.
// dG = -1/(4 pi) sum w N (ikR-1) e^{ikR} (R.v/R) |W| / R^2,
// dH = 1/(4 pi) sum w N e^{ikR}/R^3 ((kR^2 + 3(ikR-1)) (R.v/R)(R.W/R) - (ikR-1) W.v),
// dF = 1/(4 pi) sum w (3 (R.v/R)(R.W/R) - W.v)/R^3; H, F and their
// derivatives negated on the sub-domain side.
template <int ND>
__device__ __forceinline__ void quad_coef_dir(const TriGeo &g, const QuadExt &qx, const double *rt, const double (*v)[3],
                                              double kr, double ki, const QPtQ *__restrict__ rule, int nq, bool flip,
                                              cd *G, cd *H, double &F, cd (*dG)[4], cd (*dH)[4], double *dF)
{
  cd aG[4], aH[4], aDG[ND][4], aDH[ND][4];
  double f = 0.0, df[ND];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    aG[a] = mk(0, 0); aH[a] = mk(0, 0);
#pragma unroll
    for (int d = 0; d < ND; d++) { aDG[d][a] = mk(0, 0); aDH[d][a] = mk(0, 0); }
  }
#pragma unroll
  for (int d = 0; d < ND; d++) df[d] = 0.0;
  for (int p = 0; p < nq; p++) {
    const QPtQ q = rule[p];
    double r[3], W[3];
    quad_geom(g, qx, q.z, q.e, r, W);
    const double vR[3] = {r[0] - rt[0], r[1] - rt[1], r[2] - rt[2]};
    const double R = sqrt(vR[0] * vR[0] + vR[1] * vR[1] + vR[2] * vR[2]), iR = 1.0 / R, iR2 = iR * iR, iR3 = iR2 * iR;
    const double aW = sqrt(W[0] * W[0] + W[1] * W[1] + W[2] * W[2]);
    const double rdW = (vR[0] * W[0] + vR[1] * W[1] + vR[2] * W[2]) * iR;
    const cd ce = eikr(kr, ki, R);
    const cd cb = mk(-1.0 - ki * R, kr * R);
    const cd cbce = cb * ce;
    const cd gk = ce * (iR * aW), hk = cbce * (rdW * iR2), tk = cbce * (iR2 * aW);
    const cd kR = mk(kr * R, ki * R);
    const cd qq = kR * kR + cb * 3.0;
    const cd ce3 = ce * iR3;
#pragma unroll
    for (int a = 0; a < 4; a++) { fma_cr(aG[a], gk, q.w * q.n[a]); fma_cr(aH[a], hk, q.w * q.n[a]); }
    f = fma(q.w, rdW * iR2, f);
#pragma unroll
    for (int d = 0; d < ND; d++) {
      const double rd = (vR[0] * v[d][0] + vR[1] * v[d][1] + vR[2] * v[d][2]) * iR;
      const double Wdv = W[0] * v[d][0] + W[1] * v[d][1] + W[2] * v[d][2];
      const cd dh = ce3 * (qq * (rd * rdW) - cb * Wdv);
#pragma unroll
      for (int a = 0; a < 4; a++) { fma_cr(aDG[d][a], tk, q.w * q.n[a] * rd); fma_cr(aDH[d][a], dh, q.w * q.n[a]); }
      df[d] = fma(q.w, iR3 * (3.0 * rd * rdW - Wdv), df[d]);
    }
  }
  const double sh = flip ? -BEM_I_FP : BEM_I_FP;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    G[a] = aG[a] * BEM_I_FP;
    H[a] = aH[a] * sh;
#pragma unroll
    for (int d = 0; d < ND; d++) { dG[d][a] = aDG[d][a] * (-BEM_I_FP); dH[d][a] = aDH[d][a] * sh; }
  }
  F = f * sh;
#pragma unroll
  for (int d = 0; d < ND; d++) dF[d] = df[d] * sh;
}

// Self quadrangle at (zt, et) in polar coordinates about the point
// (bil_coef_bd, bil_coef.c:244-342, 426-536): 4 edges x 2 x 64 angles (this
// lane's share), 12 radial points; G always, H and F only on a warped
// quadrangle (their integrands carry the factor c1 . W2).  Partial sums of
// the calling lane, unscaled (the caller reduces and applies 1/(4 pi)).
__device__ __forceinline__ void quad_self_part(const TriGeo &g, const QuadExt &qx, double zt, double et, double kr,
                                               double ki, const double *__restrict__ xl, const double *__restrict__ wl,
                                               const double *__restrict__ xh, const double *__restrict__ wh, int lane,
                                               bool warped, cd *sg, cd *sh, double &sf)
{
  double zp = 1.0 - zt, zm = -1.0 - zt, ep = 1.0 - et, em = -1.0 - et;
  double r[5] = {sqrt(zm * zm + em * em), sqrt(zp * zp + em * em), sqrt(zp * zp + ep * ep), sqrt(zm * zm + ep * ep), 0.0};
  double th[5] = {atan2(em, zm), atan2(em, zp), atan2(ep, zp), atan2(ep, zm), 0.0};
  r[4] = r[0];
  th[4] = th[0] + 2.0 * M_PI;
  for (int a = 0; a < 4; a++) { sg[a] = mk(0, 0); sh[a] = mk(0, 0); }
  sf = 0.0;
  for (int ia = lane; ia < 512; ia += 64) {
    const int j = ia >> 7, m = (ia & 127) >> 1, half = ia & 1;
    const double dth = th[j + 1] - th[j];
    const double t = 0.25 * dth * xh[m] + (half ? 0.25 * (3.0 * th[j + 1] + th[j]) : 0.25 * (th[j + 1] + 3.0 * th[j]));
    const double rm = r[j] * r[j + 1] * sin(th[j + 1] - th[j]) / (r[j] * sin(t - th[j]) - r[j + 1] * sin(t - th[j + 1]));
    double c, sn;
    sincos(t, &sn, &c);
    const double fang = wh[m] * 0.25 * dth * 0.5 * rm;
    for (int i = 0; i < 12; i++) {
      const double rr = 0.5 * rm * (xl[i] + 1.0), z = rr * c + zt, e = rr * sn + et;
      double K[3], W[3], n[4];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        K[d] = g.c1[d] * c + g.c2[d] * sn + qx.c3[d] * (rr * c * sn + zt * sn + et * c);
        W[d] = g.W[d] + qx.W1[d] * z + qx.W2[d] * e;
      }
      const double aK = sqrt(K[0] * K[0] + K[1] * K[1] + K[2] * K[2]), iaK = 1.0 / aK;
      const double aW = sqrt(W[0] * W[0] + W[1] * W[1] + W[2] * W[2]);
      quad_N(z, e, n);
      const cd ce = eikr(kr, ki, aK * rr);
      const cd vg = ce * (iaK * aW * wl[i] * fang);
#pragma unroll
      for (int a = 0; a < 4; a++) sg[a] = sg[a] + vg * n[a];
      if (warped) {
        const double R = aK * rr, w3 = iaK * iaK * iaK * c * sn * wl[i] * fang;
        const cd vh = (mk(-1.0 - ki * R, kr * R) * ce) * w3;
#pragma unroll
        for (int a = 0; a < 4; a++) sh[a] = sh[a] + vh * n[a];
        sf += w3;
      }
    }
  }
}

// ---------------------------------------------------------------- domain id
// Og[d] accumulates the solid angle of faces whose main side is d and
// subtracts it from their sub side (setup2.c:944-978).  Loops the domains
// in order; the first face reporting "on the boundary" decides.
#define BEM_MAXDOM 8
__device__ __forceinline__ double solid_angle_tri(double a[3], double b[3], double c[3], int &flg)
{
  // unit vectors towards the vertices; signed spherical excess
  double *r[3] = {a, b, c};
  flg = 0;
  for (int i = 0; i < 3; i++) {
    double m = sqrt(r[i][0] * r[i][0] + r[i][1] * r[i][1] + r[i][2] * r[i][2]);
    if (m == 0.0) { flg = -1; return 0.0; }
    r[i][0] /= m; r[i][1] /= m; r[i][2] /= m;
  }
  double n01[3], n12[3], n20[3], t[3];
#define CRS(o, u, w) o[0] = u[1] * w[2] - u[2] * w[1]; o[1] = u[2] * w[0] - u[0] * w[2]; o[2] = u[0] * w[1] - u[1] * w[0];
  CRS(n01, a, b) CRS(n12, b, c) CRS(n20, c, a)
  CRS(t, n01, n20) double A0 = atan2(sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]), -(n01[0] * n20[0] + n01[1] * n20[1] + n01[2] * n20[2]));
  CRS(t, n12, n01) double A1 = atan2(sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]), -(n12[0] * n01[0] + n12[1] * n01[1] + n12[2] * n01[2]));
  CRS(t, n20, n12) double A2 = atan2(sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]), -(n20[0] * n12[0] + n20[1] * n12[1] + n20[2] * n12[2]));
#undef CRS
  double D = a[0] * n12[0] + a[1] * n12[1] + a[2] * n12[2];
  if (D > 0.0) return A0 + A1 + A2 - M_PI;
  if (D < 0.0) return -(A0 + A1 + A2 - M_PI);
  flg = -1;
  return 0.0;
}

// faces: per face four vertex positions (fq = 1: a quadrangle, counted as
// the triangles (0,1,2) and (0,2,3); fq = 0: a triangle, the fourth vertex
// unused) and its (md, sd); faces are grouped by main domain in segments
// segStart[d]..segStart[d+1]
__global__ void __launch_bounds__(256) k_domain(const double *__restrict__ pts, int np, const double *__restrict__ fv,
                                                const int *__restrict__ fsd, const int *__restrict__ fq,
                                                const int *__restrict__ segStart, int MN, int *__restrict__ did_out)
{
  const int p = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (p >= np) return;
  const double rt[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
  double og[BEM_MAXDOM];
  for (int d = 0; d <= MN; d++) og[d] = 0.0;
  int res = -1;
  for (int d = 0; d <= MN && res < 0; d++) {
    int onb = 0;
    for (int f = segStart[d] + lane; f < segStart[d + 1]; f += 64) {
      const double *v = fv + 12 * (size_t)f;
      double a[3] = {v[0] - rt[0], v[1] - rt[1], v[2] - rt[2]};
      double b[3] = {v[3] - rt[0], v[4] - rt[1], v[5] - rt[2]};
      double c[3] = {v[6] - rt[0], v[7] - rt[1], v[8] - rt[2]};
      int flg;
      double om = solid_angle_tri(a, b, c, flg);
      if (flg < 0) onb = 1;
      if (fq[f]) {
        double a2[3] = {v[0] - rt[0], v[1] - rt[1], v[2] - rt[2]};
        double c2[3] = {v[6] - rt[0], v[7] - rt[1], v[8] - rt[2]};
        double d2[3] = {v[9] - rt[0], v[10] - rt[1], v[11] - rt[2]};
        int flg2;
        double om2 = solid_angle_tri(a2, c2, d2, flg2);
        if (flg2 < 0) onb = 1;
        om += om2;
      }
      og[d] += om;
      og[fsd[f]] -= om;
    }
    if (__ballot(onb)) { res = d; break; }
    double od = wsum(og[d]);
    // keep the reduced value in lane 0 only so later reductions do not
    // count it 64 times
    og[d] = lane == 0 ? od : 0.0;
    if (d == 0 && fabs(od) < 2.0 * M_PI) res = 0;
    else if (od > 2.0 * M_PI) res = d;
    if (res < 0) {
      // contributions to later domains accumulated so far stay per lane
    }
  }
  if (lane == 0) did_out[p] = res;
}

// ---------------------------------------------------------------- fields at points
// what: 0 -> E,H from (E,dE,H,dH); 1 -> U from (U,dU).  Output per point:
// 6 (E,H) or 4 (U) complex values, scaled by 1/(1+F) (open domain) or 1/F.
// Quadrangles (geo[].quad != 0) take the square's rule qrule (nqq points).
// F32 (the mixed form's boundary fields of the other objects, as the own
// boundary integral k_field_bd<true>): triangles farther than 4 element
// radii from the point take the FP32 reduced far rule of the wave's far
// lanes (gl32, tri_coef32), nearer ones the FP64 rule.
template <bool F32>
__global__ void __launch_bounds__(256) k_field(const double *__restrict__ pts, const int *__restrict__ pdid, int np,
                                               int want_did, const TriGeo *__restrict__ geo, const int *__restrict__ sid,
                                               int dNe, double kr, double ki, const QPtN *__restrict__ rule, int nq,
                                               const QPt32 *__restrict__ gl32, double scale,
                                               const QuadExt *__restrict__ gq, const QPtQ *__restrict__ qrule, int nqq,
                                               int what, const cd *__restrict__ A0, const cd *__restrict__ dA0,
                                               const cd *__restrict__ A1, const cd *__restrict__ dA1,
                                               cd *__restrict__ out, int accumulate)
{
  const int p = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (p >= np) return;
  if (pdid[p] != want_did) return;
  const double rt[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
  cd acc[6] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  double F = 0.0;
  for (int s = 1 + lane; s <= dNe; s += 64) {
    int sd = sid[s], as = sd < 0 ? -sd : sd;
    const TriGeo g = geo[as];
    cd G[4], H[4];
    double f = 0.0;
    int ns = 3;
    if (g.quad != 0.0) {
      quad_coef(g, gq[as], rt, kr, ki, qrule, nqq, sd < 0, G, H, f);
      ns = 4;
    } else {
      cd dGx[3], dHx[3];
      double dfx;
      bool far = false;
      if constexpr (F32) {
        const double dx = g.c0[0] - rt[0], dy = g.c0[1] - rt[1], dz = g.c0[2] - rt[2];
        const double rad2 = elem_rad2(g);
        far = dx * dx + dy * dy + dz * dz > BEM_FAR_RADII * BEM_FAR_RADII * rad2;
        int qoff, nq32;
        far_rule(far, sqrt((kr * kr + ki * ki) * rad2), (dx * dx + dy * dy + dz * dz) / rad2, qoff, nq32);
        if (far) tri_coef32<false>(g, rt, nullptr, kr, ki, gl32 + qoff, nq32, scale, sd < 0, G, H, f, dGx, dHx, dfx);
      }
      if (!far) tri_coef<false>(g, rt, nullptr, kr, ki, rule, nq, scale, sd < 0, G, H, f, dGx, dHx, dfx);
      G[3] = H[3] = mk(0, 0);
    }
    F += f;
    if (what == 0) {
      const cd *E = A0 + (size_t)s * 12, *dE = dA0 + (size_t)s * 12, *Hh = A1 + (size_t)s * 12, *dH = dA1 + (size_t)s * 12;
      for (int n = 0; n < ns; n++)
#pragma unroll
        for (int l = 0; l < 3; l++) {
          acc[l] = acc[l] + G[n] * dE[n * 3 + l] - H[n] * E[n * 3 + l];
          acc[3 + l] = acc[3 + l] + G[n] * dH[n * 3 + l] - H[n] * Hh[n * 3 + l];
        }
    } else {
      const cd *U = A0 + (size_t)s * 16, *dU = dA0 + (size_t)s * 16;
      for (int n = 0; n < ns; n++)
#pragma unroll
        for (int l = 0; l < 4; l++) acc[l] = acc[l] + G[n] * dU[n * 4 + l] - H[n] * U[n * 4 + l];
    }
  }
  F = wsum(F);
  int nc = what == 0 ? 6 : 4;
  for (int l = 0; l < nc; l++) acc[l] = wsum(acc[l]);
  if (lane == 0) {
    double den = want_did == 0 ? 1.0 + F : F;
    for (int l = 0; l < nc; l++) {
      cd v = acc[l] * (1.0 / den);
      if (accumulate) out[(size_t)p * nc + l] = out[(size_t)p * nc + l] + v;
      else out[(size_t)p * nc + l] = v;
    }
  }
}

// ---------------------------------------------------------------- fields from potential gradients
// G[3], H[3], F and their gradients with respect to rt (x, y, z) from one
// triangle with a tabulated rule: the three dcoef_rt calls of EH_mEMP_s
// (unit directions) share one evaluation of the kernel per point.
__device__ __forceinline__ void tri_coef_grad(const TriGeo &g, const double *rt, double kr, double ki,
                                              const QPtN *__restrict__ rule, int nq, double scale, bool flip, cd *G,
                                              cd *H, double &F, cd (*dG)[3], cd (*dH)[3], double *dF)
{
  cd aG[3], aH[3], aDG[3][3], aDH[3][3];
  double f = 0.0, df[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    aG[i] = mk(0, 0); aH[i] = mk(0, 0);
#pragma unroll
    for (int d = 0; d < 3; d++) { aDG[d][i] = mk(0, 0); aDH[d][i] = mk(0, 0); }
  }
  for (int p = 0; p < nq; p++) {
    const QPtN q = rule[p];
    double r[3];
    tri_pos(g, q.z, q.e, r);
    const double vR[3] = {r[0] - rt[0], r[1] - rt[1], r[2] - rt[2]};
    const double R = sqrt(vR[0] * vR[0] + vR[1] * vR[1] + vR[2] * vR[2]), iR = 1.0 / R, iR2 = iR * iR, iR3 = iR2 * iR;
    const double rdW = (vR[0] * g.W[0] + vR[1] * g.W[1] + vR[2] * g.W[2]) * iR;
    const cd ce = eikr(kr, ki, R);
    const cd cb = mk(-1.0 - ki * R, kr * R);
    const cd cbce = cb * ce;
    const cd gk = ce * iR, hk = cbce * (rdW * iR2), tk = cbce * iR2;
    const cd kR = mk(kr * R, ki * R);
    const cd qq = kR * kR + cb * 3.0;
    const cd ce3 = ce * iR3;
    const double wn[3] = {q.w * q.n0, q.w * q.n1, q.w * q.n2};
#pragma unroll
    for (int i = 0; i < 3; i++) { fma_cr(aG[i], gk, wn[i]); fma_cr(aH[i], hk, wn[i]); }
    f = fma(q.w, rdW * iR2, f);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double rd = vR[d] * iR;   // (r' - rt).e_d / R
      const cd dh = ce3 * (qq * (rd * rdW) - cb * g.W[d]);
#pragma unroll
      for (int i = 0; i < 3; i++) { fma_cr(aDG[d][i], tk, wn[i] * rd); fma_cr(aDH[d][i], dh, wn[i]); }
      df[d] = fma(q.w, iR3 * (3.0 * rd * rdW - g.W[d]), df[d]);
    }
  }
  const double sh = flip ? -scale : scale;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    G[i] = aG[i] * (scale * g.aW);
    H[i] = aH[i] * sh;
#pragma unroll
    for (int d = 0; d < 3; d++) { dG[d][i] = aDG[d][i] * (-scale * g.aW); dH[d][i] = aDH[d][i] * sh; }
  }
  F = f * sh;
#pragma unroll
  for (int d = 0; d < 3; d++) dF[d] = df[d] * sh;
}

// E, H at points from the potentials and their gradients (EH_mEMP_s,
// d3b1_field.c:68-138): E = -grad phi + A, H = lambda0/(2 pi i) curl A.  Per
// element and direction d the rule switches to GHN when |dF_d| > 0.5
// (dcoef_rt, d3b1_elem.c:277-316); the potentials use the coefficients of
// the last direction (z), as the reference's loop leaves them.  Points with
// any |dF| > 100 contribute nothing (the reference returns -did there).
// Output added to out (6 complex per point).
__global__ void __launch_bounds__(256) k_field_memp(const double *__restrict__ pts, const int *__restrict__ pdid, int np,
                                                    int want_did, const TriGeo *__restrict__ geo,
                                                    const QuadExt *__restrict__ gq, const int *__restrict__ sid, int dNe,
                                                    double kr, double ki, const QPtN *__restrict__ rule, int nq,
                                                    double scale, int sw, const QPtN *__restrict__ gh, int ngh,
                                                    double ghscale, const QPtQ *__restrict__ qrule, int nqq,
                                                    const QPtQ *__restrict__ qgh, int nqgh, const cd *__restrict__ U,
                                                    const cd *__restrict__ dU, double lam0, cd *__restrict__ out)
{
  const int p = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (p >= np) return;
  if (pdid[p] != want_did) return;
  const double rt[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
  cd aU[4], aD[3][4];
#pragma unroll
  for (int l = 0; l < 4; l++) { aU[l] = mk(0, 0); aD[0][l] = mk(0, 0); aD[1][l] = mk(0, 0); aD[2][l] = mk(0, 0); }
  double F = 0.0, dF[3] = {0.0, 0.0, 0.0};
  const double ax[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  for (int s = 1 + lane; s <= dNe; s += 64) {
    const int sd = sid[s], as = sd < 0 ? -sd : sd;
    const TriGeo g = geo[as];
    if (g.quad != 0.0) {
      // bi-linear quadrangle: the square's rules, the same per-direction
      // GHN switch (dcoef_rt, d3b1_elem.c:277-296)
      const QuadExt qx = gq[as];
      cd G[4], H[4], DG[3][4], DH[3][4];
      double f, df[3];
      quad_coef_dir<3>(g, qx, rt, ax, kr, ki, qrule, nqq, sd < 0, G, H, f, DG, DH, df);
      if (sw) {
        const bool s0 = fabs(df[0]) > BEM_CBD_SDF, s1 = fabs(df[1]) > BEM_CBD_SDF, s2 = fabs(df[2]) > BEM_CBD_SDF;
        if (s0 || s1 || s2) {
          cd tG[4], tH[4], tDG[3][4], tDH[3][4];
          double tf, tdf[3];
          quad_coef_dir<3>(g, qx, rt, ax, kr, ki, qgh, nqgh, sd < 0, tG, tH, tf, tDG, tDH, tdf);
          const bool sws[3] = {s0, s1, s2};
#pragma unroll
          for (int d = 0; d < 3; d++)
            if (sws[d]) {
#pragma unroll
              for (int i = 0; i < 4; i++) { DG[d][i] = tDG[d][i]; DH[d][i] = tDH[d][i]; }
              df[d] = tdf[d];
            }
          if (s2) {
#pragma unroll
            for (int i = 0; i < 4; i++) { G[i] = tG[i]; H[i] = tH[i]; }
            f = tf;
          }
        }
      }
      const cd *Us = U + (size_t)s * 16, *dUs = dU + (size_t)s * 16;
#pragma unroll
      for (int n = 0; n < 4; n++)
#pragma unroll
        for (int l = 0; l < 4; l++) {
          const cd a = dUs[n * 4 + l], b = Us[n * 4 + l];
          aU[l] = aU[l] + G[n] * a - H[n] * b;
#pragma unroll
          for (int d = 0; d < 3; d++) aD[d][l] = aD[d][l] + DG[d][n] * a - DH[d][n] * b;
        }
      F += f;
#pragma unroll
      for (int d = 0; d < 3; d++) dF[d] += df[d];
      continue;
    }
    cd G[3], H[3], DG[3][3], DH[3][3];
    double f, df[3];
    tri_coef_grad(g, rt, kr, ki, rule, nq, scale, sd < 0, G, H, f, DG, DH, df);
    if (sw) {
      const bool s0 = fabs(df[0]) > BEM_CBD_SDF, s1 = fabs(df[1]) > BEM_CBD_SDF, s2 = fabs(df[2]) > BEM_CBD_SDF;
      if (s0 || s1 || s2) {
        cd tG[3], tH[3], tDG[3][3], tDH[3][3];
        double tf, tdf[3];
        tri_coef_grad(g, rt, kr, ki, gh, ngh, ghscale, sd < 0, tG, tH, tf, tDG, tDH, tdf);
        const bool sws[3] = {s0, s1, s2};
#pragma unroll
        for (int d = 0; d < 3; d++)
          if (sws[d]) {
#pragma unroll
            for (int i = 0; i < 3; i++) { DG[d][i] = tDG[d][i]; DH[d][i] = tDH[d][i]; }
            df[d] = tdf[d];
          }
        if (s2) {
#pragma unroll
          for (int i = 0; i < 3; i++) { G[i] = tG[i]; H[i] = tH[i]; }
          f = tf;
        }
      }
    }
    const cd *Us = U + (size_t)s * 16, *dUs = dU + (size_t)s * 16;
#pragma unroll
    for (int n = 0; n < 3; n++)
#pragma unroll
      for (int l = 0; l < 4; l++) {
        const cd a = dUs[n * 4 + l], b = Us[n * 4 + l];
        aU[l] = aU[l] + G[n] * a - H[n] * b;
#pragma unroll
        for (int d = 0; d < 3; d++) aD[d][l] = aD[d][l] + DG[d][n] * a - DH[d][n] * b;
      }
    F += f;
#pragma unroll
    for (int d = 0; d < 3; d++) dF[d] += df[d];
  }
  F = wsum(F);
#pragma unroll
  for (int d = 0; d < 3; d++) dF[d] = wsum(dF[d]);
#pragma unroll
  for (int l = 0; l < 4; l++) {
    aU[l] = wsum(aU[l]);
#pragma unroll
    for (int d = 0; d < 3; d++) aD[d][l] = wsum(aD[d][l]);
  }
  if (lane == 0) {
    if (fabs(dF[0]) > BEM_CBD_CDF || fabs(dF[1]) > BEM_CBD_CDF || fabs(dF[2]) > BEM_CBD_CDF) return;
    const double iden = 1.0 / (want_did == 0 ? 1.0 + F : F);
    cd u[4], du[3][4];
#pragma unroll
    for (int l = 0; l < 4; l++) {
      u[l] = aU[l] * iden;
#pragma unroll
      for (int d = 0; d < 3; d++) du[d][l] = (aD[d][l] - u[l] * dF[d]) * iden;
    }
    // lambda0 / (2 pi i) = -i lambda0 / (2 pi)
    const cd ch = mk(0.0, -lam0 / (2.0 * M_PI));
    cd *o = out + (size_t)p * 6;
#pragma unroll
    for (int i = 0; i < 3; i++) o[i] = o[i] + u[i] - du[i][3];
    o[3] = o[3] + ch * (du[1][2] - du[2][1]);
    o[4] = o[4] + ch * (du[2][0] - du[0][2]);
    o[5] = o[5] + ch * (du[0][1] - du[1][0]);
  }
}

// ---------------------------------------------------------------- coupling
// U, dU/dv at node slots of another object from this object's open-domain
// boundary values; GLN with the GHN switch per pair (dcoef_rt,
// d3b1_elem.c:277-316), added into Ui / dUi of the destination slots.
// far triangle of the coupling sums: centre farther than 4 element radii
// (the same FP64 test in both coupling kernels)
__device__ __forceinline__ bool scat_far(const TriGeo &g, const double *rt)
{
  const double dx = g.c0[0] - rt[0], dy = g.c0[1] - rt[1], dz = g.c0[2] - rt[2];
  const double rad2 = elem_rad2(g);
  return g.quad == 0.0 && dx * dx + dy * dy + dz * dz > BEM_FAR_RADII * BEM_FAR_RADII * rad2;
}

// First coupling pass: the far triangles only, in a kernel of their own so
// that none of the near-pair and quadrangle code's registers weigh on this
// loop.  F32 (mixed-precision form): FP32 with the far rule of the wave's
// far lanes; otherwise the reference's GLN rule in FP64, exactly as the
// second pass would sum it.  Per point: sums of U (4 complex), dU/dv (4),
// F, dF, and a flag (a far pair at the GHN switch level: the second pass
// then redoes the whole point, with the reference's switch).
#define BEM_SCAT_PART 19
template <bool F32>
__global__ void __launch_bounds__(256) k_scatter_far(const double *__restrict__ pts, const double *__restrict__ nrm,
                                                     int np, const TriGeo *__restrict__ geo,
                                                     const int *__restrict__ sid, int dNe, double kr, double ki,
                                                     double scale, const QPt32 *__restrict__ gl32,
                                                     const QPtN *__restrict__ gl, int ngl, const cd *__restrict__ U,
                                                     const cd *__restrict__ dU, double *__restrict__ part)
{
  const int p = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (p >= np) return;
  const double rt[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
  const double v[3] = {nrm[3 * p], nrm[3 * p + 1], nrm[3 * p + 2]};
  cd aU[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)}, aD[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  double F = 0.0, dF = 0.0;
  bool redo = false;
  for (int s = 1 + lane; s <= dNe; s += 64) {
    const int sd = sid[s], as = sd < 0 ? -sd : sd;
    const TriGeo g = geo[as];
    const bool far = scat_far(g, rt);
    cd G[3], H[3], DG[3], DH[3];
    double f, df;
    if constexpr (F32) {
      const double dx = g.c0[0] - rt[0], dy = g.c0[1] - rt[1], dz = g.c0[2] - rt[2];
      const double rad2 = elem_rad2(g);
      int qoff, nq32;
      far_rule(far, sqrt((kr * kr + ki * ki) * rad2), (dx * dx + dy * dy + dz * dz) / rad2, qoff, nq32);
      if (!far) continue;
      tri_coef32<true, true>(g, rt, v, kr, ki, gl32 + qoff, nq32, scale, sd < 0, G, H, f, DG, DH, df);
      if (fabs(df) > BEM_CBD_SDF * (1.0 - 1e-4)) { redo = true; continue; }
    } else {
      if (!far) continue;
      tri_coef<true>(g, rt, v, kr, ki, gl, ngl, scale, sd < 0, G, H, f, DG, DH, df);
      if (fabs(df) > BEM_CBD_SDF) { redo = true; continue; }
    }
    F += f; dF += df;
    const cd *Us = U + (size_t)s * 16, *dUs = dU + (size_t)s * 16;
#pragma unroll 1
    for (int n = 0; n < 3; n++)
#pragma unroll
      for (int l = 0; l < 4; l++) {
        const cd u = Us[n * 4 + l], du = dUs[n * 4 + l];
        aU[l] = aU[l] + G[n] * du - H[n] * u;
        aD[l] = aD[l] + DG[n] * du - DH[n] * u;
      }
  }
  const bool any_redo = __ballot(redo) != 0;
  F = wsum(F); dF = wsum(dF);
  for (int l = 0; l < 4; l++) { aU[l] = wsum(aU[l]); aD[l] = wsum(aD[l]); }
  if (lane == 0) {
    double *q = part + (size_t)p * BEM_SCAT_PART;
    for (int l = 0; l < 4; l++) {
      q[2 * l] = aU[l].x; q[2 * l + 1] = aU[l].y;
      q[8 + 2 * l] = aD[l].x; q[9 + 2 * l] = aD[l].y;
    }
    q[16] = F; q[17] = dF; q[18] = any_redo ? 1.0 : 0.0;
  }
}

// QUADS false: the source object has triangles only (the quadrangle branch
// compiled out, which keeps the kernel's registers down)
template <bool QUADS>
__global__ void __launch_bounds__(256) k_scatter(const double *__restrict__ pts, const double *__restrict__ nrm,
                                                 const int *__restrict__ slot, int np, const TriGeo *__restrict__ geo,
                                                 const QuadExt *__restrict__ gq, const int *__restrict__ sid, int dNe,
                                                 double kr, double ki, const QPtN *__restrict__ gl, int ngl,
                                                 const QPtN *__restrict__ gh, int ngh, double scale,
                                                 const QPtQ *__restrict__ qgl, int nqgl, const QPtQ *__restrict__ qgh,
                                                 int nqgh, const QPt32 *__restrict__ gl32, const cd *__restrict__ U,
                                                 const cd *__restrict__ dU, const double *__restrict__ part,
                                                 cd *__restrict__ Ui, cd *__restrict__ dUi)
{
  const int p = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (p >= np) return;
  const double rt[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
  const double v[3] = {nrm[3 * p], nrm[3 * p + 1], nrm[3 * p + 2]};
  cd aU[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)}, aD[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  double F = 0.0, dF = 0.0;
  const double vv[1][3] = {{v[0], v[1], v[2]}};
  // part: the far triangles' sums of k_scatter_far for this point, unless
  // it met a far pair at the GHN switch level (then everything is done here)
  const bool use_part = part && part[(size_t)p * BEM_SCAT_PART + 18] == 0.0;
  for (int s = 1 + lane; s <= dNe; s += 64) {
    int sd = sid[s], as = sd < 0 ? -sd : sd;
    const TriGeo g = geo[as];
    if (QUADS && g.quad != 0.0) {
      const QuadExt qx = gq[as];
      cd G[4], H[4], DG[1][4], DH[1][4];
      double f, df[1];
      quad_coef_dir<1>(g, qx, rt, vv, kr, ki, qgl, nqgl, sd < 0, G, H, f, DG, DH, df);
      if (fabs(df[0]) > BEM_CBD_SDF) quad_coef_dir<1>(g, qx, rt, vv, kr, ki, qgh, nqgh, sd < 0, G, H, f, DG, DH, df);
      F += f; dF += df[0];
      const cd *Us = U + (size_t)s * 16, *dUs = dU + (size_t)s * 16;
#pragma unroll
      for (int n = 0; n < 4; n++)
#pragma unroll
        for (int l = 0; l < 4; l++) {
          aU[l] = aU[l] + G[n] * dUs[n * 4 + l] - H[n] * Us[n * 4 + l];
          aD[l] = aD[l] + DG[0][n] * dUs[n * 4 + l] - DH[0][n] * Us[n * 4 + l];
        }
      continue;
    }
    cd G[3], H[3], DG[3], DH[3];
    double f, df;
    bool done = false;
    if (use_part) {
      if (scat_far(g, rt)) continue;   // in part
    } else if (gl32) {
      // mixed-precision form: far triangles in FP32 with the far rule of the
      // wave's far lanes; a result within 1e-4 of the GHN switch level is
      // redone in FP64 below (far pairs stay orders of magnitude under it)
      const double dx = g.c0[0] - rt[0], dy = g.c0[1] - rt[1], dz = g.c0[2] - rt[2];
      const double rad2 = elem_rad2(g);
      const bool far = dx * dx + dy * dy + dz * dz > BEM_FAR_RADII * BEM_FAR_RADII * rad2;
      int qoff, nq32;
      far_rule(far, sqrt((kr * kr + ki * ki) * rad2), (dx * dx + dy * dy + dz * dz) / rad2, qoff, nq32);
      if (far) {
        tri_coef32<true>(g, rt, v, kr, ki, gl32 + qoff, nq32, scale, sd < 0, G, H, f, DG, DH, df);
        done = !(fabs(df) > BEM_CBD_SDF * (1.0 - 1e-4));
      }
    }
    if (!done) {
      tri_coef<true>(g, rt, v, kr, ki, gl, ngl, scale, sd < 0, G, H, f, DG, DH, df);
      if (fabs(df) > BEM_CBD_SDF) tri_coef<true>(g, rt, v, kr, ki, gh, ngh, scale, sd < 0, G, H, f, DG, DH, df);
    }
    F += f; dF += df;
    const cd *Us = U + (size_t)s * 16, *dUs = dU + (size_t)s * 16;
#pragma unroll
    for (int n = 0; n < 3; n++)
#pragma unroll
      for (int l = 0; l < 4; l++) {
        aU[l] = aU[l] + G[n] * dUs[n * 4 + l] - H[n] * Us[n * 4 + l];
        aD[l] = aD[l] + DG[n] * dUs[n * 4 + l] - DH[n] * Us[n * 4 + l];
      }
  }
  F = wsum(F); dF = wsum(dF);
  for (int l = 0; l < 4; l++) { aU[l] = wsum(aU[l]); aD[l] = wsum(aD[l]); }
  if (use_part) {
    const double *q = part + (size_t)p * BEM_SCAT_PART;
    for (int l = 0; l < 4; l++) {
      aU[l] = aU[l] + mk(q[2 * l], q[2 * l + 1]);
      aD[l] = aD[l] + mk(q[8 + 2 * l], q[9 + 2 * l]);
    }
    F += q[16]; dF += q[17];
  }
  if (lane == 0) {
    double den = 1.0 + F;   // the target is in the open domain of this object
    size_t o = (size_t)slot[p] * 4;
    for (int l = 0; l < 4; l++) {
      cd u = aU[l] * (1.0 / den);
      cd du = (aD[l] - u * dF) * (1.0 / den);
      Ui[o + l] = Ui[o + l] + u;
      dUi[o + l] = dUi[o + l] + du;
    }
  }
}

// ---------------------------------------------------------------- boundary points
// E, H at (t, zeta, eta) on the object's own surface in domain did: GLN for
// s != t (coef_bd, no GHN switch), polar G for s == t (lit_coef_bd; H and F
// vanish on a flat element).
// F32: the sums over far triangles (farther than 4 element radii) with the
// FP32 far rules gl32 (mixed-precision form: 5x5, 6x6 or 12x12 squares by
// |k| x element radius); nearer elements, quadrangles and
// the self element stay FP64.  Quadrangles take the square's GLN rule qgl
// and, as the self element, the polar rule of bil_coef_bd.
template <bool F32>
__global__ void __launch_bounds__(256) k_field_bd(const int *__restrict__ tgt_t, const double *__restrict__ tgt_ze,
                                                  int np, int did, const TriGeo *__restrict__ geo,
                                                  const QuadExt *__restrict__ gq, const int *__restrict__ sid, int dNe,
                                                  double kr, double ki, const QPtN *__restrict__ gl,
                                                  const QPt32 *__restrict__ gl32, int ngl, double scale,
                                                  const QPtQ *__restrict__ qgl, int nqgl,
                                                  const double *__restrict__ xl, const double *__restrict__ wl,
                                                  const double *__restrict__ xh, const double *__restrict__ wh,
                                                  const cd *__restrict__ E, const cd *__restrict__ dE,
                                                  const cd *__restrict__ Hh, const cd *__restrict__ dH,
                                                  cd *__restrict__ out, double *__restrict__ rt_out)
{
  const int p = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (p >= np) return;
  const int t = tgt_t[p];
  const double zt = tgt_ze[2 * p], et = tgt_ze[2 * p + 1];
  int td = sid[t], at = td < 0 ? -td : td;
  const TriGeo gt = geo[at];
  const bool tquad = gt.quad != 0.0;
  double rt[3];
  if (tquad) {
    double Wt[3];
    quad_geom(gt, gq[at], zt, et, rt, Wt);
  } else {
    tri_pos(gt, zt, et, rt);
  }
  cd acc[6] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  double F = 0.0;
  for (int s = 1 + lane; s <= dNe; s += 64) {
    if (s == t) continue;
    int sd = sid[s], as = sd < 0 ? -sd : sd;
    const TriGeo g = geo[as];
    cd G[4], H[4];
    double f = 0.0;
    int ns = 3;
    if (g.quad != 0.0) {
      quad_coef(g, gq[as], rt, kr, ki, qgl, nqgl, sd < 0, G, H, f);
      ns = 4;
    } else {
      cd x0[3], x1[3];
      double x2;
      bool far = false;
      if constexpr (F32) {
        // FP32 only farther than 4 element radii from the element's centre,
        // with the far rule of the wave's far lanes
        const double dx = g.c0[0] - rt[0], dy = g.c0[1] - rt[1], dz = g.c0[2] - rt[2];
        const double rad2 = elem_rad2(g);
        far = dx * dx + dy * dy + dz * dz > BEM_FAR_RADII * BEM_FAR_RADII * rad2;
        int qoff, nq32;
        far_rule(far, sqrt((kr * kr + ki * ki) * rad2), (dx * dx + dy * dy + dz * dz) / rad2, qoff, nq32);
        if (far) tri_coef32<false>(g, rt, nullptr, kr, ki, gl32 + qoff, nq32, scale, sd < 0, G, H, f, x0, x1, x2);
      }
      if (!far)
        tri_coef<false>(g, rt, nullptr, kr, ki, gl, ngl, scale, sd < 0, G, H, f, x0, x1, x2);
    }
    F += f;
    const cd *Es = E + (size_t)s * 12, *dEs = dE + (size_t)s * 12, *Hs = Hh + (size_t)s * 12, *dHs = dH + (size_t)s * 12;
    for (int n = 0; n < ns; n++)
#pragma unroll
      for (int l = 0; l < 3; l++) {
        acc[l] = acc[l] + G[n] * dEs[n * 3 + l] - H[n] * Es[n * 3 + l];
        acc[3 + l] = acc[3 + l] + G[n] * dHs[n * 3 + l] - H[n] * Hs[n * 3 + l];
      }
  }
  // self element: polar rule about (zt, et), 12 radial points per angle
  cd sg[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)}, shq[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  double sfq = 0.0, bdC = 0.0;
  bool warped = false;
  if (tquad) {
    // H and F of the self element need a warped quadrangle (bil_check_plane,
    // d3b1_elem.c:640-649: planar when |c1.W2| / (|c1| |W2|) < 1e-15)
    const QuadExt qt = gq[at];
    bdC = gt.c1[0] * qt.W2[0] + gt.c1[1] * qt.W2[1] + gt.c1[2] * qt.W2[2];
    const double n1 = sqrt(gt.c1[0] * gt.c1[0] + gt.c1[1] * gt.c1[1] + gt.c1[2] * gt.c1[2]);
    const double n2 = sqrt(qt.W2[0] * qt.W2[0] + qt.W2[1] * qt.W2[1] + qt.W2[2] * qt.W2[2]);
    warped = !(fabs(bdC / (n1 * n2)) < 1e-15) && bdC != 0.0;
    quad_self_part(gt, qt, zt, et, kr, ki, xl, wl, xh, wh, lane, warped, sg, shq, sfq);
  } else {
    double zp = 1.0 - zt, zm = -0.5 - zt, ep = 0.5 * BEM_SQ3 - et, em = -0.5 * BEM_SQ3 - et;
    double r[4] = {sqrt(zm * zm + em * em), sqrt(zp * zp + et * et), sqrt(zm * zm + ep * ep), 0.0};
    double th[4] = {atan2(em, zm), atan2(-et, zp), atan2(ep, zm), 0.0};
    r[3] = r[0];
    th[3] = th[0] + 2.0 * M_PI;
    for (int ia = lane; ia < 384; ia += 64) {
      int j = ia / 128, m = (ia % 128) >> 1, half = ia & 1;
      double dth = th[j + 1] - th[j];
      double tt = 0.25 * dth * xh[m] + (half ? 0.25 * (3.0 * th[j + 1] + th[j]) : 0.25 * (th[j + 1] + 3.0 * th[j]));
      double rm = r[j] * r[j + 1] * sin(th[j + 1] - th[j]) / (r[j] * sin(tt - th[j]) - r[j + 1] * sin(tt - th[j + 1]));
      double c, sn;
      sincos(tt, &sn, &c);
      double K[3] = {gt.c1[0] * c + gt.c2[0] * sn, gt.c1[1] * c + gt.c2[1] * sn, gt.c1[2] * c + gt.c2[2] * sn};
      double aK = sqrt(K[0] * K[0] + K[1] * K[1] + K[2] * K[2]), iaK = 1.0 / aK;
      double fg = wh[m] * 0.25 * dth * 0.5 * rm;
      for (int i = 0; i < 12; i++) {
        double rr = 0.5 * rm * (xl[i] + 1.0);
        double n0, n1, n2;
        tri_N(rr * c + zt, rr * sn + et, n0, n1, n2);
        cd vv = eikr(kr, ki, aK * rr) * (iaK * wl[i] * fg);
        fma_cr(sg[0], vv, n0); fma_cr(sg[1], vv, n1); fma_cr(sg[2], vv, n2);
      }
    }
  }
  const int nst = tquad ? 4 : 3;
  const double sG = tquad ? BEM_I_FP : BEM_I_SPSQ3 * gt.aW;
  for (int a = 0; a < nst; a++) sg[a] = wsum(sg[a]) * sG;
  if (warped) {   // wave-uniform
    const double sH = (td < 0 ? -BEM_I_FP : BEM_I_FP) * bdC;
    for (int a = 0; a < 4; a++) shq[a] = wsum(shq[a]) * sH;
    sfq = wsum(sfq) * sH;
  }
  const cd *Et = E + (size_t)t * 12, *dEt = dE + (size_t)t * 12, *Ht = Hh + (size_t)t * 12, *dHt = dH + (size_t)t * 12;
  F = wsum(F);
  for (int l = 0; l < 6; l++) acc[l] = wsum(acc[l]);
  if (lane == 0) {
    for (int n = 0; n < nst; n++)
      for (int l = 0; l < 3; l++) {
        acc[l] = acc[l] + sg[n] * dEt[n * 3 + l] - shq[n] * Et[n * 3 + l];
        acc[3 + l] = acc[3 + l] + sg[n] * dHt[n * 3 + l] - shq[n] * Ht[n * 3 + l];
      }
    F += sfq;
    double den = did == 0 ? 1.0 + F : F;
    for (int l = 0; l < 6; l++) out[(size_t)p * 6 + l] = acc[l] * (1.0 / den);
    rt_out[3 * p] = rt[0]; rt_out[3 * p + 1] = rt[1]; rt_out[3 * p + 2] = rt[2];
  }
}

// ---------------------------------------------------------------- element layer
// Coefficients of one source element at a point off the element, one
// wavefront per request, lanes over the rule's points: CC (G 0-3, H 4-7,
// F 8) of coef_rt / coef_bd / coef_bd_node (t != s), and with nd = 1 or 2
// directions the derivatives dC along v0 (9-17) and v1 (18-26) of dcoef_rt /
// dcoef_bd_tz / dcoef_bd_te / dcoef_bd_t2 / dcoef_bd_node_t2
// (d3b1_elem.c:148-539): rule type 0..3 of the element's kind, and for the
// derivatives with type < 3 the 64x64 rule again when |dF| along a
// direction exceeds 0.5 (the reference's GHN switch).  out[27 idx[i] + ...].
template <int ND>
__device__ __forceinline__ void elem_sums(const TriGeo &g, const QuadExt &qx, const double *rt, const double (*v)[3],
                                          double kr, double ki, const QPtN *__restrict__ rule, int nq, double scale,
                                          const QPtQ *__restrict__ qrule, int nqq, bool flip, int lane, cd *G, cd *H,
                                          double &F, cd (*dG)[4], cd (*dH)[4], double *dF)
{
  const bool quad = g.quad != 0.0;
  const int n = quad ? nqq : nq, per = (n + 63) / 64, p0 = min(n, lane * per), cnt = min(n, p0 + per) - p0;
  for (int a = 0; a < 4; a++) {
    G[a] = H[a] = mk(0, 0);
    for (int d = 0; d < 2; d++) dG[d][a] = dH[d][a] = mk(0, 0);
  }
  F = dF[0] = dF[1] = 0.0;
  if (quad) {
    if constexpr (ND == 0) quad_coef(g, qx, rt, kr, ki, qrule + p0, cnt, flip, G, H, F);
    else quad_coef_dir<ND>(g, qx, rt, v, kr, ki, qrule + p0, cnt, flip, G, H, F, dG, dH, dF);
  } else {
    cd x0[3], x1[3];
    double x2;
    if (ND == 0) tri_coef<false>(g, rt, nullptr, kr, ki, rule + p0, cnt, scale, flip, G, H, F, x0, x1, x2);
    for (int d = 0; d < ND; d++) {
      cd tg[3], th[3], dg[3], dh[3];
      double tf, df;
      tri_coef<true>(g, rt, v[d], kr, ki, rule + p0, cnt, scale, flip, tg, th, tf, dg, dh, df);
      if (d == 0) {
        for (int a = 0; a < 3; a++) { G[a] = tg[a]; H[a] = th[a]; }
        F = tf;
      }
      for (int a = 0; a < 3; a++) { dG[d][a] = dg[a]; dH[d][a] = dh[a]; }
      dF[d] = df;
    }
  }
  for (int a = 0; a < 4; a++) {
    G[a] = wsum(G[a]);
    H[a] = wsum(H[a]);
    for (int d = 0; d < ND; d++) { dG[d][a] = wsum(dG[d][a]); dH[d][a] = wsum(dH[d][a]); }
  }
  F = wsum(F);
  for (int d = 0; d < ND; d++) dF[d] = wsum(dF[d]);
}

template <int ND>
__global__ void __launch_bounds__(256) k_elem_coef(const TriGeo *__restrict__ geo, const QuadExt *__restrict__ gq, int n,
                                                   const int *__restrict__ s, const double *__restrict__ rt,
                                                   const double *__restrict__ v0, const double *__restrict__ v1,
                                                   const int *__restrict__ idx, double kr, double ki,
                                                   const QPtN *__restrict__ rule, int nq, double scale,
                                                   const QPtQ *__restrict__ qrule, int nqq, const QPtN *__restrict__ rgh,
                                                   int ngh, double sgh, const QPtQ *__restrict__ qgh, int nqgh,
                                                   int fallback, cd *__restrict__ out)
{
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int sd = s[i], as = sd < 0 ? -sd : sd;
  const TriGeo g = geo[as];
  const QuadExt qx = g.quad != 0.0 ? gq[as] : QuadExt{};
  const double r[3] = {rt[3 * i], rt[3 * i + 1], rt[3 * i + 2]};
  double v[2][3] = {{0, 0, 0}, {0, 0, 0}};
  for (int c = 0; c < 3; c++) {
    if (ND > 0) v[0][c] = v0[3 * i + c];
    if (ND > 1) v[1][c] = v1[3 * i + c];
  }
  cd G[4], H[4], dG[2][4], dH[2][4];
  double F, dF[2];
  elem_sums<ND>(g, qx, r, v, kr, ki, rule, nq, scale, qrule, nqq, sd < 0, lane, G, H, F, dG, dH, dF);
  if (ND > 0 && fallback) {
    bool redo = false;
    for (int d = 0; d < ND; d++) redo = redo || fabs(dF[d]) > BEM_CBD_SDF;
    if (redo) elem_sums<ND>(g, qx, r, v, kr, ki, rgh, ngh, sgh, qgh, nqgh, sd < 0, lane, G, H, F, dG, dH, dF);
  }
  if (lane == 0) {
    cd *o = out + 27 * (size_t)idx[i];
    for (int a = 0; a < 4; a++) { o[a] = G[a]; o[4 + a] = H[a]; }
    o[8] = mk(F, 0.0);
    for (int d = 0; d < 2; d++) {
      for (int a = 0; a < 4; a++) {
        o[9 + 9 * d + a] = d < ND ? dG[d][a] : mk(0, 0);
        o[13 + 9 * d + a] = d < ND ? dH[d][a] : mk(0, 0);
      }
      o[17 + 9 * d] = mk(d < ND ? dF[d] : 0.0, 0.0);
    }
  }
}

// ---------------------------------------------------------------- field rows
// Rows of the E, H boundary integral of one object's domain did at np
// points, for fields of many right-hand sides as one product: the element
// coefficients G_n, H_n of EH_s / EH_s_bd (d3b1_field.c) written out
// instead of summed, R[p][4 s + n] = G_n and R[p][4 (dNe+1) + 4 s + n] = H_n
// of element s (n < 3 for a triangle, < 4 for a quadrangle, the other
// entries and element 0 zero), F[p] = the solid-angle sum.  The field of a
// right-hand side at p is sum_j R[p][j] X[j] / (did == 0 ? 1 + F : F) with
// X = [dE | dH ; -E | -H] by (element, node) rows.  SELF: points on the
// object (element tgt_t[p], local coordinates tgt_ze[p]), the self element
// by the polar rule of k_field_bd; otherwise points pts off the object.
// Triangles farther than 4 element radii take the reduced far rules (FP64,
// the tiers of the mixed-precision form, relative error below 3e-8),
// nearer ones and quadrangles the Gauss-Legendre rules.
template <bool SELF>
__global__ void __launch_bounds__(256) k_field_rows(const double *__restrict__ pts, const int *__restrict__ tgt_t,
                                                 const double *__restrict__ tgt_ze, int np, int did,
                                                 const TriGeo *__restrict__ geo, const QuadExt *__restrict__ gq,
                                                 const int *__restrict__ sid, int dNe, double kr, double ki,
                                                 const QPtN *__restrict__ gl, int ngl, const QPt *__restrict__ glr,
                                                 double scale, const QPtQ *__restrict__ qgl, int nqgl,
                                                 const double *__restrict__ xl, const double *__restrict__ wl,
                                                 const double *__restrict__ xh, const double *__restrict__ wh, size_t ld,
                                                 cd *__restrict__ R, double *__restrict__ Fo)
{
  const int p = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (p >= np) return;
  double rt[3];
  int t = 0, td = 0, at = 0;
  TriGeo gt{};
  bool tquad = false;
  double zt = 0.0, et = 0.0;
  if (SELF) {
    t = tgt_t[p];
    zt = tgt_ze[2 * p];
    et = tgt_ze[2 * p + 1];
    td = sid[t];
    at = td < 0 ? -td : td;
    gt = geo[at];
    tquad = gt.quad != 0.0;
    if (tquad) {
      double Wt[3];
      quad_geom(gt, gq[at], zt, et, rt, Wt);
    } else {
      tri_pos(gt, zt, et, rt);
    }
  } else {
    rt[0] = pts[3 * p]; rt[1] = pts[3 * p + 1]; rt[2] = pts[3 * p + 2];
  }
  cd *Rg = R + (size_t)p * ld, *Rh = Rg + 4 * ((size_t)dNe + 1);
  if (lane < 4) { Rg[lane] = mk(0, 0); Rh[lane] = mk(0, 0); }
  double F = 0.0;
  for (int s = 1 + lane; s <= dNe; s += 64) {
    if (SELF && s == t) continue;   // written below
    const int sd = sid[s], as = sd < 0 ? -sd : sd;
    const TriGeo g = geo[as];
    cd G[4], H[4];
    double f = 0.0;
    if (g.quad != 0.0) {
      quad_coef(g, gq[as], rt, kr, ki, qgl, nqgl, sd < 0, G, H, f);
    } else {
      cd x0[3], x1[3];
      double x2;
      const double dx = g.c0[0] - rt[0], dy = g.c0[1] - rt[1], dz = g.c0[2] - rt[2];
      const double rad2 = elem_rad2(g), d2 = dx * dx + dy * dy + dz * dz;
      const bool far = d2 > BEM_FAR_RADII * BEM_FAR_RADII * rad2;
      int qoff, nq;
      far_rule(far, sqrt((kr * kr + ki * ki) * rad2), d2 / rad2, qoff, nq);
      if (far) tri_coef<false, QPt>(g, rt, nullptr, kr, ki, glr + qoff, nq, scale, sd < 0, G, H, f, x0, x1, x2);
      else tri_coef<false>(g, rt, nullptr, kr, ki, gl, ngl, scale, sd < 0, G, H, f, x0, x1, x2);
      G[3] = H[3] = mk(0, 0);
    }
    F += f;
#pragma unroll
    for (int n = 0; n < 4; n++) { Rg[4 * (size_t)s + n] = G[n]; Rh[4 * (size_t)s + n] = H[n]; }
  }
  if (SELF) {
    // self element: polar rule about (zt, et), as k_field_bd
    cd sg[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)}, shq[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
    double sfq = 0.0, bdC = 0.0;
    bool warped = false;
    if (tquad) {
      const QuadExt qt = gq[at];
      bdC = gt.c1[0] * qt.W2[0] + gt.c1[1] * qt.W2[1] + gt.c1[2] * qt.W2[2];
      const double n1 = sqrt(gt.c1[0] * gt.c1[0] + gt.c1[1] * gt.c1[1] + gt.c1[2] * gt.c1[2]);
      const double n2 = sqrt(qt.W2[0] * qt.W2[0] + qt.W2[1] * qt.W2[1] + qt.W2[2] * qt.W2[2]);
      warped = !(fabs(bdC / (n1 * n2)) < 1e-15) && bdC != 0.0;
      quad_self_part(gt, qt, zt, et, kr, ki, xl, wl, xh, wh, lane, warped, sg, shq, sfq);
    } else {
      double zp = 1.0 - zt, zm = -0.5 - zt, ep = 0.5 * BEM_SQ3 - et, em = -0.5 * BEM_SQ3 - et;
      double r[4] = {sqrt(zm * zm + em * em), sqrt(zp * zp + et * et), sqrt(zm * zm + ep * ep), 0.0};
      double th[4] = {atan2(em, zm), atan2(-et, zp), atan2(ep, zm), 0.0};
      r[3] = r[0];
      th[3] = th[0] + 2.0 * M_PI;
      for (int ia = lane; ia < 384; ia += 64) {
        int j = ia / 128, m = (ia % 128) >> 1, half = ia & 1;
        double dth = th[j + 1] - th[j];
        double tt = 0.25 * dth * xh[m] + (half ? 0.25 * (3.0 * th[j + 1] + th[j]) : 0.25 * (th[j + 1] + 3.0 * th[j]));
        double rm = r[j] * r[j + 1] * sin(th[j + 1] - th[j]) / (r[j] * sin(tt - th[j]) - r[j + 1] * sin(tt - th[j + 1]));
        double c, sn;
        sincos(tt, &sn, &c);
        double K[3] = {gt.c1[0] * c + gt.c2[0] * sn, gt.c1[1] * c + gt.c2[1] * sn, gt.c1[2] * c + gt.c2[2] * sn};
        double aK = sqrt(K[0] * K[0] + K[1] * K[1] + K[2] * K[2]), iaK = 1.0 / aK;
        double fg = wh[m] * 0.25 * dth * 0.5 * rm;
        for (int i = 0; i < 12; i++) {
          double rr = 0.5 * rm * (xl[i] + 1.0);
          double n0, n1, n2;
          tri_N(rr * c + zt, rr * sn + et, n0, n1, n2);
          cd vv = eikr(kr, ki, aK * rr) * (iaK * wl[i] * fg);
          fma_cr(sg[0], vv, n0); fma_cr(sg[1], vv, n1); fma_cr(sg[2], vv, n2);
        }
      }
    }
    const int nst = tquad ? 4 : 3;
    const double sG = tquad ? BEM_I_FP : BEM_I_SPSQ3 * gt.aW;
    for (int a = 0; a < nst; a++) sg[a] = wsum(sg[a]) * sG;
    if (warped) {   // wave-uniform
      const double sH = (td < 0 ? -BEM_I_FP : BEM_I_FP) * bdC;
      for (int a = 0; a < 4; a++) shq[a] = wsum(shq[a]) * sH;
      sfq = wsum(sfq) * sH;
    }
    if (lane == 0) {
      for (int n = 0; n < 4; n++) {
        Rg[4 * (size_t)t + n] = n < nst ? sg[n] : mk(0, 0);
        Rh[4 * (size_t)t + n] = shq[n];
      }
    }
    F = wsum(F);
    F += sfq;
  } else {
    F = wsum(F);
  }
  if (lane == 0) Fo[p] = F;
}

// X of nb right-hand sides for the rows of k_field_rows: X[j][b 6 + l] with
// rows j = 4 s + n: dE (l < 3), dH (l >= 3); rows 4 (dNe+1) + 4 s + n: -E,
// -H (boundary values [s][n][3] of beam b at + b es)
__global__ void k_pack_ehbv(const cd *__restrict__ E, const cd *__restrict__ dE, const cd *__restrict__ H,
                            const cd *__restrict__ dH, size_t es, int dNe, int nb, cd *__restrict__ X)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nr = 4 * ((size_t)dNe + 1);
  if (i >= 2 * nr * nb) return;
  const size_t j = i / nb;
  const int b = (int)(i % nb);
  const bool hpart = j >= nr;
  const size_t sn = hpart ? j - nr : j;   // 4 s + n
  const size_t o = (size_t)b * es + sn * 3;
  cd *x = X + j * 6 * (size_t)nb + 6 * (size_t)b;
  for (int l = 0; l < 3; l++) {
    if (hpart) { x[l] = -E[o + l]; x[3 + l] = -H[o + l]; }
    else { x[l] = dE[o + l]; x[3 + l] = dH[o + l]; }
  }
}

// EH[b][idx[i]][l] += C[i][b 6 + l] / den_i, den = 1 + F (open domain) or F
// (idx null: i; EH of beam b at + b ldb)
__global__ void k_rows_add(const cd *__restrict__ C, int n, int nb, const double *__restrict__ F, int open,
                           const int *__restrict__ idx, size_t ldb, cd *__restrict__ EH)
{
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * nb) return;
  const int i = (int)(t / nb), b = (int)(t % nb);
  const double den = open ? 1.0 + F[i] : F[i];
  const double s = 1.0 / den;
  const cd *c = C + (size_t)i * 6 * nb + 6 * (size_t)b;
  cd *o = EH + (size_t)b * ldb + 6 * (size_t)(idx ? idx[i] : i);
  for (int l = 0; l < 6; l++) o[l] = o[l] + c[l] * s;
}

// ---------------------------------------------------------------- incident field
__device__ __forceinline__ void beam_EH(const Beam *__restrict__ bm, int nb, const double *x, const double *v, cd *e,
                                        cd *h, cd *de, cd *dh)
{
  for (int j = 0; j < 3; j++) { e[j] = mk(0, 0); h[j] = mk(0, 0); if (de) { de[j] = mk(0, 0); dh[j] = mk(0, 0); } }
  for (int i = 0; i < nb; i++) {
    const Beam b = bm[i];
    double ph = b.ki * ((x[0] - b.fx) * b.st * b.cp + (x[1] - b.fy) * b.st * b.sp + (x[2] - b.fz) * b.ct);
    double s, c;
    sincos(ph, &s, &c);
    cd a = mk(c * b.E0, s * b.E0);
    cd f[6] = {b.ex * a, b.ey * a, b.ez * a, b.hx * a, b.hy * a, b.hz * a};
    for (int j = 0; j < 3; j++) { e[j] = e[j] + f[j]; h[j] = h[j] + f[3 + j]; }
    if (de) {
      cd gcf = mk(0.0, b.ki * (b.st * b.cp * v[0] + b.st * b.sp * v[1] + b.ct * v[2]));
      for (int j = 0; j < 3; j++) { de[j] = de[j] + f[j] * gcf; dh[j] = dh[j] + f[3 + j] * gcf; }
    }
  }
}

// boundary incident data of the open-domain node slots (4 slots, phi = 0)
__global__ void k_incident_bd(const double *__restrict__ pts, const double *__restrict__ nrm,
                              const int *__restrict__ slot, int np, const Beam *__restrict__ bm, int nb,
                              cd *__restrict__ Ui, cd *__restrict__ dUi, cd *__restrict__ Ui0, cd *__restrict__ dUi0,
                              cd *__restrict__ Uip)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  cd e[3], h[3], de[3], dh[3];
  beam_EH(bm, nb, pts + 3 * p, nrm + 3 * p, e, h, de, dh);
  size_t o = (size_t)slot[p] * 4;
  for (int c = 0; c < 3; c++) {
    Ui[o + c] = e[c]; dUi[o + c] = de[c]; Ui0[o + c] = e[c]; dUi0[o + c] = de[c]; Uip[o + c] = e[c];
  }
  Ui[o + 3] = mk(0, 0); dUi[o + 3] = mk(0, 0); Ui0[o + 3] = mk(0, 0); dUi0[o + 3] = mk(0, 0); Uip[o + 3] = mk(0, 0);
}

// add the incident field to open-region points (what 0: E,H rows of 6)
__global__ void k_add_incident(const double *__restrict__ pts, const int *__restrict__ mask, int np,
                               const Beam *__restrict__ bm, int nb, cd *__restrict__ out, int what)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np || !mask[p]) return;
  cd e[3], h[3];
  beam_EH(bm, nb, pts + 3 * p, nullptr, e, h, nullptr, nullptr);
  if (what == 0)
    for (int j = 0; j < 3; j++) { out[6 * (size_t)p + j] = out[6 * (size_t)p + j] + e[j]; out[6 * (size_t)p + 3 + j] = out[6 * (size_t)p + 3 + j] + h[j]; }
  else
    for (int j = 0; j < 3; j++) out[4 * (size_t)p + j] = out[4 * (size_t)p + j] + e[j];
}

// ---------------------------------------------------------------- other beams
// E, H (and their derivatives along v) of the focused, spiral-phase, Bessel
// and radially / azimuthally polarised beams (fpw.c:3-382 and the same
// functions of lgb.c, bsb.c, bslgb.c, rab.c): a spectrum of plane waves over
// the aperture (Gauss-Legendre points in sin(theta), or the one cone angle of
// a Bessel beam) and, per spectral angle, the reference's adaptive composite
// trapezoid in phi: halving steps over [-pi, pi) until the change of
// n^2 |E|^2 + |H|^2 falls below 1e-15 sqrt(points) times its largest value,
// after at least two halvings; more than 12 halvings is an error.  One
// wavefront per point, one lane per spectral angle.
__device__ __forceinline__ void fb_ac(const FBeam &b, double ct, double st, const double *__restrict__ phit, int j,
                                      const double *x, cd *eh)
{
  const double cp = phit[3 * j], sp = phit[3 * j + 1];
  double sa, ca;
  sincos(b.ki * (st * cp * x[0] + st * sp * x[1] + ct * x[2]), &sa, &ca);
  cd ce = mk(ca, sa);
  if ((b.type == 2 || b.type == 4) && b.lg_m != 0) {
    // exp(i m phi) with phi taken into [0, 2 pi) (lgb.c:177-187)
    const double a = phit[3 * j + 2], pp = a < 0.0 ? a + 2.0 * M_PI : a;
    double sm, cm;
    sincos(fmod((double)b.lg_m * pp, 2.0 * M_PI), &sm, &cm);
    ce = ce * mk(cm, sm);
  }
  if (b.type == 5) {
    eh[0] = (b.ex * (ct * cp) - b.ey * sp) * ce;
    eh[1] = (b.ex * (ct * sp) + b.ey * cp) * ce;
    eh[2] = (b.ex * (-st)) * ce;
    eh[3] = (b.hx * (ct * cp) - b.hy * sp) * ce;
    eh[4] = (b.hx * (ct * sp) + b.hy * cp) * ce;
    eh[5] = (b.hx * (-st)) * ce;
  } else {
    const double m11 = sp * sp + ct * cp * cp, m12 = sp * cp * (ct - 1.0), m22 = ct * sp * sp + cp * cp;
    eh[0] = (b.ex * m11 + b.ey * m12) * ce;
    eh[1] = (b.ex * m12 + b.ey * m22) * ce;
    eh[2] = (b.ex * (-st * cp) - b.ey * (st * sp)) * ce;
    eh[3] = (b.hx * m11 + b.hy * m12) * ce;
    eh[4] = (b.hx * m12 + b.hy * m22) * ce;
    eh[5] = (b.hx * (-st * cp) - b.hy * (st * sp)) * ce;
  }
}

// phi integral at one spectral angle; dIt: the derivative along v (each
// component times i k (direction . v)); false when it does not converge
template <bool DV>
__device__ bool fb_phi(const FBeam &b, double ct, double st, const double *__restrict__ phit, const double *x,
                       const double *v, cd *It, cd *dIt)
{
  const double cep = b.ni * b.ni;
  cd Ia[6];
  double h = 2.0 * M_PI;
  fb_ac(b, ct, st, phit, 0, x, Ia);
  auto gfc = [&](int j) {
    const double cp = phit[3 * j], sp = phit[3 * j + 1];
    return mk(0.0, b.ki * (st * cp * v[0] + st * sp * v[1] + ct * v[2]));
  };
  double abst = 0.0;
  {
    const cd g = DV ? gfc(0) : mk(0, 0);
#pragma unroll
    for (int i = 0; i < 6; i++) {
      It[i] = Ia[i] * h;
      if (DV) dIt[i] = (Ia[i] * g) * h;
      abst += (i < 3 ? cep : 1.0) * (It[i].x * It[i].x + It[i].y * It[i].y);
    }
  }
  double f_order = abst;
  int j = 1, sc = 1, cc = 0;
  while (cc <= BEM_TRAP_MNUM) {
    h *= 0.5;
    cd Is[6], dIs[6];
#pragma unroll
    for (int i = 0; i < 6; i++) { Is[i] = mk(0, 0); dIs[i] = mk(0, 0); }
    for (int s = 1; s <= sc; s++, j++) {
      fb_ac(b, ct, st, phit, j, x, Ia);
      const cd g = DV ? gfc(j) : mk(0, 0);
#pragma unroll
      for (int i = 0; i < 6; i++) { Is[i] = Is[i] + Ia[i]; if (DV) dIs[i] = dIs[i] + Ia[i] * g; }
    }
    double abss = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
      It[i] = It[i] * 0.5 + Is[i] * h;
      if (DV) dIt[i] = dIt[i] * 0.5 + dIs[i] * h;
      abss += (i < 3 ? cep : 1.0) * (It[i].x * It[i].x + It[i].y * It[i].y);
    }
    const double crc = sqrt((double)sc) * f_order;
    if ((cc > 1 && fabs(abst - abss) < 1.0e-15 * crc) || (abss == 0.0 && abst == 0.0)) break;
    abst = abss;
    if (abst > f_order) f_order = abst;
    sc *= 2;
    cc++;
  }
  return cc < BEM_TRAP_MNUM;
}

// rotation into the beam frame (fpw.c:13-17) and back (fpw.c:21-26)
__device__ __forceinline__ void fb_rot(const FBeam &b, const double *a, double *o)
{
  const double ct = b.cos_t, st = b.sin_t, cp = b.cos_p, sp = b.sin_p;
  o[0] = a[0] * (ct * cp * cp + sp * sp) + a[1] * sp * cp * (ct - 1.0) - a[2] * st * cp;
  o[1] = a[0] * sp * cp * (ct - 1.0) + a[1] * (ct * sp * sp + cp * cp) - a[2] * st * sp;
  o[2] = a[0] * st * cp + a[1] * st * sp + a[2] * ct;
}
__device__ __forceinline__ void fb_back(const FBeam &b, const cd *a, cd *o)
{
  const double ct = b.cos_t, st = b.sin_t, cp = b.cos_p, sp = b.sin_p;
  o[0] = a[0] * (ct * cp * cp + sp * sp) + a[1] * (sp * cp * (ct - 1.0)) + a[2] * (st * cp);
  o[1] = a[0] * (sp * cp * (ct - 1.0)) + a[1] * (ct * sp * sp + cp * cp) + a[2] * (st * sp);
  o[2] = a[0] * (-st * cp) - a[1] * (st * sp) + a[2] * ct;
}

// mode 0: add E, H to rows of 6 of out; 1: add E to rows of 4 (the
// potentials' vector part); 2: add E (and dE along nrm) to the incident
// boundary data of the node slots slot[p] (components 0-2 of Ui, Ui0, Uip
// and dUi, dUi0)
template <bool DV>
__global__ void __launch_bounds__(256) k_fbeam(const double *__restrict__ pts, const double *__restrict__ nrm,
                                               const int *__restrict__ mask, int np, const FBeam *__restrict__ fb,
                                               int nf, const double *__restrict__ tht, const double *__restrict__ phit,
                                               int *__restrict__ err, int mode, cd *__restrict__ out,
                                               const int *__restrict__ slot, cd *__restrict__ Ui,
                                               cd *__restrict__ dUi, cd *__restrict__ Ui0, cd *__restrict__ dUi0,
                                               cd *__restrict__ Uip)
{
  const int p = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (p >= np) return;
  if (mask && !mask[p]) return;
  const double x[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
  const double v[3] = {DV ? nrm[3 * p] : 0.0, DV ? nrm[3 * p + 1] : 0.0, DV ? nrm[3 * p + 2] : 0.0};
  cd E[3], H[3], dE[3], dH[3];
#pragma unroll
  for (int i = 0; i < 3; i++) { E[i] = H[i] = dE[i] = dH[i] = mk(0, 0); }
  bool ok = true;
  for (int ib = 0; ib < nf; ib++) {
    const FBeam b = fb[ib];
    const double xb[3] = {x[0] - b.fx, x[1] - b.fy, x[2] - b.fz};
    double xn[3], vn[3] = {0.0, 0.0, 0.0};
    fb_rot(b, xb, xn);
    if (DV) fb_rot(b, v, vn);
    const bool bessel = b.type == 3 || b.type == 4;
    const int nth = bessel ? 1 : b.nt;
    cd s[6], ds[6];
#pragma unroll
    for (int i = 0; i < 6; i++) { s[i] = mk(0, 0); ds[i] = mk(0, 0); }
    for (int j = lane; j < nth; j += 64) {
      const double ct = bessel ? b.ct : tht[3 * (b.toff + j)];
      const double st = bessel ? b.st : tht[3 * (b.toff + j) + 1];
      const double w = bessel ? b.st : tht[3 * (b.toff + j) + 2];
      cd It[6], dIt[6];
      ok = fb_phi<DV>(b, ct, st, phit, xn, vn, It, dIt) && ok;
#pragma unroll
      for (int i = 0; i < 6; i++) { s[i] = s[i] + It[i] * w; if (DV) ds[i] = ds[i] + dIt[i] * w; }
    }
#pragma unroll
    for (int i = 0; i < 6; i++) { s[i] = wsum(s[i]); if (DV) ds[i] = wsum(ds[i]); }
    cd te[3], th[3];
    fb_back(b, s, te);
    fb_back(b, s + 3, th);
#pragma unroll
    for (int i = 0; i < 3; i++) { E[i] = E[i] + te[i] * b.E0; H[i] = H[i] + th[i] * b.E0; }
    if (DV) {
      fb_back(b, ds, te);
      fb_back(b, ds + 3, th);
#pragma unroll
      for (int i = 0; i < 3; i++) { dE[i] = dE[i] + te[i] * b.E0; dH[i] = dH[i] + th[i] * b.E0; }
    }
  }
  if (!ok) atomicOr(err, 1);
  if (lane != 0) return;
  if (mode == 0) {
    cd *o = out + 6 * (size_t)p;
#pragma unroll
    for (int i = 0; i < 3; i++) { o[i] = o[i] + E[i]; o[3 + i] = o[3 + i] + H[i]; }
  } else if (mode == 1) {
    cd *o = out + 4 * (size_t)p;
#pragma unroll
    for (int i = 0; i < 3; i++) o[i] = o[i] + E[i];
  } else {
    const size_t o = (size_t)slot[p] * 4;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      Ui[o + i] = Ui[o + i] + E[i]; Ui0[o + i] = Ui0[o + i] + E[i]; Uip[o + i] = Uip[o + i] + E[i];
      dUi[o + i] = dUi[o + i] + dE[i]; dUi0[o + i] = dUi0[o + i] + dE[i];
    }
  }
}

// ---------------------------------------------------------------- Maxwell stress
// per surface point: T = 1/2 (eps E E* + H H*) - 1/4 (eps|E|^2 + |H|^2) I;
// contributions w * (T . W) and w * (r - rc) x (T . W), scaled per element
// (-1/2 on triangles); per-element partial rows of 6 are summed in order.
__global__ void k_stress(const cd *__restrict__ EH, const double *__restrict__ rpt, const double *__restrict__ wpt,
                         const double *__restrict__ wgt, const int *__restrict__ elem, int np, double epd,
                         const double *__restrict__ rc, double *__restrict__ part)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const cd *e = EH + 6 * (size_t)p, *h = e + 3;
  double aE2 = abs2(e[0]) + abs2(e[1]) + abs2(e[2]), aH2 = abs2(h[0]) + abs2(h[1]) + abs2(h[2]);
  double Tc = 0.25 * (epd * aE2 + aH2), T[9];
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) T[3 * j + k] = 0.5 * (epd * (e[j] * conjd(e[k])).x + (h[j] * conjd(h[k])).x);
  T[0] -= Tc; T[4] -= Tc; T[8] -= Tc;
  const double *w = wpt + 3 * p, *r = rpt + 3 * p;
  double d[3] = {r[0] - rc[0], r[1] - rc[1], r[2] - rc[2]}, wt = wgt[p];
  double *o = part + 6 * (size_t)p;
  for (int j = 0; j < 3; j++) {
    o[j] = (T[3 * j] * w[0] + T[3 * j + 1] * w[1] + T[3 * j + 2] * w[2]) * wt;
    o[3 + j] = ((d[1] * T[3 * j + 2] - d[2] * T[3 * j + 1]) * w[0] + (d[2] * T[3 * j] - d[0] * T[3 * j + 2]) * w[1] +
                (d[0] * T[3 * j + 1] - d[1] * T[3 * j]) * w[2]) * wt;
  }
  (void)elem;
}

// Poynting flux per surface point: w * Re(E x H*) . W
__global__ void k_poynting(const cd *__restrict__ EH, const double *__restrict__ wpt, const double *__restrict__ wgt,
                           int np, double *__restrict__ part)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const cd *e = EH + 6 * (size_t)p, *h = e + 3;
  double vp[3] = {(e[1] * conjd(h[2]) - e[2] * conjd(h[1])).x, (e[2] * conjd(h[0]) - e[0] * conjd(h[2])).x,
                  (e[0] * conjd(h[1]) - e[1] * conjd(h[0])).x};
  const double *w = wpt + 3 * p;
  part[p] = (vp[0] * w[0] + vp[1] * w[1] + vp[2] * w[2]) * wgt[p];
}

// ---------------------------------------------------------------- launchers
extern "C" {
void bem_launch_domain(const double *pts, int np, const double *fv, const int *fsd, const int *fq, const int *segStart,
                       int MN, int *did, hipStream_t st)
{
  if (np <= 0) return;
  hipLaunchKernelGGL(k_domain, dim3((np + 3) / 4), dim3(256), 0, st, pts, np, fv, fsd, fq, segStart, MN, did);
}
// gq, qrule, nqq: quadrangle terms and the square's rule of the same type
// gl32 != nullptr: far triangles by the FP32 reduced far rules (k_field<true>)
void bem_launch_field(const double *pts, const int *pdid, int np, int want_did, const void *geo, const int *sid,
                      int dNe, double kr, double ki, const void *rule, int nq, const void *gl32, double scale,
                      const void *gq, const void *qrule, int nqq, int what, const void *A0, const void *dA0,
                      const void *A1, const void *dA1, void *out, int accumulate, hipStream_t st)
{
  if (np <= 0) return;
  bem_prof_begin(BH_K_FIELD, (double)np * dNe, st);
  if (gl32)
    hipLaunchKernelGGL(k_field<true>, dim3((np + 3) / 4), dim3(256), 0, st, pts, pdid, np, want_did, (const TriGeo *)geo,
                       sid, dNe, kr, ki, (const QPtN *)rule, nq, (const QPt32 *)gl32, scale, (const QuadExt *)gq,
                       (const QPtQ *)qrule, nqq, what, (const cd *)A0, (const cd *)dA0, (const cd *)A1, (const cd *)dA1,
                       (cd *)out, accumulate);
  else
    hipLaunchKernelGGL(k_field<false>, dim3((np + 3) / 4), dim3(256), 0, st, pts, pdid, np, want_did,
                       (const TriGeo *)geo, sid, dNe, kr, ki, (const QPtN *)rule, nq, (const QPt32 *)nullptr, scale,
                       (const QuadExt *)gq, (const QPtQ *)qrule, nqq, what, (const cd *)A0, (const cd *)dA0,
                       (const cd *)A1, (const cd *)dA1, (cd *)out, accumulate);
  bem_prof_end(BH_K_FIELD, st);
}
// qrule / qgh: the square's rules matching rule / gh, for quadrangles
void bem_launch_field_memp(const double *pts, const int *pdid, int np, int want_did, const void *geo, const void *gq,
                           const int *sid, int dNe, double kr, double ki, const void *rule, int nq, double scale, int sw,
                           const void *gh, int ngh, double ghscale, const void *qrule, int nqq, const void *qgh,
                           int nqgh, const void *U, const void *dU, double lam0, void *out, hipStream_t st)
{
  if (np <= 0) return;
  bem_prof_begin(BH_K_FIELD, (double)np * dNe, st);
  hipLaunchKernelGGL(k_field_memp, dim3((np + 3) / 4), dim3(256), 0, st, pts, pdid, np, want_did, (const TriGeo *)geo,
                     (const QuadExt *)gq, sid, dNe, kr, ki, (const QPtN *)rule, nq, scale, sw, (const QPtN *)gh, ngh,
                     ghscale, (const QPtQ *)qrule, nqq, (const QPtQ *)qgh, nqgh, (const cd *)U, (const cd *)dU, lam0,
                     (cd *)out);
  bem_prof_end(BH_K_FIELD, st);
}
// gl32 != nullptr: the mixed-precision form (FP32 far rules); part
// (np x 19 doubles) != nullptr: far triangles in a first pass
void bem_launch_scatter(const double *pts, const double *nrm, const int *slot, int np, const void *geo, const void *gq,
                        const int *sid, int dNe, double kr, double ki, const void *gl, int ngl, const void *gh, int ngh,
                        double scale, const void *qgl, int nqgl, const void *qgh, int nqgh, const void *gl32,
                        const void *U, const void *dU, void *Ui, void *dUi, int quads, double *part, hipStream_t st)
{
  if (np <= 0) return;
  const int kid = gl32 ? BH_K_SCATTER32 : BH_K_SCATTER;
  if (part) {
    // the far triangles first (their own timing record), then the rest
    const int kf = gl32 ? BH_K_SCATTER_FAR32 : BH_K_SCATTER_FAR;
    bem_prof_begin(kf, (double)np * dNe, st);
    hipLaunchKernelGGL(gl32 ? k_scatter_far<true> : k_scatter_far<false>, dim3((np + 3) / 4), dim3(256), 0, st, pts, nrm,
                       np, (const TriGeo *)geo, sid, dNe, kr, ki, scale, (const QPt32 *)gl32, (const QPtN *)gl, ngl,
                       (const cd *)U, (const cd *)dU, part);
    bem_prof_end(kf, st);
  }
  // the second pass carries the pairs only when it sums them all
  bem_prof_begin(kid, part ? 0.0 : (double)np * dNe, st);
  hipLaunchKernelGGL(quads ? k_scatter<true> : k_scatter<false>, dim3((np + 3) / 4), dim3(256), 0, st, pts, nrm, slot, np, (const TriGeo *)geo,
                     (const QuadExt *)gq, sid, dNe, kr, ki, (const QPtN *)gl, ngl, (const QPtN *)gh, ngh, scale,
                     (const QPtQ *)qgl, nqgl, (const QPtQ *)qgh, nqgh, (const QPt32 *)gl32, (const cd *)U,
                     (const cd *)dU, (const double *)part, (cd *)Ui, (cd *)dUi);
  bem_prof_end(kid, st);
}
// gl32 != nullptr: the mixed-precision form (FP32 copy of the same rule);
// gq, qgl: quadrangle terms and the square's GLN rule
void bem_launch_field_bd(const int *tgt_t, const double *tgt_ze, int np, int did, const void *geo, const void *gq,
                         const int *sid, int dNe, double kr, double ki, const void *gl, const void *gl32, int ngl,
                         double scale, const void *qgl, int nqgl, const double *xl, const double *wl, const double *xh,
                         const double *wh, const void *E, const void *dE, const void *H, const void *dH, void *out,
                         double *rt_out, hipStream_t st)
{
  if (np <= 0) return;
  const int kid = gl32 ? BH_K_FIELD_BD32 : BH_K_FIELD_BD;
  bem_prof_begin(kid, (double)np * dNe, st);
  if (gl32)
    hipLaunchKernelGGL(k_field_bd<true>, dim3((np + 3) / 4), dim3(256), 0, st, tgt_t, tgt_ze, np, did, (const TriGeo *)geo,
                       (const QuadExt *)gq, sid, dNe, kr, ki, (const QPtN *)gl, (const QPt32 *)gl32, ngl, scale,
                       (const QPtQ *)qgl, nqgl, xl, wl, xh, wh, (const cd *)E, (const cd *)dE, (const cd *)H,
                       (const cd *)dH, (cd *)out, rt_out);
  else
    hipLaunchKernelGGL(k_field_bd<false>, dim3((np + 3) / 4), dim3(256), 0, st, tgt_t, tgt_ze, np, did,
                       (const TriGeo *)geo, (const QuadExt *)gq, sid, dNe, kr, ki, (const QPtN *)gl,
                       (const QPt32 *)nullptr, ngl, scale, (const QPtQ *)qgl, nqgl, xl, wl, xh, wh, (const cd *)E,
                       (const cd *)dE, (const cd *)H, (const cd *)dH, (cd *)out, rt_out);
  bem_prof_end(kid, st);
}
void bem_launch_incident_bd(const double *pts, const double *nrm, const int *slot, int np, const void *bm, int nb,
                            void *Ui, void *dUi, void *Ui0, void *dUi0, void *Uip, hipStream_t st)
{
  if (np <= 0) return;
  hipLaunchKernelGGL(k_incident_bd, dim3((np + 255) / 256), dim3(256), 0, st, pts, nrm, slot, np, (const Beam *)bm, nb,
                     (cd *)Ui, (cd *)dUi, (cd *)Ui0, (cd *)dUi0, (cd *)Uip);
}
void bem_launch_add_incident(const double *pts, const int *mask, int np, const void *bm, int nb, void *out, int what,
                             hipStream_t st)
{
  if (np <= 0) return;
  hipLaunchKernelGGL(k_add_incident, dim3((np + 255) / 256), dim3(256), 0, st, pts, mask, np, (const Beam *)bm, nb,
                     (cd *)out, what);
}
// the other beams' fields at np points (see k_fbeam for mode); err is set
// when a phi integral does not converge
void bem_launch_fbeam(const double *pts, const double *nrm, const int *mask, int np, const void *fb, int nf,
                      const double *tht, const double *phit, int *err, int mode, void *out, const int *slot, void *Ui,
                      void *dUi, void *Ui0, void *dUi0, void *Uip, hipStream_t st)
{
  if (np <= 0 || nf <= 0) return;
  bem_prof_begin(BH_K_FIELD, (double)np * nf, st);
  if (nrm)
    hipLaunchKernelGGL(k_fbeam<true>, dim3((np + 3) / 4), dim3(256), 0, st, pts, nrm, mask, np, (const FBeam *)fb, nf,
                       tht, phit, err, mode, (cd *)out, slot, (cd *)Ui, (cd *)dUi, (cd *)Ui0, (cd *)dUi0, (cd *)Uip);
  else
    hipLaunchKernelGGL(k_fbeam<false>, dim3((np + 3) / 4), dim3(256), 0, st, pts, nrm, mask, np, (const FBeam *)fb,
                       nf, tht, phit, err, mode, (cd *)out, slot, (cd *)Ui, (cd *)dUi, (cd *)Ui0, (cd *)dUi0, (cd *)Uip);
  bem_prof_end(BH_K_FIELD, st);
}
void bem_launch_stress(const void *EH, const double *rpt, const double *wpt, const double *wgt, int np, double epd,
                       const double *rc, double *part, hipStream_t st)
{
  if (np <= 0) return;
  hipLaunchKernelGGL(k_stress, dim3((np + 255) / 256), dim3(256), 0, st, (const cd *)EH, rpt, wpt, wgt, nullptr, np, epd,
                     rc, part);
}
void bem_launch_poynting(const void *EH, const double *wpt, const double *wgt, int np, double *part, hipStream_t st)
{
  if (np <= 0) return;
  hipLaunchKernelGGL(k_poynting, dim3((np + 255) / 256), dim3(256), 0, st, (const cd *)EH, wpt, wgt, np, part);
}
}

extern "C" {
void bem_launch_field_rows(const double *pts, const int *tgt_t, const double *tgt_ze, int np, int did, const void *geo,
                        const void *gq, const int *sid, int dNe, double kr, double ki, const void *gl, int ngl,
                        const void *glr, double scale, const void *qgl, int nqgl, const double *xl, const double *wl,
                        const double *xh, const double *wh, size_t ld, void *R, double *F, hipStream_t st)
{
  if (np <= 0) return;
  bem_prof_begin(BH_K_FIELD_ROWS, (double)np * dNe, st);
  if (tgt_t)
    hipLaunchKernelGGL(k_field_rows<true>, dim3((np + 3) / 4), dim3(256), 0, st, pts, tgt_t, tgt_ze, np, did,
                       (const TriGeo *)geo, (const QuadExt *)gq, sid, dNe, kr, ki, (const QPtN *)gl, ngl,
                       (const QPt *)glr, scale, (const QPtQ *)qgl, nqgl, xl, wl, xh, wh, ld, (cd *)R, F);
  else
    hipLaunchKernelGGL(k_field_rows<false>, dim3((np + 3) / 4), dim3(256), 0, st, pts, tgt_t, tgt_ze, np, did,
                       (const TriGeo *)geo, (const QuadExt *)gq, sid, dNe, kr, ki, (const QPtN *)gl, ngl,
                       (const QPt *)glr, scale, (const QPtQ *)qgl, nqgl, xl, wl, xh, wh, ld, (cd *)R, F);
  bem_prof_end(BH_K_FIELD_ROWS, st);
}
void bem_launch_pack_ehbv(const void *E, const void *dE, const void *H, const void *dH, size_t es, int dNe, int nb,
                          void *X, hipStream_t st)
{
  const size_t n = 8 * ((size_t)dNe + 1) * nb;
  hipLaunchKernelGGL(k_pack_ehbv, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (const cd *)E, (const cd *)dE,
                     (const cd *)H, (const cd *)dH, es, dNe, nb, (cd *)X);
}
void bem_launch_rows_add(const void *C, int n, int nb, const double *F, int open, const int *idx, size_t ldb, void *EH,
                         hipStream_t st)
{
  if (n <= 0) return;
  const size_t m = (size_t)n * nb;
  hipLaunchKernelGGL(k_rows_add, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, (const cd *)C, n, nb, F, open,
                     idx, ldb, (cd *)EH);
}
void bem_launch_elem_coef(int nd, const void *geo, const void *gq, int n, const int *s, const double *rt,
                          const double *v0, const double *v1, const int *idx, double kr, double ki, const void *rule, int nq,
                          double scale, const void *qrule, int nqq, const void *rgh, int ngh, double sgh, const void *qgh,
                          int nqgh, int fallback, void *out, hipStream_t st)
{
  if (n <= 0) return;
  const dim3 grid((n + 3) / 4), blk(256);
  if (nd == 0)
    hipLaunchKernelGGL(k_elem_coef<0>, grid, blk, 0, st, (const TriGeo *)geo, (const QuadExt *)gq, n, s, rt, v0, v1, idx,
                       kr, ki, (const QPtN *)rule, nq, scale, (const QPtQ *)qrule, nqq, (const QPtN *)rgh, ngh, sgh,
                       (const QPtQ *)qgh, nqgh, fallback, (cd *)out);
  else if (nd == 1)
    hipLaunchKernelGGL(k_elem_coef<1>, grid, blk, 0, st, (const TriGeo *)geo, (const QuadExt *)gq, n, s, rt, v0, v1, idx,
                       kr, ki, (const QPtN *)rule, nq, scale, (const QPtQ *)qrule, nqq, (const QPtN *)rgh, ngh, sgh,
                       (const QPtQ *)qgh, nqgh, fallback, (cd *)out);
  else
    hipLaunchKernelGGL(k_elem_coef<2>, grid, blk, 0, st, (const TriGeo *)geo, (const QuadExt *)gq, n, s, rt, v0, v1, idx,
                       kr, ki, (const QPtN *)rule, nq, scale, (const QPtQ *)qrule, nqq, (const QPtN *)rgh, ngh, sgh,
                       (const QPtQ *)qgh, nqgh, fallback, (cd *)out);
}
}  // extern "C"
