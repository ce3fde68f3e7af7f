// bem_mo.hip -- the coupling loop of several objects (mo_solve_bv,
// mo_solve.c:10-88) on one MI355X.
//
// Each sweep adds to every object's incident data the field scattered by
// every other object at its open-domain node slots (add_scattered_mEMP and
// mEMP2_s, mo_solve.c:296-396), then solves every object again.  Two ways to
// form those coupling sums:
//
//   * direct: the reference's element quadrature over every (slot, source
//     element) pair (k_scatter_far + k_scatter, bem_field.hip);
//   * multipole (the mixed-precision form, source objects of triangles in a
//     lossless surrounding medium): the field of source object n outside
//     its bounding sphere is the series
//        U(x) = sum_{l<=L} sum_{|m|<=l} A_lm h_l(k|x-c|) Y_l^m(x-c)
//     whose coefficients collect the same quadrature points the direct sum
//     uses (the reduced far rules): with the point monopoles a (from dU) and
//     dipoles b W (from U) of the element integrals,
//        A_lm = 4 pi i k sum_q [a_q conj(R_l^m(y_q)) + b_q W . grad conj(R_l^m)(y_q)],
//     R_l^m = j_l Y_l^m.  The addition theorem reproduces the direct
//     discrete sum up to a truncation error of about rho^L, rho = the source
//     sphere's radius over the distance of the slot from its centre.  A pair
//     of objects takes the series when rho <= 0.6 for all its slots and the
//     gap to the sphere is at least 4 element radii (so every pair is a far
//     pair of the direct sum, and its near-pair switch never fires there);
//     L = max(ln 1e-10 / ln rho, kr + 6 (kr)^(1/3) + 4), at most 50.  The
//     solid-angle terms F, dF of the direct sum (a closed surface seen from
//     outside: zero up to quadrature error) are taken as zero there.
//
// Work per sweep of the series: a basis kernel (lane per source point)
// writes conj(R_l^m) of every point, a moment kernel contracts them with the
// 16 point densities (monopole and three dipole components of each of the 4
// potentials), a coefficient kernel folds the dipoles' gradients into A and
// forms the gradient coefficient sets (d_z, d_+ = d_x + i d_y, d_-), and one
// evaluation launch over all destination slots (lane per slot, the source
// object's coefficients wave-uniform) adds U and dU/dn into the incident
// data.  All series arithmetic is FP64.
//
// Spherical functions: Y_l^m orthonormal with the Condon-Shortley phase,
// Y_l^{-m} = (-1)^m conj(Y_l^m); Y_l^m = P^_l^m(cos t) w^m with w = (x+iy)/r
// and P^ = Pbar / sin^m (a polynomial in cos t: no pole).  Gradients:
//   d_z F_l^m = k [cz1 F_{l-1}^m + cz2 F_{l+1}^m],
//   d_+ F_l^m = k [cp1 F_{l-1}^{m+1} + cp2 F_{l+1}^{m+1}],
//   d_- F_l^m = k [cm1 F_{l-1}^{m-1} + cm2 F_{l+1}^{m-1}]
// for F = j Y or h Y (coefficients below).
#include "bem_obj.h"
#include "bem_prof.h"

#include <atomic>
#include <chrono>
#include <string>
#include <thread>

using namespace bem;

extern "C" {
void bem_zgemm(int, int, int, double, double, const void *, size_t, const void *, size_t, double, double, void *, size_t,
               hipStream_t);
void bem_launch_scatter(const double *, const double *, const int *, int, const void *, const void *, const int *, int,
                        double, double, const void *, int, const void *, int, double, const void *, int, const void *,
                        int, const void *, const void *, const void *, void *, void *, int, double *, hipStream_t);
}

#define MP_LMAX 50              // largest series order
#define MP_LPMAX (MP_LMAX + 1)  // moments and gradient sets run one order further
#define MP_RHO 0.6              // largest source-sphere radius / slot distance
#define MP_RHO_SPLIT 0.5        // the same for the tiles of a split pair (keeps its order near the others')
#define MP_EPS 1e-10            // truncation target (relative)
#define MP_GAP_RADII 4.0        // gap to the source sphere in element radii
#define MP_NDENS 16             // point densities: 4 potentials x (monopole, dipole z, +, -)
#define MP_QSPLIT 256           // source points per moment workgroup

// ------------------------------------------------------------------ series helpers
__host__ __device__ __forceinline__ int mp_idx(int n, int m, int Lp)
{
  if (m == 0) return n;
  const int am = m < 0 ? -m : m;
  return (Lp + 1) + (am - 1) * (2 * Lp + 2 - am) + 2 * (n - am) + (m < 0 ? 1 : 0);
}
__host__ __device__ __forceinline__ int mp_count(int Lp) { return (Lp + 1) * (Lp + 1); }

__device__ __forceinline__ double cz1(int n, int m) { return n > 0 ? sqrt(fmax(0.0, (double)((n + m) * (n - m))) / ((2.0 * n - 1) * (2.0 * n + 1))) : 0.0; }
__device__ __forceinline__ double cz2(int n, int m) { return -sqrt((double)((n + m + 1) * (n - m + 1)) / ((2.0 * n + 1) * (2.0 * n + 3))); }
__device__ __forceinline__ double cp1(int n, int m) { return n > 0 ? sqrt(fmax(0.0, (double)((n - m - 1) * (n - m))) / ((2.0 * n - 1) * (2.0 * n + 1))) : 0.0; }
__device__ __forceinline__ double cp2(int n, int m) { return sqrt((double)((n + m + 1) * (n + m + 2)) / ((2.0 * n + 1) * (2.0 * n + 3))); }
__device__ __forceinline__ double cm1(int n, int m) { return n > 0 ? -sqrt(fmax(0.0, (double)((n + m - 1) * (n + m))) / ((2.0 * n - 1) * (2.0 * n + 1))) : 0.0; }
__device__ __forceinline__ double cm2(int n, int m) { return -sqrt((double)((n - m + 1) * (n - m + 2)) / ((2.0 * n + 1) * (2.0 * n + 3))); }

// recurrence constants of P^_n^m: P^_n^m = a (u P^_{n-1}^m - b P^_{n-2}^m)
__device__ __forceinline__ double pa(int n, int m) { return sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m)); }
__device__ __forceinline__ double pb(int n, int m) { return sqrt(((n - 1.0) * (n - 1.0) - (double)m * m) / (4.0 * (n - 1.0) * (n - 1.0) - 1.0)); }

// spherical Bessel j_0..j_Lp(x) into t[n * stride]: downward recurrence
// (Miller), rescaled on the way down against overflow, normalised by the
// sum rule sum_n (2n+1) j_n^2 = 1 and signed by j_0 or j_1, whichever is the
// larger (no division by a zero of j_0)
__device__ void sph_j(double x, int Lp, double *t, int stride)
{
  if (x < 1e-12) {
    t[0] = 1.0;
    for (int n = 1; n <= Lp; n++) t[n * stride] = 0.0;
    return;
  }
  const int N = Lp + 24 + (int)(1.5 * x);
  double jp1 = 0.0, j = 1e-100, ix = 1.0 / x, sum = (2.0 * N + 1.0) * j * j, j1 = 0.0;
  for (int n = N; n >= 1; n--) {
    const double jm1 = (2.0 * n + 1.0) * ix * j - jp1;
    jp1 = j;
    j = jm1;
    sum += (2.0 * n - 1.0) * j * j;
    if (n - 1 <= Lp) t[(n - 1) * stride] = j;
    if (n == 2) j1 = j;
    if (fabs(j) > 1e100) {   // values stay within 1e+-100, their squares within 1e+-200
      j *= 1e-100;
      jp1 *= 1e-100;
      j1 *= 1e-100;
      sum *= 1e-200;
      for (int q = n - 1; q <= Lp; q++) t[q * stride] *= 1e-100;
    }
  }
  double s, c;
  sincos(x, &s, &c);
  const double j0t = s * ix, j1t = (s * ix - c) * ix;
  double f = 1.0 / sqrt(sum);
  if (fabs(j0t) >= fabs(j1t) ? (j0t * j < 0.0) : (j1t * j1 < 0.0)) f = -f;
  for (int n = 0; n <= Lp; n++) t[n * stride] *= f;
}

// spherical Hankel h_0..h_Lp(x) (first kind, upward recurrence)
__device__ void sph_h(double x, int Lp, cd *t, int stride)
{
  double s, c;
  sincos(x, &s, &c);
  const double ix = 1.0 / x;
  cd h0 = mk(s * ix, -c * ix);                                  // -i e^{ix} / x
  cd h1 = mk((s * ix - c) * ix, -(c * ix + s) * ix);           // -e^{ix} (x + i) / x^2
  t[0] = h0;
  if (Lp >= 1) t[stride] = h1;
  for (int n = 1; n < Lp; n++) {
    const cd h2 = h1 * ((2.0 * n + 1.0) * ix) - h0;
    t[(n + 1) * stride] = h2;
    h0 = h1;
    h1 = h2;
  }
}

// ------------------------------------------------------------------ source side
// One source object's series data.
struct MpSrc {
  double c[3];      // expansion centre
  double kr;        // wavenumber of the surrounding medium (real)
  double scale;     // rule scale of the element integrals
  int Lp;           // series order + 1 (moments, gradient sets)
  int L;            // series order
  int qoff, npt;    // far rule: offset in the rule table, points per element
  int dNe;          // source elements (domain 0)
  int nq;           // source points = dNe * npt
  int nsplit;       // moment workgroups over the points
  int pad;
};

// lane per source point q (element 1 + q / npt, rule point q % npt): the
// 16 densities and conj(R_n^m(y - c)) of every (n, m) (layout [q][idx])
// ROWQ: B[q][idx] (rows per point, the moment kernel's layout); otherwise
// B[idx][q] (row-major idx x points: the A operand of the many-beam moment
// product).  dens null: no densities (the many-beam path forms its own).
template <bool ROWQ, class R>
__global__ void __launch_bounds__(64) k_mp_basis(MpSrc S, const TriGeo *__restrict__ geo, const int *__restrict__ sid,
                                                 const R *__restrict__ rule, const cd *__restrict__ U,
                                                 const cd *__restrict__ dU, int q0, int nqc, cd *__restrict__ B,
                                                 size_t ldb, cd *__restrict__ dens)
{
  __shared__ double jt[(MP_LPMAX + 1) * 64];
  const int lane = threadIdx.x;
  const int ql = blockIdx.x * 64 + lane;
  if (ql >= nqc) return;
  const int q = q0 + ql;
  const int s = 1 + q / S.npt, p = q % S.npt;
  const int sd = sid[s], as = sd < 0 ? -sd : sd;
  const TriGeo g = geo[as];
  const R qp = rule[S.qoff + p];
  double y[3];
#pragma unroll
  for (int d = 0; d < 3; d++) y[d] = fma(g.c2[d], qp.e, fma(g.c1[d], qp.z, g.c0[d])) - S.c[d];
  // densities: a_c = w scale |W| du_c (monopole), b_c = -sigma w scale u_c
  // (dipole b_c W), split as W_z, (W_x - i W_y)/2, (W_x + i W_y)/2
  double n0, n1, n2;
  rule_N(qp, n0, n1, n2);
  const double wt = qp.w * S.scale;
  const double sg = sd < 0 ? -1.0 : 1.0;
  const cd *Us = U + (size_t)s * 16, *dUs = dU + (size_t)s * 16;
  cd *dq = dens ? dens + (size_t)ql * MP_NDENS : nullptr;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    if (!dq) break;
    const cd u = Us[c] * n0 + Us[4 + c] * n1 + Us[8 + c] * n2;
    const cd du = dUs[c] * n0 + dUs[4 + c] * n1 + dUs[8 + c] * n2;
    const cd a = du * (wt * g.aW), b = u * (-sg * wt);
    dq[c] = a;
    dq[4 + c] = b * g.W[2];
    dq[8 + c] = b * mk(0.5 * g.W[0], -0.5 * g.W[1]);
    dq[12 + c] = b * mk(0.5 * g.W[0], 0.5 * g.W[1]);
  }
  const double r = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
  const int Lp = S.Lp;
  sph_j(S.kr * r, Lp, jt + lane, 64);
  const double ir = r > 0.0 ? 1.0 / r : 0.0;
  const double u = r > 0.0 ? y[2] * ir : 1.0;
  const cd w = mk(y[0] * ir, y[1] * ir);
  cd *Bq = ROWQ ? B + (size_t)ql * ldb : B + ql;
  const size_t bs = ROWQ ? 1 : ldb;   // stride between coefficients
  // m-outer, n-inner; P^_m^m = -sqrt((2m+1)/2m) P^_{m-1}^{m-1}
  double pmm = 0.28209479177387814347;   // 1 / sqrt(4 pi)
  cd wm = mk(1.0, 0.0);
  for (int m = 0; m <= Lp; m++) {
    if (m > 0) {
      pmm *= -sqrt((2.0 * m + 1.0) / (2.0 * m));
      wm = wm * w;
    }
    const cd cwm = conjd(wm);
    const double sgn = (m & 1) ? -1.0 : 1.0;
    double p2 = 0.0, p1 = pmm;
    for (int n = m; n <= Lp; n++) {
      double pn;
      if (n == m) pn = pmm;
      else if (n == m + 1) pn = sqrt(2.0 * m + 3.0) * u * pmm;
      else pn = pa(n, m) * (u * p1 - pb(n, m) * p2);
      if (n > m) { p2 = p1; p1 = pn; }
      const double jr = jt[n * 64 + lane] * pn;
      const int i = mp_idx(n, m, Lp);
      Bq[i * bs] = cwm * jr;                          // conj(R_n^m)
      if (m > 0) Bq[(i + 1) * bs] = wm * (sgn * jr);  // conj(R_n^{-m}) = (-1)^m R_n^m
    }
  }
}

// T[split][idx][16] = sum over the split's points of B[q][idx] dens[q][:]
// (one wave per workgroup, lane per idx; the split's densities staged in
// LDS 64 points at a time and read as broadcasts, the next point's basis
// value loaded before the current point's 64 multiply-adds)
#define MP_MOM_CHUNK 64
__global__ void __launch_bounds__(64) k_mp_moments(int nidx, int nqc, int qper, const cd *__restrict__ B, size_t ldb,
                                                   const cd *__restrict__ dens, cd *__restrict__ T)
{
  __shared__ cd sd[MP_MOM_CHUNK * MP_NDENS];
  const int lane = threadIdx.x;
  const int idx = blockIdx.x * 64 + lane;
  const int split = blockIdx.y;
  const int qa = split * qper, qb = min(nqc, qa + qper);
  const bool live = idx < nidx;
  const int ii = live ? idx : 0;
  cd acc[MP_NDENS];
#pragma unroll
  for (int d = 0; d < MP_NDENS; d++) acc[d] = mk(0.0, 0.0);
  for (int q0 = qa; q0 < qb; q0 += MP_MOM_CHUNK) {
    const int nq = min(MP_MOM_CHUNK, qb - q0);
    __syncthreads();   // the previous chunk's densities are consumed
    for (int t = lane; t < nq * MP_NDENS; t += 64) sd[t] = dens[(size_t)q0 * MP_NDENS + t];
    __syncthreads();
    cd b = B[(size_t)q0 * ldb + ii];
    for (int q = 0; q < nq; q++) {
      const cd bn = q + 1 < nq ? B[(size_t)(q0 + q + 1) * ldb + ii] : b;
      const cd *dq = sd + q * MP_NDENS;
#pragma unroll
      for (int d = 0; d < MP_NDENS; d++) acc[d] = acc[d] + b * dq[d];
      b = bn;
    }
  }
  if (live) {
    cd *t = T + ((size_t)split * nidx + idx) * MP_NDENS;
#pragma unroll
    for (int d = 0; d < MP_NDENS; d++) t[d] = acc[d];
  }
}

__global__ void k_zero_cd(cd *__restrict__ a, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = mk(0.0, 0.0);
}

// sum of the splits' partial moments, added to Tsum (chunks of the source
// points accumulate across launches)
__global__ void k_mp_reduce(int n, int nsplit, const cd *__restrict__ T, cd *__restrict__ Tsum)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // eight partial sums (loads of eight splits in flight), added in a fixed order
  cd p[8];
#pragma unroll
  for (int u = 0; u < 8; u++) p[u] = mk(0.0, 0.0);
  int k = 0;
  for (; k + 8 <= nsplit; k += 8)
#pragma unroll
    for (int u = 0; u < 8; u++) p[u] = p[u] + T[(size_t)(k + u) * n + i];
  for (; k < nsplit; k++) p[0] = p[0] + T[(size_t)k * n + i];
  Tsum[i] = Tsum[i] + (((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7])));
}

// moment of density set d (0 monopole, 1 z, 2 +, 3 -) and potential c at (n, m)
// (T rows of ldt values; a beam's 16 moments start at column boff)
__device__ __forceinline__ cd mp_T(const cd *T, int Lp, int n, int m, int d, int c, size_t ldt = MP_NDENS, int boff = 0)
{
  if (n < 0 || n > Lp || m > n || -m > n) return mk(0.0, 0.0);
  return T[(size_t)mp_idx(n, m, Lp) * ldt + boff + d * 4 + c];
}
// series coefficient A_nm of potential c (zero beyond the order L)
__device__ cd mp_A(const cd *T, int L, int Lp, double k, int n, int m, int c, size_t ldt = MP_NDENS, int boff = 0)
{
  if (n < 0 || n > L || m > n || -m > n) return mk(0.0, 0.0);
  cd s = mp_T(T, Lp, n, m, 0, c, ldt, boff);
  s = s + (mp_T(T, Lp, n - 1, m, 1, c, ldt, boff) * cz1(n, m) + mp_T(T, Lp, n + 1, m, 1, c, ldt, boff) * cz2(n, m)) * k;
  s = s + (mp_T(T, Lp, n - 1, m - 1, 2, c, ldt, boff) * cm1(n, m) +
           mp_T(T, Lp, n + 1, m - 1, 2, c, ldt, boff) * cm2(n, m)) * k;
  s = s + (mp_T(T, Lp, n - 1, m + 1, 3, c, ldt, boff) * cp1(n, m) +
           mp_T(T, Lp, n + 1, m + 1, 3, c, ldt, boff) * cp2(n, m)) * k;
  return s * mk(0.0, 4.0 * M_PI * k);
}

// coefficient sets of the evaluation, [idx][set][c]: set 0 A (U), 1 the
// coefficients of d_z U, 2 of d_+ U, 3 of d_- U
// nb beams: T and C rows hold 16 values per beam ([idx][b 16 + ...])
__global__ void k_mp_coef(int L, int Lp, double k, const cd *__restrict__ T, cd *__restrict__ C, int nb = 1)
{
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nidx = mp_count(Lp);
  if (t >= (size_t)nidx * 4 * nb) return;
  const int i = (int)(t / (4 * (size_t)nb)), q = (int)(t % (4 * (size_t)nb)), b = q >> 2, c = q & 3;
  const size_t ldt = (size_t)MP_NDENS * nb;
  const int boff = b * MP_NDENS;
  // (n, m) of index i
  int n = 0, m = 0;
  if (i <= Lp) { n = i; m = 0; }
  else {
    int am = 1;
    while (am < Lp && mp_idx(am + 1, am + 1, Lp) <= i) am++;
    const int r = i - mp_idx(am, am, Lp);
    n = am + r / 2;
    m = (r & 1) ? -am : am;
  }
  cd *o = C + (size_t)i * ldt + boff;
  o[c] = mp_A(T, L, Lp, k, n, m, c, ldt, boff);
  cd dz = mk(0.0, 0.0), dp = mk(0.0, 0.0), dm = mk(0.0, 0.0);
  if (n + 1 <= Lp) {
    if (m <= n + 1 && -m <= n + 1) dz = dz + mp_A(T, L, Lp, k, n + 1, m, c, ldt, boff) * cz1(n + 1, m);
    if (m - 1 <= n + 1 && 1 - m <= n + 1) dp = dp + mp_A(T, L, Lp, k, n + 1, m - 1, c, ldt, boff) * cp1(n + 1, m - 1);
    if (m + 1 <= n + 1 && -m - 1 <= n + 1) dm = dm + mp_A(T, L, Lp, k, n + 1, m + 1, c, ldt, boff) * cm1(n + 1, m + 1);
  }
  if (n - 1 >= 0) {
    if (m <= n - 1 && -m <= n - 1) dz = dz + mp_A(T, L, Lp, k, n - 1, m, c, ldt, boff) * cz2(n - 1, m);
    if (m - 1 <= n - 1 && 1 - m <= n - 1) dp = dp + mp_A(T, L, Lp, k, n - 1, m - 1, c, ldt, boff) * cp2(n - 1, m - 1);
    if (m + 1 <= n - 1 && -m - 1 <= n - 1) dm = dm + mp_A(T, L, Lp, k, n - 1, m + 1, c, ldt, boff) * cm2(n - 1, m + 1);
  }
  o[4 + c] = dz * k;
  o[8 + c] = dp * k;
  o[12 + c] = dm * k;
}


// ------------------------------------------------------------------ many beams
// the 16 densities of every beam at a source point: dens[q][b 16 + d]
// (beam b's boundary values at U + b ustride)
template <class R>
__global__ void __launch_bounds__(64) k_mp_dens_multi(MpSrc S, const TriGeo *__restrict__ geo, const int *__restrict__ sid,
                                                      const R *__restrict__ rule, const cd *__restrict__ U,
                                                      const cd *__restrict__ dU, size_t ustride, int nb, int q0, int nqc,
                                                      cd *__restrict__ dens)
{
  const int ql = blockIdx.x * 64 + threadIdx.x;
  if (ql >= nqc) return;
  const int q = q0 + ql;
  const int s = 1 + q / S.npt, p = q % S.npt;
  const int sd = sid[s], as = sd < 0 ? -sd : sd;
  const TriGeo g = geo[as];
  const R qp = rule[S.qoff + p];
  double n0, n1, n2;
  rule_N(qp, n0, n1, n2);
  const double wt = qp.w * S.scale, sg = sd < 0 ? -1.0 : 1.0;
  const cd wzp = mk(0.5 * g.W[0], -0.5 * g.W[1]), wzm = mk(0.5 * g.W[0], 0.5 * g.W[1]);
  cd *dq = dens + (size_t)ql * MP_NDENS * nb;
  for (int b = 0; b < nb; b++) {
    const cd *Us = U + (size_t)b * ustride + (size_t)s * 16, *dUs = dU + (size_t)b * ustride + (size_t)s * 16;
    for (int c = 0; c < 4; c++) {
      const cd u = Us[c] * n0 + Us[4 + c] * n1 + Us[8 + c] * n2;
      const cd du = dUs[c] * n0 + dUs[4 + c] * n1 + dUs[8 + c] * n2;
      const cd bb = u * (-sg * wt);
      dq[b * MP_NDENS + c] = du * (wt * g.aW);
      dq[b * MP_NDENS + 4 + c] = bb * g.W[2];
      dq[b * MP_NDENS + 8 + c] = bb * wzp;
      dq[b * MP_NDENS + 12 + c] = bb * wzm;
    }
  }
}

// S_n^m of every index at points x_p - c: out[p][idx] (lane per point)
__global__ void __launch_bounds__(64) k_mp_sbasis(const double *__restrict__ pts, int np, double c0, double c1, double c2,
                                                  double kr, int Lp, cd *__restrict__ out)
{
  __shared__ cd ht[(MP_LPMAX + 1) * 64];
  const int lane = threadIdx.x, p = blockIdx.x * 64 + lane;
  if (p >= np) return;
  const double y[3] = {pts[3 * p] - c0, pts[3 * p + 1] - c1, pts[3 * p + 2] - c2};
  const double r = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]), ir = 1.0 / r;
  sph_h(kr * r, Lp, ht + lane, 64);
  const double u = y[2] * ir;
  const cd w = mk(y[0] * ir, y[1] * ir);
  cd *o = out + (size_t)p * mp_count(Lp);
  double pmm = 0.28209479177387814347;
  cd wm = mk(1.0, 0.0);
  for (int m = 0; m <= Lp; m++) {
    if (m > 0) {
      pmm *= -sqrt((2.0 * m + 1.0) / (2.0 * m));
      wm = wm * w;
    }
    const cd cwm = conjd(wm);
    const double sgn = (m & 1) ? -1.0 : 1.0;
    double p2 = 0.0, p1 = pmm;
    for (int n = m; n <= Lp; n++) {
      double pn;
      if (n == m) pn = pmm;
      else if (n == m + 1) pn = sqrt(2.0 * m + 3.0) * u * pmm;
      else pn = pa(n, m) * (u * p1 - pb(n, m) * p2);
      if (n > m) { p2 = p1; p1 = pn; }
      const cd hp = ht[n * 64 + lane] * pn;
      const int i = mp_idx(n, m, Lp);
      o[i] = hp * wm;
      if (m > 0) o[i + 1] = hp * cwm * sgn;
    }
  }
}

// the series products C[p][b 16 + set 4 + c] of destination slots s0 + p
// into the incident data of every beam: U += set 0, dU/dv += v_z set 1 +
// (v_x - i v_y)/2 set 2 + (v_x + i v_y)/2 set 3
__global__ void k_mp_combine(const cd *__restrict__ C, int ns, int nb, const double *__restrict__ nrm,
                             const int *__restrict__ slot, int s0, cd *__restrict__ Ui, cd *__restrict__ dUi,
                             size_t istride)
{
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)ns * nb) return;
  const int p = (int)(t / nb), b = (int)(t % nb), sl = s0 + p;
  const double *v = nrm + 3 * (size_t)sl;
  const cd vm = mk(0.5 * v[0], -0.5 * v[1]), vp = mk(0.5 * v[0], 0.5 * v[1]);
  const cd *c = C + ((size_t)p * nb + b) * MP_NDENS;
  const size_t o = (size_t)b * istride + (size_t)slot[sl] * 4;
  for (int k = 0; k < 4; k++) {
    Ui[o + k] = Ui[o + k] + c[k];
    dUi[o + k] = dUi[o + k] + c[4 + k] * v[2] + c[8 + k] * vm + c[12 + k] * vp;
  }
}

// E, H of every beam at points p0 + p from one source object's series
// products C[p][b 16 + set 4 + c] (sets U, d_z U, d_+ U, d_- U):
// E_i = -d_i U_3 + U_i, H = ch curl (U_0, U_1, U_2) (EH_mEMP_s,
// d3b1_field.c:131-136), added into EH[b][p][6] (beam b at EH + b ldb)
__global__ void k_mp_far_eh(const cd *__restrict__ C, int np, int nb, double chr, double chi, int p0, size_t ldb,
                            cd *__restrict__ EH)
{
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)np * nb) return;
  const int p = (int)(t / nb), b = (int)(t % nb);
  const cd *c = C + ((size_t)p * nb + b) * MP_NDENS;
  cd dx[4], dy[4], dz[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const cd sp = c[8 + q], sm = c[12 + q], df = sp - sm;
    dx[q] = (sp + sm) * 0.5;
    dy[q] = mk(0.5 * df.y, -0.5 * df.x);   // (d_+ - d_-) / 2i
    dz[q] = c[4 + q];
  }
  const cd ch = mk(chr, chi);
  cd *o = EH + (size_t)b * ldb + (size_t)(p0 + p) * 6;
  o[0] = o[0] + c[0] - dx[3];
  o[1] = o[1] + c[1] - dy[3];
  o[2] = o[2] + c[2] - dz[3];
  o[3] = o[3] + ch * (dy[2] - dz[1]);
  o[4] = o[4] + ch * (dz[0] - dx[2]);
  o[5] = o[5] + ch * (dx[1] - dy[0]);
}

// |E|^2 and |H|^2 of n (beam, point) rows of 6 (example4.c's ie, ih)
__global__ void k_eh_intensity(const cd *__restrict__ EH, size_t n, double *__restrict__ I)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cd *e = EH + 6 * i;
  double a = 0.0, h = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    a += e[k].x * e[k].x + e[k].y * e[k].y;
    h += e[3 + k].x * e[3 + k].x + e[3 + k].y * e[3 + k].y;
  }
  I[2 * i] = a;
  I[2 * i + 1] = h;
}

// ------------------------------------------------------------------ evaluation
// one (destination object, source object) pair of the series
struct MpPair {
  const cd *coef;     // the source's coefficient sets
  double c[3];        // its centre
  double kr;
  int Lp;             // the source's layout order
  int Le;             // evaluation order of this pair (<= Lp)
  int pad[2];
};
// one tile: 64 destination slots of one object and its list of source pairs
struct MpTile {
  const double *pts, *nrm;   // the object's slot positions and normals
  const int *slot;           // slot -> incident-data index
  cd *Ui, *dUi;
  int s0, ns;                // first slot, slots in this tile
  int p0, np;                // its pairs in the pair table
};

__global__ void __launch_bounds__(64) k_mp_eval(const MpTile *__restrict__ tiles, const MpPair *__restrict__ pairs)
{
  __shared__ cd ht[(MP_LPMAX + 1) * 64];
  const MpTile T = tiles[blockIdx.x];
  const int lane = threadIdx.x;
  const bool act = lane < T.ns;
  const int sl = T.s0 + (act ? lane : 0);
  const double x[3] = {T.pts[3 * sl], T.pts[3 * sl + 1], T.pts[3 * sl + 2]};
  const double v[3] = {T.nrm[3 * sl], T.nrm[3 * sl + 1], T.nrm[3 * sl + 2]};
  cd aU[4], aZ[4], aP[4], aM[4];
#pragma unroll
  for (int c = 0; c < 4; c++) aU[c] = aZ[c] = aP[c] = aM[c] = mk(0.0, 0.0);
  for (int pi = 0; pi < T.np; pi++) {
    const MpPair P = pairs[T.p0 + pi];
    const double y[3] = {x[0] - P.c[0], x[1] - P.c[1], x[2] - P.c[2]};
    const double r = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]), ir = 1.0 / r;
    const int Le = P.Le, Lp = P.Lp;
    sph_h(P.kr * r, Le, ht + lane, 64);
    const double u = y[2] * ir;
    const cd w = mk(y[0] * ir, y[1] * ir);
    double pmm = 0.28209479177387814347;
    cd wm = mk(1.0, 0.0);
    for (int m = 0; m <= Le; m++) {
      if (m > 0) {
        pmm *= -sqrt((2.0 * m + 1.0) / (2.0 * m));
        wm = wm * w;
      }
      const cd cwm = conjd(wm);
      const double sgn = (m & 1) ? -1.0 : 1.0;
      double p2 = 0.0, p1 = pmm;
      for (int n = m; n <= Le; n++) {
        double pn;
        if (n == m) pn = pmm;
        else if (n == m + 1) pn = sqrt(2.0 * m + 3.0) * u * pmm;
        else pn = pa(n, m) * (u * p1 - pb(n, m) * p2);
        if (n > m) { p2 = p1; p1 = pn; }
        const cd hp = ht[n * 64 + lane] * pn;
        const cd Sp = hp * wm;                   // S_n^m
        const cd *C = P.coef + (size_t)mp_idx(n, m, Lp) * 16;
#pragma unroll
        for (int c = 0; c < 4; c++) {
          aU[c] = aU[c] + C[c] * Sp;
          aZ[c] = aZ[c] + C[4 + c] * Sp;
          aP[c] = aP[c] + C[8 + c] * Sp;
          aM[c] = aM[c] + C[12 + c] * Sp;
        }
        if (m > 0) {
          const cd Sn = hp * cwm * sgn;          // S_n^{-m} = h_n (-1)^m conj(Y_n^m)
          const cd *Cn = C + 16;
#pragma unroll
          for (int c = 0; c < 4; c++) {
            aU[c] = aU[c] + Cn[c] * Sn;
            aZ[c] = aZ[c] + Cn[4 + c] * Sn;
            aP[c] = aP[c] + Cn[8 + c] * Sn;
            aM[c] = aM[c] + Cn[12 + c] * Sn;
          }
        }
      }
    }
  }
  if (!act) return;
  // d/dv = v_z d_z + (v_x - i v_y)/2 d_+ + (v_x + i v_y)/2 d_-
  const cd vm = mk(0.5 * v[0], -0.5 * v[1]), vp = mk(0.5 * v[0], 0.5 * v[1]);
  const size_t o = (size_t)T.slot[sl] * 4;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const cd du = aZ[c] * v[2] + aP[c] * vm + aM[c] * vp;
    T.Ui[o + c] = T.Ui[o + c] + aU[c];
    T.dUi[o + c] = T.dUi[o + c] + du;
  }
}

// ------------------------------------------------------------------ loop kernels
__global__ void k_reset_incident(cd *__restrict__ Ui, cd *__restrict__ dUi, const cd *__restrict__ Ui0,
                                 const cd *__restrict__ dUi0, cd *__restrict__ Uip, size_t n)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Uip[i] = Ui[i];
  Ui[i] = Ui0[i];
  dUi[i] = dUi0[i];
}

// per-slot relative change terms of ccd_f (mo_solve.c:413-455), written at
// the object's offset of the loop's slot table
__global__ void k_ccd(const cd *__restrict__ Ui, const cd *__restrict__ Uip, const int *__restrict__ slot, int np,
                      double *__restrict__ part)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const cd *c = Ui + (size_t)slot[p] * 4, *q = Uip + (size_t)slot[p] * 4;
  double ac = 0.0, ap = 0.0, dc = 0.0;
  for (int i = 0; i < 3; i++) { ac += abs2(c[i]); ap += abs2(q[i]); dc += abs2(c[i] - q[i]); }
  part[2 * p] = 2.0 * dc / (ac + ap);
  part[2 * p + 1] = 2.0 * abs2(c[3] - q[3]) / (abs2(c[3]) + abs2(q[3]));
}

// the two sums of the change terms over n slots, in a fixed order (one
// workgroup; out[0], out[1])
__global__ void __launch_bounds__(256) k_ccd_sum(const double *__restrict__ part, int n, double *__restrict__ out)
{
  __shared__ double su[256], sp[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) { a += part[2 * i]; b += part[2 * i + 1]; }
  su[threadIdx.x] = a;
  sp[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) { su[threadIdx.x] += su[threadIdx.x + s]; sp[threadIdx.x] += sp[threadIdx.x + s]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = su[0]; out[1] = sp[0]; }
}

// ------------------------------------------------------------------ host side
namespace {

int g_coupling_mode = 0;   // 0: series where the pair allows it (mixed form), 1: direct sums only
int g_last_series = 0, g_last_direct = 0, g_last_split = 0, g_last_lmax = -1;   // the last loop's plan

// grow-only work areas of the loop (bem_mo_release frees them)
Scratch *g_ws = nullptr;
enum { W_DID, W_PART, W_CCD, W_CCDSUM, W_B, W_DENS, W_T, W_TSUM, W_COEF, W_TILES, W_PAIRS, W_MS, W_MC, W_FPTS,
       W_FEH, W_FI, W_SPT, W_SNM, W_SIDX, W_SLST, W_NWS };
template <class T> T *ws(int i, size_t n)
{
  if (!g_ws) g_ws = new Scratch[W_NWS];
  return g_ws[i].get<T>(std::max<size_t>(n, 1));
}

// geometry of one object as a source of the series: centre, radius of the
// bounding sphere of its open-domain faces, largest element radius
struct SrcGeo {
  double c[3], rad, erad;
  bool ok;       // triangles only
};

SrcGeo src_geo(const bh_object *o)
{
  SrcGeo s{};
  s.ok = o->nquad == 0;
  const Dom &D = o->dom[0];
  double c[3] = {0, 0, 0};
  int nv = 0;
  for (int t = 1; t <= D.dNe; t++) {
    const int a = std::abs(D.sid[t]);
    for (int j = 0; j < 3; j++) {
      const int v = o->ed[4 * a + j];
      for (int d = 0; d < 3; d++) c[d] += o->rn[3 * (size_t)v + d];
      nv++;
    }
  }
  for (int d = 0; d < 3; d++) s.c[d] = nv ? c[d] / nv : 0.0;
  double r2 = 0.0, e2 = 0.0;
  for (int t = 1; t <= D.dNe; t++) {
    const int a = std::abs(D.sid[t]);
    double cen[3] = {0, 0, 0};
    for (int j = 0; j < 3; j++)
      for (int d = 0; d < 3; d++) cen[d] += o->rn[3 * (size_t)o->ed[4 * a + j] + d] / 3.0;
    for (int j = 0; j < 3; j++) {
      const double *p = &o->rn[3 * (size_t)o->ed[4 * a + j]];
      double q = 0.0, w = 0.0;
      for (int d = 0; d < 3; d++) { q += (p[d] - s.c[d]) * (p[d] - s.c[d]); w += (p[d] - cen[d]) * (p[d] - cen[d]); }
      r2 = std::max(r2, q);
      e2 = std::max(e2, w);
    }
  }
  s.rad = sqrt(r2);
  s.erad = sqrt(e2);
  return s;
}

// open-domain node slots of an object on the host, in the order of its
// device slot table (4 per element, centroid included)
void slot_points(const bh_object *o, std::vector<double> &pt)
{
  pt.clear();
  for (int i = 1; i <= o->Ne; i++) {
    if (o->md[i] != 0) continue;
    for (int j = 0; j < 4; j++)
      for (int q = 0; q < 3; q++) pt.push_back(o->ren[(size_t)i * 12 + j * 3 + q]);
  }
}

// the plan of one coupling loop: which pairs take the series, with what
// order, and the evaluation tiles
struct Plan {
  std::vector<SrcGeo> geo;
  std::vector<std::vector<int>> Lmn;   // [dst][src]: series order, -1 direct, -2 same object
  // split pairs (Lmn -1, the one-beam loop only): [dst][src] series order of
  // the 64-slot tiles whose slots all qualify on their own (-1: not split),
  // and per tile 1 for the series, 0 for the element sums
  std::vector<std::vector<int>> Ls;
  std::vector<std::vector<std::vector<char>>> tile_mp;
  std::vector<int> Lsrc;               // per source: largest order over its pairs (-1: none)
  std::vector<int> rule;               // per source: rule-table offset of its far rule
  std::vector<int> npt;                // and points per element
  int npairs_mp = 0, npairs_direct = 0, npairs_split = 0;
};

// split: a pair whose slots do not all qualify takes the series on the
// 64-slot tiles whose slots do (the one-beam loop; the many-beam loop sums
// such pairs by elements)
int make_plan(bh_object **objs, int N, Plan &P, bool split)
{
  P.geo.resize(N);
  for (int m = 0; m < N; m++) P.geo[m] = src_geo(objs[m]);
  P.Lmn.assign(N, std::vector<int>(N, -1));
  P.Ls.assign(N, std::vector<int>(N, -1));
  P.tile_mp.assign(N, std::vector<std::vector<char>>(N));
  P.Lsrc.assign(N, -1);
  P.rule.assign(N, 0);
  P.npt.assign(N, 0);
  std::vector<double> mingap(N, 1e300);
  std::vector<double> pt;
  for (int m = 0; m < N; m++) {
    slot_points(objs[m], pt);
    const size_t ns = pt.size() / 3;
    for (int n = 0; n < N; n++) {
      if (n == m) { P.Lmn[m][n] = -2; continue; }
      const bh_object *s = objs[n];
      const SrcGeo &G = P.geo[n];
      const zd k = s->kn[0];
      const bool eligible = g_coupling_mode == 0 && s->prec != BH_PREC_FP64 && objs[m]->prec != BH_PREC_FP64 &&
                            G.ok && k.imag() == 0.0 && k.real() > 0.0 && G.rad > 0.0;
      if (!eligible) continue;
      // the smallest distance of each 64-slot tile from the source's centre
      const size_t nt = (ns + 63) / 64;
      std::vector<double> td(nt, 1e300);
      for (size_t p = 0; p < ns; p++) {
        double q = 0.0;
        for (int d = 0; d < 3; d++) q += (pt[3 * p + d] - G.c[d]) * (pt[3 * p + d] - G.c[d]);
        td[p / 64] = std::min(td[p / 64], q);
      }
      const double kr = k.real() * G.rad;
      auto order = [&](double rho) {
        return std::max((int)ceil(log(MP_EPS) / log(rho)), (int)ceil(kr + 6.0 * cbrt(kr)) + 4);
      };
      double dmin = 1e300;
      for (double q : td) dmin = std::min(dmin, q);
      dmin = sqrt(dmin);
      const double rho = G.rad / dmin, gap = dmin - G.rad;
      if (rho <= MP_RHO && gap >= MP_GAP_RADII * G.erad) {
        const int L = order(rho);
        if (L > MP_LMAX) continue;
        P.Lmn[m][n] = L;
        P.Lsrc[n] = std::max(P.Lsrc[n], L);
        mingap[n] = std::min(mingap[n], gap);
        continue;
      }
      if (!split) continue;
      // the tiles that qualify on their own
      std::vector<char> fl(nt, 0);
      double rmax = 0.0, gmin = 1e300;
      int nq = 0;
      for (size_t t = 0; t < nt; t++) {
        const double dt = sqrt(td[t]), rt = G.rad / dt, gt = dt - G.rad;
        if (rt <= MP_RHO_SPLIT && gt >= MP_GAP_RADII * G.erad) {
          fl[t] = 1;
          nq++;
          rmax = std::max(rmax, rt);
          gmin = std::min(gmin, gt);
        }
      }
      if (nq == 0) continue;
      const int L = order(rmax);
      if (L > MP_LMAX) continue;
      P.Ls[m][n] = L;
      P.tile_mp[m][n].swap(fl);
      P.Lsrc[n] = std::max(P.Lsrc[n], L);
      mingap[n] = std::min(mingap[n], gmin);
    }
  }
  // far rule of each source (the mixed form's rules, keyed as far_rule keys
  // them, by the closest slot of any of its series pairs)
  for (int n = 0; n < N; n++) {
    if (P.Lsrc[n] < 0) continue;
    const double kR = std::abs(objs[n]->kn[0]) * P.geo[n].erad;
    const double d = mingap[n] / std::max(P.geo[n].erad, 1e-300);
    if (d >= 8.0 && kR <= 0.5) { P.rule[n] = BEM_FAR_OFF4; P.npt[n] = 48; }
    else if (kR <= 1.2) { P.rule[n] = BEM_FAR_OFF5; P.npt[n] = 75; }
    else if (kR <= 3.0) { P.rule[n] = BEM_FAR_OFF6; P.npt[n] = 108; }
    else { P.rule[n] = 0; P.npt[n] = 432; }
  }
  for (int m = 0; m < N; m++)
    for (int n = 0; n < N; n++)
      if (m != n) (P.Lmn[m][n] >= 0 ? P.npairs_mp : P.Ls[m][n] >= 0 ? P.npairs_split : P.npairs_direct)++;
  return 0;
}

// the series coefficients of source object n from its current boundary values
int mp_coefficients(bh_object *o, const Plan &P, int n, cd *coef)
{
  const Dom &D = o->dom[0];
  MpSrc S{};
  for (int d = 0; d < 3; d++) S.c[d] = P.geo[n].c[d];
  S.kr = o->kn[0].real();
  S.scale = g_rscale[2];
  S.L = P.Lsrc[n];
  S.Lp = S.L + 1;
  S.qoff = P.rule[n];
  S.npt = P.npt[n];
  S.dNe = D.dNe;
  S.nq = D.dNe * S.npt;
  const int nidx = mp_count(S.Lp);
  // the source points in chunks whose basis fits the work area (~1.5 GB)
  const size_t ldb = (size_t)nidx;
  const int qc = (int)std::max<size_t>(MP_QSPLIT, std::min<size_t>((size_t)S.nq, ((size_t)96 << 20) / ldb / MP_QSPLIT * MP_QSPLIT));
  const int nsplit = (qc + MP_QSPLIT - 1) / MP_QSPLIT;
  cd *B = ws<cd>(W_B, (size_t)qc * ldb);
  cd *dens = ws<cd>(W_DENS, (size_t)qc * MP_NDENS);
  cd *T = ws<cd>(W_T, (size_t)nsplit * nidx * MP_NDENS);
  cd *Ts = ws<cd>(W_TSUM, (size_t)nidx * MP_NDENS);
  if (!B || !dens || !T || !Ts) return fail(BH_EHIP, "series: work area (%d coefficients)", nidx);
  // the sum is cleared by a kernel (no memset node if a caller captures this)
  hipLaunchKernelGGL(k_zero_cd, g1((size_t)nidx * MP_NDENS), dim3(256), 0, g.st, Ts, (size_t)nidx * MP_NDENS);
  for (int q0 = 0; q0 < S.nq; q0 += qc) {
    const int nqc = std::min(qc, S.nq - q0);
    bem_prof_begin(BH_K_MP_BASIS, (double)nqc * nidx, g.st);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mp_basis<true, QPt>), dim3((nqc + 63) / 64), dim3(64), 0, g.st, S, o->d_geo, D.d_sid, g.d_glr, D.d_U,
                       D.d_dU, q0, nqc, B, ldb, dens);
    bem_prof_end(BH_K_MP_BASIS, g.st);
    const int ns = (nqc + MP_QSPLIT - 1) / MP_QSPLIT;
    bem_prof_begin(BH_K_MP_MOM, 8.0 * nidx * MP_NDENS * (double)nqc, g.st);
    hipLaunchKernelGGL(k_mp_moments, dim3((nidx + 63) / 64, ns), dim3(64), 0, g.st, nidx, nqc, MP_QSPLIT, B, ldb,
                       dens, T);
    bem_prof_end(BH_K_MP_MOM, g.st);
    const size_t nt = (size_t)nidx * MP_NDENS;
    hipLaunchKernelGGL(k_mp_reduce, g1(nt), dim3(256), 0, g.st, (int)nt, ns, T, Ts);
  }
  hipLaunchKernelGGL(k_mp_coef, g1((size_t)nidx * 4), dim3(256), 0, g.st, S.L, S.Lp, S.kr, Ts, coef, 1);
  HIPCHK(hipGetLastError());
  return 0;
}

// the series coefficients of one source object (centre, order and rule in
// S; its rule table) for all nb beams: coef[idx][b 16 + set 4 + c].  The
// basis of the source points is formed once and shared by every beam; the
// moments of all beams are one product (idx x points) x (points x 16 nb)
// per chunk of points.
template <class R> int series_multi(bh_object *o, const MpSrc &S, const R *rule, cd *coef, int nb)
{
  const Dom &D = o->dom[0];
  const int nidx = mp_count(S.Lp);
  const size_t ncol = (size_t)MP_NDENS * nb;
  // points per chunk: the basis (idx x points) and the densities (points x
  // 16 nb) each within about 1 GB
  const size_t cap = std::min<size_t>(((size_t)64 << 20) / nidx, ((size_t)64 << 20) / ncol);
  const int qc = (int)std::max<size_t>(64, std::min<size_t>((size_t)S.nq, cap));
  cd *B = ws<cd>(W_B, (size_t)qc * nidx);
  cd *dens = ws<cd>(W_DENS, (size_t)qc * ncol);
  cd *T = ws<cd>(W_T, (size_t)nidx * ncol);
  if (!B || !dens || !T) return fail(BH_EHIP, "series: work area (%d coefficients, %d beams)", nidx, nb);
  const size_t us = (size_t)(D.dNe + 1) * 16;
  for (int q0 = 0; q0 < S.nq; q0 += qc) {
    const int nqc = std::min(qc, S.nq - q0);
    bem_prof_begin(BH_K_MP_BASIS, (double)nqc * nidx, g.st);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mp_basis<false, R>), dim3((nqc + 63) / 64), dim3(64), 0, g.st, S, o->d_geo,
                       D.d_sid, rule, (const cd *)nullptr, (const cd *)nullptr, q0, nqc, B, (size_t)qc, (cd *)nullptr);
    bem_prof_end(BH_K_MP_BASIS, g.st);
    hipLaunchKernelGGL(k_mp_dens_multi<R>, dim3((nqc + 63) / 64), dim3(64), 0, g.st, S, o->d_geo, D.d_sid, rule,
                       (const cd *)mb_dom_ptr(o, 0, 0), (const cd *)mb_dom_ptr(o, 0, 1), us, nb, q0, nqc, dens);
    bem_zgemm(nidx, (int)ncol, nqc, 1.0, 0.0, B, (size_t)qc, dens, ncol, q0 ? 1.0 : 0.0, 0.0, T, ncol, g.st);
  }
  hipLaunchKernelGGL(k_mp_coef, g1((size_t)nidx * 4 * nb), dim3(256), 0, g.st, S.L, S.Lp, S.kr, T, coef, nb);
  HIPCHK(hipGetLastError());
  return 0;
}

// the series coefficients of source object n for all nb beams of the
// many-beam loop (its far rule)
int mp_coefficients_multi(bh_object *o, const Plan &P, int n, cd *coef, int nb)
{
  MpSrc S{};
  for (int d = 0; d < 3; d++) S.c[d] = P.geo[n].c[d];
  S.kr = o->kn[0].real();
  S.scale = g_rscale[2];
  S.L = P.Lsrc[n];
  S.Lp = S.L + 1;
  S.qoff = P.rule[n];
  S.npt = P.npt[n];
  S.dNe = o->dom[0].dNe;
  S.nq = S.dNe * S.npt;
  return series_multi(o, S, (const QPt *)g.d_glr, coef, nb);
}

// the scattered field of src at dst's node slots by direct sums, added into
// dst's incident data; check: first verify that no slot lies inside src
// (the geometry does not change during a solve: checked on the first sweep)
// n slots of dst given by their positions, normals and incident-data
// indices (device arrays; every slot of dst, or a compact subset)
int add_scattered_at(bh_object *dst, bh_object *src, const double *pts, const double *nrm, const int *sidx, int n,
                     int *d_did, bool check, double *d_part)
{
  if (n <= 0) return 0;
  if (check) {
    if (domain_ids(src, pts, n, d_did)) return BH_EHIP;
    std::vector<int> did(n);
    HIPCHK(hipMemcpyAsync(did.data(), d_did, did.size() * sizeof(int), hipMemcpyDeviceToHost, g.st));
    HIPCHK(hipStreamSynchronize(g.st));
    for (int v : did)
      if (v != 0) return fail(BH_EDOM, "scattered-field domain error (overlapping objects), domain id %d", v);
  }
  Dom &S = src->dom[0];
  bem_launch_scatter(pts, nrm, sidx, n, src->d_geo, src->d_gq, S.d_sid, S.dNe,
                     src->kn[0].real(), src->kn[0].imag(), g_rule[2], g_nrule[2], g_rule[3], g_nrule[3], g_rscale[2],
                     g_qrule[2], g_nqrule[2], g_qrule[3], g_nqrule[3], src->prec != BH_PREC_FP64 ? g.d_gl32 : nullptr,
                     S.d_U, S.d_dU, dst->d_Ui, dst->d_dUi, src->nquad > 0, d_part, g.st);
  HIPCHK(hipGetLastError());
  return 0;
}

int add_scattered(bh_object *dst, bh_object *src, int *d_did, bool check, double *d_part)
{
  return add_scattered_at(dst, src, dst->d_slotPt, dst->d_slotN, dst->d_slotIdx, dst->nslot, d_did, check, d_part);
}


// ------------------------------------------------------------------ lanes
// Objects are independent outside the coupling sums: their assembly,
// factorisation, Schur complement and solves run on several streams at
// once, one host worker thread per stream (a worker's entry-point work goes
// to its stream, set_cur_st).  Latency-bound pieces (LU panels, small
// products, triangular-solve chains) of different objects then overlap.
// Streams keep the default priority.  The tuning table's lanes entry (or
// bh_mo_set_lanes) sets the number of streams; 1 runs everything in the
// calling thread on the library's stream.
#define MO_LMAX 8
hipStream_t g_lane_st[MO_LMAX];
int g_nlanes = 0;     // streams created so far
int g_lanes_req = 0;  // bh_mo_set_lanes (0: the tuning table's lanes entry)
hipStream_t g_gpre_st = nullptr;   // the coupling loop's G1 factorisations (GPrefactor)

int lanes()
{
  int n = g_lanes_req;
  if (n <= 0) n = std::max(1, std::min(MO_LMAX, tune_value(BH_TUNE_LANES)));
  while (g_nlanes < n) {
    HIPCHK(hipStreamCreate(&g_lane_st[g_nlanes]));
    g_nlanes++;
  }
  return n;
}

// fn(o) for every object, the costliest first, spread over the lanes; the
// first failure (its message) is returned
template <class F> int run_lanes(bh_object **objs, int N, const std::vector<double> &cost, F fn)
{
  HIPCHK(hipStreamSynchronize(g.st));   // what the library's stream enqueued for these objects
  const int nl = lanes();
  if (nl <= 0) return BH_EHIP;
  std::vector<int> order(N);
  for (int i = 0; i < N; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  if (nl == 1 || N == 1) {
    for (int i : order) { int r = fn(objs[i]); if (r) return r; }
    return 0;
  }
  std::atomic<int> next(0);
  std::atomic<bool> stop(false);
  std::vector<int> rc(nl, 0);
  std::vector<std::string> msg(nl);
  std::vector<std::thread> th;
  for (int k = 0; k < std::min(nl, N); k++)
    th.emplace_back([&, k]() {
      set_cur_st(g_lane_st[k]);
      for (;;) {
        if (stop.load()) break;
        const int i = next.fetch_add(1);
        if (i >= N) break;
        const int r = fn(objs[order[i]]);
        if (r) { rc[k] = r; msg[k] = bh_last_error(); stop.store(true); break; }
      }
      if (hipStreamSynchronize(g_lane_st[k]) != hipSuccess && !rc[k]) { rc[k] = BH_EHIP; msg[k] = "lane stream"; }
      set_cur_st(nullptr);
    });
  for (std::thread &t : th) t.join();
  for (int k = 0; k < nl; k++)
    if (rc[k]) return fail(rc[k], "%s", msg[k].c_str());
  return 0;
}

std::vector<double> cost_of(bh_object **objs, int N, int power)
{
  std::vector<double> c(N);
  for (int i = 0; i < N; i++) c[i] = pow((double)objs[i]->NN, power);
  return c;
}

}  // namespace

// the work areas; the streams stay for the life of the process (the
// library's shutdown keeps every stream it created)
extern "C" void bem_mo_release(void)
{
  if (g_ws) { delete[] g_ws; g_ws = nullptr; }
}

// assembly and factorisation of N objects (create_matrix of each,
// solve_bieq2.c:10-96): the assembly of every object and domain as one
// batch (one launch per assembly pass), then the factorisations (with
// N > 1 also the Schur complements of the coupling loop) spread over the
// lanes; type and mode as bh_object_assemble and bh_object_factor.
// BEM3_HIP_ASM_BATCH=0: each object assembled on its lane before its
// factorisation (assembly and factorisations of different objects overlap).
extern "C" int bh_mo_prepare(bh_object **objs, int N, int type, int mode)
{
  API_LOCK;
  if (N <= 0) return fail(BH_EARG, "no objects");
  if (ensure_rules()) return BH_EHIP;
  // the mixed form's FP32 LU of G1 (the boundary fields' dE/dH solves,
  // once per assembly) in the same pass unless BEM3_HIP_G1_AHEAD=0
  const char *e = getenv("BEM3_HIP_G1_AHEAD");
  const bool g1 = !(e && strcmp(e, "0") == 0);
  const bool batch = tune_value(BH_TUNE_ASM_BATCH) != 0;
  if (batch) {
    int rc = objects_assemble(objs, N, type);
    if (rc) return rc;
  }
  // with several objects the coupling loop follows: each object's Schur
  // complement S and its LU are formed in the same pass (form_s, which the
  // loop then finds done), so that one pass over the lanes, costliest
  // object first, covers every per-object factorisation
  return run_lanes(objs, N, cost_of(objs, N, 3), [&](bh_object *o) {
    int r = batch ? 0 : object_assemble(o, type);
    if (!r) r = object_factor(o, mode);
    if (!r && g1) r = object_factor_g(o);
    if (!r && N > 1) r = form_s(o);
    return r;
  });
}

// 0: series where a pair allows it (the mixed-precision form), 1: the
// direct element sums for every pair
// device streams of the per-object work (1..8; 0: the tuning table's lanes
// entry).  One lane runs every kernel in order on the library's stream, so
// that per-kernel timings bracket one kernel each.
extern "C" int bh_mo_set_lanes(int n)
{
  API_LOCK;
  if (n < 0 || n > MO_LMAX) return fail(BH_EARG, "lanes %d (0..%d)", n, MO_LMAX);
  g_lanes_req = n;
  return 0;
}

extern "C" int bh_mo_set_coupling(int mode)
{
  API_LOCK;
  if (mode != 0 && mode != 1) return fail(BH_EARG, "coupling mode %d", mode);
  g_coupling_mode = mode;
  return 0;
}

// ordered object pairs of the last coupling loop summed by series and by
// element sums, and the largest series order (-1: none)
extern "C" int bh_mo_plan_stats(int *series, int *direct, int *lmax)
{
  API_LOCK;
  if (series) *series = g_last_series;
  if (direct) *direct = g_last_direct;
  if (lmax) *lmax = g_last_lmax;
  return 0;
}

// object pairs of the last coupling loop whose series takes only some
// tiles of the destination's slots (the rest by element sums)
extern "C" int bh_mo_plan_split(int *split)
{
  API_LOCK;
  if (split) *split = g_last_split;
  return 0;
}

// block-Jacobi fixed point over objects (mo_solve_bv, mo_solve.c:10-88).
// trace (optional, iter_max+1 doubles) receives cc per iteration; *iters
// receives the final iteration count.
// the coupling sums of one loop: its plan, series coefficients, evaluation
// tiles and work areas (set up once per loop; the geometry is fixed)
// compact copies of listed slots: positions, normals, incident-data indices
__global__ void k_gather_slots(const double *__restrict__ pts, const double *__restrict__ nrm,
                               const int *__restrict__ sidx, const int *__restrict__ list, int n, double *__restrict__ opt,
                               double *__restrict__ onm, int *__restrict__ oidx)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = list[i];
#pragma unroll
  for (int d = 0; d < 3; d++) { opt[3 * i + d] = pts[3 * p + d]; onm[3 * i + d] = nrm[3 * p + d]; }
  oidx[i] = sidx[p];
}

struct Coupling {
  Plan P;
  // the element-sum slots of every split pair (those of its destination's
  // tiles that the series does not take), compacted: slots off .. off+n
  struct Split { int dst, src; size_t off; int n; };
  std::vector<Split> splits;
  double *d_spt = nullptr, *d_snm = nullptr;
  int *d_sidx = nullptr;
  std::vector<int> slot0;
  int maxslot = 0, totslot = 0;
  int *d_did = nullptr;
  double *d_part = nullptr, *d_ccd = nullptr, *d_ccs = nullptr;
  std::vector<size_t> coff;
  cd *d_coef = nullptr;
  std::vector<MpTile> tiles;
  std::vector<MpPair> pairs;
  MpTile *d_tiles = nullptr;
  MpPair *d_pairs = nullptr;
  double evwork = 0.0;

  // split: pairs whose slots do not all qualify for the series take it on
  // the tiles whose slots do (the one-beam loop)
  int init(bh_object **objs, int N, bool split = false)
  {
    make_plan(objs, N, P, split);
    g_last_series = P.npairs_mp;
    g_last_direct = P.npairs_direct;
    g_last_split = P.npairs_split;
    g_last_lmax = -1;
    for (int n = 0; n < N; n++) g_last_lmax = std::max(g_last_lmax, P.Lsrc[n]);
    slot0.assign(N, 0);
    for (int m = 0; m < N; m++) { maxslot = std::max(maxslot, objs[m]->nslot); slot0[m] = totslot; totslot += objs[m]->nslot; }
    d_did = ws<int>(W_DID, maxslot);
    d_part = ws<double>(W_PART, (size_t)maxslot * 19);   // far-triangle sums of the direct pass
    d_ccd = ws<double>(W_CCD, 2 * (size_t)totslot);
    d_ccs = ws<double>(W_CCDSUM, 2);
    if (!d_did || !d_part || !d_ccd || !d_ccs) return fail(BH_EHIP, "coupling loop: work area");
    // series coefficients of every source object, and the evaluation tiles
    coff.assign(N, 0);
    size_t ncoef = 0;
    for (int n = 0; n < N; n++)
      if (P.Lsrc[n] >= 0) { coff[n] = ncoef; ncoef += (size_t)mp_count(P.Lsrc[n] + 1) * 16; }
    d_coef = ncoef ? ws<cd>(W_COEF, ncoef) : nullptr;
    auto mp_pair = [&](int n, int L) {
      MpPair q{};
      q.coef = d_coef + coff[n];
      for (int d = 0; d < 3; d++) q.c[d] = P.geo[n].c[d];
      q.kr = objs[n]->kn[0].real();
      q.Lp = P.Lsrc[n] + 1;
      q.Le = L + 1;
      return q;
    };
    for (int m = 0; m < N; m++) {
      bh_object *o = objs[m];
      bool any_split = false;
      for (int n = 0; n < N; n++) any_split = any_split || P.Ls[m][n] >= 0;
      // the object's series sources, shared by its tiles unless a pair is split
      const int p0 = (int)pairs.size();
      for (int n = 0; n < N; n++)
        if (P.Lmn[m][n] >= 0) pairs.push_back(mp_pair(n, P.Lmn[m][n]));
      const int np0 = (int)pairs.size() - p0;
      for (int s0 = 0; s0 < o->nslot; s0 += 64) {
        MpTile t{};
        t.pts = o->d_slotPt; t.nrm = o->d_slotN; t.slot = o->d_slotIdx; t.Ui = o->d_Ui; t.dUi = o->d_dUi;
        t.s0 = s0; t.ns = std::min(64, o->nslot - s0);
        t.p0 = p0; t.np = np0;
        if (any_split) {
          // this tile's own list, sources in order: the series pairs and the
          // split pairs whose series takes this tile
          t.p0 = (int)pairs.size();
          for (int n = 0; n < N; n++) {
            if (P.Lmn[m][n] >= 0) pairs.push_back(mp_pair(n, P.Lmn[m][n]));
            else if (P.Ls[m][n] >= 0 && P.tile_mp[m][n][s0 / 64]) pairs.push_back(mp_pair(n, P.Ls[m][n]));
          }
          t.np = (int)pairs.size() - t.p0;
        }
        if (t.np > 0) tiles.push_back(t);
      }
    }
    if (!tiles.empty()) {
      d_tiles = ws<MpTile>(W_TILES, tiles.size());
      d_pairs = ws<MpPair>(W_PAIRS, pairs.size());
      if (!d_coef || !d_tiles || !d_pairs) return fail(BH_EHIP, "coupling loop: series work area");
      HIPCHK(hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(MpTile), hipMemcpyHostToDevice, g.st));
      HIPCHK(hipMemcpyAsync(d_pairs, pairs.data(), pairs.size() * sizeof(MpPair), hipMemcpyHostToDevice, g.st));
      HIPCHK(hipStreamSynchronize(g.st));   // the host tables are locals of this record
    }
    for (const MpTile &t : tiles)
      for (int i = 0; i < t.np; i++) evwork += (double)t.ns * mp_count(pairs[t.p0 + i].Le);
    // split pairs: their element-sum slots as one compact list per pair
    std::vector<int> lst;
    for (int d = 0; d < N; d++)
      for (int m = 0; m < N; m++) {
        if (d == m || P.Lmn[d][m] != -1 || P.Ls[d][m] < 0) continue;
        const std::vector<char> &fl = P.tile_mp[d][m];
        Split q{d, m, lst.size(), 0};
        for (int p = 0; p < objs[d]->nslot; p++)
          if (!fl[p / 64]) lst.push_back(p);
        q.n = (int)(lst.size() - q.off);
        splits.push_back(q);
      }
    if (!lst.empty()) {
      d_spt = ws<double>(W_SPT, 3 * lst.size());
      d_snm = ws<double>(W_SNM, 3 * lst.size());
      d_sidx = ws<int>(W_SIDX, lst.size());
      int *d_lst = ws<int>(W_SLST, lst.size());
      if (!d_spt || !d_snm || !d_sidx || !d_lst) return fail(BH_EHIP, "coupling loop: slot lists of the split pairs");
      HIPCHK(hipMemcpyAsync(d_lst, lst.data(), lst.size() * sizeof(int), hipMemcpyHostToDevice, g.st));
      for (const Split &q : splits) {
        if (q.n <= 0) continue;
        const bh_object *o = objs[q.dst];
        hipLaunchKernelGGL(k_gather_slots, g1((size_t)q.n), dim3(256), 0, g.st, o->d_slotPt, o->d_slotN, o->d_slotIdx,
                           d_lst + q.off, q.n, d_spt + 3 * q.off, d_snm + 3 * q.off, d_sidx + q.off);
      }
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(g.st));   // the host list is a local
    }
    return 0;
  }

  // the field of every object at the other objects' node slots from the
  // current boundary values, added into their incident data; check: first
  // verify that no slot of an element-sum pair lies inside its source
  int sums(bh_object **objs, int N, bool check)
  {
    // series pairs: coefficients of every source, then one evaluation launch
    for (int n = 0; n < N; n++)
      if (P.Lsrc[n] >= 0) { int r = mp_coefficients(objs[n], P, n, d_coef + coff[n]); if (r) return r; }
    if (!tiles.empty()) {
      bem_prof_begin(BH_K_MP_EVAL, evwork, g.st);
      hipLaunchKernelGGL(k_mp_eval, dim3((unsigned)tiles.size()), dim3(64), 0, g.st, d_tiles, d_pairs);
      bem_prof_end(BH_K_MP_EVAL, g.st);
    }
    // the other pairs: direct sums (of a split pair, over its compact list of
    // the slots the series does not take: one launch per pair)
    for (int m = 0; m < N; m++)
      for (int d = 0; d < N; d++) {
        if (d == m || P.Lmn[d][m] != -1) continue;
        bh_object *dst = objs[d];
        if (P.Ls[d][m] < 0) {
          int r = add_scattered(dst, objs[m], d_did, check, d_part);
          if (r) return r;
          continue;
        }
        for (const Split &q : splits)
          if (q.dst == d && q.src == m && q.n > 0) {
            int r = add_scattered_at(dst, objs[m], d_spt + 3 * q.off, d_snm + 3 * q.off, d_sidx + q.off, q.n, d_did,
                                     check, d_part);
            if (r) return r;
          }
      }
    HIPCHK(hipGetLastError());
    return 0;
  }

  // the coupling sums of every beam of the many-beam loop (incident data
  // and boundary values of the beams, nb of them)
  int sums_multi(bh_object **objs, int N, int nb, bool check)
  {
    std::vector<size_t> cofm(N, 0);
    size_t ncm = 0;
    for (int n = 0; n < N; n++)
      if (P.Lsrc[n] >= 0) { cofm[n] = ncm; ncm += (size_t)mp_count(P.Lsrc[n] + 1) * MP_NDENS * nb; }
    cd *cm = ncm ? ws<cd>(W_COEF, ncm) : nullptr;
    if (ncm && !cm) return fail(BH_EHIP, "series: coefficients of %d beams", nb);
    for (int n = 0; n < N; n++)
      if (P.Lsrc[n] >= 0) { int r = mp_coefficients_multi(objs[n], P, n, cm + cofm[n], nb); if (r) return r; }
    const size_t ncol = (size_t)MP_NDENS * nb;
    for (int m = 0; m < N; m++) {
      bh_object *d = objs[m];
      const size_t is = ((size_t)d->Ne + 1) * 16;
      for (int n = 0; n < N; n++) {
        if (P.Lmn[m][n] < 0) continue;
        const int Lp = P.Lsrc[n] + 1, nidx = mp_count(Lp);
        const int sc = (int)std::max<size_t>(64, std::min<size_t>((size_t)d->nslot,
                                                                   std::min(((size_t)64 << 20) / nidx, ((size_t)64 << 20) / ncol)));
        cd *Sb = ws<cd>(W_MS, (size_t)sc * nidx), *Cm = ws<cd>(W_MC, (size_t)sc * ncol);
        if (!Sb || !Cm) return fail(BH_EHIP, "series: evaluation work area");
        for (int s0 = 0; s0 < d->nslot; s0 += sc) {
          const int ns = std::min(sc, d->nslot - s0);
          bem_prof_begin(BH_K_MP_EVAL, (double)ns * nidx * nb, g.st);
          hipLaunchKernelGGL(k_mp_sbasis, dim3((ns + 63) / 64), dim3(64), 0, g.st, d->d_slotPt + 3 * (size_t)s0, ns,
                             P.geo[n].c[0], P.geo[n].c[1], P.geo[n].c[2], objs[n]->kn[0].real(), Lp, Sb);
          bem_zgemm(ns, (int)ncol, nidx, 1.0, 0.0, Sb, (size_t)nidx, cm + cofm[n], ncol, 0.0, 0.0, Cm, ncol, g.st);
          hipLaunchKernelGGL(k_mp_combine, g1((size_t)ns * nb), dim3(256), 0, g.st, Cm, ns, nb, d->d_slotN, d->d_slotIdx,
                             s0, mb_ptr(d, 2), mb_ptr(d, 3), is);
          bem_prof_end(BH_K_MP_EVAL, g.st);
        }
      }
    }
    // the other pairs: element sums, one beam at a time (the objects'
    // one-beam state pointers lent to the beam's arrays)
    for (int m = 0; m < N; m++)
      for (int d = 0; d < N; d++) {
        if (d == m || P.Lmn[d][m] != -1) continue;
        bh_object *dst = objs[d], *src = objs[m];
        cd *keep[4] = {dst->d_Ui, dst->d_dUi, src->dom[0].d_U, src->dom[0].d_dU};
        const size_t is = ((size_t)dst->Ne + 1) * 16, us = (size_t)(src->dom[0].dNe + 1) * 16;
        int r = 0;
        for (int b = 0; b < nb && !r; b++) {
          dst->d_Ui = mb_ptr(dst, 2) + b * is;
          dst->d_dUi = mb_ptr(dst, 3) + b * is;
          src->dom[0].d_U = mb_dom_ptr(src, 0, 0) + b * us;
          src->dom[0].d_dU = mb_dom_ptr(src, 0, 1) + b * us;
          r = add_scattered(dst, src, d_did, check && b == 0, d_part);
        }
        dst->d_Ui = keep[0]; dst->d_dUi = keep[1]; src->dom[0].d_U = keep[2]; src->dom[0].d_dU = keep[3];
        if (r) return r;
      }
    HIPCHK(hipGetLastError());
    return 0;
  }

  // the change measure of every beam; the largest is returned (-1 on a failure)
  double change_multi(bh_object **objs, int N, int nb)
  {
    double *part = ws<double>(W_CCD, 2 * (size_t)totslot * nb), *out = ws<double>(W_CCDSUM, 2 * (size_t)nb);
    if (!part || !out) return -1.0;
    for (int b = 0; b < nb; b++) {
      for (int m = 0; m < N; m++) {
        bh_object *o = objs[m];
        const size_t is = ((size_t)o->Ne + 1) * 16;
        hipLaunchKernelGGL(k_ccd, g1(o->nslot), dim3(256), 0, g.st, mb_ptr(o, 2) + b * is, mb_ptr(o, 4) + b * is,
                           o->d_slotIdx, o->nslot, part + 2 * ((size_t)b * totslot + slot0[m]));
      }
      hipLaunchKernelGGL(k_ccd_sum, dim3(1), dim3(256), 0, g.st, part + 2 * (size_t)b * totslot, totslot, out + 2 * b);
    }
    std::vector<double> h(2 * (size_t)nb);
    if (hipMemcpyAsync(h.data(), out, h.size() * sizeof(double), hipMemcpyDeviceToHost, g.st) != hipSuccess ||
        hipStreamSynchronize(g.st) != hipSuccess)
      return -1.0;
    double cc = 0.0;
    for (int b = 0; b < nb; b++) cc = std::max(cc, std::max(h[2 * b], h[2 * b + 1]) / (double)totslot);
    return cc;
  }

  // ccd_f (mo_solve.c:413-455) over every object's slots; -1 on a failure
  double change(bh_object **objs, int N)
  {
    for (int m = 0; m < N; m++) {
      bh_object *o = objs[m];
      hipLaunchKernelGGL(k_ccd, g1(o->nslot), dim3(256), 0, g.st, o->d_Ui, o->d_Uip, o->d_slotIdx, o->nslot,
                         d_ccd + 2 * (size_t)slot0[m]);
    }
    hipLaunchKernelGGL(k_ccd_sum, dim3(1), dim3(256), 0, g.st, d_ccd, totslot, d_ccs);
    double h[2];
    if (hipMemcpyAsync(h, d_ccs, sizeof h, hipMemcpyDeviceToHost, g.st) != hipSuccess ||
        hipStreamSynchronize(g.st) != hipSuccess)
      return -1.0;
    return (h[0] > h[1] ? h[0] : h[1]) / (double)totslot;
  }
};

__global__ void k_copy_cd(cd *__restrict__ dst, const cd *__restrict__ src, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// the coupling sums of the loop (series or element sums, as the loop would
// take them) from the objects' current boundary values: out receives the
// scattered U and dU/dn at object dst's node slots in its incident-data
// layout, 2 x [Ne+1][4][4] complex (U, then dU); the incident data are left
// as they were
extern "C" int bh_mo_coupling_sums(bh_object **objs, int N, int dst, double *out)
{
  API_LOCK;
  if (N < 2 || dst < 0 || dst >= N) return fail(BH_EARG, "coupling sums: %d objects, destination %d", N, dst);
  if (ensure_rules()) return BH_EHIP;
  for (int m = 0; m < N; m++)
    if (!objs[m]->dom[0].d_U || !objs[m]->d_Ui) return fail(BH_ESTATE, "object %d is not solved", m);
  Coupling C;
  int rc = C.init(objs, N, true);   // the plan of the one-beam loop
  if (rc) return rc;
  // keep the incident data, clear them, add the sums, read dst's, restore
  std::vector<cd *> keep(2 * (size_t)N);
  for (int m = 0; m < N; m++) {
    const size_t n = (size_t)(objs[m]->Ne + 1) * 16;
    for (int q = 0; q < 2; q++) {
      cd *a = q ? objs[m]->d_dUi : objs[m]->d_Ui;
      HIPCHK(hipMalloc((void **)&keep[2 * m + q], n * sizeof(cd)));
      hipLaunchKernelGGL(k_copy_cd, g1(n), dim3(256), 0, g.st, keep[2 * m + q], a, n);
      hipLaunchKernelGGL(k_zero_cd, g1(n), dim3(256), 0, g.st, a, n);
    }
  }
  rc = C.sums(objs, N, true);
  const size_t nd = (size_t)(objs[dst]->Ne + 1) * 16;
  if (!rc && (hipMemcpyAsync(out, objs[dst]->d_Ui, nd * sizeof(cd), hipMemcpyDeviceToHost, g.st) != hipSuccess ||
              hipMemcpyAsync(out + 2 * nd, objs[dst]->d_dUi, nd * sizeof(cd), hipMemcpyDeviceToHost, g.st) != hipSuccess))
    rc = fail(BH_EHIP, "coupling sums: read-back");
  for (int m = 0; m < N; m++) {
    const size_t n = (size_t)(objs[m]->Ne + 1) * 16;
    for (int q = 0; q < 2; q++)
      hipLaunchKernelGGL(k_copy_cd, g1(n), dim3(256), 0, g.st, q ? objs[m]->d_dUi : objs[m]->d_Ui, keep[2 * m + q], n);
  }
  HIPCHK(hipStreamSynchronize(g.st));
  for (cd *p : keep) dfree(p);
  return rc;
}

// object_factor_g of N objects by a worker thread on a stream of its own
// (created once, default priority); finish() joins it and returns its first
// failure; the destructor joins it too, so no return path leaves it running
struct GPrefactor {
  std::thread th;
  int rc = 0;
  std::string msg;
  void start(bh_object **objs, int N, const std::vector<double> &cost)
  {
    bool any = false;
    for (int m = 0; m < N; m++) any = any || objs[m]->fmode == 3;
    const char *e = getenv("BEM3_HIP_G1_AHEAD");   // 0: the boundary fields factorise G1 themselves
    if (!any || (e && strcmp(e, "0") == 0)) return;
    if (!g_gpre_st && hipStreamCreate(&g_gpre_st) != hipSuccess) { g_gpre_st = nullptr; return; }
    std::vector<int> order(N);
    for (int i = 0; i < N; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<bh_object *> list;
    for (int i : order) list.push_back(objs[i]);
    th = std::thread([this, list]() {
      set_cur_st(g_gpre_st);
      for (bh_object *o : list) {
        const int r = object_factor_g(o);
        if (r) { rc = r; msg = bh_last_error(); break; }
      }
      if (hipStreamSynchronize(g_gpre_st) != hipSuccess && !rc) { rc = BH_EHIP; msg = "G1 factorisation stream"; }
      set_cur_st(nullptr);
    });
  }
  int finish()
  {
    if (th.joinable()) th.join();
    if (rc) return fail(rc, "%s", msg.c_str());
    return 0;
  }
  ~GPrefactor() { if (th.joinable()) th.join(); }
};

// block-Jacobi fixed point over objects (mo_solve_bv, mo_solve.c:10-88).
// trace (optional, iter_max+1 doubles) receives cc per iteration; *iters
// receives the final iteration count.
extern "C" int bh_mo_solve(bh_object **objs, int N, double ccd, int iter_min, int iter_max, double *trace, int *iters)
{
  API_LOCK;
  if (N <= 0) return fail(BH_EARG, "no objects");
  if (ensure_rules()) return BH_EHIP;
  for (int m = 0; m < N; m++)
    if (!objs[m]->fmode || !objs[m]->d_beams) return fail(BH_ESTATE, "object %d: factor and set beams first", m);
  int rc, iter = 0;
  const std::vector<double> c3 = cost_of(objs, N, 3), c2 = cost_of(objs, N, 2);
  GPrefactor gpre;   // joined on every return
  // the mixed form's objects solve in lock step (objects_solve_batch)
  // unless BEM3_HIP_BATCH_SOLVE=0; objects it does not take run on the lanes
  const char *be = getenv("BEM3_HIP_BATCH_SOLVE");
  const bool batch = !(be && strcmp(be, "0") == 0);
  auto solve_all = [&]() {
    if (batch && N > 1) {
      HIPCHK(hipStreamSynchronize(g.st));
      const int r = objects_solve_batch(objs, N);
      if (r != BH_EUNSUP) return r;
    }
    return run_lanes(objs, N, c2, [](bh_object *o) { return object_solve(o); });
  };
  // every loop starts from the incident field alone (Ui = Ui0), the state
  // mo_initialize leaves (init_boundary_data_u2, mo_setup.c:704-757): the
  // incident data of an earlier loop on the same objects hold its coupling
  // sums, which would start this one at that loop's converged state
  // and its solves start from zero, not from an earlier loop's solutions,
  // so that a loop repeats the same arithmetic whatever ran before it
  for (int m = 0; m < N; m++) {
    bh_object *o = objs[m];
    const size_t n = (size_t)(o->Ne + 1) * 16;
    hipLaunchKernelGGL(k_reset_incident, g1(n), dim3(256), 0, g.st, o->d_Ui, o->d_dUi, o->d_Ui0, o->d_dUi0, o->d_Uip, n);
    o->x_warm = false;
    o->r_valid = false;
  }
  HIPCHK(hipGetLastError());
  if (N == 1) {
    if ((rc = object_solve(objs[0]))) return rc;
  } else {
    // many solves per object: form the Schur complements once
    const auto tl0 = std::chrono::steady_clock::now();
    if ((rc = run_lanes(objs, N, c3, [](bh_object *o) { return form_s(o); }))) return rc;
    // the FP32 LUs of G1 that the boundary fields at the end solve with
    // (mixed form) are formed beside the sweeps, on a stream of their own
    // by a worker thread, largest object first: the sweeps' kernels leave
    // the device room
    gpre.start(objs, N, c3);
    Coupling C;
    if ((rc = C.init(objs, N, true))) return rc;
    bool first = true;
    // wall seconds of the last sweep's coupling sums (to their completion)
    // and of its solves, for the progress lines
    double t_sums = 0.0, t_solves = 0.0;
    // The reset of the incident data and the coupling sums of a sweep are
    // kernels only, with the same arguments every sweep once the first one
    // has checked the geometry and sized the work areas: the second sweep
    // captures them as a graph and every later sweep launches it (one
    // launch in place of about 6 per source object).  BEM3_HIP_GRAPH=0, or
    // the kernel-time registry switched on (its events do not go into a
    // graph): launched one by one.
    const char *ge = getenv("BEM3_HIP_GRAPH");
    bool use_graph = !(ge && strcmp(ge, "0") == 0) && !bem_prof_on();
    struct GraphExec {
      hipGraphExec_t x = nullptr;
      ~GraphExec() { if (x) (void)hipGraphExecDestroy(x); }
    } gx;
    auto sums_kernels = [&](bool check) -> int {
      for (int m = 0; m < N; m++) {
        bh_object *o = objs[m];
        size_t n = (size_t)(o->Ne + 1) * 16;
        hipLaunchKernelGGL(k_reset_incident, g1(n), dim3(256), 0, g.st, o->d_Ui, o->d_dUi, o->d_Ui0, o->d_dUi0, o->d_Uip, n);
      }
      return C.sums(objs, N, check);
    };
    auto sweep = [&]() -> int {
      const auto t0 = std::chrono::steady_clock::now();
      int r = 0;
      if (first || !use_graph) {
        r = sums_kernels(first);
      } else {
        if (!gx.x) {
          hipGraph_t gr = nullptr;
          HIPCHK(hipStreamBeginCapture(g.st, hipStreamCaptureModeThreadLocal));
          r = sums_kernels(false);
          const hipError_t ec = hipStreamEndCapture(g.st, &gr);
          if (r) { if (gr) (void)hipGraphDestroy(gr); return r; }
          if (ec != hipSuccess || !gr) return fail(BH_EHIP, "coupling loop: capture of the sweep's sums");
          const hipError_t ei = hipGraphInstantiate(&gx.x, gr, nullptr, nullptr, 0);
          (void)hipGraphDestroy(gr);
          if (ei != hipSuccess) { gx.x = nullptr; return fail(BH_EHIP, "coupling loop: graph of the sweep's sums"); }
        }
        HIPCHK(hipGraphLaunch(gx.x, g.st));
      }
      if (r) return r;
      first = false;
      HIPCHK(hipStreamSynchronize(g.st));   // run_lanes waits for it anyway
      const auto t1 = std::chrono::steady_clock::now();
      r = solve_all();
      t_sums = std::chrono::duration<double>(t1 - t0).count();
      t_solves = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
      return r;
    };
    // BEM3_HIP_PROGRESS=1: one line per iteration on stderr; any other
    // value but 0 names a file the lines are appended to (long solves)
    const char *pe = getenv("BEM3_HIP_PROGRESS");
    FILE *pf = nullptr;
    if (pe && *pe && strcmp(pe, "0") != 0) pf = strcmp(pe, "1") == 0 ? stderr : fopen(pe, "a");
    struct Closer { FILE *f; ~Closer() { if (f && f != stderr) fclose(f); } } closer{pf};
    if (pf) {
      fprintf(pf, "bh_mo_solve: %d objects, %d object pairs by series, %d by direct sums, %d split; Schur complements %.3f s\n",
              N, C.P.npairs_mp, C.P.npairs_direct, C.P.npairs_split, std::chrono::duration<double>(std::chrono::steady_clock::now() - tl0).count());
      fflush(pf);
    }
    {
      const auto t0 = std::chrono::steady_clock::now();
      if ((rc = solve_all())) return rc;
      if (pf) {
        fprintf(pf, "bh_mo_solve: first solves %.3f s\n", std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        fflush(pf);
      }
    }
    if ((rc = sweep())) return rc;
    if (pf) { fprintf(pf, "bh_mo_solve: count = 1, sums %.3f s, solves %.3f s\n", t_sums, t_solves); fflush(pf); }
    for (iter = 2; iter < iter_max; iter++) {
      if ((rc = sweep())) return rc;
      double cc = C.change(objs, N);
      if (cc < 0.0) return fail(BH_EHIP, "coupling loop: change measure");
      if (trace) trace[iter] = cc;
      if (pf) {
        fprintf(pf, "bh_mo_solve: count = %d, cc = %g, sums %.3f s%s, solves %.3f s\n", iter, cc, t_sums,
                gx.x ? " (graph)" : "", t_solves);
        fflush(pf);
      }
      if (cc < ccd && iter > iter_min) break;
    }
  }
  if ((rc = gpre.finish())) return rc;
  const auto te0 = std::chrono::steady_clock::now();
  if ((rc = run_lanes(objs, N, c3, [](bh_object *o) { return object_solve_EH(o); }))) return rc;
  if (const char *pe = getenv("BEM3_HIP_PROGRESS"))
    if (*pe && strcmp(pe, "0") != 0) {
      FILE *pf = strcmp(pe, "1") == 0 ? stderr : fopen(pe, "a");
      if (pf) {
        fprintf(pf, "bh_mo_solve: boundary E/H %.3f s\n", std::chrono::duration<double>(std::chrono::steady_clock::now() - te0).count());
        if (pf != stderr) fclose(pf); else fflush(pf);
      }
    }
  if (iters) *iters = iter;
  return 0;
}

// Many plane waves as separate right-hand sides (no counterpart in the
// reference, which sums every beam of a set into one incident field,
// multi_fbeam.c:764-815, so that per-beam results take one run per beam):
// nb beams (12 doubles each, as bh_object_set_beams) given to every object,
// the coupling loop (mo_solve_bv, mo_solve.c:10-88) run for all of them at
// once -- the series coefficients of all beams from one basis per source
// object, the block solves of all beams as products with 3 nb to 4 nb
// columns -- until every beam's change is below ccd after more than
// iter_min iterations, then the boundary E, H of every beam.  Every object
// must be in the mixed-precision form and factorised.
extern "C" int bh_mo_solve_beams(bh_object **objs, int N, int nb, const double *par, double ccd, int iter_min,
                                 int iter_max, int *iters)
{
  API_LOCK;
  if (N <= 0 || nb <= 0) return fail(BH_EARG, "many beams: %d objects, %d beams", N, nb);
  if (ensure_rules()) return BH_EHIP;
  for (int m = 0; m < N; m++)
    if (objs[m]->fmode != 3) return fail(BH_EUNSUP, "many beams: object %d is not factorised in the mixed-precision form", m);
  int rc, iter = 0;
  for (int m = 0; m < N; m++)
    if ((rc = object_beams_set(objs[m], nb, par))) return rc;
  const std::vector<double> c3 = cost_of(objs, N, 3), c2 = cost_of(objs, N, 2);
  if ((rc = run_lanes(objs, N, c3, [](bh_object *o) { return form_s(o); }))) return rc;
  auto solve_all = [&]() { return run_lanes(objs, N, c2, [](bh_object *o) { return object_solve_multi(o); }); };
  if ((rc = solve_all())) return rc;
  if (N > 1) {
    Coupling C;
    if ((rc = C.init(objs, N))) return rc;
    bool first = true;
    auto sweep = [&]() -> int {
      for (int m = 0; m < N; m++) {
        bh_object *o = objs[m];
        const size_t n = ((size_t)o->Ne + 1) * 16 * nb;
        hipLaunchKernelGGL(k_reset_incident, g1(n), dim3(256), 0, g.st, mb_ptr(o, 2), mb_ptr(o, 3), mb_ptr(o, 0),
                           mb_ptr(o, 1), mb_ptr(o, 4), n);
      }
      int r = C.sums_multi(objs, N, nb, first);
      if (r) return r;
      first = false;
      return solve_all();
    };
    const char *pe = getenv("BEM3_HIP_PROGRESS");
    FILE *pf = nullptr;
    if (pe && *pe && strcmp(pe, "0") != 0) pf = strcmp(pe, "1") == 0 ? stderr : fopen(pe, "a");
    struct Closer { FILE *f; ~Closer() { if (f && f != stderr) fclose(f); } } closer{pf};
    if ((rc = sweep())) return rc;
    for (iter = 2; iter < iter_max; iter++) {
      if ((rc = sweep())) return rc;
      const double cc = C.change_multi(objs, N, nb);
      if (cc < 0.0) return fail(BH_EHIP, "many beams: change measure");
      if (pf) { fprintf(pf, "bh_mo_solve_beams: count = %d, largest cc = %g\n", iter, cc); fflush(pf); }
      if (cc < ccd && iter > iter_min) break;
    }
  }
  if ((rc = run_lanes(objs, N, c3, [](bh_object *o) { return object_solve_EH_multi(o); }))) return rc;
  if (iters) *iters = iter;
  return 0;
}

// scattered E, H of every beam of the many-beam solve at np points of the
// open region (mo_EH_mEMP_s summed over the objects, mo_field.c:154-181,
// d3b1_field.c:68-138; example4.c:51-177 calls it once per point in one
// run per beam).  Each object's potentials are the series of the points of
// rule type (0..3, as dcoef_rt: at these distances no element switches
// rule) about its centre, evaluated for all beams as products
// (points x coefficients) x (coefficients x 16 nb).  Every point must lie
// at least 1/0.6 bounding radii from every object's centre (order as the
// coupling series: truncation about 1e-10); the discrete solid-angle sums
// F, dF of a closed surface seen from outside are quadrature error there
// and are taken as zero.  what 0: E, H ([nb][np][6] complex); 1: |E|^2,
// |H|^2 ([nb][np][2]).
extern "C" int bh_mo_far_field_beams(bh_object **objs, int N, int type, int np, const double *pts, int what,
                                     double *out)
{
  API_LOCK;
  if (N <= 0 || np < 0 || type < 0 || type > 3 || what < 0 || what > 1) return fail(BH_EARG, "bad far-field request");
  const int nb = objs[0]->nbm;
  for (int m = 0; m < N; m++) {
    const bh_object *o = objs[m];
    if (o->nbm != nb || nb <= 0) return fail(BH_ESTATE, "object %d: no many-beam solve of %d beams", m, nb);
    if (o->nquad != 0 || o->kn[0].imag() != 0.0 || !(o->kn[0].real() > 0.0))
      return fail(BH_EUNSUP, "far field of many beams: object %d (triangles in a lossless medium only)", m);
  }
  if (np == 0) return 0;
  if (ensure_rules()) return BH_EHIP;
  // series order of each object from the nearest point
  std::vector<SrcGeo> geo(N);
  std::vector<int> Ls(N);
  for (int m = 0; m < N; m++) {
    geo[m] = src_geo(objs[m]);
    double d2 = 1e300;
    for (int p = 0; p < np; p++) {
      double q = 0.0;
      for (int d = 0; d < 3; d++) q += (pts[3 * (size_t)p + d] - geo[m].c[d]) * (pts[3 * (size_t)p + d] - geo[m].c[d]);
      d2 = std::min(d2, q);
    }
    const double rho = geo[m].rad / sqrt(d2), kR = objs[m]->kn[0].real() * geo[m].rad;
    if (!(rho <= MP_RHO))
      return fail(BH_EUNSUP, "far field of many beams: a point lies %.4g from object %d's centre (bounding radius %.4g), "
                  "closer than the series allows (bh_mo_field per beam)", sqrt(d2), m, geo[m].rad);
    Ls[m] = std::max((int)ceil(log(MP_EPS) / log(rho)), (int)ceil(kR + 6.0 * cbrt(kR)) + 4);
    if (Ls[m] > MP_LMAX) return fail(BH_EUNSUP, "far field of many beams: object %d needs order %d", m, Ls[m]);
  }
  const size_t ncol = (size_t)MP_NDENS * nb, neh = (size_t)nb * np * 6;
  double *d_pts = ws<double>(W_FPTS, 3 * (size_t)np);
  cd *EH = ws<cd>(W_FEH, neh);
  if (!d_pts || !EH) return fail(BH_EHIP, "far field: %d points of %d beams", np, nb);
  HIPCHK(hipMemcpyAsync(d_pts, pts, 3 * (size_t)np * sizeof(double), hipMemcpyHostToDevice, g.st));
  hipLaunchKernelGGL(k_zero_cd, g1(neh), dim3(256), 0, g.st, EH, neh);
  for (int m = 0; m < N; m++) {
    bh_object *o = objs[m];
    MpSrc S{};
    for (int d = 0; d < 3; d++) S.c[d] = geo[m].c[d];
    S.kr = o->kn[0].real();
    S.scale = g_rscale[type];
    S.L = Ls[m];
    S.Lp = S.L + 1;
    S.qoff = 0;
    S.npt = g_nrule[type];
    S.dNe = o->dom[0].dNe;
    S.nq = S.dNe * S.npt;
    const int nidx = mp_count(S.Lp);
    cd *coef = ws<cd>(W_COEF, (size_t)nidx * ncol);
    if (!coef) return fail(BH_EHIP, "far field: coefficients of %d beams", nb);
    int r = series_multi(o, S, (const QPtN *)g_rule[type], coef, nb);
    if (r) return r;
    const int pc = (int)std::max<size_t>(64, std::min<size_t>((size_t)np, std::min(((size_t)64 << 20) / nidx,
                                                                                   ((size_t)64 << 20) / ncol)));
    cd *Sb = ws<cd>(W_MS, (size_t)pc * nidx), *Cm = ws<cd>(W_MC, (size_t)pc * ncol);
    if (!Sb || !Cm) return fail(BH_EHIP, "far field: evaluation work area");
    const double chi = -o->lambda0 / (2.0 * M_PI);   // ch = lambda0 / (2 pi i)
    for (int p0 = 0; p0 < np; p0 += pc) {
      const int n = std::min(pc, np - p0);
      bem_prof_begin(BH_K_MP_EVAL, (double)n * nidx * nb, g.st);
      hipLaunchKernelGGL(k_mp_sbasis, dim3((n + 63) / 64), dim3(64), 0, g.st, d_pts + 3 * (size_t)p0, n, S.c[0], S.c[1],
                         S.c[2], S.kr, S.Lp, Sb);
      bem_zgemm(n, (int)ncol, nidx, 1.0, 0.0, Sb, (size_t)nidx, coef, ncol, 0.0, 0.0, Cm, ncol, g.st);
      hipLaunchKernelGGL(k_mp_far_eh, g1((size_t)n * nb), dim3(256), 0, g.st, Cm, n, nb, 0.0, chi, p0,
                         (size_t)np * 6, EH);
      bem_prof_end(BH_K_MP_EVAL, g.st);
    }
  }
  if (what == 0) {
    HIPCHK(hipMemcpyAsync(out, EH, neh * sizeof(cd), hipMemcpyDeviceToHost, g.st));
  } else {
    double *I = ws<double>(W_FI, 2 * (size_t)nb * np);
    if (!I) return fail(BH_EHIP, "far field: intensities");
    hipLaunchKernelGGL(k_eh_intensity, g1((size_t)nb * np), dim3(256), 0, g.st, EH, (size_t)nb * np, I);
    HIPCHK(hipMemcpyAsync(out, I, 2 * (size_t)nb * np * sizeof(double), hipMemcpyDeviceToHost, g.st));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(g.st));
  return 0;
}
