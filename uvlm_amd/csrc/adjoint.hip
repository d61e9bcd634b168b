// adjoint.hip -- vector-Jacobian product of the particle rates (velocity and stretching of one
// evaluation of FMMCalculator::calculate, src/fmmCalculator.hpp:420-467, cVPM), in FP64.
//
// With cotangents ubar_i (of u_i) and abar_i (of dadt_i) the objective is
//   S = sum_i ubar_i . u_i + Jbar_i : J_i,
//   Jbar_i = alpha_i abar_i^T (transposed stretching, dadt = J^T alpha) or abar_i alpha_i^T
//   (classic, dadt = J alpha),
// and the product returns xbar = dS/dx and alphabar = dS/dalpha for a frozen tree and lists:
//   alphabar_i = (J_i abar_i or J_i^T abar_i)                        the local stretching term
//              + sum_j [near] -W_ji + [far] curl(phi)(x_i) / (4 pi)  (phi below)
//   xbar_i     = sum_j [near, i target] dS_ij/dr - sum_j [near, i source] dS_ji/dr
//              + [far, i target] J_i^T ubar_i + Jbar_i : grad J_i
//              + [far, i source] (grad F)^T alpha_i,  F = curl(phi) / (4 pi).
// phi is the far field of the adjoint sources: vortices ubar_i and vortex dipoles Jbar_i at the
// particles (P2M of far.hip), carried by the same expansions along the TRANSPOSED M2L list:
// the total-degree truncation j + n <= P - 1 of M2L is the Taylor expansion of 1/|R + d| in
// d = (x_t - X_t) - (x_s - X_s), even in (R, d), so the adjoint of the far field is exactly the
// far field with targets and sources exchanged.
// Near pair (r = x_i - x_j, core sigma_j), with e the axial vector of Jbar_i:
//   S_ij = -alpha_j . W,   W = f1 (ubar_i x r + e) + f2 (Jbar_i r) x r,
//   dS_ij/dr = -[ f2 (alpha_j . (ubar_i x r + e)) r + f1 (alpha_j x ubar_i)
//                 + f3 (alpha_j . ((Jbar_i r) x r)) r + f2 Jbar_i^T (r x alpha_j)
//                 + f2 alpha_j x (Jbar_i r) ],
//   f1 = g / (4 pi R^3), f2 = (t2 / sigma^2 - 3 g / R^2) / (4 pi R^3),
//   f3 = (15 g / R^4 - t2 (5 + 2 s^2) / (sigma^2 R^2)) / (4 pi R^3),
// s^2 = R^2 / (2 sigma^2), t2 = (2/sqrt(pi)) s e^{-s^2}, g = erf(s) - t2 (the kernel of
// multipoleKernels.hpp:223-301 and its R^2 derivatives).
//
// Adjoint fields (SoA, `cap` doubles each): 0-2 ubar, 3-11 Jbar (row-major), 12-14 xbar,
// 15-17 alphabar.
#include "common.hpp"
#include "pair_kernel.hpp"

namespace uvlm {

namespace {

constexpr int WG = 64;

enum { A_UB = 0, A_JB = 3, A_XB = 12, A_AB = 15 };

template <class T> __device__ inline void cross(const T a[3], const T b[3], T o[3])
{
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// W and dS/dr of one pair (see the header); returns false for R^2 = 0 (no interaction).
__device__ inline bool pair_adj(const double r[3], const double al[3], double sg, const double ub[3],
                                const double Jb[9], double W[3], double dS[3])
{
  const double R2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  if (R2 == 0.0) return false;
  const double is2 = 1.0 / (sg * sg);
  const double RS = 0.5 * R2 * is2, s = sqrt(RS);
  const double t2 = 2.0 * INV_SQRTPI * s * exp(-RS);
  const double g  = erf(s) - t2;
  const double q  = 0.25 * INV_PI / (R2 * sqrt(R2));
  const double f1 = g * q;
  const double f2 = (t2 * is2 - 3.0 * g / R2) * q;
  const double f3 = (15.0 * g / (R2 * R2) - t2 * (5.0 + 2.0 * RS) * is2 / R2) * q;
  const double e[3]  = {Jb[5] - Jb[7], Jb[6] - Jb[2], Jb[1] - Jb[3]};
  const double Jr[3] = {Jb[0] * r[0] + Jb[1] * r[1] + Jb[2] * r[2], Jb[3] * r[0] + Jb[4] * r[1] + Jb[5] * r[2],
                        Jb[6] * r[0] + Jb[7] * r[1] + Jb[8] * r[2]};
  double uxr[3], Jrxr[3], rxa[3], axu[3], axJr[3];
  cross(ub, r, uxr);
  cross(Jr, r, Jrxr);
  cross(r, al, rxa);
  cross(al, ub, axu);
  cross(al, Jr, axJr);
  double c1 = 0.0, c3 = 0.0;
  for (int d = 0; d < 3; d++) {
    W[d] = f1 * (uxr[d] + e[d]) + f2 * Jrxr[d];
    c1 += al[d] * (uxr[d] + e[d]);
    c3 += al[d] * Jrxr[d];
  }
  for (int d = 0; d < 3; d++) {
    const double JTrxa = Jb[d] * rxa[0] + Jb[3 + d] * rxa[1] + Jb[6 + d] * rxa[2];
    dS[d]              = -((f2 * c1 + f3 * c3) * r[d] + f1 * axu[d] + f2 * (JTrxa + axJr[d]));
  }
  return true;
}

// Near-field adjoint, one 64-lane workgroup per leaf of the (single) tree.  Like the forward
// P2P kernel, lanes = (own body) x (split s) so small leaves keep the wavefront busy, the
// bodies of all listed leaves stream through LDS in full chunks across leaf boundaries, and
// the splits are combined by cross-lane adds in a fixed order (deterministic, no atomics).
//  SIDE 0 (forward P2P list, own = target i):            xbar_i += sum_j dS_ij/dr
//  SIDE 1 (transposed list, own = source j): alphabar_j -= sum_i W_ij, xbar_j -= sum_i dS_ij/dr
// Sums over the pairs are arranged so that what belongs to the own body is applied once per
// chunk:
//  SIDE 0: sum_j dS_ij = -[X + F1 x ubar + Jbar^T V + A],  X = sum k r, F1 = sum f1 alpha_j,
//          V = sum f2 (r x alpha_j), A = sum f2 alpha_j x (Jbar r), k = f2 c1 + f3 c3;
//  SIDE 1: sum_i dS_ij = -[X + sum f2 Jbar_i^T (r x alpha) + alpha x sum (f1 ubar_i + f2 Jbar_i r)].
// MIXED: FP32 pair arithmetic on coordinates relative to the own leaf centre (formed in FP64
// while staging), f1, f2, f3 from G, G', G'' (pair_kernel.hpp), FP32 partial sums per chunk
// added into FP64 accumulators.  Otherwise FP64 with erf and exp (pair_adj).
template <bool MIXED, int SIDE>
__global__ __launch_bounds__(WG) void k_p2p_adj(const int32_t *__restrict__ leaves, int64_t nleaf,
                                                const int32_t *__restrict__ cbody, const int32_t *__restrict__ cnbody,
                                                const double *__restrict__ ccx, const double *__restrict__ ccy,
                                                const double *__restrict__ ccz, const int32_t *__restrict__ off,
                                                const int32_t *__restrict__ src, const double *__restrict__ st,
                                                double *__restrict__ aux, int64_t cap)
{
  // staged per other body: MIXED SIDE 0: rel x, q | alpha, c1 | c2, c3; SIDE 1: rel x, ubar, e, Jbar
  constexpr int NV = SIDE == 0 ? 3 : 5;   // float4 per slot (MIXED)
  constexpr int ND = SIDE == 0 ? 7 : 15;  // doubles per slot (FP64): x, alpha, sigma | x, ubar, Jbar
  __shared__ float4 s32[MIXED ? NV * WG : 1];
  __shared__ double s64[MIXED ? 1 : ND * WG];
  __shared__ int    pref[SEG + 1], lb[SEG], ln[SEG];
  const int64_t w = blockIdx.x;
  if (w >= nleaf) return;
  const int c = leaves[w], tb = cbody[c], nt = cnbody[c], lane = threadIdx.x;
  // lanes = (own body ti) x (split si), ntp = the leaf's size (at most WG), S = WG / ntp splits
  const int    ntp = near_targets_per_pass<WG>(nt), S = WG / ntp, ti = lane % ntp, si = lane / ntp;
  const bool   lane_on = si < S;
  const double cxt = ccx[c], cyt = ccy[c], czt = ccz[c];
  const int    q0 = off[c], q1 = off[c + 1];
  for (int c0 = 0; c0 < nt; c0 += ntp) {
    const int     i     = c0 + ti;
    const bool    valid = lane_on && i < nt;
    const int     jfirst = lane_on ? si : WG;  // lanes without a split skip the pair loops
    const int64_t bi    = tb + (valid ? i : 0);
    const double  xo[3] = {st[(F_X + 0) * cap + bi], st[(F_X + 1) * cap + bi], st[(F_X + 2) * cap + bi]};
    const float   ro[3] = {(float)(xo[0] - cxt), (float)(xo[1] - cyt), (float)(xo[2] - czt)};
    // own quantities: SIDE 0 ubar, Jbar, e; SIDE 1 alpha, sigma and its kernel constants
    double ubd[3] = {0, 0, 0}, Jbd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, ald[3] = {0, 0, 0}, sgd = 1.0;
    if (SIDE == 0) {
      for (int d = 0; d < 3; d++) ubd[d] = aux[(A_UB + d) * cap + bi];
      for (int k = 0; k < 9; k++) Jbd[k] = aux[(A_JB + k) * cap + bi];
    } else {
      for (int d = 0; d < 3; d++) ald[d] = st[(F_A + d) * cap + bi];
      sgd = st[F_SIG * cap + bi];
    }
    const double oq = 0.5 / (sgd * sgd), osq = sqrt(oq);
    const float  o_q = (float)oq, o_c1 = (float)(oq * osq * INV_4PI), o_c2 = (float)(2.0 * oq * oq * osq * INV_4PI),
                o_c3 = (float)(4.0 * oq * oq * oq * osq * INV_4PI);
    float ubf[3], Jbf[9], ef[3], alf[3];
    for (int d = 0; d < 3; d++) {
      ubf[d] = (float)ubd[d];
      alf[d] = (float)ald[d];
    }
    for (int k = 0; k < 9; k++) Jbf[k] = (float)Jbd[k];
    ef[0] = Jbf[5] - Jbf[7];
    ef[1] = Jbf[6] - Jbf[2];
    ef[2] = Jbf[1] - Jbf[3];
    double ax[3] = {0, 0, 0}, aa[3] = {0, 0, 0};
    for (int g0 = q0; g0 < q1; g0 += SEG) {
      const int nseg = min(SEG, q1 - g0);
      __syncthreads();
      for (int k = lane; k < nseg; k += WG) {
        const int sl = src[g0 + k];
        lb[k]        = cbody[sl];
        ln[k]        = cnbody[sl];
      }
      __syncthreads();
      lds_prefix(pref, ln, nseg);
      __syncthreads();
      const int total = pref[nseg];
      for (int j0 = 0; j0 < total; j0 += WG) {
        const int cnt = min(WG, total - j0);
        if (lane < cnt) {
          const int g  = j0 + lane;
          int       lo = 0, hi = nseg - 1;  // last q with pref[q] <= g
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pref[mid] <= g) lo = mid;
            else hi = mid - 1;
          }
          const int64_t bj = lb[lo] + (g - pref[lo]);
          const double  x = st[(F_X + 0) * cap + bj], y = st[(F_X + 1) * cap + bj], z = st[(F_X + 2) * cap + bj];
          if (SIDE == 0) {
            const double sg = st[F_SIG * cap + bj];
            const double a0 = st[(F_A + 0) * cap + bj], a1 = st[(F_A + 1) * cap + bj], a2 = st[(F_A + 2) * cap + bj];
            if (MIXED) {
              const double qq = 0.5 / (sg * sg), sq = sqrt(qq);
              s32[NV * lane + 0] = make_float4((float)(x - cxt), (float)(y - cyt), (float)(z - czt), (float)qq);
              s32[NV * lane + 1] = make_float4((float)a0, (float)a1, (float)a2, (float)(qq * sq * INV_4PI));
              s32[NV * lane + 2] = make_float4((float)(2.0 * qq * qq * sq * INV_4PI),
                                               (float)(4.0 * qq * qq * qq * sq * INV_4PI), 0.f, 0.f);
            } else {
              const double v[7] = {x, y, z, a0, a1, a2, sg};
              for (int f = 0; f < 7; f++) s64[f * WG + lane] = v[f];
            }
          } else {
            double ub[3], Jb[9];
            for (int d = 0; d < 3; d++) ub[d] = aux[(A_UB + d) * cap + bj];
            for (int k = 0; k < 9; k++) Jb[k] = aux[(A_JB + k) * cap + bj];
            if (MIXED) {
              s32[NV * lane + 0] = make_float4((float)(x - cxt), (float)(y - cyt), (float)(z - czt), (float)ub[0]);
              s32[NV * lane + 1] = make_float4((float)ub[1], (float)ub[2], (float)(Jb[5] - Jb[7]), (float)(Jb[6] - Jb[2]));
              s32[NV * lane + 2] = make_float4((float)(Jb[1] - Jb[3]), (float)Jb[0], (float)Jb[1], (float)Jb[2]);
              s32[NV * lane + 3] = make_float4((float)Jb[3], (float)Jb[4], (float)Jb[5], (float)Jb[6]);
              s32[NV * lane + 4] = make_float4((float)Jb[7], (float)Jb[8], 0.f, 0.f);
            } else {
              s64[0 * WG + lane] = x;
              s64[1 * WG + lane] = y;
              s64[2 * WG + lane] = z;
              for (int d = 0; d < 3; d++) s64[(3 + d) * WG + lane] = ub[d];
              for (int k = 0; k < 9; k++) s64[(6 + k) * WG + lane] = Jb[k];
            }
          }
        }
        __syncthreads();
        if (MIXED) {
          // SIDE 0: X, F1, V, A;  SIDE 1: X, JV, UQ, WA
          float p0[3] = {0, 0, 0}, p1[3] = {0, 0, 0}, p2[3] = {0, 0, 0}, p3[3] = {0, 0, 0};
          for (int j = jfirst; j < cnt; j += S) {
            const float4 A0 = s32[NV * j], A1 = s32[NV * j + 1], A2 = s32[NV * j + 2];
            float        r[3];
            if (SIDE == 0) {
              r[0] = ro[0] - A0.x;
              r[1] = ro[1] - A0.y;
              r[2] = ro[2] - A0.z;
            } else {
              r[0] = A0.x - ro[0];
              r[1] = A0.y - ro[1];
              r[2] = A0.z - ro[2];
            }
            const float r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            float       G, Gp, Gpp;
            kernel_g<true>(r2 * (SIDE == 0 ? A0.w : o_q), G, Gp, Gpp);
            // coincident pairs are skipped (multipoleKernels.hpp:242): every term but the f1 ones
            // carries r, so zeroing f1 is enough
            const float f1 = r2 == 0.0f ? 0.0f : G * (SIDE == 0 ? A1.w : o_c1);
            const float f2 = Gp * (SIDE == 0 ? A2.x : o_c2);
            const float f3 = Gpp * (SIDE == 0 ? A2.y : o_c3);
            if (SIDE == 0) {
              const float al[3] = {A1.x, A1.y, A1.z};
              float       v[3], Jr[3], aJ[3];
              cross(r, al, v);
              for (int d = 0; d < 3; d++) Jr[d] = Jbf[3 * d] * r[0] + Jbf[3 * d + 1] * r[1] + Jbf[3 * d + 2] * r[2];
              cross(al, Jr, aJ);
              const float c1 = ubf[0] * v[0] + ubf[1] * v[1] + ubf[2] * v[2] + al[0] * ef[0] + al[1] * ef[1] +
                               al[2] * ef[2];
              const float c3 = Jr[0] * v[0] + Jr[1] * v[1] + Jr[2] * v[2];
              const float k  = f2 * c1 + f3 * c3;
              for (int d = 0; d < 3; d++) {
                p0[d] = fmaf(k, r[d], p0[d]);
                p1[d] = fmaf(f1, al[d], p1[d]);
                p2[d] = fmaf(f2, v[d], p2[d]);
                p3[d] = fmaf(f2, aJ[d], p3[d]);
              }
            } else {
              const float4 A3 = s32[NV * j + 3], A4 = s32[NV * j + 4];
              const float  ub[3] = {A0.w, A1.x, A1.y}, e[3] = {A1.z, A1.w, A2.x};
              const float  Jb[9] = {A2.y, A2.z, A2.w, A3.x, A3.y, A3.z, A3.w, A4.x, A4.y};
              float        uxr[3], Jr[3], Jrxr[3], v[3];
              cross(ub, r, uxr);
              for (int d = 0; d < 3; d++) Jr[d] = Jb[3 * d] * r[0] + Jb[3 * d + 1] * r[1] + Jb[3 * d + 2] * r[2];
              cross(Jr, r, Jrxr);
              cross(r, alf, v);
              float c1 = 0.f, c3 = 0.f;
              for (int d = 0; d < 3; d++) {
                c1 = fmaf(alf[d], uxr[d] + e[d], c1);
                c3 = fmaf(alf[d], Jrxr[d], c3);
              }
              const float k = f2 * c1 + f3 * c3;
              for (int d = 0; d < 3; d++) {
                const float JTv = Jb[d] * v[0] + Jb[3 + d] * v[1] + Jb[6 + d] * v[2];
                p0[d]           = fmaf(k, r[d], p0[d]);
                p1[d]           = fmaf(f2, JTv, p1[d]);
                p2[d]           = fmaf(f1, ub[d], fmaf(f2, Jr[d], p2[d]));
                p3[d]           = fmaf(f1, uxr[d] + e[d], fmaf(f2, Jrxr[d], p3[d]));
              }
            }
          }
          if (SIDE == 0) {
            // sum dS = -[X + F1 x ubar + Jbar^T V + A]
            float fxu[3];
            cross(p1, ubf, fxu);
            for (int d = 0; d < 3; d++) {
              const float JTV = Jbf[d] * p2[0] + Jbf[3 + d] * p2[1] + Jbf[6 + d] * p2[2];
              ax[d] -= (double)(p0[d] + fxu[d] + JTV + p3[d]);
            }
          } else {
            // xbar_j -= sum dS = X + JV + alpha x UQ;  alphabar_j -= sum W
            float axq[3];
            cross(alf, p2, axq);
            for (int d = 0; d < 3; d++) {
              ax[d] += (double)(p0[d] + p1[d] + axq[d]);
              aa[d] -= (double)p3[d];
            }
          }
        } else {
          for (int j = jfirst; j < cnt; j += S) {
            double r[3], Wv[3], dS[3];
            if (SIDE == 0) {
              const double alj[3] = {s64[3 * WG + j], s64[4 * WG + j], s64[5 * WG + j]};
              for (int d = 0; d < 3; d++) r[d] = xo[d] - s64[d * WG + j];
              if (pair_adj(r, alj, s64[6 * WG + j], ubd, Jbd, Wv, dS))
                for (int d = 0; d < 3; d++) ax[d] += dS[d];
            } else {
              const double ubi[3] = {s64[3 * WG + j], s64[4 * WG + j], s64[5 * WG + j]};
              double       Jbi[9];
              for (int k = 0; k < 9; k++) Jbi[k] = s64[(6 + k) * WG + j];
              for (int d = 0; d < 3; d++) r[d] = s64[d * WG + j] - xo[d];
              if (pair_adj(r, ald, sgd, ubi, Jbi, Wv, dS))
                for (int d = 0; d < 3; d++) {
                  ax[d] -= dS[d];
                  aa[d] -= Wv[d];
                }
            }
          }
        }
        __syncthreads();
      }
    }
    // combine the splits (lanes ti + q * ntp) in split order
    for (int d = 0; d < 3; d++) {
      double vx = ax[d], va = aa[d];
      for (int q = 1; q < S; q++) {
        vx += __shfl(ax[d], ti + q * ntp, 64);
        if (SIDE == 1) va += __shfl(aa[d], ti + q * ntp, 64);
      }
      ax[d] = vx;
      aa[d] = va;
    }
    if (si == 0 && valid) {
      for (int d = 0; d < 3; d++) aux[(A_XB + d) * cap + bi] += ax[d];
      if (SIDE == 1)
        for (int d = 0; d < 3; d++) aux[(A_AB + d) * cap + bi] += aa[d];
    }
  }
}

// Jbar from alpha and abar (abar uploaded into the alphabar fields), the local stretching term
// of alphabar (J abar or J^T abar with the forward J), xbar cleared.
__global__ void k_vjp_setup(int64_t N, const double *__restrict__ st, double *__restrict__ aux, int64_t cap,
                            int classic)
{
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= N) return;
  double al[3], ab[3], J[9];
  for (int d = 0; d < 3; d++) {
    al[d] = st[(F_A + d) * cap + i];
    ab[d] = aux[(A_AB + d) * cap + i];
  }
  for (int k = 0; k < 9; k++) J[k] = st[(F_J + k) * cap + i];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) aux[(A_JB + 3 * a + b) * cap + i] = classic ? ab[a] * al[b] : al[a] * ab[b];
  for (int a = 0; a < 3; a++) {
    double v = 0.0;
    for (int b = 0; b < 3; b++) v += classic ? J[3 * b + a] * ab[b] : J[3 * a + b] * ab[b];
    aux[(A_AB + a) * cap + i] = v;
    aux[(A_XB + a) * cap + i] = 0.0;
  }
}

// Far source-side terms from the adjoint L2P (ast: velocity = F, dadt = (grad F)^T alpha).
__global__ void k_vjp_finish(int64_t N, const double *__restrict__ ast, double *__restrict__ aux, int64_t cap)
{
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= N) return;
  for (int d = 0; d < 3; d++) {
    aux[(A_XB + d) * cap + i] += ast[(F_DADT + d) * cap + i];
    aux[(A_AB + d) * cap + i] += ast[(F_U + d) * cap + i];
  }
}

// Cotangents of the state (X, alpha) live in the original body order: 6 SoA fields of `ccap`
// doubles (0-2 position, 3-5 strength).  orig[p] is the original index of device body p.
__global__ void k_cot_gather(int64_t N, const int32_t *__restrict__ orig, const double *__restrict__ cot,
                             int64_t ccap, double *__restrict__ aux, int64_t cap)
{
  const int64_t p = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (p >= N) return;
  const int64_t o = orig[p];
  for (int d = 0; d < 3; d++) {
    aux[(A_UB + d) * cap + p] = cot[d * ccap + o];
    aux[(A_AB + d) * cap + p] = cot[(3 + d) * ccap + o];
  }
}

__global__ void k_cot_scatter(int64_t N, const int32_t *__restrict__ orig, const double *__restrict__ aux,
                              int64_t cap, double *__restrict__ cot, int64_t ccap)
{
  const int64_t p = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (p >= N) return;
  const int64_t o = orig[p];
  for (int d = 0; d < 3; d++) {
    cot[d * ccap + o]       = aux[(A_XB + d) * cap + p];
    cot[(3 + d) * ccap + o] = aux[(A_AB + d) * cap + p];
  }
}

__global__ void k_state_to_orig(int64_t N, const int32_t *__restrict__ orig, const double *__restrict__ st,
                                int64_t cap, double *__restrict__ dst)
{
  const int64_t p = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (p >= N) return;
  const int64_t o = orig[p];
  for (int f = 0; f < 6; f++) dst[f * N + o] = st[(F_X + f) * cap + p];
  dst[6 * N + o] = st[F_SIG * cap + p];
}

struct LinTerms {
  const double *p[6];
  double        c[6];
  int           n;
};

__global__ void k_lincomb(int64_t n, double *__restrict__ out, LinTerms t)
{
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= n) return;
  double v = 0.0;
  for (int k = 0; k < t.n; k++) v += t.c[k] * t.p[k][i];
  out[i] = v;
}

}  // namespace

void cot_gather(hipStream_t s, int64_t N, const int32_t *orig, const double *cot, int64_t ccap, double *aux,
                int64_t cap)
{
  if (N <= 0) return;
  k_cot_gather<<<(int)((N + 255) / 256), 256, 0, s>>>(N, orig, cot, ccap, aux, cap);
  UVLM_LAUNCH_CHECK();
}

void cot_scatter(hipStream_t s, int64_t N, const int32_t *orig, const double *aux, int64_t cap, double *cot,
                 int64_t ccap)
{
  if (N <= 0) return;
  k_cot_scatter<<<(int)((N + 255) / 256), 256, 0, s>>>(N, orig, aux, cap, cot, ccap);
  UVLM_LAUNCH_CHECK();
}

void state_to_orig(hipStream_t s, int64_t N, const int32_t *orig, const double *st, int64_t cap, double *dst)
{
  if (N <= 0) return;
  k_state_to_orig<<<(int)((N + 255) / 256), 256, 0, s>>>(N, orig, st, cap, dst);
  UVLM_LAUNCH_CHECK();
}

void lincomb(hipStream_t s, int64_t n, double *out, int nterms, const double *const *p, const double *c)
{
  if (n <= 0) return;
  if (nterms > 6) throw Error(UVLM_ERR_INTERNAL, "lincomb: more than 6 terms");
  LinTerms t{};
  t.n = nterms;
  for (int k = 0; k < nterms; k++) {
    t.p[k] = p[k];
    t.c[k] = c[k];
  }
  k_lincomb<<<(int)((n + 255) / 256), 256, 0, s>>>(n, out, t);
  UVLM_LAUNCH_CHECK();
}

void vjp_setup(hipStream_t s, int64_t N, const double *st, double *aux, int64_t cap, int classic)
{
  if (N <= 0) return;
  k_vjp_setup<<<(int)((N + 255) / 256), 256, 0, s>>>(N, st, aux, cap, classic);
  UVLM_LAUNCH_CHECK();
}

void vjp_finish(hipStream_t s, int64_t N, const double *ast, double *aux, int64_t cap)
{
  if (N <= 0) return;
  k_vjp_finish<<<(int)((N + 255) / 256), 256, 0, s>>>(N, ast, aux, cap);
  UVLM_LAUNCH_CHECK();
}

// Cotangent of the sources' core sizes, sbar_j = dS/dsigma_j (device order of the sources), of
// S = sum over the near (P2P) pairs of ubar_i . u_i + Jbar_i : J_i: only the near field depends
// on sigma (the far field applies the singular kernel).  Per pair, r = x_i - x_j and the kernel of
// sigma_j, S_ij = -alpha_j . (f1 (ubar_i x r + e_i) + f2 (Jbar_i r) x r), e the axial vector of
// Jbar, so dS_ij/dsigma_j = -alpha_j . (f1s (ubar_i x r + e_i) + f2s (Jbar_i r) x r) with
// f1s = df1/dsigma = -2 s^2 t2 / (sigma 4 pi R^3), f2s = df2/dsigma = 2 s^2 t2 / (sigma^3 4 pi R^3),
// s^2 = R^2 / (2 sigma^2), t2 = 2 s e^{-s^2} / sqrt(pi).  FP64.  One wavefront per source leaf,
// lanes over its bodies; `off`/`ent` is CSR by source cell with target cells as entries (the
// P2P list transposed); the target bodies pass through LDS, 64 at a time, in list order.
template <bool WITH_JB>
__global__ __launch_bounds__(WG) void k_sigma_bar(const int32_t *__restrict__ leaves, int64_t nleaf,
                                                  const int32_t *__restrict__ sbody, const int32_t *__restrict__ snb,
                                                  const int32_t *__restrict__ off, const int32_t *__restrict__ ent,
                                                  const int32_t *__restrict__ tbody, const int32_t *__restrict__ tnb,
                                                  const double *__restrict__ tx, int64_t tcap,
                                                  const double *__restrict__ tcot, int64_t ccap,
                                                  const double *__restrict__ sst, int64_t scap,
                                                  double *__restrict__ sbar)
{
  __shared__ double sh[15][WG];
  const int64_t w = blockIdx.x;
  if (w >= nleaf) return;
  const int lane = threadIdx.x;
  const int sl = leaves[w], sb0 = sbody[sl], ns = snb[sl];
  for (int p0 = 0; p0 < ns; p0 += WG) {
    const int  j  = sb0 + p0 + lane;
    const bool ok = p0 + lane < ns;
    double     xj[3] = {0, 0, 0}, aj[3] = {0, 0, 0}, sg = 1.0;
    if (ok) {
      for (int d = 0; d < 3; d++) {
        xj[d] = sst[(F_X + d) * scap + j];
        aj[d] = sst[(F_A + d) * scap + j];
      }
      sg = sst[F_SIG * scap + j];
    }
    const double S2 = sg * sg;
    double       acc = 0.0;
    for (int q = off[sl]; q < off[sl + 1]; q++) {
      const int t = ent[q], tb0 = tbody[t], nt = tnb[t];
      for (int c0 = 0; c0 < nt; c0 += WG) {
        const int cnt = min(WG, nt - c0);
        __syncthreads();
        if (lane < cnt) {
          const int64_t i = tb0 + c0 + lane;
          for (int d = 0; d < 3; d++) {
            sh[d][lane]     = tx[(F_X + d) * tcap + i];
            sh[3 + d][lane] = tcot[d * ccap + i];
          }
          if (WITH_JB)
            for (int k = 0; k < 9; k++) sh[6 + k][lane] = tcot[(3 + k) * ccap + i];
        }
        __syncthreads();
        if (!ok) continue;
        for (int e = 0; e < cnt; e++) {
          const double r[3] = {sh[0][e] - xj[0], sh[1][e] - xj[1], sh[2][e] - xj[2]};
          const double R2   = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
          if (R2 == 0.0) continue;
          const double s2 = R2 / S2 * 0.5, t2 = 2.0 * INV_SQRTPI * sqrt(s2) * exp(-s2);
          const double q4 = 0.25 * INV_PI / (R2 * sqrt(R2));
          const double f1s = -2.0 * s2 * t2 * q4 / sg, f2s = 2.0 * s2 * t2 * q4 / (S2 * sg);
          const double ub[3] = {sh[3][e], sh[4][e], sh[5][e]};
          double       W[3]  = {ub[1] * r[2] - ub[2] * r[1], ub[2] * r[0] - ub[0] * r[2], ub[0] * r[1] - ub[1] * r[0]};
          for (int d = 0; d < 3; d++) W[d] *= f1s;
          if (WITH_JB) {
            const double *M = &sh[6][0];
            auto          Jb = [&](int a, int b) { return M[(3 * a + b) * WG + e]; };
            const double  ea[3] = {Jb(1, 2) - Jb(2, 1), Jb(2, 0) - Jb(0, 2), Jb(0, 1) - Jb(1, 0)};
            double        Jr[3];
            for (int a = 0; a < 3; a++) Jr[a] = Jb(a, 0) * r[0] + Jb(a, 1) * r[1] + Jb(a, 2) * r[2];
            const double Jrxr[3] = {Jr[1] * r[2] - Jr[2] * r[1], Jr[2] * r[0] - Jr[0] * r[2],
                                    Jr[0] * r[1] - Jr[1] * r[0]};
            for (int d = 0; d < 3; d++) W[d] += f1s * ea[d] + f2s * Jrxr[d];
          }
          acc -= aj[0] * W[0] + aj[1] * W[1] + aj[2] * W[2];
        }
      }
    }
    if (ok) sbar[j] = acc;
  }
}

__global__ void k_add_to_orig(int64_t N, const int32_t *__restrict__ orig, const double *__restrict__ src,
                              double *__restrict__ dst)
{
  const int64_t p = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (p >= N) return;
  dst[orig[p]] += src[p];
}

void add_to_orig(hipStream_t s, int64_t N, const int32_t *orig, const double *src, double *dst)
{
  if (N <= 0) return;
  k_add_to_orig<<<(int)((N + 255) / 256), 256, 0, s>>>(N, orig, src, dst);
  UVLM_LAUNCH_CHECK();
}

void sigma_bar(hipStream_t s, const Tree &stree, const Tree &ttree, const Lists &tl, const double *tx, int64_t tcap,
               const double *tcot, int64_t ccap, bool with_jbar, const double *sst, int64_t scap, double *sbar)
{
  if (stree.nleaf == 0) return;
  const int32_t *lv = stree.leaves.as<int32_t>(), *sb = stree.body.as<int32_t>(), *sn = stree.nbody.as<int32_t>();
  const int32_t *off = tl.p2p_off.as<int32_t>(), *ent = tl.p2p_src.as<int32_t>();
  const int32_t *tb = ttree.body.as<int32_t>(), *tn = ttree.nbody.as<int32_t>();
  const int      nb = (int)stree.nleaf;
  if (with_jbar)
    k_sigma_bar<true><<<nb, WG, 0, s>>>(lv, stree.nleaf, sb, sn, off, ent, tb, tn, tx, tcap, tcot, ccap, sst, scap, sbar);
  else
    k_sigma_bar<false><<<nb, WG, 0, s>>>(lv, stree.nleaf, sb, sn, off, ent, tb, tn, tx, tcap, tcot, ccap, sst, scap,
                                         sbar);
  UVLM_LAUNCH_CHECK();
}

void p2p_adj(hipStream_t s, const Tree &t, const Lists &l, int side, int mixed, const double *st, double *aux,
             int64_t cap)
{
  if (t.nleaf == 0) return;
  const int32_t *lv = t.leaves.as<int32_t>(), *cb = t.body.as<int32_t>(), *cn = t.nbody.as<int32_t>();
  const int32_t *off = l.p2p_off.as<int32_t>(), *src = l.p2p_src.as<int32_t>();
  const double  *cx = t.cx.as<double>(), *cy = t.cy.as<double>(), *cz = t.cz.as<double>();
  const int      nb = (int)t.nleaf;
  if (mixed) {
    if (side == 0) k_p2p_adj<true, 0><<<nb, WG, 0, s>>>(lv, t.nleaf, cb, cn, cx, cy, cz, off, src, st, aux, cap);
    else k_p2p_adj<true, 1><<<nb, WG, 0, s>>>(lv, t.nleaf, cb, cn, cx, cy, cz, off, src, st, aux, cap);
  } else {
    if (side == 0) k_p2p_adj<false, 0><<<nb, WG, 0, s>>>(lv, t.nleaf, cb, cn, cx, cy, cz, off, src, st, aux, cap);
    else k_p2p_adj<false, 1><<<nb, WG, 0, s>>>(lv, t.nleaf, cb, cn, cx, cy, cz, off, src, st, aux, cap);
  }
  UVLM_LAUNCH_CHECK();
}

}  // namespace uvlm
