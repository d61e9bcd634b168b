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

namespace uvlm {

namespace {

constexpr int WG = 64;

enum { A_UB = 0, A_JB = 3, A_XB = 12, A_AB = 15 };

__device__ inline void cross(const double a[3], const double b[3], double o[3])
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

// One 64-lane workgroup per leaf of the (single) tree; each lane owns one body of the leaf
// per chunk, the bodies of the listed leaves stream through LDS in list order (deterministic).
//  SIDE 0 (forward P2P list, own = target i):            xbar_i += sum_j dS_ij/dr
//  SIDE 1 (transposed list, own = source j): alphabar_j -= sum_i W_ij, xbar_j -= sum_i dS_ij/dr
template <int SIDE>
__global__ __launch_bounds__(WG) void k_p2p_adj(const int32_t *__restrict__ leaves, int64_t nleaf,
                                                const int32_t *__restrict__ cbody, const int32_t *__restrict__ cnbody,
                                                const int32_t *__restrict__ off, const int32_t *__restrict__ src,
                                                const double *__restrict__ st, double *__restrict__ aux, int64_t cap)
{
  constexpr int NF = SIDE == 0 ? 7 : 15;  // staged fields: x, alpha, sigma | x, ubar, Jbar
  __shared__ double sh[NF][WG];
  const int64_t w = blockIdx.x;
  if (w >= nleaf) return;
  const int c = leaves[w], tb = cbody[c], nt = cnbody[c], lane = threadIdx.x;
  const int q0 = off[c], q1 = off[c + 1];
  for (int c0 = 0; c0 < nt; c0 += WG) {
    const int     i     = c0 + lane;
    const bool    valid = i < nt;
    const int64_t bi    = tb + (valid ? i : 0);
    double        x[3], al[3] = {0, 0, 0}, sg = 1.0, ub[3] = {0, 0, 0}, Jb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int d = 0; d < 3; d++) x[d] = st[(F_X + d) * cap + bi];
    if (SIDE == 0) {
      for (int d = 0; d < 3; d++) ub[d] = aux[(A_UB + d) * cap + bi];
      for (int k = 0; k < 9; k++) Jb[k] = aux[(A_JB + k) * cap + bi];
    } else {
      for (int d = 0; d < 3; d++) al[d] = st[(F_A + d) * cap + bi];
      sg = st[F_SIG * cap + bi];
    }
    double ax[3] = {0, 0, 0}, aa[3] = {0, 0, 0};
    for (int q = q0; q < q1; q++) {
      const int sl = src[q], b0 = cbody[sl], nb = cnbody[sl];
      for (int j0 = 0; j0 < nb; j0 += WG) {
        const int cnt = min(WG, nb - j0);
        __syncthreads();
        if (lane < cnt) {
          const int64_t bj = b0 + j0 + lane;
          for (int d = 0; d < 3; d++) sh[d][lane] = st[(F_X + d) * cap + bj];
          if (SIDE == 0) {
            for (int d = 0; d < 3; d++) sh[3 + d][lane] = st[(F_A + d) * cap + bj];
            sh[6][lane] = st[F_SIG * cap + bj];
          } else {
            for (int f = 0; f < 12; f++) sh[3 + f][lane] = aux[(A_UB + f) * cap + bj];
          }
        }
        __syncthreads();
        if (!valid) continue;
        for (int j = 0; j < cnt; j++) {
          double r[3], W[3], dS[3];
          if (SIDE == 0) {
            const double alj[3] = {sh[3][j], sh[4][j], sh[5][j]};
            for (int d = 0; d < 3; d++) r[d] = x[d] - sh[d][j];
            if (pair_adj(r, alj, sh[6][j], ub, Jb, W, dS))
              for (int d = 0; d < 3; d++) ax[d] += dS[d];
          } else {
            const double ubi[3] = {sh[3][j], sh[4][j], sh[5][j]};
            double       Jbi[9];
            for (int k = 0; k < 9; k++) Jbi[k] = sh[6 + k][j];
            for (int d = 0; d < 3; d++) r[d] = sh[d][j] - x[d];
            if (pair_adj(r, al, sg, ubi, Jbi, W, dS))
              for (int d = 0; d < 3; d++) {
                ax[d] -= dS[d];
                aa[d] -= W[d];
              }
          }
        }
      }
    }
    if (valid) {
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

void p2p_adj(hipStream_t s, const Tree &t, const Lists &l, int side, const double *st, double *aux, int64_t cap)
{
  if (t.nleaf == 0) return;
  const int32_t *lv = t.leaves.as<int32_t>(), *cb = t.body.as<int32_t>(), *cn = t.nbody.as<int32_t>();
  const int32_t *off = l.p2p_off.as<int32_t>(), *src = l.p2p_src.as<int32_t>();
  if (side == 0)
    k_p2p_adj<0><<<(int)t.nleaf, WG, 0, s>>>(lv, t.nleaf, cb, cn, off, src, st, aux, cap);
  else
    k_p2p_adj<1><<<(int)t.nleaf, WG, 0, s>>>(lv, t.nleaf, cb, cn, off, src, st, aux, cap);
  UVLM_LAUNCH_CHECK();
}

}  // namespace uvlm
