// near.hip -- near-field P2P: regularised Biot-Savart velocity and velocity gradient.
//
// Kernel of src/multipoleKernels.hpp:223-301 (sensor form 303-341): for target i and
// source j with dX = x_i - x_j, R^2 != 0, s = R / (sqrt(2) sigma_j):
//   g  = erf(s) - (2/sqrt(pi)) s e^{-s^2},  f1 = g / (4 pi R^3),
//   f2 = (2 s^2 t2 - 3 g) / (4 pi R^5)      (t2 = (2/sqrt(pi)) s e^{-s^2})
//   u -= f1 (dX x a_j),   J[a][b] -= f2 (dX x a_j)_a dX_b + f1 eps_{abk} a_j,k.
// Only the source core size sigma_j enters (multipoleKernels.hpp:243).
//
// Layout: one workgroup of BLK lanes per target leaf (64, or 256 when leaves may hold more than
// 64 bodies).  Lanes = (target t) x (source split s): a pass takes ntp of the leaf's targets
// (near_targets_per_pass) and S = BLK / ntp splits, each taking every S-th source, and the
// partial sums are combined in a fixed order (cross-lane adds or LDS, no atomics).  The bodies
// of all source leaves in the target's CSR list are streamed through LDS in full chunks of BLK
// across leaf boundaries (prefix sums of the leaf sizes in LDS, binary search per slot), so
// small leaves do not pay a barrier per leaf.
//
// FP64 path: the reference formula (erf, exp, sqrt in double).
// Mixed path: FP32 arithmetic on coordinates relative to the target leaf centre (the source
// offsets are formed in FP64 while staging), two sources per step over zero-padded chunks (one
// trip count for all lanes), per-chunk FP32 partial sums added into FP64 per-target
// accumulators; G = g/s^3 and its derivative from pair_kernel.hpp.
#include "common.hpp"
#include "pair_kernel.hpp"

#include <vector>

#include <algorithm>

namespace uvlm {

namespace {

constexpr int WG = 64;

struct SrcF64 {
  double x, y, z, ax, ay, az, invS2, halfInvS2;
};

// acc[0..2] += f1 c, acc[3..10] += f2 c (x) dX with c = dX x a, and acc[12..14] += f1 a: the
// f1 eps_{abk} a_k part of J is antisymmetric, so its sum is kept as one vector and folded
// into J once per chunk (fold_antisym) instead of six multiply-adds per pair.  c is
// perpendicular to dX, so c (x) dX is traceless and its zz entry (acc[11]) is formed from the
// xx and yy sums in fold_antisym.
__device__ __forceinline__ void accumulate_pair(float *acc, float dx, float dy, float dz, float ax, float ay, float az,
                                                float f1, float f2, bool withJ)
{
  const float cx = dy * az - dz * ay, cy = dz * ax - dx * az, cz = dx * ay - dy * ax;
  acc[0] = fmaf(f1, cx, acc[0]);
  acc[1] = fmaf(f1, cy, acc[1]);
  acc[2] = fmaf(f1, cz, acc[2]);
  if (withJ) {
    const float k0 = f2 * cx, k1 = f2 * cy, k2 = f2 * cz;
    acc[3]  = fmaf(k0, dx, acc[3]);
    acc[4]  = fmaf(k0, dy, acc[4]);
    acc[5]  = fmaf(k0, dz, acc[5]);
    acc[6]  = fmaf(k1, dx, acc[6]);
    acc[7]  = fmaf(k1, dy, acc[7]);
    acc[8]  = fmaf(k1, dz, acc[8]);
    acc[9]  = fmaf(k2, dx, acc[9]);
    acc[10] = fmaf(k2, dy, acc[10]);
    acc[12] = fmaf(f1, ax, acc[12]);
    acc[13] = fmaf(f1, ay, acc[13]);
    acc[14] = fmaf(f1, az, acc[14]);
  }
}
// J[2][2] = -(J[0][0] + J[1][1]) and J[a][b] += eps_{abk} F_k with F = acc[12..14] (row-major J
// in acc[3..11]).
__device__ __forceinline__ void fold_antisym(float *acc)
{
  acc[11] = -(acc[3] + acc[7]);
  acc[4] += acc[14];
  acc[5] -= acc[13];
  acc[6] -= acc[14];
  acc[8] += acc[12];
  acc[9] += acc[13];
  acc[10] -= acc[12];
}

// TSIG: the core size comes from the TARGET body instead of the source.  That is the
// transposed near-field kernel used by the adjoint (a particle's own sigma, with the
// particle now on the target side).
// SPLIT: workgroup b takes the list entries [seg[3b+1], seg[3b+2]) of target leaf seg[3b] and
// writes its un-negated sums to part + b * 12 * maxnt (field-major, maxnt entries per field);
// k_p2p_reduce adds the segments of each target in order.
template <bool MIXED, bool WITH_J, bool TSIG, int BLK, bool SPLIT = false>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(MIXED ? 5 : 1))) void k_p2p(const int32_t *__restrict__ leaves, int64_t nleaf,
                                            const int32_t *__restrict__ tbody, const int32_t *__restrict__ tnbody,
                                            const double *__restrict__ tcx, const double *__restrict__ tcy,
                                            const double *__restrict__ tcz, const int32_t *__restrict__ off,
                                            const int32_t *__restrict__ src, const int32_t *__restrict__ sbody,
                                            const int32_t *__restrict__ snbody, const double *__restrict__ tst,
                                            int64_t tcap, const double *__restrict__ sst, int64_t scap,
                                            double *__restrict__ out, int64_t ocap,
                                            const int32_t *__restrict__ seg = nullptr, double *__restrict__ part = nullptr,
                                            int maxnt = 0)
{
  // One LDS buffer holds the staged sources during the pair loops and the split partial sums
  // after them: mixed sources take two float4 per entry for 2 BLK entries (BLK staged, the rest
  // zero padding), FP64 sources one SrcF64 per entry, the partials 12 doubles per lane.
  constexpr int SMEM = BLK == 64 ? 64 * BLK : 96 * BLK;
  __shared__ __align__(16) unsigned char smem[SMEM];
  __shared__ int pref[SEG + 1], lb[SEG], ln[SEG];
  float4 *s32 = reinterpret_cast<float4 *>(smem);
  SrcF64 *s64 = reinterpret_cast<SrcF64 *>(smem);
  double *red = reinterpret_cast<double *>(smem);
  const int64_t w = SPLIT ? seg[3 * blockIdx.x] : blockIdx.x;
  if (w >= nleaf) return;
  const int t = leaves[w], tb = tbody[t], nt = tnbody[t];
  const int lane = threadIdx.x;
  // lanes = (target ti) x (source split si): ntp targets per pass (near_targets_per_pass), S = BLK /
  // ntp splits; the BLK - ntp * S lanes left over take no sources
  const int    ntp = near_targets_per_pass<BLK>(nt), S = BLK / ntp, ti = lane % ntp, si = lane / ntp;
  const bool   lane_on = si < S;
  const double cxt = tcx[t], cyt = tcy[t], czt = tcz[t];
  const int    q0 = SPLIT ? seg[3 * blockIdx.x + 1] : off[t], q1 = SPLIT ? seg[3 * blockIdx.x + 2] : off[t + 1];
  for (int c0 = 0; c0 < nt; c0 += ntp) {
    const int  i     = c0 + ti;
    const bool valid = lane_on && i < nt;
    const int64_t bi = tb + (valid ? i : 0);
    const double  xi = tst[(F_X + 0) * tcap + bi], yi = tst[(F_X + 1) * tcap + bi], zi = tst[(F_X + 2) * tcap + bi];
    const float   rxi = (float)(xi - cxt), ryi = (float)(yi - cyt), rzi = (float)(zi - czt);
    const int     jfirst = lane_on ? si : BLK;  // lanes without a split skip the FP64 pair loop
    // target-sigma constants (TSIG only)
    const double  tsg = TSIG ? tst[F_SIG * tcap + bi] : 1.0;
    const double  tqq = 0.5 / (tsg * tsg), tsq = sqrt(tqq);
    const float   t_q = (float)tqq, t_c1 = (float)(tqq * tsq * INV_4PI), t_c2 = (float)(2.0 * tqq * tqq * tsq * INV_4PI);
    const double  t_is2 = 1.0 / (tsg * tsg), t_half = 0.5 * t_is2;
    double        acc[12];
    for (int k = 0; k < 12; k++) acc[k] = 0.0;
    for (int g0 = q0; g0 < q1; g0 += SEG) {
      const int nseg = min(SEG, q1 - g0);
      __syncthreads();
      for (int k = lane; k < nseg; k += BLK) {
        const int sl = src[g0 + k];
        lb[k]        = sbody[sl];
        ln[k]        = snbody[sl];
      }
      __syncthreads();
      if (lane < 64) lds_prefix(pref, ln, nseg);
      __syncthreads();
      const int total = pref[nseg];
      for (int j0 = 0; j0 < total; j0 += BLK) {
        const int cnt = min(BLK, total - j0);
        if (lane < cnt) {
          const int g  = j0 + lane;
          int       lo = 0, hi = nseg - 1;  // last q with pref[q] <= g
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pref[mid] <= g) lo = mid;
            else hi = mid - 1;
          }
          const int64_t bj = lb[lo] + (g - pref[lo]);
          const double  x = sst[(F_X + 0) * scap + bj], y = sst[(F_X + 1) * scap + bj], z = sst[(F_X + 2) * scap + bj];
          const double  sg = sst[F_SIG * scap + bj];
          const double  ax = sst[(F_A + 0) * scap + bj], ay = sst[(F_A + 1) * scap + bj], az = sst[(F_A + 2) * scap + bj];
          if (MIXED) {
            // (offset, q) and (alpha c1, 2 q) with q = 1 / (2 sigma^2), c1 = q^{3/2} / (4 pi): then
            // f1 = G(x) and f2 = 2 q G'(x) for the scaled strength.  TSIG keeps alpha as is.
            const double qq = 0.5 / (sg * sg), c1 = TSIG ? 1.0 : qq * sqrt(qq) * INV_4PI;
            s32[2 * lane + 0] = make_float4((float)(x - cxt), (float)(y - cyt), (float)(z - czt), (float)qq);
            s32[2 * lane + 1] = make_float4((float)(ax * c1), (float)(ay * c1), (float)(az * c1), (float)(2.0 * qq));
          } else {
            const double is2 = 1.0 / (sg * sg);
            s64[lane]        = SrcF64{x, y, z, ax, ay, az, is2, 0.5 * is2};
          }
        }
        if (MIXED) {
          // entries of no strength past the chunk: every lane runs the same trip count.  Their
          // offset and q put x = |r - (1024, 0, 0)|^2 / 2^20 near 1 for any target within a few
          // hundred units of the leaf centre, so they stay on the pair loop's common path (a pad
          // at x = 0 would take its coincident-pair test); TSIG takes q from the target, where
          // a pad at the centre gives x = |r|^2 q of the same order as the real pairs.
          const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
          const float4 pad  = TSIG ? zero : make_float4(1024.f, 0.f, 0.f, 0x1p-20f);
          if (lane >= cnt) s32[2 * lane] = pad, s32[2 * lane + 1] = zero;
          s32[2 * (BLK + lane)] = pad, s32[2 * (BLK + lane) + 1] = zero;
        }
        __syncthreads();
        if (MIXED) {
          // lane (ti, si) takes entries si, si + S, ... two per step; T2 * S <= 2 BLK entries
          // are read (S <= BLK), those from cnt on are the zero padding
          float        fa[15];
          const int    T2 = 2 * ((cnt + 2 * S - 1) / (2 * S));
          const float4 *ps = s32 + 2 * (lane_on ? si : 0);
          // FP32 partial sums cover at most FLUSH entries per lane before they go into the FP64 sums
          constexpr int FLUSH = 64;
          for (int k0 = 0; k0 < T2; k0 += FLUSH) {
            for (int k = 0; k < 15; k++) fa[k] = 0.f;
            for (int k = k0; k < min(T2, k0 + FLUSH); k += 2, ps += 4 * S) {
              const float4 A0 = ps[0], B0 = ps[1], A1 = ps[2 * S], B1 = ps[2 * S + 1];
              const float  dx0 = rxi - A0.x, dy0 = ryi - A0.y, dz0 = rzi - A0.z;
              const float  dx1 = rxi - A1.x, dy1 = ryi - A1.y, dz1 = rzi - A1.z;
              const float  r20 = dx0 * dx0 + dy0 * dy0 + dz0 * dz0, r21 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
              const float  q0 = TSIG ? t_q : A0.w, q1 = TSIG ? t_q : A1.w;
              const float  tq0 = fmaf(r20, q0, -1.0f), tq1 = fmaf(r21, q1, -1.0f);  // x - 1
              float        G0, Gp0, G1, Gp1;  // G(x) and G'(x) of both pairs
              g_near(tq0, tq1, G0, Gp0, G1, Gp1);
              // x >= 2 (t >= 1) takes the far form, and a coincident pair (R^2 = 0, so t = -1
              // exactly) is skipped (multipoleKernels.hpp:242).  t >= -1 always, so one test
              // |t| >= 1 over both pairs sends the rare steps that need either fix-up off the
              // common path.
              if (fmaxf(fabsf(tq0), fabsf(tq1)) >= XPOLY - 1.0f) {
                if (fmaxf(tq0, tq1) >= XPOLY - 1.0f) {  // lanes with a far pair take both far forms
                  float Gf0, Gpf0, Gf1, Gpf1;
                  g_far(r20 * q0, Gf0, Gpf0);
                  g_far(r21 * q1, Gf1, Gpf1);
                  if (tq0 >= XPOLY - 1.0f) G0 = Gf0, Gp0 = Gpf0;
                  if (tq1 >= XPOLY - 1.0f) G1 = Gf1, Gp1 = Gpf1;
                }
                // only the f1 alpha terms of J survive a zero dX, so zeroing f1 is enough
                if (r20 == 0.0f) G0 = 0.0f;
                if (r21 == 0.0f) G1 = 0.0f;
              }
              const float f10 = TSIG ? G0 * t_c1 : G0;
              const float f11 = TSIG ? G1 * t_c1 : G1;
              accumulate_pair(fa, dx0, dy0, dz0, B0.x, B0.y, B0.z, f10, Gp0 * (TSIG ? t_c2 : B0.w), WITH_J);
              accumulate_pair(fa, dx1, dy1, dz1, B1.x, B1.y, B1.z, f11, Gp1 * (TSIG ? t_c2 : B1.w), WITH_J);
            }
            if (WITH_J) fold_antisym(fa);
            for (int k = 0; k < (WITH_J ? 12 : 3); k++) acc[k] += (double)fa[k];
          }
        } else {
          for (int j = jfirst; j < cnt; j += S) {
            const SrcF64 &sj = s64[j];
            const double  dx = xi - sj.x, dy = yi - sj.y, dz = zi - sj.z;
            const double  R2 = dx * dx + dy * dy + dz * dz;
            if (R2 == 0.0) continue;
            const double RS     = R2 * (TSIG ? t_half : sj.halfInvS2);
            const double sqrtRS = sqrt(RS);
            const double erfS   = erf(sqrtRS);
            const double expm   = exp(-RS);
            const double invR   = 1.0 / sqrt(R2);
            const double t2     = 2.0 * INV_SQRTPI * sqrtRS * expm;
            const double o4pr3  = 0.25 * INV_PI / R2 * invR;
            const double g      = erfS - t2;
            const double f1     = g * o4pr3;
            const double cx = dy * sj.az - dz * sj.ay, cy = dz * sj.ax - dx * sj.az, cz = dx * sj.ay - dy * sj.ax;
            acc[0] += f1 * cx;
            acc[1] += f1 * cy;
            acc[2] += f1 * cz;
            if (WITH_J) {
              const double f2 = (t2 * (TSIG ? t_is2 : sj.invS2) - 3.0 * g / R2) * o4pr3;
              const double k0 = f2 * cx, k1 = f2 * cy, k2 = f2 * cz;
              acc[3] += k0 * dx;
              acc[4] += k0 * dy + f1 * sj.az;
              acc[5] += k0 * dz - f1 * sj.ay;
              acc[6] += k1 * dx - f1 * sj.az;
              acc[7] += k1 * dy;
              acc[8] += k1 * dz + f1 * sj.ax;
              acc[9] += k2 * dx + f1 * sj.ay;
              acc[10] += k2 * dy - f1 * sj.ax;
              acc[11] += k2 * dz;
            }
          }
        }
        __syncthreads();
      }
    }
    // combine the S source splits (threads ti + q*ntp) in split order
    if constexpr (BLK == 64) {
      if (S > 1)
        for (int k = 0; k < (WITH_J ? 12 : 3); k++) {
          double v = acc[k];
          for (int q = 1; q < S; q++) v += __shfl(acc[k], ti + q * ntp, 64);
          acc[k] = v;
        }
    } else if (S > 1) {
      __syncthreads();
      for (int k = 0; k < 12; k++) red[k * BLK + lane] = acc[k];
      __syncthreads();
      // not unrolled: unrolled, the compiler kept twelve loop-invariant LDS addresses from the
      // kernel start and spilled them to scratch (an extra 3 GB of HBM traffic per 4M launch)
      if (si == 0)
#pragma unroll 1
        for (int k = 0; k < (WITH_J ? 12 : 3); k++) {
          double v = 0.0;
          for (int q = 0; q < S; q++) v += red[k * BLK + ti + q * ntp];
          acc[k] = v;
        }
    }
    if (si == 0 && valid) {
      if (SPLIT) {
        double *pb = part + (int64_t)blockIdx.x * 12 * maxnt;
        for (int k = 0; k < (WITH_J ? 12 : 3); k++) pb[k * maxnt + i] = acc[k];
      } else {
        for (int d = 0; d < 3; d++) out[(F_U + d) * ocap + bi] = -acc[d];
        if (WITH_J)
          for (int k = 0; k < 9; k++) out[(F_J + k) * ocap + bi] = -acc[3 + k];
      }
    }
  }
}

// O(M*N) direct sum, FP64, reference formula; one thread per target, sources tiled
// through LDS.  Output: u (m,3), J (m,9) and, when target strengths are given, the
// classic stretching J.alpha_t (the quantity exafmm::direct produces, traverse.hpp:88-98).
__global__ __launch_bounds__(256) void k_direct(int64_t m, const double *__restrict__ xt,
                                                const double *__restrict__ at, const double *__restrict__ sst,
                                                int64_t scap, int64_t n, double *__restrict__ u,
                                                double *__restrict__ J, double *__restrict__ dadt)
{
  __shared__ SrcF64 sh[256];
  const int64_t i  = blockIdx.x * (int64_t)256 + threadIdx.x;
  const bool    ok = i < m;
  const double  xi = ok ? xt[3 * i] : 0.0, yi = ok ? xt[3 * i + 1] : 0.0, zi = ok ? xt[3 * i + 2] : 0.0;
  double        acc[12];
  for (int k = 0; k < 12; k++) acc[k] = 0.0;
  for (int64_t j0 = 0; j0 < n; j0 += 256) {
    const int64_t j = j0 + threadIdx.x;
    if (j < n) {
      const double sg  = sst[F_SIG * scap + j];
      const double is2 = 1.0 / (sg * sg);
      sh[threadIdx.x]  = SrcF64{sst[(F_X + 0) * scap + j], sst[(F_X + 1) * scap + j], sst[(F_X + 2) * scap + j],
                                sst[(F_A + 0) * scap + j], sst[(F_A + 1) * scap + j], sst[(F_A + 2) * scap + j], is2,
                                0.5 * is2};
    }
    __syncthreads();
    const int cnt = (int)min<int64_t>(256, n - j0);
    for (int q = 0; q < cnt; q++) {
      const SrcF64 &sj = sh[q];
      const double  dx = xi - sj.x, dy = yi - sj.y, dz = zi - sj.z;
      const double  R2 = dx * dx + dy * dy + dz * dz;
      if (R2 == 0.0) continue;
      const double RS     = R2 * sj.halfInvS2;
      const double sqrtRS = sqrt(RS);
      const double t2     = 2.0 * INV_SQRTPI * sqrtRS * exp(-RS);
      const double invR   = 1.0 / sqrt(R2);
      const double o4pr3  = 0.25 * INV_PI / R2 * invR;
      const double g      = erf(sqrtRS) - t2;
      const double f1     = g * o4pr3;
      const double f2     = (t2 * sj.invS2 - 3.0 * g / R2) * o4pr3;
      const double cx = dy * sj.az - dz * sj.ay, cy = dz * sj.ax - dx * sj.az, cz = dx * sj.ay - dy * sj.ax;
      acc[0] += f1 * cx;
      acc[1] += f1 * cy;
      acc[2] += f1 * cz;
      const double k0 = f2 * cx, k1 = f2 * cy, k2 = f2 * cz;
      acc[3] += k0 * dx;
      acc[4] += k0 * dy + f1 * sj.az;
      acc[5] += k0 * dz - f1 * sj.ay;
      acc[6] += k1 * dx - f1 * sj.az;
      acc[7] += k1 * dy;
      acc[8] += k1 * dz + f1 * sj.ax;
      acc[9] += k2 * dx + f1 * sj.ay;
      acc[10] += k2 * dy - f1 * sj.ax;
      acc[11] += k2 * dz;
    }
    __syncthreads();
  }
  if (!ok) return;
  for (int d = 0; d < 3; d++) u[3 * i + d] = -acc[d];
  if (J)
    for (int k = 0; k < 9; k++) J[9 * i + k] = -acc[3 + k];
  if (dadt && at)
    for (int a = 0; a < 3; a++)
      dadt[3 * i + a] = -(acc[3 + 3 * a] * at[3 * i] + acc[4 + 3 * a] * at[3 * i + 1] + acc[5 + 3 * a] * at[3 * i + 2]);
}

}  // namespace

void direct(hipStream_t s, int64_t m, const double *xt, const double *at, const double *sst, int64_t scap, int64_t n,
            double *u, double *J, double *dadt)
{
  if (m <= 0) return;
  k_direct<<<(int)((m + 255) / 256), 256, 0, s>>>(m, xt, at, sst, scap, n, u, J, dadt);
  UVLM_LAUNCH_CHECK();
}

// Sum of the SPLIT partials of each target leaf, segments in order, stored like k_p2p does.
template <bool WITH_J>
__global__ __launch_bounds__(WG) void k_p2p_reduce(const int32_t *__restrict__ leaves, int64_t nleaf,
                                                   const int32_t *__restrict__ tbody,
                                                   const int32_t *__restrict__ tnbody,
                                                   const int32_t *__restrict__ segoff, const double *__restrict__ part,
                                                   int maxnt, double *__restrict__ out, int64_t ocap)
{
  const int64_t w = blockIdx.x;
  if (w >= nleaf) return;
  const int t = leaves[w], tb = tbody[t], nt = tnbody[t];
  for (int i = threadIdx.x; i < nt; i += WG) {
    const int64_t bi = tb + i;
    for (int k = 0; k < (WITH_J ? 12 : 3); k++) {
      double v = 0.0;
      for (int b = segoff[w]; b < segoff[w + 1]; b++) v += part[(int64_t)b * 12 * maxnt + k * maxnt + i];
      if (k < 3) out[(F_U + k) * ocap + bi] = -v;
      else out[(F_J + k - 3) * ocap + bi] = -v;
    }
  }
}

// Gaussian near-field sums over the P2P lists, FP64 (the viscous and vorticity kernels of
// multipoleKernels.hpp:343-450).  One 64-lane workgroup per target leaf, lane <-> target;
// the source bodies of each list entry pass through LDS; sums in list order (deterministic).
//  GAUSS_PSE : dadt_i += prefix * sum_j (vol_i a_j - vol_j a_i) e^{-R^2 / (2 s_ij^2)} / s_ij^5,
//              s_ij^2 = sigma_i^2 + sigma_j^2, vol = 4/3 pi sigma^3, pairs with R^2 = 0 skipped
//  GAUSS_VORT: out[field + d]_i -= sum_j a_j,d e^{-R^2 / (2 sigma_j^2)} / ((2 pi)^{3/2} sigma_j^3)
enum { GAUSS_PSE = 0, GAUSS_VORT = 1 };
template <int MODE>
__global__ __launch_bounds__(WG) void k_gauss(const int32_t *__restrict__ leaves, int64_t nleaf,
                                              const int32_t *__restrict__ tbody, const int32_t *__restrict__ tnbody,
                                              const int32_t *__restrict__ off, const int32_t *__restrict__ src,
                                              const int32_t *__restrict__ sbody, const int32_t *__restrict__ snbody,
                                              const double *__restrict__ tst, int64_t tcap,
                                              const double *__restrict__ sst, int64_t scap, double *__restrict__ out,
                                              int64_t ocap, int field, double prefix)
{
  __shared__ double sx[WG], sy[WG], sz[WG], ss[WG], sa0[WG], sa1[WG], sa2[WG];
  const int64_t w = blockIdx.x;
  if (w >= nleaf) return;
  const int t = leaves[w], tb = tbody[t], nt = tnbody[t], lane = threadIdx.x;
  const int q0 = off[t], q1 = off[t + 1];
  for (int c0 = 0; c0 < nt; c0 += WG) {
    const int     i     = c0 + lane;
    const bool    valid = i < nt;
    const int64_t bi    = tb + (valid ? i : 0);
    const double  xi = tst[(F_X + 0) * tcap + bi], yi = tst[(F_X + 1) * tcap + bi], zi = tst[(F_X + 2) * tcap + bi];
    double        sgi = 0.0, voli = 0.0, ai[3] = {0.0, 0.0, 0.0};
    if (MODE == GAUSS_PSE) {
      sgi  = tst[F_SIG * tcap + bi];
      voli = 4. / 3. * M_PI * sgi * sgi * sgi;
      for (int d = 0; d < 3; d++) ai[d] = tst[(F_A + d) * tcap + bi];
    }
    double acc[3] = {0.0, 0.0, 0.0};
    for (int q = q0; q < q1; q++) {
      const int sl = src[q], b0 = sbody[sl], nb = snbody[sl];
      for (int j0 = 0; j0 < nb; j0 += WG) {
        const int cnt = min(WG, nb - j0);
        __syncthreads();
        if (lane < cnt) {
          const int64_t bj = b0 + j0 + lane;
          sx[lane]         = sst[(F_X + 0) * scap + bj];
          sy[lane]         = sst[(F_X + 1) * scap + bj];
          sz[lane]         = sst[(F_X + 2) * scap + bj];
          ss[lane]         = sst[F_SIG * scap + bj];
          sa0[lane]        = sst[(F_A + 0) * scap + bj];
          sa1[lane]        = sst[(F_A + 1) * scap + bj];
          sa2[lane]        = sst[(F_A + 2) * scap + bj];
        }
        __syncthreads();
        if (!valid) continue;
        for (int j = 0; j < cnt; j++) {
          const double dx = xi - sx[j], dy = yi - sy[j], dz = zi - sz[j];
          const double R2 = dx * dx + dy * dy + dz * dz;
          const double sj = ss[j];
          if (MODE == GAUSS_PSE) {
            if (R2 == 0.0) continue;
            const double o2   = 1.0 / (sgi * sgi + sj * sj);
            const double s5   = o2 * o2 * sqrt(o2);
            const double volj = 4. / 3. * M_PI * sj * sj * sj;
            const double e    = exp(-0.5 * R2 * o2) * s5;
            acc[0] += (voli * sa0[j] - volj * ai[0]) * e;
            acc[1] += (voli * sa1[j] - volj * ai[1]) * e;
            acc[2] += (voli * sa2[j] - volj * ai[2]) * e;
          } else {
            const double S2 = 2.0 * sj * sj;
            const double z  = exp(-R2 / S2) * (INV_PI * INV_SQRTPI * 0.70710678118654752440) / S2 / sj;
            acc[0] += sa0[j] * z;
            acc[1] += sa1[j] * z;
            acc[2] += sa2[j] * z;
          }
        }
      }
    }
    if (valid) {
      if (MODE == GAUSS_PSE)
        for (int d = 0; d < 3; d++) out[(F_DADT + d) * ocap + bi] += acc[d] * prefix;
      else
        for (int d = 0; d < 3; d++) out[(field + d) * ocap + bi] -= acc[d];
    }
  }
}

// Below this many target leaves (probe trees), long lists are cut into segments of E list
// entries that run as separate workgroups, so the GPU stays full.
constexpr int64_t SPLIT_TARGETS = 4096, SPLIT_SEGMENTS = 8192;

void p2p(hipStream_t s, const Tree &tt, const Tree &stree, const Lists &l, const double *tgt, int64_t tcap,
         const double *srcst, int64_t scap, double *out, int64_t ocap, int precision, int sensor_mode,
         int64_t max_leaf, Scratch *sc)
{
  if (tt.nleaf == 0) return;
  const int nb = (int)tt.nleaf;
  if (sc && tt.nleaf < SPLIT_TARGETS && l.np2p >= 4 * tt.nleaf) {
    // segment table on the host: the target tree is small
    const int64_t        E = std::max<int64_t>(2, (l.np2p + SPLIT_SEGMENTS - 1) / SPLIT_SEGMENTS);
    std::vector<int32_t> lv(tt.nleaf), off(tt.C + 1);
    UVLM_HIP(hipMemcpyAsync(lv.data(), tt.leaves.p, 4 * lv.size(), hipMemcpyDeviceToHost, s));
    UVLM_HIP(hipMemcpyAsync(off.data(), l.p2p_off.p, 4 * off.size(), hipMemcpyDeviceToHost, s));
    UVLM_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> seg, segoff(tt.nleaf + 1, 0);
    for (int64_t w = 0; w < tt.nleaf; w++) {
      const int32_t q0 = off[lv[w]], q1 = off[lv[w] + 1];
      int32_t       e0 = q0;
      do {
        const int32_t e1 = (int32_t)std::min<int64_t>(q1, e0 + E);
        seg.push_back((int32_t)w);
        seg.push_back(e0);
        seg.push_back(e1);
        e0 = e1;
      } while (e0 < q1);
      segoff[w + 1] = (int32_t)(seg.size() / 3);
    }
    const int64_t nsg = segoff[tt.nleaf];
    sc->p2p_seg.reserve(4 * (seg.size() + segoff.size()));
    int32_t *dseg = sc->p2p_seg.as<int32_t>(), *dsegoff = dseg + seg.size();
    UVLM_HIP(hipMemcpyAsync(dseg, seg.data(), 4 * seg.size(), hipMemcpyHostToDevice, s));
    UVLM_HIP(hipMemcpyAsync(dsegoff, segoff.data(), 4 * segoff.size(), hipMemcpyHostToDevice, s));
    UVLM_HIP(hipStreamSynchronize(s));
    const int maxnt = (int)max_leaf;
    sc->p2p_part.reserve(sizeof(double) * nsg * 12 * maxnt);
    double    *part = sc->p2p_part.as<double>();
    const bool mixed = precision == UVLM_MIXED, wj = sensor_mode != 1, ts = sensor_mode == 2;
#define UVLM_P2P_SPLIT(MX, WJ, TS)                                                                                \
  k_p2p<MX, WJ, TS, 64, true><<<(int)nsg, WG, 0, s>>>(                                                             \
      tt.leaves.as<int32_t>(), tt.nleaf, tt.body.as<int32_t>(), tt.nbody.as<int32_t>(), tt.cx.as<double>(),        \
      tt.cy.as<double>(), tt.cz.as<double>(), l.p2p_off.as<int32_t>(), l.p2p_src.as<int32_t>(),                   \
      stree.body.as<int32_t>(), stree.nbody.as<int32_t>(), tgt, tcap, srcst, scap, out, ocap, dseg, part, maxnt)
    if (ts) {
      if (mixed) UVLM_P2P_SPLIT(true, true, true);
      else UVLM_P2P_SPLIT(false, true, true);
    } else if (wj) {
      if (mixed) UVLM_P2P_SPLIT(true, true, false);
      else UVLM_P2P_SPLIT(false, true, false);
    } else {
      if (mixed) UVLM_P2P_SPLIT(true, false, false);
      else UVLM_P2P_SPLIT(false, false, false);
    }
#undef UVLM_P2P_SPLIT
    UVLM_LAUNCH_CHECK();
    if (wj)
      k_p2p_reduce<true><<<nb, WG, 0, s>>>(tt.leaves.as<int32_t>(), tt.nleaf, tt.body.as<int32_t>(),
                                           tt.nbody.as<int32_t>(), dsegoff, part, maxnt, out, ocap);
    else
      k_p2p_reduce<false><<<nb, WG, 0, s>>>(tt.leaves.as<int32_t>(), tt.nleaf, tt.body.as<int32_t>(),
                                            tt.nbody.as<int32_t>(), dsegoff, part, maxnt, out, ocap);
    UVLM_LAUNCH_CHECK();
    return;
  }
#define UVLM_P2P(MX, WJ, TS)                                                                                      \
  if (big)                                                                                                         \
    k_p2p<MX, WJ, TS, 256><<<nb, 256, 0, s>>>(tt.leaves.as<int32_t>(), tt.nleaf, tt.body.as<int32_t>(),           \
                                              tt.nbody.as<int32_t>(), tt.cx.as<double>(), tt.cy.as<double>(),       \
                                              tt.cz.as<double>(), l.p2p_off.as<int32_t>(), l.p2p_src.as<int32_t>(), \
                                              stree.body.as<int32_t>(), stree.nbody.as<int32_t>(), tgt, tcap, srcst,\
                                              scap, out, ocap);                                                    \
  else                                                                                                             \
  k_p2p<MX, WJ, TS, 64><<<nb, WG, 0, s>>>(tt.leaves.as<int32_t>(), tt.nleaf, tt.body.as<int32_t>(), tt.nbody.as<int32_t>(), \
                                  tt.cx.as<double>(), tt.cy.as<double>(), tt.cz.as<double>(),                     \
                                  l.p2p_off.as<int32_t>(), l.p2p_src.as<int32_t>(), stree.body.as<int32_t>(),      \
                                  stree.nbody.as<int32_t>(), tgt, tcap, srcst, scap, out, ocap)
  // sensor_mode: 0 particles (u and J), 1 probes (u only), 2 adjoint (u and J, target sigma).
  // Leaves of more than 64 bodies get 256-thread workgroups, so all their targets share one
  // pass over the sources.
  const bool mixed = precision == UVLM_MIXED;
  const bool big   = max_leaf > 64;
  if (sensor_mode == 2) {
    if (mixed) UVLM_P2P(true, true, true);
    else UVLM_P2P(false, true, true);
  } else if (mixed && !sensor_mode) UVLM_P2P(true, true, false);
  else if (mixed) UVLM_P2P(true, false, false);
  else if (!sensor_mode) UVLM_P2P(false, true, false);
  else UVLM_P2P(false, false, false);
#undef UVLM_P2P
  UVLM_LAUNCH_CHECK();
}

// PSE (mode 0, adds prefix * sum to dadt) or Gaussian vorticity (mode 1, subtracts the sum
// from fields field..field+2) of the source tree at the target tree's leaves, over the P2P lists.
void gauss_near(hipStream_t s, const Tree &tt, const Tree &stree, const Lists &l, const double *tgt, int64_t tcap,
                const double *srcst, int64_t scap, double *out, int64_t ocap, int mode, int field, double prefix)
{
  if (tt.nleaf == 0) return;
  const int nb = (int)tt.nleaf;
  if (mode == 0)
    k_gauss<GAUSS_PSE><<<nb, WG, 0, s>>>(tt.leaves.as<int32_t>(), tt.nleaf, tt.body.as<int32_t>(),
                                         tt.nbody.as<int32_t>(), l.p2p_off.as<int32_t>(), l.p2p_src.as<int32_t>(),
                                         stree.body.as<int32_t>(), stree.nbody.as<int32_t>(), tgt, tcap, srcst, scap,
                                         out, ocap, field, prefix);
  else
    k_gauss<GAUSS_VORT><<<nb, WG, 0, s>>>(tt.leaves.as<int32_t>(), tt.nleaf, tt.body.as<int32_t>(),
                                          tt.nbody.as<int32_t>(), l.p2p_off.as<int32_t>(), l.p2p_src.as<int32_t>(),
                                          stree.body.as<int32_t>(), stree.nbody.as<int32_t>(), tgt, tcap, srcst, scap,
                                          out, ocap, field, prefix);
  UVLM_LAUNCH_CHECK();
}

}  // namespace uvlm
