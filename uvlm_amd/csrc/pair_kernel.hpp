// pair_kernel.hpp -- FP32 evaluation of the regularised Biot-Savart kernel factor and its
// derivatives (shared by the near-field kernels and their adjoint), and the LDS prefix sum
// the near-field kernels use to stream the bodies of several leaves in full chunks.
//
// With s = R / (sqrt(2) sigma), x = s^2, g = erf(s) - t2, t2 = (2/sqrt(pi)) s e^{-s^2}
// (multipoleKernels.hpp:243-283), the pair factors are functions of G(x) = g / s^3:
//   f1 = g / (4 pi R^3)   = G(x)   q^{3/2} / (4 pi),
//   f2 = 2 df1/dR^2       = 2 G'(x)  q^{5/2} / (4 pi),
//   f3 = 2 df2/dR^2       = 4 G''(x) q^{7/2} / (4 pi),      q = 1 / (2 sigma^2),
// because 2 s^2 t2 - 3 g = s g'(s) - 3 g = 2 s^5 G'(x), and
//   G(x) = (2/sqrt(pi)) sum_{i>=0} (-1)^i 2(i+1) / ((2i+3) (i+1)!) x^i.
// For x < 2, G is approximated on [0, 2] by its degree-9 Chebyshev interpolant (kernel_g, which
// also needs G''; the forward kernels interpolate G and G' separately, c_gnear and c_gpnear
// below) (interpolation at the Chebyshev nodes of [0, 2], re-expanded in powers of t = x - 1),
// with no cancellation between erf and t2: max error 8e-11 (G), 5e-9 (G') and 7.6e-8 (G'',
// |G''| <= 0.33) in exact arithmetic, 1.4e-7 and 1.1e-7 (G, G') when evaluated by Horner in FP32.
// For x >= 2 the hardware rsq/exp2 and a short polynomial: with w = x^{-1/2} = 1/s, e = e^{-x},
//   g = erf(s) - (2/sqrt(pi)) s e = 1 - e (erfcx(s) + (2/sqrt(pi)) s),
//   G = g w^3,   G' = w^2 ((2/sqrt(pi)) e - (3/2) G),
//   G'' = -(2/sqrt(pi)) e (w^2 + w^4) - (3/2) w^2 (G' - G w^2)
// (from G' = (2/sqrt(pi)) s e x^{-3/2} - (3/2) G / x), where erfcx(s) = e^{s^2} erfc(s) is a
// degree-7 least-squares polynomial in w - 0.4647 on w in [1/4.5, 1/sqrt(2)] (max error
// 1e-8; beyond s = 4.5 the factor e < 2e-9 makes it irrelevant, so w is clamped there).
// In FP32 this form is as accurate as erff/expf (4.3e-7 on G, 8.1e-7 on G') for about a
// third of the instructions.
#pragma once
#include <hip/hip_runtime.h>

namespace uvlm {
namespace {

// Chebyshev interpolant of G on x in [0, 2], coefficients of t^i, t = x - 1.
constexpr int   NPOLY              = 10;
constexpr float XPOLY              = 2.0f;
__constant__ float c_gpoly[NPOLY] = {4.27593295550599495e-01f,  -2.26282445882890593e-01f, 7.52993075578484716e-02f,
                                     -1.86646103035026413e-02f, 3.70154966273965411e-03f,  -6.12467145832412121e-04f,
                                     8.69427232794296300e-05f,  -1.08135023163257920e-05f, 1.22398144588401693e-06f,
                                     -1.21730778572683114e-07f};

// The forward near-field kernels need G and G' only: each has its own degree-7 interpolant at
// the Chebyshev nodes of [0, 2] (coefficients of t^i, t = x - 1), accurate to 1.1e-8 (G) and
// 9.5e-9 (G') in exact arithmetic and to 7.3e-8 and 4.3e-8 by fused Horner in FP32.  Two chains
// of 7 multiply-adds, independent of each other, instead of one degree-8 polynomial and its
// derivative (15 multiply-adds, the derivative's chain waiting on the value's).
constexpr int   NNEAR              = 8;
__constant__ float c_gnear[NNEAR]  = {4.27593286009723905e-01f,  -2.26282444927409232e-01f, 7.52996131662352769e-02f,
                                      -1.86646407540406734e-02f, 3.70002071831887491e-03f,  -6.12314968091220990e-04f,
                                      8.93899977314482896e-05f,  -1.10569638636093259e-05f};
__constant__ float c_gpnear[NNEAR] = {-2.26282437333498510e-01f, 1.50598616404860219e-01f,  -5.59941046870113882e-02f,
                                      1.48061917699502943e-02f,  -3.06096641491382382e-03f, 5.21663223968849010e-04f,
                                      -7.78856703251840088e-05f, 9.79185156779749377e-06f};

// erfcx(1/w) ~ sum_k c_far[k] (w - FAR_WC)^k, w in [FAR_WMIN, 1/sqrt(2)].
constexpr int   NFAR     = 8;
constexpr float FAR_WC   = 0.46466450170438484f;
constexpr float FAR_WMIN = 0.2222222222222222f;
__constant__ float c_far[NFAR] = {2.400683016e-01f,  4.403707087e-01f, -1.874301285e-01f, 1.258202083e-02f,
                                  8.160093427e-02f, -1.016800180e-01f, 5.416155234e-02f,  2.011496387e-02f};

// G, G' and (D2) G'' at x = xq >= 0.
template <bool D2>
__device__ __forceinline__ void kernel_g(float xq, float &G, float &Gp, float &Gpp)
{
  if (xq < XPOLY) {
    const float tq = xq - 1.0f;
    float       g = c_gpoly[NPOLY - 1], gp = 0.f, gpp = 0.f;
#pragma unroll
    for (int k = NPOLY - 2; k >= 0; k--) {
      if (D2) gpp = fmaf(gpp, tq, gp);
      gp = fmaf(gp, tq, g);
      g  = fmaf(g, tq, c_gpoly[k]);
    }
    G   = g;
    Gp  = gp;
    Gpp = 2.0f * gpp;
  } else {
    const float w = __builtin_amdgcn_rsqf(xq);
    const float e = __builtin_amdgcn_exp2f(xq * -1.4426950408889634f);
    const float t = fmaxf(w, FAR_WMIN) - FAR_WC;
    float       p = c_far[NFAR - 1];
#pragma unroll
    for (int k = NFAR - 2; k >= 0; k--) p = fmaf(p, t, c_far[k]);
    const float g  = fmaf(-e, fmaf(1.1283791670955126f, xq * w, p), 1.0f);
    const float w2 = w * w;
    G              = g * w2 * w;
    Gp             = w2 * fmaf(1.1283791670955126f, e, -1.5f * G);
    if (D2) Gpp = -1.1283791670955126f * e * fmaf(w2, w2, w2) - 1.5f * w2 * fmaf(-G, w2, Gp);
  }
}

// G, G' at t0 = x0 - 1 and t1 = x1 - 1 (x < XPOLY) by the degree-7 interpolants: four
// independent Horner chains, each started after its first step.
__device__ __forceinline__ void g_near(float t0, float t1, float &G0, float &Gp0, float &G1, float &Gp1)
{
  float g0 = fmaf(c_gnear[NNEAR - 1], t0, c_gnear[NNEAR - 2]), g1 = fmaf(c_gnear[NNEAR - 1], t1, c_gnear[NNEAR - 2]);
  float p0 = fmaf(c_gpnear[NNEAR - 1], t0, c_gpnear[NNEAR - 2]), p1 = fmaf(c_gpnear[NNEAR - 1], t1, c_gpnear[NNEAR - 2]);
#pragma unroll
  for (int k = NNEAR - 3; k >= 0; k--) {
    g0 = fmaf(g0, t0, c_gnear[k]);
    g1 = fmaf(g1, t1, c_gnear[k]);
    p0 = fmaf(p0, t0, c_gpnear[k]);
    p1 = fmaf(p1, t1, c_gpnear[k]);
  }
  G0 = g0, Gp0 = p0, G1 = g1, Gp1 = p1;
}

// G, G' at x >= XPOLY from rsq/exp2 and the erfcx polynomial (the x >= XPOLY branch of kernel_g).
__device__ __forceinline__ void g_far(float xq, float &G, float &Gp)
{
  const float w = __builtin_amdgcn_rsqf(xq);
  const float e = __builtin_amdgcn_exp2f(xq * -1.4426950408889634f);
  const float t = fmaxf(w, FAR_WMIN) - FAR_WC;
  float       p = c_far[NFAR - 1];
#pragma unroll
  for (int k = NFAR - 2; k >= 0; k--) p = fmaf(p, t, c_far[k]);
  const float g  = fmaf(-e, fmaf(1.1283791670955126f, xq * w, p), 1.0f);
  const float w2 = w * w;
  G              = g * w2 * w;
  Gp             = w2 * fmaf(1.1283791670955126f, e, -1.5f * G);
}

// Targets per pass of a near-field workgroup of BLK lanes over a leaf of nt targets: lanes are
// (target, source split) with S = BLK / ntp splits, and the leaf takes ceil(nt / ntp) passes over
// its source list.  With BLK = 256 a pass costs about 1/S of the list per lane plus the staging
// (about 1/128), so ntp = ceil(nt / np) with np in 1..8 minimising np (1/S + 1/128); e.g. nt = 200
// takes 4 passes of 50 targets and 5 splits (98% of the lanes busy) instead of one pass with 56
// idle lanes.  With BLK = 64 a 64-source chunk is too short to pay for its staging again: one pass.
template <int BLK>
__device__ __forceinline__ int near_targets_per_pass(int nt)
{
  int best = max(1, min(nt, BLK));
  if (BLK <= 64) return best;
  float cost = 3.0e38f;
  for (int np = 1; np <= 8; np++) {  // np = 1 gives best, with cost 1/S + 1/128
    const int ntp = max(1, (nt + np - 1) / np);
    if (ntp > BLK) continue;
    const float c = np * (1.0f / (float)(BLK / ntp) + 1.0f / 128.0f);
    if (c < cost) cost = c, best = ntp;
  }
  return best;
}

constexpr int SEG = 256;  // source leaves whose prefix sums sit in LDS at once

// Exclusive prefix sum of v[0..n) (n <= SEG) into pref[0..n]; one wavefront, 4 per lane.
__device__ __forceinline__ void lds_prefix(int *pref, const int *v, int n)
{
  const int lane = threadIdx.x;
  int       loc[4], tot = 0;
  for (int k = 0; k < 4; k++) {
    const int i = 4 * lane + k;
    loc[k]      = i < n ? v[i] : 0;
    tot += loc[k];
  }
  int incl = tot;
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  int run = incl - tot;
  for (int k = 0; k < 4; k++) {
    const int i = 4 * lane + k;
    if (i <= n) pref[i] = run;
    run += loc[k];
  }
  if (lane == 63 && 256 <= n) pref[256] = run;
}

}  // namespace
}  // namespace uvlm
