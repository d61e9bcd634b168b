// Host-side work plan of the n-synchronous M2L kernel (k_m2l_ns in far.hip).
//
// The M2L sum of one (target, source) pair (multipoleKernels.hpp:514-565) is
//   L[j,k] += sum_{n <= P-1-j} sum_{|m| <= n} M[n,m] Y[j+n, m-k]
// over P(P+1)/2 outputs (j, k >= 0).  Terms m and -m are taken together, so output (j,k)
// at source degree n costs n+1 steps (m = 0..n).  A group of G lanes (16 or 32) shares one
// pair and walks the source degrees n in lock step: at each step every lane reads the same
// M[n,m] (an LDS broadcast) and its own harmonics.  The outputs active at degree n are
// w = tri(j)+k < tri(P-n), a prefix in (j, k) order.
//
// The W wavefronts of a workgroup (1 or 2) share the staged sources and split the source
// degrees between them; each has its own plan for its degrees.  A lane of a wavefront owns a
// few accumulator slots:
//   * home slots: output w lives in lane w % G, slot w / G, for the outputs active at the
//     wavefront's lowest degree, and is summed there over every one of its degrees where the
//     output is active and the degree not split;
//   * helper slots, one per "split" degree n: when the outputs still active at degree n
//     leave lanes idle in their last round, the m-range 0..n of each of those outputs is cut
//     into chunks of equal length, one chunk per lane, in a slot reserved for that degree.
// Every (output, n, m) term is then owned by exactly one (wavefront, lane, slot); the kernel
// adds the slots of an output in the fixed order of `out_src`, so the result is deterministic.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace uvlm {

struct M2LPlan {
  static constexpr int HOME = 31;          // level field of a home slot
  int P = 0, S = 0, H = 0, G = 16, W = 1;  // slots per lane (kernel template), home slots of
                                           // the wavefront with degree 0, group lanes, wavefronts
  std::vector<int32_t> slots;              // [W][G][S]: pack_slot(j, k, m0, level) or -1
  std::vector<int32_t> lev;                // [W][P]: chunk length | home rounds << 8 | active
                                           // slots << 16; 0 for the other wavefront's degrees
  std::vector<int32_t> out_off;            // [P(P+1)/2 + 1]
  std::vector<int32_t> out_src;            // pieces of each output: (wavefront * G + lane) * S + slot
};

inline int32_t pack_slot(int j, int k, int m0, int level) { return j | (k << 5) | (m0 << 10) | (level << 15); }
inline int     plan_tri(int n) { return n * (n + 1) / 2; }

// Slot counts the kernel is instantiated for.
inline int plan_template_slots(int need)
{
  for (int s : {4, 5, 6, 8, 12, 16})
    if (need <= s) return s;
  return -1;
}

// G: lanes per group (16 or 32).  budget: accumulator slots per lane to aim for (0 = default:
// 6, or the home slots + 1 when those alone need more; each slot holds 3 complex FP64 sums).
// W: wavefronts (1 or 2).  Returns false if P is outside 1..16, G is not 16 or 32 or W not 1 or 2.
inline bool make_m2l_plan(int P, int budget, M2LPlan &pl, int G = 16, int W = 1)
{
  if (P < 1 || P > 16 || (G != 16 && G != 32) || W < 1 || W > 2) return false;
  const int L = G, ny = plan_tri(P);
  const int H = (ny + L - 1) / L;
  if (budget <= 0) budget = std::max(6, H + 1);
  budget = std::max(budget, H);
  const int S = plan_template_slots(budget);
  if (S < 0 || S < H) return false;
  budget = std::min(budget, S);
  // window w = tri(j) + k in (j, k) order: the outputs active at degree n are w < tri(P - n)
  std::vector<int> wj(ny), wk(ny);
  for (int j = 0; j < P; j++)
    for (int k = 0; k <= j; k++) {
      wj[plan_tri(j) + k] = j;
      wk[plan_tri(j) + k] = k;
    }
  // split option of each degree: c chunks of equal length cn = (n+1)/c per output of the last
  // round (every lane of a split degree then runs the same trip count), as many as the idle
  // lanes allow; it saves n + 1 - cn steps
  std::vector<int> sp_cn(P, 0), sp_c(P, 0), sp_save(P, 0);
  for (int n = 0; n < P; n++) {
    const int A = plan_tri(P - n), rem = A % L;
    if (!rem) continue;
    int c = std::min(L / rem, n + 1);
    while (c > 1 && (n + 1) % c) c--;
    if (c > 1) sp_c[n] = c, sp_cn[n] = (n + 1) / c, sp_save[n] = n + 1 - (n + 1) / c;
  }
  // The plan of one wavefront for a set of degrees: its home slots hold the outputs active at
  // its lowest degree, and its remaining slots go to the splits that save the most steps.
  // Cost: lock-step steps plus about half a step of set-up per active slot and degree.
  struct WavePlan {
    int              Hw = 0;
    std::vector<int> split;  // degrees split, in degree order
    double           cost = 0;
  };
  auto wave_plan = [&](int mask) {
    WavePlan wp;
    int      n0 = -1;
    for (int n = 0; n < P && n0 < 0; n++)
      if ((mask >> n) & 1) n0 = n;
    if (n0 < 0) return wp;
    wp.Hw = (plan_tri(P - n0) + L - 1) / L;
    std::vector<int> opts;
    for (int n = 0; n < P; n++)
      if (((mask >> n) & 1) && sp_c[n] > 1) opts.push_back(n);
    std::stable_sort(opts.begin(), opts.end(), [&](int a, int b) { return sp_save[a] > sp_save[b]; });
    for (int n : opts)
      if (wp.Hw + (int)wp.split.size() < budget) wp.split.push_back(n);
    std::sort(wp.split.begin(), wp.split.end());
    for (int n = 0; n < P; n++) {
      if (!((mask >> n) & 1)) continue;
      const int  A     = plan_tri(P - n);
      const bool split = std::find(wp.split.begin(), wp.split.end(), n) != wp.split.end();
      if (split) wp.cost += (A / L) * (n + 1) + sp_cn[n] + 0.5 * (A / L + 1);
      else wp.cost += ((A + L - 1) / L) * (n + 1.0) + 0.5 * ((A + L - 1) / L);
    }
    return wp;
  };
  // The degrees' partition: all of them for one wavefront; for two, every split with degree 0
  // in the first is tried and the one with the least total cost kept (then the least larger
  // share: the two wavefronts meet at the end of every batch).
  std::vector<int> masks(W, 0);
  if (W == 1) masks[0] = (1 << P) - 1;
  else {
    double best_tot = 1e300, best_max = 1e300;
    for (int m = 1; m < (1 << P); m += 2) {
      const int    m1 = ((1 << P) - 1) & ~m;
      const double c0 = wave_plan(m).cost, c1 = wave_plan(m1).cost, tot = c0 + c1, mx = std::max(c0, c1);
      if (tot < best_tot - 1e-9 || (tot < best_tot + 1e-9 && mx < best_max - 1e-9)) {
        best_tot = tot, best_max = mx;
        masks[0] = m, masks[1] = m1;
      }
    }
  }

  pl.P = P;
  pl.S = S;
  pl.H = H;
  pl.G = G;
  pl.W = W;
  pl.slots.assign((size_t)W * L * S, -1);
  pl.lev.assign((size_t)W * P, 0);
  std::vector<std::vector<int>> pieces(ny);
  for (int w = 0; w < W; w++) {
    if (!masks[w]) continue;
    const WavePlan wp = wave_plan(masks[w]);
    int32_t       *sl = pl.slots.data() + (size_t)w * L * S;
    int32_t       *lv = pl.lev.data() + (size_t)w * P;
    int            n0 = 0;
    while (!((masks[w] >> n0) & 1)) n0++;
    for (int o = 0; o < plan_tri(P - n0); o++) sl[(o % L) * S + o / L] = pack_slot(wj[o], wk[o], 0, M2LPlan::HOME);
    for (int n = 0; n < P; n++)
      if ((masks[w] >> n) & 1) lv[n] = (n + 1) | (((plan_tri(P - n) + L - 1) / L) << 8);
    for (size_t t = 0; t < wp.split.size(); t++) {
      const int n = wp.split[t], A = plan_tri(P - n), full = A / L, slot = wp.Hw + (int)t;
      int       lane = 0;
      for (int o = full * L; o < A; o++)
        for (int i = 0; i < sp_c[n]; i++) sl[(lane++) * S + slot] = pack_slot(wj[o], wk[o], i * sp_cn[n], n);
      lv[n] = sp_cn[n] | (full << 8);
    }
    // slots that any lane uses at degree n (a uniform skip test in the kernel)
    for (int n = 0; n < P; n++) {
      if (!((masks[w] >> n) & 1)) continue;
      const int hr   = (lv[n] >> 8) & 255;
      int       mask = 0;
      for (int s = 0; s < wp.Hw && s < hr; s++) mask |= 1 << s;
      for (size_t t = 0; t < wp.split.size(); t++)
        if (wp.split[t] == n) mask |= 1 << (wp.Hw + (int)t);
      lv[n] |= mask << 16;
    }
    for (int lane = 0; lane < L; lane++)
      for (int s = 0; s < S; s++) {
        const int v = sl[lane * S + s];
        if (v >= 0) pieces[plan_tri(v & 31) + ((v >> 5) & 31)].push_back((w * L + lane) * S + s);
      }
  }
  pl.out_off.assign(ny + 1, 0);
  pl.out_src.clear();
  for (int o = 0; o < ny; o++) {
    std::sort(pieces[o].begin(), pieces[o].end());
    for (int v : pieces[o]) pl.out_src.push_back(v);
    pl.out_off[o + 1] = (int)pl.out_src.size();
  }
  return true;
}

}  // namespace uvlm
