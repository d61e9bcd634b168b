// Host-side work plan of the n-synchronous M2L kernel (k_m2l_ns in far.hip).
//
// The M2L sum of one (target, source) pair (multipoleKernels.hpp:514-565) is
//   L[j,k] += sum_{n <= P-1-j} sum_{|m| <= n} M[n,m] Y[j+n, m-k]
// over P(P+1)/2 outputs (j, k >= 0).  Terms m and -m are taken together, so output (j,k)
// at source degree n costs n+1 steps (m = 0..n).  A group of G lanes (16 or 32) shares one
// pair and walks the source degrees n = 0..P-1 in lock step: at each step every lane reads
// the same M[n,m] (an LDS broadcast) and its own harmonics.  A lane owns a few accumulator
// slots:
//   * home slots: output w = tri(j)+k lives in lane w % G, slot w / G and is summed there
//     over every degree n where it is active (j + n <= P-1) and not split;
//   * helper slots, one per "split" degree n: when the outputs still active at degree n
//     leave lanes idle in their last round, the m-range 0..n of each of those outputs is cut
//     into chunks of equal length, one chunk per lane, in a slot reserved for that degree.
// Every (output, n, m) term is then owned by exactly one (lane, slot); the kernel adds the
// slots of an output in the fixed order of `out_src`, so the result is deterministic.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace uvlm {

struct M2LPlan {
  static constexpr int HOME = 31;   // level field of a home slot
  int P = 0, S = 0, H = 0, G = 16;  // slots per lane (kernel template), home slots, group lanes
  std::vector<int32_t> slots;       // [G][S]: pack_slot(j, k, m0, level) or -1
  std::vector<int32_t> lev;         // [P]: chunk length | home rounds << 8 | wavefront << 15 | active slots << 16
  std::vector<int32_t> out_off;     // [P(P+1)/2 + 1]
  std::vector<int32_t> out_src;     // pieces of each output: lane * S + slot
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
// Returns false if P is outside 1..16 or G is not 16 or 32.
inline bool make_m2l_plan(int P, int budget, M2LPlan &pl, int G = 16)
{
  if (P < 1 || P > 16 || (G != 16 && G != 32)) return false;
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
  struct Split {
    int saving, n, cn, c;
  };
  std::vector<Split> opts;
  for (int n = 0; n < P; n++) {
    const int A = plan_tri(P - n), rem = A % L;
    if (!rem) continue;
    // c chunks of equal length cn = (n+1)/c per output (every lane of a split degree then runs
    // the same trip count), as many as the idle lanes allow
    int c = std::min(L / rem, n + 1);
    while (c > 1 && (n + 1) % c) c--;
    const int cn = (n + 1) / c;
    if (c > 1) opts.push_back({n + 1 - cn, n, cn, c});
  }
  std::stable_sort(opts.begin(), opts.end(), [](const Split &a, const Split &b) { return a.saving > b.saving; });
  std::vector<Split> chosen;
  for (const Split &o : opts)
    if (H + (int)chosen.size() < budget) chosen.push_back(o);
  std::sort(chosen.begin(), chosen.end(), [](const Split &a, const Split &b) { return a.n < b.n; });

  pl.P = P;
  pl.S = S;
  pl.H = H;
  pl.G = G;
  pl.slots.assign((size_t)L * S, -1);
  pl.lev.assign(P, 0);
  std::vector<std::vector<int>> pieces(ny);
  for (int w = 0; w < ny; w++) {
    const int lane = w % L, slot = w / L;
    pl.slots[lane * S + slot] = pack_slot(wj[w], wk[w], 0, M2LPlan::HOME);
    pieces[w].push_back(lane * S + slot);
  }
  for (int n = 0; n < P; n++) {
    const int A = plan_tri(P - n);
    pl.lev[n]   = (n + 1) | (((A + L - 1) / L) << 8);
  }
  for (size_t t = 0; t < chosen.size(); t++) {
    const Split &o    = chosen[t];
    const int    A    = plan_tri(P - o.n), full = A / L;
    const int    slot = H + (int)t;
    int          lane = 0;
    for (int w = full * L; w < A; w++)
      for (int i = 0; i < o.c; i++) {
        const int m0 = i * o.cn;
        if (m0 > o.n) break;
        pl.slots[lane * S + slot] = pack_slot(wj[w], wk[w], m0, o.n);
        pieces[w].push_back(lane * S + slot);
        lane++;
      }
    pl.lev[o.n] = o.cn | (full << 8);
  }
  // slots that any lane uses at degree n (a uniform skip test in the kernel)
  for (int n = 0; n < P; n++) {
    const int hr   = (pl.lev[n] >> 8) & 127;
    int       mask = 0;
    for (int s = 0; s < H && s < hr; s++) mask |= 1 << s;
    for (size_t t = 0; t < chosen.size(); t++)
      if (chosen[t].n == n) mask |= 1 << (H + (int)t);
    pl.lev[n] |= mask << 16;
  }
  // A workgroup of two wavefronts splits the degrees between them (bit 15: the second one) so
  // that the larger of the two lock-step costs is least (per degree: n+1 steps per home slot,
  // cn per helper, and about half a step of set-up per slot; all 2^P splits are tried, ties
  // keep the degree-parity split's neighbourhood by preferring the smaller mask).
  std::vector<double> dc(P, 0.0);
  for (int n = 0; n < P; n++) {
    const int lv = pl.lev[n], cn = lv & 255, hr = (lv >> 8) & 127, used = lv >> 16;
    for (int s = 0; s < S; s++)
      if ((used >> s) & 1) dc[n] += (s < hr ? n + 1 : cn) + 0.5;
  }
  double total = 0;
  for (double c : dc) total += c;
  int    best_mask = 0;
  double best      = total;
  for (int mask = 0; mask < (1 << P); mask++) {
    double c1 = 0;
    for (int n = 0; n < P; n++)
      if ((mask >> n) & 1) c1 += dc[n];
    const double worst = std::max(c1, total - c1);
    if (worst < best - 1e-9) best = worst, best_mask = mask;
  }
  for (int n = 0; n < P; n++)
    if ((best_mask >> n) & 1) pl.lev[n] |= 1 << 15;
  pl.out_off.assign(ny + 1, 0);
  pl.out_src.clear();
  for (int w = 0; w < ny; w++) {
    std::sort(pieces[w].begin(), pieces[w].end());
    for (int v : pieces[w]) pl.out_src.push_back(v);
    pl.out_off[w + 1] = (int)pl.out_src.size();
  }
  return true;
}

}  // namespace uvlm
