// Exhaustive host check of the M2L work plan (uvlm_amd/csrc/m2l_plan.hpp) for G = 16 and 32
// lanes per group, one and two wavefronts, P = 1..16 and every slot budget: every (output j,k;
// degree n; order m) term of the truncated M2L sum is owned by exactly one (wavefront, lane,
// slot), every harmonic and multipole index the kernel forms stays inside its row, the two
// wavefronts take disjoint degrees, and the per-output piece lists name each used (wavefront,
// lane, slot) once.  Two rules the kernel's uniform trip counts rest on are checked too: a home
// slot s of lane l is active at degree n exactly when G s + l < tri(P - n), and every helper
// chunk holds exactly cn orders.  Exit status 0 when all plans pass.
#include "../../uvlm_amd/csrc/m2l_plan.hpp"

#include <cstdio>
using namespace uvlm;
int main()
{
  int bad = 0;
  for (int W : {1, 2})
  for (int G : {16, 32})
  for (int P = 1; P <= 16; P++)
    for (int budget : {0, 4, 5, 6, 8, 9, 12, 16}) {
      M2LPlan pl;
      if (!make_m2l_plan(P, budget, pl, G, W)) { printf("W=%d G=%d P=%d budget=%d: no plan\n", W, G, P, budget); continue; }
      const int ny = plan_tri(P), S = pl.S, NT = 3 * ny;
      if (pl.W != W || (int)pl.slots.size() != W * G * S || (int)pl.lev.size() != W * P) { printf("bad sizes\n"); bad++; continue; }
      std::vector<int> cnt(ny * 16 * 16, 0);
      std::vector<int> owner_ok(W * G * S, 0);
      for (int n = 0; n < P; n++) {
        int owners = 0;
        for (int w = 0; w < W; w++) owners += pl.lev[w * P + n] != 0;
        if (owners != 1) { printf("W=%d P=%d: degree %d in %d wavefronts\n", W, P, n, owners); bad++; }
      }
      for (int w = 0; w < W; w++)
      for (int n = 0; n < P; n++) {
        const int lv = pl.lev[w * P + n], cn = lv & 255, hr = (lv >> 8) & 255, used = lv >> 16;
        if (!lv) continue;
        for (int s = 0; s < S; s++) {
          if (!((used >> s) & 1)) continue;
          for (int l = 0; l < G; l++) {
            const int v = pl.slots[(w * G + l) * S + s];
            if (v < 0) continue;
            const int j = v & 31, k = (v >> 5) & 31, m0 = (v >> 10) & 31, lvl = (v >> 15) & 31;
            bool act; int a, b;
            if (lvl == M2LPlan::HOME) {
              act = s < hr && j + n <= P - 1;
              a = 0; b = n + 1;
              if (act != (s < hr && G * s + l < plan_tri(P - n))) { printf("P=%d home activity rule\n", P); bad++; }
            } else {
              act = lvl == n;
              a = m0; b = std::min(m0 + cn, n + 1);
              if (act && m0 + cn > n + 1) { printf("P=%d budget=%d: unequal chunk at n=%d\n", P, budget, n); bad++; }
            }
            if (!act) continue;
            const int J = j + n;
            if (J > P - 1 || k > j) { printf("P=%d bad J/k\n", P); bad++; }
            for (int m = a; m < b; m++) {
              const int mu = m - k < 0 ? k - m : m - k;
              if (mu > J || m + k > J || 3 * (plan_tri(n) + m) + 2 >= NT) { printf("P=%d oob\n", P); bad++; }
              cnt[(plan_tri(j) + k) * 256 + n * 16 + m]++;
            }
          }
        }
      }
      for (int j = 0; j < P; j++)
        for (int k = 0; k <= j; k++)
          for (int n = 0; n + j <= P - 1; n++)
            for (int m = 0; m <= n; m++)
              if (cnt[(plan_tri(j) + k) * 256 + n * 16 + m] != 1) {
                printf("W=%d P=%d budget=%d: term (%d,%d,%d,%d) count %d\n", W, P, budget, j, k, n, m,
                       cnt[(plan_tri(j) + k) * 256 + n * 16 + m]);
                bad++;
              }
      // pieces: every used (wavefront, lane, slot) is listed under its output exactly once
      for (int o = 0; o < ny; o++)
        for (int e = pl.out_off[o]; e < pl.out_off[o + 1]; e++) {
          const int pc = pl.out_src[e];
          if (pc < 0 || pc >= W * G * S) { printf("bad piece\n"); bad++; continue; }
          const int v = pl.slots[pc];
          if (v < 0 || plan_tri(v & 31) + ((v >> 5) & 31) != o) { printf("bad piece\n"); bad++; }
          owner_ok[pc]++;
        }
      for (int pc = 0; pc < W * G * S; pc++)
        if ((pl.slots[pc] >= 0) != (owner_ok[pc] == 1)) { printf("P=%d piece %d listed %d\n", P, pc, owner_ok[pc]); bad++; }
    }
  printf("%s\n", bad ? "FAILED" : "all terms owned exactly once, in bounds");
  return bad != 0;
}
