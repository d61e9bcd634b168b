// C++ caller of the derivative API: the FMMCalculator method pairs (rateVJP, stepVJP,
// forwardTrajectory / adjointTrajectory, probeAdjoint), the autograd-style functions of
// include/uvlm/autograd.hpp, and exafmm::getList on exafmm::buildTree cells.
//
// Input (argv[1]): doubles n, m, X[3n], alpha[3n], sigma[n], ubar[3n], abar[3n], probes[3m], w[3m].
// Output (argv[2]): doubles, in this order (all [n][3] unless noted):
//   rates forward u, dadt; rates backward xbar, alphabar                 (caller order)
//   rates backward again after the calculator was reused                (caller order)
//   rateVJP xbar, alphabar (tree order), the permutation [n]
//   stepVJP(dt = 0.01) of (ubar, abar) xbar, alphabar                     (caller order)
//   trajectory forward X, alpha; backward xbar, alphabar; backward again after reuse
//   probe J [1]; probe backward (Jbar = 2) xbar, alphabar                 (caller order)
// Pairs (argv[3]): int64 n_p2p, n_m2l, then the pairs of getList(cells, cells, 0.5) of
// buildTree(bodies, 32), and the tree-ordered positions [3n] as doubles in argv[4].
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/uvlm/autograd.hpp"

struct HostTag {};
struct DeviceTag {};
typedef Vortex::FMMCalculator<HostTag, DeviceTag, Vortex::RK2, Vortex::Inviscid, Vortex::cVPM, Vortex::Transposed>
    Calc;

static std::vector<double> read_all(const char *path)
{
  FILE *f = std::fopen(path, "rb");
  if (!f) std::exit(3);
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<double> v(bytes / 8);
  if (std::fread(v.data(), 8, v.size(), f) != v.size()) std::exit(4);
  std::fclose(f);
  return v;
}

int main(int argc, char **argv)
{
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s in.bin out.bin pairs.bin tree_x.bin\n", argv[0]);
    return 2;
  }
  const std::vector<double> in = read_all(argv[1]);
  const size_t              n = (size_t)in[0], m = (size_t)in[1];
  size_t                    o = 2;
  auto take = [&](size_t k) {
    std::vector<double> v(in.begin() + o, in.begin() + o + k);
    o += k;
    return v;
  };
  const std::vector<double> X = take(3 * n), A = take(3 * n), S = take(n), ub = take(3 * n), ab = take(3 * n),
                            probes = take(3 * m), w = take(3 * m);
  std::vector<double> out;
  auto put = [&](const std::vector<double> &v) { out.insert(out.end(), v.begin(), v.end()); };
  try {
    Calc calc(8, 0.5, 64);
    // rates
    Vortex::RatesFunction<Calc>::Context rctx;
    std::vector<double>                  u, dadt, xb, alb, xb2, alb2;
    Vortex::RatesFunction<Calc>::forward(rctx, calc, X, A, S, u, dadt);
    Vortex::RatesFunction<Calc>::backward(rctx, ub, ab, xb, alb);
    put(u);
    put(dadt);
    put(xb);
    put(alb);
    calc.setArrays(probes, probes, std::vector<double>(m, 0.1));  // reuse: the forward state is gone
    calc.build();
    Vortex::RatesFunction<Calc>::backward(rctx, ub, ab, xb2, alb2);
    put(xb2);
    put(alb2);
    // method pair in tree order
    calc.setArrays(X, A, S);
    calc.build();
    const std::vector<int64_t> perm = calc.permutation();
    std::vector<double>        ubt(3 * n), abt(3 * n), xbt, albt;
    for (size_t i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) {
        ubt[3 * i + d] = ub[3 * perm[i] + d];
        abt[3 * i + d] = ab[3 * perm[i] + d];
      }
    calc.rateVJP(ubt, abt, xbt, albt);
    put(xbt);
    put(albt);
    out.insert(out.end(), perm.begin(), perm.end());
    // one step backwards (original order in and out)
    calc.setArrays(X, A, S);
    std::vector<double> sx = ub, sa = ab;
    calc.stepVJP(0.01, sx, sa);
    put(sx);
    put(sa);
    // checkpointed trajectory
    Vortex::TrajectoryFunction<Calc>::Context tctx;
    std::vector<double>                       Xf, Af, tx, ta, tx2, ta2;
    Vortex::TrajectoryFunction<Calc>::forward(tctx, calc, X, A, S, 3, 0.01, 2, Xf, Af);
    Vortex::TrajectoryFunction<Calc>::backward(tctx, ub, ab, tx, ta);
    calc.setArrays(X, A, S);  // reuse, then backward again (recomputes the forward)
    Vortex::TrajectoryFunction<Calc>::backward(tctx, ub, ab, tx2, ta2);
    put(Xf);
    put(Af);
    put(tx);
    put(ta);
    put(tx2);
    put(ta2);
    // probe objective
    Vortex::ProbeFunction<Calc>::Context pctx;
    const double                         J = Vortex::ProbeFunction<Calc>::forward(pctx, calc, X, A, S, probes, w);
    std::vector<double>                  px, pa;
    Vortex::ProbeFunction<Calc>::backward(pctx, 2.0, px, pa);
    out.push_back(J);
    put(px);
    put(pa);
    // traversal of the reference's free-function API
    exafmm::Bodies bodies(n);
    for (size_t i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) {
        bodies[i].X[d]     = X[3 * i + d];
        bodies[i].alpha[d] = A[3 * i + d];
      }
    for (size_t i = 0; i < n; i++) bodies[i].radius = S[i];
    exafmm::Cells cells = exafmm::buildTree(bodies, 32);
    auto          lists = exafmm::getList(&cells[0], &cells[0], 0.5);
    std::vector<int64_t> pairs = {(int64_t)lists.first.size(), (int64_t)lists.second.size()};
    for (auto &p : lists.first) pairs.insert(pairs.end(), {(int64_t)p.first, (int64_t)p.second});
    for (auto &p : lists.second) pairs.insert(pairs.end(), {(int64_t)p.first, (int64_t)p.second});
    std::vector<double> tree_x(3 * n);
    for (size_t i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) tree_x[3 * i + d] = bodies[i].X[d];
    FILE *f = std::fopen(argv[3], "wb");
    if (!f) return 5;
    std::fwrite(pairs.data(), 8, pairs.size(), f);
    std::fclose(f);
    f = std::fopen(argv[4], "wb");
    if (!f) return 5;
    std::fwrite(tree_x.data(), 8, tree_x.size(), f);
    std::fclose(f);
    std::printf("adjoint demo ok: n %zu, J %.6e, %zu P2P and %zu M2L pairs\n", n, J, lists.first.size(),
                lists.second.size());
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  FILE *f = std::fopen(argv[2], "wb");
  if (!f) return 5;
  std::fwrite(out.data(), 8, out.size(), f);
  std::fclose(f);
  return 0;
}
