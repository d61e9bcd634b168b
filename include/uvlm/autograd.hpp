// uvlm/autograd.hpp -- forward/backward functions over Vortex::FMMCalculator, in the manner of an
// autograd Function: forward(ctx, ...) computes the outputs and saves in ctx what backward needs;
// backward(ctx, ...) maps the cotangents of the outputs to those of the inputs (vector-Jacobian
// products).  The reference solver has no derivatives, so none of this has a counterpart there;
// the gradients are those of the FMM approximation with each evaluation's tree and lists held
// fixed (include/uvlm_fmm.h: uvlm_eval_vjp, uvlm_trajectory_vjp, uvlm_probe_adjoint).
//
// All arrays are row-major doubles in the CALLER's body order: positions, strengths and their
// cotangents [n][3], core sizes [n].  The functions undo the permutation of the tree.
//
// The calculator keeps the forward's device state (tree, lists, checkpoints).  If the calculator
// was used for anything else between forward and backward, backward first re-runs the forward
// from the inputs saved in ctx (the builds are deterministic, so the result is the same).
#ifndef UVLM_AUTOGRAD_HPP
#define UVLM_AUTOGRAD_HPP

#include <cstdint>
#include <stdexcept>
#include <vector>

#include "fmm_calculator.hpp"

namespace Vortex {

template <class Calculator> struct FunctionContext {
  Calculator         *calc    = nullptr;
  uint64_t            version = 0;  // calc->version() when the saved device state was last valid
  std::vector<double> X, alpha, sigma;
  std::vector<double> probes, weights;
  size_t              steps = 0, every = 1;
  exafmm::real_t      dt    = 0.0;
  bool                holds_forward() const { return calc && calc->version() == version; }
};

namespace detail {
// tree order -> caller order: out[perm[i]] = in[i] (rows of w doubles)
inline void to_caller(const std::vector<int64_t> &perm, const std::vector<double> &in, std::vector<double> &out,
                      int w = 3)
{
  out.assign(in.size(), 0.0);
  for (size_t i = 0; i < perm.size(); i++)
    for (int d = 0; d < w; d++) out[w * (size_t)perm[i] + d] = in[w * i + d];
}
// caller order -> tree order: out[i] = in[perm[i]]
inline std::vector<double> to_tree(const std::vector<int64_t> &perm, const std::vector<double> &in, int w = 3)
{
  if (in.size() != w * perm.size()) throw std::runtime_error("cotangent size does not match the bodies");
  std::vector<double> out(in.size());
  for (size_t i = 0; i < perm.size(); i++)
    for (int d = 0; d < w; d++) out[w * i + d] = in[w * (size_t)perm[i] + d];
  return out;
}
}  // namespace detail

// (X, alpha, sigma) -> (u, dadt): tree, lists and one FMM evaluation of the velocity and the
// stretching (the calculator's Stretching tag).  backward: (ubar, abar) -> (xbar, alphabar) of
// S = sum_i ubar_i . u_i + abar_i . dadt_i.
template <class Calculator> struct RatesFunction {
  typedef FunctionContext<Calculator> Context;

  static void forward(Context &ctx, Calculator &calc, const std::vector<double> &X, const std::vector<double> &alpha,
                      const std::vector<double> &sigma, std::vector<double> &u, std::vector<double> &dadt)
  {
    ctx       = Context{};
    ctx.calc  = &calc;
    ctx.X     = X;
    ctx.alpha = alpha;
    ctx.sigma = sigma;
    run(ctx);
    calc.calculate();
    std::vector<double> ut, dt;
    calc.getArrays(nullptr, nullptr, nullptr, &ut, &dt);
    const std::vector<int64_t> perm = calc.permutation();
    detail::to_caller(perm, ut, u);
    detail::to_caller(perm, dt, dadt);
    ctx.version = calc.version();
  }

  static void backward(Context &ctx, const std::vector<double> &ubar, const std::vector<double> &abar,
                       std::vector<double> &xbar, std::vector<double> &alphabar)
  {
    Calculator &calc = *ctx.calc;
    if (!ctx.holds_forward()) run(ctx);
    const std::vector<int64_t> perm = calc.permutation();
    std::vector<double>        xb, ab;
    calc.rateVJP(detail::to_tree(perm, ubar), detail::to_tree(perm, abar), xb, ab);
    detail::to_caller(perm, xb, xbar);
    detail::to_caller(perm, ab, alphabar);
    ctx.version = calc.version();  // the VJP keeps the forward's tree and lists
  }

private:
  static void run(Context &ctx)
  {
    ctx.calc->setArrays(ctx.X, ctx.alpha, ctx.sigma);
    ctx.calc->build();
  }
};

// (X, alpha, sigma) at the start -> (X, alpha) after `steps` advances of dt (the calculator's
// IntegrationScheme), with (X, alpha, sigma) checkpointed on the device every `every` steps.
// backward: cotangents of the final (X, alpha) -> those of the initial (X, alpha), each segment
// recomputed from its checkpoint.
template <class Calculator> struct TrajectoryFunction {
  typedef FunctionContext<Calculator> Context;

  static void forward(Context &ctx, Calculator &calc, const std::vector<double> &X, const std::vector<double> &alpha,
                      const std::vector<double> &sigma, size_t steps, exafmm::real_t dt, size_t every,
                      std::vector<double> &X_out, std::vector<double> &alpha_out)
  {
    ctx       = Context{};
    ctx.calc  = &calc;
    ctx.X     = X;
    ctx.alpha = alpha;
    ctx.sigma = sigma;
    ctx.steps = steps;
    ctx.dt    = dt;
    ctx.every = every;
    run(ctx);
    std::vector<double> xt, at;
    calc.getArrays(&xt, &at, nullptr);
    const std::vector<int64_t> perm = calc.permutation();
    detail::to_caller(perm, xt, X_out);
    detail::to_caller(perm, at, alpha_out);
    ctx.version = calc.version();
  }

  static void backward(Context &ctx, const std::vector<double> &xbar_out, const std::vector<double> &alphabar_out,
                       std::vector<double> &xbar, std::vector<double> &alphabar)
  {
    if (!ctx.holds_forward()) run(ctx);
    xbar     = xbar_out;  // the trajectory VJP takes and returns the caller's (original) order
    alphabar = alphabar_out;
    ctx.calc->adjointTrajectory(xbar, alphabar);
    ctx.version = ctx.calc->version();  // the particles end at the final state again
  }

private:
  static void run(Context &ctx)
  {
    ctx.calc->setArrays(ctx.X, ctx.alpha, ctx.sigma);
    ctx.calc->forwardTrajectory(ctx.steps, ctx.dt, ctx.every);
  }
};

// (X, alpha, sigma) -> J = sum_p w_p . u(probe_p), the velocity at the probes from one FMM
// evaluation (getSensorData).  backward: Jbar -> (xbar, alphabar) = Jbar * (dJ/dX, dJ/dalpha).
template <class Calculator> struct ProbeFunction {
  typedef FunctionContext<Calculator> Context;

  static double forward(Context &ctx, Calculator &calc, const std::vector<double> &X, const std::vector<double> &alpha,
                        const std::vector<double> &sigma, const std::vector<double> &probes,
                        const std::vector<double> &weights)
  {
    if (probes.size() != weights.size() || probes.size() % 3) throw std::runtime_error("probes and w must be [m][3]");
    ctx         = Context{};
    ctx.calc    = &calc;
    ctx.X       = X;
    ctx.alpha   = alpha;
    ctx.sigma   = sigma;
    ctx.probes  = probes;
    ctx.weights = weights;
    run(ctx);
    calc.calculate();
    const size_t   m = probes.size() / 3;
    exafmm::Bodies s(m);
    for (size_t p = 0; p < m; p++) {
      for (int d = 0; d < 3; d++) s[p].X[d] = probes[3 * p + d];
      s[p].drdt_old = (double)p;  // carries the probe's index through the probe tree's permutation
    }
    calc.getSensorData(s);
    double J = 0.0;
    for (const exafmm::Body &b : s) {
      const size_t p = (size_t)b.drdt_old;
      for (int d = 0; d < 3; d++) J += weights[3 * p + d] * b.velocity[d];
    }
    ctx.version = calc.version();
    return J;
  }

  static void backward(Context &ctx, double Jbar, std::vector<double> &xbar, std::vector<double> &alphabar)
  {
    Calculator &calc = *ctx.calc;
    if (!ctx.holds_forward()) run(ctx);
    std::vector<double> da, dx;
    calc.probeAdjoint(ctx.probes, ctx.weights, da, dx);
    const std::vector<int64_t> perm = calc.permutation();
    detail::to_caller(perm, dx, xbar);
    detail::to_caller(perm, da, alphabar);
    for (double &v : xbar) v *= Jbar;
    for (double &v : alphabar) v *= Jbar;
    ctx.version = calc.version();
  }

private:
  static void run(Context &ctx)
  {
    ctx.calc->setArrays(ctx.X, ctx.alpha, ctx.sigma);
    ctx.calc->build();
  }
};

}  // namespace Vortex

#endif  // UVLM_AUTOGRAD_HPP
