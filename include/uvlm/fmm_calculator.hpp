// uvlm/fmm_calculator.hpp -- C++ drop-in for the FMM path of kursatyurt/UVLM.
//
// Same spellings as the reference so that a caller switches by changing includes:
//   exafmm::Body / Bodies / Cell / Cells      (src/exafmmTypes.hpp:18-47, same Body layout)
//   exafmm::buildTree(Bodies&)                (src/tree.hpp:101-113)
//   exafmm::direct(Bodies&, Bodies&)          (src/traverse.hpp:88-98)
//   exafmm::writeVTK(Bodies, path)            (the driver's writeTovtk, src/main.cpp:127-164)
//   Vortex tags Inviscid/PSE/CSM, Euler/RK2/RK3/RK4, cVPM/rVPM, Transposed/Classic
//                                             (src/fmmCalculator.hpp:9-23)
//   Vortex::FMMCalculator<Host, Device, Integration, Viscous, Formulation, Stretching>
//     init / calculate / advance / getSensorData / calculateViscosity /
//     copyBodies2Host / copyBodies2Device     (src/fmmCalculator.hpp:25-1006)
// Differences: P, theta, ncrit and nu are constructor arguments (the reference has the first
// three as constexpr in exafmmTypes.hpp:12-15 and leaves nu uninitialised); Host/Device tags
// are accepted and ignored; failures throw std::runtime_error
// carrying uvlm_last_error().  All compute runs on the GPU through include/uvlm_fmm.h.
#ifndef UVLM_FMM_CALCULATOR_HPP
#define UVLM_FMM_CALCULATOR_HPP

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../uvlm_fmm.h"

namespace exafmm {
typedef double real_t;

// The reference's compile-time defaults (exafmmTypes.hpp:12-15).  Here they are only defaults:
// FMMCalculator, buildTree and getList take P, ncrit and theta as arguments.
static constexpr int    P         = 4;
static constexpr int    NTERM     = 3 * (P * (P + 1)) / 2;
static constexpr size_t ncrit     = 128;
static constexpr real_t MAC_theta = 0.3;

struct Body {
  real_t J[3][3]{{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  real_t X[3]{0.0, 0.0, 0.0};
  real_t alpha[3]{0.0, 0.0, 0.0};
  real_t velocity[3]{0.0, 0.0, 0.0};
  real_t dadt[3]{0.0, 0.0, 0.0};
  real_t velocity_old[3]{0.0, 0.0, 0.0};
  real_t dadt_old[3]{0.0, 0.0, 0.0};
  real_t X_old[3]{0.0, 0.0, 0.0};
  real_t alpha_old[3]{0.0, 0.0, 0.0};
  real_t radius{0.0};
  real_t drdt{0.0};
  real_t radius_old{0.0};
  real_t drdt_old{0.0};
};
static_assert(sizeof(Body) == 296, "exafmm::Body must keep the reference layout (37 doubles)");
static_assert(offsetof(Body, X) == 72 && offsetof(Body, alpha) == 96 && offsetof(Body, radius) == 264,
              "exafmm::Body field offsets");
typedef std::vector<Body> Bodies;

struct Cell {
  size_t NCHILD;
  size_t NBODY;
  size_t ChildOffset;
  size_t BodyOffset;
  real_t X[3];
  Cell  *CHILD;
  Body  *BODY;
  real_t R;
  size_t level;
};
typedef std::vector<Cell> Cells;

namespace detail {
inline void check(int rc)
{
  if (rc != UVLM_OK) throw std::runtime_error(std::string("uvlm: ") + uvlm_last_error());
}

struct Handle {
  uvlm_handle h = nullptr;
  explicit Handle(const uvlm_params &p) { check(uvlm_create(&p, &h)); }
  ~Handle() { uvlm_destroy(h); }
  Handle(const Handle &)            = delete;
  Handle &operator=(const Handle &) = delete;
};

inline uvlm_params params(int P, double theta, size_t ncrit, int precision, int stretching, int formulation,
                          int scheme, int device)
{
  uvlm_params p{};
  p.P           = P;
  p.theta       = theta;
  p.ncrit       = (int64_t)ncrit;
  p.precision   = precision;
  p.stretching  = stretching;
  p.formulation = formulation;
  p.scheme      = scheme;
  p.device      = device;
  return p;
}

// Fills a Cells vector (with CHILD/BODY pointers into `cells` / `bodies`) from the device tree.
inline Cells fetch_cells(uvlm_handle h, Bodies &bodies)
{
  int64_t C = 0;
  check(uvlm_num_cells(h, &C));
  std::vector<int64_t> ints(5 * C);
  std::vector<double>  reals(4 * C);
  check(uvlm_get_tree(h, C, ints.data(), reals.data()));
  Cells cells(C);
  for (int64_t c = 0; c < C; c++) {
    Cell &q       = cells[c];
    q.NCHILD      = (size_t)ints[5 * c + 0];
    q.NBODY       = (size_t)ints[5 * c + 1];
    q.ChildOffset = (size_t)ints[5 * c + 2];
    q.BodyOffset  = (size_t)ints[5 * c + 3];
    q.level       = (size_t)ints[5 * c + 4];
    for (int d = 0; d < 3; d++) q.X[d] = reals[4 * c + d];
    q.R     = reals[4 * c + 3];
    q.CHILD = q.NCHILD ? &cells[0] + q.ChildOffset : nullptr;
    q.BODY  = bodies.data() + q.BodyOffset;
  }
  return cells;
}
}  // namespace detail

// Octree of `bodies` built on the GPU; permutes `bodies` into tree order like tree.hpp:101.
inline Cells buildTree(Bodies &bodies, size_t ncrit = 128, int device = 0)
{
  detail::Handle h(detail::params(4, 0.3, ncrit, UVLM_FP64, UVLM_TRANSPOSED, UVLM_CVPM, UVLM_RK3, device));
  detail::check(uvlm_set_bodies_aos(h.h, bodies.data(), (int64_t)bodies.size(), sizeof(Body)));
  detail::check(uvlm_build(h.h));
  detail::check(uvlm_get_bodies_aos(h.h, bodies.data(), (int64_t)bodies.size(), sizeof(Body)));
  return detail::fetch_cells(h.h, bodies);
}

// Dual-tree traversal of traverse.hpp:32-85 from (targetRoot, sourceRoot): the P2P and M2L pair
// lists as (target cell index, source cell index), sorted by target then source (the reference
// returns the same pairs in its stack order).  The roots are the first cells of Cells vectors such
// as buildTree returns (children after their parent, CHILD pointers into the same vector).  The
// traversal runs on the GPU.
typedef std::vector<std::pair<size_t, size_t>> PairList;
inline std::pair<PairList, PairList> getList(const Cell *targetRoot, const Cell *sourceRoot, real_t theta = MAC_theta,
                                             int device = 0)
{
  auto pack = [](const Cell *root, std::vector<int64_t> &ints, std::vector<double> &reals) {
    // cell count: the root plus every child group reachable from it
    size_t C = 1;
    for (size_t c = 0; c < C; c++)
      if (root[c].NCHILD) C = std::max(C, (size_t)(root[c].CHILD - root) + root[c].NCHILD);
    ints.resize(5 * C);
    reals.resize(4 * C);
    for (size_t c = 0; c < C; c++) {
      const Cell &q   = root[c];
      ints[5 * c + 0] = (int64_t)q.NCHILD;
      ints[5 * c + 1] = (int64_t)q.NBODY;
      ints[5 * c + 2] = q.NCHILD ? (int64_t)(q.CHILD - root) : 0;
      ints[5 * c + 3] = (int64_t)q.BodyOffset;
      ints[5 * c + 4] = (int64_t)q.level;
      for (int d = 0; d < 3; d++) reals[4 * c + d] = q.X[d];
      reals[4 * c + 3] = q.R;
    }
    return (int64_t)C;
  };
  std::vector<int64_t> ti, si;
  std::vector<double>  tr, sr;
  const int64_t        ct = pack(targetRoot, ti, tr), cs = pack(sourceRoot, si, sr);
  int64_t np = 0, nm = 0;  // sizes first, then the pairs (the traversal runs twice)
  detail::check(uvlm_get_list(device, theta, ct, ti.data(), tr.data(), cs, si.data(), sr.data(), 0, 0, &np, &nm,
                              nullptr, nullptr));
  std::vector<int64_t> p2p(2 * np + 2), m2l(2 * nm + 2);
  detail::check(uvlm_get_list(device, theta, ct, ti.data(), tr.data(), cs, si.data(), sr.data(), np, nm, &np, &nm,
                              p2p.data(), m2l.data()));
  PairList P2P((size_t)np), M2L((size_t)nm);
  for (int64_t i = 0; i < np; i++) P2P[i] = {(size_t)p2p[2 * i], (size_t)p2p[2 * i + 1]};
  for (int64_t i = 0; i < nm; i++) M2L[i] = {(size_t)m2l[2 * i], (size_t)m2l[2 * i + 1]};
  return {std::move(P2P), std::move(M2L)};
}

// FP64 direct sum of jbodies onto bodies: velocity += u, dadt += J.alpha (traverse.hpp:88-98).
inline void direct(Bodies &bodies, Bodies &jbodies, int device = 0)
{
  detail::Handle h(detail::params(4, 0.3, 128, UVLM_FP64, UVLM_CLASSIC, UVLM_CVPM, UVLM_RK3, device));
  detail::check(uvlm_set_bodies_aos(h.h, jbodies.data(), (int64_t)jbodies.size(), sizeof(Body)));
  const size_t        m = bodies.size();
  std::vector<double> xt(3 * m), at(3 * m), u(3 * m), da(3 * m);
  for (size_t i = 0; i < m; i++)
    for (int d = 0; d < 3; d++) {
      xt[3 * i + d] = bodies[i].X[d];
      at[3 * i + d] = bodies[i].alpha[d];
    }
  detail::check(uvlm_direct(h.h, (int64_t)m, xt.data(), at.data(), u.data(), nullptr, da.data()));
  for (size_t i = 0; i < m; i++)
    for (int d = 0; d < 3; d++) {
      bodies[i].velocity[d] += u[3 * i + d];
      bodies[i].dadt[d] += da[3 * i + d];
    }
}
// Legacy-ASCII VTK POLYDATA of the particles (positions, alpha, velocity, radius), the file
// the reference's driver writes per step (src/main.cpp:127-164): same sections, same order,
// default ostream number formatting.  Host-only; returns false if the file cannot be opened.
inline bool writeVTK(const Bodies &particles, const std::string &path)
{
  std::ofstream file(path);
  if (!file) return false;
  const size_t n = particles.size();
  file << "# vtk DataFile Version 3.0\n" << "vtk output\n" << "ASCII\n" << "DATASET POLYDATA\n";
  file << "POINTS " << n << " double\n";
  for (const Body &b : particles) file << b.X[0] << " " << b.X[1] << " " << b.X[2] << "\n";
  file << "VERTICES " << n << " " << 2 * n << "\n";
  for (size_t i = 0; i < n; i++) file << "1 " << i << "\n";
  file << "POINT_DATA " << n << "\n";
  file << "VECTORS alpha double\n";
  for (const Body &b : particles) file << b.alpha[0] << " " << b.alpha[1] << " " << b.alpha[2] << "\n";
  file << "VECTORS velocity double\n";
  for (const Body &b : particles) file << b.velocity[0] << " " << b.velocity[1] << " " << b.velocity[2] << "\n";
  file << "SCALARS radius double\n" << "LOOKUP_TABLE default\n";
  for (const Body &b : particles) file << b.radius << "\n";
  return (bool)file;
}
}  // namespace exafmm

namespace Vortex {
struct Inviscid {};
struct PSE {};
struct CSM {};
struct Euler {};
struct RK2 {};
struct RK3 {};
struct RK4 {};
struct cVPM {};
struct rVPM {};
struct Transposed {};
struct Classic {};

enum class Precision { FP64 = UVLM_FP64, Mixed = UVLM_MIXED };

template <typename Host, typename Device, typename IntegrationScheme, typename ViscousScheme, typename Formulation,
          typename Stretching>
struct FMMCalculator {
  static_assert(std::is_same<ViscousScheme, Inviscid>::value || std::is_same<ViscousScheme, PSE>::value ||
                    std::is_same<ViscousScheme, CSM>::value,
                "ViscousScheme should be Inviscid, PSE or CSM");
  static_assert(std::is_same<Stretching, Transposed>::value || std::is_same<Stretching, Classic>::value,
                "Stretching should be either Transposed or Classic");

  static constexpr int scheme()
  {
    return std::is_same<IntegrationScheme, Euler>::value ? UVLM_EULER
           : std::is_same<IntegrationScheme, RK2>::value ? UVLM_RK2
           : std::is_same<IntegrationScheme, RK3>::value ? UVLM_RK3
                                                          : UVLM_RK4;
  }

  // nu: the kinematic viscosity of the PSE / CSM schemes (the reference's _nu, which it never
  // initialises, fmmCalculator.hpp:56); ignored for Inviscid.
  explicit FMMCalculator(int P = 4, double theta = 0.3, size_t ncrit = 128, Precision precision = Precision::FP64,
                         int device = 0, exafmm::real_t nu = 0.0)
      : h_(detail_params(P, theta, ncrit, precision, device))
  {
    const int v = std::is_same<ViscousScheme, PSE>::value   ? UVLM_PSE
                  : std::is_same<ViscousScheme, CSM>::value ? UVLM_CSM
                                                            : UVLM_INVISCID;
    exafmm::detail::check(uvlm_set_viscosity(h_.h, v, nu));
  }

  // fmmCalculator.hpp:282-288: tree + lists; permutes `bodies` into tree order.
  void init(exafmm::Bodies &bodies)
  {
    copyBodies2Device(bodies);
    call(uvlm_build(h_.h));
    copyBodies2Host(bodies);
  }
  // fmmCalculator.hpp:420-467
  void calculate() { call(uvlm_eval(h_.h)); }
  // fmmCalculator.hpp:713-724 (inviscid: nothing to do)
  // (PSE: dadt += exchange over the near pairs; CSM: drdt += core spreading over the dt of the
  //  last advance(), or setTimeStep())
  void calculateViscosity() { call(uvlm_viscosity(h_.h, dt_)); }
  void setTimeStep(exafmm::real_t dt) { dt_ = dt; }
  // calculateVorticityDualTree (fmmCalculator.hpp:746-753) at `sensors` (after init()):
  // velocity_old = -(Gaussian vorticity); the sensors come back in their tree order.
  void calculateVorticity(exafmm::Bodies &sensors)
  {
    call(uvlm_vorticity_aos(h_.h, sensors.data(), (int64_t)sensors.size(), sizeof(exafmm::Body)));
  }
  // fmmCalculator.hpp:855-917: init(particles), then the strengths of newPoints fitted to the
  // particles' vorticity; newPoints keeps the points above the vorticity threshold (tree order).
  void RBF(exafmm::Bodies &particles, exafmm::Bodies &newPoints)
  {
    init(particles);
    int64_t kept = 0;
    call(uvlm_rbf_aos(h_.h, newPoints.data(), (int64_t)newPoints.size(), sizeof(exafmm::Body), &kept,
                                       nullptr, nullptr));
    newPoints.resize((size_t)kept);
  }
  // fmmCalculator.hpp:919-1005: every stage rebuilds the tree; the state stays on the GPU.
  void advance(exafmm::Bodies &bodies, exafmm::real_t dt)
  {
    dt_ = dt;
    copyBodies2Device(bodies);
    call(uvlm_advance(h_.h, dt));
    copyBodies2Host(bodies);
  }
  // fmmCalculator.hpp:726-731: uses the multipoles of the last calculate().
  void getSensorData(exafmm::Bodies &sensors)
  {
    call(uvlm_sensors_aos(h_.h, sensors.data(), (int64_t)sensors.size(), sizeof(exafmm::Body)));
  }
  void copyBodies2Device(exafmm::Bodies &bodies)
  {
    call(
        uvlm_set_bodies_aos(h_.h, bodies.data(), (int64_t)bodies.size(), sizeof(exafmm::Body)));
    n_ = bodies.size();
  }
  void copyBodies2Host(exafmm::Bodies &bodies)
  {
    int64_t n = 0;
    exafmm::detail::check(uvlm_num_bodies(h_.h, &n));  // the wing's shedding appends on the device
    n_ = (size_t)n;
    bodies.resize(n_);
    exafmm::detail::check(uvlm_get_bodies_aos(h_.h, bodies.data(), (int64_t)n_, sizeof(exafmm::Body)));
  }
  // Array form of copyBodies2Device: positions and strengths [n][3], core sizes [n], in the
  // caller's order (the "original" order of the calls below that use it).
  void setArrays(const std::vector<double> &X, const std::vector<double> &alpha, const std::vector<double> &sigma)
  {
    if (X.size() != alpha.size() || X.size() != 3 * sigma.size()) throw std::runtime_error("X, alpha must be [n][3]");
    call(uvlm_set_bodies_soa(h_.h, (int64_t)sigma.size(), X.data(), alpha.data(), sigma.data()));
    n_ = sigma.size();
  }
  // The current state in tree order ([n][3] arrays, sigma [n]); null outputs are skipped.
  void getArrays(std::vector<double> *X, std::vector<double> *alpha, std::vector<double> *sigma,
                 std::vector<double> *velocity = nullptr, std::vector<double> *dadt = nullptr) const
  {
    for (std::vector<double> *v : {X, alpha, velocity, dadt})
      if (v) v->assign(3 * n_, 0.0);
    if (sigma) sigma->assign(n_, 0.0);
    auto p = [](std::vector<double> *v) { return v ? v->data() : nullptr; };
    exafmm::detail::check(uvlm_get_bodies_soa(h_.h, (int64_t)n_, p(X), p(alpha), p(sigma), p(velocity), nullptr,
                                              p(dadt), nullptr));
  }
  // perm[i] = caller index of the body at tree position i.
  std::vector<int64_t> permutation() const
  {
    std::vector<int64_t> perm(n_);
    exafmm::detail::check(uvlm_get_permutation(h_.h, (int64_t)n_, perm.data()));
    return perm;
  }
  // Tree + lists of the bodies on the device (init() without the host copies).
  void build() { call(uvlm_build(h_.h)); }
  // Interaction lists as (target cell, source cell) pairs (the reference keeps them in
  // _P2PPairsView / _M2LPairsView, fmmCalculator.hpp:50-51).
  std::pair<std::vector<std::pair<size_t, size_t>>, std::vector<std::pair<size_t, size_t>>> lists() const
  {
    int64_t np = 0, nm = 0;
    exafmm::detail::check(uvlm_get_lists(h_.h, &np, &nm, nullptr, nullptr));
    std::vector<int64_t> p(2 * np), m(2 * nm);
    exafmm::detail::check(uvlm_get_lists(h_.h, nullptr, nullptr, p.data(), m.data()));
    std::vector<std::pair<size_t, size_t>> P2P(np), M2L(nm);
    for (int64_t i = 0; i < np; i++) P2P[i] = {(size_t)p[2 * i], (size_t)p[2 * i + 1]};
    for (int64_t i = 0; i < nm; i++) M2L[i] = {(size_t)m[2 * i], (size_t)m[2 * i + 1]};
    return {std::move(P2P), std::move(M2L)};
  }
  uvlm_timing timings() const
  {
    uvlm_timing t{};
    exafmm::detail::check(uvlm_timings(h_.h, &t));
    return t;
  }
  uvlm_handle handle() const { return h_.h; }

  // Calls with no counterpart in the reference class (include/uvlm_fmm.h gives the reasons).
  // Pedrizzetti relaxation toward curl u of the last calculate(), and the invariants
  // {impulse[3], angular impulse[3], energy, sum alpha[3]} of the current particles.
  void relax(exafmm::real_t f) { call(uvlm_relax(h_.h, f)); }
  std::array<double, 10> diagnostics() const
  {
    std::array<double, 10> out{};
    exafmm::detail::check(uvlm_diagnostics(h_.h, out.data()));
    return out;
  }
  // dJ/dalpha and dJ/dx ([n][3] each, tree order) of J = sum_p w_p . u(probe_p).
  void probeAdjoint(const std::vector<double> &probes, const std::vector<double> &w, std::vector<double> &dJ_dalpha,
                    std::vector<double> &dJ_dx)
  {
    if (probes.size() != w.size() || probes.size() % 3) throw std::runtime_error("probes and w must be [m][3]");
    dJ_dalpha.assign(3 * n_, 0.0);
    dJ_dx.assign(3 * n_, 0.0);
    call(uvlm_probe_adjoint(h_.h, (int64_t)(probes.size() / 3), probes.data(), w.data(),
                                             dJ_dalpha.data(), dJ_dx.data()));
  }
  // xbar, alphabar ([n][3], tree order) of S = sum ubar.u + abar.dadt on the current tree
  // (calculate() is run first); the tree of the last init() is held fixed.
  void rateVJP(const std::vector<double> &ubar, const std::vector<double> &abar, std::vector<double> &xbar,
               std::vector<double> &alphabar)
  {
    if (ubar.size() != 3 * n_ || abar.size() != 3 * n_) throw std::runtime_error("ubar and abar must be [n][3]");
    xbar.assign(3 * n_, 0.0);
    alphabar.assign(3 * n_, 0.0);
    call(uvlm_eval_vjp(h_.h, (int64_t)n_, ubar.data(), abar.data(), xbar.data(), alphabar.data()));
  }
  // One advance(dt) backwards from the device state: cotangents ([n][3], ORIGINAL body order)
  // of the state after the step in, those of the state before it out.
  void stepVJP(exafmm::real_t dt, std::vector<double> &xbar, std::vector<double> &alphabar)
  {
    if (xbar.size() != 3 * n_ || alphabar.size() != 3 * n_) throw std::runtime_error("cotangents must be [n][3]");
    call(uvlm_step_vjp(h_.h, dt, (int64_t)n_, xbar.data(), alphabar.data()));
  }
  // Checkpointed trajectory: `steps` advances of dt from the device state with checkpoints every
  // `every` steps, then (adjointTrajectory) the cotangent of the final state mapped back to the
  // initial one, in place (original body order).
  void forwardTrajectory(size_t steps, exafmm::real_t dt, size_t every)
  {
    call(uvlm_trajectory_forward(h_.h, (int64_t)steps, dt, (int64_t)every));
  }
  void adjointTrajectory(std::vector<double> &xbar, std::vector<double> &alphabar)
  {
    if (xbar.size() != 3 * n_ || alphabar.size() != 3 * n_) throw std::runtime_error("cotangents must be [n][3]");
    call(uvlm_trajectory_vjp(h_.h, (int64_t)n_, xbar.data(), alphabar.data()));
  }
  // Near-field option (uvlm_set_near_field): pairs whose bodies can come closer than c * max sigma
  // use the regularised kernel (P2P); 0, the default, is the reference's acceptance test alone.
  void setNearField(exafmm::real_t c) { call(uvlm_set_near_field(h_.h, c)); }
  // Leaf size (and M2L plan size) for the bodies last passed to init()/copyBodies2Device();
  // returns the chosen ncrit.  init() again before calculate().
  size_t autotune(const std::vector<int64_t> &candidates, int reps = 2)
  {
    int64_t best = 0;
    double  ms   = 0.0;
    call(uvlm_autotune(h_.h, candidates.data(), (int)candidates.size(), reps, &best, &ms));
    return (size_t)best;
  }

  // Incremented by every call that changes the device state (the autograd-style functions use it
  // to tell whether the calculator still holds their forward pass).
  uint64_t version() const { return version_; }

private:
  void call(int rc)
  {
    ++version_;
    exafmm::detail::check(rc);
  }
  static uvlm_params detail_params(int P, double theta, size_t ncrit, Precision precision, int device)
  {
    return exafmm::detail::params(P, theta, ncrit, (int)precision,
                                  std::is_same<Stretching, Transposed>::value ? UVLM_TRANSPOSED : UVLM_CLASSIC,
                                  std::is_same<Formulation, rVPM>::value ? UVLM_RVPM : UVLM_CVPM, scheme(), device);
  }
  exafmm::detail::Handle h_;
  size_t                 n_       = 0;
  exafmm::real_t         dt_      = 0.0;
  uint64_t               version_ = 0;
};
}  // namespace Vortex

#endif  // UVLM_FMM_CALCULATOR_HPP
