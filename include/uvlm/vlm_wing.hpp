// uvlm/vlm_wing.hpp -- C++ caller-side wing of the reference's driver (src/main.cpp:27-74,
// 168-295) over the C ABI's uvlm_wing_* entry points; the particles are a
// Vortex::FMMCalculator's (its handle()).
//
//   uvlm::Wing wing(le, te);                 // vertex couples, as Wing::addVertexCouple
//   auto F = wing.step(fmm.handle(), U, dt); // solve + shed + advance + move by -U dt
//   fmm.copyBodies2Host(particles);          // the wake, including the shed particles
//
// The reference's wing is one chordwise row (nchord = 1); the lattice constructor takes
// [nchord+1][nspan+1][3] vertices.  Failures throw std::runtime_error with uvlm_last_error().
#ifndef UVLM_VLM_WING_HPP
#define UVLM_VLM_WING_HPP

#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../uvlm_fmm.h"

namespace uvlm {

class Wing {
public:
  using Vec3 = std::array<double, 3>;

  // Leading- and trailing-edge vertex couples (main.cpp:39-43): one row of le.size() - 1 panels.
  Wing(const std::vector<Vec3> &le, const std::vector<Vec3> &te, int device = 0)
  {
    if (le.size() != te.size() || le.size() < 2) throw std::runtime_error("need >= 2 matching vertex couples");
    std::vector<double> v;
    for (const auto *row : {&le, &te})
      for (const Vec3 &p : *row) v.insert(v.end(), p.begin(), p.end());
    create(device, (int64_t)le.size() - 1, 1, v);
  }
  // A lattice of vertices [nchord + 1][nspan + 1][3], row 0 the leading edge.
  Wing(int64_t nspan, int64_t nchord, const std::vector<double> &vertices, int device = 0)
  {
    if ((int64_t)vertices.size() != 3 * (nspan + 1) * (nchord + 1)) throw std::runtime_error("vertex count");
    create(device, nspan, nchord, vertices);
  }
  ~Wing() { uvlm_wing_destroy(w_); }
  Wing(const Wing &)            = delete;
  Wing &operator=(const Wing &) = delete;

  size_t getPanelCount() const { return (size_t)n_; }
  // calculateTopology's results (main.cpp:45-58): control points, unit normals, areas.
  void topology(std::vector<Vec3> &controlPoints, std::vector<Vec3> &normals, std::vector<double> &areas) const
  {
    std::vector<double> cp(3 * n_), nr(3 * n_);
    areas.assign(n_, 0.0);
    check(uvlm_wing_get(w_, cp.data(), nr.data(), areas.data(), nullptr));
    controlPoints.resize(n_);
    normals.resize(n_);
    for (int64_t i = 0; i < n_; i++)
      for (int d = 0; d < 3; d++) controlPoints[i][d] = cp[3 * i + d], normals[i][d] = nr[3 * i + d];
  }
  void translate(const Vec3 &d) { check(uvlm_wing_translate(w_, d.data())); }
  // Circulations (gamma, may be null) and force at the wing's position; particles may be null.
  Vec3 solve(uvlm_handle particles, const Vec3 &U, std::vector<double> *gamma = nullptr)
  {
    Vec3 F{};
    if (gamma) gamma->assign(n_, 0.0);
    check(uvlm_wing_solve(w_, particles, U.data(), gamma ? gamma->data() : nullptr, F.data()));
    return F;
  }
  // One pass of the reference's loop (main.cpp:168-295); returns the force of its solve.
  Vec3 step(uvlm_handle particles, const Vec3 &U, double dt, bool referenceOffset = true,
            std::vector<double> *gamma = nullptr)
  {
    Vec3 F{};
    if (gamma) gamma->assign(n_, 0.0);
    check(uvlm_wing_step(w_, particles, U.data(), dt, referenceOffset ? 1 : 0, gamma ? gamma->data() : nullptr,
                         F.data()));
    return F;
  }
  // CL after `steps` loop steps and its gradient with respect to the particles now (original
  // order, [n][3] each).
  double liftAdjoint(uvlm_handle particles, const Vec3 &U, double dt, int64_t steps, const Vec3 &liftDirection,
                     std::vector<double> &xbar, std::vector<double> &alphabar, bool referenceOffset = true)
  {
    int64_t n = 0;
    check(uvlm_num_bodies(particles, &n));
    xbar.assign(3 * n, 0.0);
    alphabar.assign(3 * n, 0.0);
    double CL = 0.0;
    check(uvlm_wing_lift_adjoint(w_, particles, U.data(), dt, steps, referenceOffset ? 1 : 0, liftDirection.data(),
                                 &CL, n, xbar.data(), alphabar.data()));
    return CL;
  }

  // The design gradient through a window of `steps` loop steps (checkpoints every `every`): CL
  // at its end and dCL/d(vertices) [(nchord+1)(nspan+1)][3], the wake at the window start held
  // fixed; xbar / alphabar (may be null) receive the gradient with respect to the particles.
  double windowAdjoint(uvlm_handle particles, const Vec3 &U, double dt, int64_t steps, int64_t every,
                       const Vec3 &liftDirection, std::vector<double> &vertexBar, std::vector<double> *xbar = nullptr,
                       std::vector<double> *alphabar = nullptr, bool referenceOffset = true)
  {
    int64_t n = 0;
    check(uvlm_num_bodies(particles, &n));
    if (xbar) xbar->assign(3 * n, 0.0);
    if (alphabar) alphabar->assign(3 * n, 0.0);
    vertexBar.assign(3 * (nspan_ + 1) * (nchord_ + 1), 0.0);
    double CL = 0.0;
    check(uvlm_wing_window_adjoint(w_, particles, U.data(), dt, steps, every, referenceOffset ? 1 : 0,
                                   liftDirection.data(), &CL, n, xbar ? xbar->data() : nullptr,
                                   alphabar ? alphabar->data() : nullptr, vertexBar.data()));
    return CL;
  }

private:
  uvlm_wing w_ = nullptr;
  int64_t   n_ = 0, nspan_ = 0, nchord_ = 0;
  static void check(int rc)
  {
    if (rc != UVLM_OK) throw std::runtime_error(std::string("uvlm: ") + uvlm_last_error());
  }
  void create(int device, int64_t nspan, int64_t nchord, const std::vector<double> &v)
  {
    check(uvlm_wing_create(device, nspan, nchord, v.data(), &w_));
    n_      = nspan * nchord;
    nspan_  = nspan;
    nchord_ = nchord;
  }
};

// The benchmark wing's design variables: a rectangular lattice (stations x_k = k span / nspan,
// the chord along y in nchord rows) whose station k is rotated about the x axis through the
// leading edge by -(alpha + t_k), as the reference rotates its wing by the angle of attack
// (main.cpp:172-184).  t_k, the station twist, comes from one twist value per spanwise strip:
// the mean of the two strips at an interior station, the strip's value at a tip.  Radians.
struct WingDesign {
  int64_t nspan = 40, nchord = 10;
  double  span = 6.0, chord = 1.0;

  std::vector<double> stationAngles(double alpha, const std::vector<double> &twist) const
  {
    std::vector<double> th(nspan + 1, alpha);
    if (twist.empty()) return th;
    if ((int64_t)twist.size() != nspan) throw std::runtime_error("one twist value per spanwise strip");
    for (int64_t k = 0; k <= nspan; k++)
      th[k] += k == 0 ? twist[0] : k == nspan ? twist[nspan - 1] : 0.5 * (twist[k - 1] + twist[k]);
    return th;
  }
  // Vertices [nchord + 1][nspan + 1][3], translated by `offset`.
  std::vector<double> vertices(double alpha, const std::vector<double> &twist, const Wing::Vec3 &offset = {0, 0, 0}) const
  {
    const std::vector<double> th = stationAngles(alpha, twist);
    std::vector<double>       V(3 * (nspan + 1) * (nchord + 1));
    for (int64_t r = 0; r <= nchord; r++)
      for (int64_t k = 0; k <= nspan; k++) {
        const double y = chord * (double)r / (double)nchord;
        double      *v = &V[3 * (r * (nspan + 1) + k)];
        v[0]           = (double)k * span / (double)nspan + offset[0];
        v[1]           = y * std::cos(th[k]) + offset[1];
        v[2]           = -y * std::sin(th[k]) + offset[2];
      }
    return V;
  }
  // dCL/dalpha and dCL/dtwist [nspan] from Wing::windowAdjoint's vertex cotangent.
  double gradient(const std::vector<double> &vertexBar, double alpha, const std::vector<double> &twist,
                  std::vector<double> &dtwist) const
  {
    const std::vector<double> th = stationAngles(alpha, twist);
    std::vector<double>       dth(nspan + 1, 0.0);
    for (int64_t r = 0; r <= nchord; r++)
      for (int64_t k = 0; k <= nspan; k++) {
        const double  y = chord * (double)r / (double)nchord;
        const double *g = &vertexBar[3 * (r * (nspan + 1) + k)];
        dth[k] += g[1] * (-y * std::sin(th[k])) + g[2] * (-y * std::cos(th[k]));
      }
    double dalpha = 0.0;
    for (double d : dth) dalpha += d;
    dtwist.assign(nspan, 0.0);
    for (int64_t c = 0; c < nspan; c++)
      dtwist[c] = (c == 0 ? dth[0] : 0.5 * dth[c]) + (c == nspan - 1 ? dth[nspan] : 0.5 * dth[c + 1]);
    return dalpha;
  }
};

}  // namespace uvlm

#endif
