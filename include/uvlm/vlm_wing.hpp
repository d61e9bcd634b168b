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

private:
  uvlm_wing w_ = nullptr;
  int64_t   n_ = 0;
  static void check(int rc)
  {
    if (rc != UVLM_OK) throw std::runtime_error(std::string("uvlm: ") + uvlm_last_error());
  }
  void create(int device, int64_t nspan, int64_t nchord, const std::vector<double> &v)
  {
    check(uvlm_wing_create(device, nspan, nchord, v.data(), &w_));
    n_ = nspan * nchord;
  }
};

}  // namespace uvlm

#endif
