// The reference driver's time loop (src/main.cpp:168-295) written against the C++ facades:
// a Vortex::FMMCalculator for the wake and a uvlm::Wing built from vertex couples the way
// main.cpp builds its Wing (rotated LE/TE couples).  argv: npanels steps out.bin; writes
// [steps][3] forces followed by the final particle records.  argv: design out.bin runs the
// design gradient instead (design() below).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/uvlm/fmm_calculator.hpp"
#include "../../include/uvlm/vlm_wing.hpp"

struct HostTag {};
struct DeviceTag {};

// The design gradient from C++ (uvlm::WingDesign + Wing::windowAdjoint): a 6 x 2 lattice of
// span 3 at alpha = 5 degrees, 4 loop steps, then the window of 6 steps with checkpoints every 2.
// Writes CL, dCL/dalpha and dCL/dtwist [6].
static int design(const char *path)
{
  using Calc = Vortex::FMMCalculator<HostTag, DeviceTag, Vortex::RK2, Vortex::Inviscid, Vortex::cVPM, Vortex::Transposed>;
  uvlm::WingDesign d;
  d.nspan  = 6;
  d.nchord = 2;
  d.span   = 3.0;
  const double           alpha = 5.0 * M_PI / 180.0, dt = 0.05;
  const uvlm::Wing::Vec3 U{0.0, 1.0, 0.0}, e{0.0, 0.0, 1.0};
  Calc                   fmm(6, 0.5, 1024);
  uvlm::Wing             wing(d.nspan, d.nchord, d.vertices(alpha, {}));
  for (int t = 0; t < 4; t++) wing.step(fmm.handle(), U, dt);
  std::vector<double> vbar, dtwist;
  const double        CL     = wing.windowAdjoint(fmm.handle(), U, dt, 6, 2, e, vbar);
  const double        dalpha = d.gradient(vbar, alpha, {}, dtwist);
  std::vector<double> out{CL, dalpha};
  out.insert(out.end(), dtwist.begin(), dtwist.end());
  FILE *f = std::fopen(path, "wb");
  if (!f) return 5;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  std::printf("design: CL %.9f dCL/dalpha %.9f\n", CL, dalpha);
  return 0;
}

int main(int argc, char **argv)
{
  if (argc == 3 && std::string(argv[1]) == "design") {
    try {
      return design(argv[2]);
    } catch (const std::exception &e) {
      std::fprintf(stderr, "error: %s\n", e.what());
      return 1;
    }
  }
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s npanels steps out.bin\n", argv[0]);
    return 2;
  }
  const int    npanels = std::atoi(argv[1]), steps = std::atoi(argv[2]);
  const double span = 10.0, chord = 0.4, angle = -20.0 * M_PI / 180.0, dt = 0.05;
  const double dx = span / (npanels + 1);
  std::vector<uvlm::Wing::Vec3> le, te;
  for (int i = 0; i <= npanels; i++) {  // rotation about x by -20 degrees (main.cpp:172-188)
    const double y0 = 0.0, y1 = chord;
    le.push_back({i * dx, std::cos(angle) * y0, std::sin(angle) * y0});
    te.push_back({i * dx, std::cos(angle) * y1, std::sin(angle) * y1});
  }
  using Calc = Vortex::FMMCalculator<HostTag, DeviceTag, Vortex::RK2, Vortex::Inviscid, Vortex::cVPM, Vortex::Transposed>;
  try {
    Calc       fmm(6, 0.5, 256);
    uvlm::Wing wing(le, te);
    const uvlm::Wing::Vec3 U{0.0, 1.0, 0.0};
    std::vector<double>    forces;
    for (int t = 0; t < steps; t++) {
      const auto F = wing.step(fmm.handle(), U, dt);
      forces.insert(forces.end(), F.begin(), F.end());
    }
    exafmm::Bodies particles;
    fmm.copyBodies2Host(particles);
    FILE *f = std::fopen(argv[3], "wb");
    if (!f) return 5;
    std::fwrite(forces.data(), sizeof(double), forces.size(), f);
    std::fwrite(particles.data(), sizeof(exafmm::Body), particles.size(), f);
    std::fclose(f);
    std::printf("%d steps, %zu particles, last force %.9f %.9f %.9f\n", steps, particles.size(), forces[3 * steps - 3],
                forces[3 * steps - 2], forces[3 * steps - 1]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
