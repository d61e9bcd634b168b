// The reference driver's time loop (src/main.cpp:168-295) written against the C++ facades:
// a Vortex::FMMCalculator for the wake and a uvlm::Wing built from vertex couples the way
// main.cpp builds its Wing (rotated LE/TE couples).  argv: npanels steps out.bin; writes
// [steps][3] forces followed by the final particle records.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/uvlm/fmm_calculator.hpp"
#include "../../include/uvlm/vlm_wing.hpp"

struct HostTag {};
struct DeviceTag {};

int main(int argc, char **argv)
{
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
