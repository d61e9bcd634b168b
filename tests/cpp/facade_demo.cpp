// Drop-in check of include/uvlm/fmm_calculator.hpp, written the way the reference's
// driver uses Vortex::FMMCalculator (src/main.cpp:119, 292): instantiate with tags,
// init + calculate, then advance.  Reads exafmm::Body records from argv[1], writes the
// bodies after calculate() to argv[2] and after `steps` RK3 steps to argv[3].
#include <array>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/uvlm/fmm_calculator.hpp"

struct HostTag {};
struct DeviceTag {};

static exafmm::Bodies read_bodies(const char *path)
{
  FILE *f = std::fopen(path, "rb");
  if (!f) std::exit(3);
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  exafmm::Bodies b(bytes / sizeof(exafmm::Body));
  if (std::fread(b.data(), sizeof(exafmm::Body), b.size(), f) != b.size()) std::exit(4);
  std::fclose(f);
  return b;
}

static void write_bodies(const char *path, const exafmm::Bodies &b)
{
  FILE *f = std::fopen(path, "wb");
  if (!f) std::exit(5);
  std::fwrite(b.data(), sizeof(exafmm::Body), b.size(), f);
  std::fclose(f);
}

int main(int argc, char **argv)
{
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s in.bin eval.bin advance.bin steps\n", argv[0]);
    return 2;
  }
  typedef Vortex::FMMCalculator<HostTag, DeviceTag, Vortex::RK3, Vortex::Inviscid, Vortex::cVPM, Vortex::Transposed>
                 FMMCalculator;
  exafmm::Bodies particles = read_bodies(argv[1]);
  try {
    FMMCalculator fmm(8, 0.5, 64);
    fmm.init(particles);
    fmm.calculate();
    fmm.copyBodies2Host(particles);
    write_bodies(argv[2], particles);
    const int steps = std::atoi(argv[4]);
    for (int s = 0; s < steps; s++) fmm.advance(particles, 0.01);
    write_bodies(argv[3], particles);
    // calls beyond the reference class: invariants and a no-op relaxation (f = 0)
    fmm.init(particles);
    fmm.calculate();
    const std::array<double, 10> inv = fmm.diagnostics();
    fmm.relax(0.0);
    exafmm::Cells cells = exafmm::buildTree(particles, 64);
    std::printf("facade ok: %zu bodies, %zu cells, root NBODY %zu, energy %.6e\n", particles.size(), cells.size(),
                cells[0].NBODY, inv[6]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
