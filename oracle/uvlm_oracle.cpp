// uvlm_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// A CPU restatement, in FP64 C++17 + OpenMP, of the FMM vortex-particle path of
// kursatyurt/UVLM: octree build, dual-tree traversal, the exafmm spherical-harmonic
// translations (P2M/M2M/M2L/L2L/L2P), the regularised Biot-Savart P2P (velocity and
// velocity gradient), stretching, the rVPM correction, the RK integrators and the
// sensor (probe) evaluation.  Every function cites the reference file:line it follows.
//
// Role: the checker.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load this library; the product (uvlm_amd/) never links or calls it.
//
// Semantics kept on purpose (they define parity):
//  * compile-time constants P / ncrit / MAC_theta of the reference are runtime here;
//  * the root multipole is never formed (fmmCalculator.hpp:427 loop stops at level 1);
//  * M2L only sees singular harmonics of degree < P (multipoleKernels.hpp:521,527:
//    the 4P^2 buffer is value-initialised and evalLocal fills degrees < P only);
//  * L2P computes the velocity gradient by complex step, h = 1e-32, on bicomplex
//    harmonics (multipoleKernels.hpp:630-742);
//  * RK4 stage 3 advances with 0.5*dt (fmmCalculator.hpp:683-689), as written;
//  * the sensor P2P adds velocity only, the sensor J is not reset (fmmCalculator.hpp:380-389).
// Differences that are NOT semantic: FP atomics of the reference are replaced by
// per-target gathers (deterministic sums; order differs at the ULP level); the traversal
// stack is a growable vector instead of a fixed 128-entry array (traverse.hpp:46).
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off -fopenmp).  -ffp-contract=off keeps
// every expression rounded as written, which the bit-exact tree and list checks rely on.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>
#include <omp.h>

namespace orc {

using real_t    = double;
using complex_t = std::complex<double>;

// exafmm::Body (exafmmTypes.hpp:18-32): 37 doubles, 296 bytes, same field order.
struct Body {
  real_t J[3][3];
  real_t X[3];
  real_t alpha[3];
  real_t velocity[3];
  real_t dadt[3];
  real_t velocity_old[3];
  real_t dadt_old[3];
  real_t X_old[3];
  real_t alpha_old[3];
  real_t radius;
  real_t drdt;
  real_t radius_old;
  real_t drdt_old;
};
static_assert(sizeof(Body) == 296, "Body layout must match exafmm::Body");

// exafmm::Cell (exafmmTypes.hpp:36-46) without the raw pointers.
struct Cell {
  int64_t NCHILD;
  int64_t NBODY;
  int64_t ChildOffset;
  int64_t BodyOffset;
  real_t  X[3];
  real_t  R;
  int64_t level;
};

enum { ORC_OK = 0, ORC_ERR_ARG = 1, ORC_ERR_DEPTH = 2, ORC_ERR_STATE = 3 };

// ---------------------------------------------------------------------------------
// Small helpers: multipoleKernels.hpp:10-25
static inline real_t norm3(const real_t *X) { return X[0] * X[0] + X[1] * X[1] + X[2] * X[2]; }
static inline int    oddOrEven(int n) { return ((n & 1) == 1) ? -1 : 1; }
static inline int    ipow2n(int n) { return (n >= 0) ? 1 : oddOrEven(n); }

// ---------------------------------------------------------------------------------
// Bounds: tree.hpp:7-26
static void get_bounds(const Body *bodies, int64_t n, real_t &R0, real_t *X0)
{
  real_t Xmin[3], Xmax[3];
  for (int d = 0; d < 3; d++) Xmin[d] = Xmax[d] = bodies[0].X[d];
  for (int64_t b = 0; b < n; b++) {
    for (int d = 0; d < 3; d++) Xmin[d] = std::fmin(bodies[b].X[d], Xmin[d]);
    for (int d = 0; d < 3; d++) Xmax[d] = std::fmax(bodies[b].X[d], Xmax[d]);
  }
  for (int d = 0; d < 3; d++) X0[d] = (Xmax[d] + Xmin[d]) / 2;
  R0 = 0;
  for (int d = 0; d < 3; d++) {
    R0 = std::fmax(X0[d] - Xmin[d], R0);
    R0 = std::fmax(Xmax[d] - X0[d], R0);
  }
  R0 *= 1.00001;
}

// ---------------------------------------------------------------------------------
// Octree: tree.hpp:29-113.  Adaptive top-down split, leaf iff NBODY <= ncrit (43),
// octant = (x>cx) + 2(y>cy) + 4(z>cz) (57), stable counting sort per level (70-78),
// child centre = parent centre + (R0 / 2^(l+1)) * (+-1) (86-92), child groups appended
// in DFS pre-order of their parents (81-96).  Implemented on an index permutation that
// is applied to the bodies at the end; the resulting order is the same.
struct TreeBuilder {
  const Body          *bodies;
  std::vector<int64_t> perm, tmp;
  std::vector<Cell>   *cells;
  int64_t              ncrit;
  real_t               R0;
  int                  status = ORC_OK;

  void build(int64_t ci, int64_t begin, int64_t end, const real_t *X, int level)
  {
    Cell &c  = (*cells)[ci];
    c.NBODY  = end - begin;
    c.NCHILD = 0;
    c.ChildOffset = 0;
    c.BodyOffset  = begin;
    for (int d = 0; d < 3; d++) c.X[d] = X[d];
    c.R     = R0 / (real_t)(int64_t(1) << level);
    c.level = level;
    if (end - begin <= ncrit) return;
    if (level >= 60) { status = ORC_ERR_DEPTH; return; }
    int64_t size[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    auto octant_of = [&](int64_t b) {
      const real_t *x = bodies[b].X;
      return (x[0] > X[0]) + ((x[1] > X[1]) << 1) + ((x[2] > X[2]) << 2);
    };
    for (int64_t i = begin; i < end; i++) size[octant_of(perm[i])]++;
    int64_t offsets[8], counter[8], off = begin, nchild = 0;
    for (int i = 0; i < 8; i++) {
      offsets[i] = off;
      off += size[i];
      if (size[i]) nchild++;
    }
    for (int i = 0; i < 8; i++) counter[i] = offsets[i];
    for (int64_t i = begin; i < end; i++) {
      const int64_t b = perm[i];
      tmp[counter[octant_of(b)]++] = b;
    }
    for (int64_t i = begin; i < end; i++) perm[i] = tmp[i];
    const int64_t first = (int64_t)cells->size();
    cells->resize(first + nchild);
    (*cells)[ci].NCHILD      = nchild;
    (*cells)[ci].ChildOffset = first;
    int64_t c_idx = 0;
    for (int i = 0; i < 8; i++) {
      real_t       Xc[3];
      const real_t r = R0 / (real_t)(int64_t(1) << (level + 1));
      for (int d = 0; d < 3; d++) Xc[d] = X[d];
      for (int d = 0; d < 3; d++) Xc[d] += r * (real_t)(((i & 1 << d) >> d) * 2 - 1);
      if (size[i]) {
        build(first + c_idx, offsets[i], offsets[i] + size[i], Xc, level + 1);
        if (status != ORC_OK) return;
        c_idx++;
      }
    }
  }
};

// Builds the tree and permutes `bodies` in place into tree order (tree.hpp:101-113).
static int build_tree(Body *bodies, int64_t n, int64_t ncrit, std::vector<Cell> &cells, real_t *X0_out = nullptr,
                      real_t *R0_out = nullptr)
{
  if (n <= 0) return ORC_ERR_ARG;
  real_t R0, X0[3];
  get_bounds(bodies, n, R0, X0);
  if (X0_out) for (int d = 0; d < 3; d++) X0_out[d] = X0[d];
  if (R0_out) *R0_out = R0;
  TreeBuilder tb;
  tb.bodies = bodies;
  tb.perm.resize(n);
  tb.tmp.resize(n);
  for (int64_t i = 0; i < n; i++) tb.perm[i] = i;
  cells.assign(1, Cell{});
  cells.reserve(n);
  tb.cells = &cells;
  tb.ncrit = ncrit;
  tb.R0    = R0;
  tb.build(0, 0, n, X0, 0);
  if (tb.status != ORC_OK) return tb.status;
  std::vector<Body> sorted(n);
  for (int64_t i = 0; i < n; i++) sorted[i] = bodies[tb.perm[i]];
  std::memcpy(bodies, sorted.data(), sizeof(Body) * n);
  return ORC_OK;
}

// ---------------------------------------------------------------------------------
// Dual-tree traversal: traverse.hpp:32-85.  Near iff |dX|^2 theta^2 < (Ri+Rj)^2 (56-57);
// near leaf-leaf -> P2P, near otherwise -> push, far -> M2L.  Split the target if the
// source is a leaf or (Rt >= Rs and the target is not a leaf), else split the source (74).
// Two leaf roots give the single P2P pair (0,0) (41-44).
// gap > 0 (an option of this build, off in the reference): a pair is also near when its bodies
// can come closer than gap, |dX| < gap + sqrt(3) (Ri + Rj) (cells are cubes of half-width R),
// tested on squares so the decision rounds the same on every machine.
struct Pair { int64_t t, s; };

static constexpr real_t SQRT3 = 1.7320508075688772;

static void get_list(const std::vector<Cell> &tc, const std::vector<Cell> &sc, real_t theta, real_t gap,
                     std::vector<Pair> &P2P, std::vector<Pair> &M2L)
{
  P2P.clear();
  M2L.clear();
  if (tc[0].NCHILD == 0 && sc[0].NCHILD == 0) {
    P2P.push_back({0, 0});
    return;
  }
  std::vector<Pair> stack;
  stack.push_back({0, 0});
  auto interact = [&](int64_t i, int64_t j) {
    const Cell &Ci = tc[i];
    const Cell &Cj = sc[j];
    real_t      dX[3];
    for (int d = 0; d < 3; d++) dX[d] = Ci.X[d] - Cj.X[d];
    const real_t D2 = norm3(dX), R2 = D2 * theta * theta, Rs = Ci.R + Cj.R;
    const real_t Gs = gap + SQRT3 * Rs;
    if (R2 < Rs * Rs || (gap > 0 && D2 < Gs * Gs)) {
      if (Ci.NCHILD == 0 && Cj.NCHILD == 0) P2P.push_back({i, j});
      else stack.push_back({i, j});
    } else {
      M2L.push_back({i, j});
    }
  };
  while (!stack.empty()) {
    const Pair p = stack.back();
    stack.pop_back();
    const Cell &T = tc[p.t];
    const Cell &S = sc[p.s];
    if (S.NCHILD == 0 || (T.R >= S.R && T.NCHILD != 0)) {
      for (int64_t c = T.ChildOffset; c < T.ChildOffset + T.NCHILD; c++) interact(c, p.s);
    } else {
      for (int64_t c = S.ChildOffset; c < S.ChildOffset + S.NCHILD; c++) interact(p.t, c);
    }
  }
}

// ---------------------------------------------------------------------------------
// Spherical coordinates and solid harmonics: multipoleKernels.hpp:27-172
static inline void cart2sph(const real_t *dX, real_t &r, real_t &theta, real_t &phi)
{
  r     = std::sqrt(norm3(dX));
  theta = r == 0.0 ? real_t(0) : std::acos(dX[2] / r);
  phi   = std::atan2(dX[1], dX[0]);
}

// Regular solid harmonics r^n Y_n^m and their theta derivative (multipoleKernels.hpp:53-92).
// Y[n*n+n+m], m in [-n,n]; Y_n^m = (-1)^n rho^n P_n^m(cos a) e^{i m b} / (n+m)!.
static void evalMultipole(int P, real_t rho, real_t alpha, real_t beta, complex_t *Ynm, complex_t *YnmTheta)
{
  const real_t x    = std::cos(alpha);
  const real_t y    = std::sin(alpha);
  const real_t invY = y == 0 ? 0 : 1 / y;
  real_t       fact = 1, pn = 1, rhom = 1;
  const complex_t ei(std::cos(beta), std::sin(beta));
  complex_t       eim = 1.0;
  for (int m = 0; m < P; m++) {
    real_t    p   = pn;
    const int npn = m * m + 2 * m;
    const int nmn = m * m;
    Ynm[npn]      = rhom * p * eim;
    Ynm[nmn]      = std::conj(Ynm[npn]);
    real_t p1     = p;
    p             = x * (2 * m + 1) * p1;
    if (YnmTheta) YnmTheta[npn] = rhom * (p - (m + 1) * x * p1) * invY * eim;
    rhom *= rho;
    real_t rhon = rhom;
    for (int n = m + 1; n < P; n++) {
      const int npm = n * n + n + m;
      const int nmm = n * n + n - m;
      rhon /= -(n + m);
      Ynm[npm]  = rhon * p * eim;
      Ynm[nmm]  = std::conj(Ynm[npm]);
      real_t p2 = p1;
      p1        = p;
      p         = (x * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
      if (YnmTheta) YnmTheta[npm] = rhon * ((n - m + 1) * p - (n + 1) * x * p1) * invY * eim;
      rhon *= rho;
    }
    rhom /= -(2 * m + 2) * (2 * m + 1);
    pn = -pn * fact * y;
    fact += 2;
    eim *= ei;
  }
}

// Singular solid harmonics r^{-n-1} Y_n^m, degrees < P only (multipoleKernels.hpp:137-172).
// Y_n^m = (-1)^n (n-m)! P_n^m(cos a) e^{i m b} / rho^{n+1}.
static void evalLocal(int P, real_t rho, real_t alpha, real_t beta, complex_t *Ynm)
{
  const real_t    x    = std::cos(alpha);
  const real_t    y    = std::sin(alpha);
  real_t          fact = 1, pn = 1;
  const real_t    invR = -1.0 / rho;
  real_t          rhom = -invR;
  const complex_t ei(std::cos(beta), std::sin(beta));
  complex_t       eim = 1.0;
  for (int m = 0; m < P; m++) {
    real_t p   = pn;
    int    npn = m * m + 2 * m;
    int    nmn = m * m;
    Ynm[npn]   = rhom * p * eim;
    Ynm[nmn]   = std::conj(Ynm[npn]);
    real_t p1  = p;
    p          = x * (2 * m + 1) * p1;
    rhom *= invR;
    real_t rhon = rhom;
    for (int n = m + 1; n < P; n++) {
      int npm   = n * n + n + m;
      int nmm   = n * n + n - m;
      Ynm[npm]  = rhon * p * eim;
      Ynm[nmm]  = std::conj(Ynm[npm]);
      real_t p2 = p1;
      p1        = p;
      p         = (x * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
      rhon *= invR * (n - m + 1);
    }
    pn = -pn * fact * y;
    fact += 2;
    eim *= ei;
  }
}

// ---------------------------------------------------------------------------------
// Complex-step scalar and bicomplex numbers used by L2P (complexify.hpp:28-519):
// complex arithmetic on the step axis, comparisons on real parts, first-order
// acos (complexify.hpp:437-441) and atan2 (448-453).
using cplx = std::complex<double>;
static inline cplx cs_acos(const cplx &z)
{
  return cplx(std::acos(z.real()), -z.imag() / std::sqrt(1.0 - z.real() * z.real() + 1.e-24));
}
static inline cplx cs_atan2(const cplx &z1, const cplx &z2)
{
  return cplx(std::atan2(z1.real(), z2.real()),
              (z2.real() * z1.imag() - z1.real() * z2.imag()) / (z1.real() * z1.real() + z2.real() * z2.real()));
}
struct multicomplex { cplx A, B; };  // A + jB; j is the harmonic-phase axis (complexify.hpp:470-474)
static inline multicomplex mc(cplx A, cplx B) { return {A, B}; }
static inline multicomplex product(const multicomplex &V, const multicomplex &W)
{
  return mc(V.A * W.A - V.B * W.B, V.A * W.B + V.B * W.A);
}
static inline multicomplex product(const multicomplex &V, const cplx &X) { return mc(V.A * X, V.B * X); }
static inline multicomplex conjugate(const multicomplex &V) { return mc(V.A, -V.B); }
static inline multicomplex from_C1(const complex_t &C) { return mc(cplx(C.real(), 0), cplx(C.imag(), 0)); }

static inline void cart2sph_cs(const cplx *dX, cplx &r, cplx &theta, cplx &phi)
{
  r     = std::sqrt(dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2]);
  theta = r.real() == 0.0 ? cplx(0) : cs_acos(dX[2] / r);
  phi   = cs_atan2(dX[1], dX[0]);
}

// Bicomplex regular harmonics (multipoleKernels.hpp:94-134).
static void evalMultipole_cs(int P, cplx rho, cplx alpha, cplx beta, multicomplex *Ynm, multicomplex *YnmTheta)
{
  const double       EPS  = 1e-16;
  const cplx         x    = std::cos(alpha);
  const cplx         y    = std::sin(alpha);
  const cplx         invY = (std::sin(alpha).real() <= EPS) ? cplx(0, 0) : 1.0 / y;
  cplx               fact = 1, pn = 1, rhom = 1;
  const multicomplex ei  = mc(std::cos(beta), std::sin(beta));
  multicomplex       eim = mc(1.0, 0.0);
  for (int m = 0; m < P; m++) {
    cplx p        = pn;
    int  npn      = m * m + 2 * m;
    int  nmn      = m * m;
    Ynm[npn]      = product(eim, rhom * p);
    Ynm[nmn]      = conjugate(Ynm[npn]);
    cplx p1       = p;
    p             = x * (double)(2 * m + 1) * p1;
    YnmTheta[npn] = product(eim, rhom * (p - (double)(m + 1) * x * p1) * invY);
    rhom *= rho;
    cplx rhon = rhom;
    for (int n = m + 1; n < P; n++) {
      int npm = n * n + n + m;
      int nmm = n * n + n - m;
      rhon /= (double)(-(n + m));
      Ynm[npm]      = product(eim, rhon * p);
      Ynm[nmm]      = conjugate(Ynm[npm]);
      cplx p2       = p1;
      p1            = p;
      p             = (x * (double)(2 * n + 1) * p1 - (double)(n + m) * p2) / (double)(n - m + 1);
      YnmTheta[npm] = product(eim, rhon * ((double)(n - m + 1) * p - (double)(n + 1) * x * p1) * invY);
      rhon *= rho;
    }
    rhom /= (double)(-(2 * m + 2) * (2 * m + 1));
    pn = -pn * fact * y;
    fact += 2.0;
    eim = product(eim, ei);
  }
}

template <class T1, class T2>
static inline void sph2cart(T1 r, T1 theta, T1 phi, const T2 *s, T2 *c)
{
  // multipoleKernels.hpp:41-50
  c[0] = std::sin(theta) * std::cos(phi) * s[0] + std::cos(theta) * std::cos(phi) / r * s[1] -
         std::sin(phi) / r / std::sin(theta) * s[2];
  c[1] = std::sin(theta) * std::sin(phi) * s[0] + std::cos(theta) * std::sin(phi) / r * s[1] +
         std::cos(phi) / r / std::sin(theta) * s[2];
  c[2] = std::cos(theta) * s[0] - std::sin(theta) / r * s[1];
}

// ---------------------------------------------------------------------------------
// Translation operators.  Expansion storage: E[c*NTERM + 3*(n(n+1)/2+m) + d], m in [0,n].
struct Ops {
  int P, NTERM;
  explicit Ops(int P_) : P(P_), NTERM(3 * (P_ * (P_ + 1) / 2)) {}

  // P2M: multipoleKernels.hpp:452-477
  void p2m(const real_t *Xc, const Body *b, int64_t nb, complex_t *M) const
  {
    std::vector<complex_t> Y(P * P), Yt(P * P);
    for (int64_t i = 0; i < nb; i++) {
      real_t dX[3];
      for (int d = 0; d < 3; d++) dX[d] = b[i].X[d] - Xc[d];
      real_t rho, alpha, beta;
      cart2sph(dX, rho, alpha, beta);
      evalMultipole(P, rho, alpha, beta, Y.data(), Yt.data());
      for (int n = 0; n < P; n++)
        for (int m = 0; m <= n; m++) {
          const int nm = n * n + n - m, nms = n * (n + 1) / 2 + m;
          for (int d = 0; d < 3; d++) M[3 * nms + d] += b[i].alpha[d] * Y[nm];
        }
    }
  }

  // M2M of one child into its parent: multipoleKernels.hpp:479-516
  void m2m(const real_t *Xp, const real_t *Xch, const complex_t *Mc, complex_t *Mp) const
  {
    std::vector<complex_t> Y(P * P), Yt(P * P);
    real_t                 dX[3];
    for (int d = 0; d < 3; d++) dX[d] = Xp[d] - Xch[d];
    real_t rho, alpha, beta;
    cart2sph(dX, rho, alpha, beta);
    evalMultipole(P, rho, alpha, beta, Y.data(), Yt.data());
    for (int j = 0; j < P; j++)
      for (int k = 0; k <= j; k++) {
        const int jks = j * (j + 1) / 2 + k;
        complex_t acc[3] = {0, 0, 0};
        for (int n = 0; n <= j; n++) {
          for (int m = std::max(-n, -j + k + n); m <= std::min(k - 1, n); m++) {
            const int jnkms = (j - n) * (j - n + 1) / 2 + k - m, nm = n * n + n - m;
            for (int d = 0; d < 3; d++) acc[d] += Mc[3 * jnkms + d] * Y[nm] * real_t(ipow2n(m) * oddOrEven(n));
          }
          for (int m = k; m <= std::min(n, j + k - n); m++) {
            const int jnkms = (j - n) * (j - n + 1) / 2 - k + m, nm = n * n + n - m;
            for (int d = 0; d < 3; d++)
              acc[d] += std::conj(Mc[3 * jnkms + d]) * Y[nm] * real_t(oddOrEven(k + n + m));
          }
        }
        for (int d = 0; d < 3; d++) Mp[3 * jks + d] += acc[d];
      }
  }

  // M2L: multipoleKernels.hpp:518-552.  Ynm2 has 4P^2 entries, zero beyond degree P-1.
  // harm_order > 0 fills the singular harmonics up to degree harm_order-1 instead (at most
  // 2P-1): a deliberately different truncation that only the tests use, to show that the
  // truncation check tells the two apart.
  void m2l(const real_t *Xt, const real_t *Xs, const complex_t *Ms, complex_t *Lt, int harm_order = 0) const
  {
    std::vector<complex_t> Y(4 * P * P, complex_t(0, 0));
    real_t                 dX[3];
    for (int d = 0; d < 3; d++) dX[d] = Xt[d] - Xs[d];
    real_t rho, alpha, beta;
    cart2sph(dX, rho, alpha, beta);
    evalLocal(harm_order > 0 ? std::min(harm_order, 2 * P - 1) : P, rho, alpha, beta, Y.data());
    for (int j = 0; j < P; j++) {
      const real_t Cnm = oddOrEven(j);
      for (int k = 0; k <= j; k++) {
        const int jks = j * (j + 1) / 2 + k;
        complex_t acc[3] = {0, 0, 0};
        for (int n = 0; n < P; n++) {
          for (int m = -n; m < 0; m++) {
            const int nms = n * (n + 1) / 2 - m, jnkm = (j + n) * (j + n) + j + n + m - k;
            for (int d = 0; d < 3; d++) acc[d] += std::conj(Ms[3 * nms + d]) * Cnm * Y[jnkm];
          }
          for (int m = 0; m <= n; m++) {
            const int    nms = n * (n + 1) / 2 + m, jnkm = (j + n) * (j + n) + j + n + m - k;
            const real_t Cnm2 = Cnm * oddOrEven((k - m) * (k < m) + m);
            for (int d = 0; d < 3; d++) acc[d] += Ms[3 * nms + d] * Cnm2 * Y[jnkm];
          }
        }
        for (int d = 0; d < 3; d++) Lt[3 * jks + d] += acc[d];
      }
    }
  }

  // L2L from a parent to one child: multipoleKernels.hpp:591-628
  void l2l(const real_t *Xp, const real_t *Xch, const complex_t *Lp, complex_t *Lc) const
  {
    std::vector<complex_t> Y(P * P), Yt(P * P);
    real_t                 dX[3];
    for (int d = 0; d < 3; d++) dX[d] = Xch[d] - Xp[d];
    real_t rho, alpha, beta;
    cart2sph(dX, rho, alpha, beta);
    evalMultipole(P, rho, alpha, beta, Y.data(), Yt.data());
    for (int j = 0; j < P; j++)
      for (int k = 0; k <= j; k++) {
        const int jks = j * (j + 1) / 2 + k;
        complex_t acc[3] = {0, 0, 0};
        for (int n = j; n < P; n++) {
          for (int m = j + k - n; m < 0; m++) {
            const int jnkm = (n - j) * (n - j) + n - j + m - k, nms = n * (n + 1) / 2 - m;
            for (int d = 0; d < 3; d++) acc[d] += std::conj(Lp[3 * nms + d]) * Y[jnkm] * real_t(oddOrEven(k));
          }
          for (int m = 0; m <= n; m++) {
            if (n - j >= std::abs(m - k)) {
              const int jnkm = (n - j) * (n - j) + n - j + m - k, nms = n * (n + 1) / 2 + m;
              for (int d = 0; d < 3; d++) acc[d] += Lp[3 * nms + d] * Y[jnkm] * real_t(oddOrEven((m - k) * (m < k)));
            }
          }
        }
        for (int d = 0; d < 3; d++) Lc[3 * jks + d] += acc[d];
      }
  }

  // L2P with complex-step velocity gradient: multipoleKernels.hpp:630-742.
  // Adds (1/4pi) curl(psi) to velocity and its x-derivative (J[a][b] = du_a/dx_b) to J.
  void l2p(const real_t *Xc, const complex_t *L, Body &body) const
  {
    const real_t              h = 1e-32;
    std::vector<multicomplex> Y(P * P), Yt(P * P);
    real_t                    vel[3] = {0, 0, 0};
    real_t                    J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const multicomplex        Iunit   = from_C1(complex_t(0., 1.));
    for (int H = 0; H < 3; H++) {
      cplx dX[3];
      for (int d = 0; d < 3; d++) dX[d] = body.X[d] - Xc[d];
      dX[H] += cplx(0, h);
      cplx s[3][3];
      for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++) s[a][d] = 0;
      cplx r, theta, phi;
      cart2sph_cs(dX, r, theta, phi);
      evalMultipole_cs(P, r, theta, phi, Y.data(), Yt.data());
      for (int n = 0; n < P; n++) {
        int nm  = n * n + n;
        int nms = n * (n + 1) / 2;
        for (int a = 0; a < 3; a++) {
          const multicomplex Lc = from_C1(L[3 * nms + a]);
          s[a][0] += product(Lc, Y[nm]).A / r * (double)n;
          s[a][1] += product(Lc, Yt[nm]).A;
        }
        for (int m = 1; m <= n; m++) {
          nm  = n * n + n + m;
          nms = n * (n + 1) / 2 + m;
          for (int a = 0; a < 3; a++) {
            const multicomplex Lc = from_C1(L[3 * nms + a]);
            s[a][0] += 2.0 * product(Lc, Y[nm]).A / r * (double)n;
            s[a][1] += 2.0 * product(Lc, Yt[nm]).A;
            s[a][2] += 2.0 * product(product(Lc, Y[nm]), Iunit).A * (double)m;
          }
        }
      }
      cplx c[3];
      // grad psi_x -> (u_z -= d_y psi_x, u_y += d_z psi_x), multipoleKernels.hpp:684-697
      sph2cart(r, theta, phi, s[0], c);
      if (H == 0) { vel[2] -= c[1].real(); vel[1] += c[2].real(); }
      J[1][H] += c[2].imag() / h;
      J[2][H] -= c[1].imag() / h;
      // grad psi_y, multipoleKernels.hpp:701-714
      sph2cart(r, theta, phi, s[1], c);
      if (H == 0) { vel[2] += c[0].real(); vel[0] -= c[2].real(); }
      J[0][H] -= c[2].imag() / h;
      J[2][H] += c[0].imag() / h;
      // grad psi_z, multipoleKernels.hpp:718-731
      sph2cart(r, theta, phi, s[2], c);
      if (H == 0) { vel[1] -= c[0].real(); vel[0] += c[1].real(); }
      J[0][H] += c[1].imag() / h;
      J[1][H] -= c[0].imag() / h;
    }
    const real_t inv4pi = 0.25 / M_PI;
    for (int d = 0; d < 3; d++) body.velocity[d] += inv4pi * vel[d];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) body.J[a][b] += inv4pi * J[a][b];
  }
};

// ---------------------------------------------------------------------------------
// Near field.  Device P2P with velocity and gradient: multipoleKernels.hpp:223-301.
// Accumulates into the 12-wide partial sum `acc` (u[3], J[3][3] row-major); the caller
// subtracts it, as the reference does (287-298).
static inline void p2p_pair_sum(const real_t *xi, const Body &bj, real_t *acc)
{
  real_t dX[3];
  for (int d = 0; d < 3; d++) dX[d] = xi[d] - bj.X[d];
  const real_t R2 = norm3(dX);
  if (R2 == 0) return;
  const real_t S2            = bj.radius * bj.radius;
  const real_t RS            = R2 / S2 / 2.0;
  const real_t sqrtRS        = std::sqrt(RS);
  const real_t erfSqrtRS     = std::erf(sqrtRS);
  const real_t expmRS        = std::exp(-RS);
  const real_t invSqrtR2     = 1.0 / std::sqrt(R2);
  const real_t inv_sqrtpi    = 0.564189583547756286948079451561;
  const real_t inv_pi        = 0.318309886183790671537767526745;
  const real_t t2            = 2.0 * inv_sqrtpi * sqrtRS * expmRS;
  const real_t one_4piR3     = 0.25 * inv_pi / R2 * invSqrtR2;
  const real_t *a            = bj.alpha;
  const real_t c1            = one_4piR3 * (dX[1] * a[2] - dX[2] * a[1]);
  const real_t c2            = one_4piR3 * (dX[2] * a[0] - dX[0] * a[2]);
  const real_t c3            = one_4piR3 * (dX[0] * a[1] - dX[1] * a[0]);
  const real_t g             = erfSqrtRS - t2;
  acc[0] += c1 * g;
  acc[1] += c2 * g;
  acc[2] += c3 * g;
  real_t aux = t2 / S2 - 3.0 * g / R2;
  acc[3] += aux * c1 * dX[0];
  acc[4] += aux * c1 * dX[1];
  acc[5] += aux * c1 * dX[2];
  acc[6] += aux * c2 * dX[0];
  acc[7] += aux * c2 * dX[1];
  acc[8] += aux * c2 * dX[2];
  acc[9] += aux * c3 * dX[0];
  acc[10] += aux * c3 * dX[1];
  acc[11] += aux * c3 * dX[2];
  aux = g * one_4piR3;
  acc[6] -= aux * a[2];
  acc[9] += aux * a[1];
  acc[4] += aux * a[2];
  acc[10] -= aux * a[0];
  acc[5] -= aux * a[1];
  acc[8] += aux * a[0];
}

// Sensor P2P, velocity only: multipoleKernels.hpp:303-341.
static inline void p2p_sensor_pair_sum(const real_t *xi, const Body &bj, real_t *acc)
{
  real_t dX[3];
  for (int d = 0; d < 3; d++) dX[d] = xi[d] - bj.X[d];
  const real_t R2 = norm3(dX);
  if (R2 == 0) return;
  const real_t S2        = 2.0 * bj.radius * bj.radius;
  const real_t RS        = R2 / S2;
  const real_t sqrtRS    = std::sqrt(RS);
  const real_t erfSqrtRS = std::erf(sqrtRS);
  const real_t expmRS    = std::exp(-RS);
  const real_t invSqrtR2 = 1.0 / std::sqrt(R2);
  const real_t t2        = 2.0 * 0.564189583547756286948079451561 * sqrtRS * expmRS;
  const real_t cutoff    = 0.25 * 0.318309886183790671537767526745 / R2 * invSqrtR2 * (erfSqrtRS - t2);
  const real_t *a        = bj.alpha;
  acc[0] += (dX[1] * a[2] - dX[2] * a[1]) * cutoff;
  acc[1] += (dX[2] * a[0] - dX[0] * a[2]) * cutoff;
  acc[2] += (dX[0] * a[1] - dX[1] * a[0]) * cutoff;
}

// Host direct sum with classic stretching: multipoleKernels.hpp:174-221 (the reference's
// second, independent formula for u and J*alpha).
static inline void p2p_classic_pair(const Body &bi, const Body &bj, real_t *u, real_t *s)
{
  real_t dX[3];
  for (int d = 0; d < 3; d++) dX[d] = bi.X[d] - bj.X[d];
  const real_t R2 = norm3(dX);
  if (R2 == 0) return;
  const real_t S2     = 2 * bj.radius * bj.radius;
  const real_t RS     = R2 / S2;
  real_t       cutoff = 0.25 / M_PI / R2 / std::sqrt(R2) * (std::erf(std::sqrt(RS)) - std::sqrt(4 / M_PI * RS) * std::exp(-RS));
  const real_t *a = bj.alpha, *ai = bi.alpha;
  u[0] += (dX[1] * a[2] - dX[2] * a[1]) * cutoff;
  u[1] += (dX[2] * a[0] - dX[0] * a[2]) * cutoff;
  u[2] += (dX[0] * a[1] - dX[1] * a[0]) * cutoff;
  s[0] += (ai[1] * a[2] - ai[2] * a[1]) * cutoff;
  s[1] += (ai[2] * a[0] - ai[0] * a[2]) * cutoff;
  s[2] += (ai[0] * a[1] - ai[1] * a[0]) * cutoff;
  cutoff = 0.25 / M_PI / R2 / R2 / std::sqrt(R2) *
           (3 * std::erf(std::sqrt(RS)) - (2 * RS + 3) * std::sqrt(4 / M_PI * RS) * std::exp(-RS)) *
           (ai[0] * dX[0] + ai[1] * dX[1] + ai[2] * dX[2]);
  s[0] += (a[1] * dX[2] - a[2] * dX[1]) * cutoff;
  s[1] += (a[2] * dX[0] - a[0] * dX[2]) * cutoff;
  s[2] += (a[0] * dX[1] - a[1] * dX[0]) * cutoff;
}

// ---------------------------------------------------------------------------------
// The calculator: fmmCalculator.hpp:25-1006 (Host/Device tags dropped; viscous schemes,
// RBF re-initialisation and the VLM caller are outside this path).
enum { STRETCH_TRANSPOSED = 0, STRETCH_CLASSIC = 1 };
enum { FORM_CVPM = 0, FORM_RVPM = 1 };
enum { SCHEME_EULER = 0, SCHEME_RK2 = 1, SCHEME_RK3 = 2, SCHEME_RK4 = 3 };

struct CSR {
  std::vector<int64_t> off, idx;
  void build(int64_t ntarget, const std::vector<Pair> &pairs)
  {
    off.assign(ntarget + 1, 0);
    for (auto &p : pairs) off[p.t + 1]++;
    for (int64_t i = 0; i < ntarget; i++) off[i + 1] += off[i];
    idx.resize(pairs.size());
    std::vector<int64_t> cur(off.begin(), off.end() - 1);
    for (auto &p : pairs) idx[cur[p.t]++] = p.s;  // stable: list order kept per target
  }
};

struct Calculator {
  int               P, NTERM;
  real_t            theta;
  int64_t           ncrit;
  int               stretching, formulation;
  real_t            dt = 0;
  int               viscous = 0;  // 0 Inviscid, 1 PSE, 2 CSM (fmmCalculator.hpp:9-13)
  real_t            nu      = 0;  // _nu (fmmCalculator.hpp:56; never initialised there)
  std::vector<Body> bodies;
  std::vector<Cell> cells;
  std::vector<Pair> p2p, m2l;
  std::vector<complex_t>               M, L;
  std::vector<std::vector<int64_t>>    levels;  // fmmCalculator.hpp:75-79
  bool              ready = false;
  real_t            near_c = 0, gap = 0;  // near-field option: gap = near_c * max sigma of the bodies
  // sensor state (fmmCalculator.hpp:59-64)
  std::vector<Body>      sensors;
  std::vector<Cell>      scells;
  std::vector<Pair>      sp2p, sm2l;
  std::vector<complex_t> Ls;

  Calculator(int P_, real_t th, int64_t nc, int st, int fo)
      : P(P_), NTERM(3 * (P_ * (P_ + 1) / 2)), theta(th), ncrit(nc), stretching(st), formulation(fo) {}

  // init / createTree: fmmCalculator.hpp:71-115, 262-267, 282-288
  int init(Body *caller, int64_t n)
  {
    int st = build_tree(caller, n, ncrit, cells);
    if (st != ORC_OK) return st;
    levels.clear();
    for (int64_t i = 0; i < (int64_t)cells.size(); i++) {
      const int64_t l = cells[i].level;
      if ((int64_t)levels.size() <= l) levels.resize(l + 1);
      levels[l].push_back(i);
    }
    gap = 0;
    if (near_c > 0) {
      real_t smax = 0;
      for (int64_t i = 0; i < n; i++) smax = std::max(smax, caller[i].radius);
      gap = near_c * smax;
    }
    get_list(cells, cells, theta, gap, p2p, m2l);
    bodies.assign(caller, caller + n);
    M.assign(cells.size() * NTERM, complex_t(0, 0));
    L.assign(cells.size() * NTERM, complex_t(0, 0));
    ready = true;
    return ORC_OK;
  }

  // calculate(): fmmCalculator.hpp:420-467
  int calculate()
  {
    if (!ready) return ORC_ERR_STATE;
    Ops          ops(P);
    const int64_t C = cells.size(), N = bodies.size();
    std::fill(M.begin(), M.end(), complex_t(0, 0));
    std::fill(L.begin(), L.end(), complex_t(0, 0));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) {  // resetRates (359-378)
      Body &b = bodies[i];
      for (int d = 0; d < 3; d++) b.velocity[d] = b.dadt[d] = 0;
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) b.J[a][c] = 0;
      b.drdt = 0;
    }
    // upward pass, levels deepest..1; P2M at leaves then M2M gather (290-298, 427-437)
    for (int64_t l = (int64_t)levels.size() - 1; l >= 1; l--) {
      const auto &lv = levels[l];
#pragma omp parallel for schedule(dynamic, 16)
      for (int64_t q = 0; q < (int64_t)lv.size(); q++) {
        const int64_t ci = lv[q];
        const Cell   &c  = cells[ci];
        if (c.NCHILD == 0) ops.p2m(c.X, &bodies[c.BodyOffset], c.NBODY, &M[ci * NTERM]);
        for (int64_t ch = c.ChildOffset; ch < c.ChildOffset + c.NCHILD; ch++)
          ops.m2m(c.X, cells[ch].X, &M[ch * NTERM], &M[ci * NTERM]);
      }
    }
    // M2L (439), grouped by target
    CSR m2lcsr;
    m2lcsr.build(C, m2l);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < C; t++)
      for (int64_t q = m2lcsr.off[t]; q < m2lcsr.off[t + 1]; q++) {
        const int64_t s = m2lcsr.idx[q];
        ops.m2l(cells[t].X, cells[s].X, &M[s * NTERM], &L[t * NTERM]);
      }
    // P2P (440), grouped by target leaf
    CSR p2pcsr;
    p2pcsr.build(C, p2p);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t t = 0; t < C; t++) {
      const Cell &Ct = cells[t];
      for (int64_t q = p2pcsr.off[t]; q < p2pcsr.off[t + 1]; q++) {
        const Cell &Cs = cells[p2pcsr.idx[q]];
        for (int64_t i = Ct.BodyOffset; i < Ct.BodyOffset + Ct.NBODY; i++) {
          real_t acc[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
          for (int64_t j = Cs.BodyOffset; j < Cs.BodyOffset + Cs.NBODY; j++) p2p_pair_sum(bodies[i].X, bodies[j], acc);
          Body &b = bodies[i];
          for (int d = 0; d < 3; d++) b.velocity[d] -= acc[d];
          for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) b.J[a][c] -= acc[3 + 3 * a + c];
        }
      }
    }
    // downward pass, levels 0..deepest: L2L scatter, then L2P at leaves (339-347, 442-452)
    for (int64_t l = 0; l < (int64_t)levels.size(); l++) {
      const auto &lv = levels[l];
#pragma omp parallel for schedule(dynamic, 16)
      for (int64_t q = 0; q < (int64_t)lv.size(); q++) {
        const int64_t ci = lv[q];
        const Cell   &c  = cells[ci];
        for (int64_t ch = c.ChildOffset; ch < c.ChildOffset + c.NCHILD; ch++)
          ops.l2l(c.X, cells[ch].X, &L[ci * NTERM], &L[ch * NTERM]);
        if (c.NCHILD == 0)
          for (int64_t i = c.BodyOffset; i < c.BodyOffset + c.NBODY; i++) ops.l2p(c.X, &L[ci * NTERM], bodies[i]);
      }
    }
    // stretching (391-405, 454-460) and rVPM (407-418, 462-465)
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) {
      Body &b = bodies[i];
      if (stretching == STRETCH_TRANSPOSED) {
        for (int d = 0; d < 3; d++) b.dadt[d] += b.J[0][d] * b.alpha[0] + b.J[1][d] * b.alpha[1] + b.J[2][d] * b.alpha[2];
      } else {
        for (int d = 0; d < 3; d++) b.dadt[d] += b.J[d][0] * b.alpha[0] + b.J[d][1] * b.alpha[1] + b.J[d][2] * b.alpha[2];
      }
      if (formulation == FORM_RVPM) {
        const real_t na = std::sqrt(b.alpha[0] * b.alpha[0] + b.alpha[1] * b.alpha[1] + b.alpha[2] * b.alpha[2]);
        const real_t ah[3] = {b.alpha[0] / na, b.alpha[1] / na, b.alpha[2] / na};
        const real_t dot   = b.dadt[0] * ah[0] + b.dadt[1] * ah[1] + b.dadt[2] * ah[2];
        for (int d = 0; d < 3; d++) b.dadt[d] -= 3. / 5. * dot * ah[d];
        b.drdt -= 1.0 / 5.0 * b.radius / na * dot;
      }
    }
    return ORC_OK;
  }

  // RK stage functors: fmmCalculator.hpp:493-711
  template <class F> void each(F f)
  {
    const int64_t N = bodies.size();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) f(bodies[i]);
  }
  void store_initial_state()
  {
    each([](Body &b) {
      for (int d = 0; d < 3; d++) {
        b.velocity_old[d] = 0;
        b.dadt_old[d]     = 0;
        b.X_old[d]        = b.X[d];
        b.alpha_old[d]    = b.alpha[d];
      }
      b.radius_old = b.radius;
    });
  }
  void step(int scheme, int stage)
  {
    const real_t dt_ = dt;
    if (scheme == SCHEME_EULER) {
      each([dt_](Body &b) {
        for (int d = 0; d < 3; d++) b.X[d] += b.velocity[d] * dt_;
        for (int d = 0; d < 3; d++) b.alpha[d] += b.dadt[d] * dt_;
        b.radius += b.drdt * dt_;
      });
    } else if (scheme == SCHEME_RK2) {
      if (stage == 0)
        each([dt_](Body &b) {
          for (int d = 0; d < 3; d++) { b.velocity_old[d] = b.velocity[d]; b.dadt_old[d] = b.dadt[d]; }
          b.drdt_old = b.drdt;
          for (int d = 0; d < 3; d++) b.X[d] += dt_ * b.velocity_old[d];
          for (int d = 0; d < 3; d++) b.alpha[d] += dt_ * b.dadt_old[d];
          b.radius += dt_ * b.drdt_old;
        });
      else
        each([dt_](Body &b) {
          for (int d = 0; d < 3; d++) {
            b.velocity_old[d] = 0.5 * b.velocity[d] + 0.5 * b.velocity_old[d];
            b.dadt_old[d]     = 0.5 * b.dadt[d] + 0.5 * b.dadt_old[d];
          }
          b.drdt_old = 0.5 * b.drdt + 0.5 * b.drdt_old;
          for (int d = 0; d < 3; d++) b.X[d] = b.X_old[d] + dt_ * b.velocity_old[d];
          for (int d = 0; d < 3; d++) b.alpha[d] = b.alpha_old[d] + dt_ * b.dadt_old[d];
          b.radius = b.radius_old + dt_ * b.drdt_old;
        });
    } else if (scheme == SCHEME_RK3) {
      // Williamson 2N-storage: z <- A z + dt f; y <- y + B z (564-627)
      const real_t A[3] = {0.0, -(5.0 / 9.0), (-153.0 / 128.0)};
      const real_t B[3] = {(1.0 / 3.0), (15.0 / 16.0), (8.0 / 15.0)};
      const real_t a = A[stage], bb = B[stage];
      if (stage == 0)
        each([dt_, bb](Body &b) {
          for (int d = 0; d < 3; d++) { b.velocity_old[d] = dt_ * b.velocity[d]; b.dadt_old[d] = dt_ * b.dadt[d]; }
          b.drdt_old = dt_ * b.drdt;
          for (int d = 0; d < 3; d++) b.X[d] += bb * b.velocity_old[d];
          for (int d = 0; d < 3; d++) b.alpha[d] += bb * b.dadt_old[d];
          b.radius += bb * b.drdt_old;
        });
      else
        each([dt_, a, bb](Body &b) {
          for (int d = 0; d < 3; d++) {
            b.velocity_old[d] = a * b.velocity_old[d] + dt_ * b.velocity[d];
            b.dadt_old[d]     = a * b.dadt_old[d] + dt_ * b.dadt[d];
          }
          b.drdt_old = a * b.drdt_old + dt_ * b.drdt;
          for (int d = 0; d < 3; d++) b.X[d] += bb * b.velocity_old[d];
          for (int d = 0; d < 3; d++) b.alpha[d] += bb * b.dadt_old[d];
          b.radius += bb * b.drdt_old;
        });
    } else {  // RK4 (629-711), stage 3 uses 0.5*dt as written (683-689)
      if (stage == 0)
        each([dt_](Body &b) {
          for (int d = 0; d < 3; d++) { b.velocity_old[d] = b.velocity[d]; b.dadt_old[d] = b.dadt[d]; }
          b.drdt_old = b.drdt;
          for (int d = 0; d < 3; d++) b.X[d] += 0.5 * dt_ * b.velocity_old[d];
          for (int d = 0; d < 3; d++) b.alpha[d] += 0.5 * dt_ * b.dadt_old[d];
          b.radius += 0.5 * dt_ * b.drdt_old;
        });
      else if (stage == 1 || stage == 2)
        each([dt_](Body &b) {
          for (int d = 0; d < 3; d++) { b.velocity_old[d] += 2.0 * b.velocity[d]; b.dadt_old[d] += 2.0 * b.dadt[d]; }
          b.drdt_old += 2.0 * b.drdt;
          for (int d = 0; d < 3; d++) b.X[d] = b.X_old[d] + 0.5 * dt_ * b.velocity[d];
          for (int d = 0; d < 3; d++) b.alpha[d] = b.alpha_old[d] + 0.5 * dt_ * b.dadt[d];
          b.radius = b.radius_old + 0.5 * dt_ * b.drdt;
        });
      else
        each([dt_](Body &b) {
          for (int d = 0; d < 3; d++) { b.velocity_old[d] += b.velocity[d]; b.dadt_old[d] += b.dadt[d]; }
          b.drdt_old += b.drdt;
          for (int d = 0; d < 3; d++) b.X[d] = b.X_old[d] + (dt_ / 6.0) * b.velocity_old[d];
          for (int d = 0; d < 3; d++) b.alpha[d] = b.alpha_old[d] + (dt_ / 6.0) * b.dadt_old[d];
          b.radius = b.radius_old + (dt_ / 6.0) * b.drdt_old;
        });
    }
  }

  // advance(): fmmCalculator.hpp:919-1005.  Each stage: init (tree rebuild, permutes the
  // caller's bodies) -> calculate -> step -> copy back to the caller.
  int advance(Body *caller, int64_t n, real_t dt_, int scheme)
  {
    dt = dt_;
    const int stages = scheme == SCHEME_EULER ? 1 : scheme == SCHEME_RK2 ? 2 : scheme == SCHEME_RK3 ? 3 : 4;
    for (int s = 0; s < stages; s++) {
      int st = init(caller, n);
      if (st != ORC_OK) return st;
      if (s == 0 && (scheme == SCHEME_RK2 || scheme == SCHEME_RK4)) store_initial_state();
      calculate();
      calculate_viscosity();
      step(scheme, s);
      std::memcpy(caller, bodies.data(), sizeof(Body) * n);
    }
    return ORC_OK;
  }

  // One stage of advance() from the full body state the caller holds (the *_old fields
  // included), so a stage can be checked in isolation from the rebuilds before it.
  int stage(Body *caller, int64_t n, real_t dt_, int scheme, int s)
  {
    const int stages = scheme == SCHEME_EULER ? 1 : scheme == SCHEME_RK2 ? 2 : scheme == SCHEME_RK3 ? 3 : 4;
    if (s < 0 || s >= stages) return ORC_ERR_ARG;
    dt = dt_;
    int st = init(caller, n);
    if (st != ORC_OK) return st;
    if (s == 0 && (scheme == SCHEME_RK2 || scheme == SCHEME_RK4)) store_initial_state();
    calculate();
    calculate_viscosity();
    step(scheme, s);
    std::memcpy(caller, bodies.data(), sizeof(Body) * n);
    return ORC_OK;
  }

  // calculateViscosity (fmmCalculator.hpp:713-724) after calculate().
  //  PSE (multipoleKernels.hpp:392-450, symmetric kernel of Zhao & He): for every P2P pair
  //  of the particle lists and every target i, the partial sum over the source cell's j
  //  (R^2 != 0) of (vol_i a_j - vol_j a_i) exp(-R^2 / (2 s_ij^2)) / s_ij^5, s_ij^2 =
  //  sigma_i^2 + sigma_j^2, vol = 4/3 pi sigma^3, is added to dadt_i times
  //  nu / (sqrt(2) pi^{3/2}).
  //  CSM (fmmCalculator.hpp:334-337): drdt_i += (sqrt(sigma_i^2 + 2 nu dt) - sigma_i) / dt.
  void calculate_viscosity()
  {
    if (viscous == 1) {
      CSR csr;
      csr.build(cells.size(), p2p);
      const real_t prefix = nu / std::sqrt(2.0) * (1.0 / M_PI) * (1.0 / std::sqrt(M_PI));
#pragma omp parallel for schedule(dynamic, 4)
      for (int64_t t = 0; t < (int64_t)cells.size(); t++) {
        const Cell &Ct = cells[t];
        for (int64_t q = csr.off[t]; q < csr.off[t + 1]; q++) {
          const Cell &Cs = cells[csr.idx[q]];
          for (int64_t i = Ct.BodyOffset; i < Ct.BodyOffset + Ct.NBODY; i++) {
            Body        &bi = bodies[i];
            real_t       sum[3] = {0, 0, 0};
            const real_t vol_i  = 4. / 3. * M_PI * bi.radius * bi.radius * bi.radius;
            for (int64_t j = Cs.BodyOffset; j < Cs.BodyOffset + Cs.NBODY; j++) {
              const Body &bj = bodies[j];
              real_t      dX[3];
              for (int d = 0; d < 3; d++) dX[d] = bi.X[d] - bj.X[d];
              const real_t R2 = dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2];
              if (R2 != 0) {
                const real_t o2    = 1. / (bi.radius * bi.radius + bj.radius * bj.radius);
                const real_t s5    = o2 * o2 * std::sqrt(o2);
                const real_t vol_j = 4. / 3. * M_PI * bj.radius * bj.radius * bj.radius;
                const real_t e     = std::exp(-0.5 * R2 * o2);
                for (int d = 0; d < 3; d++) sum[d] += (vol_i * bj.alpha[d] - vol_j * bi.alpha[d]) * e * s5;
              }
            }
            for (int d = 0; d < 3; d++) bi.dadt[d] += sum[d] * prefix;
          }
        }
      }
    } else if (viscous == 2) {
      for (auto &b : bodies) b.drdt += (std::sqrt(b.radius * b.radius + 2.0 * nu * dt) - b.radius) / dt;
    }
  }

  // Gaussian vorticity of the sources at the targets over near (P2P) pairs only
  // (multipoleKernels.hpp:343-390): out_i -= sum_j alpha_j exp(-R^2 / (2 sigma_j^2)) /
  // ((2 pi)^{3/2} sigma_j^3), no R^2 test.  `field` selects velocity_old (dual tree) or
  // velocity (single tree).
  static void gauss_vorticity(std::vector<Body> &tb, const std::vector<Cell> &tc, const std::vector<Body> &sb,
                              const std::vector<Cell> &sc, const std::vector<Pair> &pairs, bool to_old)
  {
    CSR csr;
    csr.build(tc.size(), pairs);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t t = 0; t < (int64_t)tc.size(); t++) {
      const Cell &Ct = tc[t];
      for (int64_t q = csr.off[t]; q < csr.off[t + 1]; q++) {
        const Cell &Cs = sc[csr.idx[q]];
        for (int64_t i = Ct.BodyOffset; i < Ct.BodyOffset + Ct.NBODY; i++) {
          real_t sum[3] = {0, 0, 0};
          for (int64_t j = Cs.BodyOffset; j < Cs.BodyOffset + Cs.NBODY; j++) {
            const Body &bj = sb[j];
            real_t      dX[3];
            for (int d = 0; d < 3; d++) dX[d] = tb[i].X[d] - bj.X[d];
            const real_t R2 = dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2];
            const real_t S2 = 2.0 * bj.radius * bj.radius;
            const real_t z  = std::exp(-R2 / S2) * (1.0 / M_PI) * (1.0 / std::sqrt(M_PI)) / std::sqrt(2.0) / S2 / bj.radius;
            for (int d = 0; d < 3; d++) sum[d] += bj.alpha[d] * z;
          }
          for (int d = 0; d < 3; d++) (to_old ? tb[i].velocity_old[d] : tb[i].velocity[d]) -= sum[d];
        }
      }
    }
  }

  // calculateVorticityDualTree (fmmCalculator.hpp:746-753) on probes: probe tree, lists
  // against the particle tree (createSensorTree 117-161), velocity_old reset and filled.
  int vorticity(Body *caller, int64_t m)
  {
    if (!ready) return ORC_ERR_STATE;
    int st = build_tree(caller, m, ncrit, scells);
    if (st != ORC_OK) return st;
    get_list(scells, cells, theta, gap, sp2p, sm2l);
    sensors.assign(caller, caller + m);
    for (auto &b : sensors)
      for (int d = 0; d < 3; d++) b.velocity_old[d] = 0;
    gauss_vorticity(sensors, scells, bodies, cells, sp2p, true);
    std::memcpy(caller, sensors.data(), sizeof(Body) * m);
    return ORC_OK;
  }

  // RBF re-initialisation (fmmCalculator.hpp:855-917).  The particles must be set with
  // init().  On return pts[0..*kept) holds the kept points (sensor-tree order) with the
  // fitted alpha; *iters = CG iterations run, res[3] = last squared residual norms.
  int rbf(Body *pts, int64_t m, int64_t *kept, int *iters, real_t *res_out)
  {
    if (!ready) return ORC_ERR_STATE;
    const real_t tol = 1e-10, eps = 2.220446049250313e-16;
    int          st  = vorticity(pts, m);
    if (st != ORC_OK) return st;
    std::vector<Body> np;
    for (int64_t i = 0; i < m; i++)
      if (!(std::abs(pts[i].velocity_old[0]) < 1e-6 && std::abs(pts[i].velocity_old[1]) < 1e-6 &&
            std::abs(pts[i].velocity_old[2]) < 1e-6))
        np.push_back(pts[i]);
    const int64_t k = np.size();
    *kept           = k;
    *iters          = 0;
    for (int d = 0; d < 3; d++) res_out[d] = 0;
    if (k == 0) return ORC_OK;
    st = build_tree(np.data(), k, ncrit, scells);  // createSensorTree(sensorOnly) 164-201
    if (st != ORC_OK) return st;
    std::vector<Pair> self_p2p, self_m2l;
    get_list(scells, scells, theta, gap, self_p2p, self_m2l);
    auto single = [&]() {  // calculateVorticitySingleTree (788-795): resetSensors, then P2P
      for (auto &b : np)
        for (int d = 0; d < 3; d++) b.velocity[d] = b.dadt[d] = 0;
      gauss_vorticity(np, scells, np, scells, self_p2p, false);
    };
    for (auto &b : np) {  // initAlphaValues (765-773)
      const real_t vol = 4. / 3. * M_PI * b.radius * b.radius * b.radius;
      for (int d = 0; d < 3; d++) b.alpha[d] = b.velocity_old[d] * vol;
    }
    single();
    real_t res[3] = {0, 0, 0};
    for (auto &b : np)  // setInitialResidualAndUpdateAlpha (797-819)
      for (int d = 0; d < 3; d++) {
        const real_t r = b.velocity_old[d] - b.velocity[d];
        res[d] += r * r;
        b.dadt_old[d] = b.alpha[d];
        b.alpha[d]    = r;
        b.X_old[d]    = r;
      }
    for (int it = 0; it < 100; it++) {
      single();
      real_t pAp[3] = {0, 0, 0}, a[3], nr[3] = {0, 0, 0};
      for (auto &b : np)
        for (int d = 0; d < 3; d++) pAp[d] += b.alpha[d] * b.velocity[d];
      for (int d = 0; d < 3; d++) a[d] = (res[d] > eps && std::abs(pAp[d]) > eps) ? res[d] / pAp[d] : 0.0;
      for (auto &b : np)
        for (int d = 0; d < 3; d++) {
          b.dadt_old[d] += a[d] * b.alpha[d];
          b.X_old[d] -= a[d] * b.velocity[d];
        }
      for (auto &b : np)
        for (int d = 0; d < 3; d++) nr[d] += b.X_old[d] * b.X_old[d];
      *iters = it + 1;
      for (int d = 0; d < 3; d++) res_out[d] = nr[d];
      if (nr[0] < tol && nr[1] < tol && nr[2] < tol) break;
      real_t beta[3];
      for (int d = 0; d < 3; d++) beta[d] = (nr[d] > eps && res[d] > eps) ? nr[d] / res[d] : 0.0;
      for (int d = 0; d < 3; d++) res[d] = nr[d];
      for (auto &b : np)
        for (int d = 0; d < 3; d++) b.alpha[d] = b.X_old[d] + beta[d] * b.alpha[d];
    }
    for (auto &b : np)  // storeAlphaValues (848-853)
      for (int d = 0; d < 3; d++) b.alpha[d] = b.dadt_old[d];
    std::memcpy(pts, np.data(), sizeof(Body) * k);
    return ORC_OK;
  }

  // getSensorData(): fmmCalculator.hpp:117-161, 274-280, 469-491, 726-731.  Uses the
  // multipoles of the last calculate(); permutes the caller's sensors into tree order.
  int sensor_data(Body *caller, int64_t m)
  {
    if (!ready) return ORC_ERR_STATE;
    int st = build_tree(caller, m, ncrit, scells);
    if (st != ORC_OK) return st;
    std::vector<std::vector<int64_t>> slevels;
    for (int64_t i = 0; i < (int64_t)scells.size(); i++) {
      const int64_t l = scells[i].level;
      if ((int64_t)slevels.size() <= l) slevels.resize(l + 1);
      slevels[l].push_back(i);
    }
    get_list(scells, cells, theta, gap, sp2p, sm2l);
    sensors.assign(caller, caller + m);
    const int64_t C = scells.size();
    Ls.assign(C * NTERM, complex_t(0, 0));
    Ops ops(P);
    for (auto &b : sensors)
      for (int d = 0; d < 3; d++) b.velocity[d] = b.dadt[d] = 0;  // resetSensors (380-389)
    CSR m2lcsr;
    m2lcsr.build(C, sm2l);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < C; t++)
      for (int64_t q = m2lcsr.off[t]; q < m2lcsr.off[t + 1]; q++) {
        const int64_t s = m2lcsr.idx[q];
        ops.m2l(scells[t].X, cells[s].X, &M[s * NTERM], &Ls[t * NTERM]);
      }
    CSR p2pcsr;
    p2pcsr.build(C, sp2p);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t t = 0; t < C; t++) {
      const Cell &Ct = scells[t];
      for (int64_t q = p2pcsr.off[t]; q < p2pcsr.off[t + 1]; q++) {
        const Cell &Cs = cells[p2pcsr.idx[q]];
        for (int64_t i = Ct.BodyOffset; i < Ct.BodyOffset + Ct.NBODY; i++) {
          real_t acc[3] = {0, 0, 0};
          for (int64_t j = Cs.BodyOffset; j < Cs.BodyOffset + Cs.NBODY; j++)
            p2p_sensor_pair_sum(sensors[i].X, bodies[j], acc);
          for (int d = 0; d < 3; d++) sensors[i].velocity[d] -= acc[d];
        }
      }
    }
    for (int64_t l = 0; l < (int64_t)slevels.size(); l++) {
      const auto &lv = slevels[l];
#pragma omp parallel for schedule(dynamic, 16)
      for (int64_t q = 0; q < (int64_t)lv.size(); q++) {
        const int64_t ci = lv[q];
        const Cell   &c  = scells[ci];
        for (int64_t ch = c.ChildOffset; ch < c.ChildOffset + c.NCHILD; ch++)
          ops.l2l(c.X, scells[ch].X, &Ls[ci * NTERM], &Ls[ch * NTERM]);
        if (c.NCHILD == 0)
          for (int64_t i = c.BodyOffset; i < c.BodyOffset + c.NBODY; i++) ops.l2p(c.X, &Ls[ci * NTERM], sensors[i]);
      }
    }
    std::memcpy(caller, sensors.data(), sizeof(Body) * m);
    return ORC_OK;
  }
};

}  // namespace orc

// =================================================================================
// C ABI for the ctypes wrapper in oracle/oracle.py.  Array conventions: positions,
// strengths and velocities are [n][3] row-major doubles; J is [n][3][3]; expansions
// are [C][NTERM] complex doubles (interleaved re, im).
using namespace orc;

extern "C" {

int orc_body_size(void) { return (int)sizeof(Body); }

void *orc_create(int P, double theta, int64_t ncrit, int stretching, int formulation)
{
  if (P < 1 || P > 40 || ncrit < 1 || !(theta > 0)) return nullptr;
  return new (std::nothrow) Calculator(P, theta, ncrit, stretching, formulation);
}
void orc_destroy(void *h) { delete (Calculator *)h; }
// Near-field option (see get_list): pairs whose bodies can come closer than c * max sigma are
// near; 0 (the default) is the reference's acceptance test alone.  Applies from the next init.
void orc_set_near_gap(void *h, double c) { ((Calculator *)h)->near_c = c > 0 ? c : 0; }

int orc_init(void *h, Body *bodies, int64_t n) { return ((Calculator *)h)->init(bodies, n); }
int orc_calculate(void *h) { return ((Calculator *)h)->calculate(); }
int orc_get_bodies(void *h, Body *out, int64_t n)
{
  Calculator *c = (Calculator *)h;
  if (n != (int64_t)c->bodies.size()) return ORC_ERR_ARG;
  std::memcpy(out, c->bodies.data(), sizeof(Body) * n);
  return ORC_OK;
}
int orc_advance(void *h, Body *bodies, int64_t n, double dt, int scheme)
{
  return ((Calculator *)h)->advance(bodies, n, dt, scheme);
}
int orc_stage(void *h, Body *bodies, int64_t n, double dt, int scheme, int stage)
{
  return ((Calculator *)h)->stage(bodies, n, dt, scheme, stage);
}
int orc_sensor_data(void *h, Body *sensors, int64_t m) { return ((Calculator *)h)->sensor_data(sensors, m); }
void orc_set_viscosity(void *h, int scheme, double nu)
{
  ((Calculator *)h)->viscous = scheme;
  ((Calculator *)h)->nu      = nu;
}
int orc_viscosity(void *h, double dt)
{
  Calculator *c = (Calculator *)h;
  if (!c->ready) return ORC_ERR_STATE;
  c->dt = dt;
  c->calculate_viscosity();
  return ORC_OK;
}
int orc_vorticity(void *h, Body *sensors, int64_t m) { return ((Calculator *)h)->vorticity(sensors, m); }
int orc_rbf(void *h, Body *pts, int64_t m, int64_t *kept, int *iters, double *res)
{
  return ((Calculator *)h)->rbf(pts, m, kept, iters, res);
}

int64_t orc_num_cells(void *h) { return (int64_t)((Calculator *)h)->cells.size(); }
// ints: [C][5] = NCHILD, NBODY, ChildOffset, BodyOffset, level; reals: [C][4] = X, R
static void dump_cells(const std::vector<Cell> &cells, int64_t *ints, double *reals)
{
  for (size_t i = 0; i < cells.size(); i++) {
    const Cell &c = cells[i];
    ints[5 * i + 0] = c.NCHILD;
    ints[5 * i + 1] = c.NBODY;
    ints[5 * i + 2] = c.ChildOffset;
    ints[5 * i + 3] = c.BodyOffset;
    ints[5 * i + 4] = c.level;
    for (int d = 0; d < 3; d++) reals[4 * i + d] = c.X[d];
    reals[4 * i + 3] = c.R;
  }
}
void orc_get_cells(void *h, int64_t *ints, double *reals) { dump_cells(((Calculator *)h)->cells, ints, reals); }
int64_t orc_num_sensor_cells(void *h) { return (int64_t)((Calculator *)h)->scells.size(); }
void orc_get_sensor_cells(void *h, int64_t *ints, double *reals) { dump_cells(((Calculator *)h)->scells, ints, reals); }

void orc_num_pairs(void *h, int sensor, int64_t *np2p, int64_t *nm2l)
{
  Calculator *c = (Calculator *)h;
  *np2p = sensor ? c->sp2p.size() : c->p2p.size();
  *nm2l = sensor ? c->sm2l.size() : c->m2l.size();
}
void orc_get_pairs(void *h, int sensor, int64_t *p2p, int64_t *m2l)
{
  Calculator *c  = (Calculator *)h;
  auto       &pp = sensor ? c->sp2p : c->p2p;
  auto       &mm = sensor ? c->sm2l : c->m2l;
  for (size_t i = 0; i < pp.size(); i++) { p2p[2 * i] = pp[i].t; p2p[2 * i + 1] = pp[i].s; }
  for (size_t i = 0; i < mm.size(); i++) { m2l[2 * i] = mm[i].t; m2l[2 * i + 1] = mm[i].s; }
}
void orc_get_expansions(void *h, double *M, double *L)
{
  Calculator *c = (Calculator *)h;
  if (M) std::memcpy(M, c->M.data(), sizeof(complex_t) * c->M.size());
  if (L) std::memcpy(L, c->L.data(), sizeof(complex_t) * c->L.size());
}

// Standalone tree build on AoS bodies (permutes in place); returns #cells or <0 on error.
int64_t orc_build_tree(Body *bodies, int64_t n, int64_t ncrit, int64_t *ints, double *reals, int64_t cap,
                       double *X0R0)
{
  std::vector<Cell> cells;
  real_t            X0[3], R0;
  int               st = build_tree(bodies, n, ncrit, cells, X0, &R0);
  if (st != ORC_OK) return -st;
  if (X0R0) { X0R0[0] = X0[0]; X0R0[1] = X0[1]; X0R0[2] = X0[2]; X0R0[3] = R0; }
  if ((int64_t)cells.size() <= cap && ints) dump_cells(cells, ints, reals);
  return (int64_t)cells.size();
}

// Harmonics, for unit tests against closed forms.
void orc_eval_multipole(int P, double rho, double alpha, double beta, double *Y, double *Yth)
{
  evalMultipole(P, rho, alpha, beta, (complex_t *)Y, (complex_t *)Yth);
}
void orc_eval_local(int P, double rho, double alpha, double beta, double *Y)
{
  evalLocal(P, rho, alpha, beta, (complex_t *)Y);
}

// Standalone translation operators (accumulate into the output arrays).
void orc_p2m(int P, const double *Xc, int64_t n, const double *X, const double *alpha, double *M)
{
  std::vector<Body> b(n);
  std::memset(b.data(), 0, sizeof(Body) * n);
  for (int64_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) { b[i].X[d] = X[3 * i + d]; b[i].alpha[d] = alpha[3 * i + d]; }
  Ops(P).p2m(Xc, b.data(), n, (complex_t *)M);
}
void orc_m2m(int P, const double *Xp, const double *Xc, const double *Mc, double *Mp)
{
  Ops(P).m2m(Xp, Xc, (const complex_t *)Mc, (complex_t *)Mp);
}
void orc_m2l(int P, const double *Xt, const double *Xs, const double *Ms, double *Lt)
{
  Ops(P).m2l(Xt, Xs, (const complex_t *)Ms, (complex_t *)Lt);
}
// M2L with the singular harmonics filled up to degree harm_order-1 (tests only, see Ops::m2l).
void orc_m2l_order(int P, int harm_order, const double *Xt, const double *Xs, const double *Ms, double *Lt)
{
  Ops(P).m2l(Xt, Xs, (const complex_t *)Ms, (complex_t *)Lt, harm_order);
}
void orc_l2l(int P, const double *Xp, const double *Xc, const double *Lp, double *Lc)
{
  Ops(P).l2l(Xp, Xc, (const complex_t *)Lp, (complex_t *)Lc);
}
void orc_l2p(int P, const double *Xc, const double *L, int64_t n, const double *X, double *u, double *J)
{
  Ops ops(P);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    Body b;
    std::memset(&b, 0, sizeof(Body));
    for (int d = 0; d < 3; d++) b.X[d] = X[3 * i + d];
    ops.l2p(Xc, (const complex_t *)L, b);
    for (int d = 0; d < 3; d++) u[3 * i + d] += b.velocity[d];
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) J[9 * i + 3 * a + c] += b.J[a][c];
  }
}

// O(M*N) direct sum with the device P2P formula: u and J at targets (sources use their own
// sigma; coincident points are skipped, multipoleKernels.hpp:242).  The reference's
// exafmm::direct (traverse.hpp:88-98) is this loop over all pairs.
void orc_direct(int64_t m, const double *xt, int64_t n, const double *xs, const double *as, const double *ss, double *u,
                double *J)
{
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < m; i++) {
    real_t acc[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    Body   bj;
    for (int64_t j = 0; j < n; j++) {
      for (int d = 0; d < 3; d++) { bj.X[d] = xs[3 * j + d]; bj.alpha[d] = as[3 * j + d]; }
      bj.radius = ss[j];
      p2p_pair_sum(&xt[3 * i], bj, acc);
    }
    for (int d = 0; d < 3; d++) u[3 * i + d] = -acc[d];
    if (J)
      for (int k = 0; k < 9; k++) J[9 * i + k] = -acc[3 + k];
  }
}

// Kernel functions of one pair with the source core sigma: f1 = g / (4 pi R^3) (u = -f1 r x a),
// f2 = 2 df1/dR^2 = (t2/sigma^2 - 3 g/R^2) / (4 pi R^3) (the J factor of multipoleKernels.hpp:
// 283-296), f3 = 2 df2/dR^2 = (15 g/R^4 - t2 (5 + 2 s^2) / (sigma^2 R^2)) / (4 pi R^3),
// s^2 = R^2 / (2 sigma^2), g = erf(s) - t2, t2 = 2 s e^{-s^2} / sqrt(pi).
static inline void pair_f123(real_t R2, real_t sigma, real_t &f1, real_t &f2, real_t &f3)
{
  const real_t S2 = sigma * sigma, s2 = R2 / S2 / 2.0, s = std::sqrt(s2);
  const real_t t2 = 2.0 / std::sqrt(M_PI) * s * std::exp(-s2);
  const real_t g  = std::erf(s) - t2;
  const real_t q  = 0.25 / M_PI / R2 / std::sqrt(R2);
  f1              = g * q;
  f2              = (t2 / S2 - 3.0 * g / R2) * q;
  f3              = (15.0 * g / (R2 * R2) - t2 * (5.0 + 2.0 * s2) / (S2 * R2)) * q;
}

// Vector-Jacobian product of the direct-sum rates u_i = -sum_j f1 (r_ij x a_j) and the
// stretching dadt_i = J_i^T a_i (transposed; classic: J_i a_i) (r_ij = x_i - x_j, kernel of
// sigma_j): with cotangents ubar,
// abar it returns xbar = d/dx and alphabar = d/dalpha of S = sum_i ubar_i.u_i + abar_i.dadt_i.
// Per pair, with Jbar_i = a_i abar_i^T (classic: abar_i a_i^T) and e = axial vector of Jbar
// (e_k = eps_abk Jbar_ab):
//   S_ij = -a_j . W,   W = f1 (ubar x r + e) + f2 (Jbar r) x r,
//   dS_ij/dr = -[ f2 (a.(ubar x r + e)) r + f1 (a x ubar) + f3 (a.((Jbar r) x r)) r
//                 + f2 Jbar^T (r x a) + f2 a x (Jbar r) ],
// so alphabar_j -= W, xbar_i += dS/dr, xbar_j -= dS/dr; plus alphabar_i += J_i abar_i
// (classic: J_i^T abar_i).
void orc_direct_vjp(int64_t n, const double *X, const double *A, const double *S, const double *ub, const double *ab,
                    int classic, double *xbar, double *alphabar)
{
  std::vector<real_t> J(9 * n), u(3 * n);
  orc_direct(n, X, n, X, A, S, u.data(), J.data());
  std::vector<real_t> Jb(9 * n), e(3 * n);
  for (int64_t i = 0; i < n; i++) {
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        Jb[9 * i + 3 * a + b] = classic ? ab[3 * i + a] * A[3 * i + b] : A[3 * i + a] * ab[3 * i + b];
    const real_t *M = &Jb[9 * i];
    e[3 * i + 0]    = M[3 * 1 + 2] - M[3 * 2 + 1];
    e[3 * i + 1]    = M[3 * 2 + 0] - M[3 * 0 + 2];
    e[3 * i + 2]    = M[3 * 0 + 1] - M[3 * 1 + 0];
    for (int a = 0; a < 3; a++) {
      alphabar[3 * i + a] = 0;
      xbar[3 * i + a]     = 0;
    }
  }
  for (int64_t i = 0; i < n; i++)
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        alphabar[3 * i + a] += (classic ? J[9 * i + 3 * b + a] : J[9 * i + 3 * a + b]) * ab[3 * i + b];
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j < n; j++) {
      real_t r[3];
      for (int d = 0; d < 3; d++) r[d] = X[3 * i + d] - X[3 * j + d];
      const real_t R2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      if (R2 == 0) continue;
      real_t f1, f2, f3;
      pair_f123(R2, S[j], f1, f2, f3);
      const real_t *M = &Jb[9 * i], *uu = &ub[3 * i], *ee = &e[3 * i], *a = &A[3 * j];
      const real_t  Jr[3]  = {M[0] * r[0] + M[1] * r[1] + M[2] * r[2], M[3] * r[0] + M[4] * r[1] + M[5] * r[2],
                              M[6] * r[0] + M[7] * r[1] + M[8] * r[2]};
      const real_t  uxr[3] = {uu[1] * r[2] - uu[2] * r[1], uu[2] * r[0] - uu[0] * r[2], uu[0] * r[1] - uu[1] * r[0]};
      const real_t  Jrxr[3] = {Jr[1] * r[2] - Jr[2] * r[1], Jr[2] * r[0] - Jr[0] * r[2], Jr[0] * r[1] - Jr[1] * r[0]};
      real_t        W[3];
      for (int d = 0; d < 3; d++) W[d] = f1 * (uxr[d] + ee[d]) + f2 * Jrxr[d];
      const real_t rxa[3] = {r[1] * a[2] - r[2] * a[1], r[2] * a[0] - r[0] * a[2], r[0] * a[1] - r[1] * a[0]};
      const real_t axu[3] = {a[1] * uu[2] - a[2] * uu[1], a[2] * uu[0] - a[0] * uu[2], a[0] * uu[1] - a[1] * uu[0]};
      const real_t axJr[3] = {a[1] * Jr[2] - a[2] * Jr[1], a[2] * Jr[0] - a[0] * Jr[2], a[0] * Jr[1] - a[1] * Jr[0]};
      const real_t c1 = a[0] * (uxr[0] + ee[0]) + a[1] * (uxr[1] + ee[1]) + a[2] * (uxr[2] + ee[2]);
      const real_t c3 = a[0] * Jrxr[0] + a[1] * Jrxr[1] + a[2] * Jrxr[2];
      real_t       dS[3];
      for (int d = 0; d < 3; d++) {
        const real_t JTrxa = M[0 * 3 + d] * rxa[0] + M[1 * 3 + d] * rxa[1] + M[2 * 3 + d] * rxa[2];
        dS[d] = -(f2 * c1 * r[d] + f1 * axu[d] + f3 * c3 * r[d] + f2 * JTrxa + f2 * axJr[d]);
      }
      for (int d = 0; d < 3; d++) {
        alphabar[3 * j + d] -= W[d];
        xbar[3 * i + d] += dS[d];
        xbar[3 * j + d] -= dS[d];
      }
    }
}

// O(M*N) direct sum with the host formula (multipoleKernels.hpp:174-221): u and classic
// stretching dadt = J*alpha_i at the targets, with target strengths at_.
void orc_direct_classic(int64_t m, const double *xt, const double *at, int64_t n, const double *xs, const double *as,
                        const double *ss, double *u, double *dadt)
{
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < m; i++) {
    Body bi, bj;
    for (int d = 0; d < 3; d++) { bi.X[d] = xt[3 * i + d]; bi.alpha[d] = at[3 * i + d]; }
    real_t uu[3] = {0, 0, 0}, s[3] = {0, 0, 0};
    for (int64_t j = 0; j < n; j++) {
      for (int d = 0; d < 3; d++) { bj.X[d] = xs[3 * j + d]; bj.alpha[d] = as[3 * j + d]; }
      bj.radius = ss[j];
      p2p_classic_pair(bi, bj, uu, s);
    }
    for (int d = 0; d < 3; d++) { u[3 * i + d] = -uu[d]; dadt[3 * i + d] = -s[d]; }
  }
}

// Direct sum with the TARGET's core size (xs/as are the sources, st the targets' sigma):
// the exact adjoint field of the probe objective J = sum_p w_p . u(x_p).  With the probes
// as sources (strength w_p) and the particles as targets, u = dJ/dalpha_j and
// J^T alpha_j = dJ/dx_j (no reference counterpart; used to check the GPU adjoint).
void orc_direct_tsigma(int64_t m, const double *xt, const double *st, int64_t n, const double *xs, const double *as,
                       double *u, double *J)
{
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < m; i++) {
    real_t acc[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    Body   bj;
    for (int64_t j = 0; j < n; j++) {
      for (int d = 0; d < 3; d++) { bj.X[d] = xs[3 * j + d]; bj.alpha[d] = as[3 * j + d]; }
      bj.radius = st[i];
      p2p_pair_sum(&xt[3 * i], bj, acc);
    }
    for (int d = 0; d < 3; d++) u[3 * i + d] = -acc[d];
    for (int k = 0; k < 9; k++) J[9 * i + k] = -acc[3 + k];
  }
}

int orc_max_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

}  // extern "C"
