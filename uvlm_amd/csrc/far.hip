// far.hip -- far-field kernels: P2M, M2M, M2L, L2L and L2P (+ stretching, rVPM), FP64.
//
// Solid harmonics use the exafmm normalisation of src/multipoleKernels.hpp:53-172:
//   regular   Y_n^m(r)  = (-1)^n rho^n     P_n^m(cos a) e^{imb} / (n+m)!
//   singular  Y_n^m(r)  = (-1)^n (n-m)! P_n^m(cos a) e^{imb} / rho^{n+1}
// with the Condon-Shortley phase in P_n^m and Y_n^{-m} = conj(Y_n^m).  Expansions are
// stored for m >= 0 only, complex index 3*(n(n+1)/2 + m) + d; degrees 0..P-1.
//
// The translation index/sign rules are those of the reference: M2M multipoleKernels.hpp:493-514,
// M2L 528-550 (singular harmonics of degree < P only, so j+n <= P-1), L2L 603-621.
// L2P evaluates the local expansion's value, gradient and Hessian at each body by
// re-centring it on the body with the L2L rule restricted to output degrees <= 2 (exact
// polynomial algebra); u = curl(psi)/(4 pi) and J = grad u follow from those 3+6 numbers.
// This replaces the reference's bicomplex complex-step L2P (multipoleKernels.hpp:630-742),
// which computes the same derivatives.
//
// Work decomposition (one wavefront = one 64-lane workgroup unless noted):
//   P2M   : one workgroup per leaf; chunks of bodies: lanes (body, column pair) fill the
//           harmonics in LDS, then lane <-> (n,m) coefficient sums over the chunk
//   M2M   : one workgroup per parent cell of a level; lane <-> output (j,k), loop over children
//   M2L   : one workgroup per target cell (or per list segment when a probe tree has few
//           cells); 4 groups of 16 lanes take every 4th source of the list; the singular
//           harmonics and the source multipole sit in LDS; the lanes of a group walk (n, m)
//           in lock step, each adding the terms of the outputs bound to its accumulator
//           slots (plan in m2l_plan.hpp).  Partial sums are combined in a fixed order at
//           the end, so results are deterministic and no atomics are used
//   L2L   : one workgroup per child cell of a level (gather from the parent)
//   L2P   : one workgroup per leaf, lane <-> body
#include "common.hpp"

#include <algorithm>
#include <vector>

namespace uvlm {

namespace {

constexpr int WG = 64;
// Expansion coefficients (n, m >= 0) per lane in the lane <-> coefficient kernels:
// tri(16) = 136 <= QPL * WG.
constexpr int QPL = 3;

inline int nblocks(int64_t n, int b) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// Hand-over of LDS data between the lanes of a one-wavefront workgroup: a wavefront's LDS
// operations execute in order, so waiting for its own (lgkmcnt 0) is enough, and global loads
// issued before this point stay in flight (__syncthreads would also wait for those).  The
// signal fences keep the compiler from moving memory accesses across it.
static_assert(WG == 64, "wave_lds_sync is for one-wavefront (64-lane) workgroups");
// The same for a workgroup of several wavefronts: each waits for its own LDS operations, then
// all meet at the barrier (global loads stay in flight).
__device__ __forceinline__ void block_lds_sync()
{
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_s_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ __forceinline__ void wave_lds_sync()
{
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xc07f);  // vmcnt 63, expcnt 7, lgkmcnt 0
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ void cfma(double2 &acc, double2 a, double2 b)
{
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double  sgn(int e) { return (e & 1) ? -1.0 : 1.0; }
__device__ __forceinline__ int     tri(int n) { return n * (n + 1) / 2; }

// Spherical angles of dX as the quantities the harmonics need: rho, x = cos(a), y = sin(a),
// e^{ib}.  cos/sin of the polar angle and of the azimuth come from the Cartesian
// components; at rho == 0 the reference's cart2sph gives a = 0, b = atan2(0,0) = 0.
// rho, 1/rho, cos/sin of the polar angle (x, y) and of the azimuth (cb, sb) of (dx, dy, dz),
// from two reciprocal square roots and products (no divisions).
struct Sph {
  double rho, inv, x, y, cb, sb;
};
__device__ __forceinline__ Sph sph(double dx, double dy, double dz)
{
  Sph          s;
  const double rxy2 = dx * dx + dy * dy, r2 = rxy2 + dz * dz;
  if (r2 == 0.0) {
    s.rho = 0.0, s.inv = 0.0, s.x = 1.0, s.y = 0.0, s.cb = 1.0, s.sb = 0.0;
    return s;
  }
  s.inv            = rsqrt(r2);
  s.rho            = r2 * s.inv;
  s.x              = dz * s.inv;
  if (rxy2 == 0.0) {
    s.y = 0.0, s.cb = 1.0, s.sb = 0.0;
  } else {
    const double ixy = rsqrt(rxy2);
    s.y              = rxy2 * ixy * s.inv;
    s.cb             = dx * ixy;
    s.sb             = dy * ixy;
  }
  return s;
}

// 1/k for k = 0..63 (entry 0 unused), so the recurrences below multiply instead of divide.
__constant__ double c_inv[64] = {
    0.0,        1.0 / 1,  1.0 / 2,  1.0 / 3,  1.0 / 4,  1.0 / 5,  1.0 / 6,  1.0 / 7,  1.0 / 8,  1.0 / 9,
    1.0 / 10, 1.0 / 11, 1.0 / 12, 1.0 / 13, 1.0 / 14, 1.0 / 15, 1.0 / 16, 1.0 / 17, 1.0 / 18, 1.0 / 19,
    1.0 / 20, 1.0 / 21, 1.0 / 22, 1.0 / 23, 1.0 / 24, 1.0 / 25, 1.0 / 26, 1.0 / 27, 1.0 / 28, 1.0 / 29,
    1.0 / 30, 1.0 / 31, 1.0 / 32, 1.0 / 33, 1.0 / 34, 1.0 / 35, 1.0 / 36, 1.0 / 37, 1.0 / 38, 1.0 / 39,
    1.0 / 40, 1.0 / 41, 1.0 / 42, 1.0 / 43, 1.0 / 44, 1.0 / 45, 1.0 / 46, 1.0 / 47, 1.0 / 48, 1.0 / 49,
    1.0 / 50, 1.0 / 51, 1.0 / 52, 1.0 / 53, 1.0 / 54, 1.0 / 55, 1.0 / 56, 1.0 / 57, 1.0 / 58, 1.0 / 59,
    1.0 / 60, 1.0 / 61, 1.0 / 62, 1.0 / 63};

// 1/k! for k = 0..31.
__constant__ double c_invfact[32] = {
    1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040, 1.0 / 40320, 1.0 / 362880,
    1.0 / 3628800, 1.0 / 39916800, 1.0 / 479001600, 1.0 / 6227020800.0, 1.0 / 87178291200.0,
    1.0 / 1307674368000.0, 1.0 / 20922789888000.0, 1.0 / 355687428096000.0, 1.0 / 6402373705728000.0,
    1.0 / 121645100408832000.0, 1.0 / 2432902008176640000.0, 1.0 / 51090942171709440000.0,
    1.0 / 1124000727777607680000.0, 1.0 / 25852016738884976640000.0, 1.0 / 620448401733239439360000.0,
    1.0 / 15511210043330985984000000.0, 1.0 / 403291461126605635584000000.0,
    1.0 / 10888869450418352160768000000.0, 1.0 / 304888344611713860501504000000.0,
    1.0 / 8841761993739701954543616000000.0, 1.0 / 265252859812191058636308480000000.0,
    1.0 / 8222838654177922817725562880000000.0};

// Column m of the regular (singular=false) or singular harmonics, degrees m..P-1, written
// to Y[tri(n) + m].  Associated Legendre recurrence in n at fixed m:
// (n-m) P_n^m = x (2n-1) P_{n-1}^m - (n+m-1) P_{n-2}^m, from P_m^m = (-1)^m (2m-1)!! y^m.
// inv: the table 1/k (c_inv, or a copy of it in LDS where m differs between the lanes: read from
// __constant__ memory at a per-lane index it is a vector load, waited for inside the recurrence).
template <bool SINGULAR>
__device__ void harmonic_column(int P, int m, const Sph &s, double2 *Y, const double *__restrict__ inv_k)
{
  // e^{imb}
  double2 e = make_double2(1.0, 0.0);
  for (int i = 0; i < m; i++) e = cmul(e, make_double2(s.cb, s.sb));
  // P_m^m = (-1)^m (2m-1)!! y^m
  double pmm = 1.0;
  for (int i = 1; i <= m; i++) pmm *= -(double)(2 * i - 1) * s.y;
  // scale at n = m
  double       f;
  const double inv = SINGULAR ? s.inv : 0.0;
  if (SINGULAR) {
    f = inv;
    for (int i = 0; i < m; i++) f *= -inv;  // (-1)^m / rho^{m+1}
  } else {
    f = 1.0;
    for (int i = 1; i <= m; i++) f *= -s.rho * inv_k[2 * i - 1] * inv_k[2 * i];  // (-1)^m rho^m / (2m)!
  }
  double p_prev = 0.0, p = pmm;
  for (int n = m; n < P; n++) {
    if (n > m) {
      const double pn = (n == m + 1) ? s.x * (2 * m + 1) * p
                                     : (s.x * (2 * n - 1) * p - (n + m - 1) * p_prev) * inv_k[n - m];
      p_prev = p;
      p      = pn;
      if (SINGULAR) f *= -(double)(n - m) * inv;  // (-1)^n (n-m)! / rho^{n+1}
      else f *= -s.rho * inv_k[n + m];            // (-1)^n rho^n / (n+m)!
    }
    const double a = f * p;
    Y[tri(n) + m]  = make_double2(a * e.x, a * e.y);
  }
}

// Column m of the singular harmonics written to the "extended" rows used by k_m2l_ns: row n
// holds orders -n..n at Yx[n*n + n + mu] with Yx[n, mu] = Y_n^mu for mu >= 0 and
// Yx[n, -mu] = (-1)^mu conj(Y_n^mu).  With this table the M2L sign rule becomes a per-output
// factor: (-1)^min(k,m) Y[J, m-k] = (-1)^k Yx[J, m-k].
__device__ void singular_column_ext(int P, int m, const Sph &s, double2 *Yx)
{
  double2 e = make_double2(1.0, 0.0);
  for (int i = 0; i < m; i++) e = cmul(e, make_double2(s.cb, s.sb));
  double pmm = 1.0;
  for (int i = 1; i <= m; i++) pmm *= -(double)(2 * i - 1) * s.y;
  const double inv = s.inv;
  double       f   = inv;
  for (int i = 0; i < m; i++) f *= -inv;
  const double sm     = (m & 1) ? -1.0 : 1.0;
  double       p_prev = 0.0, p = pmm;
  for (int n = m; n < P; n++) {
    if (n > m) {
      const double pn = (n == m + 1) ? s.x * (2 * m + 1) * p
                                     : (s.x * (2 * n - 1) * p - (n + m - 1) * p_prev) * c_inv[n - m];
      p_prev = p;
      p      = pn;
      f *= -(double)(n - m) * inv;
    }
    const double a      = f * p;
    Yx[n * n + n + m]   = make_double2(a * e.x, a * e.y);
    if (m > 0) Yx[n * n + n - m] = make_double2(sm * a * e.x, -sm * a * e.y);
  }
}

// Regular harmonics of the offset between a cell and its parent or child, from the table of the
// 8 unit octant vectors (sx, sy, sz) = (+-1, +-1, +-1): the offset is r (sx, sy, sz) with r the
// child's half-width, and Y_n^m(r k) = r^n Y_n^m(k).  Octant o has bit 0/1/2 set for a positive
// x/y/z component.  Lanes stride over the tri(P) entries.
__device__ __forceinline__ void octant_harmonics(int P, const double2 *__restrict__ octY, int o, double r, int lane,
                                                 int nlanes, double2 *Y)
{
  for (int q = lane; q < tri(P); q += nlanes) {
    int n = 0;
    while (tri(n + 1) <= q) n++;
    double f = 1.0;
    for (int i = 0; i < n; i++) f *= r;
    const double2 y = octY[o * tri(P) + q];
    Y[q]            = make_double2(y.x * f, y.y * f);
  }
}

// Y[n, mu] for any mu, from the m >= 0 table.
__device__ __forceinline__ double2 ylook(const double2 *Y, int n, int mu)
{
  return mu >= 0 ? Y[tri(n) + mu] : conj2(Y[tri(n) - mu]);
}

// ------------------------------------------------------------------ P2M
// PC > 0 fixes the expansion order at compile time; PC = 0 takes it from Prt.
template <int PC>
__global__ __launch_bounds__(WG) void k_p2m(int Prt, int NT, const int32_t *__restrict__ leaves, int64_t nleaf,
                                            const int32_t *__restrict__ cbody, const int32_t *__restrict__ cnbody,
                                            const double *__restrict__ cx, const double *__restrict__ cy,
                                            const double *__restrict__ cz, const double *__restrict__ st, int64_t cap,
                                            double2 *__restrict__ M)
{
  // Chunks of B bodies.  Phase 1: lanes (body, column pair) write the regular harmonics of
  // (body - centre) to LDS; lane c of a body takes columns c and P-1-c, so every lane walks
  // about P+1 degrees.  Phase 2: lane <-> coefficient (n,m) adds alpha conj(Y_n^m) over the
  // chunk, in body order (multipoleKernels.hpp:305-327).
  const int P = PC > 0 ? PC : Prt;
  extern __shared__ double2 smq[];
  __shared__ double tinv[64];  // 1/k: the columns' recurrences index it per lane
  const int64_t w = blockIdx.x;
  if (w >= nleaf) return;
  tinv[threadIdx.x] = c_inv[threadIdx.x];
  wave_lds_sync();
  const int    c = leaves[w], b0 = cbody[c], nb = cnbody[c];
  const double X = cx[c], Yc = cy[c], Z = cz[c];
  const int    lane = threadIdx.x, ny = tri(P);
  const int    C = (P + 1) / 2, B = WG / C;
  double2     *Yb = smq;                         // [B][ny]
  double      *al = (double *)(smq + B * ny);    // [B][3]
  double2      acc[QPL][3];
  for (int a = 0; a < QPL; a++)
    for (int d = 0; d < 3; d++) acc[a][d] = make_double2(0, 0);
  const int pb = lane / C, pc = lane % C;
  // the body of the next chunk is loaded during the current chunk's sums
  double xb[3] = {0, 0, 0}, ab[3] = {0, 0, 0};
  if (pb < B && pb < nb) {
    const int64_t bi = b0 + pb;
    for (int d = 0; d < 3; d++) xb[d] = st[(F_X + d) * cap + bi], ab[d] = st[(F_A + d) * cap + bi];
  }
  for (int c0 = 0; c0 < nb; c0 += B) {
    const int i = c0 + pb;
    if (pb < B && i < nb) {
      const Sph sp = sph(xb[0] - X, xb[1] - Yc, xb[2] - Z);
      harmonic_column<false>(P, pc, sp, Yb + pb * ny, tinv);
      if (P - 1 - pc > pc) harmonic_column<false>(P, P - 1 - pc, sp, Yb + pb * ny, tinv);
      if (pc == 0)
        for (int d = 0; d < 3; d++) al[pb * 3 + d] = ab[d];
    }
    if (pb < B && i + B < nb) {
      const int64_t bi = b0 + i + B;
      for (int d = 0; d < 3; d++) xb[d] = st[(F_X + d) * cap + bi], ab[d] = st[(F_A + d) * cap + bi];
    }
    __syncthreads();
    const int cnt = min(B, nb - c0);
    for (int a = 0; a < QPL; a++) {
      const int q = lane + a * WG;
      if (q >= ny) continue;
      for (int j = 0; j < cnt; j++) {
        const double2 y = Yb[j * ny + q];
        for (int d = 0; d < 3; d++) {
          const double ad = al[j * 3 + d];
          acc[a][d].x     = fma(ad, y.x, acc[a][d].x);
          acc[a][d].y     = fma(-ad, y.y, acc[a][d].y);  // conj(Y_n^m) = Y_n^{-m}
        }
      }
    }
    __syncthreads();
  }
  for (int a = 0; a < QPL; a++) {
    const int q = lane + a * WG;
    if (q < ny)
      for (int d = 0; d < 3; d++) M[(int64_t)c * NT + 3 * q + d] = acc[a][d];
  }
}

// ------------------------------------------------------------------ M2M
// One workgroup of nwv wavefronts per parent cell (nwv = 8, or 4 when 8 LDS slices would pass
// 64 KB): wavefront c translates children c, c + nwv, ... (their harmonics and multipoles in its
// own LDS slice, lane <-> output coefficient (n,m)), so the children are translated side by side
// instead of one after another; the wavefronts' sums are then added in wavefront order through LDS.
constexpr int M2M_MAXWV = 8;
// PC > 0 fixes the expansion order at compile time; PC = 0 takes it from Prt.
template <int PC>
__global__ __launch_bounds__(M2M_MAXWV * WG) void k_m2m(int Prt, int NT, const int32_t *__restrict__ cells,
                                                        int64_t ncell, const int32_t *__restrict__ cnchild,
                                                        const int32_t *__restrict__ cchild,
                                                        const double *__restrict__ cx, const double *__restrict__ cy,
                                                        const double *__restrict__ cz, const double *__restrict__ cR,
                                                        const double2 *__restrict__ octY, double2 *__restrict__ M)
{
  const int P = PC > 0 ? PC : Prt;
  extern __shared__ double2 sm2[];
  const int64_t w = blockIdx.x;
  if (w >= ncell) return;
  const int p = cells[w], nc = cnchild[p];
  if (nc == 0) return;
  const int nwv = blockDim.x / WG, wv = threadIdx.x / WG, lane = threadIdx.x % WG, nout = tri(P);
  double2  *Y  = sm2 + wv * (tri(P) + NT);  // tri(P)
  double2  *Mc = Y + tri(P);                // NT
  double2   acc[QPL][3];
  for (int a = 0; a < QPL; a++)
    for (int d = 0; d < 3; d++) acc[a][d] = make_double2(0, 0);
  for (int c0 = 0; c0 < nc; c0 += nwv) {
    const int  ci = c0 + wv;
    const bool on = ci < nc;  // uniform over the wavefront
    if (on) {
      const int ch = cchild[p] + ci;
      const int o  = (cx[p] > cx[ch]) | ((cy[p] > cy[ch]) << 1) | ((cz[p] > cz[ch]) << 2);
      octant_harmonics(P, octY, o, cR[ch], lane, WG, Y);
      for (int q = lane; q < NT; q += WG) Mc[q] = M[(int64_t)ch * NT + q];
    }
    __syncthreads();
    if (on)
      for (int a = 0; a < QPL; a++) {
        const int q = lane + a * WG;
        if (q >= nout) continue;
        int j = 0;
        while (tri(j + 1) <= q) j++;
        const int k = q - tri(j);
        for (int n = 0; n <= j; n++) {
          for (int m = max(-n, -j + k + n); m <= min(k - 1, n); m++) {
            const int     jnkms = tri(j - n) + k - m;
            const double2 y     = ylook(Y, n, -m);
            const double  sg    = (m >= 0 ? 1.0 : sgn(m)) * sgn(n);
            const double2 ys    = make_double2(y.x * sg, y.y * sg);
            for (int d = 0; d < 3; d++) cfma(acc[a][d], Mc[3 * jnkms + d], ys);
          }
          for (int m = k; m <= min(n, j + k - n); m++) {
            const int     jnkms = tri(j - n) - k + m;
            const double2 y     = ylook(Y, n, -m);
            const double  sg    = sgn(k + n + m);
            const double2 ys    = make_double2(y.x * sg, y.y * sg);
            for (int d = 0; d < 3; d++) cfma(acc[a][d], conj2(Mc[3 * jnkms + d]), ys);
          }
        }
      }
    __syncthreads();
  }
  // sum over the wavefronts in order, one coefficient slot a at a time: part[wv][lane][3]
  double2  *part = sm2;
  const int nw   = min(nwv, nc);
  for (int a = 0; a < QPL && a * WG < nout; a++) {
    if (wv < nw)
      for (int d = 0; d < 3; d++) part[(wv * WG + lane) * 3 + d] = acc[a][d];
    __syncthreads();
    const int q = lane + a * WG;
    if (wv == 0 && q < nout)
      for (int d = 0; d < 3; d++) {
        double2 r = part[lane * 3 + d];
        for (int v = 1; v < nw; v++) {
          r.x += part[(v * WG + lane) * 3 + d].x;
          r.y += part[(v * WG + lane) * 3 + d].y;
        }
        M[(int64_t)p * NT + 3 * q + d] = r;
      }
    __syncthreads();
  }
}

// M2M for P <= M2M_TAB_MAXP from a term table (make_m2m_table): the translation of k_m2m with
// the terms of output q = (j,k) listed once on the host, in k_m2m's (n, m) order, as 16-bit
// entries q' | v << 6: Mc entry q' = (j-n, |k-m|) and v = c * nq + (n, |m|) in one of four
// copies of the child's octant harmonics (c: Y, -Y, conj Y, -conj Y), so that each term is a
// table read, four LDS reads and three complex multiply-adds, with no index arithmetic, sign or
// branch.  The terms k_m2m takes with conj(Mc) (m >= k) are summed as B = sum Mc conj(Y') and
// added as conj(B) after those with Mc (m < k), so each lane runs two loops of its own lengths
// (cnt: plain | conjugated << 8, rows RP and RC of 64 lanes).  One wavefront per child as in
// k_m2m; the wavefronts' sums are added in wavefront order.
constexpr int M2M_TAB_MAXP = 10;
__global__ __launch_bounds__(M2M_MAXWV * WG) void k_m2m_tab(int P, int NT, const int32_t *__restrict__ cells,
                                                            int64_t ncell, const int32_t *__restrict__ cnchild,
                                                            const int32_t *__restrict__ cchild,
                                                            const double *__restrict__ cx, const double *__restrict__ cy,
                                                            const double *__restrict__ cz, const double *__restrict__ cR,
                                                            const double2 *__restrict__ octY,
                                                            const int32_t *__restrict__ tcnt,
                                                            const uint32_t *__restrict__ tab, int RP, int RC,
                                                            double2 *__restrict__ M)
{
  extern __shared__ double2 sm2[];
  const int64_t w = blockIdx.x;
  if (w >= ncell) return;
  const int p = cells[w], nc = cnchild[p];
  if (nc == 0) return;
  const int nwv = blockDim.x / WG, wv = threadIdx.x / WG, lane = threadIdx.x % WG, nq = tri(P);
  double2  *Yv  = sm2 + wv * (4 * nq + NT);  // [4][nq]
  double2  *Mc  = Yv + 4 * nq;               // [NT]
  uint32_t *ts  = reinterpret_cast<uint32_t *>(sm2 + nwv * (4 * nq + NT));  // (RP + RC) rows of 64 halves
  for (int i = threadIdx.x; i < (RP + RC) * WG / 2; i += blockDim.x) ts[i] = tab[i];
  const unsigned short *tl = reinterpret_cast<const unsigned short *>(ts) + lane;
  const int cn = tcnt[lane], cp = cn & 255, cc = cn >> 8;
  const char *yb = reinterpret_cast<const char *>(Yv), *mb = reinterpret_cast<const char *>(Mc);
  double2 A[3], B[3];
  for (int d = 0; d < 3; d++) A[d] = B[d] = make_double2(0, 0);
  for (int c0 = 0; c0 < nc; c0 += nwv) {
    const int  ci = c0 + wv;
    const bool on = ci < nc;  // uniform over the wavefront
    if (on) {
      const int    ch = cchild[p] + ci;
      const int    o  = (cx[p] > cx[ch]) | ((cy[p] > cy[ch]) << 1) | ((cz[p] > cz[ch]) << 2);
      const double r  = cR[ch];
      for (int q = lane; q < nq; q += WG) {
        int n = 0;
        while (tri(n + 1) <= q) n++;
        double f = 1.0;
        for (int i = 0; i < n; i++) f *= r;
        const double2 y = octY[o * nq + q];
        const double  a = y.x * f, b = y.y * f;
        Yv[q]          = make_double2(a, b);
        Yv[nq + q]     = make_double2(-a, -b);
        Yv[2 * nq + q] = make_double2(a, -b);
        Yv[3 * nq + q] = make_double2(-a, b);
      }
      for (int q = lane; q < NT; q += WG) Mc[q] = M[(int64_t)ch * NT + q];
    }
    __syncthreads();
    if (on) {
      for (int i = 0; i < cp; i++) {
        const int      e  = tl[i * WG];
        const double2  y  = *reinterpret_cast<const double2 *>(yb + ((e >> 6) << 4));
        const double2 *mc = reinterpret_cast<const double2 *>(mb + (e & 63) * 48);
        for (int d = 0; d < 3; d++) cfma(A[d], mc[d], y);
      }
      for (int i = 0; i < cc; i++) {
        const int      e  = tl[(RP + i) * WG];
        const double2  y  = *reinterpret_cast<const double2 *>(yb + ((e >> 6) << 4));
        const double2 *mc = reinterpret_cast<const double2 *>(mb + (e & 63) * 48);
        for (int d = 0; d < 3; d++) cfma(B[d], mc[d], y);
      }
    }
    __syncthreads();
  }
  // sum over the wavefronts in order: part[wv][lane][3]
  double2  *part = sm2;
  const int nw   = min(nwv, nc);
  if (wv < nw)
    for (int d = 0; d < 3; d++) part[(wv * WG + lane) * 3 + d] = make_double2(A[d].x + B[d].x, A[d].y - B[d].y);
  __syncthreads();
  if (wv == 0 && lane < nq)
    for (int d = 0; d < 3; d++) {
      double2 r = part[lane * 3 + d];
      for (int v = 1; v < nw; v++) {
        r.x += part[(v * WG + lane) * 3 + d].x;
        r.y += part[(v * WG + lane) * 3 + d].y;
      }
      M[(int64_t)p * NT + 3 * lane + d] = r;
    }
}

// ------------------------------------------------------------------ L2L

// L2L with the work balanced over the lanes.  Output (j, k) of the child takes
// terms (n, m) with n = j + d, m = k - d .. k + d (2d + 1 of them) for d = 0 .. P-1-j, so one lane
// per output would run (P - j)^2 steps while the average is far lower.  Here the terms are taken
// in groups of equal d: group d holds the tri(P - d) outputs with j < P - d, and when they are
// few each output's 2d + 1 terms are split over s lanes (a power of two, aligned), summed by an
// xor butterfly and handed to the output's own lane.  Groups run in ascending d (ascending n per
// output), so every sum is taken in a fixed order.
// PC > 0 fixes the expansion order at compile time; PC = 0 takes it from Prt.
template <int PC>
__global__ __launch_bounds__(WG) void k_l2l_bal(int Prt, int NT, const int32_t *__restrict__ cells, int64_t ncell,
                                                const int32_t *__restrict__ cparent, const double *__restrict__ cx,
                                                const double *__restrict__ cy, const double *__restrict__ cz,
                                                const double *__restrict__ cR, const double2 *__restrict__ octY,
                                                double2 *__restrict__ L)
{
  const int P = PC > 0 ? PC : Prt;
  extern __shared__ double2 sm2[];
  double2      *Y  = sm2;
  double2      *Lp = sm2 + tri(P);
  const int64_t w  = blockIdx.x;
  if (w >= ncell) return;
  const int c = cells[w], p = cparent[c];
  if (p < 0) return;
  const int lane = threadIdx.x;
  const int o    = (cx[c] > cx[p]) | ((cy[c] > cy[p]) << 1) | ((cz[c] > cz[p]) << 2);
  octant_harmonics(P, octY, o, cR[c], lane, WG, Y);
  for (int q = lane; q < NT; q += WG) Lp[q] = L[(int64_t)p * NT + q];
  __syncthreads();
  // terms i = first .. nterm-1 step 'step' of output q in group d, added to acc
  auto terms = [&](int q, int d, int first, int step, double2 *acc) {
    int j = 0;
    while (tri(j + 1) <= q) j++;
    const int k = q - tri(j), n = j + d;
    for (int i = first; i < 2 * d + 1; i += step) {
      const int     m = k - d + i;
      const double2 y = ylook(Y, d, m - k);
      if (m < 0) {
        const double  sg = sgn(k);
        const double2 ys = make_double2(y.x * sg, y.y * sg);
        for (int e = 0; e < 3; e++) cfma(acc[e], conj2(Lp[3 * (tri(n) - m) + e]), ys);
      } else {
        const double  sg = (m < k) ? sgn(m - k) : 1.0;
        const double2 ys = make_double2(y.x * sg, y.y * sg);
        for (int e = 0; e < 3; e++) cfma(acc[e], Lp[3 * (tri(n) + m) + e], ys);
      }
    }
  };
  double2 acc[QPL][3];
  for (int a = 0; a < QPL; a++)
    for (int e = 0; e < 3; e++) acc[a][e] = make_double2(0, 0);
  for (int d = 0; d < P; d++) {
    const int ntask = tri(P - d), nterm = 2 * d + 1;
    int       s     = 1;
    while (s < 16 && 2 * s * ntask <= WG && s < nterm) s *= 2;
    if (s == 1) {
      for (int a = 0; a < QPL; a++) {
        const int q = lane + a * WG;
        if (q < ntask) terms(q, d, 0, 1, acc[a]);
      }
    } else {  // ntask <= WG / 2: outputs q < 32, all in slot a = 0
      const int t = lane / s, sub = lane % s;
      double2   part[3];
      for (int e = 0; e < 3; e++) part[e] = make_double2(0, 0);
      if (t < ntask) terms(t, d, sub, s, part);
      for (int off = s / 2; off > 0; off /= 2)
        for (int e = 0; e < 3; e++) {
          part[e].x += __shfl_xor(part[e].x, off);
          part[e].y += __shfl_xor(part[e].y, off);
        }
      const int src = lane < ntask ? lane * s : lane;
      for (int e = 0; e < 3; e++) {
        const double vx = __shfl(part[e].x, src), vy = __shfl(part[e].y, src);
        if (lane < ntask) acc[0][e].x += vx, acc[0][e].y += vy;
      }
    }
  }
  for (int a = 0; a < QPL; a++) {
    const int q = lane + a * WG;
    if (q >= tri(P)) continue;
    for (int e = 0; e < 3; e++) {
      double2 &out = L[(int64_t)c * NT + 3 * q + e];
      out.x += acc[a][e].x;
      out.y += acc[a][e].y;
    }
  }
}

// k_l2l_bal with its terms read from a host-built table (make_l2l_table) for the fixed orders
// P = 6, 8, 10: the same groups d, lane split s and butterfly, but each lane's terms over all
// groups (at most L2L_TAB_ROWS; 31 at P = 10) are 16-bit entries loaded into registers at kernel
// start: q' | v << 6 | conj << 14, with Lp entry q' (conjugated when conj is set) times copy v
// of the child's octant harmonics (Y, -Y, conj Y, -conj Y in LDS).  A term is then three LDS
// reads of Lp, one of Y and three complex multiply-adds, without the per-term row search, sign
// and conjugation branches of k_l2l_bal.
constexpr int L2L_TAB_ROWS = 32;
// v + v of the partner lane at distance OFF in a butterfly over aligned groups of up to 16 lanes,
// by DPP moves instead of LDS permutes: OFF = 1 and 2 swap within quads; OFF = 4 and 8 take
// the mirror lane of the 8- and 16-lane half row, which pairs the same partial sums once the
// smaller steps are done (every lane then holds its group's sum, as after the xor butterfly).
template <int OFF>
__device__ __forceinline__ double dpp_pair_add(double v)
{
  constexpr int ctrl = OFF == 1 ? 0xB1 : OFF == 2 ? 0x4E : OFF == 4 ? 0x141 : 0x140;
  const int     lo = __double2loint(v), hi = __double2hiint(v);
  const int     plo = __builtin_amdgcn_mov_dpp(lo, ctrl, 0xF, 0xF, false);
  const int     phi = __builtin_amdgcn_mov_dpp(hi, ctrl, 0xF, 0xF, false);
  return v + __hiloint2double(phi, plo);
}
// Lane split of group d (k_l2l_bal): s aligned lanes per output, s a power of two, at most 16.
__host__ __device__ __forceinline__ int l2l_split(int P, int d)
{
  const int ntask = (P - d) * (P - d + 1) / 2, nterm = 2 * d + 1;
  int       s     = 1;
  while (s < 16 && 2 * s * ntask <= WG && s < nterm) s *= 2;
  return s;
}
template <int PC>
__global__ __launch_bounds__(WG) void k_l2l_tab(int NT, const int32_t *__restrict__ cells, int64_t ncell,
                                                const int32_t *__restrict__ cparent, const double *__restrict__ cx,
                                                const double *__restrict__ cy, const double *__restrict__ cz,
                                                const double *__restrict__ cR, const double2 *__restrict__ octY,
                                                const uint4 *__restrict__ tab, double2 *__restrict__ L)
{
  constexpr int P = PC, nq = PC * (PC + 1) / 2;
  static_assert(nq <= WG, "one output per lane");
  extern __shared__ double2 sm2[];
  double2      *Yv = sm2;           // [4][nq]
  double2      *Lp = sm2 + 4 * nq;  // [NT]
  const int64_t w  = blockIdx.x;
  if (w >= ncell) return;
  const int c = cells[w], p = cparent[c];
  if (p < 0) return;
  const int lane = threadIdx.x;
  uint4     tw[L2L_TAB_ROWS / 8];  // this lane's entries, two per 32-bit word
  for (int i = 0; i < L2L_TAB_ROWS / 8; i++) tw[i] = tab[lane * (L2L_TAB_ROWS / 8) + i];
  const int    o = (cx[c] > cx[p]) | ((cy[c] > cy[p]) << 1) | ((cz[c] > cz[p]) << 2);
  const double r = cR[c];
  for (int q = lane; q < nq; q += WG) {
    int n = 0;
    while (tri(n + 1) <= q) n++;
    double f = 1.0;
    for (int i = 0; i < n; i++) f *= r;
    const double2 y = octY[o * nq + q];
    const double  a = y.x * f, b = y.y * f;
    Yv[q]          = make_double2(a, b);
    Yv[nq + q]     = make_double2(-a, -b);
    Yv[2 * nq + q] = make_double2(a, -b);
    Yv[3 * nq + q] = make_double2(-a, b);
  }
  for (int q = lane; q < NT; q += WG) Lp[q] = L[(int64_t)p * NT + q];
  __syncthreads();
  auto entry = [&](int row) -> int {
    const uint4    v = tw[row >> 3];
    const unsigned x = ((row >> 1) & 3) == 0 ? v.x : ((row >> 1) & 3) == 1 ? v.y : ((row >> 1) & 3) == 2 ? v.z : v.w;
    return (row & 1) ? (int)(x >> 16) : (int)(x & 0xffff);
  };
  auto term = [&](int e, double2 *acc) {
    const double2  y  = Yv[(e >> 6) & 255];
    const double2 *lp = Lp + 3 * (e & 63);
    const double   sy = (e >> 14) ? -1.0 : 1.0;
    for (int k = 0; k < 3; k++) {
      const double2 l = lp[k];
      cfma(acc[k], make_double2(l.x, l.y * sy), y);
    }
  };
  double2 acc[3];
  for (int e = 0; e < 3; e++) acc[e] = make_double2(0, 0);
  int row = 0;
#pragma unroll
  for (int d = 0; d < P; d++) {
    const int ntask = tri(P - d), nterm = 2 * d + 1, s = l2l_split(P, d);
    if (s == 1) {
#pragma unroll
      for (int i = 0; i < nterm; i++)
        if (lane < ntask) term(entry(row + i), acc);
      row += nterm;
    } else {  // ntask <= WG / 2: outputs q < 32
      const int t = lane / s, sub = lane % s, nr = (nterm + s - 1) / s;
      double2   part[3];
      for (int e = 0; e < 3; e++) part[e] = make_double2(0, 0);
#pragma unroll
      for (int i = 0; i < nr; i++)
        if (t < ntask && sub + i * s < nterm) term(entry(row + i), part);
      row += nr;
      for (int e = 0; e < 3; e++) {
        if (s >= 2) part[e].x = dpp_pair_add<1>(part[e].x), part[e].y = dpp_pair_add<1>(part[e].y);
        if (s >= 4) part[e].x = dpp_pair_add<2>(part[e].x), part[e].y = dpp_pair_add<2>(part[e].y);
        if (s >= 8) part[e].x = dpp_pair_add<4>(part[e].x), part[e].y = dpp_pair_add<4>(part[e].y);
        if (s >= 16) part[e].x = dpp_pair_add<8>(part[e].x), part[e].y = dpp_pair_add<8>(part[e].y);
      }
      const int src = lane < ntask ? lane * s : lane;
      for (int e = 0; e < 3; e++) {
        const double vx = __shfl(part[e].x, src), vy = __shfl(part[e].y, src);
        if (lane < ntask) acc[e].x += vx, acc[e].y += vy;
      }
    }
  }
  if (lane < nq)
    for (int e = 0; e < 3; e++) {
      double2 &out = L[(int64_t)c * NT + 3 * lane + e];
      out.x += acc[e].x;
      out.y += acc[e].y;
    }
}

// ------------------------------------------------------------------ adjoint of the particle rates
// Adjoint sources of the far field (see uvlm_eval_vjp): at each body a vortex ubar (3) and
// vortex dipoles Jbar (3x3, row d = component, column c = direction), so that
//   M[n,m][d] += ubar_d conj(Y_n^m(x - X)) + sum_c Jbar_{dc} conj(d_c Y_n^m)(x - X),
// the derivative taken with respect to the source position.  Regular-harmonic ladder:
//   d_z Y_n^m = -Y_{n-1}^m,  (d_x + i d_y) Y_n^m = -Y_{n-1}^{m+1},
//   (d_x - i d_y) Y_n^m = Y_{n-1}^{m-1} (m >= 1);  for m = 0 (Y_n^0 real): d_x Y = -Re Y_{n-1}^1,
//   d_y Y = -Im Y_{n-1}^1.   Same layout as k_p2m: harmonics of a chunk of bodies in LDS, then
// lane <-> coefficient.  aux holds ubar at field 0 and Jbar (row-major) at fields 3..11.
__global__ __launch_bounds__(WG) void k_p2m_adj(int P, int NT, const int32_t *__restrict__ leaves, int64_t nleaf,
                                                const int32_t *__restrict__ cbody, const int32_t *__restrict__ cnbody,
                                                const double *__restrict__ cx, const double *__restrict__ cy,
                                                const double *__restrict__ cz, const double *__restrict__ st,
                                                const double *__restrict__ aux, int64_t cap, double2 *__restrict__ M)
{
  extern __shared__ double2 smq[];
  __shared__ double tinv[64];  // 1/k: the columns' recurrences index it per lane
  const int64_t w = blockIdx.x;
  if (w >= nleaf) return;
  tinv[threadIdx.x] = c_inv[threadIdx.x];
  wave_lds_sync();
  const int    c = leaves[w], b0 = cbody[c], nb = cnbody[c];
  const double X = cx[c], Yc = cy[c], Z = cz[c];
  const int    lane = threadIdx.x, ny = tri(P);
  const int    C = (P + 1) / 2, B = WG / C;
  double2     *Yb = smq;                       // [B][ny]
  double      *wt = (double *)(smq + B * ny);  // [B][12]: ubar, Jbar
  double2      acc[QPL][3];
  for (int a = 0; a < QPL; a++)
    for (int d = 0; d < 3; d++) acc[a][d] = make_double2(0, 0);
  const int pb = lane / C, pc = lane % C;
  for (int c0 = 0; c0 < nb; c0 += B) {
    const int i = c0 + pb;
    if (pb < B && i < nb) {
      const int64_t bi = b0 + i;
      const Sph     sp = sph(st[(F_X + 0) * cap + bi] - X, st[(F_X + 1) * cap + bi] - Yc, st[(F_X + 2) * cap + bi] - Z);
      harmonic_column<false>(P, pc, sp, Yb + pb * ny, tinv);
      if (P - 1 - pc > pc) harmonic_column<false>(P, P - 1 - pc, sp, Yb + pb * ny, tinv);
      if (pc == 0)
        for (int f = 0; f < 12; f++) wt[pb * 12 + f] = aux[f * cap + bi];
    }
    __syncthreads();
    const int cnt = min(B, nb - c0);
    for (int a = 0; a < QPL; a++) {
      const int q = lane + a * WG;
      if (q >= ny) continue;
      int n = 0;
      while (tri(n + 1) <= q) n++;
      const int m = q - tri(n);
      for (int j = 0; j < cnt; j++) {
        const double2 *Yj = Yb + j * ny;
        const double2  y  = Yj[q];
        double2        gx = make_double2(0, 0), gy = make_double2(0, 0), gz = make_double2(0, 0);
        if (n > 0) {
          const double2 up = (m + 1 <= n - 1) ? Yj[tri(n - 1) + m + 1] : make_double2(0, 0);
          if (m >= 1) {
            const double2 dn = Yj[tri(n - 1) + m - 1];
            gx = make_double2(0.5 * (dn.x - up.x), 0.5 * (dn.y - up.y));
            gy = make_double2(-0.5 * (up.y + dn.y), 0.5 * (up.x + dn.x));  // (i/2)(up + dn)
          } else {
            gx = make_double2(-up.x, 0.0);
            gy = make_double2(-up.y, 0.0);
          }
          if (m <= n - 1) {
            const double2 z0 = Yj[tri(n - 1) + m];
            gz               = make_double2(-z0.x, -z0.y);
          }
        }
        const double *wj = wt + j * 12;
        for (int d = 0; d < 3; d++) {
          const double ud = wj[d], jx = wj[3 + 3 * d], jy = wj[4 + 3 * d], jz = wj[5 + 3 * d];
          // conj of (ubar_d Y + Jbar_dx gx + Jbar_dy gy + Jbar_dz gz)
          acc[a][d].x += ud * y.x + jx * gx.x + jy * gy.x + jz * gz.x;
          acc[a][d].y -= ud * y.y + jx * gx.y + jy * gy.y + jz * gz.y;
        }
      }
    }
    __syncthreads();
  }
  for (int a = 0; a < QPL; a++) {
    const int q = lane + a * WG;
    if (q < ny)
      for (int d = 0; d < 3; d++) M[(int64_t)c * NT + 3 * q + d] = acc[a][d];
  }
}

// Target side of the far-field adjoint: with the forward local expansion L of each leaf, at every
// body x_i the L2L to the body up to degree 3 gives grad, Hessian and third derivatives of psi;
// then u = curl psi / (4 pi), J = grad u, and
//   xbar_i += J^T ubar_i + sum_ab Jbar_ab grad J_ab     (into aux fields 12..14).
// Third derivatives of Re(w_k o[3,k] Y_3^k(h)) with o[3,k] = A_k + i B_k (w_0 = 1, w_k = 2):
//   xxx = -3/4 A1 + 1/4 A3, xxy = 1/4 B1 - 1/4 B3, xxz = 1/2 A0 - 1/2 A2, xyy = -1/4 A1 - 1/4 A3,
//   xyz = 1/2 B2, xzz = A1, yyy = 3/4 B1 + 1/4 B3, yyz = 1/2 A0 + 1/2 A2, yzz = -B1, zzz = -A0.
// PC > 0 fixes the expansion order at compile time (fully unrolled sums); PC = 0 takes it from Prt.
template <int PC>
__global__ __launch_bounds__(WG) void k_l2p_adj_tgt(int Prt, int NT, const int32_t *__restrict__ leaves, int64_t nleaf,
                                                    const int32_t *__restrict__ cbody,
                                                    const int32_t *__restrict__ cnbody, const double *__restrict__ cx,
                                                    const double *__restrict__ cy, const double *__restrict__ cz,
                                                    const double2 *__restrict__ L, const double *__restrict__ st,
                                                    double *__restrict__ aux, int64_t cap)
{
  const int P = PC > 0 ? PC : Prt;
  extern __shared__ double2 sm2[];
  double2      *Lc = sm2;
  const int64_t w  = blockIdx.x;
  if (w >= nleaf) return;
  const int c = leaves[w], b0 = cbody[c], nb = cnbody[c];
  for (int q = threadIdx.x; q < NT; q += WG) Lc[q] = L[(int64_t)c * NT + q];
  __syncthreads();
  const double X = cx[c], Yc = cy[c], Z = cz[c];
  constexpr int NO = 10;
  for (int i = threadIdx.x; i < nb; i += WG) {
    const int64_t bi = b0 + i;
    const Sph     s  = sph(st[(F_X + 0) * cap + bi] - X, st[(F_X + 1) * cap + bi] - Yc, st[(F_X + 2) * cap + bi] - Z);
    double2       o[NO][3];
    for (int a = 0; a < NO; a++)
      for (int d = 0; d < 3; d++) o[a][d] = make_double2(0, 0);
    const int OJ[NO] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3}, OK[NO] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};
#pragma unroll(PC > 0 ? PC : 1)
    for (int mu = 0; mu < P; mu++) {
      // keeps each column's coefficient reads inside it (hoisted out of the body loop they would
      // need more registers than a wavefront has)
      asm volatile("" ::: "memory");
      double2 e = make_double2(1.0, 0.0);
      for (int t = 0; t < mu; t++) e = cmul(e, make_double2(s.cb, s.sb));
      double p = 1.0;
      for (int t = 1; t <= mu; t++) p *= -(double)(2 * t - 1) * s.y;
      double f = c_invfact[2 * mu];
      for (int t = 1; t <= mu; t++) f *= -s.rho;
      double pp = 0.0;
#pragma unroll(PC > 0 ? PC : 1)
      for (int nu = mu; nu < P; nu++) {
        if (nu > mu) {
          const double pn = (nu == mu + 1) ? s.x * (2 * mu + 1) * p
                                           : (s.x * (2 * nu - 1) * p - (nu + mu - 1) * pp) * c_inv[nu - mu];
          pp = p;
          p  = pn;
          f *= -s.rho * c_inv[nu + mu];
        }
        const double  v  = f * p;
        const double2 yp = make_double2(v * e.x, v * e.y);
        for (int sgm = 0; sgm < (mu > 0 ? 2 : 1); sgm++) {
          const int     mup = sgm ? -mu : mu;
          const double2 yv  = sgm ? conj2(yp) : yp;
          for (int a = 0; a < NO; a++) {
            const int j = OJ[a], k = OK[a];
            const int n = nu + j;
            if (n >= P) continue;
            const int m = mup + k;
            double    sg;
            double2   lv[3];
            if (m < 0) {
              sg = sgn(k);
              for (int d = 0; d < 3; d++) lv[d] = conj2(Lc[3 * (tri(n) - m) + d]);
            } else {
              sg = (m < k) ? sgn(m - k) : 1.0;
              for (int d = 0; d < 3; d++) lv[d] = Lc[3 * (tri(n) + m) + d];
            }
            const double2 ys = make_double2(yv.x * sg, yv.y * sg);
            for (int d = 0; d < 3; d++) cfma(o[a][d], lv[d], ys);
          }
        }
      }
    }
    double H[3][3][3], T[3][3][3][3];
    for (int d = 0; d < 3; d++) {
      const double a20 = o[3][d].x, a21 = o[4][d].x, b21 = o[4][d].y, a22 = o[5][d].x, b22 = o[5][d].y;
      H[d][0][0] = 0.5 * (a22 - a20);
      H[d][1][1] = -0.5 * (a20 + a22);
      H[d][2][2] = a20;
      H[d][0][1] = H[d][1][0] = -0.5 * b22;
      H[d][0][2] = H[d][2][0] = -a21;
      H[d][1][2] = H[d][2][1] = b21;
      const double A0 = o[6][d].x, A1 = o[7][d].x, B1 = o[7][d].y, A2 = o[8][d].x, B2 = o[8][d].y, A3 = o[9][d].x,
                   B3 = o[9][d].y;
      const double t[10] = {-0.75 * A1 + 0.25 * A3, 0.25 * B1 - 0.25 * B3, 0.5 * A0 - 0.5 * A2, -0.25 * A1 - 0.25 * A3,
                            0.5 * B2,               A1,                    0.75 * B1 + 0.25 * B3, 0.5 * A0 + 0.5 * A2,
                            -B1,                    -A0};
      // t: xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          for (int cc = 0; cc < 3; cc++) {
            const int nx = (a == 0) + (b == 0) + (cc == 0), ny_ = (a == 1) + (b == 1) + (cc == 1);
            const int nz  = 3 - nx - ny_;
            int       idx = 0;
            if (nx == 3) idx = 0;
            else if (nx == 2 && ny_ == 1) idx = 1;
            else if (nx == 2) idx = 2;
            else if (nx == 1 && ny_ == 2) idx = 3;
            else if (nx == 1 && ny_ == 1) idx = 4;
            else if (nx == 1) idx = 5;
            else if (ny_ == 3) idx = 6;
            else if (ny_ == 2) idx = 7;
            else if (ny_ == 1) idx = 8;
            else idx = 9;
            (void)nz;
            T[d][a][b][cc] = t[idx];
          }
    }
    // curl structure: u_a = sum eps_{a p q} d_p psi_q
    const int    P1[3] = {1, 2, 0}, P2[3] = {2, 0, 1};  // u_0 = d1 psi2 - d2 psi1, ...
    double       Jf[3][3], dJ[3][3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        Jf[a][b] = INV_4PI * (H[P2[a]][P1[a]][b] - H[P1[a]][P2[a]][b]);
        for (int e2 = 0; e2 < 3; e2++) dJ[a][b][e2] = INV_4PI * (T[P2[a]][P1[a]][b][e2] - T[P1[a]][P2[a]][b][e2]);
      }
    double ub[3], Jb[3][3];
    for (int a = 0; a < 3; a++) {
      ub[a] = aux[a * cap + bi];
      for (int b = 0; b < 3; b++) Jb[a][b] = aux[(3 + 3 * a + b) * cap + bi];
    }
    for (int e2 = 0; e2 < 3; e2++) {
      double v = 0.0;
      for (int a = 0; a < 3; a++) {
        v += ub[a] * Jf[a][e2];
        for (int b = 0; b < 3; b++) v += Jb[a][b] * dJ[a][b][e2];
      }
      aux[(12 + e2) * cap + bi] += v;
    }
  }
}

// ------------------------------------------------------------------ M2L

// Regular harmonics of the 8 unit octant vectors (octant_harmonics).
__global__ void k_octant_table(int P, double2 *__restrict__ Y)
{
  const int o = blockIdx.x;
  if (threadIdx.x < P)
    harmonic_column<false>(P, threadIdx.x, sph((o & 1) ? 1.0 : -1.0, (o & 2) ? 1.0 : -1.0, (o & 4) ? 1.0 : -1.0),
                           Y + o * tri(P), c_inv);
}

// Singular harmonics of the integer lattice vectors k in [-K, K]^3 (k = 0 left zero), in the
// extended-row layout of singular_column_ext: entry (J, mu) of vector k at Y[idx(k) P^2 + J^2 + J + mu].
__global__ __launch_bounds__(WG) void k_m2l_table(int P, double2 *__restrict__ Y)
{
  const int     w   = 2 * M2L_TABLE_K + 1;
  const int64_t idx = blockIdx.x;
  const int     kx = (int)(idx / (w * w)) - M2L_TABLE_K, ky = (int)((idx / w) % w) - M2L_TABLE_K,
            kz = (int)(idx % w) - M2L_TABLE_K;
  double2 *Yk = Y + idx * P * P;
  if (kx == 0 && ky == 0 && kz == 0) {
    for (int e = threadIdx.x; e < P * P; e += WG) Yk[e] = make_double2(0, 0);
    return;
  }
  if (threadIdx.x < P) singular_column_ext(P, threadIdx.x, sph(kx, ky, kz), Yk);
}

// The n-synchronous M2L (plan in m2l_plan.hpp).  One workgroup (W wavefronts) per target
// cell; the 4 groups of 16 lanes of a wavefront take every 4th source.  All lanes of a group walk the source
// degrees n and orders m in the same sequence, so M[n,m] is one LDS broadcast per group, and
// each lane adds the terms of the outputs bound to its S accumulator slots.  Blocks are
// mapped so that each of the 8 XCDs takes one contiguous range of targets: spatially close
// targets share most of their sources, and those multipoles then hit the XCD's own L2.

// Order m >= 1 of output (j,k) at degree n: terms m and -m taken together (multipoleKernels.hpp:534-546),
//   conj(M[n,m]) Y[J,-m-k] + s M[n,m] Y[J,m-k] = (-1)^k (A S' + i B D'),  M = A + iB,
//   S' = Yx[J,m-k] + (-1)^k conj(Y[J,m+k]),  D' = Yx[J,m-k] - (-1)^k conj(Y[J,m+k]),
// with ya = Yx[J,m-k], yb = Yx[J,m+k], Mm = M[n,m] (3 components) and sk = (-1)^k.
__device__ __forceinline__ void m2l_step(double2 (&a)[3], double2 ya, double2 yb, const double2 *__restrict__ Mm,
                                         double sk)
{
  const double2 Sv = make_double2(fma(sk, yb.x, ya.x), fma(-sk, yb.y, ya.y));
  const double2 Dv = make_double2(fma(-sk, yb.x, ya.x), fma(sk, yb.y, ya.y));
  for (int d = 0; d < 3; d++) {
    const double2 mv = Mm[d];
    a[d].x           = fma(mv.x, Sv.x, a[d].x);
    a[d].x           = fma(-mv.y, Dv.y, a[d].x);
    a[d].y           = fma(mv.x, Sv.y, a[d].y);
    a[d].y           = fma(mv.y, Dv.x, a[d].y);
  }
}

// Sums (-1)^k times the terms of output (j,k) at degree n over the orders m0 .. m0 + T, with
// Ya = Yx[J,m0-k], Yb = Yx[J,m0+k], J = j + n, and Mm = M[n,m0]; T is the same in every lane of
// the wavefront, so the loop needs no per-lane control.  The first order is m0 = 0 in a home
// slot: M[n,0] Y[J,-k] = (-1)^k M[n,0] Yx[J,-k] (Y_J^0 is real).  In a helper chunk it is a
// step with sign sk0: sk for m0 > 0, and 0 for m0 = 0, which makes S' = D' = Yx[J,-k] and the
// step's four multiply-adds those of cfma(M[n,0], Yx[J,-k]), in the same order.
__device__ __forceinline__ void m2l_terms_run(double2 (&a)[3], const double2 *__restrict__ Ya,
                                              const double2 *__restrict__ Yb, const double2 *__restrict__ Mm,
                                              double sk, bool home, double sk0, int T)
{
  // the first step's reads are issued ahead of the branch on the slot kind
  const double2 ya = Ya[0], m0v[3] = {Mm[0], Mm[1], Mm[2]};
  double2       Sv = ya, Dv = ya;  // home: cfma(M[n,0], ya) is the step with S' = D' = ya
  if (!home) {
    const double2 yb = Yb[0];
    Sv               = make_double2(fma(sk0, yb.x, ya.x), fma(-sk0, yb.y, ya.y));
    Dv               = make_double2(fma(-sk0, yb.x, ya.x), fma(sk0, yb.y, ya.y));
  }
  for (int d = 0; d < 3; d++) {
    a[d].x = fma(m0v[d].x, Sv.x, a[d].x);
    a[d].x = fma(-m0v[d].y, Dv.y, a[d].x);
    a[d].y = fma(m0v[d].x, Sv.y, a[d].y);
    a[d].y = fma(m0v[d].y, Dv.x, a[d].y);
  }
  // unrolled by two, so the LDS reads of step i+1 are in flight during the multiply-adds of i
#pragma unroll 2
  for (int i = 1; i <= T; i++) m2l_step(a, Ya[i], Yb[i], Mm + 3 * i, sk);
}

// SPLIT: the unit of work is a list segment (seg[3u] = target, [seg[3u+1], seg[3u+2]) = list
// entries), ncell counts segments, and the segment's local expansion goes to L + u * NT;
// k_m2l_reduce adds the segments of each target in order.
// With two wavefronts, up to 5 slots are held to 128 VGPRs (4 waves per SIMD, a few bytes of
// spill at 5); more slots keep their registers.
template <int S, bool SPLIT = false, int W = 1>
__global__ __launch_bounds__(W * WG) __attribute__((amdgpu_waves_per_eu(W == 2 && S <= 5 ? 4 : 1))) void k_m2l_ns(
                                               int P, int NT, int64_t ncell, int64_t per_xcd,
                                               const int32_t *__restrict__ off, const int32_t *__restrict__ src,
                                               const double *__restrict__ tx, const double *__restrict__ ty,
                                               const double *__restrict__ tz, const double *__restrict__ sx,
                                               const double *__restrict__ sy, const double *__restrict__ sz,
                                               const double2 *__restrict__ M, double2 *__restrict__ L,
                                               const int32_t *__restrict__ slots, const int32_t *__restrict__ lev,
                                               const int32_t *__restrict__ out_off,
                                               const int32_t *__restrict__ out_src, const M2LTable tab,
                                               const int32_t *__restrict__ seg = nullptr)
{
  extern __shared__ double2 sm2[];
  __shared__ double sinv[16];  // R0^-(J+1), J = 0..P-1 (lattice table only)
  const int64_t b = blockIdx.x;
  const int64_t u = (b & 7) * per_xcd + (b >> 3);
  if (u >= ncell) return;
  const int64_t t = SPLIT ? seg[3 * u] : u;  // target cell; the output row is u
  // W wavefronts share the staged sources: with W = 2 each takes the source degrees the plan
  // gives it, with its own 16-lane plan for them, and the output loop adds both wavefronts'
  // pieces of each output in a fixed order (W = 2 halves the LDS per wavefront, so up to 5
  // slots run at 4 waves per SIMD); pl numbers the W * 16 lanes of a group for the staging
  const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, sl = lane & 15;
  const int wv = W == 1 ? 0 : __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform in the wavefront
  const int pl = sl + 16 * wv;
  const int ny = tri(P), per = P * P + NT;
  auto      sync = [] {
    if (W == 1) wave_lds_sync();
    else block_lds_sync();
  };
  double2  *Yx = sm2 + g * per;  // extended singular rows, P*P entries
  double2  *Mg = Yx + P * P;
  // This wavefront's plan: slots [16][S] and per-degree words [P] (0: a degree of the other).
  const int32_t *wslots = slots + wv * 16 * S, *wlev = lev + wv * P;
  // Slot s of lane sl, decoded once per workgroup into an int4 of this wavefront's table after
  // the 4 groups' areas (byte offsets; hi(x) = high word of the double x).  At degree n the
  // slot adds orders m0 .. m0 + c - 1 of output (j, k), J = j + n, from Ya = Yx[J, m0 - k] and
  // Yb = Yx[J, m0 + k] = Ya + 32 k, with M[n, m0] = Mn + 48 m0:
  //  * home slot, output w = 16 s + sl = tri(j) + k, active where w < tri(P - n), m0 = 0,
  //    c = n + 1: {16 (j(j+1) - k), 16 (2j + 1), 32 k, hi((-1)^k)}, Ya = Yx + x + y n + 16 n^2;
  //  * helper slot of degree d (a chunk of c = cn orders):
  //    {16 (J(J+1) - k + m0 - d^2), 48 m0, 32 k, hi((-1)^k)}, Ya = Yx + x + 16 d^2;
  //    its first step's sign is 0 for m0 = 0 (the step is then the order-0 term, see
  //    m2l_terms_run);
  //  * unused helper lane: {0, 0, 0, 0} (a sign word of 0 marks it idle).
  int4 *htab = reinterpret_cast<int4 *>(sm2 + 4 * per) + wv * 16 * S;
  for (int q = lane; q < 16 * S; q += WG) {
    const int s = q >> 4, l = q & 15, v = wslots[l * S + s];
    const int j = v & 31, k = (v >> 5) & 31, m0 = (v >> 10) & 31, d = (v >> 15) & 31;
    const int hk = (k & 1) ? (int)0xBFF00000 : 0x3FF00000;
    int4      e  = make_int4(0, 0, 0, 0);
    if (v >= 0 && d == 31) e = make_int4(16 * (j * (j + 1) - k), 16 * (2 * j + 1), 32 * k, hk);
    else if (v >= 0) {
      const int J = j + d;
      e = make_int4(16 * (J * (J + 1) - k + m0 - d * d), 48 * m0, 32 * k, hk);
    }
    htab[q] = e;
  }
  double2 acc[S][3];
  for (int s = 0; s < S; s++)
    for (int d = 0; d < 3; d++) acc[s][d] = make_double2(0, 0);
  int wdeg = 0;  // the source degrees of this wavefront (bit n)
  for (int n = 0; n < P; n++)
    if (wlev[n]) wdeg |= 1 << n;
  const int    s0 = SPLIT ? seg[3 * u + 1] : off[t], nsrc = (SPLIT ? seg[3 * u + 2] : off[t + 1]) - s0;
  const double X = tx[t], Yc = ty[t], Z = tz[t];
  const int    lt = tab.Y ? tab.level[t] : 0;
  // lattice table only: R0^-(J+1) in sinv, and the row J of each harmonic entry this lane copies,
  // e = min(pl + 16 W i, P*P - 1) (J*J <= e <= J*J + 2J), 4 bits per entry i
  using JRows = typename std::conditional<W == 1, uint64_t, uint32_t>::type;
  JRows jrow = 0;
  if (tab.Y) {
    const double inv0 = 1.0 / tab.X0R0[3];
    if (tid < P) {
      double q = inv0;
      for (int i = 0; i < tid; i++) q *= inv0;
      sinv[tid] = q;
    }
    for (int i = 0; i < 2 * (int)sizeof(JRows); i++) {
      const int e = min(pl + 16 * W * i, P * P - 1);
      int       J = 0;
      while ((J + 1) * (J + 1) <= e) J++;
      jrow |= (JRows)J << (4 * i);
    }
    sync();
  }
  // the next batch's source index is loaded a batch ahead, so its latency overlaps the current
  // batch's multiply-adds; the source's centre and level are loaded with its multipole
  int s_cur = g < nsrc ? src[s0 + g] : 0;
  for (int base = 0; base < nsrc; base += 4) {
    const int  si     = base + g;
    const bool active = si < nsrc;
    if (active) {
      const double cx_cur = sx[s_cur], cy_cur = sy[s_cur], cz_cur = sz[s_cur];
      const int    ls_cur = tab.Y ? tab.level[s_cur] : 0;
      // four loads in flight per lane before the first LDS store: at immediate offsets from the
      // lane's first entry while every lane's four entries exist, then with indices past the end
      // clamped to the last entry, whose owner writes the same value (no branches)
      const double2 *Ms = M + (int64_t)s_cur * NT;
      // (the scheduling barriers keep the four loads ahead of the stores)
      int            q0 = 0;
#pragma unroll 1
      for (; q0 + 64 * W <= NT; q0 += 64 * W) {
        double2 v[4];
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] = Ms[pl + q0 + 16 * W * i];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 4; i++) Mg[pl + q0 + 16 * W * i] = v[i];
      }
      if (q0 < NT) {
        double2 v[4];
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] = Ms[min(pl + q0 + 16 * W * i, NT - 1)];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 4; i++) Mg[min(pl + q0 + 16 * W * i, NT - 1)] = v[i];
      }
      // same tree: X_t - X_s = r k with r = R0 / 2^l the smaller cell's half-width and k an
      // integer vector, and Y_n(r k) = r^-(n+1) Y_n(k); offsets inside the table's cube are read
      // from it and scaled per row, the rest (and other trees) are computed here
      bool from_table = false;
      if (tab.Y) {
        const int    l  = max(lt, ls_cur);
        const double sc = ldexp(sinv[0], l);  // 1 / r
        const double qx = (X - cx_cur) * sc, qy = (Yc - cy_cur) * sc, qz = (Z - cz_cur) * sc;
        const double kx = rint(qx), ky = rint(qy), kz = rint(qz);
        const double dev = fmax(fabs(qx - kx), fmax(fabs(qy - ky), fabs(qz - kz)));
        const double km  = fmax(fabs(kx), fmax(fabs(ky), fabs(kz)));
        if (dev < 1e-6 && km <= M2L_TABLE_K && km > 0.0) {
          const int      w   = 2 * M2L_TABLE_K + 1;  // |k| <= K: 32-bit index arithmetic is exact
          const int      idx = ((int)kx + M2L_TABLE_K) * w * w + ((int)ky + M2L_TABLE_K) * w + ((int)kz + M2L_TABLE_K);
          const double2 *Yk  = tab.Y + (int64_t)idx * (P * P);
          // indices past the end are clamped to the last entry (its row is in jrow too)
          for (int e0 = 0, b0 = 0; e0 < P * P; e0 += 64 * W, b0 += 16) {
            double2 y[4];
#pragma unroll
            for (int i = 0; i < 4; i++) y[i] = Yk[min(pl + e0 + 16 * W * i, P * P - 1)];
#pragma unroll
            for (int i = 0; i < 4; i++) {
              const int    e = min(pl + e0 + 16 * W * i, P * P - 1);
              const int    J = (int)(jrow >> (b0 + 4 * i)) & 15;
              const double f = ldexp(sinv[J], l * (J + 1));
              Yx[e]          = make_double2(y[i].x * f, y[i].y * f);
            }
          }
          from_table = true;
        }
      }
      if (!from_table && pl < P) {
        const Sph sp = sph(X - cx_cur, Yc - cy_cur, Z - cz_cur);
        singular_column_ext(P, pl, sp, Yx);
      }
    }
    const int s_next = si + 4 < nsrc ? src[s0 + si + 4] : 0;
    sync();
    if (active) {
      for (int dm = wdeg; dm; dm &= dm - 1) {
        const int      n = __builtin_ctz(dm), lv = wlev[n];
        const int      cn = lv & 255, hr = (lv >> 8) & 255, used = lv >> 16;
        const double2 *Mn = Mg + 3 * tri(n);
        const int      A = tri(P - n), nn = 16 * n * n;  // outputs w < A are active at degree n
        // the table is read per degree: the empty asm on a (zero) uniform offset keeps the
        // compiler from hoisting the S invariant reads out of the degree loop (8 S registers);
        // the slots' entries are then immediate offsets from one lane address
        int tq = 0;
        asm volatile("" : "+s"(tq));
        const int4 *hl = htab + tq + sl;
#pragma unroll
        for (int s = 0; s < S; s++) {
          if (!((used >> s) & 1)) continue;  // no lane has work in slot s at this degree
          // home slot (s < hr): orders 0..n of the active outputs (a prefix of the lanes);
          // otherwise a helper slot of this degree: one chunk of cn orders
          // (a home slot's activity needs no table read: w = 16 s + sl < A)
          const bool home = s < hr;
          const int4 e    = hl[16 * s];
          if (home ? 16 * s + sl < A : e.w != 0) {
            const double2 *Ya = reinterpret_cast<const double2 *>(
                reinterpret_cast<const char *>(Yx) + e.x + (home ? __umul24(e.y, n) : 0) + nn);
            const double2 *Yb = reinterpret_cast<const double2 *>(reinterpret_cast<const char *>(Ya) + e.z);
            const int      em = home ? 0 : e.y;
            const double2 *Mm = reinterpret_cast<const double2 *>(reinterpret_cast<const char *>(Mn) + em);
            m2l_terms_run(acc[s], Ya, Yb, Mm, __hiloint2double(e.w, 0), home, __hiloint2double(em ? e.w : 0, 0),
                          home ? n : cn - 1);
          }
        }
      }
    }
    sync();
    s_cur = s_next;
  }
  // combine the 4 source groups (same slot in lanes sl, sl+16, sl+32, sl+48)
  for (int s = 0; s < S; s++)
    for (int d = 0; d < 3; d++) {
      double2 &a = acc[s][d];
      a.x += __shfl_xor(a.x, 16, 64);
      a.y += __shfl_xor(a.y, 16, 64);
      a.x += __shfl_xor(a.x, 32, 64);
      a.y += __shfl_xor(a.y, 32, 64);
    }
  double2 *part = sm2;  // [W][16][S][3]: every wavefront's slots, read by the output loop
  if (W > 1) sync();
  if (g == 0)
    for (int s = 0; s < S; s++)
      for (int d = 0; d < 3; d++) part[((wv * 16 + sl) * S + s) * 3 + d] = acc[s][d];
  __syncthreads();
  for (int q = tid; q < ny; q += W * WG) {
    int j = 0;
    while (tri(j + 1) <= q) j++;
    double2 r[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    for (int e = out_off[q]; e < out_off[q + 1]; e++) {
      const int pc = out_src[e];
      for (int d = 0; d < 3; d++) {
        r[d].x += part[pc * 3 + d].x;
        r[d].y += part[pc * 3 + d].y;
      }
    }
    const double sjk = sgn(j + q - tri(j));  // (-1)^j of the M2L rule, (-1)^k of the folded rows
    for (int d = 0; d < 3; d++) L[u * NT + 3 * q + d] = make_double2(sjk * r[d].x, sjk * r[d].y);
  }
}

// L of each target = the sum of its segments' expansions, in segment order.
__global__ __launch_bounds__(WG) void k_m2l_reduce(int NT, int64_t ncell, const int32_t *__restrict__ segoff,
                                                   const double2 *__restrict__ part, double2 *__restrict__ L)
{
  const int64_t t = blockIdx.x;
  if (t >= ncell) return;
  for (int q = threadIdx.x; q < NT; q += WG) {
    double2 r = make_double2(0, 0);
    for (int u = segoff[t]; u < segoff[t + 1]; u++) {
      r.x += part[(int64_t)u * NT + q].x;
      r.y += part[(int64_t)u * NT + q].y;
    }
    L[t * NT + q] = r;
  }
}

// ------------------------------------------------------------------ L2P + stretching
// PC > 0 fixes the expansion order at compile time (fully unrolled sums with constant LDS offsets and
// signs); PC = 0 takes it from Prt.
// The L2P's U/V table holds, for output order kk in {1, 2}, the entries (n, mu) with n in
// kk+1 .. P-1 and mu in 1 .. n-kk: tri(P-1-kk) of them (none below P = kk + 2).
__host__ __device__ __forceinline__ int l2p_uv_count(int P, int kk)
{
  const int r = P - 1 - kk;
  return r > 0 ? r * (r + 1) / 2 : 0;
}
// Offset of entry (n, mu) of order kk: the kk = 1 entries come first, each in (n, mu) order.
__device__ __forceinline__ int l2p_uv_index(int P, int kk, int n, int mu)
{
  return (kk == 2 ? l2p_uv_count(P, 1) : 0) + tri(n - kk - 1) + mu - 1;
}

template <int PC>
__global__ __launch_bounds__(WG) void k_l2p(int Prt, int NT, const int32_t *__restrict__ leaves, int64_t nleaf,
                                            const int32_t *__restrict__ cbody, const int32_t *__restrict__ cnbody,
                                            const double *__restrict__ cx, const double *__restrict__ cy,
                                            const double *__restrict__ cz, const double2 *__restrict__ L,
                                            double *__restrict__ st, int64_t cap, int stretching, int formulation,
                                            int sensor_mode)
{
  const int P = PC > 0 ? PC : Prt;
  extern __shared__ double2 sm2[];
  double2      *Lc = sm2;       // [NT]
  double2      *UV = sm2 + NT;  // [l2p_uv_count(P, 1) + l2p_uv_count(P, 2)][3][U, V]
  const int64_t w  = blockIdx.x;
  if (w >= nleaf) return;
  const int c = leaves[w], b0 = cbody[c], nb = cnbody[c];
  for (int q = threadIdx.x; q < NT; q += WG) Lc[q] = L[(int64_t)c * NT + q];
  __syncthreads();
  // Output (j,k), k >= 1, takes the terms mu and -mu of Y[nu, mu] = (p, q) together:
  //   L[n, mu+k] Y + W conj(Y) = (p U.x - q V.y, q V.x + p U.y),  n = nu + j,
  // with W the sign-adjusted entry of the -mu term, U = L[n, mu+k] + W and V = L[n, mu+k] - W.
  // U and V depend on the leaf only, so they are formed here once for all its bodies, which then
  // take 4 multiply-adds per (nu, mu) instead of 8.
  const int nuv1 = l2p_uv_count(P, 1), nuv = nuv1 + l2p_uv_count(P, 2);
  for (int e = threadIdx.x; e < 3 * nuv; e += WG) {
    const int ent = e / 3, d = e % 3;
    const int kk  = ent < nuv1 ? 1 : 2;
    int       r = kk == 1 ? ent : ent - nuv1, n = kk + 1;
    while (r >= n - kk) r -= n - kk, n++;
    const int     mu = r + 1, m2 = kk - mu;
    const double2 l1 = Lc[3 * (tri(n) + mu + kk) + d];
    double2       wv;
    if (m2 < 0) {  // conj(L[n, mu-k]) (-1)^k
      const double2 l = Lc[3 * (tri(n) - m2) + d];
      wv              = make_double2(l.x * sgn(kk), -l.y * sgn(kk));
    } else {  // L[n, k-mu] (-1)^mu
      const double2 l = Lc[3 * (tri(n) + m2) + d];
      wv              = make_double2(l.x * sgn(mu), l.y * sgn(mu));
    }
    UV[2 * e + 0] = make_double2(l1.x + wv.x, l1.y + wv.y);
    UV[2 * e + 1] = make_double2(l1.x - wv.x, l1.y - wv.y);
  }
  __syncthreads();
  const double X = cx[c], Yc = cy[c], Z = cz[c];
  for (int i = threadIdx.x; i < nb; i += WG) {
    const int64_t bi = b0 + i;
    const double  xb[3] = {st[(F_X + 0) * cap + bi], st[(F_X + 1) * cap + bi], st[(F_X + 2) * cap + bi]};
    const Sph     s     = sph(xb[0] - X, xb[1] - Yc, xb[2] - Z);
    // outputs (j,k) = (1,0),(1,1),(2,0),(2,1),(2,2) in slots 1..5 for each component d; the
    // potential itself (slot 0) is not needed
    double2 o[6][3];
    for (int a = 0; a < 6; a++)
      for (int d = 0; d < 3; d++) o[a][d] = make_double2(0, 0);
    const int OJ[6] = {0, 1, 1, 2, 2, 2}, OK[6] = {0, 0, 1, 0, 1, 2};
#pragma unroll(PC > 0 ? PC : 1)
    for (int mu = 0; mu < P; mu++) {
      // keeps each column's coefficient reads inside it (hoisting them all out of the body loop
      // needs more registers than a wavefront has)
      asm volatile("" ::: "memory");
      double2 e = make_double2(1.0, 0.0);
      for (int t = 0; t < mu; t++) e = cmul(e, make_double2(s.cb, s.sb));
      double p = 1.0;
      for (int t = 1; t <= mu; t++) p *= -(double)(2 * t - 1) * s.y;
      double f = c_invfact[2 * mu];
      for (int t = 1; t <= mu; t++) f *= -s.rho;
      double pp = 0.0;
#pragma unroll(PC > 0 ? PC : 1)
      for (int nu = mu; nu < P; nu++) {
        if (nu > mu) {
          const double pn = (nu == mu + 1) ? s.x * (2 * mu + 1) * p
                                           : (s.x * (2 * nu - 1) * p - (nu + mu - 1) * pp) * c_inv[nu - mu];
          pp = p;
          p  = pn;
          f *= -s.rho * c_inv[nu + mu];
        }
        const double  v  = f * p;
        const double2 yp = make_double2(v * e.x, v * e.y);  // Y[nu, +mu]
        for (int a = 1; a < 6; a++) {
          const int j = OJ[a], k = OK[a];
          const int n = nu + j;
          if (n >= P) continue;
          if (mu == 0) {  // the single term m = k (sign +1)
            for (int d = 0; d < 3; d++) cfma(o[a][d], Lc[3 * (tri(n) + k) + d], yp);
          } else if (k == 0) {  // the -mu term is the conjugate of the +mu term: 2 Re(L[n, mu] Y)
            for (int d = 0; d < 3; d++) {
              const double2 l = Lc[3 * (tri(n) + mu) + d];
              o[a][d].x       = fma(l.x, 2.0 * yp.x, o[a][d].x);
              o[a][d].x       = fma(-l.y, 2.0 * yp.y, o[a][d].x);
            }
          } else {
            const double2 *uv = UV + 6 * l2p_uv_index(P, k, n, mu);
            for (int d = 0; d < 3; d++) {
              const double2 U = uv[2 * d], V = uv[2 * d + 1];
              o[a][d].x       = fma(yp.x, U.x, o[a][d].x);
              o[a][d].x       = fma(-yp.y, V.y, o[a][d].x);
              o[a][d].y       = fma(yp.y, V.x, o[a][d].y);
              o[a][d].y       = fma(yp.x, U.y, o[a][d].y);
            }
          }
        }
      }
    }
    // gradient g[d][b] = d psi_d / d x_b, Hessian H[d][b][c]
    double gr[3][3], H[3][3][3];
    for (int d = 0; d < 3; d++) {
      const double a10 = o[1][d].x, a11 = o[2][d].x, b11 = o[2][d].y;
      const double a20 = o[3][d].x, a21 = o[4][d].x, b21 = o[4][d].y, a22 = o[5][d].x, b22 = o[5][d].y;
      gr[d][0] = a11;
      gr[d][1] = -b11;
      gr[d][2] = -a10;
      H[d][0][0] = 0.5 * (a22 - a20);
      H[d][1][1] = -0.5 * (a20 + a22);
      H[d][2][2] = a20;
      H[d][0][1] = H[d][1][0] = -0.5 * b22;
      H[d][0][2] = H[d][2][0] = -a21;
      H[d][1][2] = H[d][2][1] = b21;
    }
    // u = curl(psi)/(4 pi); J[a][b] = d u_a / d x_b
    double u[3], Jf[3][3];
    u[0] = gr[2][1] - gr[1][2];
    u[1] = gr[0][2] - gr[2][0];
    u[2] = gr[1][0] - gr[0][1];
    for (int b = 0; b < 3; b++) {
      Jf[0][b] = H[2][1][b] - H[1][2][b];
      Jf[1][b] = H[0][2][b] - H[2][0][b];
      Jf[2][b] = H[1][0][b] - H[0][1][b];
    }
    double uu[3], JJ[3][3];
    for (int d = 0; d < 3; d++) {
      uu[d] = st[(F_U + d) * cap + bi] + INV_4PI * u[d];
      st[(F_U + d) * cap + bi] = uu[d];
    }
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        JJ[a][b] = st[(F_J + 3 * a + b) * cap + bi] + INV_4PI * Jf[a][b];
        st[(F_J + 3 * a + b) * cap + bi] = JJ[a][b];
      }
    if (sensor_mode) continue;
    // stretching (fmmCalculator.hpp:391-405) and rVPM (407-418)
    const double al[3] = {st[(F_A + 0) * cap + bi], st[(F_A + 1) * cap + bi], st[(F_A + 2) * cap + bi]};
    double       da[3];
    for (int d = 0; d < 3; d++)
      da[d] = stretching == UVLM_TRANSPOSED ? JJ[0][d] * al[0] + JJ[1][d] * al[1] + JJ[2][d] * al[2]
                                            : JJ[d][0] * al[0] + JJ[d][1] * al[1] + JJ[d][2] * al[2];
    double dr = 0.0;
    if (formulation == UVLM_RVPM) {
      const double na    = sqrt(al[0] * al[0] + al[1] * al[1] + al[2] * al[2]);
      const double ah[3] = {al[0] / na, al[1] / na, al[2] / na};
      const double dot   = da[0] * ah[0] + da[1] * ah[1] + da[2] * ah[2];
      for (int d = 0; d < 3; d++) da[d] -= 3. / 5. * dot * ah[d];
      dr -= 1.0 / 5.0 * st[F_SIG * cap + bi] / na * dot;
    }
    for (int d = 0; d < 3; d++) st[(F_DADT + d) * cap + bi] = da[d];
    st[F_DSIG * cap + bi] = dr;
  }
}

__global__ void k_zero(double *p, int64_t n)
{
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

}  // namespace

// ---------------------------------------------------------------------------------
// M2L schedule: the work plan of k_m2l_ns (m2l_plan.hpp) for the handle's slot budget, copied
// to the device, and the lattice and octant harmonic tables for P.
// Two wavefronts per workgroup when one wavefront's LDS (16 (P^2 + NT) bytes per staged source,
// 4 sources) would cap the occupancy below what its registers allow: from P = 9 on (measured:
// P = 8 is 12% faster with one, P = 10 10% faster with two, P = 12 and 14 a third).
int m2l_waves(int P, int waves) { return waves == 1 || waves == 2 ? waves : (P <= 8 ? 1 : 2); }

// Term table of k_m2m_tab for P <= M2M_TAB_MAXP: per lane (output q = (j,k) = lane, tri(P) <= 64)
// the counts of its plain and conjugated terms, then RP rows of plain terms and RC rows of
// conjugated terms, 64 lanes per row, in the order in which k_m2m takes them (n, then m).
static void make_m2m_table(int P, std::vector<int32_t> &cnt, std::vector<uint16_t> &tab, int &RP, int &RC)
{
  const int nq = P * (P + 1) / 2;
  auto tr  = [](int n) { return n * (n + 1) / 2; };
  auto neg = [](int e) { return (e & 1) != 0; };  // (-1)^e < 0
  std::vector<std::vector<uint16_t>> pl(WG), cj(WG);
  // copy c of the octant harmonics: 0 Y, 1 -Y, 2 conj Y, 3 -conj Y
  auto entry = [&](int qs, int idx, bool conjY, bool minus) {
    return (uint16_t)(qs | (((conjY ? 2 : 0) + (minus ? 1 : 0)) * nq + idx) << 6);
  };
  for (int j = 0; j < P; j++)
    for (int k = 0; k <= j; k++) {
      const int q = tr(j) + k;
      for (int n = 0; n <= j; n++) {
        // Mc[j-n, k-m] Y[n,-m] (-1)^{m<0 ? m : 0} (-1)^n for m < k
        for (int m = std::max(-n, -j + k + n); m <= std::min(k - 1, n); m++) {
          const bool minus = (m < 0 && neg(m)) != neg(n);
          pl[q].push_back(m <= 0 ? entry(tr(j - n) + k - m, tr(n) - m, false, minus)
                                 : entry(tr(j - n) + k - m, tr(n) + m, true, minus));
        }
        // conj(Mc[j-n, m-k]) Y[n,-m] (-1)^{k+n+m} for m >= k, summed as Mc conj(Y[n,-m] (-1)^{k+n+m})
        for (int m = k; m <= std::min(n, j + k - n); m++) {
          const bool minus = neg(k + n + m);
          cj[q].push_back(m == 0 ? entry(tr(j - n) - k + m, tr(n), true, minus)
                                 : entry(tr(j - n) - k + m, tr(n) + m, false, minus));
        }
      }
    }
  RP = RC = 0;
  cnt.assign(WG, 0);
  for (int q = 0; q < nq; q++) {
    RP     = std::max(RP, (int)pl[q].size());
    RC     = std::max(RC, (int)cj[q].size());
    cnt[q] = (int32_t)pl[q].size() | ((int32_t)cj[q].size() << 8);
  }
  tab.assign((size_t)(RP + RC) * WG, 0);
  for (int q = 0; q < nq; q++) {
    for (size_t i = 0; i < pl[q].size(); i++) tab[i * WG + q] = pl[q][i];
    for (size_t i = 0; i < cj[q].size(); i++) tab[(RP + i) * WG + q] = cj[q][i];
  }
}

// Term table of k_l2l_tab: for each lane, its terms over the groups d of k_l2l_bal in the order
// k_l2l_bal takes them (row offsets fixed by P), L2L_TAB_ROWS 16-bit entries per lane.
static bool make_l2l_table(int P, std::vector<uint16_t> &tab)
{
  const int nq = P * (P + 1) / 2;
  auto tr  = [](int n) { return n * (n + 1) / 2; };
  auto neg = [](int e) { return (e & 1) != 0; };
  tab.assign((size_t)WG * L2L_TAB_ROWS, 0);
  // term i of output q in group d: n = j + d, m = k - d + i, Y[d, m - k] (k_l2l_bal's terms)
  auto entry = [&](int q, int d, int i) {
    int j = 0;
    while (tr(j + 1) <= q) j++;
    const int  k = q - tr(j), n = j + d, m = k - d + i, mu = m - k;
    const int  yidx  = tr(d) + (mu >= 0 ? mu : -mu);
    const bool conjY = mu < 0;  // ylook(Y, d, mu)
    bool       minus, conjL;
    int        lq;
    if (m < 0) minus = neg(k), conjL = true, lq = tr(n) - m;
    else minus = m < k ? neg(m - k) : false, conjL = false, lq = tr(n) + m;
    const int v = ((conjY ? 2 : 0) + (minus ? 1 : 0)) * nq + yidx;
    return (uint16_t)(lq | (v << 6) | ((conjL ? 1 : 0) << 14));
  };
  int row = 0;
  for (int d = 0; d < P; d++) {
    const int ntask = tr(P - d), nterm = 2 * d + 1, s = l2l_split(P, d);
    const int nr = s == 1 ? nterm : (nterm + s - 1) / s;
    if (row + nr > L2L_TAB_ROWS) return false;
    for (int lane = 0; lane < WG; lane++) {
      const int q = s == 1 ? lane : lane / s, first = s == 1 ? 0 : lane % s;
      if (q >= ntask) continue;
      for (int i = 0; i < nr; i++) {
        const int t = first + i * s;
        if (t < nterm) tab[(size_t)lane * L2L_TAB_ROWS + row + i] = entry(q, d, t);
      }
    }
    row += nr;
  }
  return true;
}

void make_m2l_schedule(int P, M2LSchedule &s)
{
  // Default budget: with two wavefronts, 5 slots run at 4 waves per SIMD at P = 10.
  s.W              = m2l_waves(P, s.waves);
  const int budget = s.budget > 0 ? s.budget : (s.W == 2 ? 5 : 0);
  if (!make_m2l_plan(P, budget, s.plan, 16, s.W)) throw Error(UVLM_ERR_INTERNAL, "no M2L plan");
  const M2LPlan &pl = s.plan;
  s.dev_slots       = 0;
  s.dev_lev         = pl.slots.size();
  s.dev_out_off     = s.dev_lev + pl.lev.size();
  s.dev_out_src     = s.dev_out_off + pl.out_off.size();
  std::vector<int32_t> all;
  all.insert(all.end(), pl.slots.begin(), pl.slots.end());
  all.insert(all.end(), pl.lev.begin(), pl.lev.end());
  all.insert(all.end(), pl.out_off.begin(), pl.out_off.end());
  all.insert(all.end(), pl.out_src.begin(), pl.out_src.end());
  s.dev.reserve(4 * all.size());
  UVLM_HIP(hipMemcpy(s.dev.p, all.data(), 4 * all.size(), hipMemcpyHostToDevice));
  if (s.ytab_P != P) {
    const int w = 2 * M2L_TABLE_K + 1;
    s.ytab.reserve(sizeof(double2) * (size_t)w * w * w * P * P);
    k_m2l_table<<<w * w * w, WG>>>(P, s.ytab.as<double2>());
    UVLM_LAUNCH_CHECK();
    s.octY.reserve(sizeof(double2) * 8 * P * (P + 1) / 2);
    k_octant_table<<<8, WG>>>(P, s.octY.as<double2>());
    UVLM_LAUNCH_CHECK();
    s.l2l_tab = false;
    if (P == 6 || P == 8 || P == 10) {
      std::vector<uint16_t> tab;
      if (make_l2l_table(P, tab)) {
        s.l2ltab.reserve(2 * tab.size());
        UVLM_HIP(hipMemcpy(s.l2ltab.p, tab.data(), 2 * tab.size(), hipMemcpyHostToDevice));
        s.l2l_tab = true;
      }
    }
    s.m2m_rp = s.m2m_rc = 0;
    if (P <= M2M_TAB_MAXP) {
      std::vector<int32_t>  cnt;
      std::vector<uint16_t> tab;
      make_m2m_table(P, cnt, tab, s.m2m_rp, s.m2m_rc);
      s.m2mtab.reserve(4 * cnt.size() + 2 * tab.size());
      UVLM_HIP(hipMemcpy(s.m2mtab.p, cnt.data(), 4 * cnt.size(), hipMemcpyHostToDevice));
      UVLM_HIP(hipMemcpy((char *)s.m2mtab.p + 4 * cnt.size(), tab.data(), 2 * tab.size(), hipMemcpyHostToDevice));
    }
    UVLM_HIP(hipDeviceSynchronize());
    s.ytab_P = P;
  }
}

void zero_doubles(hipStream_t s, double *p, int64_t n)
{
  if (n <= 0) return;
  k_zero<<<nblocks(n, 256), 256, 0, s>>>(p, n);
  UVLM_LAUNCH_CHECK();
}

void p2m(hipStream_t s, int P, const Tree &t, const double *st, int64_t cap, double2 *M)
{
  const int NT = 3 * P * (P + 1) / 2;
  if (t.nleaf == 0) return;
  const int    B   = WG / ((P + 1) / 2), ny = P * (P + 1) / 2;  // bodies per chunk (k_p2m)
  const size_t lds = sizeof(double2) * B * ny + sizeof(double) * 3 * B;
  auto launch = [&](auto kern) {
    kern<<<(int)t.nleaf, WG, lds, s>>>(P, NT, t.leaves.as<int32_t>(), t.nleaf, t.body.as<int32_t>(),
                                       t.nbody.as<int32_t>(), t.cx.as<double>(), t.cy.as<double>(), t.cz.as<double>(),
                                       st, cap, M);
  };
  switch (P) {
  case 6: launch(k_p2m<6>); break;
  case 8: launch(k_p2m<8>); break;
  case 10: launch(k_p2m<10>); break;
  case 12: launch(k_p2m<12>); break;
  default: launch(k_p2m<0>);
  }
  UVLM_LAUNCH_CHECK();
}

void m2m_level(hipStream_t s, int P, const Tree &t, int level, double2 *M, const M2LSchedule &sch)
{
  const int     NT = 3 * P * (P + 1) / 2;
  const int64_t o = t.lvl_off[level], n = t.lvl_off[level + 1] - o;
  if (n <= 0) return;
  const double2 *octY = sch.octY.as<double2>();
  if (P <= M2M_TAB_MAXP && sch.m2m_rp + sch.m2m_rc > 0) {
    const size_t slice = sizeof(double2) * (4 * P * (P + 1) / 2 + NT);
    const size_t lds   = std::max(slice * M2M_MAXWV, sizeof(double2) * 3 * WG * M2M_MAXWV) +
                       sizeof(uint16_t) * WG * (sch.m2m_rp + sch.m2m_rc);
    const int32_t *cnt = reinterpret_cast<const int32_t *>(sch.m2mtab.p);
    k_m2m_tab<<<(int)n, M2M_MAXWV * WG, lds, s>>>(
        P, NT, t.lvl_cells.as<int32_t>() + o, n, t.nchild.as<int32_t>(), t.child.as<int32_t>(), t.cx.as<double>(),
        t.cy.as<double>(), t.cz.as<double>(), t.R.as<double>(), octY, cnt, reinterpret_cast<const uint32_t *>(cnt + WG),
        sch.m2m_rp, sch.m2m_rc, M);
    UVLM_LAUNCH_CHECK();
    return;
  }
  // wavefronts per workgroup: 8 LDS slices of (harmonics + multipole) up to 64 KB, else 4
  const size_t slice = sizeof(double2) * (P * (P + 1) / 2 + NT);
  const int    nwv   = slice * M2M_MAXWV <= 65536 ? M2M_MAXWV : M2M_MAXWV / 2;
  const size_t lds   = std::max(slice * nwv, sizeof(double2) * 3 * WG * nwv);
  auto launch = [&](auto kern) {
    kern<<<(int)n, nwv * WG, lds, s>>>(P, NT, t.lvl_cells.as<int32_t>() + o, n, t.nchild.as<int32_t>(),
                                       t.child.as<int32_t>(), t.cx.as<double>(), t.cy.as<double>(), t.cz.as<double>(),
                                       t.R.as<double>(), octY, M);
  };
  switch (P) {
  case 6: launch(k_m2m<6>); break;
  case 8: launch(k_m2m<8>); break;
  case 10: launch(k_m2m<10>); break;
  case 12: launch(k_m2m<12>); break;
  default: launch(k_m2m<0>);
  }
  UVLM_LAUNCH_CHECK();
}

void l2l_level(hipStream_t s, int P, const Tree &t, int level, double2 *L, const M2LSchedule &sch)
{
  const int     NT = 3 * P * (P + 1) / 2;
  const int64_t o = t.lvl_off[level], n = t.lvl_off[level + 1] - o;
  if (n <= 0) return;
  const double2 *octY = sch.octY.as<double2>();
  if (sch.l2l_tab && (P == 6 || P == 8 || P == 10)) {
    const size_t lds = sizeof(double2) * (4 * P * (P + 1) / 2 + NT);
    auto launch = [&](auto kern) {
      kern<<<(int)n, WG, lds, s>>>(NT, t.lvl_cells.as<int32_t>() + o, n, t.parent.as<int32_t>(), t.cx.as<double>(),
                                   t.cy.as<double>(), t.cz.as<double>(), t.R.as<double>(), octY,
                                   reinterpret_cast<const uint4 *>(sch.l2ltab.p), L);
    };
    if (P == 6) launch(k_l2l_tab<6>);
    else if (P == 8) launch(k_l2l_tab<8>);
    else launch(k_l2l_tab<10>);
    UVLM_LAUNCH_CHECK();
    return;
  }
  const size_t lds = sizeof(double2) * (P * (P + 1) / 2 + NT);
  auto launch = [&](auto kern) {
    kern<<<(int)n, WG, lds, s>>>(P, NT, t.lvl_cells.as<int32_t>() + o, n, t.parent.as<int32_t>(), t.cx.as<double>(),
                                 t.cy.as<double>(), t.cz.as<double>(), t.R.as<double>(), octY, L);
  };
  switch (P) {
  case 6: launch(k_l2l_bal<6>); break;
  case 8: launch(k_l2l_bal<8>); break;
  case 10: launch(k_l2l_bal<10>); break;
  case 12: launch(k_l2l_bal<12>); break;
  default: launch(k_l2l_bal<0>);
  }
  UVLM_LAUNCH_CHECK();
}

// Below this many target cells (probe trees), long lists are cut into segments that run as
// separate workgroups (see the P2P launcher for the same rule).
constexpr int64_t M2L_SPLIT_TARGETS = 4096, M2L_SPLIT_SEGMENTS = 8192;

void m2l(hipStream_t s, int P, const Tree &tt, const Tree &st, const Lists &l, const M2LSchedule &sch, const double2 *M,
         double2 *L, Scratch *sc)
{
  const int NT = 3 * P * (P + 1) / 2;
  if (tt.C == 0) return;
  const int32_t *d = sch.dev.as<int32_t>();
  {
    const int    S = sch.plan.S, W = sch.W;
    const size_t lds_ns = std::max(sizeof(double2) * 4 * (P * P + NT) + sizeof(int4) * W * 16 * S,
                                   sizeof(double2) * W * 16 * S * 3);
    // few targets with long lists: one workgroup per list segment, partial expansions reduced
    const bool     split = sc && tt.C < M2L_SPLIT_TARGETS && l.nm2l >= 4 * tt.C;
    int64_t        units = tt.C;
    const int32_t *dseg = nullptr, *dsegoff = nullptr;
    double2       *out  = L;
    if (split) {
      const int64_t        E = std::max<int64_t>(2, (l.nm2l + M2L_SPLIT_SEGMENTS - 1) / M2L_SPLIT_SEGMENTS);
      std::vector<int32_t> off(tt.C + 1), seg, segoff(tt.C + 1, 0);
      UVLM_HIP(hipMemcpyAsync(off.data(), l.m2l_off.p, 4 * off.size(), hipMemcpyDeviceToHost, s));
      UVLM_HIP(hipStreamSynchronize(s));
      for (int64_t t = 0; t < tt.C; t++) {
        int32_t e0 = off[t];
        do {
          const int32_t e1 = (int32_t)std::min<int64_t>(off[t + 1], e0 + E);
          seg.push_back((int32_t)t);
          seg.push_back(e0);
          seg.push_back(e1);
          e0 = e1;
        } while (e0 < off[t + 1]);
        segoff[t + 1] = (int32_t)(seg.size() / 3);
      }
      units = segoff[tt.C];
      sc->m2l_seg.reserve(4 * (seg.size() + segoff.size()));
      int32_t *ds = sc->m2l_seg.as<int32_t>();
      UVLM_HIP(hipMemcpyAsync(ds, seg.data(), 4 * seg.size(), hipMemcpyHostToDevice, s));
      UVLM_HIP(hipMemcpyAsync(ds + seg.size(), segoff.data(), 4 * segoff.size(), hipMemcpyHostToDevice, s));
      UVLM_HIP(hipStreamSynchronize(s));
      dseg    = ds;
      dsegoff = ds + seg.size();
      sc->m2l_part.reserve(sizeof(double2) * units * NT);
      out = sc->m2l_part.as<double2>();
    }
    const int64_t per_xcd = (units + 7) / 8;
    const int     grid    = (int)(8 * per_xcd);
    // the lattice table needs one tree (one root, one unit): particle evaluations and their adjoints
    M2LTable tab{};
    if (&tt == &st && sch.ytab_P == P)
      tab = M2LTable{sch.ytab.as<double2>(), tt.level.as<int32_t>(), tt.X0R0.as<double>()};
#define UVLM_M2L_LAUNCH(SS, WW)                                                                              \
  if (split)                                                                                               \
    k_m2l_ns<SS, true, WW><<<grid, WW * WG, lds_ns, s>>>(                                                  \
        P, NT, units, per_xcd, l.m2l_off.as<int32_t>(), l.m2l_src.as<int32_t>(), tt.cx.as<double>(),       \
        tt.cy.as<double>(), tt.cz.as<double>(), st.cx.as<double>(), st.cy.as<double>(), st.cz.as<double>(), M, \
        out, d + sch.dev_slots, d + sch.dev_lev, d + sch.dev_out_off, d + sch.dev_out_src, tab, dseg);    \
  else                                                                                                     \
    k_m2l_ns<SS, false, WW><<<grid, WW * WG, lds_ns, s>>>(                                                 \
        P, NT, tt.C, per_xcd, l.m2l_off.as<int32_t>(), l.m2l_src.as<int32_t>(), tt.cx.as<double>(),        \
        tt.cy.as<double>(), tt.cz.as<double>(), st.cx.as<double>(), st.cy.as<double>(), st.cz.as<double>(), M, \
        L, d + sch.dev_slots, d + sch.dev_lev, d + sch.dev_out_off, d + sch.dev_out_src, tab);
#define UVLM_M2L_NS(SS)                                                                                    \
  if (S == SS) {                                                                                           \
    if (W == 2) {                                                                                          \
      UVLM_M2L_LAUNCH(SS, 2)                                                                               \
    } else {                                                                                               \
      UVLM_M2L_LAUNCH(SS, 1)                                                                               \
    }                                                                                                      \
    UVLM_LAUNCH_CHECK();                                                                                   \
    if (split) {                                                                                           \
      k_m2l_reduce<<<(int)tt.C, WG, 0, s>>>(NT, tt.C, dsegoff, out, L);                                    \
      UVLM_LAUNCH_CHECK();                                                                                 \
    }                                                                                                      \
    return;                                                                                                \
  }
    UVLM_M2L_NS(4)
    UVLM_M2L_NS(5)
    UVLM_M2L_NS(6)
    UVLM_M2L_NS(8)
    UVLM_M2L_NS(12)
    UVLM_M2L_NS(16)
#undef UVLM_M2L_NS
#undef UVLM_M2L_LAUNCH
    throw Error(UVLM_ERR_INTERNAL, "unsupported M2L plan size");
  }
}

void l2p_stretch(hipStream_t s, int P, const Tree &t, const double2 *L, double *st, int64_t cap, int stretching,
                 int formulation, int sensor_mode)
{
  const int NT = 3 * P * (P + 1) / 2;
  if (t.nleaf == 0) return;
  // local expansion, then the U/V table of k_l2p: 3 components x (U, V) per (order, n, mu)
  const size_t lds = sizeof(double2) * (NT + 6 * (l2p_uv_count(P, 1) + l2p_uv_count(P, 2)));
  auto launch = [&](auto kern) {
    kern<<<(int)t.nleaf, WG, lds, s>>>(P, NT, t.leaves.as<int32_t>(), t.nleaf, t.body.as<int32_t>(),
                                                       t.nbody.as<int32_t>(), t.cx.as<double>(), t.cy.as<double>(),
                                                       t.cz.as<double>(), L, st, cap, stretching, formulation,
                                                       sensor_mode);
  };
  switch (P) {
  case 6: launch(k_l2p<6>); break;
  case 8: launch(k_l2p<8>); break;
  case 10: launch(k_l2p<10>); break;
  case 12: launch(k_l2p<12>); break;
  default: launch(k_l2p<0>);
  }
  UVLM_LAUNCH_CHECK();
}

void p2m_adj(hipStream_t s, int P, const Tree &t, const double *st, const double *aux, int64_t cap, double2 *M)
{
  const int NT = 3 * P * (P + 1) / 2;
  if (t.nleaf == 0) return;
  const int    B   = WG / ((P + 1) / 2), ny = P * (P + 1) / 2;
  const size_t lds = sizeof(double2) * B * ny + sizeof(double) * 12 * B;
  k_p2m_adj<<<(int)t.nleaf, WG, lds, s>>>(P, NT, t.leaves.as<int32_t>(), t.nleaf, t.body.as<int32_t>(),
                                          t.nbody.as<int32_t>(), t.cx.as<double>(), t.cy.as<double>(),
                                          t.cz.as<double>(), st, aux, cap, M);
  UVLM_LAUNCH_CHECK();
}

void l2p_adj_tgt(hipStream_t s, int P, const Tree &t, const double2 *L, const double *st, double *aux, int64_t cap)
{
  const int NT = 3 * P * (P + 1) / 2;
  if (t.nleaf == 0) return;
  auto launch = [&](auto kern) {
    kern<<<(int)t.nleaf, WG, sizeof(double2) * NT, s>>>(P, NT, t.leaves.as<int32_t>(), t.nleaf, t.body.as<int32_t>(),
                                                       t.nbody.as<int32_t>(), t.cx.as<double>(), t.cy.as<double>(),
                                                       t.cz.as<double>(), L, st, aux, cap);
  };
  switch (P) {
  case 6: launch(k_l2p_adj_tgt<6>); break;
  case 8: launch(k_l2p_adj_tgt<8>); break;
  case 10: launch(k_l2p_adj_tgt<10>); break;
  case 12: launch(k_l2p_adj_tgt<12>); break;
  default: launch(k_l2p_adj_tgt<0>);
  }
  UVLM_LAUNCH_CHECK();
}

}  // namespace uvlm
