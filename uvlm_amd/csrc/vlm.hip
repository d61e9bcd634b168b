// vlm.hip -- the vortex-lattice wing that drives the particle wake (src/main.cpp:11-295), FP64.
//
// Lattice: vertices V[r][c] (r = 0 .. nchord, row 0 the leading edge; c = 0 .. nspan), AoS
// doubles.  nchord = 1 is the reference's wing (LE_Vertices[c] = V[0][c], TE_Vertices[c] =
// V[1][c], main.cpp:27-74).  Panel p = r * nspan + c has corners LE_c = V[r][c],
// LE_c1 = V[r][c+1], TE_c = V[r+1][c], TE_c1 = V[r+1][c+1] and the ring of lines
// (getPanelVortexLine, main.cpp:60-73) TE_c1->TE_c, TE_c->LE_c, LE_c->LE_c1, LE_c1->TE_c1.
//
// Kernels: geometry (calculateTopology, 45-58), ring influences B[t][j] at points (one lane per
// (point, panel), the segment formula of 11-25), the system matrix A[i][j] = B[i][j] . n_i,
// a one-workgroup LU with partial pivoting and the inverse from it (in place of the reference's
// fullPivLu solve, 214; the matrix is fixed under the loop's rigid motion),
// the right-hand side (207-212), the Kutta-Joukowski force (223-244) and the shed particles
// (246-289).  Every sum runs in a fixed order (deterministic).
#include "common.hpp"

namespace uvlm {

namespace {

constexpr int LU_WG = 1024;

struct V3 {
  double x, y, z;
};
__device__ inline V3 ld3(const double *p) { return V3{p[0], p[1], p[2]}; }
__device__ inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
__device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct Corners {
  V3 le, le1, te, te1;
};
__device__ inline Corners corners(const double *V, int nspan, int p)
{
  const int r = p / nspan, c = p % nspan, w = nspan + 1;
  return Corners{ld3(V + 3 * (r * w + c)), ld3(V + 3 * (r * w + c + 1)), ld3(V + 3 * ((r + 1) * w + c)),
                 ld3(V + 3 * ((r + 1) * w + c + 1))};
}

// Unit vortex segment a -> b at point x (main.cpp:11-25); flags the reference's singular case
// (|r1|, |r2| or |r1 x r2| below 1e-6, where it calls std::exit) and returns zero there.
__device__ inline V3 segment(V3 a, V3 b, V3 x, int *singular)
{
  const V3     r1 = x - a, r2 = x - b, r0 = b - a;
  const V3     c  = cross(r1, r2);
  const double n1 = sqrt(dot(r1, r1)), n2 = sqrt(dot(r2, r2)), c2 = dot(c, c);
  if (n1 < 1e-6 || n2 < 1e-6 || sqrt(c2) < 1e-6) {
    atomicAdd(singular, 1);
    return V3{0, 0, 0};
  }
  const double k = (r0.x * (r1.x / n1 - r2.x / n2) + r0.y * (r1.y / n1 - r2.y / n2) + r0.z * (r1.z / n1 - r2.z / n2)) /
                   c2 * INV_4PI;
  return k * c;
}

__device__ inline V3 ring(const Corners &q, V3 x, int *singular)
{
  return segment(q.te1, q.te, x, singular) + segment(q.te, q.le, x, singular) + segment(q.le, q.le1, x, singular) +
         segment(q.le1, q.te1, x, singular);
}

// calculateTopology (main.cpp:45-58): control point, unit normal, area of each panel.
__global__ void k_wing_geom(int n, int nspan, const double *__restrict__ V, double *__restrict__ cp,
                            double *__restrict__ nrm, double *__restrict__ area)
{
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const Corners q = corners(V, nspan, p);
  const V3      m = 0.25 * (q.le + q.le1 + q.te + q.te1);
  V3            v = cross(q.le1 - q.te, q.le - q.te1);
  const double  l = sqrt(dot(v, v));
  v             = (1.0 / l) * v;
  cp[3 * p] = m.x, cp[3 * p + 1] = m.y, cp[3 * p + 2] = m.z;
  nrm[3 * p] = v.x, nrm[3 * p + 1] = v.y, nrm[3 * p + 2] = v.z;
  area[p] = 0.5 * l;
}

// B[t][j] (3 doubles) = velocity at pts[t] of ring j with unit circulation; with nrm given (the
// points are the control points), also A[t][j] = B[t][j] . n_t.
__global__ void k_wing_influence(int m, int n, int nspan, const double *__restrict__ V, const double *__restrict__ pts,
                                 const double *__restrict__ nrm, double *__restrict__ B, double *__restrict__ A,
                                 int *singular)
{
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= (int64_t)m * n) return;
  const int t = (int)(e / n), j = (int)(e % n);
  const V3  u = ring(corners(V, nspan, j), ld3(pts + 3 * t), singular);
  B[3 * e] = u.x, B[3 * e + 1] = u.y, B[3 * e + 2] = u.z;
  if (A) A[e] = dot(u, ld3(nrm + 3 * t));
}

// In-place LU with partial pivoting of the row-major n x n matrix A by one workgroup: pivot
// = largest |A[i][k]|, i >= k (smallest i on ties); piv[k] = the row swapped with k;
// *info = 1 + the first k with a zero pivot (0 if none).
__global__ __launch_bounds__(LU_WG) void k_lu(int n, double *__restrict__ A, int *__restrict__ piv, int *info)
{
  __shared__ double sv[LU_WG];
  __shared__ int    si[LU_WG];
  const int         tid = threadIdx.x;
  if (tid == 0) *info = 0;
  for (int k = 0; k < n; k++) {
    double bv = -1.0;
    int    bi = n;
    for (int i = k + tid; i < n; i += LU_WG) {
      const double v = fabs(A[(int64_t)i * n + k]);
      if (v > bv) {
        bv = v;
        bi = i;
      }
    }
    sv[tid] = bv;
    si[tid] = bi;
    __syncthreads();
    for (int o = LU_WG / 2; o > 0; o >>= 1) {
      if (tid < o) {
        const double v = sv[tid + o];
        const int    i = si[tid + o];
        if (v > sv[tid] || (v == sv[tid] && i < si[tid])) {
          sv[tid] = v;
          si[tid] = i;
        }
      }
      __syncthreads();
    }
    const int p = si[0];
    __syncthreads();
    if (tid == 0) piv[k] = p;
    if (p != k)
      for (int j = tid; j < n; j += LU_WG) {
        const double a = A[(int64_t)k * n + j];
        A[(int64_t)k * n + j] = A[(int64_t)p * n + j];
        A[(int64_t)p * n + j] = a;
      }
    __syncthreads();
    const double d = A[(int64_t)k * n + k];
    if (d == 0.0) {
      if (tid == 0 && *info == 0) *info = k + 1;
      __syncthreads();
      continue;
    }
    for (int i = k + 1 + tid; i < n; i += LU_WG) A[(int64_t)i * n + k] /= d;
    __syncthreads();
    const int r = n - k - 1;
    for (int64_t e = tid; e < (int64_t)r * r; e += LU_WG) {
      const int i = k + 1 + (int)(e / r), j = k + 1 + (int)(e % r);
      A[(int64_t)i * n + j] -= A[(int64_t)i * n + k] * A[(int64_t)k * n + j];
    }
    __syncthreads();
  }
}

// Columns of A^{-1} from k_lu's factors, one workgroup per column c: e_c permuted by the pivots,
// then the unit-lower and the upper triangular solves; Xc[c * n + i] = (A^{-1})[i][c]
// (column-major).
__global__ __launch_bounds__(256) void k_lu_inverse(int n, const double *__restrict__ LU, const int *__restrict__ piv,
                                                    double *__restrict__ Xc)
{
  const int c = blockIdx.x, tid = threadIdx.x;
  if (c >= n) return;
  double *b = Xc + (int64_t)c * n;
  for (int i = tid; i < n; i += 256) b[i] = (i == c) ? 1.0 : 0.0;
  __syncthreads();
  if (tid == 0)
    for (int k = 0; k < n; k++) {
      const int p = piv[k];
      if (p != k) {
        const double t = b[k];
        b[k]           = b[p];
        b[p]           = t;
      }
    }
  __syncthreads();
  for (int j = 0; j < n; j++) {
    const double y = b[j];
    for (int i = j + 1 + tid; i < n; i += 256) b[i] -= LU[(int64_t)i * n + j] * y;
    __syncthreads();
  }
  for (int j = n - 1; j >= 0; j--) {
    __syncthreads();
    const double x = b[j] / LU[(int64_t)j * n + j];
    __syncthreads();
    if (tid == 0) b[j] = x;
    for (int i = tid; i < j; i += 256) b[i] -= LU[(int64_t)i * n + j] * x;
    __syncthreads();
  }
}

// out_i = sum_c Xc[c * n + i] b_c (= A^{-1} b), one lane per row, columns in order.
__global__ void k_matvec_inv(int n, const double *__restrict__ Xc, const double *__restrict__ b,
                             double *__restrict__ out)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double r = 0.0;
  for (int c = 0; c < n; c++) r += Xc[(int64_t)c * n + i] * b[c];
  out[i] = r;
}

// out_c = sum_i Xc[c * n + i] g_i (= A^{-T} g), one wavefront per c: lanes over i, then a
// fixed-order butterfly.
__global__ __launch_bounds__(64) void k_matvec_inv_t(int n, const double *__restrict__ Xc,
                                                     const double *__restrict__ g, double *__restrict__ out)
{
  const int c = blockIdx.x, lane = threadIdx.x;
  if (c >= n) return;
  double r = 0.0;
  for (int i = lane; i < n; i += 64) r += Xc[(int64_t)c * n + i] * g[i];
  for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
  if (lane == 0) out[c] = r;
}

// u_i = U + v_i + sum_j B[i][j] gamma_j: the velocity at control point i (main.cpp:229-239), one
// wavefront per i, lanes over j, fixed-order butterfly.
__global__ __launch_bounds__(64) void k_wing_velocity(int n, const double *__restrict__ B,
                                                      const double *__restrict__ gamma, double Ux, double Uy,
                                                      double Uz, const double *__restrict__ v, double *__restrict__ u)
{
  const int i = blockIdx.x, lane = threadIdx.x;
  if (i >= n) return;
  V3 a{0, 0, 0};
  for (int j = lane; j < n; j += 64) a = a + gamma[j] * ld3(B + 3 * ((int64_t)i * n + j));
  for (int o = 32; o > 0; o >>= 1) {
    a.x += __shfl_xor(a.x, o, 64);
    a.y += __shfl_xor(a.y, o, 64);
    a.z += __shfl_xor(a.z, o, 64);
  }
  if (lane == 0) {
    a = a + V3{Ux, Uy, Uz};
    if (v) a = a + ld3(v + 3 * i);
    u[3 * i] = a.x, u[3 * i + 1] = a.y, u[3 * i + 2] = a.z;
  }
}

// rhs_i = -U . n_i - v_i . n_i (main.cpp:207-212); v may be null (no particles).
__global__ void k_wing_rhs(int n, const double *__restrict__ nrm, double Ux, double Uy, double Uz,
                           const double *__restrict__ v, double *__restrict__ rhs)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const V3 nn = ld3(nrm + 3 * i);
  double   r  = -(Ux * nn.x + Uy * nn.y + Uz * nn.z);
  if (v) r -= dot(ld3(v + 3 * i), nn);
  rhs[i] = r;
}

// Kutta-Joukowski force (main.cpp:223-244), one workgroup: u_i = sum_j B[i][j] gamma_j + U + v_i
// at control point i, F = sum_i Gamma_i (u_i x (LE_c1 - LE_c)), Gamma_i = gamma_i minus the
// circulation of the panel upstream (gamma_i on the first row).  Per-lane partial sums in panel
// order, then a fixed-order tree over the lanes.
__global__ __launch_bounds__(256) void k_wing_force(int n, int nspan, const double *__restrict__ V,
                                                    const double *__restrict__ uc, const double *__restrict__ gamma,
                                                    double *__restrict__ F)
{
  __shared__ double red[3][256];
  const int         tid = threadIdx.x;
  V3                acc{0, 0, 0};
  for (int i = tid; i < n; i += 256) {
    const V3      u  = ld3(uc + 3 * i);
    const Corners q  = corners(V, nspan, i);
    const double  Gn = gamma[i] - (i >= nspan ? gamma[i - nspan] : 0.0);
    acc              = acc + Gn * cross(u, q.le1 - q.le);
  }
  red[0][tid] = acc.x, red[1][tid] = acc.y, red[2][tid] = acc.z;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o)
      for (int d = 0; d < 3; d++) red[d][tid] += red[d][tid + o];
    __syncthreads();
  }
  if (tid < 3) F[tid] = red[tid][0];
}

// Particles shed from the trailing-edge row (main.cpp:246-289), one lane per spanwise panel c:
// X = midpoint of TE_c1 -> TE_c + U dt (z offset U[1] dt as the reference writes it when
// ref_offset, else U[2] dt), alpha = (TE_c - TE_c1) g_c + (LE_c - TE_c) (g_c - g_{c-1}) / 2
// + (TE_c1 - LE_c1) (g_c - g_{c+1}) / 2 with the full g_c at the tips, sigma = 2.5 |TE_c - TE_c1|.
__global__ void k_wing_shed(int nspan, int nchord, const double *__restrict__ V, const double *__restrict__ gamma,
                            double Ux, double Uy, double Uz, double dt, int ref_offset, double *__restrict__ X,
                            double *__restrict__ Al, double *__restrict__ S)
{
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= nspan) return;
  const int     p  = (nchord - 1) * nspan + c;
  const Corners q  = corners(V, nspan, p);
  const double *g  = gamma + (nchord - 1) * nspan;
  const V3      d  = q.te - q.te1;
  const V3      mp = 0.5 * (q.te1 + q.te);
  X[3 * c]         = mp.x + Ux * dt;
  X[3 * c + 1]     = mp.y + Uy * dt;
  X[3 * c + 2]     = mp.z + (ref_offset ? Uy : Uz) * dt;
  V3 a             = g[c] * d;
  a                = a + (c > 0 ? 0.5 * (g[c] - g[c - 1]) : g[c]) * (q.le - q.te);
  a                = a + (c < nspan - 1 ? 0.5 * (g[c] - g[c + 1]) : g[c]) * (q.te1 - q.le1);
  Al[3 * c] = a.x, Al[3 * c + 1] = a.y, Al[3 * c + 2] = a.z;
  S[c] = 2.5 * sqrt(dot(d, d));
}

__global__ void k_translate(int64_t n, double *__restrict__ p, double dx, double dy, double dz)
{
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= n) return;
  p[3 * i] += dx;
  p[3 * i + 1] += dy;
  p[3 * i + 2] += dz;
}

// Probe records from points: X from pts (AoS [m][3]), every other field zero.
__global__ void k_points_to_soa(int64_t m, const double *__restrict__ pts, double *__restrict__ st, int64_t cap)
{
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= m) return;
  for (int f = 0; f < N_FIELDS; f++) st[f * cap + i] = 0.0;
  for (int d = 0; d < 3; d++) st[(F_X + d) * cap + i] = pts[3 * i + d];
}

// out[perm[i]] = velocity of probe i (tree order -> the order the points were given in).
__global__ void k_probe_velocity_out(int64_t m, const int32_t *__restrict__ perm, const double *__restrict__ st,
                                     int64_t cap, double *__restrict__ out)
{
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= m) return;
  const int64_t o = perm[i];
  for (int d = 0; d < 3; d++) out[3 * o + d] = st[(F_U + d) * cap + i];
}

// New bodies N .. N+k-1: X, alpha, sigma from AoS arrays, every other field zero, original
// index = position.
__global__ void k_append_bodies(int64_t N, int64_t k, const double *__restrict__ X, const double *__restrict__ A,
                                const double *__restrict__ S, double *__restrict__ st, int64_t cap,
                                int32_t *__restrict__ orig)
{
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= k) return;
  const int64_t b = N + i;
  for (int f = 0; f < N_FIELDS; f++) st[f * cap + b] = 0.0;
  for (int d = 0; d < 3; d++) {
    st[(F_X + d) * cap + b] = X[3 * i + d];
    st[(F_A + d) * cap + b] = A[3 * i + d];
  }
  st[F_SIG * cap + b] = S[i];
  orig[b]             = (int32_t)b;
}

// ---- adjoint of the lift coefficient through the loop ----
// Force VJP, part 1 (per panel i), for J = e . F with e already scaled: u_i as in k_wing_force,
// q_i = dl_i x e, c_i = u_i . q_i and w_i = Gamma_i q_i (= dJ/dv_i).
__global__ void k_wing_force_vjp1(int n, int nspan, const double *__restrict__ V, const double *__restrict__ uc,
                                  const double *__restrict__ gamma, double ex, double ey, double ez,
                                  double *__restrict__ c, double *__restrict__ w)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const V3      u  = ld3(uc + 3 * i);
  const Corners q  = corners(V, nspan, i);
  const V3      qe = cross(q.le1 - q.le, V3{ex, ey, ez});
  const double  Gn = gamma[i] - (i >= nspan ? gamma[i - nspan] : 0.0);
  c[i]             = dot(u, qe);
  w[3 * i] = Gn * qe.x, w[3 * i + 1] = Gn * qe.y, w[3 * i + 2] = Gn * qe.z;
}

// Part 2 (per panel j): dJ/dgamma_j = c_j - c_{j+nspan} (the net circulations) + sum_i w_i . B[i][j].
__global__ __launch_bounds__(64) void k_wing_force_vjp2(int n, int nspan, const double *__restrict__ B,
                                                        const double *__restrict__ c, const double *__restrict__ w,
                                                        double *__restrict__ gbar)
{
  const int j = blockIdx.x, lane = threadIdx.x;
  if (j >= n) return;
  double g = 0.0;
  for (int i = lane; i < n; i += 64) g += dot(ld3(w + 3 * i), ld3(B + 3 * ((int64_t)i * n + j)));
  for (int o = 32; o > 0; o >>= 1) g += __shfl_xor(g, o, 64);
  if (lane == 0) gbar[j] = g + c[j] - (j + nspan < n ? c[j + nspan] : 0.0);
}

// vbar_i += -rbar_i n_i (rhs_i = -U . n_i - v_i . n_i).
__global__ void k_wing_rhs_vjp(int n, const double *__restrict__ nrm, const double *__restrict__ rbar,
                               double *__restrict__ vbar)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < 3; d++) vbar[3 * i + d] -= rbar[i] * nrm[3 * i + d];
}

// gbar (all n panels) from the cotangent abar of the shed strengths (k_wing_shed): zero off the
// trailing-edge row; abar_c(d) = abar[d * stride + c].
__global__ void k_wing_shed_vjp(int nspan, int nchord, const double *__restrict__ V, const double *__restrict__ abar,
                                int64_t stride, double *__restrict__ gbar)
{
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= nspan * nchord) return;
  const int r0 = (nchord - 1) * nspan;
  if (c < r0) {
    gbar[c] = 0.0;
    return;
  }
  const int ci = c - r0;
  auto      ab = [&](int k) { return V3{abar[k], abar[stride + k], abar[2 * stride + k]}; };
  auto      cr = [&](int k) { return corners(V, nspan, r0 + k); };
  const Corners q = cr(ci);
  const V3      a = ab(ci);
  double        g = dot(a, q.te - q.te1);
  g += (ci > 0 ? 0.5 : 1.0) * dot(a, q.le - q.te);
  g += (ci < nspan - 1 ? 0.5 : 1.0) * dot(a, q.te1 - q.le1);
  if (ci + 1 < nspan) {
    const Corners qn = cr(ci + 1);
    g -= 0.5 * dot(ab(ci + 1), qn.le - qn.te);
  }
  if (ci > 0) {
    const Corners qp = cr(ci - 1);
    g -= 0.5 * dot(ab(ci - 1), qp.te1 - qp.le1);
  }
  gbar[c] = g;
}

// Probe "vortices" for the probe adjoint: X = pts, alpha = weights, sigma = 1, the rest zero.
__global__ void k_probe_sources(int64_t m, const double *__restrict__ pts, const double *__restrict__ wts,
                                double *__restrict__ st, int64_t cap)
{
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= m) return;
  for (int f = 0; f < N_FIELDS; f++) st[f * cap + i] = 0.0;
  for (int d = 0; d < 3; d++) {
    st[(F_X + d) * cap + i] = pts[3 * i + d];
    st[(F_A + d) * cap + i] = wts[3 * i + d];
  }
  st[F_SIG * cap + i] = 1.0;
}

// cot[d][orig[p]] += ast[F_DADT + d][p], cot[3 + d][orig[p]] += ast[F_U + d][p] (probe adjoint
// results, device order -> original order).
__global__ void k_cot_add_probe(int64_t N, const int32_t *__restrict__ orig, const double *__restrict__ ast,
                                int64_t cap, double *__restrict__ cot, int64_t ccap)
{
  const int64_t p = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (p >= N) return;
  const int64_t o = orig[p];
  for (int d = 0; d < 3; d++) {
    cot[d * ccap + o] += ast[(F_DADT + d) * cap + p];
    cot[(3 + d) * ccap + o] += ast[(F_U + d) * cap + p];
  }
}

// ---- design gradient: cotangents of the wing's vertices ----
// Per-panel corner cotangents are written as [n][12] (le, le1, te, te1, 3 doubles each) and
// summed into the vertices by k_wing_vertex_gather, so every sum runs in a fixed order.

// Vector-Jacobian product of the unit segment velocity w = segment(a, b, x) (main.cpp:11-25):
// adds (dw/da)^T wb, (dw/db)^T wb, (dw/dx)^T wb.  w = k c with c = r1 x r2, r1 = x - a,
// r2 = x - b, k = r0 . (r1/|r1| - r2/|r2|) / (4 pi |c|^2), r0 = b - a.  Zero in the reference's
// singular case, where the forward returns zero.
__device__ inline void segment_vjp(V3 a, V3 b, V3 x, V3 wb, V3 &ab, V3 &bb, V3 &xb)
{
  const V3     r1 = x - a, r2 = x - b, r0 = b - a;
  const V3     c  = cross(r1, r2);
  const double n1 = sqrt(dot(r1, r1)), n2 = sqrt(dot(r2, r2)), c2 = dot(c, c);
  if (n1 < 1e-6 || n2 < 1e-6 || sqrt(c2) < 1e-6) return;
  const V3     u1 = (1.0 / n1) * r1, u2 = (1.0 / n2) * r2;
  const V3     d  = u1 - u2;
  const double k  = dot(r0, d) / c2 * INV_4PI;
  const double kb = dot(wb, c);
  const double Nb = kb * INV_4PI / c2;    // cotangent of r0 . d
  const double cb2 = -kb * k / c2;        // cotangent of |c|^2
  const V3     cb  = k * wb + (2.0 * cb2) * c;
  const V3     r0b = Nb * d, db = Nb * r0;
  V3           r1b = (1.0 / n1) * (db - dot(db, u1) * u1);
  V3           r2b = (-1.0 / n2) * (db - dot(db, u2) * u2);
  r1b              = r1b + cross(r2, cb);
  r2b              = r2b + cross(cb, r1);
  xb               = xb + r1b + r2b;
  ab               = ab - r1b - r0b;
  bb               = bb - r2b + r0b;
}

// Cotangents of the corners of ring j at control point x for a cotangent wb of the ring's
// velocity (its four segments, in the order of ring()); xb collects the point's cotangent.
__device__ inline void ring_vjp(const Corners &q, V3 x, V3 wb, V3 cb[4], V3 &xb)
{
  segment_vjp(q.te1, q.te, x, wb, cb[3], cb[2], xb);
  segment_vjp(q.te, q.le, x, wb, cb[2], cb[0], xb);
  segment_vjp(q.le, q.le1, x, wb, cb[0], cb[1], xb);
  segment_vjp(q.le1, q.te1, x, wb, cb[1], cb[3], xb);
}

// Cotangent of ring j's velocity at control point i from A[i][j] = B[i][j] . n_i (Abar) and from
// the force's u_i = sum_j B[i][j] gamma_j (ub_i gamma_j); either part may be absent.
__device__ inline V3 influence_cot(int n, int i, int j, const double *Abar, const double *nrm, const double *ub,
                                   const double *gamma)
{
  V3 wb{0, 0, 0};
  if (Abar) wb = Abar[(int64_t)i * n + j] * ld3(nrm + 3 * i);
  if (ub) wb = wb + gamma[j] * ld3(ub + 3 * i);
  return wb;
}

// One workgroup per ring j, threads over the control points: corner cotangents of ring j.
__global__ __launch_bounds__(256) void k_wing_infl_vjp_ring(int n, int nspan, const double *__restrict__ V,
                                                            const double *__restrict__ cp,
                                                            const double *__restrict__ nrm,
                                                            const double *__restrict__ Abar,
                                                            const double *__restrict__ ub,
                                                            const double *__restrict__ gamma, double *__restrict__ pb)
{
  __shared__ double red[12][256];
  const int         j = blockIdx.x, tid = threadIdx.x;
  const Corners     q = corners(V, nspan, j);
  V3                cb[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = tid; i < n; i += 256) {
    const V3 wb = influence_cot(n, i, j, Abar, nrm, ub, gamma);
    V3       xb{0, 0, 0};
    ring_vjp(q, ld3(cp + 3 * i), wb, cb, xb);
  }
  for (int k = 0; k < 4; k++) {
    red[3 * k][tid]     = cb[k].x;
    red[3 * k + 1][tid] = cb[k].y;
    red[3 * k + 2][tid] = cb[k].z;
  }
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o)
      for (int k = 0; k < 12; k++) red[k][tid] += red[k][tid + o];
    __syncthreads();
  }
  if (tid < 12) pb[12 * j + tid] = red[tid][0];
}

// One wavefront per control point i, lanes over the rings: the point's cotangent and that of its
// normal (A[i][j] = B[i][j] . n_i).
__global__ __launch_bounds__(64) void k_wing_infl_vjp_point(int n, int nspan, const double *__restrict__ V,
                                                            const double *__restrict__ cp,
                                                            const double *__restrict__ nrm,
                                                            const double *__restrict__ B,
                                                            const double *__restrict__ Abar,
                                                            const double *__restrict__ ub,
                                                            const double *__restrict__ gamma,
                                                            double *__restrict__ cpb, double *__restrict__ nb)
{
  const int i = blockIdx.x, lane = threadIdx.x;
  if (i >= n) return;
  const V3 x = ld3(cp + 3 * i);
  V3       xb{0, 0, 0}, nnb{0, 0, 0};
  for (int j = lane; j < n; j += 64) {
    const V3 wb = influence_cot(n, i, j, Abar, nrm, ub, gamma);
    V3       cb[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    ring_vjp(corners(V, nspan, j), x, wb, cb, xb);
    if (Abar) nnb = nnb + Abar[(int64_t)i * n + j] * ld3(B + 3 * ((int64_t)i * n + j));
  }
  for (int o = 32; o > 0; o >>= 1) {
    xb.x += __shfl_xor(xb.x, o, 64);
    xb.y += __shfl_xor(xb.y, o, 64);
    xb.z += __shfl_xor(xb.z, o, 64);
    nnb.x += __shfl_xor(nnb.x, o, 64);
    nnb.y += __shfl_xor(nnb.y, o, 64);
    nnb.z += __shfl_xor(nnb.z, o, 64);
  }
  if (lane == 0) {
    cpb[3 * i] = xb.x, cpb[3 * i + 1] = xb.y, cpb[3 * i + 2] = xb.z;
    nb[3 * i] = nnb.x, nb[3 * i + 1] = nnb.y, nb[3 * i + 2] = nnb.z;
  }
}

// Corner cotangents of panel p from the cotangents of its control point (cpb1 + cpb2), unit
// normal (nb1 + nb2), area (areab, the same for every panel) and leading-edge segment
// LE_c1 - LE_c (dlb, may be null): the VJP of k_wing_geom and of the force's dl.
__global__ void k_wing_geom_vjp(int n, int nspan, const double *__restrict__ V, const double *__restrict__ cpb1,
                                const double *__restrict__ cpb2, const double *__restrict__ nb1,
                                const double *__restrict__ nb2, double areab, const double *__restrict__ dlb,
                                double *__restrict__ pb)
{
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const Corners q  = corners(V, nspan, p);
  const V3      v1 = q.le1 - q.te, v2 = q.le - q.te1;
  const V3      v  = cross(v1, v2);
  const double  l  = sqrt(dot(v, v));
  const V3      nn = (1.0 / l) * v;
  const V3      mb = 0.25 * (ld3(cpb1 + 3 * p) + ld3(cpb2 + 3 * p));
  const V3      ng = ld3(nb1 + 3 * p) + ld3(nb2 + 3 * p);
  const V3      vb = (1.0 / l) * (ng - dot(ng, nn) * nn) + (0.5 * areab) * nn;
  const V3      v1b = cross(v2, vb), v2b = cross(vb, v1);
  V3            le = mb + v2b, le1 = mb + v1b, te = mb - v1b, te1 = mb - v2b;
  if (dlb) {
    const V3 g = ld3(dlb + 3 * p);
    le1        = le1 + g;
    le         = le - g;
  }
  double *o = pb + 12 * p;
  o[0] = le.x, o[1] = le.y, o[2] = le.z, o[3] = le1.x, o[4] = le1.y, o[5] = le1.z;
  o[6] = te.x, o[7] = te.y, o[8] = te.z, o[9] = te1.x, o[10] = te1.y, o[11] = te1.z;
}

// Corner cotangents (added into pb) of the trailing-edge panels from the cotangents of the
// particles they shed (k_wing_shed): position xb, strength ab (AoS-by-field with `stride`,
// first shed particle at index `base`) and core size sb (may be null).
__global__ void k_wing_shed_geom_vjp(int nspan, int nchord, const double *__restrict__ V,
                                     const double *__restrict__ gamma, const double *__restrict__ lam, int64_t stride,
                                     int64_t base, const double *__restrict__ sb, double *__restrict__ pb)
{
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= nspan) return;
  const int     p  = (nchord - 1) * nspan + c;
  const Corners q  = corners(V, nspan, p);
  const double *g  = gamma + (nchord - 1) * nspan;
  const int64_t k  = base + c;
  const V3      xb{lam[k], lam[stride + k], lam[2 * stride + k]};
  const V3      abv{lam[3 * stride + k], lam[4 * stride + k], lam[5 * stride + k]};
  const double  wl = c > 0 ? 0.5 * (g[c] - g[c - 1]) : g[c];
  const double  wr = c < nspan - 1 ? 0.5 * (g[c] - g[c + 1]) : g[c];
  V3            te = 0.5 * xb + g[c] * abv - wl * abv;
  V3            te1 = 0.5 * xb - g[c] * abv + wr * abv;
  const V3      le = wl * abv, le1 = -wr * abv;
  if (sb) {
    const V3 d  = q.te - q.te1;
    const V3 dd = (2.5 * sb[k] / sqrt(dot(d, d))) * d;
    te          = te + dd;
    te1         = te1 - dd;
  }
  double *o = pb + 12 * p;
  o[0] += le.x, o[1] += le.y, o[2] += le.z, o[3] += le1.x, o[4] += le1.y, o[5] += le1.z;
  o[6] += te.x, o[7] += te.y, o[8] += te.z, o[9] += te1.x, o[10] += te1.y, o[11] += te1.z;
}

// Vertex cotangents: vertex (r, k) is corner le of panel (r, k), le1 of (r, k-1), te of (r-1, k)
// and te1 of (r-1, k-1), where they exist; nb buffers of corner cotangents added in order.
__global__ void k_wing_vertex_gather(int nspan, int nchord, const double *__restrict__ pb0,
                                     const double *__restrict__ pb1, const double *__restrict__ pb2,
                                     double *__restrict__ vb)
{
  const int w = nspan + 1, v = blockIdx.x * 256 + threadIdx.x;
  if (v >= w * (nchord + 1)) return;
  const int r = v / w, k = v % w;
  double    acc[3] = {0, 0, 0};
  auto      add    = [&](int pr, int pc, int corner) {
    if (pr < 0 || pr >= nchord || pc < 0 || pc >= nspan) return;
    const int64_t o = 12 * (int64_t)(pr * nspan + pc) + 3 * corner;
    for (const double *pb : {pb0, pb1, pb2})
      if (pb)
        for (int d = 0; d < 3; d++) acc[d] += pb[o + d];
  };
  add(r, k, 0);
  add(r, k - 1, 1);
  add(r - 1, k, 2);
  add(r - 1, k - 1, 3);
  for (int d = 0; d < 3; d++) vb[3 * v + d] = acc[d];
}

// Abar[i][j] -= rbar_i gamma_j (A gamma = rhs), nb_i -= rbar_i (U + v_i) (rhs_i = -(U + v_i) . n_i).
__global__ void k_wing_solve_geom_acc(int n, const double *__restrict__ rbar, const double *__restrict__ gamma,
                                      double Ux, double Uy, double Uz, const double *__restrict__ v,
                                      double *__restrict__ Abar, double *__restrict__ nb)
{
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int i = (int)(e / n), j = (int)(e % n);
  Abar[e] -= rbar[i] * gamma[j];
  if (j == 0) {
    V3 uv{Ux, Uy, Uz};
    if (v) uv = uv + ld3(v + 3 * i);
    nb[3 * i] -= rbar[i] * uv.x, nb[3 * i + 1] -= rbar[i] * uv.y, nb[3 * i + 2] -= rbar[i] * uv.z;
  }
}

// cpb_b += sum_a vb_a J_i[a][b]: the control point's cotangent through v(cp), J = dv/dx there.
__global__ void k_wing_cp_acc(int n, const double *__restrict__ vb, const double *__restrict__ J,
                              double *__restrict__ cpb)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  for (int b = 0; b < 3; b++)
    cpb[3 * i + b] += vb[3 * i] * J[9 * i + b] + vb[3 * i + 1] * J[9 * i + 3 + b] + vb[3 * i + 2] * J[9 * i + 6 + b];
}

// dlb_i = Gamma_i (e x u_i): the cotangent of panel i's leading-edge segment in J = e . F.
__global__ void k_wing_force_dl_vjp(int n, int nspan, const double *__restrict__ uc, const double *__restrict__ gamma,
                                    double ex, double ey, double ez, double *__restrict__ dlb)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double Gn = gamma[i] - (i >= nspan ? gamma[i - nspan] : 0.0);
  const V3     g  = Gn * cross(V3{ex, ey, ez}, ld3(uc + 3 * i));
  dlb[3 * i] = g.x, dlb[3 * i + 1] = g.y, dlb[3 * i + 2] = g.z;
}

// Velocity and its gradient at probe i, tree order -> the given order (out AoS [m][3], Jout [m][9]).
__global__ void k_probe_velocity_grad_out(int64_t m, const int32_t *__restrict__ perm, const double *__restrict__ st,
                                          int64_t cap, double *__restrict__ out, double *__restrict__ Jout)
{
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i >= m) return;
  const int64_t o = perm[i];
  for (int d = 0; d < 3; d++) out[3 * o + d] = st[(F_U + d) * cap + i];
  for (int k = 0; k < 9; k++) Jout[9 * o + k] = st[(F_J + k) * cap + i];
}

inline int nb256(int64_t n) { return (int)std::max<int64_t>(1, (n + 255) / 256); }

}  // namespace

void wing_geom(hipStream_t s, int n, int nspan, const double *V, double *cp, double *nrm, double *area)
{
  k_wing_geom<<<nb256(n), 256, 0, s>>>(n, nspan, V, cp, nrm, area);
  UVLM_LAUNCH_CHECK();
}

void wing_influence(hipStream_t s, int m, int n, int nspan, const double *V, const double *pts, const double *nrm,
                    double *B, double *A, int *singular)
{
  k_wing_influence<<<nb256((int64_t)m * n), 256, 0, s>>>(m, n, nspan, V, pts, nrm, B, A, singular);
  UVLM_LAUNCH_CHECK();
}

void lu_factor(hipStream_t s, int n, double *A, int *piv, int *info)
{
  k_lu<<<1, LU_WG, 0, s>>>(n, A, piv, info);
  UVLM_LAUNCH_CHECK();
}

void lu_inverse(hipStream_t s, int n, const double *LU, const int *piv, double *Xc)
{
  k_lu_inverse<<<n, 256, 0, s>>>(n, LU, piv, Xc);
  UVLM_LAUNCH_CHECK();
}

void matvec_inv(hipStream_t s, int n, const double *Xc, const double *b, double *out)
{
  k_matvec_inv<<<nb256(n), 256, 0, s>>>(n, Xc, b, out);
  UVLM_LAUNCH_CHECK();
}

void matvec_inv_t(hipStream_t s, int n, const double *Xc, const double *g, double *out)
{
  k_matvec_inv_t<<<n, 64, 0, s>>>(n, Xc, g, out);
  UVLM_LAUNCH_CHECK();
}

void wing_velocity(hipStream_t s, int n, const double *B, const double *gamma, const double U[3], const double *v,
                   double *u)
{
  k_wing_velocity<<<n, 64, 0, s>>>(n, B, gamma, U[0], U[1], U[2], v, u);
  UVLM_LAUNCH_CHECK();
}

void wing_rhs(hipStream_t s, int n, const double *nrm, const double U[3], const double *v, double *rhs)
{
  k_wing_rhs<<<nb256(n), 256, 0, s>>>(n, nrm, U[0], U[1], U[2], v, rhs);
  UVLM_LAUNCH_CHECK();
}

void wing_force(hipStream_t s, int n, int nspan, const double *V, const double *u, const double *gamma, double *F)
{
  k_wing_force<<<1, 256, 0, s>>>(n, nspan, V, u, gamma, F);
  UVLM_LAUNCH_CHECK();
}

void wing_shed(hipStream_t s, int nspan, int nchord, const double *V, const double *gamma, const double U[3], double dt,
               int ref_offset, double *X, double *A, double *S)
{
  k_wing_shed<<<nb256(nspan), 256, 0, s>>>(nspan, nchord, V, gamma, U[0], U[1], U[2], dt, ref_offset, X, A, S);
  UVLM_LAUNCH_CHECK();
}

void points_to_soa(hipStream_t s, int64_t m, const double *pts, double *st, int64_t cap)
{
  k_points_to_soa<<<nb256(m), 256, 0, s>>>(m, pts, st, cap);
  UVLM_LAUNCH_CHECK();
}

void probe_velocity_out(hipStream_t s, int64_t m, const int32_t *perm, const double *st, int64_t cap, double *out)
{
  k_probe_velocity_out<<<nb256(m), 256, 0, s>>>(m, perm, st, cap, out);
  UVLM_LAUNCH_CHECK();
}

void append_bodies(hipStream_t s, int64_t N, int64_t k, const double *X, const double *A, const double *S, double *st,
                   int64_t cap, int32_t *orig)
{
  if (k <= 0) return;
  k_append_bodies<<<nb256(k), 256, 0, s>>>(N, k, X, A, S, st, cap, orig);
  UVLM_LAUNCH_CHECK();
}

void wing_force_vjp(hipStream_t s, int n, int nspan, const double *V, const double *B, const double *u,
                    const double *gamma, const double e[3], double *c, double *vbar, double *gbar)
{
  k_wing_force_vjp1<<<nb256(n), 256, 0, s>>>(n, nspan, V, u, gamma, e[0], e[1], e[2], c, vbar);
  UVLM_LAUNCH_CHECK();
  k_wing_force_vjp2<<<n, 64, 0, s>>>(n, nspan, B, c, vbar, gbar);
  UVLM_LAUNCH_CHECK();
}

void wing_rhs_vjp(hipStream_t s, int n, const double *nrm, const double *rbar, double *vbar)
{
  k_wing_rhs_vjp<<<nb256(n), 256, 0, s>>>(n, nrm, rbar, vbar);
  UVLM_LAUNCH_CHECK();
}

void wing_shed_vjp(hipStream_t s, int nspan, int nchord, const double *V, const double *abar, int64_t stride,
                   double *gbar)
{
  k_wing_shed_vjp<<<nb256((int64_t)nspan * nchord), 256, 0, s>>>(nspan, nchord, V, abar, stride, gbar);
  UVLM_LAUNCH_CHECK();
}

void probe_sources(hipStream_t s, int64_t m, const double *pts, const double *wts, double *st, int64_t cap)
{
  k_probe_sources<<<nb256(m), 256, 0, s>>>(m, pts, wts, st, cap);
  UVLM_LAUNCH_CHECK();
}

void cot_add_probe(hipStream_t s, int64_t N, const int32_t *orig, const double *ast, int64_t cap, double *cot,
                   int64_t ccap)
{
  if (N <= 0) return;
  k_cot_add_probe<<<nb256(N), 256, 0, s>>>(N, orig, ast, cap, cot, ccap);
  UVLM_LAUNCH_CHECK();
}

void wing_influence_vjp(hipStream_t s, int n, int nspan, const double *V, const double *cp, const double *nrm,
                        const double *B, const double *Abar, const double *ub, const double *gamma, double *pb_ring,
                        double *cpb, double *nb)
{
  k_wing_infl_vjp_ring<<<n, 256, 0, s>>>(n, nspan, V, cp, nrm, Abar, ub, gamma, pb_ring);
  UVLM_LAUNCH_CHECK();
  k_wing_infl_vjp_point<<<n, 64, 0, s>>>(n, nspan, V, cp, nrm, B, Abar, ub, gamma, cpb, nb);
  UVLM_LAUNCH_CHECK();
}

void wing_geom_vjp(hipStream_t s, int n, int nspan, const double *V, const double *cpb1, const double *cpb2,
                   const double *nb1, const double *nb2, double areab, const double *dlb, double *pb)
{
  k_wing_geom_vjp<<<nb256(n), 256, 0, s>>>(n, nspan, V, cpb1, cpb2, nb1, nb2, areab, dlb, pb);
  UVLM_LAUNCH_CHECK();
}

void wing_shed_geom_vjp(hipStream_t s, int nspan, int nchord, const double *V, const double *gamma, const double *lam,
                        int64_t stride, int64_t base, const double *sb, double *pb)
{
  k_wing_shed_geom_vjp<<<nb256(nspan), 256, 0, s>>>(nspan, nchord, V, gamma, lam, stride, base, sb, pb);
  UVLM_LAUNCH_CHECK();
}

void wing_vertex_gather(hipStream_t s, int nspan, int nchord, const double *pb0, const double *pb1, const double *pb2,
                        double *vb)
{
  k_wing_vertex_gather<<<nb256((int64_t)(nspan + 1) * (nchord + 1)), 256, 0, s>>>(nspan, nchord, pb0, pb1, pb2, vb);
  UVLM_LAUNCH_CHECK();
}

void wing_solve_geom_acc(hipStream_t s, int n, const double *rbar, const double *gamma, const double U[3],
                         const double *v, double *Abar, double *nb)
{
  k_wing_solve_geom_acc<<<nb256((int64_t)n * n), 256, 0, s>>>(n, rbar, gamma, U[0], U[1], U[2], v, Abar, nb);
  UVLM_LAUNCH_CHECK();
}

void wing_cp_acc(hipStream_t s, int n, const double *vb, const double *J, double *cpb)
{
  k_wing_cp_acc<<<nb256(n), 256, 0, s>>>(n, vb, J, cpb);
  UVLM_LAUNCH_CHECK();
}

void wing_force_dl_vjp(hipStream_t s, int n, int nspan, const double *u, const double *gamma, const double e[3],
                       double *dlb)
{
  k_wing_force_dl_vjp<<<nb256(n), 256, 0, s>>>(n, nspan, u, gamma, e[0], e[1], e[2], dlb);
  UVLM_LAUNCH_CHECK();
}

void probe_velocity_grad_out(hipStream_t s, int64_t m, const int32_t *perm, const double *st, int64_t cap, double *out,
                             double *Jout)
{
  k_probe_velocity_grad_out<<<nb256(m), 256, 0, s>>>(m, perm, st, cap, out, Jout);
  UVLM_LAUNCH_CHECK();
}

void translate3(hipStream_t s, int64_t n, double *p, const double d[3])
{
  if (n <= 0) return;
  k_translate<<<nb256(n), 256, 0, s>>>(n, p, d[0], d[1], d[2]);
  UVLM_LAUNCH_CHECK();
}

}  // namespace uvlm
