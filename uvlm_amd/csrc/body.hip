// body.hip -- per-body kernels: exafmm::Body AoS <-> field-major SoA, RK stage updates.
//
// RK stages follow src/fmmCalculator.hpp:493-711 term by term (state y = (X, alpha, sigma),
// registers (velocity_old, dadt_old, drdt_old), initial state (X_old, alpha_old,
// radius_old)), including the RK4 third stage that advances with 0.5*dt (683-689).
#include "common.hpp"

#include <algorithm>
#include <vector>

namespace uvlm {

namespace {

constexpr int BLK = 256;
inline int    nblocks(int64_t n) { return (int)std::max<int64_t>(1, (n + BLK - 1) / BLK); }

__global__ void k_aos_to_soa(const double *__restrict__ aos, int64_t n, int64_t stride, double *__restrict__ st,
                             int64_t cap)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i >= n) return;
  for (int f = 0; f < N_FIELDS; f++) st[f * cap + i] = aos[i * stride + f];
}

__global__ void k_soa_to_aos(const double *__restrict__ st, int64_t cap, double *__restrict__ aos, int64_t n,
                             int64_t stride)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i >= n) return;
  for (int f = 0; f < N_FIELDS; f++) aos[i * stride + f] = st[f * cap + i];
}

__global__ void k_rk(double *__restrict__ st, int64_t cap, int64_t N, int scheme, int stage, double dt)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i >= N) return;
  auto F = [&](int f) -> double & { return st[f * cap + i]; };
  if (scheme == UVLM_EULER) {  // stepEuler (511-521)
    for (int d = 0; d < 3; d++) F(F_X + d) += F(F_U + d) * dt;
    for (int d = 0; d < 3; d++) F(F_A + d) += F(F_DADT + d) * dt;
    F(F_SIG) += F(F_DSIG) * dt;
    return;
  }
  if ((scheme == UVLM_RK2 || scheme == UVLM_RK4) && stage == 0) {  // storeInitialState (493-509)
    for (int d = 0; d < 3; d++) {
      F(F_UOLD + d)  = 0.0;
      F(F_DAOLD + d) = 0.0;
      F(F_XOLD + d)  = F(F_X + d);
      F(F_AOLD + d)  = F(F_A + d);
    }
    F(F_SIGOLD) = F(F_SIG);
  }
  if (scheme == UVLM_RK2) {
    if (stage == 0) {  // stepRK21 (523-541)
      for (int d = 0; d < 3; d++) {
        F(F_UOLD + d)  = F(F_U + d);
        F(F_DAOLD + d) = F(F_DADT + d);
      }
      F(F_DSIGOLD) = F(F_DSIG);
      for (int d = 0; d < 3; d++) F(F_X + d) += dt * F(F_UOLD + d);
      for (int d = 0; d < 3; d++) F(F_A + d) += dt * F(F_DAOLD + d);
      F(F_SIG) += dt * F(F_DSIGOLD);
    } else {  // stepRK22 (543-562)
      for (int d = 0; d < 3; d++) {
        F(F_UOLD + d)  = 0.5 * F(F_U + d) + 0.5 * F(F_UOLD + d);
        F(F_DAOLD + d) = 0.5 * F(F_DADT + d) + 0.5 * F(F_DAOLD + d);
      }
      F(F_DSIGOLD) = 0.5 * F(F_DSIG) + 0.5 * F(F_DSIGOLD);
      for (int d = 0; d < 3; d++) F(F_X + d) = F(F_XOLD + d) + dt * F(F_UOLD + d);
      for (int d = 0; d < 3; d++) F(F_A + d) = F(F_AOLD + d) + dt * F(F_DAOLD + d);
      F(F_SIG) = F(F_SIGOLD) + dt * F(F_DSIGOLD);
    }
    return;
  }
  if (scheme == UVLM_RK3) {  // Williamson low-storage, stepRK31/32/33 (564-627)
    const double A = stage == 0 ? 0.0 : stage == 1 ? -(5.0 / 9.0) : (-153.0 / 128.0);
    const double B = stage == 0 ? (1.0 / 3.0) : stage == 1 ? (15.0 / 16.0) : (8.0 / 15.0);
    if (stage == 0) {
      for (int d = 0; d < 3; d++) {
        F(F_UOLD + d)  = dt * F(F_U + d);
        F(F_DAOLD + d) = dt * F(F_DADT + d);
      }
      F(F_DSIGOLD) = dt * F(F_DSIG);
    } else {
      for (int d = 0; d < 3; d++) {
        F(F_UOLD + d)  = A * F(F_UOLD + d) + dt * F(F_U + d);
        F(F_DAOLD + d) = A * F(F_DAOLD + d) + dt * F(F_DADT + d);
      }
      F(F_DSIGOLD) = A * F(F_DSIGOLD) + dt * F(F_DSIG);
    }
    for (int d = 0; d < 3; d++) F(F_X + d) += B * F(F_UOLD + d);
    for (int d = 0; d < 3; d++) F(F_A + d) += B * F(F_DAOLD + d);
    F(F_SIG) += B * F(F_DSIGOLD);
    return;
  }
  // RK4 (629-711)
  if (stage == 0) {
    for (int d = 0; d < 3; d++) {
      F(F_UOLD + d)  = F(F_U + d);
      F(F_DAOLD + d) = F(F_DADT + d);
    }
    F(F_DSIGOLD) = F(F_DSIG);
    for (int d = 0; d < 3; d++) F(F_X + d) += 0.5 * dt * F(F_UOLD + d);
    for (int d = 0; d < 3; d++) F(F_A + d) += 0.5 * dt * F(F_DAOLD + d);
    F(F_SIG) += 0.5 * dt * F(F_DSIGOLD);
  } else if (stage == 1 || stage == 2) {
    for (int d = 0; d < 3; d++) {
      F(F_UOLD + d) += 2.0 * F(F_U + d);
      F(F_DAOLD + d) += 2.0 * F(F_DADT + d);
    }
    F(F_DSIGOLD) += 2.0 * F(F_DSIG);
    for (int d = 0; d < 3; d++) F(F_X + d) = F(F_XOLD + d) + 0.5 * dt * F(F_U + d);
    for (int d = 0; d < 3; d++) F(F_A + d) = F(F_AOLD + d) + 0.5 * dt * F(F_DADT + d);
    F(F_SIG) = F(F_SIGOLD) + 0.5 * dt * F(F_DSIG);
  } else {
    for (int d = 0; d < 3; d++) {
      F(F_UOLD + d) += F(F_U + d);
      F(F_DAOLD + d) += F(F_DADT + d);
    }
    F(F_DSIGOLD) += F(F_DSIG);
    for (int d = 0; d < 3; d++) F(F_X + d) = F(F_XOLD + d) + (dt / 6.0) * F(F_UOLD + d);
    for (int d = 0; d < 3; d++) F(F_A + d) = F(F_AOLD + d) + (dt / 6.0) * F(F_DAOLD + d);
    F(F_SIG) = F(F_SIGOLD) + (dt / 6.0) * F(F_DSIGOLD);
  }
}

// Pedrizzetti relaxation, src/relaxation.hpp:8-18 in 0-based form with J[a][b] = du_a/dx_b:
// w = curl u = (J21 - J12, J02 - J20, J10 - J01); alpha <- (1-f) alpha + f |alpha| w/|w|.
// Bodies with |w| == 0 are left unchanged.
__global__ void k_relax(double *__restrict__ st, int64_t cap, int64_t N, double f)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i >= N) return;
  auto         Jm = [&](int a, int b) { return st[(F_J + 3 * a + b) * cap + i]; };
  const double w[3] = {Jm(2, 1) - Jm(1, 2), Jm(0, 2) - Jm(2, 0), Jm(1, 0) - Jm(0, 1)};
  const double nw   = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  if (nw == 0.0) return;
  double a[3];
  for (int d = 0; d < 3; d++) a[d] = st[(F_A + d) * cap + i];
  const double na = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  for (int d = 0; d < 3; d++) st[(F_A + d) * cap + i] = (1.0 - f) * a[d] + f * na * w[d] / nw;
}

// Invariants of the particle set, per block partial sums in a fixed order:
// [0..2] linear impulse 1/2 sum x × alpha, [3..5] angular impulse 1/3 sum x × (x × alpha),
// [6] kinetic energy sum u . (x × alpha), [7..9] total strength sum alpha.
constexpr int NDIAG = 10;
__global__ void k_diag_partial(const double *__restrict__ st, int64_t cap, int64_t N, double *__restrict__ part)
{
  __shared__ double sh[NDIAG][BLK];
  double v[NDIAG];
  for (int k = 0; k < NDIAG; k++) v[k] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < N; i += (int64_t)gridDim.x * BLK) {
    double x[3], a[3], u[3];
    for (int d = 0; d < 3; d++) {
      x[d] = st[(F_X + d) * cap + i];
      a[d] = st[(F_A + d) * cap + i];
      u[d] = st[(F_U + d) * cap + i];
    }
    const double c[3]  = {x[1] * a[2] - x[2] * a[1], x[2] * a[0] - x[0] * a[2], x[0] * a[1] - x[1] * a[0]};
    const double cc[3] = {x[1] * c[2] - x[2] * c[1], x[2] * c[0] - x[0] * c[2], x[0] * c[1] - x[1] * c[0]};
    for (int d = 0; d < 3; d++) {
      v[d] += 0.5 * c[d];
      v[3 + d] += cc[d] / 3.0;
      v[7 + d] += a[d];
    }
    v[6] += u[0] * c[0] + u[1] * c[1] + u[2] * c[2];
  }
  for (int k = 0; k < NDIAG; k++) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < NDIAG; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < NDIAG; k++) part[blockIdx.x * NDIAG + k] = sh[k][0];
}

__global__ void k_reset_sensor(double *__restrict__ st, int64_t cap, int64_t N)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i >= N) return;
  for (int d = 0; d < 3; d++) {
    st[(F_U + d) * cap + i]    = 0.0;
    st[(F_DADT + d) * cap + i] = 0.0;
  }
}

// CSM core spreading (fmmCalculator.hpp:334-337): drdt += (sqrt(sigma^2 + 2 nu dt) - sigma) / dt.
__global__ void k_csm(double *__restrict__ st, int64_t cap, int64_t N, double nu, double dt)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i >= N) return;
  const double sg = st[F_SIG * cap + i];
  st[F_DSIG * cap + i] += (sqrt(sg * sg + 2.0 * nu * dt) - sg) / dt;
}

__global__ void k_zero_fields(double *__restrict__ st, int64_t cap, int64_t N, int f0, int nf)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i >= N) return;
  for (int f = 0; f < nf; f++) st[(f0 + f) * cap + i] = 0.0;
}

// Per-block partial sums of a_d * b_d (d = 0..2) over fields fa, fb; fixed reduction order.
__global__ void k_dot3_partial(const double *__restrict__ st, int64_t cap, int64_t N, int fa, int fb,
                               double *__restrict__ part)
{
  __shared__ double sh[3][BLK];
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < N; i += (int64_t)gridDim.x * BLK)
    for (int d = 0; d < 3; d++) v[d] += st[(fa + d) * cap + i] * st[(fb + d) * cap + i];
  for (int d = 0; d < 3; d++) sh[d][threadIdx.x] = v[d];
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int d = 0; d < 3; d++) sh[d][threadIdx.x] += sh[d][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int d = 0; d < 3; d++) part[blockIdx.x * 3 + d] = sh[d][0];
}

// RBF conjugate gradients (fmmCalculator.hpp:765-853), fields as in the reference:
// b = velocity_old (target), Ap = velocity, p = alpha, x = dadt_old, r = X_old.
// mode 0 initAlphaValues: alpha = velocity_old * 4/3 pi sigma^3
// mode 1 setInitialResidualAndUpdateAlpha: r = b - velocity; x = alpha; alpha = r; X_old = r
// mode 2 updateAlphasAndResiduals: x += c_d p_d; r -= c_d Ap_d
// mode 3 updateP: p = r + c_d p
// mode 4 storeAlphaValues: alpha = x
__global__ void k_rbf_vec(double *__restrict__ st, int64_t cap, int64_t N, int mode, double c0, double c1, double c2)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i >= N) return;
  const double c[3] = {c0, c1, c2};
  for (int d = 0; d < 3; d++) {
    double &al = st[(F_A + d) * cap + i];
    double &x  = st[(F_DAOLD + d) * cap + i];
    double &r  = st[(F_XOLD + d) * cap + i];
    const double b = st[(F_UOLD + d) * cap + i], Ap = st[(F_U + d) * cap + i];
    if (mode == 0) {
      const double sg = st[F_SIG * cap + i];
      al              = b * (4. / 3. * M_PI * sg * sg * sg);
    } else if (mode == 1) {
      const double rr = b - Ap;
      x               = al;
      al              = rr;
      r               = rr;
    } else if (mode == 2) {
      x += c[d] * al;
      r -= c[d] * Ap;
    } else if (mode == 3) {
      al = r + c[d] * al;
    } else {
      al = x;
    }
  }
}

}  // namespace

void csm(hipStream_t s, double *st, int64_t cap, int64_t N, double nu, double dt)
{
  if (N <= 0) return;
  k_csm<<<nblocks(N), BLK, 0, s>>>(st, cap, N, nu, dt);
  UVLM_LAUNCH_CHECK();
}

void zero_fields(hipStream_t s, double *st, int64_t cap, int64_t N, int f0, int nf)
{
  if (N <= 0) return;
  k_zero_fields<<<nblocks(N), BLK, 0, s>>>(st, cap, N, f0, nf);
  UVLM_LAUNCH_CHECK();
}

void dot3(hipStream_t s, const double *st, int64_t cap, int64_t N, int fa, int fb, double *part_dev, double out[3])
{
  const int nb = (int)std::min<int64_t>(512, std::max<int64_t>(1, (N + BLK - 1) / BLK));
  k_dot3_partial<<<nb, BLK, 0, s>>>(st, cap, N, fa, fb, part_dev);
  UVLM_LAUNCH_CHECK();
  std::vector<double> h((size_t)nb * 3);
  UVLM_HIP(hipMemcpyAsync(h.data(), part_dev, 8 * h.size(), hipMemcpyDeviceToHost, s));
  UVLM_HIP(hipStreamSynchronize(s));
  out[0] = out[1] = out[2] = 0.0;
  for (int b = 0; b < nb; b++)
    for (int d = 0; d < 3; d++) out[d] += h[(size_t)b * 3 + d];
}

void rbf_vec(hipStream_t s, double *st, int64_t cap, int64_t N, int mode, const double c[3])
{
  if (N <= 0) return;
  k_rbf_vec<<<nblocks(N), BLK, 0, s>>>(st, cap, N, mode, c ? c[0] : 0.0, c ? c[1] : 0.0, c ? c[2] : 0.0);
  UVLM_LAUNCH_CHECK();
}

void aos_to_soa(hipStream_t s, const void *aos, int64_t n, int64_t stride_doubles, double *st, int64_t cap)
{
  if (n <= 0) return;
  k_aos_to_soa<<<nblocks(n), BLK, 0, s>>>((const double *)aos, n, stride_doubles, st, cap);
  UVLM_LAUNCH_CHECK();
}

void soa_to_aos(hipStream_t s, const double *st, int64_t cap, void *aos, int64_t n, int64_t stride_doubles)
{
  if (n <= 0) return;
  k_soa_to_aos<<<nblocks(n), BLK, 0, s>>>(st, cap, (double *)aos, n, stride_doubles);
  UVLM_LAUNCH_CHECK();
}

void rk_stage(hipStream_t s, double *st, int64_t cap, int64_t N, int scheme, int stage, double dt, int)
{
  k_rk<<<nblocks(N), BLK, 0, s>>>(st, cap, N, scheme, stage, dt);
  UVLM_LAUNCH_CHECK();
}

void relax(hipStream_t s, double *st, int64_t cap, int64_t N, double f)
{
  k_relax<<<nblocks(N), BLK, 0, s>>>(st, cap, N, f);
  UVLM_LAUNCH_CHECK();
}

void diagnostics(hipStream_t s, const double *st, int64_t cap, int64_t N, double *part_dev, double out[10])
{
  const int nb = (int)std::min<int64_t>(512, std::max<int64_t>(1, (N + BLK - 1) / BLK));
  k_diag_partial<<<nb, BLK, 0, s>>>(st, cap, N, part_dev);
  UVLM_LAUNCH_CHECK();
  std::vector<double> h((size_t)nb * NDIAG);
  UVLM_HIP(hipMemcpyAsync(h.data(), part_dev, 8 * h.size(), hipMemcpyDeviceToHost, s));
  UVLM_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < NDIAG; k++) out[k] = 0.0;
  for (int b = 0; b < nb; b++)
    for (int k = 0; k < NDIAG; k++) out[k] += h[(size_t)b * NDIAG + k];
}

void reset_sensor_rates(hipStream_t s, double *st, int64_t cap, int64_t N)
{
  k_reset_sensor<<<nblocks(N), BLK, 0, s>>>(st, cap, N);
  UVLM_LAUNCH_CHECK();
}

}  // namespace uvlm
