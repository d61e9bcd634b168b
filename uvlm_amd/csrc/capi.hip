// capi.cpp -- the C ABI of libuvlm_fmm.so (include/uvlm_fmm.h): handles, device buffers,
// orchestration of the tree / list / FMM kernels on one HIP stream, per-phase timing.
//
// Call sequence of one evaluation (the work of FMMCalculator::calculate,
// src/fmmCalculator.hpp:420-467):
//   P2M (all leaves) -> M2M (levels deepest-1 .. 1) -> root multipole cleared (the
//   reference never forms it, 427-437) -> M2L (every cell, stores L) -> P2P (every leaf,
//   stores u and J) -> L2L (levels 1 .. deepest, adds the parent) -> L2P fused with
//   stretching and the rVPM terms (adds the far field, writes dadt and drdt).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "common.hpp"

using namespace uvlm;

namespace {
thread_local std::string g_last_error;

// roctx range with the label the reference gives the same work (Kokkos::Profiling regions of
// fmmCalculator.hpp), so rocprofv3 --marker-trace shows the reference's phases.
struct Region {
  explicit Region(const char *name) { roctxRangePushA(name); }
  ~Region() { roctxRangePop(); }
  Region(const Region &)            = delete;
  Region &operator=(const Region &) = delete;
};

enum Ev { EV_B0, EV_TREE, EV_LISTS, EV_E0, EV_UP, EV_M2L, EV_P2P, EV_DOWN, EV_COUNT };
}  // namespace

struct uvlm_handle_s {
  uvlm_params prm{};
  int         NT = 0;
  hipStream_t s  = nullptr;
  int64_t     N = 0, cap = 0;
  DevBuf      st, st2, orig, orig2, perm, staging;
  Scratch     sc;
  Tree        tree;
  Lists       lists;
  DevBuf      M, L;
  M2LSchedule sched;
  bool        built = false, evaluated = false;
  // uvlm_build permutes the fields the evaluation overwrites (J, velocity, dadt, drdt) only when
  // something other than uvlm_eval needs them: until then they stay in st2 in the previous order
  bool        out_deferred = false;
  hipEvent_t  ev[EV_COUNT]{};
  uvlm_timing tm{};
  // probe (sensor) evaluation and its adjoint
  Tree    stree;
  Lists   slists, alists;
  DevBuf  sst, sst2, Ls, Ms, sperm, ast;
  int64_t scap = 0;
  hipEvent_t aev[2]{};
  // near-field option (uvlm_set_near_field): near_c, and the gap of the last build = near_c * max sigma
  double near_c = 0.0, gap = 0.0;
  // viscous scheme applied after each eval of uvlm_advance (UVLM_INVISCID / UVLM_PSE / UVLM_CSM)
  int    viscous = UVLM_INVISCID;
  double nu      = 0.0;
  // adjoint of the rates and of the time step: transposed lists, adjoint fields, stage states
  // (X, alpha, sigma + original index, per stage), state cotangents in the original order
  Lists   tlists;
  bool    tlists_ok = false;  // tlists is the transpose of the current build's lists
  DevBuf  vaux, vstage, vorig, vcot, vlam, vsig;
  // trajectory checkpoints (X, alpha, sigma in the original order, every `ck_every` steps)
  DevBuf  ck;
  int64_t ck_n = 0, ck_steps = 0, ck_every = 0, ck_N = 0;
  double  ck_dt = 0.0;

  double *F(int f) { return st.as<double>() + (int64_t)f * cap; }
  void    ensure_cap(int64_t n)
  {
    if (n <= cap) return;
    cap = n;
    st.reserve(sizeof(double) * N_FIELDS * cap);
    st2.reserve(sizeof(double) * N_FIELDS * cap);
    orig.reserve(4 * cap);
    orig2.reserve(4 * cap);
    perm.reserve(4 * cap);
  }
};

// A vortex-lattice wing on one device (vlm.hip): geometry, ring influences at the control
// points, the factored system matrix, and the last solve.
struct uvlm_wing_s {
  int         device = 0;
  hipStream_t s      = nullptr;
  int         nspan = 0, nchord = 0, n = 0;
  DevBuf      V, cp, nrm, area, B, LU, Ainv, piv, info, rhs, gamma, v, u, F, shedX, shedA, shedS;
};

namespace {

int fail(int code, const std::string &msg)
{
  g_last_error = msg;
  return code;
}

template <class Fn> int guarded(Fn &&fn)
{
  try {
    fn();
    return UVLM_OK;
  } catch (const Error &e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc &) {
    return fail(UVLM_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception &e) {
    return fail(UVLM_ERR_INTERNAL, e.what());
  }
}

void set_device(uvlm_handle h) { UVLM_HIP(hipSetDevice(h->prm.device)); }

float ev_ms2(uvlm_handle h, int a, int b)
{
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, h->aev[a], h->aev[b]) != hipSuccess) return 0.f;
  return ms;
}

float ev_ms(uvlm_handle h, int a, int b)
{
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, h->ev[a], h->ev[b]) != hipSuccess) return 0.f;
  return ms;
}

__global__ void k_iota(int32_t *p, int64_t n)
{
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i < n) p[i] = (int32_t)i;
}

void iota(hipStream_t s, int32_t *p, int64_t n)
{
  k_iota<<<(int)std::max<int64_t>(1, (n + 255) / 256), 256, 0, s>>>(p, n);
  UVLM_LAUNCH_CHECK();
}

// Fields of the state that an evaluation overwrites for every body (J, velocity, dadt, drdt)
// and the rest, as [first, last) field ranges.
constexpr int OUT_FIELDS[3][2]   = {{F_J, F_J + 9}, {F_U, F_DADT + 3}, {F_DSIG, F_DSIG + 1}};
constexpr int STATE_FIELDS[3][2] = {{F_X, F_A + 3}, {F_UOLD, F_SIG + 1}, {F_SIGOLD, N_FIELDS}};

// Applies a deferred build permutation to the output fields (from the previous order in st2).
void resolve_outputs(uvlm_handle h)
{
  if (!h->out_deferred) return;
  UVLM_HIP(hipSetDevice(h->prm.device));
  for (const auto &r : OUT_FIELDS)
    gather_fields(h->s, h->st2.as<double>() + (int64_t)r[0] * h->cap, h->st.as<double>() + (int64_t)r[0] * h->cap,
                  h->cap, h->cap, r[1] - r[0], h->perm.as<int32_t>(), h->N);
  h->out_deferred = false;
}

// defer_outputs: permute only the fields the next evaluation does not overwrite (uvlm_build);
// resolve_outputs permutes the rest if anything else touches the state first.
void do_build(uvlm_handle h, bool defer_outputs = false)
{
  Region      region("Tree Construction on CPU");  // fmmCalculator.hpp:73
  if (h->N <= 0) throw Error(UVLM_ERR_STATE, "no bodies set");
  hipStream_t s = h->s;
  UVLM_HIP(hipEventRecord(h->ev[EV_B0], s));
  resolve_outputs(h);
  build_tree(s, h->sc, h->F(F_X), h->F(F_X + 1), h->F(F_X + 2), h->N, h->prm.ncrit, h->tree, h->perm.as<int32_t>());
  if (defer_outputs)
    for (const auto &r : STATE_FIELDS)
      gather_fields(s, h->st.as<double>() + (int64_t)r[0] * h->cap, h->st2.as<double>() + (int64_t)r[0] * h->cap,
                    h->cap, h->cap, r[1] - r[0], h->perm.as<int32_t>(), h->N);
  else
    gather_fields(s, h->st.as<double>(), h->st2.as<double>(), h->cap, h->cap, N_FIELDS, h->perm.as<int32_t>(), h->N);
  gather_i32(s, h->orig.as<int32_t>(), h->orig2.as<int32_t>(), h->perm.as<int32_t>(), h->N);
  swap_buf(h->st, h->st2);
  swap_buf(h->orig, h->orig2);
  h->out_deferred = defer_outputs;
  UVLM_HIP(hipEventRecord(h->ev[EV_TREE], s));
  h->gap = h->near_c > 0.0 ? h->near_c * max_value(s, h->sc, h->F(F_SIG), h->N) : 0.0;
  traverse(s, h->sc, h->tree, h->tree, h->prm.theta, h->lists, h->gap);
  UVLM_HIP(hipEventRecord(h->ev[EV_LISTS], s));
  count_interactions(s, h->sc, h->tree, h->tree, h->lists, &h->lists.n_interactions);
  UVLM_HIP(hipStreamSynchronize(s));
  h->tm.tree_ms        = ev_ms(h, EV_B0, EV_TREE);
  h->tm.lists_ms       = ev_ms(h, EV_TREE, EV_LISTS);
  h->tm.n_bodies       = h->N;
  h->tm.n_cells        = h->tree.C;
  h->tm.n_leaves       = h->tree.nleaf;
  h->tm.n_levels       = h->tree.nlevels;
  h->tm.n_p2p_pairs    = h->lists.np2p;
  h->tm.n_m2l_pairs    = h->lists.nm2l;
  h->tm.n_interactions = h->lists.n_interactions;
  h->built             = true;
  h->evaluated         = false;
  h->tlists_ok         = false;
}

void do_eval(uvlm_handle h, bool sync)
{
  Region      region("Calculation of FMM");  // fmmCalculator.hpp:422
  if (!h->built) throw Error(UVLM_ERR_STATE, "uvlm_eval before uvlm_build");
  h->out_deferred = false;  // P2P stores and L2P completes every output field of every body
  hipStream_t s = h->s;
  const int   P = h->prm.P, NT = h->NT;
  const Tree &t = h->tree;
  h->M.reserve(sizeof(double2) * NT * t.C);
  h->L.reserve(sizeof(double2) * NT * t.C);
  double2 *M = h->M.as<double2>(), *L = h->L.as<double2>();
  UVLM_HIP(hipEventRecord(h->ev[EV_E0], s));
  p2m(s, P, t, h->st.as<double>(), h->cap, M);
  for (int l = t.nlevels - 2; l >= 1; l--) m2m_level(s, P, t, l, M, h->sched);
  zero_doubles(s, (double *)M, 2 * NT);
  UVLM_HIP(hipEventRecord(h->ev[EV_UP], s));
  m2l(s, P, t, t, h->lists, h->sched, M, L);
  UVLM_HIP(hipEventRecord(h->ev[EV_M2L], s));
  p2p(s, t, t, h->lists, h->st.as<double>(), h->cap, h->st.as<double>(), h->cap, h->st.as<double>(), h->cap,
      h->prm.precision, 0, h->prm.ncrit);
  UVLM_HIP(hipEventRecord(h->ev[EV_P2P], s));
  for (int l = 1; l < t.nlevels; l++) l2l_level(s, P, t, l, L, h->sched);
  l2p_stretch(s, P, t, L, h->st.as<double>(), h->cap, h->prm.stretching, h->prm.formulation, 0);
  UVLM_HIP(hipEventRecord(h->ev[EV_DOWN], s));
  h->evaluated = true;
  if (sync) {
    UVLM_HIP(hipStreamSynchronize(s));
    h->tm.upward_ms   = ev_ms(h, EV_E0, EV_UP);
    h->tm.m2l_ms      = ev_ms(h, EV_UP, EV_M2L);
    h->tm.p2p_ms      = ev_ms(h, EV_M2L, EV_P2P);
    h->tm.downward_ms = ev_ms(h, EV_P2P, EV_DOWN);
    h->tm.eval_ms     = ev_ms(h, EV_E0, EV_DOWN);
  }
}

// calculateViscosity (fmmCalculator.hpp:713-724) on the results of the last eval.
void apply_viscosity(uvlm_handle h, double dt)
{
  if (h->viscous == UVLM_PSE)
    gauss_near(h->s, h->tree, h->tree, h->lists, h->st.as<double>(), h->cap, h->st.as<double>(), h->cap,
               h->st.as<double>(), h->cap, 0, 0, h->nu / std::sqrt(2.0) * INV_PI * INV_SQRTPI);
  else if (h->viscous == UVLM_CSM)
    csm(h->s, h->st.as<double>(), h->cap, h->N, h->nu, dt);
}

// m probe records (AoS, `stride` bytes each) into h->sst, then the probe tree over them (the
// probes are permuted into its order, like the reference's createSensorTree).
void upload_probes(uvlm_handle h, const void *probes, int64_t m, int64_t stride)
{
  hipStream_t s = h->s;
  if (m > h->scap) {
    h->scap = m;
    h->sst.reserve(sizeof(double) * N_FIELDS * m);
    h->sst2.reserve(sizeof(double) * N_FIELDS * m);
    h->sperm.reserve(4 * m);
  }
  const int64_t sc = h->scap;
  h->staging.reserve((size_t)m * stride);
  UVLM_HIP(hipMemcpyAsync(h->staging.p, probes, (size_t)m * stride, hipMemcpyHostToDevice, s));
  aos_to_soa(s, h->staging.p, m, stride / 8, h->sst.as<double>(), sc);
  double *sx = h->sst.as<double>();
  build_tree(s, h->sc, sx + F_X * sc, sx + (F_X + 1) * sc, sx + (F_X + 2) * sc, m, h->prm.ncrit, h->stree,
             h->sperm.as<int32_t>());
  gather_fields(s, h->sst.as<double>(), h->sst2.as<double>(), sc, sc, N_FIELDS, h->sperm.as<int32_t>(), m);
  swap_buf(h->sst, h->sst2);
}

// The first m probe records of h->sst as 296-byte exafmm::Body records.
std::vector<double> download_probes(uvlm_handle h, int64_t m)
{
  h->staging.reserve((size_t)m * 296);
  soa_to_aos(h->s, h->sst.as<double>(), h->scap, h->staging.p, m, N_FIELDS);
  std::vector<double> out((size_t)m * N_FIELDS);
  UVLM_HIP(hipMemcpyAsync(out.data(), h->staging.p, 8 * out.size(), hipMemcpyDeviceToHost, h->s));
  UVLM_HIP(hipStreamSynchronize(h->s));
  return out;
}

void copy_out_records(const std::vector<double> &rec, int64_t m, void *dst, int64_t stride)
{
  for (int64_t i = 0; i < m; i++) std::memcpy((char *)dst + i * stride, rec.data() + i * N_FIELDS, 296);
}

// Gaussian vorticity of the particles at the probes already in h->sst (tree order): lists
// against the particle tree, velocity_old reset and filled (calculateVorticityDualTree,
// fmmCalculator.hpp:746-753, after the lists of createSensorTree 117-161).
void probe_vorticity(uvlm_handle h, int64_t m)
{
  Region region("Calculation of Vorticity");  // fmmCalculator.hpp:735
  traverse(h->s, h->sc, h->stree, h->tree, h->prm.theta, h->slists, h->gap);
  zero_fields(h->s, h->sst.as<double>(), h->scap, m, F_UOLD, 3);
  gauss_near(h->s, h->stree, h->tree, h->slists, h->sst.as<double>(), h->scap, h->st.as<double>(), h->cap,
             h->sst.as<double>(), h->scap, 1, F_UOLD, 0.0);
}

// Reserve the probe state for m probes.
void reserve_probes(uvlm_handle h, int64_t m)
{
  if (m > h->scap) {
    h->scap = m;
    h->sst.reserve(sizeof(double) * N_FIELDS * m);
    h->sst2.reserve(sizeof(double) * N_FIELDS * m);
    h->sperm.reserve(4 * m);
  }
}

// getSensorData / calculateSensorPoints (fmmCalculator.hpp:469-491, 726-731) on the m probe
// records already in h->sst: the probe tree (probes permuted into its order, h->sperm), lists
// against the particle tree, M2L from the particle multipoles, P2P velocity, L2L, L2P.
// with_j: also the near-field velocity gradient (the reference's sensor P2P gives velocity only,
// multipoleKernels.hpp:303-341; the wing's design gradient needs the whole J at its points).
void sensor_pass(uvlm_handle h, int64_t m, bool with_j = false)
{
  Region        region("Calculation of Sensors");  // fmmCalculator.hpp:471
  hipStream_t   s  = h->s;
  const int64_t sc = h->scap;
  double       *sx = h->sst.as<double>();
  build_tree(s, h->sc, sx + F_X * sc, sx + (F_X + 1) * sc, sx + (F_X + 2) * sc, m, h->prm.ncrit, h->stree,
             h->sperm.as<int32_t>());
  gather_fields(s, h->sst.as<double>(), h->sst2.as<double>(), sc, sc, N_FIELDS, h->sperm.as<int32_t>(), m);
  swap_buf(h->sst, h->sst2);
  traverse(s, h->sc, h->stree, h->tree, h->prm.theta, h->slists, h->gap);
  const int P = h->prm.P;
  h->Ls.reserve(sizeof(double2) * h->NT * h->stree.C);
  double2 *Ls = h->Ls.as<double2>();
  reset_sensor_rates(s, h->sst.as<double>(), sc, m);
  m2l(s, P, h->stree, h->tree, h->slists, h->sched, h->M.as<double2>(), Ls, &h->sc);
  p2p(s, h->stree, h->tree, h->slists, h->sst.as<double>(), sc, h->st.as<double>(), h->cap, h->sst.as<double>(), sc,
      h->prm.precision, with_j ? 0 : 1, h->prm.ncrit, &h->sc);
  for (int l = 1; l < h->stree.nlevels; l++) l2l_level(s, P, h->stree, l, Ls, h->sched);
  l2p_stretch(s, P, h->stree, Ls, h->sst.as<double>(), sc, h->prm.stretching, h->prm.formulation, 1);
}

// Velocity of the particles (their FMM at the current state: build + eval) at m device points
// (AoS [m][3]) into out (AoS [m][3], same order).  No particles: zero.
void velocity_at_points(uvlm_handle h, int64_t m, const double *pts, double *out)
{
  if (h->N <= 0) {
    zero_doubles(h->s, out, 3 * m);
    return;
  }
  do_build(h);
  do_eval(h, false);
  reserve_probes(h, m);
  points_to_soa(h->s, m, pts, h->sst.as<double>(), h->scap);
  sensor_pass(h, m);
  probe_velocity_out(h->s, m, h->sperm.as<int32_t>(), h->sst.as<double>(), h->scap, out);
}

// Velocity and its gradient J (dv/dx, near field included) of the particles at m device points
// (AoS [m][3]): out [m][3], Jout [m][9], in the points' order.  No particles: zero.
void velocity_grad_at_points(uvlm_handle h, int64_t m, const double *pts, double *out, double *Jout)
{
  if (h->N <= 0) {
    zero_doubles(h->s, out, 3 * m);
    zero_doubles(h->s, Jout, 9 * m);
    return;
  }
  do_build(h);
  do_eval(h, false);
  reserve_probes(h, m);
  points_to_soa(h->s, m, pts, h->sst.as<double>(), h->scap);
  sensor_pass(h, m, true);
  probe_velocity_grad_out(h->s, m, h->sperm.as<int32_t>(), h->sst.as<double>(), h->scap, out, Jout);
}

// Append k bodies (device AoS X [k][3], alpha [k][3], sigma [k]) after the current N, growing
// the state (fields keep their contents) when needed; new bodies get the next original indices.
// Capacity for at least `want` particles; the fields and original indices of the current ones
// are kept (the field stride changes).
void grow_particles(uvlm_handle h, int64_t want)
{
  resolve_outputs(h);
  const int64_t N = h->N;
  if (want > h->cap) {
    const int64_t ncap = std::max<int64_t>(want, h->cap + h->cap / 2);
    DevBuf        nst, norig;
    nst.reserve(sizeof(double) * N_FIELDS * ncap);
    norig.reserve(4 * ncap);
    for (int f = 0; f < N_FIELDS && N > 0; f++)
      UVLM_HIP(hipMemcpyAsync(nst.as<double>() + f * ncap, h->F(f), 8 * N, hipMemcpyDeviceToDevice, h->s));
    if (N > 0) UVLM_HIP(hipMemcpyAsync(norig.p, h->orig.p, 4 * N, hipMemcpyDeviceToDevice, h->s));
    UVLM_HIP(hipStreamSynchronize(h->s));
    swap_buf(h->st, nst);
    swap_buf(h->orig, norig);
    h->cap = ncap;
    h->st2.reserve(sizeof(double) * N_FIELDS * ncap);
    h->orig2.reserve(4 * ncap);
    h->perm.reserve(4 * ncap);
  }
}

void append_bodies_dev(uvlm_handle h, int64_t k, const double *X, const double *A, const double *S)
{
  const int64_t N = h->N, want = N + k;
  grow_particles(h, want);
  append_bodies(h->s, N, k, X, A, S, h->st.as<double>(), h->cap, h->orig.as<int32_t>());
  h->N     = want;
  h->built = h->evaluated = false;
}

// Adjoint of the probe objective sum_p w_p . u(x_p) for the m probe "vortices" already in h->sst
// (X = probe, alpha = w_p, sigma = 1) on the particles' current tree: the forward probe lists
// transposed, probe multipoles -> particle locals (M2L), L2L, P2P with the particle's sigma, and
// L2P with transposed stretching.  Leaves dJ/dalpha in h->ast velocity and dJ/dx in h->ast dadt
// (device order).
void probe_adjoint_core(uvlm_handle h, int64_t m)
{
  hipStream_t   s  = h->s;
  const int     P  = h->prm.P, NT = h->NT;
  const int64_t sc = h->scap;
  double       *sx = h->sst.as<double>();
  build_tree(s, h->sc, sx + F_X * sc, sx + (F_X + 1) * sc, sx + (F_X + 2) * sc, m, h->prm.ncrit, h->stree,
             h->sperm.as<int32_t>());
  gather_fields(s, h->sst.as<double>(), h->sst2.as<double>(), sc, sc, N_FIELDS, h->sperm.as<int32_t>(), m);
  swap_buf(h->sst, h->sst2);
  // the forward probe lists, transposed: targets = particle cells, sources = probe cells
  traverse(s, h->sc, h->stree, h->tree, h->prm.theta, h->slists, h->gap);
  transpose_lists(s, h->sc, h->stree, h->tree, h->slists, h->alists);
  // probe multipoles (root included: a probe cell is a source here)
  h->Ms.reserve(sizeof(double2) * NT * h->stree.C);
  double2 *Ms = h->Ms.as<double2>();
  p2m(s, P, h->stree, h->sst.as<double>(), sc, Ms);
  for (int l = h->stree.nlevels - 2; l >= 0; l--) m2m_level(s, P, h->stree, l, Ms, h->sched);
  // adjoint state of the particles: positions, strengths and sigma copied in
  const int64_t cap = h->cap;
  h->ast.reserve(sizeof(double) * N_FIELDS * cap);
  double *a = h->ast.as<double>();
  zero_doubles(s, a, N_FIELDS * cap);
  UVLM_HIP(hipMemcpyAsync(a + F_X * cap, h->F(F_X), 8 * 6 * cap, hipMemcpyDeviceToDevice, s));
  UVLM_HIP(hipMemcpyAsync(a + F_SIG * cap, h->F(F_SIG), 8 * cap, hipMemcpyDeviceToDevice, s));
  h->L.reserve(sizeof(double2) * NT * h->tree.C);
  double2 *La = h->L.as<double2>();
  m2l(s, P, h->tree, h->stree, h->alists, h->sched, Ms, La);
  zero_doubles(s, (double *)La, 2 * NT);  // the particle root multipole is never formed forward
  for (int l = 1; l < h->tree.nlevels; l++) l2l_level(s, P, h->tree, l, La, h->sched);
  p2p(s, h->tree, h->stree, h->alists, a, cap, h->sst.as<double>(), sc, a, cap, h->prm.precision, 2, h->prm.ncrit);
  l2p_stretch(s, P, h->tree, La, a, cap, UVLM_TRANSPOSED, UVLM_CVPM, 0);
  h->evaluated = false;  // L now holds adjoint locals
}

int stages_of(int scheme) { return scheme == UVLM_EULER ? 1 : scheme == UVLM_RK2 ? 2 : scheme == UVLM_RK3 ? 3 : 4; }

// The stages of one step as a Butcher tableau (y1 = y0 + dt sum_s b_s f_s, stage states
// Y_s = y0 + dt sum_{r<s} a_sr f_r), derived from the register updates of the rk_stage kernel
// (fmmCalculator.hpp:511-711): RK2 Heun; RK3 Williamson 2N with A = (0, -5/9, -153/128),
// B = (1/3, 15/16, 8/15); RK4 as the reference writes it (every intermediate state at dt/2).
struct Tableau {
  int    S = 1;
  double a[4][4]{}, b[4]{};
};
Tableau tableau_of(int scheme)
{
  Tableau t;
  t.S = stages_of(scheme);
  if (scheme == UVLM_EULER) {
    t.b[0] = 1.0;
  } else if (scheme == UVLM_RK2) {
    t.a[1][0] = 1.0;
    t.b[0] = t.b[1] = 0.5;
  } else if (scheme == UVLM_RK3) {
    // z_1 = dt f_0; z_k = A_k z_{k-1} + dt f_{k-1}; y <- y + B_k z_k after each stage
    const double A[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0}, B[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
    double       z[3] = {0, 0, 0}, y[3] = {0, 0, 0};  // coefficients of f_0..f_2 (per dt)
    for (int k = 0; k < 3; k++) {
      for (int r = 0; r < 3; r++) z[r] *= A[k];
      z[k] += 1.0;
      for (int r = 0; r < 3; r++) y[r] += B[k] * z[r];
      if (k < 2)
        for (int r = 0; r < 3; r++) t.a[k + 1][r] = y[r];
    }
    for (int r = 0; r < 3; r++) t.b[r] = y[r];
  } else {
    t.a[1][0] = t.a[2][1] = t.a[3][2] = 0.5;
    t.b[0] = t.b[3] = 1.0 / 6.0;
    t.b[1] = t.b[2] = 1.0 / 3.0;
  }
  return t;
}

// Vector-Jacobian product of the last do_eval (forward J in the state, forward L in h->L):
// on entry the adjoint fields hold ubar (0-2) and abar (15-17) in the device order, on exit
// xbar (12-14) and alphabar (15-17).  M and L are overwritten by the adjoint expansions.
// sbar (optional, device order): also the cotangent of the core sizes (near pairs only).
void vjp_core(uvlm_handle h, double *sbar = nullptr)
{
  hipStream_t s = h->s;
  const int   P = h->prm.P, NT = h->NT;
  const Tree &t = h->tree;
  const int64_t cap = h->cap, N = h->N;
  double       *aux = h->vaux.as<double>(), *st = h->st.as<double>();
  vjp_setup(s, N, st, aux, cap, h->prm.stretching == UVLM_CLASSIC);
  // target side: far field from the forward local expansions, then the near pairs
  l2p_adj_tgt(s, P, t, h->L.as<double2>(), st, aux, cap);
  p2p_adj(s, t, h->lists, 0, h->prm.precision == UVLM_MIXED, st, aux, cap);
  // source side: adjoint sources up the tree, across the transposed M2L list, down to the bodies
  if (!h->tlists_ok) {
    transpose_lists(s, h->sc, t, t, h->lists, h->tlists);
    h->tlists_ok = true;
  }
  double2 *M = h->M.as<double2>(), *L = h->L.as<double2>();
  p2m_adj(s, P, t, st, aux, cap, M);
  for (int l = t.nlevels - 2; l >= 1; l--) m2m_level(s, P, t, l, M, h->sched);
  zero_doubles(s, (double *)M, 2 * NT);
  m2l(s, P, t, t, h->tlists, h->sched, M, L);
  for (int l = 1; l < t.nlevels; l++) l2l_level(s, P, t, l, L, h->sched);
  h->ast.reserve(sizeof(double) * N_FIELDS * cap);
  double *a = h->ast.as<double>();
  zero_doubles(s, a + F_J * cap, 9 * cap);
  zero_doubles(s, a + F_U * cap, 3 * cap);
  UVLM_HIP(hipMemcpyAsync(a + F_X * cap, st + F_X * cap, 8 * 6 * cap, hipMemcpyDeviceToDevice, s));
  UVLM_HIP(hipMemcpyAsync(a + F_SIG * cap, st + F_SIG * cap, 8 * cap, hipMemcpyDeviceToDevice, s));
  l2p_stretch(s, P, t, L, a, cap, UVLM_TRANSPOSED, UVLM_CVPM, 0);
  p2p_adj(s, t, h->tlists, 1, h->prm.precision == UVLM_MIXED, st, aux, cap);
  vjp_finish(s, N, a, aux, cap);
  if (sbar) sigma_bar(s, t, t, h->tlists, st, cap, aux, cap, true, st, cap, sbar);
  h->evaluated = false;  // M and L hold adjoint expansions
}

void check_vjp_config(uvlm_handle h)
{
  if (h->prm.formulation != UVLM_CVPM) throw Error(UVLM_ERR_ARG, "the rate VJP covers the cVPM formulation");
  if (h->viscous != UVLM_INVISCID) throw Error(UVLM_ERR_ARG, "the rate VJP covers inviscid steps");
}

// State of the current bodies (X, alpha, sigma and original index, device order) into slot k
// of a store of 7-field records, and back.
void save_state(uvlm_handle h, DevBuf &store, DevBuf &ostore, int k)
{
  const int64_t cap = h->cap;
  double       *dst = store.as<double>() + (int64_t)k * 7 * cap;
  UVLM_HIP(hipMemcpyAsync(dst, h->F(F_X), 8 * 6 * cap, hipMemcpyDeviceToDevice, h->s));
  UVLM_HIP(hipMemcpyAsync(dst + 6 * cap, h->F(F_SIG), 8 * cap, hipMemcpyDeviceToDevice, h->s));
  UVLM_HIP(hipMemcpyAsync(ostore.as<int32_t>() + (int64_t)k * cap, h->orig.p, 4 * cap, hipMemcpyDeviceToDevice, h->s));
}
void load_state(uvlm_handle h, const DevBuf &store, const DevBuf &ostore, int k)
{
  const int64_t cap = h->cap;
  const double *src = store.as<double>() + (int64_t)k * 7 * cap;
  UVLM_HIP(hipMemcpyAsync(h->F(F_X), src, 8 * 6 * cap, hipMemcpyDeviceToDevice, h->s));
  UVLM_HIP(hipMemcpyAsync(h->F(F_SIG), src + 6 * cap, 8 * cap, hipMemcpyDeviceToDevice, h->s));
  UVLM_HIP(hipMemcpyAsync(h->orig.p, ostore.as<int32_t>() + (int64_t)k * cap, 4 * cap, hipMemcpyDeviceToDevice, h->s));
  h->built = h->evaluated = false;
}

// Cotangent of one step of uvlm_advance: on entry cot (6 fields of cap doubles, original order)
// holds the cotangent of the state after the step, on exit that of the state before it.  The
// bodies are left at the step's start.  Stage states are recomputed, and in reverse each stage
// rebuilds its own tree (the build is deterministic, so it is the forward's tree).
// sbar (optional, original order, cap doubles): the core sizes' cotangent, accumulated in place
// (sigma does not change over a cVPM inviscid step, so each stage adds its rates' part).
void step_vjp(uvlm_handle h, double dt, double *cot, double *sbar = nullptr)
{
  hipStream_t   s   = h->s;
  const int64_t cap = h->cap, N = h->N;
  const Tableau T   = tableau_of(h->prm.scheme);
  h->vaux.reserve(sizeof(double) * N_ADJ_FIELDS * cap);
  h->vstage.reserve(sizeof(double) * 7 * cap * T.S);
  h->vorig.reserve(sizeof(int32_t) * cap * T.S);
  h->vcot.reserve(sizeof(double) * 6 * cap * (T.S + 1));
  // slot s: the state stage s builds and evaluates at (slot 0 = the step's start)
  save_state(h, h->vstage, h->vorig, 0);
  for (int k = 0; k + 1 < T.S; k++) {
    do_build(h);
    do_eval(h, false);
    rk_stage(s, h->st.as<double>(), cap, N, h->prm.scheme, k, dt, 0);
    save_state(h, h->vstage, h->vorig, k + 1);
  }
  double *Yb = h->vcot.as<double>();  // Ybar_s at Yb + s * 6 cap; fbar at Yb + S * 6 cap
  double *fb = Yb + (int64_t)T.S * 6 * cap;
  for (int k = T.S - 1; k >= 0; k--) {
    const double *p[6];
    double        c[6];
    int           n = 0;
    p[n]            = cot;
    c[n++]          = dt * T.b[k];
    for (int r = k + 1; r < T.S; r++)
      if (T.a[r][k] != 0.0) {
        p[n]   = Yb + (int64_t)r * 6 * cap;
        c[n++] = dt * T.a[r][k];
      }
    lincomb(s, 6 * cap, fb, n, p, c);
    load_state(h, h->vstage, h->vorig, k);
    do_build(h);
    do_eval(h, false);
    cot_gather(s, N, h->orig.as<int32_t>(), fb, cap, h->vaux.as<double>(), cap);
    if (sbar) h->vsig.reserve(sizeof(double) * cap);
    vjp_core(h, sbar ? h->vsig.as<double>() : nullptr);
    cot_scatter(s, N, h->orig.as<int32_t>(), h->vaux.as<double>(), cap, Yb + (int64_t)k * 6 * cap, cap);
    if (sbar) add_to_orig(s, N, h->orig.as<int32_t>(), h->vsig.as<double>(), sbar);
  }
  const double *p[6];
  double        c[6];
  p[0] = cot;
  c[0] = 1.0;
  for (int k = 0; k < T.S; k++) {
    p[k + 1] = Yb + (int64_t)k * 6 * cap;
    c[k + 1] = 1.0;
  }
  lincomb(s, 6 * cap, fb, T.S + 1, p, c);
  UVLM_HIP(hipMemcpyAsync(cot, fb, 8 * 6 * cap, hipMemcpyDeviceToDevice, s));
  load_state(h, h->vstage, h->vorig, 0);
}

// The particle state (X, alpha, sigma) as 7 fields of N doubles in the ORIGINAL body order,
// and back with the original index as the device order.
void save_state_orig(uvlm_handle h, double *dst) { state_to_orig(h->s, h->N, h->orig.as<int32_t>(), h->st.as<double>(), h->cap, dst); }
// Like load_state_orig with fields of `stride` doubles, of which the first `count` are loaded.
void load_state_orig_n(uvlm_handle h, const double *src, int64_t stride, int64_t count)
{
  if (count > h->cap) throw Error(UVLM_ERR_INTERNAL, "state larger than the particle capacity");
  h->N = count;
  for (int f = 0; f < 6; f++)
    UVLM_HIP(hipMemcpyAsync(h->F(F_X + f), src + f * stride, 8 * count, hipMemcpyDeviceToDevice, h->s));
  UVLM_HIP(hipMemcpyAsync(h->F(F_SIG), src + 6 * stride, 8 * count, hipMemcpyDeviceToDevice, h->s));
  iota(h->s, h->orig.as<int32_t>(), count);
  h->built = h->evaluated = false;
}
void load_state_orig(uvlm_handle h, const double *src)
{
  const int64_t N = h->N;
  for (int f = 0; f < 6; f++)
    UVLM_HIP(hipMemcpyAsync(h->F(F_X + f), src + f * N, 8 * N, hipMemcpyDeviceToDevice, h->s));
  UVLM_HIP(hipMemcpyAsync(h->F(F_SIG), src + 6 * N, 8 * N, hipMemcpyDeviceToDevice, h->s));
  iota(h->s, h->orig.as<int32_t>(), N);
  h->built = h->evaluated = false;
}

// Host AoS (n x 3) pairs <-> 6 SoA fields of `cap` doubles on the device.
void upload_cot(uvlm_handle h, const double *xb, const double *ab, double *dev)
{
  const int64_t       n = h->N, cap = h->cap;
  std::vector<double> buf((size_t)6 * cap, 0.0);
  for (int64_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      buf[(size_t)d * cap + i]       = xb ? xb[3 * i + d] : 0.0;
      buf[(size_t)(3 + d) * cap + i] = ab ? ab[3 * i + d] : 0.0;
    }
  UVLM_HIP(hipMemcpyAsync(dev, buf.data(), 8 * buf.size(), hipMemcpyHostToDevice, h->s));
  UVLM_HIP(hipStreamSynchronize(h->s));
}
void download_cot(uvlm_handle h, const double *dev, double *xb, double *ab)
{
  const int64_t       n = h->N, cap = h->cap;
  std::vector<double> buf((size_t)6 * cap);
  UVLM_HIP(hipMemcpyAsync(buf.data(), dev, 8 * buf.size(), hipMemcpyDeviceToHost, h->s));
  UVLM_HIP(hipStreamSynchronize(h->s));
  for (int64_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      if (xb) xb[3 * i + d] = buf[(size_t)d * cap + i];
      if (ab) ab[3 * i + d] = buf[(size_t)(3 + d) * cap + i];
    }
}

// The handle check of the entry points that read or change the state: a permutation that
// uvlm_build left pending on the output fields is applied first.
void check_handle(uvlm_handle h)
{
  if (!h) throw Error(UVLM_ERR_ARG, "null handle");
  resolve_outputs(h);
}
// The handle check of uvlm_build, uvlm_eval and the queries that leave the state alone.
void check_handle_keep(uvlm_handle h)
{
  if (!h) throw Error(UVLM_ERR_ARG, "null handle");
}

// Device staging of one direct sum: targets, their strengths, u, J and dadt ([m][3], [m][3],
// [m][3], [m][9], [m][3]) in consecutive blocks, the absent ones skipped.  The kernel writes
// every element of the outputs it is given, so nothing in `d` has to start at zero.
void direct_staged(uvlm_handle h, int64_t m, const double *xt, const double *at, double *u, double *J, double *dadt,
                   double *d)
{
  const int64_t nin = 3 * m + (at ? 3 * m : 0);
  double       *dxt = d, *dat = at ? d + 3 * m : nullptr;
  double *du = d + nin, *dJ = J ? du + 3 * m : nullptr, *dd = dadt ? du + 3 * m + (J ? 9 * m : 0) : nullptr;
  UVLM_HIP(hipMemcpyAsync(dxt, xt, 24 * m, hipMemcpyHostToDevice, h->s));
  if (at) UVLM_HIP(hipMemcpyAsync(dat, at, 24 * m, hipMemcpyHostToDevice, h->s));
  direct(h->s, m, dxt, dat, h->st.as<double>(), h->cap, h->N, du, dJ, dd);
  UVLM_HIP(hipMemcpyAsync(u, du, 24 * m, hipMemcpyDeviceToHost, h->s));
  if (J) UVLM_HIP(hipMemcpyAsync(J, dJ, 72 * m, hipMemcpyDeviceToHost, h->s));
  if (dadt) UVLM_HIP(hipMemcpyAsync(dadt, dd, 24 * m, hipMemcpyDeviceToHost, h->s));
  UVLM_HIP(hipStreamSynchronize(h->s));
}

void check_direct_args(uvlm_handle h, int64_t m, const double *xt, const double *at, const double *u,
                       const double *dadt)
{
  check_handle(h);
  if (h->N <= 0) throw Error(UVLM_ERR_STATE, "no bodies set");
  if (m <= 0 || !xt || !u) throw Error(UVLM_ERR_ARG, "null or empty targets");
  if (dadt && !at) throw Error(UVLM_ERR_ARG, "dadt needs target strengths");
}

}  // namespace

// Staging of uvlm_direct_noalloc, one copy per GPU for the life of the process.  In host code
// `scratch` only names the symbol; its device address on each GPU comes from
// hipGetSymbolAddress (scratch_setup, called by uvlm_create with the handle's device current).
static __device__ __align__(256) double scratch[(size_t)UVLM_DIRECT_NOALLOC_MAX * 21];

namespace {
constexpr int kMaxDevices = 64;
std::mutex    g_scratch_mu;
double       *g_scratch_dev[kMaxDevices] = {};

void scratch_setup(int device)
{
  if (device < 0 || device >= kMaxDevices) throw Error(UVLM_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return;
  void *p = nullptr;
  UVLM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Error(UVLM_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratch_dev[device] = static_cast<double *>(p);
}

double *scratch_on(int device)
{
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  double *p = (device >= 0 && device < kMaxDevices) ? g_scratch_dev[device] : nullptr;
  if (!p) throw Error(UVLM_ERR_STATE, "scratch array not set up on this device (create a handle on it first)");
  return p;
}
}  // namespace

extern "C" {

const char *uvlm_last_error(void) { return g_last_error.c_str(); }
int         uvlm_abi_version(void) { return UVLM_ABI_VERSION; }

int uvlm_create(const uvlm_params *p, uvlm_handle *out)
{
  if (!p || !out) return fail(UVLM_ERR_ARG, "null argument");
  *out = nullptr;
  if (p->P < 2 || p->P > 16) return fail(UVLM_ERR_ARG, "P must be in [2, 16]");
  if (!(p->theta > 0.0)) return fail(UVLM_ERR_ARG, "theta must be > 0");
  if (p->ncrit < 1) return fail(UVLM_ERR_ARG, "ncrit must be >= 1");
  if (p->precision != UVLM_FP64 && p->precision != UVLM_MIXED) return fail(UVLM_ERR_ARG, "bad precision");
  if (p->stretching != UVLM_TRANSPOSED && p->stretching != UVLM_CLASSIC) return fail(UVLM_ERR_ARG, "bad stretching");
  if (p->formulation != UVLM_CVPM && p->formulation != UVLM_RVPM) return fail(UVLM_ERR_ARG, "bad formulation");
  if (p->scheme < UVLM_EULER || p->scheme > UVLM_RK4) return fail(UVLM_ERR_ARG, "bad scheme");
  uvlm_handle h = new (std::nothrow) uvlm_handle_s();
  if (!h) return fail(UVLM_ERR_NOMEM, "handle allocation failed");
  int rc = guarded([&] {
    h->prm = *p;
    h->NT  = 3 * p->P * (p->P + 1) / 2;
    set_device(h);
    scratch_setup(p->device);
    UVLM_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    for (auto &e : h->ev) UVLM_HIP(hipEventCreate(&e));
    for (auto &e : h->aev) UVLM_HIP(hipEventCreate(&e));
    // UVLM_M2L_SLOTS overrides the M2L plan's accumulator-slot budget (uvlm_autotune also sets
    // it); UVLM_M2L_WAVES (1 or 2) the wavefronts per M2L workgroup, chosen by P otherwise
    if (const char *sb = std::getenv("UVLM_M2L_SLOTS")) h->sched.budget = std::atoi(sb);
    if (const char *sw = std::getenv("UVLM_M2L_WAVES")) h->sched.waves = std::atoi(sw);
    make_m2l_schedule(p->P, h->sched);
  });
  if (rc != UVLM_OK) {
    uvlm_destroy(h);
    return rc;
  }
  *out = h;
  return UVLM_OK;
}

int uvlm_destroy(uvlm_handle h)
{
  if (!h) return UVLM_OK;
  (void)hipSetDevice(h->prm.device);
  if (h->s) (void)hipStreamSynchronize(h->s);
  for (auto &e : h->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : h->aev)
    if (e) (void)hipEventDestroy(e);
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
  return UVLM_OK;
}

int uvlm_set_bodies_aos(uvlm_handle h, const void *bodies, int64_t n, int64_t stride)
{
  Region region("copyBodies2Device");  // fmmCalculator.hpp:216
  return guarded([&] {
    check_handle(h);
    if (n <= 0 || !bodies) throw Error(UVLM_ERR_ARG, "empty or null bodies");
    if (stride < 296 || stride % 8) throw Error(UVLM_ERR_ARG, "stride must be >= 296 and a multiple of 8");
    set_device(h);
    h->N = n;
    h->ensure_cap(n);
    h->staging.reserve((size_t)n * stride);
    UVLM_HIP(hipMemcpyAsync(h->staging.p, bodies, (size_t)n * stride, hipMemcpyHostToDevice, h->s));
    aos_to_soa(h->s, h->staging.p, n, stride / 8, h->st.as<double>(), h->cap);
    iota(h->s, h->orig.as<int32_t>(), n);
    UVLM_HIP(hipStreamSynchronize(h->s));
    h->built = h->evaluated = false;
  });
}

int uvlm_get_bodies_aos(uvlm_handle h, void *bodies, int64_t n, int64_t stride)
{
  Region region("copyBodies2Host");  // fmmCalculator.hpp:237
  return guarded([&] {
    check_handle(h);
    if (n != h->N || !bodies) throw Error(UVLM_ERR_ARG, "size mismatch");
    if (stride < 296 || stride % 8) throw Error(UVLM_ERR_ARG, "stride must be >= 296 and a multiple of 8");
    set_device(h);
    h->staging.reserve((size_t)n * stride);
    if (stride > 296) UVLM_HIP(hipMemcpyAsync(h->staging.p, bodies, (size_t)n * stride, hipMemcpyHostToDevice, h->s));
    soa_to_aos(h->s, h->st.as<double>(), h->cap, h->staging.p, n, stride / 8);
    UVLM_HIP(hipMemcpyAsync(bodies, h->staging.p, (size_t)n * stride, hipMemcpyDeviceToHost, h->s));
    UVLM_HIP(hipStreamSynchronize(h->s));
  });
}

int uvlm_set_bodies_soa(uvlm_handle h, int64_t n, const double *X, const double *alpha, const double *sigma)
{
  return guarded([&] {
    check_handle(h);
    if (n <= 0 || !X || !alpha || !sigma) throw Error(UVLM_ERR_ARG, "null or empty input");
    set_device(h);
    h->N = n;
    h->ensure_cap(n);
    std::vector<double> host((size_t)N_FIELDS * n, 0.0);
    for (int64_t i = 0; i < n; i++) {
      for (int d = 0; d < 3; d++) {
        host[(F_X + d) * n + i] = X[3 * i + d];
        host[(F_A + d) * n + i] = alpha[3 * i + d];
      }
      host[F_SIG * n + i] = sigma[i];
    }
    for (int f = 0; f < N_FIELDS; f++)
      UVLM_HIP(hipMemcpyAsync(h->F(f), host.data() + (size_t)f * n, 8 * n, hipMemcpyHostToDevice, h->s));
    iota(h->s, h->orig.as<int32_t>(), n);
    UVLM_HIP(hipStreamSynchronize(h->s));
    h->built = h->evaluated = false;
  });
}

int uvlm_get_bodies_soa(uvlm_handle h, int64_t n, double *X, double *alpha, double *sigma, double *velocity,
                        double *J, double *dadt, double *drdt)
{
  return guarded([&] {
    check_handle(h);
    if (n != h->N) throw Error(UVLM_ERR_ARG, "size mismatch");
    set_device(h);
    std::vector<double> buf((size_t)9 * n);
    auto fetch = [&](int f0, int nf, double *out) {
      if (!out) return;
      for (int f = 0; f < nf; f++)
        UVLM_HIP(hipMemcpyAsync(buf.data() + (size_t)f * n, h->F(f0 + f), 8 * n, hipMemcpyDeviceToHost, h->s));
      UVLM_HIP(hipStreamSynchronize(h->s));
      for (int64_t i = 0; i < n; i++)
        for (int f = 0; f < nf; f++) out[nf * i + f] = buf[(size_t)f * n + i];
    };
    fetch(F_X, 3, X);
    fetch(F_A, 3, alpha);
    fetch(F_SIG, 1, sigma);
    fetch(F_U, 3, velocity);
    fetch(F_J, 9, J);
    fetch(F_DADT, 3, dadt);
    fetch(F_DSIG, 1, drdt);
  });
}

int uvlm_get_permutation(uvlm_handle h, int64_t n, int64_t *perm)
{
  return guarded([&] {
    check_handle(h);
    if (n != h->N || !perm) throw Error(UVLM_ERR_ARG, "size mismatch");
    set_device(h);
    std::vector<int32_t> p(n);
    UVLM_HIP(hipMemcpyAsync(p.data(), h->orig.p, 4 * n, hipMemcpyDeviceToHost, h->s));
    UVLM_HIP(hipStreamSynchronize(h->s));
    for (int64_t i = 0; i < n; i++) perm[i] = p[i];
  });
}

int uvlm_build(uvlm_handle h)
{
  Region region("Initialization of FMM");  // fmmCalculator.hpp:284
  return guarded([&] {
    check_handle_keep(h);
    set_device(h);
    do_build(h, true);
  });
}

int uvlm_eval(uvlm_handle h)
{
  return guarded([&] {
    check_handle_keep(h);
    set_device(h);
    do_eval(h, true);
  });
}

int uvlm_stage(uvlm_handle h, int stage, double dt)
{
  return guarded([&] {
    check_handle(h);
    if (!h->evaluated) throw Error(UVLM_ERR_STATE, "uvlm_stage before uvlm_eval");
    if (stage < 0 || stage >= stages_of(h->prm.scheme)) throw Error(UVLM_ERR_ARG, "bad stage");
    set_device(h);
    rk_stage(h->s, h->st.as<double>(), h->cap, h->N, h->prm.scheme, stage, dt, 0);
    UVLM_HIP(hipStreamSynchronize(h->s));
    h->evaluated = false;
  });
}

int uvlm_advance(uvlm_handle h, double dt)
{
  return guarded([&] {
    check_handle(h);
    set_device(h);
    for (int k = 0; k < stages_of(h->prm.scheme); k++) {
      do_build(h);
      do_eval(h, true);
      apply_viscosity(h, dt);
      rk_stage(h->s, h->st.as<double>(), h->cap, h->N, h->prm.scheme, k, dt, 0);
    }
    UVLM_HIP(hipStreamSynchronize(h->s));
    h->evaluated = false;
  });
}

int uvlm_num_bodies(uvlm_handle h, int64_t *n)
{
  if (!h || !n) return fail(UVLM_ERR_ARG, "null argument");
  *n = h->N;
  return UVLM_OK;
}

int uvlm_num_cells(uvlm_handle h, int64_t *n_cells)
{
  return guarded([&] {
    check_handle_keep(h);
    if (!h->built) throw Error(UVLM_ERR_STATE, "no tree");
    *n_cells = h->tree.C;
  });
}

int uvlm_get_tree(uvlm_handle h, int64_t n_cells, int64_t *ints, double *reals)
{
  return guarded([&] {
    check_handle(h);
    if (!h->built) throw Error(UVLM_ERR_STATE, "no tree");
    if (n_cells != h->tree.C) throw Error(UVLM_ERR_ARG, "size mismatch");
    set_device(h);
    const int64_t        C = n_cells;
    std::vector<int32_t> iv(6 * C);
    std::vector<double>  rv(4 * C);
    Tree                &t = h->tree;
    DevBuf *ib[5] = {&t.nchild, &t.nbody, &t.child, &t.body, &t.level};
    for (int k = 0; k < 5; k++)
      UVLM_HIP(hipMemcpyAsync(iv.data() + k * C, ib[k]->p, 4 * C, hipMemcpyDeviceToHost, h->s));
    DevBuf *rb[4] = {&t.cx, &t.cy, &t.cz, &t.R};
    for (int k = 0; k < 4; k++)
      UVLM_HIP(hipMemcpyAsync(rv.data() + k * C, rb[k]->p, 8 * C, hipMemcpyDeviceToHost, h->s));
    UVLM_HIP(hipStreamSynchronize(h->s));
    for (int64_t c = 0; c < C; c++) {
      for (int k = 0; k < 5; k++) ints[5 * c + k] = iv[k * C + c];
      for (int k = 0; k < 4; k++) reals[4 * c + k] = rv[k * C + c];
    }
  });
}

int uvlm_get_lists(uvlm_handle h, int64_t *n_p2p, int64_t *n_m2l, int64_t *p2p_out, int64_t *m2l_out)
{
  return guarded([&] {
    check_handle(h);
    if (!h->built) throw Error(UVLM_ERR_STATE, "no lists");
    if (n_p2p) *n_p2p = h->lists.np2p;
    if (n_m2l) *n_m2l = h->lists.nm2l;
    set_device(h);
    const int64_t C = h->tree.C;
    auto          expand = [&](DevBuf &off, DevBuf &src, int64_t n, int64_t *out) {
      if (!out) return;
      std::vector<int32_t> o(C + 1), sv(std::max<int64_t>(n, 1));
      UVLM_HIP(hipMemcpyAsync(o.data(), off.p, 4 * (C + 1), hipMemcpyDeviceToHost, h->s));
      if (n) UVLM_HIP(hipMemcpyAsync(sv.data(), src.p, 4 * n, hipMemcpyDeviceToHost, h->s));
      UVLM_HIP(hipStreamSynchronize(h->s));
      for (int64_t c = 0; c < C; c++)
        for (int q = o[c]; q < o[c + 1]; q++) {
          out[2 * q]     = c;
          out[2 * q + 1] = sv[q];
        }
    };
    expand(h->lists.p2p_off, h->lists.p2p_src, h->lists.np2p, p2p_out);
    expand(h->lists.m2l_off, h->lists.m2l_src, h->lists.nm2l, m2l_out);
  });
}

int uvlm_get_list(int device, double theta, int64_t ct, const int64_t *t_ints, const double *t_reals, int64_t cs,
                  const int64_t *s_ints, const double *s_reals, int64_t cap_p2p, int64_t cap_m2l, int64_t *n_p2p,
                  int64_t *n_m2l, int64_t *p2p_out, int64_t *m2l_out)
{
  hipStream_t s = nullptr;
  const int   rc = guarded([&] {
    if (!(theta > 0.0)) throw Error(UVLM_ERR_ARG, "theta must be > 0");
    if (!t_ints || !t_reals || !s_ints || !s_reals || !n_p2p || !n_m2l) throw Error(UVLM_ERR_ARG, "null argument");
    if (ct < 1 || cs < 1 || ct >= (int64_t(1) << 31) || cs >= (int64_t(1) << 31))
      throw Error(UVLM_ERR_ARG, "cell counts must be in [1, 2^31)");
    // Children must lie after their parent and inside the array: then every descent of the
    // traversal visits strictly larger cell indices and ends.
    auto check_tree = [](int64_t C, const int64_t *ints, const char *which) {
      for (int64_t c = 0; c < C; c++) {
        const int64_t nch = ints[5 * c], ch = ints[5 * c + 2];
        if (nch < 0 || nch > 8 || (nch > 0 && (ch <= c || ch + nch > C)))
          throw Error(UVLM_ERR_ARG, std::string("malformed ") + which + " tree at cell " + std::to_string(c));
      }
    };
    check_tree(ct, t_ints, "target");
    check_tree(cs, s_ints, "source");
    UVLM_HIP(hipSetDevice(device));
    UVLM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    Tree    tt, st;
    Scratch sc;
    Lists   lists;
    auto    upload = [&](Tree &t, int64_t C, const int64_t *ints, const double *reals) {
      std::vector<int32_t> nch(C), ch(C), nb(C);
      std::vector<double>  x(C), y(C), z(C), r(C);
      t.nleaf = 0;
      for (int64_t c = 0; c < C; c++) {
        nch[c] = (int32_t)ints[5 * c];
        nb[c]  = (int32_t)ints[5 * c + 1];
        ch[c]  = (int32_t)ints[5 * c + 2];
        x[c]   = reals[4 * c];
        y[c]   = reals[4 * c + 1];
        z[c]   = reals[4 * c + 2];
        r[c]   = reals[4 * c + 3];
        t.nleaf += nch[c] == 0;
      }
      t.C = C;
      for (DevBuf *b : {&t.nchild, &t.child, &t.nbody}) b->reserve(4 * C);
      for (DevBuf *b : {&t.cx, &t.cy, &t.cz, &t.R}) b->reserve(8 * C);
      UVLM_HIP(hipMemcpyAsync(t.nchild.p, nch.data(), 4 * C, hipMemcpyHostToDevice, s));
      UVLM_HIP(hipMemcpyAsync(t.child.p, ch.data(), 4 * C, hipMemcpyHostToDevice, s));
      UVLM_HIP(hipMemcpyAsync(t.nbody.p, nb.data(), 4 * C, hipMemcpyHostToDevice, s));
      UVLM_HIP(hipMemcpyAsync(t.cx.p, x.data(), 8 * C, hipMemcpyHostToDevice, s));
      UVLM_HIP(hipMemcpyAsync(t.cy.p, y.data(), 8 * C, hipMemcpyHostToDevice, s));
      UVLM_HIP(hipMemcpyAsync(t.cz.p, z.data(), 8 * C, hipMemcpyHostToDevice, s));
      UVLM_HIP(hipMemcpyAsync(t.R.p, r.data(), 8 * C, hipMemcpyHostToDevice, s));
      UVLM_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
    };
    upload(tt, ct, t_ints, t_reals);
    upload(st, cs, s_ints, s_reals);
    traverse(s, sc, tt, st, theta, lists);
    *n_p2p = lists.np2p;
    *n_m2l = lists.nm2l;
    auto expand = [&](DevBuf &off, DevBuf &src, int64_t n, int64_t cap, int64_t *out) {
      if (!out || n > cap) return;
      std::vector<int32_t> o(ct + 1), sv(std::max<int64_t>(n, 1));
      UVLM_HIP(hipMemcpyAsync(o.data(), off.p, 4 * (ct + 1), hipMemcpyDeviceToHost, s));
      if (n) UVLM_HIP(hipMemcpyAsync(sv.data(), src.p, 4 * n, hipMemcpyDeviceToHost, s));
      UVLM_HIP(hipStreamSynchronize(s));
      for (int64_t c = 0; c < ct; c++)
        for (int q = o[c]; q < o[c + 1]; q++) {
          out[2 * q]     = c;
          out[2 * q + 1] = sv[q];
        }
    };
    expand(lists.p2p_off, lists.p2p_src, lists.np2p, cap_p2p, p2p_out);
    expand(lists.m2l_off, lists.m2l_src, lists.nm2l, cap_m2l, m2l_out);
    UVLM_HIP(hipStreamSynchronize(s));
  });
  if (s) (void)hipStreamDestroy(s);
  return rc;
}

int uvlm_get_expansions(uvlm_handle h, double *M, double *L)
{
  return guarded([&] {
    check_handle(h);
    if (!h->evaluated) throw Error(UVLM_ERR_STATE, "no expansions before uvlm_eval");
    set_device(h);
    const size_t bytes = sizeof(double2) * h->NT * h->tree.C;
    if (M) UVLM_HIP(hipMemcpyAsync(M, h->M.p, bytes, hipMemcpyDeviceToHost, h->s));
    if (L) UVLM_HIP(hipMemcpyAsync(L, h->L.p, bytes, hipMemcpyDeviceToHost, h->s));
    UVLM_HIP(hipStreamSynchronize(h->s));
  });
}

int uvlm_sensors_aos(uvlm_handle h, void *probes, int64_t m, int64_t stride)
{
  return guarded([&] {
    check_handle(h);
    if (!h->evaluated) throw Error(UVLM_ERR_STATE, "uvlm_sensors_aos needs the multipoles of a uvlm_eval");
    if (m <= 0 || !probes) throw Error(UVLM_ERR_ARG, "empty or null probes");
    if (stride < 296 || stride % 8) throw Error(UVLM_ERR_ARG, "stride must be >= 296 and a multiple of 8");
    set_device(h);
    hipStream_t s = h->s;
    reserve_probes(h, m);
    const int64_t sc = h->scap;
    h->staging.reserve((size_t)m * stride);
    UVLM_HIP(hipMemcpyAsync(h->staging.p, probes, (size_t)m * stride, hipMemcpyHostToDevice, s));
    aos_to_soa(s, h->staging.p, m, stride / 8, h->sst.as<double>(), sc);
    UVLM_HIP(hipEventRecord(h->aev[0], s));
    sensor_pass(h, m);
    UVLM_HIP(hipEventRecord(h->aev[1], s));
    if (stride > 296) UVLM_HIP(hipMemcpyAsync(h->staging.p, probes, (size_t)m * stride, hipMemcpyHostToDevice, s));
    soa_to_aos(s, h->sst.as<double>(), sc, h->staging.p, m, stride / 8);
    UVLM_HIP(hipMemcpyAsync(probes, h->staging.p, (size_t)m * stride, hipMemcpyDeviceToHost, s));
    UVLM_HIP(hipStreamSynchronize(s));
    h->tm.sensor_ms = ev_ms2(h, 0, 1);
  });
}

int uvlm_probe_adjoint(uvlm_handle h, int64_t m, const double *probes, const double *weights, double *dJ_dalpha,
                       double *dJ_dx)
{
  return guarded([&] {
    check_handle(h);
    if (!h->built) throw Error(UVLM_ERR_STATE, "uvlm_probe_adjoint needs a uvlm_build");
    if (m <= 0 || !probes || !weights) throw Error(UVLM_ERR_ARG, "empty or null probes");
    set_device(h);
    hipStream_t s = h->s;
    reserve_probes(h, m);
    const int64_t sc = h->scap;
    // probes as sources: X = probe, alpha = weight, sigma = 1 (the kernel takes the target's)
    std::vector<double> host((size_t)N_FIELDS * sc, 0.0);
    for (int64_t i = 0; i < m; i++) {
      for (int d = 0; d < 3; d++) {
        host[(F_X + d) * sc + i] = probes[3 * i + d];
        host[(F_A + d) * sc + i] = weights[3 * i + d];
      }
      host[F_SIG * sc + i] = 1.0;
    }
    UVLM_HIP(hipMemcpyAsync(h->sst.p, host.data(), 8 * host.size(), hipMemcpyHostToDevice, s));
    UVLM_HIP(hipStreamSynchronize(s));
    UVLM_HIP(hipEventRecord(h->aev[0], s));
    probe_adjoint_core(h, m);
    UVLM_HIP(hipEventRecord(h->aev[1], s));
    const int64_t cap = h->cap, N = h->N;
    double       *a   = h->ast.as<double>();
    std::vector<double> buf((size_t)3 * N);
    auto fetch = [&](int f0, double *out) {
      if (!out) return;
      for (int d = 0; d < 3; d++)
        UVLM_HIP(hipMemcpyAsync(buf.data() + (size_t)d * N, a + (f0 + d) * cap, 8 * N, hipMemcpyDeviceToHost, s));
      UVLM_HIP(hipStreamSynchronize(s));
      for (int64_t i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) out[3 * i + d] = buf[(size_t)d * N + i];
    };
    fetch(F_U, dJ_dalpha);
    fetch(F_DADT, dJ_dx);
    UVLM_HIP(hipStreamSynchronize(s));
    h->tm.adjoint_ms = ev_ms2(h, 0, 1);
  });
}

int uvlm_autotune(uvlm_handle h, const int64_t *ncrit, int n, int reps, int64_t *best_ncrit, double *best_ms)
{
  return guarded([&] {
    check_handle(h);
    if (!ncrit || n <= 0 || reps <= 0) throw Error(UVLM_ERR_ARG, "empty candidate list or reps <= 0");
    for (int i = 0; i < n; i++)
      if (ncrit[i] < 1) throw Error(UVLM_ERR_ARG, "ncrit candidates must be >= 1");
    if (h->N <= 0) throw Error(UVLM_ERR_STATE, "no bodies set");
    set_device(h);
    const int64_t keep_ncrit = h->prm.ncrit;
    const int     keep_budget = h->sched.budget;
    double        best = 1e300;
    int64_t       bn = keep_ncrit;
    int           bb = keep_budget;
    hipEvent_t    e1 = h->ev[EV_DOWN];
    try {
      M2LPlan prev;
      bool    have_prev = false;
      for (int b : {4, 5, 6, 8}) {
        M2LPlan probe;
        if (!make_m2l_plan(h->prm.P, b, probe, 16, m2l_waves(h->prm.P, h->sched.waves))) continue;
        // skip budgets whose plan equals the previous one
        const bool same = have_prev && prev.S == probe.S && prev.slots == probe.slots && prev.lev == probe.lev;
        prev      = probe;
        have_prev = true;
        if (same) continue;
        h->sched.budget = b;
        make_m2l_schedule(h->prm.P, h->sched);
        for (int i = 0; i < n; i++) {
          h->prm.ncrit = ncrit[i];
          do_build(h);
          do_eval(h, true);
          double tot = 0.0;
          for (int r = 0; r < reps; r++) {
            do_build(h);
            do_eval(h, true);
            UVLM_HIP(hipEventSynchronize(e1));
            tot += ev_ms(h, EV_B0, EV_DOWN);
          }
          const double ms = tot / reps;
          if (ms < best) {
            best = ms;
            bn   = ncrit[i];
            bb   = b;
          }
        }
      }
    } catch (...) {
      h->prm.ncrit    = keep_ncrit;
      h->sched.budget = keep_budget;
      make_m2l_schedule(h->prm.P, h->sched);
      throw;
    }
    h->prm.ncrit    = bn;
    h->sched.budget = bb;
    make_m2l_schedule(h->prm.P, h->sched);
    h->built     = false;
    h->evaluated = false;
    if (best_ncrit) *best_ncrit = bn;
    if (best_ms) *best_ms = best;
  });
}

int uvlm_set_near_field(uvlm_handle h, double c)
{
  return guarded([&] {
    check_handle(h);
    if (!(c >= 0.0)) throw Error(UVLM_ERR_ARG, "the near-field factor must be >= 0");
    h->near_c = c;
    h->built = h->evaluated = false;
  });
}

int uvlm_set_viscosity(uvlm_handle h, int scheme, double nu)
{
  return guarded([&] {
    check_handle(h);
    if (scheme != UVLM_INVISCID && scheme != UVLM_PSE && scheme != UVLM_CSM) throw Error(UVLM_ERR_ARG, "bad viscous scheme");
    if (!(nu >= 0.0) || !std::isfinite(nu)) throw Error(UVLM_ERR_ARG, "nu must be finite and >= 0");
    h->viscous = scheme;
    h->nu      = nu;
  });
}

int uvlm_viscosity(uvlm_handle h, double dt)
{
  return guarded([&] {
    check_handle(h);
    if (!h->evaluated) throw Error(UVLM_ERR_STATE, "uvlm_viscosity needs the rates of a uvlm_eval");
    if (h->viscous == UVLM_CSM && !(dt > 0.0)) throw Error(UVLM_ERR_ARG, "CSM needs dt > 0");
    set_device(h);
    apply_viscosity(h, dt);
    UVLM_HIP(hipStreamSynchronize(h->s));
  });
}

int uvlm_vorticity_aos(uvlm_handle h, void *probes, int64_t m, int64_t stride)
{
  return guarded([&] {
    check_handle(h);
    if (!h->built) throw Error(UVLM_ERR_STATE, "uvlm_vorticity_aos needs a uvlm_build");
    if (m <= 0 || !probes) throw Error(UVLM_ERR_ARG, "empty or null probes");
    if (stride < 296 || stride % 8) throw Error(UVLM_ERR_ARG, "stride must be >= 296 and a multiple of 8");
    set_device(h);
    upload_probes(h, probes, m, stride);
    probe_vorticity(h, m);
    copy_out_records(download_probes(h, m), m, probes, stride);
  });
}

int uvlm_rbf_aos(uvlm_handle h, void *points, int64_t m, int64_t stride, int64_t *kept, int *iterations,
                 double *residuals)
{
  return guarded([&] {
    check_handle(h);
    if (!h->built) throw Error(UVLM_ERR_STATE, "uvlm_rbf_aos needs a uvlm_build of the particles");
    if (m <= 0 || !points || !kept) throw Error(UVLM_ERR_ARG, "empty or null points");
    if (stride < 296 || stride % 8) throw Error(UVLM_ERR_ARG, "stride must be >= 296 and a multiple of 8");
    set_device(h);
    hipStream_t s = h->s;
    // target vorticity at the new points, then drop those below 1e-6 in every component
    upload_probes(h, points, m, stride);
    probe_vorticity(h, m);
    const std::vector<double> all = download_probes(h, m);
    std::vector<double>       keep;
    for (int64_t i = 0; i < m; i++) {
      const double *r = all.data() + i * N_FIELDS;
      if (!(std::abs(r[F_UOLD]) < 1e-6 && std::abs(r[F_UOLD + 1]) < 1e-6 && std::abs(r[F_UOLD + 2]) < 1e-6))
        keep.insert(keep.end(), r, r + N_FIELDS);
    }
    const int64_t k = (int64_t)(keep.size() / N_FIELDS);
    *kept           = k;
    if (iterations) *iterations = 0;
    if (residuals) residuals[0] = residuals[1] = residuals[2] = 0.0;
    if (k == 0) return;
    // tree over the kept points and its self lists (createSensorTree(sensorOnly), 164-201)
    upload_probes(h, keep.data(), k, 296);
    traverse(s, h->sc, h->stree, h->stree, h->prm.theta, h->slists, h->gap);
    double       *st = h->sst.as<double>();
    const int64_t sc = h->scap;
    h->sc.partial.reserve(sizeof(double) * 3 * 512);
    double *part = h->sc.partial.as<double>();
    auto single = [&] {  // calculateVorticitySingleTree (788-795): velocity = -A alpha
      zero_fields(s, st, sc, k, F_U, 3);
      zero_fields(s, st, sc, k, F_DADT, 3);
      gauss_near(s, h->stree, h->stree, h->slists, st, sc, st, sc, st, sc, 1, F_U, 0.0);
    };
    const double tol = 1e-10, eps = 2.220446049250313e-16;
    rbf_vec(s, st, sc, k, 0, nullptr);  // initAlphaValues
    single();
    rbf_vec(s, st, sc, k, 1, nullptr);  // setInitialResidualAndUpdateAlpha
    double res[3];
    dot3(s, st, sc, k, F_XOLD, F_XOLD, part, res);
    for (int it = 0; it < 100; it++) {
      single();
      double pAp[3], a[3], nr[3], beta[3];
      dot3(s, st, sc, k, F_A, F_U, part, pAp);
      for (int d = 0; d < 3; d++) a[d] = (res[d] > eps && std::abs(pAp[d]) > eps) ? res[d] / pAp[d] : 0.0;
      rbf_vec(s, st, sc, k, 2, a);
      dot3(s, st, sc, k, F_XOLD, F_XOLD, part, nr);
      if (iterations) *iterations = it + 1;
      if (residuals)
        for (int d = 0; d < 3; d++) residuals[d] = nr[d];
      if (nr[0] < tol && nr[1] < tol && nr[2] < tol) break;
      for (int d = 0; d < 3; d++) beta[d] = (nr[d] > eps && res[d] > eps) ? nr[d] / res[d] : 0.0;
      for (int d = 0; d < 3; d++) res[d] = nr[d];
      rbf_vec(s, st, sc, k, 3, beta);
    }
    rbf_vec(s, st, sc, k, 4, nullptr);  // storeAlphaValues
    copy_out_records(download_probes(h, k), k, points, stride);
  });
}

int uvlm_relax(uvlm_handle h, double f)
{
  return guarded([&] {
    check_handle(h);
    if (!h->evaluated) throw Error(UVLM_ERR_STATE, "uvlm_relax needs the J of a uvlm_eval");
    if (!(f >= 0.0 && f <= 1.0)) throw Error(UVLM_ERR_ARG, "relaxation factor must be in [0, 1]");
    set_device(h);
    relax(h->s, h->st.as<double>(), h->cap, h->N, f);
    UVLM_HIP(hipStreamSynchronize(h->s));
  });
}

int uvlm_diagnostics(uvlm_handle h, double *out)
{
  return guarded([&] {
    check_handle(h);
    if (!out) throw Error(UVLM_ERR_ARG, "null output");
    if (h->N <= 0) throw Error(UVLM_ERR_STATE, "no bodies set");
    set_device(h);
    h->sc.partial.reserve(sizeof(double) * 10 * 512);
    diagnostics(h->s, h->st.as<double>(), h->cap, h->N, h->sc.partial.as<double>(), out);
  });
}

int uvlm_eval_vjp(uvlm_handle h, int64_t n, const double *ubar, const double *abar, double *xbar, double *alphabar)
{
  return guarded([&] {
    check_handle(h);
    if (!h->built) throw Error(UVLM_ERR_STATE, "uvlm_eval_vjp needs a uvlm_build");
    if (n != h->N) throw Error(UVLM_ERR_ARG, "size mismatch");
    check_vjp_config(h);
    set_device(h);
    const int64_t cap = h->cap;
    do_eval(h, true);
    h->vaux.reserve(sizeof(double) * N_ADJ_FIELDS * cap);
    h->vlam.reserve(sizeof(double) * 6 * cap);
    upload_cot(h, ubar, abar, h->vlam.as<double>());
    double *aux = h->vaux.as<double>(), *lam = h->vlam.as<double>();
    UVLM_HIP(hipMemcpyAsync(aux, lam, 8 * 3 * cap, hipMemcpyDeviceToDevice, h->s));
    UVLM_HIP(hipMemcpyAsync(aux + 15 * cap, lam + 3 * cap, 8 * 3 * cap, hipMemcpyDeviceToDevice, h->s));
    UVLM_HIP(hipEventRecord(h->aev[0], h->s));
    vjp_core(h);
    UVLM_HIP(hipEventRecord(h->aev[1], h->s));
    UVLM_HIP(hipMemcpyAsync(lam, aux + 12 * cap, 8 * 6 * cap, hipMemcpyDeviceToDevice, h->s));
    download_cot(h, lam, xbar, alphabar);
    h->tm.vjp_ms = ev_ms2(h, 0, 1);
  });
}

int uvlm_step_vjp(uvlm_handle h, double dt, int64_t n, double *xbar, double *alphabar)
{
  return guarded([&] {
    check_handle(h);
    if (n != h->N || !xbar || !alphabar) throw Error(UVLM_ERR_ARG, "size mismatch or null cotangent");
    check_vjp_config(h);
    set_device(h);
    h->vlam.reserve(sizeof(double) * 6 * h->cap);
    upload_cot(h, xbar, alphabar, h->vlam.as<double>());
    step_vjp(h, dt, h->vlam.as<double>());
    download_cot(h, h->vlam.as<double>(), xbar, alphabar);
  });
}

int uvlm_trajectory_forward(uvlm_handle h, int64_t steps, double dt, int64_t every)
{
  return guarded([&] {
    check_handle(h);
    if (h->N <= 0) throw Error(UVLM_ERR_STATE, "no bodies set");
    if (steps <= 0 || every <= 0) throw Error(UVLM_ERR_ARG, "steps and every must be positive");
    check_vjp_config(h);
    set_device(h);
    hipStream_t   s = h->s;
    const int64_t N = h->N, cap = h->cap, nck = (steps + every - 1) / every;
    h->ck.reserve(sizeof(double) * 7 * N * nck);
    double *ck = h->ck.as<double>();
    for (int64_t k = 0; k < steps; k++) {
      if (k % every == 0) {
        // checkpoint in the original order: scatter positions/strengths by the original index
        save_state_orig(h, ck + (k / every) * 7 * N);
      }
      for (int st = 0; st < stages_of(h->prm.scheme); st++) {
        do_build(h);
        do_eval(h, false);
        rk_stage(s, h->st.as<double>(), cap, N, h->prm.scheme, st, dt, 0);
      }
    }
    UVLM_HIP(hipStreamSynchronize(s));
    h->evaluated = false;
    h->ck_n      = nck;
    h->ck_steps  = steps;
    h->ck_every  = every;
    h->ck_N      = N;
    h->ck_dt     = dt;
  });
}

int uvlm_trajectory_vjp(uvlm_handle h, int64_t n, double *xbar, double *alphabar)
{
  return guarded([&] {
    check_handle(h);
    if (h->ck_n == 0 || h->ck_N != h->N) throw Error(UVLM_ERR_STATE, "uvlm_trajectory_vjp needs a uvlm_trajectory_forward");
    if (n != h->N || !xbar || !alphabar) throw Error(UVLM_ERR_ARG, "size mismatch or null cotangent");
    check_vjp_config(h);
    set_device(h);
    hipStream_t   s = h->s;
    const int64_t N = h->N, cap = h->cap, E = h->ck_every;
    const double  dt = h->ck_dt;
    // final state (original order) kept to restore it at the end
    DevBuf fin, seg;
    fin.reserve(sizeof(double) * 7 * N);
    save_state_orig(h, fin.as<double>());
    seg.reserve(sizeof(double) * 7 * N * E);
    h->vlam.reserve(sizeof(double) * 6 * cap);
    upload_cot(h, xbar, alphabar, h->vlam.as<double>());
    for (int64_t c = h->ck_n - 1; c >= 0; c--) {
      const int64_t k0 = c * E, k1 = std::min(h->ck_steps, k0 + E);
      // recompute the segment's step starts from its checkpoint
      load_state_orig(h, h->ck.as<double>() + c * 7 * N);
      for (int64_t k = k0; k < k1; k++) {
        save_state_orig(h, seg.as<double>() + (k - k0) * 7 * N);
        if (k + 1 == k1) break;
        for (int st = 0; st < stages_of(h->prm.scheme); st++) {
          do_build(h);
          do_eval(h, false);
          rk_stage(s, h->st.as<double>(), cap, N, h->prm.scheme, st, dt, 0);
        }
      }
      for (int64_t k = k1 - 1; k >= k0; k--) {
        load_state_orig(h, seg.as<double>() + (k - k0) * 7 * N);
        step_vjp(h, dt, h->vlam.as<double>());
      }
    }
    download_cot(h, h->vlam.as<double>(), xbar, alphabar);
    load_state_orig(h, fin.as<double>());
  });
}

int uvlm_append_bodies_soa(uvlm_handle h, int64_t n, const double *X, const double *alpha, const double *sigma)
{
  return guarded([&] {
    check_handle(h);
    if (n <= 0 || !X || !alpha || !sigma) throw Error(UVLM_ERR_ARG, "null or empty input");
    set_device(h);
    DevBuf buf;
    buf.reserve(sizeof(double) * 7 * n);
    double *d = buf.as<double>();
    UVLM_HIP(hipMemcpyAsync(d, X, 24 * n, hipMemcpyHostToDevice, h->s));
    UVLM_HIP(hipMemcpyAsync(d + 3 * n, alpha, 24 * n, hipMemcpyHostToDevice, h->s));
    UVLM_HIP(hipMemcpyAsync(d + 6 * n, sigma, 8 * n, hipMemcpyHostToDevice, h->s));
    append_bodies_dev(h, n, d, d + 3 * n, d + 6 * n);
    UVLM_HIP(hipStreamSynchronize(h->s));
  });
}

int uvlm_wing_create(int device, int64_t nspan, int64_t nchord, const double *vertices, uvlm_wing *out)
{
  uvlm_wing w = nullptr;
  const int rc = guarded([&] {
    if (!out || !vertices) throw Error(UVLM_ERR_ARG, "null argument");
    if (nspan < 1 || nchord < 1 || nspan * nchord > UVLM_WING_MAX_PANELS)
      throw Error(UVLM_ERR_ARG, "the lattice needs 1 <= nspan * nchord <= " + std::to_string(UVLM_WING_MAX_PANELS) +
                                    " panels");
    UVLM_HIP(hipSetDevice(device));
    w         = new uvlm_wing_s;
    w->device = device;
    w->nspan  = (int)nspan;
    w->nchord = (int)nchord;
    w->n      = (int)(nspan * nchord);
    UVLM_HIP(hipStreamCreate(&w->s));
    const int     n  = w->n;
    const int64_t nv = (nspan + 1) * (nchord + 1);
    w->V.reserve(24 * nv);
    w->cp.reserve(24 * n);
    w->nrm.reserve(24 * n);
    w->area.reserve(8 * n);
    w->B.reserve(24 * (size_t)n * n);
    w->LU.reserve(8 * (size_t)n * n);
    w->Ainv.reserve(8 * (size_t)n * n);
    w->rhs.reserve(8 * n);
    w->u.reserve(24 * n);
    w->piv.reserve(4 * n);
    w->info.reserve(8);
    w->gamma.reserve(8 * n);
    w->v.reserve(24 * n);
    w->F.reserve(24);
    w->shedX.reserve(24 * nspan);
    w->shedA.reserve(24 * nspan);
    w->shedS.reserve(8 * nspan);
    UVLM_HIP(hipMemcpyAsync(w->V.p, vertices, 24 * nv, hipMemcpyHostToDevice, w->s));
    // geometry, ring influences at the control points, system matrix, LU (main.cpp:190-214)
    wing_geom(w->s, n, w->nspan, w->V.as<double>(), w->cp.as<double>(), w->nrm.as<double>(), w->area.as<double>());
    int *flags = w->info.as<int>();  // [0] singular segments, [1] LU info
    UVLM_HIP(hipMemsetAsync(flags, 0, 8, w->s));
    wing_influence(w->s, n, n, w->nspan, w->V.as<double>(), w->cp.as<double>(), w->nrm.as<double>(),
                   w->B.as<double>(), w->LU.as<double>(), flags);
    lu_factor(w->s, n, w->LU.as<double>(), w->piv.as<int>(), flags + 1);
    int hf[2] = {0, 0};
    UVLM_HIP(hipMemcpyAsync(hf, flags, 8, hipMemcpyDeviceToHost, w->s));
    UVLM_HIP(hipStreamSynchronize(w->s));
    if (hf[0]) throw Error(UVLM_ERR_ARG, "a control point on a vortex line or a zero-length vortex line (main.cpp:19-21)");
    if (hf[1]) throw Error(UVLM_ERR_ARG, "singular system matrix (zero pivot at column " + std::to_string(hf[1] - 1) + ")");
    // the matrix is fixed for the wing (rigid motion only): its inverse turns every solve of
    // the loop, and of the adjoint, into one product
    lu_inverse(w->s, n, w->LU.as<double>(), w->piv.as<int>(), w->Ainv.as<double>());
    UVLM_HIP(hipStreamSynchronize(w->s));
    *out = w;
    w    = nullptr;
  });
  if (w) {
    if (w->s) (void)hipStreamDestroy(w->s);
    delete w;
  }
  return rc;
}

int uvlm_wing_destroy(uvlm_wing w)
{
  if (!w) return UVLM_OK;
  (void)hipSetDevice(w->device);
  if (w->s) (void)hipStreamSynchronize(w->s);
  if (w->s) (void)hipStreamDestroy(w->s);
  delete w;
  return UVLM_OK;
}

int uvlm_wing_get(uvlm_wing w, double *cp, double *normal, double *area, double *aic_lu)
{
  return guarded([&] {
    if (!w) throw Error(UVLM_ERR_ARG, "null wing");
    UVLM_HIP(hipSetDevice(w->device));
    const int n = w->n;
    if (cp) UVLM_HIP(hipMemcpyAsync(cp, w->cp.p, 24 * n, hipMemcpyDeviceToHost, w->s));
    if (normal) UVLM_HIP(hipMemcpyAsync(normal, w->nrm.p, 24 * n, hipMemcpyDeviceToHost, w->s));
    if (area) UVLM_HIP(hipMemcpyAsync(area, w->area.p, 8 * n, hipMemcpyDeviceToHost, w->s));
    if (aic_lu) UVLM_HIP(hipMemcpyAsync(aic_lu, w->LU.p, 8 * (size_t)n * n, hipMemcpyDeviceToHost, w->s));
    UVLM_HIP(hipStreamSynchronize(w->s));
  });
}

int uvlm_wing_translate(uvlm_wing w, const double d[3])
{
  return guarded([&] {
    if (!w || !d) throw Error(UVLM_ERR_ARG, "null argument");
    UVLM_HIP(hipSetDevice(w->device));
    translate3(w->s, (int64_t)(w->nspan + 1) * (w->nchord + 1), w->V.as<double>(), d);
    translate3(w->s, w->n, w->cp.as<double>(), d);
    UVLM_HIP(hipStreamSynchronize(w->s));
  });
}

namespace {
void check_pair(uvlm_wing w, uvlm_handle h)
{
  if (!w) throw Error(UVLM_ERR_ARG, "null wing");
  if (h && h->prm.device != w->device) throw Error(UVLM_ERR_ARG, "wing and particles on different devices");
  if (h) resolve_outputs(h);
}

// main.cpp:196-244 at the wing's current position: particle velocity at the control points,
// right-hand side, circulations, force.
// Jcp (optional, [n][9]): also the gradient of the particle velocity at the control points.
void wing_solve(uvlm_wing w, uvlm_handle h, const double U[3], double *Jcp = nullptr)
{
  const int n    = w->n;
  double   *vptr = nullptr;
  if (h && h->N > 0) {
    if (Jcp) velocity_grad_at_points(h, n, w->cp.as<double>(), w->v.as<double>(), Jcp);
    else velocity_at_points(h, n, w->cp.as<double>(), w->v.as<double>());
    UVLM_HIP(hipStreamSynchronize(h->s));
    vptr = w->v.as<double>();
  }
  wing_rhs(w->s, n, w->nrm.as<double>(), U, vptr, w->rhs.as<double>());
  matvec_inv(w->s, n, w->Ainv.as<double>(), w->rhs.as<double>(), w->gamma.as<double>());
  wing_velocity(w->s, n, w->B.as<double>(), w->gamma.as<double>(), U, vptr, w->u.as<double>());
  wing_force(w->s, n, w->nspan, w->V.as<double>(), w->u.as<double>(), w->gamma.as<double>(), w->F.as<double>());
}

void wing_outputs(uvlm_wing w, double *gamma, double *force)
{
  if (gamma) UVLM_HIP(hipMemcpyAsync(gamma, w->gamma.p, 8 * w->n, hipMemcpyDeviceToHost, w->s));
  if (force) UVLM_HIP(hipMemcpyAsync(force, w->F.p, 24, hipMemcpyDeviceToHost, w->s));
  UVLM_HIP(hipStreamSynchronize(w->s));
}

// main.cpp:246-289: one particle per trailing-edge panel, appended to the particles.
void wing_shed_into(uvlm_wing w, uvlm_handle h, const double U[3], double dt, int ref_offset)
{
  wing_shed(w->s, w->nspan, w->nchord, w->V.as<double>(), w->gamma.as<double>(), U, dt, ref_offset,
            w->shedX.as<double>(), w->shedA.as<double>(), w->shedS.as<double>());
  UVLM_HIP(hipStreamSynchronize(w->s));
  append_bodies_dev(h, w->nspan, w->shedX.as<double>(), w->shedA.as<double>(), w->shedS.as<double>());
  UVLM_HIP(hipStreamSynchronize(h->s));
}
}  // namespace

int uvlm_wing_solve(uvlm_wing w, uvlm_handle particles, const double U[3], double *gamma, double force[3])
{
  return guarded([&] {
    check_pair(w, particles);
    if (!U) throw Error(UVLM_ERR_ARG, "null freestream");
    UVLM_HIP(hipSetDevice(w->device));
    wing_solve(w, particles, U);
    wing_outputs(w, gamma, force);
  });
}

int uvlm_wing_shed(uvlm_wing w, uvlm_handle particles, const double U[3], double dt, int reference_offset)
{
  return guarded([&] {
    check_pair(w, particles);
    if (!particles || !U) throw Error(UVLM_ERR_ARG, "null particles or freestream");
    UVLM_HIP(hipSetDevice(w->device));
    wing_shed_into(w, particles, U, dt, reference_offset);
  });
}

int uvlm_wing_step(uvlm_wing w, uvlm_handle particles, const double U[3], double dt, int reference_offset,
                   double *gamma, double force[3])
{
  return guarded([&] {
    check_pair(w, particles);
    if (!particles || !U) throw Error(UVLM_ERR_ARG, "null particles or freestream");
    UVLM_HIP(hipSetDevice(w->device));
    uvlm_handle h = particles;
    wing_solve(w, h, U);
    wing_outputs(w, gamma, force);
    wing_shed_into(w, h, U, dt, reference_offset);
    for (int k = 0; k < stages_of(h->prm.scheme); k++) {
      do_build(h);
      do_eval(h, false);
      apply_viscosity(h, dt);
      rk_stage(h->s, h->st.as<double>(), h->cap, h->N, h->prm.scheme, k, dt, 0);
    }
    UVLM_HIP(hipStreamSynchronize(h->s));
    h->evaluated = false;
    const double d[3] = {-U[0] * dt, -U[1] * dt, -U[2] * dt};
    translate3(w->s, (int64_t)(w->nspan + 1) * (w->nchord + 1), w->V.as<double>(), d);
    translate3(w->s, w->n, w->cp.as<double>(), d);
    UVLM_HIP(hipStreamSynchronize(w->s));
  });
}

namespace {
// The wing's position at one step of a window: vertices and control points (the loop moves both
// by -U dt per step; the geometry that does not depend on the position is kept once).
struct WingPose {
  DevBuf V, cp;
};
void save_pose(uvlm_wing w, WingPose &p)
{
  const int64_t nv = (int64_t)(w->nspan + 1) * (w->nchord + 1);
  p.V.reserve(24 * nv);
  p.cp.reserve(24 * w->n);
  UVLM_HIP(hipMemcpyAsync(p.V.p, w->V.p, 24 * nv, hipMemcpyDeviceToDevice, w->s));
  UVLM_HIP(hipMemcpyAsync(p.cp.p, w->cp.p, 24 * w->n, hipMemcpyDeviceToDevice, w->s));
}
void load_pose(uvlm_wing w, const WingPose &p)
{
  const int64_t nv = (int64_t)(w->nspan + 1) * (w->nchord + 1);
  UVLM_HIP(hipMemcpyAsync(w->V.p, p.V.p, 24 * nv, hipMemcpyDeviceToDevice, w->s));
  UVLM_HIP(hipMemcpyAsync(w->cp.p, p.cp.p, 24 * w->n, hipMemcpyDeviceToDevice, w->s));
}

// Adjoint of CL = e . F / (|U|^2 S / 2) at the end of `steps` passes of the loop (solve, shed,
// RK stages, move the wing) from the current particles and wing position.  Outputs (each may be
// null): CL; the cotangents of the particles' X and alpha at the start (original order); the
// cotangent of the wing's vertices (vbar [nv][3]: the lattice moves rigidly, so one set of
// vertices fixes every step's geometry).  The particle state and the wing position are kept
// every `every` steps; each segment between two checkpoints is recomputed from its checkpoint
// (bit for bit the forward's operations) with the per-step states it needs, then run backwards.
// Backwards, each solve contributes through A^{-T} to the circulations' cotangent, to the
// accumulated cotangents of A (-rbar gamma^T) and of the normals (rhs), and through the wake
// velocity v at the control points to the particles (probe adjoint) and to the control points
// (J^T vbar, J the velocity gradient there); each shedding contributes through the shed
// particles' cotangents.  At the end the accumulated cotangents go through the ring influences
// (segment VJPs), the panel geometry (control point, normal, area) and the force's segments
// to the vertices.  The trees are held fixed per evaluation, as in the other adjoints.
void wing_window_adjoint(uvlm_wing w, uvlm_handle h, const double U[3], double dt, int64_t steps, int64_t every,
                         int ref_offset, const double e[3], double *CL, double *xbar, double *alphabar,
                         double *vertex_bar)
{
  hipStream_t   s = h->s, ws = w->s;
  const int     np = w->n, ns = w->nspan, nc = w->nchord;
  const int64_t nv = (int64_t)(ns + 1) * (nc + 1);
  const int64_t N0 = h->N;
  const bool    geom = vertex_bar != nullptr;
  const int64_t E = std::max<int64_t>(1, every), nck = (steps + E - 1) / E;
  // particle capacity for the whole window first, so the cotangent stride stays fixed
  grow_particles(h, N0 + steps * ns);
  const int64_t cap = h->cap;
  auto          keep = [&](DevBuf &b, int64_t &cnt) {
    cnt = h->N;
    b.reserve(sizeof(double) * 7 * std::max<int64_t>(cnt, 1));
    if (cnt) save_state_orig(h, b.as<double>());
  };
  auto restore = [&](const DevBuf &b, int64_t cnt) {
    if (cnt) load_state_orig_n(h, b.as<double>(), cnt, cnt);
    else h->N = 0;
  };
  std::vector<int64_t> nstart(steps + 1, 0);
  // per-step record of a solve: circulations [np], wake velocity [3 np] and (design gradient)
  // its gradient [9 np] at the control points
  const int64_t RS = geom ? 13 * np : 4 * np;
  // one pass of the loop from the current state; rec receives the solve's record, yshed the
  // state right after the shedding
  auto loop_step = [&](int64_t k, double *rec, DevBuf *yshed) {
    nstart[k] = h->N;
    double *gam = rec, *vel = rec ? rec + np : nullptr, *Jr = (rec && geom) ? rec + 4 * np : nullptr;
    wing_solve(w, h, U, Jr);
    if (Jr && h->N == 0) zero_doubles(ws, Jr, 9 * np);
    if (gam) UVLM_HIP(hipMemcpyAsync(gam, w->gamma.p, 8 * np, hipMemcpyDeviceToDevice, ws));
    if (vel) {
      if (h->N > 0) UVLM_HIP(hipMemcpyAsync(vel, w->v.p, 24 * np, hipMemcpyDeviceToDevice, ws));
      else zero_doubles(ws, vel, 3 * np);
    }
    UVLM_HIP(hipStreamSynchronize(ws));
    wing_shed_into(w, h, U, dt, ref_offset);
    if (yshed) {
      int64_t cnt = 0;
      keep(*yshed, cnt);
    }
    for (int st = 0; st < stages_of(h->prm.scheme); st++) {
      do_build(h);
      do_eval(h, false);
      rk_stage(s, h->st.as<double>(), cap, h->N, h->prm.scheme, st, dt, 0);
    }
    const double d[3] = {-U[0] * dt, -U[1] * dt, -U[2] * dt};
    translate3(ws, nv, w->V.as<double>(), d);
    translate3(ws, np, w->cp.as<double>(), d);
    UVLM_HIP(hipStreamSynchronize(ws));
    UVLM_HIP(hipStreamSynchronize(s));
  };
  // forward: checkpoints (restored in the original order, as the recomputation starts from them);
  // the last segment keeps its per-step records right away, so it is not recomputed
  std::vector<DevBuf>   ck(nck);
  std::vector<int64_t>  ckn(nck);
  std::vector<WingPose> ckpose(nck);
  std::vector<DevBuf>   yshed(E);
  std::vector<WingPose> pose(E);
  DevBuf                segrec;
  segrec.reserve(sizeof(double) * RS * E);
  const int64_t last0 = (nck - 1) * E;
  for (int64_t k = 0; k < steps; k++) {
    if (k % E == 0) {
      keep(ck[k / E], ckn[k / E]);
      restore(ck[k / E], ckn[k / E]);
      save_pose(w, ckpose[k / E]);
    }
    if (k >= last0) {
      save_pose(w, pose[k - last0]);
      loop_step(k, segrec.as<double>() + RS * (k - last0), &yshed[k - last0]);
    } else {
      loop_step(k, nullptr, nullptr);
    }
  }
  nstart[steps] = h->N;
  DevBuf  yfinal;
  int64_t nfinal = 0;
  keep(yfinal, nfinal);
  WingPose final_pose;
  save_pose(w, final_pose);
  // the lift coefficient at the final position
  DevBuf fK;  // the final solve's record
  fK.reserve(sizeof(double) * RS);
  wing_solve(w, h, U, (geom && h->N > 0) ? fK.as<double>() + 4 * np : nullptr);
  std::vector<double> area(np);
  double              F[3];
  UVLM_HIP(hipMemcpyAsync(area.data(), w->area.p, 8 * np, hipMemcpyDeviceToHost, ws));
  UVLM_HIP(hipMemcpyAsync(F, w->F.p, 24, hipMemcpyDeviceToHost, ws));
  UVLM_HIP(hipStreamSynchronize(ws));
  double S = 0.0;
  for (double a : area) S += a;
  const double q     = 0.5 * (U[0] * U[0] + U[1] * U[1] + U[2] * U[2]) * S;
  const double es[3] = {e[0] / q, e[1] / q, e[2] / q};
  const double cl    = (e[0] * F[0] + e[1] * F[1] + e[2] * F[2]) / q;
  if (CL) *CL = cl;
  // backward
  DevBuf abuf;
  abuf.reserve(sizeof(double) * (size_t)(np + 3 * np + np + np));
  double *c = abuf.as<double>(), *vbar = c + np, *gbar = vbar + 3 * np, *rbar = gbar + np;
  h->vlam.reserve(sizeof(double) * 7 * cap);
  double *lam = h->vlam.as<double>();
  zero_doubles(s, lam, 7 * cap);
  // the particles' core-size cotangent (original order): the shed core sizes follow the geometry
  double *sbar = geom ? lam + 6 * cap : nullptr;
  // geometry accumulators: A's cotangent [np][np], normals, control points, shed corners,
  // and (final solve) the force's velocity cotangent, circulations and segment cotangents
  DevBuf gacc;
  double *Abar = nullptr, *nacc = nullptr, *cpacc = nullptr, *pbshed = nullptr, *ubfin = nullptr, *gfin = nullptr,
         *dlb = nullptr, *pbring = nullptr, *pbgeom = nullptr, *cpb2 = nullptr, *nb2 = nullptr, *vbdev = nullptr;
  if (geom) {
    gacc.reserve(sizeof(double) * ((size_t)np * np + 3 * np + 3 * np + 12 * np + 3 * np + np + 3 * np + 12 * np +
                                   12 * np + 3 * np + 3 * np + 3 * nv));
    Abar   = gacc.as<double>();
    nacc   = Abar + (size_t)np * np;
    cpacc  = nacc + 3 * np;
    pbshed = cpacc + 3 * np;
    ubfin  = pbshed + 12 * np;
    gfin   = ubfin + 3 * np;
    dlb    = gfin + np;
    pbring = dlb + 3 * np;
    pbgeom = pbring + 12 * np;
    cpb2   = pbgeom + 12 * np;
    nb2    = cpb2 + 3 * np;
    vbdev  = nb2 + 3 * np;
    zero_doubles(ws, Abar, (int64_t)np * np + 3 * np + 3 * np + 12 * np);
  }
  wing_force_vjp(ws, np, ns, w->V.as<double>(), w->B.as<double>(), w->u.as<double>(), w->gamma.as<double>(), es, c,
                 vbar, gbar);
  if (geom) {
    UVLM_HIP(hipMemcpyAsync(ubfin, vbar, 24 * np, hipMemcpyDeviceToDevice, ws));
    UVLM_HIP(hipMemcpyAsync(gfin, w->gamma.p, 8 * np, hipMemcpyDeviceToDevice, ws));
    wing_force_dl_vjp(ws, np, ns, w->u.as<double>(), w->gamma.as<double>(), es, dlb);
  }
  // the solve of one step (circulations gam, wake velocity vel at the control points, npart
  // particles on the device in that step's state, wing at that step's pose): gbar -> rbar =
  // A^{-T} gbar -> the geometry accumulators, vbar -= rbar n, then the wake
  auto through_solve = [&](const double *rec, int64_t npart) {
    const double *gam = rec, *vel = rec + np, *Jr = rec + 4 * np;
    matvec_inv_t(ws, np, w->Ainv.as<double>(), gbar, rbar);
    if (geom) wing_solve_geom_acc(ws, np, rbar, gam, U, npart ? vel : nullptr, Abar, nacc);
    wing_rhs_vjp(ws, np, w->nrm.as<double>(), rbar, vbar);
    if (geom && npart) wing_cp_acc(ws, np, vbar, Jr, cpacc);
    UVLM_HIP(hipStreamSynchronize(ws));
    if (npart == 0) return;
    do_build(h);
    reserve_probes(h, np);
    probe_sources(s, np, w->cp.as<double>(), vbar, h->sst.as<double>(), h->scap);
    probe_adjoint_core(h, np);
    cot_add_probe(s, h->N, h->orig.as<int32_t>(), h->ast.as<double>(), cap, lam, cap);
    if (sbar) {
      // the near field of the wake velocity at the control points: probes (weights vbar) as targets
      h->vsig.reserve(sizeof(double) * cap);
      const int64_t sc = h->scap;
      sigma_bar(s, h->tree, h->stree, h->alists, h->sst.as<double>(), sc, h->sst.as<double>() + F_A * sc, sc, false,
                h->st.as<double>(), cap, h->vsig.as<double>());
      add_to_orig(s, h->N, h->orig.as<int32_t>(), h->vsig.as<double>(), sbar);
    }
    UVLM_HIP(hipStreamSynchronize(s));
  };
  UVLM_HIP(hipMemcpyAsync(fK.p, w->gamma.p, 8 * np, hipMemcpyDeviceToDevice, ws));
  if (nfinal > 0) UVLM_HIP(hipMemcpyAsync(fK.as<double>() + np, w->v.p, 24 * np, hipMemcpyDeviceToDevice, ws));
  UVLM_HIP(hipStreamSynchronize(ws));
  through_solve(fK.as<double>(), nstart[steps]);
  // segments backwards: recompute each (the last was recorded by the forward pass), then its
  // steps in reverse
  for (int64_t g = nck - 1; g >= 0; g--) {
    const int64_t k0 = g * E, k1 = std::min(steps, k0 + E);
    if (g != nck - 1) {
      restore(ck[g], ckn[g]);
      load_pose(w, ckpose[g]);
      UVLM_HIP(hipStreamSynchronize(ws));
      for (int64_t k = k0; k < k1; k++) {
        save_pose(w, pose[k - k0]);
        loop_step(k, segrec.as<double>() + RS * (k - k0), &yshed[k - k0]);
      }
    }
    for (int64_t k = k1 - 1; k >= k0; k--) {
      const double *rec = segrec.as<double>() + RS * (k - k0), *gam = rec;
      const int64_t nshed = nstart[k] + ns;
      load_pose(w, pose[k - k0]);
      UVLM_HIP(hipStreamSynchronize(ws));
      restore(yshed[k - k0], nshed);
      step_vjp(h, dt, lam, sbar);
      UVLM_HIP(hipStreamSynchronize(s));
      // the shed particles' cotangents are entries nstart[k] .. nstart[k] + ns - 1
      wing_shed_vjp(ws, ns, nc, w->V.as<double>(), lam + 3 * cap + nstart[k], cap, gbar);
      if (geom) wing_shed_geom_vjp(ws, ns, nc, w->V.as<double>(), gam, lam, cap, nstart[k], sbar, pbshed);
      zero_doubles(ws, vbar, 3 * np);
      UVLM_HIP(hipStreamSynchronize(ws));
      if (nstart[k]) load_state_orig_n(h, yshed[k - k0].as<double>(), nshed, nstart[k]);
      else h->N = 0;
      through_solve(rec, nstart[k]);
    }
  }
  h->N = N0;
  download_cot(h, lam, xbar, alphabar);
  if (geom) {
    // the accumulated cotangents through the ring influences, the panel geometry and the
    // force's segments to the vertices (any pose: the influences are translation invariant)
    const double areab = -cl / S;
    wing_influence_vjp(ws, np, ns, w->V.as<double>(), w->cp.as<double>(), w->nrm.as<double>(), w->B.as<double>(), Abar,
                       ubfin, gfin, pbring, cpb2, nb2);
    wing_geom_vjp(ws, np, ns, w->V.as<double>(), cpacc, cpb2, nacc, nb2, areab, dlb, pbgeom);
    wing_vertex_gather(ws, ns, nc, pbgeom, pbring, pbshed, vbdev);
    UVLM_HIP(hipMemcpyAsync(vertex_bar, vbdev, 24 * nv, hipMemcpyDeviceToHost, ws));
    UVLM_HIP(hipStreamSynchronize(ws));
  }
  // particles and wing back at the final state
  load_pose(w, final_pose);
  UVLM_HIP(hipStreamSynchronize(ws));
  restore(yfinal, nfinal);
  UVLM_HIP(hipStreamSynchronize(s));
}
}  // namespace

int uvlm_wing_lift_adjoint(uvlm_wing w, uvlm_handle particles, const double U[3], double dt, int64_t steps,
                           int reference_offset, const double e[3], double *CL, int64_t n, double *xbar,
                           double *alphabar)
{
  return guarded([&] {
    check_pair(w, particles);
    if (!particles || !U || !e) throw Error(UVLM_ERR_ARG, "null particles, freestream or direction");
    if (steps < 0) throw Error(UVLM_ERR_ARG, "negative step count");
    if (n != particles->N) throw Error(UVLM_ERR_ARG, "size mismatch");
    check_vjp_config(particles);
    UVLM_HIP(hipSetDevice(w->device));
    // one segment over the whole window: every state kept, nothing recomputed
    wing_window_adjoint(w, particles, U, dt, steps, std::max<int64_t>(steps, 1), reference_offset, e, CL, xbar,
                        alphabar, nullptr);
  });
}

int uvlm_wing_window_adjoint(uvlm_wing w, uvlm_handle particles, const double U[3], double dt, int64_t steps,
                             int64_t every, int reference_offset, const double e[3], double *CL, int64_t n,
                             double *xbar, double *alphabar, double *vertex_bar)
{
  return guarded([&] {
    check_pair(w, particles);
    if (!particles || !U || !e) throw Error(UVLM_ERR_ARG, "null particles, freestream or direction");
    if (steps < 0 || every < 1) throw Error(UVLM_ERR_ARG, "steps must be >= 0 and every >= 1");
    if ((xbar || alphabar) && n != particles->N) throw Error(UVLM_ERR_ARG, "size mismatch");
    check_vjp_config(particles);
    UVLM_HIP(hipSetDevice(w->device));
    wing_window_adjoint(w, particles, U, dt, steps, every, reference_offset, e, CL, xbar, alphabar, vertex_bar);
  });
}

int uvlm_update_bodies_soa(uvlm_handle h, int64_t n, const double *X, const double *alpha)
{
  return guarded([&] {
    check_handle(h);
    if (n != h->N) throw Error(UVLM_ERR_ARG, "size mismatch");
    set_device(h);
    std::vector<double> buf((size_t)3 * n);
    auto put = [&](int f0, const double *in) {
      if (!in) return;
      for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) buf[(size_t)d * n + i] = in[3 * i + d];
      for (int d = 0; d < 3; d++)
        UVLM_HIP(hipMemcpyAsync(h->F(f0 + d), buf.data() + (size_t)d * n, 8 * n, hipMemcpyHostToDevice, h->s));
      UVLM_HIP(hipStreamSynchronize(h->s));
    };
    put(F_X, X);
    put(F_A, alpha);
    h->evaluated = false;
  });
}

int uvlm_direct(uvlm_handle h, int64_t m, const double *xt, const double *at, double *u, double *J, double *dadt)
{
  return guarded([&] {
    check_direct_args(h, m, xt, at, u, dadt);
    set_device(h);
    DevBuf        buf;
    const int64_t nin = 3 * m + (at ? 3 * m : 0), nout = 3 * m + (J ? 9 * m : 0) + (dadt ? 3 * m : 0);
    buf.reserve(sizeof(double) * (nin + nout));
    direct_staged(h, m, xt, at, u, J, dadt, buf.as<double>());
  });
}

int uvlm_direct_noalloc(uvlm_handle h, int64_t m, const double *xt, const double *at, double *u, double *J,
                        double *dadt)
{
  return guarded([&] {
    check_direct_args(h, m, xt, at, u, dadt);
    if (m > UVLM_DIRECT_NOALLOC_MAX)
      throw Error(UVLM_ERR_ARG, "uvlm_direct_noalloc takes at most " + std::to_string(UVLM_DIRECT_NOALLOC_MAX) +
                                    " targets, got " + std::to_string(m));
    set_device(h);
    direct_staged(h, m, xt, at, u, J, dadt, scratch_on(h->prm.device));
  });
}

int uvlm_timings(uvlm_handle h, uvlm_timing *out)
{
  if (!h || !out) return fail(UVLM_ERR_ARG, "null argument");
  *out           = h->tm;
  out->ncrit     = h->prm.ncrit;
  out->m2l_slots = h->sched.plan.S;
  return UVLM_OK;
}

int uvlm_sync(uvlm_handle h)
{
  return guarded([&] {
    check_handle_keep(h);
    UVLM_HIP(hipStreamSynchronize(h->s));
  });
}

}  // extern "C"
