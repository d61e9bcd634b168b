// common.hpp -- internal declarations shared by the HIP sources of libuvlm_fmm.so.
//
// Device data layout (all arrays live in HBM, SoA):
//   bodies : 37 field arrays of N doubles, field f at st + f*cap, fields in exafmm::Body
//            order (src/exafmmTypes.hpp:18-32): J 0-8, X 9-11, alpha 12-14, velocity 15-17,
//            dadt 18-20, velocity_old 21-23, dadt_old 24-26, X_old 27-29, alpha_old 30-32,
//            radius 33, drdt 34, radius_old 35, drdt_old 36.
//   cells  : int32 nchild/nbody/child/body/level/parent arrays and double cx/cy/cz/R arrays
//            in the reference's cell order (root 0, child groups in DFS pre-order).
//   expansions : [C][NTERM] complex doubles, index 3*(n(n+1)/2+m)+d (multipoleKernels.hpp:468-471).
//   lists  : CSR by target cell (offsets C+1, sources int32), sorted by (target, source).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/uvlm_fmm.h"

#include "m2l_plan.hpp"

namespace uvlm {

// Body field indices (exafmm::Body order).
enum Field : int {
  F_J = 0, F_X = 9, F_A = 12, F_U = 15, F_DADT = 18, F_UOLD = 21, F_DAOLD = 24, F_XOLD = 27, F_AOLD = 30,
  F_SIG = 33, F_DSIG = 34, F_SIGOLD = 35, F_DSIGOLD = 36, N_FIELDS = 37
};

constexpr int    KEY_LEVELS = 21;  // 3 bits per level in a 63-bit key: cells at levels 0..21
constexpr double INV_4PI    = 0.0795774715459476678844418816863;
constexpr double INV_PI     = 0.318309886183790671537767526745;
constexpr double INV_SQRTPI = 0.564189583547756286948079451561;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define UVLM_HIP(call)                                                                                          \
  do {                                                                                                          \
    hipError_t e_ = (call);                                                                                     \
    if (e_ != hipSuccess)                                                                                       \
      throw ::uvlm::Error(UVLM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + " at " __FILE__ ":" + \
                                            std::to_string(__LINE__));                                          \
  } while (0)

#define UVLM_LAUNCH_CHECK() UVLM_HIP(hipGetLastError())

// Growable device buffer (bytes).
struct DevBuf {
  void  *p     = nullptr;
  size_t bytes = 0;
  DevBuf()     = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release()
  {
    if (p) (void)hipFree(p);
    p     = nullptr;
    bytes = 0;
  }
  // Ensures capacity; contents are NOT preserved unless keep is true.
  void reserve(size_t want, bool keep = false, hipStream_t s = nullptr)
  {
    if (want <= bytes) return;
    size_t nb = want + want / 4 + 256;
    void  *q  = nullptr;
    if (hipMalloc(&q, nb) != hipSuccess) throw Error(UVLM_ERR_NOMEM, "hipMalloc of " + std::to_string(nb) + " bytes failed");
    if (keep && p && bytes) {
      UVLM_HIP(hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, s));
      UVLM_HIP(hipStreamSynchronize(s));
    }
    release();
    p     = q;
    bytes = nb;
  }
  template <class T> T *as() const { return (T *)p; }
};
inline void swap_buf(DevBuf &a, DevBuf &b)
{
  std::swap(a.p, b.p);
  std::swap(a.bytes, b.bytes);
}

// An octree in the reference's cell order.
struct Tree {
  int64_t C = 0, nleaf = 0;
  int     nlevels = 0;
  int     sort_digits = 21;       // key digits the next build of this tree sorts (its last depth + 1)
  std::vector<int64_t> lvl_off;   // host: level l cells are lvl_cells[lvl_off[l] .. lvl_off[l+1])
  DevBuf nchild, nbody, child, body, level, parent;  // int32 [C]
  DevBuf cx, cy, cz, R;                              // double [C]
  DevBuf lvl_cells;                                  // int32 [C], grouped by level, ascending index
  DevBuf leaves;                                     // int32 [nleaf]
  DevBuf X0R0;                                       // double[4]
  int32_t *i(DevBuf &b) { return b.as<int32_t>(); }
};

// Interaction lists in CSR form by target cell.
struct Lists {
  int64_t np2p = 0, nm2l = 0, n_interactions = 0;
  DevBuf  p2p_off, p2p_src, m2l_off, m2l_src;  // int32
};

// Scratch shared by tree builds and traversals.
struct Scratch {
  DevBuf keys0, keys1, idx0, idx1, sort_tmp, scan_tmp;
  DevBuf rec_begin, rec_end, rec_parent, rec_rank, rec_nchild, rec_scan, rec_pfx, rec_pre, rec_gidx, rec_childoff;
  DevBuf rec_cx, rec_cy, rec_cz;
  DevBuf partial, counters, perm;
  DevBuf front0, front1, out_p2p, out_m2l;
  DevBuf p2p_seg, p2p_part;  // P2P list segments and their partial sums (few-target case)
  DevBuf m2l_seg, m2l_part;  // the same for M2L
  int64_t rec_cap = 0;
};

// ---------------------------------------------------------------------------------
// Launchers (defined in the .hip files).
// tree.hip
void build_tree(hipStream_t s, Scratch &sc, const double *x, const double *y, const double *z, int64_t N, int64_t ncrit,
                Tree &tree, int32_t *perm_out);
// gap > 0: pairs whose bodies can come closer than gap are near too (the near-field option)
void traverse(hipStream_t s, Scratch &sc, const Tree &tt, const Tree &st, double theta, Lists &lists,
              double gap = 0.0);
double max_value(hipStream_t s, Scratch &sc, const double *x, int64_t n);
void gather_fields(hipStream_t s, const double *src, double *dst, int64_t src_cap, int64_t dst_cap, int nfields,
                   const int32_t *perm, int64_t N);
void gather_i32(hipStream_t s, const int32_t *src, int32_t *dst, const int32_t *perm, int64_t N);
// Lists of (sources as targets): CSR by the cells of `in_sources`, entries = cells of `in_targets`.
void transpose_lists(hipStream_t s, Scratch &sc, const Tree &in_targets, const Tree &in_sources, const Lists &in,
                     Lists &out);
void count_interactions(hipStream_t s, Scratch &sc, const Tree &tt, const Tree &st, const Lists &l, int64_t *out);

// far.hip
struct M2LSchedule {
  int     budget = 0;  // accumulator slots per lane (0: default)
  int     waves  = 0;  // wavefronts per M2L workgroup (1 or 2; 0: 1 for P <= 8, else 2)
  int     W      = 1;  // the wavefronts per workgroup of the current schedule
  M2LPlan plan;
  DevBuf  dev;         // slots, lev, out_off, out_src
  int64_t dev_slots = 0, dev_lev = 0, dev_out_off = 0, dev_out_src = 0;
  DevBuf  ytab;        // singular harmonics of the lattice vectors [-K, K]^3 (M2L_TABLE_K)
  DevBuf  octY;        // regular harmonics of the 8 unit octant vectors (M2M, L2L)
  DevBuf  m2mtab;      // term table of the M2M for P <= 10: 64 counts, then (m2m_rp + m2m_rc) rows
  int     m2m_rp = 0, m2m_rc = 0;
  DevBuf  l2ltab;      // term table of the L2L for P = 6, 8, 10: 32 entries per lane
  bool    l2l_tab = false;
  int     ytab_P = 0;
};
// Half-width of the cube of integer offsets (in units of the smaller cell's half-width) whose
// singular harmonics the M2L reads from a table; 99.9% of the 4M leapfrog tree's M2L pairs.
constexpr int M2L_TABLE_K = 12;
// M2L inputs for the lattice table (Y null: compute every pair's harmonics).
struct M2LTable {
  const double2 *Y;
  const int32_t *level;
  const double  *X0R0;
};
int  m2l_waves(int P, int waves);  // wavefronts per M2L workgroup for order P (waves: 0 = by P)
void make_m2l_schedule(int P, M2LSchedule &s);
void zero_doubles(hipStream_t s, double *p, int64_t n);
void p2m(hipStream_t s, int P, const Tree &t, const double *st, int64_t cap, double2 *M);
void m2m_level(hipStream_t s, int P, const Tree &t, int level, double2 *M, const M2LSchedule &sch);
void m2l(hipStream_t s, int P, const Tree &tt, const Tree &st, const Lists &l, const M2LSchedule &sch, const double2 *M,
         double2 *L, Scratch *sc = nullptr);
void l2l_level(hipStream_t s, int P, const Tree &t, int level, double2 *L, const M2LSchedule &sch);
// adjoint of the particle rates: vortex + vortex-dipole P2M (aux fields 0..11) and the far-field
// target-side term (adds into aux fields 12..14)
void p2m_adj(hipStream_t s, int P, const Tree &t, const double *st, const double *aux, int64_t cap, double2 *M);
void l2p_adj_tgt(hipStream_t s, int P, const Tree &t, const double2 *L, const double *st, double *aux, int64_t cap);
void l2p_stretch(hipStream_t s, int P, const Tree &t, const double2 *L, double *st, int64_t cap, int stretching,
                 int formulation, int sensor_mode);

// near.hip
// max_leaf: an upper bound of the target leaves' body counts (ncrit for a built tree)
void p2p(hipStream_t s, const Tree &tt, const Tree &stree, const Lists &l, const double *tgt, int64_t tcap,
         const double *src, int64_t scap, double *out, int64_t ocap, int precision, int sensor_mode,
         int64_t max_leaf, Scratch *sc = nullptr);

void gauss_near(hipStream_t s, const Tree &tt, const Tree &stree, const Lists &l, const double *tgt, int64_t tcap,
                const double *srcst, int64_t scap, double *out, int64_t ocap, int mode, int field, double prefix);
void direct(hipStream_t s, int64_t m, const double *xt, const double *at, const double *sst, int64_t scap, int64_t n,
            double *u, double *J, double *dadt);

// adjoint.hip: vector-Jacobian product of the rates; adjoint fields 0-2 ubar, 3-11 Jbar,
// 12-14 xbar, 15-17 alphabar (abar on entry to vjp_setup)
constexpr int N_ADJ_FIELDS = 18;
void vjp_setup(hipStream_t s, int64_t N, const double *st, double *aux, int64_t cap, int classic);
void vjp_finish(hipStream_t s, int64_t N, const double *ast, double *aux, int64_t cap);
// side 0: forward P2P list (xbar of targets); side 1: transposed list (xbar, alphabar of sources)
// mixed: FP32 pair arithmetic (FP64 accumulation), else FP64
void p2p_adj(hipStream_t s, const Tree &t, const Lists &l, int side, int mixed, const double *st, double *aux,
             int64_t cap);
// sbar_j (device order of the sources) = d/dsigma_j of the near-pair sum of ubar_i . u_i +
// Jbar_i : J_i (Jbar when with_jbar), over tl = the P2P list CSR by source cell (entries = target
// cells); tcot holds ubar (3 fields) then Jbar (9 fields) with stride ccap, tx the targets' X.
void sigma_bar(hipStream_t s, const Tree &stree, const Tree &ttree, const Lists &tl, const double *tx, int64_t tcap,
               const double *tcot, int64_t ccap, bool with_jbar, const double *sst, int64_t scap, double *sbar);
// dst[orig[p]] += src[p] (device order -> original order)
void add_to_orig(hipStream_t s, int64_t N, const int32_t *orig, const double *src, double *dst);
// state cotangents in the original body order (6 fields of ccap doubles) <-> adjoint fields
void cot_gather(hipStream_t s, int64_t N, const int32_t *orig, const double *cot, int64_t ccap, double *aux,
                int64_t cap);
void cot_scatter(hipStream_t s, int64_t N, const int32_t *orig, const double *aux, int64_t cap, double *cot,
                 int64_t ccap);
// X, alpha (fields F_X..F_A+2) and sigma of the device bodies into 7 fields of N doubles,
// original order
void state_to_orig(hipStream_t s, int64_t N, const int32_t *orig, const double *st, int64_t cap, double *dst);
// out[i] = sum_k c[k] p[k][i] (at most 6 terms)
void lincomb(hipStream_t s, int64_t n, double *out, int nterms, const double *const *p, const double *c);

// vlm.hip: vortex-lattice wing (vertices V AoS [(nchord+1)(nspan+1)][3], n = nspan * nchord panels)
void wing_geom(hipStream_t s, int n, int nspan, const double *V, double *cp, double *nrm, double *area);
void wing_influence(hipStream_t s, int m, int n, int nspan, const double *V, const double *pts, const double *nrm,
                    double *B, double *A, int *singular);
void lu_factor(hipStream_t s, int n, double *A, int *piv, int *info);
// A^{-1} column-major from the factors (one workgroup per column) and products with it
void lu_inverse(hipStream_t s, int n, const double *LU, const int *piv, double *Xc);
void matvec_inv(hipStream_t s, int n, const double *Xc, const double *b, double *out);
void matvec_inv_t(hipStream_t s, int n, const double *Xc, const double *g, double *out);
// u_i = U + v_i + sum_j B[i][j] gamma_j at the control points (v may be null)
void wing_velocity(hipStream_t s, int n, const double *B, const double *gamma, const double U[3], const double *v,
                   double *u);
void wing_rhs(hipStream_t s, int n, const double *nrm, const double U[3], const double *v, double *rhs);
void wing_force(hipStream_t s, int n, int nspan, const double *V, const double *u, const double *gamma, double *F);
void wing_shed(hipStream_t s, int nspan, int nchord, const double *V, const double *gamma, const double U[3], double dt,
               int ref_offset, double *X, double *A, double *S);
void translate3(hipStream_t s, int64_t n, double *p, const double d[3]);
void points_to_soa(hipStream_t s, int64_t m, const double *pts, double *st, int64_t cap);
void probe_velocity_out(hipStream_t s, int64_t m, const int32_t *perm, const double *st, int64_t cap, double *out);
void append_bodies(hipStream_t s, int64_t N, int64_t k, const double *X, const double *A, const double *S, double *st,
                   int64_t cap, int32_t *orig);
// adjoint of the lift through the loop: force VJP (c [n], vbar [n][3], gbar [n]),
// rhs and shedding VJPs, probe "vortices", probe-adjoint results into original-order cotangents
void wing_force_vjp(hipStream_t s, int n, int nspan, const double *V, const double *B, const double *u,
                    const double *gamma, const double e[3], double *c, double *vbar, double *gbar);
void wing_rhs_vjp(hipStream_t s, int n, const double *nrm, const double *rbar, double *vbar);
void wing_shed_vjp(hipStream_t s, int nspan, int nchord, const double *V, const double *abar, int64_t stride,
                   double *gbar);
void probe_sources(hipStream_t s, int64_t m, const double *pts, const double *wts, double *st, int64_t cap);
void cot_add_probe(hipStream_t s, int64_t N, const int32_t *orig, const double *ast, int64_t cap, double *cot,
                   int64_t ccap);
// design gradient: cotangents of the wing's vertices.  Corner cotangents per panel are
// [n][12] (le, le1, te, te1); wing_vertex_gather sums up to three such buffers into [nv][3].
void wing_influence_vjp(hipStream_t s, int n, int nspan, const double *V, const double *cp, const double *nrm,
                        const double *B, const double *Abar, const double *ub, const double *gamma, double *pb_ring,
                        double *cpb, double *nb);
void wing_geom_vjp(hipStream_t s, int n, int nspan, const double *V, const double *cpb1, const double *cpb2,
                   const double *nb1, const double *nb2, double areab, const double *dlb, double *pb);
void wing_shed_geom_vjp(hipStream_t s, int nspan, int nchord, const double *V, const double *gamma, const double *lam,
                        int64_t stride, int64_t base, const double *sb, double *pb);
void wing_vertex_gather(hipStream_t s, int nspan, int nchord, const double *pb0, const double *pb1, const double *pb2,
                        double *vb);
void wing_solve_geom_acc(hipStream_t s, int n, const double *rbar, const double *gamma, const double U[3],
                         const double *v, double *Abar, double *nb);
void wing_cp_acc(hipStream_t s, int n, const double *vb, const double *J, double *cpb);
void wing_force_dl_vjp(hipStream_t s, int n, int nspan, const double *u, const double *gamma, const double e[3],
                       double *dlb);
void probe_velocity_grad_out(hipStream_t s, int64_t m, const int32_t *perm, const double *st, int64_t cap, double *out,
                             double *Jout);

// body.hip
void aos_to_soa(hipStream_t s, const void *aos, int64_t n, int64_t stride_doubles, double *st, int64_t cap);
void soa_to_aos(hipStream_t s, const double *st, int64_t cap, void *aos, int64_t n, int64_t stride_doubles);
void rk_stage(hipStream_t s, double *st, int64_t cap, int64_t N, int scheme, int stage, double dt, int store_initial);
void reset_sensor_rates(hipStream_t s, double *st, int64_t cap, int64_t N);
void relax(hipStream_t s, double *st, int64_t cap, int64_t N, double f);
void diagnostics(hipStream_t s, const double *st, int64_t cap, int64_t N, double *part_dev, double out[10]);
void csm(hipStream_t s, double *st, int64_t cap, int64_t N, double nu, double dt);
void zero_fields(hipStream_t s, double *st, int64_t cap, int64_t N, int f0, int nf);
// out[d] = sum_i st[fa+d][i] * st[fb+d][i]; part_dev holds 3 * 512 doubles
void dot3(hipStream_t s, const double *st, int64_t cap, int64_t N, int fa, int fb, double *part_dev, double out[3]);
void rbf_vec(hipStream_t s, double *st, int64_t cap, int64_t N, int mode, const double c[3]);

}  // namespace uvlm
