// tree.hip -- octree build and dual-tree traversal on the device.
//
// Compiled with -ffp-contract=off: the octant tests, the child-centre recurrence and the
// multipole acceptance test must round exactly as the reference writes them, because the
// tree and the interaction lists are compared bit-exactly.
//
// Octree (semantics of src/tree.hpp:7-113):
//   1. bounds: fmin/fmax reduction, X0 = (max+min)/2, R0 = max half-extent * 1.00001 (tree.hpp:7-26)
//   2. a 63-bit key per body: the octant digits of the 21 levels below the root, each
//      digit (x>cx)+2(y>cy)+4(z>cz) against the FP64 centre recurrence
//      c_{l+1} = c_l + (R0/2^(l+1))*(+-1) (tree.hpp:57, 86-92)
//   3. radix sort of (key, index): stable, so equal keys keep the input order.  Only the top
//      digits the tree can use are sorted: the previous build's depth plus one level; a cell
//      that must split below them flags the build, which then sorts all 21 digits again
//   4. cells level by level: a cell splits iff NBODY > ncrit (tree.hpp:43); its children are
//      the non-empty digit ranges, found by binary search in the sorted keys
//   5. cell numbering of the reference: root 0, child groups appended in DFS pre-order of
//      their parents (tree.hpp:81-96), i.e. ChildOffset(p) = 1 + sum of NCHILD over internal
//      cells before p in pre-order; pre-order = order by (BodyOffset, level)
//   6. inside a leaf the bodies keep their input order (the reference stops sorting at the
//      leaf), so a second, stable sort of the input indices by their leaf's first body fixes the
//      final permutation.
// Traversal (semantics of src/traverse.hpp:32-85): breadth-first over (target, source)
// pairs with the reference's split rule and MAC; outputs are appended with one atomic per
// wavefront and category, then sorted by (target, source).
#include <cstdlib>
#include <cstring>

#include "common.hpp"

#include <rocprim/rocprim.hpp>

#include <algorithm>

namespace uvlm {

namespace {

constexpr int BLK = 256;

inline int nblocks(int64_t n, int b = BLK) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

__global__ void k_bounds_partial(const double *__restrict__ x, const double *__restrict__ y,
                                 const double *__restrict__ z, int64_t N, double *part)
{
  __shared__ double s[6][BLK];
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < N; i += (int64_t)gridDim.x * BLK) {
    const double v[3] = {x[i], y[i], z[i]};
    for (int d = 0; d < 3; d++) {
      mn[d] = fmin(v[d], mn[d]);
      mx[d] = fmax(v[d], mx[d]);
    }
  }
  for (int d = 0; d < 3; d++) {
    s[d][threadIdx.x]     = mn[d];
    s[3 + d][threadIdx.x] = mx[d];
  }
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int d = 0; d < 3; d++) {
        s[d][threadIdx.x]     = fmin(s[d][threadIdx.x + w], s[d][threadIdx.x]);
        s[3 + d][threadIdx.x] = fmax(s[3 + d][threadIdx.x + w], s[3 + d][threadIdx.x]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int d = 0; d < 6; d++) part[blockIdx.x * 6 + d] = s[d][0];
}

// X0 and R0 exactly as tree.hpp:19-25 computes them from Xmin/Xmax (one workgroup of BLK;
// min and max are exact in any order, so the parallel reduction gives the serial result).
__global__ __launch_bounds__(BLK) void k_bounds_final(const double *part, int nb, double *X0R0)
{
  __shared__ double s[6][BLK];
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < nb; b += BLK)
    for (int d = 0; d < 3; d++) {
      mn[d] = fmin(part[b * 6 + d], mn[d]);
      mx[d] = fmax(part[b * 6 + 3 + d], mx[d]);
    }
  for (int d = 0; d < 3; d++) {
    s[d][threadIdx.x]     = mn[d];
    s[3 + d][threadIdx.x] = mx[d];
  }
  __syncthreads();
  for (int w = BLK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int d = 0; d < 3; d++) {
        s[d][threadIdx.x]     = fmin(s[d][threadIdx.x + w], s[d][threadIdx.x]);
        s[3 + d][threadIdx.x] = fmax(s[3 + d][threadIdx.x + w], s[3 + d][threadIdx.x]);
      }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  for (int d = 0; d < 3; d++) {
    mn[d] = s[d][0];
    mx[d] = s[3 + d][0];
  }
  double X0[3], R0 = 0;
  for (int d = 0; d < 3; d++) X0[d] = (mx[d] + mn[d]) / 2;
  for (int d = 0; d < 3; d++) {
    R0 = fmax(X0[d] - mn[d], R0);
    R0 = fmax(mx[d] - X0[d], R0);
  }
  R0 *= 1.00001;
  X0R0[0] = X0[0];
  X0R0[1] = X0[1];
  X0R0[2] = X0[2];
  X0R0[3] = R0;
}

__global__ void k_keys(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                       int64_t N, const double *__restrict__ X0R0, uint64_t *__restrict__ keys,
                       uint32_t *__restrict__ idx)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i >= N) return;
  const double xi = x[i], yi = y[i], zi = z[i];
  double       cx = X0R0[0], cy = X0R0[1], cz = X0R0[2];
  const double R0 = X0R0[3];
  uint64_t     key = 0;
  for (int l = 0; l < KEY_LEVELS; l++) {
    const int o = (xi > cx) + ((yi > cy) << 1) + ((zi > cz) << 2);
    key         = (key << 3) | (uint64_t)o;
    const double r = R0 / (double)(1LL << (l + 1));
    cx += r * (double)(((o & 1) >> 0) * 2 - 1);
    cy += r * (double)(((o & 2) >> 1) * 2 - 1);
    cz += r * (double)(((o & 4) >> 2) * 2 - 1);
  }
  keys[i] = key;
  idx[i]  = (uint32_t)i;
}

// The root record (begin 0, end N, no parent, rank 0, centre X0) and the error flag, on the
// device: no host round trip.
__global__ void k_root_record(int32_t n, const double *__restrict__ X0R0, int32_t *rb, int32_t *re, int32_t *rp,
                              int32_t *rr, double *cx, double *cy, double *cz, int32_t *err)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  rb[0] = 0;
  re[0] = n;
  rp[0] = -1;
  rr[0] = 0;
  cx[0] = X0R0[0];
  cy[0] = X0R0[1];
  cz[0] = X0R0[2];
  *err  = 0;
}

// The three numbers the host needs after counting a level: last exclusive-scan entry, last
// child count, error flag.
__global__ void k_level_summary(const int32_t *__restrict__ scan_last, const int32_t *__restrict__ nchild_last,
                                const int32_t *__restrict__ err, int32_t *__restrict__ out)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  out[0] = *scan_last;
  out[1] = *nchild_last;
  out[2] = *err;
}

__device__ __forceinline__ int digit_at(uint64_t key, int l) { return (int)((key >> (3 * (KEY_LEVELS - 1 - l))) & 7); }

__device__ __forceinline__ void octant_bounds(const uint64_t *__restrict__ keys, int b, int e, int l, int bnd[9])
{
  bnd[0] = b;
  bnd[8] = e;
  for (int o = 1; o < 8; o++) {
    int lo = bnd[o - 1], hi = e;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (digit_at(keys[mid], l) < o) lo = mid + 1;
      else hi = mid;
    }
    bnd[o] = lo;
  }
}

// err bit 0: deeper than the key holds; bit 1: a split needs a digit below the dlim sorted ones.
__global__ void k_count_children(int64_t off, int64_t cnt, int l, int64_t ncrit, int dlim, const int32_t *__restrict__ rb,
                                 const int32_t *__restrict__ re, const uint64_t *__restrict__ keys,
                                 int32_t *__restrict__ nchild, int32_t *__restrict__ err)
{
  const int64_t t = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (t >= cnt) return;
  const int64_t k = off + t;
  const int     b = rb[k], e = re[k];
  int           nc = 0;
  if (e - b > ncrit) {
    if (l >= KEY_LEVELS) {
      atomicOr(err, 1);
    } else if (l >= dlim) {
      atomicOr(err, 2);
    } else {
      int bnd[9];
      octant_bounds(keys, b, e, l, bnd);
      for (int o = 0; o < 8; o++) nc += bnd[o + 1] > bnd[o];
    }
  }
  nchild[k] = nc;
}

__global__ void k_write_children(int64_t off, int64_t cnt, int l, const int32_t *__restrict__ rb,
                                 const int32_t *__restrict__ re, const uint64_t *__restrict__ keys,
                                 const int32_t *__restrict__ nchild, const int32_t *__restrict__ scan,
                                 int64_t child_base, const double *__restrict__ X0R0, int32_t *__restrict__ ob,
                                 int32_t *__restrict__ oe, int32_t *__restrict__ opar, int32_t *__restrict__ orank,
                                 double *__restrict__ cx, double *__restrict__ cy, double *__restrict__ cz)
{
  const int64_t t = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (t >= cnt) return;
  const int64_t k = off + t;
  if (nchild[k] == 0) return;
  int bnd[9];
  octant_bounds(keys, rb[k], re[k], l, bnd);
  const double r = X0R0[3] / (double)(1LL << (l + 1));
  const double px = cx[k], py = cy[k], pz = cz[k];
  int          c  = 0;
  for (int o = 0; o < 8; o++) {
    if (bnd[o + 1] == bnd[o]) continue;
    const int64_t q = child_base + scan[t] + c;
    ob[q]           = bnd[o];
    oe[q]           = bnd[o + 1];
    opar[q]         = (int32_t)k;
    orank[q]        = c;
    double x = px, y = py, z = pz;
    x += r * (double)(((o & 1) >> 0) * 2 - 1);
    y += r * (double)(((o & 2) >> 1) * 2 - 1);
    z += r * (double)(((o & 4) >> 2) * 2 - 1);
    cx[q] = x;
    cy[q] = y;
    cz[q] = z;
    c++;
  }
}

__global__ void k_flags(int64_t n, const int32_t *__restrict__ nchild, int32_t *__restrict__ flag)
{
  const int64_t k = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (k < n) flag[k] = nchild[k] > 0;
}

struct LevelOffsets {
  int32_t off[KEY_LEVELS + 3];
  int     nlev;
};

// Pre-order rank of each internal record = #internal cells with (begin, level) before it.
__global__ void k_preorder_rank(int64_t C, LevelOffsets lo, const int32_t *__restrict__ rb,
                                const int32_t *__restrict__ nchild, const int32_t *__restrict__ pfx,
                                int32_t *__restrict__ pre_nchild, int32_t *__restrict__ rank_out)
{
  const int64_t k = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (k >= C || nchild[k] == 0) return;
  int lk = 0;
  while (lk + 1 < lo.nlev && lo.off[lk + 1] <= k) lk++;
  const int b    = rb[k];
  int       rank = 0;
  for (int l = 0; l < lo.nlev; l++) {
    int lo_ = lo.off[l], hi = lo.off[l + 1];
    const int base = lo_;
    // count records of level l with begin < b (l >= lk) or begin <= b (l < lk)
    while (lo_ < hi) {
      const int mid = (lo_ + hi) >> 1;
      const bool before = (l < lk) ? (rb[mid] <= b) : (rb[mid] < b);
      if (before) lo_ = mid + 1;
      else hi = mid;
    }
    rank += pfx[lo_] - pfx[base];
  }
  pre_nchild[rank] = nchild[k];
  rank_out[k]      = rank;
}

__global__ void k_assign(int64_t C, const int32_t *__restrict__ par, const int32_t *__restrict__ rnk,
                         const int32_t *__restrict__ pre_rank, const int32_t *__restrict__ pre_scan,
                         int32_t *__restrict__ gidx)
{
  const int64_t k = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (k >= C) return;
  if (k == 0) {
    gidx[0] = 0;
    return;
  }
  const int p = par[k];
  gidx[k]     = 1 + pre_scan[pre_rank[p]] + rnk[k];
}

__global__ void k_scatter_cells(int64_t C, LevelOffsets lo, const int32_t *__restrict__ rb,
                                const int32_t *__restrict__ re, const int32_t *__restrict__ nchild,
                                const int32_t *__restrict__ par, const int32_t *__restrict__ pre_rank,
                                const int32_t *__restrict__ pre_scan, const int32_t *__restrict__ gidx,
                                const int32_t *__restrict__ pfx, const double *__restrict__ rcx,
                                const double *__restrict__ rcy, const double *__restrict__ rcz,
                                const double *__restrict__ X0R0, int32_t *__restrict__ c_nchild,
                                int32_t *__restrict__ c_nbody, int32_t *__restrict__ c_child,
                                int32_t *__restrict__ c_body, int32_t *__restrict__ c_level,
                                int32_t *__restrict__ c_parent, double *__restrict__ cx, double *__restrict__ cy,
                                double *__restrict__ cz, double *__restrict__ cR, int32_t *__restrict__ lvl_cells,
                                int32_t *__restrict__ leaves)
{
  const int64_t k = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (k >= C) return;
  int lk = 0;
  while (lk + 1 < lo.nlev && lo.off[lk + 1] <= k) lk++;
  const int g  = gidx[k];
  const int nc = nchild[k];
  c_nchild[g]  = nc;
  c_nbody[g]   = re[k] - rb[k];
  c_child[g]   = nc ? 1 + pre_scan[pre_rank[k]] : 0;
  c_body[g]    = rb[k];
  c_level[g]   = lk;
  c_parent[g]  = k == 0 ? -1 : gidx[par[k]];
  cx[g]        = rcx[k];
  cy[g]        = rcy[k];
  cz[g]        = rcz[k];
  cR[g]        = X0R0[3] / (double)(1LL << lk);
  lvl_cells[k] = g;
  if (nc == 0) leaves[k - pfx[k]] = g;
}

// lb[input index] = first body of its leaf (one wavefront per leaf; lanes stride over its bodies)
__global__ void k_leaf_begin(int64_t nleafrec, const int32_t *__restrict__ leaves, const int32_t *__restrict__ c_body,
                             const int32_t *__restrict__ c_nbody, const uint32_t *__restrict__ idx,
                             uint32_t *__restrict__ lb)
{
  const int64_t w    = (blockIdx.x * (int64_t)BLK + threadIdx.x) >> 6;
  const int     lane = threadIdx.x & 63;
  if (w >= nleafrec) return;
  const int g = leaves[w], b = c_body[g], n = c_nbody[g];
  for (int i = lane; i < n; i += 64) lb[idx[b + i]] = (uint32_t)b;
}

__global__ void k_iota_u32(int64_t N, uint32_t *__restrict__ v)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i < N) v[i] = (uint32_t)i;
}

// GATHER_FIELDS fields per grid row (blockIdx.y): the state moves with full memory parallelism
// and each permutation entry is read once per group of fields.  Two fields per row moved 37
// fields of 4M bodies in 0.45 ms (5.6 TB/s, a plain copy 5.8 TB/s), four in 0.54 ms.
constexpr int GATHER_FIELDS = 2;
__global__ void k_gather_fields(const double *__restrict__ src, double *__restrict__ dst, int64_t scap, int64_t dcap,
                                const int32_t *__restrict__ perm, int64_t N, int nfields)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i >= N) return;
  const int64_t j  = perm[i];
  const int     f0 = blockIdx.y * GATHER_FIELDS, f1 = min(nfields, f0 + GATHER_FIELDS);
  double        v[GATHER_FIELDS];
  for (int f = f0; f < f1; f++) v[f - f0] = src[f * scap + j];
  for (int f = f0; f < f1; f++) dst[f * dcap + i] = v[f - f0];
}

__global__ void k_gather_i32(const int32_t *__restrict__ src, int32_t *__restrict__ dst,
                             const int32_t *__restrict__ perm, int64_t N)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i < N) dst[i] = src[perm[i]];
}

// ------------------------------- traversal -------------------------------------
struct TreeView {
  const int32_t *nchild, *child, *nbody;
  const double  *cx, *cy, *cz, *R;
};

__device__ __forceinline__ int wave_excl_scan(int v, int &total)
{
  const int lane = threadIdx.x & 63;
  int       s    = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(s, o, 64);
    if (lane >= o) s += t;
  }
  total = __shfl(s, 63, 64);
  return s - v;
}

// One BFS step.  counters: [0] next frontier, [1] P2P, [2] M2L (64-bit).
// gap > 0 (near-field option, off in the reference): a pair is also near when its bodies can come
// closer than gap, |dX| < gap + sqrt(3) (Rt + Rs), tested on squares (rounds as the oracle does).
__device__ __forceinline__ void traverse_round(const int32_t *__restrict__ ft, int64_t nf, int64_t i, TreeView T,
                                               TreeView S, double theta, double gap, int32_t *__restrict__ nxt,
                                               int64_t cap_f,
                                               uint64_t *__restrict__ p2p, int64_t cap_p, uint64_t *__restrict__ m2l,
                                               int64_t cap_m, unsigned long long *__restrict__ counters, int sbits,
                                               unsigned long long (*blk)[BLK / 64][3])
{
  int           cls[8];  // 0 push, 1 P2P, 2 M2L
  int           ct[8], cs[8];
  int           nout = 0, npush = 0, np2p = 0, nm2l = 0;
  if (i < nf) {
    const int t = ft[2 * i], s = ft[2 * i + 1];
    const bool split_t = (S.nchild[s] == 0) || (T.R[t] >= S.R[s] && T.nchild[t] != 0);
    const int  first   = split_t ? T.child[t] : S.child[s];
    const int  cnt     = split_t ? T.nchild[t] : S.nchild[s];
    for (int c = 0; c < cnt; c++) {
      const int    a = split_t ? first + c : t;
      const int    b = split_t ? s : first + c;
      const double dx = T.cx[a] - S.cx[b], dy = T.cy[a] - S.cy[b], dz = T.cz[a] - S.cz[b];
      const double D2 = dx * dx + dy * dy + dz * dz, R2 = D2 * theta * theta;
      const double Rs = T.R[a] + S.R[b];
      const double Gs = gap + 1.7320508075688772 * Rs;
      int          k;
      if (R2 < Rs * Rs || (gap > 0.0 && D2 < Gs * Gs)) k = (T.nchild[a] == 0 && S.nchild[b] == 0) ? 1 : 0;
      else k = 2;
      cls[nout] = k;
      ct[nout]  = a;
      cs[nout]  = b;
      nout++;
      npush += k == 0;
      np2p += k == 1;
      nm2l += k == 2;
    }
  }
  int tf, tp, tm;
  int of = wave_excl_scan(npush, tf), op = wave_excl_scan(np2p, tp), om = wave_excl_scan(nm2l, tm);
  // one atomic per class and block: the wavefronts' totals meet in LDS, three lanes reserve the
  // block's ranges and hand each wavefront its base (the counters are the traversal's only
  // serialisation point)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    blk[0][wv][0] = tf;
    blk[0][wv][1] = tp;
    blk[0][wv][2] = tm;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int          cl  = threadIdx.x;
    unsigned long long tot = 0;
    for (int w = 0; w < BLK / 64; w++) tot += blk[0][w][cl];
    unsigned long long base = tot ? atomicAdd(&counters[cl], tot) : 0;
    for (int w = 0; w < BLK / 64; w++) {
      blk[1][w][cl] = base;
      base += blk[0][w][cl];
    }
  }
  __syncthreads();
  const unsigned long long bf = blk[1][wv][0], bp = blk[1][wv][1], bm = blk[1][wv][2];
  for (int c = 0; c < nout; c++) {
    const uint64_t key = ((uint64_t)(uint32_t)ct[c] << sbits) | (uint64_t)(uint32_t)cs[c];
    if (cls[c] == 0) {
      const int64_t q = (int64_t)bf + of++;
      if (q < cap_f) {
        nxt[2 * q]     = ct[c];
        nxt[2 * q + 1] = cs[c];
      }
    } else if (cls[c] == 1) {
      const int64_t q = (int64_t)bp + op++;
      if (q < cap_p) p2p[q] = key;
    } else {
      const int64_t q = (int64_t)bm + om++;
      if (q < cap_m) m2l[q] = key;
    }
  }
}

// One breadth-first step.  counters: [0] next frontier size, [1] P2P and [2] M2L list sizes so far,
// [3] this step's frontier size (read here, so a batch of steps runs without the host).  The
// blocks stride over the frontier; every lane of a block takes part in each round.
__global__ void k_traverse(const int32_t *__restrict__ ft, TreeView T, TreeView S, double theta, double gap,
                           int32_t *__restrict__ nxt, int64_t cap_f, uint64_t *__restrict__ p2p, int64_t cap_p,
                           uint64_t *__restrict__ m2l, int64_t cap_m, unsigned long long *__restrict__ counters,
                           int sbits)
{
  __shared__ unsigned long long blk[2][BLK / 64][3];  // per wavefront: totals, then bases
  const int64_t nf = (int64_t)counters[3];
  for (int64_t base = blockIdx.x * (int64_t)BLK; base < nf; base += (int64_t)gridDim.x * BLK)
    traverse_round(ft, nf, base + threadIdx.x, T, S, theta, gap, nxt, cap_f, p2p, cap_p, m2l, cap_m, counters,
                   sbits, blk);
}

// Start of a traversal: the root pair (0, 0) is the first frontier, or, when both roots are leaves,
// the single P2P pair (traverse.hpp:41-44); the counters are cleared.
__global__ void k_traverse_start(TreeView T, TreeView S, int32_t *__restrict__ front, uint64_t *__restrict__ p2p,
                                 unsigned long long *__restrict__ c)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const bool leaves = T.nchild[0] == 0 && S.nchild[0] == 0;
  if (leaves) p2p[0] = 0;
  front[0] = 0;
  front[1] = 0;
  c[0] = 0;
  c[1] = leaves ? 1 : 0;
  c[2] = 0;
  c[3] = leaves ? 0 : 1;
  c[4] = 0;
  c[5] = 0;
}

// End of a step: the next frontier becomes the current one.  counters[4] (sticky) flags a buffer
// that overflowed (1 frontier, 2 P2P, 4 M2L); the traversal then stops (frontier 0) and the host
// starts it again with larger buffers.  counters[5] keeps the largest frontier seen.
__global__ void k_step_end(unsigned long long *__restrict__ c, int64_t cap_f, int64_t cap_p, int64_t cap_m)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long f = c[4];
  if ((int64_t)c[0] > cap_f) f |= 1;
  if ((int64_t)c[1] > cap_p) f |= 2;
  if ((int64_t)c[2] > cap_m) f |= 4;
  c[5] = c[0] > c[5] ? c[0] : c[5];
  c[4] = f;
  c[3] = f ? 0 : c[0];
  c[0] = 0;
}

// List keys are (row << lbits) | entry, sorted: off[c] = the first key of row c.
__global__ void k_csr(int64_t C, const uint64_t *__restrict__ keys, int64_t n, int lbits, int32_t *__restrict__ off)
{
  const int64_t c = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (c > C) return;
  const uint64_t want = (uint64_t)c << lbits;
  int64_t        lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < want) lo = mid + 1;
    else hi = mid;
  }
  off[c] = (int32_t)lo;
}

__global__ void k_low_bits(int64_t n, const uint64_t *__restrict__ keys, int lbits, int32_t *__restrict__ out)
{
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i < n) out[i] = (int32_t)(uint32_t)(keys[i] & ((1ull << lbits) - 1));
}

// One lane per list entry q: its row t (the last t with off[t] <= q) by binary search, key
// (source, target) for the transposed CSR.
__global__ void k_transpose_keys(int64_t C, int64_t n, const int32_t *__restrict__ off,
                                 const int32_t *__restrict__ src, int tbits, uint64_t *__restrict__ keys)
{
  const int64_t q = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (q >= n) return;
  int64_t lo = 0, hi = C - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (off[mid] <= q) lo = mid;
    else hi = mid - 1;
  }
  keys[q] = ((uint64_t)(uint32_t)src[q] << tbits) | (uint64_t)(uint32_t)lo;
}

// Sum over the P2P list's sorted keys (target << sbits | source) of nbody[target] * nbody[source]:
// lanes stride over the list entries (a lane per target cell walked its list serially: 98 us at
// 4M), and a grid of at most COUNT_BLOCKS blocks keeps the same-address atomics few.
constexpr int COUNT_BLOCKS = 512;
__global__ void k_count_inter(int64_t n, const uint64_t *__restrict__ keys, int sbits,
                              const int32_t *__restrict__ tn, const int32_t *__restrict__ sn,
                              unsigned long long *__restrict__ out)
{
  unsigned long long acc = 0;
  for (int64_t q = blockIdx.x * (int64_t)BLK + threadIdx.x; q < n; q += (int64_t)gridDim.x * BLK) {
    const uint64_t k = keys[q];
    acc += (unsigned long long)tn[k >> sbits] * (unsigned long long)sn[k & ((uint64_t(1) << sbits) - 1)];
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  // one atomic per block (the wavefronts' sums meet in LDS)
  __shared__ unsigned long long wsum[BLK / 64];
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < BLK / 64; w++) tot += wsum[w];
    if (tot) atomicAdd(out, tot);
  }
}

void h2d(hipStream_t s, void *d, const void *h, size_t bytes)
{
  UVLM_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
  UVLM_HIP(hipStreamSynchronize(s));
}

template <class T> void to_host(hipStream_t s, T *h, const void *d, size_t n = 1)
{
  UVLM_HIP(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, s));
  UVLM_HIP(hipStreamSynchronize(s));
}

void sort_keys_u64(hipStream_t s, Scratch &sc, uint64_t *in, uint64_t *out, int64_t n, int end_bit)
{
  size_t bytes = 0;
  UVLM_HIP(rocprim::radix_sort_keys(nullptr, bytes, in, out, (size_t)n, 0, end_bit, s));
  sc.sort_tmp.reserve(bytes);
  UVLM_HIP(rocprim::radix_sort_keys(sc.sort_tmp.p, bytes, in, out, (size_t)n, 0, end_bit, s));
}

void excl_scan_i32(hipStream_t s, Scratch &sc, const int32_t *in, int32_t *out, int64_t n)
{
  size_t bytes = 0;
  UVLM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)n, rocprim::plus<int32_t>(), s));
  sc.scan_tmp.reserve(bytes);
  UVLM_HIP(rocprim::exclusive_scan(sc.scan_tmp.p, bytes, in, out, 0, (size_t)n, rocprim::plus<int32_t>(), s));
}

int bits_for(int64_t v)
{
  int b = 1;
  while (b < 63 && (int64_t(1) << b) <= v) b++;
  return b;
}

}  // namespace

// ---------------------------------------------------------------------------------
void build_tree(hipStream_t s, Scratch &sc, const double *x, const double *y, const double *z, int64_t N,
                int64_t ncrit, Tree &tree, int32_t *perm_out)
{
  if (N <= 0) throw Error(UVLM_ERR_ARG, "empty particle set");
  if (N >= (int64_t(1) << 31)) throw Error(UVLM_ERR_ARG, "more than 2^31-1 particles");
  // 1. bounds
  const int nb = std::min(1024, nblocks(N));
  sc.partial.reserve(sizeof(double) * 6 * nb);
  tree.X0R0.reserve(sizeof(double) * 4);
  k_bounds_partial<<<nb, BLK, 0, s>>>(x, y, z, N, sc.partial.as<double>());
  k_bounds_final<<<1, BLK, 0, s>>>(sc.partial.as<double>(), nb, tree.X0R0.as<double>());
  UVLM_LAUNCH_CHECK();
  // 2-3. keys and stable sort
  sc.keys0.reserve(8 * N);
  sc.keys1.reserve(8 * N);
  sc.idx0.reserve(4 * N);
  sc.idx1.reserve(4 * N);
  k_keys<<<nblocks(N), BLK, 0, s>>>(x, y, z, N, tree.X0R0.as<double>(), sc.keys0.as<uint64_t>(),
                                     sc.idx0.as<uint32_t>());
  UVLM_LAUNCH_CHECK();
  const uint64_t *keys = sc.keys1.as<uint64_t>();
  const uint32_t *idx  = sc.idx1.as<uint32_t>();
  std::vector<int64_t> off;
  int                  dlim = std::min(KEY_LEVELS, std::max(1, tree.sort_digits));
  for (;;) {
    // stable sort of (key, index) on the top dlim digits
    {
      size_t    bytes = 0;
      const int b0    = 3 * (KEY_LEVELS - dlim);
      UVLM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, sc.keys0.as<uint64_t>(), sc.keys1.as<uint64_t>(),
                                         sc.idx0.as<uint32_t>(), sc.idx1.as<uint32_t>(), (size_t)N, b0,
                                         3 * KEY_LEVELS, s));
      sc.sort_tmp.reserve(bytes);
      UVLM_HIP(rocprim::radix_sort_pairs(sc.sort_tmp.p, bytes, sc.keys0.as<uint64_t>(), sc.keys1.as<uint64_t>(),
                                         sc.idx0.as<uint32_t>(), sc.idx1.as<uint32_t>(), (size_t)N, b0,
                                         3 * KEY_LEVELS, s));
    }
    // 4. records level by level
    auto ensure_rec = [&](int64_t want, int64_t used) {
      if (want <= sc.rec_cap) return;
      int64_t cap = std::max<int64_t>(want, 2 * sc.rec_cap);
      cap         = std::max<int64_t>(cap, 1024);
      for (DevBuf *b : {&sc.rec_begin, &sc.rec_end, &sc.rec_parent, &sc.rec_rank, &sc.rec_nchild})
        b->reserve(4 * cap, used > 0, s);
      for (DevBuf *b : {&sc.rec_cx, &sc.rec_cy, &sc.rec_cz}) b->reserve(8 * cap, used > 0, s);
      for (DevBuf *b : {&sc.rec_scan, &sc.rec_pfx, &sc.rec_pre, &sc.rec_gidx, &sc.rec_childoff})
        b->reserve(4 * (cap + 1), used > 0, s);
      sc.rec_cap = cap;
    };
    ensure_rec(std::max<int64_t>(1024, 8 * (N / std::max<int64_t>(1, ncrit) + 1)), 0);
    sc.counters.reserve(64);
    int32_t *err = sc.counters.as<int32_t>(), *summary = err + 8;
    k_root_record<<<1, 64, 0, s>>>((int32_t)N, tree.X0R0.as<double>(), sc.rec_begin.as<int32_t>(),
                                   sc.rec_end.as<int32_t>(), sc.rec_parent.as<int32_t>(), sc.rec_rank.as<int32_t>(),
                                   sc.rec_cx.as<double>(), sc.rec_cy.as<double>(), sc.rec_cz.as<double>(), err);
    UVLM_LAUNCH_CHECK();
    off = {0, 1};
    bool deeper = false;  // a cell must split below the sorted digits
    for (int l = 0;; l++) {
      const int64_t o = off[l], cnt = off[l + 1] - off[l];
      k_count_children<<<nblocks(cnt), BLK, 0, s>>>(o, cnt, l, ncrit, dlim, sc.rec_begin.as<int32_t>(),
                                                    sc.rec_end.as<int32_t>(), keys, sc.rec_nchild.as<int32_t>(), err);
      UVLM_LAUNCH_CHECK();
      excl_scan_i32(s, sc, sc.rec_nchild.as<int32_t>() + o, sc.rec_scan.as<int32_t>(), cnt);
      k_level_summary<<<1, 64, 0, s>>>(sc.rec_scan.as<int32_t>() + cnt - 1, sc.rec_nchild.as<int32_t>() + o + cnt - 1,
                                       err, summary);
      UVLM_LAUNCH_CHECK();
      int32_t sm3[3] = {0, 0, 0};
      to_host(s, sm3, summary, 3);
      if (sm3[2] & 1) throw Error(UVLM_ERR_DEPTH, "more than ncrit coincident bodies: octree deeper than 21 levels");
      if (sm3[2] & 2) {
        deeper = true;
        break;
      }
      const int64_t total = (int64_t)sm3[0] + sm3[1];
      if (total == 0) break;
      ensure_rec(off[l + 1] + total + 1, off[l + 1]);
      k_write_children<<<nblocks(cnt), BLK, 0, s>>>(
          o, cnt, l, sc.rec_begin.as<int32_t>(), sc.rec_end.as<int32_t>(), keys, sc.rec_nchild.as<int32_t>(),
          sc.rec_scan.as<int32_t>(), off[l + 1], tree.X0R0.as<double>(), sc.rec_begin.as<int32_t>(),
          sc.rec_end.as<int32_t>(), sc.rec_parent.as<int32_t>(), sc.rec_rank.as<int32_t>(), sc.rec_cx.as<double>(),
          sc.rec_cy.as<double>(), sc.rec_cz.as<double>());
      UVLM_LAUNCH_CHECK();
      off.push_back(off[l + 1] + total);
    }
    if (deeper) {  // sort all digits and build again
      dlim = KEY_LEVELS;
      continue;
    }
    break;
  }
  tree.sort_digits = (int)off.size() - 1;  // digits 0 .. nlev-1: this tree's splits plus one spare level

  const int64_t C    = off.back();
  const int     nlev = (int)off.size() - 1;
  if (nlev > KEY_LEVELS + 1) throw Error(UVLM_ERR_INTERNAL, "level count");
  if (C >= (int64_t(1) << 31)) throw Error(UVLM_ERR_ARG, "too many cells");

  // 5. numbering
  int32_t *nchild = sc.rec_nchild.as<int32_t>();
  k_flags<<<nblocks(C), BLK, 0, s>>>(C, nchild, sc.rec_scan.as<int32_t>());
  // pfx has C+1 entries: exclusive scan of C+1 flags with a zero appended
  UVLM_HIP(hipMemsetAsync(sc.rec_scan.as<int32_t>() + C, 0, 4, s));
  excl_scan_i32(s, sc, sc.rec_scan.as<int32_t>(), sc.rec_pfx.as<int32_t>(), C + 1);
  int32_t n_internal = 0;
  to_host(s, &n_internal, sc.rec_pfx.as<int32_t>() + C);
  LevelOffsets lo{};
  lo.nlev = nlev;
  for (int l = 0; l <= nlev; l++) lo.off[l] = (int32_t)off[l];
  int32_t *pre_nchild = sc.rec_pre.as<int32_t>();
  int32_t *pre_rank   = sc.rec_childoff.as<int32_t>();
  k_preorder_rank<<<nblocks(C), BLK, 0, s>>>(C, lo, sc.rec_begin.as<int32_t>(), nchild, sc.rec_pfx.as<int32_t>(),
                                             pre_nchild, pre_rank);
  UVLM_LAUNCH_CHECK();
  int32_t *pre_scan = sc.rec_scan.as<int32_t>();
  if (n_internal > 0) excl_scan_i32(s, sc, pre_nchild, pre_scan, n_internal);
  k_assign<<<nblocks(C), BLK, 0, s>>>(C, sc.rec_parent.as<int32_t>(), sc.rec_rank.as<int32_t>(), pre_rank, pre_scan,
                                      sc.rec_gidx.as<int32_t>());
  UVLM_LAUNCH_CHECK();
  tree.C       = C;
  tree.nlevels = nlev;
  tree.nleaf   = C - n_internal;
  tree.lvl_off = off;
  for (DevBuf *b : {&tree.nchild, &tree.nbody, &tree.child, &tree.body, &tree.level, &tree.parent, &tree.lvl_cells})
    b->reserve(4 * C);
  for (DevBuf *b : {&tree.cx, &tree.cy, &tree.cz, &tree.R}) b->reserve(8 * C);
  tree.leaves.reserve(4 * tree.nleaf);
  k_scatter_cells<<<nblocks(C), BLK, 0, s>>>(
      C, lo, sc.rec_begin.as<int32_t>(), sc.rec_end.as<int32_t>(), nchild, sc.rec_parent.as<int32_t>(), pre_rank,
      pre_scan, sc.rec_gidx.as<int32_t>(), sc.rec_pfx.as<int32_t>(), sc.rec_cx.as<double>(), sc.rec_cy.as<double>(),
      sc.rec_cz.as<double>(), tree.X0R0.as<double>(), tree.nchild.as<int32_t>(), tree.nbody.as<int32_t>(),
      tree.child.as<int32_t>(), tree.body.as<int32_t>(), tree.level.as<int32_t>(), tree.parent.as<int32_t>(),
      tree.cx.as<double>(), tree.cy.as<double>(), tree.cz.as<double>(), tree.R.as<double>(),
      tree.lvl_cells.as<int32_t>(), tree.leaves.as<int32_t>());
  UVLM_LAUNCH_CHECK();

  // 6. final permutation: the input indices 0..N-1, stably sorted by their leaf's first body
  uint32_t *lb = sc.keys0.as<uint32_t>(), *lb_sorted = sc.keys1.as<uint32_t>(), *iota = sc.idx0.as<uint32_t>();
  k_leaf_begin<<<nblocks(tree.nleaf * 64), BLK, 0, s>>>(tree.nleaf, tree.leaves.as<int32_t>(),
                                                        tree.body.as<int32_t>(), tree.nbody.as<int32_t>(), idx, lb);
  k_iota_u32<<<nblocks(N), BLK, 0, s>>>(N, iota);
  UVLM_LAUNCH_CHECK();
  {
    size_t bytes = 0;
    UVLM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, lb, lb_sorted, iota, (uint32_t *)perm_out, (size_t)N, 0,
                                       bits_for(N), s));
    sc.sort_tmp.reserve(bytes);
    UVLM_HIP(rocprim::radix_sort_pairs(sc.sort_tmp.p, bytes, lb, lb_sorted, iota, (uint32_t *)perm_out, (size_t)N, 0,
                                       bits_for(N), s));
  }
}

void gather_fields(hipStream_t s, const double *src, double *dst, int64_t src_cap, int64_t dst_cap, int nfields,
                   const int32_t *perm, int64_t N)
{
  if (N <= 0 || nfields <= 0) return;
  k_gather_fields<<<dim3(nblocks(N), (nfields + GATHER_FIELDS - 1) / GATHER_FIELDS), BLK, 0, s>>>(
      src, dst, src_cap, dst_cap, perm, N, nfields);
  UVLM_LAUNCH_CHECK();
}

void gather_i32(hipStream_t s, const int32_t *src, int32_t *dst, const int32_t *perm, int64_t N)
{
  k_gather_i32<<<nblocks(N), BLK, 0, s>>>(src, dst, perm, N);
  UVLM_LAUNCH_CHECK();
}

// Largest value of x[0..n) (non-negative values): per-block maxima, then one block over them.
__global__ void k_max_partial(const double *__restrict__ x, int64_t n, double *__restrict__ part)
{
  __shared__ double red[BLK];
  double            m = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) m = fmax(m, x[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = BLK / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

double max_value(hipStream_t s, Scratch &sc, const double *x, int64_t n)
{
  if (n <= 0) return 0.0;
  const int nb = (int)std::min<int64_t>(512, (n + BLK - 1) / BLK);
  sc.partial.reserve(sizeof(double) * (nb + 1));
  double *part = sc.partial.as<double>();
  k_max_partial<<<nb, BLK, 0, s>>>(x, n, part);
  k_max_partial<<<1, BLK, 0, s>>>(part, nb, part + nb);
  UVLM_LAUNCH_CHECK();
  double m = 0.0;
  to_host(s, &m, part + nb);
  return m;
}

static TreeView view(const Tree &t)
{
  return TreeView{(const int32_t *)t.nchild.p, (const int32_t *)t.child.p, (const int32_t *)t.nbody.p,
                  (const double *)t.cx.p,      (const double *)t.cy.p,      (const double *)t.cz.p,
                  (const double *)t.R.p};
}

void traverse(hipStream_t s, Scratch &sc, const Tree &tt, const Tree &st, double theta, Lists &lists, double gap)
{
  TreeView  T = view(tt), S = view(st);
  const int sbits = bits_for(std::max<int64_t>(st.C, 1));  // list keys: (target << sbits) | source
  int64_t cap_f = std::max<int64_t>(4096, 8 * (tt.C + st.C));
  int64_t cap_p = std::max<int64_t>(4096, 32 * tt.nleaf);
  int64_t cap_m = std::max<int64_t>(4096, 64 * tt.C);
  if (const char *e = std::getenv("UVLM_LIST_CAPACITY")) {  // first sizes of fresh buffers (tests of the regrowth)
    const long long v = std::atoll(e);
    if (v > 0) cap_f = cap_p = cap_m = v;
  }
  sc.front0.reserve(8 * cap_f);
  sc.front1.reserve(8 * cap_f);
  sc.out_p2p.reserve(8 * cap_p);
  sc.out_m2l.reserve(8 * cap_m);
  cap_f = std::min(sc.front0.bytes, sc.front1.bytes) / 8;
  cap_p = sc.out_p2p.bytes / 8;
  cap_m = sc.out_m2l.bytes / 8;
  unsigned long long cnt[6] = {0, 0, 0, 0, 0, 0};
  sc.counters.reserve(64);
  unsigned long long *dcnt = (unsigned long long *)sc.counters.p;
  {
    // steps run in batches of STEPS without the host; a batch ends with one read of the counters
    constexpr int STEPS = 6;
    for (;;) {
      k_traverse_start<<<1, 64, 0, s>>>(T, S, sc.front0.as<int32_t>(), sc.out_p2p.as<uint64_t>(), dcnt);
      const int grid = nblocks(std::min<int64_t>(cap_f, 2048 * (int64_t)BLK));
      int       par  = 0;
      for (;;) {
        for (int k = 0; k < STEPS; k++, par ^= 1) {
          DevBuf &fin = par ? sc.front1 : sc.front0, &fout = par ? sc.front0 : sc.front1;
          k_traverse<<<grid, BLK, 0, s>>>(fin.as<int32_t>(), T, S, theta, gap, fout.as<int32_t>(), cap_f,
                                          sc.out_p2p.as<uint64_t>(), cap_p, sc.out_m2l.as<uint64_t>(), cap_m, dcnt,
                                          sbits);
          k_step_end<<<1, 64, 0, s>>>(dcnt, cap_f, cap_p, cap_m);
          UVLM_LAUNCH_CHECK();
        }
        to_host(s, cnt, dcnt, 6);
        if (cnt[4] || cnt[3] == 0) break;
      }
      if (!cnt[4]) break;
      // a buffer overflowed: grow what did (at least doubled) and traverse again
      if (cnt[4] & 1) {
        const int64_t want = std::max<int64_t>(2 * cap_f, 2 * (int64_t)cnt[5]);
        sc.front0.reserve(8 * want);
        sc.front1.reserve(8 * want);
      }
      if (cnt[4] & 2) sc.out_p2p.reserve(8 * std::max<int64_t>(2 * cap_p, 2 * (int64_t)cnt[1]));
      if (cnt[4] & 4) sc.out_m2l.reserve(8 * std::max<int64_t>(2 * cap_m, 2 * (int64_t)cnt[2]));
      cap_f = std::min(sc.front0.bytes, sc.front1.bytes) / 8;
      cap_p = sc.out_p2p.bytes / 8;
      cap_m = sc.out_m2l.bytes / 8;
    }
  }
  lists.np2p = (int64_t)cnt[1];
  lists.nm2l = (int64_t)cnt[2];
  const int64_t C  = tt.C;
  const int     eb = sbits + bits_for(std::max<int64_t>(tt.C, 1));
  for (int which = 0; which < 2; which++) {
    const int64_t n   = which == 0 ? lists.np2p : lists.nm2l;
    DevBuf       &raw = which == 0 ? sc.out_p2p : sc.out_m2l;
    DevBuf       &off = which == 0 ? lists.p2p_off : lists.m2l_off;
    DevBuf       &src = which == 0 ? lists.p2p_src : lists.m2l_src;
    off.reserve(4 * (C + 1));
    src.reserve(4 * std::max<int64_t>(n, 1));
    sc.keys1.reserve(8 * std::max<int64_t>(n, 1));
    if (n > 0) sort_keys_u64(s, sc, raw.as<uint64_t>(), sc.keys1.as<uint64_t>(), n, eb);
    if (which == 0) {  // interactions of the P2P list, read by count_interactions
      unsigned long long *d = (unsigned long long *)sc.counters.p + 7;  // slots 0..5: the traversal's
      UVLM_HIP(hipMemsetAsync(d, 0, 8, s));
      if (n > 0)
        k_count_inter<<<std::min(nblocks(n), COUNT_BLOCKS), BLK, 0, s>>>(n, sc.keys1.as<uint64_t>(), sbits,
                                                                         tt.nbody.as<int32_t>(),
                                                                         st.nbody.as<int32_t>(), d);
    }
    k_csr<<<nblocks(C + 1), BLK, 0, s>>>(C, sc.keys1.as<uint64_t>(), n, sbits, off.as<int32_t>());
    k_low_bits<<<nblocks(n), BLK, 0, s>>>(n, sc.keys1.as<uint64_t>(), sbits, src.as<int32_t>());
    UVLM_LAUNCH_CHECK();
  }
}

void transpose_lists(hipStream_t s, Scratch &sc, const Tree &in_targets, const Tree &in_sources, const Lists &in,
                     Lists &out)
{
  const int64_t Cs    = in_sources.C;
  const int     tbits = bits_for(std::max<int64_t>(in_targets.C, 1));  // entry (target cell) bits of a key
  const int     eb    = tbits + bits_for(std::max<int64_t>(Cs, 1));
  out.np2p         = in.np2p;
  out.nm2l         = in.nm2l;
  for (int which = 0; which < 2; which++) {
    const int64_t n    = which == 0 ? in.np2p : in.nm2l;
    const DevBuf &ioff = which == 0 ? in.p2p_off : in.m2l_off;
    const DevBuf &isrc = which == 0 ? in.p2p_src : in.m2l_src;
    DevBuf       &off  = which == 0 ? out.p2p_off : out.m2l_off;
    DevBuf       &src  = which == 0 ? out.p2p_src : out.m2l_src;
    off.reserve(4 * (Cs + 1));
    src.reserve(4 * std::max<int64_t>(n, 1));
    sc.keys0.reserve(8 * std::max<int64_t>(n, 1));
    sc.keys1.reserve(8 * std::max<int64_t>(n, 1));
    if (n > 0) {
      k_transpose_keys<<<nblocks(n), BLK, 0, s>>>(in_targets.C, n, ioff.as<int32_t>(), isrc.as<int32_t>(), tbits,
                                                  sc.keys0.as<uint64_t>());
      UVLM_LAUNCH_CHECK();
    }
    if (n > 0) sort_keys_u64(s, sc, sc.keys0.as<uint64_t>(), sc.keys1.as<uint64_t>(), n, eb);
    k_csr<<<nblocks(Cs + 1), BLK, 0, s>>>(Cs, sc.keys1.as<uint64_t>(), n, tbits, off.as<int32_t>());
    k_low_bits<<<nblocks(n), BLK, 0, s>>>(n, sc.keys1.as<uint64_t>(), tbits, src.as<int32_t>());
    UVLM_LAUNCH_CHECK();
  }
}

void count_interactions(hipStream_t s, Scratch &sc, const Tree &tt, const Tree &st, const Lists &l, int64_t *out)
{
  // summed by traverse() from the sorted P2P keys of these trees' lists
  (void)tt, (void)st, (void)l;
  const unsigned long long *d = (const unsigned long long *)sc.counters.p + 7;
  unsigned long long h = 0;
  to_host(s, &h, d);
  *out = (int64_t)h;
}

}  // namespace uvlm
