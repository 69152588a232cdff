// coarse.hip -- exact solve of a multigrid level of any size: the level's
// assembled 3x3-block stencil factorised by a multifrontal Cholesky on a
// geometric nested-dissection ordering of the Q2 node lattice.
//
// Replaces the reference's assembled coarse operator and direct solve
// (MatAssemble_StokesA_AUU, stokes_assembly.c:115-297, with
// -fieldsplit_u_mg_coarse_pc_type lu, ptatin_driver_linear_ts.c:629-671).
//
// Nodes whose three dofs are all Dirichlet have identity rows and zero
// columns: they are left out of every front (x = b there), which keeps the
// factor to the free part of the operator.
//
// Ordering.  Nodes of the Q2 lattice couple only inside an element, so a
// plane of nodes at an even index (an element boundary) separates the nodes
// on its two sides.  The lattice box is bisected recursively by such planes
// (largest extent first, the even index nearest the middle) down to leaves
// of at most kLeafNodes nodes; every separator plane and every leaf is a
// front.  A front's dofs are its own nodes S followed by the boundary B of
// its subtree's box (the lattice nodes of the box grown by one, minus the
// box), which lie on ancestor separators.
//
// Factorisation (set-up, once): the fronts of one tree depth are padded to a
// common size and processed together, deepest first:
//   F = [A_SS; A_BS] from the stencil, plus the children's update matrices
//   (extend-add); L_SS = chol(F_SS); L_BS = F_BS L_SS^-T;
//   U = F_BB - L_BS L_BS^T (passed to the parent);
//   G = [L_SS^-1; L_BS L_SS^-1] kept for the solve.
// Dense kernels are rocSOLVER/rocBLAS strided-batched calls.
//
// Solve (every V-cycle; project kernels only, no allocation, capturable):
//   forward, deepest depth first: [y_S; u_B] = G w_S, w_S = y_S,
//     w_B -= u_B (summed over the depth's fronts in a fixed order);
//   backward, root first: x_S = G^T [y_S; -x_B].
// Each depth is three launches forward and two backward.  All sums have a
// fixed order, so repeated solves are bit-identical.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <unordered_map>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "ptmf_internal.h"

namespace ptmf {
namespace {

constexpr int kLeafNodes = 64;
constexpr int kFwdRows = 256;   // rows per work-group of the forward product
constexpr int kFwdCols = 256;   // largest column chunk of the forward product
constexpr long kFwdMinGroups = 2048;  // work-groups a depth's forward product aims for
constexpr int kBS = 256;

inline unsigned grid1(long n) {
  long g = (n + kBS - 1) / kBS;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

// ---------------------------------------------------------------------------
// kernels: assembly and extend-add (set-up)
// ---------------------------------------------------------------------------
// padded separator dofs (sdof < 0) get a unit diagonal
__global__ void k_mf_pad_diag(int nfr, int nf, int ns, const long *__restrict__ sdof, double *__restrict__ F) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (long)nfr * ns; t += (long)gridDim.x * blockDim.x) {
    const long f = t / ns;
    const int i = (int)(t % ns);
    if (sdof[t] < 0) F[f * (long)nf * nf + i + (long)i * nf] = 1.0;
  }
}

// F[f][i, j] += A(row i, column j) for the separator columns j = 3p + c of
// every front: row dofs of the 125 neighbours of the column's node that lie
// in the front (nbr[(f * nsn + p) * 125 + o], a local row base or -1).  The
// stencil holds rows: S[((o * 9) + 3 r + c) * N + m] = A(3m + r, 3(m + off_o) + c),
// and A is symmetric, so A(3m' + r, 3n + c) = S[((o * 9) + 3c + r) * N + n]
// for the neighbour m' = n + off_o.  Each entry has one writer (distinct
// neighbours have distinct rows); += keeps the children's updates already
// added into the front.
__global__ void k_mf_assemble(long N, const double *__restrict__ S, int nfr, int nf, int ns,
                              const long *__restrict__ sdof, const int *__restrict__ nbr,
                              double *__restrict__ F) {
  const int nsn = ns / 3;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (long)nfr * nsn * 125;
       t += (long)gridDim.x * blockDim.x) {
    const int o = (int)(t % 125);
    const long fp = t / 125;  // f * nsn + p
    const long f = fp / nsn;
    const int p = (int)(fp % nsn);
    const long g = sdof[f * ns + 3 * p];
    const int li = nbr[t];
    if (g < 0 || li < 0) continue;
    const long n = g / 3;
    double *Ff = F + f * (long)nf * nf;
    for (int c = 0; c < 3; c++)
      for (int r = 0; r < 3; r++) Ff[(li + r) + (long)(3 * p + c) * nf] += S[((long)o * 9 + 3 * c + r) * N + n];
  }
}

// add the update matrices U = F_c[nsc:, nsc:] (lower triangle) of the child
// fronts in child slot `which` into their parents' fronts, in the parents'
// lower triangle (c2p: the parent-local index of every child boundary row)
__global__ void k_mf_extend_add(int nfc, int nfc_f, int nsc, int nbc, const double *__restrict__ Fc,
                                const int *__restrict__ c2p, const int *__restrict__ parent,
                                const int *__restrict__ slot, int which, int nfp, double *__restrict__ Fp) {
  const long per = (long)nbc * nbc;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (long)nfc * per; t += (long)gridDim.x * blockDim.x) {
    const long f = t / per;
    const int a = (int)((t % per) % nbc), b = (int)((t % per) / nbc);
    if (b > a || slot[f] != which) continue;
    const int la = c2p[f * nbc + a], lb = c2p[f * nbc + b];
    if (la < 0 || lb < 0) continue;
    const double v = Fc[f * (long)nfc_f * nfc_f + (nsc + a) + (long)(nsc + b) * nfc_f];
    const int i = la > lb ? la : lb, j = la > lb ? lb : la;
    Fp[(long)parent[f] * nfp * nfp + i + (long)j * nfp] += v;
  }
}

// G keeps L_SS^-1 (lower) on top: clear the strictly upper part of that block
__global__ void k_mf_zero_upper(int nfr, int nf, int ns, double *__restrict__ G) {
  const long per = (long)ns * ns;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (long)nfr * per; t += (long)gridDim.x * blockDim.x) {
    const long f = t / per;
    const int i = (int)((t % per) % ns), j = (int)((t % per) / ns);
    if (i < j) G[f * (long)nf * ns + i + (long)j * nf] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// kernels: solve
// ---------------------------------------------------------------------------
// w = b and x = b (x keeps b on the dofs of all-Dirichlet nodes, which no
// front holds; every other entry of x is written by the backward pass)
__global__ void k_mf_copy(long n, const double *__restrict__ b, double *__restrict__ w, double *__restrict__ x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double v = b[i];
    w[i] = v;
    x[i] = v;
  }
}

// partial sums of [y_S; u_B] = G w_S: work-group (row block, column chunk
// of cw <= kFwdCols columns, front); rows of the triangular top block stop at
// the diagonal.  Every (chunk, row) partial is written, zero where the chunk
// lies right of the row's diagonal.
__global__ void __launch_bounds__(kFwdRows)
    k_mf_fwd_part(int nf, int ns, int ncc, int cw, const double *__restrict__ G, const long *__restrict__ sdof,
                  const double *__restrict__ w, double *__restrict__ part) {
  __shared__ double xs[kFwdCols];
  const long f = blockIdx.z;
  const int j0 = blockIdx.y * cw;
  const int jn = min(cw, ns - j0);
  const long *sd = sdof + f * ns;
  for (int t = threadIdx.x; t < jn; t += kFwdRows) {
    const long g = sd[j0 + t];
    xs[t] = g >= 0 ? w[g] : 0.0;
  }
  __syncthreads();
  const int i = blockIdx.x * kFwdRows + threadIdx.x;
  if (i >= nf) return;
  int je = jn;
  if (i < ns) je = min(jn, i - j0 + 1);  // G[i, j] = 0 for j > i in the top block
  const double *a = G + f * (long)nf * ns + i + (long)j0 * nf;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int j = 0;
  for (; j + 4 <= je; j += 4) {
    s0 += a[(long)j * nf] * xs[j];
    s1 += a[(long)(j + 1) * nf] * xs[j + 1];
    s2 += a[(long)(j + 2) * nf] * xs[j + 2];
    s3 += a[(long)(j + 3) * nf] * xs[j + 3];
  }
  for (; j < je; j++) s0 += a[(long)j * nf] * xs[j];
  part[(f * ncc + blockIdx.y) * (long)nf + i] = (s0 + s1) + (s2 + s3);
}

// y_S into w, u_B into the front's boundary buffer
__global__ void k_mf_fwd_sum(int nfr, int nf, int ns, int nb, int ncc, const double *__restrict__ part,
                             const long *__restrict__ sdof, double *__restrict__ w, double *__restrict__ ub) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (long)nfr * nf; t += (long)gridDim.x * blockDim.x) {
    const long f = t / nf;
    const int i = (int)(t % nf);
    double s = 0.0;
    for (int c = 0; c < ncc; c++) s += part[(f * ncc + c) * (long)nf + i];
    if (i < ns) {
      const long g = sdof[f * ns + i];
      if (g >= 0) w[g] = s;
    } else {
      ub[f * nb + (i - ns)] = s;
    }
  }
}

// w_B -= sum of the fronts' u_B, one thread per boundary dof of the depth
__global__ void k_mf_fwd_scatter(long nd, const long *__restrict__ dof, const int *__restrict__ ptr,
                                 const int *__restrict__ ent, const double *__restrict__ ub, double *__restrict__ w) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nd; t += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = ptr[t]; k < ptr[t + 1]; k++) s += ub[ent[k]];
    w[dof[t]] -= s;
  }
}

// v = [y_S; -x_B] of every front of a depth
__global__ void k_mf_bwd_gather(int nfr, int nf, int ns, int nb, const long *__restrict__ sdof,
                                const long *__restrict__ bdof, const double *__restrict__ w,
                                const double *__restrict__ x, double *__restrict__ v) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (long)nfr * nf; t += (long)gridDim.x * blockDim.x) {
    const long f = t / nf;
    const int i = (int)(t % nf);
    double val = 0.0;
    if (i < ns) {
      const long g = sdof[f * ns + i];
      if (g >= 0) val = w[g];
    } else {
      const long g = bdof[f * nb + (i - ns)];
      if (g >= 0) val = -x[g];
    }
    v[t] = val;
  }
}

// x_S = G^T v: one wave per column (rows from the diagonal down, four
// independent partial sums per lane), four columns per work-group
__global__ void __launch_bounds__(256)
    k_mf_bwd(int nf, int ns, const double *__restrict__ G, const long *__restrict__ sdof,
             const double *__restrict__ v, double *__restrict__ x) {
  const long f = blockIdx.y;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= ns) return;
  const double *g = G + f * (long)nf * ns + (long)j * nf;
  const double *vf = v + f * (long)nf;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int i = j + lane;
  for (; i + 192 < nf; i += 256) {
    s0 += g[i] * vf[i];
    s1 += g[i + 64] * vf[i + 64];
    s2 += g[i + 128] * vf[i + 128];
    s3 += g[i + 192] * vf[i + 192];
  }
  for (; i < nf; i += 64) s0 += g[i] * vf[i];
  double s = (s0 + s1) + (s2 + s3);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) {
    const long gd = sdof[f * ns + j];
    if (gd >= 0) x[gd] = s;
  }
}

// ---------------------------------------------------------------------------
// host: nested-dissection tree
// ---------------------------------------------------------------------------
struct NDFront {
  int lo[3], hi[3];    // box of the subtree (inclusive node ranges)
  int slo[3], shi[3];  // the front's own nodes (separator plane or leaf box)
  int parent = -1, slot = 0, depth = 0;
  int child[2] = {-1, -1};
  std::vector<long> snodes, bnodes;  // lattice node ids, x fastest
};

long box_count(const int lo[3], const int hi[3]) {
  return (long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
}

int nd_build(std::vector<NDFront> &fr, const int lo[3], const int hi[3], int depth, int parent, int slot) {
  const int id = (int)fr.size();
  fr.emplace_back();
  {
    NDFront &F = fr[id];
    memcpy(F.lo, lo, sizeof F.lo);
    memcpy(F.hi, hi, sizeof F.hi);
    F.depth = depth;
    F.parent = parent;
    F.slot = slot;
  }
  int dsplit = -1, best_ext = 0, s = -1;
  if (box_count(lo, hi) > kLeafNodes) {
    for (int d = 0; d < 3; d++) {
      const int ext = hi[d] - lo[d] + 1;
      // an even index strictly inside the range, nearest the middle
      int cand = -1, bal = 1 << 30;
      for (int v = lo[d] + 1; v < hi[d]; v++) {
        if (v % 2) continue;
        const int b = abs((v - lo[d]) - (hi[d] - v));
        if (b < bal) {
          bal = b;
          cand = v;
        }
      }
      if (cand >= 0 && ext > best_ext) {
        best_ext = ext;
        dsplit = d;
        s = cand;
      }
    }
  }
  if (dsplit < 0) {  // leaf: the whole box is the front's own node set
    memcpy(fr[id].slo, lo, sizeof(int) * 3);
    memcpy(fr[id].shi, hi, sizeof(int) * 3);
    return id;
  }
  int slo[3], shi[3], l1[3], h1[3], l2[3], h2[3];
  for (int d = 0; d < 3; d++) {
    slo[d] = lo[d]; shi[d] = hi[d];
    l1[d] = lo[d]; h1[d] = hi[d];
    l2[d] = lo[d]; h2[d] = hi[d];
  }
  slo[dsplit] = shi[dsplit] = s;
  h1[dsplit] = s - 1;
  l2[dsplit] = s + 1;
  memcpy(fr[id].slo, slo, sizeof slo);
  memcpy(fr[id].shi, shi, sizeof shi);
  const int c0 = nd_build(fr, l1, h1, depth + 1, id, 0);
  const int c1 = nd_build(fr, l2, h2, depth + 1, id, 1);
  fr[id].child[0] = c0;
  fr[id].child[1] = c1;
  return id;
}

}  // namespace

// One depth of the tree: its fronts padded to common sizes.
struct MFGroup {
  int nfr = 0, ns = 0, nb = 0, nf = 0, ncc = 0, cw = kFwdCols;
  std::vector<int> ids;  // front ids of this depth
  double *F = nullptr;   // nfr x nf x nf, set-up only
  double *G = nullptr;   // nfr x nf x ns
  long *sdof = nullptr, *bdof = nullptr;
  int *nbr = nullptr;  // nfr x (ns/3) x 125 local row bases (set-up only)
  int *c2p = nullptr, *parent = nullptr, *slot = nullptr;  // as a child of depth - 1 (set-up only)
  long nscat = 0;
  long *sc_dof = nullptr;
  int *sc_ptr = nullptr, *sc_ent = nullptr;
  double *ub = nullptr, *part = nullptr, *v = nullptr;
};

struct MFFactor {
  long n = 0;  // 3 * lattice nodes
  std::vector<MFGroup> groups;  // by depth, root first
  double *w = nullptr;
  long factor_doubles = 0;
  double setup_ms = 0.0;
};

namespace {

template <typename T>
int upload(T **d, const std::vector<T> &h) {
  *d = nullptr;
  if (h.empty()) return PTMF_OK;
  PTMF_TRY(ptmf_malloc((void **)d, sizeof(T) * h.size()));
  PTMF_HIP(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return PTMF_OK;
}

void free_group_setup(MFGroup &g) {
  (void)hipFree(g.F);
  (void)hipFree(g.nbr);
  (void)hipFree(g.c2p);
  (void)hipFree(g.parent);
  (void)hipFree(g.slot);
  g.F = nullptr;
  g.nbr = g.c2p = g.parent = g.slot = nullptr;
}

#define MF_BLAS(call)                                                                              \
  do {                                                                                             \
    rocblas_status st_ = (call);                                                                   \
    if (st_ != rocblas_status_success) {                                                           \
      rc = ptmf_set_error(PTMF_ERR_DEVICE, "coarse factor: %s: rocblas status %d", #call, (int)st_); \
      goto done;                                                                                   \
    }                                                                                              \
  } while (0)

}  // namespace

void mf_destroy(MFFactor *m) {
  if (!m) return;
  for (MFGroup &g : m->groups) {
    free_group_setup(g);
    void *ps[] = {g.G, g.sdof, g.bdof, g.sc_dof, g.sc_ptr, g.sc_ent, g.ub, g.part, g.v};
    for (void *p : ps) (void)hipFree(p);
  }
  (void)hipFree(m->w);
  delete m;
}

long mf_factor_doubles(const MFFactor *m) { return m ? m->factor_doubles : 0; }
double mf_setup_ms(const MFFactor *m) { return m ? m->setup_ms : 0.0; }
int mf_depths(const MFFactor *m) { return m ? (int)m->groups.size() : 0; }

int mf_factor(const Box &box, const double *S, const uint8_t *bc_host, hipStream_t stream, MFFactor **out) {
  *out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PTMF_HIP(hipEventCreate(&e0));
  PTMF_HIP(hipEventCreate(&e1));
  PTMF_HIP(hipEventRecord(e0, stream));
  const long X = box.X, Y = box.Y, Z = box.Z, N = box.nnodes;
  // ---- tree and front node sets ------------------------------------------
  std::vector<NDFront> fr;
  fr.reserve(4 * N / kLeafNodes + 64);
  {
    const int lo[3] = {0, 0, 0}, hi[3] = {(int)X - 1, (int)Y - 1, (int)Z - 1};
    nd_build(fr, lo, hi, 0, -1, 0);
  }
  int maxd = 0;
  auto kept = [&](long n) { return !bc_host || (bc_host[n] & 7u) != 7u; };
  for (NDFront &F : fr) {
    maxd = std::max(maxd, F.depth);
    for (int k = F.slo[2]; k <= F.shi[2]; k++)
      for (int j = F.slo[1]; j <= F.shi[1]; j++)
        for (int i = F.slo[0]; i <= F.shi[0]; i++) {
          const long n = i + X * (j + Y * (long)k);
          if (kept(n)) F.snodes.push_back(n);
        }
    int el[3], eh[3];
    for (int d = 0; d < 3; d++) {
      const long ext = d == 0 ? X : (d == 1 ? Y : Z);
      el[d] = std::max(F.lo[d] - 1, 0);
      eh[d] = (int)std::min<long>(F.hi[d] + 1, ext - 1);
    }
    for (int k = el[2]; k <= eh[2]; k++)
      for (int j = el[1]; j <= eh[1]; j++)
        for (int i = el[0]; i <= eh[0]; i++) {
          const bool in = i >= F.lo[0] && i <= F.hi[0] && j >= F.lo[1] && j <= F.hi[1] && k >= F.lo[2] &&
                          k <= F.hi[2];
          const long n = i + X * (j + Y * (long)k);
          if (!in && kept(n)) F.bnodes.push_back(n);
        }
  }
  MFFactor *m = new MFFactor();
  m->n = 3 * N;
  m->groups.resize(maxd + 1);
  std::vector<int> pos_in_group(fr.size());
  for (int id = 0; id < (int)fr.size(); id++) {
    MFGroup &g = m->groups[fr[id].depth];
    pos_in_group[id] = (int)g.ids.size();
    g.ids.push_back(id);
    g.ns = std::max(g.ns, 3 * (int)fr[id].snodes.size());
    g.nb = std::max(g.nb, 3 * (int)fr[id].bnodes.size());
  }
  int rc = PTMF_OK;
  rocblas_handle blas = nullptr;
  int *dinfo = nullptr;
  for (MFGroup &g : m->groups) {
    g.nfr = (int)g.ids.size();
    g.nf = g.ns + g.nb;
    // narrower column chunks where a depth has few or small fronts, so the
    // forward product has enough work-groups to stream its factor block
    g.cw = kFwdCols;
    const long rows = (long)g.nfr * ((g.nf + kFwdRows - 1) / kFwdRows);
    while (g.cw > 32 && rows * ((g.ns + g.cw - 1) / g.cw) < kFwdMinGroups) g.cw /= 2;
    g.ncc = (g.ns + g.cw - 1) / g.cw;
  }
  // ---- index tables --------------------------------------------------------
  // local dof base of every node of a front: S node p -> 3p, B node q -> ns + 3q
  auto local_map = [&](int id) {
    std::unordered_map<long, int> mp;
    const NDFront &F = fr[id];
    const int ns = m->groups[F.depth].ns;
    mp.reserve(2 * (F.snodes.size() + F.bnodes.size()));
    for (size_t p = 0; p < F.snodes.size(); p++) mp[F.snodes[p]] = 3 * (int)p;
    for (size_t q = 0; q < F.bnodes.size(); q++) mp[F.bnodes[q]] = ns + 3 * (int)q;
    return mp;
  };
  for (int d = 0; d <= maxd && rc == PTMF_OK; d++) {
    MFGroup &g = m->groups[d];
    const int nsn = g.ns / 3;
    std::vector<long> sdof((size_t)g.nfr * g.ns, -1), bdof((size_t)g.nfr * g.nb, -1);
    std::vector<int> nbr((size_t)g.nfr * nsn * 125, -1);
    std::vector<int> c2p((size_t)g.nfr * g.nb, -1), par(g.nfr, 0), slt(g.nfr, 0);
    std::vector<std::pair<long, int>> scat;
    for (int f = 0; f < g.nfr; f++) {
      const int id = g.ids[f];
      const NDFront &F = fr[id];
      for (size_t p = 0; p < F.snodes.size(); p++)
        for (int c = 0; c < 3; c++) sdof[(size_t)f * g.ns + 3 * p + c] = 3 * F.snodes[p] + c;
      for (size_t q = 0; q < F.bnodes.size(); q++)
        for (int c = 0; c < 3; c++) {
          bdof[(size_t)f * g.nb + 3 * q + c] = 3 * F.bnodes[q] + c;
          scat.emplace_back(3 * F.bnodes[q] + c, f * g.nb + 3 * (int)q + c);
        }
      const std::unordered_map<long, int> mp = local_map(id);
      for (size_t p = 0; p < F.snodes.size(); p++) {
        const long n = F.snodes[p];
        const long i = n % X, j = (n / X) % Y, k = n / (X * Y);
        for (int o = 0; o < 125; o++) {
          const int oi = o % 5 - 2, oj = (o / 5) % 5 - 2, ok = o / 25 - 2;
          if (i + oi < 0 || i + oi >= X || j + oj < 0 || j + oj >= Y || k + ok < 0 || k + ok >= Z) continue;
          const auto it = mp.find(n + oi + X * (oj + Y * (long)ok));
          if (it != mp.end()) nbr[((size_t)f * nsn + p) * 125 + o] = it->second;
        }
      }
      if (F.parent >= 0) {
        const std::unordered_map<long, int> pm = local_map(F.parent);
        par[f] = pos_in_group[F.parent];
        slt[f] = F.slot;
        for (size_t q = 0; q < F.bnodes.size(); q++) {
          const auto it = pm.find(F.bnodes[q]);
          if (it == pm.end()) {
            rc = ptmf_set_error(PTMF_ERR_DEVICE, "coarse factor: boundary node %ld outside the parent front",
                                F.bnodes[q]);
            break;
          }
          for (int c = 0; c < 3; c++) c2p[(size_t)f * g.nb + 3 * q + c] = it->second + c;
        }
      }
    }
    if (rc != PTMF_OK) break;
    std::sort(scat.begin(), scat.end());
    std::vector<long> sc_dof;
    std::vector<int> sc_ptr, sc_ent;
    for (size_t t = 0; t < scat.size(); t++) {
      if (t == 0 || scat[t].first != scat[t - 1].first) {
        sc_dof.push_back(scat[t].first);
        sc_ptr.push_back((int)t);
      }
      sc_ent.push_back(scat[t].second);
    }
    sc_ptr.push_back((int)scat.size());
    g.nscat = (long)sc_dof.size();
    if ((rc = upload(&g.sdof, sdof)) != PTMF_OK) break;
    if ((rc = upload(&g.bdof, bdof)) != PTMF_OK) break;
    if ((rc = upload(&g.nbr, nbr)) != PTMF_OK) break;
    if ((rc = upload(&g.c2p, c2p)) != PTMF_OK) break;
    if ((rc = upload(&g.parent, par)) != PTMF_OK) break;
    if ((rc = upload(&g.slot, slt)) != PTMF_OK) break;
    if ((rc = upload(&g.sc_dof, sc_dof)) != PTMF_OK) break;
    if ((rc = upload(&g.sc_ptr, sc_ptr)) != PTMF_OK) break;
    if ((rc = upload(&g.sc_ent, sc_ent)) != PTMF_OK) break;
    // a depth whose fronts hold no free node (ns == 0) only passes its
    // children's updates on: no solve factor, no solve work
    if (g.ns > 0) {
      if ((rc = ptmf_malloc((void **)&g.G, sizeof(double) * (size_t)g.nfr * g.nf * g.ns)) != PTMF_OK) break;
      if (g.nb > 0 && (rc = ptmf_malloc((void **)&g.ub, sizeof(double) * (size_t)g.nfr * g.nb)) != PTMF_OK) break;
      if ((rc = ptmf_malloc((void **)&g.part, sizeof(double) * (size_t)g.nfr * g.ncc * g.nf)) != PTMF_OK) break;
      if ((rc = ptmf_malloc((void **)&g.v, sizeof(double) * (size_t)g.nfr * g.nf)) != PTMF_OK) break;
    }
    m->factor_doubles += (long)g.nfr * g.nf * g.ns;
  }
  if (rc == PTMF_OK) rc = ptmf_malloc((void **)&m->w, sizeof(double) * m->n);
  if (rc != PTMF_OK) {
    mf_destroy(m);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
  }
  // ---- numeric factorisation, deepest depth first --------------------------
  {
    rocblas_status bs = rocblas_create_handle(&blas);
    if (bs != rocblas_status_success) {
      rc = ptmf_set_error(PTMF_ERR_DEVICE, "coarse factor: rocblas_create_handle: %d", (int)bs);
      goto done;
    }
    rocblas_set_stream(blas, stream);
    int maxfr = 1;
    for (MFGroup &g : m->groups) maxfr = std::max(maxfr, g.nfr);
    if ((rc = ptmf_malloc((void **)&dinfo, sizeof(int) * maxfr)) != PTMF_OK) goto done;
    const double one = 1.0, mone = -1.0;
    for (int d = maxd; d >= 0; d--) {
      MFGroup &g = m->groups[d];
      const long nff = (long)g.nf * g.nf;
      // F = [A_SS; A_BS] (+ unit diagonal on padded separator dofs); the
      // children's update matrices were added below when their depth ran
      if (d == maxd) {
        if ((rc = ptmf_malloc((void **)&g.F, sizeof(double) * g.nfr * nff)) != PTMF_OK) goto done;
        if (hipMemsetAsync(g.F, 0, sizeof(double) * g.nfr * nff, stream) != hipSuccess) {
          rc = ptmf_set_error(PTMF_ERR_DEVICE, "coarse factor: memset");
          goto done;
        }
      }
      k_mf_pad_diag<<<grid1((long)g.nfr * g.ns), kBS, 0, stream>>>(g.nfr, g.nf, g.ns, g.sdof, g.F);
      k_mf_assemble<<<grid1((long)g.nfr * (g.ns / 3) * 125), kBS, 0, stream>>>(N, S, g.nfr, g.nf, g.ns, g.sdof,
                                                                                   g.nbr, g.F);
      if (hipGetLastError() != hipSuccess) {
        rc = ptmf_set_error(PTMF_ERR_DEVICE, "coarse factor: assembly launch");
        goto done;
      }
      // L_SS = chol(F_SS)
      if (g.ns > 0) {
      MF_BLAS(rocsolver_dpotrf_strided_batched(blas, rocblas_fill_lower, g.ns, g.F, g.nf, nff, dinfo, g.nfr));
      {
        std::vector<int> info(g.nfr);
        if (hipMemcpyAsync(info.data(), dinfo, sizeof(int) * g.nfr, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
          rc = ptmf_set_error(PTMF_ERR_DEVICE, "coarse factor: info copy");
          goto done;
        }
        for (int f = 0; f < g.nfr; f++)
          if (info[f] != 0) {
            rc = ptmf_set_error(PTMF_ERR_DIVERGED,
                                "coarse factor: front %d of depth %d not positive definite (pivot %d): "
                                "the level operator is not SPD (a rigid-body mode left free?)",
                                f, d, info[f]);
            goto done;
          }
      }
      if (g.nb > 0) {
        // L_BS = F_BS L_SS^-T, U = F_BB - L_BS L_BS^T, then M_BS = L_BS L_SS^-1
        MF_BLAS(rocblas_dtrsm_strided_batched(blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                              rocblas_diagonal_non_unit, g.nb, g.ns, &one, g.F, g.nf, nff, g.F + g.ns,
                                              g.nf, nff, g.nfr));
        MF_BLAS(rocblas_dsyrk_strided_batched(blas, rocblas_fill_lower, rocblas_operation_none, g.nb, g.ns, &mone,
                                              g.F + g.ns, g.nf, nff, &one, g.F + g.ns + (long)g.ns * g.nf, g.nf, nff,
                                              g.nfr));
        MF_BLAS(rocblas_dtrsm_strided_batched(blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_none,
                                              rocblas_diagonal_non_unit, g.nb, g.ns, &one, g.F, g.nf, nff, g.F + g.ns,
                                              g.nf, nff, g.nfr));
      }
      // L_SS^-1 in place
      MF_BLAS(rocsolver_dtrtri_strided_batched(blas, rocblas_fill_lower, rocblas_diagonal_non_unit, g.ns, g.F, g.nf,
                                               nff, dinfo, g.nfr));
      // G = first ns columns, strictly upper part of its top block cleared
      if (hipMemcpy2DAsync(g.G, sizeof(double) * g.nf * g.ns, g.F, sizeof(double) * nff,
                           sizeof(double) * g.nf * g.ns, g.nfr, hipMemcpyDeviceToDevice, stream) != hipSuccess) {
        rc = ptmf_set_error(PTMF_ERR_DEVICE, "coarse factor: copy of the solve factor");
        goto done;
      }
      k_mf_zero_upper<<<grid1((long)g.nfr * g.ns * g.ns), kBS, 0, stream>>>(g.nfr, g.nf, g.ns, g.G);
      }  // ns > 0
      // extend-add into the parents (depth d - 1), one child slot at a time
      if (d > 0) {
        MFGroup &p = m->groups[d - 1];
        // fronts are allocated one depth ahead and freed once their parents
        // hold the updates: two depths are resident at a time
        if ((rc = ptmf_malloc((void **)&p.F, sizeof(double) * p.nfr * (long)p.nf * p.nf)) != PTMF_OK) goto done;
        if (hipMemsetAsync(p.F, 0, sizeof(double) * p.nfr * (long)p.nf * p.nf, stream) != hipSuccess) {
          rc = ptmf_set_error(PTMF_ERR_DEVICE, "coarse factor: memset");
          goto done;
        }
        for (int which = 0; which < 2; which++)
          k_mf_extend_add<<<grid1((long)g.nfr * g.nb * g.nb), kBS, 0, stream>>>(
              g.nfr, g.nf, g.ns, g.nb, g.F, g.c2p, g.parent, g.slot, which, p.nf, p.F);
      }
      if (hipGetLastError() != hipSuccess) {
        rc = ptmf_set_error(PTMF_ERR_DEVICE, "coarse factor: kernel launch");
        goto done;
      }
      if (hipStreamSynchronize(stream) != hipSuccess) {
        rc = ptmf_set_error(PTMF_ERR_DEVICE, "coarse factor: depth %d", d);
        goto done;
      }
      free_group_setup(g);  // the parent has its update; keep G only
    }
  }
done:
  if (blas) rocblas_destroy_handle(blas);
  (void)hipFree(dinfo);
  if (rc == PTMF_OK) {
    (void)hipEventRecord(e1, stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    m->setup_ms = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != PTMF_OK) {
    mf_destroy(m);
    return rc;
  }
  *out = m;
  return PTMF_OK;
}

// x = A^-1 b with the factor (kernels only: capturable, allocates nothing)
int mf_solve(const MFFactor *m, const double *b, double *x, hipStream_t s) {
  const int nd = (int)m->groups.size();
  k_mf_copy<<<grid1(m->n), kBS, 0, s>>>(m->n, b, m->w, x);
  for (int d = nd - 1; d >= 0; d--) {
    const MFGroup &g = m->groups[d];
    if (g.ns == 0) continue;
    const dim3 grid((unsigned)((g.nf + kFwdRows - 1) / kFwdRows), (unsigned)g.ncc, (unsigned)g.nfr);
    k_mf_fwd_part<<<grid, kFwdRows, 0, s>>>(g.nf, g.ns, g.ncc, g.cw, g.G, g.sdof, m->w, g.part);
    k_mf_fwd_sum<<<grid1((long)g.nfr * g.nf), kBS, 0, s>>>(g.nfr, g.nf, g.ns, g.nb, g.ncc, g.part, g.sdof, m->w,
                                                          g.ub);
    if (g.nscat > 0)
      k_mf_fwd_scatter<<<grid1(g.nscat), kBS, 0, s>>>(g.nscat, g.sc_dof, g.sc_ptr, g.sc_ent, g.ub, m->w);
  }
  for (int d = 0; d < nd; d++) {
    const MFGroup &g = m->groups[d];
    if (g.ns == 0) continue;
    k_mf_bwd_gather<<<grid1((long)g.nfr * g.nf), kBS, 0, s>>>(g.nfr, g.nf, g.ns, g.nb, g.sdof, g.bdof, m->w, x,
                                                             g.v);
    k_mf_bwd<<<dim3((unsigned)((g.ns + 3) / 4), (unsigned)g.nfr), 256, 0, s>>>(g.nf, g.ns, g.G, g.sdof, g.v, x);
  }
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

}  // namespace ptmf
