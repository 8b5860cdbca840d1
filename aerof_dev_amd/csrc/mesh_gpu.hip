// mesh_gpu.hip -- mesh orderings on the GPU: edge colouring and RCM.
//
// Edge colouring (no reference counterpart; the rule is the project's):
// key(l) = (lowbias32(l) << 32) | l; each edge takes the smallest colour not
// used by an incident edge of higher key.  That is the sequential greedy
// colouring in decreasing key order, and the Jones-Plassmann rounds below
// produce it exactly: an edge is coloured in the first round in which every
// incident edge of higher key is coloured, and two incident edges are never
// coloured in the same round, so the result does not depend on scheduling.
//
// RCM (Connectivity::renumRCM / findPseudoDiam / rootLS,
// ConnectivityCore.C:324-635): the pseudo-peripheral search and the level
// bookkeeping run on the host, each breadth-first level structure on the GPU.
// A level is built in the order the sequential search appends it: the nodes
// of the next level are claimed by the earliest frontier position that
// reaches them, and each frontier node appends its claimed neighbours in
// ascending order (the graph rows are sorted and hold the node itself; the
// reference's degree sort does not change the order, ConnectivityCore.C:620-628).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "context.h"
#include "host_internal.h"
#include "host_mesh.h"

namespace aerofx {
namespace {

constexpr int kB = 256;

__device__ __forceinline__ unsigned long long edge_key(unsigned l) {
  unsigned x = l;
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return ((unsigned long long)x << 32) | l;
}

// one Jones-Plassmann round; remaining counts the edges still uncoloured
__global__ __launch_bounds__(kB) void k_color_round(int E, const int2 *__restrict__ ends,
                                                    const int *__restrict__ rowptr, const int *__restrict__ inc,
                                                    int *__restrict__ color, int *__restrict__ remaining,
                                                    int *__restrict__ overflow) {
  const int l = blockIdx.x * kB + threadIdx.x;
  if (l >= E || color[l] >= 0) return;
  const unsigned long long kl = edge_key((unsigned)l);
  const int2 ij = ends[l];
  unsigned long long used = 0;
  bool ready = true;
  const int nd[2] = {ij.x, ij.y};
  for (int s = 0; s < 2 && ready; ++s)
    for (int q = rowptr[nd[s]]; q < rowptr[nd[s] + 1]; ++q) {
      const int e = inc[q];
      if (e == l) continue;
      const int ce = color[e];
      if (ce < 0) {
        if (edge_key((unsigned)e) > kl) {
          ready = false;
          break;
        }
      } else if (ce < 64) {
        used |= 1ull << ce;
      } else {
        *overflow = 1;
      }
    }
  if (!ready) {
    atomicAdd(remaining, 1);
    return;
  }
  if (used == ~0ull) {
    *overflow = 1;
    color[l] = 64;
    return;
  }
  color[l] = __builtin_ctzll(~used);
}

// --- breadth-first levels --------------------------------------------------
// claim[v] = min frontier position reaching v (v unvisited and in the mask)
__global__ __launch_bounds__(kB) void k_bfs_claim(int lb, int le, const int *__restrict__ ls,
                                                  const int *__restrict__ ptr, const int *__restrict__ tgt,
                                                  const int *__restrict__ loc, int *__restrict__ claim) {
  const int q = lb + blockIdx.x * kB + threadIdx.x;
  if (q >= le) return;
  const int u = ls[q];
  for (int k = ptr[u]; k < ptr[u + 1]; ++k) {
    const int v = tgt[k];
    if (loc[v]) atomicMin(&claim[v], q);
  }
}

// number of nodes each frontier position appends
__global__ __launch_bounds__(kB) void k_bfs_count(int lb, int le, const int *__restrict__ ls,
                                                  const int *__restrict__ ptr, const int *__restrict__ tgt,
                                                  const int *__restrict__ loc, const int *__restrict__ claim,
                                                  int *__restrict__ cnt) {
  const int q = lb + blockIdx.x * kB + threadIdx.x;
  if (q >= le) return;
  const int u = ls[q];
  int c = 0;
  for (int k = ptr[u]; k < ptr[u + 1]; ++k) {
    const int v = tgt[k];
    if (loc[v] && claim[v] == q) ++c;
  }
  cnt[q - lb] = c;
}

// exclusive scan of cnt[0..n) by one block (sequential over 256-wide tiles)
__global__ __launch_bounds__(kB) void k_scan(int n, const int *__restrict__ cnt, int *__restrict__ off,
                                             int *__restrict__ total) {
  __shared__ int sh[kB];
  int carry = 0;
  for (int base = 0; base < n; base += kB) {
    const int i = base + threadIdx.x;
    const int v = i < n ? cnt[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kB; o <<= 1) {
      const int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < n) off[i] = carry + sh[threadIdx.x] - v;
    const int tile_sum = sh[kB - 1];
    __syncthreads();
    carry += tile_sum;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kB) void k_bfs_append(int lb, int le, int nf, int *__restrict__ ls,
                                                   const int *__restrict__ ptr, const int *__restrict__ tgt,
                                                   int *__restrict__ loc, const int *__restrict__ claim,
                                                   const int *__restrict__ off) {
  const int q = lb + blockIdx.x * kB + threadIdx.x;
  if (q >= le) return;
  const int u = ls[q];
  int pos = nf + off[q - lb];
  for (int k = ptr[u]; k < ptr[u + 1]; ++k) {
    const int v = tgt[k];
    if (loc[v] && claim[v] == q) ls[pos++] = v;
  }
}

// loc[v] = 0 and claim reset for the nodes appended in [a, b)
__global__ __launch_bounds__(kB) void k_bfs_visit(int a, int b, const int *__restrict__ ls, int *__restrict__ loc,
                                                  int *__restrict__ claim) {
  const int q = a + blockIdx.x * kB + threadIdx.x;
  if (q >= b) return;
  const int v = ls[q];
  loc[v] = 0;
  claim[v] = INT_MAX;
}

__global__ __launch_bounds__(kB) void k_set_loc(int N, const int *__restrict__ mask, int *__restrict__ loc,
                                                int *__restrict__ claim) {
  const int v = blockIdx.x * kB + threadIdx.x;
  if (v >= N) return;
  loc[v] = mask ? (mask[v] > 0 ? 1 : 0) : 1;
  claim[v] = INT_MAX;
}

__global__ void k_set_root(int root, int *ls, int *loc) {
  if (threadIdx.x == 0) {
    ls[0] = root;
    loc[root] = 0;
  }
}

inline dim3 grid(long long n) { return dim3((unsigned)std::max<long long>(1, (n + kB - 1) / kB)); }

struct DevGraph {
  int N = 0;
  int *ptr = nullptr, *tgt = nullptr, *ls = nullptr, *loc = nullptr, *claim = nullptr, *cnt = nullptr,
      *off = nullptr, *total = nullptr, *mask = nullptr;
  hipStream_t s = nullptr;
  int *h_total = nullptr;
  ~DevGraph() {
    int *bufs[] = {ptr, tgt, ls, loc, claim, cnt, off, total, mask};
    for (int *b : bufs)
      if (b) (void)hipFree(b);
    if (h_total) (void)hipHostFree(h_total);
    if (s) (void)hipStreamDestroy(s);
  }
};

// rooted level structure on the GPU: returns the number of levels; xls, ls on the host
int dev_rootLS(DevGraph &g, int root, bool use_mask, std::vector<int> &xls, std::vector<int> &ls, int &width,
               hipError_t &err) {
  k_set_loc<<<grid(g.N), kB, 0, g.s>>>(g.N, use_mask ? g.mask : nullptr, g.loc, g.claim);
  k_set_root<<<1, 64, 0, g.s>>>(root, g.ls, g.loc);
  xls.assign(1, 0);
  int nlvl = 1, nf = 1, lb = 0, le = 1;
  width = 0;
  while (true) {
    xls.push_back(nf);
    // next level from [lb, le)
    const int nq = le - lb;
    k_bfs_claim<<<grid(nq), kB, 0, g.s>>>(lb, le, g.ls, g.ptr, g.tgt, g.loc, g.claim);
    k_bfs_count<<<grid(nq), kB, 0, g.s>>>(lb, le, g.ls, g.ptr, g.tgt, g.loc, g.claim, g.cnt);
    k_scan<<<1, kB, 0, g.s>>>(nq, g.cnt, g.off, g.total);
    k_bfs_append<<<grid(nq), kB, 0, g.s>>>(lb, le, nf, g.ls, g.ptr, g.tgt, g.loc, g.claim, g.off);
    err = hipMemcpyAsync(g.h_total, g.total, sizeof(int), hipMemcpyDeviceToHost, g.s);
    if (err == hipSuccess) err = hipStreamSynchronize(g.s);
    if (err != hipSuccess) return -1;
    const int added = *g.h_total;
    if (added > 0) k_bfs_visit<<<grid(added), kB, 0, g.s>>>(nf, nf + added, g.ls, g.loc, g.claim);
    lb = le;
    le = nf + added;
    nf = le;
    if (added == 0) break;
    width = std::max(width, added);
    nlvl++;
  }
  // xls has nlvl + 1 entries: level k = [xls[k], xls[k+1])
  xls.resize(nlvl + 1);
  ls.resize(nf);
  err = hipMemcpy(ls.data(), g.ls, sizeof(int) * nf, hipMemcpyDeviceToHost);
  if (err != hipSuccess) return -1;
  // the reference's width counts only levels after the root's
  return nlvl;
}

}  // namespace
}  // namespace aerofx

using namespace aerofx;

extern "C" int aerofx_mesh_color_edges_gpu(int num_nodes, int num_edges, const int *ptr, int *color) {
  if (num_nodes < 0 || num_edges < 0 || (num_edges > 0 && (!ptr || !color)))
    return set_error(AEROFX_ERR_ARG, "color_edges_gpu: bad argument");
  if (num_edges == 0) return 0;
  const int N = num_nodes, E = num_edges;
  // incident-edge lists per node (edge ids ascending)
  std::vector<int> rowptr(N + 1, 0), inc(2 * (size_t)E);
  for (int l = 0; l < E; ++l) {
    const int a = ptr[2 * (size_t)l], b = ptr[2 * (size_t)l + 1];
    if (a < 0 || b < 0 || a >= N || b >= N) return set_error(AEROFX_ERR_MESH, "color_edges_gpu: bad edge");
    rowptr[a + 1]++;
    rowptr[b + 1]++;
  }
  for (int n = 0; n < N; ++n) rowptr[n + 1] += rowptr[n];
  {
    std::vector<int> fill(rowptr.begin(), rowptr.end() - 1);
    for (int l = 0; l < E; ++l) {
      inc[fill[ptr[2 * (size_t)l]]++] = l;
      inc[fill[ptr[2 * (size_t)l + 1]]++] = l;
    }
  }
  int2 *d_ends = nullptr;
  int *d_row = nullptr, *d_inc = nullptr, *d_col = nullptr, *d_cnt = nullptr, *h_cnt = nullptr;
  hipStream_t s = nullptr;
  int rc = 0;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&d_ends, sizeof(int2) * E);
  if (e == hipSuccess) e = hipMalloc(&d_row, sizeof(int) * (N + 1));
  if (e == hipSuccess) e = hipMalloc(&d_inc, sizeof(int) * 2 * (size_t)E);
  if (e == hipSuccess) e = hipMalloc(&d_col, sizeof(int) * E);
  if (e == hipSuccess) e = hipMalloc(&d_cnt, sizeof(int) * 2);
  if (e == hipSuccess) e = hipHostMalloc((void **)&h_cnt, sizeof(int) * 2, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemcpy(d_ends, ptr, sizeof(int2) * E, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_row, rowptr.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_inc, inc.data(), sizeof(int) * 2 * (size_t)E, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(d_col, 0xff, sizeof(int) * E);  // -1: uncoloured
  int ncol = 0;
  // every round colours at least the edge of largest remaining key, so at
  // most E rounds; the loop also stops on overflow
  for (int round = 0; e == hipSuccess && round <= E; ++round) {
    e = hipMemsetAsync(d_cnt, 0, sizeof(int) * 2, s);
    if (e != hipSuccess) break;
    k_color_round<<<grid(E), kB, 0, s>>>(E, d_ends, d_row, d_inc, d_col, d_cnt, d_cnt + 1);
    e = hipMemcpyAsync(h_cnt, d_cnt, sizeof(int) * 2, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) break;
    if (h_cnt[1]) {
      rc = set_error(AEROFX_ERR_MESH, "color_edges_gpu: more than 64 colours");
      break;
    }
    if (h_cnt[0] == 0) break;
  }
  if (e == hipSuccess && rc == 0) e = hipMemcpy(color, d_col, sizeof(int) * E, hipMemcpyDeviceToHost);
  if (e != hipSuccess && rc == 0) rc = set_hip_error(e, "color_edges_gpu");
  if (rc == 0)
    for (int l = 0; l < E; ++l) ncol = std::max(ncol, color[l] + 1);
  (void)hipFree(d_ends);
  (void)hipFree(d_row);
  (void)hipFree(d_inc);
  (void)hipFree(d_col);
  (void)hipFree(d_cnt);
  if (h_cnt) (void)hipHostFree(h_cnt);
  if (s) (void)hipStreamDestroy(s);
  return rc ? rc : ncol;
}

// RCM with the breadth-first searches on the GPU; identical result to aerofx_mesh_rcm.
extern "C" int aerofx_mesh_rcm_gpu(int num_nodes, int num_tets, const int *tets, int *renum) {
  if (num_nodes <= 0 || num_tets < 0 || !renum || (num_tets > 0 && !tets))
    return set_error(AEROFX_ERR_ARG, "rcm_gpu: bad argument");
  const int N = num_nodes;
  std::vector<int> eptr;
  try {
    build_edges(N, num_tets, tets, eptr, nullptr);
  } catch (const std::exception &ex) {
    return set_error(AEROFX_ERR_MESH, std::string("rcm_gpu: ") + ex.what());
  }
  const long long E = (long long)eptr.size() / 2;
  // graph rows: the node itself plus its neighbours, sorted (SubDomainCore.C:253-320)
  std::vector<int> pointer(N + 1, 0), target;
  for (int i = 0; i < N; ++i) pointer[i + 1] = 1;
  for (long long l = 0; l < E; ++l) {
    pointer[eptr[2 * l] + 1]++;
    pointer[eptr[2 * l + 1] + 1]++;
  }
  for (int i = 0; i < N; ++i) pointer[i + 1] += pointer[i];
  target.assign(pointer[N], 0);
  {
    std::vector<int> fill(pointer.begin(), pointer.end() - 1);
    for (int i = 0; i < N; ++i) target[fill[i]++] = i;
    for (long long l = 0; l < E; ++l) {
      const int a = eptr[2 * l], b = eptr[2 * l + 1];
      target[fill[a]++] = b;
      target[fill[b]++] = a;
    }
    for (int i = 0; i < N; ++i) std::sort(target.begin() + pointer[i], target.begin() + pointer[i + 1]);
  }
  const int numtarget = pointer[N];
  DevGraph g;
  g.N = N;
  hipError_t e = hipStreamCreateWithFlags(&g.s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&g.ptr, sizeof(int) * (N + 1));
  if (e == hipSuccess) e = hipMalloc(&g.tgt, sizeof(int) * std::max(numtarget, 1));
  if (e == hipSuccess) e = hipMalloc(&g.ls, sizeof(int) * N);
  if (e == hipSuccess) e = hipMalloc(&g.loc, sizeof(int) * N);
  if (e == hipSuccess) e = hipMalloc(&g.claim, sizeof(int) * N);
  if (e == hipSuccess) e = hipMalloc(&g.cnt, sizeof(int) * N);
  if (e == hipSuccess) e = hipMalloc(&g.off, sizeof(int) * N);
  if (e == hipSuccess) e = hipMalloc(&g.total, sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&g.mask, sizeof(int) * N);
  if (e == hipSuccess) e = hipHostMalloc((void **)&g.h_total, sizeof(int), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemcpy(g.ptr, pointer.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g.tgt, target.data(), sizeof(int) * numtarget, hipMemcpyHostToDevice);
  if (e != hipSuccess) return set_hip_error(e, "rcm_gpu set-up");
  auto num = [&](int nd, const int *mask) {
    int r = 0;
    for (int j = pointer[nd]; j < pointer[nd + 1]; ++j)
      if (mask[target[j]]) r++;
    return r;
  };
  std::vector<int> mark(N), locMask(N, 0), lrenum(N), order(N), xls, ls, pxls, pls;
  for (int i = 0; i < N; ++i) {
    mark[i] = (pointer[i + 1] != pointer[i]) ? 1 : 0;
    renum[i] = -1;
  }
  auto upload_mask = [&]() {
    return hipMemcpy(g.mask, locMask.data(), sizeof(int) * N, hipMemcpyHostToDevice);
  };
  int nextNum = 0;
  for (int inode = 0; inode < N; ++inode) {
    if (mark[inode] != 1) continue;
    int w;
    const int h = dev_rootLS(g, inode, false, xls, ls, w, e);
    if (h < 0) return set_hip_error(e, "rcm_gpu bfs");
    for (int j = 0; j < xls[h]; ++j) {
      locMask[ls[j]] = 1;
      mark[ls[j]] = 0;
    }
    const std::vector<int> comp(ls.begin(), ls.begin() + xls[h]);
    if ((e = upload_mask()) != hipSuccess) return set_hip_error(e, "rcm_gpu mask");
    // findPseudoDiam (ConnectivityCore.C:500-560)
    int s = 0, en = 0, cmin = numtarget + 1, cmax = 0, nw;
    for (int i = 0; i < N; ++i)
      if (locMask[i]) {
        if ((nw = num(i, locMask.data())) < cmin) {
          cmin = nw;
          s = i;
        }
        if (nw > cmax) cmax = nw;
      }
    std::vector<int> sorted(cmax + 1);
    int ww;
    int hh = dev_rootLS(g, s, true, pxls, pls, ww, e);
    if (hh < 0) return set_hip_error(e, "rcm_gpu bfs");
    const int subconsize = pxls[hh];
    en = pls[pxls[hh - 1]];
    while (hh < subconsize) {
      int k;
      for (k = 0; k <= cmax; ++k) sorted[k] = -1;
      for (k = pxls[hh - 1]; k < pxls[hh]; ++k) sorted[num(pls[k], locMask.data())] = pls[k];
      int w_e = subconsize;
      for (k = 0; k <= cmax; ++k)
        if (sorted[k] >= 0) {
          const int nh = dev_rootLS(g, sorted[k], true, pxls, pls, ww, e);
          if (nh < 0) return set_hip_error(e, "rcm_gpu bfs");
          if (ww < w_e) {
            if (nh > hh) {
              s = sorted[k];
              hh = nh;
              break;
            }
            en = sorted[k];
            w_e = ww;
          }
        }
      if (k > cmax) break;
    }
    // renumRCM: breadth-first from en in ascending-neighbour order, reversed
    int h2 = dev_rootLS(g, en, true, xls, ls, w, e);
    if (h2 < 0) return set_hip_error(e, "rcm_gpu bfs");
    const int last = xls[h2];
    for (int i = last; i--;) lrenum[ls[i]] = nextNum++;
    for (int v : comp) locMask[v] = 0;
    for (int v : comp) renum[v] = lrenum[v];
  }
  return AEROFX_OK;
}
