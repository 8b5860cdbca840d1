// residual.hip -- the FP64 edge residual on gfx950 and its context.
//
// Layout in HBM (per context):
//   X4     [N][4]  coordinates padded to 32 B
//   lsq6   [N][6]  per-node LSQ factors in the form the weights use:
//                  (1/r11, 1/r22, 1/r33, r12/r11, r23/r22, psi)
//   grad   [N][16] primitive gradients, [d*5+k] for d = x,y,z (128 B/node)
//   rowptr [N+1], ent [nnz] (int2): per node, its incident edges in edge-id
//                  order (= ascending neighbour), entry = (neighbour, edge),
//                  then its boundary-face slots in face order,
//                  entry = (-(face+1), code)
//   normal [E][3], fnorm3 [F][3] (face normal / 3), invcv [N]
//
// Both passes are CSR-entry parallel: one lane per (node, incident edge or
// face) entry computes that entry's contribution, the block stages the
// contributions in LDS, and one lane per (node, component) sums its row in
// entry order.  A node's sum therefore runs over its incident edges in
// ascending edge id and then its faces, which is the order of the
// reference's sequential scatter (Edge.C:391-435 then SubDomain.C:1391-1394):
// results do not depend on scheduling and match the reference's summation.
// Blocks cover whole rows (<= kBlock entries); blockIdx is remapped so that
// each XCD sweeps a contiguous node range (its L2 keeps the neighbour window).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "aerofx.h"
#include "device_physics.h"
#include "host_internal.h"
#include "host_mesh.h"

namespace aerofx {

constexpr int kBlock = 256;
constexpr int kNumXcd = 8;

struct ResidualTables {
  const int *chunk_rows;
  const int *rowptr;
  const int2 *ent;
  const double *X4;
  const double *lsq6;
  const double *normal;
  const double *fnorm3;
  const double *invcv;
  int nchunks;
};

// blockIdx -> chunk so that XCD x (which receives blocks b with b % 8 == x)
// processes one contiguous run of chunks.
__device__ __forceinline__ int xcd_chunk(int b, int nchunks) {
  int q = nchunks / kNumXcd, r = nchunks % kNumXcd;
  int x = b % kNumXcd, k = b / kNumXcd;
  return x * q + min(x, r) + k;
}

__device__ __forceinline__ void load5(const double *p, double v[5]) {
#pragma unroll
  for (int k = 0; k < 5; ++k) v[k] = p[k];
}

// Row assignment of the block's entries into LDS.
__device__ __forceinline__ void assign_rows(const int *rowptr, int r0, int r1, int e0, int *srow) {
  for (int r = r0 + (int)threadIdx.x; r < r1; r += kBlock)
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) srow[e - e0] = r;
}

// ---------------------------------------------------------------------------
// Pass 1: least-squares gradients of the primitive variables
// (SubDomain.C:137-159, 297-374).  Per entry (n, m):
//   W = W(R_n, X_m - X_n), contribution = W (V_m - V_n)
// which is bit-identical to the reference's +Wi dV at i and -Wj dV at j.
// FROM_U: input is conservative U (converted per read).  OUT3: write three
// [N][5] arrays instead of the internal [N][16] layout.
// ---------------------------------------------------------------------------
template <bool FROM_U, bool OUT3>
__global__ __launch_bounds__(kBlock) void k_gradients(ResidualTables t, const double *__restrict__ W,
                                                      Gas gas, double *__restrict__ grad,
                                                      double *__restrict__ ddx, double *__restrict__ ddy,
                                                      double *__restrict__ ddz, int *__restrict__ err) {
  __shared__ double sc[15 * kBlock];
  __shared__ int srow[kBlock];
  const int c = xcd_chunk(blockIdx.x, t.nchunks);
  const int r0 = t.chunk_rows[c], r1 = t.chunk_rows[c + 1];
  const int e0 = t.rowptr[r0], e1 = t.rowptr[r1];
  if (err && blockIdx.x == 0 && threadIdx.x == 0) *err = 0;  // flux pass follows in stream order
  assign_rows(t.rowptr, r0, r1, e0, srow);
  __syncthreads();
  const int tid = threadIdx.x;
  if (tid < e1 - e0) {
    const int n = srow[tid];
    const int2 en = t.ent[e0 + tid];
    double cq[15];
#pragma unroll
    for (int q = 0; q < 15; ++q) cq[q] = 0.0;
    if (en.x >= 0) {
      const int m = en.x;
      double Vn[5], Vm[5];
      if (FROM_U) {
        double Un[5], Um[5];
        load5(W + 5 * (size_t)n, Un);
        load5(W + 5 * (size_t)m, Um);
        prim_from_cons(Un, gas, Vn);
        prim_from_cons(Um, gas, Vm);
      } else {
        load5(W + 5 * (size_t)n, Vn);
        load5(W + 5 * (size_t)m, Vm);
      }
      const double4 xn = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)n);
      const double4 xm = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)m);
      const double *L = t.lsq6 + 6 * (size_t)n;
      const double or11 = L[0], or22 = L[1], or33 = L[2], r12or11 = L[3], r23or22 = L[4], psi = L[5];
      const double dx0 = xm.x - xn.x, dx1 = xm.y - xn.y, dx2 = xm.z - xn.z;
      const double alpha1 = dx0 * or11 * or11;
      const double alpha2 = (dx1 - r12or11 * dx0) * or22 * or22;
      const double alpha3 = (dx2 - r23or22 * dx1 + psi * dx0) * or33 * or33;
      const double w0 = alpha1 - r12or11 * alpha2 + psi * alpha3;
      const double w1 = alpha2 - r23or22 * alpha3;
      const double w2 = alpha3;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const double d = Vm[k] - Vn[k];
        cq[k] = w0 * d;
        cq[5 + k] = w1 * d;
        cq[10 + k] = w2 * d;
      }
    }
#pragma unroll
    for (int q = 0; q < 15; ++q) sc[q * kBlock + tid] = cq[q];
  }
  __syncthreads();
  const int nr = r1 - r0;
  for (int idx = tid; idx < nr * 16; idx += kBlock) {
    const int rr = idx >> 4, q = idx & 15;
    const int r = r0 + rr;
    double s = 0.0;
    if (q < 15) {
      const int b = t.rowptr[r] - e0, eend = t.rowptr[r + 1] - e0;
      for (int e = b; e < eend; ++e) s += sc[q * kBlock + e];
    }
    if (!OUT3) {
      grad[16 * (size_t)r + q] = s;
    } else if (q < 15) {
      const int d = q / 5, k = q - 5 * d;
      double *o = d == 0 ? ddx : (d == 1 ? ddy : ddz);
      o[5 * (size_t)r + k] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Pass 2: edge fluxes (Edge.C:372-439, MUSCL + Van Albada + ROEFLUX7) and
// boundary-face fluxes (Face.C:356-389: slip wall, Steger-Warming farfield),
// summed per node in reference order, then scaled by 1/V
// (SpaceOperator.C:599-624).
// ---------------------------------------------------------------------------
struct FluxParams {
  Gas gas;
  double beta;
  double ub_in[5], ub_out[5];
  int second_order;
  int scale;
};

template <bool SECOND>
__global__ __launch_bounds__(kBlock) void k_fluxes(ResidualTables t, const double *__restrict__ U,
                                                   const double *__restrict__ grad, FluxParams fp,
                                                   double *__restrict__ R, int *__restrict__ err) {
  __shared__ double sc[5 * kBlock];
  __shared__ int srow[kBlock];
  const int c = xcd_chunk(blockIdx.x, t.nchunks);
  const int r0 = t.chunk_rows[c], r1 = t.chunk_rows[c + 1];
  const int e0 = t.rowptr[r0], e1 = t.rowptr[r1];
  assign_rows(t.rowptr, r0, r1, e0, srow);
  __syncthreads();
  const int tid = threadIdx.x;
  const Gas &g = fp.gas;
  if (tid < e1 - e0) {
    const int n = srow[tid];
    const int2 en = t.ent[e0 + tid];
    double out[5];
    double Un[5], Vn[5];
    load5(U + 5 * (size_t)n, Un);
    prim_from_cons(Un, g, Vn);
    if (en.x >= 0) {
      const int m = en.x, l = en.y;
      const bool nlow = n < m;
      const int i = nlow ? n : m, j = nlow ? m : n;
      double Um[5], Vm[5];
      load5(U + 5 * (size_t)m, Um);
      prim_from_cons(Um, g, Vm);
      double Vi[5], Vj[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        Vi[k] = nlow ? Vn[k] : Vm[k];
        Vj[k] = nlow ? Vm[k] : Vn[k];
      }
      double Vij[5], Vji[5];
      if (SECOND) {
        const double4 xi = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)i);
        const double4 xj = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)j);
        const double dx0 = xj.x - xi.x, dx1 = xj.y - xi.y, dx2 = xj.z - xi.z;
        const double *gi = grad + 16 * (size_t)i;
        const double *gj = grad + 16 * (size_t)j;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          const double ddVij = dx0 * gi[k] + dx1 * gi[5 + k] + dx2 * gi[10 + k];
          const double ddVji = dx0 * gj[k] + dx1 * gj[5 + k] + dx2 * gj[10 + k];
          vanalbada(fp.beta, Vi[k], ddVij, Vj[k], ddVji, Vij[k], Vji[k]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) { Vij[k] = Vi[k]; Vji[k] = Vj[k]; }
      }
      // checkReconstructedValues (EdgeCore.C:137-166): count once per edge
      if (nlow) {
        int bad = (Vij[0] <= 0.0 || Vij[4] + g.pstiff <= 0.0) + (Vji[0] <= 0.0 || Vji[4] + g.pstiff <= 0.0);
        if (bad) atomicAdd(err, bad);
      }
      const double nv[3] = {t.normal[3 * (size_t)l], t.normal[3 * (size_t)l + 1], t.normal[3 * (size_t)l + 2]};
      double f[5];
      if (sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]) < 1e-18) {
#pragma unroll
        for (int k = 0; k < 5; ++k) f[k] = 0.0;
      } else {
        roeflux7(g, nv, 0.0, Vij, Vji, f);
      }
#pragma unroll
      for (int k = 0; k < 5; ++k) out[k] = nlow ? f[k] : -f[k];
    } else {
      const int face = -en.x - 1, code = en.y;
      const double nv[3] = {t.fnorm3[3 * (size_t)face], t.fnorm3[3 * (size_t)face + 1],
                            t.fnorm3[3 * (size_t)face + 2]};
      if (code == AEROFX_BC_SLIP_WALL || code == -AEROFX_BC_SLIP_WALL) {
        wallflux(nv, 0.0, Vn, out);
      } else {
        const double *ub = (code == AEROFX_BC_INLET || code == -AEROFX_BC_INLET) ? fp.ub_in : fp.ub_out;
        boundflux5(g, nv, 0.0, Vn, ub, out);
      }
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) sc[k * kBlock + tid] = out[k];
  }
  __syncthreads();
  const int nr = r1 - r0;
  for (int idx = tid; idx < nr * 5; idx += kBlock) {
    const int rr = idx / 5, q = idx - 5 * rr;
    const int r = r0 + rr;
    const int b = t.rowptr[r] - e0, eend = t.rowptr[r + 1] - e0;
    double s = 0.0;
    for (int e = b; e < eend; ++e) s += sc[q * kBlock + e];
    if (fp.scale) s *= t.invcv[r];
    R[5 * (size_t)r + q] = s;
  }
}

__global__ void k_zero_int(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_primitive(int N, const double *__restrict__ U, Gas g, double *__restrict__ V) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double u[5], v[5];
  load5(U + 5 * (size_t)n, u);
  prim_from_cons(u, g, v);
#pragma unroll
  for (int k = 0; k < 5; ++k) V[5 * (size_t)n + k] = v[k];
}

}  // namespace aerofx

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
using namespace aerofx;

struct aerofx_ctx {
  int N = 0, T = 0, F = 0, E = 0;
  aerofx_params p{};
  Gas gas{};
  double ub_in[5], ub_out[5];
  hipStream_t own_stream = nullptr, stream = nullptr;
  // host copies of set-up tables
  std::vector<int> ptr;
  std::vector<double> cv, normals, fnorm, lsq;
  // device tables
  ResidualTables tab{};
  std::vector<void *> allocs;
  double *d_grad = nullptr;
  int *d_err = nullptr;
  int *h_err = nullptr;
  // per-kernel timers (the reference's Timer buckets, Timer.h:12-15)
  bool prof = false;
  struct Pending { int bucket; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  double timer_ms[AEROFX_NUM_TIMERS] = {0};
  long long timer_n[AEROFX_NUM_TIMERS] = {0};
};

namespace {

int hip_fail(hipError_t e, const char *what) {
  return set_error(AEROFX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
int dev_upload(aerofx_ctx *c, const T *src, size_t n, const T **dst) {
  void *p = nullptr;
  size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  c->allocs.push_back(p);
  if (n) {
    e = hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
  }
  *dst = static_cast<const T *>(p);
  return 0;
}

hipEvent_t take_event(aerofx_ctx *c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one kernel launch with events when profiling is on.
struct ProfScope {
  aerofx_ctx *c;
  int bucket;
  hipEvent_t a = nullptr;
  ProfScope(aerofx_ctx *ctx, int b) : c(ctx), bucket(b) {
    if (c->prof && (a = take_event(c))) (void)hipEventRecord(a, c->stream);
  }
  ~ProfScope() {
    if (!a) return;
    hipEvent_t b = take_event(c);
    if (!b) return;
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({bucket, a, b});
  }
};

void collect_timers(aerofx_ctx *c) {
  for (auto &p : c->pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      c->timer_ms[p.bucket] += ms;
      c->timer_n[p.bucket] += 1;
    }
    c->pool.push_back(p.a);
    c->pool.push_back(p.b);
  }
  c->pending.clear();
}

void destroy(aerofx_ctx *c) {
  collect_timers(c);
  for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->h_err) (void)hipHostFree(c->h_err);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

}  // namespace

extern "C" {

int aerofx_ctx_create(const aerofx_mesh *m, const aerofx_params *p, aerofx_ctx **out) {
  if (!m || !p || !out) return set_error(AEROFX_ERR_ARG, "ctx_create: null argument");
  *out = nullptr;
  const int N = m->num_nodes, T = m->num_tets, F = m->num_faces;
  if (N <= 0 || T <= 0 || F < 0 || !m->X || !m->tets || (F > 0 && (!m->faces || !m->face_codes)))
    return set_error(AEROFX_ERR_ARG, "ctx_create: bad mesh");
  if (!(p->gamma > 1.0) || !(p->rho_inf > 0.0) || !(p->p_inf > 0.0))
    return set_error(AEROFX_ERR_ARG, "ctx_create: bad parameters");
  for (long long k = 0; k < 4LL * T; ++k)
    if (m->tets[k] < 0 || m->tets[k] >= N) return set_error(AEROFX_ERR_MESH, "ctx_create: tet node out of range");
  for (long long k = 0; k < 3LL * F; ++k)
    if (m->faces[k] < 0 || m->faces[k] >= N) return set_error(AEROFX_ERR_MESH, "ctx_create: face node out of range");
  for (int f = 0; f < F; ++f) {
    int cd = m->face_codes[f] < 0 ? -m->face_codes[f] : m->face_codes[f];
    if (cd != AEROFX_BC_SLIP_WALL && cd != AEROFX_BC_INLET && cd != AEROFX_BC_OUTLET)
      return set_error(AEROFX_ERR_MESH, "ctx_create: unsupported face code (2, 4, 5 only)");
  }
  aerofx_ctx *c = new aerofx_ctx();
  c->N = N; c->T = T; c->F = F; c->p = *p;
  c->gas.gam = p->gamma;
  c->gas.gam1 = p->gamma - 1.0;
  c->gas.invgam1 = 1.0 / c->gas.gam1;
  c->gas.pstiff = p->pstiff;
  double Uf[5];
  freestream_state(*p, Uf);
  face_state(Uf, c->ub_in);
  face_state(Uf, c->ub_out);
  try {
    std::vector<int> tet_edges;
    build_edges(N, T, m->tets, c->ptr, &tet_edges);
    c->E = (int)(c->ptr.size() / 2);
    const int E = c->E;
    c->cv.resize(N);
    c->normals.resize(3 * (size_t)E);
    c->fnorm.resize(3 * (size_t)F);
    c->lsq.resize(6 * (size_t)N);
    control_volumes(N, T, m->X, m->tets, c->cv.data());
    edge_normals(T, m->X, m->tets, tet_edges.data(), E, c->normals.data());
    face_normals(F, m->X, m->faces, c->fnorm.data());
    lsq_factors(N, E, m->X, c->ptr.data(), c->lsq.data());
    for (int i = 0; i < N; ++i)
      if (!(c->cv[i] > 0.0)) { destroy(c); return set_error(AEROFX_ERR_MESH, "ctx_create: non-positive control volume"); }
    // CSR of incident edges, then face slots in face order
    std::vector<int> rowptr_e, col, eid;
    node_csr(N, E, c->ptr.data(), rowptr_e, col, eid);
    std::vector<int> fcount(N + 1, 0);
    for (long long k = 0; k < 3LL * F; ++k) fcount[m->faces[k] + 1]++;
    std::vector<int> rowptr(N + 1, 0);
    for (int n = 0; n < N; ++n) rowptr[n + 1] = rowptr[n] + (rowptr_e[n + 1] - rowptr_e[n]) + fcount[n + 1];
    std::vector<int2> ent(rowptr[N]);
    std::vector<int> fill(N);
    for (int n = 0; n < N; ++n) {
      int o = rowptr[n];
      for (int k = rowptr_e[n]; k < rowptr_e[n + 1]; ++k) ent[o++] = make_int2(col[k], eid[k]);
      fill[n] = o;
    }
    for (int f = 0; f < F; ++f)
      for (int a = 0; a < 3; ++a) ent[fill[m->faces[3 * f + a]]++] = make_int2(-(f + 1), m->face_codes[f]);
    // whole-row chunks of at most kBlock entries
    std::vector<int> chunk_rows(1, 0);
    int acc = 0;
    for (int n = 0; n < N; ++n) {
      int cnt = rowptr[n + 1] - rowptr[n];
      if (cnt > kBlock) { destroy(c); return set_error(AEROFX_ERR_MESH, "ctx_create: node degree exceeds 256"); }
      if (acc + cnt > kBlock) { chunk_rows.push_back(n); acc = 0; }
      acc += cnt;
    }
    chunk_rows.push_back(N);
    // derived device tables
    std::vector<double> X4(4 * (size_t)N), lsq6(6 * (size_t)N), fnorm3(3 * (size_t)F), invcv(N);
    for (int n = 0; n < N; ++n) {
      for (int a = 0; a < 3; ++a) X4[4 * (size_t)n + a] = m->X[3 * (size_t)n + a];
      X4[4 * (size_t)n + 3] = 0.0;
      const double *R = c->lsq.data() + 6 * (size_t)n;
      // computeLocalWeightsLeastSquares (SubDomain.C:137-159) node factors
      double or11 = 1.0 / R[0], or22 = 1.0 / R[3], or33 = 1.0 / R[5];
      double *L = lsq6.data() + 6 * (size_t)n;
      L[0] = or11; L[1] = or22; L[2] = or33;
      L[3] = R[1] * or11;
      L[4] = R[4] * or22;
      L[5] = (R[1] * R[4] - R[2] * R[3]) * or11 * or22;
      invcv[n] = 1.0 / c->cv[n];
    }
    const double third = 1.0 / 3.0;
    for (size_t k = 0; k < 3 * (size_t)F; ++k) fnorm3[k] = third * c->fnorm[k];
    int rc = 0;
    const int *d_chunk = nullptr, *d_rowptr = nullptr;
    const int2 *d_ent = nullptr;
    const double *d_X4 = nullptr, *d_lsq6 = nullptr, *d_normal = nullptr, *d_fn3 = nullptr, *d_invcv = nullptr;
    if ((rc = dev_upload(c, chunk_rows.data(), chunk_rows.size(), &d_chunk)) ||
        (rc = dev_upload(c, rowptr.data(), rowptr.size(), &d_rowptr)) ||
        (rc = dev_upload(c, ent.data(), ent.size(), &d_ent)) ||
        (rc = dev_upload(c, X4.data(), X4.size(), &d_X4)) ||
        (rc = dev_upload(c, lsq6.data(), lsq6.size(), &d_lsq6)) ||
        (rc = dev_upload(c, c->normals.data(), c->normals.size(), &d_normal)) ||
        (rc = dev_upload(c, fnorm3.data(), fnorm3.size(), &d_fn3)) ||
        (rc = dev_upload(c, invcv.data(), invcv.size(), &d_invcv))) {
      destroy(c);
      return rc;
    }
    c->tab.chunk_rows = d_chunk;
    c->tab.rowptr = d_rowptr;
    c->tab.ent = d_ent;
    c->tab.X4 = d_X4;
    c->tab.lsq6 = d_lsq6;
    c->tab.normal = d_normal;
    c->tab.fnorm3 = d_fn3;
    c->tab.invcv = d_invcv;
    c->tab.nchunks = (int)chunk_rows.size() - 1;
    void *pg = nullptr;
    hipError_t e = hipMalloc(&pg, 16 * sizeof(double) * (size_t)N);
    if (e != hipSuccess) { destroy(c); return hip_fail(e, "hipMalloc grad"); }
    c->allocs.push_back(pg);
    c->d_grad = static_cast<double *>(pg);
    e = hipMalloc(&pg, sizeof(int) * 4);
    if (e != hipSuccess) { destroy(c); return hip_fail(e, "hipMalloc err"); }
    c->allocs.push_back(pg);
    c->d_err = static_cast<int *>(pg);
    e = hipMemset(c->d_err, 0, sizeof(int) * 4);
    if (e != hipSuccess) { destroy(c); return hip_fail(e, "hipMemset"); }
    e = hipHostMalloc((void **)&c->h_err, sizeof(int) * 4, hipHostMallocDefault);
    if (e != hipSuccess) { destroy(c); return hip_fail(e, "hipHostMalloc"); }
    e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) { destroy(c); return hip_fail(e, "hipStreamCreate"); }
    c->stream = c->own_stream;
  } catch (std::exception &ex) {
    destroy(c);
    return set_error(AEROFX_ERR_MESH, ex.what());
  }
  *out = c;
  return AEROFX_OK;
}

int aerofx_ctx_destroy(aerofx_ctx *c) {
  if (!c) return AEROFX_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  destroy(c);
  return AEROFX_OK;
}

int aerofx_ctx_sizes(const aerofx_ctx *c, int *N, int *E, int *F) {
  if (!c) return set_error(AEROFX_ERR_ARG, "ctx_sizes: null ctx");
  if (N) *N = c->N;
  if (E) *E = c->E;
  if (F) *F = c->F;
  return AEROFX_OK;
}

int aerofx_ctx_set_stream(aerofx_ctx *c, void *stream) {
  if (!c) return set_error(AEROFX_ERR_ARG, "set_stream: null ctx");
  c->stream = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
  return AEROFX_OK;
}

int aerofx_ctx_sync(aerofx_ctx *c) {
  if (!c) return set_error(AEROFX_ERR_ARG, "sync: null ctx");
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? AEROFX_OK : hip_fail(e, "hipStreamSynchronize");
}

int aerofx_ctx_profile(aerofx_ctx *c, int enable) {
  if (!c) return set_error(AEROFX_ERR_ARG, "ctx_profile: null ctx");
  collect_timers(c);
  for (int k = 0; k < AEROFX_NUM_TIMERS; ++k) { c->timer_ms[k] = 0.0; c->timer_n[k] = 0; }
  c->prof = enable != 0;
  return AEROFX_OK;
}

int aerofx_ctx_timers(aerofx_ctx *c, double *ms, long long *counts) {
  if (!c) return set_error(AEROFX_ERR_ARG, "ctx_timers: null ctx");
  collect_timers(c);
  for (int k = 0; k < AEROFX_NUM_TIMERS; ++k) {
    if (ms) ms[k] = c->timer_ms[k];
    if (counts) counts[k] = c->timer_n[k];
  }
  return AEROFX_OK;
}

int aerofx_ctx_tables(const aerofx_ctx *c, int *ptr, double *cv, double *en, double *fn, double *lsq) {
  if (!c) return set_error(AEROFX_ERR_ARG, "ctx_tables: null ctx");
  if (ptr) std::memcpy(ptr, c->ptr.data(), c->ptr.size() * sizeof(int));
  if (cv) std::memcpy(cv, c->cv.data(), c->cv.size() * sizeof(double));
  if (en) std::memcpy(en, c->normals.data(), c->normals.size() * sizeof(double));
  if (fn) std::memcpy(fn, c->fnorm.data(), c->fnorm.size() * sizeof(double));
  if (lsq) std::memcpy(lsq, c->lsq.data(), c->lsq.size() * sizeof(double));
  return AEROFX_OK;
}

int aerofx_residual_ex(aerofx_ctx *c, const double *U, double *R, int flags) {
  if (!c || !U || !R) return set_error(AEROFX_ERR_ARG, "residual: null argument");
  hipError_t e;
  const int nb = c->tab.nchunks;
  if (c->p.second_order) {
    ProfScope ps(c, AEROFX_TIMER_NODAL_GRAD);
    hipLaunchKernelGGL((k_gradients<true, false>), dim3(nb), dim3(kBlock), 0, c->stream, c->tab, U, c->gas,
                       c->d_grad, nullptr, nullptr, nullptr, c->d_err);
  } else {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->d_err, 1);
  }
  FluxParams fp;
  fp.gas = c->gas;
  fp.beta = c->p.beta;
  for (int k = 0; k < 5; ++k) { fp.ub_in[k] = c->ub_in[k]; fp.ub_out[k] = c->ub_out[k]; }
  fp.second_order = c->p.second_order;
  fp.scale = (flags & 1) ? 0 : 1;
  {
    ProfScope ps(c, AEROFX_TIMER_FV_TERM);
    if (c->p.second_order)
      hipLaunchKernelGGL((k_fluxes<true>), dim3(nb), dim3(kBlock), 0, c->stream, c->tab, U, c->d_grad, fp, R,
                         c->d_err);
    else
      hipLaunchKernelGGL((k_fluxes<false>), dim3(nb), dim3(kBlock), 0, c->stream, c->tab, U, c->d_grad, fp, R,
                         c->d_err);
  }
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "residual launch");
  if (flags & 2) return AEROFX_OK;
  int cnt = 0;
  int rc = aerofx_error_count(c, &cnt);
  if (rc) return rc;
  return cnt;
}

int aerofx_residual(aerofx_ctx *c, const double *U, double *R) { return aerofx_residual_ex(c, U, R, 0); }

int aerofx_error_count(aerofx_ctx *c, int *count) {
  if (!c || !count) return set_error(AEROFX_ERR_ARG, "error_count: null argument");
  hipError_t e = hipMemcpyAsync(c->h_err, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
  e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  *count = c->h_err[0];
  return AEROFX_OK;
}

int aerofx_gradients(aerofx_ctx *c, const double *V, double *ddx, double *ddy, double *ddz) {
  if (!c || !V || !ddx || !ddy || !ddz) return set_error(AEROFX_ERR_ARG, "gradients: null argument");
  hipLaunchKernelGGL((k_gradients<false, true>), dim3(c->tab.nchunks), dim3(kBlock), 0, c->stream, c->tab, V,
                     c->gas, nullptr, ddx, ddy, ddz, nullptr);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AEROFX_OK : hip_fail(e, "gradients launch");
}

int aerofx_primitive(aerofx_ctx *c, const double *U, double *V) {
  if (!c || !U || !V) return set_error(AEROFX_ERR_ARG, "primitive: null argument");
  hipLaunchKernelGGL(k_primitive, dim3((c->N + 255) / 256), dim3(256), 0, c->stream, c->N, U, c->gas, V);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? AEROFX_OK : hip_fail(e, "primitive launch");
}

}  // extern "C"
