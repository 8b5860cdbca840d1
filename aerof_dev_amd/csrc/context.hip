// context.hip -- context set-up and the context part of the C ABI.
//
// aerofx_ctx_create builds, on the host, the reference's tables for the mesh
// (edge table, dual normals, control volumes, LSQ factors: host_mesh.cpp),
// derives the kernel-facing layouts (adjacency CSR, edge rows, lower-edge
// slot lists, per-node face lists, block chunking) and uploads them once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <exception>
#include <mutex>
#include <string>
#include <vector>

#include "context.h"
#include "host_internal.h"
#include "host_mesh.h"

namespace aerofx {

int set_hip_error(hipError_t e, const char *what) {
  return set_error(AEROFX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int ctx_alloc(aerofx_ctx *c, size_t bytes, void **out) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 256));
  if (e != hipSuccess) return set_hip_error(e, "hipMalloc");
  c->allocs.push_back(p);
  *out = p;
  return 0;
}

void ctx_release(aerofx_ctx *c, void *p) {
  auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
  if (it == c->allocs.end()) return;
  (void)hipFree(p);
  c->allocs.erase(it);
}

namespace {

// ---------------------------------------------------------------------------
// Scratch of aerofx_residual_noalloc: one static array per GPU, holding the
// residual passes' V [N][5], gradients [N][16], upper-edge sums P [N][5] and
// edge fluxes Fm [E][5], each part at a 256-byte boundary.  Host code never
// passes the array's name to a kernel: it uses the device address that
// hipGetSymbolAddress returns for the current GPU, looked up once per GPU.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

// The current GPU's scratch address (nullptr + message on failure).
char *scratch_address(std::string &err) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    err = std::string("hipGetDevice: ") + hipGetErrorString(e);
    return nullptr;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    err = "device ordinal " + std::to_string(dev) + " beyond the scratch table";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<char *>(p);
  }
  return g_scratch_ptr[dev];
}

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

template <class T>
int upload(aerofx_ctx *c, const std::vector<T> &v, const T **dst) {
  void *p = nullptr;
  int rc = ctx_alloc(c, v.size() * sizeof(T), &p);
  if (rc) return rc;
  if (!v.empty()) {
    hipError_t e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_hip_error(e, "hipMemcpy");
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
  free_krylov(c);
  free_adjoint(c);
  free_tangent(c);
  free_ilu(c);
  free_fgmres(c);
  delete c->elem;  // device memory is in c->allocs
  free_block(c);
  for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->h_err) (void)hipHostFree(c->h_err);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

// Whole-row chunks with at most `limit` items and kMaxRows rows; a row with
// more than `limit` items is a chunk of its own (the kernels hand such
// chunks to their long-row variants) and is listed in *longrows.
std::vector<int> chunk_rows(const std::vector<int> &rowptr, int N, int limit, std::vector<int> *longrows) {
  std::vector<int> ch(1, 0);
  int acc = 0, rows = 0;
  for (int n = 0; n < N; ++n) {
    int cnt = rowptr[n + 1] - rowptr[n];
    if (cnt > limit) {
      if (rows) ch.push_back(n);
      ch.push_back(n + 1);
      longrows->push_back(n);
      acc = 0;
      rows = 0;
      continue;
    }
    if (acc + cnt > limit || rows == kMaxRows) {
      ch.push_back(n);
      acc = 0;
      rows = 0;
    }
    acc += cnt;
    rows += 1;
  }
  if (rows || ch.size() == 1) ch.push_back(N);
  return ch;
}

// Flux chunks (first row, end row, first edge, end edge): whole rows with at
// most `limit` upper edges and kMaxRows rows; a row with more upper edges is
// cut into pieces of `limit` edges with no rows (first row = end row): its
// flux sum is then formed by the node pass, in the same edge order.
std::vector<int4> flux_chunks(const std::vector<int> &eup, int N, int limit) {
  std::vector<int4> out;
  int r0 = 0, acc = 0;
  auto close = [&](int r1) {
    if (r1 > r0) out.push_back(make_int4(r0, r1, eup[r0], eup[r1]));
    r0 = r1;
    acc = 0;
  };
  for (int n = 0; n < N; ++n) {
    const int cnt = eup[n + 1] - eup[n];
    if (cnt > limit) {
      close(n);
      for (int l = eup[n]; l < eup[n + 1]; l += limit) out.push_back(make_int4(n, n, l, std::min(l + limit, eup[n + 1])));
      r0 = n + 1;
      continue;
    }
    if (acc + cnt > limit || n - r0 == kMaxRows) close(n);
    acc += cnt;
  }
  close(N);
  return out;
}

}  // namespace

ProfScope::ProfScope(aerofx_ctx *ctx, int b) : c(ctx), bucket(b) {
  if (c->prof && (a = take_event(c))) (void)hipEventRecord(a, c->stream);
}

ProfScope::~ProfScope() {
  if (!a) return;
  hipEvent_t b = take_event(c);
  if (!b) return;
  (void)hipEventRecord(b, c->stream);
  c->pending.push_back({bucket, a, b});
}

}  // namespace aerofx

using namespace aerofx;

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
    if (cd != AEROFX_BC_SLIP_WALL && cd != AEROFX_BC_INLET && cd != AEROFX_BC_OUTLET &&
        cd != AEROFX_BC_ADIABATIC_WALL && cd != AEROFX_BC_ISOTHERMAL_WALL)
      return set_error(AEROFX_ERR_MESH, "ctx_create: unsupported face code (1, 2, 3, 4, 5 only)");
  }
  aerofx_ctx *c = new aerofx_ctx();
  c->N = N;
  c->T = T;
  c->F = F;
  c->p = *p;
  c->gas.gam = p->gamma;
  c->gas.gam1 = p->gamma - 1.0;
  c->gas.invgam1 = 1.0 / c->gas.gam1;
  c->gas.pstiff = p->pstiff;
  c->fc.gas = c->gas;
  c->fc.beta = p->beta;
  {
    double Uf[5];
    freestream_state(*p, Uf);
    face_state(Uf, c->fc.ub_in);
    face_state(Uf, c->fc.ub_out);
  }
  int rc = 0;
  try {
    std::vector<int> &tet_edges = c->tet_edges;
    build_edges(N, T, m->tets, c->ptr, &tet_edges);
    c->tets.assign(m->tets, m->tets + 4 * (size_t)T);
    c->faces.assign(m->faces, m->faces + 3 * (size_t)F);
    c->fcodes.assign(m->face_codes, m->face_codes + (size_t)F);
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
      if (!(c->cv[i] > 0.0)) {
        destroy(c);
        return set_error(AEROFX_ERR_MESH, "ctx_create: non-positive control volume");
      }
    const int *ptr = c->ptr.data();
    // gradient pass: node adjacency in edge-id order
    std::vector<int> rowptr, adj, eid;
    node_csr(N, E, ptr, rowptr, adj, eid);
    std::vector<int> mirror(2 * (size_t)E);
    {
      std::vector<int> at_i(E), at_j(E);
      for (int n = 0; n < N; ++n)
        for (int e = rowptr[n]; e < rowptr[n + 1]; ++e) (n < adj[e] ? at_i : at_j)[eid[e]] = e;
      for (int n = 0; n < N; ++n)
        for (int e = rowptr[n]; e < rowptr[n + 1]; ++e) mirror[e] = n < adj[e] ? at_j[eid[e]] : at_i[eid[e]];
    }
    // flux pass A: edge rows by lower node; pass B: lower-edge slot lists
    std::vector<int> eup(N + 1, 0), ledge(E), lrow(N + 1, 0);
    std::vector<int2> eij(E);
    for (int l = 0; l < E; ++l) {
      eup[ptr[2 * (size_t)l] + 1]++;
      lrow[ptr[2 * (size_t)l + 1] + 1]++;
    }
    for (int n = 0; n < N; ++n) {
      eup[n + 1] += eup[n];
      lrow[n + 1] += lrow[n];
    }
    {
      std::vector<int> fill(lrow.begin(), lrow.end() - 1);
      for (int l = 0; l < E; ++l) {
        eij[l] = make_int2(ptr[2 * (size_t)l], ptr[2 * (size_t)l + 1]);
        ledge[fill[eij[l].y]++] = l;  // ascending l = ascending lower node
      }
    }
    // per-node boundary faces in face order
    std::vector<int> frow(N + 1, 0), fidx(3 * (size_t)F), fcode(F);
    for (long long k = 0; k < 3LL * F; ++k) frow[m->faces[k] + 1]++;
    for (int n = 0; n < N; ++n) frow[n + 1] += frow[n];
    {
      std::vector<int> fill(frow.begin(), frow.end() - 1);
      for (int f = 0; f < F; ++f) {
        fcode[f] = m->face_codes[f];
        for (int a = 0; a < 3; ++a) fidx[fill[m->faces[3 * f + a]]++] = f;
      }
    }
    // rows with more than kBlock neighbours: chunks of their own, served by
    // the long-row kernels of the Jacobian and H1 passes
    std::vector<int> glong;
    std::vector<int> gchunk = chunk_rows(rowptr, N, kBlock, &glong);
    for (int n = 0; n < N; ++n) c->max_degree = std::max(c->max_degree, rowptr[n + 1] - rowptr[n]);
    int maxup = 0;
    for (int n = 0; n < N; ++n) maxup = std::max(maxup, eup[n + 1] - eup[n]);
    const int fblock = maxup <= kFluxBlockSmall ? kFluxBlockSmall : kFluxBlock;
    // one 16-B record per flux chunk: a block reads its rows and edges at once
    std::vector<int4> fchunk4 = flux_chunks(eup, N, fblock);
    // node-derived tables
    std::vector<double> X4(4 * (size_t)N), lsq6(6 * (size_t)N), invcv(N), fnorm3(3 * (size_t)F);
    for (int n = 0; n < N; ++n) {
      for (int a = 0; a < 3; ++a) X4[4 * (size_t)n + a] = m->X[3 * (size_t)n + a];
      X4[4 * (size_t)n + 3] = 0.0;
      // computeLocalWeightsLeastSquares (SubDomain.C:137-159): node factors
      const double *R = c->lsq.data() + 6 * (size_t)n;
      double or11 = 1.0 / R[0], or22 = 1.0 / R[3], or33 = 1.0 / R[5];
      double *L = lsq6.data() + 6 * (size_t)n;
      L[0] = or11;
      L[1] = or22;
      L[2] = or33;
      L[3] = R[1] * or11;
      L[4] = R[4] * or22;
      L[5] = (R[1] * R[4] - R[2] * R[3]) * or11 * or22;
      invcv[n] = 1.0 / c->cv[n];  // SpaceOperator.C:615
    }
    const double third = 1.0 / 3.0;  // FaceTria::getNormal (FaceTriaCore.C:288-292)
    for (size_t k = 0; k < 3 * (size_t)F; ++k) fnorm3[k] = third * c->fnorm[k];
    DevTables &t = c->tab;
    t.N = N;
    t.E = E;
    t.F = F;
    if ((rc = upload(c, X4, &t.X4)) || (rc = upload(c, lsq6, &t.lsq6)) || (rc = upload(c, invcv, &t.invcv)) || (rc = upload(c, c->cv, &t.cv)) ||
        (rc = upload(c, gchunk, &t.gchunk)) || (rc = upload(c, glong, &t.glong)) || (rc = upload(c, rowptr, &t.rowptr)) ||
        (rc = upload(c, adj, &t.adj)) || (rc = upload(c, eid, &t.eidx)) || (rc = upload(c, mirror, &t.mirror)) || (rc = upload(c, fchunk4, &t.fchunk4)) || (rc = upload(c, eup, &t.eup)) ||
        (rc = upload(c, eij, &t.eij)) || (rc = upload(c, ledge, &t.ledge)) ||
        (rc = upload(c, c->normals, &t.normal)) || (rc = upload(c, lrow, &t.lrow)) ||
        (rc = upload(c, frow, &t.frow)) || (rc = upload(c, fidx, &t.fidx)) || (rc = upload(c, fcode, &t.fcode)) ||
        (rc = upload(c, fnorm3, &t.fnorm3)) || (rc = upload(c, c->fnorm, &t.fnorm))) {
      destroy(c);
      return rc;
    }
    t.ngchunks = (int)gchunk.size() - 1;
    t.nglong = (int)glong.size();
    t.nfchunks = (int)fchunk4.size();
    t.fblock = fblock;
    // LDS-staged strips (both passes or neither)
    {
      HaloStrips gs, fs;
      std::vector<int> jcol(E);
      for (int l = 0; l < E; ++l) jcol[l] = ptr[2 * (size_t)l + 1];
      if (halo_strips(N, rowptr.data(), adj.data(), kGradStrip, 1 << 30, kGradHalo, gs) &&
          halo_strips(N, eup.data(), jcol.data(), kFluxStrip, kFluxStrip, kFluxHalo, fs)) {
        const int ngb = (int)gs.own.size(), nfb = (int)fs.own.size();
        std::vector<int4> gb4(ngb), fb4(nfb);
        std::vector<int2> fbh(nfb);
        for (int b = 0; b < ngb; ++b) gb4[b] = make_int4(gs.start[b], gs.start[b + 1], gs.hptr[b], gs.hptr[b + 1]);
        for (int b = 0; b < nfb; ++b) {
          fb4[b] = make_int4(fs.start[b], fs.start[b + 1], eup[fs.start[b]], eup[fs.start[b + 1]]);
          fbh[b] = make_int2(fs.hptr[b], fs.hptr[b + 1]);
        }
        // position of each edge's lower node: rows are consecutive from own[b]
        std::vector<uint32_t> floc(E);
        for (int b = 0; b < nfb; ++b)
          for (int r = fs.start[b]; r < fs.start[b + 1]; ++r)
            for (int l = eup[r]; l < eup[r + 1]; ++l)
              floc[l] = (uint32_t)(fs.own[b] + (r - fs.start[b])) | ((uint32_t)fs.loc[l] << 16);
        if ((rc = upload(c, gb4, &t.gb4)) || (rc = upload(c, gs.own, &t.gown)) ||
            (rc = upload(c, gs.halo, &t.ghalo)) || (rc = upload(c, gs.loc, &t.gloc)) ||
            (rc = upload(c, fb4, &t.fb4)) || (rc = upload(c, fbh, &t.fbh)) || (rc = upload(c, fs.halo, &t.fhalo)) ||
            (rc = upload(c, floc, &t.floc))) {
          destroy(c);
          return rc;
        }
        t.ngb = ngb;
        t.nfb = nfb;
      }
    }
    void *pv = nullptr;
    if ((rc = ctx_alloc(c, 5 * sizeof(double) * (size_t)N, &pv))) { destroy(c); return rc; }
    c->d_V = static_cast<double *>(pv);
    if ((rc = ctx_alloc(c, 16 * sizeof(double) * (size_t)N, &pv))) { destroy(c); return rc; }
    c->d_grad = static_cast<double *>(pv);
    if ((rc = ctx_alloc(c, 5 * sizeof(double) * (size_t)E, &pv))) { destroy(c); return rc; }
    c->d_Fm = static_cast<double *>(pv);
    if ((rc = ctx_alloc(c, 5 * sizeof(double) * (size_t)N, &pv))) { destroy(c); return rc; }
    c->d_P = static_cast<double *>(pv);
    if ((rc = ctx_alloc(c, 5 * sizeof(double) * (size_t)N, &pv))) { destroy(c); return rc; }
    c->d_dt = static_cast<double *>(pv);
    if ((rc = ctx_alloc(c, 4 * sizeof(int), &pv))) { destroy(c); return rc; }
    c->d_err = static_cast<int *>(pv);
    hipError_t e = hipMemset(c->d_err, 0, 4 * sizeof(int));
    if (e != hipSuccess) { destroy(c); return set_hip_error(e, "hipMemset"); }
    e = hipHostMalloc((void **)&c->h_err, 4 * sizeof(int), hipHostMallocDefault);
    if (e != hipSuccess) { destroy(c); return set_hip_error(e, "hipHostMalloc"); }
    e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) { destroy(c); return set_hip_error(e, "hipStreamCreate"); }
    c->stream = c->own_stream;
    // the no-allocation residual's scratch on this GPU (a failure here only
    // disables that entry point, which then reports it)
    c->noalloc_scratch = scratch_address(c->noalloc_err);
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

int aerofx_ctx_set_stream(aerofx_ctx *c, void *stream, int use_context_stream) {
  if (!c) return set_error(AEROFX_ERR_ARG, "set_stream: null ctx");
  c->stream = use_context_stream ? c->own_stream : static_cast<hipStream_t>(stream);
  return AEROFX_OK;
}

int aerofx_ctx_set_residual_kernels(aerofx_ctx *c, int family) {
  if (!c) return set_error(AEROFX_ERR_ARG, "set_residual_kernels: null ctx");
  if (family != AEROFX_RESIDUAL_AUTO && family != AEROFX_RESIDUAL_PER_LANE && family != AEROFX_RESIDUAL_STAGED)
    return set_error(AEROFX_ERR_ARG,
                     "set_residual_kernels: AEROFX_RESIDUAL_AUTO, AEROFX_RESIDUAL_PER_LANE or AEROFX_RESIDUAL_STAGED");
  c->residual_kernels = family == AEROFX_RESIDUAL_STAGED && c->tab.ngb > 0 ? AEROFX_RESIDUAL_STAGED
                                                                            : AEROFX_RESIDUAL_PER_LANE;
  return c->residual_kernels;
}

int aerofx_ctx_set_low_mach(aerofx_ctx *c, const aerofx_low_mach *lm) {
  if (!c) return set_error(AEROFX_ERR_ARG, "set_low_mach: null ctx");
  if (!lm) {
    c->fc.low_mach = 0;
    return AEROFX_OK;
  }
  if (c->viscous) return set_error(AEROFX_ERR_STATE, "set_low_mach: not with the Navier-Stokes terms");
  if (!(lm->min_mach > 0.0) || !(lm->cutoff >= lm->min_mach) || !(lm->slope >= 0.0) ||
      !(lm->shock_reducer >= 0.0) || !(lm->irey >= 0.0))
    return set_error(AEROFX_ERR_ARG,
                     "set_low_mach: need 0 < min_mach <= cutoff, slope >= 0, shock_reducer >= 0, irey >= 0");
  c->fc.low_mach = 1;
  c->fc.lm = LowMach{lm->min_mach, lm->slope, lm->cutoff, lm->shock_reducer, lm->irey};
  return AEROFX_OK;
}

int aerofx_ctx_set_failsafe(aerofx_ctx *c, int mode) {
  if (!c) return set_error(AEROFX_ERR_ARG, "set_failsafe: null ctx");
  if (mode < 0 || mode > 2) return set_error(AEROFX_ERR_ARG, "set_failsafe: 0 (off), 1 (on) or 2 (always on)");
  if (mode && !c->d_fs) {
    void *p = nullptr;
    int rc = ctx_alloc(c, 3 * sizeof(int) * (size_t)c->N, &p);
    if (rc) return rc;
    c->d_fs = static_cast<int *>(p);
    hipError_t e = hipMemsetAsync(c->d_fs, 0, 3 * sizeof(int) * (size_t)c->N, c->stream);
    if (e != hipSuccess) return set_hip_error(e, "set_failsafe");
  }
  c->failsafe = mode;
  c->tab.fs_mask = mode ? c->d_fs + 2 * (size_t)c->N : nullptr;
  return AEROFX_OK;
}

int aerofx_ctx_set_krylov_check(aerofx_ctx *c, int check_final_res) {
  if (!c) return set_error(AEROFX_ERR_ARG, "set_krylov_check: null ctx");
  c->check_final_res = check_final_res ? 1 : 0;
  return AEROFX_OK;
}

int aerofx_ctx_reset_failsafe(aerofx_ctx *c) {
  if (!c) return set_error(AEROFX_ERR_ARG, "reset_failsafe: null ctx");
  if (c->d_fs) {
    hipError_t e = hipMemsetAsync(c->d_fs, 0, 3 * sizeof(int) * (size_t)c->N, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_hip_error(e, "reset_failsafe");
  }
  c->fs_any = false;
  return AEROFX_OK;
}

int aerofx_ctx_failsafe_nodes(aerofx_ctx *c, int *tags) {
  if (!c || !tags) return set_error(AEROFX_ERR_ARG, "failsafe_nodes: null argument");
  if (!c->d_fs) {
    std::fill(tags, tags + 3 * (size_t)c->N, 0);
    return AEROFX_OK;
  }
  hipError_t e = hipMemcpyAsync(tags, c->d_fs, 3 * sizeof(int) * (size_t)c->N, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? AEROFX_OK : set_hip_error(e, "failsafe_nodes");
}

int aerofx_ctx_sync(aerofx_ctx *c) {
  if (!c) return set_error(AEROFX_ERR_ARG, "sync: null ctx");
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? AEROFX_OK : set_hip_error(e, "hipStreamSynchronize");
}

int aerofx_ctx_profile(aerofx_ctx *c, int enable) {
  if (!c) return set_error(AEROFX_ERR_ARG, "ctx_profile: null ctx");
  collect_timers(c);
  for (int k = 0; k < AEROFX_NUM_TIMERS; ++k) {
    c->timer_ms[k] = 0.0;
    c->timer_n[k] = 0;
  }
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

int aerofx_error_count(aerofx_ctx *c, int *count) {
  if (!c || !count) return set_error(AEROFX_ERR_ARG, "error_count: null argument");
  hipError_t e = hipMemcpyAsync(c->h_err, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e != hipSuccess) return set_hip_error(e, "hipMemcpyAsync");
  e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "hipStreamSynchronize");
  *count = c->h_err[0];
  return AEROFX_OK;
}

int aerofx_residual_ex(aerofx_ctx *c, const double *U, double *R, int flags) {
  if (!c || !U || !R) return set_error(AEROFX_ERR_ARG, "residual: null argument");
  const ResidualBufs b{c->d_V, c->d_grad, c->d_Fm, c->d_P};
  if (flags & 2) return launch_residual_with(c, b, U, R, (flags & 1) == 0);
  int cnt = 0;
  int rc = residual_failsafe(c, b, U, R, (flags & 1) == 0, &cnt);
  return rc ? rc : cnt;
}

int aerofx_residual(aerofx_ctx *c, const double *U, double *R) { return aerofx_residual_ex(c, U, R, 0); }

int aerofx_residual_noalloc(aerofx_ctx *c, const double *U, double *R, int flags) {
  if (!c || !U || !R) return set_error(AEROFX_ERR_ARG, "residual_noalloc: null argument");
  if (!c->noalloc_scratch)
    return set_error(AEROFX_ERR_STATE, "residual_noalloc: no scratch address on this GPU: " + c->noalloc_err);
  const size_t oV = 0, oG = align256(5 * sizeof(double) * (size_t)c->N);
  const size_t oP = oG + align256(16 * sizeof(double) * (size_t)c->N);
  const size_t oF = oP + align256(5 * sizeof(double) * (size_t)c->N);
  const size_t need = oF + align256(5 * sizeof(double) * (size_t)c->E);
  if (need > kNoAllocScratchBytes)
    return set_error(AEROFX_ERR_ARG, "residual_noalloc: the mesh (" + std::to_string(c->N) + " nodes, " +
                                         std::to_string(c->E) + " edges) needs " + std::to_string(need) +
                                         " bytes of scratch; the static scratch holds " +
                                         std::to_string(kNoAllocScratchBytes) + " (use aerofx_residual_ex)");
  char *base = c->noalloc_scratch;
  const ResidualBufs b{reinterpret_cast<double *>(base + oV), reinterpret_cast<double *>(base + oG),
                       reinterpret_cast<double *>(base + oF), reinterpret_cast<double *>(base + oP)};
  if (flags & 2) return launch_residual_with(c, b, U, R, (flags & 1) == 0);
  int cnt = 0;
  int rc = residual_failsafe(c, b, U, R, (flags & 1) == 0, &cnt);
  return rc ? rc : cnt;
}

int aerofx_gradients(aerofx_ctx *c, const double *V, double *ddx, double *ddy, double *ddz) {
  if (!c || !V || !ddx || !ddy || !ddz) return set_error(AEROFX_ERR_ARG, "gradients: null argument");
  return launch_gradients_v(c, V, ddx, ddy, ddz);
}

int aerofx_primitive(aerofx_ctx *c, const double *U, double *V) {
  if (!c || !U || !V) return set_error(AEROFX_ERR_ARG, "primitive: null argument");
  return launch_primitive(c, U, V);
}

}  // extern "C"
