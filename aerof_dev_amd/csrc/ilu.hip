// ilu.hip -- restricted additive Schwarz with ILU(0) subdomain solves on gfx950
// (IluPrec<Scalar,5,double> with fill 0 and type RAS, KspPrec.C:127-320).
//
// Subdomains are ranges of B consecutive nodes of the context's numbering (a
// mesh numbered by RCM, as the reference's IluPrec renumbers its subdomains);
// a subdomain's matrix holds the H1 blocks that couple two of its nodes.
// SparseMat::numericILU (SparseMatrix.C:353-396) factors it in place, the
// pivot blocks inverted; SparseMat::lusol (SparseMatrix.C:406-466) solves; RAS
// keeps each subdomain's solution on its own nodes.
//
// One workgroup per subdomain.  Rows whose lower (forward) or upper
// (backward) neighbours inside the subdomain are all done form a level; the
// rows of a level are independent, so a level runs in parallel and levels
// are separated by workgroup barriers.  Every block operation keeps the
// reference's order (each entry a left-to-right sum of five products; a row's
// lower entries in ascending column order), so the factors and the solve are
// those of the sequential algorithm.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "context.h"
#include "device_jacobian.h"
#include "host_internal.h"
#include "host_mesh.h"

namespace aerofx {

namespace {

constexpr int kIluThreads = 256;
constexpr int kIluGroup = 32;                            // lanes per row in the factorization
constexpr int kIluMaxRow = 32;                           // entries per row (diagonal included)
constexpr int kIluMaxB = 1024;                           // nodes per subdomain

}  // namespace

struct Ilu {
  int B = 0, nsub = 0;
  size_t nnz = 0;
  int *rowptr = nullptr;   // [N+1] local pattern, global rows
  int *ju = nullptr;       // [N] entry of the diagonal block
  int *col = nullptr;      // [nnz] column, local to the subdomain
  int *src = nullptr;      // [nnz] adjacency entry of the H1 block (-1: the diagonal block)
  int *flev = nullptr;     // forward levels: per subdomain, offsets into frow (nlev+1 each, concatenated)
  int *flevptr = nullptr;  // [nsub+1] into flev
  int *frow = nullptr;     // [N] rows (global) in forward level order
  int *blev = nullptr, *blevptr = nullptr, *brow = nullptr;  // the same for the backward solve
  // per position of the forward / backward schedule: (row - subdomain start,
  // first entry, end entry, diagonal entry) of the solve's sweep over the row
  int4 *fdesc = nullptr, *bdesc = nullptr;
  // the factorization's per-entry tables: pivot block of the column's row
  // (lower entries), and the updates [pptr[k], pptr[k+1]) of lower entry k:
  // upper block pjj of the column's row into entry pmm of the row
  int *ldiag = nullptr, *pptr = nullptr, *pjj = nullptr, *pmm = nullptr;
  void *LU[2] = {nullptr, nullptr};  // [nnz][25] factors, FP64 / FP32
  int lu_precision = -1;
  long long lu_version = -1;
};

void free_ilu(aerofx_ctx *c) {
  delete c->ilu;  // device memory belongs to c->allocs
  c->ilu = nullptr;
}

namespace {

template <class T>
int upload_vec(aerofx_ctx *c, const std::vector<T> &v, T **dst) {
  void *p = nullptr;
  int rc = ctx_alloc(c, std::max<size_t>(v.size(), 1) * sizeof(T), &p);
  if (rc) return rc;
  if (!v.empty()) {
    hipError_t e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_hip_error(e, "ilu table upload");
  }
  *dst = static_cast<T *>(p);
  return 0;
}

// ---------------------------------------------------------------------------
// numericILU, level by level.  A group of 32 lanes factors one row: its
// blocks are staged in LDS (converted to T, as IluPrec::getData copies the
// H1 blocks into the Scalar matrix), the lower entries are processed in
// ascending column order (25 lanes, one block entry each), the pivot block is
// inverted by one lane, and the row goes back to global memory.  For each
// lower entry k the position of the pivot block of its column's row
// (ldiag[k]) and the updates it makes -- the upper blocks jj of that row
// whose column is present further along row i, and where (pmm) -- are
// listed on the host, so a step reads them directly instead of searching.
// FP32 factors take 16 rows per pass, FP64 8 (the LDS of the row stages).
// ---------------------------------------------------------------------------
template <class T>
struct IluFactorShape {
  static constexpr int groups = sizeof(T) == 4 ? 16 : 8;
  static constexpr int threads = groups * kIluGroup;
};

template <class T>
__global__ __launch_bounds__(IluFactorShape<T>::threads) void k_ilu_factor(
    int B, const int *__restrict__ rowptr, const int *__restrict__ ju, const int *__restrict__ src,
    const int *__restrict__ ldiag, const int *__restrict__ pptr, const int *__restrict__ pjj,
    const int *__restrict__ pmm, const int *__restrict__ flev, const int *__restrict__ flevptr,
    const int *__restrict__ frow, const T *__restrict__ D, const T *__restrict__ Boff, T *__restrict__ LU,
    int *__restrict__ err) {
  constexpr int G = IluFactorShape<T>::groups;
  __shared__ T rowb[G][kIluMaxRow * 25];
  __shared__ T mult[G][25];
  __shared__ int maxd;
  const int s = blockIdx.x;
  const int g = threadIdx.x / kIluGroup, lane = threadIdx.x % kIluGroup;
  const int r5 = lane / 5, c5 = lane - 5 * (lane / 5);
  const int l0 = flevptr[s], l1 = flevptr[s + 1];
  (void)B;
  for (int lev = l0; lev + 1 < l1; ++lev) {
    const int q0 = flev[lev], q1 = flev[lev + 1];
    // rows of this level, G at a time; all loop bounds below are uniform over
    // the workgroup, so every thread meets every barrier
    for (int base = q0; base < q1; base += G) {
      const int q = base + g;
      const bool active = q < q1;
      int gi = 0, r0 = 0, nr = 0, dpos = 0;
      if (threadIdx.x == 0) maxd = 0;
      __syncthreads();
      if (active) {
        gi = frow[q];
        r0 = rowptr[gi];
        nr = rowptr[gi + 1] - r0;
        dpos = ju[gi] - r0;
        for (int idx = lane; idx < nr * 25; idx += kIluGroup) {
          const int e = idx / 25, kk = idx - 25 * e;
          const int sq = src[r0 + e];
          rowb[g][idx] = sq < 0 ? D[blk64_base(gi) + (size_t)kk * kBlkStride]
                                : Boff[blk64_base(sq) + (size_t)kk * kBlkStride];
        }
        if (lane == 0) atomicMax(&maxd, dpos);
      }
      __syncthreads();
      const int nd = maxd;
      for (int e = 0; e < nd; ++e) {
        const bool go = active && e < dpos;
        const int k = r0 + e;
        T m = T(0);
        if (go && lane < 25) {
          // mult = a_k inv(U_ll)  (applyToDenseMatrix, DENSEMATRIXTIMESDENSEMATRIX5)
          const T *Ul = LU + 25 * (size_t)ldiag[k];
          const T *pa = rowb[g] + 25 * e + 5 * r5;
          m = pa[0] * Ul[c5] + pa[1] * Ul[5 + c5] + pa[2] * Ul[10 + c5] + pa[3] * Ul[15 + c5] + pa[4] * Ul[20 + c5];
        }
        __syncthreads();
        if (go && lane < 25) {
          rowb[g][25 * e + lane] = m;
          mult[g][lane] = m;
        }
        __syncthreads();
        if (go && lane < 25) {
          // a_m -= mult a_j over the upper entries j of row l present in row i
          const T *pm = mult[g] + 5 * r5;
          const int p1 = pptr[k + 1];
          for (int pp = pptr[k]; pp < p1; ++pp) {
            const T *pj = LU + 25 * (size_t)pjj[pp];
            const T res = pm[0] * pj[c5] + pm[1] * pj[5 + c5] + pm[2] * pj[10 + c5] + pm[3] * pj[15 + c5] +
                          pm[4] * pj[20 + c5];
            rowb[g][25 * pmm[pp] + lane] -= res;
          }
        }
        __syncthreads();
      }
      if (active && lane == 0) {
        T piv[25];
        for (int kk = 0; kk < 25; ++kk) piv[kk] = rowb[g][25 * dpos + kk];
        if (invert5<T>(piv) != 0) atomicAdd(err, 1);
        for (int kk = 0; kk < 25; ++kk) rowb[g][25 * dpos + kk] = piv[kk];
      }
      __syncthreads();
      if (active)
        for (int idx = lane; idx < nr * 25; idx += kIluGroup) LU[25 * (size_t)r0 + idx] = rowb[g][idx];
      // the next level (and the next batch of this level) reads these rows
      __threadfence_block();
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// lusol, level by level: five lanes per row (one per component), the
// subdomain's x in LDS.  Forward: x_i = y_i, then x_i -= a_k x_col(k) over the
// lower entries in order; backward: x_i -= a_k x_col(k) over the upper
// entries in order, then x_i = inv(U_ii) x_i.  A row's schedule record gives
// its entry range in one load, and its entries' columns and block rows are
// fetched eight at a time before any is used, so a level costs about two
// memory latencies rather than one per entry; the subtractions still run one
// entry at a time in ascending order.
// ---------------------------------------------------------------------------
constexpr int kIluBatch = 8;

template <class T>
__device__ __forceinline__ void ilu_row_sweep(int k0, int k1, int r5, const int *__restrict__ col,
                                              const T *__restrict__ LU, const double *xs, double &v) {
  for (int kb = k0; kb < k1; kb += kIluBatch) {
    int cl[kIluBatch];
    T a[kIluBatch][5];
#pragma unroll
    for (int u = 0; u < kIluBatch; ++u) {
      const int k = kb + u < k1 ? kb + u : kb;
      cl[u] = col[k];
      const T *pa = LU + 25 * (size_t)k + 5 * r5;
#pragma unroll
      for (int cc = 0; cc < 5; ++cc) a[u][cc] = pa[cc];
    }
    // past the row's end the batch repeats its first entry and subtracts
    // 0 (v - 0 is v): the loads stay ahead of every use
#pragma unroll
    for (int u = 0; u < kIluBatch; ++u) {
      const double *xb = xs + 5 * cl[u];
      const double t = (double)a[u][0] * xb[0] + (double)a[u][1] * xb[1] + (double)a[u][2] * xb[2] +
                       (double)a[u][3] * xb[3] + (double)a[u][4] * xb[4];
      v -= kb + u < k1 ? t : 0.0;
    }
  }
}

template <class T>
__global__ __launch_bounds__(kIluThreads) void k_ilu_solve(int N, int B, const int *__restrict__ col,
                                                           const int *__restrict__ flev,
                                                           const int *__restrict__ flevptr,
                                                           const int4 *__restrict__ fdesc,
                                                           const int *__restrict__ blev,
                                                           const int *__restrict__ blevptr,
                                                           const int4 *__restrict__ bdesc, const T *__restrict__ LU,
                                                           const double *__restrict__ y, double *__restrict__ x,
                                                           const int *__restrict__ done) {
  if (done && *done) return;
  __shared__ double xs[5 * kIluMaxB];
  __shared__ double tmp[kIluThreads];
  const int s = blockIdx.x;
  const int a0 = s * B, n = min(B, N - a0);
  const int tid = threadIdx.x;
  constexpr int kRows = kIluThreads / 5;  // rows per pass: 51 (lanes 255 idle)
  const int rr = tid / 5, r5 = tid - 5 * (tid / 5);
  for (int idx = tid; idx < 5 * n; idx += kIluThreads) xs[idx] = y[5 * (size_t)a0 + idx];
  __syncthreads();
  // forward: unit lower triangle
  for (int lev = flevptr[s]; lev + 1 < flevptr[s + 1]; ++lev) {
    const int q0 = flev[lev], q1 = flev[lev + 1];
    for (int base = q0; base < q1; base += kRows) {
      const int q = base + rr;
      if (rr < kRows && q < q1) {
        const int4 d = fdesc[q];
        double v = xs[5 * d.x + r5];
        ilu_row_sweep<T>(d.y, d.z, r5, col, LU, xs, v);
        xs[5 * d.x + r5] = v;
      }
      __syncthreads();
    }
  }
  // backward: upper triangle, then the inverted pivot
  for (int lev = blevptr[s]; lev + 1 < blevptr[s + 1]; ++lev) {
    const int q0 = blev[lev], q1 = blev[lev + 1];
    for (int base = q0; base < q1; base += kRows) {
      const int q = base + rr;
      const bool act = rr < kRows && q < q1;
      int4 d = make_int4(0, 0, 0, 0);
      T piv[5];
      if (act) {
        d = bdesc[q];
        const T *pd = LU + 25 * (size_t)d.w + 5 * r5;
#pragma unroll
        for (int cc = 0; cc < 5; ++cc) piv[cc] = pd[cc];
        double v = xs[5 * d.x + r5];
        ilu_row_sweep<T>(d.y, d.z, r5, col, LU, xs, v);
        tmp[tid] = v;
      }
      __syncthreads();
      if (act) {
        const double *tv = tmp + 5 * rr;
        xs[5 * d.x + r5] = (double)piv[0] * tv[0] + (double)piv[1] * tv[1] + (double)piv[2] * tv[2] +
                           (double)piv[3] * tv[3] + (double)piv[4] * tv[4];
      }
      __syncthreads();
    }
  }
  for (int idx = tid; idx < 5 * n; idx += kIluThreads) x[5 * (size_t)a0 + idx] = xs[idx];
}

}  // namespace

int ensure_ilu(aerofx_ctx *c, int B) {
  if (B < 1 || B > kIluMaxB) return set_error(AEROFX_ERR_ARG, "ilu: subdomain size must be in [1, 1024]");
  if (c->ilu && c->ilu->B == B) return 0;
  free_ilu(c);
  const int N = c->N;
  std::vector<int> rp, adj, eid;
  node_csr(N, c->E, c->ptr.data(), rp, adj, eid);
  Ilu *u = new Ilu();
  u->B = B;
  u->nsub = (N + B - 1) / B;
  std::vector<int> rowptr(N + 1), ju(N), col, src;
  col.reserve(rp[N] + N);
  src.reserve(rp[N] + N);
  for (int i = 0; i < N; ++i) {
    const int a0 = (i / B) * B, a1 = std::min(N, a0 + B);
    rowptr[i] = (int)col.size();
    bool diag = false;
    // adjacency rows are in ascending neighbour order: insert the diagonal in place
    for (int q = rp[i]; q < rp[i + 1]; ++q) {
      const int j = adj[q];
      if (j < a0 || j >= a1) continue;
      if (!diag && j > i) {
        ju[i] = (int)col.size();
        col.push_back(i - a0);
        src.push_back(-1);
        diag = true;
      }
      col.push_back(j - a0);
      src.push_back(q);
    }
    if (!diag) {
      ju[i] = (int)col.size();
      col.push_back(i - a0);
      src.push_back(-1);
    }
    if ((int)col.size() - rowptr[i] > kIluMaxRow) {
      delete u;
      return set_error(AEROFX_ERR_MESH, "ilu: a node couples to more than 31 nodes of its subdomain");
    }
  }
  rowptr[N] = (int)col.size();
  u->nnz = col.size();
  // level schedules per subdomain: forward on the lower entries, backward on
  // the upper entries (rows of a level in ascending order)
  std::vector<int> flev, flevptr(u->nsub + 1), frow, blev, blevptr(u->nsub + 1), brow;
  std::vector<int> depth(N);
  for (int s = 0; s < u->nsub; ++s) {
    const int a0 = s * B, a1 = std::min(N, a0 + B);
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<int> &lev = pass ? blev : flev, &rows = pass ? brow : frow;
      int nlev = 0;
      if (pass == 0) {
        for (int i = a0; i < a1; ++i) {
          int d = 0;
          for (int k = rowptr[i]; k < ju[i]; ++k) d = std::max(d, depth[a0 + col[k]] + 1);
          depth[i] = d;
          nlev = std::max(nlev, d + 1);
        }
      } else {
        for (int i = a1 - 1; i >= a0; --i) {
          int d = 0;
          for (int k = ju[i] + 1; k < rowptr[i + 1]; ++k) d = std::max(d, depth[a0 + col[k]] + 1);
          depth[i] = d;
          nlev = std::max(nlev, d + 1);
        }
      }
      (pass ? blevptr : flevptr)[s] = (int)lev.size();
      std::vector<int> cnt(nlev + 1, 0);
      for (int i = a0; i < a1; ++i) cnt[depth[i] + 1]++;
      for (int l = 0; l < nlev; ++l) cnt[l + 1] += cnt[l];
      const int base = (int)rows.size();
      for (int l = 0; l <= nlev; ++l) lev.push_back(base + cnt[l]);
      rows.resize(base + (a1 - a0));
      std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
      for (int i = a0; i < a1; ++i) rows[base + fillp[depth[i]]++] = i;
    }
  }
  flevptr[u->nsub] = (int)flev.size();
  blevptr[u->nsub] = (int)blev.size();
  // every row once per schedule, each after the rows it reads
  {
    std::vector<int> fl(N, -1), bl(N, -1);
    for (int s = 0; s < u->nsub; ++s) {
      for (int l = flevptr[s]; l + 1 < flevptr[s + 1]; ++l)
        for (int q = flev[l]; q < flev[l + 1]; ++q) fl[frow[q]] = fl[frow[q]] < 0 ? l : -2;
      for (int l = blevptr[s]; l + 1 < blevptr[s + 1]; ++l)
        for (int q = blev[l]; q < blev[l + 1]; ++q) bl[brow[q]] = bl[brow[q]] < 0 ? l : -2;
    }
    bool ok = (int)frow.size() == N && (int)brow.size() == N;
    for (int i = 0; i < N && ok; ++i) {
      const int a0 = (i / B) * B;
      ok = fl[i] >= 0 && bl[i] >= 0;
      for (int k = rowptr[i]; k < rowptr[i + 1] && ok; ++k) {
        const int j = a0 + col[k];
        if (k < ju[i]) ok = j < i && fl[j] < fl[i];
        else if (k > ju[i]) ok = j > i && bl[j] < bl[i];
        else ok = j == i;
      }
    }
    if (!ok) {
      delete u;
      return set_error(AEROFX_ERR_STATE, "ilu: inconsistent level schedule");
    }
  }
  // numericILU's updates, listed once: for lower entry k = (i, l) in order,
  // the upper entries jj of row l whose column is also a later entry of row i
  std::vector<int> ldiag(u->nnz, -1), pptr(u->nnz + 1, 0), pjj, pmm;
  for (int i = 0; i < N; ++i) {
    const int a0 = (i / B) * B;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      pptr[k] = (int)pjj.size();
      if (k >= ju[i]) continue;
      const int l = a0 + col[k];
      ldiag[k] = ju[l];
      for (int jj = ju[l] + 1; jj < rowptr[l + 1]; ++jj)
        for (int t = k + 1; t < rowptr[i + 1]; ++t)
          if (col[t] == col[jj]) {
            pjj.push_back(jj);
            pmm.push_back(t - rowptr[i]);
          }
    }
  }
  pptr[u->nnz] = (int)pjj.size();
  std::vector<int4> fdesc(N), bdesc(N);
  for (int q = 0; q < N; ++q) {
    const int fi = frow[q], bi = brow[q];
    fdesc[q] = make_int4(fi - (fi / B) * B, rowptr[fi], ju[fi], ju[fi]);
    bdesc[q] = make_int4(bi - (bi / B) * B, ju[bi] + 1, rowptr[bi + 1], ju[bi]);
  }
  int rc;
  if ((rc = upload_vec(c, fdesc, &u->fdesc)) || (rc = upload_vec(c, bdesc, &u->bdesc)) ||
      (rc = upload_vec(c, ldiag, &u->ldiag)) || (rc = upload_vec(c, pptr, &u->pptr)) ||
      (rc = upload_vec(c, pjj, &u->pjj)) || (rc = upload_vec(c, pmm, &u->pmm)) ||
      (rc = upload_vec(c, rowptr, &u->rowptr)) || (rc = upload_vec(c, ju, &u->ju)) || (rc = upload_vec(c, col, &u->col)) ||
      (rc = upload_vec(c, src, &u->src)) || (rc = upload_vec(c, flev, &u->flev)) ||
      (rc = upload_vec(c, flevptr, &u->flevptr)) || (rc = upload_vec(c, frow, &u->frow)) ||
      (rc = upload_vec(c, blev, &u->blev)) || (rc = upload_vec(c, blevptr, &u->blevptr)) ||
      (rc = upload_vec(c, brow, &u->brow))) {
    delete u;
    return rc;
  }
  c->ilu = u;
  return 0;
}

// numericILU of the current H1 (its precision) into the ILU factors; the
// zero-pivot count goes to d_err[1] (added to the block-Jacobi's).
int launch_ilu_factor(aerofx_ctx *c) {
  Ilu *u = c->ilu;
  if (!u) return set_error(AEROFX_ERR_STATE, "ilu: not set up");
  const int prec = c->jac_precision;
  if (!u->LU[prec]) {
    int rc = ctx_alloc(c, (prec ? sizeof(float) : sizeof(double)) * 25 * std::max<size_t>(u->nnz, 1), &u->LU[prec]);
    if (rc) return rc;
  }
  if (prec)
    hipLaunchKernelGGL(k_ilu_factor<float>, dim3(u->nsub), dim3(IluFactorShape<float>::threads), 0, c->stream, u->B,
                       u->rowptr, u->ju, u->src, u->ldiag, u->pptr, u->pjj, u->pmm, u->flev, u->flevptr, u->frow,
                       (const float *)c->d_jacD, (const float *)c->d_jacB, (float *)u->LU[1], c->d_err + 1);
  else
    hipLaunchKernelGGL(k_ilu_factor<double>, dim3(u->nsub), dim3(IluFactorShape<double>::threads), 0, c->stream,
                       u->B, u->rowptr, u->ju, u->src, u->ldiag, u->pptr, u->pjj, u->pmm, u->flev, u->flevptr,
                       u->frow, (const double *)c->d_jacD, (const double *)c->d_jacB, (double *)u->LU[0],
                       c->d_err + 1);
  u->lu_precision = prec;
  u->lu_version = c->jac_version;
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "ilu factor launch");
}

int launch_ilu_solve(aerofx_ctx *c, const double *y, double *x, const int *done) {
  Ilu *u = c->ilu;
  if (!u || u->lu_version != c->jac_version || u->lu_precision != c->jac_precision)
    return set_error(AEROFX_ERR_STATE, "ilu: factors do not belong to the current Jacobian");
  if (u->lu_precision)
    hipLaunchKernelGGL(k_ilu_solve<float>, dim3(u->nsub), dim3(kIluThreads), 0, c->stream, c->N, u->B, u->col,
                       u->flev, u->flevptr, u->fdesc, u->blev, u->blevptr, u->bdesc, (const float *)u->LU[1], y, x,
                       done);
  else
    hipLaunchKernelGGL(k_ilu_solve<double>, dim3(u->nsub), dim3(kIluThreads), 0, c->stream, c->N, u->B, u->col,
                       u->flev, u->flevptr, u->fdesc, u->blev, u->blevptr, u->bdesc, (const double *)u->LU[0], y, x,
                       done);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "ilu solve launch");
}

}  // namespace aerofx
