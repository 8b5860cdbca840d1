// jacobian.hip -- implicit operators on gfx950: local time step, the
// first-order approximate Jacobian H1, its product with a vector (the
// MatVecProdH1 surface), and the 5x5 block-Jacobi preconditioner (the
// JacobiPrec surface), with FP64 or FP32 block storage.
//
// H1 layout (per context, storage type T = double or float):
//   D  [N]        diagonal blocks: sum of edge and face Jacobians, scaled by
//                 1/V, plus alpha/dt on the diagonal
//   B  [2E]       one off-diagonal block per adjacency entry, in the node's
//                 entry order (= ascending edge id): entry (n, m) holds
//                 A_nm, i.e. A_ij of the edge for n = i and A_ji for n = j
//   Di [N]        inverse of D (block-Jacobi), LINPACK in precision T
// Each array holds 5x5 row-major blocks interleaved in groups of 64
// (blk64_base): thread-per-node and thread-per-entry kernels then read 64
// consecutive blocks' element q as one contiguous segment.
// All kernels are node-parallel with each node's sums in the reference's
// order; arithmetic is FP64 (FP32 blocks are widened on load, as
// DenseMatrixOp<float,5,25> does with double vectors).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "context.h"
#include "device_jacobian.h"
#include "host_internal.h"

namespace aerofx {

// ---------------------------------------------------------------------------
// local time step (Edge.C:87-143, Face.C:242-290, Domain.C:32-100), beta = 1
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sound_speed(const Gas &g, const double V[5]) {
  return sqrt(g.gam * (V[4] + g.pstiff) / V[0]);
}

__global__ __launch_bounds__(kBlock) void k_timestep(DevTables t, const double *__restrict__ V, Gas g, double cfl,
                                                     const double *__restrict__ idtv, double viscous,
                                                     double *__restrict__ dt) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= t.N) return;
  double Vn[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) Vn[k] = V[5 * (size_t)n + k];
  const double beta2 = 1.0;
  double idti = 0.0;
  for (int e = t.rowptr[n]; e < t.rowptr[n + 1]; ++e) {
    const int m = t.adj[e], l = t.eidx[e];
    const bool nlow = n < m;
    const double *nl = t.normal + 3 * (size_t)l;
    double S = sqrt(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]);
    double invS = 1.0 / S;
    double nn0 = invS * nl[0], nn1 = invS * nl[1], nn2 = invS * nl[2];
    double ndot = invS * 0.0;
    double Vmid[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double vm = V[5 * (size_t)m + k];
      Vmid[k] = 0.5 * (nlow ? Vn[k] + vm : vm + Vn[k]);
    }
    double a = sound_speed(g, Vmid);
    double un = Vmid[1] * nn0 + Vmid[2] * nn1 + Vmid[3] * nn2 - ndot;
    double coeff1 = (1.0 + beta2) * un;
    double q = (1.0 - beta2) * un, r = 2.0 * a;
    double coeff2 = sqrt(q * q + r * r);
    double c = nlow ? 0.5 * (coeff1 - coeff2) : 0.5 * (-coeff1 - coeff2);
    idti += (c < 0.0 ? c : 0.0) * S;
  }
  for (int q = t.frow[n]; q < t.frow[n + 1]; ++q) {
    // Face::computeTimeStep: full face normal, 1/3 of the face per node
    const double *nf = t.fnorm + 3 * (size_t)t.fidx[q];
    double S = sqrt(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]);
    double invS = 1.0 / S;
    double un = Vn[1] * (invS * nf[0]) + Vn[2] * (invS * nf[1]) + Vn[3] * (invS * nf[2]) - invS * 0.0;
    double a = sound_speed(g, Vn);
    double coeff1 = (1.0 + beta2) * un;
    double qq = (1.0 - beta2) * un, r = 2.0 * a;
    double coeff2 = sqrt(qq * qq + r * r);
    double c = 0.5 * (coeff1 - coeff2);
    idti += (c < 0.0 ? c : 0.0) * S / 3;
  }
  // Domain.C:91 with the viscous term weighted by Time.ViscosityParameter
  dt[n] = cfl * t.cv[n] / (-1.0 * idti + (idtv ? viscous * idtv[n] : 0.0 * 0.0));
}

// ---------------------------------------------------------------------------
// H1 assembly (SpaceOperator.C:2088-2150, Edge.C:4472-4526, Face.C:816-847,
// SubDomain.C:1744-1767, TimeState.C:731-756)
// ---------------------------------------------------------------------------
// Entry-parallel: a block covers whole rows (<= 256 adjacency entries, <= 64
// rows).  Phase 1: lane k evaluates the two ROEJAC6 blocks of entry e0 + k,
// stores the off-diagonal block (64 consecutive entries -> contiguous
// segments) and its diagonal contribution in LDS; lanes also take the rows'
// boundary-face slots and store each face Jacobian in a per-slot buffer.
// Phase 2: one lane per (row, element) adds the entries in entry order, then
// the faces in face order, scales by 1/V and adds 1/dt -- the reference's
// per-node order.
// Two roejac6 evaluations inline take 256 VGPRs + 18 AGPRs, one wave per
// SIMD; asking for two waves spills ~18 registers to scratch and is faster
// (assembly 50 -> 38 ms over a 96x32x32 FP64 steady solve; three waves spill
// 116 and are slower).
// Adjacency entry e = (n, m) of H1: its off-diagonal block A_nm (scaled by
// 1/V_n) into B, and its contribution to the diagonal block of n into dg
// (dF/dU_i for n = i, -dF/dU_j for n = j).  LM: the low-Mach (Roe-Turkel)
// edge blocks, with the edge length |X_j - X_i| (Edge.C:4493-4494).
template <class T, bool LM>
__device__ __forceinline__ void jac_entry(const DevTables &t, const double *__restrict__ V, const FluxConsts &fc,
                                          int n, int e, T *__restrict__ B, double dg[25]) {
  const Gas &g = fc.gas;
  const int m = t.adj[e], l = t.eidx[e];
  const double nl[3] = {t.normal[3 * (size_t)l], t.normal[3 * (size_t)l + 1], t.normal[3 * (size_t)l + 2]};
  T *Be = B + blk64_base(e);
  if (sqrt(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]) < 1e-18) {
#pragma unroll
    for (int k = 0; k < 25; ++k) {
      Be[k * kBlkStride] = (T)0;
      dg[k] = 0.0;
    }
    return;
  }
  double Vn[5], Vm[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    Vn[k] = V[5 * (size_t)n + k];
    Vm[k] = V[5 * (size_t)m + k];
  }
  const double voln = t.invcv[n];
  double length = 1.0;
  if (LM) {
    const double4 xn = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)n);
    const double4 xm = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)m);
    const double d0 = n < m ? xm.x - xn.x : xn.x - xm.x, d1 = n < m ? xm.y - xn.y : xn.y - xm.y,
                 d2 = n < m ? xm.z - xn.z : xn.z - xm.z;
    length = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  }
  double blk[25];
  if (n < m) {
    // n = i: A_ii += dF/dU_i, A_ij = dF/dU_j / V_i
    roejac6(1.0, g, nl, 0.0, Vn, Vm, 1, blk, LM, fc.lm, length);
#pragma unroll
    for (int k = 0; k < 25; ++k) dg[k] = blk[k];
    roejac6(1.0, g, nl, 0.0, Vm, Vn, 2, blk, LM, fc.lm, length);
#pragma unroll
    for (int k = 0; k < 25; ++k) Be[k * kBlkStride] = (T)(0.0 + blk[k] * voln);
  } else {
    // n = j: A_jj -= dF/dU_j, A_ji = -dF/dU_i / V_j
    roejac6(1.0, g, nl, 0.0, Vn, Vm, 2, blk, LM, fc.lm, length);
#pragma unroll
    for (int k = 0; k < 25; ++k) dg[k] = -blk[k];
    roejac6(1.0, g, nl, 0.0, Vm, Vn, 1, blk, LM, fc.lm, length);
#pragma unroll
    for (int k = 0; k < 25; ++k) Be[k * kBlkStride] = (T)(0.0 - blk[k] * voln);
  }
}

// Boundary-face Jacobian of face slot q (a face of node n) into fjac.
__device__ __forceinline__ void jac_face(const DevTables &t, const double *__restrict__ V, const FluxConsts &fc,
                                         int n, int q, double *__restrict__ fjac) {
  const int face = t.fidx[q];
  const int code = t.fcode[face];
  const double nv[3] = {t.fnorm3[3 * (size_t)face], t.fnorm3[3 * (size_t)face + 1], t.fnorm3[3 * (size_t)face + 2]};
  double Vn[5], jac[25];
#pragma unroll
  for (int k = 0; k < 5; ++k) Vn[k] = V[5 * (size_t)n + k];
  if (is_wall_code(code)) {
    wall_jac(fc.gas, nv, 0.0, Vn, jac);
  } else {
    const double *ub = (code == AEROFX_BC_INLET || code == -AEROFX_BC_INLET) ? fc.ub_in : fc.ub_out;
    farfield_jac(fc.gas, nv, 0.0, Vn, ub, jac);
  }
#pragma unroll
  for (int k = 0; k < 25; ++k) fjac[25 * (size_t)q + k] = jac[k];
}

// Two roejac6 evaluations inline take 256 VGPRs + 18 AGPRs, one wave per
// SIMD; asking for two waves spills ~18 registers to scratch and is faster
// (assembly 50 -> 38 ms over a 96x32x32 FP64 steady solve; three waves spill
// 116 and are slower).  A chunk that is one row of more than kBlock entries
// is left to k_jacobian_long.
template <class T, bool LM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_jacobian(DevTables t, const double *__restrict__ V, FluxConsts fc,
                                                     const double *__restrict__ dt, double *__restrict__ fjac,
                                                     T *__restrict__ D, T *__restrict__ B) {
  __shared__ double sd[25 * kPad];
  __shared__ int srow[kMaxRows + 1];
  const int c = xcd_chunk(blockIdx.x, t.ngchunks);
  const int r0 = t.gchunk[c], r1 = t.gchunk[c + 1];
  const int e0 = t.rowptr[r0], e1 = t.rowptr[r1];
  if (e1 - e0 > kBlock) return;
  const int tid = threadIdx.x, nr = r1 - r0;
  if (tid <= nr) srow[tid] = t.rowptr[r0 + tid] - e0;
  __syncthreads();
  if (tid < e1 - e0) {
    const int e = e0 + tid;
    // row of entry tid: the last row start <= tid
    int lo = 0, hi = nr;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (srow[mid] <= tid) lo = mid; else hi = mid;
    }
    double dg[25];
    jac_entry<T, LM>(t, V, fc, r0 + lo, e, B, dg);
#pragma unroll
    for (int k = 0; k < 25; ++k) sd[k * kPad + tid] = dg[k];
  }
  // boundary-face Jacobians of the block's rows, one per face slot
  const int q0 = t.frow[r0], q1 = t.frow[r1];
  for (int q = q0 + tid; q < q1; q += kBlock) {
    int lo = 0, hi = nr;  // row owning slot q
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (t.frow[r0 + mid] <= q) lo = mid; else hi = mid;
    }
    jac_face(t, V, fc, r0 + lo, q, fjac);
  }
  __syncthreads();
  for (int idx = tid; idx < nr * 25; idx += kBlock) {
    const int rr = idx / 25, k = idx - 25 * rr;
    const int n = r0 + rr;
    double s = 0.0;
    for (int q = srow[rr]; q < srow[rr + 1]; ++q) s += sd[k * kPad + q];
    for (int q = t.frow[n]; q < t.frow[n + 1]; ++q) s += fjac[25 * (size_t)q + k];
    s *= t.invcv[n];
    if (dt && k % 6 == 0) s += 1.0 / dt[n];
    D[blk64_base(n) + k * kBlkStride] = (T)s;
  }
}

// One row of more than kBlock entries per block: its entries in passes of
// kBlock, the diagonal sums carried across passes in the entries' order --
// the sums of k_jacobian, term for term.
template <class T, bool LM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_jacobian_long(DevTables t, const double *__restrict__ V, FluxConsts fc,
                                                          const double *__restrict__ dt, double *__restrict__ fjac,
                                                          T *__restrict__ D, T *__restrict__ B) {
  __shared__ double sd[25 * kPad];
  const int n = t.glong[blockIdx.x];
  const int e0 = t.rowptr[n], e1 = t.rowptr[n + 1];
  const int tid = threadIdx.x;
  double s = 0.0;  // lanes 0..24: element tid of the diagonal block
  for (int base = e0; base < e1; base += kBlock) {
    if (base + tid < e1) {
      double dg[25];
      jac_entry<T, LM>(t, V, fc, n, base + tid, B, dg);
#pragma unroll
      for (int k = 0; k < 25; ++k) sd[k * kPad + tid] = dg[k];
    }
    __syncthreads();
    const int cnt = min(kBlock, e1 - base);
    if (tid < 25)
      for (int q = 0; q < cnt; ++q) s += sd[tid * kPad + q];
    __syncthreads();
  }
  for (int q = t.frow[n] + tid; q < t.frow[n + 1]; q += kBlock) jac_face(t, V, fc, n, q, fjac);
  __syncthreads();
  if (tid < 25) {
    for (int q = t.frow[n]; q < t.frow[n + 1]; ++q) s += fjac[25 * (size_t)q + tid];
    s *= t.invcv[n];
    if (dt && tid % 6 == 0) s += 1.0 / dt[n];
    D[blk64_base(n) + tid * kBlkStride] = (T)s;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_invert(int N, const T *__restrict__ D, T *__restrict__ Di,
                                                   int *__restrict__ err) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  T a[25];
#pragma unroll
  for (int k = 0; k < 25; ++k) a[k] = D[blk64_base(n) + k * kBlkStride];
  // DiagMat::invert (DiagMatrix.C:44-82): blocks with a zero diagonal entry
  // are left as they are
  bool inv = true;
#pragma unroll
  for (int k = 0; k < 5; ++k)
    if (a[6 * k] == (T)0) inv = false;
  if (inv && invert5<T>(a) != 0) atomicAdd(err, 1);
#pragma unroll
  for (int k = 0; k < 25; ++k) Di[blk64_base(n) + k * kBlkStride] = a[k];
}

// c += A b (DENSEMATRIXTIMESVECTOR5, DenseMatrixOps.h:441-446); a points at
// element 0 of an interleaved block
template <class T>
__device__ __forceinline__ void amv5_add(const T *__restrict__ a, const double b[5], double c[5]) {
  double r[25];
#pragma unroll
  for (int k = 0; k < 25; ++k) r[k] = (double)a[k * kBlkStride];
#pragma unroll
  for (int i = 0; i < 5; ++i)
    c[i] += r[5 * i] * b[0] + r[5 * i + 1] * b[1] + r[5 * i + 2] * b[2] + r[5 * i + 3] * b[3] + r[5 * i + 4] * b[4];
}

// c += A^T b (DENSEMATRIXTRANSTIMESVECTOR5, DenseMatrixOps.h:448-453)
template <class T>
__device__ __forceinline__ void amtv5_add(const T *__restrict__ a, const double b[5], double c[5]) {
  double r[25];
#pragma unroll
  for (int k = 0; k < 25; ++k) r[k] = (double)a[k * kBlkStride];
#pragma unroll
  for (int i = 0; i < 5; ++i)
    c[i] += r[i] * b[0] + r[5 + i] * b[1] + r[10 + i] * b[2] + r[15 + i] * b[3] + r[20 + i] * b[4];
}

// y = H1 p (SubDomain::computeMatVecProdH1, SubDomain.C:3917-3953), or with
// TRANS y = H1^T p (SubDomain::computeMatVecProdH1transpose, SubDomain.C:3958-3990):
// row n gathers A_mn^T p_m from the mirror entry (m, n), in edge order.
// A block covers whole rows (<= 256 entries, <= 64 rows, the gradient-pass
// chunks): lane k multiplies entry e0 + k (64 consecutive entries read as
// contiguous segments), the products go through LDS, and each row sums its
// diagonal product and then its entries in order -- the sums of the
// reference's node loop followed by its edge loop.
template <class T, bool TRANS, bool MIRRORED = false>
__global__ __launch_bounds__(kBlock) void k_h1_apply(DevTables t, const T *__restrict__ D, const T *__restrict__ B,
                                                     const double *__restrict__ p, double *__restrict__ y,
                                                     const int *__restrict__ done) {
  if (done && *done) return;
  __shared__ double sc[5 * kPad];
  __shared__ double sd[5 * kMaxRows];
  const int c = xcd_chunk(blockIdx.x, t.ngchunks);
  const int r0 = t.gchunk[c], r1 = t.gchunk[c + 1];
  const int e0 = t.rowptr[r0], e1 = t.rowptr[r1];
  if (e1 - e0 > kBlock) return;  // a long row: k_h1_apply_long
  const int tid = threadIdx.x;
  if (tid < e1 - e0) {
    const int e = e0 + tid;
    const int m = t.adj[e];
    double pm[5], cc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 5; ++k) pm[k] = p[5 * (size_t)m + k];
    if (TRANS)
      amtv5_add(B + blk64_base(MIRRORED ? e : t.mirror[e]), pm, cc);
    else
      amv5_add(B + blk64_base(e), pm, cc);
#pragma unroll
    for (int k = 0; k < 5; ++k) sc[k * kPad + tid] = cc[k];
  }
  if (tid < r1 - r0) {
    const int n = r0 + tid;
    double pn[5], dd[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 5; ++k) pn[k] = p[5 * (size_t)n + k];
    if (TRANS)
      amtv5_add(D + blk64_base(n), pn, dd);
    else
      amv5_add(D + blk64_base(n), pn, dd);
#pragma unroll
    for (int k = 0; k < 5; ++k) sd[k * kMaxRows + tid] = dd[k];
  }
  __syncthreads();
  const int nr = r1 - r0;
  for (int idx = tid; idx < nr * 5; idx += kBlock) {
    const int rr = idx / 5, q = idx - 5 * rr;
    const int n = r0 + rr;
    double s = sd[q * kMaxRows + rr];
    const int b = t.rowptr[n] - e0, eend = t.rowptr[n + 1] - e0;
    for (int e = b; e < eend; ++e) s += sc[q * kPad + e];
    y[5 * (size_t)n + q] = s;
  }
}

// k_h1_apply for one row of more than kBlock entries per block: the entries'
// products in passes of kBlock through LDS, summed in entry order after the
// diagonal product.
template <class T, bool TRANS, bool MIRRORED = false>
__global__ __launch_bounds__(kBlock) void k_h1_apply_long(DevTables t, const T *__restrict__ D,
                                                          const T *__restrict__ B, const double *__restrict__ p,
                                                          double *__restrict__ y, const int *__restrict__ done) {
  if (done && *done) return;
  __shared__ double sc[5 * kPad];
  const int n = t.glong[blockIdx.x];
  const int e0 = t.rowptr[n], e1 = t.rowptr[n + 1];
  const int tid = threadIdx.x;
  double s = 0.0;  // lanes 0..4: component tid
  if (tid < 5) {
    double pn[5], dd[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 5; ++k) pn[k] = p[5 * (size_t)n + k];
    if (TRANS)
      amtv5_add(D + blk64_base(n), pn, dd);
    else
      amv5_add(D + blk64_base(n), pn, dd);
#pragma unroll
    for (int k = 0; k < 5; ++k)
      if (k == tid) s = dd[k];
  }
  for (int base = e0; base < e1; base += kBlock) {
    const int e = base + tid;
    if (e < e1) {
      const int m = t.adj[e];
      double pm[5], cc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 5; ++k) pm[k] = p[5 * (size_t)m + k];
      if (TRANS)
        amtv5_add(B + blk64_base(MIRRORED ? e : t.mirror[e]), pm, cc);
      else
        amv5_add(B + blk64_base(e), pm, cc);
#pragma unroll
      for (int k = 0; k < 5; ++k) sc[k * kPad + tid] = cc[k];
    }
    __syncthreads();
    const int cnt = min(kBlock, e1 - base);
    if (tid < 5)
      for (int q = 0; q < cnt; ++q) s += sc[tid * kPad + q];
    __syncthreads();
  }
  if (tid < 5) y[5 * (size_t)n + tid] = s;
}

// x = Di y (DiagMat::apply, DiagMatrix.C:86-100), or x = Di^T y
template <class T, bool TRANS>
__global__ __launch_bounds__(kBlock) void k_jacobi_apply(int N, const T *__restrict__ Di,
                                                         const double *__restrict__ y, double *__restrict__ x,
                                                         const int *__restrict__ done) {
  if (done && *done) return;
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  double yn[5], acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 5; ++k) yn[k] = y[5 * (size_t)n + k];
  if (TRANS)
    amtv5_add(Di + blk64_base(n), yn, acc);
  else
    amv5_add(Di + blk64_base(n), yn, acc);
#pragma unroll
  for (int k = 0; k < 5; ++k) x[5 * (size_t)n + k] = acc[k];
}

// V = (src - sum_{k < nsub} h_k Vb_k) * (*scale), x = Di V (or Di^T V): the
// Arnoldi step's last Gram-Schmidt update and normalisation of the previous
// vector, fused with the preconditioner apply
template <class T, bool TRANS>
__global__ __launch_bounds__(kBlock) void k_jacobi_norm(int N, const T *__restrict__ Di,
                                                        const double *__restrict__ src,
                                                        const double *__restrict__ scale,
                                                        const double *__restrict__ Vb, size_t ldv, int nsub,
                                                        const double *__restrict__ h, double *__restrict__ Vout,
                                                        double *__restrict__ x, const int *__restrict__ done) {
  if (done && *done) return;
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  const double sc = *scale;
  double yn[5], acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 5; ++k) yn[k] = src[5 * (size_t)n + k];
  for (int b = 0; b < nsub; ++b) {
    const double hb = h[b];
#pragma unroll
    for (int k = 0; k < 5; ++k) yn[k] -= hb * Vb[(size_t)b * ldv + 5 * (size_t)n + k];
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    yn[k] = sc * yn[k];
    Vout[5 * (size_t)n + k] = yn[k];
  }
  if (TRANS)
    amtv5_add(Di + blk64_base(n), yn, acc);
  else
    amv5_add(Di + blk64_base(n), yn, acc);
#pragma unroll
  for (int k = 0; k < 5; ++k) x[5 * (size_t)n + k] = acc[k];
}

// Gram-Schmidt update pass of Arnoldi step j with delayed reorthogonalisation,
// fused with the preconditioner (coef: a at [0, 32), d at [32, 65), 1/tau at
// [65], 1/(tau rho) at [66]):
//   v_j     = (u_j - sum_{k<j} a_k V_k) / tau                      (slot j, in place)
//   u_{j+1} = (w - sum_{k<j} d_k V_k - d_j v_j) / (tau rho)        (slot j+1)
//   x       = P u_{j+1}  (or P^T)
// When the cycle has stopped, v_j is still finished if column j belongs to
// the solution (jlast >= j), and nothing else is written.
template <class T, bool TRANS>
__global__ __launch_bounds__(kBlock) void k_dcgs_update(int N, const T *__restrict__ Di, double *__restrict__ V,
                                                        size_t ldv, int j, const double *__restrict__ w,
                                                        const double *__restrict__ coef,
                                                        const int *__restrict__ done, const int *__restrict__ jlast,
                                                        double *__restrict__ x) {
  const bool stop = *done != 0;
  if (stop && *jlast < j) return;
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  double *u = V + (size_t)j * ldv + 5 * (size_t)n;
  double vj[5], un[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    vj[k] = u[k];
    un[k] = stop ? 0.0 : w[5 * (size_t)n + k];
  }
  // basis vectors four at a time, loads of a group in flight together;
  // both combinations keep ascending k
  int b = 0;
  for (; b + 4 <= j; b += 4) {
    double vb[4][5];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int k = 0; k < 5; ++k) vb[q][k] = V[(size_t)(b + q) * ldv + 5 * (size_t)n + k];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double ab = coef[b + q], db = coef[kMaxBasisCoef + b + q];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        vj[k] -= ab * vb[q][k];
        un[k] -= db * vb[q][k];
      }
    }
  }
  for (; b < j; ++b) {
    const double ab = coef[b], db = coef[kMaxBasisCoef + b];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double vbk = V[(size_t)b * ldv + 5 * (size_t)n + k];
      vj[k] -= ab * vbk;
      un[k] -= db * vbk;
    }
  }
  const double it = coef[2 * kMaxBasisCoef + 1];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    vj[k] *= it;
    u[k] = vj[k];
  }
  if (stop) return;
  const double dj = coef[kMaxBasisCoef + j], itr = coef[2 * kMaxBasisCoef + 2];
  double *un_out = V + (size_t)(j + 1) * ldv + 5 * (size_t)n;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    un[k] = (un[k] - dj * vj[k]) * itr;
    un_out[k] = un[k];
  }
  if (TRANS)
    amtv5_add(Di + blk64_base(n), un, acc);
  else
    amv5_add(Di + blk64_base(n), un, acc);
#pragma unroll
  for (int k = 0; k < 5; ++k) x[5 * (size_t)n + k] = acc[k];
}

// BT block e = B block mirror[e] (entries (n, m) and (m, n) swap places), so
// that lane e of the transposed apply reads slot e: 64 consecutive slots per
// 512-B (FP64) segment instead of 64 scattered blocks
template <class T>
__global__ __launch_bounds__(kBlock) void k_mirror_blocks(int nent, const int *__restrict__ mirror,
                                                          const T *__restrict__ B, T *__restrict__ BT) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= nent) return;
  const T *src = B + blk64_base(mirror[e]);
  T *dst = BT + blk64_base(e);
#pragma unroll
  for (int k = 0; k < 25; ++k) dst[k * kBlkStride] = src[k * kBlkStride];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, what);
}

int ensure_jacobian_storage(aerofx_ctx *c, int precision) {
  const size_t esz = precision ? sizeof(float) : sizeof(double);
  const size_t need_b = esz * blk64_count(2 * (size_t)c->E), need_d = esz * blk64_count(c->N);
  void **st = c->jac_store[precision];
  if (!c->d_fjac) {
    void *p = nullptr;
    int rc = ctx_alloc(c, sizeof(double) * 25 * 3 * (size_t)std::max(c->F, 1), &p);
    if (rc) return rc;
    c->d_fjac = (double *)p;
  }
  if (!st[0]) {
    int rc;
    if ((rc = ctx_alloc(c, need_d, &st[0])) || (rc = ctx_alloc(c, need_b, &st[1])) ||
        (rc = ctx_alloc(c, need_d, &st[2])))
      return rc;
  }
  c->d_jacD = st[0];
  c->d_jacB = st[1];
  c->d_jacDi = st[2];
  c->jac_precision = precision;
  return 0;
}

int launch_timestep(aerofx_ctx *c, double cfl, double *dt) {
  const int nb = (c->N + kBlock - 1) / kBlock;
  double *idtv = nullptr;
  int rc = launch_visc_idtv(c, &idtv);
  if (rc) return rc;
  hipLaunchKernelGGL(k_timestep, dim3(nb), dim3(kBlock), 0, c->stream, c->tab, c->d_V, c->gas, cfl, idtv,
                     c->ns.viscous_cst, dt);
  return check_launch("timestep launch");
}

template <class T, bool LM>
static void jacobian_launch(aerofx_ctx *c, int nchunks, const double *dt) {
  hipLaunchKernelGGL((k_jacobian<T, LM>), dim3(nchunks), dim3(kBlock), 0, c->stream, c->tab, c->d_V, c->fc, dt,
                     c->d_fjac, (T *)c->d_jacD, (T *)c->d_jacB);
  if (c->tab.nglong)
    hipLaunchKernelGGL((k_jacobian_long<T, LM>), dim3(c->tab.nglong), dim3(kBlock), 0, c->stream, c->tab, c->d_V,
                       c->fc, dt, c->d_fjac, (T *)c->d_jacD, (T *)c->d_jacB);
}

int launch_jacobian(aerofx_ctx *c, const double *dt, int precision, const double *U) {
  int rc = ensure_jacobian_storage(c, precision);
  if (rc) return rc;
  c->jac_version++;
  const int nb = (c->N + kBlock - 1) / kBlock;
  const int nchunks = c->tab.ngchunks;
  {
    ProfScope ps(c, AEROFX_TIMER_FV_JAC);
    const bool lm = c->fc.low_mach != 0;
    if (precision && lm)
      jacobian_launch<float, true>(c, nchunks, dt);
    else if (precision)
      jacobian_launch<float, false>(c, nchunks, dt);
    else if (lm)
      jacobian_launch<double, true>(c, nchunks, dt);
    else
      jacobian_launch<double, false>(c, nchunks, dt);
  }
  if ((rc = check_launch("jacobian launch"))) return rc;
  // the Navier-Stokes Galerkin blocks, then the wall conditions
  // (MatVecProdH1::evaluate, MatVecProd.C:1080-1092)
  if ((rc = launch_visc_jacobian(c, U, precision))) return rc;
  {
    ProfScope ps(c, AEROFX_TIMER_PREC_SETUP);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->d_err + 1, 1, nullptr);
    if (precision)
      hipLaunchKernelGGL((k_invert<float>), dim3(nb), dim3(kBlock), 0, c->stream, c->N, (const float *)c->d_jacD,
                         (float *)c->d_jacDi, c->d_err + 1);
    else
      hipLaunchKernelGGL((k_invert<double>), dim3(nb), dim3(kBlock), 0, c->stream, c->N,
                         (const double *)c->d_jacD, (double *)c->d_jacDi, c->d_err + 1);
    if ((rc = check_launch("invert launch"))) return rc;
    // RAS / ILU(0): factors of the same assembly, in its precision
    if (c->prec_type == AEROFX_PREC_RAS_ILU0) {
      if ((rc = ensure_ilu(c, c->prec_subdomain)) || (rc = launch_ilu_factor(c))) return rc;
    }
  }
  return 0;
}

int launch_invert_f32(aerofx_ctx *c, hipStream_t stream, const float *D, float *Di) {
  const int nb = (c->N + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, stream, c->d_err + 1, 1, nullptr);
  hipLaunchKernelGGL((k_invert<float>), dim3(nb), dim3(kBlock), 0, stream, c->N, D, Di, c->d_err + 1);
  return check_launch("invert launch");
}

int ensure_mirror(aerofx_ctx *c) {
  if (!c->d_jacB) return 0;
  const int prec = c->jac_precision;
  if (c->mirror_version == c->jac_version && c->mirror_precision == prec) return 0;
  const size_t nent = 2 * (size_t)c->E;
  if (!c->jac_mirror[prec]) {
    const size_t esz = prec ? sizeof(float) : sizeof(double);
    int rc = ctx_alloc(c, esz * blk64_count(nent), &c->jac_mirror[prec]);
    if (rc) return rc;
  }
  if (nent) {
    const dim3 g((unsigned)((nent + kBlock - 1) / kBlock));
    if (prec)
      hipLaunchKernelGGL((k_mirror_blocks<float>), g, dim3(kBlock), 0, c->stream, (int)nent, c->tab.mirror,
                         (const float *)c->d_jacB, (float *)c->jac_mirror[prec]);
    else
      hipLaunchKernelGGL((k_mirror_blocks<double>), g, dim3(kBlock), 0, c->stream, (int)nent, c->tab.mirror,
                         (const double *)c->d_jacB, (double *)c->jac_mirror[prec]);
    int rc = check_launch("mirror launch");
    if (rc) return rc;
  }
  c->mirror_version = c->jac_version;
  c->mirror_precision = prec;
  return 0;
}

// The transposed apply reads the mirrored blocks when they belong to the
// current assembly (slot e, coalesced), else the blocks of the mirror
// entries in place.
template <class T>
static void h1_launch(aerofx_ctx *c, const double *p, double *y, const int *done, bool trans) {
  const int nb = c->tab.ngchunks;
  const bool mirrored = c->mirror_version == c->jac_version && c->mirror_precision == c->jac_precision;
  const int nl = c->tab.nglong;
  if (trans && mirrored) {
    hipLaunchKernelGGL((k_h1_apply<T, true, true>), dim3(nb), dim3(kBlock), 0, c->stream, c->tab,
                       (const T *)c->d_jacD, (const T *)c->jac_mirror[c->jac_precision], p, y, done);
    if (nl)
      hipLaunchKernelGGL((k_h1_apply_long<T, true, true>), dim3(nl), dim3(kBlock), 0, c->stream, c->tab,
                         (const T *)c->d_jacD, (const T *)c->jac_mirror[c->jac_precision], p, y, done);
  } else if (trans) {
    hipLaunchKernelGGL((k_h1_apply<T, true>), dim3(nb), dim3(kBlock), 0, c->stream, c->tab,
                       (const T *)c->d_jacD, (const T *)c->d_jacB, p, y, done);
    if (nl)
      hipLaunchKernelGGL((k_h1_apply_long<T, true>), dim3(nl), dim3(kBlock), 0, c->stream, c->tab,
                         (const T *)c->d_jacD, (const T *)c->d_jacB, p, y, done);
  } else {
    hipLaunchKernelGGL((k_h1_apply<T, false>), dim3(nb), dim3(kBlock), 0, c->stream, c->tab,
                       (const T *)c->d_jacD, (const T *)c->d_jacB, p, y, done);
    if (nl)
      hipLaunchKernelGGL((k_h1_apply_long<T, false>), dim3(nl), dim3(kBlock), 0, c->stream, c->tab,
                         (const T *)c->d_jacD, (const T *)c->d_jacB, p, y, done);
  }
}

int launch_h1_apply_flag(aerofx_ctx *c, const double *p, double *y, const int *done, bool trans) {
  if (!c->d_jacD) return set_error(AEROFX_ERR_STATE, "jvp: Jacobian not set up");
  if (c->jac_precision)
    h1_launch<float>(c, p, y, done, trans);
  else
    h1_launch<double>(c, p, y, done, trans);
  return check_launch("jvp launch");
}

int launch_h1_apply(aerofx_ctx *c, const double *p, double *y, bool trans) {
  ProfScope ps(c, AEROFX_TIMER_JAC_APPLY);
  return launch_h1_apply_flag(c, p, y, nullptr, trans);
}

template <class T>
static void jacobi_launch(aerofx_ctx *c, const double *y, double *x, const int *done, bool trans) {
  const int nb = (c->N + kBlock - 1) / kBlock;
  if (trans)
    hipLaunchKernelGGL((k_jacobi_apply<T, true>), dim3(nb), dim3(kBlock), 0, c->stream, c->N,
                       (const T *)c->d_jacDi, y, x, done);
  else
    hipLaunchKernelGGL((k_jacobi_apply<T, false>), dim3(nb), dim3(kBlock), 0, c->stream, c->N,
                       (const T *)c->d_jacDi, y, x, done);
}

int launch_prec_apply_flag(aerofx_ctx *c, const double *y, double *x, const int *done, bool trans) {
  if (!c->d_jacDi) return set_error(AEROFX_ERR_STATE, "prec_apply: preconditioner not set up");
  if (c->jac_precision)
    jacobi_launch<float>(c, y, x, done, trans);
  else
    jacobi_launch<double>(c, y, x, done, trans);
  return check_launch("prec launch");
}

template <class T>
static void jacobi_norm_launch(aerofx_ctx *c, const double *src, const double *scale, const double *Vb,
                               size_t ldv, int nsub, const double *h, double *V, double *x, const int *done,
                               bool trans) {
  const int nb = (c->N + kBlock - 1) / kBlock;
  if (trans)
    hipLaunchKernelGGL((k_jacobi_norm<T, true>), dim3(nb), dim3(kBlock), 0, c->stream, c->N, (const T *)c->d_jacDi,
                       src, scale, Vb, ldv, nsub, h, V, x, done);
  else
    hipLaunchKernelGGL((k_jacobi_norm<T, false>), dim3(nb), dim3(kBlock), 0, c->stream, c->N,
                       (const T *)c->d_jacDi, src, scale, Vb, ldv, nsub, h, V, x, done);
}

int launch_prec_norm_flag(aerofx_ctx *c, const double *src, const double *scale, const double *Vb, size_t ldv,
                          int nsub, const double *h, double *V, double *x, const int *done, bool trans) {
  if (!c->d_jacDi) return set_error(AEROFX_ERR_STATE, "prec_apply: preconditioner not set up");
  if (c->jac_precision)
    jacobi_norm_launch<float>(c, src, scale, Vb, ldv, nsub, h, V, x, done, trans);
  else
    jacobi_norm_launch<double>(c, src, scale, Vb, ldv, nsub, h, V, x, done, trans);
  return check_launch("prec launch");
}

template <class T>
static void dcgs_update_launch(aerofx_ctx *c, double *V, size_t ldv, int j, const double *w, const double *coef,
                               const int *done, const int *jlast, double *x, bool trans) {
  const int nb = (c->N + kBlock - 1) / kBlock;
  if (trans)
    hipLaunchKernelGGL((k_dcgs_update<T, true>), dim3(nb), dim3(kBlock), 0, c->stream, c->N,
                       (const T *)c->d_jacDi, V, ldv, j, w, coef, done, jlast, x);
  else
    hipLaunchKernelGGL((k_dcgs_update<T, false>), dim3(nb), dim3(kBlock), 0, c->stream, c->N,
                       (const T *)c->d_jacDi, V, ldv, j, w, coef, done, jlast, x);
}

int launch_dcgs_update(aerofx_ctx *c, double *V, size_t ldv, int j, const double *w, const double *coef,
                       const int *done, const int *jlast, double *x, bool trans) {
  if (!c->d_jacDi) return set_error(AEROFX_ERR_STATE, "prec_apply: preconditioner not set up");
  if (c->jac_precision)
    dcgs_update_launch<float>(c, V, ldv, j, w, coef, done, jlast, x, trans);
  else
    dcgs_update_launch<double>(c, V, ldv, j, w, coef, done, jlast, x, trans);
  return check_launch("dcgs update launch");
}

int launch_prec_apply(aerofx_ctx *c, const double *y, double *x, bool trans) {
  ProfScope ps(c, AEROFX_TIMER_PREC_APPLY);
  if (!trans && c->prec_type == AEROFX_PREC_RAS_ILU0) return launch_ilu_solve(c, y, x, nullptr);
  return launch_prec_apply_flag(c, y, x, nullptr, trans);
}

}  // namespace aerofx
