// block_gmres.hip -- batched right-hand sides: block GMRES(m) on H1 with
// block-Jacobi, s = 16 columns at a time, for many sensitivity systems
// A x_k = b_k sharing one operator.
//
// The reference declares a multi-RHS GmresSolver::solve(VecSet&, VecSet&)
// (KspSolver.C:988-1135) that does not compile (it uses members the class
// does not have) and solves sensitivities one at a time
// (FluidShapeOptimizationHandler.C:2378-2428); this is a design of its own.
//
// Layout: a block vector is [n_pad][16] FP64 (dof-major, the 16 columns of a
// dof contiguous: 128 B rows), n = 5N rounded up to a multiple of 64 with
// zero rows.  The operator reads each H1 block once for all 16 columns.
//
// Per cycle (one hipGraph replay when use_graph != 0):
//   V_0 S_0 = R                      CholQR2
//   for j < m:
//     Z = P V_j,  W = A Z            batched block-Jacobi and H1 apply
//     H_j = V^T W; W -= V H_j        block CGS2 (two passes) with FP64 MFMA
//     H_j' = V^T W; W -= V H_j'      tall-skinny GEMMs (v_mfma_f64_16x16x4f64)
//     V_{j+1} H_{j+1,j} = W          CholQR2
// then on the host the small least-squares problem min ||E_1 S_0 - Hbar Y||
// (Givens on the block-Hessenberg band), and X += P (V Y).
//
// Precision: inner 0 = FP64 blocks; 1 = FP32 copies of H1 and of the
// block-Jacobi inverse for the inner cycles, while every cycle starts from
// the FP64 residual R = B - A64 X (iterative refinement) and the basis,
// Gram-Schmidt and update stay FP64.
//
// Determinism: V^T W is summed over a fixed split of the rows into chunks
// (one wave per chunk, rows in order), and the chunk partials in a fixed
// tree, so the tuned launch shape changes which unit does the work, not the
// arithmetic.  The autotuner times each candidate shape once per context
// and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.h"
#include "host_internal.h"

namespace aerofx {

constexpr int kS = 16;          // block width
constexpr int kTile = kS * kS;  // one 16x16 block
constexpr int kChunks = 1024;   // fixed row split of the column norms
constexpr int kMaxM = 16;       // block steps per cycle

typedef double f64x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// layout conversion
// ---------------------------------------------------------------------------
__global__ void k_to_block(size_t n, size_t npad, int nrhs, const double *__restrict__ src, double *__restrict__ dst,
                           unsigned seed) {
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= npad * kS) return;
  const size_t i = idx / kS;
  const int c = (int)(idx % kS);
  double v = 0.0;
  if (i < n) {
    if (c < nrhs) {
      v = src[(size_t)c * n + i];
    } else {
      // padding columns: a fixed pseudo-random vector keeps the block full rank
      unsigned h = (unsigned)(i * 2654435761u) ^ (unsigned)(c * 40503u) ^ seed;
      h ^= h >> 16;
      h *= 0x7feb352du;
      h ^= h >> 15;
      h *= 0x846ca68bu;
      h ^= h >> 16;
      v = (double)(h & 0xffffff) / 16777216.0 - 0.5;
    }
  }
  dst[idx] = v;
}

__global__ void k_from_block(size_t n, int nrhs, const double *__restrict__ src, double *__restrict__ dst) {
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= n * (size_t)nrhs) return;
  const int c = (int)(idx / n);
  const size_t i = idx % n;
  dst[idx] = src[i * kS + c];
}

// ---------------------------------------------------------------------------
// batched operators: one thread per (node, column)
// ---------------------------------------------------------------------------
// acc += A p[:, c] for one interleaved 5x5 block (element q at a[q * 64]),
// in the single-vector operators' order
template <class T>
__device__ __forceinline__ void bmv5_add(const T *__restrict__ a, const double *__restrict__ p, int c,
                                         double acc[5]) {
  double b[5], r[25];
#pragma unroll
  for (int k = 0; k < 5; ++k) b[k] = p[k * kS + c];
#pragma unroll
  for (int k = 0; k < 25; ++k) r[k] = (double)a[k * kBlkStride];
#pragma unroll
  for (int i = 0; i < 5; ++i)
    acc[i] += r[5 * i] * b[0] + r[5 * i + 1] * b[1] + r[5 * i + 2] * b[2] + r[5 * i + 3] * b[3] + r[5 * i + 4] * b[4];
}

// Block H1 apply, entry-parallel: a block covers whole rows (<= NE entries,
// <= NR rows); kS/NC lanes take one entry, NC columns each, so a wave reads
// the interleaved blocks of 64/(kS/NC) consecutive entries at once and the
// 16-column rows of P in NC-wide pieces.  Products go through LDS and each
// (row, component, column) sums its diagonal product and then its entries in
// order: per column exactly the single-vector operator's sums.  Two shapes:
// 32 entries / 8 rows / 2 columns per lane (25 KB of LDS, six blocks per CU)
// when every node has at most 32 neighbours, else 64 / 16 / 4 (50 KB).
constexpr int kBEntSmall = 32, kBRowsSmall = 8, kBEnt = 64, kBRows = 16;

template <int NC>
__device__ __forceinline__ void loadc(const double *p, double b[NC]) {
  if (NC == 4) {
    const double4 v = *reinterpret_cast<const double4 *>(p);
    b[0] = v.x;
    b[1] = v.y;
    b[2] = v.z;
    b[3] = v.w;
  } else {
    const double2 v = *reinterpret_cast<const double2 *>(p);
    b[0] = v.x;
    b[1] = v.y;
  }
}
template <int NC>
__device__ __forceinline__ void loadc(const float *p, double b[NC]) {
  if (NC == 4) {
    const float4 v = *reinterpret_cast<const float4 *>(p);
    b[0] = v.x;
    b[1] = v.y;
    b[2] = v.z;
    b[3] = v.w;
  } else {
    const float2 v = *reinterpret_cast<const float2 *>(p);
    b[0] = v.x;
    b[1] = v.y;
  }
}

template <int NC, class T, class TP>
__device__ __forceinline__ void bmv5xc(const T *__restrict__ a, const TP *__restrict__ p, int cg,
                                       double acc[5][NC]) {
  double b[5][NC], r[25];
#pragma unroll
  for (int k = 0; k < 5; ++k) loadc<NC>(p + k * kS + NC * cg, b[k]);
#pragma unroll
  for (int k = 0; k < 25; ++k) r[k] = (double)a[k * kBlkStride];
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int u = 0; u < NC; ++u)
      acc[i][u] = r[5 * i] * b[0][u] + r[5 * i + 1] * b[1][u] + r[5 * i + 2] * b[2][u] + r[5 * i + 3] * b[3][u] +
                  r[5 * i + 4] * b[4][u];
}

template <class T, class TP, int NE, int NR, int NC>
__global__ __launch_bounds__(256) void k_bh1_apply(DevTables t, int nchunks, const int *__restrict__ chunk,
                                                   const T *__restrict__ D, const T *__restrict__ B,
                                                   const TP *__restrict__ P, double *__restrict__ Y) {
  static_assert(NE * kS / NC == 256, "kS/NC lanes per entry, 256 lanes per block");
  __shared__ double sc[NE * 5 * kS];
  __shared__ double sd[NR * 5 * kS];
  const int c = xcd_chunk(blockIdx.x, nchunks);
  const int r0 = chunk[c], r1 = chunk[c + 1];
  const int e0 = t.rowptr[r0], e1 = t.rowptr[r1];
  const int tid = threadIdx.x, ent = tid / (kS / NC), cg = tid % (kS / NC);
  const int ne = e1 - e0, nr = r1 - r0;
  double acc[5][NC];
  if (ent < ne) {
    const int e = e0 + ent;
    bmv5xc<NC>(B + blk64_base(e), P + (size_t)t.adj[e] * 5 * kS, cg, acc);
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
      for (int u = 0; u < NC; ++u) sc[(ent * 5 + k) * kS + NC * cg + u] = acc[k][u];
  }
  if (ent < nr) {
    const int n = r0 + ent;
    bmv5xc<NC>(D + blk64_base(n), P + (size_t)n * 5 * kS, cg, acc);
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
      for (int u = 0; u < NC; ++u) sd[(ent * 5 + k) * kS + NC * cg + u] = 0.0 + acc[k][u];
  }
  __syncthreads();
  for (int idx = tid; idx < nr * 5 * kS; idx += 256) {
    const int rr = idx / (5 * kS), q = idx - rr * 5 * kS;
    const int n = r0 + rr;
    double sum = sd[rr * 5 * kS + q];
    const int b = t.rowptr[n] - e0, eend = t.rowptr[n + 1] - e0;
    for (int e = b; e < eend; ++e) sum += sc[e * 5 * kS + q];
    Y[(size_t)n * 5 * kS + q] = sum;
  }
}

template <class T, class TO>
__global__ __launch_bounds__(256) void k_bjacobi_apply(int N, const T *__restrict__ Di, const double *__restrict__ P,
                                                       TO *__restrict__ Y) {
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int n = gid / kS, c = gid % kS;
  if (n >= N) return;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  bmv5_add(Di + blk64_base(n), P + (size_t)n * 5 * kS, c, acc);
#pragma unroll
  for (int k = 0; k < 5; ++k) Y[((size_t)n * 5 + k) * kS + c] = (TO)acc[k];
}

// Y = B - Y (refinement residual), over the padded block
__global__ void k_bsub_from(size_t len, const double *__restrict__ B, double *__restrict__ Y) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < len) Y[i] = B[i] - Y[i];
}

__global__ void k_badd(size_t len, const double *__restrict__ Z, double *__restrict__ X) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < len) X[i] += Z[i];
}

__global__ void k_f64_to_f32(size_t len, const double *__restrict__ a, float *__restrict__ b) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < len) b[i] = (float)a[i];
}

// ---------------------------------------------------------------------------
// tall-skinny GEMMs on the FP64 matrix cores
// ---------------------------------------------------------------------------
// partial[c][b] = sum over the rows of chunk c of V_b^T W (16x16 tiles).
// v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16] and B[l/16][l%16] and
// holds D[l/16 + 4r][l%16], r < 4 (checked on gfx950).  Here A = V_b^T over 4 rows, B = W over
// the same rows: both are the 4x16 row slab [k0 + l/16][l%16].
// With TWO, the same sweep also forms V_b^T W2 into partial2 (one read of
// the V_b for two right-hand blocks).  NBMAX (>= nb) sizes the accumulators:
// fewer registers, more waves per SIMD for the early steps of a cycle.  The
// 12-block instance fits 254 VGPRs with no accumulation registers when asked
// for two waves per SIMD (it used 254 + 16, one wave): 16-RHS mixed solve at
// 128^3 3.92 -> 3.76 s.  The 17-block one would spill.
template <bool TWO, int NBMAX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NBMAX == 12 ? 2 : 1))) void k_tsg_dots(size_t npad, size_t rows_per_chunk, int nchunks, int mp1, int nb,
                           const double *__restrict__ V, size_t ldv, const double *__restrict__ W,
                           const double *__restrict__ W2, double *__restrict__ partial,
                           double *__restrict__ partial2) {
  const int lane = threadIdx.x & 63;
  const int chunk = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (chunk >= nchunks) return;
  const size_t r0 = (size_t)chunk * rows_per_chunk;
  const size_t r1 = min(npad, r0 + rows_per_chunk);
  f64x4 acc[NBMAX], acc2[TWO ? NBMAX : 1];
#pragma unroll
  for (int b = 0; b < NBMAX; ++b) acc[b] = (f64x4){0.0, 0.0, 0.0, 0.0};
  if (TWO) {
#pragma unroll
    for (int b = 0; b < NBMAX; ++b) acc2[TWO ? b : 0] = (f64x4){0.0, 0.0, 0.0, 0.0};
  }
  const int rr = lane >> 4, cc = lane & 15;
  // rows_per_chunk is a multiple of 16: four MFMA steps per iteration, their
  // loads issued together
  for (size_t k0 = r0; k0 < r1; k0 += 16) {
    double w[4], w2[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      w[u] = W[(k0 + 4 * u + rr) * kS + cc];
      if (TWO) w2[u] = W2[(k0 + 4 * u + rr) * kS + cc];
    }
#pragma unroll
    for (int b = 0; b < NBMAX; ++b)
      if (b < nb) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = V[(size_t)b * ldv + (k0 + 4 * u + rr) * kS + cc];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[u], w[u], acc[b], 0, 0, 0);
          if (TWO) acc2[TWO ? b : 0] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[u], w2[u], acc2[TWO ? b : 0], 0, 0, 0);
        }
      }
  }
#pragma unroll
  for (int b = 0; b < NBMAX; ++b)
    if (b < nb) {
      double *dst = partial + ((size_t)chunk * mp1 + b) * kTile;
#pragma unroll
      for (int r = 0; r < 4; ++r) dst[(rr + 4 * r) * kS + cc] = acc[b][r];
      if (TWO) {
        double *dst2 = partial2 + ((size_t)chunk * mp1 + b) * kTile;
#pragma unroll
        for (int r = 0; r < 4; ++r) dst2[(rr + 4 * r) * kS + cc] = acc2[TWO ? b : 0][r];
      }
    }
}

// Two-level fixed-order reduction of the chunk partials.  Level 1: block
// (b, g) sums chunks [g*nc/G, (g+1)*nc/G) of tile b into part2[g][b] (lane e
// owns element e); level 2: one block per tile sums the G group sums.
constexpr int kRedGroups = 64;
__global__ __launch_bounds__(256) void k_tsg_reduce1(const double *__restrict__ partial, int nchunks, int mp1,
                                                     double *__restrict__ part2) {
  const int b = blockIdx.x, g = blockIdx.y, e = threadIdx.x;
  const int per = (nchunks + kRedGroups - 1) / kRedGroups;
  const int c0 = g * per, c1 = min(nchunks, c0 + per);
  double s = 0.0;
  for (int c = c0; c < c1; ++c) s += partial[((size_t)c * mp1 + b) * kTile + e];
  part2[((size_t)g * mp1 + b) * kTile + e] = s;
}

__global__ __launch_bounds__(256) void k_tsg_reduce2(const double *__restrict__ part2, int mp1,
                                                     double *__restrict__ out) {
  const int b = blockIdx.x, e = threadIdx.x;
  double s = 0.0;
  for (int g = 0; g < kRedGroups; ++g) s += part2[((size_t)g * mp1 + b) * kTile + e];
  out[(size_t)b * kTile + e] = s;
}

// Out = beta Out + alpha sum_b V_b H_b over 16-row tiles, one wave per tile.
// A = V_b rows (16 x 4 per step), B = H_b rows (4 x 16).
// Out may be one of the V_b (in place): a wave reads all of its 16 rows of
// every V_b before it writes them.
__global__ __launch_bounds__(512) void k_tsg_update(size_t npad, int nb, const double *V, size_t ldv,
                             const double *__restrict__ H, double alpha, double beta, double *Out) {
  const int lane = threadIdx.x & 63;
  const size_t tile = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (tile * 16 >= npad) return;
  const size_t r0 = tile * 16;
  const int ri = lane & 15, kk = lane >> 4;
  f64x4 acc = (f64x4){0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < nb; ++b) {
    const double *Vb = V + (size_t)b * ldv + (r0 + ri) * kS;
    const double *Hb = H + (size_t)b * kTile;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Vb[4 * q + kk], Hb[(4 * q + kk) * kS + ri], acc, 0, 0, 0);
  }
  // D[kk + 4 r][ri] -> row r0 + kk + 4 r, column ri
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double *o = Out + (r0 + kk + 4 * r) * kS + ri;
    *o = (beta == 0.0 ? 0.0 : beta * *o) + alpha * acc[r];
  }
}

// Both updates of a block Arnoldi step in one sweep, one wave per 16-row tile:
//   V_j     = sum_{b<j} V_b cV_b + U_j cV_j                  (slot j, in place)
//   U_{j+1} = sum_{b<j} V_b cU_b + V_j cU_j + W cU_{j+1}      (slot j+1 holds W, in place)
// The slots 0..j-1 are read once for both sums; the new V_j tile goes
// through LDS from the MFMA output layout to the input layout.
__global__ __launch_bounds__(512) void k_tsg_update2(size_t npad, int j, double *V, size_t ldv,
                                                     const double *__restrict__ cV, const double *__restrict__ cU) {
  __shared__ double sv[8][kTile];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t tile = (size_t)blockIdx.x * (blockDim.x >> 6) + wave;
  if (tile * 16 >= npad) return;
  const size_t r0 = tile * 16;
  const int ri = lane & 15, kk = lane >> 4;
  // MFMA step q takes k = 4 kk + q from lane (ri, kk): a lane's four A
  // values are one contiguous 32-B piece of its row
  f64x4 a1 = (f64x4){0.0, 0.0, 0.0, 0.0}, a2 = (f64x4){0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < j; ++b) {
    const double4 x4 = *reinterpret_cast<const double4 *>(V + (size_t)b * ldv + (r0 + ri) * kS + 4 * kk);
    const double x[4] = {x4.x, x4.y, x4.z, x4.w};
    const double *Hv = cV + (size_t)b * kTile, *Hu = cU + (size_t)b * kTile;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(x[q], Hv[(4 * kk + q) * kS + ri], a1, 0, 0, 0);
      a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(x[q], Hu[(4 * kk + q) * kS + ri], a2, 0, 0, 0);
    }
  }
  double *Uj = V + (size_t)j * ldv, *Wn = V + (size_t)(j + 1) * ldv;
  const double4 w4 = *reinterpret_cast<const double4 *>(Wn + (r0 + ri) * kS + 4 * kk);
  const double4 u4 = *reinterpret_cast<const double4 *>(Uj + (r0 + ri) * kS + 4 * kk);
  const double w[4] = {w4.x, w4.y, w4.z, w4.w}, uu[4] = {u4.x, u4.y, u4.z, u4.w};
#pragma unroll
  for (int q = 0; q < 4; ++q)
    a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(uu[q], cV[(size_t)j * kTile + (4 * kk + q) * kS + ri], a1, 0, 0, 0);
  // V_j: D[kk + 4 r][ri] -> row r0 + kk + 4 r, column ri
  double *t = sv[wave];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    t[(kk + 4 * r) * kS + ri] = a1[r];
    Uj[(r0 + kk + 4 * r) * kS + ri] = a1[r];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int q = 0; q < 4; ++q)
    a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(t[ri * kS + 4 * kk + q], cU[(size_t)j * kTile + (4 * kk + q) * kS + ri],
                                              a2, 0, 0, 0);
#pragma unroll
  for (int q = 0; q < 4; ++q)
    a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(w[q], cU[(size_t)(j + 1) * kTile + (4 * kk + q) * kS + ri], a2, 0, 0,
                                              0);
#pragma unroll
  for (int r = 0; r < 4; ++r) Wn[(r0 + kk + 4 * r) * kS + ri] = a2[r];
}

// ---------------------------------------------------------------------------
// 16x16 helpers (one lane): Cholesky G = R^T R, inverse of R, products
// ---------------------------------------------------------------------------
// R upper triangular with G = R^T R; Rinv = R^-1.  A non-positive pivot (a
// numerically rank-deficient block) restarts the factorisation of G + s I
// with s = 1e-13 max diag(G) (shifted Cholesky-QR; the second Cholesky-QR
// pass restores orthogonality); flag = 1 only if that fails too.
__device__ void chol16(const double *G, double *R, double *Rinv, int *flag) {
  double a[kTile], r[kTile], ri[kTile];
  double dmax = 0.0;
  for (int e = 0; e < kTile; ++e) {
    a[e] = G[e];
    ri[e] = 0.0;
  }
  for (int j = 0; j < kS; ++j) dmax = fmax(dmax, a[j * kS + j]);
  bool ok = false;
  for (int attempt = 0; attempt < 2 && !ok; ++attempt) {
    const double shift = attempt == 0 ? 0.0 : 1e-13 * dmax + 1e-300;
    for (int e = 0; e < kTile; ++e) r[e] = 0.0;
    ok = true;
    for (int j = 0; j < kS && ok; ++j) {
      double d = a[j * kS + j] + shift;
      for (int k = 0; k < j; ++k) d -= r[k * kS + j] * r[k * kS + j];
      if (!(d > 0.0)) {
        ok = false;
        break;
      }
      const double rjj = sqrt(d);
      r[j * kS + j] = rjj;
      for (int i = j + 1; i < kS; ++i) {
        double s = a[j * kS + i];
        for (int k = 0; k < j; ++k) s -= r[k * kS + j] * r[k * kS + i];
        r[j * kS + i] = s / rjj;
      }
    }
  }
  if (!ok) {
    *flag = 1;
    for (int j = 0; j < kS; ++j)
      if (!(r[j * kS + j] > 0.0)) r[j * kS + j] = 1.0;
  }
  // back substitution for R^-1 (upper)
  for (int j = kS - 1; j >= 0; --j) {
    ri[j * kS + j] = 1.0 / r[j * kS + j];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += r[i * kS + k] * ri[k * kS + j];
      ri[i * kS + j] = -s / r[i * kS + i];
    }
  }
  for (int e = 0; e < kTile; ++e) {
    R[e] = r[e];
    Rinv[e] = ri[e];
  }
}

__global__ void k_chol16(const double *__restrict__ G, double *__restrict__ R, double *__restrict__ Rinv,
                         int *__restrict__ flag) {
  if (threadIdx.x == 0) chol16(G, R, Rinv, flag);
}

// C = A B (16x16), and optionally D += E
__global__ void k_mm16(const double *__restrict__ A, const double *__restrict__ B, double *__restrict__ C) {
  const int e = threadIdx.x;
  if (e >= kTile) return;
  const int i = e / kS, j = e % kS;
  double s = 0.0;
  for (int k = 0; k < kS; ++k) s += A[i * kS + k] * B[k * kS + j];
  C[e] = s;
}

// ---------------------------------------------------------------------------
// block Gram-Schmidt with delayed reorthogonalisation (block DCGS2): the
// 16x16 tile algebra of one Arnoldi step, one block of 256 threads (one per
// tile element; thread 0 runs the Cholesky factorisations)
// ---------------------------------------------------------------------------
// c[e] = sum_l op(a)[i][l] b[l][k] for the thread's element e = (i, k)
__device__ __forceinline__ double tdot(const double *a, const double *b, bool ta) {
  const int i = threadIdx.x / kS, k = threadIdx.x % kS;
  double s = 0.0;
#pragma unroll
  for (int l = 0; l < kS; ++l) s += (ta ? a[l * kS + i] : a[i * kS + l]) * b[l * kS + k];
  return s;
}

// At step j the slots hold V_0..V_{j-1} (final), U_j = W1_{j-1} P^-1 (the
// previous step's once-projected block, P from its Gram estimate), and
// W = A' U_j in slot j+1.  G = V^T U_j and G2 = V^T W over slots 0..j+1 give
// A_b, alpha = U^T U, B_b, beta = U^T W, gamma = W^T W.  Then (tiles):
//   T^T T = alpha - sum A_b^T A_b;  V_j = (U_j - sum V_b A_b) T^-1
//   column j-1 of H: C_b + A_b P (b < j), T P (row j)     [S0 <- T S0 at j = 0]
//   G_k = sum_c H_kc A_c;  p_k = B_k (k < j), p_j = T^-T (beta - sum A_b^T B_b)
//   h_k = (p_k - G_k) T^-1
//   Gram = T^-T (gamma - sum (G_k^T p_k + p_k^T G_k - G_k^T G_k)) T^-1 - sum h_k^T h_k = P'^T P'
//   U_{j+1} = W M - sum_{k<=j} V_k D_k,  M = T^-1 P'^-1,  D_k = (G_k T^-1 + h_k) P'^-1
// and (C, P) <- (h, P').  The update coefficients go to cV (V_j over slots
// 0..j) and cU (U_{j+1} over slots 0..j+1).  FINAL (j = m): only T and column
// m-1.  Work tiles: wk[0..7] scalars-like tiles, then G_k, p_j and h_k.
constexpr int kWkTiles = 8 + 3 * (kMaxM + 2);
template <bool FINAL>
__global__ __launch_bounds__(256) void k_bdcgs_tiles(const double *__restrict__ G, const double *__restrict__ G2,
                                                     double *__restrict__ Hbar, int m, int j, double *__restrict__ Cp,
                                                     double *__restrict__ Pp, double *__restrict__ S0,
                                                     double *__restrict__ cV, double *__restrict__ cU,
                                                     double *__restrict__ wk, int *__restrict__ flag) {
  const int e = threadIdx.x;
  double *S = wk, *T = wk + kTile, *Ti = wk + 2 * kTile, *Pn = wk + 3 * kTile, *Pi = wk + 4 * kTile;
  double *tmp = wk + 6 * kTile, *tmp2 = wk + 7 * kTile;
  double *Gk = wk + 8 * kTile, *pj = Gk + (size_t)(kMaxM + 2) * kTile, *hk = pj + kTile;
  auto At = [&](int b) { return G + (size_t)b * kTile; };
  auto Bt = [&](int b) { return G2 + (size_t)b * kTile; };
  auto Ht = [&](int r, int c) { return Hbar + ((size_t)r * m + c) * kTile; };
  // T^T T = alpha - sum A_b^T A_b
  {
    double sv = At(j)[e];
    for (int b = 0; b < j; ++b) sv -= tdot(At(b), At(b), true);
    S[e] = sv;
  }
  __syncthreads();
  if (e == 0) chol16(S, T, Ti, flag);
  __syncthreads();
  if (j >= 1) {
    for (int b = 0; b < j; ++b) Ht(b, j - 1)[e] = Cp[(size_t)b * kTile + e] + tdot(At(b), Pp, false);
    Ht(j, j - 1)[e] = tdot(T, Pp, false);
  } else {
    tmp[e] = tdot(T, S0, false);
    __syncthreads();
    S0[e] = tmp[e];
  }
  if (FINAL) return;
  __syncthreads();
  // G_k = sum_{c<j} H_kc A_c (H_kc = 0 for k > c + 1)
  for (int k = 0; k <= j; ++k) {
    double sv = 0.0;
    for (int c = k > 0 ? k - 1 : 0; c < j; ++c) sv += tdot(Ht(k, c), At(c), false);
    Gk[(size_t)k * kTile + e] = sv;
  }
  // p_j = T^-T (beta - sum A_b^T B_b)
  {
    double sv = Bt(j)[e];
    for (int b = 0; b < j; ++b) sv -= tdot(At(b), Bt(b), true);
    tmp[e] = sv;
  }
  __syncthreads();
  pj[e] = tdot(Ti, tmp, true);
  __syncthreads();
  auto pk = [&](int k) { return k < j ? Bt(k) : pj; };
  // h_k = (p_k - G_k) T^-1, and Q = gamma - sum (G_k^T p_k + p_k^T G_k - G_k^T G_k)
  {
    double q = Bt(j + 1)[e];
    for (int k = 0; k <= j; ++k) {
      const double *g = Gk + (size_t)k * kTile, *pp = pk(k);
      q -= tdot(g, pp, true) + tdot(pp, g, true) - tdot(g, g, true);
    }
    tmp[e] = q;
  }
  for (int k = 0; k <= j; ++k) {
    const double *g = Gk + (size_t)k * kTile, *pp = pk(k);
    const int i = e / kS, c = e % kS;
    double sv = 0.0;
#pragma unroll
    for (int l = 0; l < kS; ++l) sv += (pp[i * kS + l] - g[i * kS + l]) * Ti[l * kS + c];
    hk[(size_t)k * kTile + e] = sv;
  }
  __syncthreads();
  // Gram = T^-T Q T^-1 - sum h_k^T h_k
  tmp2[e] = tdot(tmp, Ti, false);
  __syncthreads();
  {
    double sv = tdot(Ti, tmp2, true);
    for (int k = 0; k <= j; ++k) sv -= tdot(hk + (size_t)k * kTile, hk + (size_t)k * kTile, true);
    S[e] = sv;
  }
  __syncthreads();
  if (e == 0) chol16(S, Pn, Pi, flag);
  __syncthreads();
  // M = T^-1 P'^-1 -> cU[j+1]; D_k = (G_k T^-1 + h_k) P'^-1 -> cU[k] = -D_k
  cU[(size_t)(j + 1) * kTile + e] = tdot(Ti, Pi, false);
  for (int k = 0; k <= j; ++k) {
    tmp[e] = tdot(Gk + (size_t)k * kTile, Ti, false) + hk[(size_t)k * kTile + e];
    __syncthreads();
    cU[(size_t)k * kTile + e] = -tdot(tmp, Pi, false);
    __syncthreads();
  }
  // V_j = U_j T^-1 - sum V_b (A_b T^-1)
  for (int b = 0; b < j; ++b) cV[(size_t)b * kTile + e] = -tdot(At(b), Ti, false);
  cV[(size_t)j * kTile + e] = Ti[e];
  // (C, P) <- (h, P')
  for (int k = 0; k <= j; ++k) Cp[(size_t)k * kTile + e] = hk[(size_t)k * kTile + e];
  Pp[e] = Pn[e];
}

// Y[:, c] = X[:, c] / d_c with d_c = sqrt(nrm2[c]) (1 for a zero column)
__global__ void k_scale_cols(size_t len, const double *__restrict__ X, const double *__restrict__ nrm2,
                             double *__restrict__ Y) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= len) return;
  const double d = sqrt(nrm2[i % kS]);
  Y[i] = d > 0.0 ? X[i] / d : X[i];
}

// S[:, c] *= d_c (undoes the column scaling in the triangular factor)
__global__ void k_scale_tile_cols(double *__restrict__ S, const double *__restrict__ nrm2) {
  const int e = threadIdx.x;
  if (e >= kTile) return;
  const double d = sqrt(nrm2[e % kS]);
  if (d > 0.0) S[e] *= d;
}

// column sums of squares of a block vector: partial over a fixed row split
__global__ __launch_bounds__(256) void k_colnorm_partial(size_t npad, const double *__restrict__ X,
                                                         double *__restrict__ partial) {
  // 256 threads = 16 row lanes x 16 columns; block b covers a fixed row range
  const int c = threadIdx.x % kS, rl = threadIdx.x / kS;
  const size_t per = (npad + kChunks - 1) / kChunks;
  const size_t r0 = blockIdx.x * per, r1 = min(npad, r0 + per);
  double s = 0.0;
  for (size_t r = r0 + rl; r < r1; r += 16) {
    const double v = X[r * kS + c];
    s += v * v;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < kS) {
    double t = 0.0;
    for (int q = 0; q < 16; ++q) t += sh[q * kS + threadIdx.x];
    partial[(size_t)blockIdx.x * kS + threadIdx.x] = t;
  }
}

__global__ void k_colnorm_reduce(const double *__restrict__ partial, double *__restrict__ out) {
  const int c = threadIdx.x;
  if (c >= kS) return;
  double s = 0.0;
  for (int b = 0; b < kChunks; ++b) s += partial[(size_t)b * kS + c];
  out[c] = s;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct TsgShape {
  int waves_dots = 4;    // waves per workgroup of k_tsg_dots
  int waves_update = 4;  // waves per workgroup of k_tsg_update
  bool tuned = false;
};

struct BlockSolver {
  int m = 0;
  size_t n = 0, npad = 0;
  double *V = nullptr;        // [m+1][npad][16]
  double *W = nullptr, *Z = nullptr, *R = nullptr, *X = nullptr, *B = nullptr, *T = nullptr;
  int nchunks = 0;            // row split of V^T W (fixed by npad)
  double *partial = nullptr;  // [nchunks][m+1][256]
  double *partial2 = nullptr; // [nchunks][m+1][256], second right-hand block of a fused sweep
  double *part2 = nullptr;    // [kRedGroups][m+1][256]
  double *Hbar = nullptr;     // [(m+1)][m] tiles: block (i, j) at (i * m + j)
  double *S0 = nullptr;       // tile
  double *G = nullptr;        // scratch tiles [kMaxM+1]
  double *G2 = nullptr;       // scratch tiles [kMaxM+1]
  double *Rt = nullptr, *Ri = nullptr, *Rt2 = nullptr, *Ri2 = nullptr;  // tiles
  double *Y = nullptr;        // [m] tiles
  double *Cp = nullptr, *Pp = nullptr;   // once-projected coefficients (kMaxM+1 tiles) and scale tile of the newest block
  double *cV = nullptr, *cU = nullptr;   // update coefficients (kMaxM+2 tiles each)
  double *wk = nullptr;                  // kWkTiles work tiles of the step algebra
  double *nrm = nullptr;      // [16]
  int *flag = nullptr;
  float *D32 = nullptr, *B32 = nullptr, *Di32 = nullptr;
  float *Zf = nullptr;        // [npad][16] FP32 preconditioned block (input of the FP32 H1 apply)
  // row chunks of the block H1 apply: [0] <= kBEnt entries, <= kBRows rows;
  // [1] <= kBEntSmall, <= kBRowsSmall (built when every node fits)
  int *bchunk[2] = {nullptr, nullptr};
  int nbchunks[2] = {0, 0};
  bool bsmall = false;
  long long f32_version = -1;     // assembly the FP32 copies were made from
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_prec = -1;
  TsgShape shape;
  std::vector<double> hH, hS0, hY;
  double *h_nrm = nullptr;
  int *h_flag = nullptr;
};

void free_block(aerofx_ctx *c);

namespace {

inline dim3 g1(size_t n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

int alloc_d(aerofx_ctx *c, size_t count, double **p) {
  void *v = nullptr;
  int rc = ctx_alloc(c, count * sizeof(double), &v);
  *p = (double *)v;
  return rc;
}

BlockSolver *get_block(aerofx_ctx *c, int m, int *rc);

// out[b] = V_b^T W (and out2[b] = V_b^T W2 when W2 is given), b < nb
void tsg_dots(BlockSolver *s, int nb, const double *V, const double *W, double *out, const double *W2 = nullptr,
              double *out2 = nullptr) {
  // rows per chunk: a multiple of 16 (four MFMA steps), fixed by npad alone
  const size_t rows = std::max<size_t>(16, ((s->npad + s->nchunks - 1) / s->nchunks + 15) / 16 * 16);
  const int wpb = s->shape.waves_dots;
  const dim3 grid((s->nchunks + wpb - 1) / wpb), block(64 * wpb);
#define TSG_DOTS(TWO, NBM)                                                                                     \
  hipLaunchKernelGGL((k_tsg_dots<TWO, NBM>), grid, block, 0, s->stream, s->npad, rows, s->nchunks, s->m + 1, nb, V, \
                     s->npad * kS, W, W2, s->partial, s->partial2)
  if (W2) {
    if (nb <= 4) TSG_DOTS(true, 4);
    else if (nb <= 8) TSG_DOTS(true, 8);
    else if (nb <= 12) TSG_DOTS(true, 12);
    else TSG_DOTS(true, kMaxM + 1);
  } else {
    if (nb <= 4) TSG_DOTS(false, 4);
    else if (nb <= 8) TSG_DOTS(false, 8);
    else if (nb <= 12) TSG_DOTS(false, 12);
    else TSG_DOTS(false, kMaxM + 1);
  }
#undef TSG_DOTS
  hipLaunchKernelGGL(k_tsg_reduce1, dim3(nb, kRedGroups), dim3(256), 0, s->stream, s->partial, s->nchunks, s->m + 1,
                     s->part2);
  hipLaunchKernelGGL(k_tsg_reduce2, dim3(nb), dim3(256), 0, s->stream, s->part2, s->m + 1, out);
  if (W2) {
    hipLaunchKernelGGL(k_tsg_reduce1, dim3(nb, kRedGroups), dim3(256), 0, s->stream, s->partial2, s->nchunks,
                       s->m + 1, s->part2);
    hipLaunchKernelGGL(k_tsg_reduce2, dim3(nb), dim3(256), 0, s->stream, s->part2, s->m + 1, out2);
  }
}

// the step's two in-place updates (V_j, U_{j+1}) in one sweep
void tsg_update2(BlockSolver *s, int j) {
  const int wpb = s->shape.waves_update;
  const size_t tiles = s->npad / 16;
  hipLaunchKernelGGL(k_tsg_update2, dim3((unsigned)((tiles + wpb - 1) / wpb)), dim3(64 * wpb), 0, s->stream, s->npad,
                     j, s->V, s->npad * kS, s->cV, s->cU);
}

void tsg_update(BlockSolver *s, int nb, const double *V, const double *H, double alpha, double beta, double *Out) {
  const int wpb = s->shape.waves_update;
  const size_t tiles = s->npad / 16;
  hipLaunchKernelGGL(k_tsg_update, dim3((unsigned)((tiles + wpb - 1) / wpb)), dim3(64 * wpb), 0, s->stream, s->npad,
                     nb, V, s->npad * kS, H, alpha, beta, Out);
}

// W = Q Rfinal with CholQR2 (two Cholesky-QR passes); Q must not alias W or
// s->T; Rfinal = R2 R1 goes to the tile Rout.
void cholqr2(BlockSolver *s, const double *W, double *Q, double *Rout) {
  tsg_dots(s, 1, W, W, s->G);
  hipLaunchKernelGGL(k_chol16, dim3(1), dim3(64), 0, s->stream, s->G, s->Rt, s->Ri, s->flag);
  tsg_update(s, 1, W, s->Ri, 1.0, 0.0, s->T);
  tsg_dots(s, 1, s->T, s->T, s->G);
  hipLaunchKernelGGL(k_chol16, dim3(1), dim3(64), 0, s->stream, s->G, s->Rt2, s->Ri2, s->flag);
  tsg_update(s, 1, s->T, s->Ri2, 1.0, 0.0, Q);
  hipLaunchKernelGGL(k_mm16, dim3(1), dim3(256), 0, s->stream, s->Rt2, s->Rt, Rout);
}

// W = H1 P over the solver's chunks, in the shape the chunks were built for
template <class T, class TP>
static void bh1_launch(aerofx_ctx *c, BlockSolver *s, const T *D, const T *B, const TP *P, double *W) {
  // the narrow shape pays at FP32 (6 workgroups per CU); at FP64 its
  // register count leaves 4 waves per SIMD and the wide shape is faster
  if constexpr (sizeof(T) == sizeof(float)) {
    if (s->bsmall) {
      hipLaunchKernelGGL((k_bh1_apply<T, TP, kBEntSmall, kBRowsSmall, 2>), dim3(s->nbchunks[1]), dim3(256), 0,
                         s->stream, c->tab, s->nbchunks[1], s->bchunk[1], D, B, P, W);
      return;
    }
  }
  hipLaunchKernelGGL((k_bh1_apply<T, TP, kBEnt, kBRows, 4>), dim3(s->nbchunks[0]), dim3(256), 0, s->stream, c->tab,
                     s->nbchunks[0], s->bchunk[0], D, B, P, W);
}

// Z = P V_j, W = H1 Z: FP64 operators on FP64 vectors, or FP32 operators on
// an FP32 Z (the products are formed in FP64)
void apply_ops64(aerofx_ctx *c, BlockSolver *s, const double *Vj, double *W) {
  const int nb = (int)((c->N * (size_t)kS + 255) / 256);
  hipLaunchKernelGGL((k_bjacobi_apply<double, double>), dim3(nb), dim3(256), 0, s->stream, c->N,
                     (const double *)c->d_jacDi, Vj, s->Z);
  bh1_launch(c, s, (const double *)c->d_jacD, (const double *)c->d_jacB, (const double *)s->Z, W);
}

void apply_ops32(aerofx_ctx *c, BlockSolver *s, const double *Vj, double *W) {
  const int nb = (int)((c->N * (size_t)kS + 255) / 256);
  hipLaunchKernelGGL((k_bjacobi_apply<float, float>), dim3(nb), dim3(256), 0, s->stream, c->N, s->Di32, Vj, s->Zf);
  bh1_launch(c, s, (const float *)s->D32, (const float *)s->B32, (const float *)s->Zf, W);
}

// one inner cycle: from s->R (the residual block), builds V, Hbar, S0
void enqueue_cycle(aerofx_ctx *c, BlockSolver *s, int prec) {
  const size_t blk = s->npad * kS;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s->stream, s->flag, 1, nullptr);
  // equilibrate the residual columns first: they converge at different
  // rates, and Cholesky-QR squares the condition number of the block
  double *cn = s->nrm + (size_t)kChunks * kS + kS;
  hipLaunchKernelGGL(k_colnorm_partial, dim3(kChunks), dim3(256), 0, s->stream, s->npad, s->R, s->nrm);
  hipLaunchKernelGGL(k_colnorm_reduce, dim3(1), dim3(64), 0, s->stream, s->nrm, cn);
  hipLaunchKernelGGL(k_scale_cols, g1(blk), dim3(256), 0, s->stream, blk, s->R, cn, s->Z);
  cholqr2(s, s->Z, s->V, s->S0);
  hipLaunchKernelGGL(k_scale_tile_cols, dim3(1), dim3(256), 0, s->stream, s->S0, cn);
  // block Arnoldi with delayed reorthogonalisation: per step one operator
  // apply into slot j+1, one sweep for the Gram tiles of slots 0..j+1
  // against U_j and W, the tile algebra, and one sweep for the two in-place
  // updates (V_j, then U_{j+1})
  for (int j = 0; j < s->m; ++j) {
    double *Uj = s->V + (size_t)j * blk, *Wn = s->V + (size_t)(j + 1) * blk;
    if (prec)
      apply_ops32(c, s, Uj, Wn);
    else
      apply_ops64(c, s, Uj, Wn);
    tsg_dots(s, j + 2, s->V, Uj, s->G, Wn, s->G2);
    hipLaunchKernelGGL((k_bdcgs_tiles<false>), dim3(1), dim3(256), 0, s->stream, s->G, s->G2, s->Hbar, s->m, j, s->Cp,
                       s->Pp, s->S0, s->cV, s->cU, s->wk, s->flag);
    tsg_update2(s, j);
  }
  // the last block's reorthogonalisation completes column m-1
  tsg_dots(s, s->m + 1, s->V, s->V + (size_t)s->m * blk, s->G);
  hipLaunchKernelGGL((k_bdcgs_tiles<true>), dim3(1), dim3(256), 0, s->stream, s->G, s->G2, s->Hbar, s->m, s->m, s->Cp,
                     s->Pp, s->S0, s->cV, s->cU, s->wk, s->flag);
}

// host least squares: min || E1 S0 - Hbar Y ||_F, Hbar ((m+1)16 x m16) block
// upper Hessenberg with upper-triangular subdiagonal blocks.  Givens by column.
void host_lsq(int m, const std::vector<double> &Ht, const std::vector<double> &S0, std::vector<double> &Y,
              std::vector<double> &resid) {
  const int rows = (m + 1) * kS, cols = m * kS;
  std::vector<double> H((size_t)rows * cols, 0.0), rhs((size_t)rows * kS, 0.0);
  for (int bi = 0; bi <= m; ++bi)
    for (int bj = 0; bj < m; ++bj) {
      if (bi > bj + 1) continue;
      const double *t = Ht.data() + ((size_t)bi * m + bj) * kTile;
      for (int a = 0; a < kS; ++a)
        for (int b = 0; b < kS; ++b) H[(size_t)(bi * kS + a) * cols + bj * kS + b] = t[a * kS + b];
    }
  for (int a = 0; a < kS; ++a)
    for (int b = 0; b < kS; ++b) rhs[(size_t)a * kS + b] = S0[a * kS + b];
  for (int col = 0; col < cols; ++col) {
    const int last = std::min(rows - 1, col + kS);
    for (int r = last; r > col; --r) {
      double a = H[(size_t)(r - 1) * cols + col], b = H[(size_t)r * cols + col];
      if (b == 0.0) continue;
      double rr = std::hypot(a, b), cs = a / rr, sn = b / rr;
      for (int k = col; k < cols; ++k) {
        double x = H[(size_t)(r - 1) * cols + k], y = H[(size_t)r * cols + k];
        H[(size_t)(r - 1) * cols + k] = cs * x + sn * y;
        H[(size_t)r * cols + k] = -sn * x + cs * y;
      }
      for (int k = 0; k < kS; ++k) {
        double x = rhs[(size_t)(r - 1) * kS + k], y = rhs[(size_t)r * kS + k];
        rhs[(size_t)(r - 1) * kS + k] = cs * x + sn * y;
        rhs[(size_t)r * kS + k] = -sn * x + cs * y;
      }
    }
  }
  Y.assign((size_t)cols * kS, 0.0);
  for (int k = 0; k < kS; ++k)
    for (int i = cols - 1; i >= 0; --i) {
      double v = rhs[(size_t)i * kS + k];
      for (int j = i + 1; j < cols; ++j) v -= H[(size_t)i * cols + j] * Y[(size_t)j * kS + k];
      const double d = H[(size_t)i * cols + i];
      Y[(size_t)i * kS + k] = d != 0.0 ? v / d : 0.0;
    }
  resid.assign(kS, 0.0);
  for (int k = 0; k < kS; ++k) {
    double s = 0.0;
    for (int i = cols; i < rows; ++i) s += rhs[(size_t)i * kS + k] * rhs[(size_t)i * kS + k];
    resid[k] = std::sqrt(s);
  }
}

// times each candidate launch shape on the current buffers; keeps the fastest
// Times the candidate waves per workgroup of the two sweeps a step runs (the
// two-right-hand-block Gram sweep and the fused update) at a mid-cycle
// width; each candidate is timed twice over three launches and its best is
// kept, so one noisy measurement does not decide.
void autotune(aerofx_ctx *c, BlockSolver *s) {
  if (s->shape.tuned) return;
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  const int j = std::min(s->m, 8) - 1;  // step whose sweeps are timed
  const size_t blk = s->npad * kS;
  double *U = s->V + (size_t)j * blk, *W = s->V + (size_t)(j + 1) * blk;
  // defined coefficients for the update's timing runs (the slots are
  // rewritten by every cycle before they are read)
  (void)hipMemsetAsync(s->cV, 0, (size_t)(kMaxM + 2) * kTile * sizeof(double), s->stream);
  (void)hipMemsetAsync(s->cU, 0, (size_t)(kMaxM + 2) * kTile * sizeof(double), s->stream);
  auto time3 = [&](auto &&launch) {
    launch();  // warm
    (void)hipEventRecord(a, s->stream);
    for (int r = 0; r < 3; ++r) launch();
    (void)hipEventRecord(b, s->stream);
    (void)hipEventSynchronize(b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
  };
  const int cand[] = {1, 2, 4};  // waves per workgroup (<= 256 threads)
  float best[4] = {1e30f, 1e30f, 1e30f, 1e30f};
  for (int pass = 0; pass < 2; ++pass)
    for (int k = 0; k < 3; ++k) {
      s->shape.waves_dots = cand[k];
      best[k] = std::min(best[k], time3([&] { tsg_dots(s, j + 2, s->V, U, s->G, W, s->G2); }));
    }
  int kb = 0;
  for (int k = 1; k < 3; ++k)
    if (best[k] < best[kb]) kb = k;
  c->blk_tune[0] = s->shape.waves_dots = cand[kb];
  const int cand_u[] = {1, 2, 4, 8};
  for (float &x : best) x = 1e30f;
  for (int pass = 0; pass < 2; ++pass)
    for (int k = 0; k < 4; ++k) {
      s->shape.waves_update = cand_u[k];
      best[k] = std::min(best[k], time3([&] { tsg_update2(s, j); }));
    }
  kb = 0;
  for (int k = 1; k < 4; ++k)
    if (best[k] < best[kb]) kb = k;
  c->blk_tune[1] = s->shape.waves_update = cand_u[kb];
  s->shape.tuned = true;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
}

BlockSolver *get_block(aerofx_ctx *c, int m, int *rc) {
  *rc = 0;
  if (c->blk && c->blk->m == m) return c->blk;
  free_block(c);
  BlockSolver *s = new BlockSolver();
  s->m = m;
  s->n = 5 * (size_t)c->N;
  s->npad = (s->n + 63) / 64 * 64;  // also a multiple of 16 rows: whole MFMA row groups
  // about 512 rows per wave of V^T W, between 64 and 8192 chunks
  s->nchunks = (int)std::min<size_t>(8192, std::max<size_t>(64, s->npad / 512));
  const size_t blk = s->npad * kS;
  int r = 0;
  if ((r = alloc_d(c, (size_t)(m + 1) * blk, &s->V)) || (r = alloc_d(c, blk, &s->W)) || (r = alloc_d(c, blk, &s->Z)) ||
      (r = alloc_d(c, blk, &s->R)) || (r = alloc_d(c, blk, &s->X)) || (r = alloc_d(c, blk, &s->B)) ||
      (r = alloc_d(c, blk, &s->T)) || (r = alloc_d(c, (size_t)s->nchunks * (m + 1) * kTile, &s->partial)) ||
      (r = alloc_d(c, (size_t)s->nchunks * (m + 1) * kTile, &s->partial2)) ||
      (r = alloc_d(c, (size_t)kRedGroups * (m + 1) * kTile, &s->part2)) ||
      (r = alloc_d(c, (size_t)(m + 1) * m * kTile, &s->Hbar)) || (r = alloc_d(c, kTile, &s->S0)) ||
      (r = alloc_d(c, (size_t)(kMaxM + 1) * kTile, &s->G)) || (r = alloc_d(c, (size_t)(kMaxM + 1) * kTile, &s->G2)) ||
      (r = alloc_d(c, kTile, &s->Rt)) || (r = alloc_d(c, kTile, &s->Ri)) || (r = alloc_d(c, kTile, &s->Rt2)) ||
      (r = alloc_d(c, kTile, &s->Ri2)) || (r = alloc_d(c, (size_t)m * kTile, &s->Y)) ||
      (r = alloc_d(c, (size_t)(kMaxM + 1) * kTile, &s->Cp)) || (r = alloc_d(c, kTile, &s->Pp)) ||
      (r = alloc_d(c, (size_t)(kMaxM + 2) * kTile, &s->cV)) || (r = alloc_d(c, (size_t)(kMaxM + 2) * kTile, &s->cU)) ||
      (r = alloc_d(c, (size_t)kWkTiles * kTile, &s->wk)) ||
      (r = alloc_d(c, (size_t)kChunks * kS + 2 * kS, &s->nrm))) {
    delete s;
    *rc = r;
    return nullptr;
  }
  void *v = nullptr;
  if ((r = ctx_alloc(c, sizeof(int) * 4, &v))) {
    delete s;
    *rc = r;
    return nullptr;
  }
  s->flag = (int *)v;
  {
    // row chunks for the block H1 apply, from the adjacency row pointers
    std::vector<int> rp(c->N + 1);
    hipError_t e0 = hipMemcpy(rp.data(), c->tab.rowptr, sizeof(int) * rp.size(), hipMemcpyDeviceToHost);
    if (e0 != hipSuccess) {
      delete s;
      *rc = set_hip_error(e0, "block solver row pointers");
      return nullptr;
    }
    int maxdeg = 0;
    for (int n = 0; n < c->N; ++n) maxdeg = std::max(maxdeg, rp[n + 1] - rp[n]);
    if (maxdeg > kBEnt) {
      delete s;
      *rc = set_error(AEROFX_ERR_MESH, "block_gmres: a node has more than 64 neighbours");
      return nullptr;
    }
    s->bsmall = maxdeg <= kBEntSmall;
    for (int shape = 0; shape < (s->bsmall ? 2 : 1); ++shape) {
      const int lim_ent = shape ? kBEntSmall : kBEnt, lim_rows = shape ? kBRowsSmall : kBRows;
      std::vector<int> ch(1, 0);
      int acc = 0, rows = 0;
      for (int n = 0; n < c->N; ++n) {
        const int cnt = rp[n + 1] - rp[n];
        if (acc + cnt > lim_ent || rows == lim_rows) {
          ch.push_back(n);
          acc = 0;
          rows = 0;
        }
        acc += cnt;
        rows += 1;
      }
      ch.push_back(c->N);
      s->nbchunks[shape] = (int)ch.size() - 1;
      if ((r = ctx_alloc(c, sizeof(int) * ch.size(), &v))) {
        delete s;
        *rc = r;
        return nullptr;
      }
      s->bchunk[shape] = (int *)v;
      e0 = hipMemcpy(s->bchunk[shape], ch.data(), sizeof(int) * ch.size(), hipMemcpyHostToDevice);
      if (e0 != hipSuccess) {
        delete s;
        *rc = set_hip_error(e0, "block solver chunks");
        return nullptr;
      }
    }
  }
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void **)&s->h_nrm, sizeof(double) * kS, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void **)&s->h_flag, sizeof(int) * 4, hipHostMallocDefault);
  if (e != hipSuccess) {
    delete s;
    *rc = set_hip_error(e, "block solver set-up");
    return nullptr;
  }
  // zero padded rows once: every block vector keeps zero rows beyond n
  (void)hipMemsetAsync(s->V, 0, (size_t)(m + 1) * blk * sizeof(double), s->stream);
  double *vecs[6] = {s->W, s->Z, s->R, s->X, s->B, s->T};
  for (double *p : vecs) (void)hipMemsetAsync(p, 0, blk * sizeof(double), s->stream);
  (void)hipMemsetAsync(s->Hbar, 0, (size_t)(m + 1) * m * kTile * sizeof(double), s->stream);
  if (c->blk_tune[0] > 0) {
    s->shape.waves_dots = c->blk_tune[0];
    s->shape.waves_update = c->blk_tune[1];
    s->shape.tuned = true;
  }
  c->blk = s;
  return s;
}

int ensure_f32(aerofx_ctx *c, BlockSolver *s) {
  if (s->D32 && s->f32_version == c->jac_version) return 0;
  const size_t nd = blk64_count(c->N), nbk = blk64_count(2 * (size_t)c->E);
  void *v = nullptr;
  int rc;
  if (!s->D32) {
    if ((rc = ctx_alloc(c, nd * sizeof(float), &v))) return rc;
    s->D32 = (float *)v;
    if ((rc = ctx_alloc(c, nbk * sizeof(float), &v))) return rc;
    s->B32 = (float *)v;
    if ((rc = ctx_alloc(c, nd * sizeof(float), &v))) return rc;
    s->Di32 = (float *)v;
    if ((rc = ctx_alloc(c, s->npad * kS * sizeof(float), &v))) return rc;
    s->Zf = (float *)v;
  }
  hipLaunchKernelGGL(k_f64_to_f32, g1(nd), dim3(256), 0, s->stream, nd, (const double *)c->d_jacD, s->D32);
  hipLaunchKernelGGL(k_f64_to_f32, g1(nbk), dim3(256), 0, s->stream, nbk, (const double *)c->d_jacB, s->B32);
  // the FP32 block-Jacobi inverse is LINPACK in float on the FP32 blocks,
  // as the single-RHS mixed path builds it
  if ((rc = launch_invert_f32(c, s->stream, s->D32, s->Di32))) return rc;
  hipError_t e = hipMemcpyAsync(c->h_err + 1, c->d_err + 1, sizeof(int), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return set_hip_error(e, "block_gmres FP32 inverse");
  if (c->h_err[1]) return set_error(AEROFX_ERR_STATE, "block_gmres: zero pivot in an FP32 diagonal block");
  s->f32_version = c->jac_version;
  return 0;
}

}  // namespace

void free_block(aerofx_ctx *c) {
  BlockSolver *s = c->blk;
  if (!s) return;
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  if (s->h_nrm) (void)hipHostFree(s->h_nrm);
  if (s->h_flag) (void)hipHostFree(s->h_flag);
  delete s;
  c->blk = nullptr;
}

}  // namespace aerofx

using namespace aerofx;

extern "C" int aerofx_block_gmres(aerofx_ctx *c, int nrhs, const double *Bv, double *Xv, int m, double eps,
                                  int max_cycles, int precision, int use_graph, aerofx_block_stats *st) {
  if (!c || !Bv || !Xv || !st) return set_error(AEROFX_ERR_ARG, "block_gmres: null argument");
  if (nrhs < 1 || nrhs > kS) return set_error(AEROFX_ERR_ARG, "block_gmres: 1 <= nrhs <= 16");
  if (m < 1 || m > kMaxM) return set_error(AEROFX_ERR_ARG, "block_gmres: 1 <= m <= 16");
  if (precision != 0 && precision != 1) return set_error(AEROFX_ERR_ARG, "block_gmres: precision 0 or 1");
  if (!c->d_jacD || c->jac_precision != 0)
    return set_error(AEROFX_ERR_STATE, "block_gmres: needs FP64 operators (aerofx_jacobian_setup precision 0)");
  auto t0 = std::chrono::steady_clock::now();
  int rc = 0;
  BlockSolver *s = get_block(c, m, &rc);
  if (!s) return rc;
  *st = aerofx_block_stats{};
  // order after the caller's work on the context stream; the event is
  // destroyed on every return
  struct EventGuard {
    hipEvent_t ev = nullptr;
    ~EventGuard() {
      if (ev) (void)hipEventDestroy(ev);
    }
  } guard;
  hipError_t e = hipEventCreateWithFlags(&guard.ev, hipEventDisableTiming);
  hipEvent_t ev = guard.ev;
  if (e == hipSuccess) e = hipEventRecord(ev, c->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(s->stream, ev, 0);
  if (e != hipSuccess) return set_hip_error(e, "block_gmres stream order");
  if (precision == 1 && (rc = ensure_f32(c, s))) return rc;
  const size_t blk = s->npad * kS;
  hipLaunchKernelGGL(k_to_block, g1(blk), dim3(256), 0, s->stream, s->n, s->npad, nrhs, Bv, s->B, 0x5eedu);
  (void)hipMemsetAsync(s->X, 0, blk * sizeof(double), s->stream);
  autotune(c, s);
  // ||b_k||
  hipLaunchKernelGGL(k_colnorm_partial, dim3(kChunks), dim3(256), 0, s->stream, s->npad, s->B, s->nrm);
  hipLaunchKernelGGL(k_colnorm_reduce, dim3(1), dim3(64), 0, s->stream, s->nrm, s->nrm + (size_t)kChunks * kS);
  e = hipMemcpyAsync(s->h_nrm, s->nrm + (size_t)kChunks * kS, sizeof(double) * kS, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return set_hip_error(e, "block_gmres norms");
  std::vector<double> bnorm(kS);
  for (int k = 0; k < kS; ++k) bnorm[k] = std::sqrt(s->h_nrm[k]);
  const int nbk = (int)((c->N * (size_t)kS + 255) / 256);
  double worst = 0.0;
  int cycle = 0;
  for (;; ++cycle) {
    // R = B - A64 X (FP64 refinement residual)
    bh1_launch(c, s, (const double *)c->d_jacD, (const double *)c->d_jacB, (const double *)s->X, s->R);
    hipLaunchKernelGGL(k_bsub_from, g1(blk), dim3(256), 0, s->stream, blk, s->B, s->R);
    hipLaunchKernelGGL(k_colnorm_partial, dim3(kChunks), dim3(256), 0, s->stream, s->npad, s->R, s->nrm);
    hipLaunchKernelGGL(k_colnorm_reduce, dim3(1), dim3(64), 0, s->stream, s->nrm, s->nrm + (size_t)kChunks * kS);
    e = hipMemcpyAsync(s->h_nrm, s->nrm + (size_t)kChunks * kS, sizeof(double) * kS, hipMemcpyDeviceToHost,
                       s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return set_hip_error(e, "block_gmres residual");
    worst = 0.0;
    for (int k = 0; k < nrhs; ++k) {
      const double rel = bnorm[k] > 0.0 ? std::sqrt(s->h_nrm[k]) / bnorm[k] : 0.0;
      worst = std::max(worst, rel);
    }
    if (worst <= eps) {
      st->converged = 1;
      break;
    }
    if (cycle >= max_cycles) break;
    // inner cycle: graph replay or eager launches
    if (use_graph) {
      if (!s->exec || s->graph_prec != precision) {
        if (s->exec) (void)hipGraphExecDestroy(s->exec);
        if (s->graph) (void)hipGraphDestroy(s->graph);
        s->exec = nullptr;
        s->graph = nullptr;
        e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return set_hip_error(e, "block_gmres capture");
        enqueue_cycle(c, s, precision);
        e = hipStreamEndCapture(s->stream, &s->graph);
        if (e == hipSuccess) e = hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) return set_hip_error(e, "block_gmres graph");
        s->graph_prec = precision;
      }
      e = hipGraphLaunch(s->exec, s->stream);
      if (e != hipSuccess) return set_hip_error(e, "block_gmres graph launch");
    } else {
      enqueue_cycle(c, s, precision);
    }
    // small least-squares problem on the host
    s->hH.resize((size_t)(m + 1) * m * kTile);
    s->hS0.resize(kTile);
    e = hipMemcpyAsync(s->hH.data(), s->Hbar, s->hH.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s->hS0.data(), s->S0, kTile * sizeof(double), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->h_flag, s->flag, sizeof(int), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return set_hip_error(e, "block_gmres Hessenberg copy");
    if (s->h_flag[0]) return set_error(AEROFX_ERR_STATE, "block_gmres: rank-deficient block (CholQR breakdown)");
    std::vector<double> resid;
    host_lsq(m, s->hH, s->hS0, s->hY, resid);
    e = hipMemcpyAsync(s->Y, s->hY.data(), s->hY.size() * sizeof(double), hipMemcpyHostToDevice, s->stream);
    if (e != hipSuccess) return set_hip_error(e, "block_gmres Y copy");
    // X += P (V Y)
    tsg_update(s, m, s->V, s->Y, 1.0, 0.0, s->W);
    if (precision)
      hipLaunchKernelGGL((k_bjacobi_apply<float, double>), dim3(nbk), dim3(256), 0, s->stream, c->N, s->Di32, s->W,
                         s->Z);
    else
      hipLaunchKernelGGL((k_bjacobi_apply<double, double>), dim3(nbk), dim3(256), 0, s->stream, c->N,
                         (const double *)c->d_jacDi, s->W, s->Z);
    hipLaunchKernelGGL(k_badd, g1(blk), dim3(256), 0, s->stream, blk, s->Z, s->X);
    st->inner_its += m;
  }
  hipLaunchKernelGGL(k_from_block, g1(s->n * (size_t)nrhs), dim3(256), 0, s->stream, s->n, nrhs, s->X, Xv);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipEventRecord(ev, s->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, ev, 0);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return set_hip_error(e, "block_gmres finish");
  st->cycles = cycle;
  st->res_max = worst;
  st->tune_waves_dots = s->shape.waves_dots;
  st->tune_waves_update = s->shape.waves_update;
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return AEROFX_OK;
}
