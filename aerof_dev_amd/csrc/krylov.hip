// krylov.hip -- GMRES and the pseudo-transient Newton driver.
//
// GmresSolver::solve (KspSolver.C:353-492): right-preconditioned GMRES(m)
// with Givens rotations and back-substitution (KspSolver.C:919-982).  The
// whole cycle runs on the GPU as a fixed sequence of kernels with no host
// synchronisation: the Hessenberg matrix, rotations, residual estimate and
// the stop flag live in device memory, a one-lane kernel applies each
// rotation, and every vector kernel of a later Arnoldi step returns at once
// when the flag is set.  The host synchronises once per cycle.
//
// Orthogonalisation is classical Gram-Schmidt with one reorthogonalisation
// pass (CGS2) in FP64, the reorthogonalisation of each new vector delayed to
// the next Arnoldi step (DCGS2) so that a step reads the basis twice and
// reduces once: one sweep takes V^T [u_j, w] and the norms, one update pass
// finishes v_j and forms the next once-projected vector u_{j+1} (see
// k_dcgs_reduce for the algebra), instead of the reference's j+1 dependent
// dot/axpy pairs (modified Gram-Schmidt, KspSolver.C:404-411).  Reductions
// are deterministic: a fixed grid of blocks sums contiguous slices (wave
// butterflies in a fixed pattern), and one block sums the partials in order.
//
// The Newton driver follows TsSolver/NewtonSolver for a steady problem with
// one Newton iteration per pseudo-time step (NewtonSolver.h:70-231,
// ImplicitTsDesc.C:225-241): F = R(U) (+ (U - Un)/dt = 0 at the first
// iteration), H1 + diag(1/dt), block-Jacobi, GMRES, U += dQ; the CFL number
// follows CflLaw "Old" (TsParameters.C:198-200) and the run stops when
// ||R||_2 <= eps ||R_0||_2 (TsDesc.C:1261-1278).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "context.h"
#include "host_internal.h"
#include "host_mesh.h"

namespace aerofx {

constexpr int kDotBlocks = 1024;  // fixed: the reduction order never depends on the device
constexpr int kSweepBlocks = 256;  // blocks of the basis sweeps: one per CU
constexpr int kMaxBasis = 32;     // basis vectors handled by one sweep (m <= 31)
static_assert(kMaxBasis == kMaxBasisCoef, "coefficient layout shared with the update kernel");
constexpr int kWave = 64;

// Device-resident GMRES state (one per right-hand side).
// Coefficients of the Gram-Schmidt update kernel, in GmresState::coef:
// a at [0, 32), d at [32, 65), 1/tau at [65], 1/(tau rho) at [66].
constexpr int kCoefD = kMaxBasis, kCoefInvTau = 2 * kMaxBasis + 1, kCoefInvTauRho = 2 * kMaxBasis + 2;
struct GmresState {
  double H[(kMaxBasis + 1) * kMaxBasis];   // rotated columns (R), row-major (i, j) -> H[i * kMaxBasis + j]
  double Hh[(kMaxBasis + 1) * kMaxBasis];  // the Hessenberg columns themselves, same layout
  double cs[2 * kMaxBasis];                // (sin, cos) per rotation
  double g[kMaxBasis + 1];
  double y[kMaxBasis];
  double cpart[kMaxBasis + 1];             // once-projected coefficients of the newest column
  double coef[2 * kMaxBasis + 4];
  double beta, target, l2res, invhn, rho;
  int iter, jlast, done, maxits, m;
};

// ---------------------------------------------------------------------------
// deterministic block reductions
// ---------------------------------------------------------------------------
// K = kMaxBasis values per lane: a transposing butterfly halves the values a
// lane carries at each of the first five exchange steps (a lane keeps the
// half selected by its lane bit and adds the partner's copy of that half),
// so after steps 32..2 lane l holds one index, summed over the lanes that
// share bit 0, and the step 1 exchange completes the wave sum: 32 shuffles
// instead of 6 per value.  The fixed pattern keeps the sums deterministic.
template <int HALF, int OFF>
__device__ __forceinline__ void halve_step(double (&v)[kMaxBasis], int lane) {
  const bool up = (lane & OFF) != 0;
#pragma unroll
  for (int i = 0; i < HALF; ++i) {
    const double keep = up ? v[HALF + i] : v[i];
    const double give = up ? v[i] : v[HALF + i];
    v[i] = keep + __shfl_xor(give, OFF, kWave);
  }
}

__device__ __forceinline__ void block_reduce_store_k(double (&v)[kMaxBasis], int nk, int nparts,
                                                     double *__restrict__ partial) {
  static_assert(kMaxBasis == 32 && kWave == 64, "the butterfly below is written for 32 values, 64 lanes");
  __shared__ double sh[kBlock / kWave][kMaxBasis];
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  halve_step<16, 32>(v, lane);
  halve_step<8, 16>(v, lane);
  halve_step<4, 8>(v, lane);
  halve_step<2, 4>(v, lane);
  halve_step<1, 2>(v, lane);
  const double tot = v[0] + __shfl_xor(v[0], 1, kWave);
  // index held by this lane: bits 5..1 of the lane select halves 16, 8, 4, 2, 1
  const int idx = ((lane >> 5) & 1) * 16 + ((lane >> 4) & 1) * 8 + ((lane >> 3) & 1) * 4 + ((lane >> 2) & 1) * 2 +
                  ((lane >> 1) & 1);
  if ((lane & 1) == 0) sh[wave][idx] = tot;
  __syncthreads();
  if ((int)threadIdx.x < nk) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) s += sh[w][threadIdx.x];
    partial[(size_t)threadIdx.x * kDotBlocks + blockIdx.x] = s;
  }
  (void)nparts;
}

// sum over lanes 0..31 (the values of lanes 32..63 are summed separately),
// fixed butterfly; every lane of the half-wave gets the total
__device__ __forceinline__ double wsum32(double v) {
#pragma unroll
  for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 32);
  return v;
}

__device__ __forceinline__ void block_reduce_store_1(double v, double *__restrict__ partial) {
  __shared__ double sh1[kBlock / kWave];
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  if (lane == 0) sh1[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) s += sh1[w];
    partial[blockIdx.x] = s;
  }
}

// Givens update of Hessenberg column jc (KspSolver.C:404-447, 919-962):
// rotations, residual estimate, stop flag.  The column (jc + 2 entries, the
// last one the subdiagonal) is in LDS; the block's threads stage the stored
// rotations and g[jc] in LDS in one parallel round, then thread 0 runs the
// rotation recurrence from LDS with unconditional loads.  Every thread of the
// block must call it (it synchronises).
struct GivensStage {
  double rot[2 * kMaxBasis];
  double gj;
  int done;
};

// Loads what the update reads into the stage (the caller may do this early,
// overlapping it with its own loads, and must synchronise before the update).
__device__ __forceinline__ void givens_stage(const GmresState *s, int jc, GivensStage &g) {
  const int t = threadIdx.x;
  if (t < 2 * kMaxBasis) g.rot[t] = s->cs[t];
  if (t == 0) g.gj = s->g[jc];
}

__device__ void givens_col(GmresState *s, int jc, const double *col, GivensStage &g, bool staged = false) {
  const int t = threadIdx.x;
  if (!staged) givens_stage(s, jc, g);
  if (t == 0) g.done = s->done;
  __syncthreads();
  if (t != 0 || g.done) return;
  const int M = kMaxBasis;
  double h[kMaxBasis + 1];
#pragma unroll
  for (int i = 0; i <= kMaxBasis; ++i) h[i] = i <= jc + 1 ? col[i] : 0.0;
  const double hn = col[jc + 1];
#pragma unroll
  for (int i = 0; i < kMaxBasis - 1; ++i) {
    const double sn = g.rot[2 * i], cs = g.rot[2 * i + 1];
    const double hj = cs * h[i] + sn * h[i + 1];
    const double hjp1 = -sn * h[i] + cs * h[i + 1];
    if (i < jc) {
      h[i] = hj;
      h[i + 1] = hjp1;
    }
  }
  double hjj = 0.0, hj1 = 0.0;
#pragma unroll
  for (int i = 0; i < kMaxBasis; ++i)
    if (i == jc) {
      hjj = h[i];
      hj1 = h[i + 1];
    }
  double sn, cs;
  if (hjj * hjj == 0.0) {
    cs = 0.0;
    sn = 1.0;
  } else {
    double coef = 1.0 / sqrt(hjj * hjj + hj1 * hj1);
    cs = sqrt(hjj * hjj) * coef;
    sn = cs * hj1 / hjj;
  }
  s->cs[2 * jc + 1] = cs;
  s->cs[2 * jc] = sn;
#pragma unroll
  for (int i = 0; i < kMaxBasis; ++i)
    if (i < jc) s->H[i * M + jc] = h[i];
  s->H[jc * M + jc] = cs * hjj + sn * hj1;
  s->H[(jc + 1) * M + jc] = 0.0;
  const double gj = g.gj;
  const double gj1 = -sn * gj;
  s->g[jc + 1] = gj1;
  s->g[jc] = cs * gj;
  s->iter += 1;
  s->jlast = jc;
  if (hn == 0.0) {
    s->done = 1;
    return;
  }
  s->l2res = fabs(gj1);
  if (s->l2res <= s->target || s->iter >= s->maxits) s->done = 1;
}

// Row pairs of a block's slice: the grid has nparts blocks (fixed by n, at
// most one per CU); block b covers a contiguous slice of even length, a lane
// takes two consecutive rows per step (16-B loads) and handles them in row
// order; an odd last row goes to the lane whose next pair would start there.
struct RowPairs {
  size_t b0, b1, even_end;
  __device__ RowPairs(size_t n) {
    const size_t slice = ((n + gridDim.x - 1) / gridDim.x + 1) & ~(size_t)1;
    b0 = min(n, blockIdx.x * slice);  // rounding the slice up can leave the last blocks empty
    b1 = min(n, b0 + slice);
    even_end = b0 + ((b1 - b0) & ~(size_t)1);
  }
  __device__ bool has_tail() const { return even_end < b1 && threadIdx.x == ((b1 - 1 - b0) / 2) % kBlock; }
};

// partial[b] = sum over block b's slice of w_i^2
__global__ __launch_bounds__(kBlock) void k_sumsq(size_t n, const double *__restrict__ w,
                                                  double *__restrict__ partial, const int *__restrict__ done) {
  if (done && *done) return;
  const RowPairs rp(n);
  double acc = 0.0;
  for (size_t i = rp.b0 + 2 * threadIdx.x; i < rp.even_end; i += 2 * kBlock) {
    const double2 w2 = *reinterpret_cast<const double2 *>(w + i);
    acc += w2.x * w2.x;
    acc += w2.y * w2.y;
  }
  if (rp.has_tail()) acc += w[rp.b1 - 1] * w[rp.b1 - 1];
  block_reduce_store_1(acc, partial);
}

// Gram-Schmidt with delayed reorthogonalisation (DCGS2: classical
// Gram-Schmidt twice, the second pass for a vector folded into the next
// Arnoldi step).  At step j the basis slots hold v_0 .. v_{j-1} (final,
// orthonormal) and u_j = (A' v_{j-1} - V c) / rho, projected once (c, rho
// from step j-1; u_0 = r / ||r||), and w = A' u_j, A' = H1 P.  One sweep
// over the slots gives, per block,
//   a = V_{j-1}^T u_j,  b = V_{j-1}^T w,  alpha = u.u,  beta = u.w,  gamma = w.w
// (partial rows 0.., 32.., 64, 65, 66).  The final sweep of a cycle (FINAL)
// needs a and alpha only.
template <bool FINAL>
__global__ __launch_bounds__(kBlock) void k_dots2(size_t n, const double *__restrict__ V, size_t ldv, int j,
                                                  const double *__restrict__ w, double *__restrict__ partial,
                                                  const int *__restrict__ done) {
  if (*done) return;
  const RowPairs rp(n);
  const double *u = V + (size_t)j * ldv;
  double aa[kMaxBasis], bb[kMaxBasis], al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
  for (int k = 0; k < kMaxBasis; ++k) aa[k] = bb[k] = 0.0;
  auto row = [&](double ui, double wi, const double *vk) {
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k) {
      aa[k] += vk[k] * ui;
      if (!FINAL) bb[k] += vk[k] * wi;
    }
    al += ui * ui;
    if (!FINAL) {
      be += ui * wi;
      ga += wi * wi;
    }
  };
  // software-pipelined: the next row pair's loads are in flight while the
  // current pair is added (one wave per SIMD: nothing else hides latency)
  auto load = [&](size_t i, double2 (&v)[kMaxBasis], double2 &u2, double2 &w2) {
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k)
      v[k] = k < j ? *reinterpret_cast<const double2 *>(V + (size_t)k * ldv + i) : make_double2(0.0, 0.0);
    u2 = *reinterpret_cast<const double2 *>(u + i);
    w2 = FINAL ? make_double2(0.0, 0.0) : *reinterpret_cast<const double2 *>(w + i);
  };
  size_t i = rp.b0 + 2 * threadIdx.x;
  double2 v[kMaxBasis], u2, w2;
  if (i < rp.even_end) load(i, v, u2, w2);
  while (i < rp.even_end) {
    const size_t inext = i + 2 * kBlock;
    double2 vn[kMaxBasis], un2, wn2;
    if (inext < rp.even_end) load(inext, vn, un2, wn2);
    double vk[kMaxBasis];
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k) vk[k] = v[k].x;
    row(u2.x, w2.x, vk);
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k) vk[k] = v[k].y;
    row(u2.y, w2.y, vk);
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k) v[k] = vn[k];
    u2 = un2;
    w2 = wn2;
    i = inext;
  }
  if (rp.has_tail()) {
    const size_t i = rp.b1 - 1;
    double vk[kMaxBasis];
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k) vk[k] = k < j ? V[(size_t)k * ldv + i] : 0.0;
    row(u[i], FINAL ? 0.0 : w[i], vk);
  }
  block_reduce_store_k(aa, j, gridDim.x, partial);
  __syncthreads();  // the reduction's LDS is reused
  if (!FINAL) {
    block_reduce_store_k(bb, j, gridDim.x, partial + (size_t)kMaxBasis * kDotBlocks);
    __syncthreads();
  }
  double sc[kMaxBasis];
#pragma unroll
  for (int k = 0; k < kMaxBasis; ++k) sc[k] = k == 0 ? al : k == 1 ? be : k == 2 ? ga : 0.0;
  block_reduce_store_k(sc, 3, gridDim.x, partial + (size_t)2 * kMaxBasis * kDotBlocks);
}

// out[k] = sum_b partial[k][b] in a fixed tree, one block per k
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double *__restrict__ partial, int nparts,
                                                            double *__restrict__ out, const int *__restrict__ done) {
  if (done && *done) return;
  __shared__ double sh[kBlock];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nparts; b += kBlock) s += partial[(size_t)k * kDotBlocks + b];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = kBlock / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = sh[0];
}

// After the sweep of step j (one block of kReduceThreads): sums the partials
// (16 lanes per value, fixed butterfly), then
//  1. tau = ||u_j - V a|| = sqrt(alpha - a.a): v_j = (u_j - V a) / tau.
//  2. Column j-1 is complete: A' v_{j-1} = V (c + rho a) + rho tau v_j; it goes
//     to Hh and through the Givens update (stop test).  At j = 0 instead the
//     right-hand side g_0 = ||r|| becomes ||r|| tau (v_0 = u_0 / tau).
//  3. Unless stopped (or FINAL): with g = Hbar_{j-1} a (A' V_{j-1} = V_j Hbar),
//     A' v_j = (w - V_j g) / tau, so h = V_j^T A' v_j = (p - g) / tau with
//     p = (b, (beta - a.b) / tau), and ||A' v_j||^2 = (gamma - 2 g.p + g.g) / tau^2.
//     The next vector u_{j+1} = (A' v_j - V_j h) / rho' = (w - V_j d) / (tau rho'),
//     d = g + tau h, rho'^2 = ||A' v_j||^2 - h.h (its exact norm is taken at
//     the next step, so rho' only scales).  If rho'^2 is not above 1e-28 of
//     ||A' v_j||^2 the Krylov space is exhausted: column j is final as it
//     is, and the cycle stops.
// coef (a, d, 1/tau, 1/(tau rho')) feeds k_dcgs_update.
constexpr int kReduceThreads = 1024;
template <bool FINAL>
__global__ __launch_bounds__(kReduceThreads) void k_dcgs_reduce(const double *__restrict__ partial, int nparts,
                                                                GmresState *st, int j) {
  if (st->done) return;
  __shared__ double sa[kMaxBasis], sb[kMaxBasis], ssc[3], col[kMaxBasis + 2], sg[kMaxBasis + 1];
  __shared__ double sH[(kMaxBasis + 1) * kMaxBasis];  // Hbar_{j-1}, rows 0..j
  __shared__ double stau;
  __shared__ int sbreak;
  __shared__ GivensStage gstage;
  const int t = threadIdx.x, grp = t / 16, sub = t % 16;
  // the Givens inputs of column j-1, loaded with the partials
  if (j >= 1) givens_stage(st, j - 1, gstage);
  // columns 0..j-2 of the Hessenberg matrix (final), loaded with the partials
  if (!FINAL)
    for (int e = t; e < (j + 1) * kMaxBasis; e += kReduceThreads)
      if (e % kMaxBasis + 1 < j) sH[e] = st->Hh[e];
  const int nval = FINAL ? j + 1 : 2 * j + 3;
  if (grp < nval) {
    int row;
    if (grp < j) row = grp;                                           // a
    else if (!FINAL && grp < 2 * j) row = kMaxBasis + (grp - j);      // b
    else row = 2 * kMaxBasis + (grp - (FINAL ? j : 2 * j));           // alpha, beta, gamma
    const double *pr = partial + (size_t)row * kDotBlocks;
    double s = 0.0;
    for (int b0 = sub; b0 < nparts; b0 += 16 * 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = b0 + 16 * u < nparts ? pr[b0 + 16 * u] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    s += __shfl_xor(s, 8, 16);
    s += __shfl_xor(s, 4, 16);
    s += __shfl_xor(s, 2, 16);
    s += __shfl_xor(s, 1, 16);
    if (sub == 0) {
      if (grp < j) sa[grp] = s;
      else if (!FINAL && grp < 2 * j) sb[grp - j] = s;
      else ssc[grp - (FINAL ? j : 2 * j)] = s;
    }
  }
  __syncthreads();
  if (t < 32) {
    // a.a over one half-wave, fixed butterfly
    const double asq = wsum32(t < j ? sa[t] * sa[t] : 0.0);
    if (t == 0) {
      const double r = ssc[0] - asq;
      stau = r > 0.0 ? sqrt(r) : 0.0;
    }
  }
  __syncthreads();
  const double tau = stau;
  if (j >= 1) {
    // column j-1, complete now
    const double rho = st->rho;
    if (t <= kMaxBasis + 1) {
      double v = 0.0;
      if (t < j) v = st->cpart[t] + rho * sa[t];
      if (t == j) v = rho * tau;
      col[t] = v;
      if (t <= j) {
        st->Hh[t * kMaxBasis + (j - 1)] = v;
        sH[t * kMaxBasis + (j - 1)] = v;
      }
    }
    __syncthreads();
    givens_col(st, j - 1, col, gstage, true);
    __syncthreads();
    if (FINAL || st->done) return;
  } else if (t == 0) {
    st->g[0] *= tau;
  }
  if (FINAL) return;
  // g = Hbar_{j-1} a (rows 0..j), in ascending column order
  if (t <= j) {
    double gt = 0.0;
    for (int cc = 0; cc < j; ++cc) gt += sH[t * kMaxBasis + cc] * sa[cc];
    sg[t] = gt;
  }
  __syncthreads();
  if (t < 32) {
    // one lane per basis index k <= j; sums over the half-wave, fixed butterfly
    const int k = t;
    const double ab = wsum32(k < j ? sa[k] * sb[k] : 0.0);
    const double inv_tau = 1.0 / tau;
    double p = 0.0, h = 0.0, gk = 0.0;
    if (k <= j) {
      gk = sg[k];
      p = k < j ? sb[k] : (ssc[1] - ab) * inv_tau;
      h = (p - gk) * inv_tau;
      st->cpart[k] = h;
      st->coef[kCoefD + k] = gk + tau * h;
      col[k] = h;
      if (k < j) st->coef[k] = sa[k];
    }
    const double gp = wsum32(gk * p), gg = wsum32(gk * gk), hh = wsum32(h * h);
    if (k == 0) {
      const double av2 = (ssc[2] - 2.0 * gp + gg) * inv_tau * inv_tau;
      const double r2 = av2 - hh;
      st->coef[kCoefInvTau] = inv_tau;
      if (r2 > 1e-28 * av2) {
        const double rho = sqrt(r2);
        st->rho = rho;
        st->coef[kCoefInvTauRho] = inv_tau / rho;
        sbreak = 0;
      } else {
        sbreak = 1;
      }
    }
  }
  __syncthreads();
  if (sbreak && t <= j + 1) {
    // exhausted Krylov space: column j is h with a zero subdiagonal
    if (t == j + 1) col[t] = 0.0;
    st->Hh[t * kMaxBasis + j] = t <= j ? col[t] : 0.0;
  }
  __syncthreads();
  if (sbreak) givens_col(st, j, col, gstage);  // sets the stop flag (zero subdiagonal)
}

// y = a x + b y, a possibly scaled by a device scalar (*as).  With b = 0, y
// is not read (it may be fresh memory holding NaN, and 0 * NaN is NaN).
__global__ __launch_bounds__(kBlock) void k_axpby(size_t n, double a, const double *__restrict__ as,
                                                  const double *__restrict__ x, double b, double *__restrict__ y,
                                                  const int *__restrict__ done) {
  if (done && *done) return;
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  const double aa = as ? a * (*as) : a;
  y[i] = b == 0.0 ? aa * x[i] : aa * x[i] + b * y[i];
}

__global__ __launch_bounds__(kBlock) void k_fill(size_t n, double v, double *__restrict__ y) {
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i < n) y[i] = v;
}

// w = sum_{k <= jlast} y_k V_k
__global__ __launch_bounds__(kBlock) void k_combine(size_t n, double *__restrict__ w, const double *__restrict__ V,
                                                    size_t ldv, const GmresState *__restrict__ s) {
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  const int nk = s->jlast + 1;
  double acc = 0.0;
  for (int k = 0; k < nk; ++k) acc += s->y[k] * V[(size_t)k * ldv + i];
  w[i] = acc;
}

// start of a cycle: beta = ||r|| (r = b - A x), target from ||b|| on the first
__global__ void k_gmres_start(GmresState *s, const double *__restrict__ nrm2, int first, double eps) {
  if (threadIdx.x != 0) return;
  const double beta = sqrt(nrm2[0]);
  s->beta = beta;
  if (first) {
    s->target = eps * beta;  // x0 = 0: r = b
    s->iter = 0;
  }
  for (int k = 0; k <= s->m; ++k) s->g[k] = 0.0;
  s->g[0] = beta;
  s->done = (beta == 0.0) ? 2 : 0;
  s->invhn = beta == 0.0 ? 0.0 : 1.0 / beta;
  s->jlast = -1;
}

// back-substitution (KspSolver.C:968-982)
// (the triangle and g are staged in LDS by the block first, then one lane
// runs the recurrence in the reference's order)
__global__ __launch_bounds__(256) void k_gmres_backsolve(GmresState *s) {
  constexpr int M = kMaxBasis;
  __shared__ double sH[M * M], sg[M], sy[M];
  const int j = s->jlast;
  for (int e = threadIdx.x; e < M * M; e += blockDim.x) sH[e] = s->H[e];
  if (threadIdx.x < M) sg[threadIdx.x] = s->g[threadIdx.x];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = j; i >= 0; --i) {
      double yi = sg[i];
      for (int k = j; k >= i + 1; --k) yi -= sH[i * M + k] * sy[k];
      sy[i] = yi / sH[i * M + i];
    }
  }
  __syncthreads();
  if ((int)threadIdx.x <= j) s->y[threadIdx.x] = sy[threadIdx.x];
}

// count nodes whose density or pressure is not positive
__global__ __launch_bounds__(kBlock) void k_check_states(int N, const double *__restrict__ U, Gas g,
                                                         int *__restrict__ bad) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  const double *u = U + 5 * (size_t)n;
  double rho = u[0];
  double vel2 = (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / (rho * rho);
  double p = (g.gam - 1.0) * (u[4] - 0.5 * rho * vel2) - g.gam * g.pstiff;
  if (!(rho > 0.0) || !(p + g.pstiff > 0.0)) atomicAdd(bad, 1);
}

// ImplicitTsDesc::checkFailSafe -> SubDomain::checkFailSafe (SubDomain.C:7123-7153):
// tag0 = node non-physical; the first-order nodes gain every node with a
// non-physical neighbour (tag1 = OR of the neighbours' tag0, SpaceOperator::fix)
__global__ __launch_bounds__(kBlock) void k_fs_state_tags(int N, const double *__restrict__ U, Gas g,
                                                          int *__restrict__ tag0) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  const double *u = U + 5 * (size_t)n;
  double rho = u[0];
  double vel2 = (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / (rho * rho);
  double p = (g.gam - 1.0) * (u[4] - 0.5 * rho * vel2) - g.gam * g.pstiff;
  tag0[n] = (!(rho > 0.0) || !(p + g.pstiff > 0.0)) ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_fs_state_fix(DevTables t, const int *__restrict__ tag0,
                                                         int *__restrict__ mask) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= t.N) return;
  int tag1 = 0;
  for (int e = t.rowptr[n]; e < t.rowptr[n + 1]; ++e) tag1 |= tag0[t.adj[e]];
  if (tag1) mask[n] = 1;
}

// ---------------------------------------------------------------------------
// MatVecProdFD (MatVecProd.C:88-118, 262-372, 879-929)
// ---------------------------------------------------------------------------
// partial sums of eps0 |U| + eps0 over all entries (computeEpsilon)
__global__ __launch_bounds__(kBlock) void k_eps_sum(size_t n, const double *__restrict__ U,
                                                    double *__restrict__ partial) {
  const size_t slice = (n + kDotBlocks - 1) / kDotBlocks;
  const size_t b0 = blockIdx.x * slice, b1 = min(n, b0 + slice);
  const double eps0 = 1.e-6;
  double acc = 0.0;
  for (size_t i = b0 + threadIdx.x; i < b1; i += kBlock) acc += eps0 * fabs(U[i]) + eps0;
  block_reduce_store_1(acc, partial);
}

// scal[0] = p.p, scal[1] = sum(eps0 |U| + eps0) -> scal[2] = eps, [3] = 1/eps, [4] = 0.5/eps
__global__ void k_fd_eps(double *scal, double size, const int *__restrict__ done) {
  if (done && *done) return;
  if (threadIdx.x != 0) return;
  const double norm = sqrt(scal[0]);
  const double eps = norm < 1.0e-14 ? 1.e-6 : scal[1] / (size * norm);
  scal[2] = eps;
  scal[3] = 1.0 / eps;
  scal[4] = 0.5 / eps;
}

// Q = U + s eps p (s = +1 or -1)
__global__ __launch_bounds__(kBlock) void k_fd_shift(size_t n, const double *__restrict__ U,
                                                     const double *__restrict__ p, const double *__restrict__ scal,
                                                     double s, double *__restrict__ Q, const int *__restrict__ done) {
  if (done && *done) return;
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  const double ep = scal[2] * p[i];
  Q[i] = s > 0.0 ? U[i] + ep : U[i] - ep;
}

// TimeState::add_dAW_dt, backward Euler, Un = U (TimeState.C:40-73):
// F += (1/dt) (1 Q + (-1) Un)
__global__ __launch_bounds__(kBlock) void k_fd_time(size_t n, const double *__restrict__ Q,
                                                    const double *__restrict__ U, const double *__restrict__ dt,
                                                    double *__restrict__ F, const int *__restrict__ done) {
  if (done && *done) return;
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  const double invDt = 1.0 / dt[i / 5];
  F[i] += invDt * (Q[i] - U[i]);
}

// prod = scal[idx] (Fa - Fb)
__global__ __launch_bounds__(kBlock) void k_fd_diff(size_t n, const double *__restrict__ Fa,
                                                    const double *__restrict__ Fb, const double *__restrict__ scal,
                                                    int idx, double *__restrict__ prod, const int *__restrict__ done) {
  if (done && *done) return;
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  prod[i] = scal[idx] * (Fa[i] - Fb[i]);
}

// err[3] += err[0]: every residual pass restarts err[0]; the products of
// MatVecProdFD sum the counts of all their residual evaluations in err[3]
__global__ void k_err_accumulate(int *err, const int *__restrict__ done) {
  if (done && *done) return;
  if (threadIdx.x == 0) err[3] += err[0];
}

// Blocks of the row-pair sweeps over n values: about 8 entries per lane per
// sweep, between 64 blocks and one per CU (fixed by n, so the reduction
// order never depends on the call).
static int sweep_parts(size_t n) {
  return (int)std::min<size_t>(kSweepBlocks, std::max<size_t>(64, (n + 8 * kBlock - 1) / (8 * kBlock)));
}

static inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// Scratch of the finite-difference products.
struct FdScratch {
  double *Q = nullptr, *Fa = nullptr, *Fb = nullptr;  // [5N]
  double *partial = nullptr;                           // [kDotBlocks]
  double *scal = nullptr;                              // [8], layout of k_fd_eps
};

// computeEpsilon, the state part: scal[1] = sum(eps0 |U| + eps0).
static void enqueue_fd_state_sum(aerofx_ctx *c, const FdScratch &s, const double *U) {
  const size_t n = 5 * (size_t)c->N;
  hipLaunchKernelGGL(k_eps_sum, dim3(kDotBlocks), dim3(kBlock), 0, c->stream, n, U, s.partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, s.partial, kDotBlocks, s.scal + 1,
                     nullptr);
}

// One MatVecProdFD::apply (MatVecProd.C:262-372) with the state sum already
// in scal[1]: eps from ||p|| (computeEpsilon), then
//   order 1: prod = (F(U + eps p) - F0) / eps
//   order 2: prod = (F(U + eps p) - F(U - eps p)) / (2 eps)
// with F(Q) = R(Q) + (Q - U) / dt (dt NULL: R only).  The residual passes'
// error counts are added to err[3].  Every kernel returns at once when
// *done != 0 (done may be NULL).
static int enqueue_fd_product(aerofx_ctx *c, const FdScratch &s, int nparts, const double *U, const double *dt,
                              const double *F0, const double *p, double *prod, int order, const int *done) {
  const size_t n = 5 * (size_t)c->N;
  int rc;
  hipLaunchKernelGGL(k_sumsq, dim3(nparts), dim3(kBlock), 0, c->stream, n, p, s.partial, done);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, s.partial, nparts, s.scal, done);
  hipLaunchKernelGGL(k_fd_eps, dim3(1), dim3(64), 0, c->stream, s.scal, (double)n, done);
  hipLaunchKernelGGL(k_fd_shift, grid_for(n), dim3(kBlock), 0, c->stream, n, U, p, s.scal, 1.0, s.Q, done);
  if ((rc = launch_residual(c, s.Q, s.Fa, true, done))) return rc;
  hipLaunchKernelGGL(k_err_accumulate, dim3(1), dim3(64), 0, c->stream, c->d_err, done);
  if (dt) hipLaunchKernelGGL(k_fd_time, grid_for(n), dim3(kBlock), 0, c->stream, n, s.Q, U, dt, s.Fa, done);
  if ((rc = launch_bcs_residual(c, s.Q, s.Fa, done))) return rc;
  if (order == 1) {
    hipLaunchKernelGGL(k_fd_diff, grid_for(n), dim3(kBlock), 0, c->stream, n, s.Fa, F0, s.scal, 3, prod, done);
  } else {
    hipLaunchKernelGGL(k_fd_shift, grid_for(n), dim3(kBlock), 0, c->stream, n, U, p, s.scal, -1.0, s.Q, done);
    if ((rc = launch_residual(c, s.Q, s.Fb, true, done))) return rc;
    hipLaunchKernelGGL(k_err_accumulate, dim3(1), dim3(64), 0, c->stream, c->d_err, done);
    if (dt) hipLaunchKernelGGL(k_fd_time, grid_for(n), dim3(kBlock), 0, c->stream, n, s.Q, U, dt, s.Fb, done);
    if ((rc = launch_bcs_residual(c, s.Q, s.Fb, done))) return rc;
    hipLaunchKernelGGL(k_fd_diff, grid_for(n), dim3(kBlock), 0, c->stream, n, s.Fa, s.Fb, s.scal, 4, prod, done);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "finite-difference product launch");
}

// The whole of aerofx_jvp_fd: MatVecProdFD::evaluate at U (F0 = F(U) in
// s.Fb for order 1) and one apply; the summed error count is left in err[3].
static int enqueue_jvp_fd(aerofx_ctx *c, const FdScratch &s, int nparts, const double *U, const double *p,
                          double *prod, int order, const double *dt) {
  int rc;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->d_err + 3, 1, nullptr);
  enqueue_fd_state_sum(c, s, U);
  if (order == 1) {
    // F(U): the time term vanishes at Q = Un
    if ((rc = launch_residual(c, U, s.Fb, true))) return rc;
    hipLaunchKernelGGL(k_err_accumulate, dim3(1), dim3(64), 0, c->stream, c->d_err, nullptr);
    if ((rc = launch_bcs_residual(c, U, s.Fb, nullptr))) return rc;
  }
  return enqueue_fd_product(c, s, nparts, U, dt, s.Fb, p, prod, order, nullptr);
}

struct Krylov {
  int m = 0;
  size_t n = 0;
  size_t ldv = 0;             // basis stride: n rounded up to 256 B, so every V_k is 16-B aligned
  int nparts = 0;             // blocks of the dot-product sweeps (fixed by n)
  double *V = nullptr;        // [m+1][n] basis
  double *w = nullptr, *z = nullptr, *r = nullptr, *x = nullptr, *b = nullptr, *R = nullptr;
  double *partial = nullptr;  // [2 kMaxBasis + 3][kDotBlocks]
  double *scal = nullptr;     // device scalars
  GmresState *st = nullptr;   // device
  GmresState *h_st = nullptr; // pinned host mirror
  // the Arnoldi sequence of one cycle, captured once per (m, precision, transpose)
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t order_ev = nullptr;
  long long graph_key = -1;
  const void *graph_fd[3] = {nullptr, nullptr, nullptr};  // Q, F0, dt captured with a matrix-free operator
  double *h_scal = nullptr;   // pinned
};

void free_krylov(aerofx_ctx *c) {
  if (!c->ksp) return;
  if (c->ksp->exec) (void)hipGraphExecDestroy(c->ksp->exec);
  if (c->ksp->graph) (void)hipGraphDestroy(c->ksp->graph);
  if (c->ksp->cap_stream) (void)hipStreamDestroy(c->ksp->cap_stream);
  if (c->ksp->order_ev) (void)hipEventDestroy(c->ksp->order_ev);
  if (c->ksp->h_st) (void)hipHostFree(c->ksp->h_st);
  if (c->ksp->h_scal) (void)hipHostFree(c->ksp->h_scal);
  delete c->ksp;
  c->ksp = nullptr;  // device buffers belong to c->allocs
}

static int ensure_krylov(aerofx_ctx *c, int m) {
  if (m < 1 || m > kMaxBasis - 1) return set_error(AEROFX_ERR_ARG, "gmres: Krylov dimension must be in [1, 31]");
  if (c->ksp && c->ksp->m >= m) return 0;
  free_krylov(c);
  Krylov *k = new Krylov();
  k->m = m;
  k->n = 5 * (size_t)c->N;
  // the sweeps hold up to 32 basis rows per lane: one wave per SIMD
  k->nparts = sweep_parts(k->n);
  void *p = nullptr;
  int rc;
  k->ldv = (k->n + 31) & ~(size_t)31;
  if ((rc = ctx_alloc(c, sizeof(double) * (size_t)(m + 1) * k->ldv, &p))) { delete k; return rc; }
  k->V = (double *)p;
  double **vecs[6] = {&k->w, &k->z, &k->r, &k->x, &k->b, &k->R};
  for (double **v : vecs) {
    if ((rc = ctx_alloc(c, sizeof(double) * k->n, &p))) { delete k; return rc; }
    *v = (double *)p;
  }
  if ((rc = ctx_alloc(c, sizeof(double) * (2 * kMaxBasis + 3) * kDotBlocks, &p))) { delete k; return rc; }
  k->partial = (double *)p;
  if ((rc = ctx_alloc(c, sizeof(double) * 16, &p))) { delete k; return rc; }
  k->scal = (double *)p;

  if ((rc = ctx_alloc(c, sizeof(GmresState), &p))) { delete k; return rc; }
  k->st = (GmresState *)p;
  hipError_t e = hipHostMalloc((void **)&k->h_st, sizeof(GmresState), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void **)&k->h_scal, sizeof(double) * 16, hipHostMallocDefault);
  if (e != hipSuccess) { delete k; return set_hip_error(e, "hipHostMalloc"); }
  c->ksp = k;
  return 0;
}

// out[0] = w.w (fixed reduction order)
static void norm2(aerofx_ctx *c, const double *w, double *out) {
  Krylov *k = c->ksp;
  hipLaunchKernelGGL(k_sumsq, dim3(k->nparts), dim3(kBlock), 0, c->stream, k->n, w, k->partial, nullptr);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, k->partial, k->nparts, out, nullptr);
}

// The finite-difference scratch of the context (allocated on first use).
static int ensure_fd(aerofx_ctx *c, FdScratch *s) {
  const size_t n = 5 * (size_t)c->N;
  if (!c->d_fd) {
    void *p = nullptr;
    int rc = ctx_alloc(c, sizeof(double) * (4 * n + kDotBlocks + 16), &p);
    if (rc) return rc;
    c->d_fd = static_cast<double *>(p);
  }
  s->Q = c->d_fd;
  s->Fa = s->Q + n;
  s->Fb = s->Fa + n;
  s->partial = c->d_fd + 4 * n;
  s->scal = s->partial + kDotBlocks;
  return 0;
}
static double *fd_own_F0(aerofx_ctx *c) { return c->d_fd + 3 * 5 * (size_t)c->N; }

// y = A p with the selected MatVecProd (transposed solves: always H1^T).
static int launch_operator(aerofx_ctx *c, const double *p, double *y, const int *done, bool trans) {
  if (trans || c->mvp_type == AEROFX_MVP_H1) return launch_h1_apply_flag(c, p, y, done, trans);
  if (c->mvp_type == AEROFX_MVP_H2) return launch_h2_apply(c, c->mvp_Q, p, c->mvp_dt, y, done);
  FdScratch s;
  int rc = ensure_fd(c, &s);
  return rc ? rc
            : enqueue_fd_product(c, s, sweep_parts(5 * (size_t)c->N), c->mvp_Q, c->mvp_dt, c->mvp_F0, p, y,
                                 c->mvp_order, done);
}

// MatVecProdFD::evaluate: selects the finite-difference operator about Q.
// F0 NULL: F(Q) = R(Q) is evaluated into the context's own buffer.  Enqueues
// the state part of computeEpsilon and clears the error sum err[3].
static int select_fd(aerofx_ctx *c, const double *Q, const double *F0, const double *dt, int order) {
  FdScratch s;
  int rc = ensure_fd(c, &s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->d_err + 3, 1, nullptr);
  if (!F0) {
    if ((rc = launch_residual(c, Q, fd_own_F0(c), true))) return rc;
    hipLaunchKernelGGL(k_err_accumulate, dim3(1), dim3(64), 0, c->stream, c->d_err, nullptr);
    if ((rc = launch_bcs_residual(c, Q, fd_own_F0(c), nullptr))) return rc;
    F0 = fd_own_F0(c);
  }
  enqueue_fd_state_sum(c, s, Q);
  c->mvp_type = AEROFX_MVP_FD;
  c->mvp_Q = Q;
  c->mvp_F0 = F0;
  c->mvp_dt = dt;
  c->mvp_order = order;
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "mvp_select launch");
}

// MatVecProdH2::evaluate: selects the exact product about Q (dt NULL: no
// time term).
static int select_h2(aerofx_ctx *c, const double *Q, const double *dt) {
  int rc = launch_h2_select(c, Q);
  if (rc) return rc;
  c->mvp_type = AEROFX_MVP_H2;
  c->mvp_Q = Q;
  c->mvp_F0 = nullptr;
  c->mvp_dt = dt;
  c->mvp_order = 1;
  return 0;
}

static int fetch_state(aerofx_ctx *c) {
  hipError_t e = hipMemcpyAsync(c->ksp->h_st, c->ksp->st, sizeof(GmresState), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? 0 : set_hip_error(e, "gmres state copy");
}

// The m Arnoldi steps of one cycle (no host decisions: a converged cycle
// leaves the remaining kernels to return at once on the stop flag).  Per
// step: H1 apply, one sweep over the basis (k_dots2), its reduction with the
// small Gram-Schmidt algebra and the Givens update of the previous column
// (k_dcgs_reduce), and one update pass over the basis fused with the
// preconditioner (k_dcgs_update); a last sweep and reduction complete the
// cycle's final column.
static int enqueue_arnoldi(aerofx_ctx *c, int m, bool trans) {
  Krylov *k = c->ksp;
  const size_t n = k->n;
  int *done = &k->st->done;
  int rc;
  {
    // u_0 = r / ||r||, z = P u_0
    ProfScope ps(c, AEROFX_TIMER_PREC_APPLY);
    if ((rc = launch_prec_norm_flag(c, k->r, &k->st->invhn, k->V, k->ldv, 0, nullptr, k->V, k->z, done, trans)))
      return rc;
    // RAS / ILU(0) replaces the block-Jacobi result (forward solves only)
    if (!trans && c->prec_type == AEROFX_PREC_RAS_ILU0 && (rc = launch_ilu_solve(c, k->V, k->z, done))) return rc;
  }
  for (int j = 0; j < m; ++j) {
    {
      ProfScope ps(c, AEROFX_TIMER_JAC_APPLY);
      if ((rc = launch_operator(c, k->z, k->w, done, trans))) return rc;
    }
    {
      ProfScope ps(c, AEROFX_TIMER_GRAM_SCHMIDT);
      hipLaunchKernelGGL((k_dots2<false>), dim3(k->nparts), dim3(kBlock), 0, c->stream, n, k->V, k->ldv, j, k->w,
                         k->partial, done);
      hipLaunchKernelGGL((k_dcgs_reduce<false>), dim3(1), dim3(kReduceThreads), 0, c->stream, k->partial,
                         k->nparts, k->st, j);
    }
    {
      ProfScope ps(c, AEROFX_TIMER_PREC_APPLY);
      if ((rc = launch_dcgs_update(c, k->V, k->ldv, j, k->w, k->st->coef, &k->st->done, &k->st->jlast, k->z,
                                   trans)))
        return rc;
      if (!trans && c->prec_type == AEROFX_PREC_RAS_ILU0 &&
          (rc = launch_ilu_solve(c, k->V + (size_t)(j + 1) * k->ldv, k->z, done)))
        return rc;
    }
  }
  {
    ProfScope ps(c, AEROFX_TIMER_GRAM_SCHMIDT);
    hipLaunchKernelGGL((k_dots2<true>), dim3(k->nparts), dim3(kBlock), 0, c->stream, n, k->V, k->ldv, m, nullptr,
                       k->partial, done);
    hipLaunchKernelGGL((k_dcgs_reduce<true>), dim3(1), dim3(kReduceThreads), 0, c->stream, k->partial, k->nparts,
                       k->st, m);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "arnoldi launch");
}

// Runs the cycle's Arnoldi steps: eagerly when profiling (per-kernel events),
// otherwise as one hipGraph captured on a private stream, ordered after the
// context stream by events.  The graph holds kernel nodes only.
static int arnoldi_cycle(aerofx_ctx *c, int m, bool trans) {
  Krylov *k = c->ksp;
  if (c->prof) return enqueue_arnoldi(c, m, trans);
  const int mvp = trans ? AEROFX_MVP_H1 : c->mvp_type;
  const bool fd = mvp == AEROFX_MVP_FD, mf = mvp != AEROFX_MVP_H1;
  const long long key =
      ((((((((long long)m * 2 + (trans ? 1 : 0)) * 2 + c->jac_precision) * 3 + mvp) * 3 + (fd ? c->mvp_order : 0)) *
             3 +
         c->residual_kernels) *
            2 +
        c->prec_type) *
           2 +
       c->fc.low_mach) *
          2048 +
      (c->prec_type == AEROFX_PREC_RAS_ILU0 ? c->prec_subdomain : 0);
  // the captured products read these buffers: a new selection re-captures
  const void *fdp[3] = {mf ? c->mvp_Q : nullptr, fd ? c->mvp_F0 : nullptr, mf ? c->mvp_dt : nullptr};
  const bool same_fd = k->graph_fd[0] == fdp[0] && k->graph_fd[1] == fdp[1] && k->graph_fd[2] == fdp[2];
  hipError_t e = hipSuccess;
  if (!k->cap_stream) e = hipStreamCreateWithFlags(&k->cap_stream, hipStreamNonBlocking);
  if (e == hipSuccess && !k->order_ev) e = hipEventCreateWithFlags(&k->order_ev, hipEventDisableTiming);
  if (e != hipSuccess) return set_hip_error(e, "arnoldi stream");
  if (!k->exec || k->graph_key != key || !same_fd) {
    if (k->exec) (void)hipGraphExecDestroy(k->exec);
    if (k->graph) (void)hipGraphDestroy(k->graph);
    k->exec = nullptr;
    k->graph = nullptr;
    hipStream_t saved = c->stream;
    c->stream = k->cap_stream;
    e = hipStreamBeginCapture(k->cap_stream, hipStreamCaptureModeThreadLocal);
    int rc = e == hipSuccess ? enqueue_arnoldi(c, m, trans) : 0;
    hipError_t e2 = hipStreamEndCapture(k->cap_stream, &k->graph);
    c->stream = saved;
    if (rc) return rc;
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&k->exec, k->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return set_hip_error(e, "arnoldi graph");
    k->graph_key = key;
    for (int q = 0; q < 3; ++q) k->graph_fd[q] = fdp[q];
  }
  hipEvent_t ev = k->order_ev;
  e = hipEventRecord(ev, c->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(k->cap_stream, ev, 0);
  if (e == hipSuccess) e = hipGraphLaunch(k->exec, k->cap_stream);
  if (e == hipSuccess) e = hipEventRecord(ev, k->cap_stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, ev, 0);
  return e == hipSuccess ? 0 : set_hip_error(e, "arnoldi graph launch");
}

struct GmresOut {
  int its = 0;
  double res = 0, target = 0, true_res = -1.0;
};

// Right-preconditioned GMRES(m) on H1 with block-Jacobi: A x = b, x0 = 0
// (ImplicitCoupledTsDesc::solveLinearSystem sets dQ = 0).  Enqueues one
// cycle at a time and synchronises once per cycle.
static int gmres(aerofx_ctx *c, const double *b, double *x, int m, double eps, int maxits, GmresOut *out,
                 bool trans = false) {
  int rc = ensure_krylov(c, m);
  if (rc) return rc;
  // the captured transposed cycle reads the mirrored blocks: bring them up
  // to date before any capture
  if (trans && (rc = ensure_mirror(c))) return rc;
  Krylov *k = c->ksp;
  const size_t n = k->n;
  *k->h_st = GmresState{};
  k->h_st->maxits = maxits;
  k->h_st->m = m;
  hipError_t e = hipMemcpyAsync(k->st, k->h_st, sizeof(GmresState), hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return set_hip_error(e, "gmres init copy");
  int *done = &k->st->done;
  hipLaunchKernelGGL(k_fill, grid_for(n), dim3(kBlock), 0, c->stream, n, 0.0, x);
  for (int cycle = 0;; ++cycle) {
    // r = b - A x
    if (cycle == 0) {
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, b, 0.0, k->r, nullptr);
    } else {
      if ((rc = launch_operator(c, x, k->w, nullptr, trans))) return rc;
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, b, 0.0, k->r, nullptr);
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, -1.0, nullptr, k->w, 1.0, k->r,
                         nullptr);
    }
    norm2(c, k->r, k->scal);
    hipLaunchKernelGGL(k_gmres_start, dim3(1), dim3(64), 0, c->stream, k->st, k->scal, cycle == 0 ? 1 : 0, eps);
    if ((rc = arnoldi_cycle(c, m, trans))) return rc;
    if ((rc = fetch_state(c))) return rc;
    const GmresState &hs = *k->h_st;
    if (hs.jlast >= 0) {
      hipLaunchKernelGGL(k_gmres_backsolve, dim3(1), dim3(256), 0, c->stream, k->st);
      hipLaunchKernelGGL(k_combine, grid_for(n), dim3(kBlock), 0, c->stream, n, k->w, k->V, k->ldv, k->st);
      if ((rc = launch_prec_apply(c, k->w, k->r, trans))) return rc;
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, k->r, 1.0, x, nullptr);
    }
    if (hs.done) break;  // converged, maxits, zero residual or lucky breakdown
  }
  e = hipGetLastError();
  if (e != hipSuccess) return set_hip_error(e, "gmres launch");
  out->its = k->h_st->iter;
  out->res = k->h_st->l2res;
  out->target = k->h_st->target;
  if (c->check_final_res) {
    // GmresSolver::solve with checkFinalRes (KspSolver.C:455-469): the true
    // residual ||b - A x|| against the target; above it the solve reports -999
    if ((rc = launch_operator(c, x, k->w, nullptr, trans))) return rc;
    hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, b, -1.0, k->w, nullptr);
    norm2(c, k->w, k->scal);
    e = hipMemcpyAsync(k->h_scal, k->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_hip_error(e, "gmres final residual");
    out->true_res = std::sqrt(k->h_scal[0]);
    if (out->true_res > out->target) out->its = -999;
  }
  return 0;
}

}  // namespace aerofx

using namespace aerofx;

extern "C" {

int aerofx_jvp_fd(aerofx_ctx *c, const double *U, const double *p, double *Jp, int order, const double *dt) {
  if (!c || !U || !p || !Jp) return set_error(AEROFX_ERR_ARG, "jvp_fd: null argument");
  if (order != 1 && order != 2) return set_error(AEROFX_ERR_ARG, "jvp_fd: order 1 or 2");
  FdScratch s;
  int rc = ensure_fd(c, &s);
  if (rc) return rc;
  if ((rc = enqueue_jvp_fd(c, s, sweep_parts(5 * (size_t)c->N), U, p, Jp, order, dt))) return rc;
  hipError_t e = hipMemcpyAsync(c->h_err + 3, c->d_err + 3, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "jvp_fd sync");
  return c->h_err[3];
}

int aerofx_mvp_select(aerofx_ctx *c, int type, const double *Q, const double *F0, const double *dt, int order) {
  if (!c) return set_error(AEROFX_ERR_ARG, "mvp_select: null ctx");
  if (type == AEROFX_MVP_H1) {
    c->mvp_type = AEROFX_MVP_H1;
    c->mvp_Q = c->mvp_F0 = c->mvp_dt = nullptr;
    return AEROFX_OK;
  }
  if (type != AEROFX_MVP_FD && type != AEROFX_MVP_H2)
    return set_error(AEROFX_ERR_ARG, "mvp_select: AEROFX_MVP_H1, AEROFX_MVP_FD or AEROFX_MVP_H2");
  if (!Q) return set_error(AEROFX_ERR_ARG, "mvp_select: null state");
  if (type == AEROFX_MVP_H2 && c->fc.low_mach)
    return set_error(AEROFX_ERR_STATE, "mvp_select: exact products not available with low-Mach preconditioning");
  if (type == AEROFX_MVP_H2)
    if (int rf = refuse_failsafe_nodes(c, "mvp_select")) return rf;
  if (type == AEROFX_MVP_H2) {
    int rc = select_h2(c, Q, dt);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(c->h_err, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_hip_error(e, "mvp_select sync");
    return c->h_err[0];
  }
  if (order != 1 && order != 2) return set_error(AEROFX_ERR_ARG, "mvp_select: order 1 or 2");
  int rc = select_fd(c, Q, F0, dt, order);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(c->h_err + 3, c->d_err + 3, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "mvp_select sync");
  return c->h_err[3];
}

int aerofx_prec_select(aerofx_ctx *c, int type, int subdomain_nodes) {
  if (!c) return set_error(AEROFX_ERR_ARG, "prec_select: null ctx");
  if (type != AEROFX_PREC_JACOBI && type != AEROFX_PREC_RAS_ILU0)
    return set_error(AEROFX_ERR_ARG, "prec_select: AEROFX_PREC_JACOBI or AEROFX_PREC_RAS_ILU0");
  if (type == AEROFX_PREC_RAS_ILU0 && (subdomain_nodes < 1 || subdomain_nodes > 1024))
    return set_error(AEROFX_ERR_ARG, "prec_select: subdomain size in [1, 1024]");
  c->prec_type = type;
  if (type == AEROFX_PREC_RAS_ILU0) c->prec_subdomain = subdomain_nodes;
  return AEROFX_OK;
}

int aerofx_mvp_apply(aerofx_ctx *c, const double *p, double *prod) {
  if (!c || !p || !prod) return set_error(AEROFX_ERR_ARG, "mvp_apply: null argument");
  if (c->mvp_type == AEROFX_MVP_H1) {
    if (!c->d_jacD) return set_error(AEROFX_ERR_STATE, "mvp_apply: operators not set up");
    int rc = launch_h1_apply(c, p, prod);
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? AEROFX_OK : set_hip_error(e, "mvp_apply sync");
  }
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->d_err + 3, 1, nullptr);
  int rc = launch_operator(c, p, prod, nullptr, false);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(c->h_err + 3, c->d_err + 3, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "mvp_apply sync");
  return c->h_err[3];
}

// Outer basis size of the flexible GMRES of the derivative solves.
constexpr int kDerivRestart = 20;

// The inner operator of the derivative solves: H1 + diag(1/dt) at U (the
// pseudo-time step of the given CFL; cfl <= 0: none) and its block-Jacobi
// inverse in `precision`; the zero-pivot count is checked.
static int derivative_operators(aerofx_ctx *c, const double *U, double cfl, int precision, const char *who) {
  int rc;
  if ((rc = launch_primitive(c, U, c->d_V))) return rc;
  if (cfl > 0.0 && (rc = launch_timestep(c, cfl, c->d_dt))) return rc;
  if ((rc = launch_jacobian(c, cfl > 0.0 ? c->d_dt : nullptr, precision, U))) return rc;
  hipError_t e = hipMemcpyAsync(c->h_err + 1, c->d_err + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, who);
  if (c->h_err[1]) return set_error(AEROFX_ERR_STATE, std::string(who) + ": zero pivot in a diagonal block");
  return 0;
}

int aerofx_adjoint_solve(aerofx_ctx *c, const double *U, const double *rhs, double *psi,
                         const aerofx_adjoint_opts *o, aerofx_adjoint_stats *st) {
  if (!c || !U || !rhs || !psi || !o || !st) return set_error(AEROFX_ERR_ARG, "adjoint_solve: null argument");
  if (c->fc.low_mach)
    return set_error(AEROFX_ERR_STATE, "adjoint_solve: not available with low-Mach preconditioning");
  if (int rf = refuse_failsafe_nodes(c, "adjoint_solve")) return rf;
  if (o->precision != 0 && o->precision != 1) return set_error(AEROFX_ERR_ARG, "adjoint_solve: precision 0 or 1");
  int rc = ensure_krylov(c, o->gmres_m);
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  *st = aerofx_adjoint_stats{};
  if ((rc = derivative_operators(c, U, o->cfl, o->precision, "adjoint_solve"))) return rc;
  // with the Navier-Stokes wall conditions the system is A^T psi = rhs with
  // A = P dR/dU + B (the wall rows replaced): A^T psi = (dR/dU)^T P psi + B^T psi
  const bool walls = c->viscous && c->nwall;
  if (walls && !c->d_adj_tmp) {
    void *p = nullptr;
    if ((rc = ctx_alloc(c, sizeof(double) * 5 * (size_t)c->N, &p))) return rc;
    c->d_adj_tmp = (double *)p;
  }
  // (dR/dU)^T psi = rhs: flexible GMRES on the exact adjoint products, each
  // step preconditioned by transposed GMRES on H1^T with block-Jacobi^T
  FgOps ops;
  ops.apply = [&](const double *x, double *y, int nq) -> int {
    for (int q = 0; q < nq; ++q) {
      const size_t off = (size_t)q * c->ksp->n;
      int r;
      if (walls) {
        if ((r = launch_bcs_adjoint(c, U, x + off, c->d_adj_tmp, nullptr, true)) ||
            (r = launch_adjoint(c, U, c->d_adj_tmp, y + off, nullptr)) ||
            (r = launch_bcs_adjoint(c, U, x + off, nullptr, y + off, false)))
          return r;
      } else if ((r = launch_adjoint(c, U, x + off, y + off, nullptr))) {
        return r;
      }
    }
    return 0;
  };
  ops.prec = [&](const double *v, double *z, int nq, int *inner) -> int {
    for (int q = 0; q < nq; ++q) {
      const size_t off = (size_t)q * c->ksp->n;
      GmresOut go;
      int r = gmres(c, v + off, z + off, o->gmres_m, o->gmres_eps, o->gmres_maxits, &go, true);
      if (r) return r;
      *inner += go.its;
    }
    return 0;
  };
  FgResult fr;
  if ((rc = fgmres_solve(c, 1, rhs, psi, kDerivRestart, o->eps, o->max_its, ops, &fr))) return rc;
  st->its = fr.outer_its;
  st->krylov_its = fr.inner_its;
  st->converged = fr.converged;
  st->res0 = fr.res0;
  st->res = fr.res;
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return AEROFX_OK;
}

int aerofx_direct_solve(aerofx_ctx *c, const double *U, int nrhs, const double *rhs, double *x,
                        const aerofx_adjoint_opts *o, aerofx_adjoint_stats *st) {
  if (!c || !U || !rhs || !x || !o || !st) return set_error(AEROFX_ERR_ARG, "direct_solve: null argument");
  if (c->fc.low_mach)
    return set_error(AEROFX_ERR_STATE, "direct_solve: not available with low-Mach preconditioning");
  if (int rf = refuse_failsafe_nodes(c, "direct_solve")) return rf;
  if (nrhs < 1 || nrhs > 16) return set_error(AEROFX_ERR_ARG, "direct_solve: 1 <= nrhs <= 16");
  if (o->precision != 0 && o->precision != 1) return set_error(AEROFX_ERR_ARG, "direct_solve: precision 0 or 1");
  int rc = ensure_krylov(c, o->gmres_m);
  if (rc) return rc;
  // the inner solves use the assembled H1, whatever operator is selected
  struct Restore {
    aerofx_ctx *c;
    int type;
    ~Restore() { c->mvp_type = type; }
  } restore{c, c->mvp_type};
  c->mvp_type = AEROFX_MVP_H1;
  auto t0 = std::chrono::steady_clock::now();
  *st = aerofx_adjoint_stats{};
  // block GMRES takes FP64 blocks and makes its own FP32 copies (precision 1)
  const bool blocked = nrhs > 1 && c->max_degree <= 64;
  if ((rc = derivative_operators(c, U, o->cfl, blocked ? 0 : o->precision, "direct_solve"))) return rc;
  const int bm = std::min(o->gmres_m, 16);
  // with the Navier-Stokes wall conditions the system is A x = P rhs with
  // A = P dR/dU + B: the wall rows of the right-hand sides become the
  // derivative of the conditions, zero for fixed walls
  // (applyBCsToDerivativeOfResidual)
  const bool walls = c->viscous && c->nwall;
  double *prhs = nullptr;
  struct FreeRhs {
    aerofx_ctx *c;
    double *&p;
    ~FreeRhs() {
      if (p) ctx_release(c, p);
    }
  } free_rhs{c, prhs};
  if (walls) {
    void *p = nullptr;
    const size_t n = 5 * (size_t)c->N;
    if ((rc = ctx_alloc(c, sizeof(double) * n * nrhs, &p))) return rc;
    prhs = (double *)p;
    for (int q = 0; q < nrhs; ++q)
      if ((rc = launch_bcs_adjoint(c, U, rhs + q * n, prhs + q * n, nullptr, true))) return rc;
    rhs = prhs;
  }
  FgOps ops;
  ops.apply = [&](const double *xv, double *y, int nq) -> int {
    for (int q = 0; q < nq; ++q) {
      const size_t off = (size_t)q * c->ksp->n;
      int r = launch_tangent(c, U, xv + off, nullptr, y + off);
      if (r) return r;
      if (walls && (r = launch_bcs_product(c, U, xv + off, y + off, nullptr))) return r;
    }
    return 0;
  };
  ops.prec = [&](const double *v, double *z, int nq, int *inner) -> int {
    // one GMRES per right-hand side, or one block GMRES for all of them
    // (whose batched H1 apply takes nodes of at most 64 neighbours)
    if (nq == 1 || !blocked) {
      for (int q = 0; q < nq; ++q) {
        const size_t off = (size_t)q * c->ksp->n;
        GmresOut go;
        int r = gmres(c, v + off, z + off, o->gmres_m, o->gmres_eps, o->gmres_maxits, &go);
        if (r) return r;
        *inner += go.its;
      }
      return 0;
    }
    aerofx_block_stats bs;
    // as many block steps per right-hand side as the single-vector inner
    // solve may take (gmres_maxits), in cycles of bm
    int r = aerofx_block_gmres(c, nq, v, z, bm, o->gmres_eps, std::max(1, (o->gmres_maxits + bm - 1) / bm),
                               o->precision, 1, &bs);
    if (r < 0) return r;
    *inner += bs.inner_its;
    return 0;
  };
  FgResult fr;
  if ((rc = fgmres_solve(c, nrhs, rhs, x, kDerivRestart, o->eps, o->max_its, ops, &fr))) return rc;
  // a non-physical reconstruction in the exact products (state U) voids x
  hipError_t e = hipMemcpyAsync(c->h_err, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "direct_solve sync");
  if (c->h_err[0]) return set_error(AEROFX_ERR_STATE, "direct_solve: non-physical reconstructed state");
  st->its = fr.outer_its;
  st->krylov_its = fr.inner_its;
  st->converged = fr.converged;
  st->res0 = fr.res0;
  st->res = fr.res;
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return AEROFX_OK;
}

int aerofx_jvp_transpose(aerofx_ctx *c, const double *p, double *JTp) {
  if (!c || !p || !JTp) return set_error(AEROFX_ERR_ARG, "jvp_transpose: null argument");
  int rc = ensure_mirror(c);
  return rc ? rc : launch_h1_apply(c, p, JTp, true);
}

int aerofx_prec_apply_transpose(aerofx_ctx *c, const double *y, double *x) {
  if (!c || !y || !x) return set_error(AEROFX_ERR_ARG, "prec_apply_transpose: null argument");
  return launch_prec_apply(c, y, x, true);
}

int aerofx_gmres_transpose(aerofx_ctx *c, const double *b, double *x, int m, double eps, int maxits, int *its,
                           double *final_res) {
  if (!c || !b || !x) return set_error(AEROFX_ERR_ARG, "gmres_transpose: null argument");
  if (!c->d_jacD) return set_error(AEROFX_ERR_STATE, "gmres_transpose: operators not set up");
  GmresOut o;
  int rc = gmres(c, b, x, m, eps, maxits, &o, true);
  if (rc) return rc;
  if (its) *its = o.its;
  if (final_res) *final_res = o.res;
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? AEROFX_OK : set_hip_error(e, "gmres_transpose sync");
}

int aerofx_timestep(aerofx_ctx *c, const double *U, double cfl, double *dt) {
  if (!c || !U || !dt) return set_error(AEROFX_ERR_ARG, "timestep: null argument");
  int rc = launch_primitive(c, U, c->d_V);
  return rc ? rc : launch_timestep(c, cfl, dt);
}

int aerofx_jacobian_setup(aerofx_ctx *c, const double *U, const double *dt, int precision) {
  if (!c || !U) return set_error(AEROFX_ERR_ARG, "jacobian_setup: null argument");
  if (precision != 0 && precision != 1) return set_error(AEROFX_ERR_ARG, "jacobian_setup: precision 0 or 1");
  int rc = launch_primitive(c, U, c->d_V);
  if (rc) return rc;
  if ((rc = launch_jacobian(c, dt, precision, U))) return rc;
  hipError_t e = hipMemcpyAsync(c->h_err + 1, c->d_err + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "jacobian_setup sync");
  if (c->h_err[1]) return set_error(AEROFX_ERR_STATE, "jacobian_setup: zero pivot in a diagonal block");
  return AEROFX_OK;
}

int aerofx_jacobian_blocks(aerofx_ctx *c, double *D, double *Aedge) {
  if (!c || !c->d_jacD) return set_error(AEROFX_ERR_STATE, "jacobian_blocks: no Jacobian");
  const size_t esz = c->jac_precision ? sizeof(float) : sizeof(double);
  std::vector<char> hd(esz * blk64_count(c->N)), hb(esz * blk64_count(2 * (size_t)c->E));
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(hd.data(), c->d_jacD, hd.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hb.data(), c->d_jacB, hb.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return set_hip_error(e, "jacobian_blocks copy");
  auto at = [&](const std::vector<char> &v, size_t idx) -> double {
    return c->jac_precision ? (double)((const float *)v.data())[idx] : ((const double *)v.data())[idx];
  };
  if (D)
    for (size_t n = 0; n < (size_t)c->N; ++n)
      for (int kk = 0; kk < 25; ++kk) D[25 * n + kk] = at(hd, blk64_base(n) + kk * kBlkStride);
  if (Aedge) {
    // adjacency entry q of node n -> edge layout [E][2][25] = (A_ij, A_ji)
    std::vector<int> rowptr, adj, eid;
    node_csr(c->N, c->E, c->ptr.data(), rowptr, adj, eid);
    for (int n = 0; n < c->N; ++n)
      for (int q = rowptr[n]; q < rowptr[n + 1]; ++q) {
        int l = eid[q];
        int side = n < adj[q] ? 0 : 1;
        for (int kk = 0; kk < 25; ++kk)
          Aedge[50 * (size_t)l + 25 * side + kk] = at(hb, blk64_base(q) + kk * kBlkStride);
      }
  }
  return AEROFX_OK;
}

int aerofx_jvp(aerofx_ctx *c, const double *p, double *Jp) {
  if (!c || !p || !Jp) return set_error(AEROFX_ERR_ARG, "jvp: null argument");
  return launch_h1_apply(c, p, Jp);
}

int aerofx_prec_apply(aerofx_ctx *c, const double *y, double *x) {
  if (!c || !y || !x) return set_error(AEROFX_ERR_ARG, "prec_apply: null argument");
  return launch_prec_apply(c, y, x);
}

int aerofx_gmres(aerofx_ctx *c, const double *b, double *x, int m, double eps, int maxits, int *its,
                 double *final_res) {
  if (!c || !b || !x) return set_error(AEROFX_ERR_ARG, "gmres: null argument");
  if (!c->d_jacD) return set_error(AEROFX_ERR_STATE, "gmres: operators not set up");
  const bool fd = c->mvp_type == AEROFX_MVP_FD;
  if (fd) hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->d_err + 3, 1, nullptr);
  GmresOut o;
  int rc = gmres(c, b, x, m, eps, maxits, &o);
  if (rc) return rc;
  if (its) *its = o.its;
  if (final_res) *final_res = o.res;
  hipError_t e = fd ? hipMemcpyAsync(c->h_err + 3, c->d_err + 3, sizeof(int), hipMemcpyDeviceToHost, c->stream)
                    : hipSuccess;
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "gmres sync");
  return fd ? c->h_err[3] : AEROFX_OK;
}

int aerofx_steady_solve(aerofx_ctx *c, double *U, const aerofx_newton_opts *o, aerofx_newton_stats *st,
                        double *history, int hist_cap) {
  if (!c || !U || !o || !st) return set_error(AEROFX_ERR_ARG, "steady_solve: null argument");
  if (o->precision != 0 && o->precision != 1) return set_error(AEROFX_ERR_ARG, "steady_solve: precision 0 or 1");
  if (o->mvp != AEROFX_MVP_H1 && o->mvp != AEROFX_MVP_FD && o->mvp != AEROFX_MVP_H2)
    return set_error(AEROFX_ERR_ARG, "steady_solve: mvp AEROFX_MVP_H1, AEROFX_MVP_FD or AEROFX_MVP_H2");
  if (o->mvp == AEROFX_MVP_H2 && c->fc.low_mach)
    return set_error(AEROFX_ERR_STATE, "steady_solve: exact products not available with low-Mach preconditioning");
  if (o->mvp == AEROFX_MVP_FD && o->fd_order != 1 && o->fd_order != 2)
    return set_error(AEROFX_ERR_ARG, "steady_solve: fd_order 1 or 2");
  if (o->prec != AEROFX_PREC_JACOBI && o->prec != AEROFX_PREC_RAS_ILU0)
    return set_error(AEROFX_ERR_ARG, "steady_solve: prec AEROFX_PREC_JACOBI or AEROFX_PREC_RAS_ILU0");
  if (o->prec == AEROFX_PREC_RAS_ILU0 && (o->prec_subdomain < 1 || o->prec_subdomain > 1024))
    return set_error(AEROFX_ERR_ARG, "steady_solve: prec_subdomain in [1, 1024]");
  int rc = ensure_krylov(c, o->gmres_m);
  if (rc) return rc;
  Krylov *k = c->ksp;
  const size_t n = k->n;
  // the caller's operator selection is restored on return
  struct Restore {
    aerofx_ctx *c;
    int type, order;
    const double *Q, *F0, *dt;
    ~Restore() {
      c->mvp_type = type;
      c->mvp_order = order;
      c->mvp_Q = Q;
      c->mvp_F0 = F0;
      c->mvp_dt = dt;
    }
  } restore{c, c->mvp_type, c->mvp_order, c->mvp_Q, c->mvp_F0, c->mvp_dt};
  c->mvp_type = AEROFX_MVP_H1;
  struct RestorePrec {
    aerofx_ctx *c;
    int type, sub;
    ~RestorePrec() {
      c->prec_type = type;
      c->prec_subdomain = sub;
    }
  } restore_prec{c, c->prec_type, c->prec_subdomain};
  c->prec_type = o->prec;
  if (o->prec == AEROFX_PREC_RAS_ILU0) c->prec_subdomain = o->prec_subdomain;
  // Newton.FailSafe and checkFinalRes for the solve (the caller's restored)
  if (o->failsafe < 0 || o->failsafe > 2) return set_error(AEROFX_ERR_ARG, "steady_solve: failsafe 0, 1 or 2");
  if (o->eps_formula != 0 && o->eps_formula != 1) return set_error(AEROFX_ERR_ARG, "steady_solve: eps_formula 0 or 1");
  struct RestoreChecks {
    aerofx_ctx *c;
    int failsafe, check;
    ~RestoreChecks() {
      (void)aerofx_ctx_set_failsafe(c, failsafe);
      c->check_final_res = check;
    }
  } restore_checks{c, c->failsafe, c->check_final_res};
  if ((rc = aerofx_ctx_set_failsafe(c, o->failsafe))) return rc;
  c->check_final_res = o->check_final_res ? 1 : 0;
  int *fs_tmp = nullptr;  // [N] state tags of the Newton failsafe
  if (c->failsafe) {
    void *p = nullptr;
    if ((rc = ctx_alloc(c, sizeof(int) * (size_t)c->N, &p))) return rc;
    fs_tmp = static_cast<int *>(p);
  }
  struct FreeTmp {
    aerofx_ctx *c;
    int *p;
    ~FreeTmp() {
      if (p) ctx_release(c, p);
    }
  } free_tmp{c, fs_tmp};
  const ResidualBufs rbuf{c->d_V, c->d_grad, c->d_Fm, c->d_P};
  auto t0 = std::chrono::steady_clock::now();
  *st = aerofx_newton_stats{};
  double res0 = 0.0, res = 0.0, cfl = o->cfl0;
  // Eisenstat (KspConvCriterion::computeEpsilonEisenstadt): one Newton
  // iteration per pseudo-time step, so its history runs over the steps
  int ew_its = 0;
  double ew_fprev = 0.0, ew_eprev = 0.0;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->d_err, 4, nullptr);
  int step;
  for (step = 0; step <= o->max_steps; ++step) {
    // spatial residual and its norm (TsDesc::computeResidualNorm); one
    // synchronisation per step reads the norm and the error counters of the
    // previous step's block inversion [1], update [2] and finite-difference
    // products [3]
    if (c->failsafe) {
      int cnt = 0;
      if ((rc = residual_failsafe(c, rbuf, U, k->R, true, &cnt))) return rc;
    } else if ((rc = launch_residual(c, U, k->R, true))) {
      return rc;
    }
    // the wall conditions of the Navier-Stokes equations (ImplicitTsDesc.C:140)
    if ((rc = launch_bcs_residual(c, U, k->R, nullptr))) return rc;
    norm2(c, k->R, k->scal);
    hipError_t e = hipMemcpyAsync(k->h_scal, k->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->h_err, c->d_err, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_hip_error(e, "steady_solve sync");
    if (step > 0 && c->h_err[1]) {
      st->steps = step;
      return set_error(AEROFX_ERR_STATE, "steady_solve: zero pivot in a diagonal block");
    }
    if (step > 0 && c->h_err[3]) {
      st->steps = step;
      return set_error(AEROFX_ERR_STATE, "steady_solve: non-physical reconstruction in a finite-difference product");
    }
    if (step > 0 && c->h_err[2]) {
      st->steps = step;
      return set_error(AEROFX_ERR_STATE, "steady_solve: update produced non-physical states");
    }
    if (c->h_err[0]) {
      st->steps = step;
      return set_error(AEROFX_ERR_STATE, "steady_solve: non-physical reconstructed state");
    }
    res = std::sqrt(k->h_scal[0]);
    if (step == 0) res0 = res;
    if (history && step < hist_cap) history[step] = res;
    if (res == 0.0 || res <= o->eps * res0) {
      st->converged = 1;
      break;
    }
    if (step == o->max_steps) break;
    // CflLaw Old (TsParameters.C:198-200) with its = step
    double rr = res / res0;
    cfl = std::min(std::max(std::max(o->cfl_coef1, std::min(step * o->cfl_coef2, cfl + o->cfl_coef2)),
                            o->cfl0 / std::pow(rr, o->ser)),
                   o->cfl_max);
    if ((rc = launch_timestep(c, cfl, c->d_dt))) return rc;
    if ((rc = launch_jacobian(c, c->d_dt, o->precision, U))) return rc;
    // the Krylov tolerance of this step, from ||b|| = ||R||
    double geps = o->gmres_eps;
    if (o->eps_formula == 1) {
      if (ew_its > 0) {
        const double eg = 0.1 * std::pow(res / ew_fprev, 2.0), e1 = 0.1 * std::pow(ew_eprev, 2.0);
        geps = e1 < 0.1 ? std::min(o->gmres_eps, eg) : std::min(o->gmres_eps, std::max(eg, e1));
      }
      ew_fprev = res;
      ew_eprev = geps;
      ++ew_its;
    }
    for (int redo = 0;; ++redo) {
      if (o->mvp == AEROFX_MVP_FD) {
        // MatVecProdFD::evaluate at the Newton state: F(U) = R(U) (the time
        // term vanishes at Q = Un); one residual per Krylov product
        if ((rc = select_fd(c, U, k->R, c->d_dt, o->fd_order))) return rc;
      } else if (o->mvp == AEROFX_MVP_H2) {
        // MatVecProdH2::evaluate at the Newton state: exact products of the
        // second-order residual plus the time term
        if ((rc = select_h2(c, U, c->d_dt))) return rc;
      }
      // b = -F
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, -1.0, nullptr, k->R, 0.0, k->b,
                         nullptr);
      GmresOut go;
      if ((rc = gmres(c, k->b, k->x, o->gmres_m, geps, o->gmres_maxits, &go))) return rc;
      if (go.its == -999) {
        ++st->final_res_failures;
        go.its = k->h_st->iter;
      }
      st->krylov_its += go.its;
      if (go.its == o->gmres_maxits) ++st->saturated;
      // Q += dQ (the failsafe keeps Q for a redo); the state check is read
      // at the next step's synchronisation
      if (c->failsafe)
        hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, U, 0.0, k->z, nullptr);
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, k->x, 1.0, U, nullptr);
      hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->d_err + 2, 1, nullptr);
      hipLaunchKernelGGL(k_check_states, dim3((c->N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->N, U,
                         c->gas, c->d_err + 2);
      if (!c->failsafe) break;
      e = hipMemcpyAsync(c->h_err + 2, c->d_err + 2, sizeof(int), hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) return set_hip_error(e, "steady_solve failsafe sync");
      if (c->h_err[2] == 0 || redo == 5) {
        // NewtonSolver.h:216: after a redo, FailSafe = On resets the fixes
        if (redo > 0 && c->failsafe == 1 && (rc = aerofx_ctx_reset_failsafe(c))) return rc;
        break;
      }
      // NewtonSolver.h:193-205: tag the neighbours of the non-physical
      // nodes first order, restore Q and redo the iteration
      const int nb = (c->N + kBlock - 1) / kBlock;
      hipLaunchKernelGGL(k_fs_state_tags, dim3(nb), dim3(kBlock), 0, c->stream, c->N, U, c->gas, fs_tmp);
      hipLaunchKernelGGL(k_fs_state_fix, dim3(nb), dim3(kBlock), 0, c->stream, c->tab, fs_tmp,
                         c->d_fs + 2 * (size_t)c->N);
      c->fs_any = true;
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, k->z, 0.0, U, nullptr);
      ++st->failsafe_redos;
      int cnt = 0;
      if ((rc = residual_failsafe(c, rbuf, U, k->R, true, &cnt))) return rc;
      if ((rc = launch_bcs_residual(c, U, k->R, nullptr))) return rc;
    }
    st->final_cfl = cfl;
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "steady_solve sync");
  st->steps = step;
  st->res0 = res0;
  st->res = res;
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return AEROFX_OK;
}

}  // extern "C"
