// residual.hip -- the FP64 edge residual on gfx950.
//
// SpaceOperator<5>::computeResidual (SpaceOperator.C:524-634) in three
// kernels, all deterministic (no atomics on floating-point data):
//
//  1. k_gradients  one lane per node n walks its neighbours m in ascending
//     order (= ascending edge id) accumulating W(R_n, X_m - X_n) (V_m - V_n):
//     the reference's LSQ edge loop (SubDomain.C:137-159, 297-374) with the
//     same per-node addition order.  The pass also writes V = V(U).
//  2. k_flux_edges edge-parallel, edges grouped by their lower node i (the
//     reference's edge order, EdgeCore.C:92-119): each edge flux is computed
//     once (MUSCL + Van Albada + ROEFLUX7, Edge.C:372-439); the block sums the
//     +f contributions of each of its rows in edge order (P_i) and writes f
//     contiguously in edge order (Fm).
//  3. k_flux_nodes node-parallel: R_n = (sum over lower edges (m, n), m
//     ascending, of -f, gathered from Fm) + P_n + boundary-face fluxes in face order (Face.C:356-389), times
//     1/V_n (SpaceOperator.C:599-624).
//
// Node data are gathered through L2 (the RCM neighbour window is small).  blockIdx is remapped so that each XCD sweeps one
// contiguous node range, keeping the RCM neighbour window in its L2.
#include <hip/hip_runtime.h>

#include "context.h"
#include "host_internal.h"

namespace aerofx {

// neighbours per load batch of the gradient pass (measured at 128^3: 2 ->
// 0.317 ms, 4 -> 0.310, 6 -> 0.307, 8 -> 0.301; from 10 on the batch no
// longer stays in registers)
constexpr int kGradChunk = 8;

// ---------------------------------------------------------------------------
// 1. gradients (+ primitive state): one lane per node, neighbours in
//    ascending order (= the reference's edge order), so every node's sums are
//    formed exactly as in the reference's sequential edge loop.
// ---------------------------------------------------------------------------
template <bool FROM_U>
__device__ __forceinline__ void load_prim(const double *__restrict__ W, int n, const Gas &gas, double v[5]) {
  double w[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) w[k] = W[5 * (size_t)n + k];
  if (FROM_U) {
    prim_from_cons(w, gas, v);
  } else {
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] = w[k];
  }
}

template <bool FROM_U, bool OUT3>
__global__ __launch_bounds__(kBlock) void k_gradients(DevTables t, const double *__restrict__ W, Gas gas,
                                                      double *__restrict__ Vout, double *__restrict__ grad,
                                                      double *__restrict__ ddx, double *__restrict__ ddy,
                                                      double *__restrict__ ddz, int *__restrict__ err,
                                                      const int *__restrict__ done) {
  if (done && *done) return;
  const int nb = (t.N + kBlock - 1) / kBlock;
  const int n = xcd_chunk(blockIdx.x, nb) * kBlock + (int)threadIdx.x;
  if (err && blockIdx.x == 0 && threadIdx.x == 0) *err = 0;  // the flux pass follows in stream order
  if (n >= t.N) return;
  double Vn[5];
  load_prim<FROM_U>(W, n, gas, Vn);
  if (Vout) {
#pragma unroll
    for (int k = 0; k < 5; ++k) Vout[5 * (size_t)n + k] = Vn[k];
  }
  const double4 xn = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)n);
  const double *L = t.lsq6 + 6 * (size_t)n;
  const double or11 = L[0], or22 = L[1], or33 = L[2], r12or11 = L[3], r23or22 = L[4], psi = L[5];
  double acc[15];
#pragma unroll
  for (int q = 0; q < 15; ++q) acc[q] = 0.0;
  // Neighbours in chunks of kGradChunk: the chunk's states and coordinates
  // are requested together and the next chunk's indices are prefetched, so a
  // row costs about one memory latency per chunk.  The sums are still formed
  // one neighbour at a time in ascending order.
  const int e0 = t.rowptr[n], e1 = t.rowptr[n + 1];
  int mi[kGradChunk];
#pragma unroll
  for (int k = 0; k < kGradChunk; ++k) mi[k] = e0 + k < e1 ? t.adj[e0 + k] : n;
  for (int e = e0; e < e1; e += kGradChunk) {
    double Wm[kGradChunk][5];
    double4 xm[kGradChunk];
#pragma unroll
    for (int k = 0; k < kGradChunk; ++k) {
#pragma unroll
      for (int q = 0; q < 5; ++q) Wm[k][q] = W[5 * (size_t)mi[k] + q];
      xm[k] = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)mi[k]);
    }
#pragma unroll
    for (int k = 0; k < kGradChunk; ++k) mi[k] = e + kGradChunk + k < e1 ? t.adj[e + kGradChunk + k] : n;
#pragma unroll
    for (int k = 0; k < kGradChunk; ++k) {
      if (e + k >= e1) break;
      double Vm[5];
      if (FROM_U) {
        prim_from_cons(Wm[k], gas, Vm);
      } else {
#pragma unroll
        for (int q = 0; q < 5; ++q) Vm[q] = Wm[k][q];
      }
      const double dx0 = xm[k].x - xn.x, dx1 = xm[k].y - xn.y, dx2 = xm[k].z - xn.z;
      const double alpha1 = dx0 * or11 * or11;
      const double alpha2 = (dx1 - r12or11 * dx0) * or22 * or22;
      const double alpha3 = (dx2 - r23or22 * dx1 + psi * dx0) * or33 * or33;
      const double w0 = alpha1 - r12or11 * alpha2 + psi * alpha3;
      const double w1 = alpha2 - r23or22 * alpha3;
      const double w2 = alpha3;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const double d = Vm[q] - Vn[q];
        acc[q] += w0 * d;
        acc[5 + q] += w1 * d;
        acc[10 + q] += w2 * d;
      }
    }
  }
  if (!OUT3 && t.fs_mask && t.fs_mask[n]) {
    // DistNodalGrad::limit with failsafe tags (DistNodalGrad.C:1083-1104)
#pragma unroll
    for (int q = 0; q < 15; ++q) acc[q] = 0.0;
  }
  if (!OUT3) {
    double2 *g2 = reinterpret_cast<double2 *>(grad + 16 * (size_t)n);
#pragma unroll
    for (int q = 0; q < 7; ++q) g2[q] = make_double2(acc[2 * q], acc[2 * q + 1]);
    g2[7] = make_double2(acc[14], 0.0);
  } else {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      ddx[5 * (size_t)n + k] = acc[k];
      ddy[5 * (size_t)n + k] = acc[5 + k];
      ddz[5 * (size_t)n + k] = acc[10 + k];
    }
  }
}

// ---------------------------------------------------------------------------
// 2. edge fluxes, computed once per edge.  A block covers whole rows (all
//    edges whose lower node lies in [r0, r1), <= FB edges); lane k of the
//    block takes edge l0 + k.  Node i data are read directly (the ~7 lanes of
//    a row hit the same lines); only the fluxes go through LDS, for the
//    per-row sums P_i in edge order.
// ---------------------------------------------------------------------------
//    LM: the spatial low-Mach flux (ROEFLUX5) with the edge length
//    |X_j - X_i| in Turkel's beta.
template <bool SECOND, int FB, bool LM>
__global__ __launch_bounds__(FB) void k_flux_edges(DevTables t, const double *__restrict__ V,
                                                       const double *__restrict__ grad, FluxConsts fc,
                                                       double *__restrict__ Fm, double *__restrict__ P,
                                                       int *__restrict__ err, int *__restrict__ tag0,
                                                       const int *__restrict__ done) {
  if (done && *done) return;
  __shared__ double sf[5 * (FB + 1)];
  const int4 ch = t.fchunk4[xcd_chunk(blockIdx.x, t.nfchunks)];
  const int r0 = ch.x, r1 = ch.y, l0 = ch.z, l1 = ch.w;
  const int tid = threadIdx.x;
  const Gas &g = fc.gas;
  if (tid < l1 - l0) {
    const int l = l0 + tid;
    const int2 ij = t.eij[l];
    const int i = ij.x, j = ij.y;
    double Vi[5], Vj[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      Vi[k] = V[5 * (size_t)i + k];
      Vj[k] = V[5 * (size_t)j + k];
    }
    double Vij[5], Vji[5];
    double length = 1.0;
    if (LM) {
      const double4 xi = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)i);
      const double4 xj = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)j);
      const double dx0 = xj.x - xi.x, dx1 = xj.y - xi.y, dx2 = xj.z - xi.z;
      length = sqrt(dx0 * dx0 + dx1 * dx1 + dx2 * dx2);
    }
    if (SECOND) {
      const double4 xi = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)i);
      const double4 xj = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)j);
      const double dx0 = xj.x - xi.x, dx1 = xj.y - xi.y, dx2 = xj.z - xi.z;
      const double2 *gi2 = reinterpret_cast<const double2 *>(grad + 16 * (size_t)i);
      const double2 *gj2 = reinterpret_cast<const double2 *>(grad + 16 * (size_t)j);
      double gi[16], gj[16];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const double2 a = gi2[q], b = gj2[q];
        gi[2 * q] = a.x;
        gi[2 * q + 1] = a.y;
        gj[2 * q] = b.x;
        gj[2 * q + 1] = b.y;
      }
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const double ddVij = dx0 * gi[k] + dx1 * gi[5 + k] + dx2 * gi[10 + k];
        const double ddVji = dx0 * gj[k] + dx1 * gj[5 + k] + dx2 * gj[10 + k];
        vanalbada(fc.beta, Vi[k], ddVij, Vj[k], ddVji, Vij[k], Vji[k]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        Vij[k] = Vi[k];
        Vji[k] = Vj[k];
      }
    }
    // checkReconstructedValues (EdgeCore.C:137-166, VarFcnSGEuler.h:90-110);
    // with the failsafe the failing sides are tagged
    const bool bi = Vij[0] <= 0.0 || Vij[4] + g.pstiff <= 0.0, bj = Vji[0] <= 0.0 || Vji[4] + g.pstiff <= 0.0;
    if (bi || bj) {
      atomicAdd(err, (int)bi + (int)bj);
      if (tag0) {
        if (bi) tag0[i] = 1;
        if (bj) tag0[j] = 1;
      }
    }
    const double nv[3] = {t.normal[3 * (size_t)l], t.normal[3 * (size_t)l + 1], t.normal[3 * (size_t)l + 2]};
    double f[5];
    if (sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]) < 1e-18) {
#pragma unroll
      for (int k = 0; k < 5; ++k) f[k] = 0.0;
    } else if (LM) {
      roe_turkel_flux(g, fc.lm, true, nv, length, Vij, Vji, f);
    } else {
      roeflux7(g, nv, 0.0, Vij, Vji, f);
    }
    double *dst = Fm + 5 * (size_t)l;  // contiguous in edge order
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      sf[k * (FB + 1) + tid] = f[k];
      dst[k] = f[k];
    }
  }
  __syncthreads();
  const int nr = r1 - r0;
  for (int idx = tid; idx < nr * 5; idx += FB) {
    const int rr = idx / 5, q = idx - 5 * rr;
    const int r = r0 + rr;
    const int b = t.eup[r] - l0, eend = t.eup[r + 1] - l0;
    double s = 0.0;
    for (int e = b; e < eend; ++e) s += sf[q * (FB + 1) + e];
    P[5 * (size_t)r + q] = s;
  }
}

// ---------------------------------------------------------------------------
// 1'. gradients with the strip's node data staged in LDS.  A block covers a
//     strip of consecutive nodes [n0, n1) (one lane each); every node the
//     strip touches (the strip and all neighbours, ascending ids) is loaded
//     once: consecutive lanes take consecutive staged nodes, so the loads
//     coalesce, and U -> V is converted once per staged node instead of once
//     per adjacency entry.  The sums are the per-lane kernel's, term for term.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kGradStrip) void k_gradients_staged(DevTables t, const double *__restrict__ U, Gas gas,
                                                                 double *__restrict__ Vout,
                                                                 double *__restrict__ grad, int *__restrict__ err,
                                                                 const int *__restrict__ done) {
  if (done && *done) return;
  __shared__ double sv[8 * kGradHalo];  // field f of staged node h at sv[f * kGradHalo + h]: V (5), X (3)
  const int b = xcd_chunk(blockIdx.x, t.ngb);
  const int4 s = t.gb4[b];
  const int tid = threadIdx.x;
  if (err && blockIdx.x == 0 && tid == 0) *err = 0;  // the flux pass follows in stream order
  const int H = s.w - s.z;
  // every staged node's loads are issued before any is used: one memory
  // latency per block instead of one per staged node
  constexpr int PER = (kGradHalo + kGradStrip - 1) / kGradStrip;
  double u[PER][5];
  double4 x[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int h = tid + r * kGradStrip;
    const int g = h < H ? t.ghalo[s.z + h] : t.ghalo[s.z];
#pragma unroll
    for (int k = 0; k < 5; ++k) u[r][k] = U[5 * (size_t)g + k];
    x[r] = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)g);
  }
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int h = tid + r * kGradStrip;
    if (h < H) {
      double v[5];
      prim_from_cons(u[r], gas, v);
#pragma unroll
      for (int k = 0; k < 5; ++k) sv[k * kGradHalo + h] = v[k];
      sv[5 * kGradHalo + h] = x[r].x;
      sv[6 * kGradHalo + h] = x[r].y;
      sv[7 * kGradHalo + h] = x[r].z;
    }
  }
  __syncthreads();
  const int n = s.x + tid;
  if (n >= s.y) return;
  const int hn = t.gown[b] + tid;
  double Vn[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    Vn[k] = sv[k * kGradHalo + hn];
    Vout[5 * (size_t)n + k] = Vn[k];
  }
  const double xn0 = sv[5 * kGradHalo + hn], xn1 = sv[6 * kGradHalo + hn], xn2 = sv[7 * kGradHalo + hn];
  const double *L = t.lsq6 + 6 * (size_t)n;
  const double or11 = L[0], or22 = L[1], or33 = L[2], r12or11 = L[3], r23or22 = L[4], psi = L[5];
  double acc[15];
#pragma unroll
  for (int q = 0; q < 15; ++q) acc[q] = 0.0;
  const int e0 = t.rowptr[n], e1 = t.rowptr[n + 1];
  for (int e = e0; e < e1; ++e) {
    const int hm = t.gloc[e];
    const double dx0 = sv[5 * kGradHalo + hm] - xn0, dx1 = sv[6 * kGradHalo + hm] - xn1,
                 dx2 = sv[7 * kGradHalo + hm] - xn2;
    const double alpha1 = dx0 * or11 * or11;
    const double alpha2 = (dx1 - r12or11 * dx0) * or22 * or22;
    const double alpha3 = (dx2 - r23or22 * dx1 + psi * dx0) * or33 * or33;
    const double w0 = alpha1 - r12or11 * alpha2 + psi * alpha3;
    const double w1 = alpha2 - r23or22 * alpha3;
    const double w2 = alpha3;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const double d = sv[q * kGradHalo + hm] - Vn[q];
      acc[q] += w0 * d;
      acc[5 + q] += w1 * d;
      acc[10 + q] += w2 * d;
    }
  }
  double2 *g2 = reinterpret_cast<double2 *>(grad + 16 * (size_t)n);
#pragma unroll
  for (int q = 0; q < 7; ++q) g2[q] = make_double2(acc[2 * q], acc[2 * q + 1]);
  g2[7] = make_double2(acc[14], 0.0);
}

// ---------------------------------------------------------------------------
// 2'. edge fluxes with the strip's node data staged in LDS.  A block covers
//     the edges of consecutive rows (lower nodes) [r0, r1); the rows and the
//     upper ends of their edges (ascending ids) are loaded once, field by
//     field with consecutive lanes on consecutive doubles of a record
//     (V: 5, X: 3 of 4, gradients: 15 of 16), so the loads coalesce.  Lane k
//     computes edge l0 + k from LDS exactly as the per-lane kernel does; the
//     fluxes go back through the same LDS for coalesced Fm stores and the
//     per-row sums P.
// ---------------------------------------------------------------------------
template <bool SECOND>
__global__ __launch_bounds__(kFluxStrip) void k_flux_staged(DevTables t, const double *__restrict__ V,
                                                            const double *__restrict__ grad, FluxConsts fc,
                                                            double *__restrict__ Fm, double *__restrict__ P,
                                                            int *__restrict__ err, const int *__restrict__ done) {
  if (done && *done) return;
  // staged node h: record sh[h * NF ...] = V (5), X (3), gradients (15); the
  // odd record length spreads the lanes' reads of random records over the
  // LDS banks, and consecutive lanes staging one record write consecutive
  // addresses
  constexpr int NF = SECOND ? 23 : 5;
  constexpr int HS = kFluxHalo;
  constexpr int FS = kFluxStrip + 1;
  // the flux staging [5][FS] reuses the node staging [HS][NF]
  __shared__ double sh[NF * HS > 5 * FS ? NF * HS : 5 * FS];
  __shared__ int sidx[HS];
  const int b = xcd_chunk(blockIdx.x, t.nfb);
  const int4 ch = t.fb4[b];
  const int2 hr = t.fbh[b];
  const int r0 = ch.x, r1 = ch.y, l0 = ch.z, l1 = ch.w;
  const int H = hr.y - hr.x;
  const int tid = threadIdx.x;
  const Gas &g = fc.gas;
  for (int h = tid; h < H; h += kFluxStrip) sidx[h] = t.fhalo[hr.x + h];
  __syncthreads();
  // all of a lane's staging loads are in flight together: consecutive lanes
  // take consecutive doubles of the records (V: 5, X: 3 of 4, gradients: 15
  // of 16), so the loads coalesce, and a block waits one memory latency
  constexpr int PV = (5 * HS + kFluxStrip - 1) / kFluxStrip;
  constexpr int PX = (4 * HS + kFluxStrip - 1) / kFluxStrip;
  constexpr int PG = (16 * HS + kFluxStrip - 1) / kFluxStrip;
  double rv[PV], rx[SECOND ? PX : 1], rg[SECOND ? PG : 1];
#pragma unroll
  for (int r = 0; r < PV; ++r) {
    const int q = tid + r * kFluxStrip;
    const int h = q / 5, f = q - 5 * h;
    rv[r] = q < 5 * H ? V[5 * (size_t)sidx[min(h, H - 1)] + f] : 0.0;
  }
  if (SECOND) {
#pragma unroll
    for (int r = 0; r < PX; ++r) {
      const int q = tid + r * kFluxStrip;
      rx[r] = q < 4 * H ? t.X4[4 * (size_t)sidx[min(q >> 2, H - 1)] + (q & 3)] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < PG; ++r) {
      const int q = tid + r * kFluxStrip;
      rg[r] = q < 16 * H ? grad[16 * (size_t)sidx[min(q >> 4, H - 1)] + (q & 15)] : 0.0;
    }
  }
#pragma unroll
  for (int r = 0; r < PV; ++r) {
    const int q = tid + r * kFluxStrip;
    const int h = q / 5, f = q - 5 * h;
    if (q < 5 * H) sh[h * NF + f] = rv[r];
  }
  if (SECOND) {
#pragma unroll
    for (int r = 0; r < PX; ++r) {
      const int q = tid + r * kFluxStrip;
      if (q < 4 * H && (q & 3) < 3) sh[(q >> 2) * NF + 5 + (q & 3)] = rx[r];
    }
#pragma unroll
    for (int r = 0; r < PG; ++r) {
      const int q = tid + r * kFluxStrip;
      if (q < 16 * H && (q & 15) < 15) sh[(q >> 4) * NF + 8 + (q & 15)] = rg[r];
    }
  }
  __syncthreads();
  const int ne = l1 - l0;
  double f[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (tid < ne) {
    const int l = l0 + tid;
    const uint32_t loc = t.floc[l];
    const double *si = sh + (int)(loc & 0xffffu) * NF, *sj = sh + (int)(loc >> 16) * NF;
    double Vi[5], Vj[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      Vi[k] = si[k];
      Vj[k] = sj[k];
    }
    double Vij[5], Vji[5];
    if (SECOND) {
      const double dx0 = sj[5] - si[5], dx1 = sj[6] - si[6], dx2 = sj[7] - si[7];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const double ddVij = dx0 * si[8 + k] + dx1 * si[13 + k] + dx2 * si[18 + k];
        const double ddVji = dx0 * sj[8 + k] + dx1 * sj[13 + k] + dx2 * sj[18 + k];
        vanalbada(fc.beta, Vi[k], ddVij, Vj[k], ddVji, Vij[k], Vji[k]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        Vij[k] = Vi[k];
        Vji[k] = Vj[k];
      }
    }
    // checkReconstructedValues (EdgeCore.C:137-166, VarFcnSGEuler.h:90-110)
    const int bad = (Vij[0] <= 0.0 || Vij[4] + g.pstiff <= 0.0) + (Vji[0] <= 0.0 || Vji[4] + g.pstiff <= 0.0);
    if (bad) atomicAdd(err, bad);
    const double nv[3] = {t.normal[3 * (size_t)l], t.normal[3 * (size_t)l + 1], t.normal[3 * (size_t)l + 2]};
    if (sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]) >= 1e-18) roeflux7(g, nv, 0.0, Vij, Vji, f);
  }
  __syncthreads();  // the node staging is dead: reuse it for the fluxes
  double *sf = sh;
#pragma unroll
  for (int k = 0; k < 5; ++k) sf[k * FS + tid] = f[k];
  __syncthreads();
  for (int q = tid; q < 5 * ne; q += kFluxStrip) {
    const int e = q / 5, k = q - 5 * e;
    Fm[5 * (size_t)l0 + q] = sf[k * FS + e];
  }
  const int nr = r1 - r0;
  for (int idx = tid; idx < nr * 5; idx += kFluxStrip) {
    const int rr = idx / 5, q = idx - 5 * rr;
    const int r = r0 + rr;
    const int eb = t.eup[r] - l0, eend = t.eup[r + 1] - l0;
    double s = 0.0;
    for (int e = eb; e < eend; ++e) s += sf[q * FS + e];
    P[5 * (size_t)r + q] = s;
  }
}

// ---------------------------------------------------------------------------
// 3. node assembly: lower edges, upper-edge partial sum (formed here from the
//    edge fluxes for a row of more than fblock upper edges), faces, 1/V
// ---------------------------------------------------------------------------
// vrow / vc (or null): the Navier-Stokes Galerkin contributions of the
// node's elements, one contiguous run per node in element order, added
// before the 1/V scaling
__global__ __launch_bounds__(kBlock) void k_flux_nodes(DevTables t, const double *__restrict__ V,
                                                       const double *__restrict__ Fm,
                                                       const double *__restrict__ P, FluxConsts fc, int scale,
                                                       const int *__restrict__ vrow, const double *__restrict__ vc,
                                                       double *__restrict__ R, const int *__restrict__ done) {
  if (done && *done) return;
  const int nb = (t.N + kBlock - 1) / kBlock;
  const int n = xcd_chunk(blockIdx.x, nb) * kBlock + (int)threadIdx.x;
  if (n >= t.N) return;
  double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const int q0 = t.lrow[n], q1 = t.lrow[n + 1];
#pragma unroll 4
  for (int q = q0; q < q1; ++q) {
    const size_t base = 5 * (size_t)t.ledge[q];
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] -= Fm[base + k];
  }
  const int u0 = t.eup[n], u1 = t.eup[n + 1];
  if (u1 - u0 > t.fblock) {
    // a row cut over several flux blocks: its upper-edge sum, in edge order
    double p[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int l = u0; l < u1; ++l)
#pragma unroll
      for (int k = 0; k < 5; ++k) p[k] += Fm[5 * (size_t)l + k];
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] += p[k];
  } else {
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] += P[5 * (size_t)n + k];
  }
  const int f0 = t.frow[n], f1 = t.frow[n + 1];
  if (f0 < f1) {
    double Vn[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Vn[k] = V[5 * (size_t)n + k];
    for (int q = f0; q < f1; ++q) {
      const int face = t.fidx[q];
      const int code = t.fcode[face];
      const double nv[3] = {t.fnorm3[3 * (size_t)face], t.fnorm3[3 * (size_t)face + 1],
                            t.fnorm3[3 * (size_t)face + 2]};
      double fl[5];
      if (is_wall_code(code)) {
        wallflux(nv, 0.0, Vn, fl);
      } else {
        const double *ub = (code == AEROFX_BC_INLET || code == -AEROFX_BC_INLET) ? fc.ub_in : fc.ub_out;
        boundflux5(fc.gas, nv, 0.0, Vn, ub, fl);
      }
#pragma unroll
      for (int k = 0; k < 5; ++k) s[k] += fl[k];
    }
  }
  if (vc) {
    double g[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = vrow[n]; q < vrow[n + 1]; ++q)
#pragma unroll
      for (int k = 0; k < 5; ++k) g[k] += vc[5 * (size_t)q + k];
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] += g[k];
  }
  const double ic = t.invcv[n];
#pragma unroll
  for (int k = 0; k < 5; ++k) R[5 * (size_t)n + k] = scale ? s[k] * ic : s[k];
}

__global__ void k_primitive(int N, const double *__restrict__ U, Gas g, double *__restrict__ V,
                            const int *__restrict__ done) {
  if (done && *done) return;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double u[5], v[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) u[k] = U[5 * (size_t)n + k];
  prim_from_cons(u, g, v);
#pragma unroll
  for (int k = 0; k < 5; ++k) V[5 * (size_t)n + k] = v[k];
}

__global__ void k_zero_int(int *p, int n, const int *done) {
  if (done && *done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <bool SECOND, int FB>
static void launch_flux_edges(aerofx_ctx *c, const ResidualBufs &b, const int *done) {
  const DevTables &t = c->tab;
  if (c->fc.low_mach)
    hipLaunchKernelGGL((k_flux_edges<SECOND, FB, true>), dim3(t.nfchunks), dim3(FB), 0, c->stream, t, b.V, b.grad,
                       c->fc, b.Fm, b.P, c->d_err, b.tag0, done);
  else
    hipLaunchKernelGGL((k_flux_edges<SECOND, FB, false>), dim3(t.nfchunks), dim3(FB), 0, c->stream, t, b.V, b.grad,
                       c->fc, b.Fm, b.P, c->d_err, b.tag0, done);
}

int launch_residual_with(aerofx_ctx *c, const ResidualBufs &b, const double *U, double *R, bool scale,
                         const int *done) {
  const DevTables &t = c->tab;
  // the staged kernels evaluate ROEFLUX7 only, without the failsafe
  const bool staged = t.ngb > 0 && c->residual_kernels == AEROFX_RESIDUAL_STAGED && !c->fc.low_mach && !c->failsafe;
  if (c->p.second_order) {
    ProfScope ps(c, AEROFX_TIMER_NODAL_GRAD);
    if (staged)
      hipLaunchKernelGGL(k_gradients_staged, dim3(t.ngb), dim3(kGradStrip), 0, c->stream, t, U, c->gas, b.V,
                         b.grad, c->d_err, done);
    else
      hipLaunchKernelGGL((k_gradients<true, false>), dim3((t.N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream,
                         t, U, c->gas, b.V, b.grad, nullptr, nullptr, nullptr, c->d_err, done);
  } else {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->d_err, 1, done);
    hipLaunchKernelGGL(k_primitive, dim3((t.N + 255) / 256), dim3(256), 0, c->stream, t.N, U, c->gas, b.V, done);
  }
  {
    ProfScope ps(c, AEROFX_TIMER_FV_TERM);
    if (staged && c->p.second_order)
      hipLaunchKernelGGL(k_flux_staged<true>, dim3(t.nfb), dim3(kFluxStrip), 0, c->stream, t, b.V, b.grad,
                         c->fc, b.Fm, b.P, c->d_err, done);
    else if (staged)
      hipLaunchKernelGGL(k_flux_staged<false>, dim3(t.nfb), dim3(kFluxStrip), 0, c->stream, t, b.V, b.grad,
                         c->fc, b.Fm, b.P, c->d_err, done);
    else if (c->p.second_order && t.fblock == kFluxBlockSmall)
      launch_flux_edges<true, kFluxBlockSmall>(c, b, done);
    else if (c->p.second_order)
      launch_flux_edges<true, kFluxBlock>(c, b, done);
    else if (t.fblock == kFluxBlockSmall)
      launch_flux_edges<false, kFluxBlockSmall>(c, b, done);
    else
      launch_flux_edges<false, kFluxBlock>(c, b, done);
  }
  // the Navier-Stokes Galerkin term (SpaceOperator.C:570-573): its element
  // pass, summed by the node pass with the fluxes (or, without the
  // contribution buffer, the per-node kernel after it)
  const bool fold = c->viscous && c->d_visc_c;
  if (fold) {
    int rc = launch_visc_elements(c, b.V, done);
    if (rc) return rc;
  }
  {
    ProfScope ps(c, AEROFX_TIMER_FV_NODES);
    hipLaunchKernelGGL(k_flux_nodes, dim3((t.N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, t, b.V,
                       b.Fm, b.P, c->fc, scale ? 1 : 0, fold ? c->elem->ntrow : nullptr,
                       fold ? c->d_visc_c : nullptr, R, done);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_hip_error(e, "residual launch");
  return fold ? 0 : launch_visc_residual(c, b.V, R, scale, done);
}

int launch_residual(aerofx_ctx *c, const double *U, double *R, bool scale, const int *done) {
  const ResidualBufs b{c->d_V, c->d_grad, c->d_Fm, c->d_P};
  return launch_residual_with(c, b, U, R, scale, done);
}

// SubDomain::finalizeTags (SubDomainCore.C:5177-5188) in its edge order: for
// each edge (i, j) tag1[i] = tag1[j] = tag0[i] + tag0[j], so a node keeps the
// value of its last edge (the last entry of its adjacency row, which is in
// edge order); then DistNodalGrad::fix (DistNodalGrad.C:1265-1284): the node
// is first order from now on if tag1 is set.
__global__ void k_fs_finalize(DevTables t, const int *__restrict__ tag0, int *__restrict__ tag1,
                              int *__restrict__ mask) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= t.N) return;
  const int e0 = t.rowptr[n], e1 = t.rowptr[n + 1];
  if (e1 > e0) tag1[n] = tag0[n] + tag0[t.adj[e1 - 1]];
  if (tag1[n]) mask[n] = 1;
}

static int read_error_count(aerofx_ctx *c, int *count) {
  hipError_t e = hipMemcpyAsync(c->h_err, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "residual error count");
  *count = c->h_err[0];
  return 0;
}

int residual_failsafe(aerofx_ctx *c, const ResidualBufs &b, const double *U, double *R, bool scale, int *count) {
  const bool fs = c->failsafe && c->p.second_order;
  ResidualBufs b1 = b;
  const int N = c->N;
  b1.tag0 = fs ? c->d_fs : nullptr;
  int rc = launch_residual_with(c, b1, U, R, scale);
  if (rc || (rc = read_error_count(c, count))) return rc;
  if (!fs || *count == 0) return 0;
  // Domain::computeFiniteVolumeTerm with failsafe (Domain.C:884-917)
  hipLaunchKernelGGL(k_fs_finalize, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->tab, c->d_fs,
                     c->d_fs + N, c->d_fs + 2 * (size_t)N);
  c->fs_any = true;
  if ((rc = launch_residual_with(c, b, U, R, scale)) || (rc = read_error_count(c, count))) return rc;
  // FailSafe = On clears the tags after the fix; AlwaysOn keeps them
  if (c->failsafe == 1) hipLaunchKernelGGL(k_zero_int, dim3((2 * N + 63) / 64), dim3(64), 0, c->stream, c->d_fs,
                                           2 * N, nullptr);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "failsafe launch");
}

int refuse_failsafe_nodes(aerofx_ctx *c, const char *who) {
  if (!c->fs_any) return 0;
  return set_error(AEROFX_ERR_STATE, std::string(who) +
                                         ": failsafe first-order nodes are set (aerofx_ctx_reset_failsafe clears them)");
}

int launch_gradients_v(aerofx_ctx *c, const double *V, double *ddx, double *ddy, double *ddz) {
  hipLaunchKernelGGL((k_gradients<false, true>), dim3((c->N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->tab, V,
                     c->gas, nullptr, nullptr, ddx, ddy, ddz, nullptr, nullptr);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "gradients launch");
}

int launch_primitive(aerofx_ctx *c, const double *U, double *V) {
  hipLaunchKernelGGL(k_primitive, dim3((c->N + 255) / 256), dim3(256), 0, c->stream, c->N, U, c->gas, V, nullptr);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "primitive launch");
}

}  // namespace aerofx
