// host_mesh.cpp -- host-side mesh tables (see host_mesh.h).
//
// Geometry follows the reference's operation order exactly so that the
// tables are bit-identical to what AERO-F would compute for the same mesh:
//   edge table      EdgeSet::find/createPointers   (EdgeCore.C:55-119)
//   control volumes ElemTet::computeControlVolumes (ElemTetCore.C:31-44, 102-112)
//   edge normals    ElemTet::computeEdgeNormalsGCL1, Xn == Xnp1
//                   (ElemTetCore.C:281-330, DistGeoState.C:499)
//   face normals    FaceTria::computeNormal        (FaceTriaCore.C:31-37)
//   LSQ factors     SubDomain::computeWeightsLeastSquares{Edge,Node}Part
//                   (SubDomainCore.C:1334-1381, 1755-1777)
//   RCM             Connectivity::renumByComponent/renumRCM/findPseudoDiam/
//                   rootLS (ConnectivityCore.C:311-635) over the graph of
//                   SubDomain::createEdgeBasedConnectivity (SubDomainCore.C:253-320)
#include "host_mesh.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace aerofx {

// ---------------------------------------------------------------------------
// Philox4x32-10
// ---------------------------------------------------------------------------
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(0xD2511F53u, c0, hi0, lo0);
    mulhilo32(0xCD9E8D57u, c2, hi1, lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double u01_53(uint32_t hi, uint32_t lo) {
  return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// Kuhn-split box
// ---------------------------------------------------------------------------
void kuhn_box_sizes(const BoxSpec &s, int &N, int &T, int &F) {
  long long n = (long long)(s.nx + 1) * (s.ny + 1) * (s.nz + 1);
  long long t = 6LL * s.nx * s.ny * s.nz;
  long long f = 4LL * ((long long)s.nx * s.ny + (long long)s.ny * s.nz + (long long)s.nx * s.nz);
  if (s.nx < 1 || s.ny < 1 || s.nz < 1 || n > 0x7fffffffLL || t > 0x7fffffffLL)
    throw std::invalid_argument("kuhn_box: bad sizes");
  N = (int)n; T = (int)t; F = (int)f;
}

// Lattice position of node (i, j, k) with the interior jitter, before the
// wall bumps.
static void lattice_node(const BoxSpec &s, int i, int j, int k, double x[3]) {
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int n = i + (nx + 1) * (j + (ny + 1) * k);
  const double h[3] = {(s.hi[0] - s.lo[0]) / nx, (s.hi[1] - s.lo[1]) / ny, (s.hi[2] - s.lo[2]) / nz};
  const uint32_t key[2] = {(uint32_t)s.seed, (uint32_t)(s.seed >> 32)};
  x[0] = s.lo[0] + h[0] * i;
  x[1] = s.lo[1] + h[1] * j;
  x[2] = s.lo[2] + h[2] * k;
  if (i == nx) x[0] = s.hi[0];
  if (j == ny) x[1] = s.hi[1];
  if (k == nz) x[2] = s.hi[2];
  bool interior = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
  if (interior && s.jitter != 0.0) {
    for (int a = 0; a < 3; ++a) {
      uint32_t ctr[4] = {(uint32_t)n, (uint32_t)a, 0x4a495454u /* "JITT" */, 0u}, o[4];
      philox4x32_10(ctr, key, o);
      double u = u01_53(o[0], o[1]);
      x[a] += (2.0 * u - 1.0) * s.jitter * h[a];
    }
  }
}

void kuhn_box_bump_derivative(const BoxSpec &s, int b, int param, double *dX) {
  if (b < 0 || b >= s.nbumps || (param != 0 && param != 1))
    throw std::invalid_argument("kuhn_box_bump_derivative: bad bump or parameter");
  for (int k = 0; k <= s.nz; ++k)
    for (int j = 0; j <= s.ny; ++j)
      for (int i = 0; i <= s.nx; ++i) {
        const int n = i + (s.nx + 1) * (j + (s.ny + 1) * k);
        double x[3];
        lattice_node(s, i, j, k, x);
        // z += a exp(-t^2) (1 - zeta), t = (x - c) / w
        const double zeta = (x[2] - s.lo[2]) / (s.hi[2] - s.lo[2]);
        const double t = (x[0] - s.bump_center[b]) / s.bump_width[b];
        const double g = std::exp(-t * t) * (1.0 - zeta);
        dX[3 * (size_t)n] = 0.0;
        dX[3 * (size_t)n + 1] = 0.0;
        dX[3 * (size_t)n + 2] = param == 0 ? g : s.bump_amp[b] * g * 2.0 * t / s.bump_width[b];
      }
}

void kuhn_box(const BoxSpec &s, double *X, int *tets, int *faces, int *codes) {
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  auto id = [&](int i, int j, int k) { return i + (nx + 1) * (j + (ny + 1) * k); };
  // node coordinates: lattice, interior jitter, then the wall bumps
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        int n = id(i, j, k);
        double x[3];
        lattice_node(s, i, j, k, x);
        if (s.nbumps > 0) {
          double zeta = (x[2] - s.lo[2]) / (s.hi[2] - s.lo[2]);
          double dz = 0.0;
          for (int b = 0; b < s.nbumps; ++b) {
            double t = (x[0] - s.bump_center[b]) / s.bump_width[b];
            dz += s.bump_amp[b] * std::exp(-t * t);
          }
          x[2] += dz * (1.0 - zeta);
        }
        X[3 * n] = x[0]; X[3 * n + 1] = x[1]; X[3 * n + 2] = x[2];
      }
  // tets: one per permutation of the axes, walking (0,0,0) -> (1,1,1)
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  static const int parity[6] = {+1, -1, -1, +1, +1, -1};
  int t = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        for (int p = 0; p < 6; ++p) {
          int c[3] = {i, j, k};
          int v[4];
          v[0] = id(c[0], c[1], c[2]);
          c[perm[p][0]]++;
          v[1] = id(c[0], c[1], c[2]);
          c[perm[p][1]]++;
          v[2] = id(c[0], c[1], c[2]);
          c[perm[p][2]]++;
          v[3] = id(c[0], c[1], c[2]);
          if (parity[p] < 0) std::swap(v[1], v[2]);
          for (int q = 0; q < 4; ++q) tets[4 * t + q] = v[q];
          ++t;
        }
  // boundary triangles: per box face, per quad, split along the Kuhn
  // diagonal (local (0,0)-(1,1)); oriented outward on the lattice.
  int f = 0;
  for (int side = 0; side < 6; ++side) {
    int ax = side / 2, hiSide = side % 2;
    int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
    if (a1 > a2) std::swap(a1, a2);          // in-plane axes in increasing order
    int n1 = (a1 == 0 ? nx : a1 == 1 ? ny : nz), n2 = (a2 == 0 ? nx : a2 == 1 ? ny : nz);
    int fixedIdx = hiSide ? (ax == 0 ? nx : ax == 1 ? ny : nz) : 0;
    double outward = hiSide ? 1.0 : -1.0;
    for (int q2 = 0; q2 < n2; ++q2)
      for (int q1 = 0; q1 < n1; ++q1) {
        auto node = [&](int d1, int d2) {
          int c[3];
          c[ax] = fixedIdx; c[a1] = q1 + d1; c[a2] = q2 + d2;
          return id(c[0], c[1], c[2]);
        };
        int tri[2][3] = {{node(0, 0), node(1, 0), node(1, 1)}, {node(0, 0), node(1, 1), node(0, 1)}};
        for (int tt = 0; tt < 2; ++tt) {
          // orientation from lattice vectors: (P2-P0)^(P1-P0) . e_ax
          // P1-P0 and P2-P0 in (a1,a2) coordinates
          int d1[2], d2[2];
          if (tt == 0) { d1[0] = 1; d1[1] = 0; d2[0] = 1; d2[1] = 1; }
          else { d1[0] = 1; d1[1] = 1; d2[0] = 0; d2[1] = 1; }
          double v1[3] = {0, 0, 0}, v2[3] = {0, 0, 0};
          v1[a1] = d1[0]; v1[a2] = d1[1];
          v2[a1] = d2[0]; v2[a2] = d2[1];
          double cr[3] = {v2[1] * v1[2] - v2[2] * v1[1], v2[2] * v1[0] - v2[0] * v1[2], v2[0] * v1[1] - v2[1] * v1[0]};
          int a = tri[tt][0], b = tri[tt][1], cc = tri[tt][2];
          if (cr[ax] * outward < 0) std::swap(b, cc);
          faces[3 * f] = a; faces[3 * f + 1] = b; faces[3 * f + 2] = cc;
          codes[f] = s.face_code[side];
          ++f;
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Edge table
// ---------------------------------------------------------------------------
static const int kEdgeEnd[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
static const int kEdgeFace[6][2] = {{0, 1}, {2, 0}, {1, 2}, {0, 3}, {3, 1}, {2, 3}};

void build_edges(int N, int T, const int *tets, std::vector<int> &ptr,
                 std::vector<int> *tet_edges) {
  // bucket the (min,max) pairs by their min node, sort and unique each bucket
  std::vector<long long> start(N + 1, 0);
  for (long long t = 0; t < T; ++t)
    for (int l = 0; l < 6; ++l) {
      int a = tets[4 * t + kEdgeEnd[l][0]], b = tets[4 * t + kEdgeEnd[l][1]];
      if (a < 0 || b < 0 || a >= N || b >= N || a == b) throw std::invalid_argument("build_edges: bad tet");
      start[std::min(a, b) + 1]++;
    }
  for (int n = 0; n < N; ++n) start[n + 1] += start[n];
  std::vector<int> bucket(start[N]);
  std::vector<long long> fill(start.begin(), start.end() - 1);
  for (long long t = 0; t < T; ++t)
    for (int l = 0; l < 6; ++l) {
      int a = tets[4 * t + kEdgeEnd[l][0]], b = tets[4 * t + kEdgeEnd[l][1]];
      if (a > b) std::swap(a, b);
      bucket[fill[a]++] = b;
    }
  std::vector<int> ustart(N + 1, 0);
  for (int n = 0; n < N; ++n) {
    auto b0 = bucket.begin() + start[n], b1 = bucket.begin() + start[n + 1];
    std::sort(b0, b1);
    ustart[n + 1] = ustart[n] + (int)(std::unique(b0, b1) - b0);
  }
  int E = ustart[N];
  ptr.assign(2 * (size_t)E, 0);
  for (int n = 0; n < N; ++n)
    for (int k = 0; k < ustart[n + 1] - ustart[n]; ++k) {
      ptr[2 * (size_t)(ustart[n] + k)] = n;
      ptr[2 * (size_t)(ustart[n] + k) + 1] = bucket[start[n] + k];
    }
  if (tet_edges) {
    tet_edges->assign(6 * (size_t)T, -1);
    for (long long t = 0; t < T; ++t)
      for (int l = 0; l < 6; ++l) {
        int a = tets[4 * t + kEdgeEnd[l][0]], b = tets[4 * t + kEdgeEnd[l][1]];
        if (a > b) std::swap(a, b);
        auto b0 = bucket.begin() + start[a], b1 = b0 + (ustart[a + 1] - ustart[a]);
        (*tet_edges)[6 * t + l] = ustart[a] + (int)(std::lower_bound(b0, b1, b) - b0);
      }
  }
}

// ---------------------------------------------------------------------------
// Geometry (reference operation order; Vec3D ops evaluate left to right)
// ---------------------------------------------------------------------------
namespace {
struct V3 { double v[3]; };
inline V3 ld(const double *X, int i) { return V3{{X[3 * i], X[3 * i + 1], X[3 * i + 2]}}; }
inline V3 sub(V3 a, V3 b) { return V3{{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]}}; }
inline V3 add(V3 a, V3 b) { return V3{{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]}}; }
inline V3 scl(double c, V3 a) { return V3{{c * a.v[0], c * a.v[1], c * a.v[2]}}; }
inline V3 cross(V3 a, V3 b) {
  return V3{{a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2],
             a.v[0] * b.v[1] - a.v[1] * b.v[0]}};
}
inline double dot(V3 a, V3 b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
}  // namespace

void control_volumes(int N, int T, const double *X, const int *tets, double *cv) {
  const double sixth = 1.0 / 6.0;
  std::fill(cv, cv + N, 0.0);
  for (long long t = 0; t < T; ++t) {
    const int *nd = tets + 4 * t;
    V3 x0 = ld(X, nd[0]);
    V3 v1 = sub(ld(X, nd[1]), x0), v2 = sub(ld(X, nd[2]), x0), v3 = sub(ld(X, nd[3]), x0);
    double volume = sixth * dot(v3, cross(v1, v2));
    for (int j = 0; j < 4; ++j) cv[nd[j]] += 0.25 * volume;
  }
}

void edge_normals(int T, const double *X, const int *tets, const int *tet_edges, int E,
                  double *normals) {
  const double third = 1.0 / 3.0;
  std::fill(normals, normals + 3 * (size_t)E, 0.0);
  for (long long t = 0; t < T; ++t) {
    const int *nd = tets + 4 * t;
    V3 x[4] = {ld(X, nd[0]), ld(X, nd[1]), ld(X, nd[2]), ld(X, nd[3])};
    V3 f[4] = {scl(third, add(add(x[0], x[1]), x[2])), scl(third, add(add(x[0], x[1]), x[3])),
               scl(third, add(add(x[0], x[2]), x[3])), scl(third, add(add(x[1], x[2]), x[3]))};
    V3 g = scl(0.25, add(add(add(x[0], x[1]), x[2]), x[3]));
    for (int l = 0; l < 6; ++l) {
      V3 a = cross(sub(f[kEdgeFace[l][1]], g), sub(f[kEdgeFace[l][0]], g));
      V3 n = add(scl(0.5, add(a, a)), scl(0.25, add(a, a)));
      double *dst = normals + 3 * (size_t)tet_edges[6 * t + l];
      if (nd[kEdgeEnd[l][0]] < nd[kEdgeEnd[l][1]]) {
        dst[0] += n.v[0]; dst[1] += n.v[1]; dst[2] += n.v[2];
      } else {
        dst[0] -= n.v[0]; dst[1] -= n.v[1]; dst[2] -= n.v[2];
      }
    }
  }
}

void face_normals(int F, const double *X, const int *faces, double *fnorm) {
  for (int f = 0; f < F; ++f) {
    const int *nd = faces + 3 * f;
    V3 x0 = ld(X, nd[0]);
    V3 n = scl(0.5, cross(sub(ld(X, nd[2]), x0), sub(ld(X, nd[1]), x0)));
    fnorm[3 * f] = n.v[0]; fnorm[3 * f + 1] = n.v[1]; fnorm[3 * f + 2] = n.v[2];
  }
}

void lsq_factors(int N, int E, const double *X, const int *ptr, double *R) {
  std::fill(R, R + 6 * (size_t)N, 0.0);
  for (long long l = 0; l < E; ++l) {
    int i = ptr[2 * l], j = ptr[2 * l + 1];
    double dx0 = X[3 * j] - X[3 * i], dx1 = X[3 * j + 1] - X[3 * i + 1], dx2 = X[3 * j + 2] - X[3 * i + 2];
    double m[6] = {dx0 * dx0, dx0 * dx1, dx0 * dx2, dx1 * dx1, dx1 * dx2, dx2 * dx2};
    for (int q = 0; q < 6; ++q) { R[6 * (size_t)i + q] += m[q]; R[6 * (size_t)j + q] += m[q]; }
  }
  for (int i = 0; i < N; ++i) {
    double *r = R + 6 * (size_t)i;
    double r11 = std::sqrt(r[0]);
    double or11 = 1.0 / r11;
    double r12 = r[1] * or11;
    double r13 = r[2] * or11;
    double r22 = std::sqrt(r[3] - r12 * r12);
    double r23 = (r[4] - r12 * r13) / r22;
    double r33 = std::sqrt(r[5] - (r13 * r13 + r23 * r23));
    r[0] = r11; r[1] = r12; r[2] = r13; r[3] = r22; r[4] = r23; r[5] = r33;
  }
}

void node_csr(int N, int E, const int *ptr, std::vector<int> &rowptr, std::vector<int> &col,
              std::vector<int> &eid) {
  rowptr.assign(N + 1, 0);
  for (long long l = 0; l < E; ++l) { rowptr[ptr[2 * l] + 1]++; rowptr[ptr[2 * l + 1] + 1]++; }
  for (int n = 0; n < N; ++n) rowptr[n + 1] += rowptr[n];
  col.assign(2 * (size_t)E, 0);
  eid.assign(2 * (size_t)E, 0);
  std::vector<int> fill(rowptr.begin(), rowptr.end() - 1);
  // ascending edge id => each row receives its incident edges in id order,
  // which is also ascending neighbour order (lower neighbours first).
  for (int l = 0; l < E; ++l) {
    int i = ptr[2 * (size_t)l], j = ptr[2 * (size_t)l + 1];
    col[fill[i]] = j; eid[fill[i]++] = l;
    col[fill[j]] = i; eid[fill[j]++] = l;
  }
}

bool halo_strips(int N, const int *rowptr, const int *col, int max_rows, int max_items, int cap,
                 HaloStrips &hs) {
  hs = HaloStrips();
  if (cap > 65536) return false;  // positions are 16-bit
  hs.start.push_back(0);
  hs.hptr.push_back(0);
  hs.loc.assign((size_t)rowptr[N], 0);
  std::vector<int> stamp(N, -1), pos(N, 0), cur;
  int blk = 0, r0 = 0;
  auto close = [&](int r1) {
    std::sort(cur.begin(), cur.end());
    for (size_t k = 0; k < cur.size(); ++k) pos[cur[k]] = (int)k;
    for (int e = rowptr[r0]; e < rowptr[r1]; ++e) hs.loc[e] = (uint16_t)pos[col[e]];
    hs.own.push_back(pos[r0]);
    hs.halo.insert(hs.halo.end(), cur.begin(), cur.end());
    hs.max_halo = std::max(hs.max_halo, (int)cur.size());
    hs.start.push_back(r1);
    hs.hptr.push_back((int)hs.halo.size());
    cur.clear();
    ++blk;
    r0 = r1;
  };
  for (int r = 0; r < N; ++r) {
    int fresh = stamp[r] != blk;
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) fresh += stamp[col[e]] != blk;
    const bool fits = r - r0 < max_rows && rowptr[r + 1] - rowptr[r0] <= max_items &&
                      (int)cur.size() + fresh <= cap;
    if (!fits) {
      if (r == r0) return false;  // this row alone exceeds a limit
      close(r);
      // the row starts a new block: count again against the empty halo
      fresh = 1 + (rowptr[r + 1] - rowptr[r]);
      if (1 > max_rows || rowptr[r + 1] - rowptr[r] > max_items || fresh > cap) return false;
    }
    if (stamp[r] != blk) { stamp[r] = blk; cur.push_back(r); }
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e)
      if (stamp[col[e]] != blk) { stamp[col[e]] = blk; cur.push_back(col[e]); }
  }
  if (r0 < N) close(N);
  return true;
}

// ---------------------------------------------------------------------------
// RCM
// ---------------------------------------------------------------------------
namespace {
struct Graph {
  int size;
  std::vector<int> pointer, target;
  int num(int nd, const int *mask) const {
    int r = 0;
    for (int j = pointer[nd]; j < pointer[nd + 1]; ++j) if (mask[target[j]]) r++;
    return r;
  }
  // rooted level structure; returns the number of levels
  int rootLS(int root, int *xls, int *ls, int &w, const int *mask, std::vector<int> &loc) const {
    w = 0;
    if (mask) std::copy(mask, mask + size, loc.begin());
    else std::fill(loc.begin(), loc.end(), 1);
    loc[root] = 0;
    ls[0] = root;
    xls[0] = 0;
    int nlvl = 1, nf = 1;
    while (nf > xls[nlvl - 1]) {
      xls[nlvl] = nf;
      int lb = xls[nlvl - 1], le = xls[nlvl];
      for (int i = lb; i < le; ++i) {
        int n1 = ls[i];
        for (int j = pointer[n1]; j < pointer[n1 + 1]; ++j) {
          int n2 = target[j];
          if (loc[n2]) { loc[n2] = 0; ls[nf++] = n2; }
        }
      }
      if (nf - xls[nlvl] > w) w = nf - xls[nlvl];
      nlvl++;
    }
    return nlvl - 1;
  }
};
}  // namespace

void rcm(int N, int E, const int *ptr, int *renum) {
  Graph g;
  g.size = N;
  g.pointer.assign(N + 1, 0);
  for (int i = 0; i < N; ++i) g.pointer[i + 1] = 1;
  for (long long l = 0; l < E; ++l) { g.pointer[ptr[2 * l] + 1]++; g.pointer[ptr[2 * l + 1] + 1]++; }
  for (int i = 0; i < N; ++i) g.pointer[i + 1] += g.pointer[i];
  g.target.assign(g.pointer[N], 0);
  {
    std::vector<int> fill(g.pointer.begin(), g.pointer.end() - 1);
    for (int i = 0; i < N; ++i) g.target[fill[i]++] = i;
    for (long long l = 0; l < E; ++l) {
      int a = ptr[2 * l], b = ptr[2 * l + 1];
      g.target[fill[a]++] = b;
      g.target[fill[b]++] = a;
    }
    for (int i = 0; i < N; ++i) std::sort(g.target.begin() + g.pointer[i], g.target.begin() + g.pointer[i + 1]);
  }
  const int numtarget = g.pointer[N];
  std::vector<int> mark(N), ls(std::max(N, 1)), xls(N + 1), locMask(N, 0), lrenum(N), loc(N);
  std::vector<int> pls(numtarget), pxls(numtarget + 1), order(N);
  for (int i = 0; i < N; ++i) { mark[i] = (g.pointer[i + 1] != g.pointer[i]) ? 1 : 0; renum[i] = -1; }
  int nextNum = 0;
  for (int inode = 0; inode < N; ++inode) {
    if (mark[inode] != 1) continue;
    int w;
    int h = g.rootLS(inode, xls.data(), ls.data(), w, nullptr, loc);
    for (int j = 0; j < xls[h]; ++j) { locMask[ls[j]] = 1; mark[ls[j]] = 0; }
    // findPseudoDiam
    int *mask = locMask.data();
    int s = 0, e = 0, cmin = numtarget + 1, cmax = 0, nw;
    for (int i = 0; i < N; ++i)
      if (mask[i]) {
        if ((nw = g.num(i, mask)) < cmin) { cmin = nw; s = i; }
        if (nw > cmax) cmax = nw;
      }
    std::vector<int> sorted(cmax + 1);
    int ww;
    int hh = g.rootLS(s, pxls.data(), pls.data(), ww, mask, loc);
    int subconsize = pxls[hh];
    e = pls[pxls[hh - 1]];
    while (hh < subconsize) {
      int k;
      for (k = 0; k <= cmax; ++k) sorted[k] = -1;
      for (k = pxls[hh - 1]; k < pxls[hh]; ++k) sorted[g.num(pls[k], mask)] = pls[k];
      int w_e = subconsize;
      for (k = 0; k <= cmax; ++k)
        if (sorted[k] >= 0) {
          int nh = g.rootLS(sorted[k], pxls.data(), pls.data(), ww, mask, loc);
          if (ww < w_e) {
            if (nh > hh) { s = sorted[k]; hh = nh; break; }
            e = sorted[k];
            w_e = ww;
          }
        }
      if (k > cmax) break;
    }
    // renumRCM: BFS from e in ascending-neighbour order (the reference's
    // degree sort leaves the order unchanged), then reverse.
    order[0] = e;
    mask[e] = -mask[e];
    int last = 1;
    for (int i = 0; i < last; ++i) {
      int cur = order[i];
      for (int j = g.pointer[cur]; j < g.pointer[cur + 1]; ++j) {
        int nb = g.target[j];
        if (mask[nb] > 0) { order[last++] = nb; mask[nb] = -mask[nb]; }
      }
    }
    for (int i = last; i--;) lrenum[order[i]] = nextNum++;
    for (int j = 0; j < xls[h]; ++j) locMask[ls[j]] = 0;
    for (int j = 0; j < xls[h]; ++j) renum[ls[j]] = lrenum[ls[j]];
  }
}

// ---------------------------------------------------------------------------
// Edge colouring
// ---------------------------------------------------------------------------
uint32_t edge_priority(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

int color_edges(int N, int E, const int *ptr, int *color) {
  std::vector<uint64_t> key(E);
  for (int l = 0; l < E; ++l) key[l] = ((uint64_t)edge_priority((uint32_t)l) << 32) | (uint32_t)l;
  std::sort(key.begin(), key.end(), [](uint64_t a, uint64_t b) { return a > b; });
  std::vector<uint64_t> used(N, 0);
  int ncol = 0;
  for (int k = 0; k < E; ++k) {
    int l = (int)(key[k] & 0xffffffffu);
    int i = ptr[2 * (size_t)l], j = ptr[2 * (size_t)l + 1];
    uint64_t u = used[i] | used[j];
    if (u == ~0ULL) return -1;
    int c = __builtin_ctzll(~u);
    color[l] = c;
    used[i] |= 1ULL << c;
    used[j] |= 1ULL << c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

void renumber_mesh(int N, int T, int F, const int *renum, double *X, int *tets, int *faces) {
  std::vector<double> Xo(X, X + 3 * (size_t)N);
  for (int i = 0; i < N; ++i)
    for (int a = 0; a < 3; ++a) X[3 * (size_t)renum[i] + a] = Xo[3 * (size_t)i + a];
  for (long long k = 0; k < 4LL * T; ++k) tets[k] = renum[tets[k]];
  for (long long k = 0; k < 3LL * F; ++k) faces[k] = renum[faces[k]];
}

}  // namespace aerofx
