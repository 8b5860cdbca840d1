// host_mesh.h -- host-side mesh tables for the edge-based FV operator.
//
// Synthetic Kuhn-split meshes, the edge table, dual-cell geometry, LSQ
// factors, the reference RCM ordering and the edge colouring rule.  Plain
// C++ (no HIP): set-up work that runs once per mesh.
#pragma once

#include <cstdint>
#include <vector>

namespace aerofx {

// Counter-based PRNG: Philox4x32-10 (Salmon et al., SC'11).
void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
// 53-bit uniform in [0,1) from two 32-bit words.
double u01_53(uint32_t hi, uint32_t lo);

struct BoxSpec {
  int nx, ny, nz;
  double lo[3], hi[3];
  double jitter;        // +-jitter*h_axis, uniform, interior nodes only
  uint64_t seed;
  int face_code[6];     // x-, x+, y-, y+, z-, z+
  int nbumps;
  const double *bump_amp, *bump_center, *bump_width;  // Gaussian bumps on z-
};

// Kuhn split: every hex -> 6 tets around its (0,0,0)-(1,1,1) diagonal.
// Node id = i + (nx+1)*(j + (ny+1)*k).  Tets positively oriented.  Boundary
// triangles oriented so that 0.5*(x2-x0)^(x1-x0) points outward.
void kuhn_box_sizes(const BoxSpec &s, int &N, int &T, int &F);
void kuhn_box(const BoxSpec &s, double *X, int *tets, int *faces, int *codes);
// Closed-form dX/dp [N][3] of bump b's amplitude (param 0) or centre (1).
void kuhn_box_bump_derivative(const BoxSpec &s, int b, int param, double *dX);

// Unique edges sorted lexicographically by (min id, max id); optional
// tet -> edge map [T][6] in ElemTet local-edge order.
void build_edges(int N, int T, const int *tets, std::vector<int> &ptr,
                 std::vector<int> *tet_edges);

// Dual-cell geometry: control volumes [N], edge normals [E][3] oriented
// low id -> high id, face normals [F][3] (full triangle, not /3).
void control_volumes(int N, int T, const double *X, const int *tets, double *cv);
void edge_normals(int T, const double *X, const int *tets, const int *tet_edges,
                  int E, double *normals);
void face_normals(int F, const double *X, const int *faces, double *fnorm);

// LSQ normal-matrix Cholesky factors [N][6] = (r11,r12,r13,r22,r23,r33).
void lsq_factors(int N, int E, const double *X, const int *ptr, double *R);

// Node adjacency in CSR form; each row sorted ascending (which is the order
// of the incident edges' ids).  eid[k] = edge id of entry k.
void node_csr(int N, int E, const int *ptr, std::vector<int> &rowptr,
              std::vector<int> &col, std::vector<int> &eid);

// Strips of consecutive rows for kernels that stage the data of every node a
// block touches in LDS once.  Row r owns the entries [rowptr[r], rowptr[r+1])
// with target nodes col[e].  Block b covers rows [start[b], start[b+1]); its
// halo is the sorted set of the rows themselves and all targets of their
// entries, halo[hptr[b] .. hptr[b+1]).  A block closes before it would exceed
// max_rows rows, max_items entries or cap halo nodes.  loc[e] is the position
// of col[e] in its block's halo and own[b] the position of row start[b] (the
// rows are consecutive ids, so row r sits at own[b] + r - start[b]).
// Returns false if a single row alone exceeds a limit.
struct HaloStrips {
  std::vector<int> start, hptr, halo, own;
  std::vector<uint16_t> loc;
  int max_halo = 0;
};
bool halo_strips(int N, const int *rowptr, const int *col, int max_rows, int max_items, int cap,
                 HaloStrips &hs);

// Reference RCM rule over the edge graph with self loops: renum[old] = new.
void rcm(int N, int E, const int *ptr, int *renum);

// Edge-colouring rule: priority key (hash(l) << 32 | l); each edge takes the
// smallest colour unused by conflicting edges of larger key.
uint32_t edge_priority(uint32_t l);
int color_edges(int N, int E, const int *ptr, int *color);

// Applies renum[old] = new to a mesh (X permuted, tets/faces relabelled).
void renumber_mesh(int N, int T, int F, const int *renum, double *X, int *tets,
                   int *faces);

}  // namespace aerofx
