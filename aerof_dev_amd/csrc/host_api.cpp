// host_api.cpp -- host-only part of the C ABI (mesh preparation, synthetic
// states, error reporting).  Needs no GPU.
#include <cmath>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "aerofx.h"
#include "host_internal.h"
#include "host_mesh.h"

namespace aerofx {
static thread_local std::string g_last_error;
int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

// DistBcDataEuler::setBoundaryConditionsGas (DistBcData.C:652-735) with zero
// gravity and identical inlet/outlet data: the per-node farfield state.
void freestream_state(const aerofx_params &p, double U[5]) {
  double gam = p.gamma, Pstiff = p.pstiff;
  double rhoin = p.rho_inf, pressurein = p.p_inf;
  double velin2 = gam * (pressurein + Pstiff) * p.mach * p.mach / rhoin;
  double velin = std::sqrt(velin2);
  U[0] = rhoin;
  U[1] = U[0] * velin * std::cos(p.alpha) * std::cos(p.yaw);
  U[2] = U[0] * velin * std::cos(p.alpha) * std::sin(p.yaw);
  U[3] = U[0] * velin * std::sin(p.alpha);
  double ptempin = pressurein + rhoin * 0.0;
  U[4] = 0.5 * rhoin * velin2 + (ptempin + gam * Pstiff) / (gam - 1.0);
}

// Face::assignFreeStreamValues2 (Face.C:38-70) for a triangle whose three
// nodes carry the same farfield state.
void face_state(const double Unode[5], double Ub[5]) {
  for (int k = 0; k < 5; ++k) {
    Ub[k] = 0.0;
    for (int j = 0; j < 3; ++j) Ub[k] += Unode[k];
    Ub[k] /= 3;
  }
}
}  // namespace aerofx

using namespace aerofx;

extern "C" {

const char *aerofx_last_error(void) { return g_last_error.c_str(); }
int aerofx_version(void) { return 1; }

static BoxSpec to_spec(const aerofx_box_spec *s) {
  BoxSpec b;
  b.nx = s->nx; b.ny = s->ny; b.nz = s->nz;
  for (int a = 0; a < 3; ++a) { b.lo[a] = s->lo[a]; b.hi[a] = s->hi[a]; }
  b.jitter = s->jitter;
  b.seed = s->seed;
  for (int f = 0; f < 6; ++f) b.face_code[f] = s->face_code[f];
  b.nbumps = s->nbumps;
  b.bump_amp = s->bump_amp; b.bump_center = s->bump_center; b.bump_width = s->bump_width;
  return b;
}

int aerofx_mesh_box_sizes(const aerofx_box_spec *spec, int *N, int *T, int *F) {
  if (!spec || !N || !T || !F) return set_error(AEROFX_ERR_ARG, "mesh_box_sizes: null argument");
  try {
    kuhn_box_sizes(to_spec(spec), *N, *T, *F);
  } catch (std::exception &e) {
    return set_error(AEROFX_ERR_ARG, e.what());
  }
  return AEROFX_OK;
}

int aerofx_mesh_box(const aerofx_box_spec *spec, double *X, int *tets, int *faces, int *codes) {
  if (!spec || !X || !tets || !faces || !codes) return set_error(AEROFX_ERR_ARG, "mesh_box: null argument");
  if (spec->nbumps < 0 || (spec->nbumps > 0 && (!spec->bump_amp || !spec->bump_center || !spec->bump_width)))
    return set_error(AEROFX_ERR_ARG, "mesh_box: bad bump arrays");
  try {
    BoxSpec b = to_spec(spec);
    int N, T, F;
    kuhn_box_sizes(b, N, T, F);
    kuhn_box(b, X, tets, faces, codes);
  } catch (std::exception &e) {
    return set_error(AEROFX_ERR_ARG, e.what());
  }
  return AEROFX_OK;
}

int aerofx_mesh_box_bump_derivative(const aerofx_box_spec *spec, int bump, int param, double *dX) {
  if (!spec || !dX) return set_error(AEROFX_ERR_ARG, "mesh_box_bump_derivative: null argument");
  if (spec->nbumps < 0 || (spec->nbumps > 0 && (!spec->bump_amp || !spec->bump_center || !spec->bump_width)))
    return set_error(AEROFX_ERR_ARG, "mesh_box_bump_derivative: bad bump arrays");
  try {
    BoxSpec b = to_spec(spec);
    int N, T, F;
    kuhn_box_sizes(b, N, T, F);
    kuhn_box_bump_derivative(b, bump, param, dX);
  } catch (std::exception &e) {
    return set_error(AEROFX_ERR_ARG, e.what());
  }
  return AEROFX_OK;
}

int aerofx_mesh_edges(int N, int T, const int *tets, int *num_edges, int *ptr) {
  if (N <= 0 || T <= 0 || !tets || !num_edges) return set_error(AEROFX_ERR_ARG, "mesh_edges: bad argument");
  try {
    std::vector<int> p;
    build_edges(N, T, tets, p, nullptr);
    *num_edges = (int)(p.size() / 2);
    if (ptr) std::memcpy(ptr, p.data(), p.size() * sizeof(int));
  } catch (std::exception &e) {
    return set_error(AEROFX_ERR_MESH, e.what());
  }
  return AEROFX_OK;
}

int aerofx_mesh_rcm(int N, int T, const int *tets, int *renum) {
  if (N <= 0 || T <= 0 || !tets || !renum) return set_error(AEROFX_ERR_ARG, "mesh_rcm: bad argument");
  try {
    std::vector<int> p;
    build_edges(N, T, tets, p, nullptr);
    rcm(N, (int)(p.size() / 2), p.data(), renum);
  } catch (std::exception &e) {
    return set_error(AEROFX_ERR_MESH, e.what());
  }
  return AEROFX_OK;
}

int aerofx_mesh_renumber(int N, int T, int F, const int *renum, double *X, int *tets, int *faces) {
  if (N <= 0 || T < 0 || F < 0 || !renum || !X || !tets || (F > 0 && !faces))
    return set_error(AEROFX_ERR_ARG, "mesh_renumber: bad argument");
  std::vector<char> seen(N, 0);
  for (int i = 0; i < N; ++i) {
    if (renum[i] < 0 || renum[i] >= N || seen[renum[i]]) return set_error(AEROFX_ERR_ARG, "mesh_renumber: not a permutation");
    seen[renum[i]] = 1;
  }
  renumber_mesh(N, T, F, renum, X, tets, faces);
  return AEROFX_OK;
}

int aerofx_mesh_color_edges(int N, int E, const int *ptr, int *color, int *num_colors) {
  if (N <= 0 || E < 0 || !ptr || !color || !num_colors) return set_error(AEROFX_ERR_ARG, "color_edges: bad argument");
  int nc = color_edges(N, E, ptr, color);
  if (nc < 0) return set_error(AEROFX_ERR_MESH, "color_edges: more than 64 colours needed");
  *num_colors = nc;
  return AEROFX_OK;
}

int aerofx_state_init(const aerofx_params *p, int kind, double amp, uint64_t seed, int N,
                      const double *X, double *U) {
  if (!p || N <= 0 || !U || (kind != 0 && !X)) return set_error(AEROFX_ERR_ARG, "state_init: bad argument");
  double Uf[5];
  freestream_state(*p, Uf);
  const double gam = p->gamma;
  const double rho0 = Uf[0];
  const double vel[3] = {Uf[1] / Uf[0], Uf[2] / Uf[0], Uf[3] / Uf[0]};
  const double vel2 = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
  const double p0 = p->p_inf;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  double phase[3] = {0, 0, 0};
  if (kind == 2) {
    uint32_t ctr[4] = {0u, 0u, 0x50484153u /* "PHAS" */, 0u}, o[4];
    philox4x32_10(ctr, key, o);
    phase[0] = u01_53(o[0], o[1]);
    phase[1] = u01_53(o[2], o[3]);
    ctr[1] = 1u;
    philox4x32_10(ctr, key, o);
    phase[2] = u01_53(o[0], o[1]);
  }
  const double twopi = 6.283185307179586476925286766559;
  for (int n = 0; n < N; ++n) {
    double *u = U + 5 * (size_t)n;
    if (kind == 0) {
      for (int k = 0; k < 5; ++k) u[k] = Uf[k];
      continue;
    }
    double rho = rho0, pr = p0;
    if (kind == 1) {
      uint32_t ctr[4] = {(uint32_t)n, 0u, 0x4e4f4953u /* "NOIS" */, 0u}, o[4];
      philox4x32_10(ctr, key, o);
      double u1 = 1.0 - u01_53(o[0], o[1]);  // (0,1]
      double u2 = u01_53(o[2], o[3]);
      double rr = std::sqrt(-2.0 * std::log(u1));
      double z1 = rr * std::cos(twopi * u2), z2 = rr * std::sin(twopi * u2);
      rho = rho0 * (1.0 + amp * z1);
      pr = p0 * (1.0 + amp * z2);
    } else if (kind == 2) {
      const double *x = X + 3 * (size_t)n;
      rho = rho0 * (1.0 + amp * std::sin(twopi * (x[0] + phase[0])) * std::sin(twopi * (x[1] + phase[1])) *
                              std::sin(twopi * (x[2] + phase[2])));
    } else {
      return set_error(AEROFX_ERR_ARG, "state_init: unknown kind");
    }
    u[0] = rho;
    u[1] = rho * vel[0];
    u[2] = rho * vel[1];
    u[3] = rho * vel[2];
    u[4] = pr / (gam - 1.0) + 0.5 * rho * vel2;
  }
  return AEROFX_OK;
}

}  // extern "C"
