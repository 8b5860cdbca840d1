/*
 * aerof_oracle.c -- CPU restatement of the AERO-F edge-based FV residual.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 * Compiled with -O2 -ffp-contract=off (see oracle/Makefile).
 *
 * Each function cites the reference file:line whose loop order and operation
 * order it follows (paths relative to the reference tree).  Where the
 * reference evaluates an expression left to right (C++ operator chains,
 * Fortran statement order) the same association is written out here.
 */
#include "aerof_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Edge table                                                                */
/* ------------------------------------------------------------------------ */

/* ElemTet::edgeEndTet (ElemTetCore.C:19) */
static const int kEdgeEnd[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
/* ElemTet::edgeFaceTet (ElemTetCore.C:20) */
static const int kEdgeFace[6][2] = {{0, 1}, {2, 0}, {1, 2}, {0, 3}, {3, 1}, {2, 3}};

static int cmp_pair(const void *a, const void *b) {
  const int *x = (const int *)a, *y = (const int *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
  return 0;
}

static int find_pair(const int *ptr, int E, int a, int b) {
  int lo = 0, hi = E - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    const int *p = ptr + 2 * mid;
    if (p[0] < a || (p[0] == a && p[1] < b)) lo = mid + 1;
    else if (p[0] == a && p[1] == b) return mid;
    else hi = mid - 1;
  }
  return -1;
}

/* EdgeSet::find keys a std::map<pair<min,max>,int> (EdgeCore.C:55-70) and
 * createPointers walks that map in key order (EdgeCore.C:92-119); the final
 * numbering therefore depends only on the edge set: lexicographic (min,max).
 * Elements are visited in element order (SubDomainCore.C:168-186). */
int or_edges(int N, int T, const int *tets, int *ptr, int *tet_edges) {
  (void)N;
  int n = 0;
  for (int t = 0; t < T; ++t)
    for (int l = 0; l < 6; ++l) {
      int a = tets[4 * t + kEdgeEnd[l][0]], b = tets[4 * t + kEdgeEnd[l][1]];
      if (a > b) { int s = a; a = b; b = s; }
      ptr[2 * n] = a; ptr[2 * n + 1] = b; ++n;
    }
  qsort(ptr, (size_t)n, 2 * sizeof(int), cmp_pair);
  int E = 0;
  for (int k = 0; k < n; ++k) {
    if (E > 0 && ptr[2 * (E - 1)] == ptr[2 * k] && ptr[2 * (E - 1) + 1] == ptr[2 * k + 1]) continue;
    ptr[2 * E] = ptr[2 * k]; ptr[2 * E + 1] = ptr[2 * k + 1]; ++E;
  }
  if (tet_edges)
    for (int t = 0; t < T; ++t)
      for (int l = 0; l < 6; ++l) {
        int a = tets[4 * t + kEdgeEnd[l][0]], b = tets[4 * t + kEdgeEnd[l][1]];
        if (a > b) { int s = a; a = b; b = s; }
        tet_edges[6 * t + l] = find_pair(ptr, E, a, b);
      }
  return E;
}

/* ------------------------------------------------------------------------ */
/* Geometry                                                                  */
/* ------------------------------------------------------------------------ */

typedef struct { double v[3]; } v3;
static v3 v3_sub(v3 a, v3 b) { v3 r = {{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]}}; return r; }
static v3 v3_add(v3 a, v3 b) { v3 r = {{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]}}; return r; }
static v3 v3_scale(double c, v3 a) { v3 r = {{c * a.v[0], c * a.v[1], c * a.v[2]}}; return r; }
/* Vec3D::operator^ (Vector3D.h:266-276) */
static v3 v3_cross(v3 a, v3 b) {
  v3 r;
  r.v[0] = a.v[1] * b.v[2] - a.v[2] * b.v[1];
  r.v[1] = a.v[2] * b.v[0] - a.v[0] * b.v[2];
  r.v[2] = a.v[0] * b.v[1] - a.v[1] * b.v[0];
  return r;
}
/* Vec3D::operator* (Vector3D.h:282-287) */
static double v3_dot(v3 a, v3 b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
static v3 v3_load(const double *X, int i) { v3 r = {{X[3 * i], X[3 * i + 1], X[3 * i + 2]}}; return r; }

/* ElemTet::computeVolume (ElemTetCore.C:31-44), computeControlVolumes
 * (ElemTetCore.C:102-112). */
void or_control_volumes(int N, int T, const double *X, const int *tets, double *cv) {
  const double sixth = 1.0 / 6.0;
  for (int i = 0; i < N; ++i) cv[i] = 0.0;
  for (int t = 0; t < T; ++t) {
    const int *nd = tets + 4 * t;
    v3 x0 = v3_load(X, nd[0]), x1 = v3_load(X, nd[1]), x2 = v3_load(X, nd[2]), x3 = v3_load(X, nd[3]);
    v3 v1 = v3_sub(x1, x0), v2 = v3_sub(x2, x0), v3_ = v3_sub(x3, x0);
    double volume = sixth * v3_dot(v3_, v3_cross(v1, v2));
    for (int j = 0; j < 4; ++j) cv[nd[j]] += 0.25 * volume;
  }
}

/* ElemTet::computeEdgeNormalsGCL1 (ElemTetCore.C:281-330) as called at setup
 * with Xn == Xnp1 and Xdot == 0 (DistGeoState.C:499). */
void or_edge_normals(int T, const double *X, const int *tets, const int *tet_edges,
                     int E, double *normals, double *normal_vel) {
  const double third = 1.0 / 3.0;
  const double c0 = 13.0 / 36.0, c1 = 5.0 / 36.0;
  static const int edgeOpEnd[6][2] = {{2, 3}, {3, 1}, {1, 2}, {0, 3}, {2, 0}, {0, 1}};
  (void)edgeOpEnd;
  for (int l = 0; l < 3 * E; ++l) normals[l] = 0.0;
  for (int l = 0; l < E; ++l) normal_vel[l] = 0.0;
  for (int t = 0; t < T; ++t) {
    const int *nd = tets + 4 * t;
    v3 x[4];
    for (int k = 0; k < 4; ++k) x[k] = v3_load(X, nd[k]);
    v3 f[4];
    f[0] = v3_scale(third, v3_add(v3_add(x[0], x[1]), x[2]));
    f[1] = v3_scale(third, v3_add(v3_add(x[0], x[1]), x[3]));
    f[2] = v3_scale(third, v3_add(v3_add(x[0], x[2]), x[3]));
    f[3] = v3_scale(third, v3_add(v3_add(x[1], x[2]), x[3]));
    v3 g = v3_scale(0.25, v3_add(v3_add(v3_add(x[0], x[1]), x[2]), x[3]));
    for (int l = 0; l < 6; ++l) {
      v3 xg0 = v3_sub(f[kEdgeFace[l][0]], g);
      v3 xg1 = v3_sub(f[kEdgeFace[l][1]], g);
      v3 a = v3_cross(xg1, xg0);
      /* 0.5*((a)+(a)) + 0.25*((a)+(a)) with identical n and n+1 terms */
      v3 n = v3_add(v3_scale(0.5, v3_add(a, a)), v3_scale(0.25, v3_add(a, a)));
      /* xdot == 0: ndot = (c0*(0+0) + c1*(0+0)) * n */
      v3 z = {{0.0, 0.0, 0.0}};
      v3 xd = v3_add(v3_scale(c0, v3_add(z, z)), v3_scale(c1, v3_add(z, z)));
      double ndot = v3_dot(xd, n);
      int e = tet_edges[6 * t + l];
      if (nd[kEdgeEnd[l][0]] < nd[kEdgeEnd[l][1]]) {
        for (int k = 0; k < 3; ++k) normals[3 * e + k] += n.v[k];
        normal_vel[e] += ndot;
      } else {
        for (int k = 0; k < 3; ++k) normals[3 * e + k] -= n.v[k];
        normal_vel[e] -= ndot;
      }
    }
  }
}

/* FaceTria::computeNormal (FaceTriaCore.C:31-37). */
void or_face_normals(int F, const double *X, const int *faces, double *fnorm) {
  for (int f = 0; f < F; ++f) {
    const int *nd = faces + 3 * f;
    v3 x0 = v3_load(X, nd[0]), x1 = v3_load(X, nd[1]), x2 = v3_load(X, nd[2]);
    v3 n = v3_scale(0.5, v3_cross(v3_sub(x2, x0), v3_sub(x1, x0)));
    for (int k = 0; k < 3; ++k) fnorm[3 * f + k] = n.v[k];
  }
}

/* SubDomain::computeWeightsLeastSquaresEdgePart (SubDomainCore.C:1334-1381)
 * and computeWeightsLeastSquaresNodePart (SubDomainCore.C:1755-1777). */
void or_lsq_weights(int N, int E, const double *X, const int *ptr, double *R) {
  for (int k = 0; k < 6 * N; ++k) R[k] = 0.0;
  for (int l = 0; l < E; ++l) {
    int i = ptr[2 * l], j = ptr[2 * l + 1];
    double dx[3];
    dx[0] = X[3 * j] - X[3 * i];
    dx[1] = X[3 * j + 1] - X[3 * i + 1];
    dx[2] = X[3 * j + 2] - X[3 * i + 2];
    double dxdx = dx[0] * dx[0], dydy = dx[1] * dx[1], dzdz = dx[2] * dx[2];
    double dxdy = dx[0] * dx[1], dxdz = dx[0] * dx[2], dydz = dx[1] * dx[2];
    R[6 * i + 0] += dxdx; R[6 * j + 0] += dxdx;
    R[6 * i + 1] += dxdy; R[6 * j + 1] += dxdy;
    R[6 * i + 2] += dxdz; R[6 * j + 2] += dxdz;
    R[6 * i + 3] += dydy; R[6 * j + 3] += dydy;
    R[6 * i + 4] += dydz; R[6 * j + 4] += dydz;
    R[6 * i + 5] += dzdz; R[6 * j + 5] += dzdz;
  }
  for (int i = 0; i < N; ++i) {
    double *r = R + 6 * i;
    double r11 = sqrt(r[0]);
    double or11 = 1.0 / r11;
    double r12 = r[1] * or11;
    double r13 = r[2] * or11;
    double r22 = sqrt(r[3] - r12 * r12);
    double r23 = (r[4] - r12 * r13) / r22;
    double r33 = sqrt(r[5] - (r13 * r13 + r23 * r23));
    r[0] = r11; r[1] = r12; r[2] = r13; r[3] = r22; r[4] = r23; r[5] = r33;
  }
}

/* ------------------------------------------------------------------------ */
/* Physics                                                                   */
/* ------------------------------------------------------------------------ */

void or_cons_to_prim(int N, const double *U, double *V, double gam, double pstiff) {
  for (int n = 0; n < N; ++n) {
    const double *u = U + 5 * n;
    double *v = V + 5 * n;
    v[0] = u[0];
    double invRho = 1.0 / u[0];
    v[1] = u[1] * invRho;
    v[2] = u[2] * invRho;
    v[3] = u[3] * invRho;
    double vel2 = v[1] * v[1] + v[2] * v[2] + v[3] * v[3];
    v[4] = (gam - 1.0) * (u[4] - 0.5 * u[0] * vel2) - gam * pstiff;
  }
}

/* computeLocalWeightsLeastSquares (SubDomain.C:137-159). */
static void lsq_local_weights(const double dx[3], const double *R, double W[3]) {
  double or11 = 1.0 / R[0];
  double or22 = 1.0 / R[3];
  double or33 = 1.0 / R[5];
  double r12or11 = R[1] * or11;
  double r23or22 = R[4] * or22;
  double psi = (R[1] * R[4] - R[2] * R[3]) * or11 * or22;
  double alpha1 = dx[0] * or11 * or11;
  double alpha2 = (dx[1] - r12or11 * dx[0]) * or22 * or22;
  double alpha3 = (dx[2] - r23or22 * dx[1] + psi * dx[0]) * or33 * or33;
  W[0] = alpha1 - r12or11 * alpha2 + psi * alpha3;
  W[1] = alpha2 - r23or22 * alpha3;
  W[2] = alpha3;
}

/* SubDomain::computeGradientsLeastSquares, non-sampled branch
 * (SubDomain.C:297-374). */
void or_gradients(int N, int E, const double *X, const int *ptr, const double *R,
                  const double *V, double *ddx, double *ddy, double *ddz) {
  for (int k = 0; k < 5 * N; ++k) ddx[k] = ddy[k] = ddz[k] = 0.0;
  for (int l = 0; l < E; ++l) {
    int i = ptr[2 * l], j = ptr[2 * l + 1];
    double Wi[3], Wj[3];
    double dx[3] = {X[3 * j] - X[3 * i], X[3 * j + 1] - X[3 * i + 1], X[3 * j + 2] - X[3 * i + 2]};
    lsq_local_weights(dx, R + 6 * i, Wi);
    dx[0] = -dx[0]; dx[1] = -dx[1]; dx[2] = -dx[2];
    lsq_local_weights(dx, R + 6 * j, Wj);
    for (int k = 0; k < 5; ++k) {
      double deltaVar = V[5 * j + k] - V[5 * i + k];
      ddx[5 * i + k] += Wi[0] * deltaVar;
      ddy[5 * i + k] += Wi[1] * deltaVar;
      ddz[5 * i + k] += Wi[2] * deltaVar;
      ddx[5 * j + k] -= Wj[0] * deltaVar;
      ddy[5 * j + k] -= Wj[1] * deltaVar;
      ddz[5 * j + k] -= Wj[2] * deltaVar;
    }
  }
}

/* RecFcn::computeVanAlbadaFcn (RecFcn.h:269-281). */
static double va_fcn(double eps, double a, double b) {
  double a2 = a * a;
  double b2 = b * b;
  if (a * b > 0.0) return (a * (b2 + eps) + b * (a2 + eps)) / (a2 + b2 + 2.0 * eps);
  else if (a * b == 0.0) return 0.5 * eps * (a + b) / (a2 + b2 + 2.0 * eps);
  else return 0.0;
}

/* RecFcn::vanalbada (RecFcn.h:433-446) with the constructor constants of
 * RecFcn.h:42-50 (beta2 = 1 - 2 beta, eps0 = 1e-16). */
void or_vanalbada5(double beta, const double *Vi, const double *ddVij, const double *Vj,
                   const double *ddVji, double *Vij, double *Vji) {
  const double beta2 = 1.0 - 2.0 * beta;
  const double eps0 = 1.e-16;
  for (int k = 0; k < 5; ++k) {
    double dVji = Vj[k] - Vi[k];
    double dV = beta2 * dVji;
    double dVi = dV + 2.0 * beta * ddVij[k];
    double dVj = dV + 2.0 * beta * ddVji[k];
    Vij[k] = Vi[k] + 0.5 * va_fcn(eps0, dVi, dVji);
    Vji[k] = Vj[k] - 0.5 * va_fcn(eps0, dVj, dVji);
  }
}

/* ROEFLUX7 (f77src/roeflux7.f:26-187), type 0.  Fortran statement order and
 * left-to-right association are kept; MAX(vp4,0.) / MIN(vp5,0.) compare with
 * a zero literal. */
void or_roeflux7(double gam, double pstiff, const double *enormal, double evitno,
                 const double *Ug, const double *Ud, double *phi) {
  double gam1 = gam - 1.0;
  double invgam1 = 1.0 / gam1;
  double rnorm = sqrt(enormal[0] * enormal[0] + enormal[1] * enormal[1] + enormal[2] * enormal[2]);
  double invnorm = 1.0 / rnorm;
  double normal[3] = {enormal[0] * invnorm, enormal[1] * invnorm, enormal[2] * invnorm};
  double vitno = evitno * invnorm;
  double vitg2 = Ug[1] * Ug[1] + Ug[2] * Ug[2] + Ug[3] * Ug[3];
  double vitd2 = Ud[1] * Ud[1] + Ud[2] * Ud[2] + Ud[3] * Ud[3];
  double squsr1 = sqrt(Ug[0]);
  double squsr2 = sqrt(Ud[0]);
  double ener1 = (Ug[4] + gam * pstiff) * invgam1 + 0.5 * Ug[0] * vitg2;
  double ener2 = (Ud[4] + gam * pstiff) * invgam1 + 0.5 * Ud[0] * vitd2;
  double usro = 1.0 / (squsr1 + squsr2);
  double uar2 = (squsr1 * Ug[1] + squsr2 * Ud[1]) * usro;
  double uar3 = (squsr1 * Ug[2] + squsr2 * Ud[2]) * usro;
  double uar4 = (squsr1 * Ug[3] + squsr2 * Ud[3]) * usro;
  double uar5 = ((ener1 + Ug[4]) / squsr1 + (ener2 + Ud[4]) / squsr2) * usro;
  double VdotNt = normal[0] * uar2 + normal[1] * uar3 + normal[2] * uar4;
  double VdotN, H;
  if (VdotNt - vitno > 0) {
    VdotN = Ug[1] * normal[0] + Ug[2] * normal[1] + Ug[3] * normal[2];
    H = gam * (Ug[4] + pstiff) + 0.5 * gam1 * Ug[0] * vitg2;
    H = H / (gam1 * Ug[0]);
    phi[0] = Ug[0] * (VdotN - vitno);
    phi[1] = phi[0] * Ug[1] + Ug[4] * normal[0];
    phi[2] = phi[0] * Ug[2] + Ug[4] * normal[1];
    phi[3] = phi[0] * Ug[3] + Ug[4] * normal[2];
    phi[4] = phi[0] * H + Ug[4] * vitno;
  } else {
    VdotN = Ud[1] * normal[0] + Ud[2] * normal[1] + Ud[3] * normal[2];
    H = gam * (Ud[4] + pstiff) + 0.5 * gam1 * Ud[0] * vitd2;
    H = H / (gam1 * Ud[0]);
    phi[0] = Ud[0] * (VdotN - vitno);
    phi[1] = phi[0] * Ud[1] + Ud[4] * normal[0];
    phi[2] = phi[0] * Ud[2] + Ud[4] * normal[1];
    phi[3] = phi[0] * Ud[3] + Ud[4] * normal[2];
    phi[4] = phi[0] * H + Ud[4] * vitno;
  }
  VdotN = VdotNt;
  double qir = 0.5 * (uar2 * uar2 + uar3 * uar3 + uar4 * uar4);
  double cr2 = gam1 * (uar5 - qir);
  double cr = sqrt(cr2);
  cr2 = 1.0 / cr2;
  double vp1 = VdotN;
  double vp4 = VdotN + cr;
  double vp5 = VdotN - cr;
  double r = cr;
  double s = -cr;
  double cr2byt = -cr;
  double invtbeta2 = -1.0 / cr;
  double dif1 = -Ug[0] + Ud[0];
  double dif2 = -Ug[0] * Ug[1] + Ud[0] * Ud[1];
  double dif3 = -Ug[0] * Ug[2] + Ud[0] * Ud[2];
  double dif4 = -Ug[0] * Ug[3] + Ud[0] * Ud[3];
  double dif5 = -ener1 + ener2;
  vp1 = (vp1 - vitno);
  vp4 = (vp4 - vitno);
  vp5 = (vp5 - vitno);
  (void)vp1;
  double flur4, flur5;
  if (VdotNt - vitno <= 0) {
    double m4 = vp4 > 0.0 ? vp4 : 0.0;
    flur4 = m4 * (cr2byt * (VdotN * dif1 - normal[0] * dif2 - normal[1] * dif3 - normal[2] * dif4) +
                  gam1 * s * invtbeta2 * (qir * dif1 - uar2 * dif2 - uar3 * dif3 - uar4 * dif4 + dif5));
    flur5 = 0.0;
  } else {
    flur4 = 0.0;
    double m5 = vp5 < 0.0 ? vp5 : 0.0;
    flur5 = -m5 * (-cr2byt * (VdotN * dif1 - normal[0] * dif2 - normal[1] * dif3 - normal[2] * dif4) +
                   (-gam1 * r * invtbeta2 * (qir * dif1 - uar2 * dif2 - uar3 * dif3 - uar4 * dif4 + dif5)));
  }
  phi[0] = phi[0] - 0.5 * (flur4 + flur5) * cr2;
  phi[1] = phi[1] - (normal[0] * (r * flur4 + s * flur5) + uar2 * (flur4 + flur5)) * 0.5 * cr2;
  phi[2] = phi[2] - (normal[1] * (r * flur4 + s * flur5) + uar3 * (flur4 + flur5)) * 0.5 * cr2;
  phi[3] = phi[3] - (normal[2] * (r * flur4 + s * flur5) + uar4 * (flur4 + flur5)) * 0.5 * cr2;
  phi[4] = phi[4] - (VdotN * (r * flur4 + s * flur5) + uar5 * (flur4 + flur5)) * 0.5 * cr2;
  for (int k = 0; k < 5; ++k) phi[k] = phi[k] * rnorm;
}

/* BOUNDFLUX5 (f77src/boundflux5.f:23-98), type 0. */
void or_boundflux5(double gam, const double *normal, double sig, const double *ua,
                   const double *ub, double *flux) {
  double fgp[5], fgm[5];
  double gam1 = gam - 1.0;
  double vnx = normal[0], vny = normal[1], vnz = normal[2];
  double cap = sqrt(vnx * vnx + vny * vny + vnz * vnz);
  double capd1 = vnx / cap, capd2 = vny / cap, capd3 = vnz / cap;
  double ro = ua[0];
  double usro = 1.0 / ro;
  double u = ua[1], v = ua[2], w = ua[3];
  double vit = u * u + v * v + w * w;
  double pm = ua[4];
  double rhoE = pm / gam1 + 0.5 * ro * vit;
  double c2 = gam * pm * usro;
  double c = sqrt(c2);
  double usc = 1.0 / c;
  double cc = 0.5 * vit + c2 / gam1;
  double unx = capd1 * u + capd2 * v + capd3 * w;
  double t1 = cap * unx - sig, t4 = cap * (unx + c) - sig, t5 = cap * (unx - c) - sig;
  double vp1 = t1 > 0.0 ? t1 : 0.0;
  double vp4 = t4 > 0.0 ? t4 : 0.0;
  double vp5 = t5 > 0.0 ? t5 : 0.0;
  double vpp = vp4 + vp5;
  double vpm = vp4 - vp5;
  fgp[0] = (gam1 * ro * vp1 + 0.5 * ro * vpp) / gam;
  fgp[1] = fgp[0] * u + 0.5 * pm * usc * capd1 * vpm;
  fgp[2] = fgp[0] * v + 0.5 * pm * usc * capd2 * vpm;
  fgp[3] = fgp[0] * w + 0.5 * pm * usc * capd3 * vpm;
  fgp[4] = 0.5 * ((gam1 / gam) * ro * vit * (vp1 - 0.5 * vpp) + rhoE * vpp + ro * unx * c * vpm / gam);
  vp1 = t1 < 0.0 ? t1 : 0.0;
  vp4 = t4 < 0.0 ? t4 : 0.0;
  vp5 = t5 < 0.0 ? t5 : 0.0;
  vpp = vp1 - 0.5 * (vp4 + vp5);
  vpm = 0.5 * (vp5 - vp4);
  double xn1 = gam1 * usc * (-0.5 * vit * ub[0] + u * ub[1] + v * ub[2] + w * ub[3] - ub[4]);
  double xn2 = unx * ub[0] - capd1 * ub[1] - capd2 * ub[2] - capd3 * ub[3];
  double ff = usc * (vpp * xn1 + vpm * xn2);
  double gg = vpm * xn1 + vpp * xn2;
  fgm[0] = vp1 * ub[0] + ff;
  fgm[1] = vp1 * ub[1] + u * ff + capd1 * gg;
  fgm[2] = vp1 * ub[2] + v * ff + capd2 * gg;
  fgm[3] = vp1 * ub[3] + w * ff + capd3 * gg;
  fgm[4] = vp1 * ub[4] + cc * ff + unx * gg;
  for (int k = 0; k < 5; ++k) flux[k] = fgp[k] + fgm[k];
}

/* FluxFcnSGWallEuler3D::compute (FluxFcnDescSG.C:779-789). */
void or_wallflux(const double *normal, double normal_vel, const double *V, double *flux) {
  flux[0] = 0.0;
  flux[1] = V[4] * normal[0];
  flux[2] = V[4] * normal[1];
  flux[3] = V[4] * normal[2];
  flux[4] = V[4] * normal_vel;
}

/* ------------------------------------------------------------------------ */
/* Boundary data and residual                                                */
/* ------------------------------------------------------------------------ */

/* DistBcDataEuler::setBoundaryConditionsGas (DistBcData.C:652-735) with zero
 * gravity, identical inlet and outlet data: the per-node farfield states
 * uin/uout. */
void or_freestream(const or_params *p, double *Uin, double *Uout) {
  double gam = p->gamma, Pstiff = p->pstiff;
  double rhoin = p->rho_inf, pressurein = p->p_inf;
  double velin2 = gam * (pressurein + Pstiff) * p->mach * p->mach / rhoin;
  double velin = sqrt(velin2);
  double U[5];
  U[0] = rhoin;
  U[1] = U[0] * velin * cos(p->alpha) * cos(p->yaw);
  U[2] = U[0] * velin * cos(p->alpha) * sin(p->yaw);
  U[3] = U[0] * velin * sin(p->alpha);
  /* per-node value (DistBcData.C:722): ptempin = pressurein + rhoin*0 */
  double ptempin = pressurein + rhoin * 0.0;
  U[4] = 0.5 * rhoin * velin2 + (ptempin + gam * Pstiff) / (gam - 1.0);
  for (int k = 0; k < 5; ++k) { Uin[k] = U[k]; Uout[k] = U[k]; }
}

/* Face::assignFreeStreamValues2 (Face.C:38-70) for a 3-node face whose nodes
 * all carry the same farfield state. */
static void face_state(const double *Unode, double *Ub) {
  for (int k = 0; k < 5; ++k) {
    Ub[k] = 0.0;
    for (int j = 0; j < 3; ++j) Ub[k] += Unode[k];
    Ub[k] /= 3;
  }
}

int or_setup(or_mesh *m, const or_params *p) {
  int N = m->N, T = m->T, F = m->F;
  m->ptr = (int *)malloc(sizeof(int) * 12 * (size_t)T + 8);
  m->tet_edges = (int *)malloc(sizeof(int) * 6 * (size_t)T + 8);
  m->E = or_edges(N, T, m->tets, m->ptr, m->tet_edges);
  int E = m->E;
  m->cv = (double *)malloc(sizeof(double) * (size_t)N);
  m->normals = (double *)malloc(sizeof(double) * 3 * (size_t)E);
  m->normal_vel = (double *)malloc(sizeof(double) * (size_t)E);
  m->fnorm = (double *)malloc(sizeof(double) * 3 * (size_t)F + 8);
  m->R = (double *)malloc(sizeof(double) * 6 * (size_t)N);
  or_control_volumes(N, T, m->X, m->tets, m->cv);
  or_edge_normals(T, m->X, m->tets, m->tet_edges, E, m->normals, m->normal_vel);
  or_face_normals(F, m->X, m->faces, m->fnorm);
  or_lsq_weights(N, E, m->X, m->ptr, m->R);
  or_freestream(p, m->Uin, m->Uout);
  face_state(m->Uin, m->Ub_in);
  face_state(m->Uout, m->Ub_out);
  return 0;
}

void or_free(or_mesh *m) {
  free(m->ptr); free(m->tet_edges); free(m->cv); free(m->normals);
  free(m->normal_vel); free(m->fnorm); free(m->R);
  m->ptr = m->tet_edges = 0;
  m->cv = m->normals = m->normal_vel = m->fnorm = m->R = 0;
}

/* SpaceOperator<5>::computeResidual (SpaceOperator.C:524-634):
 *   R = 0; V = U->V (:530); LSQ gradients (:532-536); limit = no-op for
 *   Van Albada without failsafe tags (:578-580, DistNodalGrad.C:1070-1130);
 *   Domain::computeFiniteVolumeTerm (Domain.C:847-936) = EdgeSet then
 *   FaceSet (SubDomain.C:1382-1398); r *= 1/cv (:599-624). */
int or_residual(const or_mesh *m, const or_params *p, const double *U, double *R,
                int fluxes_only, double *V_out, double *ddx_out, double *ddy_out,
                double *ddz_out) {
  return or_residual_ex(m, p, U, R, 0, fluxes_only, V_out, ddx_out, ddy_out, ddz_out);
}

static int residual_pass(const or_mesh *m, const or_params *p, const double *U, double *R, double *Rabs,
                         int fluxes_only, double *V_out, double *ddx_out, double *ddy_out, double *ddz_out,
                         int *tag0, const int *mask);

int or_residual_ex(const or_mesh *m, const or_params *p, const double *U, double *R,
                   double *Rabs, int fluxes_only, double *V_out, double *ddx_out,
                   double *ddy_out, double *ddz_out) {
  return residual_pass(m, p, U, R, Rabs, fluxes_only, V_out, ddx_out, ddy_out, ddz_out, 0, 0);
}

/* SpaceOperator::computeResidual with Newton.FailSafe (failsafe 1 On, 2
 * AlwaysOn) and Domain::computeFiniteVolumeTerm (Domain.C:847-917).  tags
 * [3][N] carries Domain's tag (tag0, tag1) and DistNodalGrad's node tags
 * (first order: gradients zeroed by limit, DistNodalGrad.C:1083-1104) from
 * call to call.  A pass with non-physical reconstructions tags their sides
 * (EdgeSet::checkReconstructedValues, EdgeCore.C:137-166), finalizeTags
 * (SubDomainCore.C:5177-5188) sets tag1 over the edges in order, the nodes
 * with tag1 become first order (DistNodalGrad::fix) and the pass is
 * re-run; On clears Domain's tag afterwards.  Returns the re-run's count
 * (the first pass's in *ierr_first). */
int or_residual_fs(const or_mesh *m, const or_params *p, const double *U, double *R, int failsafe, int *tags,
                   int *ierr_first) {
  int N = m->N, E = m->E;
  int *tag0 = tags, *tag1 = tags + N, *mask = tags + 2 * N;
  int on = failsafe && p->second_order;
  int ierr = residual_pass(m, p, U, R, 0, 0, 0, 0, 0, 0, on ? tag0 : 0, mask);
  *ierr_first = ierr;
  if (!ierr || !on) return ierr;
  for (int l = 0; l < E; ++l) {
    int i = m->ptr[2 * l], j = m->ptr[2 * l + 1];
    tag1[i] = tag0[i] + tag0[j];
    tag1[j] = tag1[i];
  }
  for (int i = 0; i < N; ++i) mask[i] = mask[i] || tag1[i];
  ierr = residual_pass(m, p, U, R, 0, 0, 0, 0, 0, 0, 0, mask);
  if (failsafe == 1)
    for (int i = 0; i < 2 * N; ++i) tags[i] = 0;
  return ierr;
}

static int residual_pass(const or_mesh *m, const or_params *p, const double *U, double *R, double *Rabs,
                         int fluxes_only, double *V_out, double *ddx_out, double *ddy_out, double *ddz_out,
                         int *tag0, const int *mask) {
  int N = m->N, E = m->E, F = m->F;
  double gam = p->gamma, pstiff = p->pstiff;
  double *V = V_out ? V_out : (double *)malloc(sizeof(double) * 5 * (size_t)N);
  double *ddx = 0, *ddy = 0, *ddz = 0;
  for (int k = 0; k < 5 * N; ++k) R[k] = 0.0;
  if (Rabs) for (int k = 0; k < 5 * N; ++k) Rabs[k] = 0.0;
  or_cons_to_prim(N, U, V, gam, pstiff);
  /* FemEquationTermNS: the Galerkin term comes first (SpaceOperator.C:570-573) */
  if (p->viscous) or_galerkin_term(m, p, V, R, Rabs);
  if (p->second_order) {
    ddx = ddx_out ? ddx_out : (double *)malloc(sizeof(double) * 5 * (size_t)N);
    ddy = ddy_out ? ddy_out : (double *)malloc(sizeof(double) * 5 * (size_t)N);
    ddz = ddz_out ? ddz_out : (double *)malloc(sizeof(double) * 5 * (size_t)N);
    or_gradients(N, E, m->X, m->ptr, m->R, V, ddx, ddy, ddz);
    if (mask)
      for (int i = 0; i < N; ++i)
        if (mask[i])
          for (int k = 0; k < 5; ++k) ddx[5 * i + k] = ddy[5 * i + k] = ddz[5 * i + k] = 0.0;
  }
  /* EdgeSet::computeFiniteVolumeTerm (Edge.C:372-439) */
  int ierr = 0;
  for (int l = 0; l < E; ++l) {
    int i = m->ptr[2 * l], j = m->ptr[2 * l + 1];
    const double *X = m->X;
    double dx[3] = {X[3 * j] - X[3 * i], X[3 * j + 1] - X[3 * i + 1], X[3 * j + 2] - X[3 * i + 2]};
    double Vi[5], Vj[5], flux[5];
    const double *vi = V + 5 * i, *vj = V + 5 * j;
    if (p->second_order) {
      double ddVij[5], ddVji[5];
      for (int k = 0; k < 5; ++k) {
        ddVij[k] = dx[0] * ddx[5 * i + k] + dx[1] * ddy[5 * i + k] + dx[2] * ddz[5 * i + k];
        ddVji[k] = dx[0] * ddx[5 * j + k] + dx[1] * ddy[5 * j + k] + dx[2] * ddz[5 * j + k];
      }
      or_vanalbada5(p->beta, vi, ddVij, vj, ddVji, Vi, Vj);
    } else {
      /* RecFcnConstant: Vij = Vi, Vji = Vj */
      for (int k = 0; k < 5; ++k) { Vi[k] = vi[k]; Vj[k] = vj[k]; }
    }
    /* checkReconstructedValues (EdgeCore.C:137-166, VarFcnSGEuler.h:90-110) */
    if (Vi[0] <= 0.0 || Vi[4] + pstiff <= 0.0) { ++ierr; if (tag0) tag0[i] = 1; }
    if (Vj[0] <= 0.0 || Vj[4] + pstiff <= 0.0) { ++ierr; if (tag0) tag0[j] = 1; }
    if (ierr) continue;
    const double *nl = m->normals + 3 * l;
    if (sqrt(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]) < 1e-18) continue;
    if (p->low_mach) {
      /* FluxFcnSGApprJacRoeEuler3D::compute (FluxFcnDescSG.C:125-139) with
       * the spatial preconditioner on; length = |X_j - X_i| (Edge.C:400),
       * edge 1/Re = mean of the nodes' (Edge.C:412) */
      double length = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
      or_roeflux5(1.0, gam, pstiff, nl, m->normal_vel[l], Vi, Vj, p->lm_min_mach, p->lm_slope, p->lm_cutoff,
                  p->lm_shock, p->lm_irey, length, 1, flux);
    } else {
      or_roeflux7(gam, pstiff, nl, m->normal_vel[l], Vi, Vj, flux);
    }
    for (int k = 0; k < 5; ++k) {
      R[5 * i + k] += flux[k];
      R[5 * j + k] -= flux[k];
    }
    if (Rabs)
      for (int k = 0; k < 5; ++k) {
        Rabs[5 * i + k] += fabs(flux[k]);
        Rabs[5 * j + k] += fabs(flux[k]);
      }
  }
  /* FaceSet::computeFiniteVolumeTerm (Face.C:1186-1215) and
   * Face::computeFiniteVolumeTerm (Face.C:356-389): per face node,
   * normal = third*faceNorm (FaceTriaCore.C:288-292). */
  const double third = 1.0 / 3.0;
  for (int f = 0; f < F; ++f) {
    int code = m->codes[f];
    double n[3] = {third * m->fnorm[3 * f], third * m->fnorm[3 * f + 1], third * m->fnorm[3 * f + 2]};
    double nvel = third * 0.0;
    for (int a = 0; a < 3; ++a) {
      int node = m->faces[3 * f + a];
      double flux[5];
      /* every wall code takes the wall flux (FluxFcn.h:270-276) */
      if (code == 1 || code == -1 || code == 2 || code == -2 || code == 3 || code == -3)
        or_wallflux(n, nvel, V + 5 * node, flux);
      else if (code == 4 || code == -4) or_boundflux5(gam, n, nvel, V + 5 * node, m->Ub_in, flux);
      else if (code == 5 || code == -5) or_boundflux5(gam, n, nvel, V + 5 * node, m->Ub_out, flux);
      else continue;
      for (int k = 0; k < 5; ++k) R[5 * node + k] += flux[k];
      if (Rabs) for (int k = 0; k < 5; ++k) Rabs[5 * node + k] += fabs(flux[k]);
    }
  }
  if (!fluxes_only)
    for (int i = 0; i < N; ++i) {
      double invcv = 1.0 / m->cv[i];
      for (int k = 0; k < 5; ++k) R[5 * i + k] *= invcv;
      if (Rabs) for (int k = 0; k < 5; ++k) Rabs[5 * i + k] *= invcv;
    }
  if (!V_out) free(V);
  if (ddx && !ddx_out) free(ddx);
  if (ddy && !ddy_out) free(ddy);
  if (ddz && !ddz_out) free(ddz);
  return ierr;
}

/* ------------------------------------------------------------------------ */
/* RCM (ConnectivityCore.C:311-635)                                          */
/* ------------------------------------------------------------------------ */

typedef struct { int size; int *pointer; int *target; } conn;

/* SubDomain::createEdgeBasedConnectivity (SubDomainCore.C:253-320): rows
 * hold the node itself plus its edge neighbours, sorted ascending. */
static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}
static void build_conn(int N, int E, const int *ptr, conn *c) {
  c->size = N;
  c->pointer = (int *)malloc(sizeof(int) * (size_t)(N + 1));
  c->target = (int *)malloc(sizeof(int) * (size_t)(2 * E + N));
  int *cnt = (int *)malloc(sizeof(int) * (size_t)N);
  for (int i = 0; i < N; ++i) cnt[i] = 1;
  for (int l = 0; l < E; ++l) { ++cnt[ptr[2 * l]]; ++cnt[ptr[2 * l + 1]]; }
  c->pointer[0] = 0;
  for (int i = 0; i < N; ++i) c->pointer[i + 1] = c->pointer[i] + cnt[i];
  for (int i = 0; i < N; ++i) { c->target[c->pointer[i]] = i; cnt[i] = 1; }
  for (int l = 0; l < E; ++l) {
    int a = ptr[2 * l], b = ptr[2 * l + 1];
    c->target[c->pointer[a] + cnt[a]++] = b;
    c->target[c->pointer[b] + cnt[b]++] = a;
  }
  for (int i = 0; i < N; ++i)
    qsort(c->target + c->pointer[i], (size_t)(c->pointer[i + 1] - c->pointer[i]), sizeof(int), cmp_int);
  free(cnt);
}

static int conn_num_mask(const conn *c, int nd, const int *mask) {
  int res = 0;
  for (int j = c->pointer[nd]; j < c->pointer[nd + 1]; ++j)
    if (mask[c->target[j]]) res++;
  return res;
}

/* Connectivity::rootLS (ConnectivityCore.C:376-413). */
static int root_ls(const conn *c, int root, int *xls, int *ls, int *w, const int *mask, int *locMask) {
  int size = c->size;
  *w = 0;
  if (mask) for (int i = 0; i < size; ++i) locMask[i] = mask[i];
  else for (int i = 0; i < size; ++i) locMask[i] = 1;
  locMask[root] = 0;
  ls[0] = root;
  xls[0] = 0;
  int nlvl = 1, nf = 1;
  while (nf > xls[nlvl - 1]) {
    xls[nlvl] = nf;
    int lbegin = xls[nlvl - 1], lend = xls[nlvl];
    for (int i = lbegin; i < lend; ++i) {
      int n1 = ls[i];
      for (int j = c->pointer[n1]; j < c->pointer[n1 + 1]; ++j) {
        int n2 = c->target[j];
        if (locMask[n2]) { locMask[n2] = 0; ls[nf++] = n2; }
      }
    }
    if (nf - xls[nlvl] > *w) *w = nf - xls[nlvl];
    nlvl++;
  }
  return nlvl - 1;
}

/* Connectivity::findPseudoDiam (ConnectivityCore.C:323-374). */
static void find_pseudo_diam(const conn *c, int *s, int *e, const int *mask, int *work) {
  int size = c->size;
  int numtarget = c->pointer[size];
  int cmin = numtarget + 1, cmax = 0, nw;
  for (int i = 0; i < size; ++i)
    if (mask[i]) {
      if ((nw = conn_num_mask(c, i, mask)) < cmin) { cmin = nw; *s = i; }
      if (nw > cmax) cmax = nw;
    }
  int *ls = (int *)malloc(sizeof(int) * (size_t)numtarget);
  int *xls = (int *)malloc(sizeof(int) * (size_t)(numtarget + 1));
  int *sorted = (int *)malloc(sizeof(int) * (size_t)(cmax + 1));
  int w;
  int h = root_ls(c, *s, xls, ls, &w, mask, work);
  int subconsize = xls[h];
  *e = ls[xls[h - 1]];
  while (h < subconsize) {
    int k;
    for (k = 0; k <= cmax; ++k) sorted[k] = -1;
    int kstrt = xls[h - 1], kstop = xls[h];
    for (k = kstrt; k < kstop; ++k) sorted[conn_num_mask(c, ls[k], mask)] = ls[k];
    int w_e = subconsize;
    for (k = 0; k <= cmax; ++k)
      if (sorted[k] >= 0) {
        int nh = root_ls(c, sorted[k], xls, ls, &w, mask, work);
        if (w < w_e) {
          if (nh > h) { *s = sorted[k]; h = nh; break; }
          *e = sorted[k];
          w_e = w;
        }
      }
    if (k > cmax) break;
  }
  free(ls); free(xls); free(sorted);
}

/* Connectivity::renumRCM (ConnectivityCore.C:590-635).  The degree "sort"
 * at :620-628 assigns order[k] twice and leaves the order unchanged, so it
 * is omitted here. */
static void renum_rcm(const conn *c, int *mask, int *nextNum, int *renum, int *work) {
  int s_node = 0, e_node = 0;
  find_pseudo_diam(c, &s_node, &e_node, mask, work);
  int size = c->size;
  int *order = (int *)malloc(sizeof(int) * (size_t)size);
  order[0] = e_node;
  mask[e_node] = -mask[e_node];
  int lastNode = 1;
  for (int i = 0; i < lastNode; ++i) {
    int curNode = order[i];
    for (int j = c->pointer[curNode]; j < c->pointer[curNode + 1]; ++j) {
      int nb = c->target[j];
      if (mask[nb] > 0) { order[lastNode] = nb; mask[nb] = -mask[nb]; lastNode += 1; }
    }
  }
  for (int i = lastNode; i--;) renum[order[i]] = (*nextNum)++;
  free(order);
}

/* Connectivity::renumByComponent(renumAlg = 2) (ConnectivityCore.C:415-588). */
void or_rcm(int N, int E, const int *ptr, int *renum) {
  conn c;
  build_conn(N, E, ptr, &c);
  int *mark = (int *)malloc(sizeof(int) * (size_t)N);
  int *ls = (int *)malloc(sizeof(int) * (size_t)N);
  int *xls = (int *)malloc(sizeof(int) * (size_t)(N + 1));
  int *locMask = (int *)malloc(sizeof(int) * (size_t)N);
  int *lrenum = (int *)malloc(sizeof(int) * (size_t)N);
  int *work = (int *)malloc(sizeof(int) * (size_t)N);
  for (int i = 0; i < N; ++i) {
    mark[i] = (c.pointer[i + 1] - c.pointer[i] != 0) ? 1 : 0;
    renum[i] = -1;
    locMask[i] = 0;
  }
  int nextNum = 0;
  for (int inode = 0; inode < N; ++inode) {
    if (mark[inode] == 1) {
      int w;
      int h = root_ls(&c, inode, xls, ls, &w, 0, work);
      for (int j = 0; j < xls[h]; ++j) { locMask[ls[j]] = 1; mark[ls[j]] = 0; }
      renum_rcm(&c, locMask, &nextNum, lrenum, work);
      for (int j = 0; j < xls[h]; ++j) locMask[ls[j]] = 0;
      for (int j = 0; j < xls[h]; ++j) { int k = ls[j]; renum[k] = lrenum[k]; }
    }
  }
  free(mark); free(ls); free(xls); free(locMask); free(lrenum); free(work);
  free(c.pointer); free(c.target);
}

/* ------------------------------------------------------------------------ */
/* Edge colouring (no reference counterpart)                                 */
/* ------------------------------------------------------------------------ */

/* Priority hash of an edge id: the "lowbias32" integer mixer
 * (x ^= x>>16; x *= 0x7feb352d; x ^= x>>15; x *= 0x846ca68b; x ^= x>>16). */
uint32_t or_edge_priority(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

typedef struct { uint64_t key; int l; } keyed;
static int cmp_keyed_desc(const void *a, const void *b) {
  uint64_t x = ((const keyed *)a)->key, y = ((const keyed *)b)->key;
  return (x < y) - (x > y);
}

/* Rule: edges conflict when they share a node.  Each edge's key is
 * (priority(l) << 32 | l); an edge takes the smallest colour not used by any
 * conflicting edge of larger key.  This equals greedy first-fit colouring in
 * decreasing key order, which is what this sequential restatement does; a
 * parallel Jones-Plassmann sweep yields the same table regardless of
 * scheduling.  Returns the colour count, or -1 if more than 64 are needed. */
int or_color_edges(int N, int E, const int *ptr, int *color) {
  keyed *ord = (keyed *)malloc(sizeof(keyed) * (size_t)(E > 0 ? E : 1));
  for (int l = 0; l < E; ++l) {
    ord[l].key = ((uint64_t)or_edge_priority((uint32_t)l) << 32) | (uint32_t)l;
    ord[l].l = l;
  }
  qsort(ord, (size_t)E, sizeof(keyed), cmp_keyed_desc);
  uint64_t *used = (uint64_t *)calloc((size_t)N, sizeof(uint64_t));
  int ncol = 0;
  for (int k = 0; k < E; ++k) {
    int l = ord[k].l;
    int i = ptr[2 * l], j = ptr[2 * l + 1];
    uint64_t u = used[i] | used[j];
    if (u == ~(uint64_t)0) { free(ord); free(used); return -1; }
    int cidx = 0;
    while (u & ((uint64_t)1 << cidx)) ++cidx;
    color[l] = cidx;
    used[i] |= (uint64_t)1 << cidx;
    used[j] |= (uint64_t)1 << cidx;
    if (cidx + 1 > ncol) ncol = cidx + 1;
  }
  free(ord);
  free(used);
  return ncol;
}

/* ------------------------------------------------------------------------ */
/* Implicit operators                                                        */
/* ------------------------------------------------------------------------ */

/* ROEJAC6 (f77src/roejac6.f:1-316) with type 0 and prec 0 (beta = 1):
 * 0.5*|n|*(A(Ug) +/- |A_roe|), row-major 5x5 into jac. gamma_diss is the
 * Schemes.NavierStokes.Gamma dissipation factor, gam the heat ratio. */
void or_roejac6(double gamma_diss, double gam, double pstiff, const double *enormal,
                double evitno, const double *Ug, const double *Ud, int sw, double *jac) {
  or_roejac6_lm(gamma_diss, gam, pstiff, enormal, evitno, Ug, Ud, sw, 1.0, 0.0, 1.0, 0.0, 0.0, 1.0, 0, jac);
}

/* ROEJAC6 with the spatial low-Mach constants (roejac6.f:130-160): with prec
 * the Roe-Turkel beta of the edge replaces 1 (min Mach "mach", slope k1,
 * cut-off cmach, shock reducer, 1/Re irey, edge length). */
void or_roejac6_lm(double gamma_diss, double gam, double pstiff, const double *enormal, double evitno,
                   const double *Ug, const double *Ud, int sw, double mach, double k1, double cmach,
                   double shockreducer, double irey, double length, int prec, double *jac) {
  double H[5][5], tmw[5][5], dmat[5][5];
  double gam1 = gam - 1.0;
  double rnorm = sqrt(enormal[0] * enormal[0] + enormal[1] * enormal[1] + enormal[2] * enormal[2]);
  double invnorm = 1.0 / rnorm;
  double n1 = enormal[0] * invnorm, n2 = enormal[1] * invnorm, n3 = enormal[2] * invnorm;
  double vitno = evitno * invnorm;
  double VdotN = Ug[1] * n1 + Ug[2] * n2 + Ug[3] * n3;
  double qir = 0.5 * (Ug[1] * Ug[1] + Ug[2] * Ug[2] + Ug[3] * Ug[3]);
  double enth = gam * (Ug[4] + pstiff) + gam1 * Ug[0] * qir;
  enth = enth / (gam1 * Ug[0]);
  H[0][0] = -vitno; H[0][1] = n1; H[0][2] = n2; H[0][3] = n3; H[0][4] = 0.0;
  H[1][0] = gam1 * n1 * qir - Ug[1] * VdotN;
  H[1][1] = (1.0 - gam1) * n1 * Ug[1] + VdotN - vitno;
  H[1][2] = -gam1 * n1 * Ug[2] + Ug[1] * n2;
  H[1][3] = -gam1 * n1 * Ug[3] + Ug[1] * n3;
  H[1][4] = gam1 * n1;
  H[2][0] = gam1 * n2 * qir - Ug[2] * VdotN;
  H[2][1] = -gam1 * n2 * Ug[1] + Ug[2] * n1;
  H[2][2] = (1.0 - gam1) * n2 * Ug[2] + VdotN - vitno;
  H[2][3] = -gam1 * n2 * Ug[3] + Ug[2] * n3;
  H[2][4] = gam1 * n2;
  H[3][0] = gam1 * n3 * qir - Ug[3] * VdotN;
  H[3][1] = -gam1 * n3 * Ug[1] + Ug[3] * n1;
  H[3][2] = -gam1 * n3 * Ug[2] + Ug[3] * n2;
  H[3][3] = (1.0 - gam1) * n3 * Ug[3] + VdotN - vitno;
  H[3][4] = gam1 * n3;
  H[4][0] = (gam1 * qir - enth) * VdotN;
  H[4][1] = enth * n1 - gam1 * Ug[1] * VdotN;
  H[4][2] = enth * n2 - gam1 * Ug[2] * VdotN;
  H[4][3] = enth * n3 - gam1 * Ug[3] * VdotN;
  H[4][4] = (1.0 + gam1) * VdotN - vitno;
  double squsr1 = sqrt(Ug[0]);
  double squsr2 = sqrt(Ud[0]);
  double ener1 = (Ug[4] + gam * pstiff) / gam1 + 0.5 * Ug[0] * (Ug[1] * Ug[1] + Ug[2] * Ug[2] + Ug[3] * Ug[3]);
  double ener2 = (Ud[4] + gam * pstiff) / gam1 + 0.5 * Ud[0] * (Ud[1] * Ud[1] + Ud[2] * Ud[2] + Ud[3] * Ud[3]);
  double usro = 1.0 / (squsr1 + squsr2);
  double uar2 = (squsr1 * Ug[1] + squsr2 * Ud[1]) * usro;
  double uar3 = (squsr1 * Ug[2] + squsr2 * Ud[2]) * usro;
  double uar4 = (squsr1 * Ug[3] + squsr2 * Ud[3]) * usro;
  double uar5 = ((ener1 + Ug[4]) / squsr1 + (ener2 + Ud[4]) / squsr2) * usro;
  enth = uar5;
  VdotN = n1 * uar2 + n2 * uar3 + n3 * uar4;
  qir = 0.5 * (uar2 * uar2 + uar3 * uar3 + uar4 * uar4);
  double cr2 = gam1 * (uar5 - qir);
  double cr = sqrt(cr2);
  cr2 = 1.0 / cr2;
  double gc = gam1 / (cr * cr);
  double usc2 = 0.5 / (cr * cr);
  double beta = 1.0;
  if (prec) {
    double shock = fabs(Ug[4] - Ud[4]) / (Ug[4] + Ud[4]) / length;
    double locMach = sqrt(2.0 * qir * cr2);
    beta = fmax(k1 * locMach, mach);
    beta = (1.0 + sqrt(irey)) * beta + shockreducer * shock;
    beta = fmin(beta, cmach);
  }
  double betaRT2 = beta * beta;
  double vp1 = VdotN;
  double t = (1.0 - betaRT2) * VdotN;
  double root = sqrt(t * t + 4.0 * betaRT2 * (cr * cr));
  double vp4 = 0.5 * ((1.0 + betaRT2) * VdotN + root);
  double vp5 = 0.5 * ((1.0 + betaRT2) * VdotN - root);
  double rRT = vp4 - vp1 * betaRT2;
  double sRT = vp5 - vp1 * betaRT2;
  double tRT = 0.5 * (vp5 - vp4);
  vp1 = fabs(vp1 - vitno) * gamma_diss;
  vp4 = fabs(vp4 - vitno) * gamma_diss;
  vp5 = fabs(vp5 - vitno) * gamma_diss;
  double bt = betaRT2 * tRT;
  double cc = cr * cr;
  tmw[0][0] = n1 + n2 * uar4 - uar3 * n3 - gc * n1 * qir;
  tmw[0][1] = gc * n1 * uar2;
  tmw[0][2] = gc * n1 * uar3 + n3;
  tmw[0][3] = gc * n1 * uar4 - n2;
  tmw[0][4] = -gc * n1;
  tmw[1][0] = n2 + n3 * uar2 - uar4 * n1 - gc * n2 * qir;
  tmw[1][1] = gc * n2 * uar2 - n3;
  tmw[1][2] = gc * n2 * uar3;
  tmw[1][3] = gc * n2 * uar4 + n1;
  tmw[1][4] = -gc * n2;
  tmw[2][0] = n3 + n1 * uar3 - uar2 * n2 - gc * n3 * qir;
  tmw[2][1] = gc * n3 * uar2 + n2;
  tmw[2][2] = gc * n3 * uar3 - n1;
  tmw[2][3] = gc * n3 * uar4;
  tmw[2][4] = -gc * n3;
  tmw[3][0] = gam1 * qir * sRT / bt + VdotN * cc / tRT;
  tmw[3][1] = -gam1 * uar2 * sRT / bt - cc * n1 / tRT;
  tmw[3][2] = -gam1 * uar3 * sRT / bt - cc * n2 / tRT;
  tmw[3][3] = -gam1 * uar4 * sRT / bt - cc * n3 / tRT;
  tmw[3][4] = gam1 * sRT / bt;
  tmw[4][0] = -gam1 * qir * rRT / bt - VdotN * cc / tRT;
  tmw[4][1] = gam1 * uar2 * rRT / bt + cc * n1 / tRT;
  tmw[4][2] = gam1 * uar3 * rRT / bt + cc * n2 / tRT;
  tmw[4][3] = gam1 * uar4 * rRT / bt + cc * n3 / tRT;
  tmw[4][4] = -gam1 * rRT / bt;
  for (int j = 0; j < 5; ++j) {
    dmat[0][j] = vp1 * (n1 * tmw[0][j] + n2 * tmw[1][j] + n3 * tmw[2][j]) + vp4 * usc2 * tmw[3][j] +
                 vp5 * usc2 * tmw[4][j];
    dmat[1][j] = vp1 * (n1 * uar2 * tmw[0][j] + (n2 * uar2 - n3) * tmw[1][j] + (n3 * uar2 + n2) * tmw[2][j]) +
                 vp4 * (usc2 * (uar2 + rRT * n1)) * tmw[3][j] + vp5 * (usc2 * (uar2 + sRT * n1)) * tmw[4][j];
    dmat[2][j] = vp1 * (n2 * uar3 * tmw[1][j] + (n1 * uar3 + n3) * tmw[0][j] + (n3 * uar3 - n1) * tmw[2][j]) +
                 vp4 * (usc2 * (uar3 + rRT * n2)) * tmw[3][j] + vp5 * (usc2 * (uar3 + sRT * n2)) * tmw[4][j];
    dmat[3][j] = vp1 * (n3 * uar4 * tmw[2][j] + (n1 * uar4 - n2) * tmw[0][j] + (n2 * uar4 + n1) * tmw[1][j]) +
                 vp4 * (usc2 * (uar4 + rRT * n3)) * tmw[3][j] + vp5 * (usc2 * (uar4 + sRT * n3)) * tmw[4][j];
    dmat[4][j] = vp1 * ((n1 * qir + (n3 * uar3 - n2 * uar4)) * tmw[0][j] +
                        (n2 * qir + (n1 * uar4 - n3 * uar2)) * tmw[1][j] +
                        (n3 * qir + (n2 * uar2 - n1 * uar3)) * tmw[2][j]) +
                 vp4 * (usc2 * (enth + VdotN * rRT)) * tmw[3][j] + vp5 * (usc2 * (enth + VdotN * sRT)) * tmw[4][j];
  }
  for (int i = 0; i < 5; ++i)
    for (int k = 0; k < 5; ++k) {
      if (sw == 1) H[i][k] = H[i][k] + dmat[i][k];
      else if (sw == 2) H[i][k] = H[i][k] - dmat[i][k];
    }
  for (int i = 0; i < 5; ++i)
    for (int k = 0; k < 5; ++k) jac[5 * i + k] = 0.5 * H[i][k] * rnorm;
}

/* VarFcnSGEuler::computedVdU (VarFcnSGEuler.h:588-603) with
 * VarFcnBase::postMultiplyBydVdU (VarFcnBase.h:383-411): res = mat * dV/dU. */
void or_post_dVdU(double gam, const double *V, const double *mat, double *res) {
  double gam1 = gam - 1.0;
  double d[25];
  double invrho = 1.0 / V[0];
  d[0] = 1.0;
  d[5] = -invrho * V[1];
  d[6] = invrho;
  d[10] = -invrho * V[2];
  d[12] = invrho;
  d[15] = -invrho * V[3];
  d[18] = invrho;
  d[20] = gam1 * 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]);
  d[21] = -gam1 * V[1];
  d[22] = -gam1 * V[2];
  d[23] = -gam1 * V[3];
  d[24] = gam1;
  for (int r = 0; r < 5; ++r) {
    const double *m = mat + 5 * r;
    double *o = res + 5 * r;
    o[0] = m[0] * d[0] + m[1] * d[5] + m[2] * d[10] + m[3] * d[15] + m[4] * d[20];
    o[1] = m[1] * d[6] + m[4] * d[21];
    o[2] = m[2] * d[12] + m[4] * d[22];
    o[3] = m[3] * d[18] + m[4] * d[23];
    o[4] = m[4] * d[24];
  }
}

/* FluxFcnSGWallEuler3D::computeJacobian (FluxFcnDescSG.C:838-886). */
void or_wall_jac(double gam, const double *n, double nvel, const double *V, double *jac) {
  double dfdV[25];
  for (int k = 0; k < 25; ++k) dfdV[k] = 0.0;
  dfdV[9] = n[0];
  dfdV[14] = n[1];
  dfdV[19] = n[2];
  dfdV[24] = nvel;
  or_post_dVdU(gam, V, dfdV, jac);
}

/* FluxFcnFD<5>::computeJacobian (FluxFcnBase.h:112-159) over BOUNDFLUX5
 * (FluxFcnSGInflow/OutflowEuler3D), then postMultiplyBydVdU. */
void or_farfield_jac(double gam, const double *n, double sig, const double *VL, const double *Ub,
                     double *jac) {
  const double eps0 = 1.e-6;
  double Veps[5], flux[5], fluxeps[5], dfdVL[25];
  or_boundflux5(gam, n, sig, VL, Ub, flux);
  for (int k = 0; k < 5; ++k) Veps[k] = VL[k];
  for (int k = 0; k < 5; ++k) {
    double eps;
    if (fabs(Veps[k]) < 1.e-10) eps = eps0;
    else eps = eps0 * Veps[k];
    double inveps = 1.0 / eps;
    Veps[k] += eps;
    if (k != 0) Veps[k - 1] = VL[k - 1];
    or_boundflux5(gam, n, sig, Veps, Ub, fluxeps);
    for (int j = 0; j < 5; ++j) dfdVL[5 * j + k] = (fluxeps[j] - flux[j]) * inveps;
  }
  or_post_dVdU(gam, VL, dfdVL, jac);
}

/* Local time step (Edge.C:87-143, Face.C:242-290, Domain.C:32-100) with
 * the default time low-Mach constants (beta = 1, no viscous term):
 * idti_i = sum of min(0.5(+-coeff1 - coeff2), 0)*S, dt_i = cfl*V_i/(-idti_i). */
void or_timestep(const or_mesh *m, const or_params *p, const double *V, double cfl, double *dt) {
  int N = m->N, E = m->E, F = m->F;
  double gam = p->gamma, pstiff = p->pstiff;
  double *idti = (double *)malloc(sizeof(double) * (size_t)N);
  for (int i = 0; i < N; ++i) idti[i] = 0.0;
  const double locbeta = 1.0, beta2 = locbeta * locbeta;
  for (int l = 0; l < E; ++l) {
    int i = m->ptr[2 * l], j = m->ptr[2 * l + 1];
    const double *nl = m->normals + 3 * l;
    double S = sqrt(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]);
    double invS = 1.0 / S;
    double n[3] = {invS * nl[0], invS * nl[1], invS * nl[2]};
    double ndot = invS * m->normal_vel[l];
    double Vmid[5];
    for (int k = 0; k < 5; ++k) Vmid[k] = 0.5 * (V[5 * i + k] + V[5 * j + k]);
    double a = sqrt(gam * (Vmid[4] + pstiff) / Vmid[0]);
    double un = Vmid[1] * n[0] + Vmid[2] * n[1] + Vmid[3] * n[2] - ndot;
    double coeff1 = (1.0 + beta2) * un;
    double coeff2 = pow(pow((1.0 - beta2) * un, 2.0) + pow(2.0 * locbeta * a, 2.0), 0.5);
    double ci = 0.5 * (coeff1 - coeff2), cj = 0.5 * (-coeff1 - coeff2);
    idti[i] += (ci < 0.0 ? ci : 0.0) * S;
    idti[j] += (cj < 0.0 ? cj : 0.0) * S;
  }
  for (int f = 0; f < F; ++f) {
    const double *nf = m->fnorm + 3 * f;
    double S = sqrt(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]);
    double invS = 1.0 / S;
    double n[3] = {invS * nf[0], invS * nf[1], invS * nf[2]};
    double ndot = invS * 0.0;
    for (int q = 0; q < 3; ++q) {
      int node = m->faces[3 * f + q];
      const double *v = V + 5 * node;
      double a = sqrt(gam * (v[4] + pstiff) / v[0]);
      double un = v[1] * n[0] + v[2] * n[1] + v[3] * n[2] - ndot;
      double coeff1 = (1.0 + beta2) * un;
      double coeff2 = pow(pow((1.0 - beta2) * un, 2.0) + pow(2.0 * locbeta * a, 2.0), 0.5);
      double ci = 0.5 * (coeff1 - coeff2);
      idti[node] += (ci < 0.0 ? ci : 0.0) * S / 3;
    }
  }
  if (p->viscous) {
    double *idtv = (double *)calloc((size_t)N, sizeof(double));
    or_viscous_idtv(m, p, V, idtv);
    for (int i = 0; i < N; ++i) dt[i] = cfl * m->cv[i] / (-1.0 * idti[i] + p->viscous_cst * idtv[i]);
    free(idtv);
  } else {
    for (int i = 0; i < N; ++i) dt[i] = cfl * m->cv[i] / (-1.0 * idti[i] + 0.0 * 0.0);
  }
  free(idti);
}

/* First-order approximate Jacobian H1 (SpaceOperator.C:2088-2150,
 * Edge.C:4472-4526, Face.C:816-847, SubDomain.C:1744-1767) plus the
 * backward-Euler term alpha_np1/dt_i (TimeState.C:731-756).
 * Adiag[N][25]; Aedge[E][2][25] = (A_ij, A_ji) per edge. */
void or_jacobian_h1(const or_mesh *m, const or_params *p, const double *V, const double *dt,
                    double *Adiag, double *Aedge) {
  int N = m->N, E = m->E, F = m->F;
  double gam = p->gamma, pstiff = p->pstiff;
  for (size_t k = 0; k < 25 * (size_t)N; ++k) Adiag[k] = 0.0;
  for (size_t k = 0; k < 50 * (size_t)E; ++k) Aedge[k] = 0.0;
  /* the Galerkin Jacobian comes before the finite-volume one (SpaceOperator.C:2131-2136) */
  if (p->viscous) or_galerkin_jacobian(m, p, V, Adiag, Aedge);
  double dfdUi[25], dfdUj[25];
  for (int l = 0; l < E; ++l) {
    int i = m->ptr[2 * l], j = m->ptr[2 * l + 1];
    const double *nl = m->normals + 3 * l;
    if (sqrt(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]) < 1e-18) continue;
    /* roejacappr3Dgas (FluxFcnDescSG.C:157-175), conservative Jacobians;
     * length = |X_j - X_i| (Edge.C:4493-4494) */
    const double *X = m->X;
    double dx[3] = {X[3 * j] - X[3 * i], X[3 * j + 1] - X[3 * i + 1], X[3 * j + 2] - X[3 * i + 2]};
    double length = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
    int lm = p->low_mach;
    or_roejac6_lm(1.0, gam, pstiff, nl, m->normal_vel[l], V + 5 * i, V + 5 * j, 1, lm ? p->lm_min_mach : 1.0,
                  lm ? p->lm_slope : 0.0, lm ? p->lm_cutoff : 1.0, lm ? p->lm_shock : 0.0, lm ? p->lm_irey : 0.0,
                  length, lm, dfdUi);
    or_roejac6_lm(1.0, gam, pstiff, nl, m->normal_vel[l], V + 5 * j, V + 5 * i, 2, lm ? p->lm_min_mach : 1.0,
                  lm ? p->lm_slope : 0.0, lm ? p->lm_cutoff : 1.0, lm ? p->lm_shock : 0.0, lm ? p->lm_irey : 0.0,
                  length, lm, dfdUj);
    double *Aii = Adiag + 25 * (size_t)i, *Ajj = Adiag + 25 * (size_t)j;
    for (int k = 0; k < 25; ++k) { Aii[k] += dfdUi[k]; Ajj[k] -= dfdUj[k]; }
    double *Aij = Aedge + 50 * (size_t)l, *Aji = Aij + 25;
    double voli = 1.0 / m->cv[i], volj = 1.0 / m->cv[j];
    for (int k = 0; k < 25; ++k) { Aij[k] += dfdUj[k] * voli; Aji[k] -= dfdUi[k] * volj; }
  }
  const double third = 1.0 / 3.0;
  for (int f = 0; f < F; ++f) {
    int code = m->codes[f];
    double n[3] = {third * m->fnorm[3 * f], third * m->fnorm[3 * f + 1], third * m->fnorm[3 * f + 2]};
    double nvel = third * 0.0;
    for (int a = 0; a < 3; ++a) {
      int node = m->faces[3 * f + a];
      double jac[25];
      if (code == 1 || code == -1 || code == 2 || code == -2 || code == 3 || code == -3)
        or_wall_jac(gam, n, nvel, V + 5 * node, jac);
      else if (code == 4 || code == -4) or_farfield_jac(gam, n, nvel, V + 5 * node, m->Ub_in, jac);
      else if (code == 5 || code == -5) or_farfield_jac(gam, n, nvel, V + 5 * node, m->Ub_out, jac);
      else continue;
      double *Aii = Adiag + 25 * (size_t)node;
      for (int k = 0; k < 25; ++k) Aii[k] += jac[k];
    }
  }
  for (int i = 0; i < N; ++i) {
    double voli = 1.0 / m->cv[i];
    double *Aii = Adiag + 25 * (size_t)i;
    for (int k = 0; k < 25; ++k) Aii[k] *= voli;
    if (dt) {
      double c_np1 = 1.0 / dt[i];
      for (int k = 0; k < 5; ++k) Aii[k + 5 * k] += c_np1;
    }
  }
}

/* SubDomain::computeMatVecProdH1 (SubDomain.C:3917-3953) with
 * DENSEMATRIXTIMESVECTOR5 (DenseMatrixOps.h:441-446). */
static void amv5_add(const double *a, const double *b, double *c) {
  c[0] += a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4];
  c[1] += a[5] * b[0] + a[6] * b[1] + a[7] * b[2] + a[8] * b[3] + a[9] * b[4];
  c[2] += a[10] * b[0] + a[11] * b[1] + a[12] * b[2] + a[13] * b[3] + a[14] * b[4];
  c[3] += a[15] * b[0] + a[16] * b[1] + a[17] * b[2] + a[18] * b[3] + a[19] * b[4];
  c[4] += a[20] * b[0] + a[21] * b[1] + a[22] * b[2] + a[23] * b[3] + a[24] * b[4];
}

void or_h1_apply(const or_mesh *m, const double *Adiag, const double *Aedge, const double *p, double *y) {
  int N = m->N, E = m->E;
  for (size_t k = 0; k < 5 * (size_t)N; ++k) y[k] = 0.0;
  for (int i = 0; i < N; ++i) amv5_add(Adiag + 25 * (size_t)i, p + 5 * (size_t)i, y + 5 * (size_t)i);
  for (int l = 0; l < E; ++l) {
    int i = m->ptr[2 * l], j = m->ptr[2 * l + 1];
    amv5_add(Aedge + 50 * (size_t)l, p + 5 * (size_t)j, y + 5 * (size_t)i);
    amv5_add(Aedge + 50 * (size_t)l + 25, p + 5 * (size_t)i, y + 5 * (size_t)j);
  }
}

/* LINPACKdgefa / LINPACKdgedi on the transpose (DenseMatrixOps.h:1006-1181)
 * for one row-major 5x5 block, in place; returns the dgefa info (0 = ok).
 * The column-major, 1-based LINPACK indexing a[i + j*n] is kept by storing
 * the transposed block at buf[6..30].  The pivot search compares |a|, which
 * is what sqrt(sqNorm(.)) comparisons select (Vector.h:56). */
#define OR_DEFINE_INVERT(NAME, T)                                                     \
  int NAME(T *blk) {                                                                  \
    enum { n = 5 };                                                                   \
    T buf[n * n + n + 1];                                                             \
    for (int r = 0; r < 5; ++r)                                                       \
      for (int c = 0; c < 5; ++c) buf[n + 1 + 5 * c + r] = blk[5 * r + c];            \
    T *a = buf;                                                                       \
    int ipvt[n + 1];                                                                  \
    T work[n + 1];                                                                    \
    for (int k = 1; k <= n - 1; ++k) {                                                \
      int kp1 = k + 1, kn = k * n, knp1 = k * n + k;                                  \
      T *aa = &a[knp1];                                                               \
      double mx = fabs((double)aa[0]);                                                \
      int l = 1;                                                                      \
      for (int ll = 1; ll < n - k + 1; ll++) {                                        \
        double t2 = fabs((double)aa[ll]);                                             \
        if (t2 > mx) { mx = t2; l = ll + 1; }                                         \
      }                                                                               \
      l += k - 1;                                                                     \
      ipvt[k] = l;                                                                    \
      if (a[l + kn] == 0.) return k;                                                  \
      if (l != k) { T t = a[l + kn]; a[l + kn] = a[knp1]; a[knp1] = t; }              \
      T t = (T)(-1. / a[knp1]);                                                       \
      aa = &a[1 + knp1];                                                              \
      for (int ll = 0; ll < n - k; ll++) aa[ll] *= t;                                 \
      T *ax = aa;                                                                     \
      for (int j = kp1; j <= n; ++j) {                                                \
        int jn1 = j * n;                                                              \
        t = a[l + jn1];                                                               \
        if (l != k) { a[l + jn1] = a[k + jn1]; a[k + jn1] = t; }                      \
        T *ay = &a[1 + k + jn1];                                                      \
        for (int ll = 0; ll < n - k; ll++) ay[ll] += t * ax[ll];                      \
      }                                                                               \
    }                                                                                 \
    ipvt[n] = n;                                                                      \
    if (a[n + n * n] == 0.) return n;                                                 \
    for (int k = 1; k <= n; ++k) {                                                    \
      int kn = k * n, knp1 = kn + k;                                                  \
      a[knp1] = (T)(1.0 / a[knp1]);                                                   \
      T t = -a[knp1];                                                                 \
      T *aa = &a[1 + kn];                                                             \
      for (int ll = 0; ll < k - 1; ll++) aa[ll] *= t;                                 \
      int kp1 = k + 1;                                                                \
      if (n < kp1) continue;                                                          \
      T *ax = aa;                                                                     \
      for (int j = kp1; j <= n; ++j) {                                                \
        int jn1 = j * n;                                                              \
        t = a[k + jn1];                                                               \
        a[k + jn1] = 0.;                                                              \
        T *ay = &a[1 + jn1];                                                          \
        for (int ll = 0; ll < k; ll++) ay[ll] += t * ax[ll];                          \
      }                                                                               \
    }                                                                                 \
    for (int kb = 1; kb <= n - 1; ++kb) {                                             \
      int k = n - kb, kn = k * n, kp1 = k + 1;                                        \
      T *aa = a + kn;                                                                 \
      for (int i = kp1; i <= n; ++i) { work[i] = aa[i]; aa[i] = 0.; }                 \
      for (int j = kp1; j <= n; ++j) {                                                \
        T t = work[j];                                                                \
        T *ax = &a[j * n + 1];                                                        \
        T *ay = &a[kn + 1];                                                           \
        for (int ll = 0; ll < n; ll++) ay[ll] += t * ax[ll];                          \
      }                                                                               \
      int l = ipvt[k];                                                                \
      if (l != k) {                                                                   \
        T *ax = &a[kn + 1], *ay = &a[l * n + 1];                                      \
        for (int ll = 0; ll < n; ll++) { T t3 = ax[ll]; ax[ll] = ay[ll]; ay[ll] = t3; } \
      }                                                                               \
    }                                                                                 \
    for (int r = 0; r < 5; ++r)                                                       \
      for (int c = 0; c < 5; ++c) blk[5 * r + c] = buf[n + 1 + 5 * c + r];            \
    return 0;                                                                         \
  }
OR_DEFINE_INVERT(or_invert5, double)
OR_DEFINE_INVERT(or_invert5f, float)

/* DiagMat::apply (DiagMatrix.C:86-100): x_i = A_i y_i (float or double
 * blocks, double arithmetic). */
void or_apply5(const double *a, const double *y, double *x) {
  x[0] = a[0] * y[0] + a[1] * y[1] + a[2] * y[2] + a[3] * y[3] + a[4] * y[4];
  x[1] = a[5] * y[0] + a[6] * y[1] + a[7] * y[2] + a[8] * y[3] + a[9] * y[4];
  x[2] = a[10] * y[0] + a[11] * y[1] + a[12] * y[2] + a[13] * y[3] + a[14] * y[4];
  x[3] = a[15] * y[0] + a[16] * y[1] + a[17] * y[2] + a[18] * y[3] + a[19] * y[4];
  x[4] = a[20] * y[0] + a[21] * y[1] + a[22] * y[2] + a[23] * y[3] + a[24] * y[4];
}
void or_apply5f(const float *a, const double *y, double *x) {
  x[0] = a[0] * y[0] + a[1] * y[1] + a[2] * y[2] + a[3] * y[3] + a[4] * y[4];
  x[1] = a[5] * y[0] + a[6] * y[1] + a[7] * y[2] + a[8] * y[3] + a[9] * y[4];
  x[2] = a[10] * y[0] + a[11] * y[1] + a[12] * y[2] + a[13] * y[3] + a[14] * y[4];
  x[3] = a[15] * y[0] + a[16] * y[1] + a[17] * y[2] + a[18] * y[3] + a[19] * y[4];
  x[4] = a[20] * y[0] + a[21] * y[1] + a[22] * y[2] + a[23] * y[3] + a[24] * y[4];
}

/* ------------------------------------------------------------------------ */
/* GMRES and the steady Newton loop                                          */
/* ------------------------------------------------------------------------ */

typedef struct {
  const or_mesh *m;
  const double *Adiag, *Aedge;  /* H1 */
  double *Dinv;                 /* [N][25] inverse diagonal blocks (FP64) */
  float *Dinvf;                 /* or FP32 blocks (mixed precision) */
  /* MatVecProdFD instead of H1 when U != NULL (MatVecProd.C:88-118,
   * 262-372): products about U with F0 = F(U), backward-Euler term with dt */
  const or_params *p;
  const double *U, *F0, *dt;
  int fd_order;
  double loceps;                /* sum(eps0 |U| + eps0), the state part of computeEpsilon */
  int ierr;                     /* reconstruction errors met in the products */
} or_ops;

static int fd_function(const or_mesh *m, const or_params *p, const double *Q, const double *U,
                       const double *dt, double *F);

/* SpaceOperator::applyBCsToResidual (SpaceOperator.C:2389-2395): the
 * strong wall conditions of the Navier-Stokes equations; nothing for Euler
 * (no BcFcn, SpaceOperator.C:246-265). */
static void residual_bcs(const or_mesh *m, const or_params *p, const double *U, double *F) {
  if (!p->viscous) return;
  int *type = (int *)malloc(sizeof(int) * (size_t)(m->N ? m->N : 1));
  or_node_types(m, type);
  or_apply_bcs_residual(m, p, type, U, F);
  free(type);
}

/* MatVecProdFD::apply with computeEpsilon (MatVecProd.C:262-372, 879-929). */
static void ops_fd_product(or_ops *o, const double *pv, double *y) {
  size_t n = 5 * (size_t)o->m->N;
  double pp = 0.0;
  for (size_t k = 0; k < n; ++k) pp += pv[k] * pv[k];
  double norm = sqrt(pp);
  double eps = norm < 1.0e-14 ? 1.e-6 : o->loceps / ((double)n * norm);
  double *Q = (double *)calloc(n ? n : 1, sizeof(double)), *Fa = (double *)calloc(n ? n : 1, sizeof(double));
  if (!Q || !Fa) {
    free(Q);
    free(Fa);
    o->ierr = -1;
    return;
  }
  for (size_t k = 0; k < n; ++k) Q[k] = o->U[k] + eps * pv[k];
  o->ierr += fd_function(o->m, o->p, Q, o->U, o->dt, Fa);
  if (o->fd_order == 1) {
    double inv = 1.0 / eps;
    for (size_t k = 0; k < n; ++k) y[k] = inv * (Fa[k] - o->F0[k]);
  } else {
    double *Fb = (double *)malloc(sizeof(double) * n);
    for (size_t k = 0; k < n; ++k) Q[k] = o->U[k] - eps * pv[k];
    o->ierr += fd_function(o->m, o->p, Q, o->U, o->dt, Fb);
    double half = 0.5 / eps;
    for (size_t k = 0; k < n; ++k) y[k] = half * (Fa[k] - Fb[k]);
    free(Fb);
  }
  free(Q);
  free(Fa);
}

static void ops_matvec(or_ops *o, const double *p, double *y) {
  if (o->U) ops_fd_product(o, p, y);
  else or_h1_apply(o->m, o->Adiag, o->Aedge, p, y);
}

static void ops_prec(const or_ops *o, const double *y, double *x) {
  int N = o->m->N;
  for (int i = 0; i < N; ++i) {
    if (o->Dinvf) or_apply5f(o->Dinvf + 25 * (size_t)i, y + 5 * (size_t)i, x + 5 * (size_t)i);
    else or_apply5(o->Dinv + 25 * (size_t)i, y + 5 * (size_t)i, x + 5 * (size_t)i);
  }
}

static double vdot(size_t n, const double *a, const double *b) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* GmresSolver::solve (KspSolver.C:353-492), typePrec = 2 (right), modified
 * Gram-Schmidt (:404-411), rotations and back-substitution (:919-982).
 * x is set to 0 first (ImplicitCoupledTsDesc.C:355).  Returns iterations. */
int or_gmres(or_ops *o, const double *b, double *x, int m, double eps, int maxits, double *final_res) {
  size_t n = 5 * (size_t)o->m->N;
  double *V = (double *)malloc(sizeof(double) * n * (size_t)(m + 1));
  double *w = (double *)malloc(sizeof(double) * n), *r = (double *)malloc(sizeof(double) * n);
  double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = (double *)calloc(2 * (size_t)m, sizeof(double));
  double *g = (double *)calloc((size_t)m + 1, sizeof(double)), *y = (double *)calloc((size_t)m, sizeof(double));
#define HH(i, j) H[(size_t)(i) * m + (j)]
  for (size_t i = 0; i < n; ++i) x[i] = 0.0;
  int iter = 0, exitLoop = 0;
  double target = 0.0, l2res = 0.0;
  do {
    ops_matvec(o, x, w);
    for (size_t i = 0; i < n; ++i) r[i] = b[i] - w[i];
    double beta = sqrt(vdot(n, r, r));
    if (iter == 0) target = eps * sqrt(vdot(n, b, b));
    if (beta == 0.0) break;
    for (size_t i = 0; i < n; ++i) V[i] = (1.0 / beta) * r[i];
    for (int k = 0; k <= m; ++k) g[k] = 0.0;
    g[0] = beta;
    int j;
    for (j = 0; j < m; ++j) {
      ops_prec(o, V + (size_t)j * n, r);
      ops_matvec(o, r, w);
      for (int i = 0; i <= j; ++i) {
        HH(i, j) = vdot(n, w, V + (size_t)i * n);
        const double *vi = V + (size_t)i * n;
        for (size_t q = 0; q < n; ++q) w[q] -= HH(i, j) * vi[q];
      }
      HH(j + 1, j) = sqrt(vdot(n, w, w));
      double hn = HH(j + 1, j);
      for (int i = 0; i < j; ++i) {
        double hj = cs[2 * i + 1] * HH(i, j) + cs[2 * i] * HH(i + 1, j);
        double hjp1 = -cs[2 * i] * HH(i, j) + cs[2 * i + 1] * HH(i + 1, j);
        HH(i, j) = hj;
        HH(i + 1, j) = hjp1;
      }
      if (HH(j, j) * HH(j, j) == 0.0) {
        cs[2 * j + 1] = 0.0;
        cs[2 * j] = 1.0;
      } else {
        double coef = 1.0 / sqrt(HH(j, j) * HH(j, j) + HH(j + 1, j) * HH(j + 1, j));
        cs[2 * j + 1] = sqrt(HH(j, j) * HH(j, j)) * coef;
        cs[2 * j] = cs[2 * j + 1] * HH(j + 1, j) / HH(j, j);
      }
      HH(j, j) = cs[2 * j + 1] * HH(j, j) + cs[2 * j] * HH(j + 1, j);
      HH(j + 1, j) = 0.0;
      g[j + 1] = -cs[2 * j] * g[j];
      g[j] = cs[2 * j + 1] * g[j];
      if (hn == 0.0) { ++iter; exitLoop = 1; break; }
      double *vn = V + (size_t)(j + 1) * n;
      for (size_t q = 0; q < n; ++q) vn[q] = (1. / hn) * w[q];
      l2res = fabs(g[j + 1]);
      ++iter;
      if (l2res <= target || iter >= maxits) { exitLoop = 1; break; }
    }
    if (j == m) --j;
    for (int i = j; i >= 0; --i) {
      y[i] = g[i];
      for (int k = j; k >= i + 1; --k) y[i] -= HH(i, k) * y[k];
      y[i] /= HH(i, i);
    }
    for (size_t q = 0; q < n; ++q) w[q] = 0.0;
    for (int k = 0; k <= j; ++k) {
      const double *vk = V + (size_t)k * n;
      for (size_t q = 0; q < n; ++q) w[q] += y[k] * vk[q];
    }
    ops_prec(o, w, r);
    for (size_t q = 0; q < n; ++q) x[q] += r[q];
  } while (!exitLoop);
#undef HH
  if (final_res) *final_res = l2res;
  free(V); free(w); free(r); free(H); free(cs); free(g); free(y);
  return iter;
}

/* Builds the operators of one Newton step: H1 at V with the dt term and the
 * inverted diagonal blocks (FP64, or FP32 copies for precision 1, as
 * JacobiPrec<float,5>::getData + setup); with mvp 1 the Krylov products are
 * MatVecProdFD about U (conservative; F0 = F(U) = R(U)) of order fd_order,
 * the preconditioner unchanged.  *ierr (optional) receives the reconstruction
 * errors of the products. */
int or_gmres_mvp(const or_mesh *m, const or_params *p, const double *V, const double *U, const double *dt,
                 int precision, int mvp, int fd_order, const double *b, double *x, int gm, double eps, int maxits,
                 double *final_res, int *ierr) {
  int N = m->N, E = m->E;
  size_t n = 5 * (size_t)N;
  double *Ad = (double *)malloc(sizeof(double) * 25 * (size_t)N);
  double *Ae = (double *)malloc(sizeof(double) * 50 * (size_t)E);
  double *F0 = NULL;
  or_jacobian_h1(m, p, V, dt, Ad, Ae);
  /* MatVecProdH1::evaluate ends with applyBCsToJacobian (MatVecProd.C:1080-1092) */
  if (p->viscous && U) {
    int *type = (int *)malloc(sizeof(int) * (size_t)N);
    or_node_types(m, type);
    or_apply_bcs_jacobian(m, p, type, U, Ad, Ae);
    free(type);
  }
  or_ops o;
  memset(&o, 0, sizeof(o));
  o.m = m;
  o.Adiag = Ad;
  o.Aedge = Ae;
  if (mvp == 1) {
    F0 = (double *)malloc(sizeof(double) * n);
    o.ierr += or_residual(m, p, U, F0, 0, NULL, NULL, NULL, NULL);
    residual_bcs(m, p, U, F0);
    o.p = p;
    o.U = U;
    o.F0 = F0;
    o.dt = dt;
    o.fd_order = fd_order;
    const double eps0 = 1.e-6;
    double loceps = 0.0;
    for (size_t k = 0; k < n; ++k) loceps += eps0 * fabs(U[k]) + eps0;
    o.loceps = loceps;
  }
  if (precision) {
    o.Dinvf = (float *)malloc(sizeof(float) * 25 * (size_t)N);
    for (size_t k = 0; k < 25 * (size_t)N; ++k) o.Dinvf[k] = (float)Ad[k];
    for (int i = 0; i < N; ++i) or_invert5f(o.Dinvf + 25 * (size_t)i);
  } else {
    o.Dinv = (double *)malloc(sizeof(double) * 25 * (size_t)N);
    for (size_t k = 0; k < 25 * (size_t)N; ++k) o.Dinv[k] = Ad[k];
    for (int i = 0; i < N; ++i) or_invert5(o.Dinv + 25 * (size_t)i);
  }
  int its = or_gmres(&o, b, x, gm, eps, maxits, final_res);
  if (ierr) *ierr = o.ierr;
  free(Ad); free(Ae); free(o.Dinv); free(o.Dinvf); free(F0);
  return its;
}

int or_gmres_h1(const or_mesh *m, const or_params *p, const double *V, const double *dt, int precision,
                const double *b, double *x, int gm, double eps, int maxits, double *final_res) {
  return or_gmres_mvp(m, p, V, NULL, dt, precision, 0, 1, b, x, gm, eps, maxits, final_res, NULL);
}

/* TsParameters::computeCflNumber, CflLaw Old (TsParameters.C:198-200):
 * cfl = min(max(max(c1, min(its*c2, cfl + c2)), cfl0/res^ser), cflMax). */
double or_cfl_old(int its, double res, double cfl, double cfl0, double cflMax, double ser, double c1, double c2) {
  double a = its * c2 < cfl + c2 ? its * c2 : cfl + c2;
  double b = c1 > a ? c1 : a;
  double c = cfl0 / pow(res, ser);
  double d = b > c ? b : c;
  return d < cflMax ? d : cflMax;
}

/* Steady pseudo-transient Newton (one Newton iteration per step) with the
 * CflLaw Old (TsParameters.C:198-200); returns steps taken; history[k] =
 * ||R||_2 at step k; *converged set when ||R|| <= eps ||R0||.  mvp 0: Krylov
 * products with H1; 1: MatVecProdFD of order fd_order about the step's state
 * (ImplicitCoupledTsDesc.C:33-44).  Stops early (not converged) when a
 * residual or a product meets a non-physical reconstruction. */
/* KspConvCriterion::computeEpsilonEisenstadt (KspConvCriterion.C:44-74):
 * gamma 0.1, alpha 2, delta 0.1; *its counts the calls (nlits). */
double or_eps_eisenstat(double eps0, double fnorm, int *its, double *fprev, double *eprev) {
  double eps;
  if (*its == 0) {
    eps = eps0;
  } else {
    eps = 0.1 * pow(fnorm / *fprev, 2.0);
    double eps1 = 0.1 * pow(*eprev, 2.0);
    if (eps1 < 0.1) eps = fmin(eps0, eps);
    else eps = fmin(eps0, fmax(eps, eps1));
  }
  *fprev = fnorm;
  *eprev = eps;
  ++*its;
  return eps;
}

int or_steady_solve_mvp(const or_mesh *m, const or_params *p, double *U, int precision, int mvp, int fd_order,
                        int max_steps, double eps, double cfl0, double cfl_max, double ser, int gm, int gmaxits,
                        double geps, double *history, int *converged, int *krylov_its) {
  return or_steady_solve_ew(m, p, U, precision, mvp, fd_order, max_steps, eps, cfl0, cfl_max, ser, gm, gmaxits, geps,
                            0, history, converged, krylov_its);
}

/* As or_steady_solve_mvp, with the Krylov tolerance of KspData::EISENSTADT
 * when eisenstat (the criterion's history runs over the pseudo-time steps:
 * one Newton iteration each, so nlits never resets, KspConvCriterion.C:31). */
int or_steady_solve_ew(const or_mesh *m, const or_params *p, double *U, int precision, int mvp, int fd_order,
                       int max_steps, double eps, double cfl0, double cfl_max, double ser, int gm, int gmaxits,
                       double geps, int eisenstat, double *history, int *converged, int *krylov_its) {
  int N = m->N;
  int ew_its = 0;
  double ew_f = 0.0, ew_e = 0.0;
  size_t n = 5 * (size_t)N;
  double *R = (double *)malloc(sizeof(double) * n), *V = (double *)malloc(sizeof(double) * n);
  double *dt = (double *)malloc(sizeof(double) * N), *b = (double *)malloc(sizeof(double) * n);
  double *dQ = (double *)malloc(sizeof(double) * n);
  double res0 = 0.0, cfl = cfl0;
  int step;
  *converged = 0;
  *krylov_its = 0;
  for (step = 0; step <= max_steps; ++step) {
    int ierr = or_residual(m, p, U, R, 0, V, 0, 0, 0);
    if (ierr) break;
    residual_bcs(m, p, U, R);
    double res = sqrt(vdot(n, R, R));
    if (step == 0) res0 = res;
    if (history) history[step] = res;
    if (res == 0.0 || res <= eps * res0) { *converged = 1; break; }
    if (step == max_steps) break;
    double rr = res / res0;
    cfl = or_cfl_old(step, rr, cfl, cfl0, cfl_max, ser, 0.0, 0.0);
    or_timestep(m, p, V, cfl, dt);
    for (size_t k = 0; k < n; ++k) b[k] = -1.0 * R[k];
    double fr;
    int perr = 0;
    double ge = eisenstat ? or_eps_eisenstat(geps, res, &ew_its, &ew_f, &ew_e) : geps;
    *krylov_its += or_gmres_mvp(m, p, V, U, dt, precision, mvp, fd_order, b, dQ, gm, ge, gmaxits, &fr, &perr);
    if (perr) break;
    for (size_t k = 0; k < n; ++k) U[k] += dQ[k];
  }
  free(R); free(V); free(dt); free(b); free(dQ);
  return step;
}

int or_steady_solve(const or_mesh *m, const or_params *p, double *U, int precision, int max_steps,
                    double eps, double cfl0, double cfl_max, double ser, int gm, int gmaxits,
                    double geps, double *history, int *converged, int *krylov_its) {
  return or_steady_solve_mvp(m, p, U, precision, 0, 1, max_steps, eps, cfl0, cfl_max, ser, gm, gmaxits, geps,
                             history, converged, krylov_its);
}

/* ------------------------------------------------------------------------ */
/* MatVecProdFD                                                              */
/* ------------------------------------------------------------------------ */

/* MatVecProdFD::computeEpsilon (MatVecProd.C:879-929), one subdomain with all
 * nodes owned: eps = sum(eps0 |U| + eps0) / (size * ||p||), eps0 = 1e-6. */
static double fd_epsilon(size_t n, const double *U, const double *p) {
  const double eps0 = 1.e-6;
  double pp = 0.0;
  for (size_t k = 0; k < n; ++k) pp += p[k] * p[k];
  double norm = sqrt(pp);
  if (norm < 1.0e-14) return eps0;
  double loceps = 0.0;
  for (size_t k = 0; k < n; ++k) loceps += eps0 * fabs(U[k]) + eps0;
  return loceps / ((double)n * norm);
}

/* F(Q) = R(Q) + (1/dt)(1 Q + (-1) Un) with Un = U (TimeState::add_dAW_dt,
 * TimeState.C:40-73, backward Euler, NONDESCRIPTOR). */
static int fd_function(const or_mesh *m, const or_params *p, const double *Q, const double *U,
                       const double *dt, double *F) {
  int ierr = or_residual(m, p, Q, F, 0, NULL, NULL, NULL, NULL);
  if (dt)
    for (int i = 0; i < m->N; ++i) {
      double invDt = 1.0 / dt[i];
      for (int k = 0; k < 5; ++k) F[5 * i + k] += invDt * (1.0 * Q[5 * i + k] + -1.0 * U[5 * i + k]);
    }
  residual_bcs(m, p, Q, F);
  return ierr;
}

/* MatVecProdFD::evaluate + apply (MatVecProd.C:88-118, 262-372). */
int or_jvp_fd(const or_mesh *m, const or_params *p, const double *U, const double *pv, int order,
              const double *dt, double *prod) {
  size_t n = 5 * (size_t)m->N;
  double *Q = (double *)calloc(n ? n : 1, sizeof(double));
  double *Fa = (double *)malloc(sizeof(double) * n);
  double *Fb = (double *)malloc(sizeof(double) * n);
  if (!Q || !Fa || !Fb) {
    free(Q);
    free(Fa);
    free(Fb);
    return -1;
  }
  double eps = fd_epsilon(n, U, pv);
  int ierr = 0;
  for (size_t k = 0; k < n; ++k) Q[k] = U[k] + eps * pv[k];
  ierr += fd_function(m, p, Q, U, dt, Fa);
  if (order == 1) {
    ierr += fd_function(m, p, U, U, dt, Fb);
    double inv = 1.0 / eps;
    for (size_t k = 0; k < n; ++k) prod[k] = inv * (Fa[k] - Fb[k]);
  } else {
    for (size_t k = 0; k < n; ++k) Q[k] = U[k] - eps * pv[k];
    ierr += fd_function(m, p, Q, U, dt, Fb);
    double half = 0.5 / eps;
    for (size_t k = 0; k < n; ++k) prod[k] = half * (Fa[k] - Fb[k]);
  }
  free(Q); free(Fa); free(Fb);
  return ierr;
}

/* ------------------------------------------------------------------------ */
/* Partitioned residual for the CPU baseline                                 */
/* ------------------------------------------------------------------------ */

/* The residual as the reference runs it on P subdomains with OpenMP
 * (Domain.C:868-928): the nodes are split into P contiguous ranges; each
 * thread runs the edge loops (gradients, then fluxes) over the edges whose
 * lower node it owns, adds contributions to its own nodes directly and
 * keeps those to later partitions' nodes in a private buffer; after each
 * loop the owners add the buffers of the earlier partitions in partition
 * order (the shared-node assembly).  Faces and the 1/V scaling run per
 * owned node.  Sums are formed in another order than or_residual's, so the
 * result agrees with it to rounding; when ierr > 0 the result is not
 * meaningful.  Returns the non-physical reconstruction count. */
int or_residual_par(const or_mesh *m, const or_params *p, const double *U, double *R, int nthreads) {
  const int N = m->N, E = m->E, F = m->F;
  const double gam = p->gamma, pstiff = p->pstiff;
  const int P = nthreads < 1 ? 1 : nthreads;
  double *V = (double *)malloc(sizeof(double) * 5 * (size_t)N);
  double *G = (double *)malloc(sizeof(double) * 15 * (size_t)N); /* [N][3][5] */
  int *nlo = (int *)malloc(sizeof(int) * (P + 1));
  int *elo = (int *)malloc(sizeof(int) * (P + 1));
  int *hend = (int *)malloc(sizeof(int) * P);     /* one past the largest upper node of each partition */
  double **hb = (double **)calloc(P, sizeof(double *));
  int ierr_total = 0, fail = 0;
  if (!V || !G || !nlo || !elo || !hend || !hb) fail = 1;
  if (!fail) {
    for (int q = 0; q <= P; ++q) nlo[q] = (int)((long long)N * q / P);
    /* edges sorted by lower node: partition q owns [elo[q], elo[q+1]) */
    for (int q = 0, l = 0; q <= P; ++q) {
      while (l < E && m->ptr[2 * l] < nlo[q]) ++l;
      elo[q] = l;
    }
    for (int q = 0; q < P; ++q) {
      int mx = nlo[q + 1];
      for (int l = elo[q]; l < elo[q + 1]; ++l)
        if (m->ptr[2 * l + 1] + 1 > mx) mx = m->ptr[2 * l + 1] + 1;
      hend[q] = mx;
      size_t len = (size_t)(mx - nlo[q + 1]);
      hb[q] = (double *)malloc(sizeof(double) * 15 * (len ? len : 1));
      if (!hb[q]) fail = 1;
    }
  }
  if (fail) {
    free(V); free(G); free(nlo); free(elo); free(hend);
    if (hb) for (int q = 0; q < P; ++q) free(hb[q]);
    free(hb);
    return -1;
  }
#pragma omp parallel num_threads(P) reduction(+ : ierr_total)
  {
#ifdef _OPENMP
    const int q = omp_get_thread_num();
#else
    const int q = 0;
#endif
    const int a = nlo[q], b = nlo[q + 1];
    const double *X = m->X;
    double *H = hb[q];
    const size_t hlen = (size_t)(hend[q] - b);
    or_cons_to_prim(b - a, U + 5 * (size_t)a, V + 5 * (size_t)a, gam, pstiff);
    for (size_t k = 15 * (size_t)a; k < 15 * (size_t)b; ++k) G[k] = 0.0;
    for (size_t k = 0; k < 15 * hlen; ++k) H[k] = 0.0;
#pragma omp barrier
    if (p->second_order) {
      for (int l = elo[q]; l < elo[q + 1]; ++l) {
        int i = m->ptr[2 * l], j = m->ptr[2 * l + 1];
        double Wi[3], Wj[3];
        double dx[3] = {X[3 * j] - X[3 * i], X[3 * j + 1] - X[3 * i + 1], X[3 * j + 2] - X[3 * i + 2]};
        lsq_local_weights(dx, m->R + 6 * (size_t)i, Wi);
        dx[0] = -dx[0]; dx[1] = -dx[1]; dx[2] = -dx[2];
        lsq_local_weights(dx, m->R + 6 * (size_t)j, Wj);
        double *gi = G + 15 * (size_t)i;
        double *gj = j < b ? G + 15 * (size_t)j : H + 15 * (size_t)(j - b);
        for (int k = 0; k < 5; ++k) {
          double d = V[5 * (size_t)j + k] - V[5 * (size_t)i + k];
          gi[k] += Wi[0] * d; gi[5 + k] += Wi[1] * d; gi[10 + k] += Wi[2] * d;
          gj[k] -= Wj[0] * d; gj[5 + k] -= Wj[1] * d; gj[10 + k] -= Wj[2] * d;
        }
      }
#pragma omp barrier
      for (int r = 0; r < q; ++r)
        for (int n = (a > nlo[r + 1] ? a : nlo[r + 1]); n < b && n < hend[r]; ++n)
          for (int k = 0; k < 15; ++k) G[15 * (size_t)n + k] += hb[r][15 * (size_t)(n - nlo[r + 1]) + k];
#pragma omp barrier
    }
    for (size_t k = 5 * (size_t)a; k < 5 * (size_t)b; ++k) R[k] = 0.0;
    for (size_t k = 0; k < 15 * hlen; ++k) H[k] = 0.0;
    int ierr = 0;
    for (int l = elo[q]; l < elo[q + 1]; ++l) {
      int i = m->ptr[2 * l], j = m->ptr[2 * l + 1];
      double Vi[5], Vj[5], flux[5];
      const double *vi = V + 5 * (size_t)i, *vj = V + 5 * (size_t)j;
      if (p->second_order) {
        double dx[3] = {X[3 * j] - X[3 * i], X[3 * j + 1] - X[3 * i + 1], X[3 * j + 2] - X[3 * i + 2]};
        double ddVij[5], ddVji[5];
        const double *gi = G + 15 * (size_t)i, *gj = G + 15 * (size_t)j;
        for (int k = 0; k < 5; ++k) {
          ddVij[k] = dx[0] * gi[k] + dx[1] * gi[5 + k] + dx[2] * gi[10 + k];
          ddVji[k] = dx[0] * gj[k] + dx[1] * gj[5 + k] + dx[2] * gj[10 + k];
        }
        or_vanalbada5(p->beta, vi, ddVij, vj, ddVji, Vi, Vj);
      } else {
        for (int k = 0; k < 5; ++k) { Vi[k] = vi[k]; Vj[k] = vj[k]; }
      }
      if (Vi[0] <= 0.0 || Vi[4] + pstiff <= 0.0) ++ierr;
      if (Vj[0] <= 0.0 || Vj[4] + pstiff <= 0.0) ++ierr;
      const double *nl = m->normals + 3 * (size_t)l;
      if (sqrt(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]) < 1e-18) continue;
      if (p->low_mach) {
        double ex[3] = {X[3 * j] - X[3 * i], X[3 * j + 1] - X[3 * i + 1], X[3 * j + 2] - X[3 * i + 2]};
        double length = sqrt(ex[0] * ex[0] + ex[1] * ex[1] + ex[2] * ex[2]);
        or_roeflux5(1.0, gam, pstiff, nl, m->normal_vel[l], Vi, Vj, p->lm_min_mach, p->lm_slope, p->lm_cutoff,
                    p->lm_shock, p->lm_irey, length, 1, flux);
      } else {
        or_roeflux7(gam, pstiff, nl, m->normal_vel[l], Vi, Vj, flux);
      }
      double *rj = j < b ? R + 5 * (size_t)j : H + 5 * (size_t)(j - b);
      for (int k = 0; k < 5; ++k) {
        R[5 * (size_t)i + k] += flux[k];
        rj[k] -= flux[k];
      }
    }
    ierr_total += ierr;
#pragma omp barrier
    for (int r = 0; r < q; ++r)
      for (int n = (a > nlo[r + 1] ? a : nlo[r + 1]); n < b && n < hend[r]; ++n)
        for (int k = 0; k < 5; ++k) R[5 * (size_t)n + k] += hb[r][5 * (size_t)(n - nlo[r + 1]) + k];
    const double third = 1.0 / 3.0;
    for (int f = 0; f < F; ++f) {
      int code = m->codes[f];
      double nv[3] = {third * m->fnorm[3 * f], third * m->fnorm[3 * f + 1], third * m->fnorm[3 * f + 2]};
      for (int c = 0; c < 3; ++c) {
        int node = m->faces[3 * f + c];
        if (node < a || node >= b) continue;
        double flux[5];
        if (code == 1 || code == -1 || code == 2 || code == -2 || code == 3 || code == -3) or_wallflux(nv, 0.0, V + 5 * (size_t)node, flux);
        else if (code == 4 || code == -4) or_boundflux5(gam, nv, 0.0, V + 5 * (size_t)node, m->Ub_in, flux);
        else if (code == 5 || code == -5) or_boundflux5(gam, nv, 0.0, V + 5 * (size_t)node, m->Ub_out, flux);
        else continue;
        for (int k = 0; k < 5; ++k) R[5 * (size_t)node + k] += flux[k];
      }
    }
    for (int n = a; n < b; ++n) {
      double invcv = 1.0 / m->cv[n];
      for (int k = 0; k < 5; ++k) R[5 * (size_t)n + k] *= invcv;
    }
  }
  free(V); free(G); free(nlo); free(elo); free(hend);
  for (int q = 0; q < P; ++q) free(hb[q]);
  free(hb);
  return ierr_total;
}

/* ---- RAS / ILU(0) ---------------------------------------------------------
 * IluPrec<Scalar,5,double> with fill 0 and type RAS (KspPrec.C:127-320):
 * SparseMat::numericILU (SparseMatrix.C:353-396) factors each subdomain's
 * matrix in place (the L part scaled by the inverted pivot block, the pivot
 * blocks inverted by invertDenseMatrix), SparseMat::lusol (SparseMatrix.C:
 * 406-466) solves; RAS keeps each subdomain's solution on the subdomain's own
 * nodes (killSlave).  A subdomain here is the node range [start[s],
 * start[s+1]); its matrix holds the H1 blocks that couple two of its nodes,
 * rows in node order and, within a row, columns in ascending order (the
 * order of the reference's RCM-renumbered subdomain matrix when the mesh is
 * numbered by RCM).  Blocks products are DENSEMATRIXTIMESDENSEMATRIX5 and
 * DENSEMATRIXTIMESVECTOR5 (DenseMatrixOps.h:302-331, 441-446): each entry a
 * left-to-right sum of five products, then assigned or subtracted. */
#define OR_DEFINE_ILU(NAME, T, INV)                                                                         \
  static int NAME(int N, int E, const int *ptr, const double *D, const double *A, int nsub,                 \
                  const int *start, const double *y, double *x) {                                           \
    int *cnt = (int *)calloc((size_t)N + 1, sizeof(int));                                                   \
    for (int l = 0; l < E; ++l) {                                                                           \
      int i = ptr[2 * l], j = ptr[2 * l + 1];                                                               \
      cnt[i + 1]++;                                                                                         \
      cnt[j + 1]++;                                                                                         \
    }                                                                                                       \
    for (int i = 0; i < N; ++i) cnt[i + 1] += cnt[i] + 0;                                                   \
    /* full adjacency with the block of each entry (row i, column j) */                                     \
    int *adj = (int *)malloc(sizeof(int) * 2 * (size_t)(E > 0 ? E : 1));                                   \
    const double **blk = (const double **)malloc(sizeof(double *) * 2 * (size_t)(E > 0 ? E : 1));          \
    int *fill = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));                                       \
    for (int i = 0; i < N; ++i) fill[i] = cnt[i];                                                           \
    for (int l = 0; l < E; ++l) {                                                                           \
      int i = ptr[2 * l], j = ptr[2 * l + 1];                                                               \
      adj[fill[i]] = j; blk[fill[i]++] = A + 50 * (size_t)l;                                                \
      adj[fill[j]] = i; blk[fill[j]++] = A + 50 * (size_t)l + 25;                                           \
    }                                                                                                       \
    int rc = 0;                                                                                             \
    for (int s = 0; s < nsub && !rc; ++s) {                                                                 \
      const int a0 = start[s], n = start[s + 1] - start[s];                                                 \
      /* local CSR: diagonal + couplings inside the subdomain, columns ascending */                         \
      int *ia = (int *)malloc(sizeof(int) * ((size_t)n + 1)), *ju = (int *)malloc(sizeof(int) * (size_t)n); \
      int nnz = 0;                                                                                          \
      for (int r = 0; r < n; ++r) {                                                                         \
        int i = a0 + r;                                                                                     \
        nnz += 1;                                                                                           \
        for (int q = cnt[i]; q < cnt[i + 1]; ++q)                                                           \
          if (adj[q] >= a0 && adj[q] < a0 + n) nnz++;                                                       \
      }                                                                                                     \
      int *ja = (int *)malloc(sizeof(int) * (size_t)nnz);                                                   \
      T *a = (T *)malloc(sizeof(T) * 25 * (size_t)nnz);                                                     \
      int k = 0;                                                                                            \
      for (int r = 0; r < n; ++r) {                                                                         \
        int i = a0 + r;                                                                                     \
        ia[r] = k;                                                                                          \
        /* columns ascending: the neighbours of a row are not sorted here, so insert in order */           \
        int cols[512];                                                                                      \
        const double *src[512];                                                                             \
        int nc = 0;                                                                                         \
        cols[nc] = r; src[nc++] = D + 25 * (size_t)i;                                                       \
        for (int q = cnt[i]; q < cnt[i + 1]; ++q)                                                           \
          if (adj[q] >= a0 && adj[q] < a0 + n && nc < 512) { cols[nc] = adj[q] - a0; src[nc++] = blk[q]; }  \
        for (int u = 1; u < nc; ++u)                                                                        \
          for (int v = u; v > 0 && cols[v - 1] > cols[v]; --v) {                                            \
            int tc = cols[v]; cols[v] = cols[v - 1]; cols[v - 1] = tc;                                      \
            const double *ts = src[v]; src[v] = src[v - 1]; src[v - 1] = ts;                                \
          }                                                                                                 \
        for (int u = 0; u < nc; ++u, ++k) {                                                                 \
          ja[k] = cols[u];                                                                                  \
          if (cols[u] == r) ju[r] = k;                                                                      \
          for (int e = 0; e < 25; ++e) a[25 * (size_t)k + e] = (T)src[u][e];                                \
        }                                                                                                   \
      }                                                                                                     \
      ia[n] = k;                                                                                            \
      /* numericILU */                                                                                      \
      int *marker = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));                                   \
      for (int r = 0; r < n; ++r) marker[r] = -1;                                                           \
      for (int i = 0; i < n && !rc; ++i) {                                                                  \
        for (int q = ia[i]; q < ia[i + 1]; ++q) marker[ja[q]] = q;                                          \
        for (int q = ia[i]; q < ju[i]; ++q) {                                                               \
          int l = ja[q];                                                                                    \
          T mult[25], res[25];                                                                              \
          const T *pa = a + 25 * (size_t)q, *pb = a + 25 * (size_t)ju[l];                                   \
          for (int r5 = 0; r5 < 5; ++r5)                                                                    \
            for (int c5 = 0; c5 < 5; ++c5)                                                                  \
              mult[5 * r5 + c5] = pa[5 * r5] * pb[c5] + pa[5 * r5 + 1] * pb[5 + c5] +                       \
                                  pa[5 * r5 + 2] * pb[10 + c5] + pa[5 * r5 + 3] * pb[15 + c5] +             \
                                  pa[5 * r5 + 4] * pb[20 + c5];                                             \
          for (int e = 0; e < 25; ++e) a[25 * (size_t)q + e] = mult[e];                                     \
          for (int jj = ju[l] + 1; jj < ia[l + 1]; ++jj) {                                                  \
            int mm = marker[ja[jj]];                                                                        \
            if (mm == -1) continue;                                                                         \
            const T *pj = a + 25 * (size_t)jj;                                                              \
            for (int r5 = 0; r5 < 5; ++r5)                                                                  \
              for (int c5 = 0; c5 < 5; ++c5)                                                                \
                res[5 * r5 + c5] = mult[5 * r5] * pj[c5] + mult[5 * r5 + 1] * pj[5 + c5] +                  \
                                   mult[5 * r5 + 2] * pj[10 + c5] + mult[5 * r5 + 3] * pj[15 + c5] +        \
                                   mult[5 * r5 + 4] * pj[20 + c5];                                          \
            for (int e = 0; e < 25; ++e) a[25 * (size_t)mm + e] -= res[e];                                  \
          }                                                                                                 \
        }                                                                                                   \
        for (int q = ia[i]; q < ia[i + 1]; ++q) marker[ja[q]] = -1;                                         \
        if (INV(a + 25 * (size_t)ju[i])) rc = a0 + i + 1;                                                   \
      }                                                                                                     \
      /* lusol: forward (unit L), backward with the inverted pivots */                                      \
      if (!rc) {                                                                                            \
        double *xs = x + 5 * (size_t)a0;                                                                    \
        for (int i = 0; i < n; ++i) {                                                                       \
          for (int c = 0; c < 5; ++c) xs[5 * i + c] = y[5 * ((size_t)a0 + i) + c];                          \
          for (int q = ia[i]; q < ju[i]; ++q) {                                                             \
            const T *pa = a + 25 * (size_t)q;                                                               \
            const double *xb = xs + 5 * (size_t)ja[q];                                                      \
            for (int r5 = 0; r5 < 5; ++r5)                                                                  \
              xs[5 * i + r5] -= (double)pa[5 * r5] * xb[0] + (double)pa[5 * r5 + 1] * xb[1] +               \
                                (double)pa[5 * r5 + 2] * xb[2] + (double)pa[5 * r5 + 3] * xb[3] +           \
                                (double)pa[5 * r5 + 4] * xb[4];                                             \
          }                                                                                                 \
        }                                                                                                   \
        for (int i = n - 1; i >= 0; --i) {                                                                  \
          for (int q = ju[i] + 1; q < ia[i + 1]; ++q) {                                                     \
            const T *pa = a + 25 * (size_t)q;                                                               \
            const double *xb = xs + 5 * (size_t)ja[q];                                                      \
            for (int r5 = 0; r5 < 5; ++r5)                                                                  \
              xs[5 * i + r5] -= (double)pa[5 * r5] * xb[0] + (double)pa[5 * r5 + 1] * xb[1] +               \
                                (double)pa[5 * r5 + 2] * xb[2] + (double)pa[5 * r5 + 3] * xb[3] +           \
                                (double)pa[5 * r5 + 4] * xb[4];                                             \
          }                                                                                                 \
          const T *pd = a + 25 * (size_t)ju[i];                                                             \
          double res[5];                                                                                    \
          for (int r5 = 0; r5 < 5; ++r5)                                                                    \
            res[r5] = (double)pd[5 * r5] * xs[5 * i] + (double)pd[5 * r5 + 1] * xs[5 * i + 1] +             \
                      (double)pd[5 * r5 + 2] * xs[5 * i + 2] + (double)pd[5 * r5 + 3] * xs[5 * i + 3] +     \
                      (double)pd[5 * r5 + 4] * xs[5 * i + 4];                                               \
          for (int r5 = 0; r5 < 5; ++r5) xs[5 * i + r5] = res[r5];                                          \
        }                                                                                                   \
      }                                                                                                     \
      free(marker); free(ia); free(ju); free(ja); free(a);                                                  \
    }                                                                                                       \
    free(cnt); free(adj); free(blk); free(fill);                                                            \
    return rc;                                                                                              \
  }
OR_DEFINE_ILU(or_ras_ilu0_f64, double, or_invert5)
OR_DEFINE_ILU(or_ras_ilu0_f32, float, or_invert5f)

int or_ras_ilu0_apply(int N, int E, const int *ptr, const double *D, const double *A, int nsub, const int *start,
                      int single, const double *y, double *x) {
  return single ? or_ras_ilu0_f32(N, E, ptr, D, A, nsub, start, y, x)
                : or_ras_ilu0_f64(N, E, ptr, D, A, nsub, start, y, x);
}

/* ROEFLUX5, type 0 (f77src/roeflux5.f): the Roe flux with the Roe-Turkel
 * low-Mach preconditioning of the dissipation (prec != 0), with Ugr = Ug,
 * Udr = Ud (FluxFcnSGApprJacRoeEuler3D without a Galerkin shift,
 * FluxFcnDescSG.C:127-131).  gamma is the dissipation coefficient, gam the
 * specific-heat ratio; mach, k1, cmach, shockreducer are the spatial
 * preconditioner's minimum Mach, slope, cut-off Mach and shock parameter
 * (LowMachPrec.h:119-124); irey the inverse Reynolds number; length the edge
 * length. */
void or_roeflux5(double gamma, double gam, double pstiff, const double *enormal, double evitno, const double *Ug,
                 const double *Ud, double mach, double k1, double cmach, double shockreducer, double irey,
                 double length, int prec, double *phi) {
  const double *Ugr = Ug, *Udr = Ud;
  double gam1 = gam - 1.0;
  double rnorm = sqrt(enormal[0] * enormal[0] + enormal[1] * enormal[1] + enormal[2] * enormal[2]);
  double invnorm = 1.0 / rnorm;
  double normal[3] = {enormal[0] * invnorm, enormal[1] * invnorm, enormal[2] * invnorm};
  double vitno = evitno * invnorm;
  /* centred terms */
  double VdotN = Ug[1] * normal[0] + Ug[2] * normal[1] + Ug[3] * normal[2];
  double vitg2 = Ug[1] * Ug[1] + Ug[2] * Ug[2] + Ug[3] * Ug[3];
  double H = gam * (Ug[4] + pstiff) + 0.5 * gam1 * Ug[0] * vitg2;
  H = H / (gam1 * Ug[0]);
  phi[0] = Ug[0] * (VdotN - vitno);
  phi[1] = phi[0] * Ug[1] + Ug[4] * normal[0];
  phi[2] = phi[0] * Ug[2] + Ug[4] * normal[1];
  phi[3] = phi[0] * Ug[3] + Ug[4] * normal[2];
  phi[4] = phi[0] * H + Ug[4] * vitno;
  VdotN = Ud[1] * normal[0] + Ud[2] * normal[1] + Ud[3] * normal[2];
  VdotN = VdotN - vitno;
  double vitd2 = Ud[1] * Ud[1] + Ud[2] * Ud[2] + Ud[3] * Ud[3];
  H = gam * (Ud[4] + pstiff) + 0.5 * gam1 * Ud[0] * vitd2;
  H = H / (gam1 * Ud[0]);
  phi[0] = phi[0] + Ud[0] * VdotN;
  phi[1] = phi[1] + Ud[0] * Ud[1] * VdotN + Ud[4] * normal[0];
  phi[2] = phi[2] + Ud[0] * Ud[2] * VdotN + Ud[4] * normal[1];
  phi[3] = phi[3] + Ud[0] * Ud[3] * VdotN + Ud[4] * normal[2];
  phi[4] = phi[4] + Ud[0] * VdotN * H + Ud[4] * vitno;
  /* Roe-averaged state */
  double squsr1 = sqrt(Ug[0]);
  double squsr2 = sqrt(Ud[0]);
  double ener1 = (Ug[4] + gam * pstiff) / gam1 + 0.5 * Ug[0] * vitg2;
  double ener2 = (Ud[4] + gam * pstiff) / gam1 + 0.5 * Ud[0] * vitd2;
  double usro = 1.0 / (squsr1 + squsr2);
  double uar2 = (squsr1 * Ug[1] + squsr2 * Ud[1]) * usro;
  double uar3 = (squsr1 * Ug[2] + squsr2 * Ud[2]) * usro;
  double uar4 = (squsr1 * Ug[3] + squsr2 * Ud[3]) * usro;
  double uar5 = ((ener1 + Ug[4]) / squsr1 + (ener2 + Ud[4]) / squsr2) * usro;
  /* dissipation */
  VdotN = normal[0] * uar2 + normal[1] * uar3 + normal[2] * uar4;
  double qir = 0.5 * (uar2 * uar2 + uar3 * uar3 + uar4 * uar4);
  double tet1 = normal[2] * uar3 - normal[1] * uar4;
  double tet2 = normal[0] * uar4 - normal[2] * uar2;
  double tet3 = normal[1] * uar2 - normal[0] * uar3;
  double cr2 = gam1 * (uar5 - qir);
  double cr = sqrt(cr2);
  cr2 = 1.0 / cr2;
  double beta;
  if (prec == 0) {
    beta = 1.0;
  } else {
    double shock = fabs(Ugr[4] - Udr[4]) / (Ugr[4] + Udr[4]) / length;
    double locMach = sqrt(2.0 * qir * cr2);
    beta = fmax(k1 * locMach, mach);
    beta = (1.0 + sqrt(irey)) * beta + shockreducer * shock;
    beta = fmin(beta, cmach);
  }
  double beta2 = beta * beta;
  double energ = (Ugr[4] + gam * pstiff) / gam1 + 0.5 * Ugr[0] * (Ugr[1] * Ugr[1] + Ugr[2] * Ugr[2] + Ugr[3] * Ugr[3]);
  double enerd = (Udr[4] + gam * pstiff) / gam1 + 0.5 * Udr[0] * (Udr[1] * Udr[1] + Udr[2] * Udr[2] + Udr[3] * Udr[3]);
  double dif1 = -Ugr[0] + Udr[0];
  double dif2 = -Ugr[0] * Ugr[1] + Udr[0] * Udr[1];
  double dif3 = -Ugr[0] * Ugr[2] + Udr[0] * Udr[2];
  double dif4 = -Ugr[0] * Ugr[3] + Udr[0] * Udr[3];
  double dif5 = -energ + enerd;
  double vp1 = VdotN;
  double sq = ((1.0 - beta2) * VdotN) * ((1.0 - beta2) * VdotN);
  double vp4 = 0.5 * ((1.0 + beta2) * VdotN + sqrt(sq + 4.0 * beta2 * (cr * cr)));
  double vp5 = 0.5 * ((1.0 + beta2) * VdotN - sqrt(sq + 4.0 * beta2 * (cr * cr)));
  /* Roe-Turkel coefficients */
  double r = vp4 - vp1 * beta2;
  double s = vp5 - vp1 * beta2;
  double t = 0.5 * (vp5 - vp4);
  vp1 = vp1 - vitno;
  vp4 = vp4 - vitno;
  vp5 = vp5 - vitno;
  double flur1 = fabs(vp1) * ((normal[0] * (1.0 - gam1 * qir * cr2) - tet1) * dif1 + (normal[0] * gam1 * uar2 * cr2) * dif2 +
                              (normal[2] + (normal[0] * gam1 * uar3 * cr2)) * dif3 +
                              (-normal[1] + (normal[0] * gam1 * uar4 * cr2)) * dif4 - (normal[0] * gam1 * cr2) * dif5);
  double flur2 = fabs(vp1) * ((normal[1] * (1.0 - gam1 * qir * cr2) - tet2) * dif1 +
                              (-normal[2] + (normal[1] * gam1 * uar2 * cr2)) * dif2 + (normal[1] * gam1 * uar3 * cr2) * dif3 +
                              (normal[0] + (normal[1] * gam1 * uar4 * cr2)) * dif4 - (normal[1] * gam1 * cr2) * dif5);
  double flur3 = fabs(vp1) * ((normal[2] * (1.0 - gam1 * qir * cr2) - tet3) * dif1 +
                              (normal[1] + (normal[2] * gam1 * uar2 * cr2)) * dif2 +
                              (-normal[0] + (normal[2] * gam1 * uar3 * cr2)) * dif3 + (normal[2] * gam1 * uar4 * cr2) * dif4 -
                              (normal[2] * gam1 * cr2) * dif5);
  double c2 = cr * cr;
  double flur4 = fabs(vp4) * ((c2 * VdotN / t + gam1 * qir * s / (t * beta2)) * dif1 -
                              (c2 * normal[0] / t + gam1 * uar2 * s / (t * beta2)) * dif2 -
                              (c2 * normal[1] / t + gam1 * uar3 * s / (t * beta2)) * dif3 -
                              (c2 * normal[2] / t + gam1 * uar4 * s / (t * beta2)) * dif4 + gam1 * s / (t * beta2) * dif5);
  double flur5 = fabs(vp5) * (-(c2 * VdotN / t + gam1 * qir * r / (beta2 * t)) * dif1 +
                              (c2 * normal[0] / t + gam1 * uar2 * r / (beta2 * t)) * dif2 +
                              (c2 * normal[1] / t + gam1 * uar3 * r / (beta2 * t)) * dif3 +
                              (c2 * normal[2] / t + gam1 * uar4 * r / (beta2 * t)) * dif4 - gam1 * r / (beta2 * t) * dif5);
  phi[0] = phi[0] - gamma * (normal[0] * flur1 + normal[1] * flur2 + normal[2] * flur3 + 0.5 * (flur4 + flur5) * cr2);
  phi[1] = phi[1] - gamma * ((uar2 * normal[0]) * flur1 + (uar2 * normal[1] - normal[2]) * flur2 +
                             (uar2 * normal[2] + normal[1]) * flur3 +
                             (normal[0] * (r * flur4 + s * flur5) + uar2 * (flur4 + flur5)) * 0.5 * cr2);
  phi[2] = phi[2] - gamma * ((uar3 * normal[0] + normal[2]) * flur1 + (uar3 * normal[1]) * flur2 +
                             (uar3 * normal[2] - normal[0]) * flur3 +
                             (normal[1] * (r * flur4 + s * flur5) + uar3 * (flur4 + flur5)) * 0.5 * cr2);
  phi[3] = phi[3] - gamma * ((uar4 * normal[0] - normal[1]) * flur1 + (uar4 * normal[1] + normal[0]) * flur2 +
                             (uar4 * normal[2]) * flur3 +
                             (normal[2] * (r * flur4 + s * flur5) + uar4 * (flur4 + flur5)) * 0.5 * cr2);
  phi[4] = phi[4] - gamma * ((qir * normal[0] + tet1) * flur1 + (qir * normal[1] + tet2) * flur2 +
                             (qir * normal[2] + tet3) * flur3 +
                             (VdotN * (r * flur4 + s * flur5) + uar5 * (flur4 + flur5)) * 0.5 * cr2);
  for (int k = 0; k < 5; ++k) phi[k] = phi[k] * 0.5 * rnorm;
}
