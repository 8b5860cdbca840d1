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

int or_residual_ex(const or_mesh *m, const or_params *p, const double *U, double *R,
                   double *Rabs, int fluxes_only, double *V_out, double *ddx_out,
                   double *ddy_out, double *ddz_out) {
  int N = m->N, E = m->E, F = m->F;
  double gam = p->gamma, pstiff = p->pstiff;
  double *V = V_out ? V_out : (double *)malloc(sizeof(double) * 5 * (size_t)N);
  double *ddx = 0, *ddy = 0, *ddz = 0;
  for (int k = 0; k < 5 * N; ++k) R[k] = 0.0;
  if (Rabs) for (int k = 0; k < 5 * N; ++k) Rabs[k] = 0.0;
  or_cons_to_prim(N, U, V, gam, pstiff);
  if (p->second_order) {
    ddx = ddx_out ? ddx_out : (double *)malloc(sizeof(double) * 5 * (size_t)N);
    ddy = ddy_out ? ddy_out : (double *)malloc(sizeof(double) * 5 * (size_t)N);
    ddz = ddz_out ? ddz_out : (double *)malloc(sizeof(double) * 5 * (size_t)N);
    or_gradients(N, E, m->X, m->ptr, m->R, V, ddx, ddy, ddz);
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
    if (Vi[0] <= 0.0 || Vi[4] + pstiff <= 0.0) ++ierr;
    if (Vj[0] <= 0.0 || Vj[4] + pstiff <= 0.0) ++ierr;
    if (ierr) continue;
    const double *nl = m->normals + 3 * l;
    if (sqrt(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]) < 1e-18) continue;
    or_roeflux7(gam, pstiff, nl, m->normal_vel[l], Vi, Vj, flux);
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
      if (code == 2 || code == -2) or_wallflux(n, nvel, V + 5 * node, flux);
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
