// ref_harness.cpp -- C entry points into reference code compiled from the
// reference tree itself (oracle/ref/Makefile).  TEST INFRASTRUCTURE ONLY:
// used by tests/ to pin the oracle restatement.  Nothing here is shipped.
//
// Compiled with -I<reference tree>; the reference's own Vector.h comes first
// because RecFcn.h uses SVec without including it.
#include <Vector.h>
#include <RecFcnDesc.h>
#include <DenseMatrixOps.h>

extern "C" {
// Fortran kernels (f77src/*.f), compiled by amdflang into the same library.
void roeflux7_(const int *type, const double *gam, const double *pstiff,
               double *enormal, const double *evitno, double *Ug, double *Ud,
               double *phi);
void boundflux5_(const int *type, const double *gam, double *normal,
                 const double *sig, double *ua, double *ub, double *flux);
void roejac6_(const int *type, const double *gamma, const double *gam,
              const double *pstiff, double *enormal, const double *evitno,
              double *Ug, double *Ud, double *jac, const int *sw,
              const double *mach, const double *k1, const double *cmach,
              const double *shockreducer, const double *irey,
              const double *length, const int *prec);

void roeflux5_(const int *type, const double *gamma, const double *gam, const double *pstiff, double *enormal,
               const double *evitno, double *Ugr, double *Ug, double *Udr, double *Ud, double *phi,
               const double *mach, const double *k1, const double *cmach, const double *shockreducer,
               const double *irey, const double *length, const int *prec);

// ROEFLUX5 (f77src/roeflux5.f) with type = 0 and the reconstructed states
// as both argument pairs (Ugr = Ug, Udr = Ud), as FluxFcnSGApprJacRoeEuler3D
// calls it without a Galerkin shift (FluxFcnDescSG.C:127-131).
void refh_roeflux5(double gamma, double gam, double pstiff, const double *n, double vn, const double *Vl,
                   const double *Vr, double mach, double k1, double cmach, double shock, double irey, double length,
                   int prec, double *phi) {
  int type = 0;
  double nn[3] = {n[0], n[1], n[2]};
  double a[7] = {0}, b[7] = {0}, f[7] = {0};
  for (int k = 0; k < 5; ++k) { a[k] = Vl[k]; b[k] = Vr[k]; }
  roeflux5_(&type, &gamma, &gam, &pstiff, nn, &vn, a, a, b, b, f, &mach, &k1, &cmach, &shock, &irey, &length, &prec);
  for (int k = 0; k < 5; ++k) phi[k] = f[k];
}

// ROEFLUX7 (f77src/roeflux7.f) with type = 0.
void refh_roeflux7(double gam, double pstiff, const double *n, double vn,
                   const double *Vl, const double *Vr, double *phi) {
  int type = 0;
  double nn[3] = {n[0], n[1], n[2]};
  double a[7] = {0}, b[7] = {0}, f[7] = {0};
  for (int k = 0; k < 5; ++k) { a[k] = Vl[k]; b[k] = Vr[k]; }
  roeflux7_(&type, &gam, &pstiff, nn, &vn, a, b, f);
  for (int k = 0; k < 5; ++k) phi[k] = f[k];
}

// BOUNDFLUX5 (f77src/boundflux5.f) with type = 0.
void refh_boundflux5(double gam, const double *n, double sig, const double *ua,
                     const double *ub, double *flux) {
  int type = 0;
  double nn[3] = {n[0], n[1], n[2]};
  double a[7] = {0}, b[7] = {0}, f[7] = {0};
  for (int k = 0; k < 5; ++k) { a[k] = ua[k]; b[k] = ub[k]; }
  boundflux5_(&type, &gam, nn, &sig, a, b, f);
  for (int k = 0; k < 5; ++k) flux[k] = f[k];
}

// ROEJAC6 (f77src/roejac6.f) with type = 0, prec = 0, as called by
// roejacappr3Dgas (FluxFcnDescSG.C:171-172).
void refh_roejac6_lm(double gamma, double gam, double pstiff, const double *n, double vn, const double *Vl,
                     const double *Vr, int sw, double mach, double k1, double cmach, double shock, double irey,
                     double length, int prec, double *jac) {
  int type = 0;
  double nn[3] = {n[0], n[1], n[2]};
  double a[7] = {0}, b[7] = {0}, j[49] = {0};
  for (int k = 0; k < 5; ++k) { a[k] = Vl[k]; b[k] = Vr[k]; }
  roejac6_(&type, &gamma, &gam, &pstiff, nn, &vn, a, b, j, &sw, &mach, &k1, &cmach, &shock, &irey, &length, &prec);
  for (int k = 0; k < 25; ++k) jac[k] = j[k];
}

void refh_roejac6(double gamma, double gam, double pstiff, const double *n,
                  double vn, const double *Vl, const double *Vr, int sw,
                  double *jac) {
  int type = 0, prec = 0;
  double nn[3] = {n[0], n[1], n[2]};
  double a[7] = {0}, b[7] = {0}, j[49] = {0};
  double mach = 1.0, k1 = 0.0, cmach = 1.0, shock = 0.0, irey = 0.0, length = 0.0;
  for (int k = 0; k < 5; ++k) { a[k] = Vl[k]; b[k] = Vr[k]; }
  roejac6_(&type, &gamma, &gam, &pstiff, nn, &vn, a, b, j, &sw, &mach, &k1,
           &cmach, &shock, &irey, &length, &prec);
  for (int k = 0; k < 25; ++k) jac[k] = j[k];
}

// RecFcnVanAlbada<5>::compute (RecFcnDesc.h:705-713), eps input 0.1 (only
// eps3 depends on it; Van Albada uses eps0 = 1e-16).
void refh_vanalbada5(double beta, const double *Vi, const double *ddVij,
                     const double *Vj, const double *ddVji, double *Vij,
                     double *Vji) {
  RecFcnVanAlbada<5> rec(beta, 0.1);
  double a[5], b[5], c[5], d[5];
  for (int k = 0; k < 5; ++k) { a[k] = Vi[k]; b[k] = ddVij[k]; c[k] = Vj[k]; d[k] = ddVji[k]; }
  rec.compute(a, b, c, d, Vij, Vji);
}

// invertDenseMatrix<Scalar,5> (DenseMatrixOps.h:1163-1183): in place.
int refh_invert5_f32(float *a) { return invertDenseMatrix<float, 5>(a); }
int refh_invert5_f64(double *a) { return invertDenseMatrix<double, 5>(a); }

// DenseMatrixOp<Scalar,5,25>::applyToVector as used by DiagMat::apply
// (DiagMatrix.C:86-100).
void refh_apply5_f32(const float *a, const double *y, double *x) {
  float aa[1][25];
  double yy[1][5], xx[1][5];
  for (int k = 0; k < 25; ++k) aa[0][k] = a[k];
  for (int k = 0; k < 5; ++k) yy[0][k] = y[k];
  DenseMatrixOp<float, 5, 25>::applyToVector(aa, 0, yy, 0, xx, 0);
  for (int k = 0; k < 5; ++k) x[k] = xx[0][k];
}
void refh_apply5_f64(const double *a, const double *y, double *x) {
  double aa[1][25];
  double yy[1][5], xx[1][5];
  for (int k = 0; k < 25; ++k) aa[0][k] = a[k];
  for (int k = 0; k < 5; ++k) yy[0][k] = y[k];
  DenseMatrixOp<double, 5, 25>::applyToVector(aa, 0, yy, 0, xx, 0);
  for (int k = 0; k < 5; ++k) x[k] = xx[0][k];
}
}
