#!/usr/bin/env python3
"""Writes tests/golden/ref_kernels.npz: input/output vectors of the
reference's own hot-path kernels, compiled from the reference tree into
oracle/_ref (roeflux7.f, boundflux5.f, roejac6.f, RecFcnDesc.h Van Albada,
DenseMatrixOps.h LINPACK inverse).  Run in the build container, where the
reference tree exists; the fixture lets the oracle be pinned on machines
without it.  Inputs are seeded random physical states (rho, p > 0) and
dual-face normals.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402


def states(rng, n):
    V = np.empty((n, 5))
    V[:, 0] = rng.uniform(0.3, 3.0, n)
    V[:, 1:4] = rng.normal(0.0, 0.8, (n, 3))
    V[:, 4] = rng.uniform(0.2, 3.0, n)
    return V


def main():
    O.build()
    rng = np.random.default_rng(20240607)
    n = 400
    gam = 1.4
    nrm = rng.normal(size=(n, 3)) * rng.uniform(1e-3, 1.0, (n, 1))
    VL, VR = states(rng, n), states(rng, n)
    Ub = np.empty((n, 5))
    for k in range(n):  # conservative farfield states
        v = states(rng, 1)[0]
        Ub[k] = [v[0], v[0] * v[1], v[0] * v[2], v[0] * v[3], v[4] / (gam - 1) + 0.5 * v[0] * (v[1:4] @ v[1:4])]
    dd1, dd2 = rng.normal(0, 0.3, (n, 5)), rng.normal(0, 0.3, (n, 5))
    roe = np.array([O.roeflux7(gam, 0.0, nrm[k], 0.0, VL[k], VR[k], use_ref=True) for k in range(n)])
    bnd = np.array([O.boundflux5(gam, nrm[k], 0.0, VL[k], Ub[k], use_ref=True) for k in range(n)])
    va = [O.vanalbada5(1.0 / 3.0, VL[k], dd1[k], VR[k], dd2[k], use_ref=True) for k in range(n)]
    jac1 = np.array([O.roejac6(gam, 0.0, nrm[k], 0.0, VL[k], VR[k], 1, use_ref=True) for k in range(n)])
    jac2 = np.array([O.roejac6(gam, 0.0, nrm[k], 0.0, VR[k], VL[k], 2, use_ref=True) for k in range(n)])
    A = rng.normal(size=(n, 5, 5)) + 4.0 * np.eye(5)
    inv64 = np.array([O.invert5(A[k], single=False, use_ref=True)[0] for k in range(n)])
    inv32 = np.array([O.invert5(A[k], single=True, use_ref=True)[0] for k in range(n)])
    np.savez_compressed(os.path.join(HERE, "ref_kernels.npz"), gamma=gam, normal=nrm, VL=VL, VR=VR, Ub=Ub,
                        ddVij=dd1, ddVji=dd2, roeflux7=roe, boundflux5=bnd,
                        va_ij=np.array([a for a, _ in va]), va_ji=np.array([b for _, b in va]),
                        roejac6_sw1=jac1, roejac6_sw2=jac2, blocks=A, inv_f64=inv64,
                        inv_f32=inv32.astype(np.float32))
    print("wrote", os.path.join(HERE, "ref_kernels.npz"))
    write_lowmach()


def write_lowmach():
    """tests/golden/ref_lowmach.npz: ROEFLUX5 (Roe-Turkel low-Mach
    dissipation) and ROEJAC6 with the preconditioning on, from the compiled
    reference, on seeded states, normals, edge lengths and preconditioner
    parameters (ROEFLUX5 also with the preconditioning off)."""
    rng = np.random.default_rng(20241203)
    n = 400
    gam = 1.4
    nrm = rng.normal(size=(n, 3)) * rng.uniform(1e-3, 1.0, (n, 1))
    VL, VR = states(rng, n), states(rng, n)
    VL[:, 1:4] *= 0.1  # low-Mach states: |u| well below the speed of sound
    VR[:, 1:4] *= 0.1
    par = np.column_stack([rng.uniform(0.005, 0.2, n),   # minimum Mach
                           rng.uniform(0.5, 2.0, n),     # slope k1
                           rng.uniform(0.5, 1.0, n),     # cut-off Mach
                           rng.uniform(0.0, 0.5, n),     # shock parameter
                           np.where(np.arange(n) % 2 == 0, 0.0, rng.uniform(0.0, 1e-3, n)),  # 1/Re
                           rng.uniform(0.01, 0.2, n)])   # edge length
    prec = (np.arange(n) % 4 != 0).astype(np.int32)
    out = np.array([O.roeflux5(1.0, gam, 0.0, nrm[k], 0.0, VL[k], VR[k], *par[k], int(prec[k]), use_ref=True)
                    for k in range(n)])
    # ROEJAC6 with the preconditioning on (roejacappr3Dgas, FluxFcnDescSG.C:157-175)
    j1 = np.array([O.roejac6(gam, 0.0, nrm[k], 0.0, VL[k], VR[k], 1, low_mach=par[k], use_ref=True)
                   for k in range(n)])
    j2 = np.array([O.roejac6(gam, 0.0, nrm[k], 0.0, VR[k], VL[k], 2, low_mach=par[k], use_ref=True)
                   for k in range(n)])
    np.savez_compressed(os.path.join(HERE, "ref_lowmach.npz"), gamma=gam, normal=nrm, VL=VL, VR=VR, params=par,
                        prec=prec, roeflux5=out, roejac6_sw1=j1, roejac6_sw2=j2)
    print("wrote", os.path.join(HERE, "ref_lowmach.npz"))


if __name__ == "__main__":
    main()
