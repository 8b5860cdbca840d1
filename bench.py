#!/usr/bin/env python3
"""FP64 edge-residual throughput on one MI355X.

Workload: 128^3-hex Kuhn-split unit box (2,146,689 nodes, 12,582,912 tets,
14,827,904 edges, 196,608 boundary triangles), interior nodes jittered
+-0.15h (Philox, seed 1234), renumbered with the reference RCM rule;
perfect gas, freestream M = 0.8, density x (1 + 0.02 sin sin sin) with
seeded phases; second-order MUSCL (LSQ gradients, Van Albada, Roe flux),
Steger-Warming farfield on 5 faces and a slip wall at z = 0.

A step is one full residual evaluation (gradient pass + flux pass, 1/V
scaling) with U resident in HBM.  value = edges / (ms per step) in Medges/s.

roofline: one residual evaluation is three kernels (k_gradients,
k_flux_edges, k_flux_nodes).  Its algorithmic bytes are 464 B/node +
40 B/edge + 40 B/face: the gradient pass reads U, X and the LSQ factors and
writes the gradients (232 B/node) plus an edge pair (8 B/edge); the flux pass
reads U, X, gradients and 1/V and writes R (232 B/node), plus edge pair and
normal (32 B/edge) and face nodes, normal and code (40 B/face).  They are
divided by the sum of the three kernels' mean durations from HIP events
bracketing each launch on the stream the kernels run on.  traffic (HBM bytes
per evaluation) comes from the rocprofv3 FETCH_SIZE/WRITE_SIZE summary under
profiles/.

cpu_baseline: the oracle's C restatement of the same residual on the same mesh
and state, on the host's cores (OpenMP over contiguous node partitions with the
shared-node assembly, as the reference runs its subdomains), timed over a
bounded number of evaluations; the one-thread figure is reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
NX = 128


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def load_traffic(name, key="bytes_per_launch"):
    """Per-launch counter figure (HBM bytes, FP64 lane-operations) from
    profiles/traffic_<name>.json (rocprofv3 FETCH_SIZE/WRITE_SIZE and
    SQ_INSTS_VALU_*_F64 of the same workload)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(key)
    except (OSError, ValueError):
        return None


def cpu_baseline(mesh, params, U, budget_s):
    """The oracle's residual on the host: OpenMP over contiguous node
    partitions with the shared-node assembly, as the reference runs its
    subdomains (one per thread), and the same on one thread."""
    import numpy as np
    import oracle as O
    t0 = time.time()
    om = O.OracleMesh(mesh, params)
    log(f"[bench] oracle set-up {time.time() - t0:.1f} s")
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))

    def timed(threads, budget, max_evals):
        times = []
        t_start = time.time()
        while True:
            t1 = time.perf_counter()
            out = om.residual_par(U, threads)
            times.append(time.perf_counter() - t1)
            if out["ierr"] != 0:
                raise RuntimeError("oracle residual reported non-physical states")
            if time.time() - t_start > budget or len(times) >= max_evals:
                return float(np.median(times)), len(times)

    med, k = timed(cores, budget_s, 10)
    med1, k1 = timed(1, budget_s / 2, 3)
    return dict(value=om.E / med / 1e6, unit="Medges/s", cores=cores, kind="port",
                sample=f"{k} full residual evaluations on the same 128^3 mesh and state, {cores} OpenMP threads "
                       f"over {cores} node partitions with the shared-node assembly, median {med * 1e3:.0f} ms each",
                single_thread_value=om.E / med1 / 1e6,
                single_thread_sample=f"{k1} evaluations on one thread, median {med1 * 1e3:.0f} ms each")


def copy_peak(dev, gib=1.0, reps=10):
    """Measured device-to-device copy rate (bytes read + written per second)."""
    import torch
    n = int(gib * (1 << 30)) // 8
    a = torch.ones(n, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * n * 8 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    return round(gbs, 1)


def newton_bench(dev, max_steps):
    """Steady Newton-GMRES on the 96x32x32 bump channel, FP64 vs mixed."""
    import numpy as np
    import torch

    from aerof_dev_amd import FlowParams, init_state
    from aerof_dev_amd import mesh as M
    from aerof_dev_amd.operators import SpaceOperator, steady_solve

    mesh = M.channel_mesh(96, 32, 32, bump_height=0.05, bump_width=0.2)
    params = FlowParams(mach=0.675)
    U0 = init_state(params, mesh.X, "freestream")
    op = SpaceOperator(mesh, params)
    # warm-up: allocations and first launches
    Uw = torch.from_numpy(U0).to(dev)
    steady_solve(op, Uw, precision=0, max_steps=2)
    steady_solve(op, Uw, precision=1, max_steps=2)
    out = {"workload": f"steady Euler, 96x32x32 Kuhn channel with Gaussian bump ({op.num_nodes} nodes, "
                       f"{op.num_edges} edges), M=0.675, Newton-GMRES(30) + block-Jacobi, CFL 10->1e4, "
                       "stop at ||R||/||R0|| <= 1e-8"}
    states = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        Ud = torch.from_numpy(U0).to(dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = steady_solve(op, Ud, precision=prec, max_steps=max_steps, eps=1e-8, cfl0=10.0, cfl_max=1e4)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        states[name] = Ud.cpu().numpy()
        out[name] = {"seconds": round(wall, 4), "steps": r["steps"], "krylov_its": r["krylov_its"],
                     "converged": r["converged"], "rel_res": float(r["res"] / r["res0"])}
        log(f"[bench] newton {name}: {wall:.3f} s, {r['steps']} steps, {r['krylov_its']} Krylov its, "
            f"converged={r['converged']} rel_res={r['res'] / r['res0']:.2e}")
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["seconds"] / out["mixed"]["seconds"], 3)
    # matrix-free Newton-Krylov (MatVecProdFD products of the FP64 residual,
    # one residual per Krylov product; FP32 block-Jacobi of H1)
    steady_solve(op, torch.from_numpy(U0).to(dev), precision=1, max_steps=2, mvp="fd")
    Ud = torch.from_numpy(U0).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = steady_solve(op, Ud, precision=1, max_steps=max_steps, eps=1e-8, cfl0=10.0, cfl_max=1e4, mvp="fd")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    Ufd = Ud.cpu().numpy()
    out["matrix_free"] = {"seconds": round(wall, 4), "steps": r["steps"], "krylov_its": r["krylov_its"],
                          "converged": r["converged"], "rel_res": float(r["res"] / r["res0"]),
                          "preconditioner": "FP32 block-Jacobi of H1",
                          "speedup_vs_h1_fp64": round(out["fp64"]["seconds"] / wall, 3),
                          "state_rel_diff_vs_h1_fp64": float(np.linalg.norm(Ufd - states["fp64"]) /
                                                             np.linalg.norm(states["fp64"]))}
    log(f"[bench] newton matrix-free: {wall:.3f} s, {r['steps']} steps, {r['krylov_its']} Krylov its, "
        f"converged={r['converged']}")
    # exact second-order products (MatVecProdH2, forward-mode kernels), FP32
    # block-Jacobi of H1
    steady_solve(op, torch.from_numpy(U0).to(dev), precision=1, max_steps=2, mvp="h2")
    Ud = torch.from_numpy(U0).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = steady_solve(op, Ud, precision=1, max_steps=max_steps, eps=1e-8, cfl0=10.0, cfl_max=1e4, mvp="h2")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out["exact_products"] = {"seconds": round(wall, 4), "steps": r["steps"], "krylov_its": r["krylov_its"],
                             "converged": r["converged"], "rel_res": float(r["res"] / r["res0"]),
                             "preconditioner": "FP32 block-Jacobi of H1",
                             "speedup_vs_h1_fp64": round(out["fp64"]["seconds"] / wall, 3),
                             "state_rel_diff_vs_h1_fp64": float(np.linalg.norm(Ud.cpu().numpy() - states["fp64"]) /
                                                                np.linalg.norm(states["fp64"]))}
    log(f"[bench] newton exact products: {wall:.3f} s, {r['steps']} steps, {r['krylov_its']} Krylov its")
    # per-kernel-bucket breakdown from a separate profiled run of each precision
    for name, prec in (("fp64", 0), ("mixed", 1)):
        Ud = torch.from_numpy(U0).to(dev)
        op.profile(True)
        steady_solve(op, Ud, precision=prec, max_steps=max_steps, eps=1e-8, cfl0=10.0, cfl_max=1e4)
        tm = op.timers()
        op.profile(False)
        out[name]["kernel_ms"] = {k: round(v[0], 2) for k, v in tm.items()}
        log(f"[bench] newton {name} kernel ms: {out[name]['kernel_ms']}")
    out["state_rel_diff_mixed_vs_fp64"] = float(np.linalg.norm(states["mixed"] - states["fp64"]) /
                                                np.linalg.norm(states["fp64"]))
    op.close()
    # RAS with ILU(0) subdomain solves on the same channel numbered by RCM (the
    # reference's IluPrec renumbering), against block-Jacobi on that numbering
    mesh_r = M.renumber(mesh, M.rcm_gpu(mesh))
    opr = SpaceOperator(mesh_r, params)
    U0r = init_state(params, mesh_r.X, "freestream")
    ras = {"workload": "the same solve on the channel numbered by RCM, subdomains of 512 nodes"}
    for name, prec, pc in (("jacobi_mixed", 1, "jacobi"), ("ras_ilu0_mixed", 1, "ras_ilu0"),
                           ("ras_ilu0_fp64", 0, "ras_ilu0")):
        steady_solve(opr, torch.from_numpy(U0r).to(dev), precision=prec, max_steps=2, prec=pc, prec_subdomain=512)
        Ud = torch.from_numpy(U0r).to(dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = steady_solve(opr, Ud, precision=prec, max_steps=max_steps, eps=1e-8, cfl0=10.0, cfl_max=1e4, prec=pc,
                         prec_subdomain=512)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ras[name] = {"seconds": round(wall, 4), "steps": r["steps"], "krylov_its": r["krylov_its"],
                     "converged": r["converged"], "rel_res": float(r["res"] / r["res0"])}
        log(f"[bench] newton {name} (RCM numbering): {wall:.3f} s, {r['steps']} steps, {r['krylov_its']} Krylov its")
    opr.close()
    out["ras_ilu0"] = ras
    return out


def batched_bench(dev, n, nrhs=16, m=12):
    """Block GMRES over 16 right-hand sides on the 128^3 box operator
    (H1 + 1/dt at CFL 100, block-Jacobi), mixed precision in one captured
    graph per cycle, against 16 single-RHS FP64 GMRES solves."""
    import numpy as np
    import torch

    from aerof_dev_amd import FlowParams, init_state
    from aerof_dev_amd import mesh as M
    from aerof_dev_amd.operators import (BlockGmresSolver, GmresSolver, MatVecProdH1, SpaceOperator)

    mesh = M.box_mesh(n, jitter=0.15, seed=1234, rcm_order=True)
    params = FlowParams(mach=0.8)
    U = torch.from_numpy(init_state(params, mesh.X, "sine", 0.02, seed=2024)).to(dev)
    op = SpaceOperator(mesh, params)
    h1 = MatVecProdH1(op)
    dt = torch.empty(op.num_nodes, dtype=torch.float64, device=dev)
    h1.timestep(U, 100.0, dt)
    h1.evaluate(U, dt, 0)
    g = torch.Generator(device="cpu").manual_seed(7)
    B = torch.randn((nrhs, op.num_nodes, 5), generator=g, dtype=torch.float64).to(dev)
    X = torch.empty_like(B)
    eps = 1e-8
    blk = BlockGmresSolver(op, m=m, eps=eps, max_cycles=400, precision=1, use_graph=True)
    blk.solve(B, X)  # first call: set-up, autotune, graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = blk.solve(B, X)
    torch.cuda.synchronize()
    t_blk = time.perf_counter() - t0
    single = GmresSolver(op, 30, eps, 30 * 400)
    Xs = torch.empty_like(B)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = 0
    for k in range(nrhs):
        xk = torch.empty_like(B[k])
        i, _ = single.solve(B[k].contiguous(), xk)
        its += i
        Xs[k] = xk
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    err = max(float((X[k] - Xs[k]).norm() / Xs[k].norm()) for k in range(nrhs))
    op.close()
    res = {"workload": f"{nrhs} RHS on H1 + 1/dt (CFL 100) with block-Jacobi, {n}^3 box ({op.num_nodes} nodes), "
                       f"stop at ||b - A x|| <= {eps} ||b|| per RHS",
           "block_mixed_graph_s": round(t_blk, 4), "block_cycles": out["cycles"], "block_m": m,
           "block_autotune_waves": list(out["tune"]),
           "single_fp64_s": round(t_single, 4), "single_fp64_krylov_its": its,
           "speedup_block_vs_single": round(t_single / t_blk, 3),
           "max_rel_solution_diff_vs_single_fp64": err}
    log(f"[bench] batched: {res}")
    return res


def adjoint_bench(dev, reps=5):
    """Adjoint kernels on the 192x64x64 bump channel (815k nodes): time of one
    (dR/dU)^T lam + (dR/dX)^T lam evaluation against one residual."""
    import numpy as np
    import torch

    from aerof_dev_amd import FlowParams, init_state
    from aerof_dev_amd import mesh as M
    from aerof_dev_amd.operators import SpaceOperator

    mesh = M.channel_mesh(192, 64, 64, bump_height=0.05, bump_width=0.2)
    params = FlowParams(mach=0.675)
    op = SpaceOperator(mesh, params)
    U = torch.from_numpy(init_state(params, mesh.X, "sine", 0.02, seed=11)).to(dev)
    lam = torch.from_numpy(np.random.default_rng(3).standard_normal((op.num_nodes, 5))).to(dev)
    R = torch.empty_like(U)
    gu = torch.empty_like(U)
    gx = torch.empty((op.num_nodes, 3), dtype=torch.float64, device=dev)
    op.compute_residual(U, R)
    op.compute_transpose_derivative_of_residual(U, lam, gu, gx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        op.compute_residual(U, R)
    torch.cuda.synchronize()
    t_res = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        op.compute_transpose_derivative_of_residual(U, lam, gu, None)
    torch.cuda.synchronize()
    t_u = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        op.compute_transpose_derivative_of_residual(U, lam, gu, gx)
    torch.cuda.synchronize()
    t_ux = (time.perf_counter() - t0) / reps
    op.close()
    res = {"workload": f"192x64x64 bump channel ({op.num_nodes} nodes, {op.num_edges} edges), second order",
           "residual_ms": round(t_res * 1e3, 3), "dRdU_T_lambda_ms": round(t_u * 1e3, 3),
           "dRdU_and_dRdX_T_lambda_ms": round(t_ux * 1e3, 3)}
    log(f"[bench] adjoint: {res}")
    return res


def viscous_bench(dev, n=NX, steps=20):
    """Laminar Navier-Stokes terms on the 128^3 Kuhn box (no-slip adiabatic
    z- wall, M 0.5, Re 1000, Sutherland): the residual with the Galerkin term,
    the Galerkin kernel alone (feTerm), and one H1 assembly with its Galerkin
    blocks (feJac), FP64 and FP32."""
    import torch

    from aerof_dev_amd import FlowParams, ViscousParams, init_state
    from aerof_dev_amd import mesh as M
    from aerof_dev_amd.operators import MatVecProdH1, SpaceOperator

    codes = list(M.FARFIELD_BOX)
    codes[4] = 3
    mesh = M.kuhn_box(n, n, n, jitter=0.15, seed=1234, face_codes=tuple(codes))
    mesh = M.renumber(mesh, M.rcm_gpu(mesh))
    mach = 0.5
    params = FlowParams(mach=mach, p_inf=1.0 / (1.4 * mach * mach), viscous=ViscousParams(reynolds=1000.0))
    op = SpaceOperator(mesh, params)
    N, E, T = op.num_nodes, op.num_edges, mesh.num_tets
    U = torch.from_numpy(init_state(params, mesh.X, "sine", 0.02, seed=2024)).to(dev)
    R = torch.empty_like(U)
    if op.compute_residual(U, R) != 0:
        raise RuntimeError("non-physical reconstruction in the viscous bench state")
    for _ in range(3):
        op.compute_residual_async(U, R)
    torch.cuda.synchronize()
    op.profile(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        op.compute_residual_async(U, R)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) / steps * 1e3
    tm = op.timers()
    op.profile(False)
    # feTerm: the element pass; the node sums of the Galerkin term run inside
    # the finite-volume node pass (fvTermNodes)
    fe_ms = tm["feTerm"][0] / tm["feTerm"][1]
    fv_ms = sum(tm[k][0] / tm[k][1] for k in ("nodalGrad", "fvTerm", "fvTermNodes") if k in tm)
    # algorithmic bytes of the Galerkin pass: per node V, X (72 B), R read and
    # written (80 B), its incidence row (4 B); per element its 4 node ids
    # (16 B) and 4 incidence slots (16 B)
    fe_bytes = 156 * N + 32 * T
    h1 = MatVecProdH1(op)
    dt = torch.empty(N, dtype=torch.float64, device=dev)
    h1.timestep(U, 100.0, dt)
    jac = {}
    for precision in (0, 1):
        h1.evaluate(U, dt, precision)
        torch.cuda.synchronize()
        op.profile(True)
        for _ in range(3):
            h1.evaluate(U, dt, precision)
        torch.cuda.synchronize()
        tj = op.timers()
        op.profile(False)
        esz = 8 if precision == 0 else 4
        # per block read and written once (25 entries), plus the node data
        jb = 2 * 25 * esz * (2 * E + N) + 72 * N + 56 * T
        fj = tj["feJac"][0] / tj["feJac"][1]
        jac["fp64" if precision == 0 else "fp32"] = {
            "feJac_ms": round(fj, 4), "fvJac_ms": round(tj["fvJac"][0] / tj["fvJac"][1], 4),
            "algorithmic_GBs": round(jb / (fj * 1e-3) / 1e9, 1)}
    # the state derivatives with the Galerkin term: one forward product and
    # one transposed product (each with its forward residual pass)
    lam = torch.randn_like(U)
    dU = torch.randn_like(U)
    out_t, out_a = torch.empty_like(U), torch.empty_like(U)
    op.compute_derivative_of_residual(U, out_t, dU=dU)
    op.compute_transpose_derivative_of_residual(U, lam, out_a, None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        op.compute_derivative_of_residual(U, out_t, dU=dU)
    torch.cuda.synchronize()
    tan_ms = (time.perf_counter() - t0) / 3 * 1e3
    t0 = time.perf_counter()
    for _ in range(3):
        op.compute_transpose_derivative_of_residual(U, lam, out_a, None)
    torch.cuda.synchronize()
    adj_ms = (time.perf_counter() - t0) / 3 * 1e3
    op.close()
    res = {"workload": f"{n}^3 Kuhn box ({N} nodes, {T} elements, {E} edges), no-slip z- wall, M {mach}, Re 1000, "
                       "Sutherland, second order",
           "residual_ms": round(wall_ms, 4), "finite_volume_ms": round(fv_ms, 4), "feTerm_ms": round(fe_ms, 4),
           "feTerm_algorithmic_GBs": round(fe_bytes / (fe_ms * 1e-3) / 1e9, 1),
           "feTerm_Melements_per_s": round(T / (fe_ms * 1e-3) / 1e6, 1), "jacobian": jac,
           "dRdU_dU_ms": round(tan_ms, 3), "dRdU_T_lambda_ms": round(adj_ms, 3)}
    log(f"[bench] viscous: {res}")
    return res


def shape_bench(dev, order=1, ncheck=1):
    """Adjoint shape gradient on the 192x64x64 bump channel (815k nodes):
    the wall lift J (pressure force on the slip walls along z) and its
    derivatives with respect to 8 Gaussian bump amplitudes with seeded centres.
    Steady solve (mixed precision) to 1e-11, mixed-precision adjoint solve,
    (dR/dX)^T psi, dJ/da_k = (dJ/dX + (dR/dX)^T psi) . dX/da_k; the first
    ncheck amplitudes against central differences of re-converged J (step
    1e-6; each side starts from the converged unperturbed state and drops
    its own initial residual by 1e-9)."""
    import numpy as np
    import torch

    from aerof_dev_amd import FlowParams, init_state
    from aerof_dev_amd import mesh as M
    from aerof_dev_amd.operators import SpaceOperator, adjoint_solve, steady_solve

    zdir, h = (0.0, 0.0, 1.0), 1e-6
    centres = np.random.default_rng(2024).uniform(0.9, 2.1, size=8)

    def mesh(amps):
        bumps = [(0.05, 1.5, 0.2)] + [(a, c, 0.2) for a, c in zip(amps, centres)]
        return M.kuhn_box(192, 64, 64, lo=(0.0, 0.0, 0.0), hi=(3.0, 1.0, 1.0), face_codes=M.CHANNEL, bumps=bumps)

    def converge(op, U, eps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = steady_solve(op, U, precision=1, max_steps=3000, eps=eps, cfl0=10.0, cfl_max=1e4)
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    params = FlowParams(mach=0.675, second_order=order == 2)
    m = mesh([0.0] * 8)
    op = SpaceOperator(m, params)
    U = torch.from_numpy(init_state(params, m.X, "freestream")).to(dev)
    r, t_steady = converge(op, U, 1e-11)
    dJdU = torch.empty_like(U)
    dJdX = torch.empty((m.num_nodes, 3), dtype=torch.float64, device=dev)
    op.wall_force(U, zdir, dJdU, dJdX)
    psi = torch.empty_like(U)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = adjoint_solve(op, U, (-dJdU).contiguous(), psi, precision=1, cfl=1e4, eps=1e-11, max_its=3000)
    gx = torch.empty_like(dJdX)
    op.compute_transpose_derivative_of_residual(U, psi, None, gx)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    total = (dJdX + gx).cpu().numpy()
    grads = []
    for k in range(8):
        amps = [0.0] * 8
        amps[k] = 1.0
        grads.append(float(np.sum(total * (mesh(amps).X - m.X))))
    op.close()
    checks = []
    for k in range(ncheck):
        Js = []
        for sgn in (1.0, -1.0):
            amps = [0.0] * 8
            amps[k] = sgn * h
            opk = SpaceOperator(mesh(amps), params)
            Uk = U.clone()
            converge(opk, Uk, 1e-9)
            Js.append(opk.wall_force(Uk, zdir))
            opk.close()
        fd = (Js[0] - Js[1]) / (2 * h)
        checks.append({"k": k, "adjoint": grads[k], "fd": fd, "rel_diff": abs(grads[k] - fd) / abs(fd)})
    res = {"workload": f"192x64x64 bump channel ({m.num_nodes} nodes), {'second' if order == 2 else 'first'} order, "
                       "wall lift, 8 bump amplitudes",
           "steady_mixed_s": round(t_steady, 3), "steady_steps": r["steps"], "steady_krylov_its": r["krylov_its"],
           "steady_rel_res": float(r["res"] / r["res0"]), "adjoint_mixed_s": round(t_adj, 3),
           "adjoint_outer_its": st["its"], "adjoint_krylov_its": st["krylov_its"], "adjoint_converged": st["converged"],
           "dJ_da": grads, "fd_checks": checks}
    log(f"[bench] shape: {res}")
    return res


def sensitivity_bench(dev, n=NX, ndesign=16):
    """Batched direct sensitivities on the n^3 Kuhn box with 8 jittered wall
    bumps: a mixed-precision steady solve, then 16 right-hand sides
    -(dR/dX) dX/ds_k (the amplitude and the centre of each bump, closed-form
    dX/ds from the generator, forward-mode (dR/dX)), solved together (block
    GMRES with FP32 inner cycles in captured graphs, FP64 refinement with the
    exact products) and one at a time in FP64."""
    import numpy as np
    import torch

    from aerof_dev_amd import FlowParams, init_state
    from aerof_dev_amd import mesh as M
    from aerof_dev_amd.operators import SpaceOperator, direct_solve, steady_solve

    rng = np.random.default_rng(5)
    centres = rng.uniform(0.25, 0.75, size=8)
    amps = rng.uniform(0.004, 0.008, size=8)
    mesh = M.kuhn_box(n, n, n, jitter=0.15, seed=1234, bumps=[(a, c, 0.1) for a, c in zip(amps, centres)])
    mesh = M.renumber(mesh, M.rcm_gpu(mesh))
    params = FlowParams(mach=0.8)
    op = SpaceOperator(mesh, params)
    N = op.num_nodes
    U = torch.from_numpy(init_state(params, mesh.X, "freestream")).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = steady_solve(op, U, precision=1, max_steps=400, eps=1e-8, cfl0=10.0, cfl_max=1e4)
    torch.cuda.synchronize()
    t_steady = time.perf_counter() - t0
    design = [(b, prm) for b in range(8) for prm in ("amplitude", "center")][:ndesign]
    rhs = torch.empty((len(design), N, 5), dtype=torch.float64, device=dev)
    for k, (b, prm) in enumerate(design):
        dX = torch.from_numpy(M.bump_derivative(mesh, b, prm)).to(dev)
        dR = torch.empty_like(U)
        op.compute_derivative_of_residual(U, dR, None, dX)
        rhs[k] = -dR
    eps = 1e-10
    x = torch.empty_like(rhs)
    direct_solve(op, U, rhs[:1].contiguous(), x[:1].contiguous(), precision=1, eps=1e-2, max_its=2)  # set-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sb = direct_solve(op, U, rhs, x, precision=1, cfl=1e4, eps=eps, max_its=400)
    torch.cuda.synchronize()
    t_blk = time.perf_counter() - t0
    t0 = time.perf_counter()
    worst, its = 0.0, 0
    xs = []
    for k in range(len(design)):
        xk = torch.empty_like(U)
        s1 = direct_solve(op, U, rhs[k].contiguous(), xk, precision=0, cfl=1e4, eps=eps, max_its=400)
        its += s1["krylov_its"]
        xs.append(xk)
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    for k in range(len(design)):
        worst = max(worst, float(torch.linalg.norm(x[k] - xs[k]) / torch.linalg.norm(xs[k])))
    op.close()
    res = {"workload": f"{len(design)} direct sensitivities (8 bump amplitudes and centres), {n}^3 Kuhn box with "
                       f"jittered wall bumps ({N} nodes), second order, M=0.8, stop at "
                       f"||rhs - dR/dU x|| <= {eps} ||rhs|| per RHS",
           "steady_mixed_s": round(t_steady, 3), "steady_steps": r["steps"], "steady_converged": r["converged"],
           "block_mixed_s": round(t_blk, 3), "block_outer_its": sb["its"], "block_inner_its": sb["krylov_its"],
           "block_converged": sb["converged"], "single_fp64_s": round(t_single, 3),
           "single_fp64_krylov_its": its, "speedup_block_vs_single": round(t_single / t_blk, 3),
           "max_rel_solution_diff_vs_single_fp64": worst}
    log(f"[bench] sensitivity: {res}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=NX, help="hexes per axis")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of oracle timing")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-newton", action="store_true")
    ap.add_argument("--newton-steps", type=int, default=600)
    ap.add_argument("--no-viscous", action="store_true", help="skip the Navier-Stokes leg")
    ap.add_argument("--batched", action="store_true",
                    help="also time 16-RHS block GMRES against 16 single-RHS FP64 solves (128^3 box)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip timing the adjoint kernels on the 192x64x64 channel")
    ap.add_argument("--shape", action="store_true",
                    help="also run the adjoint shape gradient (8 bump amplitudes) on the 192x64x64 channel")
    ap.add_argument("--sensitivity", action="store_true",
                    help="also run 16 batched direct sensitivities on the 128^3 bump box")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(multi-GPU sharding is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_calibration:
        # A plain copy kernel of known size: 1 GiB read + 1 GiB written per
        # launch.  rocprofv3's FETCH_SIZE / WRITE_SIZE on it give the scale
        # between counter units and bytes.
        n = (1 << 30) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30, "launches": 3}))
        return

    from aerof_dev_amd import FlowParams, init_state
    from aerof_dev_amd import mesh as M
    from aerof_dev_amd.operators import SpaceOperator

    t0 = time.time()
    mesh = M.box_mesh(args.n, jitter=0.15, seed=1234)
    # RCM renumbering and edge colouring on the GPU (bit-identical to the host rules)
    t1 = time.time()
    mesh = M.renumber(mesh, M.rcm_gpu(mesh))
    t_rcm = time.time() - t1
    t1 = time.time()
    colors, ncolors = M.color_edges_gpu(mesh.num_nodes, M.edges(mesh))
    t_col = time.time() - t1
    params = FlowParams(mach=0.8)
    U = init_state(params, mesh.X, "sine", 0.02, seed=2024)
    log(f"[bench] mesh {mesh.num_nodes} nodes, {mesh.num_tets} tets ({time.time() - t0:.1f} s; GPU RCM {t_rcm:.2f} s, "
        f"GPU edge colouring {t_col:.2f} s, {ncolors} colours)")
    dev = torch.device("cuda:0")
    t0 = time.time()
    op = SpaceOperator(mesh, params)
    N, E, F = op.num_nodes, op.num_edges, op.num_faces
    log(f"[bench] context: E={E} F={F} ({time.time() - t0:.1f} s)")
    Ud = torch.from_numpy(U).to(dev)
    Rd = torch.empty_like(Ud)
    ierr = op.compute_residual(Ud, Rd)
    if ierr != 0:
        raise RuntimeError(f"non-physical reconstruction count {ierr}")
    for _ in range(args.warmup):
        op.compute_residual_async(Ud, Rd)
    torch.cuda.synchronize()
    op.profile(True)
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        op.compute_residual_async(Ud, Rd)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_total = ev0.elapsed_time(ev1)
    timers = op.timers()
    op.profile(False)
    ms_step = ms_total / args.steps
    log(f"[bench] {args.steps} steps, {ms_step:.4f} ms/step; kernels {timers}")

    def per_launch(name):
        tot, cnt = timers.get(name, (0.0, 0))
        return tot / cnt if cnt else 0.0
    grad_ms = per_launch("nodalGrad")
    edge_ms = per_launch("fvTerm")
    node_ms = per_launch("fvTermNodes")
    # Algorithmic bytes of one residual evaluation: gradient pass 232 B/node
    # + 8 B/edge, flux pass 232 B/node + 32 B/edge + 40 B/face.
    grad_bytes = 232 * N + 8 * E
    flux_bytes = 232 * N + 32 * E + 40 * F
    whole_bytes = grad_bytes + flux_bytes
    kern_ms = grad_ms + edge_ms + node_ms
    achieved = whole_bytes / (kern_ms * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=load_traffic("residual_box128"),
                    kernel="residual = k_gradients + k_flux_edges + k_flux_nodes",
                    algorithmic_bytes=whole_bytes, ms_per_launch=round(kern_ms, 5))
    out = {
        "metric": "Medge-flux evals/s (FP64 residual, 128^3 Kuhn box)",
        "value": round(E / (ms_step * 1e-3) / 1e6, 1),
        "unit": "Medges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Kuhn-split jittered box, sine density perturbation)",
        "config": {"workload": "FP64 residual, 128^3-hex Kuhn box, RCM order, M=0.8",
                   "nodes": N, "edges": E, "faces": F, "parallelism": "single GPU",
                   "gpu_rcm_s": round(t_rcm, 3), "gpu_edge_colouring_s": round(t_col, 3), "edge_colours": ncolors},
        "roofline": roofline,
        "kernels": {"k_gradients": {"ms": round(grad_ms, 5),
                                    "algorithmic_GBs": round(grad_bytes / (grad_ms * 1e-3) / 1e9, 1) if grad_ms else None},
                    "k_flux_edges+k_flux_nodes": {
                        "ms": round(edge_ms + node_ms, 5),
                        "algorithmic_GBs": round(flux_bytes / ((edge_ms + node_ms) * 1e-3) / 1e9, 1) if edge_ms + node_ms else None},
                    "k_flux_edges": {"ms": round(edge_ms, 5)},
                    "k_flux_nodes": {"ms": round(node_ms, 5)}},
    }
    roofline["copy_peak_GBs"] = copy_peak(dev)
    roofline["frac_of_copy_peak"] = round(achieved / roofline["copy_peak_GBs"], 4)
    # FP64 throughput: the counted FP64 lane-operations of one evaluation
    # (profiles/) over this run's kernel time, against the FP64 vector peak
    ops = load_traffic("residual_box128", "fp64_lane_ops_per_launch")
    if ops:
        roofline["fp64_lane_ops"] = ops
        roofline["fp64_tflops"] = round(ops / (kern_ms * 1e-3) / 1e12, 2)
        roofline["fp64_peak_tflops"] = FP64_PEAK_TFLOPS
        roofline["fp64_frac"] = round(roofline["fp64_tflops"] / FP64_PEAK_TFLOPS, 4)
    if (not args.no_newton or args.batched or not args.no_adjoint or args.shape or args.sensitivity
            or not args.no_viscous):
        op.close()
    if not args.no_newton:
        out["newton"] = newton_bench(dev, args.newton_steps)
    if args.batched:
        out["batched"] = batched_bench(dev, args.n)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench(dev)
    if not args.no_viscous:
        out["viscous"] = viscous_bench(dev, args.n)
    if args.shape:
        out["shape"] = shape_bench(dev)
    if args.sensitivity:
        out["sensitivity"] = sensitivity_bench(dev, args.n)
    if not args.no_cpu:
        log("[bench] timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(mesh, params, U, args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
