#!/usr/bin/env python3
"""Benchmark of the even-odd Wilson-clover hot path on one MI355X; prints one JSON line.

Workload (even-odd preconditioned clover, csw = 1.0, 32^3x64): Haar-random
SU(3) start + 6 isotropic stout steps (rho = 0.1), antiperiodic in time, kappa = 0.1350, Z2 source on
the odd checkerboard, mixed-precision CG (FP16 storage / FP32 arithmetic inner, FP64 reliable updates,
delta = 0.1) to an FP64 true residual of 1e-10.

value: achieved bandwidth of the hopping term D_eo (one application = one step) at FP16 storage with
12-real links, in algorithmic bytes (each input spinor and link read once, output written once):
48*2 + 8*12*2 + 8 (per-site FP32 norms of input and output) = 296 bytes per output site.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def hop_bytes(prec, recon):
    s = {0: 8, 1: 4, 2: 2}[prec]
    b = 48 * s + 8 * recon * s
    if prec == 2:
        b += 8  # FP32 norm of the input and the output spinor
    return b


def eo_clover_bytes(prec, recon):
    """Fused EO clover operator per odd site: 120 s + 16 L s_u + 144 s_c (two kernels)."""
    s = {0: 8, 1: 4, 2: 2}[prec]
    b = 120 * s + 16 * recon * s + 144 * s
    if prec == 2:
        b += 5 * 4 + 2 * 4  # spinor norms (5 spinor streams) and clover norms (2 clover streams)
    return b


def measured_traffic(key, lattice):
    """HBM bytes per launch of a kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (profiles/hop_<lattice>_traffic.md), or None when no summary exists for this lattice."""
    path = os.path.join(ROOT, "profiles", f"hop_{'x'.join(map(str, lattice))}_traffic.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith(f"traffic_bytes_per_launch {key}:"):
            return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def _sample(op, what, target):
    t1 = op.time(what, 1)  # one call to size the sample
    n = max(1, int(target / max(t1, 1e-6)))
    return n, op.time(what, n)


def cpu_hop_baseline(L, U, seconds_target=8.0):
    """The oracle's FP64 hopping term (OpenMP over all host cores) on a bounded sample of the headline
    workload: the lattice and field of the GPU figure, FP64/18 algorithmic bytes (1536 per site)."""
    import numpy as np
    import oracle
    ncores = len(os.sched_getaffinity(0))
    oracle.set_threads(ncores)
    tag = "x".join(map(str, L))
    op = oracle.Op(oracle.EO_WILSON, L, U, 0.0)  # the hop does not depend on the mass
    n, t = _sample(op, 0, seconds_target)
    del op
    Vh = int(np.prod(L)) // 2
    gbs = hop_bytes(0, 18) * Vh / t / 1e9
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": ncores, "kind": "port",
            "sample": f"{n} FP64 hopping-term applications at {tag} (oracle, {t * 1e3:.1f} ms each)",
            "gflops": round(1320.0 * Vh / t / 1e9, 3)}


def cpu_solve_baseline(L, U, mass, csw, gpu_cg_iterations=None, seconds_target=8.0):
    """The oracle (FP64, OpenMP over the host cores) on the solve lattice: the clover build, the EO clover
    operator and one CG iteration on it, and the 8^3x16 unpreconditioned Wilson CGNE solved to 1e-10 in
    full."""
    import numpy as np
    import oracle
    ncores = len(os.sched_getaffinity(0))
    oracle.set_threads(ncores)
    out = {"lattice": list(L), "cores": ncores}
    t0 = time.time()
    op = oracle.Op(oracle.EO_CLOVER, L, U, mass, csw, csw)
    out["clover_build_seconds"] = round(time.time() - t0, 2)
    n, t = _sample(op, 1, seconds_target / 4)
    out["eo_clover_op_ms"] = round(t * 1e3, 2)
    n, t = _sample(op, 2, seconds_target / 4)
    out["cg_iteration_ms"] = round(t * 1e3, 2)
    if gpu_cg_iterations:
        out["plain_cg_solve_seconds_extrapolated"] = round(t * gpu_cg_iterations, 2)
        out["extrapolation"] = (f"FP64 CG iteration time x the {gpu_cg_iterations} iterations of the GPU's plain "
                                f"mixed CG on the same operator (not a measured CPU solve)")
    del op
    L1 = (8, 8, 8, 16)
    from chroma_amd import lib as cmx
    U1 = cmx.gauge_apply_bc(L1, cmx.gauge_generate(L1, kind=1, seed=1, eps=0.3), (1, 1, 1, -1))
    op = oracle.Op(oracle.UNPREC_WILSON, L1, U1, cmx.kappa_to_mass(0.115))
    chi = np.zeros((int(np.prod(L1)), 4, 3), np.complex128)
    chi[0, 0, 0] = 1.0
    t0 = time.time()
    _, nc, resid, _ = op.solve(chi, solver=oracle.CG, rsd=1e-10, maxit=5000)
    out["small_volume_cg"] = {"lattice": list(L1), "seconds": round(time.time() - t0, 4), "iterations": nc,
                              "resid": resid}
    return out


def small_volume_solves(cmx):
    """The 8^3x16 unpreconditioned Wilson CGNE (kappa 0.115, FP64, point source at the origin, FP64 true
    residual 1e-10): launch-bound, so the iterations are replayed from captured HIP graphs and timed
    against eager launches."""
    import numpy as np
    L = (8, 8, 8, 16)
    V = int(np.prod(L))
    U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=1, seed=1, eps=0.3), (1, 1, 1, -1))
    op = cmx.Operator(cmx.Gauge(L, U), cmx.UNPREC_WILSON, cmx.kappa_to_mass(0.115))
    chi = np.zeros((V, 4, 3), np.complex128)
    chi[0, 0, 0] = 1.0
    out = {"lattice": list(L), "operator": "unpreconditioned Wilson, kappa 0.115, FP64 CGNE"}
    for name, graph in [("eager", False), ("hip_graph", True)]:
        for _ in range(2):
            psi, res = op.solve(chi, solver=cmx.CG, rsd=1e-10, maxit=5000, true_residual=True, graph=graph)
        out[name] = {"seconds": round(res["solve_seconds"], 5), "iterations": res["n_count"],
                     "true_rel_resid": res["rel_resid"]}
        print(f"# small volume CG {name}: {res['n_count']} it, {res['solve_seconds'] * 1e3:.2f} ms",
              file=sys.stderr, flush=True)
    return out


def batched_uneven_solves(cmx):
    """64 independent reliable-CG solves (FP16 storage, FP64 reliable updates, FP64 true residual 1e-10)
    on uneven volumes cycled 8^4 / 12^3x24 / 16^3x32, each with its own field (exp(0.3 i H) recipe,
    seed 1000+i) and kappa drawn uniformly in [0.120, 0.134] (seed 999): one solve after another vs
    all 64 advanced together through batched launches (one launch per kernel for every problem)."""
    import numpy as np
    rng = np.random.default_rng(999)
    kap = rng.uniform(0.120, 0.134, 64)
    lats = [(8, 8, 8, 8), (12, 12, 12, 24), (16, 16, 16, 32)]
    ops, chis = [], []
    for i in range(64):
        L = lats[i % 3]
        U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=1, seed=1000 + i, eps=0.3), (1, 1, 1, -1))
        ops.append(cmx.Operator(cmx.Gauge(L, U), cmx.EO_WILSON, cmx.kappa_to_mass(kap[i]), sloppy_prec=cmx.HALF))
        chi = np.zeros((int(np.prod(L)) // 2, 4, 3), complex)
        chi[0, 0, 0] = 1.0  # point source on the first odd site
        chis.append(chi)
    kw = dict(solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=5000, sloppy_prec=cmx.HALF, true_residual=True)
    seq = None
    for _ in range(2):  # the first pass allocates the solvers' vectors; the second is reported
        t0 = time.time()
        res = [op.solve(chi, **kw)[1] for op, chi in zip(ops, chis)]
        seq = dict(seconds=round(sum(r["solve_seconds"] for r in res), 4), wall_seconds=round(time.time() - t0, 4),
                   iterations=[min(r["n_count"] for r in res), max(r["n_count"] for r in res)],
                   worst_true_rel_resid=max(r["rel_resid"] for r in res))
    bat = None
    for _ in range(2):
        t0 = time.time()
        _, res, summ = cmx.solve_batch(ops, chis, tune=True, **kw)
        bat = dict(seconds=round(summ["solve_seconds"], 4), wall_seconds=round(time.time() - t0, 4),
                   batched_iterations=summ["n_iter"], kernels_per_iteration=summ["n_stages"],
                   launch_shape={"block": summ["hop_block"], "fast": summ["hop_fast"]},
                   worst_true_rel_resid=max(r["rel_resid"] for r in res))
    out = {"problems": 64, "volumes": [list(v) for v in lats], "operator": "EO Wilson, FP16 storage",
           "solver": "reliable CG, FP64 true residual 1e-10", "sequential": seq, "batched": bat,
           "speedup": round(seq["seconds"] / bat["seconds"], 3)}
    print(f"# batched uneven: sequential {seq['seconds']:.4f} s, batched {bat['seconds']:.4f} s", file=sys.stderr,
          flush=True)
    return out


def batch_in_groups(cmx, op, chis, group, **kw):
    """The 12 sources advanced together in groups of `group` (one batched solve per group, the memory of
    `group` solver workspaces at a time): per-source results, summed solve seconds, the last summary."""
    res, secs, summ = [], 0.0, None
    for g0 in range(0, len(chis), group):
        part = chis[g0:g0 + group]
        _, r, summ = cmx.solve_batch([op] * len(part), part, tune=True, **kw)
        res += r
        secs += summ["solve_seconds"]
    return res, secs, summ


def batched_sources(cmx, L, group=12):
    """The 12 spin-colour point sources of a propagator (quarkProp4_a) on one EO Wilson operator
    (exp(0.3 i H) field, seed 1, kappa 0.115), reliable CG with FP16 storage to an FP64 true residual of
    1e-10: 12 solves one after another vs the 12 advanced together (one launch per kernel for all 12)."""
    import numpy as np
    U = cmx.gauge_apply_bc(L, cmx.gauge_generate(L, kind=1, seed=1, eps=0.3), (1, 1, 1, -1))
    op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_WILSON, cmx.kappa_to_mass(0.115), sloppy_prec=cmx.HALF)
    del U
    Vh = int(np.prod(L)) // 2
    chis = []
    for s in range(4):
        for c in range(3):
            chi = np.zeros((Vh, 4, 3), complex)
            chi[0, s, c] = 1.0  # point source on the first odd site
            chis.append(chi)
    kw = dict(solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=5000, sloppy_prec=cmx.HALF, true_residual=True)
    op.solve(chis[0], **kw)  # allocations
    res = [op.solve(chi, **kw)[1] for chi in chis]
    seq = dict(seconds=round(sum(r["solve_seconds"] for r in res), 4),
               iterations=[min(r["n_count"] for r in res), max(r["n_count"] for r in res)],
               worst_true_rel_resid=max(r["rel_resid"] for r in res))
    res, secs, summ = batch_in_groups(cmx, op, chis, group, **kw)
    bat = dict(seconds=round(secs, 4), batched_iterations=summ["n_iter"], group=group,
               launch_shape={"block": summ["hop_block"], "fast": summ["hop_fast"]},
               worst_true_rel_resid=max(r["rel_resid"] for r in res))
    out = {"lattice": list(L), "sources": 12, "operator": "EO Wilson, kappa 0.115, FP16 storage",
           "sequential": seq, "batched": bat, "speedup": round(seq["seconds"] / bat["seconds"], 3)}
    print(f"# batched sources {L}: sequential {seq['seconds']:.4f} s, batched {bat['seconds']:.4f} s",
          file=sys.stderr, flush=True)
    return out


def stout_field(cmx, L, seed):
    """Haar SU(3) start (host generator), then 6 isotropic stout steps with rho = 0.1 on the device."""
    return cmx.gauge_stout_smear(L, cmx.gauge_generate(L, kind=2, seed=seed), 6, 0.1)


def train_net(cmx, mass, csw, a, depth=1, pairs_per_source=4, lr=1e-3):
    """The learned preconditioner's network (`depth` blocks, the first starting as the two-term Neumann
    series, the others as the identity), trained on separate fields of the benchmark recipe."""
    net = cmx.Net()
    net.set_depth(depth)
    net.init_neumann(4.0 + mass)
    TL = tuple(int(v) for v in a.train_lattice.split(","))
    train = []
    for k in range(a.train_fields):
        U = cmx.gauge_apply_bc(TL, stout_field(cmx, TL, 10 + k), (1, 1, 1, -1))
        top = cmx.Operator(cmx.Gauge(TL, U), cmx.EO_CLOVER, mass, csw, sloppy_prec=cmx.HALF)
        cmx.net_attach(top, net)
        train.append(cmx.net_train(top, n_sources=8, snapshot_every=1, pairs_per_source=pairs_per_source,
                                   normalize=False, epochs=4, lr=lr, seed=100 + k, prec=cmx.HALF, max_pairs=100000))
    return net, train


FGMRES_KW = dict(rsd=1e-10, maxit=20000, sloppy_prec=2, basis_prec=2, n_krylov=16, cycle_reduction=3e-3)
# two-level multigrid: 4^4 blocks, 6 null vectors (12 coarse degrees of freedom per block), 8 MR post-smoothing steps
# at FP16 storage, coarse GCR to 0.1 or 8 steps; the outer FGMRES-DR restarts every 8 steps
MG_KW = dict(nvec=6, nu_pre=0, nu_post=8, setup_maxit=100, coarse_tol=0.1, coarse_maxit=8, seed=3)
FGMRES_MG_KW = dict(rsd=1e-10, maxit=20000, sloppy_prec=2, n_krylov=8)


def sigma_min(cmx, op, Vh, iters=6):
    """Smallest singular value of the even-odd operator M by inverse iteration on (M^dag M)^-1 =
    M^-1 g5 M^-1 g5 (M^dag = g5 M g5), each inverse an FGMRES-DR solve with an FP32 operator and basis to
    1e-6; the Rayleigh quotient 1 / <x, (M^dag M)^-1 x> of the last unit vector x estimates sigma_min^2
    from above.  None when a solve does not converge."""
    import numpy as np
    g5 = np.array([1.0, 1.0, -1.0, -1.0])[None, :, None]
    x = cmx.z2_source((Vh, 4, 3), 77)
    x /= np.linalg.norm(x)
    kw = dict(solver=cmx.FGMRESDR, rsd=1e-6, maxit=5000, sloppy_prec=cmx.SINGLE, basis_prec=cmx.SINGLE, n_krylov=16)
    lam = None
    try:
        for _ in range(iters):
            w, _ = op.solve(g5 * x, **kw)
            y, _ = op.solve(g5 * w, **kw)
            lam = 1.0 / float(np.vdot(x, y).real)
            x = y / np.linalg.norm(y)
    except cmx.CmxError as e:
        print(f"# sigma_min: {e}", file=sys.stderr, flush=True)
        return None
    return float(np.sqrt(lam))


def critical_scan(cmx, gauge, L, kappas=(0.1355, 0.150, 0.155, 0.158, 0.160)):
    """Where the synthetic ensemble (Haar + 6 stout steps) is critical: at each kappa, the smallest
    singular value of the even-odd operator, the plain mixed CG iterations on a Z2 source (to kappa
    0.158; it needs 18 000 at 0.160 and does not converge at 0.162) and the FGMRES-DR(16) iterations; kappa_c
    is where sigma_min, linear in 1/kappa near the critical point, extrapolates to zero (last two points)."""
    import numpy as np
    Vh = int(np.prod(L)) // 2
    chi = cmx.z2_source((Vh, 4, 3), 2)
    rows = []
    for k in kappas:
        op = cmx.Operator(gauge, cmx.EO_CLOVER, cmx.kappa_to_mass(k), 1.0, sloppy_prec=cmx.HALF, autotune=True)
        sm = sigma_min(cmx, op, Vh)
        row = {"kappa": k, "sigma_min": None if sm is None else round(sm, 6)}
        _, r = op.solve(chi, solver=cmx.FGMRESDR, **FGMRES_KW)
        row["fgmres_iterations"] = r["n_count"]
        if k <= 0.158:
            _, r = op.solve(chi, solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=50000, delta=0.1, sloppy_prec=cmx.HALF,
                            true_residual=True)
            row["plain_cg_iterations"] = r["n_count"]
        rows.append(row)
        print(f"# critical scan {row}", file=sys.stderr, flush=True)
        del op
    pts = [(1.0 / r["kappa"], r["sigma_min"]) for r in rows if r["sigma_min"] is not None][-2:]
    kc = None
    if len(pts) == 2:
        (k1, s1), (k2, s2) = pts
        kc = round(1.0 / (k2 - s2 * (k2 - k1) / (s2 - s1)), 5)
    return {"lattice": list(L), "rows": rows, "kappa_c": kc,
            "method": "sigma_min(M) by inverse iteration, linear in 1/kappa through the two points nearest to it"}


def near_critical_solve(cmx, gauge, L, kappa, net):
    """One Z2 source at a near-critical kappa, all to the FP64 true residual 1e-10: plain mixed CG (FP16
    inner, FP64 reliable updates), FGMRES-DR(16) without a preconditioner, with the untrained network
    (its Neumann start), with the trained network, and FGMRES-DR(8) with the two-level multigrid V-cycle."""
    import numpy as np
    mass = cmx.kappa_to_mass(kappa)
    op = cmx.Operator(gauge, cmx.EO_CLOVER, mass, 1.0, sloppy_prec=cmx.HALF, autotune=True)
    Vh = int(np.prod(L)) // 2
    chi = cmx.z2_source((Vh, 4, 3), 2)
    out = {"lattice": list(L), "kappa": kappa}
    try:  # allocates the solver's vectors; 20 iterations do not converge here
        op.solve(chi, solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=20, delta=0.1, sloppy_prec=cmx.HALF)
    except cmx.CmxError:
        pass
    _, r = op.solve(chi, solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=50000, delta=0.1, sloppy_prec=cmx.HALF,
                    true_residual=True)
    out["plain_mixed_cg"] = {"seconds": round(r["solve_seconds"], 4), "iterations": r["n_count"],
                             "true_rel_resid": r["rel_resid"]}
    start = cmx.Net()
    start.set_depth(net.depth())
    start.init_neumann(4.0 + mass)
    for name, pc, nn in (("fgmres_none", cmx.PRECOND_NONE, None), ("fgmres_neumann_start", cmx.PRECOND_LEARNED, start),
                         ("fgmres_trained", cmx.PRECOND_LEARNED, net)):
        if nn is not None:
            cmx.net_attach(op, nn)
        op.solve(chi, solver=cmx.FGMRESDR, precond=pc, **FGMRES_KW)
        _, r = op.solve(chi, solver=cmx.FGMRESDR, precond=pc, **FGMRES_KW)
        out[name] = {"seconds": round(r["solve_seconds"], 4), "iterations": r["n_count"],
                     "true_rel_resid": r["rel_resid"]}
    # two-level multigrid as the right preconditioner (set-up reported apart: it serves every later solve)
    try:
        info = op.mg_setup(blocking=(4, 4, 4, 4), **MG_KW)
        op.solve(chi, solver=cmx.FGMRESDR, precond=cmx.PRECOND_MG, **FGMRES_MG_KW)
        _, r = op.solve(chi, solver=cmx.FGMRESDR, precond=cmx.PRECOND_MG, **FGMRES_MG_KW)
        out["fgmres_multigrid"] = {"seconds": round(r["solve_seconds"], 4), "iterations": r["n_count"],
                                   "true_rel_resid": r["rel_resid"], "setup_seconds": round(info["setup_seconds"], 3),
                                   "params": dict(MG_KW, blocking=[4, 4, 4, 4])}
    except cmx.CmxError as e:
        print(f"# multigrid: {e}", file=sys.stderr, flush=True)
    p = out["plain_mixed_cg"]["seconds"]
    out["speedup_trained_vs_plain"] = round(p / out["fgmres_trained"]["seconds"], 3)
    if "fgmres_multigrid" in out:
        out["speedup_multigrid_vs_plain"] = round(p / out["fgmres_multigrid"]["seconds"], 3)
        out["speedup_multigrid_vs_trained"] = round(out["fgmres_trained"]["seconds"] /
                                                    out["fgmres_multigrid"]["seconds"], 3)
    out["speedup_trained_vs_fgmres_none"] = round(out["fgmres_none"]["seconds"] / out["fgmres_trained"]["seconds"], 3)
    out["speedup_trained_vs_neumann_start"] = round(out["fgmres_neumann_start"]["seconds"] /
                                                    out["fgmres_trained"]["seconds"], 3)
    print(f"# near-critical {L} kappa {kappa}: {out}", file=sys.stderr, flush=True)
    return out


def learned_vs_plain_sources(cmx, L, kappa, csw, a, net=None, gauge=None, variants=False, multigrid=False):
    """The 12 spin-colour point sources of a propagator on the clover operator at kappa (Haar + 6 stout
    steps, seed 1), all to the FP64 true residual 1e-10: plain mixed-precision CG one after another and
    batched, against the learned FGMRES-DR one after another and batched (with `variants`: also FGMRES-DR
    without a preconditioner and with the network's untrained Neumann start; with `multigrid`: FGMRES-DR(8)
    with the two-level multigrid, one after another and batched)."""
    import numpy as np
    mass = cmx.kappa_to_mass(kappa)
    t0 = time.time()
    if gauge is None:
        U = cmx.gauge_apply_bc(L, stout_field(cmx, L, 1), (1, 1, 1, -1))
        gauge = cmx.Gauge(L, U)
        del U
    op = cmx.Operator(gauge, cmx.EO_CLOVER, mass, csw, sloppy_prec=cmx.HALF, autotune=True)
    Vh = int(np.prod(L)) // 2
    chis = []
    for s in range(4):
        for c in range(3):
            chi = np.zeros((Vh, 4, 3), complex)
            chi[0, s, c] = 1.0
            chis.append(chi)
    kwp = dict(solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=50000, delta=0.1, sloppy_prec=cmx.HALF, true_residual=True)
    op.solve(chis[0], **kwp)
    plain = []
    for c in chis:  # one line per solve: a long leg keeps reporting progress
        plain.append(op.solve(c, **kwp)[1])
        print(f"#   plain source {len(plain)}: {plain[-1]['n_count']} it {plain[-1]['solve_seconds']:.3f} s",
              file=sys.stderr, flush=True)
    rb, bsecs, _ = batch_in_groups(cmx, op, chis, a.batch_group, **kwp)
    print(f"#   plain batched: {bsecs:.3f} s", file=sys.stderr, flush=True)
    if net is None:
        net, _ = train_net(cmx, mass, csw, a)
    kw = dict(FGMRES_KW, solver=cmx.FGMRESDR)
    runs = [("learned_fgmres_sequential", cmx.PRECOND_LEARNED, net)]
    if variants:
        start = cmx.Net()
        start.set_depth(net.depth())
        start.init_neumann(4.0 + mass)
        runs = [("fgmres_none_sequential", cmx.PRECOND_NONE, None),
                ("neumann_start_fgmres_sequential", cmx.PRECOND_LEARNED, start)] + runs
    ps = sum(r["solve_seconds"] for r in plain)
    out = {"lattice": list(L), "kappa": kappa, "sources": 12,
           "plain_mixed_cg_sequential": {"seconds": round(ps, 4), "iterations": [r["n_count"] for r in plain],
                                         "worst_true_rel_resid": max(r["rel_resid"] for r in plain)},
           "plain_mixed_cg_batched": {"seconds": round(bsecs, 4), "group": a.batch_group,
                                      "worst_true_rel_resid": max(r["rel_resid"] for r in rb)}}
    for name, pc, nn in runs:
        if nn is not None:
            cmx.net_attach(op, nn)
        op.solve(chis[0], precond=pc, **kw)
        res = [op.solve(c, precond=pc, **kw)[1] for c in chis]
        out[name] = {"seconds": round(sum(r["solve_seconds"] for r in res), 4),
                     "iterations": [r["n_count"] for r in res],
                     "worst_true_rel_resid": max(r["rel_resid"] for r in res)}
    # the same learned solves batched: one Arnoldi step of every source per host read-back
    cmx.net_attach(op, net)
    psb, rbl, summ = cmx.solve_batch([op] * 12, chis, precond=cmx.PRECOND_LEARNED, **kw)
    out["learned_fgmres_batched"] = {"seconds": round(summ["solve_seconds"], 4),
                                     "iterations": [r["n_count"] for r in rbl],
                                     "worst_true_rel_resid": max(r["rel_resid"] for r in rbl),
                                     "arnoldi_ticks": summ["n_iter"]}
    del psb
    if multigrid:
        # the two-level multigrid on the same sources (one set-up for all twelve), sequential and batched
        info = op.mg_setup(blocking=(4, 4, 4, 4), **MG_KW)
        kwm = dict(FGMRES_MG_KW, solver=cmx.FGMRESDR)
        op.solve(chis[0], precond=cmx.PRECOND_MG, **kwm)
        res = [op.solve(c, precond=cmx.PRECOND_MG, **kwm)[1] for c in chis]
        out["multigrid_fgmres_sequential"] = {"seconds": round(sum(r["solve_seconds"] for r in res), 4),
                                              "iterations": [r["n_count"] for r in res],
                                              "worst_true_rel_resid": max(r["rel_resid"] for r in res),
                                              "setup_seconds": round(info["setup_seconds"], 3)}
        psb, rbm, summ = cmx.solve_batch([op] * 12, chis, precond=cmx.PRECOND_MG, **kwm)
        out["multigrid_fgmres_batched"] = {"seconds": round(summ["solve_seconds"], 4),
                                           "iterations": [r["n_count"] for r in rbm],
                                           "worst_true_rel_resid": max(r["rel_resid"] for r in rbm),
                                           "arnoldi_ticks": summ["n_iter"]}
        del psb
        out["speedup_batched_multigrid_vs_plain_batched"] = round(
            bsecs / out["multigrid_fgmres_batched"]["seconds"], 3)
        print(f"#   multigrid: set-up {info['setup_seconds']:.2f} s, sequential "
              f"{out['multigrid_fgmres_sequential']['seconds']:.3f} s, batched "
              f"{out['multigrid_fgmres_batched']['seconds']:.3f} s", file=sys.stderr, flush=True)
    ls = out["learned_fgmres_sequential"]["seconds"]
    lb = out["learned_fgmres_batched"]["seconds"]
    out["network_blocks"] = net.depth()
    out["speedup_vs_plain"] = round(ps / ls, 3)
    out["speedup_vs_plain_batched"] = round(bsecs / ls, 3)
    out["speedup_batched_learned_vs_plain_batched"] = round(bsecs / lb, 3)
    out["speedup_batched_learned_vs_sequential_learned"] = round(ls / lb, 3)
    out["wall_seconds"] = round(time.time() - t0, 1)
    print(f"# 12 sources {L} kappa {kappa}: {out}", file=sys.stderr, flush=True)
    return out


def learned_solve(cmx, gauge, L, kappa, csw, a):
    """Learned preconditioner at kappa 0.1355 (near critical) on the benchmark field: trained on separate
    fields of the same recipe (Haar + 6 stout steps, seeds 10, 11, ...) on a smaller lattice, then used
    as the right preconditioner of FGMRES-DR on the benchmark field (not seen in training), against plain
    mixed-precision reliable CG on the same operator, both to the FP64 true residual 1e-10.  FGMRES-DR runs
    its operator and network at FP16 storage, keeps the Krylov basis in FP16 and restarts a cycle from the
    FP64 residual once its estimate has fallen by 3e-3 (the reach of FP16 arithmetic)."""
    import numpy as np
    mass = cmx.kappa_to_mass(kappa)
    op = cmx.Operator(gauge, cmx.EO_CLOVER, mass, csw, sloppy_prec=cmx.HALF, sloppy_recon=cmx.RECONS_12,
                      autotune=True)
    Vh = int(np.prod(L)) // 2
    chi = cmx.z2_source((Vh, 4, 3), 2)
    for _ in range(2):  # the first solve allocates the solver's vectors; the second is reported
        _, rc = op.solve(chi, solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=50000, delta=0.1, sloppy_prec=cmx.HALF,
                         true_residual=True)
    plain = {"seconds": round(rc["solve_seconds"], 4), "iterations": rc["n_count"],
             "true_rel_resid": rc["rel_resid"], "reliable_updates": rc["n_reliable"]}
    print(f"# solve plain_mixed_cg_fp16 kappa {kappa}: {rc['n_count']} it, {rc['solve_seconds']:.4f} s",
          file=sys.stderr, flush=True)
    TL = tuple(int(v) for v in a.train_lattice.split(","))
    t0 = time.time()
    net, train = train_net(cmx, mass, csw, a)
    for k, t in enumerate(train):
        print(f"# train field {k}: {t}", file=sys.stderr, flush=True)
    t_train = time.time() - t0
    cmx.net_attach(op, net)
    kw = dict(solver=cmx.FGMRESDR, rsd=1e-10, maxit=5000, sloppy_prec=cmx.HALF, basis_prec=cmx.HALF, n_krylov=16,
              cycle_reduction=3e-3)
    out = {"kappa": kappa, "train_lattice": list(TL), "train_fields": a.train_fields,
           "train_seconds": round(t_train, 2),
           "train": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in t.items()} for t in train],
           "network": "2 layers, 16 channels, one-hop parallel transports of half spinors, BF16 MFMA mixing",
           "solver": "FGMRES-DR(16), FP16 operator + network, FP16 basis, cycle restart at 3e-3, FP64 residual",
           "plain_mixed_cg_fp16": plain}
    for name, pc in [("fgmres_none", cmx.PRECOND_NONE), ("fgmres_learned", cmx.PRECOND_LEARNED)]:
        op.solve(chi, precond=pc, **kw)  # warm-up (buffers, weights upload)
        psi, res = op.solve(chi, precond=pc, **kw)
        out[name] = {"seconds": round(res["solve_seconds"], 4), "iterations": res["n_count"],
                     "cycles": res["n_reliable"], "true_rel_resid": res["rel_resid"]}
        print(f"# solve {name}: {res['n_count']} it, {res['solve_seconds']:.4f} s, rel resid "
              f"{res['rel_resid']:.2e}", file=sys.stderr, flush=True)
    out["seconds"] = out["fgmres_learned"]["seconds"]
    out["speedup_vs_plain"] = round(plain["seconds"] / out["seconds"], 3)
    out["speedup_vs_fgmres_none"] = round(out["fgmres_none"]["seconds"] / out["seconds"], 3)
    out["held_out"] = held_out_fields(cmx, L, mass, csw, net, kw)
    return out, net


def learned_config(cmx, a):
    """The learned-preconditioner configuration as stated: 32^3x64, clover csw 1.0, kappa 0.1355; the network
    (one block: two channel layers, 16 channels) trained on about 20 000 (residual, error) pairs from plain
    CG solves on 8 seeded stout-smeared fields of that lattice (seeds 10-17), streamed in rounds of 128 pairs
    (one field per round, the fields in turn, 2 epochs per round), then tested inside FGMRES-DR(16) on 4
    held-out fields (seeds 20-23) against plain mixed CG, beside its untrained Neumann start."""
    import numpy as np
    L, kappa, csw = (32, 32, 32, 64), 0.1355, 1.0
    mass = cmx.kappa_to_mass(kappa)
    t0 = time.time()
    net = cmx.Net()
    net.set_depth(a.c4_depth)
    net.init_neumann(4.0 + mass)
    start = cmx.Net()
    start.set_depth(a.c4_depth)
    start.init_neumann(4.0 + mass)
    ops = []
    for k in range(8):
        U = cmx.gauge_apply_bc(L, stout_field(cmx, L, 10 + k), (1, 1, 1, -1))
        ops.append(cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, csw, sloppy_prec=cmx.HALF))
        del U
    pairs, rounds, loss = 0, 0, []
    while pairs < a.c4_pairs:
        op = ops[rounds % 8]
        cmx.net_attach(op, net)
        t = cmx.net_train(op, n_sources=8, snapshot_every=1, pairs_per_source=16, normalize=False, epochs=2,
                          lr=a.c4_lr, seed=1000 + rounds, prec=cmx.HALF, max_pairs=128)
        pairs += t["n_pairs"]
        rounds += 1
        loss.append((round(t["loss_first"], 5), round(t["loss_last"], 5)))
        if rounds % 10 == 0:
            print(f"#   c4 training: round {rounds}, {pairs} pairs, loss {t['loss_first']:.5f} -> {t['loss_last']:.5f}",
                  file=sys.stderr, flush=True)
    del ops
    t_train = time.time() - t0
    kw = dict(solver=cmx.FGMRESDR, rsd=1e-10, maxit=5000, sloppy_prec=cmx.HALF, basis_prec=cmx.HALF, n_krylov=16,
              cycle_reduction=3e-3)
    trained = held_out_fields(cmx, L, mass, csw, net, kw)
    untrained = held_out_fields(cmx, L, mass, csw, start, kw)
    out = {"lattice": list(L), "kappa": kappa, "network_blocks": a.c4_depth, "train_fields": 8,
           "train_pairs": pairs, "train_rounds": rounds, "train_seconds": round(t_train, 1),
           "loss_first_round": loss[0], "loss_last_round": loss[-1],
           "held_out_trained": trained, "held_out_neumann_start": untrained,
           "speedup_trained_vs_plain": round(float(np.mean([r["speedup"] for r in trained])), 3),
           "speedup_trained_vs_neumann_start": round(float(np.mean(
               [u["learned_seconds"] / t["learned_seconds"] for t, u in zip(trained, untrained)])), 3)}
    print(f"# learned configuration: {out}", file=sys.stderr, flush=True)
    return out


def held_out_fields(cmx, L, mass, csw, net, kw, seeds=(20, 21, 22, 23)):
    """The trained network on four more fields of the recipe that training never saw (Haar + 6 stout
    steps, seeds 20-23, the benchmark lattice and kappa): plain mixed CG against learned FGMRES-DR on a Z2
    source, both to the FP64 true residual 1e-10."""
    import numpy as np
    Vh = int(np.prod(L)) // 2
    rows = []
    for seed in seeds:
        U = cmx.gauge_apply_bc(L, stout_field(cmx, L, seed), (1, 1, 1, -1))
        op = cmx.Operator(cmx.Gauge(L, U), cmx.EO_CLOVER, mass, csw, sloppy_prec=cmx.HALF,
                          sloppy_recon=cmx.RECONS_12, autotune=True)
        del U
        cmx.net_attach(op, net)
        chi = cmx.z2_source((Vh, 4, 3), seed)
        for _ in range(2):
            _, rp = op.solve(chi, solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=50000, delta=0.1, sloppy_prec=cmx.HALF,
                             true_residual=True)
        for _ in range(2):
            _, rl = op.solve(chi, precond=cmx.PRECOND_LEARNED, **kw)
        rows.append({"seed": seed, "plain_iterations": rp["n_count"], "plain_seconds": round(rp["solve_seconds"], 4),
                     "learned_iterations": rl["n_count"], "learned_seconds": round(rl["solve_seconds"], 4),
                     "learned_true_rel_resid": rl["rel_resid"],
                     "speedup": round(rp["solve_seconds"] / rl["solve_seconds"], 3)})
        print(f"# held-out field {seed}: plain {rp['n_count']} it {rp['solve_seconds']:.4f} s, learned "
              f"{rl['n_count']} it {rl['solve_seconds']:.4f} s", file=sys.stderr, flush=True)
    return rows


def hop_kernels(cmx, op, Vh, steps, multi=True):
    """FP32/FP64 hopping terms, the fused EO clover operator at every precision, the streaming copy and
    (multi) the 12-source FP16 hop on one operator, each timed over `steps` launches with HIP events."""
    extra = {}
    r64 = op.link_recon(cmx.DOUBLE)  # 12 when the links are SU(3) to rounding
    rows = [
        ("hop_fp32_recon12", 0, cmx.SINGLE, hop_bytes(1, 12) * Vh),
        (f"hop_fp64_recon{r64}", 0, cmx.DOUBLE, hop_bytes(0, r64) * Vh),
        ("eo_clover_op_fp16", 1, cmx.HALF, eo_clover_bytes(2, 12) * Vh),
        ("eo_clover_op_fp32", 1, cmx.SINGLE, eo_clover_bytes(1, 12) * Vh),
        ("eo_clover_op_fp64", 1, cmx.DOUBLE, eo_clover_bytes(0, r64) * Vh),
        ("stream_copy_fp32", 3, cmx.SINGLE, 2 * 96 * Vh),
        # five read streams and one write stream of 16 B elements carrying the FP16 hop's bytes
        ("stream_mix_hop_fp16", 5, cmx.HALF, (hop_bytes(cmx.HALF, 12) * Vh + 95) // 96 * 96),
    ]
    if multi:
        # 12 sources: each spinor read and written once per source, each link once
        rows.append(("hop_fp16_recon12_12src", 4, cmx.HALF, (12 * (48 + 4 + 48 + 4) + 8 * 12 * 2) * Vh))
    for name, what, prec, nbytes in rows:
        op.time(what, prec, 5)
        t = op.time(what, prec, steps)
        extra[name] = {"us": round(t * 1e3, 2), "GB/s": round(nbytes / (t * 1e-3) / 1e9, 1)}
        print(f"# {name}: {t * 1e3:.1f} us  {extra[name]['GB/s']} GB/s", file=sys.stderr, flush=True)
    return extra


def fp16_hop(cmx, op, Vh, steps, warmup):
    """The FP16/12 hopping term: `warmup` untimed launches, then exactly `steps` timed ones (HIP events on
    the operator's stream); returns (ms per launch, algorithmic GB/s, record)."""
    op.time(0, cmx.HALF, max(1, warmup))
    ms = op.time(0, cmx.HALF, steps)
    gbs = hop_bytes(cmx.HALF, 12) * Vh / (ms * 1e-3) / 1e9
    # the no-reuse figure counts all 8 neighbour spinors and their norms, 8 links and the output per site
    # (the QUDA-style accounting; not the graded one)
    no_reuse = (8 * (48 + 4) + 8 * 12 * 2 + 48 + 4) * Vh / (ms * 1e-3) / 1e9
    rec = {"us": round(ms * 1e3, 2), "GB/s": round(gbs, 1), "GB/s_no_reuse": round(no_reuse, 1),
           "gflops": round(1320.0 * Vh / (ms * 1e-3) / 1e9, 1)}
    return ms, gbs, rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--lattice", type=str, default="48,48,48,96",
                    help="lattice of the headline FP16 hopping term (the 12-source propagator lattice)")
    ap.add_argument("--solve-lattice", type=str, default="32,32,32,64",
                    help="lattice of the EO clover solves (kappa 0.1350 mixed CG, the learned preconditioner)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--train-lattice", type=str, default="16,16,16,32")
    ap.add_argument("--train-fields", type=int, default=4)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--hop-shape", type=str, default="",
                    help="block,tiled,zb,prefetch,fast: fixed launch shape of the FP16 kernels instead of the "
                         "tuned one (profiling runs: the counter passes serialise kernels and skew the tuner)")
    ap.add_argument("--sources-lattice", type=str, default="48,48,48,96")
    ap.add_argument("--batch-group", type=int, default=12,
                    help="sources per batched solve of the 12 (6 fits 64^3x128 in HBM; 12 does not)")
    ap.add_argument("--no-sources-comparison", action="store_true",
                    help="skip the 12 point sources at --sources-lattice, kappa 0.1355: plain mixed CG "
                         "(sequential, batched) vs learned FGMRES-DR with the network trained for the learned solve "
                         "(about 30 s)")
    ap.add_argument("--no-solve-lattice-kernels", action="store_true",
                    help="skip the kernel table of the solve lattice")
    ap.add_argument("--learned-config", action="store_true",
                    help="also run the learned-preconditioner configuration as stated (32^3x64, 20 000 training "
                         "pairs from 8 fields, 4 held-out fields; about 25 minutes, --c4-pairs to shorten)")
    ap.add_argument("--c4-pairs", type=int, default=20000)
    ap.add_argument("--c4-depth", type=int, default=1)
    ap.add_argument("--c4-lr", type=float, default=1e-3)
    ap.add_argument("--near-sources-lattice", type=str, default="",
                    help="lattice of the 12 point sources near the critical point (default: the solve lattice)")
    ap.add_argument("--near-kappa", type=float, default=0.158,
                    help="near-critical kappa of the stout ensemble (kappa_c is about 0.161 at 32^3x64)")
    ap.add_argument("--no-near-critical", action="store_true",
                    help="skip the critical scan and the near-critical solves (about 60 s)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {a.gpus} is not supported "
              "(the multi-GPU domain decomposition is a design, not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from chroma_amd import lib as cmx

    kappa, csw = 0.1350, 1.0
    # ---- headline: the FP16 hopping term on the 48^3x96 lattice -------------------------------------
    L = tuple(int(v) for v in a.lattice.split(","))
    Vh = int(np.prod(L)) // 2
    t0 = time.time()
    U = stout_field(cmx, L, 1)
    plaq = cmx.plaquette(L, U)
    U = cmx.gauge_apply_bc(L, U, (1, 1, 1, -1))
    print(f"# gauge {L}: plaquette {plaq:.6f} ({time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    gauge = cmx.Gauge(L, U)
    t0 = time.time()
    op = cmx.Operator(gauge, cmx.EO_CLOVER, cmx.kappa_to_mass(kappa), csw, sloppy_prec=cmx.HALF,
                      sloppy_recon=cmx.RECONS_12, autotune=True)
    if a.hop_shape:
        op.set_hop_config(cmx.HALF, *[int(v) for v in a.hop_shape.split(",")])
    print(f"# operator built and launch shapes tuned ({time.time() - t0:.1f} s): "
          f"{ {p: (op.hop_config(p, 1), op.hop_config(p, 2)) for p in (0, 1, 2)} }", file=sys.stderr, flush=True)
    ms, gbs, hop16 = fp16_hop(cmx, op, Vh, a.steps, a.warmup)
    print(f"# hop FP16/12 {L}: {ms * 1e3:.1f} us, {gbs:.0f} GB/s", file=sys.stderr, flush=True)
    extra = hop_kernels(cmx, op, Vh, a.steps, multi=False)
    extra["hop_fp16_recon12"] = hop16
    shapes = {name: {"operator": op.hop_config(p, 1), "hop": op.hop_config(p, 2)}
              for name, p in (("fp64", 0), ("fp32", 1), ("fp16", 2))}
    del op
    cpu_hop = None
    if not a.no_cpu_baseline:
        cpu_hop = cpu_hop_baseline(L, U)
        print(f"# cpu baseline hop: {cpu_hop}", file=sys.stderr, flush=True)
    del U

    # ---- the solve lattice: kernels, mixed CG, the learned preconditioner ----------------------------
    LS = tuple(int(v) for v in a.solve_lattice.split(","))
    VhS = int(np.prod(LS)) // 2
    US = cmx.gauge_apply_bc(LS, stout_field(cmx, LS, 1), (1, 1, 1, -1))
    plaq_s = cmx.plaquette(LS, US)
    gauge_s = cmx.Gauge(LS, US)
    op_s = cmx.Operator(gauge_s, cmx.EO_CLOVER, cmx.kappa_to_mass(kappa), csw, sloppy_prec=cmx.HALF,
                        sloppy_recon=cmx.RECONS_12, autotune=True)
    if a.hop_shape:
        op_s.set_hop_config(cmx.HALF, *[int(v) for v in a.hop_shape.split(",")])
    kernels_s = None
    if not a.no_solve_lattice_kernels:
        ms_s, gbs_s, hop16_s = fp16_hop(cmx, op_s, VhS, a.steps, a.warmup)
        print(f"# hop FP16/12 {LS}: {ms_s * 1e3:.1f} us, {gbs_s:.0f} GB/s", file=sys.stderr, flush=True)
        kernels_s = hop_kernels(cmx, op_s, VhS, a.steps, multi=True)
        kernels_s["hop_fp16_recon12"] = hop16_s
        m12 = kernels_s["hop_fp16_recon12_12src"]
        m12["us_per_source"] = round(m12["us"] / 12, 2)
        m12["speedup_vs_12_single"] = round(12 * ms_s * 1e3 / m12["us"], 3)
        kernels_s["lattice"] = list(LS)
        kernels_s["launch_shapes"] = {name: {"operator": op_s.hop_config(p, 1), "hop": op_s.hop_config(p, 2)}
                                      for name, p in (("fp64", 0), ("fp32", 1), ("fp16", 2))}

    solves = {}
    if not a.no_solve:
        chi = cmx.z2_source((VhS, 4, 3), 2)
        for name, sp in [("plain_mixed_cg_fp16", cmx.HALF), ("plain_mixed_cg_fp32", cmx.SINGLE)]:
            for _ in range(2):  # the first solve allocates the solver's vectors; the second is reported
                psi, res = op_s.solve(chi, solver=cmx.RELIABLE_CG, rsd=1e-10, maxit=50000, delta=0.1,
                                      sloppy_prec=sp, true_residual=True)
            solves[name] = {"seconds": round(res["solve_seconds"], 4), "iterations": res["n_count"],
                            "true_rel_resid": res["rel_resid"], "reliable_updates": res["n_reliable"],
                            "restarts": res["n_restarts"], "gflops": round(res["gflops"], 1)}
            print(f"# solve {name}: {res['n_count']} it, {res['solve_seconds']:.3f} s, rel resid "
                  f"{res['rel_resid']:.2e}", file=sys.stderr, flush=True)
        solves["lattice"] = list(LS)
        solves["kappa"] = kappa
        solves["plaquette"] = round(plaq_s, 6)
    del op_s

    small = None
    if not a.no_solve:
        small = small_volume_solves(cmx)

    learned = None
    if not a.no_solve and not a.no_learned:
        learned, net = learned_solve(cmx, gauge_s, LS, 0.1355, csw, a)

    # near the critical kappa: where the ensemble is critical, one source on both lattices, the 12
    # point sources on the solve lattice; the network has two blocks and is trained at that kappa
    critical = near = sources_nc = None
    if not a.no_solve and not a.no_near_critical:
        critical = critical_scan(cmx, gauge_s, LS)
        t0 = time.time()
        net_nc, train_nc = train_net(cmx, cmx.kappa_to_mass(a.near_kappa), csw, a, depth=2, pairs_per_source=16,
                                     lr=5e-4)
        near = {"network": {"blocks": 2, "train_seconds": round(time.time() - t0, 2),
                            "train": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in t.items()}
                                      for t in train_nc]},
                "solve_lattice": near_critical_solve(cmx, gauge_s, LS, a.near_kappa, net_nc),
                "headline_lattice": near_critical_solve(cmx, gauge, L, a.near_kappa, net_nc)}
        NSL = tuple(int(v) for v in a.near_sources_lattice.split(",")) if a.near_sources_lattice else LS
        sources_nc = learned_vs_plain_sources(cmx, NSL, a.near_kappa, csw, a, net_nc,
                                              gauge=gauge_s if NSL == LS else None, variants=NSL == LS,
                                              multigrid=True)
    del gauge

    batched = None
    if not a.no_solve and not a.no_batch:
        batched = {"uneven_volumes": batched_uneven_solves(cmx),
                   "propagator_sources": batched_sources(cmx, tuple(int(v) for v in a.sources_lattice.split(",")),
                                                      a.batch_group)}

    sources_cmp = None
    if learned is not None and not a.no_sources_comparison:
        sources_cmp = learned_vs_plain_sources(cmx, tuple(int(v) for v in a.sources_lattice.split(",")), 0.1355, csw, a, net)

    learned_c4 = learned_config(cmx, a) if a.learned_config else None

    cpu = None
    if not a.no_cpu_baseline:
        cpu = dict(cpu_hop)
        cpu["solve_lattice"] = cpu_solve_baseline(LS, US, cmx.kappa_to_mass(kappa), csw,
                                                  (solves or {}).get("plain_mixed_cg_fp16", {}).get("iterations"))
        print(f"# cpu baseline: {cpu}", file=sys.stderr, flush=True)

    traffic, traffic_src = measured_traffic("hop_fp16", L)
    out = {
        "metric": "Dslash GB/s (% HBM peak); solve time to resid 1e-10, plain vs learned",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16-storage/f32",
        "data": "synthetic (Haar + 6 stout steps, seeded)",
        "config": {"workload": "EO Wilson-clover hopping term D_eo, " + "x".join(map(str, L)) +
                               ", stout-smeared Haar field (the 12-source propagator lattice)",
                   "lattice": list(L), "kappa": kappa, "csw": csw, "plaquette": round(plaq, 6),
                   "storage": "FP16 spinors + per-site FP32 norm, 12-real FP16 links",
                   "bytes_per_site": hop_bytes(cmx.HALF, 12), "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": hop_bytes(cmx.HALF, 12) * Vh, "traffic_source": traffic_src,
                     # the same bytes with the same read/write mix, perfectly streamed (no gathers)
                     "streamed_mix_GB/s": extra.get("stream_mix_hop_fp16", {}).get("GB/s"),
                     "frac_of_streamed_mix": (round(gbs / extra["stream_mix_hop_fp16"]["GB/s"], 4)
                                              if "stream_mix_hop_fp16" in extra else None)},
        "cpu_baseline": cpu,
        "solve_time_plain_s": (learned or {}).get("plain_mixed_cg_fp16", {}).get("seconds"),
        "solve_time_learned_s": (learned or {}).get("seconds"),
        "learned": learned,
        "learned_config": learned_c4,
        "small_volume": small,
        "batched": batched,
        "sources_12_learned_vs_plain": sources_cmp,
        "critical": critical,
        "near_critical": near,
        "sources_12_near_critical": sources_nc,
        "solves": solves,
        "kernels": extra,
        "launch_shapes": shapes,
        "kernels_solve_lattice": kernels_s,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
