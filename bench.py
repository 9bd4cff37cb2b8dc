"""Benchmark: full particle-flow-map steps at 3D 256^3, 8 particles/cell.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--res 256] [--no-cpu-baseline]

One "step" is one pass of the reference schedule (3D/run_3D.py:716-831): CFL
dt, reinitialisations as scheduled (every 20 steps reseed + G2P, every step
the short G2P), midpoint velocity (grid RK4, advection, projection), particle
RK4, binning, P2G and the final projection (MGPCG to a 1e-6 relative
residual).  The workload is the 3D 256^3 full step of the mixed-precision
config (134,217,728 particles), synthetic: two coaxial vortex rings as the
initial vorticity, projected.  All state is resident in HBM before timing.

Prints ONE JSON line: particle-steps/s as `value`, plus MGPCG cells/s, a
per-phase breakdown, the roofline of the dominant kernel (HIP events on the
stream the kernel runs on) and a CPU baseline (the oracle, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "particle-flow-maps_amd"))

HBM_PEAK_GBS = 8000.0


def rings_3d():
    """Config: two coaxial rings, major radius 0.15 / 0.2 at z = 0.3 / 0.4,
    core 0.02, unit circulation (add_vortex_ring, 3D/init_conditions.py:10-22)."""
    return [
        dict(w=1.0, rad=0.15, delta=0.02, c=[0.5, 0.5, 0.3], unit_x=[1.0, 0, 0], unit_y=[0, 1.0, 0]),
        dict(w=1.0, rad=0.20, delta=0.02, c=[0.5, 0.5, 0.4], unit_x=[1.0, 0, 0], unit_y=[0, 1.0, 0]),
    ]


class EventTimer:
    """Accumulates GPU time of named regions with HIP events recorded on the
    stream the work is launched on (the current torch stream)."""

    def __init__(self, torch):
        self.torch = torch
        self.pending = []
        self.total = {}
        self.count = {}

    def region(self, name):
        timer = self

        class _R:
            def __enter__(self_inner):
                self_inner.a = timer.torch.cuda.Event(enable_timing=True)
                self_inner.a.record(timer.torch.cuda.current_stream())

            def __exit__(self_inner, *exc):
                b = timer.torch.cuda.Event(enable_timing=True)
                b.record(timer.torch.cuda.current_stream())
                timer.pending.append((name, self_inner.a, b))

        return _R()

    def flush(self):
        self.torch.cuda.synchronize()
        for name, a, b in self.pending:
            self.total[name] = self.total.get(name, 0.0) + a.elapsed_time(b)
            self.count[name] = self.count.get(name, 0) + 1
        self.pending = []


def instrumented_step(sim, timer):
    """One 3D step with event regions around each phase: the launches and
    their order are those of Simulation.step (the fused RK4 + P2G record path)."""
    from pfm import core
    from pfm.simulation import reinit_decision

    sim.i += 1
    i = sim.i
    cfg = sim.cfg
    fold = sim.folds_T
    with timer.region("max_speed_dt"):
        dt = sim.compute_dt()
    sim.last_dt = dt
    long_r, reseed, short, sim.ik = reinit_decision(i, sim.ik, cfg, 3)
    if long_r:
        with timer.region("reinit_long"):
            if reseed:
                sim._seed()
            core.g2p(sim.res, sim.dx, sim.pos, sim.u, sim.init_imp, None)
            core.reset_T_identity(sim.T_init)
            core.reset_T_identity(sim.T_grad_m)
    if short:
        with timer.region("reinit_short_g2p"):
            core.g2p(sim.res, sim.dx, sim.pos, sim.u, sim.imp_grad_m, sim.init_C)
        if not fold:
            with timer.region("reinit_short_T"):
                core.update_T(sim.T_grad_m, sim.T_init, sim.T_init)
                core.reset_T_identity(sim.T_grad_m)
    if cfg.use_midpoint_vel:
        with timer.region("midpoint_advect"):
            for a in range(sim.dim):
                sim.p2g_weight[a].copy_(sim.u[a])
            core.midpoint_advect(sim.res, sim.dx, sim.p2g_weight, sim.u, sim.psi_grid, sim.T_grid, 0.5 * dt)
        with timer.region("poisson"):
            sim.project()
    if fold:
        # the binning's count pass runs in the RK4 kernel's epilogue
        with timer.region("rk4_record"):
            sim.binning.rk4_record_count(sim.dx, sim.pos, sim.u, dt, sim.T_init, sim.init_C, sim.init_imp)
        with timer.region("binning"):
            sim.binning.bin_finish(sim.pos, sim.dx)
    else:
        with timer.region("rk4_record"):
            sim.binning.rk4_record(sim.dx, sim.pos, sim.T_grad_m, sim.u, dt, sim.T_init, sim.init_C, sim.init_imp,
                                   fold=False)
        with timer.region("binning"):
            sim.binning.bin(sim.pos, sim.dx, force=True)
    with timer.region("p2g"):
        sim.binning.p2g_from_records(sim.dx, sim.n, sim.u, W=sim.p2g_weight)
    with timer.region("poisson"):
        sim.project()
    return dt


def cpu_baseline(torch):
    """The CPU oracle (C/OpenMP fp32 restatement of the reference step) on the
    box's host cores, on bounded samples of the benchmark's workloads: one full
    step of the 3D 128^3 x 8 particles/cell ring scene (16.8 M particles; the
    long and short reinitialisation, the midpoint projection, RK4, P2G and the
    final projection) -- `value` -- and the 2D 128^2 leapfrog configuration in
    full (65,536 particles, 100 steps).  The ring velocity is computed on the
    GPU and copied over (initial conditions are not timed)."""
    import numpy as np

    sys.path.insert(0, ROOT)
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    omp = int(os.environ.get("OMP_NUM_THREADS", cores))
    import oracle
    from pfm import core

    # 2D 128^2: four Gaussian-core vortices, 4 ppc, reinit 12 / 8, 100 steps
    res2 = (128, 128)
    o2 = oracle.OracleSim(res2, reinit_every=12, reinit_every_grad_m=8, CFL=0.5, particles_per_cell=4, base_level=4,
                          tol=1e-6, tol_mode=1, seed=0, visualize_dt=1e9)
    o2.set_velocity_from_central(oracle.point_vortices(res2, 1 / 128, [(0.25, 0.45), (0.25, 0.55), (0.35, 0.45),
                                                                       (0.35, 0.55)], [1.0, -1.0, 1.0, -1.0]))
    o2.initialize_2d_particles()
    t0 = time.perf_counter()
    for _ in range(100):
        o2.step()
    t2 = time.perf_counter() - t0
    # 3D 128^3 rings, one full step
    res3 = (128, 128, 128)
    vc = torch.zeros((3,) + res3, dtype=torch.float32, device="cuda")
    for r in rings_3d():
        core.vortex_ring(res3, 1 / 128, vc, r["w"], r["rad"], r["delta"], r["c"], r["unit_x"], r["unit_y"], 2000)
    o3 = oracle.OracleSim(res3, reinit_every=20, reinit_every_grad_m=1, particles_per_cell=8, tol=1e-6, tol_mode=1,
                          seed=1, particle_order=1, visualize_dt=1e9)
    o3.set_velocity_from_central(vc.cpu().numpy())
    del vc
    t0 = time.perf_counter()
    o3.step()
    t3 = time.perf_counter() - t0
    return {"value": o3.n / t3, "unit": "particle-steps/s", "cores": omp, "kind": "port",
            "sample": f"oracle (C/OpenMP fp32 restatement of the reference step) on {omp} threads: one full step of "
                      f"3D 128^3 x 8 ppc two rings ({o3.n} particles) in {t3:.1f} s; the 2D 128^2 leapfrog "
                      f"configuration in full (100 steps, {o2.n} particles) in {t2:.1f} s",
            "config1_2d_128_100_steps_s": t2, "config1_particle_steps_per_s": o2.n * 100 / t2,
            "step_3d_128_s": t3}


def traffic_probe(sim, torch):
    """Kernels for HBM-traffic counters: two steps, a 1 GiB device copy
    (known bytes: 1 GiB read + 1 GiB written), then one more full step, so
    every step kernel is dispatched after the reference copy."""
    for _ in range(2):
        sim.step()
    src = torch.empty(2 ** 28, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.neg(src, out=dst)  # the reference stream: reads 1 GiB, writes 1 GiB
    torch.cuda.synchronize()
    del src, dst
    sim.step()
    torch.cuda.synchronize()
    print(f"# traffic probe done: copy {2 ** 30} B, {sim.n} particles", file=sys.stderr)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pmc_traffic_step_256.md")


def pmc_traffic(fetch_db, write_db):
    """Markdown table of HBM bytes per dispatch per kernel from a FETCH_SIZE
    run and a WRITE_SIZE run of `bench.py --traffic-probe`; each counter is
    scaled by what it reads for the 1 GiB copy (torch.neg) of the same run."""
    import sqlite3

    def per_kernel(path, counter):
        db = sqlite3.connect(path)
        rows = db.execute("select name, avg(counter_value), count(*) from pmc_events where counter_name = ? "
                          "group by name", (counter,)).fetchall()
        return {name: (avg * 1024.0, cnt) for name, avg, cnt in rows}

    f = per_kernel(fetch_db, "FETCH_SIZE")
    w = per_kernel(write_db, "WRITE_SIZE")
    copy = max((k for k in f if "neg_kernel" in k), key=lambda k: f[k][0])
    fs, ws = 2 ** 30 / f[copy][0], 2 ** 30 / w[copy][0]
    lines = [f"copy reference: FETCH_SIZE reads {f[copy][0] / 2**20:.1f} MiB and WRITE_SIZE "
             f"{w[copy][0] / 2**20:.1f} MiB for 1024 MiB each way; scale factors fetch x{fs:.3f}, "
             f"write x{ws:.3f}", "",
             "| kernel | dispatches | fetch B/dispatch | write B/dispatch | traffic B/dispatch |",
             "|---|---|---|---|---|"]
    rows = []
    for k in f:
        if k not in w:
            continue
        fb, wb = f[k][0] * fs, w[k][0] * ws
        rows.append((fb + wb, k, f[k][1], fb, wb))
    for tot, k, cnt, fb, wb in sorted(rows, reverse=True)[:40]:
        name = k.split("(")[0].replace("void ", "")
        lines.append(f"| {name} | {cnt} | {fb:.4g} | {wb:.4g} | {tot:.4g} |")
    return "\n".join(lines)


ISSUE_PROFILE = os.path.join(ROOT, "profiles", "pmc_issue_step_256.md")
# MI355X: 256 CUs x 4 SIMDs, a wave64 fp32 VALU instruction takes 4 cycles on
# a 16-lane SIMD (no packed FP32 here), 2.4 GHz peak engine clock
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4
N_CU = 256
# one vector-memory address unit (TA) per CU, busy at most every cycle of the
# 2.4 GHz peak engine clock
TA_PEAK_CYCLES_PER_S = 2.4e9


def pmc_valu(db_path):
    """Markdown table of VALU instructions per wave per kernel from a
    `--pmc SQ_WAVES SQ_INSTS_VALU` run (SQ counters sample a fixed subset of
    the waves, so the ratio per wave is what they give)."""
    import sqlite3

    db = sqlite3.connect(db_path)
    acc = {}
    for cname, name, val in db.execute("select counter_name, name, avg(counter_value) from pmc_events "
                                       "group by counter_name, name"):
        acc.setdefault(name, {})[cname] = val
    rows = []
    for name, d in acc.items():
        if d.get("SQ_WAVES", 0) > 0 and "SQ_INSTS_VALU" in d:
            rows.append((d["SQ_INSTS_VALU"] / d["SQ_WAVES"], name.split("(")[0].replace("void ", "")))
    lines = ["| kernel | VALU instructions per wave |", "|---|---|"]
    for v, k in sorted(rows, reverse=True)[:40]:
        lines.append(f"| {k} | {v:.1f} |")
    return "\n".join(lines)


def pmc_issue(sq_db, ta_db):
    """Markdown table of instructions issued per wave by kind (VALU, vector
    memory, LDS, scalar) per kernel from a `--pmc SQ_WAVES SQ_INSTS_VALU
    SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR SQ_INSTS_LDS SQ_INSTS_SALU` run, and the
    busy fraction of the vector-memory address unit (TA) from a `--pmc
    TA_TA_BUSY_sum GRBM_GUI_ACTIVE` run: TA busy cycles summed over the CUs
    over CUs x GPU-active cycles."""
    import sqlite3

    def per_kernel(path):
        db = sqlite3.connect(path)
        acc = {}
        for cname, name, val in db.execute("select counter_name, name, avg(counter_value) from pmc_events "
                                           "group by counter_name, name"):
            acc.setdefault(name.split("(")[0].replace("void ", ""), {})[cname] = val
        return acc

    sq, ta = per_kernel(sq_db), per_kernel(ta_db)
    lines = ["| kernel | VALU/wave | VMEM/wave (rd + wr) | LDS/wave | SALU/wave | TA busy cycles/CU | TA busy | "
             "GPU-active cycles |", "|---|---|---|---|---|---|---|---|"]
    rows = []
    for k, d in sq.items():
        w = d.get("SQ_WAVES", 0)
        if w <= 0:
            continue
        t = ta.get(k, {})
        cu_busy = t["TA_TA_BUSY_sum"] / N_CU if "TA_TA_BUSY_sum" in t else None
        busy = cu_busy / t["GRBM_GUI_ACTIVE"] if cu_busy is not None and t.get("GRBM_GUI_ACTIVE") else None
        rows.append((t.get("GRBM_GUI_ACTIVE", 0), k, d.get("SQ_INSTS_VALU", 0) / w,
                     (d.get("SQ_INSTS_VMEM_RD", 0) + d.get("SQ_INSTS_VMEM_WR", 0)) / w, d.get("SQ_INSTS_LDS", 0) / w,
                     d.get("SQ_INSTS_SALU", 0) / w, cu_busy, busy))
    for cyc, k, va, vm, ld, sa, cub, busy in sorted(rows, reverse=True)[:30]:
        b = f"{busy:.2f}" if busy is not None else "-"
        c = f"{cub:.4g}" if cub is not None else "-"
        lines.append(f"| {k} | {va:.1f} | {vm:.1f} | {ld:.1f} | {sa:.1f} | {c} | {b} | {cyc:.4g} |")
    return "\n".join(lines)


def committed_issue(kernel):
    """Per-wave instruction mix and TA busy cycles of `kernel` from the
    committed instruction-issue summary (None if absent)."""
    try:
        with open(ISSUE_PROFILE) as fh:
            for line in fh:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 8 and cells[0] == kernel:
                    return {"valu": float(cells[1]), "vmem": float(cells[2]), "lds": float(cells[3]),
                            "salu": float(cells[4]), "ta_busy_cycles_per_cu": float(cells[5]),
                            "ta_busy": float(cells[6])}
    except (OSError, ValueError):
        pass
    return None


def committed_valu(kernel):
    """VALU instructions per wave of `kernel` from the committed summary."""
    d = committed_issue(kernel)
    return d["valu"] if d else None


def committed_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed PMC summary."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 5 and cells[0] == kernel:
                    return float(cells[4])
    except OSError:
        pass
    return None


def summarize(db_path, top=25, by_grid=False, match=None):
    """Text summary of a rocprofv3 results database (kernel trace or PMC).
    With by_grid, kernel-trace rows are split by launch size (workgroups),
    which separates the multigrid levels; `match` keeps names containing it."""
    import sqlite3

    db = sqlite3.connect(db_path)
    cur = db.cursor()
    lines = []
    if by_grid:
        rows = cur.execute("select name, grid_x*grid_y*grid_z/(workgroup_x*workgroup_y*workgroup_z) as g, "
                           "count(*), sum(duration), avg(duration) from kernels group by name, g "
                           "order by sum(duration) desc").fetchall()
        lines.append(f"{'calls':>7} {'total ms':>10} {'avg us':>10} {'blocks':>8}  kernel")
        for name, g, cnt, tot, avg in rows:
            if match and match not in name:
                continue
            short = name.split("(")[0].replace("void ", "")
            lines.append(f"{cnt:>7} {tot / 1e6:>10.3f} {avg / 1e3:>10.2f} {g:>8}  {short[:90]}")
            if len(lines) > top:
                break
        return "\n".join(lines)
    rows = cur.execute("select name, count(*), sum(duration), avg(duration) from kernels "
                       "group by name order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows) or 1
    pmc = cur.execute("select counter_name, name, count(*), avg(counter_value), max(counter_value) "
                      "from pmc_events group by counter_name, name").fetchall()
    if pmc:
        lines.append(f"{'counter':<12} {'dispatches':>10} {'avg KiB/disp':>14} {'max KiB/disp':>14}  kernel")
        for cname, kname, cnt, avg, mx in sorted(pmc, key=lambda r: -r[4])[:top]:
            lines.append(f"{cname:<12} {cnt:>10} {avg:>14.1f} {mx:>14.1f}  {kname[:100]}")
        return "\n".join(lines)
    lines.append(f"{'calls':>7} {'total ms':>10} {'avg us':>10} {'%':>6}  kernel")
    for name, cnt, tot, avg in rows[:top]:
        lines.append(f"{cnt:>7} {tot / 1e6:>10.3f} {avg / 1e3:>10.2f} {100 * tot / total:>6.2f}  {name[:110]}")
    lines.append(f"total GPU kernel time {total / 1e6:.3f} ms over {sum(r[1] for r in rows)} dispatches")
    return "\n".join(lines)


def _shift(t, a, s):
    """t[i + s] along axis a (s = +-1), zero outside the domain."""
    n = t.shape[a]
    out = t.new_zeros(t.shape)
    if s > 0:
        out.narrow(a, 0, n - 1).copy_(t.narrow(a, 1, n - 1))
    else:
        out.narrow(a, 1, n - 1).copy_(t.narrow(a, 0, n - 1))
    return out


def fp64_reference(solver, u_bc, div, solid, torch, max_refine=6):
    """FP64 solution of the obstacle Poisson problem the solvers see (free-slip
    walls and solid cells; b = -div of the wall- and obstacle-corrected
    velocity), by iterative refinement: the residual b - A x is formed in
    float64 (torch, the 7-point operator over fluid neighbours) and each
    correction comes from the fp32 MGPCG `solver` at a 1e-6 relative
    residual; stops at a 1e-12 relative FP64 residual.  Returns the pressure
    (zero mean over fluid cells), the projected FP64 face velocities and the
    final relative residual, and the estimated bound of the pressure's relative
    error that residual implies."""
    f64 = torch.float64
    fl = (solid == 0).to(f64)
    nb = sum(_shift(fl, a, s) for a in range(3) for s in (1, -1)) * fl

    def A(x):
        xf = x * fl
        return fl * (nb * xf - sum(_shift(xf, a, s) for a in range(3) for s in (1, -1)))

    def center(r):
        return fl * (r - (r * fl).sum() / fl.sum())

    b = center(-div.to(f64))
    nb_ = float(torch.linalg.vector_norm(b))
    x = torch.zeros_like(b)
    rel_prev, best_rel, best_x = None, float("inf"), x
    for _ in range(max_refine):
        r = center(b - A(x))
        rel = float(torch.linalg.vector_norm(r)) / nb_
        if rel < best_rel:
            best_rel, best_x = rel, x
        # done, or the fp32 corrections no longer reduce the FP64 residual
        if rel < 1e-12 or (rel_prev is not None and rel > 0.1 * rel_prev):
            break
        rel_prev = rel
        solver.init(r.float().contiguous(), 1.0)
        solver.solve(max_iters=100, tol=1e-6, tol_mode="rel")
        x = x + solver.x.to(f64) * fl
    rel, x = best_rel, center(best_x)
    # certificate: |x - x*| <= |b - A x| / lambda_1, relative to |x|
    lam1 = 4.0 * math.sin(math.pi / (2 * max(x.shape))) ** 2
    cert = rel * nb_ / (lam1 * max(float(torch.linalg.vector_norm(x)), 1e-300))
    u64 = []
    for a in range(3):
        ua = u_bc[a].to(f64).clone()
        n = x.shape[a]
        # interior faces between two fluid cells: u -= p[f] - p[f - 1]; wall and
        # obstacle faces stay zero (apply_bc after subtract_grad_p)
        inner = ua.narrow(a, 1, n - 1)
        both = fl.narrow(a, 1, n - 1) * fl.narrow(a, 0, n - 1)
        inner.copy_((inner - (x.narrow(a, 1, n - 1) - x.narrow(a, 0, n - 1))) * both)
        ua.narrow(a, 0, 1).zero_()
        ua.narrow(a, n, 1).zero_()
        u64.append(ua)
    return x, u64, rel, cert


def kernel_rooflines(sim, torch, committed=True):
    """G2P (short reinitialisation), the fused midpoint grid pass and P2G
    (tile + combine over the last step's records) of the 3D step, each timed
    alone with HIP events on the launch stream (the current torch stream),
    against HBM (their algorithmic bytes), VALU issue and the vector-memory
    address unit (the committed per-wave instruction counts and TA busy
    cycles; the P2G tile kernel runs one wave per 64 cells and face axis).  Algorithmic bytes: G2P reads pos 12 and
    writes C 36 per particle; the midpoint pass reads the velocity samples
    (counted once, 4 B per face) and writes psi 12, T 12 and u 4 per face of
    every axis (32 B / face); P2G reads each 64 B record and its permutation
    entry (68 B / particle) and writes u and W per face (8 B / face)."""
    from pfm import core

    res, n, dx = sim.res, sim.n, sim.dx
    faces = sum(math.prod(core.face_shape(res, a)) for a in range(3))
    cells = math.prod(res)

    def timed(fn, reps=3):
        fn()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(torch.cuda.current_stream())
        for _ in range(reps):
            fn()
        b.record(torch.cuda.current_stream())
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    out = {}
    specs = [
        ("g2p", "pfm::k_g2p<3, true>", 48.0 * n, n,
         lambda: core.g2p(res, dx, sim.pos, sim.u, None, sim.init_C)),
        # the step's half dt: the RK4 stages gather again only the rows of
        # stencils that moved, so the work depends on how far the faces travel
        ("midpoint_advect", "pfm::k_midpoint_advect<3>", 32.0 * faces, faces,
         lambda: core.midpoint_advect(res, dx, sim.p2g_weight, sim.u, sim.psi_grid, sim.T_grid,
                                      0.5 * sim.last_dt)),
        ("p2g", "pfm::k_p2g_tile3", 68.0 * n + 8.0 * faces, 3 * cells,
         lambda: sim.binning.p2g_from_records(dx, n, sim.u, W=sim.p2g_weight)),
    ]
    for name, kname, alg, lanes, fn in specs:
        ms = timed(fn)
        gbs = alg / (ms * 1e-3) / 1e9
        row = {"kernel": kname, "ms": ms, "alg_bytes": alg, "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
        iss = committed_issue(kname) if committed else None
        if iss is not None:
            rate = iss["valu"] * math.ceil(lanes / 64) / (ms * 1e-3)
            row.update({"valu_insts_per_wave": iss["valu"], "valu_frac": rate / VALU_PEAK_WAVE_INSTS,
                        "vmem_insts_per_wave": iss["vmem"],
                        "ta_frac": iss["ta_busy_cycles_per_cu"] / (ms * 1e-3) / TA_PEAK_CYCLES_PER_S})
        out[name] = row
    return out


def serving_bench(n_jobs, rates, slots, torch):
    """Served job stream (Poisson arrivals at each rate, 70% 2D 128^2-1024^2,
    30% 3D 64^3-192^3, 20-200 steps, seed 3) on `slots` concurrently resident
    jobs with the fp16-level MGPCG, against the same jobs run one at a time
    with the plain fp32 solver."""
    from pfm import serving

    warm = [serving.Job(0, (64, 64), 3, 1), serving.Job(1, (32, 32, 32), 3, 2)]
    serving.Server(slots=2, mg_precision=1).run(warm, realtime=False)
    serving.Server(slots=1, mg_precision=0).run(warm, realtime=False)
    top = max(rates)
    jobs = serving.job_stream(n_jobs, top, seed=3)
    base = serving.Server(slots=1, mg_precision=0).run(jobs, realtime=True).summary()
    print(f"# serving one-at-a-time fp32: {base['jobs_per_s']:.2f} jobs/s, p50 {base['step_ms_p50']:.2f} ms "
          f"p99 {base['step_ms_p99']:.2f} ms", file=sys.stderr, flush=True)
    served = {}
    for r in rates:
        s = serving.Server(slots=slots, mg_precision=1).run(serving.job_stream(n_jobs, r, seed=3),
                                                            realtime=True).summary()
        s["speedup_jobs_per_s_vs_one_at_a_time"] = s["jobs_per_s"] / base["jobs_per_s"]
        served[f"rate_{r:g}_jobs_per_s"] = s
        print(f"# served at {r:g} jobs/s: {s['jobs_per_s']:.2f} jobs/s, p50 {s['step_ms_p50']:.2f} ms "
              f"p99 {s['step_ms_p99']:.2f} ms", file=sys.stderr, flush=True)
    batched = {}
    for r in rates[-1:]:
        s = serving.Server(slots=slots, mg_precision=1, batch_2d=True).run(
            serving.job_stream(n_jobs, r, seed=3), realtime=True).summary()
        s["speedup_jobs_per_s_vs_one_at_a_time"] = s["jobs_per_s"] / base["jobs_per_s"]
        batched[f"rate_{r:g}_jobs_per_s"] = s
        print(f"# served (batched 2D lane) at {r:g} jobs/s: {s['jobs_per_s']:.2f} jobs/s, p50 "
              f"{s['step_ms_p50']:.2f} ms p99 {s['step_ms_p99']:.2f} ms", file=sys.stderr, flush=True)
    sizes = [j.res for j in jobs]
    return {
        "workload": f"{n_jobs} jobs, Poisson arrivals, 70% 2D 128^2-1024^2 (4 ppc) / 30% 3D 64^3-192^3 (8 ppc), "
                    "log-uniform sizes, 20-200 steps, vortex-blob / vortex-ring ICs, seed 3, MGPCG 1e-6 relative",
        "jobs_2d": sum(len(r) == 2 for r in sizes), "jobs_3d": sum(len(r) == 3 for r in sizes),
        "total_steps": sum(j.steps for j in jobs),
        "one_at_a_time_fp32": base,
        "served_fp16": served,
        "served_fp16_3d_batched_2d": batched,
    }


def batch_bench(n_scenes, steps, warmup, sample, torch):
    """Batched 2D step: n_scenes independent scenes (square grids log-uniform
    in 64^2-512^2, 8-32 Gaussian point vortices each, seed 2, 4 particles per
    cell, reinit every 12 steps) advanced together, one persistent launch per
    phase replayed from HIP graphs, launch shapes autotuned; against the
    same scenes run one at a time (single-scene Simulation), timed on a
    size-stratified sample and scaled to all scenes, each sampled scene also
    checked against its batched result."""
    import numpy as np

    from pfm.batch import Batch2D, random_vortex_scenes
    from pfm.simulation import Simulation, StepConfig

    cfg = StepConfig(particles_per_cell=4, reinit_every=12, reinit_every_grad_m=8, CFL=0.5, base_level=4,
                     tol=1e-6, tol_mode="rel", write_imp_grad_m=False, visualize_dt=1e9, particle_order=1)
    scenes = random_vortex_scenes(n_scenes, seed=2)
    t0 = time.perf_counter()
    b = Batch2D([sc[0] for sc in scenes], cfg, seeds=list(range(n_scenes)))
    for k, (res, c, w, rad) in enumerate(scenes):
        b.set_point_vortices(k, c, w, rad)
    b.project()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    shapes = b.tune()
    order = np.argsort([sc[0][0] for sc in scenes], kind="stable")
    picked = [int(order[i]) for i in np.linspace(0, n_scenes - 1, sample).round().astype(int)]
    ic = {k: [b.u_view(k, a).clone() for a in range(2)] for k in picked}
    b.init_particles()
    for _ in range(warmup):
        b.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lock = []
    for _ in range(steps):
        lock.append(b.step())
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    info = b.scene_info()
    batch_ms = el / steps * 1e3
    n_part = sum(b.n)
    print(f"# batch {n_scenes} scenes: {batch_ms:.1f} ms/step ({n_part} particles), lockstep iters {lock}",
          file=sys.stderr, flush=True)
    # one at a time: the sampled scenes from the same initial velocity and seed
    single_ms, rel = [], []
    for k in picked:
        sim = Simulation(scenes[k][0], StepConfig(**{**cfg.__dict__, "seed": k}))
        for a in range(2):
            sim.u[a].copy_(ic[k][a])
        sim.initialize_2d_particles()
        for _ in range(warmup):
            sim.step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            sim.step()
        torch.cuda.synchronize()
        single_ms.append((time.perf_counter() - t1) / steps * 1e3)
        num = sum(float(torch.linalg.vector_norm(b.u_view(k, a) - sim.u[a]) ** 2) for a in range(2))
        den = sum(float(torch.linalg.vector_norm(sim.u[a]) ** 2) for a in range(2))
        rel.append((num / max(den, 1e-30)) ** 0.5)
        del sim
    one_ms = float(np.sum(single_ms)) * n_scenes / len(picked)
    sizes = [sc[0][0] for sc in scenes]
    return {
        "workload": f"{n_scenes} independent 2D scenes, square grids log-uniform 64^2-512^2 (seed 2), 8-32 "
                    "Gaussian point vortices each at uniform positions in [0.2,0.8]^2, circulations N(0,1), "
                    "core radius 1.5 dx, 4 particles/cell (cell-major numbering), reinit every 12 (grad_m 8), "
                    "CFL 0.5, MGPCG 1e-6 relative, fp32",
        "cells": int(sum(s * s for s in sizes)), "particles": int(n_part),
        "size_min": int(min(sizes)), "size_max": int(max(sizes)),
        "setup_s": t_setup, "steps": steps, "warmup": warmup,
        "ms_per_step": batch_ms,
        "particle_steps_per_s": n_part * steps / el,
        "scene_steps_per_s": n_scenes * steps / el,
        "lockstep_pcg_iters": lock,
        "scene_pcg_iters_last_step": {"min": int(info["iters"].min()), "max": int(info["iters"].max()),
                                      "mean": float(info["iters"].mean())},
        "launch_shapes": shapes,
        "one_at_a_time": {"sample_scenes": len(picked), "sample": "size-stratified (every "
                          f"{n_scenes // len(picked)}th scene in size order), Simulation per scene, same IC "
                          "and seed", "ms_per_scene_step_sum_sample": float(np.sum(single_ms)),
                          "est_ms_per_step_all_scenes": one_ms},
        "speedup_vs_one_at_a_time": one_ms / batch_ms,
        "check_vs_single_u_rel_l2_max": float(max(rel)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the served job stream")
    ap.add_argument("--serve-jobs", type=int, default=24)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-only", action="store_true", help="run only the served job stream and print it")
    ap.add_argument("--profile-steps", type=int, default=2, help="instrumented steps after the timed ones")
    ap.add_argument("--batch-scenes", type=int, default=512, help="scenes of the batched 2D step")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-sample", type=int, default=32, help="scenes timed one at a time")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 2D step")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched 2D step and print it")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for rocprofv3 --pmc runs: one plain 1 GiB copy (the counter scale reference) "
                         "then one launch of each step kernel after two steps; prints no JSON")
    ap.add_argument("--summarize", metavar="DB", help="print a text summary of a rocprofv3 results db and exit")
    ap.add_argument("--by-grid", action="store_true", help="with --summarize: split rows by launch size")
    ap.add_argument("--match", default=None, help="with --summarize --by-grid: keep kernels containing this")
    ap.add_argument("--pmc-valu", metavar="DB", help="print the VALU-instructions-per-wave table of an SQ run and exit")
    ap.add_argument("--pmc-issue", nargs=2, metavar=("SQ_DB", "TA_DB"),
                    help="print the per-wave instruction mix and TA busy table of an SQ run and a TA run and exit")
    ap.add_argument("--pmc-traffic", nargs=2, metavar=("FETCH_DB", "WRITE_DB"),
                    help="print the per-kernel HBM traffic table of two --traffic-probe PMC runs and exit")
    args = ap.parse_args()
    if args.summarize:
        print(summarize(args.summarize, top=40 if args.by_grid else 25, by_grid=args.by_grid, match=args.match))
        return
    if args.pmc_traffic:
        print(pmc_traffic(*args.pmc_traffic))
        return
    if args.pmc_valu:
        print(pmc_valu(args.pmc_valu))
        return
    if args.pmc_issue:
        print(pmc_issue(*args.pmc_issue))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the step shards by slab "
              "decomposition but this build measures a single MI355X.", file=sys.stderr)
        sys.exit(2)

    import torch

    from pfm.simulation import Simulation, StepConfig

    if args.serve_only:
        print(json.dumps(serving_bench(args.serve_jobs, (200.0,), args.serve_slots, torch)))
        return
    if args.batch_only:
        print(json.dumps(batch_bench(args.batch_scenes, args.batch_steps, 1, args.batch_sample, torch)))
        return

    res = (args.res,) * 3
    cfg = StepConfig(particles_per_cell=8, reinit_every=20, reinit_every_grad_m=1, CFL=0.5, tol=1e-6,
                     tol_mode="rel", seed=1, particle_order=1, write_imp_grad_m=False, visualize_dt=1e9)
    t_setup = time.perf_counter()
    sim = Simulation(res, cfg)
    sim.init_rings_3d(rings_3d())
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    print(f"# setup {t_setup:.1f}s, {sim.n} particles, {torch.cuda.memory_allocated() / 2**30:.1f} GiB",
          file=sys.stderr, flush=True)

    if args.traffic_probe:
        traffic_probe(sim, torch)
        return

    for _ in range(args.warmup):
        sim.step()
    torch.cuda.synchronize()
    iters_before = len(sim.poisson_iters)
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_per_step = el / args.steps * 1e3
    value = sim.n * args.steps / el
    its = sim.poisson_iters[iters_before:]
    print(f"# timed {args.steps} steps: {ms_per_step:.1f} ms/step, poisson iters {its}", file=sys.stderr,
          flush=True)
    # sanity of the timed run's end state: kinetic energy, the discrete
    # divergence left by the last projection (relative to the face-difference
    # scale), particles in the domain
    sanity = {"kinetic_energy": sim.kinetic_energy()}
    with torch.no_grad():
        div = None
        for a in range(3):
            x = sim.u[a].double()
            t = x.narrow(a, 1, x.shape[a] - 1) - x.narrow(a, 0, x.shape[a] - 1)
            div = t if div is None else div + t
        scale = sum(float(torch.linalg.vector_norm(x.double())) for x in sim.u)
        sanity["div_rel_l2"] = float(torch.linalg.vector_norm(div)) / scale
        sanity["max_abs_div"] = float(div.abs().max())
        sanity["finite"] = bool(all(torch.isfinite(x).all() for x in sim.u))
        sanity["pos_min"], sanity["pos_max"] = float(sim.pos.min()), float(sim.pos.max())
        del div

    # per-phase breakdown (instrumented steps, not part of `value`)
    timer = EventTimer(torch)
    for _ in range(args.profile_steps):
        instrumented_step(sim, timer)
    timer.flush()
    phases = {k: round(v / timer.count[k] * (2 if k == "poisson" else 1), 3) for k, v in timer.total.items()}

    # standalone Poisson solves (MGPCG cells/s): divergence of a seeded
    # 64-mode Fourier velocity with free-slip walls and a solid box obstacle
    # [0.4, 0.6]^3, to a 1e-6 relative residual; plain fp32 levels and the
    # fp16 / bf16 V-cycle storage (FP32 PCG, FP64 dots and bottom solve)
    from pfm.mgpcg import MGPCG
    from pfm.workloads import box_mask, fourier_velocity

    u_f = fourier_velocity(res, seed=0)
    solid = box_mask(res, 0.4, 0.6)
    cells = math.prod(res)
    mg_runs = {}
    p_ref = None
    # the FP64 reference of this problem (pressure and projected velocity)
    ref_solver = MGPCG([[2, 2]] * 3, list(res), dim=3, base_level=3, precision=0)
    ref_solver.set_solid(solid)
    u_bc = [x.clone() for x in u_f]
    ref_solver.apply_bc(*u_bc)
    ref_solver.divergence(*u_bc)
    div = ref_solver.u_div.clone()
    p64, u64, ref_rel, ref_cert = fp64_reference(ref_solver, u_bc, div, solid, torch)
    del ref_solver, div
    fl = solid == 0
    n_u64 = math.sqrt(sum(float((t * t).sum()) for t in u64))
    n_p64 = float(torch.linalg.vector_norm(p64[fl]))
    for name, prec in (("fp32", 0), ("fp16", 1), ("bf16", 2)):
        solver = MGPCG([[2, 2]] * 3, list(res), dim=3, base_level=3, precision=prec)
        solver.set_solid(solid)
        work = [x.clone() for x in u_f]
        solver.Poisson(*work, tol=1e-6, tol_mode="rel")  # warm
        torch.cuda.synchronize()
        reps, pits, t_p = 3, [], 0.0
        for _ in range(reps):
            for a in range(3):
                work[a].copy_(u_f[a])
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            pits.append(solver.Poisson(*work, tol=1e-6, tol_mode="rel"))
            torch.cuda.synchronize()
            t_p += time.perf_counter() - t1
        t_p /= reps
        pr = solver.p.double()
        pr = pr - pr[fl].mean()
        run = {"ms_per_solve": t_p * 1e3, "cells_per_s": cells / t_p, "iters": pits,
               "converged": bool(solver.converged),
               "pressure_rel_l2_vs_fp64": float(torch.linalg.vector_norm((pr - p64)[fl])) / n_p64,
               "velocity_rel_l2_vs_fp64": math.sqrt(sum(float(((w.double() - v) ** 2).sum())
                                                        for w, v in zip(work, u64))) / n_u64}
        if p_ref is None:
            p_ref = pr
        else:
            run["pressure_rel_l2_vs_fp32"] = float(torch.linalg.vector_norm((pr - p_ref)[fl]) /
                                                   torch.linalg.vector_norm(p_ref[fl]))
            run["speedup_vs_fp32"] = mg_runs["fp32"]["ms_per_solve"] / run["ms_per_solve"]
        mg_runs[name] = run
        del solver, work
    t_poisson = mg_runs["fp32"]["ms_per_solve"] * 1e-3
    pits = mg_runs["fp32"]["iters"]
    del u_f, u_bc, u64, p64

    # the other particle / face kernels of the step, timed alone (HIP events on
    # the launch stream) with their algorithmic bytes and VALU issue
    kernels = kernel_rooflines(sim, torch, args.res == 256)

    # roofline of the step's dominant kernel, RK4 + P2G record with the
    # update_T fold and the binning count pass (k_rk4_particles<3, true, true,
    # true>): per particle it reads pos 12 B, T_init 36 B, init_C 36 B,
    # init_imp 12 B and writes pos 12 B, the folded T_init 36 B, a 64 B record,
    # its cell 4 B and its slot 4 B: 216 B/particle (T_grad_m enters as the
    # identity); per cell the count is zero-filled (4 B) and read and written by
    # the atomics (8 B).  The timed launch is the entry: the fill and the
    # kernel.  Without the fold (k_rk4_particles<3, true>) T_grad_m is also
    # read and written instead of T_init being written, and nothing is binned:
    # 244 B/particle.
    fold = sim.folds_T
    ev_a = torch.cuda.Event(enable_timing=True)
    ev_b = torch.cuda.Event(enable_timing=True)
    nrep = 3
    # the measurement launches advance state (the step benchmark is done) by
    # the step's dt: the RK4 stages gather again only the rows of stencils
    # that moved, so the kernel's work depends on how far the particles travel
    dt_m = float(sim.last_dt)
    ev_a.record(torch.cuda.current_stream())
    for _ in range(nrep):
        if fold:
            sim.binning.rk4_record_count(sim.dx, sim.pos, sim.u, dt_m, sim.T_init, sim.init_C, sim.init_imp)
        else:
            sim.binning.rk4_record(sim.dx, sim.pos, sim.T_grad_m, sim.u, dt_m, sim.T_init, sim.init_C,
                                   sim.init_imp, fold=False)
    ev_b.record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    rk4_ms = ev_a.elapsed_time(ev_b) / nrep
    alg_bytes = 216.0 * sim.n + 12.0 * cells if fold else 244.0 * sim.n
    achieved = alg_bytes / (rk4_ms * 1e-3) / 1e9
    roof_kernel = ("pfm::k_rk4_particles<3, true, true, true>" if fold
                   else "pfm::k_rk4_particles<3, true, false, false>")
    traffic = committed_traffic(roof_kernel) if args.res == 256 else None
    # The kernel is bound by the vector-memory address unit (TA): its 155
    # vector-memory instructions per wave (108 of them the 12-byte row
    # gathers of the four velocity samples) keep the TA busy 0.95 of the GPU's
    # active cycles in the committed profile.  Its TA busy cycles per CU per
    # launch over the live launch time, against one TA cycle per engine cycle
    # at the 2.4 GHz peak clock; VALU issue and HBM beside it.
    iss = committed_issue(roof_kernel) if args.res == 256 else None
    hbm = {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
           "alg_bytes_per_launch": alg_bytes}
    if iss is not None:
        waves = math.ceil(sim.n / 64)
        rate = iss["valu"] * waves / (rk4_ms * 1e-3)
        ta_rate = iss["ta_busy_cycles_per_cu"] / (rk4_ms * 1e-3)
        valu = {"achieved": rate, "peak": VALU_PEAK_WAVE_INSTS, "unit": "wave64 VALU instructions/s",
                "frac": rate / VALU_PEAK_WAVE_INSTS, "insts_per_wave": iss["valu"]}
        roofline = {"kernel": roof_kernel, "bound": "vmem", "achieved": ta_rate, "peak": TA_PEAK_CYCLES_PER_S,
                    "unit": "TA (vector-memory address unit) busy cycles/s per CU",
                    "frac": ta_rate / TA_PEAK_CYCLES_PER_S, "traffic": traffic, "kernel_ms": rk4_ms,
                    "vmem_insts_per_wave": iss["vmem"], "ta_busy_of_active_cycles": iss["ta_busy"],
                    "waves_per_launch": waves, "valu": valu, "hbm": hbm}
    else:
        roofline = {"kernel": roof_kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel_ms": rk4_ms,
                    "alg_bytes_per_launch": alg_bytes}

    out = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: two coaxial vortex rings (Biot-Savart, projected), 8 particles/cell seed 1",
        "config": {"workload": f"3D {args.res}^3 full particle-flow-map step, 8 particles/cell",
                   "particles": sim.n, "cells": cells, "reinit_every": 20, "reinit_every_grad_m": 1,
                   "CFL": 0.5, "mgpcg_tol": "1e-6 relative", "parallelism": "single GPU"},
        "mgpcg_cells_per_s": cells / t_poisson,
        "mgpcg_ms_per_solve": t_poisson * 1e3,
        "mgpcg_iters": pits,
        "mgpcg": {"workload": f"{args.res}^3 Poisson, free-slip walls + solid box [0.4,0.6]^3, "
                              "RHS = div of 64 Fourier modes (seed 0), 1e-6 relative residual",
                  "fp64_reference": {
                      "method": "iterative refinement: float64 residuals (torch, the 7-point operator over fluid "
                                "neighbours), fp32 MGPCG corrections",
                      "relative_residual": ref_rel,
                      # |e| <= |r| / lambda_1 with lambda_1 = 4 sin^2(pi / 2N) the smallest nonzero
                      # eigenvalue of the Neumann Laplacian of the box (the obstacle domain's is of
                      # the same order: an estimate, not a proof)
                      "pressure_error_bound_estimate": ref_cert},
                  **mg_runs},
        "phase_ms": phases,
        "sanity": sanity,
        "roofline": roofline,
        "kernel_rooflines": kernels,
    }
    del sim
    torch.cuda.empty_cache()
    if not args.no_serve:
        out["serving"] = serving_bench(args.serve_jobs, (10.0, 50.0, 200.0), args.serve_slots, torch)
    if not args.no_batch:
        out["batch2d"] = batch_bench(args.batch_scenes, args.batch_steps, 1, args.batch_sample, torch)
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(torch)
        except Exception as e:  # reported, never fatal for the GPU number
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
