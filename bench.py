"""Benchmark of the batched 2-D granular-MD hot path on one MI355X.

Default line (one GPU's share at its real shape):
  * value: a single system of N = 2**20 bidisperse disks (radii 0.5 / 0.7, 50:50), phi = 0.84,
    periodic box, FP64, neighbor skin 0.3; FIRE steps then velocity-Verlet (NVE) steps at
    dt = 1e-2 from Gaussian velocities at T = 1e-4.  The default schedule is the workload's
    2000 FIRE + 5000 NVE steps; --steps K keeps that 2:5 split.  value = N * (FIRE steps + NVE
    steps) / wall time of the timed region (inputs resident in HBM).
  * roofline: HIP events around the fused velocity-Verlet launches of an untimed tail of at
    least 256 launches after the timed region (whatever --steps is); fire_roofline: events
    around the fused FIRE launches of 300 untimed FIRE steps; rebuild_roofline: events
    around the cell and neighbor rebuild of every halted batch in 3000 untimed NVE steps at
    T = 1e-3.
  * batched / rigid / adjoint: bounded runs of the 512-system uneven jamming batch (batched
    against one context per system), the 256 x 256 rigid bumpy batch (with the vertex-force
    roofline) and the 128 x 1024 differentiation workload (adjoint / forward time).
  * cpu_baseline: the oracle (the C++ FP64 restatement) on a bounded sample on this host.
--workload selects the full-length runs of the other workloads instead.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/sec (batched FIRE/NVE) + %HBM roofline; adjoint/forward time"
# HBM bytes per k_vv_fused launch on the default workload (N = 2^20, phi = 0.84), from the
# rocprofv3 FETCH_SIZE / WRITE_SIZE passes in profiles/k_vv_fused_disk_N1M_fire_nve.md, each
# scaled by what the counter reads for a copy kernel of known size (FETCH x 2.0, WRITE x 1.0),
# at the profiled run's mean neighbor count.  A run reports it scaled to its own count.
FUSED_TRAFFIC_N1M = (49443.9 * (131072.0 / 65545.5) + 32768.0) * 1024.0
FUSED_TRAFFIC_K = 5.46
FUSED_TRAFFIC_SOURCE = ("profile-derived: profiles/k_vv_fused_disk_N1M_fire_nve.md (FETCH_SIZE x 2.0 + "
                        "WRITE_SIZE x 1.0), neighbor-id bytes scaled to this run's mean count")
# The same for one fused FIRE step (k_fire_fused + k_fire_settle) on that workload (its
# profiled run's mean count: 5.46), and for
# k_vertex_force on the 256 x 256 rigid workload (profiles/k_vertex_force_rigid_256x256_fire_nve.md,
# mean vertex neighbor count 3.77).
FIRE_FUSED_TRAFFIC_N1M = ((49691.7 + 450.6) * (131072.0 / 65545.5) + 32864.1 + 0.5) * 1024.0   # fused + settle
FIRE_FUSED_TRAFFIC_K = 5.46
RIGID_VF_TRAFFIC = (43563.0 * (131072.0 / 65545.5) + 29164.1) * 1024.0
RIGID_VF_TRAFFIC_K = 3.77
RIGID_VF_TRAFFIC_SOURCE = ("profile-derived: profiles/k_vertex_force_rigid_256x256_fire_nve.md (FETCH_SIZE x 2.0 + "
                           "WRITE_SIZE x 1.0), neighbor-id bytes scaled to this run's mean count")
TAIL_LAUNCHES = 256   # fused launches timed after the timed region for the roofline


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=7000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--workload", choices=("disk1m", "jam512", "adjoint", "single1024", "rigid"), default="disk1m",
                    help="disk1m: one system of 2^20 disks, FIRE + NVE (default); jam512: the "
                         "512-system uneven batch jammed by compression, batched vs one context per system; "
                         "adjoint: gradients through FIRE minima and NVE; single1024: the plain N=1024 "
                         "system, FIRE to 1e-12 then 1e4 NVE steps; rigid: 256 x 256 rigid bumpy polygons")
    ap.add_argument("--plain-sample", type=int, default=32,
                    help="jam512: systems timed on the plain (one context per system) path")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the lean step batches directly instead of replaying hipGraphs "
                         "(same kernels; for profilers that cannot follow graph launches)")
    ap.add_argument("--no-tune", action="store_true",
                    help="jam512: skip the autotuner (untimed) and run the library defaults")
    ap.add_argument("--tune-cache", default=None,
                    help="jam512: JSON autotune cache to read (and extend); default: tune afresh")
    ap.add_argument("--rigid-fire-cap", type=int, default=20000,
                    help="--workload rigid: FIRE step cap (shorter for counter passes)")
    ap.add_argument("--rigid-nve", type=int, default=10000, help="--workload rigid: NVE steps")
    ap.add_argument("--no-legs", action="store_true",
                    help="default workload: skip the bounded batched / rigid / adjoint legs of the line")
    return ap.parse_args()


def cpu_baseline(n: int, steps: int):
    """Oracle (C++ FP64 restatement of the reference routines) on this host, bounded sample:
    `steps` FIRE steps + `steps` velocity-Verlet steps of the same N-particle system."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import oracle as orc
    from dpmd_amd import workloads as W
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    orc.lib().or_set_threads(threads)
    b = W.single_disk_system(n, 0.84, seed=1)
    d = orc.OracleDisk(b)
    t0 = time.perf_counter()
    d.fire_init(1e-2)
    for _ in range(steps):
        d.fire_step()
        d.fire_converged(0.0, 0.0, 1, 1e-12)
    d.vv_init(1e-2)
    d.vv_run(steps)
    t = time.perf_counter() - t0
    return {"value": n * 2 * steps / t, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle, N={n} phi=0.84 PBC: {steps} FIRE + {steps} NVE steps, {t:.1f} s, "
                      f"{threads} OpenMP threads"}


JAM = dict(dt=1e-2, dphi=1e-3, p_target=1e-7, max_comp=100, fire_max=50000, ftol=1e-10, pe_target=1e-16, retries=3)


def cpu_baseline_jam(b, systems: int, steps: int):
    """Oracle on this host, bounded sample of the jam workload: `steps` FIRE steps of the first
    `systems` systems of the batch (the per-step work of a jamming round)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    from dpmd_amd import workloads as W
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    orc.lib().or_set_threads(threads)
    sub = W.systems_of(b, 0, systems)
    d = orc.OracleDisk(sub)
    t0 = time.perf_counter()
    d.fire_init(JAM["dt"], 2)
    for _ in range(steps):
        d.fire_step(2)
        d.fire_converged(0.0, 0.0, 1, JAM["ftol"])
    t = time.perf_counter() - t0
    n = int(sub.off[-1])
    return {"value": n * steps / t, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle, first {systems} systems of the batch ({n} disks): {steps} FIRE steps, "
                      f"{t:.1f} s, {threads} OpenMP threads"}


def settled(P, off, ftol: float, pe_target: float):
    """Per system: does the state it was decided on meet the round's stop test (max|F| < ftol
    or <pe> < pe_target)?  A system decided on a round that ran out of steps (after its
    retries, once it has been compressed) does not; its relaxation had not finished, so
    rounding-level differences between two paths have grown along it."""
    import numpy as np

    from dpmd_amd import _lib as L
    C = L.ORDER_CANONICAL
    f = np.hypot(P.get("FORCE_X", C), P.get("FORCE_Y", C))
    fmax = np.maximum.reduceat(f, off[:-1])
    P.compute_pe_total()
    pe = P.get("PE_TOTAL") / np.diff(off)
    return (fmax < ftol) | (pe < pe_target)


def jam_end_state(P):
    """Canonical positions, PE_s and the static pressure (velocities zeroed first) of every
    system after a jamming run.  The static pressure is compared because a batch's later
    rounds restart FIRE on every system, which zeroes the velocities of systems decided
    earlier (as the reference's FIRE init does, fire.cuh:90-95), while a system run alone
    keeps the velocities of its last round; the stress tensor's kinetic part then differs."""
    from dpmd_amd import _lib as L
    C = L.ORDER_CANONICAL
    P.compute_pe_total()
    P.zero_velocities()
    P.compute_pressure()
    return P.get("POS_X", C), P.get("POS_Y", C), P.get("PE_TOTAL"), P.get("PRESSURE")


def compare_jam(st, ok, end_b, off, s, Q, acc):
    """System s on the plain path against the batched run: canonical positions bit for bit;
    energy and static pressure relative for jammed systems (kept apart by whether the
    decided state met the stop test), absolute for the others (P ~ 0 from cancelling terms)."""
    import numpy as np
    xb, yb, pe_b, pr_b = end_b
    x, y, pe, pr = jam_end_state(Q)
    acc["positions_bit_identical"] &= bool(np.array_equal(x, xb[off[s]:off[s + 1]])
                                           and np.array_equal(y, yb[off[s]:off[s + 1]]))
    for got, ref in ((pe[0], pe_b[s]), (pr[0], pr_b[s])):
        if st[s] == 1:
            key = "rel_jammed_settled" if ok[s] else "rel_jammed_at_cap"
            acc[key] = max(acc[key], abs(got - ref) / max(abs(ref), 1e-300))
        else:
            acc["abs_unjammed"] = max(acc["abs_unjammed"], abs(got - ref))


def run_jam512(args):
    """512 uneven systems (N_s in [64, 4096], phi_s in [0.80, 0.86], even systems periodic, odd
    walled) jammed by dphi = 1e-3 compressions + FIRE (max|F| < 1e-10 per round) until
    P_s >= 1e-7.  Batched: one context; each FIRE round is one persistent launch with one
    workgroup per system (stopped systems cost nothing).  Plain: one context per system with
    per-step launches (graph-captured batches), timed on a sample."""
    import numpy as np

    from dpmd_amd import _lib as L
    from dpmd_amd import workloads as W
    from dpmd_amd.system import Disk, jam_compress

    b = W.uneven_batch(S=512, seed=1)
    sizes = np.diff(b.off)
    kw = dict(predicate=L.FIRE_PRED_FORCE_OR_PE, a=JAM["ftol"], b=JAM["pe_target"], mode=L.FORCE_PER_SYSTEM,
              use_graph=not args.no_graph, retries=JAM["retries"])

    def load(bb):
        P = Disk()
        W.load_into(P, bb)
        P.init_neighbors()
        P.synchronize()
        return P

    if args.warmup > 0:   # code paths and graph instantiation on a small batch of one system
        jam_compress(load(W.system_of(b, 0)), JAM["dt"], JAM["dphi"], JAM["p_target"], 2, args.warmup, **kw)
    P = load(b)
    log(f"batched: S={b.S} N={b.N}")
    tune = None
    if not args.no_tune:   # untimed, like the warm-up: picks skin, cell size, block, batch length
        from dpmd_amd import autotune as T
        t0 = time.perf_counter()
        res = T.autotune(P, T.FIRE, JAM["dt"], mode=L.FORCE_PER_SYSTEM, steps=256, cache=args.tune_cache)
        tune = {"config": res.config, "rate": res.rate, "trials": len(res.trials),
                "seconds": time.perf_counter() - t0, "signature": res.signature}
        log(f"autotune: {res.config} ({len(res.trials)} trials, {tune['seconds']:.2f} s)")
    t0 = time.perf_counter()
    st, nc, work = jam_compress(P, JAM["dt"], JAM["dphi"], JAM["p_target"], JAM["max_comp"], JAM["fire_max"], **kw)
    P.synchronize()
    t_b = time.perf_counter() - t0
    steps = P.get("FIRE_STEPS")
    ok = settled(P, b.off, JAM["ftol"], JAM["pe_target"])
    end_b = jam_end_state(P)
    rate_b = work / t_b
    log(f"batched done: {t_b:.2f} s, work {work:.3e} particle-steps, {rate_b:.3e}/s; "
        f"jammed {int((st == 1).sum())} ({int(((st == 1) & ~ok).sum())} on a capped round), "
        f"dropped {int((st == -1).sum())}, max steps/system {int(steps.max())}")

    # plain path on a deterministic sample spread over the size range
    order = np.argsort(sizes, kind="stable")
    sample = sorted(order[np.linspace(0, b.S - 1, args.plain_sample).astype(int)].tolist())
    work_p, t_p, same = 0, 0.0, True
    acc = dict(rel_jammed_settled=0.0, rel_jammed_at_cap=0.0, abs_unjammed=0.0, positions_bit_identical=True)
    for s in sample:
        Q = load(W.system_of(b, s))
        # the plain method: one system per launch sequence (per-step launches, no persistent
        # per-system kernel)
        L.check(L.lib().dpmd_set_option(Q.ctx, L.OPT_PERSISTENT_FIRE, 0), Q.ctx)
        if tune:   # the same per-system configuration, so both paths do the same arithmetic
            from dpmd_amd import autotune as T
            T.apply(Q, tune["config"])
        t1 = time.perf_counter()
        qst, qnc, qwork = jam_compress(Q, JAM["dt"], JAM["dphi"], JAM["p_target"], JAM["max_comp"], JAM["fire_max"],
                                       **kw)
        Q.synchronize()
        t_p += time.perf_counter() - t1
        work_p += qwork
        compare_jam(st, ok, end_b, b.off, s, Q, acc)
        same = same and qst[0] == st[s] and qnc[0] == nc[s]
    rate_p = work_p / t_p
    log(f"plain sample: {len(sample)} systems, {t_p:.2f} s, {rate_p:.3e}/s; same decisions {same}; {acc}")
    out = {
        "metric": "particle-steps/sec (batched FIRE jamming, 512 uneven systems)",
        "value": rate_b, "unit": "particle-steps/s", "n_gpus": 1, "steps": int(steps.sum()),
        "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "512 uneven disk systems (N_s 64..4096, phi 0.80..0.86, half PBC half walls), "
                               "jammed by 1e-3 compressions + FIRE (max|F|<1e-10 or <pe><1e-16 per round, cap "
                               "50000 steps, a capped round run again up to 3 times) until P >= 1e-7",
                   "n_particles": int(b.N), "batched_seconds": t_b, "work_particle_steps": int(work),
                   "jammed": int((st == 1).sum()), "dropped": int((st == -1).sum()),
                   "compressions_max": int(nc.max()), "fire_steps_max": int(steps.max()),
                   "plain_sample_systems": len(sample), "plain_seconds": t_p, "plain_rate": rate_p,
                   "batched_over_plain": rate_b / rate_p,
                   "jammed_on_capped_round": int(((st == 1) & ~ok).sum()), "same_decisions": bool(same),
                   "positions_bit_identical": acc["positions_bit_identical"],
                   "plain_vs_batched_max_rel_diff_jammed_settled": acc["rel_jammed_settled"],
                   "plain_vs_batched_max_rel_diff_jammed_at_cap": acc["rel_jammed_at_cap"],
                   "plain_vs_batched_max_abs_diff_unjammed": acc["abs_unjammed"],
                   "autotune": tune},
        "roofline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_jam(b, 32, 3000)
    print(json.dumps(out))


def run_single1024(args):
    """N = 1024 bidisperse disks, phi = 0.85, PBC: FIRE until max|F_i| < 1e-12, then 1e4
    velocity-Verlet steps (dt 1e-2) from Gaussian velocities at T = 1e-4 with zero net
    momentum.  Timed on the default engine (one persistent workgroup for FIRE) and with
    per-step launches; the oracle runs the same protocol on the host (bit-identical)."""
    import numpy as np

    from dpmd_amd import _lib as L
    from dpmd_amd import workloads as W
    from dpmd_amd.system import Disk, VelocityVerlet, run_fire

    b = W.single_disk_system(1024, 0.85, seed=1)

    def go(engine):
        P = Disk()
        W.load_into(P, b)
        P.init_neighbors()
        L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_PERSISTENT_FIRE, engine), P.ctx)
        P.synchronize()
        t0 = time.perf_counter()
        steps, conv = run_fire(P, 1e-2, 1000000, L.FIRE_PRED_MAX_FORCE, 1e-12, use_graph=not args.no_graph)
        P.synchronize()
        t1 = time.perf_counter()
        P.compute_pe_total(); P.compute_pressure()
        pe, pr = P.get("PE_TOTAL")[0], P.get("PRESSURE")[0]
        P.rand_normal_2d("VEL_X", "VEL_Y", 1, 1)
        P.set_average_velocity(0.0, 0.0)
        P.set_temperature(1e-4)
        vv = VelocityVerlet(P, 1e-2)
        vv.init()
        vv.run(10000, use_graph=not args.no_graph)
        P.synchronize()
        t2 = time.perf_counter()
        P.compute_ke_total(); P.compute_pe_total()
        return dict(fire_steps=steps, converged=bool(conv), fire_seconds=t1 - t0, nve_seconds=t2 - t1,
                    pe=pe, pressure=pr, e_final=P.get("KE_TOTAL")[0] + P.get("PE_TOTAL")[0])

    if args.warmup > 0:
        go(1)
    res = go(1)
    plain = go(0)
    total_steps = res["fire_steps"] + 10000
    t = res["fire_seconds"] + res["nve_seconds"]
    log(f"single1024: FIRE {res['fire_steps']} steps {res['fire_seconds']:.3f} s, NVE {res['nve_seconds']:.3f} s; "
        f"per-step launches FIRE {plain['fire_seconds']:.3f} s")
    out = {
        "metric": "particle-steps/sec (FIRE to max|F|<1e-12 + 1e4 NVE steps, N=1024)",
        "value": 1024 * total_steps / t, "unit": "particle-steps/s", "n_gpus": 1, "steps": total_steps,
        "warmup": args.warmup, "ms_per_step": t * 1e3 / total_steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "plain single system: N=1024 bidisperse disks (0.5/0.7), phi=0.85, PBC, FIRE to "
                               "max|F_i|<1e-12, then 1e4 NVE steps dt=1e-2 at T=1e-4",
                   **{k: v for k, v in res.items()},
                   "per_step_launches": {"fire_seconds": plain["fire_seconds"], "nve_seconds": plain["nve_seconds"],
                                         "fire_steps": plain["fire_steps"]},
                   "engines_bit_identical": bool(plain["pe"] == res["pe"] and plain["pressure"] == res["pressure"]
                                                 and plain["fire_steps"] == res["fire_steps"])},
        "roofline": None,
    }
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as orc
        orc.lib().or_set_threads(1)
        d = orc.OracleDisk(b)
        t0 = time.perf_counter()
        rsteps, _ = d.fire_minimize(1e-2, 1000000, pred=1, ftol=1e-12)
        tf = time.perf_counter() - t0
        pe_o, pr_o = d.pe_total()[0], d.pressure()[0]
        out["config"]["oracle_fire_steps"] = int(rsteps)
        out["config"]["oracle_pe_rel_diff"] = abs(pe_o - res["pe"]) / abs(pe_o)
        out["config"]["oracle_pressure_rel_diff"] = abs(pr_o - res["pressure"]) / abs(pr_o)
        out["cpu_baseline"] = {"value": 1024 * rsteps / tf, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                               "sample": f"oracle, the same FIRE minimisation ({rsteps} steps, {tf:.1f} s), 1 thread"}
    print(json.dumps(out))


def run_rigid(args):
    """256 wall-box systems x 256 rigid bumpy polygons (n_v uniform in [12, 26], circle radii
    0.5 / 0.7, bump radius half the vertex spacing), phi = 0.70: FIRE on centre positions and
    angles (max|F| < 1e-9, cap 20000 steps), then 1e4 velocity-Verlet steps (dt 1e-2) at
    T = 1e-4.  The roofline is taken for the vertex-force kernel over the last NVE steps."""
    import numpy as np

    from dpmd_amd import _lib as L
    from dpmd_amd import workloads as W
    from dpmd_amd.system import RigidBumpy, VelocityVerlet, run_fire

    lib = L.lib()
    b = W.rigid_batch(S=256, n_per=256, phi=0.70, seed=1)

    def load():
        P = RigidBumpy()
        W.load_rigid_into(P, b)
        P.init_neighbors()
        P.synchronize()
        return P

    if args.warmup > 0:
        Q = load()
        run_fire(Q, 1e-2, args.warmup, L.FIRE_PRED_MAX_FORCE, 1e-9, mode=L.FORCE_WALL, use_graph=not args.no_graph)
        del Q
    P = load()
    log(f"rigid: S={b.S} N={b.N} vertices={b.Nv}")
    n_nve = max(2, args.rigid_nve)
    n_prof = min(500, n_nve // 2)
    t0 = time.perf_counter()
    steps, conv = run_fire(P, 1e-2, args.rigid_fire_cap, L.FIRE_PRED_MAX_FORCE, 1e-9, mode=L.FORCE_WALL,
                           use_graph=not args.no_graph)
    P.synchronize()
    t_fire = time.perf_counter() - t0
    fire_work = int(P.get("FIRE_STEPS").astype(np.int64).sum()) * 256
    P.rand_normal_2d("VEL_X", "VEL_Y", 1, 1)
    P.set_average_velocity(0.0, 0.0)
    P.set_temperature(1e-4)
    vv = VelocityVerlet(P, 1e-2, wall=True)
    vv.init()
    t1 = time.perf_counter()
    vv.run(n_nve - n_prof, use_graph=not args.no_graph)
    # The timed segment brackets every vertex-force launch with its own HIP event pair, so it
    # launches directly; its batches halt for rebuilds (the host finishes those steps) instead
    # of running the device rebuild chain each step, whose many short launches would leave
    # the stream waiting on the host between an event and its kernel.
    # Its steps are two-launch steps, so every timed launch is the full kernel (with the vertex
    # energy, whose bytes are counted below); inside fused batches the steps skip that energy.
    L.check(lib.dpmd_set_option(P.ctx, L.OPT_DEVICE_REBUILD, 1), P.ctx)
    L.check(lib.dpmd_set_option(P.ctx, L.OPT_FUSED_VV, 0), P.ctx)
    lib.dpmd_profile(P.ctx, 1)                        # HIP events around every vertex-force launch
    vv.run(n_prof, use_graph=False)
    L.check(lib.dpmd_set_option(P.ctx, L.OPT_DEVICE_REBUILD, 0), P.ctx)
    L.check(lib.dpmd_set_option(P.ctx, L.OPT_FUSED_VV, 1), P.ctx)
    P.synchronize()
    t_nve = time.perf_counter() - t1
    tot_ms, cnt = C.c_double(), C.c_int()
    L.check(lib.dpmd_profile_read(P.ctx, C.byref(tot_ms), C.byref(cnt)), P.ctx)
    lib.dpmd_profile(P.ctx, 0)
    cap, mx, nnb = P.neighbor_info()
    k = nnb / b.Nv
    # k_vertex_force per vertex: system id (4), x, y, r (24), count (4), k ids (4k) read;
    # fx, fy, pe (24) written -> 56 + 4k bytes (neighbor gathers excluded)
    bytes_per_launch = b.Nv * (56 + 4 * k)
    avg_ms = tot_ms.value / max(1, cnt.value)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    rate_nve = b.N * n_nve / t_nve
    log(f"rigid: FIRE {steps} steps {t_fire:.2f} s (converged {bool(conv)}), NVE {t_nve:.2f} s; "
        f"k_vertex_force {avg_ms * 1e3:.1f} us")
    out = {
        "metric": "particle-steps/sec (rigid bumpy polygons, 256 x 256, FIRE + NVE)",
        "value": (fire_work + b.N * n_nve) / (t_fire + t_nve), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": steps + n_nve, "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "256 wall-box systems x 256 rigid bumpy polygons (n_v 12..26, R 0.5/0.7), phi=0.70; "
                               f"FIRE (max|F|<1e-9, cap {args.rigid_fire_cap}) then {n_nve} NVE steps dt=1e-2 "
                               "at T=1e-4",
                   "n_particles": b.N, "n_vertices": b.Nv, "fire_steps": steps, "fire_converged": bool(conv),
                   "fire_seconds": t_fire, "fire_particle_steps": fire_work, "nve_seconds": t_nve,
                   "nve_particle_steps_per_s": rate_nve, "nve_vertex_steps_per_s": b.Nv * n_nve / t_nve,
                   "mean_vertex_neighbors": k},
        "roofline": {"bound": "hbm", "kernel": "k_vertex_force (vertex contact forces, walls)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": RIGID_VF_TRAFFIC + b.Nv * 4 * (k - RIGID_VF_TRAFFIC_K),
                     "traffic_source": RIGID_VF_TRAFFIC_SOURCE, "avg_launch_us": avg_ms * 1e3, "launches": cnt.value,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
    }
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as orc
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        orc.lib().or_set_threads(threads)
        sub = W.rigid_batch(S=16, n_per=256, phi=0.70, seed=1)
        d = orc.OracleRigid(sub)
        t0 = time.perf_counter()
        d.vv_init(1e-2, 1)
        d.vv_run(100, 1)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": sub.N * 100 / tc, "unit": "particle-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"oracle, 16 of the systems (4096 bodies), 100 NVE steps, {tc:.1f} s"}
    print(json.dumps(out))


def run_adjoint(args):
    """128 periodic systems x 1024 bidisperse disks at phi = 0.86.
    (a) dP_s/dr through FIRE minima (max|F| < 1e-12): implicit differentiation on each
        system's rigid backbone with per-system MINRES to a 1e-10 relative residual;
    (b) dMSD_s/dv0 through 1000 velocity-Verlet steps (dt 1e-2, T = 1e-4) with checkpoints
        every 50 steps.
    Both are checked against central finite differences (h = 1e-6) on the GPU: (a) where the
    contact network is the same at r_i +- h, (b) through the 1000 steps as they come and through
    a 50-step horizon where the MSD is smooth.  The timing ratio is adjoint / forward."""
    import numpy as np

    from dpmd_amd import _lib as L
    from dpmd_amd import adjoint as A
    from dpmd_amd import workloads as W
    from dpmd_amd.system import Disk, VelocityVerlet, run_fire

    S, n = 128, 1024
    b = W.disk_batch([n] * S, [0.86] * S, [0] * S, seed=1)
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    mode = L.FORCE_PBC
    sync = P.synchronize
    C = L.ORDER_CANONICAL
    log(f"adjoint: S={S} N={b.N}; FIRE to max|F| < 1e-12")
    t0 = time.perf_counter()
    steps, conv = run_fire(P, 1e-2, 300000, L.FIRE_PRED_MAX_FORCE, 1e-12, mode=mode)
    P.zero_velocities()
    sync()
    t_fwd_a = time.perf_counter() - t0
    log(f"minimised: {steps} FIRE steps (all stopped: {conv}), {t_fwd_a:.2f} s")
    t0 = time.perf_counter()
    dr, dLx, dLy, it, res = A.pressure_grad(P, mode, tol=1e-10, max_iter=50000)
    sync()
    t_adj_a = time.perf_counter() - t0
    log(f"implicit dP/dr: {it} MINRES iterations, residual {res:.1e}, {t_adj_a:.3f} s")
    # Central differences through re-minimisation, one system at a time: FIRE from the batch's
    # minimum to max|F| < 1e-13, then a Newton polish where its Hessian solve converges.  The
    # pressure is differentiable in r_i only while the contact network stays the same, and the
    # differences resolve the derivative only where the response is linear over the step: a
    # probe counts when the minima at r_i, r_i +- h and r_i +- 2h have the same contacts and
    # the differences at h and 2h agree to 1e-7 relative.  The candidates of a system are its
    # particles in order of |dP/dr_i|; the first that qualifies is taken.
    xm, ym = P.get("POS_X", C), P.get("POS_Y", C)
    h = 1e-6

    def minimum(q):
        Q = Disk(); W.load_into(Q, q); Q.init_neighbors()
        run_fire(Q, 1e-2, 300000, L.FIRE_PRED_MAX_FORCE, 1e-13, mode=mode)
        try:   # the polish needs a Hessian solve that converges (rattlers can leave it singular)
            A.newton_minimize(Q, mode, max_newton=20, ftol=1e-14)
        except L.DpmdError:
            run_fire(Q, 1e-2, 300000, L.FIRE_PRED_MAX_FORCE, 1e-13, mode=mode)
        Q.zero_velocities(); Q.compute_forces_mode(mode); Q.compute_pressure()
        fres = float(np.max(np.hypot(Q.get("FORCE_X"), Q.get("FORCE_Y"))))
        x, y = Q.get("POS_X", C), Q.get("POS_Y", C)
        dx = x[:, None] - x[None, :]; dx -= q.Lx[0] * np.rint(dx / q.Lx[0])
        dy = y[:, None] - y[None, :]; dy -= q.Ly[0] * np.rint(dy / q.Ly[0])
        touch = np.hypot(dx, dy) < q.rad[:, None] + q.rad[None, :] - 1e-10   # contacts of overlap > 1e-10
        np.fill_diagonal(touch, False)
        return Q.get("PRESSURE")[0], touch, fres

    errs_a, probes_a, skipped_a, fres_a = [], [], 0, []
    for s in (0, 37, 81, S - 1, 11, 64, 100, 23):
        if len(errs_a) == 4:
            break
        sl = slice(s * n, (s + 1) * n)
        one = W.system_of(b, s)
        one.x, one.y = xm[sl].copy(), ym[sl].copy()
        _, touch0, f0 = minimum(one)
        fres_a.append(f0)
        for i in np.argsort(-np.abs(dr[sl]))[:16]:
            res_pm = []
            for delta in (h, -h, 2 * h, -2 * h):
                q = W.system_of(one, 0)
                q.rad = q.rad.copy(); q.rad[i] += delta
                res_pm.append(minimum(q))
            fres_a += [r_[2] for r_ in res_pm]
            fd = (res_pm[0][0] - res_pm[1][0]) / (2 * h)
            fd2 = (res_pm[2][0] - res_pm[3][0]) / (4 * h)
            if not all(np.array_equal(r_[1], touch0) for r_ in res_pm) or abs(fd - fd2) > 1e-7 * abs(fd):
                skipped_a += 1
                continue
            errs_a.append(abs(fd - dr[s * n + i]) / abs(dr[s * n + i]))
            probes_a.append([int(s), int(i)])
            break
    log(f"dP/dr vs central differences (h = {h:g}; same contacts at +-h, +-2h, linear response; {skipped_a} candidates "
        f"skipped): rel err {['%.1e' % e for e in errs_a]}, largest residual force {max(fres_a):.1e}")

    # (b) NVE
    P.rand_normal_2d("VEL_X", "VEL_Y", 7, 1)
    P.set_average_velocity(0.0, 0.0)
    P.set_temperature(1e-4)
    xs, ys = P.get("POS_X", C), P.get("POS_Y", C)
    vx0, vy0 = P.get("VEL_X", C), P.get("VEL_Y", C)
    nsteps, every = 1000, 50
    sid = np.repeat(np.arange(S), n)

    def reset(vx, vy):
        A.set_canonical(P, "POS_X", xs); A.set_canonical(P, "POS_Y", ys)
        A.set_canonical(P, "VEL_X", vx); A.set_canonical(P, "VEL_Y", vy)
        P.check_neighbors(True)
        VelocityVerlet(P, 1e-2, mode=mode).init()
        sync()

    def msd():
        ax, ay = P.get("POS_X", C), P.get("POS_Y", C)
        return np.bincount(sid, weights=(ax - xs) ** 2 + (ay - ys) ** 2) / n, ax, ay

    reset(vx0, vy0)
    VelocityVerlet(P, 1e-2, mode=mode).run(nsteps)   # warm the graphs
    reset(vx0, vy0)
    t0 = time.perf_counter()
    VelocityVerlet(P, 1e-2, mode=mode).run(nsteps)
    sync()
    t_plain = time.perf_counter() - t0
    reset(vx0, vy0)
    t0 = time.perf_counter()
    A.vv_run_checkpointed(P, nsteps, every)
    sync()
    t_fwd_b = time.perf_counter() - t0
    m0, ax, ay = msd()
    t0 = time.perf_counter()
    _, _, gvx, gvy = A.vv_backprop(P, 2 * (ax - xs) / n, 2 * (ay - ys) / n, None, None)
    sync()
    t_adj_b = time.perf_counter() - t0
    log(f"NVE: forward {t_plain:.3f} s (checkpointed {t_fwd_b:.3f} s), backprop {t_adj_b:.3f} s")
    # Central differences of the 1000-step MSD (h = 1e-6): contacts open and close along the
    # trajectory, where the MSD has no second derivative, so these are reported as they come.
    errs_b = []
    for i in np.random.default_rng(1).choice(b.N, 12, replace=False):
        e = np.zeros(b.N); e[i] = h
        reset(vx0 + e, vy0); VelocityVerlet(P, 1e-2, mode=mode).run(nsteps); mp = msd()[0]
        reset(vx0 - e, vy0); VelocityVerlet(P, 1e-2, mode=mode).run(nsteps); mm = msd()[0]
        fd = (mp - mm)[sid[i]] / (2 * h)
        errs_b.append(abs(fd - gvx[i]) / abs(gvx[i]))
    log(f"dMSD/dv0 (1000 steps) vs central differences: rel err median {np.median(errs_b):.1e}, max {max(errs_b):.1e}")
    # The same gradient through a 50-step horizon (one checkpoint segment), checked where the
    # MSD is smooth: a component counts when its central differences at h and 2h agree to 1e-7
    # relative (for a smooth function they differ by O(h^2); where a contact opens or closes
    # within the perturbation they do not, and the differences themselves are not that accurate).
    nfd = 50
    reset(vx0, vy0)
    A.vv_run_checkpointed(P, nfd, every)
    sync()
    _, ax, ay = msd()
    _, _, gvx_s, _ = A.vv_backprop(P, 2 * (ax - xs) / n, 2 * (ay - ys) / n, None, None)

    def fd_at(i, hh):
        e = np.zeros(b.N); e[i] = hh
        reset(vx0 + e, vy0); VelocityVerlet(P, 1e-2, mode=mode).run(nfd); mp = msd()[0]
        reset(vx0 - e, vy0); VelocityVerlet(P, 1e-2, mode=mode).run(nfd); mm = msd()[0]
        return (mp - mm)[sid[i]] / (2 * hh)

    errs_s, skipped_s = [], 0
    for i in np.random.default_rng(2).permutation(b.N):
        if len(errs_s) == 12 or skipped_s == 96:
            break
        f1, f2 = fd_at(i, h), fd_at(i, 2 * h)
        if abs(f1 - f2) > 1e-7 * abs(f1) or gvx_s[i] == 0.0:
            skipped_s += 1
            continue
        errs_s.append(abs(f1 - gvx_s[i]) / abs(gvx_s[i]))
    log(f"dMSD/dv0 ({nfd} steps) vs central differences where smooth ({skipped_s} skipped): rel err max "
        f"{max(errs_s) if errs_s else float('nan'):.1e} over {len(errs_s)} components")
    ratio = t_adj_b / t_plain
    out = {
        "metric": "adjoint time / forward time (NVE reverse mode through 1000 steps, checkpoints every 50: "
                  "backprop incl. segment replays / plain forward run)",
        "value": ratio, "unit": "x", "n_gpus": 1, "steps": nsteps, "warmup": nsteps, "ms_per_step": t_plain * 1e3 / nsteps,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "128 PBC systems x 1024 bidisperse disks, phi 0.86; (a) dP/dr through FIRE minima by "
                               "implicit differentiation (MINRES 1e-10); (b) dMSD/dv0 through 1000 NVE steps",
                   "n_particles": int(b.N),
                   "a_fire_steps": steps, "a_forward_seconds": t_fwd_a, "a_adjoint_seconds": t_adj_a,
                   "a_adjoint_over_forward": t_adj_a / t_fwd_a, "a_minres_iterations": it, "a_minres_residual": res,
                   "a_fd_rel_err": errs_a, "a_fd_probes_system_particle": probes_a,
                   "a_fd_candidates_skipped_contact_change": skipped_a, "a_fd_largest_residual_force": max(fres_a),
                   "a_fd_within_1e-6": int(sum(e <= 1e-6 for e in errs_a)),
                   "b_forward_seconds": t_plain, "b_checkpointed_forward_seconds": t_fwd_b,
                   "b_backprop_seconds": t_adj_b, "b_fd_rel_err_median": float(np.median(errs_b)),
                   "b_fd_rel_err_max": float(max(errs_b)), "b_fd_within_1e-6": int(sum(e <= 1e-6 for e in errs_b)),
                   "b_fd_components": len(errs_b),
                   "b50_fd_rel_err_max": float(max(errs_s)) if errs_s else None,
                   "b50_fd_components": len(errs_s), "b50_fd_within_1e-6": int(sum(e <= 1e-6 for e in errs_s)),
                   "b50_fd_skipped_not_smooth": skipped_s,
                   "b_total_gradient_over_forward": (t_fwd_b + t_adj_b) / t_plain},
        "roofline": None,
    }
    print(json.dumps(out))


# The rebuild chain's algorithmic bytes per particle (in-step rebuild of one system, 7 live
# arrays of 8 bytes): cell counts zeroed and scanned (~8); hash and rank: sid, x, y read, cell
# id and rank written, one atomic per run of equal cells (30); rank scatter (20); per-cell sort
# (10) and the gather of the 7 arrays and the cell id into scratch with the inverse order
# (132); tiled neighbor build: own scratch arrays, cell id, static_index and its inverse-order
# lookup, halo candidates (x, y, r; ~1.5 reads per object) read, the 7 arrays, static_index,
# the skin reference (x, y) and the count written (156), plus 4 bytes per list entry.
REBUILD_BYTES_PER_PARTICLE = 356
REBUILD_TAIL_STEPS, REBUILD_TAIL_T = 3000, 1e-3

# ---- bounded legs of the default run: the metric's batched, rigid and adjoint parts ---------
JAM_B = dict(dt=1e-2, dphi=1e-3, p_target=1e-7, max_comp=6, fire_max=20000, ftol=1e-10, pe_target=1e-16, retries=3)


def leg_batched(args):
    """The 512-system uneven batch (N_s 64..4096, phi_s 0.80..0.86, half PBC half walls)
    through the pressure-target jamming driver, bounded: at most JAM_B['max_comp']
    compressions of 1e-3 and JAM_B['fire_max'] FIRE steps per round; a round stops at
    max|F| < 1e-10 or <pe> < 1e-16 (the reference's unjammed test).  Batched (one context,
    persistent per-system FIRE) against the plain path (one context per system, per-step
    launches) on 32 systems spread over the size range, with the same autotuned configuration;
    canonical positions compared bit for bit, energies and static pressures to rounding."""
    import numpy as np

    from dpmd_amd import _lib as L
    from dpmd_amd import workloads as W
    from dpmd_amd.system import Disk, jam_compress

    b = W.uneven_batch(S=512, seed=1)
    sizes = np.diff(b.off)
    kw = dict(predicate=L.FIRE_PRED_FORCE_OR_PE, a=JAM_B["ftol"], b=JAM_B["pe_target"], mode=L.FORCE_PER_SYSTEM,
              use_graph=not args.no_graph, retries=JAM_B["retries"])
    jargs = (JAM_B["dt"], JAM_B["dphi"], JAM_B["p_target"], JAM_B["max_comp"], JAM_B["fire_max"])

    def load(bb):
        P = Disk()
        W.load_into(P, bb)
        P.init_neighbors()
        P.synchronize()
        return P

    jam_compress(load(W.system_of(b, 0)), JAM_B["dt"], JAM_B["dphi"], JAM_B["p_target"], 1, 200, **kw)  # warm-up
    P = load(b)
    from dpmd_amd import autotune as T
    t0 = time.perf_counter()
    res = T.autotune(P, T.FIRE, JAM_B["dt"], mode=L.FORCE_PER_SYSTEM, steps=256, cache=args.tune_cache)
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    st, nc, work = jam_compress(P, *jargs, **kw)
    P.synchronize()
    t_b = time.perf_counter() - t0
    steps = P.get("FIRE_STEPS")
    ok = settled(P, b.off, JAM_B["ftol"], JAM_B["pe_target"])
    end_b = jam_end_state(P)
    rate_b = work / t_b
    log(f"batched: {t_b:.2f} s, {rate_b:.3e} particle-steps/s; jammed {int((st == 1).sum())} "
        f"({int(((st == 1) & ~ok).sum())} on a capped round), dropped {int((st == -1).sum())}, "
        f"max steps/system {int(steps.max())}")
    order = np.argsort(sizes, kind="stable")
    sample = sorted(order[np.linspace(0, b.S - 1, 32).astype(int)].tolist())
    work_p, t_p, same_decisions = 0, 0.0, True
    acc = dict(rel_jammed_settled=0.0, rel_jammed_at_cap=0.0, abs_unjammed=0.0, positions_bit_identical=True)
    for s_ in sample:
        Q = load(W.system_of(b, s_))
        L.check(L.lib().dpmd_set_option(Q.ctx, L.OPT_PERSISTENT_FIRE, 0), Q.ctx)
        T.apply(Q, res.config)
        t1 = time.perf_counter()
        qst, qnc, qwork = jam_compress(Q, *jargs, **kw)
        Q.synchronize()
        t_p += time.perf_counter() - t1
        work_p += qwork
        same_decisions = same_decisions and qst[0] == st[s_] and qnc[0] == nc[s_]
        compare_jam(st, ok, end_b, b.off, s_, Q, acc)
    rate_p = work_p / t_p
    log(f"plain: 32 systems {t_p:.2f} s, {rate_p:.3e}/s; batched/plain {rate_b / rate_p:.1f}x")
    return {"workload": "512 uneven disk systems (N_s 64..4096, phi 0.80..0.86, half PBC half walls), compressions "
                        f"of 1e-3 (at most {JAM_B['max_comp']}) + FIRE (max|F|<1e-10 or <pe><1e-16, cap "
                        f"{JAM_B['fire_max']} steps per round, a capped round run again up to "
                        f"{JAM_B['retries']} times) until P >= 1e-7",
            "n_particles": int(b.N), "value": rate_b, "unit": "particle-steps/s", "seconds": t_b,
            "work_particle_steps": int(work), "jammed": int((st == 1).sum()), "dropped": int((st == -1).sum()),
            "compressing_at_end": int((st == 0).sum()), "fire_steps_max": int(steps.max()),
            "plain_sample_systems": len(sample), "plain_value": rate_p, "plain_seconds": t_p,
            "batched_over_plain": rate_b / rate_p, "same_decisions": bool(same_decisions),
            "jammed_on_capped_round": int(((st == 1) & ~ok).sum()),
            "positions_bit_identical": acc["positions_bit_identical"],
            "max_rel_energy_pressure_diff_jammed_settled": acc["rel_jammed_settled"],
            "max_rel_energy_pressure_diff_jammed_at_cap": acc["rel_jammed_at_cap"],
            "max_abs_diff_unjammed": acc["abs_unjammed"],
            "autotune": {"config": res.config, "trials": len(res.trials), "seconds": t_tune}}


def leg_rigid(args):
    """256 wall boxes x 256 rigid bumpy polygons (phi 0.70), bounded: 2000 FIRE steps then 2000
    velocity-Verlet steps at T = 1e-4 (graph-captured batches); then an untimed tail of 300 NVE
    steps with HIP events around every vertex-force launch (the roofline)."""
    import numpy as np

    from dpmd_amd import _lib as L
    from dpmd_amd import workloads as W
    from dpmd_amd.system import RigidBumpy, VelocityVerlet, run_fire

    lib = L.lib()
    b = W.rigid_batch(S=256, n_per=256, phi=0.70, seed=1)
    P = RigidBumpy()
    W.load_rigid_into(P, b)
    P.init_neighbors()
    P.synchronize()
    run_fire(P, 1e-2, 50, L.FIRE_PRED_MAX_FORCE, 1e-9, mode=L.FORCE_WALL, use_graph=not args.no_graph)   # warm-up
    P = RigidBumpy()
    W.load_rigid_into(P, b)
    P.init_neighbors()
    P.synchronize()
    n_fire, n_nve = 2000, 2000
    t0 = time.perf_counter()
    steps, conv = run_fire(P, 1e-2, n_fire, L.FIRE_PRED_MAX_FORCE, 1e-9, mode=L.FORCE_WALL, use_graph=not args.no_graph)
    P.synchronize()
    t_fire = time.perf_counter() - t0
    fire_work = int(P.get("FIRE_STEPS").astype(np.int64).sum()) * 256
    P.rand_normal_2d("VEL_X", "VEL_Y", 1, 1)
    P.set_average_velocity(0.0, 0.0)
    P.set_temperature(1e-4)
    vv = VelocityVerlet(P, 1e-2, wall=True)
    vv.init()
    P.synchronize()
    t1 = time.perf_counter()
    vv.run(n_nve, use_graph=not args.no_graph)
    P.synchronize()
    t_nve = time.perf_counter() - t1
    # the tail runs two-launch steps, so every timed launch is the full kernel (with the
    # vertex energy) whose bytes are counted below; inside fused batches the steps skip it
    L.check(lib.dpmd_set_option(P.ctx, L.OPT_DEVICE_REBUILD, 1), P.ctx)
    L.check(lib.dpmd_set_option(P.ctx, L.OPT_FUSED_VV, 0), P.ctx)
    lib.dpmd_profile(P.ctx, 1)
    vv.run(300, use_graph=False)
    P.synchronize()
    L.check(lib.dpmd_set_option(P.ctx, L.OPT_DEVICE_REBUILD, 0), P.ctx)
    L.check(lib.dpmd_set_option(P.ctx, L.OPT_FUSED_VV, 1), P.ctx)
    tot_ms, cnt = C.c_double(), C.c_int()
    L.check(lib.dpmd_profile_read(P.ctx, C.byref(tot_ms), C.byref(cnt)), P.ctx)
    lib.dpmd_profile(P.ctx, 0)
    cap, mx, nnb = P.neighbor_info()
    k = nnb / b.Nv
    bytes_per_launch = b.Nv * (56 + 4 * k)
    avg_ms = tot_ms.value / max(1, cnt.value)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    rate = (fire_work + b.N * n_nve) / (t_fire + t_nve)
    log(f"rigid: FIRE {steps} steps {t_fire:.3f} s, NVE {n_nve} steps {t_nve:.3f} s; k_vertex_force "
        f"{avg_ms * 1e3:.1f} us ({achieved / HBM_PEAK_GBS:.2f} of peak)")
    return {"workload": "256 wall boxes x 256 rigid bumpy polygons (n_v 12..26, R 0.5/0.7), phi 0.70: 2000 FIRE "
                        "steps then 2000 NVE steps dt=1e-2 at T=1e-4",
            "n_particles": b.N, "n_vertices": b.Nv, "value": rate, "unit": "particle-steps/s",
            "fire_seconds": t_fire, "fire_particle_steps": fire_work, "nve_seconds": t_nve,
            "nve_particle_steps_per_s": b.N * n_nve / t_nve, "nve_us_per_step": t_nve * 1e6 / n_nve,
            "roofline": {"bound": "hbm", "kernel": "k_vertex_force (vertex contacts, walls)", "achieved": achieved,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": RIGID_VF_TRAFFIC + b.Nv * 4 * (k - RIGID_VF_TRAFFIC_K),
                         "traffic_source": RIGID_VF_TRAFFIC_SOURCE,
                         "avg_launch_us": avg_ms * 1e3, "launches": cnt.value,
                         "algorithmic_bytes_per_launch": bytes_per_launch, "mean_vertex_neighbors": k}}


def leg_adjoint(args):
    """128 periodic systems x 1024 disks at phi 0.86.  (a) FIRE to max|F| < 1e-12 (forward)
    then d(pressure)/d(radii) by implicit differentiation (MINRES to 1e-10); (b) 1000
    velocity-Verlet steps at T = 1e-4 (forward) and the reverse mode with checkpoints every 50
    steps.  Reports adjoint / forward time for both."""
    from dpmd_amd import _lib as L
    from dpmd_amd import adjoint as A
    from dpmd_amd import workloads as W
    from dpmd_amd.system import Disk, VelocityVerlet, run_fire

    S, n = 128, 1024
    b = W.disk_batch([n] * S, [0.86] * S, [0] * S, seed=1)
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    Co = L.ORDER_CANONICAL
    t0 = time.perf_counter()
    steps, conv = run_fire(P, 1e-2, 300000, L.FIRE_PRED_MAX_FORCE, 1e-12)
    P.zero_velocities()
    P.synchronize()
    t_fa = time.perf_counter() - t0
    t0 = time.perf_counter()
    dr, dLx, dLy, it, res = A.pressure_grad(P, L.FORCE_PBC, tol=1e-10, max_iter=50000)
    P.synchronize()
    t_aa = time.perf_counter() - t0
    P.rand_normal_2d("VEL_X", "VEL_Y", 7, 1)
    P.set_average_velocity(0.0, 0.0)
    P.set_temperature(1e-4)
    xs, ys = P.get("POS_X", Co), P.get("POS_Y", Co)
    vx0, vy0 = P.get("VEL_X", Co), P.get("VEL_Y", Co)

    def reset():
        A.set_canonical(P, "POS_X", xs); A.set_canonical(P, "POS_Y", ys)
        A.set_canonical(P, "VEL_X", vx0); A.set_canonical(P, "VEL_Y", vy0)
        P.check_neighbors(True)
        VelocityVerlet(P, 1e-2).init()
        P.synchronize()

    reset(); VelocityVerlet(P, 1e-2).run(1000)          # warm the graphs
    reset()
    t0 = time.perf_counter()
    VelocityVerlet(P, 1e-2).run(1000)
    P.synchronize()
    t_fb = time.perf_counter() - t0
    reset()
    A.vv_run_checkpointed(P, 1000, 50)
    P.synchronize()
    ax, ay = P.get("POS_X", Co), P.get("POS_Y", Co)
    t0 = time.perf_counter()
    A.vv_backprop(P, 2 * (ax - xs) / n, 2 * (ay - ys) / n, None, None)
    P.synchronize()
    t_ab = time.perf_counter() - t0
    log(f"adjoint: (a) FIRE {steps} steps {t_fa:.2f} s, implicit grad {t_aa:.3f} s ({it} MINRES it); "
        f"(b) forward {t_fb:.3f} s, backprop {t_ab:.3f} s ({t_ab / t_fb:.2f}x)")
    return {"workload": "128 PBC systems x 1024 disks, phi 0.86: (a) dP/dr through FIRE minima (max|F|<1e-12), "
                        "implicit differentiation with MINRES to 1e-10; (b) dMSD/d(x0, v0) through 1000 NVE steps, "
                        "checkpoints every 50",
            "a_fire_steps": steps, "a_fire_converged": bool(conv), "a_forward_seconds": t_fa,
            "a_adjoint_seconds": t_aa, "a_adjoint_over_forward": t_aa / t_fa, "a_minres_iterations": it,
            "a_minres_residual": res, "b_forward_seconds": t_fb, "b_backprop_seconds": t_ab,
            "b_adjoint_over_forward": t_ab / t_fb, "unit": "x"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(systems shard across GPUs with no communication; this run measures one GPU's share)"}))
        sys.exit(2)
    if args.workload == "jam512":
        run_jam512(args)
        return
    if args.workload == "adjoint":
        run_adjoint(args)
        return
    if args.workload == "single1024":
        run_single1024(args)
        return
    if args.workload == "rigid":
        run_rigid(args)
        return
    import numpy as np

    from dpmd_amd import _lib as L
    from dpmd_amd import workloads as W
    from dpmd_amd.system import Disk, VelocityVerlet, run_fire

    lib = L.lib()
    N = args.n
    K = max(7, args.steps)
    n_fire = round(K * 2 / 7)
    n_nve = K - n_fire
    b = W.single_disk_system(N, 0.84, seed=1)

    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    # warmup on the same context (graph capture, code paths), then restore the input
    if args.warmup > 0:
        run_fire(P, 1e-2, args.warmup, L.FIRE_PRED_MAX_FORCE, 1e-12, use_graph=not args.no_graph)
    P.set("POS_X", b.x); P.set("POS_Y", b.y); P.set("RAD", b.rad); P.set("MASS", b.mass); P.set("AREA", b.area)
    P.sync()
    P.init_neighbors()
    P.synchronize()


    # Timed region: FIRE then NVE, graph-captured batches.  Afterwards, untimed: a tail of NVE
    # steps launched directly with HIP events around the fused launches of every batch that
    # no rebuild halted (at least TAIL_LAUNCHES launches, whatever --steps is), then FIRE steps
    # with events around every FIRE force launch.
    log(f"timed region: {n_fire} FIRE + {n_nve} NVE steps, N={N}")
    t0 = time.perf_counter()
    steps_fire, _ = run_fire(P, 1e-2, n_fire, L.FIRE_PRED_MAX_FORCE, 1e-12, use_graph=not args.no_graph)
    P.synchronize()
    t_fire = time.perf_counter() - t0
    log(f"FIRE done: {steps_fire} steps, {t_fire:.3f} s")
    P.rand_normal_2d("VEL_X", "VEL_Y", 1, 1)
    P.set_average_velocity(0.0, 0.0)
    P.set_temperature(1e-4)
    vv = VelocityVerlet(P, 1e-2)
    vv.init()
    vv.run(n_nve, use_graph=not args.no_graph)
    P.synchronize()
    t = time.perf_counter() - t0
    log(f"NVE done: {n_nve} steps, {t - t_fire:.3f} s")

    # The tails run on the workload's own state: when the timed region was shorter than the
    # workload's 2000 FIRE steps, the minimisation is completed first (untimed) and the system
    # re-thermalised, so the figures do not depend on --steps.
    settle = max(0, 2000 - steps_fire)
    if settle:
        run_fire(P, 1e-2, settle, L.FIRE_PRED_MAX_FORCE, 1e-12, use_graph=not args.no_graph)
        P.rand_normal_2d("VEL_X", "VEL_Y", 1, 1)
        P.set_average_velocity(0.0, 0.0)
        P.set_temperature(1e-4)
        vv.init()
        vv.run(500, use_graph=not args.no_graph)
    tot_ms, cnt = C.c_double(), C.c_int()
    lib.dpmd_profile(P.ctx, 2)
    tail = 0
    while cnt.value < TAIL_LAUNCHES and tail < 20 * TAIL_LAUNCHES:
        vv.run(128, use_graph=False)
        tail += 128
        P.synchronize()
        t_ms, c_ = C.c_double(), C.c_int()
        L.check(lib.dpmd_profile_read(P.ctx, C.byref(t_ms), C.byref(c_)), P.ctx)
        tot_ms.value += t_ms.value
        cnt.value += c_.value
    lib.dpmd_profile(P.ctx, 0)
    cap, mx, nnb = P.neighbor_info()
    k = nnb / N
    # fused FIRE steps (k_fire_fused: step k's forces, kick, power / energy / max|F| chunk
    # partials and step k + 1's drift; k_fire_settle: the fold of the partials) in an untimed
    # tail: HIP events around the consecutive fused steps of every batch no rebuild halted
    lib.dpmd_profile(P.ctx, 3)
    ff_ms, ff_cnt = C.c_double(), C.c_int()
    run_fire(P, 1e-2, 300, L.FIRE_PRED_MAX_FORCE, 0.0, use_graph=False)
    L.check(lib.dpmd_profile_read(P.ctx, C.byref(ff_ms), C.byref(ff_cnt)), P.ctx)
    lib.dpmd_profile(P.ctx, 0)
    kf = P.neighbor_info()[2] / N
    # the rebuild chain (cell hash, scan, scatter, per-cell sort with the gather of the 7 live
    # arrays, tiled neighbor build with the copy-back) in an untimed NVE tail: HIP events around
    # the rebuild that finishes every halted batch
    # the FIRE tail left the system at rest; at T = 1e-3 rebuilds come about ten times as often
    # as at the workload's 1e-4 (the chain's work does not depend on T)
    P.rand_normal_2d("VEL_X", "VEL_Y", 1, 2)
    P.set_average_velocity(0.0, 0.0)
    P.set_temperature(REBUILD_TAIL_T)
    lib.dpmd_profile(P.ctx, 4)
    vv.init()
    vv.run(REBUILD_TAIL_STEPS, use_graph=False)
    rb_ms, rb_cnt = C.c_double(), C.c_int()
    L.check(lib.dpmd_profile_read(P.ctx, C.byref(rb_ms), C.byref(rb_cnt)), P.ctx)
    lib.dpmd_profile(P.ctx, 0)
    kr = P.neighbor_info()[2] / N
    rb_bytes = N * (REBUILD_BYTES_PER_PARTICLE + 4 * kr)
    rb_avg = rb_ms.value / max(1, rb_cnt.value)
    rb_gbs = rb_bytes / (rb_avg * 1e-3) / 1e9 if rb_cnt.value else None
    # one fused FIRE step per particle: x, y (16), vx, vy (16), r (8), count (4), k ids (4k),
    # m (8), last-rebuild x, y (16) read; the next step's x, y, vx, vy (32) written -> 100 + 4k
    # bytes (neighbor gathers, chunk partials and their fold excluded)
    ff_bytes = N * (100 + 4 * kf)
    ff_avg = ff_ms.value / max(1, ff_cnt.value)
    ff_gbs = ff_bytes / (ff_avg * 1e-3) / 1e9
    # k_vv_fused per particle (a context of one system loads no system id): reads x,y,rad (24),
    # count (4), k ids (4k), vx,vy,mass (24), last-rebuild x,y (16); writes vx,vy (16) and the
    # new x,y (16) -> 100 + 4k bytes (neighbor coordinate gathers excluded; forces stay in
    # registers).
    bytes_per_launch = N * (100 + 4 * k)
    avg_ms = tot_ms.value / max(1, cnt.value)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9

    # measured copy ceiling (plain copy kernel, 2 x 128 MiB per launch)
    ncopy = 1 << 24
    src, dst = C.c_void_p(), C.c_void_p()
    L.check(lib.dpmd_malloc(C.byref(src), ncopy * 8)); L.check(lib.dpmd_malloc(C.byref(dst), ncopy * 8))
    lib.dpmd_fill_f64(src, ncopy, 1.0)
    ms_copy = C.c_double()
    L.check(lib.dpmd_stream_copy(dst, src, ncopy, 20, C.byref(ms_copy)))
    copy_gbs = 2 * ncopy * 8 / (ms_copy.value * 1e-3) / 1e9
    lib.dpmd_free(src); lib.dpmd_free(dst)

    P.compute_pe_total()
    total_steps = steps_fire + n_nve
    value = N * total_steps / t
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": total_steps,
        "warmup": args.warmup,
        "ms_per_step": t * 1e3 / total_steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single large disk system: N=2^20 bidisperse (0.5/0.7, 50:50), phi=0.84, "
                               "PBC, FP64, skin 0.3; FIRE then NVE dt=1e-2 (T=1e-4)",
                   "n_particles": N, "fire_steps": steps_fire, "nve_steps": n_nve,
                   "fire_seconds": t_fire, "nve_seconds": t - t_fire, "roofline_tail_nve_steps": tail,
                   "roofline_settle_fire_steps": settle,
                   "mean_neighbors": k, "neighbor_capacity": cap, "max_neighbors": mx},
        "roofline": {"bound": "hbm",
                     "kernel": "k_vv_fused (NVE: forces, closing half kick, next opening half kick, drift, skin test)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (FUSED_TRAFFIC_N1M + N * 4 * (k - FUSED_TRAFFIC_K)) if N == 1 << 20 else None,
                     "traffic_source": FUSED_TRAFFIC_SOURCE if N == 1 << 20 else None,
                     "avg_launch_us": avg_ms * 1e3, "launches": cnt.value,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "copy_kernel_gbs": copy_gbs},
        "fire_roofline": {"bound": "hbm",
                          "kernel": "one fused FIRE step: k_fire_fused (forces, kick, chunk partials, next step's "
                                    "guessed drift) + k_fire_settle (fold of the partials, FIRE update)",
                          "achieved": ff_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": ff_gbs / HBM_PEAK_GBS,
                          "traffic": (FIRE_FUSED_TRAFFIC_N1M + N * 4 * (k - FIRE_FUSED_TRAFFIC_K))
                          if (N == 1 << 20 and FIRE_FUSED_TRAFFIC_N1M) else None,
                          "traffic_source": FUSED_TRAFFIC_SOURCE.replace("neighbor-id", "k_fire_fused + k_fire_settle "
                                                                         "rows; neighbor-id")
                          if (N == 1 << 20 and FIRE_FUSED_TRAFFIC_N1M) else None,
                          "avg_step_us": ff_avg * 1e3, "steps": ff_cnt.value,
                          "algorithmic_bytes_per_step": ff_bytes},
        "rebuild_roofline": {"bound": "hbm",
                             "kernel": "the in-step rebuild chain: k_zero, k_hash_rank, k_tile_sum, k_tile_down_self, "
                                       "k_scatter_rank, k_sort_gather, k_neighbors_tiled, k_clear_flags",
                             "achieved": rb_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": rb_gbs / HBM_PEAK_GBS if rb_gbs else None,
                             "traffic": None,
                             "avg_rebuild_us": rb_avg * 1e3, "rebuilds": rb_cnt.value,
                             "tail_nve_steps": REBUILD_TAIL_STEPS, "tail_temperature": REBUILD_TAIL_T,
                             "algorithmic_bytes_per_rebuild": rb_bytes},
    }
    del P
    if not args.no_legs:
        out["batched"] = leg_batched(args)
        out["rigid"] = leg_rigid(args)
        out["adjoint"] = leg_adjoint(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
