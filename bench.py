"""Benchmark of the DEM step on one MI355X.

Workload: the 1M-sphere polydisperse Hertz/history settling case (radii
U(1.0, 1.5) mm, CDT rolling friction 0.1, box 0.3 x 0.3 x 0.6 m periodic in x
and y, zplane floor, jittered 3.3 mm lattice, dt = 5e-6 s, skin 1 mm), run in
the FP64 mode and in the mixed FP32/FP64 mode. A step is one full timestep of
all particles (pair forces with history, walls, gravity, nve/sphere, neighbor
decide and any rebuild it triggers). Inputs are resident in HBM before the
timed region.

`--workload sweep` runs a batched calibration sweep: 512 independent
heap-formation simulations (N log-uniform in 2k-50k, r = 1.5 mm, finite
0.1 x 0.1 m mesh floor with open sides, per-instance friction U(0.2, 0.8),
restitution U(0.1, 0.9), rolling friction U(0, 0.3), mixed precision,
dt 1e-5 s) advanced together by one step kernel per step, with an autotuned
tile shape and graph-replayed step batches; it reports the aggregate rate
and, for a sample of instances, the kinetic-energy and contact-count error
against the same instance run alone in FP64.

`--workload superquadric` measures the superquadric contact-point solve:
the warm-start network (18 -> 64 -> 64 -> 3, BF16 matrix cores) is trained on
the GPU on ~10M seeded overlapping pose pairs (half-axes U(1, 3) mm,
blockiness U(2, 8), random orientations) resolved by the plain solve, then
200k unseen pairs are solved plain (homotopy from spheres) and warm-started;
it reports contacts/s, Newton iterations per contact and the agreement of
the converged points.

`--workload packed` runs the same 1M spheres from a compressed lattice
(about three touching neighbors per sphere, tangential history on each), the
history-heavy regime of a settled bed.

`--workload drum` runs the rotating-drum case instead (closed cylinder
R = 0.15 m, L = 0.1 m, 64 x 16 mantle quads split into triangles + fan end
caps = 2,176 triangles turning at 30 rpm via move/mesh rotate, ~304k spheres
r = 1 mm +- 10% filling the lower half on a cubic lattice, mu_pp 0.5,
mu_pw 0.4, e 0.3, dt 5e-6 s, 5,000 steps); it is a secondary measurement,
not the headline line.

Prints one JSON line: value = particle-timesteps/s of the mixed mode (the
FP64 figure and the mixed/FP64 ratio ride along), the roofline entry of the
fused step kernel (algorithmic HBM bytes / average kernel duration measured
with HIP events on the library's stream), and the CPU baseline: the C
restatement of the reference's algorithm on one host core over a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "liggghts-public_amd"))

HBM_PEAK_GBS = 8000.0
TRI_RECORD_BYTES = 312  # one per-triangle geometry record (TriGeom)


def algorithmic_bytes(n, npairs, ncontacts, nwall, mixed, ntri=0, ncand=0, history=True,
                      rolling_history=False):
    """Unique HBM bytes one launch of the fused step kernel must move:
    per particle  read x,y,z,r + v,m (64 B FP64) + omega,type (32 B FP64 /
                  16 B mixed) + xhold (24 B) + list length (4 B);
                  write x, v (48 B) + omega (24 B FP64 / 12 B mixed);
    per half-list pair  one 4-byte neighbor index;
    per touching contact (pair or wall)  tangential history read + write
                  (2 x 24 B FP64 / 2 x 12 B mixed);
    with meshes   per particle its candidate count (4 B), per candidate one
                  4-byte triangle index, per triangle its geometry record once,
                  per touching mesh contact its 4-byte slot id (counted in nwall
                  for the history).
    history=False: a contact model without tangential history (no_history)
    moves no history bytes; rolling_history=True (rolling_friction epsd2) moves
    the rolling torque beside the shear, twice the history bytes."""
    sh = (12 if mixed else 24) if history else 0
    if rolling_history:
        sh *= 2
    per_particle = 64 + (16 if mixed else 32) + 24 + 4 + 48 + (12 if mixed else 24)
    B = n * per_particle + 4 * npairs + 2 * sh * (ncontacts + nwall)
    if ntri:
        B += 4 * n + 4 * ncand + TRI_RECORD_BYTES * ntri + 4 * nwall
    return B


# workloads with a committed FETCH_SIZE / WRITE_SIZE profile of the step kernel
PROFILED = {"settling": "1M_settling", "packed": "packed_1M"}


def traffic_source(name):
    return (f"profiles/k_step_{name}_pmc.md: FETCH_SIZE + WRITE_SIZE per launch of this kernel, from "
            "rocprofv3 --pmc runs of this workload, scaled by a plain copy's counters (not measured "
            "in this run)")


def measured_traffic(workload, mode):
    """HBM bytes per launch of the step kernel from the committed FETCH_SIZE /
    WRITE_SIZE profile of this workload (calibrated against a plain copy), or
    None when no such profile exists."""
    if workload is None:
        return None
    path = os.path.join(ROOT, "profiles", f"k_step_{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            vals = dict(kv.split("=") for kv in line.split(":", 1)[1].split())
            return float(vals[mode]) if mode in vals else None
    return None


TUNE_STEPS = 2400  # covers three passes over every candidate tile shape (8 steps each, up to 3 tries)


def tune_shape(params, mats, P, walls, precision, meshes=(), instances=None, instance_materials=None,
               warmup=0):
    """Autotunes the step kernel's tile shape on a separate system of the same
    workload, after the same warmup steps as the timed system, so that the
    timed system starts with its shape fixed and no timed step runs a candidate
    shape. Returns (variant, persist, candidates)."""
    import pydem
    g = pydem.GpuSystem(params, mats, P, walls, precision=precision, meshes=meshes,
                        instances=instances, instance_materials=instance_materials)
    if warmup:
        g.run(warmup)
        g.set_step_shape(-1)
        g.run_continue(TUNE_STEPS)
        v, f, cand = g.step_shape()
        g.close()
        return v, f, cand
    g.set_step_shape(-1)
    g.run(TUNE_STEPS)
    v, f, cand = g.step_shape()
    g.close()
    return v, f, cand


def run_mode(params, mats, P, walls, precision, steps, warmup, timing_steps, meshes=(),
             instances=None, instance_materials=None, tune=False, graphs=False, keep=False):
    import pydem
    tuned = None
    if tune:
        tuned = tune_shape(params, mats, P, walls, precision, meshes, instances, instance_materials, warmup)
    g = pydem.GpuSystem(params, mats, P, walls, precision=precision, meshes=meshes,
                        instances=instances, instance_materials=instance_materials)
    if tuned is not None:
        g.set_step_shape(tuned[0], tuned[1])
        print("# tuned shape (variant, persist): " + ", ".join(
            f"({c[0]},{c[1]}) {c[2]:.1f} us" for c in tuned[2] if c[2] > 0), file=sys.stderr, flush=True)
    if graphs:
        g.set_graph_capture(True)
    g.run(warmup)
    sync()
    t0 = time.perf_counter()
    g.run_continue(steps)
    sync()
    dt = time.perf_counter() - t0
    st0 = g.stats()
    g.set_kernel_timing(1)
    g.run_continue(timing_steps)
    st = g.stats()
    g.set_kernel_timing(0)
    shape = g.step_shape()
    extra = {}
    if keep:
        extra["system"] = g
    else:
        g.close()
    n = P["x"].shape[0]
    mixed = precision == pydem.DEM_MIXED
    ntri = sum(len(m[0]) for m in meshes)
    B = algorithmic_bytes(n, st["npairs"], st["ncontacts"], st["nwall_contacts"], mixed,
                          ntri, st["nmesh_candidates"], history=params.get("tangential") != "no_history",
                          rolling_history=bool(params.get("rolling_epsd2")))
    kern_s = st["step_kernel_ms"] * 1e-3
    return {
        "throughput": n * steps / dt,
        "ms_per_step": dt / steps * 1e3,
        "kernel_ms": st["step_kernel_ms"],
        "kernel_launches": st["step_kernel_launches"],
        "bytes_per_launch": B,
        "achieved_gbs": B / kern_s / 1e9 if kern_s > 0 else None,
        "npairs": st["npairs"], "ncontacts": st["ncontacts"], "nwall": st["nwall_contacts"],
        "nmesh_candidates": st["nmesh_candidates"],
        "nrebuilds_timed": st0["nrebuilds"],
        "rebuild_ms_total_timing_pass": st["rebuild_ms_total"],
        "step_shape": {"variant": shape[0], "persist": shape[1]},
        "tuned_candidates_us": ([[c[0], c[1], round(c[2], 2)] for c in tuned[2]]
                                if tuned is not None else None),
        **extra,
    }


def sync():
    """The library's run calls return once their kernels have finished (each
    launch batch ends in a read of the control block on the handle's stream),
    so the timed region is already bracketed by completed GPU work. torch is
    not imported here: its own ROCm runtime would be a second HIP runtime in
    the process."""


def cpu_baseline(params, mats, P, walls, nsteps, meshes=()):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.OracleSystem(params, mats, P, walls, meshes=meshes)
    o.setup()
    t0 = time.perf_counter()
    o.run_continue(nsteps)
    dt = time.perf_counter() - t0
    n = P["x"].shape[0]
    o.close()
    return {"value": n * nsteps / dt, "unit": "particle-timesteps/s", "cores": 1, "kind": "port",
            "sample": f"oracle/dem_oracle.c (FP64, 1 thread) on the same {n}-sphere workload, "
                      f"{nsteps} steps after setup ({dt:.1f} s)"}


def mixed_error_vs_oracle(params, mats, P, walls):
    """Setup + one step of the mixed GPU mode against the FP64 oracle on the
    same input: per-particle force and torque error (max_i |a_i - b_i| over
    max_i |b_i|) and norm-wise error."""
    import numpy as np
    import pydem
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.OracleSystem(params, mats, P, walls)
    o.run(1)
    os_ = o.state()
    ncont = o.contact_counts()
    o.close()
    g = pydem.GpuSystem(params, mats, P, walls, precision=pydem.DEM_MIXED)
    g.run(1)
    gs = g.state()
    g.close()
    out = {"touching_pairs": int(ncont[0]), "wall_contacts": int(ncont[1])}
    for k in ("f", "torque"):
        a, b = gs[k], os_[k]
        scale = float(np.linalg.norm(b, axis=1).max())
        den = float(np.linalg.norm(b))
        out[k] = {"per_particle": float(np.linalg.norm(a - b, axis=1).max() / scale) if scale > 0 else None,
                  "normwise": float(np.linalg.norm(a - b) / den) if den > 0 else None}
    return out


def sweep_vs_alone(params, P, walls, meshes, counts, imats, g, nsteps, sample):
    """KE and contact-count error of sampled instances of the batched (mixed)
    run against the same instance run alone in FP64 for the same steps."""
    import numpy as np
    import pydem
    import workloads as wl
    st = g.state()
    bc, bw = g.instance_stats()
    rows = []
    for q in sample:
        off = int(np.sum(counts[:q]))
        n = int(counts[q])
        Pq = wl.instance_slice(P, counts, q)
        a = pydem.GpuSystem(params, imats[q], Pq, walls, meshes=meshes)
        a.run(nsteps)
        sa = a.state()
        s2 = a.stats()
        a.close()
        ke_b = 0.5 * float(Pq["rmass"] @ (st["v"][off:off + n] ** 2).sum(1))
        ke_a = 0.5 * float(Pq["rmass"] @ (sa["v"] ** 2).sum(1))
        ca = s2["ncontacts"] + s2["nwall_contacts"]
        cb = int(bc[q] + bw[q])
        rows.append({"instance": int(q), "n": n, "ke_rel_err": abs(ke_b - ke_a) / ke_a,
                     "contacts_batched": cb, "contacts_alone_fp64": ca,
                     "contact_rel_err": abs(cb - ca) / max(ca, 1)})
    return rows


MFMA_BF16_PEAK_TFLOPS = 2500.0
NET_FLOPS_PER_SAMPLE = 2 * (18 * 64 + 64 * 64 + 64 * 3) + 2 * (2 * 64 * 3 + 2 * 64 * 64 + 18 * 64)


def main_superquadric(args):
    import numpy as np
    import pydem
    import superquadrics as sq
    t_all = time.perf_counter()
    ntrain, neval = args.sq_train, args.sq_eval
    A, B = sq.overlapping_pairs(100, ntrain)
    s = pydem.SqSolver()
    s.set_pairs(A, B)
    del A, B
    r = s.solve(pydem.DEM_SQ_FRESH)
    train_solve_ms = r["solve_ms"]
    s.net_init(seed=1)
    steps = max(1, int(args.sq_epochs * ntrain / args.sq_batch))
    curve, train_ms = s.net_train(steps=steps, batch=args.sq_batch, lr=2e-3, seed=7)
    print(f"# trained {steps} steps x {args.sq_batch} in {train_ms:.0f} ms, loss {curve[0]:.3e} -> "
          f"{curve[-100:].mean():.3e}", file=sys.stderr, flush=True)
    A, B = sq.overlapping_pairs(200, neval)
    s.set_pairs(A, B)
    res = {}
    for name, start, pol in (("plain", pydem.DEM_SQ_FRESH, 0), ("warm", pydem.DEM_SQ_NETWORK, 0),
                             ("plain_polished", pydem.DEM_SQ_FRESH, 4),
                             ("warm_polished", pydem.DEM_SQ_NETWORK, 4)):
        s.solve(start, polish=pol)  # untimed: first launch
        r = s.solve(start, polish=pol)
        ms = r["solve_ms"] + r["net_ms"]
        res[name] = {"contacts_per_s": neval / (ms * 1e-3), "ms": ms, "net_ms": r["net_ms"],
                     "newton_iterations_per_contact": float(r["iters"].mean()),
                     "fallback_fraction": float(((r["flags"] & 2) != 0).mean()),
                     "failed": int((r["flags"] & 1).sum()), "point": r["point"]}
    d = np.linalg.norm(res["warm"]["point"] - res["plain"]["point"], axis=1)
    d4 = np.linalg.norm(res["warm_polished"]["point"] - res["plain_polished"]["point"], axis=1)
    for v in res.values():
        v.pop("point")
    # network training throughput against the BF16 matrix-core peak
    train_tflops = NET_FLOPS_PER_SAMPLE * args.sq_batch * steps / (train_ms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        nc = min(args.sq_cpu, neval)
        t0 = time.perf_counter()
        pyoracle.sq_contact_batch(A[:nc], B[:nc])
        dt = time.perf_counter() - t0
        cpu = {"value": nc / dt, "unit": "contacts/s", "cores": 1, "kind": "port",
               "sample": f"oracle/sq_oracle.c plain (homotopy) solve of {nc} of the same pairs, "
                         f"1 thread ({dt:.1f} s)"}
    warm = res["warm"]
    out = {
        "metric": "superquadric contacts/s (FP64 Newton, network warm start)",
        "value": warm["contacts_per_s"],
        "unit": "contacts/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": warm["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (solve) + bf16/f32 (network)",
        "data": f"synthetic (seeded overlapping pose pairs: {ntrain} training, {neval} evaluation)",
        "config": {"workload": f"{neval} superquadric pairs, half-axes U(1,3) mm, blockiness U(2,8), "
                               "random orientations; network 18-64-64-3 trained on "
                               f"{ntrain} plain-solved pairs ({steps} Adam steps, batch {args.sq_batch})",
                   "n_pairs": neval},
        "modes": res,
        "iteration_ratio_plain_over_warm": res["plain"]["newton_iterations_per_contact"]
        / warm["newton_iterations_per_contact"],
        "agreement_m": {"median": float(np.median(d)), "q99": float(np.quantile(d, 0.99)),
                        "within_1e-9": float((d <= 1e-9).mean()),
                        "polished_q99": float(np.quantile(d4, 0.99)),
                        "polished_within_1e-9": float((d4 <= 1e-9).mean())},
        "training": {"steps": steps, "batch": args.sq_batch, "ms": train_ms,
                     "loss_first": float(curve[0]), "loss_last": float(curve[-100:].mean()),
                     "plain_solve_ms_of_training_set": train_solve_ms},
        "roofline": {"kernel": "sqk::k_mlp_grad + k_mlp_adam (network training step)", "bound": "mfma",
                     "achieved": train_tflops, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": train_tflops / MFMA_BF16_PEAK_TFLOPS, "traffic": None},
        "wall_s": time.perf_counter() - t_all,
    }
    if cpu:
        out["cpu_baseline"] = cpu
    s.close()
    print(json.dumps(out), flush=True)


def _sq_ke(g):
    st, pr = g.state(), g.props()
    return float(0.5 * (pr["mass"][:, None] * st["v"] ** 2).sum() + 0.5 * (st["omega"] * st["angmom"]).sum())


def main_sqsettle(args):
    """Superquadric settling: the superquadric DEM step (FP64) on n particles
    (half-axes U(1,3) mm, blockiness U(2,8), random orientations) started
    agitated (velocities N(0, v0), spins N(0, 30 rad/s)) above a floor in a
    0.2 x 0.2 m periodic column; timed once with the reference's homotopy for
    new contacts and once with the network warm start."""
    import numpy as np
    import pydem
    import superquadrics as sq
    import workloads as wl
    t_all = time.perf_counter()
    n = args.sq_n

    def make(nn, base):
        params, mats, P, walls = wl.superquadric_settling(11, nn, base=base, gap=0.0003)
        rng = np.random.default_rng(12)
        P["v"] = rng.normal(0, args.sq_v0, (nn, 3))
        P["omega"] = rng.normal(0, 30.0, (nn, 3))
        return params, mats, P, walls

    params, mats, P, walls = make(n, (0.2, 0.2))
    # warm-start network, trained on plain-solved random overlapping pairs of the same shape ranges
    solver = pydem.SqSolver()
    A, B = sq.overlapping_pairs(100, args.sq_ntrain)
    solver.set_pairs(A, B)
    del A, B
    solver.solve(pydem.DEM_SQ_FRESH)
    solver.net_init(seed=1)
    tsteps = max(1, int(2.0 * args.sq_ntrain / args.sq_batch))
    curve, train_ms = solver.net_train(steps=tsteps, batch=args.sq_batch, lr=2e-3, seed=7)
    print(f"# network: {tsteps} Adam steps in {train_ms:.0f} ms", file=sys.stderr, flush=True)
    res = {}
    for mode in ("plain", "warm"):
        g = pydem.SqDemSystem(params, mats, P, walls)
        if mode == "warm":
            g.set_warm_start(solver)
        g.run(args.warmup)
        s0 = g.stats()
        ms = 0.0
        done = 0
        while done < args.steps:
            k = min(50, args.steps - done)
            g.run(k)
            done += k
            ms += g.stats()["run_ms"]
            print(f"# {mode}: {done}/{args.steps} steps, {ms:.0f} ms", file=sys.stderr, flush=True)
        s1 = g.stats()
        fresh = s1["fresh"] - s0["fresh"]
        res[mode] = {"ms_per_step": ms / args.steps,
                     "particle_steps_per_s": n * args.steps / (ms * 1e-3),
                     "contact_detections_per_s": (s1["checks"] - s0["checks"]) / (ms * 1e-3),
                     "touching_pairs_last_step": s1["pair_touch"], "wall_contacts_last_step": s1["wall_touch"],
                     "fresh_solves": fresh,
                     "newton_iterations_per_fresh_solve": (s1["fresh_iters"] - s0["fresh_iters"]) / max(fresh, 1),
                     "newton_iterations_per_detection": (s1["newton"] - s0["newton"])
                     / max(s1["checks"] - s0["checks"], 1),
                     "fallbacks": s1["fallbacks"] - s0["fallbacks"], "rebuilds": s1["rebuilds"] - s0["rebuilds"],
                     "kinetic_energy": _sq_ke(g)}
        print(f"# {mode}: {res[mode]['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
        g.close()
    solver.close()
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        nc = args.sq_cpu_n
        cp, cm, cP, cw = make(nc, (0.04, 0.04))
        o = pyoracle.SqsSystem(cp, pyoracle.materials(cm), cP, cw)
        o.setup()
        o.run(args.warmup)
        t0 = time.perf_counter()
        o.run(args.sq_cpu_steps)
        dt = time.perf_counter() - t0
        o.close()
        cpu = {"value": nc * args.sq_cpu_steps / dt, "unit": "particle-timesteps/s", "cores": 1, "kind": "port",
               "sample": f"oracle/sqdem_oracle.c, {nc} particles of the same generator in a 0.04 x 0.04 m column, "
                         f"{args.sq_cpu_steps} steps after {args.warmup}, 1 thread ({dt:.1f} s)"}
    warm = res["warm"]
    out = {
        "metric": "superquadric particle-timesteps/s (FP64, network warm start)",
        "value": warm["particle_steps_per_s"], "unit": "particle-timesteps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": warm["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (+ bf16/f32 network for new contacts)",
        "data": "synthetic (seeded superquadrics on a layer lattice, agitated start)",
        "config": {"workload": f"superquadric settling, {n} particles, half-axes U(1,3) mm, blockiness U(2,8), "
                               f"random orientations, 0.2 x 0.2 m periodic column, floor, v0 {args.sq_v0} m/s, "
                               "Hertz/history, dt 1e-5", "n": n},
        "modes": res,
        "iteration_ratio_plain_over_warm": res["plain"]["newton_iterations_per_fresh_solve"]
        / max(warm["newton_iterations_per_fresh_solve"], 1e-30),
        "network": {"train_pairs": args.sq_ntrain, "steps": tsteps, "train_ms": train_ms},
        "wall_s": time.perf_counter() - t_all,
    }
    if cpu:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


def main_copy(args):
    """Calibration for hardware counters: a plain device-to-device copy of a
    known size (torch's copy kernel), repeated; its FETCH_SIZE / WRITE_SIZE per
    launch against the bytes it must move gives the counters' scale."""
    import torch
    nbytes = args.copy_mib << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 20
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"metric": "device copy bandwidth (counter calibration)", "value": 2 * nbytes / dt / 1e9,
                      "unit": "GB/s", "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                      "launches": reps + 3}), flush=True)


# pair_style gran variants for the settling and packed workloads:
# (model switches, extra material properties, description)
CONTACT_MODELS = {
    "hertz": ({}, {}, "Hertz/history"),
    "hooke-nohistory-sjkr": ({"normal": "hooke", "tangential": "no_history", "cohesion": "sjkr"},
                             {"characteristic_velocity": 1.0, "cohesion_energy_density": [[3.0e4]]},
                             "hooke + tangential no_history + sjkr cohesion (30 kJ/m^3)"),
    "hertz-stiffness": ({"normal": "hertz/stiffness"},
                        {"kn": [[4.0e6]], "kt": [[5.0e6]], "gamman": [[2.0e5]], "gammat": [[1.0e5]]},
                        "hertz/stiffness/history (kn 4e6, kt 5e6, gamman 2e5, gammat 1e5)"),
    "hertz-epsd2": ({"rolling_cdt": 0, "rolling_epsd2": 1}, {},
                    "Hertz/history + rolling_friction epsd2 (rolling-torque history)"),
}


def with_contact_model(name, params, mats):
    model, props, desc = CONTACT_MODELS[name]
    return dict(params, **model), dict(mats, **props), desc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200,
                    help="untimed steps before the timed ones (the tile shape is tuned beforehand "
                         "on a separate system of the same workload)")
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--timing-steps", type=int, default=500)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="settling: skip the packed-bed figure and the error against the oracle")
    ap.add_argument("--packed-steps", type=int, default=500)
    ap.add_argument("--modes", default=None)
    ap.add_argument("--workload", choices=("settling", "packed", "drum", "sweep", "superquadric", "sqsettle",
                                           "copy"),
                    default="settling")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--sq-train", type=int, default=10_000_000)
    ap.add_argument("--sq-eval", type=int, default=200_000)
    ap.add_argument("--sq-batch", type=int, default=8192)
    ap.add_argument("--sq-epochs", type=float, default=2.0)
    ap.add_argument("--sq-cpu", type=int, default=20_000)
    ap.add_argument("--sq-n", type=int, default=200_000)
    ap.add_argument("--sq-v0", type=float, default=0.3)
    ap.add_argument("--sq-ntrain", type=int, default=2_000_000)
    ap.add_argument("--sq-cpu-n", type=int, default=2000)
    ap.add_argument("--sq-cpu-steps", type=int, default=3000)
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--contact-model", choices=sorted(CONTACT_MODELS), default="hertz",
                    help="settling / packed: the pair and wall contact model")
    ap.add_argument("--sample", type=int, default=8,
                    help="sweep: instances rerun alone in FP64 for the error report")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: runs on one GPU only; spatial-slab sharding over several GPUs "
              "is not built, so --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "superquadric":
        main_superquadric(args)
        return
    if args.workload == "sqsettle":
        return main_sqsettle(args)
    if args.workload == "copy":
        main_copy(args)
        return
    import pydem
    import workloads as wl
    meshes = ()
    inst = imats = None
    tune = graphs = True
    if args.workload == "sweep":
        params, mats, P, walls, meshes, inst, imats = wl.heap_sweep(seed=0, ninst=args.instances)
        n = P["x"].shape[0]
        workload = (f"batched heap-formation sweep: {len(inst)} instances, {n} spheres in all "
                    f"(N per instance {int(inst.min())}-{int(inst.max())}, log-uniform 2k-50k), "
                    "r 1.5 mm, 0.1 m mesh floor with open sides, per-instance mu/e/mu_r, "
                    "dt 1e-5 s")
        data = "synthetic (seeded jittered lattices, per-instance seeded materials)"
        modes = args.modes or "mixed"
    elif args.workload == "packed":
        params, mats, P, walls = wl.packed_1m(seed=0, n=args.n)
        params, mats, law = with_contact_model(args.contact_model, params, mats)
        n = P["x"].shape[0]
        workload = (f"{n}-sphere polydisperse {law} + CDT rolling, compressed lattice "
                    "(about three touching neighbors per sphere), periodic xy, zplane floor, "
                    "dt 5e-6 s, skin 1 mm")
        data = f"synthetic (seeded compressed lattice, {n} polydisperse spheres)"
        modes = args.modes or "fp64,mixed"
    elif args.workload == "drum":
        params, mats, P, walls, meshes = wl.drum(seed=0)
        n = P["x"].shape[0]
        workload = (f"rotating drum, {n} spheres r 1 mm +-10%, 2,176-triangle mesh/surface "
                    "cylinder R 0.15 m L 0.1 m at 30 rpm (move/mesh rotate), mu_pp 0.5 mu_pw 0.4, "
                    "dt 5e-6 s")
        data = f"synthetic (seeded jittered lattice, {n} spheres, generated drum mesh)"
        modes = args.modes or "mixed,fp64"
    else:
        params, mats, P, walls = wl.settling_1m(seed=0, n=args.n)
        params, mats, law = with_contact_model(args.contact_model, params, mats)
        n = args.n
        workload = (f"{n}-sphere polydisperse {law} + CDT rolling settling, "
                    "box 0.3x0.3x0.6 m periodic xy, zplane floor, dt 5e-6 s, skin 1 mm")
        data = "synthetic (seeded jittered lattice, 1M polydisperse spheres)"
        modes = args.modes or "fp64,mixed"
    res = {}
    sweep_rows = None
    for mode in modes.split(","):
        prec = pydem.DEM_MIXED if mode == "mixed" else pydem.DEM_FP64
        res[mode] = run_mode(params, mats, P, walls, prec, args.steps, args.warmup,
                             args.timing_steps, meshes, inst, imats, tune, graphs,
                             keep=inst is not None)
        if inst is not None:
            g = res[mode].pop("system")
            import numpy as np
            order = np.argsort(inst)
            pick = order[np.linspace(0, len(order) - 1, min(args.sample, len(order))).astype(int)]
            sweep_rows = sweep_vs_alone(params, P, walls, meshes, inst, imats, g,
                                        args.warmup + args.steps + args.timing_steps, pick)
            g.close()
        print(f"# {mode}: {res[mode]}", file=sys.stderr, flush=True)
    head = res.get("mixed") or res.get("fp64")
    head_mode = "mixed" if "mixed" in res else "fp64"
    metric = "particle-timesteps/s at 1M spheres, mixed vs FP64 GPU; HBM GB/s of pair kernel"
    if args.workload == "sweep":
        metric = "aggregate particle-timesteps/s, batched heap-formation sweep"
    elif args.workload == "drum":
        metric = "particle-timesteps/s, rotating drum with mesh walls"
    elif args.workload == "packed":
        metric = "particle-timesteps/s at 1M spheres in a packed bed, mixed vs FP64 GPU"
    out = {
        "metric": metric,
        "value": head["throughput"],
        "unit": "particle-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (mixed)" if head_mode == "mixed" else "f64",
        "data": data,
        "config": {"workload": workload, "n_particles": n, "precision": head_mode},
        "modes": {m: {"value": r["throughput"], "ms_per_step": r["ms_per_step"],
                      "kernel_ms": r["kernel_ms"], "achieved_gbs": r["achieved_gbs"],
                      "npairs": r["npairs"], "ncontacts": r["ncontacts"], "nwall": r["nwall"],
                      "nmesh_candidates": r["nmesh_candidates"]}
                  for m, r in res.items()},
    }
    if "mixed" in res and "fp64" in res:
        out["mixed_over_fp64"] = res["mixed"]["throughput"] / res["fp64"]["throughput"]
    out["step_shape"] = head["step_shape"]
    if sweep_rows is not None:
        out["instances_vs_alone_fp64"] = {
            "sample": sweep_rows,
            "max_ke_rel_err": max(r["ke_rel_err"] for r in sweep_rows),
            "max_contact_rel_err": max(r["contact_rel_err"] for r in sweep_rows),
        }
    out["roofline"] = {
        "kernel": "dem::k_step (fused pair + walls + gravity + nve/sphere)"
                  + (" with mesh walls" if meshes else ""),
        "bound": "hbm",
        "achieved": head["achieved_gbs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": (head["achieved_gbs"] / HBM_PEAK_GBS) if head["achieved_gbs"] else None,
        "traffic": (measured_traffic(PROFILED.get(args.workload), head_mode)
                    if args.contact_model == "hertz" else None),
        "traffic_source": (traffic_source(PROFILED[args.workload])
                           if args.contact_model == "hertz" and args.workload in PROFILED else None),
        "bytes_per_launch": head["bytes_per_launch"],
        "kernel_ms": head["kernel_ms"],
    }
    if args.workload == "settling" and args.contact_model == "hertz" and not args.no_extras:
        # the same 1M spheres from a compressed lattice (about three touching
        # neighbors each): the contact-bearing regime, timed over fewer steps
        pp, pm, pP, pw = wl.packed_1m(seed=0, n=n)
        packed = {}
        for mode in modes.split(","):
            prec = pydem.DEM_MIXED if mode == "mixed" else pydem.DEM_FP64
            r = run_mode(pp, pm, pP, pw, prec, args.packed_steps, 100, 200, (), None, None, True, True)
            packed[mode] = {"value": r["throughput"], "ms_per_step": r["ms_per_step"],
                            "kernel_ms": r["kernel_ms"], "achieved_gbs": r["achieved_gbs"],
                            "frac": r["achieved_gbs"] / HBM_PEAK_GBS if r["achieved_gbs"] else None,
                            "traffic": measured_traffic("packed_1M", mode),
                            "traffic_source": traffic_source("packed_1M"),
                            "npairs": r["npairs"], "ncontacts": r["ncontacts"], "nwall": r["nwall"],
                            "rebuilds_in_timed_steps": r["nrebuilds_timed"], "step_shape": r["step_shape"]}
            print(f"# packed {mode}: {packed[mode]}", file=sys.stderr, flush=True)
        if "mixed" in packed and "fp64" in packed:
            packed["mixed_over_fp64"] = packed["mixed"]["value"] / packed["fp64"]["value"]
        packed["workload"] = (f"{n}-sphere packed bed (compressed polydisperse lattice), Hertz/history + CDT, "
                              f"{args.packed_steps} steps after 100")
        out["packed_bed"] = packed
        # each result's error against the FP64 oracle, at 1M on both states
        err = {"settling_1m": mixed_error_vs_oracle(params, mats, P, walls),
               "packed_1m": mixed_error_vs_oracle(pp, pm, pP, pw),
               "definition": "setup + one step, mixed GPU mode vs the FP64 CPU oracle on the same input; "
                             "per_particle = max_i |a_i - b_i| / max_i |b_i|"}
        out["mixed_error_vs_fp64_oracle"] = err
        # the contact-bearing state carries the meaningful figure (the dilute
        # lattice has no contacts: its forces are gravity alone)
        out["mixed_force_rel_err"] = err["packed_1m"]["f"]["per_particle"]
        out["mixed_torque_rel_err"] = err["packed_1m"]["torque"]["per_particle"]
    if not args.no_cpu_baseline:
        if inst is not None:
            # the oracle on the median-size instance alone (instances are independent,
            # so the sweep's CPU rate is this per-instance rate)
            import numpy as np
            q = int(np.argsort(inst)[len(inst) // 2])
            out["cpu_baseline"] = cpu_baseline(params, imats[q], wl.instance_slice(P, inst, q),
                                               walls, args.cpu_steps, meshes)
        else:
            out["cpu_baseline"] = cpu_baseline(params, mats, P, walls, args.cpu_steps, meshes)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
