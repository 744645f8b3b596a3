"""Benchmark of the DEM step on one MI355X.

Workload: the 1M-sphere polydisperse Hertz/history settling case (radii
U(1.0, 1.5) mm, CDT rolling friction 0.1, box 0.3 x 0.3 x 0.6 m periodic in x
and y, zplane floor, jittered 3.3 mm lattice, dt = 5e-6 s, skin 1 mm), run in
the FP64 mode and in the mixed FP32/FP64 mode. A step is one full timestep of
all particles (pair forces with history, walls, gravity, nve/sphere, neighbor
decide and any rebuild it triggers). Inputs are resident in HBM before the
timed region.

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


def algorithmic_bytes(n, npairs, ncontacts, nwall, mixed, ntri=0, ncand=0):
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
                  for the history)."""
    sh = 12 if mixed else 24
    per_particle = 64 + (16 if mixed else 32) + 24 + 4 + 48 + (12 if mixed else 24)
    B = n * per_particle + 4 * npairs + 2 * sh * (ncontacts + nwall)
    if ntri:
        B += 4 * n + 4 * ncand + TRI_RECORD_BYTES * ntri + 4 * nwall
    return B


def run_mode(params, mats, P, walls, precision, steps, warmup, timing_steps, meshes=()):
    import pydem
    g = pydem.GpuSystem(params, mats, P, walls, precision=precision, meshes=meshes)
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
    g.close()
    n = P["x"].shape[0]
    mixed = precision == pydem.DEM_MIXED
    ntri = sum(len(m[0]) for m in meshes)
    B = algorithmic_bytes(n, st["npairs"], st["ncontacts"], st["nwall_contacts"], mixed,
                          ntri, st["nmesh_candidates"])
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
    }


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--timing-steps", type=int, default=500)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--modes", default=None)
    ap.add_argument("--workload", choices=("settling", "drum"), default="settling")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: runs on one GPU only; spatial-slab sharding over several GPUs "
              "is not built, so --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import pydem
    import workloads as wl
    meshes = ()
    if args.workload == "drum":
        params, mats, P, walls, meshes = wl.drum(seed=0)
        n = P["x"].shape[0]
        workload = (f"rotating drum, {n} spheres r 1 mm +-10%, 2,176-triangle mesh/surface "
                    "cylinder R 0.15 m L 0.1 m at 30 rpm (move/mesh rotate), mu_pp 0.5 mu_pw 0.4, "
                    "dt 5e-6 s")
        data = f"synthetic (seeded jittered lattice, {n} spheres, generated drum mesh)"
        modes = args.modes or "mixed,fp64"
    else:
        params, mats, P, walls = wl.settling_1m(seed=0, n=args.n)
        n = args.n
        workload = (f"{n}-sphere polydisperse Hertz/history + CDT rolling settling, "
                    "box 0.3x0.3x0.6 m periodic xy, zplane floor, dt 5e-6 s, skin 1 mm")
        data = "synthetic (seeded jittered lattice, 1M polydisperse spheres)"
        modes = args.modes or "fp64,mixed"
    res = {}
    for mode in modes.split(","):
        prec = pydem.DEM_MIXED if mode == "mixed" else pydem.DEM_FP64
        res[mode] = run_mode(params, mats, P, walls, prec, args.steps, args.warmup,
                             args.timing_steps, meshes)
        print(f"# {mode}: {res[mode]}", file=sys.stderr, flush=True)
    head = res.get("mixed") or res.get("fp64")
    head_mode = "mixed" if "mixed" in res else "fp64"
    out = {
        "metric": "particle-timesteps/s at 1M spheres, mixed vs FP64 GPU; HBM GB/s of pair kernel",
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
    out["roofline"] = {
        "kernel": "dem::k_step (fused pair + walls + gravity + nve/sphere)"
                  + (" with mesh walls" if meshes else ""),
        "bound": "hbm",
        "achieved": head["achieved_gbs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": (head["achieved_gbs"] / HBM_PEAK_GBS) if head["achieved_gbs"] else None,
        "traffic": None,
        "bytes_per_launch": head["bytes_per_launch"],
        "kernel_ms": head["kernel_ms"],
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(params, mats, P, walls, args.cpu_steps, meshes)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
