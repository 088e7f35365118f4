"""Throughput of the D2Q9 time step on one MI355X.

A "step" is one lattice-Boltzmann time step (one launch of the fused step
kernel) over the whole workload.  Default workload: the 8192 x 8192 random
porous medium (discs r in [4, 16], 30 % solid, BGK tau 0.8, Zou-He pressure
inlet 1.001 / outlet 1.000), populations resident in HBM before the timed
region.  The headline value is the MLUPS of the FP32 shifted-population
path; the FP64 (drop-in) and FP16 paths are timed in the same run and
reported beside it.  The three lattices then run on to the workload's full
length (20,000 steps) and the FP32/FP16 velocity and density errors
against the FP64 run, and their FP64-reduced mass difference, are stated.

The ensemble line advances 512 independent porous lattices (rows and
columns in {128, ..., 1024}, porosity in [0.5, 0.9], discs r in [3, 12])
with one step-kernel launch per step, HIP-graph replay and the autotuned
tile shape, in FP32; its per-sample Darcy permeabilities (FP64 GPU
reductions) are checked against an FP64 ensemble.

MLUPS = wet nodes * steps / time / 1e6.  Algorithmic HBM bytes per step =
wet nodes * 18 * sizeof(storage) + NR*NC/8 (bit-packed wet mask); the
roofline fraction divides the achieved rate by MI355X's 8 TB/s; the rate of
a plain 16-byte-per-lane copy kernel over one FP32 population buffer's
volume is measured in the same run and reported beside it as the practical
ceiling of a stream (copy_ceiling, frac_of_copy).  The step
kernel is the only kernel in the timed region, so its average duration is
the event-timed region divided by the number of steps.

The CPU baseline is the repo's CPU oracle in the reference's step structure
(list-driven sweeps, three population tensors, the by-value rho / u copies
of every boundary call) with OpenMP on this host's physical cores, timed in
a child process on bounded samples of the same workload (the headline) and
of the 129^2 cavity and the 512 x 2048 channel.

Other workloads: open8192 (no obstacles), micro8192 (the porous geometry
from random populations w_d (1 + 1e-3 U(-1, 1)), seed 1: the step's time
does not depend on the flow state), cavity_trt4096, porous<n>, channel,
cavity<n>.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload porous8192]
"""
import argparse
import ctypes as C
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "latticeboltzmannmethod-fonnesu-grassi-guffanti_amd"))

import numpy as np  # noqa: E402

import lbm_amd as L  # noqa: E402
import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
STORAGE_NAMES = {L.F64: "f64", L.F32_SHIFTED: "f32", L.F16_SHIFTED: "f16"}
# timed variants: name -> (storage, lbm_desc options)
# f16: stochastic rounding (the default); f16rn: rounded to nearest
VARIANTS = {"f32": (L.F32_SHIFTED, 0), "f64": (L.F64, 0), "f16": (L.F16_SHIFTED, 0),
            "f16rn": (L.F16_SHIFTED, L.OPT_F16_ROUND_TO_NEAREST)}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on an arbitrary stream (the engine's own stream)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def time_ms(self, stream, fn):
        a, b = C.c_void_p(), C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(a)) == 0 and self.hip.hipEventCreate(C.byref(b)) == 0
        assert self.hip.hipEventRecord(a, stream) == 0
        fn()
        assert self.hip.hipEventRecord(b, stream) == 0
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        self.hip.hipEventDestroy(a)
        self.hip.hipEventDestroy(b)
        return ms.value


def copy_ceiling(nbytes, reps=20):
    """The practical HBM ceiling of a plain stream over nbytes: the library's
    16-byte-per-lane copy kernel (lbm_copy_bandwidth), read + write bytes
    per second."""
    return dict(kind="copy kernel, 16-byte lanes (lbm_copy_bandwidth)", bytes=int(nbytes),
                gbs=L.copy_bandwidth(nbytes, reps))


def access_patterns(nr, nc, reps=20):
    """The step's possible access patterns without arithmetic, on a dense
    nr x nc periodic lattice, per element size (lbm_pattern_bandwidth): the
    two-buffer pull the kernels use, the AA in-place pattern (even/odd
    alternating, and each step alone), and the pull in a blocked layout."""
    out = {}
    for name, size in (("f16", 2), ("f32", 4), ("f64", 8)):
        out[name] = {L.PATTERN_NAMES[p]: round(L.pattern_bandwidth(p, size, nr, nc, reps), 1)
                     for p in sorted(L.PATTERN_NAMES)}
    return out


def build_workload(name):
    if name == "porous8192":
        return W.porous(8192)
    if name.startswith("micro"):
        # kernel micro-bench: the porous geometry with random populations
        # w_d (1 + 1e-3 U(-1, 1)), seed 1 -- the step's bytes and time do not
        # depend on the flow state
        wl = W.porous(int(name[5:]))
        wl.name = f"micro{wl.nr}x{wl.nc}"
        wl.random_state = True
        return wl
    if name.startswith("open"):
        # the porous set-up without obstacles (pressure-driven channel)
        wl = W.porous(int(name[4:]), solid=0.0)
        wl.name = f"open{wl.nr}x{wl.nc}"
        return wl
    if name.startswith("porous"):
        return W.porous(int(name[6:]))
    if name == "channel":
        return W.channel()
    if name.startswith("cavity_trt"):
        return W.cavity_trt(int(name[10:]) if len(name) > 10 else 4096)
    if name.startswith("cavity"):
        return W.cavity(int(name[6:]) if len(name) > 6 else 129)
    raise SystemExit(f"unknown workload {name}")


def measured_traffic(key, kernel="step_kernel"):
    """HBM bytes per launch of a step kernel for a workload/storage/tile-shape
    (/occupancy) key, from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    summarised in profiles/ (calibrated against a plain copy kernel); None
    when that kernel was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    e = table.get(kernel, {}).get(key)
    return None if e is None else e["read_bytes"] + e["write_bytes"]


_RANDOM_STATE = {}


def random_state(wl):
    """Populations w_d (1 + 1e-3 U(-1, 1)) on every node, seed 1 (the
    micro-bench state), in the reference layout."""
    if wl.name not in _RANDOM_STATE:
        rng = np.random.default_rng(1)
        w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
        p = rng.uniform(-1e-3, 1e-3, size=(9, wl.nc, wl.nr))
        p += 1.0
        p *= w[:, None, None]
        _RANDOM_STATE.clear()
        _RANDOM_STATE[wl.name] = p
    return _RANDOM_STATE[wl.name]


def time_storage(wl, name, steps, warmup, ev):
    storage, options = VARIANTS[name]
    lat = wl.make_lattice(storage, options=options)
    if getattr(wl, "random_state", False):
        lat.upload_populations(random_state(wl))
    lat.step(warmup, L.STEP_CHECK)
    segments, _ = lat.autotune(steps=5)
    lat.step(2 * 32, L.STEP_CHECK)              # capture and warm the step graphs
    lat.synchronize()
    stream = lat.stream_handle()
    t0 = time.perf_counter()
    ms = ev.time_ms(stream, lambda: lat.step(steps))
    lat.synchronize()
    wall = time.perf_counter() - t0
    lat.step(0, L.STEP_CHECK)
    nwet = wl.wet_count()
    per_step_s = ms / 1e3 / steps
    mlups = nwet / per_step_s / 1e6
    alg = W.algorithmic_bytes(wl, storage)
    achieved = alg / per_step_s / 1e9
    return lat, dict(storage=name, mlups=mlups, mlups_all_nodes=wl.nr * wl.nc / per_step_s / 1e6,
                     ms_per_step=ms / steps, wall_s=wall, alg_bytes_per_step=alg, achieved_gbs=achieved,
                     frac=achieved / HBM_PEAK_GBS, segments=segments, occupancy=lat.occupancy(),
                     layout="blocked" if lat.layout() == L.LAYOUT_BLOCKED else "planes",
                     traffic_bytes_per_step=measured_traffic(f"{wl.name}/{name}/R{segments}/o{lat.occupancy()}"))


def physical_cores():
    """Physical cores among the CPUs this process may run on (lscpu's
    CORE/SOCKET pairs), capped at 16, the CPU share of one GPU on the
    benchmark machines."""
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    cores = set()
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=30).stdout
        allowed = set(cpus)
        for line in out.splitlines():
            if line.startswith("#") or not line.strip():
                continue
            cpu, core, sock = line.split(",")[:3]
            if int(cpu) in allowed:
                cores.add((sock, core))
    except (OSError, ValueError, subprocess.SubprocessError):
        cores = set()
    n = len(cores) if cores else len(cpus)
    return max(1, min(n, 16)), len(cpus)


def cpu_baseline(wl, workload, seconds=8.0, extra=("cavity129", "channel")):
    """The reference's step structure on the host cores: the oracle with the
    reference's by-value rho / u copies in every boundary call, run in a
    child process with OMP_NUM_THREADS = physical cores, OMP_PROC_BIND =
    close and OMP_PLACES = cores set before its OpenMP runtime starts, on
    bounded samples (about `seconds` each) of this workload and of the
    129^2 cavity and the 512 x 2048 channel."""
    cores, cpus = physical_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    names = [workload] + [n for n in extra if n != workload]
    script = os.path.join(ROOT, "oracle", "cpu_baseline.py")
    r = subprocess.run([sys.executable, script, "--configs", ",".join(names), "--seconds", str(seconds)],
                       env=env, capture_output=True, text=True, timeout=60 + 8 * seconds * len(names))
    sys.stderr.write(r.stderr)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed ({r.returncode})")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    head = res["configs"][names[0]]
    return dict(value=round(head["mlups"], 3), unit="MLUPS", cores=cores, kind="port",
                sample=f"{head['steps']} steps of {wl.name} ({head['wet_nodes']} wet nodes) in {head['seconds']} s: "
                       f"the oracle (reference-structured list sweeps, three population tensors, by-value rho/u "
                       f"copies per boundary call as in OMPPolicy.hpp:50-72,139), OpenMP {cores} threads on "
                       f"{cores} physical cores of {cpus} CPUs, OMP_PROC_BIND=close, OMP_PLACES=cores",
                configs={k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                         for k, v in res["configs"].items()})


def run_to(lat, total_steps, name="", chunk=10000):
    """Steps a lattice until it has done total_steps - 1 steps, then one
    more step that also writes the rho/u fields; reports progress every
    `chunk` steps."""
    t0 = time.perf_counter()
    while total_steps - lat.time > 1:
        n = min(chunk, total_steps - lat.time - 1)
        lat.step(n, L.STEP_CHECK)
        log(f"    {name}: step {lat.time} of {total_steps} ({time.perf_counter() - t0:.1f}s)")
    lat.step(1, L.STEP_MACRO | L.STEP_CHECK)


def stated_errors(wl, lat_ref, lats, total_steps):
    """L2 relative errors of the fluid-node velocity and density (rho - 1)
    fields of each shifted-storage lattice against the FP64 one after
    total_steps steps, and the relative difference of the FP64-reduced
    wet mass."""
    run_to(lat_ref, total_steps, "f64")
    r0, u0 = lat_ref.macro()
    m0 = lat_ref.reduce()[0]
    m = wl.node_class == L.FLUID
    un = max(np.sum(u0[:, m] ** 2), 1e-300)
    rn = max(np.sum((r0[m] - 1.0) ** 2), 1e-300)
    out = {}
    for s, lat in lats.items():
        run_to(lat, total_steps, s)
        r1, u1 = lat.macro()
        out[s] = {
            "velocity_l2_rel": float(np.sqrt(np.sum((u1[:, m] - u0[:, m]) ** 2) / un)),
            "density_l2_rel": float(np.sqrt(np.sum((r1[m] - r0[m]) ** 2) / rn)),
            "mass_rel_diff": float((lat.reduce()[0] - m0) / m0),
        }
    return {"steps": total_steps, "vs": "f64 GPU path (oracle-pinned)", "by_storage": out}


def bench_cavity_trt(args, ev):
    """The lid-driven TRT cavity at scale (4096 x 4096, Re 1000, magic 3/16):
    MLUPS and roofline fraction of the FP64, FP32 and FP16 storage formats
    over the same step count as the headline, then the FP32/FP16 errors
    against the FP64 run after `args.c4_error_steps` steps (the
    configuration's 200,000 by default)."""
    wl = W.cavity_trt(4096)
    log(f"cavity_trt4096: {wl.wet_count()} wet nodes, TRT tau+ {wl.tau_even:.4f}, tau- {wl.tau_odd:.4f}")
    results, lats = {}, {}
    for s in ("f64", "f32", "f16"):
        lat, r = time_storage(wl, s, args.steps, args.warmup, ev)
        results[s] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        lats[s] = lat
        log(f"  cavity_trt4096 {s}: {r['mlups']:.1f} MLUPS, {r['ms_per_step']:.4f} ms/step, "
            f"{r['achieved_gbs']:.0f} GB/s = {100 * r['frac']:.1f}% of 8 TB/s "
            f"(launch {r['segments']}, occupancy {r['occupancy'] or 'full'})")
    errors = None
    if args.c4_error_steps > 0:
        t1 = time.perf_counter()
        errors = stated_errors(wl, lats["f64"], {s: lat for s, lat in lats.items() if s != "f64"},
                               args.c4_error_steps)
        for name, e in errors["by_storage"].items():
            log(f"  cavity_trt4096 {name} vs f64 after {args.c4_error_steps} steps: velocity L2 rel "
                f"{e['velocity_l2_rel']:.3e}, density L2 rel {e['density_l2_rel']:.3e}, mass rel diff "
                f"{e['mass_rel_diff']:.3e} ({time.perf_counter() - t1:.1f}s)")
    for lat in lats.values():
        lat.close()
    return dict(workload=wl.name, wet_nodes=wl.wet_count(), collision="TRT", re=wl.info["re"],
                u_lid=wl.info["u_lid"], storages=results, error_vs_f64=errors)


def bench_open(args, ev):
    """The porous workload's 8192 x 8192 channel with no obstacles (67.1M wet
    nodes, same pressure-driven BGK set-up): MLUPS and roofline fraction of
    the FP64, FP32 and FP16 storage formats over the headline's step count;
    timing only (its errors are not stated)."""
    wl = build_workload("open8192")
    log(f"open8192: {wl.wet_count()} wet nodes")
    results = {}
    for s in ("f64", "f32", "f16"):
        lat, r = time_storage(wl, s, args.steps, args.warmup, ev)
        results[s] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        lat.close()
        log(f"  open8192 {s}: {r['mlups']:.1f} MLUPS, {r['ms_per_step']:.4f} ms/step, "
            f"{r['achieved_gbs']:.0f} GB/s = {100 * r['frac']:.1f}% of 8 TB/s "
            f"(launch {r['segments']}, occupancy {r['occupancy'] or 'full'}, {r['layout']})")
    return dict(workload=wl.name, wet_nodes=wl.wet_count(), collision="BGK", storages=results)


def bench_ensemble(args, ev):
    """Many independent porous lattices of uneven sizes advanced by one
    persistent launch per step (lbm_batch), HIP-graph step loop, tile shape
    from the autotuner."""
    t0 = time.perf_counter()
    wls = W.ensemble(n_samples=args.ensemble_samples)
    wet = sum(w.wet_count() for w in wls)
    alg = sum(W.algorithmic_bytes(w, L.F32_SHIFTED) for w in wls)
    lats = [w.make_lattice(L.F32_SHIFTED) for w in wls]
    log(f"ensemble: {len(wls)} samples, {wet} wet nodes (built in {time.perf_counter() - t0:.1f}s)")
    b = L.Batch(lats)
    b.step(args.warmup, L.STEP_CHECK)
    seg, blocks, tuned_ms = b.autotune(steps=5)
    b.step(2 * 32, L.STEP_CHECK)                 # capture and warm the graphs
    b.synchronize()
    ms = ev.time_ms(b.stream_handle(), lambda: b.step(args.steps))
    b.step(0, L.STEP_CHECK)
    per_step_s = ms / 1e3 / args.steps
    res = dict(samples=len(wls), wet_nodes=wet, storage="f32", mlups=wet / per_step_s / 1e6,
               ms_per_step=ms / args.steps, alg_bytes_per_step=alg, achieved_gbs=alg / per_step_s / 1e9,
               frac=alg / per_step_s / 1e9 / HBM_PEAK_GBS, segments=seg, blocks=blocks, graph_steps=32,
               traffic_bytes_per_step=(measured_traffic(f"ensemble{len(wls)}/f32/R{seg}", "batch_kernel")
                                       if blocks == 0 else None))
    log(f"  ensemble f32: {res['mlups']:.1f} MLUPS, {res['ms_per_step']:.4f} ms/step, {res['achieved_gbs']:.0f} GB/s "
        f"= {100 * res['frac']:.1f}% of 8 TB/s (R={seg}, "
        f"{'one wave per tile' if blocks == 0 else f'{blocks} persistent work-groups'})")
    if args.ensemble_check_steps > 0:
        # per-sample Darcy permeability (FP64 GPU reductions) of the FP32
        # ensemble against the FP64 ensemble (the oracle-pinned path)
        lats64 = [w.make_lattice(L.F64) for w in wls]
        b64 = L.Batch(lats64)
        total = args.ensemble_check_steps
        t1 = time.perf_counter()
        for name, batch, lat0 in (("f32", b, lats[0]), ("f64", b64, lats64[0])):
            while lat0.time < total:
                batch.step(min(2500, total - lat0.time), L.STEP_CHECK)
                log(f"    ensemble {name}: step {lat0.time} of {total} ({time.perf_counter() - t1:.1f}s)")
        red32, red64 = b.reduce(), b64.reduce()
        nu = (0.8 - 0.5) / 3.0
        k32 = np.array([W.darcy_permeability(red32[i], nu, w.nc - 1, 0.001, w.nr * w.nc) for i, w in enumerate(wls)])
        k64 = np.array([W.darcy_permeability(red64[i], nu, w.nc - 1, 0.001, w.nr * w.nc) for i, w in enumerate(wls)])
        rel = np.abs(k32 - k64) / np.abs(k64)
        res["permeability_check"] = dict(steps=total, max_rel_diff_f32_vs_f64=float(rel.max()),
                                         median_rel_diff=float(np.median(rel)),
                                         k_range_f64=[float(k64.min()), float(k64.max())])
        log(f"  permeability after {total} steps: FP32 vs FP64 max rel diff {rel.max():.3e} "
            f"(median {np.median(rel):.3e}), k in [{k64.min():.3e}, {k64.max():.3e}]")
        b64.close()
        for lat in lats64:
            lat.close()
    b.close()
    for lat in lats:
        lat.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="porous8192")
    ap.add_argument("--storage", default="f32", choices=list(VARIANTS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", action="store_true", help="time only the headline storage")
    ap.add_argument("--error-steps", type=int, default=None,
                    help="steps after which the shifted-storage errors are stated (default: the workload's run "
                         "length, 20000 for porous8192; 0 = skip)")
    ap.add_argument("--ensemble-samples", type=int, default=512)
    ap.add_argument("--ensemble-check-steps", type=int, default=10000,
                    help="steps after which the ensemble's FP32 permeabilities are checked against FP64 (0 = skip)")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="skip the 4096^2 TRT cavity line")
    ap.add_argument("--no-open", action="store_true", help="skip the open 8192^2 channel line")
    ap.add_argument("--c4-error-steps", type=int, default=200000,
                    help="steps after which the 4096^2 TRT cavity's FP32/FP16 errors are stated (0 = skip)")
    ap.add_argument("--no-patterns", action="store_true", help="skip the access-pattern ceilings")
    ap.add_argument("--patterns-only", action="store_true",
                    help="print the access-pattern ceilings on the workload's NR x NC and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU slab decomposition is a design, not built)", file=sys.stderr)
        sys.exit(2)

    if args.patterns_only:
        wl = build_workload(args.workload)
        nr, nc = wl.nr, wl.nc
        print(json.dumps({"access_patterns": access_patterns(nr, nc), "nr": nr, "nc": nc,
                          "unit": "GB/s (18 elements moved per node)"}), flush=True)
        return

    ev = HipEvents()
    t0 = time.perf_counter()
    wl = build_workload(args.workload)
    log(f"workload {wl.name}: {wl.nr}x{wl.nc}, wet {wl.wet_count()}, info {wl.info} "
        f"(built in {time.perf_counter() - t0:.1f}s)")
    head = args.storage
    # the copy ceiling over the FP32 step's volume (one population buffer)
    ceiling = copy_ceiling(W.algorithmic_bytes(wl, L.F32_SHIFTED) // 2)
    log(f"  device copy ceiling: {ceiling['gbs']:.0f} GB/s ({ceiling['bytes'] / 1e9:.2f} GB copies)")
    patterns = None
    if not args.no_patterns and (wl.nr * wl.nc) % 64 == 0:
        patterns = access_patterns(wl.nr, wl.nc)
        for s, p in patterns.items():
            log(f"  access patterns {s}: " + ", ".join(f"{k} {v:.0f}" for k, v in p.items()) + " GB/s")
    order = [head] if args.only else [head] + [s for s in VARIANTS if s != head]
    results, lats = {}, {}
    for s in order:
        lat, r = time_storage(wl, s, args.steps, args.warmup, ev)
        r["frac_of_copy"] = r["achieved_gbs"] / ceiling["gbs"]
        if patterns:
            # the dense pattern of the layout the lattice got
            pat = "blocked_pull" if lat.layout() == L.LAYOUT_BLOCKED else "ab_pull"
            r["frac_of_pattern"] = r["achieved_gbs"] / patterns[s[:3]][pat]
        results[s] = r
        log(f"  {r['storage']}: {r['mlups']:.1f} MLUPS, {r['ms_per_step']:.4f} ms/step, "
            f"{r['achieved_gbs']:.0f} GB/s = {100 * r['frac']:.1f}% of 8 TB/s, "
            f"{100 * r['frac_of_copy']:.0f}% of the copy ceiling")
        lats[s] = lat
    errors = None
    error_steps = wl.steps if args.error_steps is None else args.error_steps
    if getattr(wl, "random_state", False):
        error_steps = 0  # errors are stated on the physical runs, not on random states
    if error_steps > 0 and "f64" in lats and len(lats) > 1:
        t1 = time.perf_counter()
        errors = stated_errors(wl, lats["f64"], {s: lat for s, lat in lats.items() if s != "f64"}, error_steps)
        for name, e in errors["by_storage"].items():
            log(f"  {name} vs f64 after {error_steps} steps: velocity L2 rel {e['velocity_l2_rel']:.3e}, "
                f"density L2 rel {e['density_l2_rel']:.3e}, mass rel diff {e['mass_rel_diff']:.3e}")
        log(f"  (error runs took {time.perf_counter() - t1:.1f}s)")
    for lat in lats.values():
        lat.close()
    ens = None if args.no_ensemble else bench_ensemble(args, ev)
    c4 = None if args.no_c4 else bench_cavity_trt(args, ev)
    opn = None if args.no_open or args.workload == "open8192" else bench_open(args, ev)
    cpu = None
    if not args.no_cpu_baseline:
        log("  CPU baseline (reference-structured oracle) ...")
        cpu = cpu_baseline(wl, args.workload)
        log(f"  CPU: {cpu['value']:.2f} MLUPS ({cpu['sample']})")
    r = results[head]
    line = {
        "metric": "D2Q9 MLUPS and % of HBM roofline, FP64 vs FP32/FP16 at stated error",
        "value": round(r["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32", "f16": "f32", "f16rn": "f32"}[head],
        "data": "synthetic",
        "config": {"workload": wl.name, "nr": wl.nr, "nc": wl.nc, "wet_nodes": wl.wet_count(),
                   "storage": head, "collision": "BGK" if wl.collision == L.BGK else "TRT",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["frac"], 4), "traffic": r["traffic_bytes_per_step"]},
        "cpu_baseline": cpu,
        "copy_ceiling": {k: (round(v, 1) if isinstance(v, float) else v) for k, v in ceiling.items()},
        "access_patterns": patterns,
        "storages": {s: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in x.items()}
                     for s, x in results.items()},
        "error_vs_f64": errors,
        "ensemble": ens,
        "cavity_trt4096": c4,
        "open8192": opn,
    }
    if "f64" in results and head != "f64":
        line["speedup_vs_f64"] = round(r["mlups"] / results["f64"]["mlups"], 3)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
