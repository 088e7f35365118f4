"""Throughput of the D2Q9 time step on one MI355X.

A "step" is one lattice-Boltzmann time step (one launch of the fused step
kernel) over the whole workload.  Default workload: the 8192 x 8192 random
porous medium (discs r in [4, 16], 30 % solid, BGK tau 0.8, Zou-He pressure
inlet 1.001 / outlet 1.000), populations resident in HBM before the timed
region.  The headline value is the MLUPS of the FP32 shifted-population
path; the FP64 (drop-in) and FP16 paths are timed in the same run and
reported beside it, with the FP32/FP64 velocity error of the sample.

MLUPS = wet nodes * steps / time / 1e6.  Algorithmic HBM bytes per step =
wet nodes * 18 * sizeof(storage) + NR*NC/8 (bit-packed wet mask); the
roofline fraction divides the achieved rate by MI355X's 8 TB/s.  The step
kernel is the only kernel in the timed region, so its average duration is
the event-timed region divided by the number of steps.

The CPU baseline is the repo's CPU oracle (a restatement of the reference's
Serial/OpenMP policy structure) timed on this host on a few steps of the
same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload porous8192]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "latticeboltzmannmethod-fonnesu-grassi-guffanti_amd"))

import numpy as np  # noqa: E402

import lbm_amd as L  # noqa: E402
import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
STORAGE_NAMES = {L.F64: "f64", L.F32_SHIFTED: "f32", L.F16_SHIFTED: "f16"}


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


def build_workload(name):
    if name == "porous8192":
        return W.porous(8192)
    if name.startswith("open"):
        # the porous set-up without obstacles (pressure-driven channel)
        return W.porous(int(name[4:]), solid=0.0)
    if name.startswith("porous"):
        return W.porous(int(name[6:]))
    if name == "channel":
        return W.channel()
    if name.startswith("cavity_trt"):
        return W.cavity_trt(int(name[10:]) if len(name) > 10 else 4096)
    if name.startswith("cavity"):
        return W.cavity(int(name[6:]) if len(name) > 6 else 129)
    raise SystemExit(f"unknown workload {name}")


def time_storage(wl, storage, steps, warmup, ev):
    lat = wl.make_lattice(storage)
    lat.step(warmup, L.STEP_CHECK)
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
    return lat, dict(storage=STORAGE_NAMES[storage], mlups=mlups, mlups_all_nodes=wl.nr * wl.nc / per_step_s / 1e6,
                     ms_per_step=ms / steps, wall_s=wall, alg_bytes_per_step=alg, achieved_gbs=achieved,
                     frac=achieved / HBM_PEAK_GBS)


def cpu_baseline(wl, seconds_budget=20.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    orc = O.OracleLattice(wl.node_class, **wl.lattice_kwargs())
    if wl.inlet:
        orc.set_open_values(0, **wl.inlet)
    if wl.outlet:
        orc.set_open_values(1, **wl.outlet)
    orc.init()
    orc.step(1)                                 # first touch / warm-up
    t0 = time.perf_counter()
    steps = 0
    while steps < 2 or (time.perf_counter() - t0 < seconds_budget and steps < 1000):
        orc.step(1)
        steps += 1
    dt = time.perf_counter() - t0
    orc.close()
    return dict(value=wl.wet_count() * steps / dt / 1e6, unit="MLUPS", cores=threads, kind="port",
                sample=f"{steps} steps of {wl.name} ({wl.wet_count()} wet nodes), CPU oracle "
                       f"(reference-structured list sweeps, OpenMP), {threads} threads")


def velocity_error(wl, lat_ref, lat, steps):
    lat_ref.step(1, L.STEP_MACRO)
    lat.step(1, L.STEP_MACRO)
    _, u0 = lat_ref.macro()
    _, u1 = lat.macro()
    m = wl.node_class == L.FLUID
    return float(np.sqrt(np.sum((u1[:, m] - u0[:, m]) ** 2) / max(np.sum(u0[:, m] ** 2), 1e-300)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="porous8192")
    ap.add_argument("--storage", default="f32", choices=["f64", "f32", "f16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", action="store_true", help="time only the headline storage")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU slab decomposition is a design, not built)", file=sys.stderr)
        sys.exit(2)

    ev = HipEvents()
    t0 = time.perf_counter()
    wl = build_workload(args.workload)
    log(f"workload {wl.name}: {wl.nr}x{wl.nc}, wet {wl.wet_count()}, info {wl.info} "
        f"(built in {time.perf_counter() - t0:.1f}s)")
    head = {"f64": L.F64, "f32": L.F32_SHIFTED, "f16": L.F16_SHIFTED}[args.storage]
    order = [head] if args.only else [head] + [s for s in (L.F64, L.F32_SHIFTED, L.F16_SHIFTED) if s != head]
    results, lats = {}, {}
    for s in order:
        lat, r = time_storage(wl, s, args.steps, args.warmup, ev)
        results[s] = r
        log(f"  {r['storage']}: {r['mlups']:.1f} MLUPS, {r['ms_per_step']:.4f} ms/step, "
            f"{r['achieved_gbs']:.0f} GB/s = {100 * r['frac']:.1f}% of 8 TB/s")
        if s in (L.F64, head):
            lats[s] = lat
        else:
            lat.close()
    err = None
    if head != L.F64 and L.F64 in lats:
        err = velocity_error(wl, lats[L.F64], lats[head], args.warmup + args.steps)
        log(f"  {STORAGE_NAMES[head]} vs f64 velocity L2 rel error after {args.warmup + args.steps + 1} steps: {err:.3e}")
    for lat in lats.values():
        lat.close()
    cpu = None
    if not args.no_cpu_baseline:
        log("  CPU baseline (oracle) ...")
        cpu = cpu_baseline(wl)
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
        "dtype": {L.F64: "f64", L.F32_SHIFTED: "f32", L.F16_SHIFTED: "f32"}[head],
        "data": "synthetic",
        "config": {"workload": wl.name, "nr": wl.nr, "nc": wl.nc, "wet_nodes": wl.wet_count(),
                   "storage": STORAGE_NAMES[head], "collision": "BGK" if wl.collision == L.BGK else "TRT",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["frac"], 4), "traffic": None},
        "cpu_baseline": cpu,
        "storages": {STORAGE_NAMES[s]: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in x.items()}
                     for s, x in results.items()},
        "error_vs_f64": {"velocity_l2_rel": err, "steps": args.warmup + args.steps + 1} if err is not None else None,
    }
    if L.F64 in results and head != L.F64:
        line["speedup_vs_f64"] = round(r["mlups"] / results[L.F64]["mlups"], 3)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
