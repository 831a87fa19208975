"""Benchmark of the gfx950 LBMForce platform on its headline workload.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload config2|config2-fluid|periodic]
                [--storage fp32|fp16|bf16] [--no-cpu-baseline]

A step is one coupled integration step: beginStep (reflection), the extra forces, the kernel's lattice step
(coupling, fused pull-stream/collision/forcing/noise pass, walls) and the Verlet update, all on the GPU with the
particles resident in HBM.  The default workload is config2: 256^3, bounce-back walls, 20,000 beads with the Keys
stencil, fluctuating fluid.  Prints one JSON line.
"""
import argparse
import ctypes
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "openmm-lbm_amd"))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, 8 TB/s


def storage_code(name):
    from lbm_amd import LBMForce
    return {"fp32": LBMForce.FP32, "fp16": LBMForce.FP16, "bf16": LBMForce.BF16, "fp64": LBMForce.FP64}[name]


def bytes_per_node(storage, solids, touched_fraction):
    """Algorithmic HBM bytes per node of one fused pass: 19 populations read and 19 written; one bit of the solid
    mask; with coupled particles one bit of the touched mask, plus, at the touched fraction of the nodes, the three
    64-bit reaction accumulators read and cleared (48 B)."""
    size = {"fp32": 4, "fp16": 2, "bf16": 2, "fp64": 8}[storage]
    b = 2*19*size
    if solids:
        b += 0.125
    if touched_fraction > 0:
        b += 0.125 + 48*touched_fraction
    return b


def make_workload(name):
    from lbm_amd import workloads
    if name == "config2":
        return workloads.config2()
    if name == "config2-fluid":
        return workloads.config2(num_beads=0)
    if name == "periodic":
        w = workloads.config2(num_beads=0, fluctuations=False)
        w.configure = [lambda f: f.setBodyAcceleration((1e-6*w.dx/w.dt**2, 0, 0))]
        w.description = {"workload": "256^3 periodic D3Q19, tau=0.6, Guo 1e-6, no walls, no fluctuations"}
        return w
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(workload_name, budget_steps=1):
    """The FP64 CPU oracle (one core, the reference's serial loops restated) on the same workload, timed over a
    bounded number of lattice steps after setup."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    w = make_workload(workload_name)
    ctx = w.build(oracle_lib.load_oracle())
    ctx.step(1)      # warm-up (first touch of the arrays)
    t0 = time.perf_counter()
    ctx.step(budget_steps)
    elapsed = time.perf_counter() - t0
    mlups = w.num_nodes*budget_steps/elapsed/1e6
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": 1, "kind": "port",
            "sample": f"{budget_steps} lattice step(s) of the same {w.grid[0]}^3 workload after 1 warm-up step, "
                      f"{elapsed:.1f} s, FP64 oracle (oracle/lbm_oracle.cpp), one thread; "
                      f"host {platform.processor() or platform.machine()}, nproc {os.cpu_count()}"}


def copy_calibration(mib=1024, repeats=5):
    """A plain device-to-device copy of known size (hipMemcpy's copy kernel), to scale the FETCH_SIZE and WRITE_SIZE
    counters of rocprofv3: the factor is the known bytes over what the counters read for it."""
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib*1024*1024
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    for _ in range(repeats):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "repeats": repeats}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="config2")
    ap.add_argument("--storage", default="fp32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (for the counter scale of rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU (the whole lattice in one domain); --gpus {args.gpus} is not "
              f"supported", file=sys.stderr)
        return 2

    from lbm_amd import _lib
    w = make_workload(args.workload)
    print(f"bench: building {args.workload} ({w.grid}, {len(w.masses)} beads, storage {args.storage})",
          file=sys.stderr, flush=True)
    ctx = w.build(storage=storage_code(args.storage))
    dp = w.device_particles(ctx)
    lib = ctx.lib
    dp.step(args.warmup)
    _lib.check(lib, lib.lbm_synchronize(ctx.handle))
    _lib.check(lib, lib.lbm_profile_fused(ctx.handle, 1, None, None))
    t0 = time.perf_counter()
    dp.step(args.steps)
    _lib.check(lib, lib.lbm_synchronize(ctx.handle))
    elapsed = time.perf_counter() - t0
    total_ms = ctypes.c_double()
    count = ctypes.c_int()
    _lib.check(lib, lib.lbm_profile_fused(ctx.handle, 0, ctypes.byref(total_ms), ctypes.byref(count)))
    print(f"bench: {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    n = w.num_nodes
    steps_per_s = args.steps/elapsed
    mlups = n*steps_per_s/1e6
    fused_ms = total_ms.value/max(count.value, 1)
    # Upper bound of the touched nodes: every stencil node distinct (the beads are sparse).
    slots = {"nearest": 1, "trilinear": 8, "threepoint": 27, "keys": 64}.get(w.description.get("stencil"), 64)
    touched_fraction = min(1.0, len(w.masses)*slots/n) if len(w.masses) else 0.0
    bpn = bytes_per_node(args.storage, True, touched_fraction)
    achieved = bpn*n/(fused_ms*1e-3)/1e9
    # HBM bytes per launch of the fused pass, measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by a
    # copy of known size) for this workload and storage, where such a measurement is committed.
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"fusedStep_{args.workload}_{args.storage}.traffic.json")
    if os.path.exists(tfile):
        traffic = round(json.load(open(tfile))["traffic_bytes_per_launch"])
    result = {
        "metric": "MLUPS & coupled steps/s (HBM GB/s %) 256^3+20k beads; FP16/MLP vs plain @err",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3*elapsed/args.steps, 4),
        "coupled_steps_per_s": round(steps_per_s, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp16": "f16-storage/f32", "bf16": "bf16-storage/f32", "fp64": "f64"}[args.storage],
        "data": "synthetic",
        "config": dict(w.description, storage=args.storage),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved/HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "fusedStep", "kernel_ms": round(fused_ms, 4), "bytes_per_node": bpn,
                     "algorithmic_bytes_per_launch": round(bpn*n)},
    }
    if not args.no_cpu_baseline:
        print("bench: timing the CPU oracle baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
