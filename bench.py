"""Benchmark of the gfx950 LBMForce platform on its headline workload.

python bench.py [--gpus 1] [--steps K] [--warmup W]
                [--workload config2|config2-fluid|periodic|config3|config4|config5|centred]
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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def host_threads():
    """The host cores this process may use (its affinity), at most 16 (a one-GPU host's share of the cores)."""
    try:
        count = len(os.sched_getaffinity(0))
    except AttributeError:
        count = os.cpu_count() or 1
    return max(1, min(16, count))


def cpu_baseline(workload_name, warmup_steps=2, timed_steps=10):
    """The FP64 CPU oracle (the reference's serial loops restated; its node loops shared among OpenMP threads) on the
    same workload: all host cores over timed_steps lattice steps after warmup_steps, then one thread over one step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    lib = oracle_lib.load_oracle()
    lib.lbm_oracle_threads.restype = ctypes.c_int
    lib.lbm_oracle_threads.argtypes = [ctypes.c_int]
    w = make_workload(workload_name)
    ctx = w.build(lib)
    threads = lib.lbm_oracle_threads(host_threads())
    ctx.step(warmup_steps)
    t0 = time.perf_counter()
    ctx.step(timed_steps)
    elapsed = time.perf_counter() - t0
    lib.lbm_oracle_threads(1)
    t1 = time.perf_counter()
    ctx.step(1)
    serial = time.perf_counter() - t1
    n = w.num_nodes
    model = cpu_model()
    return {"value": round(n*timed_steps/elapsed/1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{timed_steps} coupled lattice steps of the same {w.grid[0]}^3 workload after {warmup_steps} "
                      f"warm-up steps, {elapsed:.1f} s, FP64 oracle (oracle/lbm_oracle.cpp) with its node loops on "
                      f"{threads} OpenMP threads; host {model}, nproc {os.cpu_count()}",
            "cpu_model": model,
            "single_thread": {"value": round(n/serial/1e6, 3), "unit": "MLUPS", "cores": 1,
                              "sample": f"1 further step on one thread, {serial:.1f} s"}}


WARMUP_SECONDS = 0.25


def warm_up(stepper, sync, steps):
    """At least `steps` untimed steps, and more until WARMUP_SECONDS of GPU work have run: an idle MI355X starts in a
    low clock state (sclk level S) and its first ~25 fused passes of config2 run 0.58 down to 0.45 ms before it settles,
    so a timed region that begins after a few steps would time the clock ramp rather than the step.  Returns the number
    of warm-up steps run."""
    stepper.step(steps)
    sync()
    done = steps
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < WARMUP_SECONDS and done < steps + 2000:
        stepper.step(10)
        sync()
        done += 10
    return done


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


def precision_study(steps, n=256, num_beads=20000, warmup=300):
    """The precision form: config2 with fluctuations off, run with FP64, FP32, FP16 and BF16 storage of the
    populations.  Reports, against the FP64 run (which agrees with the FP64 CPU oracle to 1e-12), the error of the mean
    velocity profile u_x(z) (plane averages, max over z relative to the max of the profile), of the bead forces (max
    and RMS relative to max |F|), the FP64 drift of total mass and momentum, and the speed of each storage."""
    import numpy as np
    from lbm_amd import LBMForce, _lib, workloads
    results = {}
    profiles = {}
    forces = {}
    for name, st in (("fp64", LBMForce.FP64), ("fp32", LBMForce.FP32), ("fp16", LBMForce.FP16), ("bf16", LBMForce.BF16)):
        w = workloads.config2(n=n, num_beads=num_beads, fluctuations=False)
        w.configure.append(lambda f: f.setTemperature(0.0))
        ctx = w.build(storage=st)
        dp = w.device_particles(ctx)
        m0, p0 = ctx.getFluidTotals()
        dp.step(warmup)             # the same number of untimed steps for every storage, so the runs stay comparable
        _lib.check(ctx.lib, ctx.lib.lbm_synchronize(ctx.handle))
        _lib.check(ctx.lib, ctx.lib.lbm_profile_fused(ctx.handle, 1, None, None))
        t0 = time.perf_counter()
        dp.step(steps)
        _lib.check(ctx.lib, ctx.lib.lbm_synchronize(ctx.handle))
        elapsed = time.perf_counter() - t0
        total = ctypes.c_double()
        count = ctypes.c_int()
        _lib.check(ctx.lib, ctx.lib.lbm_profile_fused(ctx.handle, 0, ctypes.byref(total), ctypes.byref(count)))
        m1, p1 = ctx.getFluidTotals()
        dp.pull()
        rho, u = ctx.getFluidFields()
        ux = u[:, 0].reshape(n, n*n).mean(axis=1)
        profiles[name] = ux
        forces[name] = ctx.getForces()[:num_beads]
        results[name] = {"ms_per_step": 1e3*elapsed/steps, "fused_ms": total.value/max(1, count.value),
                         "mass_drift": (m1 - m0)/m0, "momentum_x": p1[0]}
        print(f"precision: {name} done ({results[name]['ms_per_step']:.3f} ms/step)", file=sys.stderr, flush=True)
        del dp, ctx
    ref_u, ref_f = profiles["fp64"], forces["fp64"]
    for name in results:
        results[name]["profile_rel_err"] = float(np.abs(profiles[name] - ref_u).max()/np.abs(ref_u).max())
        df = np.linalg.norm(forces[name] - ref_f, axis=1)
        fmax = np.linalg.norm(ref_f, axis=1).max()
        results[name]["force_max_rel_err"] = float(df.max()/fmax)
        results[name]["force_rms_rel_err"] = float(np.sqrt((df**2).mean())/np.sqrt((np.linalg.norm(ref_f, axis=1)**2).mean()))
        results[name]["speedup_vs_fp32"] = results["fp32"]["ms_per_step"]/results[name]["ms_per_step"]
        results[name]["fused_speedup_vs_fp32"] = results["fp32"]["fused_ms"]/results[name]["fused_ms"]
    return {"metric": "precision form: storage speed vs FP32 and error vs FP64 (config2, fluctuations off)",
            "value": round(results["fp16"]["speedup_vs_fp32"], 3), "unit": "x FP32 coupled steps/s (FP16 storage)",
            "steps": steps, "higher_is_better": True, "dtype": "f16-storage/f32",
            "config": {"workload": f"config3 = config2 {n}^3, {num_beads} beads Keys, fluctuations off, T=0",
                       "reference": "FP64 storage run on the GPU (FP64 oracle parity 1e-12)"},
            "storages": {k: {kk: (round(vv, 8) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                         for k, v in results.items()}}


def batch_study(steps, warmup, sequential_steps=20):
    """config4: 64 uneven periodic replicas stepped together (hipGraph replay, work-group size picked by the
    autotuner).  value: aggregate MLUPS over all replicas.  Also times the same replicas run one after another
    through the single-lattice path (sequential_steps steps each), the way the reference runs them."""
    from lbm_amd import workloads
    from lbm_amd.batch import Batch
    from lbm_amd.harness import DeviceParticles
    from lbm_amd.context import Context, System, VerletIntegrator
    replicas, rows, description = workloads.config4(steps=steps)
    print(f"bench: building config4 ({len(replicas)} replicas, {description['total_nodes']} nodes, "
          f"{description['total_beads']} beads)", file=sys.stderr, flush=True)
    batch = Batch(replicas)
    tune, best = batch.autotune(20)
    print(f"bench: autotune ms/step {tune} -> {best} (fused, coupling) threads", file=sys.stderr, flush=True)
    batch.step(warmup)
    batch.synchronize()
    t0 = time.perf_counter()
    batch.step(steps)
    batch.synchronize()
    elapsed = time.perf_counter() - t0
    batch.profileFused(True)
    batch.step(20)
    fused_total, fused_count = batch.profileFused(False)
    fused_ms = fused_total/max(fused_count, 1)
    nodes = description["total_nodes"]
    beads = description["total_beads"]
    steps_per_s = steps/elapsed
    mlups = nodes*steps_per_s/1e6
    # Touched nodes of the trilinear stencil: at most 8 per bead.
    bpn = bytes_per_node("fp32", False, min(1.0, 8*beads/nodes))
    achieved = bpn*nodes/(fused_ms*1e-3)/1e9
    # HBM bytes per launch of fusedBatch at the picked work-group size, measured by rocprofv3 --pmc FETCH_SIZE /
    # WRITE_SIZE (scaled by a copy of known size) on this workload.
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "fusedBatch_config4_fp32.traffic.json")
    if os.path.exists(tfile):
        shape = json.load(open(tfile))["per_threads"].get(str(best[0]))
        traffic = round(shape["traffic_bytes_per_launch"]) if shape is not None else None

    # The same replicas one after another through the single-lattice path (one GPU, one replica at a time).
    seq_ms = 0.0
    for rep in replicas:
        system = System(rep.box)
        for m in rep.masses:
            system.addParticle(m)
        ctx = Context(system, VerletIntegrator(rep.stepSize), rep.force)
        ctx.setPositions(rep.positions)
        ctx.setVelocities(rep.velocities)
        dp = DeviceParticles(ctx)
        dp.step(5)
        ctx.lib.lbm_synchronize(ctx.handle)
        t = time.perf_counter()
        dp.step(sequential_steps)
        ctx.lib.lbm_synchronize(ctx.handle)
        seq_ms += 1e3*(time.perf_counter() - t)/sequential_steps
        del dp, ctx
    return {
        "metric": "MLUPS & coupled steps/s (HBM GB/s %) 256^3+20k beads; FP16/MLP vs plain @err",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(1e3*elapsed/steps, 4),
        "coupled_steps_per_s": round(steps_per_s, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict(description, autotune_ms_per_step={f"{f}x{c}": round(v, 4) for (f, c), v in tune.items()},
                       fused_threads=best[0], coupling_threads=best[1]),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved/HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "fusedBatch",
                     "kernel_ms": round(fused_ms, 4), "bytes_per_node": bpn,
                     "algorithmic_bytes_per_launch": round(bpn*nodes),
                     "traffic_source": ("profiles/fusedBatch_config4_fp32.traffic.json: rocprofv3 --pmc FETCH_SIZE and "
                                        "WRITE_SIZE of the same workload, scaled by a copy of known size")
                                       if traffic is not None else None},
        "sequential_single_lattice": {"ms_per_step_all_replicas": round(seq_ms, 4),
                                      "speedup_of_batch": round(seq_ms/(1e3*elapsed/steps), 3),
                                      "steps_per_replica_timed": sequential_steps},
    }


def centred_study(steps, warmup):
    """The centred drag (LBMForce::Centered) on config2 (256^3, walls, 20,000 beads, fluctuating fluid), against the
    explicit drag of the same run: coupled steps with the nearest-node and the Keys stencil.  The centred drag solves
    for the forces of the step implicitly (closed form per node at the nearest node; conjugate gradients over
    overlapping stencils), so its coupling costs more than the explicit drag's."""
    from lbm_amd import LBMForce, _lib, workloads
    out = {}
    for name, stencil in (("nearest", LBMForce.NearestNode), ("keys", LBMForce.Keys)):
        for drag_name, drag in (("explicit", LBMForce.Explicit), ("centred", LBMForce.Centered)):
            w = workloads.config2(stencil=stencil)
            w.configure.append(lambda f, drag=drag: f.setDragScheme(drag))
            print(f"bench: config2 {name} stencil, {drag_name} drag", file=sys.stderr, flush=True)
            ctx = w.build()
            dp = w.device_particles(ctx)
            lib = ctx.lib
            dp.step(warmup)
            _lib.check(lib, lib.lbm_synchronize(ctx.handle))
            t0 = time.perf_counter()
            dp.step(steps)
            _lib.check(lib, lib.lbm_synchronize(ctx.handle))
            ms = 1e3*(time.perf_counter() - t0)/steps
            out.setdefault(name, {})[drag_name] = {"ms_per_step": round(ms, 4), "coupled_steps_per_s": round(1e3/ms, 2),
                                                   "mlups": round(w.num_nodes/ms/1e3, 1)}
            del dp, ctx
        out[name]["centred_cost_vs_explicit"] = round(out[name]["centred"]["ms_per_step"]/
                                                      out[name]["explicit"]["ms_per_step"], 3)
    return {
        "metric": "MLUPS & coupled steps/s (HBM GB/s %) 256^3+20k beads; FP16/MLP vs plain @err",
        "value": out["keys"]["centred"]["mlups"],
        "unit": "MLUPS (centred drag, Keys)",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "config2 with the centred drag: 256^3 Poiseuille, bounce-back walls, 20,000 beads, "
                               "fluctuating kT=2.494", "stencils": ["nearest", "keys"]},
        "drags": out,
    }


def block_keys_floor(features, targets):
    """An estimate of the error a function of the closure's inputs must leave: the part of u_Keys - u(nearest) carried
    by the stencil nodes outside the 3x3x3 block (the Keys weights restricted to the block reproduce the rest exactly).
    The thermal fluctuations of the outer nodes are independent of the inputs; their smooth mean-flow part is not, so
    this is an estimate of the floor rather than a strict bound.  Returns rms(outer part)/rms(target)."""
    import numpy as np

    def keys(r):
        a = np.abs(r)
        return np.where(a <= 1, 1 - 2.5*a*a + 1.5*a**3, np.where(a < 2, 2 - 4*a + 2.5*a*a - 0.5*a**3, 0.0))
    off = features[:, 0:3].astype(np.float64)/2.0
    u = features[:, 3:84].astype(np.float64).reshape(-1, 27, 3)
    inner = -u[:, 13, :]
    for c in range(3):
        for b in range(3):
            for a in range(3):
                w = keys(off[:, 0] - (a - 1))*keys(off[:, 1] - (b - 1))*keys(off[:, 2] - (c - 1))
                inner = inner + w[:, None]*u[:, a + 3*b + 9*c, :]
    rms = lambda x: float(np.sqrt(np.mean(np.sum(x**2, axis=1))))
    return rms(inner - targets)/rms(targets)


def closure_study(steps, warmup, snapshots=10, interval=20, epochs=100, batch=2048):
    """config5: the learned closure (nearest node + MLP, Keys fallback) against the Keys stencil on 128^3 with 100,000
    chained beads.  Generates 10 x 100,000 training pairs from the Keys run itself, trains on the GPU, then times the
    same coupled steps with Keys and with the closure, and measures the closure's bead forces against Keys at one state
    (same random forces, so the difference is the interpolation alone)."""
    import numpy as np
    from lbm_amd import _lib, workloads
    from lbm_amd.closure import MLPClosure, pairs, stats
    w = workloads.config5()
    print(f"bench: building config5 ({w.grid}, {len(w.masses)} beads)", file=sys.stderr, flush=True)
    ctx = w.build()
    dp = w.device_particles(ctx)
    lib = ctx.lib
    dp.step(warmup)
    t0 = time.perf_counter()
    feats, targs = [], []
    for _ in range(snapshots):
        dp.step(interval)
        pos, vel = dp.pull()
        f, t = pairs(ctx, pos, vel)
        feats.append(f)
        targs.append(t)
    gen_s = time.perf_counter() - t0
    X, T = np.concatenate(feats), np.concatenate(targs)
    net = MLPClosure(seed=1)
    t0 = time.perf_counter()
    losses = net.train(X, T, epochs=epochs, batch=batch, lr=3e-3, decay=0.96, seed=1)
    train_s = time.perf_counter() - t0
    print(f"bench: {len(X)} pairs in {gen_s:.1f} s, trained {epochs} epochs in {train_s:.1f} s, loss {losses[0]:.4f} "
          f"-> {losses[-1]:.4f}", file=sys.stderr, flush=True)

    def timed(n):
        _lib.check(lib, lib.lbm_synchronize(ctx.handle))
        start = time.perf_counter()
        dp.step(n)
        _lib.check(lib, lib.lbm_synchronize(ctx.handle))
        return (time.perf_counter() - start)/n

    keys_s = timed(steps)
    net.attach(ctx)
    dp.step(10)
    stats(ctx)
    closure_s = timed(steps)
    network, fallback = stats(ctx)
    # Judge at the current state: the same step's forces with the closure and with Keys.
    pos, vel = dp.pull()
    f_closure = ctx.getForces()
    MLPClosure.detach(ctx)
    f_keys = ctx.getForces()
    f, t = pairs(ctx, pos, vel)
    scale = 10.0*100.0*w.dx/w.dt                 # friction x mass x dx/dt: lattice velocity -> drag force
    slip = (t + f[:, 42:45]) - f[:, 84:87]        # u_Keys - v, lattice units
    drag = scale*slip
    df = f_closure - f_keys
    rms = lambda a: float(np.sqrt(np.mean(np.sum(np.asarray(a, dtype=np.float64)**2, axis=1))))
    pred = net.forward(f)
    n = w.num_nodes
    return {
        "metric": "MLUPS & coupled steps/s (HBM GB/s %) 256^3+20k beads; FP16/MLP vs plain @err",
        "value": round(keys_s/closure_s, 3),
        "unit": "x Keys coupled steps/s (nearest node + MLP closure)",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "higher_is_better": True, "dtype": "f32 (MLP f16/f32-acc)",
        "data": "synthetic",
        "config": w.description,
        "keys": {"ms_per_step": round(1e3*keys_s, 4), "coupled_steps_per_s": round(1/keys_s, 2),
                 "mlups": round(n/keys_s/1e6, 1)},
        "closure": {"ms_per_step": round(1e3*closure_s, 4), "coupled_steps_per_s": round(1/closure_s, 2),
                    "mlups": round(n/closure_s/1e6, 1), "network_couplings": network, "fallback_couplings": fallback},
        # The closure's accuracy is stated on the drag it approximates, gamma m (u_Keys - v): at 300 K the total coupling
        # force is dominated by the random force, identical in both runs, which would deflate the ratio.
        "bead_force_error_vs_keys": round(rms(df)/rms(drag), 5),
        "bead_force_error_definition": "rms |F_closure - F_Keys| / rms |gamma m (u_Keys - v)| over the beads at one "
                                       "state (same random forces); target <= 0.02",
        "force_error_vs_keys": {
            "rms_rel_total_force": round(rms(df)/rms(f_keys), 5),
            "rms_rel_drag_force": round(rms(df)/rms(drag), 5),
            "nearest_only_rms_rel_drag_force": round(rms(scale*t)/rms(drag), 5),
            "network_rms_rel_target": round(rms(pred - t)/rms(t), 4),
            "inputs_floor_rms_rel_target": round(block_keys_floor(f, t), 4),
            "inputs_floor_rms_rel_drag_force": round(block_keys_floor(f, t)*rms(scale*t)/rms(drag), 5)},
        "training": {"pairs": int(len(X)), "epochs": epochs, "batch": batch, "generation_s": round(gen_s, 2),
                     "train_s": round(train_s, 2), "loss_first": round(float(losses[0]), 5),
                     "loss_last": round(float(losses[-1]), 5)},
    }


def extra_configs():
    """Bounded runs of the other benchmark configurations, summarized as extra keys of the default line: the precision
    form (config3: FP16/BF16 against FP32 and FP64 storage), the batched replicas (config4) and the learned closure
    (config5, against the Keys stencil)."""
    out = {}
    print("bench: config3 (precision form, 200 timed steps per storage)", file=sys.stderr, flush=True)
    c3 = precision_study(200)
    st = c3["storages"]
    out["config3"] = {
        "workload": c3["config"]["workload"], "steps": c3["steps"],
        "metric": "coupled steps/s of FP16/BF16 storage over FP32; u_x(z) profile and bead-force errors vs the FP64 run",
        "fp16_speedup_vs_fp32": round(st["fp16"]["speedup_vs_fp32"], 3),
        "bf16_speedup_vs_fp32": round(st["bf16"]["speedup_vs_fp32"], 3),
        "fp16_profile_rel_err": st["fp16"]["profile_rel_err"], "bf16_profile_rel_err": st["bf16"]["profile_rel_err"],
        "storages": st}
    print("bench: config4 (batched replicas, 200 timed steps)", file=sys.stderr, flush=True)
    c4 = batch_study(200, 20, sequential_steps=10)
    out["config4"] = {k: c4[k] for k in ("value", "unit", "ms_per_step", "coupled_steps_per_s", "config", "roofline",
                                         "sequential_single_lattice")}
    print("bench: config5 (learned closure vs Keys, 100 timed steps)", file=sys.stderr, flush=True)
    c5 = closure_study(100, 20)
    out["config5"] = {k: c5[k] for k in ("value", "unit", "config", "keys", "closure", "bead_force_error_vs_keys",
                                         "bead_force_error_definition", "force_error_vs_keys", "training")}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="config2")
    ap.add_argument("--storage", default="fp32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the bounded config3/config4/config5 runs that the default config2 line carries")
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
    if args.workload == "config3":
        print(json.dumps(precision_study(args.steps)), flush=True)
        return 0
    if args.workload == "config4":
        print(json.dumps(batch_study(args.steps, args.warmup)), flush=True)
        return 0
    if args.workload == "config5":
        print(json.dumps(closure_study(args.steps, args.warmup)), flush=True)
        return 0
    if args.workload == "centred":
        print(json.dumps(centred_study(args.steps, args.warmup)), flush=True)
        return 0

    from lbm_amd import _lib
    w = make_workload(args.workload)
    print(f"bench: building {args.workload} ({w.grid}, {len(w.masses)} beads, storage {args.storage})",
          file=sys.stderr, flush=True)
    ctx = w.build(storage=storage_code(args.storage))
    dp = w.device_particles(ctx)
    lib = ctx.lib
    warmup_run = warm_up(dp, lambda: _lib.check(lib, lib.lbm_synchronize(ctx.handle)), args.warmup)
    _lib.check(lib, lib.lbm_profile_fused(ctx.handle, 1, None, None))
    t0 = time.perf_counter()
    dp.step(args.steps)
    _lib.check(lib, lib.lbm_synchronize(ctx.handle))
    elapsed = time.perf_counter() - t0
    total_ms = ctypes.c_double()
    count = ctypes.c_int()
    launches = (ctypes.c_float*args.steps)()
    _lib.check(lib, lib.lbm_profile_fused_launches(ctx.handle, launches, args.steps, ctypes.byref(count)))
    launch_ms = sorted(launches[:min(count.value, args.steps)])
    _lib.check(lib, lib.lbm_profile_fused(ctx.handle, 0, ctypes.byref(total_ms), ctypes.byref(count)))
    print(f"bench: {args.steps} steps in {elapsed:.3f} s; fused pass per launch (ms): "
          f"{' '.join(f'{t:.4f}' for t in launches[:min(count.value, args.steps)])}", file=sys.stderr, flush=True)

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
        "warmup_run": warmup_run,
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
                     "algorithmic_bytes_per_launch": round(bpn*n),
                     "kernel_ms_min_median_max": [round(launch_ms[0], 4), round(launch_ms[len(launch_ms)//2], 4),
                                                  round(launch_ms[-1], 4)] if launch_ms else None,
                     "traffic_source": (f"profiles/fusedStep_{args.workload}_{args.storage}.traffic.json: rocprofv3 "
                                        f"--pmc FETCH_SIZE and WRITE_SIZE of the same workload, scaled by a copy of "
                                        f"known size") if traffic is not None else None},
    }
    del dp, ctx
    if args.workload == "config2" and args.storage == "fp32" and not args.no_extras:
        result.update(extra_configs())
    if not args.no_cpu_baseline:
        print("bench: timing the CPU oracle baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
