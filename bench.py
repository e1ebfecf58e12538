"""Throughput of the SPH step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1] [--fluid-only] [--no-cpu-baseline]
                    [--no-8m] [--no-33m] [--no-batch] [--batch-scenes S] [--no-mesh] [--no-tune]
                    [--no-plain] [--only {33m,batched,mesh}]

A step is one pass of the per-timestep pipeline over every particle of a
synthetic dam-break scene, driven through the C++ HeterogeneousSimulator and
replayed as a captured HIP graph: hash -> stable radix sort -> cell ranges +
SoA reorder -> density (+ neighbour masks) -> [bodies: world transform, body
grid, forces from the fluid] -> fluid forces + contact + walls + integrate ->
[bodies: FP64 per-body sums, body integration].  Inputs are resident in HBM
before the timed region.  `value` is particle-steps/s.

The default workload is the 1,048,576-particle dam break coupled two-way to a
rigid torus (R 0.15 m, r 0.05 m, ~20K boundary spheres, density 500) dropped
from 0.3 m above the fluid.  Further lines in `config`: the 8,388,608-particle
step with four rigid boxes (`also_8M`) and without them
(`also_8M_fluid_only`), and with the boxes in the mixed-precision form
(`mixed_8M`, with its ratio to `also_8M`); the 33,554,432-particle step with 64 rigid bodies and
a 0.5 m/s inflow (`also_33M`); the batched form over 256 independent
scenes, fixed and autotuned, beside the same scenes stepped one at a time
(`batched`); and the mesh sampler's winding-number kernel (`mesh_sampler`).
Every single-fluid line carries a `drift` report from the FP64 conservation
sums (mass, momentum against the gravity impulse, energies).

Timing: 5 runs, each a fresh simulator from the scene's initial state: W
untimed warm-up steps, then one region of exactly K steps bracketed by a
device synchronisation and HIP events on the simulator's stream; `value`
and `ms_per_step` come from the median run.

Roofline: algorithmic bytes per particle (hash 24, one sort pass 16, cells +
reorder 72, density 24, forces 60, integrate 48) plus 4 B per bucket for the
cellStart fill, 24 B per occupied cell (start/end write, two reads) and 176 B
per body sphere.  The dominant kernel's bytes are divided by its average
duration, measured with HIP events on the simulator's stream around each
stage over a second, eager pass of the same steps.

cpu_baseline: the oracle (a C++/OpenMP restatement of the reference step;
the reference itself has no CPU SPH path) timed on this host's cores on a
bounded sample of the same workload, on every OpenMP thread and on one, plus
configuration 0 (16,384 particles) for its 100 steps; `config.cfg0` times
the same 100 steps on the GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fluidgrid_amd import abi, lib, scenes  # noqa: E402

HBM_PEAK = 8.0e12
DEFAULT_SPH_BLOCK = 64  # the library's default workgroup size of the fast SPH kernels
METRIC = "particle-steps/s at 1M & 8M; % HBM peak; batched/mixed vs plain FP32"
STAGE_BYTES = dict(hash=24, sort=16, cells=72, density=24, forces=60, integrate=48)
OCC_BYTES = dict(cells=8, density=8, forces=8)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def occupied_cells(sim):
    b = sim.buffers
    start = lib.download_raw(b.gridStart, (b.numBuckets,), np.uint32)
    return int(np.count_nonzero(start != 0xFFFFFFFF))


def stage_bytes(stage, n, c_occ, fused, m=0):
    if stage == "bodies":
        return 176 * m
    by = STAGE_BYTES[stage] * n + OCC_BYTES.get(stage, 0) * c_occ
    if stage == "forces" and fused:
        by += STAGE_BYTES["integrate"] * n
    return by


REPEATS = 5


def timed_region(step, stream, sync, steps):
    """One timed region of exactly `steps` steps, the device synchronised
    before and after, HIP events on the stream the steps run on around it
    (the wall clock kept beside).  Returns (event ms, wall ms)."""
    t = lib.StreamTimer(stream)
    try:
        sync()
        w0 = time.perf_counter()
        t.start()
        step(steps)
        ev = t.stop_ms()
        sync()
        return ev, 1e3 * (time.perf_counter() - w0)
    finally:
        t.close()


def median_index(values):
    return sorted(range(len(values)), key=lambda k: values[k])[len(values) // 2]


def new_sim(scene, fused, precision=abi.PRECISION_FP32):
    sim = lib.Simulator(scene, fused=fused, use_graph=True, precision=precision)
    if scene.info.get("bodies") is not None:
        sim.add_bodies(scene.info["bodies"])
    if scene.info.get("inflow") is not None:
        sim.set_inflow(scene.info["inflow"])
    return sim


def measure(scene, steps, warmup, fused=True, precision=abi.PRECISION_FP32):
    """REPEATS runs, each a fresh simulator from the scene's initial state:
    `warmup` untimed steps, then one timed region of exactly `steps` steps;
    the median run is reported.  The last run also gives the drift report,
    the occupied cells and per-stage durations (eager steps with HIP events
    between the stages, after its timed region)."""
    bodies = scene.info.get("bodies")
    ev_all, wall_all = [], []
    for rep in range(REPEATS):
        sim = new_sim(scene, fused, precision)
        try:
            last = rep == REPEATS - 1
            if last:
                c_start = sim.conservation()
            sim.step(warmup)
            ev, wall = timed_region(sim.step, sim.stream(), sim.synchronize, steps)
            ev_all.append(ev)
            wall_all.append(wall)
            if not last:
                continue
            c_end = sim.conservation()
            drift = drift_report(scene, c_start, c_end, warmup + steps)
            c_occ = occupied_cells(sim)
            sim.set_options(fused=fused, use_graph=False, stage_timing=True)
            acc = None
            nprof = max(3, min(steps, 10))
            for _ in range(nprof):
                sim.step(1)
                t = sim.stage_times()
                acc = t if acc is None else {k: acc[k] + t[k] for k in t}
            stage_ms = {k: v / nprof for k, v in acc.items()}
            pos, _ = sim.state()
            finite = bool(np.all(np.isfinite(pos)))
            body = None
            if bodies is not None:
                bs = sim.body_state()
                finite = finite and bool(np.all(np.isfinite(bs["pos"])))
                body = {"centre_y_m": [round(float(v), 6) for v in bs["pos"][:, 1]]}
        finally:
            sim.close()
    med = median_index(ev_all)
    dt = ev_all[med] * 1e-3
    return dict(seconds=dt, ms_per_step=1e3 * dt / steps, c_occ=c_occ, stage_ms=stage_ms, finite=finite, body=body,
                drift=drift, repeats_ms=[round(x, 4) for x in ev_all], wall_ms=wall_all[med])


def drift_report(scene, c0, c1, steps):
    """Conservation over the warm-up and timed steps, from the FP64 GPU sums:
    mass (exact), the momentum change beside the gravity impulse N m g t
    (walls, bodies and the clamp account for the rest), and the energies
    (a dissipation report, not a conservation claim)."""
    t = steps * float(np.float32(scene.common.deltaTimeIntegration))
    g = np.array(tuple(scene.common.global_external_force), np.float64)
    e0 = c0["kinetic"] + c0["potential"] + c0["internal"]
    e1 = c1["kinetic"] + c1["potential"] + c1["internal"]
    return {"steps": steps, "active_start": int(c0["count"]), "active_end": int(c1["count"]),
            "mass_change": float(c1["mass"] - c0["mass"]),
            "momentum_change": [float(v) for v in c1["momentum"] - c0["momentum"]],
            "gravity_impulse": [float(v) for v in c0["mass"] * g * t],
            "energy_start": {"kinetic": c0["kinetic"], "potential": c0["potential"], "internal": c0["internal"]},
            "energy_end": {"kinetic": c1["kinetic"], "potential": c1["potential"], "internal": c1["internal"]},
            "energy_change_rel": float((e1 - e0) / abs(e0)) if e0 else None}


def measure_batched(ens, steps, warmup, tune_steps=10, tune=True, plain=True):
    """The batched form over an ensemble: the launch shape is autotuned once
    (tune=False: the library's default shape is taken instead); then the
    fixed shape (256-lane workgroups, 8-bit digits, row schedules, contact
    walked in the force kernel) and the tuned shape are each timed over
    REPEATS fresh simulators from the ensemble's initial state (median
    reported); per-stage times of the tuned shape; and (plain=True) the same
    scenes stepped one at a time through the single-fluid simulator (the
    plain form's way to run them)."""
    out = {}
    bodies = ens.bodies or [None] * len(ens.scenes)

    def fresh():
        b = lib.BatchSimulator(ens)
        for s, bs in enumerate(bodies):
            if bs is not None:
                b.add_bodies(s, bs)
        return b

    try:
        if tune:
            b = fresh()
            try:
                t = b.autotune(trial_steps=tune_steps)
            finally:
                b.close()
        else:
            t = {"sph_block": 256, "sort_digit_bits": 8, "flat_density": False, "flat_forces": False,
                 "queue_items": 0, "defer_contact": True, "trials": []}
        out["tuning"] = {"autotuned": tune, "sph_block": t["sph_block"], "sort_digit_bits": t["sort_digit_bits"],
                         "flat_density": t["flat_density"], "flat_forces": t["flat_forces"],
                         "queue_items": t["queue_items"], "defer_contact": t["defer_contact"],
                         "trials_ms": [[a, c, fd, ff, q, dc, round(m, 4)] for a, c, fd, ff, q, dc, m in t["trials"]],
                         "trials_fields": ["sph_block", "sort_digit_bits", "flat_density", "flat_forces",
                                           "queue_items (0: one workgroup per block; k: persistent, k blocks "
                                           "per work item)", "defer_contact", "ms_per_step"]}
        shapes = {"fixed": (256, 8, False, False, 0, False),
                  "tuned": (t["sph_block"], t["sort_digit_bits"], t["flat_density"], t["flat_forces"],
                            t["queue_items"], t["defer_contact"])}
        for label, shape in shapes.items():
            ev_all = []
            for rep in range(REPEATS):
                b = fresh()
                try:
                    b.set_tuning(*shape)
                    b.step(warmup)
                    ev, _ = timed_region(b.step, b.stream(), b.synchronize, steps)
                    ev_all.append(ev)
                    if label == "tuned" and rep == REPEATS - 1:
                        b.set_options(use_graph=False, stage_timing=True)
                        acc = None
                        for _ in range(5):
                            b.step(1)
                            st = b.stage_times()
                            acc = st if acc is None else {k: acc[k] + st[k] for k in st}
                        out["stage_ms"] = {k: round(v / 5, 5) for k, v in acc.items()}
                        pos, _ = b.state()
                        out["finite"] = bool(np.all(np.isfinite(pos)))
                        info = b.info()
                finally:
                    b.close()
            dt = ev_all[median_index(ev_all)] * 1e-3
            out[label] = {"ms_per_step": 1e3 * dt / steps, "value": ens.n * steps / dt,
                          "repeats_ms": [round(x, 4) for x in ev_all]}
    finally:
        lib.load().fg_set_tuning(DEFAULT_SPH_BLOCK, 8)
        lib.load().fg_set_schedule(0, 0)
    out["info"] = info
    if not plain:
        return out
    # plain form: every scene alone through the single-fluid simulator
    k = max(5, steps // 4)
    total_ms = 0.0
    for sc, bs in zip(ens.scenes, bodies):
        sim = lib.Simulator(sc, fused=True, use_graph=True)
        try:
            if bs is not None:
                sim.add_bodies(bs)
            sim.step(2)
            sim.synchronize()
            t0 = time.perf_counter()
            sim.step(k)
            sim.synchronize()
            total_ms += 1e3 * (time.perf_counter() - t0) / k
        finally:
            sim.close()
    out["plain_per_scene"] = {"ms_per_step": total_ms, "value": ens.n / (total_ms * 1e-3), "steps": k}
    return out


def measure_mesh(nu=200, nv=100, points_across=128, reps=5):
    """The mesh sampler's winding-number kernel on a 40,000-triangle torus
    mesh over its 128 x 32 x 128 lattice, inputs resident on the device;
    HIP events on the library's stream around `reps` launches."""
    import ctypes as C
    from fluidgrid_amd import mesh as fm
    L = lib.load()
    v, t = fm.torus_mesh(0.3, 0.1, nu, nv)
    spacing = 0.8 / points_across
    origin, dims = fm.lattice(v, spacing)
    npts = dims[0] * dims[1] * dims[2]
    dv = lib.DeviceArray.from_numpy(np.ascontiguousarray(v))
    dt = lib.DeviceArray.from_numpy(np.ascontiguousarray(t))
    dw = lib.DeviceArray((npts,), np.float32)
    args = (dv.ptr, len(v), dt.ptr, len(t), abi.Float3(*origin), spacing, *dims, dw.ptr)
    lib.check(L.fg_mesh_winding_lattice(*args), "mesh warm-up")
    lib.check(L.fg_synchronize(), "sync")
    e0, e1 = C.c_void_p(), C.c_void_p()
    lib.check(L.fg_event_create(C.byref(e0)), "event")
    lib.check(L.fg_event_create(C.byref(e1)), "event")
    lib.check(L.fg_event_record(e0), "event")
    for _ in range(reps):
        L.fg_mesh_winding_lattice(*args)
    lib.check(L.fg_event_record(e1), "event")
    lib.check(L.fg_synchronize(), "sync")
    ms = C.c_float()
    lib.check(L.fg_event_elapsed_ms(e0, e1, C.byref(ms)), "elapsed")
    L.fg_event_destroy(e0)
    L.fg_event_destroy(e1)
    per = ms.value / reps
    w = dw.numpy()
    return {"workload": f"winding numbers of a {len(t):,}-triangle torus mesh on its {dims[0]}x{dims[1]}x{dims[2]} "
                        "sampling lattice (fg_mesh_winding_lattice)",
            "triangles": int(len(t)), "points": int(npts), "ms": per,
            "point_triangle_pairs_per_s": npts * len(t) / (per * 1e-3),
            "inside_points": int((w > 0.5).sum()), "finite": bool(np.isfinite(w).all())}


def step_bytes(scene, c_occ):
    bodies = scene.info.get("bodies")
    m = bodies.m if bodies is not None else 0
    return 244 * scene.n + 4 * scene.num_buckets + 24 * c_occ + 176 * m


def oracle_rate(scene, seconds_budget, max_steps=50):
    """Particle-steps/s of the oracle's step (coupled when the scene has
    bodies) over at least one step and about `seconds_budget` of host time."""
    from oracle import oracle
    bodies = scene.info.get("bodies")
    steps, t_total = 0, 0.0
    if bodies is not None:
        st = oracle.CoupledState(scene, bodies)
        while (t_total < seconds_budget or steps == 0) and steps < max_steps:
            t0 = time.perf_counter()
            st = oracle.coupled_step(st)
            t_total += time.perf_counter() - t0
            steps += 1
        what = f"coupled oracle steps of the {scene.n:,}-particle scene with {bodies.m:,} body spheres"
    else:
        pos, vel = scene.pos.copy(), scene.vel.copy()
        rho_p = np.zeros((scene.n, 2), np.float32)
        while (t_total < seconds_budget or steps == 0) and steps < max_steps:
            t0 = time.perf_counter()
            st = oracle.step(pos, vel, scene.fluid, scene.common, scene.kernels, scene.key_mode, rho_p_prev=rho_p)
            t_total += time.perf_counter() - t0
            pos, vel, rho_p = st.pos, st.vel, st.rho_p
            steps += 1
        what = f"oracle steps of the {scene.n:,}-particle scene"
    return scene.n * steps / t_total, steps, t_total, what


def cpu_baseline(scene, seconds_budget=12.0, one_core_budget=6.0):
    """The oracle on this host: the headline workload on every OpenMP thread
    and on one, and configuration 0 (16,384 particles, the reference's host
    step) for its 100 steps."""
    from oracle import oracle
    threads = oracle.num_threads()
    value, steps, t_total, what = oracle_rate(scene, seconds_budget)
    oracle.set_num_threads(1)
    try:
        v1, s1, t1, _ = oracle_rate(scene, one_core_budget)
    finally:
        oracle.set_num_threads(threads)
    s0 = scenes.config_scene(0)
    v0, k0, t0, _ = oracle_rate(s0, 1e9, max_steps=s0.steps)
    return dict(value=value, unit="particle-steps/s", cores=threads, kind="port",
                sample=f"{steps} {what} ({t_total:.1f} s, {threads} OpenMP threads)",
                value_1core=v1, sample_1core=f"{s1} steps of the same on 1 thread ({t1:.1f} s)",
                cfg0_host_path={"workload": f"configuration 0, {s0.n:,} particles, {k0} oracle steps",
                                "seconds": t0, "value": v0, "cores": threads},
                cpu=cpu_model(), host_cpus=os.cpu_count())


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def line_33m(a, fused):
    """Configuration 4: 33.5M particles, 64 bodies, inflow; the configuration's 100 steps."""
    s4 = scenes.config4_scene()
    b4 = s4.info["bodies"]
    log(f"scene {s4.name}: n={s4.n:,} ({s4.info['active']:,} active), {b4.b} bodies, {b4.m:,} spheres")
    k4 = min(a.steps, s4.steps)  # the configuration's 100 steps: two inflow layers are emitted
    r4 = measure(s4, k4, 5, fused)
    act = s4.info["active"]
    return {
        "workload": f"plain FP32 dam break, {act:,} active particles (+{s4.n - act:,} inflow reserve slots), "
                    f"{s4.num_buckets:,} LINEAR buckets, h=0.003, dt=8e-5, 0.5 m/s inflow through the x-min "
                    f"face, coupled two-way with {b4.b} rigid bodies (tori and boxes, {b4.m:,} spheres)",
        "particles": act, "slots": s4.n, "steps": k4, "value": act * k4 / r4["seconds"],
        "ms_per_step": r4["ms_per_step"],
        "step_hbm_frac": step_bytes(s4, r4["c_occ"]) / (r4["ms_per_step"] * 1e-3) / HBM_PEAK,
        "stage_ms": {k: round(v, 5) for k, v in r4["stage_ms"].items()},
        "finite": r4["finite"], "drift": r4["drift"],
    }


def line_batched(a, plain=None):
    """Configuration 3: the 256-scene ensemble with its bodies, fixed and tuned, beside the
    scenes stepped one at a time."""
    ens = scenes.ensemble(num_scenes=a.batch_scenes, bodies=True)
    nbod = sum(0 if x is None else x.b for x in ens.bodies)
    nsph = sum(0 if x is None else x.m for x in ens.bodies)
    log(f"ensemble: {len(ens.scenes)} scenes, n={ens.n:,}")
    kb = max(20, a.steps // 2)
    rb = measure_batched(ens, kb, max(5, a.warmup // 2), tune=a.tune, plain=a.plain)
    out = {
        "workload": f"{len(ens.scenes)} independent FP32 dam-break scenes, {ens.n:,} particles, "
                    f"{ens.num_buckets:,} buckets; counts log-uniform in [4096, 65536], edges in [0.5, 1.5] m, "
                    f"0-4 floating rigid spheres or boxes per scene ({nbod} bodies, {nsph:,} boundary spheres), "
                    "each scene coupled two-way with its own bodies",
        "steps": kb,
        "value_fixed": rb["fixed"]["value"], "value_tuned": rb["tuned"]["value"],
        "ms_per_step_fixed": rb["fixed"]["ms_per_step"], "ms_per_step_tuned": rb["tuned"]["ms_per_step"],
        "tuning": rb["tuning"], "stage_ms": rb["stage_ms"], "finite": rb["finite"],
    }
    if "plain_per_scene" in rb:
        out["plain_per_scene_value"] = rb["plain_per_scene"]["value"]
        out["plain_per_scene_ms_per_step"] = rb["plain_per_scene"]["ms_per_step"]
        out["speedup_vs_plain_per_scene"] = rb["tuned"]["value"] / rb["plain_per_scene"]["value"]
    if plain:
        out["speedup_vs_plain_1M"] = rb["tuned"]["value"] / plain
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-8m", action="store_true")
    ap.add_argument("--no-33m", action="store_true", help="skip the 33.5M-particle config-4 line (--no-8m skips it too)")
    ap.add_argument("--unfused", action="store_true")
    ap.add_argument("--tiled", action="store_true",
                    help="run the masked density/force passes as LDS tiles (fg_set_tiled) instead of row loops")
    ap.add_argument("--fluid-only", action="store_true", help="drop the rigid bodies of the configuration")
    ap.add_argument("--no-tune", dest="tune", action="store_false",
                    help="time the main line and the batched line at the library's default launch shape "
                         "instead of autotuning them")
    ap.add_argument("--sph-block", type=int, default=0, choices=(0, 64, 128, 256, 512),
                    help="workgroup size of the fast density/force kernels (fg_set_tuning; 0: the library default)")
    ap.add_argument("--mixed", action="store_true",
                    help="the main line in the mixed-precision form (FG_PRECISION_MIXED) instead of plain FP32")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble line")
    ap.add_argument("--no-plain", dest="plain", action="store_false",
                    help="batched line without the scenes stepped one at a time")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--no-mesh", action="store_true", help="skip the mesh-sampler line")
    ap.add_argument("--no-cfg0", action="store_true",
                    help="skip configuration 0's 100-step line (for profiles of one workload)")
    ap.add_argument("--only", choices=("33m", "batched", "mesh"),
                    help="run only that line and print it alone (for profiling one workload)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB ten times (reference for HBM byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (this build runs on one GPU "
              "and shards nothing across GPUs)", file=sys.stderr)
        sys.exit(2)
    lib.load()
    if lib.gpu_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    log("device:", lib.device_name(0))
    if a.tiled:
        lib.check(lib.load().fg_set_tiled(1), "fg_set_tiled")
    if a.sph_block:
        lib.check(lib.load().fg_set_tuning(a.sph_block, 8), "fg_set_tuning")
    if a.sph_block or a.tiled:
        a.tune = False  # a shape given on the command line is the one timed
    if a.copy_calibration:
        nbytes = a.copy_calibration << 20
        src, dst = lib.DeviceArray((nbytes,), np.uint8), lib.DeviceArray((nbytes,), np.uint8)
        L = lib.load()
        for _ in range(10):
            lib.check(L.fg_copy_bytes(dst.ptr, src.ptr, nbytes), "copy")
        lib.check(L.fg_synchronize(), "sync")
        print(json.dumps({"copy_calibration_bytes": nbytes}))
        return

    if a.only:
        fn = {"33m": lambda: line_33m(a, not a.unfused), "batched": lambda: line_batched(a),
              "mesh": measure_mesh}[a.only]
        print(json.dumps({"only": a.only, a.only: fn()}), flush=True)
        return

    coupled = a.config in (1, 2) and not a.fluid_only
    scene = scenes.coupled_config_scene(a.config) if coupled else scenes.config_scene(a.config)
    bodies = scene.info.get("bodies")
    fused = not a.unfused
    log(f"scene {scene.name}: n={scene.n:,} buckets={scene.num_buckets:,} "
        f"body spheres={bodies.m if bodies is not None else 0:,}")
    prec = abi.PRECISION_MIXED if a.mixed else abi.PRECISION_FP32
    tuning = None
    if a.tune:
        # the library's own tuner picks the launch shape (no result bit depends on it)
        sim = new_sim(scene, fused, prec)
        try:
            sim.step(max(2, a.warmup // 4))
            t = sim.autotune(trial_steps=20)
        finally:
            sim.close()
        tuning = {"sph_block": t["sph_block"], "sort_digit_bits": t["sort_digit_bits"],
                  "flat_density": t["flat_density"], "flat_forces": t["flat_forces"], "tiled": t["tiled"],
                  "trials_ms": [list(x[:5]) + [round(x[5], 4)] for x in t["trials"]],
                  "trials_fields": ["sph_block", "sort_digit_bits", "flat_density", "flat_forces", "tiled",
                                    "ms_per_step"]}
    r = measure(scene, a.steps, a.warmup, fused, prec)
    if a.tune:
        # the other lines run with the library's default shape
        L = lib.load()
        lib.check(L.fg_set_tuning(DEFAULT_SPH_BLOCK, 8), "fg_set_tuning")
        L.fg_set_schedule(0, 0)
        L.fg_set_tiled(0)
    n = scene.n
    value = n * a.steps / r["seconds"]
    stage_ms = r["stage_ms"]
    dom = max((k for k in stage_ms if stage_ms[k] > 0), key=lambda k: stage_ms[k])
    dom_bytes = stage_bytes(dom, n, r["c_occ"], fused, bodies.m if bodies is not None else 0)
    achieved = dom_bytes / (stage_ms[dom] * 1e-3)
    step_gbs = step_bytes(scene, r["c_occ"]) / (r["ms_per_step"] * 1e-3)
    # HBM bytes per dispatch of the dominant kernel come from rocprofv3
    # counter passes committed under profiles/ (counters cannot be read in
    # this run); the source and the commit they were measured at are named.
    traffic, traffic_source = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            t = json.load(open(tpath))
            traffic = t.get(scene.name, {}).get(dom)
            if traffic is not None:
                traffic_source = (f"profiles/traffic.json: rocprofv3 FETCH_SIZE and WRITE_SIZE passes "
                                  f"(calibrated), measured at commit {t.get('_commit', 'unknown')}, "
                                  "not in this run")
        except Exception:
            traffic = None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain FP32 dam break: {n:,} particles, {scene.num_buckets:,} LINEAR buckets, "
                        f"h={scene.info['h']}, dt={scene.info['dt']}, "
                        + (f"coupled two-way with {bodies.b} rigid bod{'y' if bodies.b == 1 else 'ies'} "
                           f"({', '.join(bodies.names)}; {bodies.m:,} boundary spheres of radius 0.5h)"
                           if bodies is not None else "fluid only (no rigid body)"),
            "particles": n,
            "buckets": scene.num_buckets,
            "occupied_cells": r["c_occ"],
            "form": ("mixed precision density (FP16 cull, exact FP32 survivors), " if a.mixed else "plain FP32, ")
                    + ("fused forces+integrate, HIP graph replay" if fused else "unfused"),
            "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
            "tuning": tuning if tuning is not None else "library default shape (64-lane workgroups, "
                                                         "row loops, 8-bit digits)",
            "step_algorithmic_GBps": step_gbs / 1e9,
            "step_hbm_frac": step_gbs / HBM_PEAK,
            "finite": r["finite"],
            "bodies": r["body"],
            "drift": r["drift"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": (dom if not (dom == "forces" and fused) else
                       ("coupled " if bodies is not None else "") + "forces+integrate (fused)"),
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "traffic_source": traffic_source,
            "algorithmic_bytes_per_launch": dom_bytes,
        },
        "timing": {"method": f"{REPEATS} runs from the scene's initial state, each {a.warmup} untimed steps "
                             f"then HIP events on the simulator's stream around exactly {a.steps} graph-replayed "
                             "steps (device synchronised before and after); value and ms_per_step from the median "
                             "run",
                   "regions_ms": r["repeats_ms"], "median_region_wall_ms": round(r["wall_ms"], 4)},
    }
    if not a.no_8m and a.config != 2:
        k8 = max(20, a.steps // 4)
        s8c = scenes.coupled_config_scene(2)
        for key, s8, prec in (("also_8M", s8c, abi.PRECISION_FP32),
                              ("also_8M_fluid_only", scenes.config_scene(2), abi.PRECISION_FP32),
                              ("mixed_8M", s8c, abi.PRECISION_MIXED)):
            log(f"scene {s8.name}: n={s8.n:,}" + (" (mixed precision)" if prec == abi.PRECISION_MIXED else ""))
            r8 = measure(s8, k8, max(5, a.warmup // 2), fused, prec)
            b8 = s8.info.get("bodies")
            form = ("mixed precision (FP16 cell offsets, packed-FP16 conservative cull, exact FP32 survivors, "
                    "compensated density; forces FP32)" if prec == abi.PRECISION_MIXED else "plain FP32")
            out["config"][key] = {
                "workload": f"{form} dam break, {s8.n:,} particles, "
                            + (f"coupled two-way with {b8.b} rigid boxes ({b8.m:,} boundary spheres)"
                               if b8 is not None else "fluid only"),
                "particles": s8.n, "steps": k8, "value": s8.n * k8 / r8["seconds"],
                "ms_per_step": r8["ms_per_step"],
                "step_hbm_frac": step_bytes(s8, r8["c_occ"]) / (r8["ms_per_step"] * 1e-3) / HBM_PEAK,
                "stage_ms": {k: round(v, 5) for k, v in r8["stage_ms"].items()},
                "finite": r8["finite"], "drift": r8["drift"],
            }
        mx = out["config"]["mixed_8M"]
        mx["ratio_vs_plain_8M"] = mx["value"] / out["config"]["also_8M"]["value"]
    if not a.no_33m and not a.no_8m and a.config == 1:
        out["config"]["also_33M"] = line_33m(a, fused)
    if not a.no_batch:
        out["config"]["batched"] = line_batched(a, out["value"])
    if not a.no_mesh:
        log("mesh sampler ...")
        out["config"]["mesh_sampler"] = measure_mesh()
    if not a.no_cfg0:
        s0 = scenes.config_scene(0)
        r0 = measure(s0, s0.steps, 10, fused)
        out["config"]["cfg0"] = {"workload": f"configuration 0: {s0.n:,} particles, its {s0.steps} steps",
                                 "ms_per_step": r0["ms_per_step"], "value": s0.n * s0.steps / r0["seconds"],
                                 "seconds_100_steps": r0["seconds"], "finite": r0["finite"]}
    if not a.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(scene)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
