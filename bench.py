#!/usr/bin/env python3
"""COLA particle-steps/s on one MI355X.

Workload (default): 512^3 particles, 1024^3 PM mesh, periodic box
1024 Mpc/h, flat LCDM (Om 0.31, Ob 0.049, h 0.67, ns 0.96, sigma8 0.81,
Eisenstein-Hu no-wiggle P(k)), 2LPT initial conditions at a = 0.1 from
seeded white noise, 10 COLA steps linear in a to a = 1, FP32 state with FP64
time coefficients.

A step is one pm.force + cola.kick + cola.drift over all particles, all
device-resident.  --warmup steps run first on a copy of the initial state,
which is then restored, so the timed --steps are steps 0..K-1 of a K-step run
from a = 0.1.  value = Np * K / T over the timed K steps.

Prints one JSON line (see the "roofline" and "cpu_baseline" fields).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def schedule(a0, a1, nstep):
    return [(a0 + (a1 - a0) / nstep * (i + 0.5), a0 + (a1 - a0) / nstep * (i + 1.0))
            for i in range(nstep)]


def cola_step(fs, p, a_vel, a_pos):
    fs.pm.force(p)
    fs.cola.kick(p, a_vel)
    fs.cola.drift(p, a_pos)


def adjoint_leg(fs, ps, reps=5, n=128, nm=256, L=256.0):
    """Differentiable COLA: n^3 particles, nm^3 mesh (default 128^3 / 256^3,
    box 256 Mpc/h), 10 steps a = 0.1 -> 1; loss = mean((delta - target)^2)
    against the final density of a second seed.  Times the checkpointing
    forward and the adjoint sweep (gradient with respect to the n^3 white
    noise), both device-resident, for the two checkpoint policies: 'store'
    (every step's force meshes and bins kept; the default whenever it fits,
    2.6 GB at 128^3 / 256^3, 154 GB at 512^3 / 1024^3) and 'recompute'
    (positions only)."""
    import torch
    from fs2_amd.autograd import ColaModel
    c = fs._fs
    nstep = 10
    res = {"workload": f"differentiable COLA {n}^3 particles, {nm}^3 mesh, box {L:g} Mpc/h, "
                       f"{nstep} steps a=0.1->1, d MSE(delta, target) / d noise"}
    for policy in ("store", "recompute"):
        print(f"[bench] adjoint {n}^3/{nm}^3 {policy}", file=sys.stderr, flush=True)
        try:
            model = ColaModel(n, nm, L, 0.1, 1.0, nstep, ps, checkpoint=policy)
        except MemoryError as e:   # 'store' at large sizes may not fit
            res[policy] = {"error": str(e)}
            continue
        gen = torch.Generator(device="cuda").manual_seed(2)
        target = model.forward(torch.randn((n,) * 3, device="cuda", generator=gen))
        gen.manual_seed(1)
        w = torch.randn((n,) * 3, device="cuda", generator=gen)
        delta = model.forward(w)
        dbar = 2.0 * (delta - target) / delta.numel()
        model.backward(dbar)   # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            delta = model.forward(w)
        torch.cuda.synchronize()
        t_fwd = (time.perf_counter() - t0) / reps
        c._profile_reset()
        c._profile_enable(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            g = model.backward(dbar)
        torch.cuda.synchronize()
        t_bwd = (time.perf_counter() - t0) / reps
        prof = c._profile_get()
        c._profile_enable(False)
        top = sorted(prof, key=lambda r: -r[2])[:6]
        res[policy] = {
            "forward_ms": round(t_fwd * 1e3, 3),
            "backward_ms": round(t_bwd * 1e3, 3),
            "adjoint_forward_ratio": round(t_bwd / t_fwd, 3),
            "grad_norm": float(g.norm()),
            "backward_top_kernels_ms": {name: round(total / reps, 3)
                                        for name, _, total, _ in top},
        }
        del model
    res["adjoint_forward_ratio"] = res["store"].get("adjoint_forward_ratio",
                                                    res["recompute"]["adjoint_forward_ratio"])
    res["checkpoint_bytes"] = {"positions": 12 * n**3 * (nstep + 1),
                               "store_extra": nstep * (3 * 4 * nm * nm * 2 * ((nm // 2 + 1 + 15) // 16 * 16)
                                                       + 16 * n**3)}
    return res


def batch_leg(fs, ps, reps=5):
    """Batched ensemble: 96 realisations with distinct seeds, 32 x (32^3 p,
    64^3 mesh, box 32), 32 x (48^3 p, 96^3 mesh, box 48), 32 x (64^3 p, 128^3
    mesh, box 64) Mpc/h, 2LPT at a = 0.1, 10 COLA steps to a = 1, all in one
    hipGraph (one stream per size class) with autotuned paint tile shapes.
    value = total particles * 10 / T(one graph replay), ICs restored before
    each replay outside the timed region."""
    from fs2_amd.batch import Ensemble
    classes = [(32, 64, 32, 32.0), (48, 96, 32, 48.0), (64, 128, 32, 64.0)]
    nstep = 10
    ens = Ensemble(classes)
    ens.init_lpt(ps, 0.1, range(1, ens.nbox_total + 1))
    tuned = ens.autotune(reps=3)
    ens.set_schedule(0.1, 1.0, nstep)
    ens.capture()

    def timed(graph):
        ts = []
        for _ in range(reps):
            ens.reset()
            ens.synchronize()
            t0 = time.perf_counter()
            ens.run(graph=graph)
            ens.synchronize()
            ts.append(time.perf_counter() - t0)
        ts.sort()
        return ts[len(ts) // 2]

    timed(True)   # warm-up replay
    t_graph = timed(True)
    t_eager = timed(False)
    return {
        "workload": "96 boxes: 32 x 32^3 p/64^3 mesh/32 Mpc/h + 32 x 48^3 p/96^3/48 + "
                    "32 x 64^3 p/128^3/64, 10 COLA steps a=0.1->1, one hipGraph",
        "particles": ens.np_total,
        "value": ens.np_total * nstep / t_graph,
        "unit": "particle-steps/s",
        "ms_graph": round(t_graph * 1e3, 3),
        "ms_eager": round(t_eager * 1e3, 3),
        "eager_value": ens.np_total * nstep / t_eager,
        "tiles": {f"{c[0]}^3/{c[1]}^3": {"tile": list(t), "bin_paint_readout_ms": round(ms, 4)}
                  for c, (t, ms) in zip(classes, tuned)},
    }


def fof_leg(fs, ps, n_p=256, nc=512, box=512.0, nstep=10):
    """Friends-of-friends on the final positions of a COLA run (n_p^3
    particles, nc^3 mesh, 10 steps a=0.1->1), linking length 0.2 x mean
    separation.  Times the whole call: D2H of x through pinned memory, the
    kd-tree build and the linking on the host threads (OMP_NUM_THREADS, else
    all cores, at most 64), and the device reordering."""
    import numpy as np
    p = fs.lpt.init(n_p, box, 0.1, ps, 1, "cola")
    fs.pm.init(nc, nc / n_p, box)
    fs.pm.set_paint_variant("sorted")
    for a_vel, a_pos in schedule(0.1, 1.0, nstep):
        cola_step(fs, p, a_vel, a_pos)
    fs._fs._synchronize()
    ll = 0.2 * box / n_p
    t0 = time.perf_counter()
    nfof = fs.fof.find_groups(p, ll, compute_nfof=True)
    t = time.perf_counter() - t0
    sizes = nfof[nfof > 0]
    return {
        "workload": f"FoF b=0.2 on {n_p}^3 particles after 10 COLA steps ({nc}^3 mesh, box {box:g})",
        "particles": n_p**3,
        "seconds": round(t, 4),
        "value": n_p**3 / t,
        "unit": "particles/s",
        "host_threads": max(1, min(64, int(os.environ.get("OMP_NUM_THREADS", "0")) or
                                   (os.cpu_count() or 1))),
        "groups": int(len(sizes)),
        "groups_ge_20": int(np.count_nonzero(sizes >= 20)),
        "largest": int(sizes.max()),
    }


def _cpu_threads():
    """Host threads for the CPU baseline: the cores this process may use,
    at most 16 (one GPU's share of its host)."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return max(1, min(16, usable))


def _records_of(p):
    from oracle import cpu_baseline as cb
    return cb.records(p.x, p.v, p.dx1, p.dx2, p.id)


def _cpu_run(recs, nc, box, pm_factor, a_x, a_v, steps, threads):
    """Runs `steps` of the reference's CPU step (oracle/cpu_step.cpp) and
    returns seconds per step and per phase."""
    from oracle import cpu_baseline as cb
    sim = cb.CpuCola(recs, nc, box, pm_factor, 0.31, a_x, threads, a_v=a_v)
    tot, phases = 0.0, {}
    for a_vel, a_pos in steps:
        t0 = time.perf_counter()
        ph = sim.step(a_vel, a_pos)
        dt = time.perf_counter() - t0
        tot += dt
        print(f"[bench] cpu step {nc}^3 mesh, {sim.threads} thread(s): {dt:.2f} s",
              file=sys.stderr, flush=True)
        for k, v in ph.items():
            phases[k] = phases.get(k, 0.0) + v
    n = len(steps)
    return tot / n, {k: round(v / n, 3) for k, v in phases.items()}, sim.threads


def cpu_baseline(state, n_p, nc, box):
    """The reference's CPU step (oracle/cpu_step.cpp: the loops of
    py/pm.cpp and py/cola.cpp on 56-byte AoS records, OpenMP at the
    reference's pragma sites, FP32 pocketfft transforms) timed for one step
    on the final, clustered state of the GPU run, with all usable host cores
    (at most 16)."""
    from oracle import cpu_baseline as cb
    threads = _cpu_threads()
    recs, a_x, a_v = state
    t_step, phases, used = _cpu_run(recs, nc, box, nc / n_p, a_x, a_v,
                                    [(a_v + 0.09, a_x + 0.09)], threads)
    return {
        "value": n_p**3 / t_step,
        "unit": "particle-steps/s",
        "cores": used,
        "kind": "port",
        "seconds_per_step": round(t_step, 3),
        "phases_s": phases,
        "host": cb.host_info(),
        "sample": (f"1 COLA step of the reference's CPU loops (C++ port, OpenMP at its pragma "
                   f"sites, {used} threads; FP32 pocketfft, {used} workers) at {n_p}^3 "
                   f"particles / {nc}^3 mesh on the a = {a_x:g} state of the GPU run"),
    }


def cpu_full_leg(fs, ps):
    """The CPU baseline planned for the reference's loops: configurations
    A (64^3 / 128^3, box 64) and B (256^3 / 512^3, box 512) timed over all 10
    COLA steps from the 2LPT state, C (512^3 / 1024^3, box 1024) for the
    last 2 steps from the GPU state after 8, extrapolated to 10; each with 1
    thread (the reference's shipped build has no OpenMP) and with all usable
    cores (at most 16)."""
    from oracle import cpu_baseline as cb
    out = {"host": cb.host_info()}
    threads = sorted({1, _cpu_threads()})
    sched = schedule(0.1, 1.0, 10)
    for name, n_p, nc, box in (("A", 64, 128, 64.0), ("B", 256, 512, 512.0),
                               ("C", 512, 1024, 1024.0)):
        p = fs.lpt.init(n_p, box, 0.1, ps, 1, "cola")
        steps, a_x, a_v = sched, 0.1, 0.1
        if name == "C":
            fs.pm.init(nc, nc / n_p, box)
            fs.pm.set_paint_variant("sorted")
            for a_vel, a_pos in sched[:8]:
                cola_step(fs, p, a_vel, a_pos)
            steps, a_x, a_v = sched[8:], sched[7][1], sched[7][0]
        recs = _records_of(p)
        del p
        for t in threads:
            print(f"[bench] cpu_full {name} {t} thread(s)", file=sys.stderr, flush=True)
            t_step, phases, used = _cpu_run(recs.copy(), nc, box, nc / n_p, a_x, a_v, steps, t)
            out[f"{name}_{used}t"] = {
                "particles": n_p**3, "mesh": nc, "steps_timed": len(steps),
                "extrapolated_to_10_steps": len(steps) != 10,
                "seconds_per_step": round(t_step, 3),
                "seconds_10_steps": round(10 * t_step, 2),
                "value": n_p**3 / t_step, "unit": "particle-steps/s", "threads": used,
                "phases_s": phases}
    return out


PROFILE_TABLE = os.path.join(ROOT, "profiles", "cola_512p_1024mesh_kernels.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (read + write), from the rocprofv3
    FETCH_SIZE / WRITE_SIZE summary committed under profiles/ for this
    workload, or None.  Rows: | `rocprof name` | short | calls | avg us |
    read GB | write GB | total GB |."""
    try:
        with open(PROFILE_TABLE) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 7 and cells[1] == kernel:
                    return float(cells[6]) * 1e9
    except (OSError, ValueError):
        return None
    return None


def copy_probe(fs, gib=4):
    """The practical ceiling of a 1-read + 1-write streaming pass: the best
    of the library's grid-stride float4 copy over a few grid sizes
    (`fs2_copy_bandwidth`, HIP events on the library stream), with torch's
    copy kernel beside it.  `gib` GiB are read and as many written."""
    import torch
    nbytes = gib * (1 << 30)
    grids = {}
    for g in (1024, 2048, 4096, 16384, 65536):
        ms = fs._fs._copy_bandwidth(nbytes, g, 5)
        grids[g] = round(2 * nbytes / (ms * 1e-3) / 1e9, 1)
    g_best = max(grids, key=grids.get)
    n = nbytes // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 5
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms_t = e0.elapsed_time(e1) / reps
    del a, b
    return {"bytes": 2 * nbytes, "GBps": grids[g_best], "kernel": "grid-stride float4 copy",
            "grid": g_best, "GBps_by_grid": grids,
            "torch_copy_GBps": round(2 * nbytes / (ms_t * 1e-3) / 1e9, 1)}


# Compulsory bytes of one COLA step by stage (Np particles, Ng = n^3,
# Nk = n^2 (n/2+1)): the paint reads x and writes delta once; the FFT stage
# is one r2c and three Green-multiplied c2r transforms, each reading and
# writing its mesh once (4 Ng + 8 Nk); the readout reads x and three meshes
# and writes f; kick and drift each read and write 60 bytes per particle.
# Where kernels fuse the paint with the z r2c and the three z c2r with the
# readout, a transform's compulsory bytes can no longer be set against the
# time of its own passes (its z pass runs inside the paint or the readout,
# its y and x passes on their own), so the paint, the FFT and the readout
# form one force stage.  The stage bytes always sum to 156 Np + 32 Ng +
# 32 Nk.  The bin count runs inside the kick + drift kernel after the first
# step and is timed there.
def stage_table(prof):
    names = {r[0] for r in prof}
    fused = "paint_tile_r2c" in names or "readout_columns" in names
    kick_drift = {"kick_drift": (("cola_",), lambda Np, Ng, Nk: 120 * Np)}
    if fused:
        return {"paint+fft+readout": (("paint_", "fft_", "readout"),
                                      lambda Np, Ng, Nk: 36 * Np + 32 * Ng + 32 * Nk),
                **kick_drift}
    return {
        "paint": (("paint_",), lambda Np, Ng, Nk: 12 * Np + 4 * Ng),
        "fft": (("fft_",), lambda Np, Ng, Nk: 16 * Ng + 32 * Nk),
        "readout": (("readout",), lambda Np, Ng, Nk: 24 * Np + 12 * Ng),
        **kick_drift,
    }


def stage_roofline(prof, steps, npart, n):
    """Per-stage time per step (sum of the stage's kernels, HIP events) and
    the fraction of 8 TB/s its compulsory bytes reach."""
    Ng, Nk = n**3, n * n * (n // 2 + 1)
    out = {}
    for stage, (prefixes, nbytes) in stage_table(prof).items():
        ms = sum(r[2] for r in prof if r[0].startswith(prefixes)) / steps
        b = nbytes(npart, Ng, Nk)
        out[stage] = {"ms": round(ms, 4), "algorithmic_bytes": b,
                      "frac": round(b / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
                      if ms > 0 and b > 0 else None}
    return out


def main_leg(args, fs, ps):
    """The headline: COLA particle-steps/s of one box on one GPU."""
    c = fs._fs
    print(f"[bench] 2LPT {args.np}^3 box {args.box}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    p = fs.lpt.init(args.np, args.box, 0.1, ps, args.seed, "cola")
    c._synchronize()
    t_ic = time.perf_counter() - t0
    fs.pm.init(args.nc, args.nc / args.np, args.box)
    fs.pm.set_paint_variant("sorted")
    snap = fs.Particles(args.np**3, args.box)
    snap.copy_from(p)

    sched = schedule(0.1, 1.0, args.steps)
    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        a_vel, a_pos = sched[i % len(sched)]
        cola_step(fs, p, a_vel, a_pos)
    c._synchronize()
    p.copy_from(snap)
    c._synchronize()
    del snap

    c._profile_reset()
    c._profile_enable(True)
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for a_vel, a_pos in sched:
        cola_step(fs, p, a_vel, a_pos)
    # the last drift waits for the next density binning; apply it inside the
    # timed region
    c._particles_sync(p._particles)
    c._synchronize()
    T = time.perf_counter() - t0
    prof = c._profile_get()
    c._profile_enable(False)

    # CIC variants timed separately on the final (a = 1, most clustered) state
    paint_ms = {}
    for variant in ("sorted", "atomic", "global"):
        fs.pm.set_paint_variant(variant)
        fs.pm.send_positions(p)
        fs.pm.compute_density(p)
        c._synchronize()
        t1 = time.perf_counter()
        for _ in range(3):
            fs.pm.compute_density(p)
        c._synchronize()
        paint_ms[variant] = round((time.perf_counter() - t1) / 3 * 1e3, 3)
    fs.pm.set_paint_variant("sorted")
    state = None
    if args.cpu_baseline == "on":
        a_x, a_v, _ = p.times
        state = (_records_of(p), a_x, a_v)
    del p

    npart = args.np**3
    kernels = {}
    for name, launches, total_ms, bytes_per in prof:
        if launches == 0:
            continue
        avg = total_ms / launches
        gbps = bytes_per / (avg * 1e-3) / 1e9
        kernels[name] = {"launches": launches, "avg_ms": round(avg, 4),
                         "total_ms": round(total_ms, 3),
                         "algorithmic_GB": round(bytes_per / 1e9, 3),
                         "GBps": round(gbps, 1), "frac_hbm": round(gbps / PEAK_HBM_GBS, 3)}
    dom = max(prof, key=lambda r: r[2])
    dom_avg_s = dom[2] / dom[1] * 1e-3
    achieved = dom[3] / dom_avg_s / 1e9
    # Bytes every launch declares, per step: intermediate FFT passes, the
    # paint's sort traffic and the fused passes' extra meshes included.
    issued_bytes = sum(r[3] * r[1] for r in prof) / args.steps
    stages = stage_roofline(prof, args.steps, npart, args.nc)
    step_bytes = sum(v["algorithmic_bytes"] for v in stages.values())
    headline = (args.np, args.nc, args.box) == (512, 1024, 1024.0)
    return {
        "value": npart * args.steps / T,
        "ms_per_step": T / args.steps * 1e3,
        "config": {"workload": f"COLA {args.np}^3 particles, {args.nc}^3 mesh, box "
                               f"{args.box:g} Mpc/h, {args.steps} steps a=0.1->1",
                   "np": args.np, "nc": args.nc, "box": args.box,
                   "paint": "sorted (deterministic, default)", "ic_seconds": round(t_ic, 3)},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "algorithmic_bytes": dom[3], "avg_ms": round(dom_avg_s * 1e3, 4),
                     "traffic": measured_traffic(dom[0]) if headline else None},
        "step_roofline": {"algorithmic_bytes": step_bytes,
                          "definition": "156 Np + 32 Ng + 32 Nk per step (compulsory traffic)",
                          "achieved_GBps": round(step_bytes / (T / args.steps) / 1e9, 1),
                          "frac": round(step_bytes / (T / args.steps) / 1e9 / PEAK_HBM_GBS, 4),
                          "issued_bytes": issued_bytes,
                          "issued_GBps": round(issued_bytes / (T / args.steps) / 1e9, 1)},
        "stage_roofline": stages,
        "kernels": kernels,
        "paint_variants_ms_final_state": paint_ms,
        "_state": state,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--np", type=int, default=512, help="particles per dimension")
    ap.add_argument("--nc", type=int, default=1024, help="PM mesh per dimension")
    ap.add_argument("--box", type=float, default=1024.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-baseline", default="on", choices=["on", "off"])
    ap.add_argument("--legs", default="main,adjoint,batch",
                    help="comma list of main (headline COLA box), adjoint (differentiable "
                         "128^3/256^3 run), batch (96-box hipGraph ensemble), fof "
                         "(halo finder after a 256^3 COLA run), adjoint_c (the differentiable run at "
                         "512^3/1024^3), cpu_full (the reference's CPU step at A, B and C, 1 "
                         "thread and all cores; minutes); fof, adjoint_c and cpu_full are not in "
                         "the default list")
    args = ap.parse_args()
    legs = set(x.strip() for x in args.legs.split(",") if x.strip())

    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} requested; this build measures one GPU only "
              f"(the multi-GPU slab decomposition is design-only).", file=sys.stderr)
        sys.exit(2)

    from fs2_amd import build as _build
    _build.ensure_built()
    import fs2_amd.fs as fs
    fs.cosmology.init(0.31)
    ps = fs.PowerSpectrum.eisenstein_hu(0.31, 0.049, 0.67, 0.96, 0.81)

    out = {
        "metric": "COLA particle-steps/s, 512^3 p/1024^3 mesh, 1 GPU",
        "value": None,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (2LPT from seeded white noise, EH98 no-wiggle P(k))",
        "config": None,
        "roofline": None,
        "copy_peak": None,
        "adjoint": None,
        "batch": None,
        "cpu_baseline": None,
    }
    if "fof" in legs:
        out["fof"] = None
    def optional(name, fn):
        """Runs a secondary leg; a failure is recorded in the line instead of
        losing the headline measurement."""
        try:
            return fn()
        except Exception as e:   # noqa: BLE001 (reported, not hidden)
            print(f"[bench] {name} failed: {e!r}", file=sys.stderr, flush=True)
            return {"error": f"{type(e).__name__}: {e}"}

    state = None
    if "main" in legs:
        out.update(main_leg(args, fs, ps))
        state = out.pop("_state")
        out["copy_peak"] = optional("copy probe", lambda: copy_probe(fs))
        if "GBps" in out["copy_peak"]:
            out["roofline"]["frac_of_copy_peak"] = round(
                out["roofline"]["achieved"] / out["copy_peak"]["GBps"], 4)
            for k in out["kernels"].values():
                k["frac_copy_peak"] = round(k["GBps"] / out["copy_peak"]["GBps"], 3)
    if "adjoint" in legs:
        print("[bench] adjoint 128^3/256^3", file=sys.stderr, flush=True)
        out["adjoint"] = optional("adjoint", lambda: adjoint_leg(fs, ps))
    if "adjoint_c" in legs:
        out["adjoint_512p_1024mesh"] = optional(
            "adjoint 512^3", lambda: adjoint_leg(fs, ps, reps=2, n=512, nm=1024, L=1024.0))
    if "batch" in legs:
        print("[bench] batched ensemble (96 boxes, hipGraph)", file=sys.stderr, flush=True)
        out["batch"] = optional("batch", lambda: batch_leg(fs, ps))
    if "fof" in legs:
        print("[bench] friends-of-friends 256^3", file=sys.stderr, flush=True)
        out["fof"] = optional("fof", lambda: fof_leg(fs, ps))
    if "cpu_full" in legs:
        print("[bench] cpu baseline, configurations A, B, C", file=sys.stderr, flush=True)
        out["cpu_full"] = optional("cpu full", lambda: cpu_full_leg(fs, ps))
    if state is not None:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = optional(
            "cpu baseline", lambda: cpu_baseline(state, args.np, args.nc, args.box))
        del state
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
