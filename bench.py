"""FP64 6D tracking throughput (particle-turns/s) on one MI355X.

A step is one pass (turn) of the whole beam through the lattice.  Default
workload: the 40-cell storage ring (80 CSBEND with 16 slices, 240 KQUAD,
160 KSEXT, 40 KOCT, drifts, one RFCA at 352 MHz / 6 MV; 1082 elements) with a
matched 1,000,000-particle Gaussian beam, FP64 throughout.  The beam is
resident in HBM before the timed region.  Throughput counts survivors per pass
(elegant's effort, do_tracking.c:2758-2765).

Also reported on the same JSON line:
  roofline      achieved algorithmic FLOP/s of the tracking kernel (FLOPs per
                particle-turn from the reference routines' operation counts x
                particle-turns of the timed launches / summed HIP-event kernel
                time on the library's stream) against MI355X's FP64 peak;
  cpu_baseline  the CPU checker (C11 restatement of the same routines, glibc,
                -O3, no FP contraction) timed on this host's cores on a
                bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed passes")
    ap.add_argument("--warmup", type=int, default=2, help="untimed passes (the first includes the RF fiducial)")
    ap.add_argument("--workload", default="storage_ring",
                    choices=["storage_ring", "fodo_ring", "storage_ring_wake", "error_seeds", "learned_da",
                             "mlp_train"])
    ap.add_argument("--variants", type=int, default=256, help="error_seeds / learned_da: lattice variants")
    ap.add_argument("--particles", type=int, default=0, help="override the workload's particle count")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads for the CPU baseline (0 = the cores allotted to this job, at most 16)")
    ap.add_argument("--survivors", action="store_true",
                    help="report the survivors after every timed pass (from the passes of loss)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk", type=int, default=0, help="passes per kernel launch (0 = library default)")
    ap.add_argument("--block", type=int, default=0, help="threads per workgroup (0 = library default)")
    ap.add_argument("--epochs", type=int, default=6, help="learned_da: training epochs of the network")
    ap.add_argument("--tuning", default="", help="error_seeds: ppl,block,passes_per_launch,use_graph,compact "
                                                 "instead of the autotuner's choice")
    return ap.parse_args()


def workload(name, n):
    from elegant_amd import lattices as Lt
    if name == "fodo_ring":
        bl, Po, rows = Lt.fodo_workload(n=n or 10000)
        desc = "FODO ring, 100 cells (601 elements), %d particles" % rows.shape[0]
    elif name == "storage_ring_wake":
        bl, Po, rows = Lt.storage_ring_wake_workload(n=n or 2_000_000)
        desc = "storage ring 40 cells + CHARGE 5 nC + WAKE + ZLONGIT 8192 bins, %d particles" % rows.shape[0]
    else:
        bl, Po, rows = Lt.storage_ring_workload(n=n or 1_000_000)
        desc = "storage ring 40 cells (80 CSBEND x16 slices, 240 KQUAD, 160 KSEXT, 40 KOCT, RFCA; %d elements), " \
               "%d particles" % (len(bl), rows.shape[0])
    return bl, Po, rows, desc


def host_cores():
    """Cores this job may use: the CPU set the process is allowed to run on,
    at most 16 (the host share that comes with one GPU on the measurement
    box, which may show many more CPUs than are allotted to the job)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            return next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "?")
    except OSError:
        return "?"


def _cpu_run(name, threads, per_thread, passes):
    """The CPU checker (glibc build) on `threads` shards of `per_thread`
    particles of the workload for `passes` passes, one thread per shard (the
    static particle split of Pelegant, do_tracking.c:4736-4800; the shards of
    these workloads never interact).  Returns (particle-turns, seconds)."""
    import numpy as np
    from oracle import oracle_ctypes as O
    bl, Po, rows, _ = workload(name, per_thread * threads)
    shards = [np.ascontiguousarray(rows[k * per_thread:(k + 1) * per_thread]) for k in range(threads)]
    efforts = [0] * threads
    O.lib("glibc")

    def work(k):
        o = O.Oracle("glibc")
        o.do_tracking(shards[k], bl, Po, passes)
        efforts[k] = o.effort()

    ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return sum(efforts), time.perf_counter() - t0


def cpu_baseline(name, threads):
    """The CPU checker timed on this host on bounded samples of the same
    workload: all `threads` cores (the Pelegant-equivalent static split) and
    one core (the serial, elegant-equivalent run).  About 10 s of CPU work
    each for the storage rings, the FODO ring's first 30 passes in full."""
    if name == "fodo_ring":
        per_thread, passes, serial = 10000 // threads, 30, (10000, 10)
    else:
        per_thread, passes, serial = 2048, 16, (2048, 8)
    total, dt = _cpu_run(name, threads, per_thread, passes)
    stotal, sdt = _cpu_run(name, 1, *serial)
    model = _cpu_model()
    return {"value": total / dt, "unit": "particle-turns/s", "cores": threads, "kind": "port",
            "sample": "%d threads x %d particles x %d passes of the same workload (%d particle-turns, %.1f s) "
                      "on %s, nproc %d" % (threads, per_thread, passes, total, dt, model, os.cpu_count() or 0),
            "serial": {"value": stotal / sdt, "unit": "particle-turns/s", "cores": 1,
                       "sample": "1 thread x %d particles x %d passes (%d particle-turns, %.1f s)"
                                 % (serial[0], serial[1], stotal, sdt)}}


def traffic(kernel, workload):
    """HBM bytes per launch of a kernel for a workload, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (profiles/traffic.json), or None"""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t[kernel][workload]
    except (OSError, KeyError, ValueError):
        return None


def error_seeds(args):
    """Batched error-seed study: variants of the storage ring with seeded
    FSE/misalignment/roll errors, each with its own aperture grid (41 rays x
    11 momentum offsets, 1,024-65,536 particles log-uniform by seed), all
    tracked in one persistent graph-launched batch with autotuned
    particles-per-lane, block size and passes per launch."""
    import torch
    from elegant_amd import da, native, studies
    vs = studies.variants(range(args.variants))
    rays = da.Rays(studies.RAYS)
    tr = native.Tracker(0)
    b = native.Batch(tr)
    for v in vs:
        b.add_variant(v.beamline, v.Po, studies.grid_rows(rays, studies.DELTAS, v.nx))
    if args.tuning:
        b.set_tuning(*[int(v) for v in args.tuning.split(",")])
        tun = b.tuning()
    else:
        tun = b.autotune(4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = b.track(0, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    n = sum(b.sizes)
    b.close()
    tr.close()
    return {"metric": "particle-turns/s FP64 6D tracking (batched error-seed study)", "value": st.effort / wall,
            "unit": "particle-turns/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
            "ms_per_step": wall * 1e3 / max(args.steps, 1), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "%d storage-ring error-seed variants, aperture grids 41 rays x 11 offsets, "
                                   "%d particles in total" % (len(vs), n),
                       "variants": len(vs), "particles": n, "passes_timed": args.steps,
                       "survivors": int(st.n_left), "kernel_launches": int(st.kernel_launches),
                       "tuning": {"particles_per_lane": tun.particles_per_lane, "block": tun.block,
                                  "passes_per_launch": tun.passes_per_launch, "use_graph": tun.use_graph,
                                  "compact": tun.compact}}}


def learned_da(args):
    """Learned aperture proposal: train on the first half of the variants'
    full grids, then find the other half's apertures with the network's
    guesses confirmed by FP64 tracking, by bisection alone and by the plain
    grid; DA maps (variant x momentum offset) per hour of device work."""
    from elegant_amd import learned, native
    half = args.variants // 2
    tr = native.Tracker(0)
    rep = learned.study(tr, range(half), range(half, 2 * half), passes=args.steps, epochs=args.epochs,
                        log=lambda m: print(m, file=sys.stderr, flush=True))
    tr.close()
    best = max(("learned", "learned_ladder", "learned_sweep"), key=lambda k: rep[k]["maps_per_hour"])
    return {"metric": "DA maps/hour learned-proposal vs plain", "value": rep[best]["maps_per_hour"],
            "unit": "maps/hour", "n_gpus": 1, "steps": args.steps, "warmup": 0, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "learned aperture proposal: %d training + %d evaluation storage-ring error-seed "
                                   "variants, 41 rays x 11 offsets, %d passes" % (half, half, args.steps),
                       "search": best},
            "plain_grid_maps_per_hour": rep["plain_grid"]["maps_per_hour"],
            "report": rep}


def mlp_train(args):
    """Training throughput of the aperture network (19-256-256-256-1, FP32,
    Adam, minibatches of 4096) on a synthetic training set of the learned
    study's size (1.86 M samples), 6 epochs after a one-epoch warm-up."""
    import numpy as np
    from elegant_amd import learned, native
    tr = native.Tracker(0)
    rng = np.random.default_rng(0)
    n = args.particles or 1_860_000
    X = rng.normal(0, 1, (n, 19)).astype(np.float32)
    y = (X[:, 0] ** 2 + X[:, 1] ** 2 < 1.0).astype(np.float32)
    m = learned.MLP(tr, seed=2)
    m.train(X[:65536], y[:65536], epochs=1, batch=4096)
    t0 = time.perf_counter()
    losses = m.train(X, y, epochs=6, batch=4096, lr=1e-3)
    wall = time.perf_counter() - t0
    m.close()
    tr.close()
    flop = 6 * n * 3 * 2 * (19 * 256 + 256 * 256 * 2 + 256)
    return {"metric": "MLP training samples/s (aperture network)", "value": 6 * n / wall, "unit": "samples/s",
            "n_gpus": 1, "steps": 6, "warmup": 1, "ms_per_step": wall * 1e3 / 6, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "19-256-256-256-1 perceptron, %d samples x 6 epochs, batch 4096" % n},
            "tflops_3x_forward": flop / wall / 1e12, "final_loss": float(losses[-1])}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (no multi-GPU path); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "error_seeds":
        print(json.dumps(error_seeds(args)))
        return
    if args.workload == "learned_da":
        print(json.dumps(learned_da(args)))
        return
    if args.workload == "mlp_train":
        print(json.dumps(mlp_train(args)))
        return
    import torch
    from elegant_amd import lattices as Lt
    from elegant_amd import native

    bl, Po, rows, desc = workload(args.workload, args.particles)
    n = rows.shape[0]
    tr = native.Tracker(0)
    if args.chunk:
        tr.set_chunk_passes(args.chunk)
    if args.block:
        tr.set_block(args.block)
    tr.load(bl, Po)
    tr.upload(rows)
    if args.warmup:
        tr.track(0, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = tr.track(args.warmup, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = st.effort / wall
    flops_pt = Lt.flops_per_particle_turn(bl)
    flops_item = Lt.flops_per_particle_turn_itemised(bl)
    achieved = flops_pt * st.effort / (st.track_ms * 1e-3) / 1e12 if st.track_ms > 0 else None
    line = {
        "metric": "particle-turns/s FP64 6D tracking",
        "value": value,
        "unit": "particle-turns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "particles": n, "elements": len(bl), "passes_timed": args.steps,
                   "survivors": int(st.n_left), "kernel_launches": int(st.kernel_launches)},
        "roofline": {
            "bound": "fp64_valu",
            "note": "FP64 vector ALU bound (no MFMA: nothing here is a dense contraction); MI355X FP64 vector "
                    "peak 78.6 TFLOP/s; algorithmic FLOPs %d per particle-turn by the per-element formulas "
                    "(+,-,*,/,sqrt 1, libm call 20); itemised count of the routines as written %d"
                    % (flops_pt, flops_item),
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": (achieved / FP64_PEAK_TFLOPS) if achieved else None,
            "traffic": traffic("eh_track_kernel", "%s_%d" % (args.workload, n)),
            "kernel_ms_per_launch": st.track_ms / max(st.kernel_launches, 1),
        },
    }
    if args.survivors:
        import numpy as np
        lp, _ = tr.loss_records()
        lost = np.sort(lp[lp >= 0])
        passes = np.arange(args.warmup, args.warmup + args.steps)
        # survivors after pass k: particles not lost on any pass <= k
        line["survivors_after_pass"] = (n - np.searchsorted(lost, passes, side="right")).tolist()
    if args.workload == "storage_ring_wake":
        line["collective_ms"] = st.collective_ms
        line["collective_GBps"] = (Lt.collective_bytes_per_particle() * 2 * st.effort / max(st.collective_ms, 1e-9)
                                   / 1e6)
    tr.close()
    if not args.no_cpu:
        threads = args.cpu_threads or host_cores()
        line["cpu_baseline"] = cpu_baseline(args.workload, threads)
        line["speedup_vs_cpu"] = value / line["cpu_baseline"]["value"]
        line["speedup_vs_serial_cpu"] = value / line["cpu_baseline"]["serial"]["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
