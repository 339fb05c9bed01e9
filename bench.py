#!/usr/bin/env python3
"""Benchmark of the reciprocal-space S(q) path on one MI355X.

Workload (the metric configuration): cubic box L = 113 A, N = 96,000 atoms
(32,000 Si + 64,000 O) on a seeded subset of the 46^3 simple-cubic lattice
with N(0, 0.3^2) jitter per frame, neutron weights, partials on, q_max 10,
500 shells.  Positions are generated from a seed and resident in HBM before
the timed region.

A step is one pass of the path over a batch of --frames-per-step frames
(one captured graph replay; 25 frames by default, a quarter of the 100-frame
trajectory, so the pipeline of each replay fills and drains once per 25
frames).
The headline value is the FP32-spread forward (S(q) + partials, FP64 shell
sums) in frames/s.  Next to it the line reports the plain FP64 (sigma = 2)
forward, the forward + adjoint of the S(q)-matching loss on the same
geometry, per-stage times from HIP events, the HBM roofline of the dominant
kernel and the CPU restatement of the reference FINUFFT path timed on the
host cores.  An "ingest" object times frames read from a LAMMPS dump of the
same workload: the native reader alone, dump -> S(q) end to end, and the
reference's own reader on the same file.

  python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--frames-per-step 25]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_F64_PEAK_TFLOPS = 78.6
# MI355X vector FP64: 256 CUs x 64 fused multiply-adds (2 flops) per clock x 2.4 GHz
FP64_VALU_PEAK_TFLOPS = 78.6
COPY_BYTES = 1 << 30
GRAD_EPS_F32 = 1e-7
ERR_SHELLS = 40          # q < 0.8 at dq = 0.02: the oracle's direct sum finishes in seconds


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames-per-step", type=int, default=25)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip forward + adjoint")
    ap.add_argument("--fp64-frames", type=int, default=10)
    ap.add_argument("--adj-frames", type=int, default=20)
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the batched-trajectory leg (L=80, N=32,000, 1000 frames)")
    ap.add_argument("--c3-frames", type=int, default=1000)
    ap.add_argument("--no-c5", action="store_true",
                    help="skip the large triclinic leg (10^6 atoms, 4 species, X-ray, "
                         "forward and forward + adjoint)")
    ap.add_argument("--c5-frames", type=int, default=6)
    ap.add_argument("--no-ingest", action="store_true",
                    help="skip the dump-reader and dump-to-S(q) measurement")
    ap.add_argument("--ingest-frames", type=int, default=64)
    ap.add_argument("--no-dyn", action="store_true",
                    help="skip the dyn amplitude leg (96k atoms, lattice modes to |q| = 2)")
    ap.add_argument("--no-debye", action="store_true",
                    help="skip the Debye pair-histogram leg (96k atoms, rmax = L/2)")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the per-column error report against the FP64 oracle")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run an elementwise copy of 1 GiB (reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(frames would shard across GPUs with one all-reduce of the shell sums)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from sqcalc_amd import HipStructureFactor
    from sqcalc_amd.synthetic import config

    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
    wl = config("c2")
    fps = args.frames_per_step
    nbatch = max(1, wl.nframes // fps)
    print("# generating %d frames of %d atoms" % (nbatch * fps, wl.natoms), file=sys.stderr,
          flush=True)
    host = wl.frames(0, nbatch * fps)
    pos = torch.tensor(host, dtype=torch.float64, device=dev).contiguous()

    def make(precision, **kw):
        return HipStructureFactor(wl.cell, wl.origin, wl.types, wl.mapping, weight=wl.weight,
                                  norm=wl.norm, qmin=wl.qmin, qmax=wl.qmax, nq=wl.nq,
                                  precision=precision, **kw)

    # ---- FP32-spread forward (headline): one graph replay per step ----
    sf = make("single", max_batch=fps, tune=True)
    info = sf.info
    print("# fp32 plan: m=%s fine=%s ns=%d sigma=%.2f K=%d lines=%d tiles=%s bytes=%.2f GB"
          % (list(info.modes), list(info.fine), info.nspread, info.upsampfac, info.nmodes,
             info.nlines, list(info.tile), info.device_bytes / 1e9), file=sys.stderr, flush=True)
    sf.accumulate_device(pos[0:1])  # the tuner times the tiles on the first frame
    info = sf.info
    print("# tuned tiles %s" % list(info.tile), file=sys.stderr, flush=True)
    sf.reset()
    for w in range(args.warmup):
        sf.accumulate_device(pos[(w % nbatch) * fps:(w % nbatch + 1) * fps])
    torch.cuda.synchronize()
    sf.fetch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        b = k % nbatch
        sf.accumulate_device(pos[b * fps:(b + 1) * fps])  # one graph replay
    sf.fetch()  # synchronises the plan stream
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    frames_per_s = fps * args.steps / (t1 - t0)
    # The same steps again with HIP events around every stage launch on the
    # plan's own stream (event records do not go into the replayed graph, so
    # these steps launch kernel by kernel): per-kernel times for the roofline.
    sf.profile(True)
    t2 = time.perf_counter()
    for k in range(args.steps):
        b = k % nbatch
        sf.accumulate_device(pos[b * fps:(b + 1) * fps])
    stages = sf.stage_times()  # synchronises
    t3 = time.perf_counter()
    sf.profile(False)
    profiled_fps = fps * args.steps / (t3 - t2)
    nbytes = sf.stage_bytes()
    dom = max((k for k in stages if k in nbytes), key=lambda k: stages[k][0])
    dom_ms = stages[dom][0] / stages[dom][1]
    achieved = nbytes[dom] / (dom_ms * 1e-3) / 1e9
    traffic = _committed_traffic(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes": int(nbytes[dom]),
                "avg_ms": round(dom_ms, 4)}
    stage_ms = {k: round(v[0] / v[1], 4) for k, v in stages.items()}
    # the whole frame against the same byte model
    frame_bytes = sum(nbytes[k] for k in ("spread_z", "dft_y", "dft_x_bin"))
    ref_bytes = _full_grid_bytes(info, wl.natoms, 8)
    tables = {}
    if not args.no_errors:
        tables["fp32"] = _one_frame_table(sf, pos[0:1])
    sf.close()

    # ---- plain FP64 forward (sigma 2) ----
    fp64 = None
    if not args.no_fp64:
        sd = make("double", tune=True, max_batch=args.fp64_frames)
        nf = min(args.fp64_frames, pos.shape[0])
        sd.accumulate_device(pos[0:1])  # tunes on the first frame
        sd.accumulate_device(pos[0:nf])  # warm-up of the graph
        sd.fetch()
        t0 = time.perf_counter()
        sd.accumulate_device(pos[0:nf])
        sd.fetch()
        fp64 = nf / (time.perf_counter() - t0)
        fine64 = list(sd.info.fine)
        if not args.no_errors:
            tables["fp64"] = _one_frame_table(sd, pos[0:1])
        sd.close()

    # ---- forward + adjoint (S(q)-matching loss) ----
    adj = adj_graph = grad_err = None
    if not args.no_adjoint:
        from sqcalc_amd.autograd import sq_matching_loss
        # FP32 gradient plans run at eps 1e-7 (ns 12): the kernel derivative's
        # aliasing error sets the gradient's accuracy, 1e-5 relative there
        sa = make("single", partials=False, tune=True, eps=GRAD_EPS_F32)
        tgt_sf = make("single", partials=False)
        tgt_sf.accumulate_device(pos[-1:])
        num, _, _, _ = tgt_sf.fetch()
        tgt_sf.close()
        # target: S(q) of a second seed's frame
        target = torch.tensor(np.where(sa.den > 0, num / np.where(sa.den > 0, sa.den, 1), 0.0),
                              device=dev)
        nf = min(args.adj_frames, pos.shape[0])
        frames = [pos[k].clone().requires_grad_(True) for k in range(nf)]
        loss = sq_matching_loss(frames[0], sa, target)
        loss.backward()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(nf):
            loss = sq_matching_loss(frames[k], sa, target)
            loss.backward()
        torch.cuda.synchronize()
        adj = nf / (time.perf_counter() - t0)
        # the same loss step captured once with torch.cuda.graph and replayed
        # per frame (positions copied into the captured input)
        static = pos[0].clone().requires_grad_(True)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2):
                static.grad = None
                sq_matching_loss(static, sa, target).backward()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        static.grad = None
        with torch.cuda.graph(graph):
            sq_matching_loss(static, sa, target).backward()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(nf):
            with torch.no_grad():
                static.copy_(pos[k])
            graph.replay()
        torch.cuda.synchronize()
        adj_graph = nf / (time.perf_counter() - t0)
        if not args.no_errors:
            grad_err = _gradient_error(sa, make, pos[0], target)
        sa.close()

    # ---- batched trajectory: 1000 frames through graph replays, autotuned tiles ----
    c3 = None
    if not args.no_c3:
        c3 = _batched_trajectory(args.c3_frames, dev)

    # ---- large triclinic cell: 10^6 atoms, 4 species, X-ray, forward + adjoint ----
    c5 = None
    if not args.no_c5:
        c5 = _c5_leg(args.c5_frames, dev)

    # ---- frames from a LAMMPS dump: native reader, and dump -> S(q) end to end ----
    ingest = None
    if not args.no_ingest:
        ingest = _ingest(wl, host, args.ingest_frames, make, not args.no_cpu)

    # ---- dyn: reciprocal-lattice amplitudes on the FP64 matrix cores ----
    dyn = None
    if not args.no_dyn:
        dyn = _dyn_leg(wl, host, dev, not args.no_cpu)

    # ---- Debye: pair-distance histograms per type pair ----
    debye = None
    if not args.no_debye:
        debye = _debye_leg(wl, host, dev, not args.no_cpu)

    # ---- per-column error of both precisions against the FP64 oracle ----
    errors = None
    if tables:
        errors = _column_errors(wl, host[0], tables)

    # ---- CPU restatement of the reference FINUFFT path ----
    cpu = None
    if not args.no_cpu:
        cpu = _cpu_baseline(wl, host[0], int(info.nmodes))

    line = {
        "metric": "S(q) frames/s (fwd, FP32 spread, FP64 shell sums) at 96k atoms",
        "value": round(frames_per_s, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered SiO2 lattice, positions resident in HBM)",
        "config": {"workload": "c2: cubic L=113 A, N=96000 (Si 32000, O 64000), lattice + "
                               "N(0,0.3^2) jitter, neutron, partials, q_max 10, nq 500",
                   "frames_per_step": fps, "natoms": wl.natoms, "modes": list(info.modes),
                   "kept_modes": int(info.nmodes), "fine_grid": list(info.fine),
                   "ns": int(info.nspread), "upsampfac": info.upsampfac},
        "roofline": roofline,
        "frame_hbm_model_gb": round(frame_bytes / 1e9, 3),
        "frame_frac_of_hbm_peak": round(frame_bytes * frames_per_s / 1e9 / HBM_PEAK_GBS, 4),
        "full_grid_model_gb": round(ref_bytes / 1e9, 3),
        "full_grid_equiv_frac": round(ref_bytes * frames_per_s / 1e9 / HBM_PEAK_GBS, 4),
        "stage_ms": stage_ms,
        "profiled_pass_frames_per_s": round(profiled_fps, 2),
        "fp64_plain_frames_per_s": round(fp64, 2) if fp64 else None,
        "fp32_over_fp64": round(frames_per_s / fp64, 2) if fp64 else None,
        "fwd_adj_frames_per_s": round(adj, 2) if adj else None,
        "fwd_adj_graph_frames_per_s": round(adj_graph, 2) if adj else None,
        "fwd_adj_eps": GRAD_EPS_F32,
        "fwd_adj_grad_err_vs_fp64": grad_err,
        "column_max_abs_err_vs_fp64_oracle": errors,
        "dyn": dyn,
        "debye": debye,
        "batched_trajectory": c3,
        "large_triclinic": c5,
        "cpu_baseline": cpu,
        "ingest": ingest,
    }
    print(json.dumps(line), flush=True)
    return 0


def _c5_leg(nframes, dev):
    """Triclinic a = b = c = 250 A (xy 10, xz 5), N = 10^6 of 4 species (Si, O,
    Na, Ca), jittered lattice, X-ray weights, all 10 partials, q_max 4, FP32
    spread and FFT with FP64 shell sums: forward frames/s through graph
    replays, and forward + adjoint frames/s of the S(q)-matching loss on S and
    every pair column (the gradient of the loss per frame)."""
    import numpy as np
    import torch

    from sqcalc_amd import HipStructureFactor
    from sqcalc_amd.autograd import structure_factor
    from sqcalc_amd.synthetic import config
    wl = config("c5")
    print("# generating %d frames of %d atoms (c5)" % (nframes + 1, wl.natoms), file=sys.stderr,
          flush=True)
    host = wl.frames(0, nframes + 1)
    pos = torch.tensor(host, dtype=torch.float64, device=dev).contiguous()
    out = {"workload": "c5: triclinic 250 A (xy 10, xz 5), N=10^6 (Si, O, Na, Ca "
                       "0.25/0.55/0.12/0.08), jittered lattice, xray, 10 partials, q_max 4, FP32",
           "frames": nframes}

    def make(**kw):
        return HipStructureFactor(wl.cell, wl.origin, wl.types, wl.mapping, weight=wl.weight,
                                  norm=wl.norm, qmax=wl.qmax, nq=wl.nq, precision="single", **kw)
    sf = make(tune=True, max_batch=nframes)
    sf.accumulate_device(pos[0:1])
    sf.accumulate_device(pos[1:nframes + 1])  # warm-up of the graph
    sf.fetch()
    t0 = time.perf_counter()
    sf.accumulate_device(pos[1:nframes + 1])
    sf.fetch()
    out["fwd_frames_per_s"] = round(nframes / (time.perf_counter() - t0), 2)
    info = sf.info
    out["fine_grid"] = list(info.fine)
    out["kept_modes"] = int(info.nmodes)
    out["tile"] = list(info.tile)
    sf.close()
    sa = make(tune=True, eps=GRAD_EPS_F32)
    nq = sa.nq
    rng = np.random.default_rng(7)
    target = torch.tensor(rng.random(nq), device=dev)
    npair = sa.info.ncomp - 1
    wP = torch.tensor(rng.standard_normal((npair, nq)) * 1e-3, device=dev)

    def step(k):
        p = pos[k].clone().requires_grad_(True)
        S, P = structure_factor(p, sa)
        loss = torch.sum((S - target) ** 2) + torch.sum(wP * P)
        loss.backward()
        return p.grad

    step(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(1, nframes + 1):
        step(k)
    torch.cuda.synchronize()
    out["fwd_adj_frames_per_s"] = round(nframes / (time.perf_counter() - t0), 2)
    out["fwd_adj_eps"] = GRAD_EPS_F32
    sa.close()
    return out


def _batched_trajectory(nframes, dev):
    """Cubic L = 80 A, N = 32,000 (50/50), random walk sigma 0.05 A per frame,
    unit weights, partials, q_max 8, FP32: all frames resident in HBM, pushed
    through captured graphs of 50 frames with autotuned tiles, and the same
    frames launched kernel by kernel for comparison."""
    import torch

    from sqcalc_amd import HipStructureFactor
    from sqcalc_amd.synthetic import config
    wl = config("c3")
    print("# generating %d frames of %d atoms" % (nframes, wl.natoms), file=sys.stderr,
          flush=True)
    pos = torch.tensor(wl.frames(0, nframes), dtype=torch.float64, device=dev).contiguous()
    out = {"workload": "c3: cubic L=80 A, N=32000 (50/50), random walk 0.05 A, unit, "
                       "partials, q_max 8, FP32", "frames": nframes}
    for key, kw in (("graph", dict(tune=True, max_batch=50)), ("eager", dict(max_batch=1))):
        sf = HipStructureFactor(wl.cell, wl.origin, wl.types, wl.mapping, weight=wl.weight,
                                norm=wl.norm, qmax=wl.qmax, nq=wl.nq, precision="single", **kw)
        sf.accumulate_device(pos[:50])  # warm-up; the tuner runs on the first frame
        sf.fetch()
        sf.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sf.accumulate_device(pos)
        sf.fetch()
        dt = time.perf_counter() - t0
        out[key + "_frames_per_s"] = round(nframes / dt, 1)
        if key == "graph":
            info = sf.info
            out["tile"] = list(info.tile)
            out["fine_grid"] = list(info.fine)
            out["kept_modes"] = int(info.nmodes)
            # the frame against the fused byte model, and whether its working
            # set (W1 + W2) fits the 256 MB Infinity Cache: then the model's
            # "bandwidth" may exceed the HBM peak
            nb = sf.stage_bytes()
            fb = sum(nb[k] for k in ("spread_z", "dft_y", "dft_x_bin"))
            M3 = (info.modes[2] - 1) // 2
            w1 = info.nfields * (M3 + 1) * info.fine[1] * info.fine[0] * 8
            w2 = info.nfields * info.nlines * info.fine[0] * 8
            out["frame_hbm_model_gb"] = round(fb / 1e9, 4)
            out["model_GBps"] = round(fb * nframes / dt / 1e9, 1)
            out["model_frac_of_hbm_peak"] = round(fb * nframes / dt / 1e9 / HBM_PEAK_GBS, 4)
            out["working_set_mb"] = round((w1 + w2) / 1e6, 1)
            out["fits_infinity_cache_256mb"] = (w1 + w2) < 256e6
        sf.close()
    return out


def _full_grid_bytes(info, natoms, b):
    """Bytes per frame of the unfused pipeline (full fine grid written by the
    spreader, pruned row-column FFT with one HBM round trip per pass, kept
    modes stored): S b (2 Gf + 2 m1 n2 n3 + 2 m1 m2 n3 + K) + N (3*8 + 4)."""
    m1, m2, _ = info.modes
    n1, n2, n3 = info.fine
    gf = n1 * n2 * n3
    return (info.nfields * b * (2 * gf + 2 * m1 * n2 * n3 + 2 * m1 * m2 * n3 + info.nmodes)
            + natoms * 28)


def _committed_traffic(kernel):
    """HBM bytes per launch of the kernel from the committed rocprofv3 PMC
    summary (profiles/), when one exists for this workload."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_c2.json")
    try:
        with open(path) as fh:
            return json.load(fh).get(kernel)
    except (OSError, ValueError):
        return None


def _ingest(wl, host, nframes, make, with_reference):
    """A LAMMPS dump of the workload (id type x y z, %.10f, nframes frames of
    the jittered lattice, wrapped): the native reader's rate into a resident
    host batch, the rate of dump -> FP32 S(q) through sqcalc_amd.dump.feed
    (parse of batch k + 1 beside the device work of batch k), and, in the
    CPU-baseline leg, the reference's own reader (src/sqc_dump.f90 built from
    its sources) on the same file."""
    import subprocess
    import tempfile

    import numpy as np

    from sqcalc_amd.dump import DumpReader, feed
    n = wl.natoms
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    L = np.diag(wl.cell)
    row = "%d %d %.10f %.10f %.10f\n"
    tmp = tempfile.mkdtemp(prefix="sqc_ingest_")
    path = os.path.join(tmp, "c2.dump")
    blocks = {}                                # atom text of each distinct frame, formatted once
    with open(path, "w") as fh:
        ids = np.arange(1, n + 1)
        for f in range(nframes):
            h = f % min(len(host), 8)
            if h not in blocks:
                cols = np.empty((n, 5), dtype=object)
                cols[:, 0], cols[:, 1], cols[:, 2:] = ids, wl.types, np.mod(host[h], L)
                blocks[h] = (row * n) % tuple(cols.ravel())
            fh.write("ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n"
                     % (f, n))
            fh.write("".join("%.10f %.10f\n" % (wl.origin[d], wl.origin[d] + L[d])
                             for d in range(3)))
            fh.write("ITEM: ATOMS id type x y z\n")
            fh.write(blocks[h])
    size = os.path.getsize(path)
    buf = np.ones((nframes, n, 3))              # 147 MB at 64 frames of 96,000 atoms
    best = float("inf")
    for _ in range(2):
        t0 = time.perf_counter()
        with DumpReader(path, threads) as rd:
            got, _ = rd.read_frames(buf, n)
        best = min(best, time.perf_counter() - t0)
    assert got == nframes
    sf = make("single", max_batch=8, tune=True)
    sf.accumulate(buf[:1])                    # tuner and graph capture outside the timing
    sf.reset()
    t0 = time.perf_counter()
    done = feed(sf, path, batch=8)          # one worker fewer than the host threads
    e2e = done / (time.perf_counter() - t0)
    sf.close()
    out = {"sample": "%d frames of %d atoms, id type x y z %%.10f, %.1f MB" % (nframes, n, size / 1e6),
           "threads": threads, "feed_threads": max(1, threads - 1),
           "reader_frames_per_s": round(nframes / best, 2),
           "reader_MB_per_s": round(size / best / 1e6, 1),
           "dump_to_sq_frames_per_s": round(e2e, 2)}
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_dump")
    if with_reference and os.path.exists(ref):
        t0 = time.perf_counter()
        subprocess.run([ref, path, os.path.join(tmp, "ref.bin")], check=True,
                       capture_output=True)
        out["reference_reader_frames_per_s"] = round(nframes / (time.perf_counter() - t0), 3)
        out["reference_reader_cores"] = 1
    for name in os.listdir(tmp):
        os.remove(os.path.join(tmp, name))
    os.rmdir(tmp)
    return out


def _one_frame_table(sf, pos1):
    """The plan's S(q) table of one frame (accumulators reset first)."""
    sf.reset()
    sf.accumulate_device(pos1)
    return sf.table()


def _column_errors(wl, frame, tables):
    """Max |d| of every column (S, then the pair columns) of one frame's
    table against the FP64 direct-sum oracle on the lowest ERR_SHELLS shells
    (the shells are the same: both bin from q = 0 with dq = 0.02)."""
    from oracle import oracle as O
    dq = (wl.qmax - wl.qmin) / wl.nq
    run = O.Run(wl.cell, wl.origin, wl.types, wl.mapping, weight=wl.weight, norm=wl.norm,
                qmin=wl.qmin, qmax=wl.qmin + ERR_SHELLS * dq, nq=ERR_SHELLS)
    run.add_frame(frame)
    ref = run.table()
    names = ["S"] + ["S(%s)" % l for l in ("Si-Si", "Si-O", "O-O")]
    out = {"window": "%d shells, q < %.2f, %d modes" % (ERR_SHELLS, ERR_SHELLS * dq,
                                                        int(run.md["hkl"].shape[0]))}
    for key, tab in tables.items():
        out[key] = {n: float("%.3e" % np.max(np.abs(tab[:ERR_SHELLS, c + 1] - ref[:, c + 1])))
                    for c, n in enumerate(names)}
    out["max_abs_ref"] = {n: float("%.4g" % np.max(np.abs(ref[:, c + 1])))
                          for c, n in enumerate(names)}
    return out


def _gradient_error(sa, make, frame, target):
    """The FP32 gradient plan against an FP64 one (eps 1e-13) on one frame,
    max |g32 - g64| / max |g64|, two ways:
      adjoint    the same upstream gradient gS = 2 (S64 - T) into both
                 adjoints (what the hand-written adjoint adds);
      loss       each plan's own S(q)-matching loss sum_s (S_s - T_s)^2,
                 which also carries the FP32 forward's error of S into
                 gS = 2 (S - T): large where |S - T| is small next to S
                 (a target from another frame of the same lattice)."""
    import torch

    from sqcalc_amd.autograd import structure_factor
    sd = make("double", partials=False, eps=1e-13, max_batch=1)
    with torch.no_grad():
        s64, _ = structure_factor(frame, sd)
    gs = (2.0 * (s64 - target)).detach()
    out = {}
    for key, fixed in (("adjoint", True), ("loss", False)):
        g = []
        for plan in (sa, sd):
            x = frame.clone().requires_grad_(True)
            S, _ = structure_factor(x, plan)
            (torch.sum(gs * S) if fixed else torch.sum((S - target) ** 2)).backward()
            g.append(x.grad.cpu().numpy())
        out[key] = float("%.3e" % (np.abs(g[0] - g[1]).max() / np.abs(g[1]).max()))
    with torch.no_grad():
        s32, _ = structure_factor(frame, sa)
    out["max_abs_S_minus_T"] = float("%.3g" % (s64 - target).abs().max().item())
    out["max_abs_S32_minus_S64"] = float("%.3g" % (s32 - s64).abs().max().item())
    sd.close()
    return out


def _dyn_leg(wl, host, dev, with_cpu, qmax=2.0, fps=10, steps=5):
    """The dyn amplitudes rho_a(n, t) of the 96,000-atom workload for every
    reciprocal-lattice vector with 0 < |q| <= qmax (one of each +- pair, the
    --qpoints grid sampling), FP64 on the matrix cores: frames/s with the
    positions resident, the matrix kernel against the FP64 MFMA peak
    (algorithmic flops 8 N nmodes per frame), the max deviation from the
    oracle on a sample of modes, and the oracle (the reference's separable
    phase algorithm restated, all host cores) on that sample, scaled to all
    modes."""
    import torch

    from sqcalc_amd.dynamics import HipDynAmplitudes, grid_modes
    b = 2.0 * np.pi * np.linalg.inv(wl.cell).T
    hkl, _ = grid_modes(wl.cell, qmax)
    dy = HipDynAmplitudes(b, wl.types, hkl)
    info = dy.info
    pos = torch.tensor(host[:fps], dtype=torch.float64, device=dev).contiguous()
    out = dy.rho_device(pos)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dy.rho_device(pos, out=out)
    torch.cuda.synchronize()
    fr = fps * steps / (time.perf_counter() - t0)
    dy.profile(True)
    dy.rho_device(pos, out=out)
    st = dy.stage_times()
    dy.profile(False)
    k_ms = st["dyn_mfma"][0] / st["dyn_mfma"][1]
    alg = 8.0 * len(hkl) * wl.natoms
    res = {"workload": "c2 geometry (96k atoms, 2 species), lattice modes 0 < |q| <= %.1f, "
                       "one of each +- pair, FP64" % qmax,
           "nmodes": int(len(hkl)), "frames_per_s": round(fr, 2),
           "stage_ms": {k: round(v[0] / v[1], 4) for k, v in st.items()},
           "roofline": {"bound": "mfma", "kernel": "k_dyn_rho",
                        "achieved": round(alg / (k_ms * 1e-3) / 1e12, 2),
                        "peak": MFMA_F64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(alg / (k_ms * 1e-3) / 1e12 / MFMA_F64_PEAK_TFLOPS, 4),
                        "executed_TFLOPs": round(8.0 * info.macs_per_frame / (k_ms * 1e-3) / 1e12, 2),
                        "avg_ms": round(k_ms, 4), "algorithmic_flops": alg}}
    rho = out[0].cpu().numpy()
    dy.close()
    pick = np.sort(np.random.default_rng(7).choice(len(hkl), 4000, replace=False))
    from oracle import oracle as O
    sp = np.searchsorted(np.unique(wl.types), wl.types).astype(np.int32)
    t1 = time.perf_counter()
    ref = O.dyn_rho(host[0], sp, len(np.unique(wl.types)), b, hkl[pick])
    dt = time.perf_counter() - t1
    res["max_abs_err_vs_oracle"] = float("%.3e" % np.abs(rho[:, pick] - ref).max())
    res["max_abs_rho"] = float("%.4g" % np.abs(ref).max())
    if with_cpu:
        per_frame = dt * len(hkl) / len(pick)
        res["cpu_baseline"] = {"value": round(1.0 / per_frame, 4), "unit": "frames/s",
                               "cores": O.lib().sqo_num_threads(), "kind": "port",
                               "sample": "all %d atoms x %d of %d modes in %.2f s (separable "
                                         "phase tables, src/sqc_dynamics.f90:826-851), scaled "
                                         "to all modes" % (wl.natoms, len(pick), len(hkl), dt)}
    return res


def _debye_leg(wl, host, dev, with_cpu, fps=4, steps=3, dr=0.01):
    """Debye pair histograms of the 96,000-atom workload at the reference's
    default cut-off (half the box side, 56.5 A, dr 0.01 A: 5650 bins x 3 type
    pairs), FP64, positions resident: frames/s, the pair kernel's rate in
    pairs/s and in FP64 operations against the vector FP64 peak (28 per pair:
    the displacement pos_j - pos_i + bins . shift, its square, the root and
    the scale to a bin, src/sqc_debye.f90:216-221), a full-size check (every
    unordered pair once: counts symmetric, total near the ideal-gas
    estimate), and the oracle's pair loop (the reference's arithmetic,
    all host cores) on a sample of first atoms, scaled to the frame."""
    import torch

    from sqcalc_amd.debye import HipDebye
    deb = HipDebye(wl.cell, wl.origin, wl.types, mapping=wl.mapping, rmax=None, dr=dr,
                   weight=wl.weight)
    info = deb.info
    pos = torch.tensor(host[:fps], dtype=torch.float64, device=dev).contiguous()
    deb.accumulate_device(pos[:1])
    deb.histogram()
    deb.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        deb.accumulate_device(pos)
    h, nf = deb.histogram()
    fr = fps * steps / (time.perf_counter() - t0)
    deb.profile(True)
    deb.accumulate_device(pos)
    deb.histogram()
    st = deb.stage_times()
    deb.profile(False)
    k_ms = st["pairs"][0] / st["pairs"][1]
    n = wl.natoms
    pairs = n * (n - 1) // 2
    ops = 28.0 * pairs
    vol = abs(float(np.linalg.det(wl.cell)))
    ideal = n * (n - 1) * 4.0 / 3.0 * np.pi * deb.rmax ** 3 / vol
    res = {"workload": "c2 geometry (96k atoms, L=113 A, 2 types 1:2), rmax = L/2 = %.2f A, dr "
                       "%.3g A, %d bins x 3 type pairs, FP64" % (deb.rmax, dr, deb.nbins),
           "frames_per_s": round(fr, 2), "pairs_per_frame": pairs,
           "tile": int(info.tile), "workgroups": int(info.workgroups),
           "hist_in_lds": int(info.hist_in_lds),
           "stage_ms": {k: round(v[0] / v[1], 4) for k, v in st.items()},
           "pairs_per_s": float("%.4g" % (pairs / (k_ms * 1e-3))),
           "roofline": {"bound": "valu", "kernel": "k_debye_pairs", "unit": "TFLOP/s",
                        "achieved": round(ops / (k_ms * 1e-3) / 1e12, 2),
                        "peak": FP64_VALU_PEAK_TFLOPS,
                        "frac": round(ops / (k_ms * 1e-3) / 1e12 / FP64_VALU_PEAK_TFLOPS, 4),
                        "avg_ms": round(k_ms, 4), "ops_per_pair": 28},
           "check": {"symmetric": bool(np.array_equal(h, h.transpose(1, 0, 2))),
                     "ordered_pairs_per_frame": float(h.sum() / nf),
                     "ideal_gas_estimate": float("%.6g" % ideal)}}
    deb.close()
    if with_cpu:
        from oracle import oracle as O
        run = O.DebyeRun(wl.cell, wl.origin, wl.types, int(np.max(wl.types)), rmax=deb.rmax,
                         dr=dr)
        rows = 8192
        t1 = time.perf_counter()
        run.rows(host[0], 0, rows)
        dt = time.perf_counter() - t1
        per_frame = dt * n / rows
        res["cpu_baseline"] = {"value": round(1.0 / per_frame, 5), "unit": "frames/s",
                               "cores": O.lib().sqo_num_threads(), "kind": "port",
                               "sample": "%d of %d first atoms against all atoms in %.2f s (the "
                                         "pair loop of debye_accumulate over every image within "
                                         "rmax), scaled to the frame" % (rows, n, dt)}
    return res


def _cpu_baseline(wl, frame, nmodes):
    """One frame of the CPU restatement of the reference FINUFFT path (sigma
    1.25, ES kernel, eps 1e-9, FP64, scipy pocketfft), all host cores; and the
    FP64 direct sum (the reference's --method direct, src/sqc_structure_factor.f90:1398-1515)
    on every atom and a 22^3 block of modes, scaled linearly to the frame's
    kept modes (its cost is atoms x modes)."""
    import numpy as np

    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0))
    run = O.Run(wl.cell, wl.origin, wl.types, wl.mapping, weight=wl.weight, norm=wl.norm,
                qmin=wl.qmin, qmax=wl.qmax, nq=wl.nq)
    t0 = time.perf_counter()
    O.nufft_frame(run, frame, eps=1e-9, upsampfac=1.25, nthreads=cores)
    dt = time.perf_counter() - t0
    side = 22
    h = np.stack(np.meshgrid(*[np.arange(1, side + 1)] * 3, indexing="ij"), -1).reshape(-1, 3)
    b = O.reciprocal(wl.cell)
    md = {"hkl": np.ascontiguousarray(h, dtype=np.int32),
          "qvec": np.ascontiguousarray(h @ b.T), "qlen": np.zeros(len(h))}
    t1 = time.perf_counter()
    O.direct_rho(frame, wl.cell, wl.origin, run.species_of, len(run.species_type), md)
    dd = time.perf_counter() - t1
    per_frame = dd * nmodes / len(h)
    return {"value": round(1.0 / dt, 4), "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": "1 frame of the same workload in %.1f s: the CPU restatement of the "
                      "reference FINUFFT path in FP64 (sigma 1.25, eps 1e-9, ES kernel with the "
                      "cuFINUFFT width and shape rule, sorted sub-problem spreading in C with "
                      "OpenMP, scipy pocketfft, deconvolution, serial partials and shell sums)"
                      % dt,
            "direct_sum": {"frames_per_s": round(1.0 / per_frame, 6), "seconds_per_frame":
                           round(per_frame, 1),
                           "sample": "%d atoms x %d of %d kept modes in %.2f s, scaled to all "
                                     "modes" % (wl.natoms, len(h), nmodes, dd)}}


if __name__ == "__main__":
    sys.exit(main())
