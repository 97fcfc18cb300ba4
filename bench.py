#!/usr/bin/env python3
"""Throughput of the PO diffraction hot path on one MI355X.

Metric: beam.direction evaluations per second -- one unit is one (beam,
theta, phi) evaluation of the aperture integral, Jones rotation and coherent
sum, counted as sum over orientations of beams x directions.

Workload (the orientation-average configuration): hexagonal ice column
D = 50 um, L = 100 um, lambda = 0.532 um, m = 1.3116 + 1e-4 i; grid of 541
theta rows (0-10 deg at 0.05 deg, 10.5-180 deg at 0.5 deg) x 96 phi; FP32
storage with FP64 phase arguments and polygon moments (the mixed kernel), the
FP64 kernel reported beside it.  One step = one pass of the hot path over one
resident batch of ORIENT_PER_STEP orientations, each with a beam count drawn
uniformly from 2000-15000.  Beams are synthetic: resampled from the prepared
beams the reference's tracer produced for a D = 10 um column of the same
shape, rigidly rotated and scaled by 5 (mbs-fast_amd/synthetic.py).

Timing: inputs are uploaded once before the timed region; each step is one
mbs_diffract (main kernel + deterministic reduction).  The roofline figure
uses the main kernel's duration measured with HIP events recorded on the
engine's own stream (mbs_get_timing).

HBM traffic per launch ("roofline.traffic") is read from the committed
rocprofv3 summary TRAFFIC_FILE (FETCH_SIZE + WRITE_SIZE of the main kernel on
this workload, each scaled by what the same counter reads for a plain device
copy of known size, `bench.py --calibrate-copy`); counters cannot be read from
inside the timed run.

Prints one JSON line.  `--gpus N` with N > 1 exits: this build runs on one
GPU (orientations would shard across GPUs, each GPU holding its own Mueller
grid, with one sum-reduce of the grids at the end).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ORIENT_PER_STEP = 16
BEAMS_MIN, BEAMS_MAX = 2000, 15000
# MI355X published vector-ALU peaks (FLOP/s).  The mixed kernel's aperture
# integral runs FP32 on the vector ALU (its phase arguments FP64); its bound
# is the FP32 vector peak.  Without packed FP32 (not used by this build) an
# FP32 FMA issues at the FP64 rate, 78.6 TFLOP/s.
FP32_PEAK = 157.3e12
FP64_PEAK = 78.6e12
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "diffract_hbm_traffic_d50_541x96.txt")


def measured_traffic():
    """{"mixed": bytes, "fp64": bytes} per launch from TRAFFIC_FILE's
    'traffic_bytes <kernel> <bytes>' lines, or {} if it is absent."""
    out = {}
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes":
                    out[parts[1]] = float(parts[2])
    return out


def calibrate_copy():
    """Plain device-to-device copy of 1 GiB, three times (torch's elementwise
    copy kernel): the reference launch for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = 1 << 30
    a = torch.zeros(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n, "launches": 3}))


def algorithmic_bytes(batch, cells):
    """HBM bytes one launch must move at least: the device beam records read
    once (FP64 header 320 B, FP32 header 80 B, MFMA operand 256 B per beam;
    vertex records 32 + 16 + 4 B per vertex), the grid tables read once
    (15 FP64 per direction) and the Mueller grid written once (16 FP64 per
    direction)."""
    n_vert = int(batch.vert_offsets[-1])
    return float(batch.n_beams) * (320 + 80 + 256) + n_vert * 52.0 + cells * (15 + 16) * 8.0


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def unit_flops(batch):
    """Algorithmic FLOPs of one (beam, direction) unit per beam:
    169 + 4 nv + 7 nE (FMA = 2, add/mul/rcp = 1), sincos and reciprocals
    excluded; nE = mean of the two valid-edge counts (the lane uses one set,
    chosen by |A| vs |B|)."""
    nv = batch.beam_nv.astype(np.float64)
    f = batch.edge_flags
    ex = np.add.reduceat((f & 1).astype(np.float64), batch.vert_offsets[:-1])
    ey = np.add.reduceat(((f >> 1) & 1).astype(np.float64), batch.vert_offsets[:-1])
    ne = 0.5 * (ex + ey)
    return 169.0 + 4.0 * nv + 7.0 * ne, ne


REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def _ref_cmd(binary, n_orient, threads, out, tgrid):
    return [os.path.join(REF_DIR, binary), "--method", "po", "--particle", "1", "100", "50",
            "--refractive-index", "1.3116", "1e-4", "--wavelength-um", "0.532",
            "--max-reflections", "10", "--sobol-seed", str(n_orient), "1",
            "--theta-grid-file", tgrid, "--phi-points", "96", "--cutoff-profile", "off",
            "--threads", str(threads), "--close", "--output", out]


def cpu_baseline_reference(budget_s=12.0):
    """Times the reference's own CPU double path (oracle/_ref/mbs_po, built
    from the reference sources with AVX2/FMA, OpenMP on all host threads) on
    the same particle and grid: Sobol orientations (seed 1) of the D=50 um
    column, phase-2 (diffraction) seconds as the reference prints them
    (src/scattering/TracerPOTotal.cpp:6023-6025).  Units come from the
    reference's own prepared beams of the same command, dumped by
    oracle/_ref/mbs_po_dump."""
    import re
    import shutil
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_fixtures
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or os.cpu_count() or 1
    work = tempfile.mkdtemp(prefix="mbs_cpu_")
    try:
        th = [round(0.05 * i, 2) for i in range(201)] + [10.5 + 0.5 * i for i in range(340)]
        tgrid = os.path.join(work, "theta.grid")
        with open(tgrid, "w") as f:
            f.write("\n".join("%.10g" % t for t in th) + "\n")
        # the reference accepts power-of-two Sobol counts only
        n = 1 << max(2, (threads - 1).bit_length())
        while True:
            d = os.path.join(work, "n%d" % n)
            os.makedirs(d)
            p = subprocess.run(_ref_cmd("mbs_po", n, threads, "cpu", tgrid) + ["--backend", "cpu"],
                               cwd=d, stdin=subprocess.DEVNULL, capture_output=True, text=True,
                               timeout=600)
            m = re.search(r"Phase 2 \(diffraction, OpenMP\): ([0-9.]+) s", p.stdout)
            if p.returncode != 0 or not m:
                raise RuntimeError("reference CPU run failed (rc %d): %s %s"
                                   % (p.returncode, p.stdout[-300:], p.stderr[-300:]))
            t2 = float(m.group(1))
            if t2 >= budget_s / 3 or n >= 256:
                break
            grow = max(2.0, budget_s / 2 / max(t2, 0.05))
            n = min(256, n << max(1, int(np.ceil(np.log2(grow)))))
        dump = os.path.join(d, "dump.bin")
        env = dict(os.environ, MBS_FIXTURE_DUMP=dump, MBS_FIXTURE_NO_EXPECTED="1")
        subprocess.run(_ref_cmd("mbs_po_dump", n, threads, "dmp", tgrid) + ["--backend", "gpu"],
                       cwd=d, stdin=subprocess.DEVNULL, capture_output=True, env=env,
                       timeout=900, check=True)
        fx = make_fixtures.read_dump(dump)
        beams = int(fx["orient_offsets"][-1])
        units = beams * 541 * 96
        return {"value": units / t2, "unit": "beam.dir evals/s", "cores": threads,
                "kind": "reference",
                "sample": "reference CLI, %d Sobol orientations of the D=50 um column "
                          "(%d beams, %.1f per orientation) x 541x96 directions = %.3g units; "
                          "phase-2 (diffraction) %.2f s" % (n, beams, beams / n, units, t2)}, fx
    finally:
        shutil.rmtree(work, ignore_errors=True)


def bench_traced(eng, abi, fx, args):
    """The engine on the reference tracer's own beams of the CPU baseline's
    command (the same orientations, beams and grid -- the handler's own, from
    the dump -- that the reference CLI just diffracted on the host):
    autotuned mixed precision, graph replay."""
    batch = abi.BeamBatch.from_fixture(fx)
    consts = abi.consts_from_fixture(fx)
    grid = abi.Grid.from_fixture(fx)
    units = float(batch.n_beams) * grid.cells
    hip = Hip()
    ctx = eng.Context(0)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    ctx.set_timing(True)
    ctx.autotune(abi.PREC_MIXED)
    ks = []
    for _ in range(max(1, args.warmup)):
        ctx.diffract(abi.PREC_MIXED)
    for _ in range(min(args.steps, 5)):
        ctx.diffract(abi.PREC_MIXED)
        ks.append(ctx.timing()["kernel_ms"])
    ctx.set_graph(True)
    ctx.diffract(abi.PREC_MIXED)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.diffract(abi.PREC_MIXED)
    hip.sync()
    wall = time.perf_counter() - t0
    ctx.close()
    return {"value": units * args.steps / wall, "unit": "beam.dir evals/s",
            "ms_per_step": 1e3 * wall / args.steps, "kernel_ms": float(np.mean(ks)),
            "orientations": int(batch.n_orient), "beams": int(batch.n_beams),
            "units_per_step": units}


def bench_config3(pkg, args):
    """The batched non-convex aggregate at its stated batch (opt-in,
    --config3): 16,384 orientations with beam counts log-uniform over
    50..40,000 (a synthetic resample of the 8-column aggregate's traced beams,
    tests/golden/config3_aggregate.npz; ~1e8 beams, ~130 GB of device
    records) on configuration 2's 541 x 96 grid.  One eager beam-segment
    launch, then the persistent work-queue launch captured as a graph and
    replayed twice; every replay must equal the eager result bit for bit."""
    abi, syn, eng = pkg.abi, pkg.synthetic, pkg.engine
    fx = dict(np.load(os.path.join(ROOT, "tests", "golden", "config3_aggregate.npz")))
    src, consts = abi.BeamBatch.from_fixture(fx), abi.consts_from_fixture(fx)
    n_orient = 16384
    rng = np.random.default_rng(7)
    counts = np.exp(rng.uniform(np.log(50), np.log(40000), n_orient)).astype(np.int64)
    t0 = time.perf_counter()
    parts = []
    for i, chunk in enumerate(np.array_split(counts, 16)):  # with a progress line each
        parts.append(syn.resample(src, chunk, seed=7 + i))
        print("config 3: built %d of 16 parts" % (i + 1), file=sys.stderr, flush=True)
    batch = syn.concat(parts, np.full(n_orient, 1.0 / n_orient))
    del parts
    build_s = time.perf_counter() - t0
    n_beams = int(batch.n_beams)
    print("config 3: %d orientations, %d beams built in %.1f s" % (n_orient, n_beams, build_s),
          file=sys.stderr, flush=True)
    grid = syn.config2_grid()
    units = float(n_beams) * grid.cells
    hip = Hip()
    ctx = eng.Context(0)
    ctx.set_grid(grid)
    t0 = time.perf_counter()
    ctx.upload(batch, consts)
    hip.sync()
    upload_s = time.perf_counter() - t0
    del batch
    print("config 3: uploaded in %.1f s" % upload_s, file=sys.stderr, flush=True)
    ctx.set_timing(True)
    ctx.set_target_blocks(131072)
    ctx.set_work_mode(abi.WORK_SEGMENTS)
    t0 = time.perf_counter()
    want = ctx.mueller(abi.PREC_MIXED)[0]
    seg_wall = time.perf_counter() - t0
    seg_kernel = ctx.timing()["kernel_ms"]
    print("config 3: beam segments %.1f s (kernel %.1f ms)" % (seg_wall, seg_kernel),
          file=sys.stderr, flush=True)
    ctx.set_work_mode(abi.WORK_QUEUE)
    ctx.set_graph(True)
    walls, same = [], True
    for rep in range(3):  # the first call captures the graph
        hip.sync()
        t0 = time.perf_counter()
        ctx.diffract(abi.PREC_MIXED)
        hip.sync()
        walls.append(time.perf_counter() - t0)
        got = np.zeros_like(want)
        ctx.accumulate(got)  # the last result
        same = same and bool(np.array_equal(got, want))
        print("config 3: work-queue graph call %d, %.1f s, equal to the segments: %s"
              % (rep, walls[-1], same), file=sys.stderr, flush=True)
    captures, launches = ctx.graph_stats()
    ctx.set_graph(False)
    ctx.close()
    replay = float(np.mean(walls[1:]))
    return {"config3": {
        "workload": "8-column hexagonal aggregate (seed 7, D~U(5,30) um, L/D~U(1,4)), "
                    "m=1.3116+1e-4i, 6 reflections; 16384 orientations, beams per "
                    "orientation log-uniform 50..40000 (synthetic resample of the traced "
                    "beams); 541 theta x 96 phi",
        "orientations": n_orient, "beams": n_beams, "units": units,
        "value": units / replay, "unit": "beam.dir evals/s",
        "graph_replay_s": replay, "segments_wall_s": seg_wall, "segments_kernel_ms": seg_kernel,
        "graph_captures": captures, "graph_launches": launches,
        "replays_equal_segments": same, "batch_build_s": build_s, "upload_s": upload_s}}


def cpu_baseline(pkg, batch, grid, consts, budget_s=12.0):
    """Times the FP64 CPU oracle (OpenMP, all host threads) on a bounded
    sample: the first beams of the batch's first orientation on the full grid,
    growing the sample until one run takes ~1/3 of the budget."""
    from oracle import oracle
    n = 250
    last = None
    spent = 0.0
    while True:
        sub = pkg.synthetic.take_beams(batch, 0, n)
        t0 = time.perf_counter()
        oracle.diffract(grid, sub, consts)
        dt = time.perf_counter() - t0
        spent += dt
        last = (n, dt)
        if dt > budget_s / 3 or spent > budget_s or n >= batch.n_beams:
            break
        n = min(batch.n_beams, int(n * max(2.0, (budget_s / 3) / max(dt, 1e-3))))
    n, dt = last
    units = n * grid.cells
    return {"value": units / dt, "unit": "beam.dir evals/s", "cores": oracle.threads(),
            "kind": "port",
            "sample": "%d beams of one orientation x %d directions (%.3g units), FP64, "
                      "%.2f s" % (n, grid.cells, units, dt)}


class Hip:
    """Minimal HIP runtime access for device synchronisation."""

    def __init__(self):
        import ctypes
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.L = ctypes.CDLL(name)
                break
            except OSError:
                continue
        else:
            raise RuntimeError("libamdhip64.so not found")

    def sync(self):
        rc = self.L.hipDeviceSynchronize()
        if rc != 0:
            raise RuntimeError("hipDeviceSynchronize: %d" % rc)


FFT_N, FFT_BATCH, FFT_PAD = 2048, 64, 4.0


def bench_aperture_fft(ctx, batch, args):
    """Aperture FFT of FFT_BATCH beams on FFT_N^2 FP32 grids spanning FFT_PAD
    times each polygon's extent: raster + row pass (writes the rows whose
    pixel centres fall inside the polygon's y range, 1/FFT_PAD of them; the
    others are zero and skipped) and column pass (reads those rows, writes the
    whole n^2 x 8 B grid); HBM-bound, reported against 8 TB/s."""
    beams = np.arange(FFT_BATCH) % batch.n_beams
    ctx.set_timing(True)
    for _ in range(max(1, args.warmup)):
        ctx.aperture_fft(beams, FFT_N, pad=FFT_PAD)
    rows, cols = [], []
    for _ in range(args.steps):
        ctx.aperture_fft(beams, FFT_N, pad=FFT_PAD)
        t = ctx.aperture_fft_timing()
        rows.append(t["rows_ms"])
        cols.append(t["cols_ms"])
    grid_bytes = float(FFT_N) * FFT_N * 8 * FFT_BATCH
    # rows i with |2i + 1 - n| <= n / pad carry the polygon (the grid is
    # scaled to the extent, so the count is the same for every beam)
    i = np.arange(FFT_N)
    nz = float(np.count_nonzero(np.abs(2 * i + 1 - FFT_N) <= FFT_N / FFT_PAD)) / FFT_N
    r, c = float(np.mean(rows)), float(np.mean(cols))
    rows_bytes, cols_bytes = nz * grid_bytes, (nz + 1.0) * grid_bytes
    return {"n": FFT_N, "batch": FFT_BATCH, "pad": FFT_PAD, "dtype": "c64",
            "nonzero_row_fraction": nz,
            "grids_per_s": FFT_BATCH / ((r + c) * 1e-3),
            "rows_ms": r, "cols_ms": c,
            "rows_GBps": rows_bytes / (r * 1e-3) / 1e9,
            "cols_GBps": cols_bytes / (c * 1e-3) / 1e9,
            "roofline": {"bound": "hbm",
                         "achieved": (rows_bytes + cols_bytes) / ((r + c) * 1e-3) / 1e9,
                         "peak": 8000.0, "unit": "GB/s",
                         "frac": (rows_bytes + cols_bytes) / ((r + c) * 1e-3) / 8e12,
                         "algorithmic_bytes": rows_bytes + cols_bytes,
                         "traffic": measured_traffic().get("fft")}}


def bench_near_forward(eng, abi, args):
    """The large-particle near-forward workload (hexagonal column D=200 um,
    L=400 um, m=1.3116+1e-3i, 12 reflections; the reference tracer's beams of
    4 Sobol orientations, tests/golden/config5_near_forward.npz; theta 0-5 deg
    at 0.005 deg x 48 phi) through both F sources: the direct integral for
    every beam, and the spectra (mbs_spectrum_source) for the near-forward
    beams.  The spectrum path's time per call is the source's (scan, raster +
    FFT, sampling; it depends on batch and grid only) plus its diffraction
    kernel.  The crossover is the number of grid cells per selected beam at
    which the source's fixed cost per beam equals what it saves per cell."""
    fx = dict(np.load(os.path.join(ROOT, "tests", "golden", "config5_near_forward.npz")))
    batch = abi.BeamBatch.from_fixture(fx)
    consts = abi.consts_from_fixture(fx)
    theta = np.radians(np.round(np.arange(0, 1001) * 0.005, 10))
    grid = abi.Grid.build(theta, 48)
    units = float(batch.n_beams) * grid.cells
    ctx = eng.Context(0)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    ctx.set_timing(True)

    def timed(flags):
        for _ in range(max(1, args.warmup)):
            ctx.diffract(flags)
        ks = []
        for _ in range(max(3, min(args.steps, 10))):
            ctx.diffract(flags)
            ks.append(ctx.timing()["kernel_ms"])
        return float(np.median(ks))

    direct_ms = timed(abi.PREC_MIXED)
    info = ctx.spectrum_source(a_max=0.1, oversample=8.0)
    source_ms = info["scan_ms"] + info["fft_ms"] + info["sample_ms"]
    spec_ms = timed(abi.PREC_MIXED | abi.SPECTRUM)
    ctx.close()
    n_sel = max(info["n_selected"], 1)
    saved_per_cell = (direct_ms - spec_ms) / (n_sel * grid.cells)    # ms per (beam, cell)
    sample_per_cell = info["sample_ms"] / (n_sel * grid.cells)
    fixed_per_beam = (info["scan_ms"] + info["fft_ms"]) / n_sel
    gain = saved_per_cell - sample_per_cell
    return {
        "workload": "hexagonal column D=200 um L=400 um, m=1.3116+1e-3i, 12 reflections, "
                    "4 orientations (%d traced beams), theta 0-5 deg at 0.005 deg x 48 phi"
                    % batch.n_beams,
        "direct": {"value": units / (direct_ms * 1e-3), "kernel_ms": direct_ms},
        "spectrum": {"value": units / ((source_ms + spec_ms) * 1e-3),
                     "kernel_ms": spec_ms, "source_ms": source_ms,
                     "selected_beams": info["n_selected"],
                     "selected_area_share": info["selected_area"] / info["total_area"],
                     "fft_sides": {str(k): v for k, v in info["n_by_size"].items() if v},
                     "spectrum_MB": info["spectrum_bytes"] / 1e6,
                     "scan_ms": info["scan_ms"], "fft_ms": info["fft_ms"],
                     "sample_ms": info["sample_ms"]},
        "unit": "beam.dir evals/s",
        "crossover_cells_per_beam": (fixed_per_beam / gain) if gain > 0 else None,
        "note": "crossover: cells per selected beam above which the spectrum source would "
                "pay for itself (null: sampling a cell costs more than the direct integral)",
    }


def bench_size_scan(eng, abi, syn, args):
    """The configuration-5 size scan: 16 sizes, D = 10..200 um log-spaced,
    1,024 orientations each (a synthetic resample of the traced D=200 um beams:
    per-orientation counts drawn from the traced orientations', beams rotated
    at random), theta 0-5 deg at 0.005 deg x 48 phi.  One upload; every size
    re-derived on the device (mbs_rescale_beams: vertices and centres times
    s, areas times s^2, phases at the scaled path).  Per size: mixed and FP64
    evals/s from the kernel times (HIP events).  The FP64-oracle error per
    size is tests/test_config5_gpu.py's."""
    fx = dict(np.load(os.path.join(ROOT, "tests", "golden", "config5_near_forward.npz")))
    src = abi.BeamBatch.from_fixture(fx)
    consts = abi.consts_from_fixture(fx)
    rng = np.random.default_rng(5)
    counts = rng.choice(np.diff(src.orient_offsets), 1024)
    batch = syn.resample(src, counts, seed=5)
    theta = np.radians(np.round(np.arange(0, 1001) * 0.005, 10))
    grid = abi.Grid.build(theta, 48)
    nb = batch.n_beams
    units = float(nb) * grid.cells
    ctx = eng.Context(0)
    ctx.set_grid(grid)
    ctx.upload(batch, consts)
    # scaling data that keeps J_phased as uploaded (raw J = J_phased, path 0,
    # sign +1, no absorption paths)
    ctx.upload_scaling(batch.beam_scalars[:, abi.B_JP:abi.B_JP + 8], np.zeros(nb), np.ones(nb),
                       np.zeros(nb + 1, np.int64), np.zeros(0), 0.0, consts.wave_index)
    ctx.set_timing(True)
    rows = []
    for d in np.geomspace(10.0, 200.0, 16):
        ctx.rescale(d / 200.0)
        per = {}
        for name, flags, reps in (("mixed", abi.PREC_MIXED, 2), ("fp64", abi.PREC_FP64, 1)):
            ctx.diffract(flags)
            ks = []
            for _ in range(reps):
                ctx.diffract(flags)
                ks.append(ctx.timing()["kernel_ms"])
            per[name] = units / (float(np.mean(ks)) * 1e-3)
        rows.append({"D_um": round(float(d), 2), "mixed": per["mixed"], "fp64": per["fp64"]})
    ctx.close()
    return {"workload": "hexagonal column D=10..200 um (16 sizes, log-spaced), m=1.3116+1e-3i, "
                        "12 reflections; 1024 orientations, %d beams; theta 0-5 deg at 0.005 deg "
                        "x 48 phi; device rescale per size" % nb,
            "unit": "beam.dir evals/s", "units_per_size": units, "sizes": rows}


def run(args):
    import mbs_loader
    pkg = mbs_loader.load()
    abi, syn, eng = pkg.abi, pkg.synthetic, pkg.engine
    fx = dict(np.load(os.path.join(ROOT, "tests", "golden", "sobol_average.npz")))
    src = abi.BeamBatch.from_fixture(fx)
    consts = abi.consts_from_fixture(fx)
    rng = np.random.default_rng(args.seed)
    counts = rng.integers(BEAMS_MIN, BEAMS_MAX + 1, ORIENT_PER_STEP)
    batch = syn.resample(src, counts, scale=5.0, seed=args.seed)
    grid = syn.config2_grid()
    flops_b, ne = unit_flops(batch)
    flops_launch = float(flops_b.sum()) * grid.cells
    alg_bytes = algorithmic_bytes(batch, grid.cells)
    units = float(batch.n_beams) * grid.cells
    hip = Hip()
    ctx = eng.Context(0)
    ctx.set_grid(grid)
    ctx.set_timing(True)
    t_up = time.perf_counter()
    ctx.upload(batch, consts)  # H2D copy from pageable memory + pack, outside the timed region
    hip.sync()
    upload_ms = 1e3 * (time.perf_counter() - t_up)
    # the same batch through the page-locked staging (mbs_stage_beams +
    # mbs_upload_staged): host fill of the staging, asynchronous copies, pack
    ctx.upload_staged(batch, consts)   # first use allocates the staging
    hip.sync()
    t_up = time.perf_counter()
    ctx.upload_staged(batch, consts)
    hip.sync()
    staged_ms = 1e3 * (time.perf_counter() - t_up)
    staged_device_ms = ctx.timing()["upload_ms"]   # copies + pack on the stream (HIP events)
    upload_bytes = (batch.beam_scalars.nbytes + batch.beam_nv.nbytes + batch.beam_external.nbytes
                    + batch.vert_offsets.nbytes + 4 * batch.vx.nbytes + batch.edge_flags.nbytes
                    + batch.orient_offsets.nbytes + batch.weights.nbytes)
    results = {}
    traffic = measured_traffic()

    def leg(name, flags, peak, mode=None, graph=False):
        """Times args.steps calls of mbs_diffract(flags): wall per step, and
        the main kernel's mean duration from the engine's events (graph
        replay brackets the whole sequence, so the kernel time comes from an
        eager pass of the same shape first)."""
        if mode is not None:
            ctx.set_work_mode(mode)
        kernel_ms = []
        for _ in range(max(1, args.warmup)):
            ctx.diffract(flags)
        for _ in range(min(args.steps, 5)):
            ctx.diffract(flags)
            kernel_ms.append(ctx.timing()["kernel_ms"])
        tm = ctx.timing()
        if graph:
            ctx.set_graph(True)
            for _ in range(max(1, args.warmup)):
                ctx.diffract(flags)
        hip.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.diffract(flags)
            if not graph:
                kernel_ms.append(ctx.timing()["kernel_ms"])  # waits for this step
        hip.sync()
        wall = time.perf_counter() - t0
        if graph:
            ctx.set_graph(False)
        kms = float(np.mean(kernel_ms))
        # the committed counter figures of this leg's kernel and launch shape
        tb = traffic.get(name)
        results[name] = {
            "value": units * args.steps / wall,
            "ms_per_step": 1e3 * wall / args.steps,
            "kernel_ms": kms,
            "roofline": {"bound": "valu", "achieved": flops_launch / (kms * 1e-3) / 1e12,
                         "peak": peak / 1e12, "unit": "TFLOP/s",
                         "frac": flops_launch / (kms * 1e-3) / peak,
                         "traffic": tb,
                         "algorithmic_bytes": alg_bytes,
                         "traffic_over_algorithmic": tb / alg_bytes if tb else None},
            "work_items": tm["n_items"], "workgroups": tm["n_blocks"],
            "graph": graph,
        }

    # the default launch shapes (beam segments planned on the host; the
    # device work queue), then the shape the autotuner picks, replayed as a
    # graph: the headline
    leg("mixed_segments", abi.PREC_MIXED, FP32_PEAK, mode=abi.WORK_SEGMENTS)
    leg("mixed_queue", abi.PREC_MIXED, FP32_PEAK, mode=abi.WORK_QUEUE)
    leg("mixed_valu", abi.PREC_MIXED | abi.NO_MFMA, FP32_PEAK, mode=abi.WORK_SEGMENTS)
    leg("fp64", abi.PREC_FP64, FP64_PEAK, mode=abi.WORK_SEGMENTS)
    ctx.set_work_mode(0)
    t0 = time.perf_counter()
    tune = ctx.autotune(abi.PREC_MIXED)
    tune_s = time.perf_counter() - t0
    best = min(tune, key=lambda e: e[3])
    leg("mixed", abi.PREC_MIXED, FP32_PEAK, graph=True)
    results["mixed"]["tuned"] = {"target_blocks": best[0], "work_mode": best[1],
                                 "mfma": bool(best[2]), "waves_per_simd": best[4],
                                 "candidates": len(tune),
                                 "tuning_s": tune_s}
    # the counter figures of the launch shape the tuner picked on this device
    # (the traffic file lists every shape it tries)
    shape_key = ("mixed_q_t%d" % best[0] if best[1] == abi.WORK_QUEUE
                 else "mixed_s_w%d_g%d" % (best[4], results["mixed"]["workgroups"]))
    tb = traffic.get(shape_key) if best[2] else None
    if tb is not None:
        rl = results["mixed"]["roofline"]
        rl["traffic"] = tb
        rl["traffic_over_algorithmic"] = tb / rl["algorithmic_bytes"]
    results["mixed"]["roofline"]["traffic_shape"] = shape_key if tb is not None else "mixed"
    fft = bench_aperture_fft(ctx, batch, args)
    ctx.close()
    near = bench_near_forward(eng, abi, args)
    scan = bench_size_scan(eng, abi, pkg.synthetic, args)
    hist = np.bincount(batch.beam_nv, minlength=32)
    out = {
        "metric": "beam.dir evals/s",
        "value": results["mixed"]["value"],
        "unit": "beam.dir evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": results["mixed"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64-phase",
        "data": "synthetic: beams resampled from the reference tracer's prepared beams "
                "(D=10 um column) rotated and scaled to D=50 um",
        "config": {
            "workload": "orientation average, hexagonal column D=50 um L=100 um, "
                        "lambda=0.532 um, m=1.3116+1e-4i, 541 theta x 96 phi",
            "orientations_per_step": ORIENT_PER_STEP,
            "beams_per_step": int(batch.n_beams),
            "beams_per_orientation": "U(%d,%d)" % (BEAMS_MIN, BEAMS_MAX),
            "directions": grid.cells,
            "units_per_step": units,
            "vertex_histogram": {str(k): int(v) for k, v in enumerate(hist) if v},
            "mean_valid_edges": float(ne.mean()),
            "flops_per_unit_mean": float(flops_b.mean()),
        },
        "kernel_ms": results["mixed"]["kernel_ms"],
        "upload_ms": upload_ms,
        "upload": {"bytes": upload_bytes, "pageable_ms": upload_ms,
                   "pageable_GBps": upload_bytes / (upload_ms * 1e-3) / 1e9,
                   "staged_ms": staged_ms,
                   "staged_device_ms": staged_device_ms,
                   "staged_GBps": upload_bytes / (staged_device_ms * 1e-3) / 1e9,
                   "note": "pageable: wall of mbs_upload_beams (copies from pageable memory "
                           "and the device pack); staged: the batch filled into page-locked "
                           "staging by the host (staged_ms, wall) and its asynchronous copies "
                           "plus the device pack on the stream (staged_device_ms, events)"},
        "pcie_inclusive_value": units / ((staged_device_ms + results["mixed"]["ms_per_step"])
                                         * 1e-3),
        "roofline": results["mixed"]["roofline"],
        "launch": results["mixed"]["tuned"],
        "fp64": {k: results["fp64"][k] for k in ("value", "ms_per_step", "kernel_ms",
                                                 "roofline")},
        "mixed_valu": {k: results["mixed_valu"][k] for k in ("value", "ms_per_step", "kernel_ms",
                                                             "roofline")},
        "mixed_segments": {k: results["mixed_segments"][k]
                           for k in ("value", "ms_per_step", "kernel_ms", "work_items",
                                     "workgroups")},
        "mixed_queue": {k: results["mixed_queue"][k]
                        for k in ("value", "ms_per_step", "kernel_ms", "work_items",
                                  "workgroups")},
        "aperture_fft": fft,
        "near_forward": near,
        "size_scan": scan,
    }
    if not args.no_cpu:
        port = cpu_baseline(pkg, batch, grid, consts, args.cpu_budget / 2)
        ref = None
        traced_fx = None
        if os.path.exists(os.path.join(REF_DIR, "mbs_po")) and \
                os.path.exists(os.path.join(REF_DIR, "mbs_po_dump")):
            try:
                ref, traced_fx = cpu_baseline_reference(args.cpu_budget)
            except Exception as e:  # the oracle port stays the baseline
                port["reference_error"] = str(e)
        out["cpu_baseline"] = ref if ref is not None else port
        out["cpu_baseline"]["cpu_model"] = cpu_model()
        if ref is not None:
            out["cpu_oracle_port"] = port
        out["gpu_over_cpu_fp64"] = results["fp64"]["value"] / out["cpu_baseline"]["value"]
        if traced_fx is not None:
            # the same beams the reference CLI diffracted, through the engine
            tr = bench_traced(eng, abi, traced_fx, args)
            tr["over_reference_cpu_same_beams"] = tr["value"] / ref["value"]
            out["traced"] = tr
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of CPU time for the baseline sample")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    ap.add_argument("--config3", action="store_true",
                    help="only run the 16,384-orientation aggregate batch (~1e8 beams)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (orientations "
              "would shard across GPUs with one sum-reduce of the Mueller grid).  "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.config3:
        import mbs_loader
        print(json.dumps(bench_config3(mbs_loader.load(), args)))
        return
    if args.steps < 1 or args.warmup < 0:
        sys.exit("--steps must be >= 1 and --warmup >= 0")
    run(args)


if __name__ == "__main__":
    main()
