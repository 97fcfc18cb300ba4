"""Time to 1 % Mueller error for the 50 um hexagonal column, end to end.

A measurement rather than a parity test; it runs only with
MBS_CONVERGENCE=1 (about ten minutes on one GPU and 16 host cores).  The reference's
orientation-convergence method (scripts/run_po_orientation_sampling_convergence.sh,
scripts/plot_po_orientation_sampling_convergence.py:43-90): independent runs
of the CLI at growing orientation counts N, each compared with a converged
high-N run on the same theta grid, and the wall time of the first N whose
error is at most 1 % on every element P11 ... P44.

Workload: D = 50 um, L = 100 um, m = 1.3116 + 1e-4 i, lambda = 0.532 um, 10
reflections, theta rows 0:0.05:10, 10.5:0.5:180 (541) x 96 phi, Owen-Sobol
orientations (--sobol-seed N 1).  Converged reference: FP64 engine, the mean
of REF_RUNS independent scrambles (seeds 7, 8, ...) of REF_N orientations
each (4 x 32768 = 131072 by default; every run reports when it ends).

Sides:
  gpu fp64   oracle/_ref/mbs_po_hip --backend gpu, MBS_HIP_PRECISION=fp64
  gpu mixed  the same with MBS_HIP_PRECISION=fp32 (FP32 aperture sums, FP64
             phase, MFMA coherent sum)
  cpu        oracle/_ref/mbs_po --backend cpu on the host's cores: run at the
             smallest N and scaled linearly in N (tracing and diffraction are
             both per-orientation loops) to the N the GPU sides need
Both CLIs trace with the reference's own tracer; the diffraction phase is the
only part that differs.  Wall time and the CLI's own phase-1 / phase-2 times
are printed for every run.

Errors: metrics.convergence_errors (2 pi dcos-weighted: M11 relative L2 and
P_ij / P11 RMS) decides the 1 % bar; metrics.plot_element_errors (the
reference plot's unweighted form) is printed beside it.
"""
import os
import re
import subprocess
import tempfile
import time

import numpy as np
import pytest

import metrics as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
CPU_BIN = os.path.join(REF, "mbs_po")
HIP_BIN = os.path.join(REF, "mbs_po_hip")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("MBS_CONVERGENCE") != "1",
                                 reason="long measurement; set MBS_CONVERGENCE=1 to run"),
              pytest.mark.skipif(not (os.path.exists(CPU_BIN) and os.path.exists(HIP_BIN)),
                                 reason="reference CLI builds (oracle/_ref) are not present")]

THETA = [round(0.05 * i, 2) for i in range(201)] + [10.5 + 0.5 * i for i in range(340)]
REF_N = int(os.environ.get("MBS_CONVERGENCE_REF_N", "32768"))
REF_RUNS = int(os.environ.get("MBS_CONVERGENCE_REF_RUNS", "4"))
N_MAX = int(os.environ.get("MBS_CONVERGENCE_NMAX", "32768"))
NPHI = 96
BAR = 0.01


def _threads():
    return max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", 0)) or os.cpu_count() or 1))


def _run(wd, binary, backend, n, seed, out, prec=None):
    env = dict(os.environ)
    if prec:
        env["MBS_HIP_PRECISION"] = prec
    args = [binary, "--method", "po", "--particle", "1", "100", "50",
            "--refractive-index", "1.3116", "1e-4", "--wavelength-um", "0.532",
            "--max-reflections", "10", "--sobol-seed", str(n), str(seed),
            "--theta-grid-file", "theta.grid", "--phi-points", str(NPHI),
            "--cutoff-profile", "off", "--threads", str(_threads()), "--close",
            "--backend", backend, "--output", out]
    t0 = time.perf_counter()
    # The CLI's per-chunk lines are echoed as it prints them, so a long run
    # shows its progress.
    p = subprocess.Popen(args, cwd=wd, env=env, stdin=subprocess.DEVNULL, stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True)
    lines = []
    for line in p.stdout:
        lines.append(line)
        if "adaptive memory" in line:
            print("  [%s] %s" % (out, line.strip()[:160]), flush=True)
    rc = p.wait(timeout=1200)
    wall = time.perf_counter() - t0
    text = "".join(lines)
    assert rc == 0, "%s rc %d:\n%s" % (binary, rc, text[-3000:])
    ph = [float(x) for x in re.findall(r"Phase [12] \([^)]*\): ([0-9.]+) s", text)[:2]]
    rows = np.loadtxt(os.path.join(wd, out, out + ".dat"), skiprows=1)
    return rows, wall, ph


def _sweep(wd, name, prec, ref):
    """Runs N = 64, 128, ... until the bar is met (or N_MAX); returns the
    first N that meets it with its wall and phase times (None if none does)
    and the largest element error of every run."""
    n, hit, worst = 64, None, []
    while n <= N_MAX:
        rows, wall, ph = _run(wd, HIP_BIN, "gpu", n, 1, "%s_%d" % (name, n), prec)
        e, ep = M.convergence_errors(rows, ref), M.plot_element_errors(rows, ref)
        print("%-9s N=%5d wall %7.2f s (trace %.2f s, diffraction %.2f s)  max element %.4f "
              "(M%d)  P11 %.4f | plot form: max %.4f" % (name, n, wall, ph[0], ph[1], e.max(),
                                                          11 + 10 * (e.argmax() // 4) + e.argmax() % 4,
                                                          e[0], ep.max()), flush=True)
        worst.append(e.max())
        if e.max() <= BAR:
            hit = (n, wall, ph)
            break
        n *= 2
    return hit, worst


def test_time_to_one_percent():
    wd = tempfile.mkdtemp(prefix="mbs_conv_")
    with open(os.path.join(wd, "theta.grid"), "w") as f:
        f.write("\n".join("%.10g" % t for t in THETA) + "\n")
    print("converged reference: fp64 engine, %d x %d orientations (seeds 7..%d)"
          % (REF_RUNS, REF_N, 6 + REF_RUNS), flush=True)
    parts = []
    for r in range(REF_RUNS):
        rows, wall, ph = _run(wd, HIP_BIN, "gpu", REF_N, 7 + r, "ref%d" % r, "fp64")
        parts.append(rows)
        print("reference run %d: wall %.1f s (trace %.1f s, diffraction %.1f s)"
              % (r, wall, *ph), flush=True)
    ref = np.mean(parts, axis=0)
    if REF_RUNS > 1:
        # spread of the independent scrambles: the reference's own error is
        # about this divided by sqrt(REF_RUNS)
        spread = max(M.convergence_errors(p, ref).max() for p in parts)
        print("largest element error of one reference run against their mean: %.4f" % spread,
              flush=True)
    sweeps = {name: _sweep(wd, name, prec, ref) for name, prec in (("gpu fp64", "fp64"),
                                                                   ("gpu mixed", "fp32"))}
    n_cpu = 64
    _, wall_cpu, ph_cpu = _run(wd, CPU_BIN, "cpu", n_cpu, 1, "cpu_%d" % n_cpu)
    print("cpu       N=%5d wall %7.2f s (trace %.2f s, diffraction %.2f s) on %d threads"
          % (n_cpu, wall_cpu, ph_cpu[0], ph_cpu[1], _threads()), flush=True)
    for name, (hit, worst) in sweeps.items():
        # Sobol sets converge: the worst element shrinks from the first N to the last.
        assert len(worst) == 1 or worst[-1] < worst[0], (name, worst)
        if hit is None:
            print("%s: 1 %% on every element not reached by N = %d" % (name, N_MAX))
            continue
        n, wall, ph = hit
        cpu_wall = wall_cpu * n / n_cpu
        cpu_diff = ph_cpu[1] * n / n_cpu
        print("%s: 1 %% at N = %d after %.2f s wall (diffraction %.2f s); cpu reference at the "
              "same N (scaled from N=%d): %.1f s wall (diffraction %.1f s) -> %.1fx wall, "
              "%.0fx diffraction" % (name, n, wall, ph[1], n_cpu, cpu_wall, cpu_diff,
                                     cpu_wall / wall, cpu_diff / max(ph[1], 1e-9)))
