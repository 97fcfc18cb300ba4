"""The drop-in, end to end: the reference's own CLI with the MI355X backend.

oracle/_ref/mbs_po_hip is the reference CLI built from its sources with
HandlerPO's GPU members supplied by mbs-fast_amd/host/HandlerPOHip.cpp (the
product's drop-in translation unit) over libmbs_hip.so; oracle/_ref/mbs_po is
the same sources' CPU build.  Both trace the same particle and orientations
with the reference's own tracer and PrepareBeams and write the reference's
azimuth-averaged .dat (src/handler/HandlerPOTotal.cpp:23-122); only the
diffraction backend differs (`--backend cpu` runs HandleBeamsToLocal,
`--backend gpu` runs HandleOrientationsToLocalGpu on the engine).

Bars (gate_l2 = the reference release gate's 2 pi dcos-weighted L2 over all
16 elements, tests/run_cuda_release_gate.sh): FP64 1e-6 -- the reference CPU
path's polynomial sincos carries ~1e-8 phase errors, so FP64 agreement is
bounded by the reference, not by the engine (the engine meets 1e-10 against
the libm-sincos oracle in test_engine_gpu.py); mixed precision: P11 gate_l2
1e-4.
"""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import metrics as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
CPU_BIN = os.path.join(REF, "mbs_po")
HIP_BIN = os.path.join(REF, "mbs_po_hip")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(CPU_BIN) and os.path.exists(HIP_BIN)),
                                 reason="reference CLI builds (oracle/_ref) are not present")]

THETA_CONFIG2 = [round(0.05 * i, 2) for i in range(201)] + [10.5 + 0.5 * i for i in range(340)]


def _threads():
    return max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", 0)) or os.cpu_count() or 1))


def _run(binary, args, wd, out, backend, env=None):
    e = dict(os.environ)
    e.update(env or {})
    p = subprocess.run([binary] + args + ["--backend", backend, "--output", out],
                       cwd=wd, env=e, stdin=subprocess.DEVNULL, capture_output=True,
                       text=True, timeout=600)
    assert p.returncode == 0, "%s failed (rc %d):\n%s\n%s" % (
        os.path.basename(binary), p.returncode, p.stdout[-2000:], p.stderr[-2000:])
    m = re.search(r"Phase 2 \(diffraction[^)]*\): ([0-9.]+) s", p.stdout)
    return float(m.group(1)) if m else float("nan")


def _dat(wd, out, suffix=""):
    return np.loadtxt(os.path.join(wd, out, out + suffix + ".dat"), skiprows=1)


def _case(L, D, n_im, refl, n_orient, nphi):
    wd = tempfile.mkdtemp(prefix="mbs_dropin_")
    with open(os.path.join(wd, "theta.grid"), "w") as f:
        f.write("\n".join("%.10g" % t for t in THETA_CONFIG2) + "\n")
    args = ["--method", "po", "--particle", "1", str(L), str(D),
            "--refractive-index", "1.3116", str(n_im), "--wavelength-um", "0.532",
            "--max-reflections", str(refl), "--sobol-seed", str(n_orient), "1",
            "--theta-grid-file", "theta.grid", "--phi-points", str(nphi),
            "--cutoff-profile", "off", "--threads", str(_threads()), "--no-shadow-output",
            "--close"]
    return wd, args


@pytest.mark.parametrize("shape", [(20, 10, 1e-4, 6, 8, 24), (100, 50, 1e-4, 10, 16, 96)],
                         ids=["D10_8orient_24phi", "D50_16orient_96phi"])
def test_reference_cli_gpu_backend_matches_cpu_backend(shape):
    L, D, n_im, refl, n_orient, nphi = shape
    print("case D=%g L=%g refl=%d orientations=%d phi=%d" % (D, L, refl, n_orient, nphi),
          flush=True)
    wd, args = _case(L, D, n_im, refl, n_orient, nphi)
    t_cpu = _run(CPU_BIN, args, wd, "cpu", "cpu")
    cpu, cpu_ns = _dat(wd, "cpu"), _dat(wd, "cpu", "_noshadow")
    t64 = _run(HIP_BIN, args, wd, "g64", "gpu", {"MBS_HIP_PRECISION": "fp64"})
    g64, g64_ns = _dat(wd, "g64"), _dat(wd, "g64", "_noshadow")
    t32 = _run(HIP_BIN, args, wd, "g32", "gpu", {"MBS_HIP_PRECISION": "fp32"})
    g32 = _dat(wd, "g32")
    assert g64.shape == cpu.shape == g32.shape
    np.testing.assert_array_equal(g64[:, :2], cpu[:, :2])  # angles and 2 pi dcos
    e64, e64_ns = M.gate_l2(g64, cpu), M.gate_l2(g64_ns, cpu_ns)
    e32 = M.gate_l2(g32[:, :3], cpu[:, :3])
    print("phase 2: cpu %.3f s, gpu fp64 %.3f s, gpu mixed %.3f s; gate_l2 fp64 %.2e "
          "(no-shadow %.2e), mixed P11 %.2e" % (t_cpu, t64, t32, e64, e64_ns, e32))
    assert e64 < 1e-6 and e64_ns < 1e-6
    assert e32 < 1e-4
