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
from golden.particles import write_column_aggregate

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
    with open(os.path.join(wd, out + ".log"), "w") as f:
        f.write(p.stdout + p.stderr)
    m = re.search(r"Phase 2 \(diffraction[^)]*\): ([0-9.]+) s", p.stdout)
    return float(m.group(1)) if m else float("nan")


def _dat(wd, out, suffix=""):
    return np.loadtxt(os.path.join(wd, out, out + suffix + ".dat"), skiprows=1)


def _case(particle, n_im, refl, n_orient, nphi, wd=None):
    """particle: the --particle numbers, or a particle file name in wd."""
    wd = wd or tempfile.mkdtemp(prefix="mbs_dropin_")
    with open(os.path.join(wd, "theta.grid"), "w") as f:
        f.write("\n".join("%.10g" % t for t in THETA_CONFIG2) + "\n")
    source = (["--pf", particle] if isinstance(particle, str)
              else ["--particle"] + [str(a) for a in particle])
    args = ["--method", "po"] + source + [
            "--refractive-index", "1.3116", str(n_im), "--wavelength-um", "0.532",
            "--max-reflections", str(refl), "--sobol-seed", str(n_orient), "1",
            "--theta-grid-file", "theta.grid", "--phi-points", str(nphi),
            "--cutoff-profile", "off", "--threads", str(_threads()), "--no-shadow-output",
            "--close"]
    return wd, args


def _compare(wd, args):
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


@pytest.mark.parametrize("shape", [(20, 10, 1e-4, 6, 8, 24), (100, 50, 1e-4, 10, 16, 96)],
                         ids=["D10_8orient_24phi", "D50_16orient_96phi"])
def test_reference_cli_gpu_backend_matches_cpu_backend(shape):
    L, D, n_im, refl, n_orient, nphi = shape
    print("case D=%g L=%g refl=%d orientations=%d phi=%d" % (D, L, refl, n_orient, nphi),
          flush=True)
    _compare(*_case((1, L, D), n_im, refl, n_orient, nphi))


# The reference's other built-in particles (src/main.cpp ParticleType):
# bullet rosette (3: LENGTH DIAMETER, six bullets), droxtal (4: SIZE), concave
# hexagonal column (10: LENGTH DIAMETER CAVITY_DEPTH, traced as non-convex) and
# the two-column aggregate (12: LENGTH DIAMETER 2, non-convex).  Non-convex
# tracing gives beams with more vertices and shadowed apertures; the slot and
# the engine must handle them unchanged.
@pytest.mark.parametrize("particle", [(3, 20, 10), (4, 10), (10, 20, 10, 0.3), (12, 20, 10, 2)],
                         ids=["bullet_rosette", "droxtal", "concave_hexagonal",
                              "two_column_aggregate"])
def test_reference_cli_gpu_backend_other_particles(particle):
    print("case particle %s refl=6 orientations=8 phi=24" % (particle,), flush=True)
    _compare(*_case(particle, 1e-4, 6, 8, 24))


def test_reference_cli_gpu_backend_eight_column_aggregate():
    """An 8-column aggregate read from a particle file (seed 7, D ~ U(5, 30)
    um, L/D ~ U(1, 4)): non-convex tracing over 64 facets, ~4.4k beams per
    orientation with up to 13 vertices at 6 reflections."""
    wd = tempfile.mkdtemp(prefix="mbs_dropin_")
    radii = write_column_aggregate(os.path.join(wd, "aggregate.particle"))
    print("case 8-column aggregate (bounding radii %s um) refl=6 orientations=4 phi=24"
          % ", ".join("%.1f" % r for r in radii), flush=True)
    _compare(*_case("aggregate.particle", 1e-4, 6, 4, 24, wd))


def test_reference_cli_fft_phi_interpolation():
    """The reference CLI's --fft path on the HIP backend
    (HandleOrientationsToLocalGpuFftPhi: 48 direct azimuths upsampled to 96)
    against the direct HIP backend on the same run.  The two differ by the
    phi interpolation error of the coarse grid; the bound only catches a
    broken upsampling (a lost or misplaced harmonic moves the averaged rows
    by far more)."""
    wd, args = _case((1, 20, 10), 1e-4, 6, 8, 96)
    _run(HIP_BIN, args, wd, "direct", "gpu")
    _run(HIP_BIN, args + ["--fft"], wd, "fft", "gpu")
    log = open(os.path.join(wd, "fft.log")).read()
    assert "direct Nphi=48, output Nphi=96" in log
    for suffix in ("", "_noshadow"):
        a, b = _dat(wd, "fft", suffix), _dat(wd, "direct", suffix)
        g = M.gate_l2(a, b)
        print("--fft vs direct%s: gate_l2 %.2e" % (suffix or " (full)", g))
        assert g < 5e-2
    # validation: the check grid here is min(96, 2 x 48) = the full grid, so
    # a tolerance the estimate exceeds (--fft-tolerance) returns the direct
    # result, and the check alone (MBS_FFT_CHECK=1, no tolerance) reports and
    # keeps the interpolated one
    _run(HIP_BIN, args + ["--fft", "--fft-tolerance", "1e-9"], wd, "fft_refined", "gpu")
    _run(HIP_BIN, args + ["--fft", "--allow-experimental-environment"], wd, "fft_kept", "gpu",
         {"MBS_FFT_CHECK": "1"})
    for out in ("fft_refined", "fft_kept"):
        assert "GPU FFT check full" in open(os.path.join(wd, out + ".log")).read()
    for suffix in ("", "_noshadow"):
        r = M.gate_l2(_dat(wd, "fft_refined", suffix), _dat(wd, "direct", suffix))
        k = M.gate_l2(_dat(wd, "fft_kept", suffix), _dat(wd, "fft", suffix))
        print("refined vs direct%s %.2e; kept vs --fft %.2e" % (suffix, r, k))
        assert r < 1e-12 and k < 1e-12
    # adaptive rows (MBS_FFT_ADAPTIVE_PHI, threshold low enough to take the
    # maximum of 8): those theta rows are the direct ones, the others stay
    # interpolated
    _run(HIP_BIN, args + ["--fft", "--allow-experimental-environment"], wd, "fft_adaptive",
         "gpu", {"MBS_FFT_ADAPTIVE_PHI": "1", "MBS_FFT_ADAPTIVE_PHI_THRESHOLD": "1e-9"})
    assert "8 of %d theta rows computed directly" % len(THETA_CONFIG2) in open(
        os.path.join(wd, "fft_adaptive.log")).read()
    for suffix in ("", "_noshadow"):
        a, d, f = (_dat(wd, n, suffix) for n in ("fft_adaptive", "direct", "fft"))
        scale = np.abs(d[:, 2:3])
        is_d = np.all(np.abs(a - d)[:, 2:] <= 1e-12 * scale, axis=1)
        is_f = np.all(np.abs(a - f)[:, 2:] <= 1e-12 * scale, axis=1)
        print("adaptive%s: %d rows direct, %d interpolated" % (suffix, (is_d & ~is_f).sum(),
                                                              is_f.sum()))
        assert np.all(is_d | is_f) and 1 <= (is_d & ~is_f).sum() <= 8


def test_reference_cli_fft_theta_factor():
    """--fft with MBS_FFT_THETA_FACTOR=2 on a uniform 181-row grid: the
    coarse grid has 91 rows and 48 azimuths, upsampled in theta, then in phi.
    Trigonometric interpolation passes through its samples, so the even rows
    (the coarse rows) must equal plain --fft's (181 direct rows, the same phi
    upsampling) to rounding; the odd rows are interpolated.  (Their error is
    large here -- 2 degree rows undersample this column's forward peak; the
    oracle restatement of the same upsampling gives the same figure.)"""
    wd = tempfile.mkdtemp(prefix="mbs_fft_theta_")
    args = ["--method", "po", "--particle", "1", "20", "10", "--refractive-index", "1.3116",
            "1e-4", "--wavelength-um", "0.532", "--max-reflections", "6", "--sobol-seed", "8",
            "1", "--scattering-grid", "0", "180", "96", "180", "--cutoff-profile", "off",
            "--threads", str(_threads()), "--close", "--fft"]
    _run(HIP_BIN, args, wd, "fft", "gpu")
    _run(HIP_BIN, args + ["--allow-experimental-environment"], wd, "fft_theta", "gpu",
         {"MBS_FFT_THETA_FACTOR": "2"})
    log = open(os.path.join(wd, "fft_theta.log")).read()
    assert "direct Ntheta=91, output Ntheta=181 (factor=2)" in log
    a, b = _dat(wd, "fft_theta"), _dat(wd, "fft")
    np.testing.assert_array_equal(a[:, :2], b[:, :2])
    rel = np.max(np.abs(a[:, 2:] - b[:, 2:]), axis=1) / np.abs(b[:, 2])
    print("--fft theta factor 2 vs --fft: even rows max rel %.2e, odd rows max rel %.2e, "
          "gate_l2 %.2e" % (rel[0::2].max(), rel[1::2].max(), M.gate_l2(a, b)))
    assert rel[0::2].max() < 1e-9 and rel[1::2].max() > 1e-6


def test_reference_cli_multi_size_fft_phi():
    """Multi-size mode with --fft: the fused HandleOrientationsToLocalGpuFftPhiMultiK
    (MBS_GPU_MULTI_K_FULL=1) against the per-size phi-FFT calls the CLI makes
    without it.  Both run the coarse grid (18 of 36 azimuths) and the same
    upsampling, so they agree to rounding; the CPU backend has no --fft."""
    wd = tempfile.mkdtemp(prefix="mbs_multisize_fft_")
    args = ["--method", "po", "--particle", "1", "20", "10",
            "--refractive-index", "1.3116", "1e-3", "--wavelength-um", "0.532",
            "--max-reflections", "6", "--oldauto", "8", "--multigrid", "6", "10", "3",
            "--scattering-grid", "0", "180", "36", "180", "--cutoff-profile", "off",
            "--threads", str(_threads()), "--close", "--fft"]
    env = {"MBS_HIP_PRECISION": "fp64"}
    _run(HIP_BIN, args, wd, "per_size", "gpu", env)
    env["MBS_GPU_MULTI_K_FULL"] = "1"
    _run(HIP_BIN, args + ["--allow-experimental-environment"], wd, "fused", "gpu", env)
    log = open(os.path.join(wd, "fused.log")).read()
    assert "falling back to per-size GPU path" not in log
    assert "direct Nphi=18, output Nphi=36" in open(os.path.join(wd, "per_size.log")).read()
    sizes = sorted(f[len("per_size_"):-len(".dat")]
                   for f in os.listdir(os.path.join(wd, "per_size")) if f.endswith(".dat"))
    assert len(sizes) == 3, sizes
    for size in sizes:
        e = M.gate_l2(_dat(wd, "fused", "_" + size), _dat(wd, "per_size", "_" + size))
        print("  size %s: fused vs per-size phi-FFT gate_l2 %.2e" % (size, e))
        assert e < 1e-9


@pytest.mark.parametrize("fused", [False, True], ids=["per_size_scaled", "fused_multi_scale"])
def test_reference_cli_multi_size(fused):
    """The reference's shared-trace multi-size mode (--multigrid D_MIN D_MAX N,
    src/scattering/TracerPOTotal.cpp:4070-4520): one trace at the largest size,
    then every size diffracted from rescaled beams.  Per size it calls
    HandleOrientationsToLocalGpu(..., scale, waveIndex) (scale < 1 for the
    smaller sizes); with MBS_GPU_MULTI_K_FULL=1 it calls the fused
    HandleOrientationsToLocalGpuMultiK for all sizes at once.  Absorbing
    (m = 1.3116 + 1e-3 i), so the absorption factors are rescaled too."""
    wd = tempfile.mkdtemp(prefix="mbs_multisize_")
    args = ["--method", "po", "--particle", "1", "20", "10",
            "--refractive-index", "1.3116", "1e-3", "--wavelength-um", "0.532",
            "--max-reflections", "6", "--oldauto", "8", "--multigrid", "6", "10", "3",
            "--scattering-grid", "0", "180", "24", "180", "--cutoff-profile", "off",
            "--threads", str(_threads()), "--close"]
    print("case multi-size D 6..10 (3 sizes), 150 orientations, fused=%s" % fused, flush=True)
    t_cpu = _run(CPU_BIN, args, wd, "cpu", "cpu")
    env = {"MBS_HIP_PRECISION": "fp64"}
    gpu_args = list(args)
    if fused:
        # The reference CLI accepts this experimental switch only when told to.
        env["MBS_GPU_MULTI_K_FULL"] = "1"
        gpu_args.append("--allow-experimental-environment")
    t_gpu = _run(HIP_BIN, gpu_args, wd, "gpu", "gpu", env)
    log = open(os.path.join(wd, "gpu.log")).read()
    if fused:
        assert "fused multi-k kernel enabled" in log
        assert "falling back to per-size GPU path" not in log
    sizes = sorted(f[len("cpu_"):-len(".dat")] for f in os.listdir(os.path.join(wd, "cpu"))
                   if f.endswith(".dat"))
    assert len(sizes) == 3, sizes
    for size in sizes:
        cpu = _dat(wd, "cpu", "_" + size)
        gpu = _dat(wd, "gpu", "_" + size)
        assert gpu.shape == cpu.shape
        np.testing.assert_array_equal(gpu[:, :2], cpu[:, :2])
        e = M.gate_l2(gpu, cpu)
        print("  size %s: gate_l2 fp64 %.2e" % (size, e))
        assert e < 1e-6
    print("multi-size phase 2: cpu %.2f s, gpu %.2f s" % (t_cpu, t_gpu))
