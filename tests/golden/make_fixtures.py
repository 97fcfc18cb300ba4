#!/usr/bin/env python3
"""Writes the golden fixtures of the PO diffraction hot path.

Runs only in the container that holds the reference sources (it needs the
binaries that ``make -f oracle/ref.mk`` builds into ``oracle/_ref``).  For each
case it

1. runs the unmodified reference CLI (``oracle/_ref/mbs_po``) and keeps its
   Mueller output file;
2. runs the fixture writer (``oracle/_ref/ref_fixed_harness`` for a fixed
   orientation, ``oracle/_ref/mbs_po_dump`` for an orientation average), which
   serialises the reference's own ``PreparedOrientation`` batches together with
   the Mueller grid the reference's CPU routines compute from them;
3. checks that (2) reproduces (1), so the dumped beams are the reference's
   beams;
4. stores inputs and expected outputs as a compressed ``.npz`` beside this
   script.

Fixture layout (all little-endian, FP64 unless noted)::

    theta[nZen+1] (rad), az_step, vf[nAz, nZen+1, 3]
    wave_index, eps1, compl_wave[2], inv_compl_wave[2], legacy_sign (int)
    orient_offsets[nOrient+1] (int64), weights[nOrient]
    beam_scalars[nBeams, 36], beam_nv, beam_external, beam_valid (int32)
    vert_offsets[nBeams+1] (int64), vx, vy, slope_yx, slope_xy, edge_flags (u8)
    mueller_full[nAz, nZen+1, 16], mueller_noshadow[nAz, nZen+1, 16]

``beam_scalars`` columns: hor(3) ver(3) dir(3) cen(3) area NT(3) NP(3) nxDT(3)
nxDP(3) J_phased(8: 00r 00i 01r 01i 10r 10i 11r 11i) normal(3) -- the fields of
the reference's PreparedBeam (src/handler/HandlerPO.h:14-46).
"""
import os
import shutil
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(REPO, "oracle", "_ref")
# scratch directory of the reference runs (set by main)
WORK = None


class _Reader:
    def __init__(self, data):
        self.d = data
        self.o = 0

    def take(self, n):
        b = self.d[self.o:self.o + n]
        if len(b) != n:
            raise EOFError
        self.o += n
        return b

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def f64(self, n=None):
        if n is None:
            return struct.unpack("<d", self.take(8))[0]
        return np.frombuffer(self.take(8 * n), dtype="<f8").copy()

    def tag(self):
        return self.take(4).decode()

    def done(self):
        return self.o >= len(self.d)


def read_dump(path):
    """Parses the stream written by ref_dump_shim.cpp into numpy arrays.

    Beam records are variable-length (a 12-byte header, 36 doubles, then 33
    bytes per vertex: x, y, slope_yx, slope_xy and a flag byte); a scan
    over the record headers finds every record, and the fields are then
    gathered column by column in blocks."""
    buf = np.fromfile(path, dtype=np.uint8)
    data = buf.tobytes()
    r = _Reader(data)
    assert r.take(8) == b"MBSFIX01"
    n_az = r.i32()
    n_zen = r.i32()
    out = {"theta": r.f64(n_zen + 1), "az_step": r.f64()}
    out["vf"] = r.f64(n_az * (n_zen + 1) * 3).reshape(n_az, n_zen + 1, 3)
    out["wave_index"] = r.f64()
    out["eps1"] = r.f64()
    out["compl_wave"] = r.f64(2)
    out["inv_compl_wave"] = r.f64(2)
    out["legacy_sign"] = r.i32()
    r.i32()  # no-shadow flag of the handler (informational)
    weights, orient_beams, beam_off = [], [], []
    scal_rec, scal_cabs = [], None
    m_full = np.zeros((n_az, n_zen + 1, 16))
    m_ns = np.zeros((n_az, n_zen + 1, 16))
    i32 = struct.Struct("<i")
    o = r.o
    while o < len(data):
        assert data[o:o + 4] == b"BTCH", data[o:o + 4]
        n_or = i32.unpack_from(data, o + 4)[0]
        o += 8
        for _ in range(n_or):
            assert data[o:o + 4] == b"ORNT"
            weights.append(struct.unpack_from("<d", data, o + 4)[0])
            nb = i32.unpack_from(data, o + 12)[0]
            orient_beams.append(nb)
            o += 16
            for _ in range(nb):
                beam_off.append(o)
                o += 300 + 33 * i32.unpack_from(data, o)[0]
        r.o = o
        if r.done():
            break
        if data[r.o:r.o + 4] == b"SCAL":
            r.tag()
            scal_cabs = r.f64()
            o = r.o
            nb_batch = sum(orient_beams[-n_or:]) if n_or else 0
            for _ in range(nb_batch):
                scal_rec.append(o)
                o += 80 + 4 + 8 * i32.unpack_from(data, o + 80)[0]
            r.o = o
        assert r.tag() == "MFUL"
        m_full += r.f64(n_az * (n_zen + 1) * 16).reshape(n_az, n_zen + 1, 16)
        assert r.tag() == "MNSH"
        m_ns += r.f64(n_az * (n_zen + 1) * 16).reshape(n_az, n_zen + 1, 16)
        o = r.o
    boff = np.array(beam_off, dtype=np.int64)
    nb = len(boff)

    def gather(starts, width):
        idx = starts[:, None] + np.arange(width, dtype=np.int64)[None, :]
        return buf[idx]

    hdr = gather(boff, 12).view("<i4").reshape(nb, 3) if nb else np.zeros((0, 3), np.int32)
    nv = hdr[:, 0].astype(np.int32)
    scalars = np.empty((nb, 36))
    for s0 in range(0, nb, 65536):
        blk = boff[s0:s0 + 65536]
        scalars[s0:s0 + len(blk)] = gather(blk + 12, 288).view("<f8").reshape(-1, 36)
    voff = np.concatenate([[0], np.cumsum(nv)]).astype(np.int64)
    n_vert = int(voff[-1])
    # byte offset of every vertex record: its beam's block plus 33 j
    vbeam = np.repeat(np.arange(nb), nv)
    vstart = boff[vbeam] + 300 + 33 * (np.arange(n_vert) - voff[vbeam])
    vrec = np.empty((n_vert, 4))
    flags = np.empty(n_vert, dtype=np.uint8)
    for s0 in range(0, n_vert, 262144):
        blk = vstart[s0:s0 + 262144]
        raw = gather(blk, 33)
        vrec[s0:s0 + len(blk)] = np.ascontiguousarray(raw[:, :32]).view("<f8").reshape(-1, 4)
        flags[s0:s0 + len(blk)] = raw[:, 32]
    out["orient_offsets"] = np.concatenate([[0], np.cumsum(orient_beams)]).astype(np.int64)
    out["weights"] = np.array(weights)
    out["beam_scalars"] = scalars
    out["beam_nv"] = nv
    out["beam_external"] = hdr[:, 1].astype(np.int32)
    out["beam_valid"] = hdr[:, 2].astype(np.int32)
    out["vert_offsets"] = voff
    out["vx"] = vrec[:, 0].copy()
    out["vy"] = vrec[:, 1].copy()
    out["slope_yx"] = vrec[:, 2].copy()
    out["slope_xy"] = vrec[:, 3].copy()
    out["edge_flags"] = flags
    out["mueller_full"] = m_full
    out["mueller_noshadow"] = m_ns
    if scal_rec:
        # the multi-size data of every beam (MBS_FIXTURE_SCALING=1)
        so = np.array(scal_rec, dtype=np.int64)
        rec = np.empty((len(so), 10))
        for s0 in range(0, len(so), 65536):
            blk = so[s0:s0 + 65536]
            rec[s0:s0 + len(blk)] = gather(blk, 80).view("<f8").reshape(-1, 10)
        n_abs = gather(so + 80, 4).view("<i4").reshape(-1).astype(np.int64)
        aoff = np.concatenate([[0], np.cumsum(n_abs)]).astype(np.int64)
        abeam = np.repeat(np.arange(len(so)), n_abs)
        astart = so[abeam] + 84 + 8 * (np.arange(int(aoff[-1])) - aoff[abeam])
        out["raw_jones"] = rec[:, :8].copy()
        out["path"] = rec[:, 8].copy()
        out["sign"] = rec[:, 9].copy()
        out["abs_offsets"] = aoff
        out["abs_paths"] = (gather(astart, 8).view("<f8").reshape(-1).copy()
                            if len(astart) else np.zeros(0))
        out["c_abs"] = scal_cabs
    return out


def read_fixed_dat(path, n_az, n_zen):
    """Reads the per-direction .dat of the fixed path (HandlerPO.cpp:117-153):
    header 'nZen nAz', then rows 'theta phi energy M11..M44' ordered t-major."""
    rows = np.loadtxt(path, skiprows=1)
    assert rows.shape == ((n_zen + 1) * n_az, 19), rows.shape
    m = rows[:, 3:].reshape(n_zen + 1, n_az, 16)
    return np.transpose(m, (1, 0, 2)).copy(), rows[:, 0].reshape(n_zen + 1, n_az)[:, 0]


def read_total_dat(path):
    """Reads the azimuth-averaged .dat (HandlerPOTotal.cpp:23-122):
    header line, then rows 'ScAngle 2pi*dcos M11..M44'."""
    return np.loadtxt(path, skiprows=1)


def run(cmd, cwd, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    p = subprocess.run(cmd, cwd=cwd, env=e, stdin=subprocess.DEVNULL,
                       stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout[-4000:])
        raise SystemExit("command failed: %s" % " ".join(cmd))
    return p.stdout


def rel_err(a, b):
    """max |a-b| / max(|M11(b)|) per cell, over all 16 elements."""
    scale = np.abs(b[..., :1])
    scale = np.where(scale > 0, scale, 1.0)
    return float(np.max(np.abs(a - b) / scale))


def fresh_dir(name):
    # The reference appends to its .dat files, so every case starts empty.
    wd = os.path.join(WORK, name)
    if os.path.isdir(wd):
        shutil.rmtree(wd)
    os.makedirs(wd)
    return wd


def fixed_case(name, L, D, n_re, n_im, wave, refl, beta, gamma, th0, th1, nphi, nth,
               keep_rows=None):
    wd = fresh_dir(name)
    args = ["--method", "po", "--particle", "1", str(L), str(D),
            "--refractive-index", str(n_re), str(n_im), "--wavelength-um", str(wave),
            "--max-reflections", str(refl), "--fixed-orientation", str(beta), str(gamma),
            "--scattering-grid", str(th0), str(th1), str(nphi), str(nth),
            "--backend", "cpu", "--cutoff-profile", "off", "--threads", "8",
            "--close", "--output", "ref"]
    run([os.path.join(REF_BIN, "mbs_po")] + args, wd)
    dump = os.path.join(wd, "dump.bin")
    run([os.path.join(REF_BIN, "ref_fixed_harness"), str(L), str(D), str(n_re), str(n_im),
         str(wave), str(refl), str(beta), str(gamma), str(th0), str(th1), str(nphi),
         str(nth)], wd, {"MBS_FIXTURE_DUMP": dump})
    fx = read_dump(dump)
    dat_path = os.path.join(wd, "ref", "ref.dat")
    m_dat, theta_deg = read_fixed_dat(dat_path, nphi, nth)
    err = rel_err(fx["mueller_full"], m_dat)
    print("%s: %d beams, harness(HandleBeamsToLocal) vs reference .dat(HandleBeams): "
          "max|dM|/M11 = %.3e" % (name, len(fx["beam_nv"]), err))
    assert err < 1e-11, "harness does not reproduce the reference fixed path"
    fx["mueller_fixed_dat"] = m_dat
    fx["case"] = np.array(name)
    if keep_rows is not None:
        fx["kept_theta_rows"] = np.array(keep_rows, dtype=np.int64)
        for k in ("mueller_full", "mueller_noshadow", "mueller_fixed_dat"):
            fx[k] = fx[k][:, keep_rows, :]
    return fx


def sobol_case(name, L, D, n_re, n_im, wave, refl, n_orient, seed, nphi, nth,
               th0=0, th1=180, tgrid=None, chunk=None, particle_file=None):
    """particle_file: a callable writing a particle file into the run
    directory (used instead of the built-in column L, D)."""
    wd = fresh_dir(name)
    if tgrid is not None:
        with open(os.path.join(wd, "theta.grid"), "w") as f:
            f.write("\n".join("%.10g" % t for t in tgrid) + "\n")
        grid = ["--theta-grid-file", "theta.grid", "--phi-points", str(nphi)]
    else:
        grid = ["--scattering-grid", str(th0), str(th1), str(nphi), str(nth)]
    if particle_file is not None:
        particle_file(os.path.join(wd, "particle.txt"))
        source = ["--pf", "particle.txt"]
    else:
        source = ["--particle", "1", str(L), str(D)]
    args = ["--method", "po"] + source + [
            "--refractive-index", str(n_re), str(n_im), "--wavelength-um", str(wave),
            "--max-reflections", str(refl), "--sobol-seed", str(n_orient), str(seed)] + grid + [
            "--cutoff-profile", "off", "--threads", "8", "--no-shadow-output", "--close"]
    if chunk:
        args += ["--orientation-chunk", str(chunk)]
    run([os.path.join(REF_BIN, "mbs_po")] + args + ["--backend", "cpu", "--output", "cpu"], wd)
    dump = os.path.join(wd, "dump.bin")
    run([os.path.join(REF_BIN, "mbs_po_dump")] + args + ["--backend", "gpu", "--output", "dmp"],
        wd, {"MBS_FIXTURE_DUMP": dump})
    fx = read_dump(dump)
    cpu = read_total_dat(os.path.join(wd, "cpu", "cpu.dat"))
    dmp = read_total_dat(os.path.join(wd, "dmp", "dmp.dat"))
    cpu_ns = read_total_dat(os.path.join(wd, "cpu", "cpu_noshadow.dat"))
    scale = np.abs(cpu[:, 2:3])
    err = float(np.max(np.abs(dmp[:, 2:] - cpu[:, 2:]) / scale))
    print("%s: %d orientations, %d beams (min %d max %d per orientation); "
          "dump-path vs CPU-backend azimuth-averaged .dat: max|dM|/M11 = %.3e"
          % (name, len(fx["weights"]), len(fx["beam_nv"]),
             np.diff(fx["orient_offsets"]).min(), np.diff(fx["orient_offsets"]).max(), err))
    assert err < 1e-11, "dumped batches do not reproduce the reference CPU backend"
    fx["ref_total_dat"] = cpu
    fx["ref_total_dat_noshadow"] = cpu_ns
    fx["case"] = np.array(name)
    return fx


def sobol_points(n=4096, seeds=(1, 7, 42, 4096, 123456789)):
    """The reference's own Sobol2D points (x then y per seed), dumped by
    oracle/_ref/ref_sobol_dump (tests/golden/ref_sobol_dump.cpp)."""
    wd = fresh_dir("sobol_points")
    out = os.path.join(wd, "sobol.bin")
    run([os.path.join(REF_BIN, "ref_sobol_dump"), out, str(n)] + [str(s) for s in seeds], wd)
    raw = np.fromfile(out, dtype=np.float64).reshape(len(seeds), 2, n)
    print("sobol_points: %d seeds x %d points" % (len(seeds), n))
    return {"seeds": np.array(seeds, dtype=np.uint32), "x": raw[:, 0], "y": raw[:, 1]}


def save(name, fx):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **fx)
    print("wrote %s (%.1f KiB)" % (path, os.path.getsize(path) / 1024.0))


def main(only=()):
    """Writes every fixture, or only the named ones (command-line arguments)."""
    global WORK
    WORK = tempfile.mkdtemp(prefix="mbs_fixtures_")
    try:
        _main(only)
    finally:
        shutil.rmtree(WORK, ignore_errors=True)


def _main(only):
    want = lambda name: not only or name in only  # noqa: E731
    if want("sobol_points"):
        save("sobol_points", sobol_points())
    if not any(want(n) for n in ("config1_fixed", "fixed_absorbing", "sobol_average",
                                 "config5_near_forward", "config3_aggregate")):
        return
    # Configuration 1: hexagonal column L=20 D=10 um, m=1.3116, lambda=0.532,
    # 8 reflections, beta=30 gamma=10, theta 0..180 in 720 steps x 48 phi.
    # Inputs are kept whole; the expected grid every 8th theta row (+ last).
    rows = list(range(0, 721, 8)) + [720]
    if want("config1_fixed"):
        save("config1_fixed", fixed_case("config1_fixed", 20, 10, 1.3116, 0.0, 0.532, 8,
                                         30, 10, 0, 180, 48, 720, keep_rows=rows))
    # Absorbing fixed case (complex refractive index; exercises the absorption
    # factors folded into J_phased), small grid.
    if want("fixed_absorbing"):
        save("fixed_absorbing", fixed_case("fixed_absorbing", 12, 8, 1.3116, 0.01, 0.532, 6,
                                           47, 23, 0, 180, 16, 90))
    # Orientation average: reference Owen-Sobol (seed 1) orientations of the
    # configuration-2 particle shape scaled down (L=20 D=10 um) with absorption,
    # the non-uniform theta grid of configuration 2 on 24 phi bins.
    # Large-particle near-forward case: the hexagonal column D=200 L=400 um,
    # m=1.3116+1e-3i, 12 reflections, 4 Sobol orientations (seed 1); the
    # expected grids on theta 0..5 deg at 0.05 deg x 12 phi (the tests run
    # the full 0.005 deg grid against the oracle).
    if want("config5_near_forward"):
        save("config5_near_forward", sobol_case("config5_near_forward", 400, 200, 1.3116, 1e-3,
                                                0.532, 12, 4, 1, 12, 100, th0=0, th1=5))
    # Batched non-convex aggregate: the 8-column aggregate (seed 7, D ~ U(5, 30)
    # um, L/D ~ U(1, 4)), m=1.3116+1e-4i, 6 reflections, 2 Sobol orientations
    # (seed 1); expected grids on theta 0..180 in 90 steps x 12 phi.  The beams
    # seed synthetic batches of the configuration's size.
    if want("config3_aggregate"):
        from particles import write_column_aggregate
        save("config3_aggregate", sobol_case("config3_aggregate", 0, 0, 1.3116, 1e-4, 0.532, 6,
                                             2, 1, 12, 90,
                                             particle_file=write_column_aggregate))
    tg = [round(0.05 * i, 2) for i in range(201)] + [10.5 + 0.5 * i for i in range(340)]
    if want("sobol_average"):
        save("sobol_average", sobol_case("sobol_average", 20, 10, 1.3116, 1e-4, 0.532, 6,
                                         8, 1, 24, 540, tgrid=tg))


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
