"""CPU-only checks of the C-ABI boundary: the built library loads and exports
every entry point that include/mbs_hip.h declares, and the ctypes mirror has the
same struct layout as the header (checked against gcc's view of the header)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "mbs_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(mbs_[a-z_0-9]+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("mbs_ctx_create", "mbs_set_grid", "mbs_upload_beams", "mbs_diffract",
                     "mbs_accumulate_mueller", "mbs_check_runtime", "mbs_query_memory",
                     "mbs_visible_device_count", "mbs_last_error"):
        assert required in names
    assert len(names) >= 18


def test_library_exports_every_declared_symbol(pkg):
    path = pkg.engine.LIB_PATH
    assert os.path.exists(path), "libmbs_hip.so not built"
    lib = C.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.mbs_abi_version() == 1


def test_struct_layout_matches_header(pkg, tmp_path):
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "%s"
int main(void) {
  printf("%%zu %%zu %%zu %%zu\\n", sizeof(mbs_beam_batch), sizeof(mbs_grid),
         sizeof(mbs_consts), sizeof(mbs_device_info));
  printf("%%zu %%zu %%zu\\n", offsetof(mbs_beam_batch, edge_flags),
         offsetof(mbs_grid, vf), offsetof(mbs_consts, legacy_sign));
  printf("%%zu %%zu\\n", sizeof(mbs_beam_scaling), offsetof(mbs_beam_scaling, wave_index));
  return 0;
}
""" % HEADER)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    abi = pkg.abi
    want = [C.sizeof(abi.mbs_beam_batch), C.sizeof(abi.mbs_grid), C.sizeof(abi.mbs_consts),
            C.sizeof(abi.mbs_device_info), abi.mbs_beam_batch.edge_flags.offset,
            abi.mbs_grid.vf.offset, abi.mbs_consts.legacy_sign.offset,
            C.sizeof(abi.mbs_beam_scaling), abi.mbs_beam_scaling.wave_index.offset]
    assert [int(v) for v in out] == want


def test_no_device_is_reported_not_faked(pkg):
    """Without a HIP device the engine reports it; it never computes on the CPU."""
    import ctypes
    lib = pkg.engine.lib()
    if lib.mbs_visible_device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(pkg.engine.EngineError):
        pkg.engine.Context(0)
    info = pkg.abi.mbs_device_info()
    assert lib.mbs_check_runtime(ctypes.byref(info)) == pkg.abi.MBS_ERR_DEVICE
