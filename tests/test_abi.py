"""The C-ABI boundary without a GPU: the HIP library loads, exports every entry
point include/ne_gpu.h declares, and the Python ctypes mirror of the header's
structs has the C layout (checked by compiling the header with gcc)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from nativeengine_amd import _abi, gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ne_gpu.h")


def declared_entry_points():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ne_gpu_\w+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared_entry_points()
    for must in ["ne_gpu_create", "ne_gpu_destroy", "ne_gpu_set_params", "ne_gpu_set_bodies", "ne_gpu_step",
                 "ne_gpu_get_state", "ne_gpu_get_contacts", "ne_gpu_get_pairs", "ne_gpu_reset_warm_cache"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = gpu.load_library()
    missing = [n for n in declared_entry_points() if not hasattr(lib, n)]
    assert not missing, f"libnegpu.so lacks {missing}"
    assert sorted(gpu.EXPORTS) == declared_entry_points(), "Python binding list drifted from the header"


def test_library_is_gfx950_code_object():
    """The shared object carries a gfx950 code object (no other targets)."""
    blob = open(gpu.LIB_PATH, "rb").read()
    targets = set(re.findall(rb"amdhsa--(gfx\w+)", blob))
    assert targets == {b"gfx950"}, targets


def _c_layout():
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "ne_gpu.h"
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void) {
  printf("ne_world_params %zu\n", sizeof(ne_world_params));
  printf("ne_bodies %zu\n", sizeof(ne_bodies));
  printf("ne_gpu_config %zu\n", sizeof(ne_gpu_config));
  printf("ne_tune_result %zu\n", sizeof(ne_tune_result));
  O(ne_world_params, half); O(ne_world_params, periodic); O(ne_world_params, dt); O(ne_world_params, gravity);
  O(ne_world_params, contact_beta); O(ne_world_params, warm_start); O(ne_world_params, solver_iterations);
  O(ne_world_params, sleep_lin_vel); O(ne_world_params, wake_vel);
  O(ne_bodies, n); O(ne_bodies, x); O(ne_bodies, inv_mass); O(ne_bodies, shape); O(ne_bodies, flags);
  O(ne_bodies, hull_offsets); O(ne_bodies, nocollide);
  O(ne_gpu_config, cell_scale);
  O(ne_tune_result, broadphase); O(ne_tune_result, ms_per_step); O(ne_tune_result, trials);
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write(prog)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", exe, c], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return dict((k, int(v)) for k, v in (line.split() for line in out.splitlines()))


def test_ctypes_mirror_matches_c_layout():
    lay = _c_layout()
    assert lay["ne_world_params"] == C.sizeof(_abi.WorldParams)
    assert lay["ne_bodies"] == C.sizeof(_abi.Bodies)
    assert lay["ne_gpu_config"] == C.sizeof(_abi.GpuConfig)
    assert lay["ne_tune_result"] == C.sizeof(_abi.TuneResult)
    for key, off in lay.items():
        if "." not in key:
            continue
        t, f = key.split(".")
        cls = {"ne_world_params": _abi.WorldParams, "ne_bodies": _abi.Bodies, "ne_gpu_config": _abi.GpuConfig,
               "ne_tune_result": _abi.TuneResult}[t]
        assert getattr(cls, f).offset == off, key


def test_oracle_and_reference_harness_export_the_test_api(oracle_lib):
    import engines
    lib = C.CDLL(engines.ORACLE_LIB)
    for n in ["create", "destroy", "set_params", "set_bodies", "step", "get_state", "get_contacts", "pair_contacts",
              "diagnostics", "min_image", "wrap"]:
        assert hasattr(lib, "ne_oracle_" + n)
    if engines.ref_available():
        rl = C.CDLL(engines.REF_LIB)
        for n in ["create", "set_params", "set_bodies", "step", "get_state", "get_contacts", "pair_contacts"]:
            assert hasattr(rl, "ne_ref_" + n)


def test_create_without_device_fails_cleanly():
    """No GPU here: creating a context reports an error code, nothing crashes,
    and there is no CPU fallback."""
    try:
        w = gpu.GpuWorld()
    except gpu.NeGpuError as e:
        assert e.code in (_abi.NE_ERR_NO_DEVICE, _abi.NE_ERR_HIP)
        return
    w.close()
    pytest.skip("a GPU is visible")


def test_invalid_arguments_rejected_without_device():
    lib = gpu.load_library()
    assert lib.ne_gpu_create(None, None) == _abi.NE_ERR_INVALID
    assert lib.ne_gpu_step(None, 1) == _abi.NE_ERR_INVALID
    assert lib.ne_gpu_set_params(None, None) == _abi.NE_ERR_INVALID
