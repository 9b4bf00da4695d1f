"""The gfx950 code of every kernel source, compiled with the library's own flags, keeps the
floating-point behaviour the bit-exact parity rests on: f32 denormals preserved on input and
output (the reference's f32 arithmetic has no flush to zero) and no packed-FP32 instructions
(their selection is disabled; the kernels use scalar f32 operations only).  Runs on the CPU:
hipcc cross-compiles to assembly."""
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "jagua-rs_amd"))
import build as B  # noqa: E402

KERNEL_SOURCES = [s for s in B.SOURCES if s.endswith(".hip")]


def _asm(src, out_dir):
    out = os.path.join(out_dir, src + ".s")
    cmd = [B.HIPCC, "-x", "hip", *B.COMMON, *B.DEVICE, "--cuda-device-only", "-S", "-I", B.CSRC,
           os.path.join(B.CSRC, src), "-o", out]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    return src, open(out).read()


@pytest.mark.skipif(not os.path.exists(B.HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_kernels_keep_denormals_and_use_no_packed_fp32(tmp_path):
    with cf.ThreadPoolExecutor(max_workers=min(6, len(KERNEL_SOURCES))) as ex:
        outs = dict(ex.map(lambda s: _asm(s, str(tmp_path)), KERNEL_SOURCES))
    n_kernels = 0
    for src, asm in outs.items():
        modes = re.findall(r"\.amdhsa_float_denorm_mode_32\s+(\d+)", asm)
        assert modes, f"{src}: no kernel descriptors"
        assert set(modes) == {"3"}, f"{src}: f32 denormal modes {sorted(set(modes))}"
        packed = sorted(set(re.findall(r"\bv_pk_(?:add|mul|fma|mov)_f32\b", asm)))
        assert not packed, f"{src}: packed FP32 instructions {packed}"
        n_kernels += len(modes)
    assert n_kernels >= 30
