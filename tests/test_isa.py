"""The device code of libgpufem contains no scalar-cache writes (scalar
stores, scalar atomics, scalar cache write-back): every global write goes
through vector memory instructions.  Checked on the gfx950 code objects the
build produced (our own objects, extracted and disassembled here)."""
import glob
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LLVM = "/opt/rocm/llvm/bin"
# the objects of the current sources (a stale object of a removed source is skipped)
SRCS = sorted(glob.glob(os.path.join(ROOT, "gpufem_amd", "csrc", "*.hip")))
OBJS = [o for o in (os.path.join(ROOT, "gpufem_amd", "_build", os.path.basename(s) + ".o")
                    for s in SRCS) if os.path.exists(o)]
FORBIDDEN = re.compile(r"^\s*(s_store_|s_buffer_store|s_atomic_|s_buffer_atomic|s_dcache_wb|"
                       r"s_scratch_store|s_dcache_discard)")


def disassemble(obj, tmp):
    fat = os.path.join(tmp, "fat.bin")
    dev = os.path.join(tmp, "dev.co")
    subprocess.run([LLVM + "/llvm-objcopy", "--dump-section=.hip_fatbin=" + fat, obj,
                    os.path.join(tmp, "host.o")], check=True)
    subprocess.run([LLVM + "/clang-offload-bundler", "--unbundle", "--type=o",
                    "--input=" + fat, "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                    "--output=" + dev], check=True)
    out = subprocess.run([LLVM + "/llvm-objdump", "-d", dev], check=True, capture_output=True,
                         text=True).stdout
    return out


@pytest.mark.skipif(not OBJS or not os.path.exists(LLVM + "/llvm-objdump"),
                    reason="library objects or ROCm LLVM tools not present")
@pytest.mark.parametrize("obj", OBJS, ids=[os.path.basename(o) for o in OBJS])
def test_no_scalar_cache_writes(obj):
    with tempfile.TemporaryDirectory() as tmp:
        asm = disassemble(obj, tmp)
    src = os.path.join(ROOT, "gpufem_amd", "csrc", os.path.basename(obj)[:-2])
    if "__global__" in open(src).read():
        assert "s_endpgm" in asm                  # we did disassemble kernels
    bad = [ln for ln in asm.splitlines() if FORBIDDEN.match(ln)]
    assert not bad, bad[:5]
