"""CPU check of the built library's gfx950 code objects: no kernel uses scratch (private) memory.
A kernel that spills registers or keeps an array it indexes at run time in scratch memory runs
far slower (the batched SOR kernels once did, at 128 bytes per lane); the code objects' own
metadata (.private_segment_fixed_size) shows it without a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fluid-dynamics1_amd", "libcnv.so")
LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def code_objects(tmp_path):
    """The gfx950 code object of each bundle in the library's fat binary section."""
    fb = tmp_path / "fatbin.bin"
    subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", LIB, str(fb)], check=True)
    data = fb.read_bytes()
    starts = [i for i in range(len(data)) if data.startswith(MAGIC, i)]
    out = []
    for k, s in enumerate(starts):
        part = tmp_path / ("bundle%d.bin" % k)
        part.write_bytes(data[s:starts[k + 1] if k + 1 < len(starts) else len(data)])
        co = tmp_path / ("bundle%d.co" % k)
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", "--input=" + str(part),
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", "--output=" + str(co)], check=True)
        out.append(co)
    return out


@pytest.mark.skipif(not (os.path.exists(LIB) and shutil.which("objcopy")
                         and os.path.exists(os.path.join(LLVM, "llvm-readelf"))),
                    reason="needs the built library and the ROCm LLVM tools")
def test_no_kernel_uses_scratch_memory(tmp_path):
    kernels = {}
    for co in code_objects(tmp_path):
        notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", str(co)], check=True,
                               capture_output=True, text=True).stdout
        name = None
        for line in notes.splitlines():
            line = line.strip()
            if line.startswith(".name:"):
                name = line.split(":", 1)[1].strip()
            elif line.startswith(".private_segment_fixed_size:") and name:
                kernels[name] = int(line.split(":", 1)[1])
                name = None
    sor = [k for k in kernels if "sor_rb" in k]
    assert len(sor) >= 100 and any("batch" in k for k in sor), len(sor)
    assert {k: v for k, v in kernels.items() if v != 0} == {}
