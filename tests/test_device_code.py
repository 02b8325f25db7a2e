"""The engine library's gfx950 device code holds no packed FP32 arithmetic.

The build passes `-Xclang -target-feature -Xclang -packed-fp32-ops` (the
minus removes the feature) so that no kernel uses v_pk_add_f32,
v_pk_mul_f32 or v_pk_fma_f32; hipcc's host pass warns that x86-64 does not
know that feature, which says nothing about the device pass.  This test
reads the code objects of the built library itself: it splits the
.hip_fatbin section into its offload bundles (one per translation unit),
unbundles the gfx950 code object of each and disassembles it with the ROCm
LLVM tools.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "peridem_amd", "lib", "libpdem_hip.so")
LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
PACKED_FP32 = ("v_pk_add_f32", "v_pk_mul_f32", "v_pk_fma_f32")


def _tools_present():
    return all(os.path.exists(os.path.join(LLVM, t)) for t in ("llvm-objcopy", "clang-offload-bundler",
                                                              "llvm-objdump"))


def _disassemble(tmp_path):
    fat = tmp_path / "fat.bin"
    subprocess.run([os.path.join(LLVM, "llvm-objcopy"), "--dump-section", f".hip_fatbin={fat}", LIB,
                    str(tmp_path / "lib_copy.so")], check=True, capture_output=True)
    data = fat.read_bytes()
    offs = []
    i = data.find(MAGIC)
    while i >= 0:
        offs.append(i)
        i = data.find(MAGIC, i + 1)
    assert offs, "no offload bundle in .hip_fatbin"
    texts = []
    for k, o in enumerate(offs):
        end = offs[k + 1] if k + 1 < len(offs) else len(data)
        b = tmp_path / f"bundle{k}.bin"
        b.write_bytes(data[o:end])
        co = tmp_path / f"gfx950_{k}.co"
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", f"--input={b}",
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True,
                       capture_output=True)
        out = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", str(co)], check=True,
                             capture_output=True, text=True).stdout
        texts.append(out)
    return texts


@pytest.mark.skipif(not os.path.exists(LIB), reason="engine library not built")
@pytest.mark.skipif(not _tools_present(), reason="ROCm LLVM tools absent")
def test_no_packed_fp32_in_device_code(tmp_path):
    texts = _disassemble(tmp_path)
    assert len(texts) >= 2  # engine and network translation units
    n_insts = 0
    for t in texts:
        n_insts += t.count("\n\t")
        for op in PACKED_FP32:
            assert op not in t, f"{op} found in the device code"
    # the disassembly is real code: the closure's MFMA products are in it
    assert any("v_mfma_f32_16x16x16" in t for t in texts)
    assert n_insts > 10000
