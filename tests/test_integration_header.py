"""The reference-side binding (integration/gpu_dem_model.h, a model::DEMModel
subclass over include/pdem.h) compiles against the reference's own headers:
g++ -fsyntax-only on a translation unit that constructs the model and calls
run(), with no error and no warning from the binding itself.  Skipped where
the reference tree is absent."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"

TU = """#include "gpu_dem_model.h"
int main(int, char **argv) {
  inp::Input deck(argv[1]);
  pdem_binding::GpuDEMModel model(&deck, PDEM_PREC_MIXED);
  model.run(&deck);
  return 0;
}
"""


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")) or shutil.which("g++") is None,
                    reason="reference tree or g++ not available")
def test_binding_compiles_against_reference_headers(tmp_path):
    src = tmp_path / "binding_tu.cpp"
    src.write_text(TU)
    inc = [os.path.join(ROOT, "integration"), os.path.join(ROOT, "include"), os.path.join(REF, "src"),
           os.path.join(REF, "external"), os.path.join(REF, "external", "fmt", "include")]
    if os.path.isdir("/opt/conda/include"):  # mpi.h for the reference's util headers
        inc.append("/opt/conda/include")
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra"] + ["-I" + p for p in inc] + [str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    ours = [ln for ln in r.stderr.splitlines()
            if re.search(r"(gpu_dem_model\.h|binding_tu\.cpp):\d+:\d+: (error|warning)", ln)]
    assert r.returncode == 0, r.stderr[-4000:]
    assert not ours, "\n".join(ours)
