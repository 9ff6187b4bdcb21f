"""The drop-in: the reference-side shim igaos_amd/integration/pdhg_hip.cpp
(igaos::pdhg::solve over the C-ABI) compiles against the reference's own
headers (include/igaos/*.h, src/io/model.hpp, src/pdhg/pdhg.hpp), links with
libigaos_pdhg.so and solves an MPS file parsed by the reference's
io::parse_mps.  Needs the reference sources (skipped where they are
absent) and g++; without a GPU the engine reports status "error" with the
reference's no-GPU wording ("CUDA"), as src/api/engine.hpp:37 does."""
import os
import shutil
import subprocess

import pytest

from tests.conftest import netlib_path

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "src/io/model.hpp")),
                    reason="reference sources not present")
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_reference_side_shim_builds_and_solves(tmp_path, objectives):
    lib = os.path.join(ROOT, "igaos_amd")
    exe = str(tmp_path / "solve_mps")
    cmd = ["g++", "-std=c++20", "-O1",
           "-I", os.path.join(REF, "include"), "-I", os.path.join(REF, "src/io"),
           "-I", os.path.join(REF, "src/pdhg"), "-I", os.path.join(ROOT, "include"),
           os.path.join(lib, "integration", "pdhg_hip.cpp"),
           os.path.join(lib, "integration", "solve_mps.cpp"),
           "-L", lib, "-ligaos_pdhg", f"-Wl,-rpath,{lib}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe, netlib_path("afiro"), "1e-6"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    status, obj = out.stdout.split()[:2]
    import igaos_amd
    if igaos_amd.has_pdhg:
        ref = objectives["netlib/afiro"]["highs_ref"]
        assert status == "near-optimal"
        assert abs(float(obj) - ref) <= 1e-6 * (1 + abs(ref))
    else:
        assert status == "error" and "CUDA" in out.stdout
