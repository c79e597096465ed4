"""The drop-in claim of the C++ API: the reference's own driver,
src/main.cpp (17-123), unchanged, compiles against this build's operator
headers (ipcc2021-final_amd/host/include: lattice_fermion.h, lattice_gauge.h,
invert.h, check.h, load_gauge.h, ...) and links against libipcc_api.so and
libwilson.so plus MPICH, in place of the reference's src/*.cpp
(src/Makefile:1-34).  Runs where the reference tree is present (this
container); the binary the checker makefile builds the same way
(oracle/_ref/main_dropin) is run on the GPU by tests/test_gpu_main.py."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_MAIN = "/root/reference/src/main.cpp"
MPI = "/opt/conda"
API = os.path.join(ROOT, "ipcc2021-final_amd")


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="the reference tree is not present")
@pytest.mark.skipif(not os.path.exists(os.path.join(MPI, "include", "mpi.h")),
                    reason="MPICH headers not present")
def test_reference_main_compiles_and_links_against_the_api(tmp_path):
    for lib in ("libipcc_api.so", "libwilson.so"):
        assert os.path.exists(os.path.join(API, lib)), f"build {lib} first"
    exe = str(tmp_path / "main_dropin")
    cmd = ["g++", "-std=c++11", "-O2", f"-I{MPI}/include", f"-I{API}/host/include",
           f"-I{ROOT}/include", "-o", exe, REF_MAIN, f"-L{API}", "-lipcc_api", "-lwilson",
           f"{MPI}/lib/libmpi.so", f"-Wl,-rpath-link,{MPI}/lib", f"-Wl,-rpath,{API}"]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    # the operator calls resolve to this build's libraries, not the reference's
    nm = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True)
    undefined = nm.stdout
    for sym in ("CGinvert", "load_gauge", "check"):
        assert sym in undefined, sym
    p = subprocess.run(["ldd", exe], capture_output=True, text=True)
    assert os.path.join(API, "libipcc_api.so") in p.stdout
