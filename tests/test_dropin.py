"""The drop-in C++ API (include/accfft_gpu.h, accfft_gpuf.h,
accfft_operators_gpu.h, transpose_cuda.h): a program written against the
reference's function names (tests/cpp/dropin_steps.cpp, modelled on
steps/step1 to step5) must compile unchanged and pass the
reference programs' own correctness thresholds on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_steps")


def test_dropin_headers_compile(tmp_path):
    """Header-only wrappers compile as plain C++ (host compiler only)."""
    src = tmp_path / "t.cpp"
    src.write_text(
        '#include "accfft_gpu.h"\n#include "accfft_gpuf.h"\n#include "accfft_operators_gpu.h"\n'
        '#include "accfft_poisson_gpu.h"\n#include "transpose_cuda.h"\n'
        "int main(){int n[3]={8,8,8},i[3],s[3],o[3],t[3];MPI_Comm c;int d[2]={0,0};"
        "accfft_create_comm(MPI_COMM_WORLD,d,&c);"
        "return accfft_local_size_dft_r2c_gpu(n,i,s,o,t,c) == 8*8*10*8 ? 0 : 1;}\n")
    exe = tmp_path / "t"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), str(src),
                        "-L", os.path.join(ROOT, "accfft_amd"), "-laccfft_amd",
                        f"-Wl,-rpath,{os.path.join(ROOT, 'accfft_amd')}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # the size query is host-only: runs without a GPU
    assert subprocess.run([str(exe)]).returncode == 0


@pytest.mark.gpu
def test_dropin_program_on_gpu(dev):
    if not os.path.exists(BIN):
        from accfft_amd.build import build_dropin_example
        build_dropin_example()
    print("running tests/cpp/dropin_steps", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL CHECKS PASSED" in r.stdout
