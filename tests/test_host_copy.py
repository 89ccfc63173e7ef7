"""The filter layer's plane copies (csrc/host_copy.h: non-temporal AVX2 stores where the host
has AVX2, memcpy otherwise) on the CPU: a C++ check program compiled with g++ compares them with
row-by-row memcpy over random widths, row counts, strides and misalignments."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "vapoursynth-zipcu_amd", "csrc")


def test_host_plane_copy_matches_memcpy(tmp_path):
    exe = str(tmp_path / "host_copy_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", f"-I{CSRC}",
                    os.path.join(HERE, "cpp", "host_copy_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "host copy: ok" in out.stdout
