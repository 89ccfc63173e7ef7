"""The engine's cache bookkeeping (csrc/slot_cache.h, the behaviour of src/framecache.zig:21-115)
on the CPU: a C++ check program compiled with g++ and run under a time limit. It covers duplicate
keys sharing one fill, forgetting an unfilled claim, least-recently-used reuse, requests larger
than the table, and a multithreaded stress run (clamped temporal windows, 10 % of fills failing)
that asserts one filler per key, resident buffers holding their key, and no deadlock."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "vapoursynth-zipcu_amd", "csrc")


def test_lease_table_properties(tmp_path):
    exe = str(tmp_path / "lease_table_check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-pthread", "-Wall", "-Wextra", f"-I{CSRC}",
                    os.path.join(HERE, "cpp", "lease_table_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stderr
    assert "lease table: ok" in out.stdout
