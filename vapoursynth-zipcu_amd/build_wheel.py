"""Builds the VapourSynth plugin wheel: libvszipcu_amd.so under vapoursynth/plugins/vszipcu/.

The layout is the one VapourSynth autoloads from site-packages and the reference wheel uses
(hatch_build.py:49-50, 96-97): the plugin library plus a manifest.vs naming it. The library is
built first (make -C vapoursynth-zipcu_amd); the wheel is written directly (zip + METADATA,
WHEEL, RECORD), so no build backend is needed.

    python vapoursynth-zipcu_amd/build_wheel.py [--out DIR] [--no-build]
"""
import argparse
import base64
import hashlib
import os
import subprocess
import sys
import zipfile

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libvszipcu_amd.so")
NAME = "vapoursynth_vszipcu_amd"
VERSION = "0.1.0"
TAG = "py3-none-linux_x86_64"
PLUGIN_DIR = "vapoursynth/plugins/vszipcu"


def _record_line(path, data):
    digest = base64.urlsafe_b64encode(hashlib.sha256(data).digest()).rstrip(b"=").decode()
    return f"{path},sha256={digest},{len(data)}"


def build_wheel(out_dir, build=True):
    if build:
        subprocess.run(["make", "-s", "-j", str(min(os.cpu_count() or 1, 16)), "-C", HERE], check=True)
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} was not built; the wheel would not load")
    with open(LIB, "rb") as f:
        lib = f.read()
    lib_name = os.path.basename(LIB)
    files = {
        f"{PLUGIN_DIR}/{lib_name}": lib,
        f"{PLUGIN_DIR}/manifest.vs": f"[VapourSynth Manifest V1]\n{os.path.splitext(lib_name)[0]}\n".encode(),
    }
    dist = f"{NAME}-{VERSION}.dist-info"
    files[f"{dist}/METADATA"] = (
        "Metadata-Version: 2.1\n"
        f"Name: {NAME.replace('_', '-')}\n"
        f"Version: {VERSION}\n"
        "Summary: VapourSynth vszipcu BM3D / BM3Dv2 / VAggregate for AMD MI355X (gfx950, HIP)\n"
        "Requires-Python: >=3.10\n"
        "Requires-Dist: VapourSynth>=75\n").encode()
    files[f"{dist}/WHEEL"] = (
        "Wheel-Version: 1.0\nGenerator: vszipcu-amd build_wheel.py\nRoot-Is-Purelib: false\n"
        f"Tag: {TAG}\n").encode()
    record = [_record_line(p, d) for p, d in files.items()] + [f"{dist}/RECORD,,"]
    files[f"{dist}/RECORD"] = ("\n".join(record) + "\n").encode()
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, f"{NAME}-{VERSION}-{TAG}.whl")
    with zipfile.ZipFile(path, "w", zipfile.ZIP_DEFLATED) as z:
        for p, d in files.items():
            info = zipfile.ZipInfo(p, date_time=(2026, 1, 1, 0, 0, 0))
            info.external_attr = (0o755 if p.endswith(".so") else 0o644) << 16
            info.compress_type = zipfile.ZIP_DEFLATED
            z.writestr(info, d)
    return path


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "dist"))
    ap.add_argument("--no-build", action="store_true")
    args = ap.parse_args()
    print(build_wheel(args.out, build=not args.no_build))


if __name__ == "__main__":
    sys.exit(main())
