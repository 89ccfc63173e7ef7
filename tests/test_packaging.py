"""Wheel layout and the installed-plugin smoke check (reference: hatch_build.py:49-50, 96-97 and
.github/smoke.py): the wheel puts the plugin where VapourSynth autoloads it, the plugin loads
and registers from there, and creating a filter either renders a frame (GPU present) or fails
with a clean error (no GPU) instead of crashing. The VapourSynth core is played by the test host.
"""
import base64
import hashlib
import os
import sys
import zipfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "vs4host"))
sys.path.insert(0, os.path.join(REPO, "vapoursynth-zipcu_amd"))

import build_wheel  # noqa: E402
import host  # noqa: E402
from vszipcu import GRAYS  # noqa: E402


@pytest.fixture(scope="module")
def wheel(tmp_path_factory):
    out = tmp_path_factory.mktemp("wheel")
    return build_wheel.build_wheel(str(out), build=False)


def test_wheel_layout_and_record(wheel):
    assert os.path.basename(wheel) == "vapoursynth_vszipcu_amd-0.1.0-py3-none-linux_x86_64.whl"
    with zipfile.ZipFile(wheel) as z:
        names = set(z.namelist())
        assert "vapoursynth/plugins/vszipcu/libvszipcu_amd.so" in names
        assert z.read("vapoursynth/plugins/vszipcu/manifest.vs") == b"[VapourSynth Manifest V1]\nlibvszipcu_amd\n"
        record = z.read("vapoursynth_vszipcu_amd-0.1.0.dist-info/RECORD").decode().splitlines()
        for line in record:
            path, digest, size = line.split(",")
            if not digest:
                continue
            data = z.read(path)
            assert int(size) == len(data)
            want = base64.urlsafe_b64encode(hashlib.sha256(data).digest()).rstrip(b"=").decode()
            assert digest == "sha256=" + want, path
        assert {p.split(",")[0] for p in record} == names
        assert "Root-Is-Purelib: false" in z.read("vapoursynth_vszipcu_amd-0.1.0.dist-info/WHEEL").decode()


def test_installed_plugin_smoke(wheel, tmp_path):
    site = tmp_path / "site-packages"
    with zipfile.ZipFile(wheel) as z:
        z.extractall(site)
    plugin = site / "vapoursynth" / "plugins" / "vszipcu" / "libvszipcu_amd.so"
    assert plugin.is_file()
    core = host.Host()
    assert core.load(str(plugin)) == 3
    assert core.plugin()["namespace"] == "vszipcu"
    assert sorted(core.functions) == ["BM3D", "BM3Dv2", "VAggregate"]
    src = core.source([[np.full((256, 256), 0.5, np.float32)]] * 2, GRAYS)
    try:
        try:
            clip = core.invoke("BM3D", clip=src, sigma=[1.5])
        except host.HostError as e:       # no GPU: creation fails cleanly
            assert str(e) == "BM3D: HIP initialization failed (no usable GPU)."
        else:                              # GPU present: a frame renders
            try:
                planes, _ = clip.frame(0)
                assert np.allclose(planes[0], 0.5, atol=1e-6)
            finally:
                clip.free()
    finally:
        src.free()
