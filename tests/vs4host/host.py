"""Python side of the VapourSynth API v4 test host (vs4host.cpp).

    core = Host(); core.load(LIB)                       # calls VapourSynthPluginInit2
    src = core.source(frames, YUV444PS)                 # frames: list of [plane arrays]
    node = core.invoke("BM3D", clip=src, sigma=[5, 3, 3], radius=1)
    planes, props = node.frame(0)

`invoke` raises HostError with the message the plugin (or the host's signature check) put
into the output map, as VapourSynth raises vs.Error.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, "_build", "libvs4host.so")


class HostError(Exception):
    pass


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(HOST_LIB):
        build()
    L = C.CDLL(HOST_LIB)
    vp, cp, i, i64 = C.c_void_p, C.c_char_p, C.c_int, C.c_int64
    sig = {
        "h_core_new": (vp, []), "h_load_plugin": (i, [vp, cp, cp, i]),
        "h_plugin_string": (cp, [vp, i]), "h_plugin_version": (i, [vp, i]), "h_function": (cp, [vp, i, i]),
        "h_map_new": (vp, []), "h_map_free": (None, [vp]), "h_map_set_int": (None, [vp, cp, i64]),
        "h_map_set_float": (None, [vp, cp, C.c_double]), "h_map_set_data": (None, [vp, cp, cp, i]),
        "h_map_set_node": (None, [vp, cp, vp]), "h_map_set_empty": (None, [vp, cp, i]), "h_map_error": (cp, [vp]), "h_map_get_node": (vp, [vp, cp]),
        "h_invoke": (vp, [vp, cp, vp]),
        "h_source_new": (vp, [i, i, i, i, i, i, i, i, vp]), "h_node_free": (None, [vp]),
        "h_node_info": (None, [vp, C.POINTER(C.c_int)]), "h_node_name": (cp, [vp]),
        "h_node_dep": (i, [vp, i, C.POINTER(C.c_int)]),
        "h_get_frame": (vp, [vp, i, cp, i]), "h_frame_free": (None, [vp]),
        "h_frame_plane": (i, [vp, i, vp]), "h_frame_dims": (None, [vp, i, C.POINTER(C.c_int)]),
        "h_frame_prop_ints": (i, [vp, cp, C.POINTER(C.c_int64), i]),
        "h_live_frames": (i, []), "h_live_nodes": (i, []), "h_misuse": (i, []), "h_misuse_log": (cp, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


_L = None


def lib():
    global _L
    if _L is None:
        _L = _load()
    return _L


class Node:
    def __init__(self, ptr):
        self.ptr = ptr

    def free(self):
        if self.ptr:
            lib().h_node_free(self.ptr)
            self.ptr = None

    def info(self):
        a = (C.c_int * 9)()
        lib().h_node_info(self.ptr, a)
        keys = ("color_family", "sample_type", "bits", "ssw", "ssh", "num_planes", "width", "height", "num_frames")
        return dict(zip(keys, list(a)))

    @property
    def name(self):
        return lib().h_node_name(self.ptr).decode()

    def deps(self):
        out, k = [], 0
        p = C.c_int()
        while lib().h_node_dep(self.ptr, k, C.byref(p)) == 0:
            out.append(p.value)
            k += 1
        return out

    def frame(self, n, props=("BM3D_V_radius", "BM3D_V_process", "_SourceFrame")):
        err = C.create_string_buffer(512)
        f = lib().h_get_frame(self.ptr, n, err, 512)
        if not f:
            raise HostError(err.value.decode())
        try:
            planes = []
            for p in range(self.info()["num_planes"]):
                wh = (C.c_int * 2)()
                lib().h_frame_dims(f, p, wh)
                a = np.empty((wh[1], wh[0]), np.float32)
                lib().h_frame_plane(f, p, a.ctypes.data)
                planes.append(a)
            pr = {}
            for k in props:
                buf = (C.c_int64 * 8)()
                cnt = lib().h_frame_prop_ints(f, k.encode(), buf, 8)
                if cnt >= 0:
                    pr[k] = list(buf[: min(cnt, 8)])
            return planes, pr
        finally:
            lib().h_frame_free(f)


class Host:
    def __init__(self):
        self.core = lib().h_core_new()
        self.functions = {}

    def load(self, path):
        err = C.create_string_buffer(512)
        n = lib().h_load_plugin(self.core, path.encode(), err, 512)
        if n < 0:
            raise HostError(err.value.decode())
        for k in range(n):
            name, args, ret = (lib().h_function(self.core, k, w).decode() for w in range(3))
            self.functions[name] = (args, ret)
        return n

    def plugin(self):
        s = [lib().h_plugin_string(self.core, w).decode() for w in range(3)]
        return {"id": s[0], "namespace": s[1], "name": s[2], "version": lib().h_plugin_version(self.core, 0),
                "api": lib().h_plugin_version(self.core, 1)}

    def source(self, frames, fmt):
        """frames: list of per-frame plane lists (float32 2-D arrays)."""
        h, w = frames[0][0].shape
        flat = np.concatenate([np.ascontiguousarray(p, np.float32).ravel() for fr in frames for p in fr])
        ptr = lib().h_source_new(fmt.color_family, fmt.sample_type, fmt.bits_per_sample, fmt.sub_sampling_w,
                                 fmt.sub_sampling_h, w, h, len(frames), flat.ctypes.data)
        return Node(ptr)

    def invoke(self, name, **kwargs):
        m = lib().h_map_new()
        try:
            for k, v in kwargs.items():
                vals = v if isinstance(v, (list, tuple)) else [v]
                if not vals:          # an empty list still sets the key, as in VapourSynth
                    lib().h_map_set_empty(m, k.encode(), 1)
                for x in vals:
                    if isinstance(x, Node):
                        lib().h_map_set_node(m, k.encode(), x.ptr)
                    elif isinstance(x, (bool, int, np.integer)):
                        lib().h_map_set_int(m, k.encode(), int(x))
                    elif isinstance(x, (float, np.floating)):
                        lib().h_map_set_float(m, k.encode(), float(x))
                    elif isinstance(x, (str, bytes)):
                        b = x.encode() if isinstance(x, str) else x
                        lib().h_map_set_data(m, k.encode(), b, len(b))
                    else:
                        raise TypeError(f"{k}: unsupported value {x!r}")
            out = lib().h_invoke(self.core, name.encode(), m)
        finally:
            lib().h_map_free(m)
        try:
            e = lib().h_map_error(out)
            if e is not None:
                raise HostError(e.decode())
            return Node(lib().h_map_get_node(out, b"clip"))
        finally:
            lib().h_map_free(out)


def live():
    return lib().h_live_frames(), lib().h_live_nodes()


def misuse():
    return lib().h_misuse(), lib().h_misuse_log().decode()
