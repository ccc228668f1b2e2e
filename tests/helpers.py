"""Shared helpers of the test suite."""
import numpy as np

OUT_FIELDS = ["score", "tle", "gtle", "qle", "gscore", "max_off"]


def diff_mask(a, b):
    d = np.zeros(len(a), bool)
    for f in OUT_FIELDS:
        d |= a[f] != b[f]
    return d


def assert_same_outputs(got, want, label=""):
    d = diff_mask(got, want)
    if d.any():
        k = np.where(d)[0][:5]
        rows = [(int(i), [int(got[f][i]) for f in OUT_FIELDS], [int(want[f][i]) for f in OUT_FIELDS])
                for i in k]
        raise AssertionError(f"{label}: {int(d.sum())} of {len(d)} pairs differ; first: {rows}")


class DeviceArrays:
    """numpy arrays copied into device memory through libamdhip64 (tests of
    the device-pointer entry points)."""

    def __init__(self, *arrays):
        import ctypes
        self._ct = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.ptrs = []
        for a in arrays:
            d = ctypes.c_void_p()
            assert self.hip.hipMalloc(ctypes.byref(d), max(a.nbytes, 1)) == 0
            assert self.hip.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0
            self.ptrs.append(d)

    def ptr(self, k):
        return self.ptrs[k].value

    def upload(self, k, a):
        assert self.hip.hipMemcpy(self.ptrs[k], a.ctypes.data, a.nbytes, 1) == 0

    def download(self, k, a):
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptrs[k], a.nbytes, 2) == 0
        return a

    def close(self):
        for d in self.ptrs:
            self.hip.hipFree(d)
        self.ptrs = []


def check_c_layouts(tmp_path, root, header, dtypes=(), structs=()):
    """Compiles `header` as C11 with warnings as errors, together with
    static assertions that each C struct has the size and field offsets of
    its Python mirror: dtypes are (C name, numpy dtype), structs are (C name,
    ctypes.Structure subclass)."""
    import ctypes
    import os
    import subprocess
    checks = []
    for cname, dt in dtypes:
        checks.append(f"_Static_assert(sizeof({cname}) == {dt.itemsize}, \"{cname}\");")
        for name in dt.names:
            off = dt.fields[name][1]
            checks.append(f"_Static_assert(offsetof({cname}, {name}) == {off}, \"{cname}.{name}\");")
    for cname, cls in structs:
        checks.append(f"_Static_assert(sizeof({cname}) == {ctypes.sizeof(cls)}, \"{cname}\");")
        for name, _ in cls._fields_:
            off = getattr(cls, name).offset
            checks.append(f"_Static_assert(offsetof({cname}, {name}) == {off}, \"{cname}.{name}\");")
    src = tmp_path / "layout_check.c"
    src.write_text(f"#include <stddef.h>\n#include \"{header}\"\n" + "\n".join(checks) +
                   "\nint main(void) { return 0; }\n")
    subprocess.check_call(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-I",
                           os.path.join(root, "include"), "-c", str(src), "-o",
                           str(tmp_path / "layout_check.o")])
