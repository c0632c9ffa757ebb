"""CPU tests of the C ABI: the in-tree library loads and exports every
function that include/*.h declares; host-side setup entry points work
without a device."""
import ctypes as C
import os
import re

import numpy as np

import i3elvis_amd as I

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for f in sorted(os.listdir(inc)):
        if not f.endswith(".h"):
            continue
        src = open(os.path.join(inc, f)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(i3g_[a-z_0-9]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    lib = I.lib()
    names = declared_functions()
    assert len(names) >= 30, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_error_strings():
    for rc in (0, -1, -2, -3, -4, -5, -6):
        s = I.lib().i3g_error_string(rc)
        assert s and len(s) >= 2


def test_host_setup_without_device():
    h = I.hierarchy(257, 257, 257, 5)
    assert list(h.mgx[:6]) == [257, 131, 68, 36, 20, 12]
    rules, nr = I.freeslip_rules()
    bm, bv, bn = I.expand_bc(rules, nr, I.hierarchy(17, 17, 17, 1))
    assert bm.dtype == np.int32 and len(bn) * 2 == len(bv)
