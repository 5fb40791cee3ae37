"""CPU-side checks of the C-ABI library: it loads and exports every entry point
declared in include/mesoflow_amd.h, and the ctypes mirror matches the header."""
import ctypes as C
import os
import re
import subprocess

import pytest

from mesoflow_amd import abi, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mesoflow_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(runtime.LIB_PATH), "build the library first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", runtime.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(mf_[a-z0-9_]+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_library_loads_and_binds_every_symbol():
    L = runtime.lib()
    for f in declared_functions():
        assert hasattr(L, f)
        assert f in runtime.SIGNATURES, f


def test_status_strings():
    L = runtime.lib()
    assert L.mf_status_string(abi.MF_ERR_NAN_DSDT) == b"NaN detected in dsdt"
    assert L.mf_status_string(abi.MF_ERR_NEG_PRESSURE) == b"Pressure has gone negative"
    assert L.mf_status_string(abi.MF_ERR_UNKNOWN_SCHEME) == b"unknown hyperbolics scheme"


def _c_sizeof(typename):
    src = '#include "mesoflow_amd.h"\n#include <stdio.h>\n#include <stddef.h>\nint main(){printf("%%zu\\n", sizeof(%s));return 0;}\n' % typename
    tmp = "/tmp/mf_sizeof_%s" % typename
    with open(tmp + ".c", "w") as f:
        f.write(src)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), tmp + ".c", "-o", tmp], check=True)
    return int(subprocess.run([tmp], capture_output=True, text=True, check=True).stdout)


@pytest.mark.parametrize("cname,pytype", [("mf_array4", abi.Array4), ("mf_box", abi.Box), ("mf_geom", abi.Geom),
                                          ("mf_params", abi.Params), ("mf_plugin_params", abi.PluginParams)])
def test_struct_layout_matches_header(cname, pytype):
    assert _c_sizeof(cname) == C.sizeof(pytype)


def test_no_device_here_is_reported():
    if runtime.device_count() > 0:
        pytest.skip("a GPU is present")
    from mesoflow_amd import cases
    c = cases.shocktube_n2(n_cell=(8, 4, 4))
    with pytest.raises(abi.MFError) as e:
        runtime.Context(c)
    assert e.value.status == abi.MF_ERR_NO_DEVICE


def test_library_loads_after_torch():
    """runtime.lib() imports torch before it loads the library, so a process
    that uses both keeps torch's HIP runtime (the two share sonames)."""
    import sys
    code = ("import sys; from mesoflow_amd import runtime; assert 'torch' not in sys.modules; "
            "runtime.lib(); assert 'torch' in sys.modules; "
            "maps = open('/proc/self/maps').read(); "
            "import torch, os; tlib = os.path.realpath(os.path.join(os.path.dirname(torch.__file__), 'lib', 'libamdhip64.so')); "
            "hips = {l.split()[-1] for l in maps.splitlines() if 'libamdhip64' in l}; "
            "print(sorted(hips)); assert hips == {tlib}, hips")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
