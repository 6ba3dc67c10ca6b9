"""The C-ABI library loads on the CPU, exports every function the header
declares, and the ctypes structures match the C layout."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from elmm_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "elmm_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(elmm_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(capi.EXPORTS)


def test_library_exports_every_symbol():
    lib = capi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert b"gfx950" in lib.elmm_version()


def _c_layout(struct_name, fields):
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"elmm_gpu.h\"\nint main(void){\n"
    src += f'printf("%zu\\n", sizeof({struct_name}));\n'
    for f in fields:
        src += f'printf("%zu\\n", offsetof({struct_name}, {f}));\n'
    src += "return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], stdout=subprocess.PIPE, text=True, check=True).stdout.split()
    return [int(x) for x in out]


@pytest.mark.parametrize("cls,cname", [(capi.Config, "elmm_config"), (capi.Lists, "elmm_lists"),
                                      (capi.Diag, "elmm_diag"), (capi.Cuboid, "elmm_cuboid"),
                                      (capi.PointSource, "elmm_point_source"), (capi.Timers, "elmm_timers"),
                                      (capi.Request, "elmm_request"), (capi.Tiles, "elmm_tiles"),
                                      (capi.RequestStatus, "elmm_request_status")])
def test_struct_layout(cls, cname):
    names = [f[0] for f in cls._fields_]
    got = _c_layout(cname, names)
    assert got[0] == C.sizeof(cls)
    assert got[1:] == [getattr(cls, n).offset for n in names]


def test_errors_are_reported_not_raised_in_c():
    cfg = capi.make_config(nx=2, ny=2, nz=2, dx=1, dy=1, dz=1, dt=0.1)
    with pytest.raises(capi.ElmmError) as e:
        capi.Case(cfg)
    assert e.value.code == -1
    assert "at least 4" in str(e.value)
