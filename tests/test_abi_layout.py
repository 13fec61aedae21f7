"""The Julia binding (julia/FastMultipoleHIP.jl) and the Python mirror (fastmultipole_amd/_lib.py)
lay out every struct of include/fmmhip.h exactly as the C compiler does: same field order, sizes and
offsets, same total size. The C side is measured with offsetof/sizeof (tests/abi_layout.c, built
with gcc); Julia lays out isbits structs with C's alignment rules, so its offsets follow from the
field types parsed out of the module."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from fastmultipole_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# C struct -> (Julia mirror, ctypes mirror)
MIRRORS = {
    "fmmhip_options": ("FmmHipOptions", _lib.Options),
    "fmmhip_stats": ("FmmHipStats", _lib.Stats),
    "fmmhip_request": ("FmmHipRequest", _lib.Request),
    "fmmhip_request_stats": ("FmmHipRequestStats", _lib.RequestStats),
    "fmmhip_server_stats_t": ("FmmHipServerStats", _lib.ServerStats),
}


@pytest.fixture(scope="module")
def c_layout():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "abi_layout")
        subprocess.run(["gcc", "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", exe,
                        os.path.join(ROOT, "tests", "abi_layout.c")], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    lay = {}
    for line in out.splitlines():
        s, f, off, size = line.split()
        lay.setdefault(s, []).append((f, int(off), int(size)))
    return lay


def julia_structs():
    text = open(os.path.join(ROOT, "julia", "FastMultipoleHIP.jl")).read()
    out = {}
    for m in re.finditer(r"^(?:mutable\s+)?struct\s+(\w+)\s*(?:#[^\n]*)?\n(.*?)^end", text, flags=re.S | re.M):
        fields = []
        for line in m.group(2).splitlines():
            fm = re.match(r"\s*(\w+)::([\w{}]+)\s*$", line.split("#")[0])
            if fm:
                fields.append((fm.group(1), fm.group(2)))
        out[m.group(1)] = fields
    return out


_PRIM = {"Int32": 4, "Cint": 4, "Int64": 8, "Float64": 8, "Cdouble": 8}


def julia_layout(name, structs):
    """(fields with offset and size, total size, alignment) under C rules."""
    off, align, rows = 0, 1, []
    for f, t in structs[name]:
        if t in _PRIM:
            size = al = _PRIM[t]
        elif t.startswith("Ptr{"):
            size = al = 8
        else:
            _, size, al = julia_layout(t, structs)
        off = (off + al - 1) // al * al
        rows.append((f, off, size))
        off += size
        align = max(align, al)
    return rows, (off + align - 1) // align * align, align


@pytest.mark.parametrize("cname", sorted(MIRRORS))
def test_julia_mirror_matches_c_layout(cname, c_layout):
    jname, _ = MIRRORS[cname]
    structs = julia_structs()
    assert jname in structs, jname
    rows, size, _ = julia_layout(jname, structs)
    c_rows = [r for r in c_layout[cname] if r[0] != "sizeof"]
    c_size = [r[1] for r in c_layout[cname] if r[0] == "sizeof"][0]
    assert rows == c_rows, (cname, rows, c_rows)
    assert size == c_size


@pytest.mark.parametrize("cname", sorted(MIRRORS))
def test_python_mirror_matches_c_layout(cname, c_layout):
    _, cls = MIRRORS[cname]
    rows = [(f, getattr(cls, f).offset, getattr(cls, f).size) for f, _ in cls._fields_]
    c_rows = [r for r in c_layout[cname] if r[0] != "sizeof"]
    c_size = [r[1] for r in c_layout[cname] if r[0] == "sizeof"][0]
    assert rows == c_rows, (cname, rows, c_rows)
    assert ctypes.sizeof(cls) == c_size


def test_julia_binding_calls_only_declared_entry_points():
    """Every ccall of the Julia module names a function the header declares (and the library exports)."""
    text = open(os.path.join(ROOT, "julia", "FastMultipoleHIP.jl")).read()
    called = set(re.findall(r"ccall\(\(:(fmmhip_\w+),", text))
    assert called and called <= set(_lib.EXPORTED), called - set(_lib.EXPORTED)
