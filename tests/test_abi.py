"""The C ABI without a GPU: libmray_hip.so loads and exports every function
include/mray_hip.h declares, the Python binding declares the same set, and
the ctypes mirrors of the parameter structs have the C layouts (sizes and
field offsets from the header). No compute call is made."""
import ctypes as C
import pathlib
import re

import pytest

from mray_amd import hip
from mray_amd.types import Camera, GenParams, RecParams, RefParams

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "mray_hip.h"
TYPES = ROOT / "include" / "mray_types.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(mray_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    if not hip.LIB_PATH.exists():
        pytest.skip("libmray_hip.so not built (make -C mray_amd)")
    lib = C.CDLL(str(hip.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_the_header():
    assert sorted(hip._SIGNATURES) == declared_functions()


def struct_fields(name):
    """(type, field) pairs of a typedef struct in mray_types.h."""
    text = re.sub(r"/\*.*?\*/", "", TYPES.read_text(), flags=re.S)
    body = re.search(r"typedef struct " + name + r"\s*\{(.*?)\}\s*" + name + r"\s*;", text, re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if decl:
            t, f = decl.rsplit(None, 1)
            fields.append((t, re.sub(r"\[.*\]", "", f)))
    return fields


@pytest.mark.parametrize("cname, cls", [("mray_camera", Camera), ("mray_gen_params", GenParams),
                                        ("mray_rec_params", RecParams), ("mray_ref_params", RefParams)])
def test_parameter_structs_match_header(cname, cls):
    fields = struct_fields(cname)
    assert [f for _, f in fields] == [f for f, _ in cls._fields_]
    for (t, _), (_, ct) in zip(fields, cls._fields_):
        base = ct._type_ if hasattr(ct, "_length_") else ct
        expect = {"float": C.c_float, "int32_t": C.c_int32}[t]
        assert base is expect


def test_host_library_exports_every_declared_function():
    from mray_amd import moment_file

    if not moment_file.LIB_PATH.exists():
        pytest.skip("libmray_host.so not built (make -C mray_amd)")
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "mray_host.h").read_text(), flags=re.S)
    names = sorted(set(re.findall(r"\b(mray_host_[a-z0-9_]+)\s*\(", text)))
    assert len(names) >= 4
    lib = C.CDLL(str(moment_file.LIB_PATH))
    assert not [n for n in names if not hasattr(lib, n)]
