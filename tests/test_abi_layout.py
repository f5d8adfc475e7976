"""The ctypes mirrors in gds2para_amd/_capi.py against include/gds2para.h:
every struct's size and every field's offset, as gcc lays out the header.
A field added on one side only (an option appended to gp_solve_opts, say)
shifts the ones after it and fails here instead of passing garbage."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from gds2para_amd import _capi as capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = {
    "gp_mesh": capi.GpMesh, "gp_mesh_sizes": capi.GpMeshSizes, "gp_gen_params": capi.GpGenParams,
    "gp_device_info": capi.GpDeviceInfo, "gp_v0_sizes": capi.GpV0Sizes, "gp_solve_opts": capi.GpSolveOpts,
    "gp_solve_stats": capi.GpSolveStats, "gp_hierarchy_info": capi.GpHierarchyInfo,
    "gp_autotune_info": capi.GpAutotuneInfo, "gp_vh_info": capi.GpVhInfo,
    "gp_request_report": capi.GpRequestReport, "gp_server_info": capi.GpServerInfo,
    "gp_server_backend": capi.GpServerBackend,
    "gp_column_report": capi.GpColumnReport, "gp_stream_column": capi.GpStreamColumn, "gp_v0_stats": capi.GpV0Stats,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gds2para.h"', "int main(void) {"]
    for cname, py in PAIRS.items():
        lines.append(f'  printf("{cname} . %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["  return 0;", "}"]
    d = tmp_path_factory.mktemp("abi")
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run([cc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    table = {}
    for line in out.splitlines():
        name, field, value = line.split()
        table[(name, field)] = int(value)
    return table


def test_header_struct_list_is_mirrored():
    header = open(os.path.join(ROOT, "include", "gds2para.h")).read()
    structs = set(re.findall(r"^typedef struct(?: gp_[a-z0-9_]+)? \{.*?^\} (gp_[a-z0-9_]+);", header,
                             flags=re.M | re.S))
    assert structs == set(PAIRS), sorted(structs ^ set(PAIRS))


@pytest.mark.parametrize("cname", sorted(PAIRS))
def test_struct_layout_matches_header(c_layout, cname):
    py = PAIRS[cname]
    assert C.sizeof(py) == c_layout[(cname, ".")], (cname, C.sizeof(py), c_layout[(cname, ".")])
    for f in py._fields_:
        assert getattr(py, f[0]).offset == c_layout[(cname, f[0])], (cname, f[0])


def test_enum_values_match_header(tmp_path):
    """Every enumerator of the header's enums has a constant of the same name
    and value in _capi.py (gcc prints the header's values)."""
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    header = open(os.path.join(ROOT, "include", "gds2para.h")).read()
    names = []
    for body in re.findall(r"^typedef enum(?: gp_[a-z0-9_]+)? \{(.*?)^\} gp_[a-z0-9_]+;", header, flags=re.M | re.S):
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        names += re.findall(r"\b(GP_[A-Z0-9_]+)\b", body)
    assert names
    lines = ['#include <stdio.h>', '#include "gds2para.h"', "int main(void) {"]
    lines += [f'  printf("{nm} %d\\n", (int){nm});' for nm in names]
    lines += ["  return 0;", "}"]
    src, exe = tmp_path / "enums.c", tmp_path / "enums"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run([cc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        nm, value = line.split()
        assert hasattr(capi, nm), f"{nm} is not mirrored in _capi.py"
        assert getattr(capi, nm) == int(value), (nm, getattr(capi, nm), value)
