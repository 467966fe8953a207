"""The C ABI boundary without a GPU: the built library loads, exports every
function include/centrex_lindblad.h declares, and the ctypes mirrors of the
argument structs list the header's fields in the header's order with the
same sizes (a field added to one side only would shift every later field)."""

from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

from centrex_tlf_amd.lindblad import _abi
from centrex_tlf_amd.lindblad.native import library_path

HEADER = Path(__file__).resolve().parents[1] / "include" / "centrex_lindblad.h"


def _strip_comments(text: str) -> str:
    return re.sub(r"/\*.*?\*/", "", re.sub(r"//[^\n]*", "", text, flags=re.S), flags=re.S)


def _functions(text: str) -> list[str]:
    body = _strip_comments(text)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(lb_\w+)\s*\(", body, flags=re.M)))


def _struct_fields(text: str, name: str) -> list[tuple[str, str, int]]:
    """(type, name, count) per field; handles `T a, b;` and `T a[3];`."""
    body = _strip_comments(text)
    m = re.search(r"typedef struct " + name + r"\s*\{(.*?)\}\s*" + name + ";", body, flags=re.S)
    assert m, name
    fields = []
    for decl in m.group(1).split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        fm = re.match(r"(.*?)(\w+)\s*(?:\[(\d+)\])?((?:\s*,\s*\w+)*)$", decl)
        ctype = fm.group(1).strip()
        count = int(fm.group(3)) if fm.group(3) else 1
        fields.append((ctype, fm.group(2), count))
        for extra in filter(None, (x.strip() for x in fm.group(4).split(","))):
            fields.append((ctype, extra, 1))
    return fields


_CSIZE = {"int64_t": 8, "int32_t": 4, "double": 8, "uint8_t": 1}


def _size_of(ctype: str) -> int:
    return 8 if "*" in ctype else _CSIZE[ctype.replace("const ", "").strip()]


@pytest.fixture(scope="module")
def lib():
    path = library_path()
    if not path.exists():
        pytest.fail(f"{path} is not built (run __graft_entry__.build())")
    pytest.importorskip("torch")  # loads the HIP runtime the library links against
    return C.CDLL(str(path))


def test_header_declares_the_boundary():
    names = _functions(HEADER.read_text())
    for required in ("lb_plan_create", "lb_plan_destroy", "lb_solve", "lb_rhs_eval", "lb_adjoint",
                     "lb_solve_group", "lb_last_error"):
        assert required in names


def test_library_exports_every_declared_function(lib):
    missing = [n for n in _functions(HEADER.read_text()) if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("name,mirror", [("lb_solve_args", _abi.LbSolveArgs), ("lb_plan_desc", _abi.LbPlanDesc),
                                         ("lb_eff_plan_desc", _abi.LbEffPlanDesc),
                                         ("lb_eff_solve_args", _abi.LbEffSolveArgs)])
def test_struct_mirrors_match_header(name, mirror):
    fields = _struct_fields(HEADER.read_text(), name)
    assert [f for f, _ in mirror._fields_] == [n for _, n, _ in fields]
    for (ctype, fname, count), (mname, mtype) in zip(fields, mirror._fields_):
        assert C.sizeof(mtype) == _size_of(ctype) * count, (fname, ctype, mtype)
