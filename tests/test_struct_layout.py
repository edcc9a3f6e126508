"""The ctypes mirrors of the option/stat structs match include/jfem.h field
for field (names and order), so a field added on one side only fails here on
the CPU instead of as a garbled call on the GPU."""
import os
import re

import pytest

from jfem import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "jfem.h")


def header_fields(name):
    text = open(HEADER).read()
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + ";", text)
    assert m, name
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    out = []
    for decl in body.split(";"):
        words = decl.replace(",", " , ").split()
        if not words:
            continue
        # "type a, b" -> a, b
        names = " ".join(words[1:]).split(",")
        out += [n.strip() for n in names if n.strip()]
    return out


@pytest.mark.parametrize("cname,pyname", [("jfm_solve_opts", "SolveOpts"), ("jfm_solve_stats", "SolveStats"),
                                          ("jfm_newton_opts", "NewtonOpts"), ("jfm_newton_stats", "NewtonStats"),
                                          ("jfm_info", "Info")])
def test_struct_mirrors_header(cname, pyname):
    py = [f for f, _ in getattr(_lib, pyname)._fields_]
    assert py == header_fields(cname)
