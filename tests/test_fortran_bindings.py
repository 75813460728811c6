"""Static checks of the Fortran iso_c_binding layer against the C header:
every BIND(C, name=...) names a function declared in include/self_amd.h
with the same number of arguments, and the flag parameters carry the
header's values."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "self_amd.h")
F90 = os.path.join(ROOT, "self_amd", "fortran", "self_amd_cabi.f90")


def _c_decls():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:int|const char \*)\s*\*?\s*(self_\w+)\s*\(([^)]*)\)\s*;", txt):
        args = [a for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        out[m.group(1)] = len(args)
    return out


def _f_interfaces():
    txt = open(F90).read()
    out = {}
    for m in re.finditer(r"FUNCTION\s+(\w+)\s*\(([^)]*)\)\s*(?:&\s*)?BIND\(C,\s*name=\"(\w+)\"\)", txt, flags=re.S):
        args = [a for a in m.group(2).replace("&", " ").split(",") if a.strip()]
        out[m.group(3)] = len(args)
    return out


def test_every_fortran_binding_matches_a_c_declaration():
    c, f = _c_decls(), _f_interfaces()
    f.pop("strlen", None)
    assert len(f) >= 30
    for name, n in f.items():
        assert name in c, f"{name} is not declared in self_amd.h"
        assert c[name] == n, f"{name}: {n} Fortran arguments, {c[name]} in C"


def test_flag_values_match_the_header():
    hdr = open(HDR).read()
    f90 = open(F90).read()
    for name, val in re.findall(r"#define\s+(SELF_[A-Z0-9_]+)\s+\(?(-?\d+)\)?", hdr):
        m = re.search(rf"\b{name}\s*=\s*(-?\d+)", f90)
        if m:
            assert int(m.group(1)) == int(val), name
