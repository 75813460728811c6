"""Static checks of the Fortran iso_c_binding layer against the C header:
every BIND(C, name=...) names a function declared in include/self_amd.h
with the same number of arguments, and the flag parameters carry the
header's values."""
import os
import re
import subprocess

import pytest

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


REF_DRIVERS = ["conservative-motionless", "conservative-linearwaverefraction", "dipole"]


@pytest.mark.parametrize("example", REF_DRIVERS)
def test_reference_driver_compiles_unchanged_against_the_drop_in(tmp_path, example):
    """The reference's own shallow-water driver programs (examples/
    shallowwater/<example>/localmods/ShallowWater_Driver.f90, read in place)
    compile and link, unmodified, against the drop-in modules of
    self_amd/fortran/sw_dropin.f90 and libself_amd.so."""
    src = os.path.join("/root/reference/examples/shallowwater", example, "localmods", "ShallowWater_Driver.f90")
    build = os.path.join(ROOT, "self_amd", "fortran", "_build")
    fc = "/opt/rocm/bin/amdflang"
    if not (os.path.exists(src) and os.path.exists(os.path.join(build, "sw_dropin.o")) and os.path.exists(fc)):
        pytest.skip("reference tree, amdflang or the built drop-in modules absent")
    obj = [os.path.join(build, f) for f in ("self_amd_cabi.o", "shallowwater_amd.o", "sw_dropin.o")]
    exe = str(tmp_path / "driver")
    r = subprocess.run([fc, "-O2", "-I", build, src] + obj + ["-L" + os.path.join(ROOT, "self_amd"), "-lself_amd",
                                                             "-o", exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.path.exists(exe)


def test_reference_cgsem_driver_compiles_unchanged_against_the_drop_in(tmp_path):
    """The reference's CGSEM example driver (examples/cgsem-elliptic2d/
    paraboloid/localmods/CGsemElliptic2D_Driver.f90, read in place: Build,
    Solve with a Dirichlet function, ConstructMatrix, the preconditioner's
    ConstructPCMatrix, WriteTecplot, WriteResidual, Trash, and the mesh's
    GetPositionAtNode) compiles and links, unmodified, against the drop-in
    modules of self_amd/fortran/cgsem_dropin.f90 and libself_amd.so."""
    src = "/root/reference/examples/cgsem-elliptic2d/paraboloid/localmods/CGsemElliptic2D_Driver.f90"
    build = os.path.join(ROOT, "self_amd", "fortran", "_build")
    fc = "/opt/rocm/bin/amdflang"
    if not (os.path.exists(src) and os.path.exists(os.path.join(build, "cgsem_dropin.o")) and os.path.exists(fc)):
        pytest.skip("reference tree, amdflang or the built drop-in modules absent")
    obj = [os.path.join(build, f) for f in ("self_amd_cabi.o", "shallowwater_amd.o", "sw_dropin.o",
                                            "cgsem_dropin.o")]
    exe = str(tmp_path / "driver")
    r = subprocess.run([fc, "-O2", "-I", build, "-J", str(tmp_path), src] + obj +
                       ["-L" + os.path.join(ROOT, "self_amd"), "-lself_amd", "-o", exe],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.path.exists(exe)


def test_reference_adv2d_driver_compiles_unchanged_against_the_drop_in(tmp_path):
    """The reference's 2-D advection-diffusion driver (src/highend/
    advection-diffusion2d/Advection_Driver.f90, read in place: Build with the
    pickup at iterInit, GlobalTimeDerivative, WriteTecplot, ForwardStepRK3,
    WritePickup, mesh % WriteTecplot, Trash) compiles and links, unmodified,
    against the drop-in modules of self_amd/fortran/adv_dropin.f90 and
    libself_amd.so."""
    src = "/root/reference/src/highend/advection-diffusion2d/Advection_Driver.f90"
    build = os.path.join(ROOT, "self_amd", "fortran", "_build")
    fc = "/opt/rocm/bin/amdflang"
    if not (os.path.exists(src) and os.path.exists(os.path.join(build, "adv_dropin.o")) and os.path.exists(fc)):
        pytest.skip("reference tree, amdflang or the built drop-in modules absent")
    obj = [os.path.join(build, f) for f in ("self_amd_cabi.o", "shallowwater_amd.o", "sw_dropin.o", "adv_dropin.o")]
    exe = str(tmp_path / "driver")
    r = subprocess.run([fc, "-O2", "-I", build, "-J", str(tmp_path), src] + obj +
                       ["-L" + os.path.join(ROOT, "self_amd"), "-lself_amd", "-o", exe],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.path.exists(exe)


ADV3D_DIRS = ["src/highend/advection3d", "examples/advection3d/spot-advection/localmods"]


@pytest.mark.parametrize("where", ADV3D_DIRS)
@pytest.mark.parametrize("program", ["Advection3D_Driver", "Advection3D_InitialConditions"])
def test_reference_adv3d_programs_compile_unchanged_against_the_drop_in(tmp_path, where, program):
    """The reference's 3-D advection programs (Advection3D_Driver.f90 and
    Advection3D_InitialConditions.f90 of src/highend/advection3d and of the
    spot-advection example, read in
    place: Build with and without the pickup, the clocks of Timing,
    dgStorage % GetNumberOfNodes, mesh % GetPositionAtNode, SetSolution,
    SetVelocity, SetRelaxationField/Factor, GlobalTimeDerivative(t, m),
    WriteTecplot, ForwardStepRK3, WritePickup, mesh % WriteTecplot, Trash)
    compile and link, unmodified, against the drop-in modules of
    self_amd/fortran/adv3d_dropin.f90 (module directory _build3d, shared
    modules from _build) and libself_amd.so."""
    src = f"/root/reference/{where}/{program}.f90"
    build = os.path.join(ROOT, "self_amd", "fortran", "_build")
    build3 = os.path.join(ROOT, "self_amd", "fortran", "_build3d")
    fc = "/opt/rocm/bin/amdflang"
    if not (os.path.exists(src) and os.path.exists(os.path.join(build3, "adv3d_dropin.o")) and os.path.exists(fc)):
        pytest.skip("reference tree, amdflang or the built drop-in modules absent")
    obj = [os.path.join(build, f) for f in ("self_amd_cabi.o", "shallowwater_amd.o", "sw_dropin.o")]
    obj.append(os.path.join(build3, "adv3d_dropin.o"))
    exe = str(tmp_path / "prog")
    r = subprocess.run([fc, "-O2", "-I", build3, "-I", build, "-J", str(tmp_path), src] + obj +
                       ["-L" + os.path.join(ROOT, "self_amd"), "-lself_amd", "-o", exe],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.path.exists(exe)


@pytest.mark.parametrize("where", ["examples/elliptic1d/parabola/localmods", "src/cgsem"])
def test_reference_elliptic1d_driver_compiles_unchanged_against_the_drop_in(tmp_path, where):
    """The reference's 1-D elliptic driver (Elliptic1D_Driver.f90 of the
    parabola example and of src/cgsem, read in place: Build, Solve(ioerr),
    WriteTecplot, WriteResidual, Trash of TYPE Elliptic1D) compiles and
    links, unmodified, against the drop-in modules of
    self_amd/fortran/elliptic1d_dropin.f90 and libself_amd.so."""
    src = f"/root/reference/{where}/Elliptic1D_Driver.f90"
    build = os.path.join(ROOT, "self_amd", "fortran", "_build")
    fc = "/opt/rocm/bin/amdflang"
    if not (os.path.exists(src) and os.path.exists(os.path.join(build, "elliptic1d_dropin.o")) and os.path.exists(fc)):
        pytest.skip("reference tree, amdflang or the built drop-in modules absent")
    obj = [os.path.join(build, f) for f in ("self_amd_cabi.o", "shallowwater_amd.o", "sw_dropin.o",
                                            "elliptic1d_dropin.o")]
    exe = str(tmp_path / "driver")
    r = subprocess.run([fc, "-O2", "-I", build, "-J", str(tmp_path), src] + obj +
                       ["-L" + os.path.join(ROOT, "self_amd"), "-lself_amd", "-o", exe],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.path.exists(exe)
