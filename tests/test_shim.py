"""The drop-in shim (dustyfargoadsg_amd/shim/fargo_shim.c) as a link unit.

CPU checks of what the reference's host would link against:
  * the shim defines exactly the 26 symbols the host objects import from the
    removed hot-path objects (SourceEuler.o, TransportEuler.o, Viscosity.o,
    Pframeforce.o, sg*.o, Dsys.o, DustUpdate.o);
  * everything it imports is libdfargo's C ABI, the C library, or a host
    symbol declared in shim/fargo_ref.h;
  * the small C host (shim/fargo_host.c) links shim + libdfargo with no
    undefined symbol;
and, where the reference sources are present (this container, not the GPU
box), that every host symbol the shim imports is defined by the reference's
host sources (src/var.c parameters, src/global.h, src/Interpret.c,
src/main.c, the kept .c files) and not only by a removed object, and that
the structures of fargo_ref.h have src/types.h's fields in its order.  The
reference's own build (its generated headers) is not run: the reference is
read as text only.
"""
import os
import re
import subprocess

import pytest

from dustyfargoadsg_amd import build, _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "dustyfargoadsg_amd", "shim")
REF = "/root/reference/src"

BOUNDARY = {
    "AlgoGas", "InitEuler", "FillPolar1DArrays", "ComputeSoundSpeed", "ComputePressureField",
    "ComputeTemperatureField", "ComputeViscousTerms", "ComputeViscousHeating", "ComputeEntropyDiffusion",
    "ComputeRadiativeDiffusion", "CircumPlanetaryMass", "FViscosity", "AspectRatio", "DAspectRatio",
    "InitGasDensity", "InitGasEnergy", "InitDustDensity", "InitDustSystem", "FreeDust", "SolveOrbits",
    "compute_selfgravity", "init_planetarysys_withSG", "interpolation",
    "DiskOnPrimaryAcceleration", "DiskOnPrimaryAcceleration_alt", "LostMass",
}
REMOVED = ["SourceEuler.c", "TransportEuler.c", "Viscosity.c", "Pframeforce.c", "sgmain.c", "sginit.c",
           "sgdens.c", "sgkernel.c", "sgacc.c", "sgupdate.c", "sgsysinit.c", "Dsys.c", "DustUpdate.c",
           "sgzero.c"]
LIBC = {"calloc", "free", "exp", "log", "pow", "sqrt", "sincos", "cos", "sin", "fopen", "fclose", "getenv",
        "strstr", "time", "memcpy", "memset", "fprintf", "snprintf", "__fprintf_chk", "__snprintf_chk",
        "__stack_chk_fail"}


def _symbols(obj, undefined):
    out = subprocess.run(["nm", "-u" if undefined else "--defined-only", obj], capture_output=True, text=True,
                         check=True).stdout
    names = set()
    for line in out.splitlines():
        parts = line.split()
        if undefined:
            names.add(parts[-1])
        elif len(parts) == 3 and parts[1] in "TDBC":
            names.add(parts[2])
    return names


@pytest.fixture(scope="module")
def shim_obj():
    if not os.path.exists(build.OUT):
        pytest.skip("libdfargo.so is not built")
    build.build_shim()
    return os.path.join(build.BUILD, "fargo_shim.c.o")


def _ref_host_names():
    """The extern names of shim/fargo_ref.h (parameters, switches, tables,
    grids) and its helper functions."""
    text = open(os.path.join(SHIM, "fargo_ref.h")).read()
    names = set()
    for m in re.finditer(r"^extern\s+(?:real|int|boolean|char|PolarGrid)\s+([^;]+);", text, re.M):
        for item in m.group(1).split(","):
            names.add(re.sub(r"[\s*]|\[.*\]", "", item.split("[")[0]))
    for m in re.finditer(r"^\w[\w\s*]*?\b(\w+)\([^)]*\);", text, re.M):
        names.add(m.group(1))
    return names


def test_shim_defines_the_26_boundary_symbols(shim_obj):
    assert _symbols(shim_obj, False) == BOUNDARY


def test_shim_imports_only_the_abi_libc_and_host(shim_obj):
    lib_syms = set(subprocess.run(["nm", "-D", "--defined-only", build.OUT], capture_output=True, text=True,
                                  check=True).stdout.split())
    host = _ref_host_names() | {"FillSigma", "FillEnergy", "FillDSigma"}
    for name in _symbols(shim_obj, True):
        if name.startswith("dfargo_"):
            assert name in lib_syms, name
        else:
            assert name in LIBC or name in host, name


def test_host_links_with_no_undefined_symbol(shim_obj):
    # build_shim links with -Wl,--no-undefined; the binary is there
    assert os.access(build.SHIM_HOST, os.X_OK)


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference sources are not present")
def test_host_symbols_are_defined_by_the_reference_host(shim_obj):
    """Each reference symbol the shim imports has a definition the kept host
    objects provide: a src/var.c parameter (var("KEY", &NAME / NAME, ...)),
    a definition in src/global.h, src/Interpret.c or src/main.c, or a
    function defined in a kept .c file."""
    var_c = open(os.path.join(REF, "var.c")).read()
    params = set(re.findall(r'var\("\w+",\s*&?(\w+),', var_c))
    defs = ""
    for f in ("global.h", "Interpret.c", "main.c"):
        defs += open(os.path.join(REF, f), errors="replace").read() + "\n"
    kept_funcs = set()
    for f in os.listdir(REF):
        if f.endswith(".c") and f not in REMOVED:
            kept_funcs |= set(re.findall(r"^\w[\w\s*]*?\b(\w+)\s*\(", open(os.path.join(REF, f),
                                                                         errors="replace").read(), re.M))
    missing = []
    for name in sorted(_symbols(shim_obj, True)):
        if name.startswith("dfargo_") or name in LIBC:
            continue
        ok = (name in params or name in kept_funcs or
              re.search(r"^(?:boolean|int|real|PolarGrid|long)\b[^;(]*\b" + name + r"\b", defs, re.M))
        if not ok:
            missing.append(name)
    assert not missing, missing


def _fields(text, struct):
    m = re.search(r"struct\s+" + struct + r"\s*\{(.*?)\}", text, re.S)
    assert m, struct
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        words = decl.replace("*", " * ").split()
        base = words[0]
        rest = " ".join(words[1:])
        for item in rest.split(","):
            stars = item.count("*")
            fields.append((base, stars, item.replace("*", "").strip()))
    return fields


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference sources are not present")
@pytest.mark.parametrize("struct", ["pair", "force", "polargrid", "planetary_system", "DustParticlesSystem"])
def test_struct_layouts_match_types_h(struct):
    ref = _fields(open(os.path.join(REF, "types.h")).read(), struct)
    ours = _fields(open(os.path.join(SHIM, "fargo_ref.h")).read(), struct)
    assert ours == ref
