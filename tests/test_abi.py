"""CPU checks of the C-ABI library: it loads, exports every function that
include/vpmfmm.h declares, and its host-only entry points behave (no GPU
compute is called here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vpmfmm.h")


def _ensure_built():
    from lsmps_vpm_v2_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from lsmps_vpm_v2_amd import build
        build.build()
    return _lib


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vpmfmm_[a-z0-9_]+)\s*\(", txt)))


def test_header_declarations_are_exported():
    _lib = _ensure_built()
    names = declared_functions()
    assert len(names) >= 19
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == names


def test_binding_argument_counts_match_header():
    """Every function the binding declares argtypes for takes as many
    arguments as its prototype in the header."""
    _lib = _ensure_built()
    L = _lib.load()
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    checked = 0
    for name in declared_functions():
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", txt)
        params = [a for a in m.group(1).split(",") if a.strip() and a.strip() != "void"]
        fn = getattr(L, name)
        if fn.argtypes is None:
            continue
        assert len(fn.argtypes) == len(params), name
        checked += 1
    assert checked >= 30


def test_exported_symbols_are_c_linkage():
    _lib = _ensure_built()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for n in declared_functions():
        assert n in syms, n


def test_default_options_are_the_reference_values():
    _lib = _ensure_built()
    o = _lib.default_opts()
    # reference global.cpp:303-313 and :164-165
    assert (o.p, o.max_par, o.max_src, o.exp_tree) == (10, 75, 50, 1.0)
    assert (o.u_inf, o.v_inf) == (1.0, 0.0)
    assert (o.min_level, o.max_level, o.binom_mode, o.list4_mode) == (1, 0, 1, 1)
    assert o.precision == _lib.FP64 and o.stage_mask == 0
    assert o.order_select == _lib.ORDER_FIXED and o.reselect_every == 0


def test_struct_layouts_match_header(tmp_path):
    """sizeof/offsetof from the C header (compiled with gcc) equal the ctypes view."""
    _lib = _ensure_built()
    probe = tmp_path / "probe.c"
    structs = [("vpmfmm_opts", _lib.Opts), ("vpmfmm_stats", _lib.Stats),
               ("vpmfmm_selector", _lib.Selector), ("vpmfmm_selection", _lib.Selection),
               ("vpmfmm_type1_opts", _lib.Type1Opts), ("vpmfmm_fmm3d_opts", _lib.Fmm3dOpts),
               ("vpmfmm_fmm3d_stats", _lib.Fmm3dStats)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vpmfmm.h"',
             "int main(void) {"]
    for cname, _ in structs:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
    for cname, ct in structs:
        lines += [f'printf("%zu\\n", offsetof({cname}, {f}));' for f, _ in ct._fields_]
    lines += ["return 0; }"]
    probe.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(probe), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    vals = [int(v) for v in out]
    assert vals[:len(structs)] == [C.sizeof(ct) for _, ct in structs]
    offs = vals[len(structs):]
    want = [getattr(ct, f).offset for _, ct in structs for f, _ in ct._fields_]
    assert offs == want


def test_null_handle_is_rejected():
    _lib = _ensure_built()
    L = _lib.load()
    assert L.vpmfmm_destroy(None) == 0
    assert L.vpmfmm_field(None, 10, None, None, None, None, None, None) == 1
    assert L.vpmfmm_last_stats(None, None) == 1
    assert L.vpmfmm_set_selector(None, None, None) == 1
    assert L.vpmfmm_mlp_forward(None, None, None, 0, None) == 1


def test_invalid_options_are_rejected_before_any_device_call():
    _lib = _ensure_built()
    L = _lib.load()
    for field, value in (("p", 0), ("p", 41), ("max_par", 0), ("max_par", 5000),
                         ("exp_tree", 0.0), ("min_level", 0), ("max_level", 17),
                         ("binom_mode", 2), ("list4_mode", 3), ("stage_mask", 8),
                         ("order_select", 2), ("reselect_every", -1), ("judge_every", -2)):
        o = _lib.default_opts()
        setattr(o, field, value)
        h = C.c_void_p()
        assert L.vpmfmm_create(C.byref(o), C.byref(h)) == 1, field
        assert not h.value


def test_oracle_library_exports(oracle):
    L = oracle.lib()
    for n in ("orc_tree_build", "orc_fmm_field", "orc_velocity", "orc_direct_field",
              "orc_fmm_expansions", "orc_binom_ref"):
        assert hasattr(L, n)


def test_product_never_imports_the_oracle():
    """The product package and its sources must not load the checker."""
    pkg = os.path.join(ROOT, "lsmps_vpm_v2_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                src = open(os.path.join(dirpath, fn)).read()
                assert "import oracle" not in src and "from oracle" not in src, fn
                assert "fmm2d_oracle" not in src and "libref_fmm2d" not in src, fn


def build_drop_in(tmp_path, name="drop_in_velocity"):
    """Compile tests/cpp/<name>.cpp against the C++ drop-in header."""
    _lib = _ensure_built()
    exe = tmp_path / name
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I",
                    os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", name + ".cpp"), f"-L{libdir}",
                    "-lvpmfmm", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    return exe


def test_cpp_drop_in_header_compiles(tmp_path):
    """The reference's call site get_velocity(particle, grid, step) compiles
    against include/vpmfmm_velocity_calc.hpp with the reference's member types,
    with DIM 2 and with DIM 3."""
    exe = build_drop_in(tmp_path)
    assert exe.exists()
    exe3 = build_drop_in(tmp_path, "drop_in_velocity3d")
    assert exe3.exists()
