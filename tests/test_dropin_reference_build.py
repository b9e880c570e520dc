"""The drop-in recipe against the reference program's own sources (compile
only; needs the reference tree, skipped where it is absent).

A view of the reference tree is made in a temporary folder: symbolic links
to its files, except src/velocity_calculation/velocity_calc.hpp, whose place
takes include/vpmfmm_reference_velocity_calc.hpp (the recipe's one change
to the reference's sources).  Both translation units
that include that header -- main.cpp (main.cpp:44, 72, 410) and
src/penalization/penalization_calc.cpp (:3, 363, 402) -- are compiled with
the reference's flags against the reference's own struct Particle
(Utils.hpp:77-159).  Their objects must then hold one VelocityCalc: the
drop-in's constructor and get_velocity, and no reference to the replaced
class's out-of-line members (velocity_calc.cpp), which leaves the build."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"

pytestmark = pytest.mark.skipif(not os.path.isfile(os.path.join(REF, "main.cpp")),
                                reason="the reference tree is not in this container")


def _view(dst):
    """Top-level entries linked; src/ rebuilt as real folders with linked
    files (a linked folder would resolve "../" inside the reference tree)."""
    os.makedirs(dst)
    for top in os.listdir(REF):
        if top != "src":
            os.symlink(os.path.join(REF, top), os.path.join(dst, top))
    src = os.path.join(REF, "src")
    for dirpath, _, files in os.walk(src):
        out = os.path.join(dst, "src", os.path.relpath(dirpath, src))
        os.makedirs(out, exist_ok=True)
        for fn in files:
            os.symlink(os.path.join(dirpath, fn), os.path.join(out, fn))
    target = os.path.join(dst, "src", "velocity_calculation", "velocity_calc.hpp")
    os.remove(target)
    shutil.copyfile(os.path.join(ROOT, "include", "vpmfmm_reference_velocity_calc.hpp"), target)


def _compile(view, rel, obj):
    # the reference's flags (Makefile:9-21) at -O0 with inline functions kept:
    # main.cpp exits before its time loop (main.cpp:276), and without the flag
    # the unreachable call's instantiation is not emitted
    subprocess.run(["g++", "-std=c++17", "-O0", "-fopenmp", "-w", "-fkeep-inline-functions",
                    "-I", os.path.join(ROOT, "include"), "-c", os.path.join(view, rel),
                    "-o", obj], check=True, timeout=300)
    return subprocess.run(["nm", "-C", obj], capture_output=True, text=True,
                          check=True).stdout.splitlines()


def test_one_velocitycalc_across_the_reference_program(tmp_path):
    view = str(tmp_path / "ref_view")
    _view(view)
    syms = {rel: _compile(view, rel, str(tmp_path / (os.path.basename(rel) + ".o")))
            for rel in ("main.cpp", "src/penalization/penalization_calc.cpp")}
    ctors = set()
    for rel, lines in syms.items():
        vc = [ln for ln in lines if "VelocityCalc" in ln]
        # no symbol of the replaced class (global VelocityCalc) at all
        bad = [ln for ln in vc if re.search(r"(?<!vpmfmm::)\bVelocityCalc::", ln)]
        assert not bad, (rel, bad)
        ctors |= {ln.split(None, 2)[-1] for ln in vc
                  if re.search(r"\bW\b", ln) and "VelocityCalc::VelocityCalc()" in ln}
        # the drop-in's member template instantiated for the reference's Particle
        assert any(re.search(r"\bW\b.*vpmfmm::VelocityCalc::get_velocity<Particle>"
                             r"\(Particle&, int\)", ln) for ln in vc), rel
    assert ctors == {"vpmfmm::VelocityCalc::VelocityCalc()"}, ctors


def test_reference_header_is_replaced_whole():
    """The forwarding header keeps the replaced header's include guard and the
    headers it brought in, so nothing that relied on them breaks."""
    ref = open(os.path.join(REF, "src", "velocity_calculation", "velocity_calc.hpp")).read()
    ours = open(os.path.join(ROOT, "include", "vpmfmm_reference_velocity_calc.hpp")).read()
    guard = re.search(r"#ifndef\s+(\w+)", ref).group(1)
    assert f"#ifndef {guard}" in ours and f"#define {guard}" in ours
    for inc in re.findall(r'#include\s+"([^"]+)"', ref):
        assert f'#include "{inc}"' in ours, inc
