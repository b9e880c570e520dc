"""Build the reference time-loop driver (tests/cpp/reference_loop.cpp)
against the reference's own simulation sources, twice: with the GPU drop-in
VelocityCalc and with the oracle's (test infrastructure only).

The reference sources are compiled in place from /root/reference through a
view (a tree of links) in which only src/velocity_calculation/
velocity_calc.hpp is replaced; velocity_calc.cpp, which defines the replaced
class, and main.cpp are left out.  The flags are the reference's
(Makefile:9-21).  Outputs go to oracle/_ref/loop/ only.

    python oracle/build_loop.py        (needs /root/reference; a no-op without it)
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = "/root/reference"
OUT = os.path.join(HERE, "_ref", "loop")
FLAGS = ["-std=c++17", "-O2", "-fopenmp", "-w"]
SKIP = {"main.cpp", "src/velocity_calculation/velocity_calc.cpp"}
# the translation units that include velocity_calc.hpp, compiled per variant
PER_VARIANT = {"src/penalization/penalization_calc.cpp"}


def sources():
    """The reference Makefile's SRCS list (Makefile:12-68), read from it."""
    out, take = [], False
    for line in open(os.path.join(REF, "Makefile")):
        s = line.strip()
        if s.startswith("SRCS"):
            take = True
            s = s.split("=", 1)[1].strip()
        if not take:
            continue
        if s.startswith("#") or not s:
            if s.startswith("#"):
                continue
            break
        name = s.rstrip("\\").strip()
        if name.endswith(".cpp"):
            out.append(name)
        if not s.endswith("\\"):
            break
    return [s for s in out if s not in SKIP]


def view(dst, header):
    if os.path.isdir(dst):
        shutil.rmtree(dst)
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
    shutil.copyfile(header, target)


def compile_one(view_dir, rel, obj, extra):
    if os.path.exists(obj) and os.path.getmtime(obj) > os.path.getmtime(
            os.path.realpath(os.path.join(view_dir, rel))) and not extra:
        return obj
    subprocess.run(["g++", *FLAGS, *extra, "-I", view_dir, "-c", os.path.join(view_dir, rel),
                    "-o", obj], check=True)
    return obj


def main():
    if not os.path.isfile(os.path.join(REF, "main.cpp")):
        print("build_loop: the reference tree is not here; nothing built")
        return 0
    os.makedirs(os.path.join(OUT, "obj"), exist_ok=True)
    variants = {
        "gpu": (os.path.join(ROOT, "include", "vpmfmm_reference_velocity_calc.hpp"),
                ["-I", os.path.join(ROOT, "include")],
                [os.path.join(ROOT, "lsmps_vpm_v2_amd", "libvpmfmm.so"),
                 "-Wl,-rpath,$ORIGIN/../../../lsmps_vpm_v2_amd"]),
        "oracle": (os.path.join(HERE, "oracle_velocity_calc.hpp"), ["-I", HERE],
                   [os.path.join(HERE, "liboracle_fmm2d.so"), "-Wl,-rpath,$ORIGIN/../.."]),
    }
    common = [s for s in sources() if s not in PER_VARIANT]
    base_view = os.path.join(OUT, "view_oracle")
    for name, (header, inc, _) in variants.items():
        view(os.path.join(OUT, f"view_{name}"), header)
    jobs = int(os.environ.get("MAX_JOBS", "8"))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda rel: compile_one(
            base_view, rel, os.path.join(OUT, "obj", rel.replace("/", "_") + ".o"), []),
            common))
    for name, (header, inc, link) in variants.items():
        vdir = os.path.join(OUT, f"view_{name}")
        own = [compile_one(vdir, rel, os.path.join(OUT, "obj", f"{name}_" +
                                                   rel.replace("/", "_") + ".o"), inc)
               for rel in sorted(PER_VARIANT)]
        drv = os.path.join(OUT, "obj", f"{name}_reference_loop.o")
        subprocess.run(["g++", *FLAGS, *inc, "-I", vdir, "-c",
                        os.path.join(ROOT, "tests", "cpp", "reference_loop.cpp"), "-o", drv],
                       check=True)
        exe = os.path.join(OUT, f"reference_loop_{name}")
        subprocess.run(["g++", "-fopenmp", "-o", exe, drv, *own, *objs, *link], check=True)
        print("built", exe)
    return 0


if __name__ == "__main__":
    sys.exit(main())
