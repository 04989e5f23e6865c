"""The C ABI as a C or C++ caller sees it: the header compiles on its own,
the ctypes mirror in rodv_amd/engine.py has the same struct layouts, the
library exports pairhmm::schedule::schedule_pairhmm with the reference's
signature (a caller declaring it as the reference's header does links and
runs), and nothing but its API."""
import ctypes as C
import os
import subprocess
import sys

import pytest

from rodv_amd import engine as E

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin"))
import driver as D  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")

SIZES_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "rodv_phmm.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(rodv_phmm_config), sizeof(rodv_phmm_batch),
           sizeof(rodv_phmm_stats), sizeof(rodv_phmm_tuning), sizeof(rodv_phmm_tune_entry),
           sizeof(rodv_phmm_params), offsetof(rodv_phmm_config, use_tuning),
           offsetof(rodv_phmm_config, haplotype_chains), offsetof(rodv_phmm_tuning, cls));
    return 0;
}
"""

SHIM_CPP = r"""
#include <vector>
#include <cstdint>
#include <cstdio>
#include "rodv_phmm_schedule.hpp"
// the reference's declaration (pairhmm_schedule.h:14-25), restated: the
// library's function has exactly its type
namespace pairhmm { namespace schedule {
using Fn = bool (*)(const std::vector<std::vector<uint8_t>> &, const std::vector<std::vector<uint8_t>> &,
                    std::vector<std::vector<double>> &, const std::vector<std::vector<uint8_t>> &,
                    const std::vector<std::vector<uint8_t>> &, const std::vector<std::vector<uint8_t>> &,
                    const std::vector<std::vector<uint8_t>> &, bool, double, double, bool);
static Fn f = &schedule_pairhmm;
}}
int main() {
    // no haplotypes: false before any device is touched (pairhmm_schedule.cpp:130-132)
    std::vector<std::vector<uint8_t>> none, reads{{'A'}}, q{{30}};
    std::vector<std::vector<double>> res;
    const bool r = pairhmm::schedule::f(none, reads, res, q, q, q, q, false, 0.1, 0.1, false);
    std::printf("%d\n", (int)r);
    return 0;
}
"""


def _compile(tmp_path, src, name, cxx=False):
    p = tmp_path / name
    p.write_text(src)
    exe = tmp_path / (name + ".out")
    cmd = (["g++", "-std=c++17"] if cxx else ["gcc", "-std=c11"]) + ["-Wall", "-Werror", f"-I{INC}",
                                                                       str(p), "-o", str(exe)]
    if cxx:  # linked against the engine, as a reference caller would be
        cmd += [f"-L{D.LIBDIR}", "-lrodv_phmm", f"-Wl,-rpath,{D.LIBDIR}",
                "-Wl,-rpath-link,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_struct_layouts_match_ctypes(tmp_path):
    exe = _compile(tmp_path, SIZES_C, "sizes.c")
    vals = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                           text=True).stdout.split()]
    assert vals == [C.sizeof(E.Config), C.sizeof(E.CBatch), C.sizeof(E.Stats),
                    C.sizeof(E.Tuning), C.sizeof(E.TuneEntry), C.sizeof(E.Params),
                    E.Config.use_tuning.offset, E.Config.haplotype_chains.offset,
                    E.Tuning.cls.offset]


def test_schedule_pairhmm_links_with_reference_signature(tmp_path):
    exe = _compile(tmp_path, SHIM_CPP, "shim.cpp", cxx=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout
    assert out.strip() == "0"


def _dynsyms(path, *flags):
    out = subprocess.run(["nm", "-D", *flags, path], check=True, capture_output=True, text=True).stdout
    return [l.split()[-1] for l in out.splitlines() if l.strip()]


def test_reference_caller_links_unchanged(tmp_path):
    # the driver declares schedule_pairhmm exactly as the reference's header
    # does; the symbol g++ asks for is the one the library defines
    exe = D.build(str(tmp_path / "schedule_driver"))
    want = [n for n in _dynsyms(exe, "--undefined-only") if "schedule_pairhmm" in n]
    assert len(want) == 1
    assert want[0] in _dynsyms(E.LIB_PATH, "--defined-only")


def test_library_exports_only_its_api():
    # C names: the C ABI (rodv_phmm_*) and the HIP toolchain's compilation-
    # unit ids; C++ names: schedule_pairhmm alone (the version script keeps
    # standard-library template instances local)
    names = _dynsyms(E.LIB_PATH, "--defined-only")
    plain = [n for n in names if not n.startswith("_Z")]
    assert plain and all(n.startswith(("rodv_phmm_", "__hip_cuid_")) for n in plain), plain
    mangled = [n for n in names if n.startswith("_Z")]
    assert len(mangled) == 1 and mangled[0].startswith("_ZN7pairhmm8schedule16schedule_pairhmm"), mangled
