"""The C ABI as a C or C++ caller sees it: the header compiles on its own,
the ctypes mirror in rodv_amd/engine.py has the same struct layouts, and the
header-only schedule_pairhmm shim compiles with the reference's signature."""
import ctypes as C
import os
import subprocess

import pytest

from rodv_amd import engine as E

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
#include "rodv_phmm_schedule.hpp"
// the reference's declaration (pairhmm_schedule.h:14-25), restated: a caller
// compiled against it binds to the shim
namespace pairhmm { namespace schedule {
using Fn = bool (*)(const std::vector<std::vector<uint8_t>> &, const std::vector<std::vector<uint8_t>> &,
                    std::vector<std::vector<double>> &, const std::vector<std::vector<uint8_t>> &,
                    const std::vector<std::vector<uint8_t>> &, const std::vector<std::vector<uint8_t>> &,
                    const std::vector<std::vector<uint8_t>> &, bool, double, double, bool);
static Fn f = &schedule_pairhmm;
}}
int main() { return pairhmm::schedule::f == nullptr; }
"""


def _compile(tmp_path, src, name, cxx=False):
    p = tmp_path / name
    p.write_text(src)
    exe = tmp_path / (name + ".out")
    cmd = (["g++", "-std=c++17"] if cxx else ["gcc", "-std=c11"]) + ["-Wall", "-Werror", f"-I{INC}",
                                                                       str(p), "-o", str(exe)]
    if cxx:
        cmd = ["g++", "-std=c++17", "-Wall", "-Werror", f"-I{INC}", "-fsyntax-only", str(p)]
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


def test_schedule_shim_compiles_with_reference_signature(tmp_path):
    _compile(tmp_path, SHIM_CPP, "shim.cpp", cxx=True)
