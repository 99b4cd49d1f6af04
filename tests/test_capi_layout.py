"""The ctypes mirrors of the C ABI's structures (grim_amd/capi.py) against
include/grim_hip.h: a C program compiled with gcc prints sizeof and the
offset of every member the Python side names, and both must agree, so a
field added to one side only (or in another order) fails here instead of
shifting every later field the library reads."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from grim_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

PAIRS = [
    (capi.Params, "grim_params"),
    (capi.ZoneBatch, "grim_zone_batch"),
    (capi.NewtonOpts, "grim_newton_opts"),
    (capi.NewtonStats, "grim_newton_stats"),
    (capi.GridConfig, "grim_grid_config"),
    (capi.StepStats, "grim_step_stats"),
    (capi.StageArrays, "grim_stage_arrays"),
    (capi.TorusObservables, "grim_torus_observables"),
    (capi.TorusParams, "grim_torus_params"),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_ctypes_structures_match_the_header(tmp_path):
    lines = ['#include "grim_hip.h"', "#include <stddef.h>", "#include <stdio.h>", "int main(void) {"]
    expect = []
    for cls, cname in PAIRS:
        lines.append(f'  printf("%zu\\n", sizeof({cname}));')
        expect.append((f"sizeof({cname})", C.sizeof(cls)))
        for f in cls._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({cname}, {f[0]}));')
            expect.append((f"{cname}.{f[0]}", getattr(cls, f[0]).offset))
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert len(got) == len(expect)
    bad = [(name, want, g) for (name, want), g in zip(expect, got) if want != g]
    assert not bad, bad
