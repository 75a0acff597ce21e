"""The C ABI (include/glemur.h) against the built library and the Python
binding: every declared entry point is exported and bound, and the struct
layouts the binding assumes match what a C compiler makes of the header.
No GPU needed (nothing here calls into HIP)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import glemur_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "glemur.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gm_[a-z0-9_]+)\s*\(", text)))


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", g.LIB_PATH], capture_output=True, text=True, check=True)
    return {l.split()[-1] for l in out.stdout.splitlines() if l.strip()}


def test_every_declared_function_is_exported_and_bound():
    decl = declared_functions()
    assert len(decl) >= 40
    syms = exported_symbols()
    missing = [f for f in decl if f not in syms]
    assert not missing, missing
    unbound = [f for f in decl if f not in g.EXPORTED]
    assert not unbound, unbound
    for f in decl:
        assert getattr(g.lib, f) is not None


def test_no_undeclared_gm_exports():
    syms = {s for s in exported_symbols() if s.startswith("gm_")}
    extra = syms - set(declared_functions())
    assert not extra, extra


STRUCTS = {
    "gm_params": g.Params,
    "gm_reduction": g.Reduction,
    "gm_step_info": g.StepInfo,
    "gm_run_opts": g.RunOpts,
    "gm_run_stats": g.RunStats,
    "gm_step_record": g.StepRecord,
    "gm_warm_info": g.WarmInfo,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    """sizeof/offsetof of every bound struct field, as gcc lays out the header."""
    d = tmp_path_factory.mktemp("abi")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "glemur.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = d / "layout.c"
    src.write_text("\n".join(lines))
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    lay = {}
    for l in out.splitlines():
        s, f, v = l.split()
        lay[(s, f)] = int(v)
    return lay


@pytest.mark.parametrize("cname", list(STRUCTS))
def test_struct_layout_matches_header(c_layout, cname):
    py = STRUCTS[cname]
    assert C.sizeof(py) == c_layout[(cname, "sizeof")]
    for fname, _ in py._fields_:
        assert getattr(py, fname).offset == c_layout[(cname, fname)], fname


def test_oracle_params_layout_equals_product():
    """The oracle's struct is the same struct (parameters are memmoved)."""
    import oracle as o
    assert C.sizeof(o.Params) == C.sizeof(g.Params)
    for (a, ta), (b, tb) in zip(o.Params._fields_, g.Params._fields_):
        assert a == b and getattr(o.Params, a).offset == getattr(g.Params, b).offset


def test_error_convention():
    """Negative codes with a message from gm_last_error; no exit() inside."""
    assert g.lib.gm_version() > 0
    assert g.lib.gm_params_default(None) < 0
    assert len(g.lib.gm_last_error()) > 0
    ctx = C.c_void_p()
    assert g.lib.gm_create(None, 0, C.byref(ctx)) < 0
    assert g.lib.gm_destroy(None) <= 0
