"""Namelist reader and parameter derivation against golden structs written
by the reference's own reader (src/readNamelist.f90 compiled with amdflang,
see tests/golden/make_params_golden.py).  CPU only."""
import ctypes as C
import json
import os

import pytest

import glemur_amd as g

HERE = os.path.dirname(os.path.abspath(__file__))
PDIR = os.path.join(HERE, "golden", "params")

# The reference declares blobXc..blobScale without defaults
# (src/readNamelist.f90:104): entries a file does not assign are undefined
# there (this library zero-fills them).  Only assigned entries are compared.
BLOB_FIELDS = ("blobXc", "blobYc", "blobZc", "blobZl", "blobTwist", "blobScale")
ASSIGNED = {
    "defaults": {},
    "reference_commas": {},
    "braid_c1": {f: range(6) for f in BLOB_FIELDS},
    "syntax_mix": {"blobXc": range(2), "blobZc": range(3)},
}


def golden(name):
    with open(os.path.join(PDIR, name + ".json")) as f:
        return json.load(f)


def product_fields(p):
    out = {}
    for name, ty in g.Params._fields_:
        v = getattr(p, name)
        if ty is C.c_double:
            out[name] = float(v).hex()
        elif ty is C.c_int:
            out[name] = int(v)
        elif ty is C.c_bool:
            out[name] = bool(v)
        elif issubclass(ty, C.Array) and ty._type_ is C.c_char:
            out[name] = bytes(C.string_at(C.addressof(p) + getattr(g.Params, name).offset, ty._length_)).hex()
        elif issubclass(ty, C.Array) and ty._type_ is C.c_double:
            out[name] = [float(x).hex() for x in v]
    return out


def compare(name, got, want):
    bad = []
    for k, wv in want.items():
        gv = got[k]
        if k in BLOB_FIELDS:
            idx = ASSIGNED[name].get(k, ())
            if any(gv[i] != wv[i] for i in idx):
                bad.append((k, [gv[i] for i in idx], [wv[i] for i in idx]))
            if not idx:
                assert all(float.fromhex(x) == 0.0 for x in gv), k
            continue
        if gv != wv:
            bad.append((k, gv, wv))
    assert not bad, bad


@pytest.mark.parametrize("name", ["defaults", "reference_commas", "braid_c1", "syntax_mix"])
def test_namelist_matches_reference_reader(name):
    p = g.params_read(os.path.join(PDIR, name + ".in"))
    compare(name, product_fields(p), golden(name))


def test_reference_params_in_with_semicolons():
    """The reference's own src/params.in separates items with ';', which
    gfortran accepts and amdflang's runtime does not; the product reads it
    to the same struct as the comma-separated copy."""
    p = g.params_read(os.path.join(PDIR, "reference_src_params.in"))
    compare("reference_commas", product_fields(p), golden("reference_commas"))


def test_defaults_single_precision_rounding():
    """Fortran default-real literals are single precision even in the FP64
    build (src/readNamelist.f90:100-117): dt0 = 3e-4 is float(3e-4)."""
    p = g.params_default()
    import numpy as np
    assert p.dt0 == float(np.float32(3e-4)) and p.dt0 != 3e-4
    assert p.width == float(np.float32(0.6))
    assert p.ar == float(np.float32(1.414213562373095))
    assert p.maxError == float(np.float32(5e-5))
    want = golden("defaults")
    assert float(p.dt0).hex() == want["dt0"] and float(p.ar).hex() == want["ar"]


def test_derived_spacing():
    """dx = Lx/(nx-1), dx1 = 1/dx (src/readNamelist.f90:185-186);
    initDist names map to initDistCode (src/readNamelist.f90:257-262)."""
    p = g.params_read(os.path.join(PDIR, "syntax_mix.in"))
    assert p.dx == p.Lx / (p.nx - 1) and p.dx1 == 1 / p.dx
    assert p.dz == p.Lz / (p.nz - 1) and p.dz1 == 1 / p.dz
    assert p.initDistCode == 0 and p.get_str("initDist") == "initShearX"
    q = g.params_default()
    q.set_str("initDist", "centerShift")
    g.params_derive(q)
    assert q.initDistCode == 1


def test_errors(tmp_path):
    with pytest.raises(g.GlemurError) as e:
        g.params_read(tmp_path / "missing.in")
    assert e.value.code == g.GM_ERR_IO if hasattr(g, "GM_ERR_IO") else e.value.code < 0
    bad = tmp_path / "bad.in"
    bad.write_text("&comp\n nx = 12, nosuchname = 3\n/\n&start\n/\n&run\n/\n&io\n/\n")
    with pytest.raises(g.GlemurError):
        g.params_read(bad)
    q = g.params_default()
    q.nx = 1
    with pytest.raises(g.GlemurError):
        g.params_derive(q)
