#!/usr/bin/env python3
"""Write the golden parsed-parameter fixtures tests/golden/params/*.json.

Each params/<name>.in is read by the reference's own namelist reader
(src/readNamelist.f90, compiled with amdflang -DPRECISION=64 into
oracle/_ref/libreadnamelist.so by `make -C oracle ref`), in a fresh process
whose working directory holds that file as params.in (the reader opens
'params.in' in the current directory, and its namelist variables keep their
values between calls, so every file gets its own process).  The struct it
fills is dumped field by field: doubles as float.hex, strings as the raw 30
bytes.

Run in a container where the reference sources are present:
    make -C oracle ref && python tests/golden/make_params_golden.py
Files the reference reader rejects (e.g. ';' separators) get no .json.
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libreadnamelist.so")

CHILD = r"""
import ctypes as C, json, sys
sys.path.insert(0, sys.argv[1])
from oracle import Params
lib = C.CDLL(sys.argv[2])
p = Params()
lib.readnamelist(C.byref(p))
out = {}
for name, ty in Params._fields_:
    v = getattr(p, name)
    if ty is C.c_double:
        out[name] = float(v).hex()
    elif ty is C.c_int:
        out[name] = int(v)
    elif ty is C.c_bool:
        out[name] = bool(v)
    elif issubclass(ty, C.Array) and ty._type_ is C.c_char:
        out[name] = bytes(C.string_at(C.addressof(p) + getattr(Params, name).offset, ty._length_)).hex()
    elif issubclass(ty, C.Array) and ty._type_ is C.c_double:
        out[name] = [float(x).hex() for x in v]
    else:
        raise TypeError(name)
print(json.dumps(out))
"""


def main():
    if not os.path.exists(LIB):
        sys.exit(f"{LIB} missing: run `make -C oracle ref` first")
    pdir = os.path.join(HERE, "params")
    for fn in sorted(os.listdir(pdir)):
        if not fn.endswith(".in"):
            continue
        with tempfile.TemporaryDirectory() as tmp:
            shutil.copy(os.path.join(pdir, fn), os.path.join(tmp, "params.in"))
            r = subprocess.run([sys.executable, "-c", CHILD, ROOT, LIB], cwd=tmp, capture_output=True, text=True)
        out = os.path.join(pdir, fn[:-3] + ".json")
        if r.returncode != 0 or not r.stdout.strip():
            print(f"{fn}: rejected by the reference reader (rc={r.returncode}); no fixture")
            if os.path.exists(out):
                os.remove(out)
            continue
        with open(out, "w") as f:
            json.dump(json.loads(r.stdout), f, indent=1, sort_keys=True)
            f.write("\n")
        print(f"{fn}: wrote {os.path.basename(out)}")


if __name__ == "__main__":
    main()
