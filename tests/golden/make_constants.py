"""Writes tests/golden/ref_constants.json from oracle/_ref/ref_constants,
the probe compiled (by `make -C oracle ref`) from the reference's own
mod_kind.f90 and mod_con.f90: each mod_con constant as the reference's
compiler evaluates it, as the hex of its IEEE binary64 pattern."""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_constants")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    out = subprocess.check_output([exe], text=True)
    table = {}
    for line in out.splitlines():
        name, hexv = line.split()
        table[name] = hexv.lower()
    path = os.path.join(HERE, "ref_constants.json")
    with open(path, "w") as f:
        json.dump({"source": "mod_con.f90 compiled with amdflang -O2 (no -r8)",
                   "binary64_hex": table}, f, indent=1, sort_keys=True)
    print("wrote", path, len(table), "constants")


if __name__ == "__main__":
    sys.exit(main())
