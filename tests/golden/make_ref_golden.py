"""Regenerate tests/golden/ref_geometry.json from the reference's own host code.

Runs oracle/_ref/ref_geom, which oracle/Makefile compiles from the reference's
variables.h, model.h, initializationTool.h and MRT_Matrix.h (C/ folder of the
reference) with a small harness that stands in for the absent main.cu.  The
fixture holds only numbers: inputs and the reference's outputs.

    make -C oracle ref && python tests/golden/make_ref_golden.py
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_geom")


def main() -> int:
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    if not os.path.exists(EXE):
        print("reference driver unavailable (reference tree absent)", file=sys.stderr)
        return 1
    out = subprocess.run([EXE], check=True, capture_output=True, text=True).stdout
    data = json.loads(out)
    data["_source"] = ("oracle/_ref/ref_geom built from the reference's variables.h, model.h, "
                       "initializationTool.h, MRT_Matrix.h")
    with open(os.path.join(HERE, "ref_geometry.json"), "w") as fh:
        json.dump(data, fh, separators=(",", ":"))
    print("wrote", os.path.join(HERE, "ref_geometry.json"))
    return 0


if __name__ == "__main__":
    sys.exit(main())
