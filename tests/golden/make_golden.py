"""Writes the reference known-answer fixtures under tests/golden/.

Runs oracle/_ref/ref_kat -- built by `make -C oracle ref` from the reference's
own sources (camera.cpp, transform.cpp, filter.cpp, table_dist.cpp,
material.cpp, phase_function.cpp) in the authoring container -- and stores its
JSON output (seeded inputs plus the reference's outputs) as fixtures:

    python tests/golden/make_golden.py

The reference never travels to the GPU host; these fixtures do.
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
KAT = os.path.join(ROOT, "oracle", "_ref", "ref_kat")


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    for mode in ("pcg", "table", "camera", "materials", "phase", "texture"):
        out = subprocess.run([KAT, mode], check=True, capture_output=True, text=True).stdout
        data = json.loads(out)
        path = os.path.join(HERE, f"ref_{mode}.json")
        with open(path, "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print(f"wrote {path} ({os.path.getsize(path)} bytes)", file=sys.stderr)


if __name__ == "__main__":
    main()
