"""Builds tests/golden/make_xorwow_vectors.cpp with the ROCm compiler (host
code only: rocRAND's XORWOW is host and device code) and writes
tests/golden/xorwow_vectors.json."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "xorwow_vectors")
        subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", os.path.join(HERE, "make_xorwow_vectors.cpp"),
                        "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    with open(os.path.join(HERE, "xorwow_vectors.json"), "w") as f:
        f.write(out)


if __name__ == "__main__":
    main()
