"""Regenerates tests/golden/philox_rocrand.json from rocRAND's Philox header
(host build with hipcc; no GPU is touched)."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "philox_rocrand")
        subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "-o", exe,
                        os.path.join(HERE, "philox_rocrand.cpp")], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    with open(os.path.join(HERE, "philox_rocrand.json"), "w") as f:
        f.write(out)


if __name__ == "__main__":
    main()
