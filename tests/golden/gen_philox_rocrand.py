"""Regenerates tests/golden/philox_rocrand.json from rocRAND's Philox engine.

Compiles gen_philox_rocrand.cpp host-side with hipcc (ROCm headers only; no
GPU needed) and captures its JSON output.
"""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "gen")
        subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-o", exe, os.path.join(HERE, "gen_philox_rocrand.cpp")],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    with open(os.path.join(HERE, "philox_rocrand.json"), "w") as f:
        f.write(out)


if __name__ == "__main__":
    main()
