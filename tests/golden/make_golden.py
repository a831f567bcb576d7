"""Writes the golden fixtures of the advection stencils (run in a container that holds the reference).

- stencil_16.npz: the inputs the reference drivers compiler/tests/fixtures/stencils/{cdu,cvd,cdw}_main.f90
  build at 16^3 (dumped by dump_inputs.f90, which restates their initialisation loops), and the interiors
  of U2, V2 and W2 that the unchanged reference drivers print (g0.17, exact round trip), built from the
  reference sources by `make -C oracle ref`.

Only data is committed; the reference sources stay where they are.
"""
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
REF = ROOT / "oracle" / "_ref"
FC = "/opt/rocm/bin/amdflang"
N = 16


def main():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
    outputs = {}
    for name, driver in (("cdu", "cdu_main"), ("cdv", "cvd_main"), ("cdw", "cdw_main")):
        text = subprocess.run([str(REF / driver)], capture_output=True, text=True, check=True).stdout
        values = np.array([float(line) for line in text.split()], dtype=np.float64)
        assert values.size == N ** 3, values.size
        outputs[name] = values.reshape((N, N, N), order="F")
    with tempfile.TemporaryDirectory() as tmp:
        exe = Path(tmp) / "dump_inputs"
        subprocess.run([FC, "-O3", "-ffp-contract=off", str(HERE / "dump_inputs.f90"), "-o", str(exe)], check=True)
        subprocess.run([str(exe)], cwd=tmp, check=True)
        raw = np.fromfile(Path(tmp) / "stencil_inputs.bin", dtype=np.float64)
    e = N + 2
    U, V, W = (raw[n * e ** 3:(n + 1) * e ** 3].reshape((e, e, e), order="F") for n in range(3))
    np.savez_compressed(HERE / "stencil_16.npz", U=U, V=V, W=W, U2=outputs["cdu"], V2=outputs["cdv"],
                        W2=outputs["cdw"], compiler=np.array(FC))
    print("wrote", HERE / "stencil_16.npz")


if __name__ == "__main__":
    sys.exit(main())
