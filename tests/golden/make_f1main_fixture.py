"""Extract the reference's own 1-D end-to-end golden vectors into a small
numpy fixture (tests/golden/f1main_expected.npz), and copy the three HDF5
files themselves beside it (f1main_expected_*.h5, 190 KB: data the reference's
own test holds, used to check the Puffin-format HDF5 reader on files the
reference wrote).

Source: test/expected/f1main_expected_{electrons,field,integrated}.h5 of the
reference, produced by its pFUnit test test/testMPIIntegration.pf:54-127 on
the deck test/inputs/f1main.in (1-D, 30 steps, 319 macroparticles).  h5py is
not available, so the HDF5 files are read through the ``h5dump`` tool at full
precision (%.17g).  Run once in a container that has the reference mounted:

    python tests/golden/make_f1main_fixture.py /root/reference
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys

import numpy as np

H5DUMP = os.environ.get("H5DUMP", "/opt/conda/bin/h5dump")
HERE = os.path.dirname(os.path.abspath(__file__))


def _dump(path: str, obj: str, attr: bool = False) -> str:
    flag = "-a" if attr else "-d"
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-y", "-w", "0", flag, obj, path],
                         check=True, stdout=subprocess.PIPE, text=True).stdout
    return out


def _numbers(text: str) -> np.ndarray:
    body = text.split("DATA {", 1)[1]
    body = body.split("}", 1)[0]
    vals = re.findall(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?", body)
    return np.array([float(v) for v in vals])


def dataset(path: str, name: str) -> np.ndarray:
    return _numbers(_dump(path, name))


def attribute(path: str, name: str) -> float:
    return float(_numbers(_dump(path, name, attr=True))[0])


def main(ref_root: str) -> None:
    exp = os.path.join(ref_root, "test", "expected")
    fe = os.path.join(exp, "f1main_expected_electrons.h5")
    ff = os.path.join(exp, "f1main_expected_field.h5")
    fi = os.path.join(exp, "f1main_expected_integrated.h5")
    electrons = dataset(fe, "/electrons").reshape(319, 7)
    aperp = dataset(ff, "/aperp").reshape(2, 511)
    power = dataset(fi, "/power")
    bunch = dataset(fi, "/bunchingFundamental")
    info = {}
    for key in ("eta", "kappa", "rho", "aw", "gamma_r", "sLengthOfElmZ2", "sStepSize",
                "nZ2", "nSteps", "zbarTotal", "lambda_w", "transArea", "fieldMesh"):
        info[key] = attribute(fi, "/runInfo/" + key)
    out = os.path.join(HERE, "f1main_expected.npz")
    np.savez(out, electrons=electrons, aperp=aperp, power=power,
             bunching_fundamental=bunch,
             info_keys=np.array(list(info.keys())),
             info_vals=np.array(list(info.values())))
    print("wrote", out, electrons.shape, aperp.shape, power.shape, bunch.shape)
    for f in (fe, ff, fi):
        shutil.copyfile(f, os.path.join(HERE, os.path.basename(f)))
        print("copied", os.path.basename(f))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
