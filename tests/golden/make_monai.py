"""Writes the Monai valley fixture (tests/golden/monai/) from the reference's
own case files, cases/monai/ of al0vya/gpu-mwdg2:

- bed.npz: the bed elevations of bed-data.txt (third column, 244 x 393
  nodes in file order), float64, compressed;
- monai.bci, monai.bdy, monai.stage: copied as they are;
- res-<solver>-eps-<eps>.txt: the gauge depths the reference binary wrote
  (FP32, max_ref_lvl 9, 22.5 s) for hwfv1/mwdg2 at epsilon 0 and 1e-3.

    python tests/golden/make_monai.py [/root/reference]
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "cases", "monai")
    out = os.path.join(HERE, "monai")
    os.makedirs(out, exist_ok=True)
    z = np.loadtxt(os.path.join(src, "bed-data.txt"), skiprows=1, usecols=2)
    np.savez_compressed(os.path.join(out, "bed.npz"), z=z)
    for f in ("monai.bci", "monai.bdy", "monai.stage"):
        shutil.copyfile(os.path.join(src, f), os.path.join(out, f))
    for solver in ("hwfv1", "mwdg2"):
        for eps in ("0", "0.001"):
            shutil.copyfile(os.path.join(src, f"{solver}-eps-{eps}", "res.txt"),
                            os.path.join(out, f"res-{solver}-eps-{eps}.txt"))
    print(out)


if __name__ == "__main__":
    main()
