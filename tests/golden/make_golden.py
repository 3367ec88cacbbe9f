"""Collect the golden fixtures of the parity tests.

1. The reference's only numerically sharp fixtures: the N=7 P_N flux splittings
   ``code/validationData/A{x,y}{Plus,Minus}.csv`` (64 x 64, never read by the
   reference's own scripts).  They hold (A +- |A|)/2 for the x and y streaming
   matrices; copied verbatim (data, 6 significant digits).

Run where the reference tree is mounted:

    python tests/golden/make_golden.py /root/reference/code
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(code_dir):
    src = os.path.join(code_dir, "validationData")
    for name in ("AxPlus.csv", "AxMinus.csv", "AyPlus.csv", "AyMinus.csv"):
        shutil.copyfile(os.path.join(src, name), os.path.join(HERE, name))
        print("copied", name)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/code")
