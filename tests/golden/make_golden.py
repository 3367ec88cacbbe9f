"""Collect the golden fixtures of the parity tests.

1. The reference's only numerically sharp fixtures: the N=7 P_N flux splittings
   ``code/validationData/A{x,y}{Plus,Minus}.csv`` (64 x 64, never read by the
   reference's own scripts).  They hold (A +- |A|)/2 for the x and y streaming
   matrices; copied verbatim (data, 6 significant digits).
2. Two of the reference's dose outputs ``code/output/dose_csd_nx{20,50}ny{20,50}
   .{txt,vtr}`` (writedlm text and WriteVTK RectilinearGrid files; the 50 x 50
   grid also carries the "u" array): the byte-level fixtures of
   ``titus_amd/io.py``.  Copied verbatim (data).
3. ``code/Lung.png``, the image the "lung" preset reads (``settings.jl:266-283``).

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
    out = os.path.join(code_dir, "output")
    for n in (20, 50):
        for ext in ("txt", "vtr"):
            name = f"dose_csd_nx{n}ny{n}.{ext}"
            shutil.copyfile(os.path.join(out, name), os.path.join(HERE, name))
            print("copied", name)
    shutil.copyfile(os.path.join(code_dir, "Lung.png"), os.path.join(HERE, "Lung.png"))
    print("copied Lung.png")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/code")
