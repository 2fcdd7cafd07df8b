"""Copy the reference's example input decks (run.in, input_top.dat,
input_crd.dat -- data files, ~57 KB in all) into tests/golden/examples/ as
parser fixtures.  Runs in the build container, where the reference checkout
exists; the GPU box only reads the copies."""
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/examples"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "examples")

for name in sorted(os.listdir(REF)):
    src = os.path.join(REF, name)
    if not os.path.isdir(src):
        continue
    dst = os.path.join(OUT, name)
    os.makedirs(dst, exist_ok=True)
    for f in ("run.in", "input_top.dat", "input_crd.dat"):
        shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
    print("copied", name)
