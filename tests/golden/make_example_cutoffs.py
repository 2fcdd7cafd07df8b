"""Extract the Ewald cutoff lines of the reference's example runs into a small
JSON fixture (data only).  Runs in the build container, where the reference
checkout exists; the GPU box only reads the committed JSON.

For each example it records the inputs the cutoff rule depends on (box from
input_top.dat, Bjerrum length, alpha and dipole flag as printed in run.log,
which records what the run actually used) and the printed results: real-space
cutoff, k-space cutoff and the two rounded cell counts.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/examples"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "example_cutoffs.json")


def grab(text, label):
    m = re.search(re.escape(label) + r"\s*:\s*(\S+)", text)
    return m.group(1)


def main():
    rows = []
    for name in sorted(os.listdir(REF)):
        d = os.path.join(REF, name)
        if not os.path.isdir(d):
            continue
        log = open(os.path.join(d, "run.log")).read()
        top = open(os.path.join(d, "input_top.dat")).read().split()
        box = [float(top[top.index(k) + 1]) for k in
               ("Box_Length_X:", "Box_Length_Y:", "Box_Length_Z:")]
        rows.append({
            "example": name,
            "box": box,
            "lB": float(grab(log, "[EP] Bjerrum length (ul)")),
            "alpha": float(grab(log, "[EP] Ewald alpha (ul^-2)")),
            "dipole": grab(log, "[EP] Use dipole correction?") == "yes",
            "real_cutoff": float(grab(log, "[EP] Real space cutoff (ul)")),
            "cell_sum_number": int(grab(log, "[EP] (cell sum number)")),
            "repl_cutoff_printed": grab(log, "[EP] k space cutoff (ul^-2)"),
            "k_sum_number": int(grab(log, "[EP] (k sum number)")),
            "source": "examples/%s/run.log, examples/%s/input_top.dat" % (name, name),
        })
    with open(OUT, "w") as f:
        json.dump(rows, f, indent=1)
    print("wrote", OUT, len(rows))


if __name__ == "__main__":
    main()
