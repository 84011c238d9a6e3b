"""Writes tests/golden/aliev_panfilov_trace.txt: rows of the reference's own single-cell
trace tutorials/manufacturedSolutions/singleCellBidomain/postProcessing/
AlievPanfilov_myocyte_S1_1000.txt (time s, Vm mV, 7 decimals; AlievPanfilov myocyte,
RKF45 absTol 1e-6 relTol 1e-4, deltaT 5e-5 s, 1 s).  Kept: every 10th row, and every row
of 0.2575-0.2900 s (the stimulated upstroke).  Data only; run where the reference is:
    python tests/golden/make_aliev_panfilov_trace.py /root/reference
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = ("tutorials/manufacturedSolutions/singleCellBidomain/postProcessing/"
       "AlievPanfilov_myocyte_S1_1000.txt")


def main(ref):
    with open(os.path.join(ref, SRC)) as f:
        lines = f.read().splitlines()
    head, rows = lines[0], lines[1:]
    keep = [r for k, r in enumerate(rows)
            if (k + 1) % 10 == 0 or 0.2575 <= float(r.split()[0]) <= 0.29]
    with open(os.path.join(HERE, "aliev_panfilov_trace.txt"), "w") as f:
        f.write("# rows of %s (every 10th, all of 0.2575-0.2900 s)\n" % SRC)
        f.write(head + "\n")
        f.write("\n".join(keep) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
