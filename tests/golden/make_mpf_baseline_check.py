"""Extracts the check values printed by the reference's CPU baseline (cpu_baseline/cpu_*.txt in the
reference tree) into mpf_baseline_check.json: L[n-1][n-1] of the Cholesky factor and P[n-1][0] of
L^-1 B for its generated SPD matrix, at the printed precision. Run in the build container only.
    python tests/golden/make_mpf_baseline_check.py"""
import glob
import json
import os
import re

REF = "/root/reference/cpu_baseline"
out = []
for f in sorted(glob.glob(os.path.join(REF, "cpu_*.txt"))):
    txt = open(f).read()
    h = re.search(r"prec=(\d+) n=(\d+) nrhs=(\d+)", txt)
    c = re.search(r"check L\[n-1\]\[n-1\]=(\S+)\s+P\[n-1\]\[0\]=(\S+)", txt)
    if h and c:
        out.append({"file": os.path.basename(f), "prec": int(h.group(1)), "n": int(h.group(2)),
                    "nrhs": int(h.group(3)), "L_last": c.group(1), "P_last_0": c.group(2)})
json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "mpf_baseline_check.json"), "w"),
          indent=1)
print(out)
