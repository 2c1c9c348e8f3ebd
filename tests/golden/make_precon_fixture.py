"""Extract the preconditioner dump held in the reference
(qed_fermion/preconditioners_orig/inspect_precon.md, written by inspect_precon.py
for L=8, Ltau=80) into tests/golden/precon_dump_L8_Lt80.json.

The dump lists, for tau offsets 0..6 and -1..-5 (cyclic offset Ltau-r), the
value of P[row, row + offset*Vs] at rows 3 + tau*Vs for tau = 0..79 (-100 marks
an absent entry).  Run in a container that has the reference mounted:
    python tests/golden/make_precon_fixture.py /root/reference
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = os.path.join(ref, "qed_fermion", "preconditioners_orig", "inspect_precon.md")
text = open(src).read()
out = {}
for m in re.finditer(r"Rank\((-?\d+)\) diagonal of precon:\s*\[([^\]]*)\]", text):
    out[m.group(1)] = [float(v) for v in m.group(2).split()]
assert sorted(out, key=int) == [str(k) for k in range(-5, 7)], sorted(out)
assert all(len(v) == 80 for v in out.values())
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "precon_dump_L8_Lt80.json")
json.dump({"source": "qed_fermion/preconditioners_orig/inspect_precon.md", "L": 8, "Ltau": 80, "rank": out}, open(dst, "w"),
          indent=0)
print("wrote", dst)
