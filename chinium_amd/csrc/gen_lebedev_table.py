"""Generate lebedev_table.inc (compiled into the library) from
chinium_amd/data/lebedev_orbits.txt, the project's table of Lebedev-Laikov
orbit generators. Run by the build before compiling gridgen.cpp."""
import os
import sys

here = os.path.dirname(os.path.abspath(__file__))
src = os.path.join(here, "..", "data", "lebedev_orbits.txt")
dst = sys.argv[1] if len(sys.argv) > 1 else os.path.join(here, "lebedev_table.inc")
rules = []
cur = None
for line in open(src):
    line = line.strip()
    if not line or line.startswith("#"):
        continue
    tok = line.split()
    if tok[0] == "rule":
        cur = [int(tok[1]), []]
        rules.append(cur)
    else:
        cur[1].append((int(tok[0]), tok[1], tok[2], tok[3]))
out = ["// Generated from data/lebedev_orbits.txt by gen_lebedev_table.py. Do not edit.",
       "static const LebOrbit kLebOrbits[] = {"]
index = []
k = 0
for n, orbits in rules:
    index.append((n, k, len(orbits)))
    for t, a, b, v in orbits:
        out.append("  {%d, %s, %s, %s}," % (t, a, b, v))
        k += 1
out.append("};")
out.append("static const LebRule kLebRules[] = {")
for n, first, cnt in index:
    out.append("  {%d, %d, %d}," % (n, first, cnt))
out.append("};")
open(dst, "w").write("\n".join(out) + "\n")
