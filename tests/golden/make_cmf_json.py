"""Extract the colour-matching data of the reference's src/color/mod.rs (the
CMF table: CIE 1931 2-degree x, y, z at 400-710 nm in 10 nm steps, and the
D65 table at the same wavelengths) into cmf_d65.json.  Data only: the
numbers as the reference holds them.  Run once in the container that has the
reference mounted."""
import json
import os
import re
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/color/mod.rs"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cmf_d65.json")

txt = open(SRC).read()
cmf_body = re.search(r"const CMF: \[\[f32; 4\]; 32\] = \[(.*?)\];", txt, re.S).group(1)
rows = [[float(v) for v in r.split(",") if v.strip()] for r in re.findall(r"\[([^\[\]]+)\]", cmf_body)]
d65_body = re.search(r"const D65: \[f32; 32\] = \[(.*?)\];", txt, re.S).group(1)
d65 = [float(v) for v in d65_body.replace("\n", " ").split(",") if v.strip()]
assert len(rows) == 32 and all(len(r) == 4 for r in rows), len(rows)
assert len(d65) == 32, len(d65)
json.dump({"source": "src/color/mod.rs (const CMF, const D65)",
           "nm": [r[0] for r in rows], "xbar": [r[1] for r in rows], "ybar": [r[2] for r in rows],
           "zbar": [r[3] for r in rows], "d65": d65}, open(OUT, "w"), indent=1)
print(OUT)
