"""Extract the pigment data of the reference's assets/pigments.ron (a RON
text file) into pigments_ron.json: per pigment name, 16-band reflectance,
scatter, tint (default 1 when absent, src/color/mod.rs:12-25), viscosity,
yield stress.  Run once in the container that has the reference mounted."""
import json
import os
import re
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/assets/pigments.ron"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pigments_ron.json")

txt = open(SRC).read()
out = []
for block in re.findall(r"\(\s*name:\s*\"([^\"]+)\",(.*?)\n        \)", txt, re.S):
    name, body = block
    refl = [float(v) for v in re.search(r"reflectance:\s*\[(.*?)\]", body, re.S).group(1).replace("\n", " ").split(",") if v.strip()]

    def num(key, default=None):
        m = re.search(key + r":\s*([-0-9.eE+]+)", body)
        return float(m.group(1)) if m else default
    out.append({"name": name, "reflectance": refl, "scatter": num("scatter"), "tint": num("tint", 1.0),
                "viscosity": num("viscosity"), "yield_stress": num("yield_stress")})
assert len(out) == 12, len(out)
json.dump({"source": "assets/pigments.ron", "pigments": out}, open(OUT, "w"), indent=1)
print(OUT, len(out))
