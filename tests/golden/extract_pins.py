"""Extract the values the reference recorded in its validation notebooks into
reference_pins.json (run in a checkout that has the reference at REF; the
fixture is committed so tests never read the reference).

Each value is parsed from the notebook's stored cell output, and the cell is
identified by its source text; the JSON records notebook, cell index and the
printed text.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_pins.json")


def cells(nb):
    with open(os.path.join(REF, "Test", nb)) as fh:
        return json.load(fh)["cells"]


def out_text(cell):
    parts = []
    for o in cell.get("outputs", []):
        if "text" in o:
            parts.append("".join(o["text"]))
        elif "data" in o and "text/plain" in o["data"]:
            parts.append("".join(o["data"]["text/plain"]))
    return "\n".join(parts)


def find(nb, must_contain, idx=None):
    for i, c in enumerate(cells(nb)):
        if c["cell_type"] == "code" and must_contain in "".join(c["source"]) and (idx is None or i == idx):
            return i, out_text(c)
    raise KeyError((nb, must_contain))


FLOAT = r"-?\d+\.\d+(?:e[-+]?\d+)?"
pins = {}

i, t = find("beamcols_fea_validation.ipynb", "model.solve(which_solver='dense')", 4)
v = [float(x) for x in re.findall(FLOAT, t)]
pins["arch_dense"] = dict(notebook="Test/beamcols_fea_validation.ipynb", cell=i, min_uz=v[0], C=v[1], text=t.strip())
i, t = find("beamcols_fea_validation.ipynb", "model.solve(which_solver='sparse')", 5)
v = [float(x) for x in re.findall(FLOAT, t)]
pins["arch_scipy"] = dict(notebook="Test/beamcols_fea_validation.ipynb", cell=i, min_uz=v[0], C=v[1], text=t.strip())
i, t = find("beamcols_ad_validation.ipynb", "print('AD,Dense")
v = [float(x) for x in re.findall(FLOAT, t)]
pins["arch_grad_node49"] = dict(notebook="Test/beamcols_ad_validation.ipynb", cell=i, node=49,
                                dense=v[0], jax_sparse=v[1], scipy=v[2], text=t.strip())
for solver, idx in (("dense", 4), ("scipy", 5), ("jax", 6)):
    i, t = find("shells_fea_validation.ipynb", "print(np.min(uz))", idx)
    pins[f"barrel_min_uz_{solver}"] = dict(notebook="Test/shells_fea_validation.ipynb", cell=i,
                                           min_uz=float(re.findall(FLOAT, t)[0]), text=t.strip())
i, t = find("shells_fea_validation.ipynb", "print(0.5*model.u@model.get_loads())")
pins["barrel_C"] = dict(notebook="Test/shells_fea_validation.ipynb", cell=i, C=float(re.findall(FLOAT, t)[0]),
                        text=t.strip())
i, t = find("shells_ad_validation.ipynb", "print('AD,Dense")
v = [float(x) for x in re.findall(FLOAT, t)]
pins["barrel_grad_node201"] = dict(notebook="Test/shells_ad_validation.ipynb", cell=i, node=201,
                                   dense=v[0], jax_sparse=v[1], scipy=v[2], text=t.strip())
# multi-span beam chains f(n, 100): min u_z per n (Test/Frames_speed.ipynb cell 18)
i, t = find("Frames_speed.ipynb", "for i in [2,4,6,8,10,20,30,80,120,240]:", 18)
ns = [int(x) for x in re.findall(r"Each span has (\d+) elements", t)]
uz = [float(x) for x in re.findall(r"\n(" + FLOAT + r")\nDOF", t)]
pins["frames_min_uz"] = dict(notebook="Test/Frames_speed.ipynb", cell=i, m=100, n=ns, min_uz=uz)

with open(OUT, "w") as fh:
    json.dump(pins, fh, indent=1)
print(json.dumps({k: {kk: vv for kk, vv in v.items() if kk != "text"} for k, v in pins.items()}, indent=1))
