"""Write tests/golden/air11_park_tables.json from the reference's Park
mechanism source, read as text.

The data file hybrid/chemistry/air11_park.py of the reference
(superboss2300/jax-fluid-sim) holds the species (molar mass, charge, NASA-7
fits) and the 22 reactions as literal constructor calls.  This script parses
that file with ``ast`` and evaluates only literals, names bound to literals,
arithmetic on them, ``range`` and the three constructors (``make_nasa7``,
``Species``, ``Reaction``, whose defaults are those of
hybrid/chemistry/mechanism.py:27-55).  Nothing from the reference is imported
or executed.  The JSON it writes is a fixture of data: the chemistry tests
compare the oracle's and the product's tables against it.

Usage: python tests/golden/extract_park_tables.py [reference root]
(default /root/reference).
"""

import ast
import json
import operator
import sys
from pathlib import Path

SPECIES_DEFAULTS = {"charge": 0.0, "poly": None, "is_diatomic": False, "theta_v": 0.0}
REACTION_DEFAULTS = {"third_body": False, "efficiencies": (), "reversible": True,
                     "fwd_temp": "T", "rev_temp": "T"}
BINOPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
          ast.Div: operator.truediv}


class Evaluator:
    def __init__(self):
        self.env = {}

    def value(self, node):
        if isinstance(node, ast.Constant):
            return node.value
        if isinstance(node, ast.Name):
            return self.env[node.id]
        if isinstance(node, (ast.Tuple, ast.List)):
            return tuple(self.value(e) for e in node.elts)
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = self.value(node.operand)
            return -v if isinstance(node.op, ast.USub) else v
        if isinstance(node, ast.BinOp) and type(node.op) in BINOPS:
            return BINOPS[type(node.op)](self.value(node.left), self.value(node.right))
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name):
            args = [self.value(a) for a in node.args]
            kw = {k.arg: self.value(k.value) for k in node.keywords}
            return self.call(node.func.id, args, kw)
        raise ValueError(f"unsupported expression at line {node.lineno}: {ast.dump(node)[:80]}")

    @staticmethod
    def call(name, args, kw):
        if name == "range":
            return tuple(range(*args))
        if name == "make_nasa7":
            keys = ("t_low", "t_mid", "t_high", "low", "high")
            d = dict(zip(keys, args))
            d.update(kw)
            return {k: (list(map(float, d[k])) if k in ("low", "high") else float(d[k]))
                    for k in keys}
        if name == "Species":
            d = dict(SPECIES_DEFAULTS)
            d.update(dict(zip(("name", "molar_mass", "charge", "poly"), args)))
            d.update(kw)
            return d
        if name == "Reaction":
            d = dict(REACTION_DEFAULTS)
            d.update(dict(zip(("reactants", "products", "A", "b", "theta"), args)))
            d.update(kw)
            d["reactants"] = [list(map(int, r)) for r in d["reactants"]]
            d["products"] = [list(map(int, r)) for r in d["products"]]
            d["efficiencies"] = [[int(s), float(e)] for s, e in d["efficiencies"]]
            for k in ("A", "b", "theta"):
                d[k] = float(d[k])
            return d
        raise ValueError(f"unsupported call {name}")


def extract(path):
    tree = ast.parse(Path(path).read_text(), filename=str(path))
    ev = Evaluator()
    for stmt in tree.body:
        if not isinstance(stmt, ast.Assign) or len(stmt.targets) != 1:
            continue
        target = stmt.targets[0]
        try:
            val = ev.value(stmt.value)
        except (ValueError, KeyError):
            continue
        if isinstance(target, ast.Name):
            ev.env[target.id] = val
        elif isinstance(target, ast.Tuple):
            for t, v in zip(target.elts, val):
                ev.env[t.id] = v
    species = [dict(name=s["name"], molar_mass=float(s["molar_mass"]),
                    charge=float(s["charge"]), nasa7=s["poly"]) for s in ev.env["AIR11_SPECIES"]]
    return {"source": "hybrid/chemistry/air11_park.py", "species": species,
            "reactions": list(ev.env["AIR11_REACTIONS"])}


if __name__ == "__main__":
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    data = extract(root / "hybrid" / "chemistry" / "air11_park.py")
    out = Path(__file__).with_name("air11_park_tables.json")
    out.write_text(json.dumps(data, indent=1) + "\n")
    print(f"wrote {out}: {len(data['species'])} species, {len(data['reactions'])} reactions")
