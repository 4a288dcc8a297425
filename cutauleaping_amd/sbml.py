"""SBML -> input-folder converter (Python 3, standard library only).

The reference's converter (SBML2BSW.py, Python 2 + libsbml) writes left_side,
right_side, M_feed, M_0, c_vector, alphabet and boundaries
(SBML2BSW.py:287-323), but not the MX_0 / c_matrix / t_vector files its own
CLI reads (kernel.cu:143-147).  This converter follows SBML2BSW.py's model
semantics and writes both sets:

  species   name = name (or id), suffixed "_in_<compartment>" when a
            compartment is set (SBML2BSW.py:58-66, 331-337); the initial count
            is initialAmount, else initialConcentration (:72-75); species are
            numbered in document order, first occurrence of a name wins (:87-92)
  feeds     M_feed = 1 for constant species (:297-300)
  reactions one row per reaction; stoichiometries of repeated species
            overwrite (:242-262); rate constants are the kinetic law's local
            parameters (:229-231), or else the global parameters its formula
            names, a reversible reaction or two such parameters adding a
            reverse row with swapped sides (:174-201, 264-269)
Mass-action kinetics only (the simulator's propensities are mass action).

    python -m cutauleaping_amd.sbml model.xml out_dir --rows 1024 --t-end 10 --samples 100
"""
from __future__ import annotations

import argparse
import math
import os
import xml.etree.ElementTree as ET


def _local(tag: str) -> str:
    return tag.rsplit("}", 1)[-1]


def _children(el, name):
    return [c for c in el if _local(c.tag) == name]


def _find(el, *path):
    cur = [el]
    for name in path:
        cur = [c for e in cur for c in _children(e, name)]
    return cur


def _float(s, default=float("nan")):
    try:
        return float(s)
    except (TypeError, ValueError):
        return default


class SBMLModel:
    """Parsed model: species names, initial counts, feeds, stoichiometry, rates."""

    def __init__(self, path: str):
        root = ET.parse(path).getroot()
        models = _find(root, "model")
        if not models:
            raise ValueError(f"{path}: no <model> element")
        model = models[0]
        self.names, self.x0, self.feed = [], [], []
        index, id2name = {}, {}
        for sp in _find(model, "listOfSpecies", "species"):
            name = sp.get("name") or sp.get("id")
            comp = sp.get("compartment") or ""
            if comp:
                name = f"{name}_in_{comp}"
            id2name[sp.get("id")] = name
            amount = _float(sp.get("initialAmount"))
            if math.isnan(amount):
                amount = _float(sp.get("initialConcentration"), 0.0)
            if name in index:
                continue
            index[name] = len(self.names)
            self.names.append(name)
            self.x0.append(amount)
            self.feed.append(1 if sp.get("constant") == "true" else 0)
        params = {p.get("id"): _float(p.get("value"), 0.0) for p in _find(model, "listOfParameters", "parameter")}
        S = len(self.names)
        self.left, self.right, self.c, self.reactions = [], [], [], []
        for rx in _find(model, "listOfReactions", "reaction"):
            rid = rx.get("id") or rx.get("name") or f"R{len(self.reactions)}"
            lv, rv = [0] * S, [0] * S
            for side, vec in (("listOfReactants", lv), ("listOfProducts", rv)):
                for ref in _find(rx, side, "speciesReference"):
                    st = _float(ref.get("stoichiometry"), 1.0)
                    vec[index[id2name[ref.get("species")]]] = int(round(st))
            laws = _find(rx, "kineticLaw")
            rates, reverse = [], False
            if laws:
                law = laws[0]
                local = _find(law, "listOfParameters", "parameter") + _find(law, "listOfLocalParameters",
                                                                           "localParameter")
                if local:
                    rates = [_float(p.get("value"), 0.0) for p in local]
                else:
                    names = [ci.text.strip() for ci in law.iter() if _local(ci.tag) == "ci" and ci.text]
                    used = []
                    for n in names:
                        if n in params and n not in used:
                            used.append(n)
                    if not used:
                        raise ValueError(f"no kinetic parameter found for reaction {rid}")
                    if len(used) > 2:
                        raise ValueError(f"too many parameters in the kinetic law of {rid}")
                    reverse = rx.get("reversible") == "true" or len(used) == 2
                    rates = [params[n] for n in used]
            if not rates:
                raise ValueError(f"reaction {rid} has no kinetic law")
            self.reactions.append(rid)
            self.left.append(lv)
            self.right.append(rv)
            self.c.append(rates[0])
            if reverse:
                self.reactions.append(rid + " (reverse)")
                self.left.append(rv)
                self.right.append(lv)
                self.c.append(rates[1] if len(rates) > 1 else rates[0])

    def write_folder(self, out: str, rows: int = 1, t_end: float = 10.0, samples: int = 100,
                     sampled=None) -> None:
        """Writes the reference's folder: left_side, right_side, M_feed, MX_0
        and c_matrix (rows identical rows), t_vector (samples uniform times to
        t_end), cs_vector (optional), plus M_0, c_vector and alphabet."""
        os.makedirs(out, exist_ok=True)
        tab = lambda row: "\t".join(str(v) for v in row)
        w = lambda name, text: open(os.path.join(out, name), "w").write(text)
        w("left_side", "".join(tab(r) + "\n" for r in self.left))
        w("right_side", "".join(tab(r) + "\n" for r in self.right))
        w("M_feed", tab(self.feed) + "\n")
        x0 = [int(v) if float(v).is_integer() else v for v in self.x0]
        w("MX_0", (tab(x0) + "\n") * rows)
        w("c_matrix", (tab(f"{v:.17g}" for v in self.c) + "\n") * rows)
        w("t_vector", "".join(f"{t_end * (i + 1) / samples:.17g}\n" for i in range(samples)))
        if sampled is not None:
            w("cs_vector", "".join(f"{int(s)}\n" for s in sampled))
        w("M_0", "\t".join(f"{float(v):e}" for v in self.x0) + "\n")
        w("c_vector", "".join(f"{v:e}\n" for v in self.c))
        w("alphabet", "".join(n + "\t" for n in self.names))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("sbml")
    ap.add_argument("out_dir")
    ap.add_argument("--rows", type=int, default=1, help="rows of MX_0 / c_matrix (simulations)")
    ap.add_argument("--t-end", type=float, default=10.0)
    ap.add_argument("--samples", type=int, default=100)
    ap.add_argument("--sampled", type=str, default=None, help="comma-separated species indices (cs_vector)")
    a = ap.parse_args(argv)
    m = SBMLModel(a.sbml)
    sampled = [int(s) for s in a.sampled.split(",")] if a.sampled else None
    m.write_folder(a.out_dir, rows=a.rows, t_end=a.t_end, samples=a.samples, sampled=sampled)
    print(f" * {len(m.names)} species, {len(m.reactions)} reactions written to {a.out_dir}")


if __name__ == "__main__":
    main()
