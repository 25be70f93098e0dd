"""Writes tests/golden/radial_readme_energies.json: the radial-atom total
energies and iteration counts the reference publishes in its README.md
(UTF-16; tables 'Resultados PZ81' and 'Resultados PBE96', lines 173-241),
kept as printed.  Runs only where the reference checkout exists."""
import json
import os
import re

README = "/root/reference/README.md"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "radial_readme_energies.json")


def main():
    txt = open(README, "rb").read().decode("utf-16")
    out, cur = {}, None
    for ln in txt.splitlines():
        if ln.startswith("# "):
            cur = "pz81" if "Resultados PZ81" in ln else "pbe96" if "Resultados PBE96" in ln else None
        m = re.match(r"\|\s*(\d+)\s*\|\s*(\w+)\s*\|\s*(\d+)\+(\d+)\s*\|\s*(\d+)\s*\|\s*(-?[\d.]+)\s*\|", ln)
        if m and cur:
            out.setdefault(cur, {})[m.group(1)] = {"symbol": m.group(2), "alpha": int(m.group(3)),
                                                   "beta": int(m.group(4)), "iterations": int(m.group(5)),
                                                   "total": m.group(6)}
    doc = {"source": "the reference's README.md (UTF-16), tables 'Resultados PZ81' and 'Resultados PBE96' "
                     "(lines 173-241): solveAtom on RadialGrid(2000, 30.0); totals kept as printed (decimal strings)",
           "grid": {"points": 2000, "rmax": 30.0}, **out}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)


if __name__ == "__main__":
    main()
