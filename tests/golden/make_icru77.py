"""Extract the ICRU-77 total electron stopping-power rows this package uses.

Source data: ICRU Report 77 (Elastic Scattering of Electrons and Positrons,
2007), as tabulated in the reference's pn_solver/stopping_power.jl:51-154
(energies `E_sp` in MeV, `data_tot_sp[Z]` in MeV/cm).  Only the numbers are
taken (they are physical data); this script reads that file as text and writes
`pnepma_amd/data/icru77_tot_sp.json`.  It is run once, in a checkout that has
the reference mounted, and its output is committed: it is the provenance of
that table, not part of the package.

    python tests/golden/make_icru77.py /root/reference/pn_solver/stopping_power.jl
"""
import json
import os
import re
import sys

ELEMENTS = (13, 14, 24, 26, 28, 29)  # Al, Si, Cr, Fe, Ni, Cu


def _numbers(body):
    return [float(v) for v in re.findall(r"[-+]?\d+\.?\d*(?:e[-+]?\d+)?", body)]


def main(path):
    text = open(path).read()
    e_body = re.search(r"E_sp\s*=\s*\[(.*?)\]u\"MeV\"", text, re.S).group(1)
    energies = _numbers(e_body)
    tot = text[text.index("data_tot_sp"):text.index("data_col_sp")]
    rows = {}
    for z in ELEMENTS:
        m = re.search(r"\n\s*%d\s*=>\s*\[(.*?)\]" % z, tot, re.S)
        vals = _numbers(m.group(1))
        assert len(vals) == len(energies), (z, len(vals), len(energies))
        rows[str(z)] = vals
    out = {
        "source": "ICRU Report 77 total stopping power, MeV/cm, as tabulated in pn_solver/stopping_power.jl",
        "energy_MeV": energies,
        "tot_sp_MeV_per_cm": rows,
    }
    root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    dst = os.path.join(root, "pnepma_amd", "data", "icru77_tot_sp.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", dst, len(energies), "energies,", len(rows), "elements")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/pn_solver/stopping_power.jl")
