"""Write tests/golden/validation_profiles.json from the analytic profiles the
reference ships with its validation examples (run in the build container,
where the reference is at /root/reference; the GPU tests read only the JSON):

  Examples/Couette/SteadyState.dat      Couette start-up: u(z) at t = 1, 5, 15,
                                        50 s and the steady line (Z, V)
  Examples/Poiseuille/SteadyState.dat   Poiseuille start-up under a body force:
                                        u(z) at t = 0.02, 0.1, 0.2, 0.5 s and
                                        the steady parabola (Z, V)
  Examples/Standing_Column/Ideal.dat    hydrostatic pressure P / (rho g H)
                                        against (H - Y) / H

Each file is Tecplot ASCII: a VARIABLES line, then ZONEs of numbers; zones
with VARSHARELIST share the first variable (Z) of the first zone and list
only the second.
"""
import json
import os
import re
import sys

REF = os.environ.get("FJSPH_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "validation_profiles.json")


def read_tecplot(path):
    zones, cur = [], None
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("TITLE") or s.startswith("VARIABLES"):
                continue
            if s.startswith("ZONE"):
                m = re.search(r'T="([^"]*)"', s)
                cur = {"title": m.group(1) if m else "", "shared": "VARSHARELIST" in s, "values": []}
                zones.append(cur)
                continue
            cur["values"].append([float(v) for v in s.split()])
    first = zones[0]
    z = [row[0] for row in first["values"]]
    out = {first["title"]: {"z": z, "v": [row[1] for row in first["values"]]}}
    for zone in zones[1:]:
        vals = [row[0] for row in zone["values"]] if zone["shared"] else [row[1] for row in zone["values"]]
        zz = z if zone["shared"] else [row[0] for row in zone["values"]]
        out[zone["title"]] = {"z": zz, "v": vals}
    return out


def main():
    data = {
        "source": "analytic profiles of the reference's validation examples (Tecplot files under Examples/)",
        "couette": read_tecplot(os.path.join(REF, "Examples/Couette/SteadyState.dat")),
        "poiseuille": read_tecplot(os.path.join(REF, "Examples/Poiseuille/SteadyState.dat")),
        "standing_column": read_tecplot(os.path.join(REF, "Examples/Standing_Column/Ideal.dat")),
    }
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT, {k: list(v) for k, v in data.items() if isinstance(v, dict)})


if __name__ == "__main__":
    sys.exit(main())
