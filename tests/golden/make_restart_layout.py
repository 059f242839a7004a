"""Write tests/golden/restart_layout.json: the restart file layout as the
reference's writer declares it, read out of its source text.

Sources (read as text, nothing is built or run):
  src/H5IO.cpp  Write_HDF5_Attributes (root attributes), Write_Zone_Data
                (per-block datasets), Write_Inlet_Data, Write_Zone_Attributes
  src/Var.h     declarations giving the C type of each unsigned attribute
                (size_t -> u64, uint -> u32)

Usage: python tests/golden/make_restart_layout.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "restart_layout.json")


def body(src, signature, last=False):
    """Text of the function whose definition starts with `signature`: the
    first occurrence, or the last one where a forward declaration comes
    first.  (H5IO.cpp defines Write_Zone_Data twice: the restart writer in
    namespace HDF5 first, the h5part output writer later.)"""
    i = src.rindex(signature) if last else src.index(signature)
    j = src.index("{", i)
    depth, k = 0, j
    while True:
        if src[k] == "{":
            depth += 1
        elif src[k] == "}":
            depth -= 1
            if depth == 0:
                return src[j:k + 1]
        k += 1


def c_type_of(expr, var_h):
    """Declared type of the member named by the last identifier of expr."""
    ident = re.findall(r"[A-Za-z_]\w*", expr)[-1]
    m = re.search(r"\b(size_t|uint|unsigned|int|bool|real|double)\s+" + ident + r"\b", var_h)
    return m.group(1) if m else None


def main():
    h5 = open(os.path.join(REF, "src", "H5IO.cpp")).read()
    var_h = open(os.path.join(REF, "src", "Var.h")).read()
    call = re.compile(r'HDF5::Write_(Real|Uint|Int|String|Vector)_Attribute\(\s*(\w+),\s*"([^"]+)",\s*(.*?)\s*\);',
                      re.S)

    root = []
    for kind, _, name, expr in call.findall(body(h5, "void Write_HDF5_Attributes(", last=True)):
        e = {"name": name, "kind": kind.lower()}
        if kind == "Uint":
            e["ctype"] = c_type_of(expr, var_h)
        root.append(e)

    # block attributes: the literal names, tagged with the #if SIMDIM == 3
    # sections and the nTimes branch they sit in
    zone = []
    dim3 = False
    branch = "any"
    for line in body(h5, "void Write_Zone_Attributes(").splitlines():
        s = line.strip()
        if s.startswith("#if SIMDIM == 3"):
            dim3 = True
        elif s.startswith("#endif"):
            dim3 = False
        elif s.startswith("if (limits.nTimes != 0)"):
            branch = "moving"
        elif s == "else" and branch == "moving":
            branch = "fixed"
        elif s.startswith('HDF5::Write_Real_Attribute(zone, "Fixed velocity'):
            branch = "any"
        for kind, name in re.findall(r'HDF5::Write_(\w+)_Attribute\(\s*zone,\s*"([^"]+)"\s*,', s):
            zone.append({"name": name, "kind": kind.lower(), "dim3": dim3, "branch": branch})
    # per-time names built from a counter ("Time k", "Velocity k x")
    zone_timed = sorted(set(re.findall(r'"(Time |Velocity )"\s*\+\s*std::to_string\(time\)',
                                       body(h5, "void Write_Zone_Attributes("))))

    zd = body(h5, "void Write_Zone_Data(")
    decl = dict((v, t) for t, v in re.findall(r"vector<(\w+)>\s+(\w+)\(length\)", zd))
    datasets = []
    for base, var in re.findall(r'string name = "([^"]+) " \+ get_dim\(dim\);\s*Write_Variable_Scalar\(fout, name, '
                                r'dims, 0, (\w+)\)', zd):
        datasets.append({"name": base, "vector": True, "ctype": decl[var]})
    for name, var in re.findall(r'Write_Variable_Scalar\(fout, "([^"]+)", dims, 0, (\w+)\)', zd):
        datasets.append({"name": name, "vector": False, "ctype": decl[var]})

    inl = body(h5, "void Write_Inlet_Data(")
    inlet = [{"name": n} for n in re.findall(r'Write_Variable_Scalar\(fout, "([^"]+)"', inl)]
    # both hold size_t (bound_block::back, vector<size_t> buff)
    for e in inlet:
        e["ctype"] = "size_t"

    out = {"source": "src/H5IO.cpp Write_HDF5_Attributes, Write_Zone_Data, Write_Inlet_Data, "
                     "Write_Zone_Attributes; src/Var.h declarations",
           "root_attributes": root, "zone_attributes": zone, "zone_timed_prefixes": zone_timed,
           "zone_datasets": datasets, "inlet_datasets": inlet}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"{OUT}: {len(root)} root attributes, {len(zone)} block attributes, {len(datasets)} datasets")


if __name__ == "__main__":
    main()
