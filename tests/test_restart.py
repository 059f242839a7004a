"""Restart files (include/fjsph_restart.h) on the CPU: the host-level writer
and reader, checked with an independent HDF5 reader (h5dump of the image's
HDF5 1.10 tools) against the layout the reference's writer declares
(tests/golden/restart_layout.json, extracted from src/H5IO.cpp and
src/Var.h by tests/golden/make_restart_layout.py), and by round trips."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from fjsph_amd import abi

H5DUMP = "/opt/conda/bin/h5dump"
LAYOUT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "restart_layout.json")))

restart = pytest.importorskip("fjsph_amd.restart")


@pytest.fixture(scope="module")
def lib():
    try:
        restart.load()
    except Exception as e:  # not built in this checkout
        pytest.skip(str(e))
    return restart


def _planes(dim, **kw):
    return abi.make_planes(**kw)


def image(dim, seed=0, n_cols=4):
    """A synthetic state: a moving boundary block, a fluid block with a
    deletion plane and an inlet block with n_cols columns."""
    rng = np.random.default_rng(seed)
    nb, nf = 12, 28
    ni = n_cols * (1 + restart.NBUF) + 6
    n = nb + nf + ni
    up = (0, 0, 1) if dim == 3 else (0, 1)
    blocks = list(abi.make_blocks([
        dict(first=0, last=nb, is_fluid=0, bound_solver=abi.PRESSURE_G, no_slip=1),
        dict(first=nb, last=nb + nf, is_fluid=1, vel=(0.25, -0.5, 0.0)),
        dict(first=nb + nf, last=n, is_fluid=1)]))
    blocks[0].n_times = 3
    planes = [abi.make_planes(), abi.make_planes(delete=(up, 0.3)),
              abi.make_planes(inlet=1, fixed_vel_or_dynamic=1, insert=(up, 0.01), aero=(up, 0.2), delete=(up, 0.4))]
    clock = abi.Clock(current_time=0.0123, last_frame_time=0.012, delta_t=1e-6, cfl=0.85, current_frame=7,
                      n_stable=3, n_unstable=1, next_part_id=n + 17)
    s = abi.default_settings(dim)
    s.particle_step = 5e-5
    s.frame_time_interval = 1e-3
    first_in = nb + nf
    back = np.arange(first_in, first_in + n_cols)
    buf = np.arange(first_in + n_cols, first_in + n_cols * (1 + restart.NBUF))
    img = dict(dim=dim, n=n, n_bound=nb, clock=clock, settings=s, max_points=123456, blocks=blocks, planes=planes,
               times=[np.array([0.0, 0.1, 0.25]), np.zeros(0), np.zeros(0)],
               vels=[rng.standard_normal(9), np.zeros(0), np.zeros(0)],
               back=[None, None, back], buffer=[None, None, buf])
    for f in restart.FIELDS_D:
        img[f] = rng.standard_normal((n, dim))
    for f in restart.FIELDS_1:
        img[f] = rng.standard_normal(n)
    img["b"] = rng.integers(0, 6, n).astype(np.int32)
    img["pid"] = (rng.permutation(n) + 5).astype(np.int64)
    img["cell_id"] = rng.integers(-3, 10 ** 9, n).astype(np.int64)
    return img


STRINGS = ["input_file", "fluid_file", "bound_file", "tau_mesh", "tau_bmap", "tau_sol", "foam_dir", "foam_sol",
           "vlm_file", "output_prefix", "restart_prefix", "output_names"]


def full_meta(dim):
    """Every string set, so that every attribute the reference writes is in
    the file (empty strings are skipped, src/H5IO.cpp:110)."""
    return restart.default_meta(dim, **{k: f"{k}.txt" for k in STRINGS})


def h5_layout(path):
    """{(group path, kind, name): {"type", "space", "strsize", "strpad"}}
    from the header dump of h5dump."""
    txt = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    stack, objs = [], {}
    for line in txt.splitlines():
        s = line.strip()
        m = re.match(r'(GROUP|ATTRIBUTE|DATASET) "(.*)" \{$', s)
        if m:
            stack.append({"kind": m.group(1), "name": m.group(2)})
            continue
        if s.startswith("HDF5 ") and s.endswith("{"):
            stack.append({"kind": "FILE", "name": ""})
            continue
        m = re.match(r"DATATYPE\s+(\S+)( \{)?$", s)
        if m:
            stack[-1]["type"] = m.group(1)
            if m.group(2):
                stack.append({"kind": "TYPE", "owner": stack[-1]})
            continue
        m = re.match(r"STRSIZE (\d+);", s)
        if m:
            stack[-1]["owner"]["strsize"] = int(m.group(1))
            continue
        m = re.match(r"STRPAD (\S+);", s)
        if m:
            stack[-1]["owner"]["strpad"] = m.group(1)
            continue
        m = re.match(r"DATASPACE\s+(SCALAR|SIMPLE \{ \( ([\d, ]+) \))", s)
        if m:
            stack[-1]["space"] = "scalar" if m.group(1) == "SCALAR" else tuple(int(v) for v in m.group(2).split(","))
            continue
        if s == "}":
            obj = stack.pop()
            if obj["kind"] in ("ATTRIBUTE", "DATASET"):
                gpath = "/".join(o["name"] for o in stack if o["kind"] == "GROUP" and o["name"] != "/")
                objs[(gpath, obj["kind"], obj["name"])] = obj
    return objs


# HDF5 file types of the reference's writers (src/H5IO.cpp): Write_Real_Attribute
# and vectors H5T_IEEE_F64LE (:83-104); Write_Uint_Attribute NATIVE_ULONG for
# size_t, NATIVE_UINT for uint (:68-76); Write_Int_Attribute NATIVE_UINT (:78-81);
# datasets NATIVE_DOUBLE / UINT / ULONG / LONG after their element type.
ATTR_TYPE = {"real": "H5T_IEEE_F64LE", "vector": "H5T_IEEE_F64LE", "int": "H5T_STD_U32LE",
             ("uint", "size_t"): "H5T_STD_U64LE", ("uint", "uint"): "H5T_STD_U32LE", "string": "H5T_STRING"}
DSET_TYPE = {"real": "H5T_IEEE_F64LE", "uint": "H5T_STD_U32LE", "size_t": "H5T_STD_U64LE", "long": "H5T_STD_I64LE"}


def _attr_type(e):
    return ATTR_TYPE[(e["kind"], e["ctype"])] if e["kind"] == "uint" else ATTR_TYPE[e["kind"]]


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not installed")
@pytest.mark.parametrize("dim", [2, 3])
def test_layout_matches_reference_writer(lib, tmp_path, dim):
    img = image(dim)
    path = str(tmp_path / "run_particles.h5")
    meta = full_meta(dim)
    names = (restart.C.c_char_p * 3)(b"walls", b"drop", b"jet")
    meta.block_names = names
    lib.write_host(path, img, meta)
    objs = h5_layout(path)

    # root attributes: exactly the reference's set, each with its type
    root = {k[2]: v for k, v in objs.items() if k[0] == "" and k[1] == "ATTRIBUTE"}
    want = {e["name"]: e for e in LAYOUT["root_attributes"]}
    assert set(root) == set(want)
    for name, e in want.items():
        assert root[name]["type"] == _attr_type(e), name
        if e["kind"] == "vector":
            assert root[name]["space"] == (dim,), name
        else:
            assert root[name]["space"] == "scalar", name
        if e["kind"] == "string":
            assert root[name]["strpad"] == "H5T_STR_SPACEPAD"

    groups = ["Boundary 0", "Fluid 0", "Fluid 1"]
    assert {k[0] for k in objs} == {""} | set(groups)
    axes = "xyz"[:dim]
    for k, g in enumerate(groups):
        bl = img["blocks"][k]
        length = bl.last - bl.first
        dsets = {key[2]: v for key, v in objs.items() if key[0] == g and key[1] == "DATASET"}
        want = {}
        for e in LAYOUT["zone_datasets"]:
            for nm in ([f"{e['name']} {a}" for a in axes] if e["vector"] else [e["name"]]):
                want[nm] = (DSET_TYPE[e["ctype"]], (length,))
        if k == 2:
            ncol = len(img["back"][2])
            want["Back index"] = (DSET_TYPE["size_t"], (ncol,))
            want["Buffer index"] = (DSET_TYPE["size_t"], (restart.NBUF * ncol,))
        assert {nm: (v["type"], v["space"]) for nm, v in dsets.items()} == want, g

        attrs = {key[2]: v for key, v in objs.items() if key[0] == g and key[1] == "ATTRIBUTE"}
        moving = len(img["times"][k]) > 0
        want = set()
        for e in LAYOUT["zone_attributes"]:
            if e["dim3"] and dim != 3:
                continue
            if e["branch"] == ("fixed" if moving else "moving"):
                continue
            want.add(e["name"])
        if moving:
            for t in range(len(img["times"][k])):
                want.add(f"Time {t}")
                want |= {f"Velocity {t} {a}" for a in axes}
        assert set(attrs) == want, g
        for nm, v in attrs.items():
            assert v["type"] == ("H5T_STRING" if nm == "Block name" else "H5T_IEEE_F64LE"), nm


def _h5_values(path, obj, attr=True):
    """Values of one attribute or dataset, printed by h5dump."""
    txt = subprocess.run([H5DUMP, "-a" if attr else "-d", obj, path], capture_output=True, text=True,
                         check=True).stdout
    data = txt[txt.index("DATA {"):]
    return [float(v) for v in re.findall(r"(?:\(\d+\):\s*|,\s*)(-?[\d.]+(?:e[-+]?\d+)?)", data)]


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not installed")
def test_values_read_by_h5dump(lib, tmp_path):
    """The values land where the reference's reader looks for them."""
    img = image(3)
    path = str(tmp_path / "v_particles.h5")
    meta = full_meta(3)
    meta.scale = 0.5
    lib.write_host(path, img, meta)
    assert _h5_values(path, "/Simulation current time") == [0.0123]
    assert _h5_values(path, "/Particle index to add") == [img["clock"].next_part_id]
    assert _h5_values(path, "/Current frame") == [7]
    assert _h5_values(path, "/SPH CFL condition") == [0.85]
    assert _h5_values(path, "/Number of fluid points") == [img["n"] - img["n_bound"]]
    assert _h5_values(path, "/SPH maximum particle count") == [123456]
    f0 = img["blocks"][1]
    pid = _h5_values(path, "/Fluid 0/Particle ID", attr=False)
    assert pid == list(img["pid"][f0.first:f0.last].astype(float))
    # positions are written divided by the grid scale (src/H5IO.cpp:406)
    zx = np.array(_h5_values(path, "/Fluid 0/Position coordinate x", attr=False))
    # (h5dump prints 6 significant digits)
    assert np.allclose(zx, img["x"][f0.first:f0.last, 0] / 0.5, rtol=1e-5, atol=1e-5)
    assert _h5_values(path, "/Fluid 1/Back index", attr=False) == list(img["back"][2].astype(float))
    assert _h5_values(path, "/Boundary 0/Times count") == [3.0]
    assert _h5_values(path, "/Fluid 1/Block type") == [6.0]


@pytest.mark.parametrize("dim", [2, 3])
def test_host_round_trip(lib, tmp_path, dim):
    img = image(dim, seed=dim)
    path = str(tmp_path / "rt_particles.h5")
    meta = restart.default_meta(dim)
    meta.scale = 0.25  # a power of two: x / scale * scale is exact
    lib.write_host(path, img, meta)
    h = lib.read_header(path)
    assert (h.dim, h.n_bound_blocks, h.n_fluid_blocks) == (dim, 1, 2)
    assert (h.bound_points, h.fluid_points) == (img["n_bound"], img["n"] - img["n_bound"])
    assert h.particle_step == 5e-5 and h.frame_time_interval == 1e-3 and h.scale == 0.25
    r = lib.read_host(path)
    for f in restart.FIELDS_D + restart.FIELDS_1 + ("b", "pid", "cell_id"):
        assert np.array_equal(np.asarray(r[f]).reshape(np.shape(img[f])), img[f]), f
    c = r["clock"]
    # Read_HDF5: last_frame_time = current_time (src/H5IO.cpp:1248); no timestep
    assert (c.current_time, c.last_frame_time, c.delta_t, c.cfl) == (0.0123, 0.0123, 0.0, 0.85)
    assert (c.current_frame, c.n_stable, c.n_unstable, c.next_part_id) == (7, 3, 1, img["clock"].next_part_id)
    assert r["max_points"] == 123456
    for k, (a, b) in enumerate(zip(r["blocks"], img["blocks"])):
        assert (a.first, a.last, a.is_fluid, a.bound_solver, a.no_slip, a.n_times) == \
               (b.first, b.last, b.is_fluid, b.bound_solver, b.no_slip, b.n_times), k
        if b.n_times == 0:
            assert list(a.vel)[:dim] == list(b.vel)[:dim]
        assert np.array_equal(r["times"][k], img["times"][k])
        nt = len(img["times"][k])
        want_v = np.asarray(img["vels"][k]).reshape(nt, 3).copy()
        want_v[:, dim:] = 0.0  # only dim components are in the file
        assert np.array_equal(np.asarray(r["vels"][k]).reshape(nt, 3), want_v)
    for k, (a, b) in enumerate(zip(r["planes"], img["planes"])):
        assert (a.inlet, a.fixed_vel_or_dynamic) == (b.inlet, b.fixed_vel_or_dynamic)
        for f in ("delete_norm", "insert_norm", "aero_norm"):
            assert list(getattr(a, f))[:dim] == list(getattr(b, f))[:dim], (k, f)
        assert (a.delconst, a.insconst, a.aeroconst) == (b.delconst, b.insconst, b.aeroconst)
    assert r["back"][:2] == [None, None] and r["buffer"][:2] == [None, None]
    assert np.array_equal(r["back"][2], img["back"][2]) and np.array_equal(r["buffer"][2], img["buffer"][2])


def test_rejects_what_the_layout_cannot_hold(lib, tmp_path):
    path = str(tmp_path / "bad_particles.h5")
    img = image(3)
    img["buffer"][2] = img["buffer"][2][:-1]  # not 4 per column
    with pytest.raises(restart.RestartError, match="4 buffer"):
        lib.write_host(path, img, None)
    img = image(3)
    img["blocks"][1].last -= 1  # blocks no longer tile [0, n)
    with pytest.raises(restart.RestartError, match="tile"):
        lib.write_host(path, img, None)
    with pytest.raises(restart.RestartError):
        lib.read_header(str(tmp_path / "missing_particles.h5"))


RESTART_HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                              "fjsph_restart.h")


def test_restart_library_exports_every_declared_symbol(lib):
    text = re.sub(r"/\*.*?\*/", "", open(RESTART_HEADER).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(fjsph_restart_[a-z0-9_]+)\s*\(", text)))
    assert len(names) == 8
    so = lib.load()
    assert not [n for n in names if not hasattr(so, n)]
    assert set(names) == set(restart.SIGNATURES)


def test_restart_struct_layouts_match_c(lib, tmp_path):
    import textwrap
    src = tmp_path / "sizes.c"
    src.write_text(textwrap.dedent(f"""
        #include <stdio.h>
        #include <stddef.h>
        #include "{RESTART_HEADER}"
        int main(void) {{
            printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(fjsph_restart_meta), sizeof(fjsph_restart_header),
                   sizeof(fjsph_restart_host), sizeof(fjsph_clock), offsetof(fjsph_restart_meta, offset_axis),
                   offsetof(fjsph_restart_host, cell_rho), offsetof(fjsph_restart_header, single_file));
            return 0;
        }}
    """))
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    C = restart.C
    assert got == [C.sizeof(restart.Meta), C.sizeof(restart.Header), C.sizeof(restart.Host), C.sizeof(abi.Clock),
                   restart.Meta.offset_axis.offset, restart.Host.cell_rho.offset, restart.Header.single_file.offset]
