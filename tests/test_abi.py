"""The C ABI without a GPU: the library loads, exports what include/raptor_rt.h
declares, the C++ shim compiles, and the host-side generators behave."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "raptor_rt.h")).read()
    return sorted(set(re.findall(r"\b(rt_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from raptor_amd import rt

    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(rt.lib(), n)]
    assert not missing, missing
    assert sorted(rt.EXPORTS) == names


def test_cpp_shim_compiles():
    src = os.path.join(ROOT, "tests", "cpp", "shim_demo.cpp")
    res = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "include"), src],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_header_is_plain_c():
    res = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-Wall", "-x", "c",
                          os.path.join(ROOT, "include", "raptor_rt.h")], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_random_triangles_deterministic_and_bounded():
    from raptor_amd import rt

    a1, b1 = rt.gen_random_triangles(1, 1000, 0.01, 0.0, 1.0)
    a2, b2 = rt.gen_random_triangles(1, 1000, 0.01, 0.0, 1.0)
    a3, _ = rt.gen_random_triangles(2, 1000, 0.01, 0.0, 1.0)
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2) and not np.array_equal(a1, a3)
    v = a1.reshape(-1, 3, 3)
    c = v.mean(1)
    assert np.all(np.abs(v - c[:, None, :]) <= 0.02 + 1e-6)
    assert a1.min() >= -0.01 - 1e-6 and a1.max() <= 1.01 + 1e-6
    assert b1.min() >= 0.0 and b1.max() < 1.0


def test_terrain_layout():
    # RAPTor/terrain_generator/height_map.h:16-43: vertex (i, j) at (i*xs, h, j*ys);
    # two triangles per cell; |h| <= a * sum(p^k) because |noise| <= 1.
    from raptor_amd import rt

    g = 33
    abc, alb = rt.gen_terrain(g, 0.15, 0.02, 0.5, 8, 57, 2, 0.2, 0.8)
    assert abc.shape == (2 * (g - 1) ** 2, 9)
    xs = np.float32(1.0) / np.float32(g - 1)
    v = abc.reshape(-1, 3, 3)
    ii = np.rint(v[..., 0] / xs).astype(int)
    jj = np.rint(v[..., 2] / xs).astype(int)
    assert np.array_equal(v[..., 0], (ii.astype(np.float32) * xs).astype(np.float32))
    # First cell: (r+j, r+j+1, r'+j+1) and (r'+j, r+j, r'+j+1) with r = 0, r' = g.
    assert (ii[0] == [0, 0, 1]).all() and (jj[0] == [0, 1, 1]).all()
    assert (ii[1] == [1, 0, 1]).all() and (jj[1] == [0, 0, 1]).all()
    bound = 0.15 * sum(0.5 ** k for k in range(8))
    assert np.abs(v[..., 1]).max() <= bound + 1e-6
    # Shared vertices carry the same height in every triangle that uses them.
    heights = {}
    for t in range(v.shape[0]):
        for k in range(3):
            key = (ii[t, k], jj[t, k])
            heights.setdefault(key, v[t, k, 1])
            assert heights[key] == v[t, k, 1]
    assert len(heights) == g * g
    assert alb.min() >= 0.2 and alb.max() < 0.8


def test_bad_arguments_are_rejected_without_gpu():
    from raptor_amd import rt

    assert rt.lib().rt_gen_terrain(1, 0.1, 0.1, 0.5, 2, 1, 1, 0.0, 1.0, None, None) == -1
    assert rt.lib().rt_gen_random_triangles(1, -5, 0.01, 0.0, 1.0, None, None) == -1
    assert rt.lib().rt_ctx_create(0, None) == rt.RT_ERR_INVALID


def _c_layout(struct, fields):
    """sizeof and offsetof of a header struct, from the C compiler."""
    import subprocess
    import tempfile

    src = "#include <stddef.h>\n#include <stdio.h>\n#include \"raptor_rt.h\"\nint main(void) {\n"
    src += f'printf("%zu\\n", sizeof({struct}));\n'
    for f in fields:
        src += f'printf("%zu\\n", offsetof({struct}, {f}));\n'
    src += "return 0;\n}\n"
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        with open(c, "w") as fh:
            fh.write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    return out[0], dict(zip(fields, out[1:]))


@pytest.mark.parametrize("struct,dtype_name", [("rt_material", "MATERIAL_DTYPE"),
                                               ("rt_texture_map", "TEXTURE_MAP_DTYPE"),
                                               ("rt_texture_image", "TEXTURE_IMAGE_DTYPE")])
def test_python_rows_match_the_c_structs(struct, dtype_name):
    """The numpy rows the bindings pass have the header's size and offsets."""
    from raptor_amd import rt

    dt = getattr(rt, dtype_name)
    size, offs = _c_layout(struct, list(dt.names))
    assert dt.itemsize == size
    assert {k: dt.fields[k][1] for k in dt.names} == offs


def test_mesh_info_matches_the_c_struct():
    from raptor_amd import rt

    names = [k for k, _ in rt.rt_mesh_info._fields_]
    size, offs = _c_layout("rt_mesh_info", names)
    assert C.sizeof(rt.rt_mesh_info) == size
    assert {k: getattr(rt.rt_mesh_info, k).offset for k in names} == offs


def test_batch_params_match_the_c_struct():
    from raptor_amd import rt

    names = [k for k, _ in rt.rt_batch_params._fields_]
    size, offs = _c_layout("rt_batch_params", names)
    assert C.sizeof(rt.rt_batch_params) == size
    assert {k: getattr(rt.rt_batch_params, k).offset for k in names} == offs


def test_material_table_fields():
    from raptor_amd import rt

    t = rt.material_table([dict(kd=(1, 2, 3), rf=0.5), dict(kind=1, x=0.2, rs=0.4, ts=0.3, ri_r=1.2)])
    assert t["reflectance"][0] == np.float32(0.5) and t["kind"][1] == 1
    assert t["transmittance"][1] == np.float32(0.3) and t["refractive_index"][1] == np.float32(1.2)
