"""Writes regression_cases.json: the reference's image-regression cases for the
formats this library parses (OFF, PLY, OBJ, NFF), as data for
raptor_amd.regression.

Each case is read from the reference's test sources as text
(rt/regression_tests/{off,ply,obj,nff}_tests.cc): the BOOST_AUTO_TEST_CASE name
(which names the expected image test_data/<suite>/<name>_exp.png), the
regression_fixture arguments (model path under $RAPTOR_DATA, format, camera
position and axes, background colour, pan/tilt/roll, resolution) and the
add_light calls (rgb, centre, distance falloff, radius); an NFF case takes
its camera, background and lights from the file. Fixture tooling, not part
of the library: run it where the reference checkout is, from the repository
root:

    python tests/golden/make_regression_cases.py /path/to/RAPTor
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(os.path.dirname(os.path.dirname(HERE)), "raptor_amd", "regression_cases.json")
SUITES = ("off", "ply", "obj", "nff")


def split_args(s):
    """Top-level comma-separated arguments of a call's argument text."""
    out, depth, cur = [], 0, []
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    if "".join(cur).strip():
        out.append("".join(cur).strip())
    return out


def call_args(text, start):
    """The argument text of the call whose '(' is at text[start]."""
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "(":
            depth += 1
        elif text[i] == ")":
            depth -= 1
            if depth == 0:
                return text[start + 1:i]
    raise ValueError("unbalanced call")


def vec(arg):
    m = re.fullmatch(r"(?:point_t<>|ext_colour_t)\s*\((.*)\)", arg.strip(), re.S)
    if not m:
        raise ValueError(f"not a vector: {arg}")
    v = [float(x) for x in split_args(m.group(1))]
    return v * 3 if len(v) == 1 else v


def parse_suite(path, suite):
    text = open(path).read()
    cases = []
    for m in re.finditer(r"BOOST_AUTO_TEST_CASE\(\s*(\w+)\s*\)", text):
        name = m.group(1)
        end = text.find("BOOST_AUTO_TEST_CASE(", m.end())
        body = text[m.end():end if end > 0 else len(text)]
        f = re.search(r"regression_fixture\s+fixture\s*\(", body)
        if not f:
            continue
        a = split_args(call_args(body, f.end() - 1))
        fmt = a[2].split("::")[-1]
        case = dict(suite=suite, name=name, input=a[1].strip('"'), format=fmt, camera_from_file=(fmt == "nff"),
                    camera=dict(pos=[0.0, 0.0, -10.0], x=[1.0, 0.0, 0.0], y=[0.0, 1.0, 0.0], z=[0.0, 0.0, 1.0]),
                    background=[0.0, 0.0, 0.0], rotation=[0.0, 0.0, 0.0], resolution=[640, 480], lights=[],
                    unsupported=[])
        keys = ["pos", "x", "y", "z"]
        for k, arg in zip(keys, a[3:7]):
            case["camera"][k] = vec(arg)
        if len(a) > 7:
            case["background"] = vec(a[7])
        if len(a) > 10:
            case["rotation"] = [float(x) for x in a[8:11]]
        if len(a) > 12:
            case["resolution"] = [int(a[11]), int(a[12])]
        if len(a) > 13:
            case["unsupported"].append("anti-aliasing samples")
        for lm in re.finditer(r"\.add_(light|spotlight|directional_light)\s*\(", body):
            la = split_args(call_args(body, lm.end() - 1))
            if lm.group(1) != "light":
                case["unsupported"].append(lm.group(1).replace("_", " "))
                continue
            case["lights"].append(dict(rgb=vec(la[0]), pos=vec(la[1]), d=float(la[2]), r=float(la[3])))
        if any(case["rotation"]):
            case["unsupported"].append("camera pan/tilt/roll")
        cases.append(case)
    return cases


def main(ref_root):
    rt_dir = os.path.join(ref_root, "raytracer", "regression_tests")
    if not os.path.isdir(rt_dir):
        rt_dir = os.path.join(ref_root, "RAPTor", "raytracer", "regression_tests")
    cases = []
    for suite in SUITES:
        cases += parse_suite(os.path.join(rt_dir, f"{suite}_tests.cc"), suite)
    with open(OUT, "w") as fh:
        json.dump(dict(source="rt/regression_tests/{off,ply,obj,nff}_tests.cc", cases=cases), fh, indent=1)
        fh.write("\n")
    counts = {s: sum(c["suite"] == s for c in cases) for s in SUITES}
    print(f"{OUT}: {len(cases)} cases {counts}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/RAPTor")
