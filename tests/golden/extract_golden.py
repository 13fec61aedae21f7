"""Extract the golden vectors hard-coded in FastMultipole's test files.

Run in the build container (where the reference is mounted) as

    python tests/golden/extract_golden.py /root/reference

It reads the reference test sources as TEXT, pulls out only the numeric
literals the tests compare against (inputs and expected outputs), and writes
them to tests/golden/reference_vectors.json. No reference code is executed and
no source text is kept: the fixture holds numbers only, keyed by the test file
and line they come from.
"""
import json
import os
import re
import sys

NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
CPLX_SPACED = re.compile(rf"^\s*({NUM})\s*([-+])\s*({NUM})im\s*$")
CPLX_TIGHT = re.compile(rf"^({NUM})([-+])({NUM})im$")


def parse_complex(tok):
    tok = tok.strip()
    m = CPLX_SPACED.match(tok) or CPLX_TIGHT.match(tok)
    if m:
        re_, sign, im = m.groups()
        v = float(im)
        return float(re_), (-v if sign == "-" else v)
    return float(tok), 0.0


def bracket_content(text, start):
    """Content of the bracket opening at or after `start` (handles nesting)."""
    i = text.index("[", start)
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "[":
            depth += 1
        elif text[j] == "]":
            depth -= 1
            if depth == 0:
                return text[i + 1:j], j
    raise ValueError("unbalanced bracket")


def find_assignment(text, name, occurrence=1, after=0):
    pat = re.compile(rf"(?m)^\s*{re.escape(name)}\s*=\s*")
    pos = after
    m = None
    for _ in range(occurrence):
        m = pat.search(text, pos)
        if m is None:
            raise KeyError(f"{name} (occurrence {occurrence}) not found")
        pos = m.end()
    return m


def complex_list(text, name, occurrence=1, after=0):
    m = find_assignment(text, name, occurrence, after)
    head = text[m.end():m.end() + 20]
    neg = head.lstrip().startswith("-")
    content, _ = bracket_content(text, m.end())
    vals = [parse_complex(t) for t in content.split(",") if t.strip()]
    if neg:
        vals = [(-a, -b) for a, b in vals]
    return {"re": [v[0] for v in vals], "im": [v[1] for v in vals]}


def real_column(text, name, occurrence=1, after=0):
    m = find_assignment(text, name, occurrence, after)
    content, _ = bracket_content(text, m.end())
    return [float(t) for t in content.replace(";", " ").split()]


def matrix(text, name, occurrence=1, after=0):
    m = find_assignment(text, name, occurrence, after)
    content, _ = bracket_content(text, m.end())
    rows = [r for r in re.split(r"[;\n]", content) if r.strip()]
    out_re, out_im = [], []
    for r in rows:
        vals = [parse_complex(t) for t in r.split()]
        out_re.append([v[0] for v in vals])
        out_im.append([v[1] for v in vals])
    return {"re": out_re, "im": out_im}


def array3_slices(text, name, occurrence=1, after=0):
    """vec([a b; c d;;; ...]) -> list of 2x2 slices [[a, b], [c, d]]."""
    m = find_assignment(text, name, occurrence, after)
    start = text.index("vec(", m.end())
    content, _ = bracket_content(text, start + 4)
    slices = []
    for s in content.split(";;;"):
        rows = [r for r in s.split(";") if r.strip()]
        slices.append([[float(t) for t in r.split()] for r in rows])
    return slices


def scalar(text, name, occurrence=1, after=0):
    m = find_assignment(text, name, occurrence, after)
    line = text[m.end():text.index("\n", m.end())]
    return float(re.match(NUM, line.strip()).group(0))


def line_of(text, pos):
    return text.count("\n", 0, pos) + 1


def main(ref):
    T = lambda f: open(os.path.join(ref, "test", f), encoding="utf-8").read()
    out = {"_about": "numeric literals extracted from FastMultipole test files; keys name file:testset"}

    # --- rotate_test.jl
    t = T("rotate_test.jl")
    out["rotate_test.jl:H_pi2"] = {f"H{n}": matrix(t, f"H{n}")["re"] for n in range(1, 5)}
    out["rotate_test.jl:T_3pi7"] = {"beta_num": 3, "beta_den": 7, **{f"T{n}": matrix(t, f"T{n}")["re"] for n in range(1, 5)}}
    ys = find_assignment(t, "θ", 1, t.index('"multipole rotation about y axis"')).start()
    out["rotate_test.jl:multipole_y"] = {
        "theta": scalar(t, "θ", 1, ys), "p": 10,
        "in_re": real_column(t, "unrotated_weights_real", 1, ys),
        "in_im": real_column(t, "unrotated_weights_imag", 1, ys),
        "out_full": complex_list(t, "rotated_weights_test", 1, ys)}
    zs = t.index('"multipole rotation about z axis"')
    out["rotate_test.jl:multipole_z"] = {
        "phi": scalar(t, "ϕ", 1, zs), "p": 10,
        "in_full": complex_list(t, "unrotated_weights_test", 1, zs),
        "out_full": complex_list(t, "rotated_weights_test", 1, zs)}
    ls = t.index('"local rotation about y axis"')
    out["rotate_test.jl:local_y"] = {
        "theta": scalar(t, "θ", 1, ls), "p": 10,
        "in_full": complex_list(t, "weights_test", 1, ls),
        "out_full": complex_list(t, "rotated_weights_test", 1, ls)}
    lz = t.index('"local rotation about z axis"')
    out["rotate_test.jl:local_z"] = {
        "phi": scalar(t, "ϕ", 1, lz), "p": 10,
        "in_full": complex_list(t, "weights_test", 1, lz),
        "out_full": complex_list(t, "rotated_weights_test", 1, lz)}

    # --- translate_multipole_test.jl
    t = T("translate_multipole_test.jl")
    out["translate_multipole_test.jl:z"] = {
        "r": scalar(t, "r"), "p": 10,
        "in_slices": None,
        "out_full": complex_list(t, "translated_weights_test", 1)}
    # the input is written as vec(tmp) .= vec([...]); locate it directly
    i = t.index("vec(tmp) .= vec(")
    content, _ = bracket_content(t, i + len("vec(tmp) .= vec("))
    out["translate_multipole_test.jl:z"]["in_slices"] = [
        [[float(x) for x in r.split()] for r in s.split(";") if r.strip()] for s in content.split(";;;")]
    ps = t.index('"multipole-to-multipole: point-and-shoot"')
    out["translate_multipole_test.jl:point_and_shoot"] = {
        "x": [0.1, 0.2, -0.3], "dxs": [-0.2, 0.07, -0.1], "strength": 0.7, "p": 7,
        "target_offset": [0.1, 0.2, 0.14],
        "out_full": complex_list(t, "translated_weights_test", 1, ps)}

    # --- translate_multipole_to_local_test.jl
    t = T("translate_multipole_to_local_test.jl")
    out["translate_multipole_to_local_test.jl:z"] = {
        "r": scalar(t, "r"), "p": 10,
        "in_full": complex_list(t, "original_weights_test"),
        "out_full": complex_list(t, "translated_weights_test")}
    ps = t.index('"multipole-to-local: point-and-shoot"')
    out["translate_multipole_to_local_test.jl:point_and_shoot"] = {
        "p": 10, "source_center": [0.2, 0.4, -0.15999999999999998], "target_center": [2.5, -4.3999999999999995, 0.8],
        "in_full": complex_list(t, "multipole_expansion_test", 1, ps),
        "out_full": complex_list(t, "local_expansion_test", 1, ps)}

    # --- translate_local_test.jl
    t = T("translate_local_test.jl")
    out["translate_local_test.jl:z"] = {
        "r": scalar(t, "r"), "p": 10,
        "in_full": complex_list(t, "original_weights_test"),
        "out_full": complex_list(t, "translated_weights_test")}
    ps = t.index('"local-to-local: point-and-shoot"')
    out["translate_local_test.jl:point_and_shoot"] = {
        "p": 10, "source_center": [2.5, -4.3999999999999995, 0.8], "target_center": [2.4, -4.199999999999999, 0.5],
        "in_full": complex_list(t, "source_weights_test", 1, ps),
        "out_full": complex_list(t, "translated_weights_test", 1, ps)}

    # --- evaluate_expansions_test.jl
    t = T("evaluate_expansions_test.jl")
    out["evaluate_expansions_test.jl:l2p"] = {
        "p": 10, "x_target": [2.3, -4.1, 0.4], "x_source": [0.1 - 0.2, 0.2 + 0.07, -0.3 - 0.1], "mass": 0.7,
        "center": [2.4, -4.199999999999999, 0.5],
        "local_full": complex_list(t, "local_weights_test")}

    # --- lamb_helmholtz_test.jl
    t = T("lamb_helmholtz_test.jl")
    for key, title, rname in (("transform_multipole", '"Lamb-Helmholtz transform: multipole"', "r"),
                              ("transform_local", '"Lamb-Helmholtz transform: local"', "r")):
        s = t.index(title)
        out[f"lamb_helmholtz_test.jl:{key}"] = {
            "r": scalar(t, rname, 1, s), "p": 10,
            "w1": complex_list(t, "weights_1", 1, s), "w2": complex_list(t, "weights_2", 1, s),
            "new_w1": complex_list(t, "new_weights_1", 1, s), "new_w2": complex_list(t, "new_weights_2", 1, s)}
    s = t.index('"Lamb-Helmholtz transform: multipole-to-multipole"')
    out["lamb_helmholtz_test.jl:m2m"] = {
        "p": 10, "source_center": [0.01, 0.02, -0.03], "target_center": [0.04, 0.04, -0.019999999999999997],
        "w1": complex_list(t, "weights_1", 1, s), "w2": complex_list(t, "weights_2", 1, s),
        "out1": complex_list(t, "translated_weights_1", 1, s), "out2": complex_list(t, "translated_weights_2", 1, s)}
    s = t.index('"Lamb-Helmholtz transform: multipole-to-local"')
    out["lamb_helmholtz_test.jl:m2l"] = {
        "p": 10, "n_checked": 7, "source_center": [0.04, 0.04, -0.019999999999999997], "target_center": [0.96, 0.03, -0.02],
        "w1": complex_list(t, "weights_1", 1, s), "w2": complex_list(t, "weights_2", 1, s),
        "out1": complex_list(t, "translated_weights_1", 1, s), "out2": complex_list(t, "translated_weights_2", 1, s)}
    s = t.index('"Lamb-Helmholtz transform: local-to-local"')
    out["lamb_helmholtz_test.jl:l2l"] = {
        "p": 10, "n_checked": 7, "source_center": [0.96, 0.03, -0.02], "target_center": [0.98, 0.01, 0.01],
        "w1": complex_list(t, "weights_1", 1, s), "w2": complex_list(t, "weights_2", 1, s),
        "out1": complex_list(t, "translated_weights_1", 1, s), "out2": complex_list(t, "translated_weights_2", 1, s)}
    s = t.index('"Lamb-Helmholtz: vorton"')
    out["lamb_helmholtz_test.jl:vorton"] = {
        "p": 10, "x_source": [0.0, 0.0, 0.0], "x_target": [1.0, 0.0, 0.0], "gamma": [0.0, 0.0, 1.0],
        "center": [0.01, 0.02, -0.03], "local_center_offset": [-0.04, 0.03, -0.02],
        "w1": complex_list(t, "multipole_weights_1", 1, s), "w2": complex_list(t, "multipole_weights_2", 1, s)}

    # --- bodytomultipole_test.jl
    t = T("bodytomultipole_test.jl")
    s = t.index('"body-to-multipole: point dipole"')
    out["bodytomultipole_test.jl:point_dipole"] = {
        "p": 10, "x_source": [0.5, -0.3, 0.45], "x_target": [7.3, 8.4, -7.2], "q": [0.7, 0.3, 1.5],
        "center": [0.0, 0.0, 0.0], "target_offset": [0.1, 0.2, -0.3],
        "multipole_full": complex_list(t, "multipole_check", 1, s)}
    s = t.index('"body-to-multipole: point vortex"')
    out["bodytomultipole_test.jl:point_vortex"] = {
        "p": 10, "w1": complex_list(t, "multipole_weights_1", 1, s), "w2": complex_list(t, "multipole_weights_2", 1, s)}

    # --- tree_test.jl known-answer fixture (5 bodies)
    out["tree_test.jl:five_bodies"] = {
        "positions": [[1.2, 1.1, 0.8], [0.8, 0.9, 0.2], [0.1, 0.2, 0.9], [0.1, 0.3, 0.2], [0.2, 0.25, 0.4]],
        "center": [0.65, 0.65, 0.55], "n_bodies": [5, 2, 1, 1, 1, 1, 1], "leaf_size": 1,
        "octant_signs": [[0, 0, 0], [-1, -1, -1], [1, 1, -1], [-1, -1, 1], [1, 1, 1]]}

    # --- fmm_test.jl inputs (5 masses) and direct_test.jl (7 masses)
    out["fmm_test.jl:scalar_potential"] = {
        "positions": [[1.2, 1.1, 0.8], [0.8, 0.9, 0.2], [0.1, 0.2, 0.9], [0.1, 0.3, 0.2], [0.2, 0.25, 0.4]],
        "masses": [0.8, 1.1, 2.2, 0.5, 1.9], "p": 24, "multipole_acceptance": 0.5}
    out["direct_test.jl:direct"] = {
        "x": [[-5.0, -4.8, -5.1, 4.9, 5.2, 5.1, 5.3], [0.4, 0.2, 0.1, 0.3, 0.2, -0.1, 0.1], [-0.1, 0.2, 0.1, 0.1, 0.0, -0.1, -0.2]],
        "m": [1.4, 2.0, 2.3, 4.1, 1.1, 3.4, 4.5]}

    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", dst, "with", len(out) - 1, "fixtures")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
