"""Writes tests/golden/kats.json: the known-answer vectors of the
reference's own unit tests for the hot path, read from the Zig sources
under /root/reference (text only; nothing of the reference is built or
run).  Each vector keeps its source file:line.  Run in this container:

    python tests/golden/make_kats.py

The JSON holds inputs and expected outputs only; tests/test_oracle_kats.py
checks the oracle against it."""
import json
import math
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import zigexpr as Z  # noqa: E402

REF = os.environ.get("Z2D_REFERENCE", "/root/reference")
OUT = os.path.join(HERE, "kats.json")

FORMATS = ["argb", "xrgb", "rgb", "rgba", "alpha8", "alpha4", "alpha2", "alpha1"]


def src(path):
    with open(os.path.join(REF, path)) as f:
        return f.read()


def line_of(text, offset):
    return text.count("\n", 0, offset) + 1


# ---------------------------------------------------------------------------
# value evaluation
# ---------------------------------------------------------------------------
CONSTS = {"math.pi": math.pi}


def num(node, env=None):
    k = node[0]
    if k == "num":
        return node[1]
    if k == "neg":
        return -num(node[1], env)
    if k == "bin":
        a, b = num(node[2], env), num(node[3], env)
        return {"+": a + b, "-": a - b, "*": a * b, "/": a / b}[node[1]]
    if k == "id" and node[1] in CONSTS:
        return CONSTS[node[1]]
    if k == "call" and node[1] == ("id", "math.floatEps") and node[2] == [("id", "f32")]:
        return 2.0 ** -23
    if k == "id" and env and node[1] in env:
        return env[node[1]]
    raise ValueError(f"not a number: {node}")


def pixel_literal(node):
    """`.{ .fmt = .{ .r = .., ... } }` or `pixel.Pixel{ ... }` -> [fmt, r, g, b, a]."""
    f = Z.fields(node)
    assert len(f) == 1, node
    (fmt, inner), = f.items()
    assert fmt in FORMATS, fmt
    c = Z.fields(inner)
    return [fmt, int(num(c.get("r", ("num", 0)))), int(num(c.get("g", ("num", 0)))),
            int(num(c.get("b", ("num", 0)))), int(num(c.get("a", ("num", 0))))]


def color_args(node):
    """Color.InitArgs `.{ .rgb = .{r, g, b} }` etc. -> [space, [v0..v3]]."""
    f = Z.fields(node)
    (kind, inner), = f.items()
    vals = [float(num(v)) for _, v in inner[2]]
    space = {"rgb": "linear_rgb", "rgba": "linear_rgb", "srgb": "srgb", "srgba": "srgb",
             "hsl": "hsl", "hsla": "hsl"}[kind]
    if len(vals) == 3:
        vals.append(1.0)
    return [space, vals]


def component(node):
    """A color component: a number, or ["powf", x, y] for `math.pow(f32, x,
    y)` (evaluated by the checker with the oracle's std.math.pow)."""
    if node[0] == "call" and node[1] == ("id", "math.pow"):
        assert node[2][0] == ("id", "f32"), node
        return ["powf", float(num(node[2][1])), float(num(node[2][2]))]
    return float(num(node))


def color_init_call(node, env=None):
    """`color.LinearRGB.init(r, g, b, a).asColor()` / `colorpkg.HSL.init(..)`
    / `HSL.init(..)` / a named constant of those -> [space, [v0..v3]]."""
    if node[0] == "id" and env and node[1] in env:
        return env[node[1]]
    if node[0] == "call" and node[1][0] == "field" and node[1][2] == "asColor":
        node = node[1][1]
    if node[0] == "id" and node[1].endswith(".asColor") and env:
        return env[node[1][: -len(".asColor")]]
    assert node[0] == "call", node
    callee = node[1][1]
    space = ("hsl" if callee.endswith("HSL.init") else
             "srgb" if callee.endswith("SRGB.init") else
             "linear_rgb" if callee.endswith("LinearRGB.init") else None)
    assert space, callee
    return [space, [component(a) for a in node[2]]]


def interp_table(text, start, method=None):
    """Color interpolation tables (color.zig): a, b, t, the method (fixed
    for the per-space tables, from the case for InterpolationMethod) and the
    expected color; the checked output is interpolateEncode == expected
    encoded (LinearRGB.fromColor(expected).encodeRGBA())."""
    body = Z.test_block(text, start)
    env = {}
    for m in re.finditer(r"const (\w+) = ((?:color\.)?(?:LinearRGB|SRGB|HSL)\.init\([^;]*\));",
                         body):
        env[m.group(1)] = color_init_call(Z.parse(m.group(2)))
    _, items = Z.cases_table(body)
    out = []
    for it in items:
        f = Z.fields(it)
        if method is None:
            (meth, pv), = Z.fields(f["method"]).items()
            polar = enum(pv) if pv[0] == "enum" else "shorter"
        elif method == "hsl":
            meth, polar = "hsl", enum(f["method"])
        else:
            meth, polar = method, "shorter"
        out.append({"name": f["name"][1], "method": meth, "polar": polar,
                    "a": color_init_call(f["a"], env), "b": color_init_call(f["b"], env),
                    "t": float(num(f["t"])), "expected": color_init_call(f["expected"], env)})
    return out


def plotter_cases(text, start):
    """fill_plotter.plot on node lists: edges and their count."""
    body = Z.test_block(text, start)
    out, nodes = [], []
    cur = None
    for st in flatten_blocks(Z.statements(body)):
        if re.match(r"^(const alloc|var nodes|defer )", st):
            if st.startswith("var nodes"):
                nodes = []
            continue
        m = re.match(r"^try nodes\.append\(alloc, (.*)\)$", st, re.S)
        if m:
            nodes.append(path_node(Z.parse(m.group(1))))
            continue
        m = re.match(r"^var result = try plot\(alloc, nodes\.items, (.*)\)$", st, re.S)
        if m:
            sc, tol = Z.parse("f(" + m.group(1) + ")")[2]
            cur = {"nodes": list(nodes), "scale": float(num(sc)), "tolerance": float(num(tol))}
            out.append(cur)
            continue
        m = re.match(r"^try testing\.expectEqual\((\d+), result\.edges\.items\.len\)$", st)
        if m:
            cur["n_edges"] = int(m.group(1))
            continue
        m = re.match(r"^try testing\.expectEqualSlices\(Polygon\.Edge, (.*), result\.edges\.items\)$",
                     st, re.S)
        if m:
            lst = Z.parse(m.group(1))
            assert lst[0] == "addr"
            cur["edges"] = [[float(num(Z.fields(e)[k])) for k in ("y0", "y1", "x_start", "x_inc")]
                            for _, e in lst[1][2]]
            continue
        raise ValueError(f"statement not understood: {st[:80]}")
    return out


def path_node(node):
    """`.{ .move_to = .{ .point = .{ .x = 5, .y = 0 } } }` -> [tag, x, y]."""
    (tag, inner), = Z.fields(node).items()
    if tag == "close_path":
        return ["close_path"]
    pt = Z.fields(Z.fields(inner)["point"])
    return [tag, float(num(pt["x"])), float(num(pt["y"]))]


def closed_sets(text, start):
    body = Z.test_block(text, start).replace("[_]PathNode{", ".{")
    out, nodes = [], None
    for st in flatten_blocks(Z.statements(body)):
        m = re.match(r"^const nodes = (.*)$", st, re.S)
        if m:
            nodes = [path_node(v) for _, v in Z.parse(m.group(1))[2]]
            continue
        m = re.match(r"^try testing\.expectEqual\((true|false), PathNode\.isClosedNodeSet\((.*)\)\)$",
                     st, re.S)
        if m:
            arg = m.group(2).strip()
            out.append({"nodes": [] if arg == "&.{}" else nodes, "expected": m.group(1) == "true"})
            continue
        raise ValueError(f"statement not understood: {st[:80]}")
    return out


def matrix_given(node):
    """`Transformation.identity` or `Transformation.identity.scale(sx, sy)`
    -> the matrix handed to setTransformation, {ax, by, cx, dy, tx, ty}."""
    if node is None or node == ("id", "Transformation.identity"):
        return [1.0, 0.0, 0.0, 1.0, 0.0, 0.0]
    assert node[0] == "call" and node[1] == ("id", "Transformation.identity.scale"), node
    sx, sy = (float(num(a)) for a in node[2])
    return [sx, 0.0, 0.0, sy, 0.0, 0.0]


def enum(node):
    assert node[0] == "enum", node
    return node[1]


# ---------------------------------------------------------------------------
# compositor: runPixel over format pairs (src_over, dst_in)
# ---------------------------------------------------------------------------
def pixel_chain(node, env):
    """A pixel expression as {base: [fmt, r, g, b, a], ops: [...]}, where ops
    are ["convert", fmt] (T.fromPixel), ["multiply"] (RGBA.multiply) and
    ["set_a", v] (an assignment to the alpha field)."""
    k = node[0]
    if k == "init":
        return {"base": pixel_literal(node), "ops": []}
    if k == "id":
        name = node[1]
        if name.endswith(".asPixel"):
            return pixel_chain(("id", name[: -len(".asPixel")]), env)
        return json.loads(json.dumps(env[name]))
    if k == "call":
        callee, args = node[1], node[2]
        if callee[0] == "id":
            name = callee[1]
            m = re.fullmatch(r"pixel\.(\w+)\.fromPixel", name)
            if m:
                fmt = {"RGB": "rgb", "XRGB": "xrgb", "RGBA": "rgba", "ARGB": "argb",
                       "Alpha8": "alpha8", "Alpha4": "alpha4", "Alpha2": "alpha2",
                       "Alpha1": "alpha1"}[m.group(1)]
                c = pixel_chain(args[0], env)
                c["ops"].append(["convert", fmt])
                return c
            if name.endswith(".asPixel"):
                return pixel_chain(("id", name[: -len(".asPixel")]), env)
            if name.endswith(".multiply"):
                c = pixel_chain(("id", name[: -len(".multiply")]), env)
                c["ops"].append(["multiply"])
                return c
        if callee[0] == "field" and callee[2] == "asPixel":
            return pixel_chain(callee[1], env)
        if callee[0] == "field" and callee[2] == "multiply":
            c = pixel_chain(callee[1], env)
            c["ops"].append(["multiply"])
            return c
    raise ValueError(f"pixel expression not understood: {node}")


def flatten_blocks(stmts):
    out = []
    for s in stmts:
        if s.startswith("{"):
            out += flatten_blocks(Z.statements(s[1:s.rindex("}")]))
        else:
            out.append(s)
    return out


def compositor_pairs(text, start):
    body = Z.test_block(text, start)
    env, cases = {}, []
    decl = re.compile(r"^(?:const|var)\s+(\w+)(?:\s*:\s*([\w.]+))?\s*=\s*(.*)$", re.S)
    for st in flatten_blocks(Z.statements(body)):
        m = decl.match(st)
        if m:
            name, typ, rhs = m.groups()
            node = Z.parse(rhs)
            if typ == "pixel.RGBA" and node[0] == "init":
                f = Z.fields(node)
                env[name] = {"base": ["rgba"] + [int(num(f[c])) for c in "rgba"], "ops": []}
            else:
                env[name] = pixel_chain(node, env)
            continue
        m = re.match(r"^(\w+)\.a\s*=\s*(\d+)$", st)
        if m:  # a field assignment: `fg_127.a = 127`, `bg_alpha1.a = 0`
            env[m.group(1)]["ops"].append(["set_a", int(m.group(2))])
            continue
        if st.startswith("try testing.expectEqualDeep("):
            call = Z.parse(st[len("try "):])
            exp, got = call[2]
            assert got[0] == "call" and got[1] == ("id", "runPixel"), got
            prec, dst, srcp, op = got[2]
            (efmt, einner), = Z.fields(exp).items()
            expected = ({"base": pixel_literal(exp), "ops": []} if einner[0] == "init"
                        else pixel_chain(einner, env))  # `.{ .rgb = fg_rgb }`
            cases.append({"precision": enum(prec), "op": enum(op),
                          "dst": pixel_chain(dst, env), "src": pixel_chain(srcp, env),
                          "expected": expected})
            continue
        raise ValueError(f"statement not understood: {st[:80]}")
    return cases


def compositor_all_ops(text, start, precision):
    body = Z.test_block(text, start)
    _, items = Z.cases_table(body)
    out = []
    for it in items:
        f = Z.fields(it)
        out.append({"name": f["name"][1], "precision": precision, "op": enum(f["operator"]),
                    "expected": pixel_literal(f["expected"]), "bg": color_args(f["bg"]),
                    "fg": color_args(f["fg"])})
    return out


# ---------------------------------------------------------------------------
# pixels
# ---------------------------------------------------------------------------
def pixel_from_pixel(text, start):
    body = Z.test_block(text, start)
    _, items = Z.cases_table(body)
    out = []
    for it in items:
        f = Z.fields(it)
        pairs = f["pairs"]
        assert pairs[0] == "addr"
        for _, pr in pairs[1][2]:
            p = Z.fields(pr)
            out.append({"name": f["name"][1], "a": pixel_literal(p["a"]),
                        "b": pixel_literal(p["b"])})
    return out


def pixel_from_color(text, start):
    body = Z.test_block(text, start)
    _, items = Z.cases_table(body)
    return [{"name": Z.fields(it)["name"][1], "expected": pixel_literal(Z.fields(it)["expected"]),
             "color": color_args(Z.fields(it)["args"])} for it in items]


# ---------------------------------------------------------------------------
# dither, gradients, polygon, nodes
# ---------------------------------------------------------------------------
def dither_cases(text, start):
    body = Z.test_block(text, start)
    _, items = Z.cases_table(body)
    g = re.search(r"\.linear = \.\{ \.x0 = ([\d.]+), \.y0 = ([\d.]+), \.x1 = ([\d.]+), "
                  r"\.y1 = ([\d.]+) \}", body)
    grad = {"type": "linear", "geom": [float(v) for v in g.groups()], "method": "linear_rgb",
            "stops": [[float(num(Z.parse(o))), color_args(Z.parse(c))] for o, c in re.findall(
                r"g\.addStopAssumeCapacity\(([^,]+), (\.\{ \.rgb = \.\{[^}]*\} \})\)", body)]}
    out = []
    for it in items:
        f = Z.fields(it)
        s = f["source"]
        if s[0] == "enum":
            source = {"gradient": grad}
        else:
            (kind, v), = Z.fields(s).items()
            source = {"pixel": pixel_literal(v)} if kind == "pixel" else {"color": color_args(v)}
        out.append({"name": f["name"][1], "expected": pixel_literal(f["expected"]),
                    "type": enum(f["type"]), "source": source, "scale": int(num(f["scale"])),
                    "x": int(num(f["x"])), "y": int(num(f["y"]))})
    return out


def gradient_offsets(text, start, kind):
    body = Z.test_block(text, start)
    _, items = Z.cases_table(body)
    out = []
    for it in items:
        f = Z.fields(it)
        if kind == "linear":
            geom = [num(f[k]) for k in ("x0", "y0", "x1", "y1")]
        elif kind == "radial":
            geom = [num(f[k]) for k in ("inner_x", "inner_y", "inner_radius", "outer_x",
                                        "outer_y", "outer_radius")]
        else:
            geom = [num(f[k]) for k in ("center_x", "center_y", "angle")]
        out.append({"name": f["name"][1], "type": kind, "geom": [float(v) for v in geom],
                    "matrix_given": matrix_given(f.get("matrix")),
                    "x": int(num(f["x"])), "y": int(num(f["y"])),
                    "expected": float(num(f["expected"]))})
    return out


def gradient_pixels(text, start):
    """Statement walk of a getPixel test: initBuffer, addAssumeCapacity on
    its stops, then expectEqualDeep(pixel, gradient.getPixel(x, y))."""
    body = Z.test_block(text, start)
    out, cur = [], None
    for st in flatten_blocks(Z.statements(body)):
        m = re.match(r"^var gradient = (\w+)\.initBuffer\((.*)\)$", st, re.S)
        if m:
            args = Z.parse("f(" + m.group(2) + ")")[2]
            kind = m.group(1).lower()
            method = args[-1]
            if method[0] == "enum":
                meth, polar = method[1], "shorter"
            else:
                (meth, pv), = Z.fields(method).items()
                polar = enum(pv)
            cur = {"type": kind, "geom": [float(num(a)) for a in args[:-2]], "method": meth,
                   "polar": polar, "stops": []}
            continue
        m = re.match(r"^gradient\.stops\.addAssumeCapacity\((.*)\)$", st, re.S)
        if m:
            off, col = Z.parse("f(" + m.group(1) + ")")[2]
            cur["stops"].append([float(num(off)), color_args(col)])
            continue
        if st.startswith("try testing.expectEqualDeep("):
            exp, got = Z.parse(st[len("try "):])[2]
            assert got[0] == "call" and got[1] == ("id", "gradient.getPixel"), got
            out.append({"gradient": json.loads(json.dumps(cur)), "x": int(num(got[2][0])),
                        "y": int(num(got[2][1])), "expected": pixel_literal(exp)})
            continue
        if st.startswith("var stop_buffer"):
            continue
        raise ValueError(f"statement not understood: {st[:80]}")
    return out


def stop_search(text, start):
    """Stop.List.search: stops added in order, then search(t) against
    (c0, c1, offset)."""
    body = Z.test_block(text, start)
    out, stops = [], []
    pending = None
    for st in flatten_blocks(Z.statements(body)):
        m = re.match(r"^(?:stop_list|stops)\.addAssumeCapacity\((.*)\)$", st, re.S)
        if m:
            off, col = Z.parse("f(" + m.group(1) + ")")[2]
            stops.append([float(num(off)), color_args(col)])
            continue
        m = re.match(r"^try testing\.expectEqualDeep\(Stop\.List\.SearchResult\{(.*)\}, "
                     r"(\w+)\.search\((.*)\)\)$", st, re.S)
        if m:
            f = Z.fields(Z.parse(".{" + m.group(1) + "}"))
            lst = [] if m.group(2) == "stop_list_zero" else stops
            out.append({"stops": json.loads(json.dumps(lst)), "t": float(num(Z.parse(m.group(3)))),
                        "c0": color_init_call(f["c0"]), "c1": color_init_call(f["c1"]),
                        "offset": float(num(f["offset"])), "tol": 0.0})
            continue
        m = re.match(r"^(?:var )?got = stop_list\.search\((.*)\)$", st, re.S)
        if m:
            pending = {"stops": json.loads(json.dumps(stops)), "t": float(num(Z.parse(m.group(1))))}
            continue
        m = re.match(r"^(?:var )?expected(?:\s*:\s*[\w.]+)? = \.\{(.*)\}$", st, re.S)
        if m:
            f = Z.fields(Z.parse(".{" + m.group(1) + "}"))
            pending.update({"c0": color_init_call(f["c0"]), "c1": color_init_call(f["c1"]),
                            "offset": float(num(f["offset"])),
                            "tol": 2.0 ** -23})  # expectApproxEqAbs(.., floatEps(f32))
            out.append(pending)
            continue
        if re.match(r"^try testing\.expect(EqualDeep|ApproxEqAbs)\(expected", st):
            continue
        if re.match(r"^(var|const) (stops|stop_buffer|stop_list|stop_list_zero)\b", st):
            continue
        raise ValueError(f"statement not understood: {st[:80]}")
    return out


def in_box(text, start):
    body = Z.test_block(text, start)
    _, items = Z.cases_table(body)
    out = []
    for it in items:
        f = Z.fields(it)
        p = Z.fields(f["polygon"])
        # the test calls inBox(scale, box_height, box_width) on a function
        # declared inBox(scale, box_width, box_height): kept as called
        out.append({"name": f["name"][1],
                    "extents": [num(p["extent_top"]), num(p["extent_bottom"]),
                                num(p["extent_left"]), num(p["extent_right"])],
                    "scale": float(num(f["scale"])), "width_arg": int(num(f["box_height"])),
                    "height_arg": int(num(f["box_width"])),
                    "expected": f["expected"] == ("id", "true")})
    return out


def main():
    kats = {}
    comp = src("src/compositor.zig")
    kats["compositor_pairs"] = {
        "source": "src/compositor.zig:2452-3076",
        "cases": compositor_pairs(comp, 2452) + compositor_pairs(comp, 2766)}
    kats["compositor_ops"] = {
        "source": "src/compositor.zig:3078-3443 (integer), :3445-3860 (float)",
        "cases": compositor_all_ops(comp, 3078, "integer") +
        compositor_all_ops(comp, 3445, "float")}
    pix = src("src/pixel.zig")
    kats["pixel_from_pixel"] = {"source": "src/pixel.zig:845-1168",
                                "cases": pixel_from_pixel(pix, 845)}
    kats["pixel_from_color"] = {"source": "src/pixel.zig:1371-1415",
                                "cases": pixel_from_color(pix, 1371)}
    kats["dither"] = {"source": "src/Dither.zig:164-266",
                      "cases": dither_cases(src("src/Dither.zig"), 164)}
    gr = src("src/gradient.zig")
    kats["gradient_offset"] = {
        "source": "src/gradient.zig:1200-1250, :1373-1486, :1702-1766",
        "cases": gradient_offsets(gr, 1200, "linear") + gradient_offsets(gr, 1373, "radial") +
        gradient_offsets(gr, 1702, "conic")}
    kats["gradient_pixel"] = {
        "source": "src/gradient.zig:1252-1371, :1608-1619, :1768-1779",
        "cases": gradient_pixels(gr, 1252) + gradient_pixels(gr, 1608) +
        gradient_pixels(gr, 1768)}
    kats["stop_search"] = {"source": "src/gradient.zig:963-1146",
                           "cases": stop_search(gr, 963) + stop_search(gr, 1098)}
    kats["in_box"] = {"source": "src/internal/tess/Polygon.zig:452-784",
                      "cases": in_box(src("src/internal/tess/Polygon.zig"), 452)}
    col = src("src/color.zig")
    kats["interpolate"] = {
        "source": "src/color.zig:788-886 (linear), :1056-1153 (sRGB), :1334-1511 (HSL), "
                  ":1513-1779 (InterpolationMethod)",
        "cases": interp_table(col, 788, "linear_rgb") + interp_table(col, 1056, "srgb") +
        interp_table(col, 1334, "hsl") + interp_table(col, 1513)}
    fp = src("src/internal/tess/fill_plotter.zig")
    kats["fill_plotter"] = {"source": "src/internal/tess/fill_plotter.zig:121-209",
                            "cases": plotter_cases(fp, 121) + plotter_cases(fp, 153)}
    kats["closed_node_set"] = {"source": "src/internal/path_nodes.zig:63-137",
                               "cases": closed_sets(src("src/internal/path_nodes.zig"), 63)}
    with open(OUT, "w") as f:
        json.dump(kats, f, indent=1)
        f.write("\n")
    for k, v in kats.items():
        print(f"{k}: {len(v['cases'])} cases ({v['source']})")


if __name__ == "__main__":
    main()
