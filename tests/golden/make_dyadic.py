"""Write tests/golden/dyadic_coverage.json: exact expected images for scenes
of triangles whose fp32 evaluation is exact.

Every scene is drawn in a 16x16 frame with the orthographic projection
ortho(0, 16, 0, 16, -1, 1), which maps eye (x, y) to pixel (x, 16 - y) and
eye z to NDC depth -z without rounding.  Vertices lie on a 1/4-pixel grid,
twice the signed pixel area is a power of two and vertex depths are
multiples of 1/16, so every fp32 operation of Renderer.cpp:205-261 is exact
and the reference must produce exactly what rational geometry gives:

  covered      pixel centre inside the closed triangle (inclusive edges)
  culled       twice the signed area (p0-p1)^(p2-p1) <= 0
  depth        the plane's depth at the pixel centre
  winner       first triangle in draw order with the smallest depth

This is independent of the oracle: it uses orientation tests and Fractions,
not the edge-function walk.  Run: python tests/golden/make_dyadic.py
"""
import json
import random
from fractions import Fraction as Q
from pathlib import Path

W = H = 16
OUT = Path(__file__).with_name("dyadic_coverage.json")


def orient(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def twice_area_ref(p):
    # (p0 - p1) ^ (p2 - p1): positive = front face (y down in pixels)
    ax, ay = p[0][0] - p[1][0], p[0][1] - p[1][1]
    bx, by = p[2][0] - p[1][0], p[2][1] - p[1][1]
    return ax * by - ay * bx


def expected(tris):
    """tris: list of ((px,py,zz) x 3) with Fractions, pixel coordinates."""
    win = [[-1] * W for _ in range(H)]
    dep = [[None] * W for _ in range(H)]
    for t, tri in enumerate(tris):
        p = [(v[0], v[1]) for v in tri]
        if twice_area_ref(p) <= 0:
            continue
        o = orient(*p)
        for i in range(H):
            for j in range(W):
                c = (Q(2 * j + 1, 2), Q(2 * i + 1, 2))
                l0 = orient(c, p[1], p[2]) / o
                l1 = orient(p[0], c, p[2]) / o
                l2 = orient(p[0], p[1], c) / o
                if l0 < 0 or l1 < 0 or l2 < 0:
                    continue
                z = l0 * tri[0][2] + l1 * tri[1][2] + l2 * tri[2][2]
                if dep[i][j] is None or z < dep[i][j]:
                    dep[i][j] = z
                    win[i][j] = 2 * t
    return win, dep


def rand_tri(rng):
    while True:
        p = [(Q(rng.randint(-8, 72), 4), Q(rng.randint(-8, 72), 4)) for _ in range(3)]
        a = twice_area_ref(p)
        if a == 0:
            continue
        m = abs(a)
        if m.numerator & (m.numerator - 1) or m.denominator & (m.denominator - 1):
            continue  # not a power of two
        if not (Q(1, 4) <= m <= 512):
            continue
        z = [Q(rng.randint(1, 15), 16) for _ in range(3)]
        if rng.random() < 0.3:
            z = [z[0]] * 3
        return [(p[k][0], p[k][1], z[k]) for k in range(3)]


def hand_scenes():
    f = Q
    quarter = f(1, 4)
    # quad split on its diagonal, drawn twice: the diagonal and the shared
    # edges go to the first-drawn triangle
    # (8 x 8 legs: twice the area is 64, a power of two, so fp32 is exact;
    # with e.g. 12 x 12 legs the reference's own rounding leaves crack pixels
    # on the diagonal that exact geometry does not have)
    quad = [[(f(2), f(2), quarter), (f(2), f(10), quarter), (f(10), f(10), quarter)],
            [(f(2), f(2), quarter), (f(10), f(10), quarter), (f(10), f(2), quarter)]]
    right = [(f(2), f(2), quarter), (f(2), f(10), quarter), (f(10), f(2), quarter)]
    flipped = [right[0], right[2], right[1]]
    tilted = [(f(0), f(0), f(1, 16)), (f(0), f(16), f(15, 16)), (f(16), f(0), f(8, 16))]
    return [quad + quad, [right, right], [flipped, right], [tilted, right]]


def main():
    rng = random.Random(20240601)
    scenes = hand_scenes()
    for _ in range(36):
        scenes.append([rand_tri(rng) for _ in range(rng.randint(1, 6))])
    out = []
    for tris in scenes:
        win, dep = expected(tris)
        out.append({
            "triangles": [[[float(v[0]), float(v[1]), float(v[2])] for v in t] for t in tris],
            "winner": win,
            "depth": [[None if d is None else float(d) for d in row] for row in dep],
        })
    OUT.write_text(json.dumps({"width": W, "height": H, "scenes": out}, separators=(",", ":")))
    print(f"wrote {OUT} ({len(out)} scenes)")


if __name__ == "__main__":
    main()
