"""Exact area-coverage rendering of convex polygons, for closed-form tests.

Pixel (i, j) covers [i - 0.5, i + 0.5] x [j - 0.5, j + 0.5] in pixel-index
coordinates (the convention of the detector's corners).  Its value is
light + (dark - light) * (area of the pixel square inside the polygon),
rounded to u8; the area is found by clipping the square against the
polygon's half-planes (Sutherland-Hodgman), so an edge's position is
carried by the pixel values to within the u8 rounding.
"""
from __future__ import annotations

import numpy as np


def _clip(poly, a, b, c):
    """Part of polygon `poly` with a*x + b*y <= c."""
    out = []
    n = len(poly)
    for i in range(n):
        p, q = poly[i], poly[(i + 1) % n]
        fp = a * p[0] + b * p[1] - c
        fq = a * q[0] + b * q[1] - c
        if fp <= 0:
            out.append(p)
        if (fp < 0) != (fq < 0) and fp != fq:
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    return out


def _area(poly):
    s = 0.0
    for i in range(len(poly)):
        x0, y0 = poly[i]
        x1, y1 = poly[(i + 1) % len(poly)]
        s += x0 * y1 - x1 * y0
    return abs(s) * 0.5


def outward_planes(corners):
    """(nx, ny, c) per edge with n the outward unit normal: inside is n.p <= c."""
    pts = np.asarray(corners, dtype=np.float64)
    n = len(pts)
    signed = sum(pts[i, 0] * pts[(i + 1) % n, 1] - pts[(i + 1) % n, 0] * pts[i, 1] for i in range(n))
    planes = []
    for i in range(n):
        p, q = pts[i], pts[(i + 1) % n]
        ex, ey = q - p
        nx, ny = (ey, -ex) if signed > 0 else (-ey, ex)
        norm = np.hypot(nx, ny)
        planes.append((nx / norm, ny / norm, (nx * p[0] + ny * p[1]) / norm))
    return planes


def coverage(width, height, corners):
    """Fraction of every pixel inside the convex polygon `corners`."""
    planes = outward_planes(corners)
    cov = np.ones((height, width))
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    border = np.zeros((height, width), dtype=bool)
    for nx, ny, c in planes:
        d = nx * xx + ny * yy - c  # signed distance of the pixel centre
        cov[d >= 0.7072] = 0.0     # wholly outside this edge
        border |= np.abs(d) < 0.7072
    border &= cov > 0
    for j, i in zip(*np.nonzero(border)):
        poly = [(i - 0.5, j - 0.5), (i + 0.5, j - 0.5), (i + 0.5, j + 0.5), (i - 0.5, j + 0.5)]
        for nx, ny, c in planes:
            poly = _clip(poly, nx, ny, c)
            if not poly:
                break
        cov[j, i] = _area(poly) if poly else 0.0
    return cov


def render(width, height, corners, dark=40.0, light=200.0):
    img = light + (dark - light) * coverage(width, height, corners)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def unit_square_map(corners):
    """Projective map of the unit square (0,0),(1,0),(1,1),(0,1) onto the
    quad `corners` (TL, TR, BR, BL), as a function of (u, v) arrays."""
    c = np.asarray(corners, dtype=np.float64)
    A = []
    rhs = []
    for (u, v), (x, y) in zip(((0, 0), (1, 0), (1, 1), (0, 1)), c):
        A.append([u, v, 1, 0, 0, 0, -u * x, -v * x])
        rhs.append(x)
        A.append([0, 0, 0, u, v, 1, -u * y, -v * y])
        rhs.append(y)
    h = np.linalg.solve(np.array(A), np.array(rhs))

    def f(u, v):
        w = h[6] * u + h[7] * v + 1.0
        return (h[0] * u + h[1] * v + h[2]) / w, (h[3] * u + h[4] * v + h[5]) / w
    return f


def render_tag(code, bits, width, height, corners, black=30.0, white=220.0, background=None):
    """A tag (border cells black, interior cell bit 1 = white, bits read row
    by row from the most significant) with every cell an exactly covered
    polygon; outside the tag the background (default white).  Returns the
    float image before rounding."""
    dim = bits + 2
    f = unit_square_map(corners)
    bg = white if background is None else background
    img = np.full((height, width), float(bg))
    img += (black - bg) * coverage(width, height, corners)  # the whole tag black
    for cy in range(1, dim - 1):
        for cx in range(1, dim - 1):
            k = (cy - 1) * bits + (cx - 1)
            if not (int(code) >> (bits * bits - 1 - k)) & 1:
                continue
            us = np.array([cx, cx + 1, cx + 1, cx]) / dim
            vs = np.array([cy, cy, cy + 1, cy + 1]) / dim
            xs, ys = f(us, vs)
            img += (white - black) * coverage(width, height, np.stack([xs, ys], 1))
    return img
