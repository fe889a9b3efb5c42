"""Seeded Wavefront OBJ text for the ingest tests (our own data).

A deformed torus of nu x nv vertices with texture coordinates and normals,
whose faces mix the syntaxes a loader meets in the wild: quads, triangles and
pentagons (fan-triangulated), `v`, `v/vt`, `v//vn` and `v/vt/vn` corners,
negative (relative) indices, comments, blank and indented lines, CRLF line
ends, numbers with exponents, explicit signs and no leading digit, a corner
that points outside the file (dropped), a two-corner face (no triangle) and
a vertex record with extra fields.
"""
import numpy as np


def fmt(rng, x):
    """One of several textual forms of x, chosen by rng."""
    k = rng.integers(0, 6)
    if k == 0:
        return "%.6f" % x
    if k == 1:
        return "%.17g" % x
    if k == 2:
        return "%.5e" % x
    if k == 3:
        s = "%.4f" % x
        return s.replace("0.", ".", 1) if s.startswith("0.") else ("+" + s if x >= 0 else s)
    if k == 4:
        return "%.3E" % x
    return repr(float(np.float32(x)))


def torus_obj(nu=40, nv=24, seed=7, crlf_every=7):
    rng = np.random.default_rng(seed)
    R, r = 2.0, 0.7
    lines = ["# seeded torus for the ingest tests", "o torus", ""]
    vid = {}
    for j in range(nv):
        for i in range(nu):
            u, v = 2 * np.pi * i / nu, 2 * np.pi * j / nv
            rr = r * (1.0 + 0.15 * np.sin(3 * u) * np.cos(2 * v)) + rng.uniform(-0.01, 0.01)
            x = (R + rr * np.cos(v)) * np.cos(u)
            y = rr * np.sin(v) + 0.3 * np.sin(2 * u)
            z = (R + rr * np.cos(v)) * np.sin(u)
            lead = "  " if (i + j) % 11 == 0 else ""
            extra = " 1.0" if (i * 7 + j) % 29 == 0 else ""
            lines.append(lead + "v %s %s %s%s" % (fmt(rng, x), fmt(rng, y), fmt(rng, z), extra))
            vid[(i, j)] = len(vid) + 1
    for j in range(nv):
        for i in range(nu):
            u, v = i / nu, j / nv
            if (i + 2 * j) % 5 == 0:
                lines.append("vt %s %s %s" % (fmt(rng, u), fmt(rng, v), fmt(rng, 0.5)))
            else:
                lines.append("vt %s %s" % (fmt(rng, u), fmt(rng, v)))
    for j in range(nv):
        for i in range(nu):
            u, v = 2 * np.pi * i / nu, 2 * np.pi * j / nv
            n = np.array([np.cos(v) * np.cos(u), np.sin(v), np.cos(v) * np.sin(u)])
            n += rng.uniform(-0.05, 0.05, 3)
            lines.append("vn %s %s %s" % (fmt(rng, n[0]), fmt(rng, n[1]), fmt(rng, n[2])))
    lines.append("")
    lines.append("# faces")
    total = nu * nv

    def corner(k, mode, rel):
        idx = k - total - 1 if rel else k  # negative: counted back from the file's totals
        if mode == 0:
            return "%d" % idx
        if mode == 1:
            return "%d/%d" % (idx, idx)
        if mode == 2:
            return "%d//%d" % (idx, idx)
        return "%d/%d/%d" % (idx, idx, idx)

    for j in range(nv):
        for i in range(nu):
            a = vid[(i, j)]
            b = vid[((i + 1) % nu, j)]
            c = vid[((i + 1) % nu, (j + 1) % nv)]
            d = vid[(i, (j + 1) % nv)]
            mode = int(rng.integers(0, 4)) if rng.uniform() < 0.3 else 3
            rel = rng.uniform() < 0.2
            shape = (i + j) % 9
            if shape == 0:
                lines.append("f " + " ".join(corner(k, mode, rel) for k in (a, b, c)))
                lines.append("f " + " ".join(corner(k, mode, rel) for k in (a, c, d)))
            elif shape == 1:
                e = vid[((i + 2) % nu, j)]
                lines.append("f " + " ".join(corner(k, mode, rel) for k in (a, e, b, c, d)))
            elif shape == 2:
                lines.append("\tf  " + " ".join(corner(k, mode, rel) for k in (a, b, c, d)))
            else:
                lines.append("f " + " ".join(corner(k, mode, rel) for k in (a, b, c, d)))
    lines.append("f 1/1/1 %d/1/1 2/2/2" % (total + 50))  # one corner outside the file
    lines.append("f 3 4")                                # two corners: no triangle
    lines.append("# end")
    out = []
    for n, ln in enumerate(lines):
        out.append(ln + ("\r\n" if crlf_every and n % crlf_every == 0 else "\n"))
    return "".join(out).encode()


def big_obj(n_side=1000, seed=3):
    """A large height-field mesh (n_side^2 vertices, 2 (n_side-1)^2
    triangles) as OBJ text, for the ingest timing."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(-50, 50, n_side)
    X, Z = np.meshgrid(xs, xs)
    Y = 2.0 * np.sin(X * 0.1) * np.cos(Z * 0.13) + rng.uniform(-0.05, 0.05, X.shape)
    v = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    head = "\n".join("v %.6f %.6f %.6f" % tuple(p) for p in v)
    i = np.arange(n_side - 1)
    I, J = np.meshgrid(i, i)
    a = (J * n_side + I + 1).ravel()
    b, c, d = a + 1, a + n_side + 1, a + n_side
    f1 = np.stack([a, b, c], 1)
    f2 = np.stack([a, c, d], 1)
    faces = np.empty((2 * len(a), 3), np.int64)
    faces[0::2], faces[1::2] = f1, f2
    tail = "\n".join("f %d %d %d" % tuple(f) for f in faces)
    return (head + "\n" + tail + "\n").encode()
