"""Wavefront OBJ meshes, parsed as the reference's ObjLoader.Load does
without subdivision or displacement (src/RayTracer/Scene/ObjLoader.cs:99-294,
BuildTrianglesDirect :357-400, TryParseFaceVertex :430-455, TryParseIndex
:458-474, TryParseVec3/Vec2 :477-502):

  * lines are trimmed; empty lines and '#' comments are skipped; only
    ``v ``, ``vn ``, ``vt `` and ``f `` (prefix with one space) are read,
    every other directive (o, g, s, usemtl, mtllib ...) is ignored;
  * numbers are split on spaces only, parsed with .NET's invariant float
    grammar and rounded once to float32; a malformed line is a warning;
  * face vertices ``v``, ``v/vt``, ``v/vt/vn``, ``v//vn``; 1-based or
    negative (relative to the counts so far) indices; a position index out
    of range drops that vertex with a warning, a face with fewer than three
    vertices is skipped;
  * a face carries normals (UVs) only when every one of its vertices has
    one; a channel that some faces have and others lack is dropped for the
    whole file;
  * faces are fan-triangulated from their first vertex; a triangle whose
    float32 edge cross product has squared length below 1e-12 is skipped;
  * with normals (and UVs) a triangle is a SmoothTriangle, else a flat
    Triangle (UVs alone are dropped, as the flat Triangle has none).

The result is the Mesh's triangle list in the reference's order, ready for
SceneBuilder.mesh_triangles."""
import os
import re
from fractions import Fraction

import numpy as np

_WS = " \t\n\v\f\r"
_FLOAT = re.compile(r"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?")
_SPECIAL = {"infinity": np.inf, "+infinity": np.inf, "-infinity": -np.inf, "nan": np.nan, "+nan": np.nan,
            "-nan": np.nan, "∞": np.inf, "+∞": np.inf, "-∞": -np.inf}
_INT = re.compile(r"[+-]?\d+")
_LINES = re.compile(r"\r\n|\r|\n")     # File.ReadLines' line breaks


class ObjLoadError(ValueError):
    """The reference's loader would throw on this input."""


def parse_float32(tok):
    """float.TryParse(tok, NumberStyles.Float, InvariantCulture) rounded once
    to float32 (round to nearest even from the decimal), or None."""
    s = tok.strip(_WS)
    if not _FLOAT.fullmatch(s):
        v = _SPECIAL.get(s.lower())
        return None if v is None else np.float32(v)
    x = float(s)                       # correctly rounded to double
    with np.errstate(over="ignore"):
        f = np.float32(x)
    if np.isfinite(f) and float(f) != x:
        # a double that sits exactly halfway between two floats may be the
        # rounding of a decimal that is not: decide from the decimal itself
        g = np.nextafter(f, np.float32(np.inf) if x > float(f) else np.float32(-np.inf))
        mid = (float(f) + float(g)) / 2.0
        if x == mid:
            exact = Fraction(s)        # a true tie keeps f: the conversion rounded to even
            if exact != Fraction(mid) and (exact > Fraction(mid)) == (float(g) > float(f)):
                f = g
    return f


def _parse_int32(s):
    s = s.strip(_WS)
    if not _INT.fullmatch(s):
        return None
    v = int(s)
    return v if -2**31 <= v < 2**31 else None


def _parse_index(s, n):
    """TryParseIndex: (ok, index); index is set even when out of range."""
    if not s:
        return False, -1
    raw = _parse_int32(s)
    if raw is None or raw == 0:
        return False, -1
    idx = raw - 1 if raw > 0 else n + raw
    return 0 <= idx < n, idx


def _split(s, limit):
    """String.Split(' ', RemoveEmptyEntries) into at most `limit` parts, the
    last holding the remainder."""
    out = []
    i, n = 0, len(s)
    while i < n and len(out) < limit - 1:
        while i < n and s[i] == " ":
            i += 1
        if i >= n:
            break
        j = s.find(" ", i)
        j = n if j < 0 else j
        out.append(s[i:j])
        i = j
    while i < n and s[i] == " ":
        i += 1
    if i < n:
        out.append(s[i:])
    return out


def _vec(line, prefix, k):
    parts = _split(line[prefix:].strip(_WS), k + 1)
    if len(parts) < k:
        return None
    vals = [parse_float32(p) for p in parts[:k]]
    return None if any(v is None for v in vals) else vals


class ObjMesh:
    """Triangles of one OBJ file in Mesh order.

    tris: (n, 9) float32 vertex positions; normals: (n, 9) float32 vertex
    normals (zero rows for flat triangles) or None; uvs: (n, 6) float32
    vertex UVs of the smooth triangles that have them (barycentric UVs
    elsewhere) or None; vertex_count: positions in the file; warnings."""

    def __init__(self, tris, normals, uvs, vertex_count, warnings):
        self.tris, self.normals, self.uvs = tris, normals, uvs
        self.vertex_count = vertex_count
        self.warnings = warnings

    def __len__(self):
        return len(self.tris)


def parse(text, warnings=None):
    """ObjLoader.Load on the file's text; None when it yields no triangles
    (the reference returns null and warns)."""
    warnings = [] if warnings is None else warnings
    pos, nrm, uv = [], [], []
    f_pos, f_nrm, f_uv = [], [], []
    for num, raw in enumerate(_LINES.split(text), 1):
        line = raw.strip()
        if not line or line[0] == "#":
            continue
        if line.startswith("v "):
            v = _vec(line, 2, 3)
            (pos.append(v) if v is not None else warnings.append(f"OBJ line {num}: malformed vertex position"))
        elif line.startswith("vn "):
            v = _vec(line, 3, 3)
            (nrm.append(v) if v is not None else warnings.append(f"OBJ line {num}: malformed vertex normal"))
        elif line.startswith("vt "):
            v = _vec(line, 3, 2)
            (uv.append(v) if v is not None else warnings.append(f"OBJ line {num}: malformed texture coordinate"))
        elif line.startswith("f "):
            fv = []
            for tok in line[2:].split(" "):
                if not tok:
                    continue
                parts = tok.split("/", 2)
                ok, v = _parse_index(parts[0], len(pos))
                if not ok:
                    warnings.append(f"OBJ line {num}: malformed face vertex '{tok}'")
                    continue
                vt = vn = -1
                if len(parts) >= 2 and parts[1]:
                    _, vt = _parse_index(parts[1], len(uv))
                if len(parts) >= 3 and parts[2]:
                    _, vn = _parse_index(parts[2], len(nrm))
                fv.append((v, vt, vn))
            if len(fv) < 3:
                warnings.append(f"OBJ line {num}: face has fewer than 3 vertices, skipping")
                continue
            f_pos.append([a for a, _, _ in fv])
            f_nrm.append([c for _, _, c in fv] if all(c >= 0 for _, _, c in fv) else [])
            f_uv.append([b for _, b, _ in fv] if all(b >= 0 for _, b, _ in fv) else [])

    def consistent(ch):
        return any(ch) and all(ch)     # some faces have it and none lacks it

    if not consistent(f_nrm):
        f_nrm = None
    if not consistent(f_uv):
        f_uv = None
    if not f_pos:
        warnings.append("OBJ file produced zero faces.")
        return None
    P = np.asarray(pos, np.float32).reshape(-1, 3)
    N = np.asarray(nrm, np.float32).reshape(-1, 3)
    T = np.asarray(uv, np.float32).reshape(-1, 2)

    # fan triangulation (BuildTrianglesDirect)
    fi, k = [], []
    for i, f in enumerate(f_pos):
        for j in range(1, len(f) - 1):
            fi.append(i)
            k.append(j)
    idx = np.empty((len(fi), 3), np.int64)
    for r, (i, j) in enumerate(zip(fi, k)):
        f = f_pos[i]
        idx[r] = (f[0], f[j], f[j + 1])
    v0, v1, v2 = P[idx[:, 0]], P[idx[:, 1]], P[idx[:, 2]]
    a, b = v1 - v0, v2 - v0
    cx = a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1]
    cy = a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2]
    cz = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    keep = ((cx * cx + cy * cy) + cz * cz) >= np.float32(1e-12)
    rows = np.nonzero(keep)[0]
    tris = np.concatenate([v0, v1, v2], 1)[rows]
    normals = uvs = None
    if f_nrm is not None:
        normals = np.zeros((len(rows), 9), np.float32)
        if f_uv is not None:
            uvs = np.tile(np.array([0, 0, 1, 0, 0, 1], np.float32), (len(rows), 1))
        for out, r in enumerate(rows):
            i, j = fi[r], k[r]
            fn = f_nrm[i]
            if not fn:
                continue
            sel = (fn[0], fn[j], fn[j + 1])
            if max(sel) >= len(N):
                raise ObjLoadError(f"OBJ face {i + 1}: normal index out of range")
            normals[out] = N[list(sel)].ravel()
            fu = f_uv[i] if f_uv is not None else []
            if fu:
                sel = (fu[0], fu[j], fu[j + 1])
                if max(sel) >= len(T):
                    raise ObjLoadError(f"OBJ face {i + 1}: texture coordinate index out of range")
                uvs[out] = T[list(sel)].ravel()
        smooth = np.array([bool(f_nrm[fi[r]]) for r in rows])
        zero = (np.abs(normals.reshape(-1, 3, 3)).sum(2) == 0).any(1) & smooth
        if zero.any():
            raise ObjLoadError("OBJ: a smooth triangle has a zero-length vertex normal "
                               "(SmoothTriangle would normalise it to NaN)")
    if len(rows) == 0:
        warnings.append("OBJ file produced zero triangles.")
        return None
    return ObjMesh(tris, normals, uvs, len(P), warnings)


def load(path, warnings=None):
    """ObjLoader.Load(path, ...): None (with a warning) for a missing file."""
    warnings = [] if warnings is None else warnings
    if not os.path.isfile(path):
        warnings.append(f"OBJ file not found: {path}")
        return None
    with open(path, "r", encoding="utf-8-sig", errors="replace") as fh:
        return parse(fh.read(), warnings)
