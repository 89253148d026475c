"""Oracle restatement of the reference's scene-file loading (test infrastructure only).

A second, pure-Python statement of what the reference does when it reads a ``.dumb``
scene and the OBJ files it names, used by the tests to check the library's C++ loader
(dr_scene_file_*) value for value.  It follows the reference directly:

  Dson.cu:224-447            tokens -> Dict / Array / String / Number / Bool / Null
                             (token-count rules of parseToken, duplicate keys rejected,
                             numbers as `stringstream >> double`)
  MeshReader.impl.h          getToken / getLine / getFloat / getNode, `g` objects and their
                             name list, x mirrored for v and vn, vertex normals from all
                             faces, per-object index maps (extructObject)
  PolyMesh.impl.h:140-152    fan triangulation (0, prev, cur)
  Transform.impl.h           rebake (:256-281), >> (:85-87), direction (:166-168),
                             fromDson (:209-237); BakedTriMesh.impl.h:16-18 for normals
  DumbRenderContext.cu       fromDson :337-371, groups :106-271, camera :476-514,
                             renderer :515-640, include :641-659, materials / lights /
                             objects :662-787
  plugins                    DumbBasedShader.impl.cuh:5-16,125-170;
                             SimpleStochasticShader.impl.cuh:4-11,40-65;
                             FresnelShader.impl.cuh:21-25; SimpleSoftDirectionalLight
                             .impl.cuh:5-10,22-49; SphericalLightEmitter.impl.cuh:5-9,21-41;
                             SimpleStochasticLense.impl.cuh:4-9,28-43

fp32 arithmetic is done one operation at a time on numpy float32 scalars; sinf, cosf
and tanf are the C library's (called through ctypes), sqrt is numpy's correctly rounded
float32 sqrt.  Plugin types without a counterpart on the device (textures, spotlights,
spherical_segment lens, primitive meshes) raise SceneError.

The output matches ``dumbray_amd.scene_file.DumbRenderContext``: tris / normals (n, 9)
float32, mat_id (n,), materials and lights as plugin dicts (oracle.pack_material /
pack_light), camera (13 floats), settings.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os
import re

import numpy as np

f32 = np.float32
_LIBM = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
for _n in ("sinf", "cosf", "tanf"):
    getattr(_LIBM, _n).restype = ctypes.c_float
    getattr(_LIBM, _n).argtypes = [ctypes.c_float]

RADIAN = f32(57.29577951308233)


class SceneError(ValueError):
    pass


def sinf(x):
    return f32(_LIBM.sinf(float(x)))


def cosf(x):
    return f32(_LIBM.cosf(float(x)))


def tanf(x):
    return f32(_LIBM.tanf(float(x)))


# ---------------------------------------------------------------------------- Vector3
def V(x, y, z):
    return (f32(x), f32(y), f32(z))


def vadd(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def vsub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def vmul(a, s):
    return (a[0] * s, a[1] * s, a[2] * s)


def vdot(a, b):
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]


def vcross(u, v):
    """u & v, Vector3.impl.h:19-22 (this = u)."""
    return (u[1] * v[2] - v[1] * u[2], v[0] * u[2] - u[0] * v[2], u[0] * v[1] - v[0] * u[1])


def vnormalized(v):
    inv = f32(1.0) / np.sqrt(f32((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]), dtype=f32)
    return vmul(v, inv)


# ---------------------------------------------------------------------------- Transform
class Transform:
    def __init__(self):
        self.position = V(0, 0, 0)
        self.euler = V(0, 0, 0)
        self.scale = V(1, 1, 1)
        self.rebake()

    @staticmethod
    def _turn(axis, v1, v2, angle):
        """rotateTwoAxisAgainstThisRadian: DefSinCos; v = v*cos - (axis x v)*sin for v1 then v2."""
        sa, ca = sinf(angle), cosf(angle)
        k = vcross(axis, v1)
        v1 = (v1[0] * ca - k[0] * sa, v1[1] * ca - k[1] * sa, v1[2] * ca - k[2] * sa)
        k = vcross(axis, v2)
        v2 = (v2[0] * ca - k[0] * sa, v2[1] * ca - k[1] * sa, v2[2] * ca - k[2] * sa)
        return v1, v2

    def rebake(self):
        x, y, z = V(1, 0, 0), V(0, 1, 0), V(0, 0, 1)
        inv = f32(1) / RADIAN                      # Vector3 / float
        r = vmul(self.euler, inv)
        x, z = self._turn(y, x, z, -r[1])
        y, z = self._turn(x, y, z, -r[0])
        x, y = self._turn(z, x, y, -r[2])
        s = self.scale
        self.tx = (x[0] * s[0], x[1] * s[1], x[2] * s[2])
        self.ty = (y[0] * s[0], y[1] * s[1], y[2] * s[2])
        self.tz = (z[0] * s[0], z[1] * s[1], z[2] * s[2])

    def point(self, v):
        p = self.position
        return (vdot(v, self.tx) + p[0], vdot(v, self.ty) + p[1], vdot(v, self.tz) + p[2])

    def normal(self, n):
        return vsub(self.point(n), self.position)

    def _axis(self, t):
        s = self.scale
        return (t[0] / s[0], t[1] / s[1], t[2] / s[2])

    def right(self):
        return self._axis(self.tx)

    def up(self):
        return self._axis(self.ty)

    def front(self):
        return self._axis(self.tz)


# ---------------------------------------------------------------------------- Dson
_WS = " \t\n\r\f\v"
_STRUCT = "{}[],:"
_ESC = {'"': '"', "\\": "\\", "/": "/", "b": "\b", "f": "\f", "n": "\n", "r": "\r", "t": "\t"}


class Tok:
    """A token of the reference's parser network: bare text, or an entry ({ [ ") with sub-tokens."""

    def __init__(self, text="", entry="", sub=None):
        self.text, self.entry, self.sub = text, entry, sub if sub is not None else []


def _tokens(s, i, closer):
    """Tokens until `closer` (None at the top level); returns (tokens, index after closer)."""
    out = []
    while True:
        while i < len(s) and s[i] in _WS:
            i += 1
        if i >= len(s):
            if closer is not None:
                raise SceneError("Dson: unterminated " + closer)
            return out, i
        c = s[i]
        if c == closer:
            return out, i + 1
        if c in "}]":
            raise SceneError("Dson: unexpected " + c)
        if c in "{[":
            sub, i = _tokens(s, i + 1, "}" if c == "{" else "]")
            out.append(Tok(entry=c, sub=sub))
        elif c in ",:":
            out.append(Tok(text=c))
            i += 1
        elif c == '"':
            j, buf = i + 1, []
            while j < len(s) and s[j] != '"':
                if s[j] == "\\" and j + 1 < len(s) and s[j + 1] in _ESC:
                    buf.append(_ESC[s[j + 1]])
                    j += 2
                else:
                    buf.append(s[j])
                    j += 1
            if j >= len(s):
                raise SceneError("Dson: unterminated string")
            txt = "".join(buf)
            out.append(Tok(entry='"', sub=[Tok(text=txt)] if txt else []))
            i = j + 1
        else:
            j = i
            while j < len(s) and s[j] not in _WS and s[j] not in _STRUCT and s[j] != '"':
                j += 1
            out.append(Tok(text=s[i:j]))
            i = j


_NUM = re.compile(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d*)?")


def _stream_double(text):
    """`stream >> double` on "text\\n": the float-looking prefix, which must convert entirely."""
    m = _NUM.match(text)
    if not m:
        return None
    lit = m.group(0)
    if re.fullmatch(r".*[eE][+-]?", lit):
        return None                                # exponent without digits
    return float(lit)


def _value(t):
    if t.entry == "":
        if t.text == "true":
            return True
        if t.text == "false":
            return False
        if t.text == "null":
            return None
        v = _stream_double(t.text)
        if v is None:
            raise SceneError("Could not categorize token '" + t.text + "'")
        return v
    if t.entry == "{":
        n = len(t.sub)
        if n == 0:
            return {}
        if n % 4 != 3:
            raise SceneError("Dict has to be a certain amount of a:b couples")
        d = {}
        for k in range(0, n, 4):
            key = t.sub[k]
            if not (key.text == "" and key.entry == '"' and len(key.sub) <= 1):
                raise SceneError("Dict only accepts strings as keys")
            if t.sub[k + 1].text != ":" or t.sub[k + 1].entry:
                raise SceneError("Expected ':' after a key")
            if k + 3 < n and (t.sub[k + 3].text != "," or t.sub[k + 3].entry):
                raise SceneError("Expected ',' after a key : value pair")
        for k in range(0, n, 4):
            name = t.sub[k].sub[0].text if t.sub[k].sub else ""
            if name in d:
                raise SceneError("Dict has a duplicate entry for '" + name + "'")
            d[name] = _value(t.sub[k + 2])
        return d
    if t.entry == "[":
        n = len(t.sub)
        if n == 0:
            return []
        if n % 2 != 1:
            raise SceneError("Array has to be a certain amount of entries, separated by ','")
        for k in range(1, n, 2):
            if t.sub[k].text != "," or t.sub[k].entry:
                raise SceneError("Array elements should be separated by ','")
        return [_value(t.sub[k]) for k in range(0, n, 2)]
    if t.entry == '"':
        if len(t.sub) > 1:
            raise SceneError("bad string token")
        return t.sub[0].text if t.sub else ""
    raise SceneError("Dson parser failed to decode token")


class Str(str):
    """Dson string (distinguishes strings from other values in type checks)."""


def _wrap(v):
    if isinstance(v, dict):
        return {k: _wrap(x) for k, x in v.items()}
    if isinstance(v, list):
        return [_wrap(x) for x in v]
    if isinstance(v, str):
        return Str(v)
    return v


def parse_dson(text):
    toks, _ = _tokens(text, 0, None)
    if len(toks) < 1:
        raise SceneError("No in tokens")
    if len(toks) > 1:
        raise SceneError("Multiple entries in tokens")
    return _wrap(_value(toks[0]))


def _is_num(v):
    return isinstance(v, float) and not isinstance(v, bool)


def _fv(v):
    """Number::floatValue: (float)double."""
    return f32(v)


def _iv(v):
    """Number::intValue: (int)double, truncation toward zero."""
    return int(v)


def vector_from(v):
    """Vector3::fromDson: first three numbers of an array (others skipped), missing 0."""
    if not isinstance(v, list):
        raise SceneError("Vector3 can not be constructed from any Dson::Object other than Dson::Array")
    nums = [_fv(x) for x in v if _is_num(x)]
    nums = (nums + [f32(0)] * 3)[:3]
    return (nums[0], nums[1], nums[2])


def transform_from(v):
    if not isinstance(v, dict):
        raise SceneError("Transform can not be constructed from any Dson::Object other than Dson::Dict")
    t = Transform()
    if "position" in v:
        t.position = vector_from(v["position"])
    if "rotation" in v:
        t.euler = vector_from(v["rotation"])
        t.rebake()
    if "scale" in v:
        t.scale = vector_from(v["scale"])
        t.rebake()
    if "relative_position" in v:
        r = vector_from(v["relative_position"])
        d = vadd(vadd(vmul(t.right(), r[0]), vmul(t.up(), r[1])), vmul(t.front(), r[2]))
        t.position = vadd(t.position, d)
    return t


# ---------------------------------------------------------------------------- OBJ
def _isspace(c):
    return c in " \t\n\r\f\v"


class _ObjText:
    def __init__(self, text):
        self.f = text + "\0"
        self.c = 0

    def move_to_relevant(self, bp="", ib="", lb=""):
        f = self.f
        while f[self.c] != "\0" and (_isspace(f[self.c]) or f[self.c] in bp):
            if f[self.c] in ib:
                return False
            self.c += 1
        return not (f[self.c] != "\0" and f[self.c] in lb)

    def get_line(self):
        f, out = self.f, []
        if self.move_to_relevant():
            while f[self.c] not in "\n\0":
                out.append(f[self.c])
                self.c += 1
        return "".join(out).rstrip(" \t\n\r\f\v")

    def get_token(self, bp="", ib="", lb="#"):
        f, out = self.f, []
        if self.move_to_relevant(bp, ib, lb):
            while f[self.c] != "\0" and not (_isspace(f[self.c]) or f[self.c] in bp):
                out.append(f[self.c])
                self.c += 1
            return "".join(out)
        if f[self.c] != "\0" and f[self.c] in lb:
            self.get_line()
        return ""

    def get_float(self):
        t = self.get_token("", "\n")
        if t == "":
            return f32(0)
        return f32(_atof(t))


def _atof(t):
    m = re.match(r"[ \t\n\r\f\v]*([+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)", t)
    return float(m.group(1)) if m else 0.0


def _atoi(t):
    m = re.match(r"[ \t\n\r\f\v]*([+-]?\d+)", t)
    return int(m.group(1)) if m else 0


def _node(tok):
    """getNode: v/t/n, each atoi of the text before the next '/', missing = 0, minus 1."""
    parts = tok.split("/")
    vals = []
    for k in range(3):
        s = parts[k] if k < len(parts) else ""
        vals.append((_atoi(s) if s != "" else 0) - 1)
    return vals


def read_obj(text):
    """MeshReader::readObj: (list of meshes, list of names); each mesh a list of faces,
    each face a list of (position, normal) corners."""
    T = _ObjText(text)
    verts, norms, texs, faces, objects, names = [], [], [], [], [], []
    start = 0
    while T.f[T.c] != "\0":
        s = T.get_token()
        if s == "":
            continue
        if s in ("v", "vn", "vt"):
            x, y, z = T.get_float(), T.get_float(), T.get_float()
            if s == "vt":
                texs.append((x * f32(1), y * f32(1), z * f32(1)))
            else:
                (verts if s == "v" else norms).append((x * f32(-1), y * f32(1), z * f32(1)))
        elif s == "f":
            face = []
            while True:
                t = T.get_token("", "\n")
                if t == "":
                    break
                face.append(_node(t))
            faces.append(face)
        elif s == "g":
            if len(faces) > start:
                objects.append((start, len(faces)))
                start = len(faces)
            else:
                names = []
            names.append(T.get_line())
    if len(faces) > start:
        objects.append((start, len(faces)))
    else:
        names = []
    if not objects:
        return [], names

    nv = len(verts)
    vn = [V(0, 0, 0)] * nv
    for face in faces:
        n = len(face)
        for j in range(n):
            p, q, c = face[(j - 1 + n) % n][0], face[(j + 1) % n][0], face[j][0]
            if not (0 <= p < nv and 0 <= q < nv and 0 <= c < nv):
                raise SceneError("UNABLE TO CALCULATE VERTEX NORMALS")
            vn[c] = vadd(vn[c], vcross(vsub(verts[p], verts[c]), vsub(verts[q], verts[c])))
    vn = [vnormalized(x) if (x[0] != 0 or x[1] != 0 or x[2] != 0) else x for x in vn]

    meshes = []
    for a, b in objects:
        nidx = {}
        mesh = []
        for i in range(a, b):
            face = []
            for v, t, n in faces[i]:
                if v < 0 or v >= nv:
                    raise SceneError("VERTEX ERROR")
                if n not in nidx:
                    if n >= len(norms):
                        raise SceneError("NORMAL ERROR")
                    nidx[n] = norms[n] if n >= 0 else vn[v]
                if t >= len(texs):
                    raise SceneError("TEXTURE ERROR")
                face.append((verts[v], nidx[n], texs[t] if t >= 0 else V(0, 0, 0)))
            mesh.append(face)
        meshes.append(mesh)
    return meshes, names


def bake(mesh):
    """PolyMesh::bake: triangle fans (0, prev, cur)."""
    out = []
    for face in mesh:
        prev = 1
        for cur in range(2, len(face)):
            out.append((face[0], face[prev], face[cur]))
            prev = cur
    return out


# ---------------------------------------------------------------------------- plugins
def _dumb_based_ctor(fresnel, spec, diffuse, metal):
    m = max(min(f32(metal), f32(1)), f32(0))
    return {"fresnel": tuple(f32(c) for c in fresnel), "diffuse": vmul(tuple(f32(c) for c in diffuse), f32(1) - m),
            "spec": f32(spec), "metal": m}


_PRESETS = {
    "rough_gold": ((1.022, 0.782, 0.344), 16.0, (1.0, 1.0, 1.0), 1.0),
    "glossy_gold": ((1.022, 0.782, 0.344), 128.0, (1.0, 1.0, 1.0), 1.0),
    "glossy_finish": ((0.175, 0.15, 0.125), 4096.0, (0.32, 0.32, 0.32), 0.25),
    "matte_finish": ((0.175, 0.15, 0.125), 8.0, (0.24, 0.24, 0.24), 0.05),
}
_PRESETS["gold"], _PRESETS["glossy"], _PRESETS["matte"] = (_PRESETS["glossy_gold"], _PRESETS["glossy_finish"],
                                                         _PRESETS["matte_finish"])


def _colored(d, ck, tk, tl, of, color, images=None):
    """ColoredTexture::fromDson: (colour, map) with map = None or {texture, tiling, offset}."""
    if ck in d:
        color = vector_from(d[ck])
    tex, tiling, offset = -1, (1.0, 1.0), (0.0, 0.0)
    if tk in d:
        tex = images.image_id(d[tk]) if images is not None else _no_images(d[tk])
    if tl in d:
        v = vector_from(d[tl])
        tiling = (float(v[0]), float(v[1]))
    if of in d:
        v = vector_from(d[of])
        offset = (float(v[0]), float(v[1]))
    return color, ({"texture": tex, "tiling": tiling, "offset": offset} if tex >= 0 else None)


def _no_images(v):
    if isinstance(v, (bool, float)) or v is None:
        return -1
    raise SceneError("texture outside a scene file")


def _number(d, k, cur):
    if k not in d:
        return cur
    if not _is_num(d[k]):
        raise SceneError(k + " has to be a number")
    return _fv(d[k])


def material_from(d, images=None):
    if not isinstance(d, dict):
        raise SceneError("Material should be contained in Dson::Dict")
    if "type" not in d:
        raise SceneError("Material has to have a type")
    t = d["type"]
    if not isinstance(t, Str):
        raise SceneError("Material type has to be a string")
    if t == "dumb_based":
        s = _dumb_based_ctor((0.2, 0.2, 0.2), 8.0, (0.2, 0.2, 0.2), 0.5)
        if "preset" in d:
            if not isinstance(d["preset"], Str):
                raise SceneError("DumbBasedShader preset should be of a string type")
            if d["preset"] not in _PRESETS:
                raise SceneError("DumbBasedShader preset does not exist")
            s = _dumb_based_ctor(*_PRESETS[d["preset"]])
        fres, metal, spec = s["fresnel"], s["metal"], s["spec"]
        div = f32(1) if metal == f32(1) else f32(1) - metal
        diffuse = vmul(s["diffuse"], f32(1) / div)
        fres, m_color = _colored(d, "fresnel", "fresnel_texture", "fresnel_tiling", "fresnel_offset", fres, images)
        diffuse, m_diff = _colored(d, "diffuse", "diffuse_texture", "diffuse_tiling", "diffuse_offset", diffuse,
                                   images)
        ncol, m_norm = _colored(d, "normal_color", "normal_texture", "normal_tiling", "normal_offset", V(1, 1, 1),
                                images)
        _, m_alpha = _colored(d, "alpha_cutout_color", "alpha_cutout_texture", "alpha_cutout_tiling",
                              "alpha_cutout_offset", V(1, 1, 1), images)
        thresh = _number(d, "alpha_cutout_threshold", f32(0.5))
        spec = _number(d, "specular", spec)
        metal = _number(d, "metal", metal)
        s = _dumb_based_ctor(fres, spec, diffuse, metal)
        out = {"kind": "dumb_based", "albedo": tuple(float(c) for c in s["fresnel"]),
               "diffuse": tuple(float(c) for c in s["diffuse"]), "p": (float(s["spec"]), float(s["metal"]), 0.0)}
        if any(m is not None for m in (m_color, m_diff, m_norm, m_alpha)):
            out["maps"] = {"color": m_color, "diffuse": m_diff, "normal": m_norm, "alpha": m_alpha,
                           "normal_color": tuple(float(c) for c in ncol), "alpha_threshold": float(thresh)}
        return out
    if t == "simple_stochastic":
        def ctor(color, diffuse, smooth, shine):
            total = f32(diffuse) + f32(smooth)
            mul = (f32(1) / total) if total > f32(1) else f32(1)
            return tuple(f32(c) for c in color), f32(diffuse) * mul, f32(smooth) * mul, f32(shine)
        color, diffuse, smooth, shine = ctor((1, 1, 1), 0.05, 0.95, 2.0)
        if "color" in d:
            color = vector_from(d["color"])
        diffuse = _number(d, "diffuse", diffuse)
        smooth = _number(d, "smoothness", smooth)
        shine = _number(d, "shininess", shine)
        color, diffuse, smooth, shine = ctor(color, diffuse, smooth, shine)
        return {"kind": "simple_stochastic", "albedo": tuple(float(c) for c in color),
                "p": (float(diffuse), float(smooth), float(shine))}
    if t == "fresnel_color":
        color = V(1, 1, 1)
        engines = {"iron": (0.56, 0.57, 0.58), "copper": (0.95, 0.64, 0.54), "gold": (1.0, 0.71, 0.29),
                   "aluminum": (0.91, 0.92, 0.92), "silver": (0.95, 0.93, 0.88)}
        if isinstance(d.get("engine_color"), Str) and d["engine_color"] in engines:
            color = V(*engines[d["engine_color"]])
        color, m_color = _colored(d, "color", "texture", "tiling", "offset", color, images)
        out = {"kind": "fresnel", "albedo": tuple(float(c) for c in color)}
        if m_color is not None:
            out["maps"] = {"color": m_color, "diffuse": None, "normal": None, "alpha": None,
                           "normal_color": (1.0, 1.0, 1.0), "alpha_threshold": 0.5}
        return out
    raise SceneError('Unknown material type: "' + t + '"')


def light_from(d):
    if not isinstance(d, dict):
        raise SceneError("Light should be contained in Dson::Dict")
    if "type" not in d:
        raise SceneError("Light has to have a type")
    t = d["type"]
    if not isinstance(t, Str):
        raise SceneError("Light type has to be a string")
    if t == "simple_soft_directional":
        shade = V(1, 1, 1)
        direction = vnormalized(vnormalized(V(1, -1, 1)))
        dist = f32(1024)
        soft = dist * f32(0.125)
        distance = dist
        softness = soft / (f32(0) if distance == f32(0) else distance)
        if "color" in d:
            shade = vector_from(d["color"])
        if "direction" in d:
            direction = vector_from(d["direction"])
        distance = _number(d, "distance", distance)
        softness = _number(d, "softness", softness)
        u = vnormalized(direction)
        return {"kind": "soft_directional", "emission": tuple(float(c) for c in shade),
                "edge_u": tuple(float(c) for c in u), "dist": float(distance), "radius": float(distance * softness)}
    if t == "simple_spherical":
        shade, pos, radius = V(1, 1, 1), V(0, 0, 0), f32(1)
        if "color" in d:
            shade = vector_from(d["color"])
        if "position" in d:
            pos = vector_from(d["position"])
        radius = _number(d, "radius", radius)
        return {"kind": "spherical", "emission": tuple(float(c) for c in shade),
                "corner": tuple(float(c) for c in pos), "radius": float(radius)}
    if t == "spotlight":                  # Spotlight.impl.cuh:95-161, constructor :7-38
        shade, pos, direction = V(1, 1, 1), V(0, 0, 0), V(0, 0, 1)
        if "color" in d:
            shade = vector_from(d["color"])
        lum = _number(d, "luminosity", f32(1500))
        if "position" in d:
            pos = vector_from(d["position"])
        if "target" in d:
            direction = vsub(vector_from(d["target"]), pos)
        elif "direction" in d:
            direction = vector_from(d["direction"])
        inner = _number(d, "hotspot_angle", f32(16))
        outer = _number(d, "falloff_angle", f32(64))
        power = _number(d, "falloff_power", f32(1))
        size = _number(d, "emitter_size", f32(0.25))
        sphere = False
        if "emitter_shape" in d:
            if not isinstance(d["emitter_shape"], Str):
                raise SceneError("Spotlight emitter shape must be a string")
            if d["emitter_shape"] in ("disc", "sphere"):
                sphere = d["emitter_shape"] == "sphere"
        shadows = True
        if "cast_shadows" in d:
            if not isinstance(d["cast_shadows"], bool):
                raise SceneError("Spotlight shadow cast flag must be a boolean")
            shadows = d["cast_shadows"]
        samples = 1
        if "sample_count" in d:
            if not _is_num(d["sample_count"]):
                raise SceneError("Spotlight sample count must be a number")
            samples = _iv(d["sample_count"])
        eps32 = f32(32) * f32(0.000005)
        if size < eps32:
            size = eps32
        if inner < 0:
            inner = -inner
        if inner > 180:
            inner = f32(180)
        if outer < 0:
            outer = -inner
        if outer > 180:
            outer = f32(180)
        if outer < inner:
            outer = inner
        pi = f32(3.14159265359)
        flags = ((16 if samples > 16 else samples) if samples > 0 else 0) | (32 if shadows else 0) | (64 if sphere else 0)
        return {"kind": "spot", "emission": tuple(float(c * lum) for c in shade), "corner": tuple(float(c) for c in pos),
                "edge_u": tuple(float(c) for c in vnormalized(direction)),
                "edge_v": (float(cosf((inner * pi) / f32(180))), float(cosf((outer * pi) / f32(180))), float(power)),
                "radius": float(size), "flags": flags}
    raise SceneError('unsupported or unknown light type: "' + t + '"')


def lens_focal(angle):
    """SimpleStochasticLense(angle): x = 0.5 / tan(angle / (2 * RADIAN)), angle clamped to [1, 179]."""
    a = f32(angle)
    if a < 1:
        a = f32(1)
    elif a > 179:
        a = f32(179)
    return f32(0.5) / tanf(a / (f32(2) * RADIAN))


# ---------------------------------------------------------------------------- context
class SceneFile:
    def __init__(self):
        self.source_path = ""
        self.material_names = {}
        self.obj_files = {}
        self.groups = {}
        self.materials, self.lights = [], []
        self.units = []                        # (verts (3 tuples), normals (3 tuples), material id)
        self.camera = Transform()
        self.lens_x = None
        self.max_bounces, self.fsaa, self.boxing, self.ignore_backfaces = 2, (1, 1), 0, True
        self.lens, self.lens_params = 0, [f32(0)] * 8
        self.textures, self.texture_names = [], {}

    # -------------------------------------------------------------- files
    def load(self, path, depth=0):
        if depth > 32:
            raise SceneError("include depth")
        try:
            with open(path, "rb") as fh:
                text = fh.read().decode("latin-1")
        except OSError:
            raise SceneError('Could not open file: "' + path + '"') from None
        k = len(path)
        while k > 0 and path[k - 1] not in "/\\":
            k -= 1
        self.parse(text, path[:k], depth)
        return self

    def parse(self, text, base_dir="", depth=0):
        root = parse_dson(text)
        saved, self.source_path = self.source_path, base_dir
        try:
            self.from_dson(root, depth)
        finally:
            self.source_path = saved
        return self

    def from_dson(self, d, depth):
        if not isinstance(d, dict):
            raise SceneError("Render context can be constructed only from a Dson::Dict")
        if "include" in d:
            inc = d["include"]
            if not isinstance(inc, list):
                raise SceneError("Include should be a list of strings")
            for name in inc:
                if not isinstance(name, Str):
                    raise SceneError("Can not include non-string objects")
                p = self.source_path + name
                self.load(p if os.path.isfile(p) else name, depth + 1)
        for key, fn in (("materials", self.add_material), ("lights", self.add_light), ("groups", self.add_group),
                        ("objects", self.add_object)):
            if key in d:
                if not isinstance(d[key], list):
                    raise SceneError(key + " should be contained in Dson::Array")
                for e in d[key]:
                    fn(e)
        if "camera" in d:
            self.parse_camera(d["camera"])
        if self.lens_x is None:
            self.lens_x = lens_focal(64.0)
        if "renderer" in d:
            self.parse_renderer(d["renderer"])

    def load_obj(self, name):
        key = "obj::" + name
        if key in self.obj_files:
            return
        p = self.source_path + name
        path = p if os.path.isfile(p) else name
        try:
            with open(path, "rb") as fh:
                text = fh.read().decode("latin-1")
        except OSError:
            raise SceneError("Could not read file: '" + name + "'") from None
        meshes, names = read_obj(text)
        if len(meshes) != len(names):
            raise SceneError("File '" + name + "' does not have equal amount of names and meshes")
        self.obj_files[key] = dict(zip(names, meshes))

    # -------------------------------------------------------------- sections
    def image_id(self, v):
        """DumbRenderContext::getImageId: -1, a texture named before, or a new texture from a PNG
        file (decoded with Pillow to RGBA, byte / 255)."""
        if isinstance(v, (bool, float)) or v is None:
            return -1
        if isinstance(v, Str):
            for key in ("name::" + v, "png::" + v):
                if key in self.texture_names:
                    return self.texture_names[key]
            raise SceneError('Texture "' + v + '" not found')
        if not isinstance(v, dict):
            raise SceneError("Unsupported dson type for texture")
        if "png" not in v:
            raise SceneError("Image dict incomplete")
        if not isinstance(v["png"], Str):
            raise SceneError("Image 'png' entry MUST BE a string")
        from PIL import Image

        name = v["png"]
        path = self.source_path + name
        if not os.path.isfile(path):
            path = name
        try:
            rgba = np.asarray(Image.open(path).convert("RGBA"))
        except Exception:
            raise SceneError('Could not read file: "' + path + '"') from None
        tid = len(self.textures)
        self.textures.append({"rgba": rgba.astype(np.float32) / f32(255), "filter": 1})
        self.texture_names["png::" + path] = tid
        self.texture_names["png::" + name] = tid
        if "filtering" in v:
            if not isinstance(v["filtering"], Str) or v["filtering"] not in ("none", "bilinear"):
                raise SceneError("Image filter can be only none/[bilinear]")
            self.textures[tid]["filter"] = 0 if v["filtering"] == "none" else 1
        if "name" in v:
            if not isinstance(v["name"], Str):
                raise SceneError("Image name entry MUST BE a string")
            self.texture_names["name::" + v["name"]] = tid
        return tid

    def add_material(self, d):
        m = material_from(d, self)
        self.materials.append(m)
        mid = len(self.materials) - 1
        if isinstance(d.get("name"), Str):
            self.material_names[d["name"]] = mid
        return mid

    def material_id(self, v):
        if isinstance(v, Str):
            if v not in self.material_names:
                raise SceneError('Material not found: "' + v + '"')
            return self.material_names[v]
        return self.add_material(v)

    def add_light(self, d):
        self.lights.append(light_from(d))

    def item_from(self, d):
        if not isinstance(d, dict):
            raise SceneError("Group item can only be made from a dictionary")
        if "mesh" not in d:
            raise SceneError("Item has to have a 'mesh' entry")
        m = d["mesh"]
        if not isinstance(m, dict):
            raise SceneError("Item 'mesh' entry has to be a dict")
        if "obj" in m:
            if not isinstance(m["obj"], Str):
                raise SceneError(".obj file name must be a string")
            self.load_obj(m["obj"])
            item = {"file": "obj::" + m["obj"], "object": "", "t": Transform(), "mat": -1}
        elif "primitive" in m:
            raise SceneError("primitive meshes are not implemented")
        else:
            raise SceneError("Object mesh could not be parsed")
        if "object" in m:
            if not isinstance(m["object"], Str):
                raise SceneError(".obj mesh name must be a string")
            if m["object"] not in self.obj_files[item["file"]]:
                raise SceneError("Object [" + m["object"] + "] not found")
            item["object"] = m["object"]
        if "transform" in d:
            item["t"] = transform_from(d["transform"])
        if "material" in d:
            item["mat"] = self.material_id(d["material"])
        return item

    def sub_from(self, d):
        sg = {"group": None, "t": Transform(), "fallback": -1, "override": -1}
        if isinstance(d, Str):
            sg["group"] = "named::" + d
            if sg["group"] not in self.groups:
                raise SceneError("Group [" + d + "] not found")
            return sg
        if not isinstance(d, dict):
            raise SceneError("Subgroup should be expressed either as a string, or a dict")
        if "group" in d:
            if not isinstance(d["group"], Str):
                raise SceneError("Subgroup 'group' must be a string")
            sg["group"] = "named::" + d["group"]
            if sg["group"] not in self.groups:
                raise SceneError("Group [" + sg["group"] + "] not found")
            if "transform" in d:
                sg["t"] = transform_from(d["transform"])
            if "fallback_material" in d:
                sg["fallback"] = self.material_id(d["fallback_material"])
            if "override_material" in d:
                sg["override"] = self.material_id(d["override_material"])
        else:
            sg["group"] = self.add_group(d)
        return sg

    def add_group(self, d):
        if not isinstance(d, dict):
            raise SceneError("Group can only be made from a dictionary")
        g = {"items": [], "subs": [], "t": Transform(), "fallback": -1, "override": -1}
        if "elements" in d:
            if not isinstance(d["elements"], list):
                raise SceneError("Group elements should be contained in an array")
            for e in d["elements"]:
                if isinstance(e, dict) and "mesh" in e:
                    g["items"].append(self.item_from(e))
                else:
                    g["subs"].append(self.sub_from(e))
        if "transform" in d:
            g["t"] = transform_from(d["transform"])
        if "fallback_material" in d:
            g["fallback"] = self.material_id(d["fallback_material"])
        if "override_material" in d:
            g["override"] = self.material_id(d["override_material"])
        if "name" in d:
            if not isinstance(d["name"], Str):
                raise SceneError("Group name must be a string")
            name = "named::" + d["name"]
        else:
            name = "unnamed::" + str(len(self.groups))
        self.groups[name] = g
        return name

    def resolve_item(self, item):
        meshes = self.obj_files[item["file"]]
        chosen = [meshes[item["object"]]] if item["object"] else list(meshes.values())
        out, t = [], item["t"]
        for mesh in chosen:
            for tri in bake(mesh):
                out.append([[t.point(c[0]) for c in tri], [t.normal(c[1]) for c in tri], item["mat"],
                            [c[2] for c in tri]])
        return out

    @staticmethod
    def _finish(units, t, override, fallback):
        for u in units:
            if override >= 0:
                u[2] = override
            elif u[2] < 0:
                u[2] = fallback
            u[0] = [t.point(v) for v in u[0]]
            u[1] = [t.normal(n) for n in u[1]]

    def resolve_group(self, name, depth=0):
        if depth > 64:
            raise SceneError("group nesting")
        g = self.groups[name]
        out = []
        for it in g["items"]:
            out += self.resolve_item(it)
        for sg in g["subs"]:
            part = self.resolve_group(sg["group"], depth + 1)
            self._finish(part, sg["t"], sg["override"], sg["fallback"])
            out += part
        self._finish(out, g["t"], g["override"], g["fallback"])
        return out

    def add_object(self, d):
        if isinstance(d, dict) and "mesh" in d:
            item = self.item_from(d)
            if item["mat"] < 0:
                raise SceneError("Object MUST have a material attached to it")
            units = self.resolve_item(item)
        else:
            sg = self.sub_from(d)
            units = self.resolve_group(sg["group"])
            self._finish(units, sg["t"], sg["override"], sg["fallback"])
            if any(u[2] < 0 for u in units):
                raise SceneError("Resolved group must have a material attached to it")
        self.units += units

    def parse_camera(self, d):
        if not isinstance(d, dict):
            raise SceneError("Camera should be contained in Dson::Dict")
        if "lense" in d:
            lz = d["lense"]
            if not isinstance(lz, dict):
                raise SceneError("Camera lense has to be a Dson::Dict")
            if "type" not in lz:
                raise SceneError("Lense has to have a type")
            if not isinstance(lz["type"], Str):
                raise SceneError("Lense type has to be a string")
            if lz["type"] == "simple_stochastic":
                x = lens_focal(60.0)
                if "angle" in lz:
                    if not _is_num(lz["angle"]):
                        raise SceneError("SimpleStochasticLense angle can only be a number")
                    x = lens_focal(_fv(lz["angle"]))
                self.lens_x = x
                self.lens, self.lens_params = 0, [f32(0)] * 8
            elif lz["type"] == "spherical_segment":    # SphericalSegmentLense.cuh:9, .impl.cuh:38-70
                ax = ay = f32(60)
                if "angle" in lz:
                    if not _is_num(lz["angle"]):
                        raise SceneError("SimpleFocalPlaneLense angle must be a number")
                    ax = ay = _fv(lz["angle"])
                dist = _number(lz, "distance", f32(128))
                soft = _number(lz, "sensor", f32(1))
                filt = vector_from(lz["filter"]) if "filter" in lz else V(1, 1, 1)
                if self.lens_x is None:
                    self.lens_x = f32(0)
                self.lens, self.lens_params = 2, [ax, ay, dist, soft, filt[0], filt[1], filt[2], f32(0)]
            else:
                raise SceneError('unknown lense type: "' + lz["type"] + '"')
        if "transform" in d:
            self.camera = transform_from(d["transform"])

    def parse_renderer(self, d):
        if not isinstance(d, dict):
            raise SceneError("Renderer should be contained in Dson::Dict")
        if "resources" in d:
            r = d["resources"]
            if not isinstance(r, dict):
                raise SceneError("Renderer Resources has to be a Dson::Dict type")
            if "cpu" in r:
                c = r["cpu"]
                if not (_is_num(c) or (isinstance(c, Str) and c in ("all", "all_but_gpu", "all_but_one_per_gpu",
                                                                      "none", "one"))):
                    raise SceneError("Renderer Resources CPU")
            if "gpu" in r and not _is_num(r["gpu"]):
                raise SceneError("Renderer Resources GPU can only be a number")
        if "blocks" in d:
            b = d["blocks"]
            if not isinstance(b, dict):
                raise SceneError("Renderer Blocks has to be a Dson::Dict type")
            for k in ("cpu_cut", "gpu_cut"):
                if k in b and not _is_num(b[k]):
                    raise SceneError("Renderer Blocks " + k + " has to be a number")
            if "force_host_block_synchronisation" in b and not isinstance(b["force_host_block_synchronisation"], bool):
                raise SceneError("force_host_block_synchronisation has to be a boolean")
        if "pixel" in d:
            p = d["pixel"]
            if not isinstance(p, dict):
                raise SceneError("Renderer Pixel has to be a Dson::Dict type")
            if "boxing" in p:
                modes = {"height": 0, "width": 1, "min": 2, "max": 3}
                if not isinstance(p["boxing"], Str) or p["boxing"] not in modes:
                    raise SceneError("Renderer Pixel boxing")
                self.boxing = modes[p["boxing"]]
            if "bounces" in p:
                if not _is_num(p["bounces"]):
                    raise SceneError("Renderer Pixel bounces has to be a number")
                self.max_bounces = _iv(p["bounces"])
            if "samples_per_pixel" in p:
                s = p["samples_per_pixel"]
                if not (isinstance(s, list) and len(s) == 2 and _is_num(s[0]) and _is_num(s[1])):
                    raise SceneError("samples_per_pixel has to be an array, containing two numbers")
                self.fsaa = (_iv(s[0]), _iv(s[1]))
            if "pixels_per_gpu_thread" in p and not _is_num(p["pixels_per_gpu_thread"]):
                raise SceneError("pixels_per_gpu_thread has to be a number")
            if "ignore_backfaces" in p:
                if not isinstance(p["ignore_backfaces"], bool):
                    raise SceneError("ignore_backfaces has to have a boolean value")
                self.ignore_backfaces = p["ignore_backfaces"]

    # -------------------------------------------------------------- outputs
    @property
    def tris(self):
        return np.array([[c for v in u[0] for c in v] for u in self.units], np.float32).reshape(-1, 9)

    @property
    def normals(self):
        return np.array([[c for v in u[1] for c in v] for u in self.units], np.float32).reshape(-1, 9)

    @property
    def mat_id(self):
        return np.array([u[2] for u in self.units], np.int32)

    @property
    def uv(self):
        return np.array([[c for t in u[3] for c in t[:2]] for u in self.units], np.float32).reshape(-1, 6)

    @property
    def cam_rows(self):
        t = self.camera
        return np.array(list(t.tx) + list(t.ty) + list(t.tz), np.float32)

    @property
    def cam(self):
        o = self.camera.point(V(0, 0, 0))
        vals = list(o) + list(self.camera.right()) + list(self.camera.up()) + list(self.camera.front()) + [self.lens_x]
        return np.array(vals, np.float32)


def load(path):
    return SceneFile().load(path)


def parse(text, base_dir=""):
    return SceneFile().parse(text, base_dir)
