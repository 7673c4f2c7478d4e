"""CPU restatement of the reference's asset readers (test infrastructure: only tests may import it).

- `parse_obj`: ObjLoader.Parse / ParseMtl / ReadFace / TriangulateFace / ObjLexer (obj.cs:19-512),
  token for token, returning per object the fanned triangles (positions, vertex normals or None per
  triangle, UVs with the OBJ defaults) and the material record, before the Triangle constructors.
- `decode_hdr`: ImageHdr.LoadHdr (image.cs:307-397) + Color.FromPixel(PixelHdr) (math.cs:196-202)
  + Texture.FromHdr's luminance clamp (texture.cs:154-163).
- `decode_tga`: Image.LoadTga (image.cs:217-266).
Pure Python on small inputs; pinned by hand-built files with known answers in tests/test_assets.py.
"""
import math
import os
import struct

import numpy as np


class ObjError(Exception):
    pass


# ---- ObjLexer (obj.cs:19-125) ------------------------------------------------------------------
WORD, SLASH, NEWLINE, END = "Word", "Slash", "Newline", "End"


class _Lexer:
    def __init__(self, text):
        self.s = text
        self.i = 0
        self.peeked = None  # (kind, word)

    def _read(self):
        s, n = self.s, len(self.s)
        while True:
            if self.i >= n:
                return END, ""
            ch = s[self.i]
            self.i += 1
            if ch == "\n":
                return NEWLINE, ""
            if ch == "/":
                return SLASH, ""
            if ch in " \t\r":
                continue
            if ch == "#":
                while True:
                    if self.i >= n:
                        return END, ""
                    c = s[self.i]
                    self.i += 1
                    if c == "\n":
                        return NEWLINE, ""
            j = self.i
            while j < n and s[j] not in " \t\r\n/":
                j += 1
            word = ch + s[self.i:j]
            self.i = j
            return WORD, word

    def peek(self):
        if self.peeked is None:
            self.peeked = self._read()
        return self.peeked[0]

    def next(self):
        if self.peeked is not None:
            t, self.peeked = self.peeked, None
            return t
        return self._read()

    def skip_line(self):
        while self.next()[0] not in (NEWLINE, END):
            pass


def _float(lx):
    kind, w = lx.next()
    if kind != WORD:
        raise ObjError(f"OBJ: expected number, got {kind}")
    body = w[1:] if w[:1] in "+-" else w
    if body.lower() in ("infinity", "nan"):
        return float(w)
    mant, _, exp = body.replace("E", "e").partition("e")
    ip, _, fp = mant.partition(".")
    ok = (ip + fp).isdigit() and (ip == "" or ip.isdigit()) and (fp == "" or fp.isdigit())
    if "e" in body.lower():
        e = exp[1:] if exp[:1] in "+-" else exp
        ok = ok and e.isdigit()
    if not ok:
        raise ObjError(f"OBJ: invalid number '{w}'")
    return float(np.float32(float(w)))


def _int(lx):
    kind, w = lx.next()
    if kind != WORD:
        raise ObjError(f"OBJ: expected int, got {kind}")
    body = w[1:] if w[:1] in "+-" else w
    if not body.isdigit() or not (-2 ** 31 <= int(w) < 2 ** 31):
        raise ObjError(f"OBJ: invalid int '{w}'")
    return int(w)


def _index(lx, count):
    idx = _int(lx)
    idx = count + idx if idx < 0 else idx - 1
    if idx < 0 or idx >= count:
        raise ObjError(f"OBJ: index out of bounds ({idx}, count: {count})")
    return idx


def _default_material():
    return dict(albedo=[1.0, 1.0, 1.0], roughness=1.0, metallic=0.0, transparency=0.0, ior=1.5,
                radiance=[0.0, 0.0, 0.0], maps={})


def parse_mtl(text):
    """ObjLoader.ParseMtl (obj.cs:260-409): name -> material dict (maps: slot -> file name)."""
    out = {}
    lx = _Lexer(text)
    st = {"name": None}

    def reset():
        st.update(color=[1.0, 1.0, 1.0], radiance=[0.0, 0.0, 0.0], roughness=1.0, metallic=0.0, transparency=0.0,
                  ior=1.5, hasPr=False, ns=0.0, maps={})

    def build():
        if st["name"] is None:
            return
        r = st["roughness"] if st["hasPr"] else float(np.sqrt(np.float32(2) / (np.float32(st["ns"]) + np.float32(2))))
        r = min(max(r, 0.0), 1.0)
        out[st["name"]] = dict(albedo=st["color"], roughness=r, metallic=st["metallic"],
                               transparency=st["transparency"], ior=st["ior"], radiance=st["radiance"],
                               maps=dict(st["maps"]))

    reset()
    while True:
        kind, w = lx.next()
        if kind == END:
            break
        if kind != WORD:
            continue
        if w == "newmtl":
            build()
            st["name"] = lx.next()[1]
            reset()
        elif w == "Kd":
            st["color"] = [_float(lx), _float(lx), _float(lx)]
        elif w == "Ke":
            st["radiance"] = [_float(lx), _float(lx), _float(lx)]
        elif w == "Pr":
            st["roughness"], st["hasPr"] = _float(lx), True
        elif w == "Pm":
            st["metallic"] = _float(lx)
        elif w == "Ns":
            st["ns"] = _float(lx)
        elif w == "Ni":
            st["ior"] = _float(lx)
        elif w == "d":
            st["transparency"] = float(np.float32(1.0) - np.float32(_float(lx)))
        elif w == "Tr":
            st["transparency"] = _float(lx)
        elif w in ("map_Kd", "map_Pm", "map_Pr", "map_roughness", "map_bump", "norm", "map_Kn"):
            slot = {"map_Kd": "color", "map_Pm": "metallic", "map_Pr": "roughness", "map_roughness": "roughness",
                    "map_bump": "normal", "norm": "normal", "map_Kn": "normal"}[w]
            st["maps"][slot] = lx.next()[1]
        lx.skip_line()
    build()
    return out


def parse_obj(text, mtl_files=None, default_material=None):
    """ObjLoader.Parse (obj.cs:153-258). mtl_files: name -> MTL text (stands in for the files next
    to the OBJ). Returns a list of objects: dict(pos (n, 9), nrm (n, 9), has_nrm (n,), uv (n, 6),
    material, smooth) where smooth says whether the file had no vn when the object was built."""
    positions, normals, uvs = [], [], []
    tri = {"pos": [], "nrm": [], "has": [], "uv": []}
    materials = {}
    state = {"mat": None}
    objects = []

    def build_object():
        if not tri["has"]:
            return
        mat = materials.get(state["mat"]) if state["mat"] is not None else None
        if mat is None:
            mat = default_material or _default_material()
        objects.append(dict(pos=np.array(tri["pos"], np.float32).reshape(-1, 9),
                            nrm=np.array(tri["nrm"], np.float32).reshape(-1, 9),
                            has_nrm=np.array(tri["has"], bool), uv=np.array(tri["uv"], np.float32).reshape(-1, 6),
                            material=mat, smooth=len(normals) == 0))
        for v in tri.values():
            v.clear()

    lx = _Lexer(text)
    while True:
        kind, w = lx.next()
        if kind == END:
            break
        if kind != WORD:
            continue
        if w == "v":
            positions.append([_float(lx), _float(lx), _float(lx)])
            lx.skip_line()
        elif w == "vt":
            u, v = _float(lx), _float(lx)
            uvs.append([u, float(np.float32(1.0) - np.float32(v))])
            lx.skip_line()
        elif w == "vn":
            normals.append([_float(lx), _float(lx), _float(lx)])
            lx.skip_line()
        elif w == "f":
            face = []
            while lx.peek() == WORD:
                p = _index(lx, len(positions))
                n = t = -1
                if lx.peek() == SLASH:
                    lx.next()
                    if lx.peek() == WORD:
                        t = _index(lx, len(uvs))
                    if lx.peek() == SLASH:
                        lx.next()
                        if lx.peek() == WORD:
                            n = _index(lx, len(normals))
                face.append((p, n, t))
            if len(face) < 3:
                raise ObjError("OBJ: face has fewer than 3 vertices")
            a = face[0]
            for i in range(2, len(face)):
                b, c = face[i - 1], face[i]
                tri["pos"] += positions[a[0]] + positions[b[0]] + positions[c[0]]
                tri["uv"] += (uvs[a[2]] if a[2] >= 0 else [0.0, 0.0]) + (uvs[b[2]] if b[2] >= 0 else [1.0, 0.0]) + \
                    (uvs[c[2]] if c[2] >= 0 else [0.0, 1.0])
                has = a[1] >= 0 and b[1] >= 0 and c[1] >= 0
                tri["nrm"] += (normals[a[1]] + normals[b[1]] + normals[c[1]]) if has else [0.0] * 9
                tri["has"].append(has)
        elif w in ("o", "g"):
            build_object()
            if lx.peek() == WORD:
                lx.next()
            lx.skip_line()
        elif w == "mtllib":
            name = lx.next()[1]
            if mtl_files and name in mtl_files:
                materials.update(parse_mtl(mtl_files[name]))
            lx.skip_line()
        elif w == "usemtl":
            build_object()
            state["mat"] = lx.next()[1]
            lx.skip_line()
        else:
            lx.skip_line()
    build_object()
    if not objects:
        raise ObjError("OBJ: no objects found")
    return objects


# ---- Images ------------------------------------------------------------------------------------
def decode_tga(data):
    """Image.LoadTga (image.cs:217-266): (h, w, 3) uint8, top row first."""
    id_len, cmap, itype = data[0], data[1], data[2]
    w, h = struct.unpack_from("<HH", data, 12)
    bpp, desc = data[16], data[17]
    if cmap != 0:
        raise ValueError("TGA: Color-mapped images not supported")
    if itype not in (2, 3):
        raise ValueError(f"TGA: Unsupported image type {itype}")
    if (itype == 2 and bpp != 24) or (itype == 3 and bpp != 8):
        raise ValueError("TGA: unsupported bit depth")
    px = np.frombuffer(bytes(data[18:18 + w * h * (3 if itype == 2 else 1)]), np.uint8)
    img = np.repeat(px.reshape(h, w, 1), 3, 2) if itype == 3 else px.reshape(h, w, 3)[..., ::-1]
    if not desc & 0x20:
        img = img[::-1]
    return np.ascontiguousarray(img)


def decode_hdr(data, max_luminance=None):
    """ImageHdr.LoadHdr + Texture.FromHdr: (h, w, 3) float32."""
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].replace(b"\r", b"").decode("latin-1")
        pos = end + 1
        return s

    magic = line()
    if not (magic.startswith("#?RADIANCE") or magic.startswith("#?RGBE")):
        raise ValueError(f"ImageHdr: Invalid magic '{magic}'")
    while line():
        pass
    parts = line().split(" ")
    if len(parts) != 4 or parts[0] != "-Y" or parts[2] != "+X":
        raise ValueError("ImageHdr: Unsupported size/orientation")
    h, w = int(parts[1]), int(parts[3])
    rgbe = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        hd = data[pos:pos + 4]
        pos += 4
        if hd[0] == 2 and hd[1] == 2 and not hd[2] & 0x80:
            if (hd[2] << 8 | hd[3]) != w:
                raise ValueError("ImageHdr: Scanline width mismatch")
            for c in range(4):
                x = 0
                while x < w:
                    code = data[pos]
                    pos += 1
                    if code > 128:
                        rgbe[y, x:x + code - 128, c] = data[pos]
                        pos += 1
                        x += code - 128
                    else:
                        rgbe[y, x:x + code, c] = np.frombuffer(data[pos:pos + code], np.uint8)
                        pos += code
                        x += code
        else:
            if hd[0] == 1 and hd[1] == 1 and hd[2] == 1:
                raise ValueError("ImageHdr: Old RLE format not supported")
            rgbe[y] = np.frombuffer(data[pos - 4:pos - 4 + w * 4], np.uint8).reshape(w, 4)
            pos += w * 4 - 4
    out = np.zeros((h, w, 3), np.float32)
    for y in range(h):
        for x in range(w):
            r, g, b, e = (int(v) for v in rgbe[y, x])
            if e == 0:
                continue
            scale = np.float32(math.ldexp(1.0, e - 128)) / np.float32(256.0)
            c = [np.float32(r) * scale, np.float32(g) * scale, np.float32(b) * scale]
            if max_luminance is not None:
                lum = np.float32(0.2126) * c[0] + np.float32(0.7152) * c[1] + np.float32(0.0722) * c[2]
                if lum > np.float32(max_luminance):
                    k = np.float32(max_luminance) / lum
                    c = [v * k for v in c]
            out[y, x] = c
    return out


def write_files(directory, files):
    """Writes name -> text/bytes into directory (helper for tests)."""
    for name, content in files.items():
        mode = "wb" if isinstance(content, (bytes, bytearray)) else "w"
        with open(os.path.join(directory, name), mode) as f:
            f.write(content)
