"""The C# binding documented in INTEGRATION.md against include/sharpray.h: every [StructLayout]
struct has the C struct's fields in the same order with the same names and widths, and every
[DllImport] has the C prototype's parameters in the same order with compatible types. A host that
copies the documented binding then passes exactly what the library reads (a missing field would
make the library read past the managed struct)."""
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Blittable C# value types of the reference (math.cs, image.cs) as scalar lists.
CS_VALUE = {"float": ["f32"], "uint": ["u32"], "int": ["i32"], "long": ["i64"], "ulong": ["u64"],
            "Vec2": ["f32"] * 2, "Vec3": ["f32"] * 3, "Color": ["f32"] * 3, "Quat": ["f32"] * 4}
C_SCALAR = {"float": "f32", "uint32_t": "u32", "int32_t": "i32", "int64_t": "i64", "uint64_t": "u64",
            "double": "f64"}
# Element type a C# pointer parameter passes (Color / Vec3 / Vec2 arrays are float arrays).
CS_PTR_ELEM = {"float": "float", "Color": "float", "Vec3": "float", "Vec2": "float", "uint": "uint32_t",
               "int": "int32_t", "long": "int64_t", "double": "double", "Pixel": "uint8_t"}
CS_BY_VALUE = {"int": "int32_t", "uint": "uint32_t", "float": "float", "ulong": "uint64_t", "long": "int64_t"}


def snake(name):
    return re.sub(r"(?<=[a-z0-9])([A-Z])", r"_\1", name).lower()


def _csharp():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```csharp\n(.*?)```", text, re.S)
    code = "\n".join(blocks)
    return re.sub(r"//[^\n]*", "", code)


def _header():
    text = open(os.path.join(ROOT, "include", "sharpray.h")).read()
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def cs_structs():
    out = {}
    for name, body in re.findall(r"\[StructLayout\(LayoutKind\.Sequential\)\]\s*public struct (\w+)\s*\{(.*?)\}",
                                 _csharp(), re.S):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            m = re.match(r"public\s+(\w+)\s+(.*)$", decl, re.S)
            assert m, f"{name}: cannot parse '{decl}'"
            for n in m.group(2).split(","):
                fields.append((m.group(1), n.strip()))
        out[name] = fields
    return out


def c_structs():
    out = {}
    for name, body in re.findall(r"typedef struct (sr_\w+) \{(.*?)\} \1;", _header(), re.S):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            m = re.match(r"(const\s+)?(\w+)\s*(\*?)\s*(\w+)(?:\[(\d+)\])?$", decl)
            assert m, f"{name}: cannot parse '{decl}'"
            fields.append((m.group(2) + m.group(3), m.group(4), int(m.group(5) or 1)))
        out[name] = fields
    return out


def flat_cs(name, structs):
    kinds = []
    for typ, _ in structs[name]:
        if typ in CS_VALUE:
            kinds += CS_VALUE[typ]
        elif typ in structs:
            kinds += flat_cs(typ, structs)
        else:
            raise AssertionError(f"{name}: unknown C# field type {typ}")
    return kinds


def flat_c(name, structs):
    kinds = []
    for typ, _, count in structs[name]:
        if typ.endswith("*"):
            kinds += ["ptr"] * count
        elif typ in C_SCALAR:
            kinds += [C_SCALAR[typ]] * count
        else:
            kinds += flat_c(typ, structs) * count
    return kinds


def test_every_documented_struct_matches_the_header():
    cs, c = cs_structs(), c_structs()
    assert {"RenderParams", "Material", "View", "DenoiseParams", "ComposeParams"} <= set(cs)
    for name, fields in cs.items():
        cname = "sr_" + snake(name)
        assert cname in c, f"C# struct {name} has no {cname} in include/sharpray.h"
        assert [snake(n) for _, n in fields] == [n for _, n, _ in c[cname]], f"{name}: field names or order"
        assert flat_cs(name, cs) == flat_c(cname, c), f"{name}: field widths"


def test_render_params_carries_every_field():
    """The regression this test exists for: RenderParams once lacked LightSampler, so the library
    read 4 bytes past the managed struct."""
    cs = cs_structs()
    assert [n for _, n in cs["RenderParams"]][-1] == "LightSampler"
    assert len(flat_cs("RenderParams", cs)) == len(flat_c("sr_render_params", c_structs())) == 12


def c_prototypes():
    out = {}
    for ret, name, args in re.findall(r"^(int|const char\s*\*)\s+(sr_\w+)\((.*?)\);", _header(), re.S | re.M):
        params = []
        for a in args.split(","):
            a = " ".join(a.split())
            if not a or a == "void":
                continue
            m = re.match(r"(const\s+)?(\w+)\s*(\**)\s*(\w+)$", a)
            assert m, f"{name}: cannot parse '{a}'"
            params.append(((m.group(1) or "").strip(), m.group(2), m.group(3)))
        out[name] = (ret.replace(" ", ""), params)
    return out


def cs_imports():
    out = {}
    for ret, name, args in re.findall(r"\[DllImport\(Lib\)\]\s*public static extern (\w+) (sr_\w+)\((.*?)\);",
                                      _csharp(), re.S):
        params = []
        for a in args.split(","):
            a = " ".join(a.split())
            if not a:
                continue
            m = re.match(r"(?:(in|out|ref)\s+)?(\w+)(\*?)\s+(\w+)$", a)
            assert m, f"{name}: cannot parse '{a}'"
            params.append((m.group(1) or "", m.group(2), m.group(3)))
        out[name] = (ret, params)
    return out


def compatible(cs, c, structs):
    mod, typ, star = cs
    cconst, ctyp, cstars = c
    if typ == "IntPtr":
        return ctyp == "sr_ctx" and cstars == ("**" if mod == "out" else "*")
    if typ == "string":
        return cconst == "const" and ctyp == "char" and cstars == "*"
    if typ in structs:  # in / out / pointer to a documented struct
        if ctyp != "sr_" + snake(typ) or cstars != "*":
            return False
        return (mod == "in") == (cconst == "const") if mod else True
    if mod == "out":
        return cstars == "*" and ctyp == CS_BY_VALUE.get(typ) and not cconst
    if star:
        return cstars == "*" and ctyp == CS_PTR_ELEM.get(typ)
    return cstars == "" and ctyp == CS_BY_VALUE.get(typ)


def test_every_documented_import_matches_its_prototype():
    protos, imports, structs = c_prototypes(), cs_imports(), cs_structs()
    assert len(imports) >= 30
    for name, (ret, params) in imports.items():
        assert name in protos, f"{name} is not declared in include/sharpray.h"
        cret, cparams = protos[name]
        assert (ret, cret) in (("int", "int"), ("IntPtr", "constchar*")), f"{name}: return type"
        assert len(params) == len(cparams), f"{name}: {len(params)} parameters vs {len(cparams)} in C"
        for i, (a, b) in enumerate(zip(params, cparams)):
            assert compatible(a, b, structs), f"{name} parameter {i}: C# {a} vs C {b}"


def test_checker_catches_a_dropped_field(monkeypatch):
    """The comparison fails on the documented struct with its last field taken out."""
    text = _csharp()
    broken = re.sub(r"public uint LightSampler;\s*", "", text)
    assert broken != text
    monkeypatch.setattr(sys.modules[__name__], "_csharp", lambda: broken)
    cs = cs_structs()
    assert len(flat_cs("RenderParams", cs)) != len(flat_c("sr_render_params", c_structs()))
