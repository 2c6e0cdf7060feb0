"""The descriptor structs as the C compiler lays them out (include/mol.h, compiled with gcc)
equal their ctypes mirror field by field: size, and offset of every field.  This is the layout
a foreign binding (the Julia ccall mirror, the Python ctypes one) must reproduce."""
import ctypes as C
import os
import subprocess
import tempfile

import pytest

from molgpu import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = {"mol_optab": native.mol_optab, "mol_slot": native.mol_slot, "mol_closure": native.mol_closure,
           "mol_desc": native.mol_desc}


@pytest.fixture(scope="module")
def c_layout():
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "mol.h"', "int main(void) {"]
    for name, cls in STRUCTS.items():
        lines.append('  printf("%s size %%zu\\n", sizeof(%s));' % (name, name))
        for fname, _ in cls._fields_:
            lines.append('  printf("%s %s %%zu\\n", offsetof(%s, %s));' % (name, fname, name, fname))
    lines += ["  return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        with open(src, "w") as f:
            f.write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    table = {}
    for ln in out.splitlines():
        a, b, v = ln.split()
        table[(a, b)] = int(v)
    return table


@pytest.mark.parametrize("name", sorted(STRUCTS))
def test_ctypes_mirror_matches_the_c_layout(name, c_layout):
    cls = STRUCTS[name]
    assert C.sizeof(cls) == c_layout[(name, "size")], name
    for fname, _ in cls._fields_:
        assert getattr(cls, fname).offset == c_layout[(name, fname)], (name, fname)


def test_every_header_field_is_mirrored():
    """No field of the C structs is missing from the mirror (field names parsed from mol.h)."""
    import re
    text = open(os.path.join(ROOT, "include", "mol.h")).read()
    for name, cls in STRUCTS.items():
        end = text.index("} %s;" % name)
        start = text.rindex("typedef struct {", 0, end) + len("typedef struct {")
        body = re.sub(r"/\*.*?\*/", "", text[start:end], flags=re.S)
        fields = set()
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            for part in decl.split(","):
                tok = re.findall(r"[A-Za-z_]\w*", part.split("[")[0])
                if tok:
                    fields.add(tok[-1])
        assert fields == {f for f, _ in cls._fields_}, (name, fields ^ {f for f, _ in cls._fields_})


# ---- the Julia mirror documented in INTEGRATION.md ---------------------------------------

_JL_SCALARS = {"Int32": C.c_int32, "Int64": C.c_int64, "UInt32": C.c_uint32, "Float64": C.c_double}


def _jl_structs():
    """The struct definitions of the Julia binding in INTEGRATION.md, as ctypes structures built
    with Julia's isbits layout rules (each field at its natural alignment, as in C)."""
    import re
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    consts = {m.group(1): int(m.group(2)) for m in re.finditer(r"const (\w+) = (\d+)", text)}
    out = {}

    def ctype(t):
        t = t.strip()
        m = re.fullmatch(r"NTuple\{(\w+),(.+)\}", t)
        if m:
            k = m.group(1)
            return ctype(m.group(2)) * (consts[k] if k in consts else int(k))
        if t.startswith("Ptr{"):
            return C.c_void_p
        return _JL_SCALARS[t]

    for m in re.finditer(r"struct (Mol\w+)\n(.*?)\nend", text, re.S):
        fields = []
        for decl in re.split(r"[;\n]", m.group(2)):
            decl = decl.strip()
            if decl:
                name, t = decl.split("::", 1)
                fields.append((name.strip(), ctype(t)))
        out[m.group(1)] = type(m.group(1), (C.Structure,), {"_fields_": fields})
    return out


@pytest.mark.parametrize("jl,cname", [("MolOptab", "mol_optab"), ("MolSlot", "mol_slot"),
                                      ("MolClosure", "mol_closure"), ("MolDesc", "mol_desc")])
def test_julia_mirror_matches_the_c_layout(jl, cname, c_layout):
    structs = _jl_structs()
    assert jl in structs, jl
    cls = structs[jl]
    # same field names in the same order as the C struct, same offsets and size
    assert [f for f, _ in cls._fields_] == [f for f, _ in STRUCTS[cname]._fields_], jl
    assert C.sizeof(cls) == c_layout[(cname, "size")], jl
    for fname, _ in cls._fields_:
        assert getattr(cls, fname).offset == c_layout[(cname, fname)], (jl, fname)
