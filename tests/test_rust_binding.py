"""The Rust FFI crate (box2d-rs_amd/rust/b2g-sys) declares the C ABI of include/b2g.h one to one.

There is no Rust toolchain here, so the crate is checked as text: every function the header
declares has an `extern "C"` declaration with the same parameters in order (C types mapped to
their Rust equivalents: int32_t -> i32, const T* -> *const T, T* -> *mut T, ...) and the same
return type, and nothing else is declared; every struct of the header has a #[repr(C)] struct
with the same field types in order, whose size (C layout rules) equals sizeof of the C struct as
gcc compiles the header; every #define'd constant has a Rust const of the same value.
"""
import os
import re
import subprocess
import tempfile

from helpers import ROOT

HEADER = os.path.join(ROOT, "include", "b2g.h")
LIB_RS = os.path.join(ROOT, "box2d-rs_amd", "rust", "b2g-sys", "src", "lib.rs")

C_BASE = {"int32_t": "i32", "int64_t": "i64", "uint32_t": "u32", "uint16_t": "u16", "int16_t": "i16",
          "float": "f32", "double": "f64", "int": "c_int", "char": "c_char"}
SIZES = {"i32": 4, "u32": 4, "f32": 4, "i64": 8, "f64": 8, "u16": 2, "i16": 2, "u8": 1, "c_int": 4}


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def c_type_to_rust(t):
    t = " ".join(t.replace("*", " * ").split())
    stars = t.count("*")
    const = t.startswith("const ")
    base = t.replace("const ", "").replace("*", "").strip()
    r = C_BASE.get(base, base)
    if stars == 0:
        return r
    if stars == 1:
        return ("*const " if const else "*mut ") + r
    return "*mut " * (stars - 1) + ("*const " if const else "*mut ") + r


def header_functions():
    text = _strip_comments(open(HEADER).read())
    out = {}
    for ret, name, params in re.findall(r"^(int|const char\s*\*)\s*(b2g_\w+)\s*\(([^)]*)\)\s*;", text, re.M | re.S):
        ps = []
        for p in params.split(","):
            p = " ".join(p.split())
            if not p or p == "void":
                continue
            m = re.match(r"(.*?)(\w+)$", p)
            ps.append(c_type_to_rust(m.group(1).strip()))
        out[name] = ("c_int" if ret == "int" else "*const c_char", ps)
    return out


def rust_functions():
    text = open(LIB_RS).read()
    block = text[text.index('extern "C" {'):]
    out = {}
    for name, params, ret in re.findall(r"pub fn (b2g_\w+)\s*\(([^)]*)\)\s*->\s*([^;]+);", block, re.S):
        ps = []
        for p in params.split(","):
            p = " ".join(p.split())
            if p:
                ps.append(p.split(":", 1)[1].strip())
        out[name] = (" ".join(ret.split()), ps)
    return out


def test_functions_match_header():
    h, r = header_functions(), rust_functions()
    assert len(h) > 40
    assert sorted(h) == sorted(r), (set(h) ^ set(r))
    for name in h:
        assert h[name] == r[name], "%s: header %s, crate %s" % (name, h[name], r[name])


def header_structs():
    text = _strip_comments(open(HEADER).read())
    out = {}
    for body, name in re.findall(r"typedef struct \w+ \{(.*?)\}\s*(\w+);", text, re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"([\w ]+?) (.+)$", decl)
            ctype = C_BASE.get(m.group(1).strip(), m.group(1).strip())
            for var in m.group(2).split(","):
                var = var.strip()
                dims = [int(x) for x in re.findall(r"\[(\d+)\]", var)]
                n = 1
                for d in dims:
                    n *= d
                fields.append((ctype, n))
        out[name] = fields
    return out


def rust_structs():
    text = open(LIB_RS).read()
    out = {}
    for name, body in re.findall(r"pub struct (b2g_\w+)\s*\{(.*?)\}", text, re.S):
        if "_private" in body:
            continue
        fields = []
        for ty in re.findall(r"pub \w+:\s*([^,]+?)(?:,|$)", body.strip(), re.M):
            ty = ty.strip()
            m = re.match(r"\[(\w+);\s*(\d+)\]$", ty)
            fields.append((m.group(1), int(m.group(2))) if m else (ty, 1))
        out[name] = fields
    return out


def _flatten(fields, structs):
    """Struct-typed fields expanded into their scalar fields (for the order check)."""
    flat = []
    for ty, n in fields:
        if ty in structs:
            for _ in range(n):
                flat += _flatten(structs[ty], structs)
        else:
            flat.append((ty, n))
    return flat


def _layout_size(fields, structs):
    off, align = 0, 1
    for ty, n in fields:
        if ty in structs:
            sz, al = _layout_size(structs[ty], structs)
        else:
            sz = al = SIZES[ty]
        off = (off + al - 1) // al * al + sz * n
        align = max(align, al)
    return (off + align - 1) // align * align, align


def test_structs_match_header():
    h, r = header_structs(), rust_structs()
    assert sorted(h) == sorted(r), set(h) ^ set(r)
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "sizes.c")
        with open(src, "w") as f:
            f.write('#include <stdio.h>\n#include "%s"\nint main(void) {\n' % HEADER)
            for name in sorted(h):
                f.write('  printf("%s %%zu\\n", sizeof(%s));\n' % (name, name))
            f.write("  return 0;\n}\n")
        exe = os.path.join(d, "sizes")
        subprocess.check_call(["gcc", "-std=c99", "-o", exe, src])
        sizes = dict((a, int(b)) for a, b in (l.split() for l in subprocess.check_output([exe]).decode().splitlines()))
    for name in h:
        assert _flatten(h[name], h) == _flatten(r[name], r), name
        assert _layout_size(r[name], r)[0] == sizes[name], name


def test_constants_match_header():
    text = _strip_comments(open(HEADER).read())
    want = dict(re.findall(r"^#define (B2G_\w+) (-?\d+)", text, re.M))
    want.update(re.findall(r"^\s+(B2G_KERNEL_\w+) = (-?\d+)", text, re.M))
    got = dict(re.findall(r"pub const (B2G_\w+): \w+ = (-?\d+);", open(LIB_RS).read()))
    assert want == got
