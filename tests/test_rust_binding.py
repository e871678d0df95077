"""The Rust FFI crate (par2-rs_amd/rust/par2hip-sys/src/lib.rs) against the C
header (include/par2hip.h), with no Rust toolchain: every entry point is
declared on both sides with the same argument names and types (pointer levels
and their constness included), every struct has the same fields in the same
order with the same types, and every status and enum constant has the same
value.  The Python binding's ctypes signatures are checked against the same
header (names and argument counts)."""
from __future__ import annotations

import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "par2hip.h")
RUST = os.path.join(ROOT, "par2-rs_amd", "rust", "par2hip-sys", "src", "lib.rs")
STRATEGY = os.path.join(ROOT, "par2-rs_amd", "rust", "par2-rs", "src", "galois", "simd", "hip.rs")

BASE = {"int32_t": "i32", "uint32_t": "u32", "uint64_t": "u64", "uint16_t": "u16", "uint8_t": "u8",
        "size_t": "usize", "float": "f32", "double": "f64", "char": "c_char", "void": "c_void"}


def c_type(decl: str) -> tuple[str, str]:
    """'const uint16_t *const *srcs' -> ('srcs', '*const *const u16')."""
    toks = re.findall(r"\*|\w+|\[\d+\]", decl)
    arr = [t for t in toks if t.startswith("[")]
    toks = [t for t in toks if not t.startswith("[")]
    name = toks[-1]
    toks = toks[:-1]
    base_const = False
    levels = []  # constness of each pointer level, innermost first
    base = None
    for t in toks:
        if t == "const":
            if levels:
                levels[-1] = True
            else:
                base_const = True
        elif t == "*":
            levels.append(False)
        else:
            base = t
    rust = BASE.get(base, base)
    consts = [base_const] + levels
    for i in range(len(levels)):
        rust = ("*const " if consts[i] else "*mut ") + rust
    if arr:
        rust = f"[{rust}; {arr[0][1:-1]}]"
    return name, rust


def strip_c(text):
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def header_functions():
    text = strip_c(open(HEADER).read())
    out = {}
    for m in re.finditer(r"([\w\s\*]+?)\b(par2hip_\w+)\s*\(([^;{}]*?)\)\s*;", text):
        ret_decl, name, params = m.group(1), m.group(2), m.group(3)
        _, ret = c_type(ret_decl.strip() + " r")
        args = [] if params.strip() in ("", "void") else [c_type(p.strip()) for p in params.split(",")]
        out[name] = (args, ret)
    return out


def header_structs():
    text = strip_c(open(HEADER).read())
    out = {}
    for m in re.finditer(r"typedef struct \{(.*?)\}\s*(\w+);", text, re.S):
        fields = []
        for decl in m.group(1).split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            first, *rest = [d.strip() for d in decl.split(",")]
            base = first.rsplit(" ", 1)[0]
            fields.append(c_type(first))
            for r in rest:
                fields.append(c_type(base + " " + r))
        out[m.group(2)] = fields
    return out


def header_constants():
    text = strip_c(open(HEADER).read())
    out = {}
    for body in re.findall(r"enum \{(.*?)\};", text, re.S):
        for name, val in re.findall(r"(PAR2HIP_\w+)\s*=\s*(-?\d+)", body):
            out[name] = int(val)
    return out


def rust_norm(t):
    return " ".join(t.replace("*const", " *const ").replace("*mut", " *mut ").split())


def rust_functions():
    text = open(RUST).read()
    block = text[text.index('extern "C" {'):]
    out = {}
    for m in re.finditer(r"pub fn (par2hip_\w+)\((.*?)\)\s*(?:->\s*([^;]+))?;", block, re.S):
        args = []
        for a in re.split(r",\s*(?![^\[]*\])", " ".join(m.group(2).split())):
            a = a.strip().rstrip(",")
            if a:
                n, t = a.split(":", 1)
                args.append((n.strip(), rust_norm(t)))
        out[m.group(1)] = (args, rust_norm(m.group(3) or "()"))
    return out


def rust_structs():
    text = open(RUST).read()
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive[^\]]*\]\s*)?pub struct (\w+) \{(.*?)\}", text, re.S):
        fields = [(n, rust_norm(t)) for n, t in re.findall(r"pub (\w+): ([^,]+),", m.group(2))]
        out[m.group(1)] = fields
    return out


def rust_constants():
    return {n: int(v) for n, v in re.findall(r"pub const (PAR2HIP_\w+): i32 = (-?\d+);", open(RUST).read())}


def test_c_type_reader():
    assert c_type("uint16_t *const *dsts") == ("dsts", "*const *mut u16")
    assert c_type("const uint16_t *const *srcs") == ("srcs", "*const *const u16")
    assert c_type("par2hip_ctx **out") == ("out", "*mut *mut par2hip_ctx")
    assert c_type("const void *d_recovery") == ("d_recovery", "*const c_void")
    assert c_type("char path[1024]") == ("path", "[c_char; 1024]")


def test_every_entry_point_matches():
    c, r = header_functions(), rust_functions()
    assert len(c) >= 48
    assert sorted(c) == sorted(r), (set(c) ^ set(r))
    for name, (args, ret) in c.items():
        rargs, rret = r[name]
        assert rargs == args, f"{name}: header {args} vs Rust {rargs}"
        assert rret == ret, f"{name}: returns {ret} vs {rret}"


def test_every_struct_matches():
    c, r = header_structs(), rust_structs()
    for name, fields in c.items():
        assert name in r, name
        assert r[name] == fields, f"{name}: header {fields} vs Rust {r[name]}"
    opaque = {"par2hip_ctx", "par2hip_batch_plan", "par2hip_server"}
    assert set(r) == set(c) | opaque


def test_every_constant_matches():
    c, r = header_constants(), rust_constants()
    assert c == r


def test_strategy_uses_only_declared_entry_points():
    used = set(re.findall(r"ffi::(par2hip_\w+)\(", open(STRATEGY).read()))
    assert used and used <= set(header_functions())
    # the region call passes one coefficient count per row (short rows)
    assert "lens.as_ptr()" in open(STRATEGY).read()


def test_python_binding_signatures_match_header():
    import ast

    src = open(os.path.join(ROOT, "par2-rs_amd", "par2hip.py")).read()
    tree = ast.parse(src)
    sig = None
    for node in ast.walk(tree):
        if isinstance(node, ast.Assign) and getattr(node.targets[0], "id", None) == "sig":
            sig = node.value
    assert isinstance(sig, ast.Dict)
    c = header_functions()
    for k, v in zip(sig.keys, sig.values):
        name = k.value
        assert name in c, name
        nargs = len(v.elts[0].elts)
        assert nargs == len(c[name][0]), f"{name}: ctypes takes {nargs} arguments, the header {len(c[name][0])}"
