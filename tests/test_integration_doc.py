"""The Rust bindings written out in INTEGRATION.md agree with the C header:
every `#[repr(C)]` struct has the size and field offsets the host compiler
gives the header's struct, and every `fn bk_*` in an `extern "C"` block has
the header's parameter list and return type. A maintainer who copies the
document gets a binding whose calls read and write exactly what the library
does. No GPU call."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
DOC = ROOT / "INTEGRATION.md"
HEADER = ROOT / "include" / "boitata_gpu.h"

SCALARS = {"i32": (4, 4), "i64": (8, 8), "f64": (8, 8), "u32": (4, 4), "u64": (8, 8)}


def rust_blocks():
    return re.findall(r"```rust\n(.*?)```", DOC.read_text(), flags=re.S)


def rust_structs():
    out = {}
    for block in rust_blocks():
        for name, body in re.findall(r"#\[repr\(C\)\]\s*pub struct (\w+)\s*\{(.*?)\}", block, flags=re.S):
            fields = re.findall(r"pub (\w+):\s*([^,]+?)\s*(?:,|$)", body.strip())
            out[name] = fields
    return out


def c_name(rust):
    return "bk_" + re.sub(r"(?<!^)([A-Z])", r"_\1", rust[2:]).lower()


def layout(ty, structs, memo):
    """(size, align) of a Rust type under repr(C)."""
    ty = ty.strip()
    if ty in SCALARS:
        return SCALARS[ty]
    if ty.startswith("*"):
        return 8, 8
    m = re.fullmatch(r"\[(.+);\s*(\d+)\]", ty)
    if m:
        s, a = layout(m.group(1), structs, memo)
        return s * int(m.group(2)), a
    if ty in structs:
        return struct_layout(ty, structs, memo)[:2]
    raise AssertionError(f"unknown Rust type {ty}")


def struct_layout(name, structs, memo):
    if name in memo:
        return memo[name]
    off, align, offsets = 0, 1, {}
    for fname, fty in structs[name]:
        s, a = layout(fty, structs, memo)
        off = (off + a - 1) // a * a
        offsets[fname] = off
        off += s
        align = max(align, a)
    size = (off + align - 1) // align * align
    memo[name] = (size, align, offsets)
    return memo[name]


def test_documented_structs_match_the_header(tmp_path):
    structs = rust_structs()
    assert {"BkSearch", "BkVariogram", "BkStructure", "BkBlock"} <= set(structs)
    memo = {}
    lines, want = [], []
    for name in sorted(structs):
        size, _, offsets = struct_layout(name, structs, memo)
        c = c_name(name)
        lines.append(f'printf("%zu\\n", sizeof({c}));')
        want.append(size)
        for f, o in offsets.items():
            lines.append(f'printf("%zu\\n", offsetof({c}, {f}));')
            want.append(o)
    src = tmp_path / "layout.c"
    src.write_text(f'#include "{HEADER}"\n#include <stdio.h>\n#include <stddef.h>\nint main(void){{\n'
                   + "\n".join(lines) + "\nreturn 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == want


def canon_c(param):
    """('const'|'', base, pointer depth) of a C parameter or return type."""
    p = param.strip()
    p = re.sub(r"\b\w+$", "", p) if re.search(r"[\w\*]\s+\w+$", p) or re.search(r"\*\s*\w+$", p) else p
    depth = p.count("*")
    const = "const" if re.search(r"\bconst\b", p) else ""
    base = re.sub(r"\bconst\b|\*|\s", "", p)
    base = {"int": "int32_t", "int32_t": "int32_t", "int64_t": "int64_t", "double": "double", "char": "char",
            "void": "void"}.get(base, base)
    return const, base, depth


def canon_rust(ty):
    t = ty.strip()
    depth, const = 0, ""
    while t.startswith("*"):
        m = re.match(r"\*(const|mut)\s+(.*)", t)
        if depth == 0:
            const = "const" if m.group(1) == "const" else ""
        depth += 1
        t = m.group(2).strip()
    base = {"i32": "int32_t", "i64": "int64_t", "f64": "double", "std::os::raw::c_char": "char",
            "c_char": "char", "()": "void"}.get(t, t)
    if base.startswith("Bk"):
        base = "bk_ctx" if base == "BkCtx" else c_name(base)
    return const, base, depth


def header_prototypes():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    out = {}
    for ret, name, args in re.findall(r"(\b[\w ]+?\**)\s*\b(bk_\w+)\s*\(([^)]*)\)\s*;", text):
        params = [] if args.strip() in ("", "void") else [a for a in args.split(",")]
        out[name] = (canon_c(ret + " x"), [canon_c(a) for a in params])
    return out


def rust_functions():
    out = {}
    for block in rust_blocks():
        for ext in re.findall(r'extern "C"\s*\{(.*?)\n\}', block, flags=re.S):
            for name, args, ret in re.findall(r"fn (bk_\w+)\s*\(([^)]*)\)\s*(?:->\s*([^;]+))?;", ext, flags=re.S):
                params = [a.split(":", 1)[1] for a in args.split(",") if a.strip()]
                out[name] = (canon_rust(ret if ret else "()"), [canon_rust(p) for p in params])
    return out


def test_documented_functions_match_the_header():
    c = header_prototypes()
    r = rust_functions()
    assert {"bk_create", "bk_set_samples", "bk_krige", "bk_block_krige", "bk_loo", "bk_neighbors"} <= set(r)
    for name, (ret, params) in r.items():
        assert name in c, name
        cret, cparams = c[name]
        assert ret[1:] == cret[1:], (name, ret, cret)
        assert len(params) == len(cparams), (name, len(params), len(cparams))
        for k, (a, b) in enumerate(zip(params, cparams)):
            assert a == b, (name, k, a, b)


@pytest.mark.parametrize("rust,c", [("BkSearch", "bk_search"), ("BkMikOut", "bk_mik_out"),
                                    ("BkDiagOut", "bk_diag_out")])
def test_struct_names(rust, c):
    assert c_name(rust) == c
