"""The Rust FFI crate (crates/pystamps-hip/src/lib.rs) against the C header
(include/pystamps_hip.h).  CPU.

Rust is not installed here, so the crate is checked as source: every header
struct has a #[repr(C)] mirror with the same field names in the same order and
the corresponding Rust types, whose repr(C) offsets equal the offsetof() values
gcc reports for the header; every declared function has an extern "C"
declaration with the same parameter types and return type; the ABI version and
the constants agree.
"""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "pystamps_hip.h"
LIB_RS = ROOT / "crates" / "pystamps-hip" / "src" / "lib.rs"

STRUCTS = {"ps2_patch_in": "Ps2PatchIn", "ps2_patch_out": "Ps2PatchOut", "ps2_run_opts": "Ps2RunOpts",
           "ps2_tiles": "Ps2Tiles", "ps2_device_outputs": "Ps2DeviceOutputs"}
SCALARS = {"int64_t": "i64", "int32_t": "i32", "uint32_t": "u32", "double": "f64", "float": "f32",
           "size_t": "usize", "int": "c_int", "char": "c_char", "void": "c_void",
           "ps2_session": "Ps2Session", "ps2_iteration_fn": "Ps2IterationFn", **STRUCTS}
RUST_SIZE = {"i64": 8, "i32": 4, "u32": 4, "f64": 8, "f32": 4, "usize": 8, "c_int": 4}


def _strip_c_comments(text):
    return re.sub(r"/\*.*?\*/", " ", text, flags=re.S)


def _c_decl_to_rust(base, stars, const_base, const_ptr=False):
    """C declarator (base type, pointer depth, const on the pointee) -> Rust type."""
    t = SCALARS[base]
    for level in range(stars):
        inner_const = const_base if level == 0 else const_ptr
        t = f"*{'const' if inner_const else 'mut'} {t}"
    return t


def header_structs():
    text = _strip_c_comments(HEADER.read_text())
    out = {}
    for body, name in re.findall(r"typedef struct \{(.*?)\}\s*(\w+);", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"(const\s+)?(\w+)\s+(.*)$", decl)
            const, base, rest = bool(m.group(1)), m.group(2), m.group(3)
            for d in rest.split(","):
                d = d.strip()
                stars = d.count("*")
                ident = d.replace("*", "").strip()
                arr = re.match(r"(\w+)\[(\w+)\]", ident)
                if arr:
                    n = {"PS2_COH_BINS": 100}.get(arr.group(2), arr.group(2))
                    fields.append((arr.group(1), f"[{SCALARS[base]}; {n}]"))
                else:
                    fields.append((ident, _c_decl_to_rust(base, stars, const)))
        out[name] = fields
    return out


def rust_structs():
    text = LIB_RS.read_text()
    out = {}
    for name, body in re.findall(r"pub struct (\w+) \{(.*?)\n\}", text, flags=re.S):
        fields = []
        for line in body.splitlines():
            m = re.match(r"\s*pub (\w+): (.+?),\s*$", line)
            if m:
                t = m.group(2).replace("PS2_COH_BINS", "100")
                fields.append((m.group(1), t))
        out[name] = fields
    return out


def header_functions():
    text = _strip_c_comments(HEADER.read_text())
    out = {}
    for ret, name, args in re.findall(r"\n\s*((?:const\s+)?\w+\s*\**)\s*(ps2_hip_\w+)\s*\(([^)]*)\)\s*;", text):
        rconst = ret.strip().startswith("const")
        rbase = ret.replace("const", "").replace("*", "").strip()
        rtype = _c_decl_to_rust(rbase, ret.count("*"), rconst) if rbase != "void" or "*" in ret else "()"
        params = []
        for a in " ".join(args.split()).split(","):
            a = a.strip()
            if a in ("void", ""):
                continue
            m = re.match(r"(const\s+)?(\w+)\s*((?:\*\s*(?:const\s*)?)*)\s*(\w+)$", a)
            assert m, a
            const, base, ptrs = bool(m.group(1)), m.group(2), m.group(3)
            stars = ptrs.count("*")
            const_ptr = "const" in ptrs   # `T *const *`: the outer pointee is const
            if stars == 2 and const_ptr:
                params.append(f"*const *mut {SCALARS[base]}")
            else:
                params.append(_c_decl_to_rust(base, stars, const))
        out[name] = (rtype, params)
    return out


def rust_functions():
    text = LIB_RS.read_text()
    block = text[text.index('extern "C" {'):]
    block = block[:block.index("\n}\n")]
    out = {}
    for name, args, ret in re.findall(r"pub fn (ps2_hip_\w+)\((.*?)\)\s*(?:->\s*([^;]+))?;", block, flags=re.S):
        params = [a.split(":", 1)[1].strip() for a in " ".join(args.split()).split(",") if a.strip()]
        out[name] = ((ret or "()").strip(), params)
    return out


def test_structs_mirror_the_header():
    h, r = header_structs(), rust_structs()
    assert set(STRUCTS) <= set(h)
    for cname, rname in STRUCTS.items():
        assert rname in r, rname
        assert r[rname] == h[cname], (cname, [x for x in zip(h[cname], r[rname]) if x[0] != x[1]])


def _rust_offsets(fields):
    off, out = 0, []
    for name, t in fields:
        if t.startswith("["):
            elem, n = t[1:-1].split(";")
            size, align = RUST_SIZE[elem.strip()] * int(n), RUST_SIZE[elem.strip()]
        elif t.startswith("*"):
            size = align = 8
        else:
            size = align = RUST_SIZE[t]
        off = (off + align - 1) // align * align
        out.append((name, off))
        off += size
    return out


def test_repr_c_offsets_equal_gcc_offsetof(tmp_path):
    h = header_structs()
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pystamps_hip.h"', "int main(void) {"]
    for cname in STRUCTS:
        for field, _ in h[cname]:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
    lines.append("  return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        s, f, v = line.split()
        got[(s, f)] = int(v)
    r = rust_structs()
    for cname, rname in STRUCTS.items():
        for field, off in _rust_offsets(r[rname]):
            assert got[(cname, field)] == off, (cname, field, got[(cname, field)], off)


def test_extern_functions_mirror_the_header():
    h, r = header_functions(), rust_functions()
    assert len(h) == 27
    assert set(h) == set(r), (set(h) ^ set(r))
    for name, (ret, params) in h.items():
        assert r[name][1] == params, (name, params, r[name][1])
        assert r[name][0] == ret, (name, ret, r[name][0])


def test_callback_type_mirrors_the_header():
    # typedef int (*ps2_iteration_fn)(void *user, int64_t iteration, ps2_patch_out *state);
    text = _strip_c_comments(HEADER.read_text())
    m = re.search(r"typedef\s+(\w+)\s*\(\s*\*\s*ps2_iteration_fn\s*\)\s*\(([^)]*)\)\s*;", text)
    assert m, "ps2_iteration_fn typedef"
    c_ret = _c_decl_to_rust(m.group(1), 0, False)
    c_params = []
    for a in m.group(2).split(","):
        pm = re.match(r"\s*(const\s+)?(\w+)\s*(\**)\s*\w+\s*$", a)
        c_params.append(_c_decl_to_rust(pm.group(2), pm.group(3).count("*"), bool(pm.group(1))))
    rs = LIB_RS.read_text()
    rm = re.search(r"pub type Ps2IterationFn = Option<unsafe extern \"C\" fn\((.*?)\) -> (\w+)>;", rs)
    assert rm, "Ps2IterationFn alias"
    r_params = [a.split(":", 1)[1].strip() for a in rm.group(1).split(",")]
    assert (rm.group(2), r_params) == (c_ret, c_params)


def test_constants_and_abi_version():
    text = _strip_c_comments(HEADER.read_text())
    rs = LIB_RS.read_text()
    for name, value in re.findall(r"#define (PS2_\w+) \(?(-?\d+)\)?", text):
        m = re.search(rf"pub const {name}: \w+ = (-?\d+);", rs)
        assert m, name
        assert int(m.group(1)) == int(value), name
    from pystamps_amd import _native
    assert re.search(rf"PS2_HIP_ABI_VERSION: c_int = {_native.ABI_VERSION};", rs)


@pytest.mark.parametrize("path", ["Cargo.toml", "build.rs", "src/lib.rs"])
def test_crate_files_present(path):
    assert (ROOT / "crates" / "pystamps-hip" / path).is_file()
