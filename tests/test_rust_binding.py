"""The source-only Rust crate (micrographia_amd/rust/md-hip) against the C header it binds, on the
CPU (no Rust toolchain in this image): every #[repr(C)] struct has the header struct's fields in
the same order with the same types, and every function of its extern block is declared in the
header with the same parameter count."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = open(os.path.join(ROOT, "include", "rahman_hip.h")).read()
CRATE = os.path.join(ROOT, "micrographia_amd", "rust", "md-hip")
LIB = open(os.path.join(CRATE, "src", "lib.rs")).read()

RUST_TO_C = {"i64": "int64_t", "u64": "uint64_t", "u32": "uint32_t", "i32": "int32_t", "f64": "double", "f32": "float",
             "RhPair": "rh_pair", "*const u32": "const uint32_t*"}
PAIRS = {"RhSystem3": "rh_system3", "RhSystem2": "rh_system2", "RhMode": "rh_mode", "RhPair": "rh_pair",
         "RhLedger": "rh_ledger3", "RhListStats": "rh_list_stats"}


def c_fields(name):
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), HEADER, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"((?:const )?[a-z_0-9]+\*?)\s+(.*)", decl)
        ty, names = m.group(1), m.group(2)
        for part in names.split(","):
            part = part.strip()
            dims = re.findall(r"\[(\d+)\]", part)
            out.append((part.split("[")[0].strip(), ty, tuple(int(d) for d in dims)))
    return out


def rust_fields(name):
    body = re.search(r"pub struct %s \{(.*?)\n\}" % name, LIB, re.S).group(1)
    out = []
    for line in body.splitlines():
        m = re.match(r"\s*pub (\w+): (.+),$", line.strip() + ("" if line.strip().endswith(",") else ","))
        if not m:
            continue
        fname, ty = m.group(1), m.group(2).rstrip(",")
        dims = []
        while ty.startswith("["):
            inner, n = re.match(r"\[(.*); (\d+)\]$", ty).groups()
            dims.append(int(n))
            ty = inner
        out.append((fname, RUST_TO_C[ty], tuple(dims)))
    return out


def test_repr_c_structs_match_the_header():
    for rust, c in PAIRS.items():
        assert rust_fields(rust) == c_fields(c), (rust, c)


def test_every_bound_function_is_declared_with_its_arity():
    externs = re.search(r'unsafe extern "C" \{(.*?)\n\}', LIB, re.S).group(1)
    bound = re.findall(r"fn (rh_\w+)\((.*?)\)", externs)
    assert len(bound) > 40
    for name, args in bound:
        m = re.search(r"\b%s\(([^)]*)\)" % name, HEADER)
        assert m, f"{name} is not in the header"
        c_args = [a for a in m.group(1).split(",") if a.strip() and a.strip() != "void"]
        r_args = [a for a in args.split(",") if a.strip()]
        assert len(c_args) == len(r_args), name


def test_crate_files_exist():
    for f in ("Cargo.toml", "build.rs", os.path.join("src", "lib.rs")):
        assert os.path.exists(os.path.join(CRATE, f))
    toml = open(os.path.join(CRATE, "Cargo.toml")).read()
    assert 'name = "md-hip"' in toml and "md-core" in toml
