"""The Rust binding (ticit_amd/bindings/rust/ticit-hip-sys) against the C
header it binds (include/ticit_hip.h). No Rust toolchain is needed: the test
reads both sources and checks that every function the header declares is
declared in the crate with the same parameter names and count, that every
struct has the same fields in the same order, and that the status codes and
table sizes agree. A field added to the header without the binding, or a
reordered field (which would silently shift a #[repr(C)] layout), fails here.
"""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ticit_hip.h")
CRATE = os.path.join(ROOT, "ticit_amd", "bindings", "rust", "ticit-hip-sys")


def strip_c_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def c_functions(text: str) -> dict:
    out = {}
    for m in re.finditer(r"\b(?:[\w\s\*]+?)\b(ticit_hip_\w+)\s*\(([^;{]*?)\)\s*;", text):
        name, params = m.group(1), m.group(2).strip()
        if params in ("", "void"):
            out[name] = []
            continue
        names = []
        for p in params.split(","):
            p = re.sub(r"\[.*?\]", "", p).strip()
            names.append(re.findall(r"\w+", p)[-1])
        out[name] = names
    return out


def c_structs(text: str) -> dict:
    out = {}
    for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} \1;", text, flags=re.S):
        fields = []
        for decl in m.group(2).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            decl = re.sub(r"\[[^\]]*\]", "", decl)
            head, *rest = [d.strip() for d in decl.split(",")]
            fields.append(re.findall(r"\w+", head)[-1])
            fields += [re.findall(r"\w+", r)[-1] for r in rest]
        out[m.group(1)] = fields
    return out


def rust_functions(text: str) -> dict:
    body = text[text.index('extern "C" {'):]
    body = body[:body.index("\n}\n")]
    out = {}
    for m in re.finditer(r"pub fn (ticit_hip_\w+)\((.*?)\)", re.sub(r"//[^\n]*", "", body), flags=re.S):
        params = [p.strip() for p in m.group(2).split(",") if p.strip()]
        out[m.group(1)] = [p.split(":")[0].strip() for p in params]
    return out


def rust_structs(text: str) -> dict:
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\][^{]*?pub struct (\w+) \{(.*?)\n\}", text, flags=re.S):
        fields = re.findall(r"pub (\w+):", m.group(2))
        out[m.group(1)] = fields
    return out


def sources():
    with open(HEADER) as f:
        header = strip_c_comments(f.read())
    with open(os.path.join(CRATE, "src", "lib.rs")) as f:
        crate = f.read()
    return header, crate


def test_every_header_function_is_bound_with_the_same_parameters():
    header, crate = sources()
    cf, rf = c_functions(header), rust_functions(crate)
    assert len(cf) >= 30, sorted(cf)
    missing = sorted(set(cf) - set(rf))
    assert not missing, missing
    for name, params in cf.items():
        assert rf[name] == params, (name, params, rf[name])
    assert set(rf) == set(cf), sorted(set(rf) - set(cf))


def test_every_header_struct_has_the_same_fields_in_order():
    header, crate = sources()
    cs, rs = c_structs(header), rust_structs(crate)
    assert {"ticit_hip_options", "ticit_hip_result", "ticit_hip_program_info", "ticit_hip_batch_item",
            "ticit_hip_batch_options", "ticit_hip_batch_info", "ticit_hip_parity",
            "ticit_hip_reference_bits"} <= set(cs)
    for name, fields in cs.items():
        assert name in rs, name
        assert rs[name] == fields, (name, fields, rs[name])


def test_status_codes_and_table_sizes_agree():
    header, crate = sources()
    for m in re.finditer(r"(TICIT_HIP_[A-Z_]+) = (\d+)", header):
        if m.group(1).startswith("TICIT_HIP_FEATURE"):
            continue
        assert re.search(rf"pub const {m.group(1)}: ticit_hip_status = {m.group(2)};", crate), m.group(1)
    nfeat = len(re.findall(r"TICIT_HIP_FEATURE_\w+", header))
    assert f"pub const TICIT_HIP_NFEATURES: usize = {nfeat};" in crate
    for name in ("TICIT_HIP_PLANNER_FEATURES", "TICIT_HIP_MLP_PARAMS"):
        value = re.search(rf"#define {name} (\d+)", open(HEADER).read()).group(1)
        assert f"pub const {name}: usize = {value};" in crate


def test_the_crate_links_the_in_tree_library_name():
    with open(os.path.join(CRATE, "build.rs")) as f:
        build = f.read()
    with open(os.path.join(CRATE, "Cargo.toml")) as f:
        cargo = f.read()
    assert "rustc-link-lib=dylib=ticit_hip" in build and 'links = "ticit_hip"' in cargo
