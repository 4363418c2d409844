"""The C ABI library: it loads and exports every entry point include/krabgpu.h
declares (no compute call: this runs without a GPU)."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "krabgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    lib = os.path.join(ROOT, "krabmaga_amd", "libkrabgpu.so")
    assert os.path.exists(lib), "build with __graft_entry__.build()"
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (kg_[a-z0-9_]+)$", out, flags=re.M))
    declared = header_functions()
    assert declared, "no kg_ functions parsed from the header"
    missing = [f for f in declared if f not in exported]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from krabmaga_amd import _abi
    assert sorted(_abi.EXPORTS) == header_functions()


def test_library_loads_and_reports_errors_without_a_gpu_call():
    from krabmaga_amd import _abi
    L = _abi.lib()
    assert L.kg_last_error() == b""
    assert L.kg_field_create(None, None) == _abi.KG_E_INVALID
    assert b"NULL" in L.kg_last_error()


def test_rust_binding_in_integration_matches_the_header():
    """INTEGRATION.md's extern block declares every entry point, and its
    #[repr(C)] structs have the header's fields in the header's order."""
    md = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    h = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "krabgpu.h")).read(),
               flags=re.S)
    assert sorted(set(re.findall(r"pub fn (kg_[a-z0-9_]+)\(", md))) == header_functions()
    structs = re.findall(r"typedef struct (kg_\w+) \{(.*?)\} \1;", h, flags=re.S)
    assert len(structs) >= 4
    for name, body in structs:
        m = re.search(r"pub struct " + name + r" \{(.*?)\}", md, flags=re.S)
        assert m, f"{name} missing from the Rust binding"
        assert re.findall(r"pub (\w+):", m.group(1)) == re.findall(r"\b(\w+);", body), name
    for name, value in re.findall(r"#define (KG_E_\w+|KG_OK|KG_RUN_\w+) \(?(-?\d+)\)?", h):
        assert re.search(rf"pub const {name}: c_int = {value};", md), name
