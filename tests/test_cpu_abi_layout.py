"""The C ABI's structs have one layout everywhere they are declared.

A small C++ program compiled against include/lgbm_hip.h prints sizeof and
every field's offsetof; the ctypes structures of the Python binding and the
#[repr(C)] structs of the Rust stub in INTEGRATION.md must agree with it
field for field (names, order, offsets, total size). Any field added to the
header and not to a binding fails here.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

from tests._pkg import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lgbm_hip.h")
INTEGRATION = os.path.join(ROOT, "INTEGRATION.md")

CTYPES = {
    "lgbmhip_bin_params": capi.BinParams,
    "lgbmhip_feature_meta": capi.FeatureMeta,
    "lgbmhip_split_params": capi.SplitParams,
    "lgbmhip_train_opts": capi.TrainOpts,
    "lgbmhip_metric_opts": capi.MetricOpts,
}


def header_structs():
    """typedef struct { ... } name; -> {name: [field, ...]} in declaration order."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for body, name in re.findall(r"typedef struct\s*\{(.*?)\}\s*(lgbmhip_\w+)\s*;", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            m = re.match(r"(?:const\s+)?\w+\s*\**\s*(.*)$", decl, flags=re.S)
            for f in m.group(1).split(","):
                fields.append(f.strip().lstrip("*").strip())
        out[name] = fields
    return out


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    structs = header_structs()
    lines = ['#include <cstdio>', '#include <cstddef>', f'#include "{HEADER}"', "int main() {"]
    for name, fields in structs.items():
        lines.append(f'  std::printf("{name} . %zu\\n", sizeof({name}));')
        for f in fields:
            lines.append(f'  std::printf("{name} {f} %zu\\n", offsetof({name}, {f}));')
    lines += ["  return 0;", "}"]
    d = tmp_path_factory.mktemp("abi")
    src, exe = d / "layout.cpp", d / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["g++", "-std=c++17", "-o", str(exe), str(src)], check=True)
    res = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    lay = {}
    for ln in res.splitlines():
        s, f, v = ln.split()
        lay.setdefault(s, {})[f] = int(v)
    return structs, lay


def test_header_declares_the_bound_structs(c_layout):
    structs, _ = c_layout
    assert set(CTYPES) <= set(structs), set(CTYPES) - set(structs)


@pytest.mark.parametrize("name", sorted(CTYPES))
def test_ctypes_layout_equals_the_header(c_layout, name):
    structs, lay = c_layout
    cls = CTYPES[name]
    assert [f[0] for f in cls._fields_] == structs[name]
    assert C.sizeof(cls) == lay[name]["."]
    for f in structs[name]:
        assert getattr(cls, f).offset == lay[name][f], (name, f)


RUST_SIZE = {"i32": 4, "u32": 4, "i64": 8, "f32": 4, "c_double": 8, "f64": 8, "u8": 1, "i8": 1}


def rust_structs():
    """#[repr(C)] pub struct name { pub f: T, ... } of the INTEGRATION.md stub
    -> {name: [(field, size)]} (opaque handles skipped)."""
    text = open(INTEGRATION).read()
    out = {}
    for name, body in re.findall(r"#\[repr\(C\)\]\s*pub struct (lgbmhip_\w+)\s*\{(.*?)\}", text, flags=re.S):
        fields = re.findall(r"pub\s+(\w+)\s*:\s*([^,]+?)\s*(?:,|$)", body.strip())
        if not fields:
            continue
        out[name] = [(f, 8 if t.startswith("*") else RUST_SIZE[t.strip()]) for f, t in fields]
    return out


def test_rust_stub_declares_every_header_struct():
    assert set(rust_structs()) == set(CTYPES), set(rust_structs()) ^ set(CTYPES)


@pytest.mark.parametrize("name", sorted(CTYPES))
def test_rust_stub_layout_equals_the_header(c_layout, name):
    structs, lay = c_layout
    fields = rust_structs()[name]
    assert [f for f, _ in fields] == structs[name]
    # repr(C): each field at the next multiple of its own size (all are scalars or pointers)
    off, align = 0, 1
    for f, sz in fields:
        off = (off + sz - 1) // sz * sz
        assert off == lay[name][f], (name, f, off, lay[name][f])
        off += sz
        align = max(align, sz)
    assert (off + align - 1) // align * align == lay[name]["."]


def header_prototypes():
    """lgbmhip_* function prototypes of the header: {name: parameter count}."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for name, params in re.findall(r"\b\w[\w\s\*]*?\b(lgbmhip_\w+)\s*\(([^)]*)\)\s*;", text):
        params = params.strip()
        out[name] = 0 if params in ("", "void") else len(params.split(","))
    return out


def test_ctypes_signatures_match_the_header():
    """Every function the binding declares takes as many arguments as its
    header prototype, and every header function is bound."""
    protos = header_prototypes()
    table = capi.SIGNATURES
    assert set(table) == set(protos), (set(table) ^ set(protos))
    for name, (_, args) in table.items():
        assert len(args) == protos[name], f"{name}: ctypes {len(args)} arguments, header {protos[name]}"
