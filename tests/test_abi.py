"""C-ABI checks that need no GPU: the library loads (its HIP runtime
dependency resolves without a device) and exports every entry point that
include/bqcd_hip.h declares; the Fortran module's interfaces, the ctypes
argument lists and the HMC structs of both bindings match the header;
device tensors are checked before they reach the library; and nothing on
the host side falls back to the oracle when the library is absent."""
import ctypes
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
HEADER = os.path.join(REPO, "include", "bqcd_hip.h")
LIB = os.path.join(REPO, "bqcd-r399-quda_amd", "libbqcd_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(hb_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_parses():
    names = declared()
    for must in ("hb_init", "hb_cg", "hb_cg_mixed", "hb_cg_multi", "hb_dslash", "hb_apply",
                 "hb_dsf", "hb_dsd", "hb_dsg", "hb_hmc_u", "hb_clover_mult", "hb_build_clover"):
        assert must in names
    assert len(names) >= 39


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbqcd_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbqcd_hip.so not built")
def test_error_path_without_device():
    """hb_init on a host with no usable device (or a bad argument) returns an
    error code and a message instead of crashing."""
    lib = ctypes.CDLL(LIB)
    lib.hb_last_error.restype = ctypes.c_char_p
    L = (ctypes.c_int * 4)(4, 4, 4, 3)          # odd extent: rejected before any HIP call
    bc = (ctypes.c_int * 4)(1, 1, 1, -1)
    ctx = ctypes.c_void_p()
    rc = lib.hb_init(L, bc, 0, ctypes.byref(ctx))
    assert rc == -1
    assert lib.hb_last_error()


def test_harness_fails_loudly_without_library(monkeypatch):
    import bqcd_hip

    monkeypatch.setattr(bqcd_hip, "LIB_PATH", os.path.join(HERE, "no_such_lib.so"))
    monkeypatch.setattr(bqcd_hip, "_lib", None, raising=False)
    with pytest.raises(bqcd_hip.HipError):
        bqcd_hip.lib()


def test_device_tensors_are_checked_before_they_reach_the_library():
    """the harness hands the library only contiguous float64 GPU tensors with
    enough values; anything else is an error on the host, never a kernel
    reading past a buffer"""
    import numpy as np
    import torch
    import bqcd_hip

    for bad in (torch.zeros(48, dtype=torch.float64), np.zeros(48)):  # host memory
        with pytest.raises(bqcd_hip.HipError, match="float64 GPU tensor"):
            bqcd_hip._tp(bad)

    class FakeDevice:  # stands in for a GPU tensor: only its description is read
        is_cuda = True
        shape = (12, 2, 2)

        def __init__(self, dtype, contiguous=True):
            self.dtype, self.contiguous = dtype, contiguous

        def is_contiguous(self):
            return self.contiguous

        def numel(self):
            return 48

        def data_ptr(self):
            return 4096

    with pytest.raises(bqcd_hip.HipError, match="float64 GPU tensor"):
        bqcd_hip._tp(FakeDevice(torch.float32))
    with pytest.raises(bqcd_hip.HipError, match="float64 GPU tensor"):
        bqcd_hip._tp(FakeDevice(torch.float64, contiguous=False))
    with pytest.raises(bqcd_hip.HipError, match="48 values where 96"):
        bqcd_hip._tp(FakeDevice(torch.float64), 96)
    assert bqcd_hip._tp(FakeDevice(torch.float64), 48).value == 4096


FORTRAN = os.path.join(REPO, "bqcd-r399-quda_amd", "fortran", "bqcd_hip.F90")


def header_text():
    return re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)


def fortran_text():
    return re.sub(r"&[ \t]*\n\s*", " ", open(FORTRAN).read())   # continuation lines joined


def test_fortran_interfaces_match_the_header():
    """every bind(C) function of the Fortran module names a declared entry point
    and passes as many arguments as the header declares"""
    protos = {}
    for m in re.finditer(r"^\s*(?:const\s+)?[\w ]+?\**\s*(hb_\w+)\s*\(([^;]*?)\)\s*;", header_text(), flags=re.M):
        args = m.group(2).strip()
        protos[m.group(1)] = 0 if args in ("", "void") else len(args.split(","))
    binds = {}
    pat = r"(?:function|subroutine)[ \t]+\w+[ \t]*\(([^)]*)\)[^\n]*?bind\(C,\s*name=\"(hb_\w+)\"\)"
    for m in re.finditer(pat, fortran_text()):
        binds[m.group(2)] = len(m.group(1).split(",")) if m.group(1).strip() else 0
    assert len(binds) >= 30
    for name, n in binds.items():
        assert name in protos, name
        assert n == protos[name], (name, n, protos[name])


def c_struct(name):
    body = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + r"\s*;", header_text()).group(1)
    kinds = {"double": "d", "int": "i", "char": "c", "float": "f"}
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if decl:
            t, rest = decl.split(None, 1)
            out += [(n.strip(), kinds[t]) for n in rest.split(",")]
    return out


def fortran_type(name):
    body = re.search(r"type, bind\(C\) :: " + name + r"\n(.*?)end type", fortran_text(), flags=re.S).group(1)
    kinds = {"real(c_double)": "d", "integer(c_int)": "i", "character(kind=c_char)": "c", "real(c_float)": "f"}
    out = []
    for line in body.strip().splitlines():
        t, names = line.split("::")
        out += [(n.strip(), kinds[t.strip()]) for n in names.split(",")]
    return out


@pytest.mark.parametrize("name", ["hb_hmc_params", "hb_hmc_out"])
def test_struct_mirrors_match_the_header(name):
    """the HMC parameter and result structs have the header's members, in order
    and with the same C types, in the Fortran module and in the ctypes harness"""
    import ctypes as ct
    import bqcd_hip

    c = c_struct(name)
    assert fortran_type(name) == c
    py = {"hb_hmc_params": bqcd_hip.HmcParams, "hb_hmc_out": bqcd_hip.HmcOut}[name]
    kinds = {ct.c_double: "d", ct.c_int: "i", ct.c_char: "c", ct.c_float: "f"}
    rename = {"lam": "lambda"}   # a Python keyword
    assert [(rename.get(n, n), kinds[t]) for n, t in py._fields_] == c


def split_top_level(s):
    """s split at the commas outside any bracket"""
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([":
            depth += 1
        elif ch in ")]":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    return [p for p in parts + [cur] if p.strip()]


def test_ctypes_argtypes_match_the_header():
    """every argtypes list of the ctypes harness has the header's argument count"""
    protos = {}
    for m in re.finditer(r"^\s*(?:const\s+)?[\w ]+?\**\s*(hb_\w+)\s*\(([^;]*?)\)\s*;", header_text(), flags=re.M):
        args = m.group(2).strip()
        protos[m.group(1)] = 0 if args in ("", "void") else len(args.split(","))
    src = open(os.path.join(REPO, "bqcd-r399-quda_amd", "bqcd_hip.py")).read()
    seen = 0
    for m in re.finditer(r"L\.(hb_\w+)\.argtypes\s*=\s*\[", src):
        i = j = m.end()
        depth = 1
        while depth:
            if src[j] in "([":
                depth += 1
            elif src[j] in ")]":
                depth -= 1
            j += 1
        n = len(split_top_level(src[i:j - 1]))
        assert protos.get(m.group(1)) == n, (m.group(1), n, protos.get(m.group(1)))
        seen += 1
    assert seen >= len(protos) - 1   # all but the argument-less hb_last_error


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbqcd_hip.so not built")
def test_multi_block_limit_keeps_buffers_below_4gib():
    """hb_multi_block_limit (host only, no device): the largest RHS block whose
    block spinor fields (12 complex numbers per site and RHS, 8 bytes each in
    FP32, 4 in 16 bits) stay below the 4 GiB span of one buffer"""
    import bqcd_hip

    vh = {"24x48": 24 ** 3 * 48 // 2, "32x64": 32 ** 3 * 64 // 2, "48x96": 48 ** 3 * 96 // 2,
          "64x128": 64 ** 3 * 128 // 2}
    lim = bqcd_hip.multi_block_limit
    assert lim(vh["24x48"], "f32") == 12 and lim(vh["24x48"], "f16") == 12
    assert lim(vh["32x64"], "f32") == 12 and lim(vh["32x64"], "f16") == 12
    assert lim(vh["48x96"], "f32") == 8 and lim(vh["48x96"], "f16") == 12   # 6.1 GB for 12 in FP32
    assert lim(vh["64x128"], "f32") == 2 and lim(vh["64x128"], "f16") == 4
    for p in ("f32", "f16"):
        for k, v in vh.items():
            nr = lim(v, p)
            assert 12 * v * nr * (8 if p == "f32" else 4) < 2 ** 32, (k, p)
    assert lim(0, "f16") == 0 and lim(vh["24x48"], "f64") == 0


PLUGIN_ROUTINES = ["d", "d_dag", "version_of_d", "dsf_xf", "dsf_xb", "dsf_yf", "dsf_yb", "dsf_zf", "dsf_zb",
                   "dsf_tf", "dsf_tb", "clover_init", "clover_mult_a", "clover_mult_ao", "clover_mult_b",
                   "clover_dsf", "clover_dsd", "clover_action", "allocate_clover_field_a",
                   "allocate_clover_field_b"]


def test_plugin_file_provides_every_libd_and_libclover_routine():
    """fortran/hip_plugins.F90 defines each routine BQCD links from LIBD
    (d/Makefile:113-129) and LIBCLOVER (clover/Makefile:44-71), as external
    procedures with the reference's names; where the plugin-linked driver has
    been built, its object file defines them"""
    text = open(os.path.join(REPO, "bqcd-r399-quda_amd", "fortran", "hip_plugins.F90")).read()
    defined = set(re.findall(r"^\s*(?:REAL\s+|integer\s+)?(?:function|subroutine)\s+(\w+)\s*\(", text,
                             flags=re.M | re.I))
    missing = [n for n in PLUGIN_ROUTINES if n not in defined]
    assert not missing, missing
    obj = os.path.join(REPO, "oracle", "_ref", "build_seq", "hip_plugins.o")
    if os.path.exists(obj):
        nm = subprocess.run(["nm", obj], capture_output=True, text=True).stdout
        syms = {line.split()[-1] for line in nm.splitlines() if " T " in line}
        assert not [n for n in PLUGIN_ROUTINES if n + "_" not in syms]
