"""The C-ABI library builds for gfx950, loads, and exports every entry point that
include/irc_hip.h declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from ironclust_amd import libirc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "irc_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(irc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = libirc.load()
    syms = declared_symbols()
    assert len(syms) >= 13
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(syms) == sorted(libirc.EXPORTED)


def test_version_string():
    assert b"gfx950" in libirc.load().irc_version()


def test_library_is_gfx950_code_object():
    data = open(libirc.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_options_default():
    o = libirc.make_options()
    assert o.precision == libirc.PREC_AUTO and o.force_fail == 0 and o.k_over == 0


def test_ctx_create_rejects_bad_device():
    lib = libirc.load()
    h = ctypes.c_void_p()
    assert lib.irc_ctx_create(ctypes.byref(h), 1 << 20) != libirc.IRC_OK
    assert not h.value


def test_no_oracle_in_product():
    # the product package never imports or links the oracle
    pkg = os.path.join(ROOT, "ironclust_amd")
    for dp, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                src = open(os.path.join(dp, f), errors="replace").read()
                assert "oracle" not in src.replace("oracle is", ""), f
