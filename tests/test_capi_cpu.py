"""C-ABI checks that need no GPU: exported symbols, .gr I/O, error paths."""
import os
import re

import numpy as np
import pytest

import grus_framework_amd as grus
from grus_framework_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
HEADER = os.path.join(ROOT, "include", "grus_c.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(grus_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_header_function():
    L = capi.lib()
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes table covers exactly the header
    assert sorted(n for n, _, _ in capi.SIGNATURES) == names


def test_version_string():
    assert b"gfx950" in capi.lib().grus_version()


@pytest.mark.parametrize("name", ["odd_weighted", "even_weighted", "odd_unweighted", "isolated"])
def test_gr_roundtrip_byte_identical(tmp_path, name, oracle):
    src = os.path.join(GOLD, name + ".gr")
    V, E, weighted = grus.gr_header(src)
    off, col, w = grus.gr_read(src, need_weight=True)
    # the product reader agrees with the oracle reader
    g, _ = oracle.read_gr(src)
    assert (off == g.off).all() and (col == g.col).all() and (w == g.w).all()
    # unweighted files read back as weight 1 (graph_loader.cuh:95-101)
    if not weighted:
        assert (w == 1).all()
    out = tmp_path / "x.gr"
    grus.gr_write(str(out), off, col, w if weighted else None)
    assert open(src, "rb").read() == open(out, "rb").read()


def test_gr_errors(tmp_path):
    with pytest.raises(grus.GrusError) as e:
        grus.gr_header(str(tmp_path / "missing.gr"))
    assert e.value.code == capi.E_IO
    # truncated: header claims more edges than the file holds
    src = open(os.path.join(GOLD, "odd_weighted.gr"), "rb").read()
    bad = tmp_path / "trunc.gr"
    bad.write_bytes(src[:60])
    with pytest.raises(grus.GrusError) as e:
        grus.gr_read(str(bad))
    assert e.value.code == capi.E_FORMAT
    # wrong version
    ver = tmp_path / "ver.gr"
    ver.write_bytes(b"\x02" + src[1:])
    with pytest.raises(grus.GrusError) as e:
        grus.gr_header(str(ver))
    assert e.value.code == capi.E_FORMAT
    # destination out of range
    rng = bytearray(src)
    rng[64:68] = (99).to_bytes(4, "little")
    r = tmp_path / "range.gr"
    r.write_bytes(bytes(rng))
    with pytest.raises(grus.GrusError) as e:
        grus.gr_read(str(r))
    assert e.value.code == capi.E_FORMAT


def test_null_arguments_are_errors_not_crashes():
    L = capi.lib()
    assert L.grus_init(0, None) == capi.E_INVALID
    assert L.grus_graph_info_get(None, None) == capi.E_INVALID
    assert L.grus_bfs(None, 0, 0, None, None) == capi.E_INVALID
    assert L.grus_server_poll(None, None, 0) == capi.E_INVALID
    assert b"null argument" in L.grus_last_error()


def test_no_device_reported_cleanly():
    if grus.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(grus.GrusError) as e:
        grus.Context(0)
    assert e.value.code == capi.E_NODEVICE
