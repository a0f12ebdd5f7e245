"""CPU checks of the C-ABI boundary: the library loads, exports every entry
point include/rsx.h declares, and reports the reference's error codes for
argument errors that are decided before any device work."""
import ctypes
import os
import re

import pytest

import reedsolomon_amd as rs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(rsx_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert "rsx_encode" in names and "rsx_reconstruct" in names
    missing = [n for n in sorted(names) if not hasattr(rs.lib, n)]
    assert not missing, missing


def test_error_codes_match_header():
    text = open(os.path.join(ROOT, "include", "rsx.h")).read()
    codes = dict((name, int(v)) for name, v in re.findall(r"#define (RSX_ERR_\w+) (-\d+)", text))
    assert codes["RSX_ERR_TOO_FEW_SHARDS"] == rs.ErrTooFewShards.code
    assert codes["RSX_ERR_SHARD_SIZE"] == rs.ErrShardSize.code
    assert codes["RSX_ERR_INV_SHARD_NUM"] == rs.ErrInvShardNum.code
    assert codes["RSX_ERR_MAX_SHARD_NUM"] == rs.ErrMaxShardNum.code
    assert codes["RSX_ERR_INVALID_SHARD_SIZE"] == rs.ErrInvalidShardSize.code


def test_new_shape_errors_need_no_device(golden_dir):
    # TestNew (reedsolomon_test.go:2070-2095): shape errors come before any
    # device work, so they hold without a GPU.
    import json
    cases = json.load(open(os.path.join(golden_dir, "kat.json")))["new_errors"]["cases"]
    for c in cases:
        if c["err"] is None:
            continue
        with pytest.raises(getattr(rs, c["err"])):
            rs.New(c["k"], c["m"])


def test_new_without_device_fails_loudly():
    if rs.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(rs.ErrDevice):
        rs.New(10, 4)


def test_leopard8_invalid_shape_rejected():
    # k + ceilPow2(m) > 256 indexes past the reference's GF(2^8) tables.
    with pytest.raises(rs.ErrMaxShardNum):
        rs.New(200, 56, rs.WithLeopardGF(True))


@pytest.mark.parametrize("k,m", [(65000, 536), (1, 65535)])
def test_leopard16_out_of_field_shape_rejected(k, m):
    # k + m <= 65536 passes the reference's newFF16 check, but k + ceilPow2(m)
    # > 65536 indexes past its fftSkew table.
    with pytest.raises(rs.ErrMaxShardNum):
        rs.New(k, m, rs.WithLeopardGF16(True))
