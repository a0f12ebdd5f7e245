"""ctypes binding of the CPU oracle (oracle/librsref.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg as the parity checker; the product package never
imports it.

Shards are numpy uint8 arrays; ``None`` (or a zero-length array) marks an
absent shard exactly like a nil / zero-length slice in the reference.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librsref.so")

OK = 0
ERR_INV_SHARD_NUM = -1
ERR_MAX_SHARD_NUM = -2
ERR_NOT_SUPPORTED = -3
ERR_TOO_FEW_SHARDS = -4
ERR_INVALID_INPUT = -5
ERR_SHARD_NO_DATA = -6
ERR_SHARD_SIZE = -7
ERR_INVALID_SHARD_SIZE = -8
ERR_SINGULAR = -11
ERR_MATRIX = -12
ERR_NOT_SQUARE = -13

MATRIX_DEFAULT, MATRIX_CAUCHY, MATRIX_PAR1, MATRIX_JERASURE, MATRIX_XOR = range(5)


def build() -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _u8p = ctypes.POINTER(ctypes.c_uint8)
        _lib.rsref_gal_mul.restype = ctypes.c_uint8
        _lib.rsref_gal_exp.restype = ctypes.c_uint8
        _lib.rsref_gal_divide.restype = ctypes.c_uint8
        _lib.rsref_leo16_mul_log.restype = ctypes.c_uint16
        _lib.rsref_leo16_add_mod.restype = ctypes.c_uint16
        _lib.rsref_gf16_mul.restype = ctypes.c_uint16
        _lib.rsref_leo8_mul_log.restype = ctypes.c_uint8
        _lib.rsref_leo8_add_mod.restype = ctypes.c_uint8
        del _u8p
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class _Shards:
    """Pointer + length arrays for a list of shards (None = absent)."""

    def __init__(self, shards, size_for_missing: int = 0):
        self.keep = []
        n = len(shards)
        self.ptrs = (ctypes.c_void_p * max(n, 1))()
        self.lens = (ctypes.c_size_t * max(n, 1))()
        self.bufs = []
        for i, s in enumerate(shards):
            if s is None or len(s) == 0:
                buf = np.zeros(max(size_for_missing, 1), dtype=np.uint8)
                self.keep.append(buf)
                self.bufs.append(buf)
                self.ptrs[i] = buf.ctypes.data
                self.lens[i] = 0
            else:
                assert s.dtype == np.uint8 and s.flags["C_CONTIGUOUS"]
                self.bufs.append(s)
                self.ptrs[i] = s.ctypes.data
                self.lens[i] = len(s)
        self.n = n


class _NilShards:
    """Pointer + length arrays where None is Go's nil (a NULL pointer)."""

    def __init__(self, shards):
        n = len(shards)
        self.ptrs = (ctypes.c_void_p * max(n, 1))()
        self.lens = (ctypes.c_size_t * max(n, 1))()
        for i, s in enumerate(shards):
            if s is None:
                self.ptrs[i] = None
                self.lens[i] = 0
            else:
                assert s.dtype == np.uint8 and s.flags["C_CONTIGUOUS"]
                self.ptrs[i] = s.ctypes.data if len(s) else ctypes.addressof(_EMPTY)
                self.lens[i] = len(s)
        self.n = n


_EMPTY = (ctypes.c_uint8 * 1)()  # backing for non-nil empty slices


def gf8_tables():
    e = np.zeros(256, np.uint8)
    lg = np.zeros(256, np.uint8)
    inv = np.zeros(256, np.uint8)
    mul = np.zeros(65536, np.uint8)
    lo = np.zeros(4096, np.uint8)
    hi = np.zeros(4096, np.uint8)
    gfni = np.zeros(256, np.uint64)
    lib().rsref_gf8_tables(_ptr(e), _ptr(lg), _ptr(inv), _ptr(mul), _ptr(lo), _ptr(hi), _ptr(gfni))
    return dict(exp=e, log=lg, inv=inv, mul=mul.reshape(256, 256), low=lo.reshape(256, 16),
                high=hi.reshape(256, 16), gfni=gfni)


def gal_mul(a: int, b: int) -> int:
    return lib().rsref_gal_mul(ctypes.c_uint8(a), ctypes.c_uint8(b))


def gal_exp(a: int, n: int) -> int:
    return lib().rsref_gal_exp(ctypes.c_uint8(a), ctypes.c_int(n))


def gal_mul_slice(c: int, inp: np.ndarray, out: np.ndarray, xor: bool) -> None:
    lib().rsref_gal_mul_slice(ctypes.c_uint8(c), _ptr(inp), _ptr(out), ctypes.c_size_t(len(inp)),
                              ctypes.c_int(int(xor)))


def matrix_multiply(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint8)
    b = np.ascontiguousarray(b, np.uint8)
    out = np.zeros((a.shape[0], b.shape[1]), np.uint8)
    err = lib().rsref_matrix_multiply(_ptr(a), a.shape[0], a.shape[1], _ptr(b), b.shape[0],
                                      b.shape[1], _ptr(out))
    if err:
        raise ValueError(err)
    return out


def matrix_invert(m: np.ndarray):
    m = np.ascontiguousarray(m, np.uint8)
    out = np.zeros((m.shape[0], m.shape[0]), np.uint8)
    err = lib().rsref_matrix_invert(_ptr(m), m.shape[0], m.shape[1], _ptr(out))
    return err, out


def fast_isa() -> int:
    return lib().rsref_fast_isa()


def code_some_shards_fast(rows: np.ndarray, inputs, outputs, threads: int = 1, isa: int = -1) -> None:
    """The reference's SIMD path (rsref_fast.c): out = rows * in over `threads`."""
    rows = np.ascontiguousarray(rows, np.uint8)
    ins = _Shards(inputs)
    outs = _Shards(outputs)
    lib().rsref_code_some_shards_fast(_ptr(rows), rows.shape[0], rows.shape[1], ins.ptrs, outs.ptrs,
                                      ctypes.c_size_t(len(inputs[0])), ctypes.c_int(threads),
                                      ctypes.c_int(isa))


def codec_matrix(k: int, m: int, kind: int = MATRIX_DEFAULT, custom: np.ndarray | None = None):
    out = np.zeros((k + m, k), np.uint8)
    if custom is not None:
        custom = np.ascontiguousarray(custom, np.uint8)
        err = lib().rsref_codec_matrix(k, m, kind, _ptr(custom), custom.shape[0], custom.shape[1],
                                       _ptr(out))
    else:
        err = lib().rsref_codec_matrix(k, m, kind, None, 0, 0, _ptr(out))
    if err:
        raise ValueError(err)
    return out


def code_some_shards(rows: np.ndarray, inputs, outputs, clear: bool = True) -> None:
    rows = np.ascontiguousarray(rows, np.uint8)
    ins = _Shards(inputs)
    outs = _Shards(outputs)
    lib().rsref_code_some_shards(_ptr(rows), rows.shape[0], rows.shape[1], ins.ptrs, outs.ptrs,
                                 ctypes.c_size_t(len(inputs[0])), ctypes.c_int(int(clear)))


class Classic:
    """Oracle of reedSolomon (reedsolomon.go) for one (k, m, matrix)."""

    def __init__(self, k: int, m: int, kind: int = MATRIX_DEFAULT, custom=None):
        self.k, self.m = k, m
        self.matrix = codec_matrix(k, m, kind, custom)

    def encode(self, shards) -> int:
        s = _Shards(shards)
        return lib().rsref_encode(_ptr(self.matrix), self.k, self.m, s.ptrs, s.lens, s.n)

    def encode_idx(self, data: np.ndarray, idx: int, parity) -> int:
        s = _Shards(parity)
        return lib().rsref_encode_idx(_ptr(self.matrix), self.k, self.m, _ptr(data),
                                      ctypes.c_size_t(len(data)), idx, s.ptrs, s.lens, s.n)

    def verify(self, shards):
        s = _Shards(shards)
        ok = ctypes.c_int(0)
        err = lib().rsref_verify(_ptr(self.matrix), self.k, self.m, s.ptrs, s.lens, s.n,
                                 ctypes.byref(ok))
        return err, bool(ok.value)

    def _reconstruct(self, shards, data_only, required):
        size = max((len(x) for x in shards if x is not None), default=0)
        s = _Shards(shards, size)
        filled = np.zeros(max(len(shards), 1), np.uint8)
        if required is not None:
            # A k-entry mask: entries past k read as false (the reference
            # indexes past the slice and panics when a parity shard is missing).
            req = np.zeros(max(len(shards), len(required)), np.uint8)
            req[:len(required)] = np.asarray(required, np.uint8)
            err = lib().rsref_reconstruct(_ptr(self.matrix), self.k, self.m, s.ptrs, s.lens, s.n,
                                          int(data_only), _ptr(req), len(required), _ptr(filled))
        else:
            err = lib().rsref_reconstruct(_ptr(self.matrix), self.k, self.m, s.ptrs, s.lens, s.n,
                                          int(data_only), None, 0, _ptr(filled))
        if err == OK:
            for i in range(len(shards)):
                if filled[i]:
                    shards[i] = s.bufs[i][:size]
        return err

    def reconstruct(self, shards) -> int:
        return self._reconstruct(shards, False, None)

    def reconstruct_data(self, shards) -> int:
        return self._reconstruct(shards, True, None)

    def update(self, shards, new_data) -> int:
        """reedsolomon.go:762 Update (None = nil); modifies shards in place."""
        s, nd = _NilShards(shards), _NilShards(new_data)
        return lib().rsref_update(_ptr(self.matrix), self.k, self.m, s.ptrs, s.lens, s.n, nd.ptrs, nd.lens, nd.n)

    def decode_idx(self, dst, expect_input, inp) -> int:
        """reedsolomon.go:1464 DecodeIdx (None = nil; expect_input None = merge mode)."""
        d, i = _NilShards(dst), _NilShards(inp)
        if expect_input is None:
            return lib().rsref_decode_idx(_ptr(self.matrix), self.k, self.m, d.ptrs, d.lens, d.n, None, 0,
                                          i.ptrs, i.lens, i.n)
        ex = np.asarray(expect_input, np.uint8)
        return lib().rsref_decode_idx(_ptr(self.matrix), self.k, self.m, d.ptrs, d.lens, d.n, _ptr(ex), len(ex),
                                      i.ptrs, i.lens, i.n)

    def reconstruct_some(self, shards, required) -> int:
        # reedsolomon.go:1655-1660
        if len(required) != self.k and len(required) != self.k + self.m:
            return ERR_INVALID_INPUT
        return self._reconstruct(shards, len(required) == self.k, required)


class Leopard:
    """Oracle of leopardFF16 (bits=16) / leopardFF8 (bits=8)."""

    def __init__(self, k: int, m: int, bits: int = 16):
        self.k, self.m, self.bits = k, m, bits
        self._enc = lib().rsref_leo16_encode if bits == 16 else lib().rsref_leo8_encode
        self._ver = lib().rsref_leo16_verify if bits == 16 else lib().rsref_leo8_verify
        self._rec = lib().rsref_leo16_reconstruct if bits == 16 else lib().rsref_leo8_reconstruct

    def encode(self, shards) -> int:
        s = _Shards(shards)
        return self._enc(self.k, self.m, s.ptrs, s.lens, s.n)

    def verify(self, shards):
        s = _Shards(shards)
        ok = ctypes.c_int(0)
        err = self._ver(self.k, self.m, s.ptrs, s.lens, s.n, ctypes.byref(ok))
        return err, bool(ok.value)

    def _reconstruct(self, shards, recover_all):
        size = max((len(x) for x in shards if x is not None), default=0)
        s = _Shards(shards, size)
        filled = np.zeros(max(len(shards), 1), np.uint8)
        err = self._rec(self.k, self.m, s.ptrs, s.lens, s.n, int(recover_all), _ptr(filled))
        if err == OK:
            for i in range(len(shards)):
                if filled[i]:
                    shards[i] = s.bufs[i][:size]
        return err

    def reconstruct(self, shards) -> int:
        return self._reconstruct(shards, True)

    def reconstruct_data(self, shards) -> int:
        return self._reconstruct(shards, False)

    def reconstruct_some(self, shards, required) -> int:
        # leopard.go:378-383: all only when len(required) == total
        return self._reconstruct(shards, len(required) == self.k + self.m)


def leo16_tables():
    e = np.zeros(65536, np.uint16)
    lg = np.zeros(65536, np.uint16)
    sk = np.zeros(65535, np.uint16)
    w = np.zeros(65536, np.uint16)
    lib().rsref_leo16_tables(_ptr(e), _ptr(lg), _ptr(sk), _ptr(w))
    return dict(exp=e, log=lg, skew=sk, walsh=w)


def leo8_tables():
    e = np.zeros(256, np.uint8)
    lg = np.zeros(256, np.uint8)
    sk = np.zeros(255, np.uint8)
    w = np.zeros(256, np.uint8)
    lib().rsref_leo8_tables(_ptr(e), _ptr(lg), _ptr(sk), _ptr(w))
    return dict(exp=e, log=lg, skew=sk, walsh=w)


def leo16_mul_log(a: int, log_b: int) -> int:
    return lib().rsref_leo16_mul_log(ctypes.c_uint16(a), ctypes.c_uint16(log_b))


def leo16_add_mod(a: int, b: int) -> int:
    return lib().rsref_leo16_add_mod(ctypes.c_uint16(a), ctypes.c_uint16(b))


def gf16_mul(a: int, b: int) -> int:
    return lib().rsref_gf16_mul(ctypes.c_uint16(a), ctypes.c_uint16(b))


def leo8_mul_log(a: int, log_b: int) -> int:
    return lib().rsref_leo8_mul_log(ctypes.c_uint8(a), ctypes.c_uint8(log_b))


def leo8_add_mod(a: int, b: int) -> int:
    return lib().rsref_leo8_add_mod(ctypes.c_uint8(a), ctypes.c_uint8(b))


def leo16_errlocs(k: int, m: int, erased) -> np.ndarray:
    er = np.asarray(erased, np.uint8)
    out = np.zeros(65536, np.uint16)
    lib().rsref_leo16_errlocs(k, m, _ptr(er), 1, _ptr(out))
    return out


def gf16_mul_slice_xor(scalar: int, inp: np.ndarray, out: np.ndarray) -> None:
    """mulslice.go:73-95 GF16MulSliceXor on the split layout (each 64-byte
    chunk: 32 low bytes, then 32 high bytes): out ^= scalar * in, elementwise
    mulLog(x, logLUT[scalar]) (leopard.go:1000-1006)."""
    if len(inp) == 0 or scalar == 0:
        return
    t = leo16_tables()
    exp, log = t["exp"].astype(np.int64), t["log"].astype(np.int64)
    c = inp.reshape(-1, 2, 32)
    x = c[:, 0, :].astype(np.int64) | (c[:, 1, :].astype(np.int64) << 8)
    s = log[x] + int(log[scalar])
    s = (s + (s >> 16)) & 0xFFFF  # addMod: partial reduction (expLUT[modulus] == expLUT[0])
    prod = np.where(x == 0, 0, exp[s])
    o = out.reshape(-1, 2, 32)
    o[:, 0, :] ^= (prod & 0xFF).astype(np.uint8)
    o[:, 1, :] ^= (prod >> 8).astype(np.uint8)
