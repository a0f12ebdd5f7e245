"""MI355X erasure coding with the klauspost/reedsolomon Encoder API.

This module mirrors the reference's Go API (reedsolomon.go:25-165,
options.go) over the C ABI of include/rsx.h, so code and tests read like the
reference's:

    enc = New(10, 4, WithMI355X(0))
    enc.Encode(shards)            # list of uint8 numpy arrays or torch tensors
    shards[3] = None
    enc.Reconstruct(shards)
    ok = enc.Verify(shards)

Shards are numpy ``uint8`` arrays (host memory, staged through HBM) or
``torch.uint8`` tensors on the GPU (used in place).  ``None`` or a
zero-length shard is absent, like a nil slice.  All compute runs in the HIP
library; importing fails loudly if it has not been built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librsx.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `python reedsolomon_amd/build.py`")

# PyTorch-ROCm ships its own libamdhip64 (soname libamdhip64.so.7).  Loading
# torch first makes librsx.so bind to that same runtime, so torch tensors and
# this library share one HIP runtime in the process.
try:
    import torch  # noqa: F401
except ImportError:
    pass

lib = ctypes.CDLL(LIB_PATH)


class _Options(ctypes.Structure):
    _fields_ = [
        ("leopard", ctypes.c_int),
        ("matrix", ctypes.c_int),
        ("fast_one_parity", ctypes.c_int),
        ("inversion_cache", ctypes.c_int),
        ("custom_matrix", ctypes.c_void_p),
        ("custom_rows", ctypes.c_int),
        ("custom_cols", ctypes.c_int),
        ("device", ctypes.c_int),
    ]


_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
_SZ = ctypes.POINTER(ctypes.c_size_t)
_U8 = ctypes.POINTER(ctypes.c_uint8)

lib.rsx_options_default.argtypes = [ctypes.POINTER(_Options)]
lib.rsx_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Options), ctypes.POINTER(_P)]
lib.rsx_free.argtypes = [_P]
for _n in ("rsx_shard_size_multiple", "rsx_data_shards", "rsx_parity_shards", "rsx_total_shards",
           "rsx_engine", "rsx_decode_cache_size", "rsx_synchronize"):
    getattr(lib, _n).argtypes = [_P]
lib.rsx_coding_matrix.argtypes = [_P, _U8]
lib.rsx_encode.argtypes = [_P, _PP, _SZ, ctypes.c_int]
lib.rsx_encode_idx.argtypes = [_P, _P, ctypes.c_size_t, ctypes.c_int, _PP, _SZ, ctypes.c_int]
lib.rsx_verify.argtypes = [_P, _PP, _SZ, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
lib.rsx_reconstruct.argtypes = [_P, _PP, _SZ, ctypes.c_int, _U8]
lib.rsx_reconstruct_data.argtypes = [_P, _PP, _SZ, ctypes.c_int, _U8]
lib.rsx_reconstruct_some.argtypes = [_P, _PP, _SZ, ctypes.c_int, _U8, ctypes.c_int, _U8]
lib.rsx_reconstruct_noalloc.argtypes = [_P, _PP, _SZ, ctypes.c_int, ctypes.c_int, _U8]
lib.rsx_update.argtypes = [_P, _PP, _SZ, ctypes.c_int, _PP, _SZ, ctypes.c_int]
lib.rsx_decode_idx.argtypes = [_P, _PP, _SZ, ctypes.c_int, _U8, ctypes.c_int, _PP, _SZ, ctypes.c_int]
lib.rsx_gal_mul_slice.argtypes = [ctypes.c_int, ctypes.c_uint8, _P, _P, ctypes.c_size_t, ctypes.c_int]
lib.rsx_inv.argtypes = [ctypes.c_uint8]
lib.rsx_inv.restype = ctypes.c_uint8
lib.rsx_gf16_mul.argtypes = [ctypes.c_uint16, ctypes.c_uint16]
lib.rsx_gf16_mul.restype = ctypes.c_uint16
_U16 = ctypes.POINTER(ctypes.c_uint16)
lib.rsx_leopard_tables.argtypes = [ctypes.c_int, _U16, _U16, _U16, _U16]
lib.rsx_gf16_mul_slice_xor.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint16), ctypes.c_int, _P, _PP, _SZ,
                                       ctypes.c_size_t]
_READ_FN = ctypes.CFUNCTYPE(ctypes.c_longlong, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t)
_WRITE_FN = ctypes.CFUNCTYPE(ctypes.c_longlong, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t)
lib.rsx_stream_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(_P)]
lib.rsx_stream_free.argtypes = [_P]
lib.rsx_stream_free.restype = None
lib.rsx_stream_encode.argtypes = [_P, _READ_FN, _P, _WRITE_FN, _P, ctypes.POINTER(ctypes.c_int)]
lib.rsx_stream_set_concurrency.argtypes = [_P, ctypes.c_int, ctypes.c_int]
lib.rsx_stream_verify.argtypes = [_P, _READ_FN, _P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
lib.rsx_stream_reconstruct.argtypes = [_P, _U8, _U8, _READ_FN, _P, _WRITE_FN, _P, ctypes.POINTER(ctypes.c_int)]
lib.rsx_set_stream.argtypes = [_P, _P]
lib.rsx_get_stream.argtypes = [_P]
lib.rsx_get_stream.restype = _P
lib.rsx_last_error.restype = ctypes.c_char_p
lib.rsx_device_count.restype = ctypes.c_int


# --- errors: one class per reference sentinel (reedsolomon.go:216-1862) ---
class RSError(Exception):
    code = 0


def _err(name: str, code: int, doc: str):
    cls = type(name, (RSError,), {"code": code, "__doc__": doc})
    globals()[name] = cls
    return cls


_ERRORS = {
    -1: _err("ErrInvShardNum", -1, "cannot create Encoder with less than one data shard or less than zero parity shards"),
    -2: _err("ErrMaxShardNum", -2, "cannot create Encoder with more than 256 data+parity shards"),
    -3: _err("ErrNotSupported", -3, "operation not supported"),
    -4: _err("ErrTooFewShards", -4, "too few shards given"),
    -5: _err("ErrInvalidInput", -5, "invalid input"),
    -6: _err("ErrShardNoData", -6, "no shard data"),
    -7: _err("ErrShardSize", -7, "shard sizes do not match"),
    -8: _err("ErrInvalidShardSize", -8, "invalid shard size"),
    -9: _err("ErrShortData", -9, "not enough data to fill the number of requested shards"),
    -10: _err("ErrReconstructRequired", -10, "reconstruction required"),
    -14: _err("ErrReconstructMismatch", -14, "valid shards and fill shards are mutually exclusive"),
    -11: _err("ErrSingular", -11, "matrix is singular"),
    -12: _err("ErrMatrix", -12, "invalid custom coding matrix"),
    -20: _err("ErrDevice", -20, "HIP device error"),
    -21: _err("ErrNoMem", -21, "device allocation failed"),
    -22: _err("ErrMixedMemory", -22, "host and device shards mixed in one call"),
    -23: _err("ErrCancelled", -23, "request cancelled before admission"),
    -24: _err("ErrOverBudget", -24, "request rejected: larger than the HBM budget or the queue is full"),
}


def _check(code: int) -> None:
    if code != 0:
        cls = _ERRORS.get(code, RSError)
        raise cls(f"{cls.__name__}: {lib.rsx_last_error().decode(errors='replace')}")


# --- options (options.go) ---
def WithLeopardGF16(enabled: bool = True):
    def f(o):
        o.leopard = 1 if enabled else 0
    return f


def WithLeopardGF(enabled: bool = True):
    def f(o):
        o.leopard = 2 if enabled else 0
    return f


def WithCauchyMatrix():
    def f(o):
        o.matrix = 1
    return f


def WithPAR1Matrix():
    def f(o):
        o.matrix = 2
    return f


def WithJerasureMatrix():
    def f(o):
        o.matrix = 3
    return f


def WithFastOneParityMatrix():
    def f(o):
        o.fast_one_parity = 1
    return f


def WithInversionCache(enabled: bool = True):
    def f(o):
        o.inversion_cache = int(enabled)
    return f


def WithCustomMatrix(matrix):
    m = np.ascontiguousarray(np.asarray(matrix, dtype=np.uint8))

    def f(o):
        o.custom_matrix = m.ctypes.data
        o.custom_rows, o.custom_cols = m.shape
        o._keep = m
    return f


def WithMI355X(device: int = 0):
    """The new option: run the codec on HIP device `device`."""
    def f(o):
        o.device = device
    return f


def _noop(*_a, **_k):
    return lambda o: None


# CPU-tuning options of the reference have no meaning on the GPU.
WithMaxGoroutines = WithMinSplitSize = _noop
WithSSSE3 = WithAVX2 = WithSSE2 = WithAVX512 = WithGFNI = WithAVXGFNI = _noop
WithWorkAllocator = _noop


# Stream options (options.go:98-146, streaming.go:157-163, 178-183): kept on
# the option struct's Python side and applied by NewStream.
def WithAutoGoroutines(shard_size: int):
    """Also the stream block size when WithStreamBlockSize is not given."""
    def f(o):
        o._shard_size = shard_size
    return f


def WithStreamBlockSize(n: int):
    def f(o):
        o._stream_bs = n
    return f


def WithConcurrentStreams(enabled: bool = True):
    def f(o):
        o._conc_reads = o._conc_writes = bool(enabled)
    return f


def WithConcurrentStreamReads(enabled: bool = True):
    def f(o):
        o._conc_reads = bool(enabled)
    return f


def WithConcurrentStreamWrites(enabled: bool = True):
    def f(o):
        o._conc_writes = bool(enabled)
    return f


def device_count() -> int:
    return lib.rsx_device_count()


# --- shard marshalling ---
def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _shard_info(x):
    """(address, len, is_device) of one shard; None/empty -> (0, 0, None)."""
    if x is None:
        return 0, 0, None
    if _is_torch(x):
        if x.numel() == 0:
            return 0, 0, None
        assert x.dtype.itemsize == 1 and x.is_contiguous(), "shards must be contiguous uint8"
        return x.data_ptr(), x.numel(), bool(x.is_cuda)
    a = x
    if len(a) == 0:
        return 0, 0, None
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"], "shards must be contiguous uint8"
    return a.ctypes.data, len(a), False


class _Marshal:
    def __init__(self, shards):
        n = len(shards)
        self.n = n
        self.ptrs = (ctypes.c_void_p * max(n, 1))()
        self.lens = (ctypes.c_size_t * max(n, 1))()
        self.device = None
        for i, s in enumerate(shards):
            p, ln, dev = _shard_info(s)
            self.ptrs[i] = p or None
            self.lens[i] = ln
            if dev is not None and self.device is None:
                self.device = dev


_EMPTY = (ctypes.c_uint8 * 1)()  # address for non-nil empty shards


def _after_torch(stream_ptr) -> None:
    """Orders work on the library stream `stream_ptr` after everything queued
    so far on this thread's current torch stream (the producer of the device
    tensors passed in), without blocking the host."""
    import torch
    cur = torch.cuda.current_stream()
    if cur.cuda_stream != stream_ptr:
        torch.cuda.ExternalStream(stream_ptr).wait_stream(cur)


def _torch_sync() -> None:
    """Waits for this thread's current torch stream when torch has a GPU
    (device tensors handed to the library's own streams must be complete)."""
    import sys
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_available():
        torch.cuda.current_stream().synchronize()


def _zeros_like(x, n):
    if _is_torch(x):
        import torch
        return torch.zeros(n, dtype=torch.uint8, device=x.device)
    return np.zeros(n, np.uint8)


def split_shards(data, data_shards: int, total_shards: int, multiple: int = 1):
    """Encoder.Split (reedsolomon.go:1803-1860; Leopard leopard.go:319-376,
    leopard8.go:293-350, where `multiple` is 64): perShard =
    ceil(len / data_shards) rounded up to `multiple`; shard i holds bytes
    [i * perShard, (i + 1) * perShard) of `data` (views where the input has
    them), zero-padded past its end, and the remaining shards, parity
    included, are zero.  Go's reuse of spare slice capacity has no Python
    counterpart; the bytes of every shard are the reference's."""
    n = len(data)
    if n == 0:
        raise ErrShortData()
    if total_shards == 1 and n % multiple == 0:
        return [data]
    per = (n + data_shards - 1) // data_shards
    per = (per + multiple - 1) // multiple * multiple
    full = n // per
    out = [data[i * per:(i + 1) * per] for i in range(min(full, total_shards))]
    rest = data[full * per:]
    for _ in range(total_shards - len(out)):
        pad = _zeros_like(data, per)
        m = min(len(rest), per)
        if m:
            pad[:m] = rest[:m]
            rest = rest[m:]
        out.append(pad)
    return out


def join_shards(dst, shards, data_shards: int, out_size: int) -> None:
    """Encoder.Join (reedsolomon.go:1875-1910, leopard.go:276-310): writes the
    first out_size bytes of the data shards to dst; ErrTooFewShards with fewer
    than data_shards entries, ErrReconstructRequired for a None data shard
    before out_size is covered, ErrShortData when they hold fewer bytes."""
    if len(shards) < data_shards:
        raise ErrTooFewShards()
    shards = shards[:data_shards]
    size = 0
    for sh in shards:
        if sh is None:
            raise ErrReconstructRequired()
        size += len(sh)
        if size >= out_size:
            break
    if size < out_size:
        raise ErrShortData()
    write = out_size
    for sh in shards:
        if write == 0:
            return
        b = _host_bytes(sh)
        if write < len(b):
            dst.write(b[:write])
            return
        dst.write(b)
        write -= len(b)


def write_shard_files(path: str, shards, out_dir=None):
    """The reference's file layout for shards (examples/simple-encoder.go:91-101):
    shard i of `path` is written to `<out_dir or dir of path>/<file>.<i>`.
    Returns the file names."""
    d, f = os.path.split(path)
    d = out_dir if out_dir else d
    names = []
    for i, sh in enumerate(shards):
        name = os.path.join(d, f"{f}.{i}")
        with open(name, "wb") as fh:
            fh.write(_host_bytes(sh))
        names.append(name)
    return names


def read_shard_files(path: str, total_shards: int):
    """examples/simple-decoder.go:76-85: reads `<path>.<i>` for every shard;
    a file that cannot be read is None (a missing shard)."""
    shards = []
    for i in range(total_shards):
        try:
            with open(f"{path}.{i}", "rb") as fh:
                shards.append(np.frombuffer(fh.read(), dtype=np.uint8).copy())
        except OSError:
            shards.append(None)
    return shards


def encode_file(enc, path: str, out_dir=None):
    """examples/simple-encoder.go:77-101: Split the file, Encode, write
    `<file>.<i>` shard files.  Returns the file names."""
    with open(path, "rb") as fh:
        data = np.frombuffer(fh.read(), dtype=np.uint8).copy()
    shards = enc.Split(data)
    enc.Encode(shards)
    return write_shard_files(path, shards, out_dir)


def decode_files(enc, path: str, out_path: str, out_size=None) -> bool:
    """examples/simple-decoder.go:72-123: read `<path>.<i>`, Verify, Reconstruct
    if needed (then Verify again), Join into out_path.  out_size defaults to
    the example's len(shards[0]) * DataShards (the original size is not
    stored).  Returns whether a reconstruction was needed."""
    shards = read_shard_files(path, enc.TotalShards())
    rebuilt = False
    try:
        ok = enc.Verify(shards)
    except RSError:
        ok = False  # a missing shard: Verify reports it as an error
    if not ok:
        enc.Reconstruct(shards)
        rebuilt = True
        if not enc.Verify(shards):
            raise RSError("verification failed after reconstruction, data likely corrupted")
    size = next(len(s) for s in shards if s is not None)
    with open(out_path, "wb") as fh:
        enc.Join(fh, shards, out_size if out_size is not None else size * enc.DataShards())
    return rebuilt


def _host_bytes(x) -> bytes:
    if _is_torch(x):
        return x.cpu().numpy().tobytes()
    return np.asarray(x, np.uint8).tobytes()


class _NilMarshal:
    """Pointer + length arrays where None is Go's nil (a NULL pointer) and an
    empty array is a non-nil empty slice."""

    def __init__(self, shards):
        n = len(shards)
        self.n = n
        self.ptrs = (ctypes.c_void_p * max(n, 1))()
        self.lens = (ctypes.c_size_t * max(n, 1))()
        self.device = None
        for i, s in enumerate(shards):
            if s is None:
                continue
            p, ln, dev = _shard_info(s)
            self.ptrs[i] = p or ctypes.addressof(_EMPTY)
            self.lens[i] = ln
            if dev is not None and self.device is None:
                self.device = dev


class Encoder:
    """reedsolomon.Encoder + Extensions over the HIP library."""

    def __init__(self, handle, k, m):
        self._h = handle
        self._k, self._m = k, m

    def __del__(self, _free=lib.rsx_free):
        h = getattr(self, "_h", None)
        if h:
            _free(h)
            self._h = None

    # Extensions (reedsolomon.go:137-165)
    def ShardSizeMultiple(self) -> int:
        return lib.rsx_shard_size_multiple(self._h)

    def DataShards(self) -> int:
        return self._k

    def ParityShards(self) -> int:
        return self._m

    def TotalShards(self) -> int:
        return self._k + self._m

    def Engine(self) -> int:
        return lib.rsx_engine(self._h)

    def CodingMatrix(self) -> np.ndarray:
        out = np.zeros((self._k + self._m, self._k), np.uint8)
        _check(lib.rsx_coding_matrix(self._h, out.ctypes.data_as(_U8)))
        return out

    def AllocAligned(self, each: int):
        return [np.zeros(each, np.uint8) for _ in range(self._k + self._m)]

    def Synchronize(self) -> None:
        _check(lib.rsx_synchronize(self._h))

    def _before(self, device) -> None:
        if device:
            _after_torch(lib.rsx_get_stream(self._h))

    def DecodeCacheSize(self) -> int:
        return lib.rsx_decode_cache_size(self._h)

    @property
    def handle(self):
        return self._h

    # Encoder (reedsolomon.go:25-133)
    def Encode(self, shards) -> None:
        mm = _Marshal(shards)
        self._before(mm.device)
        _check(lib.rsx_encode(self._h, mm.ptrs, mm.lens, mm.n))
        if mm.device:
            self.Synchronize()

    def EncodeIdx(self, data_shard, idx: int, parity) -> None:
        mm = _Marshal(parity)
        p, ln, dev = _shard_info(data_shard)
        self._before(mm.device or dev)
        _check(lib.rsx_encode_idx(self._h, p or None, ln, idx, mm.ptrs, mm.lens, mm.n))
        if mm.device or dev:
            self.Synchronize()

    def Verify(self, shards) -> bool:
        mm = _Marshal(shards)
        self._before(mm.device)
        ok = ctypes.c_int(0)
        _check(lib.rsx_verify(self._h, mm.ptrs, mm.lens, mm.n, ctypes.byref(ok)))
        return bool(ok.value)

    def _prepare_outputs(self, shards):
        """Gives every absent shard a buffer of the shard size (the reference's
        reuse-capacity-else-allocate rule applied before the call)."""
        size, dev, like = 0, None, None
        for s in shards:
            p, ln, d = _shard_info(s)
            if ln:
                size, dev, like = ln, d, s
                break
        bufs = list(shards)
        if size == 0:
            return bufs, size, dev
        for i, s in enumerate(shards):
            if _shard_info(s)[1] == 0:
                if dev:
                    import torch
                    bufs[i] = torch.empty(size, dtype=torch.uint8, device=like.device)
                else:
                    bufs[i] = np.empty(size, np.uint8)
        return bufs, size, dev

    def _reconstruct(self, fn, shards, *extra):
        n = len(shards)
        bufs, size, dev = self._prepare_outputs(shards)
        mm = _Marshal(bufs)
        for i, s in enumerate(shards):  # absent shards keep length 0
            if _shard_info(s)[1] == 0:
                mm.lens[i] = 0
        filled = (ctypes.c_uint8 * max(n, 1))()
        self._before(dev)
        _check(fn(self._h, mm.ptrs, mm.lens, mm.n, *extra, filled))
        if dev:
            self.Synchronize()
        for i in range(n):
            if filled[i]:
                shards[i] = bufs[i]

    # Split / Join (reedsolomon.go:1803-1910, leopard.go:276-314, leopard8.go:293-350)
    # are data-format helpers with no device work; see split_shards / join_shards.
    def Split(self, data):
        """Shards of `data` (numpy array or torch tensor): full shards are
        views of `data`, the rest new zero-padded buffers."""
        return split_shards(data, self._k, self.TotalShards(), self.ShardSizeMultiple())

    def Join(self, dst, shards, outSize: int) -> None:
        """Writes the first outSize bytes of the data shards to dst (a file-like
        object with write())."""
        join_shards(dst, shards, self._k, outSize)

    def Update(self, shards, newDatashards) -> None:
        """reedsolomon.go:762.  None = nil.  Parity is updated in place and each
        changed old data shard becomes old ^ new, as in the reference."""
        a, b = _NilMarshal(shards), _NilMarshal(newDatashards)
        self._before(a.device or b.device)
        _check(lib.rsx_update(self._h, a.ptrs, a.lens, a.n, b.ptrs, b.lens, b.n))
        if a.device or b.device:
            self.Synchronize()

    def DecodeIdx(self, dst, expectInput, input) -> None:
        """reedsolomon.go:1464.  None = nil; expectInput None = merge mode."""
        d, i = _NilMarshal(dst), _NilMarshal(input)
        self._before(d.device or i.device)
        if expectInput is None:
            _check(lib.rsx_decode_idx(self._h, d.ptrs, d.lens, d.n, None, 0, i.ptrs, i.lens, i.n))
        else:
            ex = (ctypes.c_uint8 * max(len(expectInput), 1))(*[1 if e else 0 for e in expectInput])
            _check(lib.rsx_decode_idx(self._h, d.ptrs, d.lens, d.n, ex, len(expectInput), i.ptrs, i.lens, i.n))
        if d.device or i.device:
            self.Synchronize()

    def Reconstruct(self, shards) -> None:
        self._reconstruct(lib.rsx_reconstruct, shards)

    def ReconstructData(self, shards) -> None:
        self._reconstruct(lib.rsx_reconstruct_data, shards)

    def ReconstructNoAlloc(self, shards, data_only: bool = False) -> None:
        """Reconstruct (or ReconstructData) of device shards through
        rsx_reconstruct_noalloc: the library allocates nothing during the call
        (decode tables go to a fixed per-GPU scratch array).  No reference
        counterpart.  Absent shards get output buffers here, as for
        Reconstruct."""
        self._reconstruct(lib.rsx_reconstruct_noalloc, shards, 1 if data_only else 0)

    def ReconstructSome(self, shards, required) -> None:
        req = (ctypes.c_uint8 * max(len(required), 1))(*[1 if r else 0 for r in required])
        self._reconstruct(lib.rsx_reconstruct_some, shards, req, len(required))


def New(data_shards: int, parity_shards: int, *opts) -> Encoder:
    """reedsolomon.New (reedsolomon.go:425)."""
    o = _Options()
    lib.rsx_options_default(ctypes.byref(o))
    for f in opts:
        f(o)
    h = ctypes.c_void_p()
    _check(lib.rsx_new(data_shards, parity_shards, ctypes.byref(o), ctypes.byref(h)))
    return Encoder(h.value, data_shards, parity_shards)


# --- batched path (include/rsx.h rsx_batch_*) ---
OP_ENCODE, OP_RECONSTRUCT, OP_RECONSTRUCT_DATA, OP_VERIFY = 1, 2, 3, 4


class _BatchItem(ctypes.Structure):
    _fields_ = [("codec", ctypes.c_void_p), ("op", ctypes.c_int), ("nshards", ctypes.c_int),
                ("shards", ctypes.c_void_p), ("lens", ctypes.c_void_p)]


lib.rsx_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(_P)]
lib.rsx_batch_free.argtypes = [_P]
lib.rsx_batch_plan.argtypes = [_P, ctypes.POINTER(_BatchItem), ctypes.c_int,
                               ctypes.POINTER(ctypes.c_int), _U8]
lib.rsx_batch_launch.argtypes = [_P]
lib.rsx_batch_launch_on.argtypes = [_P, _P]
lib.rsx_batch_synchronize.argtypes = [_P]
lib.rsx_batch_stream.argtypes = [_P]
lib.rsx_batch_stream.restype = _P
lib.rsx_batch_num_tasks.argtypes = [_P]
lib.rsx_batch_set_tuning.argtypes = [_P, ctypes.c_int, ctypes.c_int]
lib.rsx_batch_get_tuning.argtypes = [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
lib.rsx_batch_set_graph.argtypes = [_P, ctypes.c_int]
lib.rsx_batch_verify_result.argtypes = [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]


def _marshal_item(enc, op, shards, item):
    """Fills one rsx_batch_item / rsx_request.  Absent shards of reconstruct
    items get output tensors (returned in `bufs`); the library still sees
    them as absent (length 0)."""
    import torch
    bufs = list(shards)
    size, like = 0, None
    for s in shards:
        if s is not None and s.numel():
            size, like = s.numel(), s
            break
    if op in (OP_RECONSTRUCT, OP_RECONSTRUCT_DATA) and size:
        for i, s in enumerate(shards):
            if s is None or s.numel() == 0:
                bufs[i] = torch.empty(size, dtype=torch.uint8, device=like.device)
    mm = _Marshal(bufs)
    for i, s in enumerate(shards):
        if s is None or s.numel() == 0:
            mm.lens[i] = 0
    item.codec = enc.handle
    item.op = op
    item.nshards = len(shards)
    item.shards = ctypes.cast(mm.ptrs, ctypes.c_void_p)
    item.lens = ctypes.cast(mm.lens, ctypes.c_void_p)
    return mm, bufs


class Batch:
    """Many Encode / Reconstruct / ReconstructData calls as one planned,
    graph-captured launch.  Items are (encoder, op, shards) with shards on
    the GPU; absent shards of reconstruct items get output tensors at plan
    time, holding their results after launch()."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib.rsx_batch_create(device, ctypes.byref(h)))
        self._h = h.value
        self._keep = []

    def __del__(self, _free=lib.rsx_batch_free):
        h = getattr(self, "_h", None)
        if h:
            _free(h)
            self._h = None

    def set_tuning(self, tile_bytes: int = 0, blocks_per_cu: int = 0) -> None:
        _check(lib.rsx_batch_set_tuning(self._h, tile_bytes, blocks_per_cu))

    def tuning(self):
        t, b = ctypes.c_int(), ctypes.c_int()
        lib.rsx_batch_get_tuning(self._h, ctypes.byref(t), ctypes.byref(b))
        return t.value, b.value

    def num_tasks(self) -> int:
        return lib.rsx_batch_num_tasks(self._h)

    @property
    def stream(self):
        return lib.rsx_batch_stream(self._h)

    def set_graph(self, enabled: bool) -> None:
        """Capture the plan in a HIP graph (default) or enqueue it directly."""
        _check(lib.rsx_batch_set_graph(self._h, 1 if enabled else 0))

    def plan(self, items):
        """Returns the per-item status codes (0 = ok).  An autotuning plan runs
        the work a few times, so torch work queued on the items' shards is
        waited for first."""
        _torch_sync()
        n = len(items)
        arr = (_BatchItem * max(n, 1))()
        keep, fills = [], []
        total_flags = 0
        for idx, (enc, op, shards) in enumerate(items):
            mm, bufs = _marshal_item(enc, op, shards, arr[idx])
            keep.append(mm)
            fills.append((total_flags, shards, bufs))
            total_flags += len(shards)
        status = (ctypes.c_int * max(n, 1))()
        filled = (ctypes.c_uint8 * max(total_flags, 1))()
        _check(lib.rsx_batch_plan(self._h, arr, n, status, filled))
        for off, shards, bufs in fills:
            for i in range(len(shards)):
                if filled[off + i]:
                    shards[i] = bufs[i]
        self._keep = [keep, arr, items]
        return [status[i] for i in range(n)]

    def verify_result(self, item: int) -> bool:
        """Outcome of Verify item `item` after the launch has completed."""
        ok = ctypes.c_int()
        _check(lib.rsx_batch_verify_result(self._h, item, ctypes.byref(ok)))
        return bool(ok.value)

    def launch(self, stream=None) -> None:
        """Replays the plan on the batch's stream (ordered after this thread's
        current torch stream) or on `stream` (the caller orders it)."""
        if stream is None:
            _after_torch(self.stream)
            _check(lib.rsx_batch_launch(self._h))
        else:
            _check(lib.rsx_batch_launch_on(self._h, stream))

    def synchronize(self) -> None:
        _check(lib.rsx_batch_synchronize(self._h))

    def run(self, items):
        st = self.plan(items)
        self.launch()
        self.synchronize()
        return st



# --- request loop (include/rsx.h rsx_server_*) ---
_Request = _BatchItem  # same layout: {codec, op, nshards, shards, lens}
_U64 = ctypes.POINTER(ctypes.c_uint64)
_I = ctypes.POINTER(ctypes.c_int)
lib.rsx_server_create.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(_P)]
lib.rsx_server_free.argtypes = [_P]
lib.rsx_server_free.restype = None
lib.rsx_server_submit.argtypes = [_P, ctypes.POINTER(_Request), _U64]
lib.rsx_server_poll.argtypes = [_P, ctypes.c_uint64, _I, _I, _I, _U64]
lib.rsx_server_wait.argtypes = [_P, ctypes.c_uint64, _I, _I, _U64]
lib.rsx_server_drain.argtypes = [_P]
lib.rsx_server_run_trace.argtypes = [_P, ctypes.POINTER(_Request), _U64, ctypes.c_int, _I, _I, _U64,
                                     _U64]
lib.rsx_server_stats.argtypes = [_P, _U64, _U64, _U64, _U64, _U64]
lib.rsx_server_cancel.argtypes = [_P, ctypes.c_uint64, _I]
lib.rsx_server_pause.argtypes = [_P, ctypes.c_int]
lib.rsx_server_set_queue_limit.argtypes = [_P, ctypes.c_size_t]
lib.rsx_server_admission.argtypes = [_P, _U64, _U64, _U64]


def _fill_outputs(op, k, shards, bufs):
    for i, s in enumerate(shards):
        if (s is None or s.numel() == 0) and bufs[i] is not None and (
                op == OP_RECONSTRUCT or (op == OP_RECONSTRUCT_DATA and i < k)):
            shards[i] = bufs[i]


class Server:
    """A request loop on one GPU: requests (encoder, op, shards) with shards
    in device memory are admitted in arrival order while the shard bytes in
    flight fit `hbm_budget`, and batched into launches as they arrive
    (continuous batching).  Absent shards of reconstruct requests are
    replaced by the rebuilt tensors when the request completes."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, max_batch: int = 0):
        h = ctypes.c_void_p()
        _check(lib.rsx_server_create(device, hbm_budget, max_batch, ctypes.byref(h)))
        self._h = h.value
        self._pending = {}

    def close(self, _free=lib.rsx_server_free):
        h = getattr(self, "_h", None)
        if h:
            _free(h)
            self._h = None
            self._pending.clear()

    def __del__(self):
        self.close()

    def submit(self, enc, op, shards) -> int:
        _torch_sync()  # the request loop's streams read the tensors later
        req = _Request()
        mm, bufs = _marshal_item(enc, op, shards, req)
        t = ctypes.c_uint64()
        _check(lib.rsx_server_submit(self._h, ctypes.byref(req), ctypes.byref(t)))
        self._pending[t.value] = (enc, op, shards, bufs, mm)
        return t.value

    def wait(self, ticket: int):
        """Returns (status, verify_ok, latency_ns)."""
        st, ok, lat = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64()
        _check(lib.rsx_server_wait(self._h, ticket, ctypes.byref(st), ctypes.byref(ok), ctypes.byref(lat)))
        enc, op, shards, bufs, _ = self._pending.pop(ticket)
        if st.value == 0:
            _fill_outputs(op, enc.DataShards(), shards, bufs)
        return st.value, bool(ok.value), lat.value

    def run_trace(self, requests, arrival_ns):
        """Submits requests[i] = (encoder, op, shards) at arrival_ns[i] after
        the start and waits for all.  Returns (status, verify_ok, latency_ns,
        makespan_ns); latency runs from the scheduled arrival."""
        n = len(requests)
        _torch_sync()
        arr = (_Request * max(n, 1))()
        keep = []
        for i, (enc, op, shards) in enumerate(requests):
            keep.append(_marshal_item(enc, op, shards, arr[i]))
        arrivals = (ctypes.c_uint64 * max(n, 1))(*[int(a) for a in arrival_ns])
        st = (ctypes.c_int * max(n, 1))()
        ok = (ctypes.c_int * max(n, 1))()
        lat = (ctypes.c_uint64 * max(n, 1))()
        span = ctypes.c_uint64()
        _check(lib.rsx_server_run_trace(self._h, arr, arrivals, n, st, ok, lat, ctypes.byref(span)))
        for i, (enc, op, shards) in enumerate(requests):
            if st[i] == 0:
                _fill_outputs(op, enc.DataShards(), shards, keep[i][1])
        return ([st[i] for i in range(n)], [bool(ok[i]) for i in range(n)],
                [lat[i] for i in range(n)], span.value)

    def cancel(self, ticket: int) -> bool:
        """Cancels a request not yet admitted into a batch (True); its wait()
        then reports ErrCancelled's code.  False: it runs to completion."""
        c = ctypes.c_int()
        _check(lib.rsx_server_cancel(self._h, ticket, ctypes.byref(c)))
        return bool(c.value)

    def pause(self, paused: bool = True) -> None:
        """Holds (True) or resumes (False) admission of queued requests."""
        _check(lib.rsx_server_pause(self._h, int(bool(paused))))

    def set_queue_limit(self, max_queued_bytes: int) -> None:
        """submit() raises ErrOverBudget past this many queued shard bytes
        (0: no limit)."""
        _check(lib.rsx_server_set_queue_limit(self._h, max_queued_bytes))

    def admission(self):
        """(requests rejected at submit, requests cancelled, bytes queued)."""
        v = [ctypes.c_uint64() for _ in range(3)]
        _check(lib.rsx_server_admission(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def drain(self) -> None:
        _check(lib.rsx_server_drain(self._h))

    def stats(self):
        """(batched launches, requests in them, max shard bytes in flight)."""
        return self.timing()[:3]

    def timing(self):
        """(batched launches, requests in them, max shard bytes in flight,
        host ns planning and launching batches, device ns of the batches)."""
        v = [ctypes.c_uint64() for _ in range(5)]
        _check(lib.rsx_server_stats(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)



# --- LowLevel (mulslice.go, gf16.go) ---
def Inv(e: int) -> int:
    """mulslice.go:42: multiplicative inverse in GF(2^8) (Inv(0) = 0)."""
    return lib.rsx_inv(e)


def leopard_tables(bits: int) -> dict:
    """The Leopard field tables the engine computes in (leopard.go:1104-1195
    for bits = 16, leopard8.go:1029-1117 for bits = 8): exp, log, skew
    (fftSkew) and walsh (logWalsh) as uint16 arrays.  The Go tables are
    package-private; this exposes them for checks."""
    order = 1 << bits
    t = dict(exp=np.zeros(order, np.uint16), log=np.zeros(order, np.uint16),
             skew=np.zeros(order - 1, np.uint16), walsh=np.zeros(order, np.uint16))
    ptr = {k: v.ctypes.data_as(_U16) for k, v in t.items()}
    _check(lib.rsx_leopard_tables(bits, ptr["exp"], ptr["log"], ptr["skew"], ptr["walsh"]))
    return t


class LowLevel:
    """mulslice.go LowLevel on the GPU (`device`); slices are numpy arrays or
    torch.uint8 CUDA tensors.  Length errors the reference panics on raise."""

    def __init__(self, device: int = 0):
        self.device = device

    def WithOptions(self, *opts) -> None:
        pass

    def GalMulSlice(self, c: int, inp, out) -> None:
        _torch_sync()
        p, n, _ = _shard_info(inp)
        q, nq, _ = _shard_info(out)
        if nq < n:
            raise ErrInvalidShardSize("out must be at least as long as in")
        _check(lib.rsx_gal_mul_slice(self.device, c, p or None, q or None, n, 0))

    def GalMulSliceXor(self, c: int, inp, out) -> None:
        _torch_sync()
        p, n, _ = _shard_info(inp)
        q, nq, _ = _shard_info(out)
        if nq < n:
            raise ErrInvalidShardSize("out must be at least as long as in")
        _check(lib.rsx_gal_mul_slice(self.device, c, p or None, q or None, n, 1))

    def GF16Mul(self, a: int, b: int) -> int:
        return lib.rsx_gf16_mul(a, b)

    def _gf16(self, scalars, inp, outs):
        _torch_sync()
        p, n, _ = _shard_info(inp)
        mm = _Marshal(outs)
        sc = (ctypes.c_uint16 * max(len(scalars), 1))(*scalars)
        _check(lib.rsx_gf16_mul_slice_xor(self.device, sc, len(scalars), p or None, mm.ptrs, mm.lens, n))

    def GF16MulSliceXor(self, scalar: int, inp, out) -> None:
        self._gf16([scalar], inp, [out])

    def GF16MulSliceXor8(self, scalars, inp, outs) -> None:
        assert len(scalars) == 8 and len(outs) == 8
        self._gf16(list(scalars), inp, list(outs))



# --- StreamEncoder (streaming.go) ---
class StreamReadError(RSError):
    """streaming.go:98: a read from stream `stream` failed (`err`)."""
    code = -15

    def __init__(self, stream: int, err=None):
        super().__init__(f"error reading stream {stream}: {err}")
        self.stream, self.err = stream, err


class StreamWriteError(RSError):
    """streaming.go:116: a write to stream `stream` failed (`err`)."""
    code = -16

    def __init__(self, stream: int, err=None):
        super().__init__(f"error writing stream {stream}: {err}")
        self.stream, self.err = stream, err


class _StreamIO:
    """ctypes callbacks over Python file objects (read(n) / write(bytes))."""

    def __init__(self, readers=None, writers=None):
        self.readers, self.writers = readers or [], writers or []
        self.error = None
        self.rd = _READ_FN(self._read)
        self.wr = _WRITE_FN(self._write)

    def _read(self, ctx, i, buf, n):
        try:
            view = memoryview((ctypes.c_char * n).from_address(buf)).cast("B")
            r = self.readers[i]
            got = 0
            readinto = getattr(r, "readinto", None)
            while got < n:  # io.ReadFull, straight into the pinned block
                if readinto is not None:
                    k = readinto(view[got:])
                    if not k:
                        break
                else:
                    chunk = r.read(n - got)
                    if not chunk:
                        break
                    k = len(chunk)
                    view[got:got + k] = chunk
                got += k
            return got
        except Exception as e:  # reported as StreamReadError
            self.error = e
            return -1

    def _write(self, ctx, i, buf, n):
        try:
            w = self.writers[i].write(memoryview((ctypes.c_char * n).from_address(buf)).cast("B"))
            return n if w is None else int(w)
        except Exception as e:
            self.error = e
            return -1


def _stream_check(rc, io_, err_stream):
    if rc == -15:
        raise StreamReadError(err_stream.value, io_.error)
    if rc == -16:
        raise StreamWriteError(err_stream.value, io_.error or "short write")
    _check(rc)


class StreamEncoder:
    """streaming.go StreamEncoder: shards are file-like objects (read(n) /
    write(b)); data moves block by block through pinned host buffers and the
    GPU, reads of the next block overlapping the GPU work."""

    def __init__(self, handle, k, m):
        self._h, self._k, self._m = handle, k, m

    def __del__(self, _free=lib.rsx_stream_free):
        h = getattr(self, "_h", None)
        if h:
            _free(h)
            self._h = None

    def Encode(self, data, parity) -> None:
        if len(data) != self._k or len(parity) != self._m:
            raise ErrTooFewShards()
        io_ = _StreamIO(data, parity)
        es = ctypes.c_int(-1)
        _stream_check(lib.rsx_stream_encode(self._h, io_.rd, None, io_.wr, None, ctypes.byref(es)), io_, es)

    def Verify(self, shards) -> bool:
        if len(shards) != self._k + self._m:
            raise ErrTooFewShards()
        io_ = _StreamIO(shards)
        ok, es = ctypes.c_int(0), ctypes.c_int(-1)
        _stream_check(lib.rsx_stream_verify(self._h, io_.rd, None, ctypes.byref(ok), ctypes.byref(es)), io_, es)
        return bool(ok.value)

    def Reconstruct(self, valid, fill) -> None:
        n = self._k + self._m
        if len(valid) != n or len(fill) != n:
            raise ErrTooFewShards()
        io_ = _StreamIO(valid, fill)
        v = (ctypes.c_uint8 * n)(*[0 if x is None else 1 for x in valid])
        f = (ctypes.c_uint8 * n)(*[0 if x is None else 1 for x in fill])
        es = ctypes.c_int(-1)
        _stream_check(lib.rsx_stream_reconstruct(self._h, v, f, io_.rd, None, io_.wr, None, ctypes.byref(es)),
                      io_, es)

    def Split(self, data, dst, size: int) -> None:
        """streaming.go:556-594 (host-only data movement): the data stream,
        then zero padding to TotalShards * perShard, copied perShard bytes per
        data shard writer."""
        if size == 0:
            raise ErrShortData()
        if len(dst) != self._k:
            raise ErrInvShardNum()
        for i, d in enumerate(dst):
            if d is None:
                raise StreamWriteError(i, ErrShardNoData())
        per = (size + self._k - 1) // self._k
        pad = (self._k + self._m) * per - size
        eof = False
        for d in dst:
            need = per
            while need:
                chunk = b"" if eof else data.read(need)
                if chunk:
                    d.write(chunk)
                    need -= len(chunk)
                    continue
                eof = True
                if pad == 0:
                    raise ErrShortData()
                take = min(need, pad)
                d.write(bytes(take))
                pad -= take
                need -= take

    def Join(self, dst, shards, outSize: int) -> None:
        """streaming.go:529-554 (host-only data movement)."""
        if len(shards) < self._k:
            raise ErrTooFewShards()
        shards = shards[:self._k]
        for i, sh in enumerate(shards):
            if sh is None:
                raise StreamReadError(i, ErrShardNoData())
        left = outSize
        for sh in shards:
            while left > 0:
                chunk = sh.read(min(left, 1 << 20))
                if not chunk:
                    break
                dst.write(chunk)
                left -= len(chunk)
        if left:
            raise ErrShortData()


def NewStream(data_shards: int, parity_shards: int, *opts, block_size: int = 0) -> StreamEncoder:
    """streaming.go:147 NewStream (classic engine).  The block size is
    block_size, else WithStreamBlockSize, else a positive WithAutoGoroutines
    shard size (streaming.go:157-163), else 4 MiB."""
    o = _Options()
    lib.rsx_options_default(ctypes.byref(o))
    for f in opts:
        f(o)
    bs = block_size or getattr(o, "_stream_bs", 0)
    if bs <= 0 and getattr(o, "_shard_size", 0) > 0:
        bs = o._shard_size
    h = ctypes.c_void_p()
    _check(lib.rsx_stream_create(data_shards, parity_shards, ctypes.byref(o), max(bs, 0), ctypes.byref(h)))
    st = StreamEncoder(h.value, data_shards, parity_shards)
    reads, writes = getattr(o, "_conc_reads", False), getattr(o, "_conc_writes", False)
    if reads or writes:
        _check(lib.rsx_stream_set_concurrency(st._h, int(reads), int(writes)))
    return st


def NewStreamC(data_shards: int, parity_shards: int, conc_reads: bool, conc_writes: bool, *opts,
               block_size: int = 0) -> StreamEncoder:
    """streaming.go:190 NewStreamC: NewStream with concurrent reads / writes."""
    return NewStream(data_shards, parity_shards, *opts, WithConcurrentStreamReads(conc_reads),
                     WithConcurrentStreamWrites(conc_writes), block_size=block_size)


__all__ = [n for n in dir() if n.startswith(("With", "Err", "New", "Encoder", "RSError"))] + [
    "device_count", "lib", "LIB_PATH", "Batch", "Server", "LowLevel", "Inv", "NewStream", "NewStreamC",
    "StreamEncoder",
    "StreamReadError", "StreamWriteError", "OP_ENCODE", "OP_RECONSTRUCT",
    "OP_RECONSTRUCT_DATA", "OP_VERIFY"]
