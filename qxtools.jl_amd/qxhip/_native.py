"""ctypes binding of libqxhip.so (C ABI declared in include/qxhip.h).

The library is loaded from this package directory; a missing library is an
error, never a fallback."""

from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

# QXHIP_LIB names another build of the same ABI (the sanitizer build of the host code)
LIB_PATH = os.environ.get("QXHIP_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "libqxhip.so")

PRECISIONS = {"c128": 0, "c64_4m": 1, "c64_3m": 2, "c64_bf16x3": 3}

QX_ERR_UNSUPPORTED = -8   # valid program outside what the library executes (include/qxhip.h)

SYMBOLS = [
    "qx_ctx_create", "qx_ctx_destroy", "qx_plan_from_dsl", "qx_plan_set_tensor",
    "qx_plan_set_option", "qx_plan_finalize", "qx_plan_info", "qx_plan_key",
    "qx_amplitudes", "qx_amplitudes_vjp", "qx_plan_stats", "qx_plan_destroy", "qx_last_error",
    "qx_dsl_inspect", "qx_dsl_inspect_ex", "qx_amplitudes_uniform", "qx_sample_rejection",
    "qx_measure_mfma", "qx_norm_vjp", "qx_plan_update_tensor", "qx_amplitudes_noalloc",
]


class QxError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"qxhip error {status}: {message}")
        self.status = status


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run qxtools.jl_amd/build.py")
        L = C.CDLL(LIB_PATH)
        p, i32, i64, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_size_t
        dp = C.POINTER(C.c_double)
        L.qx_ctx_create.argtypes = [C.c_int, C.POINTER(p)]
        L.qx_ctx_destroy.argtypes = [p]
        L.qx_ctx_destroy.restype = None
        L.qx_plan_from_dsl.argtypes = [p, C.c_char_p, sz, C.c_int, C.POINTER(p)]
        L.qx_plan_set_tensor.argtypes = [p, C.c_char_p, dp, i64]
        L.qx_plan_update_tensor.argtypes = [p, C.c_char_p, dp, i64]
        L.qx_plan_set_option.argtypes = [p, C.c_char_p, i64]
        L.qx_plan_finalize.argtypes = [p]
        L.qx_plan_info.argtypes = [p] + [C.POINTER(i64)] * 5
        L.qx_plan_key.argtypes = [p, i64, C.c_char_p, sz, C.POINTER(i64)]
        L.qx_amplitudes.argtypes = [p, C.c_char_p, i64, i64, i64, dp]
        L.qx_amplitudes_noalloc.argtypes = [p, C.c_char_p, i64, i64, i64, dp]
        L.qx_amplitudes_vjp.argtypes = [p, C.c_char_p, i64, dp, dp]
        L.qx_plan_stats.argtypes = [p, C.c_char_p, sz]
        L.qx_plan_destroy.argtypes = [p]
        L.qx_plan_destroy.restype = None
        L.qx_dsl_inspect.argtypes = [C.c_char_p, sz, C.c_int, i64, C.c_uint64, C.c_char_p, sz]
        L.qx_dsl_inspect_ex.argtypes = [C.c_char_p, sz, C.c_int, i64, C.c_uint64, i64, C.c_char_p, sz]
        L.qx_amplitudes_uniform.argtypes = [p, i64, C.c_uint64, C.c_char_p, dp]
        L.qx_sample_rejection.argtypes = [p, i64, C.c_uint64, C.c_double, C.c_int, i64, C.c_char_p, dp, dp,
                                          C.POINTER(i64)]
        L.qx_measure_mfma.argtypes = [p, C.c_int, dp]
        L.qx_norm_vjp.argtypes = [p, C.c_char_p, i64, dp, dp]
        L.qx_last_error.argtypes = []
        L.qx_last_error.restype = C.c_char_p
        for f in SYMBOLS:
            if f not in ("qx_ctx_destroy", "qx_plan_destroy", "qx_last_error"):
                getattr(L, f).restype = i32
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status != 0:
        raise QxError(status, lib().qx_last_error().decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def inspect(dsl: str, precision: str = "c128", batch: int = 1, fused: int = 0, batch_last: bool = False) -> dict:
    """Lower a program on the host (no device): steps, kernels, arena schedule.
    `fused`: bit mask of slice variables contracted inline; `batch_last`: the plan
    option of that name."""
    b = dsl.encode()
    cap = 1 << 16
    while True:
        buf = C.create_string_buffer(cap)
        st = lib().qx_dsl_inspect_ex(b, len(b), PRECISIONS[precision], batch, fused, int(bool(batch_last)), buf, cap)
        if st == 0:
            return json.loads(buf.value.decode())
        if st != -1 or cap > (1 << 28):
            _check(st)
        cap *= 4


class Context:
    """One device + one HIP stream (qx_ctx)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().qx_ctx_create(device, C.byref(self._h)))

    def measure_mfma(self, precision: str = "c128") -> float:
        """Measured matrix-core TFLOP/s of the device: FP64 products for "c128",
        FP32 for the complex64 precisions (`qx_measure_mfma`)."""
        t = C.c_double()
        _check(lib().qx_measure_mfma(self._h, PRECISIONS[precision], C.byref(t)))
        return t.value

    def close(self):
        if self._h:
            lib().qx_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """A lowered, device-resident .qx program (qx_plan)."""

    def __init__(self, ctx: Context, dsl: str, data: dict, precision: str = "c128", **options):
        self.ctx = ctx
        self._h = C.c_void_p()
        b = dsl.encode()
        _check(lib().qx_plan_from_dsl(ctx._h, b, len(b), PRECISIONS[precision], C.byref(self._h)))
        for k, v in options.items():
            if v is not None:
                _check(lib().qx_plan_set_option(self._h, k.encode(), int(v)))
        for key in self.keys_needed():
            arr = np.ascontiguousarray(np.asarray(data[key]).reshape(-1, order="F"), dtype=np.complex128)
            _check(lib().qx_plan_set_tensor(self._h, key.encode(), _dptr(arr.view(np.float64)), arr.size))
        _check(lib().qx_plan_finalize(self._h))
        info = self.info()
        self.num_chars, self.result_elems, self.num_slices = info[:3]

    def update_tensor(self, key: str, value) -> None:
        """New values of data key `key` after finalize, in the same device buffer
        (`qx_plan_update_tensor`): later calls, captured graphs included, see them."""
        arr = np.ascontiguousarray(np.asarray(value).reshape(-1, order="F"), dtype=np.complex128)
        _check(lib().qx_plan_update_tensor(self._h, key.encode(), _dptr(arr.view(np.float64)), arr.size))

    def keys_needed(self) -> list:
        n = C.c_int64()
        _check(lib().qx_plan_info(self._h, None, None, None, C.byref(n), None))
        out = []
        buf = C.create_string_buffer(512)
        for i in range(n.value):
            _check(lib().qx_plan_key(self._h, i, buf, 512, None))
            out.append(buf.value.decode())
        return out

    def key_sizes(self) -> list:
        """[(data key, elements)] in the gradient layout; fixed once the program is
        parsed, so queried once."""
        cached = getattr(self, "_key_sizes", None)
        if cached is not None:
            return list(cached)
        out = []
        buf = C.create_string_buffer(512)
        ne = C.c_int64()
        for i in range(len(self.keys_needed())):
            _check(lib().qx_plan_key(self._h, i, buf, 512, C.byref(ne)))
            out.append((buf.value.decode(), ne.value))
        self._key_sizes = tuple(out)
        return out

    def info(self) -> tuple:
        v = [C.c_int64() for _ in range(5)]
        _check(lib().qx_plan_info(self._h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def _bits(self, bitstrings) -> bytes:
        nc = self.num_chars
        out = []
        for s in bitstrings:
            if len(s) < nc:
                raise ValueError(f"bitstring {s!r} shorter than {nc} characters")
            out.append(s[:nc])
        return "".join(out).encode()

    def amplitudes(self, bitstrings, slice_begin: int = 0, slice_end: int | None = None) -> np.ndarray:
        """Complex128 array (nb,) for closed networks, else (nb, result_elems)."""
        bitstrings = list(bitstrings)
        nb = len(bitstrings)
        se = self.num_slices if slice_end is None else slice_end
        out = np.zeros(nb * self.result_elems, dtype=np.complex128)
        _check(lib().qx_amplitudes(self._h, self._bits(bitstrings), nb, slice_begin, se,
                                   _dptr(out.view(np.float64))))
        return out if self.result_elems == 1 else out.reshape(nb, self.result_elems)

    def _split(self, raw: bytes, n: int) -> list:
        nc = self.num_chars
        return [raw[i * nc:(i + 1) * nc].decode() for i in range(n)]

    def amplitudes_uniform(self, num: int, seed: int) -> tuple:
        """(bitstrings, amplitudes) of `num` device-generated uniform bitstrings."""
        bits = C.create_string_buffer(max(1, num * self.num_chars))
        out = np.zeros(num * self.result_elems, dtype=np.complex128)
        _check(lib().qx_amplitudes_uniform(self._h, num, seed & (2 ** 64 - 1), bits, _dptr(out.view(np.float64))))
        amps = out if self.result_elems == 1 else out.reshape(num, self.result_elems)
        return self._split(bits.raw, num), amps

    def sample_rejection(self, num_samples: int, seed: int, M: float = 1e-4, fix_M: bool = False,
                         max_candidates: int | None = None) -> tuple:
        """(bitstrings, amplitudes, final M, candidates consumed) of device-side
        empirical-supremum rejection sampling; QxError if max_candidates run out."""
        if max_candidates is None:
            max_candidates = 1000 * max(1, num_samples) + 100000
        bits = C.create_string_buffer(max(1, num_samples * self.num_chars))
        amps = np.zeros(num_samples, dtype=np.complex128)
        m_out = C.c_double(0.0)
        used = C.c_int64(0)
        _check(lib().qx_sample_rejection(self._h, num_samples, seed & (2 ** 64 - 1), float(M), int(bool(fix_M)),
                                         max_candidates, bits, _dptr(amps.view(np.float64)), C.byref(m_out),
                                         C.byref(used)))
        return self._split(bits.raw, num_samples), amps, m_out.value, used.value

    def vjp(self, bitstrings, amp_cot: np.ndarray) -> dict:
        bitstrings = list(bitstrings)
        nb = len(bitstrings)
        cot = np.ascontiguousarray(np.asarray(amp_cot, dtype=np.complex128).reshape(-1))
        if cot.size != nb * self.result_elems:
            raise ValueError("amp_cot size differs from nb * result_elems")
        sizes = self.key_sizes()
        grad = np.zeros(sum(n for _, n in sizes), dtype=np.complex128)
        _check(lib().qx_amplitudes_vjp(self._h, self._bits(bitstrings), nb, _dptr(cot.view(np.float64)),
                                       _dptr(grad.view(np.float64))))
        out, o = {}, 0
        for k, n in sizes:
            out[k] = grad[o:o + n]
            o += n
        return out

    def key_offsets(self) -> dict:
        """{data key: (offset, elements)} in the flat gradient of the pullbacks."""
        cached = getattr(self, "_key_offsets", None)
        if cached is None:
            cached, o = {}, 0
            for k, n in self.key_sizes():
                cached[k] = (o, n)
                o += n
            self._key_offsets = cached
        return cached

    def norm_vjp(self, bitstrings, flat: bool = False) -> tuple:
        """(amplitudes, {key: gradient}) of the loss sum |amp|^2 in one pass
        (`qx_norm_vjp`); with `flat`, the gradient as one array in the layout of
        `key_offsets`.  QxError with status QX_ERR_UNSUPPORTED when slices loop."""
        bitstrings = list(bitstrings)
        nb = len(bitstrings)
        sizes = self.key_sizes()
        grad = np.zeros(sum(n for _, n in sizes), dtype=np.complex128)
        amps = np.zeros(nb * self.result_elems, dtype=np.complex128)
        _check(lib().qx_norm_vjp(self._h, self._bits(bitstrings), nb, _dptr(amps.view(np.float64)),
                                 _dptr(grad.view(np.float64))))
        if flat:
            return (amps if self.result_elems == 1 else amps.reshape(nb, self.result_elems)), grad
        out, o = {}, 0
        for k, n in sizes:
            out[k] = grad[o:o + n]
            o += n
        return (amps if self.result_elems == 1 else amps.reshape(nb, self.result_elems)), out

    def set_option(self, name: str, value: int) -> None:
        _check(lib().qx_plan_set_option(self._h, name.encode(), int(value)))

    def stats(self) -> dict:
        cap = 1 << 16
        while True:
            buf = C.create_string_buffer(cap)
            st = lib().qx_plan_stats(self._h, buf, cap)
            if st == 0:
                return json.loads(buf.value.decode())
            if cap > (1 << 28):
                _check(st)
            cap *= 4

    def close(self):
        if self._h:
            lib().qx_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
