"""The ``hip`` matching backend: gfx950 kernels behind pyTME's backend API.

Arrays are ``torch.Tensor`` objects on the ROCm device; torch provides device
memory, streams and the generic (setup-time) array operations.  The
operations of the per-rotation scoring loop run hand-written HIP kernels from
``libpytme_hip.so`` (see ``include/pytme_hip.h``):

==============================  ==============================================
backend method                  kernel(s)
==============================  ==============================================
``rigid_transform``             ``tm_rotate_trilinear`` (scipy order-1 rules)
``build_fft`` / ``rfftn`` ...   ``tm_fft_r2c`` / ``tm_fft_c2r`` (Stockham)
``multiply`` (spectra)          ``tm_cmul_cc`` / ``tm_cmul_cr``
``add/subtract/multiply/...``   ``tm_binary_f32`` (float32, same shape/scalar)
``sum`` (float32, all axes)     ``tm_sum_f64`` (deterministic FP64)
``normalize_template``          ``tm_normalize_template`` (FP64 moments)
``norm_scores``                 ``tm_norm_scores``
``max_score_over_rotations``    ``tm_max_update``
``fill``                        ``tm_fill_f32`` / ``tm_fill_i32``
``max_filter_coordinates``      ``tm_peaks_max_filter`` + ``tm_peaks_write``
``topk_indices``                ``tm_topk_f32``
``greedy_distance_filter``      ``tm_peak_filter``
==============================  ==============================================

The backend mirrors ``tme/backends/pytorch_backend.py`` in its interface
(constructor accepts ``device=``; ``to_sharedarr`` returns the device array
itself) and ``tme/backends/npfftw_backend.py`` in its numerical semantics.
There is no CPU path for the compute primitives: on a CPU device (allowed
for host-side logic and tests) or without the built library they raise.
"""

from contextlib import contextmanager, nullcontext
from typing import Tuple

import numpy as np

from .. import _lib
from .matching_backend import MatchingBackend

__all__ = ["HipBackend"]

# FFT plans are cached per process: the backend is re-instantiated by every
# scan task (reference ``_wrap_backend``), plans must survive that.
_FFT_PLANS = {}
_WORKSPACES = {}


def _next_fast_len(n: int) -> int:
    """Smallest 5-smooth integer >= n (matches scipy.fft.next_fast_len(real=True))."""
    n = max(int(n), 1)
    best = None
    p5 = 1
    while p5 < 2 * n:
        p35 = p5
        while p35 < 2 * n:
            p = p35
            while p < n:
                p *= 2
            if best is None or p < best:
                best = p
            p35 *= 3
        p5 *= 5
    return best


class _Pose:
    """Deferred rigid transform (rotation about a centre plus translation).

    Returned by :meth:`HipBackend._rigid_transform_matrix` in place of the
    homogeneous matrix of the numpy backend; the HIP rotation kernel builds
    the pull matrix ``T(-t) T(c) R^-1 T(-c)`` itself in double precision.
    """

    def __init__(self, rotation, translation=None, center=None):
        self.rotation = rotation
        self.translation = translation
        self.center = center


class HipBackend(MatchingBackend):
    """Matching backend running the scoring loop on gfx950 HIP kernels.

    Parameters
    ----------
    device : str, optional
        Torch device string, ``"cuda"`` (the current ROCm device) by default.
        ``"cpu"`` is accepted for host-side logic only.
    fused : bool, optional
        Let the registered scores use the batched fused FLC/CC scan when the
        analyzer is :class:`MaxScoreOverRotations` (default True).  False
        forces the plain per-rotation loop.
    rotation_batch : int or "auto", optional
        Rotations per fused launch.  By default the largest of 64, 32, 16 ...
        whose workspace fits a quarter of the free HBM; ``"auto"`` also
        times that size and its halves on the first call of a configuration
        and keeps the fastest (``hip_fused.autotune_batch``).
    scan_graph : bool, optional
        Replay each fused batch's kernel sequence as a hipGraph (default True).
    scan_tiles : tuple of int, optional
        ``(pass_a_columns, pass_c_rows)`` of the fused passes
        (``tm_scan_plan_set_tiles``); by default the kernels' own choice, or
        with ``rotation_batch="auto"`` the autotuner's.
    scan_radix : int, optional
        Radix plan variant (0 or 1, ``tm_scan_plan_set_radix``) of the fused
        passes' compile-time FFTs; by default 0, or with
        ``rotation_batch="auto"`` the autotuner's.
    """

    #: spline order the rotation kernel implements (reference CLI default for
    #: GPU backends, scripts/match_template.py:741-747)
    default_interpolation_order = 1

    def __init__(
        self,
        device="cuda",
        float_dtype=None,
        complex_dtype=None,
        int_dtype=None,
        overflow_safe_dtype=None,
        fused: bool = True,
        rotation_batch: int = None,
        scan_graph: bool = True,
        scan_tiles=None,
        scan_radix=None,
        **kwargs,
    ):
        import torch

        super().__init__(
            array_backend=torch,
            float_dtype=torch.float32 if float_dtype is None else float_dtype,
            complex_dtype=torch.complex64 if complex_dtype is None else complex_dtype,
            int_dtype=torch.int32 if int_dtype is None else int_dtype,
            overflow_safe_dtype=torch.float32 if overflow_safe_dtype is None else overflow_safe_dtype,
        )
        self._torch = torch
        self.device = torch.device(device)
        self._gpu = self.device.type == "cuda"
        self.fused = bool(fused)
        self.rotation_batch = rotation_batch
        self.scan_graph = bool(scan_graph)
        self.scan_tiles = None if scan_tiles is None else tuple(int(x) for x in scan_tiles)
        self.scan_radix = None if scan_radix is None else int(scan_radix)
        self._init_kwargs = dict(device=device, fused=fused, rotation_batch=rotation_batch, scan_graph=scan_graph,
                                 scan_tiles=scan_tiles, scan_radix=scan_radix, **kwargs)

    # ------------------------------------------------------------------ infra
    def _require_gpu(self, what: str):
        if not self._gpu:
            raise RuntimeError(
                f"hip backend: {what} runs only on a GPU device (device={self.device}); "
                "there is no CPU implementation"
            )
        _lib.lib()

    def _stream(self):
        return self._torch.cuda.current_stream().cuda_stream

    def _workspace(self, ndoubles: int):
        key = (self._torch.cuda.current_device(), "red")
        ws = _WORKSPACES.get(key)
        if ws is None or ws.numel() < ndoubles:
            ws = self._torch.empty(int(ndoubles), dtype=self._torch.float64, device=self.device)
            _WORKSPACES[key] = ws
        return ws

    def _dev_f32(self, x) -> bool:
        t = self._torch
        return isinstance(x, t.Tensor) and x.dtype == t.float32 and x.is_cuda and x.is_contiguous()

    def _dev_c64(self, x) -> bool:
        t = self._torch
        return isinstance(x, t.Tensor) and x.dtype == t.complex64 and x.is_cuda and x.is_contiguous()

    # ------------------------------------------------------------- conversion
    def to_backend_array(self, arr, check_device: bool = True):
        t = self._torch
        if isinstance(arr, t.Tensor):
            if arr.device == self.device or not check_device:
                return arr
            return arr.to(self.device)
        arr = np.asarray(arr)
        return t.from_numpy(np.ascontiguousarray(arr)).to(self.device)

    def to_numpy_array(self, arr):
        t = self._torch
        if isinstance(arr, np.ndarray):
            return arr
        if isinstance(arr, t.Tensor):
            return arr.detach().cpu().numpy()
        return np.asarray(arr)

    def to_cpu_array(self, arr):
        return arr.cpu()

    def get_fundamental_dtype(self, arr):
        t = self._torch
        if t.is_complex(arr):
            return complex
        if t.is_floating_point(arr):
            return float
        return int

    def free_cache(self):
        """Release the kept fused-scan plan and torch's cached blocks."""
        if self._gpu:
            from . import hip_fused

            hip_fused.release()
            self._torch.cuda.empty_cache()

    # ------------------------------------------------------------- arithmetic
    def _binary(self, op: int, torch_fn, x1, x2, out):
        t = self._torch
        if self._gpu and self._dev_f32(x1):
            b, scalar, ok = None, 0.0, False
            if isinstance(x2, (int, float, np.floating, np.integer)) and not isinstance(x2, (bool, np.bool_)):
                scalar, ok = float(x2), True
            elif self._dev_f32(x2) and (
                tuple(x2.shape) == tuple(x1.shape) or (x2.numel() == 1 and x2.ndim <= x1.ndim)
            ):
                b, ok = x2, True
            if ok:
                if out is None:
                    out = t.empty_like(x1)
                if self._dev_f32(out) and tuple(out.shape) == tuple(x1.shape):
                    _lib.call(
                        "tm_binary_f32", op, x1.data_ptr(), None if b is None else b.data_ptr(), scalar,
                        out.data_ptr(), x1.numel(), 1 if b is None else b.numel(), self._stream(),
                    )
                    return out
        return self._torch_binary(torch_fn, x1, x2, out)

    def _torch_binary(self, torch_fn, x1, x2, out):
        t = self._torch
        if not isinstance(x1, t.Tensor):
            ref = x2 if isinstance(x2, t.Tensor) else None
            x1 = t.as_tensor(x1, device=None if ref is None else ref.device)
        res = torch_fn(x1, x2)
        if out is None:
            return res
        out.copy_(res if tuple(res.shape) == tuple(out.shape) else res.expand_as(out))
        return out

    def add(self, x1, x2, out=None):
        return self._binary(0, self._torch.add, x1, x2, out)

    def subtract(self, x1, x2, out=None):
        return self._binary(1, self._torch.subtract, x1, x2, out)

    def multiply(self, x1, x2, out=None):
        t = self._torch
        if self._gpu and self._dev_c64(x1):
            if self._dev_c64(x2) and tuple(x2.shape) == tuple(x1.shape):
                out = t.empty_like(x1) if out is None else out
                if self._dev_c64(out) and out.shape == x1.shape:
                    _lib.call("tm_cmul_cc", x1.data_ptr(), x2.data_ptr(), out.data_ptr(), x1.numel(),
                              x1.numel(), self._stream())
                    return out
            if self._dev_f32(x2) and tuple(x2.shape) == tuple(x1.shape):
                out = t.empty_like(x1) if out is None else out
                if self._dev_c64(out) and out.shape == x1.shape:
                    _lib.call("tm_cmul_cr", x1.data_ptr(), x2.data_ptr(), out.data_ptr(), x1.numel(),
                              x1.numel(), self._stream())
                    return out
        if self._gpu and self._dev_f32(x1) and self._dev_c64(x2) and x1.shape == x2.shape:
            return self.multiply(x2, x1, out=out)
        return self._binary(2, t.multiply, x1, x2, out)

    def divide(self, x1, x2, out=None):
        t = self._torch
        if not isinstance(x1, t.Tensor):
            res = t.divide(t.as_tensor(x1, device=x2.device, dtype=x2.dtype), x2)
            if out is None:
                return res
            out.copy_(res)
            return out
        return self._binary(3, t.divide, x1, x2, out)

    def mod(self, x1, x2, *args, **kwargs):
        return self._torch.remainder(x1, x2, *args, **kwargs)

    def sum(self, arr, axis=None, keepdims: bool = False, **kwargs):
        t = self._torch
        full = axis is None or (
            isinstance(axis, (tuple, list)) and sorted(int(a) % max(arr.ndim, 1) for a in axis) == list(range(arr.ndim))
        )
        if self._gpu and self._dev_f32(arr) and full and arr.numel() > 0:
            ws = self._workspace(_lib.lib().tm_reduce_workspace())
            out = t.empty((), dtype=t.float64, device=arr.device)
            _lib.call("tm_sum_f64", arr.data_ptr(), arr.numel(), out.data_ptr(), ws.data_ptr(), self._stream())
            if keepdims:
                out = out.reshape((1,) * arr.ndim)
            return out
        if axis is None:
            res = t.sum(arr)
            return res.reshape((1,) * arr.ndim) if keepdims else res
        return t.sum(arr, dim=axis, keepdim=keepdims)

    def mean(self, arr, axis=None, keepdims=False):
        t = self._torch
        if axis is None:
            return t.mean(arr)
        return t.mean(arr, dim=axis, keepdim=keepdims)

    def std(self, arr, axis=None, keepdims=False):
        t = self._torch
        if axis is None:
            return t.std(arr, unbiased=False)
        return t.std(arr, dim=axis, keepdim=keepdims, unbiased=False)

    def max(self, arr, *args, **kwargs):
        return self._torch.amax(arr, *args, **kwargs)

    def min(self, arr, *args, **kwargs):
        return self._torch.amin(arr, *args, **kwargs)

    def maximum(self, x1, x2, out=None):
        t = self._torch
        if not isinstance(x2, t.Tensor):
            x2 = t.as_tensor(x2, dtype=x1.dtype, device=x1.device)
        res = t.maximum(x1, x2.to(x1.device))
        if out is None:
            return res
        out.copy_(res)
        return out

    def minimum(self, x1, x2, out=None):
        t = self._torch
        if not isinstance(x2, t.Tensor):
            x2 = t.as_tensor(x2, dtype=x1.dtype, device=x1.device)
        res = t.minimum(x1, x2.to(x1.device))
        if out is None:
            return res
        out.copy_(res)
        return out

    def sqrt(self, arr, out=None):
        return self._torch.sqrt(arr, out=out) if out is not None else self._torch.sqrt(arr)

    def square(self, arr, out=None):
        if out is not None:
            return self._torch.square(arr, out=out)
        return self._torch.square(arr)

    def abs(self, arr, out=None):
        return self._torch.abs(arr, out=out) if out is not None else self._torch.abs(arr)

    def power(self, *args, **kwargs):
        return self._torch.pow(*args, **kwargs)

    def transpose(self, arr, axes=None):
        if axes is None:
            axes = tuple(range(arr.ndim - 1, -1, -1))
        return arr.permute(axes)

    # ------------------------------------------------------ allocation / info
    def size(self, arr) -> int:
        t = self._torch
        return arr.numel() if isinstance(arr, t.Tensor) else int(np.size(arr))

    def fill(self, arr, value):
        t = self._torch
        if self._gpu and self._dev_f32(arr):
            _lib.call("tm_fill_f32", arr.data_ptr(), float(value), arr.numel(), self._stream())
            return arr
        if self._gpu and isinstance(arr, t.Tensor) and arr.dtype == t.int32 and arr.is_contiguous() and arr.is_cuda:
            _lib.call("tm_fill_i32", arr.data_ptr(), int(value), arr.numel(), self._stream())
            return arr
        arr.fill_(value)
        return arr

    def zeros(self, shape, dtype=None):
        if isinstance(shape, int):
            shape = (shape,)
        return self._torch.zeros(tuple(int(x) for x in shape), dtype=dtype, device=self.device)

    def empty(self, shape, dtype=None):
        if isinstance(shape, int):
            shape = (shape,)
        return self._torch.empty(tuple(int(x) for x in shape), dtype=dtype, device=self.device)

    def zeros_like(self, arr):
        return self._torch.zeros_like(arr)

    def full(self, shape, fill_value, dtype=None):
        if isinstance(shape, int):
            shape = (shape,)
        return self._torch.full(tuple(int(x) for x in shape), fill_value, dtype=dtype, device=self.device)

    def eps(self, dtype):
        return float(self._torch.finfo(dtype).eps)

    def datatype_bytes(self, dtype) -> int:
        return self._torch.empty((), dtype=dtype).element_size()

    def astype(self, arr, dtype):
        t = self._torch
        if isinstance(arr, t.Tensor):
            if t.is_complex(arr) and not dtype.is_complex:
                arr = arr.real
            return arr.to(dtype)
        return t.as_tensor(arr, device=self.device).to(dtype)

    def copy(self, arr):
        return self._torch.clone(arr)

    def arange(self, *args, **kwargs):
        return self._torch.arange(*args, device=self.device, **kwargs)

    @staticmethod
    def at(arr, idx, value):
        arr[idx] = value
        return arr

    def tobytes(self, arr) -> bytes:
        return self.to_numpy_array(arr).tobytes()

    # ------------------------------------------------------------------ layout
    def roll(self, a, shift, axis, **kwargs):
        shift = tuple(int(s) for s in np.atleast_1d(shift))
        axis = tuple(int(x) for x in np.atleast_1d(axis))
        return self._torch.roll(a, shifts=shift, dims=axis)

    def topleft_pad(self, arr, shape: Tuple[int], padval: float = 0):
        t = self._torch
        shape = tuple(int(x) for x in shape)
        b = t.full(shape, padval, dtype=arr.dtype, device=arr.device)
        box = tuple(slice(0, min(int(x), int(y))) for x, y in zip(arr.shape, shape))
        b[box] = arr[box]
        return b

    def reverse(self, arr, axis: Tuple[int] = None):
        if axis is None:
            axis = tuple(range(arr.ndim))
        axis = tuple(int(a) for a in axis)
        if not axis:
            return arr
        return self._torch.flip(arr, dims=axis)

    # ----------------------------------------------------------------- sharing
    def to_sharedarr(self, arr, shared_memory_handler=None):
        # device arrays cannot be placed in host shared memory; GPU scans run
        # one rotation job per device (reference cupy/pytorch backends).
        return (arr,)

    def from_sharedarr(self, args):
        if isinstance(args, tuple) and len(args) == 1:
            return args[0]
        return args

    # ------------------------------------------------------------------ device
    def get_available_memory(self) -> int:
        if self._gpu:
            return int(self._torch.cuda.mem_get_info()[0])
        import psutil

        return int(psutil.virtual_memory().available)

    @contextmanager
    def set_device(self, device_index: int):
        if self._gpu:
            with self._torch.cuda.device(int(device_index)):
                yield None
        else:
            with nullcontext():
                yield None

    def device_count(self) -> int:
        return int(self._torch.cuda.device_count()) if self._gpu else 1

    def synchronize(self):
        if self._gpu:
            _lib.call("tm_stream_sync", self._stream())

    # --------------------------------------------------------------------- FFT
    def compute_convolution_shapes(self, arr1_shape, arr2_shape):
        conv = [int(x + y - 1) for x, y in zip(arr1_shape, arr2_shape)]
        fast = [_next_fast_len(x) for x in conv]
        fast_ft = list(fast[:-1]) + [fast[-1] // 2 + 1]
        return conv, fast, fast_ft

    def _plan(self, real_shape: Tuple[int], batch: int = 1):
        import ctypes

        self._require_gpu("FFT")
        real_shape = tuple(int(x) for x in real_shape)
        key = (self._torch.cuda.current_device(), real_shape, int(batch))
        plan = _FFT_PLANS.get(key)
        if plan is None:
            handle = ctypes.c_void_p()
            _lib.call(
                "tm_fft_plan_create", len(real_shape), _lib.i64_array(real_shape), int(batch),
                ctypes.byref(handle),
            )
            plan = handle.value
            _FFT_PLANS[key] = plan
        return plan

    def _check_axes(self, axes, ndim):
        if axes is None:
            return
        if sorted(int(a) % ndim for a in axes) != list(range(ndim)):
            raise ValueError("hip backend FFTs transform all axes (batched axes are not supported)")

    def _r2c(self, plan, arr, out):
        if not (self._dev_f32(arr) and self._dev_c64(out)):
            raise ValueError("rfftn expects contiguous float32 input and complex64 output on the GPU")
        _lib.call("tm_fft_r2c", plan, arr.data_ptr(), out.data_ptr(), self._stream())
        return out

    def _c2r(self, plan, arr, out):
        if not (self._dev_c64(arr) and self._dev_f32(out)):
            raise ValueError("irfftn expects contiguous complex64 input and float32 output on the GPU")
        _lib.call("tm_fft_c2r", plan, arr.data_ptr(), out.data_ptr(), self._stream())
        return out

    def build_fft(
        self,
        fwd_shape,
        inv_shape,
        real_dtype=None,
        cmpl_dtype=None,
        inv_output_shape=None,
        temp_fwd=None,
        temp_inv=None,
        fwd_axes=None,
        inv_axes=None,
        **kwargs,
    ):
        """FFT pair on ``fwd_shape`` (numpy.fft.rfftn/irfftn semantics).

        ``irfftn(arr, out)`` uses ``arr`` as scratch (like pyFFTW with
        ``overwrite_input=True`` in the reference numpy backend).
        """
        real_shape = tuple(int(x) for x in (inv_output_shape if inv_output_shape is not None else fwd_shape))
        if tuple(int(x) for x in fwd_shape) != real_shape:
            raise ValueError(f"fwd_shape {tuple(fwd_shape)} != inverse output shape {real_shape}")
        self._check_axes(fwd_axes, len(real_shape))
        self._check_axes(inv_axes, len(real_shape))
        expected_ft = list(real_shape[:-1]) + [real_shape[-1] // 2 + 1]
        if [int(x) for x in inv_shape] != expected_ft:
            raise ValueError(f"inv_shape {tuple(inv_shape)} is not the half spectrum of {real_shape}")
        plan = self._plan(real_shape)

        def rfftn(arr, out, *args, **kw):
            return self._r2c(plan, arr, out)

        def irfftn(arr, out, *args, **kw):
            return self._c2r(plan, arr, out)

        return rfftn, irfftn

    def rfftn(self, arr, s=None, axes=None, **kwargs):
        t = self._torch
        if s is not None and tuple(int(x) for x in s) != tuple(arr.shape):
            raise ValueError("hip rfftn does not pad or crop (s must equal arr.shape)")
        self._check_axes(axes, arr.ndim)
        arr = arr.to(t.float32).contiguous()
        plan = self._plan(arr.shape)
        out = t.empty(tuple(arr.shape[:-1]) + (arr.shape[-1] // 2 + 1,), dtype=t.complex64, device=arr.device)
        return self._r2c(plan, arr, out)

    def irfftn(self, arr, s=None, axes=None, **kwargs):
        t = self._torch
        self._check_axes(axes, arr.ndim)
        if s is None:
            s = tuple(arr.shape[:-1]) + (2 * (arr.shape[-1] - 1),)
        s = tuple(int(x) for x in s)
        if tuple(arr.shape) != s[:-1] + (s[-1] // 2 + 1,):
            raise ValueError(f"spectrum shape {tuple(arr.shape)} does not match output shape {s}")
        work = arr.to(t.complex64).clone(memory_format=t.contiguous_format)
        plan = self._plan(s)
        out = t.empty(s, dtype=t.float32, device=arr.device)
        return self._c2r(plan, work, out)

    # ---------------------------------------------------------------- rotation
    def _rigid_transform_matrix(self, rotation_matrix, translation=None, center=None):
        return _Pose(rotation_matrix, translation=translation, center=center)

    def center_of_mass(self, arr, cutoff: float = None):
        t = self._torch
        a = arr.to(t.float64)
        if cutoff is not None:
            a = t.where(a > cutoff, a, t.zeros_like(a))
        total = a.sum()
        grids = t.meshgrid(*[t.arange(n, dtype=t.float64, device=a.device) for n in a.shape], indexing="ij")
        return t.stack([(a * g).sum() / total for g in grids])

    def _host_vec(self, v, ndim):
        if v is None:
            return None
        return [float(x) for x in np.asarray(self.to_numpy_array(v), dtype=np.float64).reshape(-1)[:ndim]]

    def rigid_transform(
        self,
        arr,
        rotation_matrix,
        arr_mask=None,
        translation=None,
        use_geometric_center: bool = False,
        out=None,
        out_mask=None,
        order: int = 1,
        cache: bool = False,
        batched: bool = False,
    ):
        """Rotate ``arr`` (and ``arr_mask``) into the top-left box of ``out``.

        Follows ``NumpyFFTWBackend.rigid_transform``
        (``tme/backends/npfftw_backend.py:390-449``) at spline order 1 with
        ``mode="constant"``: the output box ``out[:s0, :s1, :s2]`` is fully
        written, the rest of ``out`` is untouched.
        """
        t = self._torch
        self._require_gpu("rigid_transform")
        if int(order) != 1:
            raise ValueError(f"Got {order} but the hip backend implements interpolation order 1.")
        if batched:
            for i in range(arr.shape[0]):
                self.rigid_transform(
                    arr[i], rotation_matrix, None if arr_mask is None else arr_mask[i], translation,
                    use_geometric_center, out[i], None if out_mask is None else out_mask[i], order, cache,
                )
            return out, out_mask
        if out is None:
            out = t.zeros_like(arr)
        if arr_mask is not None and out_mask is None:
            out_mask = t.zeros_like(arr_mask)

        ndim = arr.ndim
        if ndim not in (2, 3):
            raise ValueError("rigid_transform supports 2D and 3D arrays")
        shape = tuple(int(x) for x in arr.shape)

        pose = rotation_matrix
        if not isinstance(pose, _Pose):
            rm = pose
            if tuple(rm.shape) == (ndim + 1, ndim + 1):
                pose = ("matrix", rm)
            else:
                center = None
                if use_geometric_center:
                    center = [(s - 1) / 2 for s in shape]
                else:
                    center = self._host_vec(self.center_of_mass(arr, cutoff=0), ndim)
                pose = _Pose(rm, translation=translation, center=center)

        src = arr.to(t.float32).contiguous()
        srcm = None if arr_mask is None else arr_mask.to(t.float32).contiguous()
        for o in (out, out_mask):
            if o is None:
                continue
            if o.dtype != t.float32 or o.stride(-1) != 1 or not o.is_cuda:
                raise ValueError("rigid_transform output must be float32 with a contiguous last axis")
            if any(int(a) < int(b) for a, b in zip(o.shape, shape)):
                raise ValueError("rigid_transform output is smaller than the input")
        if out_mask is not None and tuple(out_mask.stride()) != tuple(out.stride()):
            raise ValueError("out and out_mask must share strides")

        if ndim == 3:
            kshape = shape
            strides = (out.stride(0), out.stride(1))
        else:
            kshape = (1,) + shape
            strides = (0, out.stride(0))

        rots_ptr = mats_ptr = None
        center = translation3 = None
        keep = []
        if isinstance(pose, tuple):
            m = np.asarray(self.to_numpy_array(pose[1]), dtype=np.float64)
            m = m / m[ndim, ndim]
            m3 = np.zeros((3, 4), dtype=np.float64)
            if ndim == 3:
                m3[:, :3] = m[:3, :3]
                m3[:, 3] = m[:3, 3]
            else:
                m3[0, 0] = 1.0
                m3[1:, 1:3] = m[:2, :2]
                m3[1:, 3] = m[:2, 2]
            md = t.from_numpy(m3.reshape(-1)).to(self.device)
            keep.append(md)
            mats_ptr = md.data_ptr()
        else:
            rot = pose.rotation
            if ndim == 2:
                r2 = np.asarray(self.to_numpy_array(rot), dtype=np.float32)
                r3 = np.eye(3, dtype=np.float32)
                r3[1:, 1:] = r2
                rot = t.from_numpy(r3).to(self.device)
            rot = self.to_backend_array(rot).to(t.float32).contiguous()
            keep.append(rot)
            rots_ptr = rot.data_ptr()
            c = pose.center if pose.center is not None else [0.0] * ndim
            c = self._host_vec(c, ndim)
            tr = self._host_vec(pose.translation, ndim) if pose.translation is not None else [0.0] * ndim
            if ndim == 2:
                c, tr = [0.0] + c, [0.0] + tr
            center, translation3 = _lib.f64_array(c), _lib.f64_array(tr)

        _lib.call(
            "tm_rotate_trilinear",
            src.data_ptr(),
            None if srcm is None else srcm.data_ptr(),
            _lib.i32_array(kshape),
            rots_ptr,
            mats_ptr,
            center,
            translation3,
            1,
            out.data_ptr(),
            None if out_mask is None else out_mask.data_ptr(),
            _lib.i64_array(strides),
            0,
            None,
            None,
            self._stream(),
        )
        return out, out_mask

    # ----------------------------------------------------------------- scoring
    def _n_obs_args(self, n_obs):
        t = self._torch
        if isinstance(n_obs, t.Tensor):
            if n_obs.numel() != 1:
                raise ValueError("n_obs must be a scalar")
            if n_obs.is_cuda:
                nd = n_obs.reshape(()).to(t.float64)
                return nd, nd.data_ptr(), 0.0
            return None, None, float(n_obs.item())
        return None, None, float(n_obs)

    def normalize_template(self, template, mask, n_observations, axis=None):
        """In-place reference ``normalize_template`` (FP64 moments)."""
        t = self._torch
        self._require_gpu("normalize_template")
        if axis is not None or not (self._dev_f32(template) and self._dev_f32(mask)) or template.shape != mask.shape:
            return None
        keep, nptr, nhost = self._n_obs_args(n_observations)
        ws = self._workspace(_lib.lib().tm_reduce_workspace())
        _lib.call(
            "tm_normalize_template", template.data_ptr(), mask.data_ptr(), template.numel(), nptr, nhost,
            ws.data_ptr(), self._stream(),
        )
        return template

    def norm_scores(self, arr, exp_sq, sq_exp, n_obs, eps, out):
        """Reference ``norm_scores`` (npfftw_backend.py:537-587), numpy tie rule."""
        t = self._torch
        self._require_gpu("norm_scores")
        scalar_arr = not isinstance(arr, t.Tensor)
        if scalar_arr and float(arr) != 1.0:
            arr = t.full_like(exp_sq, float(arr))
            scalar_arr = False
        if not (self._dev_f32(exp_sq) and self._dev_f32(sq_exp) and self._dev_f32(out)):
            raise ValueError("norm_scores expects contiguous float32 GPU arrays")
        if not scalar_arr and not self._dev_f32(arr):
            raise ValueError("norm_scores expects contiguous float32 GPU arrays")
        if exp_sq.shape != sq_exp.shape or out.shape != exp_sq.shape:
            raise ValueError("norm_scores operands must share a shape")
        keep, nptr, nhost = self._n_obs_args(n_obs)
        _lib.call(
            "tm_norm_scores", None if scalar_arr else arr.data_ptr(), exp_sq.data_ptr(), sq_exp.data_ptr(),
            nptr, nhost, float(eps), out.data_ptr(), out.numel(), out.numel(), self._stream(),
        )
        return out

    def max_score_over_rotations(self, scores, max_scores, rotations, rotation_index: int):
        """Strict-``>`` running maximum (npfftw_backend.py:503-535)."""
        t = self._torch
        self._require_gpu("max_score_over_rotations")
        if not (self._dev_f32(scores) and self._dev_f32(max_scores)):
            raise ValueError("scores must be contiguous float32 GPU arrays")
        if rotations.dtype != t.int32 or not rotations.is_contiguous():
            raise ValueError("rotations must be a contiguous int32 GPU array")
        if scores.shape != max_scores.shape or rotations.shape != scores.shape:
            raise ValueError("score arrays must share a shape")
        _lib.call(
            "tm_max_update", scores.data_ptr(), max_scores.data_ptr(), rotations.data_ptr(), int(rotation_index),
            scores.numel(), self._stream(),
        )
        return max_scores, rotations

    # --------------------------------------------------------- peak extraction
    def max_filter_coordinates(self, score_space, min_distance):
        """Coordinates (n, 3) int64, in C order, of the voxels equal to the
        maximum of their box of edge ``min_distance`` with zero outside the
        volume (npfftw_backend.py:200-206: scipy ``maximum_filter`` with
        ``mode="constant"``, then ``nonzero``); ``tm_peaks_max_filter`` and
        ``tm_peaks_write``.  3-D float32 score volumes."""
        t = self._torch
        self._require_gpu("max_filter_coordinates")
        s = self.to_backend_array(score_space)
        if s.ndim != 3:
            raise ValueError(f"hip max_filter_coordinates takes a 3-D score volume, got {s.ndim} dimensions")
        if s.dtype != t.float32 or not s.is_contiguous():
            s = s.to(t.float32).contiguous()
        # scipy treats an edge of 0 as the identity filter, the same as 1
        size = max(1, int(np.max(min_distance)))
        F = s.numel()
        nblk = -(-F // 4096)
        # edges up to 32 filter in one pass without the xy-maxima volume
        tmp = t.empty(F, dtype=t.float32, device=s.device) if size > 32 else None
        mask = t.empty(64 * nblk, dtype=t.int64, device=s.device)
        cnt = t.empty(nblk, dtype=t.int64, device=s.device)
        off = t.empty(nblk + 1, dtype=t.int64, device=s.device)
        shape = _lib.i64_array(s.shape)
        _lib.call("tm_peaks_max_filter", s.data_ptr(), shape, size, None if tmp is None else tmp.data_ptr(),
                  mask.data_ptr(), cnt.data_ptr(), off.data_ptr(), self._stream())
        n = int(off[nblk].item())
        coords = t.empty((n, 3), dtype=t.int64, device=s.device)
        if n:
            _lib.call("tm_peaks_write", mask.data_ptr(), off.data_ptr(), shape, coords.data_ptr(), self._stream())
        return coords

    def unravel_index(self, indices, shape):
        return self._torch.unravel_index(self.to_backend_array(indices), tuple(int(x) for x in shape))

    def topk_indices(self, arr, k: int):
        """Indices of the k largest entries, largest first, as unravelled
        index arrays (npfftw_backend.py:176-181); ties go to the lower flat
        index and NaN ranks above every number; ``tm_topk_f32``."""
        t = self._torch
        self._require_gpu("topk_indices")
        flat = self.to_backend_array(arr).reshape(-1)
        if flat.dtype != t.float32 or not flat.is_contiguous():
            flat = flat.to(t.float32).contiguous()
        k = int(min(int(k), flat.numel()))
        if k <= 0:
            return self.unravel_index(t.empty(0, dtype=t.int64, device=flat.device), arr.shape)
        nbytes = int(_lib.lib().tm_topk_workspace_bytes(flat.numel(), k))
        key = (t.cuda.current_device(), "topk")
        ws = _WORKSPACES.get(key)
        if ws is None or ws.numel() < nbytes:
            ws = t.empty(nbytes, dtype=t.uint8, device=flat.device)
            _WORKSPACES[key] = ws
        out = t.empty(k, dtype=t.int64, device=flat.device)
        _lib.call("tm_topk_f32", flat.data_ptr(), flat.numel(), k, out.data_ptr(), ws.data_ptr(), nbytes,
                  self._stream())
        return self.unravel_index(out, arr.shape)

    #: rows the distance filter resolves in parallel; beyond this it keeps at
    #: most ``sequential_keep`` rows
    parallel_filter_rows = 16384
    sequential_keep = 4096

    def greedy_distance_filter(self, coordinates, min_distance: int, max_keep: int = None):
        """Row indices kept by the greedy minimum-distance pass over integer
        coordinates taken in order (the int64 ``find_candidate_indices`` of
        tme/external/bindings.cpp:84-124: a row is dropped when
        ``trunc(sqrt(d2)) <= min_distance`` for a kept row), the first
        ``max_keep`` of them; ``tm_peak_filter``."""
        t = self._torch
        self._require_gpu("greedy_distance_filter")
        c = self.to_backend_array(coordinates)
        if c.ndim != 2 or c.shape[1] != 3:
            raise ValueError("greedy_distance_filter takes (n, 3) coordinates")
        c = c.to(t.int64).contiguous()
        n = int(c.shape[0])
        if n == 0:
            return t.empty(0, dtype=t.int64, device=c.device)
        limit = n if max_keep is None else min(n, int(max_keep))
        if n > self.parallel_filter_rows and limit > self.sequential_keep:
            raise ValueError(
                f"the hip distance filter keeps at most {self.sequential_keep} of more than "
                f"{self.parallel_filter_rows} rows; pass max_keep"
            )
        nbytes = int(_lib.lib().tm_peak_filter_workspace_bytes(n))
        key = (t.cuda.current_device(), "peak_filter")
        ws = _WORKSPACES.get(key)
        if ws is None or ws.numel() < nbytes:
            ws = t.empty(nbytes, dtype=t.uint8, device=c.device)
            _WORKSPACES[key] = ws
        keep = t.empty(limit, dtype=t.int64, device=c.device)
        nkeep = t.zeros(1, dtype=t.int64, device=c.device)
        _lib.call("tm_peak_filter", c.data_ptr(), n, int(min_distance), limit, keep.data_ptr(), nkeep.data_ptr(),
                  ws.data_ptr(), nbytes, self._stream())
        return keep[: int(nkeep.item())]

    #: state rows + new rows the device caller step takes
    fold_max_rows = 8192

    def fold_peaks(self, state, candidates, score_space, rotation_matrix, num_peaks: int, min_distance: int,
                   min_boundary_distance: int, min_score=None, max_score=None):
        """One step of the peak caller on the device (``tm_peak_caller_fold``):
        the candidate positions (n, 3) of ``score_space`` folded into the
        caller ``state`` exactly as ``PeakCaller.__call__`` does with its
        mask, preselection, merge, sort and distance filter
        (tme/analyzer/peaks.py:241-290, 327-384, 531-583).  Returns the new
        state, or ``None`` when the step is outside what the kernel takes
        (the caller then runs the step itself)."""
        t = self._torch
        if not self._gpu:
            return None
        s = self.to_backend_array(score_space)
        c = self.to_backend_array(candidates)
        if s.ndim != 3 or c.ndim != 2 or c.shape[1] != 3 or s.dtype != t.float32 or not s.is_contiguous():
            return None
        pos, rot, sc, det = (self.to_backend_array(x) for x in state)
        n_cand, n_state = int(c.shape[0]), int(pos.shape[0])
        if n_cand == 0:
            return None
        if (pos.ndim != 2 or pos.shape[1] != 3 or tuple(rot.shape[1:]) != (3, 3) or rot.shape[0] != n_state
                or sc.shape != (n_state,) or det.shape != (n_state,)):
            return None
        kk = min(n_cand, 2 * num_peaks) if n_cand > num_peaks else n_cand
        if n_state + kk > self.fold_max_rows:
            return None
        dev = s.device
        c = c.to(t.int64).contiguous()
        pos = pos.to(device=dev, dtype=t.int64).contiguous()
        rot = rot.to(device=dev, dtype=t.float32).contiguous()
        sc = sc.to(device=dev, dtype=t.float32).contiguous()
        det = det.to(device=dev, dtype=t.float32).contiguous()
        R = self.to_backend_array(rotation_matrix).to(device=dev, dtype=t.float32).contiguous()
        if R.numel() != 9:
            return None
        nbytes = int(_lib.lib().tm_peak_caller_fold_workspace_bytes(n_cand, n_state, int(num_peaks)))
        key = (t.cuda.current_device(), "peak_fold")
        ws = _WORKSPACES.get(key)
        if ws is None or ws.numel() < nbytes:
            ws = t.empty(nbytes, dtype=t.uint8, device=dev)
            _WORKSPACES[key] = ws
        out_pos = t.empty((num_peaks, 3), dtype=t.int64, device=dev)
        out_rot = t.empty((num_peaks, 3, 3), dtype=t.float32, device=dev)
        out_sc = t.empty(num_peaks, dtype=t.float32, device=dev)
        out_det = t.empty(num_peaks, dtype=t.float32, device=dev)
        counts = t.empty(3, dtype=t.int64, device=dev)
        ptr = lambda x: x.data_ptr() if x.numel() else None  # noqa: E731
        _lib.call(
            "tm_peak_caller_fold", c.data_ptr(), n_cand, s.data_ptr(), _lib.i64_array(s.shape), ptr(pos), ptr(rot),
            ptr(sc), ptr(det), n_state, R.data_ptr(), int(num_peaks), int(min_distance), int(min_boundary_distance),
            int(min_score is not None), float(0.0 if min_score is None else min_score), int(max_score is not None),
            float(0.0 if max_score is None else max_score), out_pos.data_ptr(), out_rot.data_ptr(),
            out_sc.data_ptr(), out_det.data_ptr(), counts.data_ptr(), ws.data_ptr(), nbytes, self._stream(),
        )
        n_valid, _, n_out = (int(x) for x in counts.tolist())
        if n_valid == 0:
            return tuple(state)
        return out_pos[:n_out], out_rot[:n_out], out_sc[:n_out], out_det[:n_out]

    # ------------------------------------------------------------- fused scan
    def fused_scoring(self, **kwargs):
        """Batched fused scan hook used by the registered scores.

        Returns ``NotImplemented`` when the fused path does not apply, in
        which case the caller runs the plain per-rotation loop.
        """
        if not (self.fused and self._gpu):
            return NotImplemented
        from . import hip_fused

        collect = getattr(self, "_fused_collect", None)
        if collect is not None:
            # fused_group: record the scoring function's arguments, scan later
            collect.append((kwargs.pop("score"), kwargs))
            return kwargs["callback"]
        return hip_fused.run(self, **kwargs)

    def fused_group(self, scoring, jobs, stats=None):
        """Several templates' scans as one schedule (no counterpart in the
        reference, which runs one template per scripts/match_template.py
        call): ``jobs`` are ``(setup, rotations, callback)`` triples of the
        registered ``scoring`` function; each round runs one rotation batch
        of every template as one replayed hipGraph (``hip_fused.run_group``;
        ``stats``: its batch sizes and timing).  Raises ``ValueError`` where
        the fused scan does not apply."""
        if not (self.fused and self._gpu):
            raise ValueError("fused_group needs the fused scan on a GPU device")
        from . import hip_fused

        collected = []
        self._fused_collect = collected
        try:
            for setup, rotations, callback in jobs:
                scoring(rotations=rotations, callback=callback, **setup)
        finally:
            self._fused_collect = None
        if len(collected) != len(jobs):
            raise ValueError("fused_group: a scoring function did not take the fused path")
        hip_fused.run_group(self, collected, stats)
