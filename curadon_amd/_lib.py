"""ctypes binding of libcuradon_mi.so, the replacement for curadon's nanobind module
``curadon_ext`` (python/standalone/src/bind.cpp:58-66).

Names, argument order and error types follow the reference binding:
``plan_2d(device, vol_prec, vol_shape, vol_spacing, vol_offset, det_prec, det_count,
det_spacing, det_offset, DSO, DSD, angles, det_pitch, COR)`` with r/w ``forward_block_x`` /
``forward_block_y`` (python/standalone/src/plan.cu:30-59), ``forward_2d(volume, sinogram,
plan)`` (forward.cu:92-136), ``backward_2d(volume, sinogram, plan)`` (backward.cu:89-133),
``forward_3d`` / ``backward_3d`` with 15 arguments (forward.cu:23-75, backward.cu:20-72).

Arrays are torch tensors on a ROCm device; work is enqueued on the tensor device's current
torch stream. There is no CPU fallback: without the native library, importing this module
raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch  # loads the HIP runtime that libcuradon_mi.so binds to (same soname)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libcuradon_mi.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"curadon_amd: native library not found at {LIB_PATH}; build it with "
        "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C curadon_amd/csrc`"
    )

_lib = C.CDLL(LIB_PATH)

# ----------------------------------------------------------------------------- enums ----
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_DTYPE = 2
ERR_DEVICE = 3
ERR_HIP = 4
ERR_OUT_OF_MEMORY = 5
ERR_NOT_SUPPORTED = 6

F16, F32, BF16 = 16, 32, 1016
BP3D_REFERENCE, BP3D_MATCHED, BP3D_FDK = 0, 1, 2
BP2D_REFERENCE, BP2D_FBP = 0, 1
PLAN2D_EXACT_GEOMETRY = 1
_BP3D_MODES = {"reference": BP3D_REFERENCE, "matched": BP3D_MATCHED, "fdk": BP3D_FDK}
_BP2D_MODES = {"reference": BP2D_REFERENCE, "fbp": BP2D_FBP}

_DTYPE_CODE = {torch.float32: F32, torch.float16: F16, torch.bfloat16: BF16}
_CODE_DTYPE = {v: k for k, v in _DTYPE_CODE.items()}
_CODE_NAME = {F32: "f32", F16: "f16", BF16: "bf16"}


class CuradonError(RuntimeError):
    """A failure reported by the native library (HIP runtime error, out of memory...)."""

    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


class CuradonOutOfMemory(CuradonError, MemoryError):
    pass


_u64 = C.c_uint64
_i32 = C.c_int
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)


class Plan2DInfo(C.Structure):
    _fields_ = [
        ("device", C.c_int), ("vol_prec", C.c_int), ("det_prec", C.c_int),
        ("nx", _u64), ("ny", _u64), ("det_count", _u64), ("nangles", _u64),
        ("total_samples", _u64), ("table_bytes", _u64),
    ]


class Cone3DDesc(C.Structure):
    _fields_ = [
        ("vol_shape", _u64 * 3), ("vol_spacing", C.c_double * 3), ("vol_offset", C.c_double * 3),
        ("det_shape", _u64 * 2), ("det_spacing", C.c_double * 2), ("det_offset", C.c_double * 2),
        ("det_rotation", C.c_double * 3), ("DSO", C.c_double), ("DSD", C.c_double),
        ("COR", C.c_double), ("nangles", _u64), ("phi", _dp), ("theta", _dp), ("psi", _dp),
    ]


class Plan3DInfo(C.Structure):
    _fields_ = [
        ("device", C.c_int), ("vol_prec", C.c_int), ("det_prec", C.c_int),
        ("nx", _u64), ("ny", _u64), ("nz", _u64), ("nu", _u64), ("nv", _u64),
        ("nangles", _u64), ("total_samples", _u64), ("table_bytes", _u64),
    ]


class BPJob(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("mode", C.c_int32), ("columns", C.c_int32), ("sino", _vp),
        ("table", _vp),
        ("volume", _vp), ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
        ("nu", C.c_uint32), ("nv", C.c_uint32), ("nangles", C.c_uint32), ("sx", C.c_float),
        ("sy", C.c_float), ("ox", C.c_float), ("oy", C.c_float), ("inv_du", C.c_float),
        ("tile_begin", C.c_uint32), ("tile_end", C.c_uint32),
    ]


JOB_FAN2D, JOB_CONE3D = 0, 1

# Every symbol the public header declares, with its ctypes signature.
SIGNATURES = {
    "cmi_last_error": (C.c_char_p, []),
    "cmi_version": (C.c_int, []),
    "cmi_plan2d_create": (_i32, [_i32, _i32, _u64p, _dp, _dp, _i32, _u64, C.c_double, C.c_double,
                                 _dp, _dp, _dp, _u64, C.c_double, C.c_double, C.POINTER(_vp)]),
    "cmi_plan2d_create_vec": (_i32, [_i32, _i32, _u64p, _dp, _dp, _i32, _u64, _dp, _u64,
                                     C.POINTER(_vp)]),
    "cmi_plan2d_create_ex": (_i32, [_i32, _i32, _u64p, _dp, _dp, _i32, _u64, C.c_double,
                                    C.c_double, _dp, _dp, _dp, _u64, C.c_double, C.c_double,
                                    C.c_uint32, C.POINTER(_vp)]),
    "cmi_plan2d_destroy": (_i32, [_vp]),
    "cmi_plan2d_set_block": (_i32, [_vp, C.c_uint32, C.c_uint32]),
    "cmi_plan2d_get_block": (_i32, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "cmi_plan2d_get_info": (_i32, [_vp, C.POINTER(Plan2DInfo)]),
    "cmi_forward_2d": (_i32, [_vp, _vp, _i32, _vp, _i32, _u64, _vp]),
    "cmi_forward_2d_noalloc": (_i32, [_vp, _vp, _i32, _vp, _i32, _u64, _vp]),
    "cmi_backward_2d": (_i32, [_vp, _vp, _i32, _vp, _i32, _u64, _vp]),
    "cmi_backward_2d_weighted": (_i32, [_vp, _vp, _i32, _vp, _i32, _u64, _i32, _vp]),
    "cmi_adjoint_2d": (_i32, [_vp, _vp, _i32, _vp, _i32, _u64, _vp]),
    "cmi_backward_2d_transpose": (_i32, [_vp, _vp, _i32, _vp, _i32, _u64, _i32, _vp]),
    "cmi_plan3d_create": (_i32, [_i32, _i32, _i32, C.POINTER(Cone3DDesc), C.POINTER(_vp)]),
    "cmi_plan3d_create_vec": (_i32, [_i32, _i32, _i32, _u64p, _dp, _dp, _u64p, _dp, _u64,
                                     C.POINTER(_vp)]),
    "cmi_plan3d_destroy": (_i32, [_vp]),
    "cmi_plan3d_get_info": (_i32, [_vp, C.POINTER(Plan3DInfo)]),
    "cmi_forward_3d": (_i32, [_vp, _vp, _i32, _vp, _i32, _u64, _vp]),
    "cmi_backward_3d": (_i32, [_vp, _vp, _i32, _vp, _i32, _u64, _i32, _vp]),
    "cmi_adjoint_3d": (_i32, [_vp, _vp, _i32, _vp, _i32, _u64, _vp]),
    "cmi_backward_3d_transpose": (_i32, [_vp, _vp, _i32, _vp, _i32, _u64, _i32, _vp]),
    "cmi_fan2d_bp_table": (_i32, [_u64, _dp, _dp, _dp, C.POINTER(C.c_float)]),
    "cmi_cone3d_bp_matrices": (_i32, [C.POINTER(Cone3DDesc), _i32, C.POINTER(C.c_float),
                                      C.POINTER(C.c_int)]),
    "cmi_stage_projections": (_i32, [_i32, _vp, _u64, _u64, _u64, _u64, C.c_uint32, _vp,
                                     C.c_float, _vp, _i32, _vp]),
    "cmi_filter_pack": (_i32, [_i32, _vp, _u64, _u64, _vp, _u64, _u64, _vp, _vp]),
    "cmi_filter_unpack": (_i32, [_i32, _vp, _u64, _u64, _u64, _vp, _vp]),
    "cmi_bp_job_tiles": (_i32, [C.POINTER(BPJob), _u64p]),
    "cmi_server_create": (_i32, [_i32, C.c_uint32, C.POINTER(_vp)]),
    "cmi_server_destroy": (_i32, [_vp]),
    "cmi_server_get_capacity": (_i32, [_vp, C.POINTER(C.c_uint32)]),
    "cmi_serve_backproject": (_i32, [_vp, C.POINTER(BPJob), C.c_uint32, _i32, _i32, _vp]),
}

for _name, (_res, _args) in SIGNATURES.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def last_error() -> str:
    return _lib.cmi_last_error().decode("utf-8", "replace")


def version() -> int:
    return _lib.cmi_version()


def check(status: int):
    """Raise the Python error the reference binding would raise for this failure class."""
    if status == OK:
        return
    msg = last_error()
    if status == ERR_DEVICE:
        raise AttributeError(msg)        # nb::attribute_error, forward.cu:95-105
    if status == ERR_DTYPE:
        raise TypeError(msg)
    if status == ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == ERR_OUT_OF_MEMORY:
        raise CuradonOutOfMemory(status, msg)
    if status == ERR_NOT_SUPPORTED:
        raise NotImplementedError(msg)
    raise CuradonError(status, msg)


def precision_code(prec) -> int:
    """Map a precision argument to the library's dtype code.

    Accepts the reference's ints (32 -> f32, 16 -> f16; plan.cu:20-28), torch dtypes and the
    strings 'f32', 'f16', 'bf16'."""
    if isinstance(prec, torch.dtype):
        if prec in _DTYPE_CODE:
            return _DTYPE_CODE[prec]
    elif isinstance(prec, str):
        m = {"f32": F32, "float32": F32, "f16": F16, "float16": F16, "bf16": BF16,
             "bfloat16": BF16}
        if prec.lower() in m:
            return m[prec.lower()]
    else:
        try:
            v = int(prec)
        except (TypeError, ValueError):
            v = None
        if v in (32, 16):
            return F32 if v == 32 else F16
        if v == BF16:
            return BF16
    raise AttributeError("Precision must either be f32, f16 or bf16")  # plan.cu:26


def dtype_of_code(code: int) -> torch.dtype:
    return _CODE_DTYPE[code]


def tensor_code(t: torch.Tensor) -> int:
    if t.dtype not in _DTYPE_CODE:
        raise TypeError("Only float32, float16 and bfloat16 are supported")  # forward.cu:134
    return _DTYPE_CODE[t.dtype]


def _require_device_tensor(t, what):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise TypeError(f"{what} must be a ROCm device tensor (got {t.device})")
    if not t.is_contiguous():
        raise TypeError(f"{what} must be C-contiguous")


def stream_of(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _f64(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    if n is not None and a.size != n:
        raise AttributeError(f"expected {n} values, got {a.size}")
    return a


# ----------------------------------------------------------------------------- 2D -------

class plan_2d:
    """Per-geometry state of the 2D fan-beam operators (replaces curad::plan_2d).

    Construction builds the per-angle geometry in FP64 on the host and the per-ray table on
    ``device``; the plan is immutable afterwards apart from the launch-shape tunables."""

    def __init__(self, device, vol_prec, vol_shape, vol_spacing, vol_offset, det_prec,
                 det_count, det_spacing, det_offset, DSO, DSD, angles, det_pitch=0.0, COR=0.0,
                 exact_geometry=False):
        """``exact_geometry`` (not in the reference): det_offset shifts the detector along its
        axis, det_pitch turns it about its centre and COR moves the rotation centre, in the
        forward, backprojection and adjoint alike (cmi_plan2d_create_ex); without it, the
        reference's semantics (the forward moves the source by det_offset, the backprojection
        ignores it, det_pitch and COR are unused)."""
        self._h = None
        angles = _f64(angles)
        A = angles.size
        DSO = _f64(DSO, A)
        DSD = _f64(DSD, A)
        vs = np.ascontiguousarray(np.asarray(vol_shape, dtype=np.uint64).reshape(-1))
        if vs.size != 2:
            raise AttributeError("vol_shape must be two values")
        sp = _f64(vol_spacing, 2)
        of = _f64(vol_offset, 2)
        self.device = int(device)
        self.vol_prec = precision_code(vol_prec)
        self.det_prec = precision_code(det_prec)
        self.vol_shape = (int(vs[0]), int(vs[1]))
        self.det_count = int(det_count)
        self.nangles = A
        self.exact_geometry = bool(exact_geometry)
        h = _vp()
        check(_lib.cmi_plan2d_create_ex(
            self.device, self.vol_prec, vs.ctypes.data_as(_u64p), sp.ctypes.data_as(_dp),
            of.ctypes.data_as(_dp), self.det_prec, self.det_count, float(det_spacing),
            float(det_offset), DSO.ctypes.data_as(_dp), DSD.ctypes.data_as(_dp),
            angles.ctypes.data_as(_dp), A, float(det_pitch), float(COR),
            PLAN2D_EXACT_GEOMETRY if exact_geometry else 0, C.byref(h)))
        self._h = h

    @classmethod
    def from_views(cls, device, vol_prec, vol_shape, vol_spacing, vol_offset, det_prec,
                   det_count, vecs):
        """Plan from ASTRA-like views ([A][6]: source x, y, detector centre x, y, cell step
        u x, y in world coordinates; cmi_plan2d_create_vec)."""
        self = cls.__new__(cls)
        self._h = None
        v = np.ascontiguousarray(np.asarray(vecs, dtype=np.float64))
        if v.ndim != 2 or v.shape[1] != 6:
            raise AttributeError("vecs must have shape (nangles, 6)")
        vs = np.ascontiguousarray(np.asarray(vol_shape, dtype=np.uint64).reshape(-1))
        if vs.size != 2:
            raise AttributeError("vol_shape must be two values")
        sp = _f64(vol_spacing, 2)
        of = _f64(vol_offset, 2)
        self.device = int(device)
        self.vol_prec = precision_code(vol_prec)
        self.det_prec = precision_code(det_prec)
        self.vol_shape = (int(vs[0]), int(vs[1]))
        self.det_count = int(det_count)
        self.nangles = int(v.shape[0])
        h = _vp()
        check(_lib.cmi_plan2d_create_vec(
            self.device, self.vol_prec, vs.ctypes.data_as(_u64p), sp.ctypes.data_as(_dp),
            of.ctypes.data_as(_dp), self.det_prec, self.det_count, v.ctypes.data_as(_dp),
            self.nangles, C.byref(h)))
        self._h = h
        return self

    @property
    def handle(self):
        return self._h

    def _block(self):
        bx, by = C.c_uint32(), C.c_uint32()
        check(_lib.cmi_plan2d_get_block(self._h, C.byref(bx), C.byref(by)))
        return bx.value, by.value

    def set_block(self, block_x, block_y):
        """Set both launch-shape tunables at once (block_x * block_y <= 256)."""
        check(_lib.cmi_plan2d_set_block(self._h, int(block_x), int(block_y)))

    @property
    def forward_block_x(self):
        return self._block()[0]

    @forward_block_x.setter
    def forward_block_x(self, v):
        check(_lib.cmi_plan2d_set_block(self._h, int(v), self._block()[1]))

    @property
    def forward_block_y(self):
        return self._block()[1]

    @forward_block_y.setter
    def forward_block_y(self, v):
        check(_lib.cmi_plan2d_set_block(self._h, self._block()[0], int(v)))

    def info(self) -> dict:
        i = Plan2DInfo()
        check(_lib.cmi_plan2d_get_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in Plan2DInfo._fields_}

    def sinogram_shape(self):
        return (self.nangles, self.det_count)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h is not None and _lib is not None:
            _lib.cmi_plan2d_destroy(h)


def _check_pair_2d(volume, sinogram, plan, which):
    _require_device_tensor(volume, "volume")
    _require_device_tensor(sinogram, "sinogram")
    if volume.device != sinogram.device:
        raise AttributeError("Volume and sinogram must be on the same device")   # forward.cu:95-97
    if volume.device.index != plan.device:
        raise AttributeError("Volume and plan are not on the same device")  # forward.cu:99-101
    vc, sc = tensor_code(volume), tensor_code(sinogram)
    if vc != plan.vol_prec:
        raise AttributeError(
            f"Plan precision does not match volume precision ({_CODE_NAME[vc]})")  # forward.cu:108-110
    if sc != plan.det_prec:
        raise AttributeError(
            f"Plan precision does not match sinogram precision ({_CODE_NAME[sc]})")
    ny, nx = plan.vol_shape
    A, D = plan.nangles, plan.det_count
    if volume.dim() < 2 or tuple(volume.shape[-2:]) != (ny, nx):
        raise ValueError(f"{which}: volume shape {tuple(volume.shape)} does not end in {(ny, nx)}")
    if sinogram.dim() < 2 or tuple(sinogram.shape[-2:]) != (A, D):
        raise ValueError(f"{which}: sinogram shape {tuple(sinogram.shape)} does not end in {(A, D)}")
    bv = volume.numel() // (ny * nx)
    bs = sinogram.numel() // (A * D)
    if bv != bs or tuple(volume.shape[:-2]) != tuple(sinogram.shape[:-2]):
        raise ValueError(f"{which}: batch shapes differ: volume {tuple(volume.shape)}, "
                         f"sinogram {tuple(sinogram.shape)}")
    return vc, sc, bv


def forward_2d(volume, sinogram, plan):
    """sinogram := A volume (overwrites). Batched over leading dimensions."""
    vc, sc, b = _check_pair_2d(volume, sinogram, plan, "forward_2d")
    if b == 0:
        return
    check(_lib.cmi_forward_2d(plan.handle, volume.data_ptr(), vc, sinogram.data_ptr(), sc, b,
                              stream_of(volume)))


# Bytes of the static staging scratch of forward_2d_noalloc (CMI_NOALLOC_SCRATCH_BYTES).
NOALLOC_SCRATCH_BYTES = 64 << 20


def forward_2d_noalloc(volume, sinogram, plan):
    """forward_2d through cmi_forward_2d_noalloc: the same result without any device allocation
    (the staging scratch is the library's static array of NOALLOC_SCRATCH_BYTES per GPU; calls
    that need more raise). Calls that overlap on one GPU race on that array."""
    vc, sc, b = _check_pair_2d(volume, sinogram, plan, "forward_2d_noalloc")
    if b == 0:
        return
    check(_lib.cmi_forward_2d_noalloc(plan.handle, volume.data_ptr(), vc, sinogram.data_ptr(),
                                      sc, b, stream_of(volume)))


def backward_2d(volume, sinogram, plan, weighting="reference"):
    """volume += B sinogram, the reference's voxel-driven backprojection; weighting="fbp"
    squares the magnification weight (fan-beam filtered backprojection)."""
    w = {"reference": BP2D_REFERENCE, "fbp": BP2D_FBP}.get(weighting)
    if w is None:
        raise ValueError(f"unknown weighting {weighting!r}")
    vc, sc, b = _check_pair_2d(volume, sinogram, plan, "backward_2d")
    if b == 0:
        return
    check(_lib.cmi_backward_2d_weighted(plan.handle, volume.data_ptr(), vc, sinogram.data_ptr(),
                                        sc, b, w, stream_of(volume)))


def backward_2d_transpose(volume, sinogram, plan, weighting="reference"):
    """sinogram := B^T volume (overwrites): the transpose of backward_2d with the same
    weighting, the gradient of the voxel-driven backprojection."""
    w = {"reference": BP2D_REFERENCE, "fbp": BP2D_FBP}.get(weighting)
    if w is None:
        raise ValueError(f"unknown weighting {weighting!r}")
    vc, sc, b = _check_pair_2d(volume, sinogram, plan, "backward_2d_transpose")
    if b == 0:
        return
    check(_lib.cmi_backward_2d_transpose(plan.handle, sinogram.data_ptr(), sc,
                                         volume.data_ptr(), vc, b, w, stream_of(volume)))


def adjoint_2d(volume, sinogram, plan):
    """volume += A^T sinogram, the exact transpose of forward_2d."""
    vc, sc, b = _check_pair_2d(volume, sinogram, plan, "adjoint_2d")
    if b == 0:
        return
    check(_lib.cmi_adjoint_2d(plan.handle, volume.data_ptr(), vc, sinogram.data_ptr(), sc, b,
                              stream_of(volume)))


# ----------------------------------------------------------------------------- 3D -------

def _cone_desc(vol_shape, vol_spacing, vol_offset, phi, theta, psi, det_shape, det_spacing,
               det_offset, det_rotation, DSO, DSD, COR):
    """A cmi_cone3d_desc and the FP64 angle arrays it points to (keep them alive with it)."""
    phi = _f64(phi)
    A = phi.size
    theta, psi = _f64(theta, A), _f64(psi, A)
    d = Cone3DDesc()
    vs = [int(v) for v in np.asarray(vol_shape).reshape(-1)]
    ds = [int(v) for v in np.asarray(det_shape).reshape(-1)]
    if len(vs) != 3 or len(ds) != 2:
        raise AttributeError("vol_shape needs 3 values and det_shape 2")
    d.vol_shape[:] = vs
    d.vol_spacing[:] = list(_f64(vol_spacing, 3))
    d.vol_offset[:] = list(_f64(vol_offset, 3))
    d.det_shape[:] = ds
    d.det_spacing[:] = list(_f64(det_spacing, 2))
    d.det_offset[:] = list(_f64(det_offset, 2))
    d.det_rotation[:] = list(_f64(det_rotation, 3))
    d.DSO, d.DSD, d.COR = float(DSO), float(DSD), float(COR)
    d.nangles = A
    d.phi = phi.ctypes.data_as(_dp)
    d.theta = theta.ctypes.data_as(_dp)
    d.psi = psi.ctypes.data_as(_dp)
    return d, (phi, theta, psi)


def _bp_matrices_of(desc, mode, with_layout):
    out = np.zeros((int(desc.nangles), 3, 4), dtype=np.float32)
    yinv = C.c_int()
    check(_lib.cmi_cone3d_bp_matrices(C.byref(desc), _BP3D_MODES[mode],
                                      out.ctypes.data_as(C.POINTER(C.c_float)),
                                      C.byref(yinv)))
    return (out, bool(yinv.value)) if with_layout else out


def cone3d_bp_matrices(vol_shape, vol_spacing, vol_offset, phi, theta, psi, det_shape,
                       det_spacing, det_offset, det_rotation, DSO, DSD, COR, mode="fdk",
                       with_layout=False):
    """The backprojection matrices of `mode` for a cone-beam description (arguments as
    plan_3d), computed on the host: no device plan and no synchronisation."""
    d, keep = _cone_desc(vol_shape, vol_spacing, vol_offset, phi, theta, psi, det_shape,
                         det_spacing, det_offset, det_rotation, DSO, DSD, COR)
    r = _bp_matrices_of(d, mode, with_layout)
    del keep
    return r


class plan_3d:
    """Per-geometry state of the 3D cone-beam operators. The reference has no 3D plan (it
    rebuilds its tables on every call, forward_3d.hpp:214-274); this one holds the per-angle
    geometry and the backprojection matrices on ``device``.

    Arguments follow curadon_ext.forward_3d (forward.cu:23-38) in Python axis order:
    vol_shape / vol_spacing / vol_offset (z, y, x), angles phi / theta / psi (A,),
    det_shape / det_spacing / det_offset (v, u), det_rotation (pitch, roll, yaw)."""

    def __init__(self, device, vol_prec, det_prec, vol_shape, vol_spacing, vol_offset, phi,
                 theta, psi, det_shape, det_spacing, det_offset, det_rotation, DSO, DSD, COR):
        self._h = None
        self.device = int(device)
        self.vol_prec = precision_code(vol_prec)
        self.det_prec = precision_code(det_prec)
        d, keep = _cone_desc(vol_shape, vol_spacing, vol_offset, phi, theta, psi, det_shape,
                             det_spacing, det_offset, det_rotation, DSO, DSD, COR)
        self._desc, self._keep = d, keep
        A = int(d.nangles)
        self.vol_shape = tuple(int(v) for v in d.vol_shape)
        self.sino_shape = (A, int(d.det_shape[0]), int(d.det_shape[1]))
        self.nangles = A
        h = _vp()
        check(_lib.cmi_plan3d_create(self.device, self.vol_prec, self.det_prec, C.byref(d),
                                     C.byref(h)))
        self._h = h

    @classmethod
    def from_views(cls, device, vol_prec, det_prec, vol_shape, vol_spacing, vol_offset,
                   det_shape, vecs):
        """Plan from ASTRA-like cone views ([A][12]: source, panel centre, u-step, v-step in
        world (x, y, z); cmi_plan3d_create_vec). Shapes in Python axis order: vol_shape /
        vol_spacing / vol_offset (z, y, x), det_shape (v, u)."""
        self = cls.__new__(cls)
        self._h = None
        v = np.ascontiguousarray(np.asarray(vecs, dtype=np.float64))
        if v.ndim != 2 or v.shape[1] != 12:
            raise AttributeError("vecs must have shape (nangles, 12)")
        vs = np.ascontiguousarray(np.asarray(vol_shape, dtype=np.uint64).reshape(-1))
        ds = np.ascontiguousarray(np.asarray(det_shape, dtype=np.uint64).reshape(-1))
        if vs.size != 3 or ds.size != 2:
            raise AttributeError("vol_shape needs 3 values and det_shape 2")
        sp = _f64(vol_spacing, 3)
        of = _f64(vol_offset, 3)
        self.device = int(device)
        self.vol_prec = precision_code(vol_prec)
        self.det_prec = precision_code(det_prec)
        self.vol_shape = tuple(int(x) for x in vs)
        A = int(v.shape[0])
        self.sino_shape = (A, int(ds[0]), int(ds[1]))
        self.nangles = A
        self._desc = None
        self._keep = (v, vs, ds, sp, of)
        h = _vp()
        check(_lib.cmi_plan3d_create_vec(
            self.device, self.vol_prec, self.det_prec, vs.ctypes.data_as(_u64p),
            sp.ctypes.data_as(_dp), of.ctypes.data_as(_dp), ds.ctypes.data_as(_u64p),
            v.ctypes.data_as(_dp), A, C.byref(h)))
        self._h = h
        return self

    @property
    def handle(self):
        return self._h

    def bp_matrices(self, mode="fdk", with_layout=False):
        """(A, 3, 4) FP32 projection matrices cmi_backward_3d uses in `mode`; with_layout also
        returns whether their U and W rows are free of y (the backprojectors then run along
        y). Plans from views have no descriptor and raise."""
        if self._desc is None:
            raise NotImplementedError("backprojection matrices of a plan from views")
        return _bp_matrices_of(self._desc, mode, with_layout)

    def info(self) -> dict:
        i = Plan3DInfo()
        check(_lib.cmi_plan3d_get_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in Plan3DInfo._fields_}

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h is not None and _lib is not None:
            _lib.cmi_plan3d_destroy(h)


def _check_pair_3d(volume, sinogram, plan, which):
    _require_device_tensor(volume, "volume")
    _require_device_tensor(sinogram, "sinogram")
    if volume.device != sinogram.device:
        raise AttributeError("Volume and sinogram must be on the same device")   # forward.cu:40-42
    if volume.device.index != plan.device:
        raise AttributeError("Volume and plan are not on the same device")
    vc, sc = tensor_code(volume), tensor_code(sinogram)
    if vc != plan.vol_prec:
        raise AttributeError(f"Plan precision does not match volume precision ({_CODE_NAME[vc]})")
    if sc != plan.det_prec:
        raise AttributeError(
            f"Plan precision does not match sinogram precision ({_CODE_NAME[sc]})")
    if volume.dim() < 3 or tuple(volume.shape[-3:]) != plan.vol_shape:
        raise ValueError(f"{which}: volume shape {tuple(volume.shape)} does not end in {plan.vol_shape}")
    if sinogram.dim() < 3 or tuple(sinogram.shape[-3:]) != plan.sino_shape:
        raise ValueError(f"{which}: sinogram shape {tuple(sinogram.shape)} does not end in "
                         f"{plan.sino_shape}")
    if tuple(volume.shape[:-3]) != tuple(sinogram.shape[:-3]):
        raise ValueError(f"{which}: batch shapes differ")
    b = 1
    for v in volume.shape[:-3]:
        b *= int(v)
    return vc, sc, b


def forward_3d_plan(volume, sinogram, plan):
    """sinogram := A volume (cone beam). Rays that miss the volume write 0."""
    vc, sc, b = _check_pair_3d(volume, sinogram, plan, "forward_3d")
    if b:
        check(_lib.cmi_forward_3d(plan.handle, volume.data_ptr(), vc, sinogram.data_ptr(), sc,
                                  b, stream_of(volume)))


def backward_3d_plan(volume, sinogram, plan, mode="reference"):
    """volume += B sinogram; mode 'reference' (the reference's projection), 'matched' (the
    forward model's projection) or 'fdk' (matched, weighted by (DSO / depth)^2)."""
    m = _BP3D_MODES.get(mode)
    if m is None:
        raise ValueError(f"unknown backprojection mode {mode!r}")
    vc, sc, b = _check_pair_3d(volume, sinogram, plan, "backward_3d")
    if b:
        check(_lib.cmi_backward_3d(plan.handle, volume.data_ptr(), vc, sinogram.data_ptr(), sc,
                                   b, m, stream_of(volume)))


def backward_3d_transpose(volume, sinogram, plan, mode="reference"):
    """sinogram := B^T volume (overwrites): the transpose of backward_3d_plan in the same mode,
    the gradient of the voxel-driven backprojection."""
    m = _BP3D_MODES.get(mode)
    if m is None:
        raise ValueError(f"unknown backprojection mode {mode!r}")
    vc, sc, b = _check_pair_3d(volume, sinogram, plan, "backward_3d_transpose")
    if b:
        check(_lib.cmi_backward_3d_transpose(plan.handle, sinogram.data_ptr(), sc,
                                             volume.data_ptr(), vc, b, m, stream_of(volume)))


def adjoint_3d_plan(volume, sinogram, plan):
    """volume += A^T sinogram (transpose of forward_3d_plan)."""
    vc, sc, b = _check_pair_3d(volume, sinogram, plan, "adjoint_3d")
    if b:
        check(_lib.cmi_adjoint_3d(plan.handle, volume.data_ptr(), vc, sinogram.data_ptr(), sc, b,
                                  stream_of(volume)))


_PLAN3D_CACHE = {}


def _plan3d_for(x, vol_shape, vol_spacing, vol_offset, sinogram, phi, theta, psi, sino_shape,
                det_spacing, det_offset, det_rotation, DSO, DSD, COR):
    key = (x.device.index, x.dtype, sinogram.dtype,
           tuple(np.asarray(vol_shape).reshape(-1).tolist()),
           tuple(_f64(vol_spacing).tolist()), tuple(_f64(vol_offset).tolist()),
           _f64(phi).tobytes(), _f64(theta).tobytes(), _f64(psi).tobytes(),
           tuple(np.asarray(sino_shape).reshape(-1).tolist()), tuple(_f64(det_spacing).tolist()),
           tuple(_f64(det_offset).tolist()), tuple(_f64(det_rotation).tolist()),
           float(DSO), float(DSD), float(COR))
    p = _PLAN3D_CACHE.get(key)
    if p is None:
        if len(_PLAN3D_CACHE) >= 8:
            _PLAN3D_CACHE.pop(next(iter(_PLAN3D_CACHE)))
        p = plan_3d(x.device.index, x.dtype, sinogram.dtype, vol_shape, vol_spacing, vol_offset,
                    phi, theta, psi, sino_shape, det_spacing, det_offset, det_rotation, DSO, DSD,
                    COR)
        _PLAN3D_CACHE[key] = p
    return p


def forward_3d(x, vol_shape, vol_spacing, vol_offset, sinogram, phi, theta, psi, sino_shape,
               det_spacing, det_offset, det_rotation, DSO, DSD, COR):
    """curadon_ext.forward_3d with its 15 arguments (forward.cu:23-75); the plan for the
    geometry is built once and cached."""
    p = _plan3d_for(x, vol_shape, vol_spacing, vol_offset, sinogram, phi, theta, psi, sino_shape,
                    det_spacing, det_offset, det_rotation, DSO, DSD, COR)
    forward_3d_plan(x, sinogram, p)


def backward_3d(x, vol_shape, vol_spacing, vol_offset, sinogram, phi, theta, psi, sino_shape,
                det_spacing, det_offset, det_rotation, DSO, DSD, COR):
    """curadon_ext.backward_3d with its 15 arguments (backward.cu:20-72), reference mode."""
    p = _plan3d_for(x, vol_shape, vol_spacing, vol_offset, sinogram, phi, theta, psi, sino_shape,
                    det_spacing, det_offset, det_rotation, DSO, DSD, COR)
    backward_3d_plan(x, sinogram, p)


# ----------------------------------------------------------------------------- serving --

def fan2d_bp_table(angles, DSO, DSD) -> np.ndarray:
    """(A, 4) FP32 backprojection table (-sin, cos, DSD, DSO), as the 2D plan builds it."""
    ang = _f64(angles)
    A = ang.size
    dso, dsd = _f64(DSO, A), _f64(DSD, A)
    out = np.zeros((A, 4), dtype=np.float32)
    check(_lib.cmi_fan2d_bp_table(A, ang.ctypes.data_as(_dp), dso.ctypes.data_as(_dp),
                                  dsd.ctypes.data_as(_dp),
                                  out.ctypes.data_as(C.POINTER(C.c_float))))
    return out


def filter_pack(src, cw, dst):
    """dst (ceil(rows / 2), L, 2) FP32 := row pairs of src (rows, D) times their weight rows cw
    (cw_rows, D), zero padded to L (cmi_filter_pack). Enqueued on the current stream."""
    for t, name in ((src, "source"), (cw, "weights"), (dst, "destination")):
        _require_device_tensor(t, name)
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise TypeError(f"{name} must be a contiguous FP32 tensor")
    D = src.shape[-1]
    rows = src.numel() // D
    if cw.shape[-1] != D or dst.dim() != 3 or dst.shape[0] != (rows + 1) // 2 or dst.shape[2] != 2:
        raise ValueError("filter_pack: shapes do not match")
    if src.device != cw.device or src.device != dst.device:
        raise ValueError("filter_pack: tensors on different devices")
    check(_lib.cmi_filter_pack(dst.device.index or 0, src.data_ptr(), rows, D, cw.data_ptr(),
                               cw.numel() // D, dst.shape[1], dst.data_ptr(),
                               torch.cuda.current_stream(dst.device).cuda_stream))


def filter_unpack(src, dst):
    """dst (rows, D) FP32 := the real and imaginary parts of the first D entries of the row
    pairs in src (ceil(rows / 2), L, 2) FP32 (cmi_filter_unpack). Enqueued on the current
    stream."""
    for t, name in ((src, "source"), (dst, "destination")):
        _require_device_tensor(t, name)
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise TypeError(f"{name} must be a contiguous FP32 tensor")
    D = dst.shape[-1]
    rows = dst.numel() // D
    if src.dim() != 3 or src.shape[0] != (rows + 1) // 2 or src.shape[2] != 2 or src.shape[1] < D:
        raise ValueError("filter_unpack: shapes do not match")
    if src.device != dst.device:
        raise ValueError("filter_unpack: tensors on different devices")
    check(_lib.cmi_filter_unpack(dst.device.index or 0, src.data_ptr(), rows, D, src.shape[1],
                                 dst.data_ptr(), torch.cuda.current_stream(dst.device).cuda_stream))


def stage_projections(src, dst, row_scale=None, scale=1.0):
    """dst (A, D + 4) or (A, V + 4, U + 4) := zero-bordered scale * row_scale[a] * src, where
    src is FP32 (A, D) / (A, V, U) or a wider buffer whose first U entries per row are used
    (for instance an inverse-FFT output of length L >= U). Enqueued on the current stream."""
    _require_device_tensor(dst, "destination")
    if src.dtype != torch.float32 or src.device != dst.device:
        raise TypeError("source must be FP32 on the destination's device")
    if src.stride(-1) != 1:
        raise TypeError("source rows must be contiguous")
    if dst.dim() == 2:
        A, Dp = dst.shape
        V, U, vpad = 1, Dp - 4, 0
        if src.dim() != 2 or src.shape[0] != A or src.shape[1] < U:
            raise ValueError("source shape does not match the staged layout")
        stride = src.stride(0)
    elif dst.dim() == 3:
        A, Vp, Up = dst.shape
        V, U, vpad = Vp - 4, Up - 4, 2
        if src.dim() != 3 or src.shape[0] != A or src.shape[1] != V or src.shape[2] < U:
            raise ValueError("source shape does not match the staged layout")
        stride = src.stride(1)
        if src.stride(0) != V * stride:
            raise ValueError("source angles must be evenly strided")
    else:
        raise ValueError("destination must be 2D or 3D")
    rs = None
    if row_scale is not None:
        rs = row_scale.to(device=dst.device, dtype=torch.float32).contiguous()
        if rs.numel() != A:
            raise ValueError("row_scale needs one value per angle")
    check(_lib.cmi_stage_projections(dst.device.index or 0, src.data_ptr(), stride, A, V, U,
                                     vpad, rs.data_ptr() if rs is not None else None,
                                     float(scale), dst.data_ptr(), tensor_code(dst),
                                     stream_of(dst)))
    return rs  # keep alive until the copy has run


def bp_job_tiles(job: BPJob) -> int:
    n = _u64()
    check(_lib.cmi_bp_job_tiles(C.byref(job), C.byref(n)))
    return int(n.value)


class Server:
    """Multi-job backprojection launcher (cmi_server): one persistent launch per call of
    `backproject` over a list of BPJob descriptors."""

    def __init__(self, device=0, max_jobs=256):
        self._h = None
        self.device = int(device)
        self.max_jobs = int(max_jobs)
        h = _vp()
        check(_lib.cmi_server_create(self.device, self.max_jobs, C.byref(h)))
        self._h = h
        cap = C.c_uint32()
        check(_lib.cmi_server_get_capacity(self._h, C.byref(cap)))
        self.capacity = int(cap.value)

    def backproject(self, jobs, vol_dtype, sino_dtype, stream=None):
        n = len(jobs)
        arr = (BPJob * n)(*jobs)
        if stream is None:
            stream = torch.cuda.current_stream(self.device).cuda_stream
        check(_lib.cmi_serve_backproject(self._h, arr, n, precision_code(vol_dtype),
                                         precision_code(sino_dtype), stream))

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h is not None and _lib is not None:
            _lib.cmi_server_destroy(h)
