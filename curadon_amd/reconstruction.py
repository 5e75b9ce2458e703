"""Filtered backprojection around the operators (the pipeline of examples/fbp.py:27-99).

Fan beam (flat detector, full 2 pi scan):
    f = (pi / A) * (DSO / DSD) * B_fbp[ (cos_w * p) (*) h ]
FDK cone beam (flat panel, circular scan):
    f = (pi / A) * DSD / (du * DSO) * B_fdk[ (cos_w * p) (*)_u h ]
where cos_w = DSD / sqrt(DSD^2 + u^2 [+ v^2]) is the cosine pre-weight, (*) the convolution
along detector rows with the band-limited ramp h of unit cell spacing (Kak and Slaney,
"Principles of Computerized Tomographic Imaging", eq. 3.61: h[0] = 1/4, h[k odd] =
-1/(pi k)^2, h[k even] = 0) times a frequency window, B_fbp the 2D backprojector with the
squared magnification weight and B_fdk the geometry-matched 3D backprojector with the
(DSO / depth)^2 weight. The convolution is that of an FFT of length >= 2 D (zero padded, so
it is linear, not circular; the length has only the prime factors 2, 3 and 5), evaluated on
the device for rows of up to 1024 cells as a product with the filter's Toeplitz matrix. The
result is in the image's density units.

Windows follow examples/fbp.py:27-48 over the normalised frequency |W| in [0, 1] (1 = Nyquist):
ram-lak 1, shepp-logan sinc(W / (2 s)), cosine cos(W pi / (2 s)) (the reference multiplies by
an undefined H there), hamming 0.54 + 0.46 cos(W pi / s), hann cos(W pi / (2 s))^2, and a
cutoff |W| <= s (the reference tests W <= s, which keeps every negative frequency).

`reference_filter` reproduces examples/fbp.py exactly (circular FFT filtering with
fftshift(|linspace(-1, 1, D)|), times pi / (2 A)), for callers that want its output.
"""
from __future__ import annotations

import functools
import threading
import math

import torch

from .geometry import ConeGeometry, FanGeometry

WINDOWS = ("ram-lak", "shepp-logan", "cosine", "hamming", "hann")


def _window(W: torch.Tensor, name, s: float) -> torch.Tensor:
    a = W.abs()
    if name is None:
        return torch.ones_like(a)
    if name == "ram-lak":
        w = torch.ones_like(a)
    elif name == "shepp-logan":
        w = torch.sinc(W / (2 * s))
    elif name == "cosine":
        w = torch.cos(W * math.pi / (2 * s))
    elif name == "hamming":
        w = 0.54 + 0.46 * torch.cos(a * math.pi / s)
    elif name == "hann":
        w = torch.cos(a * math.pi / (2 * s)) ** 2
    else:
        raise RuntimeError(f"Unknown Filter {name}")
    return w * (a <= s)


def fft_length(n_det: int) -> int:
    """Smallest even L >= 2 n_det whose only prime factors are 2, 3 and 5 (sizes rocFFT runs
    with its fastest radices); zero padding to L makes the row convolution linear."""
    L = 2 * n_det
    while True:
        m = L
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1 and L % 2 == 0:
            return L
        L += 1


@functools.lru_cache(maxsize=64)
def _ramp_response_host(n_det: int, window, frequency_scaling: float) -> torch.Tensor:
    return _ramp_response_f64(n_det, window, frequency_scaling).to(torch.float32)


@functools.lru_cache(maxsize=16)
def _ramp_response_f64(n_det: int, window, frequency_scaling: float) -> torch.Tensor:
    L = fft_length(n_det)
    k = torch.arange(L, dtype=torch.float64)
    m = torch.where(k <= L // 2, k, k - L)                   # signed index, FFT order
    h = torch.zeros(L, dtype=torch.float64)
    h[0] = 0.25
    odd = (m.remainder(2) != 0)
    h[odd] = -1.0 / (math.pi * m[odd]) ** 2
    H = torch.fft.rfft(h).real                               # symmetric kernel: real response
    W = torch.arange(L // 2 + 1, dtype=torch.float64) / (L // 2)
    return H * _window(W, window, frequency_scaling)


@functools.lru_cache(maxsize=64)
def _ramp_response_dev(n_det: int, window, frequency_scaling: float, device) -> torch.Tensor:
    return _ramp_response_host(n_det, window, frequency_scaling).to(device)


def ramp_response(n_det: int, window="ram-lak", frequency_scaling: float = 1.0,
                  device=None) -> torch.Tensor:
    """Real frequency response (rfft layout) of the windowed band-limited ramp for rows of
    n_det cells, FFT length L = fft_length(n_det). Returns (L // 2 + 1,) float32. Responses
    are computed once per (n_det, window, scaling) and kept on each device they were asked
    for; the caller gets its own copy."""
    key = (int(n_det), window, float(frequency_scaling))
    if device is None or torch.device(device).type == "cpu":
        return _ramp_response_host(*key).clone()
    return _ramp_response_dev(*key, torch.device(device)).clone()


def filter_rows(p: torch.Tensor, H: torch.Tensor) -> torch.Tensor:
    """Linear convolution of every row (last axis) of p with the filter of response H, in
    FP64 for FP64 input and FP32 otherwise."""
    D = p.shape[-1]
    L = 2 * (H.shape[0] - 1)
    dt = torch.float64 if p.dtype == torch.float64 else torch.float32
    P = torch.fft.rfft(p.to(dt), n=L, dim=-1)
    return torch.fft.irfft(P * H.to(dt), n=L, dim=-1)[..., :D]


# Longest FFT run as complex row pairs: rocFFT transforms single-precision rows up to 4096
# points in one pass through LDS; longer ones take several passes, and a real transform of
# length L uses a complex one of L / 2.
PAIR_FFT_MAX = 4096


# Rows of at most this many cells are filtered as one FP32 matrix product with the D x D
# Toeplitz matrix of the filter (a library GEMM on the matrix cores): 2 D multiply-adds per
# output instead of the FFTs' five passes over HBM. Measured on MI355X for (rows, D):
# (393216, 384) 0.58 ms against 1.78 ms by pair FFTs, (442368, 576) 2.8 against 6.0 ms,
# (2048, 1449) even, (4096, 2897) 0.77 against 0.45 ms.
TOEPLITZ_MAX = 1024


@functools.lru_cache(maxsize=16)
def _toeplitz(n_det: int, window, frequency_scaling: float, device) -> torch.Tensor:
    """(n_det, n_det) FP32 matrix T[j, k] = h[(k - j) mod L], h the impulse response of the
    windowed ramp (FP64 inverse transform of its length-L response): a row x times T is the
    first n_det entries of the circular convolution of x, zero-padded to L = fft_length(n_det),
    with h, which is the linear convolution filter_rows computes."""
    L = fft_length(n_det)
    h = torch.fft.irfft(_ramp_response_f64(n_det, window, frequency_scaling), n=L)
    j = torch.arange(n_det)[:, None]
    k = torch.arange(n_det)[None, :]
    return h[(k - j) % L].to(torch.float32).to(device).contiguous()


_MM_LOCK = threading.Lock()


def _mm_ieee(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """a @ b with full FP32 products whatever the caller's TF32 setting. The precision is
    chosen when the product is enqueued; it is switched and restored through one API family
    (torch.backends.cuda.matmul.fp32_precision), which leaves the caller's state as it was in
    every combination of the legacy and new settings, and under a lock, so concurrent callers
    of this function do not restore each other's values."""
    m = torch.backends.cuda.matmul
    with _MM_LOCK:
        old = m.fp32_precision
        if old == "ieee":
            return torch.mm(a, b)
        m.fp32_precision = "ieee"
        try:
            return torch.mm(a, b)
        finally:
            m.fp32_precision = old


@functools.lru_cache(maxsize=64)
def _full_response(n_det: int, window, frequency_scaling: float, device) -> torch.Tensor:
    H = _ramp_response_dev(n_det, window, frequency_scaling, device)
    return torch.cat([H, H[1:-1].flip(0)]).contiguous()    # length L, real and even


def filter_rows_weighted(p: torch.Tensor, cw: torch.Tensor, n_det: int, window="ram-lak",
                         frequency_scaling: float = 1.0) -> torch.Tensor:
    """Linear convolution of every row of p * cw (cw's rows repeat over p's rows: (A, D) for
    a fan beam, (V, U) for a cone beam) with the windowed ramp for rows of n_det cells, for FP32
    device tensors. Rows of at most TOEPLITZ_MAX cells: one FP32 product with the filter's
    Toeplitz matrix. Longer rows: FFTs of length L = fft_length(n_det); up to L = PAIR_FFT_MAX
    in pairs through complex FFTs (the response is real and even, so the real and imaginary
    parts stay apart), with native kernels for the weighting, padding, pairing and unpairing;
    longer ones through real FFTs, whose half-length complex transform is then the shorter one.
    Returns a contiguous FP32 tensor shaped like p."""
    from . import _lib
    D = p.shape[-1]
    p = p.contiguous()
    cw = cw.to(device=p.device, dtype=torch.float32).contiguous()
    if D == int(n_det) and D <= TOEPLITZ_MAX:
        T = _toeplitz(D, window, float(frequency_scaling), p.device)
        w = (p.reshape(-1, cw.shape[0], D) * cw).reshape(-1, D)
        return _mm_ieee(w, T).reshape(p.shape)
    L = fft_length(int(n_det))
    if L > PAIR_FFT_MAX:
        H = _ramp_response_dev(int(n_det), window, float(frequency_scaling), p.device)
        w = (p.reshape(-1, cw.shape[0], D) * cw).reshape(p.shape)
        y = torch.fft.irfft(torch.fft.rfft(w, n=L, dim=-1) * H, n=L, dim=-1)
        return y[..., :D].contiguous()
    Hf = _full_response(int(n_det), window, float(frequency_scaling), p.device)
    rows = p.numel() // D
    z = torch.empty(((rows + 1) // 2, L, 2), dtype=torch.float32, device=p.device)
    _lib.filter_pack(p, cw, z)
    Z = torch.fft.fft(torch.view_as_complex(z), dim=-1)
    del z
    Z.mul_(Hf)
    y = torch.view_as_real(torch.fft.ifft(Z, dim=-1))
    del Z
    out = torch.empty(p.shape, dtype=torch.float32, device=p.device)
    _lib.filter_unpack(y, out)
    return out


def cosine_weights_fan(geom: FanGeometry, device=None) -> torch.Tensor:
    """(A, D) cosine pre-weights DSD / sqrt(DSD^2 + u^2), u the cell centre on the detector.
    Evaluated in FP64 on `device` (correctly rounded operations: the same values on any
    device) and returned as FP32."""
    D = geom.det_shape
    u = (torch.arange(D, dtype=torch.float64, device=device) - (D - 1) / 2) * geom.det_spacing
    dsd = torch.as_tensor(geom.DSD, dtype=torch.float64).to(device)[:, None]
    return (dsd / torch.sqrt(dsd * dsd + (u * u)[None, :])).to(torch.float32)


def cosine_weights_cone(geom: ConeGeometry, device=None) -> torch.Tensor:
    """(V, U) cosine pre-weights DSD / sqrt(DSD^2 + u^2 + v^2), FP64 on `device`, as FP32."""
    V, U = (int(v) for v in geom.det_shape)
    dv, du = (float(v) for v in geom.det_spacing)
    u = (torch.arange(U, dtype=torch.float64, device=device) - (U - 1) / 2) * du
    v = (torch.arange(V, dtype=torch.float64, device=device) - (V - 1) / 2) * dv
    r2 = (v * v)[:, None] + (u * u)[None, :]
    return (geom.DSD / torch.sqrt(geom.DSD * geom.DSD + r2)).to(torch.float32)


def _native_filter(sinogram: torch.Tensor, dt) -> bool:
    # FP32 device projections go through the native pair filter; FP64 and host tensors
    # through torch.fft's real transforms
    return sinogram.is_cuda and dt == torch.float32 and sinogram.dtype == torch.float32


def fbp_filter(sinogram: torch.Tensor, geom, window="ram-lak", frequency_scaling=1.0):
    """Cosine-weighted, ramp-filtered and scaled projections (FP32), ready for the weighted
    backprojection (`backward(..., mode="fbp")` for fan beam, mode="fdk" for cone beam)."""
    dev = sinogram.device
    dt = torch.float64 if sinogram.dtype == torch.float64 else torch.float32
    if (isinstance(geom, FanGeometry) and not geom.is_centred_circular()) or (
            isinstance(geom, ConeGeometry) and geom.vecs is not None):
        raise NotImplementedError(
            "fbp: geometries built from ASTRA-like views (from_vec), or fan geometries with "
            "an exact detector offset, rotation or COR, have no FBP / FDK weighting here; use "
            "backward / adjoint, or a centred circular geometry")
    if isinstance(geom, FanGeometry):
        cw = cosine_weights_fan(geom, dev).to(dt)
        scale = (math.pi / geom.nangles) * torch.as_tensor(
            geom.DSO / geom.DSD, dtype=dt, device=dev)[:, None]
        if _native_filter(sinogram, dt):
            return filter_rows_weighted(sinogram, cw, geom.det_shape, window,
                                        frequency_scaling) * scale
        H = ramp_response(geom.det_shape, window, frequency_scaling, dev)
        return filter_rows(sinogram.to(dt) * cw, H) * scale
    if isinstance(geom, ConeGeometry):
        cw = cosine_weights_cone(geom, dev).to(dt)
        scale = (math.pi / geom.nangles) * geom.DSD / (float(geom.det_spacing[1]) * geom.DSO)
        if _native_filter(sinogram, dt):
            return filter_rows_weighted(sinogram, cw, int(geom.det_shape[1]), window,
                                        frequency_scaling) * scale
        H = ramp_response(int(geom.det_shape[1]), window, frequency_scaling, dev)
        return filter_rows(sinogram.to(dt) * cw, H) * scale
    raise ValueError("Unknown geometry type")


def fbp(sinogram: torch.Tensor, geom, window="ram-lak", frequency_scaling=1.0, volume=None):
    """Filtered backprojection (fan beam) / FDK (cone beam) into `volume` (+=) or a new one.
    The projections are filtered in FP32 (FP64 for FP64 input) and the filtered projections
    cast to the geometry's sinogram precision before the backprojection. Pass the raw
    projections in FP32: rounding them to FP16 before the ramp filter costs ~1.5e-3 relative L2
    on its own, storing the filtered ones in FP16 ~2e-5."""
    from . import _lib
    from .backward import backward
    q = fbp_filter(sinogram, geom, window, frequency_scaling)
    q = q.to(_lib.dtype_of_code(geom.plan.det_prec)).contiguous()
    mode = "fbp" if isinstance(geom, FanGeometry) else "fdk"
    return backward(q, volume, geom, mode=mode)


def reference_filter(sinogram: torch.Tensor, filter_name="ram-lak", frequency_scaling=1.0):
    """examples/fbp.py's filtering verbatim: circular FFT convolution with
    fftshift(window(|linspace(-1, 1, D)|)), times pi / (2 A)."""
    A, D = sinogram.shape[-2], sinogram.shape[-1]
    W = torch.linspace(-1, 1, D, dtype=torch.float64)
    w = W.abs() * (_window(W, filter_name, frequency_scaling) if filter_name else 1.0)
    if filter_name is None:
        w = torch.ones(D, dtype=torch.float64)
    w = torch.fft.fftshift(w).to(torch.float32).to(sinogram.device)
    out = torch.fft.ifft(torch.fft.fft(sinogram.to(torch.float32), dim=-1) * w, dim=-1).real
    return out * (math.pi / (2 * A))
