"""ctypes binding of libygrid's C ABI (include/ygrid.h)."""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
YGRID_MEM_HOST = 0
YGRID_MEM_DEVICE = 1

EXPORTED = [
    "ygrid_last_error", "ygrid_version", "ygrid_ctx_create", "ygrid_ctx_destroy", "ygrid_ctx_set_stream",
    "ygrid_ctx_synchronize", "ygrid_ctx_clear_cache", "ygrid_ctx_set_small_support", "ygrid_gridder_create", "ygrid_gridder_configure",
    "ygrid_gridder_clone", "ygrid_gridder_destroy",
    "ygrid_initialise_grid", "ygrid_grid", "ygrid_finalise_grid", "ygrid_finalise_weights",
    "ygrid_initialise_degrid", "ygrid_degrid", "ygrid_degrid_noalloc", "ygrid_finalise_degrid", "ygrid_is_model_empty",
    "ygrid_set_mfs_weight", "ygrid_get_stats", "ygrid_set_profiling", "ygrid_get_cf", "ygrid_get_grid",
    "ygrid_batch_create", "ygrid_batch_destroy", "ygrid_batch_add", "ygrid_batch_set_graph", "ygrid_batch_run",
    "ygrid_batch_autotune", "ygrid_batch_info",
    "ygrid_lcf_create", "ygrid_lcf_destroy", "ygrid_lcf_num_params", "ygrid_lcf_get_params", "ygrid_lcf_set_params",
    "ygrid_lcf_clear_cache",
    "ygrid_lcf_loss_grad", "ygrid_lcf_forward", "ygrid_lcf_train_wproject", "ygrid_gridder_set_learned_cf",
    "ygrid_imaging_chunk", "ygrid_uvw_rotate", "ygrid_ctx_set_small_band_rows", "ygrid_gridder_autotune_small",
]


class YgridError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ygrid error {code}: {msg}")
        self.code = code


class _ImageDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("npol", ctypes.c_int), ("nchan", ctypes.c_int),
                ("cellx", ctypes.c_double), ("celly", ctypes.c_double), ("has_freq_axis", ctypes.c_int),
                ("freq_start", ctypes.c_double), ("freq_end", ctypes.c_double),
                ("centre_lon", ctypes.c_double), ("centre_lat", ctypes.c_double), ("stokes", ctypes.c_void_p)]


class _Chunk(ctypes.Structure):
    _fields_ = [("nrow", ctypes.c_int), ("nchan", ctypes.c_int), ("npol", ctypes.c_int), ("mem", ctypes.c_int),
                ("uvw", ctypes.c_void_p), ("delay", ctypes.c_void_p), ("freq", ctypes.c_void_p),
                ("vis", ctypes.c_void_p), ("vis_out", ctypes.c_void_p), ("noise", ctypes.c_void_p),
                ("flag", ctypes.c_void_p), ("feed1", ctypes.c_void_p), ("pointing1", ctypes.c_void_p),
                ("stokes", ctypes.c_void_p), ("dish_pointing1", ctypes.c_void_p), ("vis_out_rows", ctypes.c_int),
                ("feed1_pa", ctypes.c_void_p)]


class _UvwRotation(ctypes.Structure):
    _fields_ = [("nrow", ctypes.c_int), ("mem", ctypes.c_int), ("uvw", ctypes.c_void_p),
                ("phase_centre", ctypes.c_void_p), ("phase_centre_per_row", ctypes.c_int),
                ("tangent_lon", ctypes.c_double), ("tangent_lat", ctypes.c_double),
                ("centre_lon", ctypes.c_double), ("centre_lat", ctypes.c_double),
                ("uvw_out", ctypes.c_void_p), ("delay_out", ctypes.c_void_p)]


class Stats(ctypes.Structure):
    _fields_ = [("samples_gridded", ctypes.c_double), ("taps_gridded", ctypes.c_double),
                ("samples_degridded", ctypes.c_double), ("taps_degridded", ctypes.c_double),
                ("vectors_flagged", ctypes.c_double), ("vectors_wflagged", ctypes.c_double),
                ("time_grid_s", ctypes.c_double), ("time_degrid_s", ctypes.c_double), ("time_cf_s", ctypes.c_double),
                ("support", ctypes.c_int), ("oversample", ctypes.c_int), ("n_cf", ctypes.c_int),
                ("ms_prep", ctypes.c_double), ("ms_bin", ctypes.c_double), ("ms_grid_kernel", ctypes.c_double),
                ("ms_degrid_kernel", ctypes.c_double), ("n_grid_launches", ctypes.c_double),
                ("n_degrid_launches", ctypes.c_double), ("n_bin_reused", ctypes.c_double),
                ("grid_waves", ctypes.c_int), ("grid_pipeline", ctypes.c_int),
                ("degrid_waves", ctypes.c_int), ("degrid_pipeline", ctypes.c_int),
                ("cf_precision", ctypes.c_int), ("vis_precision", ctypes.c_int), ("learned_cf_active", ctypes.c_int),
                ("degrid_matrix", ctypes.c_int), ("rows_rejected_pointing", ctypes.c_double),
                ("small_support_passes", ctypes.c_double),
                ("cf_iterations", ctypes.c_double), ("cf_generations", ctypes.c_double),
                ("cf_generations_pa", ctypes.c_double), ("cf_generations_freq", ctypes.c_double),
                ("snapshot_fits", ctypes.c_double), ("snapshot_regrids", ctypes.c_double),
                ("snapshot_a", ctypes.c_double), ("snapshot_b", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("grid_waves", ctypes.c_int), ("grid_pipeline", ctypes.c_int), ("degrid_waves", ctypes.c_int),
                ("degrid_pipeline", ctypes.c_int), ("item_pairs", ctypes.c_int), ("ms", ctypes.c_double),
                ("trials", ctypes.c_int)]


class LcfTrainOpts(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_int), ("batch", ctypes.c_int), ("lr", ctypes.c_float),
                ("seed", ctypes.c_ulonglong), ("holdout_every", ctypes.c_int)]


class LcfReport(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_int), ("steps", ctypes.c_int), ("examples", ctypes.c_int), ("cfs", ctypes.c_int),
                ("holdout_cfs", ctypes.c_int), ("first_loss", ctypes.c_double), ("last_loss", ctypes.c_double),
                ("holdout_max_err", ctypes.c_double), ("holdout_rms_err", ctypes.c_double),
                ("train_s", ctypes.c_double), ("support", ctypes.c_int)]


def library_path() -> str:
    return os.environ.get("YGRID_LIBRARY", os.path.join(_HERE, "libygrid.so"))


_LIB = None


def lib():
    """Load libygrid.so; raises if it has not been built (no fallback exists)."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python -m yandasoft_amd.build`")
        L = ctypes.CDLL(path)
        L.ygrid_last_error.restype = ctypes.c_char_p
        for name in EXPORTED:
            getattr(L, name)  # every declared entry point must be exported
        _LIB = L
    return _LIB


def _check(rc: int):
    if rc != 0:
        raise YgridError(rc, lib().ygrid_last_error().decode())


@dataclass
class ImageDesc:
    nx: int
    ny: int
    npol: int = 1
    nchan: int = 1
    cellx: float = 0.0
    celly: float = 0.0
    freq_axis: tuple | None = None  # (start, end) Hz
    centre: tuple = (0.0, 0.0)      # image centre (lon, lat) rad
    stokes: tuple | None = None     # casacore Stokes codes of the image pols (None: the data's frame)

    def c(self) -> _ImageDesc:
        fa = self.freq_axis
        self._stokes = None if self.stokes is None else np.asarray(self.stokes, dtype=np.int32)
        return _ImageDesc(self.nx, self.ny, self.npol, self.nchan, self.cellx, self.celly, 1 if fa else 0,
                          fa[0] if fa else 0.0, fa[1] if fa else 0.0, self.centre[0], self.centre[1],
                          None if self._stokes is None else self._stokes.ctypes.data)


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())  # torch tensor


@dataclass
class Chunk:
    """Accessor chunk.  Arrays are in casacore memory order: uvw (nrow, 3) float64;
    cubes (npol, nchan, nrow) complex64 in numpy C order (= casacore Cube(nRow,nChan,nPol)).
    Either all numpy (host) or all torch cuda tensors (device)."""
    uvw: object
    freq: object
    vis: object = None
    vis_out: object = None
    noise: object = None
    flag: object = None
    delay: object = None
    feed1: object = None
    pointing1: object = None
    stokes: tuple | None = None  # casacore Stokes codes of the visibility pols (None: the image's)
    dish_pointing1: object = None  # (nrow, 2) radians or None (= pointing1)
    vis_out_rows: int = 0          # rows of the cube vis_out points into (0: nrow)
    feed1_pa: object = None        # (nrow,) float32 parallactic angle of each row's first feed, rad
    nrow: int = 0
    nchan: int = 0
    npol: int = 0
    device: bool = False

    def __post_init__(self):
        self.nrow = int(self.uvw.shape[0])
        self.nchan = int(self.freq.shape[0])
        cube = self.vis if self.vis is not None else (self.vis_out if self.vis_out is not None else self.noise)
        self.npol = int(cube.shape[0]) if cube is not None else self.npol
        if not self.device:
            for name in ("uvw", "freq", "vis", "vis_out", "noise", "flag", "delay", "feed1", "pointing1",
                         "dish_pointing1", "feed1_pa"):
                a = getattr(self, name)
                if a is not None:
                    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"], f"{name} must be C-contiguous"
            if self.feed1_pa is not None:
                assert self.feed1_pa.dtype == np.float32 and self.feed1_pa.shape == (self.nrow,), \
                    "feed1_pa must be (nrow,) float32"

    def c(self) -> _Chunk:
        self._stokes = None if self.stokes is None else np.asarray(self.stokes, dtype=np.int32)
        return _Chunk(self.nrow, self.nchan, self.npol, YGRID_MEM_DEVICE if self.device else YGRID_MEM_HOST,
                      _ptr(self.uvw), _ptr(self.delay), _ptr(self.freq), _ptr(self.vis), _ptr(self.vis_out),
                      _ptr(self.noise), _ptr(self.flag), _ptr(self.feed1), _ptr(self.pointing1),
                      None if self._stokes is None else self._stokes.ctypes.data, _ptr(self.dish_pointing1),
                      int(self.vis_out_rows), _ptr(self.feed1_pa))


class Context:
    def __init__(self, device: int = 0):
        self.h = ctypes.c_void_p()
        _check(lib().ygrid_ctx_create(device, ctypes.byref(self.h)))

    def set_stream(self, stream_ptr: int):
        _check(lib().ygrid_ctx_set_stream(self.h, ctypes.c_void_p(stream_ptr)))

    def synchronize(self):
        _check(lib().ygrid_ctx_synchronize(self.h))

    def clear_cache(self):
        """Drop the cached sample binning; the next pass bins again."""
        _check(lib().ygrid_ctx_clear_cache(self.h))

    def set_small_support(self, max_support: int):
        """Passes with CF support <= max_support (0..8) run the small-support kernels; 0: none."""
        _check(lib().ygrid_ctx_set_small_support(self.h, int(max_support)))

    def set_small_band_rows(self, rows: int):
        """Rows (4, 8, 16; 0: by support) of the small-support gridding kernel's LDS band."""
        _check(lib().ygrid_ctx_set_small_band_rows(self.h, int(rows)))

    def close(self):
        if self.h:
            _check(lib().ygrid_ctx_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def parset_text(params: dict | None) -> bytes:
    if not params:
        return b""
    return "\n".join(f"{k}={str(v).lower() if isinstance(v, bool) else v}" for k, v in params.items()).encode()


class Gridder:
    """IVisGridder mirror over the C ABI."""

    def __init__(self, ctx: Context, name: str, params: dict | None = None, _handle=None):
        self.ctx = ctx
        if _handle is not None:
            self.h = _handle
        else:
            self.h = ctypes.c_void_p()
            _check(lib().ygrid_gridder_create(ctx.h, name.encode(), parset_text(params), ctypes.byref(self.h)))
        self.img = None

    def configure(self, options: dict):
        """ygrid_gridder_configure: the factory's global options (padding, alldatapsf,
        oversampleweight, cleargrids, MaxPointingSeparation, parotation)."""
        text = "\n".join(f"{k}={v}" for k, v in options.items()).encode()
        _check(lib().ygrid_gridder_configure(self.h, text))

    def clone(self) -> "Gridder":
        h = ctypes.c_void_p()
        _check(lib().ygrid_gridder_clone(self.h, ctypes.byref(h)))
        g = Gridder(self.ctx, "", _handle=h)
        g.img = self.img
        return g

    def initialise_grid(self, img: ImageDesc, dopsf: bool = False, dopcf: bool = False):
        self.img = img
        d = img.c()
        _check(lib().ygrid_initialise_grid(self.h, ctypes.byref(d), int(dopsf), int(dopcf)))

    def grid(self, chunk: Chunk):
        c = chunk.c()
        _check(lib().ygrid_grid(self.h, ctypes.byref(c)))

    def finalise_grid(self, out=None):
        """Returns the image as numpy (nchan, npol, ny, nx) unless a device tensor is given."""
        img = self.img
        if out is None:
            out = np.zeros((img.nchan, img.npol, img.ny, img.nx), dtype=np.float32)
            _check(lib().ygrid_finalise_grid(self.h, ctypes.c_void_p(out.ctypes.data), YGRID_MEM_HOST))
        else:
            _check(lib().ygrid_finalise_grid(self.h, ctypes.c_void_p(_ptr(out)), YGRID_MEM_DEVICE))
        return out

    def finalise_weights(self):
        img = self.img
        out = np.zeros((img.nchan, img.npol, img.ny, img.nx), dtype=np.float32)
        _check(lib().ygrid_finalise_weights(self.h, ctypes.c_void_p(out.ctypes.data), YGRID_MEM_HOST))
        return out

    def initialise_degrid(self, img: ImageDesc, model):
        self.img = img
        d = img.c()
        if isinstance(model, np.ndarray):
            m = np.ascontiguousarray(model, dtype=np.float32)
            _check(lib().ygrid_initialise_degrid(self.h, ctypes.byref(d), ctypes.c_void_p(m.ctypes.data), YGRID_MEM_HOST))
        else:
            _check(lib().ygrid_initialise_degrid(self.h, ctypes.byref(d), ctypes.c_void_p(_ptr(model)), YGRID_MEM_DEVICE))

    def degrid(self, chunk: Chunk):
        c = chunk.c()
        _check(lib().ygrid_degrid(self.h, ctypes.byref(c)))

    def autotune_small(self, img: ImageDesc, chunk: Chunk, trials: int = 2):
        """ygrid_gridder_autotune_small: (best band rows, its ms); the context keeps the choice."""
        d, c = img.c(), chunk.c()
        rows, ms = ctypes.c_int(), ctypes.c_double()
        _check(lib().ygrid_gridder_autotune_small(self.h, ctypes.byref(d), ctypes.byref(c), int(trials),
                                                  ctypes.byref(rows), ctypes.byref(ms)))
        return rows.value, ms.value

    def degrid_noalloc(self, chunk: Chunk):
        """ygrid_degrid_noalloc: degrid() with the pass's buffers in the library's fixed per-GPU
        workspace (device-resident chunk, CF cache already built)."""
        c = chunk.c()
        _check(lib().ygrid_degrid_noalloc(self.h, ctypes.byref(c)))

    def finalise_degrid(self):
        _check(lib().ygrid_finalise_degrid(self.h))

    def is_model_empty(self) -> bool:
        v = ctypes.c_int()
        _check(lib().ygrid_is_model_empty(self.h, ctypes.byref(v)))
        return bool(v.value)

    def set_mfs_weight(self, ref_freq: float, order: int):
        _check(lib().ygrid_set_mfs_weight(self.h, ctypes.c_double(ref_freq), order))

    def set_learned_cf(self, lcf, max_err: float):
        """Attach a trained LearnedCF (used while its held-out error <= max_err) or detach (None)."""
        _check(lib().ygrid_gridder_set_learned_cf(self.h, lcf.h if lcf is not None else None,
                                                  ctypes.c_double(max_err)))
        self._lcf = lcf

    def set_profiling(self, on: bool = True):
        _check(lib().ygrid_set_profiling(self.h, int(on)))

    def stats(self) -> Stats:
        s = Stats()
        _check(lib().ygrid_get_stats(self.h, ctypes.byref(s)))
        return s

    def cf_plane(self, k: int) -> np.ndarray:
        S = ctypes.c_int()
        _check(lib().ygrid_get_cf(self.h, k, None, ctypes.byref(S)))
        W = 2 * S.value + 1
        out = np.zeros((W, W), dtype=np.complex64)
        _check(lib().ygrid_get_cf(self.h, k, ctypes.c_void_p(out.ctypes.data), ctypes.byref(S)))
        return out

    def grid_array(self) -> np.ndarray:
        n = ctypes.c_size_t()
        _check(lib().ygrid_get_grid(self.h, None, ctypes.byref(n)))
        out = np.zeros(n.value // 2, dtype=np.complex64)
        _check(lib().ygrid_get_grid(self.h, ctypes.c_void_p(out.ctypes.data), ctypes.byref(n)))
        return out

    def close(self):
        if self.h:
            _check(lib().ygrid_gridder_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def imaging_chunk(degridders, gridders, chunk: Chunk):
    """One chunk of the imaging equation's data loop (ImageFFTEquation.cc:437-480) on the
    device: residual = chunk.vis - sum of the degridders' predictions, gridded by every gridder;
    chunk.vis_out (optional) receives the residual."""
    D = (ctypes.c_void_p * max(1, len(degridders)))(*[g.h.value for g in degridders])
    G = (ctypes.c_void_p * max(1, len(gridders)))(*[g.h.value for g in gridders])
    c = chunk.c()
    _check(lib().ygrid_imaging_chunk(D, len(degridders), G, len(gridders), ctypes.byref(c)))


def uvw_rotate(ctx: Context, uvw, phase_centre, tangent, centre, uvw_out=None, delay_out=None):
    """rotatedUVW(tangent) and uvwRotationDelay(tangent, centre) of an accessor chunk on the device
    (ygrid_uvw_rotate).  uvw: (nrow, 3) metres in each row's phase-centre frame; phase_centre:
    (2,) or (nrow, 2) (lon, lat) radians; tangent, centre: (lon, lat).  numpy arrays (host: the
    results are returned as new arrays unless given) or device arrays with data_ptr() (device:
    uvw_out and delay_out must be given; asynchronous on the context's stream)."""
    device = not isinstance(uvw, np.ndarray)
    n = int(uvw.shape[0])
    if not device:
        uvw = np.ascontiguousarray(uvw, dtype=np.float64)
        phase_centre = np.ascontiguousarray(phase_centre, dtype=np.float64)
        uvw_out = np.empty((n, 3)) if uvw_out is None else uvw_out
        delay_out = np.empty(n) if delay_out is None else delay_out
    per_row = 1 if (len(phase_centre.shape) == 2 and phase_centre.shape[0] == n and n != 1) else 0
    r = _UvwRotation(n, YGRID_MEM_DEVICE if device else YGRID_MEM_HOST, _ptr(uvw), _ptr(phase_centre), per_row,
                     float(tangent[0]), float(tangent[1]), float(centre[0]), float(centre[1]), _ptr(uvw_out),
                     _ptr(delay_out))
    _check(lib().ygrid_uvw_rotate(ctx.h, ctypes.byref(r)))
    return uvw_out, delay_out


class Batch:
    """Batched passes over several gridders (ygrid_batch_*): chunks in device memory (torch
    tensors); run() gives every gridder the result of grid() / degrid() on its chunk."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        self._keep = []
        _check(lib().ygrid_batch_create(ctx.h, ctypes.byref(self.h)))

    def add(self, gridder: Gridder, chunk: Chunk):
        assert chunk.device, "batched passes need device chunks"
        c = chunk.c()
        self._keep.append((gridder, chunk))
        _check(lib().ygrid_batch_add(self.h, gridder.h, ctypes.byref(c)))

    def set_graph(self, on: bool = True):
        _check(lib().ygrid_batch_set_graph(self.h, int(on)))

    def run(self):
        _check(lib().ygrid_batch_run(self.h))

    def autotune(self, trials: int = 2) -> Tuning:
        t = Tuning()
        _check(lib().ygrid_batch_autotune(self.h, trials, ctypes.byref(t)))
        return t

    def info(self):
        ms, caps = ctypes.c_double(), ctypes.c_int()
        _check(lib().ygrid_batch_info(self.h, ctypes.byref(ms), ctypes.byref(caps)))
        return ms.value, caps.value

    def close(self):
        if self.h:
            _check(lib().ygrid_batch_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LearnedCF:
    """Coordinate network for CF taps (ygrid_lcf_*)."""

    def __init__(self, ctx: Context, hidden: int = 128, layers: int = 4, seed: int = 1):
        self.ctx = ctx
        self.hidden, self.layers = hidden, layers
        self.h = ctypes.c_void_p()
        _check(lib().ygrid_lcf_create(ctx.h, hidden, layers, ctypes.c_ulonglong(seed), ctypes.byref(self.h)))

    @property
    def nparams(self) -> int:
        n = ctypes.c_int()
        _check(lib().ygrid_lcf_num_params(self.h, ctypes.byref(n)))
        return n.value

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.nparams, np.float32)
        _check(lib().ygrid_lcf_get_params(self.h, ctypes.c_void_p(out.ctypes.data)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        _check(lib().ygrid_lcf_set_params(self.h, ctypes.c_void_p(p.ctypes.data)))

    def clear_cache(self):
        """Forget the last evaluation (the next pass evaluates the network)."""
        _check(lib().ygrid_lcf_clear_cache(self.h))

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = ctypes.c_double()
        _check(lib().ygrid_lcf_loss_grad(self.h, ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(t.ctypes.data),
                                         x.shape[0], ctypes.byref(loss), ctypes.c_void_p(g.ctypes.data)))
        return loss.value, g

    def forward(self, x, half=False) -> np.ndarray:
        """half: False/0 FP32, True/1 FP16 weights, 2 the fused FP16 matrix-core inference."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], 2), np.float32)
        _check(lib().ygrid_lcf_forward(self.h, ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(y.ctypes.data),
                                       x.shape[0], int(half)))
        return y

    def train_wproject(self, gridder: Gridder, epochs=20, batch=4096, lr=2e-3, seed=1, holdout_every=8) -> LcfReport:
        o = LcfTrainOpts(epochs, batch, lr, seed, holdout_every)
        r = LcfReport()
        _check(lib().ygrid_lcf_train_wproject(self.h, gridder.h, ctypes.byref(o), ctypes.byref(r)))
        return r

    def close(self):
        if self.h:
            _check(lib().ygrid_lcf_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
