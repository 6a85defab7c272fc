"""Host-side mirror of the reference's ``ocean_redi`` module over libredi_hip.so.

``OceanRedi`` plays the role of ``type(ocean_redi_t)`` with its ``init`` /
``destroy`` type-bound procedures and the public ``redi_calc_coeffs`` /
``redi_apply_flux`` / ``redi_apply_flux_hoist`` entries
(redi/ocean_redi.F90:71-128, 466-529, 572-607, 927-1002, 1201-1262):
same argument meaning, same silent no-op guards (returned as
``REDI_SKIPPED``), same in-place update of every tracer.  ``DeviceDomain``
holds one ``RediCase`` resident in HBM: the grid (``hgrid_t``), metrics
(``ocean_metrics_t``), EOS (``ocean_eos_t``), state
(``multilayer_cgrid_state_t``), BC tags and the VarMix face diffusivities.

Torch is used only to own device memory and to name the current stream; all
compute runs in the HIP kernels of libredi_hip.so.  There is no CPU fallback:
if the library is missing the constructor raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from .inputs import RediCase

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libredi_hip.so")

REDI_OK = 0
REDI_SKIPPED = 1
REDI_NZ_MAX = 256

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


class HGrid(ctypes.Structure):
    _fields_ = [("nx_total", _I), ("ny_total", _I), ("nx_phys", _I), ("ny_phys", _I),
                ("nghost", _I), ("pad_", _I), ("dx", _D), ("dy", _D)]


class Metrics(ctypes.Structure):
    _fields_ = [(n, _P) for n in ("dy_cu", "dx_cv", "idxCu", "idyCv", "areaT", "wet_u", "wet_v")]


class Eos(ctypes.Structure):
    _fields_ = [("variant", _I), ("ts_convention", _I), ("rho0", _D), ("T_ref", _D),
                ("S_ref", _D), ("alpha_T", _D), ("beta_S", _D), ("p_ref", _D)]


class State(ctypes.Structure):
    _fields_ = [("nz_ml", _I), ("idx_temperature", _I), ("idx_salinity", _I),
                ("ntracers", _I), ("h_layer", _P), ("tracers", ctypes.POINTER(_P))]


class Bc(ctypes.Structure):
    _fields_ = [("west", _I), ("east", _I), ("south", _I), ("north", _I)]




class Coeffs(ctypes.Structure):
    _fields_ = [(n, _P) for n in ("uPoL", "uPoR", "uKoL", "uKoR", "uhEff", "vPoL", "vPoR",
                                  "vKoL", "vKoR", "vhEff", "khtr_u", "khtr_v")]


COEFF_NAMES = ("uPoL", "uPoR", "uKoL", "uKoR", "uhEff", "vPoL", "vPoR", "vKoL", "vKoR", "vhEff")

# every extern "C" symbol declared in include/redi_hip.h
EXPORTED = ("redi_init", "redi_destroy", "redi_configure", "redi_export_coeffs", "redi_set_tiles",
            "redi_state_bytes", "redi_set_recon_order", "redi_nonmonotone_faces",
            "redi_set_tracer_group",
            "redi_calc_coeffs", "redi_apply_flux", "redi_apply_flux_hoist", "redi_step",
            "redi_workspace_bytes", "redi_last_error", "redi_version", "redi_opt_launch",
            "redi_cuda_launch_", "redi_launch_nt", "redi_launch_nt_noalloc",
            "redi_timing_enable", "redi_timing_collect",
            "redi_kernel_name", "redi_init_form", "redi_calc_coeffs_fp32",
            "redi_apply_flux_fp32", "redi_step_fp32", "redi_state_bytes_form",
            "redi_set_tendency_output", "redi_set_recon_chunk", "redi_set_slope_taper",
            "redi_set_guard_output",
            "redi_set_flux_segment", "redi_set_tile_order",
            # the batched form (bound in batch.py)
            "redi_batch_create", "redi_batch_destroy", "redi_batch_domain_bytes",
            "redi_batch_add", "redi_batch_remove", "redi_batch_step", "redi_batch_step_serial",
            "redi_batch_export_coeffs", "redi_batch_info")
FORMS = {"fp64": 0, "fp32_store": 1}  # REDI_FORM_*
NKERNELS = 9  # REDI_NKERNELS


class KernelTimes(ctypes.Structure):
    _fields_ = [("ms", _D * NKERNELS), ("launches", ctypes.c_longlong * NKERNELS)]

_lib = None


class RediError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """Load libredi_hip.so (built in-tree by ``__graft_entry__.build()``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RediError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.redi_init.argtypes = [ctypes.POINTER(_P), ctypes.POINTER(HGrid), _I]
        L.redi_init_form.argtypes = [ctypes.POINTER(_P), ctypes.POINTER(HGrid), _I, _I]
        L.redi_destroy.argtypes = [_P]
        L.redi_configure.argtypes = [_P, _I, _D, _I, _D]
        L.redi_export_coeffs.argtypes = [_P, ctypes.POINTER(Coeffs), _P]
        L.redi_set_tiles.argtypes = [_P, _I, _I]
        L.redi_set_recon_order.argtypes = [_P, _I]
        L.redi_set_recon_chunk.argtypes = [_P, _I]
        L.redi_set_slope_taper.argtypes = [_P, _D]
        L.redi_set_flux_segment.argtypes = [_P, _I]
        L.redi_set_tile_order.argtypes = [_P, _I]
        L.redi_set_tendency_output.argtypes = [_P, _I, ctypes.POINTER(_P)]
        L.redi_set_guard_output.argtypes = [_P, _P, _P]
        L.redi_set_tracer_group.argtypes = [_P, _I]
        L.redi_nonmonotone_faces.argtypes = [_P, ctypes.POINTER(ctypes.c_longlong),
                                             ctypes.POINTER(ctypes.c_longlong)]
        L.redi_state_bytes.restype = ctypes.c_longlong
        L.redi_state_bytes.argtypes = [_I, _I, _I]
        L.redi_calc_coeffs.argtypes = [ctypes.POINTER(HGrid), ctypes.POINTER(Metrics),
                                       ctypes.POINTER(Eos), _P, ctypes.POINTER(State), _P]
        L.redi_apply_flux.argtypes = [ctypes.POINTER(HGrid), ctypes.POINTER(Metrics), _P,
                                      ctypes.POINTER(State), _D, _P, _P, ctypes.POINTER(Bc), _P]
        L.redi_apply_flux_hoist.argtypes = L.redi_apply_flux.argtypes
        # the FP32-store entries: redi_state32_t has redi_state_t's layout
        L.redi_calc_coeffs_fp32.argtypes = L.redi_calc_coeffs.argtypes
        L.redi_apply_flux_fp32.argtypes = L.redi_apply_flux.argtypes
        L.redi_step.argtypes = [ctypes.POINTER(HGrid), ctypes.POINTER(Metrics), ctypes.POINTER(Eos),
                                _P, ctypes.POINTER(State), _D, _P, _P, ctypes.POINTER(Bc), _I, _P]
        L.redi_step_fp32.argtypes = L.redi_step.argtypes
        L.redi_state_bytes_form.restype = ctypes.c_longlong
        L.redi_state_bytes_form.argtypes = [_I, _I, _I, _I]
        L.redi_workspace_bytes.restype = ctypes.c_longlong
        L.redi_workspace_bytes.argtypes = [_I, _I, _I]
        L.redi_last_error.restype = ctypes.c_char_p
        L.redi_launch_nt.argtypes = [_I, _I, _I, _D, _I, _I, _I, _I, _I, _I, _I,
                                     ctypes.POINTER(Eos), _P, _I, ctypes.POINTER(_P), _I, _I,
                                     _P, _P, ctypes.POINTER(Coeffs), _P, _P, _P, _P, _P, _P, _P, _P]
        L.redi_launch_nt_noalloc.argtypes = L.redi_launch_nt.argtypes
        L.redi_timing_enable.argtypes = [_P, _I]
        L.redi_timing_collect.argtypes = [_P, ctypes.POINTER(KernelTimes)]
        L.redi_kernel_name.restype = ctypes.c_char_p
        L.redi_kernel_name.argtypes = [_I]
        _lib = L
    return _lib


def last_error() -> str:
    return lib().redi_last_error().decode()


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise RediError(f"{what} failed ({rc}): {last_error()}")
    return rc


def _ptr(t) -> _P:
    return _P(t.data_ptr()) if t is not None else _P(0)


def _stream(device) -> _P:
    import torch
    return _P(torch.cuda.current_stream(device).cuda_stream)


def sanity(hTr_T) -> str:
    """The reference driver's sanity rule for a temperature field after a
    run (redi/drivers/dc_main.F90:149-153): "garbage" when its minimum is
    NaN or the field is all zero, else "ok"."""
    a = np.asarray(hTr_T)
    lo, hi = a.min(), a.max()
    if np.isnan(lo) or (lo == 0.0 and hi == 0.0):
        return "garbage"
    return "ok"


class DeviceDomain:
    """One ``RediCase`` resident in HBM plus its ctypes descriptors.  In the
    FP32-store form h and the tracers are held in FP32 (rounded from the
    case's FP64 fields); the metrics stay FP64."""

    METRIC_NAMES = ("dy_cu", "dx_cv", "idxCu", "idyCv", "areaT", "wet_u", "wet_v", "khtr_u_ext",
                    "khtr_v_ext")

    def __init__(self, case: RediCase, device: str = "cuda", form: str = "fp64"):
        import torch
        if form not in FORMS:
            raise ValueError(f"unknown form {form!r}")
        dev = torch.device(device)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        fdt = torch.float32 if form == "fp32_store" else torch.float64
        self._bind(case, form, dev, t(case.h).to(fdt), t(case.hTr).to(fdt),
                   [t(getattr(case, n)) for n in self.METRIC_NAMES])

    @classmethod
    def from_tensors(cls, case: RediCase, form: str, device, h, hTr, metrics) -> "DeviceDomain":
        """A domain over device tensors the caller filled: h (nz, ny, nx) and
        hTr (nt, nz, ny, nx) in the form's storage type, and the FP64 2-D
        arrays of METRIC_NAMES, in that order.  ``case`` supplies the
        scalars and shapes."""
        import torch
        d = cls.__new__(cls)
        d._bind(case, form, torch.device(device), h, hTr, list(metrics))
        return d

    def _bind(self, case, form, dev, h, hTr, metrics) -> None:
        self.case = case
        self.form = form
        self.device = dev
        self.h = h
        self.hTr = hTr  # (nt, nz, ny, nx)
        (self.dy_cu, self.dx_cv, self.idxCu, self.idyCv, self.areaT, self.wet_u, self.wet_v,
         self.khtr_u_ext, self.khtr_v_ext) = metrics
        self.grid = HGrid(case.nx, case.ny, case.nx_phys, case.ny_phys, case.nghost, 0, 0.0, 0.0)
        self.metrics = Metrics(*(_ptr(x).value for x in (self.dy_cu, self.dx_cv, self.idxCu,
                                                           self.idyCv, self.areaT, self.wet_u,
                                                           self.wet_v)))
        self.eos = Eos(case.eos_variant, 1, case.rho0, case.T_ref, case.S_ref, case.alpha_T,
                       case.beta_S, case.p_ref)
        self.bc = Bc(*(1 if w else 0 for w in case.walls))
        self._refresh_state()

    def _refresh_state(self):
        c = self.case
        nt = self.hTr.shape[0]
        self._trp = (_P * nt)(*(self.hTr[t].data_ptr() for t in range(nt)))
        self.state = State(c.nz, c.idx_T, c.idx_S, nt, _P(self.h.data_ptr()),
                           ctypes.cast(self._trp, ctypes.POINTER(_P)))

    def clone_tracers(self) -> "DeviceDomain":
        """A domain that shares this one's h, metrics and face diffusivities
        (a Redi step only reads them) and owns a device copy of its tracers
        (the state a step updates)."""
        import copy
        d = copy.copy(self)
        d.hTr = self.hTr.clone()
        d._refresh_state()
        return d

    def set_tracers(self, hTr) -> None:
        """Overwrite the tracers in place (host array or tensor; converted to
        the domain's storage type)."""
        import torch
        src = hTr if isinstance(hTr, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(hTr))
        self.hTr.copy_(src.to(self.hTr.dtype))

    def tracers_host(self) -> np.ndarray:
        return self.hTr.cpu().numpy()


class OceanRedi:
    """``ocean_redi_t`` on one GPU.

    The Phase-A state lives in HBM inside the library; ``coeffs()`` exports
    it in the reference layout.
    """

    def __init__(self, case_or_grid, nz: Optional[int] = None, device: str = "cuda",
                 form: str = "fp64"):
        import torch
        L = lib()
        if form not in FORMS:
            raise ValueError(f"unknown form {form!r}")
        self.form = form
        if isinstance(case_or_grid, RediCase):
            c = case_or_grid
            grid = HGrid(c.nx, c.ny, c.nx_phys, c.ny_phys, c.nghost, 0, 0.0, 0.0)
            nz = c.nz if nz is None else nz
        else:
            grid = case_or_grid
        self.device = torch.device(device)
        with torch.cuda.device(self.device):
            self.nx, self.ny = grid.nx_total, grid.ny_total
            self.nz = max(1, int(nz))
            self.ns = 2 * self.nz + 2
            h = _P()
            _check(L.redi_init_form(ctypes.byref(h), ctypes.byref(grid), int(nz), FORMS[form]),
                   "redi_init")
            self.handle = h
        self.enable = False
        self.khtr = 0.0

    # -- ocean_redi_t knobs --------------------------------------------------
    def configure(self, enable: bool = True, khtr: float = 0.0, continuous: bool = True,
                  ref_pres: float = -1.0) -> None:
        _check(lib().redi_configure(self.handle, int(bool(enable)), float(khtr),
                                    int(bool(continuous)), float(ref_pres)), "redi_configure")
        self.enable, self.khtr = bool(enable), float(khtr)

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().redi_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- public entries --------------------------------------------------------
    def _fp32(self) -> bool:
        return self.form == "fp32_store"

    def calc_coeffs(self, dom: DeviceDomain) -> int:
        """redi_calc_coeffs (redi/ocean_redi.F90:572-607)."""
        fn = lib().redi_calc_coeffs_fp32 if self._fp32() else lib().redi_calc_coeffs
        return _check(fn(ctypes.byref(dom.grid), ctypes.byref(dom.metrics), ctypes.byref(dom.eos),
                         self.handle, ctypes.byref(dom.state), _stream(self.device)),
                      "redi_calc_coeffs")

    def apply_flux(self, dom: DeviceDomain, dt: Optional[float] = None,
                   use_ext: Optional[bool] = None, bc: Optional[Bc] = None,
                   hoist: bool = False) -> int:
        """redi_apply_flux / redi_apply_flux_hoist (redi/ocean_redi.F90:927-1002,
        1201-1262).  ``use_ext`` defaults to the case's flag; ``bc`` defaults to
        the case's wall tags (pass ``False`` for "bc absent" = all walls)."""
        c = dom.case
        dt = c.dt if dt is None else dt
        use_ext = bool(c.use_ext) if use_ext is None else use_ext
        if bc is None:
            bcp = ctypes.byref(dom.bc)
        elif bc is False:
            bcp = None
        else:
            bcp = ctypes.byref(bc)
        ku = _ptr(dom.khtr_u_ext) if use_ext else _P(0)
        kv = _ptr(dom.khtr_v_ext) if use_ext else _P(0)
        if self._fp32():
            fn = lib().redi_apply_flux_fp32
        else:
            fn = lib().redi_apply_flux_hoist if hoist else lib().redi_apply_flux
        return _check(fn(ctypes.byref(dom.grid), ctypes.byref(dom.metrics), self.handle,
                         ctypes.byref(dom.state), float(dt), ku, kv, bcp,
                         _stream(self.device)), "redi_apply_flux")

    def step(self, dom: DeviceDomain, nsteps: int = 1) -> int:
        c = dom.case
        ku = _ptr(dom.khtr_u_ext) if c.use_ext else _P(0)
        kv = _ptr(dom.khtr_v_ext) if c.use_ext else _P(0)
        fn = lib().redi_step_fp32 if self._fp32() else lib().redi_step
        return _check(fn(ctypes.byref(dom.grid), ctypes.byref(dom.metrics), ctypes.byref(dom.eos),
                         self.handle, ctypes.byref(dom.state), float(c.dt), ku, kv,
                         ctypes.byref(dom.bc), int(nsteps), _stream(self.device)), "redi_step")

    def timing(self, on: bool = True) -> None:
        """Bracket every kernel launch of this object with HIP events."""
        _check(lib().redi_timing_enable(self.handle, int(bool(on))), "redi_timing_enable")

    def collect_timing(self) -> dict:
        """{kernel name: (total ms, launches)} since the last collect."""
        kt = KernelTimes()
        _check(lib().redi_timing_collect(self.handle, ctypes.byref(kt)), "redi_timing_collect")
        out = {}
        for k in range(NKERNELS):
            if kt.launches[k]:
                out[lib().redi_kernel_name(k).decode()] = (kt.ms[k], int(kt.launches[k]))
        return out

    def set_tendency_output(self, tend) -> None:
        """Write the FP64 tendency dTr / areaT of every later apply_flux into
        ``tend`` (a float64 tensor (nt, nz, ny, nx) on the device), or stop
        with ``None``."""
        if tend is None:
            self._tend_ptrs = None
            _check(lib().redi_set_tendency_output(self.handle, 0, None), "redi_set_tendency_output")
            return
        nt = tend.shape[0]
        self._tend_ptrs = (_P * nt)(*(tend[t].data_ptr() for t in range(nt)))
        _check(lib().redi_set_tendency_output(self.handle, nt,
                                              ctypes.cast(self._tend_ptrs, ctypes.POINTER(_P))),
               "redi_set_tendency_output")

    def guard_chunks(self) -> int:
        """Sublayer chunks (of 63) per face of the guard-decision output."""
        return (2 * self.nz + 1 + 62) // 63

    def set_guard_output(self, u_bits, v_bits) -> None:
        """Record the down-gradient guard's zeroed sublayers of every later
        apply_flux in ``u_bits`` / ``v_bits`` (int64 tensors (nt, nu, nch) /
        (nt, nv, nch) on the device, see redi_set_guard_output), or stop
        with ``None, None``."""
        _check(lib().redi_set_guard_output(self.handle, _ptr(u_bits), _ptr(v_bits)),
               "redi_set_guard_output")

    def set_recon_order(self, layer_major: bool) -> None:
        _check(lib().redi_set_recon_order(self.handle, int(bool(layer_major))),
               "redi_set_recon_order")

    def set_slope_taper(self, slope_max: float) -> None:
        """Opt-in neutral-slope taper of later calc_coeffs (0 = off, the
        reference's behaviour)."""
        _check(lib().redi_set_slope_taper(self.handle, float(slope_max)), "redi_set_slope_taper")

    def set_tile_order(self, vertical: bool) -> None:
        """Number sweep tiles down the rows first (schedule only, same bits)."""
        _check(lib().redi_set_tile_order(self.handle, int(bool(vertical))), "redi_set_tile_order")

    def set_flux_segment(self, cells: int) -> None:
        """Cells per run of the Phase-B flux kernels (schedule only, same bits)."""
        _check(lib().redi_set_flux_segment(self.handle, int(cells)), "redi_set_flux_segment")

    def set_recon_chunk(self, kt: int) -> None:
        """Layers per tiled-reconstruction workgroup (8, 16 or 32)."""
        _check(lib().redi_set_recon_chunk(self.handle, int(kt)), "redi_set_recon_chunk")

    def nonmonotone_faces(self) -> tuple:
        nu, nv = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().redi_nonmonotone_faces(self.handle, ctypes.byref(nu), ctypes.byref(nv)),
               "redi_nonmonotone_faces")
        return nu.value, nv.value

    def set_tracer_group(self, n: int) -> None:
        _check(lib().redi_set_tracer_group(self.handle, int(n)), "redi_set_tracer_group")

    def set_tiles(self, tx: int, ty: int) -> None:
        _check(lib().redi_set_tiles(self.handle, int(tx), int(ty)), "redi_set_tiles")

    def coeffs(self, with_khtr: bool = False) -> dict:
        """The Phase-A arrays uPoL..vhEff (and khtr_u/v) in the reference
        layout, as new tensors: (ns, ny, nx+1), (ns-1, ny, nx+1),
        (ns, ny+1, nx), (ns-1, ny+1, nx); Ko int32."""
        import torch
        nx, ny, ns, d = self.nx, self.ny, self.ns, self.device
        f64, i32 = torch.float64, torch.int32
        t = {
            "uPoL": torch.empty((ns, ny, nx + 1), dtype=f64, device=d),
            "uPoR": torch.empty((ns, ny, nx + 1), dtype=f64, device=d),
            "uKoL": torch.empty((ns, ny, nx + 1), dtype=i32, device=d),
            "uKoR": torch.empty((ns, ny, nx + 1), dtype=i32, device=d),
            "uhEff": torch.empty((ns - 1, ny, nx + 1), dtype=f64, device=d),
            "vPoL": torch.empty((ns, ny + 1, nx), dtype=f64, device=d),
            "vPoR": torch.empty((ns, ny + 1, nx), dtype=f64, device=d),
            "vKoL": torch.empty((ns, ny + 1, nx), dtype=i32, device=d),
            "vKoR": torch.empty((ns, ny + 1, nx), dtype=i32, device=d),
            "vhEff": torch.empty((ns - 1, ny + 1, nx), dtype=f64, device=d),
            "khtr_u": torch.empty((ny, nx + 1), dtype=f64, device=d),
            "khtr_v": torch.empty((ny + 1, nx), dtype=f64, device=d),
        }
        cf = Coeffs(*(t[n].data_ptr() for n in COEFF_NAMES + ("khtr_u", "khtr_v")))
        _check(lib().redi_export_coeffs(self.handle, ctypes.byref(cf), _stream(d)),
               "redi_export_coeffs")
        if not with_khtr:
            t.pop("khtr_u"); t.pop("khtr_v")
        return t


def make_redi(case: RediCase, device: str = "cuda", form: str = "fp64"):
    """(OceanRedi, DeviceDomain) configured from the case's knobs, in the
    given precision form ("fp64" or "fp32_store")."""
    dom = DeviceDomain(case, device, form)
    r = OceanRedi(case, device=device, form=form)
    r.configure(enable=bool(case.enable), khtr=case.khtr)
    return r, dom
