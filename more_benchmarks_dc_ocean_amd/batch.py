"""Batched Redi steps over many independent domains of uneven size.

``RediBatch`` wraps the library's batch object (``redi_batch_*`` in
include/redi_hip.h): every domain admitted to it is advanced by
``step()`` with one hipGraph launch per Redi step, in which one persistent
kernel per kernel kind walks a device-resident work table covering every
domain.  Each domain's result is bit-identical to ``OceanRedi.step`` on that
domain alone in the same precision form; ``step_serial()`` runs the same
steps as one launch sequence per domain, for comparison.

The reference steps one domain per call (redi/drivers/dc_main.F90:124-136);
this form has no counterpart there and reuses its per-domain contract
(redi_calc_coeffs then redi_apply_flux, redi/ocean_redi.F90:572-607,
927-1002) unchanged for each domain of the batch.

Each domain's Phase-A state and step scratch are carved from one device
arena allocated when the batch is created, so admitting a domain allocates
no device memory; a domain that does not fit raises ``BudgetError``.  The
fields (h, tracers, metrics) stay in the caller's ``DeviceDomain``, which
the batch keeps alive while the domain is in it.
"""
from __future__ import annotations

import ctypes
from typing import Dict, Optional

from .redi import (Bc, COEFF_NAMES, Coeffs, DeviceDomain, Eos, FORMS, HGrid, Metrics, RediError,
                   _P, _ptr, _stream, last_error, lib)

_I = ctypes.c_int
_D = ctypes.c_double
_LL = ctypes.c_longlong

REDI_ERR_BUDGET = -6
BATCH_EXPORTED = ("redi_batch_create", "redi_batch_destroy", "redi_batch_domain_bytes",
                  "redi_batch_add", "redi_batch_remove", "redi_batch_step",
                  "redi_batch_step_serial", "redi_batch_export_coeffs", "redi_batch_info")
SWEEP_LANES = 64  # lanes of a sweep tile in the batched kernels


class BudgetError(RediError):
    """The domain does not fit the batch's arena (or every slot is taken)."""


class DomainDesc(ctypes.Structure):
    """redi_domain_t."""
    _fields_ = [("grid", HGrid), ("nz_ml", _I), ("ntracers", _I), ("idx_temperature", _I),
                ("idx_salinity", _I), ("metrics", Metrics), ("eos", Eos), ("h_layer", _P),
                ("tracers", ctypes.POINTER(_P)), ("dt", _D), ("khtr", _D), ("khtr_u_ext", _P),
                ("khtr_v_ext", _P), ("bc", Bc), ("tile_i", _I), ("tile_j", _I), ("kchunk", _I),
                ("pad_", _I), ("slope_max", _D)]


_bound = False


def blib() -> ctypes.CDLL:
    global _bound
    L = lib()
    if not _bound:
        L.redi_batch_create.argtypes = [ctypes.POINTER(_P), _I, _I, _I, _LL, _I]
        L.redi_batch_destroy.argtypes = [_P]
        L.redi_batch_domain_bytes.restype = _LL
        L.redi_batch_domain_bytes.argtypes = [_I, _I, _I, _I, _I]
        L.redi_batch_add.argtypes = [_P, ctypes.POINTER(DomainDesc), _P]
        L.redi_batch_remove.argtypes = [_P, _I]
        L.redi_batch_step.argtypes = [_P, _I, _P]
        L.redi_batch_step_serial.argtypes = [_P, _I, _P]
        L.redi_batch_export_coeffs.argtypes = [_P, _I, ctypes.POINTER(Coeffs), _P]
        L.redi_batch_info.argtypes = [_P, ctypes.POINTER(_I), ctypes.POINTER(_LL),
                                      ctypes.POINTER(_LL)]
        _bound = True
    return L


def domain_bytes(nx: int, ny: int, nz: int, form: str = "fp32_store", group: int = 4) -> int:
    """Arena bytes one domain (nx, ny: totals incl. ghosts) takes in a batch."""
    return int(blib().redi_batch_domain_bytes(nx, ny, nz, FORMS[form], group))


class RediBatch:
    """Many independent domains stepped together on one GPU."""

    def __init__(self, form: str = "fp32_store", group: int = 4, max_domains: int = 128,
                 arena_bytes: int = 1 << 30, persist: int = 1, device: str = "cuda"):
        import torch
        if form not in FORMS:
            raise ValueError(f"unknown form {form!r}")
        self.form, self.group, self.persist = form, int(group), int(persist)
        self.device = torch.device(device)
        self.domains: Dict[int, DeviceDomain] = {}
        h = _P()
        with torch.cuda.device(self.device):
            rc = blib().redi_batch_create(ctypes.byref(h), FORMS[form], self.group, int(max_domains),
                                          int(arena_bytes), self.persist)
        if rc < 0:
            raise RediError(f"redi_batch_create failed ({rc}): {last_error()}")
        self.handle = h

    def add(self, dom: DeviceDomain, shape: Optional[Dict] = None, slope_max: float = 0.0) -> int:
        """Admit ``dom`` (its form must be the batch's); returns its slot.
        ``shape``: {"tile_i", "tile_j", "kchunk"} launch shape of its kernels;
        ``slope_max``: its opt-in slope taper (0 = off)."""
        if dom.form != self.form:
            raise ValueError(f"domain form {dom.form} != batch form {self.form}")
        c = dom.case
        shape = shape or {}
        nt = dom.hTr.shape[0]
        trp = (_P * nt)(*(dom.hTr[t].data_ptr() for t in range(nt)))
        d = DomainDesc()
        d.grid = dom.grid
        d.nz_ml, d.ntracers = c.nz, nt
        d.idx_temperature, d.idx_salinity = c.idx_T, c.idx_S
        d.metrics, d.eos = dom.metrics, dom.eos
        d.h_layer = _ptr(dom.h)
        d.tracers = ctypes.cast(trp, ctypes.POINTER(_P))
        d.dt, d.khtr = float(c.dt), float(c.khtr)
        d.khtr_u_ext = _ptr(dom.khtr_u_ext) if c.use_ext else _P(0)
        d.khtr_v_ext = _ptr(dom.khtr_v_ext) if c.use_ext else _P(0)
        d.bc = dom.bc
        d.tile_i = int(shape.get("tile_i", 0))
        d.tile_j = int(shape.get("tile_j", 0))
        d.kchunk = int(shape.get("kchunk", 0))
        d.slope_max = float(slope_max)
        slot = blib().redi_batch_add(self.handle, ctypes.byref(d), _stream(self.device))
        if slot == REDI_ERR_BUDGET:
            raise BudgetError(last_error())
        if slot < 0:
            raise RediError(f"redi_batch_add failed ({slot}): {last_error()}")
        self.domains[slot] = dom
        return slot

    def remove(self, slot: int) -> DeviceDomain:
        rc = blib().redi_batch_remove(self.handle, int(slot))
        if rc < 0:
            raise RediError(f"redi_batch_remove failed ({rc}): {last_error()}")
        return self.domains.pop(slot)

    def step(self, nsteps: int = 1) -> None:
        """``nsteps`` Redi steps of every domain (one graph launch each)."""
        rc = blib().redi_batch_step(self.handle, int(nsteps), _stream(self.device))
        if rc < 0:
            raise RediError(f"redi_batch_step failed ({rc}): {last_error()}")

    def step_serial(self, nsteps: int = 1) -> None:
        """The same steps, one launch sequence per domain."""
        rc = blib().redi_batch_step_serial(self.handle, int(nsteps), _stream(self.device))
        if rc < 0:
            raise RediError(f"redi_batch_step_serial failed ({rc}): {last_error()}")

    def coeffs(self, slot: int) -> dict:
        """The slot's Phase-A state in the reference layout (device tensors)."""
        import torch
        dom = self.domains[slot]
        nx, ny, ns = dom.case.nx, dom.case.ny, dom.case.ns
        dev = self.device
        shp = {"u": (ns, ny, nx + 1), "v": (ns, ny + 1, nx)}
        out = {}
        for n in COEFF_NAMES:
            g = n[0]
            shape = shp[g] if not n.endswith("hEff") else (ns - 1,) + shp[g][1:]
            dt = torch.int32 if "Ko" in n else torch.float64
            out[n] = torch.empty(shape, dtype=dt, device=dev)
        c = Coeffs(*([_ptr(out[n]) for n in ("uPoL", "uPoR", "uKoL", "uKoR", "uhEff", "vPoL", "vPoR",
                                              "vKoL", "vKoR", "vhEff")] + [_P(0), _P(0)]))
        rc = blib().redi_batch_export_coeffs(self.handle, int(slot), ctypes.byref(c),
                                             _stream(self.device))
        if rc < 0:
            raise RediError(f"redi_batch_export_coeffs failed ({rc}): {last_error()}")
        return out

    def info(self) -> Dict[str, int]:
        n, used, free = _I(), _LL(), _LL()
        blib().redi_batch_info(self.handle, ctypes.byref(n), ctypes.byref(used), ctypes.byref(free))
        return {"domains": n.value, "used_bytes": used.value, "largest_free_bytes": free.value}

    def close(self) -> None:
        if getattr(self, "handle", None):
            blib().redi_batch_destroy(self.handle)
            self.handle = None
        self.domains = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
