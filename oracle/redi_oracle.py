"""Python access to the oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only to check results or as the timed CPU baseline; the
product path never touches it.

* ``OracleRedi`` mirrors ``ocean_redi_t`` + ``redi_calc_coeffs`` +
  ``redi_apply_flux`` (redi/ocean_redi.F90:71-128, 572-607, 927-1002) on top
  of the C restatement in ``liboracle.so``: the host-side guards, the
  ``use_ext`` face copy, the per-tracer snapshot and the two whole-field
  phases.
* ``run_reference`` runs the compiled reference (``_ref/redi_ref``, built by
  ``make ref`` from the reference's own Fortran sources) on a case and returns
  its "REDO" outputs.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "redi_ref")

_lib = None


class OracleEos(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int), ("pad_", ctypes.c_int),
                ("rho0", ctypes.c_double), ("T_ref", ctypes.c_double),
                ("S_ref", ctypes.c_double), ("alpha_T", ctypes.c_double),
                ("beta_S", ctypes.c_double), ("p_ref", ctypes.c_double)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        D, I = ctypes.c_double, ctypes.c_int
        P = ctypes.c_void_p
        _lib.oracle_fv_diff.restype = D
        _lib.oracle_fv_diff.argtypes = [D] * 6
        _lib.oracle_ppm_edge.restype = D
        _lib.oracle_ppm_edge.argtypes = [D] * 9
        _lib.oracle_interpolate_position.restype = D
        _lib.oracle_interpolate_position.argtypes = [D] * 4
        _lib.oracle_ppm_ave.restype = D
        _lib.oracle_ppm_ave.argtypes = [D] * 5
        _lib.oracle_eos_density_point.restype = D
        _lib.oracle_eos_density_point.argtypes = [ctypes.POINTER(OracleEos), D, D, D]
        _lib.oracle_eos_specvol_derivs.restype = None
        _lib.oracle_eos_specvol_derivs.argtypes = [ctypes.POINTER(OracleEos), D, D, D, P, P]
        _lib.oracle_roquet_spv_point.restype = None
        _lib.oracle_roquet_spv_point.argtypes = [D, D, D, P, P, P]
        _lib.oracle_plm_diff.argtypes = [I, P, P, P]
        _lib.oracle_interface_scalar.argtypes = [I, P, P, P, D]
        _lib.oracle_neutral_positions_continuous.argtypes = [I] + [P] * 15
        _lib.oracle_build_column.argtypes = [I, P, P, P, ctypes.POINTER(OracleEos), P, P, P, P, P]
        _lib.oracle_face_coeffs.argtypes = [I, P, P, P, P, P, P, ctypes.POINTER(OracleEos), P, P, P, P, P]
        _lib.oracle_tracer_column.argtypes = [I, P, P, P, P, P, P]
        _lib.oracle_sublayer_dT.restype = D
        _lib.oracle_sublayer_dT.argtypes = [I, I, I, I, D, D, D, D] + [P] * 8
        _lib.oracle_calc_coeffs.argtypes = [I, I, I, ctypes.POINTER(OracleEos)] + [P] * 15
        _lib.oracle_calc_coeffs_taper.argtypes = ([I, I, I, ctypes.POINTER(OracleEos)] + [P] * 7
                                                  + [D] + [P] * 10)
        _lib.oracle_face_coeffs_taper.argtypes = ([I, P, P, P, P, P, P, ctypes.POINTER(OracleEos),
                                                   D, D] + [P] * 5)
        _lib.oracle_apply_flux_tracer.argtypes = [I, I, I, D, I, I, I, I, I, I, I] + [P] * 20
        _lib.oracle_set_threads.restype = I
        _lib.oracle_set_threads.argtypes = [I]
    return _lib


def ptr(a: np.ndarray) -> ctypes.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


def eos_of(case) -> OracleEos:
    return OracleEos(case.eos_variant, 0, case.rho0, case.T_ref, case.S_ref,
                     case.alpha_T, case.beta_S, case.p_ref)


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))


class OracleRedi:
    """``ocean_redi_t`` on the host, driven by the C restatement.

    Coefficient arrays have the reference shapes, stored C-contiguous as
    ``(ns, ny, nx+1)`` etc. (= Fortran ``(nx+1, ny, ns)``), initialised like
    ``ocean_redi_init`` (Po = hEff = 0, Ko = 1; redi/ocean_redi.F90:493-505).
    """

    def __init__(self, case, slope_max: float = 0.0):
        self.case = case
        # opt-in slope taper, not part of the reference (0 = the plain restatement)
        self.slope_max = float(slope_max)
        nx, ny, ns = case.nx, case.ny, case.ns
        self.uPoL = np.zeros((ns, ny, nx + 1)); self.uPoR = np.zeros((ns, ny, nx + 1))
        self.uKoL = np.ones((ns, ny, nx + 1), np.int32); self.uKoR = np.ones((ns, ny, nx + 1), np.int32)
        self.uhEff = np.zeros((ns - 1, ny, nx + 1))
        self.vPoL = np.zeros((ns, ny + 1, nx)); self.vPoR = np.zeros((ns, ny + 1, nx))
        self.vKoL = np.ones((ns, ny + 1, nx), np.int32); self.vKoR = np.ones((ns, ny + 1, nx), np.int32)
        self.vhEff = np.zeros((ns - 1, ny + 1, nx))
        self.khtr_u = np.zeros((ny, nx + 1)); self.khtr_v = np.zeros((ny + 1, nx))

    def coeffs(self) -> dict:
        return {k: getattr(self, k) for k in ("uPoL", "uPoR", "uKoL", "uKoR", "uhEff",
                                              "vPoL", "vPoR", "vKoL", "vKoR", "vhEff")}

    def calc_coeffs(self, h: np.ndarray, hTr: np.ndarray) -> None:
        """redi_calc_coeffs (redi/ocean_redi.F90:572-607)."""
        c = self.case
        if not c.enable or c.idx_T <= 0 or c.idx_S <= 0:
            return
        e = eos_of(c)
        if self.slope_max > 0.0:
            lib().oracle_calc_coeffs_taper(
                c.nx, c.ny, c.nz, ctypes.byref(e), ptr(h), ptr(hTr[c.idx_T - 1]),
                ptr(hTr[c.idx_S - 1]), ptr(c.wet_u), ptr(c.wet_v), ptr(c.idxCu), ptr(c.idyCv),
                self.slope_max,
                ptr(self.uPoL), ptr(self.uPoR), ptr(self.uKoL), ptr(self.uKoR), ptr(self.uhEff),
                ptr(self.vPoL), ptr(self.vPoR), ptr(self.vKoL), ptr(self.vKoR), ptr(self.vhEff))
            return
        lib().oracle_calc_coeffs(
            c.nx, c.ny, c.nz, ctypes.byref(e), ptr(h), ptr(hTr[c.idx_T - 1]),
            ptr(hTr[c.idx_S - 1]), ptr(c.wet_u), ptr(c.wet_v),
            ptr(self.uPoL), ptr(self.uPoR), ptr(self.uKoL), ptr(self.uKoR), ptr(self.uhEff),
            ptr(self.vPoL), ptr(self.vPoR), ptr(self.vKoL), ptr(self.vKoR), ptr(self.vhEff))

    def apply_flux(self, h: np.ndarray, hTr: np.ndarray) -> None:
        """redi_apply_flux (redi/ocean_redi.F90:927-1002); updates hTr."""
        c = self.case
        if not c.enable:
            return
        if not c.use_ext and c.khtr <= 0.0:
            return
        if c.use_ext:
            self.khtr_u[...] = c.khtr_u_ext
            self.khtr_v[...] = c.khtr_v_ext
        else:
            self.khtr_u[...] = c.khtr
            self.khtr_v[...] = c.khtr
        ww, we, ws, wn = (1 if w else 0 for w in c.walls)
        for t in range(hTr.shape[0]):
            snap = hTr[t].copy()
            lib().oracle_apply_flux_tracer(
                c.nx, c.ny, c.nz, c.dt, c.nghost, c.nx_phys, c.ny_phys, ww, we, ws, wn,
                ptr(self.khtr_u), ptr(self.khtr_v), ptr(c.dy_cu), ptr(c.dx_cv),
                ptr(c.idxCu), ptr(c.idyCv), ptr(c.areaT), ptr(h), ptr(snap), ptr(hTr[t]),
                ptr(self.uPoL), ptr(self.uPoR), ptr(self.uKoL), ptr(self.uKoR), ptr(self.uhEff),
                ptr(self.vPoL), ptr(self.vPoR), ptr(self.vKoL), ptr(self.vKoR), ptr(self.vhEff))

    def step(self, h: np.ndarray, hTr: np.ndarray) -> None:
        self.calc_coeffs(h, hTr)
        self.apply_flux(h, hTr)


def run_oracle(case, nsteps: int | None = None, slope_max: float = 0.0) -> dict:
    """Run the C restatement from a cold start; outputs mirror read_output."""
    n = case.nsteps if nsteps is None else nsteps
    o = OracleRedi(case, slope_max)
    h = np.ascontiguousarray(case.h)
    hTr = np.ascontiguousarray(case.hTr.copy())
    out = {}
    for s in range(n):
        o.step(h, hTr)
        if s == 0:
            out.update({k: v.copy() for k, v in o.coeffs().items()})
            out["hTr1"] = hTr.copy()
    out["hTrN"] = hTr
    return out


def have_reference() -> bool:
    return os.path.exists(REF_BIN)


def build_reference() -> bool:
    """Compile the reference's Fortran sources into _ref/ (needs the
    reference tree; returns False where it is absent)."""
    if not os.path.isdir("/root/reference/redi"):
        return have_reference()
    subprocess.run(["make", "-s", "-C", HERE, "ref", "dropin-ref"], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    # the reference's own driver on both module sets (needs the drop-in
    # modules built first)
    subprocess.run(["make", "-s", "-C", HERE, "dc-main"], check=False,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return have_reference()


DROPIN_REF_BIN = os.path.join(HERE, "_ref", "redi_dropin_ref")
DC_MAIN_REF_BIN = os.path.join(HERE, "_ref", "redi_dc_main_ref")
DC_MAIN_HIP_BIN = os.path.join(HERE, "_ref", "redi_dc_main_hip")


def run_reference(case, nsteps: int | None = None, mode: int = 0,
                  threads: int | None = None, nwarm: int = 0,
                  want_output: bool = True):
    """Run the compiled reference on ``case``.  Returns (outputs, stdout)."""
    import dataclasses
    from more_benchmarks_dc_ocean_amd.inputs import write_input, read_output
    if not have_reference():
        raise FileNotFoundError(REF_BIN)
    if nsteps is not None:
        case = dataclasses.replace(case, nsteps=nsteps)
    env = dict(os.environ)
    if threads is not None:
        env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin") if want_output else "-"
        write_input(fin, case)
        p = subprocess.run([REF_BIN, fin, fout, str(mode), str(nwarm)], env=env,
                           capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError(f"redi_ref failed ({p.returncode}): {p.stdout}\n{p.stderr}")
        out = read_output(fout) if want_output else None
    return out, p.stdout
