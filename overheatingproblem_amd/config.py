"""Host-side mirror of the step configuration (``rpfvm_config`` in include/rpfvm.h).

The fields map one-to-one onto the reference's ControlFile keys and process
globals (src/RP_FVM.c:66-84, 434-529; include/Golbal.h:55-70;
include/Boundary_Condition.h:49; include/initialize.h:1298-1301).
"""
from __future__ import annotations

import ctypes

GHOST = 4

BC_OUTFLOW, BC_REFLECTION, BC_PERIODICITY, BC_INFLOW, BC_FIXED_VALUE = range(5)
BC_DMR_WALL, BC_DMR_SHOCK = 5, 6
BC_NAMES = {"outflow": BC_OUTFLOW, "reflection": BC_REFLECTION, "periodicity": BC_PERIODICITY,
            "inflow": BC_INFLOW, "fixed": BC_FIXED_VALUE}

SCHEME_HLL, SCHEME_HLLC, SCHEME_ROE, SCHEME_LLF = 1, 2, 3, 10
SCHEME_HLLHC, SCHEME_HLLCHC, SCHEME_ROEHC = 11, 22, 33   # not in mixed batches
# Solver names used in output file names (src/RP_FVM.c:785-793).
SCHEME_NAMES = {1: "HLL", 2: "HLLC", 3: "Roe", 11: "HLLHC", 22: "HLLCHC", 33: "RoeHC"}

PRECISION_FP64, PRECISION_MIXED = 0, 1

FAULT_HLL_BAD_STATE, FAULT_NONFINITE_DT, FAULT_MIXED_FALLBACK, FAULT_POSITIVITY = 1, 2, 4, 8
DEBUG_MIXED_ALL_FP64 = 1
DEBUG_LARGE_ARRAY_FORMS = 2
DEBUG_TWO_SWEEP = 4
DEBUG_ONE_PASS = 8
DEBUG_MARCH_FORMS = 16

# The reference's default inflow / fixed-value states: a Mach-3 stream of
# rho = 1.4, u = 3, p = 1 in conserved form (include/Boundary_Condition.h:49,
# include/initialize.h:1298-1301).
_MACH3 = (1.4, 1.4 * 3.0, 0.0, 1.0 / 0.4 + 0.5 * 1.4 * 3.0 * 3.0)


class Config(ctypes.Structure):
    """ctypes image of ``struct rpfvm_config``."""

    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("time_order", ctypes.c_int32),
        ("scheme", ctypes.c_int32),
        ("recon_order", ctypes.c_int32),
        ("characteristic", ctypes.c_int32),
        ("source", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("positivity", ctypes.c_int32),
        ("debug_flags", ctypes.c_int32),
        ("bc", ctypes.c_int32 * 4),
        ("dx", ctypes.c_double),
        ("dy", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("gravity", ctypes.c_double),
        ("inflow_state", ctypes.c_double * 4),
        ("fixed_state", (ctypes.c_double * 4) * 4),
        ("lx", ctypes.c_double),
        ("bc_par", (ctypes.c_double * 4) * 4),
        ("bc_alt_state", (ctypes.c_double * 4) * 4),
    ]

    @property
    def R(self) -> int:
        return self.nx + 2 * GHOST

    @property
    def C(self) -> int:
        return self.ny + 2 * GHOST

    @property
    def shape(self) -> tuple:
        return (4, self.R, self.C)

    def copy(self) -> "Config":
        c = Config()
        ctypes.pointer(c)[0] = self
        return c

    def as_dict(self) -> dict:
        return {
            "nx": self.nx, "ny": self.ny, "time_order": self.time_order, "scheme": self.scheme,
            "recon_order": self.recon_order, "characteristic": self.characteristic,
            "source": self.source, "precision": self.precision,
            "positivity": self.positivity, "bc": list(self.bc),
            "dx": self.dx, "dy": self.dy, "gamma": self.gamma, "cfl": self.cfl,
            "gravity": self.gravity, "inflow_state": list(self.inflow_state),
            "fixed_state": [list(r) for r in self.fixed_state],
            "lx": self.lx, "bc_par": [list(r) for r in self.bc_par],
            "bc_alt_state": [list(r) for r in self.bc_alt_state],
        }


def make_config(nx: int, ny: int, *, dx: float | None = None, dy: float | None = None,
                lx: float = 1.0, ly: float = 1.0, time_order: int = 3, scheme: int = SCHEME_HLLC,
                recon_order: int = 5, characteristic: bool = False, source: bool = False,
                precision: int = PRECISION_FP64, positivity: bool = False, bc=(0, 0, 0, 0),
                gamma: float = 1.4,
                cfl: float = 0.5, gravity: float = 1.0, inflow_state=_MACH3,
                fixed_state=(_MACH3, _MACH3, _MACH3, _MACH3)) -> Config:
    """Build a Config.  Defaults are the reference's effective ones (RK3, WENO5
    component-wise, include/Golbal.h:58,60) with HLLC and outflow walls."""
    c = Config()
    c.nx, c.ny = int(nx), int(ny)
    c.time_order, c.scheme, c.recon_order = int(time_order), int(scheme), int(recon_order)
    c.characteristic, c.source, c.precision = int(bool(characteristic)), int(bool(source)), int(precision)
    c.positivity = int(bool(positivity))
    for k in range(4):
        c.bc[k] = int(bc[k])
    c.dx = float(dx) if dx is not None else lx / nx
    c.dy = float(dy) if dy is not None else ly / ny
    c.gamma, c.cfl, c.gravity = float(gamma), float(cfl), float(gravity)
    for k in range(4):
        c.inflow_state[k] = float(inflow_state[k])
        for s in range(4):
            c.fixed_state[s][k] = float(fixed_state[s][k])
    return c


def config_from_dict(d: dict) -> Config:
    c = make_config(d["nx"], d["ny"], dx=d["dx"], dy=d["dy"], time_order=d["time_order"],
                    scheme=d["scheme"], recon_order=d["recon_order"],
                    characteristic=d["characteristic"], source=d["source"],
                    precision=d.get("precision", 0), positivity=d.get("positivity", 0),
                    bc=d["bc"], gamma=d["gamma"], cfl=d["cfl"],
                    gravity=d["gravity"], inflow_state=d["inflow_state"],
                    fixed_state=d["fixed_state"])
    c.lx = float(d.get("lx", 0.0))
    for s in range(4):
        for k in range(4):
            c.bc_par[s][k] = float(d.get("bc_par", [[0.0] * 4] * 4)[s][k])
            c.bc_alt_state[s][k] = float(d.get("bc_alt_state", [[0.0] * 4] * 4)[s][k])
    return c
