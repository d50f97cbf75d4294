"""Shared helpers for parity tests: run the same case through libostc (GPU)
and the FP64 oracle, stage by stage, and compare fields."""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle_ffi as O
from openstc_amd import _lib as L
from openstc_amd.runtime import Simulation

# oracle array name for each GPU field
ORC_NAME = {
    "BV": "bv", "SV": "sv", "CV": "cv", "VEL": "vel", "SOUND": "acoustic_speed", "MACH": "mach", "MUL": "mul",
    "LAMBDA": "thermal_conductivity", "RHO_D": "rho_D", "GAMMA": "gamma", "CP": "cp", "MUT": "mut",
    "LAMBDA_T": "turb_therm_cond", "TURB_JAC": "turb_src_jac", "DQ": "dq", "SPECTR": "inv_spectr_rad",
    "DT": "dt_local", "WALL_DIST": "wall_distance", "JAC": "jac", "METRIC": "metric", "BV_LAST": "bv_last",
    "CHEM_JAC": "chem_src_jac",
}


def interior(a, g=2, halo=0):
    s = slice(g - halo, a.shape[-1] - g + halo)
    t = slice(g - halo, a.shape[-2] - g + halo)
    u = slice(g - halo, a.shape[-3] - g + halo)
    return a[..., u, t, s]


def scaled_err(gpu: np.ndarray, orc: np.ndarray) -> float:
    """max |gpu-orc| / max |orc| over each plane, worst plane."""
    gpu = np.atleast_1d(gpu)
    orc = np.atleast_1d(orc)
    worst = 0.0
    for l in range(gpu.shape[0]):
        a, b = gpu[l], orc[l]
        scale = np.max(np.abs(b))
        d = np.max(np.abs(a - b))
        if scale == 0:
            e = d
        else:
            e = d / scale
        if not np.isfinite(e):
            return float("inf")
        worst = max(worst, float(e))
    return worst


class Pair:
    """A GPU Simulation and an OracleSimulation of the same case."""

    def __init__(self, case, precision=None):
        self.case = case
        self.gpu = Simulation(case, precision)
        self.orc = O.OracleSimulation(case)
        self.lib = self.gpu.ctx.lib
        self.c = self.gpu.ctx.h
        self.p = C.byref(self.orc.param)

    def init(self):
        self.gpu.initialize_computation()
        self.orc.initialize_computation()

    def zg(self, b=0):
        return self.gpu.zones[b]

    def zo(self, b=0):
        return C.byref(self.orc.zones[b])

    def oarr(self, field, b=0):
        n = self.zg(b).planes(field)
        return self.orc.blocks[b].arr[ORC_NAME[field]][:n]

    def err(self, field, b=0, halo=0):
        g = self.case.ngg
        a = self.zg(b).download(field)
        o = self.oarr(field, b)
        return scaled_err(interior(a, g, halo), interior(o, g, halo))

    def sync_to_oracle(self, fields, b=0):
        """Copy GPU state into the oracle so a stage compares only its own arithmetic."""
        for f in fields:
            n = self.zg(b).planes(f)
            if n:
                self.orc.blocks[b].arr[ORC_NAME[f]][:n] = self.zg(b).download(f)

    def close(self):
        self.gpu.close()
