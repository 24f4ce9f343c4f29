"""Skyrme couplings (forces.py:7-156)."""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .reader import read_yaml

FORCE_FILE = Path(__file__).resolve().with_name("_forces.yml")


class HfbForce(ctypes.Structure):
    """Mirror of ``hfb_force`` (include/hfb_hip.h)."""
    _fields_ = [(n, ctypes.c_double) for n in
                ("b0", "b0p", "b1", "b1p", "b2", "b2p", "b3", "b3p", "b4", "b4p", "power", "slate")] + \
               [("h2m", ctypes.c_double * 2)] + \
               [(n, ctypes.c_double) for n in
                ("Crho0", "Crho1", "Crho0D", "Crho1D", "Cdrho0", "Cdrho1", "Ctau0", "Ctau1", "CdJ0", "CdJ1",
                 "v0neut", "v0prot", "rho0pr", "e2", "mass_number")] + \
               [(n, ctypes.c_int) for n in ("ipair", "ex", "zpe", "tcoul")]


@dataclass
class Forces:
    name: str
    ex: int
    zpe: int
    h2m: np.ndarray
    t0: float
    t1: float
    t2: float
    t3: float
    t4: float
    x0: float
    x1: float
    x2: float
    x3: float
    b4p: float
    power: float
    v0prot: float
    v0neut: float
    rho0pr: float
    ipair: int
    pair_reg: bool
    delta_fit: np.ndarray
    pair_cutoff: np.ndarray
    state_cutoff: np.ndarray
    softcut_range: float
    tbcs: bool
    h2ma: float = 0.0
    nucleon_mass: float = 0.0
    b0: float = 0.0
    b0p: float = 0.0
    b1: float = 0.0
    b1p: float = 0.0
    b2: float = 0.0
    b2p: float = 0.0
    b3: float = 0.0
    b3p: float = 0.0
    b4: float = 0.0
    slate: float = 0.0
    Crho0: float = 0.0
    Crho1: float = 0.0
    Crho0D: float = 0.0
    Crho1D: float = 0.0
    Cdrho0: float = 0.0
    Cdrho1: float = 0.0
    Ctau0: float = 0.0
    Ctau1: float = 0.0
    CdJ0: float = 0.0
    CdJ1: float = 0.0

    def derive(self, pi=math.pi, e2=1.4399784085965135):
        """b- and C-couplings from t, x, power (forces.py:95-122)."""
        t0, t1, t2, t3, t4 = self.t0, self.t1, self.t2, self.t3, self.t4
        x0, x1, x2, x3 = self.x0, self.x1, self.x2, self.x3
        self.b0 = t0 * (1 + 0.5 * x0)
        self.b0p = t0 * (0.5 + x0)
        self.b1 = (t1 + 0.5 * x1 * t1 + t2 + 0.5 * x2 * t2) / 4
        self.b1p = (t1 * (0.5 + x1) - t2 * (0.5 + x2)) / 4
        self.b2 = (3 * t1 * (1 + 0.5 * x1) - t2 * (1 + 0.5 * x2)) / 8
        self.b2p = (3 * t1 * (0.5 + x1) + t2 * (0.5 + x2)) / 8
        self.b3 = t3 * (1 + 0.5 * x3) / 4
        self.b3p = t3 * (0.5 + x3) / 4
        self.b4 = t4 / 2
        self.slate = (3 / pi) ** (1 / 3) * e2
        self.Crho0 = 0.5 * self.b0 - 0.25 * self.b0p
        self.Crho1 = -0.25 * self.b0p
        self.Crho0D = (1 / 3) * self.b3 - (1 / 6) * self.b3p
        self.Crho1D = -(1 / 6) * self.b3p
        self.Cdrho0 = -0.5 * self.b2 + 0.25 * self.b2p
        self.Cdrho1 = 0.25 * self.b2p
        self.Ctau0 = self.b1 - 0.5 * self.b1p
        self.Ctau1 = -0.5 * self.b1p
        self.CdJ0 = -self.b4 - 0.5 * self.b4p
        self.CdJ1 = -0.5 * self.b4p
        return self

    def to_c(self, params, mass_number) -> HfbForce:
        f = HfbForce()
        for n in ("b0", "b0p", "b1", "b1p", "b2", "b2p", "b3", "b3p", "b4", "b4p", "power", "slate",
                  "Crho0", "Crho1", "Crho0D", "Crho1D", "Cdrho0", "Cdrho1", "Ctau0", "Ctau1", "CdJ0", "CdJ1",
                  "v0neut", "v0prot", "rho0pr"):
            setattr(f, n, float(getattr(self, n)))
        f.h2m[0], f.h2m[1] = float(self.h2m[0]), float(self.h2m[1])
        f.e2 = float(params.e2)
        f.mass_number = float(mass_number)
        f.ipair, f.ex, f.zpe, f.tcoul = int(self.ipair), int(self.ex), int(self.zpe), int(bool(params.tcoul))
        return f


def init_forces(params, **kwargs) -> Forces:
    """forces.py:60-143: table lookup by name, derived couplings, pairing
    strengths by mode (5: VDI, 6: DDDI) with keyword overrides."""
    table = read_yaml(kwargs.get("force_file", FORCE_FILE))
    name = kwargs.get("name", "SLy4")
    force = table.get(name)
    if force is None:
        raise KeyError(f"Force '{name}' not found in the {Path(kwargs.get('force_file', FORCE_FILE)).name}.")
    ipair = kwargs.get("ipair", 0)
    f = Forces(name=name, ex=force["ex"], zpe=force["zpe"], h2m=np.array(force["h2m"], dtype=np.float64),
               t0=force["t0"], t1=force["t1"], t2=force["t2"], t3=force["t3"], t4=force["t4"],
               x0=force["x0"], x1=force["x1"], x2=force["x2"], x3=force["x3"], b4p=force["b4p"],
               power=force["power"], v0prot=0.0, v0neut=0.0, rho0pr=0.16, ipair=ipair,
               pair_reg=kwargs.get("pair_reg", False),
               delta_fit=np.array(kwargs.get("delta_fit", [-1.0, -1.0])),
               pair_cutoff=np.array(kwargs.get("pair_cutoff", [0.0, 0.0]), dtype=np.float64),
               state_cutoff=np.array(kwargs.get("state_cutoff", [0.0, 0.0]), dtype=np.float64),
               softcut_range=kwargs.get("softcut_range", 0.1), tbcs=kwargs.get("tbcs", False))
    for k in ("t0", "t1", "t2", "t3", "t4", "x0", "x1", "x2", "x3", "b4p", "power"):
        if k in kwargs:
            setattr(f, k, float(kwargs[k]))
    f.derive(params.pi, params.e2)
    if ipair == 5:
        f.v0prot, f.v0neut, f.rho0pr = force["vdi"]["v0prot"], force["vdi"]["v0neut"], force["vdi"]["rho0pr"]
    if ipair == 6:
        f.v0prot, f.v0neut, f.rho0pr = force["dddi"]["v0prot"], force["dddi"]["v0neut"], force["dddi"]["rho0pr"]
    for k in ("v0prot", "v0neut", "rho0pr"):
        if k in kwargs and kwargs[k] is not None:
            setattr(f, k, float(kwargs[k]))
    f.h2ma = float(0.5 * np.sum(f.h2m))
    f.nucleon_mass = params.hbc ** 2 / (2.0 * f.h2ma)
    return f
