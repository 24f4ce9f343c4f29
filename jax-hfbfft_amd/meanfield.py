"""Mean field: h psi and the Skyrme potentials (meanfield.py:7-391).

``Meanfield`` keeps the reference's fields as float64 device tensors.  When
``f32`` is set, ``skyrme`` also writes float32 copies that the FP32 spinor
kernels read (the mixed-precision path); FP64 spinor blocks always use the
float64 fields."""
from __future__ import annotations

from dataclasses import dataclass, field

import torch

from . import _lib, kernels


@dataclass
class Meanfield:
    upot: torch.Tensor
    bmass: torch.Tensor
    divaq_buf: torch.Tensor
    v_pair: torch.Tensor
    aq: torch.Tensor
    spot: torch.Tensor
    wlspot: torch.Tensor
    dbmass_buf: torch.Tensor
    ecorrp_dev: torch.Tensor            # device scalar view (E vector slot or own storage)
    f32: dict = field(default_factory=dict)
    derivs_ctx: object = None           # set by skyrme: divaq/dbmass are due for this context
    vpair_max: torch.Tensor = None      # (2,) device maxima of v_pair per isospin, written by skyrme

    @property
    def ecorrp(self) -> float:
        return float(self.ecorrp_dev.item())

    # divaq = div aq and dbmass = grad bmass (meanfield.py:329-344) are part of
    # the reference's Meanfield, but hpsi01 reads neither; they are computed
    # from the current aq and bmass when first read after a skyrme call, so
    # the SCF loop does not pay for them.
    def _refresh_derivs(self):
        if self.derivs_ctx is not None:
            ctx, self.derivs_ctx = self.derivs_ctx, None
            kernels.meanfield_derivs(ctx, self)

    @property
    def divaq(self) -> torch.Tensor:
        self._refresh_derivs()
        return self.divaq_buf

    @property
    def dbmass(self) -> torch.Tensor:
        self._refresh_derivs()
        return self.dbmass_buf

    def fields(self, dtype):
        """(upot, bmass, aq, spot, wlspot, v_pair) in the precision of a spinor dtype."""
        if dtype == torch.complex128:
            return self.upot, self.bmass, self.aq, self.spot, self.wlspot, self.v_pair
        g = self.f32
        if not g:
            raise RuntimeError("FP32 spinors need the float32 potentials: call init_meanfield(..., f32=True)")
        return g["upot"], g["bmass"], g["aq"], g["spot"], g["wlspot"], g["v_pair"]


def init_meanfield(grids, device=None, f32=False) -> Meanfield:
    """meanfield.py:23-38 (zeros), plus float32 copies when ``f32``."""
    dev = torch.device("cuda") if device is None else torch.device(device)
    s4 = (2, grids.nx, grids.ny, grids.nz)
    s5 = (2, 3, grids.nx, grids.ny, grids.nz)
    z = lambda s, t=torch.float64: torch.zeros(s, dtype=t, device=dev)
    mf = Meanfield(z(s4), z(s4), z(s4), z(s4), z(s5), z(s5), z(s5), z(s5), z(()))
    if f32:
        mf.f32 = {"upot": z(s4, torch.float32), "bmass": z(s4, torch.float32), "v_pair": z(s4, torch.float32),
                  "aq": z(s5, torch.float32), "spot": z(s5, torch.float32), "wlspot": z(s5, torch.float32)}
    return mf


def hpsi01(grids, meanfield, iq, weight, weightuv, pinn):
    """Single-particle Hamiltonian on a two-spinor (2,nx,ny,nz) or a block of
    them (S,2,nx,ny,nz), all of isospin iq (meanfield.py:40-159).
    Returns (weight*h psi - weightuv*v_pair psi, h psi, v_pair psi)."""
    single = pinn.dim() == 4
    x = (pinn.unsqueeze(0) if single else pinn).contiguous()
    S = x.shape[0]
    dev = x.device
    w = torch.as_tensor(weight, dtype=torch.float64, device=dev).reshape(-1).expand(S).contiguous()
    wuv = torch.as_tensor(weightuv, dtype=torch.float64, device=dev).reshape(-1).expand(S).contiguous()
    npsi0 = S if int(iq) == 0 else 0
    pout, pmf, pdel, _ = kernels.hpsi(grids.ctx, x, npsi0, *meanfield.fields(x.dtype), w, wuv, want_partial=False)
    if single:
        return pout[0], pmf[0], pdel[0]
    return pout, pmf, pdel


@dataclass
class DensityDerivs:
    """Derived density fields shared by skyrme and integ_energy."""
    divj: torch.Tensor
    lap: torch.Tensor
    grad: torch.Tensor
    curls: torch.Tensor
    curlj: torch.Tensor

    def __getitem__(self, k):
        return getattr(self, k)


def density_derivs(grids, densities, out=None, coulomb=None, e2=None) -> DensityDerivs:
    """The derived density fields; with ``coulomb`` also its Coulomb
    potential of these densities into coulomb.wcoul (coulomb.py:54-75), in
    one call that shares the launches."""
    ctx = grids.ctx
    if out is None:
        dev = densities.rho.device
        s4 = (2, grids.nx, grids.ny, grids.nz)
        s5 = (2, 3, grids.nx, grids.ny, grids.nz)
        out = DensityDerivs(*(torch.empty(s, dtype=torch.float64, device=dev) for s in (s4, s4, s5, s5, s5)))
    if coulomb is None:
        kernels.density_derivs(ctx, densities, out)
        return out
    if coulomb.wcoul is None:
        coulomb.wcoul = torch.empty((grids.nx, grids.ny, grids.nz), dtype=torch.float64, device=densities.rho.device)
    kernels.density_derivs_coulomb(ctx, densities, out, e2, coulomb.wcoul)
    return out


def skyrme(coulomb, densities, forces, grids, meanfield, params, static, constraints=None, derivs=None, E=None,
           mass_number=None, wcoul_ready=False, finish=True, with_energy=False):
    """Potentials from the densities (meanfield.py:163-391); returns (meanfield, coulomb).

    ``derivs`` may carry precomputed DensityDerivs of these densities; ``E``
    is a device energy vector whose ecorrp slot is written (a private one is
    used otherwise); ``wcoul_ready``: coulomb.wcoul already holds the
    Coulomb potential of these densities; ``finish=False`` leaves ecorrp
    and the v_pair maxima to a later kernels.energy_tail; ``with_energy``
    (with finish=False) also leaves the partials of integ_energy for it,
    from the same launch."""
    if static is not None and getattr(static, "outertype", "N") != "N":
        raise NotImplementedError(f"outertype '{static.outertype}' not implemented.")
    ctx = grids.ctx
    if derivs is None:
        derivs = density_derivs(grids, densities)
    from .coulomb import poisson
    if params.tcoul and not wcoul_ready:
        coulomb.wcoul = poisson(grids, params, coulomb, densities.rho, out=coulomb.wcoul)
    if E is None:
        E = torch.zeros(kernels.EN_LEN, dtype=torch.float64, device=densities.rho.device)
    A = mass_number if mass_number is not None else 0.0
    fc = forces.to_c(params, A)
    if with_energy:
        if finish:
            raise ValueError("with_energy leaves the finishing step to kernels.energy_tail (finish=False)")
        kernels.skyrme_energy(ctx, fc, densities, derivs, coulomb.wcoul, meanfield, meanfield.f32 or None)
    else:
        kernels.skyrme(ctx, fc, densities, derivs, coulomb.wcoul, meanfield, E, meanfield.f32 or None, finish=finish)
    if constraints is not None and constraints.tconstraint:  # meanfield.py:246-248
        from .constraint import add_constraint_potential
        add_constraint_potential(meanfield.upot, constraints, (meanfield.f32 or {}).get("upot"))
    meanfield.derivs_ctx = ctx
    meanfield.ecorrp_dev = E[kernels.EN["ECORRP"]]
    meanfield.vpair_max = E[kernels.EN["VMAX"]:kernels.EN["VMAX"] + 2]
    return meanfield, coulomb
