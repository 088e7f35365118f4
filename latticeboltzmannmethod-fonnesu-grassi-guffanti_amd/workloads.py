"""The lattice configurations the engine is measured and checked on.

Each builder returns a :class:`Workload`: a node-class plane from the
library's own geometry helpers, the collision and boundary setup, and the
prescribed inlet/outlet values in list order (row-major, as the
reference's node lists).  Sizes are parameters so that the same setups run
at full size on the GPU and scaled down where the CPU oracle checks them.

  cavity   lid-driven cavity, BGK, lid u = (0.1, 0), Re 100 (129 x 129)
  channel  channel with a cylinder, TRT magic 3/16, parabolic Zou-He
           inlet u_max = 0.04, Zou-He pressure outlet rho = 1, Re 200
           (512 x 2048, cylinder r 26 at row 260, column 410)
  porous   random discs r in [4, 16] to 30 % solid, BGK tau 0.8, Zou-He
           pressure inlet rho 1.001 / outlet rho 1.000 (8192 x 8192)
  cavity_trt  lid-driven cavity, TRT magic 3/16, u 0.1, Re 1000 (4096 x 4096)
  ensemble 512 independent porous channels, rows/columns in {128, ..., 1024},
           porosity in [0.5, 0.9], discs r in [3, 12], the porous set-up
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

import lbm_amd as L

MAGIC = 3.0 / 16.0


@dataclasses.dataclass
class Workload:
    name: str
    node_class: np.ndarray          # (NC, NR) uint8
    collision: int
    tau: float = 1.0
    tau_even: float = 1.0
    tau_odd: float = 1.0
    inlet_policy: int = L.ZOUHE_VELOCITY
    outlet_policy: int = L.OPEN_NOOP
    inlet: dict = dataclasses.field(default_factory=dict)    # ux / uy / rho arrays, inlet list order
    outlet: dict = dataclasses.field(default_factory=dict)
    steps: int = 1000
    info: dict = dataclasses.field(default_factory=dict)

    @property
    def nr(self):
        return self.node_class.shape[1]

    @property
    def nc(self):
        return self.node_class.shape[0]

    def wet_count(self):
        c = self.node_class
        return int(np.count_nonzero((c == L.FLUID) | (c == L.INLET) | (c == L.OUTLET)))

    def lattice_kwargs(self):
        return dict(collision=self.collision, tau=self.tau, tau_even=self.tau_even, tau_odd=self.tau_odd,
                    inlet_policy=self.inlet_policy, outlet_policy=self.outlet_policy)

    def make_lattice(self, storage=L.F64, **kw):
        lat = L.Lattice(self.node_class, storage=storage, **self.lattice_kwargs(), **kw)
        if self.inlet:
            lat.set_open_values(0, **self.inlet)
        if self.outlet:
            lat.set_open_values(1, **self.outlet)
        lat.init_equilibrium()
        return lat


def trt_taus(nu, magic=MAGIC):
    """tau+ from the viscosity (nu = (tau+ - 1/2)/3) and tau- from the magic
    parameter, TRTCollision.hpp:101-104,128-134."""
    te = 3.0 * nu + 0.5
    to = magic * (1.0 / (te - 0.5)) + 0.5
    return te, to


def open_rows(node_class, which):
    """Row indices of the inlet (0) / outlet (1) nodes in list order."""
    cls = node_class
    code = L.INLET if which == 0 else L.OUTLET
    jj, ii = np.nonzero(cls == code)            # (NC, NR) -> column, row
    order = np.lexsort((jj, ii))                 # row-major: row first, then column
    return ii[order], jj[order]


def cavity(n=129, u_lid=0.1, re=100.0):
    cls = L.geometry_cavity(n, n)
    nu = u_lid * (n - 2) / re
    ni = len(open_rows(cls, 0)[0])
    return Workload(name=f"cavity{n}", node_class=cls, collision=L.BGK, tau=3.0 * nu + 0.5,
                    inlet_policy=L.ZOUHE_VELOCITY, outlet_policy=L.OPEN_NOOP,
                    inlet=dict(ux=np.full(ni, u_lid), uy=np.zeros(ni)), steps=20000,
                    info=dict(re=re, u_lid=u_lid, nu=nu))


def cavity_trt(n=4096, u_lid=0.1, re=1000.0, magic=MAGIC):
    cls = L.geometry_cavity(n, n)
    nu = u_lid * (n - 2) / re
    te, to = trt_taus(nu, magic)
    ni = len(open_rows(cls, 0)[0])
    return Workload(name=f"cavity_trt{n}", node_class=cls, collision=L.TRT, tau_even=te, tau_odd=to,
                    inlet_policy=L.ZOUHE_VELOCITY, outlet_policy=L.OPEN_NOOP,
                    inlet=dict(ux=np.full(ni, u_lid), uy=np.zeros(ni)), steps=200000,
                    info=dict(re=re, u_lid=u_lid, nu=nu))


def channel(nr=512, nc=2048, radius=26, centre=(260, 410), u_max=0.04, re=200.0, magic=MAGIC):
    cls = L.geometry_channel(nr, nc)
    L.add_disc(cls, centre[0], centre[1], radius)
    u_mean = 2.0 / 3.0 * u_max
    nu = u_mean * 2 * radius / re
    te, to = trt_taus(nu, magic)
    rows, _ = open_rows(cls, 0)
    h = nr - 2                                    # half-way walls at rows 0.5 and nr - 1.5
    y = rows - 0.5
    ux = 4.0 * u_max * y * (h - y) / (h * h)
    no = len(open_rows(cls, 1)[0])
    return Workload(name=f"channel{nr}x{nc}", node_class=cls, collision=L.TRT, tau_even=te, tau_odd=to,
                    inlet_policy=L.ZOUHE_VELOCITY, outlet_policy=L.ZOUHE_PRESSURE,
                    inlet=dict(ux=ux, uy=np.zeros_like(ux)), outlet=dict(rho=np.ones(no)), steps=50000,
                    info=dict(re=re, u_max=u_max, nu=nu, lift_drag_every=100))


def porous(n=8192, seed=3, rmin=4, rmax=16, solid=0.30, tau=0.8, rho_in=1.001, rho_out=1.000, nc=None):
    nr = n
    nc = n if nc is None else nc
    cls = L.geometry_channel(nr, nc)
    frac, ndiscs = L.geometry_porous(cls, seed, rmin, rmax, solid)
    ni = len(open_rows(cls, 0)[0])
    no = len(open_rows(cls, 1)[0])
    return Workload(name=f"porous{nr}x{nc}", node_class=cls, collision=L.BGK, tau=tau,
                    inlet_policy=L.ZOUHE_PRESSURE, outlet_policy=L.ZOUHE_PRESSURE,
                    inlet=dict(rho=np.full(ni, rho_in)), outlet=dict(rho=np.full(no, rho_out)), steps=20000,
                    info=dict(solid=frac, discs=ndiscs, seed=seed, rho_in=rho_in, rho_out=rho_out))


ENSEMBLE_SIZES = tuple(range(128, 1025, 64))


def ensemble(n_samples=512, seed=5, sizes=ENSEMBLE_SIZES, porosity=(0.5, 0.9), rmin=3, rmax=12, tau=0.8,
             rho_in=1.001, rho_out=1.000):
    """Independent porous channels of uneven sizes: per sample (from one
    seeded generator) rows and columns drawn uniformly from `sizes`, a
    porosity uniform in `porosity` (solid fraction 1 - porosity), discs of
    radius uniform in [rmin, rmax]; the pressure-driven BGK set-up of
    `porous` (tau 0.8, inlet 1.001 / outlet 1.000)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_samples):
        nc = int(rng.choice(sizes))
        nr = int(rng.choice(sizes))
        phi = float(rng.uniform(*porosity))
        wl = porous(nr, seed=seed * 100003 + s, rmin=rmin, rmax=rmax, solid=1.0 - phi, tau=tau, rho_in=rho_in,
                    rho_out=rho_out, nc=nc)
        wl.name = f"sample{s}:{nr}x{nc}"
        wl.info["porosity_target"] = phi
        out.append(wl)
    return out


def darcy_permeability(reduced, nu, length, drho, n_all):
    """k = nu * <j_x> * L / (c_s^2 * drho), <j_x> = fluid x-momentum per node
    of the whole cross-section (superficial), c_s^2 = 1/3."""
    jx_superficial = reduced[1] / n_all
    return nu * jx_superficial * length / (drho / 3.0)


# HBM bytes per wet node and step (9 loads + 9 stores of the storage type).
BYTES_PER_WET_NODE = {L.F64: 144, L.F32_SHIFTED: 72, L.F16_SHIFTED: 36}


def algorithmic_bytes(wl: Workload, storage):
    return wl.wet_count() * BYTES_PER_WET_NODE[storage] + math.ceil(wl.nr * wl.nc / 8)
