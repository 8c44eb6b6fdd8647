"""Input deck of the solve step and the synthetic workloads.

``StepConfig`` carries the knobs of the 3D solve step that the reference reads from SimulationConfig3D
(src/ImpressionFEA.Core/ThreeD/SimulationConfig3D.cs:19-44) and from the static tunables of Forging3DSolver
(src/ImpressionFEA.Core/ThreeD/Forging3DSolver.cs:26-50, :1048): flow law, friction, ram speed, penalty factor,
Newton blend, tolerance, iteration limits, PCG budget.

Dies are analytic height fields (``DieSpec``): kind 0 is a flat plane, kind 1 a top die with a hexagonal cavity,
flash land and gutter, kind 2 one half of a connecting-rod closed die (two bosses joined by an I-section shank, flash
land and gutter; the bottom half mirrors the top). The same height function is evaluated by the host driver and by
the oracle.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np


@dataclass
class DieSpec:
    side: int          # 0 bottom (metal above), 1 top (metal below)
    kind: int = 0      # 0 flat, 1 hex cavity, 2 connecting-rod half
    p: tuple = (0.0,) * 8

    # connecting-rod footprint (kind 2): big boss (centre x, radius), small boss, gutter depth
    CONROD = (-42.0, 20.0, 49.5, 12.5, 5.0)

    # height of the die surface in the closed position (top dies are offset by the current gap)
    def height(self, x, y):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        if self.kind == 0:
            return np.full(np.broadcast(x, y).shape, self.p[0])

        def smooth(u):
            u = np.clip(u, 0.0, 1.0)
            return u * u * (3 - 2 * u)

        if self.kind == 2:
            # p = (flash thickness t, big-boss depth, small-boss depth, flange depth, web depth, shank width W,
            #      wall transition width, land width); depths are per half, measured from the parting face
            t, db_, ds_, df_, dw_, w_, wall, land = self.p
            xb, rb, xs, rs, g = self.CONROD
            fl = 0.25 * w_
            # distances as sqrt of the sum of squares (not hypot): the same IEEE operations in the oracle and on the
            # device, so the three evaluations of the die agree bit for bit
            db = np.sqrt((x - xb) * (x - xb) + y * y) - rb
            ds = np.sqrt((x - xs) * (x - xs) + y * y) - rs
            dsh = np.maximum(np.maximum(xb - x, x - xs), np.abs(y) - 0.5 * w_)
            shank = dw_ + (df_ - dw_) * smooth((np.abs(y) - (0.5 * w_ - fl) + 0.5 * wall) / wall)
            depth = np.maximum(np.maximum(db_ * smooth((-db + 0.5 * wall) / wall),
                                          ds_ * smooth((-ds + 0.5 * wall) / wall)),
                               shank * smooth((-dsh + 0.5 * wall) / wall))
            h = 0.5 * t + depth + g * smooth((np.minimum(np.minimum(db, ds), dsh) - land) / wall)
            return h if self.side == 1 else -h
        s3 = np.sqrt(3.0) / 2
        d = np.maximum(np.abs(x), np.maximum(np.abs(0.5 * x + s3 * y), np.abs(-0.5 * x + s3 * y))) - 0.5 * self.p[2]
        h = self.p[0] + self.p[1] * smooth((-d + 0.5 * self.p[3]) / self.p[3])
        h = h + self.p[5] * smooth((d - self.p[4]) / self.p[3])
        return h

    def clearance(self, q, off):
        h = self.height(q[..., 0], q[..., 1]) + off
        return q[..., 2] - h if self.side == 0 else h - q[..., 2]

    def project(self, q, off):
        """Vertical projection onto the (offset) surface and the normal pointing out of the die."""
        dd = 1e-5
        x, y = q[..., 0], q[..., 1]
        hx = (self.height(x + dd, y) - self.height(x - dd, y)) / (2 * dd)
        hy = (self.height(x, y + dd) - self.height(x, y - dd)) / (2 * dd)
        up = np.stack([-hx, -hy, np.ones_like(hx)], axis=-1)
        up = up / np.sqrt(up[..., 0] * up[..., 0] + up[..., 1] * up[..., 1] + up[..., 2] * up[..., 2])[..., None]
        nrm = up if self.side == 0 else -up
        p = np.stack([x, y, self.height(x, y) + off], axis=-1)
        return p, nrm


@dataclass
class StepConfig:
    # flow law sigma = C (eps0 + strain)^n rate^m (Materials/FlowStressLaw.cs:6-9)
    C: float = 120.0
    n: float = 0.1
    m: float = 0.15
    eps0: float = 0.01
    friction: float = 0.3          # shear friction factor
    ram: float = 10.0              # mm/s
    penalty_factor: float = 1e6    # relative ANP penalty
    newton_blend: float = 0.6
    tolerance: float = 1e-6
    size: float = 60.0             # billet height (reference strain rate = ram / size)
    stroke_increment: float = 0.5
    h: float = 4.0                 # nominal element size (contact tolerances)
    max_iterations: int = 70
    pcg_budget: int = 15
    quadratic_ls: bool = True
    adaptive_newton: bool = True
    extrapolate_start: bool = False
    remesh: bool = False           # remesh when the mesh degrades, retry a failed update on a new mesh (:545, :600-621)
    remesh_quality: float = 0.12   # SimulationConfig3D.RemeshQuality
    sym_factor: int = 1
    bottom: DieSpec = field(default_factory=lambda: DieSpec(0, 0, (0.0,) * 8))
    top: DieSpec = field(default_factory=lambda: DieSpec(1, 0, (20.0,) + (0.0,) * 7))

    def constants(self):
        """rateRef, rateCut, sigRef, kPen, u0 (Forging3DSolver.cs:146-150)."""
        rate_ref = self.ram / self.size
        rate_cut = 1e-3 * rate_ref
        sig_ref = self.C * (self.eps0 + 0.0) ** self.n * max(rate_ref, 1e-12) ** self.m
        k_pen = self.penalty_factor * sig_ref / rate_ref
        return dict(rateRef=rate_ref, rateCut=rate_cut, sigRef=sig_ref, kPen=k_pen, u0=5e-3 * self.ram)


@dataclass
class Workload:
    name: str
    cfg: StepConfig
    diameter: float
    height: float
    nxy: int
    nz: int
    seed: int = 0
    quarter: bool = False
    increments: int = 20
    gap0: float = 0.0
    axis: str = "z"          # billet axis: "z" upright (bottom at z = 0), "x" lying along x, centred on x = 0
    center_z: float = 0.0    # height of the axis of a lying billet

    def mesh(self):
        """The workload's jittered Kuhn billet: (x [nn,3], tets [ne,4], onX0, onY0)."""
        from .mesh import kuhn_cylinder
        x, t, ox, oy = kuhn_cylinder(self.diameter, self.height, self.nxy, self.nz, seed=self.seed,
                                     quarter=self.quarter)
        if self.axis == "x":
            # rotate the upright billet about y by 90 degrees (orientation kept): axis -> x, and centre it
            x = np.stack([x[:, 2] - 0.5 * self.height, x[:, 1], self.center_z - x[:, 0]], axis=1)
            ox = np.zeros_like(ox)
            oy = np.zeros_like(oy)
        return x, t, ox, oy


def upset_small(seed: int = 0, quarter: bool = False, nxy: int = 10, nz: int = 10, **over) -> Workload:
    """Cylinder upset between flat dies: D 40 x H 60 mm, 10x10x10 Kuhn cells (1,331 nodes, 6,000 tets, 3,993 DOF),
    C = 120 MPa, eps0 = 0.01, n = 0.1, m = 0.15, friction 0.3, penalty 1e6, ram 10 mm/s, 20 x 0.5 mm."""
    H, D = 60.0, 40.0
    shut = 20.0
    cfg = StepConfig(C=120.0, n=0.1, m=0.15, eps0=0.01, friction=0.3, ram=10.0, penalty_factor=1e6, tolerance=1e-6,
                     size=H, stroke_increment=0.5, h=max(D / nxy, H / nz), sym_factor=4 if quarter else 1,
                     bottom=DieSpec(0, 0, (0.0,) + (0.0,) * 7), top=DieSpec(1, 0, (shut,) + (0.0,) * 7))
    cfg = replace(cfg, **over)
    return Workload("upset-D40H60", cfg, D, H, nxy, nz, seed=seed, quarter=quarter, increments=20, gap0=H - shut)


def hexdie_production(seed: int = 0, nxy: int = 42, nz: int = 64, **over) -> Workload:
    """Closed-die hex-head upset: billet D 30 x H 45 mm pressed into a hexagonal cavity (across-flats 38 mm) with a
    2 mm flash land of width 6 mm (w/t = 3) and a gutter beyond it; 42x42x64 Kuhn cells = 120,185 nodes, 677,376 tets,
    360,555 DOF. C = 80 MPa, n = 0.05, m = 0.2, eps0 = 0.01, friction 0.4, penalty 1e6, ram 10 mm/s, 40 increments."""
    H, D = 45.0, 30.0
    t, af, wall = 2.0, 38.0, 2.0
    area = np.sqrt(3.0) / 2 * af * af
    depth = np.pi * (D / 2) ** 2 * H / area - t     # cavity volume ~ billet volume
    top = DieSpec(1, 1, (t, depth, af, wall, 3 * t, 5.0, 0.0, 0.0))
    cfg = StepConfig(C=80.0, n=0.05, m=0.2, eps0=0.01, friction=0.4, ram=10.0, penalty_factor=1e6, tolerance=1e-6,
                     size=H, stroke_increment=0.5, h=max(D / nxy, H / nz), sym_factor=1,
                     bottom=DieSpec(0, 0, (0.0,) + (0.0,) * 7), top=top)
    cfg = replace(cfg, **over)
    gap0 = H - (t + depth)                               # billet top touches the cavity ceiling
    return Workload("hexdie-D30H45", cfg, D, H, nxy, nz, seed=seed, increments=40, gap0=gap0)


def conrod_production(seed: int = 0, nxy: int = 40, nz: int = 192, **over) -> Workload:
    """Connecting-rod closed die: a billet D 25 x L 120 mm lying along x, pressed between two die halves with a big
    boss (diameter 40), a small boss (diameter 25) and an I-section shank (width 16, flanges 8 mm and web 3 mm deep per
    half), flash 2.5 mm with a land of w/t = 3 and a gutter; 40x40x192 Kuhn cells = 324,433 nodes, 1,843,200 tets,
    973,299 DOF, full model. Flow law of the hex-die workload; 60 increments close the initial 16.5 mm gap."""
    D, L = 25.0, 120.0
    t = 2.5
    p = (t, 12.0, 9.0, 8.0, 3.0, 16.0, 2.0, 3 * t)
    zc, gap0 = 8.25, 16.5      # the billet rests on the bottom web; the top die first touches the top web
    cfg = StepConfig(C=80.0, n=0.05, m=0.2, eps0=0.01, friction=0.4, ram=10.0, penalty_factor=1e6, tolerance=1e-6,
                     size=D, stroke_increment=gap0 / 60, h=max(D / nxy, L / nz), sym_factor=1,
                     bottom=DieSpec(0, 2, p), top=DieSpec(1, 2, p))
    cfg = replace(cfg, **over)
    return Workload("conrod-D25L120", cfg, D, L, nxy, nz, seed=seed, increments=60, gap0=gap0, axis="x",
                    center_z=zc)

