"""Core value types of the head-solve path.

Mirrors MacroFlow3D's ``src/core`` and ``src/numerics/pin_spec.hpp``:
``Grid3D`` (src/core/Grid3D.hpp:32-55), ``BCType``/``BCFace``/``BCSpec``
(src/core/BCSpec.hpp:13-88), ``PinMode``/``PinSpec``/``needs_pin``/``pin_enabled``
(src/numerics/pin_spec.hpp:56-120). Same names, field meaning and errors.
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass, field

from . import _lib


@dataclass
class Grid3D:
    """Structured cell-centred grid; cell (i,j,k) is stored at i + nx*(j + ny*k).

    The solver uses dx for every direction (isotropic grids), like the
    reference (src/core/Grid3D.hpp:9-31); dy and dz only enter the velocities.
    """
    nx: int = 0
    ny: int = 0
    nz: int = 0
    dx: float = 1.0
    dy: float | None = None
    dz: float | None = None

    def __post_init__(self):
        if self.dy is None:
            self.dy = self.dx
        if self.dz is None:
            self.dz = self.dx

    def num_cells(self) -> int:
        return self.nx * self.ny * self.nz

    def Lx(self) -> float:
        return self.nx * self.dx

    def Ly(self) -> float:
        return self.ny * self.dy

    def Lz(self) -> float:
        return self.nz * self.dz

    def idx(self, i: int, j: int, k: int) -> int:
        return i + self.nx * (j + self.ny * k)

    @property
    def shape_zyx(self):
        """numpy/torch shape of a cell field viewed as [k, j, i]."""
        return (self.nz, self.ny, self.nx)

    def c(self) -> _lib.mf3d_grid:
        return _lib.mf3d_grid(int(self.nx), int(self.ny), int(self.nz), float(self.dx),
                              float(self.dy), float(self.dz))


class BCType(enum.IntEnum):
    Dirichlet = 0
    Neumann = 1
    Periodic = 2


@dataclass
class BCFace:
    type: BCType = BCType.Dirichlet
    value: float = 0.0


def _default_face():
    return BCFace(BCType.Dirichlet, 0.0)


@dataclass
class BCSpec:
    """Boundary conditions of the six faces (default: Dirichlet 0 everywhere)."""
    xmin: BCFace = field(default_factory=_default_face)
    xmax: BCFace = field(default_factory=_default_face)
    ymin: BCFace = field(default_factory=_default_face)
    ymax: BCFace = field(default_factory=_default_face)
    zmin: BCFace = field(default_factory=_default_face)
    zmax: BCFace = field(default_factory=_default_face)

    def faces(self):
        return [self.xmin, self.xmax, self.ymin, self.ymax, self.zmin, self.zmax]

    def is_periodic_x(self) -> bool:
        return self.xmin.type == BCType.Periodic and self.xmax.type == BCType.Periodic

    def is_periodic_y(self) -> bool:
        return self.ymin.type == BCType.Periodic and self.ymax.type == BCType.Periodic

    def is_periodic_z(self) -> bool:
        return self.zmin.type == BCType.Periodic and self.zmax.type == BCType.Periodic

    def is_all_homog_neumann(self) -> bool:
        return all(f.type == BCType.Neumann and f.value == 0.0 for f in self.faces())

    def validate(self) -> None:
        """BCSpec::validate (src/core/BCSpec.hpp:63-87)."""
        for a, (lo, hi) in zip("xyz", [(self.xmin, self.xmax), (self.ymin, self.ymax),
                                       (self.zmin, self.zmax)]):
            if (lo.type == BCType.Periodic) != (hi.type == BCType.Periodic):
                raise RuntimeError(f"Periodic BC must be specified on both {a}min and {a}max")
        for f in self.faces():
            if not math.isfinite(f.value):
                raise RuntimeError("BC values must be finite")

    def c(self) -> _lib.mf3d_bc:
        o = _lib.mf3d_bc()
        for i, f in enumerate(self.faces()):
            o.type[i] = int(f.type)
            o.value[i] = float(f.value)
        return o

    @staticmethod
    def darcy_x(h_west: float = 1.0, h_east: float = 0.0) -> "BCSpec":
        """Dirichlet h_west at x=0 and h_east at x=L, no-flow on the other faces."""
        n = lambda: BCFace(BCType.Neumann, 0.0)  # noqa: E731
        return BCSpec(BCFace(BCType.Dirichlet, h_west), BCFace(BCType.Dirichlet, h_east),
                      n(), n(), n(), n())


class PinMode(enum.IntEnum):
    Auto = 0
    On = 1
    Off = 2


@dataclass
class PinSpec:
    """Diagonal doubling at cell 0 for singular systems (pin_spec.hpp:72-78)."""
    enabled: bool = False
    index: int = 0


def needs_pin(bc: BCSpec) -> bool:
    """pin_spec.hpp:93-101: singular when no face is Dirichlet."""
    return not any(f.type == BCType.Dirichlet for f in bc.faces())


def pin_enabled(mode: PinMode, bc: BCSpec) -> bool:
    """pin_spec.hpp:110-120."""
    if mode == PinMode.On:
        return True
    if mode == PinMode.Off:
        return False
    return needs_pin(bc)


def pin_mode_str(mode: PinMode) -> str:
    return {PinMode.On: "on", PinMode.Off: "off", PinMode.Auto: "auto"}.get(mode, "unknown")
