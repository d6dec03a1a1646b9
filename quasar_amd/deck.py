"""YAML input decks for the ideal-MHD solver (python/quasar/mhd/io.py).

A deck is parsed into the dataclasses below, validated, and lowered to the
solver's configuration and its seeded initial state.  The schema, defaults,
key names and error type (ValueError, naming the offending key) are those of
the reference loader, so its decks parse unchanged:

    units: normalized | SI
    domain: {nx, ny, lx_m, ly_m, origin_x_m?, origin_y_m?}
    geometry: cartesian
    numerics: {gamma, reconstruction, riemann, integrator, ct, positivity,
               rho_floor, p_floor, cfl}
    initial: {type: <token>, params: {...}}
    time: {dt_s: auto | <float>, steps, t_end?}
    diagnostics: {output_path, cadence, fields, divb}
    boundary: {fluid, field}        # scalar | 4-list | side-keyed map
    background_field: {enabled: false}

Scheme and boundary names are checked against the solver's registries
(quasar_amd.mhd.registered_*).  Two optional numerics keys select this
build's extensions: ``arithmetic`` (fp64 | fp32) and ``indicator``
(none | mlp, the learned troubled-cell form with the packaged network).

Not supported here, and rejected with a ValueError naming the key: the
cylindrical geometry and an enabled static background field (the reference's
field-split B0 form with its coil and file sources).  Those decks need the
reference solver.
"""
from __future__ import annotations

import math
import operator
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Dict, Sequence, Tuple, Union

import numpy as np
import yaml

from . import mhd

# Size ceilings for deck-supplied grids (io.py:77-80).
MAX_GRID_DIM = 1 << 16
MAX_GRID_CELLS = 1 << 30

INITIAL_TYPES = ("brio_wu", "alfven_wave", "orszag_tang", "blast", "rotor", "confined_blob")
STATE_COMPONENTS = ("rho", "mx", "my", "mz", "energy", "bx", "by", "bz")
SIDE_KEYS = ("x_lo", "x_hi", "y_lo", "y_hi")
ARITHMETIC_FORMS = ("fp64", "fp32")
INDICATORS = ("none", "mlp")

MU0 = 4.0e-7 * math.pi
SQRT_MU0 = math.sqrt(MU0)

_TOP_LEVEL_KEYS = {"units", "domain", "geometry", "numerics", "initial", "time", "diagnostics", "boundary",
                   "background_field"}
_PRIMITIVE_KEYS = ("rho", "p", "vx", "vy", "vz", "bx", "by", "bz")


# ---------------------------------------------------------------------------
# Value checks (python/quasar/_deck.py).

class _UniqueKeyLoader(yaml.SafeLoader):
    """Safe YAML loader that refuses a mapping key spelled twice."""


def _unique_mapping(loader, node, deep=False):
    seen = set()
    for key_node, _ in node.value:
        if key_node.tag == "tag:yaml.org,2002:merge":
            continue
        key = loader.construct_object(key_node, deep=deep)
        try:
            repeated = key in seen
            seen.add(key)
        except TypeError:
            continue
        if repeated:
            mark = key_node.start_mark
            raise ValueError(f"duplicate YAML key {key!r} at line {mark.line + 1}, column {mark.column + 1}")
    return loader.construct_mapping(node, deep=deep)


_UniqueKeyLoader.add_constructor(yaml.resolver.BaseResolver.DEFAULT_MAPPING_TAG, _unique_mapping)


def load_yaml(stream) -> Any:
    return yaml.load(stream, Loader=_UniqueKeyLoader)


def _require(d: dict, key: str, context: str):
    if key not in d:
        raise ValueError(f"{context}: missing required field {key!r}")
    return d[key]


def _as_finite(value, context: str) -> float:
    if isinstance(value, (bool, np.bool_)):
        raise ValueError(f"{context} must be a finite number, not a boolean")
    try:
        v = float(value)
    except (TypeError, ValueError, OverflowError):
        raise ValueError(f"{context} must be a finite number") from None
    if not math.isfinite(v):
        raise ValueError(f"{context} must be finite")
    return v


def _require_positive(value, context: str) -> float:
    v = _as_finite(value, context)
    if v <= 0.0:
        raise ValueError(f"{context} must be positive")
    return v


def _as_exact_int(value, context: str) -> int:
    """Integers only: booleans and floats (even 16.0) are refused so a typo
    cannot silently change a grid size or a step count."""
    if isinstance(value, (bool, np.bool_)):
        raise ValueError(f"{context} must be an integer, not a boolean")
    try:
        return int(operator.index(value))
    except TypeError:
        raise ValueError(f"{context} must be an integer") from None


def _as_bool(value, context: str) -> bool:
    if not isinstance(value, (bool, np.bool_)):
        raise ValueError(f"{context} must be a boolean")
    return bool(value)


def _mapping(value, context: str) -> dict:
    if not isinstance(value, dict):
        raise ValueError(f"{context} must be a mapping")
    return value


def _reject_unknown(d: dict, allowed, context: str) -> None:
    extra = set(d) - set(allowed)
    if extra:
        raise ValueError(f"{context}: unknown key(s): {', '.join(sorted(repr(k) for k in extra))}")


def _finite_pair(value, context: str) -> Tuple[float, float]:
    if isinstance(value, (str, bytes, dict)):
        raise ValueError(f"{context} must be a two-element sequence")
    try:
        if len(value) != 2:
            raise ValueError(f"{context} must contain exactly two values")
        return _as_finite(value[0], f"{context}[0]"), _as_finite(value[1], f"{context}[1]")
    except TypeError:
        raise ValueError(f"{context} must be a two-element sequence") from None


def parse_side_map(spec, default: str, what: str) -> Tuple[str, str, str, str]:
    """A per-side boundary spec as (x_lo, x_hi, y_lo, y_hi): a scalar for all
    sides, a 4-list, or a map keyed by side name (_deck.py:154-172)."""
    if spec is None:
        return (default,) * 4
    if isinstance(spec, str):
        return (spec,) * 4
    if isinstance(spec, (list, tuple)) and len(spec) == 4:
        return tuple(str(s) for s in spec)
    if isinstance(spec, dict):
        extra = sorted(set(spec) - set(SIDE_KEYS))
        if extra:
            raise ValueError(f"{what} contains unknown side key(s) {extra}")
        return tuple(str(spec.get(k, default)) for k in SIDE_KEYS)
    raise ValueError(f"{what} must be a string, 4-element list, or side-keyed map")


_parse_side_map = parse_side_map


def magnetic_to_internal(value, units: str):
    """Tesla -> sqrt(Pa) for SI decks (python/quasar/mhd/_units.py)."""
    a = np.asarray(value, dtype=np.float64)
    return a / SQRT_MU0 if units == "SI" else a


def magnetic_to_output(value, units: str):
    a = np.asarray(value, dtype=np.float64)
    return a * SQRT_MU0 if units == "SI" else a


# ---------------------------------------------------------------------------
# Schema (io.py:292-360).

@dataclass
class Domain:
    nx: int
    ny: int
    lx_m: float
    ly_m: float
    origin_x_m: float = 0.0
    origin_y_m: float = 0.0


@dataclass
class Numerics:
    gamma: float = 5.0 / 3.0
    reconstruction: str = "mp7"
    riemann: str = "hlld"
    integrator: str = "ssprk3"
    ct: str = "fd_ct_christlieb"
    positivity: str = "troubled_cell"
    rho_floor: float = 1.0e-8
    p_floor: float = 1.0e-9
    cfl: float = 0.4
    arithmetic: str = "fp64"
    indicator: str = "none"


@dataclass
class Initial:
    type: str
    params: dict = field(default_factory=dict)


@dataclass
class Time:
    dt_s: Union[float, str] = "auto"
    steps: int = 100
    t_end: Union[float, None] = None


@dataclass
class Diagnostics:
    output_path: str = "out.npz"
    cadence: int = 0
    fields: list = field(default_factory=lambda: list(STATE_COMPONENTS))
    divb: bool = True

    def __post_init__(self) -> None:
        self.fields = [str(name).lower() for name in self.fields]


@dataclass
class BoundaryConfig:
    """Per-side kinds in the order [x_min, x_max, y_min, y_max]; fluid and field
    closures are independent."""
    fluid: Tuple[str, str, str, str] = ("periodic",) * 4
    field: Tuple[str, str, str, str] = ("periodic",) * 4


@dataclass
class BackgroundConfig:
    """The static background field block.  Only the disabled default is
    supported here: the solver has no field-split B0 form."""
    enabled: bool = False
    profile: str = "uniform"
    bx0: float = 0.0
    by0: float = 0.0
    bz0: float = 0.0
    params: dict = field(default_factory=dict)
    file: Union[str, None] = None
    a_file: Union[str, None] = None
    conductors: Union[list, None] = None


@dataclass
class MhdDeck:
    domain: Domain
    numerics: Numerics = field(default_factory=Numerics)
    initial: Union[Initial, None] = None
    time: Time = field(default_factory=Time)
    diagnostics: Diagnostics = field(default_factory=Diagnostics)
    boundary: BoundaryConfig = field(default_factory=BoundaryConfig)
    background: BackgroundConfig = field(default_factory=BackgroundConfig)
    units: str = "normalized"
    geometry: str = "cartesian"
    raw: dict = field(default_factory=dict)
    source_dir: Union[Path, None] = None

    def validate(self) -> None:
        if self.units not in ("SI", "normalized"):
            raise ValueError("units must be 'SI' or 'normalized'")
        if self.geometry not in ("cartesian", "cylindrical"):
            raise ValueError("geometry must be 'cartesian' or 'cylindrical'")
        if self.geometry == "cylindrical":
            raise ValueError("geometry 'cylindrical' is not supported by this solver (cartesian only)")
        self._validate_domain()
        self._validate_numerics()
        self._validate_initial()
        self._validate_time()
        self._validate_diagnostics()
        self._validate_boundary()
        self._validate_background()

    def _validate_domain(self) -> None:
        nx = _as_exact_int(self.domain.nx, "domain.nx")
        ny = _as_exact_int(self.domain.ny, "domain.ny")
        if nx <= 0 or ny <= 0:
            raise ValueError("domain.nx and domain.ny must be positive")
        if nx > MAX_GRID_DIM or ny > MAX_GRID_DIM:
            raise ValueError(f"domain.nx/ny must be <= {MAX_GRID_DIM} (got {nx}x{ny})")
        if nx * ny > MAX_GRID_CELLS:
            raise ValueError(f"domain.nx*ny must be <= {MAX_GRID_CELLS} cells")
        lx = _as_finite(self.domain.lx_m, "domain.lx_m")
        ly = _as_finite(self.domain.ly_m, "domain.ly_m")
        if lx <= 0.0 or ly <= 0.0:
            raise ValueError("domain.lx_m and domain.ly_m must be positive")
        x0 = _as_finite(self.domain.origin_x_m, "domain.origin_x_m")
        y0 = _as_finite(self.domain.origin_y_m, "domain.origin_y_m")
        x1, y1 = x0 + lx, y0 + ly
        _as_finite(x1, "domain upper x bound")
        _as_finite(y1, "domain upper y bound")
        dx, dy = lx / nx, ly / ny
        _require_positive(dx, "domain x spacing")
        _require_positive(dy, "domain y spacing")
        # Cell centres must stay distinct from the domain edges in binary64.
        if (x0 + 0.5 * dx == x0 or x0 + (nx - 0.5) * dx == x1 or y0 + 0.5 * dy == y0
                or y0 + (ny - 0.5) * dy == y1):
            raise ValueError("domain cell coordinates collapse in floating-point precision")

    def _validate_numerics(self) -> None:
        n = self.numerics
        if _as_finite(n.gamma, "numerics.gamma") <= 1.0:
            raise ValueError("numerics.gamma must be > 1")
        _check_registered(n.reconstruction, mhd.registered_reconstructions(), "numerics.reconstruction")
        _check_registered(n.riemann, mhd.registered_riemann_solvers(), "numerics.riemann")
        _check_registered(n.integrator, mhd.registered_integrators(), "numerics.integrator")
        _check_registered(n.ct, mhd.registered_ct_schemes(), "numerics.ct")
        _check_registered(n.positivity, mhd.registered_positivity_limiters(), "numerics.positivity")
        _require_positive(n.rho_floor, "numerics.rho_floor")
        _require_positive(n.p_floor, "numerics.p_floor")
        cfl = _as_finite(n.cfl, "numerics.cfl")
        if not 0.0 < cfl <= 1.0:
            raise ValueError("numerics.cfl must be in (0, 1]")
        _check_registered(n.arithmetic, ARITHMETIC_FORMS, "numerics.arithmetic")
        _check_registered(n.indicator, INDICATORS, "numerics.indicator")
        if n.indicator == "mlp" and n.reconstruction not in ("mp5", "mp7"):
            raise ValueError("numerics.indicator 'mlp' needs reconstruction mp5 or mp7")

    def _validate_initial(self) -> None:
        if self.initial is None:
            raise ValueError("deck must define an 'initial' section")
        if self.initial.type not in INITIAL_TYPES:
            raise ValueError(f"initial.type {self.initial.type!r} must be one of {list(INITIAL_TYPES)}")
        if not isinstance(self.initial.params, dict):
            raise ValueError("initial.params must be a mapping")
        p = self.initial.params
        kind = self.initial.type
        if kind == "brio_wu":
            _check_brio_wu(p)
        elif kind == "alfven_wave":
            _check_alfven(self, p)
        elif kind == "orszag_tang":
            _reject_unknown(p, {"b0"}, "initial.params")
            _as_finite(p.get("b0", 1.0), "initial.params.b0")
        elif kind == "blast":
            _check_blast(p)
        elif kind == "rotor":
            _check_rotor(p)
        else:
            _check_blob(p, self.domain)

    def _validate_time(self) -> None:
        if isinstance(self.time.dt_s, str) and self.time.dt_s != "auto":
            raise ValueError("time.dt_s must be a float or the string 'auto'")
        if not isinstance(self.time.dt_s, str):
            _require_positive(self.time.dt_s, "time.dt_s")
        if _as_exact_int(self.time.steps, "time.steps") <= 0:
            raise ValueError("time.steps must be positive")
        if self.time.t_end is not None:
            _require_positive(self.time.t_end, "time.t_end")

    def _validate_diagnostics(self) -> None:
        d = self.diagnostics
        if not isinstance(d.output_path, str) or not d.output_path.strip():
            raise ValueError("diagnostics.output_path must be a non-empty string")
        if _as_exact_int(d.cadence, "diagnostics.cadence") < 0:
            raise ValueError("diagnostics.cadence must be >= 0")
        _as_bool(d.divb, "diagnostics.divb")
        for name in d.fields:
            if name not in STATE_COMPONENTS:
                raise ValueError(f"diagnostics.fields entry {name!r} must be one of {list(STATE_COMPONENTS)}")

    def _validate_boundary(self) -> None:
        for what, sides, allowed in (("fluid", self.boundary.fluid, mhd.registered_mhd_fluid_boundaries()),
                                     ("field", self.boundary.field, mhd.registered_mhd_field_boundaries())):
            for k, kind in enumerate(sides):
                if kind == "internal":
                    raise ValueError(f"boundary.{what}[{k}] = 'internal' is reserved for the distributed tile "
                                     "runtime")
                if kind == "axis":
                    raise ValueError("boundary kind 'axis' requires geometry 'cylindrical'")
                if kind not in allowed:
                    raise ValueError(f"boundary.{what}[{k}] = {kind!r} must be one of {sorted(allowed)}")
        fl, fd = self.boundary.fluid, self.boundary.field
        for side in range(4):
            if (fl[side] == "periodic") != (fd[side] == "periodic"):
                raise ValueError("fluid and field periodicity must match on every side")
        for axis in range(2):
            lo, hi = 2 * axis, 2 * axis + 1
            if (fl[lo] == "periodic") != (fl[hi] == "periodic") or (fd[lo] == "periodic") != (fd[hi] == "periodic"):
                raise ValueError("periodic boundaries must be selected on both sides of an axis")

    def _validate_background(self) -> None:
        if _as_bool(self.background.enabled, "background_field.enabled"):
            raise ValueError("background_field.enabled: the static background field B0 (field-split form) is "
                             "not supported by this solver")


def _check_registered(name, allowed: Sequence[str], context: str) -> None:
    if name not in set(allowed):
        raise ValueError(f"{context} {name!r} must be one of {sorted(set(allowed))}")


def _check_brio_wu(p: dict) -> None:
    _reject_unknown(p, {"interface", "left", "right"}, "initial.params")
    _as_finite(_require(p, "interface", "initial.params"), "initial.params.interface")
    states = {}
    for side in ("left", "right"):
        s = _mapping(_require(p, side, "initial.params"), f"initial.params.{side}")
        _reject_unknown(s, _PRIMITIVE_KEYS, f"initial.params.{side}")
        missing = [k for k in _PRIMITIVE_KEYS if k not in s]
        if missing:
            raise ValueError(f"initial.params.{side}: missing required field(s) {missing}")
        for k in _PRIMITIVE_KEYS:
            _as_finite(s[k], f"initial.params.{side}.{k}")
        for k in ("rho", "p"):
            if float(s[k]) <= 0.0:
                raise ValueError(f"initial.params.{side}.{k} must be positive")
        states[side] = s
    if float(states["left"]["bx"]) != float(states["right"]["bx"]):
        raise ValueError("brio_wu requires a continuous normal magnetic field: initial.params.left.bx must "
                         "equal initial.params.right.bx")


def _check_alfven(deck: MhdDeck, p: dict) -> None:
    _reject_unknown(p, {"rho", "p", "b0", "amplitude", "wavenumber", "polarization"}, "initial.params")
    _require_positive(p.get("rho", 1.0), "initial.params.rho")
    _require_positive(p.get("p", 0.1), "initial.params.p")
    _as_finite(p.get("amplitude", 1.0e-6), "initial.params.amplitude")
    if _as_finite(p.get("b0", 1.0), "initial.params.b0") == 0.0:
        raise ValueError("alfven_wave requires a nonzero total axial background field")
    n = _as_exact_int(p.get("wavenumber", 1), "initial.params.wavenumber")
    if n <= 0:
        raise ValueError("initial.params.wavenumber must be positive")
    if 2 * n >= deck.domain.nx:
        raise ValueError("initial.params.wavenumber must be below the x-grid Nyquist mode")
    pol = p.get("polarization", "circular")
    if not isinstance(pol, str) or pol != "circular":
        raise ValueError("alfven_wave currently supports only polarization='circular'")


def _check_blast(p: dict) -> None:
    _reject_unknown(p, {"rho_ambient", "p_ambient", "p_core", "r_in", "center", "bx", "by", "bz"},
                    "initial.params")
    for k, default in (("rho_ambient", 1.0), ("p_ambient", 0.1), ("p_core", 10.0), ("r_in", 0.1)):
        _require_positive(p.get(k, default), f"initial.params.{k}")
    _finite_pair(p.get("center", [0.0, 0.0]), "initial.params.center")
    for k in ("bx", "by", "bz"):
        _as_finite(p.get(k, 0.0), f"initial.params.{k}")


def _check_rotor(p: dict) -> None:
    _reject_unknown(p, {"center", "r0", "r1", "rho_in", "rho_out", "u0", "p", "bx", "by", "bz"}, "initial.params")
    _finite_pair(p.get("center", [0.5, 0.5]), "initial.params.center")
    r0 = _as_finite(p.get("r0", 0.1), "initial.params.r0")
    r1 = _as_finite(p.get("r1", 0.115), "initial.params.r1")
    if r0 <= 0.0:
        raise ValueError("initial.params.r0 must be positive")
    if r1 <= r0:
        raise ValueError("initial.params.r1 must be greater than r0")
    for k, default in (("rho_in", 10.0), ("rho_out", 1.0), ("p", 1.0)):
        _require_positive(p.get(k, default), f"initial.params.{k}")
    u0 = _as_finite(p.get("u0", 2.0), "initial.params.u0")
    if not math.isfinite(u0 / r0):
        raise ValueError("rotor angular velocity u0/r0 is not representable")
    for k in ("bx", "by", "bz"):
        _as_finite(p.get(k, 0.0), f"initial.params.{k}")


def _check_blob(p: dict, domain: Domain) -> None:
    _reject_unknown(p, {"bz", "rho_in", "rho_out", "p_in", "p_out", "blob_half"}, "initial.params")
    for k, default in (("rho_in", 10.0), ("rho_out", 1.0), ("p_in", 1.0), ("p_out", 0.1)):
        _require_positive(p.get(k, default), f"initial.params.{k}")
    _as_finite(p.get("bz", 0.1), "initial.params.bz")
    _require_positive(p.get("blob_half", 0.25 * min(domain.lx_m, domain.ly_m)), "initial.params.blob_half")


# ---------------------------------------------------------------------------
# Parsing (io.py:722-852).

def _parse_domain(d) -> Domain:
    d = _mapping(d, "domain")
    _reject_unknown(d, {"nx", "ny", "lx_m", "ly_m", "origin_x_m", "origin_y_m"}, "domain")
    return Domain(nx=_as_exact_int(_require(d, "nx", "domain"), "domain.nx"),
                  ny=_as_exact_int(_require(d, "ny", "domain"), "domain.ny"),
                  lx_m=_as_finite(_require(d, "lx_m", "domain"), "domain.lx_m"),
                  ly_m=_as_finite(_require(d, "ly_m", "domain"), "domain.ly_m"),
                  origin_x_m=_as_finite(d.get("origin_x_m", 0.0), "domain.origin_x_m"),
                  origin_y_m=_as_finite(d.get("origin_y_m", 0.0), "domain.origin_y_m"))


def _parse_numerics(d) -> Numerics:
    if d is None:
        return Numerics()
    d = _mapping(d, "numerics")
    _reject_unknown(d, {"gamma", "reconstruction", "riemann", "integrator", "ct", "positivity", "rho_floor",
                        "p_floor", "cfl", "arithmetic", "indicator"}, "numerics")
    base = Numerics()
    return Numerics(gamma=_as_finite(d.get("gamma", base.gamma), "numerics.gamma"),
                    reconstruction=str(d.get("reconstruction", base.reconstruction)),
                    riemann=str(d.get("riemann", base.riemann)),
                    integrator=str(d.get("integrator", base.integrator)),
                    ct=str(d.get("ct", base.ct)),
                    positivity=str(d.get("positivity", base.positivity)),
                    rho_floor=_as_finite(d.get("rho_floor", base.rho_floor), "numerics.rho_floor"),
                    p_floor=_as_finite(d.get("p_floor", base.p_floor), "numerics.p_floor"),
                    cfl=_as_finite(d.get("cfl", base.cfl), "numerics.cfl"),
                    arithmetic=str(d.get("arithmetic", base.arithmetic)),
                    indicator=str(d.get("indicator", base.indicator)))


def _parse_initial(d) -> Union[Initial, None]:
    if d is None:
        return None
    d = _mapping(d, "initial")
    _reject_unknown(d, {"type", "params"}, "initial")
    return Initial(type=str(_require(d, "type", "initial")),
                   params=dict(_mapping(d.get("params", {}), "initial.params")))


def _parse_time(d) -> Time:
    if d is None:
        return Time()
    d = _mapping(d, "time")
    _reject_unknown(d, {"dt_s", "steps", "t_end"}, "time")
    dt = d.get("dt_s", "auto")
    t_end = d.get("t_end")
    return Time(dt_s=dt if isinstance(dt, str) else _as_finite(dt, "time.dt_s"),
                steps=_as_exact_int(d.get("steps", 100), "time.steps"),
                t_end=None if t_end is None else _as_finite(t_end, "time.t_end"))


def _parse_diagnostics(d) -> Diagnostics:
    if d is None:
        return Diagnostics()
    d = _mapping(d, "diagnostics")
    _reject_unknown(d, {"output_path", "cadence", "fields", "divb"}, "diagnostics")
    out = d.get("output_path", "out.npz")
    if not isinstance(out, str):
        raise ValueError("diagnostics.output_path must be a string")
    return Diagnostics(output_path=out,
                       cadence=_as_exact_int(d.get("cadence", 0), "diagnostics.cadence"),
                       fields=list(d.get("fields", list(STATE_COMPONENTS))),
                       divb=_as_bool(d.get("divb", True), "diagnostics.divb"))


def _parse_boundary(d) -> BoundaryConfig:
    if d is None:
        return BoundaryConfig()
    d = _mapping(d, "boundary")
    _reject_unknown(d, {"fluid", "field"}, "boundary")
    for name in ("fluid", "field"):
        if isinstance(d.get(name), dict):
            _reject_unknown(d[name], set(SIDE_KEYS), f"boundary.{name}")
    return BoundaryConfig(fluid=parse_side_map(d.get("fluid"), "periodic", "boundary.fluid"),
                          field=parse_side_map(d.get("field"), "periodic", "boundary.field"))


def _parse_background(d) -> BackgroundConfig:
    if d is None:
        return BackgroundConfig()
    d = _mapping(d, "background_field")
    _reject_unknown(d, {"enabled", "profile", "bx0", "by0", "bz0", "params", "file", "a_file", "conductors"},
                    "background_field")
    # `enabled` is a master switch: a disabled block is the default block.
    if not _as_bool(d.get("enabled", False), "background_field.enabled"):
        return BackgroundConfig()
    return BackgroundConfig(enabled=True, profile=str(d.get("profile", "uniform")))


def parse(data) -> MhdDeck:
    data = _mapping(data, "deck")
    _reject_unknown(data, _TOP_LEVEL_KEYS, "deck")
    deck = MhdDeck(domain=_parse_domain(_require(data, "domain", "deck")),
                   numerics=_parse_numerics(data.get("numerics")),
                   initial=_parse_initial(data.get("initial")),
                   time=_parse_time(data.get("time")),
                   diagnostics=_parse_diagnostics(data.get("diagnostics")),
                   boundary=_parse_boundary(data.get("boundary")),
                   background=_parse_background(data.get("background_field")),
                   units=str(data.get("units", "normalized")),
                   geometry=str(data.get("geometry", "cartesian")),
                   raw=data)
    deck.validate()
    return deck


def load(path: Union[str, Path]) -> MhdDeck:
    with open(path) as fh:
        data = load_yaml(fh)
    if not isinstance(data, dict):
        raise ValueError(f"{path}: top-level YAML must be a mapping")
    deck = parse(data)
    deck.source_dir = Path(path).resolve().parent
    return deck


# ---------------------------------------------------------------------------
# Initial state (io.py:922-1012): the deck's scalars lowered to the seed
# kernel's parameter block; every per-cell value is computed on the device.

def initial_condition_spec(deck: MhdDeck, nghost: int) -> mhd.MhdInitialConditionSpec:
    d, p = deck.domain, deck.initial.params
    spec = mhd.MhdInitialConditionSpec()
    spec.kind = deck.initial.type
    spec.grid = mhd.Grid2D(d.nx, d.ny, d.lx_m, d.ly_m, d.origin_x_m, d.origin_y_m, nghost)
    spec.gamma = deck.numerics.gamma
    spec.scheme_order = mhd.reconstruction_order(deck.numerics.reconstruction)
    spec.magnetic_scale = 1.0 / SQRT_MU0 if deck.units == "SI" else 1.0

    def fl(key, default):
        return float(p.get(key, default))

    kind = deck.initial.type
    if kind == "brio_wu":
        spec.interface = fl("interface", 0.5)
        spec.left = [float(p["left"][k]) for k in _PRIMITIVE_KEYS]
        spec.right = [float(p["right"][k]) for k in _PRIMITIVE_KEYS]
    elif kind == "alfven_wave":
        spec.rho, spec.pressure = fl("rho", 1.0), fl("p", 0.1)
        spec.b0 = spec.total_b0 = fl("b0", 1.0)
        spec.amplitude = fl("amplitude", 1.0e-6)
        spec.wavenumber = _as_exact_int(p.get("wavenumber", 1), "initial.params.wavenumber")
        # SI: the eigenrelation's dB/sqrt(mu0 rho) and the sub-cell energy term.
        spec.magnetic_velocity_scale = 1.0 / SQRT_MU0 if deck.units == "SI" else 1.0
    elif kind == "orszag_tang":
        spec.b_uniform = [fl("b0", 1.0), 0.0, 0.0]
    elif kind == "blast":
        c = p.get("center", [0.0, 0.0])
        spec.center = [float(c[0]), float(c[1])]
        spec.rho_ambient, spec.p_ambient = fl("rho_ambient", 1.0), fl("p_ambient", 0.1)
        spec.p_core, spec.r_in = fl("p_core", 10.0), fl("r_in", 0.1)
        spec.b_uniform = [fl("bx", 0.0), fl("by", 0.0), fl("bz", 0.0)]
    elif kind == "rotor":
        c = p.get("center", [0.5, 0.5])
        spec.center = [float(c[0]), float(c[1])]
        spec.r0, spec.r1 = fl("r0", 0.1), fl("r1", 0.115)
        spec.rho_in, spec.rho_out = fl("rho_in", 10.0), fl("rho_out", 1.0)
        spec.u0, spec.p_ambient = fl("u0", 2.0), fl("p", 1.0)
        spec.b_uniform = [fl("bx", 0.0), fl("by", 0.0), fl("bz", 0.0)]
    else:  # confined_blob, centred in the domain
        spec.center = [d.origin_x_m + 0.5 * d.lx_m, d.origin_y_m + 0.5 * d.ly_m]
        spec.rho_in, spec.rho_out = fl("rho_in", 10.0), fl("rho_out", 1.0)
        spec.p_in, spec.p_out = fl("p_in", 1.0), fl("p_out", 0.1)
        spec.blob_half = fl("blob_half", 0.25 * min(d.lx_m, d.ly_m))
        spec.b_uniform = [0.0, 0.0, fl("bz", 0.1)]
    return spec


def build_initial_state(deck: MhdDeck, nghost: int) -> Dict[str, np.ndarray]:
    """{component: flat ghost-padded host array} for every STATE_COMPONENT,
    seeded on the device (needs the GPU)."""
    if _as_exact_int(nghost, "nghost") < 0:
        raise ValueError("nghost must be non-negative")
    return mhd.build_initial_state(initial_condition_spec(deck, nghost))
