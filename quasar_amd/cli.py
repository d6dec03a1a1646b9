"""``python -m quasar_amd.cli run <input.yaml>``: build the solver from a deck,
seed the initial condition, step, and write ``out.npz``
(python/quasar/mhd/cli.py).

The run loop, its time accounting (auto dt refreshed from the live state
before each step, the last step clipped exactly to t_end), the divergence
sampling policy and every key of the output archive follow the reference CLI,
so decks, post-processing and the reference's example checks read the output
unchanged.  Distributed execution (several GPUs, checkpoints) is not part of
this solver; its options are refused with an error.
"""
from __future__ import annotations

import argparse
import copy
import time as _time
from dataclasses import dataclass
from pathlib import Path
from typing import Optional, Sequence, Union

import numpy as np

from . import deck as mhd_io
from . import mhd


@dataclass
class RunResult:
    final_step: int
    final_time: float
    diagnostics_path: Path
    distributed: bool = False


def confine_output_path(base, rel: str, *, label: str = "output.path") -> Path:
    """Resolve a deck-supplied output path inside `base` (the deck's directory)
    and refuse escapes (python/quasar/_paths.py:17-29)."""
    base_dir = Path(base).resolve()
    out = (base_dir / rel).resolve()
    if not out.is_relative_to(base_dir):
        raise ValueError(f"{label} {rel!r} escapes the deck directory {base_dir}")
    out.parent.mkdir(parents=True, exist_ok=True)
    return out


def _positive_int(value: str) -> int:
    v = int(value)
    if v <= 0:
        raise argparse.ArgumentTypeError("must be a positive integer")
    return v


def _make_config(deck: mhd_io.MhdDeck) -> mhd.MhdConfig:
    """Solver configuration of a deck; nghost 0 lets the solver size the halo
    from the reconstruction (cli.py:30-120)."""
    d, n = deck.domain, deck.numerics
    cfg = mhd.MhdConfig()
    cfg.grid = mhd.Grid2D(d.nx, d.ny, d.lx_m, d.ly_m, d.origin_x_m, d.origin_y_m, 0)
    cfg.gamma = n.gamma
    cfg.geometry = deck.geometry
    cfg.reconstruction, cfg.riemann, cfg.integrator = n.reconstruction, n.riemann, n.integrator
    cfg.ct, cfg.positivity = n.ct, n.positivity
    cfg.rho_floor, cfg.p_floor, cfg.cfl = n.rho_floor, n.p_floor, n.cfl
    cfg.arithmetic, cfg.indicator = n.arithmetic, n.indicator
    for side, name in enumerate(deck.boundary.fluid):
        cfg.boundary.set_fluid_side(side, name)
    for side, name in enumerate(deck.boundary.field):
        cfg.boundary.set_field_side(side, name)
    return cfg


def prepare_run(deck: mhd_io.MhdDeck):
    """(solver, dt, dt_is_auto): the solver seeded with the deck's initial
    condition and the t = 0 timestep; an explicit dt above the CFL limit of
    the seeded state is refused before any step (cli.py:123-172)."""
    solver = mhd.MhdSolver2D(_make_config(deck))
    nghost = solver.grid().nghost
    for component, buf in mhd_io.build_initial_state(deck, nghost).items():
        solver.seed_state(component, buf)
    if deck.numerics.indicator == "mlp":
        from . import learned
        params, threshold, _ = learned.load()
        solver.set_indicator(params, threshold)
    cfl = solver.cfl_limit()
    if deck.time.dt_s == "auto":
        return solver, cfl, True
    dt = float(deck.time.dt_s)
    if dt > cfl:
        raise ValueError(f"time.dt_s ({dt:.6e}) exceeds the MHD CFL stability limit ({cfl:.6e}) for this grid and "
                         "seeded state; reduce dt_s or use 'auto'.")
    return solver, dt, False


def _interior(a2d: np.ndarray, nx: int, ny: int, g: int) -> np.ndarray:
    return a2d[g:g + ny, g:g + nx]


def _component_2d(solver, deck: mhd_io.MhdDeck, name: str) -> np.ndarray:
    g = solver.grid().nghost
    flat = np.asarray(solver.state_component_to_host(name))
    if name in ("bx", "by", "bz"):
        flat = mhd_io.magnetic_to_output(flat, deck.units)
    return flat.reshape(deck.domain.ny + 2 * g, deck.domain.nx + 2 * g)


def _fields(solver, deck: mhd_io.MhdDeck, names) -> dict:
    nx, ny, g = deck.domain.nx, deck.domain.ny, solver.grid().nghost
    return {name: _interior(_component_2d(solver, deck, name), nx, ny, g) for name in names}


def _flatten_for_npz(solver, deck: mhd_io.MhdDeck, final_step: int, final_time: float, divb_series: list,
                     snapshots: list, extra_scalars: dict) -> dict:
    """The archive's keys (cli.py:338-384): run metadata as length-1 arrays,
    the final interior state (ny, nx) per component with cell-centred B,
    the div B series, the cadence snapshots and the extra scalars."""
    flat = {"final_step": np.array([final_step]), "final_time_s": np.array([final_time]),
            "nx": np.array([deck.domain.nx]), "ny": np.array([deck.domain.ny]),
            "lx_m": np.array([deck.domain.lx_m]), "ly_m": np.array([deck.domain.ly_m]),
            "nghost": np.array([solver.grid().nghost]), "units": np.array([deck.units]),
            "geometry": np.array([deck.geometry]), "gamma": np.array([deck.numerics.gamma])}
    for name, a in _fields(solver, deck, mhd_io.STATE_COMPONENTS).items():
        flat[f"state_{name}"] = a
    if deck.diagnostics.divb:
        series = mhd_io.magnetic_to_output(np.asarray(divb_series, dtype=np.float64), deck.units)
        flat["divb_linf"] = series
        last = series[-1] if divb_series else mhd_io.magnetic_to_output(solver.divergence_b_max(), deck.units)
        flat["divb_linf_final"] = np.array([last])
    if snapshots:
        flat["snapshot_steps"] = np.array([s["step"] for s in snapshots])
        flat["snapshot_times_s"] = np.array([s["time_s"] for s in snapshots])
        for name in deck.diagnostics.fields:
            flat[f"snapshot_state_{name}"] = np.stack([s["fields"][name] for s in snapshots])
        if deck.diagnostics.divb:
            flat["snapshot_divb_linf"] = mhd_io.magnetic_to_output(
                np.array([s["divb"] for s in snapshots], dtype=np.float64), deck.units)
    for key, value in extra_scalars.items():
        flat[key] = np.array([value])
    return flat


def _run_loop(solver, deck: mhd_io.MhdDeck, dt: float, dt_is_auto: bool, out_path,
              args: argparse.Namespace):
    """Step the solver to deck.time.steps or t_end and write the archive
    (cli.py:386-491).  Returns (steps done, simulation time)."""
    snapshots, divb_series, extra = [], [], {}
    sim_time = 0.0
    t0 = _time.time()
    if deck.initial.type == "orszag_tang":
        mass, energy = solver.conserved_cell_totals()
        extra.update(mass_initial=float(mass), energy_initial=float(energy))
    initial = {f"state_{k}_initial": v for k, v in _fields(solver, deck, mhd_io.STATE_COMPONENTS).items()}
    log_every = max(0, int(args.log_every))
    divb_on = bool(deck.diagnostics.divb)
    if divb_on:
        divb_series.append(float(solver.divergence_b_max()))
    cadence, steps, t_end = deck.diagnostics.cadence, deck.time.steps, deck.time.t_end
    done = 0
    while done < steps and (t_end is None or sim_time < float(t_end)):
        limit = float(solver.cfl_limit()) if dt_is_auto else float(dt)
        if not np.isfinite(limit) or not limit > 0.0:
            raise RuntimeError("MHD timestep is not finite and strictly positive")
        dt_step, clipped = limit, False
        if t_end is not None:
            remaining = float(t_end) - sim_time
            dt_step = min(dt_step, remaining)
            clipped = dt_step == remaining
        if not np.isfinite(dt_step) or not dt_step > 0.0:
            raise RuntimeError("MHD timestep cannot make positive finite progress")
        next_time = None
        if not clipped:
            next_time = sim_time + dt_step
            if not np.isfinite(next_time) or not next_time > sim_time:
                raise RuntimeError("MHD timestep is too small to advance simulation time")
        # Auto dt was just reduced from the live state: the unchecked step
        # skips a second reduction.
        if dt_is_auto:
            solver.step_unchecked(dt_step)
        else:
            solver.step(dt_step)
        # The endpoint is assigned only after integrating the exact residual.
        sim_time = float(t_end) if clipped else next_time
        done += 1
        last = done == steps or (t_end is not None and sim_time >= float(t_end))
        on_cadence = cadence > 0 and done % cadence == 0
        logging = log_every > 0 and done % log_every == 0
        divb_now = float(solver.divergence_b_max()) if divb_on and (on_cadence or last or logging) else None
        if divb_now is not None:
            divb_series.append(divb_now)
        if on_cadence:
            snap = {"step": done, "time_s": sim_time, "fields": _fields(solver, deck, deck.diagnostics.fields)}
            if divb_on:
                snap["divb"] = divb_now
            snapshots.append(snap)
        if logging:
            elapsed = _time.time() - t0
            rate = done / elapsed if elapsed > 0 else 0.0
            eta = (steps - done) / rate if rate > 0 else float("nan")
            msg = f"step {done}/{steps}  t={sim_time:.6e}  rate={rate:.0f} step/s  eta={eta:.0f}s"
            if divb_on:
                msg += f"  |divB|inf={divb_now:.3e}"
            print(msg, flush=True)
    flat = _flatten_for_npz(solver, deck, done, sim_time, divb_series, snapshots, extra)
    flat.update(initial)
    # np.savez on a suffixless path would append ".npz": write through the
    # already-confined path so diagnostics.output_path is exact.
    with Path(out_path).open("wb") as fh:
        np.savez(fh, **flat)
    return done, sim_time


def run(input_deck: Union[mhd_io.MhdDeck, str, Path], *, options=None, steps_override: Optional[int] = None,
        verbose: bool = False, print_config: bool = False, log_every: int = 0) -> RunResult:
    """Run a deck (a path or a parsed MhdDeck) on this GPU (cli.py:287-316)."""
    if steps_override is not None and (isinstance(steps_override, bool) or not isinstance(steps_override, int)
                                       or steps_override <= 0):
        raise ValueError("steps_override must be a positive integer")
    if options is not None:
        raise ValueError("distributed execution is not supported by this solver (one GPU per run)")
    if isinstance(input_deck, mhd_io.MhdDeck):
        deck_path, deck = None, input_deck
    else:
        deck_path = Path(input_deck).resolve()
        deck = mhd_io.load(deck_path)
    if steps_override is not None:
        deck = copy.deepcopy(deck)
        deck.time = mhd_io.Time(dt_s=deck.time.dt_s, steps=steps_override, t_end=deck.time.t_end)
        deck.validate()
    directory = deck_path.parent if deck_path is not None else Path.cwd()
    out_path = confine_output_path(directory, deck.diagnostics.output_path, label="diagnostics.output_path")
    solver, dt, dt_is_auto = prepare_run(deck)
    if print_config:
        d, n = deck.domain, deck.numerics
        print(f"deck    : {deck_path if deck_path is not None else '<in-memory>'}")
        print(f"grid    : {d.nx}x{d.ny}  ({d.lx_m}x{d.ly_m})  origin=({d.origin_x_m}, {d.origin_y_m})  "
              f"nghost={solver.grid().nghost}")
        print(f"geometry: {deck.geometry}")
        print(f"gamma   : {n.gamma}")
        print(f"schemes : recon={n.reconstruction} riemann={n.riemann} integ={n.integrator} ct={n.ct} "
              f"pos={n.positivity} arithmetic={n.arithmetic} indicator={n.indicator}")
        print(f"initial : {deck.initial.type}")
        print(f"dt      : {dt:.6e}    steps: {deck.time.steps}")
    final_step, final_time = _run_loop(solver, deck, dt, dt_is_auto, out_path,
                                       argparse.Namespace(log_every=log_every))
    if print_config or verbose:
        print(f"wrote   : {out_path}")
    return RunResult(final_step=final_step, final_time=final_time, diagnostics_path=out_path)


_DISTRIBUTED_OPTIONS = ("devices", "decomposition", "transport", "diagnostics_layout", "checkpoint",
                        "checkpoint_every", "restart")


def _build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(prog="quasar-amd-mhd", description="Ideal-MHD driver (one MI355X).")
    sub = p.add_subparsers(dest="command", required=True)
    r = sub.add_parser("run", help="Run an MHD simulation from a YAML deck.")
    r.add_argument("input", help="Path to the YAML deck.")
    r.add_argument("--verbose", action="store_true", help="print informational output (default: quiet)")
    r.add_argument("--print-config", action="store_true", help="Print resolved deck + dt before running.")
    r.add_argument("--steps-override", type=_positive_int, default=None, help="Override deck.time.steps.")
    r.add_argument("--log-every", type=int, default=0, help="Print progress every N steps (0 = off).")
    # The reference's distributed options, accepted only to refuse them clearly.
    r.add_argument("--devices", default=None)
    r.add_argument("--decomposition", default=None)
    r.add_argument("--transport", default=None)
    r.add_argument("--diagnostics-layout", default=None)
    r.add_argument("--checkpoint", default=None)
    r.add_argument("--checkpoint-every", type=_positive_int, default=None)
    r.add_argument("--restart", default=None)
    return p


def main(argv: Optional[Sequence[str]] = None) -> int:
    parser = _build_parser()
    args = parser.parse_args(argv)
    if any(getattr(args, name) is not None for name in _DISTRIBUTED_OPTIONS):
        parser.error("distributed execution is not supported by this solver (one GPU per run)")
    run(args.input, steps_override=args.steps_override, verbose=args.verbose, print_config=args.print_config,
        log_every=args.log_every)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
