"""Measured choice of the real-subspace kernel's state tile.

A state tile splits an atom's S configuration orders over L lanes of SL
orders each (``tsim_tile_shape``). Which tile runs fastest depends on S, on
the tangent set carried, on the B1 map and diffusion variants and on the
device; the library's built-in choice is a cost model. ``tile_for`` instead
times every covering tile on a probe of the actual launch (its event stream,
up to 16,384 of its atoms) the first time a configuration is seen, and
caches the winner per (device, S, columns, feature bits). Every tile runs the
same per-order arithmetic, so the choice changes only the speed.

Set ``TORCHSIM_AMD_TILE_CACHE`` to a JSON file path to keep the choices across
processes.
"""

from __future__ import annotations

import json
import os
import threading
from typing import Callable

import torch

from .. import _native

PROBE_ATOMS = 16384
_FEATURES_THAT_MATTER = _native.FEATURE_DIFFUSING | _native.FEATURE_TRANSMIT

_lock = threading.Lock()
_choices: dict[str, int] = {}
_loaded = False


def _cache_path() -> str | None:
    return os.environ.get("TORCHSIM_AMD_TILE_CACHE") or None


def _load() -> None:
    global _loaded
    if _loaded:
        return
    _loaded = True
    path = _cache_path()
    if path and os.path.exists(path):
        try:
            with open(path, encoding="utf-8") as handle:
                stored = json.load(handle)
            _choices.update({str(k): int(v) for k, v in stored.items()})
        except (OSError, ValueError):
            pass


def _save() -> None:
    path = _cache_path()
    if not path:
        return
    try:
        with open(path, "w", encoding="utf-8") as handle:
            json.dump(_choices, handle, indent=1, sort_keys=True)
    except OSError:
        pass


def key(device: torch.device, states: int, columns: tuple[str, ...], features: int, tag: str = "") -> str:
    name = torch.cuda.get_device_name(device)
    text = f"{name}|S={states}|cols={','.join(columns)}|f={features & _FEATURES_THAT_MATTER}"
    return f"{text}|{tag}" if tag else text


def candidates(states: int) -> list[int]:
    """Tiles covering S orders, without those carrying more than twice the
    orders of the smallest cover (pure padding)."""
    shapes = _native.tiles()
    covering = [(sl * lanes, index) for index, (sl, lanes) in enumerate(shapes) if sl * lanes >= states]
    if not covering:
        return []
    smallest = min(cover for cover, _ in covering)
    return [index for cover, index in covering if cover <= 2 * smallest]


def choose(run: Callable[[int], None], tiles: list[int], *, repeats: int = 2) -> int:
    """The tile among ``tiles`` with the shortest ``run(tile)`` on the current
    stream (one warm-up launch each, then the best of ``repeats``)."""
    best, best_ms = tiles[0], float("inf")
    for tile in tiles:
        run(tile)
        elapsed = float("inf")
        for _ in range(repeats):
            start = torch.cuda.Event(enable_timing=True)
            stop = torch.cuda.Event(enable_timing=True)
            start.record()
            run(tile)
            stop.record()
            stop.synchronize()
            elapsed = min(elapsed, start.elapsed_time(stop))
        if elapsed < best_ms:
            best, best_ms = tile, elapsed
    return best


def tile_for(device: torch.device, states: int, columns: tuple[str, ...], features: int,
             probe: Callable[[int], None], tag: str = "") -> int:
    """The cached or freshly measured tile for this configuration; ``probe``
    runs one launch of the configuration on the probe atoms with a given tile;
    ``tag`` names what else the choice is kept for (a kernel variant, a size
    class)."""
    with _lock:
        _load()
        name = key(device, states, columns, features, tag)
        if name in _choices:
            return _choices[name]
    tiles = candidates(states)
    if not tiles:
        return -1
    chosen = tiles[0] if len(tiles) == 1 else choose(probe, tiles)
    with _lock:
        _choices[name] = chosen
        _save()
    return chosen


def forget() -> None:
    """Drop the in-memory choices (the next launch of each configuration measures again)."""
    with _lock:
        _choices.clear()
