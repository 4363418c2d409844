"""Autotuner of an ensemble's layout for a workload.

The reference fixes the cell size (Field2D::new's discretization) and it is
part of the semantics (the query window is floor(radius / d) cells), so the
tunable knobs are the device-side ones: the layout's form (tile-resident or
bin-grouped), tile width and height in cells, bin height, and the LDS budget
of a normal-class work-group of the tile-resident form.  The tuner builds an
ensemble of the workload for each candidate, times a few schedule steps
with HIP events on the ensemble's stream after a warm-up, and keeps the
fastest.  Results do not depend on the choice (the tests check this), only
speed.
"""
import ctypes
import itertools

import numpy as np

from .ensemble import Ensemble, Tuning


class _Hip:
    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")

    def event(self):
        e = ctypes.c_void_p()
        assert self.L.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def elapsed_ms(self, stream, fn):
        a, b = self.event(), self.event()
        assert self.L.hipEventRecord(a, ctypes.c_void_p(stream)) == 0
        fn()
        assert self.L.hipEventRecord(b, ctypes.c_void_p(stream)) == 0
        assert self.L.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.L.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        self.L.hipEventDestroy(a)
        self.L.hipEventDestroy(b)
        return ms.value


def candidates(mean_agents_per_cell):
    """Tile shapes around ~450 agents per tile, bins of 8-16 rows, and two
    LDS budgets (four or five normal work-groups per CU)."""
    target = 450.0 / max(mean_agents_per_cell, 1e-6)
    out = [Tuning()]
    for tc, tr_ in itertools.product((8, 16, 32), (16, 32, 64, 128, 256)):
        if not (0.35 * target <= tc * tr_ <= 2.5 * target):
            continue
        for hb in (8, 16):
            if hb > tr_:
                continue
            for lds in (24 * 1024, 32 * 1024):
                out.append(Tuning(tile_cols=tc, tile_rows=tr_, bin_rows=hb, lds_normal=lds))
    return out


PATHS = ({"binned": True}, {"tile_resident": True})


def tune(specs, steps=16, warmup=8, max_candidates=12, log=None):
    """Picks the fastest Tuning and form for the runs `specs`; returns
    (tuning, path, table): path holds the Ensemble.run keywords of the form,
    table rows are (tuning, path, ms per step).  The LDS budget only shapes
    the tile-resident form, so the bin-grouped one is timed once per tile
    shape."""
    hip = _Hip()
    cells = sum(s.width * s.height / (s.discretization ** 2) for s in specs)
    agents = sum(s.agents for s in specs)
    cands = candidates(agents / max(cells, 1.0))
    if len(cands) > max_candidates:
        # keep the automatic choice and an even spread of the rest
        idx = np.linspace(1, len(cands) - 1, max_candidates - 1).round().astype(int)
        cands = [cands[0]] + [cands[i] for i in sorted(set(idx))]
    table = []
    seen = set()
    for t in cands:
        for path in PATHS:
            if path.get("binned"):
                shape = (t.tile_cols, t.tile_rows, t.bin_rows)
                if shape in seen:
                    continue
                seen.add(shape)
            try:
                with Ensemble(t) as ens:
                    ens.admit(specs)
                    ens.run(warmup, **path)
                    ens.sync()
                    ms = hip.elapsed_ms(ens.stream, lambda: ens.run(steps, **path))
            except Exception as exc:  # a candidate the geometry rejects
                if log:
                    log(f"autotune: {t} {path} rejected ({exc})")
                continue
            table.append((t, path, ms / steps))
            if log:
                log(f"autotune: {t} {path} -> {ms / steps:.3f} ms/step")
    best = min(table, key=lambda r: r[2])
    return best[0], best[1], table
