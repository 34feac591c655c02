"""Helpers shared by the GPU parity tests: build the product grid and the
oracle on the same configuration and compare channel blocks."""
import numpy as np

import msbgx
from oracle_lib import build_oracle


def pair(cfg, walls=msbgx.DBC_SOLID, options=0):
    g = msbgx.build_grid(cfg, walls)
    if options:
        g.set_options(options)
    g.prepare_pressure_solve_flip(scene=cfg.scene)
    o = build_oracle(cfg, walls)
    o.prepare_scene(cfg.scene)
    return g, o


def blocks_of(g, mg, which="all"):
    lv, fl = g.block_info(mg)
    if which == "relax":
        return g.relax_list(mg), lv
    if which == "fluid":
        return np.nonzero((fl & msbgx.BLK_NO_FLUID) == 0)[0], lv
    return np.arange(g.n_blocks), lv


def compare_channel(g, o, ch, mg, blocks, lv, exact=True, rtol=0.0):
    """Max abs diff and max abs value over the given blocks; asserts equality
    bit-for-bit when exact."""
    maxdiff, maxval, nmis = 0.0, 0.0, 0
    for b in blocks:
        e = g.bsx >> int(lv[b])
        a = g.get_block(ch, mg, int(b), e)
        r = o.get_block(ch, mg, int(b), e)
        if r is None:
            r = np.zeros(e ** 3, a.dtype if a is not None else np.float32)
        if a is None:
            a = np.zeros_like(r)
        if exact:
            bad = a.view(np.uint32 if a.dtype == np.float32 else np.uint16) != r.view(
                np.uint32 if r.dtype == np.float32 else np.uint16)
            # +0.0 and -0.0 compare equal as values
            if a.dtype == np.float32:
                bad &= ~((a == 0) & (r == 0))
            nmis += int(bad.sum())
        d = np.abs(a.astype(np.float64) - r.astype(np.float64))
        maxdiff = max(maxdiff, float(d.max()) if d.size else 0.0)
        maxval = max(maxval, float(np.abs(r.astype(np.float64)).max()) if r.size else 0.0)
    return nmis, maxdiff, maxval
