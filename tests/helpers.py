"""Shared helpers of the tests: parity metric and small runs."""
import numpy as np


# components of one vector quantity share their scale: momentum (1-3), velocity (6-8)
VECTOR_GROUPS = ((1, 2, 3), (6, 7, 8))


def rel_err(a, b):
    """Norm-wise relative error per component: max|a_c - b_c| / scale_c, with
    scale_c = max|a_c|, or for a component of the momentum or velocity vector
    the largest max|a| over the three components of that vector (a component
    that is physically zero, e.g. w in a 2-D flow, holds only rounding noise)."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = [np.max(np.abs(a[c])) for c in range(a.shape[0])]
    for g in VECTOR_GROUPS:
        m = max(scale[c] for c in g)
        for c in g:
            scale[c] = m
    out = []
    for c in range(a.shape[0]):
        num = np.max(np.abs(a[c] - b[c]))
        out.append(num / scale[c] if scale[c] > 0 else num)
    return np.array(out)


def run_oracle(case, nsteps, state=None):
    from oracle.pyoracle import OracleLevel
    L = OracleLevel(case)
    L.set_state(case.make_state() if state is None else state)
    L.step(nsteps)
    return L.get_state(), L.info()


def run_gpu(case, nsteps, mode=1, state=None, tile=None, boxes=None):
    from mesoflow_amd import runtime
    ctx = runtime.Context(case)
    L = runtime.Level(ctx, boxes)
    if tile:
        L.set_tiles(tile)
    L.set_state(case.make_state() if state is None else state)
    L.step(nsteps, mode)
    out = L.get_state(), L.info()
    L.close()
    ctx.close()
    return out
