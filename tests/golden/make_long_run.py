"""Golden fixture of a long run: the CPU restatement advanced 1000 frames on a
seeded 48x48 blob canvas (tilt 20 deg, 0.586 mm cells) fed by a circling
nozzle for 800 frames; stores the state after 100 and 1000 frames.  The GPU test replays the same frames and compares.

python make_long_run.py            -> long_run_48.npz      (n = 1, the reference's default)
python make_long_run.py --n 0.7    -> long_run_48_n07.npz  (shear-thinning power law)"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "paint-sim-3d_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import scenes  # noqa: E402
from helpers import make_pair  # noqa: E402

W = 48
FRAMES = (100, 1000)
KW = dict(seed=1000, nblobs=7, sigma_cells=(2.0, 6.0), peak=(2e-4, 1.5e-3), tau=(0.01, 0.5), mu=(2.0, 20.0))


FIXTURES = {1.0: "long_run_48.npz", 0.7: "long_run_48_n07.npz"}


def scene(n_index=1.0):
    s = scenes.blobs_settings(W, tilt_deg=20.0)
    s.n_index = n_index
    st = scenes.blobs_state(W, W, **KW)
    return s, st


def deposits(frame, dx, make_capsule):
    """A nozzle circling the canvas centre for the first 800 frames."""
    if frame > 800:
        return []
    import math
    t0, t1 = 0.05 * (frame - 1), 0.05 * frame
    c = 0.5 * W * dx
    rad = 0.3 * W * dx
    p0 = (c + rad * math.cos(t0), c + rad * math.sin(t0))
    p1 = (c + rad * math.cos(t1), c + rad * math.sin(t1))
    return [make_capsule(p0, p1, 1.5 * dx, 4e-11, scenes.jet_paint(), dx)]


def digest(st):
    h = hashlib.sha256()
    for k in sorted(st):
        h.update(np.ascontiguousarray(st[k]).tobytes())
    return h.hexdigest()


if __name__ == "__main__":
    n_index = float(sys.argv[sys.argv.index("--n") + 1]) if "--n" in sys.argv else 1.0
    s, st = scene(n_index)
    _, o = make_pair(W, W, s, st, gpu=False)
    out = {"init_sha256": np.array(digest(st))}
    t0 = time.time()
    from helpers import oracle_capsule
    dx = o.params.dx
    for f in range(1, FRAMES[-1] + 1):
        o.step(deposits(f, dx, oracle_capsule))
        if f in FRAMES:
            for k in ("h", "strata_h", "strata_pig", "rheo", "delta", "tile_flags"):
                out[f"f{f}_{k}"] = o.buf(k).copy()
            r = o.mass_report()
            out[f"f{f}_wet_ml"] = np.array(r.wet_ml)
            out[f"f{f}_substeps"] = np.array(round(o.params.frame_dt / o.params.dt))
            print(f"frame {f}: {time.time() - t0:.1f} s, active {o.active_count()}, wet {r.wet_ml:.6f} mL", flush=True)
    np.savez_compressed(os.path.join(HERE, FIXTURES[n_index]), **out)
