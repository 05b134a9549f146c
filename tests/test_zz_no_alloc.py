"""nlpf_screen_run_noalloc (the entry point that allocates nothing, its level outputs in the
library's static device scratch) against nlpf_screen_run on the same records: the films must be
equal bit for bit.  Two calls in a row on one handle, the second with a smaller input; the first
is large enough that its slices run in two groups through the static scratch."""
import numpy as np
import pytest

import nlpf as N
from gpu_util import gen_screen, ramp_stops

pytestmark = pytest.mark.gpu


def film_pair(h, desc, rec, anc):
    W, H = desc.width, desc.height
    dt = np.float64 if desc.precision == N.FP64 else np.float32
    nb = H * W * 3 * dt().itemsize
    f_ref, f_na = h.alloc(nb), h.alloc(nb)
    h.screen_run(desc, rec, anc, f_ref, None)
    h.memset(f_na, 0xFF, nb)  # NaN everywhere: every pixel must be written
    h.screen_run_noalloc(desc, rec, anc, f_na)
    h.sync()
    a = f_ref.download(dt, H * W * 3)
    b = f_na.download(dt, H * W * 3)
    f_ref.free()
    f_na.free()
    return a, b


def test_noalloc_matches_screen_run_twice():
    pos, val = ramp_stops(np.random.default_rng(11))
    table = N.style_table([N.toon(0.2, depth_mask=N.depth_mask(1)), N.ramp(pos, val, inner=0),
                           N.hue_shift(0.1, inner=1)])
    with N.Handle({"type": "pf", "mode": "screen"}) as h:
        h.set_style(table, [2])
        cases = [
            # 1080p, 6 slices: 49.8 MB of level images per slice, 5 slices per group -> 2 groups
            (1920, 1080, 6, 3, N.FP32, 7),
            # smaller input on the same handle and the same scratch, FP16 records, 7x7 window
            (301, 163, 3, 4, N.FP16, 3),
        ]
        for W, H, S, L, prec, r in cases:
            print(f"no-alloc case {W}x{H}x{S}spp L={L} precision={prec} r={r}", flush=True)
            desc, rec, anc = gen_screen(h, W, H, S, L, prec, seed=W, radius=r)
            a, b = film_pair(h, desc, rec, anc)
            assert np.isfinite(a).all()
            assert np.array_equal(a, b), (W, H, np.abs(a - b).max())
            rec.free()
            if anc is not None:
                anc.free()


def test_noalloc_rejects_slices_beyond_the_scratch():
    """One 4K FP64 slice needs 2 x 3 x 3840 x 2160 x 8 bytes (398 MB) of level images, more
    than the static scratch: rejected before any launch."""
    with N.Handle({"type": "pf", "mode": "screen"}) as h:
        desc = h.screen_desc(3840, 2160, 1, 2, N.FP64, 7)
        dummy = h.alloc(256)
        with pytest.raises(N.NlpfError, match="static"):
            h.screen_run_noalloc(desc, dummy, None, dummy)
        # a single level needs one image per slice (199 MB): accepted limits are per level count
        assert 2 * 3 * 3840 * 2160 * 8 > N.NOALLOC_SCRATCH_BYTES >= 3 * 3840 * 2160 * 8
