"""pdem_pair_log_noalloc: the training-row log staged in the library's fixed
device array instead of buffers allocated per call.  Its rows must equal
pdem_pair_log's (compared as sorted row sets: both append rows through an
atomic counter), on two calls in a row, the second on a smaller model, and a
capacity above the staging maximum must be refused."""
import numpy as np
import pytest

from peridem_amd import capi, configs

pytestmark = pytest.mark.gpu


def _ensemble(n, seed):
    # equal radii on the tightest grid: particles touch within a few hundred steps
    return configs.ensemble_2d(n_instances=n, n_min=2, n_max=6, ratios=(6,), speed_lo=2.0, speed_hi=4.0,
                               pitch_factor=1.0, R_lo=1.0e-3, R_hi=1.0e-3, seed=seed)


def _rows(eng, noalloc, cap=4096):
    f, t, b = eng.pair_log(band_h=0.5, cap=cap, baseline=True, noalloc=noalloc)
    r = np.concatenate([f, t, b], axis=1)
    return r[np.lexsort(r.T[::-1])]


@pytest.mark.parametrize("precision", [capi.PREC_FP64])
def test_pair_log_noalloc_equals_pair_log(precision):
    counts = []
    for n_inst, seed, steps in ((6, 11, 300), (2, 5, 300)):
        spec = _ensemble(n_inst, seed)
        print(f"[noalloc] ensemble of {n_inst} (seed {seed}): {spec.n_nodes} nodes, {steps} steps", flush=True)
        e = capi.GpuEngine(spec, precision)
        e.closure_setup()
        e.integrate_begin()
        e.integrate_cd_graph(steps)
        a = _rows(e, noalloc=False)
        b = _rows(e, noalloc=True)
        print(f"[noalloc] rows: pair_log {len(a)}, pair_log_noalloc {len(b)}", flush=True)
        assert len(a) > 0, "no pair near contact in the test window"
        np.testing.assert_array_equal(a, b)
        # a capacity below the row count keeps one row; the next call starts
        # from a cleared counter again
        f1, _, _ = e.pair_log(band_h=0.5, cap=1, baseline=True, noalloc=True)
        assert len(f1) == 1
        np.testing.assert_array_equal(_rows(e, noalloc=True), a)
        with pytest.raises(RuntimeError, match="staging maximum"):
            e.pair_log(band_h=0.5, cap=capi.GpuEngine.PAIR_LOG_NOALLOC_MAX_ROWS + 1, noalloc=True)
        counts.append(len(a))
        e.close()
    assert counts[1] < counts[0], f"the second call must log fewer rows: {counts}"
