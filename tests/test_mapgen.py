"""The volume-map generator on the GPU against the host builder.

Both evaluate the same analytic SDF and the same float64 quadrature in the
same order; the stored float32 maps must agree to the last bit except where
a float64 rounding difference of the SDF's square roots crosses a float32
rounding boundary (none is expected; the bound below allows one ulp).
"""
import numpy as np
import pytest

from pbf_host import build_volume_map
from pairs import ulp_diff

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,res,lo,hi,h,r,shape", [
    (0, (24, 20, 16), (0.0, 0.0, 0.0), (1.6, 1.6, 1.2), 0.1, 0.025, None),
    (0, (20, 20, 20), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.025, 0.0, None),
    (1, (32, 32, 32), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.05, 0.0125, [0.5, 0.5, 0.5, 0.2, 0.5, 0.5, 0.5, 0.3, 0.08]),
    # the reference's own volume definition (PBF_VOLUME_REFERENCE = 0x100)
    (0x100, (20, 20, 20), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.025, 0.0, None),
    (0x101, (24, 24, 24), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.025, 0.0, [0.5, 0.45, 0.5, 0.12, 0.5, 0.45, 0.5, 0.25, 0.05]),
])
def test_gpu_map_matches_host_builder(kind, res, lo, hi, h, r, shape):
    print(f"[mapgen] kind={kind} res={res}", flush=True)
    g = build_volume_map(kind, res, lo, hi, h, shape, r, where="gpu")
    c = build_volume_map(kind, res, lo, hi, h, shape, r, where="host")
    assert np.all(np.isfinite(g))
    assert ulp_diff(g[:, 0], c[:, 0]) <= 1
    assert ulp_diff(g[:, 1], c[:, 1]) <= 1
    assert (g[:, 1] > 0).sum() == (c[:, 1] > 0).sum()
