"""The binary64 twin of the path sample (oracle/oracle_fp64.c) against the
binary32 oracle: the same estimator and random numbers with every geometric
and shading quantity in binary64, so converged pixel means must agree within
their statistical error (a binary32 effect larger than the noise - a lost
hit, a self-intersection, a cancellation - would show as a bias).  Both
estimators: the reference's (refCompat 1) and the corrected one (0)."""
import numpy as np
import pytest

import strelka_amd as sa
from oracle import Oracle


def z_stats(m32, v32, m64, v64):
    var = v32 + v64
    ok = var > 0
    z = (m32 - m64)[ok] / np.sqrt(var[ok])
    return float(np.mean(z * z)), float(np.mean(np.abs(z) > 4)), float(np.sqrt(np.mean((m32 - m64) ** 2)))


@pytest.mark.parametrize("ref_compat", [True, False])
def test_fp32_oracle_converges_to_the_fp64_twin(ref_compat):
    s = sa.SceneData.generate(6)
    o = Oracle.from_scene(s, sa.HostBvh.build(s).nodes)
    w, h = s.camera.dimension[0], s.camera.dimension[1]
    spp = 256
    p = sa.render_params(s.camera, spp=spp, max_depth=8, frame=3, sampling=1, sky=s.sky, ref_compat=ref_compat)
    smp = o.pt_frame(p, 0, 0, w, h).reshape(h, w, spp, 3).astype(np.float64)
    m32, v32 = smp.mean(axis=2), smp.var(axis=2, ddof=1) / spp
    # the twin's samples use other frame numbers: independent estimates
    p.pt.frameNumber = 1000
    m64, v64 = o.pt_frame_fp64(p, 0, 0, w, h)
    chi, tail, rmse = z_stats(m32, v32, m64, v64)
    print(f"ref_compat={ref_compat}: mean z^2 {chi:.3f}, |z| > 4 on {tail:.4f}, RMSE of the means {rmse:.5f}, "
          f"mean pixel {m64.mean():.4f}")
    assert 0.6 < chi < 1.5 and tail < 0.005
