"""bench.py: more than one GPU is refused with a message (the library runs on
one MI355X), before any GPU work; the tiled HMC start repeats links
correctly."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_request_is_refused():
    p = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_tile_links_repeats_every_plaquette():
    """bench.py's tiled HMC start: the links of a 4x4x6x4 lattice repeated to
    8x4x12x8 give every plaquette of the large lattice the value of its image,
    so the plaquette sum grows by the volume ratio, and each link sits at the
    site whose coordinates reduce to its image's."""
    import numpy as np
    import oracle_lib as ol

    sys.path.insert(0, REPO)
    import bench

    Ls, L = (4, 4, 6, 4), (8, 4, 12, 8)
    small, big = ol.Lattice(Ls), ol.Lattice(L)
    us = small.gauge_exp_random(99, 0.5)
    ub = bench.tile_links(us, Ls, L)
    assert abs(big.plaquette_sum(ub) - 8 * small.plaquette_sum(us)) < 1e-9 * abs(big.plaquette_sum(ub))
    # site (x, y, z, t) = (5, 3, 7, 6) of the large lattice is (1, 3, 1, 2) of the small one
    x, y, z, t = 5, 3, 7, 6
    xs, ys, zs, ts = 1, 3, 1, 2
    jb, js = x + L[0] * (y + L[1] * (z + L[2] * t)), xs + Ls[0] * (ys + Ls[1] * (zs + Ls[2] * ts))
    eo = (x + y + z + t) % 2
    assert np.array_equal(ub[:, eo, jb // 2], us[:, eo, js // 2])
    with pytest.raises(ValueError):
        bench.tile_links(us, Ls, (8, 4, 8, 8))   # 6 does not divide 8


def test_hmc_tile_that_does_not_divide_is_refused():
    p = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--hmc", "1",
                        "--hmc-lattice", "24,24,24,48"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 1
    assert "does not divide" in p.stderr
    assert p.stdout == ""
