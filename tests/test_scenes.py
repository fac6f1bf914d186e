"""The host scene generators and the device generator's size formula on the
CPU (the device generator itself is compared with these in
tests/test_gpu_scenes.py)."""

import numpy as np
import pytest

from reyn_amd import scenes
from reyn_amd.sim import dam_break_sizes


def _cases():
    yield scenes.config1_sesph()
    yield scenes.config3_iisph(side=24)
    for sc, _ in scenes.config5_jobs(6, seed=12, n_lo=2000, n_hi=20000):
        yield sc


@pytest.mark.parametrize("k", range(8))
def test_device_sizes_match_host_scenes(k):
    sc = list(_cases())[k]
    d = scenes.dam_break_desc(sc)
    nl, nb = dam_break_sizes(d)
    assert nl == int(np.prod(sc.meta["fluid_counts"]))
    assert nb == sc.n_bdy
    assert sc.n <= nl


def test_hash_uniform_known_values():
    """The jitter hash (lowbias32 mixer twice) against values worked by hand in
    32-bit integer arithmetic, and its range."""
    def mix(x):
        x &= 0xFFFFFFFF
        x ^= x >> 16
        x = (x * 0x7FEB352D) & 0xFFFFFFFF
        x ^= x >> 15
        x = (x * 0x846CA68B) & 0xFFFFFFFF
        x ^= x >> 16
        return x
    for seed, idx, axis in ((0, 0, 0), (1, 5, 2), (4021, 123456, 1), (2**32 - 1, 2**24, 0)):
        s = mix((seed * 0x9E3779B9 + 0x7F4A7C15) & 0xFFFFFFFF)
        h = mix(s ^ ((idx * 3 + axis) & 0xFFFFFFFF))
        assert scenes.hash_uniform(seed, np.array([idx]), axis)[0] == (h >> 8) / 2.0**24
    u = scenes.hash_uniform(7, np.arange(100000), 1)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01


def test_wall_shell_by_index():
    """Every wall point lies outside the box in at least one axis and the
    shell of a box whose sides are whole spacings has the closed-form count."""
    s, L = 0.1, 2
    xb = scenes.box_walls((0, 0, 0), (1.0, 0.5, 0.7), s, L)
    n = np.array([10, 5, 7]) + 2 * L
    assert xb.shape[0] == n.prod() - (n - 2 * L).prod()
    lo, hi = np.zeros(3), np.array([1.0, 0.5, 0.7])
    assert np.all(np.any((xb < lo) | (xb > hi), axis=1))
