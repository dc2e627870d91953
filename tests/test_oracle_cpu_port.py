"""The oracle's C++ restatement of the bulge + disk drawing (oracle/cpu_draw.cpp, the CPU
baseline) against the numpy oracle's drawImage on the same stamps: sheared and unsheared
galaxies, odd and even stamp sides, zero and sub-pixel offsets, FFT sizes 2^k and 3 2^k."""
import numpy as np

from galsim_amd import workloads
from oracle import cpu_port
from oracle import draw as od


def test_cpu_port_matches_numpy_oracle():
    n = 18
    p = workloads.bulge_disk_params(n, seed=4)
    p['g1'][5] = p['g2'][5] = 0.                       # diagonal Jacobian: quadrant fills
    profs = [workloads.bulge_disk_profile(od, p, i) for i in range(n)]
    sizes = np.array([48, 64, 33, 50, 96, 128, 41, 70, 160] * 2)
    offs = np.random.default_rng(2).uniform(-0.5, 0.5, (n, 2))
    offs[3] = 0.
    b = cpu_port.Batch(profs, sizes, sizes, 0.2, offs)
    outs = b.draw(3)
    Ns = set()
    for i in range(n):
        want = profs[i].drawImage(int(sizes[i]), int(sizes[i]), scale=0.2, offset=tuple(offs[i]))
        err = np.max(np.abs(outs[i] - want)) / np.max(np.abs(want))
        assert err < 1e-13, (i, err)
        Ns.add(b.recs[i].N)
    assert any(N % 3 == 0 for N in Ns) and any(N % 3 != 0 for N in Ns)
    # one thread gives the same pixels
    again = cpu_port.Batch(profs[:4], sizes[:4], sizes[:4], 0.2, offs[:4]).draw(1)
    for i in range(4):
        np.testing.assert_array_equal(again[i], outs[i])
