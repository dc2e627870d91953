"""GPU parity of the batched path: many stamps of uneven size and FFT size in one call equal the
per-stamp drawImage results (FP64) and the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import galsim_amd as g
from galsim_amd import workloads
from galsim_amd.batch import BatchPlan
from oracle import draw as od


def test_batch_matches_drawimage_and_oracle():
    profs, nx, ny, offs, p = workloads.build('bulge_disk', 24, seed=7)
    plan = BatchPlan(profs, nx, ny, 0.2, offs, precision='f64')
    plan.run()
    plan32 = BatchPlan(profs, nx, ny, 0.2, offs, precision='f32')
    plan32.run()
    assert len(plan.fft_sizes()) >= 2
    for i in range(len(profs)):
        b = plan.stamp(i).cpu().numpy()
        im = profs[i].drawImage(nx=int(nx[i]), ny=int(ny[i]), scale=0.2, dtype=np.float64)
        assert np.array_equal(b, im.array), i
        b32 = plan32.stamp(i).cpu().numpy()
        assert np.max(np.abs(b32 - b)) / np.max(np.abs(b)) < 1e-5, i
    for i in range(0, len(profs), 6):
        ref = workloads.bulge_disk_profile(od, p, i).drawImage(int(nx[i]), int(ny[i]), scale=0.2)
        b = plan.stamp(i).cpu().numpy()
        assert np.max(np.abs(b - ref)) / np.max(np.abs(ref)) < 1e-10, i


def test_batch_exp_moffat_offsets():
    profs, nx, ny, offs, p = workloads.build('exp_moffat', 16, seed=3)
    plan = BatchPlan(profs, nx, ny, 0.2, offs, precision='f64')
    plan.run()
    for i in range(0, 16, 5):
        ref = workloads.exp_moffat_profile(od, p, i).drawImage(64, 64, scale=0.2,
                                                               offset=tuple(offs[i]))
        b = plan.stamp(i).cpu().numpy()
        assert np.max(np.abs(b - ref)) / np.max(np.abs(ref)) < 1e-10, i
