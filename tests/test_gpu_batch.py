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


def _peak_rel_all(plan_a, plan_b, n):
    worst = 0.
    for i in range(n):
        a = plan_a.stamp(i).double()
        b = plan_b.stamp(i).double()
        worst = max(worst, float((a - b).abs().max() / b.abs().max()))
    return worst


@pytest.mark.parametrize("precision,tol", [("f32", 2e-6), ("f64", 1e-12)])
def test_work_list_schedule_matches_per_size_launches(precision, tol):
    """One fold and one row launch over a work list of every FFT size (default tiles, tuned
    tiles, a persistent grid) against the per-size launches of the same batch."""
    profs, nx, ny, offs, _ = workloads.build('bulge_disk', 40, seed=5)
    ref = BatchPlan(profs, nx, ny, 0.2, offs, precision=precision)
    ref.run()
    assert len(ref.fft_sizes()) >= 2
    plan = BatchPlan(profs, nx, ny, 0.2, offs, precision=precision)
    plan.use_schedule()
    plan.run()
    assert _peak_rel_all(plan, ref, len(profs)) < tol
    plan.use_schedule(cols={N: max(1, BatchPlan.default_cols(N) // 2) for N in ref.fft_sizes()},
                      rows={N: BatchPlan.default_rows(N) * 2 for N in ref.fft_sizes()},
                      fold_grid=37, rows_grid=11)
    plan.out.zero_()
    plan.run()
    assert _peak_rel_all(plan, ref, len(profs)) < tol
    tuned = plan.autotune(reps=2)
    print("tuned:", {k: tuned[k] for k in ('cols', 'rows', 'grid', 'ms')})
    plan.out.zero_()
    plan.run()
    assert _peak_rel_all(plan, ref, len(profs)) < tol


def test_graph_replay_matches_eager():
    """The batch captured in a HIP graph: every replay (with eager work and synchronisation
    between replays) equals the eager result."""
    import torch
    profs, nx, ny, offs, _ = workloads.build('bulge_disk', 24, seed=9)
    eager = BatchPlan(profs, nx, ny, 0.2, offs, precision='f32')
    eager.use_schedule()
    eager.run()
    ref = eager.out.clone()
    plan = BatchPlan(profs, nx, ny, 0.2, offs, precision='f32')
    plan.use_schedule()
    plan.capture()
    other = BatchPlan(profs[:5], nx[:5], ny[:5], 0.2, offs[:5], precision='f32')
    for k in range(3):
        plan.out.fill_(float('nan'))
        torch.cuda.synchronize()
        plan.replay()
        torch.cuda.synchronize()
        assert torch.equal(plan.out, ref), k
        other.run()                     # eager work between replays
        torch.cuda.synchronize()


def test_work_list_skips_malformed_items():
    """Items naming a stamp outside the batch or an empty range are skipped by the kernels."""
    import torch
    from galsim_amd import _lib
    profs, nx, ny, offs, _ = workloads.build('exp_moffat', 3, seed=1)
    plan = BatchPlan(profs, nx, ny, 0.2, offs, precision='f32')
    plan.use_schedule()
    plan.run()
    ref = plan.out.clone()
    items, sc = plan._sched
    bad = torch.tensor([[7, 0, 4, 4], [0, 0, 0, 4], [-1, 0, 1, 1]], dtype=torch.int32,
                       device='cuda')
    fold = torch.cat([bad, items[:sc.nfold]])
    rows = torch.cat([bad, items[sc.nfold:]])
    allitems = torch.cat([fold, rows]).contiguous()
    sc2 = _lib.Schedule()
    sc2.fold = allitems.data_ptr()
    sc2.rows = allitems.data_ptr() + len(fold) * 16
    sc2.nfold, sc2.nrows = len(fold), len(rows)
    sc2.fold_grid, sc2.rows_grid = len(fold), len(rows)
    sc2.fold_lds, sc2.rows_lds = sc.fold_lds, sc.rows_lds
    sc2.program = sc.program                        # the same (specialised) fold kernel
    plan._sched = (allitems, sc2)
    plan.out.zero_()
    plan.run()
    assert torch.equal(plan.out, ref)


def test_large_sersic_stamps_fp32_against_fp64_and_oracle():
    """Sersic n in 3..6 with hlr 2..6" under Kolmogorov at drawImage's own stamp sizes: FFT grids
    of 768 to 3072 (those above 2048 in FP32 only).  FP32 within 1e-5 of the peak of the FP64
    kernels; FP64 within 1e-10 of the oracle on the smallest stamp."""
    profs, nx, ny, offs, p = workloads.build('large_sersic', 6, seed=4)
    plan32 = BatchPlan(profs, nx, ny, 0.2, offs, precision='f32')
    plan32.use_schedule()
    plan32.run()
    sizes = plan32.fft_sizes()
    print("FFT sizes", sizes)
    assert any(N >= 1024 for N in sizes)
    small = [i for i in range(len(profs)) if int(plan32.recs['N'][np.nonzero(
        plan32.order == i)[0][0]]) <= 2048]
    plan64 = BatchPlan([profs[i] for i in small], nx[small], ny[small], 0.2, offs[small],
                       precision='f64')
    plan64.run()
    for k, i in enumerate(small):
        a = plan32.stamp(i).double()
        b = plan64.stamp(k)
        err = float((a - b).abs().max() / b.abs().max())
        print("stamp", i, "n", p['n'][i], "hlr %.2f" % p['hlr'][i], "peak-rel FP32 err %.2e" % err)
        assert err < 1e-5, (i, err)
    i = min(small, key=lambda i: int(nx[i]))
    ref = workloads.large_sersic_profile(od, p, i).drawImage(int(nx[i]), int(ny[i]), scale=0.2)
    b = plan64.stamp(small.index(i)).cpu().numpy()
    assert np.max(np.abs(b - ref)) / np.max(np.abs(ref)) < 1e-10


def test_plain_fp64_sweep_256_against_oracle():
    """The plain FP64 workload: 256 seeded Exponential x Moffat stamps (hlr ~ U(0.8, 2.5)",
    |g| ~ U(0, 0.5), sub-pixel offsets), 64x64 at 0.2"/px, batched on the GPU, every stamp
    within 1e-10 of the peak of the oracle's drawImage."""
    profs, nx, ny, offs, p = workloads.build('exp_moffat', 256, seed=0)
    plan = BatchPlan(profs, nx, ny, 0.2, offs, precision='f64')
    plan.use_schedule()
    plan.run()
    worst = 0.
    for i in range(256):
        ref = workloads.exp_moffat_profile(od, p, i).drawImage(64, 64, scale=0.2,
                                                               offset=tuple(offs[i]))
        b = plan.stamp(i).cpu().numpy()
        worst = max(worst, np.max(np.abs(b - ref)) / np.max(np.abs(ref)))
    print("256-stamp FP64 sweep: worst peak-relative error %.2e" % worst)
    assert worst < 1e-10


def test_batch_of_added_profiles_matches_drawimage():
    # the work-list kernels run every leaf and operator through the program interpreter: a
    # batch mixing Spergel, Airy, the inclined disks and a deconvolution equals drawImage
    def make(i):
        k = i % 5
        if k == 0:
            return g.Convolve([g.Spergel(0.3 + 0.1 * i, half_light_radius=0.6, flux=50.),
                               g.Kolmogorov(fwhm=0.7)])
        if k == 1:
            return g.Airy(lam_over_diam=0.4 + 0.02 * i, obscuration=0.2, flux=20.)
        if k == 2:
            return g.Convolve([g.InclinedExponential((0.3 + 0.1 * i) * g.radians,
                                                     half_light_radius=0.8, flux=30.),
                               g.Moffat(3., fwhm=0.7)])
        if k == 3:
            return g.Convolve([g.InclinedSersic(2.0, 1.0 * g.radians, half_light_radius=0.5,
                                                flux=30.), g.Kolmogorov(fwhm=0.6)])
        psf = g.Gaussian(sigma=0.3)
        gal = g.Exponential(half_light_radius=0.7, flux=40.).shear(g1=0.1, g2=0.2)
        return g.Convolve([gal, psf, g.Moffat(3., fwhm=0.8), g.Deconvolve(psf)])
    profs = [make(i) for i in range(10)]
    nx = np.array([40 + 4 * (i % 3) for i in range(10)])
    ny = np.array([44 - 4 * (i % 2) for i in range(10)])
    offs = np.zeros((10, 2))
    for prec, tol in (('f64', 1e-12), ('f32', 1e-5)):
        plan = BatchPlan(profs, nx, ny, 0.2, offs, precision=prec)
        plan.run()
        for i in range(10):
            im = profs[i].drawImage(nx=int(nx[i]), ny=int(ny[i]), scale=0.2, dtype=np.float64)
            b = plan.stamp(i).double().cpu().numpy()
            err = np.max(np.abs(b - im.array)) / np.max(np.abs(im.array))
            print(prec, i, err)
            assert err < tol, (prec, i, err)


def test_bad_stamps_are_flagged_not_fatal_and_overflow_reruns_in_fp64():
    # A batch with a NaN half-light radius (cannot be planned: flagged, drawn as zeros), an
    # FP32 overflow (flux 3e38: the column transform's sums exceed the FP32 range, the kernels
    # flag the stamp's non-finite pixels, rerun_flagged redraws it in FP64), and good stamps,
    # which come out as they do in a batch of their own.
    import torch
    from galsim_amd import batch
    from galsim_amd import _lib as L

    def bd(m, hlr, flux):
        gal = m.Exponential(half_light_radius=hlr, flux=flux).shear(g1=0.1, g2=0.2)
        return m.Convolve([gal, m.Kolmogorov(fwhm=0.8)])
    profs = [bd(g, 1.0, 1e3), bd(g, float('nan'), 1e3), bd(g, 0.4, 3e38), bd(g, 1.4, 2e3)]
    print("status batch, 4 stamps of 48x48, FP32", flush=True)
    plan = batch.BatchPlan(profs, 48, 48, 0.2, precision='f32')
    plan.run()
    st = plan.status()
    print("status", st)
    assert st[1] & batch.STAMP_BAD_PARAMS and not st[1] & L.STAMP_NONFINITE
    assert st[2] & L.STAMP_NONFINITE
    assert st[0] == 0 and st[3] == 0
    assert torch.count_nonzero(plan.stamp(1)) == 0
    ref = batch.BatchPlan([profs[0], profs[3]], 48, 48, 0.2, precision='f32')
    ref.run()
    assert torch.equal(plan.stamp(0), ref.stamp(0)) and torch.equal(plan.stamp(3), ref.stamp(1))
    redo = plan.rerun_flagged()
    assert list(redo) == [2]
    st = plan.status()
    assert st[2] & batch.STAMP_RERUN_FP64 and not st[2] & L.STAMP_NONFINITE
    a = plan.stamp(2).double().cpu().numpy()
    want = od.Convolve([od.Exponential(half_light_radius=0.4, flux=3e38).shear(g1=0.1, g2=0.2),
                        od.Kolmogorov(fwhm=0.8)]).drawImage(48, 48, scale=0.2)
    assert np.all(np.isfinite(a))
    assert np.max(np.abs(a - want)) < 1e-6 * np.max(np.abs(want))
    # a later run clears the device flags and the rerun marks
    plan.run()
    assert plan.status()[2] & L.STAMP_NONFINITE and not plan.status()[2] & batch.STAMP_RERUN_FP64


def _to_oracle(obj):
    """The oracle object of a galsim_amd profile tree (same parameters)."""
    if isinstance(obj, g.Transformation):
        return od.Transformation(_to_oracle(obj._original), obj._jac, (obj._dx, obj._dy),
                                 obj._flux_ratio)
    if isinstance(obj, g.Convolution):
        return od.Convolve(*[_to_oracle(o) for o in obj.obj_list])
    if isinstance(obj, g.Sum):
        return od.Add(*[_to_oracle(o) for o in obj.obj_list])
    if isinstance(obj, g.Exponential):
        return od.Exponential(scale_radius=obj.scale_radius, flux=obj.flux)
    if isinstance(obj, g.Sersic):
        return od.Sersic(obj.n, scale_radius=obj.scale_radius, flux=obj.flux)
    if isinstance(obj, g.Kolmogorov):
        return od.Kolmogorov(lam_over_r0=obj.lam_over_r0, flux=obj.flux)
    raise TypeError(type(obj))


def test_config_batch_matches_per_object_batch_and_oracle():
    # a config-driven batch (vectorised planning) draws what the per-object plan of the same
    # config values draws, and a sample of its stamps matches the oracle
    import torch
    from galsim_amd import config
    from tests.test_config_cpu import bulge_disk_config
    n = 256
    cfg = bulge_disk_config(n)
    profs, nx, ny, scale, offs, method = config.build_profiles(cfg, n, seed=3)
    for prec, tol_plan, tol_oracle in (('f64', 1e-12, 1e-10), ('f32', 1e-6, 1e-5)):
        print("config batch", prec, n, "stamps", flush=True)
        cb = config.build_batch(cfg, n, prec, seed=3)
        cb.run()
        pb = BatchPlan(profs, nx, ny, scale, offs, precision=prec)
        pb.run()
        torch.cuda.synchronize()
        assert np.all(cb.status() == 0)
        worst = max(float((cb.stamp(i) - pb.stamp(i)).abs().max() / pb.stamp(i).abs().max())
                    for i in range(n))
        print("config vs per-object plan, max peak-relative difference", worst)
        assert worst < tol_plan
        wo = 0.
        for i in range(0, n, 37):
            ref = _to_oracle(profs[i]).drawImage(int(nx[i]), int(ny[i]), scale=scale,
                                                 offset=tuple(offs[i]))
            a = cb.stamp(i).double().cpu().numpy()
            wo = max(wo, float(np.max(np.abs(a - ref)) / np.max(np.abs(ref))))
        print("config batch vs oracle, max peak-relative error", wo)
        assert wo < tol_oracle
