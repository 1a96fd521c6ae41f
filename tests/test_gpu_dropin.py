"""The drop-in ``phasing`` object on the GPU: the reference's call sequences
(examples/basic_phasing_example.py:17-43, advanced_phasing_example.py:20-60)."""
import numpy as np
import pytest

from als_speckle_amd import synth
from als_speckle_amd.phasing import align_global_phase, phasing
from oracle import phasing_c128 as O

pytestmark = pytest.mark.gpu


def test_seed_reproduces_numpy_legacy_draws():
    f = synth.make_frame(128, 0, shape="disk", radius=20, photons=1e6)
    r = phasing()
    r.load(modulus=f.modulus, support=f.support, numtrials=2)
    r.seed(17)
    got = r.get(r.psi_in)
    np.random.seed(17)
    want = (np.random.rand(128, 128) + 1j * np.random.rand(128, 128)).astype(np.complex64) * f.support
    assert np.array_equal(got, want.astype(np.complex64))


def test_basic_example_flow():
    f = synth.make_frame(128, 1, shape="disk", radius=20, photons=1e6)
    r = phasing()
    r.load(modulus=f.modulus)
    r.load(support=f.support, numtrials=3)
    assert r.can_iterate
    assert (r.r0, r.rows) == tuple(int(v) for v in (O.bounding_box(f.support, pad=10)[0],
                                                    np.diff(O.bounding_box(f.support, pad=10)[:2])[0]))
    finals = []
    for t in range(3):
        r.seed()
        r.iterate(40, order=(("hio", 30), ("er", 10)), beta=0.9)
        finals.append(r.get(r.psi_in))
        tr = r.get_error_trace()
        assert tr.shape == (40,) and np.all(np.isfinite(tr))
    buf = r.get(r.savebuffer)
    assert buf.shape == (4, r.rows, r.cols)
    for t in range(3):
        assert np.array_equal(buf[t], finals[t][r.r0:r.r0 + r.rows, r.c0:r.c0 + r.cols])
    al = align_global_phase(buf[:3])
    want = O.align_global_phase(buf[:3].astype(np.complex128))
    assert np.linalg.norm(al - want) / np.linalg.norm(want) < 1e-5
    assert r.numtrial == 3
    with pytest.raises(AssertionError):
        r.seed()
        r.iterate(1)  # fourth trial: savebuffer has numtrials + 1 = 4 slots -> ok
        r.seed()
        r.iterate(1)  # fifth: out of slots


def test_reload_support_resets_trial_counter():
    f = synth.make_frame(64, 2, shape="disk", radius=10, photons=1e6)
    r = phasing()
    r.load(modulus=f.modulus, support=f.support, numtrials=2)
    r.seed(1)
    r.iterate(5)
    assert r.numtrial == 1
    small = f.support[16:48, 16:48]
    r.load(support=small)         # embedded top-left, square bbox pad 4
    assert r.numtrial == 0
    assert r.get(r.support).shape == (64, 64)


def test_shrinkwrap_keyword_shrinks_support():
    f = synth.make_frame(256, 3, shape="polygon", fill=0.1, photons=1e7)
    r = phasing()
    r.load(modulus=f.modulus, support=f.support, numtrials=1)
    r.seed(2)
    before = r.get(r.support).sum()
    r.iterate(60, shrinkwrap=True)
    after = r.get(r.support)
    assert after.sum() <= before and np.all(after <= f.support)


def test_progress_printing(capsys):
    f = synth.make_frame(64, 4, shape="disk", radius=10, photons=1e6)
    r = phasing()
    r.load(modulus=f.modulus, support=f.support, numtrials=1)
    r.seed(3)
    r.iterate(10, silent=5)
    out = capsys.readouterr().out
    assert "iteration 5" in out and "iteration 10" in out
    assert r.get_error_trace().shape == (10,)


def test_inputs_from_fits_and_png_files(tmp_path):
    """load() with a modulus read from a FITS file and a support read from a
    PNG (0/255 greyscale) and from a ds9 region file through io.open
    (speckle/io.py:35-91): the same iterations as from the arrays."""
    from PIL import Image
    from als_speckle_amd import io as sio
    f = synth.make_frame(128, 9, shape="disk", radius=20, photons=1e6)
    sio.writefits(str(tmp_path / "m.fits"), f.modulus, overwrite=True)
    Image.fromarray((f.support * 255).astype(np.uint8), "L").save(tmp_path / "s.png")
    m = sio.open(str(tmp_path / "m.fits"))
    s = sio.open(str(tmp_path / "s.png"))
    assert s.dtype == np.uint8 and np.array_equal(s > 0, f.support > 0)
    # a disk region of radius 23 at the frame centre (1-based ds9 position)
    reg = tmp_path / "s.reg"
    reg.write_text("# Filename: %s\ncircle(65,65,23)\n" % (tmp_path / "m.fits"))
    sreg = sio.open(str(reg))
    assert sreg.shape == (128, 128)
    runs = []
    for mod, sup in ((f.modulus, f.support), (m, s), (f.modulus, sreg.astype(np.float32))):
        r = phasing()
        r.load(modulus=mod, support=sup, numtrials=1)
        r.seed(4)
        r.iterate(12, order=(("hio", 10), ("er", 2)), beta=0.9)
        runs.append((r.get(r.psi_in), r.get_error_trace()))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert np.all(np.isfinite(runs[2][1])) and runs[2][1][-1] < runs[2][1][0]


def test_debug_on_writes_the_reference_snapshots(tmp_path, monkeypatch):
    """iterate(debug_on=(2,)) writes _iteration's FITS snapshots of iteration
    2 with the reference's names (phasing.py:1193-1226, 1244-1246): Psi, the
    divisor and the replaced Psi (fftshifted, magnitude and phase), psi_out
    (real and imaginary) and the HIO update; compared with the oracle's
    transform, replacement and inverse of the same iterate, and the iterations
    themselves are those of a run without snapshots."""
    from als_speckle_amd import io as sio
    monkeypatch.chdir(tmp_path)
    f = synth.make_frame(128, 21, shape="disk", radius=20, photons=1e6)

    def hio_run(calls, dbg=None):
        r = phasing()
        r.load(modulus=f.modulus, support=f.support, numtrials=1)
        r.seed(5)
        for i, k in enumerate(calls):
            last = i == len(calls) - 1
            r.iterate(k, order=(("hio", 1),), beta=0.9, debug_on=dbg if last else None)
        return r.get(r.psi_in), r.get_error_trace()

    # the snapshot call is the second of two: its iteration 2 is the fifth
    # iteration of the run, so Psi is formed from the iterate after four
    runs = [hio_run((2, 4)), hio_run((2, 4), dbg=(2,))]
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    before = hio_run((4,))[0].astype(np.complex128)
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == sorted(["gpu i2 psi_fourier_1_mag.fits", "gpu i2 psi_fourier_1_phase.fits",
                            "gpu i2 fourier_div_mag.fits", "gpu i2 fourier_div_phase.fits",
                            "gpu i2 psi_fourier_2_mag.fits", "gpu i2 psi_fourier_2_phase.fits",
                            "gpu i2 out_real.fits", "gpu i2 out_imag.fits",
                            "gpu i2 hio update_real.fits", "gpu i2 hio update_imag.fits"])
    s = np.fft.fftshift
    Psi = O.fft2(before)
    out, _ = O.fourier_constraint(before, f.modulus)
    Pp = O.fft2(out)
    rd = lambda name: sio.openfits(str(tmp_path / ("gpu i2 " + name + ".fits")))
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(rd("psi_fourier_1_mag"), s(np.abs(Psi))) < 1e-5
    assert rel(rd("fourier_div_mag"), s(np.abs(Psi))) < 1e-5
    assert rel(rd("psi_fourier_2_mag"), s(np.abs(Pp))) < 1e-5
    assert rel(rd("out_real") + 1j * rd("out_imag"), out) < 1e-5
    # the update: the iterate after the snapshot iteration (five in all)
    assert np.array_equal(rd("hio update_real") + 1j * rd("hio update_imag"), hio_run((5,))[0])
