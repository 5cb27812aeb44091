"""Host side of the serving loop's request building (no GPU): which
reference orbit a request runs on."""
import numpy as np

from fractalshades_amd import serving, settings


def test_shallow_frames_use_the_critical_orbit():
    """Frames wider than newton_zoom_level (1e-5) run around the critical
    orbit with BLA off, as the reference does for every shallow frame
    (perturbation.py:669-672, mandelbrot_M2.py:521-524); deeper ones on the
    centre's resident nucleus orbit with BLA on.  The resident key tells
    which requests may share one device copy of their orbit."""
    c = serving.movie_centre(2000)
    specs = serving.movie_requests(6, 32, 18, 2000, log10_width=(1., -30.))
    keys = set()
    for sp in specs:
        f, cia, c_pix = serving.build_request(sp, [c], "fp64")
        shallow = float(sp.width) > settings.newton_zoom_level
        assert cia.resident_key[0] == ("critical" if shallow else 0)
        assert cia.fields["bla_on"] == int(not shallow)
        if shallow:
            assert cia.fields["ref_order"] == 100
            assert np.all(cia.arrays["Zn_path"].view(np.complex128) == 0)
        else:
            assert cia.fields["ref_len"] == 2001
        keys.add(cia.resident_key)
        assert c_pix.size == 32 * 18
    assert len(keys) == 2
