"""The LTC fitter (dr_ltc_fit, the reference's fit/fitLTC.cpp) against the
table the renderer embeds (data/ltc_tabM.f32, the reference's ltc.inc): each
entry refitted from the embedded table's neighbouring entry as first guess
reproduces the embedded entry, which fixes the BRDF, the LTC, the
Nelder-Mead search, the first-guess rule and the table layout together (a
transposed or mis-normalised matrix, a wrong alpha or a wrong view angle
fails here)."""
import numpy as np
import pytest

from delta_ray_amd import _abi, ltc_table

N = 128


@pytest.fixture(scope="module")
def ref():
    return ltc_table.embedded()


def _refit(ref, a, t, consistent=False):
    tab, amp = ltc_table.fit(N, rows=(a, a + 1), ts=(t, t + 1), guess=ref, consistent=consistent, threads=1)
    return tab[a + t * N], amp[a + t * N]


@pytest.mark.parametrize("a,t,tol", [(127, 0, 1e-6), (100, 0, 1e-6), (64, 1, 1e-5), (64, 40, 1e-5),
                                     (110, 127, 1e-5), (30, 10, 2e-4), (20, 60, 2e-4)])
def test_refit_matches_embedded_entry(ref, a, t, tol):
    m, amp = _refit(ref, a, t)
    r = ref[a + t * N]
    assert np.abs(m - r).max() <= tol, (m, r)
    assert 0.0 < amp <= 1.0 + 1e-3


def test_embedded_entries_are_isotropic_ltc(ref):
    """the zeroed entries and the unit (2, 2) entry of every stored matrix"""
    assert np.all(ref[:, [1, 3, 5, 7]] == 0.0)
    assert np.all(np.abs(ref[:, 8] - 1.0) < 1e-6)


def test_consistent_lobe_changes_rough_entries(ref):
    """fitting the lobe the renderer evaluates (exponent 2 / r rather than
    the reference's alpha = r^2) changes the table"""
    m, _ = _refit(ref, 64, 40, consistent=True)
    assert np.abs(m - ref[64 + 40 * N]).max() > 1e-2


def test_partial_fit_needs_guess():
    lib = _abi.load_library()
    tab = np.zeros((N * N, 9), np.float32)
    amp = np.zeros(N * N, np.float32)
    import ctypes
    fp = ctypes.POINTER(ctypes.c_float)
    rc = lib.dr_ltc_fit(N, 3, 4, 5, 6, None, 0, 1, tab.ctypes.data_as(fp), amp.ctypes.data_as(fp))
    assert rc == _abi.DR_ERROR_INVALID_ARGUMENT
    assert b"first guesses" in lib.dr_last_error()
