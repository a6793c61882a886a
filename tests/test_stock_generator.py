"""Encoding error with the counter-based generator against the unmodified
reference drawing from libc's rand() (oracle/_ref/libs2tc_ref_stock.so).

The bit-exact parity tests compare against the reference with the same
generator; this checks that replacing the generator costs no quality: the
total colour distance of both encodings (oracle_colordist_error) agrees
within a stated tolerance.  Skipped where the stock reference was not built.
"""
import numpy as np
import pytest

import imgutil
from oracle import StockReference, stock_reference_available

# relative difference of total colour distance allowed between the two
# generators on the images below (measured: below 0.5%)
TOL = 0.02


@pytest.mark.skipif(not stock_reference_available(), reason="stock reference (oracle/_ref) not built here")
@pytest.mark.parametrize("cd,dxt,nrandom,refine", [(5, 0, 16, 1), (7, 0, 64, 2), (3, 2, 8, 2), (0, 1, 32, 1)])
def test_error_against_stock_rand(oracle, cd, dxt, nrandom, refine):
    st = StockReference()
    rng = np.random.default_rng(cd * 100 + nrandom)
    w, h = 48, 32
    rgba = imgutil.smooth_rgba8(rng, w, h)
    q = oracle.rgb565_image(rgba, w, h, 4, imgutil.ALPHABITS[dxt], 0)
    ours = oracle.encode_quantized(q, w, h, dxt, cd, nrandom, refine, seed=0)
    st.srand(1)
    stock = st.encode_quantized(q, w, h, dxt, cd, nrandom, refine)
    e_ours = oracle.colordist_error(q, w, h, dxt, cd, ours)
    e_stock = oracle.colordist_error(q, w, h, dxt, cd, stock)
    assert abs(e_ours - e_stock) <= TOL * max(abs(e_stock), 1), (e_ours, e_stock)
