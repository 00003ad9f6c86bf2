"""cc_diag_noalloc (the bulk diagnostics with no device allocation: partial
sums, results and the file's x3 in the library's static per-GPU scratch)
against cc_diag, which allocates: two calls in a row, the second on a smaller
grid, each before the first step (the uploaded file) and after steps (the
device state read node by node)."""
import numpy as np
import pytest

from compchan3dmpi_amd import formats
from compchan3dmpi_amd.channel import CCError, Channel

pytestmark = pytest.mark.gpu


def mesh_file(n1, n2, n3):
    x1, x2, x3 = formats.mesh(n1, n2, n3)
    return np.round(x3, 6), round(x2[-1], 6)


@pytest.mark.parametrize("precision", ["fp64", "fp64-strict", "mixed"])
def test_diag_noalloc_matches_diag(precision):
    for n, steps in (((64, 32, 48), 10), ((24, 20, 16), 5)):
        print(f"\n[case] diag_noalloc {n} {precision}", flush=True)
        x3f, x2e = mesh_file(*n)
        ch = Channel(*n, dt=1e-3, istat=5, precision=precision)
        try:
            ch.upload(formats.laminar_ic(*n))
            for s in (0, steps):
                if s:
                    assert ch.step(s) == 0
                a = ch.diag_noalloc(x3f, x2e)
                b = ch.diag(x3f, x2e)
                same = int(np.count_nonzero(a != b))
                print(f"  after {s} steps: {same} of 16 values differ", flush=True)
                assert np.allclose(a, b, rtol=1e-12, atol=0.0), (a, b)
        finally:
            ch.close()


def test_diag_noalloc_rejects_tall_grids():
    ch = Channel(8, 8, 4097, dt=1e-4, precision="fp64")
    try:
        ch.upload(formats.laminar_ic(8, 8, 4097))
        with pytest.raises(CCError, match="static scratch"):
            ch.diag_noalloc(np.zeros(4097), 1.0)
    finally:
        ch.close()
