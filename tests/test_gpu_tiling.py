"""Launch tiling of the fused stage kernel (cc_tiling): the planes per block
and, for FP32, the tile shape that minimise rounds x (planes + 3), where a
round fills every resident slot (CUs x blocks per CU) once."""
import pytest

from compchan3dmpi_amd.channel import Channel

pytestmark = pytest.mark.gpu

GRIDS = [(64, 32, 48), (96, 64, 48), (160, 96, 80), (256, 128, 192), (34, 128, 192),
         (7, 9, 6), (37, 21, 11)]

# (grid, precision) -> (tile, planes per block, blocks) on a 256-CU MI355X:
# FP64 32x16 holds one block per CU, FP32 32x8 three.
EXPECTED_256 = {
    ((64, 32, 48), "fp64"): ((32, 16), 1, 184),      # 4 tiles x 46 chunks, one round
    ((64, 32, 48), "mixed"): ((32, 8), 1, 368),
    ((96, 64, 48), "fp64"): ((32, 16), 3, 192),      # 12 tiles x 16 chunks
    ((160, 96, 80), "fp64"): ((32, 16), 10, 240),    # 30 tiles x 8 chunks
    ((256, 128, 192), "fp64"): ((32, 16), 48, 256),  # 64 tiles x 4 chunks, one round
    ((256, 128, 192), "mixed"): ((32, 8), 32, 768),  # 128 tiles x 6 chunks
    ((34, 128, 192), "fp64"): ((32, 16), 6, 256),    # an 8-way x slab of the above
}


def cdiv(a, b):
    return -(-a // b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("grid", GRIDS)
def test_tiling_is_the_model_optimum(grid, precision):
    print(f"\n[case] tiling {grid} {precision}", flush=True)
    ch = Channel(*grid, dt=1e-4, precision=precision)
    try:
        t = ch.tiling()
    finally:
        ch.close()
    (tx, ty), kc = t["tile"], t["kchunk"]
    P1, P2, nint = grid[0] - 2, grid[1] - 2, grid[2] - 2
    tiles = cdiv(P1, tx) * cdiv(P2, ty)
    assert t["blocks"] == tiles * cdiv(nint, kc)
    if precision == "fp64":
        assert (tx, ty) == (32, 16)
    slots = t["cus"] * t["blocks_per_cu"]
    assert slots >= 1

    def cost(k):
        return cdiv(tiles * cdiv(nint, k), slots) * (k + 3)

    best = min(cost(k) for k in range(1, nint + 1))
    assert cost(kc) == best
    # ties go to the longer chunk
    assert all(cost(k) > best for k in range(kc + 1, nint + 1))
    if t["cus"] == 256 and (grid, precision) in EXPECTED_256:
        assert ((tx, ty), kc, t["blocks"]) == EXPECTED_256[(grid, precision)], t


@pytest.mark.parametrize("precision", ["fp64", "fp64-strict", "mixed"])
def test_profile_sampling_counts_and_leaves_results_alone(precision):
    """cc_profile: mode 1 times both stage launches of every step, mode -k
    those of the steps whose number is a multiple of k; the timing events
    change no result bit."""
    import numpy as np

    from compchan3dmpi_amd import formats

    n = (24, 20, 16)
    v0 = formats.laminar_ic(*n)
    out = []
    for mode in (0, 1, -5):
        ch = Channel(*n, dt=1e-3, istat=5, precision=precision)
        try:
            ch.upload(v0)
            assert ch.step(3) == 0
            if mode:
                ch.profile(mode)
            assert ch.step(20) == 0  # steps 4..23
            ms, launches = ch.profile(0)
            out.append(ch.download()[0])
        finally:
            ch.close()
        if mode == 1:
            assert launches == 40 and ms > 0.0
        elif mode == -5:
            assert launches == 2 * 4 and ms > 0.0  # steps 5, 10, 15, 20
        else:
            assert launches == 0
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
