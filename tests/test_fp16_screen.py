"""The distance screen of the 16-bit list build (mt_neigh_block_atoms16):
staged x and y rounded to FP16, their differences taken in FP16, the squares
summed with an FP32 z term, compared with rlist^2 widened by

    3 rlist delta + 2 delta^2,   delta = 2 h(2.6 cs) + h(rlist),

h(v) = half an FP16 ulp at v. The screen may keep pairs just beyond rlist but
must never drop one inside it. Restated here with numpy's IEEE binary16 on
worst-case geometry (pairs just inside rlist, at the largest staged
coordinates the block frame holds) for cell widths from 0.3 to 40."""
import numpy as np
import pytest


def half_ulp16(v):
    """2^(exponent(v) - 11), as mt_half_ulp16 builds it from the FP32 bits."""
    e = (np.float32(v).view(np.uint32) >> 23).astype(np.int64) - 127
    return np.float32(2.0 ** (e - 11))


def threshold(rlist, cs):
    delta = np.float32(2.0) * half_ulp16(np.float32(2.6) * np.float32(cs)) + half_ulp16(rlist)
    rl2 = np.float32(rlist) * np.float32(rlist)
    return rl2 + np.float32(3.0) * np.float32(rlist) * delta + np.float32(2.0) * delta * delta


def screen(p, c, rlist, cs):
    """The kernel's test for home positions p[i] against candidates c[i] (block frame)."""
    pxy = p[:, :2].astype(np.float32).astype(np.float16)
    cxy = c[:, :2].astype(np.float32).astype(np.float16)
    dxy = (pxy - cxy).astype(np.float16)                                # v_pk_add_f16
    dz = p[:, 2].astype(np.float32) - c[:, 2].astype(np.float32)
    d2 = dxy.astype(np.float32) ** 2                                   # exact in FP32
    r2 = (d2[:, 0] + d2[:, 1]).astype(np.float32) + dz * dz             # v_dot2_f32_f16
    return r2 < threshold(rlist, cs)


@pytest.mark.parametrize("cs", [0.3, 1.0, 2.8, 2.99, 6.5, 11.2, 40.0])
def test_no_pair_inside_rlist_is_dropped(cs):
    rng = np.random.default_rng(int(cs * 100))
    rlist = cs * 0.999                      # cells are at least rlist wide
    lim = 2.5 * cs                          # |staged x, y| within 2.5 cell widths
    n = 200000
    # home atoms anywhere in the block frame, candidates at distance just
    # below rlist in random directions: the hardest case for the screen
    p = rng.uniform(-lim + rlist, lim - rlist, (n, 3))
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    r = rlist * (1.0 - rng.uniform(0.0, 2e-3, n))
    c = p + u * r[:, None]
    keep = screen(p, c, rlist, cs)
    assert keep.all(), (cs, int((~keep).sum()))


def test_screen_rejects_far_pairs():
    """Pairs 2 % beyond rlist are rejected: the widening is a thin shell."""
    rng = np.random.default_rng(5)
    cs = rlist = 2.8
    p = rng.uniform(-4.0, 4.0, (50000, 3))
    u = rng.normal(size=(50000, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    c = p + u * (1.02 * rlist)
    assert not screen(p, c, rlist, cs).any()


def test_margin_is_small():
    """Under 1.5 % more pairs at the benchmark's cells (2.94 sigma, rlist
    2.8) and at 6.5 A water cells."""
    for cs, rlist in ((2.94, 2.8), (6.52, 6.5)):
        extra = (float(threshold(rlist, cs)) / rlist ** 2) ** 1.5 - 1.0
        assert 0.0 < extra < 0.015, (cs, extra)


@pytest.mark.parametrize("v", [0.3, 1.0, 1.5, 2.8, 3.99, 4.0, 7.28, 16.9, 100.0])
def test_half_ulp16_matches_binary16(v):
    """h(v) is half the spacing of IEEE binary16 numbers at v."""
    assert float(half_ulp16(v)) == float(np.spacing(np.float16(v))) / 2.0
