"""The LDS bank model behind the 384-point z-line exchange layouts (pyspectral_amd.lds_banks).

Pinned against what fixes it independently: a conflict-free access pattern costs one cycle per
half-wave, a pattern with every lane on one bank costs 32, and the padded layout's modelled
conflict share reproduces the 38 % that SQ counters measured for the kernel before the
swizzles (profiles/counters_2d_8192_3d_384.md records the 29 % after them)."""

from pyspectral_amd import lds_banks as B


def test_half_wave_cycles_extremes():
    assert B.half_wave_cycles(list(range(32))) == 1          # 64 dwords on 64 banks
    assert B.half_wave_cycles([32 * i for i in range(32)]) == 32  # every lane on banks 0-1
    assert B.half_wave_cycles([0] * 32) == 1                  # one address: a broadcast


def test_swizzles_are_row_permutations():
    for f in B.SWIZZLES_384.values():
        assert sorted(f(e) for e in range(384)) == list(range(384))


def test_swizzled_exchanges_are_conflict_free_and_the_pad_reproduces_the_counters():
    L, nt, e, fwd = 384, 128, 12, [4, 4, 2, 12]
    for (inv, x), f in B.SWIZZLES_384.items():
        plan = list(reversed(fwd)) if inv else fwd
        rw, ns, rr = B.plan_exchanges(plan)[x]
        w, r, ideal = B.exchange_cycles(L, nt, e, rw, ns, rr, f)
        assert w + r == ideal
        pw, pr, _ = B.exchange_cycles(L, nt, e, rw, ns, rr, B.pad32)
        assert pw + pr > ideal
    padded = B.tile_cycles(L, nt, e, fwd, 2, 3, lambda inv, x: B.pad32)
    share = 1 - padded["ideal"] / padded["cycles"]
    assert abs(share - 0.38) < 0.01
    cur = B.tile_cycles(L, nt, e, fwd, 2, 3,
                        lambda inv, x: B.SWIZZLES_384.get((inv, x), B.pad32))
    assert cur["cycles"] < 0.75 * padded["cycles"]
