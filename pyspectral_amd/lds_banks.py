"""Bank-conflict model of the LDS exchanges of the row transforms (RowsCT in psx_fft_ct.h).

    python -m pyspectral_amd.lds_banks            # the 384-point z-line tiles of the 3D step

A Stockham pass of radix R with stride Ns (the product of the earlier radices) writes the R
outputs of a butterfly Ns elements apart; the next pass of radix R' reads R' elements L / R'
apart.  LDS on gfx950 has 64 banks of 4 bytes; a wave's 64-bit access is served per 32-lane
half-wave, and a half-wave costs as many cycles as the largest number of distinct dwords it puts
on one bank (1 when its 64 dwords fall on 64 banks).  This model counts those cycles for the
writes and reads of each exchange of a plan, for a layout given as a function of the element
index within a row (rows are L apart in whole half-waves, so the row pitch does not enter).

The layouts that RowsCT<384, 4>::xaddr uses were chosen with it: the padded layout e + e / 32
costs 2304 cycles per tile of four rows (two inverse and three forward transforms), 37.5 % of
them conflicts, against the 38 % bank-conflict share of LDS time that SQ counters measured for
the kernel; the three XOR swizzles bring the exchanges to 1632 cycles, 11.8 % conflicts (the
counters then read 29 % over all of the kernel's LDS accesses, which include the pass-twiddle
reads that this model leaves out).
"""

from __future__ import annotations

import argparse
from collections import defaultdict
from collections.abc import Callable, Sequence

Layout = Callable[[int], int]


def pad32(e: int) -> int:
    """The padded row layout of RowsCT for 384- and 8192-point rows."""
    return e + (e >> 5)


# the exchange layouts of RowsCT<384, 4>::xaddr: (inverse?, exchange index) -> layout
SWIZZLES_384: dict[tuple[bool, int], Layout] = {
    (True, 0): lambda e: e ^ ((e >> 5) & 3),
    (False, 1): lambda e: e ^ ((e >> 3) & 15),
    (False, 2): lambda e: e ^ ((e >> 1) & 16),
}


def half_wave_cycles(addrs: Sequence[int]) -> int:
    """Cycles of one 32-lane half-wave of 64-bit accesses at element addresses addrs."""
    bank: dict[int, set[int]] = defaultdict(set)
    for a in addrs:
        for w in (2 * a, 2 * a + 1):
            bank[w % 64].add(w)
    return max(len(s) for s in bank.values())


def exchange_cycles(L: int, nt: int, e: int, rw: int, ns: int, rr: int,
                    f: Layout) -> tuple[int, int, int]:
    """(write, read, conflict-free) cycles of one work-group of nt threads (e elements each, rows of
    L) for the exchange between a pass of radix rw and stride ns and the next pass of radix rr."""
    w = r = ideal = 0
    for h0 in range(0, nt, 32):
        lanes = range(h0, h0 + 32)
        lw, lr = L // rw, L // rr
        for b in range(e // rw):
            for q in range(rw):
                ad = []
                for tid in lanes:
                    j = (tid + b * nt) % lw
                    k = j % ns
                    ad.append(f((j - k) * rw + k + q * ns))
                w += half_wave_cycles(ad)
                ideal += 1
        for b in range(e // rr):
            for m in range(rr):
                r += half_wave_cycles([f((tid + b * nt) % lr + m * lr) for tid in lanes])
                ideal += 1
    return w, r, ideal


def plan_exchanges(radices: Sequence[int]) -> list[tuple[int, int, int]]:
    """(rw, ns, rr) of each exchange of a plan given in pass order."""
    out, ns = [], 1
    for q in range(len(radices) - 1):
        out.append((radices[q], ns, radices[q + 1]))
        ns *= radices[q]
    return out


def tile_cycles(L: int, nt: int, e: int, forward: Sequence[int], n_inv: int, n_fwd: int,
                layout: Callable[[bool, int], Layout]) -> dict[str, int]:
    """Cycles per tile of n_inv inverse (reversed plan) and n_fwd forward transforms."""
    tot = {"cycles": 0, "ideal": 0}
    for inv, count, plan in ((True, n_inv, list(reversed(forward))), (False, n_fwd, forward)):
        for x, (rw, ns, rr) in enumerate(plan_exchanges(plan)):
            w, r, ideal = exchange_cycles(L, nt, e, rw, ns, rr, layout(inv, x))
            tot["cycles"] += count * (w + r)
            tot["ideal"] += count * ideal
    return tot


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.parse_args(argv)
    L, nt, e, fwd = 384, 128, 12, [4, 4, 2, 12]
    print("exchange | writes: radix, stride | reads: radix | padded w + r | current w + r | ideal")
    for inv, plan in ((True, list(reversed(fwd))), (False, fwd)):
        for x, (rw, ns, rr) in enumerate(plan_exchanges(plan)):
            cur = SWIZZLES_384.get((inv, x), pad32)
            pw, pr, ideal = exchange_cycles(L, nt, e, rw, ns, rr, pad32)
            cw, cr, _ = exchange_cycles(L, nt, e, rw, ns, rr, cur)
            print(f"{'inverse' if inv else 'forward'} {x} | {rw}, {ns} | {rr} | {pw} + {pr} | "
                  f"{cw} + {cr} | {ideal}")
    for name, lay in (("padded", lambda inv, x: pad32),
                      ("current", lambda inv, x: SWIZZLES_384.get((inv, x), pad32))):
        t = tile_cycles(L, nt, e, fwd, 2, 3, lay)
        print(f"{name}: {t['cycles']} cycles per tile, "
              f"{1 - t['ideal'] / t['cycles']:.1%} of them conflicts")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
