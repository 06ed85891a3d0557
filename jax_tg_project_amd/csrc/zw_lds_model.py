"""LDS bank model behind the ZwLds table of tg_passes.hpp.

The register z bodies exchange each line's points through a per-line LDS
buffer: element n of line lw sits at lw * LPL + pidx(n), pidx(n) = n +
(n >> PS).  This script counts the LDS cycles of every put and get of a
body's line transforms (plan stages as in reg_fft: rs_put with stride P,
rs_get with stride N / R) and of the Hermitian split (zw_split_pair), for
a wavefront of 64 lanes served 256 bytes of lanes per cycle (32 FP32 or 16
FP64 lanes) by 64 banks of 4 bytes: per group of lanes, the cost is the
largest number of distinct 4-byte words that map to one bank.  It prints,
per (precision, N, E), the current and the cheapest (PS, LPL).

    python zw_lds_model.py            # the FP32 shapes of the table
    python zw_lds_model.py 96 12 4    # one shape: N, E, bytes per real
"""
import math
import sys
from collections import defaultdict


def greedy_plan(n, e):
    """reg_fft's default plan: the largest of 16, 12, 8, 4, 3, 2 dividing e and the rest of n."""
    out, rem = [], n
    while rem > 1:
        for c in (16, 12, 8, 4, 3, 2):
            if e % c == 0 and rem % c == 0:
                out.append(c)
                rem //= c
                break
        else:
            raise ValueError(f"no plan for N={n}, E={e}")
    return out


def transform_accesses(n, e, plan, pidx, lines, tpl, lpl):
    acc, p = [], 1
    for i, r in enumerate(plan):
        if i + 1 < len(plan):
            nb = e // r
            for j in range(nb):
                for m in range(r):
                    row = []
                    for lw in range(lines):
                        for t in range(tpl):
                            b = t * nb + j
                            k = b % p
                            row.append(lw * lpl + pidx((b - k) * r + k + m * p))
                    acc.append(row)
            r2 = plan[i + 1]
            nb2 = e // r2
            for j in range(nb2):
                for m in range(r2):
                    acc.append([lw * lpl + pidx(t * nb2 + j + m * (n // r2)) for lw in range(lines) for t in range(tpl)])
        p *= r
    return acc


def split_accesses(n, e, plan, pidx, lines, tpl, lpl):
    r1 = plan[0]
    nb1, s1, kz = e // r1, n // r1, n // 3 + 1
    acc = [[lw * lpl + pidx(t * nb1 + j + m * s1) for lw in range(lines) for t in range(tpl)]
           for j in range(nb1) for m in range(r1)]
    for i in range((kz + tpl - 1) // tpl):
        for mirror in (False, True):
            row = []
            for lw in range(lines):
                for t in range(tpl):
                    k = min(t + i * tpl, kz - 1)
                    row.append(lw * lpl + pidx((n - k) % n if mirror else k))
            acc.append(row)
    return acc


def cycles(accesses, words):
    group = 256 // (4 * words)
    total = 0
    for row in accesses:
        for g in range(0, len(row), group):
            banks = defaultdict(set)
            for el in row[g:g + group]:
                for w in range(words * el, words * el + words):
                    banks[w % 64].add(w)
            total += max(len(v) for v in banks.values())
    return total


def cost(n, e, real_bytes, ps, lpl, n_inv=3, n_fwd=2, n_split=1):
    """Cycles of one hydro body: three inverse and two forward transforms, one split."""
    plan = greedy_plan(n, e)
    tpl = n // e
    lines = 64 // tpl
    pidx = lambda v: v + (v >> ps)
    acc = transform_accesses(n, e, plan, pidx, lines, tpl, lpl) * n_inv
    acc += transform_accesses(n, e, plan[::-1], pidx, lines, tpl, lpl) * n_fwd
    acc += split_accesses(n, e, plan, pidx, lines, tpl, lpl) * n_split
    return cycles(acc, 2 * real_bytes // 4)


def default_layout(n, e, real_bytes):
    ps = int(math.log2(e))
    used = n + (n >> ps)
    lpl = (used + 7) // 16 * 16 + 8 if real_bytes == 8 else (used + 15) // 32 * 32 + 16
    return ps, lpl


def best_layout(n, e, real_bytes):
    p0 = int(math.log2(e))
    best = None
    for ps in range(max(1, p0 - 2), p0 + 2):
        used = n + (n >> ps)
        for lpl in range(used, used + 33):
            c = cost(n, e, real_bytes, ps, lpl)
            if best is None or (c, lpl) < best[:2]:
                best = (c, lpl, ps)
    return best


def main(argv):
    shapes = [(int(argv[0]), int(argv[1]), int(argv[2]))] if argv else \
        [(32, 8, 4), (32, 16, 4), (48, 12, 4), (64, 8, 4), (64, 16, 4), (96, 12, 4), (128, 16, 4), (256, 16, 4)]
    for n, e, rb in shapes:
        ps0, lpl0 = default_layout(n, e, rb)
        c0 = cost(n, e, rb, ps0, lpl0)
        c, lpl, ps = best_layout(n, e, rb)
        print(f"FP{8 * rb} N={n} E={e}: default PS={ps0} LPL={lpl0} {c0} cycles; best PS={ps} LPL={lpl} {c} cycles")


if __name__ == "__main__":
    main(sys.argv[1:])
