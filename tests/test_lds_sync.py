"""LDS exchange synchronisation of the DGT passes, checked on the CPU.

dif_passes / dit_passes / gsi_passes (cupoly_amd/csrc/dgt_dev.hpp; the last
two share one exchange sequence) skip the workgroup barriers of an exchange
that is wave-local (each wave reads back only the slots it wrote).  They skip
the barrier before an exchange's stores when the previous exchange was also
wave-local under a map with the same padding ratio, or when the previous
exchange of the sequence used the same map (each thread then overwrites only
the words it read itself).  This test
replays the LDS addresses those functions compute (the same pass layouts,
padding maps and skip rules) for every thread of every workgroup shape the
kernels launch, and requires that any two accesses to one LDS slot from
different waves, at least one of them a store, are separated by a workgroup
barrier.  A planted mistake (a cross-wave exchange treated as wave-local)
must be reported.
"""
import bisect
import collections

import pytest

EPT, LOG_EPT = 8, 3


def lds_map_for(lst_lo):
    return (1, 3) if lst_lo == 0 else (8, 6) if (lst_lo >= 3 and lst_lo % 3 == 0) else (0, 0)


def lds_row(L):
    return (1 << L) + ((1 << L) >> 3)


def pass_index(L, LS0, R, t, i):
    M = 1 << L; T = M // EPT; U = 1 << R; ST = M >> (LS0 + R)
    ju, q = i // U, i % U
    u = t + ju * T
    return (u // ST) * (U * ST) + (u % ST) + q * ST


def rt_addrs(L, t, ls0, m):
    A, S = m
    lst = L - ls0 - LOG_EPT
    b = ((t >> lst) << (lst + LOG_EPT)) + (t & ((1 << lst) - 1))
    pb = b + A * (b >> S)
    d = (1 << lst) + ((A << (lst - S)) if lst >= S else 0)
    return [pb + i * d for i in range(EPT)]


def fixed_addrs(L, LS0, R, m, t):
    A, S = m
    out = []
    for i in range(EPT):
        l = pass_index(L, LS0, R, t, i)
        out.append(l + A * (l >> S))
    return out


def plan(LSB, LSE):
    NS = LSE - LSB; R0 = NS % LOG_EPT; NF = NS // LOG_EPT
    return R0, NF, LSB + R0


def dif_ops(L, LSB, LSE, local_max=6, same_map=True):
    R0, NF, LSF = plan(LSB, LSE)
    MP = lds_map_for(L - LSF - LOG_EPT)
    ops = []
    if R0 > 0:
        if NF == 0:
            return ops
        ops += [("sync",), ("st", lambda t: fixed_addrs(L, LSB, R0, MP, t)), ("sync",),
                ("ld", lambda t, ls=LSF: rt_addrs(L, t, ls, MP))]
    prev_local = prev_pad = False
    prev_m = MP if R0 > 0 else None
    for i in range(NF - 1):
        ls0 = LSF + LOG_EPT * i
        m = lds_map_for(L - ls0 - 2 * LOG_EPT)
        local = (L - ls0 - LOG_EPT) <= local_max; pad = m[0] != 0
        skip = (local and prev_local and pad == prev_pad) or (same_map and m == prev_m)
        ops.append(("wave",) if skip else ("sync",))
        ops.append(("st", lambda t, ls=ls0, mm=m: rt_addrs(L, t, ls, mm)))
        ops.append(("wave",) if local else ("sync",))
        ops.append(("ld", lambda t, ls=ls0 + LOG_EPT, mm=m: rt_addrs(L, t, ls, mm)))
        prev_local, prev_pad, prev_m = local, pad, m
    return ops


def dit_ops(L, LSB, LSE, local_max=6, same_map=True, after_dif=False):
    R0, NF, LSF = plan(LSB, LSE)
    MP = lds_map_for(L - LSF - LOG_EPT)
    ops = []
    prev_local = prev_pad = False
    # after_dif: the previous LDS access was dif_ops' last exchange over the
    # same stages (gsi_passes<..., AFTER_DIF = true>)
    prev_m = lds_map_for(L - (LSF + LOG_EPT * (NF - 2)) - 2 * LOG_EPT) if after_dif and NF >= 2 else None
    for i in range(NF - 2, -1, -1):
        ls0 = LSF + LOG_EPT * i
        m = lds_map_for(L - ls0 - 2 * LOG_EPT)
        local = (L - ls0 - LOG_EPT) <= local_max; pad = m[0] != 0
        skip = (local and prev_local and pad == prev_pad) or (same_map and m == prev_m)
        ops.append(("wave",) if skip else ("sync",))
        ops.append(("st", lambda t, ls=ls0 + LOG_EPT, mm=m: rt_addrs(L, t, ls, mm)))
        ops.append(("wave",) if local else ("sync",))
        ops.append(("ld", lambda t, ls=ls0, mm=m: rt_addrs(L, t, ls, mm)))
        prev_local, prev_pad, prev_m = local, pad, m
    if R0 > 0 and NF > 0:
        ops += [("wave",) if (same_map and MP == prev_m) else ("sync",),
                ("st", lambda t: rt_addrs(L, t, LSF, MP)), ("sync",),
                ("ld", lambda t: fixed_addrs(L, LSB, R0, MP, t))]
    return ops


def check(L, LSB, LSE, rows, seqs, joined=()):
    """seqs run in order with a workgroup barrier between them, except before
    the sequences whose index is in `joined` (run straight after the one
    before, as gsi_passes<AFTER_DIF> after dif_passes)."""
    T = (1 << L) // EPT
    nthreads = T * rows
    ops = [("sync",)]
    for k, s in enumerate(seqs):
        if k in joined:
            ops.pop()
        ops += s + [("sync",)]
    acc = collections.defaultdict(list)  # addr -> [(op index, wave, is_store)]
    span = lds_row(L)
    for k, op in enumerate(ops):
        if op[0] not in ("st", "ld"):
            continue
        for tid in range(nthreads):
            row, t = tid // T, tid % T
            for a in op[1](t):
                assert 0 <= a < span, (L, a)
                acc[row * span + a].append((k, tid // 64, op[0] == "st"))
    syncs = [k for k, op in enumerate(ops) if op[0] == "sync"]
    def separated(k1, k2):
        lo, hi = min(k1, k2), max(k1, k2)
        j = bisect.bisect_right(syncs, lo)
        return j < len(syncs) and syncs[j] < hi
    bad = 0
    for a, lst in acc.items():
        # per op, group waves; compare consecutive relevant accesses
        for x in range(len(lst)):
            for y in range(x + 1, len(lst)):
                k1, w1, s1 = lst[x]; k2, w2, s2 = lst[y]
                if w1 == w2 or not (s1 or s2):
                    continue
                if k1 == k2 or not separated(k1, k2):
                    bad += 1
    nwave = sum(1 for op in ops if op[0] == "wave")
    return bad, nwave


def rows_for(L):
    T = (1 << L) // EPT
    return 1 if T >= 256 else 256 // T


CASES = [(L, 0, L, rows_for(L)) for L in range(4, 14)]        # single, fused, a*s+e, tensor
CASES += [(L, 0, L, 8192 >> L) for L in range(10, 14)]         # uneven tiles, 1024 threads
CASES += [(13, 0, 7, 1), (13, 0, 8, 1), (7, 0, 7, 64)]         # column and block kernels
CASES += [(9, 0, s, 1) for s in range(3, 9)] + [(7, 0, 7, 4)]  # 64-thread split, N = 2^10 .. 2^15


@pytest.mark.parametrize("L,lsb,lse,rows", CASES)


def test_exchanges_are_race_free(L, lsb, lse, rows):
    seq = [dif_ops(L, lsb, lse), dif_ops(L, lsb, lse), dit_ops(L, lsb, lse)]
    bad, _ = check(L, lsb, lse, rows, seq)
    assert bad == 0
    # the fused product: the inverse runs straight after b's forward sequence
    seq = [dif_ops(L, lsb, lse), dif_ops(L, lsb, lse), dit_ops(L, lsb, lse, after_dif=True)]
    bad, _ = check(L, lsb, lse, rows, seq, joined=(2,))
    assert bad == 0


def bank_conflicts(L, rows, ops):
    """16-lane groups of 16-byte accesses; a group is conflict-free when its 16
    slots are distinct modulo 16 (256 bytes).  Returns (conflicted, groups)."""
    T = (1 << L) // EPT
    nthreads = T * rows
    bad = tot = 0
    for op in ops:
        if op[0] not in ("st", "ld"):
            continue
        for g0 in range(0, nthreads, 16):
            tids = range(g0, min(g0 + 16, nthreads))
            addrs = [op[1](tid % T) for tid in tids]
            base = [(tid // T) * lds_row(L) for tid in tids]
            for i in range(EPT):
                tot += 1
                if len({(b + a[i]) % 16 for b, a in zip(base, addrs)}) != len(addrs):
                    bad += 1
    return bad, tot


@pytest.mark.parametrize("L,lsb,lse,rows", [c for c in CASES if c[0] >= 10 or c[3] > 1])
def test_exchanges_are_bank_conflict_free(L, lsb, lse, rows):
    bad, tot = bank_conflicts(L, rows, dif_ops(L, lsb, lse) + dit_ops(L, lsb, lse))
    assert tot > 0 and bad == 0


def test_wave_local_exchanges_exist_at_n16384():
    # the two exchanges closest to the last stages skip barriers at N_g = 8192,
    # and every exchange whose previous one in the sequence used the same map
    # skips the barrier before its stores
    _, nwave = check(13, 0, 13, 1, [dif_ops(13, 0, 13, same_map=False),
                                     dit_ops(13, 0, 13, same_map=False)])
    assert nwave == 6
    _, nwave = check(13, 0, 13, 1, [dif_ops(13, 0, 13), dit_ops(13, 0, 13)])
    assert nwave == 10


def test_planted_map_mismatch_skip_is_reported():
    """Skipping the barrier before the stores is safe only when the previous
    exchange used the same map: pretend a padded map is the same."""
    L = 13
    R0, NF, LSF = plan(0, L)

    def bad_dif():
        ops = []
        prev_m = None
        for i in range(NF - 1):
            ls0 = LSF + LOG_EPT * i
            m = (1, 3) if i % 2 else (0, 0)  # alternate maps, skip regardless
            ops.append(("wave",) if prev_m is not None else ("sync",))
            ops.append(("st", lambda t, ls=ls0, mm=m: rt_addrs(L, t, ls, mm)))
            ops.append(("sync",))
            ops.append(("ld", lambda t, ls=ls0 + LOG_EPT, mm=m: rt_addrs(L, t, ls, mm)))
            prev_m = m
        return ops
    bad, _ = check(L, 0, L, rows_for(L), [bad_dif()])
    assert bad > 0


def test_planted_cross_wave_skip_is_reported():
    seq = [dif_ops(13, 0, 13, local_max=9), dit_ops(13, 0, 13, local_max=9)]
    bad, _ = check(13, 0, 13, 1, seq)
    assert bad > 0
