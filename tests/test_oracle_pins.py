"""Pins for the CPU oracle: each oracle function is checked against something
other than itself (closed forms, brute force, exact rational arithmetic, the
reference's own known-answer tests, or a restatement of the reference's kernel
index maps).  CPU only."""
import random
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

M64 = (1 << 64) - 1


# --------------------------------------------------------------------------
# modular core
# --------------------------------------------------------------------------
def ref_barrett_mod(a, p):
    """The reference's Barrett reduction, src/cuda/dgt.cu:394-435 (BFV branch),
    restated on Python ints with 64-bit wrap-around."""
    nb = p.bit_length()
    mu = (1 << (2 * nb)) // p  # dgt.cu:1763-1766, floor(4^NumBits(p) / p)
    z = (1 << nb) - 1
    lo, hi = a & M64, a >> 64
    m0 = lo & z
    m1 = ((hi << (64 - nb)) + (lo >> nb)) & M64
    prod = m1 * mu
    q = ((((prod >> 64) << (64 - nb)) & M64) + ((prod & M64) >> nb)) & M64
    qp = q * p
    q0 = (qp & M64) & z
    q1 = ((((qp >> 64) << (64 - nb)) & M64) + ((qp & M64) >> nb)) & M64
    r0 = (m0 - q0) & M64
    r1 = ((m1 - q1) - (1 if m0 < q0 else 0)) & M64
    r = ((r0 - r1 * p) & M64) & z
    return (r - p) & M64 if r >= p else r


def test_reference_barrett_equals_exact_mod():
    """The reference's device mulmod is exact: its Barrett restated agrees with
    `%` on random and edge products, so the oracle's exact `%` matches it."""
    P = O.tables()["primes"]
    rng = random.Random(7)
    for p in P[:17]:
        cases = [(rng.randrange(p), rng.randrange(p)) for _ in range(3000)]
        cases += [(rng.randrange(1 << 63), rng.randrange(p)) for _ in range(500)]
        cases += [(x, y) for x in (0, 1, p - 1, p - 2, (1 << 63) - 1) for y in (0, 1, p - 1)]
        for x, y in cases:
            assert ref_barrett_mod(x * y, p) == (x * y) % p
            assert O.lib().or_mulmod(x, y, p) == (x * y) % p


def test_powmod_matches_python():
    p = O.tables()["primes"][3]
    rng = random.Random(1)
    for _ in range(200):
        a, e = rng.randrange(p), rng.randrange(1 << 64)
        assert O.lib().or_powmod(a, e, p) == pow(a, e, p)


# --------------------------------------------------------------------------
# DGT
# --------------------------------------------------------------------------
def bitrev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


def gi_mul(a, b, p):
    return ((a[0] * b[0] - a[1] * b[1]) % p, (a[0] * b[1] + a[1] * b[0]) % p)


def gi_pow(a, e, p):
    r = (1, 0)
    while e:
        if e & 1:
            r = gi_mul(r, a, p)
        a = gi_mul(a, a, p)
        e >>= 1
    return r


def closed_form_dgt(A, xi, omega, p):
    """X_k = sum_n A_n xi^n omega^(nk), stored at bitrev(k)."""
    Ng = len(A)
    bits = Ng.bit_length() - 1
    tw = [gi_mul(A[n], gi_pow(xi, n, p), p) for n in range(Ng)]
    out = [None] * Ng
    for k in range(Ng):
        sr = si = 0
        for n in range(Ng):
            w = pow(omega, n * k, p)
            sr += tw[n][0] * w
            si += tw[n][1] * w
        out[bitrev(k, bits)] = (sr % p, si % p)
    return out


@pytest.mark.parametrize("Ng,rid", [(16, 0), (32, 5), (64, 11), (128, 2)])
def test_dgt_forward_closed_form(Ng, rid):
    T = O.tables()
    p = T["primes"][rid]
    si = T["sizes"].index(Ng)
    xi = T["nthroot"][rid][si]
    omega = pow(T["proot"][rid], (p - 1) // Ng, p)
    rng = random.Random(Ng + rid)
    A = [(rng.randrange(p), rng.randrange(p)) for _ in range(Ng)]
    x = np.array(A, dtype=np.uint64)
    O.lib().or_dgt_forward(O._p(x), Ng, p, O._p(O.u64(xi)), omega)
    assert [tuple(int(v) for v in row) for row in x] == closed_form_dgt(A, xi, omega, p)


# The reference's hierarchical Na x Nb split (include/cuPoly/cuda/dgt.h:37-50).
HDGT = {1024: (32, 32), 2048: (64, 32), 4096: (64, 64), 8192: (128, 64),
        16384: (128, 128), 32768: (256, 128)}


@pytest.mark.parametrize("Ng", sorted(HDGT))
def test_hdgt_placement_is_full_bitreversal(Ng):
    """bitrev_Na(k / Nb) + Na * bitrev_Nb(k % Nb) == bitrev_Ng(k)."""
    Na, Nb = HDGT[Ng]
    la, lb, lg = Na.bit_length() - 1, Nb.bit_length() - 1, Ng.bit_length() - 1
    for k in range(Ng):
        assert bitrev(k // Nb, la) + Na * bitrev(k % Nb, lb) == bitrev(k, lg)


def ref_hdgt(A, Na, Nb, p, xi, g_root, direction):
    """The reference's two-pass hierarchical DGT restated from its kernels:
    hdgt_tr<Nb> / hdgt_tr<Na> (dgt.cu:819-908) forward and hidgt<Na> /
    hidgt_tr<Nb> (dgt.cu:917-1006) inverse, with the tables of
    compute_gjk / compute_nthroot / DGTEngine::init (dgt.cu:1509-1692,
    1879-1960).  One residue; A is a list of (re, im)."""
    Ng = Na * Nb

    def gjk(S, inverse):
        g = pow(g_root, (p - 1) // (2 * S), p)
        L = S.bit_length() - 1
        tab = {}
        for s in range(L):
            for j in range(S):
                e = (2 * S - j if inverse else j) * ((2 * S) >> (L - s))
                tab[(s, j)] = pow(g, e, p)
        return tab

    def bitrev_map(M):
        return [bitrev(i, M.bit_length() - 1) for i in range(M)]

    w = pow(g_root, (p - 1) // Ng, p)
    winv = pow(w, -1, p)
    brNb = bitrev_map(Nb)
    gN = [pow(w, (c // Nb) * brNb[c % Nb], p) for c in range(Ng)]
    ninv = pow(Ng, -1, p)
    ginvN = [pow(winv, brNb[c // Na] * (c % Na), p) * ninv % p for c in range(Ng)]

    def dif(s_data, S, tab):  # doDGT (dgt.cu:755-777) with dgttransform
        L = S.bit_length() - 1
        for stride in range(L):
            m = S // (2 << stride)
            new = list(s_data)
            for cid in range(S // 2):
                j = cid * 2 * m // S
                i = j + (cid % (S // (2 * m))) * 2 * m
                a = tab[(stride, j)]
                x0, x1 = s_data[i], s_data[i + m]
                new[i] = ((x0[0] + x1[0]) % p, (x0[1] + x1[1]) % p)
                new[i + m] = (a * (x0[0] - x1[0]) % p, a * (x0[1] - x1[1]) % p)
            s_data = new
        return s_data

    def dit(s_data, S, tab):  # doIDGT (dgt.cu:781-806) with idgttransform
        L = S.bit_length() - 1
        for stride in range(L - 1, -1, -1):
            m = S // (2 << stride)
            new = list(s_data)
            for cid in range(S // 2):
                j = cid * 2 * m // S
                i = j + (cid % (S // (2 * m))) * 2 * m
                a = tab[(stride, j)]
                x0, x1 = s_data[i], s_data[i + m]
                ax = (a * x1[0] % p, a * x1[1] % p)
                new[i] = ((x0[0] + ax[0]) % p, (x0[1] + ax[1]) % p)
                new[i + m] = ((x0[0] - ax[0]) % p, (x0[1] - ax[1]) % p)
            s_data = new
        return s_data

    data = list(A)
    if direction == 0:
        nth = [gi_pow(xi, c, p) for c in range(Ng)]
        tmp = [None] * Ng
        gNb = gjk(Nb, False)
        for blk in range(Na):  # hdgt_tr<Nb>: Nr = Nb, Nc = Na
            s = [gi_mul(data[t * Na + blk], nth[t * Na + blk], p) for t in range(Nb)]
            s = dif(s, Nb, gNb)
            for t in range(Nb):
                tmp[t + blk * Nb] = s[t]
        gNa = gjk(Na, False)
        for blk in range(Nb):  # hdgt_tr<Na>: Nr = Na, Nc = Nb
            s = [(gN[t * Nb + blk] * tmp[t * Nb + blk][0] % p,
                  gN[t * Nb + blk] * tmp[t * Nb + blk][1] % p) for t in range(Na)]
            s = dif(s, Na, gNa)
            for t in range(Na):
                data[t + blk * Na] = s[t]
        return data
    invnth = [gi_pow(xi, c, p) for c in range(Ng)]  # xi here is INVNTHROOT
    tmp = [None] * Ng
    iNa = gjk(Na, True)
    for blk in range(Nb):  # hidgt<Na>: contiguous rows
        s = dit(data[blk * Na:(blk + 1) * Na], Na, iNa)
        for t in range(Na):
            c = t + blk * Na
            tmp[c] = (ginvN[c] * s[t][0] % p, ginvN[c] * s[t][1] % p)
    iNb = gjk(Nb, True)
    out = [None] * Ng
    for blk in range(Na):  # hidgt_tr<Nb>: strided columns, untwist
        s = dit([tmp[t * Na + blk] for t in range(Nb)], Nb, iNb)
        for t in range(Nb):
            out[t * Na + blk] = gi_mul(s[t], invnth[t * Na + blk], p)
    return out


@pytest.mark.parametrize("Ng,rid", [(1024, 1), (2048, 4)])
def test_oracle_dgt_matches_reference_hierarchical_kernels(Ng, rid):
    T = O.tables()
    Na, Nb = HDGT[Ng]
    p = T["primes"][rid]
    si = T["sizes"].index(Ng)
    xi, xiinv = T["nthroot"][rid][si], T["invnthroot"][rid][si]
    g = T["proot"][rid]
    rng = random.Random(Ng)
    A = [(rng.randrange(p), rng.randrange(p)) for _ in range(Ng)]
    want = ref_hdgt(A, Na, Nb, p, xi, g, 0)
    omega = pow(g, (p - 1) // Ng, p)
    x = np.array(A, dtype=np.uint64)
    O.lib().or_dgt_forward(O._p(x), Ng, p, O._p(O.u64(xi)), omega)
    assert [tuple(int(v) for v in row) for row in x] == want
    back = ref_hdgt(want, Na, Nb, p, xiinv, g, 1)
    assert back == A
    O.lib().or_dgt_inverse(O._p(x), Ng, p, O._p(O.u64(xiinv)), omega)
    assert [tuple(int(v) for v in row) for row in x] == A


def test_pointwise_matches_gaussian_integer_arithmetic():
    """or_pointwise (add, sub, GI product, GI multiply-add) against Python
    integers on random slots and the edge residues 0, 1, p - 1."""
    P = O.Params(3, 2, 64, 65537)
    rng = random.Random(21)
    Ng = P.Ng
    shape = (len(P.primes), Ng, 2)
    vals = {}
    for name in "abc":
        x = np.empty(shape, dtype=np.uint64)
        for r, p in enumerate(P.primes):
            edge = [0, 1, p - 1]
            for i in range(Ng):
                for j in range(2):
                    x[r, i, j] = edge[(i + j) % 3] if i < 6 else rng.randrange(p)
        vals[name] = x
    a, b, c = vals["a"], vals["b"], vals["c"]
    got = [O.pointwise(P, a, b, c, base=2, op=op) for op in range(4)]
    for r, p in enumerate(P.primes):
        for i in range(Ng):
            ar, ai = int(a[r, i, 0]), int(a[r, i, 1])
            br, bi = int(b[r, i, 0]), int(b[r, i, 1])
            cr, ci = int(c[r, i, 0]), int(c[r, i, 1])
            pr_, pi_ = (ar * br - ai * bi) % p, (ar * bi + ai * br) % p
            want = [((ar + br) % p, (ai + bi) % p), ((ar - br) % p, (ai - bi) % p),
                    (pr_, pi_), ((pr_ + cr) % p, (pi_ + ci) % p)]
            for op in range(4):
                assert (int(got[op][r, i, 0]), int(got[op][r, i, 1])) == want[op], (op, r, i)


def test_xi_rho_dot_word_decomposition():
    """xi / rho against Python-integer products by (Q/q_i)^-1 and Q/q_i, only
    residue i of output i touched, and the reference's identity
    sum_i rho(a)[i] * xi(b)[i] == a * b (test_bfv_mode.cu:492-546) through dot."""
    P = O.Params(3, 4, 256, 65537)
    rng = np.random.default_rng(3)
    a = np.stack([rng.integers(0, q, size=(P.Ng, 2), dtype=np.uint64) for q in P.q])
    b = np.stack([rng.integers(0, q, size=(P.Ng, 2), dtype=np.uint64) for q in P.q])
    X = O.xi_bfv(P, b)
    R = O.rho_bfv(P, a)
    for i, q in enumerate(P.q):
        Qi = P.Q // q
        assert P.Qqi[i] == Qi % q and P.invQqi[i] * Qi % q == 1
        for j in range(P.k):
            if j != i:
                assert not X[i][j].any() and not R[i][j].any()
        for s_ in range(0, P.Ng, 17):
            for h in range(2):
                assert int(X[i][i, s_, h]) == int(b[i, s_, h]) * pow(Qi, -1, q) % q
                assert int(R[i][i, s_, h]) == int(a[i, s_, h]) * Qi % q
    assert np.array_equal(O.dot(P, R, X), O.pointwise(P, a, b, op=2))


def test_crt_reconstruction_identity():
    """oracle CRT (poly_icrt restated) returns x for the residues of any
    x in [0, Q), including 0 and Q - 1 (Chinese remainder theorem)."""
    for k in (1, 3, 8):
        P = O.Params(k, 1, 64, 65537)
        rng = random.Random(k)
        xs = [0, P.Q - 1] + [rng.randrange(P.Q) for _ in range(P.n - 2)]
        assert O.crt(P, residues_of(xs, P.q, P.Ng)) == xs


def test_dgt_roundtrip_and_constant():
    P = O.Params(3, 4, 4096, 65537)
    rng = np.random.default_rng(3)
    x = np.stack([rng.integers(0, p, size=(P.Ng, 2), dtype=np.uint64) for p in P.q])
    f = O.dgt(P, x, 0, 0)
    assert not np.array_equal(f, x)
    assert np.array_equal(O.dgt(P, f, 0, 1), x)
    # DGT(constant) = constant in every slot (cudaengine.cu op_by_int relies on it)
    c = np.zeros_like(x)
    c[:, 0, 0] = [12345, 777, 5]
    fc = O.dgt(P, c, 0, 0)
    for r, val in enumerate([12345, 777, 5]):
        assert np.all(fc[r, :, 0] == val) and np.all(fc[r, :, 1] == 0)


@pytest.mark.parametrize("n,k", [(32, 1), (256, 2), (2048, 2)])
def test_polymul_matches_schoolbook(n, k):
    P = O.Params(k, k + 1, n, 256)
    rng = np.random.default_rng(n)
    a = np.stack([rng.integers(0, p, size=n, dtype=np.uint64) for p in P.q])
    b = np.stack([rng.integers(0, p, size=n, dtype=np.uint64) for p in P.q])
    c = O.polymul_batch(P, O.fold(a, P.Ng)[None], O.fold(b, P.Ng)[None])[0]
    for r, p in enumerate(P.q):
        assert np.array_equal(O.unfold(c[r]), O.schoolbook(a[r], b[r], p))


def test_schoolbook_matches_python_bigints():
    p = O.tables()["primes"][0]
    n = 64
    rng = random.Random(5)
    a = [rng.randrange(p) for _ in range(n)]
    b = [rng.randrange(p) for _ in range(n)]
    want = [0] * n
    for i in range(n):
        for j in range(n):
            if i + j < n:
                want[i + j] += a[i] * b[j]
            else:
                want[i + j - n] -= a[i] * b[j]
    got = O.schoolbook(a, b, p)
    assert [int(x) for x in got] == [w % p for w in want]


def test_polymul_n2048_golden_vector():
    """n=2048, k=2 product against the committed fixture (inputs from the
    Philox uniform sampler, seed 2048); the fixture itself was checked
    against the schoolbook product when it was generated."""
    import os
    path = os.path.join(O.REPO, "tests", "golden", "polymul_n2048_k2.npz")
    g = np.load(path)
    P = O.Params(2, 3, 2048, 256)
    a = O.sample_uniform(P.q, P.n, int(g["seed"]), 0)
    b = O.sample_uniform(P.q, P.n, int(g["seed"]), 1)
    assert np.array_equal(a, g["a"]) and np.array_equal(b, g["b"])
    c = O.polymul_batch(P, O.fold(a, P.Ng)[None], O.fold(b, P.Ng)[None])[0]
    assert np.array_equal(O.unfold(c), g["c"])


# --------------------------------------------------------------------------
# FP64 with directed rounding
# --------------------------------------------------------------------------
def rd_of(fr):
    """Largest double <= the exact rational fr (fr > 0)."""
    d = float(fr)  # round to nearest
    if Fraction(d) > fr:
        d = np.nextafter(d, -np.inf)
    return float(d)


def test_rcp_rd_exact():
    for q in O.tables()["primes"]:
        qd = float(q)  # RN conversion
        assert O.rcp_rd(q) == rd_of(Fraction(1) / Fraction(qd))


def test_fma_rd_exact():
    rng = random.Random(11)
    for _ in range(3000):
        a = float(rng.randrange(1 << 63))
        b = O.rcp_rd(rng.choice(O.tables()["primes"]))
        c = rng.random() * rng.choice([1, 4, 30])
        exact = Fraction(a) * Fraction(b) + Fraction(c)
        assert O.fma_rd(a, b, c) == rd_of(exact)


# --------------------------------------------------------------------------
# HPS
# --------------------------------------------------------------------------
def residues_of(vals, primes, Ng):
    flat = np.array([[v % p for v in vals] for p in primes], dtype=np.uint64)
    return O.fold(flat, Ng)


@pytest.mark.parametrize("k", [1, 2, 4])
def test_basis_ext_q_to_b_exact_crt(k):
    P = O.Params(k, k + 1, 2048, 65537)
    rng = random.Random(k)
    x = [rng.randrange(P.Q) for _ in range(P.n)]
    x[:4] = [0, 1, P.Q - 1, P.Q // 2]
    out, v = O.basis_ext_q_to_b(P, residues_of(x, P.q, P.Ng), want_v=True)
    got = O.unfold(out)
    for c in range(P.n):
        frac = Fraction(x[c], P.Q)
        near_edge = frac < Fraction(1, 10**8) or frac > 1 - Fraction(1, 10**8)
        for i, bi in enumerate(P.b):
            want = x[c] % bi
            if near_edge and c == 2:  # x = Q - 1 snaps to x - Q (documented quirk)
                assert int(got[i, c]) in (want, (x[c] - P.Q) % bi)
            else:
                assert int(got[i, c]) == want, (c, i)


def test_basis_ext_q_to_b_kat_squares():
    """Known-answer test of the reference (test_bfv_mode.cu:548-638): x_i = i^2 mod t."""
    t = 256
    P = O.Params(2, 3, 2048, t)
    x = [(i * i) % t for i in range(P.n)]
    got = O.unfold(O.basis_ext_q_to_b(P, residues_of(x, P.q, P.Ng)))
    for i in range(P.kl):
        assert [int(v) for v in got[i]] == x


@pytest.mark.parametrize("k", [2, 4])
def test_basis_ext_b_to_q_exact_crt(k):
    P = O.Params(k, k + 1, 2048, 65537)
    rng = random.Random(100 + k)
    y = [rng.randrange(P.B // 1000, P.B - P.B // 1000) for _ in range(P.n)]
    flat = np.array([[v % bi for v in y] for bi in P.b], dtype=np.uint64)
    got = O.unfold(O.basis_ext_b_to_q(P, flat))
    for j, qj in enumerate(P.q):
        assert [int(v) for v in got[j]] == [v % qj for v in y]


@pytest.mark.parametrize("k,t", [(2, 256), (4, 65537)])
def test_complex_scaling_within_one_of_exact(k, t):
    """round(t*x/Q) mod b_i within +-1 (the reference's own tolerance,
    test_bfv_mode.cu:1092-1095), exact for almost all coefficients."""
    P = O.Params(k, k + 1, 1024, t)
    rng = random.Random(200 + k)
    x = [rng.randrange(P.Q * P.B) for _ in range(P.n)]
    aq = residues_of(x, P.q, P.Ng)
    ab = residues_of(x, P.b, P.Ng)
    out = O.complex_scaling(P, aq, ab)
    exact_hits = 0
    for c in range(P.n):
        r = (2 * t * x[c] + P.Q) // (2 * P.Q)  # round half up of t*x/Q
        for i, bi in enumerate(P.b):
            d = (int(out[i, c]) - r) % bi
            assert d in (0, 1, bi - 1), (c, i)
            exact_hits += d == 0
    assert exact_hits >= 0.99 * P.n * P.kl


def test_complex_scaling_then_b_to_q_small_inputs():
    """The reference's ComplexBasisScalingUniform shape: x in [0, Q),
    then complex scaling and B -> Q give floor(t*x/Q) within 1."""
    t = 65537
    P = O.Params(4, 5, 1024, t)
    rng = random.Random(9)
    x = [rng.randrange(P.Q) for _ in range(P.n)]
    aq = residues_of(x, P.q, P.Ng)
    ab = O.basis_ext_q_to_b(P, aq)
    out = O.basis_ext_b_to_q(P, O.complex_scaling(P, aq, ab))
    flat = O.unfold(out)
    for c in range(P.n):
        assert abs(int(flat[0, c]) - t * x[c] // P.Q) < 2


@pytest.mark.parametrize("t,k", [(256, 3), (1 << 20, 2), (65537, 3), (65537, 4), (3, 2),
                                 ((1 << 31) - 1, 3)])
def test_simple_scaling_matches_exact_rounding(t, k):
    """round(t*x/Q) mod t exactly (test_bfv_mode.cu:938-983); powers of two as
    the reference, other t < 2^32 through the exact-mod form.  The FP64 sum of
    the fractional parts may round the other way only where t*x/Q lies within
    about k*2^-52 of a half-integer; random x hit that essentially never."""
    P = O.Params(k, k + 1, 1024, t)
    rng = random.Random(17 + t)
    x = [rng.randrange(P.Q) for _ in range(P.n)]
    x[:4] = [0, 1, P.Q - 1, P.Q // 2]
    got = O.unfold(O.simple_scaling(P, residues_of(x, P.q, P.Ng)))
    want = [((2 * t * v + P.Q) // (2 * P.Q)) % t for v in x]
    mism = sum(int(g) != w for g, w in zip(got, want))
    assert mism <= 2, mism


def test_simple_scaling_recovers_scaled_messages():
    """floor(Q/t) m + e with small |e| scales back to m, t = 65537 and 256."""
    for t, k in ((65537, 4), (256, 3)):
        P = O.Params(k, k + 1, 2048, t)
        rng = random.Random(t)
        m = [rng.randrange(t) for _ in range(P.n)]
        x = [((P.Q // t) * mi + rng.randrange(-1000, 1001)) % P.Q for mi in m]
        got = O.unfold(O.simple_scaling(P, residues_of(x, P.q, P.Ng)))
        assert [int(g) for g in got] == m


# --------------------------------------------------------------------------
# Philox and samplers
# --------------------------------------------------------------------------
def test_philox4x32_10_known_answers():
    """Random123 known-answer vectors for philox4x32-10."""
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox([0xffffffff] * 4, [0xffffffff] * 2) == [
        0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                    [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def chi2_stat(counts, probs):
    n = sum(counts)
    return sum((c - n * p) ** 2 / (n * p) for c, p in zip(counts, probs) if n * p > 0)


def test_ternary_chi_square():
    vals = O.sample_small(1 << 16, 2, 99, 3)
    counts = [int(np.sum(vals == v)) for v in (-1, 0, 1)]
    # 2 degrees of freedom: P(chi2 > 13.8) = 0.001
    assert chi2_stat(counts, [1 / 3] * 3) < 13.8


def test_narrow_and_binary_chi_square():
    vals = O.sample_small(1 << 16, 4, 98, 3)
    assert vals.min() == 0 and vals.max() == 2
    assert chi2_stat([int(np.sum(vals == v)) for v in (0, 1, 2)], [1 / 3] * 3) < 13.8
    vals = O.sample_small(1 << 16, 5, 97, 3)
    assert set(np.unique(vals)) == {0, 1}
    assert chi2_stat([int(np.sum(vals == v)) for v in (0, 1)], [0.5, 0.5]) < 10.8  # df 1


def test_fixed_weight_exact_count_uniform_positions_and_signs():
    """ZO (h = n/2) and Hamming weight (h = 64): exactly h nonzeros of +-1;
    over many streams each position is chosen with probability h/n and the
    signs are fair (chi-square, p = 0.001); the chosen set equals a brute-force
    selection of the h smallest keys."""
    import ctypes
    L = O.lib()
    L.or_fixed_weight_key.restype = ctypes.c_uint64
    L.or_fixed_weight_key.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                      ctypes.POINTER(ctypes.c_int)]
    n = 256
    for h in (64, 128):
        hits = np.zeros(n, dtype=np.int64)
        pos = neg = 0
        for stream in range(800):
            v = O.sample_fixed_weight(n, h, 4242, stream)
            assert int(np.count_nonzero(v)) == h and set(np.unique(v)) <= {-1, 0, 1}
            hits += v != 0
            pos += int(np.sum(v == 1))
            neg += int(np.sum(v == -1))
        # df 255: 0.999 quantile ~ 330; positions are not independent across
        # coefficients of one draw, but the sum over draws is multinomial-like
        assert chi2_stat(list(hits), [1 / n] * n) < 330
        assert chi2_stat([pos, neg], [0.5, 0.5]) < 10.8
    neg_ = ctypes.c_int()
    keys = [(L.or_fixed_weight_key(1, 2, j, ctypes.byref(neg_)), j) for j in range(512)]
    assert len({k for k, _ in keys}) == 512  # distinct
    chosen = {j for _, j in sorted(keys)[:64]}
    assert chosen == set(np.nonzero(O.sample_fixed_weight(512, 64, 1, 2))[0].tolist())


def test_gaussian_chi_square_and_bound():
    table, bound, probs = O.cdt()
    vals = O.sample_small(1 << 17, 3, 1234, 5)
    assert vals.min() >= -bound and vals.max() <= bound
    # pool the tails where expected counts are small
    counts, ps = [], []
    n = len(vals)
    tail_c, tail_p = 0, 0.0
    for e, p in zip(range(-bound, bound + 1), probs):
        c = int(np.sum(vals == e))
        if n * p < 20:
            tail_c += c
            tail_p += p
        else:
            counts.append(c)
            ps.append(p)
    counts.append(tail_c)
    ps.append(tail_p)
    df = len(counts) - 1
    # chi-square 0.999 quantile ~ df + 3.1*sqrt(2 df) + 2 (loose, df ~ 20)
    assert chi2_stat(counts, ps) < df + 3.1 * (2 * df) ** 0.5 + 2
    assert abs(float(np.std(vals)) - 3.2) < 0.05


def test_uniform_mod_q_range_and_chi_square():
    P = O.Params(2, 3, 4096, 256)
    u = O.sample_uniform(P.q, P.n, 5, 6)
    for r, p in enumerate(P.q):
        assert int(u[r].max()) < p
        bins = (u[r] // np.uint64(p // 16 + 1)).astype(np.int64)
        counts = np.bincount(bins, minlength=16)
        assert chi2_stat(list(counts), [1 / 16] * 16) < 37.7  # df 15, p = 0.001


def test_hom_mul_composition_matches_exact_rescaled_tensor():
    """HPS multiplication core against exact integers: with inputs lifted to
    [0, Q) and z the signed negacyclic tensor product over Z, the result is
    round(t * z / Q) mod q_j within +-1.  (z < 0 is held as z + QB; scaling
    gives a value just below B, which the B -> Q snap maps back to t z / Q.)"""
    k, n, t = 2, 64, 65537
    P = O.Params(k, k + 1, n, t)
    rng = random.Random(77)
    polys = [[rng.randrange(P.Q) for _ in range(n)] for _ in range(4)]

    def res(vals):
        return O.fold(np.array([[v % q for v in vals] for q in P.q], dtype=np.uint64), P.Ng)

    ct1 = np.stack([res(polys[0]), res(polys[1])])
    ct2 = np.stack([res(polys[2]), res(polys[3])])
    got = O.hom_mul(P, ct1, ct2)

    def negamul(a, b):
        c = [0] * n
        for i in range(n):
            for j in range(n):
                if i + j < n:
                    c[i + j] += a[i] * b[j]
                else:
                    c[i + j - n] -= a[i] * b[j]
        return c

    c0, c1, d0, d1 = polys
    z = [negamul(c0, d0), [x + y for x, y in zip(negamul(c0, d1), negamul(c1, d0))],
         negamul(c1, d1)]
    for m in range(3):
        flat = O.unfold(got[m])
        for c in range(n):
            y = (2 * t * z[m][c] + P.Q) // (2 * P.Q)
            for j, q in enumerate(P.q):
                d = (int(flat[j, c]) - y) % q
                assert d in (0, 1, q - 1), (m, c, j)


def test_oracle_bfv_multiply_relinearize_decrypts():
    """The oracle's relinearization key and relinearization, pinned by what they
    are for: encrypt m1 and m2 (textbook BFV, ternary secret), multiply with the
    oracle's hom_mul (tensor + HPS rescale), relinearize with the key, decrypt
    with the secret alone: m1 * m2 mod (X^n + 1, t).  Also: every key pair
    satisfies b_i + a_i s = rho(s^2)_i - e_i with e_i the Gaussian draw of its
    counters, and the unrelinearized tensor decrypts with s and s^2."""
    n, k, t = 1024, 3, 65537
    P = O.Params(k, k + 1, n, t)
    rng = np.random.default_rng(31)
    s_vals = O.sample_small(n, 2, 17, 0)
    m1, m2 = rng.integers(0, t, n), rng.integers(0, t, n)
    ct1 = O.bfv_encrypt(P, m1, s_vals, 17, 10)
    ct2 = O.bfv_encrypt(P, m2, s_vals, 17, 20)
    assert np.array_equal(O.bfv_decrypt(P, ct1, s_vals), m1)
    key = O.relin_key(P, s_vals, 17, 100)
    sh = O.dgt(P, O.fold(O.small_to_residues(s_vals, P.q), P.Ng), 0, 0)
    rho = O.rho_bfv(P, O.pointwise(P, sh, sh, op=2))
    for i in range(k):
        lhs = O.pointwise(P, key[i, 0], O.pointwise(P, key[i, 1], sh, op=2), op=0)
        e = O.dgt(P, O.pointwise(P, rho[i], lhs, op=1), 0, 1)  # = e_i
        want = O.small_to_residues(O.sample_small(n, 3, 17, 100 + 2 * i + 1), P.q)
        assert np.array_equal(O.unfold(e), want), i
    y = O.hom_mul(P, ct1, ct2)
    want = O.negacyclic_mod_t(m1, m2, t)
    assert np.array_equal(O.bfv_decrypt(P, O.relinearize(P, y, key), s_vals), want)
    # planted mistake: digits present in their own residue only (the
    # reference's xi as it is) satisfy sum_i rho(a)_i xi(b)_i = a b but do
    # not relinearize: the key's noise terms need the digit in every residue
    D = O.relin_digits(P, y[2])
    wrong = []
    for m in range(2):
        acc = np.zeros((k, P.Ng, 2), dtype=np.uint64)
        for i in range(k):
            Di = np.zeros_like(D[i])
            Di[i] = D[i, i]
            acc = O.pointwise(P, O.dgt(P, Di, 0, 0), key[i, m], acc, op=3)
        wrong.append(O.pointwise(P, y[m], O.dgt(P, acc, 0, 1), op=0))
    assert not np.array_equal(O.bfv_decrypt(P, np.stack(wrong), s_vals), want)


@pytest.mark.parametrize("n,k,batch", [(64, 2, 3), (2048, 3, 2), (16384, 4, 1)])
def test_cpu_port_matches_oracle(n, k, batch):
    """The CPU baseline (oracle/cpu_port.c, precomputed Shoup tables, no
    division in its loops) computes the same products as the oracle."""
    P = O.Params(k, k + 1, n, 65537)
    rng = np.random.default_rng(n + k)
    a = np.empty((batch, k, P.Ng, 2), dtype=np.uint64)
    b = np.empty_like(a)
    for j, qj in enumerate(P.q):
        a[:, j] = rng.integers(0, qj, size=(batch, P.Ng, 2), dtype=np.uint64)
        b[:, j] = rng.integers(0, qj, size=(batch, P.Ng, 2), dtype=np.uint64)
    assert np.array_equal(O.port_polymul_batch(P, a, b), O.polymul_batch(P, a, b))
