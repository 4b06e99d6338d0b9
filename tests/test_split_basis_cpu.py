"""The split basis of the product kernels (cupoly_amd/csrc/dgt_dev.hpp), restated
in Python integers and checked against the schoolbook negacyclic product.

For the 63-bit bucket primes omega of order N = n/2 lies in Z_p, so p = 1 mod 4
and Z_p[i] = Z_p x Z_p through a + ib -> (a + Jb, a - Jb), J = omega^(N/4).  The
kernels keep the reference's forward network (twisted DIF with the real
omega) on the two components, multiply slot-wise with two products mod p,
undo the split tree of X^N -+ J with Gentleman-Sande butterflies whose roots
come from the leaves xi+-^-1 omega^-bitrev(l) by squaring (the engine's table
construction), and convert back with 1/2N and 1/2JN.  These tests pin that
construction: the product equals the schoolbook product, the DGT-domain values
map slot-wise onto the split-basis spectra (which the relinearization key's
split copy relies on), and planted mistakes are caught."""
import json
import os
import random

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _tables():
    with open(os.path.join(HERE, "golden", "tables63.json")) as f:
        d = json.load(f)

    def conv(x):
        return [conv(y) for y in x] if isinstance(x, list) else int(x)

    return {k: conv(v) for k, v in d.items()}


T = _tables()


def brv(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


class Split:
    """Tables of one (prime, N) as the engine builds them."""

    def __init__(self, r, N, wrong_root=False):
        p = T["primes"][r]
        self.p, self.N, self.S = p, N, N.bit_length() - 1
        si = T["sizes"].index(N)
        self.om = pow(T["proot"][r], (p - 1) // N, p)
        omi = pow(self.om, p - 2, p)
        self.J = pow(self.om, N // 4, p)
        xr, xim = T["nthroot"][r][si]
        self.xi = (xr, xim)
        self.xs = [(xr + self.J * xim) % p, (xr - self.J * xim) % p]
        # inverse roots at node 2^s + g, per component: leaves xi^-1 omi^bitrev(l)
        self.root = [[0] * N, [0] * N]
        for c in range(2):
            xinv = pow(self.xs[c], p - 2, p)
            cur = [xinv * pow(omi, brv(l, self.S), p) % p for l in range(N)]
            for s in range(self.S - 1, -1, -1):
                nxt = []
                for g in range(1 << s):
                    ri = cur[2 * g]
                    self.root[c][(1 << s) + g] = ri
                    nxt.append(ri * ri % p)
                cur = nxt
        if wrong_root:  # a plausible slip: the root of the sibling group
            for c in range(2):
                self.root[c][2], self.root[c][3] = self.root[c][3], self.root[c][2]

    def dif(self, x):
        """The forward network of the kernels and the reference's DGT: natural
        order in, bitrev out, twiddle omega^(j 2^s)."""
        p, N = self.p, self.N
        x = list(x)
        for s in range(self.S):
            m = N >> (s + 1)
            for g in range(1 << s):
                for j in range(m):
                    a, b = x[g * 2 * m + j], x[g * 2 * m + j + m]
                    x[g * 2 * m + j] = (a + b) % p
                    x[g * 2 * m + j + m] = (a - b) * pow(self.om, j << s, p) % p
        return x

    def forward(self, coef):
        """coef: n = 2N coefficients -> (U, V) spectra."""
        p, N = self.p, self.N
        t = [self.J * coef[g + N] % p for g in range(N)]
        u = [(coef[g] + t[g]) * pow(self.xs[0], g, p) % p for g in range(N)]
        v = [(coef[g] - t[g]) * pow(self.xs[1], g, p) % p for g in range(N)]
        return self.dif(u), self.dif(v)

    def inverse(self, U, V):
        p, N = self.p, self.N
        out = []
        for c, X in enumerate((U, V)):
            x = list(X)
            for s in range(self.S - 1, -1, -1):
                m = N >> (s + 1)
                for g in range(1 << s):
                    ri = self.root[c][(1 << s) + g]
                    for j in range(m):
                        a, b = x[g * 2 * m + j], x[g * 2 * m + j + m]
                        x[g * 2 * m + j] = (a + b) % p
                        x[g * 2 * m + j + m] = (a - b) * ri % p
            out.append(x)
        U, V = out
        h = pow(2 * N, p - 2, p)
        hj = h * pow(self.J, p - 2, p) % p
        return [(U[g] + V[g]) * h % p for g in range(N)] + [(U[g] - V[g]) * hj % p for g in range(N)]

    def product(self, a, b):
        Ua, Va = self.forward(a)
        Ub, Vb = self.forward(b)
        return self.inverse([x * y % self.p for x, y in zip(Ua, Ub)],
                            [x * y % self.p for x, y in zip(Va, Vb)])


def schoolbook(a, b, p):
    n = len(a)
    c = [0] * n
    for i in range(n):
        for j in range(n):
            if i + j < n:
                c[i + j] = (c[i + j] + a[i] * b[j]) % p
            else:
                c[i + j - n] = (c[i + j - n] - a[i] * b[j]) % p
    return c


@pytest.mark.parametrize("N", [16, 32, 64])
@pytest.mark.parametrize("r", [0, 7, 29])
def test_split_basis_product_equals_schoolbook(N, r):
    sp = Split(r, N)
    p = sp.p
    assert sp.J * sp.J % p == p - 1
    assert pow(sp.xs[0], N, p) == sp.J and pow(sp.xs[1], N, p) == p - sp.J
    rng = random.Random(N * 31 + r)
    for trial in range(3):
        if trial == 0:  # edge residues
            a = [p - 1] * (2 * N)
            b = [p - 1 if i % 3 else 1 for i in range(2 * N)]
        else:
            a = [rng.randrange(p) for _ in range(2 * N)]
            b = [rng.randrange(p) for _ in range(2 * N)]
        assert sp.product(a, b) == schoolbook(a, b, p)


def test_planted_root_mistake_is_caught():
    sp = Split(3, 32, wrong_root=True)
    rng = random.Random(5)
    a = [rng.randrange(sp.p) for _ in range(64)]
    b = [rng.randrange(sp.p) for _ in range(64)]
    assert sp.product(a, b) != schoolbook(a, b, sp.p)


@pytest.mark.parametrize("r", [1, 12])
def test_dgt_domain_maps_slotwise_to_split_spectra(r):
    """The relinearization key is converted slot-wise, X -> (X.re + J X.im,
    X.re - J X.im): the reference's DGT spectrum of a (twist by the Gaussian xi,
    same DIF network) maps onto the split-basis spectra position by position."""
    N = 32
    sp = Split(r, N)
    p = sp.p
    rng = random.Random(r)
    a = [rng.randrange(p) for _ in range(2 * N)]
    # reference DGT: A_g = a_g + i a_{g+N}, twisted by xi^g, then the DIF network
    xr, xim = sp.xi
    pr, pi = 1, 0
    re, im = [], []
    for g in range(N):
        ar, ai = a[g], a[g + N]
        re.append((ar * pr - ai * pi) % p)
        im.append((ar * pi + ai * pr) % p)
        pr, pi = (pr * xr - pi * xim) % p, (pr * xim + pi * xr) % p
    Xr, Xi = sp.dif(re), sp.dif(im)
    U, V = sp.forward(a)
    assert [(x + sp.J * y) % p for x, y in zip(Xr, Xi)] == U
    assert [(x - sp.J * y) % p for x, y in zip(Xr, Xi)] == V
    # and the mapping with the wrong sign of J on V is not
    assert [(x + sp.J * y) % p for x, y in zip(Xr, Xi)] != V
