"""Python side of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker or as the timed CPU
baseline.  The product (``cupoly_amd``) never imports it.

It wraps ``oracle/liboracle.so`` (built from ``cupoly_oracle.c`` by
``oracle/Makefile``) with numpy, and computes the engine's RNS constants with
exact Python integers, restating ``CUDAEngine::gen_rns_primes`` and
``CUDAEngine::precompute`` (``src/cuda/cudaengine.cu:1249-1511``) and the DGT
root choice of ``DGTEngine::init`` (``src/cuda/dgt.cu:1694-1982``).

Parity status of each oracle function is documented in the C file's header
and in the tests that pin it.
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "liboracle.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_dblp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)

_lib = None


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        u64, i32, dbl = ctypes.c_uint64, ctypes.c_int, ctypes.c_double
        L.or_mulmod.restype = u64
        L.or_mulmod.argtypes = [u64, u64, u64]
        L.or_powmod.restype = u64
        L.or_powmod.argtypes = [u64, u64, u64]
        L.or_dgt_forward.argtypes = [_u64p, i32, u64, _u64p, u64]
        L.or_dgt_inverse.argtypes = [_u64p, i32, u64, _u64p, u64]
        L.or_dgt_poly.argtypes = [_u64p, i32, i32, _u64p, _u64p, _u64p, i32]
        L.or_pointwise.argtypes = [_u64p, _u64p, _u64p, _u64p, i32, i32, _u64p, i32]
        L.or_polymul_batch.argtypes = [_u64p, _u64p, _u64p, i32, i32, i32, _u64p,
                                       _u64p, _u64p, _u64p]
        L.or_negacyclic_schoolbook.argtypes = [_u64p, _u64p, _u64p, i32, u64]
        L.or_fma_rd.restype = dbl
        L.or_fma_rd.argtypes = [dbl, dbl, dbl]
        L.or_rcp_rd.restype = dbl
        L.or_rcp_rd.argtypes = [u64]
        L.or_basis_ext_q_to_b.argtypes = [_u64p, _u64p, i32, i32, i32, _u64p, _u64p,
                                          _u64p, _u64p, _u64p, _dblp, _u64p]
        L.or_complex_scaling.argtypes = [_u64p, _u64p, _u64p, i32, i32, i32, _u64p,
                                         _u64p, _u64p, _u64p]
        L.or_basis_ext_b_to_q.argtypes = [_u64p, _u64p, i32, i32, i32, _u64p, _u64p,
                                          _u64p, _u64p, _u64p, _u64p]
        L.or_simple_scaling.argtypes = [_u64p, _u64p, i32, i32, u64, _u64p, _u64p]
        L.or_scale_residue.argtypes = [_u64p, _u64p, i32, u64, u64]
        L.or_sample_fixed_weight.argtypes = [_i64p, i32, i32, u64, u64]
        L.or_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.or_sample_uniform_poly.argtypes = [_u64p, i32, i32, _u64p, u64, u64]
        L.or_sample_small_poly.argtypes = [_i64p, i32, i32, u64, u64, _u64p, i32]
        _lib = L
    return _lib


def _p(a, t=_u64p):
    return a.ctypes.data_as(t)


def u64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


# --------------------------------------------------------------------------
# Tables and parameters
# --------------------------------------------------------------------------
_tables = None
_cdt = None


def tables():
    global _tables
    if _tables is None:
        with open(os.path.join(REPO, "tests", "golden", "tables63.json")) as f:
            t = json.load(f)
        t["primes"] = [int(p) for p in t["primes"]]
        t["nthroot"] = [[(int(a), int(b)) for a, b in row] for row in t["nthroot"]]
        t["invnthroot"] = [[(int(a), int(b)) for a, b in row] for row in t["invnthroot"]]
        _tables = t
    return _tables


def cdt():
    global _cdt
    if _cdt is None:
        with open(os.path.join(REPO, "tests", "golden", "cdt.json")) as f:
            c = json.load(f)
        _cdt = (np.array([int(x) for x in c["cdt"]], dtype=np.uint64), c["bound"],
                [float(x) for x in c["probs"]])
    return _cdt


def prod(xs):
    r = 1
    for x in xs:
        r *= x
    return r


class Params:
    """RNS parameters for (k, kl, n, t): Q = first k bucket primes, B = next kl
    (cudaengine.cu:1265-1289), roots for Ng = n/2 (dgt.cu:1526-1538, 1900-1904)."""

    def __init__(self, k, kl, n, t):
        T = tables()
        self.k, self.kl, self.n, self.t = k, kl, n, t
        self.Ng = n // 2
        si = T["sizes"].index(self.Ng)
        allp = T["primes"][:k + kl]
        self.q = allp[:k]
        self.b = allp[k:]
        self.primes = allp
        self.xi = [T["nthroot"][r][si] for r in range(k + kl)]
        self.xiinv = [T["invnthroot"][r][si] for r in range(k + kl)]
        self.omega = [pow(T["proot"][r], (p - 1) // self.Ng, p) for r, p in enumerate(allp)]
        Q, B = prod(self.q), prod(self.b)
        self.Q, self.B = Q, B
        q, b = self.q, self.b
        # cudaengine.cu:1363-1411
        self.invQqi = [pow((Q // qj) % qj, -1, qj) for qj in q]
        self.Qqi = [(Q // qj) % qj for qj in q]
        self.Qqibi = [(Q // qj) % bi for bi in b for qj in q]          # [i][j]
        self.Qbi = [Q % bi for bi in b]
        self.invBbi = [pow((B // bi) % bi, -1, bi) for bi in b]
        self.Bbiqi = [(B // bi) % qj for qj in q for bi in b]          # [j][i]
        self.Bqi = [B % qj for qj in q]
        # cudaengine.cu:1463-1511 (exact; the reference uses 512-bit RR)
        self.lam = [t * pow((B * Q // bi) % bi, -1, bi) * (B // bi) % bi for bi in b]
        self.wint, self.wfrac = [], []
        for bi in b:
            for qj in q:
                aux = t * pow((Q * B // qj) % qj, -1, qj) * B
                self.wint.append((aux // qj) % bi)
                fr = ((aux % qj) << 128) // qj
                self.wfrac += [fr & (2**64 - 1), fr >> 64]
        self.twint, self.twfrac = [], []
        for qj in q:
            aux = t * pow((Q // qj) % qj, -1, qj)
            self.twint.append(aux // qj)
            fr = ((aux % qj) << 128) // qj
            self.twfrac += [fr & (2**64 - 1), fr >> 64]

    def base_primes(self, base):
        return {0: self.q, 1: self.b, 2: self.primes}[base]

    def base_roots(self, base):
        off = {0: 0, 1: self.k, 2: 0}[base]
        nres = len(self.base_primes(base))
        return (self.xi[off:off + nres], self.xiinv[off:off + nres],
                self.omega[off:off + nres])


# --------------------------------------------------------------------------
# numpy wrappers.  A "GI array" is uint64 [..., Ng, 2] (re, imag).
# --------------------------------------------------------------------------
def fold(coefs, Ng):
    """Flat coefficients [..., nres, 2Ng] -> GI [..., nres, Ng, 2]
    (poly_crt fold, polynomial.cu:206-217)."""
    c = np.asarray(coefs, dtype=np.uint64)
    return np.ascontiguousarray(np.stack([c[..., :Ng], c[..., Ng:]], axis=-1))


def unfold(gi):
    g = np.asarray(gi)
    return np.ascontiguousarray(np.concatenate([g[..., 0], g[..., 1]], axis=-1))


def dgt(P, data, base, direction):
    """DGTEngine::execute_dgt restated: data GI [nres, Ng, 2] (copy returned)."""
    x = u64(data).copy()
    pr = P.base_primes(base)
    xi, xiinv, om = P.base_roots(base)
    roots = xi if direction == 0 else xiinv
    lib().or_dgt_poly(_p(x), P.Ng, len(pr), _p(u64(pr)), _p(u64(roots).reshape(-1)),
                      _p(u64(om)), direction)
    return x


def pointwise(P, a, b, c=None, base=0, op=2):
    pr = P.base_primes(base)
    a, b = u64(a), u64(b)
    out = np.empty_like(a)
    cc = u64(c) if c is not None else a
    lib().or_pointwise(_p(out), _p(a), _p(b), _p(cc), P.Ng, len(pr), _p(u64(pr)), op)
    return out


def polymul_batch(P, a, b, base=0):
    """Coefficient-domain products; a, b GI [batch, nres, Ng, 2]."""
    pr = P.base_primes(base)
    xi, xiinv, om = P.base_roots(base)
    a, b = u64(a), u64(b)
    c = np.empty_like(a)
    batch = a.shape[0]
    lib().or_polymul_batch(_p(c), _p(a), _p(b), batch, P.Ng, len(pr), _p(u64(pr)),
                           _p(u64(xi).reshape(-1)), _p(u64(xiinv).reshape(-1)),
                           _p(u64(om)))
    return c


def xi_bfv(P, a, outs=None):
    """poly_xi_bfv restated (cudaengine.cu:992-1035): residue i of output i is
    a_i (Q/q_i)^-1 mod q_i; other residues keep outs' contents (zeros)."""
    a = u64(a)
    outs = [np.zeros_like(a) for _ in range(P.k)] if outs is None else [u64(o).copy() for o in outs]
    for i in range(P.k):
        lib().or_scale_residue(_p(outs[i][i]), _p(np.ascontiguousarray(a[i])), P.Ng, P.q[i],
                               P.invQqi[i])
    return outs


def rho_bfv(P, a, outs=None):
    """poly_rho_bfv restated (cudaengine.cu:1037-1062): residue i < k of output
    i is a_i (Q/q_i) mod q_i; a in base Q or QB."""
    a = u64(a)
    outs = [np.zeros_like(a) for _ in range(P.k)] if outs is None else [u64(o).copy() for o in outs]
    for i in range(P.k):
        lib().or_scale_residue(_p(outs[i][i]), _p(np.ascontiguousarray(a[i])), P.Ng, P.q[i],
                               P.Qqi[i])
    return outs


def dot(P, a_list, b_list, base=0):
    """poly_dot (polynomial.cu:1439-1451): poly_mul then poly_mul_add."""
    acc = pointwise(P, a_list[0], b_list[0], base=base, op=2)
    for x, y in zip(a_list[1:], b_list[1:]):
        acc = pointwise(P, x, y, acc, base=base, op=3)
    return acc


def schoolbook(a, b, p):
    a, b = u64(a), u64(b)
    c = np.empty_like(a)
    lib().or_negacyclic_schoolbook(_p(c), _p(a), _p(b), len(a), p)
    return c


def rcp_rd(q):
    return lib().or_rcp_rd(q)


def fma_rd(a, b, c):
    return lib().or_fma_rd(a, b, c)


def basis_ext_q_to_b(P, x_q, want_v=False):
    """x_q GI [k, Ng, 2] coefficient domain -> GI [kl, Ng, 2]."""
    out = np.zeros((P.kl, P.Ng, 2), dtype=np.uint64)
    v = np.zeros(P.n, dtype=np.uint64)
    rcp = np.array([rcp_rd(qj) for qj in P.q], dtype=np.float64)
    lib().or_basis_ext_q_to_b(_p(out), _p(u64(x_q)), P.Ng, P.k, P.kl, _p(u64(P.q)),
                              _p(u64(P.b)), _p(u64(P.invQqi)), _p(u64(P.Qqibi)),
                              _p(u64(P.Qbi)), _p(rcp, _dblp), _p(v))
    return (out, v) if want_v else out


def complex_scaling(P, a_q, a_b):
    """a_q GI [k, Ng, 2], a_b GI [kl, Ng, 2] -> uint64 [kl, 2Ng]."""
    out = np.zeros((P.kl, P.n), dtype=np.uint64)
    lib().or_complex_scaling(_p(out), _p(u64(a_q)), _p(u64(a_b)), P.Ng, P.k, P.kl,
                             _p(u64(P.b)), _p(u64(P.wint)), _p(u64(P.wfrac)),
                             _p(u64(P.lam)))
    return out


def basis_ext_b_to_q(P, x_b, want_v=False):
    """x_b uint64 [kl, 2Ng] -> GI [k, Ng, 2]."""
    out = np.zeros((P.k, P.Ng, 2), dtype=np.uint64)
    v = np.zeros(P.n, dtype=np.uint64)
    lib().or_basis_ext_b_to_q(_p(out), _p(u64(x_b)), P.Ng, P.k, P.kl, _p(u64(P.q)),
                              _p(u64(P.b)), _p(u64(P.invBbi)), _p(u64(P.Bbiqi)),
                              _p(u64(P.Bqi)), _p(v))
    return (out, v) if want_v else out


def simple_scaling(P, x_q):
    out = np.zeros((P.Ng, 2), dtype=np.uint64)
    lib().or_simple_scaling(_p(out), _p(u64(x_q)), P.Ng, P.k, P.t, _p(u64(P.twint)),
                            _p(u64(P.twfrac)))
    return out


def philox(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_p(c, _u32p), _p(k, _u32p), _p(o, _u32p))
    return [int(x) for x in o]


def sample_uniform(primes, n, seed, stream):
    out = np.zeros((len(primes), n), dtype=np.uint64)
    lib().or_sample_uniform_poly(_p(out), n, len(primes), _p(u64(primes)), seed, stream)
    return out


def sample_fixed_weight(n, h, seed, stream):
    """Exactly min(h, n) coefficients +-1 at the smallest keys.  int64 [n]."""
    out = np.zeros(n, dtype=np.int64)
    lib().or_sample_fixed_weight(_p(out, _i64p), n, h, seed, stream)
    return out


def sample_small(n, kind, seed, stream):
    """kind 2 = ternary, 3 = discrete Gaussian, 4 = narrow {0,1,2},
    5 = binary {0,1}.  Returns int64 [n]."""
    table, bound, _ = cdt()
    out = np.zeros(n, dtype=np.int64)
    lib().or_sample_small_poly(_p(out, _i64p), n, kind, seed, stream, _p(table), bound)
    return out


def small_to_residues(vals, primes):
    """Signed small coefficients -> flat residues [nres, n]."""
    v = np.asarray(vals, dtype=np.int64)
    return np.stack([np.where(v < 0, np.uint64(p) - (-v).astype(np.uint64),
                              v.astype(np.uint64)) for p in primes])


def crt(P, x_q_gi):
    """GI [k, Ng, 2] residues -> list of n Python ints in [0, Q)."""
    flat = unfold(x_q_gi)
    out = []
    for c in range(P.n):
        acc = 0
        for j, qj in enumerate(P.q):
            Mj = P.Q // qj
            acc += int(flat[j, c]) * P.invQqi[j] % qj * Mj
        out.append(acc % P.Q)
    return out


def hom_mul(P, ct1, ct2):
    """BFV multiplication core restated from the reference's pieces
    (polynomial.cu:350-457, 951-999): Q->B of the four inputs, DGT in QB,
    dr2 tensor, inverse DGT, complex scaling and B->Q.  ct1, ct2 GI
    [2, k, Ng, 2] base Q coefficient domain -> [3, k, Ng, 2]."""
    k = P.k
    spec = []
    for x in (ct1[0], ct1[1], ct2[0], ct2[1]):
        xb = basis_ext_q_to_b(P, x)
        spec.append(dgt(P, np.concatenate([u64(x), xb]), 2, 0))
    c0, c1, d0, d1 = spec
    z0 = pointwise(P, c0, d0, base=2, op=2)
    z2 = pointwise(P, c1, d1, base=2, op=2)
    z1 = pointwise(P, pointwise(P, c0, d1, base=2, op=2), pointwise(P, c1, d0, base=2, op=2),
                   base=2, op=0)
    out = []
    for z in (z0, z1, z2):
        zc = dgt(P, z, 2, 1)
        out.append(basis_ext_b_to_q(P, complex_scaling(P, zc[:k], zc[k:])))
    return np.stack(out)


def sample_muladd(P, a, seed, stream0):
    """a GI [batch, k, Ng, 2] (DGT domain) -> a*s + e with s ternary (stream
    stream0 + 2i) and e discrete Gaussian (stream0 + 2i + 1), the reference's
    Sampler::sample + poly_mul_add (sampler.cu:576-635, polynomial.cu:916-948)."""
    out = np.empty_like(u64(a))
    for i in range(a.shape[0]):
        s = fold(small_to_residues(sample_small(P.n, 2, seed, stream0 + 2 * i), P.q), P.Ng)
        e = fold(small_to_residues(sample_small(P.n, 3, seed, stream0 + 2 * i + 1), P.q), P.Ng)
        out[i] = pointwise(P, a[i], dgt(P, s, 0, 0), dgt(P, e, 0, 0), base=0, op=3)
    return out


def relin_key(P, s_vals, seed, stream0):
    """Relinearization key for the secret with signed coefficients s_vals [n],
    as the library defines it (BFV relinearization over the reference's
    rho, polynomial.cu:610-693): a_i = DGT(uniform, stream stream0 + 2i),
    e_i = DGT(Gaussian, stream0 + 2i + 1), b_i = -(a_i s + e_i) + rho(s^2)_i,
    all in the DGT domain of base Q.  -> [k, 2, k, Ng, 2] ([i][0] = b_i)."""
    n, k = P.n, P.k
    sh = dgt(P, fold(small_to_residues(s_vals, P.q), P.Ng), 0, 0)
    rho = rho_bfv(P, pointwise(P, sh, sh, op=2))
    key = np.empty((k, 2, k, P.Ng, 2), dtype=np.uint64)
    for i in range(k):
        a = dgt(P, fold(sample_uniform(P.q, n, seed, stream0 + 2 * i), P.Ng), 0, 0)
        e = dgt(P, fold(small_to_residues(sample_small(n, 3, seed, stream0 + 2 * i + 1), P.q),
                        P.Ng), 0, 0)
        key[i, 1] = a
        key[i, 0] = pointwise(P, rho[i], pointwise(P, a, sh, e, op=3), op=1)
    return key


def relin_digits(P, y2):
    """RNS digits of y2 (GI [k, Ng, 2], coefficient domain): digit i is the
    integer [y2_i (Q/q_i)^-1]_{q_i} (residue i of the reference's xi,
    cudaengine.cu:992-1035), written to every residue j as its value mod q_j.
    -> [k (digit), k (residue), Ng, 2]."""
    k = P.k
    y2 = u64(y2)
    out = np.empty((k, k, P.Ng, 2), dtype=np.uint64)
    for i in range(k):
        d = np.zeros((P.Ng, 2), dtype=np.uint64)
        lib().or_scale_residue(_p(d), _p(np.ascontiguousarray(y2[i])), P.Ng, P.q[i], P.invQqi[i])
        for j, qj in enumerate(P.q):
            out[i, j] = d % np.uint64(qj)
    return out


def relinearize(P, y, key):
    """y = (y0, y1, y2) GI [3, k, Ng, 2] coefficient domain, key [k, 2, k, Ng, 2]
    (DGT domain) -> (y0 + sum_i D_i b_i, y1 + sum_i D_i a_i) [2, k, Ng, 2]."""
    y = u64(y)
    D = relin_digits(P, y[2])
    out = []
    for m in range(2):
        acc = np.zeros((P.k, P.Ng, 2), dtype=np.uint64)
        for i in range(P.k):
            acc = pointwise(P, dgt(P, D[i], 0, 0), key[i, m], acc, op=3)
        out.append(pointwise(P, y[m], dgt(P, acc, 0, 1), op=0))
    return np.stack(out)


def negacyclic_mod_t(m1, m2, t):
    """m1 * m2 mod (X^n + 1, t) for integer vectors in [0, t), t < 2^17."""
    n = len(m1)
    full = np.convolve(np.asarray(m1, dtype=np.int64), np.asarray(m2, dtype=np.int64))
    res = full[:n].copy()
    res[:n - 1] -= full[n:]
    return np.mod(res, t)


def bfv_encrypt(P, m, s_vals, seed, stream):
    """Textbook BFV encryption of m (integers in [0, t), length n) under the
    secret with signed coefficients s_vals: c1 = a uniform, c0 = Delta m -
    a s + e, Delta = floor(Q / t), e discrete Gaussian.  -> GI [2, k, Ng, 2]."""
    a = fold(sample_uniform(P.q, P.n, seed, stream), P.Ng)
    e = fold(small_to_residues(sample_small(P.n, 3, seed, stream + 1), P.q), P.Ng)
    s = fold(small_to_residues(s_vals, P.q), P.Ng)
    delta = P.Q // P.t
    dm = fold(np.array([[(int(x) * (delta % qj)) % qj for x in m] for qj in P.q],
                       dtype=np.uint64), P.Ng)
    a_s = polymul_batch(P, a[None], s[None])[0]
    c0 = pointwise(P, pointwise(P, dm, e, op=0), a_s, op=1)
    return np.stack([c0, a])


def bfv_decrypt(P, ct, s_vals):
    """round(t/Q [c0 + c1 s]_Q) mod t for a 2-polynomial ciphertext -> int64 [n]."""
    s = fold(small_to_residues(s_vals, P.q), P.Ng)
    x = pointwise(P, u64(ct[0]), polymul_batch(P, u64(ct[1])[None], s[None])[0], op=0)
    return unfold(simple_scaling(P, x)).astype(np.int64)


# --------------------------------------------------------------------------
# CPU baseline: the optimised CPU port (cpu_port.c), timed by the benchmark.
# --------------------------------------------------------------------------
_port = None
_port_plans = {}


def port_lib():
    global _port
    if _port is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcpuport.so")
        L = ctypes.CDLL(path)
        L.cpu_plan_create.restype = ctypes.c_void_p
        L.cpu_plan_create.argtypes = [ctypes.c_int, ctypes.c_int, _u64p, _u64p, _u64p, _u64p]
        L.cpu_plan_destroy.argtypes = [ctypes.c_void_p]
        L.cpu_polymul_batch.restype = ctypes.c_int
        L.cpu_polymul_batch.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p, ctypes.c_int]
        _port = L
    return _port


def port_polymul_batch(P, a, b):
    """Coefficient-domain products c[i] = a[i] * b[i] (GI [batch, k, Ng, 2]) by
    the optimised CPU port."""
    key = (P.k, P.Ng)
    if key not in _port_plans:
        xi, xiinv, om = P.base_roots(0)
        h = port_lib().cpu_plan_create(P.Ng, P.k, _p(u64(P.q)), _p(u64(xi).reshape(-1)),
                                       _p(u64(xiinv).reshape(-1)), _p(u64(om)))
        if not h:
            raise MemoryError("cpu_port: plan allocation failed")
        _port_plans[key] = h
    a, b = np.ascontiguousarray(u64(a)), np.ascontiguousarray(u64(b))
    c = np.empty_like(a)
    if port_lib().cpu_polymul_batch(_port_plans[key], _p(c), _p(a), _p(b), a.shape[0]) != 0:
        raise MemoryError("cpu_port: working buffer allocation failed")
    return c
