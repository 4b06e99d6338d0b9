"""Argument checks of the Python bindings, on the CPU.

Every wrapper hands the library raw device addresses, so a tensor of the
wrong element type, a host tensor or a non-contiguous view must be refused
with ValueError before any library entry point runs.  The library is replaced
by a recorder here: a call that reaches it fails the test.
"""
import types

import pytest
import torch

import cupoly_amd as cp


class _Recorder:
    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def fn(*args):
            self.calls.append(name)
            return 0
        return fn


@pytest.fixture
def fake(monkeypatch):
    rec = _Recorder()
    monkeypatch.setattr(cp, "lib", lambda: rec)
    k, kl, n = 2, 3, 64
    eng = types.SimpleNamespace(k=k, kl=kl, n=n, N=n // 2, t=65537,
                                get_n_residues=lambda base: {0: k, 1: kl, 2: k + kl}[base])
    ctx = types.SimpleNamespace(engine=eng, handle=None)
    return rec, ctx, eng


def _bad_variants(shape):
    """(label, tensor) for each defect: wrong dtype (right shape), non-contiguous
    int64, contiguous int64 on the host."""
    good_shape_i32 = torch.zeros(shape, dtype=torch.int32)
    big = torch.zeros(shape[:-1] + (shape[-1] * 2,), dtype=torch.int64)
    noncontig = big[..., ::2]
    assert tuple(noncontig.shape) == tuple(shape) and not noncontig.is_contiguous()
    host = torch.zeros(shape, dtype=torch.int64)
    return [("dtype", good_shape_i32, "int64 or uint64"),
            ("noncontig", noncontig, "contiguous"),
            ("host", host, "device")]


def _wrappers(ctx, eng):
    """(name, callable taking one operand tensor) for wrappers whose operands
    all have the shape of that tensor's slot."""
    k, N = eng.k, eng.N
    q = (k, N, 2)
    b = (3, k, N, 2)
    return [
        ("execute_dgt", q, lambda t: cp.execute_dgt(ctx, t, cp.QBase, cp.FORWARD)),
        ("execute_add_dgt", q, lambda t: cp.execute_add_dgt(ctx, t, t, t)),
        ("execute_mul_dgt_gi", q, lambda t: cp.execute_mul_dgt_gi(ctx, t, t, t)),
        ("execute_polynomial_negate", q, lambda t: cp.execute_polynomial_negate(ctx, t, t)),
        ("execute_polynomial_double_op_by_int", q,
         lambda t: cp.execute_polynomial_double_op_by_int(ctx, t, t, 3, t, t, 5, cp.MULMUL)),
        ("execute_polynomial_simple_scaling", q, lambda t: cp.execute_polynomial_simple_scaling(ctx, t)),
        ("sample", q, lambda t: cp.Sampler.sample(ctx, t, cp.UNIFORM, 1, 2)),
        ("polymul_batch", b, lambda t: cp.polymul_batch(ctx, t, t, t)),
        ("polymul_plain", b, lambda t: cp.polymul_plain(ctx, t, t, t)),
        ("sample_muladd_batch", b, lambda t: cp.sample_muladd_batch(ctx, t, t, 1, 0)),
        ("dot", q, lambda t: cp.dot(ctx, t, [t], [t])),
    ]


def test_pointer_checks_refuse_before_any_library_call(fake):
    rec, ctx, eng = fake
    for name, shape, call in _wrappers(ctx, eng):
        for label, t, msg in _bad_variants(shape):
            with pytest.raises(ValueError, match=msg):
                call(t)
            assert rec.calls == [], "%s/%s reached the library: %s" % (name, label, rec.calls)


def test_hom_mul_gather_checks_every_pointer(fake):
    rec, ctx, eng = fake
    k, N = eng.k, eng.N
    for label, bad_out, msg in _bad_variants((3, k, N, 2)):
        # good inputs cannot be built without a GPU; the output list is
        # converted first, so a bad output alone must stop the call
        ct = torch.zeros((2, k, N, 2), dtype=torch.int64)
        with pytest.raises(ValueError, match=msg):
            cp.hom_mul_gather(ctx, [bad_out], [ct], [ct])
        assert rec.calls == [], label


def test_uneven_plan_checks_every_pointer(fake):
    rec, ctx, eng = fake
    n, k = 64, 2
    for label, bad, msg in _bad_variants((k, n // 2, 2)):
        engine = types.SimpleNamespace(handle=types.SimpleNamespace(value=None))
        with pytest.raises(ValueError, match=msg):
            cp.UnevenPlan(ctx, [engine], [(n, k, bad, bad, bad)])
        assert rec.calls == [], label


def test_hom_mul_plan_checks_every_pointer(fake):
    rec, ctx, eng = fake
    eng.handle = types.SimpleNamespace(value=None)
    k, N = eng.k, eng.N
    ct = torch.zeros((2, k, N, 2), dtype=torch.int64)
    for label, bad, msg in _bad_variants((3, k, N, 2)):
        # outputs are converted first: a bad output alone stops the call
        with pytest.raises(ValueError, match=msg):
            cp.HomMulPlan(ctx, [eng], [(0, bad, ct, ct)])
        assert rec.calls == [], label
    with pytest.raises(ValueError, match="names no shape"):
        cp.HomMulPlan(ctx, [eng], [(1, bad, bad, bad)])
    assert rec.calls == []


def test_shape_mismatch_is_refused(fake):
    rec, ctx, eng = fake
    t = torch.zeros((eng.k + 1, eng.N, 2), dtype=torch.int64)
    with pytest.raises(ValueError, match="shape"):
        cp.execute_add_dgt(ctx, t, t, t)
    assert rec.calls == []
