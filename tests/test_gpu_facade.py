"""The C++ facade (include/cupoly_amd.hpp) used as a cuPoly caller would use
it: tests/cpp/facade_test imports residues, multiplies and adds with
poly_mul / poly_add, extends and scales with the HPS wrappers, selects and
discards residues, and compares with oracle results passed in a file."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_test")


@pytest.mark.parametrize("k,n", [(2, 2048), (4, 16384)])
def test_cpp_facade_drop_in(gpu, tmp_path, k, n):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/facade_test not built (run __graft_entry__.build())")
    P = O.Params(k, k + 1, n, 65537)
    a = O.sample_uniform(P.q, n, 31, 0)
    b = O.sample_uniform(P.q, n, 31, 1)
    sid = int(np.argmax(P.q))  # the largest prime: plant values the others must reduce
    a[sid, :4] = P.q[sid] - 1 - np.arange(4, dtype=np.uint64)
    a[sid, n - 1] = P.q[sid] - 1
    prod = O.unfold(O.polymul_batch(P, O.fold(a, P.Ng)[None], O.fold(b, P.Ng)[None])[0])
    s = np.stack([(a[r] + b[r]) % np.uint64(p) for r, p in enumerate(P.q)])
    a_b = O.unfold(O.basis_ext_q_to_b(P, O.fold(a, P.Ng)))            # [kl][n]
    scaled = O.unfold(O.simple_scaling(P, O.fold(a, P.Ng)))           # [n], mod t
    sel = np.stack([a[sid] % np.uint64(p) for p in P.q])
    assert any(int(a[sid].max()) >= p for p in P.q)
    case = tmp_path / "case.bin"
    with open(case, "wb") as f:
        f.write(np.array([k, k + 1, n, 65537], dtype=np.uint64).tobytes())
        for arr in (a, b, prod, s, a_b, scaled, np.array([sid]), sel):
            f.write(np.ascontiguousarray(arr, dtype=np.uint64).tobytes())
    r = subprocess.run([BIN, str(case)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "facade ok" in r.stdout
