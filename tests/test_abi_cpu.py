"""CPU-only checks of the C ABI library: it loads, exports every symbol the
header declares, rejects bad arguments before touching the device, and its
host-side RNS constants equal the oracle's exact Python-integer values."""
import ctypes
import os
import re
import struct

import pytest

import cupoly_amd as cp
from oracle import oracle as O

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "cupoly_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cupoly_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = cp.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(cp.SIGNATURES), set(names) ^ set(cp.SIGNATURES)


def test_version_string():
    assert b"gfx950" in cp.lib().cupoly_version()


@pytest.mark.parametrize("k,kl,n,t,code", [
    (0, 1, 2048, 256, -1),      # k < 1
    (4, 15, 2048, 256, -1),     # kl too large (MAX_COPRIMES_IN_A_BASE)
    (15, 3, 2048, 256, -1),     # k too large
    (2, 3, 2047, 256, -2),      # n not a power of two
    (2, 3, 16, 256, -2),        # no root of i tabulated for N = 8
    (2, 3, 131072, 256, -2),    # above the largest tabulated size
    (2, 3, 2048, 1, -2),        # t < 2
])
def test_engine_create_rejects_bad_parameters(k, kl, n, t, code):
    h = ctypes.c_void_p()
    rc = cp.lib().cupoly_engine_create(k, kl, n, t, ctypes.byref(h))
    assert rc == code
    assert not h.value
    assert cp.lib().cupoly_last_error()


def test_null_arguments_are_errors():
    L = cp.lib()
    assert L.cupoly_dgt(None, None, 0, 0) == -1
    assert L.cupoly_polymul_batch(None, None, None, None, 1) == -1
    assert L.cupoly_engine_create(2, 3, 2048, 256, None) == -1
    assert L.cupoly_get_n_residues(None, 0) == -4
    assert L.cupoly_ctx_destroy(None) == -1


@pytest.mark.parametrize("k,kl,t", [(1, 2, 256), (2, 3, 256), (4, 5, 65537), (3, 4, 65537),
                                    (8, 9, 65537), (5, 5, 7)])
def test_host_constants_match_exact_integers(k, kl, t):
    P = O.Params(k, kl, 2048, t)
    expect = {
        "invQqi": P.invQqi, "Qqi": P.Qqi, "Qqibi": P.Qqibi, "Qbi": P.Qbi, "invBbi": P.invBbi,
        "Bbiqi": P.Bbiqi, "Bqi": P.Bqi, "lambda": P.lam, "wint": P.wint,
        "twint": P.twint, "twfrac": P.twfrac,
    }
    for name, want in expect.items():
        assert cp.hps_constant(k, kl, t, name) == [int(x) for x in want], name
    wf = cp.hps_constant(k, kl, t, "wfrac")
    assert wf == [int(x) for x in P.wfrac]


def test_host_rcp_rd_matches_oracle():
    got = cp.hps_constant(8, 9, 65537, "rcp_q")
    for bits, q in zip(got, O.tables()["primes"]):
        assert struct.unpack("<d", struct.pack("<Q", bits))[0] == O.rcp_rd(q)
