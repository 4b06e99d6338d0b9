"""Host-side checks of the serving loop's guards (no device work)."""
import pytest

from cupoly_amd import serving


class _StubShape:
    """A ring shape with a footprint and no contexts: Server reserves nothing for it."""

    def __init__(self, footprint):
        self.footprint = footprint
        self.ctxs = []

    def __repr__(self):
        return "stub"


def test_budget_below_one_request_is_rejected():
    # admission could never take such a request, so run() would wait forever
    with pytest.raises(ValueError, match="below the footprint"):
        serving.Server([_StubShape(1 << 20)], hbm_budget=(1 << 20) - 1)


def test_batch_cap_follows_the_budget():
    sh = _StubShape(1000)
    srv = serving.Server([sh], hbm_budget=3500, max_batch={sh: 64})
    assert srv.max_batch[sh] == 3


def test_empty_request_list_gives_an_empty_report():
    rep = serving.Server([_StubShape(1000)], hbm_budget=4000).run([])
    assert rep["requests"] == 0 and rep["batches"] == 0 and rep["p99_ms"] is None
