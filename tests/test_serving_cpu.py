"""Host-side checks of the serving loop's guards (no device work)."""
import pytest

from cupoly_amd import serving


class _StubShape:
    """A ring shape with a footprint and `lanes` fake lane contexts whose
    workspace the Server charges (it reserves nothing on a device: the
    reserve call is replaced)."""

    def __init__(self, footprint, lanes=0, ws=0):
        self.footprint = footprint
        self.ctxs = [None] * lanes
        self.ws = ws

    def workspace_bytes(self, batch):
        return self.ws * batch

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


def test_reserved_workspaces_are_charged_first(monkeypatch):
    monkeypatch.setattr(serving.cp, "hom_mul_reserve", lambda ctx, batch: None)
    sh = _StubShape(1000, lanes=2, ws=100)
    # 2 lanes x 4 x 100 reserved = 800; 10000 - 800 leaves 9 outputs
    srv = serving.Server([sh], hbm_budget=10000, max_batch={sh: 4})
    assert srv.reserved == 800
    assert srv.max_batch[sh] == 4
    with pytest.raises(ValueError, match="reserved lane workspaces"):
        serving.Server([sh], hbm_budget=1500, max_batch={sh: 4})
