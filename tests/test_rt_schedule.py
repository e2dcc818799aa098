"""The learned predictor's attempt schedule (RTPropagator._attempt_due), host logic only."""
from types import SimpleNamespace

from salmon2_amd.rt import RTPropagator


def _schedule(backoff, verdicts, nt):
    """Steps that attempt, given the accept (True) / reject (False) verdict of each attempt."""
    st = SimpleNamespace(backoff=backoff, fails=0, last_attempt=0)
    out = []
    it = iter(verdicts)
    for itt in range(4, nt + 1):  # attempts start once four potentials are known
        if RTPropagator._attempt_due(st, itt):
            out.append(itt)
            st.last_attempt = itt
            st.fails = 0 if next(it) else st.fails + 1
    return out


def test_no_backoff_attempts_every_step():
    assert _schedule(None, [False] * 20, 12) == list(range(4, 13))


def test_backoff_after_rejections():
    # 2 rejections in a row, then one attempt every 3 steps
    assert _schedule((2, 3), [False] * 20, 16) == [4, 5, 8, 11, 14]


def test_acceptance_resumes_every_step():
    # rejections at 4, 5 -> back-off; accepted at 8 -> every step again; rejections at 9, 10 -> back-off
    verdicts = [False, False, True, False, False, False, False]
    assert _schedule((2, 3), verdicts, 16) == [4, 5, 8, 9, 10, 13, 16]
