"""The asynchronous result hand-off on the GPU: queries submitted to the serving
loop under a token (evaluate_query's asynchronous form, CalciteInterpreter.cpp:
920-960) are parked in the result repository when they have run (update_token,
ResultSetRepository.cpp:45-112) and read back with get_result (:186-205). Every
parked frame is compared with the oracle; freeing the last reference returns the
query's HBM budget to the server."""
import numpy as np
import pytest

from oracle import planner as P
from oracle import pyoracle as O
from tests.helpers import assert_column_equal, c2_mix_table, to_device

pytestmark = pytest.mark.gpu

MIX = [P.GDF_INT32, P.GDF_INT64, P.GDF_FLOAT32, P.GDF_FLOAT64]
QUERIES = [
    ("LogicalFilter(condition=[AND(>($0, -200000), <($3, 1.2))])",
     "LogicalProject(A=[+($0, $1)], B=[*($3, 2.5)], C=[$2])"),
    (None, "LogicalProject(X=[-($1, MOD($0, 13))], Y=[>($2, 0.25)])"),
    ("LogicalFilter(condition=[<($2, 0.5)])", None),
    ("LogicalFilter(condition=[>=($1, 0)])", "LogicalProject(S=[CASE(IS NOT NULL($3), $3, 0.0)])"),
]


@pytest.fixture(scope="module")
def B():
    import blz_ral
    if blz_ral.device_count() < 1:
        pytest.fail("no HIP device visible")
    return blz_ral


def _oracle(host, filt, proj):
    mid = O.process_filter(host, filt) if filt else host
    return O.execute_project(mid, proj) if proj else mid


def _names(proj, n):
    if proj is None:
        return [f"${i}" for i in range(n)]
    return [e.split("=[")[0].strip() for e in proj[len("LogicalProject("):-1].split("], ")]


def test_served_results_parked_under_tokens(B):
    RS = B.ResultSetRepository
    rng = np.random.default_rng(46_000)
    n = 150_001
    host = c2_mix_table(rng, MIX, [0.05, 0.0, 0.1, 0.05], n)
    frame = [to_device(B, h) for h in host]
    srv = B.Server(4 << 30, max_batch=8, n_streams=2, r=4)
    conn = RS.init_session()
    tokens = [7_000 + i for i in range(len(QUERIES))]
    print(f"CASE result_sets queries={len(QUERIES)} rows={n}", flush=True)
    for tok, (filt, proj) in zip(tokens, QUERIES):
        RS.register_query(conn, tok)
        srv.submit_token(frame, filt, proj, tok)
    # read in reverse order: each get_result blocks until its token is parked
    for tok, (filt, proj) in reversed(list(zip(tokens, QUERIES))):
        rs = RS.get_result(conn, tok)
        ref = _oracle(host, filt, proj)
        assert rs.errorMsg == "", rs.errorMsg
        assert rs.duration > 0
        assert rs.names == _names(proj, len(ref))
        assert len(rs.columns) == len(ref)
        for c, r in zip(rs.columns, ref):
            assert_column_equal(c, r, what=f"token {tok}: {filt} {proj}")
        for t, c in zip(rs.column_tokens, rs.columns):
            g = RS.get_column(conn, t)
            assert g.col.size == c.col.size and g.data().tobytes() == c.data().tobytes()
    assert srv.stats()["used_bytes"] > 0
    # a second reader of the first token: the first free keeps the frame
    RS.get_result(conn, tokens[0])
    assert RS.try_free_result(conn, tokens[0])
    assert RS.get_result(conn, tokens[0]).columns[0].col.size > 0  # a third reference
    RS.try_free_result(conn, tokens[0])
    for tok in tokens:
        assert RS.try_free_result(conn, tok)
    assert srv.stats()["used_bytes"] == 0, srv.stats()
    RS.remove_all_connection_tokens(conn)
    srv.close()


def test_query_errors_are_parked(B):
    """A query that cannot be planned parks its error text under its token; the
    call that submits it succeeds, as evaluate_query does."""
    RS = B.ResultSetRepository
    col = B.DeviceColumn.from_numpy(np.arange(1000, dtype=np.int64))
    srv = B.Server(1 << 30, max_batch=4, n_streams=1, r=4)
    conn = RS.init_session()
    RS.register_query(conn, 8_001)
    print("CASE result_sets error", flush=True)
    srv.submit_token([col], None, "LogicalProject(E=[+($7, 1)])", 8_001)
    rs = RS.get_result(conn, 8_001)
    assert rs.errorMsg != "" and rs.columns == []
    assert RS.try_free_result(conn, 8_001)
    # an unregistered token is refused at submission
    with pytest.raises(B.BlzError, match="Token does not exist"):
        srv.submit_token([col], None, "LogicalProject(E=[+($0, 1)])", 8_002)
    RS.remove_all_connection_tokens(conn)
    srv.close()
