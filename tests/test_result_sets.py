"""The result repository's bookkeeping (result_set_repository,
src/ResultSetRepository.cpp), without a device: frames of zero-row columns with
no buffers, so no free reaches HIP.

Each case follows the reference's method: register_query (:36-44) lists a token
under its connection, not ready; update_token (:45-112) parks the frame, gives
every column a column token and wakes readers; get_result (:186-205) blocks until
then and counts a reference; try_free_result (:166-184) drops one reference and
frees at the last; remove_all_connection_tokens (:152-164) frees what no reader
holds; get_column (:207-226) finds a parked column by its token. Error texts are
the reference's."""
import threading
import time

import pytest

from oracle import planner as P


@pytest.fixture(scope="module")
def RS():
    import blz_ral
    return blz_ral.ResultSetRepository


def _cols(B, dtypes):
    out = []
    for dt in dtypes:
        g = B.GdfColumn()
        g.data = None
        g.valid = None
        g.size = 0
        g.dtype = dt
        g.null_count = 0
        out.append(B.DeviceColumn.adopt(g, owned=False))
    return out


def _err(fn, *args):
    import blz_ral
    with pytest.raises(blz_ral.BlzError) as e:
        fn(*args)
    return str(e.value)


def test_sessions_are_distinct(RS):
    ids = {RS.init_session() for _ in range(64)}
    assert len(ids) == 64


def test_update_then_get_result(RS):
    import blz_ral as B
    conn = RS.init_session()
    tok = 1001
    RS.register_query(conn, tok)
    RS.update_token(tok, _cols(B, [P.GDF_INT64, P.GDF_FLOAT64]), duration=12.5, names=["E0", "E1"])
    rs = RS.get_result(conn, tok)
    assert rs.errorMsg == "" and rs.duration == 12.5 and rs.names == ["E0", "E1"]
    assert [c.col.dtype for c in rs.columns] == [P.GDF_INT64, P.GDF_FLOAT64]
    assert len(set(rs.column_tokens)) == 2 and all(t != 0 for t in rs.column_tokens)
    # get_column finds each parked column by its column token
    for t, c in zip(rs.column_tokens, rs.columns):
        assert RS.get_column(conn, t).col.dtype == c.col.dtype
    assert RS.try_free_result(conn, tok)
    assert not RS.try_free_result(conn, tok)


def test_get_result_blocks_until_update(RS):
    import blz_ral as B
    conn = RS.init_session()
    tok = 1002
    RS.register_query(conn, tok)
    t_update = []

    def producer():
        time.sleep(0.3)
        t_update.append(time.monotonic())
        RS.update_token(tok, _cols(B, [P.GDF_INT32]), duration=1.0)

    th = threading.Thread(target=producer)
    th.start()
    rs = RS.get_result(conn, tok)
    t_back = time.monotonic()
    th.join()
    assert t_update and t_back >= t_update[0]
    assert len(rs.columns) == 1
    RS.try_free_result(conn, tok)


def test_error_texts(RS):
    conn = RS.init_session()
    assert "Connection does not exist" in _err(RS.get_result, conn ^ 0x5A5A, 1)
    assert "Result set does not exist" in _err(RS.get_result, conn, 987654321)
    assert "Token does not exist" in _err(RS.update_token, 987654321, [])
    assert "Column does not exist" in _err(RS.get_column, conn, 123)
    assert "Connection does not exist" in _err(RS.get_column, conn ^ 0x5A5A, 123)
    assert "Connection does not exist" in _err(RS.try_free_result, conn ^ 0x5A5A, 1)
    assert RS.try_free_result(conn, 987654321) is False


def test_error_result_is_parked(RS):
    conn = RS.init_session()
    RS.register_query(conn, 1003)
    RS.update_token(1003, [], error_msg="evaluate_split_query failed")
    rs = RS.get_result(conn, 1003)
    assert rs.errorMsg == "evaluate_split_query failed" and rs.columns == []
    RS.try_free_result(conn, 1003)


def test_references_counted(RS):
    """Two readers: the first free drops a reference, the second frees."""
    import blz_ral as B
    conn = RS.init_session()
    RS.register_query(conn, 1004)
    RS.update_token(1004, _cols(B, [P.GDF_INT8]))
    a = RS.get_result(conn, 1004)
    RS.get_result(conn, 1004)
    assert RS.try_free_result(conn, 1004)
    assert RS.get_column(conn, a.column_tokens[0]).col.dtype == P.GDF_INT8  # still parked
    assert RS.try_free_result(conn, 1004)
    assert "Result set does not exist" in _err(RS.get_result, conn, 1004)
    assert "Column does not exist" in _err(RS.get_column, conn, a.column_tokens[0])


def test_result_never_read_is_freed_at_once(RS):
    import blz_ral as B
    conn = RS.init_session()
    RS.register_query(conn, 1005)
    RS.update_token(1005, _cols(B, [P.GDF_INT16]))
    assert RS.try_free_result(conn, 1005)
    assert "Result set does not exist" in _err(RS.get_result, conn, 1005)


def test_remove_all_connection_tokens(RS):
    """Results no reader holds are freed and the connection is forgotten; a held
    result stays until its reader frees it."""
    import blz_ral as B
    conn = RS.init_session()
    for tok in (1006, 1007):
        RS.register_query(conn, tok)
        RS.update_token(tok, _cols(B, [P.GDF_INT32]))
    held = RS.get_result(conn, 1007)
    RS.remove_all_connection_tokens(conn)
    assert "Connection does not exist" in _err(RS.get_result, conn, 1006)
    conn2 = RS.init_session()
    assert "Result set does not exist" in _err(RS.get_result, conn2, 1006)
    # 1007 is still parked (a reader holds it): reachable from any connection
    assert RS.get_column(conn2, held.column_tokens[0]).col.dtype == P.GDF_INT32
    RS.register_query(conn2, 1007)  # re-registering drops the old frame
    assert "Column does not exist" in _err(RS.get_column, conn2, held.column_tokens[0])
    RS.remove_all_connection_tokens(conn2)


def test_reader_of_a_freed_token_is_woken(RS):
    """A reader blocked on a token that is freed before it is parked returns with
    "Result set does not exist" rather than waiting for ever."""
    conn = RS.init_session()
    RS.register_query(conn, 1008)
    out = []

    def reader():
        import blz_ral
        try:
            RS.get_result(conn, 1008)
            out.append("ready")
        except blz_ral.BlzError as e:
            out.append(str(e))

    th = threading.Thread(target=reader)
    th.start()
    time.sleep(0.2)
    RS.remove_all_connection_tokens(conn)
    th.join(timeout=10)
    assert not th.is_alive()
    assert out and "Result set does not exist" in out[0]
