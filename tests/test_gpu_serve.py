"""The continuous-batching engine of the server: a request's a3m is the same
alone and in a mix of concurrent requests, and equals msa()'s whole-database
cascade text for that query; the ticket route on the engine."""
import io
import tarfile
import threading

import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

LENS = [60, 150, 230, 310, 480, 700, 950]


@pytest.fixture(scope="module")
def sworld(dev):
    from sseqs_amd import synth
    from sseqs_amd.db import SeqDB
    queries = [synth.make_query(L, 70 + i, seed=12) for i, L in enumerate(LENS)]
    db = SeqDB.synthetic(20000, queries, seed=23, device=dev)
    return dict(db=db, queries=queries)


def test_engine_alone_equals_mix(sworld, tmp_path):
    import sseqs_amd
    from sseqs_amd.serve import SearchEngine
    db, queries = sworld["db"], sworld["queries"]
    announce("serving engine: each query alone, then all at once")
    with SearchEngine(db, top_prefilter=2000, top_hits=200, max_batch=4, max_wait_s=0.02) as eng:
        alone = [eng.search(q, timeout=120) for q in queries]
        assert eng.stats.batch_sizes == [1] * len(queries)
        # the mix: every query twice, submitted together from several threads
        futs = [None] * (2 * len(queries))

        def sub(i):
            futs[i] = eng.submit(queries[i % len(queries)])
        ts = [threading.Thread(target=sub, args=(i,)) for i in range(len(futs))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        mixed = [f.result(timeout=120) for f in futs]
        assert max(eng.stats.batch_sizes) > 1  # requests shared sweeps
        assert max(eng.stats.batch_sizes) <= 4
    for i, text in enumerate(mixed):
        assert text == alone[i % len(queries)], i
    # and the drop-in call's whole-database cascade gives the same text
    for i in (0, 3, 6):
        fn = tmp_path / f"q{i}.a3m"
        sseqs_amd.msa(queries[i], str(fn), num_msas=200, verbose=False, db=db,
                      cascade="global", top_prefilter=2000)
        assert fn.read_text() == alone[i], i
    assert alone[0].count(">") > 20


def test_engine_hbm_budget_splits_sweeps(sworld):
    """A budget that no two consecutive queries' buffers (plus their sweep's
    own) fit admits one query per sweep (the first query of a sweep is
    always admitted)."""
    from sseqs_amd.serve import SearchEngine, query_bytes, sweep_bytes
    db, queries = sworld["db"], sworld["queries"]
    announce("serving engine: HBM budget admission")
    ls = [len(q) for q in queries[:5]]
    pair = [sweep_bytes(db, max(a, b), 200) + query_bytes(db, a, 2000, 200) +
            query_bytes(db, b, 2000, 200) for a, b in zip(ls, ls[1:])]
    budget = min(pair) - 1
    with SearchEngine(db, top_prefilter=2000, top_hits=200, max_batch=8,
                      hbm_budget=budget) as eng:
        futs = [eng.submit(q) for q in queries[:5]]
        texts = [f.result(timeout=120) for f in futs]
        assert max(eng.stats.batch_sizes) == 1
    with SearchEngine(db, top_prefilter=2000, top_hits=200, max_batch=8) as eng:
        assert [eng.search(q, timeout=120) for q in queries[:5]] == texts


def test_ticket_route_on_engine(sworld, tmp_path):
    pytest.importorskip("fastapi")
    from fastapi.testclient import TestClient

    from sseqs_amd.serve import SearchEngine
    from sseqs_amd.server import create_app
    db, queries = sworld["db"], sworld["queries"]
    announce("ticket route on the serving engine")
    with SearchEngine(db, top_prefilter=2000, top_hits=200) as eng:
        want = eng.search(queries[1], timeout=120)
        app = create_app(root=tmp_path, chunks=4, engine=eng)
        out = {}

        def post(i):
            with TestClient(app) as c:
                r = c.post("/ticket/msa", data={"q": f">r{i}\n{queries[1]}\n>s{i}\n{queries[2]}"})
                blob = c.get(f"/result/download/{r.json()['id']}").content
                with tarfile.open(fileobj=io.BytesIO(blob), mode="r:gz") as tar:
                    out[i] = tar.extractfile("uniref.a3m").read().decode()
        ts = [threading.Thread(target=post, args=(i,)) for i in range(3)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
    rows = want.replace("@", "X").split("\n")
    for i, text in out.items():
        first = text.split("\x00\n")[0].split("\n")
        assert first[1:] == rows[1:]
