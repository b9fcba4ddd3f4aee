"""Server routes (server.py:30-110,133-169) with a stub search, on the CPU."""
from __future__ import annotations

import hashlib
import io
import json
import tarfile

import pytest

pytest.importorskip("fastapi")
pytest.importorskip("httpx")
from fastapi.testclient import TestClient  # noqa: E402

from sseqs_amd.server import ENV_A3M, create_app  # noqa: E402


class StubSearch:
    """Writes a fixed a3m per query and records every call."""

    def __init__(self):
        self.calls = []

    def __call__(self, queries, filenames, **opts):
        self.calls.append((list(queries), list(filenames), opts))
        for q, fn in zip(queries, filenames):
            with open(fn, "w") as fh:
                fh.write(f">original_query\n{q}\n>target_0 score=7\n{q[:2]}@{q[3:]}\n")


def sha(s: str) -> str:
    return hashlib.sha256(s.encode()).hexdigest()


def untar(blob: bytes) -> dict[str, bytes]:
    with tarfile.open(fileobj=io.BytesIO(blob), mode="r:gz") as tar:
        return {m.name: tar.extractfile(m).read() for m in tar.getmembers()}


@pytest.fixture
def stub_app(tmp_path, monkeypatch):
    for k in ("MSAS_NUM", "MSAS_FFT_RANK", "MSAS_TOP_FFT", "MSAS_TOP_SW", "MSAS_TOP_SW_AFFINE"):
        monkeypatch.delenv(k, raising=False)
    stub = StubSearch()
    app = create_app(msa_fn=stub, root=tmp_path, chunks=4)
    return TestClient(app), stub, tmp_path


def test_ticket_msa_and_download(stub_app):
    client, stub, root = stub_app
    q = ">101\nMKTAYIAK\n>102\nGGSLLV"
    r = client.post("/ticket/msa", data={"q": q})
    assert r.status_code == 200
    body = r.json()
    assert body == {"status": "COMPLETE", "id": sha("GGSLLV")}
    # one search for the distinct sequences, with the reference's options
    assert len(stub.calls) == 1
    queries, files, opts = stub.calls[0]
    assert queries == ["MKTAYIAK", "GGSLLV"]
    assert opts == dict(fft_rank=4, top_fft=200, top_sw=10, top_sw_affine=2, num_msas=4096,
                        bs=100_000_000)
    # written under a temporary name in the cache folder, then renamed
    assert f"cache_msa/4_4096_200_4_10_2/{sha('MKTAYIAK')}.a3m." in files[0]
    assert (root / "cache_msa/4_4096_200_4_10_2" / f"{sha('MKTAYIAK')}.a3m").exists()

    d = client.get(f"/result/download/{body['id']}")
    assert d.status_code == 200 and d.headers["content-type"] == "application/gzip"
    members = untar(d.content)
    assert set(members) == {"uniref.a3m", ENV_A3M}
    assert members[ENV_A3M] == b""
    parts = members["uniref.a3m"].split(b"\x00\n")
    # first line replaced by the request header, '@' replaced by 'X'
    assert parts[0].decode() == ">101\nMKTAYIAK\n>target_0 score=7\nMKXAYIAK\n"
    assert parts[1].decode() == ">102\nGGSLLV\n>target_0 score=7\nGGXLLV\n"


def test_ticket_cache_and_duplicates(stub_app):
    client, stub, root = stub_app
    client.post("/ticket/msa", data={"q": ">a\nMKTAYIAK"})
    assert len(stub.calls) == 1
    # cached sequence: no new search; duplicates are searched once
    r = client.post("/ticket/msa", data={"q": ">a\nMKTAYIAK\n>b\nPPPQQ\n>c\nPPPQQ"})
    assert len(stub.calls) == 2 and stub.calls[1][0] == ["PPPQQ"]
    members = untar(client.get(f"/result/download/{r.json()['id']}").content)
    parts = members["uniref.a3m"].split(b"\x00\n")
    assert len(parts) == 3
    assert parts[1].startswith(b">b\n") and parts[2].startswith(b">c\n")


def test_ticket_pair_name_and_options(stub_app, monkeypatch):
    client, stub, root = stub_app
    monkeypatch.setenv("MSAS_NUM", "100")
    monkeypatch.setenv("MSAS_TOP_FFT", "50")
    r = client.post("/ticket/pair", data={"q": ">x\nMKTAYIAK"})
    assert stub.calls[0][2]["num_msas"] == 100 and stub.calls[0][2]["top_fft"] == 50
    assert "4_100_50_4_10_2" in stub.calls[0][1][0]
    members = untar(client.get(f"/result/download/{r.json()['id']}").content)
    assert set(members) == {"pair.a3m", ENV_A3M}


def test_download_unknown_is_404(stub_app):
    client, _, _ = stub_app
    assert client.get("/result/download/nope").status_code == 404


def test_json_msa_route(stub_app):
    client, stub, root = stub_app
    r = client.post("/msa", json={"sequences": ["MKTAYIAK", "GGSLLV"], "msas_num": 64,
                                  "dataset": "uniref30_2302+bfd_mgy_colabfold"})
    assert r.status_code == 200
    assert r.headers["content-type"].startswith("application/x-ndjson")
    out = json.loads(r.text)
    assert out["type"] == "data_msa" and out["num_msas"] == 2
    assert out["cache_folder"] == "4_64_10_1_2_2"
    assert out["data"]["0"].startswith(">original_query\nMKTAYIAK\n")
    assert stub.calls[0][2] == dict(num_msas=64, bs=100_000_000, fft_rank=1, top_fft=10, top_sw=2,
                                    top_sw_affine=2)


def test_json_msa_reports_errors(tmp_path):
    def boom(*a, **k):
        raise RuntimeError("no database")

    client = TestClient(create_app(msa_fn=boom, root=tmp_path, chunks=4))
    r = client.post("/msa", json={"sequences": ["MK"], "dataset": "uniref30_2302+bfd_mgy_colabfold"})
    assert json.loads(r.text) == {"error": "MSA did not succeed: no database."}


def test_json_msa_rejects_unknown_dataset(stub_app):
    client, _, _ = stub_app
    r = client.post("/msa", json={"sequences": ["MK"], "dataset": "other"})
    assert r.status_code == 422


def test_ticket_requires_q(stub_app):
    client, stub, _ = stub_app
    assert client.post("/ticket/msa", data={"x": "1"}).status_code == 422
    assert not stub.calls


def test_ticket_status_poll(stub_app):
    client, _, _ = stub_app
    r = client.post("/ticket/msa", data={"q": ">a\nMKTAYIAK", "mode": "env"})
    tid = r.json()["id"]
    assert client.get(f"/ticket/{tid}").json() == {"status": "COMPLETE", "id": tid}
    assert client.get("/ticket/nope").json() == {"status": "UNKNOWN", "id": "nope"}


def test_concurrent_tickets_search_once(tmp_path, monkeypatch):
    """Two requests for the same uncached sequence at once: one search, both
    answered from the same complete file (no request sees a missing or
    partial a3m)."""
    import threading
    import time

    for k in ("MSAS_NUM", "MSAS_FFT_RANK", "MSAS_TOP_FFT", "MSAS_TOP_SW", "MSAS_TOP_SW_AFFINE"):
        monkeypatch.delenv(k, raising=False)
    calls = []

    def slow(queries, filenames, **opts):
        calls.append(list(queries))
        for q, fn in zip(queries, filenames):
            with open(fn, "w") as fh:
                fh.write(f">original_query\n{q}\n")
                fh.flush()
                time.sleep(0.3)  # a reader of a half-written file would see this line only
                fh.write(f">hit\n{q}\n")

    app = create_app(msa_fn=slow, root=tmp_path, chunks=4)
    out = [None, None]

    def post(i):
        with TestClient(app) as c:
            r = c.post("/ticket/msa", data={"q": f">r{i}\nMKTAYIAK"})
            out[i] = untar(c.get(f"/result/download/{r.json()['id']}").content)

    ts = [threading.Thread(target=post, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert calls == [["MKTAYIAK"]]
    # both tickets name the same tarball (its id is the sequence's hash, as
    # in server.py), written whole by each request
    for i in range(2):
        text = out[i]["uniref.a3m"].decode()
        assert text in (f">r{j}\nMKTAYIAK\n>hit\nMKTAYIAK\n" for j in range(2))


def test_environmental_database(tmp_path, monkeypatch):
    """With env_msa_kwargs the env member holds a second search's a3m (the
    reference's is always empty); both searches are cached."""
    for k in ("MSAS_NUM", "MSAS_FFT_RANK", "MSAS_TOP_FFT", "MSAS_TOP_SW", "MSAS_TOP_SW_AFFINE"):
        monkeypatch.delenv(k, raising=False)
    calls = []

    def stub(queries, filenames, db=None, **opts):
        calls.append(db)
        for q, fn in zip(queries, filenames):
            with open(fn, "w") as fh:
                fh.write(f">original_query\n{q}\n>{db}_hit\n{q}\n")

    client = TestClient(create_app(msa_fn=stub, root=tmp_path, chunks=4,
                                   msa_kwargs={"db": "UNIREF"}, env_msa_kwargs={"db": "ENV"}))
    r = client.post("/ticket/msa", data={"q": ">a\nMKTAYIAK"})
    members = untar(client.get(f"/result/download/{r.json()['id']}").content)
    assert members["uniref.a3m"].decode() == ">a\nMKTAYIAK\n>UNIREF_hit\nMKTAYIAK\n"
    assert members[ENV_A3M].decode() == ">a\nMKTAYIAK\n>ENV_hit\nMKTAYIAK\n"
    assert calls == ["UNIREF", "ENV"]
    client.post("/ticket/msa", data={"q": ">b\nMKTAYIAK"})
    assert calls == ["UNIREF", "ENV"]  # both cached
