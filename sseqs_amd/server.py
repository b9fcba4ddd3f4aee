"""MSA server: the ColabFold/boltz ticket API and the JSON ``/msa`` route
(mirror of server.py).

Routes, as in server.py:30-110,133-169:

* ``POST /ticket/msa`` and ``POST /ticket/pair``. The form field ``q`` holds
  FASTA: header and sequence lines alternate. Sequences not yet in the cache
  (``cache_msa/<CHUNKS>_<num>_<top_fft>_<rank>_<top_sw>_<top_sw_affine>/
  <sha256>.a3m``) are searched with one ``msa()`` call. Each a3m gets '@'
  replaced by 'X' and its first line replaced by the request's header; the
  a3m files are joined with ``b"\\x00\\n"`` and packed as ``uniref.a3m``
  (``pair.a3m`` on the pair route) plus
  ``bfd.mgnify30.metaeuk30.smag30.a3m`` (empty, as in server.py, unless an
  environmental database is configured: ``ENVDB`` / ``env_msa_kwargs``)
  into a tar.gz. The route answers
  ``{"status": "COMPLETE", "id": <tar id>}``.
* ``GET /ticket/{tar_id}``: ``{"status": "COMPLETE" | "UNKNOWN", "id": ...}``,
  the status poll of the ColabFold client API (absent from server.py).
* ``GET /result/download/{tar_id}``: that tar.gz, or 404.
* ``POST /msa`` (JSON :class:`MSARequest`): one ndjson line,
  ``{"type": "data_msa", "data": {i: a3m}, "num_msas": n, "cache_folder": f}``,
  or ``{"error": ...}``.

Differences from server.py:

* The app is built by :func:`create_app`; ``main()`` serves it. The
  reference starts uvicorn at import.
* Searches run in a worker thread, one at a time, so the event loop keeps
  serving downloads meanwhile. The reference called ``msa()`` inside the
  event loop. Each a3m is written under a temporary name and renamed into
  the cache, and the cache is checked again under the search lock.
* ``/msa`` passes the file names as ``filenames``, the name ``msa()``
  takes. The reference passed ``filenames_a3m=`` (server.py:147), so its
  ``/msa`` always answered with the error line.
* The tar id is the hash of the last distinct sequence in request order.
  The reference took the last one in Python ``set`` order.
"""

import argparse
import hashlib
import io
import json
import os
import tarfile
import threading
import uuid
from pathlib import Path
from typing import Callable, List, Literal

ENV_A3M = "bfd.mgnify30.metaeuk30.smag30.a3m"


def _sha(s: str) -> str:
    return hashlib.sha256(s.encode("utf-8")).hexdigest()


def create_app(msa_fn: Callable | None = None, root: str | os.PathLike = ".",
               chunks: int | None = None, msa_kwargs: dict | None = None, engine=None,
               env_msa_kwargs: dict | None = None):
    """FastAPI app serving the reference routes.

    msa_fn: the search, called as ``msa_fn(queries, filenames, **options)``;
    default :func:`sseqs_amd.msa.msa`. root: directory holding ``cache_msa/``
    and the tarballs. chunks: the database size tag of the cache folders
    (default ``$CHUNKS`` or 4, server.py:25). msa_kwargs: extra keyword
    arguments for every search (for example ``db=``).

    engine: a :class:`sseqs_amd.serve.SearchEngine`.  Searches then go to
    its continuous-batching device loop instead of ``msa_fn``: concurrent
    requests share sweeps, no lock serialises them, and a sequence that
    several requests wait for is searched once.  The engine's whole-database
    cascade (its top_prefilter and top_hits) replaces the per-batch
    cascade options of the request, and its cache folders are named
    ``engine_<top_prefilter>_<top_hits>`` so the two never mix.

    env_msa_kwargs: keyword arguments (for example ``db=`` an environmental
    database) of a second search of every ticket's sequences, whose a3m
    texts fill the ``bfd.mgnify30.metaeuk30.smag30.a3m`` member the same
    way ``uniref.a3m`` is filled (server.py writes it empty, which is what
    happens without this argument).
    """
    from fastapi import FastAPI, HTTPException, Request
    from fastapi.concurrency import run_in_threadpool
    from fastapi.middleware.cors import CORSMiddleware
    from fastapi.responses import FileResponse, Response, StreamingResponse
    from pydantic import BaseModel, Field

    if msa_fn is None:
        from .msa import msa as msa_fn
    root = Path(root)
    if chunks is None:
        from .db import chunks_tag
        chunks = chunks_tag()
    extra = dict(msa_kwargs or {})
    lock = threading.Lock()
    result_store: dict[str, Path] = {}

    app = FastAPI()
    app.add_middleware(CORSMiddleware, allow_origins=["*"], allow_credentials=True,
                       allow_methods=["*"], allow_headers=["*"])
    app.state.result_store = result_store

    inflight: dict[str, object] = {}
    inflight_lock = threading.Lock()

    def write_atomic(path: str, text: str) -> None:
        tmp = f"{path}.{uuid.uuid4().hex}.tmp"
        with open(tmp, "w") as fh:
            fh.write(text)
        os.replace(tmp, path)

    def engine_search(queries: list[str], filenames: list[str]) -> None:
        waits = []
        with inflight_lock:
            for q, f in dict(zip(filenames, queries)).items():
                if os.path.exists(f):
                    continue
                fut = inflight.get(f)
                if fut is None:
                    fut = engine.submit(q)
                    inflight[f] = fut
                waits.append((f, fut))
        try:
            for f, fut in waits:
                write_atomic(f, fut.result())
        finally:
            with inflight_lock:
                for f, fut in waits:
                    if inflight.get(f) is fut and fut.done():
                        del inflight[f]

    def search(queries: list[str], filenames: list[str], kwargs: dict | None = None,
               **opts) -> None:
        if engine is not None and kwargs is None:
            engine_search(queries, filenames)
            return
        kw = extra if kwargs is None else kwargs
        """Search the queries whose a3m is not yet in the cache.  The check
        is repeated under the lock, so a sequence that two requests both
        found missing is searched once; each a3m is written to a file of its
        own and renamed into place, so a reader never sees a partial file
        and a search never deletes a cached file another request reads."""
        with lock:
            pending = dict(zip(filenames, queries))  # one search per file
            todo = [(q, f) for f, q in pending.items() if not os.path.exists(f)]
            if not todo:
                return
            tag = uuid.uuid4().hex
            tmps = [f"{f}.{tag}.tmp" for _, f in todo]
            try:
                msa_fn([q for q, _ in todo], tmps, **opts, **kw)
                for t, (_, f) in zip(tmps, todo):
                    os.replace(t, f)
            finally:
                for t in tmps:
                    if os.path.exists(t):
                        os.remove(t)

    async def form_q(request: Request) -> str:
        """The form field q.  URL-encoded forms (what ColabFold/boltz clients
        send) are parsed here; multipart needs python-multipart."""
        ctype = request.headers.get("content-type", "")
        if ctype.startswith("multipart/"):
            fields = await request.form()
            q = fields.get("q")
        else:
            from urllib.parse import parse_qs
            q = parse_qs((await request.body()).decode("utf-8"), keep_blank_values=True).get("q")
            q = q[0] if q else None
        if q is None:
            raise HTTPException(status_code=422, detail="form field 'q' is required")
        return str(q)

    @app.post("/ticket/msa", include_in_schema=False)
    @app.post("/ticket/pair", include_in_schema=False)
    async def ticket_msa(request: Request):
        # the ColabFold client also sends `mode` (env / all / pair...); as in
        # server.py every mode searches the one database
        lines = (await form_q(request)).split("\n")
        num = int(os.getenv("MSAS_NUM", 4096))
        rank = int(os.getenv("MSAS_FFT_RANK", 4))
        top_fft = int(os.getenv("MSAS_TOP_FFT", 200))
        top_sw = int(os.getenv("MSAS_TOP_SW", 10))
        top_sw_affine = int(os.getenv("MSAS_TOP_SW_AFFINE", 2))
        folder = root / "cache_msa" / (
            f"{chunks}_{num}_{top_fft}_{rank}_{top_sw}_{top_sw_affine}" if engine is None else
            f"{chunks}_engine_{engine.kc}_{engine.kh}")
        folder.mkdir(parents=True, exist_ok=True)

        heads, proteins = lines[::2], lines[1::2]
        out_paths = [folder / f"{_sha(p)}.a3m" for p in proteins]
        distinct = list(dict.fromkeys(proteins))
        todo = [p for p in distinct if not (folder / f"{_sha(p)}.a3m").exists()]
        if todo:
            await run_in_threadpool(search, todo, [str(folder / f"{_sha(p)}.a3m") for p in todo],
                                    fft_rank=rank, top_fft=top_fft, top_sw=top_sw,
                                    top_sw_affine=top_sw_affine, num_msas=num, bs=100_000_000)

        def joined(paths: list[Path]) -> bytes:
            parts = []
            for i, path in enumerate(paths):
                text = path.read_text().replace("@", "X")
                rows = text.split("\n")
                rows[0] = heads[i]
                parts.append("\n".join(rows).encode())
            return b"\x00\n".join(parts)

        content = joined(out_paths)
        env_content = b""
        if env_msa_kwargs is not None:
            env_folder = root / "cache_msa" / (
                f"env_{chunks}_{num}_{top_fft}_{rank}_{top_sw}_{top_sw_affine}")
            env_folder.mkdir(parents=True, exist_ok=True)
            env_paths = [env_folder / f"{_sha(p)}.a3m" for p in proteins]
            env_todo = [p for p in distinct if not (env_folder / f"{_sha(p)}.a3m").exists()]
            if env_todo:
                await run_in_threadpool(search, env_todo,
                                        [str(env_folder / f"{_sha(p)}.a3m") for p in env_todo],
                                        dict(env_msa_kwargs), fft_rank=rank, top_fft=top_fft,
                                        top_sw=top_sw, top_sw_affine=top_sw_affine, num_msas=num,
                                        bs=100_000_000)
            env_content = joined(env_paths)

        tar_id = _sha(distinct[-1]) if distinct else _sha("")
        tar_dir = root / f"{tar_id}_tar"
        tar_dir.mkdir(parents=True, exist_ok=True)
        tar_file = tar_dir / f"{tar_id}.tar.gz"
        name = "pair.a3m" if request.url.path.endswith("/ticket/pair") else "uniref.a3m"
        tmp = tar_dir / f"{tar_id}.{uuid.uuid4().hex}.tmp"
        with tarfile.open(tmp, "w:gz") as tar:
            ti = tarfile.TarInfo(name=name)
            ti.size = len(content)
            tar.addfile(ti, io.BytesIO(content))
            ti2 = tarfile.TarInfo(name=ENV_A3M)
            ti2.size = len(env_content)
            tar.addfile(ti2, io.BytesIO(env_content))
        os.replace(tmp, tar_file)
        result_store[tar_id] = tar_file
        return {"status": "COMPLETE", "id": tar_id}

    @app.get("/ticket/{tar_id}", include_in_schema=False)
    async def ticket_status(tar_id: str):
        """Status poll of the ColabFold client API.  Searches finish before
        the POST returns, so a known id is COMPLETE."""
        path = result_store.get(tar_id)
        status = "COMPLETE" if path is not None and path.exists() else "UNKNOWN"
        return {"status": status, "id": tar_id}

    @app.get("/result/download/{tar_id}", include_in_schema=False)
    async def result_download(tar_id: str):
        path = result_store.get(tar_id)
        if path is None or not path.exists():
            return Response(status_code=404)
        return FileResponse(path, media_type="application/gzip", filename="out.tar.gz")

    class MSARequest(BaseModel):  # server.py:113-124
        sequences: List[str] = Field(["ADAM", "ADA"])
        msas_num: int = Field(4096)
        fft_rank: int = Field(1)
        top_fft: int = Field(10)
        top_sw: int = Field(2)
        top_sw_affine: int = Field(2)
        dataset: Literal["uniref30_2302+bfd_mgy_colabfold"] = Field(...)

    @app.post("/msa")
    def get_msa(req: MSARequest):
        folder = root / (f"{chunks}_{req.msas_num}_{req.top_fft}_{req.fft_rank}_{req.top_sw}_"
                         f"{req.top_sw_affine}" if engine is None else
                         f"{chunks}_engine_{engine.kc}_{engine.kh}")
        folder.mkdir(parents=True, exist_ok=True)
        files = [str(folder / f"{_sha(p)}.a3m") for p in req.sequences]

        def event_stream():
            try:
                search(list(req.sequences), files, num_msas=req.msas_num, bs=100_000_000,
                       fft_rank=req.fft_rank, top_fft=req.top_fft, top_sw=req.top_sw,
                       top_sw_affine=req.top_sw_affine)
                data = {i: Path(f).read_text() for i, f in enumerate(files)}
                yield json.dumps({"type": "data_msa", "data": data,
                                  "num_msas": len(req.sequences), "cache_folder": folder.name})
            except Exception as e:  # reported to the client, as server.py:160-161
                yield json.dumps({"error": f"MSA did not succeed: {e}."})

        return StreamingResponse(event_stream(), media_type="application/x-ndjson")

    return app


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description="Run MSA Server")
    ap.add_argument("-port", "-p", default=8000, type=int, help="Port for server. ")
    args, _ = ap.parse_known_args(argv)
    import uvicorn
    env_kwargs = None
    if os.environ.get("ENVDB"):  # an environmental database (.xbit), searched like DBPATH
        from .db import SeqDB, env_chunks
        env_kwargs = {"db": SeqDB.from_xbit(os.environ["ENVDB"], env_chunks())}
    uvicorn.run(create_app(env_msa_kwargs=env_kwargs), host="0.0.0.0", port=args.port,
                reload=False, access_log=False, workers=1)


if __name__ == "__main__":
    main()
