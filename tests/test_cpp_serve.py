"""The served form from a C++ host (tests/cpp/serve_client.cpp over the library's request loop,
ahead_serve_*): random queries of the four serve shapes, all submitted at once to one and to four
slots, each result equal to the oracle's fused form of the query (rows, results, groups, sums,
keys, every error vector), with flips in every fact column and attribute."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_bat as OB
from ahead_amd import serve, ssb
from test_cpp_ssb import SMALL, _pred, read_result

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "cpp" / "_build" / "serve_client"
KIND = {"key": "int", "meas": "int", "tiny": "tiny", "short": "short"}
KEYS = {"d_datekey", "c_custkey", "s_suppkey", "p_partkey"}


def write_inputs(spec, d: Path, queries: dict) -> None:
    host = ssb.host_tables(spec)
    lines = []
    for name, (codes, meta, _) in host.items():
        codes.tofile(d / f"{name}.bin")
        lines.append(f"{name} {KIND[ssb.COLUMNS[name][1]]} {meta.a} {codes.size}")
    (d / "tables.txt").write_text("\n".join(lines) + "\n")
    q = []
    for name, qq in queries.items():
        q.append(f"query {name} {qq.family}")
        q.append(f"fused {qq.kind} {qq.pre_mask} {' '.join(qq.meas)}")
        q.append(" ".join(["groups"] + [str(g) for g in qq.groups]))
        for col, p in qq.preds:
            q.append(f"pred {col} {_pred(p)}")
        for dim in qq.dims:
            q.append(f"dim {dim.name} {dim.attr or '-'} {dim.flags}")
            for col, p in dim.sels:
                q.append(f"sel {col} {_pred(p)}")
    (d / "queries.txt").write_text("\n".join(q) + "\n")


def test_cpp_serve_client_links_without_gpu():
    if not EXE.exists():
        pytest.skip("tests/cpp not built (run __graft_entry__.build())")
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the gpu test runs the program")
    res = subprocess.run([str(EXE), "/nonexistent", "1"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 2 and "no GPU" in res.stdout, (res.returncode, res.stdout, res.stderr)


@pytest.mark.gpu
@pytest.mark.parametrize("slots", [1, 4])
def test_cpp_serve_client_matches_the_oracle(tmp_path, slots):
    if not EXE.exists():
        pytest.skip("tests/cpp not built (run __graft_entry__.build())")
    spec = ssb.SsbSpec(rate=1e-3, seed=91, flip_cols=tuple(c for c in ssb.COLUMNS if c not in KEYS), **SMALL)
    rq = serve.random_queries(40, 0, 92)
    queries = {f"q{r.qid}": r.spec() for r in rq}
    write_inputs(spec, tmp_path, queries)
    print(f"C++ serve client, {len(queries)} queries, {slots} slot(s)", flush=True)
    res = subprocess.run([str(EXE), str(tmp_path), str(slots)], capture_output=True, text=True, timeout=300)
    print(res.stdout[-2000:])
    assert res.returncode == 0, res.stdout + res.stderr
    host = OB.tables(spec)
    flipped = 0
    for name, q in queries.items():
        got = read_result(tmp_path / f"serve_{name}.txt")
        ref = ssb.fused(host, q, OB, fallback=False)
        assert ref["status"] == 0 and got["status"] == 0, name
        assert got["rows"] == ref["rows"], name
        if "result" in ref:
            assert got["result"] == ref["result"], name
        else:
            assert got["groups"] == ref["groups"], name
            assert np.array_equal(got["sums"], ref["sums"]), name
            assert len(got["keys"]) == len(ref["keys"])
            for a, b in zip(got["keys"], ref["keys"]):
                assert np.array_equal(a, b), name
        re = dict(ref["errors"].items)
        assert sorted(got["errors"]) == sorted(re), (name, sorted(got["errors"]), sorted(re))
        for k, v in re.items():
            assert np.array_equal(got["errors"][k], v), (name, k)
        flipped += len(re)
    assert flipped > 0
