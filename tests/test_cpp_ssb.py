"""The thirteen SSB queries through the C++ operator layer (tests/cpp/ssb_chains.cpp, the chains
of ahead_amd.ssb written against include/ahead/column_operators/OperatorsAN.hpp as the
reference's drivers call the templates) against the oracle's chains: rows, results, groups,
sums, keys and every error vector, with flips in the selected columns."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_bat as OB
from ahead_amd import ssb

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "cpp" / "_build" / "ssb_chains"
SMALL = dict(n_fact=200_000, n_date=2556, n_cust=3_000, n_supp=200, n_part=1_400)
KIND = {"key": "int", "meas": "int", "tiny": "tiny", "short": "short"}


def _need_exe():
    if not EXE.exists():
        pytest.skip("tests/cpp not built (run __graft_entry__.build())")


def _pred(p) -> str:
    return " ".join(str(int(x)) for x in p)


def write_inputs(spec, d: Path) -> None:
    host = ssb.host_tables(spec)
    lines = []
    for name, (codes, meta, _) in host.items():
        codes.tofile(d / f"{name}.bin")
        lines.append(f"{name} {KIND[ssb.COLUMNS[name][1]]} {meta.a} {codes.size}")
    (d / "tables.txt").write_text("\n".join(lines) + "\n")
    q = []
    for name, qq in ssb.QUERIES.items():
        q.append(f"query {name} {qq.family}")
        for col, p in qq.preds:
            q.append(f"pred {col} {_pred(p)}")
        for dim in qq.dims:
            q.append(f"dim {dim.name} {dim.attr or '-'}")
            for col, p in dim.sels:
                q.append(f"sel {col} {_pred(p)}")
    (d / "queries.txt").write_text("\n".join(q) + "\n")


def read_result(path: Path) -> dict:
    r = {"keys": [], "errors": {}}
    for line in path.read_text().splitlines():
        f = line.split()
        if f[0] in ("rows", "result", "groups"):
            r[f[0]] = int(f[1])
        elif f[0] == "sums":
            r["sums"] = np.array([int(x) for x in f[1:]], dtype=np.uint64)
        elif f[0] == "keys":
            r["keys"].append(np.array([int(x) for x in f[1:]], dtype=np.uint64))
        elif f[0] == "err":
            r["errors"][f[1]] = np.array([int(x) for x in f[2:]], dtype=np.uint64)
    return r


def test_cpp_ssb_program_links_without_gpu():
    _need_exe()
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the gpu test runs the program")
    res = subprocess.run([str(EXE), "/nonexistent"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 2 and "no GPU" in res.stdout, (res.returncode, res.stdout, res.stderr)


@pytest.mark.gpu
def test_cpp_ssb_chains_match_the_oracle(tmp_path):
    _need_exe()
    # flips in the selected columns, except the group attributes (p_category is Q4.2's): a corrupt
    # fetched group attribute makes the reference's chain throw (groupbyAN.tcc:105-107)
    attrs = {d.attr for q in ssb.QUERIES.values() for d in q.dims}
    flips = tuple(c for c, v in ssb.COLUMNS.items() if v[4] and c not in attrs)
    spec = ssb.SsbSpec(rate=2e-3, seed=81, flip_cols=flips, **SMALL)
    write_inputs(spec, tmp_path)
    print("C++ chains of the thirteen queries", flush=True)
    res = subprocess.run([str(EXE), str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(res.stdout[-2000:])
    assert res.returncode == 0, res.stdout + res.stderr
    host = OB.tables(spec)
    flipped = 0
    for name, q in ssb.QUERIES.items():
        got = read_result(tmp_path / f"out_{name}.txt")
        ref = ssb.chain(host, q, OB)
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
