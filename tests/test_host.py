"""CPU-side checks of the product library: exports, edge-list reader and
writer, CLI argument handling, workload generators.  No GPU calls."""

import hashlib
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT


def _declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(kcl_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(kcl):
    declared = _declared_functions()
    assert len(declared) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", kcl.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (kcl_\w+)", out))
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    for name in declared:  # resolvable through ctypes as well
        getattr(kcl.lib(), name)


def test_python_wrapper_binds_declared_symbols(kcl):
    assert set(kcl.EXPORTS) <= set(_declared_functions())


def test_reader_text_semantics(kcl, tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("0 1\n  2\t3\n\n4 5 6 7\n8 x 9 10\n")
    pairs, n_in = kcl.read_edge_list(str(p))
    # fscanf("%u %u") stops at the first token that does not parse
    assert pairs.tolist() == [[0, 1], [2, 3], [4, 5], [6, 7]]
    assert n_in == 8


def test_reader_signs_wrap_like_scanf(kcl, tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("+1 2\n-5 3\n")
    pairs, n_in = kcl.read_edge_list(str(p))
    assert pairs.tolist() == [[1, 2], [2**32 - 5, 3]]
    assert n_in == 2**32 - 4


def test_reader_rejects_max_id(kcl, tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("0 4294967295\n")
    with pytest.raises(kcl.KclError) as e:
        kcl.read_edge_list(str(p))
    assert e.value.name == "KCL_E_UNSUPPORTED"


def test_reader_empty_and_missing(kcl, tmp_path):
    p = tmp_path / "empty.txt"
    p.write_text("")
    pairs, n_in = kcl.read_edge_list(str(p))
    assert pairs.shape == (0, 2) and n_in == 1  # the reference's n = 0 + 1
    with pytest.raises(kcl.KclError) as e:
        kcl.read_edge_list(str(tmp_path / "nope.txt"))
    assert e.value.name == "KCL_E_IO"


def test_writer_round_trip(kcl, tmp_path):
    rng = np.random.default_rng(0)
    pairs = rng.integers(0, 10_000, (5000, 2)).astype(np.uint32)
    for name in ("a.txt", "a.u32"):
        path = str(tmp_path / name)
        kcl.write_edge_list(path, pairs)
        back, n_in = kcl.read_edge_list(path)
        assert np.array_equal(back, pairs) and n_in == int(pairs.max()) + 1
    bad = tmp_path / "bad.u32"
    bad.write_bytes(b"\x00" * 12)
    with pytest.raises(kcl.KclError) as e:
        kcl.read_edge_list(str(bad))
    assert e.value.name == "KCL_E_PARSE"


def test_reader_matches_oracle_reader(kcl, oracle, tmp_path):
    from kclistpp_amd import graphgen as gg
    pairs = gg.workload("er20k", seed=3, scale_down=4)
    path = str(tmp_path / "g.txt")
    gg.write_text(pairs, path)
    got, n_in = kcl.read_edge_list(path)
    assert np.array_equal(got, pairs)
    assert n_in == int(pairs.max()) + 1


def _fscanf_text(rng, tokens, stop_at=None):
    """Random numbers joined by assorted white space, with signed and glued
    forms; a text that does not parse at token index `stop_at`."""
    out = []
    for i in range(tokens):
        r = rng.random()
        if stop_at is not None and i == stop_at:
            t = str(int(rng.integers(0, 9))) + "q"
        elif r < 0.02:
            t = "-" + str(int(rng.integers(2, 100)))
        elif r < 0.03:
            t = str(int(rng.integers(0, 9))) + "-" + str(int(rng.integers(2, 9)))
        elif r < 0.04:
            t = "+" + str(int(rng.integers(0, 9)))
        else:
            t = str(int(rng.integers(0, 10**7)))
        out.append(t + (" ", "\n", "\t", "  ", " \r\n")[int(rng.integers(0, 5))])
    return "".join(out)


@pytest.mark.parametrize("stop", [None, 0.37, 0.999])
def test_reader_threaded_matches_fscanf(kcl, oracle, tmp_path, stop):
    # > 4 MiB of text: the reader cuts it into chunks parsed by several
    # threads; the result must equal the fscanf("%u %u") loop, including
    # where text that does not parse sits deep inside one chunk.
    rng = np.random.default_rng(11)
    tokens = 900_001
    text = _fscanf_text(rng, tokens, None if stop is None else int(stop * tokens))
    assert len(text) > (1 << 22)
    path = str(tmp_path / "big.txt")
    with open(path, "w") as f:
        f.write(text)
    got, n_in = kcl.read_edge_list(path)
    want, n_want = oracle.read_edge_list(path)
    assert np.array_equal(got, want)
    assert n_in == n_want


CLI = os.path.join(ROOT, "kclistpp_amd", "bin", "kCDensestCmpSync")


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
@pytest.mark.parametrize("args,code", [
    ([], 2), (["8", "10", "4", "x"], 2), (["8", "ten", "4", "f", "t"], 2),
    (["8", "10", "1", "f", "t"], 2), (["8", "10", "256", "f", "t"], 2),
    (["8", "10", "4", "/nonexistent/file.txt", "t"], 1),
])
def test_cli_argument_errors(args, code, tmp_path):
    r = subprocess.run([CLI, *args], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == code, r.stderr


def test_generators_are_deterministic():
    from kclistpp_amd import graphgen as gg
    a = gg.workload("er20k", seed=1)
    b = gg.workload("er20k", seed=1)
    assert np.array_equal(a, b)
    assert a.shape == (100435, 2)
    # no loops, no repeats in either orientation
    assert (a[:, 0] != a[:, 1]).all()
    key = np.minimum(a[:, 0], a[:, 1]).astype(np.uint64) * 20000 + np.maximum(a[:, 0], a[:, 1])
    assert np.unique(key).size == a.shape[0]
    c = gg.workload("cl1m", seed=2, scale_down=256)
    assert hashlib.sha256(c.tobytes()).hexdigest() == \
        hashlib.sha256(gg.workload("cl1m", seed=2, scale_down=256).tobytes()).hexdigest()


def test_certificate_max_flow_matches_edmonds_karp(tmp_path):
    """The exact pipeline's implicit-arc max-flow (clique_flow.h) against a
    textbook Edmonds-Karp on the explicit network, 400 random instances,
    then a 60-node community (k = 5) end to end."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "cfc")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "kclistpp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "clique_flow_check.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, CLIQUE_FLOW_COMMUNITY="60"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches: 0" in r.stdout and "saturated 1" in r.stdout


@pytest.mark.parametrize("case", ["threaded", "sparse_ids", "small"])
def test_host_degeneracy_order_matches_oracle(oracle, tmp_path, case):
    """The host replay of the reference's heap (host_order.cpp) against the
    oracle's restatement, on a graph large enough for the threaded
    adjacency build (m >= 2^22, ids dense), one whose sparse ids keep it
    sequential, and a small one."""
    import shutil
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    from kclistpp_amd import graphgen as gg
    if case == "threaded":
        pairs = gg.workload("cl1m", seed=4, scale_down=3)
        assert pairs.shape[0] >= (1 << 22)
    elif case == "sparse_ids":
        rng = np.random.default_rng(9)
        ids = np.sort(rng.choice(200_000_000, 100_000, replace=False)).astype(np.uint32)
        pairs = gg.workload("er20k", seed=3, scale_down=1)
        pairs = np.ascontiguousarray(ids[pairs % ids.size])
        pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    else:
        pairs = gg.workload("er20k", seed=5, scale_down=1)
    exe = str(tmp_path / "dgc")
    subprocess.run([gxx, "-O2", "-std=c++17", "-pthread",
                    "-I", os.path.join(ROOT, "kclistpp_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                    "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                    os.path.join(ROOT, "tests", "native", "degeneracy_check.cpp"),
                    os.path.join(ROOT, "kclistpp_amd", "csrc", "host_order.cpp"), "-o", exe],
                   check=True)
    src, dst = str(tmp_path / "pairs.u32"), str(tmp_path / "rank.u32")
    np.ascontiguousarray(pairs, "<u4").tofile(src)
    subprocess.run([exe, src, dst], check=True, timeout=300)
    got = np.fromfile(dst, "<u4")
    n = int(pairs.max()) + 1
    assert got.size == n
    assert np.array_equal(got, oracle.rank(pairs, n))
