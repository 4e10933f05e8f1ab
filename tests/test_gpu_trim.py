"""GPU parity of k-clique core trimming against the reference's kClistCoreTrim
outputs (golden vectors) and the oracle: k-clique degrees, the peel order
and running core, the densest-prefix line, the trimmed graph, the stat line."""

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden_index, golden_pairs, load_golden
from kclistpp_amd import graphgen as gg

pytestmark = pytest.mark.gpu

TRIM = golden_index()["trim"]
CLI = os.path.join(ROOT, "kclistpp_amd", "bin", "kClistCoreTrim")


def _say(*a):
    print(*a, flush=True)


def _dens_line(info, mp, pk):
    n2, size = info["n2"], info["size"]
    ids = [int(mp[pk[i]]) for i in range(n2 - 1, n2 - size - 1, -1)] if size != n2 else []
    return "%f %f %u" % (info["ckrho"], info["rho"], size) + "".join(f" {v}" for v in ids)


@pytest.mark.parametrize("gpu_min", [0, 1])
@pytest.mark.parametrize("case", TRIM)
def test_trim_matches_reference(kcl, case, gpu_min):
    """gpu_min = 1: every peel pop through at least one clique counts its
    cliques on the GPU (the default sends only heavy pops there)."""
    gd = load_golden(case)
    k = int(gd["k"])
    _say("trim", case, "gpu_min", gpu_min)
    with kcl.Trim(golden_pairs(gd), k, gpu_peel_min_cliques=gpu_min) as t:
        if gpu_min == 1 and t.info["ncktot"] > 0:
            assert t.info["gpu_pops"] > 0
        mp, nck, pk, pc, kept = t.nodes()
        assert np.array_equal(np.stack([mp, nck]).T.astype(np.uint64), gd["ckdeg"])
        assert np.array_equal(np.stack([mp[pk], pc]).T.astype(np.uint64), gd["ckcore"])
        assert _dens_line(t.info, mp, pk) == str(gd["ckdens"][0])
        assert np.array_equal(t.edges(), gd["trimmed"])
        assert [t.info["ntrim"], t.info["mtrim"], t.info["ncktrim"]] == gd["stat"].tolist()


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
@pytest.mark.parametrize("case", ["trim_cl3k_k4", "trim_k7_k5_bridge_k4"])
def test_trim_cli_files(kcl, case, tmp_path):
    gd = load_golden(case)
    k = int(gd["k"])
    path = str(tmp_path / "g.txt")
    gg.write_text(gd["pairs"], path)
    _say("trim cli", case)
    r = subprocess.run([CLI, str(k), path, "deg.txt", "core.txt", "dens.txt", "tg"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    deg = np.loadtxt(tmp_path / "deg.txt", dtype=np.uint64, ndmin=2)
    core = np.loadtxt(tmp_path / "core.txt", dtype=np.uint64, ndmin=2)
    assert np.array_equal(deg, gd["ckdeg"]) and np.array_equal(core, gd["ckcore"])
    assert open(tmp_path / "dens.txt").read().strip() == str(gd["ckdens"][0])
    trimmed = np.loadtxt(tmp_path / f"trimmed_tg_{k}.txt", dtype=np.uint32, ndmin=2)
    assert np.array_equal(trimmed.reshape(-1, 2), gd["trimmed"])
    stat = open(tmp_path / f"stat_trim_tg_{k}.txt").read().splitlines()
    assert stat[0].startswith("[Number of Nodes]")
    assert [int(x) for x in stat[1].split("\t")[:3]] == gd["stat"].tolist()


@pytest.mark.parametrize("gpu_min", [0, 1, 1000])
@pytest.mark.parametrize("k", [3, 4, 5, 6, 7])
def test_trim_vs_oracle(kcl, oracle, k, gpu_min):
    pairs, rng = gg.chung_lu(2500, 20000, 2.2, 250.0, 40 + k)
    pairs = gg.finalize(gg.plant_community(pairs, 2500, 36, 0.85, rng), rng)
    info_o, mp_o, nck_o, pk_o, pc_o, trimmed_o = oracle.trim(pairs, k)
    _say("trim vs oracle k", k, "gpu_min", gpu_min)
    with kcl.Trim(pairs, k, gpu_peel_min_cliques=gpu_min) as t:
        mp, nck, pk, pc, kept = t.nodes()
        assert np.array_equal(mp, mp_o) and np.array_equal(nck, nck_o)
        assert np.array_equal(pk, pk_o) and np.array_equal(pc, pc_o)
        assert np.array_equal(t.edges(), trimmed_o)
        for f in ("n2", "e2", "core", "ncktot", "ckcore", "ckrho", "rho", "size", "ntrim",
                  "ncktrim", "mtrim"):
            assert t.info[f] == info_o[f], f


def test_trim_degrees_sum_to_k_times_cliques(kcl, oracle):
    pairs = gg.workload("cl1m", seed=4, scale_down=128)
    _say("trim cl1m/128 k=5")
    with kcl.Trim(pairs, 5) as t:
        _, nck, _, _, _ = t.nodes()
        assert int(nck.sum()) == 5 * t.info["ncktot"]
        og = oracle.dag(pairs)
        assert t.info["ncktot"] == oracle.count(og, 5)
