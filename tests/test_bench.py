"""bench.py host logic on the CPU: the one-GPU guard and the table-run
roofline arithmetic (no GPU is touched)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_tt_roofline_takes_the_larger_phase():
    sys.path.insert(0, ROOT)
    import bench
    from perft_gpu_amd.workloads import C2_OPS_PER_MADE_NODE
    ceil = {"random16_per_s": 2e10, "copy_bytes_per_s": 5e12}
    st = {"tt_probes": 1000, "tt_stores": 500, "tt_hits": 100, "frontier_nodes": 4000, "dup_nodes": 2000,
          "made_nodes": 10000}
    r = bench.tt_roofline(st, 1e-3, ceil)
    rnd = 1000 + 500 + 4000 + 3 * 2000
    t_mem = rnd / 2e10 + 2 * bench.FRONTIER_NODE_BYTES * 4000 / 5e12
    t_valu = (10000 - 2000 - 100) * C2_OPS_PER_MADE_NODE / (bench.VALU_PEAK_TOPS * 1e12)
    assert r["random_accesses"] == rnd and r["counted_nodes"] == 7900
    assert abs(r["t_memory_s"] - t_mem) < 1e-15 and abs(r["t_valu_s"] - t_valu) < 1e-15
    assert abs(r["frac"] - max(t_mem, t_valu) / 1e-3) < 1e-12
    assert abs(r["serial_frac"] - (t_mem + t_valu) / 1e-3) < 1e-12
    assert r["bound"] == ("valu" if t_valu >= t_mem else "hbm-random")
    json.dumps(r)


def test_issue_view_divides_counted_instructions_by_simd_cycles():
    """a C2 step's issue rate: instructions per leaf x leaves / 64 lanes over
    1,024 SIMDs x 2.4 GHz x the kernel time, against the loop mix's limit;
    only Kogge-Stone has a measured instruction count"""
    sys.path.insert(0, ROOT)
    import bench
    from perft_gpu_amd import workloads as W
    r = bench.issue_view(W.C2_LEAVES, 5.0, "ks")
    want = W.C2_VALU_LANE_INSTR_PER_LEAF * W.C2_LEAVES / 64 / (5e-3 * 1024 * 2.4e9)
    assert abs(r["wave_instr_per_simd_cycle"] - want) < 1e-12
    assert abs(r["frac"] - want / W.C2_LOOP_ISSUE_LIMIT) < 1e-12
    assert 0.25 < W.C2_LOOP_ISSUE_LIMIT < 0.42  # between all single-rate and all double-rate issue
    assert bench.issue_view(W.C2_LEAVES, 5.0, "fancy") is None
    json.dumps(r)


def test_op_model_per_made_node_matches_the_fixture():
    """the C2 tree's made nodes are its perft 1..4 sums, and the op total is
    the fixture's"""
    with open(os.path.join(ROOT, "tests", "golden", "workloads.json")) as f:
        c2 = json.load(f)["c2"]
    from perft_gpu_amd import workloads as W
    made = sum(sum(c2["perft"][str(d)]) for d in range(1, 5))
    assert W.C2_MADE_NODES == made
    assert abs(W.C2_OPS_PER_MADE_NODE - c2["op_model"]["int32_lane_ops"] / made) < 1e-9


def test_cpu_c1_reports_the_initial_position_count():
    """the C1 CPU leg: Kogge-Stone oracle on one thread and split by root move
    over several, median / p10 / p90 of the runs"""
    sys.path.insert(0, ROOT)
    import bench
    from perft_gpu_amd import START_FEN
    r = bench.cpu_c1(START_FEN, 4, 197281, 3, reps=3)
    assert r["counts_match"] and r["cores"] == 1 and r["backend"] == "ks" and r["runs"] == 3
    assert abs(r["value"] - 197281 / r["median_s"]) < 1e-6 * r["value"]
    assert r["p10_s"] <= r["median_s"] <= r["p90_s"]
    a = r["all_cores"]
    assert a["cores"] == 3 and abs(a["value"] - 197281 / a["median_s"]) < 1e-6 * a["value"]


def test_host_cpus_reports_the_share_used():
    sys.path.insert(0, ROOT)
    import bench
    h = bench.host_cpus()
    assert 1 <= h["threads"] <= 16 and h["threads"] <= h["allowed"] and h["nproc"] >= 1


def test_cpu_subset_legs_check_against_the_deep_goldens():
    """the C3 / C5 CPU legs on tiny slices: counts equal to the golden ones
    (the deep golden file holds the C3 divide and the C5 perft 6 counts)"""
    sys.path.insert(0, ROOT)
    import bench
    from perft_gpu_amd import workloads as W
    g = bench.load_goldens()
    if "c5_deep" in g:
        r = bench.cpu_c5_subset(W.positions(W.C5), g["c5_deep"], 4, n=1)
        assert r["counts_match_golden"] is True and r["nodes"] == g["c5_deep"]["perft6"][0]
    assert "c3_deep" in g and g["c3_deep"]["perft8"] == sum(g["c3_deep"]["divide_d8"].values())
    # the golden divide's root moves are the C3 position's legal moves
    assert len(g["c3_deep"]["divide_d8"]) == len(g["c3_deep"]["moves"])
