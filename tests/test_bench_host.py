"""bench.py host logic (no GPU): the algorithmic byte model and the choice of
the committed profile summary whose measured traffic a run reports."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

bench = pytest.importorskip("bench")


def test_comm_bytes_per_stage_model():
    """RK4 stages move (6, 6, 8, 6) S at n_E = 1: 6.5 S on average; AB4 with
    n_E = 2 and exchange: 4 + 2 + 4 + 1 = 11 S; Peierls: 4.5 S."""
    npl, nk = 196, 14400
    s = 16.0 * npl * nk
    assert bench.kernel_bytes("comm", npl, nk, 14, 1) == pytest.approx(6.5 * s)
    assert bench.kernel_bytes("comm", 64, 8100, 8, 2, "AB", 4, True) == pytest.approx(11 * 16.0 * 64 * 8100)
    assert bench.kernel_bytes("comm", npl, nk, 14, 1, peierls=True) == pytest.approx(4.5 * s)
    assert bench.kernel_bytes("grad_col", npl, nk, 14, 1) == pytest.approx(2 * s)
    assert bench.kernel_bytes("grad_row", npl, nk, 14, 1) == pytest.approx(3 * s)
    # mixed precision reads H0 and the one xi from FP32 copies: 5.5 S
    assert bench.kernel_bytes("comm", npl, nk, 14, 1, f32_fields=True) == pytest.approx(5.5 * s)
    # upper triangles of H0 and xi: 105 of 196 planes each, 6.5 - 2 (91 / 196) = 5.571 S;
    # in FP32 as well: 6.5 - 2 (1 - 105 / 392) = 5.036 S
    assert bench.kernel_bytes("comm", npl, nk, 14, 1, tri_fields=True) == pytest.approx((6.5 - 2 * 91 / 196) * s)
    assert bench.kernel_bytes("comm", npl, nk, 14, 1, f32_fields=True, tri_fields=True) == \
        pytest.approx((6.5 - 2 * (1 - 105 / 392)) * s)
    # exchange with the forward row pass taken from the gradient's: grad_row writes it too
    assert bench.kernel_bytes("grad_row", 64, 8100, 8, 2, "AB", 4, True, hf_fused=True) == \
        pytest.approx(4 * 16.0 * 64 * 8100)
    assert bench.kernel_bytes("grad_row", npl, nk, 14, 1, hf_fused=True) == pytest.approx(3 * s)
    # the Peierls gauge reads S_A[H] in place of H0 and xi whatever their storage
    assert bench.kernel_bytes("comm", npl, nk, 14, 1, peierls=True, tri_fields=True) == pytest.approx(4.5 * s)


def test_traffic_summary_matches_the_run_variant():
    t, src = bench.measured_traffic("comm", "ipa14_hex120", "full", "fp64")
    assert src == os.path.join("profiles", "comm_kernel_ipa14_hex120_rk4_full_fp64.txt")
    assert t == pytest.approx(bench.kernel_bytes("comm", 196, 14400, 14, 1), rel=0.05)
    # each variant reads its own summary, or none
    t, src = bench.measured_traffic("comm", "ipa14_hex120", "full", "fp64", gauge="peierls")
    assert src.endswith("_peierls.txt")
    assert t == pytest.approx(bench.kernel_bytes("comm", 196, 14400, 14, 1, peierls=True), rel=0.05)
    t, src = bench.measured_traffic("comm", "ipa14_hex120", "full", "fp64", gradient="k")
    assert src.endswith("_kgradient.txt")
    t, src = bench.measured_traffic("kgrad", "ipa14_hex120", "full", "fp64", gradient="k")
    assert src is not None and t == pytest.approx(2 * 16.0 * 196 * 14400, rel=0.05)
    assert bench.measured_traffic("kgrad", "ipa14_hex120", "full", "fp64") == (None, None)
    t, src = bench.measured_traffic("comm", "ipa14_hex120", "hermitian", "fp64")
    assert t == pytest.approx(bench.kernel_bytes("comm", 105, 14400, 14, 1), rel=0.05)
    t, src = bench.measured_traffic("comm", "ipa14_hex120", "full", "mixed")
    assert src.endswith("_mixed.txt")
    assert t == pytest.approx(bench.kernel_bytes("comm", 196, 14400, 14, 1, f32_fields=True), rel=0.05)
    t, src = bench.measured_traffic("grad_fused", "ipa14_hex120", "full", "fp64")
    assert src.endswith("grad_fused_kernel_ipa14_hex120_rk4_full_fp64.txt")
    for prec in ("fp64", "mixed"):
        t, src = bench.measured_traffic("rotate", "rotate_ipa14_hex120", "full", prec)
        assert src.endswith(f"rotate_kernel_rotate_ipa14_hex120_{prec}.txt")
        assert t == pytest.approx(3 * 16.0 * 196 * 14400, rel=0.05)
    assert bench.measured_traffic("comm", "hsex8_hex90", "hermitian", "fp64") == (None, None)
    assert bench.measured_traffic("comm", "ipa14_hex120", "full", "fp64", gradient="k", gauge="peierls") == (None, None)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_graph_capture_stays_out_of_the_timed_region():
    """The timed steps are whole graphs, and the warm-up runs at least one
    whole graph, so capture and instantiation happen before t0."""
    for steps, graph, warmup in ((20, 20, 5), (200, 20, 20), (25, 20, 3), (7, 20, 0), (30, 8, 1)):
        g = bench.graph_chunk(steps, graph)
        assert 1 <= g <= graph and steps % g == 0
        w = bench.warmup_steps(warmup, g)
        assert w >= warmup and w >= g and w % g == 0
    assert bench.graph_chunk(20, 0) == 0 and bench.warmup_steps(5, 0) == 5
