"""The algorithmic-bytes contract of one RK stage (roofline.py), pinned by hand counts."""
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import roofline as RL

from .common import case_mesh


def test_stage_bytes_hand_count():
    # n=10, nf=20, ncsr=40, nt=50, nw=3, SA (nv=5, g=6), state 8 B, work/geometry 4 B:
    # state 8*3*5*10 = 1200, gradients 4*2*6*10 = 480, faces 20*(12+24) = 720,
    # CSR 4*(40+11) = 204, cells 10*(12+8) = 200, terms 50*(4+8) + 4*11 = 644, walls 3*(8+16) = 72.
    assert RL.stage_bytes(10, 20, 40, 50, 3, 5, True, True, wu=8, ws=4, wg=4) == 3520
    # Euler: no gradients and no least-squares terms; base stage adds the FP64 local dt round trip.
    assert RL.stage_bytes(10, 20, 40, 0, 3, 4, False, False, wu=8, ws=8, wg=8) == (
        8 * 3 * 4 * 10 + 20 * 60 + 4 * 51 + 10 * 32 + 3 * 40)
    assert RL.stage_bytes(10, 20, 40, 0, 3, 4, False, False, 8, 8, 8, base=True) - \
        RL.stage_bytes(10, 20, 40, 0, 3, 4, False, False, 8, 8, 8) == 160


def test_lf2_sa_fp32_example_is_2p7_mb():
    # The worked example of the contract: LF2 SA FP32, n = 9533, nf ~ 19.8k, ncsr ~ 39.7k,
    # nt ~ 40.4k, nw = 686 -> about 2.7 MB per stage with one 4-byte width.
    b = RL.stage_bytes(9533, 19800, 39700, 40400, 686, 5, True, True, 4, 4, 4)
    assert abs(b - 2.7e6) < 0.05e6


def test_design_bytes_exceed_contract_by_the_fp64_state():
    _, mesh, _ = case_mesh("G0001", "LF2")
    n = mesh.n_cells
    d, c = RL.design_stage_bytes(mesh, 32), RL.contract_stage_bytes(mesh, 32)
    assert d - c == 4 * 3 * 5 * n
    assert RL.design_stage_bytes(mesh, 64) == RL.contract_stage_bytes(mesh, 64)


def test_bench_reads_the_profile_summary_of_its_own_workload_only():
    # bench.py reports measured HBM traffic only from the committed rocprofv3 summary whose
    # workload line matches (batch, tile, seed); the summaries carry the figure they state.
    import importlib.util
    import re
    from pathlib import Path

    root = Path(__file__).resolve().parents[1]
    spec = importlib.util.spec_from_file_location("bench_module", root / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for path in sorted((root / "profiles").glob("stage_pass_*.md")):
        text = path.read_text()
        w = re.search(r"^stage_pass_profile: batch=(\d+) tile=(\d+) seed=(\d+)$", text, re.M)
        if not w:
            continue
        batch, tile, seed = (int(x) for x in w.groups())
        stated = int(re.search(r"^stage_pass_hbm_bytes: (\d+)$", text, re.M).group(1))
        assert bench.profiled_traffic(batch, tile, seed) == stated
        # The summary's MB figure is the same measurement.
        mb = float(re.search(r"Measured HBM traffic: ([\d.]+) MB per stage pass", text).group(1))
        assert abs(mb - stated / 1e6) < 0.1
    assert bench.profiled_traffic(256, 128, 1) is None
    assert bench.profiled_traffic(255, 128, 2026) is None
    assert bench.profiled_traffic(256, 256, 2026) is None
