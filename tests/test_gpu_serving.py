"""Serving loop on the GPU: requests served in a mix give bitwise their results alone.

Four requests (different airfoils and sizes) arrive as a Poisson stream into a two-slot FP32
arena and run the time-accurate check loop to t* = 0.03; each result (steps, time, CL, CD,
CM_z, raw residuals) must equal the same request solved in a session of its own.
"""
import pytest

from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import batch as B
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import config as CF
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd import serve as S
from quick_gpu_cfd_for_neural_network_with_converging_inputs_airfoil2d_version_amd.session import Session

from .common import LAMINAR, example_polygon, naca_polygon

pytestmark = pytest.mark.gpu


def test_serving_mix_equals_alone_bitwise():
    polys = [example_polygon("G0001"), naca_polygon(0.12), example_polygon("G0200"), naca_polygon(0.09, 0.03)]
    presets = ["LF1", "LF2", "LF1", "LF1"]
    arrivals = S.poisson_arrivals(40.0, len(polys), seed=7)
    reqs = [S.Request(i, polys[i], CF.make_case(presets[i], LAMINAR), float(arrivals[i])) for i in range(4)]
    built = [S.build_request_mesh(r) for r in reqs]
    caps, feats = zip(*[B.capacity(m) for m, _, _ in built])
    backend = S.DeviceBackend([B.envelope(caps)] * 2, B.union(feats), precision=32)
    try:
        rep = S.Server(backend, mode="time", stop_time=0.03, budget=120, workers=2).serve(reqs)
    finally:
        backend.close()
    print({k: v for k, v in rep.items() if k != "outcomes"}, flush=True)
    assert rep["completed"] == 4
    for req, (mesh, params, u0), out in zip(reqs, built, rep["outcomes"]):
        with Session(mesh, params, u0, precision=32) as s:
            ref = s.solve(req.case, mode="time", stop_time=0.03, budget=120, history=0)
        got = out.result
        print(req.rid, out.slot, got["steps"], ref["steps"], got["cl"], ref["cl"], flush=True)
        for key in ("steps", "rejected", "time", "cl", "cd", "cm", "residual", "checks", "reason"):
            assert got[key] == ref[key], key
