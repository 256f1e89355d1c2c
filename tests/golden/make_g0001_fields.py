"""Write G0001_fields.npz: the reference's published SI samples of its converged G0001 steady
solves (docs/data/G0001_{LF1,LF2}.npz): on LF1 per control volume p, temperature, mach and
speed; on LF1 and LF2 per wall fragment x, patch and Cp.  For the export parity test.  Run once in a container that holds the
reference; the fixture travels, this script does not need to."""
from pathlib import Path

import numpy as np

SOURCE = Path("/root/reference/docs/data")
out = {}
for level in ("LF1", "LF2"):
    with np.load(SOURCE / f"G0001_{level}.npz", allow_pickle=False) as z:
        keys = ("surface_x", "cp", "patch") + (("p", "temperature", "mach", "speed") if level == "LF1" else ())
        for key in keys:
            out[f"{level}_{key}"] = np.asarray(z[key], dtype=np.float64)
        out[f"{level}_time"] = np.asarray(float(z["time"]))
        out[f"{level}_cells"] = np.asarray(int(z["cells"]))
np.savez_compressed(Path(__file__).with_name("G0001_fields.npz"), **out)
