"""Mesh and control variants shared by the oracle pins and the GPU parity tests: every domain-side
kind, the reference's published control runs, fixed-dt and rejection schedules, fatal cases and
the SRD-grouped viscous limit."""
import math

from .common import EULER, LAMINAR, merge

GAMMA = 1.4
GM = GAMMA - 1
KINDS = {"transmissive": 1, "reflect": 2, "farfield": 3, "total-inlet": 4, "pressure-outlet": 5, "no-slip": 6}

# The reference's control cases (validation/exercise_controls.py:5-7), as configuration overrides.
SUBSONIC_OUTLET = {
    "freestream": {"Mach": 0.7, "AOA_degrees": 3.0, "p_Pa": 30000.0, "T_K": 250.0, "chord_m": 0.8},
    "physics": {"turbulence": "none", "viscosity": "constant", "constant_mu_Pa_s": 1.7e-5},
    "boundaries": {"xmax": "pressure-outlet", "outlet_pressure_Pa": 30000.0},
}
FIXED_DT = {"solver": {"time_step": {"mode": "fixed", "dt_seconds": 1e-7}}}
AUSM_CFL01 = {"physics": {"turbulence": "none", "viscosity": "sutherland"}, "solver": {"flux": "ausm", "CFL": 0.1}}
CONTROLS = {"subsonic_outlet_LF2_G0001": SUBSONIC_OUTLET, "fixed_dt_1e-7s_LF2_G0001": FIXED_DT,
            "cpu_ausm_LF2_G0001": AUSM_CFL01}

SUBSONIC = {"freestream": {"Mach": 0.5, "AOA_degrees": 2.0}, "physics": {"turbulence": "none"}}


def total_inlet_side(mach=0.5, aoa=2.0):
    """xmin total inlet whose total state is that of the nondimensional freestream (T* = p/rho)."""
    t = 1.0 / GAMMA
    factor = 1 + 0.5 * GM * mach * mach
    a = math.radians(aoa)
    return {0: dict(kind=KINDS["total-inlet"], total_pressure=t * factor ** (GAMMA / GM),
                    total_temperature=t * factor, gas_constant=1.0, direction=(math.cos(a), math.sin(a)))}


# Mesh variants with every domain-side kind: (airfoil, preset, overrides, side overrides).
VARIANTS = {
    "outlet_subsonic": ("G0001", "LF1", SUBSONIC_OUTLET, None),
    "reflect_laminar": ("G0002", "LF1", LAMINAR, {2: dict(kind=2), 3: dict(kind=2)}),
    "reflect_euler": ("G0002", "LF1", EULER, {2: dict(kind=2), 3: dict(kind=2)}),
    "transmissive_laminar": ("G0010", "LF1", LAMINAR, {1: dict(kind=1), 2: dict(kind=1), 3: dict(kind=1)}),
    "total_inlet": ("G0100", "LF1", SUBSONIC, total_inlet_side()),
    "noslip_side_laminar": ("G0001", "LF1", LAMINAR, {2: dict(kind=6, wall_velocity=(0.3, 0.0))}),
    "noslip_side_sa": ("G0001", "LF1", None, {3: dict(kind=6, wall_velocity=(-0.2, 0.0))}),
    "inherit_legacy": ("G0200", "LF1", LAMINAR, {3: dict(kind=-1)}),
}

# Fixed-dt schedules on G0001 LF1 laminar (dt, t*, steps, rejected trials): a plain fixed step,
# a step whose first trial always fails once (dt halves, then every step is a retry), and one
# that fails twice per step.  The counts are the compiled reference's.
FIXED_SCHEDULES = [(2.0e-4, 0.01, 50, 0), (0.0128, 0.2, 31, 31), (0.0256, 0.4, 63, 127)]

# Low Reynolds number (constant mu 1 Pa s, Re_c about 150) on a non-agglomerated laminar mesh
# with SRD viscous stabilization, so the SRD-grouped viscous limit binds.
VISCOUS_SRD = merge(LAMINAR, {"physics": {"viscosity": "constant", "constant_mu_Pa_s": 1.0}},
                    {"mesh": {"agglomerate_small": False}, "physics": {"viscous_stabilization": "srd"}})
