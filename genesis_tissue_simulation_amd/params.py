"""Solver parameters of the XPBD substep, shared by the HIP path and its checkers.

Each field is the value the reference derives from `PhysicsDesc` / `XpbdParams` / `XpbdGraspParams` and the contact
model (`xpbd/xpbd_solver.h:38-75`, `physics.cu:281-295`, `xpbd/xpbd_solver.cu:2200-2210`).
"""
from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np

# tet constraint forms of k_solve_tet (`xpbd/xpbd_solver.cu:749-829`)
TET_VOLUME = 0      # hydrostatic row only, C = J - 1, one compliance per tet
TET_NEOHOOKEAN = 1  # coupled deviatoric / hydrostatic 2x2 block, C_H = J - 1 (legacy)
TET_GUARDED = 2     # coupled block with C_H = log J + min(0, (J - .05)/.05) and the per-update trust region


@dataclass(frozen=True)
class SolverParams:
    substeps: int = 10
    iterations: int = 1
    tet_mode: int = TET_VOLUME
    omega: float = 1.0
    edge_ks: float = 1e7          # > 0: every live edge, compliance 1/ks; <= 0: fallback edges E*L only
    young: float = 0.0            # native Young's modulus (Pa * 0.01) of the fallback edge law
    beta_s: float = 0.0           # edge Kelvin-Voigt damping stiffness
    strain_damping: float = 0.0   # tet Kelvin damping time (s)
    bulk_damping: float = 0.0     # guarded bulk damping time (s)
    rayleigh: float = 0.0         # global velocity damping (1/s)
    damping: float = 0.0          # legacy per-substep velocity kill
    contact_enabled: bool = False
    mu_s: float = 0.6
    mu_d: float = 0.4
    jaw_thickness: float = 0.1
    contact_radius: float = 0.01
    jaw_gap_closed: float = 0.0
    contact_compliance: float = 1e-8
    precise_strain: bool = False  # tet invariants from F - I (closer to FP64); False: the reference's direct forms
    # tissue guard and material (LCS_SetTissueGuard / LCS_SetTissueMaterial, xpbd_solver.cu:1895-1915, :2165-2171)
    guard_min_volume: float = 0.0  # minimum J of a live tet kept by the end-of-substep safeguard, 0..0.95
    guard_passes: int = 0          # tensile reinforcement passes after every tet sweep, 0..4; 0 turns the guard off
    tensile_strength: float = 80.0  # strength of the s^6 fibre energy (the reference's default)
    # bending angle constraints (PhysicsDesc kb / betaB, xpbd_solver.cu:2496-2499, :534-573); kb <= 0 turns them off
    bend_kb: float = 0.0            # bend stiffness (native force units per radian)
    beta_b: float = 0.0             # bend Kelvin-Voigt damping stiffness

    def with_(self, **changes):
        return replace(self, **changes)


# contact models of xpbd_set_contact_model (`xpbd/xpbd_solver.h:291-293`)
CONTACT_LEGACY = 0    # grasp capture with slab ring, core pins and break check; CCD gate against shaft and jaws
CONTACT_FRICTION = 1  # swept friction contact of every node with the tool shapes


@dataclass(frozen=True)
class GraspParams:
    """XpbdGraspParams (`xpbd/xpbd_solver.h:250-289`), named as `native_backend.py`'s GraspParams fields."""
    capture: float = 0.5              # grasp01 above which the jaws capture
    release: float = 0.3              # grasp01 below which they release
    closed_gap: float = 0.0           # floor of the jaw gap
    core_fraction: float = 0.55       # normalised-ellipse core fraction
    friction: float = 0.6             # Coulomb mu of the captured ring
    contact_compliance: float = 1e-8  # slab and CCD gate compliance (scaled by 1/h^2)
    skin: float = 0.0                 # capture band widening
    ccd_velocity: float = 0.5         # approach speed above which the CCD gate fires
    ccd_displacement: float = 0.0     # per-substep cap of a CCD correction
    ccd_omega: float = 0.25
    grasp_omega: float = 0.25         # under-relaxation of the ring and core pins
    core_compliance: float = 1e-5     # core pin compliance (scaled by 1/h^2)
    break_force: float = 15000.0      # core force EMA above which a core corner lets go
    grasp_displacement: float = 0.0   # per-substep cap of a grasp correction
    exempt_rings: int = 1             # tear-law ring dilation (no ring on this path)

    def with_(self, **changes):
        return replace(self, **changes)


def legacy_grasp_params(spacing):
    """The grasp parameters `NativeTissue.initialize` sends for a lattice of `spacing` (`native_backend.py:314-317`)."""
    s = float(spacing)
    return GraspParams(skin=s * 0.5, ccd_displacement=s * 0.2, grasp_displacement=s * 0.2)


def volume_stiffness(compliance):
    """Per-tet stiffness k = 1/compliance (float32), so that alpha~ = (1/h^2)/k as in k_ComputeTetAlpha
    (`xpbd/xpbd_solver.cu:578-587`)."""
    c = np.asarray(compliance, dtype=np.float64)
    return (1.0 / c).astype(np.float32)


def neohookean_stiffness(young, poisson, type_volume, tet_type):
    """mu*V and kappa*V per tet as `xpbd_build_tets` bakes them (`xpbd/xpbd_solver.cu:1636-1644, :1730-1731`)."""
    nu = np.float32(poisson)
    E = np.float32(young)
    if E > 0 and -1 < nu < 0.5:
        mu = np.float32(E / (np.float32(2) * (np.float32(1) + nu)))
        kappa = np.float32(E / (np.float32(3) * (np.float32(1) - np.float32(2) * nu)))
    else:
        mu = kappa = np.float32(1)
    v = np.asarray(type_volume, dtype=np.float64).astype(np.float32)[np.asarray(tet_type)]
    return (mu * v).astype(np.float32), (kappa * v).astype(np.float32)
