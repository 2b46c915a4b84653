"""A small tissue stepped as one persistent workgroup (Tissue(persistent=True): the batched kernel over the tissue's
own buffers, colours separated by workgroup barriers) against the FP32 oracle: the C1 block with the blade cutting
every step and gravity ramping in, in the direct and the precise tet forms, with one and two iterations and strain
damping, then compacted and stepped on; positions, velocities, multipliers, particle frames and cut topology bit-exact after every step. A
tissue over the colour-size limit, or with tool contact armed, takes the per-colour launches."""
import numpy as np
import pytest

from genesis_tissue_simulation_amd.workloads import block_6144

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precise,iterations,damping", [(False, 1, 0.0), (True, 2, 0.01), (True, 1, 0.0)])
def test_persistent_small_tissue_matches_oracle(precise, iterations, damping):
    import torch

    from genesis_tissue_simulation_amd.tissue import Tissue, swept_cut_tool
    from oracle.oracle import Oracle

    w = block_6144(precise=precise)
    p = w.params.with_(iterations=iterations, strain_damping=damping)
    kd, kh = w.stiffness()
    print(f"persistent block-6144: precise {precise}, iterations {iterations}, strain damping {damping}", flush=True)
    t = Tissue(w.grid, p, kd, kh, w.initial_positions, persistent=True)
    o = Oracle(w.grid, p, w.initial_positions)
    o.set_tet_stiffness(kd, kh)
    for sim in (t, o):
        sim.set_gravity(w.gravity)
    for s in range(6, 20):
        tool = swept_cut_tool(w.blade(s), w.blade(s + 1), w.blade_thickness)
        ramp = min(1.0, (s - 5) / 6.0)  # gravity ramping in, as NativeTissue.step does
        for sim in (t, o):
            sim.set_gravity(tuple(ramp * g for g in w.gravity))
            sim.step(w.dt)
            sim.detect_cut(tool)
        assert t._solo is not None, "the block qualifies for the persistent step"
        if s == 12:
            t.compact(apply=True)  # the persistent view follows the shortened lists
        torch.cuda.synchronize()
        po, vo = o.state()
        np.testing.assert_array_equal(t.positions(), po, err_msg=f"positions after step {s}")
        np.testing.assert_array_equal(t.velocities(), vo, err_msg=f"velocities after step {s}")
    np.testing.assert_array_equal(t.rotations(), o.rotations())
    ct, co = t.cut_state(), o.cut_state()
    assert co["revision"] > 0
    np.testing.assert_array_equal(ct["neighbors"], co["neighbors"])
    np.testing.assert_array_equal(ct["cut_points"], co["cut_points"])
    if iterations > 1:
        for a, b in zip(t.multipliers(), o.multipliers()):
            np.testing.assert_array_equal(a, b)


def test_large_or_contacting_tissue_takes_the_colour_launches():
    from genesis_tissue_simulation_amd.tissue import Tissue
    from genesis_tissue_simulation_amd.workloads import kidney_250k

    w = block_6144()
    kd, kh = w.stiffness()
    t = Tissue(w.grid, w.params.with_(contact_enabled=True), kd, kh, w.initial_positions, persistent=True)
    t.set_gravity(w.gravity)
    tool = dict(shaftA=(-1.0, 3.0, 1.0), shaftB=(5.0, 3.0, 1.0), shaftR=0.1)
    t.step(w.dt, tool=tool, tool_armed=True)
    assert t._solo is None
    t.step(w.dt)
    assert t._solo is not None
    k = kidney_250k(spacing=0.5)
    kd, kh = k.stiffness()
    big = Tissue(k.grid, k.params.with_(contact_enabled=False), kd, kh, k.initial_positions, persistent=True)
    big.step(k.dt)
    assert big._solo is None  # colours of the kidney lattice exceed the persistent limit
