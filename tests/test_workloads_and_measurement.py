"""Workload definitions and the measurement helpers of bench.py (no GPU)."""
import json

import numpy as np

from genesis_tissue_simulation_amd.workloads import (GRAVITY, adjoint_instance, all_blades, block_6144,
                                                     cutting_block, tet_count)


def test_block_workload_sizes_and_schedule():
    w = block_6144()
    assert tet_count(w.grid) == 6144 and len(w.grid.positions) == 17 * 9 * 9
    assert w.params.substeps == 10 and w.params.iterations == 1 and w.steps == 200
    assert w.gravity_at(0) == GRAVITY and w.gravity_at(500) == GRAVITY
    S, E = w.blade(0)
    assert abs(float(np.linalg.norm(E - S)) - 4 * w.info["spacing"]) < 1e-6


def test_cutting_block_sizes_blades_and_gravity_ramp():
    w = cutting_block()
    assert w.grid.dims == (111, 55, 55) and tet_count(w.grid) == 6 * 111 * 55 * 55 == 2_014_650
    assert len(all_blades(w)) == 3
    assert w.params.substeps == 20 and w.params.iterations == 2
    # NativeTissue.step's ramp: min(1, (tick + 1) dt / 0.5)
    assert np.isclose(w.gravity_at(0)[1], GRAVITY[1] * (1 / 60) / 0.5)
    assert w.gravity_at(29) == GRAVITY
    hi = w.grid.positions.max(axis=0)
    for blade in all_blades(w):
        S, E = blade(0)
        assert S[1] > hi[1] > E[1]           # vertical, entering from above the top face
        S1, _ = blade(w.steps)
        assert not np.allclose(S, S1)        # the path moves


def test_adjoint_instance_loads_and_mask():
    inst = adjoint_instance(seed=2)
    T = tet_count(inst.grid)
    assert 20_000 <= T <= 21_000 and len(inst.compliance_hidden) == T
    assert np.allclose(inst.compliance_init, 1e-7)
    assert np.isclose(np.median(inst.compliance_hidden), 1e-7, rtol=0.2)
    g = np.asarray(GRAVITY, dtype=np.float64)[None, :] * inst.grid.mass[:, None]
    loaded = np.abs(inst.force - g).max(axis=1) > 1e-3
    top = np.isclose(inst.grid.positions[:, 1], inst.grid.positions[:, 1].max())
    assert loaded.any() and not (loaded & ~top).any()
    assert abs(loaded.sum() - 0.05 * top.sum()) <= 1
    assert inst.mask[inst.grid.pinned != 0].sum() == 0 and inst.mask.sum() > 0


def _counter_csv(path, counter, rows):
    head = ("Correlation_Id,Dispatch_Id,Agent_Id,Queue_Id,Process_Id,Thread_Id,Grid_Size,Kernel_Id,Kernel_Name,"
            "Workgroup_Size,LDS_Block_Size,Scratch_Size,VGPR_Count,Accum_VGPR_Count,SGPR_Count,Counter_Name,"
            "Counter_Value,Start_Timestamp,End_Timestamp\n")
    with open(path, "w") as f:
        f.write(head)
        for i, (name, value) in enumerate(rows):
            f.write(f'{i},{i},"Agent 1",1,1,1,64,1,"{name}",64,0,0,8,0,16,"{counter}",{value},0,1\n')


def test_traffic_file_scales_each_counter_by_the_copy(tmp_path, monkeypatch):
    import bench

    fetch, write = tmp_path / "f.csv", tmp_path / "w.csv"
    copy = "(anonymous namespace)::k_stream_copy(long long, HIP_vector_type<float, 4u> const*, HIP_vector_type<float, 4u>*)"
    tet = "void (anonymous namespace)::k_tets<1, false, true, false>(int)"
    # the copy of 1 MiB reads half its size on FETCH_SIZE and all of it on WRITE_SIZE
    _counter_csv(fetch, "FETCH_SIZE", [(copy, 512.0), (tet, 10.0), (tet, 30.0)])
    _counter_csv(write, "WRITE_SIZE", [(copy, 1024.0), (tet, 4.0), (tet, 8.0)])
    monkeypatch.setattr(bench, "TRAFFIC_FILE", tmp_path / "traffic.json")
    bench.write_traffic(str(fetch), str(write), 1, "test")
    out = json.loads((tmp_path / "traffic.json").read_text())
    assert out["calibration"]["fetch_factor"] == 2.0 and out["calibration"]["write_factor"] == 1.0
    k = out["kernels"]["k_tets"]
    assert k["dispatches"] == 2
    assert k["read_bytes_per_launch"] == 20.0 * 2.0 * 1024 and k["write_bytes_per_launch"] == 6.0 * 1024
    assert bench.load_traffic("k_tets")["bytes_per_launch"] == (40.0 + 6.0) * 1024


def test_position_error_is_node_distance_over_largest_coordinate():
    import bench

    x64 = np.array([[1.0, 0.0, 0.0], [0.0, -4.0, 0.0]])
    x = x64 + np.array([[0.0, 3e-4, 4e-4], [0.0, 0.0, 0.0]])
    e = bench.position_errors(x, x64)
    assert np.isclose(e["rel_err_vs_fp64"], 5e-4 / 4.0)
    assert np.isclose(e["rms_rel_err_vs_fp64"], np.sqrt((5e-4) ** 2 / 2) / 4.0)
