"""bench.py's roofline fields read two per-dispatch figures of the stage kernels from the
committed counter profile: the counter-measured HBM bytes (roofline.traffic) and the VALU
instructions (roofline.valu_issue_frac).  CPU only."""
import bench


def test_profile_lines_read_by_the_bench():
    for kernel in bench.LAUNCH_NAMES:
        t = bench.profiled_traffic(kernel)
        v = bench.profiled_value('valu_insts', kernel)
        assert t is not None and v is not None, kernel
        # bytes per dispatch between the algorithmic bytes and 3x them at 385x385x115
        ncell = bench.NX * bench.NY * bench.NZ
        algo = bench.ALGO_BYTES[True][bench.LAUNCH_NAMES.index(kernel)] * ncell
        assert algo < t < 3 * algo, (kernel, t, algo)
        # VALU instructions per 64 cells within the range the profile states (3,300 - 3,700)
        per64 = v / (ncell / 64.0)
        assert 3000 < per64 < 4000, (kernel, per64)
    assert bench.profiled_value('no_such_key', 'visc_stage2') is None
